"""GPU parity of the HIP engine against the CPU oracle (run on an MI355X).

Bars (FP32 path): E-step statistics within 2e-5 (relative Frobenius norm
per statistic) and the log-likelihood sum within 1e-7 of the FP64 oracle on
identical float32 inputs; the LL trajectory of a full run within 1e-6
relative, identical branch sequences (Newton / natural gradient, LL
decrease) and the final W within an Amari distance of 1e-4, wherever the
trajectory itself allows it: a full run is held to max(1e-6, 4x the
oracle's own spread when its input is rounded to float32 or carries
float32-sized noise), iteration by iteration.  BF16 path: E-step
statistics within 1e-3 and the first LL within 1e-6; later iterations
within max(1e-6, 4x the oracle's spread under input perturbations the size
of the BF16 unmixing's rounding, 2^-17).
"""
import numpy as np
import pytest

import cudamica_amd as ca
from oracle import pyoracle as po
from tests.conftest import amari_distance

pytestmark = pytest.mark.gpu

STAT_NAMES = ("usum", "kappa_numer", "lambda_numer", "mu_numer", "mu_denom", "beta_numer",
              "beta_denom", "rho_numer", "sigma2_numer", "cnew", "phi")


def _say(msg):
    print(msg, flush=True)


def _random_state(n, M, m, seed):
    st = ca.synth_init(n, M, m, seed)
    rng = np.random.default_rng(seed)
    st["rho"] = rng.uniform(1.0, 2.0, size=st["rho"].shape)
    st["mu"] = st["mu"] + rng.normal(scale=0.2, size=st["mu"].shape)
    st["beta"] = rng.uniform(0.6, 1.6, size=st["beta"].shape)
    a = rng.uniform(0.2, 1.0, size=(M, m, n))
    st["alpha"] = (a / a.sum(axis=1, keepdims=True)).ravel()
    A = st["A"].reshape(M, n, n)
    A += rng.normal(scale=0.05, size=A.shape)
    st["A"] = A.ravel()
    g = rng.uniform(0.5, 1.0, size=M)
    st["gm"] = g / g.sum()
    return st


def _compare_stats(gpu, ora, n, m, M, tol=2e-5, ll_tol=1e-7, skip=(), vsum_tol=1e-6):
    """Relative Frobenius error per statistic; cnew (sum_t v x, close to 0 for
    centred data and M = 1) is compared in absolute terms, as cnew/vsum in
    units of the whitened data's standard deviation (2)."""
    gv, gll = po.stats_views(gpu, n, m, M)
    ov, oll = po.stats_views(ora, n, m, M)
    worst = {}
    for h in range(M):
        assert abs(gv[h]["vsum"] - ov[h]["vsum"]) <= vsum_tol * abs(ov[h]["vsum"]), (h, gv[h]["vsum"], ov[h]["vsum"])
        for k in STAT_NAMES:
            a, b = np.asarray(gv[h][k]), np.asarray(ov[h][k])
            if k == "cnew":
                rel = np.max(np.abs(a - b)) / ov[h]["vsum"] / 2.0
            else:
                rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)
            worst[k] = max(worst.get(k, 0.0), rel)
    _say(f"  worst relative errors: { {k: f'{v:.2e}' for k, v in worst.items()} }; "
         f"LL sum {abs(gll - oll) / abs(oll):.2e}")
    for k, v in worst.items():
        if k not in skip:
            assert v < tol, f"{k}: relative error {v:.3e} >= {tol}"
    assert abs(gll - oll) <= ll_tol * abs(oll), (gll, oll)


@pytest.mark.parametrize("n,T,M,m", [
    (16, 4000, 1, 3), (64, 20000, 1, 3), (71, 6000, 2, 2), (32, 5000, 1, 1),
    (128, 20000, 4, 3), (256, 8000, 2, 4)])
def test_estep_stats_fp32(built, n, T, M, m):
    _say(f"[estep fp32] n={n} T={T} M={M} m={m}")
    spec = ca.synth_spec(n, T, seed=100 + n + M, cond_max=0.0)
    x = ca.synth_data(spec)
    xw, S, mean, ldetS = po.preprocess(x, 1)
    x32 = xw.astype(np.float32).astype(np.float64)
    st = _random_state(n, M, m, seed=7 + n)
    ora = po.estep(x32, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                   st["gm"], M, m)
    with ca.Engine(ca.options(n, T, M, m, precision=ca.PREC_FP32)) as eng:
        eng.load_white(x32, ldetS)
        eng.set_state(st)
        gpu = eng.estep_stats()
    _compare_stats(gpu, ora, n, m, M)


@pytest.mark.parametrize("n,T,M,m", [(128, 20000, 4, 3), (71, 20011, 2, 3), (96, 12345, 2, 4)])
def test_estep_stats_bf16(built, n, T, M, m):
    """Split-BF16 unmixing (~2^-16 relative b): statistics within 1e-3, the
    LL sum within 1e-5; mu_denom is reported, not bounded (its weight
    |y|^(rho-2) is singular at y = 0, where b's rounding dominates).  The
    odd sample counts end in a partial tile and a partial 8-sample group of
    the packed bf16 g, and n = 71 and 96 leave padding channels in the last
    64-channel block; m = 4 runs the one-sample tiles."""
    _say(f"[estep bf16] n={n} T={T} M={M} m={m}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=300 + n, cond_max=0.0))
    xw, S, mean, ldetS = po.preprocess(x, 1)
    x32 = xw.astype(np.float32).astype(np.float64)
    st = _random_state(n, M, m, seed=301 + n)
    ora = po.estep(x32, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                   st["gm"], M, m)
    with ca.Engine(ca.options(n, T, M, m, precision=ca.PREC_BF16)) as eng:
        eng.load_white(x32, ldetS)
        eng.set_state(st)
        gpu = eng.estep_stats()
    _compare_stats(gpu, ora, n, m, M, tol=1e-3, ll_tol=1e-5, skip=("mu_denom",), vsum_tol=1e-4)


@pytest.mark.parametrize("sph", [1, 2, 0])
def test_preprocess_parity(built, sph):
    n, T = 48, 30000
    _say(f"[preprocess] do_sphere={sph}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=55, cond_max=0.0))
    x = x + np.linspace(-3, 3, n, dtype=np.float32)[None]     # non-zero channel means
    xw, S, mean, ldetS = po.preprocess(x, sph)
    with ca.Engine(ca.options(n, T, 1, 3)) as eng:
        eng.load_raw(x)
        eng.preprocess(sph)
        Sg, mg, lg = eng.sphere()
        wg = eng.white()
    np.testing.assert_allclose(mg, mean, rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(Sg, S, rtol=1e-9, atol=1e-11 * np.abs(S).max())
    assert abs(lg - ldetS) < 1e-9 * max(1.0, abs(ldetS))
    assert np.max(np.abs(wg - xw)) <= 2e-6 * np.max(np.abs(xw))


def _run_both(w, T, iters, precision, newt=None, graph=False, flags=None):
    n, M, m = w.n, w.M, w.m
    x = ca.synth_data(w.spec(T=T))
    st0 = w.init_state()
    xw, S, mean, ldetS = po.preprocess(x, w.do_sphere)
    nsi = w.newt_start_iter if newt is None else newt
    flags = flags or {}
    cfg = po.default_config(n, T, M, m, maxiter=iters, newt_start_iter=nsi, **flags)
    fin, logs = po.run(cfg, xw, ldetS, st0)
    with ca.Engine(ca.options(n, T, M, m, precision=precision, max_iter=iters,
                              newt_start_iter=nsi, **flags)) as eng:
        eng.load_raw(x)
        eng.preprocess(w.do_sphere)
        eng.set_state(st0)
        eng.iterate(iters, graph=graph)
        eng.sync()
        glog = eng.log(iters)
        gst = eng.get_state()

    def margin(k):
        """min |kappa_i s2_i kappa_k s2_k - 1| of the oracle's Hessian test at
        0-based iteration k of its own trajectory (kernels.cu:975-979)."""
        if k == 0:
            sk, llp = {kk: np.asarray(v, np.float64).ravel() for kk, v in st0.items()}, 0.0
        else:
            ck = po.default_config(n, T, M, m, maxiter=k, newt_start_iter=nsi, **flags)
            sk, lk = po.run(ck, xw, ldetS, st0)
            llp = lk["ll"][-1]
        lr = logs["lrate"][k]
        stats = po.estep(xw, ldetS, sk["A"], sk["c"], sk["alpha"], sk["mu"], sk["beta"], sk["rho"],
                         sk["gm"], M, m, update_rho=cfg.update_rho)
        W = np.concatenate([np.linalg.inv(sk["A"][h * n * n:(h + 1) * n * n].reshape(n, n).T).T.ravel()
                            for h in range(M)])
        _, _, _, tr = po.mstep(cfg, k + 1, stats, W, sk, lr, ll_prev=llp, trace=True)
        ks = (tr["kappa"] * tr["sigma2"]).reshape(M, n)
        return float(min(np.min(np.abs(np.outer(v, v) - 1.0)) for v in ks))

    return fin, logs, gst, glog, margin


def test_bf16_mu_denom_at_exact_cancellation(built):
    """At the workload's initial state mu is exactly -1, 0 or 1, so over
    ~10^9 elements some b round onto mu in FP32 and y cancels to 0; with
    rho = 1.5 the weight |y|^(rho-2) of such an element must stay at the
    BF16 path's resolution (2^-17) instead of 2^63.  The split-BF16
    statistics are compared with the FP32 path (FP64 unmixing, held to the
    oracle by the tests above): mu_denom within 5e-2 (its error scales like
    sqrt of b's rounding there), mu after one iteration within 1e-2."""
    w = ca.WORKLOADS["n128_T2M_M4"]
    n, M, m, T = w.n, w.M, w.m, 400000
    _say(f"[bf16 mu_denom] n={n} T={T} M={M} m={m}, initial state of the workload")
    x = ca.synth_data(w.spec(T=T))
    st0 = w.init_state()
    res = {}
    for p in (ca.PREC_BF16, ca.PREC_FP32):
        with ca.Engine(w.options(T=T, precision=p, max_iter=2)) as eng:
            eng.load_raw(x)
            eng.preprocess(w.do_sphere)
            eng.set_state(st0)
            stats = eng.estep_stats()
            eng.iterate(1)
            eng.sync()
            res[p] = (stats, eng.get_state())
    a, _ = po.stats_views(res[ca.PREC_BF16][0], n, m, M)
    b, _ = po.stats_views(res[ca.PREC_FP32][0], n, m, M)
    md = max(np.linalg.norm(a[h]["mu_denom"] - b[h]["mu_denom"]) / np.linalg.norm(b[h]["mu_denom"])
             for h in range(M))
    mu = (np.linalg.norm(res[ca.PREC_BF16][1]["mu"] - res[ca.PREC_FP32][1]["mu"])
          / np.linalg.norm(res[ca.PREC_FP32][1]["mu"]))
    _say(f"  mu_denom relative difference {md:.2e}; mu after one iteration {mu:.2e}")
    assert md <= 5e-2
    assert mu <= 1e-2


def _oracle_envelope(w, T, iters, newt, fin, logs, seeds=(1, 2), eps=6e-8, flags=None):
    """Conditioning of the EM trajectory itself: the FP64 oracle rerun on its
    whitened input perturbed by float32-rounding-sized noise (once rounded to
    float32, once with relative N(0, eps) noise per seed).  Returns the
    cumulative maximum over the reruns of the relative LL difference per
    iteration and the largest Amari distance between the final W's.  Where
    every rho sits at rho = 1 the mu update divides by sum(u/|y|), which the
    sample nearest y = 0 dominates, so a 1e-15 relative input change already
    grows to ~1e-4 in LL within 30 iterations."""
    n, M, m = w.n, w.M, w.m
    x = ca.synth_data(w.spec(T=T))
    st0 = w.init_state()
    xw, _, _, ldetS = po.preprocess(x, w.do_sphere)
    cfg = po.default_config(n, T, M, m, maxiter=iters, newt_start_iter=newt, **(flags or {}))
    inputs = [xw.astype(np.float32).astype(np.float64)]
    for s in seeds:
        rng = np.random.default_rng(s)
        inputs.append(xw * (1.0 + eps * rng.standard_normal(xw.shape)))
    env = np.zeros(iters)
    wenv = 0.0
    for xp in inputs:
        fp, lp = po.run(cfg, xp, ldetS, st0)
        env = np.maximum(env, np.abs(lp["ll"] - logs["ll"]) / np.abs(logs["ll"]))
        for h in range(M):
            Wp = fp["W"][h * n * n:(h + 1) * n * n].reshape(n, n).T
            Wo = fin["W"][h * n * n:(h + 1) * n * n].reshape(n, n).T
            wenv = max(wenv, amari_distance(Wp, np.linalg.inv(Wo)))
    return np.maximum.accumulate(env), wenv


def _check_trajectory(fin, logs, gst, glog, n, M, ll_tol, w_tol, env=None, wenv=0.0, margin=None):
    """Branch-aware trajectory check.  The LL-decrease test and the Newton
    test are discrete decisions; a near-tie can go either way under
    rounding, after which the trajectories legitimately differ.  So: the LL
    must agree within ll_tol up to and including the first iteration whose
    decisions differ, and a differing LL-decrease decision must be a
    near-tie of the oracle's own trajectory.  Without a differing decision
    the whole trajectory, the lrate schedule and the final W are checked.
    With env (the oracle's own sensitivity, _oracle_envelope) the LL bar per
    iteration is max(ll_tol, 4 env) and the W bar max(w_tol, 4 wenv)."""
    K = len(logs["ll"])
    tol = np.full(K, ll_tol) if env is None else np.maximum(ll_tol, 4.0 * env)
    w_tol = max(w_tol, 4.0 * wenv)
    same = ((glog["newton"] == logs["newton"]).all(axis=1)
            & ((glog["bflag"] > 0) == (logs["bflag"] > 0)).all(axis=1)
            & (glog["lldec"] == logs["lldec"]))
    kdiv = K if same.all() else int(np.argmin(same))
    rel = np.abs(glog["ll"] - logs["ll"]) / np.abs(logs["ll"])
    _say(f"  LL max rel err {rel[:kdiv + 1].max():.3e} over {min(kdiv + 1, K)} iterations; "
         f"final LL gpu {glog['ll'][-1]:.12f} oracle {logs['ll'][-1]:.12f}")
    if env is not None:
        _say(f"  oracle envelope at the worst iteration: {tol[np.argmax(rel[:kdiv + 1] / tol[:kdiv + 1])]:.3e}")
    assert (rel[:kdiv + 1] <= tol[:kdiv + 1]).all()
    dists = []
    for h in range(M):
        Wg = gst["W"][h * n * n:(h + 1) * n * n].reshape(n, n).T
        Wo = fin["W"][h * n * n:(h + 1) * n * n].reshape(n, n).T
        dists.append(amari_distance(Wg, np.linalg.inv(Wo)))
    _say(f"  final Amari(W_gpu, W_oracle) per model: {[f'{d:.3e}' for d in dists]}")
    if kdiv < K:
        # the first differing decision must be a near-tie of the oracle's own
        # test: the LL step for the decrease rule, the Hessian-test margin for
        # the Newton / bflag decisions
        step = abs(logs["ll"][kdiv] - logs["ll"][kdiv - 1]) / abs(logs["ll"][kdiv]) if kdiv > 0 else np.inf
        _say(f"  decisions first differ at iteration {kdiv + 1}: oracle LL step {step:.2e}")
        if glog["lldec"][kdiv] != logs["lldec"][kdiv]:
            assert step <= 10 * tol[kdiv], "an LL-decrease decision differs without a near-tie"
        else:
            assert margin is not None, "a Newton / Hessian-test decision differs"
            mg = margin(kdiv)
            _say(f"  oracle Hessian-test margin there: {mg:.2e}")
            assert mg <= 1e-4, "a Newton / Hessian-test decision differs without a near-tie"
        return
    if env is None:
        np.testing.assert_allclose(glog["lrate"], logs["lrate"], rtol=1e-12)
    assert max(dists) <= w_tol


def test_full_run_fp32_n16(built):
    w = ca.WORKLOADS["n16_T20k"]
    _say("[run fp32] n16_T20k 50 iterations")
    fin, logs, gst, glog, margin = _run_both(w, w.T, 50, ca.PREC_FP32)
    # the oracle itself moves by up to 3.4e-6 (iteration 5) when its input is
    # rounded to float32, so 1e-6 holds only where the trajectory allows it
    env, wenv = _oracle_envelope(w, w.T, 50, w.newt_start_iter, fin, logs)
    _check_trajectory(fin, logs, gst, glog, w.n, w.M, 1e-6, 1e-4, env=env, wenv=wenv, margin=margin)


def test_full_run_fp32_newton(built):
    w = ca.WORKLOADS["n64_T1M"]
    _say("[run fp32] n64 subsample T=60000, 30 iterations, Newton from 10")
    fin, logs, gst, glog, margin = _run_both(w, 60000, 30, ca.PREC_FP32, newt=10)
    assert glog["newton"].sum() > 0
    env, wenv = _oracle_envelope(w, 60000, 30, 10, fin, logs)
    _check_trajectory(fin, logs, gst, glog, w.n, w.M, 1e-6, 1e-4, env=env, wenv=wenv, margin=margin)


@pytest.mark.parametrize("flags", [
    dict(do_reparm=0),
    dict(update_A=0),
    dict(update_mu=0, update_beta=0, update_rho=0),
    dict(update_alpha=0, update_gm=0, do_newton=0),
], ids=["no_reparm", "no_A", "no_mu_beta_rho", "no_alpha_gm_newton"])
def test_full_run_update_flags(built, flags):
    """The M-step's switches (config.c keys update_A, update_mu, ...,
    do_reparm, do_newton): 12 iterations of a 2-model run with Newton from
    iteration 4, each switch set off against the oracle with the same
    switches."""
    w = ca.WORKLOADS["n16_T20k"]
    T, iters, newt = 20000, 12, 4
    wf = ca.Workload("n16_M2_flags", n=16, T=T, M=2, m=3, iters=iters, seed=w.seed,
                     do_sphere=w.do_sphere, newt_start_iter=newt)
    _say(f"[run fp32 flags] n16 M2 T={T} {flags}")
    fin, logs, gst, glog, margin = _run_both(wf, T, iters, ca.PREC_FP32, newt=newt, flags=flags)
    env, wenv = _oracle_envelope(wf, T, iters, newt, fin, logs, flags=flags)
    _check_trajectory(fin, logs, gst, glog, wf.n, wf.M, 1e-6, 1e-4, env=env, wenv=wenv, margin=margin)
    if flags.get("update_A", 1) == 0:
        np.testing.assert_array_equal(gst["A"], wf.init_state()["A"])


def test_full_run_bf16_models(built):
    """BF16 GEMMs round W x to ~2^-17 and g, x to 2^-9 in Phi.  This
    trajectory is ill-conditioned (the oracle itself moves by 3e-5 in LL
    within 4 iterations under float32-sized input perturbations, 6e-5 under
    2^-17-sized ones), so the bar per iteration is max(1e-6, 4x the oracle's
    spread under 2^-17 perturbations of its input), and the first
    iteration (the E-step alone) must agree within 1e-6."""
    w = ca.WORKLOADS["n128_T2M_M4"]
    T = 100000
    _say(f"[run bf16] n128 M4 subsample T={T}, 8 iterations")
    fin, logs, gst, glog, _ = _run_both(w, T, 8, ca.PREC_BF16)
    rel = np.abs(glog["ll"] - logs["ll"]) / np.abs(logs["ll"])
    env, _ = _oracle_envelope(w, T, 8, w.newt_start_iter, fin, logs, eps=2.0 ** -17)
    fmt = dict(formatter={"float_kind": lambda v: f"{v:.2e}"})
    _say(f"  bf16 LL rel err per iteration: {np.array2string(rel, **fmt)}")
    _say(f"  oracle spread under 2^-17 input perturbation: {np.array2string(env, **fmt)}")
    assert rel[0] <= 1e-6
    assert np.all(rel <= np.maximum(1e-6, 4.0 * env))


def test_graph_replay_matches_eager(built):
    n, T, M, m = 32, 12000, 2, 3
    _say("[graph] replay vs eager")
    x = ca.synth_data(ca.synth_spec(n, T, seed=77, cond_max=0.0))
    st0 = ca.synth_init(n, M, m, 78)
    res = []
    for graph in (False, True):
        with ca.Engine(ca.options(n, T, M, m, max_iter=6, newt_start_iter=3)) as eng:
            eng.load_raw(x)
            eng.preprocess(1)
            eng.set_state(st0)
            for k in range(3):       # replays separated by host reads of the logs
                eng.iterate(2, graph=graph)
                eng.sync()
                eng.log(2 * (k + 1))
            res.append((eng.log(6), eng.get_state()))
    (la, sa), (lb, sb) = res
    np.testing.assert_array_equal(la["ll"], lb["ll"])
    for k in sa:
        np.testing.assert_array_equal(sa[k], sb[k], err_msg=k)


def test_ltall_and_output_transform(built):
    w = ca.WORKLOADS["n16_T20k"]
    n, M, m, T = w.n, w.M, w.m, 8000
    _say("[ltall/output] n16 T=8000 3 iterations")
    x = ca.synth_data(w.spec(T=T))
    st0 = w.init_state()
    xw, S, mean, ldetS = po.preprocess(x, 1)
    cfg = po.default_config(n, T, M, m, maxiter=3)
    fin, logs = po.run(cfg, xw, ldetS, st0, want_ltall=True)
    Ao, co = po.output_transform(S, mean, fin["A"], fin["c"], n, M)
    with ca.Engine(ca.options(n, T, M, m, max_iter=3)) as eng:
        eng.load_raw(x)
        eng.preprocess(1)
        eng.set_state(st0)
        eng.iterate(2)
        eng.record_ltall(True)
        eng.iterate(1)
        lt = eng.ltall()
        Ag, cg = eng.output()
    np.testing.assert_allclose(lt, logs["ltall"], rtol=2e-6, atol=2e-6 * np.abs(logs["ltall"]).max())
    np.testing.assert_allclose(Ag, Ao, rtol=1e-4, atol=1e-5)
    np.testing.assert_allclose(cg, co, rtol=1e-4, atol=1e-5)


def test_ldetS_reaches_state_set_before_the_data(built):
    """ldetS enters the log-likelihood constant through the device state:
    whitened data loaded (or sphered) after set_state must still be counted
    with their own ldetS."""
    n, T, M, m = 24, 6000, 2, 3
    _say(f"[ldetS order] n={n} T={T} M={M}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=515, cond_max=0.0))
    xw, _, _, ldetS = po.preprocess(x, 1)
    x32 = xw.astype(np.float32).astype(np.float64)
    st = _random_state(n, M, m, seed=516)
    ora = po.estep(x32, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                   st["gm"], M, m)
    with ca.Engine(ca.options(n, T, M, m)) as eng:
        eng.load_white(x32, ldetS + 3.0)
        eng.set_state(st)
        eng.load_white(x32, ldetS)
        a = eng.estep_stats()
        eng.load_raw(x)
        eng.preprocess(1)
        b = eng.estep_stats()
    _compare_stats(a, ora, n, m, M)
    # the GPU's own sphering: float32-rounded data of its own, so mu_denom (singular
    # weight at y = 0) is reported, not bounded
    _compare_stats(b, ora, n, m, M, tol=1e-4, ll_tol=1e-6, skip=("mu_denom",))
