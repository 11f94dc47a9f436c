"""GPU parity of the unmixing and E-step variants that only wider channel
counts reach, and of the 256-channel FP8 workload's own generator.

Which kernel each shape runs (NP = 64 ceil(n / 64)):
* split-BF16 and FP8 unmixing with register-resident W: NP = 192 runs
  unmix_lds_kernel<2, 6, *>, NP = 256 unmix_lds_kernel<2, 8, *> (three
  prefetch stages for BF16, two waves per SIMD for FP8);
* NP > 256 (BF16): the LDS-tiled unmix_bf16x3_kernel;
* M = 1 recordings: the single-model E-step, M >= 2 the fused multi-model
  E-step (m = 4: one-sample tiles).

Bars: E-step statistics within 1e-3 of the FP64 oracle on the same float32
(BF16) or dequantised (FP8) input, the log-likelihood sum within 1e-5
(mu_denom reported, not bounded: its weight |y|^(rho-2) is singular at
y = 0, where the unmixing's 2^-17 rounding dominates); runs within
max(1e-6, 4x the oracle's own spread under input perturbations of the
unmixing's rounding) per iteration, the first iteration within 1e-6.
"""
import numpy as np
import pytest

import cudamica_amd as ca
from oracle import pyoracle as po
from tests.test_gpu_engine import _compare_stats, _random_state

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("n,T,M,m", [(150, 20011, 2, 3), (256, 12000, 2, 4), (400, 6000, 1, 3),
                                     (512, 5003, 1, 2)])
def test_estep_stats_bf16_wide(built, n, T, M, m):
    _say(f"[estep bf16 wide] n={n} T={T} M={M} m={m}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=900 + n, cond_max=0.0))
    xw, _, _, ldetS = po.preprocess(x, 1)
    x32 = xw.astype(np.float32).astype(np.float64)
    st = _random_state(n, M, m, seed=901 + n)
    ora = po.estep(x32, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                   st["gm"], M, m)
    with ca.Engine(ca.options(n, T, M, m, precision=ca.PREC_BF16)) as eng:
        eng.load_white(x32, ldetS)
        eng.set_state(st)
        gpu = eng.estep_stats()
    _compare_stats(gpu, ora, n, m, M, tol=1e-3, ll_tol=1e-5, skip=("mu_denom",), vsum_tol=1e-4)


@pytest.mark.parametrize("n,T,M,m", [(150, 20000, 2, 3), (256, 16000, 2, 4)])
def test_estep_stats_fp8_wide(built, n, T, M, m):
    _say(f"[estep fp8 wide] n={n} T={T} M={M} m={m}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=910 + n, cond_max=0.0))
    st = _random_state(n, M, m, seed=911 + n)
    with ca.Engine(ca.options(n, T, M, m, precision=ca.PREC_FP8)) as eng:
        eng.load_raw(x)
        eng.preprocess(1)
        xh = eng.white().astype(np.float64)
        _, _, ldetS = eng.sphere()
        eng.set_state(st)
        gpu = eng.estep_stats()
    ora = po.estep(xh, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                   st["gm"], M, m)
    _compare_stats(gpu, ora, n, m, M, tol=1e-3, ll_tol=1e-5, skip=("mu_denom",), vsum_tol=1e-4)


def _envelope(cfg, xin, ldetS, st0, ref_ll, eps, seeds=(1, 2)):
    env = np.zeros(cfg.maxiter)
    for s in seeds:
        xp = xin * (1.0 + eps * np.random.default_rng(s).standard_normal(xin.shape))
        _, lp = po.run(cfg, xp, ldetS, st0)
        env = np.maximum(env, np.abs(lp["ll"] - ref_ll) / np.abs(ref_ll))
    return np.maximum.accumulate(env)


def _c4_subsample(T):
    """The 256-channel, 2-model, 4-component workload's generator and seeds
    (models alternating block by block) at T samples, with blocks of T / 3
    so that both mixing matrices occur."""
    w = ca.WORKLOADS["n256_T4M_M2"]
    spec = w.spec(T=T)
    spec.block = T // 3
    return w, ca.synth_data(spec), w.init_state()


@pytest.mark.parametrize("precision", [ca.PREC_FP8, ca.PREC_BF16], ids=["fp8", "bf16"])
def test_c4_workload_run_follows_oracle(built, precision):
    """256 ch, M = 2 alternating blocks, m = 4: FP8 e4m3 storage with
    per-channel scales (unmix_lds_kernel<2, 8, true>) and split BF16
    (unmix_lds_kernel<2, 8, false>), 5 EM iterations against the FP64
    oracle on the data the path runs on (FP8: the dequantised x^)."""
    T, iters = 45000, 5
    w, x, st0 = _c4_subsample(T)
    n, M, m = w.n, w.M, w.m
    _say(f"[run c4 {'fp8' if precision == ca.PREC_FP8 else 'bf16'}] n={n} T={T} M={M} m={m}, "
         f"{iters} iterations")
    with ca.Engine(w.options(T=T, precision=precision, max_iter=iters)) as eng:
        eng.load_raw(x)
        eng.preprocess(1)
        xh = eng.white().astype(np.float64)
        _, _, ldetS = eng.sphere()
        eng.set_state(st0)
        eng.iterate(iters)
        eng.sync()
        glog = eng.log(iters)
    if precision == ca.PREC_BF16:
        xw, _, _, ldetS = po.preprocess(x, 1)
        xin = xw.astype(np.float32).astype(np.float64)
    else:
        xin = xh
    cfg = po.default_config(n, T, M, m, maxiter=iters, newt_start_iter=w.newt_start_iter)
    _, lq = po.run(cfg, xin, ldetS, st0)
    env = _envelope(cfg, xin, ldetS, st0, lq["ll"], 2.0 ** -17)
    rel = np.abs(glog["ll"] - lq["ll"]) / np.abs(lq["ll"])
    fmt = dict(formatter={"float_kind": lambda v: f"{v:.2e}"})
    _say(f"  LL rel. error per iteration: {np.array2string(rel, **fmt)}")
    _say(f"  oracle spread under 2^-17 perturbation: {np.array2string(env, **fmt)}")
    assert rel[0] <= 1e-6
    assert np.all(rel <= np.maximum(1e-6, 4.0 * env))


@pytest.mark.parametrize("precision,policy", [(ca.PREC_FP32, 0), (ca.PREC_BF16, 0), (ca.PREC_BF16, 8)],
                         ids=["fp32", "bf16", "bf16-grouped"])
def test_batch_workload_each_recording_follows_oracle(built, precision, policy):
    """The batched-recordings workload's generator (n drawn from {32, 64,
    128}, M = 1, m = 3, each recording its own seed), shortened to
    T in [20k, 50k]: six recordings iterated as one batch, each one's LL
    trajectory checked against its own FP64 oracle run: the first iteration
    within 1e-7 (FP32) / 1e-6 (BF16), later ones within max(1e-6, 4x the
    oracle's spread under float32-sized (FP32) or 2^-17 (BF16) input
    perturbations); BF16 also with the grouped launches over all recordings
    (each recording's own partition)."""
    iters = 4
    recs = ca.batch_workload(count=6, precision=precision, t_range=(20_000, 50_000), iters=iters)
    data, states, opts = [], [], []
    for w in recs:
        data.append(ca.synth_data(w.spec()))
        states.append(w.init_state())
        opts.append(w.options(max_iter=iters))
    _say(f"[batch vs oracle] {len(recs)} recordings {[(w.n, w.T) for w in recs]}, "
         f"precision {precision}, {iters} iterations")
    with ca.Batch(opts) as b:
        for e, x, st in zip(b.engines, data, states):
            e.load_raw(x)
            e.preprocess(1)
            e.set_state(st)
        b.set_policy(policy)
        b.iterate(iters)
        b.sync()
        logs = [e.log(iters) for e in b.engines]
    eps = 6e-8 if precision == ca.PREC_FP32 else 2.0 ** -17
    first_tol = 1e-7 if precision == ca.PREC_FP32 else 1e-6
    fmt = dict(formatter={"float_kind": lambda v: f"{v:.2e}"})
    for r, (w, x, st0, lg) in enumerate(zip(recs, data, states, logs)):
        xw, _, _, ldetS = po.preprocess(x, 1)
        xin = xw.astype(np.float32).astype(np.float64)
        cfg = po.default_config(w.n, w.T, w.M, w.m, maxiter=iters, newt_start_iter=w.newt_start_iter)
        _, lo = po.run(cfg, xin, ldetS, st0)
        env = _envelope(cfg, xin, ldetS, st0, lo["ll"], eps)
        rel = np.abs(lg["ll"] - lo["ll"]) / np.abs(lo["ll"])
        _say(f"  recording {r} n={w.n} T={w.T}: LL rel. error {np.array2string(rel, **fmt)}, "
             f"envelope {np.array2string(env, **fmt)}")
        assert rel[0] <= first_tol, r
        assert np.all(rel <= np.maximum(1e-6, 4.0 * env)), r


# The split E-step: shapes whose models' 64-channel blocks do not fit one
# 8-wave workgroup (M ceil(n/64) > 8) run three passes -- per-group channel
# sums of q, the model posteriors, then the statistics -- with workgroups of
# model groups x channel-block groups (the reference allows 1024 channels
# and 16 models, config.h:15-16).  (n, T, M, m, precision): 8 models of 128
# channels (2 model groups), 2 models of 512 (one model per group), 1024
# channels in two channel groups (m = 1), 16 models of 64 channels, and 300
# channels x 3 models in FP8 (the tiled FP8 unmixing kernel, NP = 320).
@pytest.mark.parametrize("n,T,M,m,prec", [
    (128, 12000, 8, 3, ca.PREC_FP32), (512, 5000, 2, 3, ca.PREC_BF16),
    (1024, 3001, 1, 1, ca.PREC_FP32), (64, 8000, 16, 2, ca.PREC_BF16),
    (300, 6000, 3, 4, ca.PREC_FP8)], ids=["n128M8", "n512M2", "n1024M1", "n64M16", "n300M3fp8"])
def test_estep_stats_split(built, n, T, M, m, prec):
    _say(f"[estep split] n={n} T={T} M={M} m={m} precision {prec}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=950 + n + M, cond_max=0.0))
    st = _random_state(n, M, m, seed=951 + n + M)
    with ca.Engine(ca.options(n, T, M, m, precision=prec)) as eng:
        if prec == ca.PREC_FP8:
            eng.load_raw(x)
            eng.preprocess(1)
            xin = eng.white().astype(np.float64)
            ldetS = eng.sphere()[2]
        elif n > 512:
            # the raw mixtures as the data (statistics need no sphering; the
            # oracle's Jacobi eigensolver is slow at 1024 channels)
            xin, ldetS = x.astype(np.float64), 0.0
            eng.load_white(xin, ldetS)
        else:
            xw, _, _, ldetS = po.preprocess(x, 1)
            xin = xw.astype(np.float32).astype(np.float64)
            eng.load_white(xin, ldetS)
        eng.set_state(st)
        gpu = eng.estep_stats()
    ora = po.estep(xin, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                   st["gm"], M, m)
    if prec == ca.PREC_FP32:
        # vsum of a model with 6 % of the samples (8 models): FP32 posteriors
        _compare_stats(gpu, ora, n, m, M, vsum_tol=2e-6)
    else:
        _compare_stats(gpu, ora, n, m, M, tol=1e-3, ll_tol=1e-5, skip=("mu_denom",), vsum_tol=1e-4)


def test_split_run_and_ltall_follow_oracle(built):
    """8 models of 96 channels (split E-step), FP32 path: 5 EM iterations
    against the FP64 oracle (first iteration within 1e-7, later ones within
    max(1e-6, 4x the oracle's spread under float32-sized perturbations)),
    then the per-sample model log-likelihoods of one more E-step (ltall)."""
    n, T, M, m, iters = 96, 16000, 8, 3, 5
    _say(f"[run split] n={n} T={T} M={M} m={m}, {iters} iterations")
    x = ca.synth_data(ca.synth_spec(n, T, seed=960, nmix=M, block=T // (2 * M), alternate=1,
                                    cond_max=100.0))
    st0 = ca.synth_init(n, M, m, 961)
    xw, _, _, ldetS = po.preprocess(x, 1)
    xin = xw.astype(np.float32).astype(np.float64)
    with ca.Engine(ca.options(n, T, M, m, max_iter=iters + 1)) as eng:
        eng.load_white(xin, ldetS)
        eng.set_state(st0)
        eng.iterate(iters)
        eng.sync()
        glog = eng.log(iters)
        gst = eng.get_state()
        eng.record_ltall(True)
        eng.iterate(1)
        lt = eng.ltall()
    cfg = po.default_config(n, T, M, m, maxiter=iters)
    fin, lq = po.run(cfg, xin, ldetS, st0)
    env = _envelope(cfg, xin, ldetS, st0, lq["ll"], 6e-8)
    rel = np.abs(glog["ll"] - lq["ll"]) / np.abs(lq["ll"])
    fmt = dict(formatter={"float_kind": lambda v: f"{v:.2e}"})
    _say(f"  LL rel. error per iteration: {np.array2string(rel, **fmt)}; "
         f"envelope {np.array2string(env, **fmt)}")
    assert rel[0] <= 1e-7
    assert np.all(rel <= np.maximum(1e-6, 4.0 * env))
    # ltall of one E-step from the GPU's final state against the oracle's
    cfg1 = po.default_config(n, T, M, m, maxiter=1)
    st = {k: gst[k] for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm")}
    _, l1 = po.run(cfg1, xin, ldetS, st, want_ltall=True)
    err = np.max(np.abs(lt - l1["ltall"]))
    _say(f"  ltall max abs error {err:.2e} (|L| up to {np.abs(l1['ltall']).max():.1f})")
    assert err <= 2e-6 * np.abs(l1["ltall"]).max()
