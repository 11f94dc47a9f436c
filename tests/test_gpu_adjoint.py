"""The per-sample log-likelihood and its adjoint (include/amica_adjoint.h) on
the GPU, through a thin torch.autograd.Function over the C ABI:

* forward: sum_t ll_t equals the FP64 oracle's E-step log-likelihood sum,
  for shapes within one workgroup and for the three-pass split (up to 16
  models, 1024 channels);
* adjoint: torch.autograd.gradcheck in FP64 against finite differences
  (16 channels x 5000 samples, M = 2, m = 4);
* adjoint with dll = 1: the reference's own statistics (FP64 oracle) --
  dW W^T = vsum I - Phi, dalpha = usum/alpha, dmu = sqrt(beta) mu_numer,
  dbeta = (usum - beta_denom)/(2 beta), drho = usum psi(1+1/rho)/rho^2 -
  rho_numer/rho, dgm = vsum/gm;
* the FP32 variant against the FP64 one (relative Frobenius error 1e-4).
"""
import numpy as np
import pytest
import scipy.special as sp

import cudamica_amd as ca
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _say(msg):
    print(msg, flush=True)


class AmicaLL(torch.autograd.Function):
    """ll_t = amica_ll_forward(x; W, c, alpha, mu, beta, rho, gm); backward
    calls amica_ll_adjoint.  x is [T][n] float32 or float64 on the GPU, the
    parameters FP64 in the engine's layouts."""

    @staticmethod
    def forward(ctx, x, W, c, alpha, mu, beta, rho, gm, ldetS, n, M, m):
        T = x.shape[0]
        bits = 64 if x.dtype == torch.float64 else 32
        ps = [p.detach().contiguous() for p in (W, c, alpha, mu, beta, rho, gm)]
        ll = torch.empty(T, dtype=torch.float64, device=x.device)
        stream = torch.cuda.current_stream().cuda_stream
        ca.ll_forward(bits, x.data_ptr(), T, n, M, m,
                      {k: p.data_ptr() for k, p in zip(ca.PARAM_NAMES, ps)}, ldetS,
                      ll.data_ptr(), stream)
        ctx.save_for_backward(x, *ps)
        ctx.meta = (ldetS, n, M, m, bits)
        return ll

    @staticmethod
    def backward(ctx, dll):
        x, *ps = ctx.saved_tensors
        ldetS, n, M, m, bits = ctx.meta
        dll = dll.detach().to(torch.float64).contiguous()
        grads = [torch.empty_like(p) for p in ps]
        stream = torch.cuda.current_stream().cuda_stream
        ca.ll_adjoint(bits, x.data_ptr(), x.shape[0], n, M, m,
                      {k: p.data_ptr() for k, p in zip(ca.PARAM_NAMES, ps)}, ldetS,
                      dll.data_ptr(), {k: g.data_ptr() for k, g in zip(ca.PARAM_NAMES, grads)},
                      0, stream)
        return (None, *grads, None, None, None, None)


def _problem(n, T, M, m, seed):
    """Whitened synthetic data and a state away from the rho bounds.  Beyond
    64 channels the mixing matrix is not drawn to a condition bound (the
    rejection draw would dominate the test's time); beyond 128 the data are
    only centred and scaled per channel (ldetS = 0), since the likelihood
    takes any x and the oracle's sphering is slow there."""
    x = ca.synth_data(ca.synth_spec(n, T, seed=seed, nmix=M, block=max(1, T // (2 * M)),
                                    alternate=1, rho=(1.0, 2.0),
                                    cond_max=50.0 if n <= 64 else 0.0))
    if n <= 128:
        xw, _, _, ldetS = po.preprocess(x, 1)
    else:
        x = x.astype(np.float64)
        xw, ldetS = (x - x.mean(axis=0)) / x.std(axis=0), 0.0
    st = ca.synth_init(n, M, m, seed + 1)
    rng = np.random.default_rng(seed)
    st["rho"] = rng.uniform(1.2, 1.9, size=st["rho"].shape)
    st["beta"] = st["beta"] * rng.uniform(0.7, 1.4, size=st["beta"].shape)
    st["mu"] = st["mu"] + rng.normal(0.0, 0.1, size=st["mu"].shape)
    a = rng.uniform(0.5, 1.5, size=(M, m, n))
    st["alpha"] = (a / a.sum(axis=1, keepdims=True)).ravel()
    g = rng.uniform(0.5, 1.5, size=M)
    st["gm"] = g / g.sum()
    W = np.zeros(M * n * n)
    for h in range(M):
        Ah = st["A"][h * n * n:(h + 1) * n * n].reshape(n, n).T
        W[h * n * n:(h + 1) * n * n] = np.linalg.inv(Ah).T.ravel()
    return xw, ldetS, st, W


def _tensors(st, W, dev):
    vals = {"W": W, "c": st["c"], "alpha": st["alpha"], "mu": st["mu"], "beta": st["beta"],
            "rho": st["rho"], "gm": st["gm"]}
    return [torch.tensor(np.asarray(vals[k], np.float64), device=dev) for k in ca.PARAM_NAMES]


# (n, T, M, m): all (model, 64-channel block) pairs in one workgroup; then
# shapes beyond 8 pairs, which run the three-pass split (10 pairs over two
# groups of 8, the second part-empty; 9 single-block models)
SHAPES = [(16, 5000, 2, 4), (300, 6000, 2, 3), (100, 5000, 5, 2), (40, 4000, 9, 2)]


@pytest.mark.parametrize("n,T,M,m", SHAPES, ids=[f"n{s[0]}M{s[2]}" for s in SHAPES])
def test_forward_matches_oracle_ll(n, T, M, m):
    _say(f"[ll forward] n={n} T={T} M={M} m={m}")
    xw, ldetS, st, W = _problem(n, T, M, m, 7101)
    dev = torch.device("cuda:0")
    ps = _tensors(st, W, dev)
    st_or = po.estep(xw, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                     st["gm"], M, m)
    _, llsum = po.stats_views(st_or, n, m, M)
    x64 = torch.tensor(xw, device=dev)
    ll = AmicaLL.apply(x64, *ps, ldetS, n, M, m)
    got = float(ll.sum())
    _say(f"  sum ll gpu {got:.12f} oracle {llsum:.12f}")
    assert abs(got - llsum) <= 1e-12 * abs(llsum)
    ll32 = AmicaLL.apply(x64.float(), *ps, ldetS, n, M, m)
    rel = float(((ll32 - ll).abs() / ll.abs()).max())
    _say(f"  FP32 variant: max per-sample relative error {rel:.2e}")
    assert rel <= 1e-5


def test_adjoint_gradcheck_fp64():
    n, T, M, m = 16, 5000, 2, 4
    _say(f"[ll adjoint gradcheck] n={n} T={T} M={M} m={m}, FP64")
    xw, ldetS, st, W = _problem(n, T, M, m, 7202)
    dev = torch.device("cuda:0")
    ps = [p.requires_grad_() for p in _tensors(st, W, dev)]
    x64 = torch.tensor(xw, device=dev)
    ok = torch.autograd.gradcheck(lambda *p: AmicaLL.apply(x64, *p, ldetS, n, M, m), ps,
                                  eps=1e-6, atol=1e-5, rtol=1e-4, fast_mode=True)
    assert ok


@pytest.mark.parametrize("n,T,M,m", SHAPES, ids=[f"n{s[0]}M{s[2]}" for s in SHAPES])
def test_adjoint_equals_reference_statistics(n, T, M, m):
    _say(f"[ll adjoint vs oracle statistics] n={n} T={T} M={M} m={m}")
    xw, ldetS, st, W = _problem(n, T, M, m, 7303)
    dev = torch.device("cuda:0")
    ps = [p.requires_grad_() for p in _tensors(st, W, dev)]
    x64 = torch.tensor(xw, device=dev)
    ll = AmicaLL.apply(x64, *ps, ldetS, n, M, m)
    ll.sum().backward()
    g = {k: p.grad.cpu().numpy() for k, p in zip(ca.PARAM_NAMES, ps)}
    sts, _ = po.stats_views(po.estep(xw, ldetS, st["A"], st["c"], st["alpha"], st["mu"],
                                     st["beta"], st["rho"], st["gm"], M, m), n, m, M)

    def rel(a, b):
        return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / np.linalg.norm(np.ravel(b)))

    nm = n * m
    for h in range(M):
        s = sts[h]
        sl = slice(h * nm, (h + 1) * nm)
        al = st["alpha"][sl].reshape(m, n)
        be = st["beta"][sl].reshape(m, n)
        rh = st["rho"][sl].reshape(m, n)
        Wh = W[h * n * n:(h + 1) * n * n].reshape(n, n).T
        dWh = g["W"][h * n * n:(h + 1) * n * n].reshape(n, n).T
        checks = {
            "dW W^T = vsum I - phi": rel(dWh @ Wh.T, s["vsum"] * np.eye(n) - s["phi"]),
            "dalpha = usum/alpha": rel(g["alpha"][sl].reshape(m, n), s["usum"] / al),
            "dmu = sqrt(beta) mu_numer": rel(g["mu"][sl].reshape(m, n), np.sqrt(be) * s["mu_numer"]),
            "dbeta = (usum - beta_denom)/(2 beta)":
                rel(g["beta"][sl].reshape(m, n), (s["usum"] - s["beta_denom"]) / (2 * be)),
            "drho = usum psi/rho^2 - rho_numer/rho":
                rel(g["rho"][sl].reshape(m, n),
                    s["usum"] * sp.digamma(1 + 1 / rh) / rh ** 2 - s["rho_numer"] / rh),
            "dgm = vsum/gm": abs(g["gm"][h] - s["vsum"] / st["gm"][h]) / (s["vsum"] / st["gm"][h]),
        }
        for name, err in checks.items():
            _say(f"  model {h}: {name}: {err:.2e}")
            assert err <= 1e-10, name


@pytest.mark.parametrize("n,T,M,m", [(64, 20000, 1, 3), (300, 6000, 2, 3), (100, 5000, 5, 2)],
                         ids=["n64M1", "n300M2", "n100M5"])
def test_adjoint_fp32_matches_fp64(n, T, M, m):
    """The FP32 variant (FP32 MFMA GEMMs, FP32 density math) against the FP64
    one, within one workgroup and through the three-pass split (the last
    two shapes)."""
    _say(f"[ll adjoint fp32 vs fp64] n={n} T={T} M={M} m={m}")
    xw, ldetS, st, W = _problem(n, T, M, m, 7404)
    dev = torch.device("cuda:0")
    out = {}
    for dt in (torch.float64, torch.float32):
        ps = [p.requires_grad_() for p in _tensors(st, W, dev)]
        x = torch.tensor(xw, device=dev, dtype=dt)
        AmicaLL.apply(x, *ps, ldetS, n, M, m).sum().backward()
        out[dt] = [p.grad.cpu().numpy() for p in ps]
    for k, a, b in zip(ca.PARAM_NAMES, out[torch.float32], out[torch.float64]):
        err = float(np.linalg.norm(a - b) / np.linalg.norm(b))
        _say(f"  d{k}: relative error {err:.2e}")
        assert err <= 1e-4, k
