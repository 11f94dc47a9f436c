"""amica_ll_forward_noalloc (the allocation-free variant of amica_ll_forward,
include/amica_adjoint.h) against amica_ll_forward through the C ABI, on two
calls in a row (the second on a smaller problem, after the first left its
workspace filled), in FP64 and FP32 density math.  The allocation-free
variant forms b = W x - W c per lane (it has no room for a [T][M NP]
buffer) where amica_ll_forward reads it from the MFMA unmixing GEMM, so the
per-sample log-likelihoods agree to the rounding of b: 1e-12 relative in
FP64, 2e-6 in FP32.  The first call is also held to the FP64 oracle's
log-likelihood sum, and models beyond the workspace are rejected before any
GPU work."""
import numpy as np
import pytest
import torch

import cudamica_amd as ca
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _problem(n, T, M, m, seed):
    x = ca.synth_data(ca.synth_spec(n, T, seed=seed, nmix=M, block=max(1, T // (2 * M)),
                                    alternate=1, rho=(1.0, 2.0), cond_max=50.0))
    xw, _, _, ldetS = po.preprocess(x, 1)
    st = ca.synth_init(n, M, m, seed + 1)
    rng = np.random.default_rng(seed)
    st["rho"] = rng.uniform(1.2, 1.9, size=st["rho"].shape)
    st["beta"] = st["beta"] * rng.uniform(0.7, 1.4, size=st["beta"].shape)
    W = np.zeros(M * n * n)
    for h in range(M):
        Ah = st["A"][h * n * n:(h + 1) * n * n].reshape(n, n).T
        W[h * n * n:(h + 1) * n * n] = np.linalg.inv(Ah).T.ravel()
    vals = {"W": W, "c": st["c"], "alpha": st["alpha"], "mu": st["mu"], "beta": st["beta"],
            "rho": st["rho"], "gm": st["gm"]}
    return xw, ldetS, st, vals


def _both(xw, ldetS, vals, n, M, m, bits):
    dev = torch.device("cuda:0")
    ps = {k: torch.tensor(np.asarray(vals[k], np.float64), device=dev) for k in ca.PARAM_NAMES}
    addr = {k: p.data_ptr() for k, p in ps.items()}
    x = torch.tensor(xw, device=dev, dtype=torch.float64 if bits == 64 else torch.float32)
    T = x.shape[0]
    ll_a = torch.full((T,), np.nan, dtype=torch.float64, device=dev)
    ll_b = torch.full((T,), np.nan, dtype=torch.float64, device=dev)
    stream = torch.cuda.current_stream().cuda_stream
    # the allocation-free call first, so that in a fresh process it is the
    # library's first call on the device
    ca.ll_forward_noalloc(bits, x.data_ptr(), T, n, M, m, addr, ldetS, ll_a.data_ptr(), stream)
    torch.cuda.current_stream().synchronize()
    ca.ll_forward(bits, x.data_ptr(), T, n, M, m, addr, ldetS, ll_b.data_ptr(), stream)
    torch.cuda.current_stream().synchronize()
    return ll_a.cpu().numpy(), ll_b.cpu().numpy()


@pytest.mark.parametrize("bits", [64, 32])
def test_noalloc_forward_matches_allocating_forward(bits):
    first = (96, 9000, 2, 3)
    second = (40, 3001, 1, 4)
    for k, (n, T, M, m) in enumerate((first, second)):
        _say(f"[ll forward noalloc] call {k + 1}: n={n} T={T} M={M} m={m}, FP{bits}")
        xw, ldetS, st, vals = _problem(n, T, M, m, 8100 + 17 * k + bits)
        a, b = _both(xw, ldetS, vals, n, M, m, bits)
        assert np.all(np.isfinite(a))
        rel = float(np.max(np.abs(a - b) / np.abs(b)))
        _say(f"  max per-sample relative difference {rel:.2e}")
        assert rel <= (1e-12 if bits == 64 else 2e-6)
        if k == 0 and bits == 64:
            ora = po.estep(xw, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"],
                           st["rho"], st["gm"], M, m)
            _, llsum = po.stats_views(ora, n, m, M)
            _say(f"  sum ll {a.sum():.12f}, oracle {llsum:.12f}")
            assert abs(a.sum() - llsum) <= 1e-12 * abs(llsum)


def test_noalloc_rejects_models_beyond_its_workspace():
    n, M, m = 300, 1, 3          # M n n = 90000 > 65536
    with pytest.raises(ca.EngineError, match="allocation-free workspace"):
        ca.ll_forward_noalloc(64, 1, 10, n, M, m, {k: 1 for k in ca.PARAM_NAMES}, 0.0, 1, 0)
