"""FP8 e4m3 storage of the whitened data with per-channel scales (precision
fp8) on the GPU.

* The stored data are on the e4m3 grid: x^ / s_i is an OCP e4m3 value
  (s_i = max_t |x_ti| / 448), within half an e4m3 step of x / s_i.
* Given the dequantized data x^ (what the algorithm runs on), the FP8
  path's E-step statistics match the FP64 oracle on x^ to the split-BF16
  bars (1e-3, LL sum 1e-5; mu_denom reported, not bounded).
* A short run follows the FP64 oracle on x^ (first LL within 1e-6, later
  ones within 4x the oracle's own spread under input perturbations the
  size of the path's rounding); its distance to the oracle on the
  unquantized data is reported.
"""
import numpy as np
import pytest

import cudamica_amd as ca
from oracle import pyoracle as po
from tests.test_gpu_engine import _compare_stats, _random_state

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _e4m3_values():
    """All finite OCP e4m3 (E4M3FN) values: bias 7, 3 mantissa bits, max 448,
    subnormals m 2^-9."""
    vals = set()
    for e in range(16):
        for mnt in range(8):
            if e == 15 and mnt == 7:
                continue                                   # NaN
            v = mnt * 2.0 ** -9 if e == 0 else (1 + mnt / 8) * 2.0 ** (e - 7)
            vals.add(v)
            vals.add(-v)
    return np.array(sorted(vals))


def test_fp8_storage_is_e4m3_with_channel_scales(built):
    n, T = 64, 20000
    _say(f"[fp8 storage] n={n} T={T}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=800, cond_max=50.0))
    xw, _, _, _ = po.preprocess(x, 1)
    with ca.Engine(ca.options(n, T, 1, 3, precision=ca.PREC_FP8)) as eng:
        eng.load_raw(x)
        eng.preprocess(1)
        xh = eng.white().astype(np.float64)
    s = np.abs(xh).max(axis=0) / 448.0
    q = xh / s
    grid = _e4m3_values()
    idx = np.clip(np.searchsorted(grid, q), 1, len(grid) - 1)
    near = np.where(np.abs(grid[idx] - q) < np.abs(grid[idx - 1] - q), grid[idx], grid[idx - 1])
    assert np.max(np.abs(near - q) / np.maximum(np.abs(q), 2.0 ** -9)) < 1e-6, "values off the e4m3 grid"
    # round to nearest: within half an e4m3 step of the float32 whitened data
    ref = xw.astype(np.float32).astype(np.float64) / s
    step = np.maximum(np.abs(ref), 2.0 ** -6) * 2.0 ** -3
    err = np.abs(q - ref) / step
    _say(f"  max |x^ - x| in e4m3 steps: {err.max():.3f}; rms relative error "
         f"{np.sqrt(np.mean((xh - xw) ** 2) / np.mean(xw ** 2)):.2e}")
    assert err.max() <= 0.5 + 1e-3


def test_fp8_estep_stats_match_oracle_on_dequantized_data(built):
    n, T, M, m = 128, 20000, 2, 4
    _say(f"[estep fp8] n={n} T={T} M={M} m={m}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=801, cond_max=0.0))
    st = _random_state(n, M, m, seed=802)
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


def test_fp8_run_follows_oracle(built):
    """The FP8 path rounds W diag(s) to a pair of BF16 values, so b = W x^
    carries relative errors of ~2^-17.  This trajectory is ill-conditioned
    (every run of it, the oracle's included, moves by ~1e-4 in LL within 8
    iterations under input perturbations of that size), so the bar per
    iteration is max(1e-6, 4x the oracle's spread over two runs on
    x^ (1 + 2^-17 N(0, 1)))."""
    w = ca.WORKLOADS["n256_T4M_M2"]
    n, M, m, T, iters = 64, 2, 4, 60000, 8
    _say(f"[run fp8] n={n} T={T} M={M} m={m}, {iters} iterations")
    x = ca.synth_data(ca.synth_spec(n, T, seed=803, nmix=2, block=10000, alternate=1, cond_max=0.0))
    st0 = ca.synth_init(n, M, m, 804)
    with ca.Engine(ca.options(n, T, M, m, precision=ca.PREC_FP8, max_iter=iters)) as eng:
        eng.load_raw(x)
        eng.preprocess(1)
        xh = eng.white().astype(np.float64)
        _, _, ldetS = eng.sphere()
        eng.set_state(st0)
        eng.iterate(iters)
        eng.sync()
        ll = eng.log(iters)["ll"]
    cfg = po.default_config(n, T, M, m, maxiter=iters, newt_start_iter=w.newt_start_iter)
    _, lq = po.run(cfg, xh, ldetS, st0)
    env = np.zeros(iters)
    for s in (1, 2):
        xp = xh * (1.0 + 2.0 ** -17 * np.random.default_rng(s).standard_normal(xh.shape))
        _, lp = po.run(cfg, xp, ldetS, st0)
        env = np.maximum(env, np.abs(lp["ll"] - lq["ll"]) / np.abs(lq["ll"]))
    env = np.maximum.accumulate(env)
    xw, _, _, ldetS0 = po.preprocess(x, 1)
    _, l0 = po.run(cfg, xw, ldetS0, st0)
    rel_q = np.abs(ll - lq["ll"]) / np.abs(lq["ll"])
    rel_0 = np.abs(ll - l0["ll"]) / np.abs(l0["ll"])
    fmt = dict(formatter={"float_kind": lambda v: f"{v:.2e}"})
    _say(f"  LL rel. error vs oracle on the dequantized data: {np.array2string(rel_q, **fmt)}")
    _say(f"  oracle spread under 2^-17 perturbation: {np.array2string(env, **fmt)}")
    _say(f"  LL rel. error vs oracle on the unquantized data: {np.array2string(rel_0, **fmt)}")
    assert rel_q[0] <= 1e-6
    assert np.all(rel_q <= np.maximum(1e-6, 4 * env))
