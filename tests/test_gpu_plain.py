"""The plain FP32 path (AMICA_PREC_FP32_PLAIN) on the GPU against the FP64
oracle: both GEMMs on FP32 operands (v_mfma_f32_16x16x4f32), b and g in
FP32, FP32 density math.  It is the path the mixed-precision ones are
compared with in speed, so it is held to the oracle like them:

* one E-step's statistics for the register-resident unmixing kernel at
  NP = 64, 128, 192 and 256 and the generic one beyond (NP = 320), with
  padding channels, partial tiles, one to four models and m = 2 .. 4;
* 50 iterations of the 16-channel workload, branch-aware against the oracle's
  own spread under float32-sized input perturbations;
* 8 iterations of the 128-channel, 4-model workload on 100k samples.
"""
import numpy as np
import pytest

import cudamica_amd as ca
from oracle import pyoracle as po
from tests.test_gpu_engine import (_check_trajectory, _compare_stats, _oracle_envelope, _random_state,
                                   _run_both)

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


SHAPES = [(128, 20000, 4, 3), (71, 20011, 2, 3), (40, 9000, 3, 3), (150, 8000, 2, 2),
          (200, 9001, 2, 4), (300, 6000, 1, 3)]


@pytest.mark.parametrize("n,T,M,m", SHAPES, ids=[f"n{s[0]}M{s[2]}m{s[3]}" for s in SHAPES])
def test_estep_stats_fp32plain(built, n, T, M, m):
    """FP32-MFMA unmixing rounds b to ~1e-7 relative: statistics within 1e-4,
    the LL sum within 1e-6 (mu_denom reported, not bounded: its weight
    |y|^(rho-2) is singular at y = 0).  Beyond 128 channels the data are
    centred and scaled per channel (ldetS = 0) instead of sphered."""
    _say(f"[estep fp32plain] n={n} T={T} M={M} m={m}")
    x = ca.synth_data(ca.synth_spec(n, T, seed=500 + n + M, cond_max=0.0))
    if n <= 128:
        xw, _, _, ldetS = po.preprocess(x, 1)
    else:
        xd = x.astype(np.float64)
        xw, ldetS = (xd - xd.mean(axis=0)) / xd.std(axis=0), 0.0
    x32 = xw.astype(np.float32).astype(np.float64)
    st = _random_state(n, M, m, seed=501 + n)
    ora = po.estep(x32, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                   st["gm"], M, m)
    with ca.Engine(ca.options(n, T, M, m, precision=ca.PREC_FP32_PLAIN)) as eng:
        eng.load_white(x32, ldetS)
        eng.set_state(st)
        gpu = eng.estep_stats()
    _compare_stats(gpu, ora, n, m, M, tol=1e-4, ll_tol=1e-6, skip=("mu_denom",), vsum_tol=1e-5)


def test_full_run_fp32plain_n16(built):
    w = ca.WORKLOADS["n16_T20k"]
    _say("[run fp32plain] n16_T20k 50 iterations")
    fin, logs, gst, glog, margin = _run_both(w, w.T, 50, ca.PREC_FP32_PLAIN)
    env, wenv = _oracle_envelope(w, w.T, 50, w.newt_start_iter, fin, logs)
    _check_trajectory(fin, logs, gst, glog, w.n, w.M, 1e-6, 1e-4, env=env, wenv=wenv, margin=margin)


def test_full_run_fp32plain_models(built):
    """The headline's shape on a 100k-sample subsample: the first iteration
    (the E-step alone) within 1e-6, then max(1e-6, 4x the oracle's spread
    under 2^-20 relative perturbations of its input -- the size of b's
    rounding, a 128-term FP32 dot product, sqrt(128) 2^-24 ~ 2^-20.5;
    near y = 0 the mu_denom weight |y|^(rho-2) amplifies it, as on the BF16
    path)."""
    w = ca.WORKLOADS["n128_T2M_M4"]
    T = 100000
    _say(f"[run fp32plain] n128 M4 subsample T={T}, 8 iterations")
    fin, logs, gst, glog, _ = _run_both(w, T, 8, ca.PREC_FP32_PLAIN)
    rel = np.abs(glog["ll"] - logs["ll"]) / np.abs(logs["ll"])
    env, _ = _oracle_envelope(w, T, 8, w.newt_start_iter, fin, logs, eps=2.0 ** -20)
    fmt = dict(formatter={"float_kind": lambda v: f"{v:.2e}"})
    _say(f"  fp32plain LL rel err per iteration: {np.array2string(rel, **fmt)}")
    _say(f"  oracle spread under 2^-20 input perturbation: {np.array2string(env, **fmt)}")
    assert rel[0] <= 1e-6
    assert np.all(rel <= np.maximum(1e-6, 4.0 * env))
