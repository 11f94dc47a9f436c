"""ctypes binding of the C++ CPU oracle (TEST INFRASTRUCTURE ONLY).

Loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
The product package (cudamica_amd) never imports this module.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OrcConfig(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("T", ctypes.c_long), ("M", ctypes.c_int), ("m", ctypes.c_int),
        ("maxiter", ctypes.c_int), ("do_newton", ctypes.c_int), ("update_rho", ctypes.c_int),
        ("rho_start_iter", ctypes.c_int), ("newt_start_iter", ctypes.c_int),
        ("update_gm", ctypes.c_int), ("update_alpha", ctypes.c_int), ("update_A", ctypes.c_int),
        ("update_mu", ctypes.c_int), ("update_beta", ctypes.c_int), ("do_reparm", ctypes.c_int),
        ("lrate0", ctypes.c_double), ("lratemax", ctypes.c_double),
        ("lnatrate", ctypes.c_double), ("lratefact", ctypes.c_double),
        ("rholrate", ctypes.c_double), ("rhomin", ctypes.c_double), ("rhomax", ctypes.c_double),
        ("maxdec", ctypes.c_int),
    ]


class MStepTrace(ctypes.Structure):
    _fields_ = [(k, ctypes.POINTER(ctypes.c_double))
                for k in ("kappa", "lambda_", "sigma2", "cnew", "phi", "B")]


_STATE_KEYS = ("A", "c", "alpha", "mu", "beta", "rho", "gm")

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        fp = ctypes.POINTER(ctypes.c_float)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_default_config.argtypes = [ctypes.POINTER(OrcConfig)]
        L.orc_stats_stride.argtypes = [ctypes.c_int, ctypes.c_int]
        L.orc_stats_stride.restype = ctypes.c_long
        L.orc_preprocess.argtypes = [ctypes.c_int, ctypes.c_long, fp, ctypes.c_int, dp, dp, dp, dp]
        est = [ctypes.c_int, ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp,
               ctypes.c_double, dp, dp, dp, dp, dp, dp, dp, dp, ctypes.c_int]
        L.orc_estep.argtypes = est
        L.orc_estep_f32.argtypes = est
        L.orc_run.argtypes = [ctypes.POINTER(OrcConfig), dp, ctypes.c_double, dp, dp, dp, dp, dp,
                              dp, dp, dp, dp, dp, ip, ip, ip, dp, ctypes.c_int, ctypes.c_int]
        L.orc_mstep.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_int, dp, dp, dp, dp, dp, dp, dp, dp,
                                dp, dp, dp, ip, ip, ip, ctypes.POINTER(MStepTrace)]
        L.orc_step.argtypes = [ctypes.POINTER(OrcConfig), ctypes.c_int, dp, ctypes.c_double, dp, dp, dp,
                               dp, dp, dp, dp, dp, ctypes.c_double, dp, dp, ip, ip, ip, ctypes.c_int,
                               ctypes.c_int]
        L.orc_plant.argtypes = [ctypes.c_int]
        L.orc_pinv.argtypes = [ctypes.c_int, dp, dp]
        L.orc_logabsdet.argtypes = [ctypes.c_int, dp]
        L.orc_logabsdet.restype = ctypes.c_double
        L.orc_eigh.argtypes = [ctypes.c_int, dp, dp, dp]
        L.orc_psi.argtypes = [ctypes.c_double]
        L.orc_psi.restype = ctypes.c_double
        L.orc_output_transform.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _f(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _i(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def default_config(n, T, M=1, m=3, maxiter=200, **kw):
    cfg = OrcConfig()
    lib().orc_default_config(ctypes.byref(cfg))
    cfg.n, cfg.T, cfg.M, cfg.m, cfg.maxiter = n, T, M, m, maxiter
    for k, v in kw.items():
        setattr(cfg, k, v)
    return cfg


def stats_stride(n, m):
    return lib().orc_stats_stride(n, m)


def preprocess(x_raw, do_sphere=1):
    """x_raw: float32 (T, n). Returns x_white (T, n) float64, S (n, n) colmajor-as-array, mean, ldetS."""
    x_raw = np.ascontiguousarray(x_raw, dtype=np.float32)
    T, n = x_raw.shape
    xw = np.empty((T, n), np.float64)
    S = np.empty(n * n, np.float64)
    mean = np.empty(n, np.float64)
    ldetS = np.zeros(1, np.float64)
    lib().orc_preprocess(n, T, _f(x_raw), do_sphere, _d(xw), _d(S), _d(mean), _d(ldetS))
    return xw, S, mean, float(ldetS[0])


def estep(x_white, ldetS, A, c, alpha, mu, beta, rho, gm, M, m, update_rho=1, f32=False, nthreads=0):
    T, n = x_white.shape
    st = np.zeros(M * stats_stride(n, m) + 1, np.float64)
    fn = lib().orc_estep_f32 if f32 else lib().orc_estep
    args = [np.ascontiguousarray(a, dtype=np.float64) for a in (A, c, alpha, mu, beta, rho, gm)]
    fn(n, T, M, m, update_rho, _d(np.ascontiguousarray(x_white)), ldetS, *[_d(a) for a in args], _d(st),
       nthreads)
    return st


def run(cfg, x_white, ldetS, state, precision=64, nthreads=0, want_ltall=False):
    """Runs cfg.maxiter EM iterations; state is a dict of float64 arrays (A, c, alpha, mu, beta,
    rho, gm) that is copied, not modified.  Returns (final state, logs)."""
    st = {k: np.array(v, dtype=np.float64, copy=True).ravel() for k, v in state.items()}
    n, M, K = cfg.n, cfg.M, cfg.maxiter
    W = np.zeros(n * n * M, np.float64)
    ll = np.zeros(K, np.float64)
    lr = np.zeros(K, np.float64)
    bf = np.zeros(K * M, np.int32)
    nw = np.zeros(K * M, np.int32)
    dec = np.zeros(K, np.int32)
    lt = np.zeros(cfg.T * M, np.float64) if want_ltall else None
    ltp = _d(lt) if want_ltall else ctypes.POINTER(ctypes.c_double)()
    rc = lib().orc_run(ctypes.byref(cfg), _d(np.ascontiguousarray(x_white)), ldetS,
                       _d(st["A"]), _d(st["c"]), _d(st["alpha"]), _d(st["mu"]), _d(st["beta"]),
                       _d(st["rho"]), _d(st["gm"]), _d(W), _d(ll), _d(lr), _i(bf), _i(nw), _i(dec),
                       ltp, precision, nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_run failed: {rc}")
    st["W"] = W
    logs = dict(ll=ll, lrate=lr, bflag=bf.reshape(K, M), newton=nw.reshape(K, M), lldec=dec)
    if want_ltall:
        logs["ltall"] = lt.reshape(M, cfg.T)
    return st, logs


def mstep(cfg, it, stats, W, state, lrate, ll_prev=0.0, ll_cur=None, trace=False):
    """One M-step (orc_mstep) at 1-based iteration `it` from canonical
    statistics.  ll_cur defaults to the statistics' own log-likelihood.
    Returns (new state, new lrate, decisions dict, trace dict or None)."""
    n, M = cfg.n, cfg.M
    st = {k: np.array(state[k], dtype=np.float64, copy=True).ravel() for k in _STATE_KEYS}
    stats = np.ascontiguousarray(stats, dtype=np.float64)
    if ll_cur is None:
        ll_cur = stats[-1] / (n * cfg.T)
    hist = np.array([ll_prev, ll_cur], np.float64)
    lr = np.array([lrate], np.float64)
    bf = np.zeros(M, np.int32)
    nw = np.zeros(M, np.int32)
    dec = np.zeros(1, np.int32)
    tr = None
    trp = ctypes.POINTER(MStepTrace)()
    if trace:
        tr = dict(kappa=np.zeros(M * n), lambda_=np.zeros(M * n), sigma2=np.zeros(M * n),
                  cnew=np.zeros(M * n), phi=np.zeros(M * n * n), B=np.zeros(M * n * n))
        t = MStepTrace(*[_d(tr[k]) for k in ("kappa", "lambda_", "sigma2", "cnew", "phi", "B")])
        trp = ctypes.pointer(t)
    Wc = np.ascontiguousarray(W, dtype=np.float64).ravel()
    rc = lib().orc_mstep(ctypes.byref(cfg), it, _d(stats), _d(Wc), _d(hist),
                         *[_d(st[k]) for k in _STATE_KEYS], _d(lr), _i(bf), _i(nw), _i(dec), trp)
    if rc != 0:
        raise RuntimeError(f"orc_mstep failed: {rc}")
    return st, float(lr[0]), dict(bflag=bf, newton=nw, lldec=int(dec[0])), tr


def step(cfg, it, x_white, ldetS, state, lrate, ll_prev=0.0, precision=64, nthreads=0):
    """One EM iteration (orc_step) at 1-based iteration `it` from `state`
    (not modified).  Returns (new state with W of this iteration, LL of this
    iteration, new lrate, decisions dict)."""
    n, M = cfg.n, cfg.M
    st = {k: np.array(state[k], dtype=np.float64, copy=True).ravel() for k in _STATE_KEYS}
    lr = np.array([lrate], np.float64)
    W = np.zeros(n * n * M)
    ll = np.zeros(1)
    bf = np.zeros(M, np.int32)
    nw = np.zeros(M, np.int32)
    dec = np.zeros(1, np.int32)
    rc = lib().orc_step(ctypes.byref(cfg), it, _d(np.ascontiguousarray(x_white)), ldetS,
                        *[_d(st[k]) for k in _STATE_KEYS], _d(lr), float(ll_prev), _d(W), _d(ll),
                        _i(bf), _i(nw), _i(dec), precision, nthreads)
    if rc != 0:
        raise RuntimeError(f"orc_step failed: {rc}")
    st["W"] = W
    return st, float(ll[0]), float(lr[0]), dict(bflag=bf, newton=nw, lldec=int(dec[0]))


# planted-mistake codes of orc_plant (oracle.h), test-only
PLANTS = {
    "kappa_numer_no_beta": 1, "lambda_numer_plus": 2, "mu_denom_no_sqrt_beta": 3,
    "rho_numer_no_log": 4, "sigma2_from_x": 5, "cnew_no_v": 6, "g_no_sqrt_beta": 7,
    "kappa_old_alpha": 8, "lambda_no_mu": 9, "beta_sqrt_ratio": 10, "mu_step_sign": 11,
    "rho_step_sign": 12, "rho_psi_arg": 13, "newton_transpose": 14, "ab_order": 15,
    "natgrad_sign": 16, "reparam_beta_nu": 17, "mu_shift_wt": 18, "lrate_growth": 19,
    "lldec_no_halve": 20, "bflag_no_diag": 21, "alpha_over_t": 22, "phi_no_norm": 23,
}


def plant(name):
    """Plants the named mistake in the oracle (None removes it).  Test-only."""
    lib().orc_plant(0 if name is None else PLANTS[name])


def pinv(A):
    n = A.shape[0]
    a = np.ascontiguousarray(np.asarray(A, np.float64).T).ravel()  # column-major
    W = np.zeros(n * n)
    lib().orc_pinv(n, _d(a), _d(W))
    return W.reshape(n, n).T


def logabsdet(A):
    n = A.shape[0]
    a = np.ascontiguousarray(np.asarray(A, np.float64).T).ravel()
    return lib().orc_logabsdet(n, _d(a))


def eigh(C):
    n = C.shape[0]
    a = np.ascontiguousarray(np.asarray(C, np.float64).T).ravel()
    w = np.zeros(n)
    V = np.zeros(n * n)
    lib().orc_eigh(n, _d(a), _d(w), _d(V))
    return w, V.reshape(n, n).T


def psi(x):
    return lib().orc_psi(float(x))


def output_transform(S, mean, A, c, n, M):
    A = np.array(A, np.float64, copy=True).ravel()
    c = np.array(c, np.float64, copy=True).ravel()
    lib().orc_output_transform(n, M, _d(np.ascontiguousarray(S, np.float64).ravel()),
                               _d(np.ascontiguousarray(mean, np.float64)), _d(A), _d(c))
    return A, c


def stats_views(st, n, m, M):
    """Splits a canonical statistics buffer into named per-model arrays."""
    stride = stats_stride(n, m)
    nm = n * m
    out = []
    for h in range(M):
        a = st[h * stride:(h + 1) * stride]
        d = {"vsum": a[0]}
        off = 1
        for name in ("usum", "kappa_numer", "lambda_numer", "mu_numer", "mu_denom", "beta_numer",
                     "beta_denom", "rho_numer"):
            d[name] = a[off:off + nm].reshape(m, n)
            off += nm
        d["sigma2_numer"] = a[off:off + n]; off += n
        d["cnew"] = a[off:off + n]; off += n
        d["phi"] = a[off:off + n * n].reshape(n, n).T  # row i, col k
        out.append(d)
    return out, st[M * stride]
