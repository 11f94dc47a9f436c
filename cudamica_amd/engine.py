"""ctypes binding of libcudamica_amd (include/amica_engine.h, include/amica_synth.h).

The engine runs on the GPU only: loading fails loudly when the library is
missing, and every engine call raises on a HIP error.  There is no CPU
fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcudamica_amd.so")
CLI_PATH = os.path.join(_HERE, "bin", "cudaamica")

PREC_FP32 = 32
PREC_FP32_PLAIN = 33      # FP32 MFMA GEMMs, FP32 b and g: the plain FP32 comparator
PREC_BF16 = 16
PREC_FP8 = 8


class Options(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("T", ctypes.c_long), ("M", ctypes.c_int), ("m", ctypes.c_int),
        ("precision", ctypes.c_int), ("max_iter", ctypes.c_int), ("device", ctypes.c_int),
        ("do_newton", ctypes.c_int), ("update_rho", ctypes.c_int),
        ("rho_start_iter", ctypes.c_int), ("newt_start_iter", ctypes.c_int),
        ("update_gm", ctypes.c_int), ("update_alpha", ctypes.c_int), ("update_A", ctypes.c_int),
        ("update_mu", ctypes.c_int), ("update_beta", ctypes.c_int), ("do_reparm", ctypes.c_int),
        ("lrate0", ctypes.c_double), ("lratemax", ctypes.c_double),
        ("lnatrate", ctypes.c_double), ("lratefact", ctypes.c_double),
        ("rholrate", ctypes.c_double), ("rhomin", ctypes.c_double), ("rhomax", ctypes.c_double),
        ("maxdec", ctypes.c_int),
    ]


class SynthSpec(ctypes.Structure):
    _fields_ = [
        ("n", ctypes.c_int), ("T", ctypes.c_long), ("nmix", ctypes.c_int), ("block", ctypes.c_long),
        ("alternate", ctypes.c_int), ("rho_lo", ctypes.c_double), ("rho_hi", ctypes.c_double),
        ("sub_frac", ctypes.c_double), ("sub_lo", ctypes.c_double), ("sub_hi", ctypes.c_double),
        ("cond_max", ctypes.c_double), ("noise_std", ctypes.c_double),
        ("seed", ctypes.c_ulonglong),
    ]


# (name, restype, argtypes) of every C entry point this binding uses
_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I = ctypes.POINTER(ctypes.c_int)
_V = ctypes.c_void_p
_SIGNATURES = [
    ("amica_options_default", None, [ctypes.POINTER(Options), ctypes.c_int, ctypes.c_long, ctypes.c_int, ctypes.c_int]),
    ("amica_engine_create", ctypes.c_int, [ctypes.POINTER(Options), ctypes.POINTER(_V)]),
    ("amica_engine_destroy", ctypes.c_int, [_V]),
    ("amica_engine_load_raw", ctypes.c_int, [_V, _V, ctypes.c_int]),
    ("amica_engine_center", ctypes.c_int, [_V]),
    ("amica_engine_whiten", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_engine_preprocess", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_engine_get_sphere", ctypes.c_int, [_V, _D, _D, _D]),
    ("amica_engine_load_white", ctypes.c_int, [_V, _D, ctypes.c_double]),
    ("amica_engine_get_white", ctypes.c_int, [_V, _F]),
    ("amica_engine_set_state", ctypes.c_int, [_V, _D, _D, _D, _D, _D, _D, _D]),
    ("amica_engine_get_state", ctypes.c_int, [_V, _D, _D, _D, _D, _D, _D, _D, _D]),
    ("amica_engine_iterate", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_engine_iterate_graph", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_engine_sync", ctypes.c_int, [_V]),
    ("amica_engine_iterations", ctypes.c_int, [_V]),
    ("amica_engine_get_log", ctypes.c_int, [_V, ctypes.c_int, _D, _D, _I, _I, _I, _D]),
    ("amica_engine_estep_stats", ctypes.c_int, [_V, _D]),
    ("amica_engine_record_ltall", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_engine_get_ltall", ctypes.c_int, [_V, _D]),
    ("amica_engine_output", ctypes.c_int, [_V, _D, _D]),
    ("amica_engine_set_timing", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_engine_get_timing", ctypes.c_int, [_V, _D, _I]),
    ("amica_last_error", ctypes.c_char_p, []),
    ("amica_engine_debug_buffer", ctypes.c_int, [_V, ctypes.c_char_p, _V, ctypes.c_size_t]),
    ("amica_engine_cu_count", ctypes.c_int, [_V]),
    ("amica_engine_set_cu_budget", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_hbm_probe", ctypes.c_int, [ctypes.c_int, ctypes.c_size_t, ctypes.c_int, _D, _D]),
    ("amica_synth_data", ctypes.c_int, [ctypes.POINTER(SynthSpec), _F, _D, _D, ctypes.c_int]),
    ("amica_synth_init", ctypes.c_int, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong,
                                        ctypes.c_int, ctypes.c_double, _D, _D, _D, _D, _D, _D, _D]),
    ("amica_synth_cond", ctypes.c_double, [ctypes.c_int, _D]),
    ("amica_batch_create", ctypes.c_int, [ctypes.POINTER(Options), ctypes.c_int, ctypes.POINTER(_V)]),
    ("amica_batch_destroy", ctypes.c_int, [_V]),
    ("amica_batch_size", ctypes.c_int, [_V]),
    ("amica_batch_engine", ctypes.c_int, [_V, ctypes.c_int, ctypes.POINTER(_V)]),
    ("amica_batch_set_budgets", ctypes.c_int, [_V, _I]),
    ("amica_batch_set_policy", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_batch_tune", ctypes.c_int, [_V, ctypes.c_int, _I, _D]),
    ("amica_batch_iterate", ctypes.c_int, [_V, ctypes.c_int]),
    ("amica_batch_sync", ctypes.c_int, [_V]),
    ("amica_ll_forward", ctypes.c_int, [ctypes.c_int, _V, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int] + [_V] * 7 + [ctypes.c_double, _V, _V]),
    ("amica_ll_forward_noalloc", ctypes.c_int, [ctypes.c_int, _V, ctypes.c_long, ctypes.c_int,
                                                ctypes.c_int, ctypes.c_int] + [_V] * 7
     + [ctypes.c_double, _V, _V]),
    ("amica_ll_adjoint", ctypes.c_int, [ctypes.c_int, _V, ctypes.c_long, ctypes.c_int, ctypes.c_int,
                                        ctypes.c_int] + [_V] * 7 + [ctypes.c_double] + [_V] * 10),
]

_lib = None


def load_library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _d(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], "float64 C-contiguous array expected"
    return a.ctypes.data_as(_D)


def _f(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "float32 C-contiguous array expected"
    return a.ctypes.data_as(_F)


def _i(a):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_I)


class EngineError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        msg = load_library().amica_last_error()
        raise EngineError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def options(n, T, M=1, m=3, precision=PREC_FP32, max_iter=200, **kw):
    o = Options()
    load_library().amica_options_default(ctypes.byref(o), n, T, M, m)
    o.precision = precision
    o.max_iter = max_iter
    for k, v in kw.items():
        if not hasattr(o, k):
            raise KeyError(k)
        setattr(o, k, v)
    return o


class Engine:
    """One recording on one GPU.  Arrays use the flat layouts of amica_engine.h."""

    def __init__(self, opts, handle=None):
        """Creates an engine, or (handle given) wraps one a Batch owns."""
        self.L = load_library()
        self.o = opts
        self.owned = handle is None
        if handle is None:
            handle = _V()
            _check(self.L.amica_engine_create(ctypes.byref(opts), ctypes.byref(handle)),
                   "amica_engine_create")
        self.h = handle

    def close(self):
        if getattr(self, "h", None):
            if self.owned:
                self.L.amica_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # data -------------------------------------------------------------------
    def load_raw(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        assert x.shape == (self.o.T, self.o.n)
        _check(self.L.amica_engine_load_raw(self.h, x.ctypes.data_as(_V), 0), "load_raw")

    def load_raw_device(self, ptr):
        _check(self.L.amica_engine_load_raw(self.h, ctypes.c_void_p(ptr), 1), "load_raw")

    def preprocess(self, do_sphere=1):
        _check(self.L.amica_engine_preprocess(self.h, do_sphere), "preprocess")

    def sphere(self):
        n = self.o.n
        S = np.zeros(n * n)
        mean = np.zeros(n)
        l = np.zeros(1)
        _check(self.L.amica_engine_get_sphere(self.h, _d(S), _d(mean), _d(l)), "get_sphere")
        return S, mean, float(l[0])

    def load_white(self, xw, ldetS):
        xw = np.ascontiguousarray(xw, dtype=np.float64)
        _check(self.L.amica_engine_load_white(self.h, _d(xw), float(ldetS)), "load_white")

    def white(self):
        x = np.zeros((self.o.T, self.o.n), np.float32)
        _check(self.L.amica_engine_get_white(self.h, _f(x)), "get_white")
        return x

    # state ------------------------------------------------------------------
    def set_state(self, st):
        a = [np.ascontiguousarray(st[k], dtype=np.float64).ravel()
             for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm")]
        _check(self.L.amica_engine_set_state(self.h, *[_d(v) for v in a]), "set_state")

    def get_state(self):
        n, M, m = self.o.n, self.o.M, self.o.m
        st = dict(A=np.zeros(n * n * M), c=np.zeros(n * M), alpha=np.zeros(n * m * M),
                  mu=np.zeros(n * m * M), beta=np.zeros(n * m * M), rho=np.zeros(n * m * M),
                  gm=np.zeros(M), W=np.zeros(n * n * M))
        _check(self.L.amica_engine_get_state(self.h, *[_d(st[k]) for k in
                                                        ("A", "c", "alpha", "mu", "beta", "rho", "gm", "W")]),
               "get_state")
        return st

    # iterations -------------------------------------------------------------
    def iterate(self, k=1, graph=False):
        fn = self.L.amica_engine_iterate_graph if graph else self.L.amica_engine_iterate
        _check(fn(self.h, k), "iterate")

    def sync(self):
        _check(self.L.amica_engine_sync(self.h), "sync")

    def iterations(self):
        r = self.L.amica_engine_iterations(self.h)
        if r < 0:
            _check(r, "iterations")
        return r

    def log(self, n_iter):
        M = self.o.M
        ll = np.zeros(n_iter)
        lr = np.zeros(n_iter)
        bf = np.zeros(n_iter * M, np.int32)
        nw = np.zeros(n_iter * M, np.int32)
        dec = np.zeros(n_iter, np.int32)
        vs = np.zeros(n_iter * M)
        _check(self.L.amica_engine_get_log(self.h, n_iter, _d(ll), _d(lr), _i(bf), _i(nw), _i(dec), _d(vs)),
               "get_log")
        return dict(ll=ll, lrate=lr, bflag=bf.reshape(n_iter, M), newton=nw.reshape(n_iter, M),
                    lldec=dec, vsum=vs.reshape(n_iter, M))

    def estep_stats(self):
        n, M, m = self.o.n, self.o.M, self.o.m
        stride = 1 + 8 * n * m + 2 * n + n * n
        st = np.zeros(M * stride + 1)
        _check(self.L.amica_engine_estep_stats(self.h, _d(st)), "estep_stats")
        return st

    def record_ltall(self, on=True):
        _check(self.L.amica_engine_record_ltall(self.h, 1 if on else 0), "record_ltall")

    def ltall(self):
        lt = np.zeros(self.o.M * self.o.T)
        _check(self.L.amica_engine_get_ltall(self.h, _d(lt)), "get_ltall")
        return lt.reshape(self.o.M, self.o.T)

    def output(self):
        n, M = self.o.n, self.o.M
        A = np.zeros(n * n * M)
        c = np.zeros(n * M)
        _check(self.L.amica_engine_output(self.h, _d(A), _d(c)), "output")
        return A, c

    def set_timing(self, on=True):
        _check(self.L.amica_engine_set_timing(self.h, 1 if on else 0), "set_timing")

    def timing(self):
        """Average ms per timed iteration: inverse + constants, unmixing GEMM,
        E-step, statistic GEMM, reductions + M-step."""
        ms = np.zeros(5)
        cnt = np.zeros(1, np.int32)
        _check(self.L.amica_engine_get_timing(self.h, _d(ms), _i(cnt)), "get_timing")
        return ms, int(cnt[0])

    def debug_buffer(self, name, dtype, shape):
        out = np.zeros(shape, dtype)
        _check(self.L.amica_engine_debug_buffer(self.h, name.encode(), out.ctypes.data_as(_V), out.nbytes),
               f"debug_buffer({name})")
        return out

    def cu_count(self):
        return self.L.amica_engine_cu_count(self.h)

    def set_cu_budget(self, cus):
        """Grids sized for `cus` compute units (0: the whole device)."""
        _check(self.L.amica_engine_set_cu_budget(self.h, int(cus)), "set_cu_budget")


def hbm_probe(device=0, mib=2048, reps=5):
    """Achievable HBM bandwidth (GB/s) of this library's streaming copy
    (bytes read + written) and read-only kernels over `mib` MiB buffers."""
    c, r = np.zeros(1), np.zeros(1)
    _check(load_library().amica_hbm_probe(device, mib << 20, reps, _d(c), _d(r)), "amica_hbm_probe")
    return float(c[0]), float(r[0])


def synth_spec(n, T, seed, nmix=1, block=0, alternate=0, rho=(1.0, 2.0), sub_frac=0.0,
               sub=(2.5, 4.0), cond_max=0.0, noise_std=0.0):
    s = SynthSpec()
    s.n, s.T, s.nmix, s.block, s.alternate = n, T, nmix, block, alternate
    s.rho_lo, s.rho_hi = rho
    s.sub_frac = sub_frac
    s.sub_lo, s.sub_hi = sub
    s.cond_max, s.noise_std, s.seed = cond_max, noise_std, seed
    return s


def synth_data(spec, nthreads=0, want_truth=False):
    L = load_library()
    x = np.zeros((spec.T, spec.n), np.float32)
    A = np.zeros(spec.nmix * spec.n * spec.n)
    rho = np.zeros(spec.n)
    _check(L.amica_synth_data(ctypes.byref(spec), _f(x), _d(A), _d(rho), nthreads), "amica_synth_data")
    if want_truth:
        return x, A, rho
    return x


def synth_init(n, M, m, seed, fix_init=1, rho0=1.5):
    L = load_library()
    st = dict(A=np.zeros(n * n * M), c=np.zeros(n * M), alpha=np.zeros(n * m * M),
              mu=np.zeros(n * m * M), beta=np.zeros(n * m * M), rho=np.zeros(n * m * M), gm=np.zeros(M))
    _check(L.amica_synth_init(n, M, m, seed, fix_init, rho0, *[_d(st[k]) for k in
                                                                 ("A", "c", "alpha", "mu", "beta", "rho", "gm")]),
           "amica_synth_init")
    return st


def cond(A):
    A = np.ascontiguousarray(np.asarray(A, np.float64).T)  # column-major flat
    return load_library().amica_synth_cond(A.shape[0], _d(A.ravel()))


BATCH_POLICIES = tuple(f"{mode}, {bud}" for mode in ("one graph", "graph per stream")
                       for bud in ("whole GPU each", "work-proportional CUs", "twice proportional",
                                   "half GPU each")) + tuple(
    f"grouped launches, {k}" for k in ("each recording's own partition", "512 samples per work item",
                                        "1024 samples per work item", "2048 samples per work item"))


class Batch:
    """Many recordings per EM iteration (include/amica_batch.h).  engines[r]
    is recording r's engine (load data and a state through it)."""

    def __init__(self, opts_list):
        self.L = load_library()
        arr = (Options * len(opts_list))(*opts_list)
        h = _V()
        _check(self.L.amica_batch_create(arr, len(opts_list), ctypes.byref(h)), "amica_batch_create")
        self.h = h
        self.engines = []
        for r, o in enumerate(opts_list):
            eh = _V()
            _check(self.L.amica_batch_engine(self.h, r, ctypes.byref(eh)), "amica_batch_engine")
            self.engines.append(Engine(o, handle=eh))

    def close(self):
        if getattr(self, "h", None):
            for e in self.engines:
                e.close()
            self.L.amica_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_budgets(self, budgets=None):
        arr = None if budgets is None else _i(np.ascontiguousarray(budgets, dtype=np.int32))
        _check(self.L.amica_batch_set_budgets(self.h, arr), "amica_batch_set_budgets")

    def set_policy(self, p):
        _check(self.L.amica_batch_set_policy(self.h, p), "amica_batch_set_policy")

    def tune(self, iters=3):
        """Returns (chosen policy, ms per batched iteration for every policy)."""
        p = ctypes.c_int(0)
        ms = np.zeros(len(BATCH_POLICIES))
        _check(self.L.amica_batch_tune(self.h, iters, ctypes.byref(p), _d(ms)), "amica_batch_tune")
        return p.value, ms

    def iterate(self, k=1):
        _check(self.L.amica_batch_iterate(self.h, k), "amica_batch_iterate")

    def sync(self):
        _check(self.L.amica_batch_sync(self.h), "amica_batch_sync")


# per-sample log-likelihood and its adjoint (include/amica_adjoint.h) -------
PARAM_NAMES = ("W", "c", "alpha", "mu", "beta", "rho", "gm")


def ll_forward(real_bits, x, T, n, M, m, params, ldetS, ll, stream=0):
    """amica_ll_forward on device buffers given as integer addresses: x, ll
    and params[name] for name in PARAM_NAMES (FP64 device arrays)."""
    L = load_library()
    _check(L.amica_ll_forward(real_bits, x, T, n, M, m, *[params[k] for k in PARAM_NAMES],
                              ldetS, ll, stream), "amica_ll_forward")


def ll_forward_noalloc(real_bits, x, T, n, M, m, params, ldetS, ll, stream=0):
    """amica_ll_forward_noalloc: as ll_forward with the library's static
    workspace (allocates nothing, does not synchronise `stream`)."""
    L = load_library()
    _check(L.amica_ll_forward_noalloc(real_bits, x, T, n, M, m, *[params[k] for k in PARAM_NAMES],
                                      ldetS, ll, stream), "amica_ll_forward_noalloc")


def ll_adjoint(real_bits, x, T, n, M, m, params, ldetS, dll, grads, ll=0, stream=0):
    """amica_ll_adjoint on device addresses; grads[name] receives d/d name."""
    L = load_library()
    _check(L.amica_ll_adjoint(real_bits, x, T, n, M, m, *[params[k] for k in PARAM_NAMES], ldetS,
                              dll, ll, *[grads[k] for k in PARAM_NAMES], stream),
           "amica_ll_adjoint")
