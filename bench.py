#!/usr/bin/env python3
"""Benchmark: AMICA EM iterations per second on one MI355X.

Workload (default): 128 channels x 2,000,000 samples, M=4 models, m=3
mixture components, non-stationary synthetic data (4 mixing matrices over
50k-sample blocks), BF16 GEMM operands with FP32 density math and FP64
statistics / M-step.  A step is one full EM iteration (unmixing GEMM, fused
E-step, statistic GEMM, reductions, M-step with Newton/natural gradient and
reparameterisation) over all samples, which stay resident in HBM.

Prints one JSON line.  Fields beyond the driver's contract:
  roofline      HBM roofline of the fused E-step, from HIP events recorded
                around it on the engine stream during the timed steps; also
                against the bandwidth of this library's own streaming copy
                kernel (amica_hbm_probe)
  valu          the E-step's VALU-issue bound (the one it runs against): SIMD
                cycles its VALU instructions occupy per launch (rocprofv3
                SQ_ACTIVE_INST_VALU x 4, from profiles/valu_<workload>_<prec>.json)
                over the SIMD cycles of 1024 SIMDs at the 2400 MHz maximum engine
                clock during the live average launch time; frac_under_counters
                uses the cycles the counter run itself counted (GRBM_GUI_ACTIVE)
  gemm          algorithmic GEMM FLOPs (4 M n^2 T) over the two GEMMs' time,
                against the dense MFMA peak of their operand type
  cpu_baseline  the FP64 CPU oracle (C++, OpenMP on all host cores) on the
                whole recording, 2 iterations (median), with its FP32
                variant beside it
  accuracy_vs_fp64  after 50 iterations on 60k samples of the workload's
                generator (its 40 blocks scaled to 1.5k samples), from the
                same initial state: LL error against the FP64 oracle, per
                model the Amari distance between the GPU's and the
                oracle's W, the relative errors of alpha, mu, beta, rho,
                and each run's Amari distance to the true mixing matrices;
                for a BF16 / FP8 run the same for the FP32 path (fp32_path)
  adjoint       amica_ll_forward / amica_ll_adjoint (FP32 density math) on the
                whole whitened recording at the state the iterations reached:
                milliseconds per call
  plain_fp32    BF16 / FP8 runs: the same workload on the plain FP32 path (both
                GEMMs on FP32 MFMA, FP32 b and g), 5 iterations after 2, and
                speedup_vs_plain_fp32 = value / its rate
  ll_vs_fp32_path     BF16 / FP8 runs: the relative LL difference to the FP32
                path (FP64 unmixing, held to the oracle by the parity tests)
                after 50 iterations on the whole recording
  ll_rel_err_vs_fp64  max over the CPU baseline's iterations on the whole
                recording (same data, same initial state) of
                |LL_gpu - LL_fp64| / |LL_fp64|;
                ll_rel_err_per_iteration lists them (the first is the
                E-step alone, the second follows the first M-step)
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "EM iters/s @128ch×2M samples,M=4; % HBM roofline; rel. ΔLL vs FP64"
HBM_PEAK_GBS = 8000.0


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


PRECISIONS = {"fp32": 32, "fp32plain": 33, "bf16": 16, "fp8": 8}
PREC_NAME = {v: k for k, v in PRECISIONS.items()}
# arithmetic types of the path: GEMM operands + density math
DTYPE = {32: "f64+f32", 33: "f32", 16: "bf16+f32", 8: "fp8+bf16+f32"}


def estep_bytes(n, T, M, precision):
    """Algorithmic HBM bytes of one E-step launch: M n T (b read + g written)."""
    g_bytes = 4 if precision in (32, 33) else 2
    return M * n * T * (4 + g_bytes)


def _cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def cpu_baseline(w, x, st0, precision, iters=2):
    """The reference's update math on the host: the FP64 oracle (and its
    FP32 variant) on the whole recording, `iters` EM iterations each on all
    host cores, timed per iteration (median); and the GPU engine on the same
    data for the per-iteration LL check against FP64."""
    import cudamica_amd as ca
    from oracle import pyoracle as po
    cores = _cores()
    n, M, m, T = w.n, w.M, w.m, w.T
    xw, S, mean, ldetS = po.preprocess(x, w.do_sphere)
    res = {}
    for bits in (64, 32):
        log(f"cpu baseline: oracle FP{bits}, {cores} threads, {iters} iterations on all {T} samples")
        st = {k: np.array(v, copy=True) for k, v in st0.items()}
        per = []
        logs = None
        for k in range(iters):
            cfg = po.default_config(n, T, M, m, maxiter=1, newt_start_iter=w.newt_start_iter)
            t0 = time.perf_counter()
            nst, ll, lr, _ = po.step(cfg, k + 1, xw, ldetS, st,
                                     cfg.lrate0 if k == 0 else lr, ll_prev=0.0 if k == 0 else ll,
                                     precision=bits, nthreads=cores)
            per.append(time.perf_counter() - t0)
            st = nst
            if bits == 64:
                logs = (logs or []) + [ll]
        res[bits] = (float(np.median(per)), per, logs)
    del xw
    with ca.Engine(w.options(precision=precision, max_iter=iters)) as eng:
        eng.load_raw(x)
        eng.preprocess(w.do_sphere)
        eng.set_state(st0)
        eng.iterate(iters)
        eng.sync()
        ll_gpu = eng.log(iters)["ll"]
    ll64 = np.array(res[64][2])
    rel = np.abs(ll_gpu - ll64) / np.abs(ll64)
    med64, per64, _ = res[64]
    med32, per32, _ = res[32]
    return {"value": 1.0 / med64, "unit": "EM iters/s", "cores": cores, "kind": "port",
            "sample": f"FP64 oracle (C++, OpenMP), {iters} EM iterations on all {T} samples, "
                      f"median of {[round(t, 2) for t in per64]} s",
            "fp32_variant": {"value": 1.0 / med32, "unit": "EM iters/s",
                             "iteration_s": [round(t, 3) for t in per32]}}, rel


def _amari(W, A):
    """Amari index of P = W A, normalised to [0, 1]: 0 when P is a scaled
    permutation (W unmixes A up to order and scale)."""
    P = np.abs(W @ A)
    n = P.shape[0]
    r = (P / P.max(axis=1, keepdims=True)).sum(axis=1) - 1.0
    c = (P / P.max(axis=0, keepdims=True)).sum(axis=0) - 1.0
    return float((r.sum() + c.sum()) / (2.0 * n * (n - 1)))


def accuracy_vs_fp64(w, x, st0, precision, T_acc=60000, iters=50):
    """Error of the GPU path against the FP64 oracle after `iters` iterations
    on T_acc samples of the workload's generator (its blocks scaled down),
    both from the same data and initial state:
    LL (last and max over the run), per model the Amari distance between the
    two final W and the relative errors of alpha, mu, beta, rho; and each
    model's Amari distance from the nearest true mixing matrix (of the ones
    that occur in those samples), for both runs.  A reduced-precision run
    also reports the FP32 path on the same problem (fp32_path)."""
    import cudamica_amd as ca
    from oracle import pyoracle as po
    n, M, m = w.n, w.M, w.m
    T_acc = min(T_acc, w.T)
    # the workload's generator at T_acc samples with its block structure
    # scaled down (the same number of blocks, each T_acc / T as long), so that
    # every mixing matrix and so every model has data
    spec = w.spec(T=T_acc)
    if w.nmix > 1 and w.block > 0:
        spec.block = max(1, T_acc // max(1, w.T // w.block))
    xs, A_true, _ = ca.synth_data(spec, want_truth=True)
    A_true = A_true.reshape(spec.nmix, n, n).transpose(0, 2, 1)
    t0 = time.perf_counter()
    xw, S, mean, ldetS = po.preprocess(xs, w.do_sphere)
    cfg = po.default_config(n, T_acc, M, m, maxiter=iters, newt_start_iter=w.newt_start_iter)
    fin, logs = po.run(cfg, xw, ldetS, st0, nthreads=_cores())
    log(f"accuracy: FP64 oracle, {iters} iterations on {T_acc} samples in {time.perf_counter() - t0:.1f} s")
    So = S.reshape(n, n).T
    nn, nm = n * n, n * m

    def gpu_run(prec):
        with ca.Engine(w.options(T=T_acc, precision=prec, max_iter=iters)) as eng:
            eng.load_raw(xs)
            eng.preprocess(w.do_sphere)
            Sg = eng.sphere()[0].reshape(n, n).T
            eng.set_state(st0)
            eng.iterate(iters)
            eng.sync()
            glog = eng.log(iters)
            gst = eng.get_state()
        rel = np.abs(glog["ll"] - logs["ll"]) / np.abs(logs["ll"])
        res = {"ll_rel_last": float(rel[-1]), "ll_rel_max": float(rel.max()),
               "ll_rel_max_iteration": int(np.argmax(rel)) + 1,
               "decisions_equal": bool(np.array_equal(glog["lldec"], logs["lldec"])
                                       and np.array_equal(glog["newton"], logs["newton"])),
               "models": []}
        for h in range(M):
            Wg = gst["W"][h * nn:(h + 1) * nn].reshape(n, n).T
            Wo = fin["W"][h * nn:(h + 1) * nn].reshape(n, n).T
            sl = slice(h * nm, (h + 1) * nm)
            d = {"amari_W_vs_oracle": float(_amari(Wg, np.linalg.inv(Wo))),
                 "amari_truth_gpu": float(min(_amari(Wg @ Sg, At) for At in A_true)),
                 "amari_truth_oracle": float(min(_amari(Wo @ So, At) for At in A_true))}
            for k in ("alpha", "mu", "beta", "rho"):
                d[f"{k}_rel"] = float(np.linalg.norm(gst[k][sl] - fin[k][sl]) / np.linalg.norm(fin[k][sl]))
            res["models"].append(d)
        worst = max(mm["amari_W_vs_oracle"] for mm in res["models"])
        log(f"accuracy ({PREC_NAME[prec]}) after {iters} iterations on {T_acc} samples: LL rel "
            f"{rel[-1]:.2e} (max {rel.max():.2e}), worst W Amari vs oracle {worst:.2e}")
        return res

    out = {"samples": T_acc, "iterations": iters, **gpu_run(precision)}
    if precision != 32:
        # the FP32 path on the same problem, against the same oracle run
        out["fp32_path"] = gpu_run(32)
    return out


# dense MFMA peaks of the GEMMs' operand type (MI355X, TFLOP/s): split-BF16
# and FP8 paths run bf16 MFMAs, the FP32 path FP64 (unmixing) and FP32, the
# plain FP32 path FP32
GEMM_PEAK_TFLOPS = {16: 2500.0, 8: 2500.0, 32: 78.6, 33: 157.3}


def measured_hbm_gbs(mib=2048, reps=5):
    """Achievable HBM bandwidth from this library's own streaming kernels
    (amica_hbm_probe): a copy (bytes read + written per second) and a
    read-only pass, best of `reps` over `mib` MiB buffers."""
    import cudamica_amd as ca
    try:
        return ca.hbm_probe(0, mib, reps)
    except Exception as ex:  # noqa: BLE001 -- reported, not fatal to the bench line
        log(f"hbm probe failed: {ex}")
        return None, None


def ll_vs_fp32_path(w, x, st0, precision, iters=50):
    """The reduced-precision path against the FP32 path (FP64 unmixing, held
    to the FP64 oracle by the parity tests) on the whole recording, both from
    the same state for `iters` iterations.  The FP8 path is compared with the
    FP32 path on its own dequantised data (the data it runs on).  Returns the
    relative LL difference at the last iteration, its maximum over the run
    and where that was, and both runs' counts of LL-decrease iterations."""
    import cudamica_amd as ca
    logs = []
    with ca.Engine(w.options(precision=precision, max_iter=iters)) as eng:
        eng.load_raw(x)
        eng.preprocess(w.do_sphere)
        if precision == 8:
            xh = eng.white().astype(np.float64)
            ldetS = eng.sphere()[2]
        eng.set_state(st0)
        eng.iterate(iters, graph=True)
        eng.sync()
        logs.append(eng.log(iters))
    with ca.Engine(w.options(precision=32, max_iter=iters)) as eng:
        if precision == 8:
            eng.load_white(xh, ldetS)
            del xh
        else:
            eng.load_raw(x)
            eng.preprocess(w.do_sphere)
        eng.set_state(st0)
        eng.iterate(iters, graph=True)
        eng.sync()
        logs.append(eng.log(iters))
    rel = np.abs(logs[0]["ll"] - logs[1]["ll"]) / np.abs(logs[1]["ll"])
    imax = int(np.argmax(rel))
    log(f"LL vs the FP32 path over {iters} iterations: last {rel[-1]:.2e}, max {rel[imax]:.2e} "
        f"(iteration {imax + 1})")
    return {"iterations": iters, "rel_last": float(rel[-1]), "rel_max": float(rel[imax]),
            "rel_max_iteration": imax + 1,
            "reference": "fp32 path on the dequantised data" if precision == 8 else "fp32 path",
            "lldec_iterations": [int(np.sum(l["lldec"])) for l in logs],
            "ll_last": float(logs[0]["ll"][-1]), "ll_last_fp32": float(logs[1]["ll"][-1])}


def plain_fp32_rate(w, x, st0, steps=5, warmup=2):
    """The same workload on the plain FP32 path (both GEMMs on FP32 MFMA, FP32
    b and g): EM iterations per second over `steps` iterations after
    `warmup`, timed as the headline is."""
    import cudamica_amd as ca
    with ca.Engine(w.options(precision=ca.PREC_FP32_PLAIN, max_iter=warmup + steps)) as eng:
        eng.load_raw(x)
        eng.preprocess(w.do_sphere)
        eng.set_state(st0)
        eng.iterate(warmup)
        eng.sync()
        t0 = time.perf_counter()
        eng.iterate(steps)
        eng.sync()
        dt = time.perf_counter() - t0
    log(f"plain FP32 path: {dt * 1e3 / steps:.3f} ms/iter")
    return {"value": steps / dt, "unit": "EM iters/s", "ms_per_step": dt * 1e3 / steps,
            "steps": steps, "warmup": warmup}


def adjoint_rate(xw, ldetS, st, n, M, m, reps=3):
    """amica_ll_forward and amica_ll_adjoint (real_bits 32: FP32 density math,
    FP64 sums and gradients) over the whole whitened recording at the state
    the timed iterations reached, dll = 1: median milliseconds per call of
    `reps` after one untimed call.  Device buffers come from torch."""
    import torch
    import cudamica_amd as ca
    dev = torch.device("cuda:0")
    T = xw.shape[0]
    x = torch.from_numpy(np.ascontiguousarray(xw, np.float32)).to(dev)
    ps = {k: torch.tensor(np.asarray(st[k], np.float64).ravel(), device=dev) for k in ca.PARAM_NAMES}
    gs = {k: torch.empty_like(p) for k, p in ps.items()}
    ll = torch.empty(T, dtype=torch.float64, device=dev)
    dll = torch.ones(T, dtype=torch.float64, device=dev)
    pa = {k: p.data_ptr() for k, p in ps.items()}
    ga = {k: g.data_ptr() for k, g in gs.items()}
    stream = torch.cuda.current_stream().cuda_stream

    def fwd():
        ca.ll_forward(32, x.data_ptr(), T, n, M, m, pa, ldetS, ll.data_ptr(), stream)

    def adj():
        ca.ll_adjoint(32, x.data_ptr(), T, n, M, m, pa, ldetS, dll.data_ptr(), ga, 0, stream)

    out = {}
    for name, fn in (("forward_ms", fwd), ("adjoint_ms", adj)):
        fn()
        torch.cuda.synchronize()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()                      # each call synchronises its stream
            ts.append((time.perf_counter() - t0) * 1e3)
        out[name] = float(np.median(ts))
    log(f"ll forward {out['forward_ms']:.2f} ms, adjoint {out['adjoint_ms']:.2f} ms per call")
    return {"real_bits": 32, "samples": T, **out,
            "note": "per call on the whole recording, workspace allocation and synchronisation "
                    "included; gradients w.r.t. W, c, alpha, mu, beta, rho, gm with dll = 1"}


def copy_probe(mib):
    """One hipMemset and one device-to-device hipMemcpy of `mib` MiB (the HIP
    runtime's fill and copy kernels), so that a rocprofv3 --pmc pass over this
    run can scale FETCH_SIZE and WRITE_SIZE by what they read for traffic of
    known size."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    nbytes = ctypes.c_size_t(mib << 20)
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), nbytes) != 0:
            raise RuntimeError("hipMalloc failed in the copy probe")
    ok = (hip.hipMemset(a, 1, nbytes) == 0 and hip.hipMemcpy(b, a, nbytes, 3) == 0
          and hip.hipDeviceSynchronize() == 0)
    hip.hipFree(a)
    hip.hipFree(b)
    if not ok:
        raise RuntimeError("copy probe failed")
    log(f"copy probe: memset + copy of {mib} MiB")


def measured_traffic(workload, prec):
    """HBM bytes per E-step launch measured with rocprofv3 --pmc FETCH_SIZE and
    WRITE_SIZE passes (scaled by the copy probe), kept in profiles/ as JSON."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"traffic_{workload}_{PREC_NAME[prec]}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        return json.load(f).get("traffic_bytes")


def valu_roofline(workload, prec, avg_ms):
    """The E-step's VALU-issue bound: the SIMD cycles its VALU instructions
    occupy per launch (rocprofv3 SQ_ACTIVE_INST_VALU quad-cycles x 4, kept in
    profiles/ as JSON) over the SIMD cycles the GPU offers at its maximum
    engine clock during the live average launch time."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        f"valu_{workload}_{PREC_NAME[prec]}.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        v = json.load(f)
    cyc = v["valu_issue_cycles"]
    peak = v["simds"] * v["max_engine_clock_mhz"] * 1e6      # SIMD cycles per second
    achieved = cyc / (avg_ms * 1e-3)
    c = v["counters"]
    return {"bound": "valu-issue", "kernel": "estep_kernel", "achieved": achieved / 1e9,
            "peak": peak / 1e9, "unit": "G SIMD-cycles/s", "frac": achieved / peak,
            "issue_cycles_per_launch": cyc,
            "valu_insts_per_launch": c["SQ_INSTS_VALU"],
            "transcendental_insts_per_launch": c["SQ_INSTS_VALU_TRANS_F32"],
            "frac_under_counters": v["valu_issue_frac_under_counters"],
            "clock_mhz_under_counters": v["clock_mhz_under_counters"]}


def bench_batch(args):
    """Many recordings per EM iteration (one hipGraph): the batch workload of
    32 recordings, n in {32, 64, 128}, T ~ U[100k, 1.5M], M = 1, m = 3."""
    import cudamica_amd as ca
    prec = PRECISIONS[args.precision or "fp32"]
    ws = ca.batch_workload(precision=prec)
    t0 = time.perf_counter()
    opts = [w.options(max_iter=args.warmup + args.steps + 8 + 3 * len(ca.BATCH_POLICIES)) for w in ws]
    b = ca.Batch(opts)
    for w, e in zip(ws, b.engines):
        e.load_raw(ca.synth_data(w.spec()))
        e.preprocess(w.do_sphere)
        e.set_state(w.init_state())
    log(f"batch of {len(ws)} recordings ({sum(w.n * w.T for w in ws) / 1e9:.2f} G channel-samples) "
        f"ready in {time.perf_counter() - t0:.1f} s")
    p, ms = b.tune(iters=2)
    log("tuner (ms per batched iteration): " +
        ", ".join(f"{nm} {t:.2f}" for nm, t in zip(ca.BATCH_POLICIES, ms)) +
        f"; using {ca.BATCH_POLICIES[p]}")
    b.iterate(args.warmup)
    b.sync()
    t0 = time.perf_counter()
    b.iterate(args.steps)
    b.sync()
    dt = time.perf_counter() - t0
    lls = [float(e.log(args.warmup + args.steps)["ll"][-1]) for e in b.engines]
    # each recording alone on the GPU (its own one-iteration graph), the
    # batched iteration's yardstick: the sum of the solo iteration times
    solo = []
    for e in b.engines:
        e.iterate(1, graph=True)
        e.sync()
        t1 = time.perf_counter()
        e.iterate(2, graph=True)
        e.sync()
        solo.append((time.perf_counter() - t1) / 2)
    b.close()
    log(f"batched iteration {dt * 1e3 / args.steps:.2f} ms; sum of the recordings' solo iterations "
        f"{sum(solo) * 1e3:.2f} ms")
    out = {
        "metric": "batched EM iters/s (32 recordings per iteration, one hipGraph)",
        "value": args.steps / dt, "unit": "batched EM iters/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": DTYPE[prec], "data": "synthetic",
        "config": {"workload": "batch32_n32-128_T100k-1.5M_M1", "recordings": len(ws),
                   "precision": PREC_NAME[prec], "parallelism": "single-gpu"},
        "tuner_ms": dict(zip(ca.BATCH_POLICIES, [float(t) for t in ms])),
        "policy": ca.BATCH_POLICIES[p],
        "recording_iters_per_s": len(ws) * args.steps / dt,
        "solo_ms_sum": sum(solo) * 1e3,
        "solo_ms": [t * 1e3 for t in solo],
        "ll_finite": bool(np.all(np.isfinite(lls))),
    }
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="n128_T2M_M4",
                    help="a workload of cudamica_amd.WORKLOADS, or batch32 (32 recordings)")
    ap.add_argument("--precision", choices=sorted(PRECISIONS), default=None)
    ap.add_argument("--T", type=int, default=None, help="override the sample count")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true",
                    help="skip the 50-iteration error report against the FP64 oracle")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="skip timing the likelihood and its adjoint on the recording")
    ap.add_argument("--no-plain", action="store_true",
                    help="skip timing the plain FP32 path on the same workload")
    ap.add_argument("--no-ll-check", action="store_true",
                    help="skip the 50-iteration LL comparison with the FP32 path")
    ap.add_argument("--cu-budget", type=int, default=0, metavar="CUS",
                    help="size the grids for CUS compute units (occupancy studies; 0: all)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="run a memset + copy of MIB MiB first (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this engine does not split samples across GPUs "
              "(one GPU owns all samples); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch32":
        bench_batch(args)
        return
    if not args.no_adjoint:
        # torch provides the adjoint timing's device buffers; it must bring up
        # its HIP context before this library's first HIP call
        import torch
        torch.cuda.init()
    import cudamica_amd as ca
    w = ca.WORKLOADS[args.workload]
    if args.T:
        w.T = args.T
    prec = w.precision if args.precision is None else PRECISIONS[args.precision]
    n, T, M, m = w.n, w.T, w.M, w.m

    if args.copy_probe:
        copy_probe(args.copy_probe)
    t0 = time.perf_counter()
    x = ca.synth_data(w.spec())
    st0 = w.init_state()
    log(f"data {T}x{n} generated in {time.perf_counter() - t0:.1f} s")

    timing_steps = min(args.steps, 10)
    eng = ca.Engine(w.options(precision=prec, max_iter=args.warmup + args.steps + timing_steps + 1))
    if args.cu_budget:
        eng.set_cu_budget(args.cu_budget)
    eng.load_raw(x)
    t0 = time.perf_counter()
    eng.preprocess(w.do_sphere)
    log(f"preprocessing on the GPU: {time.perf_counter() - t0:.2f} s")
    eng.set_state(st0)
    eng.iterate(args.warmup)
    eng.sync()
    log(f"warmup {args.warmup} iterations done")

    t0 = time.perf_counter()
    eng.iterate(args.steps)
    eng.sync()
    dt = time.perf_counter() - t0
    # per-kernel times from HIP events on the engine stream, in further
    # iterations (the timed ones run without events)
    eng.set_timing(True)
    eng.iterate(timing_steps)
    eng.sync()
    kms, cnt = eng.timing()
    eng.set_timing(False)
    lg = eng.log(args.warmup + args.steps + timing_steps)
    adj_in = None
    if not args.no_adjoint:
        adj_in = (eng.white(), eng.sphere()[2], eng.get_state())
    eng.close()
    its = args.steps / dt
    log(f"timed {args.steps} iterations: {dt * 1e3 / args.steps:.3f} ms/iter; kernels (ms) "
        f"inverse {kms[0]:.3f} unmix {kms[1]:.3f} estep {kms[2]:.3f} gx {kms[3]:.3f} "
        f"reduce+mstep {kms[4]:.3f}; LL {lg['ll'][-1]:.10f}")

    nbytes = estep_bytes(n, T, M, prec)
    achieved = nbytes / (kms[2] * 1e-3) / 1e9
    hbm, hbm_read = measured_hbm_gbs()
    gemm_flop = 4.0 * M * n * n * T          # GEMM1 + GEMM2, algorithmic
    gemm_s = (kms[1] + kms[3]) * 1e-3
    gemm_peak = GEMM_PEAK_TFLOPS[prec]
    out = {
        "metric": METRIC,
        "value": its,
        "unit": "EM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": DTYPE[prec],
        "data": "synthetic",
        "config": {"workload": w.name, "n": n, "T": T, "M": M, "m": m,
                   "precision": PREC_NAME[prec], "parallelism": "single-gpu",
                   **({"cu_budget": args.cu_budget} if args.cu_budget else {})},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(w.name, prec),
                     "kernel": "estep_kernel", "bytes_per_launch": nbytes,
                     "avg_ms": kms[2], "launches": cnt,
                     "measured_copy_gbs": hbm, "measured_read_gbs": hbm_read,
                     "frac_of_measured": achieved / hbm if hbm else None},
        "valu": valu_roofline(w.name, prec, kms[2]),
        "gemm": {"flop_per_iter": gemm_flop, "ms": gemm_s * 1e3,
                 "tflops": gemm_flop / gemm_s / 1e12,
                 "peak_tflops": gemm_peak, "mfma_frac": gemm_flop / gemm_s / 1e12 / gemm_peak},
        "kernel_ms": {"inverse": kms[0], "unmix": kms[1], "estep": kms[2], "stat_gemm": kms[3],
                      "reduce_mstep": kms[4]},
        "ll_last": float(lg["ll"][-1]),
    }
    if adj_in is not None:
        out["adjoint"] = adjoint_rate(adj_in[0], adj_in[1], adj_in[2], n, M, m)
        adj_in = None
    if prec in (16, 8) and not args.no_plain:
        pl = plain_fp32_rate(w, x, st0)
        out["plain_fp32"] = pl
        out["speedup_vs_plain_fp32"] = its / pl["value"]
    if prec != 32 and not args.no_ll_check:
        out["ll_vs_fp32_path"] = ll_vs_fp32_path(w, x, st0, prec)
    if not args.no_accuracy:
        out["accuracy_vs_fp64"] = accuracy_vs_fp64(w, x, st0, prec)
    if not args.no_cpu_baseline:
        cb, rel = cpu_baseline(w, x, st0, prec)
        out["cpu_baseline"] = cb
        out["ll_rel_err_vs_fp64"] = float(np.max(rel))
        out["ll_rel_err_per_iteration"] = [float(r) for r in rel]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
