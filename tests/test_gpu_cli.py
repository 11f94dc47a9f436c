"""The drop-in command line program on the GPU: parameter file in, the
reference's stdout lines and raw float64 result files out.  The files must
be exactly what the engine's C ABI computes from the same inputs, and the
LL trajectory must follow the CPU oracle on the same data and initial
state.  The program replays one captured hipGraph per iteration and reads
only that iteration's log entry; with `jsonlogfile` it runs each iteration
eagerly between HIP events and writes one JSON line per iteration (LL,
lrate, per-model vsum / bflag / Newton flag, stage times, E-step bytes);
both ways give the same trajectory bit for bit."""
import os
import re
import subprocess

import numpy as np
import pytest

import cudamica_amd as ca
from oracle import pyoracle as po

pytestmark = pytest.mark.gpu


def test_cli_end_to_end(built, tmp_path):
    n, T, M, m, iters = 12, 6000, 2, 3, 6
    print("[cli] n=12 T=6000 M=2 m=3, 6 iterations", flush=True)
    x = ca.synth_data(ca.synth_spec(n, T, seed=501, nmix=2, block=1500, alternate=1, cond_max=0.0))
    st0 = ca.synth_init(n, M, m, 502)
    d = str(tmp_path)
    x.tofile(os.path.join(d, "data.fdt"))
    st0["A"].tofile(os.path.join(d, "A0.bin"))
    st0["c"].tofile(os.path.join(d, "c0.bin"))
    outs = ["weightsfile", "afile", "cfile", "llfile", "ltallfile", "gmfile", "alphafile", "mufile",
            "betafile", "rhofile"]
    lines = [f"datafile {d}/data.fdt", f"nchannels {n}", f"nsamples {T}", f"nmmodels {M}",
             f"nsdm {m}", f"maxiter {iters}", f"spherefile {d}/out.sph", "do_sphere 1",
             "update_rho 1", "do_newton on", "newt_start_iter 3", f"ainitfile {d}/A0.bin",
             f"cinitfile {d}/c0.bin", "# a comment line", "unknownkey 5", f"jsonlogfile {d}/iters.jsonl"]
    lines += [f"{k} {d}/out.{k}" for k in outs]
    with open(os.path.join(d, "run.param"), "w") as f:
        f.write("\n".join(lines) + "\n")
    r = subprocess.run([ca.CLI_PATH, "-f", os.path.join(d, "run.param")], capture_output=True,
                       text=True, timeout=180)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    got = [(int(a), float(b), float(c)) for a, b, c in
           re.findall(r"Iteration (\d+): lrate ([-0-9.e+]+) LL ([-0-9.e+]+)", r.stdout)]
    assert [g[0] for g in got] == list(range(1, iters + 1))

    xw, S, mean, ldetS = po.preprocess(x, 1)
    cfg = po.default_config(n, T, M, m, maxiter=iters, newt_start_iter=3)
    fin, logs = po.run(cfg, xw, ldetS, st0)
    ll_cli = np.array([g[2] for g in got])

    # 1. the program is the engine: the same data, options and initial state
    #    through the C ABI give the same iterations and the same files
    with ca.Engine(ca.options(n, T, M, m, max_iter=iters, do_newton=1, newt_start_iter=3)) as eng:
        eng.load_raw(x)
        eng.preprocess(1)
        eng.set_state(st0)
        eng.iterate(iters)
        eng.sync()
        elog = eng.log(iters)
        est = eng.get_state()
        Aout, cout = eng.output()
    np.testing.assert_allclose(ll_cli, elog["ll"], rtol=1e-14)
    np.testing.assert_allclose([g[1] for g in got], elog["lrate"], rtol=1e-14)

    def rd(k, cnt):
        a = np.fromfile(os.path.join(d, f"out.{k}"), dtype=np.float64)
        assert a.size == cnt, (k, a.size, cnt)
        return a

    np.testing.assert_allclose(rd("llfile", iters), ll_cli, rtol=1e-15)
    # one JSON line per iteration with the same LL, lrate and per-model vsum
    import json
    with open(os.path.join(d, "iters.jsonl")) as f:
        js = [json.loads(line) for line in f]
    assert [j["iteration"] for j in js] == list(range(1, iters + 1))
    np.testing.assert_array_equal([j["ll"] for j in js], elog["ll"])
    np.testing.assert_array_equal([j["lrate"] for j in js], elog["lrate"])
    np.testing.assert_array_equal([[mm["vsum"] for mm in j["models"]] for j in js], elog["vsum"])
    assert all(j["ms"] > 0 for j in js)
    stages = ("inverse", "unmix", "estep", "stat_gemm", "reduce_mstep")
    assert all(j["stages_ms"][k] > 0 for j in js for k in stages)
    assert all(sum(j["stages_ms"].values()) <= j["ms"] for j in js)
    assert all(j["estep_bytes"] == M * n * T * (8 + 4) for j in js)     # FP32 path: b FP64, g FP32
    # without the JSON log every iteration replays the captured graph: the
    # same trajectory, bit for bit
    with open(os.path.join(d, "run_graph.param"), "w") as f:
        f.write("\n".join(l for l in lines if not l.startswith("jsonlogfile")) + "\n")
    r2 = subprocess.run([ca.CLI_PATH, "-f", os.path.join(d, "run_graph.param")], capture_output=True,
                        text=True, timeout=180)
    assert r2.returncode == 0, r2.stdout[-3000:] + r2.stderr[-3000:]
    got2 = re.findall(r"Iteration (\d+): lrate ([-0-9.e+]+) LL ([-0-9.e+]+)", r2.stdout)
    assert got2 == re.findall(r"Iteration (\d+): lrate ([-0-9.e+]+) LL ([-0-9.e+]+)", r.stdout)
    np.testing.assert_allclose(rd("weightsfile", n * n * M), est["W"], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(rd("afile", n * n * M), Aout, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(rd("cfile", n * M), cout, rtol=1e-12, atol=1e-14)
    for k, key, cnt in (("gmfile", "gm", M), ("alphafile", "alpha", n * m * M),
                        ("mufile", "mu", n * m * M), ("betafile", "beta", n * m * M),
                        ("rhofile", "rho", n * m * M)):
        np.testing.assert_allclose(rd(k, cnt), est[key], rtol=1e-12, atol=1e-14, err_msg=k)
    np.testing.assert_allclose(np.fromfile(os.path.join(d, "out.sph")), S, rtol=1e-9, atol=1e-12)
    lt = rd("ltallfile", T * M)
    assert np.isfinite(lt).all()

    # 2. the numbers are the reference's: iteration 1 (a function of the data
    #    and the initial state only) within 1e-6 of the FP64 oracle, and the
    #    rest of the trajectory within 4x the oracle's own spread when its
    #    whitened input is rounded to float32 (with Newton from iteration 3
    #    on 6000 samples, rounding grows by orders of magnitude per iteration)
    rel = np.abs(ll_cli - logs["ll"]) / np.abs(logs["ll"])
    assert rel[0] <= 1e-6, rel
    _, lp = po.run(cfg, xw.astype(np.float32).astype(np.float64), ldetS, st0)
    env = np.maximum.accumulate(np.abs(lp["ll"] - logs["ll"]) / np.abs(logs["ll"]))
    print(f"[cli] LL rel err {np.array2string(rel, precision=2)}; "
          f"oracle spread {np.array2string(env, precision=2)}", flush=True)
    assert (rel <= np.maximum(1e-6, 4.0 * env)).all()
