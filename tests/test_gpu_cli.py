"""The drop-in command line program on the GPU: parameter file in, the
reference's stdout lines and raw float64 result files out, checked against
the CPU oracle on the same data and initial state."""
import os
import re
import subprocess

import numpy as np
import pytest

import cudamica_amd as ca
from oracle import pyoracle as po
from tests.conftest import amari_distance

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
             f"cinitfile {d}/c0.bin", "# a comment line", "unknownkey 5"]
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
    fin, logs = po.run(cfg, xw, ldetS, st0, want_ltall=True)
    # plumbing test on a tiny recording (T = 6000, Newton from iteration 3),
    # where FP32 rounding is amplified faster than on realistic sizes; the
    # numeric bar (1e-6) is held by tests/test_gpu_engine.py
    ll_cli = np.array([g[2] for g in got])
    np.testing.assert_allclose(ll_cli, logs["ll"], rtol=2e-5)
    np.testing.assert_allclose([g[1] for g in got], logs["lrate"], rtol=1e-12)

    def rd(k, cnt):
        a = np.fromfile(os.path.join(d, f"out.{k}"), dtype=np.float64)
        assert a.size == cnt, (k, a.size, cnt)
        return a

    np.testing.assert_allclose(rd("llfile", iters), ll_cli, rtol=1e-15)
    np.testing.assert_allclose(np.fromfile(os.path.join(d, "out.sph")), S, rtol=1e-9, atol=1e-12)
    W = rd("weightsfile", n * n * M)
    for h in range(M):
        Wg = W[h * n * n:(h + 1) * n * n].reshape(n, n).T
        Wo = fin["W"][h * n * n:(h + 1) * n * n].reshape(n, n).T
        assert amari_distance(Wg, np.linalg.inv(Wo)) < 1e-3
    Ao, co = po.output_transform(S, mean, fin["A"], fin["c"], n, M)
    np.testing.assert_allclose(rd("afile", n * n * M), Ao, rtol=1e-2, atol=1e-3)
    np.testing.assert_allclose(rd("cfile", n * M), co, rtol=1e-3, atol=1e-4)
    np.testing.assert_allclose(rd("gmfile", M), fin["gm"], rtol=1e-4)
    for k, key in (("alphafile", "alpha"), ("mufile", "mu"), ("betafile", "beta"), ("rhofile", "rho")):
        np.testing.assert_allclose(rd(k, n * m * M), fin[key], rtol=1e-3, atol=1e-4, err_msg=k)
    lt = rd("ltallfile", T * M)
    np.testing.assert_allclose(lt, logs["ltall"].ravel(), rtol=1e-4, atol=1e-4 * np.abs(lt).max())
