"""The built artefacts on the CPU: the C-ABI library loads and exports every
function declared in include/*.h; the command line program handles a
missing argument; the synthetic generator is deterministic."""
import os
import re
import subprocess

import numpy as np

import cudamica_amd as ca

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    names = set()
    decl = re.compile(r"^\s*(?:extern\s+)?[A-Za-z_][\w\s\*]*?\b([a-zA-Z_]\w*)\s*\(")
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        text = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", text)
        for line in text.splitlines():
            m = decl.match(line)
            if m and not line.strip().startswith(("typedef", "return", "if", "#")):
                names.add(m.group(1))
    return names - {"sizeof"}


def test_library_exports_every_declared_symbol(built):
    out = subprocess.run(["nm", "-D", "--defined-only", ca.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if l.strip()}
    declared = _declared_functions()
    assert len(declared) > 60
    missing = sorted(declared - exported)
    assert not missing, missing
    for g in ("currentConfig", "programName", "gpus", "gpuCount", "blocks", "nblocks"):
        assert g in exported


def test_cli_requires_parameter_file(built):
    r = subprocess.run([ca.CLI_PATH], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0
    assert "configuration file is mandatory" in r.stdout


def test_engine_reports_errors_without_touching_gpu(built):
    import pytest
    with pytest.raises(ca.EngineError):
        ca.Engine(ca.options(0, 100, 1, 3))            # n out of range
    with pytest.raises(ca.EngineError):
        ca.Engine(ca.options(1025, 100, 1, 3))         # beyond the reference's 1024 channels
    with pytest.raises(ca.EngineError):
        ca.Engine(ca.options(64, 100, 17, 3))          # beyond its 16 models
    with pytest.raises(ca.EngineError):
        ca.Engine(ca.options(64, 100, 2, 9))           # beyond 8 mixture components


def test_synth_deterministic_and_unit_variance(built):
    spec = ca.synth_spec(8, 30000, seed=42, rho=(1.0, 2.0), sub_frac=0.25, cond_max=20.0)
    x1, A, rho = ca.synth_data(spec, nthreads=1, want_truth=True)
    x4 = ca.synth_data(spec, nthreads=4)
    assert np.array_equal(x1, x4)
    A = A.reshape(8, 8).T
    assert ca.cond(A) <= 20.0
    assert np.all((rho[:2] >= 2.5) & (rho[:2] <= 4.0)) and np.all((rho[2:] >= 1.0) & (rho[2:] <= 2.0))
    s = np.linalg.solve(A, x1.astype(np.float64).T)
    np.testing.assert_allclose(s.var(axis=1), 1.0, atol=0.05)
    x2 = ca.synth_data(ca.synth_spec(8, 30000, seed=43, cond_max=20.0))
    assert not np.array_equal(x1, x2)


def test_synth_nonstationary_blocks(built):
    spec = ca.synth_spec(4, 4000, seed=5, nmix=2, block=1000, alternate=1)
    x, A, _ = ca.synth_data(spec, want_truth=True)
    A0, A1 = A[:16].reshape(4, 4).T, A[16:].reshape(4, 4).T
    s0 = np.linalg.solve(A0, x[:1000].astype(np.float64).T)
    s1 = np.linalg.solve(A1, x[1000:2000].astype(np.float64).T)
    np.testing.assert_allclose(s0.var(axis=1), 1.0, atol=0.2)
    np.testing.assert_allclose(s1.var(axis=1), 1.0, atol=0.2)


def test_init_state_recipe(built):
    n, M, m = 6, 2, 3
    st = ca.synth_init(n, M, m, 11)
    A = st["A"].reshape(M, n, n)
    np.testing.assert_allclose(np.sqrt((A ** 2).sum(axis=2)), 1.0, rtol=1e-14)   # unit columns
    assert np.all(np.abs(A.transpose(0, 2, 1) - np.eye(n)) < 0.06)
    np.testing.assert_allclose(st["mu"].reshape(M, m, n)[:, :, 0], [[-1, 0, 1]] * M)
    assert np.all(st["beta"] == 1.0) and np.all(st["rho"] == 1.5)
    np.testing.assert_allclose(st["alpha"], 1.0 / m)
    np.testing.assert_allclose(st["gm"], 0.5)


def test_ll_entry_points_validate_before_touching_gpu(built):
    """amica_ll_forward / amica_ll_adjoint reject bad arguments with
    AMICA_E_INVALID (-3) before any device call (dummy addresses are never
    dereferenced)."""
    L = ca.load_library()
    p = [1] * 7
    assert L.amica_ll_forward(16, 1, 10, 4, 1, 3, *p, 0.0, 1, None) == -3        # real_bits
    assert L.amica_ll_forward(32, 1, 10, 4, 1, 9, *p, 0.0, 1, None) == -3        # m > 8
    assert L.amica_ll_forward(32, 1, 10, 1025, 1, 3, *p, 0.0, 1, None) == -3     # n > 1024
    assert L.amica_ll_forward(32, 1, 10, 16, 17, 3, *p, 0.0, 1, None) == -3      # M > 16
    assert L.amica_ll_forward_noalloc(32, 1, 10, 300, 2, 3, *p, 0.0, 1, None) == -3  # M NP > 512
    assert L.amica_ll_forward(64, 1, 10, 4, 1, 3, *p, 0.0, None, None) == -3     # no ll output
    assert b"ll output" in L.amica_last_error()
    assert L.amica_ll_adjoint(64, 1, 10, 4, 1, 3, *p, 0.0, None, None, *p, None) == -3  # no dll


REF_MAIN = "/root/reference/src/cudaamica.cu"


def test_reference_main_builds_against_these_headers(built, tmp_path):
    """The reference's own main (src/cudaamica.cu:17-43), unchanged, compiles
    as C99 against include/ and links against libcudamica_amd.so: the
    boundary keeps its names and signatures.  Skipped where the reference
    tree is absent."""
    import pytest
    if not os.path.exists(REF_MAIN):
        pytest.skip("reference sources not present")
    exe = str(tmp_path / "cudaamica_ref_main")
    r = subprocess.run(["gcc", "-x", "c", "-std=c99", "-Wall", "-I", os.path.join(ROOT, "include"),
                        REF_MAIN, "-x", "none", "-o", exe, "-L", os.path.dirname(ca.LIB_PATH),
                        "-lcudamica_amd", f"-Wl,-rpath,{os.path.dirname(ca.LIB_PATH)}",
                        "-Wl,-rpath,/opt/rocm/lib"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert "configuration file is mandatory" in r.stdout
