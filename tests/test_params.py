"""The parameter-file boundary (include/config.h, include/params.h) on the CPU.

Follows fraimondo/cudamica src/config.c:12-213 and src/params.c:7-297: the
same keys, defaults, mandatory keys, on/off/none booleans and the
update_rho quirk (a missing update_rho sets do_newton, config.c:146-153);
fixtures are the reference's own example files test.in and jah.in.
"""
import ctypes
import os

import pytest

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

N = ctypes.c_ulong
R = ctypes.c_double
S = ctypes.c_char_p


class Config(ctypes.Structure):
    _fields_ = [("devs", N * 8), ("devcount", N), ("datafile", S), ("nchannels", N),
                ("nsamples", N), ("nmmodels", N), ("nsdm", N), ("maxiter", N), ("do_sphere", N),
                ("do_newton", N), ("ainitfile", S), ("cinitfile", S), ("kinitfile", S),
                ("spherefile", S), ("afile", S), ("weightsfile", S), ("khindsfile", S),
                ("cfile", S), ("llfile", S), ("ltallfile", S), ("gmfile", S), ("alphafile", S),
                ("mufile", S), ("betafile", S), ("rhofile", S), ("comp_thresh", R),
                ("ident_intvl", N), ("update_gm", N), ("update_alpha", N), ("update_A", N),
                ("update_mu", N), ("update_beta", N), ("update_rho", N), ("do_reparm", N),
                ("fix_init", N), ("lrate0", R), ("lratemax", R), ("lnatrate", R),
                ("newt_start_iter", N), ("rho_start_iter", N), ("lratefact", R),
                ("block_size", N), ("blocks_per_gpu_per_model", N), ("rho0", R), ("rholrate", R),
                ("rhomin", R), ("rhomax", R), ("showLL", N), ("numdec", N), ("maxdec", N),
                ("precision", N), ("seed", N)]


@pytest.fixture
def cfg(built):
    L = built
    L.getConfig.argtypes = [ctypes.c_char_p]
    L.getConfig.restype = ctypes.c_int
    L.freeConfig.restype = ctypes.c_int
    conf = Config.in_dll(L, "currentConfig")

    def parse(text_or_path, tmp=None):
        path = text_or_path
        if tmp is not None:
            path = os.path.join(tmp, "p.in")
            with open(path, "w") as f:
                f.write(text_or_path)
        rc = L.getConfig(path.encode())
        return rc, conf
    yield parse
    L.freeConfig()


def test_reference_test_in(cfg):
    rc, c = cfg(os.path.join(GOLDEN, "test.in"))
    assert rc == 0
    assert c.devcount == 1 and c.devs[0] == 0
    assert c.datafile == b"datasets/Memorize.fdt"
    assert c.spherefile == b"datasets/Memorize.sph"
    assert c.weightsfile == b"datasets/Memorize.wts"
    assert c.afile == b"datasets/Memorize.A" and c.cfile == b"datasets/Memorize.c"
    assert c.llfile == b"datasets/Memorize.ll" and c.ltallfile == b"datasets/Memorize.ltall"
    assert c.rhofile == b"datasets/Memorize.rho"
    assert (c.nchannels, c.nsamples, c.nmmodels, c.nsdm) == (71, 319500, 1, 3)
    assert c.update_rho == 1 and c.do_newton == 1 and c.maxiter == 2
    assert c.block_size == 256 and c.fix_init == 0
    assert c.rho_start_iter == 1 and c.newt_start_iter == 50   # last line, no trailing newline
    assert c.do_sphere == 1
    assert c.precision == 32


def test_reference_jah_in_unknown_keys(cfg):
    """jah.in uses key names the reference does not parse; they are ignored,
    and the missing update_rho sets do_newton (config.c:146-153)."""
    rc, c = cfg(os.path.join(GOLDEN, "jah.in"))
    assert rc == 0
    assert (c.nchannels, c.nsamples, c.nmmodels, c.nsdm, c.maxiter) == (128, 51200, 2, 2, 1)
    assert c.do_sphere == 1            # "sphering 1" is not a key: default
    assert c.update_rho == 0           # "updateRho" is not a key: quirk path
    assert c.do_newton == 1
    assert c.newt_start_iter == 50 and c.block_size == 5000


def test_defaults_and_fixed_values(cfg, tmp_path):
    rc, c = cfg("datafile d.fdt\nnchannels 4\nnsamples 100\nspherefile s.sph\nupdate_rho 0\n", str(tmp_path))
    assert rc == 0
    assert (c.maxiter, c.nsdm, c.nmmodels, c.block_size, c.fix_init) == (200, 3, 1, 5000, 1)
    assert c.do_newton == 1 and c.update_rho == 0
    assert (c.lrate0, c.lratemax, c.lnatrate, c.lratefact) == (0.1, 1.0, 0.1, 0.5)
    assert (c.rho0, c.rholrate, c.rhomin, c.rhomax) == (1.5, 0.1, 1.0, 2.0)
    assert c.maxdec == 3 and c.do_reparm == 1 and c.afile is None and c.seed == 1234


def test_missing_mandatory_keys(cfg, tmp_path):
    rc, c = cfg("nchannels 4\nnsamples 100\n", str(tmp_path))
    assert rc == -2      # ERRORNOPARAM for datafile / spherefile


def test_booleans_and_invalid_values(cfg, tmp_path):
    base = "datafile d\nnchannels 4\nnsamples 9\nspherefile s\nupdate_rho 1\n"
    rc, c = cfg(base + "do_newton off\nfix_init none\n", str(tmp_path))
    assert rc == 0 and c.do_newton == 0 and c.fix_init == 2
    rc, c = cfg(base + "do_newton maybe\n", str(tmp_path))
    assert rc == -3      # ERRORINVALIDPARAM
    rc, c = cfg(base + "maxiter\n", str(tmp_path))
    assert rc == -3


def test_exact_key_match_and_comments(cfg, tmp_path):
    """`afile` must not be taken from the `datafile` line (the reference
    matches keys by substring first, params.c:12-17)."""
    text = "# datafile commented.fdt\ndatafile real.fdt\nnchannels 3\nnsamples 10\nspherefile s\n"
    rc, c = cfg(text, str(tmp_path))
    assert rc == 0 and c.datafile == b"real.fdt" and c.afile is None


def test_devices_list_and_precision(cfg, tmp_path):
    base = "datafile d\nnchannels 4\nnsamples 9\nspherefile s\n"
    rc, c = cfg(base + "devices 2 0 1\nprecision bf16\nseed 99\n", str(tmp_path))
    assert rc == 0 and c.devcount == 3 and list(c.devs[:3]) == [2, 0, 1]
    assert c.precision == 16 and c.seed == 99
    rc, c = cfg(base + "precision fp8\n", str(tmp_path))
    assert rc == 0 and c.precision == 8
    rc, c = cfg(base + "precision fp4\n", str(tmp_path))
    assert rc == -3
    rc, c = cfg(base + "precision fp32plain\n", str(tmp_path))
    assert rc == 0 and c.precision == 33
    # more device ordinals than the list holds: an error, nothing written past it
    rc, c = cfg(base + "devices" + " 1" * 70 + "\n", str(tmp_path))
    assert rc == -3


def test_missing_file(cfg):
    rc, c = cfg("/nonexistent/param.in")
    assert rc == -4      # ERRORINVALIDCONFIG
