"""bench.py's contract pieces that need no GPU: the algorithmic E-step bytes
per launch (M n T (b read + g written)) against the worked values of the
measurement plan (64 ch x 1M: 0.51 GB FP32; 128 ch x 2M x 4 models: 8.19 GB
FP32, 6.14 GB with bf16 g; 256 ch x 4M x 2 models: 16.4 / 12.3 GB), and the
single-GPU refusal of --gpus N > 1."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_estep_bytes_match_the_worked_values():
    assert bench.estep_bytes(64, 1_000_000, 1, 32) == 512_000_000
    assert bench.estep_bytes(128, 2_000_000, 4, 32) == 8_192_000_000
    assert bench.estep_bytes(128, 2_000_000, 4, 16) == 6_144_000_000
    assert bench.estep_bytes(128, 2_000_000, 4, 33) == 8_192_000_000     # plain FP32: b and g FP32
    assert bench.estep_bytes(256, 4_000_000, 2, 32) == 16_384_000_000
    assert bench.estep_bytes(256, 4_000_000, 2, 8) == 12_288_000_000


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_amari_index_is_zero_exactly_for_scaled_permutations():
    import numpy as np
    rng = np.random.default_rng(3)
    A = rng.normal(size=(6, 6))
    P = np.eye(6)[rng.permutation(6)] * rng.uniform(0.5, 2.0, size=6) * rng.choice([-1, 1], size=6)
    W = P @ np.linalg.inv(A)
    assert bench._amari(W, A) < 1e-12
    assert bench._amari(np.linalg.inv(A + 0.1 * rng.normal(size=(6, 6))), A) > 1e-3


def test_valu_roofline_from_the_committed_counters():
    """The valu block recomputes its fractions from the committed counter
    profile: 4 x SQ_ACTIVE_INST_VALU SIMD cycles per launch over 1024 SIMDs
    at 2400 MHz times the launch time."""
    import json
    v = bench.valu_roofline("n128_T2M_M4", 16, 3.0)
    prof = json.load(open(os.path.join(ROOT, "profiles", "valu_n128_T2M_M4_bf16.json")))
    cyc = 4 * prof["counters"]["SQ_ACTIVE_INST_VALU"]
    assert v["issue_cycles_per_launch"] == cyc
    assert abs(v["frac"] - cyc / (1024 * 2.4e9 * 3.0e-3)) < 1e-12
    grbm_per_xcd = prof["counters"]["GRBM_GUI_ACTIVE"] / 8
    assert abs(v["frac_under_counters"] - cyc / 1024 / grbm_per_xcd) < 1e-12
    assert 0.0 < v["frac_under_counters"] <= 1.0
    assert bench.valu_roofline("n128_T2M_M4", 8, 3.0) is None      # no profile for that path
