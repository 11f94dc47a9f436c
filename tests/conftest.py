import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP engine)")


def flat_to_np(st, n, M, m):
    """Flat engine/oracle layouts -> numpy shapes used by oracle/amica_np.py."""
    return dict(
        A=st["A"].reshape(M, n, n).transpose(0, 2, 1).copy(),
        c=st["c"].reshape(M, n).copy(),
        alpha=st["alpha"].reshape(M, m, n).copy(),
        mu=st["mu"].reshape(M, m, n).copy(),
        beta=st["beta"].reshape(M, m, n).copy(),
        rho=st["rho"].reshape(M, m, n).copy(),
        gm=np.asarray(st["gm"], np.float64).copy(),
    )


def amari_distance(W, A):
    """Amari index of P = W A (0 when P is a scaled permutation), normalised to [0, 1]."""
    P = np.abs(W @ A)
    n = P.shape[0]
    r = (P / P.max(axis=1, keepdims=True)).sum(axis=1) - 1.0
    c = (P / P.max(axis=0, keepdims=True)).sum(axis=0) - 1.0
    return (r.sum() + c.sum()) / (2.0 * n * (n - 1))


@pytest.fixture(scope="session")
def built():
    from cudamica_amd import engine
    if not os.path.exists(engine.LIB_PATH):
        pytest.fail("libcudamica_amd.so not built (run __graft_entry__.build())")
    return engine.load_library()


@pytest.fixture(scope="session")
def oracle_lib():
    from oracle import pyoracle
    return pyoracle.lib()
