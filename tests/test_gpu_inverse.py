"""The FP64 unmixing matrices W_h = A_h^-1 and log|det W_h| that every EM
iteration derives from the mixing matrices on the GPU (Gauss-Jordan with
partial pivoting: one register-resident workgroup per model up to 192
channels, blocked by panels beyond), against numpy's LAPACK inverse and
slogdet, for sizes on both sides of every kernel shape (n <= 64, 128, 192,
256, 512) and for matrices that force row exchanges at every step."""
import numpy as np
import pytest

import cudamica_amd as ca

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _matrices(n, M, kind, rng):
    out = []
    for h in range(M):
        if kind == "random":
            A = rng.standard_normal((n, n))
        elif kind == "permuted":      # zero diagonal: the natural pivot order fails at step 0
            P = np.eye(n)[rng.permutation(n)]
            while np.any(np.diag(P) != 0):
                P = np.eye(n)[rng.permutation(n)]
            A = P * rng.uniform(1.0, 2.0, (n, 1)) + 1e-3 * rng.standard_normal((n, n)) * (P == 0)
        else:                         # graded: condition number ~1e8
            U, _ = np.linalg.qr(rng.standard_normal((n, n)))
            V, _ = np.linalg.qr(rng.standard_normal((n, n)))
            A = U @ np.diag(np.logspace(0, -8, n)) @ V.T
        out.append(A)
    return out


@pytest.mark.parametrize("n,M,kind", [(1, 2, "random"), (5, 2, "random"), (64, 2, "random"),
                                      (65, 2, "permuted"), (128, 4, "random"), (129, 1, "graded"),
                                      (150, 2, "graded"), (192, 3, "permuted"), (193, 1, "random"),
                                      (200, 2, "permuted"),
                                      (256, 2, "random"), (300, 1, "random"), (512, 1, "graded")])
def test_inverse_and_logdet(built, n, M, kind):
    rng = np.random.default_rng(1000 + n)
    _say(f"[inverse] n={n} M={M} {kind}")
    As = _matrices(n, M, kind, rng)
    T = 2048
    x = ca.synth_data(ca.synth_spec(n, T, seed=n, cond_max=0.0))
    st = ca.synth_init(n, M, 3, 7)
    # column-major storage: A_h[i][k] at h n n + k n + i
    st["A"] = np.concatenate([A.T.ravel() for A in As])
    with ca.Engine(ca.options(n, T, M, 3)) as eng:
        eng.load_raw(x)
        eng.preprocess(1)
        eng.set_state(st)
        eng.estep_stats()
        W = eng.get_state()["W"]
        ldet = eng.debug_buffer("ldet", np.float64, (M,))
    for h, A in enumerate(As):
        Wg = W[h * n * n:(h + 1) * n * n].reshape(n, n).T
        Wn = np.linalg.inv(A)
        cond = np.linalg.cond(A)
        err = np.linalg.norm(Wg - Wn) / np.linalg.norm(Wn)
        res = np.linalg.norm(Wg @ A - np.eye(n)) / np.sqrt(n)
        _, lad = np.linalg.slogdet(A)
        _say(f"  model {h}: cond {cond:.1e}, |W - inv(A)|/|inv(A)| {err:.2e}, "
             f"|W A - I| {res:.2e}, log|det W| {ldet[h]:.12f} vs {-lad:.12f}")
        assert err <= 1e-15 * n * cond
        assert res <= 1e-15 * n * cond
        assert abs(ldet[h] + lad) <= 1e-12 * max(1.0, abs(lad)) * n
