"""Pins the CPU oracle (oracle/oracle.cpp) before it is trusted as the checker.

The reference ships no golden vectors or known-answer tests for this path
(its author compared dumps against MATLAB scripts that are not in the
repository), so the oracle is pinned by:
  * an independent numpy/scipy restatement (oracle/amica_np.py) on tiny
    inputs -- two restatements agreeing to ~1e-12;
  * closed forms: the generalized-Gaussian log-density against
    scipy.stats.gennorm, digamma against scipy, pinv / log|det| / eigh
    against numpy.linalg, the sphering matrix against 2 C^-1/2;
  * invariants of the E-step and M-step;
  * the natural-gradient identity sum_t dl_t/dW W^T = vsum I - phi, by
    finite differences of the log-likelihood;
  * brute force with mpmath on the smallest case;
  * planted mistakes in the numpy restatement that the comparison must catch.
"""
import numpy as np
import pytest
from scipy import special, stats

from oracle import amica_np as an
from oracle import pyoracle as po
from tests.conftest import flat_to_np

LOG2_REF = 0.693147180559945


def _tiny(n=4, T=300, M=2, m=3, seed=0):
    rng = np.random.default_rng(seed)
    S = rng.laplace(size=(T, n))
    x = (S @ rng.normal(size=(n, n)).T).astype(np.float32)
    A0 = np.stack([np.eye(n) + 0.05 * (rng.random((n, n)) - 0.5) for _ in range(M)])
    A0 /= np.sqrt((A0 ** 2).sum(1, keepdims=True))
    st = dict(A=A0, c=rng.normal(size=(M, n)), alpha=np.full((M, m, n), 1.0 / m),
              mu=np.tile((np.arange(m) - (m - 1) / 2.0)[None, :, None], (M, 1, n)).astype(float),
              beta=np.ones((M, m, n)), rho=np.full((M, m, n), 1.5), gm=np.full(M, 1.0 / M))
    flat = dict(A=np.ascontiguousarray(A0.transpose(0, 2, 1)).ravel(), c=st["c"].ravel(),
                alpha=st["alpha"].ravel(), mu=st["mu"].ravel(), beta=st["beta"].ravel(),
                rho=st["rho"].ravel(), gm=st["gm"].copy())
    return x, st, flat


def _stats_np_to_flat(stats, psum, n, m):
    parts = []
    for d in stats:
        parts.append([d["vsum"]])
        for k in ("usum", "kappa_numer", "lambda_numer", "mu_numer", "mu_denom", "beta_numer",
                  "beta_denom", "rho_numer"):
            parts.append(np.asarray(d[k]).ravel())
        parts.append(d["sigma2_numer"])
        parts.append(d["cnew"])
        parts.append(d["phi"].T.ravel())
    parts.append([psum])
    return np.concatenate([np.asarray(p, np.float64).ravel() for p in parts])


@pytest.mark.parametrize("n,T,M,m", [(3, 200, 1, 3), (4, 300, 2, 2), (5, 257, 3, 3), (6, 128, 1, 1)])
def test_estep_matches_numpy(n, T, M, m):
    x, st, flat = _tiny(n, T, M, m, seed=n + M)
    xw, S, mean, ldetS = po.preprocess(x, 1)
    st["rho"] = np.random.default_rng(1).uniform(1.0, 2.0, size=st["rho"].shape)
    flat["rho"] = st["rho"].ravel().copy()
    ora = po.estep(xw, ldetS, flat["A"], flat["c"], flat["alpha"], flat["mu"], flat["beta"],
                   flat["rho"], flat["gm"], M, m)
    stats, psum, _ = an.estep(xw, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"],
                              st["rho"], st["gm"])
    ref = _stats_np_to_flat(stats, psum, n, m)
    np.testing.assert_allclose(ora, ref, rtol=1e-11, atol=1e-11)


def test_preprocess_matches_numpy():
    x, _, _ = _tiny(6, 500, 1, 3, seed=3)
    for sph in (0, 1, 2):
        xw, S, mean, ldetS = po.preprocess(x, sph)
        xw2, S2, mean2, ldetS2 = an.preprocess(x, sph)
        np.testing.assert_allclose(xw, xw2, rtol=1e-11, atol=1e-11)
        np.testing.assert_allclose(S.reshape(6, 6).T, S2, rtol=1e-11, atol=1e-12)
        assert abs(ldetS - ldetS2) < 1e-11


def test_sphering_closed_form():
    """S = 2 C^-1/2, so the whitened covariance is 4 I (preprocess.cu:173-229)."""
    x, _, _ = _tiny(5, 2000, 1, 3, seed=4)
    xw, S, mean, ldetS = po.preprocess(x, 1)
    C = np.cov(xw.T, bias=True)
    np.testing.assert_allclose(C, 4.0 * np.eye(5), atol=1e-10)
    xc = x.astype(np.float64) - x.astype(np.float64).mean(0)
    Craw = xc.T @ xc / len(xc)
    w, V = np.linalg.eigh(Craw)
    np.testing.assert_allclose(S.reshape(5, 5).T, 2.0 * (V * w ** -0.5) @ V.T, rtol=1e-10)
    assert abs(ldetS + np.linalg.slogdet(2.0 * (V * w ** -0.5) @ V.T)[1]) < 1e-10


@pytest.mark.parametrize("n,M,m,iters,newt", [(3, 1, 3, 6, 50), (4, 2, 2, 6, 2), (4, 1, 3, 8, 3)])
def test_em_iterations_match_numpy(n, M, m, iters, newt):
    x, st, flat = _tiny(n, 300, M, m, seed=10 + n)
    xw, S, mean, ldetS = po.preprocess(x, 1)
    cfg = po.default_config(n, 300, M, m, maxiter=iters, newt_start_iter=newt)
    fin, logs = po.run(cfg, xw, ldetS, flat)
    st2, lls, lrs, logs2 = an.run(an.default_cfg(n, 300, M, m, newt_start_iter=newt), xw, ldetS, st, iters)
    np.testing.assert_allclose(logs["ll"], lls, rtol=1e-12)
    np.testing.assert_allclose(logs["lrate"], lrs, rtol=1e-15)
    assert [list(r) for r in logs["newton"]] == [l["newton"] for l in logs2]
    assert [list(r) for r in logs["bflag"]] == [l["bflag"] for l in logs2]
    fnp = flat_to_np(fin, n, M, m)
    for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm"):
        np.testing.assert_allclose(fnp[k], st2[k], rtol=1e-7, atol=1e-9, err_msg=k)


def test_update_A_off_still_moves_c():
    """With update_A off the reference skips the whole A update (Newton test,
    lrate growth, reparameterisation, amica.cu:2156-2447) but still copies
    cnew into c after that block (amica.cu:2450-2462): A stays, c becomes the
    posterior-weighted data mean of the first E-step."""
    n, M, m = 4, 2, 3
    x, st, flat = _tiny(n, 300, M, m, seed=31)
    xw, S, mean, ldetS = po.preprocess(x, 1)
    cfg = po.default_config(n, 300, M, m, maxiter=1, update_A=0)
    fin, logs = po.run(cfg, xw, ldetS, flat)
    np.testing.assert_array_equal(fin["A"], flat["A"])
    stats, _ = po.stats_views(po.estep(xw, ldetS, flat["A"], flat["c"], flat["alpha"], flat["mu"],
                                       flat["beta"], flat["rho"], flat["gm"], M, m), n, m, M)
    want = np.concatenate([d["cnew"] / d["vsum"] for d in stats])
    np.testing.assert_allclose(fin["c"], want, rtol=1e-13)
    assert not np.allclose(fin["c"], flat["c"])
    assert (logs["newton"] == 0).all() and (logs["bflag"] == 0).all()


def test_one_iteration_agrees_to_1e11():
    x, st, flat = _tiny(5, 400, 2, 3, seed=21)
    xw, S, mean, ldetS = po.preprocess(x, 1)
    cfg = po.default_config(5, 400, 2, 3, maxiter=1, newt_start_iter=1)
    fin, logs = po.run(cfg, xw, ldetS, flat)
    st2, lls, _, _ = an.run(an.default_cfg(5, 400, 2, 3, newt_start_iter=1), xw, ldetS, st, 1)
    fnp = flat_to_np(fin, 5, 2, 3)
    for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm"):
        np.testing.assert_allclose(fnp[k], st2[k], rtol=1e-11, atol=1e-12, err_msg=k)


def test_density_closed_form():
    """Q of logalogblogpfun (kernels.cu:282-307) is the generalized-Gaussian
    log-density (scipy.stats.gennorm) plus log alpha + .5 log beta; rho = 2
    is a Gaussian of variance 1/2 and rho = 1 the Laplace density."""
    y = np.linspace(-4, 4, 41)
    for rho in (1.0, 1.3, 1.5, 2.0, 3.1):
        q = -LOG2_REF - special.gammaln(1.0 + 1.0 / rho) - np.abs(y) ** rho
        np.testing.assert_allclose(q, stats.gennorm.logpdf(y, rho), atol=5e-15)
    np.testing.assert_allclose(-LOG2_REF - special.gammaln(1.5) - y ** 2,
                               stats.norm.logpdf(y, scale=np.sqrt(0.5)), atol=5e-15)
    np.testing.assert_allclose(-LOG2_REF - special.gammaln(2.0) - np.abs(y),
                               stats.laplace.logpdf(y), atol=5e-15)


def test_psi_against_scipy():
    for r in np.linspace(1.0, 4.0, 61):
        x = 1.0 + 1.0 / r
        assert abs(po.psi(x) - special.digamma(x)) <= 1e-14 * max(1.0, abs(special.digamma(x))) + 1e-15
    for x in (0.3, 1.0, 2.5, 10.0):
        assert abs(po.psi(x) - special.digamma(x)) < 1e-13


def test_linear_algebra():
    rng = np.random.default_rng(5)
    for n in (1, 3, 16, 40):
        A = rng.normal(size=(n, n))
        np.testing.assert_allclose(po.pinv(A), np.linalg.pinv(A), rtol=1e-9, atol=1e-11)
        assert abs(po.logabsdet(A) - np.linalg.slogdet(A)[1]) < 1e-10
        C = A @ A.T + n * np.eye(n)
        w, V = po.eigh(C)
        np.testing.assert_allclose(w, np.linalg.eigvalsh(C), rtol=1e-12)
        np.testing.assert_allclose((V * w) @ V.T, C, rtol=1e-11, atol=1e-10)
    B = np.diag([2.0, 0.0, 3.0])   # exact-zero singular value: truncated, as amica.cu:64-76
    np.testing.assert_allclose(po.pinv(B), np.diag([0.5, 0.0, 1.0 / 3.0]), atol=1e-15)


def test_estep_invariants():
    n, T, M, m = 5, 400, 3, 3
    x, st, flat = _tiny(n, T, M, m, seed=7)
    xw, S, mean, ldetS = po.preprocess(x, 1)
    ora = po.estep(xw, ldetS, flat["A"], flat["c"], flat["alpha"], flat["mu"], flat["beta"],
                   flat["rho"], flat["gm"], M, m)
    views, psum = po.stats_views(ora, n, m, M)
    vs = sum(v["vsum"] for v in views)
    assert abs(vs - T) < 1e-9                          # sum_h v = 1 per sample
    for v in views:                                    # sum_j z = 1 per (h, i, t)
        np.testing.assert_allclose(v["usum"].sum(axis=0), v["vsum"], rtol=1e-12)
    cfg = po.default_config(n, T, M, m, maxiter=1)
    fin, logs = po.run(cfg, xw, ldetS, flat)
    np.testing.assert_allclose(fin["alpha"].reshape(M, m, n).sum(axis=1), 1.0, rtol=1e-12)
    assert abs(fin["gm"].sum() - 1.0) < 1e-12


def _ll_sum(x, ldetS, st):
    _, psum, _ = an.estep(x, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"], st["rho"],
                          st["gm"], update_rho=False)
    return psum


def test_natural_gradient_identity():
    """sum_t dl_t/dW_h W_h^T = vsum_h I - phi_h (the geteyemphi term,
    kernels.cu:1003-1014), by central differences in W = A^-1."""
    n, T, M, m = 3, 200, 2, 2
    x, st, flat = _tiny(n, T, M, m, seed=8)
    xw, _, _, ldetS = po.preprocess(x, 1)
    st["rho"] = np.random.default_rng(2).uniform(1.2, 1.9, size=st["rho"].shape)
    stats, psum, aux = an.estep(xw, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"],
                                st["rho"], st["gm"])
    h = 1
    W = aux["W"][h]
    G = np.zeros((n, n))
    eps = 1e-6
    for i in range(n):
        for k in range(n):
            for sgn in (1, -1):
                Wp = W.copy()
                Wp[i, k] += sgn * eps
                s2 = {kk: v.copy() for kk, v in st.items()}
                s2["A"][h] = np.linalg.inv(Wp)
                G[i, k] += sgn * _ll_sum(xw, ldetS, s2) / (2 * eps)
    lhs = G @ W.T
    rhs = stats[h]["vsum"] * np.eye(n) - stats[h]["phi"]
    np.testing.assert_allclose(lhs, rhs, rtol=2e-5, atol=2e-5)


def test_ll_brute_force_mpmath():
    mp = pytest.importorskip("mpmath")
    mp.mp.dps = 40
    n, T, M, m = 2, 8, 2, 2
    x, st, flat = _tiny(n, T, M, m, seed=9)
    xw, _, _, ldetS = po.preprocess(x, 1)
    ora = po.estep(xw, ldetS, flat["A"], flat["c"], flat["alpha"], flat["mu"], flat["beta"],
                   flat["rho"], flat["gm"], M, m)
    tot = mp.mpf(0)
    for t in range(T):
        terms = []
        for h in range(M):
            Wm = mp.inverse(mp.matrix(st["A"][h].tolist()))
            b = Wm * (mp.matrix(xw[t].tolist()) - mp.matrix(st["c"][h].tolist()))
            L = mp.log(st["gm"][h]) + mp.log(abs(mp.det(Wm))) + ldetS
            for i in range(n):
                s = mp.mpf(0)
                for j in range(m):
                    a, mu, be, r = (st[k][h, j, i] for k in ("alpha", "mu", "beta", "rho"))
                    y = mp.sqrt(be) * (b[i] - mu)
                    s += mp.exp(mp.log(a) + mp.log(be) / 2 - mp.mpf(LOG2_REF) - mp.loggamma(1 + 1 / mp.mpf(r))
                                - abs(y) ** r)
                L += mp.log(s)
            terms.append(L)
        mx = max(terms)
        tot += mx + mp.log(sum(mp.exp(L - mx) for L in terms))
    assert abs(float(tot) - ora[-1]) < 1e-11 * abs(float(tot))


@pytest.mark.parametrize("mistake", ["drop_half_log_beta", "mu_shift_sign", "phi_transpose",
                                     "kappa_old_alpha"])
def test_planted_mistakes_are_caught(mistake, monkeypatch):
    """A plausible error planted in the numpy restatement must break its
    agreement with the C++ oracle within two iterations."""
    n, T, M, m = 4, 300, 2, 3
    x, st, flat = _tiny(n, T, M, m, seed=31)
    xw, _, _, ldetS = po.preprocess(x, 1)
    cfg = po.default_config(n, T, M, m, maxiter=2, newt_start_iter=1)
    fin, logs = po.run(cfg, xw, ldetS, flat)
    orig_estep, orig_mstep = an.estep, an.mstep

    if mistake == "drop_half_log_beta":
        def bad_estep(x_, l_, A, c, alpha, mu, beta, rho, gm, update_rho=True):
            return orig_estep(x_, l_, A, c, alpha ** 1.0 / np.sqrt(beta), mu, beta, rho, gm, update_rho)
        monkeypatch.setattr(an, "estep", bad_estep)
    elif mistake == "mu_shift_sign":
        def bad_mstep(cfg_, it, stats, s, lrate, numdec, hist):
            c_old = s["c"].copy()
            r = orig_mstep(cfg_, it, stats, s, lrate, numdec, hist)
            for h in range(M):
                s["mu"][h] += 2 * (s["W"][h] @ (s["c"][h] - c_old[h]))[None]
            return r
        monkeypatch.setattr(an, "mstep", bad_mstep)
    elif mistake == "phi_transpose":
        def bad_estep(*a, **k):
            stats, psum, aux = orig_estep(*a, **k)
            for d in stats:
                d["phi"] = d["phi"].T.copy()
            return stats, psum, aux
        monkeypatch.setattr(an, "estep", bad_estep)
    elif mistake == "kappa_old_alpha":
        def bad_mstep(cfg_, it, stats, s, lrate, numdec, hist):
            saved = s["alpha"].copy()
            for h in range(M):
                stats[h]["kappa_numer"] = stats[h]["kappa_numer"] * saved[h] * m
            return orig_mstep(cfg_, it, stats, s, lrate, numdec, hist)
        monkeypatch.setattr(an, "mstep", bad_mstep)
    st2, lls, _, _ = an.run(an.default_cfg(n, T, M, m, newt_start_iter=1), xw, ldetS, st, 2)
    fnp = flat_to_np(fin, n, M, m)
    diffs = [np.max(np.abs(fnp[k] - st2[k])) for k in ("A", "mu", "beta", "alpha")]
    diffs.append(np.max(np.abs(logs["ll"] - lls)))
    assert max(diffs) > 1e-6, f"planted mistake {mistake} went undetected"
