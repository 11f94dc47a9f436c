"""Pins the oracle's M-step and curvature statistics by what the mathematics
fixes, independently of the oracle's own formulas.

Every expected value here comes from library routines and numerics, not
from a retyped oracle formula:
  * responsibilities from scipy.stats.gennorm log-densities and
    scipy.special.logsumexp;
  * score functions f = -d/dy log p(y) and the mixture score by
    Richardson-extrapolated finite differences of those library densities;
  * the gradient of the total log-likelihood by finite differences (rho,
    W) of the same library log-likelihood;
  * maximisers of the expected complete-data log-likelihood by
    scipy.optimize;
  * invariances of the log-likelihood (column scaling of A with the
    matching mu, beta rescaling; the c / mu shift);
  * hand-computed known answers for the learning-rate rules.
Each check runs once on the clean oracle (it must pass) and once for each
mistake planted in the oracle that it is meant to catch (it must fail),
through orc_plant.  References: fraimondo/cudamica src/kernels.cu:583-1065
(statistics, kappa/lambda, rho, Newton B), src/amica.cu:1963-2447 (lrate,
M-step, A update, reparameterisation).
"""
import contextlib

import numpy as np
import pytest
from scipy import optimize, special, stats

from oracle import pyoracle as po
from tests.conftest import flat_to_np

STAT_NAMES = ("usum", "kappa_numer", "lambda_numer", "mu_numer", "mu_denom", "beta_numer",
              "beta_denom", "rho_numer", "sigma2_numer", "cnew", "phi")


@contextlib.contextmanager
def planted(name):
    po.plant(name)
    try:
        yield
    finally:
        po.plant(None)


def _expect(ok, plant):
    """The clean oracle must pass the check, a planted one must fail it."""
    if plant is None:
        assert ok, "the clean oracle fails the check"
    else:
        assert not ok, f"planted mistake {plant} went undetected"


# ---------------------------------------------------------------- fixtures --
def _data(n, T, seed, src_rho=1.3):
    rng = np.random.default_rng(seed)
    S = stats.gennorm.rvs(src_rho, size=(T, n), random_state=rng)
    x = S @ rng.normal(size=(n, n)).T
    return x - x.mean(axis=0)


def _state(n, M, m, seed, rho=(1.15, 1.9), mu_scale=0.4, c_scale=0.3):
    rng = np.random.default_rng(seed)
    A = np.stack([np.eye(n) + 0.3 * rng.normal(size=(n, n)) for _ in range(M)])
    a = rng.uniform(0.3, 1.0, size=(M, m, n))
    g = rng.uniform(0.5, 1.0, size=M)
    st = dict(A=A, c=c_scale * rng.normal(size=(M, n)), alpha=a / a.sum(axis=1, keepdims=True),
              mu=mu_scale * rng.normal(size=(M, m, n)), beta=rng.uniform(0.5, 2.0, size=(M, m, n)),
              rho=rng.uniform(rho[0], rho[1], size=(M, m, n)), gm=g / g.sum())
    return st


def _flat(st):
    return dict(A=np.ascontiguousarray(st["A"].transpose(0, 2, 1)).ravel(), c=st["c"].ravel(),
                alpha=st["alpha"].ravel(), mu=st["mu"].ravel(), beta=st["beta"].ravel(),
                rho=st["rho"].ravel(), gm=np.asarray(st["gm"], np.float64).ravel())


def _np(flat, n, M, m):
    return flat_to_np({k: np.asarray(v) for k, v in flat.items()}, n, M, m)


def _cfg(n, T, M, m, **kw):
    base = dict(update_rho=0, update_A=0, do_reparm=0, do_newton=0)
    base.update(kw)
    return po.default_config(n, T, M, m, maxiter=1, **base)


# --------------------------------------------- library-density machinery --
def _d_dy(fun, y, rel=1e-3):
    """d fun / dy by central differences with a step relative to |y| (no
    step crosses y = 0), Richardson-extrapolated: error O(rel^4)."""
    h = rel * np.abs(y)
    d1 = (fun(y + h) - fun(y - h)) / (2.0 * h)
    d2 = (fun(y + 2.0 * h) - fun(y - 2.0 * h)) / (4.0 * h)
    return (4.0 * d1 - d2) / 3.0


def _score(y, r):
    """f(y) = -d/dy log gennorm(r).pdf(y)."""
    return -_d_dy(lambda t: stats.gennorm.logpdf(t, r), y)


def _component_logpdf(b, al, mu, be, r):
    """log alpha + log density of b under component (mu, beta, rho): the
    generalized Gaussian with scale beta^-1/2 (scipy.stats.gennorm)."""
    return np.log(al) + stats.gennorm.logpdf(b, r, loc=mu, scale=1.0 / np.sqrt(be))


def _library_estep(x, ldetS, st):
    """Responsibilities and every statistic of the E-step from library
    densities and finite-difference scores (no oracle formula)."""
    A, c, alpha, mu, beta, rho, gm = (st[k] for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm"))
    M, m, n = alpha.shape
    per, L = [], []
    for h in range(M):
        W = np.linalg.inv(A[h])
        b = (x - c[h]) @ W.T                                               # (T, n)
        lp = _component_logpdf(b[:, None, :], alpha[h][None], mu[h][None], beta[h][None], rho[h][None])
        q = special.logsumexp(lp, axis=1)                                   # (T, n)
        per.append(dict(W=W, b=b, z=np.exp(lp - q[:, None, :])))
        L.append(np.log(gm[h]) + np.linalg.slogdet(W)[1] + ldetS + q.sum(axis=1))
    L = np.array(L)
    P = special.logsumexp(L, axis=0)
    out = []
    for h in range(M):
        d = per[h]
        v = np.exp(L[h] - P)
        b = d["b"]
        u = v[:, None, None] * d["z"]                                       # (T, m, n)
        sb = np.sqrt(beta[h])
        y = sb[None] * (b[:, None, :] - mu[h][None])
        f = _score(y, rho[h][None])
        yr = stats.gennorm.logpdf(0.0, rho[h][None]) - stats.gennorm.logpdf(y, rho[h][None])  # |y|^rho

        def mix_logpdf(bb, h=h):
            return special.logsumexp(_component_logpdf(bb[:, None, :], alpha[h][None], mu[h][None],
                                                       beta[h][None], rho[h][None]), axis=1)
        # mixture score in b, with a step below the distance to every component centre
        gap = np.min(np.abs(b[:, None, :] - mu[h][None]), axis=1)
        hb = 1e-3 * gap
        d1 = (mix_logpdf(b + hb) - mix_logpdf(b - hb)) / (2 * hb)
        d2 = (mix_logpdf(b + 2 * hb) - mix_logpdf(b - 2 * hb)) / (4 * hb)
        g = -v[:, None] * (4 * d1 - d2) / 3
        s = dict(v=v, u=u, y=y, f=f, b=b, W=d["W"])
        s["vsum"] = v.sum()
        s["usum"] = u.sum(axis=0)
        s["kappa_numer"] = beta[h] * (u * f * f).sum(axis=0)
        s["lambda_numer"] = (u * (f * y - 1.0) ** 2).sum(axis=0)
        s["mu_numer"] = (u * f).sum(axis=0)
        s["mu_denom"] = sb * (u * f / y).sum(axis=0)
        s["beta_numer"] = s["usum"]
        s["beta_denom"] = (u * f * y).sum(axis=0)
        s["rho_numer"] = (u * yr * np.log(yr)).sum(axis=0)
        s["sigma2_numer"] = (v[:, None] * b * b).sum(axis=0)
        s["cnew"] = (v[:, None] * x).sum(axis=0)
        s["phi"] = g.T @ b
        out.append(s)
    return out, P


def _library_ll(x, ldetS, st):
    return _library_estep_ll_only(x, ldetS, st)


def _library_estep_ll_only(x, ldetS, st):
    A, c, alpha, mu, beta, rho, gm = (st[k] for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm"))
    L = []
    for h in range(alpha.shape[0]):
        W = np.linalg.inv(A[h])
        b = (x - c[h]) @ W.T
        lp = _component_logpdf(b[:, None, :], alpha[h][None], mu[h][None], beta[h][None], rho[h][None])
        L.append(np.log(gm[h]) + np.linalg.slogdet(W)[1] + ldetS + special.logsumexp(lp, axis=1).sum(axis=1))
    return special.logsumexp(np.array(L), axis=0).sum()


def _rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(np.asarray(b)), 1e-300)


def _oracle_ll(x, ldetS, flat, n, M, m):
    st = po.estep(x, ldetS, flat["A"], flat["c"], flat["alpha"], flat["mu"], flat["beta"], flat["rho"],
                  flat["gm"], M, m)
    return st[-1]


# ------------------------------------------------------------------ tests --
@pytest.mark.parametrize("plant", [None, "kappa_numer_no_beta", "lambda_numer_plus",
                                   "mu_denom_no_sqrt_beta", "rho_numer_no_log", "sigma2_from_x",
                                   "cnew_no_v", "g_no_sqrt_beta"])
def test_estep_statistics_from_library_densities(plant):
    """Every E-step statistic (kernels.cu:473-827, amica.cu:1280-1550)
    equals its definition evaluated with scipy's generalized-Gaussian
    densities and finite-difference scores: usum = sum u,
    kappa_numer = beta sum u f^2, lambda_numer = sum u (f y - 1)^2,
    mu_numer = sum u f, mu_denom = sqrt(beta) sum u f / y,
    beta_denom = sum u f y, rho_numer = sum u |y|^rho log |y|^rho,
    sigma2_numer = sum v b^2, cnew = sum v x, phi = sum g b^T with g the
    model posterior times the mixture score of b; and the LL sum."""
    n, T, M, m = 3, 4000, 2, 3
    x = _data(n, T, seed=1)
    st = _state(n, M, m, seed=2)
    ldetS = 0.37
    lib, P = _library_estep(x, ldetS, st)
    with planted(plant):
        ora = po.estep(x, ldetS, **{k: v for k, v in _flat(st).items()}, M=M, m=m)
    views, oll = po.stats_views(ora, n, m, M)
    errs = {}
    for h in range(M):
        for k in STAT_NAMES:
            errs[k] = max(errs.get(k, 0.0), _rel(views[h][k], lib[h][k]))
        errs["vsum"] = max(errs.get("vsum", 0.0), abs(views[h]["vsum"] - lib[h]["vsum"]) / lib[h]["vsum"])
    errs["ll"] = abs(oll - P.sum()) / abs(P.sum())
    print({k: f"{v:.1e}" for k, v in errs.items()})
    _expect(all(v < 1e-6 for v in errs.values()), plant)


@pytest.mark.parametrize("plant", [None, "beta_sqrt_ratio", "mu_step_sign", "alpha_over_t", "cnew_no_v"])
def test_gaussian_components_take_the_textbook_em_step(plant):
    """With rho = 2 every component is Gaussian (variance 1/(2 beta)) and
    one M-step (amica.cu:1995-2137) is the textbook EM step of a Gaussian
    mixture per source: gm = mean v, alpha = sum u / sum v,
    mu = the u-weighted mean of b, c = the v-weighted mean of x, and beta
    the maximiser (scipy.optimize) of sum u log N(b; mu_old, 1/(2 beta))."""
    n, T, M, m = 3, 3000, 2, 2
    x = _data(n, T, seed=3, src_rho=1.0)
    st = _state(n, M, m, seed=4, rho=(2.0, 2.0))
    ldetS = -0.2
    lib, _ = _library_estep(x, ldetS, st)
    cfg = _cfg(n, T, M, m)
    with planted(plant):
        new, _, _, _ = po.step(cfg, 1, x, ldetS, _flat(st), 0.1)
    nw = _np(new, n, M, m)
    errs = {"gm": _rel(nw["gm"], [lib[h]["vsum"] / T for h in range(M)])}
    for h in range(M):
        L = lib[h]
        errs["alpha"] = max(errs.get("alpha", 0), _rel(nw["alpha"][h], L["usum"] / L["vsum"]))
        mu_em = (L["u"] * L["b"][:, None, :]).sum(0) / L["usum"]
        errs["mu"] = max(errs.get("mu", 0), _rel(nw["mu"][h], mu_em))
        errs["c"] = max(errs.get("c", 0), _rel(nw["c"][h], L["cnew"] / L["vsum"]))
        bstar = np.zeros((m, n))
        for j in range(m):
            for i in range(n):
                w, bb, mu0 = L["u"][:, j, i], L["b"][:, i], st["mu"][h, j, i]
                obj = lambda s: -np.sum(w * stats.norm.logpdf(bb, mu0, np.sqrt(0.5 / np.exp(s))))  # noqa: E731
                bstar[j, i] = np.exp(optimize.minimize_scalar(obj, bracket=(-3, 3), tol=1e-12).x)
        errs["beta"] = max(errs.get("beta", 0), _rel(nw["beta"][h], bstar))
    print({k: f"{v:.1e}" for k, v in errs.items()})
    _expect(all(v < 1e-7 for v in errs.values()), plant)


@pytest.mark.parametrize("plant", [None, "beta_sqrt_ratio", "mu_step_sign"])
def test_mu_beta_steps_for_general_shapes(plant):
    """For rho in (1, 2): the mu step (amica.cu:2095-2115) is the iteratively
    reweighted least-squares step (Lawson) for the u-weighted L_rho location,
    mu_new = sum u w b / sum u w with w = |b - mu|^(rho-2), and it does not
    increase sum u |sqrt(beta) (b - mu)|^rho; the beta step
    (amica.cu:2117-2137) satisfies beta_new / beta = (beta* / beta)^(rho/2),
    where beta* maximises (scipy.optimize) sum u log p(b; mu_old, beta, rho)."""
    n, T, M, m = 3, 3000, 1, 2
    x = _data(n, T, seed=5)
    st = _state(n, M, m, seed=6)
    ldetS = 0.0
    lib, _ = _library_estep(x, ldetS, st)[0], None
    cfg = _cfg(n, T, M, m)
    with planted(plant):
        new, _, _, _ = po.step(cfg, 1, x, ldetS, _flat(st), 0.1)
    nw = _np(new, n, M, m)
    L = lib[0]
    u, b = L["u"], L["b"]
    errs, descent = {}, True
    mu0, be0, r0 = st["mu"][0], st["beta"][0], st["rho"][0]
    w = np.abs(b[:, None, :] - mu0[None]) ** (r0[None] - 2.0)
    errs["mu_irls"] = _rel(nw["mu"][0], (u * w * b[:, None, :]).sum(0) / (u * w).sum(0))
    for j in range(m):
        for i in range(n):
            J = lambda mu: np.sum(u[:, j, i] * np.abs(np.sqrt(be0[j, i]) * (b[:, i] - mu)) ** r0[j, i])  # noqa: E731
            descent &= J(nw["mu"][0, j, i]) <= J(mu0[j, i]) * (1 + 1e-12)
    lhs, rhs = np.zeros((m, n)), np.zeros((m, n))
    for j in range(m):
        for i in range(n):
            ww, bb, mu, r = u[:, j, i], b[:, i], mu0[j, i], r0[j, i]
            obj = lambda s: -np.sum(ww * stats.gennorm.logpdf(bb, r, loc=mu, scale=np.exp(-0.5 * s)))  # noqa: E731
            bstar = np.exp(optimize.minimize_scalar(obj, bracket=(-3, 3), tol=1e-12).x)
            lhs[j, i] = np.log(nw["beta"][0, j, i] / be0[j, i])
            rhs[j, i] = 0.5 * r * np.log(bstar / be0[j, i])
    errs["beta_ratio"] = np.max(np.abs(lhs - rhs))
    print({k: f"{v:.1e}" for k, v in errs.items()}, "descent", descent)
    _expect(errs["mu_irls"] < 1e-9 and errs["beta_ratio"] < 1e-7 and descent, plant)


@pytest.mark.parametrize("plant", [None, "rho_step_sign", "rho_psi_arg", "rho_numer_no_log"])
def test_rho_step_is_the_scaled_likelihood_gradient(plant):
    """updaterho (kernels.cu:906-954): rho_new - rho = rholrate rho^2 /
    (usum psi(1 + 1/rho)) dLL/drho, with dLL/drho a finite difference of
    the library log-likelihood and psi scipy's digamma (clamp bounds
    widened so that no step is clipped)."""
    n, T, M, m = 2, 3000, 2, 2
    x = _data(n, T, seed=7)
    st = _state(n, M, m, seed=8, rho=(1.2, 1.8))
    ldetS = 0.1
    lib, _ = _library_estep(x, ldetS, st)
    cfg = _cfg(n, T, M, m, update_rho=1, rho_start_iter=1, rhomin=-1e9, rhomax=1e9, rholrate=0.1)
    with planted(plant):
        new, _, _, _ = po.step(cfg, 1, x, ldetS, _flat(st), 0.1)
    nw = _np(new, n, M, m)
    err = 0.0
    for h in range(M):
        for j in range(m):
            for i in range(n):
                r = st["rho"][h, j, i]
                dr = 1e-4 * r

                def ll_at(rv):
                    s2 = {k: np.array(v, copy=True) for k, v in st.items()}
                    s2["rho"][h, j, i] = rv
                    return _library_estep_ll_only(x, ldetS, s2)
                d1 = (ll_at(r + dr) - ll_at(r - dr)) / (2 * dr)
                d2 = (ll_at(r + 2 * dr) - ll_at(r - 2 * dr)) / (4 * dr)
                grad = (4 * d1 - d2) / 3
                expect = 0.1 * r * r / (lib[h]["usum"][j, i] * special.digamma(1 + 1 / r)) * grad
                err = max(err, abs((nw["rho"][h, j, i] - r) - expect) / max(abs(expect), 1e-3))
    print(f"rho step error {err:.2e}")
    _expect(err < 1e-6, plant)


@pytest.mark.parametrize("plant", [None, "kappa_old_alpha", "lambda_no_mu", "sigma2_from_x",
                                   "kappa_numer_no_beta"])
def test_kappa_lambda_sigma2_are_fisher_terms(plant):
    """The Newton step's curvature terms (kernels.cu:848-867, amica.cu:
    2055-2087).  For one model, one component and mu = 0 they are the
    textbook quantities of Newton ICA: kappa = mean phi(s)^2,
    lambda = mean (phi(s) s - 1)^2, sigma2 = mean s^2, with phi the score
    -d/ds log p(s) of the library density (finite differences).  With
    several models and components they are the posterior expectations
    kappa = sum_t sum_j u beta f^2 / vsum and
    lambda = sum_t sum_j u (mu^2 beta f^2 + (f y - 1)^2) / vsum (the new alpha
    weights usum / vsum); sigma2 = sum v b^2 / vsum."""
    ok = True
    for (n, M, m, seed, mu0) in ((3, 1, 1, 9, True), (3, 2, 3, 10, False)):
        T = 3000
        x = _data(n, T, seed=seed)
        st = _state(n, M, m, seed=seed + 100)
        if mu0:
            st["mu"][:] = 0.0
        ldetS = 0.0
        lib, _ = _library_estep(x, ldetS, st)
        cfg = _cfg(n, T, M, m, update_A=1)
        fl = _flat(st)
        ora = po.estep(x, ldetS, **fl, M=M, m=m)
        W = np.concatenate([np.ascontiguousarray(lib[h]["W"].T).ravel() for h in range(M)])
        with planted(plant):
            ora = po.estep(x, ldetS, **fl, M=M, m=m)
            _, _, _, tr = po.mstep(cfg, 1, ora, W, fl, 0.1, trace=True)
        errs = {}
        for h in range(M):
            L = lib[h]
            kap = tr["kappa"][h * n:(h + 1) * n]
            lam = tr["lambda_"][h * n:(h + 1) * n]
            s2 = tr["sigma2"][h * n:(h + 1) * n]
            if mu0 and m == 1:
                s = L["b"]
                phi = -_d_dy(lambda t: stats.gennorm.logpdf(t, st["rho"][h, 0][None], loc=0.0,
                                                             scale=1 / np.sqrt(st["beta"][h, 0][None])), s)
                ek, el, es = (phi ** 2).mean(0), ((phi * s - 1) ** 2).mean(0), (s ** 2).mean(0)
            else:
                be, mu = st["beta"][h], st["mu"][h]
                ek = (L["u"] * be[None] * L["f"] ** 2).sum((0, 1)) / L["vsum"]
                el = (L["u"] * (mu[None] ** 2 * be[None] * L["f"] ** 2 + (L["f"] * L["y"] - 1) ** 2)
                      ).sum((0, 1)) / L["vsum"]
                es = L["sigma2_numer"] / L["vsum"]
            errs[f"kappa{h}"], errs[f"lambda{h}"], errs[f"sigma2{h}"] = _rel(kap, ek), _rel(lam, el), _rel(s2, es)
        print({k: f"{v:.1e}" for k, v in errs.items()})
        ok &= all(v < 1e-6 for v in errs.values())
    _expect(ok, plant)


def _gaussian_newton_state(x, s_scale):
    """One model, one Gaussian component per source (rho = 2, mu = 0) whose
    beta makes kappa_i sigma2_i = s_scale_i^2 exactly: for a Gaussian
    component kappa = 4 beta^2 E[b^2] and sigma2 = E[b^2]."""
    n = x.shape[1]
    rng = np.random.default_rng(11)
    A = np.eye(n) + 0.1 * rng.normal(size=(n, n))
    b = x @ np.linalg.inv(A).T
    eb2 = (b ** 2).mean(0)
    st = dict(A=A[None], c=np.zeros((1, n)), alpha=np.ones((1, 1, n)), mu=np.zeros((1, 1, n)),
              beta=(np.asarray(s_scale) / (2 * eb2))[None, None], rho=np.full((1, 1, n), 2.0),
              gm=np.ones(1))
    return st


def _grad_W(x, ldetS, st, h=0):
    """G = dLL/dW W^T of the library log-likelihood, by finite differences
    over the entries of W = A^-1 (c fixed)."""
    n = st["A"].shape[1]
    W0 = np.linalg.inv(st["A"][h])
    dW = np.zeros((n, n))
    for i in range(n):
        for k in range(n):
            def ll_at(e):
                Wp = W0.copy()
                Wp[i, k] += e
                s2 = {kk: np.array(v, copy=True) for kk, v in st.items()}
                s2["A"][h] = np.linalg.inv(Wp)
                return _library_estep_ll_only(x, ldetS, s2)
            e = 1e-5
            d1 = (ll_at(e) - ll_at(-e)) / (2 * e)
            d2 = (ll_at(2 * e) - ll_at(-2 * e)) / (4 * e)
            dW[i, k] = (4 * d1 - d2) / 3
    return dW @ W0.T


@pytest.mark.parametrize("plant", [None, "newton_transpose", "bflag_no_diag", "phi_no_norm"])
def test_newton_matrix_and_test(plant):
    """The Newton direction (kernels.cu:961-1014, amica.cu:2156-2279) solves
    the pairwise 2x2 systems of the ICA Hessian: for i != k,
    kappa_i sigma2_k B_ik + B_ki = phi_ik and lambda_i B_ii = 1 - phi_ii,
    where phi = I - G / vsum with G = dLL/dW W^T by finite differences of
    the library log-likelihood; the natural-gradient branch has B = I - phi;
    bflag counts the pairs whose 2x2 block [[k_i s_k, 1], [1, k_k s_i]] is
    not positive definite (numpy eigvalsh) plus the sources with
    kappa sigma2 <= 1."""
    n, T = 3, 3000
    x = _data(n, T, seed=12)
    ok = True
    for scales, newton in (((1.2, 1.4, 1.6), True), ((0.7, 1.3, 1.5), False)):
        st = _gaussian_newton_state(x, scales)
        fl = _flat(st)
        lib, _ = _library_estep(x, 0.0, st)
        L = lib[0]
        cfg = _cfg(n, T, 1, 1, update_A=1, do_newton=1, newt_start_iter=1)
        W = np.ascontiguousarray(L["W"].T).ravel()
        with planted(plant):
            ora = po.estep(x, 0.0, **fl, M=1, m=1)
            _, _, dec, tr = po.mstep(cfg, 1, ora, W, fl, 0.1, trace=True)
        phi = np.eye(n) - _grad_W(x, 0.0, st) / L["vsum"]
        B = tr["B"].reshape(n, n).T
        kap = (L["u"] * st["beta"][0][None] * L["f"] ** 2).sum((0, 1)) / L["vsum"]
        lam = (L["u"] * (L["f"] * L["y"] - 1) ** 2).sum((0, 1)) / L["vsum"]
        s2 = L["sigma2_numer"] / L["vsum"]
        ks = kap * s2
        nonpd = sum(1 for i in range(n) for k in range(n) if i != k and
                    np.linalg.eigvalsh(np.array([[kap[i] * s2[k], 1.0], [1.0, kap[k] * s2[i]]])).min() <= 0)
        bflag = nonpd + int((ks <= 1.0).sum())
        ok &= int(dec["bflag"][0]) == bflag and int(dec["newton"][0]) == int(newton) and (bflag == 0) == newton
        if newton:
            err = 0.0
            for i in range(n):
                for k in range(n):
                    if i != k:
                        err = max(err, abs(kap[i] * s2[k] * B[i, k] + B[k, i] - phi[i, k]))
                err = max(err, abs(lam[i] * B[i, i] - (1 - phi[i, i])))
        else:
            err = np.max(np.abs(B - (np.eye(n) - phi)))
        print(f"newton={newton} bflag oracle {dec['bflag'][0]} library {bflag}; B error {err:.1e}")
        ok &= err < 1e-6
    _expect(ok, plant)


@pytest.mark.parametrize("plant", [None, "ab_order", "natgrad_sign"])
def test_a_update_moves_along_the_chosen_direction(plant):
    """The A update (amica.cu:2287-2364): A - lrate A B (natural gradient) or
    A + lrate A B (Newton) changes the log-likelihood at the rate
    <G, B> resp. -<G, B>, G = dLL/dW W^T by finite differences of the
    library log-likelihood; measured by finite differences in lrate
    (Richardson, lrate 1e-4 and 2e-4).  On the natural gradient the rate
    is vsum |I - phi|^2 >= 0: the likelihood does not decrease."""
    n, T = 3, 3000
    x = _data(n, T, seed=13)
    ok = True
    for scales, newton in (((1.2, 1.4, 1.6), True), ((0.7, 1.3, 1.5), False)):
        st = _gaussian_newton_state(x, scales)
        fl = _flat(st)
        G = _grad_W(x, 0.0, st)
        ll0 = _oracle_ll(x, 0.0, fl, n, 1, 1)
        D = []
        for lr in (1e-4, 2e-4):
            cfg = _cfg(n, T, 1, 1, update_A=1, do_newton=1 if newton else 0, newt_start_iter=1,
                       lrate0=lr, lratemax=lr, lnatrate=lr, update_mu=0, update_beta=0,
                       update_alpha=0, update_gm=0)
            with planted(plant):
                ora = po.estep(x, 0.0, **fl, M=1, m=1)
                W = np.ascontiguousarray(np.linalg.inv(st["A"][0]).T).ravel()
                new, lr_out, dec, tr = po.mstep(cfg, 1, ora, W, fl, lr, trace=True)
            assert abs(lr_out - lr) < 1e-15 and int(dec["newton"][0]) == int(newton)
            D.append((_oracle_ll(x, 0.0, new, n, 1, 1) - ll0) / lr)
            B = tr["B"].reshape(n, n).T
        rate = 2 * D[0] - D[1]
        expect = (-1.0 if newton else 1.0) * np.sum(G * B)
        err = abs(rate - expect) / abs(expect)
        print(f"newton={newton}: dLL/dlrate {rate:.6e}, expected {expect:.6e} (rel {err:.1e})")
        ok &= err < 1e-4
        if not newton:
            ok &= rate > 0
    _expect(ok, plant)


@pytest.mark.parametrize("plant", [None, "lrate_growth", "lldec_no_halve"])
def test_learning_rate_rules_known_answers(plant):
    """The lrate rules (amica.cu:1985-1993, 2281-2285) on hand-computed
    cases: a decrease of the LL halves lrate (lratefact 0.5) before the
    growth lrate + min(0.1, lrate), capped by lratemax when every model took
    the Newton step and by lnatrate otherwise."""
    n, T = 3, 3000
    x = _data(n, T, seed=14)
    cases = [  # (Newton state, iteration, LL history, lrate in, expected lrate out, expected lldec)
        (False, 1, (0.0, -1.0), 0.1, 0.1, 0),
        (False, 3, (-1.0, -1.1), 0.1, 0.1, 1),    # 0.05 -> min(0.1, 0.05 + 0.05)
        (False, 3, (-1.1, -1.0), 0.02, 0.04, 0),
        (False, 3, (-1.0, -1.1), 0.02, 0.02, 1),  # 0.01 -> 0.02
        (True, 3, (-1.1, -1.0), 0.3, 0.4, 0),     # Newton: min(1, 0.3 + 0.1)
        (True, 3, (-1.0, -1.1), 0.3, 0.25, 1),    # 0.15 -> 0.15 + 0.1
        (True, 3, (-1.1, -1.0), 0.05, 0.1, 0),
        (True, 3, (-1.1, -1.0), 0.95, 1.0, 0),    # capped by lratemax
    ]
    ok = True
    with planted(plant):
        for newton, it, hist, lr_in, lr_out, lldec in cases:
            st = _gaussian_newton_state(x, (1.2, 1.4, 1.6) if newton else (0.7, 1.3, 1.5))
            fl = _flat(st)
            cfg = _cfg(n, T, 1, 1, update_A=1, do_newton=1, newt_start_iter=1)
            ora = po.estep(x, 0.0, **fl, M=1, m=1)
            W = np.ascontiguousarray(np.linalg.inv(st["A"][0]).T).ravel()
            _, got, dec, _ = po.mstep(cfg, it, ora, W, fl, lr_in, ll_prev=hist[0], ll_cur=hist[1])
            good = abs(got - lr_out) < 1e-15 and dec["lldec"] == lldec and int(dec["newton"][0]) == int(newton)
            print(f"newton={newton} it={it} hist={hist} lrate {lr_in} -> {got} (expected {lr_out})")
            ok &= good
        # before newt_start_iter the natural-gradient cap applies
        cfg = _cfg(n, T, 1, 1, update_A=1, do_newton=1, newt_start_iter=5)
        st = _gaussian_newton_state(x, (1.2, 1.4, 1.6))
        fl = _flat(st)
        ora = po.estep(x, 0.0, **fl, M=1, m=1)
        W = np.ascontiguousarray(np.linalg.inv(st["A"][0]).T).ravel()
        _, got, dec, _ = po.mstep(cfg, 3, ora, W, fl, 0.3, ll_prev=-1.1, ll_cur=-1.0)
        ok &= abs(got - 0.1) < 1e-15 and int(dec["newton"][0]) == 0
    _expect(ok, plant)


def test_loglikelihood_invariant_under_column_scaling():
    """The E-step's LL (amica.cu:1009-1606) is unchanged when A's columns
    are scaled by 1/nu, mu by nu and beta by 1/nu^2 (the rescaling the
    reparameterisation applies, amica.cu:2367-2420)."""
    n, T, M, m = 3, 2000, 2, 3
    x = _data(n, T, seed=15)
    st = _state(n, M, m, seed=16)
    nu = np.random.default_rng(17).uniform(0.5, 2.0, size=(M, n))
    sc = {k: np.array(v, copy=True) for k, v in st.items()}
    sc["A"] = st["A"] / nu[:, None, :]
    sc["mu"] = st["mu"] * nu[:, None, :]
    sc["beta"] = st["beta"] / nu[:, None, :] ** 2
    a = _oracle_ll(x, 0.2, _flat(st), n, M, m)
    b = _oracle_ll(x, 0.2, _flat(sc), n, M, m)
    assert abs(a - b) <= 1e-12 * abs(a)
    assert abs(a - _library_estep_ll_only(x, 0.2, st)) <= 1e-12 * abs(a)


@pytest.mark.parametrize("plant", [None, "reparam_beta_nu", "mu_shift_wt"])
def test_reparameterisation_keeps_the_likelihood(plant):
    """Reparameterisation (amica.cu:2367-2462) with lrate 0 and the other
    updates off: (a) one model, c = 0: A's columns come out with unit norm
    and the LL is unchanged; (b) two models with unit-norm columns, c away
    from cnew: mu -= W (cnew - c) with c <- cnew leaves the LL unchanged."""
    n, T = 3, 2000
    x = _data(n, T, seed=18)
    ok = True
    for M, normalise in ((1, False), (2, True)):
        m = 2
        st = _state(n, M, m, seed=19 + M, c_scale=0.0 if M == 1 else 0.5)
        if normalise:
            st["A"] = st["A"] / np.sqrt((st["A"] ** 2).sum(axis=1, keepdims=True))
        fl = _flat(st)
        cfg = _cfg(n, T, M, m, update_A=1, do_reparm=1, lrate0=0.0, lnatrate=0.0, lratemax=0.0,
                   update_mu=0, update_beta=0, update_alpha=0, update_gm=0)
        with planted(plant):
            new, _, lr, _ = po.step(cfg, 1, x, 0.0, fl, 0.0)
        a = _oracle_ll(x, 0.0, fl, n, M, m)
        b = _oracle_ll(x, 0.0, new, n, M, m)
        An = _np(new, n, M, m)["A"]
        norms = np.sqrt((An ** 2).sum(axis=1))
        print(f"M={M}: LL {a:.15e} -> {b:.15e}; column norms within {np.max(np.abs(norms - 1)):.1e} of 1")
        ok &= abs(a - b) <= 1e-11 * abs(a) and np.max(np.abs(norms - 1)) < 1e-13 and lr == 0.0
        if M == 2:
            ok &= np.max(np.abs(np.asarray(new["c"]) - np.asarray(fl["c"]))) > 1e-3   # c did move
    _expect(ok, plant)


@pytest.mark.parametrize("plant", [None, "phi_no_norm"])
def test_phi_identity_is_a_fixed_point(plant):
    """When sum_t g b^T = vsum I (phi = I after the normalisation) both the
    Newton and the natural-gradient direction vanish and A is unchanged
    (amica.cu:2156-2364, do_reparm off)."""
    n, T = 3, 3000
    x = _data(n, T, seed=20)
    ok = True
    for scales, newton in (((1.2, 1.4, 1.6), True), ((0.7, 1.3, 1.5), False)):
        st = _gaussian_newton_state(x, scales)
        fl = _flat(st)
        ora = po.estep(x, 0.0, **fl, M=1, m=1)
        v, _ = po.stats_views(ora, n, 1, 1)
        stride = po.stats_stride(n, 1)
        phi_off = stride - n * n
        ora[phi_off:stride] = (v[0]["vsum"] * np.eye(n)).T.ravel()
        cfg = _cfg(n, T, 1, 1, update_A=1, do_newton=1, newt_start_iter=1)
        W = np.ascontiguousarray(np.linalg.inv(st["A"][0]).T).ravel()
        with planted(plant):
            new, _, dec, _ = po.mstep(cfg, 1, ora, W, fl, 0.3)
        ok &= int(dec["newton"][0]) == int(newton)
        ok &= np.max(np.abs(np.asarray(new["A"]) - fl["A"])) < 1e-14
    _expect(ok, plant)
