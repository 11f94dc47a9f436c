"""Independent numpy restatement of the AMICA EM iteration (TEST INFRASTRUCTURE ONLY).

A second, vectorised FP64 restatement of the reference's equations, written
separately from oracle.cpp (numpy.linalg for pinv / slogdet / eigh,
scipy.special for gammaln / digamma / logsumexp) so that the two restatements
cross-check each other on tiny inputs.  Follows fraimondo/cudamica:
  preprocessing   src/preprocess.cu:31-420
  E-step          src/amica.cu:996-1611, src/kernels.cu:251-827
  M-step, Newton  src/amica.cu:1963-2447, src/kernels.cu:848-1065
Arrays use natural numpy shapes: A (M, n, n) with A[h][i, k]; c (M, n);
alpha/mu/beta/rho (M, m, n) indexed [h, j, i]; x (T, n).
Only tests/ import this module.
"""
import numpy as np
from scipy.special import digamma, gammaln, logsumexp

LOG2_REF = 0.693147180559945  # config.h:25


def preprocess(x_raw, do_sphere=1):
    x = np.asarray(x_raw, np.float64)
    T, n = x.shape
    mean = x.mean(axis=0)
    xc = x - mean
    if do_sphere == 1:
        C = xc.T @ xc / T
        d, V = np.linalg.eigh(C)
        S = V @ np.diag(2.0 / np.sqrt(d)) @ V.T
    elif do_sphere == 2:
        mn = xc.mean(axis=0)
        S = np.diag(1.0 / np.sqrt(((xc - mn) ** 2).sum(0) / (T - 1)))
    else:
        S = np.eye(n)
    ldetS = -np.linalg.slogdet(S)[1]
    return xc @ S.T, S, mean, ldetS


def estep(x, ldetS, A, c, alpha, mu, beta, rho, gm, update_rho=True):
    T, n = x.shape
    M, m, _ = alpha.shape
    W = np.stack([np.linalg.pinv(A[h]) for h in range(M)])
    ldet = np.array([np.linalg.slogdet(W[h])[1] for h in range(M)])
    b = np.stack([x @ W[h].T - W[h] @ c[h] for h in range(M)])          # (M, T, n)
    y = np.sqrt(beta)[:, None] * (b[:, :, None, :] - mu[:, None])        # (M, T, m, n)
    ay = np.abs(y)
    yr = ay ** rho[:, None]
    Q = (np.log(alpha) + 0.5 * np.log(beta) - LOG2_REF - gammaln(1.0 + 1.0 / rho))[:, None] - yr
    q = logsumexp(Q, axis=2)                                             # (M, T, n)
    z = np.exp(Q - q[:, :, None, :])
    L = np.log(gm)[:, None] + ldet[:, None] + ldetS + q.sum(axis=2)     # (M, T)
    P = logsumexp(L, axis=0)                                             # (T,)
    v = np.exp(L - P[None])
    u = v[:, :, None, None] * z
    fp = rho[:, None] * np.sign(y) * ay ** (rho[:, None] - 1.0)
    ufp = u * fp
    g = (np.sqrt(beta)[:, None] * ufp).sum(axis=2)                       # (M, T, n)
    low = (rho <= 2.0)
    out = []
    for h in range(M):
        d = {}
        d["vsum"] = v[h].sum()
        d["usum"] = u[h].sum(0)
        d["kappa_numer"] = beta[h] * (ufp[h] * fp[h]).sum(0)
        d["lambda_numer"] = (u[h] * (fp[h] * y[h] - 1.0) ** 2).sum(0)
        d["mu_numer"] = ufp[h].sum(0)
        md_lo = (ufp[h] / y[h]).sum(0)
        md_hi = (ufp[h] * fp[h]).sum(0)
        d["mu_denom"] = np.sqrt(beta[h]) * np.where(low[h], md_lo, md_hi)
        bd_lo = (ufp[h] * y[h]).sum(0)
        bd_hi = rho[h] * (u[h] * yr[h]).sum(0)
        d["beta_denom"] = np.where(low[h], bd_lo, bd_hi)
        d["beta_numer"] = np.where(low[h], d["usum"], 0.0)
        if update_rho:
            with np.errstate(divide="ignore", invalid="ignore"):
                t = np.where(yr[h] > 0, yr[h] * np.log(yr[h]), 0.0)
            d["rho_numer"] = (u[h] * t).sum(0)
        else:
            d["rho_numer"] = np.zeros((m, n))
        d["sigma2_numer"] = (v[h][:, None] * b[h] ** 2).sum(0)
        d["cnew"] = (v[h][:, None] * x).sum(0)
        d["phi"] = g[h].T @ b[h]
        out.append(d)
    return out, P.sum(), dict(W=W, ldet=ldet, L=L, v=v)


def mstep(cfg, it, stats, st, lrate, numdec, ll_hist):
    """One M-step; st is a dict of arrays modified in place.  Returns (lrate, numdec, log)."""
    n, M, m, T = cfg["n"], cfg["M"], cfg["m"], cfg["T"]
    log = {"lldec": 0, "bflag": [0] * M, "newton": [0] * M}
    if it > 1 and ll_hist[-1] < ll_hist[-2]:
        log["lldec"] = 1
        lrate *= cfg["lratefact"]
        numdec += 1
        if numdec > cfg["maxdec"]:
            numdec = 0
    A, c, alpha, mu, beta, rho, gm = (st[k] for k in ("A", "c", "alpha", "mu", "beta", "rho", "gm"))
    W = st["W"]
    vsum = np.array([s["vsum"] for s in stats])
    gm[:] = vsum / T
    for h in range(M):
        alpha[h] = stats[h]["usum"] / vsum[h]
    sigma2 = np.stack([s["sigma2_numer"] for s in stats]) / vsum[:, None]
    cnew = np.stack([s["cnew"] for s in stats]) / vsum[:, None]
    phi = np.stack([s["phi"] for s in stats]) / vsum[:, None, None]
    usum = np.stack([s["usum"] for s in stats])
    dk = np.stack([s["kappa_numer"] for s in stats]) / usum
    lam_n = np.stack([s["lambda_numer"] for s in stats]) / usum
    kappa = (alpha * dk).sum(1)
    lam = (alpha * (mu ** 2 * dk + lam_n)).sum(1)
    mu += np.stack([s["mu_numer"] / s["mu_denom"] for s in stats])
    beta *= np.stack([s["beta_numer"] / s["beta_denom"] for s in stats])
    if cfg["update_rho"] and it >= cfg["rho_start_iter"]:
        rn = np.stack([s["rho_numer"] for s in stats])
        r = rho + cfg["rholrate"] * (1.0 - (rho / digamma(1.0 + 1.0 / rho)) * rn / usum)
        rho[:] = np.clip(r, cfg["rhomin"], cfg["rhomax"])
    no_newt = 0
    Bs = []
    for h in range(M):
        r = kappa[h] * sigma2[h]
        D = np.outer(r, r)
        bflag = int((D <= 1.0).sum())
        log["bflag"][h] = bflag
        no_newt |= bflag > 0
        newt = bflag == 0 and cfg["do_newton"] == 1 and it >= cfg["newt_start_iter"]
        log["newton"][h] = int(newt)
        if newt:
            B = (np.outer(sigma2[h], kappa[h]) * phi[h] - phi[h].T) / (D - 1.0)
            B[np.diag_indices(n)] = (1.0 - np.diag(phi[h])) / lam[h]
        else:
            B = np.eye(n) - phi[h]
        Bs.append(B)
    if it >= cfg["newt_start_iter"] and cfg["do_newton"] and not no_newt:
        lrate = min(cfg["lratemax"], lrate + min(0.1, lrate))
    else:
        lrate = min(cfg["lnatrate"], lrate + min(0.1, lrate))
    for h in range(M):
        f = lrate if log["newton"][h] else -lrate
        A[h] = A[h] + f * A[h] @ Bs[h]
    for h in range(M):
        nu = np.sqrt((A[h] ** 2).sum(0))
        A[h] /= nu[None]
        mu[h] *= nu[None]
        beta[h] /= nu[None] ** 2
    for h in range(M):
        mu[h] -= (W[h] @ (cnew[h] - c[h]))[None]
    c[:] = cnew
    return lrate, numdec, log


def run(cfg, x, ldetS, state, maxiter):
    """cfg: dict with n, M, m, T and the M-step options; state: dict of (M,...) arrays."""
    st = {k: np.array(v, np.float64, copy=True) for k, v in state.items()}
    lrate, numdec = cfg["lrate0"], 0
    lls, lrates, logs = [], [], []
    for it in range(1, maxiter + 1):
        stats, psum, aux = estep(x, ldetS, st["A"], st["c"], st["alpha"], st["mu"], st["beta"],
                                 st["rho"], st["gm"], cfg["update_rho"])
        st["W"] = aux["W"]
        lls.append(psum / (cfg["n"] * cfg["T"]))
        lrates.append(lrate)
        lrate, numdec, log = mstep(cfg, it, stats, st, lrate, numdec, lls)
        logs.append(log)
    return st, np.array(lls), np.array(lrates), logs


def default_cfg(n, T, M, m, **kw):
    cfg = dict(n=n, T=T, M=M, m=m, do_newton=1, update_rho=1, rho_start_iter=1, newt_start_iter=50,
               lrate0=0.1, lratemax=1.0, lnatrate=0.1, lratefact=0.5, rholrate=0.1, rhomin=1.0,
               rhomax=2.0, maxdec=3)
    cfg.update(kw)
    return cfg
