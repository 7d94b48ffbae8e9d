"""Restated Krylov solvers in numpy -- TEST INFRASTRUCTURE ONLY (the oracle).

Follows the reference's algorithms step by step so that iteration counts and
convergence histories of the GPU solvers can be compared:
  gmres_mgs ... source/numerical_algos/lin_solvers/gmres.h:376-494 (restarted,
                modified Gram-Schmidt, Givens rotations, check at restarts)
  bicgstabl ... source/numerical_algos/lin_solvers/bicgstabl.h:151-274
  bicgstab .... source/numerical_algos/lin_solvers/bicgstab.h:134-253 (no
                preconditioner; residual recomputed every recalc steps)
Dots are plain numpy (np.dot, pairwise order); these are references for the
iteration behaviour, not for bits.  Pinned in tests/test_oracle_krylov.py against
scipy's direct solver on small nonsymmetric problems.
"""
import numpy as np


def _rot(dx, dy):
    if dy == 0.0:
        return 1.0, 0.0
    if abs(dy) > abs(dx):
        t = dx / dy
        sn = 1.0 / np.sqrt(1.0 + t * t)
        return t * sn, sn
    t = dy / dx
    cs = 1.0 / np.sqrt(1.0 + t * t)
    return cs, t * cs


def gmres_mgs(A, b, x0, m, rel_tol, max_iters, early_exit=False):
    """Returns x, iterations, [(iteration, ||r||/||b||) at every check]."""
    x = x0.copy()
    bn = np.linalg.norm(b)
    tol = rel_tol * bn
    r = b - A @ x
    it = 0
    hist = [(0, np.linalg.norm(r) / bn)]
    while not (np.linalg.norm(r) <= tol or it >= max_iters):
        beta = np.linalg.norm(r)
        V = [r / beta]
        H = np.zeros((m + 1, m))
        cs = np.zeros(m)
        sn = np.zeros(m)
        s = np.zeros(m + 1)
        s[0] = beta
        i = -1
        while True:
            i += 1
            it += 1
            w = A @ V[i]
            for k in range(i + 1):
                H[k, i] = np.dot(V[k], w)
                w = w - H[k, i] * V[k]
            h = np.linalg.norm(w)
            H[i + 1, i] = h
            V.append(w / h if h != 0 else w)
            for k in range(i):
                t = cs[k] * H[k, i] + sn[k] * H[k + 1, i]
                H[k + 1, i] = -sn[k] * H[k, i] + cs[k] * H[k + 1, i]
                H[k, i] = t
            cs[i], sn[i] = _rot(H[i, i], H[i + 1, i])
            H[i, i] = cs[i] * H[i, i] + sn[i] * H[i + 1, i]
            H[i + 1, i] = 0.0
            t = cs[i] * s[i] + sn[i] * s[i + 1]
            s[i + 1] = -sn[i] * s[i] + cs[i] * s[i + 1]
            s[i] = t
            if h == 0 or i + 1 >= m or it >= max_iters or (early_exit and abs(s[i + 1]) <= tol):
                break
        y = np.zeros(i + 1)
        for j in range(i, -1, -1):
            y[j] = (s[j] - np.dot(H[j, j + 1:i + 1], y[j + 1:i + 1])) / H[j, j]
        for j in range(i + 1):
            x = x + y[j] * V[j]
        r = b - A @ x
        hist.append((it, np.linalg.norm(r) / bn))
    return x, it, hist


def bicgstabl(A, b, x0, L, rel_tol, max_iters, recalc=1):
    x = x0.copy()
    bn = np.linalg.norm(b)
    tol = rel_tol * bn
    r = [np.zeros_like(b) for _ in range(L + 1)]
    u = [np.zeros_like(b) for _ in range(L + 1)]
    r[0] = b - A @ x
    rt = r[0] / np.linalg.norm(r[0])
    rho, alpha, omega = 1.0, 0.0, 1.0
    it = 0
    hist = []
    tau = np.zeros((L + 1, L + 1))
    sigma = np.zeros(L + 1)
    g = np.zeros(L + 1)
    gp = np.zeros(L + 1)
    gpp = np.zeros(L + 1)
    while True:
        rn = np.linalg.norm(r[0])
        hist.append((it, rn / bn))
        if rn <= tol or it >= max_iters:
            break
        rho = -omega * rho
        for j in range(L):
            rho1 = np.dot(r[j], rt)
            beta = alpha * rho1 / rho
            rho = rho1
            for i in range(j + 1):
                u[i] = r[i] - beta * u[i]
            u[j + 1] = A @ u[j]
            alpha = rho / np.dot(u[j + 1], rt)
            for i in range(j + 1):
                r[i] = r[i] - alpha * u[i + 1]
            r[j + 1] = A @ r[j]
            x = x + alpha * u[0]
        for j in range(1, L + 1):
            for i in range(1, j):
                tau[i, j] = np.dot(r[j], r[i]) / sigma[i]
                r[j] = r[j] - tau[i, j] * r[i]
            sigma[j] = np.dot(r[j], r[j])
            gp[j] = np.dot(r[0], r[j]) / sigma[j]
        g[L] = gp[L]
        omega = g[L]
        for j in range(L - 1, 0, -1):
            g[j] = gp[j] - sum(tau[j, i] * g[i] for i in range(j + 1, L + 1))
        for j in range(1, L):
            gpp[j] = g[j + 1] + sum(tau[j, i] * g[i + 1] for i in range(j + 1, L))
        x = x + g[1] * r[0]
        r[0] = r[0] - gp[L] * r[L]
        u[0] = u[0] - g[L] * u[L]
        for j in range(1, L):
            x = x + gpp[j] * r[j]
            r[0] = r[0] - gp[j] * r[j]
            u[0] = u[0] - g[j] * u[j]
        if recalc and ((it // L + 1) % recalc == 0):
            r[0] = b - A @ x
        it += L
    return x, it, hist


def bicgstab(A, b, x0, rel_tol, max_iters, recalc=0):
    """Plain BiCGStab; returns x, iterations, [(iteration, ||r||/||b||)]."""
    x = x0.copy()
    bn = np.linalg.norm(b)
    tol = rel_tol * bn
    r = b - A @ x
    rs = r.copy()
    rho_prev = alpha = omega_prev = 1.0
    nu = np.zeros_like(b)
    p = np.zeros_like(b)
    it = 0
    hist = []
    while True:
        rn = np.linalg.norm(r)
        hist.append((it, rn / bn))
        if rn <= tol or it >= max_iters:
            break
        rho = np.dot(r, rs)
        beta = (rho / rho_prev) * (alpha / omega_prev)
        p = r - beta * omega_prev * nu + beta * p
        nu = A @ p
        alpha = rho / np.dot(nu, rs)
        s = r - alpha * nu
        t = A @ s
        omega = np.dot(t, s) / np.dot(t, t)
        x = x + alpha * p + omega * s
        if recalc and (it + 1) % recalc == 0:
            r = b - A @ x
        else:
            r = s - omega * t
        omega_prev, rho_prev = omega, rho
        it += 1
    return x, it, hist
