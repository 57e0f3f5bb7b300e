"""Brute-force solutions of one-contact Coulomb problems, from the
definition of the problem: r in K, u = W r + q, u~ = u + mu |u_t| e_n in K*,
r . u~ = 0.  Used to pin the oracle's local solvers (test helper)."""
import numpy as np


def closed_form_diag(a, b, q, mu):
    """W = diag(a, b, b): take-off, stick or slide."""
    if q[0] >= 0:
        return np.zeros(3)
    rn = -q[0] / a
    qt = np.asarray(q[1:])
    nq = np.linalg.norm(qt)
    if nq / b <= mu * rn:
        return np.array([rn, -qt[0] / b, -qt[1] / b])
    return np.array([rn, -mu * rn * qt[0] / nq, -mu * rn * qt[1] / nq])


def one_contact_solutions(W, q, mu, ngrid=20000):
    """All solutions for a general SPD W (take-off, stick, or slide along
    direction theta found by scanning and bisection)."""
    W = np.asarray(W, dtype=float)
    q = np.asarray(q, dtype=float)
    sols = []
    if q[0] >= 0:
        sols.append(np.zeros(3))
    r = -np.linalg.solve(W, q)
    if r[0] > 0 and np.hypot(r[1], r[2]) <= mu * r[0]:
        sols.append(r)

    def slide(th):
        d = np.array([1.0, -mu * np.cos(th), -mu * np.sin(th)])
        wd = W @ d
        rn = -q[0] / wd[0]
        u = q + rn * wd
        return rn, u, d

    def cross(th):
        rn, u, _ = slide(th)
        return u[1] * np.sin(th) - u[2] * np.cos(th)

    ths = np.linspace(0, 2 * np.pi, ngrid + 1)
    cs = np.array([cross(t) for t in ths])
    for k in range(ngrid):
        if cs[k] == 0 or cs[k] * cs[k + 1] < 0:
            lo, hi = ths[k], ths[k + 1]
            for _ in range(200):
                mid = 0.5 * (lo + hi)
                if np.sign(cross(mid)) == np.sign(cross(lo)):
                    lo = mid
                else:
                    hi = mid
            th = 0.5 * (lo + hi)
            rn, u, d = slide(th)
            if rn > 0 and u[1] * np.cos(th) + u[2] * np.sin(th) > 0:
                sols.append(rn * d)
    return sols
