"""Local (one-contact) solvers of the oracle against closed forms and brute
force (onecontact...c:221-1015, fc3d_projection.c:350-603)."""
import numpy as np
import pytest

from brute_force import closed_form_diag, one_contact_solutions
from conftest import load_fixture

LOCAL = [565, 551, 550]


def solve_local(oracle, sid, W, q, mu, r0=None, tol=None, max_iter=None):
    L = oracle.load()
    o = oracle.Options.local(sid)
    if tol is not None:
        o.dparam[0] = tol
    if max_iter is not None:
        o.iparam[0] = max_iter
    w = np.ascontiguousarray(np.asarray(W, dtype=float).reshape(-1, order="F"))
    st = np.zeros(4)
    assert L.oracle_local_state_init(oracle.d(w), o.p, oracle.d(st)) == 0
    r = np.zeros(3) if r0 is None else np.array(r0, dtype=float)
    qq = np.ascontiguousarray(q, dtype=float)
    info = L.oracle_local_solve(oracle.d(w), oracle.d(qq), mu, oracle.d(r), o.p, oracle.d(st))
    return info, r


@pytest.mark.parametrize("sid", LOCAL)
def test_closed_form_diagonal_blocks(oracle, sid):
    rng = np.random.default_rng(11)
    cases = {"takeoff": 0, "stick": 0, "slide": 0}
    for _ in range(300):
        a, b = rng.uniform(0.5, 5), rng.uniform(0.5, 10)
        mu = rng.uniform(0.1, 0.9)
        q = rng.normal(size=3)
        exact = closed_form_diag(a, b, q, mu)
        info, r = solve_local(oracle, sid, np.diag([a, b, b]), q, mu)
        if q[0] >= 0:
            cases["takeoff"] += 1
        elif np.hypot(*exact[1:]) < mu * exact[0] * (1 - 1e-12):
            cases["stick"] += 1
        else:
            cases["slide"] += 1
        np.testing.assert_allclose(r, exact, rtol=1e-9, atol=1e-12)
    assert min(cases.values()) > 20


def test_projection_with_local_iteration_closed_form(oracle):
    rng = np.random.default_rng(12)
    for _ in range(200):
        a, b = rng.uniform(0.5, 5), rng.uniform(0.5, 10)
        mu = rng.uniform(0.1, 0.9)
        q = rng.normal(size=3)
        q[0] = -abs(q[0]) - 0.1
        info, r = solve_local(oracle, 553, np.diag([a, b, b]), q, mu, tol=1e-28, max_iter=20000)
        np.testing.assert_allclose(r, closed_form_diag(a, b, q, mu), rtol=1e-7, atol=1e-9)


def test_fixture_one_contact_brute_force(oracle):
    """FrictionContact3D_1c.dat: W = [[15,-1,-2],[-1,14,1],[-2,1,16]],
    q = (-1, 2, 3), mu = 0.1 — solved by a brute-force angle search."""
    W, q, mu = load_fixture("FrictionContact3D_1c.dat")
    Wd = W.to_dense()
    np.testing.assert_array_equal(Wd, [[15, -1, -2], [-1, 14, 1], [-2, 1, 16]])
    sols = one_contact_solutions(Wd, q, mu[0])
    assert len(sols) == 1
    for sid in LOCAL:
        info, r = solve_local(oracle, sid, Wd, q, mu[0])
        np.testing.assert_allclose(r, sols[0], rtol=1e-10, atol=1e-12)


def test_random_spd_one_contact_brute_force(oracle):
    rng = np.random.default_rng(13)
    n = 0
    for _ in range(60):
        M = rng.normal(size=(3, 3))
        W = M @ M.T + 2 * np.eye(3)
        q = rng.normal(size=3)
        mu = rng.uniform(0.1, 0.6)
        sols = one_contact_solutions(W, q, mu, ngrid=4000)
        if len(sols) != 1:
            continue
        n += 1
        info, r = solve_local(oracle, 565, W, q, mu)
        np.testing.assert_allclose(r, sols[0], rtol=1e-8, atol=1e-10)
    assert n > 40


def _moreau_projection_holds(x, p, mu, tol=1e-12):
    """p is the Euclidean projection of x onto K = {|t| <= mu n} iff p in K,
    x - p in the polar cone {mu |t| <= -n} and (x - p) . p = 0 (Moreau)."""
    d = x - p
    s = tol * max(1.0, np.linalg.norm(x))
    return (np.hypot(p[1], p[2]) <= mu * p[0] + s and mu * np.hypot(d[1], d[2]) <= -d[0] + s
            and abs(d @ p) <= s * max(1.0, np.linalg.norm(x)))


def _oc_proj_fixed_points(oracle):
    """One OC_PROJ step (552) from a brute-force solution of a random SPD
    one-contact problem: a solution satisfies r = P_K(r - rho (u + mu|u_t| e_n))
    for every rho > 0, so the step must return it unchanged.  Returns the
    largest relative move and the number of sliding solutions seen."""
    rng = np.random.default_rng(21)
    worst, slides, n = 0.0, 0, 0
    for _ in range(80):
        M = rng.normal(size=(3, 3))
        W = M @ M.T + 2 * np.eye(3)
        q = rng.normal(size=3)
        mu = rng.uniform(0.1, 0.6)
        sols = one_contact_solutions(W, q, mu, ngrid=4000)
        if len(sols) != 1 or sols[0][0] <= 0:
            continue
        rs = sols[0]
        n += 1
        if np.hypot(rs[1], rs[2]) > mu * rs[0] * (1 - 1e-9):
            slides += 1
        info, r = solve_local(oracle, 552, W, q, mu, r0=rs)
        worst = max(worst, np.linalg.norm(r - rs) / np.linalg.norm(rs))
    assert n > 40
    return worst, slides


def test_oc_proj_keeps_solutions_fixed(oracle):
    worst, slides = _oc_proj_fixed_points(oracle)
    assert slides > 10
    assert worst < 1e-9


def _oc_proj_diag_from_zero(oracle):
    """W = diag(a, b, b), r = 0: u = q, so the step is the Moreau projection
    of x = -(q_n + mu |q_t|, q_t) / a.  Returns the number of failures."""
    rng = np.random.default_rng(22)
    bad = 0
    for _ in range(300):
        a, b = rng.uniform(0.5, 5), rng.uniform(0.5, 10)
        mu = rng.uniform(0.1, 0.9)
        q = rng.normal(size=3)
        x = -np.array([q[0] + mu * np.hypot(q[1], q[2]), q[1], q[2]]) / a
        info, r = solve_local(oracle, 552, np.diag([a, b, b]), q, mu)
        if not _moreau_projection_holds(x, r, mu):
            bad += 1
    return bad


def test_oc_proj_closed_form_from_zero(oracle):
    assert _oc_proj_diag_from_zero(oracle) == 0


def test_oc_proj_mutations_are_caught(oracle):
    try:
        oracle.mutation(7)  # mu |u_t| dropped: sliding solutions move
        worst, _ = _oc_proj_fixed_points(oracle)
        assert worst > 1e-3
        oracle.mutation(8)  # rho = 1 / W11: the diagonal closed form fails
        assert _oc_proj_diag_from_zero(oracle) > 100
    finally:
        oracle.mutation(0)
