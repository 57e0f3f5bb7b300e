"""Pins of the oracle's 3x3 algebra and cone projection (op3x3.h:342-981,
projectionOnCone.c:32-73) against numpy and mathematical invariants."""
import numpy as np
import pytest

from conftest import cone_projection


def test_gepp_matches_numpy_solve(oracle):
    L = oracle.load()
    rng = np.random.default_rng(1)
    for t in range(400):
        A = rng.normal(size=(3, 3)) + (3 * np.eye(3) if t % 2 else 0)
        # permute rows so that every pivot branch (0, 1, 2 first) is taken
        A = A[np.roll(np.arange(3), t % 3)]
        b = rng.normal(size=3)
        a = np.ascontiguousarray(A.reshape(-1, order="F"))
        x = b.copy()
        info = L.oracle_solve_3x3_gepp(oracle.d(a), oracle.d(x))
        assert info == 0
        np.testing.assert_allclose(x, np.linalg.solve(A, b), rtol=1e-9, atol=1e-11)


def test_gepp_singular_returns_info(oracle):
    """info = 1 when the first or second pivot is below DBL_EPSILON; a zero
    third pivot is not tested by the reference (op3x3.h:955-978) and yields
    non-finite values with info = 0."""
    L = oracle.load()
    z = np.zeros(9)
    assert L.oracle_solve_3x3_gepp(oracle.d(z), oracle.d(np.ones(3))) == 1
    A2 = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 1.0], [3.0, 6.0, 2.0]])  # 2nd pivot zero
    a2 = np.ascontiguousarray(A2.reshape(-1, order="F"))
    assert L.oracle_solve_3x3_gepp(oracle.d(a2), oracle.d(np.ones(3))) == 1
    A3 = np.array([[1.0, 2.0, 3.0], [2.0, 4.0, 6.0], [1.0, 0.0, 1.0]])  # 3rd pivot zero
    a3 = np.ascontiguousarray(A3.reshape(-1, order="F"))
    x = np.ones(3)
    assert L.oracle_solve_3x3_gepp(oracle.d(a3), oracle.d(x)) == 0
    assert not np.all(np.isfinite(x))


def test_gepp_pivoting_mutation_is_caught(oracle):
    L = oracle.load()
    A = np.array([[1e-17, 1.0, 0.0], [1.0, 1.0, 0.0], [0.0, 0.0, 1.0]])
    b = np.array([1.0, 2.0, 3.0])
    a = np.ascontiguousarray(A.reshape(-1, order="F"))
    try:
        oracle.mutation(6)
        x = b.copy()
        info = L.oracle_solve_3x3_gepp(oracle.d(a), oracle.d(x))
        bad = info != 0 or not np.allclose(x, np.linalg.solve(A, b), rtol=1e-9)
    finally:
        oracle.mutation(0)
    assert bad


def test_mvp_and_mm_column_major(oracle):
    L = oracle.load()
    rng = np.random.default_rng(2)
    A, B = rng.normal(size=(3, 3)), rng.normal(size=(3, 3))
    v, r = rng.normal(size=3), rng.normal(size=3)
    a = np.ascontiguousarray(A.reshape(-1, order="F"))
    b = np.ascontiguousarray(B.reshape(-1, order="F"))
    y = r.copy()
    L.oracle_mvp3x3(oracle.d(a), oracle.d(v), oracle.d(y))
    np.testing.assert_allclose(y, r + A @ v, rtol=1e-14)
    c = np.zeros(9)
    L.oracle_mm3x3(oracle.d(a), oracle.d(b), oracle.d(c))
    np.testing.assert_allclose(c.reshape(3, 3, order="F"), A @ B, rtol=1e-13)


def test_ortho_base(oracle):
    L = oracle.load()
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = rng.normal(size=3)
        t1, t2 = np.zeros(3), np.zeros(3)
        nn = n.copy()
        assert L.oracle_orthoBaseFromVector(oracle.d(nn), oracle.d(t1), oracle.d(t2)) == 0
        R = np.vstack([nn, t1, t2])
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        np.testing.assert_allclose(nn, n / np.linalg.norm(n), rtol=1e-14)


def _proj(oracle, x, mu):
    y = np.ascontiguousarray(x, dtype=float).copy()
    st = oracle.load().oracle_projectionOnCone(oracle.d(y), mu)
    return y, st


def test_projection_invariants(oracle):
    """Idempotence, orthogonality and the variational inequality of a
    Euclidean projection onto a closed convex cone (the pattern of the
    reference tools/test/test_projection.c)."""
    rng = np.random.default_rng(4)
    seen = set()
    for _ in range(3000):
        mu = rng.uniform(0.05, 1.5)
        x = rng.normal(size=3) * rng.uniform(0.1, 10)
        p, st = _proj(oracle, x, mu)
        seen.add(st)
        assert np.hypot(p[1], p[2]) <= mu * p[0] * (1 + 1e-12) + 1e-14
        p2, st2 = _proj(oracle, p, mu)
        np.testing.assert_allclose(p2, p, rtol=1e-12, atol=1e-14)
        assert abs(np.dot(x - p, p)) <= 1e-12 * max(1.0, np.dot(x, x))
        for _ in range(3):  # (x - Px).(y - Px) <= 0 for y in the cone
            yn = rng.uniform(0, 3)
            ang = rng.uniform(0, 2 * np.pi)
            rad = rng.uniform(0, mu * yn)
            y = np.array([yn, rad * np.cos(ang), rad * np.sin(ang)])
            assert np.dot(x - p, y - p) <= 1e-10 * max(1.0, np.dot(x, x))
    assert seen == {0, 1, 2}


def test_projection_matches_definition_and_mutation(oracle):
    rng = np.random.default_rng(5)
    xs = [rng.normal(size=3) for _ in range(500)]
    for x in xs:
        mu = 0.3
        p, _ = _proj(oracle, x, mu)
        np.testing.assert_allclose(p, cone_projection(x, mu), rtol=1e-12, atol=1e-14)
    try:
        oracle.mutation(4)
        bad = any(not np.allclose(_proj(oracle, x, 0.3)[0], cone_projection(x, 0.3)) for x in xs)
    finally:
        oracle.mutation(0)
    assert bad


def test_dual_cone_projection(oracle):
    """Moreau decomposition: P_K(x) + P_{-K*}(x) = x with K* the dual cone
    of slope 1/mu; projectionOnDualCone(u) projects on K*."""
    L = oracle.load()
    rng = np.random.default_rng(6)
    for _ in range(500):
        mu = rng.uniform(0.1, 1.0)
        x = rng.normal(size=3)
        pk, _ = _proj(oracle, x, mu)
        y = -x.copy()
        L.oracle_projectionOnDualCone(oracle.d(y), mu)  # P_{K*}(-x)
        np.testing.assert_allclose(pk - y, x, rtol=1e-10, atol=1e-12)


def _eig3(oracle, A):
    a = np.ascontiguousarray(np.asarray(A, dtype=float).reshape(-1, order="F"))
    e = np.zeros(3)
    oracle.load().oracle_eig_3x3(oracle.d(a), oracle.d(e))
    return e


def _random_symmetric(rng, kind):
    if kind == "spd":
        M = rng.normal(size=(3, 3))
        return M @ M.T + 0.1 * np.eye(3)
    if kind == "indefinite":
        M = rng.normal(size=(3, 3))
        return M + M.T
    # repeated eigenvalues: Q diag(l, l, m) Q^T
    Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    lam = rng.uniform(0.5, 5)
    return Q @ np.diag([lam, lam, rng.uniform(0.5, 5)]) @ Q.T


def _eig_max_error(oracle, seed=31):
    """Largest |eig - eigvalsh| / |A| over random symmetric matrices (SPD,
    indefinite, repeated eigenvalues); the oracle's order is descending."""
    rng = np.random.default_rng(seed)
    worst = 0.0
    for kind in ("spd", "indefinite", "repeated"):
        for _ in range(300):
            A = _random_symmetric(rng, kind)
            ref = np.sort(np.linalg.eigvalsh(A))[::-1]
            worst = max(worst, np.max(np.abs(_eig3(oracle, A) - ref)) / np.linalg.norm(A))
    return worst


def test_eig_3x3_against_eigvalsh(oracle):
    """op3x3.h:1065-1111 against LAPACK's symmetric eigensolver."""
    assert _eig_max_error(oracle) < 1e-7  # the trigonometric method loses digits
    # near repeated eigenvalues (acos near +-1); well separated ones are tight
    rng = np.random.default_rng(32)
    for _ in range(200):
        Q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        lam = np.sort(rng.uniform(1, 10, size=3))[::-1] + np.array([2.0, 1.0, 0.0])
        A = Q @ np.diag(lam) @ Q.T
        np.testing.assert_allclose(_eig3(oracle, A), lam, rtol=1e-12)


def test_eig_3x3_diagonal_is_returned_unsorted(oracle):
    """The reference returns a diagonal matrix's entries in stored order, so
    the spectral-norm rho of diag(2, 7, 7) is 1/2 (not 1/7)."""
    np.testing.assert_array_equal(_eig3(oracle, np.diag([2.0, 7.0, 7.0])), [2.0, 7.0, 7.0])
    rho = np.zeros(3)
    w = np.ascontiguousarray(np.diag([2.0, 7.0, 7.0]).reshape(-1, order="F"))
    oracle.load().oracle_rho_spectral_norm(oracle.d(w), oracle.d(rho))
    np.testing.assert_array_equal(rho, [0.5, 0.5, 0.5])
    A = np.array([[15.0, -1, -2], [-1, 14, 1], [-2, 1, 16]])
    w = np.ascontiguousarray(A.reshape(-1, order="F"))
    oracle.load().oracle_rho_spectral_norm(oracle.d(w), oracle.d(rho))
    np.testing.assert_allclose(rho, 1.0 / np.max(np.linalg.eigvalsh(A)), rtol=1e-13)


def test_eig_3x3_mutation_is_caught(oracle):
    try:
        oracle.mutation(9)
        assert _eig_max_error(oracle) > 1e-2
    finally:
        oracle.mutation(0)
