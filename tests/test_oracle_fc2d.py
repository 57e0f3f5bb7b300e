"""The fc2d oracle (oracle/fc2d_oracle.c) and the product's reading of the
2-D fixtures, on the CPU.

Pins of the oracle:
- the local solve against closed forms (diagonal block) and against the 2-D
  Coulomb law checked with an independent numpy residual;
- the residual and the row products against numpy on dense W;
- the NSGS against the reference's own test: test_fc2d_1.c runs FC2D_NSGS
  at tol 1e-5 and 10000 iterations on data_collection_fc2d_1.c and expects
  every run to succeed.
"""
import numpy as np
import pytest

from conftest import FC2D_COLLECTION, fixture_file
from oracle import pyoracle as O
from siconos_amd import numerics as N


def natural_map_error(W, q, mu, r):
    """sqrt(sum_c |r_c - P_K(r_c - (u_n + mu |u_t|, u_t))|^2) with u = W r + q,
    scaled by max(|r|, |u|, |q|) when that exceeds DBL_EPSILON: the fc2d
    natural-map residual written from its definition (projection onto the
    2-D cone {|x_t| <= mu x_n})."""
    u = W @ r + q
    e = 0.0
    for c in range(len(mu)):
        x = np.array([r[2 * c] - u[2 * c] - mu[c] * abs(u[2 * c + 1]), r[2 * c + 1] - u[2 * c + 1]])
        m = mu[c]
        if abs(x[1]) <= m * x[0]:
            p = x
        elif m * abs(x[1]) <= -x[0]:
            p = np.zeros(2)
        else:
            pn = (x[0] + m * abs(x[1])) / (1 + m * m)
            p = np.array([pn, m * pn * np.sign(x[1])])
        e += np.sum((r[2 * c:2 * c + 2] - p) ** 2)
    scale = max(np.linalg.norm(r), np.linalg.norm(u), np.linalg.norm(q))
    return np.sqrt(e) / scale if scale > np.finfo(float).eps else np.sqrt(e)


@pytest.mark.parametrize("name", FC2D_COLLECTION)
def test_product_reader_matches_the_oracle_parser(lib, name, tmp_path):
    """Two independent parsers of the same .dat file agree, for dense, SBM
    and CSparse storage; the product writes what it reads."""
    p = N.FrictionContactProblem.from_file(fixture_file(name))
    o = O.Fc2dProblem(fixture_file(name))
    W, q, mu = o.dense()
    assert p.dimension == 2 and p.numberOfContacts == o.nc and p.storage == o.storage
    np.testing.assert_array_equal(p.dense(), W)
    np.testing.assert_array_equal(p.q, q)
    np.testing.assert_array_equal(p.mu, mu)
    out = str(tmp_path / "p.dat")
    assert p.write(out) == 0
    p2 = N.FrictionContactProblem.from_file(out)
    np.testing.assert_array_equal(p2.dense(), W)
    np.testing.assert_array_equal(p2.q, q)


def test_local_solve_closed_forms():
    a, b, mu = 2.0, 5.0, 0.3
    W = np.array([a, 0.0, 0.0, b])  # column-major diag(a, b)
    # take-off
    np.testing.assert_array_equal(O.fc2d_local_solve(W, [0.5, -1.0], mu), [0.0, 0.0])
    # stick: r = -W^-1 q inside the cone
    r = O.fc2d_local_solve(W, [-1.0, 0.2], mu)
    np.testing.assert_allclose(r, [0.5, -0.04], rtol=1e-15)
    # slide in both tangential directions: r_t = -+ mu r_n, r_n = -q_n / a
    r = O.fc2d_local_solve(W, [-1.0, 3.0], mu)
    np.testing.assert_allclose(r, [0.5, -0.15], rtol=1e-15)
    r = O.fc2d_local_solve(W, [-1.0, -3.0], mu)
    np.testing.assert_allclose(r, [0.5, 0.15], rtol=1e-15)


def _random_blocks(seed, n):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        A = rng.normal(size=(2, 2))
        W = A @ A.T + 0.5 * np.eye(2)
        q = rng.normal(size=2) * 2
        out.append((W, q, rng.uniform(0.05, 1.0)))
    return out


def test_local_solve_satisfies_the_coulomb_law():
    """On random symmetric positive blocks the one-contact solution has zero
    natural-map residual; a planted swap of the slide directions fails it."""
    cases = _random_blocks(7, 400)
    errs = []
    for W, q, mu in cases:
        r = O.fc2d_local_solve(W, q, mu)
        errs.append(natural_map_error(W, q, np.array([mu]), r))
    assert max(errs) < 1e-13
    O.fc2d_mutation(1)
    try:
        bad = [natural_map_error(W, q, np.array([mu]), O.fc2d_local_solve(W, q, mu))
               for W, q, mu in cases]
    finally:
        O.fc2d_mutation(0)
    assert max(bad) > 1e-3


@pytest.mark.parametrize("name", FC2D_COLLECTION)
def test_row_products_and_residual_match_numpy(name):
    o = O.Fc2dProblem(fixture_file(name))
    W, q, mu = o.dense()
    rng = np.random.default_rng(3)
    r = rng.normal(size=2 * o.nc)
    for c in sorted(set(rng.integers(0, o.nc, 20).tolist()) | {0, o.nc - 1}):
        y = o.local_q(c, r)
        x = r.copy()
        x[2 * c:2 * c + 2] = 0.0
        ref = q[2 * c:2 * c + 2] + W[2 * c:2 * c + 2] @ x
        np.testing.assert_allclose(y, ref, rtol=1e-12, atol=1e-12 * np.abs(W).max())
    e, u = o.compute_error(r)
    np.testing.assert_allclose(u, W @ r + q, rtol=1e-12, atol=1e-12 * np.abs(W).max())
    assert abs(e - natural_map_error(W, q, mu, r)) <= 1e-12 * max(e, 1e-300)


@pytest.mark.parametrize("name", FC2D_COLLECTION)
def test_oracle_nsgs_passes_the_reference_test(name):
    """test_fc2d_1.c: FC2D_NSGS, tol 1e-5, 10000 iterations, every fixture of
    data_collection_fc2d_1.c succeeds."""
    o = O.Fc2dProblem(fixture_file(name))
    info, r, u, it, res = o.nsgs(tol=1e-5, max_iter=10000, mode=2)
    W, q, mu = o.dense()
    assert info == 0 and res <= 1e-5 and it >= 1
    assert natural_map_error(W, q, mu, r) <= 1e-5
    np.testing.assert_allclose(u, W @ r + q, rtol=1e-12, atol=1e-12 * np.abs(W).max())


def test_oracle_nsgs_order_and_modes():
    """Any sweep order reaches the tolerance; LIGHT mode stops on the light
    error alone; an unevaluated mode (FULL) sweeps max_iter times and reports
    info 1 with an infinite residual, as fc2d_nsgs does; max_iter bounds the
    sweeps and returns info 1."""
    o = O.Fc2dProblem(fixture_file("FrictionContactProblem00237.dat"))
    W, q, mu = o.dense()
    rng = np.random.default_rng(1)
    order = rng.permutation(o.nc).astype(np.int32)
    info, r, u, it, res = o.nsgs(tol=1e-5, max_iter=10000, mode=2, order=order)
    assert info == 0 and natural_map_error(W, q, mu, r) <= 1e-5
    info, r, u, it, res = o.nsgs(tol=1e-5, max_iter=10000, mode=1)
    assert info == 0 and res < 1e-5
    info, r, u, it, res = o.nsgs(tol=1e-5, max_iter=7, mode=0)
    assert info == 1 and it == 7 and res == np.inf
    info, r, u, it, res = o.nsgs(tol=1e-12, max_iter=3, mode=2)
    assert info == 1 and it == 3
