"""Edge cases of the GPU solve through the C ABI, against the oracle run in
the same sweep order (fc3d_nsgs.c:560-917): a single contact, contacts that
share no body (W block diagonal, one colour), frictionless contacts
(mu = 0), and a star, one body carrying every contact, so that each row
holds all the other contacts and each contact is a colour of its own (more
colours than the staged focus passes take).  The matrix-form FP64 results
are compared bit for bit; the velocity form and the FP32 sweeps are judged
by the oracle's fc3d_compute_error."""
import numpy as np
import pytest

from siconos_amd import numerics as N
from test_gpu_velocity import _dense_problem

pytestmark = pytest.mark.gpu


def random_bodies(nb, ba, bb, seed):
    """Random 3x6 blocks, diagonal M^-1 and free velocities; W and q
    assembled densely on the host."""
    rng = np.random.default_rng(seed)
    nc = len(ba)
    H = rng.normal(size=36 * nc)
    minv = rng.uniform(0.5, 2.0, size=6 * nb)
    vfree = rng.normal(size=6 * nb)
    W, q, _ = _dense_problem(nb, ba, bb, H, minv, vfree)
    return W, q, H, minv, vfree


def plain_matches_oracle(oracle, W, q, mu, tol, max_iter=2000, schedule=0):
    P = N.HipProblem(W, q, mu)
    P.set_reaction(np.zeros(3 * W.nc))
    s = P.solve(N.hip_params(tol=tol, max_iter=max_iter, schedule=schedule))
    info, ro, uo, it, res = oracle.nsgs(W, q, mu, tol=tol, max_iter=max_iter, order=P.order())
    assert s.info == info == 0 and s.iterations == it
    np.testing.assert_array_equal(P.reaction(), ro)
    np.testing.assert_array_equal(P.velocity(), uo)
    np.testing.assert_allclose(s.residual, res, rtol=1e-12)
    return P, s


def velocity_form_converges(oracle, P, W, q, mu, nb, ba, bb, H, minv, vfree, tol, **extra):
    P.set_bodies(nb, ba, bb, H, minv, vfree)
    for precision in (0, 1):
        s = P.solve(N.hip_params(tol=tol, max_iter=5000, warm_start=1, velocity_form=1,
                                 precision=precision, fp32_tol=max(tol, 1e-6),
                                 fp32_local_solver=550 if precision else 0, **extra))
        assert s.info == 0 and s.residual <= tol, precision
        e, _ = oracle.compute_error(W, q, mu, P.reaction())
        assert e <= tol and e == pytest.approx(s.residual, rel=1e-9), precision


@pytest.mark.parametrize("one_body", [False, True])
def test_a_single_contact(oracle, one_body):
    """One contact, in a slice of 64 slots whose other 63 are padding."""
    ba = np.array([0], dtype=np.int32)
    bb = np.array([-1 if one_body else 1], dtype=np.int32)
    W, q, H, minv, vfree = random_bodies(2, ba, bb, 11)
    if q[0] >= 0.0:  # approaching, not the take-off case: q = H v_free with v_free negated
        vfree = -vfree
        q = -q
    assert q[0] < 0.0
    mu = np.array([0.4])
    P, s = plain_matches_oracle(oracle, W, q, mu, tol=1e-12)
    assert P.num_colours == 1
    velocity_form_converges(oracle, P, W, q, mu, 2, ba, bb, H, minv, vfree, tol=1e-10)


def test_contacts_that_share_no_body(oracle):
    """W block diagonal: one colour, each contact's local problem is the
    whole problem, so the first sweep solves it and the second confirms."""
    nc = 300
    ba = np.arange(0, 2 * nc, 2, dtype=np.int32)
    bb = ba + 1
    W, q, H, minv, vfree = random_bodies(2 * nc, ba, bb, 12)
    assert W.nnzb == nc
    mu = np.random.default_rng(1).uniform(0.1, 0.9, nc)
    P, s = plain_matches_oracle(oracle, W, q, mu, tol=1e-10)
    assert P.num_colours == 1 and s.iterations <= 3
    velocity_form_converges(oracle, P, W, q, mu, 2 * nc, ba, bb, H, minv, vfree, tol=1e-8)


def test_frictionless_contacts(oracle):
    """mu = 0: the cone is the half-line of normal reactions, the tangential
    reactions stay zero."""
    g = N.generate_random_pairs(300, 1000, 0.0, 0.0, 3)
    mu = np.asarray(g.mu)
    assert np.all(mu == 0.0)
    P, s = plain_matches_oracle(oracle, g.W, g.q, mu, tol=1e-8)
    r = P.reaction().reshape(-1, 3)
    assert np.all(r[:, 1:] == 0.0) and np.all(r[:, 0] >= 0.0)
    velocity_form_converges(oracle, P, g.W, g.q, mu, g.nbodies, g.body_a, g.body_b, g.H,
                            g.minv, g.vfree, tol=1e-7)


@pytest.fixture(scope="module")
def star():
    n = 100
    ba = np.zeros(n, dtype=np.int32)
    bb = np.arange(1, n + 1, dtype=np.int32)
    W, q, H, minv, vfree = random_bodies(n + 1, ba, bb, 13)
    mu = np.random.default_rng(2).uniform(0.1, 0.9, n)
    return n, ba, bb, W, q, H, minv, vfree, mu


@pytest.mark.parametrize("schedule", [0, 1, 7, 8])
def test_a_star_of_long_rows(oracle, star, schedule):
    """Every row holds 99 off-diagonal blocks (the gathers' longest rows)
    and each contact is a colour of its own (100 colours, every slice of the
    single-launch schedule 7 waiting on the one before)."""
    n, ba, bb, W, q, H, minv, vfree, mu = star
    assert W.nnzb == n * n
    P, s = plain_matches_oracle(oracle, W, q, mu, tol=1e-8, schedule=schedule)
    assert P.num_colours == n


def test_a_star_in_the_velocity_form_with_focused_passes(oracle, star):
    """The velocity form on the star, with focused passes over more colours
    than the staged passes take (their unstaged form runs)."""
    n, ba, bb, W, q, H, minv, vfree, mu = star
    P = N.HipProblem(W, q, mu)
    velocity_form_converges(oracle, P, W, q, mu, n + 1, ba, bb, H, minv, vfree, tol=1e-8,
                            focus_passes=10, focus_every=16, focus_after=16, focus_count=16,
                            focus_share=0.0)


@pytest.mark.parametrize("schedule", [1, 2])
def test_the_edge_cases_in_one_batch(oracle, star, schedule):
    """The same problems side by side in one persistent batch launch
    (fc3d_hip_batch_*), each bit-exact with the oracle in its own sweep
    order: a one-contact problem, a block-diagonal one, a frictionless one
    and the star, between two ordinary problems."""
    ba1 = np.array([0], dtype=np.int32)
    W1, q1, *_ = random_bodies(2, ba1, np.array([1], dtype=np.int32), 11)
    if q1[0] >= 0.0:
        q1 = -q1
    bad = np.arange(0, 120, 2, dtype=np.int32)
    Wd, qd, *_ = random_bodies(120, bad, bad + 1, 12)
    g0 = N.generate_random_pairs(80, 200, 0.0, 0.0, 3)
    n, ba, bb, Ws, qs, H, minv, vfree, mus = star
    rng = np.random.default_rng(5)
    probs = N.generate_batch(2, 40, 80, 0.1, 0.9, 17)
    probs = [probs[0], (W1, q1, np.array([0.4])), (Wd, qd, rng.uniform(0.1, 0.9, 60)),
             (g0.W, g0.q, np.asarray(g0.mu)), (Ws, qs, mus), probs[1]]
    B = N.Batch(probs)
    B.set_reaction(np.zeros(3 * B.ntotal))
    s, info, iters, res = B.solve(N.hip_params(tol=1e-8, max_iter=2000, schedule=schedule))
    r, u, order = B.reaction(), B.velocity(), B.order()
    for k, (W, q, mu) in enumerate(probs):
        c0 = B.offsets[k]
        sel = order[(order >= c0) & (order < c0 + W.nc)] - c0
        io, ro, uo, it, reso = oracle.nsgs(W, q, mu, tol=1e-8, max_iter=2000, order=sel)
        assert io == 0 and info[k] == io and iters[k] == it, k
        np.testing.assert_array_equal(r[3 * c0:3 * (c0 + W.nc)], ro, err_msg=f"problem {k}")
        np.testing.assert_array_equal(u[3 * c0:3 * (c0 + W.nc)], uo, err_msg=f"problem {k}")
        np.testing.assert_allclose(res[k], reso, rtol=1e-12)
    assert s.converged == len(probs)
