"""The oracle's fc3d_compute_error and fc3d_nsgs (fc3d_compute_error.c:70-122,
fc3d_nsgs.c:560-917) against closed forms, an independent residual, the
reference fixtures under the option sets of test/test_nsgs_1.c, and the
colour-order invariance that the GPU solver relies on."""
import numpy as np
import pytest

from conftest import load_fixture, natural_residual
from siconos_amd import numerics as N

# files of the reference's data_collection_1.c that are committed here; the
# five option sets of test_nsgs_1.c are expected to converge on all of them
# the reference's data_collection_1.c, all 25 files (used by test_nsgs_1.c)
COLLECTION1 = ["FC3D_Example1.dat", "FC3D_Example1_SBM.dat", "FrictionContact3D_1c.dat",
               "FrictionContact3D_RR_1c.dat", "Capsules-i100-1090.dat", "Capsules-i100-889.dat",
               "Capsules-i101-404.dat", "Capsules-i103-990.dat", "Capsules-i122-1617.dat",
               "Rover1039.dat", "Rover1040.dat", "Rover1041.dat", "Rover11035.dat",
               "Rover11211.dat", "Rover3865.dat", "Rover4144.dat", "Rover4493.dat",
               "Rover4516.dat", "Rover4609.dat", "Rover4613.dat", "Rover4622.dat",
               "Rover9770.dat", "NESpheres_10_1.dat", "NESpheres_30_1.dat",
               "Confeti-ex13-Fc3D-SBM.dat"]


def example1_solution():
    s = np.sqrt(10.0)
    r = np.tile([1.0, -0.1 / s, -0.3 / s], 3)
    u = np.tile([0.0, 1 - 0.1 / s, 3 - 0.3 / s], 3)
    return r, u


@pytest.mark.parametrize("name", ["FC3D_Example1_SBM.dat", "FC3D_Example1.dat"])
def test_example1_closed_form(oracle, name):
    W, q, mu = load_fixture(name)
    np.testing.assert_array_equal(W.to_dense(), np.eye(9))
    r, u = example1_solution()
    e, uo = oracle.compute_error(W, q, mu, r)
    assert e < 1e-15
    np.testing.assert_allclose(uo, u, rtol=1e-15)
    info, rr, uu, it, res = oracle.nsgs(W, q, mu, tol=1e-12)
    assert info == 0 and res < 1e-12
    np.testing.assert_allclose(rr, r, rtol=1e-14)
    np.testing.assert_allclose(uu, u, rtol=1e-14)


def test_compute_error_matches_independent_residual(oracle):
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    rng = np.random.default_rng(0)
    for _ in range(3):
        r = rng.normal(size=3000)
        e, u = oracle.compute_error(g.W, g.q, g.mu, r)
        e2, u2 = natural_residual(g.W, g.q, g.mu, r)
        np.testing.assert_allclose(u, u2, rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(e, e2, rtol=1e-10)


@pytest.mark.parametrize("mut", [2, 3])
def test_residual_mutations_are_caught(oracle, mut):
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    r = np.random.default_rng(0).normal(size=3000)
    e2, _ = natural_residual(g.W, g.q, g.mu, r)
    try:
        oracle.mutation(mut)
        e, _ = oracle.compute_error(g.W, g.q, g.mu, r)
    finally:
        oracle.mutation(0)
    assert abs(e - e2) > 1e-6 * e2


def _options(oracle, variant):
    """The five option sets of the reference test/test_nsgs_1.c:26-114."""
    o = oracle.Options()
    o.dparam[0] = 1e-5
    if variant == 1:
        o.set_local_solver(551)
        o.internal().iparam[10] = 1  # JeanMoreau formulation
    elif variant == 2:
        o.iparam[0] = 10000
        o.set_local_solver(553)
        o.internal().dparam[0] = 1e-12
        o.internal().iparam[0] = 10
    elif variant == 3:
        o.iparam[0] = 10000
        o.set_local_solver(551)
        o.internal().iparam[14] = 2
    elif variant == 4:
        o.iparam[0] = 10000
        o.set_local_solver(551)
        o.internal().iparam[14] = 1
    return o


@pytest.mark.parametrize("variant", range(5))
@pytest.mark.parametrize("name", COLLECTION1)
def test_fixture_collections_converge(oracle, name, variant):
    W, q, mu = load_fixture(name)
    o = _options(oracle, variant)
    info, r, u, it, res = oracle.nsgs(W, q, mu, options=o)
    assert info == 0, (name, variant, it, res)
    assert np.all(np.isfinite(r)) and np.all(np.isfinite(u))
    e, _ = natural_residual(W, q, mu, r)
    assert e <= 1e-5 * (1 + 1e-6)


def test_config1_converges_to_1e8(oracle):
    """300 spheres, 1000 contacts, tol 1e-8, 1000 sweeps (the CPU case)."""
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    info, r, u, it, res = oracle.nsgs(g.W, g.q, g.mu, tol=1e-8, max_iter=1000)
    assert info == 0 and res < 1e-8
    e, _ = natural_residual(g.W, g.q, g.mu, r)
    assert e < 1e-8


def test_order_inside_a_colour_does_not_matter(oracle):
    """Contacts of one colour share no body (W_ij = 0), so any order inside a
    colour gives the same reactions bit for bit."""
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 3)
    colour, nco = oracle.greedy_colouring(g.W)
    rng = np.random.default_rng(1)
    orders = []
    for t in range(2):
        order = []
        for c in range(nco):
            members = np.flatnonzero(colour == c)
            order.extend(rng.permutation(members) if t else members)
        orders.append(np.array(order, dtype=np.int32))
    out = [oracle.nsgs(g.W, g.q, g.mu, tol=1e-30, max_iter=40, order=o)[1] for o in orders]
    np.testing.assert_array_equal(out[0], out[1])
    natural = oracle.nsgs(g.W, g.q, g.mu, tol=1e-30, max_iter=40)[1]
    assert not np.array_equal(natural, out[0])


def test_max_iter_returns_minus_20(oracle):
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    info, r, u, it, res = oracle.nsgs(g.W, g.q, g.mu, tol=1e-8, max_iter=1)
    assert info == -20 and it == 1


def test_trivial_case(oracle):
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    q = np.abs(g.q)
    info, r, u, it, res = oracle.nsgs(g.W, q, g.mu, tol=1e-8)
    assert info == 0 and it == 0
    np.testing.assert_array_equal(r, 0.0)
    np.testing.assert_array_equal(u, q)


def test_solve_many_equals_one_solve_per_problem(oracle):
    """The batched CPU baseline (oracle_solve_many, OpenMP over problems)
    returns what one fc3d_driver call per problem returns."""
    probs = N.generate_batch(12, 8, 60, 0.1, 0.9, 4)
    ok, iters, res, r = oracle.solve_many(probs, 1e-6, 1000, 4)
    off = 0
    for k, (W, q, mu) in enumerate(probs):
        info, rk, uk, it, rs = oracle.nsgs(W, q, mu, tol=1e-6, max_iter=1000)
        assert it == iters[k] and rs == res[k]
        np.testing.assert_array_equal(r[3 * off:3 * (off + W.nc)], rk)
        off += W.nc
    assert ok == sum(1 for x in res if x <= 1e-6)


@pytest.mark.parametrize("name", COLLECTION1)
def test_spectral_norm_rho_converges(oracle, name):
    """Reference defaults with the SPECTRAL_NORM rho strategy (local
    iparam[9] = 1, fc3d_AlartCurnier_functions.c:814-829) on the collection."""
    W, q, mu = load_fixture(name)
    o = _options(oracle, 0)
    o.internal().iparam[9] = 1
    info, r, u, it, res = oracle.nsgs(W, q, mu, options=o)
    assert info == 0, (name, it, res)
    e, _ = natural_residual(W, q, mu, r)
    assert e <= 1e-5 * (1 + 1e-6)


FREEZE_COLLECTION = ["Capsules-i100-1090.dat", "Capsules-i100-889.dat", "Capsules-i101-404.dat",
                     "Capsules-i103-990.dat", "Capsules-i122-1617.dat"]


def freeze_options(oracle, variant):
    """The five option sets of the reference test/test_nsgs_freeze_1.c:26-118
    (run on data_collection_2.c): the test_nsgs_1.c sets with freezing 10."""
    o = _options(oracle, variant)
    o.iparam[15] = 10
    return o


@pytest.mark.parametrize("variant", range(5))
@pytest.mark.parametrize("name", FREEZE_COLLECTION)
def test_freeze_collection_converges(oracle, name, variant):
    W, q, mu = load_fixture(name)
    info, r, u, it, res = oracle.nsgs(W, q, mu, options=freeze_options(oracle, variant))
    assert info == 0, (name, variant, it, res)
    e, _ = natural_residual(W, q, mu, r)
    assert e <= 1e-5 * (1 + 1e-6)


def _freeze_replay(oracle, nfreeze=3, sweeps=60):
    """Runs the oracle with freezing in LIGHT mode (fixed tolerance) and
    replays fc3d_nsgs.c:719-800 from the reactions after every sweep: a
    contact frozen by the criteria keeps its reaction bit for bit for nfreeze
    sweeps, the light error includes its last term, and the counters reset
    when nc - 1 are frozen.  Returns (largest relative light-error mismatch,
    freezes seen, resets seen)."""
    g = N.generate_random_pairs(60, 150, 0.2, 0.5, 3)
    nc = g.nc
    tol = 1e-3
    o = oracle.Options()
    o.iparam[13] = 1  # LIGHT
    o.iparam[15] = nfreeze
    o.iparam[0] = sweeps
    o.dparam[0] = 1e-300
    rec = []
    o.set_callback(lambda r, u, err: rec.append((r, err)))
    oracle.nsgs(g.W, g.q, g.mu, options=o)
    o.set_callback(None)
    frz = np.zeros(nc, dtype=int)
    le2 = np.zeros(nc)
    norm_r = 1e24
    prev = np.zeros(3 * nc)
    worst, freezes, resets = 0.0, 0, 0
    for k, (R, err) in enumerate(rec, start=1):
        ftmp = float(nc) * float(nc) / 1000.0
        crit1 = 1e-300 * 1e-300 / ftmp
        crit2 = norm_r * norm_r / ftmp
        if np.count_nonzero(frz) >= nc - 1:
            frz[:] = 0
            resets += 1
        light = 0.0
        for c in range(nc):
            o_, l_ = prev[3 * c:3 * c + 3], R[3 * c:3 * c + 3]
            if frz[c] > 0:
                assert np.array_equal(l_, o_), (k, c)
                frz[c] -= 1
                light += le2[c]
                continue
            d = [float(o_[0] - l_[0]), float(o_[1] - l_[1]), float(o_[2] - l_[2])]
            le2[c] = d[0] * d[0] + d[1] * d[1] + d[2] * d[2]
            light += le2[c]
            sq = float(l_[0] * l_[0] + l_[1] * l_[1] + l_[2] * l_[2])
            if (le2[c] <= crit1 * sq or sq <= crit2) and k >= 10:
                frz[c] = nfreeze
                freezes += 1
        s = 0.0
        for x in R:
            s += float(x) * float(x)
        norm_r = np.sqrt(s)
        incr = np.sqrt(light) / norm_r
        worst = max(worst, abs(incr - err) / max(abs(incr), 1e-300))
        prev = R
    del tol
    return worst, freezes, resets, len(rec)


def test_freeze_bookkeeping_replayed_sweep_by_sweep(oracle):
    worst, freezes, resets, n = _freeze_replay(oracle)
    assert n == 60 and freezes > 100 and resets > 0
    assert worst < 1e-13


def test_freeze_stale_light_error_mutation_is_caught(oracle):
    try:
        oracle.mutation(10)
        worst, freezes, _, _ = _freeze_replay(oracle)
        assert freezes > 100 and worst > 1e-3
    finally:
        oracle.mutation(0)
