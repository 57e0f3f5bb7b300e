"""Anderson acceleration of the coloured sweep (fc3d_hip_params.accel_*).

The GPU iterates are checked against a numpy restatement of the same
Anderson rule driven by single oracle sweeps in the GPU's colour order. The
oracle sweeps are bit-identical to the GPU's plain sweeps, so the two differ
only in the summation order of the Gram dot products. Convergence and the
sweep savings are checked on the granular pile.
"""
import os

import numpy as np
import pytest

from siconos_amd import numerics as N

pytestmark = pytest.mark.gpu


def _local_params(o, **kw):
    return N.hip_params(local_solver=551, local_max_iter=int(o.iparam[0]),
                        local_tol=float(o.dparam[0]), local_formulation=int(o.iparam[10]),
                        rho_strategy=int(o.iparam[9]), rho_constant=float(o.dparam[3]),
                        ls_max_iter=int(o.iparam[12]), **kw)


def anderson_mirror(sweep, K, m, every, reg, n3, counts=None):
    """x_{k+1} = G(x_k) - dG gamma on every `every`-th sweep, history of m
    differences restarted when |f| grows; a step whose next sweep has a larger
    |f| than the sweep it extrapolated is undone (back to G(x_k)), the history
    restarts and steps pause for a doubling back-off; the last sweep takes no
    step (the solve stops there)."""
    r = np.zeros(n3)
    F, G = [], []
    fprev = gprev = None
    nf_prev = np.inf
    sweeps = steps = rejects = 0
    pending, nf_ref, backoff, wait_until = False, 0.0, 0, 0
    for k in range(1, K + 1):
        gx = sweep(r)
        f = gx - r
        nf = np.sqrt(np.sum(f * f))
        if k == K:
            if counts is not None:
                counts.update(steps=steps, rejects=rejects)
            return gx
        sweeps += 1
        if pending and nf > nf_ref:
            pending = False
            rejects += 1
            r = gprev  # G(x_k) of the rejected step
            F, G = [], []
            fprev = gprev = None
            backoff = min(max(2, 2 * backoff), 64)
            wait_until = sweeps + backoff
            continue
        if pending:
            pending = False
            backoff //= 2
        reset = len(F) > 0 and nf > nf_prev
        form = fprev is not None and not reset
        if reset:
            F, G = [], []
        elif form:
            F.append(f - fprev)
            G.append(gx - gprev)
            if len(F) > m:
                F.pop(0)
                G.pop(0)
        fprev, gprev, nf_prev = f, gx, nf
        r = gx
        if form and sweeps % every == 0 and sweeps >= wait_until:
            dF, dG = np.stack(F, 1), np.stack(G, 1)
            A = dF.T @ dF
            A += reg * np.trace(A) * np.eye(len(F))
            gam = np.linalg.solve(A, dF.T @ f)
            r = gx - dG @ gam
            steps += 1
            pending, nf_ref = True, nf
    return r


@pytest.mark.parametrize("K", [2, 5, 12, 25, 60])
def test_iterates_match_the_restated_anderson_rule(oracle, K):
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    o = oracle.Options()
    o.set_local_solver(551)
    lo = o.internal()
    H = N.HipProblem(g.W, g.q, g.mu)
    order = H.order()

    def sweep(r):
        info, r1, u, it, res = oracle.nsgs(g.W, g.q, g.mu, reaction=r, options=o, tol=1e-300,
                                           max_iter=1, order=order)
        return r1

    s = H.solve(_local_params(lo, tol=1e-300, max_iter=K, warm_start=1, accel_depth=3,
                              accel_every=2))
    assert s.iterations == K
    counts = {}
    ref = anderson_mirror(sweep, K, 3, 2, 1e-10, 3 * g.nc, counts)
    r = H.reaction()
    assert np.linalg.norm(r - ref) <= 1e-9 * np.linalg.norm(ref)
    assert (s.accel_steps, s.accel_rejects) == (counts["steps"], counts["rejects"])
    if K >= 5:
        assert s.accel_steps >= 1


def test_plain_sweeps_are_unchanged_when_acceleration_is_off(oracle):
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    H = N.HipProblem(g.W, g.q, g.mu)
    s = H.solve(N.hip_params(tol=1e-8, max_iter=1000, warm_start=1, accel_depth=0))
    info, ro, uo, it, res = oracle.nsgs(g.W, g.q, g.mu, tol=1e-8, max_iter=1000,
                                        order=H.order())
    assert s.iterations == it and s.accel_steps == 0
    np.testing.assert_array_equal(H.reaction(), ro)


@pytest.fixture(scope="module")
def pile50k(lib):
    g = N.generate_granular_pile(20000, 50000, 0.1, 0.9, 5)
    return g, N.HipProblem(g.W, g.q, g.mu)


def test_pile_converges_in_fewer_sweeps(pile50k):
    g, H = pile50k
    plain = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1))
    acc = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, accel_depth=3))
    assert plain.info == 0 and plain.residual <= 1e-6
    assert acc.info == 0 and acc.residual <= 1e-6
    print(f"\n[anderson] pile 50k: plain {plain.iterations} sweeps, accelerated "
          f"{acc.iterations} sweeps ({acc.accel_steps} steps, {acc.accel_restarts} restarts)")
    assert acc.iterations < 0.7 * plain.iterations  # measured: 579 -> 349


def test_fp32_sweeps_with_acceleration_meet_the_fp64_tolerance(pile50k):
    g, H = pile50k
    s = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, precision=1,
                             fp32_tol=1e-6, accel_depth=3))
    assert s.info == 0 and s.residual <= 1e-6 and s.fp32_sweeps > 0
    # the residual is the FP64 fc3d_compute_error of the returned reaction
    assert abs(H.compute_error() - s.residual) <= 1e-12 + 1e-9 * s.residual


def test_graph_and_direct_launches_agree(pile50k):
    g, H = pile50k
    p = dict(tol=1e-300, max_iter=40, warm_start=1, accel_depth=3, sweeps_per_launch=8)
    H.solve(N.hip_params(**p))
    r_graph = H.reaction()
    os.environ["FC3D_HIP_NO_GRAPH"] = "1"
    try:
        H.solve(N.hip_params(**p))
    finally:
        del os.environ["FC3D_HIP_NO_GRAPH"]
    np.testing.assert_array_equal(H.reaction(), r_graph)
    H.solve(N.hip_params(instrument=1, **p))
    np.testing.assert_array_equal(H.reaction(), r_graph)


def test_invalid_acceleration_options_are_refused(pile50k):
    g, H = pile50k
    for kw in (dict(accel_depth=9), dict(accel_depth=-1), dict(accel_depth=2, accel_every=0),
               dict(accel_depth=2, accel_reg=-1.0)):
        with pytest.raises(RuntimeError, match="-2"):
            H.solve(N.hip_params(**kw))


def test_driver_maps_the_acceleration_options(lib):
    g = N.generate_granular_pile(4000, 10000, 0.1, 0.9, 7)
    prob = N.FrictionContactProblem(3, g.W, g.q, g.mu)
    o = N.SolverOptions(N.SICONOS_FRICTION_3D_NSGS_HIP)
    o.dparam[N.SICONOS_DPARAM_TOL] = 1e-6
    o.iparam[N.SICONOS_IPARAM_MAX_ITER] = 3000
    o.iparam[28] = 3
    r = np.zeros(3 * g.nc)
    u = np.zeros(3 * g.nc)
    assert N.fc3d_driver(prob, r, u, o) == 0
    assert o.dparam[N.SICONOS_DPARAM_RESIDU] <= 1e-6
    it_acc = o.iparam[N.SICONOS_IPARAM_ITER_DONE]
    o.iparam[28] = 0
    r[:] = 0.0
    assert N.fc3d_driver(prob, r, u, o) == 0
    assert it_acc < o.iparam[N.SICONOS_IPARAM_ITER_DONE]
