"""Time-stepping use of a resident problem: in-place value updates
(fc3d_hip_problem_update) and the previous step's reaction as the start, as
the reference's LinearOSNS keeps it between steps (LinearOSNS.cpp:873-891)."""
import numpy as np
import pytest

from siconos_amd import numerics as N

pytestmark = pytest.mark.gpu


def _step(g, k, scale=1e-3):
    """A small change of the free velocity, as between two time steps."""
    rng = np.random.default_rng(100 + k)
    return g.q + scale * np.linalg.norm(g.q) / np.sqrt(len(g.q)) * rng.normal(size=len(g.q))


@pytest.mark.parametrize("precision", [0, 1])
def test_updated_problem_equals_a_fresh_one(lib, precision):
    g = N.generate_granular_pile(8000, 20000, 0.1, 0.9, 11)
    prm = N.hip_params(tol=1e-6, max_iter=2000, warm_start=1, precision=precision,
                       fp32_tol=1e-6)
    H = N.HipProblem(g.W, g.q, g.mu)
    H.solve(prm)  # allocates the FP32 copies when precision = 1
    values2 = g.W.values * 1.01
    q2 = _step(g, 0, 0.1)
    mu2 = np.clip(g.mu * 0.9, 0.05, None)
    H.update(values2, q2, mu2)
    s1 = H.solve(prm)
    W2 = N.SparseBlock3(g.W.row_ptr, g.W.col_idx, values2)  # (nnzb, 9): column-major as given
    np.testing.assert_array_equal(W2.values, values2)
    F = N.HipProblem(W2, q2, mu2)
    s2 = F.solve(prm)
    assert s1.iterations == s2.iterations and s1.info == s2.info == 0
    np.testing.assert_array_equal(H.reaction(), F.reaction())
    np.testing.assert_array_equal(H.velocity(), F.velocity())
    # partial updates keep the rest
    H.update(q=g.q)
    H.update(values=g.W.values, mu=g.mu)
    s3 = H.solve(prm)
    G = N.HipProblem(g.W, g.q, g.mu)
    s4 = G.solve(prm)
    assert s3.iterations == s4.iterations
    np.testing.assert_array_equal(H.reaction(), G.reaction())


def test_previous_step_start_and_its_learned_fallback(lib):
    from siconos_amd import warmstart
    g = N.generate_granular_pile(20000, 50000, 0.1, 0.9, 12)
    H = N.HipProblem(g.W, g.q, g.mu)
    s0 = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1))
    assert s0.info == 0
    r_prev = H.reaction()
    # next step: slightly different free velocity, same contacts
    H.update(q=_step(g, 1))
    zero = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1))
    H.set_reaction(r_prev)
    prev = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=0))
    assert zero.info == 0 and prev.info == 0 and prev.residual <= 1e-6
    print(f"\n[timestep] zero start {zero.iterations} sweeps, previous step {prev.iterations}")
    assert prev.iterations < 0.5 * zero.iterations
    # learned start against the previous step's reaction: the previous step wins
    x, t, _ = warmstart.collect([(g.W, g.q, g.mu)], tol=1e-6)
    model, _ = warmstart.train(x, t, seed=3, epochs=5, batch=1024)
    H.set_model(model)
    H.set_reaction(r_prev)
    s = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=3))
    assert s.info == 0 and s.residual <= 1e-6
    assert s.warm_start_used == 0
    assert s.start_residual_fallback < s.start_residual_learned
    assert s.iterations <= prev.iterations + 1
