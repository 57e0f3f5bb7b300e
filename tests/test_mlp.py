"""The learned warm start: the network's kernels against an FP64 CPU reference,
training, and what the trained start gains inside the solver.

CPU tests pin the reference itself, with finite differences and a planted
mistake. GPU tests call the library through its C ABI.
"""
import numpy as np
import pytest

import mlp_reference as R


def _random_params(rng, scale=1.0):
    Ws, bs = [], []
    for i, (n, k) in enumerate(R.SHAPES):
        fan = R.NF if i == 0 else R.H
        W = rng.uniform(-1, 1, (n, k)) * np.sqrt(6.0 / fan) * scale
        if i == 0:
            W[:, R.NF:] = 0.0
        if i == 3:
            W[R.NO:, :] = 0.0
        b = rng.normal(0, 0.1, n)
        if i == 3:
            b[R.NO:] = 0.0
        Ws.append(W)
        bs.append(b)
    return R.pack(Ws, bs)


def _fd_check(mutation=None, seed=0):
    rng = np.random.default_rng(seed)
    p = _random_params(rng)
    x = rng.normal(0, 1, (6, R.NF))
    t = rng.normal(0, 1, (6, R.NO))
    mean, std = rng.normal(0, 0.3, R.NF), rng.uniform(0.5, 2.0, R.NF)
    _, g = R.loss_and_grads(p, x, t, mean, std, mutation=mutation)
    worst = 0.0
    for name, sl in R.block_slices():
        idx = rng.choice(np.arange(sl.start, sl.stop), 12, replace=False)
        for i in idx:
            h = 1e-5
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            fd = (R.loss_and_grads(pp, x, t, mean, std)[0] -
                  R.loss_and_grads(pm, x, t, mean, std)[0]) / (2 * h)
            worst = max(worst, abs(fd - g[i]) / (1e-6 + abs(fd)))
    return worst


def test_reference_gradients_match_finite_differences():
    assert _fd_check() < 1e-5


def test_finite_differences_catch_a_wrong_gelu_derivative():
    assert _fd_check(mutation="gelu_grad") > 1e-2


def test_bf16_rounding_is_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7  # BF16 spacing at 1
    assert R.bf16(one + ulp / 2) == one  # tie -> even (1.0)
    assert R.bf16(one + 3 * ulp / 2) == one + 2 * ulp  # tie -> even (1 + 2 ulp)
    assert R.bf16(one + ulp / 2 + 2 ** -20) == one + ulp
    assert R.bf16(-3.0) == -3.0


# ---------------------------------------------------------------- GPU -----

def _gpu_model_with(params, mean, std, seed=0):
    from siconos_amd import numerics as N
    m = N.Mlp(seed=seed)
    m.set_params(params.astype(np.float32), mean.astype(np.float32), std.astype(np.float32))
    return m


def _inputs(rng, n):
    x = rng.normal(0, 1, (n, R.NF)) * rng.uniform(0.5, 3.0, R.NF) + rng.normal(0, 1, R.NF)
    mean = x.mean(axis=0)
    std = x.std(axis=0)
    return x.astype(np.float32), mean.astype(np.float32), std.astype(np.float32)


@pytest.mark.gpu
def test_forward_matches_the_reference(lib):
    rng = np.random.default_rng(1)
    p = _random_params(rng).astype(np.float32)
    x, mean, std = _inputs(rng, 1000)  # not a multiple of 64: padded rows
    m = _gpu_model_with(p, mean, std)
    y = m.forward(x).astype(np.float64)
    ye, _ = R.forward(p, x, mean, std, emulate=True)
    y64, _ = R.forward(p, x, mean, std, emulate=False)
    rms = np.sqrt((ye ** 2).mean())
    # BF16 rounding emulated: only FP32 accumulation order and rare ties differ
    assert np.abs(y - ye).max() < 2e-3 * rms + 1e-6
    # against plain FP64: the BF16 operands' own error
    assert np.sqrt(((y - y64) ** 2).mean()) < 3e-2 * rms


@pytest.mark.gpu
@pytest.mark.parametrize("n", [64, 500])
def test_gradients_match_the_cpu_fp64_reference(lib, n):
    rng = np.random.default_rng(2 + n)
    p = _random_params(rng).astype(np.float32)
    x, mean, std = _inputs(rng, n)
    t = rng.normal(0, 1, (n, R.NO)).astype(np.float32)
    m = _gpu_model_with(p, mean, std)
    g, loss = m.gradients(x, t)
    le, ge = R.loss_and_grads(p, x, t, mean, std, emulate=True)
    l64, g64 = R.loss_and_grads(p, x, t, mean, std, emulate=False)
    assert abs(loss - le) < 1e-4 * le
    assert abs(loss - l64) < 2e-2 * l64
    for name, sl in R.block_slices():
        ref = np.linalg.norm(ge[sl])
        ref64 = np.linalg.norm(g64[sl])
        if name in ("b4",):
            assert np.abs(g[sl][R.NO:]).max() == 0.0  # unused outputs get no gradient
        if name == "W1":
            w1 = g[sl].reshape(R.H, R.K1)
            assert np.abs(w1[:, R.NF:]).max() == 0.0  # padded inputs get no gradient
        assert np.linalg.norm(g[sl] - ge[sl]) <= 1e-2 * ref + 1e-9, name
        assert np.linalg.norm(g[sl] - g64[sl]) <= 5e-2 * ref64 + 1e-9, name


@pytest.mark.gpu
def test_training_fits_a_smooth_map(lib):
    from siconos_amd import numerics as N
    rng = np.random.default_rng(5)
    n = 8192
    x = rng.normal(0, 1, (n, R.NF)).astype(np.float32)
    A = rng.normal(0, 0.4, (R.NF, R.NO))
    t = np.tanh(x @ A).astype(np.float32)
    m = N.Mlp(seed=7)
    _, l0 = m.gradients(x[:4096], t[:4096])
    m.train(x, t, epochs=25, batch=256, lr=2e-3, seed=3)
    _, l1 = m.gradients(x[:4096], t[:4096])
    assert l1 < l0 / 20, (l0, l1)
    # the trained weights of padded inputs and unused outputs stay zero
    p, _, _ = m.params()
    W1 = p[:R.H * R.K1].reshape(R.H, R.K1)
    assert np.abs(W1[:, R.NF:]).max() == 0.0


def _cfg2_problem(seed, nc=50_000, nbodies=20_000):
    from siconos_amd import numerics as N
    g = N.generate_random_pairs(nbodies, nc, 0.2, 0.5, seed)
    return g.W, g.q, g.mu


@pytest.fixture(scope="module")
def trained_cfg2_model(lib):
    """Trained on 200k converged contacts of four 50k-contact problems."""
    from siconos_amd import warmstart
    x, t, sweeps = warmstart.collect([_cfg2_problem(100 + k) for k in range(4)], tol=1e-6)
    assert len(x) == 200_000
    model, loss = warmstart.train(x, t, seed=11, epochs=30, batch=1024, lr=1e-3)
    return model, loss, sweeps


@pytest.mark.gpu
@pytest.mark.parametrize("trial", [1, 8])
def test_learned_warm_start_ablation(trained_cfg2_model, trial):
    """Ablation on fresh seeded 50k-contact problems: the same solver from the
    zero start and with the learned start, each start given `trial` sweeps
    before the lower FP64 residual is kept.

    What the network gains is a lower residual over the first sweeps on every
    fresh problem, so the trial keeps it.  Measured here and on the 1M pile,
    that gain does not carry to the sweep count at 1e-6: the network removes
    local, fast-decaying error, while the sweeps are set by a few slowly
    drifting contact clusters that no per-contact prediction reaches.  The
    test therefore asserts the residual gain and the trial's bookkeeping, and
    prints the sweep counts rather than asserting a gain there is not."""
    from siconos_amd import numerics as N
    model, loss, train_sweeps = trained_cfg2_model
    report = []
    for seed in (201, 202):
        W, q, mu = _cfg2_problem(seed)
        H = N.HipProblem(W, q, mu)
        H.set_model(model)
        zero = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1))
        learned = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=2,
                                       start_trial_sweeps=trial))
        assert zero.info == 0 and zero.residual <= 1e-6
        assert learned.info == 0 and learned.residual <= 1e-6
        report.append((seed, zero.iterations, learned.iterations, learned.warm_start_used,
                       learned.start_residual_fallback, learned.start_residual_learned))
        # the kept start is the one with the lower residual after the trial,
        # and the other start's trial sweeps are reported apart
        kept_learned = learned.start_residual_learned < learned.start_residual_fallback
        assert learned.warm_start_used == (2 if kept_learned else 1)
        assert learned.start_trial_sweeps == trial
    print("\n[ablation] trial %d sweeps, train loss %.4f, training sweeps %s"
          % (trial, loss, train_sweeps))
    for r in report:
        print("[ablation] seed %d: zero start %d sweeps, learned start %d sweeps (used %d), "
              "residual after the trial: zero %.3e learned %.3e" % r)
    assert all(r[3] == 2 and r[5] < 0.9 * r[4] for r in report)


@pytest.mark.gpu
def test_learned_start_in_the_velocity_form(trained_cfg2_model):
    """The learned start with the velocity form and FP32 sweeps: the trial
    sweeps run the velocity-form kernels, the body velocities are rebuilt
    from the kept reaction, and the FP64 residual judges the result."""
    from siconos_amd import numerics as N
    model = trained_cfg2_model[0]
    g = N.generate_random_pairs(20_000, 50_000, 0.2, 0.5, 205)
    H = N.HipProblem(g.W, g.q, g.mu)
    H.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
    H.set_model(model)
    kw = dict(tol=1e-6, max_iter=3000, velocity_form=1, precision=1, fp32_tol=1e-6,
              fp32_local_solver=550, fp32_local_tol=1e-7)
    zero = H.solve(N.hip_params(warm_start=1, **kw))
    s = H.solve(N.hip_params(warm_start=2, start_trial_sweeps=4, **kw))
    assert s.info == 0 and s.residual <= 1e-6
    assert s.warm_start_used in (1, 2) and s.start_trial_sweeps == 4
    assert s.iterations >= 4
    print(f"\nvelocity form: zero start {zero.iterations} sweeps, learned start "
          f"{s.iterations} (used {s.warm_start_used})")


@pytest.mark.gpu
def test_fallback_keeps_the_zero_start_bit_for_bit(lib):
    """A network whose start is worse is discarded after one sweep; the solve
    then equals the zero-start solve exactly."""
    from siconos_amd import numerics as N
    W, q, mu = _cfg2_problem(300, nc=5000, nbodies=2000)
    m = N.Mlp(seed=1)
    p, mean, std = m.params()
    p[-R.NOP:-R.NOP + 3] = [40.0, 25.0, -25.0]  # b4: a huge constant reaction
    m.set_params(p)
    H = N.HipProblem(W, q, mu)
    H.set_model(m)
    zero = H.solve(N.hip_params(tol=1e-8, max_iter=1000, warm_start=1))
    r0 = H.reaction()
    fb = H.solve(N.hip_params(tol=1e-8, max_iter=1000, warm_start=2))
    assert fb.warm_start_used == 1
    assert fb.start_residual_learned > fb.start_residual_fallback
    assert fb.iterations == zero.iterations
    assert np.array_equal(H.reaction(), r0)


@pytest.mark.gpu
def test_learned_start_with_fp32_sweeps(trained_cfg2_model):
    from siconos_amd import numerics as N
    model = trained_cfg2_model[0]
    W, q, mu = _cfg2_problem(203)
    H = N.HipProblem(W, q, mu)
    H.set_model(model)
    s = H.solve(N.hip_params(tol=1e-6, max_iter=2000, warm_start=2, precision=1))
    assert s.info == 0 and s.residual <= 1e-6
    assert s.fp32_sweeps >= 1


@pytest.mark.gpu
def test_warm_start_2_needs_a_model(lib):
    from siconos_amd import numerics as N
    W, q, mu = _cfg2_problem(301, nc=500, nbodies=200)
    H = N.HipProblem(W, q, mu)
    with pytest.raises(RuntimeError, match="-3"):
        H.solve(N.hip_params(warm_start=2))


@pytest.mark.gpu
def test_driver_uses_the_default_model(trained_cfg2_model):
    from siconos_amd import numerics as N
    model = trained_cfg2_model[0]
    W, q, mu = _cfg2_problem(204, nc=20_000, nbodies=8_000)
    prob = N.FrictionContactProblem(3, W, q, mu)
    o = N.SolverOptions(N.SICONOS_FRICTION_3D_NSGS_HIP)
    o.dparam[N.SICONOS_DPARAM_TOL] = 1e-6
    o.iparam[N.SICONOS_IPARAM_MAX_ITER] = 2000
    o.iparam[N.SICONOS_FRICTION_3D_NSGS_HIP_WARM_START] = 2
    r = np.zeros(3 * W.nc)
    u = np.zeros(3 * W.nc)
    model.set_default()
    try:
        info = N.fc3d_driver(prob, r, u, o)
    finally:
        N.capi.load().fc3d_hip_set_default_model(None)
    assert info == 0
    assert o.iparam[27] in (1, 2)
    assert o.dparam[N.SICONOS_DPARAM_RESIDU] <= 1e-6
