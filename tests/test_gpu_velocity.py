"""Velocity-form sweeps (fc3d_hip_params.velocity_form with the factors of W
attached by fc3d_hip_problem_set_bodies).

In exact arithmetic the velocity form is the same Gauss-Seidel iteration as
the matrix form, because u_c - W_cc r_c = q_c + sum_{j != c} W_cj r_j when
v = v_free + M^-1 H^T r. The tests compare the two forms' iterates, which
differ only by rounding, and check that solves end at the FP64 tolerance
judged with W.
"""
import numpy as np
import pytest

from siconos_amd import numerics as N

pytestmark = pytest.mark.gpu


def _with_bodies(g):
    H = N.HipProblem(g.W, g.q, g.mu)
    H.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
    return H


@pytest.mark.parametrize("K", [1, 3, 10, 30])
def test_iterates_match_the_matrix_form(oracle, K):
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    H = _with_bodies(g)
    H.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=1))
    rv = H.reaction()
    H.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=0))
    rm = H.reaction()
    info, ro, uo, it, res = oracle.nsgs(g.W, g.q, g.mu, tol=1e-300, max_iter=K, order=H.order())
    np.testing.assert_array_equal(rm, ro)  # the matrix form is the oracle, bit for bit
    assert np.linalg.norm(rv - ro) <= 1e-10 * np.linalg.norm(ro)


@pytest.mark.parametrize("kind", ["pairs", "pile"])
def test_solves_end_at_the_fp64_tolerance(lib, kind):
    if kind == "pairs":
        g = N.generate_random_pairs(20000, 50000, 0.2, 0.5, 2)
    else:
        g = N.generate_granular_pile(20000, 50000, 0.1, 0.9, 5)
    H = _with_bodies(g)
    m = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1))
    v = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, velocity_form=1))
    assert v.info == 0 and v.residual <= 1e-6
    assert abs(v.iterations - m.iterations) <= max(3, 0.02 * m.iterations)
    assert abs(H.compute_error() - v.residual) <= 1e-12 + 1e-9 * v.residual
    f = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, velocity_form=1, precision=1,
                             fp32_tol=1e-6, accel_depth=3))
    assert f.info == 0 and f.residual <= 1e-6 and f.fp32_sweeps > 0
    print(f"\n[velocity] {kind}: matrix {m.iterations} sweeps, velocity {v.iterations}, "
          f"fp32+anderson velocity {f.iterations} ({f.accel_steps} steps)")


def test_factors_are_checked_and_dropped_by_updates(lib):
    g = N.generate_granular_pile(4000, 10000, 0.1, 0.9, 3)
    H = N.HipProblem(g.W, g.q, g.mu)
    with pytest.raises(RuntimeError, match="-3"):
        H.solve(N.hip_params(velocity_form=1))  # no factors yet
    bad = g.H.copy()
    bad[5] *= 1.5
    with pytest.raises(RuntimeError, match="-2"):
        H.set_bodies(g.nbodies, g.body_a, g.body_b, bad, g.minv, g.vfree)
    with pytest.raises(RuntimeError, match="-2"):
        H.set_bodies(g.nbodies, g.body_a, g.body_b + g.nbodies, g.H, g.minv, g.vfree)
    # v_free must give the resident q = H v_free: one body's free velocity
    # moved is refused, a rounding-level change is not
    vf = g.vfree.copy()
    vf[6 * int(g.body_a[0]) + 1] += 1e-3 * (1.0 + abs(vf[6 * int(g.body_a[0]) + 1]))
    with pytest.raises(RuntimeError, match="H v_free differs from q"):
        H.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, vf)
    H.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree * (1.0 + 1e-14))
    H.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
    with pytest.raises(RuntimeError, match="-3"):  # the learned start needs a model
        H.solve(N.hip_params(velocity_form=1, warm_start=2))
    s = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, velocity_form=1))
    assert s.info == 0
    H.update(q=g.q)  # a new q: the factors (and v_free) may no longer match
    with pytest.raises(RuntimeError, match="-3"):
        H.solve(N.hip_params(velocity_form=1))
    # attached again (new device buffers): the captured sweeps are rebuilt
    H.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
    again = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, velocity_form=1))
    assert again.iterations == s.iterations
    np.testing.assert_array_equal(H.reaction(), _fresh_reaction(g, velocity_form=1))


def _fresh_reaction(g, **kw):
    F = _with_bodies(g)
    F.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, **kw))
    return F.reaction()


@pytest.mark.parametrize("precision", [0, 1])
def test_anderson_steps_move_the_body_velocities_with_r(lib, precision):
    """An Anderson step on r moves the body velocities by the same combination
    of their own differences; after the solve, v must still equal
    v_free + M^-1 H^T r, so u = H v matches W r + q up to rounding.  The
    sweep counts of the two forms are compared over three piles: with FP32
    sweeps, Anderson's count on one pile moves by tens of percent either way
    with the last bits of the iterates (seeds 6-10: velocity/matrix from
    0.98 to 1.46 on single piles), so one pile says little."""
    tot_s = tot_m = 0
    for seed in (6, 7, 8):
        g = N.generate_granular_pile(20000, 50000, 0.1, 0.9, seed)
        H = _with_bodies(g)
        prm = N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, velocity_form=1,
                           accel_depth=3, precision=precision, fp32_tol=1e-6)
        s = H.solve(prm)
        assert s.info == 0 and s.residual <= 1e-6 and s.accel_steps > 0
        assert abs(H.compute_error() - s.residual) <= 1e-12 + 1e-9 * s.residual
        m = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, accel_depth=3,
                                 precision=precision, fp32_tol=1e-6))
        assert m.info == 0
        print(f"\n[velocity] anderson precision {precision} pile {seed}: velocity {s.iterations} "
              f"sweeps ({s.accel_steps} steps), matrix {m.iterations} ({m.accel_steps} steps)")
        tot_s += s.iterations
        tot_m += m.iterations
        # repeatable: the same solve again gives the same iterates
        s2 = H.solve(prm)
        assert s2.iterations == s.iterations
    assert tot_s <= 1.25 * tot_m + 10


@pytest.mark.parametrize("velocity_form", [0, 1])
def test_fp32_sweeps_with_their_own_local_solver(lib, velocity_form):
    """fp32_local_solver changes the local solver of the FP32 sweeps only: the
    FP64 sweeps keep the reference's hybrid solver and the FP64 tolerance
    still holds; an FP64-only solve ignores the option bit for bit."""
    g = N.generate_granular_pile(20000, 50000, 0.1, 0.9, 5)
    H = _with_bodies(g)
    base = dict(tol=1e-6, max_iter=3000, warm_start=1, velocity_form=velocity_form)
    ref = H.solve(N.hip_params(**base))
    r_ref = H.reaction()
    same = H.solve(N.hip_params(fp32_local_solver=550, **base))
    assert same.iterations == ref.iterations
    np.testing.assert_array_equal(H.reaction(), r_ref)
    for ls in (550, 551):
        s = H.solve(N.hip_params(precision=1, fp32_tol=1e-6, fp32_local_solver=ls, **base))
        assert s.info == 0 and s.residual <= 1e-6 and s.fp32_sweeps > 0
        assert abs(H.compute_error() - s.residual) <= 1e-12 + 1e-9 * s.residual
        assert s.iterations <= 1.1 * ref.iterations + 10
    with pytest.raises(RuntimeError, match="-4"):
        H.solve(N.hip_params(precision=1, fp32_local_solver=123, **base))


def test_driver_maps_the_fp32_local_solver(lib):
    g = N.generate_granular_pile(4000, 10000, 0.1, 0.9, 7)
    prob = N.FrictionContactProblem(3, g.W, g.q, g.mu)
    o = N.SolverOptions(N.SICONOS_FRICTION_3D_NSGS_HIP)
    o.dparam[N.SICONOS_DPARAM_TOL] = 1e-6
    o.iparam[N.SICONOS_IPARAM_MAX_ITER] = 3000
    o.iparam[20] = 1  # FP32 sweeps
    o.dparam[26] = 550.0
    r = np.zeros(3 * g.nc)
    u = np.zeros(3 * g.nc)
    assert N.fc3d_driver(prob, r, u, o) == 0
    assert o.dparam[N.SICONOS_DPARAM_RESIDU] <= 1e-6
    # dparam[27]: the FP32 sweeps' local tolerance reaches the solver (other
    # FP32 iterates), and the FP64 tolerance still holds
    r_default = r.copy()
    o.dparam[27] = 1e-7
    r[:] = 0.0
    assert N.fc3d_driver(prob, r, u, o) == 0
    assert o.dparam[N.SICONOS_DPARAM_RESIDU] <= 1e-6
    assert not np.array_equal(r, r_default)
    o.dparam[27] = 0.0
    o.dparam[26] = 123.0
    r[:] = 0.0
    assert N.fc3d_driver(prob, r, u, o) != 0


@pytest.mark.parametrize("precision", [0, 1])
def test_tail_launch_gives_the_same_iterates(lib, precision):
    """The last small colours run in one workgroup (k_sweep_vel_tail) instead
    of one launch each: same per-slice computation, so the same iterates and
    sweep counts bit for bit (FC3D_HIP_VEL_TAIL=0 turns the tail off).  The
    tail launch also ends the sweep with 512 or 256 threads standing for
    k_finalize_sweep's 1024: the same light errors bit for bit."""
    import os
    g = N.generate_granular_pile(20000, 50000, 0.1, 0.9, 5)
    H = _with_bodies(g)
    prm = N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, velocity_form=1,
                       precision=precision, fp32_tol=1e-6, fp32_local_solver=550)
    a = H.solve(prm)
    ra = H.reaction()
    os.environ["FC3D_HIP_VEL_TAIL"] = "0"
    try:
        b = H.solve(prm)
    finally:
        del os.environ["FC3D_HIP_VEL_TAIL"]
    assert a.iterations == b.iterations and a.info == b.info == 0
    assert a.incr_error == b.incr_error and a.residual == b.residual
    np.testing.assert_array_equal(H.reaction(), ra)
    # instrumented mode times the tail as one launch
    i = H.solve(N.hip_params(tol=1e-300, max_iter=5, warm_start=1, velocity_form=1,
                             precision=precision, fp32_tol=1e-300, instrument=1))
    assert i.sweep_kernel_launches < 5 * H.num_colours
    if precision:
        # one tail launch a sweep, reported apart from the per-colour launches
        assert i.sweep_tail_launches_fp32 == i.fp32_sweeps == 5
        assert 0 < i.sweep_tail_ms_fp32 < i.sweep_kernel_ms_fp32
        assert 0 < i.sweep_tail_bytes_fp32 < i.sweep_kernel_bytes_fp32


def test_replay_length_does_not_change_a_mixed_solve(lib):
    """Sweeps are replayed in graphs of sweeps_per_launch sweeps (the FP64
    sweeps that end a mixed solve at most 4 at a time); the sweeps launched
    after the stop do nothing, so every replay length gives the same
    iterates, hand-over sweep and result bit for bit."""
    g = N.generate_granular_pile(20000, 50000, 0.1, 0.9, 5)
    H = _with_bodies(g)
    out = []
    for spl in (1, 3, 16, 32):
        s = H.solve(N.hip_params(tol=1e-6, max_iter=3000, warm_start=1, velocity_form=1,
                                 precision=1, fp32_tol=1e-6, fp32_local_solver=550,
                                 fp32_local_tol=1e-7, sweeps_per_launch=spl))
        assert s.info == 0 and s.residual <= 1e-6
        out.append((s.iterations, s.fp32_sweeps, s.residual, H.reaction()))
    for it, f32, res, r in out[1:]:
        assert (it, f32, res) == out[0][:3]
        np.testing.assert_array_equal(r, out[0][3])


def test_isotropic_bodies_keep_two_values_and_the_same_bits(lib):
    """Bodies whose diagonal M^-1 is (m_t, m_t, m_t, m_r, m_r, m_r) (the
    generator's spheres) keep 2 values a body in the FP32 copy; the sweeps,
    the velocity set-up and the focus stage expand them to the same six
    numbers, so FC3D_HIP_MINV_ISO=0 (six values a body) gives the same bits."""
    import os
    g = N.generate_granular_pile(20000, 50000, 0.1, 0.9, 101)
    out = []
    for iso in ("1", "0"):
        os.environ["FC3D_HIP_MINV_ISO"] = iso
        try:
            H = _with_bodies(g)
        finally:
            del os.environ["FC3D_HIP_MINV_ISO"]
        rs = []
        for K in (1, 5):
            H.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=1,
                                 precision=1, fp32_tol=1e-300, fp32_local_solver=550,
                                 fp32_local_tol=1e-7))
            rs.append(H.reaction())
        s = H.solve(N.hip_params(tol=1e-6, max_iter=6000, warm_start=1, velocity_form=1,
                                 precision=1, fp32_tol=1e-6, fp32_local_solver=550,
                                 fp32_local_tol=1e-7, focus_passes=50, focus_every=16,
                                 focus_after=64, focus_count=128, focus_share=0.5))
        assert s.info == 0 and s.focus_rounds > 0
        rs.append(H.reaction())
        out.append((s.iterations, rs))
    assert out[0][0] == out[1][0]
    for a, b in zip(out[0][1], out[1][1]):
        np.testing.assert_array_equal(a, b)


def test_rigid_body_blocks_use_the_frames_layout_and_others_do_not(oracle):
    """Generator piles (rigid-body blocks) take the frames layout (T and the
    lever arms beside the resident W_cc) unless FC3D_HIP_FRAMES=0; random
    3x6 blocks keep the general layout.  Both layouts track the rows-of-W
    iterates (the oracle's) to rounding."""
    import os
    g = N.generate_granular_pile(3000, 8000, 0.1, 0.9, 4)
    H = _with_bodies(g)
    assert H.body_layout() == 2
    os.environ["FC3D_HIP_FRAMES"] = "0"
    try:
        H0 = _with_bodies(g)
    finally:
        del os.environ["FC3D_HIP_FRAMES"]
    assert H0.body_layout() == 1
    rng = np.random.default_rng(11)
    nb, nc = 300, 1000
    pairs = set()
    while len(pairs) < nc:
        a, b = sorted(rng.choice(nb, 2, replace=False).tolist())
        pairs.add((a, b))
    pairs = sorted(pairs)
    ba = np.array([p[0] for p in pairs], dtype=np.int32)
    bb = np.array([p[1] for p in pairs], dtype=np.int32)
    Hr = rng.normal(size=36 * nc)
    minv = rng.uniform(0.5, 2.0, size=6 * nb)
    vfree = rng.normal(size=6 * nb)
    mu = rng.uniform(0.1, 0.9, size=nc)
    W, q, _ = N.assemble(nb, ba, bb, Hr, minv, vfree)
    Pr = N.HipProblem(W, q, mu)
    Pr.set_bodies(nb, ba, bb, Hr, minv, vfree)
    assert Pr.body_layout() == 1
    for prob, kind in ((g, "pile"), (g, "pile, general layout"), (None, "random")):
        if prob is None:
            P = Pr
            WW, qq, mm = W, q, mu
        else:
            P, WW, qq, mm = (H0 if "general" in kind else H), g.W, g.q, g.mu
        for K in (3, 20):
            P.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=1))
            rv = P.reaction()
            info, ro, uo, it, res = oracle.nsgs(WW, qq, mm, tol=1e-300, max_iter=K,
                                                order=P.order())
            assert np.linalg.norm(rv - ro) <= 1e-10 * np.linalg.norm(ro), (kind, K)


def _dense_problem(nb, ba, bb, H, minv, vfree, minv_diagonal=True):
    """W = J M^-1 J^T and q = J v_free assembled densely on the host, J the
    contact Jacobian with the 3x6 blocks of H (a contact with the same body
    on both sides adds both blocks to that body's columns); M^-1 is given per
    body as 6 diagonal entries or as a 6x6 column-major block."""
    nc = len(ba)
    if minv_diagonal:
        Mi = np.diag(minv)
    else:
        Mi = np.zeros((6 * nb, 6 * nb))
        for k in range(nb):
            Mi[6 * k:6 * k + 6, 6 * k:6 * k + 6] = minv[36 * k:36 * k + 36].reshape(6, 6, order="F")
    J = np.zeros((3 * nc, 6 * nb))
    for c in range(nc):
        for side, k in enumerate((ba[c], bb[c])):
            if k >= 0:
                J[3 * c:3 * c + 3, 6 * k:6 * k + 6] += H[36 * c + 18 * side:
                                                        36 * c + 18 * side + 18].reshape(3, 6, order="F")
    W = J @ Mi @ J.T
    return N.SparseBlock3.from_dense(W), J @ vfree, W


def test_same_body_and_one_body_contacts(oracle):
    """Contacts with one body (side b = -1) and with the same body on both
    sides: the velocity form adds side b's update to side a's in registers
    (FP32) or reads it back (FP64); both must track the oracle's iterates."""
    rng = np.random.default_rng(21)
    nb, nc = 150, 450
    ba = rng.integers(0, nb, nc).astype(np.int32)
    bb = rng.integers(0, nb, nc).astype(np.int32)
    bb[:40] = -1          # one body
    bb[40:70] = ba[40:70]  # the same body on both sides
    same = bb == ba
    bb[70:][same[70:]] = (ba[70:][same[70:]] + 1) % nb
    H = rng.normal(size=36 * nc)
    minv = rng.uniform(0.5, 2.0, size=6 * nb)
    vfree = rng.normal(size=6 * nb)
    mu = rng.uniform(0.1, 0.9, size=nc)
    Wsb, q, _ = _dense_problem(nb, ba, bb, H, minv, vfree)
    P = N.HipProblem(Wsb, q, mu)
    P.set_bodies(nb, ba, bb, H, minv, vfree)
    for K in (3, 15):
        info, ro, uo, it, res = oracle.nsgs(Wsb, q, mu, tol=1e-300, max_iter=K, order=P.order())
        P.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=1))
        assert np.linalg.norm(P.reaction() - ro) <= 1e-10 * np.linalg.norm(ro), K
        P.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=1, precision=1,
                             fp32_tol=1e-300))
        assert np.linalg.norm(P.reaction() - ro) <= 1e-4 * np.linalg.norm(ro), K
    s = P.solve(N.hip_params(tol=1e-6, max_iter=5000, warm_start=1, velocity_form=1, precision=1,
                             fp32_tol=1e-6, fp32_local_solver=550))
    assert s.info == 0 and s.residual <= 1e-6
    assert abs(P.compute_error() - s.residual) <= 1e-12 + 1e-9 * s.residual


def test_frames_layout_with_a_same_body_contact(oracle):
    """A contact with the same body on both sides: its W_cc holds the cross
    terms of the two sides, which the frames layout reads with the other
    diagonal blocks (it does not rebuild W_cc), so such a problem takes the
    frames layout too and its iterates track the oracle's."""
    g = N.generate_granular_pile(300, 800, 0.1, 0.9, 4)
    ba = np.array(g.body_a, dtype=np.int32)
    bb = np.array(g.body_b, dtype=np.int32)
    c = int(np.nonzero((ba >= 0) & (bb >= 0))[0][0])
    bb[c] = ba[c]
    Wsb, q, _ = _dense_problem(g.nbodies, ba, bb, np.asarray(g.H), np.asarray(g.minv),
                               np.asarray(g.vfree))
    P = N.HipProblem(Wsb, q, g.mu)
    P.set_bodies(g.nbodies, ba, bb, g.H, g.minv, g.vfree)
    assert P.body_layout() == 2
    for K in (1, 10):
        P.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=1))
        info, ro, uo, it, res = oracle.nsgs(Wsb, q, g.mu, tol=1e-300, max_iter=K,
                                            order=P.order())
        assert np.linalg.norm(P.reaction() - ro) <= 1e-10 * np.linalg.norm(ro), K


def test_full_mass_blocks(oracle):
    """Bodies with full 6x6 M^-1 blocks (minv_diagonal = 0): the update reads
    the block after the local solve instead of six diagonal entries; the
    iterates track the oracle's in FP64 and in FP32."""
    rng = np.random.default_rng(23)
    nb, nc = 120, 360
    ba = rng.integers(0, nb, nc).astype(np.int32)
    bb = rng.integers(0, nb, nc).astype(np.int32)
    bb[:30] = -1
    bb[30:50] = ba[30:50]
    same = bb == ba
    bb[50:][same[50:]] = (ba[50:][same[50:]] + 1) % nb
    H = rng.normal(size=36 * nc)
    minv = np.empty(36 * nb)
    for k in range(nb):
        A = rng.normal(size=(6, 6))
        minv[36 * k:36 * k + 36] = (A @ A.T + 0.5 * np.eye(6)).reshape(-1, order="F")
    vfree = rng.normal(size=6 * nb)
    mu = rng.uniform(0.1, 0.9, size=nc)
    Wsb, q, _ = _dense_problem(nb, ba, bb, H, minv, vfree, minv_diagonal=False)
    P = N.HipProblem(Wsb, q, mu)
    P.set_bodies(nb, ba, bb, H, minv, vfree, minv_diagonal=False)
    for K in (3, 15):
        info, ro, uo, it, res = oracle.nsgs(Wsb, q, mu, tol=1e-300, max_iter=K, order=P.order())
        P.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=1))
        assert np.linalg.norm(P.reaction() - ro) <= 1e-10 * np.linalg.norm(ro), K
        P.solve(N.hip_params(tol=1e-300, max_iter=K, warm_start=1, velocity_form=1, precision=1,
                             fp32_tol=1e-300))
        assert np.linalg.norm(P.reaction() - ro) <= 1e-4 * np.linalg.norm(ro), K
