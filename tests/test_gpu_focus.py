"""Focused passes of the velocity form (fc3d_hip_params.focus_*): extra
Gauss-Seidel passes over the few contact clusters that still move, on a pile
whose plain sweeps spend most of their time on them.  Judged by the oracle's
fc3d_compute_error; the selection uses atomics, so the results must not
depend on their order (two solves give the same bits)."""
import ctypes as C
import os

import numpy as np
import pytest

from siconos_amd import capi
from siconos_amd import numerics as N

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pile():
    # 50k contacts whose plain sweeps need ~2400 sweeps, most of them for a
    # few clusters of 5-7 bodies
    g = N.generate_granular_pile(20_000, 50_000, 0.1, 0.9, 101)
    H = N.HipProblem(g.W, g.q, g.mu)
    H.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
    return g, H


FOCUS = dict(focus_passes=50, focus_every=16, focus_after=64, focus_count=128, focus_share=0.5)


@pytest.mark.parametrize("precision", [0, 1])
def test_focused_passes_converge_in_fewer_sweeps(pile, oracle, precision):
    g, H = pile
    base = dict(tol=1e-6, max_iter=6000, warm_start=1, velocity_form=1, precision=precision,
                fp32_tol=1e-6, fp32_local_solver=550 if precision else 0, fp32_local_tol=1e-7)
    s0 = H.solve(N.hip_params(**base))
    assert s0.info == 0
    s1 = H.solve(N.hip_params(**base, **FOCUS))
    r1 = H.reaction()
    print(f"precision {precision}: {s0.iterations} sweeps plain, {s1.iterations} focused "
          f"({s1.focus_rounds} rounds of {s1.focus_due} due, {s1.focus_listed} contacts listed, "
          f"last share {s1.focus_last_share:.3f})", flush=True)
    assert s1.info == 0 and s1.residual <= 1e-6
    assert s1.focus_rounds > 0
    assert s1.iterations < 0.5 * s0.iterations
    e, _ = oracle.compute_error(g.W, g.q, g.mu, r1)
    assert e <= 1e-6 and e == pytest.approx(s1.residual, rel=1e-9)
    s2 = H.solve(N.hip_params(**base, **FOCUS))
    assert s2.iterations == s1.iterations and s2.focus_listed == s1.focus_listed
    np.testing.assert_array_equal(H.reaction(), r1)


@pytest.mark.parametrize("omega", [1.0, 1.3])
def test_focus_relaxation(pile, oracle, omega):
    """focus_omega relaxes the focus passes' steps only; without it (1.0) or
    with more of it the solve still converges, and gives the same bits twice."""
    g, H = pile
    base = dict(tol=1e-6, max_iter=6000, warm_start=1, velocity_form=1, precision=1,
                fp32_tol=1e-6, fp32_local_solver=550, fp32_local_tol=1e-7)
    s1 = H.solve(N.hip_params(**base, **FOCUS, focus_omega=omega))
    r1 = H.reaction()
    print(f"focus_omega {omega}: {s1.iterations} sweeps, {s1.focus_rounds} rounds", flush=True)
    assert s1.info == 0 and s1.residual <= 1e-6 and s1.focus_rounds > 0
    e, _ = oracle.compute_error(g.W, g.q, g.mu, r1)
    assert e <= 1e-6 and e == pytest.approx(s1.residual, rel=1e-9)
    s2 = H.solve(N.hip_params(**base, **FOCUS, focus_omega=omega))
    assert s2.iterations == s1.iterations
    np.testing.assert_array_equal(H.reaction(), r1)


@pytest.mark.parametrize("precision", [0, 1])
def test_velocities_in_lds_give_the_same_bits(pile, precision):
    """The passes keep the round's body velocities in LDS; with
    FC3D_HIP_FOCUS_HBM=1 they read and write them in HBM instead.  The
    arithmetic is the same, so are the results."""
    g, H = pile
    base = dict(tol=1e-6, max_iter=6000, warm_start=1, velocity_form=1, precision=precision,
                fp32_tol=1e-6, fp32_local_solver=550 if precision else 0, fp32_local_tol=1e-7)
    s1 = H.solve(N.hip_params(**base, **FOCUS))
    r1 = H.reaction()
    os.environ["FC3D_HIP_FOCUS_HBM"] = "1"
    try:
        s2 = H.solve(N.hip_params(**base, **FOCUS))
    finally:
        del os.environ["FC3D_HIP_FOCUS_HBM"]
    r2 = H.reaction()
    assert s1.info == 0 and s1.focus_rounds > 0
    assert (s2.iterations, s2.focus_rounds, s2.focus_listed) == \
        (s1.iterations, s1.focus_rounds, s1.focus_listed)
    np.testing.assert_array_equal(r2, r1)


def test_rounds_beyond_the_lds_keep_their_velocities_in_hbm(pile):
    """A round whose bodies exceed the LDS capacity (6144 FP32 velocities)
    runs its passes with the velocities in HBM: with 1024 seed contacts the
    rounds list about ten thousand bodies.  Those rounds give the same bits
    as the forced HBM path, and the default selection stays in LDS."""
    g, H = pile
    base = dict(tol=1e-6, max_iter=6000, warm_start=1, velocity_form=1, precision=1,
                fp32_tol=1e-6, fp32_local_solver=550, fp32_local_tol=1e-7)
    big = dict(FOCUS, focus_count=1024)
    s1 = H.solve(N.hip_params(**base, **big))
    r1 = H.reaction()
    print(f"1024 seeds: {s1.focus_rounds} rounds, {s1.focus_lds_rounds} in LDS, "
          f"{s1.focus_listed} contacts listed", flush=True)
    assert s1.info == 0 and s1.focus_rounds > 0
    assert s1.focus_lds_rounds < s1.focus_rounds
    os.environ["FC3D_HIP_FOCUS_HBM"] = "1"
    try:
        s2 = H.solve(N.hip_params(**base, **big))
    finally:
        del os.environ["FC3D_HIP_FOCUS_HBM"]
    assert s2.focus_lds_rounds == 0 and s2.iterations == s1.iterations
    np.testing.assert_array_equal(H.reaction(), r1)
    s3 = H.solve(N.hip_params(**base, **FOCUS))
    assert s3.focus_rounds > 0 and s3.focus_lds_rounds == s3.focus_rounds


def test_bodies_set_again_with_more_bodies(pile):
    """set_bodies again with a larger body count (the extra bodies carry no
    contact): the focus rounds' per-body arrays follow the new count, and
    the solve gives the bits of the first set-up."""
    g, H = pile
    base = dict(tol=1e-6, max_iter=6000, warm_start=1, velocity_form=1, precision=1,
                fp32_tol=1e-6, fp32_local_solver=550, fp32_local_tol=1e-7)
    s1 = H.solve(N.hip_params(**base, **FOCUS))
    r1 = H.reaction()
    extra = 50_000
    nb = g.nbodies + extra
    minv = np.concatenate([g.minv, np.ones(6 * extra)])
    vfree = np.concatenate([g.vfree, np.zeros(6 * extra)])
    H2 = N.HipProblem(g.W, g.q, g.mu)
    H2.set_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
    H2.solve(N.hip_params(**base, **FOCUS))
    H2.set_bodies(nb, g.body_a, g.body_b, g.H, minv, vfree)
    s2 = H2.solve(N.hip_params(**base, **FOCUS))
    assert s2.info == 0 and s2.focus_rounds == s1.focus_rounds
    assert s2.iterations == s1.iterations
    np.testing.assert_array_equal(H2.reaction(), r1)


@pytest.fixture(scope="module")
def mixed_pile(pile):
    """The pile with one contact in ten given a single body (side b = -1)
    and another one in ten the same body on both sides, W and q assembled on
    the device from the bodies (fc3d_hip_assemble)."""
    g, _ = pile
    ba = np.array(g.body_a, dtype=np.int32)
    bb = np.array(g.body_b, dtype=np.int32)
    bb[0::10] = -1
    bb[5::10] = ba[5::10]
    W, q, _ = N.assemble(g.nbodies, ba, bb, g.H, g.minv, g.vfree)
    H = N.HipProblem(W, q, g.mu)
    H.set_bodies(g.nbodies, ba, bb, g.H, g.minv, g.vfree)
    return W, q, g.mu, H


@pytest.mark.parametrize("precision", [0, 1])
def test_one_body_and_same_body_contacts(mixed_pile, oracle, precision):
    """Focused passes over contacts with one body and with the same body on
    both sides (a fifth of the contacts, spread evenly, so every round's
    hundreds of listed contacts hold both kinds): the step skips side b or
    adds its update to side a's, in LDS as in HBM.  The solve converges by
    the oracle's error and both paths give the same bits."""
    W, q, mu, H = mixed_pile
    base = dict(tol=1e-6, max_iter=6000, warm_start=1, velocity_form=1, precision=precision,
                fp32_tol=1e-6, fp32_local_solver=550 if precision else 0, fp32_local_tol=1e-7)
    s1 = H.solve(N.hip_params(**base, **FOCUS))
    r1 = H.reaction()
    print(f"precision {precision}: {s1.iterations} sweeps, {s1.focus_rounds} rounds "
          f"({s1.focus_lds_rounds} in LDS), {s1.focus_listed} contacts listed", flush=True)
    assert s1.info == 0 and s1.residual <= 1e-6
    assert s1.focus_rounds > 0 and s1.focus_lds_rounds == s1.focus_rounds
    assert s1.focus_listed >= 100 * s1.focus_rounds
    # the FP32 local solver leaves a contact within its tolerance untouched,
    # so late rounds see fewer than focus_count contacts move; they take
    # those that did (a selection that skips such rounds runs 61 of 346 in
    # FP32 here, and the solve takes 5588 sweeps instead of about 1800)
    assert s1.focus_rounds >= 0.9 * s1.focus_due
    assert s1.iterations <= 2500
    e, _ = oracle.compute_error(W, q, mu, r1)
    assert e <= 1e-6 and e == pytest.approx(s1.residual, rel=1e-9)
    os.environ["FC3D_HIP_FOCUS_HBM"] = "1"
    try:
        s2 = H.solve(N.hip_params(**base, **FOCUS))
    finally:
        del os.environ["FC3D_HIP_FOCUS_HBM"]
    assert s2.focus_lds_rounds == 0
    assert (s2.iterations, s2.focus_rounds, s2.focus_listed) == \
        (s1.iterations, s1.focus_rounds, s1.focus_listed)
    np.testing.assert_array_equal(H.reaction(), r1)


@pytest.mark.parametrize("omega", [0.0, 2.0, float("nan")])
def test_focus_omega_outside_0_2_is_refused(pile, omega):
    g, H = pile
    p = N.hip_params(velocity_form=1, **FOCUS, focus_omega=omega)
    rc = capi.load().fc3d_hip_solve(H._h, C.byref(p), C.byref(capi.HipStats()))
    assert rc == -2
    assert b"focus_omega" in capi.load().fc3d_hip_last_error()


def test_focused_passes_need_the_velocity_form(pile):
    g, H = pile
    rc = capi.load().fc3d_hip_solve(H._h, C.byref(N.hip_params(**FOCUS)),
                                    C.byref(capi.HipStats()))
    assert rc == -2
