"""The drop-in boundary through fc3d_driver (solver FC3D_NSGS_HIP): contact
freezing, the per-sweep callback, NM_SPARSE input and the device problem kept
resident between calls, against the oracle run in the GPU's colour order."""
import ctypes as C

import numpy as np
import pytest

from conftest import load_fixture
from siconos_amd import capi
from siconos_amd import numerics as N

pytestmark = pytest.mark.gpu


def hip_options_like(ref):
    """FC3D_NSGS_HIP options carrying the oracle options' NSGS and local
    parameters (slots 0-19 keep the reference meaning)."""
    o = N.SolverOptions(540)
    o.iparam[:20] = ref.iparam[:20]
    o.dparam[:20] = ref.dparam[:20]
    o.update_internal(0, ref.internal()._s.solverId)
    o.internal().iparam[:] = ref.internal().iparam[:]
    o.internal().dparam[:] = ref.internal().dparam[:]
    return o


def colour_order(W, q, mu):
    return N.HipProblem(W, q, mu).order()


@pytest.mark.parametrize("variant", range(5))
@pytest.mark.parametrize("name", ["Capsules-i100-1090.dat", "Capsules-i100-889.dat",
                                  "Capsules-i101-404.dat", "Capsules-i103-990.dat",
                                  "Capsules-i122-1617.dat"])
def test_freeze_collection_through_fc3d_driver(oracle, name, variant):
    """test_nsgs_freeze_1.c (freezing 10, five option sets, data_collection_2.c)
    bit for bit against the oracle in colour order."""
    from test_oracle_nsgs import freeze_options
    W, q, mu = load_fixture(name)
    o = hip_options_like(freeze_options(oracle, variant))
    assert o.iparam[15] == 10
    p = N.FrictionContactProblem(3, W, q, mu)
    r, u = np.zeros(3 * W.nc), np.zeros(3 * W.nc)
    info = N.fc3d_driver(p, r, u, o)
    info_o, ro, uo, it, res = oracle.nsgs(W, q, mu, options=freeze_options(oracle, variant),
                                          order=colour_order(W, q, mu))
    assert info == info_o == 0
    assert o.iparam[1] == it
    np.testing.assert_array_equal(r, ro)
    np.testing.assert_array_equal(u, uo)


@pytest.mark.parametrize("schedule", [1, 7])
def test_freeze_on_a_generated_problem(oracle, schedule):
    """Freezing 5 on 20k contacts (random pairs), per-colour and single-launch
    schedules, against the oracle in colour order."""
    g = N.generate_random_pairs(8000, 20_000, 0.2, 0.5, 7)
    H = N.HipProblem(g.W, g.q, g.mu)
    H.set_reaction(np.zeros(3 * g.nc))
    s = H.solve(N.hip_params(tol=1e-6, max_iter=2000, freeze_contacts=5, schedule=schedule))
    o = oracle.Options()
    o.iparam[15] = 5
    info, ro, uo, it, res = oracle.nsgs(g.W, g.q, g.mu, options=o, tol=1e-6, max_iter=2000,
                                        order=H.order())
    assert s.info == info == 0 and s.iterations == it
    np.testing.assert_array_equal(H.reaction(), ro)
    np.testing.assert_array_equal(H.velocity(), uo)
    # freezing changes the iteration: the plain solve takes another path
    s2 = H.solve(N.hip_params(tol=1e-6, max_iter=2000, warm_start=1, schedule=schedule))
    assert not np.array_equal(H.reaction(), ro)
    assert s2.info == 0


def test_freeze_refused_where_not_restated():
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    H = N.HipProblem(g.W, g.q, g.mu)
    rc = capi.load().fc3d_hip_solve(H._h, C.byref(N.hip_params(freeze_contacts=3, precision=1)),
                                    C.byref(capi.HipStats()))
    assert rc == -5


def test_callback_sees_every_sweep(oracle):
    """collectStatsIteration (fc3d_nsgs.c:103-110) after every sweep: the same
    reactions, bit for bit, as the oracle's callback in colour order."""
    W, q, mu = load_fixture("Capsules-i101-404.dat")
    from test_oracle_nsgs import _options
    got = []

    def cb(env, size, r, u, err, extra):
        got.append((np.ctypeslib.as_array(r, shape=(size,)).copy(), err))

    fn = capi.CALLBACK_FN(cb)
    cbs = capi.Callback(None, fn)
    o = hip_options_like(_options(oracle, 0))
    o._p.contents.callback = C.pointer(cbs)
    p = N.FrictionContactProblem(3, W, q, mu)
    r, u = np.zeros(3 * W.nc), np.zeros(3 * W.nc)
    assert N.fc3d_driver(p, r, u, o) == 0
    o._p.contents.callback = None
    ref = _options(oracle, 0)
    want = []
    ref.set_callback(lambda rr, uu, err: want.append((rr, err)))
    info, ro, uo, it, res = oracle.nsgs(W, q, mu, options=ref, order=colour_order(W, q, mu))
    ref.set_callback(None)
    assert len(got) == len(want) == it == o.iparam[1]
    for k, ((rg, eg), (rw, ew)) in enumerate(zip(got, want)):
        np.testing.assert_array_equal(rg, rw, err_msg=f"sweep {k + 1}")
        assert eg == pytest.approx(ew, rel=1e-12)


def _write_sparse_dat(path, Wd, q, mu):
    """A 3-D problem with W in NM_SPARSE (CSC) storage, in the .dat format of
    FrictionContactProblem.c:67-158 and NumericsSparseMatrix.c:555-583."""
    import scipy.sparse as sp
    n = Wd.shape[0]
    A = sp.csc_matrix(Wd)
    A.eliminate_zeros()
    A.sort_indices()
    lines = ["3", str(n // 3), "2", str(n), str(n), "2",
             "CSparse Version 3.1.9, May 4, 2016.  Copyright (c) Timothy A. Davis, 2006-2016",
             f"{n}-by-{n}, nzmax: {A.nnz} nnz: {A.nnz}, 1-norm: {abs(A).sum(axis=0).max():.6g}"]
    for j in range(n):
        a, b = A.indptr[j], A.indptr[j + 1]
        lines.append(f"    col {j} : locations {a} to {b - 1}")
        for k in range(a, b):
            lines.append(f"      {A.indices[k]} : {A.data[k]:.17g}")
    lines += [f"{x:.17g}" for x in q] + [f"{x:.17g}" for x in mu]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def _blocks_ascending(Wd):
    """CSR over 3x3 blocks of a dense W: the non-zero blocks of each block row
    by ascending block column, the diagonal block always."""
    nc = Wd.shape[0] // 3
    rp, ci, vals = [0], [], []
    for i in range(nc):
        for j in range(nc):
            b = Wd[3 * i:3 * i + 3, 3 * j:3 * j + 3]
            if i == j or np.any(b != 0):
                ci.append(j)
                vals.append(b.reshape(-1, order="F"))
        rp.append(len(ci))
    return N.SparseBlock3(np.array(rp, dtype=np.int64), np.array(ci, dtype=np.int64),
                          np.array(vals))


@pytest.mark.parametrize("name", ["Capsules-i101-404.dat", "Capsules-i100-889.dat",
                                  "FC3D_Example1_SBM.dat", "Rover4396.dat"])
def test_sparse_storage_through_fc3d_driver(oracle, tmp_path, name):
    """NM_SPARSE input (fc3d_nsgs.c:613-620): the reference's sparse row
    product adds each row's stored entries by ascending column, which is the
    order of column-sorted 3x3 blocks, so the result equals the oracle on
    those blocks bit for bit."""
    W, q, mu = load_fixture(name)
    Wd = W.to_dense()
    path = tmp_path / "sparse.dat"
    _write_sparse_dat(path, Wd, q, mu)
    p = N.FrictionContactProblem.from_file(path)
    assert p.storage == 2
    o = N.SolverOptions(540)
    o.dparam[0] = 1e-5
    r, u = np.zeros(3 * W.nc), np.zeros(3 * W.nc)
    info = N.fc3d_driver(p, r, u, o)
    B = _blocks_ascending(Wd)
    ref = oracle.Options()
    ref.dparam[0] = 1e-5
    info_o, ro, uo, it, res = oracle.nsgs(B, q, mu, options=ref, order=colour_order(B, q, mu))
    assert info == info_o == 0 and o.iparam[1] == it
    np.testing.assert_array_equal(r, ro)
    np.testing.assert_array_equal(u, uo)
    # fc3d_compute_error on the same sparse problem
    e = C.c_double()
    rc = capi.load().fc3d_compute_error(p._p, N.dptr(r), N.dptr(u), 1e-5, None,
                                        float(np.linalg.norm(q)), C.byref(e))
    assert rc == 0 and e.value == pytest.approx(res, rel=1e-12)


def _stats():
    h, c, v = C.c_int64(), C.c_int64(), C.c_int64()
    capi.load().fc3d_hip_resident_stats(C.byref(h), C.byref(c), C.byref(v))
    return h.value, c.value, v.value


def test_driver_keeps_the_device_problem_between_calls(oracle):
    """Two fc3d_driver calls on one 50k-contact problem with another q: the
    second reuses the device problem (no colouring, layout or tuning) and
    matches a fresh problem bit for bit; new values are picked up; with
    residency 1, W's values are kept while NumericsDataVersion is unchanged."""
    g = N.generate_random_pairs(20_000, 50_000, 0.2, 0.5, 2)
    p = N.FrictionContactProblem(3, g.W, g.q, g.mu)
    o = N.SolverOptions(540)
    o.dparam[0] = 1e-6
    r, u = np.zeros(3 * g.nc), np.zeros(3 * g.nc)
    h0, c0, v0 = _stats()
    assert N.fc3d_driver(p, r, u, o) == 0
    t_first = o.dparam[20]
    h1, c1, v1 = _stats()
    assert (h1 - h0, c1 - c0) == (0, 1)
    # another q: the same contact set at a new step
    q2 = g.q * 0.9 + 0.05
    qp = p._p.contents.q
    for k in range(3 * g.nc):
        qp[k] = q2[k]
    r2, u2 = np.zeros(3 * g.nc), np.zeros(3 * g.nc)
    assert N.fc3d_driver(p, r2, u2, o) == 0
    h2, c2, v2 = _stats()
    assert (h2 - h1, c2 - c1, v2 - v1) == (1, 0, 1)
    assert o.dparam[20] < 0.5 * t_first  # set-up without colouring, layout, tuning
    fresh = N.SolverOptions(540)
    fresh.dparam[0] = 1e-6
    fresh.dparam[28] = -1
    p2 = N.FrictionContactProblem(3, g.W, q2, g.mu)
    r3, u3 = np.zeros(3 * g.nc), np.zeros(3 * g.nc)
    assert N.fc3d_driver(p2, r3, u3, fresh) == 0
    assert fresh.iparam[1] == o.iparam[1]
    np.testing.assert_array_equal(r2, r3)
    np.testing.assert_array_equal(u2, u3)
    # new W values are uploaded (default residency)
    S = p._p.contents.M.contents.matrix1.contents
    for k in range(S.nbblocks):
        for e in range(9):
            S.block[k][e] *= 1.5
    r4, u4 = np.zeros(3 * g.nc), np.zeros(3 * g.nc)
    assert N.fc3d_driver(p, r4, u4, o) == 0
    W15 = N.SparseBlock3(g.W.row_ptr, g.W.col_idx, g.W.values * 1.5)
    H = N.HipProblem(W15, q2, g.mu)
    H.set_reaction(np.zeros(3 * g.nc))
    H.solve(N.hip_params(tol=1e-6))
    np.testing.assert_array_equal(r4, H.reaction())
    # residency 1: an unchanged non-zero version keeps the device values
    o.dparam[28] = 1
    S.version = 7
    assert N.fc3d_driver(p, r4, u4, o) == 0  # uploads (version seen first)
    _, _, va = _stats()
    vals = g.W.values.reshape(-1, 9)
    for k in range(S.nbblocks):
        for e in range(9):
            S.block[k][e] = vals[k, e]
    r5, u5 = np.zeros(3 * g.nc), np.zeros(3 * g.nc)
    assert N.fc3d_driver(p, r5, u5, o) == 0
    _, _, vb = _stats()
    assert vb == va  # version 7 unchanged: values not re-uploaded
    np.testing.assert_array_equal(r5, H.reaction())
    S.version = 8
    r5[:] = 0.0
    assert N.fc3d_driver(p, r5, u5, o) == 0
    np.testing.assert_array_equal(r5, r3)
    capi.load().fc3d_hip_resident_clear(None)
