"""Host logic of the product library (no GPU needed): registry, options,
driver checks, .dat I/O, generators, and the exported C-ABI."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import REPO, fixture_path, load_fixture
from siconos_amd import capi
from siconos_amd import numerics as N


def declared_symbols():
    names = set()
    for h in sorted(os.listdir(os.path.join(REPO, "include"))):
        text = open(os.path.join(REPO, "include", h)).read()
        for m in re.finditer(r"^FC3D_API\s+[^;(]*?\**\s*(\w+)\s*\(", text, re.M):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol(lib):
    names = declared_symbols()
    assert len(names) > 60
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(capi.SIGNATURES), set(names) ^ set(capi.SIGNATURES)


def test_registry_contains_the_hip_solver_and_local_solvers(lib):
    for sid, name in [(540, b"FC3D_NSGS_HIP"), (565, b"ONECONE_NSN_GP_HYBRID"),
                      (551, b"ONECONE_NSN_GP"), (550, b"ONECONE_NSN"), (552, b"OC_PROJ"),
                      (553, b"OC_PROJ_LI")]:
        assert lib.solver_registry_exists(sid)
        assert lib.solver_options_id_to_name(sid) == name
        assert lib.solver_options_name_to_id(name) == sid
    assert not lib.solver_registry_exists(500)  # the CPU solver is not in the product
    assert lib.solver_options_id_to_name(12345) == b"NONAME"


def test_duplicate_registration_is_rejected(lib):
    entry = C.c_void_p(lib.solver_registry_lookup(540))
    n = lib.solver_registry_count()
    assert lib.solver_registry_register(entry) == -1
    assert lib.solver_registry_count() == n


def test_options_defaults_equal_the_reference_nsgs_defaults(oracle):
    """FC3D_NSGS_HIP keeps fc3d_nsgs_set_default (fc3d_nsgs.c:919-950) and the
    hybrid local defaults (onecontact...c:1037-1056); compared with the
    oracle's independent restatement."""
    o = N.SolverOptions(540)
    ref = oracle.Options()
    np.testing.assert_array_equal(o.iparam[:20], ref.iparam[:20])
    np.testing.assert_array_equal(o.dparam[:20], ref.dparam[:20])
    li, lr = o.internal(), ref.internal()
    assert li.solverId == 565
    np.testing.assert_array_equal(li.iparam, lr.iparam)
    np.testing.assert_array_equal(li.dparam[:2], lr.dparam[:2])
    np.testing.assert_array_equal(li.dparam[3], lr.dparam[3])
    assert o.iparam[0] == 1000 and o.dparam[0] == 1e-4 and o.iparam[13] == 2
    assert o.iparam[23] == 8


def test_update_internal_and_copy(lib):
    o = N.SolverOptions(540)
    o.update_internal(0, 553)
    assert o.internal().solverId == 553
    assert o.internal().iparam[0] == 1000 and o.internal().dparam[0] == 1e-14
    c = lib.solver_options_copy(o._p)
    assert c.contents.solverId == 540 and c.contents.internalSolvers[0].contents.solverId == 553
    lib.solver_options_delete(c)


def test_driver_argument_checks_and_trivial_case(lib):
    g = N.generate_random_pairs(30, 50, 0.2, 0.5, 9)
    p = N.FrictionContactProblem(3, g.W, np.abs(g.q), g.mu)
    o = N.SolverOptions(540)
    r = np.full(150, 7.0)
    u = np.zeros(150)
    # every q_n >= 0: take-off everywhere, no GPU involved (fc3d_driver.c:51-71)
    assert N.fc3d_driver(p, r, u, o) == 0
    np.testing.assert_array_equal(r, 0.0)
    np.testing.assert_array_equal(u, np.abs(g.q))
    assert lib.fc3d_driver(None, N.dptr(r), N.dptr(u), o._p) == -1
    bad = N.SolverOptions(540)
    bad._p.contents.solverId = 12345
    p2 = N.FrictionContactProblem(3, g.W, g.q, g.mu)
    assert N.fc3d_driver(p2, r, u, bad) == -4
    bad._p.contents.solverId = 540


def test_no_cpu_fallback_without_a_gpu(lib):
    """On a machine without a GPU every compute entry point fails loudly."""
    if lib.fc3d_hip_device_count() > 0:
        pytest.skip("a GPU is present")
    g = N.generate_random_pairs(30, 50, 0.2, 0.5, 9)
    p = N.FrictionContactProblem(3, g.W, g.q, g.mu)
    o = N.SolverOptions(540)
    r, u = np.zeros(150), np.zeros(150)
    assert N.fc3d_driver(p, r, u, o) == -60
    info, e = N.fc3d_compute_error(p, r, u, 1e-6, o, 1.0)
    assert info == -60
    with pytest.raises(RuntimeError):
        N.colour_contacts(g.W)


@pytest.mark.parametrize("name", ["FC3D_Example1_SBM.dat", "FC3D_Example1.dat",
                                  "Capsules-i122-1617.dat", "FrictionContact3D_1c.dat"])
def test_dat_roundtrip(tmp_path, name):
    p = N.FrictionContactProblem.from_file(fixture_path(name))
    out = tmp_path / "copy.dat"
    assert p.write(out) == 0
    p2 = N.FrictionContactProblem.from_file(out)
    W1, W2 = p.sparse_block(), p2.sparse_block()
    np.testing.assert_array_equal(W1.row_ptr, W2.row_ptr)
    np.testing.assert_array_equal(W1.col_idx, W2.col_idx)
    np.testing.assert_array_equal(W1.values, W2.values)
    np.testing.assert_array_equal(p.q, p2.q)
    np.testing.assert_array_equal(p.mu, p2.mu)


def test_dense_and_sbm_example_are_the_same_problem():
    W1, q1, mu1 = load_fixture("FC3D_Example1.dat")
    W2, q2, mu2 = load_fixture("FC3D_Example1_SBM.dat")
    np.testing.assert_array_equal(W1.to_dense(), W2.to_dense())
    np.testing.assert_array_equal(q1, q2)
    np.testing.assert_array_equal(mu1, mu2)


def test_unreadable_file_is_refused():
    with pytest.raises(IOError):
        N.FrictionContactProblem.from_file(os.path.join(REPO, "README.md"))


@pytest.mark.parametrize("gen", ["pairs", "pile"])
def test_generator_invariants(gen):
    if gen == "pairs":
        g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    else:
        g = N.generate_granular_pile(400, 1000, 0.1, 0.9, 1)
    W = g.W
    D = W.to_dense()
    np.testing.assert_allclose(D, D.T, atol=1e-14)
    ev = np.linalg.eigvalsh(D)
    assert ev.min() > -1e-12
    assert np.sum(ev > 1e-9) <= 6 * g.nbodies
    for i in range(W.nc):
        ks = [k for k in range(W.row_ptr[i], W.row_ptr[i + 1]) if W.col_idx[k] == i]
        assert len(ks) == 1
        np.testing.assert_allclose(W.block(ks[0]), np.diag([2.0, 7.0, 7.0]), atol=1e-13)
        cols = W.col_idx[W.row_ptr[i]:W.row_ptr[i + 1]]
        assert np.all(np.diff(cols) > 0)
    # row i couples exactly the contacts that share a body with i
    for i in range(0, W.nc, 37):
        bodies = {g.body_a[i], g.body_b[i]}
        expect = sorted(j for j in range(W.nc) if {g.body_a[j], g.body_b[j]} & bodies)
        np.testing.assert_array_equal(W.col_idx[W.row_ptr[i]:W.row_ptr[i + 1]], expect)
    lo, hi = (0.2, 0.5) if gen == "pairs" else (0.1, 0.9)
    assert g.mu.min() >= lo and g.mu.max() <= hi


def test_generator_is_seeded():
    a = N.generate_random_pairs(300, 1000, 0.2, 0.5, 7)
    b = N.generate_random_pairs(300, 1000, 0.2, 0.5, 7)
    c = N.generate_random_pairs(300, 1000, 0.2, 0.5, 8)
    np.testing.assert_array_equal(a.W.values, b.W.values)
    np.testing.assert_array_equal(a.q, b.q)
    assert not np.array_equal(a.q, c.q)


def test_generator_q_is_H_vfree_free_velocity_scale():
    """q = H v_free with v_free ~ N(0,1): for the random-pair generator the
    normal component of q has variance 2 (two bodies, unit translational
    part, no rotational part along n) and the tangential ones 2 + 2 r^2."""
    g = N.generate_random_pairs(20000, 50000, 0.2, 0.5, 2)
    qn = g.q[0::3]
    qt = g.q[1::3]
    assert abs(np.var(qn) - 2.0) < 0.1
    assert abs(np.var(qt) - 4.0) < 0.2


def test_sbm_from_dense_keeps_nonzero_blocks(lib):
    D = np.zeros((9, 9))
    D[0:3, 0:3] = np.eye(3)
    D[3:6, 6:9] = np.arange(9).reshape(3, 3)
    D[6:9, 3:6] = np.arange(9).reshape(3, 3).T
    S = lib.SBM_from_dense_3x3(3, N.dptr(np.ascontiguousarray(D.reshape(-1, order="F"))))
    s = S.contents
    assert s.nbblocks == 5  # 3 diagonal (kept even if zero) + 2 coupling
    lib.SBM_free_all(S)


def test_fc2d_row_builder_keeps_the_heap_intact():
    """Regression: fc2d_driver builds the scalar rows of every storage on the
    host before any device call (an SBM block puts four entries in its two
    rows).  Run in a child process with glibc's heap checking on; without a
    GPU the solve itself returns NUMERICS_ERR_DEVICE."""
    import subprocess
    import sys
    from conftest import FC2D_COLLECTION, fixture_file
    paths = [fixture_file(n) for n in FC2D_COLLECTION]
    code = (
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {REPO!r})\n"
        "from siconos_amd import numerics as N\n"
        f"for path in {paths!r}:\n"
        "    p = N.FrictionContactProblem.from_file(path)\n"
        "    r = np.zeros(2 * p.numberOfContacts)\n"
        "    rc = N.fc2d_driver(p, r, r.copy(), N.SolverOptions(440))\n"
        "    assert rc in (0, -60), rc\n"
        "print('ok')\n")
    env = dict(os.environ, MALLOC_CHECK_="3")
    out = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                         timeout=300)
    assert out.returncode == 0 and out.stdout.strip().endswith("ok"), out.stderr[-2000:]


def test_options_copy_keeps_the_pointer_links():
    """solver_options_copy keeps callback, solverData and solverParameters
    (SolverOptions.c:249-255)."""
    o = N.SolverOptions(540)
    fn = capi.CALLBACK_FN(lambda *a: None)
    cbs = capi.Callback(None, fn)
    o._p.contents.callback = C.pointer(cbs)
    o._p.contents.solverData = C.c_void_p(0x1234)
    o._p.contents.solverParameters = C.c_void_p(0x5678)
    lib = capi.load()
    cp = lib.solver_options_copy(o._p)
    try:
        assert C.addressof(cp.contents.callback.contents) == C.addressof(cbs)
        assert cp.contents.solverData == 0x1234 and cp.contents.solverParameters == 0x5678
    finally:
        cp.contents.callback = None
        cp.contents.solverData = None
        cp.contents.solverParameters = None
        lib.solver_options_delete(cp)
        o._p.contents.callback = None
        o._p.contents.solverData = None
        o._p.contents.solverParameters = None
