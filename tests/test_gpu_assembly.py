"""Assembly of W = H M^-1 H^T and q = H v_free (fc3d_hip_assemble) against a
dense product and against the host generator's own assembly."""
import numpy as np
import pytest

from siconos_amd import numerics as N


def dense_from_bodies(nb, ba, bb, H, minv, vfree, diagonal):
    nc = len(ba)
    Hg = np.zeros((3 * nc, 6 * nb))
    for c in range(nc):
        for side, k in enumerate((ba[c], bb[c])):
            if k < 0:
                continue
            blk = H[36 * c + 18 * side:36 * c + 18 * side + 18].reshape(3, 6, order="F")
            Hg[3 * c:3 * c + 3, 6 * k:6 * k + 6] += blk
    Mi = np.zeros((6 * nb, 6 * nb))
    for k in range(nb):
        if diagonal:
            Mi[6 * k:6 * k + 6, 6 * k:6 * k + 6] = np.diag(minv[6 * k:6 * k + 6])
        else:
            Mi[6 * k:6 * k + 6, 6 * k:6 * k + 6] = minv[36 * k:36 * k + 36].reshape(6, 6, order="F")
    return Hg @ Mi @ Hg.T, Hg @ vfree


def test_generator_exports_the_inputs_of_its_assembly():
    """CPU: the exported H, M^-1 and v_free reproduce the generator's W and q."""
    g = N.generate_random_pairs(30, 60, 0.2, 0.5, 7)
    Wd, q = dense_from_bodies(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree, True)
    np.testing.assert_allclose(g.W.to_dense(), Wd, rtol=0, atol=1e-13)
    np.testing.assert_allclose(g.q, q, rtol=0, atol=1e-13)


@pytest.mark.gpu
@pytest.mark.parametrize("kind", ["pairs", "pile"])
def test_assembly_matches_the_generator(lib, kind):
    if kind == "pairs":
        g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    else:
        g = N.generate_granular_pile(8000, 20000, 0.1, 0.9, 5)
    W, q, ms = N.assemble(g.nbodies, g.body_a, g.body_b, g.H, g.minv, g.vfree)
    np.testing.assert_array_equal(W.row_ptr, g.W.row_ptr)
    np.testing.assert_array_equal(W.col_idx, g.W.col_idx)
    scale = np.abs(g.W.values).max()
    np.testing.assert_allclose(W.values, g.W.values, rtol=0, atol=4e-16 * 16 * scale)
    np.testing.assert_allclose(q, g.q, rtol=0, atol=4e-16 * 16 * np.abs(g.q).max())
    assert ms > 0


@pytest.mark.gpu
def test_assembly_with_full_mass_blocks_and_fixed_bodies(lib):
    rng = np.random.default_rng(3)
    nb, nc = 30, 90
    ba = rng.integers(0, nb, nc).astype(np.int32)
    bb = rng.integers(0, nb, nc).astype(np.int32)
    bb[rng.random(nc) < 0.15] = -1  # contact with a fixed obstacle
    bb[(bb == ba)] = -1
    ba[:3] = -1  # both sides fixed for a few: a zero diagonal block
    bb[:3] = -1
    H = rng.normal(size=36 * nc)
    Ms = []
    for k in range(nb):
        A = rng.normal(size=(6, 6))
        Ms.append((A @ A.T + 6 * np.eye(6)).ravel(order="F"))
    minv = np.concatenate(Ms)
    vfree = rng.normal(size=6 * nb)
    W, q, ms = N.assemble(nb, ba, bb, H, minv, vfree, minv_diagonal=False)
    Wd, qd = dense_from_bodies(nb, ba, bb, H, minv, vfree, False)
    np.testing.assert_allclose(W.to_dense(), Wd, rtol=0, atol=1e-12 * np.abs(Wd).max())
    np.testing.assert_allclose(q, qd, rtol=0, atol=1e-12 * np.abs(qd).max())
    # the structure is exactly "shares a body" plus the diagonal
    for c in range(nc):
        row = set(W.col_idx[W.row_ptr[c]:W.row_ptr[c + 1]])
        expect = {c} | {d for d in range(nc)
                        if ({ba[c], bb[c]} - {-1}) & ({ba[d], bb[d]} - {-1})}
        assert row == expect


@pytest.mark.gpu
def test_assembly_refuses_bad_body_indices(lib):
    H = np.zeros(36)
    with pytest.raises(RuntimeError, match="-2"):
        N.assemble(2, [0], [5], H, np.ones(12), np.zeros(12))


@pytest.mark.gpu
@pytest.mark.parametrize("diagonal", [True, False])
def test_assembly_of_blocks_with_several_shared_sides(lib, diagonal):
    """Blocks that sum more than one pair of sides: two contacts between the
    same two bodies (in either order), and contacts with one body on both
    sides (four pairs on their diagonal block), against the dense product."""
    rng = np.random.default_rng(21)
    nb, nc = 12, 60
    ba = rng.integers(0, nb, nc).astype(np.int32)
    bb = ((ba + 1 + rng.integers(0, nb - 1, nc)) % nb).astype(np.int32)
    ba[40:50], bb[40:50] = ba[0:10], bb[0:10]      # the same pair, same order
    ba[50:55], bb[50:55] = bb[10:15], ba[10:15]    # the same pair, swapped
    bb[55:60] = ba[55:60]                          # one body on both sides
    bb[30:33] = -1
    H = rng.normal(size=36 * nc)
    if diagonal:
        minv = rng.uniform(0.5, 2.0, size=6 * nb)
    else:
        Ms = []
        for _ in range(nb):
            A = rng.normal(size=(6, 6))
            Ms.append((A @ A.T + 6 * np.eye(6)).ravel(order="F"))
        minv = np.concatenate(Ms)
    vfree = rng.normal(size=6 * nb)
    W, q, _ = N.assemble(nb, ba, bb, H, minv, vfree, minv_diagonal=diagonal)
    Wd, qd = dense_from_bodies(nb, ba, bb, H, minv, vfree, diagonal)
    np.testing.assert_allclose(W.to_dense(), Wd, rtol=0, atol=1e-12 * np.abs(Wd).max())
    np.testing.assert_allclose(q, qd, rtol=0, atol=1e-12 * np.abs(qd).max())
    # the case mix the test is for
    pairs = lambda c, d: sum(1 for x in (ba[c], bb[c]) for y in (ba[d], bb[d]) if x >= 0 and x == y)
    assert pairs(0, 40) == 2 and pairs(10, 50) == 2 and pairs(55, 55) == 4
