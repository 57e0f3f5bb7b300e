"""fc3d_hip_compute_error_device (no allocation, caller-owned device arrays)
against fc3d_hip_compute_error on the same problems: velocities bit for bit
(both add each row's blocks in the stored order), the error to 1e-12 relative
(the partial sums are added in another order).  Two calls in a row, the second
on a smaller problem, as the static scratch array keeps the first call's
partials."""
import ctypes as C

import numpy as np
import pytest

from siconos_amd import capi
from siconos_amd import numerics as N

pytestmark = pytest.mark.gpu


class _Hip:
    """Device buffers through the HIP runtime (the test's own allocations)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.bufs = []

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        assert self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0  # host to device
        self.bufs.append(p)
        return p.value

    def download(self, p, like):
        out = np.empty_like(like)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(out.ctypes.data, C.c_void_p(p), out.nbytes, 2) == 0
        return out

    def free(self):
        for p in self.bufs:
            self.lib.hipFree(p)
        self.bufs = []


def _device_call(W, q, mu, r, norm):
    d = _Hip()
    try:
        u0 = np.full(3 * W.nc, -1.0)
        e0 = np.full(1, -1.0)
        ptrs = [d.upload(np.asarray(W.row_ptr, dtype=np.int64)),
                d.upload(np.asarray(W.col_idx, dtype=np.int64)),
                d.upload(np.asarray(W.values, dtype=np.float64)),
                d.upload(np.asarray(q, dtype=np.float64)), d.upload(np.asarray(mu, dtype=np.float64)),
                d.upload(np.asarray(r, dtype=np.float64))]
        pu, pe = d.upload(u0), d.upload(e0)
        rc = capi.load().fc3d_hip_compute_error_device(W.nc, int(W.nnzb), *ptrs, norm, pu, pe, 0,
                                                       None)
        u = d.download(pu, u0)
        e = d.download(pe, e0)
        return rc, u, float(e[0])
    finally:
        d.free()


def test_no_alloc_residual_matches_the_resident_one_twice():
    for nb, nc, seed in ((20_000, 50_000, 3), (3_000, 7_000, 4)):
        g = N.generate_random_pairs(nb, nc, 0.2, 0.5, seed)
        H = N.HipProblem(g.W, g.q, g.mu)
        # a reaction after a few sweeps: inside the cone, not a solution
        H.set_reaction(np.zeros(3 * nc))
        H.solve(N.hip_params(tol=1e-12, max_iter=5))
        r = H.reaction()
        norm = float(np.linalg.norm(g.q))
        e_ref = H.compute_error(norm)
        u_ref = H.velocity()
        rc, u, e = _device_call(g.W, g.q, g.mu, r, norm)
        print(f"{nc} contacts: error {e:.17g} (resident {e_ref:.17g})", flush=True)
        assert rc == 0, capi.last_error()
        np.testing.assert_array_equal(u, u_ref)
        assert e == pytest.approx(e_ref, rel=1e-12)


def test_no_alloc_residual_refuses_sizes_beyond_its_scratch():
    g = N.generate_random_pairs(300, 1000, 0.2, 0.5, 1)
    d = _Hip()
    p = d.upload(np.zeros(16))
    lib = capi.load()
    rc = lib.fc3d_hip_compute_error_device((1 << 22) + 1, 0, p, p, p, p, p, p, 1.0, p, p, 0, None)
    assert rc == -2 and "scratch" in capi.last_error()
    rc = lib.fc3d_hip_compute_error_device(g.nc, 0, p, p, p, p, p, p, -1.0, p, p, 0, None)
    assert rc == -2
    d.free()
