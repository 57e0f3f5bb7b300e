"""Python surface of the fc3d path, mirroring the reference's numerics Python
module (numerics/python/friction_contact_pybind11.cpp:160-166): a
``FrictionContactProblem`` built from W (3x3 sparse-block or dense), q and mu,
``SolverOptions(solver_id)`` and ``fc3d_driver``.  Everything calls the C-ABI
of ``libsiconos_amd.so``; arithmetic runs on the GPU.
"""
import ctypes as C

import numpy as np

from . import capi

SICONOS_FRICTION_3D_NSGS = 500
SICONOS_FRICTION_3D_NSGS_HIP = 540
SICONOS_FRICTION_3D_ONECONTACT_NSN = 550
SICONOS_FRICTION_3D_ONECONTACT_NSN_GP = 551
SICONOS_FRICTION_3D_ONECONTACT_ProjectionOnCone = 552
SICONOS_FRICTION_3D_ONECONTACT_ProjectionOnConeWithLocalIteration = 553
SICONOS_FRICTION_3D_ONECONTACT_NSN_GP_HYBRID = 565

SICONOS_IPARAM_MAX_ITER = 0
SICONOS_IPARAM_ITER_DONE = 1
SICONOS_DPARAM_TOL = 0
SICONOS_DPARAM_RESIDU = 1
SICONOS_NSGS_ERROR_EVALUATION_TYPE = 13
SICONOS_FRICTION_3D_NSN_FORMULATION = 10
SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY = 14
SICONOS_FRICTION_3D_NSGS_HIP_PRECISION = 20
SICONOS_FRICTION_3D_NSGS_HIP_WARM_START = 21
SICONOS_FRICTION_3D_NSGS_HIP_SWEEPS_PER_LAUNCH = 23
SICONOS_FRICTION_3D_NSGS_HIP_NUM_COLOURS = 25
SICONOS_NSGS_FREEZING_CONTACT = 15
SICONOS_FRICTION_2D_NSGS_HIP = 440
NUMERICS_ERR_INVALID_ARGUMENT = -2
NUMERICS_ERR_INVALID_SOLVER = -4
NUMERICS_ERR_NOT_IMPLEMENTED = -5
NUMERICS_ERR_MAX_ITER = -20

_libc = C.CDLL(None)
_libc.malloc.restype = C.c_void_p
_libc.malloc.argtypes = [C.c_size_t]


def dptr(a):
    return a.ctypes.data_as(capi.c_double_p)


def fptr(a):
    return a.ctypes.data_as(capi.c_float_p)


def iptr(a):
    return a.ctypes.data_as(capi.c_int_p)


def lptr(a):
    return a.ctypes.data_as(capi.c_int64_p)


class SparseBlock3:
    """CSR over 3x3 blocks: row_ptr (nc+1), col_idx (nnzb), blocks (nnzb, 3, 3).

    ``blocks[k]`` is the k-th block as a 3x3 array; it is stored column-major
    (Fortran order), as the reference stores ``block[k]``."""

    def __init__(self, row_ptr, col_idx, blocks):
        self.row_ptr = np.ascontiguousarray(row_ptr, dtype=np.int64)
        self.col_idx = np.ascontiguousarray(col_idx, dtype=np.int64)
        b = np.asarray(blocks, dtype=np.float64)
        if b.ndim == 3:  # (nnzb, 3, 3) in natural (row, col) indexing -> column-major storage
            b = np.ascontiguousarray(b.transpose(0, 2, 1)).reshape(-1, 9)
        self.values = np.ascontiguousarray(b.reshape(-1, 9))  # column-major per block
        self.nc = len(self.row_ptr) - 1

    @property
    def nnzb(self):
        return int(self.row_ptr[-1])

    def block(self, k):
        return self.values[k].reshape(3, 3, order="F")

    def to_dense(self):
        n = self.nc
        D = np.zeros((3 * n, 3 * n))
        for i in range(n):
            for k in range(self.row_ptr[i], self.row_ptr[i + 1]):
                j = self.col_idx[k]
                D[3 * i:3 * i + 3, 3 * j:3 * j + 3] = self.block(k)
        return D

    @staticmethod
    def from_dense(D, keep_zero_diagonal=True):
        n = D.shape[0] // 3
        rp, ci, vals = [0], [], []
        for i in range(n):
            for j in range(n):
                b = D[3 * i:3 * i + 3, 3 * j:3 * j + 3]
                if i == j or np.any(b != 0.0):
                    ci.append(j)
                    vals.append(np.asarray(b, dtype=np.float64).reshape(-1, order="F"))
            rp.append(len(ci))
        return SparseBlock3(np.array(rp), np.array(ci), np.array(vals).reshape(-1, 9))


class FrictionContactProblem:
    """A friction-contact problem (FrictionContactProblem.h:33-46).

    Built from Python for dimension 3: ``M`` is a SparseBlock3 or a dense
    (3nc x 3nc) array.  ``from_file`` also reads dimension 2 and compressed
    (NM_SPARSE) matrices.  The C struct is owned by this object."""

    def __init__(self, dim=3, M=None, q=None, mu=None, _cptr=None):
        self._lib = capi.load()
        if _cptr is not None:
            self._p = _cptr
            return
        if dim != 3:
            raise ValueError("only dimension 3 is handled")
        if isinstance(M, np.ndarray):
            M = SparseBlock3.from_dense(M)
        q = np.ascontiguousarray(q, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        sbm = self._lib.SBM_new_3x3_from_csr(M.nc, lptr(M.row_ptr), lptr(M.col_idx),
                                             dptr(M.values))
        nm = self._lib.NM_create_from_data(1, 3 * M.nc, 3 * M.nc, C.cast(sbm, C.c_void_p))
        # q and mu are released by frictionContactProblem_free: allocate with malloc
        pq = _libc.malloc(8 * len(q))
        pmu = _libc.malloc(8 * len(mu))
        C.memmove(pq, q.ctypes.data, 8 * len(q))
        C.memmove(pmu, mu.ctypes.data, 8 * len(mu))
        self._p = self._lib.frictionContactProblem_new_with_data(
            3, M.nc, nm, C.cast(pq, capi.c_double_p), C.cast(pmu, capi.c_double_p))

    @staticmethod
    def from_file(path):
        lib = capi.load()
        p = lib.frictionContact_new_from_filename(str(path).encode())
        if not p:
            raise IOError(f"cannot read friction-contact problem from {path}")
        return FrictionContactProblem(_cptr=p)

    def write(self, path):
        return self._lib.frictionContact_printInFilename(self._p, str(path).encode())

    @property
    def numberOfContacts(self):
        return self._p.contents.numberOfContacts

    @property
    def dimension(self):
        return self._p.contents.dimension

    @property
    def storage(self):
        return self._p.contents.M.contents.storageType

    @property
    def q(self):
        n = self.dimension * self.numberOfContacts
        return np.ctypeslib.as_array(self._p.contents.q, shape=(n,)).copy()

    def dense(self):
        """W as a dense array, whatever its storage (duplicates summed)."""
        M = self._p.contents.M.contents
        n0, n1 = M.size0, M.size1
        if M.storageType == 0:
            D = np.ctypeslib.as_array(M.matrix0, shape=(n0 * n1,))
            return D.reshape(n0, n1, order="F").copy()
        out = np.zeros((n0, n1))
        if M.storageType == 1:
            S = M.matrix1.contents
            d = self.dimension
            for r in range(S.filled1 - 1):
                for k in range(S.index1_data[r], S.index1_data[r + 1]):
                    c = S.index2_data[k]
                    b = np.ctypeslib.as_array(S.block[k], shape=(d * d,)).reshape(d, d, order="F")
                    out[d * r:d * r + d, d * c:d * c + d] += b
            return out
        A = self._lib.NM_csc(self._p.contents.M).contents
        nnz = A.p[A.n]
        p = np.ctypeslib.as_array(A.p, shape=(A.n + 1,))
        i = np.ctypeslib.as_array(A.i, shape=(max(nnz, 1),))[:nnz]
        x = np.ctypeslib.as_array(A.x, shape=(max(nnz, 1),))[:nnz]
        for j in range(A.n):
            np.add.at(out[:, j], i[p[j]:p[j + 1]], x[p[j]:p[j + 1]])
        return out

    @property
    def mu(self):
        return np.ctypeslib.as_array(self._p.contents.mu, shape=(self.numberOfContacts,)).copy()

    def sparse_block(self):
        """W as SparseBlock3 (dense storage is converted)."""
        M = self._p.contents.M.contents
        if M.storageType == 0:
            D = np.ctypeslib.as_array(M.matrix0, shape=(M.size0 * M.size1,))
            return SparseBlock3.from_dense(D.reshape(M.size0, M.size1, order="F"))
        S = M.matrix1.contents
        n = S.filled1 - 1
        rp = np.array([S.index1_data[i] for i in range(n + 1)], dtype=np.int64)
        ci = np.array([S.index2_data[k] for k in range(S.nbblocks)], dtype=np.int64)
        vals = np.empty((S.nbblocks, 9))
        for k in range(S.nbblocks):
            vals[k] = np.ctypeslib.as_array(S.block[k], shape=(9,))
        return SparseBlock3(rp, ci, vals)

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            self._lib.frictionContactProblem_free(p)
            self._p = None


class SolverOptions:
    """Owning wrapper of ``SolverOptions*`` (SolverOptions.h:65-91)."""

    def __init__(self, solver_id=None, _cptr=None, _owner=None):
        self._lib = capi.load()
        self._owner = _owner
        if _cptr is not None:
            self._p = _cptr
        else:
            self._p = self._lib.solver_options_create(solver_id)
            if not self._p:
                raise ValueError(f"solver id {solver_id} is not registered")

    @property
    def solverId(self):
        return self._p.contents.solverId

    @property
    def iparam(self):
        return np.ctypeslib.as_array(self._p.contents.iparam, shape=(30,))

    @property
    def dparam(self):
        return np.ctypeslib.as_array(self._p.contents.dparam, shape=(30,))

    def internal(self, n=0):
        p = self._lib.solver_options_get_internal_solver(self._p, n)
        return SolverOptions(_cptr=p, _owner=self) if p else None

    def update_internal(self, n, solver_id):
        self._lib.solver_options_update_internal(self._p, n, solver_id)

    def __del__(self):
        p = getattr(self, "_p", None)
        if p and self._owner is None:
            self._lib.solver_options_delete(p)
        self._p = None


def fc3d_driver(problem, reaction, velocity, options):
    """numerics fc3d_driver (fc3d_driver.c:77-145); reaction is in/out."""
    assert reaction.dtype == np.float64 and velocity.dtype == np.float64
    return capi.load().fc3d_driver(problem._p, dptr(reaction), dptr(velocity), options._p)


def fc2d_driver(problem, reaction, velocity, options):
    """numerics fc2d_driver (fc2d_driver.c:38-150); reaction is in/out."""
    assert reaction.dtype == np.float64 and velocity.dtype == np.float64
    return capi.load().fc2d_driver(problem._p, dptr(reaction), dptr(velocity), options._p)


def fc2d_compute_error(problem, z, w, tolerance, norm):
    """fc2d_compute_error (fc2d_compute_error.c:104-148), on the GPU;
    returns (info, error)."""
    err = C.c_double(0.0)
    info = capi.load().fc2d_compute_error(problem._p, dptr(z), dptr(w), tolerance, norm,
                                           C.byref(err))
    return info, err.value


class Fc2dHipProblem:
    """Resident fc2d problem of the C-ABI (fc2d_hip.h), uploaded from a
    dimension-2 FrictionContactProblem."""

    def __init__(self, problem, device=0):
        self._lib = capi.load()
        self.nc = problem.numberOfContacts
        h = C.c_void_p()
        rc = self._lib.fc2d_problem_to_device(problem._p, device, C.byref(h))
        if rc:
            raise RuntimeError(f"fc2d_problem_to_device: {rc} {capi.last_error()}")
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.fc2d_hip_problem_destroy(h)
            self._h = None

    @property
    def num_colours(self):
        return self._lib.fc2d_hip_problem_num_colours(self._h)

    def order(self):
        o = np.zeros(self.nc, dtype=np.int32)
        self._lib.fc2d_hip_problem_order(self._h, iptr(o))
        return o

    def set_reaction(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._lib.fc2d_hip_set_reaction(self._h, dptr(r)), "fc2d_hip_set_reaction")

    def reaction(self):
        r = np.zeros(2 * self.nc)
        self._check(self._lib.fc2d_hip_get_reaction(self._h, dptr(r)), "fc2d_hip_get_reaction")
        return r

    def velocity(self):
        u = np.zeros(2 * self.nc)
        self._check(self._lib.fc2d_hip_get_velocity(self._h, dptr(u)), "fc2d_hip_get_velocity")
        return u

    def solve(self, max_iter=1000, tol=1e-4, error_evaluation=2, schedule=0):
        """schedule: 0 automatic, 1 one workgroup, 2 one launch per colour."""
        p = capi.Fc2dParams(max_iter, tol, error_evaluation, schedule)
        s = capi.Fc2dStats()
        self._check(self._lib.fc2d_hip_solve(self._h, C.byref(p), C.byref(s)), "fc2d_hip_solve")
        return s

    def compute_error(self, r=None, norm=None):
        e = C.c_double()
        u = np.zeros(2 * self.nc)
        rr = None if r is None else np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._lib.fc2d_hip_compute_error(self._h, dptr(rr) if rr is not None else None,
                                                      dptr(u), 0.0 if norm is None else norm,
                                                      C.byref(e)), "fc2d_hip_compute_error")
        return e.value, u

    @staticmethod
    def _check(rc, what):
        if rc:
            raise RuntimeError(f"{what}: {rc} {capi.last_error()}")


def fc3d_compute_error(problem, z, w, tolerance, options, norm):
    """fc3d_compute_error (fc3d_compute_error.c:70-122); returns (info, error)."""
    err = C.c_double(0.0)
    info = capi.load().fc3d_compute_error(problem._p, dptr(z), dptr(w), tolerance,
                                           options._p if options else None, norm, C.byref(err))
    return info, err.value


class GeneratedProblem:
    """Seeded synthetic problem (arrays copied out of the C generator)."""

    def __init__(self, g):
        n, nb = g.nc, g.nnzb
        self.nc = n
        self.nbodies = g.nbodies
        self.W = SparseBlock3(np.ctypeslib.as_array(g.row_ptr, shape=(n + 1,)).copy(),
                              np.ctypeslib.as_array(g.col_idx, shape=(nb,)).copy(),
                              np.ctypeslib.as_array(g.blocks, shape=(nb * 9,)).copy())
        self.q = np.ctypeslib.as_array(g.q, shape=(3 * n,)).copy()
        self.mu = np.ctypeslib.as_array(g.mu, shape=(n,)).copy()
        self.body_a = np.ctypeslib.as_array(g.body_a, shape=(n,)).copy()
        self.body_b = np.ctypeslib.as_array(g.body_b, shape=(n,)).copy()
        # the inputs of the assembly: per contact the 3x6 blocks of both bodies
        # (column-major), per body the diagonal of M^-1 and the free velocity
        self.H = np.ctypeslib.as_array(g.H, shape=(n * 36,)).copy() if g.H else None
        self.minv = (np.ctypeslib.as_array(g.minv, shape=(g.nbodies * 6,)).copy()
                     if g.minv else None)
        self.vfree = (np.ctypeslib.as_array(g.vfree, shape=(g.nbodies * 6,)).copy()
                      if g.vfree else None)

    def problem(self):
        return FrictionContactProblem(3, self.W, self.q, self.mu)


def _generate(fn, nbodies, nc, mu_lo, mu_hi, seed):
    lib = capi.load()
    g = capi.Generated()
    rc = getattr(lib, fn)(nbodies, nc, mu_lo, mu_hi, seed, C.byref(g))
    if rc:
        raise RuntimeError(f"{fn} failed with {rc}")
    try:
        return GeneratedProblem(g)
    finally:
        lib.fc3d_generated_free(C.byref(g))


def generate_random_pairs(nbodies, nc, mu_lo, mu_hi, seed):
    return _generate("fc3d_generate_random_pairs", nbodies, nc, mu_lo, mu_hi, seed)


def generate_granular_pile(nbodies, nc, mu_lo, mu_hi, seed):
    return _generate("fc3d_generate_granular_pile", nbodies, nc, mu_lo, mu_hi, seed)


def assemble(nbodies, body_a, body_b, H, minv, vfree, minv_diagonal=True, device=0):
    """W = H M^-1 H^T and q = H v_free on the GPU (fc3d_hip_assemble).
    Returns (SparseBlock3 W, q, device milliseconds)."""
    lib = capi.load()
    body_a = np.ascontiguousarray(body_a, dtype=np.int32)
    body_b = np.ascontiguousarray(body_b, dtype=np.int32)
    H = np.ascontiguousarray(H, dtype=np.float64)
    minv = np.ascontiguousarray(minv, dtype=np.float64)
    vfree = np.ascontiguousarray(vfree, dtype=np.float64)
    nc = len(body_a)
    g = capi.Generated()
    ms = C.c_double()
    rc = lib.fc3d_hip_assemble(nbodies, nc, iptr(body_a), iptr(body_b), dptr(H), dptr(minv),
                               int(minv_diagonal), dptr(vfree), device, C.byref(g), C.byref(ms))
    if rc:
        raise RuntimeError(f"fc3d_hip_assemble: {rc} {capi.last_error()}")
    try:
        nb = g.nnzb
        W = SparseBlock3(np.ctypeslib.as_array(g.row_ptr, shape=(nc + 1,)).copy(),
                         np.ctypeslib.as_array(g.col_idx, shape=(nb,)).copy(),
                         np.ctypeslib.as_array(g.blocks, shape=(nb * 9,)).copy())
        q = np.ctypeslib.as_array(g.q, shape=(3 * nc,)).copy()
    finally:
        lib.fc3d_generated_free(C.byref(g))
    return W, q, ms.value


def colour_contacts(W):
    """Greedy colouring on the GPU; returns (colours, ncolours)."""
    col = np.empty(W.nc, dtype=np.int32)
    n = capi.load().fc3d_colour_contacts(W.nc, lptr(W.row_ptr), lptr(W.col_idx), iptr(col))
    if n < 0:
        raise RuntimeError(f"colouring failed: {n} {capi.last_error()}")
    return col, n


def colouring_is_valid(W, colour):
    colour = np.ascontiguousarray(colour, dtype=np.int32)
    return bool(capi.load().fc3d_colouring_is_valid(W.nc, lptr(W.row_ptr), lptr(W.col_idx),
                                                    iptr(colour)))


def hip_params(**kw):
    p = capi.HipParams()
    capi.load().fc3d_hip_params_default(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class HipProblem:
    """Resident problem of the C-ABI (fc3d_hip.h)."""

    def __init__(self, W, q, mu, device=0):
        self._lib = capi.load()
        self.nc = W.nc
        self.W = W
        h = C.c_void_p()
        q = np.ascontiguousarray(q, dtype=np.float64)
        mu = np.ascontiguousarray(mu, dtype=np.float64)
        rc = self._lib.fc3d_hip_problem_create(W.nc, lptr(W.row_ptr), lptr(W.col_idx),
                                               dptr(W.values), dptr(q), dptr(mu), device,
                                               C.byref(h))
        if rc:
            raise RuntimeError(f"fc3d_hip_problem_create: {rc} {capi.last_error()}")
        self._h = h

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what}: {rc} {capi.last_error()}")

    @property
    def num_colours(self):
        return self._lib.fc3d_hip_problem_num_colours(self._h)

    def colours(self):
        c = np.empty(self.nc, dtype=np.int32)
        self._check(self._lib.fc3d_hip_problem_colours(self._h, iptr(c)), "colours")
        return c

    def order(self):
        o = np.empty(self.nc, dtype=np.int32)
        self._check(self._lib.fc3d_hip_problem_order(self._h, iptr(o)), "order")
        return o

    def update(self, values=None, q=None, mu=None):
        """New W block values (same CSR order and structure), q and/or mu."""
        a = [None if v is None else np.ascontiguousarray(v, dtype=np.float64)
             for v in (values, q, mu)]
        self._check(self._lib.fc3d_hip_problem_update(
            self._h, *[None if v is None else dptr(v) for v in a]), "update")

    def body_layout(self):
        """0 no factors, 1 general 3x6 blocks, 2 frames layout."""
        return self._lib.fc3d_hip_problem_body_layout(self._h)

    def set_bodies(self, nbodies, body_a, body_b, H, minv, vfree, minv_diagonal=True):
        """Factors of W for velocity-form sweeps (fc3d_hip_problem_set_bodies)."""
        a = np.ascontiguousarray(body_a, dtype=np.int32)
        b = np.ascontiguousarray(body_b, dtype=np.int32)
        H = np.ascontiguousarray(H, dtype=np.float64)
        minv = np.ascontiguousarray(minv, dtype=np.float64)
        vfree = np.ascontiguousarray(vfree, dtype=np.float64)
        self._check(self._lib.fc3d_hip_problem_set_bodies(
            self._h, nbodies, iptr(a), iptr(b), dptr(H), dptr(minv), int(minv_diagonal),
            dptr(vfree)), "set_bodies")

    def set_reaction(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._lib.fc3d_hip_set_reaction(self._h, dptr(r)), "set_reaction")

    def reaction(self):
        r = np.empty(3 * self.nc)
        self._check(self._lib.fc3d_hip_get_reaction(self._h, dptr(r)), "get_reaction")
        return r

    def velocity(self):
        u = np.empty(3 * self.nc)
        self._check(self._lib.fc3d_hip_get_velocity(self._h, dptr(u)), "get_velocity")
        return u

    def solve(self, params):
        s = capi.HipStats()
        self._check(self._lib.fc3d_hip_solve(self._h, C.byref(params), C.byref(s)), "solve")
        return s

    def compute_error(self, norm=-1.0):
        e = C.c_double()
        self._check(self._lib.fc3d_hip_compute_error(self._h, norm, C.byref(e)), "compute_error")
        return e.value

    def gemv(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.array(y, dtype=np.float64)
        self._check(self._lib.fc3d_hip_gemv(self._h, dptr(x), dptr(y)), "gemv")
        return y

    def features(self, target=True):
        """(features [nc, 22], scaled resident reaction [nc, 3] or None)."""
        f = np.empty((self.nc, MLP_FEATURES), dtype=np.float32)
        t = np.empty((self.nc, 3), dtype=np.float32) if target else None
        self._check(self._lib.fc3d_hip_problem_features(
            self._h, fptr(f), fptr(t) if target else None), "features")
        return f, t

    def set_model(self, model):
        """Model for warm_start=2 (kept alive by this object)."""
        self._model = model
        self._check(self._lib.fc3d_hip_problem_set_model(
            self._h, model._h if model is not None else None), "set_model")

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.fc3d_hip_problem_destroy(h)
            self._h = None


MLP_FEATURES = 22
MLP_NPARAMS = 39312


class Mlp:
    """The warm-start network of fc3d_hip.h (fc3d_hip_mlp_*)."""

    def __init__(self, seed=0, device=0):
        self._lib = capi.load()
        h = C.c_void_p()
        rc = self._lib.fc3d_hip_mlp_create(seed, device, C.byref(h))
        if rc:
            raise RuntimeError(f"fc3d_hip_mlp_create: {rc} {capi.last_error()}")
        self._h = h

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what}: {rc} {capi.last_error()}")

    def params(self):
        p = np.empty(MLP_NPARAMS, dtype=np.float32)
        m = np.empty(MLP_FEATURES, dtype=np.float32)
        sd = np.empty(MLP_FEATURES, dtype=np.float32)
        self._check(self._lib.fc3d_hip_mlp_get_params(self._h, fptr(p), fptr(m), fptr(sd)),
                    "mlp_get_params")
        return p, m, sd

    def set_params(self, params=None, mean=None, std=None):
        a = [None if v is None else np.ascontiguousarray(v, dtype=np.float32)
             for v in (params, mean, std)]
        self._check(self._lib.fc3d_hip_mlp_set_params(
            self._h, *[None if v is None else fptr(v) for v in a]), "mlp_set_params")

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.empty((len(x), 3), dtype=np.float32)
        self._check(self._lib.fc3d_hip_mlp_forward(self._h, len(x), fptr(x), fptr(y)),
                    "mlp_forward")
        return y

    def gradients(self, x, t):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(MLP_NPARAMS, dtype=np.float32)
        loss = C.c_float()
        self._check(self._lib.fc3d_hip_mlp_gradients(self._h, len(x), fptr(x), fptr(t), fptr(g),
                                                     C.byref(loss)), "mlp_gradients")
        return g, loss.value

    def train(self, x, t, epochs=30, batch=1024, lr=1e-3, seed=0, set_normalisation=True):
        x = np.ascontiguousarray(x, dtype=np.float32)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = C.c_float()
        self._check(self._lib.fc3d_hip_mlp_train(self._h, len(x), fptr(x), fptr(t), epochs,
                                                 batch, lr, seed, int(set_normalisation),
                                                 C.byref(loss)), "mlp_train")
        return loss.value

    def set_default(self):
        """Model used by fc3d_driver / fc3d_nsgs_hip when iparam[21] = 2."""
        self._lib.fc3d_hip_set_default_model(self._h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.fc3d_hip_mlp_destroy(h)
            self._h = None


class Batch:
    """Many independent problems, solved in one persistent launch
    (fc3d_hip_batch_*).  ``problems`` is a list of (SparseBlock3, q, mu)."""

    def __init__(self, problems, device=0):
        self._lib = capi.load()
        self.sizes = np.array([W.nc for W, _, _ in problems], dtype=np.int32)
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        rps, cis, vals, qs, mus = [0], [], [], [], []
        base_b = 0
        for (W, q, mu), c0 in zip(problems, self.offsets[:-1]):
            rps.append(W.row_ptr[1:] + base_b)
            cis.append(W.col_idx + c0)
            vals.append(W.values)
            qs.append(np.asarray(q, dtype=np.float64))
            mus.append(np.asarray(mu, dtype=np.float64))
            base_b += W.nnzb
        self.row_ptr = np.concatenate([np.array([0], dtype=np.int64)] + [np.asarray(x, dtype=np.int64) for x in rps[1:]])
        self.col_idx = np.ascontiguousarray(np.concatenate(cis), dtype=np.int64)
        self.values = np.ascontiguousarray(np.concatenate(vals))
        self.q = np.ascontiguousarray(np.concatenate(qs))
        self.mu = np.ascontiguousarray(np.concatenate(mus))
        self.ntotal = int(self.offsets[-1])
        h = C.c_void_p()
        rc = self._lib.fc3d_hip_batch_create(len(problems), iptr(self.sizes), lptr(self.row_ptr),
                                             lptr(self.col_idx), dptr(self.values), dptr(self.q),
                                             dptr(self.mu), device, C.byref(h))
        if rc:
            raise RuntimeError(f"fc3d_hip_batch_create: {rc} {capi.last_error()}")
        self._h = h

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what}: {rc} {capi.last_error()}")

    def order(self):
        o = np.empty(self.ntotal, dtype=np.int32)
        self._check(self._lib.fc3d_hip_batch_order(self._h, iptr(o)), "order")
        return o

    def set_reaction(self, r):
        r = np.ascontiguousarray(r, dtype=np.float64)
        self._check(self._lib.fc3d_hip_batch_set_reaction(self._h, dptr(r)), "set_reaction")

    def reaction(self):
        r = np.empty(3 * self.ntotal)
        self._check(self._lib.fc3d_hip_batch_get_reaction(self._h, dptr(r)), "get_reaction")
        return r

    def velocity(self):
        u = np.empty(3 * self.ntotal)
        self._check(self._lib.fc3d_hip_batch_get_velocity(self._h, dptr(u)), "get_velocity")
        return u

    def solve(self, params):
        n = len(self.sizes)
        s = capi.HipBatchStats()
        info = np.empty(n, dtype=np.int32)
        iters = np.empty(n, dtype=np.int32)
        res = np.empty(n)
        self._check(self._lib.fc3d_hip_batch_solve(self._h, C.byref(params), C.byref(s),
                                                   iptr(info), iptr(iters), dptr(res)), "solve")
        return s, info, iters, res

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.fc3d_hip_batch_destroy(h)
            self._h = None


def generate_batch(nprob, nc_lo, nc_hi, mu_lo, mu_hi, seed, bodies_per_contact=0.4):
    """Independent random-pair problems with nc ~ U[nc_lo, nc_hi] and
    bodies = nc / 2.5 (seeded)."""
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nc_lo, nc_hi + 1, size=nprob)
    out = []
    for k, n in enumerate(sizes):
        nb = max(3, int(round(n * bodies_per_contact)))
        while n > nb * (nb - 1) // 2:
            nb += 1
        g = generate_random_pairs(nb, int(n), mu_lo, mu_hi, seed * 1000003 + k)
        out.append((g.W, g.q, g.mu))
    return out
