"""ctypes loader of the CPU oracle (test infrastructure only).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the
reference results are compared against; the product (siconos_amd) never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "liboracle_fc3d.so")

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
lp = C.POINTER(C.c_int64)
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    sig = {
        "oracle_set_mutation": (None, [C.c_int]),
        "oracle_mvp3x3": (None, [dp, dp, dp]),
        "oracle_mm3x3": (None, [dp, dp, dp]),
        "oracle_solve_3x3_gepp": (C.c_int, [dp, dp]),
        "oracle_orthoBaseFromVector": (C.c_int, [dp, dp, dp]),
        "oracle_projectionOnCone": (C.c_uint, [dp, C.c_double]),
        "oracle_projectionOnDualCone": (C.c_uint, [dp, C.c_double]),
        "oracle_computeAlartCurnierSTD": (None, [dp, dp, C.c_double, dp, dp, dp, dp]),
        "oracle_computeAlartCurnierJeanMoreau": (None, [dp, dp, C.c_double, dp, dp, dp, dp]),
        "oracle_alart_curnier_dual": (None, [dp, dp, C.c_double, dp, dp, dp, dp]),
        "oracle_rho_split_spectral_norm": (None, [dp, dp]),
        "oracle_rho_split_spectral_norm_cond": (None, [dp, dp]),
        "oracle_eig_3x3": (None, [dp, dp]),
        "oracle_rho_spectral_norm": (None, [dp, dp]),
        "oracle_unitary_compute_and_add_error": (None, [dp, dp, C.c_double, dp]),
        "oracle_fc3d_compute_error": (C.c_int, [C.c_void_p, dp, dp, C.c_double, C.c_double, dp]),
        "oracle_nsgs_options_new": (C.c_void_p, []),
        "oracle_local_options_new": (C.c_void_p, [C.c_int]),
        "oracle_set_local_solver": (None, [C.c_void_p, C.c_int]),
        "oracle_options_free": (None, [C.c_void_p]),
        "oracle_local_solve": (C.c_int, [dp, dp, C.c_double, dp, C.c_void_p, dp]),
        "oracle_local_state_init": (C.c_int, [dp, C.c_void_p, dp]),
        "oracle_fc3d_nsgs": (None, [C.c_void_p, dp, dp, ip, C.c_void_p, ip]),
        "oracle_fc3d_driver": (C.c_int, [C.c_void_p, dp, dp, C.c_void_p, ip]),
        "oracle_greedy_colouring": (C.c_int, [C.c_int, lp, lp, ip]),
        "oracle_problem_from_csr": (C.c_void_p, [C.c_int, lp, lp, dp, dp, dp]),
        "oracle_problem_free": (None, [C.c_void_p]),
        "oracle_solve_many": (C.c_int, [C.c_int, ip, lp, lp, lp, lp, dp, dp, dp, dp, C.c_double,
                                        C.c_int, C.c_int, ip, dp]),
        "oracle_fc2d_read": (C.c_void_p, [C.c_char_p]),
        "oracle_fc2d_free": (None, [C.c_void_p]),
        "oracle_fc2d_info": (C.c_int, [C.c_void_p, ip]),
        "oracle_fc2d_dense": (C.c_int, [C.c_void_p, dp, dp, dp]),
        "oracle_fc2d_local_q": (C.c_int, [C.c_void_p, C.c_int, dp, dp]),
        "oracle_fc2d_local_solve": (None, [dp, C.c_double, dp, C.c_double, dp]),
        "oracle_fc2d_compute_error": (C.c_int, [C.c_void_p, dp, dp, dp]),
        "oracle_fc2d_nsgs": (C.c_int, [C.c_void_p, dp, dp, C.c_double, C.c_int, C.c_int, ip,
                                       dp]),
        "oracle_fc2d_set_mutation": (None, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def d(a):
    return a.ctypes.data_as(dp)


def i32(a):
    return a.ctypes.data_as(ip)


def i64(a):
    return a.ctypes.data_as(lp)


CALLBACK_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, dp, dp, C.c_double, C.c_void_p)


class _Callback(C.Structure):
    _fields_ = [("env", C.c_void_p), ("collectStatsIteration", CALLBACK_FN)]


class _SO(C.Structure):
    # SolverOptions (fc3d_numerics.h)
    _fields_ = [("solverId", C.c_int), ("isSet", C.c_bool), ("iSize", C.c_size_t),
                ("iparam", ip), ("dSize", C.c_size_t), ("dparam", dp), ("filterOn", C.c_bool),
                ("dWorkSize", C.c_size_t), ("dWork", dp), ("iWorkSize", C.c_size_t),
                ("iWork", ip), ("numberOfInternalSolvers", C.c_size_t),
                ("internalSolvers", C.POINTER(C.c_void_p)), ("callback", C.POINTER(_Callback)),
                ("solverParameters", C.c_void_p), ("solverData", C.c_void_p)]


class Options:
    """Reference-default NSGS options (owned)."""

    def __init__(self, ptr=None, owner=None):
        self.lib = load()
        self.p = ptr if ptr is not None else self.lib.oracle_nsgs_options_new()
        self.owner = owner

    @staticmethod
    def local(solver_id):
        lib = load()
        return Options(lib.oracle_local_options_new(solver_id))

    @property
    def _s(self):
        return C.cast(self.p, C.POINTER(_SO)).contents

    @property
    def iparam(self):
        return np.ctypeslib.as_array(self._s.iparam, shape=(30,))

    @property
    def dparam(self):
        return np.ctypeslib.as_array(self._s.dparam, shape=(30,))

    def internal(self):
        return Options(self._s.internalSolvers[0], owner=self)

    def set_local_solver(self, sid):
        self.lib.oracle_set_local_solver(self.p, sid)

    def set_callback(self, fn):
        """fn(reaction, velocity, error) after every sweep (copies of the
        arrays), as the reference's collectStatsIteration is called."""
        if fn is None:
            self._s.callback = None
            self._cb = None
            return

        def tramp(env, size, r, u, err, extra):
            fn(np.ctypeslib.as_array(r, shape=(size,)).copy(),
               np.ctypeslib.as_array(u, shape=(size,)).copy(), err)

        f = CALLBACK_FN(tramp)
        cb = _Callback(None, f)
        self._cb = (f, cb)
        self._s.callback = C.pointer(cb)

    def __del__(self):
        if getattr(self, "p", None) and self.owner is None:
            if getattr(self, "_cb", None):
                self._s.callback = None
            self.lib.oracle_options_free(self.p)
            self.p = None


class Problem:
    def __init__(self, W, q, mu):
        self.lib = load()
        self.nc = W.nc
        self.q = np.ascontiguousarray(q, dtype=np.float64)
        self.mu = np.ascontiguousarray(mu, dtype=np.float64)
        self.p = self.lib.oracle_problem_from_csr(W.nc, i64(W.row_ptr), i64(W.col_idx),
                                                  d(W.values), d(self.q), d(self.mu))

    def __del__(self):
        if getattr(self, "p", None):
            self.lib.oracle_problem_free(self.p)
            self.p = None


def nsgs(W, q, mu, reaction=None, options=None, order=None, tol=None, max_iter=None):
    """Runs the oracle fc3d_driver; returns (info, r, u, iterations, residual)."""
    lib = load()
    pb = Problem(W, q, mu)
    o = options or Options()
    if tol is not None:
        o.dparam[0] = tol
    if max_iter is not None:
        o.iparam[0] = max_iter
    r = np.zeros(3 * W.nc) if reaction is None else np.array(reaction, dtype=np.float64)
    u = np.zeros(3 * W.nc)
    ordp = None
    if order is not None:
        order = np.ascontiguousarray(order, dtype=np.int32)
        ordp = i32(order)
    info = lib.oracle_fc3d_driver(pb.p, d(r), d(u), o.p, ordp)
    return info, r, u, int(o.iparam[1]), float(o.dparam[1])


def compute_error(W, q, mu, r, norm=None):
    lib = load()
    pb = Problem(W, q, mu)
    u = np.zeros(3 * W.nc)
    e = C.c_double()
    if norm is None:
        norm = float(np.sqrt(np.sum(pb.q * pb.q)))
    r = np.ascontiguousarray(r, dtype=np.float64)
    lib.oracle_fc3d_compute_error(pb.p, d(r), d(u), 1.0, norm, C.byref(e))
    return e.value, u


def greedy_colouring(W):
    lib = load()
    c = np.empty(W.nc, dtype=np.int32)
    n = lib.oracle_greedy_colouring(W.nc, i64(W.row_ptr), i64(W.col_idx), i32(c))
    return c, n


def solve_many(problems, tol, max_iter, nthreads):
    """oracle_solve_many: every (W, q, mu) from a zero start in natural order,
    one problem per thread; returns (converged, iterations, residuals, r)."""
    lib = load()
    nc = np.array([W.nc for W, _, _ in problems], dtype=np.int32)
    rp_off = np.zeros(len(problems), dtype=np.int64)
    b_off = np.zeros(len(problems), dtype=np.int64)
    rps, cis, bls = [], [], []
    ro = bo = 0
    for k, (W, _, _) in enumerate(problems):
        rp_off[k] = ro
        b_off[k] = bo
        rps.append(np.asarray(W.row_ptr, dtype=np.int64) + bo)
        cis.append(np.asarray(W.col_idx, dtype=np.int64))
        bls.append(np.asarray(W.values, dtype=np.float64).ravel())
        ro += W.nc + 1
        bo += W.nnzb
    rp = np.ascontiguousarray(np.concatenate(rps))
    ci = np.ascontiguousarray(np.concatenate(cis))
    bl = np.ascontiguousarray(np.concatenate(bls))
    q = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for _, x, _ in problems]))
    mu = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for _, _, x in problems]))
    r = np.zeros(len(q))
    iters = np.zeros(len(problems), dtype=np.int32)
    res = np.zeros(len(problems))
    ok = lib.oracle_solve_many(len(problems), i32(nc), i64(rp_off), i64(b_off), i64(rp), i64(ci),
                               d(bl), d(q), d(mu), d(r), tol, max_iter, nthreads, i32(iters),
                               d(res))
    return ok, iters, res, r


def mutation(m):
    load().oracle_set_mutation(m)


class Fc2dProblem:
    """A dimension-2 .dat problem read by the oracle's own parser."""

    def __init__(self, path):
        self.lib = load()
        self.h = self.lib.oracle_fc2d_read(path.encode())
        if not self.h:
            raise ValueError(f"oracle_fc2d_read failed on {path}")
        info = np.zeros(3, dtype=np.int32)
        self.lib.oracle_fc2d_info(self.h, info.ctypes.data_as(ip))
        self.nc, self.storage, self.origin = (int(x) for x in info)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.oracle_fc2d_free(self.h)
            self.h = None

    def dense(self):
        n = 2 * self.nc
        W = np.zeros((n, n), order="F")
        q = np.zeros(n)
        mu = np.zeros(self.nc)
        self.lib.oracle_fc2d_dense(self.h, W.ctypes.data_as(dp), d(q), d(mu))
        return W, q, mu

    def local_q(self, c, r):
        x = np.array(r, dtype=np.float64)
        y = np.zeros(2)
        self.lib.oracle_fc2d_local_q(self.h, c, d(x), d(y))
        return y

    def compute_error(self, r):
        z = np.ascontiguousarray(r, dtype=np.float64)
        w = np.zeros(2 * self.nc)
        e = C.c_double()
        self.lib.oracle_fc2d_compute_error(self.h, d(z), d(w), C.byref(e))
        return e.value, w

    def nsgs(self, tol=1e-5, max_iter=10000, mode=2, order=None, reaction=None):
        """oracle_fc2d_nsgs; mode: 1 LIGHT, 2 LIGHT_WITH_FULL_FINAL (others
        are not evaluated); returns (info, r, u, iterations, residual)."""
        r = np.zeros(2 * self.nc) if reaction is None else np.array(reaction, dtype=np.float64)
        u = np.zeros(2 * self.nc)
        out = np.zeros(2)
        op = None
        if order is not None:
            order = np.ascontiguousarray(order, dtype=np.int32)
            op = order.ctypes.data_as(ip)
        info = self.lib.oracle_fc2d_nsgs(self.h, d(r), d(u), tol, max_iter, mode, op, d(out))
        return info, r, u, int(out[0]), float(out[1])


def fc2d_local_solve(W, q, mu):
    W = np.ascontiguousarray(W, dtype=np.float64).ravel(order="F") if np.ndim(W) == 2 else \
        np.ascontiguousarray(W, dtype=np.float64)
    D = W[0] * W[3] - W[1] * W[2]
    P = np.zeros(2)
    load().oracle_fc2d_local_solve(d(W), D, d(np.ascontiguousarray(q, dtype=np.float64)), mu,
                                   d(P))
    return P


def fc2d_mutation(m):
    load().oracle_fc2d_set_mutation(m)
