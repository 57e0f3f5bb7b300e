"""ctypes declarations of the C-ABI (include/fc3d_numerics.h, include/fc3d_hip.h).

Loading fails loudly when ``siconos_amd/lib/libsiconos_amd.so`` is absent:
there is no fallback implementation.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libsiconos_amd.so")

c_int_p = C.POINTER(C.c_int)
c_float_p = C.POINTER(C.c_float)
c_double_p = C.POINTER(C.c_double)
c_int64_p = C.POINTER(C.c_int64)
c_size_t_p = C.POINTER(C.c_size_t)


class SparseBlockStructuredMatrix(C.Structure):
    _fields_ = [("nbblocks", C.c_size_t), ("block", C.POINTER(c_double_p)),
                ("blocknumber0", C.c_size_t), ("blocknumber1", C.c_size_t),
                ("blocksize0", c_size_t_p), ("blocksize1", c_size_t_p),
                ("filled1", C.c_size_t), ("filled2", C.c_size_t),
                ("index1_data", c_size_t_p), ("index2_data", c_size_t_p),
                ("diagonal_blocks", c_size_t_p), ("version", C.c_uint64)]


class CSparseMatrix(C.Structure):
    _fields_ = [("nzmax", C.c_int64), ("m", C.c_int64), ("n", C.c_int64), ("p", c_int64_p),
                ("i", c_int64_p), ("x", c_double_p), ("nz", C.c_int64)]


class NumericsSparseMatrix(C.Structure):
    _fields_ = [("triplet", C.POINTER(CSparseMatrix)), ("half_triplet", C.POINTER(CSparseMatrix)),
                ("csc", C.POINTER(CSparseMatrix)), ("trans_csc", C.POINTER(CSparseMatrix)),
                ("csr", C.POINTER(CSparseMatrix)), ("diag_indx", c_int64_p),
                ("origin", C.c_int), ("linearSolverParams", C.c_void_p),
                ("versions", C.c_uint64 * 5)]


class NumericsMatrix(C.Structure):
    pass


NumericsMatrix._fields_ = [("storageType", C.c_int), ("size0", C.c_int), ("size1", C.c_int),
                           ("matrix0", c_double_p),
                           ("matrix1", C.POINTER(SparseBlockStructuredMatrix)),
                           ("matrix2", C.POINTER(NumericsSparseMatrix)),
                           ("internalData", C.c_void_p),
                           ("version", C.c_uint64), ("destructible", C.POINTER(NumericsMatrix))]


class FrictionContactProblem(C.Structure):
    _fields_ = [("dimension", C.c_int), ("numberOfContacts", C.c_int),
                ("M", C.POINTER(NumericsMatrix)), ("q", c_double_p), ("mu", c_double_p)]


CALLBACK_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_size_t, c_double_p, c_double_p, C.c_double,
                          C.c_void_p)


# fc3d_hip_sweep_observer: (env, sweeps_done, incr_error)
OBSERVER_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.c_double)


class Callback(C.Structure):
    _fields_ = [("env", C.c_void_p), ("collectStatsIteration", CALLBACK_FN)]


class SolverOptions(C.Structure):
    pass


SolverOptions._fields_ = [("solverId", C.c_int), ("isSet", C.c_bool), ("iSize", C.c_size_t),
                          ("iparam", c_int_p), ("dSize", C.c_size_t), ("dparam", c_double_p),
                          ("filterOn", C.c_bool), ("dWorkSize", C.c_size_t),
                          ("dWork", c_double_p), ("iWorkSize", C.c_size_t), ("iWork", c_int_p),
                          ("numberOfInternalSolvers", C.c_size_t),
                          ("internalSolvers", C.POINTER(C.POINTER(SolverOptions))),
                          ("callback", C.POINTER(Callback)), ("solverParameters", C.c_void_p),
                          ("solverData", C.c_void_p)]


class HipParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol", C.c_double), ("error_evaluation", C.c_int),
                ("error_frequency", C.c_int), ("filter_local_solution", C.c_int),
                ("relaxation", C.c_int), ("omega", C.c_double), ("local_solver", C.c_int),
                ("local_max_iter", C.c_int), ("local_tol", C.c_double),
                ("local_formulation", C.c_int), ("rho_strategy", C.c_int),
                ("rho_constant", C.c_double), ("ls_max_iter", C.c_int),
                ("hybrid_strategy", C.c_int), ("hybrid_max_loop", C.c_int),
                ("hybrid_max_iter", C.c_int), ("precision", C.c_int), ("warm_start", C.c_int),
                ("schedule", C.c_int), ("sweeps_per_launch", C.c_int), ("fp32_tol", C.c_double),
                ("fp32_local_tol", C.c_double), ("instrument", C.c_int),
                ("accel_depth", C.c_int), ("accel_every", C.c_int), ("accel_reg", C.c_double),
                ("velocity_form", C.c_int), ("fp32_local_solver", C.c_int),
                ("freeze_contacts", C.c_int), ("focus_passes", C.c_int),
                ("focus_every", C.c_int), ("focus_after", C.c_int), ("focus_count", C.c_int),
                ("focus_share", C.c_double), ("start_trial_sweeps", C.c_int),
                ("focus_omega", C.c_double)]


class HipStats(C.Structure):
    _fields_ = [("info", C.c_int), ("iterations", C.c_int), ("fp32_sweeps", C.c_int),
                ("residual", C.c_double), ("incr_error", C.c_double), ("num_colours", C.c_int),
                ("warm_start_used", C.c_int), ("solve_ms", C.c_double),
                ("sweep_kernel_ms", C.c_double), ("sweep_kernel_launches", C.c_int64),
                ("sweep_kernel_bytes", C.c_double), ("schedule", C.c_int),
                ("start_residual_fallback", C.c_double), ("start_residual_learned", C.c_double),
                ("accel_steps", C.c_int), ("accel_restarts", C.c_int),
                ("accel_rejects", C.c_int), ("sweep_kernel_ms_fp32", C.c_double),
                ("sweep_kernel_launches_fp32", C.c_int64),
                ("sweep_kernel_bytes_fp32", C.c_double),
                ("sweep_tail_ms_fp32", C.c_double), ("sweep_tail_launches_fp32", C.c_int64),
                ("sweep_tail_bytes_fp32", C.c_double), ("start_trial_sweeps", C.c_int),
                ("focus_rounds", C.c_int),
                ("focus_listed", C.c_int64), ("focus_due", C.c_int),
                ("focus_last_share", C.c_double), ("focus_lds_rounds", C.c_int)]


class Fc2dParams(C.Structure):
    _fields_ = [("max_iter", C.c_int), ("tol", C.c_double), ("error_evaluation", C.c_int),
                ("schedule", C.c_int)]


class Fc2dStats(C.Structure):
    _fields_ = [("info", C.c_int), ("iterations", C.c_int), ("residual", C.c_double),
                ("num_colours", C.c_int), ("velocity_set", C.c_int), ("solve_ms", C.c_double),
                ("schedule", C.c_int)]


class HipBatchStats(C.Structure):
    _fields_ = [("solve_ms", C.c_double), ("converged", C.c_int),
                ("converged_contacts", C.c_int64), ("total_sweeps", C.c_int64),
                ("fp32_sweeps", C.c_int64), ("max_residual", C.c_double), ("schedule", C.c_int)]


class Generated(C.Structure):
    _fields_ = [("nc", C.c_int), ("nbodies", C.c_int), ("nnzb", C.c_int64),
                ("row_ptr", c_int64_p), ("col_idx", c_int64_p), ("blocks", c_double_p),
                ("q", c_double_p), ("mu", c_double_p), ("body_a", c_int_p), ("body_b", c_int_p),
                ("H", c_double_p), ("minv", c_double_p), ("vfree", c_double_p)]


FCP_p = C.POINTER(FrictionContactProblem)
SO_p = C.POINTER(SolverOptions)
NM_p = C.POINTER(NumericsMatrix)
SBM_p = C.POINTER(SparseBlockStructuredMatrix)
H_p = C.c_void_p  # fc3d_hip_problem*

# name: (restype, argtypes)
SIGNATURES = {
    "numerics_error_string": (C.c_char_p, [C.c_int]),
    "solver_registry_register": (C.c_int, [C.c_void_p]),
    "solver_registry_lookup": (C.c_void_p, [C.c_int]),
    "solver_registry_lookup_by_name": (C.c_void_p, [C.c_char_p]),
    "solver_registry_count": (C.c_size_t, []),
    "solver_registry_exists": (C.c_int, [C.c_int]),
    "solver_options_create": (SO_p, [C.c_int]),
    "solver_options_create_by_name": (SO_p, [C.c_char_p]),
    "solver_options_delete": (None, [SO_p]),
    "solver_options_copy": (SO_p, [SO_p]),
    "solver_options_update_internal": (None, [SO_p, C.c_size_t, C.c_int]),
    "solver_options_id_to_name": (C.c_char_p, [C.c_int]),
    "solver_options_name_to_id": (C.c_int, [C.c_char_p]),
    "solver_options_get_internal_solver": (SO_p, [SO_p, C.c_size_t]),
    "NM_create_from_data": (NM_p, [C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "NM_free": (NM_p, [NM_p]),
    "SBM_new": (SBM_p, []),
    "SBM_free_all": (None, [SBM_p]),
    "SBM_new_3x3_from_csr": (SBM_p, [C.c_size_t, c_int64_p, c_int64_p, c_double_p]),
    "SBM_compute_diagonal_block_indices": (c_size_t_p, [SBM_p]),
    "SBM_diagonal_block_index": (C.c_size_t, [SBM_p, C.c_size_t]),
    "SBM_from_dense_3x3": (SBM_p, [C.c_int, c_double_p]),
    "SBM_gemv_3x3": (C.c_int, [C.c_size_t, C.c_size_t, SBM_p, c_double_p, c_double_p]),
    "NM_prod_mv_3x3": (C.c_int, [C.c_int, C.c_int, NM_p, c_double_p, c_double_p]),
    "frictionContactProblem_new": (FCP_p, []),
    "frictionContactProblem_new_with_data": (FCP_p, [C.c_int, C.c_size_t, NM_p, c_double_p,
                                                     c_double_p]),
    "frictionContactProblem_free": (None, [FCP_p]),
    "frictionContact_printInFile": (C.c_int, [FCP_p, C.c_void_p]),
    "frictionContact_printInFilename": (C.c_int, [FCP_p, C.c_char_p]),
    "frictionContact_newFromFile": (FCP_p, [C.c_void_p]),
    "frictionContact_new_from_filename": (FCP_p, [C.c_char_p]),
    "fc3d_checkTrivialCase": (C.c_int, [FCP_p, c_double_p, c_double_p, SO_p]),
    "fc3d_driver": (C.c_int, [FCP_p, c_double_p, c_double_p, SO_p]),
    "fc2d_driver": (C.c_int, [FCP_p, c_double_p, c_double_p, SO_p]),
    "fc2d_compute_error": (C.c_int, [FCP_p, c_double_p, c_double_p, C.c_double, C.c_double,
                                     c_double_p]),
    "NM_csc": (C.POINTER(CSparseMatrix), [C.POINTER(NumericsMatrix)]),
    "NSM_new_from_file": (C.POINTER(NumericsSparseMatrix), [C.c_void_p]),
    "NSM_write_in_file": (C.c_int, [C.POINTER(NumericsSparseMatrix), C.c_void_p]),
    "NSM_free": (C.POINTER(NumericsSparseMatrix), [C.POINTER(NumericsSparseMatrix)]),
    "fc2d_nsgs_hip": (None, [FCP_p, c_double_p, c_double_p, c_int_p, SO_p]),
    "fc2d_nsgs_hip_set_default": (None, [SO_p]),
    "fc3d_hip_problem_body_layout": (C.c_int, [C.c_void_p]),
    "fc2d_hip_params_default": (None, [C.c_void_p]),
    "fc2d_problem_to_device": (C.c_int, [FCP_p, C.c_int, C.POINTER(C.c_void_p)]),
    "fc2d_hip_problem_create": (C.c_int, [C.c_int, c_int64_p, c_int64_p, c_double_p, C.c_int,
                                          C.c_int, c_double_p, c_double_p, c_double_p, C.c_int,
                                          C.POINTER(C.c_void_p)]),
    "fc2d_hip_problem_destroy": (None, [C.c_void_p]),
    "fc2d_hip_problem_num_colours": (C.c_int, [C.c_void_p]),
    "fc2d_hip_problem_order": (C.c_int, [C.c_void_p, c_int_p]),
    "fc2d_hip_set_reaction": (C.c_int, [C.c_void_p, c_double_p]),
    "fc2d_hip_get_reaction": (C.c_int, [C.c_void_p, c_double_p]),
    "fc2d_hip_get_velocity": (C.c_int, [C.c_void_p, c_double_p]),
    "fc2d_hip_solve": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    "fc2d_hip_compute_error": (C.c_int, [C.c_void_p, c_double_p, c_double_p, C.c_double,
                                         c_double_p]),
    "fc3d_solver_options_create": (SO_p, [C.c_int]),
    "fc3d_compute_error": (C.c_int, [FCP_p, c_double_p, c_double_p, C.c_double, SO_p,
                                     C.c_double, c_double_p]),
    "fc3d_nsgs_hip": (None, [FCP_p, c_double_p, c_double_p, c_int_p, SO_p]),
    "fc3d_nsgs_hip_set_default": (None, [SO_p]),
    "fc3d_generate_random_pairs": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double,
                                             C.c_uint64, C.POINTER(Generated)]),
    "fc3d_generate_granular_pile": (C.c_int, [C.c_int, C.c_int, C.c_double, C.c_double,
                                              C.c_uint64, C.POINTER(Generated)]),
    "fc3d_generated_free": (None, [C.POINTER(Generated)]),
    "fc3d_generated_to_problem": (FCP_p, [C.POINTER(Generated)]),
    "fc3d_hip_assemble": (C.c_int, [C.c_int, C.c_int, c_int_p, c_int_p, c_double_p, c_double_p,
                                    C.c_int, c_double_p, C.c_int, C.POINTER(Generated),
                                    c_double_p]),
    "fc3d_colour_contacts": (C.c_int, [C.c_int, c_int64_p, c_int64_p, c_int_p]),
    "fc3d_colouring_is_valid": (C.c_int, [C.c_int, c_int64_p, c_int64_p, c_int_p]),
    # fc3d_hip.h
    "fc3d_hip_device_count": (C.c_int, []),
    "fc3d_hip_last_error": (C.c_char_p, []),
    "fc3d_hip_params_default": (None, [C.POINTER(HipParams)]),
    "fc3d_hip_problem_create": (C.c_int, [C.c_int, c_int64_p, c_int64_p, c_double_p, c_double_p,
                                          c_double_p, C.c_int, C.POINTER(H_p)]),
    "fc3d_hip_problem_destroy": (None, [H_p]),
    "fc3d_hip_problem_num_colours": (C.c_int, [H_p]),
    "fc3d_hip_problem_nnzb": (C.c_int64, [H_p]),
    "fc3d_hip_problem_colours": (C.c_int, [H_p, c_int_p]),
    "fc3d_hip_problem_order": (C.c_int, [H_p, c_int_p]),
    "fc3d_hip_problem_update": (C.c_int, [H_p, c_double_p, c_double_p, c_double_p]),
    "fc3d_hip_problem_set_bodies": (C.c_int, [H_p, C.c_int, c_int_p, c_int_p, c_double_p,
                                              c_double_p, C.c_int, c_double_p]),
    "fc3d_hip_set_reaction": (C.c_int, [H_p, c_double_p]),
    "fc3d_hip_get_reaction": (C.c_int, [H_p, c_double_p]),
    "fc3d_hip_get_velocity": (C.c_int, [H_p, c_double_p]),
    "fc3d_hip_solve": (C.c_int, [H_p, C.POINTER(HipParams), C.POINTER(HipStats)]),
    "fc3d_hip_solve_observed": (C.c_int, [H_p, C.POINTER(HipParams), C.POINTER(HipStats),
                                          OBSERVER_FN, C.c_void_p]),
    "fc3d_hip_compute_error": (C.c_int, [H_p, C.c_double, c_double_p]),
    "fc3d_hip_compute_error_device": (C.c_int, [C.c_int, C.c_int64, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                C.c_double, C.c_void_p, C.c_void_p, C.c_int,
                                                C.c_void_p]),
    "fc3d_hip_resident_clear": (None, [NM_p]),
    "fc3d_hip_resident_stats": (None, [c_int64_p, c_int64_p, c_int64_p]),
    "fc3d_hip_gemv": (C.c_int, [H_p, c_double_p, c_double_p]),
    "fc3d_hip_alart_curnier": (C.c_int, [C.c_int, c_double_p, c_double_p, c_double_p, c_double_p,
                                         C.c_int, c_double_p, c_double_p, c_double_p]),
    "fc3d_hip_projection_on_cone": (C.c_int, [C.c_int, c_double_p, c_double_p]),
    "fc3d_hip_local_solve": (C.c_int, [C.c_int, c_double_p, c_double_p, c_double_p, c_double_p,
                                       C.POINTER(HipParams), c_double_p, C.c_int, c_int_p]),
    "fc3d_hip_colour": (C.c_int, [C.c_int, c_int64_p, c_int64_p, c_int_p]),
    "fc3d_hip_mlp_create": (C.c_int, [C.c_uint64, C.c_int, C.POINTER(H_p)]),
    "fc3d_hip_mlp_destroy": (None, [H_p]),
    "fc3d_hip_mlp_get_params": (C.c_int, [H_p, c_float_p, c_float_p, c_float_p]),
    "fc3d_hip_mlp_set_params": (C.c_int, [H_p, c_float_p, c_float_p, c_float_p]),
    "fc3d_hip_mlp_forward": (C.c_int, [H_p, C.c_int, c_float_p, c_float_p]),
    "fc3d_hip_mlp_gradients": (C.c_int, [H_p, C.c_int, c_float_p, c_float_p, c_float_p,
                                         c_float_p]),
    "fc3d_hip_mlp_train": (C.c_int, [H_p, C.c_int, c_float_p, c_float_p, C.c_int, C.c_int,
                                     C.c_float, C.c_uint64, C.c_int, c_float_p]),
    "fc3d_hip_problem_features": (C.c_int, [H_p, c_float_p, c_float_p]),
    "fc3d_hip_problem_set_model": (C.c_int, [H_p, H_p]),
    "fc3d_hip_set_default_model": (None, [H_p]),
    "fc3d_hip_default_model": (H_p, []),
    "fc3d_hip_batch_create": (C.c_int, [C.c_int, c_int_p, c_int64_p, c_int64_p, c_double_p,
                                        c_double_p, c_double_p, C.c_int, C.POINTER(H_p)]),
    "fc3d_hip_batch_destroy": (None, [H_p]),
    "fc3d_hip_batch_order": (C.c_int, [H_p, c_int_p]),
    "fc3d_hip_batch_set_reaction": (C.c_int, [H_p, c_double_p]),
    "fc3d_hip_batch_get_reaction": (C.c_int, [H_p, c_double_p]),
    "fc3d_hip_batch_get_velocity": (C.c_int, [H_p, c_double_p]),
    "fc3d_hip_batch_solve": (C.c_int, [H_p, C.POINTER(HipParams), C.POINTER(HipBatchStats),
                                       c_int_p, c_int_p, c_double_p]),
}

_lib = None


def load():
    """Returns the loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m siconos_amd.build` "
                           "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def last_error():
    return load().fc3d_hip_last_error().decode()
