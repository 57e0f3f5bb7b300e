/* fc3d_numerics.h — host-side drop-in interface for the 3-D friction-contact
 * NSGS hot path (C99).
 *
 * The struct layouts below keep the field order, types and meaning of the
 * Siconos numerics structs they stand for, so that a caller compiled against
 * Siconos hands the same memory to this library.  Every declaration cites the
 * reference interface it replaces (paths are relative to the Siconos tree).
 *
 * All exported functions carry FC3D_API.  Functions that do arithmetic on the
 * problem (fc3d_compute_error, SBM_gemv_3x3, NM_prod_mv_3x3, the FC3D_NSGS_HIP
 * solver) run on the GPU through the C-ABI declared in fc3d_hip.h; they return
 * a negative error code when no GPU or no HIP runtime is present.  There is no
 * CPU fallback.
 */
#ifndef FC3D_NUMERICS_H
#define FC3D_NUMERICS_H

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FC3D_API __attribute__((visibility("default")))

/* ------------------------------------------------------------------------ */
/* Error codes — numerics/src/utils/numerics_errors.h:41-67                 */
/* ------------------------------------------------------------------------ */
enum NUMERICS_ERROR_CODES {
  NUMERICS_OK = 0,
  NUMERICS_ERR_NULL_POINTER = -1,
  NUMERICS_ERR_INVALID_ARGUMENT = -2,
  NUMERICS_ERR_INVALID_OPTION = -3,
  NUMERICS_ERR_INVALID_SOLVER = -4,
  NUMERICS_ERR_NOT_IMPLEMENTED = -5,
  NUMERICS_ERR_MEMORY = -6,
  NUMERICS_ERR_FILE_IO = -7,
  NUMERICS_ERR_MAX_ITER = -20,
  NUMERICS_ERR_DIVERGENCE = -21,
  NUMERICS_ERR_STAGNATION = -22,
  NUMERICS_ERR_LINEAR_SOLVER = -23,
  NUMERICS_ERR_LOCAL_SOLVER = -24,
  NUMERICS_ERR_PROJECTION = -25,
  NUMERICS_ERR_INFEASIBLE = -40,
  NUMERICS_ERR_UNBOUNDED = -41,
  NUMERICS_ERR_ILL_CONDITIONED = -42,
  NUMERICS_ERR_SINGULAR_MATRIX = -43,
  NUMERICS_ERR_UNKNOWN = -99,
  /* no counterpart in the reference: the GPU runtime is missing or failed */
  NUMERICS_ERR_DEVICE = -60
};

/* numerics/src/utils/numerics_errors.h:70-116 */
FC3D_API const char* numerics_error_string(int code);

/* ------------------------------------------------------------------------ */
/* Matrices — numerics/src/linalg/NM_types.h:29-34, NumericsDataVersion.h:25-30,
 * SparseBlockMatrix.h:82-124, NumericsMatrix.h:73-92                        */
/* ------------------------------------------------------------------------ */
typedef enum NumericsMatrix_types {
  NM_DENSE = 0,
  NM_SPARSE_BLOCK = 1,
  NM_SPARSE = 2,
  NM_UNKNOWN = 3
} NM_types;

typedef uint64_t version_t;
typedef struct {
  version_t number;
} NumericsDataVersion;

typedef struct SparseBlockStructuredMatrix {
  size_t nbblocks;      /* number of non-null blocks                         */
  double** block;       /* block[k]: column-major values of block k          */
  size_t blocknumber0;  /* number of block rows                              */
  size_t blocknumber1;  /* number of block columns                           */
  size_t* blocksize0;   /* cumulated row sizes (3, 6, 9, ...)                */
  size_t* blocksize1;   /* cumulated column sizes                            */
  size_t filled1;       /* number of entries of index1_data (rows + 1)       */
  size_t filled2;       /* number of entries of index2_data (= nbblocks)     */
  size_t* index1_data;  /* row pointers                                      */
  size_t* index2_data;  /* block column indices                              */
  size_t* diagonal_blocks; /* cached index of the diagonal block of each row */
  NumericsDataVersion version;
} SparseBlockStructuredMatrix;

/* Compressed storage.  CSparseMatrix keeps the layout of CXSparse's
 * cs_dl_sparse (CSparseMatrix.h:37 with CS_LONG): p holds n + 1 column
 * pointers (CSC) or the column index of each entry (triplet), i the row
 * indices; nz = -1 marks CSC, -2 CSR, >= 0 a triplet of nz entries.
 * NumericsSparseMatrix keeps NumericsSparseMatrix.h:95-112.  On this path
 * an NM_SPARSE matrix is read from a .dat file (NSM_new_from_file,
 * NumericsSparseMatrix.c:555-583) and used by the fc2d solver. */
typedef struct CSparseMatrix {
  int64_t nzmax;
  int64_t m;
  int64_t n;
  int64_t* p;
  int64_t* i;
  double* x;
  int64_t nz;
} CSparseMatrix;

typedef enum { NSM_UNKNOWN, NSM_TRIPLET, NSM_CSC, NSM_CSR, NSM_HALF_TRIPLET } NSM_t;

typedef struct NumericsSparseMatrix {
  CSparseMatrix* triplet;
  CSparseMatrix* half_triplet;
  CSparseMatrix* csc;
  CSparseMatrix* trans_csc;
  CSparseMatrix* csr;
  int64_t* diag_indx;
  NSM_t origin;
  void* linearSolverParams; /* not used on this path */
  NumericsDataVersion versions[5];
} NumericsSparseMatrix;

typedef struct NumericsMatrix {
  NM_types storageType;
  int size0;
  int size1;
  double* matrix0;                      /* dense, column-major             */
  SparseBlockStructuredMatrix* matrix1; /* sparse block                    */
  NumericsSparseMatrix* matrix2;        /* compressed (NM_SPARSE)          */
  void* internalData;
  NumericsDataVersion version;
  struct NumericsMatrix* destructible;
} NumericsMatrix;

/* numerics/src/FrictionContact/FrictionContactProblem.h:33-46 */
typedef struct FrictionContactProblem {
  int dimension;
  int numberOfContacts;
  NumericsMatrix* M;
  double* q;
  double* mu;
} FrictionContactProblem;

/* ------------------------------------------------------------------------ */
/* Solver options — numerics/src/core/SolverOptions.h:36-121                */
/* ------------------------------------------------------------------------ */
typedef struct {
  void* env;
  void (*collectStatsIteration)(void* env, size_t size, double* reaction, double* velocity,
                                double error, void* extra_data);
} Callback;

#define OPTIONS_PARAM_SIZE 30

typedef struct SolverOptions {
  int solverId;
  bool isSet;
  size_t iSize;
  int* iparam;
  size_t dSize;
  double* dparam;
  bool filterOn;
  size_t dWorkSize;
  double* dWork;
  size_t iWorkSize;
  int* iWork;
  size_t numberOfInternalSolvers;
  struct SolverOptions** internalSolvers;
  Callback* callback;
  void* solverParameters;
  void* solverData;
} SolverOptions;

/* Common indices — SolverOptions.h:94-121 and naming_conventions.h:100-125 */
enum SICONOS_IPARAM {
  SICONOS_IPARAM_MAX_ITER = 0,
  SICONOS_IPARAM_ITER_DONE = 1,
  SICONOS_IPARAM_PREALLOC = 2,
  SICONOS_IPARAM_ERROR_EVALUATION = 3,
  SICONOS_IPARAM_NSGS_SHUFFLE = 5,
  SICONOS_IPARAM_PATHSEARCH_STACKSIZE = 19
};
enum SICONOS_DPARAM {
  SICONOS_DPARAM_TOL = 0,
  SICONOS_DPARAM_RESIDU = 1,
  SICONOS_DPARAM_TIME_BEFORE_LOOP = 20,
  SICONOS_DPARAM_TIME_IN_LOOP = 21,
  SICONOS_DPARAM_TIME_AFTER_LOOP = 22
};

/* ------------------------------------------------------------------------ */
/* Solver identifiers — FrictionContact_options.h:83-171                    */
/* ------------------------------------------------------------------------ */
enum FRICTION_SOLVER {
  SICONOS_FRICTION_3D_NSGS = 500,
  SICONOS_FRICTION_3D_ONECONTACT_NSN = 550,
  SICONOS_FRICTION_3D_ONECONTACT_NSN_GP = 551,
  SICONOS_FRICTION_3D_ONECONTACT_ProjectionOnCone = 552,
  SICONOS_FRICTION_3D_ONECONTACT_ProjectionOnConeWithLocalIteration = 553,
  SICONOS_FRICTION_3D_ONECONTACT_NSN_GP_HYBRID = 565,
  /* No counterpart in the reference: coloured NSGS on one MI355X (gfx950).
   * 540 is unused in the reference's 500-549 local-formulation range. */
  SICONOS_FRICTION_3D_NSGS_HIP = 540
};

/* Effective NSGS iparam / dparam indices.  The reference declares
 * SICONOS_FRICTION_3D_IPARAM_ERROR_EVALUATION = 7 in an enum
 * (FrictionContact_options.h:233) but then redefines the name as a macro for
 * index 13 (NonSmoothGaussSeidel_options.h:275); every use in fc3d_nsgs.c
 * therefore reads iparam[13].  The values below are the effective ones. */
enum SICONOS_FRICTION_3D_IPARAM {
  SICONOS_FRICTION_3D_IPARAM_INTERNAL_ERROR_STRATEGY = 2,
  SICONOS_FRICTION_3D_CURRENT_CONTACT_NUMBER = 5
};
enum SICONOS_NSGS_IPARAM { /* NonSmoothGaussSeidel_options.h:58-112 */
  SICONOS_NSGS_RELAXATION = 10,
  SICONOS_NSGS_SHUFFLE = 11,
  SICONOS_NSGS_SHUFFLE_SEED = 12,
  SICONOS_NSGS_ERROR_EVALUATION_TYPE = 13,
  SICONOS_NSGS_ERROR_EVALUATION_FREQUENCY = 14,
  SICONOS_NSGS_FREEZING_CONTACT = 15,
  SICONOS_NSGS_FILTER_LOCAL_SOLUTION = 16,
  SICONOS_NSGS_PRINTING_STYLE = 17
};
enum SICONOS_NSGS_DPARAM {
  SICONOS_NSGS_RELAXATION_VALUE = 10,
  SICONOS_NSGS_INTERNAL_ERROR_RATIO = 11
};
enum SICONOS_NSGS_ERROR_EVALUATION { /* NonSmoothGaussSeidel_options.h:121-140 */
  SICONOS_NSGS_ERROR_EVALUATION_FULL = 0,
  SICONOS_NSGS_ERROR_EVALUATION_LIGHT = 1,
  SICONOS_NSGS_ERROR_EVALUATION_LIGHT_WITH_FULL_FINAL = 2,
  SICONOS_NSGS_ERROR_EVALUATION_ADAPTIVE = 3
};
enum SICONOS_NSGS_INTERNAL_ERROR_STRATEGY {
  SICONOS_NSGS_INTERNAL_ERROR_STRATEGY_ADAPTIVE = 0,
  SICONOS_NSGS_INTERNAL_ERROR_STRATEGY_GIVEN_VALUE = 1,
  SICONOS_NSGS_INTERNAL_ERROR_STRATEGY_ADAPTIVE_N_CONTACT = 2
};
#define SICONOS_FRICTION_3D_IPARAM_ERROR_EVALUATION SICONOS_NSGS_ERROR_EVALUATION_TYPE
#define SICONOS_FRICTION_3D_DPARAM_INTERNAL_ERROR_RATIO SICONOS_NSGS_INTERNAL_ERROR_RATIO

/* One-contact Newton parameters — FrictionContact_options.h:295-366 */
enum SICONOS_FRICTION_3D_NSN_IPARAM {
  SICONOS_FRICTION_3D_NSN_RHO_STRATEGY = 9,
  SICONOS_FRICTION_3D_NSN_FORMULATION = 10,
  SICONOS_FRICTION_3D_NSN_LINESEARCH = 11,
  SICONOS_FRICTION_3D_NSN_LINESEARCH_MAX_ITER = 12,
  SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY = 14,
  SICONOS_FRICTION_3D_NSN_HYBRID_MAX_LOOP = 15,
  SICONOS_FRICTION_3D_NSN_HYBRID_MAX_ITER = 16
};
enum SICONOS_FRICTION_3D_NSN_DPARAM { SICONOS_FRICTION_3D_NSN_RHO = 3 };
enum SICONOS_FRICTION_3D_NSN_RHO_STRATEGY {
  SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_CONSTANT = 0,
  SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPECTRAL_NORM = 1,
  SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM_COND = 2,
  SICONOS_FRICTION_3D_NSN_FORMULATION_RHO_STRATEGY_SPLIT_SPECTRAL_NORM = 3
};
enum SICONOS_FRICTION_3D_NSN_FORMULATION {
  SICONOS_FRICTION_3D_NSN_FORMULATION_ALARTCURNIER_STD = 0,
  SICONOS_FRICTION_3D_NSN_FORMULATION_JEANMOREAU_STD = 1
};
enum SICONOS_FRICTION_3D_NSN_LINESEARCH {
  SICONOS_FRICTION_3D_NSN_LINESEARCH_GOLDSTEINPRICE = 0,
  SICONOS_FRICTION_3D_NSN_LINESEARCH_NO = -1
};
enum SICONOS_FRICTION_3D_NSN_HYBRID {
  SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_NO = 0,
  SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_PLI_NSN_LOOP = 1,
  SICONOS_FRICTION_3D_NSN_HYBRID_STRATEGY_NSN_AND_PLI_NSN_LOOP = 2
};

/* Extra parameters of SICONOS_FRICTION_3D_NSGS_HIP (slots 20-29 are unused by
 * the reference NSGS).  dparam[20..22] keep the reference meaning of
 * SICONOS_DPARAM_TIME_{BEFORE,IN,AFTER}_LOOP (set-up, sweeps, final check). */
enum SICONOS_FRICTION_3D_NSGS_HIP_IPARAM {
  SICONOS_FRICTION_3D_NSGS_HIP_PRECISION = 20,   /* 0 FP64, 1 FP32 sweeps + FP64 refinement */
  SICONOS_FRICTION_3D_NSGS_HIP_WARM_START = 21,  /* 0 caller's reaction, 1 zero, 2 learned */
  SICONOS_FRICTION_3D_NSGS_HIP_SCHEDULE = 22,    /* 0 autotuned, 1.. explicit schedule   */
  SICONOS_FRICTION_3D_NSGS_HIP_SWEEPS_PER_LAUNCH = 23, /* sweeps per graph replay       */
  SICONOS_FRICTION_3D_NSGS_HIP_DEVICE = 24,      /* HIP device ordinal                  */
  SICONOS_FRICTION_3D_NSGS_HIP_NUM_COLOURS = 25, /* out: colours of the contact graph   */
  SICONOS_FRICTION_3D_NSGS_HIP_FP32_SWEEPS = 26, /* out: sweeps done in FP32             */
  SICONOS_FRICTION_3D_NSGS_HIP_WARM_START_USED = 27, /* out: start actually used        */
  SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_DEPTH = 28,  /* Anderson history columns, 0 = off  */
  SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_EVERY = 29   /* Anderson step every n-th sweep      */
};
enum SICONOS_FRICTION_3D_NSGS_HIP_DPARAM {
  SICONOS_FRICTION_3D_NSGS_HIP_FP32_TOL = 23,    /* residual at which FP32 hands over   */
  SICONOS_FRICTION_3D_NSGS_HIP_TIME_SOLVE = 24,  /* out: device time of sweeps, ms      */
  SICONOS_FRICTION_3D_NSGS_HIP_ACCEL_REG = 25,   /* Anderson regularisation (x trace)    */
  SICONOS_FRICTION_3D_NSGS_HIP_FP32_LOCAL_SOLVER = 26, /* local solver id of the FP32 sweeps,
                                                         0 = the internal solver's         */
  SICONOS_FRICTION_3D_NSGS_HIP_FP32_LOCAL_TOL = 27,    /* local tolerance of the FP32 sweeps,
                                                         0 = the default (1e-6)             */
  /* Device residency across fc3d_driver / fc3d_compute_error calls on the same
   * NumericsMatrix: 0 (default) keeps the device problem (colouring, layout,
   * tuned schedule, graphs) while the block structure is unchanged and
   * re-uploads W's values, q and mu on every call; 1 also skips W's values
   * while the storage's NumericsDataVersion is unchanged and non-zero
   * (NumericsMatrix.c:112-135; the caller bumps it when values change);
   * -1 builds and frees a device problem per call. */
  SICONOS_FRICTION_3D_NSGS_HIP_RESIDENCY = 28,
  /* warm start 2: sweeps each start runs before the better one is kept
   * (0 = the default, one sweep) */
  SICONOS_FRICTION_3D_NSGS_HIP_START_TRIAL_SWEEPS = 29
};

/* ------------------------------------------------------------------------ */
/* Solver registry — numerics/src/utils/solver_registry.h:38-95             */
/* ------------------------------------------------------------------------ */
typedef int solver_id_t;
typedef int (*SolverInitFunc)(void* problem, SolverOptions* options);
typedef int (*SolverSolveFunc)(void* problem, double* var1, double* var2, SolverOptions* options);
typedef int (*SolverSolveFunc3)(void* problem, double* var1, double* var2, double* var3,
                                SolverOptions* options);
typedef void (*SolverFreeFunc)(void* problem, SolverOptions* options);
typedef double (*SolverErrorFunc)(void* problem, double* reaction, double* velocity,
                                  double* work);
typedef void (*SolverSetDefaultFunc)(SolverOptions* options);

typedef struct {
  solver_id_t id;
  const char* name;
  const char* description;
  SolverInitFunc init;
  union {
    SolverSolveFunc solve;
    SolverSolveFunc3 solve3;
  };
  SolverFreeFunc free;
  SolverErrorFunc error;
  SolverSetDefaultFunc set_default;
  int default_max_iter;
  double default_tol;
  bool is_local_solver;
  bool has_3var;
} SolverEntry;

#define SOLVER_REGISTRY_MAX 256

/* solver_registry.c:28-95 */
FC3D_API int solver_registry_register(const SolverEntry* entry);
FC3D_API const SolverEntry* solver_registry_lookup(solver_id_t id);
FC3D_API const SolverEntry* solver_registry_lookup_by_name(const char* name);
FC3D_API size_t solver_registry_count(void);
FC3D_API int solver_registry_exists(solver_id_t id);

/* Registration at load time, as solver_registry.h:75-95 does it. */
#define FC3D_REGISTER_SOLVER(solver_id, name_str, desc_str, init_fn, solve_fn, free_fn,      \
                             err_fn, set_default_fn, max_iter_val, tol_val, local_val)      \
  static int fc3d_solver_reg_##solver_id(void) __attribute__((constructor, used));         \
  static int fc3d_solver_reg_##solver_id(void) {                                           \
    static const SolverEntry entry = {.id = solver_id,                                     \
                                      .name = name_str,                                    \
                                      .description = desc_str,                             \
                                      .init = init_fn,                                     \
                                      .solve = solve_fn,                                   \
                                      .free = free_fn,                                     \
                                      .error = err_fn,                                     \
                                      .set_default = set_default_fn,                       \
                                      .default_max_iter = max_iter_val,                    \
                                      .default_tol = tol_val,                              \
                                      .is_local_solver = local_val,                        \
                                      .has_3var = false};                                  \
    return solver_registry_register(&entry);                                               \
  }

/* ------------------------------------------------------------------------ */
/* SolverOptions — numerics/src/core/SolverOptions.c:44-94, 176-268, 310-379 */
/* ------------------------------------------------------------------------ */
FC3D_API SolverOptions* solver_options_create(int solver_id);
FC3D_API SolverOptions* solver_options_create_by_name(const char* solver_name);
FC3D_API void solver_options_delete(SolverOptions* options);
FC3D_API SolverOptions* solver_options_copy(SolverOptions* source);
FC3D_API void solver_options_update_internal(SolverOptions* parent, size_t internal_solver_number,
                                             int solver_id);
FC3D_API const char* solver_options_id_to_name(int id);
FC3D_API int solver_options_name_to_id(const char* name);
FC3D_API SolverOptions* solver_options_get_internal_solver(SolverOptions* options, size_t n);

/* ------------------------------------------------------------------------ */
/* Matrices (host) — NumericsMatrix.c, SparseBlockMatrix.c                  */
/* ------------------------------------------------------------------------ */
/* NumericsMatrix.c NM_create_from_data: wraps data (double* for NM_DENSE,
 * SBM* for NM_SPARSE_BLOCK, NumericsSparseMatrix* for NM_SPARSE) */
FC3D_API NumericsMatrix* NM_create_from_data(int storageType, int size0, int size1, void* data);
/* NumericsMatrix.c NM_free: releases the matrix and its storage */
FC3D_API NumericsMatrix* NM_free(NumericsMatrix* m);
/* NumericsSparseMatrix.c:555-583 and CSparseMatrix.c:651-790: an NM_SPARSE
 * matrix as NM_new_from_file reads it (origin, then a cs_print listing) */
FC3D_API NumericsSparseMatrix* NSM_new_from_file(FILE* file);
FC3D_API int NSM_write_in_file(const NumericsSparseMatrix* A, FILE* file);
FC3D_API NumericsSparseMatrix* NSM_free(NumericsSparseMatrix* A);
/* NumericsMatrix.c NM_csc: the CSC form of an NM_SPARSE matrix (compressed
 * from the triplet on first use, as NM_csc does with cs_compress) */
FC3D_API CSparseMatrix* NM_csc(NumericsMatrix* A);
/* SparseBlockMatrix.c SBM_new / SBM_free(SBM_FREE_ALL) */
FC3D_API SparseBlockStructuredMatrix* SBM_new(void);
FC3D_API void SBM_free_all(SparseBlockStructuredMatrix* m);
/* Builds an SBM of 3x3 blocks from CSR arrays (copies the data); no
 * reference counterpart under this name — the COO->CSR route of
 * SparseBlockMatrix.c:2657-2703 (SBCM_to_SBM) ends in the same layout. */
FC3D_API SparseBlockStructuredMatrix* SBM_new_3x3_from_csr(size_t nrows, const int64_t* row_ptr,
                                                          const int64_t* col_idx,
                                                          const double* values);
/* SparseBlockMatrix.c:1941-1965 */
FC3D_API size_t* SBM_compute_diagonal_block_indices(SparseBlockStructuredMatrix* M);
FC3D_API size_t SBM_diagonal_block_index(SparseBlockStructuredMatrix* M, size_t row);
/* Dense column-major 3n x 3n -> SBM of its non-zero 3x3 blocks (diagonal
 * blocks always kept); the dense input branch of fc3d_local_problem_fill_M
 * (fc3d_local_problem_tools.c:65-71) reads the same blocks. */
FC3D_API SparseBlockStructuredMatrix* SBM_from_dense_3x3(int n, const double* dense);
/* SparseBlockMatrix.c:360-419 — y += A x on the GPU */
FC3D_API int SBM_gemv_3x3(size_t sizeX, size_t sizeY, const SparseBlockStructuredMatrix* A,
                          double* x, double* y);
/* NumericsMatrix.c:280-312 — y += A x on the GPU (dense inputs converted) */
FC3D_API int NM_prod_mv_3x3(int sizeX, int sizeY, NumericsMatrix* A, double* x, double* y);

/* ------------------------------------------------------------------------ */
/* Problems — FrictionContactProblem.c:67-200                               */
/* ------------------------------------------------------------------------ */
FC3D_API FrictionContactProblem* frictionContactProblem_new(void);
FC3D_API FrictionContactProblem* frictionContactProblem_new_with_data(int dim, size_t nc,
                                                                       NumericsMatrix* M,
                                                                       double* q, double* mu);
FC3D_API void frictionContactProblem_free(FrictionContactProblem* problem);
FC3D_API int frictionContact_printInFile(FrictionContactProblem* problem, FILE* file);
FC3D_API int frictionContact_printInFilename(FrictionContactProblem* problem, char* filename);
FC3D_API FrictionContactProblem* frictionContact_newFromFile(FILE* file);
FC3D_API FrictionContactProblem* frictionContact_new_from_filename(const char* filename);

/* ------------------------------------------------------------------------ */
/* Driver — fc3d_driver.c:51-145, NonSmoothDrivers.h:90                      */
/* ------------------------------------------------------------------------ */
FC3D_API int fc3d_checkTrivialCase(FrictionContactProblem* problem, double* velocity,
                                   double* reaction, SolverOptions* options);
FC3D_API int fc3d_driver(FrictionContactProblem* problem, double* reaction, double* velocity,
                         SolverOptions* options);
FC3D_API SolverOptions* fc3d_solver_options_create(solver_id_t solver_id);

/* fc3d_compute_error.c:70-122 — computed on the GPU; returns 1 if error > tol,
 * 0 otherwise, negative on failure. */
FC3D_API int fc3d_compute_error(FrictionContactProblem* problem, double* z, double* w,
                                double tolerance, SolverOptions* options, double norm,
                                double* error);

/* The FC3D_NSGS_HIP solver itself (registered as SICONOS_FRICTION_3D_NSGS_HIP);
 * fc3d_nsgs.c:560-917 is the behaviour it keeps, in colour order. */
FC3D_API void fc3d_nsgs_hip(FrictionContactProblem* problem, double* reaction, double* velocity,
                            int* info, SolverOptions* options);
FC3D_API void fc3d_nsgs_hip_set_default(SolverOptions* options);
/* Device problems FC3D_NSGS_HIP and fc3d_compute_error keep resident between
 * calls (dparam[SICONOS_FRICTION_3D_NSGS_HIP_RESIDENCY]; no reference
 * counterpart: the reference keeps no device copy).
 * clear drops those of M (NULL: all) that no call is using; NM_free drops
 * those of the matrix it frees.  stats: calls that reused a resident problem,
 * device problems built, and reuses that re-uploaded W's values. */
FC3D_API void fc3d_hip_resident_clear(const NumericsMatrix* M);
FC3D_API void fc3d_hip_resident_stats(int64_t* hits, int64_t* creates, int64_t* value_uploads);

/* ------------------------------------------------------------------------ */
/* Seeded problem generators (no reference counterpart: the reference gets W
 * from its mechanics kernel, LinearOSNS.cpp:522-700).  Outputs are malloc'ed
 * CSR arrays over 3x3 blocks (column-major values), owned by the caller and
 * released with fc3d_generated_free.                                       */
/* ------------------------------------------------------------------------ */
typedef struct {
  int nc;            /* contacts                                          */
  int nbodies;       /* bodies                                            */
  int64_t nnzb;      /* stored 3x3 blocks, diagonal included              */
  int64_t* row_ptr;  /* nc + 1                                            */
  int64_t* col_idx;  /* nnzb, ascending within each row                   */
  double* blocks;    /* nnzb * 9, column-major per block                  */
  double* q;         /* 3 nc                                              */
  double* mu;        /* nc                                                */
  int* body_a;       /* nc: first body of each contact                    */
  int* body_b;       /* nc: second body                                   */
  double* H;         /* nc * 36: per contact the 3x6 blocks for body_a and
                        body_b, column-major (generators only)             */
  double* minv;      /* nbodies * 6: diagonal of M_k^-1 (generators only)  */
  double* vfree;     /* nbodies * 6 (generators only)                     */
} fc3d_generated;

/* Spheres r=1, m=1, I=0.4; random distinct body pairs; random unit normals;
 * v_free ~ N(0,1) per DOF; mu ~ U[mu_lo, mu_hi]. */
FC3D_API int fc3d_generate_random_pairs(int nbodies, int nc, double mu_lo, double mu_hi,
                                        uint64_t seed, fc3d_generated* out);
/* Spheres on a jittered lattice pile; contacts between lattice neighbours,
 * normal along the line of centres; mu ~ U[mu_lo, mu_hi]; v_free ~ N(0,1)
 * scaled by 0.1 on a 10% subset of bodies. */
FC3D_API int fc3d_generate_granular_pile(int nbodies, int nc, double mu_lo, double mu_hi,
                                         uint64_t seed, fc3d_generated* out);
FC3D_API void fc3d_generated_free(fc3d_generated* g);
/* Wraps a generated problem as a FrictionContactProblem with SBM storage
 * (copies the data). */
FC3D_API FrictionContactProblem* fc3d_generated_to_problem(const fc3d_generated* g);

/* W = H M^-1 H^T and q = H v_free assembled on the GPU: the interaction block
 * of contacts c and d is the sum, over the bodies k they share, of
 * H_{c,k} M_k^-1 H_{d,k}^T (LinearOSNS.cpp:522-700, numerics
 * GlobalFrictionContactProblem.c:399-446 for the same product on SBM
 * matrices), stored as block CSR like BlockCSRMatrix.cpp:234-263 builds it.
 * Contact c touches bodies body_a[c] and body_b[c] (-1: a fixed body with no
 * degrees of freedom) through the 3x6 column-major blocks H + 36 c and
 * H + 36 c + 18.  Minv holds per body a 6x6 column-major block, or its
 * diagonal (6 values) when minv_diagonal != 0.  On return `out` holds nc,
 * nbodies, nnzb, row_ptr, col_idx (ascending, diagonal present), blocks and
 * q; the other pointers are NULL.  Release with fc3d_generated_free.
 * device_ms (may be NULL) receives the device time of the numeric kernels. */
FC3D_API int fc3d_hip_assemble(int nbodies, int nc, const int* body_a, const int* body_b,
                               const double* H, const double* Minv, int minv_diagonal,
                               const double* vfree, int device, fc3d_generated* out,
                               double* device_ms);

/* ------------------------------------------------------------------------ */
/* Colouring (no reference counterpart; replaces PETSc MatColoring GREEDY of
 * tools/parallel/graph_tools.c:73-83 with a deterministic definition):
 * contacts are coloured in index order, each taking the smallest colour not
 * used by a lower-index neighbour in the block graph of W.  Computed on the
 * GPU; returns the number of colours or a negative error code.             */
/* ------------------------------------------------------------------------ */
FC3D_API int fc3d_colour_contacts(int nc, const int64_t* row_ptr, const int64_t* col_idx,
                                  int* colour);
/* Validity of a colouring: 1 if no block W_ij (i != j) joins two contacts of
 * one colour. */
FC3D_API int fc3d_colouring_is_valid(int nc, const int64_t* row_ptr, const int64_t* col_idx,
                                     const int* colour);

#ifdef __cplusplus
}
#endif
#endif /* FC3D_NUMERICS_H */
