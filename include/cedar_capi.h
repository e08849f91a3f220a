/*
 * Handle-based C API of the HIP BoxMG solvers: one process, one GPU.
 *
 * The same names, handle encoding, error codes and argument meaning as the reference's C API
 * (include/cedar/capi.h:12-403, implemented in src/interface/object.cc and the files of
 * src/2d/interface/c and src/3d/interface/c), so that a program written against that header
 * links against libcedar_hip.so instead.  The reference binds every topology to an MPI
 * communicator; here a topology is one process's whole grid on the current GPU:
 *
 *   - cedar_topo_create2d/3d take the communicator as an opaque pointer that is ignored, and
 *     accept only nprocx = nprocy (= nprocz) = 1 with lnx[0] = ngx (etc.), else CEDAR_ERR_TOPO;
 *   - a vector's storage (cedar_vec_baseptr) is host memory in the reference's layout: Fortran
 *     order, one ghost layer, lengths from cedar_vec_len2d/3d (ghosts included).  cedar_matvec
 *     and cedar_solver_run copy their inputs to the GPU, run there, and copy the results back;
 *     callers that keep their data in HBM use the device-pointer API of cedar_hip.h instead;
 *   - cedar_solver_create runs the setup on the GPU with the settings of the matrix's config
 *     (keys and defaults of src/multilevel_settings.cc:15-61 and src/kernel_params.cc:72-89);
 *     the solver keeps its own copy of the operator taken at creation.
 *
 * Handles: an int whose top four bits are the kind (topo 1, solver 2, mat 3, vec 4, config 5)
 * and whose low 28 bits are a 1-based slot (include/cedar/interface/object.h:8-30).  A freed
 * handle is set to the kind's NULL value.  Handles are not thread-safe.
 */
#ifndef CEDAR_CAPI_H
#define CEDAR_CAPI_H

#ifdef __cplusplus
extern "C" {
#endif

#define CEDAR_SUCCESS     0
#define CEDAR_ERR_TOPO    1
#define CEDAR_ERR_KIND    2
#define CEDAR_ERR_HANDLE  3
#define CEDAR_ERR_MAT     4
#define CEDAR_ERR_VEC     5
#define CEDAR_ERR_DIM     6
#define CEDAR_ERR_STENCIL 7
#define CEDAR_ERR_SOLVER  8
#define CEDAR_ERR_CONFIG  9
#define CEDAR_ERR_FNAME   10
#define CEDAR_ERR_CONFIG_PATH 11
#define CEDAR_ERR_OTHER   12

typedef double cedar_real;       /* include/cedar/ctypes.h:8-9 */
typedef unsigned int cedar_len;

typedef int cedar_config;
typedef int cedar_topo;
typedef int cedar_mat;
typedef int cedar_vec;
typedef int cedar_solver;

#define CEDAR_TOPO_NULL ((cedar_topo) 0x10000000)
#define CEDAR_MAT_NULL ((cedar_mat) 0x30000000)
#define CEDAR_SOLVER_NULL ((cedar_solver) 0x20000000)
#define CEDAR_VEC_NULL ((cedar_vec) 0x40000000)
#define CEDAR_CONFIG_NULL ((cedar_config) 0x50000000)

/* stencil directions, vertex-based (include/cedar/2d/base_types.h, include/cedar/3d/base_types.h) */
typedef enum {
    BMG2_C = 0, BMG2_W = 1, BMG2_S = 2, BMG2_SW = 3, BMG2_NW = 4,
    BMG2_SE = 5, BMG2_N = 6, BMG2_NE = 7, BMG2_E = 8
} bmg2_dir;
typedef enum {
    BMG3_P = 0, BMG3_PW = 1, BMG3_PS = 2, BMG3_B = 3, BMG3_PSW = 4, BMG3_PNW = 5, BMG3_BW = 6,
    BMG3_BNW = 7, BMG3_BN = 8, BMG3_BNE = 9, BMG3_BE = 10, BMG3_BSE = 11, BMG3_BS = 12,
    BMG3_BSW = 13
} cdr3_dir;

typedef struct {
    unsigned int i;   /* 0-based interior index */
    unsigned int j;
    bmg2_dir dir;
} cedar_coord_2d;
typedef struct {
    unsigned int i;
    unsigned int j;
    unsigned int k;
    cdr3_dir dir;
} cedar_coord_3d;

typedef enum { CEDAR_STENCIL_FIVE_PT, CEDAR_STENCIL_NINE_PT } cedar_stencil_2d;
typedef enum { CEDAR_STENCIL_SEVEN_PT, CEDAR_STENCIL_XXVII_PT } cedar_stencil_3d;

/* Log levels from the config's "log" array ("status", "info", "error", "memory", "debug",
 * "timer"; default status and error), src/util/log.cc:53-95.  Replaces cedar_log_init
 * (include/cedar/capi.h:47). */
int cedar_log_init(cedar_config conf);

/* JSON config file (include/cedar/config.h:12-111).  cedar_config_create: CEDAR_ERR_FNAME when
 * the file cannot be read, CEDAR_ERR_CONFIG when it is not valid JSON (the reference throws).
 * cedar_config_get_int: value at a dotted path ("solver.max-iter"); numbers are truncated and
 * booleans read as 0/1 like nlohmann::json's get<int>; CEDAR_ERR_CONFIG_PATH when absent or not
 * a number.  include/cedar/capi.h:59-84. */
int cedar_config_create(const char *fname, cedar_config *newconfig);
int cedar_config_get_int(cedar_config conf, const char *confpath, int *val);
int cedar_config_free(cedar_config *conf);

/* include/cedar/capi.h:101-142; comm is ignored (see above) */
int cedar_topo_create2d(void *comm, unsigned int ngx, unsigned int ngy, unsigned int lnx[],
                        unsigned int lny[], int nprocx, int nprocy, cedar_topo *newtopo);
int cedar_topo_create3d(void *comm, unsigned int ngx, unsigned int ngy, unsigned int ngz,
                        unsigned int lnx[], unsigned int lny[], unsigned int lnz[], int nprocx,
                        int nprocy, int nprocz, cedar_topo *newtopo);
int cedar_topo_free(cedar_topo *topo);

/* Grid functions, zero-initialised (include/cedar/capi.h:154-233).  cedar_vec_free checks the
 * vector kind (the reference's checks the topology kind, src/2d/interface/c/vec.cc:104, and so
 * rejects every vector). */
int cedar_vec_create2d(cedar_topo topo, cedar_vec *vec);
int cedar_vec_create3d(cedar_topo topo, cedar_vec *vec);
int cedar_vec_baseptr(cedar_vec vec, cedar_real **base);
int cedar_vec_len2d(cedar_vec vec, cedar_len *ilen, cedar_len *jlen);
int cedar_vec_len3d(cedar_vec vec, cedar_len *ilen, cedar_len *jlen, cedar_len *klen);
int cedar_vec_getdim(cedar_vec vec, int *nd);
int cedar_vec_free(cedar_vec *vec);

/* Matrices (include/cedar/capi.h:272-367, src/2d/interface/c/mat.cc:60-144,
 * src/3d/interface/c/mat.cc).  cedar_mat_set2d takes vertex-based couplings: entry (i, j, dir)
 * is the coefficient of the dir neighbour in row (i, j); off-diagonal values are stored negated,
 * and N, E, NE, SE couplings are stored as the S, W, SW, NW couplings of the neighbour, as the
 * reference does.  A five-point matrix rejects SW, NW, NE and SE with CEDAR_ERR_STENCIL (the
 * reference rejects NW, NE, SE and writes SW out of bounds).  cedar_mat_set3d stores dir as
 * given (symmetric storage).  Unlike the reference, vals is not negated in place, and a
 * coordinate outside the grid (or a coupling landing outside the ghost-inclusive array) gives
 * CEDAR_ERR_DIM; entries before a rejected one stay set, as in the reference. */
int cedar_mat_create2d(cedar_config conf, cedar_topo topo, cedar_stencil_2d sten_type,
                       cedar_mat *mat);
int cedar_mat_create3d(cedar_config conf, cedar_topo topo, cedar_stencil_3d sten_type,
                       cedar_mat *mat);
int cedar_mat_set2d(cedar_mat mat, unsigned int nvals, cedar_coord_2d coords[], cedar_real vals[]);
int cedar_mat_set3d(cedar_mat mat, unsigned int nvals, cedar_coord_3d coords[], cedar_real vals[]);
/* y = mat * x on the GPU (BMG2/3_SymStd_UTILS_matvec order); ghosts of y unchanged */
int cedar_matvec(cedar_mat mat, cedar_vec x, cedar_vec y);
int cedar_mat_gettopo(cedar_mat mat, cedar_topo *topo);
/* 2D: "op-0-0.txt" in the working directory in the stencil operator<< layout
 * (src/2d/mpi/stencil_op.cc:6-59); 3D: "so-0.0.0" holding the number of directions, the number
 * of values and one value per line (src/3d/util/io.cc:6-25), here the ghost-inclusive
 * (nx+2)(ny+2)(nz+2) points of each direction plane in turn. */
int cedar_mat_dump(cedar_mat mat);
int cedar_mat_free(cedar_mat *mat);

/* Solver (include/cedar/capi.h:380-403, src/2d/interface/c/solver.cc:30-130).  Settings from
 * the matrix's config: solver.relaxation, solver.cycle.{type, nrelax-pre, nrelax-post},
 * solver.num-levels, solver.max-iter, solver.tol, solver.min_coarse, solver.cg-solver ("LU";
 * "cedar" and "redist" name MPI coarse solvers and run the same band Cholesky on one GPU), and for
 * 3D plane relaxation the "plane-config" subtree (default line-xy, max-iter 1).  GPU keys, not in
 * the reference: gpu.precision ("f64" default, or "mixed"), gpu.graph (true), gpu.fused-coarse
 * (true), gpu.autotune (false).  grid.periodic (Cedar's per_mask) is honoured for 2D problems;
 * for 3D a true entry gives CEDAR_ERR_CONFIG, as do unknown relaxation or cycle names.
 * cedar_solver_run solves from the x given (multilevel::solve(b, x),
 * include/cedar/multilevel.h:278-298) and logs "Iteration i relative l2 norm: ..." at the status
 * level; CEDAR_ERR_OTHER when the GPU reports an error. */
int cedar_solver_create(cedar_mat mat, cedar_solver *solver);
int cedar_solver_run(cedar_solver solver, cedar_vec x, cedar_vec b);
int cedar_solver_destroy(cedar_solver *solver);
int cedar_solver_free(cedar_solver *solver);   /* the name the reference implements */

#ifdef __cplusplus
}
#endif
#endif /* CEDAR_CAPI_H */
