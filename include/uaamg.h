/*
 * uaamg.h -- C-ABI of the MI355X (gfx950) UAAMG-preconditioned CG pressure
 * solve.  Plain pointers and sizes only; every entry point returns a status
 * code (UAAMG_OK == 0).  Thread safety: one context per host thread.
 *
 * Layout contract (reference include/core/mesh.h:32-35, mesh_3d.h:32-38):
 * cell arrays are ghost-padded, i fastest,
 *     idx = i + j*(nx+2) [+ k*(nx+2)*(ny+2)],   interior i,j,k = 1..n,
 * spacing h = L/n (mesh.h:21-22).  For 2D pass dim = 2 and n[2] = 1.
 *
 * Which reference interface each entry point replaces is stated above it
 * (paths relative to the reference root).
 */
#ifndef UAAMG_H
#define UAAMG_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
enum {
    UAAMG_OK = 0,
    UAAMG_ERR_ARG = 1,         /* bad argument (null pointer, bad dims, no grid set) */
    UAAMG_ERR_HIP = 2,         /* HIP runtime error (message in uaamg_last_error) */
    UAAMG_ERR_OOM = 3,         /* device allocation failed */
    UAAMG_ERR_NODEVICE = 4,    /* no usable gfx950 device */
    UAAMG_ERR_ADMISSION = 5,   /* request exceeds the context's device-memory budget */
    UAAMG_NOT_CONVERGED = 10,  /* max_iter reached (reference: silent loop exit) */
    UAAMG_BREAKDOWN = 11,      /* p.Ap < 1e-15 (reference pcg.cpp:100: silent return) */
    UAAMG_NONFINITE = 12       /* NaN/Inf in a CG scalar */
};

/* ---- precision of the device path ---- */
enum {
    UAAMG_FP64 = 0,            /* everything FP64: parity mode */
    UAAMG_FP32 = 1,            /* plain form: FP32 vectors and levels, FP64 reductions/scalars */
    UAAMG_MIXED_BF16 = 2,      /* FP32 CG; 3D coarse levels stored BF16 (FP32 arithmetic); 2D levels stay FP32 */
    UAAMG_MIXED_FP16 = 3,      /* reserved, not built: 1/h^2 couplings exceed the FP16 range at 256^3;
                                  uaamg_create / uaamg_set_opts reject it with UAAMG_ERR_ARG */
    UAAMG_MIXED_BF16R = 4,     /* FP32 CG and level vectors; 3D coarse-level stencil records stored BF16 */
    UAAMG_MIXED_F64 = 5        /* FP64 CG vectors, scalars and level-0 smoothing; coarse levels FP32
                                  (CudaPCG3D::solve_mixed, cuda_pcg_3d_opt.cu:319-330) */
};

enum { UAAMG_WALL_NEUMANN = 0, UAAMG_WALL_DIRICHLET = 1 };
enum { UAAMG_TOL_ABS = 0, UAAMG_TOL_REL = 1 };

typedef struct {
    int dim;                   /* 2 or 3 */
    int n[3];                  /* interior cells per axis; n[2] = 1 for 2D */
    double h[3];               /* spacing per axis; h[2] ignored for 2D */
    const uint8_t* solid;      /* padded mask, 1 = solid (host pointer, or device pointer for *_device calls) */
    int wall[6];               /* -x,+x,-y,+y,-z,+z; reference walls are all Neumann */
} uaamg_grid;

typedef struct {
    int precision;             /* UAAMG_FP64 | UAAMG_FP32 | UAAMG_MIXED_* */
    int nu_pre, nu_post;       /* RBGS sweeps before/after the coarse correction (reference 1,1) */
    int coarsest_side;         /* hierarchy stops once any side <= this (reference 4) */
    int coarsest_sweeps;       /* symmetric sweep pairs on the coarsest level (reference 10) */
    int tol_mode;              /* UAAMG_TOL_ABS (pcg.cpp:112) | UAAMG_TOL_REL (cuda_pcg_3d_opt.cu:410) */
    int max_iter;              /* default for uaamg_solve* when max_iter < 0 is passed */
    double tol;                /* default tolerance when tol < 0 is passed */
    int project_mean;          /* -1 auto (on iff all walls Neumann), 0 off, 1 on */
    int fused;                 /* 1: fused LDS V-cycle kernels (needs nu_pre == nu_post == 1); 0: per-pass kernels */
    int check_every;           /* CG iterations queued between host convergence checks (>=1) */
    int use_graph;             /* 1: replay queued iterations from a captured hipGraph */
    int timing;                /* bitmask of UAAMG_K_* kernel classes timed with HIP events (0 off, 0xff all) */
    int flexible;              /* 0: beta = (r,z)/rsold (pcg_3d.cpp:115); 1: Polak-Ribiere (flexible CG) */
    /* Kernel forms and launch shapes.  Every form computes the same
     * arithmetic; the defaults (uaamg_default_opts) are the fastest measured
     * at 256^3.  Changing a structural form (UAAMG_FORM_ROWS, _ROWS_COARSE,
     * _TAIL, _PMV) or rows_kc / march_* through uaamg_set_opts drops the grid
     * (set it again). */
    int forms;                 /* bitmask of UAAMG_FORM_* allowed (default UAAMG_FORM_DEFAULT) */
    int xfork_level;           /* side-stream x update forked after this level's down kernel
                                  (>= the tail's first level: right before the tail); default 2 */
    int xblocks, xthreads;     /* x update launch: at most xblocks workgroups of xthreads (default 256 x 256) */
    int rows_kc;               /* planes per level-0 row-kernel tile, multiple of 4 (0: auto) */
    int march_kc_min;          /* fewest planes per march-kernel chunk, even (default 2) */
    int march_target;          /* workgroups a march-kernel launch aims at (default 1024) */
    int trace;                 /* 1: relative residual history of each solve to stderr */
    int preconditioner;        /* UAAMG_PRECOND_UAAMG (key "pcg_uaamg", factory.cpp:23),
                                  UAAMG_PRECOND_NONE (key "cg": IdentityPreconditioner, z = r,
                                  factory.cpp:19, factory_3d.cpp:19), UAAMG_PRECOND_GMG (keys "pcg"
                                  and "pcg_gmg": GMGPreconditioner, factory.cpp:20-21) or
                                  UAAMG_PRECOND_AMG (key "pcg_amg": AMGPreconditioner,
                                  factory.cpp:22).  GMG and AMG exist for 2D grids only (the 3D
                                  factory has no such keys, factory_3d.cpp:21-23: a 3D grid is
                                  UAAMG_ERR_ARG) and run V(2,2) whatever nu_pre / nu_post say
                                  (gmg_preconditioner.cpp:163-174); changing it drops the grid */
    int hierarchy_mode;        /* UAAMG_HIERARCHY_REFERENCE (default): a coarse cell is solid when
                                  at least 2 of its 4 (2D) or 4 of its 8 (3D) children are
                                  (restrictSolid, uaamg_preconditioner.cpp:43-53, _3d.cpp:34-45;
                                  gmg/amg_preconditioner.cpp:29-39 / :141-155), which drops fluid
                                  from the coarse grids and leaves one-sided couplings into
                                  coarse-solid cells.  UAAMG_HIERARCHY_KEEP_FLUID: solid only when
                                  every child is; the Galerkin coarse operators are then exactly
                                  the aggregation operators R^T A P.  No counterpart in the
                                  reference: results then differ from the reference's (fewer
                                  iterations where solids are present).  Changing it drops the grid */
} uaamg_opts;

enum { UAAMG_PRECOND_UAAMG = 0, UAAMG_PRECOND_NONE = 1, UAAMG_PRECOND_GMG = 2, UAAMG_PRECOND_AMG = 3 };
enum { UAAMG_HIERARCHY_REFERENCE = 0, UAAMG_HIERARCHY_KEEP_FLUID = 1 };

/* kernel forms (uaamg_opts.forms, uaamg_stats.forms_used) */
enum {
    UAAMG_FORM_ROWS = 1 << 0,        /* level-0 row kernels (3D, FP32 vectors, nx <= 256) */
    UAAMG_FORM_ROWS_COARSE = 1 << 1, /* record row kernels on levels with 64 < nx <= 128 */
    UAAMG_FORM_TAIL = 1 << 2,        /* small coarse levels in one LDS-resident workgroup */
    UAAMG_FORM_PMV = 1 << 3,         /* p update fused into the operator pass; with the fused
                                        V-cycle also r update in the level-0 down kernel and
                                        x update on a second stream */
    UAAMG_FORM_XMASK = 1 << 4,       /* x update reads the mask (default: it does not need to) */
    UAAMG_FORM_NT_X = 1 << 5,        /* nontemporal loads/stores in the x update */
    UAAMG_FORM_NT_P = 1 << 6,        /* nontemporal loads of z, p_old in the operator pass */
    UAAMG_FORM_NT_D = 1 << 7,        /* nontemporal loads of r_old, Ap in the level-0 down kernel */
    UAAMG_FORM_NT_U = 1 << 8,        /* nontemporal loads of r in the level-0 up kernel */
    UAAMG_FORM_SPLIT = 1 << 9,       /* (report only) the split CG update ran */
    UAAMG_FORM_DEFAULT = UAAMG_FORM_ROWS | UAAMG_FORM_ROWS_COARSE | UAAMG_FORM_TAIL | UAAMG_FORM_PMV |
                         UAAMG_FORM_NT_X | UAAMG_FORM_NT_P | UAAMG_FORM_NT_D | UAAMG_FORM_NT_U
};

typedef struct {
    int iters;                 /* CG loop iterations (matvecs) performed */
    int status;                /* UAAMG_OK | UAAMG_NOT_CONVERGED | UAAMG_BREAKDOWN | UAAMG_NONFINITE */
    int nlevels;
    double rnorm;              /* sqrt(r.r) at exit */
    double r0norm;             /* sqrt(r.r) of the initial (projected, negated) residual */
    double ms_total;           /* wall time of the call (host clock) */
    double ms_setup;           /* reserved */
    double ms_kernel[8];       /* device time per kernel class timed (opts.timing bit set), see UAAMG_K_* */
    int    n_kernel[8];        /* launches per timed class */
    double alg_bytes;          /* canonical algorithmic HBM bytes of this solve */
    int    kernel_set;         /* level-0 CG kernels used: 0 separate passes; 1 p update fused into the
                                  operator; 2 in addition r update in the level-0 down kernel and the
                                  x update on a second stream, beside the coarse levels */
    int    forms_used;         /* UAAMG_FORM_* bits of the kernel forms that ran */
} uaamg_stats;

/* kernel classes for uaamg_stats.ms_kernel */
enum { UAAMG_K_MATVEC = 0, UAAMG_K_UPDATE = 1, UAAMG_K_DOWN0 = 2, UAAMG_K_UP0 = 3,
       UAAMG_K_PUPDATE = 4, UAAMG_K_COARSE = 5, UAAMG_K_INIT = 6, UAAMG_K_OTHER = 7 };

typedef struct uaamg_ctx uaamg_ctx;

/* Defaults reproduce the reference: FP64, V(1,1), coarsest side 4, 10 coarsest
 * pairs, absolute tolerance, mean projection auto. */
int uaamg_default_opts(uaamg_opts* o);

/* Replaces constructing PCG(UAAMGPreconditioner) / CudaPCG3D
 * (src/solver/factory.cpp:23, include/solver/cuda/cuda_pcg_3d.h:7-49):
 * one context owns the device hierarchy and CG buffers. */
int uaamg_create(uaamg_ctx** out, const uaamg_opts* o);
void uaamg_destroy(uaamg_ctx* ctx);
const char* uaamg_last_error(const uaamg_ctx* ctx);
int uaamg_set_opts(uaamg_ctx* ctx, const uaamg_opts* o);

/* Uploads the mask and builds the hierarchy on device: coarse masks,
 * Galerkin couplings (CudaUAAMGPreconditioner3DT::setupLevels,
 * cuda_uaamg_preconditioner_3d.cu:324-376).  Re-uploads on every call, so a
 * mask change without a resize is honoured (the reference wrapper's
 * stale-mask case, cuda_pcg_solver.cu:45-49). g->solid is a host pointer. */
int uaamg_set_grid(uaamg_ctx* ctx, const uaamg_grid* g);
/* Same, g->solid is a device pointer. */
int uaamg_set_grid_device(uaamg_ctx* ctx, const uaamg_grid* g);

/* Variable Laplacian of the current 2D grid: replaces Grid::lap_diag,
 * lap_off_x, lap_off_y (include/core/grid.h:21-23, allocated by
 * Grid::init_variable_lap, src/core/grid.cpp:12-17, filled for instance by
 * SmoothCylinder::applyToLaplacian, src/geometry/cylinder_model.cpp:50-71),
 * which PCG::solve's matvec reads when present (src/solver/pcg.cpp:36-46):
 * Av = lap_diag v + lap_off_x (vL + vR) + lap_off_y (vB + vT) on fluid
 * cells, a wall or solid neighbour taking the cell's own value.  The V-cycle
 * keeps the uniform coefficients, as the reference's preconditioner does.
 * Host FP64 padded arrays, all three or all NULL (back to the uniform
 * operator); 2D grids only (UAAMG_ERR_ARG otherwise).  Kept until the next
 * uaamg_set_grid, which drops them. */
int uaamg_set_variable_lap(uaamg_ctx* ctx, const double* lap_diag, const double* lap_off_x,
                           const double* lap_off_y);

/* Replaces Solver::solve(Grid&, rhs, max_iter, tol) for key "pcg_uaamg"
 * (include/solver/solver.h:29-30, src/solver/pcg.cpp:18-126,
 * pcg_3d.cpp:18-125): solves lap(p) = rhs; p[fluid interior] += x where x is
 * the PCG iterate from x0 = 0 (pcg.cpp:107).  rhs_host, p_host: padded FP64
 * host arrays.  max_iter < 0 / tol < 0 take the context defaults. */
int uaamg_solve(uaamg_ctx* ctx, const double* rhs_host, double* p_host,
                int max_iter, double tol, uaamg_stats* st);

/* Replaces CudaPCG3D::solve / solve_optimized (cuda_pcg_3d.h:42-45) on device
 * pointers: rhs_dev, p_dev are padded FP64 device arrays.  warm_start = 0:
 * reference semantics (p += x, x from 0).  warm_start = 1: p_dev holds x0,
 * the solve starts from r = b - A x0 and p_dev receives the solution. */
int uaamg_solve_device(uaamg_ctx* ctx, const double* rhs_dev, double* p_dev,
                       int warm_start, int max_iter, double tol, uaamg_stats* st);

/* One V-cycle z = M^-1 r with a fresh hierarchy on the current grid
 * (Preconditioner::apply, include/solver/preconditioner/preconditioner.h:28-29;
 * uaamg_preconditioner.cpp:177-199). Host FP64 padded arrays; interior of z written. */
int uaamg_apply_precond(uaamg_ctx* ctx, const double* r_host, double* z_host);

/* The PCG operator Av (pcg.cpp:35-61, pcg_3d.cpp:38-55) and p.Av over fluid
 * cells, as the HIP matvec+dot kernel computes them. Host FP64 arrays. */
int uaamg_matvec(uaamg_ctx* ctx, const double* v_host, double* Av_host, double* vAv);

/* Level count and per-level dims (n per axis), as built by set_grid. */
int uaamg_hierarchy(const uaamg_ctx* ctx, int* nlevels, int* dims, int max_levels);

/* Download one level's mask and stencil channels (FP64 copies) for checking
 * against restrictSolid/setupFineCoeffs/galerkinCoarsen
 * (uaamg_preconditioner_3d.cpp:34-116).  Any pointer may be NULL. */
int uaamg_level_data(uaamg_ctx* ctx, int level, uint8_t* solid, double* diag,
                     double* cx, double* cy, double* cz, double* excess);

/* Standalone Ax=b solve with the semantics of the reference's amgpcg
 * (test/solver/test_amgpcg.cpp:29-70 with test/test_common.h): N x N
 * node-centred Dirichlet Poisson, h = 1/(N+1), row-major b[i*(N+2)+j];
 * GMG V-cycle with 2+2 RBGS sweeps, 20 at the coarsest; absolute tolerance.
 * x receives the (N+2)^2 solution; *iters as the reference's out_iters.
 * precision: UAAMG_FP64 or UAAMG_FP32. */
int uaamg_amgpcg(int N, const double* b, double* x, int max_iter, double tol,
                 int* iters, int precision);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code that must not call hipMalloc: tight serving loops, harnesses
 * that forbid allocation): uaamg_amgpcg on device arrays, with every
 * intermediate buffer in one fixed 32 MiB device array that the library holds
 * on each GPU it runs on, from its first call there until the process ends.
 * b_dev, x_dev: FP64 device arrays of (N+2)^2 values in uaamg_amgpcg's layout
 * (x_dev is written whole); stream: a hipStream_t, or NULL for the default
 * stream; the call returns when x_dev is written.  Sizes whose buffers exceed
 * the array (N > 737 in FP64, N > 1043 in FP32) return UAAMG_ERR_ADMISSION.
 * Allocates nothing: no device or pinned host memory, no stream or event.
 * Every call on one GPU uses that GPU's one array and clears it first: calls
 * that overlap in time on one GPU (from other host threads or streams) race,
 * so run them one after another. */
int uaamg_amgpcg_noalloc(int N, const double* b_dev, double* x_dev, int max_iter, double tol,
                         int* iters, int precision, void* stream);

/* Pressure projection around the solve: replaces PressureProjection::project
 * (src/pressure/pressure.cpp:5-32) and PressureProjection3D::project
 * (src/pressure/pressure_3d.cpp:5-39) with the pcg_uaamg solver.  On the grid
 * set by uaamg_set_grid: rhs = div(u)/dt on non-solid interior cells
 * (fvc::divergence, operators.h:18-21, operators_3d.h:17-21), p = 0, the
 * solve (as uaamg_solve_device, cold start), then u -= dt (p+ - p-)/dx on
 * interior faces between two non-solid cells (and v, w).  u, v, w are the
 * reference's MAC arrays (mesh.h:27-35, mesh_3d.h:29-37): u (nx+1)(ny+2)[(nz+2)],
 * v (nx+2)(ny+1)[(nz+2)], w (nx+2)(ny+2)(nz+1) (ignored in 2D); p is the
 * padded cell array.  The correction is applied whatever the solve status, as
 * the reference does.  _device: device pointers; otherwise host arrays
 * (copied in and out). */
int uaamg_project_device(uaamg_ctx* ctx, double* u, double* v, double* w, double* p, double dt,
                         int max_iter, double tol, uaamg_stats* st);
int uaamg_project(uaamg_ctx* ctx, double* u, double* v, double* w, double* p, double dt,
                  int max_iter, double tol, uaamg_stats* st);

/* Device memory (bytes) a context would hold for grid g under opts o; 0 when
 * uaamg_set_grid would reject the pair (invalid grid or options). */
size_t uaamg_bytes_required(const uaamg_grid* g, const uaamg_opts* o);

/* ---- batches: many independent solves in one set of launches ----
 * A batch holds grids of any size, mask and walls (2D and 3D).  Each member
 * is solved with the semantics of uaamg_solve_device (one Solver::solve per
 * member, include/solver/solver.h:29-30 / pcg_3d.cpp:18-125); the batch runs
 * every member's PCG iteration as one fixed sequence of launches over a
 * (grid, level, tile) work-list, captured as a hipGraph when opts.use_graph,
 * with launch shapes picked by a built-in autotuner on the first solve.
 * precision: UAAMG_FP32 or UAAMG_MIXED_BF16; V(1,1) (nu_pre = nu_post = 1),
 * fused = 1.  One host thread per batch. */
typedef struct uaamg_batch uaamg_batch;
int uaamg_batch_create(uaamg_batch** out, const uaamg_opts* o);
void uaamg_batch_destroy(uaamg_batch* b);
const char* uaamg_batch_last_error(const uaamg_batch* b);
/* Add a member and build its hierarchy on device (as uaamg_set_grid);
 * *id = member index (order of addition). g->solid is a host pointer. */
int uaamg_batch_add(uaamg_batch* b, const uaamg_grid* g, int* id);
/* Same, g->solid is a device pointer. */
int uaamg_batch_add_device(uaamg_batch* b, const uaamg_grid* g, int* id);
/* Remove every member (device memory is released). */
int uaamg_batch_clear(uaamg_batch* b);
int uaamg_batch_size(const uaamg_batch* b);
size_t uaamg_batch_bytes(const uaamg_batch* b);
/* Launch shapes the autotuner chose for the batch's 2D or 3D members (set by
 * their first solve): march work-list density (workgroups aimed at per level
 * and grid), fewest planes per march chunk, planes per level-0 row tile (0:
 * the engine's choice).  No counterpart in the reference (it has no batched
 * form).  UAAMG_ERR_ARG when the batch has no member of that dimension. */
int uaamg_batch_tuning(const uaamg_batch* b, int dim, int* march_target, int* march_kc_min, int* rows_kc);
/* Solve every member: rhs_dev[i], p_dev[i] are padded FP64 device arrays of
 * member i; warm_start[i] as in uaamg_solve_device (NULL: all cold).  st
 * (NULL or an array of uaamg_batch_size entries) receives per-member stats;
 * ms_total is the wall time of the whole call.  Returns UAAMG_OK or the
 * first member's non-OK status. */
int uaamg_batch_solve(uaamg_batch* b, const double* const* rhs_dev, double* const* p_dev,
                      const int* warm_start, int max_iter, double tol, uaamg_stats* st);

/* ---- serving: pressure solves of many concurrent simulations ----
 * Each request is one Solver::solve of one simulation frame
 * (include/solver/solver.h:29-30); the loop keeps each simulation's hierarchy
 * resident while the HBM budget allows (least recently used idle simulations
 * are evicted), runs the active requests in batches and admits new requests
 * between chunks of iterations (continuous batching).  Simulations fall into
 * four size classes (2D / 3D, fewer / at least 2^18 padded cells), each
 * batched on a stream of its own, so small requests complete at the pace of
 * their own chunks, not of the largest grid's.  With warm_start = 1 the
 * request's p_dev holds the previous frame's pressure and receives the new
 * one.  precision: UAAMG_FP32 or UAAMG_MIXED_BF16. */
typedef struct uaamg_server uaamg_server;
int uaamg_server_create(uaamg_server** out, const uaamg_opts* o, size_t hbm_budget, int max_active);
void uaamg_server_destroy(uaamg_server* s);
const char* uaamg_server_last_error(const uaamg_server* s);
/* Register simulation `sim` with its grid (host mask, copied). */
int uaamg_server_register(uaamg_server* s, int sim, const uaamg_grid* g);
/* Queue one solve of `sim` (device arrays); *req = request id.  The server
 * reads rhs_dev and p_dev on its own stream from the step() that admits the
 * request: work that writes them on another stream must be complete before
 * submit (synchronize that stream).  p_dev holds the result once step()
 * reports the request done. */
int uaamg_server_submit(uaamg_server* s, int sim, const double* rhs_dev, double* p_dev, int warm_start,
                        int max_iter, double tol, int* req);
/* Complete the size classes whose chunk of iterations has finished, admit
 * queued requests into classes with no chunk running, launch the next chunk
 * of every such class; if no class had finished, wait until one has (and
 * refill it).  done[0..*n_done) = finished request ids.  Requests that finish
 * while done[] is full are returned first by the next step (their latency is
 * taken when they finish). */
int uaamg_server_step(uaamg_server* s, int* done, int max_done, int* n_done);
/* Stats of a finished request (ms_total = latency from submit to completion). */
int uaamg_server_request(const uaamg_server* s, int req, uaamg_stats* st);
int uaamg_server_status(const uaamg_server* s, int* queued, int* active, int* resident, size_t* bytes_used);

#ifdef __cplusplus
}
#endif
#endif
