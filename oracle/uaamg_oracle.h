/*
 * uaamg_oracle.h -- CPU restatement of the reference UAAMG-PCG pressure solve.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP path:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it.  The product (uaamg_amd/, include/uaamg.h) never calls it.
 *
 * Layout contract (include/core/mesh.h:32-35, mesh_3d.h:32-38): cell arrays
 * are ghost-padded, i fastest: idx = i + j*(nx+2) [+ k*(nx+2)*(ny+2)],
 * interior indices 1..n.  For dim == 2 pass nz = 1 and use k = 0.
 *
 * Every function follows the reference loop and summation order so that, in
 * FP64 and with the default options, results match the compiled reference
 * bit for bit (checked against oracle/_ref in tests/test_oracle.py).
 */
#ifndef UAAMG_ORACLE_H
#define UAAMG_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NEUMANN   0
#define ORC_DIRICHLET 1

typedef struct {
    int dim;                     /* 2 or 3 */
    int nx, ny, nz;              /* interior cells; nz = 1 when dim == 2 */
    double dx, dy, dz;           /* spacing, L/n as mesh.h:21-22 */
    const unsigned char* solid;  /* padded mask, 1 = solid */
    int wall[6];                 /* -x,+x,-y,+y,-z,+z : ORC_NEUMANN (reference) or ORC_DIRICHLET */
    /* Variable Laplacian of Grid (include/core/grid.h:21-23), 2D only: padded
     * per-cell diagonal and x/y off-diagonal coefficients read by the PCG
     * matvec (pcg.cpp:36-46).  All NULL: the uniform operator. */
    const double* lap_diag;
    const double* lap_off_x;
    const double* lap_off_y;
} orc_grid;

typedef struct {
    int nu_pre, nu_post;         /* RBGS sweeps per side of the V-cycle (reference 1,1) */
    int coarsest_side;           /* hierarchy stops once a side <= this (reference 4) */
    int coarsest_sweeps;         /* symmetric sweep pairs at the coarsest level (reference 10) */
    int tol_mode;                /* 0: sqrt(r.r) < tol (pcg.cpp:112); 1: r.r < tol^2 r0.r0 */
    int project_mean;            /* -1 auto (on iff every wall Neumann), 0 off, 1 on */
    int precond;                 /* ORC_PRECOND_UAAMG (key "pcg_uaamg"), ORC_PRECOND_IDENTITY
                                    (key "cg": IdentityPreconditioner, z = r; factory.cpp:19),
                                    ORC_PRECOND_GMG (keys "pcg", "pcg_gmg": GMGPreconditioner, 2D) or
                                    ORC_PRECOND_AMG (key "pcg_amg": AMGPreconditioner, 2D) */
    int hierarchy_mode;          /* 0: the reference's coarse-solid rule (at least 2 of 4 / 4 of 8
                                    children solid, restrictSolid); 1: coarse cell solid only when
                                    every child is (no fluid dropped from the coarse grids) */
} orc_opts;

#define ORC_PRECOND_UAAMG    0
#define ORC_PRECOND_IDENTITY 1
#define ORC_PRECOND_GMG      2
#define ORC_PRECOND_AMG      3

typedef struct {
    int iters;       /* loop iterations entered (matvecs); = CudaPCG3D::last_iters when converged */
    int status;      /* 0 converged, 1 max_iter reached, 2 breakdown (pAp < 1e-15) */
    int applies;     /* preconditioner applications */
    double rnorm;    /* sqrt(r.r) when the loop ended */
    double r0norm;   /* sqrt(r.r) of the initial residual */
} orc_result;

void orc_default_opts(orc_opts* o);

/* PCG(UAAMG) solve: p[fluid interior] += x, x the PCG iterate from 0
 * (pcg.cpp:18-126, pcg_3d.cpp:18-125). rhs/p are padded arrays. */
int orc_pcg(const orc_grid* g, const orc_opts* o, const double* rhs, double* p,
            int max_iter, double tol, orc_result* res);

/* One UAAMG V-cycle application z = M^-1 r, hierarchy set up from scratch
 * (uaamg_preconditioner.cpp:177-199, uaamg_preconditioner_3d.cpp:198-229).
 * Writes interior cells of z only. */
int orc_uaamg_apply(const orc_grid* g, const orc_opts* o, const double* r, double* z);

/* One application of the preconditioner o->precond names: UAAMG as above,
 * identity z = r, or one GMG / AMG V(2,2) cycle on a hierarchy set up from
 * scratch (GMGPreconditioner::apply, gmg_preconditioner.cpp:176-194;
 * AMGPreconditioner::apply, amg_preconditioner.cpp:190-209; 2D only). */
int orc_precond_apply(const orc_grid* g, const orc_opts* o, const double* r, double* z);

/* Hierarchy shape: number of levels and (nx,ny,nz) per level (buildHierarchy). */
int orc_hierarchy_dims(const orc_grid* g, const orc_opts* o, int* nlevels, int* dims, int max_levels);

/* Per-level mask and stencil channels after set-up (restrictSolid,
 * setupFineCoeffs, galerkinCoarsen). Any output pointer may be NULL.
 * cz is untouched for dim == 2. Arrays are the padded size of that level. */
int orc_level_data(const orc_grid* g, const orc_opts* o, int level,
                   unsigned char* solid, double* diag, double* cx, double* cy,
                   double* cz, double* excess);

/* PCG matvec with uniform constants (pcg.cpp:47-59, pcg_3d.cpp:38-55), or
 * with the grid's variable coefficients (pcg.cpp:36-46) when set, extended
 * with ghost = -v on Dirichlet walls. Writes interior cells. */
int orc_matvec(const orc_grid* g, const double* v, double* Av);

/* Standalone amgpcg (test/solver/test_amgpcg.cpp:6-70 with
 * test/test_common.h): node-centred Dirichlet Poisson, h = 1/(N+1),
 * row-major v[i*(N+2)+j]. b, x: (N+2)^2 arrays. */
int orc_amgpcg(int N, const double* b, double* x, int max_iter, double tol, int* iters);

/* Pressure projection (pressure.cpp:5-32, pressure_3d.cpp:5-39) on MAC
 * velocities: u (nx+1)(ny+2)[(nz+2)], v (nx+2)(ny+1)[(nz+2)],
 * w (nx+2)(ny+2)(nz+1) (mesh.h:27-35, mesh_3d.h:29-37; w unused in 2D),
 * p padded.  rhs = div(u)/dt, p = 0, orc_pcg, then u -= dt grad p on faces
 * between two non-solid cells.  The two halves are exported for tests. */
int orc_project(const orc_grid* g, const orc_opts* o, double* u, double* v, double* w, double* p, double dt,
                int max_iter, double tol, orc_result* res);
void orc_project_rhs(const orc_grid* g, const double* u, const double* v, const double* w, double dt, double* rhs);
void orc_project_correct(const orc_grid* g, double* u, double* v, double* w, const double* p, double dt);

#ifdef __cplusplus
}
#endif
#endif
