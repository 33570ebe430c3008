/*
 * uaamg_oracle.c -- CPU restatement (FP64, single thread) of the reference
 * matrix-free UAAMG-preconditioned CG.  TEST INFRASTRUCTURE ONLY: the checker
 * for the HIP path and the timed CPU baseline; never linked into the product.
 *
 * Followed, function by function (all paths relative to the reference root):
 *   hierarchy      src/solver/preconditioner/uaamg_preconditioner.cpp:23-41
 *                  src/solver/preconditioner/uaamg_preconditioner_3d.cpp:11-32
 *   restrictSolid  uaamg_preconditioner.cpp:43-53,  _3d.cpp:34-45
 *   setupFine      uaamg_preconditioner.cpp:56-75,  _3d.cpp:50-76
 *   galerkin       uaamg_preconditioner.cpp:79-102, _3d.cpp:81-116
 *   smooth (RBGS)  uaamg_preconditioner.cpp:105-121, _3d.cpp:123-141
 *   restrictResid  uaamg_preconditioner.cpp:124-144, _3d.cpp:144-165
 *   prolongateAdd  uaamg_preconditioner.cpp:147-160, _3d.cpp:168-181
 *   vCycle         uaamg_preconditioner.cpp:162-175, _3d.cpp:183-196
 *   apply          uaamg_preconditioner.cpp:177-199, _3d.cpp:198-229
 *   PCG            src/solver/pcg.cpp:18-126, src/solver/pcg_3d.cpp:18-125
 *   relative tol   src/solver/cuda/cuda_pcg_3d_opt.cu:376-380,410
 *   amgpcg         test/solver/test_amgpcg.cpp:6-70, test/test_common.h:16-97
 *   identity       include/solver/preconditioner/identity_preconditioner.h:19
 *   GMG / AMG      src/solver/preconditioner/gmg_preconditioner.cpp:13-194,
 *                  src/solver/preconditioner/amg_preconditioner.cpp:13-209
 *   variable lap   src/solver/pcg.cpp:36-46 (Grid::lap_*, grid.h:21-23)
 *
 * Extensions beyond the reference, all inactive with default options and
 * Neumann walls (then every result is the reference's, bit for bit):
 *   - Dirichlet walls: ghost = -interior in the matvec; the smoother/Galerkin
 *     operator carries a per-cell diagonal excess 2/h^2 per Dirichlet face,
 *     coarsened as the sum over children (R^T A P of a diagonal term);
 *     mean projection off.
 *   - nu_pre/nu_post sweeps, coarsest side and coarsest sweep count.
 *   - hierarchy_mode 1: a coarse cell is solid only when every child is
 *     (restrict_solid; the reference's threshold is mode 0).
 *   - relative tolerance mode (the GPU solve_optimized rule).
 */
#include "uaamg_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 40

typedef struct {
    int nx, ny, nz;          /* nz = 1 for 2D */
    double dx, dy, dz;
    size_t N;
    double *p, *b, *diag, *cx, *cy, *cz, *ex;
    unsigned char* solid;
} Level;

typedef struct {
    int dim, nl;
    int wall[6];
    orc_opts o;
    Level L[MAXL];
} Hier;

void orc_default_opts(orc_opts* o) {
    o->nu_pre = 1; o->nu_post = 1;
    o->coarsest_side = 4;
    o->coarsest_sweeps = 10;
    o->tol_mode = 0;
    o->project_mean = -1;
    o->precond = ORC_PRECOND_UAAMG;
    o->hierarchy_mode = 0;
}

static inline size_t IX(const Level* L, int i, int j, int k) {
    size_t si = (size_t)L->nx + 2;
    return (size_t)i + (size_t)j * si + (size_t)k * si * ((size_t)L->ny + 2);
}

/* interior k range: 1..nz in 3D, the single plane k = 0 in 2D */
#define KLO(d, L) ((d) == 3 ? 1 : 0)
#define KHI(d, L) ((d) == 3 ? (L)->nz : 0)

static void free_hier(Hier* H) {
    for (int l = 0; l < H->nl; l++) {
        Level* L = &H->L[l];
        free(L->p); free(L->b); free(L->diag); free(L->cx); free(L->cy);
        free(L->cz); free(L->ex); free(L->solid);
    }
    H->nl = 0;
}

static int build_hier(Hier* H, const orc_grid* g, const orc_opts* o) {
    memset(H, 0, sizeof(*H));
    H->dim = g->dim;
    H->o = *o;
    for (int f = 0; f < 6; f++) H->wall[f] = g->wall[f];
    int nx = g->nx, ny = g->ny, nz = (g->dim == 3) ? g->nz : 1;
    double dx = g->dx, dy = g->dy, dz = g->dz;
    int S = o->coarsest_side;
    while (nx >= 2 && ny >= 2 && (g->dim == 2 || nz >= 2)) {
        if (H->nl >= MAXL) return -1;
        Level* L = &H->L[H->nl];
        L->nx = nx; L->ny = ny; L->nz = nz;
        L->dx = dx; L->dy = dy; L->dz = dz;
        L->N = (size_t)(nx + 2) * (ny + 2) * (g->dim == 3 ? (nz + 2) : 1);
        L->p = calloc(L->N, sizeof(double));
        L->b = calloc(L->N, sizeof(double));
        L->diag = calloc(L->N, sizeof(double));
        L->cx = calloc(L->N, sizeof(double));
        L->cy = calloc(L->N, sizeof(double));
        L->cz = calloc(L->N, sizeof(double));
        L->ex = calloc(L->N, sizeof(double));
        L->solid = calloc(L->N, 1);
        H->nl++;
        if (!L->p || !L->b || !L->diag || !L->cx || !L->cy || !L->cz || !L->ex || !L->solid) {
            free_hier(H);
            return -2;
        }
        if (nx <= S || ny <= S || (g->dim == 3 && nz <= S)) break;
        nx /= 2; ny /= 2; if (g->dim == 3) nz /= 2;
        dx *= 2.0; dy *= 2.0; dz *= 2.0;
    }
    return H->nl > 0 ? 0 : -3;
}

static void restrict_solid(int dim, const Level* F, Level* C, int mode) {
    int thr = mode ? (dim == 3 ? 8 : 4) : (dim == 3) ? 4 : 2;
    for (int ic = 1; ic <= C->nx; ic++)
        for (int jc = 1; jc <= C->ny; jc++)
            for (int kc = KLO(dim, C); kc <= KHI(dim, C); kc++) {
                int i_f = 2 * ic - 1, j_f = 2 * jc - 1, k_f = (dim == 3) ? 2 * kc - 1 : 0;
                int sc = 0;
                for (int di = 0; di < 2; di++)
                    for (int dj = 0; dj < 2; dj++)
                        for (int dk = 0; dk < (dim == 3 ? 2 : 1); dk++)
                            if (F->solid[IX(F, i_f + di, j_f + dj, k_f + dk)]) sc++;
                C->solid[IX(C, ic, jc, kc)] = (sc >= thr);
            }
}

static double wall_excess(const Hier* H, const Level* L, int i, int j, int k) {
    /* 2/h^2 per Dirichlet face touched by the cell (cell-centred p = 0 on the face) */
    double e = 0.0;
    double idx2 = 1.0 / (L->dx * L->dx), idy2 = 1.0 / (L->dy * L->dy), idz2 = 1.0 / (L->dz * L->dz);
    if (H->wall[0] == ORC_DIRICHLET && i == 1) e += 2.0 * idx2;
    if (H->wall[1] == ORC_DIRICHLET && i == L->nx) e += 2.0 * idx2;
    if (H->wall[2] == ORC_DIRICHLET && j == 1) e += 2.0 * idy2;
    if (H->wall[3] == ORC_DIRICHLET && j == L->ny) e += 2.0 * idy2;
    if (H->dim == 3) {
        if (H->wall[4] == ORC_DIRICHLET && k == 1) e += 2.0 * idz2;
        if (H->wall[5] == ORC_DIRICHLET && k == L->nz) e += 2.0 * idz2;
    }
    return e;
}

static double diag_sum(int dim, const Level* L, size_t id, size_t im, size_t jm, size_t km) {
    if (dim == 3)
        return L->cx[id] + L->cx[im] + L->cy[id] + L->cy[jm] + L->cz[id] + L->cz[km];
    return L->cx[id] + L->cx[im] + L->cy[id] + L->cy[jm];
}

static void setup_fine(const Hier* H, Level* L) {
    int dim = H->dim, nx = L->nx, ny = L->ny, nz = L->nz;
    double idx2 = 1.0 / (L->dx * L->dx), idy2 = 1.0 / (L->dy * L->dy), idz2 = 1.0 / (L->dz * L->dz);
    memset(L->cx, 0, L->N * sizeof(double));
    memset(L->cy, 0, L->N * sizeof(double));
    memset(L->cz, 0, L->N * sizeof(double));
    memset(L->diag, 0, L->N * sizeof(double));
    memset(L->ex, 0, L->N * sizeof(double));
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = KLO(dim, L); k <= KHI(dim, L); k++) {
                size_t id = IX(L, i, j, k);
                if (L->solid[id]) continue;
                if (i < nx && !L->solid[IX(L, i + 1, j, k)]) L->cx[id] = idx2;
                if (j < ny && !L->solid[IX(L, i, j + 1, k)]) L->cy[id] = idy2;
                if (dim == 3 && k < nz && !L->solid[IX(L, i, j, k + 1)]) L->cz[id] = idz2;
            }
    for (int i = 1; i <= nx; i++)
        for (int j = 1; j <= ny; j++)
            for (int k = KLO(dim, L); k <= KHI(dim, L); k++) {
                size_t id = IX(L, i, j, k);
                if (L->solid[id]) continue;
                double d = diag_sum(dim, L, id, IX(L, i - 1, j, k), IX(L, i, j - 1, k),
                                    dim == 3 ? IX(L, i, j, k - 1) : id);
                double e = wall_excess(H, L, i, j, k);
                L->ex[id] = e;
                if (e != 0.0) d += e;
                L->diag[id] = d;
            }
}

static void galerkin(int dim, const Level* F, Level* C) {
    memset(C->cx, 0, C->N * sizeof(double));
    memset(C->cy, 0, C->N * sizeof(double));
    memset(C->cz, 0, C->N * sizeof(double));
    memset(C->diag, 0, C->N * sizeof(double));
    memset(C->ex, 0, C->N * sizeof(double));
    for (int ic = 1; ic <= C->nx; ic++)
        for (int jc = 1; jc <= C->ny; jc++)
            for (int kc = KLO(dim, C); kc <= KHI(dim, C); kc++) {
                size_t cidx = IX(C, ic, jc, kc);
                if (C->solid[cidx]) continue;
                int i_f = 2 * ic - 1, j_f = 2 * jc - 1;
                double sx = 0, sy = 0, sz = 0, e = 0;
                if (dim == 3) {
                    int k_f = 2 * kc - 1;
                    for (int dj = 0; dj < 2; dj++) for (int dk = 0; dk < 2; dk++)
                        sx += F->cx[IX(F, i_f + 1, j_f + dj, k_f + dk)];
                    for (int di = 0; di < 2; di++) for (int dk = 0; dk < 2; dk++)
                        sy += F->cy[IX(F, i_f + di, j_f + 1, k_f + dk)];
                    for (int di = 0; di < 2; di++) for (int dj = 0; dj < 2; dj++)
                        sz += F->cz[IX(F, i_f + di, j_f + dj, k_f + 1)];
                    for (int di = 0; di < 2; di++) for (int dj = 0; dj < 2; dj++) for (int dk = 0; dk < 2; dk++)
                        e += F->ex[IX(F, i_f + di, j_f + dj, k_f + dk)];
                } else {
                    for (int dj = 0; dj < 2; dj++) sx += F->cx[IX(F, i_f + 1, j_f + dj, 0)];
                    for (int di = 0; di < 2; di++) sy += F->cy[IX(F, i_f + di, j_f + 1, 0)];
                    for (int di = 0; di < 2; di++) for (int dj = 0; dj < 2; dj++)
                        e += F->ex[IX(F, i_f + di, j_f + dj, 0)];
                }
                C->cx[cidx] = sx; C->cy[cidx] = sy; C->cz[cidx] = sz; C->ex[cidx] = e;
            }
    for (int ic = 1; ic <= C->nx; ic++)
        for (int jc = 1; jc <= C->ny; jc++)
            for (int kc = KLO(dim, C); kc <= KHI(dim, C); kc++) {
                size_t cidx = IX(C, ic, jc, kc);
                if (C->solid[cidx]) continue;
                double d = diag_sum(dim, C, cidx, IX(C, ic - 1, jc, kc), IX(C, ic, jc - 1, kc),
                                    dim == 3 ? IX(C, ic, jc, kc - 1) : cidx);
                if (C->ex[cidx] != 0.0) d += C->ex[cidx];
                C->diag[cidx] = d;
            }
}

static void sweep(int dim, Level* L, int parity) {
    for (int i = 1; i <= L->nx; i++)
        for (int j = 1; j <= L->ny; j++)
            for (int k = KLO(dim, L); k <= KHI(dim, L); k++) {
                if (((i + j + k) & 1) != parity) continue;   /* k = 0 in 2D: (i+j)&1 */
                size_t id = IX(L, i, j, k);
                if (L->solid[id] || L->diag[id] < 1e-30) continue;
                size_t im = IX(L, i - 1, j, k), jm = IX(L, i, j - 1, k);
                double nb;
                if (dim == 3) {
                    size_t km = IX(L, i, j, k - 1);
                    nb = L->cx[id] * L->p[IX(L, i + 1, j, k)] + L->cx[im] * L->p[im]
                       + L->cy[id] * L->p[IX(L, i, j + 1, k)] + L->cy[jm] * L->p[jm]
                       + L->cz[id] * L->p[IX(L, i, j, k + 1)] + L->cz[km] * L->p[km];
                } else {
                    nb = L->cx[id] * L->p[IX(L, i + 1, j, k)] + L->cx[im] * L->p[im]
                       + L->cy[id] * L->p[IX(L, i, j + 1, k)] + L->cy[jm] * L->p[jm];
                }
                L->p[id] = (L->b[id] + nb) / L->diag[id];
            }
}

static void smooth(int dim, Level* L, int sweeps, int reverse) {
    int p0 = reverse ? 0 : 1;
    for (int sw = 0; sw < sweeps; sw++) { sweep(dim, L, p0); sweep(dim, L, 1 - p0); }
}

static void restrict_residual(int dim, const Level* F, Level* C) {
    for (int ic = 1; ic <= C->nx; ic++)
        for (int jc = 1; jc <= C->ny; jc++)
            for (int kc = KLO(dim, C); kc <= KHI(dim, C); kc++) {
                size_t cidx = IX(C, ic, jc, kc);
                if (C->solid[cidx]) continue;
                int i_f = 2 * ic - 1, j_f = 2 * jc - 1, k_f = (dim == 3) ? 2 * kc - 1 : 0;
                double sum = 0;
                for (int di = 0; di < 2; di++)
                    for (int dj = 0; dj < 2; dj++)
                        for (int dk = 0; dk < (dim == 3 ? 2 : 1); dk++) {
                            int fi = i_f + di, fj = j_f + dj, fk = k_f + dk;
                            size_t fid = IX(F, fi, fj, fk);
                            if (F->solid[fid]) continue;
                            size_t im = IX(F, fi - 1, fj, fk), jm = IX(F, fi, fj - 1, fk);
                            double Ax;
                            if (dim == 3) {
                                size_t km = IX(F, fi, fj, fk - 1);
                                Ax = F->diag[fid] * F->p[fid]
                                   - F->cx[fid] * F->p[IX(F, fi + 1, fj, fk)] - F->cx[im] * F->p[im]
                                   - F->cy[fid] * F->p[IX(F, fi, fj + 1, fk)] - F->cy[jm] * F->p[jm]
                                   - F->cz[fid] * F->p[IX(F, fi, fj, fk + 1)] - F->cz[km] * F->p[km];
                            } else {
                                Ax = F->diag[fid] * F->p[fid]
                                   - F->cx[fid] * F->p[IX(F, fi + 1, fj, fk)] - F->cx[im] * F->p[im]
                                   - F->cy[fid] * F->p[IX(F, fi, fj + 1, fk)] - F->cy[jm] * F->p[jm];
                            }
                            sum += F->b[fid] - Ax;
                        }
                C->b[cidx] = sum;
            }
}

static void prolongate_add(int dim, const Level* C, Level* F) {
    for (int ic = 1; ic <= C->nx; ic++)
        for (int jc = 1; jc <= C->ny; jc++)
            for (int kc = KLO(dim, C); kc <= KHI(dim, C); kc++) {
                size_t cidx = IX(C, ic, jc, kc);
                if (C->solid[cidx]) continue;
                double val = 2.0 * C->p[cidx];
                int i_f = 2 * ic - 1, j_f = 2 * jc - 1, k_f = (dim == 3) ? 2 * kc - 1 : 0;
                for (int di = 0; di < 2; di++)
                    for (int dj = 0; dj < 2; dj++)
                        for (int dk = 0; dk < (dim == 3 ? 2 : 1); dk++) {
                            size_t fid = IX(F, i_f + di, j_f + dj, k_f + dk);
                            if (!F->solid[fid]) F->p[fid] += val;
                        }
            }
}

static void vcycle(Hier* H, int level) {
    int dim = H->dim;
    Level* L = &H->L[level];
    if (level == H->nl - 1) {
        for (int s = 0; s < H->o.coarsest_sweeps; s++) { smooth(dim, L, 1, 0); smooth(dim, L, 1, 1); }
        return;
    }
    smooth(dim, L, H->o.nu_pre, 0);
    Level* C = &H->L[level + 1];
    restrict_residual(dim, L, C);
    memset(C->p, 0, C->N * sizeof(double));
    vcycle(H, level + 1);
    prolongate_add(dim, C, L);
    smooth(dim, L, H->o.nu_post, 1);
}

static void setup_levels(Hier* H, const orc_grid* g) {
    Level* L0 = &H->L[0];
    for (size_t c = 0; c < L0->N; c++) L0->solid[c] = g->solid[c] ? 1 : 0;
    for (int l = 1; l < H->nl; l++) restrict_solid(H->dim, &H->L[l - 1], &H->L[l], H->o.hierarchy_mode);
    setup_fine(H, L0);
    for (int l = 1; l < H->nl; l++) galerkin(H->dim, &H->L[l - 1], &H->L[l]);
}

static void apply_hier(Hier* H, const orc_grid* g, const double* r, double* z) {
    int dim = H->dim;
    Level* L0 = &H->L[0];
    setup_levels(H, g);   /* every apply, as the reference does */
    for (int i = 1; i <= L0->nx; i++)
        for (int j = 1; j <= L0->ny; j++)
            for (int k = KLO(dim, L0); k <= KHI(dim, L0); k++) {
                size_t id = IX(L0, i, j, k);
                L0->b[id] = r[id];
            }
    for (int l = 0; l < H->nl; l++) memset(H->L[l].p, 0, H->L[l].N * sizeof(double));
    vcycle(H, 0);
    for (int i = 1; i <= L0->nx; i++)
        for (int j = 1; j <= L0->ny; j++)
            for (int k = KLO(dim, L0); k <= KHI(dim, L0); k++) {
                size_t id = IX(L0, i, j, k);
                z[id] = L0->p[id];
            }
}

static int check_grid(const orc_grid* g) {
    if (!g || !g->solid) return -1;
    if (g->dim != 2 && g->dim != 3) return -1;
    if (g->nx < 1 || g->ny < 1 || (g->dim == 3 && g->nz < 1)) return -1;
    const int nvar = (g->lap_diag != 0) + (g->lap_off_x != 0) + (g->lap_off_y != 0);
    if (nvar != 0 && (nvar != 3 || g->dim != 2)) return -1;   /* all three, 2D only (grid.h:21-23) */
    return 0;
}

int orc_uaamg_apply(const orc_grid* g, const orc_opts* o, const double* r, double* z) {
    orc_opts d; if (!o) { orc_default_opts(&d); o = &d; }
    if (check_grid(g)) return -1;
    Hier* H = malloc(sizeof(Hier));
    if (!H) return -2;
    int rc = build_hier(H, g, o);
    if (rc == 0) apply_hier(H, g, r, z);
    free_hier(H); free(H);
    return rc;
}

int orc_hierarchy_dims(const orc_grid* g, const orc_opts* o, int* nlevels, int* dims, int max_levels) {
    orc_opts d; if (!o) { orc_default_opts(&d); o = &d; }
    if (check_grid(g)) return -1;
    Hier* H = malloc(sizeof(Hier));
    if (!H) return -2;
    int rc = build_hier(H, g, o);
    if (rc == 0) {
        *nlevels = H->nl;
        for (int l = 0; l < H->nl && l < max_levels; l++) {
            dims[3 * l] = H->L[l].nx; dims[3 * l + 1] = H->L[l].ny; dims[3 * l + 2] = H->L[l].nz;
        }
    }
    free_hier(H); free(H);
    return rc;
}

int orc_level_data(const orc_grid* g, const orc_opts* o, int level,
                   unsigned char* solid, double* diag, double* cx, double* cy,
                   double* cz, double* excess) {
    orc_opts d; if (!o) { orc_default_opts(&d); o = &d; }
    if (check_grid(g)) return -1;
    Hier* H = malloc(sizeof(Hier));
    if (!H) return -2;
    int rc = build_hier(H, g, o);
    if (rc == 0 && (level < 0 || level >= H->nl)) rc = -4;
    if (rc == 0) {
        setup_levels(H, g);
        const Level* L = &H->L[level];
        if (solid) memcpy(solid, L->solid, L->N);
        if (diag) memcpy(diag, L->diag, L->N * sizeof(double));
        if (cx) memcpy(cx, L->cx, L->N * sizeof(double));
        if (cy) memcpy(cy, L->cy, L->N * sizeof(double));
        if (cz && g->dim == 3) memcpy(cz, L->cz, L->N * sizeof(double));
        if (excess) memcpy(excess, L->ex, L->N * sizeof(double));
    }
    free_hier(H); free(H);
    return rc;
}

/* ---------------- GMG and AMG preconditioners (2D; keys "pcg"/"pcg_gmg", "pcg_amg") ----------------
 * GMGPreconditioner (src/solver/preconditioner/gmg_preconditioner.cpp:13-194) and
 * AMGPreconditioner (amg_preconditioner.cpp:13-209): the same hierarchy shape
 * (:13-27 / :157-173), coarse solid iff at least half of the 2x2 children
 * are (:29-39 / :141-155), a smoother that adds (b - diag x + lap) / d_eff
 * in place per colour (:41-81 / :13-50), V(2,2) with 10 x (forward +
 * reverse) sweeps on the coarsest level (:163-174 / :175-188; here
 * coarsest_sweeps pairs, 10 by default), and
 * restriction by the mean of the fluid children's residuals; AMG forms the
 * residual as b - (diag x - lap) and accumulates it through its aggregate
 * map (:91-111), GMG as (b - diag x) + lap per coarse cell (:121-142). */

typedef struct {
    int nx, ny;
    double dx, dy, idx2, idy2, diag;
    double *p, *b;
    unsigned char* solid;
    size_t N;
} GLevel;

typedef struct { int nl, amg, csweeps, keep_fluid; GLevel L[MAXL]; } GHier;

#define GX2(L, i, j) ((size_t)(i) + (size_t)(j) * (size_t)((L)->nx + 2))

static void gmg_free(GHier* H) {
    for (int l = 0; l < H->nl; l++) { free(H->L[l].p); free(H->L[l].b); free(H->L[l].solid); }
    H->nl = 0;
}

static int gmg_build(GHier* H, const orc_grid* g, const orc_opts* o, int amg) {
    int nx = g->nx, ny = g->ny;
    double dx = g->dx, dy = g->dy;
    H->nl = 0; H->amg = amg; H->csweeps = o->coarsest_sweeps; H->keep_fluid = o->hierarchy_mode;
    while (nx >= 2 && ny >= 2 && H->nl < MAXL) {
        GLevel* L = &H->L[H->nl++];
        L->nx = nx; L->ny = ny; L->dx = dx; L->dy = dy;
        L->idx2 = 1.0 / (dx * dx); L->idy2 = 1.0 / (dy * dy);
        L->diag = amg ? 2.0 * L->idx2 + 2.0 * L->idy2 : 2.0 * (L->idx2 + L->idy2);
        L->N = (size_t)(nx + 2) * (ny + 2);
        L->p = calloc(L->N, sizeof(double)); L->b = calloc(L->N, sizeof(double)); L->solid = calloc(L->N, 1);
        if (!L->p || !L->b || !L->solid) return -2;
        if (nx <= o->coarsest_side || ny <= o->coarsest_side) break;
        nx /= 2; ny /= 2; dx *= 2.0; dy *= 2.0;
    }
    return 0;
}

/* one colour ((i + j) & 1 == odd) of the in-place smoother */
static void gmg_colour(GLevel* L, int odd) {
    const unsigned char* S = L->solid;
    const double idx2 = L->idx2, idy2 = L->idy2, diag = L->diag;
    for (int i = 1; i <= L->nx; i++)
        for (int j = 1 + ((i + odd + 1) % 2); j <= L->ny; j += 2) {
            size_t id = GX2(L, i, j);
            if (S[id]) continue;
            double* P = L->p;
            double pL = (i > 1 && !S[GX2(L, i - 1, j)]) ? P[GX2(L, i - 1, j)] : P[id];
            double pR = (i < L->nx && !S[GX2(L, i + 1, j)]) ? P[GX2(L, i + 1, j)] : P[id];
            double pB = (j > 1 && !S[GX2(L, i, j - 1)]) ? P[GX2(L, i, j - 1)] : P[id];
            double pT = (j < L->ny && !S[GX2(L, i, j + 1)]) ? P[GX2(L, i, j + 1)] : P[id];
            double lap = (pL + pR) * idx2 + (pB + pT) * idy2;
            double eff_d = diag;
            if (i == 1 || S[GX2(L, i - 1, j)]) eff_d -= idx2;
            if (i == L->nx || S[GX2(L, i + 1, j)]) eff_d -= idx2;
            if (j == 1 || S[GX2(L, i, j - 1)]) eff_d -= idy2;
            if (j == L->ny || S[GX2(L, i, j + 1)]) eff_d -= idy2;
            P[id] += ((eff_d < 1e-15) ? 0.0 : 1.0 / eff_d) * (L->b[id] - diag * P[id] + lap);
        }
}
/* smooth: odd then even colour; smoothReverse: even then odd */
static void gmg_smooth(GLevel* L, int sweeps, int reverse) {
    for (int s = 0; s < sweeps; s++) { gmg_colour(L, reverse ? 0 : 1); gmg_colour(L, reverse ? 1 : 0); }
}

static double gmg_resid(const GLevel* F, int i, int j, int amg) {
    const unsigned char* S = F->solid;
    const double* P = F->p;
    size_t id = GX2(F, i, j);
    double pL = (i > 1 && !S[GX2(F, i - 1, j)]) ? P[GX2(F, i - 1, j)] : P[id];
    double pR = (i < F->nx && !S[GX2(F, i + 1, j)]) ? P[GX2(F, i + 1, j)] : P[id];
    double pB = (j > 1 && !S[GX2(F, i, j - 1)]) ? P[GX2(F, i, j - 1)] : P[id];
    double pT = (j < F->ny && !S[GX2(F, i, j + 1)]) ? P[GX2(F, i, j + 1)] : P[id];
    if (amg) {
        double Ax = F->diag * P[id] - (pL + pR) * F->idx2 - (pB + pT) * F->idy2;
        return F->b[id] - Ax;
    }
    double lap = (pL + pR) * F->idx2 + (pB + pT) * F->idy2;
    return F->b[id] - F->diag * P[id] + lap;
}

/* mean of the fluid children's residuals (children in the order (0,0),(0,1),(1,0),(1,1)) */
static void gmg_restrict(const GLevel* F, GLevel* C, int amg) {
    for (int ic = 1; ic <= C->nx; ic++)
        for (int jc = 1; jc <= C->ny; jc++) {
            double sum = 0; int cnt = 0;
            for (int di = 0; di < 2; di++)
                for (int dj = 0; dj < 2; dj++) {
                    int fi = 2 * ic - 1 + di, fj = 2 * jc - 1 + dj;
                    if (F->solid[GX2(F, fi, fj)]) continue;
                    sum += gmg_resid(F, fi, fj, amg);
                    cnt++;
                }
            size_t cid = GX2(C, ic, jc);
            if (amg) C->b[cid] = (cnt > 0) ? sum / cnt : 0.0;
            else if (!C->solid[cid] && cnt > 0) C->b[cid] = sum / cnt;
        }
}

/* GMG adds x_c to the fluid children of fluid coarse cells (gmg :144-156); AMG
 * adds x_c of the aggregate to every fluid child (amg :113-123), which for a
 * solid coarse cell is its zero-filled x */
static void gmg_prolong(const GLevel* C, GLevel* F, int amg) {
    for (int ic = 1; ic <= C->nx; ic++)
        for (int jc = 1; jc <= C->ny; jc++) {
            size_t cid = GX2(C, ic, jc);
            if (C->solid[cid] && !amg) continue;
            for (int di = 0; di < 2; di++)
                for (int dj = 0; dj < 2; dj++) {
                    size_t fid = GX2(F, 2 * ic - 1 + di, 2 * jc - 1 + dj);
                    if (!F->solid[fid]) F->p[fid] += C->p[cid];
                }
        }
}

static void gmg_vcycle(GHier* H, int l) {
    GLevel* L = &H->L[l];
    if (l == H->nl - 1) {
        for (int s = 0; s < H->csweeps; s++) { gmg_smooth(L, 1, 0); gmg_smooth(L, 1, 1); }
        return;
    }
    gmg_smooth(L, 2, 0);
    GLevel* C = &H->L[l + 1];
    gmg_restrict(L, C, H->amg);
    memset(C->p, 0, C->N * sizeof(double));
    gmg_vcycle(H, l + 1);
    gmg_prolong(C, L, H->amg);
    gmg_smooth(L, 2, 1);
}

static void gmg_apply(GHier* H, const orc_grid* g, const double* r, double* z) {
    GLevel* L0 = &H->L[0];
    for (size_t c = 0; c < L0->N; c++) L0->solid[c] = g->solid[c] ? 1 : 0;
    for (int l = 1; l < H->nl; l++) {   /* at least half of the 2x2 children solid (all with hierarchy_mode 1) */
        GLevel *F = &H->L[l - 1], *C = &H->L[l];
        for (int ic = 1; ic <= C->nx; ic++)
            for (int jc = 1; jc <= C->ny; jc++) {
                int sc = 0;
                for (int di = 0; di < 2; di++)
                    for (int dj = 0; dj < 2; dj++) sc += F->solid[GX2(F, 2 * ic - 1 + di, 2 * jc - 1 + dj)] ? 1 : 0;
                C->solid[GX2(C, ic, jc)] = sc >= (H->keep_fluid ? 4 : 2);
            }
    }
    for (int i = 1; i <= L0->nx; i++)
        for (int j = 1; j <= L0->ny; j++) L0->b[GX2(L0, i, j)] = r[GX2(L0, i, j)];
    for (int l = 0; l < H->nl; l++) memset(H->L[l].p, 0, H->L[l].N * sizeof(double));
    gmg_vcycle(H, 0);
    for (int i = 1; i <= L0->nx; i++)
        for (int j = 1; j <= L0->ny; j++) z[GX2(L0, i, j)] = L0->p[GX2(L0, i, j)];
}

/* ---------------- PCG (pcg.cpp / pcg_3d.cpp) ---------------- */

typedef struct { int dim, nx, ny, nz; const unsigned char* s; } Dom;

static inline size_t GX(const Dom* d, int i, int j, int k) {
    size_t si = (size_t)d->nx + 2;
    return (size_t)i + (size_t)j * si + (size_t)k * si * ((size_t)d->ny + 2);
}

#define FOR_INT(d) \
    for (int i = 1; i <= (d)->nx; i++) \
        for (int j = 1; j <= (d)->ny; j++) \
            for (int k = ((d)->dim == 3 ? 1 : 0); k <= ((d)->dim == 3 ? (d)->nz : 0); k++)

static double dotf(const Dom* d, const double* a, const double* b) {
    double s = 0;
    FOR_INT(d) { size_t id = GX(d, i, j, k); if (!d->s[id]) s += a[id] * b[id]; }
    return s;
}

static void subtract_mean(const Dom* d, double* v) {
    double sum = 0; int count = 0;
    FOR_INT(d) { size_t id = GX(d, i, j, k); if (!d->s[id]) { sum += v[id]; count++; } }
    double mean = (count > 0) ? sum / count : 0.0;
    FOR_INT(d) { size_t id = GX(d, i, j, k); if (!d->s[id]) v[id] -= mean; }
}

static void matvec_u(const orc_grid* g, const double* v, double* Av) {
    Dom dd = {g->dim, g->nx, g->ny, g->dim == 3 ? g->nz : 1, g->solid}; const Dom* d = &dd;
    double idx2 = 1.0 / (g->dx * g->dx), idy2 = 1.0 / (g->dy * g->dy), idz2 = 1.0 / (g->dz * g->dz);
    double diag = (g->dim == 3) ? 2.0 * (idx2 + idy2 + idz2) : 2.0 * (idx2 + idy2);
    const unsigned char* S = g->solid;
    const int* w = g->wall;
    int nx = g->nx, ny = g->ny, nz = g->nz;
    FOR_INT(d) {
        size_t idx = GX(d, i, j, k);
        if (S[idx]) { Av[idx] = 0.0; continue; }
        double vc = v[idx];
        double vL = (i > 1 && !S[GX(d, i - 1, j, k)]) ? v[GX(d, i - 1, j, k)] : ((i == 1 && w[0]) ? -vc : vc);
        double vR = (i < nx && !S[GX(d, i + 1, j, k)]) ? v[GX(d, i + 1, j, k)] : ((i == nx && w[1]) ? -vc : vc);
        double vB = (j > 1 && !S[GX(d, i, j - 1, k)]) ? v[GX(d, i, j - 1, k)] : ((j == 1 && w[2]) ? -vc : vc);
        double vT = (j < ny && !S[GX(d, i, j + 1, k)]) ? v[GX(d, i, j + 1, k)] : ((j == ny && w[3]) ? -vc : vc);
        if (g->dim == 3) {
            double vF = (k > 1 && !S[GX(d, i, j, k - 1)]) ? v[GX(d, i, j, k - 1)] : ((k == 1 && w[4]) ? -vc : vc);
            double vK = (k < nz && !S[GX(d, i, j, k + 1)]) ? v[GX(d, i, j, k + 1)] : ((k == nz && w[5]) ? -vc : vc);
            Av[idx] = diag * vc - (vL + vR) * idx2 - (vB + vT) * idy2 - (vF + vK) * idz2;
        } else if (g->lap_diag) {
            /* variable coefficients, pcg.cpp:44: same neighbour rule, one off-diagonal per axis */
            Av[idx] = g->lap_diag[idx] * vc + g->lap_off_x[idx] * (vL + vR) + g->lap_off_y[idx] * (vB + vT);
        } else {
            Av[idx] = diag * vc - (vL + vR) * idx2 - (vB + vT) * idy2;
        }
    }
}

int orc_matvec(const orc_grid* g, const double* v, double* Av) {
    if (check_grid(g)) return -1;
    matvec_u(g, v, Av);
    return 0;
}

static size_t padded_size(const orc_grid* g) {
    return (size_t)(g->nx + 2) * (g->ny + 2) * (g->dim == 3 ? (size_t)(g->nz + 2) : 1);
}

/* one application of the preconditioner that o->precond names (z zeroed first) */
int orc_precond_apply(const orc_grid* g, const orc_opts* o, const double* r, double* z) {
    orc_opts d; if (!o) { orc_default_opts(&d); o = &d; }
    if (check_grid(g)) return -1;
    const size_t N = padded_size(g);
    if (o->precond == ORC_PRECOND_UAAMG) return orc_uaamg_apply(g, o, r, z);
    if (o->precond == ORC_PRECOND_IDENTITY) { memcpy(z, r, N * sizeof(double)); return 0; }
    if (o->precond != ORC_PRECOND_GMG && o->precond != ORC_PRECOND_AMG) return -1;
    if (g->dim != 2) return -1;
    GHier* G = calloc(1, sizeof(GHier));
    if (!G) return -2;
    int rc = gmg_build(G, g, o, o->precond == ORC_PRECOND_AMG);
    if (rc == 0) { memset(z, 0, N * sizeof(double)); gmg_apply(G, g, r, z); }
    gmg_free(G); free(G);
    return rc;
}

int orc_pcg(const orc_grid* g, const orc_opts* o, const double* rhs_in, double* p,
            int max_iter, double tol, orc_result* res) {
    orc_opts dflt; if (!o) { orc_default_opts(&dflt); o = &dflt; }
    if (check_grid(g)) return -1;
    Dom dd = {g->dim, g->nx, g->ny, g->dim == 3 ? g->nz : 1, g->solid}; const Dom* d = &dd;
    int all_neumann = 1;
    for (int f = 0; f < (g->dim == 3 ? 6 : 4); f++) if (g->wall[f] != ORC_NEUMANN) all_neumann = 0;
    int project = (o->project_mean < 0) ? all_neumann : o->project_mean;
    size_t N = padded_size(g);
    double* rhs = malloc(N * sizeof(double));
    double* r = malloc(N * sizeof(double));
    double* z = calloc(N, sizeof(double));
    double* pd = malloc(N * sizeof(double));
    double* Ap = calloc(N, sizeof(double));
    Hier* H = malloc(sizeof(Hier));
    GHier* G = calloc(1, sizeof(GHier));
    orc_result rr = {0, 1, 0, 0.0, 0.0};
    int rc = 0;
    const int gmg = o->precond == ORC_PRECOND_GMG || o->precond == ORC_PRECOND_AMG;
    if (!rhs || !r || !z || !pd || !Ap || !H || !G) { rc = -2; goto out_nohier; }
    if (gmg && g->dim != 2) { rc = -1; goto out_nohier; }
    rc = build_hier(H, g, o);
    if (rc) goto out_nohier;
    if (gmg && (rc = gmg_build(G, g, o, o->precond == ORC_PRECOND_AMG)) != 0) goto out;

    memcpy(rhs, rhs_in, N * sizeof(double));
    if (project) {
        double sum = 0; int count = 0;
        FOR_INT(d) { size_t id = GX(d, i, j, k); if (!d->s[id]) { sum += rhs[id]; count++; } }
        double mean = (count > 0) ? sum / count : 0.0;
        FOR_INT(d) { size_t id = GX(d, i, j, k); if (!d->s[id]) rhs[id] = -(rhs[id] - mean); }
    } else {
        FOR_INT(d) { size_t id = GX(d, i, j, k); if (!d->s[id]) rhs[id] = -rhs[id]; }
    }
    memcpy(r, rhs, N * sizeof(double));
    if (project) subtract_mean(d, r);
    const int ident = o->precond == ORC_PRECOND_IDENTITY;   /* z = r: identity_preconditioner.h:19 */
    if (ident) memcpy(z, r, N * sizeof(double)); else if (gmg) gmg_apply(G, g, r, z); else apply_hier(H, g, r, z);
    rr.applies++;
    if (project) subtract_mean(d, z);
    memcpy(pd, z, N * sizeof(double));
    double rsold = dotf(d, r, z);
    double r0sq = dotf(d, r, r);
    rr.r0norm = sqrt(r0sq);
    rr.rnorm = rr.r0norm;
    double tolsq = (r0sq > 0) ? r0sq * tol * tol : tol * tol;

    for (int it = 0; it < max_iter; it++) {
        matvec_u(g, pd, Ap);
        double pAp = dotf(d, pd, Ap);
        if (pAp < 1e-15) { rr.status = 2; rr.iters = it + 1; goto out; }
        double alpha = rsold / pAp;
        FOR_INT(d) {
            size_t id = GX(d, i, j, k);
            if (!d->s[id]) { p[id] += alpha * pd[id]; r[id] -= alpha * Ap[id]; }
        }
        double rsnew = dotf(d, r, r);
        rr.rnorm = sqrt(rsnew);
        int conv = (o->tol_mode == 1) ? (rsnew < tolsq) : (sqrt(rsnew) < tol);
        if (conv) { rr.status = 0; rr.iters = it + 1; goto out; }
        if (ident) memcpy(z, r, N * sizeof(double)); else if (gmg) gmg_apply(G, g, r, z); else apply_hier(H, g, r, z);
        rr.applies++;
        if (project) subtract_mean(d, z);
        double beta = dotf(d, r, z) / rsold;
        FOR_INT(d) { size_t id = GX(d, i, j, k); if (!d->s[id]) pd[id] = z[id] + beta * pd[id]; }
        rsold = dotf(d, r, z);
    }
    rr.status = 1; rr.iters = max_iter;
out:
    free_hier(H);
out_nohier:
    if (G) gmg_free(G);
    free(G); free(H); free(rhs); free(r); free(z); free(pd); free(Ap);
    if (res) *res = rr;
    return rc;
}

/* ---------------- standalone amgpcg (test_amgpcg.cpp + test_common.h) ---------------- */

typedef struct { int N; double h; double* v; } AG;

static int ag_new(AG* a, int n) {
    a->N = n; a->h = 1.0 / (n + 1);
    a->v = calloc((size_t)(n + 2) * (n + 2), sizeof(double));
    return a->v ? 0 : -1;
}
#define AV(a, i, j) ((a)->v[(size_t)(i) * ((a)->N + 2) + (j)])

static double ag_Ax(const AG* x, int i, int j) {
    double inv_h2 = 1.0 / (x->h * x->h);
    return (4 * AV(x, i, j) - AV(x, i + 1, j) - AV(x, i - 1, j) - AV(x, i, j + 1) - AV(x, i, j - 1)) * inv_h2;
}
static double ag_dot(const AG* a, const AG* b) {
    double s = 0;
    for (int i = 1; i <= a->N; i++) for (int j = 1; j <= a->N; j++) s += AV(a, i, j) * AV(b, i, j);
    return s;
}
static void ag_axpy(double al, const AG* x, AG* y) {
    for (int i = 1; i <= x->N; i++) for (int j = 1; j <= x->N; j++) AV(y, i, j) += al * AV(x, i, j);
}
static void ag_rbgs(AG* x, const AG* b, int iterations) {
    double inv_diag = x->h * x->h / 4.0;
    int N = x->N;
    for (int sw = 0; sw < iterations; sw++) {
        for (int i = 1; i <= N; i++)
            for (int j = 1 + (i % 2); j <= N; j += 2)
                AV(x, i, j) += inv_diag * (AV(b, i, j) - ag_Ax(x, i, j));
        for (int i = 1; i <= N; i++)
            for (int j = 1 + ((i + 1) % 2); j <= N; j += 2)
                AV(x, i, j) += inv_diag * (AV(b, i, j) - ag_Ax(x, i, j));
    }
}

static int ag_levels(int N, int* lv, int maxl) {
    int n = N, c = 0;
    while (n >= 2) {
        if (c >= maxl) return -1;
        lv[c++] = n;
        if (n <= 3) break;
        n = n / 2; if (n < 2) n = 2;
    }
    return c;
}

typedef struct { int nl; AG x[MAXL], b[MAXL]; } AH;

static void ag_vcycle(AH* A, int level) {
    AG* x = &A->x[level];
    const AG* b = &A->b[level];
    if (level == A->nl - 1) { ag_rbgs(x, b, 20); return; }
    ag_rbgs(x, b, 2);
    AG r; ag_new(&r, x->N);
    for (int i = 1; i <= x->N; i++) for (int j = 1; j <= x->N; j++) AV(&r, i, j) = AV(b, i, j) - ag_Ax(x, i, j);
    /* restrict(): coarse size Nf/2 (>= 2), made even when > 2 */
    int Nc = r.N / 2; if (Nc < 2) Nc = 2; if (Nc % 2 != 0 && Nc > 2) Nc--;
    AG rc; ag_new(&rc, Nc);
    for (int ic = 1; ic <= Nc; ic++)
        for (int jc = 1; jc <= Nc; jc++) {
            int i_f = 2 * ic - 1, j_f = 2 * jc - 1;
            AV(&rc, ic, jc) = 0.25 * (AV(&r, i_f, j_f) + AV(&r, i_f + 1, j_f) + AV(&r, i_f, j_f + 1) + AV(&r, i_f + 1, j_f + 1));
        }
    AG* bc = &A->b[level + 1];
    /* b_c(i,j) = r_c(i,j) read with r_c's own stride: rows/cols past Nc hit r_c's zero ghosts */
    for (int i = 1; i <= bc->N; i++)
        for (int j = 1; j <= bc->N; j++) AV(bc, i, j) = rc.v[(size_t)i * (rc.N + 2) + j];
    AG* xc = &A->x[level + 1];
    memset(xc->v, 0, (size_t)(xc->N + 2) * (xc->N + 2) * sizeof(double));
    ag_vcycle(A, level + 1);
    /* prolongate(x_c, Nf, 1.0) */
    AG corr; ag_new(&corr, x->N);
    for (int ic = 1; ic <= xc->N; ic++)
        for (int jc = 1; jc <= xc->N; jc++) {
            double val = AV(xc, ic, jc) * 1.0;
            int i_f = 2 * ic - 1, j_f = 2 * jc - 1;
            AV(&corr, i_f, j_f) = val; AV(&corr, i_f + 1, j_f) = val;
            AV(&corr, i_f, j_f + 1) = val; AV(&corr, i_f + 1, j_f + 1) = val;
        }
    for (int i = 1; i <= x->N; i++) for (int j = 1; j <= x->N; j++) AV(x, i, j) += AV(&corr, i, j);
    ag_rbgs(x, b, 2);
    free(r.v); free(rc.v); free(corr.v);
}

int orc_amgpcg(int N, const double* bin, double* xout, int max_iter, double tol, int* iters) {
    if (N < 2) return -1;
    int lv[MAXL];
    AH* A = calloc(1, sizeof(AH));
    if (!A) return -2;
    A->nl = ag_levels(N, lv, MAXL);
    for (int l = 0; l < A->nl; l++) { ag_new(&A->x[l], lv[l]); ag_new(&A->b[l], lv[l]); }
    AG x, r, z, p, Ap;
    ag_new(&x, N); ag_new(&r, N); ag_new(&z, N); ag_new(&p, N); ag_new(&Ap, N);
    size_t NN = (size_t)(N + 2) * (N + 2);
    for (int i = 1; i <= N; i++) for (int j = 1; j <= N; j++) AV(&r, i, j) = bin[(size_t)i * (N + 2) + j];
    int nl = A->nl;
    /* precondition(z, r) */
#define PRECOND() do { \
        for (int i = 1; i <= N; i++) for (int j = 1; j <= N; j++) AV(&A->b[0], i, j) = AV(&r, i, j); \
        for (int l = 0; l < nl; l++) memset(A->x[l].v, 0, (size_t)(A->x[l].N + 2) * (A->x[l].N + 2) * sizeof(double)); \
        ag_vcycle(A, 0); \
        for (int i = 1; i <= N; i++) for (int j = 1; j <= N; j++) AV(&z, i, j) = AV(&A->x[0], i, j); \
    } while (0)
    PRECOND();
    memcpy(p.v, z.v, NN * sizeof(double));
    double rsold = ag_dot(&r, &z);
    int done = 0;
    for (int k = 0; k < max_iter; k++) {
        for (int i = 1; i <= N; i++) for (int j = 1; j <= N; j++) AV(&Ap, i, j) = ag_Ax(&p, i, j);
        double alpha = rsold / ag_dot(&p, &Ap);
        ag_axpy(alpha, &p, &x); ag_axpy(-alpha, &Ap, &r);
        double rnorm = sqrt(ag_dot(&r, &r));
        if (rnorm < tol) { *iters = k + 1; done = 1; break; }
        PRECOND();
        double rsnew = ag_dot(&r, &z);
        double beta = rsnew / rsold;
        for (int i = 1; i <= N; i++) for (int j = 1; j <= N; j++) AV(&p, i, j) = AV(&z, i, j) + beta * AV(&p, i, j);
        rsold = rsnew;
    }
#undef PRECOND
    if (!done) *iters = max_iter;
    memcpy(xout, x.v, NN * sizeof(double));
    for (int l = 0; l < nl; l++) { free(A->x[l].v); free(A->b[l].v); }
    free(A); free(x.v); free(r.v); free(z.v); free(p.v); free(Ap.v);
    return 0;
}

/* ---------------------------------------------------------------- projection
 * PressureProjection::project (src/pressure/pressure.cpp:5-32) and
 * PressureProjection3D::project (src/pressure/pressure_3d.cpp:5-39):
 *   1. rhs = div(u)/dt on non-solid interior cells (fvc::divergence,
 *      include/ops/operators.h:18-21, operators_3d.h:17-21), 0 elsewhere;
 *   2. p = 0 everywhere, then the PCG(UAAMG) solve (orc_pcg);
 *   3. on interior faces between two non-solid cells, u -= dt*(p+ - p-)/dx
 *      (and v, w), in the reference's loop ranges.
 * MAC layouts of mesh.h:27-35 / mesh_3d.h:29-37. */
static size_t mu(const orc_grid* g, int i, int j, int k) {
    return (size_t)i + (size_t)j * (g->nx + 1) + (size_t)k * (g->nx + 1) * (g->ny + 2);
}
static size_t mv(const orc_grid* g, int i, int j, int k) {
    return (size_t)i + (size_t)j * (g->nx + 2) + (size_t)k * (g->nx + 2) * (g->ny + 1);
}
static size_t mw(const orc_grid* g, int i, int j, int k) {
    return (size_t)i + (size_t)j * (g->nx + 2) + (size_t)k * (g->nx + 2) * (g->ny + 2);
}
static size_t mp(const orc_grid* g, int i, int j, int k) {
    return (size_t)i + (size_t)j * (g->nx + 2) + (size_t)k * (g->nx + 2) * (g->ny + 2);
}

void orc_project_rhs(const orc_grid* g, const double* u, const double* v, const double* w, double dt, double* rhs) {
    const int k0 = g->dim == 3 ? 1 : 0, k1 = g->dim == 3 ? g->nz : 0;
    memset(rhs, 0, padded_size(g) * sizeof(double));
    for (int k = k0; k <= k1; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i <= g->nx; i++) {
                if (g->solid[mp(g, i, j, k)]) continue;
                double div = (u[mu(g, i, j, k)] - u[mu(g, i - 1, j, k)]) / g->dx
                           + (v[mv(g, i, j, k)] - v[mv(g, i, j - 1, k)]) / g->dy;
                if (g->dim == 3) div = div + (w[mw(g, i, j, k)] - w[mw(g, i, j, k - 1)]) / g->dz;
                rhs[mp(g, i, j, k)] = div / dt;
            }
}

void orc_project_correct(const orc_grid* g, double* u, double* v, double* w, const double* p, double dt) {
    const int k0 = g->dim == 3 ? 1 : 0, k1 = g->dim == 3 ? g->nz : 0;
    const unsigned char* s = g->solid;
    for (int k = k0; k <= k1; k++)
        for (int j = 1; j <= g->ny; j++)
            for (int i = 1; i < g->nx; i++) {
                if (s[mp(g, i, j, k)] || s[mp(g, i + 1, j, k)]) continue;
                u[mu(g, i, j, k)] -= dt * (p[mp(g, i + 1, j, k)] - p[mp(g, i, j, k)]) / g->dx;
            }
    for (int k = k0; k <= k1; k++)
        for (int j = 1; j < g->ny; j++)
            for (int i = 1; i <= g->nx; i++) {
                if (s[mp(g, i, j, k)] || s[mp(g, i, j + 1, k)]) continue;
                v[mv(g, i, j, k)] -= dt * (p[mp(g, i, j + 1, k)] - p[mp(g, i, j, k)]) / g->dy;
            }
    if (g->dim == 3)
        for (int k = 1; k < g->nz; k++)
            for (int j = 1; j <= g->ny; j++)
                for (int i = 1; i <= g->nx; i++) {
                    if (s[mp(g, i, j, k)] || s[mp(g, i, j, k + 1)]) continue;
                    w[mw(g, i, j, k)] -= dt * (p[mp(g, i, j, k + 1)] - p[mp(g, i, j, k)]) / g->dz;
                }
}

int orc_project(const orc_grid* g, const orc_opts* o, double* u, double* v, double* w, double* p, double dt,
                int max_iter, double tol, orc_result* res) {
    if (check_grid(g)) return -1;
    size_t N = padded_size(g);
    double* rhs = malloc(N * sizeof(double));
    if (!rhs) return -2;
    orc_project_rhs(g, u, v, w, dt, rhs);
    memset(p, 0, N * sizeof(double));
    int rc = orc_pcg(g, o, rhs, p, max_iter, tol, res);
    free(rhs);
    if (rc) return rc;
    orc_project_correct(g, u, v, w, p, dt);
    return 0;
}
