// kernels_cg.hpp -- device-resident PCG on level 0: FP64 reductions written
// as one partial per workgroup and summed in block order by one-workgroup
// finish kernels, which also decide alpha/beta/mean/convergence on device, so
// the host only polls Ctrl::done every few iterations.
//
// The vector kernels run flat over the padded array (4 consecutive cells per
// thread, 16/32-byte vector loads; cells whose neighbour byte has no fluid bit
// are left alone).  Each kernel body takes its (virtual) block index and
// block count, so the same body serves a single grid (blockIdx) and one grid's
// slice of a batched work-list.
//
// Reference behaviour (paths relative to the reference root):
//   src/solver/pcg_3d.cpp:18-125 / pcg.cpp:18-126
//     rhs' = -(rhs - mean_fluid(rhs)); r = rhs'; subtract_mean(r)        (:150-162, :206-207)
//     z = M r; subtract_mean(z); p = z; rsold = (r,z)                    (:210-215)
//     loop: Ap; pAp < 1e-15 -> return; alpha; x += alpha p; r -= alpha Ap;
//           sqrt(r.r) < tol -> return; z = M r; subtract_mean(z);
//           beta = (r,z)/rsold; p = z + beta p; rsold = (r,z)             (:219-250)
//   matvec with uniform constants and ghost = self at walls/solids       (:164-181)
//   relative tolerance r.r < tol^2 r0.r0                                 (cuda_pcg_3d_opt.cu:376-380)
//
// Precision: vectors are stored as T (FP32 or FP64) but every update and the
// operator A p are evaluated in FP64 and rounded once.  In FP32 the stencil
// difference diag*v - sum(c*v_nb) cancels to ~1/kappa of its terms once p is
// smooth, so evaluating it in FP32 loses most digits and stalls the solve.
#pragma once
#include "kernels_vcycle.hpp"

namespace uaamg {

template <typename T> using vec4 = T __attribute__((ext_vector_type(4)));
// four consecutive values at an element-aligned (not vec4-aligned) address:
// one unaligned vector load on gfx950
template <typename T> using vec4u = T __attribute__((ext_vector_type(4), aligned(sizeof(T))));

// Grid-stride loop over 4-cell vectors.  fv_b / fv_n are the (virtual)
// block index and block count: blockIdx.x / gridDim.x for a single grid, or a
// work-list item's slice of one grid in the batched kernels.
#define FLAT_LOOP(N4)                                                                          \
    for (long v_ = (long)fv_b * blockDim.x + threadIdx.x; v_ < (N4) / 4;                      \
         v_ += (long)fv_n * blockDim.x)

// Stencil form: with fv_n a multiple of 8, blocks with fv_b % 8 == x share the
// contiguous slab x of 8 (block b of a launch runs on XCD b % 8), so the
// neighbour planes a block reads were mostly fetched into that XCD's L2 by
// the same XCD; otherwise the plain grid-stride order.  Either way every
// vec4 index is visited by exactly one thread.
#define FLAT_LOOP_XCD(N4)                                                                      \
    const long xv_n_ = (N4) / 4;                                                               \
    const bool xv_on_ = (fv_n & 7) == 0;                                                       \
    const long xv_len_ = (xv_n_ + 7) / 8;                                                      \
    const long xv_beg_ = xv_on_ ? (fv_b & 7) * xv_len_ + (long)(fv_b >> 3) * blockDim.x        \
                                : (long)fv_b * blockDim.x;                                     \
    const long xv_end_ = xv_on_ ? min(xv_n_, ((fv_b & 7) + 1) * xv_len_) : xv_n_;              \
    const long xv_step_ = (long)(xv_on_ ? (fv_n >> 3) : fv_n) * blockDim.x;                    \
    for (long v_ = xv_beg_ + threadIdx.x; v_ < xv_end_; v_ += xv_step_)

__device__ __forceinline__ uint8_t nbyte(uint32_t w, int q) { return (uint8_t)(w >> (8 * q)); }

// (i, j, k) of a padded index (only needed for Dirichlet walls)
template <int D>
__device__ __forceinline__ void coords(const Geo& g, long c, int& i, int& j, int& k) {
    i = (int)(c % g.sx);
    long rest = c / g.sx;
    if (D == 3) { j = (int)(rest % (g.ny + 2)); k = (int)(rest / (g.ny + 2)); }
    else { j = (int)rest; k = 0; }
}

__global__ void k_ctrl_init(Ctrl* ctrl, const double* count, int max_iter, double tol, int tol_mode, int project,
                            int flexible, double* hist) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        Ctrl c = {};
        c.flexible = flexible;
        c.hist = hist;
        c.count = *count;
        c.max_iter = max_iter;
        c.tol = tol;
        c.tol_mode = tol_mode;
        c.project = project;
        c.first = 1;
        c.done = (max_iter <= 0) ? 1 : 0;
        c.status = (max_iter <= 0) ? ST_MAXITER : ST_RUNNING;
        *ctrl = c;
    }
}

// Sum of rhs over fluid cells -> m1 (the mean removed from the rhs).
__device__ __forceinline__ void d_init_sum(const double* __restrict__ rhs, const uint8_t* __restrict__ nbr, long N4,
                           double* partials, Ctrl* ctrl, int fv_b, int fv_n) {
    double v[1] = {0.0};
    FLAT_LOOP(N4) {
        long c = 4 * v_;
        uint32_t w = *(const uint32_t*)(nbr + c);
        if (!(w & 0x01010101u)) continue;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (nbyte(w, q) & NB_FLUID) v[0] += rhs[c + q];
    }
    block_store<1>(v, partials);
}
__global__ void k_init_sum(const double* __restrict__ rhs, const uint8_t* __restrict__ nbr, long N4,
                           double* partials, Ctrl* ctrl) {
    d_init_sum(rhs, nbr, N4, partials, ctrl, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void d_fin_init_sum(const double* partials, int nbt, int base, int nb, Ctrl* ctrl) {
    double tot[1];
    reduce_partials_at<1>(partials, nbt, base, nb, tot);
    if (threadIdx.x == 0) ctrl->m1 = (ctrl->project && ctrl->count > 0) ? tot[0] / ctrl->count : 0.0;
}
__global__ void k_fin_init_sum(const double* partials, int nb, Ctrl* ctrl) {
    d_fin_init_sum(partials, nb, 0, nb, ctrl);
}

// r = -(rhs - m1) on fluid cells (-rhs without projection), 0 elsewhere; m2 = mean(r).
template <typename T>
__device__ __forceinline__ void d_init_r(const double* __restrict__ rhs, T* __restrict__ r, const uint8_t* __restrict__ nbr,
                         long N4, double* partials, Ctrl* ctrl, int fv_b, int fv_n) {
    const double m1 = ctrl->m1;
    const int project = ctrl->project;
    double v[1] = {0.0};
    FLAT_LOOP(N4) {
        long c = 4 * v_;
        uint32_t w = *(const uint32_t*)(nbr + c);
        vec4<T> out = (T)0;
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (nbyte(w, q) & NB_FLUID) {
                T val = project ? (T)(-(rhs[c + q] - m1)) : (T)(-rhs[c + q]);
                out[q] = val;
                v[0] += (double)val;
            }
        *(vec4<T>*)(r + c) = out;
    }
    block_store<1>(v, partials);
}
template <typename T>
__global__ void k_init_r(const double* __restrict__ rhs, T* __restrict__ r, const uint8_t* __restrict__ nbr,
                         long N4, double* partials, Ctrl* ctrl) {
    d_init_r<T>(rhs, r, nbr, N4, partials, ctrl, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void d_fin_init_r(const double* partials, int nbt, int base, int nb, Ctrl* ctrl) {
    double tot[1];
    reduce_partials_at<1>(partials, nbt, base, nb, tot);
    if (threadIdx.x == 0) ctrl->m2 = (ctrl->project && ctrl->count > 0) ? tot[0] / ctrl->count : 0.0;
}
__global__ void k_fin_init_r(const double* partials, int nb, Ctrl* ctrl) {
    d_fin_init_r(partials, nb, 0, nb, ctrl);
}

// r -= m2 on fluid cells; r0.r0 and the squared stopping threshold.
template <typename T>
__device__ __forceinline__ void d_init_sub(T* __restrict__ r, const uint8_t* __restrict__ nbr, long N4, double* partials, Ctrl* ctrl, int fv_b, int fv_n) {
    const double m2 = ctrl->m2;
    const int project = ctrl->project;
    double v[1] = {0.0};
    FLAT_LOOP(N4) {
        long c = 4 * v_;
        uint32_t w = *(const uint32_t*)(nbr + c);
        if (!(w & 0x01010101u)) continue;
        vec4<T> rv = *(const vec4<T>*)(r + c);
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (nbyte(w, q) & NB_FLUID) {
                T val = rv[q];
                if (project) { val = (T)((double)val - m2); rv[q] = val; }
                v[0] += (double)val * (double)val;
            }
        *(vec4<T>*)(r + c) = rv;
    }
    block_store<1>(v, partials);
}
template <typename T>
__global__ void k_init_sub(T* __restrict__ r, const uint8_t* __restrict__ nbr, long N4, double* partials, Ctrl* ctrl) {
    d_init_sub<T>(r, nbr, N4, partials, ctrl, blockIdx.x, gridDim.x);
}
__device__ __forceinline__ void d_fin_init_sub(const double* partials, int nbt, int base, int nb, Ctrl* ctrl) {
    double tot[1];
    reduce_partials_at<1>(partials, nbt, base, nb, tot);
    if (threadIdx.x == 0) {
        ctrl->r0sq = tot[0];
        ctrl->rsnew = tot[0];
        ctrl->tolsq = (tot[0] > 0) ? tot[0] * ctrl->tol * ctrl->tol : ctrl->tol * ctrl->tol;
    }
}
__global__ void k_fin_init_sub(const double* partials, int nb, Ctrl* ctrl) {
    d_fin_init_sub(partials, nb, 0, nb, ctrl);
}

// The PCG operator at one fluid cell, in FP64: uniform constants, ghost = self
// on Neumann walls and solid neighbours, ghost = -self on Dirichlet walls.
// vl..vk are the six neighbour values (loaded unconditionally by the caller).
template <int D>
__device__ __forceinline__ double apply_A_at(int i, int j, int k, uint8_t nb, double vc, double vl, double vr, double vb,
                                             double vt, double vf, double vk, const Geo& g, const Consts<double>& c) {
    double vL = (nb & NB_XM) ? vl : ((c.any_dirichlet && i == 1 && c.wall[0]) ? -vc : vc);
    double vR = (nb & NB_XP) ? vr : ((c.any_dirichlet && i == g.nx && c.wall[1]) ? -vc : vc);
    double vB = (nb & NB_YM) ? vb : ((c.any_dirichlet && j == 1 && c.wall[2]) ? -vc : vc);
    double vT = (nb & NB_YP) ? vt : ((c.any_dirichlet && j == g.ny && c.wall[3]) ? -vc : vc);
    double Av = c.diag_u * vc - (vL + vR) * c.idx2 - (vB + vT) * c.idy2;
    if (D == 3) {
        double vF = (nb & NB_ZM) ? vf : ((c.any_dirichlet && k == 1 && c.wall[4]) ? -vc : vc);
        double vK = (nb & NB_ZP) ? vk : ((c.any_dirichlet && k == g.nz && c.wall[5]) ? -vc : vc);
        Av = Av - (vF + vK) * c.idz2;
    }
    return Av;
}
template <int D>
__device__ __forceinline__ double apply_A(long id, uint8_t nb, double vc, double vl, double vr, double vb, double vt,
                                          double vf, double vk, const Geo& g, const Consts<double>& c) {
    int i = 0, j = 0, k = 0;
    if (c.any_dirichlet) coords<D>(g, id, i, j, k);
    return apply_A_at<D>(i, j, k, nb, vc, vl, vr, vb, vt, vf, vk, g, c);
}

// Variable Laplacian (2D): Grid::lap_diag / lap_off_x / lap_off_y
// (grid.h:21-23), padded FP64 per-cell arrays; d == nullptr: uniform operator.
struct VarLap {
    const double* d = nullptr;
    const double* ox = nullptr;
    const double* oy = nullptr;
};

// The variable-coefficient operator (pcg.cpp:36-46): the same neighbour rule
// as the uniform one, one off-diagonal coefficient per axis,
// Av = diag v + off_x (vL + vR) + off_y (vB + vT), in that order.
__device__ __forceinline__ double apply_A_var(int i, int j, long id, uint8_t nb, double vc, double vl, double vr,
                                              double vb, double vt, const Geo& g, const Consts<double>& c,
                                              const VarLap& L) {
    double vL = (nb & NB_XM) ? vl : ((c.any_dirichlet && i == 1 && c.wall[0]) ? -vc : vc);
    double vR = (nb & NB_XP) ? vr : ((c.any_dirichlet && i == g.nx && c.wall[1]) ? -vc : vc);
    double vB = (nb & NB_YM) ? vb : ((c.any_dirichlet && j == 1 && c.wall[2]) ? -vc : vc);
    double vT = (nb & NB_YP) ? vt : ((c.any_dirichlet && j == g.ny && c.wall[3]) ? -vc : vc);
    return L.d[id] * vc + L.ox[id] * (vL + vR) + L.oy[id] * (vB + vT);
}

// Operator on a 4-cell vector at c0 (any fluid cell in it is interior, so
// every neighbour index below is inside the padded array).  VAR (2D only):
// the variable-coefficient operator.
template <int D, typename T, bool VAR = false>
__device__ __forceinline__ void apply_A4(const T* __restrict__ p, long c0, uint32_t w, const vec4<T>& pc,
                                         const Geo& g, const Consts<double>& c, double (&out)[4],
                                         const VarLap& L = VarLap{}) {
    const vec4<T> pb = *(const vec4u<T>*)(p + c0 - g.sx);
    const vec4<T> pt = *(const vec4u<T>*)(p + c0 + g.sx);
    vec4<T> pf = (T)0, pk = (T)0;
    if (D == 3) { pf = *(const vec4u<T>*)(p + c0 - g.sy); pk = *(const vec4u<T>*)(p + c0 + g.sy); }
    const double left = (double)p[c0 - 1], right = (double)p[c0 + 4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        uint8_t nb = nbyte(w, q);
        double vl = (q == 0) ? left : (double)pc[q == 0 ? 0 : q - 1];
        double vr = (q == 3) ? right : (double)pc[q == 3 ? 3 : q + 1];
        if constexpr (VAR && D == 2) {
            int i = 0, j = 0, k = 0;
            if (c.any_dirichlet) coords<D>(g, c0 + q, i, j, k);
            out[q] = (nb & NB_FLUID) ? apply_A_var(i, j, c0 + q, nb, (double)pc[q], vl, vr, (double)pb[q],
                                                   (double)pt[q], g, c, L)
                                     : 0.0;
        } else {
            out[q] = (nb & NB_FLUID) ? apply_A<D>(c0 + q, nb, (double)pc[q], vl, vr, (double)pb[q], (double)pt[q],
                                                  (double)pf[q], (double)pk[q], g, c)
                                     : 0.0;
        }
    }
}

// Ap = A p; pAp over fluid; alpha = rsold / pAp.  Non-fluid cells get Ap = 0.
template <int D, typename T, bool VAR = false>
__device__ __forceinline__ void d_matvec_dot(const T* __restrict__ p, T* __restrict__ Ap, const uint8_t* __restrict__ nbr, Geo g,
                             Consts<double> c, long N4, double* partials, Ctrl* ctrl, int fv_b, int fv_n,
                             VarLap L = VarLap{}) {
    if (ctrl->done) return;
    double v[1] = {0.0};
    FLAT_LOOP_XCD(N4) {
        long c0 = 4 * v_;
        uint32_t w = *(const uint32_t*)(nbr + c0);
        vec4<T> out = (T)0;
        if (w & 0x01010101u) {
            vec4<T> pc = *(const vec4<T>*)(p + c0);
            double o[4];
            apply_A4<D, T, VAR>(p, c0, w, pc, g, c, o, L);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                if (!(nbyte(w, q) & NB_FLUID)) continue;
                T oq = (T)o[q];
                out[q] = oq;
                v[0] += (double)pc[q] * (double)oq;
            }
        }
        *(vec4<T>*)(Ap + c0) = out;
    }
    block_store<1>(v, partials);
}
template <int D, typename T, bool VAR = false>
__global__ void k_matvec_dot(const T* __restrict__ p, T* __restrict__ Ap, const uint8_t* __restrict__ nbr, Geo g,
                             Consts<double> c, long N4, double* partials, Ctrl* ctrl, VarLap L) {
    d_matvec_dot<D, T, VAR>(p, Ap, nbr, g, c, N4, partials, ctrl, blockIdx.x, gridDim.x, L);
}
__device__ __forceinline__ void d_fin_matvec(const double* partials, int nbt, int base, int nb, Ctrl* ctrl) {
    if (ctrl->done) { if (threadIdx.x == 0) ctrl->xalpha = 0.0; return; }
    double tot[1];
    reduce_partials_at<1>(partials, nbt, base, nb, tot);
    if (threadIdx.x == 0) {
        double pAp = tot[0];
        ctrl->pAp = pAp;
        // alpha = 0 when the solve stops here: no x update is pending
        if (!isfinite(pAp)) { ctrl->done = 1; ctrl->status = ST_NONFINITE; ctrl->iter += 1; ctrl->alpha = 0.0; }
        else if (pAp < 1e-15) { ctrl->done = 1; ctrl->status = ST_BREAKDOWN; ctrl->iter += 1; ctrl->alpha = 0.0; }
        else ctrl->alpha = ctrl->rsold / pAp;
        ctrl->xalpha = ctrl->alpha;
    }
}
__global__ void k_fin_matvec(const double* partials, int nb, Ctrl* ctrl) {
    d_fin_matvec(partials, nb, 0, nb, ctrl);
}

// Warm start: r -= A x0 (after the rhs projection); r.r recomputed.
template <int D, typename T, bool VAR = false>
__device__ __forceinline__ void d_init_warm(T* __restrict__ r, const T* __restrict__ x0, const uint8_t* __restrict__ nbr, Geo g,
                            Consts<double> c, long N4, double* partials, Ctrl* ctrl, int fv_b, int fv_n,
                            VarLap L = VarLap{}) {
    double v[1] = {0.0};
    FLAT_LOOP(N4) {
        long c0 = 4 * v_;
        uint32_t w = *(const uint32_t*)(nbr + c0);
        if (!(w & 0x01010101u)) continue;
        vec4<T> pc = *(const vec4<T>*)(x0 + c0);
        vec4<T> rv = *(const vec4<T>*)(r + c0);
        double o[4];
        apply_A4<D, T, VAR>(x0, c0, w, pc, g, c, o, L);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (!(nbyte(w, q) & NB_FLUID)) continue;
            T val = (T)((double)rv[q] - o[q]);
            rv[q] = val;
            v[0] += (double)val * (double)val;
        }
        *(vec4<T>*)(r + c0) = rv;
    }
    block_store<1>(v, partials);
}
template <int D, typename T, bool VAR = false>
__global__ void k_init_warm(T* __restrict__ r, const T* __restrict__ x0, const uint8_t* __restrict__ nbr, Geo g,
                            Consts<double> c, long N4, double* partials, Ctrl* ctrl, VarLap L) {
    d_init_warm<D, T, VAR>(r, x0, nbr, g, c, N4, partials, ctrl, blockIdx.x, gridDim.x, L);
}
__device__ __forceinline__ void d_fin_init_warm(const double* partials, int nbt, int base, int nb, Ctrl* ctrl) {
    double tot[1];
    reduce_partials_at<1>(partials, nbt, base, nb, tot);
    if (threadIdx.x == 0) {
        ctrl->rsnew = tot[0];
        // the stopping threshold stays relative to the cold-start residual
        if (ctrl->tol_mode == 0 && sqrt(tot[0]) < ctrl->tol) { ctrl->done = 1; ctrl->status = ST_CONVERGED; }
        if (ctrl->tol_mode == 1 && tot[0] < ctrl->tolsq) { ctrl->done = 1; ctrl->status = ST_CONVERGED; }
    }
}
__global__ void k_fin_init_warm(const double* partials, int nb, Ctrl* ctrl) {
    d_fin_init_warm(partials, nb, 0, nb, ctrl);
}

// x += alpha p; r -= alpha Ap (fluid cells); r.r; convergence / iteration cap.
// FLEX: also (r_new, z_old - mean_old) for the Polak-Ribiere beta.
template <typename T, bool FLEX>
__device__ __forceinline__ void d_update(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ Ap,
                         const T* __restrict__ z, const uint8_t* __restrict__ nbr, long N4, double* partials,
                         Ctrl* ctrl, int fv_b, int fv_n) {
    if (ctrl->done) return;
    const double alpha = ctrl->alpha;
    double v[3] = {0.0, 0.0, 0.0};
    FLAT_LOOP(N4) {
        long c = 4 * v_;
        uint32_t w = *(const uint32_t*)(nbr + c);
        if (!(w & 0x01010101u)) continue;
        vec4<T> xv = *(const vec4<T>*)(x + c), rv = *(const vec4<T>*)(r + c);
        vec4<T> pv = *(const vec4<T>*)(p + c), av = *(const vec4<T>*)(Ap + c);
        vec4<T> zv;
        if (FLEX) zv = *(const vec4<T>*)(z + c);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (!(nbyte(w, q) & NB_FLUID)) continue;
            xv[q] = (T)((double)xv[q] + alpha * (double)pv[q]);
            T rn = (T)((double)rv[q] - alpha * (double)av[q]);
            rv[q] = rn;
            v[0] += (double)rn * (double)rn;
            if (FLEX) { v[1] += (double)rn * (double)zv[q]; v[2] += (double)rn; }
        }
        *(vec4<T>*)(x + c) = xv;
        *(vec4<T>*)(r + c) = rv;
    }
    if (FLEX) block_store<3>(v, partials);
    else { double v1[1] = {v[0]}; block_store<1>(v1, partials); }
}
template <typename T, bool FLEX>
__global__ void k_update(T* __restrict__ x, T* __restrict__ r, const T* __restrict__ p, const T* __restrict__ Ap,
                         const T* __restrict__ z, const uint8_t* __restrict__ nbr, long N4, double* partials,
                         Ctrl* ctrl) {
    d_update<T, FLEX>(x, r, p, Ap, z, nbr, N4, partials, ctrl, blockIdx.x, gridDim.x);
}
template <bool FLEX>
__device__ __forceinline__ void d_fin_update(const double* partials, int nbt, int base, int nb, Ctrl* ctrl) {
    if (ctrl->done) return;
    double tot[3] = {0.0, 0.0, 0.0};
    if (FLEX) reduce_partials_at<3>(partials, nbt, base, nb, tot);
    else { double t1[1]; reduce_partials_at<1>(partials, nbt, base, nb, t1); tot[0] = t1[0]; }
    if (threadIdx.x == 0) {
        double rs = tot[0];
        ctrl->rsnew = rs;
        if (FLEX) ctrl->rz_cross = tot[1] - ctrl->mean_z * tot[2];
        if (ctrl->hist) ctrl->hist[ctrl->iter] = rs;
        ctrl->iter += 1;
        bool conv = (ctrl->tol_mode == 1) ? (rs < ctrl->tolsq) : (sqrt(rs) < ctrl->tol);
        if (conv) { ctrl->done = 1; ctrl->status = ST_CONVERGED; }
        else if (!isfinite(rs)) { ctrl->done = 1; ctrl->status = ST_NONFINITE; }
        else if (ctrl->iter >= ctrl->max_iter) { ctrl->done = 1; ctrl->status = ST_MAXITER; }
    }
}
template <bool FLEX>
__global__ void k_fin_update(const double* partials, int nb, Ctrl* ctrl) {
    d_fin_update<FLEX>(partials, nb, 0, nb, ctrl);
}

// Shared tail of the z reductions: mean of z, (r, z - mean), beta, rsold.
// beta = (r_new, z_new)/rsold (pcg_3d.cpp:115), or with ctrl->flexible the
// Polak-Ribiere form (r_new, z_new - z_old)/rsold, which tolerates a
// preconditioner that is not exactly symmetric (FP32/BF16 levels).
__device__ __forceinline__ void finish_zsums(Ctrl* ctrl, double sz, double srz, double sr) {
    double mean = (ctrl->project && ctrl->count > 0) ? sz / ctrl->count : 0.0;
    double rz = srz - mean * sr;
    ctrl->sum_z = sz; ctrl->sum_rz = srz; ctrl->sum_r = sr;
    ctrl->mean_z = mean;
    ctrl->rz = rz;
    if (ctrl->first) { ctrl->rsold = rz; ctrl->beta = 0.0; ctrl->first = 0; }
    else {
        double num = ctrl->flexible ? rz - ctrl->rz_cross : rz;
        ctrl->beta = num / ctrl->rsold;
        ctrl->rsold = rz;
    }
    if (!isfinite(rz)) { ctrl->done = 1; ctrl->status = ST_NONFINITE; }
}

// sum z, sum r.z, sum r over fluid cells (per-pass V-cycle path).
template <typename T>
__device__ __forceinline__ void d_zsums(const T* __restrict__ z, const T* __restrict__ r, const uint8_t* __restrict__ nbr, long N4,
                        double* partials, Ctrl* ctrl, int fv_b, int fv_n) {
    if (ctrl->done) return;
    double v[3] = {0.0, 0.0, 0.0};
    FLAT_LOOP(N4) {
        long c = 4 * v_;
        uint32_t w = *(const uint32_t*)(nbr + c);
        if (!(w & 0x01010101u)) continue;
        vec4<T> zv = *(const vec4<T>*)(z + c), rv = *(const vec4<T>*)(r + c);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            if (!(nbyte(w, q) & NB_FLUID)) continue;
            double zz = (double)zv[q], rr = (double)rv[q];
            v[0] += zz; v[1] += rr * zz; v[2] += rr;
        }
    }
    block_store<3>(v, partials);
}
template <typename T>
__global__ void k_zsums(const T* __restrict__ z, const T* __restrict__ r, const uint8_t* __restrict__ nbr, long N4,
                        double* partials, Ctrl* ctrl) {
    d_zsums<T>(z, r, nbr, N4, partials, ctrl, blockIdx.x, gridDim.x);
}
// finish of the z sums (k_zsums or the fused level-0 up kernel)
__device__ __forceinline__ void d_fin_zsums(const double* partials, int nbt, int base, int nb, Ctrl* ctrl) {
    if (ctrl->done) return;
    double tot[3];
    reduce_partials_at<3>(partials, nbt, base, nb, tot);
    if (threadIdx.x == 0) finish_zsums(ctrl, tot[0], tot[1], tot[2]);
}
__global__ void k_fin_zsums(const double* partials, int nb, Ctrl* ctrl) {
    d_fin_zsums(partials, nb, 0, nb, ctrl);
}

// x += alpha p on fluid cells, alpha = *xalpha (a Ctrl::xa slot; the x half of k_update,
// launched on its own; it does not read the done flag, which the r half may
// set while it runs).  With NB = false the mask is not read: for a p that is
// exactly 0 on every non-fluid cell and ghost (the operator pass's p), the
// update leaves those cells of x unchanged by itself.
template <typename T, bool NB = true, bool NT = false>
__global__ void k_xupdate(T* __restrict__ x, const T* __restrict__ p, const uint8_t* __restrict__ nbr, long N4,
                          const double* xalpha) {
    const double alpha = *xalpha;
    if (alpha == 0.0) return;
    const int fv_b = blockIdx.x, fv_n = gridDim.x;
    FLAT_LOOP(N4) {
        long c = 4 * v_;
        uint32_t w = 0x01010101u;
        if constexpr (NB) {
            w = *(const uint32_t*)(nbr + c);
            if (!(w & 0x01010101u)) continue;
        }
        vec4<T> xv, pv;
        if constexpr (NT) {   // streaming: no reuse, keep the caches for the kernels beside it
            xv = __builtin_nontemporal_load((const vec4<T>*)(x + c));
            pv = __builtin_nontemporal_load((const vec4<T>*)(p + c));
        } else {
            xv = *(const vec4<T>*)(x + c); pv = *(const vec4<T>*)(p + c);
        }
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (!NB || (nbyte(w, q) & NB_FLUID)) xv[q] = (T)((double)xv[q] + alpha * (double)pv[q]);
        if constexpr (NT) __builtin_nontemporal_store(xv, (vec4<T>*)(x + c));
        else *(vec4<T>*)(x + c) = xv;
    }
}

// p = (z - mean) + beta p on fluid cells.
template <typename T>
__device__ __forceinline__ void d_pupdate(T* __restrict__ p, const T* __restrict__ z, const uint8_t* __restrict__ nbr, long N4,
                          Ctrl* ctrl, int fv_b, int fv_n) {
    if (ctrl->done) return;
    const double mean = ctrl->mean_z, beta = ctrl->beta;
    FLAT_LOOP(N4) {
        long c = 4 * v_;
        uint32_t w = *(const uint32_t*)(nbr + c);
        if (!(w & 0x01010101u)) continue;
        vec4<T> zv = *(const vec4<T>*)(z + c), pv = *(const vec4<T>*)(p + c);
#pragma unroll
        for (int q = 0; q < 4; q++)
            if (nbyte(w, q) & NB_FLUID) pv[q] = (T)(((double)zv[q] - mean) + beta * (double)pv[q]);
        *(vec4<T>*)(p + c) = pv;
    }
}
template <typename T>
__global__ void k_pupdate(T* __restrict__ p, const T* __restrict__ z, const uint8_t* __restrict__ nbr, long N4,
                          Ctrl* ctrl) {
    d_pupdate<T>(p, z, nbr, N4, ctrl, blockIdx.x, gridDim.x);
}

// p_out += x on fluid cells (cold start) or p_out = x (warm start).
template <typename T>
__global__ void k_finalize(const T* __restrict__ x, double* __restrict__ pout, const uint8_t* __restrict__ nbr,
                           long N, int assign) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x, s = (long)gridDim.x * blockDim.x;
    for (; t < N; t += s)
        if (nbr[t] & NB_FLUID) pout[t] = assign ? (double)x[t] : pout[t] + (double)x[t];
}

// Cast FP64 -> T on fluid cells (warm-start x0 for the FP32 path).
template <typename T>
__global__ void k_cast_in(const double* __restrict__ src, T* __restrict__ dst, const uint8_t* __restrict__ nbr, long N) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x, s = (long)gridDim.x * blockDim.x;
    for (; t < N; t += s) dst[t] = (nbr[t] & NB_FLUID) ? (T)src[t] : T(0);
}

}  // namespace uaamg
