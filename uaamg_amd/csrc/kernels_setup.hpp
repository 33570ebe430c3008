// kernels_setup.hpp -- hierarchy set-up on device: level-0 neighbour bytes,
// coarse solid masks, Galerkin coarse couplings.
//
// Reference behaviour (paths relative to the reference root):
//   restrictSolid   src/solver/preconditioner/uaamg_preconditioner.cpp:43-53 (2D, >= 2 of 4)
//                   src/solver/preconditioner/uaamg_preconditioner_3d.cpp:34-45 (3D, >= 4 of 8)
//   setupFineCoeffs uaamg_preconditioner_3d.cpp:50-76
//   galerkinCoarsen uaamg_preconditioner_3d.cpp:81-116 (coarse coupling = sum of the
//                   2 (2D) / 4 (3D) fine couplings across the shared face; only for
//                   fluid coarse cells, even when the neighbour is coarse-solid)
#pragma once
#include "common.hpp"

namespace uaamg {

// Interior loop helper: block (64,4,1); grid.z covers k in chunks of kz planes.
#define UAAMG_CELL_LOOP(g, D, kz)                                                   \
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;                        \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;                        \
    const bool inb = (i <= (g).nx && j <= (g).ny);                                  \
    const int k0 = (D == 3) ? (int)blockIdx.z * (kz) + 1 : 0;                       \
    const int k1 = (D == 3) ? min(k0 + (kz) - 1, (g).nz) : 0;                       \
    for (int k = k0; inb && k <= k1; k++)

// Level-0 neighbour bytes from the solid mask, and the fluid-cell count.
template <int D>
__global__ void k_nbr0(const uint8_t* __restrict__ solid, uint8_t* __restrict__ nbr, Geo g, int kz,
                       double* partials) {
    double cnt[1] = {0.0};
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const bool inb = (i <= g.nx && j <= g.ny);
    const int k0 = (D == 3) ? (int)blockIdx.z * kz + 1 : 0;
    const int k1 = (D == 3) ? min(k0 + kz - 1, g.nz) : 0;
    // the seven bytes of a cell are read unconditionally (an interior cell's
    // neighbours are all inside the padded grid) and the planes of a thread
    // are unrolled, so the loads of several cells are in flight together:
    // reading the neighbours only behind the cell's own byte made every cell
    // two dependent memory round trips (72 us at 256^3)
#pragma unroll 4
    for (int k = k0; k <= k1; k++) {
        if (!inb) break;
        const long id = gidx(g, i, j, k);
        const uint8_t s0 = solid[id], sxp = solid[id + 1], sxm = solid[id - 1];
        const uint8_t syp = solid[id + g.sx], sym = solid[id - g.sx];
        uint8_t szp = 1, szm = 1;
        if constexpr (D == 3) { szp = solid[id + g.sy]; szm = solid[id - g.sy]; }
        uint8_t b = 0;
        if (!s0) {
            b = NB_FLUID;
            if (i < g.nx && !sxp) b |= NB_XP;
            if (i > 1 && !sxm) b |= NB_XM;
            if (j < g.ny && !syp) b |= NB_YP;
            if (j > 1 && !sym) b |= NB_YM;
            if (D == 3) {
                if (k < g.nz && !szp) b |= NB_ZP;
                if (k > 1 && !szm) b |= NB_ZM;
            }
            cnt[0] += 1.0;
        }
        nbr[id] = b;
    }
    block_store<1>(cnt, partials);
}

// fluid-cell count from the k_nbr0 partials (one workgroup)
__global__ void k_fin_count(const double* partials, int nb, double* count_out) {
    double tot[1];
    reduce_partials<1>(partials, nb, tot);
    if (threadIdx.x == 0) *count_out = tot[0];
}

// Coarse solid mask: solid iff >= 4 of 8 (3D) / >= 2 of 4 (2D) children are
// solid.  fine_is_nbr: the fine level is level 0 (solid = !(nbr & NB_FLUID)).
template <int D>
__global__ void k_coarsen_mask(const uint8_t* __restrict__ fine, uint8_t* __restrict__ coarse,
                               Geo fg, Geo cg, int kz, int fine_is_nbr, int all_solid) {
    UAAMG_CELL_LOOP(cg, D, kz) {
        int i_f = 2 * i - 1, j_f = 2 * j - 1, k_f = (D == 3) ? 2 * k - 1 : 0;
        int sc = 0;
#pragma unroll
        for (int di = 0; di < 2; di++)
#pragma unroll
            for (int dj = 0; dj < 2; dj++)
#pragma unroll
                for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++) {
                    uint8_t v = fine[gidx(fg, i_f + di, j_f + dj, k_f + dk)];
                    sc += fine_is_nbr ? !(v & NB_FLUID) : (v != 0);
                }
        // the reference's threshold (restrictSolid), or every child solid
        const int thr = all_solid ? (D == 3 ? 8 : 4) : (D == 3 ? 4 : 2);
        coarse[gidx(cg, i, j, k)] = (sc >= thr) ? 1 : 0;
    }
}

// Dirichlet diagonal excess of a level-0 cell (zero for the reference's Neumann walls).
template <int D, typename CT>
__device__ __forceinline__ CT wall_excess0(const Consts<CT>& c, const Geo& g, int i, int j, int k) {
    CT e = CT(0);
    if (c.wall[0] && i == 1) e += c.ex2[0];
    if (c.wall[1] && i == g.nx) e += c.ex2[1];
    if (c.wall[2] && j == 1) e += c.ex2[2];
    if (c.wall[3] && j == g.ny) e += c.ex2[3];
    if (D == 3) {
        if (c.wall[4] && k == 1) e += c.ex2[4];
        if (c.wall[5] && k == g.nz) e += c.ex2[5];
    }
    return e;
}

// Fine-level couplings in FP64 for the Galerkin sums.
struct FineL0 {
    const uint8_t* nbr; Consts<double> c; Geo g;
    __device__ double cx(long id) const { return (nbr[id] & NB_XP) ? c.idx2 : 0.0; }
    __device__ double cy(long id) const { return (nbr[id] & NB_YP) ? c.idy2 : 0.0; }
    __device__ double cz(long id) const { return (nbr[id] & NB_ZP) ? c.idz2 : 0.0; }
    template <int D> __device__ double ex(long id, int i, int j, int k) const {
        if (!c.any_dirichlet || !(nbr[id] & NB_FLUID)) return 0.0;
        return wall_excess0<D, double>(c, g, i, j, k);
    }
};
template <typename C4, typename S>
struct FineLc {
    const C4* coef;
    const S* exc;
    __device__ double cx(long id) const { return (double)coef[id].x; }
    __device__ double cy(long id) const { return (double)coef[id].y; }
    __device__ double cz(long id) const { return (double)coef[id].z; }
    template <int D> __device__ double ex(long id, int, int, int) const { return (double)exc[id]; }
};

// Galerkin coarse couplings (cx, cy, cz) and diagonal excess for every
// interior coarse cell; solid coarse cells get zeros.  Sums run in FP64 in
// the reference's order and are rounded once to the level's storage type.
// The diagonal (record .w) is filled by k_level_diag once all couplings exist.
template <int D, typename FineSrc, typename C4, typename S>
__global__ void k_galerkin(FineSrc F, Geo fg, const uint8_t* __restrict__ cmask, C4* __restrict__ ccoef,
                           S* __restrict__ cex, Geo cg, int kz) {
    UAAMG_CELL_LOOP(cg, D, kz) {
        long cid = gidx(cg, i, j, k);
        C4 out;
        out.x = 0; out.y = 0; out.z = 0; out.w = 0;
        double e = 0;
        if (!cmask[cid]) {
            int i_f = 2 * i - 1, j_f = 2 * j - 1;
            double sx = 0, sy = 0, sz = 0;
            if (D == 3) {
                int k_f = 2 * k - 1;
                for (int dj = 0; dj < 2; dj++) for (int dk = 0; dk < 2; dk++) sx += F.cx(gidx(fg, i_f + 1, j_f + dj, k_f + dk));
                for (int di = 0; di < 2; di++) for (int dk = 0; dk < 2; dk++) sy += F.cy(gidx(fg, i_f + di, j_f + 1, k_f + dk));
                for (int di = 0; di < 2; di++) for (int dj = 0; dj < 2; dj++) sz += F.cz(gidx(fg, i_f + di, j_f + dj, k_f + 1));
                for (int di = 0; di < 2; di++) for (int dj = 0; dj < 2; dj++) for (int dk = 0; dk < 2; dk++)
                    e += F.template ex<D>(gidx(fg, i_f + di, j_f + dj, k_f + dk), i_f + di, j_f + dj, k_f + dk);
            } else {
                for (int dj = 0; dj < 2; dj++) sx += F.cx(gidx(fg, i_f + 1, j_f + dj, 0));
                for (int di = 0; di < 2; di++) sy += F.cy(gidx(fg, i_f + di, j_f + 1, 0));
                for (int di = 0; di < 2; di++) for (int dj = 0; dj < 2; dj++)
                    e += F.template ex<D>(gidx(fg, i_f + di, j_f + dj, 0), i_f + di, j_f + dj, 0);
            }
            out.x = sx; out.y = sy; out.z = sz;
        }
        ccoef[cid] = out;
        cex[cid] = (S)e;
    }
}

// Record diagonal: .w = cx + cx(-x) + cy + cy(-y) [+ cz + cz(-z)] (+ excess)
// in FP64 in galerkinCoarsen's order (uaamg_preconditioner_3d.cpp:107-115),
// or -1 for a solid cell (the V-cycle kernels test .w < 0 for "solid").
template <int D, typename C4, typename S>
__global__ void k_level_diag(C4* __restrict__ coef, const S* __restrict__ cex, const uint8_t* __restrict__ mask,
                             Geo g, int kz) {
    UAAMG_CELL_LOOP(g, D, kz) {
        long id = gidx(g, i, j, k);
        double d;
        if (mask[id]) d = -1.0;
        else {
            d = (double)coef[id].x + (double)coef[id - 1].x + (double)coef[id].y + (double)coef[id - g.sx].y;
            if (D == 3) d = d + (double)coef[id].z + (double)coef[id - g.sy].z;
            double e = (double)cex[id];
            if (e != 0.0) d += e;
        }
        using RE = typename std::remove_cv<typename std::remove_reference<decltype(coef[id].w)>::type>::type;
        if constexpr (sizeof(RE) == 2) {
            // 16-bit records: round the diagonal up, so that it never drops below
            // the sum of the (rounded) couplings -- the coarse operator stays
            // diagonally dominant, hence positive definite, and the V-cycle a
            // valid preconditioner
            RE dd = (RE)d;
            if (d > 0.0 && (double)dd < d) {
                uint16_t u = __builtin_bit_cast(uint16_t, dd);
                dd = __builtin_bit_cast(RE, (uint16_t)(u + 1));
            }
            coef[id].w = dd;
        } else {
            coef[id].w = d;
        }
    }
}

template <typename V>
__global__ void k_fill(V* __restrict__ x, long n, V v, const int* skip) {
    if (skip && *skip) return;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) x[t] = v;
}

// z = r: the identity preconditioner of factory key "cg"
// (identity_preconditioner.h:19, IdentityPreconditioner::apply)
template <typename V>
__global__ void k_copy(const V* __restrict__ src, V* __restrict__ dst, long n, const int* skip) {
    if (skip && *skip) return;
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    long stride = (long)gridDim.x * blockDim.x;
    for (; t < n; t += stride) dst[t] = src[t];
}

}  // namespace uaamg
