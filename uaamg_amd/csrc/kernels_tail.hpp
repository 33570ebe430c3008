// kernels_tail.hpp -- the small coarse levels of the V-cycle in one
// workgroup.
//
// Below a few thousand cells a level is latency-bound: each separate launch
// costs more than its arithmetic.  k_tail runs the whole lower part of the
// V(1,1) cycle -- forward sweep, residual restriction, the coarsest
// 10 x (forward + reverse) sweeps, prolongation and reverse sweep for every
// level from `first` down to the coarsest -- in one 1024-thread workgroup
// with x, b and the residual of every tail level in LDS and the stencil
// records of the cells a thread owns in registers.  Arithmetic and order are
// those of the per-pass kernels (k_smooth, k_restrict, k_prolong), i.e. of
// the reference smooth / restrictResidual / prolongateAdd
// (uaamg_preconditioner_3d.cpp:123-196); the FP32 form divides by the
// hardware reciprocal, as the march kernels do.
//
// A thread owns pairs of x-adjacent cells (i odd, i+1): one cell of each
// colour, so every RBGS pass keeps all lanes busy.
#pragma once
#include "kernels_vcycle.hpp"

namespace uaamg {

constexpr int TAIL_THREADS = 1024;
constexpr int TAIL_MAXL = 5;      // tail levels (3D: 16^3..4^3 is 3; 2D: 64^2..4^2 is 5)
constexpr int TAIL_MAXP = 2;      // pairs per thread on the first tail level
constexpr int TAIL_LDS_BYTES = 96 * 1024;   // x, b of every tail level + one residual array

template <typename C4>
struct TailArgs {
    int nl;                        // levels in the tail; the last one is the coarsest
    Geo g[TAIL_MAXL];
    const C4* coef[TAIL_MAXL];
    int xo[TAIL_MAXL], bo[TAIL_MAXL];   // LDS offsets (elements) of x and b per level
    int ro;                        // LDS offset of the residual scratch
    int npass;                     // coarsest passes (2 * sweeps + 1)
};

// Stencils of one owned pair of cells, in the reference channel order.
template <typename CT>
struct TailPair {
    CT d[2], xp[2], xm[2], yp[2], ym[2], zp[2], zm[2];
    int id0;          // padded index of the first cell
    int i0, j, k;
    int par0;         // colour of the first cell
    bool valid, ok1;  // pair exists; second cell inside the level
};

template <int D, typename CT, typename C4>
__device__ __forceinline__ void tail_load_pair(TailPair<CT>& P, const C4* __restrict__ coef, const Geo& g, int p) {
    const int npx = (g.nx + 1) >> 1;
    const int npairs = npx * g.ny * (D == 3 ? g.nz : 1);
    P.valid = p < npairs;
    if (!P.valid) { P.id0 = 0; P.ok1 = false; P.par0 = 0; P.i0 = P.j = P.k = 0; return; }
    const int ip = p % npx;
    P.i0 = 2 * ip + 1;
    P.j = (p / npx) % g.ny + 1;
    P.k = (D == 3) ? p / (npx * g.ny) + 1 : 0;
    P.ok1 = P.i0 + 1 <= g.nx;
    P.par0 = (P.i0 + P.j + P.k) & 1;
    P.id0 = (int)gidx(g, P.i0, P.j, P.k);
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const long id = P.id0 + q;
        if (q == 1 && !P.ok1) {
            P.d[q] = CT(-1); P.xp[q] = P.xm[q] = P.yp[q] = P.ym[q] = P.zp[q] = P.zm[q] = CT(0);
            continue;
        }
        Sten<CT> s = sten_lc<D, CT>(coef, g, id);
        P.d[q] = s.d; P.xp[q] = s.xp; P.xm[q] = s.xm; P.yp[q] = s.yp; P.ym[q] = s.ym; P.zp[q] = s.zp; P.zm[q] = s.zm;
    }
}

// (b + nb) / d: exact division in FP64 (the parity path); in FP32 the
// hardware reciprocal, as in the march kernels of the same levels, which
// shortens the dependent chain of every pass
template <typename CT>
__device__ __forceinline__ CT tdiv(CT num, CT d) {
    if constexpr (std::is_same<CT, double>::value) return num / d;
    else return num * __builtin_amdgcn_rcpf(d);
}

// One RBGS colour on the owned pairs of one level.
template <int D, typename CT, int NP>
__device__ __forceinline__ void tail_smooth(const TailPair<CT> (&P)[NP], CT* X, const CT* Bv, const Geo& g, int colour) {
#pragma unroll
    for (int v = 0; v < NP; v++) {
        const TailPair<CT>& p = P[v];
        if (!p.valid) continue;
        const int q = (p.par0 == colour) ? 0 : 1;
        if (q == 1 && !p.ok1) continue;
        const CT d = q ? p.d[1] : p.d[0];
        if (!(d >= CT(1e-30))) continue;           // solid (d < 0) or isolated
        const int id = p.id0 + q;
        const CT xp = q ? p.xp[1] : p.xp[0], xm = q ? p.xm[1] : p.xm[0];
        const CT yp = q ? p.yp[1] : p.yp[0], ym = q ? p.ym[1] : p.ym[0];
        CT nb = xp * X[id + 1] + xm * X[id - 1] + yp * X[id + g.sx] + ym * X[id - g.sx];
        if (D == 3) {
            const CT zp = q ? p.zp[1] : p.zp[0], zm = q ? p.zm[1] : p.zm[0];
            nb = nb + zp * X[id + g.sy] + zm * X[id - g.sy];
        }
        X[id] = tdiv<CT>(Bv[id] + nb, d);
    }
}

// b - A x on the owned fluid cells (0 elsewhere), formed as k_restrict forms it.
template <int D, typename CT, int NP>
__device__ __forceinline__ void tail_residual(const TailPair<CT> (&P)[NP], const CT* X, const CT* Bv, CT* Rs, const Geo& g) {
#pragma unroll
    for (int v = 0; v < NP; v++) {
        const TailPair<CT>& p = P[v];
        if (!p.valid) continue;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (q == 1 && !p.ok1) continue;
            const int id = p.id0 + q;
            CT res = CT(0);
            if (p.d[q] >= CT(0)) {
                double Ax = (double)p.d[q] * (double)X[id] - (double)p.xp[q] * (double)X[id + 1]
                          - (double)p.xm[q] * (double)X[id - 1] - (double)p.yp[q] * (double)X[id + g.sx]
                          - (double)p.ym[q] * (double)X[id - g.sx];
                if (D == 3) Ax = Ax - (double)p.zp[q] * (double)X[id + g.sy] - (double)p.zm[q] * (double)X[id - g.sy];
                res = (CT)((double)Bv[id] - Ax);
            }
            Rs[id] = res;
        }
    }
}

// x_f += 2 x_c on the owned fluid cells whose parent lies in the coarse level.
template <int D, typename CT, int NP>
__device__ __forceinline__ void tail_prolong(const TailPair<CT> (&P)[NP], CT* X, const CT* XC, const Geo& g, const Geo& cg) {
#pragma unroll
    for (int v = 0; v < NP; v++) {
        const TailPair<CT>& p = P[v];
        if (!p.valid) continue;
        const int jc = (p.j + 1) >> 1, kc = (D == 3) ? (p.k + 1) >> 1 : 0;
        const int ic = (p.i0 + 1) >> 1;              // both cells of the pair share the parent
        if (ic > cg.nx || jc > cg.ny || (D == 3 && kc > cg.nz)) continue;
        const CT xc = XC[gidx(cg, ic, jc, kc)];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (q == 1 && !p.ok1) continue;
            if (!(p.d[q] >= CT(0))) continue;
            X[p.id0 + q] = X[p.id0 + q] + CT(2) * xc;
        }
    }
    (void)g;
}

// b_c = sum of the children's residuals; x_c = 0 (every coarse interior cell).
template <int D, typename CT>
__device__ __forceinline__ void tail_restrict(const CT* Rs, CT* BC, CT* XC, const Geo& fg, const Geo& cg) {
    const int n = cg.nx * cg.ny * (D == 3 ? cg.nz : 1);
    for (int c = threadIdx.x; c < n; c += TAIL_THREADS) {
        const int i = c % cg.nx + 1, j = (c / cg.nx) % cg.ny + 1, k = (D == 3) ? c / (cg.nx * cg.ny) + 1 : 0;
        const int fi = 2 * i - 1, fj = 2 * j - 1, fk = (D == 3) ? 2 * k - 1 : 0;
        double sum = 0;
        for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++)
                for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++)
                    sum += (double)Rs[gidx(fg, fi + di, fj + dj, fk + dk)];
        const long cid = gidx(cg, i, j, k);
        BC[cid] = (CT)sum;
        XC[cid] = CT(0);
    }
}

template <int D, typename S, typename C4, int NLMAX>
__device__ __forceinline__ void d_tail(const TailArgs<C4>& A, const S* __restrict__ b_first, S* __restrict__ x_first,
                                       const int* skip) {
    static_assert(NLMAX >= 2 && NLMAX <= TAIL_MAXL, "tail depth");
    using CT = typename CTOf<S>::type;
    if (skip && *skip) return;
    __shared__ CT lds[TAIL_LDS_BYTES / sizeof(CT)];
    const int tid = threadIdx.x;
    const int nl = A.nl;   // 1 <= nl <= NLMAX (checked on the host); 1: the coarsest level only

    // registers: the first tail level has up to 2 pairs per thread, the others 1
    TailPair<CT> P0[TAIL_MAXP], P1[1], P2[1], P3[1], P4[1];
#pragma unroll
    for (int v = 0; v < TAIL_MAXP; v++) tail_load_pair<D, CT, C4>(P0[v], A.coef[0], A.g[0], tid + TAIL_THREADS * v);
    if (nl > 1) tail_load_pair<D, CT, C4>(P1[0], A.coef[1], A.g[1], tid); else P1[0].valid = false;
    if constexpr (NLMAX > 2) { if (nl > 2) tail_load_pair<D, CT, C4>(P2[0], A.coef[2], A.g[2], tid); }
    if constexpr (NLMAX > 3) { if (nl > 3) tail_load_pair<D, CT, C4>(P3[0], A.coef[3], A.g[3], tid); }
    if constexpr (NLMAX > 4) { if (nl > 4) tail_load_pair<D, CT, C4>(P4[0], A.coef[4], A.g[4], tid); }

    // x = 0 on every level (ghosts included); b of the first level from memory
    for (int l = 0; l < nl; l++) {
        CT* X = lds + A.xo[l];
        for (int e = tid; e < A.g[l].N; e += TAIL_THREADS) X[e] = CT(0);
    }
    {
        CT* B0 = lds + A.bo[0];
        for (int e = tid; e < A.g[0].N; e += TAIL_THREADS) B0[e] = (CT)b_first[e];
    }
    __syncthreads();

    auto level_down = [&](auto& P, int l) {
        CT* X = lds + A.xo[l];
        const CT* Bv = lds + A.bo[l];
        CT* Rs = lds + A.ro;
        tail_smooth<D>(P, X, Bv, A.g[l], 1);
        __syncthreads();
        tail_smooth<D>(P, X, Bv, A.g[l], 0);
        __syncthreads();
        tail_residual<D>(P, X, Bv, Rs, A.g[l]);
        __syncthreads();
        tail_restrict<D, CT>(Rs, lds + A.bo[l + 1], lds + A.xo[l + 1], A.g[l], A.g[l + 1]);
        __syncthreads();
    };
    auto level_coarsest = [&](auto& P, int l) {
        CT* X = lds + A.xo[l];
        const CT* Bv = lds + A.bo[l];
        const Geo& g = A.g[l];
        const int npairs = ((g.nx + 1) >> 1) * g.ny * (D == 3 ? g.nz : 1);
        if (npairs <= 64) {
            // every pair is owned by a lane of wave 0 (thread = pair): the
            // passes need no workgroup barrier, only wave order (LDS accesses
            // of one wave complete in program order; the fences keep the
            // compiler from moving them across passes)
            if (tid < 64) {
                for (int s = 0; s < A.npass; s++) {
                    tail_smooth<D>(P, X, Bv, g, (s & 1) ? 0 : 1);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
            __syncthreads();
            return;
        }
        for (int s = 0; s < A.npass; s++) {     // odd, even, odd, ... (forward/reverse pairs collapsed)
            tail_smooth<D>(P, X, Bv, g, (s & 1) ? 0 : 1);
            __syncthreads();
        }
    };
    auto level_up = [&](auto& P, int l) {
        CT* X = lds + A.xo[l];
        const CT* Bv = lds + A.bo[l];
        tail_prolong<D>(P, X, lds + A.xo[l + 1], A.g[l], A.g[l + 1]);
        __syncthreads();
        tail_smooth<D>(P, X, Bv, A.g[l], 0);
        __syncthreads();
        tail_smooth<D>(P, X, Bv, A.g[l], 1);
        __syncthreads();
    };
    // down
    if (nl > 1) level_down(P0, 0);
    if constexpr (NLMAX > 2) { if (nl > 2) level_down(P1, 1); }
    if constexpr (NLMAX > 3) { if (nl > 3) level_down(P2, 2); }
    if constexpr (NLMAX > 4) { if (nl > 4) level_down(P3, 3); }
    // coarsest
    if (nl == 1) level_coarsest(P0, 0);
    if (nl == 2) level_coarsest(P1, 1);
    if constexpr (NLMAX > 2) { if (nl == 3) level_coarsest(P2, 2); }
    if constexpr (NLMAX > 3) { if (nl == 4) level_coarsest(P3, 3); }
    if constexpr (NLMAX > 4) { if (nl == 5) level_coarsest(P4, 4); }
    // up
    if constexpr (NLMAX > 4) { if (nl > 4) level_up(P3, 3); }
    if constexpr (NLMAX > 3) { if (nl > 3) level_up(P2, 2); }
    if constexpr (NLMAX > 2) { if (nl > 2) level_up(P1, 1); }
    if (nl > 1) level_up(P0, 0);
    {
        const CT* X0 = lds + A.xo[0];
        for (int e = tid; e < A.g[0].N; e += TAIL_THREADS) x_first[e] = (S)X0[e];
    }
}

template <int D, typename S, typename C4, int NLMAX>
__global__ __launch_bounds__(TAIL_THREADS) void k_tail(TailArgs<C4> A, const S* __restrict__ b_first,
                                                        S* __restrict__ x_first, const int* skip) {
    d_tail<D, S, C4, NLMAX>(A, b_first, x_first, skip);
}

}  // namespace uaamg
