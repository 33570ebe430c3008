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
constexpr int TAIL_BMAX = 8;     // b elements per thread of the first tail level (padded cells <= 8192)
constexpr int TAIL_NSLOT = TAIL_MAXP + TAIL_MAXL - 1;   // record slots: TAIL_MAXP on the first level, 1 below

// Stencil records of the owned pairs, built at set-up in the compute type:
// 14 values per pair (d, xp, xm, yp, ym, zp, zm of both cells), NQ 16-byte
// quads per thread, stored [slot][quad][thread] so that every quad load of a
// wave is one contiguous 1 KiB read.
template <typename CT> struct TailRec;
template <> struct TailRec<float> { static constexpr int NQ = 4; using Q = float4; };
template <> struct TailRec<double> { static constexpr int NQ = 7; using Q = double2; };
template <typename CT> constexpr int tail_rec_per_q() { return 16 / (int)sizeof(CT); }
template <typename CT> constexpr long tail_rec_bytes(int nslots) { return (long)nslots * TailRec<CT>::NQ * TAIL_THREADS * 16; }
__host__ __device__ constexpr int tail_slot(int l, int v) { return l == 0 ? v : TAIL_MAXP + l - 1; }

template <typename C4>
struct TailArgs {
    int nl;                        // levels in the tail; the last one is the coarsest
    Geo g[TAIL_MAXL];
    const C4* coef[TAIL_MAXL];
    int xo[TAIL_MAXL], bo[TAIL_MAXL];   // LDS offsets (elements) of x and b per level
    int ro;                        // LDS offset of the residual scratch
    int npass;                     // coarsest passes (2 * sweeps + 1)
    const void* b;                 // b of the first tail level (level storage type)
    void* x;                       // x of the first tail level (written at the end)
    const int* skip;               // stop flag (nullptr: never skip)
    const void* rec;               // pair records (compute type), TAIL_NSLOT slots at most
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

// Record element type of a coefficient vector type.
template <typename C4> struct TailE { using E = float; };
template <> struct TailE<double4> { using E = double; };
template <> struct TailE<BF16x4> { using E = __bf16; };

// The record loads of one pair as they come from memory.  Every pair of
// every tail level issues its loads before any of them is used, so that the
// start of the kernel waits for one memory round trip instead of one per
// pair; an absent pair reads the first interior cell (in bounds, discarded)
// and the second cell of a pair at i = nx reads the ghost at nx + 1 (inside
// the padded level, discarded).
template <typename C4>
struct TailRaw {
    C4 c[2];
    typename TailE<C4>::E xm[2], ym[2], zm[2];
};

template <int D, typename C4>
__device__ __forceinline__ void tail_issue_pair(TailRaw<C4>& R, const C4* __restrict__ coef, const Geo& g, int p) {
    const int npx = (g.nx + 1) >> 1;
    const int npairs = npx * g.ny * (D == 3 ? g.nz : 1);
    const int pp = p < npairs ? p : 0;
    const int ip = pp % npx;
    const int j = (pp / npx) % g.ny + 1, k = (D == 3) ? pp / (npx * g.ny) + 1 : 0;
    const int id0 = (int)gidx(g, 2 * ip + 1, j, k);
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const int id = id0 + q;
        using E = typename TailE<C4>::E;
        const E* e = reinterpret_cast<const E*>(coef);   // {x, y, z, w} per record
        R.c[q] = gload(coef + id);
        R.xm[q] = gload(e + 4 * (id - 1));
        R.ym[q] = gload(e + 4 * (id - g.sx) + 1);
        if (D == 3) R.zm[q] = gload(e + 4 * (id - g.sy) + 2);
    }
}

// The stencils of pair p from its loaded records, as sten_lc forms them.
template <int D, typename CT, typename C4>
__device__ __forceinline__ void tail_load_pair(TailPair<CT>& P, const TailRaw<C4>& R, const Geo& g, int p) {
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
        if (q == 1 && !P.ok1) {
            P.d[q] = CT(-1); P.xp[q] = P.xm[q] = P.yp[q] = P.ym[q] = P.zp[q] = P.zm[q] = CT(0);
            continue;
        }
        const C4 c0 = R.c[q];
        CT d = (CT)c0.w;
        P.xp[q] = (CT)c0.x; P.yp[q] = (CT)c0.y;
        P.xm[q] = (CT)R.xm[q]; P.ym[q] = (CT)R.ym[q];
        if (D == 3) { P.zp[q] = (CT)c0.z; P.zm[q] = (CT)R.zm[q]; }
        else { P.zp[q] = CT(0); P.zm[q] = CT(0); }
        if (d >= CT(0)) d = snap_diag<C4, CT>(d, P.xp[q] + P.xm[q] + P.yp[q] + P.ym[q] + P.zp[q] + P.zm[q]);
        P.d[q] = d;
    }
}

// The 14 record values of one pair, as loaded.
template <typename CT>
struct TailRecV { CT v[TailRec<CT>::NQ * tail_rec_per_q<CT>()]; };

template <typename CT>
__device__ __forceinline__ void tail_issue_rec(TailRecV<CT>& R, const void* rec, int slot, int t) {
    using Q = typename TailRec<CT>::Q;
    constexpr int NQ = TailRec<CT>::NQ;
    const Q* base = static_cast<const Q*>(rec) + (long)slot * NQ * TAIL_THREADS + t;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const Q x = gload(base + q * TAIL_THREADS);
        if constexpr (sizeof(CT) == 4) { R.v[4 * q] = x.x; R.v[4 * q + 1] = x.y; R.v[4 * q + 2] = x.z; R.v[4 * q + 3] = x.w; }
        else { R.v[2 * q] = x.x; R.v[2 * q + 1] = x.y; }
    }
}

// Geometry of pair p (computed) and its stencils (from the record).
template <int D, typename CT>
__device__ __forceinline__ void tail_pair_from_rec(TailPair<CT>& P, const TailRecV<CT>& R, const Geo& g, int p) {
    const int npx = (g.nx + 1) >> 1;
    const int npairs = npx * g.ny * (D == 3 ? g.nz : 1);
    P.valid = p < npairs;
    const int pp = P.valid ? p : 0;
    const int ip = pp % npx;
    P.i0 = 2 * ip + 1;
    P.j = (pp / npx) % g.ny + 1;
    P.k = (D == 3) ? pp / (npx * g.ny) + 1 : 0;
    P.ok1 = P.valid && P.i0 + 1 <= g.nx;
    P.par0 = (P.i0 + P.j + P.k) & 1;
    P.id0 = (int)gidx(g, P.i0, P.j, P.k);
#pragma unroll
    for (int q = 0; q < 2; q++) {
        P.d[q] = R.v[q]; P.xp[q] = R.v[2 + q]; P.xm[q] = R.v[4 + q]; P.yp[q] = R.v[6 + q];
        P.ym[q] = R.v[8 + q]; P.zp[q] = R.v[10 + q]; P.zm[q] = R.v[12 + q];
    }
}

// Set-up: the records of every slot of the tail (one workgroup per slot).
template <int D, typename CT, typename C4>
__global__ __launch_bounds__(TAIL_THREADS) void k_tail_records(TailArgs<C4> A, CT* __restrict__ rec) {
    const int slot = blockIdx.x;
    const int l = slot < TAIL_MAXP ? 0 : slot - TAIL_MAXP + 1;
    const int v = slot < TAIL_MAXP ? slot : 0;
    if (l >= A.nl) return;
    const int t = threadIdx.x, p = t + TAIL_THREADS * v;
    TailRaw<C4> R;
    tail_issue_pair<D, C4>(R, A.coef[l], A.g[l], p);
    TailPair<CT> P;
    tail_load_pair<D, CT, C4>(P, R, A.g[l], p);
    constexpr int PER = tail_rec_per_q<CT>(), NQ = TailRec<CT>::NQ;
    CT f[NQ * PER];
#pragma unroll
    for (int e = 0; e < NQ * PER; e++) f[e] = CT(0);
    if (P.valid) {
#pragma unroll
        for (int q = 0; q < 2; q++) {
            f[q] = P.d[q]; f[2 + q] = P.xp[q]; f[4 + q] = P.xm[q]; f[6 + q] = P.yp[q];
            f[8 + q] = P.ym[q]; f[10 + q] = P.zp[q]; f[12 + q] = P.zm[q];
        }
    }
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int e = 0; e < PER; e++) rec[(((long)slot * NQ + q) * TAIL_THREADS + t) * PER + e] = f[q * PER + e];
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
__device__ __forceinline__ void d_tail(const TailArgs<C4>& Ag) {
    static_assert(NLMAX >= 2 && NLMAX <= TAIL_MAXL, "tail depth");
    using CT = typename CTOf<S>::type;
    __shared__ __align__(16) CT lds[TAIL_LDS_BYTES / sizeof(CT)];
    const int tid = threadIdx.x;
    // The arguments sit in memory no cache has seen yet (a fresh kernel
    // argument block, or the batch descriptor): scalar loads would fetch them
    // line by line as the code first touches them, one memory round trip
    // each.  One vector load per dword brings the block in one round trip.
    // The skip flag is read beside the first data loads and tested before
    // the first store to memory, not waited for up front.
    __shared__ TailArgs<C4> A;
    {
        constexpr int NW = (int)(sizeof(TailArgs<C4>) / 4);
        static_assert(NW <= TAIL_THREADS && sizeof(TailArgs<C4>) % 4 == 0, "tail arguments");
        if (tid < NW) reinterpret_cast<uint32_t*>(&A)[tid] = reinterpret_cast<const uint32_t*>(&Ag)[tid];
    }
    __syncthreads();
    const S* __restrict__ b_first = static_cast<const S*>(A.b);
    S* __restrict__ x_first = static_cast<S*>(A.x);
    const int nl = A.nl;   // 1 <= nl <= NLMAX (checked on the host); 1: the coarsest level only

    // every global load of the kernel's start first: the records of the
    // owned pairs (the first tail level has up to 2 pairs per thread, the
    // others 1) and b of the first level (at most TAIL_BMAX per thread,
    // checked on the host)
    TailRecV<CT> R0[TAIL_MAXP], R1, R2, R3, R4;
#pragma unroll
    for (int v = 0; v < TAIL_MAXP; v++) tail_issue_rec<CT>(R0[v], A.rec, tail_slot(0, v), tid);
    if (nl > 1) tail_issue_rec<CT>(R1, A.rec, tail_slot(1, 0), tid);
    if constexpr (NLMAX > 2) { if (nl > 2) tail_issue_rec<CT>(R2, A.rec, tail_slot(2, 0), tid); }
    if constexpr (NLMAX > 3) { if (nl > 3) tail_issue_rec<CT>(R3, A.rec, tail_slot(3, 0), tid); }
    if constexpr (NLMAX > 4) { if (nl > 4) tail_issue_rec<CT>(R4, A.rec, tail_slot(4, 0), tid); }
    S bt[TAIL_BMAX];
    const int N0 = (int)A.g[0].N;
#pragma unroll
    for (int u = 0; u < TAIL_BMAX; u++) {
        const int e = tid + TAIL_THREADS * u;
        if (e < N0) bt[u] = gload(b_first + e);
    }
    // a vector load (the index lives in a VGPR), so that the flag is waited
    // for where it is tested, not at once by a scalar branch
    int skipped = 0;
    if (A.skip) {
        int z = 0;
        asm volatile("" : "+v"(z));
        skipped = gload(A.skip + z);
    }

    // x = 0 on every level (ghosts included) while the loads are in flight
    for (int l = 0; l < nl; l++) {
        CT* X = lds + A.xo[l];
        for (int e = tid; e < A.g[l].N; e += TAIL_THREADS) X[e] = CT(0);
    }
    TailPair<CT> P0[TAIL_MAXP], P1[1], P2[1], P3[1], P4[1];
#pragma unroll
    for (int v = 0; v < TAIL_MAXP; v++) tail_pair_from_rec<D, CT>(P0[v], R0[v], A.g[0], tid + TAIL_THREADS * v);
    if (nl > 1) tail_pair_from_rec<D, CT>(P1[0], R1, A.g[1], tid); else P1[0].valid = false;
    if constexpr (NLMAX > 2) { if (nl > 2) tail_pair_from_rec<D, CT>(P2[0], R2, A.g[2], tid); }
    if constexpr (NLMAX > 3) { if (nl > 3) tail_pair_from_rec<D, CT>(P3[0], R3, A.g[3], tid); }
    if constexpr (NLMAX > 4) { if (nl > 4) tail_pair_from_rec<D, CT>(P4[0], R4, A.g[4], tid); }
    {
        CT* B0 = lds + A.bo[0];
#pragma unroll
        for (int u = 0; u < TAIL_BMAX; u++) {
            const int e = tid + TAIL_THREADS * u;
            if (e < N0) B0[e] = (CT)bt[u];
        }
    }
    __syncthreads();
    if (skipped) return;   // uniform: every thread read the same flag

    auto level_down = [&](auto& P, int l) {
        // geometry and offsets copied to registers once: read through the
        // LDS copy of the arguments they would be fetched again after every
        // barrier
        const Geo g = A.g[l], gc = A.g[l + 1];
        CT* X = lds + A.xo[l];
        const CT* Bv = lds + A.bo[l];
        CT* Rs = lds + A.ro;
        CT* BC = lds + A.bo[l + 1];
        CT* XC = lds + A.xo[l + 1];
        tail_smooth<D>(P, X, Bv, g, 1);
        __syncthreads();
        tail_smooth<D>(P, X, Bv, g, 0);
        __syncthreads();
        tail_residual<D>(P, X, Bv, Rs, g);
        __syncthreads();
        tail_restrict<D, CT>(Rs, BC, XC, g, gc);
        __syncthreads();
    };
    auto level_coarsest = [&](auto& P, int l) {
        CT* X = lds + A.xo[l];
        const CT* Bv = lds + A.bo[l];
        const Geo g = A.g[l];
        const int npass = A.npass;
        const int npairs = ((g.nx + 1) >> 1) * g.ny * (D == 3 ? g.nz : 1);
        if (npairs <= 64) {
            // every pair is owned by a lane of wave 0 (thread = pair): the
            // passes need no workgroup barrier, only wave order (LDS accesses
            // of one wave complete in program order; the fences keep the
            // compiler from moving them across passes)
            if (tid < 64) {
                for (int s = 0; s < npass; s++) {
                    tail_smooth<D>(P, X, Bv, g, (s & 1) ? 0 : 1);
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                }
            }
            __syncthreads();
            return;
        }
        for (int s = 0; s < npass; s++) {     // odd, even, odd, ... (forward/reverse pairs collapsed)
            tail_smooth<D>(P, X, Bv, g, (s & 1) ? 0 : 1);
            __syncthreads();
        }
    };
    auto level_up = [&](auto& P, int l) {
        const Geo g = A.g[l], gc = A.g[l + 1];
        CT* X = lds + A.xo[l];
        const CT* Bv = lds + A.bo[l];
        const CT* XC = lds + A.xo[l + 1];
        tail_prolong<D>(P, X, XC, g, gc);
        __syncthreads();
        tail_smooth<D>(P, X, Bv, g, 0);
        __syncthreads();
        tail_smooth<D>(P, X, Bv, g, 1);
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
        for (int e = tid; e < A.g[0].N; e += TAIL_THREADS) gstore(x_first + e, (S)X0[e]);
    }
}

template <int D, typename S, typename C4, int NLMAX>
__global__ __launch_bounds__(TAIL_THREADS) void k_tail(TailArgs<C4> A) {
    d_tail<D, S, C4, NLMAX>(A);
}

}  // namespace uaamg
