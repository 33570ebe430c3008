// kernels_gmg.hpp -- the GMG and AMG preconditioners of the reference's
// other factory keys ("pcg" / "pcg_gmg": GMGPreconditioner; "pcg_amg":
// AMGPreconditioner), 2D.
//
// Both build the same 2x2 hierarchy with a rediscretised 5-point operator
// per level (h doubles per level), a coarse cell solid when at least two of
// its four children are, and V(2,2) cycles with 10 x (forward + reverse)
// sweeps on the coarsest level (gmg_preconditioner.cpp:13-194,
// amg_preconditioner.cpp:13-209).  The smoother updates one colour in place:
//   x += (b - diag x + lap) / d_eff,  lap = (xL + xR)/dx^2 + (xB + xT)/dy^2
// with a wall or solid neighbour taking the cell's own value and d_eff the
// diagonal without those faces.  Restriction is the mean of the fluid
// children's residuals; AMG writes the residual as b - (diag x - lap) and
// every coarse cell, GMG as (b - diag x) + lap and only fluid coarse cells.
// Prolongation adds the parent's x to the fluid children.  The arithmetic
// and its order are the reference's, so the FP64 form matches it bit for bit.
#pragma once
#include "common.hpp"

namespace uaamg {

// per-level constants of the rediscretised operator (host computes them as
// the reference does: GMG diag = 2 (1/dx^2 + 1/dy^2), AMG 2/dx^2 + 2/dy^2)
template <typename CT>
struct GmgC {
    CT idx2, idy2, diag;
};

// ---------------------------------------------------------------- fused passes
//
// One kernel per level on the way down (both forward sweeps, the residual
// and the restriction), one on the way up (the prolongation and both reverse
// sweeps), and one workgroup for the small levels: 2 launches per level and
// V-cycle instead of 10 pass by pass (at 1024 x 512, FP64: 208 instead of
// 451 us per PCG iteration).  A workgroup owns a GT_X x GT_Y tile of the level and keeps the
// tile and a halo of H cells in LDS; each colour pass updates the region
// inset by one more cell, so after four passes the values on the tile (and,
// on the way down, one ring more for the residual) are those of the
// whole-level passes: each cell's update is the same arithmetic on the same
// neighbour values, so the results do not depend on the tiling.

constexpr int GT_X = 64, GT_Y = 16, GT_THREADS = 256;   // 32 x 8 coarse cells per tile: one per thread
constexpr int GMG_TAIL_THREADS = 1024, GMG_TAIL_MAXL = 8;
constexpr int GMG_TAIL_LDS = 60 * 1024;                  // bytes of x, b and mask of the tail levels

// the neighbour rule of gmg_nb on an LDS array of row stride RX: e is the
// cell, (gi, gj) its level coordinates; wall / solid neighbours give the
// cell's own value and drop their face from the diagonal
template <typename CT>
__device__ __forceinline__ void gmg_nb_lds(const CT* X, const uint8_t* M, int e, int RX, int gi, int gj, const Geo& g,
                                           CT& xc, CT& xL, CT& xR, CT& xB, CT& xT, bool& wl, bool& wr, bool& wb, bool& wt) {
    xc = X[e];
    wl = gi == 1 || M[e - 1];
    wr = gi == g.nx || M[e + 1];
    wb = gj == 1 || M[e - RX];
    wt = gj == g.ny || M[e + RX];
    xL = wl ? xc : X[e - 1];
    xR = wr ? xc : X[e + 1];
    xB = wb ? xc : X[e - RX];
    xT = wt ? xc : X[e + RX];
}

// one cell of a colour pass, the reference smoother's arithmetic in its order
// (gmg_preconditioner.cpp:41-81, amg_preconditioner.cpp:13-50)
template <typename CT>
__device__ __forceinline__ CT gmg_update_lds(const CT* X, const CT* B, const uint8_t* M, int e, int RX, int gi, int gj,
                                             const Geo& g, const GmgC<CT>& c) {
    CT xc, xL, xR, xB, xT;
    bool wl, wr, wb, wt;
    gmg_nb_lds<CT>(X, M, e, RX, gi, gj, g, xc, xL, xR, xB, xT, wl, wr, wb, wt);
    const CT lap = (xL + xR) * c.idx2 + (xB + xT) * c.idy2;
    CT eff_d = c.diag;
    if (wl) eff_d -= c.idx2;
    if (wr) eff_d -= c.idx2;
    if (wb) eff_d -= c.idy2;
    if (wt) eff_d -= c.idy2;
    return xc + ((eff_d < CT(1e-15)) ? CT(0) : CT(1) / eff_d) * (B[e] - c.diag * xc + lap);
}

// the residual of one fine cell as the restriction forms it (GMG
// (b - diag x) + lap, gmg_preconditioner.cpp:121-142; AMG b - (diag x - lap),
// amg_preconditioner.cpp:91-111)
template <bool AMG, typename CT>
__device__ __forceinline__ CT gmg_resid_lds(const CT* X, const CT* B, const uint8_t* M, int e, int RX, int gi, int gj,
                                            const Geo& g, const GmgC<CT>& c) {
    CT xc, xL, xR, xB, xT;
    bool wl, wr, wb, wt;
    gmg_nb_lds<CT>(X, M, e, RX, gi, gj, g, xc, xL, xR, xB, xT, wl, wr, wb, wt);
    if constexpr (AMG) {
        const CT Ax = c.diag * xc - (xL + xR) * c.idx2 - (xB + xT) * c.idy2;
        return B[e] - Ax;
    } else {
        const CT lap = (xL + xR) * c.idx2 + (xB + xT) * c.idy2;
        return B[e] - c.diag * xc + lap;
    }
}

// restriction of the four children of coarse cell (ic, jc) (children (0,0),
// (0,1), (1,0), (1,1); the restriction's sums and writes); (gi0, gj0) are the
// level coordinates of LDS cell 0
template <bool AMG, typename CT, typename VC>
__device__ __forceinline__ void gmg_restrict_one(const CT* X, const CT* B, const uint8_t* M, int RX, int gi0, int gj0,
                                                 const Geo& g, const GmgC<CT>& c, int ic, int jc, VC* bc,
                                                 const uint8_t* SC, const Geo& cg) {
    CT sum = CT(0);
    int cnt = 0;
#pragma unroll
    for (int di = 0; di < 2; di++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++) {
            const int fi = 2 * ic - 1 + di, fj = 2 * jc - 1 + dj;
            const int e = (fi - gi0) + (fj - gj0) * RX;
            if (M[e]) continue;
            sum += gmg_resid_lds<AMG, CT>(X, B, M, e, RX, fi, fj, g, c);
            cnt++;
        }
    const long cid = (long)ic + (long)jc * cg.sx;
    if constexpr (AMG) bc[cid] = (VC)((cnt > 0) ? sum / (CT)cnt : CT(0));
    else if (!SC[cid] && cnt > 0) bc[cid] = (VC)(sum / (CT)cnt);
}

// Per-cell smoother data of a tile region, formed once per launch (the mask
// does not change between passes): F = wall / solid flags of the four
// faces (bit 0 -x, 1 +x, 2 -y, 3 +y), D = (eff_d < 1e-15) ? 0 : 1 / eff_d
// with eff_d formed as the smoother forms it.  The update is then the
// smoother's arithmetic on the same values.
template <typename CT>
__device__ __forceinline__ void gmg_cell_data(const uint8_t* M, int e, int RX, int gi, int gj, const Geo& g,
                                              const GmgC<CT>& c, uint8_t& f, CT& d) {
    const bool wl = gi == 1 || M[e - 1], wr = gi == g.nx || M[e + 1];
    const bool wb = gj == 1 || M[e - RX], wt = gj == g.ny || M[e + RX];
    CT eff_d = c.diag;
    if (wl) eff_d -= c.idx2;
    if (wr) eff_d -= c.idx2;
    if (wb) eff_d -= c.idy2;
    if (wt) eff_d -= c.idy2;
    f = (uint8_t)((wl ? 1 : 0) | (wr ? 2 : 0) | (wb ? 4 : 0) | (wt ? 8 : 0));
    d = (eff_d < CT(1e-15)) ? CT(0) : CT(1) / eff_d;
}
template <typename CT>
__device__ __forceinline__ CT gmg_update_fd(const CT* X, const CT* B, uint8_t f, CT d, int e, int RX, const GmgC<CT>& c) {
    const CT xc = X[e];
    const CT xL = (f & 1) ? xc : X[e - 1];
    const CT xR = (f & 2) ? xc : X[e + 1];
    const CT xB = (f & 4) ? xc : X[e - RX];
    const CT xT = (f & 8) ? xc : X[e + RX];
    const CT lap = (xL + xR) * c.idx2 + (xB + xT) * c.idy2;
    return xc + d * (B[e] - c.diag * xc + lap);
}

// One colour pass over a tile region inset by p cells: the region's cells
// of that colour, two to a row pair of lanes apart, so every lane works.
template <typename CT, int RX, int RY>
__device__ __forceinline__ void gmg_tile_pass(CT* X, const CT* B, const uint8_t* M, const uint8_t* F, const CT* Dv,
                                              int gi0, int gj0, const GmgC<CT>& c, int p, int colour) {
    constexpr int HX = RX / 2;   // RX even: HX cells of each colour per row
    for (int k = threadIdx.x; k < RY * HX; k += GT_THREADS) {
        const int rj = k / HX;
        const int ri = 2 * (k % HX) + ((colour - gi0 - gj0 - rj) & 1);
        if (ri < p || ri >= RX - p || rj < p || rj >= RY - p) continue;
        const int e = ri + rj * RX;
        if (M[e]) continue;
        X[e] = gmg_update_fd<CT>(X, B, F[e], Dv[e], e, RX, c);
    }
    __syncthreads();
}

// Down: x = 0, smooth twice (odd, even colours), restrict the residual; x
// of the tile and the coarse b written.  Halo 5: four passes, then the
// residual's neighbours.
template <bool AMG, typename VX, typename VC>
__global__ __launch_bounds__(GT_THREADS) void k_gmg_down(VX* __restrict__ x, const VX* __restrict__ b,
                                                         const uint8_t* __restrict__ S, Geo g,
                                                         GmgC<typename CTOf<VX>::type> c, VC* __restrict__ bc,
                                                         const uint8_t* __restrict__ SC, Geo cg, const int* skip) {
    using CT = typename CTOf<VX>::type;
    if (skip && *skip) return;
    constexpr int H = 5, RX = GT_X + 2 * H, RY = GT_Y + 2 * H, N = RX * RY;
    static_assert(RX % 2 == 0, "two colours per row");
    __shared__ CT X[N];
    __shared__ CT B[N];
    __shared__ CT Dv[N];
    __shared__ uint8_t M[N];
    __shared__ uint8_t F[N];
    const int gi0 = 1 + (int)blockIdx.x * GT_X - H, gj0 = 1 + (int)blockIdx.y * GT_Y - H;
    {
        // every load of the region issued before the first LDS store
        constexpr int NT = (N + GT_THREADS - 1) / GT_THREADS;
        uint8_t ms[NT];
        CT bs[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int e = (int)threadIdx.x + t * GT_THREADS;
            const int gi = gi0 + e % RX, gj = gj0 + e / RX;
            const bool in = e < N && gi >= 1 && gi <= g.nx && gj >= 1 && gj <= g.ny;
            const long id = in ? (long)gi + (long)gj * g.sx : 0;
            ms[t] = in ? (S[id] ? 1 : 0) : 1;   // outside the level: never updated, never a neighbour value
            bs[t] = in ? (CT)b[id] : CT(0);
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int e = (int)threadIdx.x + t * GT_THREADS;
            if (e < N) { M[e] = ms[t]; B[e] = bs[t]; X[e] = CT(0); }
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < N; e += GT_THREADS) {
        const int ri = e % RX, rj = e / RX;
        if (ri < 1 || ri >= RX - 1 || rj < 1 || rj >= RY - 1 || M[e]) continue;
        gmg_cell_data<CT>(M, e, RX, gi0 + ri, gj0 + rj, g, c, F[e], Dv[e]);
    }
    __syncthreads();
    for (int p = 1; p <= 4; p++) gmg_tile_pass<CT, RX, RY>(X, B, M, F, Dv, gi0, gj0, c, p, p & 1);   // odd, even, twice
    {
        const int ic = 1 + (int)blockIdx.x * (GT_X / 2) + (int)threadIdx.x % (GT_X / 2);
        const int jc = 1 + (int)blockIdx.y * (GT_Y / 2) + (int)threadIdx.x / (GT_X / 2);
        if (ic <= cg.nx && jc <= cg.ny) gmg_restrict_one<AMG, CT, VC>(X, B, M, RX, gi0, gj0, g, c, ic, jc, bc, SC, cg);
    }
    for (int e = threadIdx.x; e < GT_X * GT_Y; e += GT_THREADS) {
        const int ti = e % GT_X, tj = e / GT_X;
        const int gi = gi0 + H + ti, gj = gj0 + H + tj;
        if (gi <= g.nx && gj <= g.ny) x[(long)gi + (long)gj * g.sx] = (VX)X[(ti + H) + (tj + H) * RX];
    }
}

// Up: x += the parent's x on fluid cells (GMG: fluid parents only), smooth
// twice in reverse (even, odd colours); x of the tile written.  Halo 4.
template <bool AMG, typename VX, typename VC>
__global__ __launch_bounds__(GT_THREADS) void k_gmg_up(VX* __restrict__ x, const VX* __restrict__ b,
                                                       const uint8_t* __restrict__ S, Geo g,
                                                       GmgC<typename CTOf<VX>::type> c, const VC* __restrict__ xcoarse,
                                                       const uint8_t* __restrict__ SC, Geo cg, const int* skip) {
    using CT = typename CTOf<VX>::type;
    if (skip && *skip) return;
    constexpr int H = 4, RX = GT_X + 2 * H, RY = GT_Y + 2 * H, N = RX * RY;
    static_assert(RX % 2 == 0, "two colours per row");
    __shared__ CT X[N];
    __shared__ CT B[N];
    __shared__ CT Dv[N];
    __shared__ uint8_t M[N];
    __shared__ uint8_t F[N];
    const int gi0 = 1 + (int)blockIdx.x * GT_X - H, gj0 = 1 + (int)blockIdx.y * GT_Y - H;
    {
        // every load of the region (fine mask, b, x; the parent's mask and x)
        // issued before the first LDS store
        constexpr int NT = (N + GT_THREADS - 1) / GT_THREADS;
        uint8_t ms[NT], pm[NT];
        CT bs[NT], xs[NT], xp[NT];
        bool hp[NT];
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int e = (int)threadIdx.x + t * GT_THREADS;
            const int gi = gi0 + e % RX, gj = gj0 + e / RX;
            const bool in = e < N && gi >= 1 && gi <= g.nx && gj >= 1 && gj <= g.ny;
            const long id = in ? (long)gi + (long)gj * g.sx : 0;
            const int ic = (gi + 1) >> 1, jc = (gj + 1) >> 1;
            hp[t] = in && ic <= cg.nx && jc <= cg.ny;
            const long cid = hp[t] ? (long)ic + (long)jc * cg.sx : 0;
            ms[t] = in ? (S[id] ? 1 : 0) : 1;
            bs[t] = in ? (CT)b[id] : CT(0);
            xs[t] = in ? (CT)x[id] : CT(0);
            pm[t] = hp[t] ? SC[cid] : 1;
            xp[t] = hp[t] ? (CT)xcoarse[cid] : CT(0);
        }
#pragma unroll
        for (int t = 0; t < NT; t++) {
            const int e = (int)threadIdx.x + t * GT_THREADS;
            if (e >= N) continue;
            CT v = xs[t];
            if (!ms[t] && hp[t] && (AMG || !pm[t])) v = (CT)(VX)(v + xp[t]);   // prolongateAdd (gmg :144-156, amg :113-123)
            M[e] = ms[t]; B[e] = bs[t]; X[e] = v;
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < N; e += GT_THREADS) {
        const int ri = e % RX, rj = e / RX;
        if (ri < 1 || ri >= RX - 1 || rj < 1 || rj >= RY - 1 || M[e]) continue;
        gmg_cell_data<CT>(M, e, RX, gi0 + ri, gj0 + rj, g, c, F[e], Dv[e]);
    }
    __syncthreads();
    for (int p = 1; p <= 4; p++) gmg_tile_pass<CT, RX, RY>(X, B, M, F, Dv, gi0, gj0, c, p, (p & 1) ? 0 : 1);   // even, odd, twice
    for (int e = threadIdx.x; e < GT_X * GT_Y; e += GT_THREADS) {
        const int ti = e % GT_X, tj = e / GT_X;
        const int gi = gi0 + H + ti, gj = gj0 + H + tj;
        if (gi <= g.nx && gj <= g.ny) x[(long)gi + (long)gj * g.sx] = (VX)X[(ti + H) + (tj + H) * RX];
    }
}

// The small levels in one workgroup: x, b and mask of every tail level in
// LDS, the V-cycle below the first tail level run pass by pass with
// workgroup barriers (gmg_vcycle from the first tail level down).
struct GmgTailArgs {
    int nl;                              // levels in the tail; the last is the coarsest
    Geo g[GMG_TAIL_MAXL];
    GmgC<double> c[GMG_TAIL_MAXL];
    const uint8_t* mask[GMG_TAIL_MAXL];  // padded solid masks (device)
    int off[GMG_TAIL_MAXL];              // LDS element offset of each level's arrays
    int total;                           // padded cells of all tail levels
    int csweeps;                         // coarsest (forward + reverse) pairs
};

// one colour pass over a whole tail level, the level's cells of that colour
// only; threads [t0, t0 + nt) take part
template <typename CT>
__device__ __forceinline__ void gmg_tail_cells(CT* X, const CT* B, const uint8_t* M, const Geo& g, const GmgC<CT>& c,
                                               int colour, int t0, int nt) {
    const int RX = (int)g.sx, HX = (g.nx + 1) / 2;
    for (int k = (int)threadIdx.x - t0; k < g.ny * HX; k += nt) {
        const int gj = 1 + k / HX;
        const int gi = 1 + ((colour - 1 - gj) & 1) + 2 * (k % HX);
        if (gi > g.nx) continue;
        const int e = gi + gj * RX;
        if (M[e]) continue;
        X[e] = gmg_update_lds<CT>(X, B, M, e, RX, gi, gj, g, c);
    }
}
// End of a tail phase: the whole workgroup, or (levels with at most 64
// cells of a colour, run by wave 0 alone) wave order only -- the LDS
// accesses of one wave complete in program order; the fences keep the
// compiler from moving them across phases.
__device__ __forceinline__ void gmg_tail_sync(bool wave) {
    if (wave) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    } else {
        __syncthreads();
    }
}

template <bool AMG, typename CT>
struct GmgTailLevels {
    const GmgTailArgs& A;
    CT *Xa, *Ba;
    uint8_t* Ma;
    __device__ GmgC<CT> c(int l) const { return GmgC<CT>{(CT)A.c[l].idx2, (CT)A.c[l].idy2, (CT)A.c[l].diag}; }
    __device__ void smooth(int l, bool reverse, int sweeps, bool wave) const {
        const int nt = wave ? 64 : (int)blockDim.x;
        for (int s = 0; s < sweeps; s++)
            for (int q = 0; q < 2; q++) {
                gmg_tail_cells<CT>(Xa + A.off[l], Ba + A.off[l], Ma + A.off[l], A.g[l], c(l), (q == 0) == reverse ? 0 : 1, 0, nt);
                gmg_tail_sync(wave);
            }
    }
    __device__ void restrict_to(int l, bool wave) const {   // level l -> l + 1
        const int nt = wave ? 64 : (int)blockDim.x;
        const Geo& cg = A.g[l + 1];
        for (int t = threadIdx.x; t < cg.nx * cg.ny; t += nt) {
            const int ic = 1 + t % cg.nx, jc = 1 + t / cg.nx;
            gmg_restrict_one<AMG, CT, CT>(Xa + A.off[l], Ba + A.off[l], Ma + A.off[l], (int)A.g[l].sx, 0, 0, A.g[l], c(l),
                                          ic, jc, Ba + A.off[l + 1], Ma + A.off[l + 1], cg);
        }
        gmg_tail_sync(wave);
    }
    __device__ void prolong_from(int l, bool wave) const {   // level l + 1 -> l
        const int nt = wave ? 64 : (int)blockDim.x;
        CT* X = Xa + A.off[l];
        const uint8_t* M = Ma + A.off[l];
        const Geo& g = A.g[l];
        const Geo& cg = A.g[l + 1];
        const CT* XC = Xa + A.off[l + 1];
        const uint8_t* MC = Ma + A.off[l + 1];
        for (int e = threadIdx.x; e < (int)g.N; e += nt) {
            const int gi = e % (int)g.sx, gj = e / (int)g.sx;
            if (gi < 1 || gi > g.nx || gj < 1 || gj > g.ny || M[e]) continue;
            const int ic = (gi + 1) >> 1, jc = (gj + 1) >> 1;
            if (ic > cg.nx || jc > cg.ny) continue;
            const int cid = ic + jc * (int)cg.sx;
            if (!AMG && MC[cid]) continue;
            X[e] = X[e] + XC[cid];
        }
        gmg_tail_sync(wave);
    }
};

template <bool AMG, typename VF, typename CT = typename CTOf<VF>::type>
__global__ __launch_bounds__(GMG_TAIL_THREADS) void k_gmg_tail(GmgTailArgs A, VF* __restrict__ x0,
                                                               const VF* __restrict__ b0, const int* skip) {
    if (skip && *skip) return;
    extern __shared__ __align__(16) unsigned char gmg_lds[];
    CT* Xa = reinterpret_cast<CT*>(gmg_lds);
    CT* Ba = Xa + A.total;
    uint8_t* Ma = reinterpret_cast<uint8_t*>(Ba + A.total);
    for (int l = 0; l < A.nl; l++) {
        const Geo& g = A.g[l];
        for (int e = threadIdx.x; e < (int)g.N; e += blockDim.x) {
            const int gi = e % (int)g.sx, gj = e / (int)g.sx;
            const bool in = gi >= 1 && gi <= g.nx && gj >= 1 && gj <= g.ny;
            Ma[A.off[l] + e] = in ? (A.mask[l][e] ? 1 : 0) : 1;
            Xa[A.off[l] + e] = CT(0);
            Ba[A.off[l] + e] = (l == 0 && in) ? (CT)b0[e] : CT(0);
        }
    }
    __syncthreads();
    const GmgTailLevels<AMG, CT> T{A, Xa, Ba, Ma};
    const int nl = A.nl;
    // first level small enough for wave 0 alone (levels shrink downwards)
    int lw = nl;
    for (int l = 0; l < nl; l++)
        if (A.g[l].ny * ((A.g[l].nx + 1) / 2) <= 64) { lw = l; break; }
    const int top = lw < nl - 1 ? lw : nl - 1;   // levels [top, nl) run in the inner part
    // down through the workgroup levels
    for (int l = 0; l < top; l++) { T.smooth(l, false, 2, false); T.restrict_to(l, false); }
    // the small levels and the coarsest (csweeps x (forward, reverse))
    {
        const bool wave = lw <= nl - 1;
        if (!wave || threadIdx.x < 64) {
            for (int l = top; l < nl - 1; l++) { T.smooth(l, false, 2, wave); T.restrict_to(l, wave); }
            for (int s = 0; s < A.csweeps; s++) { T.smooth(nl - 1, false, 1, wave); T.smooth(nl - 1, true, 1, wave); }
            for (int l = nl - 2; l >= top; l--) { T.prolong_from(l, wave); T.smooth(l, true, 2, wave); }
        }
        __syncthreads();
    }
    // up through the workgroup levels
    for (int l = top - 1; l >= 0; l--) { T.prolong_from(l, false); T.smooth(l, true, 2, false); }
    const Geo& g0 = A.g[0];
    for (int e = threadIdx.x; e < (int)g0.N; e += blockDim.x) {
        const int gi = e % (int)g0.sx, gj = e / (int)g0.sx;
        if (gi >= 1 && gi <= g0.nx && gj >= 1 && gj <= g0.ny) x0[e] = (VF)Xa[e];
    }
}

}  // namespace uaamg
