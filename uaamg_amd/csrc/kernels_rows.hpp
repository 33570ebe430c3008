// kernels_rows.hpp -- level-0 V(1,1) kernels of the FP32 path (3D).
//
// The same preconditioner as the reference's level-0 smooth /
// restrictResidual / prolongateAdd (uaamg_preconditioner_3d.cpp:123-181).
// The level's x starts at 0 in every V-cycle, so the forward sweep and the
// residual are local formulas of r (the level-0 b):
//   odd  cells: x0 = r / d                      (forward sweep, odd colour)
//   even cells: x0 = (r + sum_nb c x0_nb) / d   (forward sweep, even colour)
//   residual:   odd: sum_nb c x0_nb  (r - d x0 = 0),  even: 0,
//               isolated fluid cells (d < 1e-30, skipped by the sweeps): r
// and on the way up, with x' = x0 + 2 x_coarse on fluid cells:
//   even cells: z = (r + sum_nb c x'_nb) / d    (reverse sweep, even colour)
//   odd  cells: z = (r + sum_nb c z_nb) / d     (reverse sweep, odd colour)
//   isolated cells keep x'.
// These are identities in exact arithmetic; the FP32 path rounds them
// differently from the FP64 parity kernels (which keep the reference order).
// Level-0 couplings are 1/h^2 on every face between two fluid interior cells
// and 0 otherwise, and x is 0 on every non-fluid cell, so the neighbour sums
// use the per-axis constants.
//
// Mapping (wave64): a wave owns whole x-rows, CPL consecutive cells per lane
// (nx <= 64*CPL, no x halo); RW rows per wave, NW waves stacked in y (2 halo
// rows on each side of the tile); the workgroup marches along z over a chunk
// of KC planes (+2 halo planes each side).  Colours are compile-time per
// (cell, row, plane slot), so no lane idles on the other colour.  x-neighbours
// across lanes move by wave shuffle, y-neighbours across waves through LDS
// (one barrier per plane step), z-neighbours stay in a 4-plane register
// queue.  The 2x2x2 restriction children all live in one thread.  Workgroups
// are numbered XCD-first (consecutive ids go to different XCDs), so the tiles
// one XCD runs are neighbours in y and share their halo rows in its L2.
#pragma once
#include <type_traits>
#include "common.hpp"

namespace uaamg {

// Level-0 stencil constants of the FP32 path: 1/d and d per neighbour pattern
// (nb >> 1 = +x,-x,+y,-y,+z,-z present; 1/d = 0 where d < 1e-30).  Dirichlet
// walls add their excess per cell.
struct RowTab {
    float rinv[64];
    float d[64];
    float idx2, idy2, idz2;
    float ex[6];
    int any_dirichlet;
};

struct RowGeo {
    int nty;       // tiles in y
    int nkc;       // chunks in z
    int KC;        // planes per chunk (multiple of 4)
    int ntiles;    // nty * nkc
    int per_xcd;   // ceil(ntiles / 8)
};

// Whole-wave shifts by one lane with DPP (wave_shr:1 = 0x138, wave_shl:1 =
// 0x130; all rows and banks enabled, lanes with no source keep old = 0):
// one VALU move instead of a ds_bpermute round trip through the LDS unit.
__device__ __forceinline__ float lane_from_left(float v) {   // lane l gets lane l-1; lane 0 gets 0
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x138, 0xF, 0xF, false));
}
__device__ __forceinline__ float lane_from_right(float v) {  // lane l gets lane l+1; lane 63 gets 0
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xF, 0xF, false));
}

// tile of block b of a single-grid launch: consecutive block ids go
// round-robin to the 8 XCDs, so block b runs on XCD b % 8; XCD x gets tiles
// [x*per_xcd, (x+1)*per_xcd)
__device__ __forceinline__ int row_tile_of_block(const RowGeo& rg, int b) {
    return (b & 7) * rg.per_xcd + (b >> 3);
}

// 1/d of the RW x CPL cells of a lane's plane (0 on non-fluid cells and
// where d < 1e-30): without Dirichlet walls the table value; with them d =
// the table's d plus each touching wall's excess, divided in FP32.  Every
// table read is issued before the first is used (one LDS round trip per
// plane, not one per cell) and the Dirichlet test is one uniform branch.
template <int RW, int CPL>
__device__ __forceinline__ void rows_rinv_plane(const uint8_t (&nb)[RW][CPL], float (&ri)[RW][CPL], const float* srinv,
                                                const float* sd, const RowTab& T, const Geo& g, int i0,
                                                const int (&jrow)[RW], int k) {
    float v[RW][CPL];
    if (!T.any_dirichlet) {
#pragma unroll
        for (int t = 0; t < RW; t++)
#pragma unroll
            for (int q = 0; q < CPL; q++) v[t][q] = srinv[nb[t][q] >> 1];
#pragma unroll
        for (int t = 0; t < RW; t++)
#pragma unroll
            for (int q = 0; q < CPL; q++) ri[t][q] = (nb[t][q] & NB_FLUID) ? v[t][q] : 0.0f;
    } else {
#pragma unroll
        for (int t = 0; t < RW; t++)
#pragma unroll
            for (int q = 0; q < CPL; q++) v[t][q] = sd[nb[t][q] >> 1];
#pragma unroll
        for (int t = 0; t < RW; t++)
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                float d = v[t][q];
                const int i = i0 + q, j = jrow[t];
                if (i == 1) d += T.ex[0];
                if (i == g.nx) d += T.ex[1];
                if (j == 1) d += T.ex[2];
                if (j == g.ny) d += T.ex[3];
                if (k == 1) d += T.ex[4];
                if (k == g.nz) d += T.ex[5];
                ri[t][q] = (nb[t][q] & NB_FLUID) ? (d >= 1e-30f ? 1.0f / d : 0.0f) : 0.0f;
            }
    }
}

// The CPL values and neighbour bytes of one lane's cells in one vector load
// each.  Row starts are not 16-byte aligned in the padded layout, so the
// types carry 4-byte (values) and 1-byte (bytes) alignment; gfx950 global
// loads take unaligned addresses.  Lanes past nx read the ghost cell and what
// follows it (inside the array and its margin) and are masked with okx.
template <int CPL> struct RowVec;
template <> struct RowVec<4> {
    typedef float V __attribute__((ext_vector_type(4), aligned(4)));
    typedef uint32_t B __attribute__((aligned(1)));
};
template <> struct RowVec<2> {
    typedef float V __attribute__((ext_vector_type(2), aligned(4)));
    typedef uint16_t B __attribute__((aligned(1)));
};
template <int CPL, bool NT = false>
__device__ __forceinline__ void row_load(const float* __restrict__ r, const uint8_t* __restrict__ nbr, int id,
                                         const bool (&okx)[CPL], float (&vr)[CPL], uint8_t (&vn)[CPL]) {
    typename RowVec<CPL>::V v;
    if constexpr (NT) v = __builtin_nontemporal_load((const typename RowVec<CPL>::V*)(r + id));
    else v = *(const typename RowVec<CPL>::V*)(r + id);
    const uint32_t b = *(const typename RowVec<CPL>::B*)(nbr + id);
#pragma unroll
    for (int q = 0; q < CPL; q++) {
        vr[q] = okx[q] ? v[q] : 0.0f;
        vn[q] = okx[q] ? (uint8_t)(b >> (8 * q)) : (uint8_t)0;
    }
}

template <int CPL, int NW, int RW>
struct RowLDS {
    float tab[2][64];
    // rows shared between waves: [buffer][kind 0: S1 output, 1: S2 output]
    // [wave][0: first row, 1: last row][colour cell q/2][lane]
    float ex[2][2][NW][2][CPL / 2][64];
};

template <int V> using IC = std::integral_constant<int, V>;

// Sum over the 6 neighbours of cell (T, Q) of plane slot s: in-plane values
// from X[s] (other waves' rows from the LDS exchange eb/kind), z-neighbours
// from slots sp and sm; edge = the neighbour across the lane boundary.
template <int CPL, int NW, int RW, int T, int Q, int s, int sp, int sm>
__device__ __forceinline__ float row_nsum(const float (&X)[4][RW][CPL], float edge, const RowLDS<CPL, NW, RW>& L,
                                          int eb, int kind, int w, int lane, const RowTab& K) {
    float xl, xr, xu, xd;
    if constexpr (Q > 0) xl = X[s][T][Q - 1]; else xl = edge;
    if constexpr (Q < CPL - 1) xr = X[s][T][Q + 1]; else xr = edge;
    if constexpr (T > 0) xu = X[s][T - 1][Q]; else xu = L.ex[eb][kind][w > 0 ? w - 1 : 0][1][Q / 2][lane];
    if constexpr (T < RW - 1) xd = X[s][T + 1][Q]; else xd = L.ex[eb][kind][w < NW - 1 ? w + 1 : NW - 1][0][Q / 2][lane];
    return K.idx2 * (xl + xr) + K.idy2 * (xu + xd) + K.idz2 * (X[sp][T][Q] + X[sm][T][Q]);
}

template <int N, typename F>
__device__ __forceinline__ void sfor(F&& f) {
    if constexpr (N > 0) {
        sfor<N - 1>(f);
        f(IC<N - 1>{});
    }
}

// Shared set-up of the row kernels (H halo rows on each side of the tile).
// Addresses are 32-bit element offsets: a row's cells are roff[t] + q.  Rows
// and planes outside the grid are clamped to the ghost layer (neighbour byte
// 0: read as non-fluid), lanes past nx are masked with okx.  FULL: the
// launch guarantees nx == 64 CPL, so okx and full are compile-time true and
// every lane mask folds away (the masks cost about a tenth of the level-0
// kernels' VALU instructions).
template <int CPL, int NW, int RW, int H = 2, bool FULL = false>
struct RowThread {
    static constexpr int TY = NW * RW - 2 * H;
    int lane, w, i0, b0, m0, last;
    int jrow[RW];
    int roff[RW];
    bool okx[CPL];
    bool tile_row[RW];
    bool full;        // every lane of every row is inside the grid (nx == 64 CPL)
    __device__ void init(const Geo& g, const RowGeo& rg, int ty, int kc) {
        lane = threadIdx.x & 63;
        // the wave index is uniform: in a scalar register, the row tests
        // below (jrow, tile_row) are scalar branches, not exec-mask ones
        w = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
        full = FULL || g.nx == 64 * CPL;
        i0 = 1 + CPL * lane;
        b0 = 1 + ty * TY;
        m0 = 1 + kc * rg.KC;
        last = min(m0 + rg.KC - 1, g.nz);
#pragma unroll
        for (int q = 0; q < CPL; q++) okx[q] = FULL || i0 + q <= g.nx;
#pragma unroll
        for (int t = 0; t < RW; t++) {
            jrow[t] = b0 - H + w * RW + t;
            tile_row[t] = (w * RW + t >= H) && (w * RW + t < TY + H) && jrow[t] <= g.ny;
            const int jc = min(max(jrow[t], 0), g.ny + 1);
            roff[t] = min(i0, g.nx + 1) + jc * (int)g.sx;
        }
    }
    // element offset of plane P (clamped to the ghost planes)
    __device__ __forceinline__ int plane(const Geo& g, int P) const {
        return min(max(P, 0), g.nz + 1) * (int)g.sy;
    }
};

// ------------------------------------------------------------------ down
// r -> coarse b (restricted residual after the forward sweep).  A workgroup
// writes the coarse cells whose 8 children lie in its tile.
//
// With FR (fused CG residual update, pcg_3d.cpp:106 for r) the kernel
// reads r_old and Ap and works on r = r_old - alpha Ap (recomputed on the
// halo), writes that r on its tile to r_new (another buffer: neighbouring
// workgroups still read r_old) and stores the partial sums of r.r over its
// tile's fluid cells.  It also takes over the operator's finish
// (k_fin_matvec): every workgroup sums the operator's p.Ap partials in the
// same order and forms alpha = (r, z) / p.Ap itself (pcg_3d.cpp:99; (r, z)
// as the z-sums finish stored it); workgroup 0 records p.Ap, alpha, the x
// update's alpha, a breakdown, and rsold = (r, z) for the next beta, in
// scalars that no other workgroup of this launch reads.
// the convergence test (k_fin_update) folded into the next kernel: the r.r
// partials it sums and the scalars (ctrl == nullptr: not folded)
struct FinUpd {
    const double* partials = nullptr;
    int nb = 0;
    Ctrl* ctrl = nullptr;
};

struct RUpd {
    const float* Ap = nullptr;
    float* rnew = nullptr;
    double* partials = nullptr;      // r.r partials (this kernel's)
    const double* pap = nullptr;     // p.Ap partials of the operator pass
    int npap = 0;
    Ctrl* ctrl = nullptr;
    int xslot = 0;                   // ctrl->xa slot of this iteration's x pass (its p buffer)
};

template <int CPL, int NW, int RW, typename VC, bool FR = false, bool NT = false, bool FULL = false>
__device__ __forceinline__ void d_rows_down0(const float* __restrict__ r, const uint8_t* __restrict__ nbr, const Geo& g,
                                             const RowTab& K, const RowGeo& rg, VC* __restrict__ bc, const Geo& cg,
                                             const int* skip, int tile, const RUpd& u = RUpd{}) {
    static_assert(CPL % 2 == 0 && RW % 2 == 0, "pairs of cells and rows");
    if (skip && *skip) {
        // a stopped solve: no x update for this iteration
        if constexpr (FR) { if (blockIdx.x == 0 && threadIdx.x == 0) u.ctrl->xa[u.xslot] = 0.0; }
        return;
    }
    __shared__ RowLDS<CPL, NW, RW> L;
    float alpha = 0.0f;   // r update in FP32 (one FMA), as in k_rows_pmv
    if constexpr (FR) {
        __shared__ double s_pap;
        double tot[1];
        reduce_partials_at<1>(u.pap, u.npap, 0, u.npap, tot);
        if (threadIdx.x == 0) s_pap = tot[0];
        __syncthreads();
        const double pAp = s_pap;
        const bool ok = isfinite(pAp) && pAp >= 1e-15;   // pcg.cpp:100 breakdown test
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            Ctrl* c = u.ctrl;
            c->pAp = pAp;
            if (!isfinite(pAp)) { c->done = 1; c->status = ST_NONFINITE; c->iter += 1; c->alpha = 0.0; }
            else if (pAp < 1e-15) { c->done = 1; c->status = ST_BREAKDOWN; c->iter += 1; c->alpha = 0.0; }
            else c->alpha = c->rz / pAp;
            c->xa[u.xslot] = c->alpha;
            // the operator pass read rsold and the first flag; nothing in
            // this launch does
            c->rsold = c->rz;
            c->first = 0;
        }
        if (!ok) return;   // the whole solve stops; the finish after this kernel sees done
        alpha = (float)(u.ctrl->rz / pAp);
    }
    if (tile >= rg.ntiles) {   // the whole workgroup leaves together
        if constexpr (FR) { double z0[1] = {0.0}; block_store<1>(z0, u.partials); }
        return;
    }
    const int ty = tile % rg.nty, kc = tile / rg.nty;
    double rr[1] = {0.0};
    RowThread<CPL, NW, RW, 2, FULL> th;
    th.init(g, rg, ty, kc);
    const int lane = th.lane, w = th.w;
    for (int e = threadIdx.x; e < 64; e += NW * 64) { L.tab[0][e] = K.rinv[e]; L.tab[1][e] = K.d[e]; }
    for (int e = threadIdx.x; e < 2 * 2 * NW * 2 * (CPL / 2) * 64; e += NW * 64) (&L.ex[0][0][0][0][0][0])[e] = 0.0f;
    __syncthreads();

    float X[4][RW][CPL], Rq[2][RW][CPL], Iq[4][RW][CPL];
    float pfr[2][RW][CPL];
    float pfa[FR ? 2 : 1][RW][CPL];
    uint8_t pfn[2][RW][CPL];
    float acc[RW / 2][CPL / 2];
#pragma unroll
    for (int s = 0; s < 4; s++)
#pragma unroll
        for (int t = 0; t < RW; t++)
#pragma unroll
            for (int q = 0; q < CPL; q++) { X[s][t][q] = 0.0f; Rq[s & 1][t][q] = 0.0f; Iq[s][t][q] = 0.0f; }
#pragma unroll
    for (int a = 0; a < RW / 2; a++)
#pragma unroll
        for (int c = 0; c < CPL / 2; c++) acc[a][c] = 0.0f;

    auto issue = [&](int slot, int P) {
        // unconditional loads from clamped addresses: straight-line code lets
        // the compiler wait for exactly the loads a step consumes
        const int po = th.plane(g, P);
#pragma unroll
        for (int t = 0; t < RW; t++) {
            row_load<CPL, NT>(r, nbr, th.roff[t] + po, th.okx, pfr[slot][t], pfn[slot][t]);
            if constexpr (FR) {
                typename RowVec<CPL>::V v;
                if constexpr (NT) v = __builtin_nontemporal_load((const typename RowVec<CPL>::V*)(u.Ap + th.roff[t] + po));
                else v = *(const typename RowVec<CPL>::V*)(u.Ap + th.roff[t] + po);
#pragma unroll
                for (int q = 0; q < CPL; q++) pfa[slot][t][q] = v[q];
            }
        }
    };
    issue(1, th.m0 - 2);
    issue(0, th.m0 - 1);

    // one plane step: plane P enters the window; S = P & 3
    auto step = [&](auto Sc, int P) {
        constexpr int S = decltype(Sc)::value;
        constexpr int S1 = (S + 3) & 3, S2 = (S + 2) & 3, S3 = (S + 1) & 3;   // planes P-1, P-2, P-3
        constexpr int PB = S & 1;        // prefetch buffer of plane P
        constexpr int EW = S & 1;        // LDS buffer written in this step
        constexpr int ER = EW ^ 1;       // LDS buffer written in the previous step
        // S1: plane P enters; odd cells x0 = r / d.  Parity of (i,j,k) is
        // (q + t + P) & 1 (i0 and the first region row are odd, RW even).
        if constexpr (FR) {   // r = r_old - alpha Ap; r_new and r.r on the tile
            const bool pin = P >= th.m0 && P <= th.last;
            sfor<RW>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                float rs = 0.0f;   // the row's r.r in FP32, accumulated in FP64
                sfor<CPL>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    const bool fl = pfn[PB][t][q] & NB_FLUID;
                    const float rn = fl ? fmaf(-alpha, pfa[PB][t][q], pfr[PB][t][q]) : 0.0f;
                    pfr[PB][t][q] = rn;
                    rs = fmaf(th.okx[q] ? rn : 0.0f, rn, rs);
                });
                if (pin && th.tile_row[t]) rr[0] += (double)rs;
                if (pin && th.tile_row[t]) {
                    const int id = th.roff[t] + P * (int)g.sy;
                    if (th.full || th.okx[CPL - 1]) {
                        typename RowVec<CPL>::V v;
#pragma unroll
                        for (int q = 0; q < CPL; q++) v[q] = pfr[PB][t][q];
                        *(typename RowVec<CPL>::V*)(u.rnew + id) = v;
                    } else {
#pragma unroll
                        for (int q = 0; q < CPL; q++)
                            if (th.okx[q]) u.rnew[id + q] = pfr[PB][t][q];
                    }
                }
            });
        }
        float riP[RW][CPL];
        rows_rinv_plane<RW, CPL>(pfn[PB], riP, L.tab[0], L.tab[1], K, g, th.i0, th.jrow, P);
        sfor<RW>([&](auto tc) {
            constexpr int t = decltype(tc)::value;
            sfor<CPL>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const uint8_t nb = pfn[PB][t][q];
                const float rv = (nb & NB_FLUID) ? pfr[PB][t][q] : 0.0f;
                const float ri = riP[t][q];
                Rq[S & 1][t][q] = rv;
                Iq[S][t][q] = ri;
                // isolated fluid cells (ri == 0, no fluid neighbour: their x is
                // never read) carry r, their residual; non-fluid cells carry 0
                if constexpr (((q + t + S) & 1) == 1) X[S][t][q] = (ri != 0.0f) ? rv * ri : rv;
                else X[S][t][q] = (ri != 0.0f) ? 0.0f : rv;
            });
        });
        issue(PB, min(P + 2, th.last + 2));   // past last + 2 (never used): the cached plane last + 2 again, no memory traffic
        sfor<CPL>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (((q + S) & 1) == 1) L.ex[EW][0][w][0][q / 2][lane] = X[S][0][q];
            if constexpr (((q + RW - 1 + S) & 1) == 1) L.ex[EW][0][w][1][q / 2][lane] = X[S][RW - 1][q];
        });
        // S2: plane P-1, even cells: x0 = (r + sum c x0_odd) / d
        if (P - 1 >= th.m0 - 1) {
            sfor<RW>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr bool q0 = ((t + S1) & 1) == 0;   // cell 0 of this row is even
                const float edge = q0 ? lane_from_left(X[S1][t][CPL - 1]) : lane_from_right(X[S1][t][0]);
                sfor<CPL>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    if constexpr (((q + t + S1) & 1) == 0) {
                        const float s = row_nsum<CPL, NW, RW, t, q, S1, S, S2>(X, edge, L, ER, 0, w, lane, K);
                        // (r + s) / d on cells with 1/d != 0; isolated and non-fluid
                        // cells keep X.  Written as a blend so it stays branch-free.
                        const float ri = Iq[S1][t][q];
                        X[S1][t][q] = (Rq[S1 & 1][t][q] + s) * ri + ((ri != 0.0f) ? 0.0f : X[S1][t][q]);
                    }
                });
            });
            sfor<CPL>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (((q + S1) & 1) == 0) L.ex[EW][1][w][0][q / 2][lane] = X[S1][0][q];
                if constexpr (((q + RW - 1 + S1) & 1) == 0) L.ex[EW][1][w][1][q / 2][lane] = X[S1][RW - 1][q];
            });
        }
        // S3: plane R = P-2: residual of the odd cells, restriction over the
        // plane pair (R-1, R) with R even
        const int R = P - 2;
        if (R >= th.m0 && R <= th.last) {
            if (R & 1) {
#pragma unroll
                for (int a = 0; a < RW / 2; a++)
#pragma unroll
                    for (int c = 0; c < CPL / 2; c++) acc[a][c] = 0.0f;
            }
            sfor<RW>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr bool q0 = ((t + S2) & 1) == 1;   // cell 0 of this row is odd
                const float edge = q0 ? lane_from_left(X[S2][t][CPL - 1]) : lane_from_right(X[S2][t][0]);
                sfor<CPL>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    if constexpr (((q + t + S2) & 1) == 1) {
                        const float s = row_nsum<CPL, NW, RW, t, q, S2, S1, S3>(X, edge, L, ER, 1, w, lane, K);
                        const bool iso = Iq[S2][t][q] == 0.0f;
                        acc[t / 2][q / 2] += (iso ? 0.0f : s) + (iso ? X[S2][t][q] : 0.0f);
                    } else {
                        // even cells: residual 0 after the even update, r on isolated cells
                        acc[t / 2][q / 2] += (Iq[S2][t][q] != 0.0f) ? 0.0f : X[S2][t][q];
                    }
                });
            });
            if (!(R & 1)) {
                const int kcc = R >> 1;
#pragma unroll
                for (int a = 0; a < RW / 2; a++) {
                    if (!th.tile_row[2 * a]) continue;
                    const int jc = (th.jrow[2 * a] + 1) >> 1;
#pragma unroll
                    for (int c = 0; c < CPL / 2; c++) {
                        const int ic = (th.i0 + 2 * c + 1) >> 1;
                        if (ic <= cg.nx && jc <= cg.ny && kcc <= cg.nz)
                            bc[ic + jc * (int)cg.sx + kcc * (int)cg.sy] = (VC)acc[a][c];
                    }
                }
            }
        }
        __syncthreads();
    };
    for (int P = th.m0 - 2; P <= th.last + 2; P += 4) {   // m0 - 2 = 3 (mod 4)
        step(IC<3>{}, P);
        if (P + 1 > th.last + 2) break;
        step(IC<0>{}, P + 1);
        if (P + 2 > th.last + 2) break;
        step(IC<1>{}, P + 2);
        if (P + 3 > th.last + 2) break;
        step(IC<2>{}, P + 3);
    }
    if constexpr (FR) block_store<1>(rr, u.partials);
}

template <int CPL, int NW, int RW, typename VC>
__global__ __launch_bounds__(NW * 64) void k_rows_down0(const float* __restrict__ r, const uint8_t* __restrict__ nbr,
                                                        Geo g, RowTab K, RowGeo rg, VC* __restrict__ bc, Geo cg,
                                                        const int* skip) {
    d_rows_down0<CPL, NW, RW, VC>(r, nbr, g, K, rg, bc, cg, skip, row_tile_of_block(rg, blockIdx.x));
}

// down0 with the fused residual update (r_old -> r_new, r.r partials) and
// the operator's finish (alpha from the p.Ap partials)
template <int CPL, int NW, int RW, typename VC, bool NT = false, bool FULL = false>
__global__ __launch_bounds__(NW * 64) void k_rows_down0r(const float* __restrict__ rold, const uint8_t* __restrict__ nbr,
                                                         Geo g, RowTab K, RowGeo rg, VC* __restrict__ bc, Geo cg,
                                                         const int* skip, const float* __restrict__ Ap,
                                                         float* __restrict__ rnew, double* partials,
                                                         const double* pap, int npap, Ctrl* ctrl, int xslot) {
    RUpd u;
    u.Ap = Ap; u.rnew = rnew; u.partials = partials; u.pap = pap; u.npap = npap; u.ctrl = ctrl; u.xslot = xslot;
    d_rows_down0<CPL, NW, RW, VC, true, NT, FULL>(rold, nbr, g, K, rg, bc, cg, skip, row_tile_of_block(rg, blockIdx.x), u);
}

// ------------------------------------------------------------------ up
// r, coarse x -> z (prolongation + reverse sweep), with sum z, sum r z, sum r
// over fluid cells reduced per workgroup into partials.
template <int CPL, int NW, int RW, typename VCX, bool NT = false, bool FULL = false>
__device__ __forceinline__ void d_rows_up0(const float* __restrict__ r, const uint8_t* __restrict__ nbr, const Geo& g,
                                           const RowTab& K, const RowGeo& rg, const VCX* __restrict__ xc, const Geo& cg,
                                           float* __restrict__ z, const int* skip, double* partials, int tile) {
    static_assert(CPL % 2 == 0 && RW % 2 == 0, "pairs of cells and rows");
    if (skip && *skip) return;
    __shared__ RowLDS<CPL, NW, RW> L;
    const bool active = tile < rg.ntiles;   // idle workgroups still store zero partials
    const int ty = tile % rg.nty, kc = tile / rg.nty;
    RowThread<CPL, NW, RW, 2, FULL> th;
    th.init(g, rg, active ? ty : 0, active ? kc : 0);
    const int lane = th.lane, w = th.w;
    for (int e = threadIdx.x; e < 64; e += NW * 64) { L.tab[0][e] = K.rinv[e]; L.tab[1][e] = K.d[e]; }
    for (int e = threadIdx.x; e < 2 * 2 * NW * 2 * (CPL / 2) * 64; e += NW * 64) (&L.ex[0][0][0][0][0][0])[e] = 0.0f;
    __syncthreads();
    double red[3] = {0.0, 0.0, 0.0};

    if (active) {
        float X[4][RW][CPL], Rq[4][RW][CPL], Iq[4][RW][CPL];
        float pfr[2][RW][CPL];
        uint8_t pfn[2][RW][CPL];
        float pfx[2][RW / 2][CPL / 2];
        bool pcin[2];
        int coff[RW / 2][CPL / 2];
        bool cok[RW / 2][CPL / 2];
#pragma unroll
        for (int a = 0; a < RW / 2; a++)
#pragma unroll
            for (int c = 0; c < CPL / 2; c++) {
                const int ic = (th.i0 + 2 * c + 1) >> 1, jc = (th.jrow[2 * a] + 1) >> 1;
                cok[a][c] = th.okx[2 * c] && th.jrow[2 * a] >= 1 && th.jrow[2 * a] <= g.ny && ic <= cg.nx && jc >= 1 && jc <= cg.ny;
                coff[a][c] = min(ic, cg.nx + 1) + min(max(jc, 0), cg.ny + 1) * (int)cg.sx;
            }
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int t = 0; t < RW; t++)
#pragma unroll
                for (int q = 0; q < CPL; q++) { X[s][t][q] = 0.0f; Rq[s][t][q] = 0.0f; Iq[s][t][q] = 0.0f; }

        auto issue = [&](int slot, int P) {
            const int po = th.plane(g, P);
#pragma unroll
            for (int t = 0; t < RW; t++) row_load<CPL, NT>(r, nbr, th.roff[t] + po, th.okx, pfr[slot][t], pfn[slot][t]);
            // coarse x of the parents; masked where it is used (S1), so the
            // loads stay in flight
            const int kcc = (P + 1) >> 1;
            pcin[slot] = P >= 1 && P <= g.nz && kcc >= 1 && kcc <= cg.nz;
            const int cpo = min(max(kcc, 0), cg.nz + 1) * (int)cg.sy;
#pragma unroll
            for (int a = 0; a < RW / 2; a++)
#pragma unroll
                for (int c = 0; c < CPL / 2; c++) pfx[slot][a][c] = (float)xc[coff[a][c] + cpo];
        };
        issue(1, th.m0 - 2);
        issue(0, th.m0 - 1);

        auto step = [&](auto Sc, int P) {
            constexpr int S = decltype(Sc)::value;
            constexpr int S1 = (S + 3) & 3, S2 = (S + 2) & 3, S3 = (S + 1) & 3;
            constexpr int PB = S & 1, EW = S & 1, ER = EW ^ 1;
            // S1: plane P: x' = x0 + 2 x_coarse (x0 = r/d on odd cells; the
            // even x0 is replaced by the reverse sweep, except on isolated cells
            // where it is 0)
            float riP[RW][CPL];
            rows_rinv_plane<RW, CPL>(pfn[PB], riP, L.tab[0], L.tab[1], K, g, th.i0, th.jrow, P);
            sfor<RW>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                sfor<CPL>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    const uint8_t nb = pfn[PB][t][q];
                    const bool fl = nb & NB_FLUID;
                    const float rv = fl ? pfr[PB][t][q] : 0.0f;
                    const float ri = riP[t][q];
                    const float w2 = (fl && pcin[PB] && cok[t / 2][q / 2]) ? 2.0f * pfx[PB][t / 2][q / 2] : 0.0f;
                    Rq[S][t][q] = rv;
                    Iq[S][t][q] = ri;
                    if constexpr (((q + t + S) & 1) == 1) X[S][t][q] = rv * ri + w2;
                    else X[S][t][q] = w2;
                });
            });
            issue(PB, min(P + 2, th.last + 2));   // past last + 2 (never used): the cached plane last + 2 again, no memory traffic
            sfor<CPL>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                if constexpr (((q + S) & 1) == 1) L.ex[EW][0][w][0][q / 2][lane] = X[S][0][q];
                if constexpr (((q + RW - 1 + S) & 1) == 1) L.ex[EW][0][w][1][q / 2][lane] = X[S][RW - 1][q];
            });
            // S2: plane P-1, even cells: z = (r + sum c x'_odd) / d
            if (P - 1 >= th.m0 - 1) {
                sfor<RW>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr bool q0 = ((t + S1) & 1) == 0;
                    const float edge = q0 ? lane_from_left(X[S1][t][CPL - 1]) : lane_from_right(X[S1][t][0]);
                    sfor<CPL>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        if constexpr (((q + t + S1) & 1) == 0) {
                            const float s = row_nsum<CPL, NW, RW, t, q, S1, S, S2>(X, edge, L, ER, 0, w, lane, K);
                            const float ri = Iq[S1][t][q];
                            X[S1][t][q] = (Rq[S1][t][q] + s) * ri + ((ri != 0.0f) ? 0.0f : X[S1][t][q]);
                        }
                    });
                });
                sfor<CPL>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    if constexpr (((q + S1) & 1) == 0) L.ex[EW][1][w][0][q / 2][lane] = X[S1][0][q];
                    if constexpr (((q + RW - 1 + S1) & 1) == 0) L.ex[EW][1][w][1][q / 2][lane] = X[S1][RW - 1][q];
                });
            }
            // S3: plane R = P-2, odd cells: z = (r + sum c z_even) / d; output
            const int R = P - 2;
            if (R >= th.m0 && R <= th.last) {
                sfor<RW>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    constexpr bool q0 = ((t + S2) & 1) == 1;
                    const float edge = q0 ? lane_from_left(X[S2][t][CPL - 1]) : lane_from_right(X[S2][t][0]);
                    sfor<CPL>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        if constexpr (((q + t + S2) & 1) == 1) {
                            const float s = row_nsum<CPL, NW, RW, t, q, S2, S1, S3>(X, edge, L, ER, 1, w, lane, K);
                            const float ri = Iq[S2][t][q];
                            X[S2][t][q] = (Rq[S2][t][q] + s) * ri + ((ri != 0.0f) ? 0.0f : X[S2][t][q]);
                        }
                    });
                    if (th.tile_row[t]) {
                        // the row's CPL cells summed in FP32 (one FMA each), the
                        // row sums accumulated in FP64
                        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
                        sfor<CPL>([&](auto qc) {
                            constexpr int q = decltype(qc)::value;
                            const float zv = th.okx[q] ? X[S2][t][q] : 0.0f, rv = th.okx[q] ? Rq[S2][t][q] : 0.0f;
                            s0 += zv;
                            s1 = fmaf(rv, zv, s1);
                            s2 += rv;
                        });
                        red[0] += (double)s0;
                        red[1] += (double)s1;
                        red[2] += (double)s2;
                        const int id = th.roff[t] + R * (int)g.sy;
                        if (th.full || th.okx[CPL - 1]) {
                            typename RowVec<CPL>::V v;
#pragma unroll
                            for (int q = 0; q < CPL; q++) v[q] = X[S2][t][q];
                            *(typename RowVec<CPL>::V*)(z + id) = v;
                        } else {
#pragma unroll
                            for (int q = 0; q < CPL; q++)
                                if (th.okx[q]) z[id + q] = X[S2][t][q];
                        }
                    }
                });
            }
            __syncthreads();
        };
        for (int P = th.m0 - 2; P <= th.last + 2; P += 4) {
            step(IC<3>{}, P);
            if (P + 1 > th.last + 2) break;
            step(IC<0>{}, P + 1);
            if (P + 2 > th.last + 2) break;
            step(IC<1>{}, P + 2);
            if (P + 3 > th.last + 2) break;
            step(IC<2>{}, P + 3);
        }
    }
    block_store<3>(red, partials);
}

template <int CPL, int NW, int RW, typename VCX, bool NT = false, bool FULL = false>
__global__ __launch_bounds__(NW * 64) void k_rows_up0(const float* __restrict__ r, const uint8_t* __restrict__ nbr,
                                                      Geo g, RowTab K, RowGeo rg, const VCX* __restrict__ xc, Geo cg,
                                                      float* __restrict__ z, const int* skip, double* partials) {
    d_rows_up0<CPL, NW, RW, VCX, NT, FULL>(r, nbr, g, K, rg, xc, cg, z, skip, partials, row_tile_of_block(rg, blockIdx.x));
}

// ================================================================== record levels
// Levels l >= 1 (3D) with the same schedule: per-cell records {cx, cy, cz, d}
// (couplings to the +x/+y/+z neighbours; d < 0: inactive), level values b / x
// in SV (float, or __bf16 for the BF16 level storage), arithmetic in FP32.
// The coupling to a -x/-y/-z neighbour is that neighbour's own +coupling:
// -x from the thread's previous cell or, for cell 0, the previous lane (DPP);
// -y from the thread's previous row or, for a wave's first row, a load of cy
// one row up (an L1/L2 hit: the wave above loads that row); -z from the
// previous plane's slot.  1/d is the hardware reciprocal, as in the march
// kernels of these levels.  Cells outside the grid and lanes past nx read as
// inactive.

template <int CPL, typename SV>
__device__ __forceinline__ void vals_load(const SV* __restrict__ p, int id, const bool (&ok)[CPL], float (&v)[CPL]) {
    if constexpr (std::is_same<SV, float>::value) {
        const typename RowVec<CPL>::V u = *(const typename RowVec<CPL>::V*)(p + id);
#pragma unroll
        for (int q = 0; q < CPL; q++) v[q] = ok[q] ? u[q] : 0.0f;
    } else {
        static_assert(sizeof(SV) == 2, "16-bit level values");
        typedef unsigned short H __attribute__((ext_vector_type(CPL), aligned(2)));
        const H u = *(const H*)((const unsigned short*)p + id);
#pragma unroll
        for (int q = 0; q < CPL; q++) v[q] = ok[q] ? __builtin_bit_cast(float, (uint32_t)u[q] << 16) : 0.0f;
    }
}

// planes the record kernels (levels 64 < nx <= 128) keep in flight: the
// loads of a plane are issued REC_PD steps before it enters (a power of 2).
// 4 measured slower at 128^3 (45 / 38 us instead of 19 / 21 us): the
// extra buffers take the kernels past 168 VGPRs, into scratch.
constexpr int REC_PD = 2;

template <typename C4> struct RecE;
template <> struct RecE<float4> { using E = float; };
template <> struct RecE<BF16x4> { using E = __bf16; };

// BF16 records: the diagonal is formed from the six couplings (snap_diag)
template <typename C4> constexpr bool rec_xd() { return std::is_same<C4, BF16x4>::value; }

template <int CPL, int RW, typename C4>
struct RecPf {            // one plane's loads
    float b[RW][CPL];
    C4 rec[RW][CPL];
    C4 up[rec_xd<C4>() ? CPL : 1];   // BF16: the row above the first row (its cy is the -y coupling)
};

// LDS of the record kernels: the row exchange, and the cy of each wave's last
// row per plane slot (the -y coupling of the next wave's first row)
template <int CPL, int NW, int RW>
struct RecLDS {
    RowLDS<CPL, NW, RW> r;
    float cy[4][NW][CPL][64];
};

template <int CPL, int NW, int RW>
struct RecThread : RowThread<CPL, NW, RW> {
    bool rowok[RW];
    __device__ void init_rec(const Geo& g, const RowGeo& rg, int ty, int kc) {
        this->init(g, rg, ty, kc);
#pragma unroll
        for (int t = 0; t < RW; t++) rowok[t] = this->jrow[t] >= 1 && this->jrow[t] <= g.ny;
    }
};

template <int CPL, int NW, int RW, typename SV, typename C4>
__device__ __forceinline__ void rec_issue(RecPf<CPL, RW, C4>& f, const SV* __restrict__ b, const C4* __restrict__ rec,
                                          const RecThread<CPL, NW, RW>& th, const Geo& g, int P) {
    const int po = th.plane(g, P);
#pragma unroll
    for (int t = 0; t < RW; t++) {
        vals_load<CPL, SV>(b, th.roff[t] + po, th.okx, f.b[t]);
#pragma unroll
        for (int q = 0; q < CPL; q++) f.rec[t][q] = rec[th.roff[t] + po + q];
    }
    if constexpr (rec_xd<C4>()) {
        const int ju = min(max(th.jrow[0] - 1, 0), g.ny + 1);
        const int ru = min(th.i0, g.nx + 1) + ju * (int)g.sx;
#pragma unroll
        for (int q = 0; q < CPL; q++) f.up[q] = rec[ru + po + q];
    }
}


// per-plane register state of the record kernels
template <int CPL, int RW>
struct RecQ {
    float X[4][RW][CPL], Iq[4][RW][CPL], Rq[4][RW][CPL];
    float Cx[4][RW][CPL], Cy[4][RW][CPL], Cz[4][RW][CPL];
};

// BF16 records: the -z coupling of the first halo plane (plane m0 - 3's cz)
// into the slot before it (slot 2: plane m0 - 2 has slot 3)
template <int CPL, int NW, int RW, typename C4>
__device__ __forceinline__ void rec_prev_cz(RecQ<CPL, RW>& Q, const C4* __restrict__ rec,
                                            const RecThread<CPL, NW, RW>& th, const Geo& g) {
    if constexpr (rec_xd<C4>()) {
        const int P = th.m0 - 3;
        const int po = th.plane(g, P);
#pragma unroll
        for (int t = 0; t < RW; t++)
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                const bool ok = P >= 1 && th.rowok[t] && th.okx[q];
                Q.Cz[2][t][q] = ok ? (float)rec[th.roff[t] + po + q].z : 0.0f;
            }
    } else {
        (void)Q; (void)rec; (void)th; (void)g;
    }
}

// plane P (slot S) enters: unpack the prefetched loads; the last row's cy
// goes to the LDS slot S for the wave below
template <int S, int CPL, int NW, int RW, typename C4>
__device__ __forceinline__ void rec_enter(RecQ<CPL, RW>& Q, const RecPf<CPL, RW, C4>& f, const RecThread<CPL, NW, RW>& th,
                                          bool planeok, RecLDS<CPL, NW, RW>& L) {
    float dd[RW][CPL];
#pragma unroll
    for (int t = 0; t < RW; t++)
#pragma unroll
        for (int q = 0; q < CPL; q++) {
            const bool ok = planeok && th.rowok[t] && th.okx[q];
            const C4 r = f.rec[t][q];
            const float d = ok ? (float)r.w : -1.0f;
            dd[t][q] = d;
            Q.Cx[S][t][q] = ok ? (float)r.x : 0.0f;
            Q.Cy[S][t][q] = ok ? (float)r.y : 0.0f;
            Q.Cz[S][t][q] = ok ? (float)r.z : 0.0f;
            Q.Rq[S][t][q] = d >= 0.0f ? f.b[t][q] : 0.0f;
        }
    if constexpr (rec_xd<C4>()) {
        // the diagonal from the six couplings: -x from the previous cell (the
        // previous lane for cell 0), -y from the row above (loaded for the
        // first row), -z from the previous plane's slot
        constexpr int SP = (S + 3) & 3;
#pragma unroll
        for (int t = 0; t < RW; t++) {
            const float cl0 = lane_from_left(Q.Cx[S][t][CPL - 1]);
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                const float cxm = (q > 0) ? Q.Cx[S][t][q > 0 ? q - 1 : 0] : cl0;
                const float cym = (t > 0) ? Q.Cy[S][t > 0 ? t - 1 : 0][q] : (float)f.up[q].y;
                const float sum = Q.Cx[S][t][q] + cxm + Q.Cy[S][t][q] + cym + Q.Cz[S][t][q] + Q.Cz[SP][t][q];
                if (dd[t][q] >= 0.0f) dd[t][q] = snap_diag<C4, float>(dd[t][q], sum);
            }
        }
    }
#pragma unroll
    for (int t = 0; t < RW; t++)
#pragma unroll
        for (int q = 0; q < CPL; q++) Q.Iq[S][t][q] = dd[t][q] >= 1e-30f ? __builtin_amdgcn_rcpf(dd[t][q]) : 0.0f;
#pragma unroll
    for (int q = 0; q < CPL; q++) L.cy[S][th.w][q][th.lane] = Q.Cy[S][RW - 1][q];
}

// sum over the 6 neighbours of cell (T, Qc) of slot s (z+: slot sp, z-: slot
// sm); edge / cedge: the neighbour value / its coupling across the lane boundary
// (wave 0's first row is the outer halo row: its -y terms are never output)
template <int CPL, int NW, int RW, int T, int Qc, int s, int sp, int sm>
__device__ __forceinline__ float rec_nsum(const RecQ<CPL, RW>& Q, float edge, float cedge, const RecLDS<CPL, NW, RW>& L,
                                          int eb, int kind, int w, int lane) {
    float xl, xr, xu, xd, cl, cu;
    if constexpr (Qc > 0) { xl = Q.X[s][T][Qc - 1]; cl = Q.Cx[s][T][Qc - 1]; } else { xl = edge; cl = cedge; }
    if constexpr (Qc < CPL - 1) xr = Q.X[s][T][Qc + 1]; else xr = edge;
    if constexpr (T > 0) { xu = Q.X[s][T - 1][Qc]; cu = Q.Cy[s][T - 1][Qc]; }
    else { xu = L.r.ex[eb][kind][w > 0 ? w - 1 : 0][1][Qc / 2][lane]; cu = L.cy[s][w > 0 ? w - 1 : 0][Qc][lane]; }
    if constexpr (T < RW - 1) xd = Q.X[s][T + 1][Qc]; else xd = L.r.ex[eb][kind][w < NW - 1 ? w + 1 : NW - 1][0][Qc / 2][lane];
    return Q.Cx[s][T][Qc] * xr + cl * xl + Q.Cy[s][T][Qc] * xd + cu * xu + Q.Cz[s][T][Qc] * Q.X[sp][T][Qc]
         + Q.Cz[sm][T][Qc] * Q.X[sm][T][Qc];
}

// one colour of plane slot s: x = (b + sum) / d on the cells of parity PAR
// with 1/d != 0 (others keep x); neighbour rows of other waves from LDS
// buffer er / kind rk; with PUB, the wave's first / last row go to buffer ew /
// kind 1 for the next colour
template <int PAR, bool PUB, int CPL, int NW, int RW, int s, int sp, int sm>
__device__ __forceinline__ void rec_colour(RecQ<CPL, RW>& Q, RecLDS<CPL, NW, RW>& L, int er, int ew, int rk, int w,
                                           int lane) {
    sfor<RW>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        constexpr bool q0 = ((t + s) & 1) == PAR;   // cell 0 of this row has parity PAR
        const float edge = q0 ? lane_from_left(Q.X[s][t][CPL - 1]) : lane_from_right(Q.X[s][t][0]);
        const float cedge = q0 ? lane_from_left(Q.Cx[s][t][CPL - 1]) : 0.0f;
        sfor<CPL>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (((q + t + s) & 1) == PAR) {
                const float sum = rec_nsum<CPL, NW, RW, t, q, s, sp, sm>(Q, edge, cedge, L, er, rk, w, lane);
                const float ri = Q.Iq[s][t][q];
                Q.X[s][t][q] = (Q.Rq[s][t][q] + sum) * ri + ((ri != 0.0f) ? 0.0f : Q.X[s][t][q]);
            }
        });
    });
    if constexpr (PUB) {
        sfor<CPL>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (((q + s) & 1) == PAR) L.r.ex[ew][1][w][0][q / 2][lane] = Q.X[s][0][q];
            if constexpr (((q + RW - 1 + s) & 1) == PAR) L.r.ex[ew][1][w][1][q / 2][lane] = Q.X[s][RW - 1][q];
        });
    }
}

// b (level l) -> coarse b: forward sweep and restricted residual
template <int CPL, int NW, int RW, typename SV, typename C4>
__device__ __forceinline__ void d_rows_downc(const SV* __restrict__ b, const C4* __restrict__ rec, const Geo& g,
                                             const RowGeo& rg, SV* __restrict__ bc, const Geo& cg, const int* skip, int tile,
                                             const FinUpd& fin = FinUpd{}) {
    static_assert(CPL % 2 == 0 && RW % 2 == 0, "pairs of cells and rows");
    if (fin.ctrl) {
        // the convergence test of the level-0 down kernel's r.r partials
        // (k_fin_update) in every workgroup: each sums them in the same order
        // and stops on the same r.r; workgroup 0 records it.  Only workgroup
        // 0 reads the iteration count (it increments it): a max_iter stop is
        // seen by the following kernels, and any work this launch does after
        // it is unused.
        // (done is read once per workgroup: workgroup 0 of this launch may
        // set it while others start, and the whole workgroup must agree)
        __shared__ int s_done;
        __shared__ double s_rs;
        if (threadIdx.x == 0) s_done = fin.ctrl->done;
        __syncthreads();
        if (s_done) return;
        double tot[1];
        reduce_partials_at<1>(fin.partials, fin.nb, 0, fin.nb, tot);
        if (threadIdx.x == 0) s_rs = tot[0];
        __syncthreads();
        const double rs = s_rs;
        Ctrl* c = fin.ctrl;
        const bool conv = (c->tol_mode == 1) ? (rs < c->tolsq) : (sqrt(rs) < c->tol);
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            c->rsnew = rs;
            if (c->hist) c->hist[c->iter] = rs;
            c->iter += 1;
            if (conv) { c->done = 1; c->status = ST_CONVERGED; }
            else if (!isfinite(rs)) { c->done = 1; c->status = ST_NONFINITE; }
            else if (c->iter >= c->max_iter) { c->done = 1; c->status = ST_MAXITER; }
        }
        if (conv || !isfinite(rs)) return;
    } else if (skip && *skip) {
        return;
    }
    __shared__ RecLDS<CPL, NW, RW> L;
    if (tile >= rg.ntiles) return;
    const int ty = tile % rg.nty, kc = tile / rg.nty;
    RecThread<CPL, NW, RW> th;
    th.init_rec(g, rg, ty, kc);
    const int lane = th.lane, w = th.w;
    for (int e = threadIdx.x; e < 2 * 2 * NW * 2 * (CPL / 2) * 64; e += NW * 64) (&L.r.ex[0][0][0][0][0][0])[e] = 0.0f;
    for (int e = threadIdx.x; e < 4 * NW * CPL * 64; e += NW * 64) (&L.cy[0][0][0][0])[e] = 0.0f;
    __syncthreads();

    RecQ<CPL, RW> Q;
    RecPf<CPL, RW, C4> pf[REC_PD];
    float acc[RW / 2][CPL / 2];
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
#pragma unroll
        for (int t = 0; t < RW; t++)
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                Q.X[s4][t][q] = 0.0f; Q.Iq[s4][t][q] = 0.0f; Q.Rq[s4][t][q] = 0.0f;
                Q.Cx[s4][t][q] = 0.0f; Q.Cy[s4][t][q] = 0.0f; Q.Cz[s4][t][q] = 0.0f;
            }
    }
#pragma unroll
    for (int a = 0; a < RW / 2; a++)
#pragma unroll
        for (int c = 0; c < CPL / 2; c++) acc[a][c] = 0.0f;
    rec_prev_cz<CPL, NW, RW, C4>(Q, rec, th, g);
    // planes m0 - 2 .. m0 + REC_PD - 3 in flight before the first step, each
    // in the buffer of its slot (m0 = 1 mod 4: plane m0 - 2 + d has slot 3 + d)
    sfor<REC_PD>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        rec_issue<CPL, NW, RW, SV, C4>(pf[(3 + d) & (REC_PD - 1)], b, rec, th, g, th.m0 - 2 + d);
    });

    auto step = [&](auto Sc, int P) {
        constexpr int S = decltype(Sc)::value;
        constexpr int S1 = (S + 3) & 3, S2 = (S + 2) & 3, S3 = (S + 1) & 3;
        constexpr int PB = S & (REC_PD - 1), EW = S & 1, ER = EW ^ 1;
        // S1: plane P enters; odd cells x0 = b / d; isolated cells carry b
        rec_enter<S, CPL, NW, RW, C4>(Q, pf[PB], th, P >= 1 && P <= g.nz, L);
        sfor<RW>([&](auto tc) {
            constexpr int t = decltype(tc)::value;
            sfor<CPL>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const float ri = Q.Iq[S][t][q], rv = Q.Rq[S][t][q];
                if constexpr (((q + t + S) & 1) == 1) Q.X[S][t][q] = (ri != 0.0f) ? rv * ri : rv;
                else Q.X[S][t][q] = (ri != 0.0f) ? 0.0f : rv;
            });
        });
        rec_issue<CPL, NW, RW, SV, C4>(pf[PB], b, rec, th, g, min(P + REC_PD, th.last + 2));   // past last + 2 (never used): the cached plane last + 2 again
        sfor<CPL>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (((q + S) & 1) == 1) L.r.ex[EW][0][w][0][q / 2][lane] = Q.X[S][0][q];
            if constexpr (((q + RW - 1 + S) & 1) == 1) L.r.ex[EW][0][w][1][q / 2][lane] = Q.X[S][RW - 1][q];
        });
        // S2: plane P-1, even cells
        if (P - 1 >= th.m0 - 1) rec_colour<0, true, CPL, NW, RW, S1, S, S2>(Q, L, ER, EW, 0, w, lane);
        // S3: plane R = P-2: residual (odd: sum c x0_even; even: 0; isolated: b)
        // and restriction over the plane pair (R-1, R), R even
        const int R = P - 2;
        if (R >= th.m0 && R <= th.last) {
            if (R & 1) {
#pragma unroll
                for (int a = 0; a < RW / 2; a++)
#pragma unroll
                    for (int c = 0; c < CPL / 2; c++) acc[a][c] = 0.0f;
            }
            sfor<RW>([&](auto tc) {
                constexpr int t = decltype(tc)::value;
                constexpr bool q0 = ((t + S2) & 1) == 1;
                const float edge = q0 ? lane_from_left(Q.X[S2][t][CPL - 1]) : lane_from_right(Q.X[S2][t][0]);
                const float cedge = q0 ? lane_from_left(Q.Cx[S2][t][CPL - 1]) : 0.0f;
                sfor<CPL>([&](auto qc) {
                    constexpr int q = decltype(qc)::value;
                    const bool iso = Q.Iq[S2][t][q] == 0.0f;
                    if constexpr (((q + t + S2) & 1) == 1) {
                        const float sum = rec_nsum<CPL, NW, RW, t, q, S2, S1, S3>(Q, edge, cedge, L, ER, 1, w, lane);
                        acc[t / 2][q / 2] += (iso ? 0.0f : sum) + (iso ? Q.X[S2][t][q] : 0.0f);
                    } else {
                        acc[t / 2][q / 2] += iso ? Q.X[S2][t][q] : 0.0f;
                    }
                });
            });
            if (!(R & 1)) {
                const int kcc = R >> 1;
#pragma unroll
                for (int a = 0; a < RW / 2; a++) {
                    if (!th.tile_row[2 * a]) continue;
                    const int jc = (th.jrow[2 * a] + 1) >> 1;
#pragma unroll
                    for (int c = 0; c < CPL / 2; c++) {
                        const int ic = (th.i0 + 2 * c + 1) >> 1;
                        if (ic <= cg.nx && jc <= cg.ny && kcc <= cg.nz)
                            bc[ic + jc * (int)cg.sx + kcc * (int)cg.sy] = (SV)acc[a][c];
                    }
                }
            }
        }
        __syncthreads();
    };
    for (int P = th.m0 - 2; P <= th.last + 2; P += 4) {
        step(IC<3>{}, P);
        if (P + 1 > th.last + 2) break;
        step(IC<0>{}, P + 1);
        if (P + 2 > th.last + 2) break;
        step(IC<1>{}, P + 2);
        if (P + 3 > th.last + 2) break;
        step(IC<2>{}, P + 3);
    }
}

template <int CPL, int NW, int RW, typename SV, typename C4>
__global__ __launch_bounds__(NW * 64) void k_rows_downc(const SV* __restrict__ b, const C4* __restrict__ rec, Geo g,
                                                        RowGeo rg, SV* __restrict__ bc, Geo cg, const int* skip,
                                                        FinUpd fin) {
    d_rows_downc<CPL, NW, RW, SV, C4>(b, rec, g, rg, bc, cg, skip, row_tile_of_block(rg, blockIdx.x), fin);
}

// b (level l), coarse x -> x (level l): prolongation and reverse sweep
template <int CPL, int NW, int RW, typename SV, typename C4>
__device__ __forceinline__ void d_rows_upc(const SV* __restrict__ b, const C4* __restrict__ rec, const Geo& g,
                                           const RowGeo& rg, const SV* __restrict__ xc, const Geo& cg,
                                           SV* __restrict__ xout, const int* skip, int tile) {
    static_assert(CPL % 2 == 0 && RW % 2 == 0, "pairs of cells and rows");
    if (skip && *skip) return;
    __shared__ RecLDS<CPL, NW, RW> L;
    if (tile >= rg.ntiles) return;
    const int ty = tile % rg.nty, kc = tile / rg.nty;
    RecThread<CPL, NW, RW> th;
    th.init_rec(g, rg, ty, kc);
    const int lane = th.lane, w = th.w;
    for (int e = threadIdx.x; e < 2 * 2 * NW * 2 * (CPL / 2) * 64; e += NW * 64) (&L.r.ex[0][0][0][0][0][0])[e] = 0.0f;
    for (int e = threadIdx.x; e < 4 * NW * CPL * 64; e += NW * 64) (&L.cy[0][0][0][0])[e] = 0.0f;
    __syncthreads();

    RecQ<CPL, RW> Q;
    RecPf<CPL, RW, C4> pf[REC_PD];
    float pfx[REC_PD][RW / 2][CPL / 2];
    bool pcin[REC_PD];
    int coff[RW / 2][CPL / 2];
    bool cok[RW / 2][CPL / 2];
#pragma unroll
    for (int a = 0; a < RW / 2; a++)
#pragma unroll
        for (int c = 0; c < CPL / 2; c++) {
            const int ic = (th.i0 + 2 * c + 1) >> 1, jc = (th.jrow[2 * a] + 1) >> 1;
            cok[a][c] = th.okx[2 * c] && th.jrow[2 * a] >= 1 && th.jrow[2 * a] <= g.ny && ic <= cg.nx && jc >= 1 && jc <= cg.ny;
            coff[a][c] = min(ic, cg.nx + 1) + min(max(jc, 0), cg.ny + 1) * (int)cg.sx;
        }
#pragma unroll
    for (int s4 = 0; s4 < 4; s4++) {
#pragma unroll
        for (int t = 0; t < RW; t++)
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                Q.X[s4][t][q] = 0.0f; Q.Iq[s4][t][q] = 0.0f; Q.Rq[s4][t][q] = 0.0f;
                Q.Cx[s4][t][q] = 0.0f; Q.Cy[s4][t][q] = 0.0f; Q.Cz[s4][t][q] = 0.0f;
            }
    }
    auto issue = [&](int slot, int P) {
        rec_issue<CPL, NW, RW, SV, C4>(pf[slot], b, rec, th, g, P);
        const int kcc = (P + 1) >> 1;
        pcin[slot] = P >= 1 && P <= g.nz && kcc >= 1 && kcc <= cg.nz;
        const int cpo = min(max(kcc, 0), cg.nz + 1) * (int)cg.sy;
#pragma unroll
        for (int a = 0; a < RW / 2; a++)
#pragma unroll
            for (int c = 0; c < CPL / 2; c++) pfx[slot][a][c] = (float)xc[coff[a][c] + cpo];
    };
    rec_prev_cz<CPL, NW, RW, C4>(Q, rec, th, g);
    // planes m0 - 2 .. m0 + REC_PD - 3 in flight before the first step (plane
    // m0 - 2 + d has slot 3 + d)
    sfor<REC_PD>([&](auto dc) {
        constexpr int d = decltype(dc)::value;
        issue((3 + d) & (REC_PD - 1), th.m0 - 2 + d);
    });

    auto step = [&](auto Sc, int P) {
        constexpr int S = decltype(Sc)::value;
        constexpr int S1 = (S + 3) & 3, S2 = (S + 2) & 3, S3 = (S + 1) & 3;
        constexpr int PB = S & (REC_PD - 1), EW = S & 1, ER = EW ^ 1;
        // S1: plane P: x' = x0 + 2 x_coarse on active cells
        rec_enter<S, CPL, NW, RW, C4>(Q, pf[PB], th, P >= 1 && P <= g.nz, L);
        sfor<RW>([&](auto tc) {
            constexpr int t = decltype(tc)::value;
            sfor<CPL>([&](auto qc) {
                constexpr int q = decltype(qc)::value;
                const bool act = th.rowok[t] && th.okx[q] && P >= 1 && P <= g.nz && pf[PB].rec[t][q].w >= 0.0f;
                const float w2 = (act && pcin[PB] && cok[t / 2][q / 2]) ? 2.0f * pfx[PB][t / 2][q / 2] : 0.0f;
                if constexpr (((q + t + S) & 1) == 1) Q.X[S][t][q] = Q.Rq[S][t][q] * Q.Iq[S][t][q] + w2;
                else Q.X[S][t][q] = w2;
            });
        });
        issue(PB, min(P + REC_PD, th.last + 2));   // past last + 2 (never used): the cached plane last + 2 again, no memory traffic
        sfor<CPL>([&](auto qc) {
            constexpr int q = decltype(qc)::value;
            if constexpr (((q + S) & 1) == 1) L.r.ex[EW][0][w][0][q / 2][lane] = Q.X[S][0][q];
            if constexpr (((q + RW - 1 + S) & 1) == 1) L.r.ex[EW][0][w][1][q / 2][lane] = Q.X[S][RW - 1][q];
        });
        // S2: plane P-1, even cells: x = (b + sum c x'_odd) / d
        if (P - 1 >= th.m0 - 1) rec_colour<0, true, CPL, NW, RW, S1, S, S2>(Q, L, ER, EW, 0, w, lane);
        // S3: plane R = P-2, odd cells: x = (b + sum c x_even) / d; output
        const int R = P - 2;
        if (R >= th.m0 && R <= th.last) {
            rec_colour<1, false, CPL, NW, RW, S2, S1, S3>(Q, L, ER, EW, 1, w, lane);
#pragma unroll
            for (int t = 0; t < RW; t++) {
                if (!th.tile_row[t]) continue;
                const int id = th.roff[t] + R * (int)g.sy;
                if (std::is_same<SV, float>::value && (th.full || th.okx[CPL - 1])) {
                    typename RowVec<CPL>::V v;
#pragma unroll
                    for (int q = 0; q < CPL; q++) v[q] = Q.X[S2][t][q];
                    *(typename RowVec<CPL>::V*)((float*)xout + id) = v;
                } else {
#pragma unroll
                    for (int q = 0; q < CPL; q++)
                        if (th.okx[q]) xout[id + q] = (SV)Q.X[S2][t][q];
                }
            }
        }
        __syncthreads();
    };
    for (int P = th.m0 - 2; P <= th.last + 2; P += 4) {
        step(IC<3>{}, P);
        if (P + 1 > th.last + 2) break;
        step(IC<0>{}, P + 1);
        if (P + 2 > th.last + 2) break;
        step(IC<1>{}, P + 2);
        if (P + 3 > th.last + 2) break;
        step(IC<2>{}, P + 3);
    }
}

template <int CPL, int NW, int RW, typename SV, typename C4>
__global__ __launch_bounds__(NW * 64) void k_rows_upc(const SV* __restrict__ b, const C4* __restrict__ rec, Geo g,
                                                      RowGeo rg, const SV* __restrict__ xc, Geo cg,
                                                      SV* __restrict__ xout, const int* skip) {
    d_rows_upc<CPL, NW, RW, SV, C4>(b, rec, g, rg, xc, cg, xout, skip, row_tile_of_block(rg, blockIdx.x));
}

}  // namespace uaamg
