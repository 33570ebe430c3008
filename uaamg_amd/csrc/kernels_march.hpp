// kernels_march.hpp -- fused V(1,1) level kernels, 2.5D form.
//
// Same arithmetic as the per-pass kernels (and the reference smooth /
// restrictResidual / prolongateAdd, uaamg_preconditioner_3d.cpp:123-181),
// organised for throughput:
//   * a workgroup owns an in-plane region of RX x RY cells (the tile plus a
//     2-cell ring) and marches along the slowest axis (k in 3D, j in 2D);
//   * every thread owns PPT fixed in-plane positions, so address and parity
//     work is done once, and keeps the values of its positions on the last
//     few planes in registers (the vertical neighbours of the stencil);
//   * only in-plane neighbour values go through LDS (one slot per plane of
//     the 3-plane window);
//   * the loads of plane P+2 are issued at step P (register prefetch), so
//     memory latency overlaps two steps of arithmetic.
// Step P: odd colour of plane P (pre-smooth b/d, or + the prolongation on the
// way up), even colour of plane P-1, then the residual + restriction (down) or
// the odd colour of the reverse sweep + output (up) of plane P-2.
#pragma once
#include "kernels_fused.hpp"

namespace uaamg {

// Per-level stencil state of one cell as carried in registers.
template <bool L0, typename CT> struct MCell;
template <typename CT> struct MCell<true, CT> { uint8_t nb; };
template <typename CT> struct MCell<false, CT> { CT cx, cy, cz, d; };   // cz: march coupling

template <int D, bool L0, typename CT>
struct MSten { CT ap, am, bp, bm, mp, mm, d; bool fluid; };

// Stencil of a level-0 cell from its neighbour byte (channels a = x,
// b = y (3D), m = march: z in 3D, y in 2D).
template <int D, typename CT>
__device__ __forceinline__ MSten<D, true, CT> msten0(uint8_t nb, const Consts<CT>& c, const Geo& g, int i, int j, int k) {
    MSten<D, true, CT> t;
    t.fluid = nb & NB_FLUID;
    t.ap = (nb & NB_XP) ? c.idx2 : CT(0);
    t.am = (nb & NB_XM) ? c.idx2 : CT(0);
    if (D == 3) {
        t.bp = (nb & NB_YP) ? c.idy2 : CT(0);
        t.bm = (nb & NB_YM) ? c.idy2 : CT(0);
        t.mp = (nb & NB_ZP) ? c.idz2 : CT(0);
        t.mm = (nb & NB_ZM) ? c.idz2 : CT(0);
    } else {
        t.bp = CT(0); t.bm = CT(0);
        t.mp = (nb & NB_YP) ? c.idy2 : CT(0);
        t.mm = (nb & NB_YM) ? c.idy2 : CT(0);
    }
    // reference order x+, x-, y+, y-, z+, z-
    CT dd = (D == 3) ? t.ap + t.am + t.bp + t.bm + t.mp + t.mm : t.ap + t.am + t.mp + t.mm;
    if (c.any_dirichlet && t.fluid) {
        CT ex = wall_excess0<D, CT>(c, g, i, j, k);
        if (ex != CT(0)) dd += ex;
    }
    t.d = t.fluid ? dd : CT(-1);
    return t;
}

// (b + nb) / d: exact IEEE division in the FP64 parity path; in FP32 the
// hardware reciprocal (1 ulp, deterministic) keeps the V-cycle a fixed
// linear operator at a fraction of the instruction cost
template <typename CT>
__device__ __forceinline__ CT mdiv(CT num, CT d) {
    if constexpr (std::is_same<CT, double>::value) return num / d;
    else return num * __builtin_amdgcn_rcpf(d);
}

// neighbour sum, reference order
template <int D, typename ST, typename CT>
__device__ __forceinline__ CT msum(const ST& t, CT xap, CT xam, CT xbp, CT xbm, CT xmp, CT xmm) {
    if (D == 3) return t.ap * xap + t.am * xam + t.bp * xbp + t.bm * xbm + t.mp * xmp + t.mm * xmm;
    return t.ap * xap + t.am * xam + t.mp * xmp + t.mm * xmm;
}

template <int D, bool L0, int RX, int RY, typename CT>
struct MarchLDS {
    static constexpr int RN = RX * RY;
    CT x[3][RN];
    CT ca[L0 ? 1 : 3][L0 ? 1 : RN];   // Lc: +a couplings of planes in the window
    CT cb[L0 ? 1 : 3][L0 ? 1 : RN];   // Lc: +b couplings (3D)
};

// Positions of one thread in the region and their global in-plane data.
template <int D, int RX, int RY, int PPT>
struct MPos {
    int e[PPT];        // region index ri + rj*RX
    int gi[PPT], gj[PPT];
    long off[PPT];     // in-plane global offset (i + j*sx in 3D, i in 2D)
    bool in[PPT];      // inside the grid in-plane
    bool r1[PPT], r0[PPT];
    __device__ void init(const Geo& g, int a0, int b0) {
        constexpr int NT = RX * RY / PPT;
#pragma unroll
        for (int u = 0; u < PPT; u++) {
            int t = threadIdx.x;
            int ri = t % RX;
            int rj = (D == 3) ? t / RX + (NT / RX) * u : 0;
            if (D == 2) ri = t + NT * u;
            e[u] = ri + rj * RX;
            gi[u] = a0 - 2 + ri;
            gj[u] = (D == 3) ? b0 - 2 + rj : 0;
            in[u] = gi[u] >= 1 && gi[u] <= g.nx && (D == 2 || (gj[u] >= 1 && gj[u] <= g.ny));
            off[u] = (long)gi[u] + ((D == 3) ? (long)gj[u] * g.sx : 0);
            r1[u] = ri >= 1 && ri <= RX - 2 && (D == 2 || (rj >= 1 && rj <= RY - 2));
            r0[u] = ri >= 2 && ri <= RX - 3 && (D == 2 || (rj >= 2 && rj <= RY - 3));
        }
    }
};

template <int D> __device__ __forceinline__ long mstride(const Geo& g) { return (D == 3) ? g.sy : g.sx; }
template <int D> __device__ __forceinline__ int mcount(const Geo& g) { return (D == 3) ? g.nz : g.ny; }

template <bool L0, typename CT, typename C4, int D>
__device__ __forceinline__ MCell<L0, CT> mload_cell(const uint8_t* __restrict__ nbr, const C4* __restrict__ rec,
                                                    long id, bool valid) {
    MCell<L0, CT> c;
    if constexpr (L0) c.nb = valid ? nbr[id] : (uint8_t)0;
    else {
        if (valid) {
            C4 r = rec[id];
            c.cx = (CT)r.x; c.cy = (D == 3) ? (CT)r.y : CT(0); c.cz = (D == 3) ? (CT)r.z : (CT)r.y; c.d = (CT)r.w;
        } else { c.cx = CT(0); c.cy = CT(0); c.cz = CT(0); c.d = CT(-1); }
    }
    return c;
}

// Stencil of cell (position u) at a plane: own cell state c, -a/-b couplings
// from the LDS slot of that plane (Lc), -m coupling from the previous plane's
// own state cprev (Lc).
template <int D, bool L0, int RX, int RY, typename CT, bool XD = false>
__device__ __forceinline__ MSten<D, L0, CT> mst(const MCell<L0, CT>& c, const MCell<L0, CT>& cprev,
                                                const MarchLDS<D, L0, RX, RY, CT>& S, int slot, int e,
                                                const Consts<CT>& k0, const Geo& g, int i, int j, int k) {
    if constexpr (L0) {
        (void)cprev; (void)S; (void)slot; (void)e;
        MSten<D, true, CT> t0 = msten0<D, CT>(c.nb, k0, g, i, j, k);
        MSten<D, L0, CT> t;
        t.ap = t0.ap; t.am = t0.am; t.bp = t0.bp; t.bm = t0.bm; t.mp = t0.mp; t.mm = t0.mm; t.d = t0.d; t.fluid = t0.fluid;
        return t;
    } else {
        MSten<D, L0, CT> t;
        t.d = c.d;
        t.fluid = c.d >= CT(0);
        t.ap = c.cx; t.am = S.ca[slot][e - 1];
        if (D == 3) { t.bp = c.cy; t.bm = S.cb[slot][e - RX]; }
        else { t.bp = CT(0); t.bm = CT(0); }
        t.mp = c.cz; t.mm = cprev.cz;
        // BF16 records: the exact diagonal where it carries no excess (snap_diag)
        if constexpr (XD) { if (t.fluid) t.d = snap_diag<BF16x4, CT>(c.d, t.ap + t.am + t.bp + t.bm + t.mp + t.mm); }
        return t;
    }
}

// ------------------------------------------------------------------ down
template <int D, bool L0, int RX, int RY, int PPT, typename VB, typename C4, typename VC, typename CC4>
__device__ __forceinline__ void d_down_m(const VB* __restrict__ b, FusedArgs<D, L0, typename CTOf<VB>::type, C4> A,
                                                VC* __restrict__ bc, const CC4* __restrict__ crec, Geo cg, int KC,
                                                const int* skip, int bx, int by, int bz) {
    using CT = typename CTOf<VB>::type;
    using Cell = MCell<L0, CT>;
    // BF16 records: the diagonal of a plane's cells is formed once its -x/-y
    // couplings are in LDS (after the plane's barrier), so the odd cells'
    // b/d of an entering plane moves behind that barrier
    constexpr bool XD = !L0 && std::is_same<C4, BF16x4>::value;
    constexpr int TX = RX - 4, TY = (D == 3) ? RY - 4 : 1, RN = RX * RY;
    static_assert(RX * RY / PPT == 256, "256 threads");
    if (skip && *skip) return;
    __shared__ MarchLDS<D, L0, RX, RY, CT> S;
    __shared__ typename std::conditional<std::is_same<CT, double>::value, double, float>::type sres[2][RN];
    __shared__ uint8_t sfl[2][RN];
    const Geo g = A.g;
    const int a0 = bx * TX + 1, b0 = (D == 3) ? by * TY + 1 : 0, m0 = bz * KC + 1;
    const int nm = mcount<D>(g);
    const long ms = mstride<D>(g);
    const int last = min(m0 + KC - 1, nm);
    MPos<D, RX, RY, PPT> pos;
    pos.init(g, a0, b0);

    // two prefetch slots, each consumed two steps after its loads are issued
    VB pfbA[PPT], pfbB[PPT];
    Cell pfcA[PPT], pfcB[PPT];
    auto issue = [&](VB (&pfb)[PPT], Cell (&pfc)[PPT], int m) {
#pragma unroll
        for (int u = 0; u < PPT; u++) {
            bool valid = pos.in[u] && m >= 1 && m <= nm;
            long id = pos.off[u] + (long)m * ms;
            pfb[u] = valid ? b[id] : VB(0);
            pfc[u] = mload_cell<L0, CT, C4, D>(A.nbr, A.rec, id, valid);
        }
    };
    issue(pfbA, pfcA, m0 - 2);
    issue(pfbB, pfcB, m0 - 1);
    CT qb[PPT][3];        // b at planes P-2, P-1, P
    Cell qc[PPT][4];      // cell state at planes P-3 .. P
    CT qx[PPT][4];        // x at planes P-3 .. P
#pragma unroll
    for (int u = 0; u < PPT; u++) {
        for (int q = 0; q < 3; q++) qb[u][q] = CT(0);
        for (int q = 0; q < 4; q++) { qx[u][q] = CT(0); qc[u][q] = mload_cell<L0, CT, C4, D>(A.nbr, A.rec, 0, false); }
        // the -z coupling of the first halo plane: plane m0 - 3's record
        if constexpr (XD) {
            const bool v3 = pos.in[u] && m0 - 3 >= 1 && m0 - 3 <= nm;
            qc[u][3] = mload_cell<L0, CT, C4, D>(A.nbr, A.rec, pos.off[u] + (long)(m0 - 3) * ms, v3);
        }
    }
    using RT = typename std::conditional<std::is_same<CT, double>::value, double, float>::type;

    auto body = [&](int P, VB (&pfb)[PPT], Cell (&pfc)[PPT]) {
        const int s = ((P % 3) + 3) % 3;
        // S1: plane P enters the window; odd cells get the pre-smooth b/d
#pragma unroll
        for (int u = 0; u < PPT; u++) {
            CT bv = (CT)pfb[u];
            Cell cv = pfc[u];
            qb[u][0] = qb[u][1]; qb[u][1] = qb[u][2]; qb[u][2] = bv;
            qc[u][0] = qc[u][1]; qc[u][1] = qc[u][2]; qc[u][2] = qc[u][3]; qc[u][3] = cv;
            qx[u][0] = qx[u][1]; qx[u][1] = qx[u][2]; qx[u][2] = qx[u][3];
            int gk = (D == 3) ? P : 0, gjj = (D == 3) ? pos.gj[u] : P;
            if constexpr (!XD) {
                MSten<D, L0, CT> t;
                if constexpr (L0) t = mst<D, L0, RX, RY, CT>(cv, cv, S, s, pos.e[u], A.c, g, pos.gi[u], gjj, gk);
                else { t.d = cv.d; }
                CT xo = (((pos.gi[u] + gjj + gk) & 1) == 1 && t.d >= CT(1e-30)) ? mdiv<CT>(bv + CT(0), t.d) : CT(0);
                qx[u][3] = xo;
                S.x[s][pos.e[u]] = xo;
            }
            if constexpr (!L0) { S.ca[s][pos.e[u]] = cv.cx; if (D == 3) S.cb[s][pos.e[u]] = cv.cy; }
        }
        issue(pfb, pfc, min(P + 2, last + 2));   // past last + 2 (never used): the cached plane last + 2 again
        __syncthreads();
        if constexpr (XD) {   // odd cells of plane P: b/d with the formed diagonal
#pragma unroll
            for (int u = 0; u < PPT; u++) {
                int gk = (D == 3) ? P : 0, gjj = (D == 3) ? pos.gj[u] : P;
                // the outer ring has no -x/-y neighbour in the region: stored diagonal
                const CT d = pos.r1[u] ? mst<D, L0, RX, RY, CT, XD>(qc[u][3], qc[u][2], S, s, pos.e[u], A.c, g,
                                                                    pos.gi[u], gjj, gk).d
                                       : qc[u][3].d;
                CT xo = (((pos.gi[u] + gjj + gk) & 1) == 1 && d >= CT(1e-30)) ? mdiv<CT>(qb[u][2] + CT(0), d) : CT(0);
                qx[u][3] = xo;
                S.x[s][pos.e[u]] = xo;
            }
        }
        // S2: forward sweep, even colour, plane Q = P-1 on the 1-ring region
        const int Q = P - 1;
        if (Q >= m0 - 1) {
            const int sq = ((Q % 3) + 3) % 3;
#pragma unroll
            for (int u = 0; u < PPT; u++) {
                if (!pos.r1[u]) continue;
                int gk = (D == 3) ? Q : 0, gjj = (D == 3) ? pos.gj[u] : Q;
                if (((pos.gi[u] + gjj + gk) & 1) != 0) continue;
                MSten<D, L0, CT> t = mst<D, L0, RX, RY, CT, XD>(qc[u][2], qc[u][1], S, sq, pos.e[u], A.c, g, pos.gi[u], gjj, gk);
                if (!(t.d >= CT(1e-30))) continue;
                const int e = pos.e[u];
                CT nb = msum<D>(t, S.x[sq][e + 1], S.x[sq][e - 1], D == 3 ? S.x[sq][e + RX] : CT(0),
                                D == 3 ? S.x[sq][e - RX] : CT(0), qx[u][3], qx[u][1]);
                CT xe = mdiv<CT>(qb[u][1] + nb, t.d);
                qx[u][2] = xe;
                S.x[sq][e] = xe;
            }
        }
        __syncthreads();
        // S3: residual of plane R = P-2 on the tile
        const int R = P - 2;
        const bool doR = (R >= m0 && R <= last);
        if (doR) {
            const int sr = ((R % 3) + 3) % 3;
#pragma unroll
            for (int u = 0; u < PPT; u++) {
                if (!pos.r0[u]) continue;
                int gk = (D == 3) ? R : 0, gjj = (D == 3) ? pos.gj[u] : R;
                MSten<D, L0, CT> t = mst<D, L0, RX, RY, CT, XD>(qc[u][1], qc[u][0], S, sr, pos.e[u], A.c, g, pos.gi[u], gjj, gk);
                const int e = pos.e[u];
                RT res = RT(0);
                if (t.fluid) {
                    RT Ax = (RT)t.d * (RT)qx[u][1] - (RT)t.ap * (RT)S.x[sr][e + 1] - (RT)t.am * (RT)S.x[sr][e - 1];
                    if (D == 3) Ax = Ax - (RT)t.bp * (RT)S.x[sr][e + RX] - (RT)t.bm * (RT)S.x[sr][e - RX];
                    Ax = Ax - (RT)t.mp * (RT)qx[u][2] - (RT)t.mm * (RT)qx[u][0];
                    res = (RT)qb[u][0] - Ax;
                }
                sres[R & 1][e] = res;
                sfl[R & 1][e] = t.fluid;
            }
        }
        __syncthreads();
        // S4: restrict the plane pair (R-1, R) into coarse plane R/2 (solid
        // coarse cells get the sum too; no level ever reads b there)
        if (doR && (R & 1) == 0 && R - 1 >= m0) {
            constexpr int CA = TX / 2, CB = (D == 3) ? TY / 2 : 1;
            const int mc = R >> 1;
            for (int q = threadIdx.x; q < CA * CB; q += 256) {
                int ca = q % CA, cb = q / CA;
                int ic = (a0 + 1) / 2 + ca, jc = (D == 3) ? (b0 + 1) / 2 + cb : 0;
                if (ic > cg.nx || (D == 3 && jc > cg.ny)) continue;
                long cid = (long)ic + ((D == 3) ? (long)jc * cg.sx + (long)mc * cg.sy : (long)mc * cg.sx);
                RT sum = RT(0);
                for (int di = 0; di < 2; di++)
                    for (int dj = 0; dj < (D == 3 ? 2 : 1); dj++)
                        for (int dk = 0; dk < 2; dk++) {
                            int e = (2 + 2 * ca + di) + ((D == 3) ? (2 + 2 * cb + dj) * RX : 0);
                            int sl = (R - 1 + dk) & 1;
                            if (sfl[sl][e]) sum += sres[sl][e];
                        }
                bc[cid] = (VC)sum;
            }
        }
    };
    for (int P = m0 - 2; P <= last + 2; P += 2) {
        body(P, pfbA, pfcA);
        if (P + 1 > last + 2) break;
        body(P + 1, pfbB, pfcB);
    }
    (void)crec;
}
template <int D, bool L0, int RX, int RY, int PPT, typename VB, typename C4, typename VC, typename CC4>
__global__ __launch_bounds__(256) void k_down_m(const VB* __restrict__ b, FusedArgs<D, L0, typename CTOf<VB>::type, C4> A,
                                                VC* __restrict__ bc, const CC4* __restrict__ crec, Geo cg, int KC,
                                                const int* skip) {
    d_down_m<D, L0, RX, RY, PPT, VB, C4, VC, CC4>(b, A, bc, crec, cg, KC, skip, blockIdx.x, blockIdx.y, blockIdx.z);
}

// ------------------------------------------------------------------ up
template <int D, bool L0, int RX, int RY, int PPT, bool REDUCE, typename VB, typename C4, typename VX, typename VCX>
__device__ __forceinline__ void d_up_m(const VB* __restrict__ b, FusedArgs<D, L0, typename CTOf<VB>::type, C4> A,
                                              const VCX* __restrict__ xc, Geo cg, VX* __restrict__ xout, int KC,
                                              const int* skip, double* partials, int bx, int by, int bz) {
    using CT = typename CTOf<VB>::type;
    using Cell = MCell<L0, CT>;
    // BF16 records: as in d_down_m, the entering plane's odd b/d and the
    // prolongation move behind the plane's barrier
    constexpr bool XD = !L0 && std::is_same<C4, BF16x4>::value;
    constexpr int TX = RX - 4, TY = (D == 3) ? RY - 4 : 1;
    static_assert(RX * RY / PPT == 256, "256 threads");
    if (skip && *skip) return;
    __shared__ MarchLDS<D, L0, RX, RY, CT> S;
    const Geo g = A.g;
    const int a0 = bx * TX + 1, b0 = (D == 3) ? by * TY + 1 : 0, m0 = bz * KC + 1;
    const int nm = mcount<D>(g), cnm = mcount<D>(cg);
    const long ms = mstride<D>(g), cms = mstride<D>(cg);
    const int last = min(m0 + KC - 1, nm);
    MPos<D, RX, RY, PPT> pos;
    pos.init(g, a0, b0);
    long coff[PPT];
    bool cin[PPT];
#pragma unroll
    for (int u = 0; u < PPT; u++) {
        int ic = (pos.gi[u] + 1) >> 1, jc = (D == 3) ? (pos.gj[u] + 1) >> 1 : 0;
        cin[u] = pos.in[u] && ic >= 1 && ic <= cg.nx && (D == 2 || (jc >= 1 && jc <= cg.ny));
        coff[u] = (long)ic + ((D == 3) ? (long)jc * cg.sx : 0);
    }

    VB pfbA[PPT], pfbB[PPT];
    Cell pfcA[PPT], pfcB[PPT];
    CT pfxA[PPT], pfxB[PPT];
    auto issue = [&](VB (&pfb)[PPT], Cell (&pfc)[PPT], CT (&pfx)[PPT], int m) {
#pragma unroll
        for (int u = 0; u < PPT; u++) {
            bool valid = pos.in[u] && m >= 1 && m <= nm;
            long id = pos.off[u] + (long)m * ms;
            pfb[u] = valid ? b[id] : VB(0);
            pfc[u] = mload_cell<L0, CT, C4, D>(A.nbr, A.rec, id, valid);
            int mc = (m + 1) >> 1;
            bool cv = valid && cin[u] && mc >= 1 && mc <= cnm;
            pfx[u] = cv ? (CT)xc[coff[u] + (long)mc * cms] : CT(0);
        }
    };
    issue(pfbA, pfcA, pfxA, m0 - 2);
    issue(pfbB, pfcB, pfxB, m0 - 1);
    CT qb[PPT][3];
    Cell qc[PPT][4];
    CT qx[PPT][4];
#pragma unroll
    for (int u = 0; u < PPT; u++) {
        for (int q = 0; q < 3; q++) qb[u][q] = CT(0);
        for (int q = 0; q < 4; q++) { qx[u][q] = CT(0); qc[u][q] = mload_cell<L0, CT, C4, D>(A.nbr, A.rec, 0, false); }
        if constexpr (XD) {   // the -z coupling of the first halo plane
            const bool v3 = pos.in[u] && m0 - 3 >= 1 && m0 - 3 <= nm;
            qc[u][3] = mload_cell<L0, CT, C4, D>(A.nbr, A.rec, pos.off[u] + (long)(m0 - 3) * ms, v3);
        }
    }
    CT xcs[PPT];   // coarse x of the entering plane (XD: used behind the barrier)
    double red[3] = {0.0, 0.0, 0.0};

    auto body = [&](int P, VB (&pfb)[PPT], Cell (&pfc)[PPT], CT (&pfx)[PPT]) {
        const int s = ((P % 3) + 3) % 3;
        // S1: plane P: pre-smooth odd values + prolongation on fluid cells
#pragma unroll
        for (int u = 0; u < PPT; u++) {
            CT bv = (CT)pfb[u];
            Cell cv = pfc[u];
            CT xcv = pfx[u];
            qb[u][0] = qb[u][1]; qb[u][1] = qb[u][2]; qb[u][2] = bv;
            qc[u][0] = qc[u][1]; qc[u][1] = qc[u][2]; qc[u][2] = qc[u][3]; qc[u][3] = cv;
            qx[u][0] = qx[u][1]; qx[u][1] = qx[u][2]; qx[u][2] = qx[u][3];
            int gk = (D == 3) ? P : 0, gjj = (D == 3) ? pos.gj[u] : P;
            if constexpr (!XD) {
                MSten<D, L0, CT> t;
                if constexpr (L0) t = mst<D, L0, RX, RY, CT>(cv, cv, S, s, pos.e[u], A.c, g, pos.gi[u], gjj, gk);
                else { t.d = cv.d; t.fluid = cv.d >= CT(0); }
                CT v = (((pos.gi[u] + gjj + gk) & 1) == 1 && t.d >= CT(1e-30)) ? mdiv<CT>(bv + CT(0), t.d) : CT(0);
                if (t.d >= CT(0)) v = v + CT(2) * xcv;
                qx[u][3] = v;
                S.x[s][pos.e[u]] = v;
            } else {
                xcs[u] = xcv;
            }
            if constexpr (!L0) { S.ca[s][pos.e[u]] = cv.cx; if (D == 3) S.cb[s][pos.e[u]] = cv.cy; }
        }
        issue(pfb, pfc, pfx, min(P + 2, last + 2));   // past last + 2 (never used): the cached plane last + 2 again
        __syncthreads();
        if constexpr (XD) {   // plane P: odd b/d with the formed diagonal, + 2 x_coarse on fluid cells
#pragma unroll
            for (int u = 0; u < PPT; u++) {
                int gk = (D == 3) ? P : 0, gjj = (D == 3) ? pos.gj[u] : P;
                // the outer ring has no -x/-y neighbour in the region: stored diagonal
                const CT d = pos.r1[u] ? mst<D, L0, RX, RY, CT, XD>(qc[u][3], qc[u][2], S, s, pos.e[u], A.c, g,
                                                                    pos.gi[u], gjj, gk).d
                                       : qc[u][3].d;
                CT v = (((pos.gi[u] + gjj + gk) & 1) == 1 && d >= CT(1e-30)) ? mdiv<CT>(qb[u][2] + CT(0), d) : CT(0);
                if (d >= CT(0)) v = v + CT(2) * xcs[u];
                qx[u][3] = v;
                S.x[s][pos.e[u]] = v;
            }
        }
        // S2: reverse sweep, even colour, plane Q = P-1 on the 1-ring region
        const int Q = P - 1;
        if (Q >= m0 - 1) {
            const int sq = ((Q % 3) + 3) % 3;
#pragma unroll
            for (int u = 0; u < PPT; u++) {
                if (!pos.r1[u]) continue;
                int gk = (D == 3) ? Q : 0, gjj = (D == 3) ? pos.gj[u] : Q;
                if (((pos.gi[u] + gjj + gk) & 1) != 0) continue;
                MSten<D, L0, CT> t = mst<D, L0, RX, RY, CT, XD>(qc[u][2], qc[u][1], S, sq, pos.e[u], A.c, g, pos.gi[u], gjj, gk);
                if (!(t.d >= CT(1e-30))) continue;
                const int e = pos.e[u];
                CT nb = msum<D>(t, S.x[sq][e + 1], S.x[sq][e - 1], D == 3 ? S.x[sq][e + RX] : CT(0),
                                D == 3 ? S.x[sq][e - RX] : CT(0), qx[u][3], qx[u][1]);
                CT xe = mdiv<CT>(qb[u][1] + nb, t.d);
                qx[u][2] = xe;
                S.x[sq][e] = xe;
            }
        }
        __syncthreads();
        // S3: reverse sweep, odd colour, plane R = P-2 on the tile; output
        const int R = P - 2;
        if (R >= m0 && R <= last) {
            const int sr = ((R % 3) + 3) % 3;
#pragma unroll
            for (int u = 0; u < PPT; u++) {
                if (!pos.r0[u] || !pos.in[u]) continue;
                int gk = (D == 3) ? R : 0, gjj = (D == 3) ? pos.gj[u] : R;
                MSten<D, L0, CT> t = mst<D, L0, RX, RY, CT, XD>(qc[u][1], qc[u][0], S, sr, pos.e[u], A.c, g, pos.gi[u], gjj, gk);
                const int e = pos.e[u];
                CT v = qx[u][1];
                if (((pos.gi[u] + gjj + gk) & 1) == 1 && t.d >= CT(1e-30)) {
                    CT nb = msum<D>(t, S.x[sr][e + 1], S.x[sr][e - 1], D == 3 ? S.x[sr][e + RX] : CT(0),
                                    D == 3 ? S.x[sr][e - RX] : CT(0), qx[u][2], qx[u][0]);
                    v = mdiv<CT>(qb[u][0] + nb, t.d);
                }
                VX out = (VX)v;
                xout[pos.off[u] + (long)R * ms] = out;
                if (REDUCE && t.fluid) {
                    double zz = (double)out, rr = (double)qb[u][0];
                    red[0] += zz; red[1] += rr * zz; red[2] += rr;
                }
            }
        }
        __syncthreads();
    };
    for (int P = m0 - 2; P <= last + 2; P += 2) {
        body(P, pfbA, pfcA, pfxA);
        if (P + 1 > last + 2) break;
        body(P + 1, pfbB, pfcB, pfxB);
    }
    if constexpr (REDUCE) block_store<3>(red, partials);
    (void)TY;
}
template <int D, bool L0, int RX, int RY, int PPT, bool REDUCE, typename VB, typename C4, typename VX, typename VCX>
__global__ __launch_bounds__(256) void k_up_m(const VB* __restrict__ b, FusedArgs<D, L0, typename CTOf<VB>::type, C4> A,
                                              const VCX* __restrict__ xc, Geo cg, VX* __restrict__ xout, int KC,
                                              const int* skip, double* partials) {
    d_up_m<D, L0, RX, RY, PPT, REDUCE, VB, C4, VX, VCX>(b, A, xc, cg, xout, KC, skip, partials, blockIdx.x, blockIdx.y, blockIdx.z);
}

}  // namespace uaamg
