// kernels_fused.hpp -- fused V(1,1) kernels: one launch per level on the way
// down (pre-smooth + residual + restriction) and one on the way up
// (prolongation + post-smooth), plus a single-workgroup coarsest solve.
//
// Exactness.  Every level's solution starts at zero in the reference V-cycle
// (uaamg_preconditioner_3d.cpp:192, :221), so the forward pre-smooth is a
// local function of b: odd cells get b/diag (their neighbours are still 0),
// even cells (b + sum c*x_odd)/diag.  The kernels recompute these values in a
// 2-cell halo instead of storing them, apply the same operations in the same
// order as the reference (uaamg_preconditioner_3d.cpp:123-181), and colour
// passes read only the other colour, so the result equals the per-pass
// sequence cell by cell.
//
// Layout of the work.  A workgroup owns a TX x TY x TZ tile of fine cells
// (TX x TY in 2D), aligned to coarse cells.  It loads the tile plus a 2-cell
// ring in one burst (every thread issues its loads back to back, so many
// requests are in flight), keeps x, b and the stencil in LDS, and runs the
// colour passes on shrinking regions: odd cells on the 2-ring, even cells on
// the 1-ring, then the residual + restriction (down) or the final odd pass
// (up) on the tile.  Two barriers per tile.  HBM traffic is one read of b and
// of the stencil per cell plus the coarse input/output; halo re-reads hit L2.
#pragma once
#include "kernels_cg.hpp"

namespace uaamg {

// Stencil source of a level for the fused kernels.
//   L0: neighbour bytes + uniform constants (level 0)
//   Lc: records {cx, cy, cz, diag} (levels >= 1)
template <int D, bool L0, typename CT, typename C4>
struct FusedArgs {
    Geo g;
    const uint8_t* nbr;    // L0
    const C4* rec;         // Lc
    Consts<CT> c;          // L0 constants
};

template <int D, bool L0, int TX, int TY, int TZ, typename CT>
struct TileLDS {
    static constexpr int RX = TX + 4, RY = TY + 4, RZ = (D == 3) ? TZ + 4 : 1;
    static constexpr int RN = RX * RY * RZ;
    static constexpr int SY = RX, SZ = RX * RY;     // LDS strides
    CT x[RN];
    CT b[RN];
    uint8_t n[L0 ? RN : 1];          // L0 neighbour bytes
    CT d[L0 ? 1 : RN];               // Lc diagonal (< 0: solid / outside)
    CT cp[3][L0 ? 1 : RN];           // Lc couplings +x, +y, +z (2D: +x, +y)
};

// Diagonal and couplings of LDS cell e (a, b, m = tile-region coordinates
// shifted to global i, j, k for the Dirichlet excess).
template <typename CT>
struct St { CT xp, xm, yp, ym, zp, zm, d; bool fluid; };

template <int D, bool L0, int TX, int TY, int TZ, typename CT>
__device__ __forceinline__ St<CT> tile_sten(const TileLDS<D, L0, TX, TY, TZ, CT>& S, int e, const Consts<CT>& c,
                                            const Geo& g, int i, int j, int k) {
    using L = TileLDS<D, L0, TX, TY, TZ, CT>;
    St<CT> t;
    if constexpr (L0) {
        uint8_t nb = S.n[e];
        t.fluid = nb & NB_FLUID;
        t.xp = (nb & NB_XP) ? c.idx2 : CT(0);
        t.xm = (nb & NB_XM) ? c.idx2 : CT(0);
        t.yp = (nb & NB_YP) ? c.idy2 : CT(0);
        t.ym = (nb & NB_YM) ? c.idy2 : CT(0);
        t.zp = (D == 3 && (nb & NB_ZP)) ? c.idz2 : CT(0);
        t.zm = (D == 3 && (nb & NB_ZM)) ? c.idz2 : CT(0);
        CT dd = t.xp + t.xm + t.yp + t.ym;
        if (D == 3) dd = dd + t.zp + t.zm;
        if (c.any_dirichlet && t.fluid) {
            CT ex = wall_excess0<D, CT>(c, g, i, j, k);
            if (ex != CT(0)) dd += ex;
        }
        t.d = t.fluid ? dd : CT(-1);
    } else {
        t.d = S.d[e];
        t.fluid = t.d >= CT(0);
        t.xp = S.cp[0][e]; t.xm = S.cp[0][e - 1];
        t.yp = S.cp[1][e]; t.ym = S.cp[1][e - L::SY];
        if (D == 3) { t.zp = S.cp[2][e]; t.zm = S.cp[2][e - L::SZ]; }
        else { t.zp = CT(0); t.zm = CT(0); }
    }
    return t;
}

// Neighbour sum in the reference order: x+, x-, y+, y-, z+, z-.
template <int D, int SY, int SZ, typename CT>
__device__ __forceinline__ CT tile_nb(const St<CT>& t, const CT* x, int e) {
    CT s = t.xp * x[e + 1] + t.xm * x[e - 1] + t.yp * x[e + SY] + t.ym * x[e - SY];
    if (D == 3) s = s + t.zp * x[e + SZ] + t.zm * x[e - SZ];
    return s;
}

// Load the tile + 2-ring region: b, stencil; x = pre-smoothed odd value b/d
// (even cells 0).  Cells outside the grid: b = 0, not fluid.  All global
// loads of a thread are issued before any of them is used (register staging),
// so a workgroup has its whole region in flight at once.
template <int D, bool L0, int TX, int TY, int TZ, typename CT, typename VB, typename C4>
__device__ __forceinline__ void tile_load(TileLDS<D, L0, TX, TY, TZ, CT>& S, const VB* __restrict__ bsrc,
                                          const FusedArgs<D, L0, CT, C4>& A, int i0, int j0, int k0) {
    using L = TileLDS<D, L0, TX, TY, TZ, CT>;
    constexpr int NT = 256, PER = (L::RN + NT - 1) / NT;
    const Geo& g = A.g;
    VB bv[PER];
    uint8_t nbv[L0 ? PER : 1];
    C4 rv[L0 ? 1 : PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
        int e = threadIdx.x + u * NT;
        int ri = e % L::RX, rest = e / L::RX;
        int rj = rest % L::RY, rk = rest / L::RY;
        int i = i0 - 2 + ri, j = j0 - 2 + rj, k = (D == 3) ? k0 - 2 + rk : 0;
        bool valid = e < L::RN && i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && (D == 2 || (k >= 1 && k <= g.nz));
        long id = valid ? gidx(g, i, j, k) : 0;
        bv[u] = valid ? bsrc[id] : VB(0);
        if constexpr (L0) nbv[u] = valid ? A.nbr[id] : (uint8_t)0;
        else {
            if (valid) rv[u] = A.rec[id];
            else { rv[u].x = 0; rv[u].y = 0; rv[u].z = 0; rv[u].w = -1; }
        }
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
        int e = threadIdx.x + u * NT;
        if (e >= L::RN) break;
        int ri = e % L::RX, rest = e / L::RX;
        int rj = rest % L::RY, rk = rest / L::RY;
        int i = i0 - 2 + ri, j = j0 - 2 + rj, k = (D == 3) ? k0 - 2 + rk : 0;
        CT b = (CT)bv[u];
        CT dv;
        if constexpr (L0) {
            S.n[e] = nbv[u];
            St<CT> t = tile_sten<D, L0, TX, TY, TZ, CT>(S, e, A.c, g, i, j, k);
            dv = t.d;
        } else {
            dv = (CT)rv[u].w;
            S.d[e] = dv;
            S.cp[0][e] = (CT)rv[u].x; S.cp[1][e] = (CT)rv[u].y; S.cp[2][e] = (D == 3) ? (CT)rv[u].z : CT(0);
        }
        S.b[e] = b;
        S.x[e] = (((i + j + k) & 1) == 1 && dv >= CT(1e-30)) ? (b + CT(0)) / dv : CT(0);
    }
}

// ------------------------------------------------------------------ down
// b_{l+1} = R (b_l - A_l S_fwd(b_l)), S_fwd = one forward RBGS sweep from 0.
template <int D, bool L0, int TX, int TY, int TZ, typename VB, typename C4, typename VC, typename CC4>
__global__ __launch_bounds__(256) void k_down(const VB* __restrict__ b, FusedArgs<D, L0, typename CTOf<VB>::type, C4> A,
                                              VC* __restrict__ bc, const CC4* __restrict__ crec, Geo cg,
                                              const int* skip) {
    using CT = typename CTOf<VB>::type;
    using L = TileLDS<D, L0, TX, TY, TZ, CT>;
    constexpr int RX = L::RX, RY = L::RY, RZ = L::RZ, SY = L::SY, SZ = L::SZ;
    if (skip && *skip) return;
    __shared__ L S;
    const Geo g = A.g;
    const int i0 = blockIdx.x * TX + 1, j0 = blockIdx.y * TY + 1, k0 = (D == 3) ? blockIdx.z * TZ + 1 : 0;
    tile_load<D, L0, TX, TY, TZ, CT, VB, C4>(S, b, A, i0, j0, k0);
    __syncthreads();
    // forward sweep, even colour, on the 1-ring region
    constexpr int QX = TX + 2, QY = TY + 2, QZ = (D == 3) ? TZ + 2 : 1;
    for (int q = threadIdx.x; q < QX * QY * QZ; q += blockDim.x) {
        int ri = q % QX + 1, rest = q / QX;
        int rj = rest % QY + 1, rk = (D == 3) ? rest / QY + 1 : 0;
        int i = i0 - 2 + ri, j = j0 - 2 + rj, k = (D == 3) ? k0 - 2 + rk : 0;
        if (((i + j + k) & 1) != 0) continue;
        int e = ri + rj * SY + rk * SZ;
        St<CT> t = tile_sten<D, L0, TX, TY, TZ, CT>(S, e, A.c, g, i, j, k);
        if (!(t.d >= CT(1e-30))) continue;
        S.x[e] = (S.b[e] + tile_nb<D, SY, SZ, CT>(t, S.x, e)) / t.d;
    }
    __syncthreads();
    // residual of the 8 (4) children of each coarse cell, summed in the
    // reference order (di, dj, dk), in FP64
    constexpr int CX = TX / 2, CY = TY / 2, CZ = (D == 3) ? TZ / 2 : 1;
    for (int q = threadIdx.x; q < CX * CY * CZ; q += blockDim.x) {
        int ci = q % CX, rest = q / CX;
        int cj = rest % CY, ck = rest / CY;
        int ic = (i0 + 1) / 2 + ci, jc = (j0 + 1) / 2 + cj, kc = (D == 3) ? (k0 + 1) / 2 + ck : 0;
        if (ic > cg.nx || jc > cg.ny || (D == 3 && kc > cg.nz)) continue;
        long cid = gidx(cg, ic, jc, kc);
        if ((double)crec[cid].w < 0) { bc[cid] = (VC)0; continue; }
        double sum = 0.0;
        for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++)
                for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++) {
                    int ri = 2 + 2 * ci + di, rj = 2 + 2 * cj + dj, rk = (D == 3) ? 2 + 2 * ck + dk : 0;
                    int e = ri + rj * SY + rk * SZ;
                    int i = i0 - 2 + ri, j = j0 - 2 + rj, k = (D == 3) ? k0 - 2 + rk : 0;
                    St<CT> t = tile_sten<D, L0, TX, TY, TZ, CT>(S, e, A.c, g, i, j, k);
                    if (!t.fluid) continue;
                    double Ax = (double)t.d * (double)S.x[e] - (double)t.xp * (double)S.x[e + 1]
                              - (double)t.xm * (double)S.x[e - 1] - (double)t.yp * (double)S.x[e + SY]
                              - (double)t.ym * (double)S.x[e - SY];
                    if (D == 3) Ax = Ax - (double)t.zp * (double)S.x[e + SZ] - (double)t.zm * (double)S.x[e - SZ];
                    sum += (double)S.b[e] - Ax;
                }
        bc[cid] = (VC)sum;
    }
    (void)RX; (void)RY; (void)RZ;
}

// ------------------------------------------------------------------ up
// x_l = S_rev(S_fwd(b_l) + 2 P x_{l+1}); level 0 also reduces sum z, sum r.z, sum r.
template <int D, bool L0, int TX, int TY, int TZ, bool REDUCE, typename VB, typename C4, typename VX, typename VCX>
__global__ __launch_bounds__(256) void k_up(const VB* __restrict__ b, FusedArgs<D, L0, typename CTOf<VB>::type, C4> A,
                                            const VCX* __restrict__ xc, Geo cg, VX* __restrict__ xout,
                                            const int* skip, double* partials, Ctrl* ctrl) {
    using CT = typename CTOf<VB>::type;
    using L = TileLDS<D, L0, TX, TY, TZ, CT>;
    constexpr int RX = L::RX, RY = L::RY, SY = L::SY, SZ = L::SZ;
    constexpr int PX = TX / 2 + 2, PY = TY / 2 + 2, PZ = (D == 3) ? TZ / 2 + 2 : 1;
    if (skip && *skip) return;
    __shared__ L S;
    __shared__ CT sxc[PX * PY * PZ];
    const Geo g = A.g;
    const int i0 = blockIdx.x * TX + 1, j0 = blockIdx.y * TY + 1, k0 = (D == 3) ? blockIdx.z * TZ + 1 : 0;
    // coarse solution on the parents of the region (coarse tile + 1 ring); 0 outside
    const int pi0 = (i0 + 1) / 2 - 1, pj0 = (j0 + 1) / 2 - 1, pk0 = (D == 3) ? (k0 + 1) / 2 - 1 : 0;
    {
        constexpr int PN = PX * PY * PZ, PPER = (PN + 255) / 256;
        VCX cv[PPER];
#pragma unroll
        for (int u = 0; u < PPER; u++) {
            int q = threadIdx.x + u * 256;
            int pi = q % PX, rest = q / PX;
            int pj = rest % PY, pk = rest / PY;
            int ic = pi0 + pi, jc = pj0 + pj, kc = (D == 3) ? pk0 + pk : 0;
            bool valid = q < PN && ic >= 1 && ic <= cg.nx && jc >= 1 && jc <= cg.ny && (D == 2 || (kc >= 1 && kc <= cg.nz));
            cv[u] = valid ? xc[gidx(cg, ic, jc, kc)] : VCX(0);
        }
#pragma unroll
        for (int u = 0; u < PPER; u++) {
            int q = threadIdx.x + u * 256;
            if (q < PN) sxc[q] = (CT)cv[u];
        }
    }
    tile_load<D, L0, TX, TY, TZ, CT, VB, C4>(S, b, A, i0, j0, k0);
    __syncthreads();
    // prolongation x += 2 x_c on fluid cells of the 2-ring region (the value
    // of even active cells is recomputed below without being read)
    for (int e = threadIdx.x; e < L::RN; e += blockDim.x) {
        int ri = e % RX, rest = e / RX;
        int rj = rest % RY, rk = rest / RY;
        int i = i0 - 2 + ri, j = j0 - 2 + rj, k = (D == 3) ? k0 - 2 + rk : 0;
        bool fluid;
        if constexpr (L0) fluid = S.n[e] & NB_FLUID;
        else fluid = S.d[e] >= CT(0);
        if (!fluid) continue;
        int pi = ((i + 1) >> 1) - pi0, pj = ((j + 1) >> 1) - pj0, pk = (D == 3) ? ((k + 1) >> 1) - pk0 : 0;
        S.x[e] = S.x[e] + CT(2) * sxc[pi + PX * (pj + PY * pk)];
    }
    __syncthreads();
    // reverse sweep, even colour, on the 1-ring region
    constexpr int QX = TX + 2, QY = TY + 2, QZ = (D == 3) ? TZ + 2 : 1;
    for (int q = threadIdx.x; q < QX * QY * QZ; q += blockDim.x) {
        int ri = q % QX + 1, rest = q / QX;
        int rj = rest % QY + 1, rk = (D == 3) ? rest / QY + 1 : 0;
        int i = i0 - 2 + ri, j = j0 - 2 + rj, k = (D == 3) ? k0 - 2 + rk : 0;
        if (((i + j + k) & 1) != 0) continue;
        int e = ri + rj * SY + rk * SZ;
        St<CT> t = tile_sten<D, L0, TX, TY, TZ, CT>(S, e, A.c, g, i, j, k);
        if (!(t.d >= CT(1e-30))) continue;
        S.x[e] = (S.b[e] + tile_nb<D, SY, SZ, CT>(t, S.x, e)) / t.d;
    }
    __syncthreads();
    // reverse sweep, odd colour, on the tile; write out (and reduce on level 0)
    double red[3] = {0.0, 0.0, 0.0};
    constexpr int TZo = (D == 3) ? TZ : 1;
    for (int q = threadIdx.x; q < TX * TY * TZo; q += blockDim.x) {
        int ri = q % TX + 2, rest = q / TX;
        int rj = rest % TY + 2, rk = (D == 3) ? rest / TY + 2 : 0;
        int i = i0 - 2 + ri, j = j0 - 2 + rj, k = (D == 3) ? k0 - 2 + rk : 0;
        if (i > g.nx || j > g.ny || (D == 3 && k > g.nz)) continue;
        int e = ri + rj * SY + rk * SZ;
        St<CT> t = tile_sten<D, L0, TX, TY, TZ, CT>(S, e, A.c, g, i, j, k);
        CT v = S.x[e];
        if (((i + j + k) & 1) == 1 && t.d >= CT(1e-30)) v = (S.b[e] + tile_nb<D, SY, SZ, CT>(t, S.x, e)) / t.d;
        VX out = (VX)v;
        xout[gidx(g, i, j, k)] = out;
        if (REDUCE && t.fluid) {
            double zz = (double)out, rr = (double)S.b[e];
            red[0] += zz; red[1] += rr * zz; red[2] += rr;
        }
    }
    if constexpr (REDUCE) block_store<3>(red, partials);
    (void)ctrl;
}

// ------------------------------------------------------------------ coarsest
// 2*sweeps+1 alternating colour passes (odd first) from x = 0: the
// reference's sweeps x (forward, reverse) pairs with the repeated colour at
// each pair boundary dropped (recomputing a colour from an unchanged other
// colour reproduces the same values).  Whole level in LDS, one workgroup.
template <int D, int CMAX, typename VB, typename C4>
__global__ __launch_bounds__(256) void k_coarsest(VB* __restrict__ xout, const VB* __restrict__ b,
                                                  const C4* __restrict__ rec, Geo g, int npass, const int* skip) {
    using CT = typename CTOf<VB>::type;
    if (skip && *skip) return;
    __shared__ CT sx[CMAX], sb[CMAX], sd[CMAX], sa[CMAX], sbb[CMAX], sm[CMAX];
    const int N = (int)g.N;
    for (int c = threadIdx.x; c < N; c += blockDim.x) {
        int i = c % g.sx, rest = c / g.sx;
        int j = rest % (g.ny + 2), k = rest / (g.ny + 2);
        bool interior = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && (D == 2 || (k >= 1 && k <= g.nz));
        C4 r = rec[c];
        sb[c] = (CT)b[c];
        sd[c] = interior ? (CT)r.w : CT(-1);
        sa[c] = (CT)r.x; sbb[c] = (CT)r.y; sm[c] = (D == 3) ? (CT)r.z : CT(0);
        sx[c] = CT(0);
    }
    __syncthreads();
    const int ni = g.nx * g.ny * (D == 3 ? g.nz : 1);
    for (int p = 0; p < npass; p++) {
        const int parity = (p & 1) ? 0 : 1;
        for (int q = threadIdx.x; q < ni; q += blockDim.x) {
            int i = q % g.nx + 1, rest = q / g.nx;
            int j = rest % g.ny + 1, k = (D == 3) ? rest / g.ny + 1 : 0;
            if (((i + j + k) & 1) != parity) continue;
            long c = (long)i + (long)j * g.sx + (long)k * g.sy;
            CT dv = sd[c];
            if (!(dv >= CT(1e-30))) continue;
            CT nb = sa[c] * sx[c + 1] + sa[c - 1] * sx[c - 1] + sbb[c] * sx[c + g.sx] + sbb[c - g.sx] * sx[c - g.sx];
            if (D == 3) nb = nb + sm[c] * sx[c + g.sy] + sm[c - g.sy] * sx[c - g.sy];
            sx[c] = (sb[c] + nb) / dv;
        }
        __syncthreads();
    }
    for (int q = threadIdx.x; q < ni; q += blockDim.x) {
        int i = q % g.nx + 1, rest = q / g.nx;
        int j = rest % g.ny + 1, k = (D == 3) ? rest / g.ny + 1 : 0;
        long c = (long)i + (long)j * g.sx + (long)k * g.sy;
        xout[c] = (VB)sx[c];
    }
}

}  // namespace uaamg
