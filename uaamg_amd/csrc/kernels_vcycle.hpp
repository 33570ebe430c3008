// kernels_vcycle.hpp -- per-pass V-cycle kernels (one RBGS colour, residual
// restriction, prolongation), used for any nu_pre/nu_post and as the
// cross-check of the fused kernels.
//
// Reference behaviour (paths relative to the reference root):
//   smooth          src/solver/preconditioner/uaamg_preconditioner_3d.cpp:123-141
//                   (parity (i+j+k)&1 on 1-based indices, skip solid and diag < 1e-30,
//                    x = (b + sum c*x_nb) / diag in place)
//   restrictResidual uaamg_preconditioner_3d.cpp:144-165 (sum of fluid children's b - Ax)
//   prolongateAdd   uaamg_preconditioner_3d.cpp:168-181 (fluid children += 2 x_c)
#pragma once
#include "kernels_setup.hpp"

namespace uaamg {

// The 7-point (5-point in 2D) stencil of one cell in compute type CT.
template <typename CT>
struct Sten {
    CT xp, xm, yp, ym, zp, zm, d;
    bool fluid;
};

// Level 0: couplings from the neighbour byte.
template <int D, typename CT>
__device__ __forceinline__ Sten<CT> sten_l0(uint8_t nb, const Consts<CT>& c, const Geo& g, int i, int j, int k) {
    Sten<CT> s;
    s.fluid = nb & NB_FLUID;
    s.xp = (nb & NB_XP) ? c.idx2 : CT(0);
    s.xm = (nb & NB_XM) ? c.idx2 : CT(0);
    s.yp = (nb & NB_YP) ? c.idy2 : CT(0);
    s.ym = (nb & NB_YM) ? c.idy2 : CT(0);
    s.zp = (D == 3 && (nb & NB_ZP)) ? c.idz2 : CT(0);
    s.zm = (D == 3 && (nb & NB_ZM)) ? c.idz2 : CT(0);
    CT d = s.xp + s.xm + s.yp + s.ym;
    if (D == 3) d = d + s.zp + s.zm;
    if (c.any_dirichlet && s.fluid) {
        CT e = wall_excess0<D, CT>(c, g, i, j, k);
        if (e != CT(0)) d += e;
    }
    s.d = d;
    return s;
}

// BF16 records hold the diagonal rounded up to the next BF16 value (the
// coarse operator stays diagonally dominant; k_level_diag).  On a cell with no
// Dirichlet excess that is the rounded-up sum of its six stored couplings, a
// small lift that makes a pure-Neumann level non-singular and puts the
// residual a BF16 V-cycle leaves in its smoothest modes.  Where the six
// couplings are at hand, a stored diagonal equal to their rounded-up sum is
// replaced by the sum itself (exact in FP32: six BF16 values of close
// magnitude), so such a level keeps its zero row sums; a diagonal with
// Dirichlet excess differs from that sum by far more than one BF16 step and
// is kept.  Other record types store the diagonal unrounded.
template <typename C4, typename CT>
__device__ __forceinline__ CT snap_diag(CT dstored, CT csum) {
    if constexpr (std::is_same<C4, BF16x4>::value) {
        if (!(dstored > CT(0))) return dstored;
        const float cs = (float)csum;
        __bf16 r = (__bf16)cs;
        if ((float)r < cs) r = __builtin_bit_cast(__bf16, (uint16_t)(__builtin_bit_cast(uint16_t, r) + 1));
        return ((float)r == (float)dstored) ? csum : dstored;
    } else {
        (void)csum;
        return dstored;
    }
}

// Levels >= 1: stored record {cx, cy, cz, diag} (diag < 0 marks a solid
// cell); the -x/-y/-z couplings of a cell live at its neighbour.
template <int D, typename CT, typename C4>
__device__ __forceinline__ Sten<CT> sten_lc(const C4* __restrict__ coef, const Geo& g, long id) {
    Sten<CT> s;
    C4 c0 = coef[id];
    s.d = (CT)c0.w;
    s.fluid = s.d >= CT(0);
    s.xp = (CT)c0.x; s.yp = (CT)c0.y;
    s.xm = (CT)coef[id - 1].x;
    s.ym = (CT)coef[id - g.sx].y;
    if (D == 3) { s.zp = (CT)c0.z; s.zm = (CT)coef[id - g.sy].z; }
    else { s.zp = CT(0); s.zm = CT(0); }
    if (s.fluid) s.d = snap_diag<C4, CT>(s.d, s.xp + s.xm + s.yp + s.ym + s.zp + s.zm);
    return s;
}

// Arguments of one level for the per-pass kernels.
template <typename V, typename C4>
struct LevelArgs {
    Geo g;
    V* b;                    // level right-hand side (level 0: the CG residual r)
    V* x;                    // level solution
    const uint8_t* mask;     // level 0: neighbour bytes (unused above level 0)
    const C4* coef;          // levels >= 1: {cx, cy, cz, diag (< 0: solid)}
};

template <int D, bool L0, typename V, typename C4, typename CT>
__device__ __forceinline__ Sten<CT> sten_at(const LevelArgs<V, C4>& L, const Consts<CT>& c, int i, int j, int k, long id) {
    if constexpr (L0) return sten_l0<D, CT>(L.mask[id], c, L.g, i, j, k);
    else return sten_lc<D, CT>(L.coef, L.g, id);
}

// One RBGS colour pass: cells with ((i+j+k)&1) == parity.
template <int D, bool L0, typename V, typename C4>
__global__ void k_smooth(LevelArgs<V, C4> L, Consts<typename CTOf<V>::type> c, int parity, int kz, const int* skip) {
    using CT = typename CTOf<V>::type;
    if (skip && *skip) return;
    const Geo g = L.g;
    UAAMG_CELL_LOOP(g, D, kz) {
        if (((i + j + k) & 1) != parity) continue;
        long id = gidx(g, i, j, k);
        Sten<CT> s = sten_at<D, L0>(L, c, i, j, k, id);
        if (!s.fluid || s.d < CT(1e-30)) continue;
        CT nb = s.xp * (CT)L.x[id + 1] + s.xm * (CT)L.x[id - 1] + s.yp * (CT)L.x[id + g.sx] + s.ym * (CT)L.x[id - g.sx];
        if (D == 3) nb = nb + s.zp * (CT)L.x[id + g.sy] + s.zm * (CT)L.x[id - g.sy];
        L.x[id] = (V)(((CT)L.b[id] + nb) / s.d);
    }
}

// Residual restriction b_c = sum over fluid children of (b - A x); solid
// coarse cells get 0 (never read: the smoother skips them).
template <int D, bool L0, typename V, typename C4, typename VC, typename CC4>
__global__ void k_restrict(LevelArgs<V, C4> F, LevelArgs<VC, CC4> Cl, Consts<typename CTOf<V>::type> c, int kz, const int* skip) {
    using CT = typename CTOf<V>::type;
    if (skip && *skip) return;
    const Geo cg = Cl.g, fg = F.g;
    UAAMG_CELL_LOOP(cg, D, kz) {
        long cid = gidx(cg, i, j, k);
        if ((double)Cl.coef[cid].w < 0) { Cl.b[cid] = (VC)0; continue; }
        int i_f = 2 * i - 1, j_f = 2 * j - 1, k_f = (D == 3) ? 2 * k - 1 : 0;
        // the residual b - A x cancels strongly; it is formed in FP64 whatever the storage
        double sum = 0;
        for (int di = 0; di < 2; di++)
            for (int dj = 0; dj < 2; dj++)
                for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++) {
                    int fi = i_f + di, fj = j_f + dj, fk = k_f + dk;
                    long fid = gidx(fg, fi, fj, fk);
                    Sten<CT> s = sten_at<D, L0>(F, c, fi, fj, fk, fid);
                    if (!s.fluid) continue;
                    double Ax = (double)s.d * (double)F.x[fid] - (double)s.xp * (double)F.x[fid + 1]
                              - (double)s.xm * (double)F.x[fid - 1] - (double)s.yp * (double)F.x[fid + fg.sx]
                              - (double)s.ym * (double)F.x[fid - fg.sx];
                    if (D == 3) Ax = Ax - (double)s.zp * (double)F.x[fid + fg.sy] - (double)s.zm * (double)F.x[fid - fg.sy];
                    sum += (double)F.b[fid] - Ax;
                }
        Cl.b[cid] = (VC)sum;
    }
}

// Prolongation: fluid fine cells of a fluid coarse parent get += 2 x_c.
template <int D, bool L0, typename V, typename C4, typename VC, typename CC4>
__global__ void k_prolong(LevelArgs<V, C4> F, LevelArgs<VC, CC4> Cl, int kz, const int* skip) {
    using CT = typename CTOf<V>::type;
    if (skip && *skip) return;
    const Geo fg = F.g, cg = Cl.g;
    UAAMG_CELL_LOOP(fg, D, kz) {
        long fid = gidx(fg, i, j, k);
        bool fluid;
        if constexpr (L0) fluid = F.mask[fid] & NB_FLUID;
        else fluid = (double)F.coef[fid].w >= 0;
        if (!fluid) continue;
        int ic = (i + 1) >> 1, jc = (j + 1) >> 1, kc = (D == 3) ? (k + 1) >> 1 : 0;
        if (ic > cg.nx || jc > cg.ny || (D == 3 && kc > cg.nz)) continue;
        long cid = gidx(cg, ic, jc, kc);
        if ((double)Cl.coef[cid].w < 0) continue;
        F.x[fid] = (V)((CT)F.x[fid] + CT(2) * (CT)Cl.x[cid]);
    }
}

}  // namespace uaamg
