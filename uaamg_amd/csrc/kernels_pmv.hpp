// kernels_pmv.hpp -- level-0 search direction and operator in one pass
// (FP32 vectors, 3D).
//
// One pass replaces the p update of iteration k-1 and the operator of
// iteration k (pcg_3d.cpp:116-122 and 94-96):
//   p_new = (z - mean_z) + beta p_old      on fluid cells (0 elsewhere)
//   Ap    = d p_new - sum_nb c p_new_nb     (FP64 arithmetic on the FP32
//                                            values, rounded once; d from the
//                                            neighbour byte + Dirichlet excess)
//   partial sums of p_new . Ap             (deterministic, per workgroup)
// p_old and p_new are separate buffers: a workgroup recomputes p_new on its
// halo row and planes from z and p_old, so no workgroup reads what another
// one writes.  p_new is k_pupdate's formula (in FP32); the operator is k_matvec_dot's
// (apply_A) in the stencil form of the row kernels: p_new is 0 on every
// non-fluid cell, so a face to a solid cell or a Neumann wall drops out of
// both d and the sum, and a Dirichlet wall face adds 2/h^2 to d (apply_A's
// ghost value -vc).  The operator stays in FP64 like apply_A: for smooth p,
// d p and the neighbour sum nearly cancel, and an FP32 difference would lose
// the digits CG needs at relres 1e-6.
//
// Mapping as in the row kernels: a wave owns whole x-rows (CPL cells per
// lane), NW waves of RW rows per workgroup with one halo row on each side,
// a march over KC planes (+1 halo plane on each side) with the three planes
// of the stencil in registers, x-neighbours by DPP, y-neighbours of other
// waves through LDS (one barrier per plane).
#pragma once
#include "kernels_cg.hpp"
#include "kernels_rows.hpp"

namespace uaamg {

constexpr int PMV_NW = 9, PMV_RW = 2;   // 18 rows: 16 output rows and one halo row on each side

// FP64 stencil constants: d per neighbour pattern (nb >> 1), 1/h^2 per axis,
// Dirichlet excess per face
struct PmvTab {
    double d[64];
    double idx2, idy2, idz2;
    double ex[6];
    int any_dirichlet;
};

template <int CPL, int NW>
struct PmvLDS {
    float ex[2][NW][2][CPL][64];   // [step parity][wave][first / last row][cell][lane]
};

// With FZ the pass also does the finish of the z sums (finish_zsums,
// pcg_3d.cpp:113-115): every workgroup adds the level-0 up kernel's partials
// of sum z, sum r.z, sum r in the same order and forms the same mean_z, r.z
// and beta; workgroup 0 stores them.  rsold and the first-iteration flag are
// read here and only written by the next kernel (the level-0 down kernel),
// so no workgroup can see a value of this launch.
template <int CPL, int NW, int RW, bool FZ = false, bool NT = false, bool FULL = false>
__device__ __forceinline__ void d_rows_pmv(const float* __restrict__ z, const float* __restrict__ pold,
                                           float* __restrict__ pnew, float* __restrict__ Ap,
                                           const uint8_t* __restrict__ nbr, const Geo& g, const PmvTab& K,
                                           const RowGeo& rg, double* partials, Ctrl* ctrl, int tile,
                                           const double* zp = nullptr, int nzp = 0) {
    if (ctrl->done) return;
    __shared__ PmvLDS<CPL, NW> L;
    __shared__ double sd[64];
    __shared__ double s_zs[3];
    double acc[1] = {0.0};
    const bool active = tile < rg.ntiles;   // idle workgroups still store a zero partial
    const int ty = tile % rg.nty, kc = tile / rg.nty;
    // odd z-chunks march downwards, so the two tiles on either side of a
    // chunk boundary read the planes they share at about the same time (both
    // first or both last), and the second read can hit the L2: at 256^3 the
    // pass fetches 195.9 instead of 199.3 MB (28.44 vs 28.54 ms per solve)
    const int dir = (kc & 1) ? -1 : 1;
    RowThread<CPL, NW, RW, 1, FULL> th;
    th.init(g, rg, active ? ty : 0, active ? kc : 0);
    const int lane = th.lane, w = th.w;

    float PN[4][RW][CPL];        // p_new of planes (slot = plane & 3)
    uint8_t NB[4][RW][CPL];
    float pfz[2][RW][CPL], pfp[2][RW][CPL];
    uint8_t pfn[2][RW][CPL];
    auto issue = [&](int slot, int P) {
        const int po = th.plane(g, P);
#pragma unroll
        for (int t = 0; t < RW; t++) {
            row_load<CPL, NT>(z, nbr, th.roff[t] + po, th.okx, pfz[slot][t], pfn[slot][t]);
            typename RowVec<CPL>::V v;
            if constexpr (NT) v = __builtin_nontemporal_load((const typename RowVec<CPL>::V*)(pold + th.roff[t] + po));
            else v = *(const typename RowVec<CPL>::V*)(pold + th.roff[t] + po);
#pragma unroll
            for (int q = 0; q < CPL; q++) pfp[slot][t][q] = v[q];
        }
    };
    // the z-sum partials first, then the first two planes: both sets of
    // loads are in flight while the partials are added
    double zs[3] = {0.0, 0.0, 0.0};
    if constexpr (FZ) {
        for (int x = lin_thread(); x < nzp; x += n_threads())
#pragma unroll
            for (int q = 0; q < 3; q++) zs[q] += zp[(long)q * nzp + x];
    }
    const int Pfirst = dir > 0 ? th.m0 - 1 : th.last + 1;
    const int nst = th.last - th.m0 + 3;
    if (active) {
        issue(0, Pfirst);
        issue(1, Pfirst + dir);
    }
    for (int e = threadIdx.x; e < 64; e += NW * 64) sd[e] = K.d[e];
    if constexpr (FZ) {
        block_sum<3>(zs);
        if (threadIdx.x == 0) { s_zs[0] = zs[0]; s_zs[1] = zs[1]; s_zs[2] = zs[2]; }
    }
    __syncthreads();
    // the p update runs in FP32 (one FMA) with the FP64 scalars rounded once:
    // it differs from k_pupdate's FP64 form by the rounding of the scalars,
    // far below the FP32 storage of the vectors
    float mean, beta;
    if constexpr (FZ) {
        const double sz = s_zs[0], srz = s_zs[1], sr = s_zs[2];
        const double m = (ctrl->project && ctrl->count > 0) ? sz / ctrl->count : 0.0;
        const double rz = srz - m * sr;
        const double b = ctrl->first ? 0.0 : rz / ctrl->rsold;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            ctrl->sum_z = sz; ctrl->sum_rz = srz; ctrl->sum_r = sr;
            ctrl->mean_z = m; ctrl->rz = rz; ctrl->beta = b;
            if (!isfinite(rz)) { ctrl->done = 1; ctrl->status = ST_NONFINITE; }
        }
        if (!isfinite(rz)) return;   // the whole solve stops; the down kernel sees done
        mean = (float)m; beta = (float)b;
    } else {
        mean = (float)ctrl->mean_z; beta = (float)ctrl->beta;
    }

    if (active) {
#pragma unroll
        for (int s = 0; s < 4; s++)
#pragma unroll
            for (int t = 0; t < RW; t++)
#pragma unroll
                for (int q = 0; q < CPL; q++) { PN[s][t][q] = 0.0f; NB[s][t][q] = 0; }

        // step: plane P enters (slot S); the operator on plane P-1
        auto step = [&](auto Sc, int P) {
            constexpr int S = decltype(Sc)::value;
            constexpr int S1 = (S + 3) & 3, S2 = (S + 2) & 3;   // planes P-1, P-2
            constexpr int PB = S & 1;
            const bool pin = P >= 1 && P <= g.nz;
#pragma unroll
            for (int t = 0; t < RW; t++)
#pragma unroll
                for (int q = 0; q < CPL; q++) {
                    const uint8_t nb = pin ? pfn[PB][t][q] : (uint8_t)0;
                    NB[S][t][q] = nb;
                    PN[S][t][q] = (nb & NB_FLUID)
                        ? fmaf(beta, pfp[PB][t][q], pfz[PB][t][q] - mean) : 0.0f;
                }
            // the last plane to enter is last + 1 (upwards) or m0 - 1 (downwards);
            // beyond it the loads repeat that cached plane (no memory traffic), so
            // they stay unconditional and the waits stay per step
            issue(PB, dir > 0 ? min(P + 2, th.last + 1) : max(P - 2, th.m0 - 1));
#pragma unroll
            for (int q = 0; q < CPL; q++) {
                L.ex[PB][w][0][q][lane] = PN[S][0][q];
                L.ex[PB][w][1][q][lane] = PN[S][RW - 1][q];
            }
            const int R = P - dir;
            if (R >= th.m0 && R <= th.last) {
                constexpr int RB = S1 & 1;   // LDS buffer written when plane R entered
                // the diagonals of the plane's cells: every table read before
                // the first use, the Dirichlet excess in one uniform branch
                double dP[RW][CPL];
#pragma unroll
                for (int t = 0; t < RW; t++)
#pragma unroll
                    for (int q = 0; q < CPL; q++) dP[t][q] = sd[NB[S1][t][q] >> 1];
                if (K.any_dirichlet) {
#pragma unroll
                    for (int t = 0; t < RW; t++)
#pragma unroll
                        for (int q = 0; q < CPL; q++) {
                            const int i = th.i0 + q, j = th.jrow[t];
                            double d = dP[t][q];
                            if (i == 1) d += K.ex[0];
                            if (i == g.nx) d += K.ex[1];
                            if (j == 1) d += K.ex[2];
                            if (j == g.ny) d += K.ex[3];
                            if (R == 1) d += K.ex[4];
                            if (R == g.nz) d += K.ex[5];
                            dP[t][q] = d;
                        }
                }
                sfor<RW>([&](auto tc) {
                    constexpr int t = decltype(tc)::value;
                    const float el = lane_from_left(PN[S1][t][CPL - 1]);
                    const float er = lane_from_right(PN[S1][t][0]);
                    float out[CPL];
                    float pap = 0.0f;   // the row's p.Ap in FP32, accumulated in FP64
                    sfor<CPL>([&](auto qc) {
                        constexpr int q = decltype(qc)::value;
                        const float vc = PN[S1][t][q];
                        const float vl = (q > 0) ? PN[S1][t][q > 0 ? q - 1 : 0] : el;
                        const float vr = (q < CPL - 1) ? PN[S1][t][q < CPL - 1 ? q + 1 : 0] : er;
                        float vb, vt;
                        if constexpr (t > 0) vb = PN[S1][t - 1][q]; else vb = L.ex[RB][w > 0 ? w - 1 : 0][1][q][lane];
                        if constexpr (t < RW - 1) vt = PN[S1][t + 1][q]; else vt = L.ex[RB][w < NW - 1 ? w + 1 : NW - 1][0][q][lane];
                        const uint8_t nb = NB[S1][t][q];
                        const double d = dP[t][q];
                        const double sum = K.idx2 * ((double)vl + (double)vr) + K.idy2 * ((double)vb + (double)vt)
                                         + K.idz2 * ((double)PN[S2][t][q] + (double)PN[S][t][q]);
                        const bool fl = nb & NB_FLUID;
                        const float o = fl ? (float)(d * (double)vc - sum) : 0.0f;
                        pap = fmaf(th.okx[q] ? vc : 0.0f, o, pap);
                        out[q] = o;
                    });
                    if (th.tile_row[t]) acc[0] += (double)pap;
                    if (th.tile_row[t]) {
                        const int id = th.roff[t] + R * (int)g.sy;
                        if (th.full || th.okx[CPL - 1]) {
                            typename RowVec<CPL>::V vp, va;
#pragma unroll
                            for (int q = 0; q < CPL; q++) { vp[q] = PN[S1][t][q]; va[q] = out[q]; }
                            *(typename RowVec<CPL>::V*)(pnew + id) = vp;
                            *(typename RowVec<CPL>::V*)(Ap + id) = va;
                        } else {
#pragma unroll
                            for (int q = 0; q < CPL; q++)
                                if (th.okx[q]) { pnew[id + q] = PN[S1][t][q]; Ap[id + q] = out[q]; }
                        }
                    }
                });
            }
            __syncthreads();
        };
        for (int k = 0; k < nst; k += 4) {
            step(IC<0>{}, Pfirst + dir * k);
            if (k + 1 >= nst) break;
            step(IC<1>{}, Pfirst + dir * (k + 1));
            if (k + 2 >= nst) break;
            step(IC<2>{}, Pfirst + dir * (k + 2));
            if (k + 3 >= nst) break;
            step(IC<3>{}, Pfirst + dir * (k + 3));
        }
    }
    block_store<1>(acc, partials);
}

template <int CPL, int NW, int RW, bool FZ = false, bool NT = false, bool FULL = false>
__global__ __launch_bounds__(NW * 64) void k_rows_pmv(const float* __restrict__ z, const float* __restrict__ pold,
                                                      float* __restrict__ pnew, float* __restrict__ Ap,
                                                      const uint8_t* __restrict__ nbr, Geo g, PmvTab K,
                                                      RowGeo rg, double* partials, Ctrl* ctrl,
                                                      const double* zp, int nzp) {
    d_rows_pmv<CPL, NW, RW, FZ, NT, FULL>(z, pold, pnew, Ap, nbr, g, K, rg, partials, ctrl, row_tile_of_block(rg, blockIdx.x),
                                zp, nzp);
}

}  // namespace uaamg
