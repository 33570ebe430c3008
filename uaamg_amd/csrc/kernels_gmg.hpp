// kernels_gmg.hpp -- the GMG and AMG preconditioners of the reference's
// other factory keys ("pcg" / "pcg_gmg": GMGPreconditioner; "pcg_amg":
// AMGPreconditioner), 2D, as per-pass device kernels.
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

#define GMG_LOOP2(g)                                                  \
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;          \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;          \
    if (i > (g).nx || j > (g).ny) return;                             \
    const long id = (long)i + (long)j * (g).sx;

// neighbour values with the wall / solid rule and the effective diagonal
template <typename V, typename CT>
__device__ __forceinline__ void gmg_nb(const V* __restrict__ x, const uint8_t* __restrict__ S, const Geo& g, int i,
                                       int j, long id, CT& xL, CT& xR, CT& xB, CT& xT) {
    const CT xc = (CT)x[id];
    xL = (i > 1 && !S[id - 1]) ? (CT)x[id - 1] : xc;
    xR = (i < g.nx && !S[id + 1]) ? (CT)x[id + 1] : xc;
    xB = (j > 1 && !S[id - g.sx]) ? (CT)x[id - g.sx] : xc;
    xT = (j < g.ny && !S[id + g.sx]) ? (CT)x[id + g.sx] : xc;
}

// one colour ((i + j) & 1 == odd) of the smoother (gmg :41-81, amg :13-50)
template <typename V>
__global__ void k_gmg_colour(V* __restrict__ x, const V* __restrict__ b, const uint8_t* __restrict__ S, Geo g,
                             GmgC<typename CTOf<V>::type> c, int odd, const int* skip) {
    using CT = typename CTOf<V>::type;
    if (skip && *skip) return;
    GMG_LOOP2(g)
    if (((i + j) & 1) != odd || S[id]) return;
    CT xL, xR, xB, xT;
    gmg_nb<V, CT>(x, S, g, i, j, id, xL, xR, xB, xT);
    const CT lap = (xL + xR) * c.idx2 + (xB + xT) * c.idy2;
    CT eff_d = c.diag;
    if (i == 1 || S[id - 1]) eff_d -= c.idx2;
    if (i == g.nx || S[id + 1]) eff_d -= c.idx2;
    if (j == 1 || S[id - g.sx]) eff_d -= c.idy2;
    if (j == g.ny || S[id + g.sx]) eff_d -= c.idy2;
    const CT xc = (CT)x[id];
    x[id] = (V)(xc + ((eff_d < CT(1e-15)) ? CT(0) : CT(1) / eff_d) * ((CT)b[id] - c.diag * xc + lap));
}

// coarse b = mean of the fluid children's residuals (children (0,0), (0,1),
// (1,0), (1,1) in that order; gmg :121-142, amg :91-111)
template <bool AMG, typename VF, typename VC>
__global__ void k_gmg_restrict(const VF* __restrict__ x, const VF* __restrict__ b, const uint8_t* __restrict__ S,
                               Geo fg, GmgC<typename CTOf<VF>::type> c, VC* __restrict__ bc,
                               const uint8_t* __restrict__ SC, Geo cg, const int* skip) {
    using CT = typename CTOf<VF>::type;
    if (skip && *skip) return;
    const int ic = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int jc = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (ic > cg.nx || jc > cg.ny) return;
    CT sum = CT(0);
    int cnt = 0;
#pragma unroll
    for (int di = 0; di < 2; di++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++) {
            const int i = 2 * ic - 1 + di, j = 2 * jc - 1 + dj;
            const long id = (long)i + (long)j * fg.sx;
            if (S[id]) continue;
            CT xL, xR, xB, xT;
            gmg_nb<VF, CT>(x, S, fg, i, j, id, xL, xR, xB, xT);
            const CT xc = (CT)x[id];
            CT res;
            if constexpr (AMG) {
                const CT Ax = c.diag * xc - (xL + xR) * c.idx2 - (xB + xT) * c.idy2;
                res = (CT)b[id] - Ax;
            } else {
                const CT lap = (xL + xR) * c.idx2 + (xB + xT) * c.idy2;
                res = (CT)b[id] - c.diag * xc + lap;
            }
            sum += res;
            cnt++;
        }
    const long cid = (long)ic + (long)jc * cg.sx;
    if constexpr (AMG) bc[cid] = (VC)((cnt > 0) ? sum / (CT)cnt : CT(0));
    else if (!SC[cid] && cnt > 0) bc[cid] = (VC)(sum / (CT)cnt);
}

// x_f += x_c on fluid fine cells that have a parent; GMG skips solid parents
// (gmg :144-156), AMG adds their zero-filled x (amg :113-123)
template <bool AMG, typename VF, typename VC>
__global__ void k_gmg_prolong(VF* __restrict__ x, const uint8_t* __restrict__ S, Geo fg, const VC* __restrict__ xc,
                              const uint8_t* __restrict__ SC, Geo cg, const int* skip) {
    using CT = typename CTOf<VF>::type;
    if (skip && *skip) return;
    GMG_LOOP2(fg)
    if (S[id]) return;
    const int ic = (i + 1) >> 1, jc = (j + 1) >> 1;
    if (ic > cg.nx || jc > cg.ny) return;
    const long cid = (long)ic + (long)jc * cg.sx;
    if (!AMG && SC[cid]) return;
    x[id] = (VF)((CT)x[id] + (CT)xc[cid]);
}

}  // namespace uaamg
