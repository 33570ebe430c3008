// kernels_batch.hpp -- one launch per phase for many independent grids.
//
// A batch holds grids of uneven size, mask and walls.  Each phase of a PCG
// iteration (matvec, update, each V-cycle level, the tail, the p update) is
// one launch over a work-list of (grid, tile) items built on the host; the
// workgroup reads its grid's descriptor and runs the same device body as the
// single-grid kernel.  Reductions write one partial per item; the finish
// kernels run one workgroup per grid over that grid's contiguous item slice.
// Converged grids skip their items (their done flag), so a batch runs until
// its slowest member converges while the work shrinks.
#pragma once
#include "kernels_cg.hpp"
#include "kernels_march.hpp"
#include "kernels_rows.hpp"
#include "kernels_tail.hpp"

namespace uaamg {

constexpr int BATCH_MAXL = 20;

// Everything a batched kernel needs about one grid (device memory, one per grid).
template <typename S>
struct BGrid {
    using C4 = typename Vec4<S>::type;
    Geo g0;
    long N4;
    float *r, *z, *p, *Ap, *x;
    const uint8_t* nbr0;
    Consts<double> c0d;
    Consts<float> c0;
    Ctrl* ctrl;
    const double* count;
    // the solve's input/output (set per solve)
    const double* rhs;
    double* pout;
    int warm;
    // V-cycle
    int nl, lt;                        // levels; first level of the tail kernel
    Geo lg[BATCH_MAXL];
    S* lb[BATCH_MAXL];
    S* lx[BATCH_MAXL];
    C4* lcoef[BATCH_MAXL];
    int lKC[BATCH_MAXL];               // march chunk per level (level 0 too, when not on rows)
    int cpl;                           // level-0 row kernels: cells per lane (0: march)
    RowTab rtab;
    RowGeo rgeo;
    TailArgs<C4> tail;
    // slices of the reducing launches
    int flat_base, flat_cnt;
    const double* up_partials;
    int up_nbt, up_base, up_cnt;
};

// Work item: grid, and three launch coordinates (flat: local block, block count).
struct BItem { int g, a, b, c; };

// ---------------------------------------------------------------- CG
template <typename S>
__global__ void kb_ctrl_init(BGrid<S>* G, int ng, int max_iter, double tol, int tol_mode, int flexible,
                             int project_mean) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    Ctrl c = {};
    c.flexible = flexible;
    c.hist = nullptr;
    c.count = *G[g].count;
    c.max_iter = max_iter;
    c.tol = tol;
    c.tol_mode = tol_mode;
    c.project = project_mean >= 0 ? project_mean : (G[g].c0d.any_dirichlet ? 0 : 1);   // auto: on iff all walls Neumann
    c.first = 1;
    c.done = (max_iter <= 0) ? 1 : 0;
    c.status = (max_iter <= 0) ? ST_MAXITER : ST_RUNNING;
    *G[g].ctrl = c;
}

// copy every member's CG scalars into one array (one device-to-host copy)
template <typename S>
__global__ void kb_gather_ctrl(BGrid<S>* G, int ng, Ctrl* out) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ng) out[g] = *G[g].ctrl;
}

// flat phases: item = (grid, local block, block count)
template <typename S>
__global__ __launch_bounds__(256) void kb_init_sum(BGrid<S>* G, const BItem* items, double* partials) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    d_init_sum(d.rhs, d.nbr0, d.N4, partials, d.ctrl, it.a, it.b);
}
template <typename S>
__global__ __launch_bounds__(256) void kb_init_r(BGrid<S>* G, const BItem* items, double* partials) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    d_init_r<float>(d.rhs, d.r, d.nbr0, d.N4, partials, d.ctrl, it.a, it.b);
}
template <typename S>
__global__ __launch_bounds__(256) void kb_init_sub(BGrid<S>* G, const BItem* items, double* partials) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    d_init_sub<float>(d.r, d.nbr0, d.N4, partials, d.ctrl, it.a, it.b);
}
template <int D, typename S>
__global__ __launch_bounds__(256) void kb_cast_warm(BGrid<S>* G, const BItem* items) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    const long n = d.g0.N;
    for (long t = (long)it.a * blockDim.x + threadIdx.x; t < n; t += (long)it.b * blockDim.x) {
        if (d.warm) d.x[t] = (d.nbr0[t] & NB_FLUID) ? (float)d.pout[t] : 0.0f;
        else d.x[t] = 0.0f;
    }
    for (long t = (long)it.a * blockDim.x + threadIdx.x; t < n; t += (long)it.b * blockDim.x) d.p[t] = 0.0f;
}
template <int D, typename S>
__global__ __launch_bounds__(256) void kb_init_warm(BGrid<S>* G, const BItem* items, double* partials) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    if (!d.warm) { double v[1] = {0.0}; block_store<1>(v, partials); return; }
    d_init_warm<D, float>(d.r, d.x, d.nbr0, d.g0, d.c0d, d.N4, partials, d.ctrl, it.a, it.b);
}
template <int D, typename S>
__global__ __launch_bounds__(256) void kb_matvec(BGrid<S>* G, const BItem* items, double* partials) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    if (d.ctrl->done) { double v[1] = {0.0}; block_store<1>(v, partials); return; }
    d_matvec_dot<D, float>(d.p, d.Ap, d.nbr0, d.g0, d.c0d, d.N4, partials, d.ctrl, it.a, it.b);
}
template <typename S, bool FLEX>
__global__ __launch_bounds__(256) void kb_update(BGrid<S>* G, const BItem* items, double* partials) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    if (d.ctrl->done) {
        double v[FLEX ? 3 : 1] = {};
        block_store<FLEX ? 3 : 1>(v, partials);
        return;
    }
    d_update<float, FLEX>(d.x, d.r, d.p, d.Ap, d.z, d.nbr0, d.N4, partials, d.ctrl, it.a, it.b);
}
template <typename S>
__global__ __launch_bounds__(256) void kb_pupdate(BGrid<S>* G, const BItem* items) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    d_pupdate<float>(d.p, d.z, d.nbr0, d.N4, d.ctrl, it.a, it.b);
}
template <typename S>
__global__ __launch_bounds__(256) void kb_finalize(BGrid<S>* G, const BItem* items) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    const long n = d.g0.N;
    for (long t = (long)it.a * blockDim.x + threadIdx.x; t < n; t += (long)it.b * blockDim.x)
        if (d.nbr0[t] & NB_FLUID) d.pout[t] = d.warm ? (double)d.x[t] : d.pout[t] + (double)d.x[t];
}

// finish kernels: one workgroup per grid over its slice of the flat partials
template <typename S>
__global__ __launch_bounds__(1024) void kb_fin_init_sum(BGrid<S>* G, const double* partials, int nbt) {
    const BGrid<S>& d = G[blockIdx.x];
    d_fin_init_sum(partials, nbt, d.flat_base, d.flat_cnt, d.ctrl);
}
template <typename S>
__global__ __launch_bounds__(1024) void kb_fin_init_r(BGrid<S>* G, const double* partials, int nbt) {
    const BGrid<S>& d = G[blockIdx.x];
    d_fin_init_r(partials, nbt, d.flat_base, d.flat_cnt, d.ctrl);
}
template <typename S>
__global__ __launch_bounds__(1024) void kb_fin_init_sub(BGrid<S>* G, const double* partials, int nbt) {
    const BGrid<S>& d = G[blockIdx.x];
    d_fin_init_sub(partials, nbt, d.flat_base, d.flat_cnt, d.ctrl);
}
template <typename S>
__global__ __launch_bounds__(1024) void kb_fin_init_warm(BGrid<S>* G, const double* partials, int nbt) {
    const BGrid<S>& d = G[blockIdx.x];
    if (!d.warm) return;
    d_fin_init_warm(partials, nbt, d.flat_base, d.flat_cnt, d.ctrl);
}
template <typename S>
__global__ __launch_bounds__(1024) void kb_fin_matvec(BGrid<S>* G, const double* partials, int nbt) {
    const BGrid<S>& d = G[blockIdx.x];
    d_fin_matvec(partials, nbt, d.flat_base, d.flat_cnt, d.ctrl);
}
template <typename S, bool FLEX>
__global__ __launch_bounds__(1024) void kb_fin_update(BGrid<S>* G, const double* partials, int nbt) {
    const BGrid<S>& d = G[blockIdx.x];
    d_fin_update<FLEX>(partials, nbt, d.flat_base, d.flat_cnt, d.ctrl);
}
template <typename S>
__global__ __launch_bounds__(1024) void kb_fin_zsums(BGrid<S>* G) {
    const BGrid<S>& d = G[blockIdx.x];
    d_fin_zsums(d.up_partials, d.up_nbt, d.up_base, d.up_cnt, d.ctrl);
}

// ---------------------------------------------------------------- V-cycle
// level-0 row kernels (3D): item = (grid, tile)
template <int CPL, typename S>
__global__ __launch_bounds__(640) void kb_rows_down0(BGrid<S>* G, const BItem* items) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    d_rows_down0<CPL, 10, 2, S>(d.r, d.nbr0, d.g0, d.rtab, d.rgeo, d.lb[1], d.lg[1], &d.ctrl->done, it.a);
}
template <int CPL, typename S>
__global__ __launch_bounds__(640) void kb_rows_up0(BGrid<S>* G, const BItem* items, double* partials) {
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    if (d.ctrl->done) { double v[3] = {0.0, 0.0, 0.0}; block_store<3>(v, partials); return; }
    d_rows_up0<CPL, 10, 2, S>(d.r, d.nbr0, d.g0, d.rtab, d.rgeo, d.lx[1], d.lg[1], d.z, &d.ctrl->done, partials, it.a);
}

// marching kernels: item = (grid, bx, by, bz) of the level's tile grid
template <int D, int RX, int RY, int PPT, typename S>
__global__ __launch_bounds__(256) void kb_march_down0(BGrid<S>* G, const BItem* items) {
    using C4 = typename Vec4<S>::type;
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    FusedArgs<D, true, float, C4> A{d.g0, d.nbr0, nullptr, d.c0};
    d_down_m<D, true, RX, RY, PPT, float, C4, S, C4>(d.r, A, d.lb[1], d.lcoef[1], d.lg[1], d.lKC[0], &d.ctrl->done,
                                                     it.a, it.b, it.c);
}
template <int D, int RX, int RY, int PPT, typename S>
__global__ __launch_bounds__(256) void kb_march_up0(BGrid<S>* G, const BItem* items, double* partials) {
    using C4 = typename Vec4<S>::type;
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    if (d.ctrl->done) { double v[3] = {0.0, 0.0, 0.0}; block_store<3>(v, partials); return; }
    FusedArgs<D, true, float, C4> A{d.g0, d.nbr0, nullptr, d.c0};
    d_up_m<D, true, RX, RY, PPT, true, float, C4, float, S>(d.r, A, d.lx[1], d.lg[1], d.z, d.lKC[0], &d.ctrl->done,
                                                           partials, it.a, it.b, it.c);
}
template <int D, int RX, int RY, int PPT, typename S>
__global__ __launch_bounds__(256) void kb_march_down(BGrid<S>* G, const BItem* items, int l) {
    using C4 = typename Vec4<S>::type;
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    FusedArgs<D, false, float, C4> A{d.lg[l], nullptr, d.lcoef[l], Consts<float>{}};
    d_down_m<D, false, RX, RY, PPT, S, C4, S, C4>(d.lb[l], A, d.lb[l + 1], d.lcoef[l + 1], d.lg[l + 1], d.lKC[l],
                                                  &d.ctrl->done, it.a, it.b, it.c);
}
template <int D, int RX, int RY, int PPT, typename S>
__global__ __launch_bounds__(256) void kb_march_up(BGrid<S>* G, const BItem* items, int l) {
    using C4 = typename Vec4<S>::type;
    const BItem it = items[blockIdx.x];
    const BGrid<S>& d = G[it.g];
    FusedArgs<D, false, float, C4> A{d.lg[l], nullptr, d.lcoef[l], Consts<float>{}};
    d_up_m<D, false, RX, RY, PPT, false, S, C4, S, S>(d.lb[l], A, d.lx[l + 1], d.lg[l + 1], d.lx[l], d.lKC[l],
                                                      &d.ctrl->done, nullptr, it.a, it.b, it.c);
}

// the small levels of every grid: one workgroup per grid
template <int D, typename S, int NLMAX>
__global__ __launch_bounds__(TAIL_THREADS) void kb_tail(BGrid<S>* G) {
    const BGrid<S>& d = G[blockIdx.x];
    d_tail<D, S, typename Vec4<S>::type, NLMAX>(d.tail);
}

}  // namespace uaamg
