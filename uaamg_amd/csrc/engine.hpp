// engine.hpp -- host-side orchestration of one grid's hierarchy and PCG on a
// HIP stream.  Template over dimension D, CG vector type T and coarse-level
// storage type S.  All launches are asynchronous on the context's stream; the
// host waits only when it polls Ctrl::done between batches of iterations.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>
#include "../../include/uaamg.h"
#include "kernels_cg.hpp"
#include "kernels_march.hpp"
#include "kernels_rows.hpp"
#include "kernels_tail.hpp"
#include "kernels_pmv.hpp"
#include "kernels_batch.hpp"
#include "kernels_gmg.hpp"

namespace uaamg {

struct HipError {
    hipError_t e; std::string where;
};

#define HIPCHK(call)                                                                   \
    do {                                                                               \
        hipError_t _e = (call);                                                        \
        if (_e != hipSuccess) throw HipError{_e, std::string(#call) + " @" + std::to_string(__LINE__)}; \
    } while (0)

struct Launch {
    dim3 grid, block;
    int kz;
    long nblocks() const { return (long)grid.x * grid.y * grid.z; }
};

inline Launch cells_launch(const Geo& g, int D) {
    Launch L;
    L.block = dim3(64, 4, 1);
    unsigned gx = (g.nx + 63) / 64, gy = (g.ny + 3) / 4;
    long planes = (D == 3) ? g.nz : 1;
    long per_plane = (long)gx * gy;
    int kz = (int)std::max<long>(1, (per_plane * planes + 8191) / 8192);
    if (D != 3) kz = 1;
    unsigned gz = (D == 3) ? (unsigned)((g.nz + kz - 1) / kz) : 1u;
    L.grid = dim3(gx, gy, gz);
    L.kz = kz;
    return L;
}

// elements allocated past the end of level buffers read by the row kernels
constexpr long kRowSlack = 64;

// Per-class device timing with events (opts.timing).
struct Timer {
    bool on = false;
    unsigned mask = 0;        // bit c set: time kernel class c
    hipStream_t s = nullptr;
    struct Rec { int cls; hipEvent_t a, b; };
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    hipEvent_t get() {
        if (used == pool.size()) { hipEvent_t e; HIPCHK(hipEventCreate(&e)); pool.push_back(e); }
        return pool[used++];
    }
    bool want(int cls) const { return on && ((mask >> cls) & 1u); }
    void begin(int cls, hipEvent_t& a) { if (!want(cls)) { a = nullptr; return; } a = get(); HIPCHK(hipEventRecord(a, s)); recs.push_back({cls, a, nullptr}); }
    void end(hipEvent_t a) { if (!a) return; hipEvent_t b = get(); HIPCHK(hipEventRecord(b, s)); recs.back().b = b; }
    void reset() { recs.clear(); used = 0; }
    void collect(uaamg_stats* st) {
        if (!on || !st) return;
        for (auto& r : recs) {
            float ms = 0; HIPCHK(hipEventElapsedTime(&ms, r.a, r.b));
            st->ms_kernel[r.cls] += ms; st->n_kernel[r.cls] += 1;
        }
    }
    ~Timer() { for (auto e : pool) (void)hipEventDestroy(e); }
};

struct TScope {
    Timer& t; hipEvent_t a{};
    TScope(Timer& t_, int cls) : t(t_) { t.begin(cls, a); }
    ~TScope() { t.end(a); }
};

// Launch of the main kernel of a level-0 class.  When the class is timed the
// events ride on the dispatch itself (hipExtLaunchKernelGGL), so the time is
// that kernel's execution alone, without event-record gaps or the small
// finishing kernels that follow it.
template <typename... KP, typename... Args>
inline void tlaunch(Timer& t, int cls, void (*kern)(KP...), dim3 grid, dim3 block, hipStream_t s, Args... args) {
    if (t.want(cls)) {
        hipEvent_t a = t.get(), b = t.get();
        hipExtLaunchKernelGGL(kern, grid, block, 0, s, a, b, 0, args...);
        t.recs.push_back({cls, a, b});
    } else {
        kern<<<grid, block, 0, s>>>(args...);
    }
}

// forms fixed when an engine is built (buffers and launch shapes depend on them)
constexpr int kStructForms = UAAMG_FORM_ROWS | UAAMG_FORM_ROWS_COARSE | UAAMG_FORM_TAIL | UAAMG_FORM_PMV;

class EngineBase {
public:
    virtual ~EngineBase() = default;
    virtual void set_grid(const uint8_t* solid_dev) = 0;
    virtual int solve(const double* rhs_dev, double* p_dev, int warm, int max_iter, double tol, uaamg_stats* st) = 0;
    virtual void apply(const double* r_dev, double* z_dev) = 0;
    virtual double matvec(const double* v_dev, double* Av_dev) = 0;
    virtual void level_data(int level, uint8_t* solid, double* diag, double* cx, double* cy, double* cz, double* ex) = 0;
    virtual int nlevels() const = 0;
    virtual Geo geo(int l) const = 0;
    virtual size_t bytes() const = 0;
    virtual void set_opts(const uaamg_opts& o) = 0;
    virtual void set_stream(hipStream_t s) = 0;
    virtual const uint8_t* solid_dev() const = 0;   // the level-0 solid mask (padded, device)
    // variable Laplacian (2D): padded FP64 device arrays, copied; all null: uniform operator
    virtual void set_variable_lap(const double* d, const double* ox, const double* oy) = 0;
};

// Hierarchy shape, shared by engines and the byte estimator (buildHierarchy,
// uaamg_preconditioner_3d.cpp:11-32): push, stop once a side <= S, halve.
inline std::vector<Geo> hierarchy_shape(int dim, int nx, int ny, int nz, int S) {
    std::vector<Geo> v;
    while (nx >= 2 && ny >= 2 && (dim == 2 || nz >= 2)) {
        v.push_back(make_geo(dim, nx, ny, nz));
        if (nx <= S || ny <= S || (dim == 3 && nz <= S)) break;
        nx /= 2; ny /= 2; if (dim == 3) nz /= 2;
        if (v.size() > 40) break;
    }
    return v;
}

template <typename CT>
Consts<CT> make_consts(int dim, const double h[3], const int wall[6]) {
    Consts<CT> c{};
    double ix = 1.0 / (h[0] * h[0]), iy = 1.0 / (h[1] * h[1]), iz = (dim == 3) ? 1.0 / (h[2] * h[2]) : 0.0;
    c.idx2 = (CT)ix; c.idy2 = (CT)iy; c.idz2 = (CT)iz;
    c.diag_u = (CT)((dim == 3) ? 2.0 * (ix + iy + iz) : 2.0 * (ix + iy));
    double e[6] = {2.0 * ix, 2.0 * ix, 2.0 * iy, 2.0 * iy, 2.0 * iz, 2.0 * iz};
    c.any_dirichlet = 0;
    for (int f = 0; f < 6; f++) {
        int on = (f < 2 * dim) ? (wall[f] == UAAMG_WALL_DIRICHLET) : 0;
        c.wall[f] = on;
        c.ex2[f] = on ? (CT)e[f] : CT(0);
        c.any_dirichlet |= on;
    }
    return c;
}

// R: storage of the coarse levels' stencil records (default: S, the level
// vectors' storage)
template <int D, typename T, typename S, typename R = S>
class Engine final : public EngineBase {
    using CT = typename CTOf<T>::type;
    using SC = typename CTOf<S>::type;
    using C4 = typename Vec4<R>::type;

    struct Level {
        Geo g;
        Launch lc;
        S* b = nullptr; S* x = nullptr;
        uint8_t* mask = nullptr;     // solid mask (set-up only)
        C4* coef = nullptr;          // {cx, cy, cz, diag (< 0: solid)}
        S* ex = nullptr;             // Dirichlet diagonal excess (set-up only)
    };

public:
    Engine(const uaamg_opts& o, const int n[3], const double h[3], const int wall[6], hipStream_t s)
        : o_(o), stream_(s) {
        for (int q = 0; q < 3; q++) { n_[q] = n[q]; h_[q] = h[q]; }
        for (int f = 0; f < 6; f++) wall_[f] = wall[f];
        shape_ = hierarchy_shape(D, n[0], n[1], D == 3 ? n[2] : 1, o.coarsest_side);
        if (shape_.empty()) throw HipError{hipErrorInvalidValue, "grid too small for a hierarchy"};
        c0_ = make_consts<CT>(D, h, wall);
        c0d_ = make_consts<double>(D, h, wall);
        g0_ = shape_[0];
        N4_ = (g0_.N + 3) / 4 * 4;
        margin_ = ((g0_.sx + 8 + 63) / 64) * 64;
        lc0_ = cells_launch(g0_, D);
        int all_neumann = 1;
        for (int f = 0; f < 2 * D; f++) if (wall[f] != UAAMG_WALL_NEUMANN) all_neumann = 0;
        project_ = (o.project_mean < 0) ? all_neumann : o.project_mean;
        ident_ = o.preconditioner == UAAMG_PRECOND_NONE;
        gmg_ = o.preconditioner == UAAMG_PRECOND_GMG ? 1 : (o.preconditioner == UAAMG_PRECOND_AMG ? 2 : 0);
        if (gmg_ && (D != 2 || sizeof(S) < 4))
            throw HipError{hipErrorInvalidValue, "GMG / AMG preconditioners: 2D grids, FP64 or FP32 levels"};
        if (gmg_) { gmg_setup(); gmg_tail_setup(); }
        if (!ident_ && !gmg_) rows_setup(h, wall);   // the identity and GMG / AMG forms run the flat kernels only
        timer_.s = stream_;
        if (pmv_on_) {   // the x update's stream, created before any capture
            HIPCHK(hipStreamCreateWithFlags(&xs_, hipStreamNonBlocking));
            HIPCHK(hipEventCreateWithFlags(&xfork_, hipEventDisableTiming));
            HIPCHK(hipEventCreateWithFlags(&xjoin_, hipEventDisableTiming));
        }
        allocate();
        tail_setup();
        rows_coarse_setup();
    }
    ~Engine() override { destroy_graph(); destroy_xstream(); release(); }

    void set_opts(const uaamg_opts& o) override {
        int all_neumann = 1;
        for (int f = 0; f < 2 * D; f++) if (wall_[f] != UAAMG_WALL_NEUMANN) all_neumann = 0;
        o_.nu_pre = o.nu_pre; o_.nu_post = o.nu_post; o_.coarsest_sweeps = o.coarsest_sweeps;
        o_.tol_mode = o.tol_mode; o_.max_iter = o.max_iter; o_.tol = o.tol; o_.fused = o.fused;
        o_.check_every = o.check_every; o_.use_graph = o.use_graph; o_.timing = o.timing; o_.flexible = o.flexible;
        o_.project_mean = o.project_mean;
        // launch-time forms only: the structural ones (kStructForms, rows_kc,
        // march_*) were fixed at construction, a change rebuilds the engine
        o_.forms = (o_.forms & kStructForms) | (o.forms & ~kStructForms);
        o_.xfork_level = o.xfork_level; o_.xblocks = o.xblocks; o_.xthreads = o.xthreads; o_.trace = o.trace;
        project_ = (o.project_mean < 0) ? all_neumann : o.project_mean;
        destroy_graph();
    }
    void set_stream(hipStream_t s) override { stream_ = s; timer_.s = s; }
    const uint8_t* solid_dev() const override { return solid0_; }
    int nlevels() const override { return (int)shape_.size(); }
    Geo geo(int l) const override { return shape_[l]; }
    size_t bytes() const override { return bytes_; }

    // ---------------- set-up ----------------
    void set_variable_lap(const double* d, const double* ox, const double* oy) override {
        if (D != 2) throw HipError{hipErrorInvalidValue, "variable Laplacian: 2D grids only (grid.h:21-23)"};
        destroy_graph();
        if (!d) { vlap_ = VarLap{}; return; }
        if (!vlap_buf_) vlap_buf_ = dalloc<double>(3 * (size_t)g0_.N);
        const size_t n = (size_t)g0_.N * sizeof(double);
        HIPCHK(hipMemcpyAsync(vlap_buf_, d, n, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipMemcpyAsync(vlap_buf_ + g0_.N, ox, n, hipMemcpyDeviceToDevice, stream_));
        HIPCHK(hipMemcpyAsync(vlap_buf_ + 2 * g0_.N, oy, n, hipMemcpyDeviceToDevice, stream_));
        vlap_.d = vlap_buf_; vlap_.ox = vlap_buf_ + g0_.N; vlap_.oy = vlap_buf_ + 2 * g0_.N;
    }

    void set_grid(const uint8_t* solid_dev) override {
        const int nl = nlevels();
        vlap_ = VarLap{};   // a new grid drops the variable Laplacian
        HIPCHK(hipMemcpyAsync(solid0_, solid_dev, g0_.N, hipMemcpyDeviceToDevice, stream_));
        k_nbr0<D><<<lc0_.grid, lc0_.block, 0, stream_>>>(solid0_, nbr0_, g0_, lc0_.kz, partials_);
        k_fin_count<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)lc0_.nblocks(), count_);
        HIPCHK(hipGetLastError());
        for (int l = 1; l < nl; l++) {
            Level& F = lv_[l - 1]; Level& C = lv_[l];
            const uint8_t* fm = (l == 1) ? nbr0_ : F.mask;
            k_coarsen_mask<D><<<C.lc.grid, C.lc.block, 0, stream_>>>(fm, C.mask, F.g, C.g, C.lc.kz, l == 1,
                                                                     o_.hierarchy_mode == UAAMG_HIERARCHY_KEEP_FLUID);
            HIPCHK(hipGetLastError());
            if (l == 1) {
                FineL0 src{nbr0_, c0d_, g0_};
                k_galerkin<D, FineL0, C4, S><<<C.lc.grid, C.lc.block, 0, stream_>>>(src, F.g, C.mask, C.coef, C.ex, C.g, C.lc.kz);
            } else {
                FineLc<C4, S> src{F.coef, F.ex};
                k_galerkin<D, FineLc<C4, S>, C4, S><<<C.lc.grid, C.lc.block, 0, stream_>>>(src, F.g, C.mask, C.coef, C.ex, C.g, C.lc.kz);
            }
            k_level_diag<D, C4, S><<<C.lc.grid, C.lc.block, 0, stream_>>>(C.coef, C.ex, C.mask, C.g, C.lc.kz);
            HIPCHK(hipGetLastError());
        }
        tail_build();
        destroy_graph();
    }

    // ---------------- V-cycle ----------------
    // Level 0 right-hand side is the CG residual r, its solution is z.
    void vcycle(const T* r, T* z, const int* skip) {
        if (ident_) {   // key "cg": z = r (identity_preconditioner.h:19)
            zsums_fused_ = false;
            k_copy<T><<<flat_blocks(2048), 256, 0, stream_>>>(r, z, N4_, skip);
            HIPCHK(hipGetLastError());
            return;
        }
        if (gmg_) {
            zsums_fused_ = false;
            vcycle_gmg(r, z, skip);
            return;
        }
        zsums_fused_ = o_.fused && o_.nu_pre == 1 && o_.nu_post == 1 && fused_ok_;
        if (zsums_fused_) vcycle_fused(r, z, skip);
        else vcycle_pass(r, z, skip);
    }

    void vcycle_pass(const T* r, T* z, const int* skip) {
        const int nl = nlevels();
        fill_async(z, g0_.N, skip);
        for (int l = 1; l < nl; l++) fill_async(lv_[l].x, lv_[l].g.N, skip);
        LevelArgs<T, C4> L0{g0_, const_cast<T*>(r), z, nbr0_, nullptr};
        if (nl == 1) { coarsest_pass<true>(L0, skip); return; }
        // down
        for (int l = 0; l < nl - 1; l++) {
            TScope ts(timer_, l == 0 ? UAAMG_K_DOWN0 : UAAMG_K_COARSE);
            Level& C = lv_[l + 1];
            LevelArgs<S, C4> LC{C.g, C.b, C.x, C.mask, C.coef};
            if (l == 0) {
                smooth_pass<true>(L0, o_.nu_pre, false, skip);
                k_restrict_skip<D, true>(L0, LC, C.lc, skip);
            } else {
                Level& F = lv_[l];
                LevelArgs<S, C4> LF{F.g, F.b, F.x, F.mask, F.coef};
                smooth_pass<false>(LF, o_.nu_pre, false, skip);
                k_restrict_skip<D, false>(LF, LC, C.lc, skip);
            }
        }
        {
            TScope ts(timer_, UAAMG_K_COARSE);
            Level& C = lv_[nl - 1];
            LevelArgs<S, C4> LC{C.g, C.b, C.x, C.mask, C.coef};
            coarsest_pass<false>(LC, skip);
        }
        for (int l = nl - 2; l >= 0; l--) {
            TScope ts(timer_, l == 0 ? UAAMG_K_UP0 : UAAMG_K_COARSE);
            Level& C = lv_[l + 1];
            LevelArgs<S, C4> LC{C.g, C.b, C.x, C.mask, C.coef};
            if (l == 0) {
                prolong_skip<true>(L0, LC, skip);
                smooth_pass<true>(L0, o_.nu_post, true, skip);
            } else {
                Level& F = lv_[l];
                LevelArgs<S, C4> LF{F.g, F.b, F.x, F.mask, F.coef};
                prolong_skip<false>(LF, LC, skip);
                smooth_pass<false>(LF, o_.nu_post, true, skip);
            }
        }
    }

    // ---- GMG / AMG (keys "pcg", "pcg_gmg", "pcg_amg"; 2D) ----
    // The same hierarchy shape and coarse solid masks as the UAAMG levels
    // (gmg_preconditioner.cpp:13-39), a rediscretised operator per level
    // (h doubles per level) and V(2,2) with coarsest_sweeps (forward +
    // reverse) pairs on the coarsest level (:163-174).  Level 0 smooths z in
    // place against r; the coarse levels use their own b and x.
    std::vector<GmgC<double>> gmgc_;
    void gmg_setup() {
        gmgc_.clear();
        double dx = h_[0], dy = h_[1];
        for (size_t l = 0; l < shape_.size(); l++) {
            GmgC<double> c;
            c.idx2 = 1.0 / (dx * dx); c.idy2 = 1.0 / (dy * dy);
            // the two classes' diagonals, each as its source writes it
            // (gmg_preconditioner.cpp:43, amg_preconditioner.cpp:15)
            c.diag = gmg_ == 2 ? 2.0 * c.idx2 + 2.0 * c.idy2 : 2.0 * (c.idx2 + c.idy2);
            gmgc_.push_back(c);
            dx *= 2.0; dy *= 2.0;
        }
    }
    template <typename V>
    GmgC<typename CTOf<V>::type> gmg_c(int l) const {
        using X = typename CTOf<V>::type;
        return GmgC<X>{(X)gmgc_[l].idx2, (X)gmgc_[l].idy2, (X)gmgc_[l].diag};
    }
    // levels from gmg_lt_ on run in one workgroup (k_gmg_tail), the ones
    // above it with one tiled launch down and one up per level
    int gmg_lt_ = 0;
    GmgTailArgs gmg_targs_{};
    size_t gmg_tail_lds_ = 0;
    void gmg_tail_setup() {
        const int nl = nlevels();
        const size_t cb = (sizeof(T) == 8 || sizeof(S) == 8) ? 8 : 4;   // LDS element size, the larger form
        size_t bytes = 0;
        int lt = nl;
        while (lt > 0 && nl - (lt - 1) <= GMG_TAIL_MAXL) {
            const size_t add = (size_t)shape_[lt - 1].N * (2 * cb + 1);
            if (bytes + add > (size_t)GMG_TAIL_LDS) break;
            bytes += add;
            lt--;
        }
        if (lt == nl)   // a coarsest level of side <= coarsest_side in one dimension only, long in the other
            throw HipError{hipErrorInvalidValue, "GMG / AMG: coarsest level too large for the one-workgroup solve"};
        gmg_lt_ = lt;
    }
    void gmg_tail_args(int csweeps) {
        const int nl = nlevels();
        GmgTailArgs& A = gmg_targs_;
        A.nl = nl - gmg_lt_;
        int off = 0;
        for (int q = 0; q < A.nl; q++) {
            const int l = gmg_lt_ + q;
            A.g[q] = shape_[l];
            A.c[q] = gmgc_[l];
            A.mask[q] = l == 0 ? solid0_ : lv_[l].mask;
            A.off[q] = off;
            off += (int)shape_[l].N;
        }
        A.total = off;
        A.csweeps = csweeps;
        const size_t cb = gmg_lt_ == 0 ? sizeof(typename CTOf<T>::type) : sizeof(typename CTOf<S>::type);
        gmg_tail_lds_ = (size_t)off * (2 * cb + 1);
    }
    static dim3 gmg_tiles(const Geo& g) { return dim3((unsigned)((g.nx + GT_X - 1) / GT_X), (unsigned)((g.ny + GT_Y - 1) / GT_Y)); }
    template <typename VX>
    void gmg_down(VX* x, const VX* b, const uint8_t* m, int l, const int* skip) {
        Level& C = lv_[l + 1];
        auto kern = gmg_ == 2 ? k_gmg_down<true, VX, S> : k_gmg_down<false, VX, S>;
        kern<<<gmg_tiles(shape_[l]), GT_THREADS, 0, stream_>>>(x, b, m, shape_[l], gmg_c<VX>(l), C.b, C.mask, C.g, skip);
        HIPCHK(hipGetLastError());
    }
    template <typename VX>
    void gmg_up(VX* x, const VX* b, const uint8_t* m, int l, const int* skip) {
        Level& C = lv_[l + 1];
        auto kern = gmg_ == 2 ? k_gmg_up<true, VX, S> : k_gmg_up<false, VX, S>;
        kern<<<gmg_tiles(shape_[l]), GT_THREADS, 0, stream_>>>(x, b, m, shape_[l], gmg_c<VX>(l), C.x, C.mask, C.g, skip);
        HIPCHK(hipGetLastError());
    }
    template <typename VF>
    void gmg_tail(VF* x, const VF* b, const int* skip) {
        gmg_tail_args(std::max(0, o_.coarsest_sweeps));
        auto kern = gmg_ == 2 ? k_gmg_tail<true, VF> : k_gmg_tail<false, VF>;
        kern<<<1, GMG_TAIL_THREADS, gmg_tail_lds_, stream_>>>(gmg_targs_, x, b, skip);
        HIPCHK(hipGetLastError());
    }
    // z = M r: levels 0 .. lt-1 down (x of each level and b of the next
    // written), the tail from lt, levels lt-1 .. 0 up; every cell of every
    // x the V-cycle reads is written earlier in the same V-cycle, so nothing
    // is zero-filled
    void vcycle_gmg(const T* r, T* z, const int* skip) {
        if constexpr (D == 2 && sizeof(S) >= 4) {
            TScope ts(timer_, UAAMG_K_DOWN0);
            const int lt = gmg_lt_;
            for (int l = 0; l < lt; l++) {
                if (l == 0) gmg_down<T>(z, r, solid0_, 0, skip);
                else gmg_down<S>(lv_[l].x, lv_[l].b, lv_[l].mask, l, skip);
            }
            if (lt == 0) gmg_tail<T>(z, r, skip);
            else gmg_tail<S>(lv_[lt].x, lv_[lt].b, skip);
            for (int l = lt - 1; l >= 0; l--) {
                if (l == 0) gmg_up<T>(z, r, solid0_, 0, skip);
                else gmg_up<S>(lv_[l].x, lv_[l].b, lv_[l].mask, l, skip);
            }
        }
    }

    // ---- fused V-cycle: one down and one up launch per level, LDS coarsest ----
    // 2.5D marching kernels: in-plane region RX x RY (tile + 2-ring), PPT
    // positions per thread, march over KC planes of the slowest axis
    static constexpr int RX0 = (D == 3) ? 64 : 256, RY0 = (D == 3) ? 16 : 1, PPT0 = (D == 3) ? 4 : 1;
    static constexpr int RXC = (D == 3) ? 32 : 256, RYC = (D == 3) ? 8 : 1, PPTC = 1;
    static constexpr int CMAX = 2048;

    struct MLaunch { dim3 grid; int KC; };
    static MLaunch march_launch(const Geo& g, int RX, int RY, long target, int kc_min = 2) {
        int TX = RX - 4, TY = (D == 3) ? RY - 4 : 1;
        unsigned ta = (unsigned)((g.nx + TX - 1) / TX), tb = (D == 3) ? (unsigned)((g.ny + TY - 1) / TY) : 1u;
        int nm = (D == 3) ? g.nz : g.ny;
        long tiles = (long)ta * tb;
        // small levels are latency-bound: short chunks (down to kc_min planes)
        // give more workgroups and fewer sequential plane steps
        kc_min = std::max(2, kc_min / 2 * 2);
        long chunks = std::max<long>(1, target / std::max<long>(1, tiles));
        int KC = (int)((nm + chunks - 1) / chunks);
        KC = std::max(KC, kc_min);
        KC += (KC & 1);
        MLaunch m;
        m.KC = KC;
        m.grid = dim3(ta, tb, (unsigned)((nm + KC - 1) / KC));
        return m;
    }
    MLaunch mlaunch(const Geo& g, int RX, int RY) const {
        return march_launch(g, RX, RY, std::max(1, o_.march_target), o_.march_kc_min);
    }
    // ---- batch membership: this grid's buffers and launch shapes ----
    bool batch_ready() const {
        return std::is_same<T, float>::value && !ident_ && !gmg_ && fused_ok_ && tail_first_ > 0 && o_.nu_pre == 1 && o_.nu_post == 1
            && nlevels() <= BATCH_MAXL;
    }
    double alg_bytes_of(int iters) const { return alg_bytes(iters); }
    long padded_cells() const { return g0_.N; }
    long n4() const { return N4_; }
    int rows_cpl() const { return rows_cpl_; }
    int tail_first() const { return tail_first_; }
    // fill the batch descriptor (ctrl: the batch's scalar slot for this grid)
    void fill_batch(BGrid<S>& d, Ctrl* ctrl) const {
        if constexpr (std::is_same<T, float>::value) {
            d.g0 = g0_; d.N4 = N4_;
            d.r = r_; d.z = z_; d.p = p_; d.Ap = Ap_; d.x = x_;
            d.nbr0 = nbr0_;
            d.c0d = c0d_;
            d.c0 = c0_;
            d.ctrl = ctrl;
            d.count = count_;
            d.nl = nlevels();
            d.lt = tail_first_;
            for (int l = 0; l < d.nl && l < BATCH_MAXL; l++) {
                d.lg[l] = lv_[l].g;
                d.lb[l] = lv_[l].b; d.lx[l] = lv_[l].x; d.lcoef[l] = lv_[l].coef;
                d.lKC[l] = 8;
            }
            d.cpl = rows_cpl_;
            d.rtab = rtab_;
            d.rgeo = rgeo_;
            d.tail = targs_;
            d.tail.npass = (o_.coarsest_sweeps > 0) ? 2 * o_.coarsest_sweeps + 1 : 0;
            d.tail.skip = &ctrl->done;
        }
    }

    // ---- tail: the small coarse levels in one workgroup ----
    static constexpr int TAIL_NLMAX = (D == 3) ? (sizeof(S) == 8 ? 2 : 3) : (sizeof(S) == 8 ? 3 : 5);
    int tail_first_ = 0;        // first level of the tail (0: no tail)
    TailArgs<C4> targs_{};
    // First level of the tail kernel for a hierarchy (0: none fits): the
    // largest tail whose first level has <= 2048 cell pairs and <= 8192
    // padded cells, deeper levels <= 1024 pairs, at most TAIL_NLMAX levels and
    // x, b, residual within the LDS.
    static int choose_tail(const std::vector<Geo>& shape) {
        const int nl = (int)shape.size();
        auto pairs = [&](int l) {
            const Geo& g = shape[l];
            return (long)((g.nx + 1) / 2) * g.ny * (D == 3 ? g.nz : 1);
        };
        for (int lt = 1; lt <= nl - 1; lt++) {
            if (nl - lt > TAIL_NLMAX) continue;
            if (pairs(lt) > (long)TAIL_MAXP * TAIL_THREADS) continue;
            if (shape[lt].N > (long)TAIL_BMAX * TAIL_THREADS) continue;
            bool fit = true;
            for (int l = lt + 1; l < nl; l++) fit = fit && pairs(l) <= TAIL_THREADS;
            long elems = 0;
            for (int l = lt; l < nl; l++) elems += 2 * shape[l].N;
            elems += shape[lt].N;
            if (!fit || elems * (long)sizeof(SC) > TAIL_LDS_BYTES) continue;
            return lt;
        }
        return 0;
    }
    void tail_setup() {
        tail_first_ = 0;
        if (!(o_.forms & UAAMG_FORM_TAIL) || ident_ || gmg_) return;
        const int nl = nlevels();
        tail_first_ = choose_tail(shape_);
        if (!tail_first_) return;
        targs_.nl = nl - tail_first_;
        int off = 0;
        for (int q = 0; q < targs_.nl; q++) {
            const Level& L = lv_[tail_first_ + q];
            targs_.g[q] = L.g;
            targs_.coef[q] = L.coef;
            targs_.xo[q] = off; off += (int)L.g.N;
            targs_.bo[q] = off; off += (int)L.g.N;
        }
        targs_.ro = off;
        targs_.npass = (o_.coarsest_sweeps > 0) ? 2 * o_.coarsest_sweeps + 1 : 0;
        targs_.b = lv_[tail_first_].b;
        targs_.x = lv_[tail_first_].x;
        targs_.skip = nullptr;
        // per-thread stencil records (filled by tail_build once the level
        // records exist)
        tail_rec_ = dalloc<char>((size_t)tail_rec_bytes<SC>(TAIL_MAXP + targs_.nl - 1));
        targs_.rec = tail_rec_;
    }
    char* tail_rec_ = nullptr;
    // after the level records of a grid are built
    void tail_build() {
        if (!tail_first_) return;
        k_tail_records<D, SC, C4><<<TAIL_MAXP + targs_.nl - 1, TAIL_THREADS, 0, stream_>>>(targs_, reinterpret_cast<SC*>(tail_rec_));
        HIPCHK(hipGetLastError());
    }
    void tail_launch(const int* skip) {
        targs_.npass = (o_.coarsest_sweeps > 0) ? 2 * o_.coarsest_sweeps + 1 : 0;
        targs_.skip = skip;
        k_tail<D, S, C4, TAIL_NLMAX><<<1, TAIL_THREADS, 0, stream_>>>(targs_);
    }

    // ---- level-0 row kernels (3D, FP32 vectors, nx <= 256) ----
    static constexpr int ROWS_NW = 10, ROWS_RW = 2;
    int rows_cpl_ = 0;          // cells per lane (0: row kernels not used)
    RowTab rtab_{};
    RowGeo rgeo_{};
    unsigned rows_blocks_ = 0;
    void rows_setup(const double h[3], const int wall[6]) {
        rows_cpl_ = 0;
        if constexpr (D == 3 && std::is_same<T, float>::value) {
            if (!(o_.forms & UAAMG_FORM_ROWS)) return;
            if (nlevels() < 2) return;
            const int nx = n_[0];
            rows_cpl_ = nx <= 128 ? 2 : (nx <= 256 ? 4 : 0);
            if (!rows_cpl_) return;
            const double ix = 1.0 / (h[0] * h[0]), iy = 1.0 / (h[1] * h[1]), iz = 1.0 / (h[2] * h[2]);
            for (int m = 0; m < 64; m++) {
                double d = ix * (((m >> 0) & 1) + ((m >> 1) & 1)) + iy * (((m >> 2) & 1) + ((m >> 3) & 1))
                         + iz * (((m >> 4) & 1) + ((m >> 5) & 1));
                rtab_.d[m] = (float)d;
                rtab_.rinv[m] = d >= 1e-30 ? (float)(1.0 / d) : 0.0f;
            }
            rtab_.idx2 = (float)ix; rtab_.idy2 = (float)iy; rtab_.idz2 = (float)iz;
            for (int m = 0; m < 64; m++)
                ptab_.d[m] = ix * (((m >> 0) & 1) + ((m >> 1) & 1)) + iy * (((m >> 2) & 1) + ((m >> 3) & 1))
                           + iz * (((m >> 4) & 1) + ((m >> 5) & 1));
            ptab_.idx2 = ix; ptab_.idy2 = iy; ptab_.idz2 = iz;
            const double e[6] = {2 * ix, 2 * ix, 2 * iy, 2 * iy, 2 * iz, 2 * iz};
            rtab_.any_dirichlet = 0;
            for (int f = 0; f < 6; f++) {
                const bool on = wall[f] == UAAMG_WALL_DIRICHLET;
                rtab_.ex[f] = on ? (float)e[f] : 0.0f;
                ptab_.ex[f] = on ? e[f] : 0.0;
                rtab_.any_dirichlet |= on;
            }
            ptab_.any_dirichlet = rtab_.any_dirichlet;
            rgeo_ = make_rowgeo(n_[1], n_[2], ROWS_NW * ROWS_RW - 4, o_.rows_kc);
            rows_blocks_ = 8u * (unsigned)rgeo_.per_xcd;
            pmv_on_ = (o_.forms & UAAMG_FORM_PMV) != 0;
            pgeo_ = make_rowgeo(n_[1], n_[2], PMV_NW * PMV_RW - 2, o_.rows_kc);
            pmv_blocks_ = 8u * (unsigned)pgeo_.per_xcd;
        }
    }
    // Tiles of the row kernels: TY rows by KC planes, KC a multiple of 4 chosen
    // so that about 256 tiles (one per CU) cover the level (KC <= 16; more
    // planes per tile amortise the 4 halo plane steps, more tiles fill the chip)
    // unless kc_force > 0 sets it
    static RowGeo make_rowgeo(int ny, int nz, int TY = ROWS_NW * ROWS_RW - 4, int kc_force = 0) {
        RowGeo rg{};
        rg.nty = (ny + TY - 1) / TY;
        int KC = (int)std::min<long>(16, std::max<long>(4, (long)nz * rg.nty / 256 / 4 * 4));
        if (kc_force > 0) KC = std::max(4, kc_force / 4 * 4);
        rg.KC = KC;
        rg.nkc = (nz + KC - 1) / KC;
        rg.ntiles = rg.nty * rg.nkc;
        rg.per_xcd = (rg.ntiles + 7) / 8;
        return rg;
    }
    // record levels on the row kernels: cells per lane per level (0: march kernels)
    int rcpl_[32] = {};
    RowGeo rgl_[32] = {};
    void rows_coarse_setup() {
        for (int l = 0; l < 32; l++) rcpl_[l] = 0;
        if constexpr (D == 3 && std::is_same<T, float>::value) {
            if (!(o_.forms & UAAMG_FORM_ROWS_COARSE)) return;
            const int lt = tail_first_ ? tail_first_ : nlevels() - 1;
            for (int l = 1; l < lt && l < 32; l++) {
                const Geo& g = shape_[l];
                // two cells per lane (64 < nx <= 128): smaller levels stay on the
                // march kernels (too few rows to fill the chip), wider ones would
                // need four cells per lane and more registers than the workgroup has
                if (g.nx <= 64 || g.nx > 128) continue;
                rcpl_[l] = 2;
                rgl_[l] = make_rowgeo(g.ny, g.nz);
            }
        }
    }
    void rows_downc(int l, const int* skip) {
        if constexpr (D == 3 && std::is_same<T, float>::value) {
            Level& F = lv_[l];
            Level& C = lv_[l + 1];
            const RowGeo& rg = rgl_[l];
            FinUpd fin{};
            if (fin_fold_ && l == 1) { fin = FinUpd{partials_, (int)rows_blocks_, ctrl_}; fin_fold_ = false; }
            k_rows_downc<2, ROWS_NW, ROWS_RW, S, C4><<<8u * (unsigned)rg.per_xcd, ROWS_NW * 64, 0, stream_>>>(
                F.b, F.coef, F.g, rg, C.b, C.g, skip, fin);
        }
    }
    // the convergence test of the split update runs in the level-1 row
    // kernel instead of its own one-workgroup launch (k_fin_update, 5.2 us
    // per iteration at 256^3) when level 1 is on the row kernels; set by
    // rows_down0, taken by rows_downc
    bool fin_fold_ = false;
    bool fin_foldable() const {
        const int lt = tail_first_ ? tail_first_ : nlevels() - 1;
        return nlevels() > 2 && lt > 1 && rcpl_[1] != 0;
    }
    void rows_upc(int l, const int* skip) {
        if constexpr (D == 3 && std::is_same<T, float>::value) {
            Level& F = lv_[l];
            Level& C = lv_[l + 1];
            const RowGeo& rg = rgl_[l];
            k_rows_upc<2, ROWS_NW, ROWS_RW, S, C4><<<8u * (unsigned)rg.per_xcd, ROWS_NW * 64, 0, stream_>>>(
                F.b, F.coef, F.g, rg, C.x, C.g, F.x, skip);
        }
    }
    // every lane's cells lie inside the row (nx == 64 cells-per-lane): the
    // row kernels' instantiations without lane masks
    bool rows_full() const { return rows_cpl_ > 0 && g0_.nx == 64 * rows_cpl_; }
    void rows_down0(const T* r, Level& C, const int* skip) {
        if constexpr (D == 3 && std::is_same<T, float>::value) {
            if (fr_old_) {   // alpha from the operator's partials; r = r_old - alpha Ap written to r; r.r -> convergence
                // nontemporal loads of r_old and Ap (last use of both; without
                // UAAMG_FORM_NT_D plain): 34.5 vs 33.5 solves/s at 256^3 in three paired runs
                const bool dnt = (o_.forms & UAAMG_FORM_NT_D) && rows_cpl_ == 4;
                forms_used_ |= dnt ? UAAMG_FORM_NT_D : 0;
                tlaunch(timer_, UAAMG_K_DOWN0, rows_cpl_ == 4 ? (dnt ? (rows_full() ? k_rows_down0r<4, ROWS_NW, ROWS_RW, S, true, true>
                                                                                    : k_rows_down0r<4, ROWS_NW, ROWS_RW, S, true>)
                                                                     : k_rows_down0r<4, ROWS_NW, ROWS_RW, S>)
                                                              : k_rows_down0r<2, ROWS_NW, ROWS_RW, S>,
                        dim3(rows_blocks_), dim3(ROWS_NW * 64), stream_, (const float*)fr_old_, (const uint8_t*)nbr0_, g0_, rtab_,
                        rgeo_, C.b, C.g, skip, (const float*)Ap_, (float*)r, partials_, (const double*)partials_m_,
                        (int)pmv_blocks_, ctrl_, xslot_);
                fin_fold_ = fin_foldable();
                if (!fin_fold_) k_fin_update<false><<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)rows_blocks_, ctrl_);
                return;
            }
            tlaunch(timer_, UAAMG_K_DOWN0, rows_cpl_ == 4 ? k_rows_down0<4, ROWS_NW, ROWS_RW, S> : k_rows_down0<2, ROWS_NW, ROWS_RW, S>,
                    dim3(rows_blocks_), dim3(ROWS_NW * 64), stream_, r, nbr0_, g0_, rtab_, rgeo_, C.b, C.g, skip);
        }
    }
    void rows_up0(const T* r, Level& C, T* z, const int* skip) {
        if constexpr (D == 3 && std::is_same<T, float>::value) {
            // nontemporal loads of r (without UAAMG_FORM_NT_U plain): 34.6 vs
            // 33.9 solves/s at 256^3 in three paired runs
            const bool unt = (o_.forms & UAAMG_FORM_NT_U) && rows_cpl_ == 4;
            forms_used_ |= unt ? UAAMG_FORM_NT_U : 0;
            tlaunch(timer_, UAAMG_K_UP0, rows_cpl_ == 4 ? (unt ? (rows_full() ? k_rows_up0<4, ROWS_NW, ROWS_RW, S, true, true>
                                                                              : k_rows_up0<4, ROWS_NW, ROWS_RW, S, true>)
                                                               : k_rows_up0<4, ROWS_NW, ROWS_RW, S>)
                                                        : k_rows_up0<2, ROWS_NW, ROWS_RW, S>,
                    dim3(rows_blocks_), dim3(ROWS_NW * 64), stream_, r, nbr0_, g0_, rtab_, rgeo_, C.x, C.g, z, skip, partials_);
            // with the split update the next operator pass does this finish
            if (!fr_) k_fin_zsums<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)rows_blocks_, ctrl_);
        }
    }

    long fused_max_blocks() const {
        MLaunch m = mlaunch(g0_, RX0, RY0);
        return (long)m.grid.x * m.grid.y * m.grid.z;
    }
    bool fused_supported() const { return nlevels() >= 2; }

    void vcycle_fused(const T* r, T* z, const int* skip) {
        const int nl = nlevels();
        FusedArgs<D, true, CT, C4> A0{g0_, nbr0_, nullptr, c0_};
        const int lt = tail_first_ ? tail_first_ : nl - 1;   // levels >= lt: tail kernel / coarsest solve
        for (int l = 0; l < lt; l++) {
            Level& C = lv_[l + 1];
            if (l == 0 && rows_cpl_) {
                rows_down0(r, C, skip);
                if (fr_old_ && xfork_level() == 0) fork_x();
            } else if (l == 0) {
                MLaunch m = mlaunch(g0_, RX0, RY0);
                tlaunch(timer_, UAAMG_K_DOWN0, k_down_m<D, true, RX0, RY0, PPT0, T, C4, S, C4>, m.grid, dim3(256), stream_,
                        r, A0, C.b, (const C4*)C.coef, C.g, m.KC, skip);
            } else if (rcpl_[l]) {
                TScope ts(timer_, UAAMG_K_COARSE);
                rows_downc(l, skip);
            } else {
                TScope ts(timer_, UAAMG_K_COARSE);
                Level& F = lv_[l];
                FusedArgs<D, false, SC, C4> A{F.g, nullptr, F.coef, Consts<SC>{}};
                MLaunch m = mlaunch(F.g, RXC, RYC);
                auto kern = k_down_m<D, false, RXC, RYC, PPTC, S, C4, S, C4>;
                if (fr_old_ && xfork_level() == l && !timer_.on && !capturing_) {
                    // the fork event rides on this dispatch's completion: no
                    // marker packet between it and the next kernel on stream_
                    hipExtLaunchKernelGGL(kern, m.grid, dim3(256), 0, stream_, nullptr, xfork_, 0,
                                          F.b, A, C.b, C.coef, C.g, m.KC, skip);
                    fork_recorded_ = true;
                } else {
                    kern<<<m.grid, 256, 0, stream_>>>(F.b, A, C.b, C.coef, C.g, m.KC, skip);
                }
            }
            HIPCHK(hipGetLastError());
            if (fr_old_ && l > 0 && xfork_level() == l) fork_x();
        }
        if (fr_old_ && !xforked_ && xfork_level() >= lt) fork_x();   // before the tail / coarsest solve
        if (tail_first_) {
            TScope ts(timer_, UAAMG_K_COARSE);
            tail_launch(skip);
        } else {
            TScope ts(timer_, UAAMG_K_COARSE);
            Level& C = lv_[nl - 1];
            int npass = (o_.coarsest_sweeps > 0) ? 2 * o_.coarsest_sweeps + 1 : 0;
            if (C.g.N <= CMAX) {
                k_coarsest<D, CMAX, S, C4><<<1, 256, 0, stream_>>>(C.x, C.b, C.coef, C.g, npass, skip);
            } else {
                fill_async(C.x, C.g.N, skip);
                LevelArgs<S, C4> LC{C.g, C.b, C.x, C.mask, C.coef};
                coarsest_pass<false>(LC, skip);
            }
            HIPCHK(hipGetLastError());
        }
        for (int l = lt - 1; l >= 0; l--) {
            Level& C = lv_[l + 1];
            if (l == 0 && rows_cpl_) {
                rows_up0(r, C, z, skip);
            } else if (l == 0) {
                MLaunch m = mlaunch(g0_, RX0, RY0);
                tlaunch(timer_, UAAMG_K_UP0, k_up_m<D, true, RX0, RY0, PPT0, true, T, C4, T, S>, m.grid, dim3(256), stream_,
                        r, A0, (const S*)C.x, C.g, z, m.KC, skip, partials_);
                k_fin_zsums<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)(m.grid.x * m.grid.y * m.grid.z), ctrl_);
            } else if (rcpl_[l]) {
                TScope ts(timer_, UAAMG_K_COARSE);
                rows_upc(l, skip);
            } else {
                TScope ts(timer_, UAAMG_K_COARSE);
                Level& F = lv_[l];
                FusedArgs<D, false, SC, C4> A{F.g, nullptr, F.coef, Consts<SC>{}};
                MLaunch m = mlaunch(F.g, RXC, RYC);
                k_up_m<D, false, RXC, RYC, PPTC, false, S, C4, S, S><<<m.grid, 256, 0, stream_>>>(
                    F.b, A, C.x, C.g, F.x, m.KC, skip, nullptr);
            }
            HIPCHK(hipGetLastError());
        }
    }

    // ---------------- solve ----------------
    // Initial residual, its projection, the first V-cycle and p: everything
    // before the first CG iteration (pcg_3d.cpp:22-89), on the bound Ctrl.
    // With the fused p update (pmv_on_ and fused_p), the first p = z - mean is
    // formed by the first iteration's k_rows_pmv instead of k_pupdate.
    void start(const double* rhs_dev, double* p_dev, int warm, int max_iter, double tol, bool fused_p = false) {
        T* x = x_is_p() ? reinterpret_cast<T*>(p_dev) : x_;
        {
            TScope ts(timer_, UAAMG_K_INIT);
            double* hist = nullptr;
            if (trace()) {
                if (hist_n_ < max_iter) {
                    if (hist_) HIPCHK(hipFree(hist_));
                    HIPCHK(hipMalloc(&hist_, sizeof(double) * std::max(1, max_iter)));
                    hist_n_ = max_iter;
                }
                hist = hist_;
            }
            k_ctrl_init<<<1, 64, 0, stream_>>>(ctrl_, count_, max_iter, tol, o_.tol_mode, project_, o_.flexible, hist);
            if (!x_is_p()) {
                if (warm) k_cast_in<T><<<flat_blocks(2048), 256, 0, stream_>>>(p_dev, x_, nbr0_, g0_.N);
                else fill_async(x_, N4_, nullptr);
            }
            fill_async(p_, N4_, nullptr);
            if (fused_p) fill_async(p2_, N4_, nullptr);
            pcur_ = 0;
            fr_old_ = nullptr;
            const unsigned fb = flat_blocks(2048);
            k_init_sum<<<fb, 256, 0, stream_>>>(rhs_dev, nbr0_, N4_, partials_, ctrl_);
            k_fin_init_sum<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)fb, ctrl_);
            k_init_r<T><<<fb, 256, 0, stream_>>>(rhs_dev, r_, nbr0_, N4_, partials_, ctrl_);
            k_fin_init_r<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)fb, ctrl_);
            k_init_sub<T><<<fb, 256, 0, stream_>>>(r_, nbr0_, N4_, partials_, ctrl_);
            k_fin_init_sub<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)fb, ctrl_);
            if (warm) {
                if (vlap_.d) k_init_warm<D, T, true><<<fb, 256, 0, stream_>>>(r_, x, nbr0_, g0_, c0d_, N4_, partials_, ctrl_, vlap_);
                else k_init_warm<D, T><<<fb, 256, 0, stream_>>>(r_, x, nbr0_, g0_, c0d_, N4_, partials_, ctrl_, vlap_);
                k_fin_init_warm<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)fb, ctrl_);
            }
            HIPCHK(hipGetLastError());
        }
        const int* skip = &ctrl_->done;
        vcycle(r_, z_, skip);
        zsums(skip);
        if (!fused_p) pupdate();
    }
    // p += x (cold start) or p = x (warm start) on fluid cells
    void finalize_into(double* p_dev, int warm) {
        if (!x_is_p()) k_finalize<T><<<flat_blocks(2048), 256, 0, stream_>>>(x_, p_dev, nbr0_, g0_.N, warm);
        HIPCHK(hipGetLastError());
    }
    Ctrl* ctrl_ptr() const { return ctrl_; }
    // mark the solve state idle (done): batched kernels skip this grid
    void set_idle() {
        k_ctrl_init<<<1, 64, 0, stream_>>>(ctrl_, count_, 0, 0.0, o_.tol_mode, project_, o_.flexible, nullptr);
        HIPCHK(hipGetLastError());
    }

    int solve(const double* rhs_dev, double* p_dev, int warm, int max_iter, double tol, uaamg_stats* st) override {
        auto t0 = std::chrono::steady_clock::now();
        timer_.on = (o_.timing != 0);
        timer_.mask = (unsigned)o_.timing;
        timer_.reset();
        T* x = x_is_p() ? reinterpret_cast<T*>(p_dev) : x_;
        forms_used_ = 0;
        // the split update needs the fused V-cycle (its level-0 row kernels)
        fr_ = pmv_on_ && !o_.flexible && o_.fused && o_.nu_pre == 1 && o_.nu_post == 1 && fused_ok_;
        start(rhs_dev, p_dev, warm, max_iter, tol, pmv_on_);
        const int* skip = &ctrl_->done;
        int it_cap = max_iter;
        int every = std::max(1, o_.check_every);
        int queued = 0;
        int nchunks = 1;   // chunks queued before the next host check
        int it0 = -1; double rs0 = 0.0;   // iteration count and r.r at the first check
        Ctrl hc{};
        while (true) {
            for (int c = 0; c < nchunks && queued < it_cap; c++) {
                int chunk = std::min(every, it_cap - queued);
                run_chunk(x, skip, chunk, every);
                queued += chunk;
            }
            HIPCHK(hipMemcpyAsync(ctrl_host_, ctrl_, sizeof(Ctrl), hipMemcpyDeviceToHost, stream_));
            HIPCHK(hipStreamSynchronize(stream_));
            std::memcpy(&hc, ctrl_host_, sizeof(Ctrl));
            if (hc.done || queued >= it_cap) break;
            if (g0_.N < kAdaptCells) continue;
            if (it0 < 0) { it0 = hc.iter; rs0 = hc.rsnew; continue; }
            nchunks = chunks_ahead(hc, every, it0, rs0);
        }
        join_x();
        if (!x_is_p()) {
            TScope ts(timer_, UAAMG_K_OTHER);
            k_finalize<T><<<flat_blocks(2048), 256, 0, stream_>>>(x_, p_dev, nbr0_, g0_.N, warm);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(stream_));
        auto t1 = std::chrono::steady_clock::now();
        if (trace() && hc.iter > 0) {
            std::vector<double> h(hc.iter);
            HIPCHK(hipMemcpy(h.data(), hist_, sizeof(double) * hc.iter, hipMemcpyDeviceToHost));
            std::fprintf(stderr, "[uaamg] r0=%.6e", std::sqrt(hc.r0sq));
            for (int q = 0; q < hc.iter; q++) std::fprintf(stderr, "%s%.3e", q % 10 ? " " : "\n[uaamg] ", std::sqrt(h[q] / hc.r0sq));
            std::fprintf(stderr, "\n");
        }
        if (st) {
            st->iters = hc.iter;
            st->status = hc.status == ST_CONVERGED ? UAAMG_OK
                       : hc.status == ST_BREAKDOWN ? UAAMG_BREAKDOWN
                       : hc.status == ST_NONFINITE ? UAAMG_NONFINITE : UAAMG_NOT_CONVERGED;
            st->nlevels = nlevels();
            st->rnorm = std::sqrt(std::max(0.0, hc.rsnew));
            st->r0norm = std::sqrt(std::max(0.0, hc.r0sq));
            st->ms_total = std::chrono::duration<double, std::milli>(t1 - t0).count();
            timer_.collect(st);
            st->alg_bytes = alg_bytes(hc.iter);
            st->kernel_set = fr_ ? 2 : (pmv_on_ ? 1 : 0);
            st->forms_used = forms_used();
        }
        return st ? st->status : 0;
    }

    // Chunks to queue before the next host check: 3/4 of the iterations still
    // needed if r.r keeps falling at its average rate since the first check
    // (log-linear in the iteration count), at least one chunk.  The device
    // stops by itself; the host check only ends the queueing, so a chunk
    // queued past the convergence costs its kernels' early exits, and a host
    // check costs a round trip (at 256^3, 106 iterations: 6 checks instead of
    // 27 with chunks of 4, same overshoot; 32.7 vs 32.4 solves/s).  Only on
    // grids of at least kAdaptCells padded cells: on small grids an
    // iteration's early exits cost about as much as the iteration, and a
    // mixed stream of 2D / 3D grids up to 1024^2 / 160^3 ran 5 % slower with
    // every grid scheduled this way.
    static constexpr long kAdaptCells = 1L << 21;
    static int chunks_ahead(const Ctrl& hc, int every, int it0, double rs0) {
        if (hc.iter <= it0 || !(hc.rsnew > 0.0) || !(rs0 > 0.0)) return 1;
        const double target = (hc.tol_mode == 1) ? hc.tolsq : hc.tol * hc.tol;
        const double rate = std::log(hc.rsnew / rs0) / (hc.iter - it0);   // per iteration, < 0 when falling
        if (!(rate < 0.0) || !(target > 0.0) || !(target < hc.rsnew)) return 1;
        const double rem = std::log(target / hc.rsnew) / rate;
        if (!std::isfinite(rem)) return 1;
        return (int)std::max(1.0, std::min(1e6, std::floor(0.75 * rem / every)));
    }

    // UAAMG_FORM_* bits of the forms the last solve ran
    int forms_used() const {
        int f = forms_used_ | (pmv_on_ ? UAAMG_FORM_PMV : 0) | (fr_ ? UAAMG_FORM_SPLIT : 0);
        if (zsums_fused_) {
            f |= (rows_cpl_ ? UAAMG_FORM_ROWS : 0) | (tail_first_ ? UAAMG_FORM_TAIL : 0);
            const int lt = tail_first_ ? tail_first_ : nlevels() - 1;
            for (int l = 1; l < lt && l < 32; l++) f |= rcpl_[l] ? UAAMG_FORM_ROWS_COARSE : 0;
        }
        return f;
    }

    // persistent grid for the flat level-0 kernels: at most `cap` workgroups of 256
    unsigned flat_blocks(unsigned cap) const {
        long need = (N4_ / 4 + 255) / 256;
        return (unsigned)std::max<long>(1, std::min<long>(need, cap));
    }

    // `chunk` CG iterations; with use_graph the full-size chunk is captured
    // once into a hipGraph and replayed (the graph holds kernel nodes only:
    // the kernels read the device-resident scalars and the done flag)
    // (with the fused p update the captured sequence depends on which buffer
    // holds p when it starts: one graph per starting buffer)
    hipGraph_t graph_[2] = {nullptr, nullptr};
    hipGraphExec_t gexec_[2] = {nullptr, nullptr};
    int graph_chunk_[2] = {0, 0};
    int graph_forms_[2] = {0, 0};
    void destroy_graph() {
        for (int k = 0; k < 2; k++) {
            if (gexec_[k]) (void)hipGraphExecDestroy(gexec_[k]);
            if (graph_[k]) (void)hipGraphDestroy(graph_[k]);
            gexec_[k] = nullptr; graph_[k] = nullptr; graph_chunk_[k] = 0;
        }
        graph_valid_ = false;
    }
    void run_chunk(T* x, const int* skip, int chunk, int every) {
        // Eager chunks need no join here: each iteration joins before the
        // operator pass that overwrites the p buffer an unjoined x pass reads
        // (the alternation continues across chunks).  A graph's own forks and
        // joins assume nothing is pending when it starts.
        if (o_.use_graph && !timer_.on && chunk == every) {
            join_x();   // an x update forked by an eager chunk precedes the graph's own
            const int k = pcur_;
            if (!graph_valid_ || graph_chunk_[k] != chunk) {
                if (gexec_[k]) { (void)hipGraphExecDestroy(gexec_[k]); gexec_[k] = nullptr; }
                if (graph_[k]) { (void)hipGraphDestroy(graph_[k]); graph_[k] = nullptr; }
                const int f0 = forms_used_;
                forms_used_ = 0;
                HIPCHK(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                capturing_ = true;
                for (int c = 0; c < chunk; c++) iteration(x, skip);
                join_x();   // every fork joins the capturing stream before the capture ends
                capturing_ = false;
                HIPCHK(hipStreamEndCapture(stream_, &graph_[k]));
                HIPCHK(hipGraphInstantiate(&gexec_[k], graph_[k], nullptr, nullptr, 0));
                graph_valid_ = true;
                graph_chunk_[k] = chunk;
                graph_forms_[k] = forms_used_;   // the forms the graph's kernels use
                forms_used_ = f0;
                pcur_ = k;   // capture does not run the iterations
            }
            HIPCHK(hipGraphLaunch(gexec_[k], stream_));
            forms_used_ |= graph_forms_[k];
            if (pmv_on_ && (chunk & 1)) pcur_ ^= 1;
            return;
        }
        for (int c = 0; c < chunk; c++) iteration(x, skip);
    }

    void iteration(T* x, const int* skip) {
        const unsigned fb = flat_blocks(2048);
        const T* pv = p_;        // the p of this iteration's operator and update
        unsigned mb = fb;        // workgroups of the operator launch (its partials)
        // With the split update, the x pass of iteration k reads p_k (one of
        // the two p buffers) and its alpha slot Ctrl::xa[buffer]; the operator
        // pass of iteration k+2 is the first to overwrite either.  So the main
        // stream joins the x stream only before the operator passes that write
        // buffer 0: that join covers the x passes of the two iterations before
        // it, and a join costs a barrier packet on the main queue (about 6 us
        // at 256^3).  Other forms join every iteration.
        xforked_ = false;
        if (!fr_ || pcur_ == 1) join_x();
        if (pmv_on_) {
            const T* pold = pbuf(pcur_);
            T* pnew = pbuf(pcur_ ^ 1);
            // p.Ap partials in their own buffer: the down kernel that sums
            // them writes its r.r partials while other workgroups still read
            // (with the split update it also finishes the level-0 up kernel's z sums)
            if constexpr (D == 3 && std::is_same<T, float>::value) {
                // nontemporal loads of z and p_old (last use of both; without
                // UAAMG_FORM_NT_P plain): 33.6 vs 33.0 solves/s at 256^3 in three paired runs
                const bool pnt = (o_.forms & UAAMG_FORM_NT_P) && fr_ && rows_cpl_ == 4;
                forms_used_ |= pnt ? UAAMG_FORM_NT_P : 0;
                auto kern = fr_ ? (rows_cpl_ == 4 ? (pnt ? (rows_full() ? k_rows_pmv<4, PMV_NW, PMV_RW, true, true, true>
                                                                        : k_rows_pmv<4, PMV_NW, PMV_RW, true, true>)
                                                         : k_rows_pmv<4, PMV_NW, PMV_RW, true>)
                                                  : k_rows_pmv<2, PMV_NW, PMV_RW, true>)
                                : (rows_cpl_ == 4 ? k_rows_pmv<4, PMV_NW, PMV_RW, false> : k_rows_pmv<2, PMV_NW, PMV_RW, false>);
                tlaunch(timer_, UAAMG_K_MATVEC, kern, dim3(pmv_blocks_), dim3(PMV_NW * 64), stream_, (const float*)z_,
                        (const float*)pold, (float*)pnew, (float*)Ap_, (const uint8_t*)nbr0_, g0_, ptab_, pgeo_,
                        partials_m_, ctrl_, (const double*)partials_, (int)rows_blocks_);
            }
            pv = pnew;
            mb = pmv_blocks_;
            if (!fr_) k_fin_matvec<<<1, FIN_THREADS, 0, stream_>>>(partials_m_, (int)mb, ctrl_);
        } else {
            tlaunch(timer_, UAAMG_K_MATVEC, vlap_.d ? k_matvec_dot<D, T, true> : k_matvec_dot<D, T>, dim3(fb), dim3(256), stream_,
                    (const T*)p_, Ap_, (const uint8_t*)nbr0_, g0_, c0d_, N4_, partials_, ctrl_, vlap_);
            k_fin_matvec<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)mb, ctrl_);
        }
        if (fr_) {   // r update and convergence test inside the level-0 down kernel
            fr_old_ = rbuf(pcur_);
            xfork_x_ = x;
            xfork_p_ = pv;
            xslot_ = pcur_ ^ 1;   // the buffer pv is in
            T* rnew = rbuf(pcur_ ^ 1);
            pcur_ ^= 1;
            vcycle(rnew, z_, skip);   // forks the x update after the level-0 down kernel
            fr_old_ = nullptr;
            zsums(skip);
            return;
        }
        if (pmv_on_) pcur_ ^= 1;
        if (o_.flexible) {
            tlaunch(timer_, UAAMG_K_UPDATE, k_update<T, true>, dim3(fb), dim3(256), stream_,
                    x, r_, pv, (const T*)Ap_, (const T*)z_, (const uint8_t*)nbr0_, N4_, partials_, ctrl_);
            k_fin_update<true><<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)fb, ctrl_);
        } else {
            tlaunch(timer_, UAAMG_K_UPDATE, k_update<T, false>, dim3(fb), dim3(256), stream_,
                    x, r_, pv, (const T*)Ap_, (const T*)z_, (const uint8_t*)nbr0_, N4_, partials_, ctrl_);
            k_fin_update<false><<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)fb, ctrl_);
        }
        vcycle(r_, z_, skip);
        zsums(skip);
        if (!pmv_on_) pupdate();
    }

    void zsums(const int*) {
        if (zsums_fused_) return;   // the fused level-0 up kernel already reduced
        TScope ts(timer_, UAAMG_K_PUPDATE);
        const unsigned fb = flat_blocks(2048);
        k_zsums<T><<<fb, 256, 0, stream_>>>(z_, r_, nbr0_, N4_, partials_, ctrl_);
        k_fin_zsums<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)fb, ctrl_);
    }
    void pupdate() {
        tlaunch(timer_, UAAMG_K_PUPDATE, k_pupdate<T>, dim3(flat_blocks(2048)), dim3(256), stream_,
                p_, (const T*)z_, (const uint8_t*)nbr0_, N4_, ctrl_);
    }

    // ---------------- test hooks ----------------
    void apply(const double* r_dev, double* z_dev) override {
        k_ctrl_init<<<1, 64, 0, stream_>>>(ctrl_, count_, 1, 0.0, 0, project_, 0, nullptr);
        cast_in_all(r_dev, r_);
        vcycle(r_, z_, nullptr);   // the fused level-0 reductions land in ctrl, unused here
        cast_out_all(z_, z_dev);
        HIPCHK(hipStreamSynchronize(stream_));
    }

    double matvec(const double* v_dev, double* Av_dev) override {
        k_ctrl_init<<<1, 64, 0, stream_>>>(ctrl_, count_, 1, 0.0, 0, project_, 0, nullptr);
        cast_in_all(v_dev, p_);
        if (vlap_.d) k_matvec_dot<D, T, true><<<flat_blocks(2048), 256, 0, stream_>>>(p_, Ap_, nbr0_, g0_, c0d_, N4_, partials_, ctrl_, vlap_);
        else k_matvec_dot<D, T><<<flat_blocks(2048), 256, 0, stream_>>>(p_, Ap_, nbr0_, g0_, c0d_, N4_, partials_, ctrl_, vlap_);
        k_fin_matvec<<<1, FIN_THREADS, 0, stream_>>>(partials_, (int)flat_blocks(2048), ctrl_);
        cast_out_all(Ap_, Av_dev);
        HIPCHK(hipMemcpyAsync(ctrl_host_, ctrl_, sizeof(Ctrl), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        Ctrl hc; std::memcpy(&hc, ctrl_host_, sizeof(Ctrl));
        return hc.pAp;
    }

    void level_data(int level, uint8_t* solid, double* diag, double* cx, double* cy, double* cz, double* ex) override;

private:
    // the iterate always lives in an internal buffer (guard margins, aligned);
    // the caller's p receives p += x (cold start) or p = x (warm start) at the end
    bool x_is_p() const { return false; }

    // zero-fill with a kernel (no memset nodes, so the sequence can be graph-captured)
    template <typename V>
    void fill_async(V* p, long n, const int* skip) {
        long blocks = std::min<long>((n + 255) / 256, 4096);
        k_fill<V><<<(unsigned)std::max<long>(blocks, 1), 256, 0, stream_>>>(p, n, V(0), skip);
        HIPCHK(hipGetLastError());
    }

    template <bool L0, typename V>
    void smooth_pass(LevelArgs<V, C4>& A, int sweeps, bool reverse, const int* skip) {
        Launch lc = cells_launch(A.g, D);
        int p0 = reverse ? 0 : 1;
        for (int sw = 0; sw < sweeps; sw++) {
            k_smooth_skip<L0>(A, p0, lc, skip);
            k_smooth_skip<L0>(A, 1 - p0, lc, skip);
        }
    }
    template <bool L0, typename V>
    void coarsest_pass(LevelArgs<V, C4>& A, const int* skip) {
        for (int s = 0; s < o_.coarsest_sweeps; s++) {
            smooth_pass<L0>(A, 1, false, skip);
            smooth_pass<L0>(A, 1, true, skip);
        }
    }
    template <bool L0, typename V>
    void k_smooth_skip(LevelArgs<V, C4>& A, int parity, const Launch& lc, const int* skip) {
        k_smooth<D, L0, V, C4><<<lc.grid, lc.block, 0, stream_>>>(A, consts_for<V>(), parity, lc.kz, skip);
        HIPCHK(hipGetLastError());
    }
    template <int DD, bool L0, typename V>
    void k_restrict_skip(LevelArgs<V, C4>& F, LevelArgs<S, C4>& Cl, const Launch& lc, const int* skip) {
        k_restrict<DD, L0, V, C4, S, C4><<<lc.grid, lc.block, 0, stream_>>>(F, Cl, consts_for<V>(), lc.kz, skip);
        HIPCHK(hipGetLastError());
    }
    template <bool L0, typename V>
    void prolong_skip(LevelArgs<V, C4>& F, LevelArgs<S, C4>& Cl, const int* skip) {
        Launch lc = cells_launch(F.g, D);
        k_prolong<D, L0, V, C4, S, C4><<<lc.grid, lc.block, 0, stream_>>>(F, Cl, lc.kz, skip);
        HIPCHK(hipGetLastError());
    }
    template <typename V>
    Consts<typename CTOf<V>::type> consts_for() const {
        if constexpr (std::is_same<typename CTOf<V>::type, CT>::value) return c0_;
        else return make_consts<typename CTOf<V>::type>(D, h_, wall_);
    }

    void cast_in_all(const double* src, T* dst);
    void cast_out_all(const T* src, double* dst);

    double alg_bytes(int iters) const {
        // canonical count of the reference algorithm's logical passes, each array read or written once per pass:
        // per iteration 13 s_v + 2 bytes (CG body) + level-0 V-cycle + coarse levels.
        double sv = sizeof(T), sl = sizeof(S), sc = sizeof(R);
        double rho = (D == 3) ? 1.0 / 8.0 : 1.0 / 4.0, ch = (D == 3) ? 4 : 3;
        double N0 = (double)n_[0] * n_[1] * (D == 3 ? n_[2] : 1);
        double cg = 13 * sv + 2;
        double l0 = 3 * sv + 2 * sl + 2 + 2 * rho * sl;
        double coarse = (5 * sl + 2 + 2 * ch * sc + 2 * rho * sl) * (rho / (1 - rho));
        double per_iter = ident_ ? cg + 2 * sv : cg + l0 + coarse;   // identity: z = r copy instead of the V-cycle
        if (gmg_) {
            // GMG / AMG V(2,2): per level 8 colour passes (x and mask read,
            // half of b read and of x written), restriction (fine x, b, mask
            // read, coarse b written), prolongation (fine x and mask read, x
            // written, coarse x read), the x fill
            auto lev = [&](double s, double sn) { return 8 * (2 * s + 1) + (2 * s + 1 + rho * sn) + (2 * s + 1 + rho * sn) + s; };
            per_iter = cg + lev(sv, sl) + lev(sl, sl) * (rho / (1 - rho));
        }
        if (vlap_.d) per_iter += 3 * 8;                              // the FP64 coefficient arrays of the operator
        double setup = 3 * sv + 3 + (1 + ch * sc) * (rho / (1 - rho));
        double init = 48;
        return N0 * (setup + init + iters * per_iter);
    }

    void allocate();
    void release();

    uaamg_opts o_;
    hipStream_t stream_;
    int n_[3]; double h_[3]; int wall_[6];
    int project_ = 1;
    std::vector<Geo> shape_;
    Geo g0_;
    long N4_ = 0;          // level-0 padded length rounded up to a multiple of 4 (flat vector kernels)
    Launch lc0_;
    Consts<CT> c0_;
    Consts<double> c0d_;
    std::vector<Level> lv_;
    uint8_t* solid0_ = nullptr; uint8_t* nbr0_ = nullptr;
    T *r_ = nullptr, *z_ = nullptr, *p_ = nullptr, *Ap_ = nullptr, *x_ = nullptr;
    // fused p update + operator (k_rows_pmv): p alternates between p_ and p2_
    // With fr_ (not flexible) the CG update is split: r -= alpha Ap moves into
    // the level-0 down kernel (r alternates between r_ and r2_ like p) and
    // x += alpha p runs on a second stream.
    T* p2_ = nullptr;
    T* r2_ = nullptr;
    double* partials_m_ = nullptr;       // p.Ap partials of k_rows_pmv
    bool pmv_on_ = false;
    bool fr_ = false;                    // this solve uses the split update
    const T* fr_old_ = nullptr;          // r_old of the fused down kernel (nullptr: plain)
    // The x update (x += alpha p, k_xupdate) runs on a second stream, forked
    // inside the V-cycle so it overlaps the latency-bound coarse levels, and
    // joined before the next operator pass (or the end of a captured chunk /
    // of the solve).  Default priority, like stream_.
    hipStream_t xs_ = nullptr;
    // Fork point (opts.xfork_level): after the down kernel of that level, or
    // (any value >= the tail level) right before the tail.  The update runs
    // as one 256-thread workgroup per CU (opts.xblocks x opts.xthreads): it
    // then takes longer but leaves the coarse kernels room.  Measured at
    // 256^3 (solves/s): 2048 workgroups forked before the tail 31.5; 256
    // workgroups forked after level 2 32.4; 192 or 128 workgroups 30.3;
    // 256 x 512 or 1024 threads 31.8 / 30.4; forked after level 0 with 128
    // workgroups 26.7.  The update reads no mask unless UAAMG_FORM_XMASK is
    // set: the operator pass writes p = 0 on every non-fluid cell and p's
    // ghosts stay 0, so x + alpha p leaves those cells of x as they are.
    int xfork_level() const { return std::max(0, o_.xfork_level); }
    hipEvent_t xfork_ = nullptr, xjoin_ = nullptr;
    bool xpend_ = false;
    T* xfork_x_ = nullptr;
    const T* xfork_p_ = nullptr;
    bool fork_recorded_ = false;   // xfork_ was recorded by the preceding dispatch itself
    bool xforked_ = false;         // this iteration's x pass was forked
    int xslot_ = 0;                // Ctrl::xa slot (p buffer) of this iteration's x pass
    bool capturing_ = false;       // a hipGraph capture of stream_ is open
    void fork_x() {
        if (!fork_recorded_) HIPCHK(hipEventRecord(xfork_, stream_));
        fork_recorded_ = false;
        HIPCHK(hipStreamWaitEvent(xs_, xfork_, 0));
        const unsigned xthreads = (unsigned)std::min(1024, std::max(64, o_.xthreads / 64 * 64));
        const long xneed = (N4_ / 4 + xthreads - 1) / xthreads;
        const unsigned xb = (unsigned)std::max<long>(1, std::min<long>(xneed, std::max(8, o_.xblocks)));
        // nontemporal loads and stores (without UAAMG_FORM_NT_X plain): x and p
        // are used once here, and the kernels running beside the update keep
        // their working set in the caches (33.3 vs 32.9 solves/s at 256^3)
        const bool xnb = (o_.forms & UAAMG_FORM_XMASK) != 0;
        const bool xnt = !xnb && (o_.forms & UAAMG_FORM_NT_X);
        forms_used_ |= (xnb ? UAAMG_FORM_XMASK : 0) | (xnt ? UAAMG_FORM_NT_X : 0);
        tlaunch(timer_, UAAMG_K_UPDATE, xnb ? k_xupdate<T, true> : (xnt ? k_xupdate<T, false, true> : k_xupdate<T, false>), dim3(xb),
                dim3(xthreads), xs_, xfork_x_, xfork_p_, (const uint8_t*)nbr0_, N4_, (const double*)&ctrl_->xa[xslot_]);
        HIPCHK(hipEventRecord(xjoin_, xs_));
        xpend_ = true;
        xforked_ = true;
    }
    void join_x() {
        if (!xpend_) return;
        HIPCHK(hipStreamWaitEvent(stream_, xjoin_, 0));
        xpend_ = false;
    }
    void destroy_xstream() {
        if (xs_) { (void)hipStreamSynchronize(xs_); (void)hipStreamDestroy(xs_); xs_ = nullptr; }
        if (xfork_) { (void)hipEventDestroy(xfork_); xfork_ = nullptr; }
        if (xjoin_) { (void)hipEventDestroy(xjoin_); xjoin_ = nullptr; }
    }
    RowGeo pgeo_{};
    PmvTab ptab_{};
    unsigned pmv_blocks_ = 0;
    int pcur_ = 0;                       // buffer holding the current p and r (0: p_ / r_, 1: p2_ / r2_)
    T* pbuf(int k) const { return k ? p2_ : p_; }
    T* rbuf(int k) const { return k ? r2_ : r_; }
    double* partials_ = nullptr; double* count_ = nullptr;
    Ctrl* ctrl_ = nullptr; Ctrl* ctrl_host_ = nullptr;
    size_t bytes_ = 0;
    bool graph_valid_ = false;
    bool fused_ok_ = false;
    bool zsums_fused_ = false;
    Timer timer_;
    std::vector<void*> allocs_;
    bool trace() const { return o_.trace != 0; }   // relres history to stderr (cf. PCG_DBG, cuda_pcg_3d_opt.cu:384)
    int forms_used_ = 0;                           // UAAMG_FORM_* bits that ran in the current solve
    bool ident_ = false;                           // identity preconditioner (key "cg")
    int gmg_ = 0;                                  // 1: GMG (keys "pcg", "pcg_gmg"), 2: AMG ("pcg_amg")
    VarLap vlap_{};                                // variable Laplacian (2D; d == nullptr: uniform)
    double* vlap_buf_ = nullptr;
    double* hist_ = nullptr; int hist_n_ = 0;

    // level-0 vectors carry a zeroed guard margin on both sides: the 4-cell
    // stencil loads of k_matvec_dot / k_init_warm touch up to sx+4 cells past
    // the first/last interior cell (2D first/last rows)
    long margin_ = 0;
    template <typename P> P* dalloc_margin(size_t n) {
        P* base = dalloc<P>(n + 2 * (size_t)margin_);
        return base + margin_;
    }
    template <typename P> P* dalloc(size_t n) {
        void* p = nullptr;
        size_t bytes = n * sizeof(P);
        if (bytes == 0) bytes = 8;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) throw HipError{e, "hipMalloc"};
        allocs_.push_back(p);
        bytes_ += bytes;
        HIPCHK(hipMemsetAsync(p, 0, bytes, stream_));
        return reinterpret_cast<P*>(p);
    }
};

}  // namespace uaamg
