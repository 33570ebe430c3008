// uaamg.hip -- C-ABI (include/uaamg.h) over the gfx950 engines.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <cmath>
#include <deque>
#include <unordered_map>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <string>
#include "../../include/uaamg.h"
#include "kernels_proj.hpp"
#include "batch.hpp"
#include "engine.hpp"
#include "amgpcg.hpp"

using namespace uaamg;

namespace uaamg {

template <typename A, typename B>
__global__ void k_cast_all(const A* __restrict__ src, B* __restrict__ dst, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x, s = (long)gridDim.x * blockDim.x;
    for (; t < n; t += s) dst[t] = (B)src[t];
}

template <int D, typename T, typename S, typename R>
void Engine<D, T, S, R>::cast_in_all(const double* src, T* dst) {
    long n = g0_.N;
    k_cast_all<double, T><<<(unsigned)std::min<long>((n + 255) / 256, 4096), 256, 0, stream_>>>(src, dst, n);
    HIPCHK(hipGetLastError());
}
template <int D, typename T, typename S, typename R>
void Engine<D, T, S, R>::cast_out_all(const T* src, double* dst) {
    long n = g0_.N;
    k_cast_all<T, double><<<(unsigned)std::min<long>((n + 255) / 256, 4096), 256, 0, stream_>>>(src, dst, n);
    HIPCHK(hipGetLastError());
}

template <int D, typename T, typename S, typename R>
void Engine<D, T, S, R>::allocate() {
    const int nl = nlevels();
    lv_.resize(nl);
    lv_[0].g = g0_;
    lv_[0].lc = lc0_;
    solid0_ = dalloc<uint8_t>(g0_.N);
    // kRowSlack: the row kernels load CPL consecutive cells per lane from a
    // clamped row start, up to CPL - 1 cells past the last ghost cell
    nbr0_ = dalloc<uint8_t>(N4_ + kRowSlack);
    count_ = dalloc<double>(1);
    ctrl_ = dalloc<Ctrl>(1);
    HIPCHK(hipHostMalloc((void**)&ctrl_host_, sizeof(Ctrl), hipHostMallocDefault));
    r_ = dalloc_margin<T>(N4_); z_ = dalloc_margin<T>(N4_); p_ = dalloc_margin<T>(N4_); Ap_ = dalloc_margin<T>(N4_);
    if (!x_is_p()) x_ = dalloc_margin<T>(N4_);
    if (pmv_on_) { p2_ = dalloc_margin<T>(N4_); r2_ = dalloc_margin<T>(N4_); partials_m_ = dalloc<double>(pmv_blocks_ + 64); }
    long maxb = std::max<long>(lc0_.nblocks(), 2048);
    for (int l = 1; l < nl; l++) {
        Level& L = lv_[l];
        L.g = shape_[l];
        L.lc = cells_launch(L.g, D);
        L.b = dalloc<S>(L.g.N + kRowSlack); L.x = dalloc<S>(L.g.N + kRowSlack);
        L.mask = dalloc<uint8_t>(L.g.N);
        L.coef = dalloc<C4>(L.g.N + kRowSlack);
        L.ex = dalloc<S>(L.g.N);
        maxb = std::max(maxb, L.lc.nblocks());
    }
    maxb = std::max(maxb, fused_max_blocks());
    maxb = std::max(maxb, (long)rows_blocks_);
    partials_ = dalloc<double>(3 * maxb + 64);
    fused_ok_ = fused_supported();
    HIPCHK(hipStreamSynchronize(stream_));
}

template <int D, typename T, typename S, typename R>
void Engine<D, T, S, R>::release() {
    (void)hipStreamSynchronize(stream_);
    for (void* p : allocs_) (void)hipFree(p);
    allocs_.clear();
    if (ctrl_host_) (void)hipHostFree(ctrl_host_);
    ctrl_host_ = nullptr;
    if (hist_) (void)hipFree(hist_);
    hist_ = nullptr;
}

template <int D, typename T, typename S, typename R>
void Engine<D, T, S, R>::level_data(int level, uint8_t* solid, double* diag, double* cx, double* cy, double* cz, double* ex) {
    const Geo g = shape_[level];
    const long N = g.N;
    std::vector<uint8_t> m(N);
    std::vector<double> vx(N, 0.0), vy(N, 0.0), vz(N, 0.0), ve(N, 0.0), vd(N, 0.0);
    auto IX = [&](int i, int j, int k) { return (long)i + (long)j * g.sx + (long)k * g.sy; };
    const int klo = (D == 3) ? 1 : 0, khi = (D == 3) ? g.nz : 0;
    if (level == 0) {
        std::vector<uint8_t> nb(N);
        HIPCHK(hipMemcpyAsync(m.data(), solid0_, N, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(nb.data(), nbr0_, N, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        const Consts<double>& c = c0d_;
        for (int i = 1; i <= g.nx; i++)
            for (int j = 1; j <= g.ny; j++)
                for (int k = klo; k <= khi; k++) {
                    long id = IX(i, j, k);
                    if (!(nb[id] & NB_FLUID)) continue;
                    vx[id] = (nb[id] & NB_XP) ? c.idx2 : 0.0;
                    vy[id] = (nb[id] & NB_YP) ? c.idy2 : 0.0;
                    if (D == 3) vz[id] = (nb[id] & NB_ZP) ? c.idz2 : 0.0;
                    double e = 0.0;
                    if (c.wall[0] && i == 1) e += c.ex2[0];
                    if (c.wall[1] && i == g.nx) e += c.ex2[1];
                    if (c.wall[2] && j == 1) e += c.ex2[2];
                    if (c.wall[3] && j == g.ny) e += c.ex2[3];
                    if (D == 3) {
                        if (c.wall[4] && k == 1) e += c.ex2[4];
                        if (c.wall[5] && k == g.nz) e += c.ex2[5];
                    }
                    ve[id] = e;
                }
        for (int i = 1; i <= g.nx; i++)
            for (int j = 1; j <= g.ny; j++)
                for (int k = klo; k <= khi; k++) {
                    long id = IX(i, j, k);
                    if (!(nb[id] & NB_FLUID)) continue;
                    double d = (D == 3) ? vx[id] + vx[id - 1] + vy[id] + vy[id - g.sx] + vz[id] + vz[id - g.sy]
                                        : vx[id] + vx[id - 1] + vy[id] + vy[id - g.sx];
                    if (ve[id] != 0.0) d += ve[id];
                    vd[id] = d;
                }
    } else {
        const Level& L = lv_[level];
        std::vector<C4> cf(N);
        std::vector<S> exv(N);
        HIPCHK(hipMemcpyAsync(m.data(), L.mask, N, hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(cf.data(), L.coef, N * sizeof(C4), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipMemcpyAsync(exv.data(), L.ex, N * sizeof(S), hipMemcpyDeviceToHost, stream_));
        HIPCHK(hipStreamSynchronize(stream_));
        for (int i = 1; i <= g.nx; i++)
            for (int j = 1; j <= g.ny; j++)
                for (int k = klo; k <= khi; k++) {
                    long id = IX(i, j, k);
                    vx[id] = (double)cf[id].x; vy[id] = (double)cf[id].y;
                    if (D == 3) vz[id] = (double)cf[id].z;
                    ve[id] = (double)exv[id];
                    // the record's diagonal as stored (solid cells carry -1, reported as 0)
                    vd[id] = m[id] ? 0.0 : (double)cf[id].w;
                }
    }
    if (solid) std::memcpy(solid, m.data(), N);
    if (diag) std::memcpy(diag, vd.data(), N * sizeof(double));
    if (cx) std::memcpy(cx, vx.data(), N * sizeof(double));
    if (cy) std::memcpy(cy, vy.data(), N * sizeof(double));
    if (cz && D == 3) std::memcpy(cz, vz.data(), N * sizeof(double));
    if (ex) std::memcpy(ex, ve.data(), N * sizeof(double));
}

}  // namespace uaamg

// ------------------------------------------------------------------ C-ABI

struct uaamg_ctx {
    uaamg_opts o{};
    std::string err;
    hipStream_t stream = nullptr;
    std::unique_ptr<EngineBase> eng;
    int dim = 0, n[3] = {0, 0, 0}, wall[6] = {0, 0, 0, 0, 0, 0};
    double h[3] = {0, 0, 0};
    int eng_precision = -1, eng_side = -1;
    uint8_t* solid_stage = nullptr; long solid_n = 0;
    double* rhs_stage = nullptr; double* p_stage = nullptr; long stage_n = 0;
};

static int fail(uaamg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    return code;
}

static int from_hip(uaamg_ctx* c, const HipError& e) {
    std::string m = std::string(hipGetErrorString(e.e)) + " (" + e.where + ")";
    return fail(c, e.e == hipErrorOutOfMemory ? UAAMG_ERR_OOM : UAAMG_ERR_HIP, m);
}

#define CABI_TRY(ctx, ...)                                                       \
    try { __VA_ARGS__ }                                                          \
    catch (const HipError& e) { return from_hip(ctx, e); }                       \
    catch (const std::bad_alloc&) { return fail(ctx, UAAMG_ERR_OOM, "host allocation failed"); }

extern "C" int uaamg_default_opts(uaamg_opts* o) {
    if (!o) return UAAMG_ERR_ARG;
    std::memset(o, 0, sizeof(*o));
    o->precision = UAAMG_FP64;
    o->nu_pre = 1; o->nu_post = 1;
    o->coarsest_side = 4;
    o->coarsest_sweeps = 10;
    o->tol_mode = UAAMG_TOL_ABS;
    o->max_iter = 2000;         // Config::solve_iters (include/config/config.h:30)
    o->tol = 1e-6;              // Config::solve_tol (config.h:31)
    o->project_mean = -1;
    o->fused = 1;
    o->check_every = 4;
    o->use_graph = 0;
    o->timing = 0;
    o->flexible = 0;
    o->forms = UAAMG_FORM_DEFAULT;
    o->xfork_level = 2;
    o->xblocks = 256;
    o->xthreads = 256;
    o->rows_kc = 0;
    o->march_kc_min = 2;
    o->march_target = 1024;
    o->trace = 0;
    o->preconditioner = UAAMG_PRECOND_UAAMG;
    o->hierarchy_mode = UAAMG_HIERARCHY_REFERENCE;
    return UAAMG_OK;
}

static bool opts_valid(const uaamg_opts* o, std::string& why) {
    if (o->precision < UAAMG_FP64 || o->precision > UAAMG_MIXED_F64) { why = "precision"; return false; }
    if (o->precision == UAAMG_MIXED_FP16) {
        why = "precision UAAMG_MIXED_FP16 is not built (1/h^2 couplings exceed the FP16 range; UAAMG_MIXED_BF16 is the 16-bit form)";
        return false;
    }
    if (o->forms & ~(UAAMG_FORM_DEFAULT | UAAMG_FORM_XMASK)) { why = "forms: unknown bits"; return false; }
    if (o->xfork_level < 0 || o->xblocks < 1 || o->xthreads < 64 || o->xthreads > 1024) { why = "x update launch"; return false; }
    if (o->rows_kc < 0 || o->march_kc_min < 0 || o->march_target < 1) { why = "rows_kc / march_*"; return false; }
    if (o->preconditioner < UAAMG_PRECOND_UAAMG || o->preconditioner > UAAMG_PRECOND_AMG) { why = "preconditioner"; return false; }
    if (o->hierarchy_mode != UAAMG_HIERARCHY_REFERENCE && o->hierarchy_mode != UAAMG_HIERARCHY_KEEP_FLUID) { why = "hierarchy_mode"; return false; }
    if (o->nu_pre < 0 || o->nu_post < 0 || o->nu_pre > 16 || o->nu_post > 16) { why = "nu_pre/nu_post"; return false; }
    if (o->coarsest_side < 1) { why = "coarsest_side"; return false; }
    if (o->coarsest_sweeps < 0) { why = "coarsest_sweeps"; return false; }
    if (o->tol_mode != UAAMG_TOL_ABS && o->tol_mode != UAAMG_TOL_REL) { why = "tol_mode"; return false; }
    return true;
}

static int check_device(std::string& why) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) { why = "no HIP device"; return UAAMG_ERR_NODEVICE; }
    int dev = 0;
    (void)hipGetDevice(&dev);
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, dev) != hipSuccess) { why = "hipGetDeviceProperties failed"; return UAAMG_ERR_NODEVICE; }
    if (std::strncmp(p.gcnArchName, "gfx950", 6) != 0) {
        why = std::string("device is ") + p.gcnArchName + ", this library is built for gfx950";
        return UAAMG_ERR_NODEVICE;
    }
    return UAAMG_OK;
}

extern "C" int uaamg_create(uaamg_ctx** out, const uaamg_opts* o) {
    if (!out) return UAAMG_ERR_ARG;
    *out = nullptr;
    uaamg_opts d;
    uaamg_default_opts(&d);
    if (!o) o = &d;
    std::string why;
    if (!opts_valid(o, why)) { std::fprintf(stderr, "uaamg_create: invalid option: %s\n", why.c_str()); return UAAMG_ERR_ARG; }
    int rc = check_device(why);
    if (rc) { std::fprintf(stderr, "uaamg_create: %s\n", why.c_str()); return rc; }
    uaamg_ctx* c = new (std::nothrow) uaamg_ctx();
    if (!c) return UAAMG_ERR_OOM;
    c->o = *o;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return UAAMG_ERR_HIP; }
    *out = c;
    return UAAMG_OK;
}

extern "C" void uaamg_destroy(uaamg_ctx* c) {
    if (!c) return;
    c->eng.reset();
    if (c->solid_stage) (void)hipFree(c->solid_stage);
    if (c->rhs_stage) (void)hipFree(c->rhs_stage);
    if (c->p_stage) (void)hipFree(c->p_stage);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" const char* uaamg_last_error(const uaamg_ctx* c) { return c ? c->err.c_str() : "null context"; }

extern "C" int uaamg_set_opts(uaamg_ctx* c, const uaamg_opts* o) {
    if (!c || !o) return UAAMG_ERR_ARG;
    std::string why;
    if (!opts_valid(o, why)) return fail(c, UAAMG_ERR_ARG, "invalid option: " + why);
    bool rebuild = (o->precision != c->o.precision) || (o->coarsest_side != c->o.coarsest_side) ||
                   ((o->forms ^ c->o.forms) & uaamg::kStructForms) || o->rows_kc != c->o.rows_kc ||
                   o->march_kc_min != c->o.march_kc_min || o->march_target != c->o.march_target ||
                   o->preconditioner != c->o.preconditioner || o->hierarchy_mode != c->o.hierarchy_mode;
    c->o = *o;
    if (rebuild) c->eng.reset();     // a grid must be set again
    else if (c->eng) c->eng->set_opts(*o);
    return UAAMG_OK;
}

static bool grid_valid(const uaamg_grid* g, std::string& why) {
    if (!g) { why = "null grid"; return false; }
    if (g->dim != 2 && g->dim != 3) { why = "dim must be 2 or 3"; return false; }
    for (int q = 0; q < g->dim; q++) {
        if (g->n[q] < 2) { why = "each side needs >= 2 cells"; return false; }
        if (!(g->h[q] > 0)) { why = "spacing must be positive"; return false; }
    }
    for (int f = 0; f < 2 * g->dim; f++)
        if (g->wall[f] != UAAMG_WALL_NEUMANN && g->wall[f] != UAAMG_WALL_DIRICHLET) { why = "wall type"; return false; }
    if (!g->solid) { why = "null solid mask"; return false; }
    long N = (long)(g->n[0] + 2) * (g->n[1] + 2) * (g->dim == 3 ? (g->n[2] + 2) : 1);
    if (N > (1L << 31)) { why = "grid too large for 32-bit block indexing"; return false; }
    return true;
}

static long padded_n(const uaamg_ctx* c) {
    return (long)(c->n[0] + 2) * (c->n[1] + 2) * (c->dim == 3 ? (c->n[2] + 2) : 1);
}

static std::unique_ptr<EngineBase> make_engine(const uaamg_opts& o, const uaamg_grid& g, hipStream_t s) {
    int n[3] = {g.n[0], g.n[1], g.dim == 3 ? g.n[2] : 1};
    double h[3] = {g.h[0], g.h[1], g.dim == 3 ? g.h[2] : 1.0};
    if (g.dim == 2) {
        switch (o.precision) {
            case UAAMG_FP64: return std::make_unique<Engine<2, double, double>>(o, n, h, g.wall, s);
            case UAAMG_FP32: return std::make_unique<Engine<2, float, float>>(o, n, h, g.wall, s);
            // 2D hierarchies keep FP32 levels under MIXED_BF16: with 8-bit mantissas in
            // the level state, 2D solves with solids take 2-3x the iterations or
            // stagnate (measured on the batched mixed-size stream); 3D loses ~10 %
            case UAAMG_MIXED_BF16: return std::make_unique<Engine<2, float, float>>(o, n, h, g.wall, s);
            case UAAMG_MIXED_BF16R: return std::make_unique<Engine<2, float, float>>(o, n, h, g.wall, s);
            case UAAMG_MIXED_F64: return std::make_unique<Engine<2, double, float>>(o, n, h, g.wall, s);
            default: throw HipError{hipErrorInvalidValue, "precision not built"};   // opts_valid rejects it first
        }
    }
    switch (o.precision) {
        case UAAMG_FP64: return std::make_unique<Engine<3, double, double>>(o, n, h, g.wall, s);
        case UAAMG_FP32: return std::make_unique<Engine<3, float, float>>(o, n, h, g.wall, s);
        case UAAMG_MIXED_BF16: return std::make_unique<Engine<3, float, __bf16>>(o, n, h, g.wall, s);
        case UAAMG_MIXED_BF16R: return std::make_unique<Engine<3, float, float, __bf16>>(o, n, h, g.wall, s);
        case UAAMG_MIXED_F64: return std::make_unique<Engine<3, double, float>>(o, n, h, g.wall, s);
        default: throw HipError{hipErrorInvalidValue, "precision not built"};   // opts_valid rejects it first
    }
}

static int set_grid_impl(uaamg_ctx* c, const uaamg_grid* g, bool device_mask) {
    if (!c) return UAAMG_ERR_ARG;
    std::string why;
    if (!grid_valid(g, why)) return fail(c, UAAMG_ERR_ARG, why);
    if (g->dim != 2 && (c->o.preconditioner == UAAMG_PRECOND_GMG || c->o.preconditioner == UAAMG_PRECOND_AMG))
        return fail(c, UAAMG_ERR_ARG, "GMG / AMG preconditioners: 2D grids only (factory_3d.cpp:21-23)");
    CABI_TRY(c, {
        bool same = c->eng && c->dim == g->dim && c->eng_precision == c->o.precision && c->eng_side == c->o.coarsest_side;
        for (int q = 0; q < 3 && same; q++) same = (q < g->dim ? c->n[q] == g->n[q] && c->h[q] == g->h[q] : true);
        for (int f = 0; f < 6 && same; f++) same = (f < 2 * g->dim ? c->wall[f] == g->wall[f] : true);
        if (!same) {
            c->eng.reset();
            c->dim = g->dim;
            for (int q = 0; q < 3; q++) { c->n[q] = (q < g->dim) ? g->n[q] : 1; c->h[q] = (q < g->dim) ? g->h[q] : 1.0; }
            for (int f = 0; f < 6; f++) c->wall[f] = (f < 2 * g->dim) ? g->wall[f] : 0;
            c->eng = make_engine(c->o, *g, c->stream);
            c->eng_precision = c->o.precision;
            c->eng_side = c->o.coarsest_side;
        }
        long N = padded_n(c);
        const uint8_t* src = g->solid;
        if (!device_mask) {
            if (c->solid_n < N) {
                if (c->solid_stage) HIPCHK(hipFree(c->solid_stage));
                c->solid_stage = nullptr;
                HIPCHK(hipMalloc(&c->solid_stage, N));
                c->solid_n = N;
            }
            HIPCHK(hipMemcpyAsync(c->solid_stage, g->solid, N, hipMemcpyHostToDevice, c->stream));
            src = c->solid_stage;
        }
        c->eng->set_grid(src);
        HIPCHK(hipStreamSynchronize(c->stream));
    })
    return UAAMG_OK;
}

extern "C" int uaamg_set_grid(uaamg_ctx* c, const uaamg_grid* g) { return set_grid_impl(c, g, false); }
extern "C" int uaamg_set_grid_device(uaamg_ctx* c, const uaamg_grid* g) { return set_grid_impl(c, g, true); }

static int ensure_stage(uaamg_ctx* c) {
    long N = padded_n(c);
    if (c->stage_n >= N) return UAAMG_OK;
    if (c->rhs_stage) HIPCHK(hipFree(c->rhs_stage));
    if (c->p_stage) HIPCHK(hipFree(c->p_stage));
    c->rhs_stage = c->p_stage = nullptr;
    HIPCHK(hipMalloc(&c->rhs_stage, N * sizeof(double)));
    HIPCHK(hipMalloc(&c->p_stage, N * sizeof(double)));
    c->stage_n = N;
    return UAAMG_OK;
}

extern "C" int uaamg_set_variable_lap(uaamg_ctx* c, const double* lap_diag, const double* lap_off_x,
                                      const double* lap_off_y) {
    if (!c) return UAAMG_ERR_ARG;
    if (!c->eng) return fail(c, UAAMG_ERR_ARG, "no grid set");
    const int nset = (lap_diag != nullptr) + (lap_off_x != nullptr) + (lap_off_y != nullptr);
    if (nset != 0 && nset != 3) return fail(c, UAAMG_ERR_ARG, "variable Laplacian: pass all three arrays or none");
    if (c->dim != 2) return fail(c, UAAMG_ERR_ARG, "variable Laplacian: 2D grids only (grid.h:21-23)");
    CABI_TRY(c, {
        if (nset == 0) {
            c->eng->set_variable_lap(nullptr, nullptr, nullptr);
        } else {
            const long N = padded_n(c);
            double* buf = nullptr;
            HIPCHK(hipMalloc(&buf, 3 * N * sizeof(double)));
            struct Free { double* p; ~Free() { (void)hipFree(p); } } fr{buf};
            HIPCHK(hipMemcpyAsync(buf, lap_diag, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(buf + N, lap_off_x, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
            HIPCHK(hipMemcpyAsync(buf + 2 * N, lap_off_y, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
            c->eng->set_variable_lap(buf, buf + N, buf + 2 * N);
            HIPCHK(hipStreamSynchronize(c->stream));
        }
    })
    return UAAMG_OK;
}

extern "C" int uaamg_solve_device(uaamg_ctx* c, const double* rhs_dev, double* p_dev, int warm_start,
                                  int max_iter, double tol, uaamg_stats* st) {
    if (!c) return UAAMG_ERR_ARG;
    if (!c->eng) return fail(c, UAAMG_ERR_ARG, "no grid set");
    if (!rhs_dev || !p_dev) return fail(c, UAAMG_ERR_ARG, "null array");
    if (max_iter < 0) max_iter = c->o.max_iter;
    if (tol < 0) tol = c->o.tol;
    if (st) std::memset(st, 0, sizeof(*st));
    int rc = UAAMG_OK;
    CABI_TRY(c, { rc = c->eng->solve(rhs_dev, p_dev, warm_start ? 1 : 0, max_iter, tol, st); })
    return rc;
}

// ---------------------------------------------------------------- projection
static size_t mac_n(const uaamg_ctx* c, int comp) {
    const long nx = c->n[0], ny = c->n[1], nz = c->dim == 3 ? c->n[2] : 0;
    if (c->dim == 2) return comp == 0 ? (nx + 1) * (ny + 2) : comp == 1 ? (nx + 2) * (ny + 1) : 0;
    return comp == 0 ? (nx + 1) * (ny + 2) * (nz + 2) : comp == 1 ? (nx + 2) * (ny + 1) * (nz + 2) : (nx + 2) * (ny + 2) * (nz + 1);
}

static int project_impl(uaamg_ctx* c, double* u, double* v, double* w, double* p, double dt, int max_iter,
                        double tol, uaamg_stats* st) {
    const Geo g = c->eng->geo(0);
    const MacGeo m = make_mac(g);
    const uint8_t* solid = c->eng->solid_dev();
    const unsigned blocks = (unsigned)std::min<long>((g.N + 255) / 256, 8192);
    const double dz = c->dim == 3 ? c->h[2] : 1.0;
    ensure_stage(c);
    if (c->dim == 3)
        k_proj_rhs<3><<<blocks, 256, 0, c->stream>>>(u, v, w, solid, c->rhs_stage, g, m, c->h[0], c->h[1], dz, dt);
    else
        k_proj_rhs<2><<<blocks, 256, 0, c->stream>>>(u, v, w, solid, c->rhs_stage, g, m, c->h[0], c->h[1], dz, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemsetAsync(p, 0, g.N * sizeof(double), c->stream));   // std::fill(g.p, 0), pressure.cpp:17
    int rc = c->eng->solve(c->rhs_stage, p, 0, max_iter, tol, st);
    if (c->dim == 3)
        k_proj_correct<3><<<blocks, 256, 0, c->stream>>>(u, v, w, p, solid, g, m, c->h[0], c->h[1], dz, dt);
    else
        k_proj_correct<2><<<blocks, 256, 0, c->stream>>>(u, v, w, p, solid, g, m, c->h[0], c->h[1], dz, dt);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    return rc;
}

extern "C" int uaamg_project_device(uaamg_ctx* c, double* u, double* v, double* w, double* p, double dt,
                                    int max_iter, double tol, uaamg_stats* st) {
    if (!c) return UAAMG_ERR_ARG;
    if (!c->eng) return fail(c, UAAMG_ERR_ARG, "no grid set");
    if (!u || !v || !p || (c->dim == 3 && !w)) return fail(c, UAAMG_ERR_ARG, "null array");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, UAAMG_ERR_ARG, "dt must be positive");
    if (max_iter < 0) max_iter = c->o.max_iter;
    if (tol < 0) tol = c->o.tol;
    if (st) std::memset(st, 0, sizeof(*st));
    int rc = UAAMG_OK;
    CABI_TRY(c, { rc = project_impl(c, u, v, w, p, dt, max_iter, tol, st); })
    return rc;
}

extern "C" int uaamg_project(uaamg_ctx* c, double* u, double* v, double* w, double* p, double dt,
                             int max_iter, double tol, uaamg_stats* st) {
    if (!c) return UAAMG_ERR_ARG;
    if (!c->eng) return fail(c, UAAMG_ERR_ARG, "no grid set");
    if (!u || !v || !p || (c->dim == 3 && !w)) return fail(c, UAAMG_ERR_ARG, "null array");
    if (!(dt > 0.0) || !std::isfinite(dt)) return fail(c, UAAMG_ERR_ARG, "dt must be positive");
    if (max_iter < 0) max_iter = c->o.max_iter;
    if (tol < 0) tol = c->o.tol;
    if (st) std::memset(st, 0, sizeof(*st));
    int rc = UAAMG_OK;
    double* buf[4] = {nullptr, nullptr, nullptr, nullptr};
    const size_t nb[4] = {mac_n(c, 0), mac_n(c, 1), mac_n(c, 2), (size_t)padded_n(c)};
    double* host[4] = {u, v, c->dim == 3 ? w : nullptr, p};
    auto release = [&]() { for (auto& b : buf) if (b) { (void)hipFree(b); b = nullptr; } };
    try {
        for (int q = 0; q < 4; q++) {
            if (!host[q]) continue;
            HIPCHK(hipMalloc(&buf[q], nb[q] * sizeof(double)));
            if (q < 3) HIPCHK(hipMemcpyAsync(buf[q], host[q], nb[q] * sizeof(double), hipMemcpyHostToDevice, c->stream));
        }
        rc = project_impl(c, buf[0], buf[1], buf[2], buf[3], dt, max_iter, tol, st);
        for (int q = 0; q < 4; q++)
            if (host[q]) HIPCHK(hipMemcpyAsync(host[q], buf[q], nb[q] * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    } catch (const HipError& e) { release(); return from_hip(c, e); }
    catch (const std::bad_alloc&) { release(); return fail(c, UAAMG_ERR_OOM, "host allocation failed"); }
    release();
    return rc;
}

extern "C" int uaamg_solve(uaamg_ctx* c, const double* rhs_host, double* p_host, int max_iter, double tol,
                           uaamg_stats* st) {
    if (!c) return UAAMG_ERR_ARG;
    if (!c->eng) return fail(c, UAAMG_ERR_ARG, "no grid set");
    if (!rhs_host || !p_host) return fail(c, UAAMG_ERR_ARG, "null array");
    if (max_iter < 0) max_iter = c->o.max_iter;
    if (tol < 0) tol = c->o.tol;
    if (st) std::memset(st, 0, sizeof(*st));
    int rc = UAAMG_OK;
    CABI_TRY(c, {
        ensure_stage(c);
        long N = padded_n(c);
        HIPCHK(hipMemcpyAsync(c->rhs_stage, rhs_host, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->p_stage, p_host, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
        rc = c->eng->solve(c->rhs_stage, c->p_stage, 0, max_iter, tol, st);
        HIPCHK(hipMemcpyAsync(p_host, c->p_stage, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    })
    return rc;
}

extern "C" int uaamg_apply_precond(uaamg_ctx* c, const double* r_host, double* z_host) {
    if (!c) return UAAMG_ERR_ARG;
    if (!c->eng) return fail(c, UAAMG_ERR_ARG, "no grid set");
    if (!r_host || !z_host) return fail(c, UAAMG_ERR_ARG, "null array");
    CABI_TRY(c, {
        ensure_stage(c);
        long N = padded_n(c);
        HIPCHK(hipMemcpyAsync(c->rhs_stage, r_host, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
        c->eng->apply(c->rhs_stage, c->p_stage);
        std::vector<double> z(N);
        HIPCHK(hipMemcpyAsync(z.data(), c->p_stage, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        // interior cells only, as the reference apply writes them
        const long sx = c->n[0] + 2, sy = sx * (c->n[1] + 2);
        const int klo = c->dim == 3 ? 1 : 0, khi = c->dim == 3 ? c->n[2] : 0;
        for (int k = klo; k <= khi; k++)
            for (int j = 1; j <= c->n[1]; j++)
                for (int i = 1; i <= c->n[0]; i++) { long id = i + j * sx + k * sy; z_host[id] = z[id]; }
    })
    return UAAMG_OK;
}

extern "C" int uaamg_matvec(uaamg_ctx* c, const double* v_host, double* Av_host, double* vAv) {
    if (!c) return UAAMG_ERR_ARG;
    if (!c->eng) return fail(c, UAAMG_ERR_ARG, "no grid set");
    if (!v_host || !Av_host) return fail(c, UAAMG_ERR_ARG, "null array");
    CABI_TRY(c, {
        ensure_stage(c);
        long N = padded_n(c);
        HIPCHK(hipMemcpyAsync(c->rhs_stage, v_host, N * sizeof(double), hipMemcpyHostToDevice, c->stream));
        double d = c->eng->matvec(c->rhs_stage, c->p_stage);
        std::vector<double> a(N);
        HIPCHK(hipMemcpyAsync(a.data(), c->p_stage, N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const long sx = c->n[0] + 2, sy = sx * (c->n[1] + 2);
        const int klo = c->dim == 3 ? 1 : 0, khi = c->dim == 3 ? c->n[2] : 0;
        for (int k = klo; k <= khi; k++)
            for (int j = 1; j <= c->n[1]; j++)
                for (int i = 1; i <= c->n[0]; i++) { long id = i + j * sx + k * sy; Av_host[id] = a[id]; }
        if (vAv) *vAv = d;
    })
    return UAAMG_OK;
}

extern "C" int uaamg_hierarchy(const uaamg_ctx* c, int* nlevels, int* dims, int max_levels) {
    if (!c || !nlevels) return UAAMG_ERR_ARG;
    if (!c->eng) return UAAMG_ERR_ARG;
    *nlevels = c->eng->nlevels();
    for (int l = 0; l < *nlevels && l < max_levels && dims; l++) {
        Geo g = c->eng->geo(l);
        dims[3 * l] = g.nx; dims[3 * l + 1] = g.ny; dims[3 * l + 2] = g.nz;
    }
    return UAAMG_OK;
}

extern "C" int uaamg_level_data(uaamg_ctx* c, int level, uint8_t* solid, double* diag, double* cx, double* cy,
                                double* cz, double* excess) {
    if (!c) return UAAMG_ERR_ARG;
    if (!c->eng) return fail(c, UAAMG_ERR_ARG, "no grid set");
    if (level < 0 || level >= c->eng->nlevels()) return fail(c, UAAMG_ERR_ARG, "level out of range");
    CABI_TRY(c, { c->eng->level_data(level, solid, diag, cx, cy, cz, excess); })
    return UAAMG_OK;
}

extern "C" size_t uaamg_bytes_required(const uaamg_grid* g, const uaamg_opts* o) {
    std::string why;
    if (!grid_valid(g, why)) return 0;
    uaamg_opts d;
    uaamg_default_opts(&d);
    if (!o) o = &d;
    if (!opts_valid(o, why)) return 0;
    if (g->dim != 2 && (o->preconditioner == UAAMG_PRECOND_GMG || o->preconditioner == UAAMG_PRECOND_AMG)) return 0;
    auto shape = hierarchy_shape(g->dim, g->n[0], g->n[1], g->dim == 3 ? g->n[2] : 1, o->coarsest_side);
    if (shape.empty()) return 0;
    // mirrors Engine::allocate (device bytes requested by a context)
    const bool v64 = o->precision == UAAMG_FP64 || o->precision == UAAMG_MIXED_F64;
    const size_t sv = v64 ? 8 : 4;
    // coarse-level vectors (b, x, excess) and stencil records: 16-bit only in
    // 3D (2D hierarchies keep FP32 levels under the 16-bit forms)
    const bool d3 = g->dim == 3;
    const size_t sl = (o->precision == UAAMG_FP64) ? 8 : ((o->precision == UAAMG_MIXED_BF16 && d3) ? 2 : 4);
    const size_t sr = (o->precision == UAAMG_FP64) ? 8
                    : (((o->precision == UAAMG_MIXED_BF16 || o->precision == UAAMG_MIXED_BF16R) && d3) ? 2 : 4);
    const Geo& g0 = shape[0];
    const size_t N0 = (size_t)g0.N, N4 = (N0 + 3) / 4 * 4;
    const size_t margin = (size_t)((g0.sx + 8 + 63) / 64) * 64;
    // level-0 vectors r, z, p, Ap, x and the second p and r buffers of the
    // fused CG kernels (3D FP32 forms with rows of at most 256 cells)
    const bool pmv = g->dim == 3 && !v64 && g0.nx <= 256 && shape.size() >= 2 &&
                     (o->forms & UAAMG_FORM_ROWS) && (o->forms & UAAMG_FORM_PMV);
    size_t bytes = N0 + N4 + kRowSlack + 8 + sizeof(uaamg::Ctrl) + (pmv ? 7 : 5) * (N4 + 2 * margin) * sv;
    if (pmv) bytes += (size_t)(8 * ((((g0.ny + 15) / 16) * ((g0.nz + 3) / 4) + 7) / 8) + 64) * 8;   // p.Ap partials
    long maxb = std::max<long>(cells_launch(g0, g->dim).nblocks(), 2048);
    for (size_t l = 1; l < shape.size(); l++) {
        bytes += (size_t)shape[l].N * (3 * sl + 4 * sr + 1) + (size_t)kRowSlack * 6 * std::max(sl, sr);
        maxb = std::max(maxb, cells_launch(shape[l], g->dim).nblocks());
    }
    // fused level-0 launch (at most) and the row kernels' workgroups
    {
        using E3 = Engine<3, float, float>;
        using E2 = Engine<2, float, float>;
        const dim3 mg = (g->dim == 3) ? E3::march_launch(g0, E3::RX0, E3::RY0, o->march_target, o->march_kc_min).grid
                                      : E2::march_launch(g0, E2::RX0, E2::RY0, o->march_target, o->march_kc_min).grid;
        maxb = std::max(maxb, (long)mg.x * mg.y * mg.z);
        if (g->dim == 3) {
            const long tiles = (long)((g0.ny + 15) / 16) * ((g0.nz + 3) / 4);
            maxb = std::max(maxb, 8 * ((tiles + 7) / 8));
        }
    }
    bytes += (3 * (size_t)maxb + 64) * 8;
    // the tail kernel's pair records (Engine::tail_setup)
    if ((o->forms & UAAMG_FORM_TAIL) && o->preconditioner == UAAMG_PRECOND_UAAMG) {
        const bool l64 = sl == 8;
        int lt;
        if (g->dim == 3) lt = l64 ? Engine<3, double, double>::choose_tail(shape) : Engine<3, float, float>::choose_tail(shape);
        else lt = l64 ? Engine<2, double, double>::choose_tail(shape) : Engine<2, float, float>::choose_tail(shape);
        if (lt > 0) {
            const int nslots = uaamg::TAIL_MAXP + (int)shape.size() - lt - 1;
            bytes += (size_t)(l64 ? uaamg::tail_rec_bytes<double>(nslots) : uaamg::tail_rec_bytes<float>(nslots));
        }
    }
    return bytes;
}

extern "C" int uaamg_amgpcg(int N, const double* b, double* x, int max_iter, double tol, int* iters, int precision) {
    if (N < 2 || !b || !x || !iters) return UAAMG_ERR_ARG;
    if (precision != UAAMG_FP64 && precision != UAAMG_FP32) return UAAMG_ERR_ARG;
    std::string why;
    int rc = check_device(why);
    if (rc) return rc;
    try {
        return precision == UAAMG_FP64 ? amgpcg_run<double>(N, b, x, max_iter, tol, iters)
                                       : amgpcg_run<float>(N, b, x, max_iter, tol, iters);
    } catch (const HipError& e) {
        std::fprintf(stderr, "uaamg_amgpcg: %s (%s)\n", hipGetErrorString(e.e), e.where.c_str());
        return e.e == hipErrorOutOfMemory ? UAAMG_ERR_OOM : UAAMG_ERR_HIP;
    }
}

// ------------------------------------------------------------------ amgpcg without allocation
//
// Every intermediate buffer of uaamg_amgpcg_noalloc lives in `scratch`, a
// fixed device array each GPU holds from the library's load until the process
// ends.  In host code `scratch` names a host-side stand-in, not the device
// memory: the device address comes from hipGetSymbolAddress, once per device
// (with that device current), and only that address is passed to kernels and
// copies.  Calls on one GPU share its array: calls that overlap in time on one
// GPU race.
namespace {
constexpr size_t kAmgScratchBytes = 32u << 20;
__device__ __align__(256) unsigned char scratch[kAmgScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

// device address of `scratch` on the current device (looked up on its first use there)
int scratch_on_device(unsigned char** out, std::string& why) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) { why = "hipGetDevice failed"; return UAAMG_ERR_HIP; }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            why = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return UAAMG_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return UAAMG_OK;
}
}  // namespace

extern "C" int uaamg_amgpcg_noalloc(int N, const double* b_dev, double* x_dev, int max_iter, double tol,
                                    int* iters, int precision, void* stream) {
    if (N < 2 || !b_dev || !x_dev || !iters) return UAAMG_ERR_ARG;
    if (precision != UAAMG_FP64 && precision != UAAMG_FP32) return UAAMG_ERR_ARG;
    const size_t need = precision == UAAMG_FP64 ? amgpcg_arena_bytes<double>(N) : amgpcg_arena_bytes<float>(N);
    if (need > kAmgScratchBytes) {
        std::fprintf(stderr, "uaamg_amgpcg_noalloc: N = %d needs %zu bytes of scratch, the fixed device array holds %zu "
                     "(N <= 737 in FP64, N <= 1043 in FP32); use uaamg_amgpcg\n", N, need, kAmgScratchBytes);
        return UAAMG_ERR_ADMISSION;
    }
    std::string why;
    int rc = check_device(why);
    if (rc) { std::fprintf(stderr, "uaamg_amgpcg_noalloc: %s\n", why.c_str()); return rc; }
    unsigned char* ptr = nullptr;
    rc = scratch_on_device(&ptr, why);
    if (rc) { std::fprintf(stderr, "uaamg_amgpcg_noalloc: %s\n", why.c_str()); return rc; }
    hipStream_t s = static_cast<hipStream_t>(stream);
    double hscal[8] = {0.0};   // dot products land in host stack memory (no pinned allocation)
    try {
        return precision == UAAMG_FP64
            ? amgpcg_core<double>(N, b_dev, x_dev, max_iter, tol, iters, ptr, kAmgScratchBytes, s, hscal)
            : amgpcg_core<float>(N, b_dev, x_dev, max_iter, tol, iters, ptr, kAmgScratchBytes, s, hscal);
    } catch (const HipError& e) {
        std::fprintf(stderr, "uaamg_amgpcg_noalloc: %s (%s)\n", hipGetErrorString(e.e), e.where.c_str());
        return UAAMG_ERR_HIP;
    }
}

// ------------------------------------------------------------------ batch C-ABI

struct uaamg_batch {
    uaamg_opts o{};
    std::string err;
    hipStream_t stream = nullptr;
    std::unique_ptr<uaamg::BatchBase> part[2];       // [0]: 2D members, [1]: 3D members
    std::vector<std::unique_ptr<EngineBase>> solo;   // members a batch cannot take (a one-level hierarchy)
    std::vector<std::pair<int, int>> members;        // (part 0/1, index in part) or (2, index in solo)
    uint8_t* stage = nullptr; long stage_n = 0;
};

static int bfail(uaamg_batch* b, int code, const std::string& msg) {
    if (b) b->err = msg;
    return code;
}

#define BATCH_TRY(b, ...)                                                                          \
    try { __VA_ARGS__ }                                                                            \
    catch (const HipError& e) {                                                                    \
        return bfail(b, e.e == hipErrorOutOfMemory ? UAAMG_ERR_OOM : UAAMG_ERR_HIP,               \
                     std::string(hipGetErrorString(e.e)) + " (" + e.where + ")");                 \
    }                                                                                              \
    catch (const std::bad_alloc&) { return bfail(b, UAAMG_ERR_OOM, "host allocation failed"); }

static std::unique_ptr<uaamg::BatchBase> make_batch_part(int dim, const uaamg_opts& o, hipStream_t s) {
    if (dim == 2) return std::make_unique<uaamg::BatchEngine<2, float>>(o, s);   // FP32 levels in 2D (see make_engine)
    if (o.precision == UAAMG_FP32) return std::make_unique<uaamg::BatchEngine<3, float>>(o, s);
    return std::make_unique<uaamg::BatchEngine<3, __bf16>>(o, s);
}

extern "C" int uaamg_batch_create(uaamg_batch** out, const uaamg_opts* o) {
    if (!out) return UAAMG_ERR_ARG;
    *out = nullptr;
    uaamg_opts d;
    uaamg_default_opts(&d);
    d.precision = UAAMG_FP32;
    if (!o) o = &d;
    std::string why;
    if (!opts_valid(o, why)) return UAAMG_ERR_ARG;
    if (o->precision != UAAMG_FP32 && o->precision != UAAMG_MIXED_BF16) return UAAMG_ERR_ARG;
    if (o->nu_pre != 1 || o->nu_post != 1 || !o->fused || !(o->forms & UAAMG_FORM_TAIL) ||
        o->preconditioner != UAAMG_PRECOND_UAAMG) return UAAMG_ERR_ARG;
    int rc = check_device(why);
    if (rc) { std::fprintf(stderr, "uaamg_batch_create: %s\n", why.c_str()); return rc; }
    uaamg_batch* b = new (std::nothrow) uaamg_batch();
    if (!b) return UAAMG_ERR_OOM;
    b->o = *o;
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return UAAMG_ERR_HIP; }
    *out = b;
    return UAAMG_OK;
}

extern "C" void uaamg_batch_destroy(uaamg_batch* b) {
    if (!b) return;
    b->part[0].reset();
    b->part[1].reset();
    if (b->stage) (void)hipFree(b->stage);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

extern "C" const char* uaamg_batch_last_error(const uaamg_batch* b) { return b ? b->err.c_str() : "null batch"; }

static int batch_add_impl(uaamg_batch* b, const uaamg_grid* g, int* id, bool device_mask) {
    if (!b) return UAAMG_ERR_ARG;
    std::string why;
    if (!grid_valid(g, why)) return bfail(b, UAAMG_ERR_ARG, why);
    BATCH_TRY(b, {
        const int pi = g->dim == 3 ? 1 : 0;
        if (!b->part[pi]) b->part[pi] = make_batch_part(g->dim, b->o, b->stream);
        const long N = (long)(g->n[0] + 2) * (g->n[1] + 2) * (g->dim == 3 ? (g->n[2] + 2) : 1);
        const uint8_t* src = g->solid;
        if (!device_mask) {
            if (b->stage_n < N) {
                if (b->stage) HIPCHK(hipFree(b->stage));
                b->stage = nullptr;
                HIPCHK(hipMalloc(&b->stage, N));
                b->stage_n = N;
            }
            HIPCHK(hipMemcpyAsync(b->stage, g->solid, N, hipMemcpyHostToDevice, b->stream));
            src = b->stage;
        }
        int n[3] = {g->n[0], g->n[1], g->dim == 3 ? g->n[2] : 1};
        double h[3] = {g->h[0], g->h[1], g->dim == 3 ? g->h[2] : 1.0};
        int wall[6] = {0, 0, 0, 0, 0, 0};
        for (int f = 0; f < 2 * g->dim; f++) wall[f] = g->wall[f];
        auto shape = hierarchy_shape(g->dim, n[0], n[1], n[2], b->o.coarsest_side);
        const bool batchable = shape.size() >= 2 &&
            (g->dim == 3 ? Engine<3, float, float>::choose_tail(shape) : Engine<2, float, float>::choose_tail(shape)) > 0;
        if (batchable) {
            const int local = b->part[pi]->add(n, h, wall, src);
            b->members.emplace_back(pi, local);
        } else {   // solved on its own inside uaamg_batch_solve
            b->solo.push_back(make_engine(b->o, *g, b->stream));
            b->solo.back()->set_grid(src);
            b->members.emplace_back(2, (int)b->solo.size() - 1);
        }
        HIPCHK(hipStreamSynchronize(b->stream));
        if (id) *id = (int)b->members.size() - 1;
    })
    return UAAMG_OK;
}

extern "C" int uaamg_batch_add(uaamg_batch* b, const uaamg_grid* g, int* id) { return batch_add_impl(b, g, id, false); }
extern "C" int uaamg_batch_add_device(uaamg_batch* b, const uaamg_grid* g, int* id) { return batch_add_impl(b, g, id, true); }

extern "C" int uaamg_batch_clear(uaamg_batch* b) {
    if (!b) return UAAMG_ERR_ARG;
    BATCH_TRY(b, {
        for (auto& p : b->part) if (p) p->clear();
        b->solo.clear();
        b->members.clear();
    })
    return UAAMG_OK;
}

extern "C" int uaamg_batch_size(const uaamg_batch* b) { return b ? (int)b->members.size() : 0; }

extern "C" int uaamg_batch_tuning(const uaamg_batch* b, int dim, int* march_target, int* march_kc_min, int* rows_kc) {
    if (!b || (dim != 2 && dim != 3)) return UAAMG_ERR_ARG;
    const auto& p = b->part[dim == 3 ? 1 : 0];
    if (!p) return UAAMG_ERR_ARG;
    if (march_target) *march_target = p->tuned_target();
    if (march_kc_min) *march_kc_min = p->tuned_kc_min();
    if (rows_kc) *rows_kc = p->tuned_rows_kc();
    return UAAMG_OK;
}

extern "C" size_t uaamg_batch_bytes(const uaamg_batch* b) {
    if (!b) return 0;
    size_t s = 0;
    for (auto& p : b->part) if (p) s += p->bytes();
    for (auto& e : b->solo) s += e->bytes();
    return s;
}

extern "C" int uaamg_batch_solve(uaamg_batch* b, const double* const* rhs_dev, double* const* p_dev,
                                 const int* warm_start, int max_iter, double tol, uaamg_stats* st) {
    if (!b || !rhs_dev || !p_dev) return UAAMG_ERR_ARG;
    const int n = (int)b->members.size();
    for (int i = 0; i < n; i++)
        if (!rhs_dev[i] || !p_dev[i]) return bfail(b, UAAMG_ERR_ARG, "null rhs or p pointer");
    if (max_iter < 0) max_iter = b->o.max_iter;
    if (tol < 0) tol = b->o.tol;
    BATCH_TRY(b, {
        auto t0 = std::chrono::steady_clock::now();
        std::vector<const double*> rhs[2];
        std::vector<double*> pout[2];
        std::vector<int> warm[2];
        for (int i = 0; i < n; i++) {
            const int pi = b->members[i].first;
            if (pi == 2) continue;
            rhs[pi].push_back(rhs_dev[i]);
            pout[pi].push_back(p_dev[i]);
            warm[pi].push_back(warm_start ? warm_start[i] : 0);
        }
        for (int pi = 0; pi < 2; pi++)
            if (b->part[pi] && b->part[pi]->size())
                b->part[pi]->begin(rhs[pi].data(), pout[pi].data(), warm[pi].data(), max_iter, tol);
        auto busy = [&]() {
            for (auto& p : b->part) if (p && p->size() && !p->done()) return true;
            return false;
        };
        while (busy()) {
            for (auto& p : b->part) if (p && p->size()) p->chunk();
            HIPCHK(hipStreamSynchronize(b->stream));
            for (auto& p : b->part) if (p && p->size()) p->poll();
        }
        for (auto& p : b->part) if (p && p->size()) p->finish();
        HIPCHK(hipStreamSynchronize(b->stream));
        std::vector<uaamg_stats> solo_st(b->solo.size());
        for (int i = 0; i < n; i++)
            if (b->members[i].first == 2)
                b->solo[b->members[i].second]->solve(rhs_dev[i], p_dev[i], warm_start ? warm_start[i] : 0, max_iter, tol,
                                                     &solo_st[b->members[i].second]);
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        int worst = UAAMG_OK;
        for (int i = 0; i < n; i++) {
            uaamg_stats tmp;
            if (b->members[i].first == 2) { tmp = solo_st[b->members[i].second]; tmp.ms_total = ms; }
            else b->part[b->members[i].first]->stats(b->members[i].second, &tmp, ms);
            if (st) st[i] = tmp;
            if (tmp.status != UAAMG_OK && worst == UAAMG_OK) worst = tmp.status;
        }
        return worst;
    })
    return UAAMG_OK;
}

// ------------------------------------------------------------------ serving C-ABI
//
// A request loop over batches with continuous admission: each simulation
// (registered once with its grid) keeps its hierarchy resident on the device
// while memory allows; requests (one pressure solve of one frame) queue in
// arrival order; every step admits what fits -- a simulation's next frame once
// its previous one is done, new hierarchies within the HBM budget, evicting
// the least recently used idle simulations -- starts the admitted requests,
// runs one chunk of iterations of every active request as one batch, and
// completes the converged ones.  Requests that arrive while others run join
// at the next chunk.
//
// Simulations are split into two size classes (below / from 2^18 padded
// cells), each with a stream of its own and one batch per dimension: a class
// whose chunk has finished is completed, refilled and relaunched while the
// other runs, so a 64^2 request does not wait for the chunks of 256^3
// members (a chunk of the large class takes milliseconds, one of the small
// class tens of microseconds).  Two streams besides the caller's keep the
// classes on hardware queues of their own (a process has four).

namespace {
struct SimState {
    int dim = 0, n[3] = {1, 1, 1}, wall[6] = {0, 0, 0, 0, 0, 0};
    double h[3] = {1, 1, 1};
    std::vector<uint8_t> mask;
    int part = -1, slot = -1;          // resident: batch part (size class) and member index
    size_t bytes = 0;
    long last_used = 0;
    bool busy = false;
};
struct Request {
    int sim = -1;
    const double* rhs = nullptr;
    double* p = nullptr;
    int warm = 0, max_iter = 0;
    double tol = 0;
    int state = 0;                     // 0 queued, 1 running, 2 done
    uaamg_stats st{};
    std::chrono::steady_clock::time_point t_submit, t_done;
};
}  // namespace

// batch part of a simulation: 2D small, 2D large, 3D small, 3D large; the
// size class (small 0, large 1) is the part's low bit
constexpr int kServeParts = 4;
constexpr long kServeLarge = 1L << 18;   // padded cells from which a grid is large
static int serve_part_of(int dim, long cells) { return (dim == 3 ? 2 : 0) + (cells >= kServeLarge ? 1 : 0); }
static int serve_class_of(int part) { return part & 1; }

struct uaamg_server {
    uaamg_opts o{};
    std::string err;
    size_t budget = 0, used = 0;
    int max_active = 0;
    // one batch per part; one stream and completion event per size class;
    // inflight: a chunk of the class was launched and its event has not been
    // seen complete
    std::unique_ptr<uaamg::BatchBase> part[kServeParts];
    hipStream_t cstream[2] = {};
    hipEvent_t cev[2] = {};
    bool inflight[2] = {};
    std::vector<int> slot_sim[kServeParts];
    std::unordered_map<int, SimState> sims;
    std::vector<Request> reqs;
    std::deque<int> queue;
    std::vector<int> active;
    std::deque<int> undelivered;       // finished requests not yet returned by a step (done[] was full)
    long tick = 0;
    uint8_t* stage = nullptr; long stage_n = 0;
};

static int sfail(uaamg_server* s, int code, const std::string& msg) {
    if (s) s->err = msg;
    return code;
}

#define SERVER_TRY(s, ...)                                                                         \
    try { __VA_ARGS__ }                                                                            \
    catch (const HipError& e) {                                                                    \
        return sfail(s, e.e == hipErrorOutOfMemory ? UAAMG_ERR_OOM : UAAMG_ERR_HIP,               \
                     std::string(hipGetErrorString(e.e)) + " (" + e.where + ")");                 \
    }                                                                                              \
    catch (const std::bad_alloc&) { return sfail(s, UAAMG_ERR_OOM, "host allocation failed"); }

extern "C" int uaamg_server_create(uaamg_server** out, const uaamg_opts* o, size_t hbm_budget, int max_active) {
    if (!out || max_active < 1) return UAAMG_ERR_ARG;
    *out = nullptr;
    uaamg_opts d;
    uaamg_default_opts(&d);
    d.precision = UAAMG_FP32;
    if (!o) o = &d;
    std::string why;
    if (!opts_valid(o, why)) return UAAMG_ERR_ARG;
    if (o->precision != UAAMG_FP32 && o->precision != UAAMG_MIXED_BF16) return UAAMG_ERR_ARG;
    if (o->nu_pre != 1 || o->nu_post != 1 || !o->fused || !(o->forms & UAAMG_FORM_TAIL) ||
        o->preconditioner != UAAMG_PRECOND_UAAMG) return UAAMG_ERR_ARG;
    int rc = check_device(why);
    if (rc) { std::fprintf(stderr, "uaamg_server_create: %s\n", why.c_str()); return rc; }
    uaamg_server* s = new (std::nothrow) uaamg_server();
    if (!s) return UAAMG_ERR_OOM;
    s->o = *o;
    s->budget = hbm_budget;
    s->max_active = max_active;
    // the class streams are created together (default priority), so that
    // they land on different hardware queues
    for (int c = 0; c < 2; c++)
        if (hipStreamCreateWithFlags(&s->cstream[c], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&s->cev[c], hipEventDisableTiming) != hipSuccess) {
            uaamg_server_destroy(s);
            return UAAMG_ERR_HIP;
        }
    *out = s;
    return UAAMG_OK;
}

extern "C" void uaamg_server_destroy(uaamg_server* s) {
    if (!s) return;
    for (int c = 0; c < 2; c++)
        if (s->cstream[c]) (void)hipStreamSynchronize(s->cstream[c]);
    for (int pi = 0; pi < kServeParts; pi++) s->part[pi].reset();
    if (s->stage) (void)hipFree(s->stage);
    for (int c = 0; c < 2; c++) {
        if (s->cev[c]) (void)hipEventDestroy(s->cev[c]);
        if (s->cstream[c]) (void)hipStreamDestroy(s->cstream[c]);
    }
    delete s;
}

extern "C" const char* uaamg_server_last_error(const uaamg_server* s) { return s ? s->err.c_str() : "null server"; }

extern "C" int uaamg_server_register(uaamg_server* s, int sim, const uaamg_grid* g) {
    if (!s) return UAAMG_ERR_ARG;
    std::string why;
    if (!grid_valid(g, why)) return sfail(s, UAAMG_ERR_ARG, why);
    if (s->sims.count(sim)) return sfail(s, UAAMG_ERR_ARG, "simulation already registered");
    {
        auto shape = hierarchy_shape(g->dim, g->n[0], g->n[1], g->dim == 3 ? g->n[2] : 1, s->o.coarsest_side);
        const bool ok = shape.size() >= 2 &&
            (g->dim == 3 ? Engine<3, float, float>::choose_tail(shape) : Engine<2, float, float>::choose_tail(shape)) > 0;
        if (!ok) return sfail(s, UAAMG_ERR_ARG, "grid too small or too flat for the serving batches (needs 2 levels)");
    }
    SimState st;
    st.dim = g->dim;
    for (int q = 0; q < 3; q++) { st.n[q] = q < g->dim ? g->n[q] : 1; st.h[q] = q < g->dim ? g->h[q] : 1.0; }
    for (int f = 0; f < 6; f++) st.wall[f] = f < 2 * g->dim ? g->wall[f] : 0;
    const long N = (long)(st.n[0] + 2) * (st.n[1] + 2) * (g->dim == 3 ? st.n[2] + 2 : 1);
    st.mask.assign(g->solid, g->solid + N);
    s->sims.emplace(sim, std::move(st));
    return UAAMG_OK;
}

extern "C" int uaamg_server_submit(uaamg_server* s, int sim, const double* rhs_dev, double* p_dev, int warm_start,
                                   int max_iter, double tol, int* req) {
    if (!s || !rhs_dev || !p_dev) return UAAMG_ERR_ARG;
    if (!s->sims.count(sim)) return sfail(s, UAAMG_ERR_ARG, "unknown simulation");
    Request r;
    r.sim = sim; r.rhs = rhs_dev; r.p = p_dev; r.warm = warm_start;
    r.max_iter = max_iter < 0 ? s->o.max_iter : max_iter;
    r.tol = tol < 0 ? s->o.tol : tol;
    r.t_submit = std::chrono::steady_clock::now();
    s->reqs.push_back(r);
    s->queue.push_back((int)s->reqs.size() - 1);
    if (req) *req = (int)s->reqs.size() - 1;
    return UAAMG_OK;
}

static size_t sim_estimate(const uaamg_server* s, const SimState& st) {
    uaamg_grid g{};
    g.dim = st.dim;
    for (int q = 0; q < 3; q++) { g.n[q] = st.n[q]; g.h[q] = st.h[q]; }
    g.solid = st.mask.data();
    for (int f = 0; f < 6; f++) g.wall[f] = st.wall[f];
    uaamg_opts mo = s->o;
    mo.forms &= ~UAAMG_FORM_PMV;   // as a batch member is built (BatchEngine::add)
    return uaamg_bytes_required(&g, &mo);
}

// drop the least recently used idle resident simulation of a class with no
// chunk in flight; false if none
static bool server_evict_one(uaamg_server* s) {
    int victim = -1;
    long best = 0;
    for (auto& kv : s->sims) {
        const SimState& st = kv.second;
        if (st.slot < 0 || st.busy || s->inflight[serve_class_of(st.part)]) continue;
        if (victim < 0 || st.last_used < best) { victim = kv.first; best = st.last_used; }
    }
    if (victim < 0) return false;
    SimState& v = s->sims[victim];
    const int pi = v.part, slot = v.slot;
    s->part[pi]->remove(slot);
    s->slot_sim[pi].erase(s->slot_sim[pi].begin() + slot);
    for (int k = slot; k < (int)s->slot_sim[pi].size(); k++) s->sims[s->slot_sim[pi][k]].slot = k;
    s->used -= std::min(s->used, v.bytes);
    v.slot = -1; v.part = -1; v.bytes = 0;
    return true;
}

// completions of the classes whose chunk has finished; with wait, blocks until
// at least one in-flight class has finished (if any is in flight).  Returns
// whether a class finished.
static bool server_collect(uaamg_server* s, bool wait, std::vector<int>& fin) {
    bool any = false;
    for (;;) {
        bool inflight = false;
        for (int cl = 0; cl < 2; cl++) {
            if (!s->inflight[cl]) continue;
            inflight = true;
            const hipError_t q = hipEventQuery(s->cev[cl]);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) throw HipError{q, "hipEventQuery"};
            s->inflight[cl] = false;
            any = true;
            std::vector<int> still, mine;
            for (int id : s->active) {
                Request& r = s->reqs[id];
                SimState& st = s->sims[r.sim];
                if (serve_class_of(st.part) != cl) { still.push_back(id); continue; }
                const uaamg::Ctrl& c = s->part[st.part]->member_ctrl(st.slot);
                if (!c.done) { still.push_back(id); continue; }
                s->part[st.part]->finish_member(st.slot, r.p, r.warm);
                r.st.iters = c.iter;
                r.st.status = c.status == uaamg::ST_CONVERGED ? UAAMG_OK
                            : c.status == uaamg::ST_BREAKDOWN ? UAAMG_BREAKDOWN
                            : c.status == uaamg::ST_NONFINITE ? UAAMG_NONFINITE : UAAMG_NOT_CONVERGED;
                r.st.rnorm = std::sqrt(std::max(0.0, c.rsnew));
                r.st.r0norm = std::sqrt(std::max(0.0, c.r0sq));
                r.state = 2;
                st.busy = false;
                st.last_used = ++s->tick;
                mine.push_back(id);
            }
            s->active.swap(still);
            if (!mine.empty()) HIPCHK(hipStreamSynchronize(s->cstream[cl]));   // p written
            const auto now = std::chrono::steady_clock::now();
            for (int id : mine) {   // latency taken when the result is in p
                Request& r = s->reqs[id];
                r.t_done = now;
                r.st.ms_total = std::chrono::duration<double, std::milli>(r.t_done - r.t_submit).count();
                fin.push_back(id);
            }
        }
        if (any || !wait || !inflight) return any;
        std::this_thread::yield();
    }
}

// admission, in arrival order, into classes with no chunk in flight; new
// members join their batch before any starts; then one chunk of every idle
// class with active members is launched on that class's stream
static void server_admit_and_launch(uaamg_server* s) {
    std::vector<int> starting;
    for (auto it = s->queue.begin(); it != s->queue.end() && (int)(s->active.size() + starting.size()) < s->max_active;) {
        Request& r = s->reqs[*it];
        SimState& st = s->sims[r.sim];
        if (st.busy) { ++it; continue; }              // its previous frame is still running
        const int pi = st.slot >= 0 ? st.part : serve_part_of(st.dim, (long)st.mask.size());
        const int cl = serve_class_of(pi);
        if (s->inflight[cl]) { ++it; continue; }      // its class is running a chunk
        if (st.slot < 0) {
            const size_t need = sim_estimate(s, st);
            while (s->used + need > s->budget && server_evict_one(s)) {}
            if (s->used + need > s->budget) { ++it; continue; }   // waits for memory
            if (!s->part[pi]) s->part[pi] = make_batch_part(st.dim, s->o, s->cstream[cl]);
            const long N = (long)st.mask.size();
            if (s->stage_n < N) {
                if (s->stage) HIPCHK(hipFree(s->stage));
                s->stage = nullptr;
                HIPCHK(hipMalloc(&s->stage, N));
                s->stage_n = N;
            }
            HIPCHK(hipMemcpyAsync(s->stage, st.mask.data(), N, hipMemcpyHostToDevice, s->cstream[cl]));
            st.slot = s->part[pi]->add(st.n, st.h, st.wall, s->stage);
            HIPCHK(hipStreamSynchronize(s->cstream[cl]));   // the staged mask is read before its next use
            st.part = pi;
            s->slot_sim[pi].push_back(r.sim);
            st.bytes = s->part[pi]->member_bytes(st.slot);
            s->used += st.bytes;
        }
        st.busy = true;
        st.last_used = ++s->tick;
        r.state = 1;
        starting.push_back(*it);
        it = s->queue.erase(it);
    }
    for (int pi = 0; pi < kServeParts; pi++)
        if (s->part[pi] && !s->inflight[serve_class_of(pi)]) s->part[pi]->prepare();
    for (int id : starting) {
        Request& r = s->reqs[id];
        SimState& st = s->sims[r.sim];
        s->part[st.part]->start_member(st.slot, r.rhs, r.p, r.warm, r.max_iter, r.tol);
        s->active.push_back(id);
    }
    bool run[kServeParts] = {};
    for (int id : s->active) run[s->sims[s->reqs[id].sim].part] = true;
    for (int cl = 0; cl < 2; cl++) {
        if (s->inflight[cl]) continue;
        bool any = false;
        for (int pi = cl; pi < kServeParts; pi += 2)
            if (run[pi]) { s->part[pi]->run_chunk(); any = true; }
        if (!any) continue;
        HIPCHK(hipEventRecord(s->cev[cl], s->cstream[cl]));
        s->inflight[cl] = true;
    }
}

extern "C" int uaamg_server_step(uaamg_server* s, int* done, int max_done, int* n_done) {
    if (!s) return UAAMG_ERR_ARG;
    if (n_done) *n_done = 0;
    SERVER_TRY(s, {
        std::vector<int> fin;
        // finished classes first (no wait), then admissions and launches;
        // if no class had finished, wait for one and refill it at once
        const bool got = server_collect(s, false, fin);
        server_admit_and_launch(s);
        if (!got && server_collect(s, true, fin)) server_admit_and_launch(s);
        // finished requests a previous step could not return (its done[] was full) come first
        int nd = 0;
        while (!s->undelivered.empty() && nd < max_done && done) { done[nd++] = s->undelivered.front(); s->undelivered.pop_front(); }
        for (int id : fin) {
            if (done && nd < max_done) done[nd++] = id;
            else s->undelivered.push_back(id);
        }
        if (n_done) *n_done = nd;
    })
    return UAAMG_OK;
}

extern "C" int uaamg_server_request(const uaamg_server* s, int req, uaamg_stats* st) {
    if (!s || req < 0 || req >= (int)s->reqs.size() || !st) return UAAMG_ERR_ARG;
    const Request& r = s->reqs[req];
    if (r.state != 2) return UAAMG_NOT_CONVERGED;
    *st = r.st;
    return UAAMG_OK;
}

extern "C" int uaamg_server_status(const uaamg_server* s, int* queued, int* active, int* resident, size_t* bytes_used) {
    if (!s) return UAAMG_ERR_ARG;
    if (queued) *queued = (int)s->queue.size();
    if (active) *active = (int)s->active.size();
    if (resident) {
        int n = 0;
        for (auto& kv : s->sims) n += kv.second.slot >= 0;
        *resident = n;
    }
    if (bytes_used) *bytes_used = s->used;
    return UAAMG_OK;
}
