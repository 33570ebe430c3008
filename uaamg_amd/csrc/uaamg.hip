// uaamg.hip -- C-ABI (include/uaamg.h) over the gfx950 engines.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include "../../include/uaamg.h"
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

template <int D, typename T, typename S>
void Engine<D, T, S>::cast_in_all(const double* src, T* dst) {
    long n = g0_.N;
    k_cast_all<double, T><<<(unsigned)std::min<long>((n + 255) / 256, 4096), 256, 0, stream_>>>(src, dst, n);
    HIPCHK(hipGetLastError());
}
template <int D, typename T, typename S>
void Engine<D, T, S>::cast_out_all(const T* src, double* dst) {
    long n = g0_.N;
    k_cast_all<T, double><<<(unsigned)std::min<long>((n + 255) / 256, 4096), 256, 0, stream_>>>(src, dst, n);
    HIPCHK(hipGetLastError());
}

template <int D, typename T, typename S>
void Engine<D, T, S>::allocate() {
    const int nl = nlevels();
    lv_.resize(nl);
    lv_[0].g = g0_;
    lv_[0].lc = lc0_;
    solid0_ = dalloc<uint8_t>(g0_.N);
    nbr0_ = dalloc<uint8_t>(N4_);
    count_ = dalloc<double>(1);
    ctrl_ = dalloc<Ctrl>(1);
    HIPCHK(hipHostMalloc((void**)&ctrl_host_, sizeof(Ctrl), hipHostMallocDefault));
    r_ = dalloc_margin<T>(N4_); z_ = dalloc_margin<T>(N4_); p_ = dalloc_margin<T>(N4_); Ap_ = dalloc_margin<T>(N4_);
    if (!x_is_p()) x_ = dalloc_margin<T>(N4_);
    long maxb = std::max<long>(lc0_.nblocks(), 2048);
    for (int l = 1; l < nl; l++) {
        Level& L = lv_[l];
        L.g = shape_[l];
        L.lc = cells_launch(L.g, D);
        L.b = dalloc<S>(L.g.N); L.x = dalloc<S>(L.g.N);
        L.mask = dalloc<uint8_t>(L.g.N);
        L.coef = dalloc<C4>(L.g.N);
        L.ex = dalloc<S>(L.g.N);
        maxb = std::max(maxb, L.lc.nblocks());
    }
    maxb = std::max(maxb, fused_max_blocks());
    maxb = std::max(maxb, (long)rows_blocks_);
    partials_ = dalloc<double>(3 * maxb + 64);
    fused_ok_ = fused_supported();
    HIPCHK(hipStreamSynchronize(stream_));
}

template <int D, typename T, typename S>
void Engine<D, T, S>::release() {
    (void)hipStreamSynchronize(stream_);
    for (void* p : allocs_) (void)hipFree(p);
    allocs_.clear();
    if (ctrl_host_) (void)hipHostFree(ctrl_host_);
    ctrl_host_ = nullptr;
    if (hist_) (void)hipFree(hist_);
    hist_ = nullptr;
}

template <int D, typename T, typename S>
void Engine<D, T, S>::level_data(int level, uint8_t* solid, double* diag, double* cx, double* cy, double* cz, double* ex) {
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
    return UAAMG_OK;
}

static bool opts_valid(const uaamg_opts* o, std::string& why) {
    if (o->precision < UAAMG_FP64 || o->precision > UAAMG_MIXED_FP16) { why = "precision"; return false; }
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
    if (!opts_valid(o, why)) return UAAMG_ERR_ARG;
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
    bool rebuild = (o->precision != c->o.precision) || (o->coarsest_side != c->o.coarsest_side);
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
            case UAAMG_MIXED_BF16: return std::make_unique<Engine<2, float, __bf16>>(o, n, h, g.wall, s);
            default: throw HipError{hipErrorInvalidValue, "FP16 level storage is not built (MIXED_BF16 is the 16-bit form)"};
        }
    }
    switch (o.precision) {
        case UAAMG_FP64: return std::make_unique<Engine<3, double, double>>(o, n, h, g.wall, s);
        case UAAMG_FP32: return std::make_unique<Engine<3, float, float>>(o, n, h, g.wall, s);
        case UAAMG_MIXED_BF16: return std::make_unique<Engine<3, float, __bf16>>(o, n, h, g.wall, s);
        default: throw HipError{hipErrorInvalidValue, "FP16 level storage is not built (MIXED_BF16 is the 16-bit form)"};
    }
}

static int set_grid_impl(uaamg_ctx* c, const uaamg_grid* g, bool device_mask) {
    if (!c) return UAAMG_ERR_ARG;
    std::string why;
    if (!grid_valid(g, why)) return fail(c, UAAMG_ERR_ARG, why);
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
    auto shape = hierarchy_shape(g->dim, g->n[0], g->n[1], g->dim == 3 ? g->n[2] : 1, o->coarsest_side);
    if (shape.empty()) return 0;
    size_t sv = (o->precision == UAAMG_FP64) ? 8 : 4;
    size_t sl = (o->precision == UAAMG_FP64) ? 8 : (o->precision == UAAMG_FP32 ? 4 : 2);
    size_t N0 = (size_t)shape[0].N;
    size_t bytes = N0 * (2 + 4 * sv + (o->precision == UAAMG_FP64 ? 0 : sv));
    for (size_t l = 1; l < shape.size(); l++) bytes += (size_t)shape[l].N * (2 * sl + 1 + 4 * sl);
    Launch lc = cells_launch(shape[0], g->dim);
    bytes += (3 * (size_t)lc.nblocks() + 64) * 8 + 4096;
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

// ------------------------------------------------------------------ batch C-ABI

struct uaamg_batch {
    uaamg_opts o{};
    std::string err;
    hipStream_t stream = nullptr;
    std::unique_ptr<uaamg::BatchBase> part[2];       // [0]: 2D members, [1]: 3D members
    std::vector<std::pair<int, int>> members;        // (part, index in part) in order of addition
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
    if (dim == 2) {
        if (o.precision == UAAMG_FP32) return std::make_unique<uaamg::BatchEngine<2, float>>(o, s);
        return std::make_unique<uaamg::BatchEngine<2, __bf16>>(o, s);
    }
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
    if (o->nu_pre != 1 || o->nu_post != 1 || !o->fused) return UAAMG_ERR_ARG;
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
        const int local = b->part[pi]->add(n, h, wall, src);
        HIPCHK(hipStreamSynchronize(b->stream));
        b->members.emplace_back(pi, local);
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
        b->members.clear();
    })
    return UAAMG_OK;
}

extern "C" int uaamg_batch_size(const uaamg_batch* b) { return b ? (int)b->members.size() : 0; }

extern "C" size_t uaamg_batch_bytes(const uaamg_batch* b) {
    if (!b) return 0;
    size_t s = 0;
    for (auto& p : b->part) if (p) s += p->bytes();
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
        const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        int worst = UAAMG_OK;
        for (int i = 0; i < n; i++) {
            uaamg_stats tmp;
            b->part[b->members[i].first]->stats(b->members[i].second, &tmp, ms);
            if (st) st[i] = tmp;
            if (tmp.status != UAAMG_OK && worst == UAAMG_OK) worst = tmp.status;
        }
        return worst;
    })
    return UAAMG_OK;
}
