// amgpcg.hpp -- the standalone Ax=b amgpcg entry point on the GPU.
//
// Semantics of the reference's standalone amgpcg (test/solver/test_amgpcg.cpp:6-70
// with test/test_common.h:7-97): N x N node-centred Dirichlet Poisson
// (h = 1/(N+1), ghost = 0), row-major v[i*(N+2)+j]; levels N, N/2, ... down to
// <= 3; V-cycle: 2 RBGS sweeps, residual, restrict 0.25*sum(2x2) into a coarse
// grid of size N/2 made even (read back with that grid's own row stride, so an
// odd level gets a zero last row/column), recurse, prolong x1.0, 2 RBGS sweeps;
// 20 sweeps on the coarsest.  RBGS: first (i+j) odd, then (i+j) even,
// x += h^2/4 (b - A x).  PCG with absolute tolerance, out_iters = k+1.
#pragma once
#include <functional>
#include "engine.hpp"

namespace uaamg {

template <typename T>
__global__ void k_cast_all_ag(const double* __restrict__ src, T* __restrict__ dst, long n, int N) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    long i = t / (N + 2), j = t % (N + 2);
    dst[t] = (i >= 1 && i <= N && j >= 1 && j <= N) ? (T)src[t] : T(0);
}
template <typename T>
__global__ void k_cast_all_ag_out(const T* __restrict__ src, double* __restrict__ dst, long n) {
    long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) dst[t] = (double)src[t];
}

#define AG_LOOP(N)                                                          \
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;                \
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;                \
    if (i > (N) || j > (N)) return;                                         \
    const long id = (long)i * ((N) + 2) + j;

template <typename T>
__device__ __forceinline__ T ag_Ax(const T* x, long id, long s, T inv_h2) {
    return (T(4) * x[id] - x[id + s] - x[id - s] - x[id + 1] - x[id - 1]) * inv_h2;
}

template <typename T>
__global__ void k_ag_rbgs(T* __restrict__ x, const T* __restrict__ b, int N, T inv_h2, T inv_diag, int odd) {
    AG_LOOP(N)
    if (((i + j) & 1) != odd) return;
    long s = N + 2;
    x[id] = x[id] + inv_diag * (b[id] - ag_Ax(x, id, s, inv_h2));
}

// b_c(ic,jc) = 0.25 * sum of the 2x2 fine residuals for ic,jc <= Nc, else 0.
template <typename T>
__global__ void k_ag_restrict(const T* __restrict__ x, const T* __restrict__ b, int N, T inv_h2,
                              T* __restrict__ bc, int Nb, int Nc) {
    const int jc = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int ic = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (ic > Nb || jc > Nb) return;
    T v = T(0);
    if (ic <= Nc && jc <= Nc) {
        long s = N + 2;
        int i_f = 2 * ic - 1, j_f = 2 * jc - 1;
        long a = (long)i_f * s + j_f;
        T r00 = b[a] - ag_Ax(x, a, s, inv_h2);
        T r10 = b[a + s] - ag_Ax(x, a + s, s, inv_h2);
        T r01 = b[a + 1] - ag_Ax(x, a + 1, s, inv_h2);
        T r11 = b[a + s + 1] - ag_Ax(x, a + s + 1, s, inv_h2);
        v = T(0.25) * (r00 + r10 + r01 + r11);
    }
    bc[(long)ic * (Nb + 2) + jc] = v;
}

template <typename T>
__global__ void k_ag_prolong(T* __restrict__ x, int N, const T* __restrict__ xc, int Nb) {
    AG_LOOP(N)
    int ic = (i + 1) >> 1, jc = (j + 1) >> 1;
    T c = (ic <= Nb && jc <= Nb) ? xc[(long)ic * (Nb + 2) + jc] * T(1) : T(0);
    x[id] = x[id] + c;
}

template <typename T>
__global__ void k_ag_matvec(const T* __restrict__ p, T* __restrict__ Ap, int N, T inv_h2) {
    AG_LOOP(N)
    Ap[id] = ag_Ax(p, id, N + 2, inv_h2);
}

// sum over interior of a*b into one FP64 value (deterministic last-block reduction)
template <typename T>
__global__ void k_ag_dot(const T* __restrict__ a, const T* __restrict__ b, int N, double* partials) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int i = blockIdx.y * blockDim.y + threadIdx.y + 1;
    double v[1] = {0.0};
    if (i <= N && j <= N) { long id = (long)i * (N + 2) + j; v[0] = (double)a[id] * (double)b[id]; }
    block_store<1>(v, partials);
}
__global__ void k_ag_dot_fin(const double* partials, int nb, double* out) {
    double tot[1];
    reduce_partials<1>(partials, nb, tot);
    if (threadIdx.x == 0) *out = tot[0];
}

template <typename T>
__global__ void k_ag_axpy(T* __restrict__ y, const T* __restrict__ x, int N, double a) {
    AG_LOOP(N)
    y[id] = (T)((double)y[id] + a * (double)x[id]);
}

template <typename T>
__global__ void k_ag_pupd(T* __restrict__ p, const T* __restrict__ z, int N, double beta) {
    AG_LOOP(N)
    p[id] = (T)((double)z[id] + beta * (double)p[id]);
}

// Level sizes of the amgpcg hierarchy: N, N/2, ... down to <= 3 (test_amgpcg.cpp:66-71)
inline std::vector<int> amgpcg_levels(int N) {
    std::vector<int> lv;
    for (int n = N; n >= 2;) { lv.push_back(n); if (n <= 3) break; n = n / 2; if (n < 2) n = 2; }
    return lv;
}

// Device buffers of one amgpcg solve, carved in order from one arena (each
// start 256-byte aligned): per level x and b, then x, r, z, p, Ap, the dot
// partials and the scalar slot.
constexpr size_t kAgAlign = 256;
inline size_t ag_round(size_t b) { return (b + kAgAlign - 1) / kAgAlign * kAgAlign; }
inline dim3 ag_grid(int n) { return dim3((n + 63) / 64, (n + 3) / 4); }

template <typename T>
size_t amgpcg_arena_bytes(int N) {
    size_t bytes = 0;
    for (int n : amgpcg_levels(N)) bytes += 2 * ag_round((size_t)(n + 2) * (n + 2) * sizeof(T));
    const size_t NN = (size_t)(N + 2) * (N + 2);
    bytes += 5 * ag_round(NN * sizeof(T));
    const long maxb = (long)ag_grid(N).x * ag_grid(N).y;
    bytes += ag_round((maxb + 64) * sizeof(double)) + ag_round(64);
    return bytes;
}

// The solve on device arrays: b_dev (FP64, (N+2)^2 row-major, interior read)
// and x_dev (FP64, (N+2)^2, written whole).  `arena` holds at least
// amgpcg_arena_bytes<T>(N) bytes; it is cleared here.  The host reads each dot
// product (hscal: a host double, pinned or not) as the reference loop does.
template <typename T>
int amgpcg_core(int N, const double* b_dev, double* x_dev, int max_iter, double tol, int* iters,
                unsigned char* arena, size_t arena_bytes, hipStream_t s, double* hscal) {
    const std::vector<int> lv = amgpcg_levels(N);
    const int nl = (int)lv.size();
    const size_t need = amgpcg_arena_bytes<T>(N);
    if (need > arena_bytes) throw HipError{hipErrorInvalidValue, "amgpcg arena too small"};
    HIPCHK(hipMemsetAsync(arena, 0, need, s));
    size_t used = 0;
    auto take = [&](size_t bytes) {
        void* p = arena + used;
        used += ag_round(bytes);
        return p;
    };
    const size_t NN = (size_t)(N + 2) * (N + 2);
    std::vector<T*> X(nl), B(nl);
    for (int l = 0; l < nl; l++) {
        size_t n2 = (size_t)(lv[l] + 2) * (lv[l] + 2);
        X[l] = (T*)take(n2 * sizeof(T)); B[l] = (T*)take(n2 * sizeof(T));
    }
    T *x = (T*)take(NN * sizeof(T)), *r = (T*)take(NN * sizeof(T)), *z = (T*)take(NN * sizeof(T));
    T *p = (T*)take(NN * sizeof(T)), *Ap = (T*)take(NN * sizeof(T));
    dim3 blk(64, 4);
    auto grd = ag_grid;
    long maxb = (long)grd(N).x * grd(N).y;
    double* partials = (double*)take((maxb + 64) * sizeof(double));
    double* scal = (double*)take(64);

    k_cast_all_ag<T><<<(unsigned)((NN + 255) / 256), 256, 0, s>>>(b_dev, r, (long)NN, N);

    auto h_of = [](int n) { return 1.0 / (n + 1); };
    auto rbgs = [&](int l, int sweeps) {
        int n = lv[l];
        double h = h_of(n);
        T ih2 = (T)(1.0 / (h * h)), idg = (T)(h * h / 4.0);
        for (int sw = 0; sw < sweeps; sw++) {
            k_ag_rbgs<T><<<grd(n), blk, 0, s>>>(X[l], B[l], n, ih2, idg, 1);
            k_ag_rbgs<T><<<grd(n), blk, 0, s>>>(X[l], B[l], n, ih2, idg, 0);
        }
    };
    std::function<void(int)> vc = [&](int l) {
        if (l == nl - 1) { rbgs(l, 20); return; }
        rbgs(l, 2);
        int n = lv[l], nb = lv[l + 1];
        int nc = n / 2; if (nc < 2) nc = 2; if (nc % 2 != 0 && nc > 2) nc--;
        double h = h_of(n);
        k_ag_restrict<T><<<grd(nb), blk, 0, s>>>(X[l], B[l], n, (T)(1.0 / (h * h)), B[l + 1], nb, nc);
        HIPCHK(hipMemsetAsync(X[l + 1], 0, (size_t)(nb + 2) * (nb + 2) * sizeof(T), s));
        vc(l + 1);
        k_ag_prolong<T><<<grd(n), blk, 0, s>>>(X[l], n, X[l + 1], nb);
        rbgs(l, 2);
    };
    auto precond = [&]() {
        HIPCHK(hipMemcpyAsync(B[0], r, NN * sizeof(T), hipMemcpyDeviceToDevice, s));
        for (int l = 0; l < nl; l++)
            HIPCHK(hipMemsetAsync(X[l], 0, (size_t)(lv[l] + 2) * (lv[l] + 2) * sizeof(T), s));
        vc(0);
        HIPCHK(hipMemcpyAsync(z, X[0], NN * sizeof(T), hipMemcpyDeviceToDevice, s));
    };
    auto dot = [&](const T* a, const T* bb) {
        k_ag_dot<T><<<grd(N), blk, 0, s>>>(a, bb, N, partials);
        k_ag_dot_fin<<<1, 1024, 0, s>>>(partials, (int)(grd(N).x * grd(N).y), scal);
        HIPCHK(hipMemcpyAsync(hscal, scal, sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(hipStreamSynchronize(s));
        return hscal[0];
    };
    const double h0 = h_of(N);
    const T ih2_0 = (T)(1.0 / (h0 * h0));
    precond();
    HIPCHK(hipMemcpyAsync(p, z, NN * sizeof(T), hipMemcpyDeviceToDevice, s));
    double rsold = dot(r, z);
    int it_out = max_iter;
    for (int k = 0; k < max_iter; k++) {
        k_ag_matvec<T><<<grd(N), blk, 0, s>>>(p, Ap, N, ih2_0);
        double alpha = rsold / dot(p, Ap);
        k_ag_axpy<T><<<grd(N), blk, 0, s>>>(x, p, N, alpha);
        k_ag_axpy<T><<<grd(N), blk, 0, s>>>(r, Ap, N, -alpha);
        double rnorm = std::sqrt(dot(r, r));
        if (rnorm < tol) { it_out = k + 1; break; }
        precond();
        double rsnew = dot(r, z);
        double beta = rsnew / rsold;
        k_ag_pupd<T><<<grd(N), blk, 0, s>>>(p, z, N, beta);
        rsold = rsnew;
    }
    HIPCHK(hipGetLastError());
    k_cast_all_ag_out<T><<<(unsigned)((NN + 255) / 256), 256, 0, s>>>(x, x_dev, (long)NN);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s));
    *iters = it_out;
    return UAAMG_OK;
}

// Host arrays in and out: the arena and a b/x staging array are allocated
// per call on a stream of its own.
template <typename T>
int amgpcg_run(int N, const double* b_host, double* x_host, int max_iter, double tol, int* iters) {
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    std::vector<void*> allocs;
    auto dal = [&](size_t bytes) {
        void* p = nullptr;
        hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) throw HipError{e, "hipMalloc"};
        allocs.push_back(p);
        return p;
    };
    struct Cleanup {
        std::vector<void*>& a; hipStream_t s;
        ~Cleanup() { (void)hipStreamSynchronize(s); for (void* p : a) (void)hipFree(p); (void)hipStreamDestroy(s); }
    } cleanup{allocs, s};
    const size_t NN = (size_t)(N + 2) * (N + 2);
    const size_t ab = amgpcg_arena_bytes<T>(N);
    unsigned char* arena = (unsigned char*)dal(ab);
    double* bd = (double*)dal(NN * sizeof(double));
    double* hscal = nullptr;
    HIPCHK(hipHostMalloc((void**)&hscal, 64, hipHostMallocDefault));
    struct HostFree { double* p; ~HostFree() { if (p) (void)hipHostFree(p); } } hf{hscal};
    HIPCHK(hipMemcpyAsync(bd, b_host, NN * sizeof(double), hipMemcpyHostToDevice, s));
    // bd holds b until the core has cast it into r, then receives x
    amgpcg_core<T>(N, bd, bd, max_iter, tol, iters, arena, ab, s, hscal);
    HIPCHK(hipMemcpyAsync(x_host, bd, NN * sizeof(double), hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return UAAMG_OK;
}

}  // namespace uaamg
