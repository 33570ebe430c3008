// common.hpp -- shared device-side types for the gfx950 UAAMG-PCG kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <type_traits>

namespace uaamg {

constexpr int WAVE = 64;

// Geometry of one level: interior n, padded strides (i fastest).
// 2D levels use nz = 1 and k = 0 (sy is then unused).
struct Geo {
    int nx, ny, nz;
    long sx, sy;   // sx = nx+2, sy = (nx+2)*(ny+2)
    long N;        // padded cell count
};

inline Geo make_geo(int dim, int nx, int ny, int nz) {
    Geo g;
    g.nx = nx; g.ny = ny; g.nz = (dim == 3) ? nz : 1;
    g.sx = nx + 2; g.sy = (long)(nx + 2) * (ny + 2);
    g.N = (dim == 3) ? g.sy * (nz + 2) : g.sy;
    return g;
}

__device__ __forceinline__ long gidx(const Geo& g, int i, int j, int k) {
    return (long)i + (long)j * g.sx + (long)k * g.sy;
}

// Load / store through a pointer the compiler cannot prove to be global (one
// read back from LDS or from a descriptor in memory).  Through a generic
// pointer it emits flat instructions, which also count against the LDS wait
// counter: every later LDS wait would then wait for memory as well.
template <typename T>
__device__ __forceinline__ T gload(const T* p) {
    typedef unsigned u4 __attribute__((ext_vector_type(4)));
    typedef unsigned u2 __attribute__((ext_vector_type(2)));
    if constexpr (sizeof(T) == 32) {
        struct R { u4 a, b; } r;
        r.a = *(const __attribute__((address_space(1))) u4*)p;
        r.b = *((const __attribute__((address_space(1))) u4*)p + 1);
        return __builtin_bit_cast(T, r);
    } else if constexpr (sizeof(T) == 16) {
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) u4*)p);
    } else if constexpr (sizeof(T) == 8) {
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) u2*)p);
    } else if constexpr (sizeof(T) == 4) {
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) unsigned*)p);
    } else {
        static_assert(sizeof(T) == 2, "gload: 2, 4, 8, 16 or 32 bytes");
        return __builtin_bit_cast(T, *(const __attribute__((address_space(1))) unsigned short*)p);
    }
}
template <typename T>
__device__ __forceinline__ void gstore(T* p, T v) {
    if constexpr (sizeof(T) == 8) {
        *(__attribute__((address_space(1))) unsigned long long*)p = __builtin_bit_cast(unsigned long long, v);
    } else if constexpr (sizeof(T) == 4) {
        *(__attribute__((address_space(1))) unsigned*)p = __builtin_bit_cast(unsigned, v);
    } else {
        static_assert(sizeof(T) == 2, "gstore: 2, 4 or 8 bytes");
        *(__attribute__((address_space(1))) unsigned short*)p = __builtin_bit_cast(unsigned short, v);
    }
}

// Level-0 neighbour byte: which couplings of the fine stencil exist.
// bit0 fluid, bit1 +x, bit2 -x, bit3 +y, bit4 -y, bit5 +z, bit6 -z.
// A coupling exists iff both cells are fluid and the face is interior
// (setupFineCoeffs, uaamg_preconditioner_3d.cpp:58-66); for a fluid cell the
// -x bit equals the PCG matvec's "(i>1 && !solid(i-1))" test (pcg_3d.cpp:46).
enum : uint8_t { NB_FLUID = 1, NB_XP = 2, NB_XM = 4, NB_YP = 8, NB_YM = 16, NB_ZP = 32, NB_ZM = 64 };

// Per-level uniform constants in compute type CT.
template <typename CT>
struct Consts {
    CT idx2, idy2, idz2;       // 1/h^2 per axis at this level (level 0 only: fine constants)
    CT diag_u;                 // 2*(idx2+idy2[+idz2]) of the PCG matvec (pcg_3d.cpp:40)
    CT ex2[6];                 // Dirichlet diagonal excess per face (2/h^2), 0 for Neumann
    int wall[6];               // 1 = Dirichlet
    int any_dirichlet;
};

// Device-resident CG scalars; written by the last block of each reducing kernel.
struct Ctrl {
    double rsold, pAp, alpha, rsnew, rz, beta, mean_z, r0sq, tolsq, tol;
    double m1, m2, count;
    double sum_r, sum_z, sum_rz;
    double rz_cross;       // (r_new, z_old - mean_old) for the flexible beta
    double xalpha;         // alpha of this iteration's x update; 0 when the operator finish found the solve stopped
    double xa[2];          // the same for the split update's x pass, per p buffer: the x pass of one
                           // iteration may still run while the next iteration forms its alpha
    int iter, max_iter, done, status, first, tol_mode, project, flexible;
    double* hist;          // optional per-iteration r.r history (nullptr: off)
};

enum { ST_RUNNING = 0, ST_CONVERGED = 1, ST_MAXITER = 2, ST_BREAKDOWN = 3, ST_NONFINITE = 4 };

// Compute type for a storage type: FP64 stays FP64, everything else is FP32.
template <typename S> struct CTOf { using type = float; };
template <> struct CTOf<double> { using type = double; };

// Level record {cx, cy, cz, diag} of a storage type; 16-bit levels use an
// 8-byte record (one load per cell).
struct alignas(8) BF16x4 { __bf16 x, y, z, w; };
struct alignas(8) F16x4 { _Float16 x, y, z, w; };
template <typename S> struct Vec4;
template <> struct Vec4<double> { using type = double4; };
template <> struct Vec4<float> { using type = float4; };
template <> struct Vec4<__bf16> { using type = BF16x4; };
template <> struct Vec4<_Float16> { using type = F16x4; };

// linear block id / count for 3D grids
__device__ __forceinline__ int lin_block() {
    return blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
}
__device__ __forceinline__ int n_blocks() { return gridDim.x * gridDim.y * gridDim.z; }
__device__ __forceinline__ int lin_thread() { return threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z); }
__device__ __forceinline__ int n_threads() { return blockDim.x * blockDim.y * blockDim.z; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
    return v;
}

// Block-wide sum of NV doubles; result valid in thread 0. Fixed order.
template <int NV>
__device__ __forceinline__ void block_sum(double (&v)[NV]) {
    __shared__ double red[NV][16];
    int t = lin_thread(), nt = n_threads();
    int lane = t & (WAVE - 1), w = t / WAVE, nw = (nt + WAVE - 1) / WAVE;
#pragma unroll
    for (int q = 0; q < NV; q++) v[q] = wave_sum(v[q]);
    __syncthreads();
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < NV; q++) red[q][w] = v[q];
    __syncthreads();
    if (t == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) {
            double s = red[q][0];
            for (int x = 1; x < nw; x++) s += red[q][x];
            v[q] = s;
        }
    }
}

// Workgroup size of the one-workgroup finish kernels: partial counts are a
// few hundred to 2048, so four waves load them in a few rounds and the
// block-level sum crosses four waves instead of sixteen.
constexpr int FIN_THREADS = 256;

// Deterministic two-kernel reduction.  Producer kernels call block_store():
// each block writes its NV partial sums to partials[q * nblocks + block].  A
// one-workgroup finish kernel then adds the partials in block order
// (reduce_partials) and runs the scalar logic.  No cross-block fence or
// atomic is needed inside the producer: the kernel boundary makes the
// partials visible (an agent-scope release per block would write back the
// whole L2 on this chip).
template <int NV>
__device__ __forceinline__ void block_store(double (&v)[NV], double* partials) {
    int b = lin_block(), nb = n_blocks();
    block_sum<NV>(v);
    if (lin_thread() == 0) {
#pragma unroll
        for (int q = 0; q < NV; q++) partials[(long)q * nb + b] = v[q];
    }
}

// In one workgroup: totals of the partials of blocks [base, base + nb) of a
// producer launch of nbt blocks; valid in thread 0.
template <int NV>
__device__ __forceinline__ void reduce_partials_at(const double* partials, int nbt, int base, int nb, double (&tot)[NV]) {
    int t = lin_thread(), nt = n_threads();
#pragma unroll
    for (int q = 0; q < NV; q++) {
        double s = 0.0;
        for (int x = t; x < nb; x += nt) s += partials[(long)q * nbt + base + x];
        tot[q] = s;
    }
    block_sum<NV>(tot);
}
template <int NV>
__device__ __forceinline__ void reduce_partials(const double* partials, int nb, double (&tot)[NV]) {
    reduce_partials_at<NV>(partials, nb, 0, nb, tot);
}

}  // namespace uaamg
