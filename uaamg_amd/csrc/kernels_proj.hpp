// kernels_proj.hpp -- the pressure projection around the solve, on device.
//
// PressureProjection::project (pressure.cpp:5-32) and
// PressureProjection3D::project (pressure_3d.cpp:5-39) on the reference's
// MAC arrays (FP64, mesh.h:27-35 / mesh_3d.h:29-37):
//   rhs = div(u) / dt on non-solid interior cells, 0 elsewhere
//         (fvc::divergence, operators.h:18-21 / operators_3d.h:17-21);
//   after the solve, on each interior face between two non-solid cells,
//         u -= dt (p+ - p-) / dx   (and v, w).
// One thread per padded cell; the cell (i, j, k) owns its +x, +y, +z faces,
// so every face is written by one thread.  Same operations and order as the
// reference loops (divisions, not reciprocal products), so FP64 results are
// bit-identical.  Both passes stream u, v, w, p once (HBM-bound).
#pragma once
#include "common.hpp"

namespace uaamg {

struct MacGeo {
    long usx, usy;   // u: i + j (nx+1) + k (nx+1)(ny+2)
    long vsx, vsy;   // v: i + j (nx+2) + k (nx+2)(ny+1)
    long wsx, wsy;   // w: i + j (nx+2) + k (nx+2)(ny+2)
};

inline MacGeo make_mac(const Geo& g) {
    MacGeo m;
    m.usx = g.nx + 1; m.usy = (long)(g.nx + 1) * (g.ny + 2);
    m.vsx = g.nx + 2; m.vsy = (long)(g.nx + 2) * (g.ny + 1);
    m.wsx = g.nx + 2; m.wsy = (long)(g.nx + 2) * (g.ny + 2);
    return m;
}

template <int D>
__global__ void k_proj_rhs(const double* __restrict__ u, const double* __restrict__ v, const double* __restrict__ w,
                           const uint8_t* __restrict__ solid, double* __restrict__ rhs, Geo g, MacGeo m,
                           double dx, double dy, double dz, double dt) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N; c += (long)gridDim.x * blockDim.x) {
        const int i = (int)(c % g.sx);
        const long rest = c / g.sx;
        const int j = (int)(D == 3 ? rest % (g.ny + 2) : rest);
        const int k = (int)(D == 3 ? rest / (g.ny + 2) : 0);
        const bool in = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && (D == 2 || (k >= 1 && k <= g.nz));
        double r = 0.0;
        if (in && !solid[c]) {
            const long iu = i + j * m.usx + k * m.usy, iv = i + j * m.vsx + k * m.vsy;
            double div = (u[iu] - u[iu - 1]) / dx + (v[iv] - v[iv - m.vsx]) / dy;
            if (D == 3) {
                const long iw = i + j * m.wsx + k * m.wsy;
                div = div + (w[iw] - w[iw - m.wsy]) / dz;
            }
            r = div / dt;
        }
        rhs[c] = r;
    }
}

template <int D>
__global__ void k_proj_correct(double* __restrict__ u, double* __restrict__ v, double* __restrict__ w,
                               const double* __restrict__ p, const uint8_t* __restrict__ solid, Geo g, MacGeo m,
                               double dx, double dy, double dz, double dt) {
    for (long c = (long)blockIdx.x * blockDim.x + threadIdx.x; c < g.N; c += (long)gridDim.x * blockDim.x) {
        const int i = (int)(c % g.sx);
        const long rest = c / g.sx;
        const int j = (int)(D == 3 ? rest % (g.ny + 2) : rest);
        const int k = (int)(D == 3 ? rest / (g.ny + 2) : 0);
        const bool in = i >= 1 && i <= g.nx && j >= 1 && j <= g.ny && (D == 2 || (k >= 1 && k <= g.nz));
        if (!in || solid[c]) continue;
        if (i < g.nx && !solid[c + 1]) {
            const long iu = i + j * m.usx + k * m.usy;
            u[iu] -= dt * (p[c + 1] - p[c]) / dx;
        }
        if (j < g.ny && !solid[c + g.sx]) {
            const long iv = i + j * m.vsx + k * m.vsy;
            v[iv] -= dt * (p[c + g.sx] - p[c]) / dy;
        }
        if (D == 3 && k < g.nz && !solid[c + g.sy]) {
            const long iw = i + j * m.wsx + k * m.wsy;
            w[iw] -= dt * (p[c + g.sy] - p[c]) / dz;
        }
    }
}

}  // namespace uaamg
