// uaamg_vtk.hpp -- ASCII legacy-VTK frames of the host mirror's grids, the
// on-disk format the reference writes for its visual checks:
//   VtkWriter::write    src/io/vtk_writer.cpp:9-83   (include/io/vtk_writer.h)
//   VtkWriter3D::write  src/io/vtk_writer_3d.cpp:10-85 (include/io/vtk_writer_3d.h)
// A frame is <out_dir>/frame_<5-digit frame>.vtk: STRUCTURED_POINTS on the
// (nx+1) x (ny+1) [x (nz+1)] cell corners, spacing (dx, dy[, dz]), with
// POINT_DATA velocity (each component the mean of the MAC faces that touch
// the corner inside the domain, 0 where none), vorticity (2D: dv/dx - du/dy
// from the faces around the corner; 3D: magnitude of the centred-difference
// curl at the clamped cell), divergence and solid marker at the clamped
// cell, every number as the reference prints it (scientific, 6 digits), so
// that a file equals the reference's byte for byte for the same fields.
// Host-side output: after HipPCGSolver[3D]::project the grid's u, v, (w), p
// are on the host.
#pragma once
#include <algorithm>
#include <cmath>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>

#include "uaamg_solver.hpp"

namespace uaamg_host {

namespace vtk_detail {
inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

inline std::string frame_path(const std::string& out_dir, int frame) {
    std::ostringstream ss;
    ss << out_dir << "/frame_" << std::setw(5) << std::setfill('0') << frame << ".vtk";
    return ss.str();
}

inline void header(std::ostream& f, const char* what, int frame, int nx, int ny, int nz, double dx, double dy,
                   double dz, bool three) {
    f << std::scientific << std::setprecision(6);
    f << "# vtk DataFile Version 2.0\n";
    f << what << frame << "\n";
    f << "ASCII\n";
    f << "DATASET STRUCTURED_POINTS\n";
    if (three) {
        f << "DIMENSIONS " << (nx + 1) << " " << (ny + 1) << " " << (nz + 1) << "\n";
        f << "ORIGIN 0 0 0\n";
        f << "SPACING " << dx << " " << dy << " " << dz << "\n";
        f << "POINT_DATA " << (nx + 1) * (ny + 1) * (nz + 1) << "\n";
    } else {
        f << "DIMENSIONS " << (nx + 1) << " " << (ny + 1) << " 1\n";
        f << "ORIGIN 0 0 0\n";
        f << "SPACING " << dx << " " << dy << " 1\n";
        f << "POINT_DATA " << (nx + 1) * (ny + 1) << "\n";
    }
}
}  // namespace vtk_detail

struct VtkWriter {
    // Returns false (and writes nothing) when the file cannot be opened, where
    // the reference prints "Cannot write <path>" and returns.
    static bool write(const Grid& g, int frame, const std::string& out_dir) {
        using vtk_detail::clampi;
        const std::string path = vtk_detail::frame_path(out_dir, frame);
        std::ofstream f(path);
        if (!f) return false;
        vtk_detail::header(f, "LFM 2D Fluid - Frame ", frame, g.nx, g.ny, 1, g.dx, g.dy, 1.0, false);
        auto U = [&](int i, int j) { return g.u[g.iu(i, j)]; };
        auto V = [&](int i, int j) { return g.v[g.iv(i, j)]; };
        f << "VECTORS velocity float\n";
        for (int j = 0; j <= g.ny; j++)
            for (int i = 0; i <= g.nx; i++) {
                double su = 0.0, sv = 0.0;
                int nu = 0, nv = 0;
                const bool jin = j > 0 && j <= g.ny, iin = i > 0 && i <= g.nx;
                if (i > 0 && jin) { su += U(i - 1, j); nu++; }
                if (i < g.nx && jin) { su += U(i, j); nu++; }
                if (j > 0 && iin) { sv += V(i, j - 1); nv++; }
                if (j < g.ny && iin) { sv += V(i, j); nv++; }
                f << (nu > 0 ? su / nu : 0.0) << " " << (nv > 0 ? sv / nv : 0.0) << " 0\n";
            }
        f << "SCALARS vorticity float 1\n";
        f << "LOOKUP_TABLE default\n";
        for (int j = 0; j <= g.ny; j++)
            for (int i = 0; i <= g.nx; i++) {
                double dvdx = 0.0, dudy = 0.0;
                if (i > 0 && i < g.nx && j > 0 && j <= g.ny) dvdx = (V(i + 1, j) - V(i, j)) / g.dx;
                if (i > 0 && i <= g.nx && j > 0 && j < g.ny) dudy = (U(i, j + 1) - U(i, j)) / g.dy;
                f << (dvdx - dudy) << "\n";
            }
        f << "SCALARS divergence float 1\n";
        f << "LOOKUP_TABLE default\n";
        for (int j = 0; j <= g.ny; j++)
            for (int i = 0; i <= g.nx; i++) {
                const int ci = clampi(i, 1, g.nx), cj = clampi(j, 1, g.ny);
                f << ((U(ci, cj) - U(ci - 1, cj)) / g.dx + (V(ci, cj) - V(ci, cj - 1)) / g.dy) << "\n";
            }
        f << "SCALARS solid float 1\n";
        f << "LOOKUP_TABLE default\n";
        for (int j = 0; j <= g.ny; j++)
            for (int i = 0; i <= g.nx; i++)
                f << (g.is_solid(clampi(i, 1, g.nx), clampi(j, 1, g.ny)) ? 1.0 : 0.0) << "\n";
        return bool(f);
    }
    static bool write(const Grid& g, int frame, const char* out_dir) { return write(g, frame, std::string(out_dir)); }
    // the reference's signature takes its Config and reads out_dir
    template <typename Cfg>
    static bool write(const Grid& g, int frame, const Cfg& cfg) { return write(g, frame, std::string(cfg.out_dir)); }
};

struct VtkWriter3D {
    static bool write(const Grid3D& g, int frame, const std::string& out_dir) {
        using vtk_detail::clampi;
        const std::string path = vtk_detail::frame_path(out_dir, frame);
        std::ofstream f(path);
        if (!f) return false;
        vtk_detail::header(f, "LFM 3D Fluid - Frame ", frame, g.nx, g.ny, g.nz, g.dx, g.dy, g.dz, true);
        auto U = [&](int i, int j, int k) { return g.u[g.iu(i, j, k)]; };
        auto V = [&](int i, int j, int k) { return g.v[g.iv(i, j, k)]; };
        auto W = [&](int i, int j, int k) { return g.w[g.iw(i, j, k)]; };
        f << "VECTORS velocity float\n";
        for (int k = 0; k <= g.nz; k++)
            for (int j = 0; j <= g.ny; j++)
                for (int i = 0; i <= g.nx; i++) {
                    double su = 0.0, sv = 0.0, sw = 0.0;
                    int nu = 0, nv = 0, nw = 0;
                    const bool iin = i > 0 && i <= g.nx, jin = j > 0 && j <= g.ny, kin = k > 0 && k <= g.nz;
                    if (i > 0 && jin && kin) { su += U(i - 1, j, k); nu++; }
                    if (i < g.nx && jin && kin) { su += U(i, j, k); nu++; }
                    if (j > 0 && iin && kin) { sv += V(i, j - 1, k); nv++; }
                    if (j < g.ny && iin && kin) { sv += V(i, j, k); nv++; }
                    if (k > 0 && iin && jin) { sw += W(i, j, k - 1); nw++; }
                    if (k < g.nz && iin && jin) { sw += W(i, j, k); nw++; }
                    f << (nu > 0 ? su / nu : 0.0) << " " << (nv > 0 ? sv / nv : 0.0) << " "
                      << (nw > 0 ? sw / nw : 0.0) << "\n";
                }
        // |curl u| at the clamped cell, centred differences clamped to the
        // interior (fvc::vorticity, include/ops/operators_3d.h:44-60)
        f << "SCALARS vorticity_magnitude float 1\nLOOKUP_TABLE default\n";
        for (int k = 0; k <= g.nz; k++)
            for (int j = 0; j <= g.ny; j++)
                for (int i = 0; i <= g.nx; i++) {
                    const int ci = clampi(i, 1, g.nx), cj = clampi(j, 1, g.ny), ck = clampi(k, 1, g.nz);
                    const int ip = std::min(ci + 1, g.nx), im = std::max(ci - 1, 1);
                    const int jp = std::min(cj + 1, g.ny), jm = std::max(cj - 1, 1);
                    const int kp = std::min(ck + 1, g.nz), km = std::max(ck - 1, 1);
                    const double dw_dy = (W(ci, jp, ck) - W(ci, jm, ck)) / (2.0 * g.dy);
                    const double dv_dz = (V(ci, cj, kp) - V(ci, cj, km)) / (2.0 * g.dz);
                    const double du_dz = (U(ci, cj, kp) - U(ci, cj, km)) / (2.0 * g.dz);
                    const double dw_dx = (W(ip, cj, ck) - W(im, cj, ck)) / (2.0 * g.dx);
                    const double dv_dx = (V(ip, cj, ck) - V(im, cj, ck)) / (2.0 * g.dx);
                    const double du_dy = (U(ci, jp, ck) - U(ci, jm, ck)) / (2.0 * g.dy);
                    const double a = dw_dy - dv_dz, b = du_dz - dw_dx, c = dv_dx - du_dy;
                    f << std::sqrt(a * a + b * b + c * c) << "\n";
                }
        f << "SCALARS divergence float 1\nLOOKUP_TABLE default\n";
        for (int k = 0; k <= g.nz; k++)
            for (int j = 0; j <= g.ny; j++)
                for (int i = 0; i <= g.nx; i++) {
                    const int ci = clampi(i, 1, g.nx), cj = clampi(j, 1, g.ny), ck = clampi(k, 1, g.nz);
                    f << ((U(ci, cj, ck) - U(ci - 1, cj, ck)) / g.dx + (V(ci, cj, ck) - V(ci, cj - 1, ck)) / g.dy
                          + (W(ci, cj, ck) - W(ci, cj, ck - 1)) / g.dz) << "\n";
                }
        f << "SCALARS solid float 1\nLOOKUP_TABLE default\n";
        for (int k = 0; k <= g.nz; k++)
            for (int j = 0; j <= g.ny; j++)
                for (int i = 0; i <= g.nx; i++)
                    f << (g.is_solid(clampi(i, 1, g.nx), clampi(j, 1, g.ny), clampi(k, 1, g.nz)) ? 1.0 : 0.0) << "\n";
        return bool(f);
    }
    static bool write(const Grid3D& g, int frame, const char* out_dir) { return write(g, frame, std::string(out_dir)); }
    template <typename Cfg>
    static bool write(const Grid3D& g, int frame, const Cfg& cfg) { return write(g, frame, std::string(cfg.out_dir)); }
};

}  // namespace uaamg_host
