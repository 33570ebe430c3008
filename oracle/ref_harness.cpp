// ref_harness.cpp -- C-ABI shim over the reference's own CPU solver, compiled
// from the reference sources where they lie (see oracle/Makefile, target ref).
// Test infrastructure only: used to pin the oracle restatement and to write
// golden fixtures; output goes to oracle/_ref/ and is never shipped.
//
// Wraps: PCG/PCG3D (src/solver/pcg.cpp, pcg_3d.cpp) with
// UAAMGPreconditioner/UAAMGPreconditioner3D; Grid/Grid3D built with
// lengths lx,ly,lz so dx = lx/nx exactly as mesh.h:21-22.
#include "core/grid.h"
#include "core/grid_3d.h"
#include "solver/pcg.h"
#include "solver/pcg_3d.h"
#include "solver/preconditioner/uaamg_preconditioner.h"
#include "solver/preconditioner/uaamg_preconditioner_3d.h"
#include "solver/preconditioner/identity_preconditioner.h"
#include "solver/preconditioner/identity_preconditioner_3d.h"
#include "solver/preconditioner/gmg_preconditioner.h"
#include "solver/preconditioner/amg_preconditioner.h"
#include "geometry/cylinder_model.h"
#include "pressure/pressure.h"
#include "pressure/pressure_3d.h"
#include "io/vtk_writer.h"
#include "io/vtk_writer_3d.h"
#include "config/config.h"
#include <memory>
#include <vector>

namespace {
// Counts preconditioner applications: PCG applies once before its loop and
// once per loop iteration that does not exit on the tolerance test.
struct Count2D : Preconditioner {
    UAAMGPreconditioner inner; int n = 0;
    void apply(const Grid& g, const std::vector<double>& r, std::vector<double>& z) override { n++; inner.apply(g, r, z); }
    std::string name() const override { return inner.name(); }
};
struct Count3D : Preconditioner3D {
    UAAMGPreconditioner3D inner; int n = 0;
    void apply(const Grid3D& g, const std::vector<double>& r, std::vector<double>& z) override { n++; inner.apply(g, r, z); }
    std::string name() const override { return inner.name(); }
};
}

extern "C" int ref_pcg(int dim, int nx, int ny, int nz, double lx, double ly, double lz,
                       const unsigned char* solid, const double* rhs, double* p,
                       int max_iter, double tol, int* applies) {
    if (dim == 2) {
        Grid g(nx, ny, lx, ly);
        size_t N = g.p.size();
        for (size_t c = 0; c < N; c++) { g.solid[c] = solid[c] != 0; g.p[c] = p[c]; }
        std::vector<double> r(rhs, rhs + N);
        auto pc = std::make_unique<Count2D>(); Count2D* raw = pc.get();
        PCG s(std::move(pc));
        s.solve(g, r, max_iter, tol);
        for (size_t c = 0; c < N; c++) p[c] = g.p[c];
        *applies = raw->n;
        return 0;
    }
    Grid3D g(nx, ny, nz, lx, ly, lz);
    size_t N = g.p.size();
    for (size_t c = 0; c < N; c++) { g.solid[c] = solid[c] != 0; g.p[c] = p[c]; }
    std::vector<double> r(rhs, rhs + N);
    auto pc = std::make_unique<Count3D>(); Count3D* raw = pc.get();
    PCG3D s(std::move(pc));
    s.solve(g, r, max_iter, tol);
    for (size_t c = 0; c < N; c++) p[c] = g.p[c];
    *applies = raw->n;
    return 0;
}

// PCG with the preconditioner of a factory key (factory.cpp:19-23,
// factory_3d.cpp:19): precond 0 = UAAMGPreconditioner[3D] ("pcg_uaamg"),
// 1 = IdentityPreconditioner[3D] ("cg"), 2 = GMGPreconditioner ("pcg",
// "pcg_gmg", 2D), 3 = AMGPreconditioner ("pcg_amg", 2D); lap_diag/lap_off_x/lap_off_y
// (2D, all or none) fill Grid's variable Laplacian (grid.h:21-23,
// init_variable_lap, grid.cpp) that PCG::solve's matvec reads (pcg.cpp:36-46).
extern "C" int ref_pcg_ex(int dim, int nx, int ny, int nz, double lx, double ly, double lz,
                          const unsigned char* solid, const double* rhs, double* p,
                          int max_iter, double tol, int precond,
                          const double* lap_diag, const double* lap_off_x, const double* lap_off_y) {
    if (dim == 2) {
        Grid g(nx, ny, lx, ly);
        size_t N = g.p.size();
        for (size_t c = 0; c < N; c++) { g.solid[c] = solid[c] != 0; g.p[c] = p[c]; }
        if (lap_diag) {
            g.init_variable_lap();
            for (size_t c = 0; c < N; c++) { g.lap_diag[c] = lap_diag[c]; g.lap_off_x[c] = lap_off_x[c]; g.lap_off_y[c] = lap_off_y[c]; }
        }
        std::vector<double> r(rhs, rhs + N);
        std::unique_ptr<Preconditioner> pc;
        if (precond == 1) pc = std::make_unique<IdentityPreconditioner>();
        else if (precond == 2) pc = std::make_unique<GMGPreconditioner>();
        else if (precond == 3) pc = std::make_unique<AMGPreconditioner>();
        else pc = std::make_unique<UAAMGPreconditioner>();
        PCG s(std::move(pc));
        s.solve(g, r, max_iter, tol);
        for (size_t c = 0; c < N; c++) p[c] = g.p[c];
        return 0;
    }
    if (lap_diag || precond >= 2) return -1;
    Grid3D g(nx, ny, nz, lx, ly, lz);
    size_t N = g.p.size();
    for (size_t c = 0; c < N; c++) { g.solid[c] = solid[c] != 0; g.p[c] = p[c]; }
    std::vector<double> r(rhs, rhs + N);
    std::unique_ptr<Preconditioner3D> pc;
    if (precond == 1) pc = std::make_unique<IdentityPreconditioner3D>();
    else pc = std::make_unique<UAAMGPreconditioner3D>();
    PCG3D s(std::move(pc));
    s.solve(g, r, max_iter, tol);
    for (size_t c = 0; c < N; c++) p[c] = g.p[c];
    return 0;
}

extern "C" int ref_uaamg_apply(int dim, int nx, int ny, int nz, double lx, double ly, double lz,
                               const unsigned char* solid, const double* r, double* z) {
    if (dim == 2) {
        Grid g(nx, ny, lx, ly);
        size_t N = g.p.size();
        for (size_t c = 0; c < N; c++) g.solid[c] = solid[c] != 0;
        std::vector<double> rv(r, r + N), zv(z, z + N);
        UAAMGPreconditioner pc; pc.apply(g, rv, zv);
        for (size_t c = 0; c < N; c++) z[c] = zv[c];
        return 0;
    }
    Grid3D g(nx, ny, nz, lx, ly, lz);
    size_t N = g.p.size();
    for (size_t c = 0; c < N; c++) g.solid[c] = solid[c] != 0;
    std::vector<double> rv(r, r + N), zv(z, z + N);
    UAAMGPreconditioner3D pc; pc.apply(g, rv, zv);
    for (size_t c = 0; c < N; c++) z[c] = zv[c];
    return 0;
}

// SmoothCylinder::applyToLaplacian (src/geometry/cylinder_model.cpp:50-71) on a
// fresh Grid: the solid mask it marks and the variable Laplacian it writes.
extern "C" int ref_smooth_cylinder_lap(int nx, int ny, double lx, double ly, double cx, double cy, double R,
                                       double tw, unsigned char* solid, double* diag, double* off_x, double* off_y) {
    Grid g(nx, ny, lx, ly);
    g.init_variable_lap();
    SmoothCylinder cyl(cx, cy, R, g.dx, g.dy, tw);
    cyl.applyToLaplacian(g, g.lap_diag, g.lap_off_x, g.lap_off_y);
    for (size_t c = 0; c < g.p.size(); c++) {
        solid[c] = g.solid[c] ? 1 : 0;
        diag[c] = g.lap_diag[c]; off_x[c] = g.lap_off_x[c]; off_y[c] = g.lap_off_y[c];
    }
    return 0;
}

// dx exactly as the reference's Grid computes it, for building matching inputs.
extern "C" double ref_spacing(double l, int n) { Grid g(n, 2, l, 1.0); return g.dx; }

// PressureProjection[3D]::project (src/pressure/pressure.cpp:5-32,
// pressure_3d.cpp:5-39) with the pcg_uaamg solver on MAC velocities in the
// Grid's own layout; u, v, w, p are copied in and out.
extern "C" int ref_project(int dim, int nx, int ny, int nz, double lx, double ly, double lz,
                           const unsigned char* solid, double* u, double* v, double* w, double* p,
                           double dt, int max_iter, double tol) {
    if (dim == 2) {
        Grid g(nx, ny, lx, ly);
        for (size_t c = 0; c < g.p.size(); c++) { g.solid[c] = solid[c] != 0; g.p[c] = p[c]; }
        for (size_t c = 0; c < g.u.size(); c++) g.u[c] = u[c];
        for (size_t c = 0; c < g.v.size(); c++) g.v[c] = v[c];
        PCG s(std::make_unique<UAAMGPreconditioner>());
        PressureProjection::project(g, dt, s, max_iter, tol);
        for (size_t c = 0; c < g.p.size(); c++) p[c] = g.p[c];
        for (size_t c = 0; c < g.u.size(); c++) u[c] = g.u[c];
        for (size_t c = 0; c < g.v.size(); c++) v[c] = g.v[c];
        return 0;
    }
    Grid3D g(nx, ny, nz, lx, ly, lz);
    for (size_t c = 0; c < g.p.size(); c++) { g.solid[c] = solid[c] != 0; g.p[c] = p[c]; }
    for (size_t c = 0; c < g.u.size(); c++) g.u[c] = u[c];
    for (size_t c = 0; c < g.v.size(); c++) g.v[c] = v[c];
    for (size_t c = 0; c < g.w.size(); c++) g.w[c] = w[c];
    PCG3D s(std::make_unique<UAAMGPreconditioner3D>());
    PressureProjection3D::project(g, dt, s, max_iter, tol);
    for (size_t c = 0; c < g.p.size(); c++) p[c] = g.p[c];
    for (size_t c = 0; c < g.u.size(); c++) u[c] = g.u[c];
    for (size_t c = 0; c < g.v.size(); c++) v[c] = g.v[c];
    for (size_t c = 0; c < g.w.size(); c++) w[c] = g.w[c];
    return 0;
}

// VtkWriter::write / VtkWriter3D::write (src/io/vtk_writer.cpp:9-83,
// vtk_writer_3d.cpp:10-85) on a grid built from the given MAC arrays and
// padded solid mask: writes <out_dir>/frame_<frame>.vtk.
extern "C" int ref_write_vtk(int dim, int nx, int ny, int nz, double lx, double ly, double lz,
                             const double* u, const double* v, const double* w, const unsigned char* solid,
                             int frame, const char* out_dir) {
    Config cfg;
    cfg.out_dir = out_dir;
    if (dim == 2) {
        Grid g(nx, ny, lx, ly);
        g.u.assign(u, u + g.u.size());
        g.v.assign(v, v + g.v.size());
        for (size_t c = 0; c < g.solid.size(); c++) g.solid[c] = solid[c] != 0;
        VtkWriter::write(g, frame, cfg);
    } else {
        Grid3D g(nx, ny, nz, lx, ly, lz);
        g.u.assign(u, u + g.u.size());
        g.v.assign(v, v + g.v.size());
        g.w.assign(w, w + g.w.size());
        for (size_t c = 0; c < g.solid.size(); c++) g.solid[c] = solid[c] != 0;
        VtkWriter3D::write(g, frame, cfg);
    }
    return 0;
}
