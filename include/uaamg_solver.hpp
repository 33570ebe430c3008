// uaamg_solver.hpp -- C++ host-side mirror of the reference's solver plugin
// interface for the pressure solve, implemented over the C-ABI (uaamg.h).
//
// Same names and argument meaning as the reference:
//   Mesh2D/Grid, Mesh3D/Grid3D   include/core/mesh.h:17-44, grid.h:16-48,
//                                mesh_3d.h:16-47, grid_3d.h:14-38
//   Solver / Solver3D            include/solver/solver.h:18-34, solver_3d.h:7-12
//   Factory::create(key)         src/solver/factory.cpp:16-25 (string keys),
//                                with SolverKind mirroring those keys
//   CudaPCGSolver[3D]            include/solver/cuda_pcg_solver.h:14-32 (the
//                                GPU drop-in precedent) -> HipPCGSolver[3D]
//   amgpcg_solve                 test/solver/test_amgpcg.cpp:29-70
//   PressureProjection[3D]       include/pressure/pressure.h, pressure_3d.h;
//                                src/pressure/pressure.cpp:5-32, pressure_3d.cpp:5-39
//
// Behaviour kept from the reference: solve() writes g.p in place, adding the
// PCG iterate on fluid interior cells (pcg.cpp:107); breakdown and
// non-convergence return silently (pcg.cpp:100,112) -- the status is kept in
// last_status() for callers that want it.  Kept from the CUDA wrapper: the
// device context is reused while dims stay the same.  Changed on purpose: the
// solid mask is re-sent on every solve (cuda_pcg_solver.cu:45-49 sends it only
// on reallocation, so a moving solid with fixed dims went stale).
#pragma once
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "uaamg.h"

namespace uaamg_host {

struct Mesh2D {
    int nx, ny;
    double dx, dy;
    Mesh2D(int nx_, int ny_, double lx, double ly) : nx(nx_), ny(ny_), dx(lx / nx_), dy(ly / ny_) {}
    int u_size() const { return (nx + 1) * (ny + 2); }
    int v_size() const { return (nx + 2) * (ny + 1); }
    int p_size() const { return (nx + 2) * (ny + 2); }
    int iu(int i, int j) const { return i + j * (nx + 1); }
    int iv(int i, int j) const { return i + j * (nx + 2); }
    int ip(int i, int j) const { return i + j * (nx + 2); }
};

struct Mesh3D {
    int nx, ny, nz;
    double dx, dy, dz;
    Mesh3D(int nx_, int ny_, int nz_, double lx, double ly, double lz)
        : nx(nx_), ny(ny_), nz(nz_), dx(lx / nx_), dy(ly / ny_), dz(lz / nz_) {}
    int u_size() const { return (nx + 1) * (ny + 2) * (nz + 2); }
    int v_size() const { return (nx + 2) * (ny + 1) * (nz + 2); }
    int w_size() const { return (nx + 2) * (ny + 2) * (nz + 1); }
    int p_size() const { return (nx + 2) * (ny + 2) * (nz + 2); }
    int iu(int i, int j, int k) const { return i + j * (nx + 1) + k * (nx + 1) * (ny + 2); }
    int iv(int i, int j, int k) const { return i + j * (nx + 2) + k * (nx + 2) * (ny + 1); }
    int iw(int i, int j, int k) const { return i + j * (nx + 2) + k * (nx + 2) * (ny + 2); }
    int ip(int i, int j, int k) const { return i + j * (nx + 2) + k * (nx + 2) * (ny + 2); }
};

class Grid : public Mesh2D {
public:
    std::vector<double> u, v, p;
    std::vector<bool> solid;
    // Variable Laplacian coefficients (grid.h:21-23); empty: the uniform operator
    std::vector<double> lap_diag, lap_off_x, lap_off_y;
    Grid(int nx_, int ny_, double lx, double ly)
        : Mesh2D(nx_, ny_, lx, ly), u(u_size(), 0.0), v(v_size(), 0.0), p(p_size(), 0.0), solid(p_size(), false) {}
    bool has_variable_lap() const { return !lap_diag.empty(); }
    void init_variable_lap() {   // grid.cpp:12-17
        lap_diag.assign(p_size(), 0.0);
        lap_off_x.assign(p_size(), 0.0);
        lap_off_y.assign(p_size(), 0.0);
    }
    bool is_solid(int i, int j) const { return solid[ip(i, j)]; }
    void set_solid(int i, int j) { solid[ip(i, j)] = true; }
    double& u_at(int i, int j) { return u[iu(i, j)]; }
    double& v_at(int i, int j) { return v[iv(i, j)]; }
};

struct Grid3D : public Mesh3D {
    std::vector<double> u, v, w, p;
    std::vector<bool> solid;
    Grid3D(int nx_, int ny_, int nz_, double lx, double ly, double lz)
        : Mesh3D(nx_, ny_, nz_, lx, ly, lz), u(u_size(), 0.0), v(v_size(), 0.0), w(w_size(), 0.0),
          p(p_size(), 0.0), solid(p_size(), false) {}
    bool is_solid(int i, int j, int k) const { return solid[ip(i, j, k)]; }
    void set_solid(int i, int j, int k) { solid[ip(i, j, k)] = true; }
    double& u_at(int i, int j, int k) { return u[iu(i, j, k)]; }
    double& v_at(int i, int j, int k) { return v[iv(i, j, k)]; }
    double& w_at(int i, int j, int k) { return w[iw(i, j, k)]; }
};

class Solver {
public:
    virtual ~Solver() = default;
    virtual void solve(Grid& g, const std::vector<double>& rhs, int max_iter, double tol) = 0;
    virtual std::string name() const = 0;
};

class Solver3D {
public:
    virtual ~Solver3D() = default;
    virtual void solve(Grid3D& g, const std::vector<double>& rhs, int max_iter, double tol) = 0;
    virtual std::string name() const = 0;
};

// The reference's dispatch keys (factory.cpp:17-23, config.h:29).
enum class SolverKind { Jacobi, RBGS, CG, PCG, PCG_GMG, PCG_AMG, PCG_UAAMG, Unknown };
enum class Backend { CPU, HIP };

inline SolverKind parse_solver_kind(const std::string& key) {
    if (key == "jacobi") return SolverKind::Jacobi;
    if (key == "rbgs") return SolverKind::RBGS;
    if (key == "cg") return SolverKind::CG;
    if (key == "pcg") return SolverKind::PCG;          // the reference maps "pcg" to GMG
    if (key == "pcg_gmg") return SolverKind::PCG_GMG;
    if (key == "pcg_amg") return SolverKind::PCG_AMG;
    if (key == "pcg_uaamg") return SolverKind::PCG_UAAMG;
    return SolverKind::Unknown;
}

inline const char* solver_key(SolverKind k) {
    switch (k) {
        case SolverKind::Jacobi: return "jacobi";
        case SolverKind::RBGS: return "rbgs";
        case SolverKind::CG: return "cg";
        case SolverKind::PCG: return "pcg";
        case SolverKind::PCG_GMG: return "pcg_gmg";
        case SolverKind::PCG_AMG: return "pcg_amg";
        case SolverKind::PCG_UAAMG: return "pcg_uaamg";
        default: return "unknown";
    }
}

namespace detail {
struct Ctx {
    uaamg_ctx* h = nullptr;
    explicit Ctx(const uaamg_opts& o) {
        int rc = uaamg_create(&h, &o);
        if (rc != UAAMG_OK) throw std::runtime_error("uaamg_create failed: status " + std::to_string(rc));
    }
    ~Ctx() { uaamg_destroy(h); }
    Ctx(const Ctx&) = delete;
    Ctx& operator=(const Ctx&) = delete;
};
inline void check(uaamg_ctx* h, int rc, const char* what) {
    if (rc == UAAMG_OK || rc == UAAMG_NOT_CONVERGED || rc == UAAMG_BREAKDOWN || rc == UAAMG_NONFINITE) return;
    throw std::runtime_error(std::string(what) + ": " + uaamg_last_error(h));
}
}  // namespace detail

// PCG with the UAAMG preconditioner on the GPU, behind Solver.  opts default
// to the reference (FP64, V(1,1), absolute tolerance); pass UAAMG_FP32 etc.
// HipPCGSolver(false) is plain CG (the identity preconditioner), as
// CudaPCGSolver(use_precond) (cuda_pcg_solver.h:16) and factory key "cg";
// opts.preconditioner = UAAMG_PRECOND_GMG / _AMG gives keys "pcg" /
// "pcg_gmg" / "pcg_amg" (factory.cpp:20-22).
// A Grid with a variable Laplacian (has_variable_lap) is solved with it, as
// PCG::solve does (pcg.cpp:36-46).
class HipPCGSolver : public Solver {
public:
    explicit HipPCGSolver(const uaamg_opts* o = nullptr) {
        uaamg_default_opts(&opts_);
        if (o) opts_ = *o;
    }
    explicit HipPCGSolver(bool use_precond, const uaamg_opts* o = nullptr) : HipPCGSolver(o) {
        opts_.preconditioner = use_precond ? UAAMG_PRECOND_UAAMG : UAAMG_PRECOND_NONE;
    }
    void solve(Grid& g, const std::vector<double>& rhs, int max_iter, double tol) override {
        if ((int)rhs.size() != g.p_size()) throw std::invalid_argument("rhs size != (nx+2)(ny+2)");
        set_grid(g);
        uaamg_stats st{};
        int rc = uaamg_solve(ctx_->h, rhs.data(), g.p.data(), max_iter, tol, &st);
        detail::check(ctx_->h, rc, "uaamg_solve");
        last_ = st;
    }
    // rhs = div(u)/dt, p = 0, solve, u -= dt grad p, all on the device
    void project(Grid& g, double dt, int max_iter, double tol) {
        set_grid(g);
        uaamg_stats st{};
        int rc = uaamg_project(ctx_->h, g.u.data(), g.v.data(), nullptr, g.p.data(), dt, max_iter, tol, &st);
        detail::check(ctx_->h, rc, "uaamg_project");
        last_ = st;
    }
    std::string name() const override {
        switch (opts_.preconditioner) {
            case UAAMG_PRECOND_NONE: return "CG(HIP)";
            case UAAMG_PRECOND_GMG: return "PCG(HIP-GMG)";
            case UAAMG_PRECOND_AMG: return "PCG(HIP-AMG)";
            default: return "PCG(HIP-UAAMG)";
        }
    }
    const uaamg_stats& last_stats() const { return last_; }
    int last_status() const { return last_.status; }

private:
    void set_grid(const Grid& g) {
        if (!ctx_) ctx_ = std::make_unique<detail::Ctx>(opts_);
        mask_.assign(g.solid.begin(), g.solid.end());
        uaamg_grid gd{};
        gd.dim = 2; gd.n[0] = g.nx; gd.n[1] = g.ny; gd.n[2] = 1;
        gd.h[0] = g.dx; gd.h[1] = g.dy; gd.h[2] = 1.0;
        gd.solid = mask_.data();
        detail::check(ctx_->h, uaamg_set_grid(ctx_->h, &gd), "uaamg_set_grid");
        if (g.has_variable_lap()) {
            const size_t n = (size_t)g.p_size();
            if (g.lap_diag.size() != n || g.lap_off_x.size() != n || g.lap_off_y.size() != n)
                throw std::invalid_argument("variable Laplacian arrays must hold (nx+2)(ny+2) values");
            detail::check(ctx_->h, uaamg_set_variable_lap(ctx_->h, g.lap_diag.data(), g.lap_off_x.data(), g.lap_off_y.data()),
                          "uaamg_set_variable_lap");
        }
    }
    uaamg_opts opts_;
    std::unique_ptr<detail::Ctx> ctx_;
    std::vector<uint8_t> mask_;
    uaamg_stats last_{};
};

class HipPCGSolver3D : public Solver3D {
public:
    explicit HipPCGSolver3D(const uaamg_opts* o = nullptr) {
        uaamg_default_opts(&opts_);
        if (o) opts_ = *o;
    }
    explicit HipPCGSolver3D(bool use_precond, const uaamg_opts* o = nullptr) : HipPCGSolver3D(o) {
        opts_.preconditioner = use_precond ? UAAMG_PRECOND_UAAMG : UAAMG_PRECOND_NONE;
    }
    void solve(Grid3D& g, const std::vector<double>& rhs, int max_iter, double tol) override {
        if ((int)rhs.size() != g.p_size()) throw std::invalid_argument("rhs size != (nx+2)(ny+2)(nz+2)");
        set_grid(g);
        uaamg_stats st{};
        int rc = uaamg_solve(ctx_->h, rhs.data(), g.p.data(), max_iter, tol, &st);
        detail::check(ctx_->h, rc, "uaamg_solve");
        last_ = st;
    }
    // rhs = div(u)/dt, p = 0, solve, u -= dt grad p, all on the device
    void project(Grid3D& g, double dt, int max_iter, double tol) {
        set_grid(g);
        uaamg_stats st{};
        int rc = uaamg_project(ctx_->h, g.u.data(), g.v.data(), g.w.data(), g.p.data(), dt, max_iter, tol, &st);
        detail::check(ctx_->h, rc, "uaamg_project");
        last_ = st;
    }
    std::string name() const override {
        return opts_.preconditioner == UAAMG_PRECOND_NONE ? "CG3D(HIP)" : "PCG3D(HIP-UAAMG3D)";
    }
    const uaamg_stats& last_stats() const { return last_; }
    int last_status() const { return last_.status; }

private:
    void set_grid(const Grid3D& g) {
        if (!ctx_) ctx_ = std::make_unique<detail::Ctx>(opts_);
        mask_.assign(g.solid.begin(), g.solid.end());
        uaamg_grid gd{};
        gd.dim = 3; gd.n[0] = g.nx; gd.n[1] = g.ny; gd.n[2] = g.nz;
        gd.h[0] = g.dx; gd.h[1] = g.dy; gd.h[2] = g.dz;
        gd.solid = mask_.data();
        detail::check(ctx_->h, uaamg_set_grid(ctx_->h, &gd), "uaamg_set_grid");
    }
    uaamg_opts opts_;
    std::unique_ptr<detail::Ctx> ctx_;
    std::vector<uint8_t> mask_;
    uaamg_stats last_{};
};

// Factory::create of this library: the PCG keys "pcg_uaamg", "cg", "pcg",
// "pcg_gmg" and "pcg_amg" (factory.cpp:19-23) on the HIP backend have
// back-ends here; every other key (or the CPU backend) throws instead of the
// reference's silent fallback to Jacobi (factory.cpp:24), so a misspelled key
// cannot quietly select a different solver.  Factory3D has "pcg_uaamg" and
// "cg" (the reference's factory_3d.cpp:26-27 falls back to Jacobi3D for
// "pcg_uaamg" and has no GMG / AMG).
inline int preconditioner_of(SolverKind k) {
    switch (k) {
        case SolverKind::CG: return UAAMG_PRECOND_NONE;
        case SolverKind::PCG: case SolverKind::PCG_GMG: return UAAMG_PRECOND_GMG;   // factory.cpp:20-21
        case SolverKind::PCG_AMG: return UAAMG_PRECOND_AMG;
        case SolverKind::PCG_UAAMG: return UAAMG_PRECOND_UAAMG;
        default: return -1;
    }
}
struct Factory {
    static std::unique_ptr<Solver> create(const std::string& key, Backend be = Backend::HIP,
                                          const uaamg_opts* o = nullptr) {
        const int pc = preconditioner_of(parse_solver_kind(key));
        if (be == Backend::HIP && pc >= 0) {
            uaamg_opts oo;
            uaamg_default_opts(&oo);
            if (o) oo = *o;
            oo.preconditioner = pc;
            return std::make_unique<HipPCGSolver>(&oo);
        }
        throw std::invalid_argument("solver key '" + key + "' has no back-end in this library");
    }
};
struct Factory3D {
    static std::unique_ptr<Solver3D> create(const std::string& key, Backend be = Backend::HIP,
                                            const uaamg_opts* o = nullptr) {
        const SolverKind k = parse_solver_kind(key);
        if (be == Backend::HIP && (k == SolverKind::PCG_UAAMG || k == SolverKind::CG))
            return std::make_unique<HipPCGSolver3D>(k == SolverKind::PCG_UAAMG, o);
        throw std::invalid_argument("solver key '" + key + "' has no back-end in this library");
    }
};

// PressureProjection / PressureProjection3D::project (pressure.cpp:5-32,
// pressure_3d.cpp:5-39): same signature; the solver must be the HIP one (its
// context runs the rhs assembly, the solve and the correction on the device).
struct PressureProjection {
    static void project(Grid& g, double dt, Solver& solver, int max_iter, double tol) {
        auto* s = dynamic_cast<HipPCGSolver*>(&solver);
        if (!s) throw std::invalid_argument("PressureProjection::project needs HipPCGSolver (the HIP solver)");
        s->project(g, dt, max_iter, tol);
    }
};
struct PressureProjection3D {
    static void project(Grid3D& g, double dt, Solver3D& solver, int max_iter, double tol) {
        auto* s = dynamic_cast<HipPCGSolver3D*>(&solver);
        if (!s) throw std::invalid_argument("PressureProjection3D::project needs HipPCGSolver3D (the HIP solver)");
        s->project(g, dt, max_iter, tol);
    }
};

// Standalone amgpcg (test/test_common.h:7-14 grid: N x N, h = 1/(N+1),
// row-major v[i*(N+2)+j]).
namespace amg {
struct Grid {
    int N;
    double h;
    std::vector<double> v;
    explicit Grid(int n) : N(n), h(1.0 / (n + 1)), v((size_t)(n + 2) * (n + 2), 0.0) {}
    double operator()(int i, int j) const { return v[(size_t)i * (N + 2) + j]; }
    double& operator()(int i, int j) { return v[(size_t)i * (N + 2) + j]; }
};
inline Grid amgpcg_solve(const Grid& b, int max_iter, double tol, int& out_iters, int precision = UAAMG_FP64) {
    Grid x(b.N);
    int rc = uaamg_amgpcg(b.N, b.v.data(), x.v.data(), max_iter, tol, &out_iters, precision);
    if (rc != UAAMG_OK) throw std::runtime_error("uaamg_amgpcg failed: status " + std::to_string(rc));
    return x;
}
}  // namespace amg

}  // namespace uaamg_host
