// Drives the C++ host mirror (include/uaamg_solver.hpp) the way the
// reference's application drives its Solver plugins (Factory::create(key),
// solver->solve(grid, rhs, max_iter, tol), g.p accumulated in place), and
// writes inputs and results as raw float64 files for the Python parity test.
//
//   host_solver_test keys          key parsing / factory errors (no GPU)
//   host_solver_test solve OUTDIR  GPU solves through the Solver interface
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "uaamg_solver.hpp"
#include "uaamg_vtk.hpp"

using namespace uaamg_host;

static int fails = 0;
#define CHECK(c) do { if (!(c)) { std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #c, __LINE__); ++fails; } } while (0)

static void write_f64(const std::string& path, const std::vector<double>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); ++fails; return; }
    std::fwrite(v.data(), sizeof(double), v.size(), f);
    std::fclose(f);
}

static void write_u8(const std::string& path, const std::vector<bool>& v) {
    std::vector<unsigned char> b(v.begin(), v.end());
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); ++fails; return; }
    std::fwrite(b.data(), 1, b.size(), f);
    std::fclose(f);
}

static int run_keys() {
    const char* keys[] = {"jacobi", "rbgs", "cg", "pcg", "pcg_gmg", "pcg_amg", "pcg_uaamg"};
    for (const char* k : keys) CHECK(std::string(solver_key(parse_solver_kind(k))) == k);
    CHECK(parse_solver_kind("pcg-uaamg") == SolverKind::Unknown);
    bool threw = false;
    try { Factory::create("jacobi"); } catch (const std::invalid_argument&) { threw = true; }
    CHECK(threw);
    threw = false;
    try { Factory3D::create("pcg_uaamg", Backend::CPU); } catch (const std::invalid_argument&) { threw = true; }
    CHECK(threw);
    threw = false;
    try { Factory3D::create("pcg_gmg"); } catch (const std::invalid_argument&) { threw = true; }
    CHECK(threw);   // the 3D factory has no GMG / AMG (factory_3d.cpp:21-23)
    // the PCG keys have HIP back-ends (contexts are created on the first solve)
    CHECK(Factory::create("cg")->name() == "CG(HIP)");
    CHECK(Factory::create("pcg")->name() == "PCG(HIP-GMG)");
    CHECK(Factory::create("pcg_gmg")->name() == "PCG(HIP-GMG)");
    CHECK(Factory::create("pcg_amg")->name() == "PCG(HIP-AMG)");
    CHECK(Factory3D::create("cg")->name() == "CG3D(HIP)");
    CHECK(Factory::create("pcg_uaamg")->name() == "PCG(HIP-UAAMG)");
    CHECK(Factory3D::create("pcg_uaamg")->name() == "PCG3D(HIP-UAAMG3D)");
    std::printf("keys %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}

static int run_solve(const std::string& out) {
    // 3D: 24x16x12 on 1.5 x 1.0 x 0.75 with a solid ball, Gaussian source.
    {
        Grid3D g(24, 16, 12, 1.5, 1.0, 0.75);
        std::vector<double> rhs(g.p_size(), 0.0);
        for (int k = 1; k <= g.nz; ++k)
            for (int j = 1; j <= g.ny; ++j)
                for (int i = 1; i <= g.nx; ++i) {
                    double dx = i - 16.0, dy = j - 8.0, dz = k - 6.0;
                    if (dx * dx + dy * dy + dz * dz < 9.0) g.set_solid(i, j, k);
                    double ex = i - 6.0, ey = j - 9.0, ez = k - 5.0;
                    rhs[g.ip(i, j, k)] = 1e3 * std::exp(-0.05 * (ex * ex + ey * ey + ez * ez));
                }
        uaamg_opts o;
        uaamg_default_opts(&o);
        o.tol_mode = UAAMG_TOL_REL;
        auto s = Factory3D::create("pcg_uaamg", Backend::HIP, &o);
        CHECK(s->name() == "PCG3D(HIP-UAAMG3D)");
        s->solve(g, rhs, 300, 1e-10);
        auto* hs = dynamic_cast<HipPCGSolver3D*>(s.get());
        CHECK(hs && hs->last_status() == UAAMG_OK);
        std::printf("3d iters %d status %d\n", hs->last_stats().iters, hs->last_status());
        write_f64(out + "/g3_rhs.f64", rhs);
        write_u8(out + "/g3_solid.u8", g.solid);
        write_f64(out + "/g3_p1.f64", g.p);
        // second solve of the same system: p accumulates (p += x)
        s->solve(g, rhs, 300, 1e-10);
        write_f64(out + "/g3_p2.f64", g.p);
    }
    // 2D: 40x24 on 2.0 x 1.2 with a solid box.
    {
        Grid g(40, 24, 2.0, 1.2);
        std::vector<double> rhs(g.p_size(), 0.0);
        for (int j = 1; j <= g.ny; ++j)
            for (int i = 1; i <= g.nx; ++i) {
                if (i >= 20 && i < 26 && j >= 8 && j < 15) g.set_solid(i, j);
                rhs[g.ip(i, j)] = 1e3 * std::sin(0.3 * i) * std::cos(0.2 * j);
            }
        uaamg_opts o;
        uaamg_default_opts(&o);
        o.tol_mode = UAAMG_TOL_REL;
        auto s = Factory::create("pcg_uaamg", Backend::HIP, &o);
        s->solve(g, rhs, 300, 1e-10);
        auto* hs = dynamic_cast<HipPCGSolver*>(s.get());
        CHECK(hs && hs->last_status() == UAAMG_OK);
        std::printf("2d iters %d status %d\n", hs->last_stats().iters, hs->last_status());
        write_f64(out + "/g2_rhs.f64", rhs);
        write_u8(out + "/g2_solid.u8", g.solid);
        write_f64(out + "/g2_p1.f64", g.p);
        // key "cg": the same system with the identity preconditioner
        Grid gc(40, 24, 2.0, 1.2);
        gc.solid = g.solid;
        auto c = Factory::create("cg", Backend::HIP, &o);
        c->solve(gc, rhs, 2000, 1e-10);
        auto* hc = dynamic_cast<HipPCGSolver*>(c.get());
        CHECK(hc && hc->last_status() == UAAMG_OK);
        std::printf("2d cg iters %d status %d\n", hc->last_stats().iters, hc->last_status());
        write_f64(out + "/g2_cg_p.f64", gc.p);
        // keys "pcg_gmg" and "pcg_amg": the same system with the GMG / AMG preconditioners
        for (const char* key : {"pcg_gmg", "pcg_amg"}) {
            Grid gm(40, 24, 2.0, 1.2);
            gm.solid = g.solid;
            auto m = Factory::create(key, Backend::HIP, &o);
            m->solve(gm, rhs, 300, 1e-10);
            auto* hm = dynamic_cast<HipPCGSolver*>(m.get());
            CHECK(hm && hm->last_status() == UAAMG_OK);
            std::printf("2d %s iters %d status %d\n", key, hm->last_stats().iters, hm->last_status());
            write_f64(out + "/g2_" + key + "_p.f64", gm.p);
        }
        // variable Laplacian (grid.h:21-23): x couplings varying with y, y couplings with x
        Grid gv(40, 24, 2.0, 1.2);
        gv.solid = g.solid;
        gv.init_variable_lap();
        for (int j = 0; j < gv.ny + 2; ++j)
            for (int i = 0; i < gv.nx + 2; ++i) {
                const double ox = -(1.0 + 0.5 * std::sin(0.3 * j)) / (gv.dx * gv.dx);
                const double oy = -(1.0 + 0.4 * std::cos(0.2 * i)) / (gv.dy * gv.dy);
                gv.lap_off_x[gv.ip(i, j)] = ox;
                gv.lap_off_y[gv.ip(i, j)] = oy;
                gv.lap_diag[gv.ip(i, j)] = -2.0 * ox - 2.0 * oy;
            }
        auto v = Factory::create("pcg_uaamg", Backend::HIP, &o);
        v->solve(gv, rhs, 500, 1e-10);
        auto* hv = dynamic_cast<HipPCGSolver*>(v.get());
        CHECK(hv && hv->last_status() == UAAMG_OK);
        std::printf("2d variable-Laplacian iters %d status %d\n", hv->last_stats().iters, hv->last_status());
        write_f64(out + "/g2v_diag.f64", gv.lap_diag);
        write_f64(out + "/g2v_ox.f64", gv.lap_off_x);
        write_f64(out + "/g2v_oy.f64", gv.lap_off_y);
        write_f64(out + "/g2v_p.f64", gv.p);
    }
    // projection frames with a moving solid at fixed dims on one solver
    // (the mask changes between frames; cuda_pcg_solver.cu:45-49 would keep the first)
    {
        Grid3D g(20, 16, 12, 1.25, 1.0, 0.75);
        auto s = Factory3D::create("pcg_uaamg", Backend::HIP);
        for (int f = 0; f < 3; ++f) {
            std::fill(g.solid.begin(), g.solid.end(), false);
            for (int k = 1; k <= g.nz; ++k)
                for (int j = 1; j <= g.ny; ++j)
                    for (int i = 1; i <= g.nx; ++i) {
                        double dx = i - (7.0 + 3 * f), dy = j - 8.0, dz = k - 6.0;
                        if (dx * dx + dy * dy + dz * dz < 7.0) g.set_solid(i, j, k);
                    }
            // the previous frame's projected velocity plus a frame-dependent perturbation
            for (size_t q = 0; q < g.u.size(); ++q) g.u[q] += 0.5 * std::sin(0.37 * q + f);
            for (size_t q = 0; q < g.v.size(); ++q) g.v[q] += 0.5 * std::cos(0.23 * q + 2 * f);
            for (size_t q = 0; q < g.w.size(); ++q) g.w[q] += 0.5 * std::sin(0.11 * q - f);
            const std::string pre = out + "/fr" + std::to_string(f);
            write_f64(pre + "_u0.f64", g.u);
            write_f64(pre + "_v0.f64", g.v);
            write_f64(pre + "_w0.f64", g.w);
            write_u8(pre + "_solid.u8", g.solid);
            // the reference's options (FP64, absolute tolerance 1e-8): the solve ends at its
            // p.Ap < 1e-15 breakdown return (pcg.cpp:100), as the reference's does
            PressureProjection3D::project(g, 0.01, *s, 300, 1e-8);
            auto* hs = dynamic_cast<HipPCGSolver3D*>(s.get());
            CHECK(hs && hs->last_stats().iters > 5);
            std::printf("frame %d projection iters %d status %d\n", f, hs->last_stats().iters, hs->last_status());
            write_f64(pre + "_u.f64", g.u);
            write_f64(pre + "_v.f64", g.v);
            write_f64(pre + "_w.f64", g.w);
            write_f64(pre + "_p.f64", g.p);
        }
    }
    // projection, 3D: the grid of the first case with a smooth velocity field
    {
        Grid3D g(24, 16, 12, 1.5, 1.0, 0.75);
        for (int k = 1; k <= g.nz; ++k)
            for (int j = 1; j <= g.ny; ++j)
                for (int i = 1; i <= g.nx; ++i) {
                    double dx = i - 16.0, dy = j - 8.0, dz = k - 6.0;
                    if (dx * dx + dy * dy + dz * dz < 9.0) g.set_solid(i, j, k);
                }
        for (int k = 0; k < g.nz + 2; ++k)
            for (int j = 0; j < g.ny + 2; ++j)
                for (int i = 0; i <= g.nx; ++i) g.u_at(i, j, k) = std::sin(0.4 * i + 0.3 * j) * std::cos(0.2 * k);
        for (int k = 0; k < g.nz + 2; ++k)
            for (int j = 0; j <= g.ny; ++j)
                for (int i = 0; i < g.nx + 2; ++i) g.v_at(i, j, k) = std::cos(0.25 * i) * std::sin(0.5 * j + 0.1 * k);
        for (int k = 0; k <= g.nz; ++k)
            for (int j = 0; j < g.ny + 2; ++j)
                for (int i = 0; i < g.nx + 2; ++i) g.w_at(i, j, k) = std::sin(0.15 * i * k + 0.2 * j);
        write_f64(out + "/pj_u0.f64", g.u);
        write_f64(out + "/pj_v0.f64", g.v);
        write_f64(out + "/pj_w0.f64", g.w);
        write_u8(out + "/pj_solid.u8", g.solid);
        auto s = Factory3D::create("pcg_uaamg", Backend::HIP);   // reference options: FP64, absolute tol
        PressureProjection3D::project(g, 0.01, *s, 300, 1e-8);
        auto* hs = dynamic_cast<HipPCGSolver3D*>(s.get());
        std::printf("projection iters %d status %d\n", hs->last_stats().iters, hs->last_status());
        write_f64(out + "/pj_u.f64", g.u);
        write_f64(out + "/pj_v.f64", g.v);
        write_f64(out + "/pj_w.f64", g.w);
        write_f64(out + "/pj_p.f64", g.p);
    }
    // standalone amgpcg, N = 64, b = sin(pi x) sin(pi y) (test_amgpcg.cpp:29-70)
    {
        amg::Grid b(64);
        for (int i = 1; i <= b.N; ++i)
            for (int j = 1; j <= b.N; ++j)
                b(i, j) = std::sin(M_PI * i * b.h) * std::sin(M_PI * j * b.h);
        int iters = 0;
        amg::Grid x = amg::amgpcg_solve(b, 100, 1e-10, iters);
        std::printf("amgpcg iters %d\n", iters);
        write_f64(out + "/amg_b.f64", b.v);
        write_f64(out + "/amg_x.f64", x.v);
        std::vector<double> it(1, (double)iters);
        write_f64(out + "/amg_iters.f64", it);
    }
    std::printf("solve %s\n", fails ? "FAILED" : "ok");
    return fails ? 1 : 0;
}

template <typename T>
static bool read_raw(const std::string& path, std::vector<T>& out, size_t n) {
    FILE* f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    out.resize(n);
    const size_t got = std::fread(out.data(), sizeof(T), n, f);
    std::fclose(f);
    return got == n;
}

// VtkWriter[3D]::write of a grid read from DIR (spec.txt: dim nx ny nz lx ly
// lz frame; u.f64, v.f64, w.f64 MAC arrays; solid.u8 padded mask) into DIR/ours
static int run_vtk(const std::string& dir) {
    FILE* f = std::fopen((dir + "/spec.txt").c_str(), "r");
    if (!f) return 2;
    int dim, nx, ny, nz, frame;
    double lx, ly, lz;
    const int got = std::fscanf(f, "%d %d %d %d %lf %lf %lf %d", &dim, &nx, &ny, &nz, &lx, &ly, &lz, &frame);
    std::fclose(f);
    if (got != 8) return 2;
    std::vector<uint8_t> s;
    bool ok;
    if (dim == 2) {
        Grid g(nx, ny, lx, ly);
        ok = read_raw(dir + "/u.f64", g.u, g.u.size()) && read_raw(dir + "/v.f64", g.v, g.v.size()) &&
             read_raw(dir + "/solid.u8", s, g.solid.size());
        if (!ok) return 2;
        for (size_t c = 0; c < s.size(); c++) g.solid[c] = s[c] != 0;
        return VtkWriter::write(g, frame, dir + "/ours") ? 0 : 3;
    }
    Grid3D g(nx, ny, nz, lx, ly, lz);
    ok = read_raw(dir + "/u.f64", g.u, g.u.size()) && read_raw(dir + "/v.f64", g.v, g.v.size()) &&
         read_raw(dir + "/w.f64", g.w, g.w.size()) && read_raw(dir + "/solid.u8", s, g.solid.size());
    if (!ok) return 2;
    for (size_t c = 0; c < s.size(); c++) g.solid[c] = s[c] != 0;
    return VtkWriter3D::write(g, frame, dir + "/ours") ? 0 : 3;
}

int main(int argc, char** argv) {
    if (argc >= 2 && std::strcmp(argv[1], "keys") == 0) return run_keys();
    if (argc >= 3 && std::strcmp(argv[1], "vtk") == 0) return run_vtk(argv[2]);
    if (argc >= 3 && std::strcmp(argv[1], "solve") == 0) {
        try {
            return run_solve(argv[2]);
        } catch (const std::exception& e) {
            std::fprintf(stderr, "exception: %s\n", e.what());
            return 2;
        }
    }
    std::fprintf(stderr, "usage: %s keys | solve OUTDIR | vtk DIR\n", argv[0]);
    return 64;
}
