"""GPU: the C++ host mirror (include/uaamg_solver.hpp) driven like the
reference application drives its Solver plugins, checked against the oracle.
The program runs as a child process; it links the HIP library through the
C-ABI only."""
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TOOL = os.path.join(ROOT, "build", "host_solver_test")

pytestmark = pytest.mark.gpu


def _rd(d, name, dt=np.float64):
    return np.fromfile(os.path.join(d, name), dtype=dt)


@pytest.fixture(scope="module")
def host_run(tmp_path_factory, gpu):
    if not os.path.exists(HOST_TOOL):
        pytest.fail("build/host_solver_test not built (run __graft_entry__.build())")
    d = str(tmp_path_factory.mktemp("host"))
    print(f"host mirror: {HOST_TOOL} solve {d}", flush=True)
    r = subprocess.run([HOST_TOOL, "solve", d], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    return d, r.stdout


def _fluid_err(pb, p, po):
    m = pb.fluid_interior_mask()
    a, b = p[m] - p[m].mean(), po[m] - po[m].mean()
    return float(np.abs(a - b).max() / np.abs(b).max())


def test_host_mirror_3d_matches_oracle(host_run):
    d, _ = host_run
    pb = ob.Problem((24, 16, 12), (1.5, 1.0, 0.75), solid=_rd(d, "g3_solid.u8", np.uint8))
    rhs = _rd(d, "g3_rhs.f64")
    po, res = ob.pcg(pb, rhs, 300, 1e-10, tol_mode=1)
    assert res.status == 0
    p1 = _rd(d, "g3_p1.f64")
    assert _fluid_err(pb, p1, po) <= 1e-9
    # solid and ghost cells untouched (pcg.cpp:107 writes fluid interior only)
    assert np.all(p1[~pb.fluid_interior_mask()] == 0.0)
    # a second solve accumulates: p2 = p1 + x, x solving the same system
    p2 = _rd(d, "g3_p2.f64")
    assert _fluid_err(pb, p2, 2 * po) <= 1e-9


def test_host_mirror_2d_matches_oracle(host_run):
    d, _ = host_run
    pb = ob.Problem((40, 24), (2.0, 1.2), solid=_rd(d, "g2_solid.u8", np.uint8))
    rhs = _rd(d, "g2_rhs.f64")
    po, res = ob.pcg(pb, rhs, 300, 1e-10, tol_mode=1)
    assert res.status == 0
    assert _fluid_err(pb, _rd(d, "g2_p1.f64"), po) <= 1e-9


def test_host_mirror_amgpcg_matches_oracle(host_run):
    d, _ = host_run
    b = _rd(d, "amg_b.f64")
    xo, it = ob.amgpcg(64, b, 100, 1e-10)
    x = _rd(d, "amg_x.f64")
    assert int(_rd(d, "amg_iters.f64")[0]) == it
    assert np.abs(x - xo).max() <= 1e-10 * np.abs(xo).max()


def test_host_mirror_projection_matches_oracle(host_run):
    """PressureProjection3D::project through the host mirror (FP64, the
    reference's options) equals the oracle's projection."""
    d, _ = host_run
    pb = ob.Problem((24, 16, 12), (1.5, 1.0, 0.75), solid=_rd(d, "pj_solid.u8", np.uint8))
    u0, v0, w0 = _rd(d, "pj_u0.f64"), _rd(d, "pj_v0.f64"), _rd(d, "pj_w0.f64")
    uo, vo, wo, po, res = ob.project(pb, u0, v0, w0, 0.01, 300, 1e-8)
    assert np.abs(_rd(d, "pj_p.f64") - po).max() <= 1e-9 * np.abs(po).max()
    vmax = max(np.abs(u0).max(), np.abs(v0).max(), np.abs(w0).max())
    for name, ref in (("pj_u.f64", uo), ("pj_v.f64", vo), ("pj_w.f64", wo)):
        assert np.abs(_rd(d, name) - ref).max() <= 1e-9 * vmax


def test_host_mirror_cg_matches_oracle(host_run):
    """Factory::create("cg"): PCG with the identity preconditioner
    (factory.cpp:19) equals the oracle's plain CG."""
    d, _ = host_run
    pb = ob.Problem((40, 24), (2.0, 1.2), solid=_rd(d, "g2_solid.u8", np.uint8))
    rhs = _rd(d, "g2_rhs.f64")
    po, res = ob.pcg(pb, rhs, 2000, 1e-10, tol_mode=1, precond=ob.PRECOND_IDENTITY)
    assert res.status == 0
    assert _fluid_err(pb, _rd(d, "g2_cg_p.f64"), po) <= 1e-9


@pytest.mark.parametrize("key,precond", [("pcg_gmg", ob.PRECOND_GMG), ("pcg_amg", ob.PRECOND_AMG)])
def test_host_mirror_gmg_amg_match_oracle(host_run, key, precond):
    """Factory::create("pcg_gmg") / ("pcg_amg"): PCG with the GMG / AMG
    preconditioner (factory.cpp:21-22) equals the oracle's solve."""
    d, _ = host_run
    pb = ob.Problem((40, 24), (2.0, 1.2), solid=_rd(d, "g2_solid.u8", np.uint8))
    rhs = _rd(d, "g2_rhs.f64")
    po, res = ob.pcg(pb, rhs, 300, 1e-10, tol_mode=1, precond=precond)
    assert res.status == 0
    assert _fluid_err(pb, _rd(d, f"g2_{key}_p.f64"), po) <= 1e-9


def test_host_mirror_variable_lap_matches_oracle(host_run):
    """A Grid with a variable Laplacian (init_variable_lap) solved through
    the factory's pcg_uaamg equals the oracle's PCG on the same coefficients."""
    d, _ = host_run
    lap = (_rd(d, "g2v_diag.f64"), _rd(d, "g2v_ox.f64"), _rd(d, "g2v_oy.f64"))
    pb = ob.Problem((40, 24), (2.0, 1.2), solid=_rd(d, "g2_solid.u8", np.uint8), lap=lap)
    rhs = _rd(d, "g2_rhs.f64")
    po, res = ob.pcg(pb, rhs, 500, 1e-10, tol_mode=1)
    assert res.status == 0
    assert _fluid_err(pb, _rd(d, "g2v_p.f64"), po) <= 1e-9


def test_host_mirror_projection_frames_moving_solid(host_run):
    """Three PressureProjection3D::project frames on one solver, the solid
    ball moving between frames at fixed dims: every frame equals the
    oracle's projection with that frame's mask (a stale mask would not)."""
    d, _ = host_run
    masks = []
    for f in range(3):
        solid = _rd(d, f"fr{f}_solid.u8", np.uint8)
        masks.append(solid)
        pb = ob.Problem((20, 16, 12), (1.25, 1.0, 0.75), solid=solid)
        u0, v0, w0 = (_rd(d, f"fr{f}_{c}0.f64") for c in "uvw")
        uo, vo, wo, po, res = ob.project(pb, u0, v0, w0, 0.01, 300, 1e-8)
        assert np.abs(_rd(d, f"fr{f}_p.f64") - po).max() <= 1e-9 * np.abs(po).max(), f
        vmax = max(np.abs(u0).max(), np.abs(v0).max(), np.abs(w0).max())
        for c, ref in zip("uvw", (uo, vo, wo)):
            assert np.abs(_rd(d, f"fr{f}_{c}.f64") - ref).max() <= 1e-9 * vmax, (f, c)
    assert not np.array_equal(masks[0], masks[1]) and not np.array_equal(masks[1], masks[2])
