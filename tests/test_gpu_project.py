"""GPU: the pressure projection around the solve (PressureProjection[3D]::project,
pressure.cpp:5-32, pressure_3d.cpp:5-39) through the C-ABI against the oracle."""
import numpy as np
import pytest

import oracle_bind as ob
import cases
import uaamg_amd as U

pytestmark = pytest.mark.gpu

NAMES = ("2d_32x16", "2d_64x16_cylinder", "2d_33x17_odd", "2d_48x40_aniso_solids", "2d_40x30_dirichlet",
         "3d_16_sphere", "3d_17x9x5_odd", "3d_24x20x18_aniso_solids", "3d_40x36x34_walls_solids",
         "3d_200x70x40_walls_solids")
PROJ = [c for c in cases.parity_cases() if c[0] in NAMES]


def ctx_for(pb, **opts):
    c = U.Context(**opts)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    return c


@pytest.mark.parametrize("name,pb", PROJ, ids=[c[0] for c in PROJ])
def test_project_fp64_matches_oracle(gpu, name, pb):
    """FP64: the same iterations as the oracle's projection, p within 1e-9 of
    max|p|, and the corrected velocities within 1e-9 of max|u|."""
    rng = np.random.default_rng(21)
    u, v, w = cases.mac_fields(pb, rng, closed=False)
    dt = 0.01
    uo, vo, wo, po, res = ob.project(pb, u, v, w, dt, 300, 1e-8)
    c = ctx_for(pb, precision=U.FP64)
    ug, vg, wg, pg, st = c.project(u, v, w, dt, 300, 1e-8)
    c.close()
    assert st.status in (U.OK, U.BREAKDOWN) and abs(st.iters - res.iters) <= 1, (st.iters, res.iters)
    assert np.abs(pg - po).max() <= 1e-9 * np.abs(po).max()
    vmax = max(np.abs(u).max(), np.abs(v).max())
    assert np.abs(ug - uo).max() <= 1e-9 * vmax
    assert np.abs(vg - vo).max() <= 1e-9 * vmax
    if pb.dim == 3:
        assert np.abs(wg - wo).max() <= 1e-9 * vmax


@pytest.mark.parametrize("name,pb", [c for c in PROJ if "dirichlet" not in c[0] and "walls" not in c[0]],
                         ids=[c[0] for c in PROJ if "dirichlet" not in c[0] and "walls" not in c[0]])
def test_project_fp32_divergence_free(gpu, name, pb):
    """FP32 at relres 1e-6 on a closed field: the divergence left on fluid
    cells is below 1e-4 of the initial one (relres bounds the 2-norm of the
    residual; its largest entry sits higher: 1.1e-5 measured on the sphere)."""
    rng = np.random.default_rng(22)
    u, v, w = cases.mac_fields(pb, rng, closed=True)
    dt = 0.02
    c = ctx_for(pb, precision=U.FP32, tol_mode=U.TOL_REL)
    ug, vg, wg, pg, st = c.project(u, v, w, dt, 300, 1e-6)
    c.close()
    assert st.status == U.OK, U.status_name(st.status)
    fl = pb.fluid_interior_mask()
    d0 = cases.mac_divergence(pb, u, v, w)
    d1 = cases.mac_divergence(pb, ug, vg, wg)
    assert np.abs(d1[fl]).max() <= 1e-4 * np.abs(d0[fl]).max()


def test_project_device_equals_host(gpu):
    """The device entry point on torch buffers gives the host entry point's
    result bit for bit (3D, FP32)."""
    import torch
    name, pb = [c for c in PROJ if c[0] == "3d_24x20x18_aniso_solids"][0]
    rng = np.random.default_rng(23)
    u, v, w = cases.mac_fields(pb, rng, closed=False)
    dt = 0.05
    c = ctx_for(pb, precision=U.FP32, tol_mode=U.TOL_REL)
    uh, vh, wh, ph, st_h = c.project(u, v, w, dt, 300, 1e-6)
    dev = torch.device("cuda:0")
    ud, vd, wd = (torch.from_numpy(a.copy()).to(dev) for a in (u, v, w))
    pd = torch.full((pb.N,), 7.0, dtype=torch.float64, device=dev)   # overwritten with 0 before the solve
    st_d = c.project_device(ud.data_ptr(), vd.data_ptr(), wd.data_ptr(), pd.data_ptr(), dt, 300, 1e-6)
    torch.cuda.synchronize()
    c.close()
    assert st_d.iters == st_h.iters
    np.testing.assert_array_equal(pd.cpu().numpy(), ph)
    np.testing.assert_array_equal(ud.cpu().numpy(), uh)
    np.testing.assert_array_equal(vd.cpu().numpy(), vh)
    np.testing.assert_array_equal(wd.cpu().numpy(), wh)


def test_project_bad_arguments(gpu):
    name, pb = PROJ[0]
    c = ctx_for(pb, precision=U.FP64)
    u, v, w = cases.mac_fields(pb, np.random.default_rng(0))
    with pytest.raises(RuntimeError):
        c.project(u, v, w, 0.0, 10, 1e-6)
    c.close()
