"""GPU: error and edge paths through the C-ABI -- a non-finite rhs, an
all-solid mask, a mask that changes at fixed dims on one context (with a
captured graph), sizes that do not match the grid."""
import numpy as np
import pytest

import cases
import oracle_bind as ob
import uaamg_amd as U

pytestmark = pytest.mark.gpu

CASES = {c[0]: c[1] for c in cases.parity_cases()}


def _ctx(pb, **opts):
    c = U.Context(**opts)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    return c


@pytest.mark.parametrize("prec", [U.FP64, U.FP32, U.MIXED_BF16, U.MIXED_F64])
@pytest.mark.parametrize("bad", [np.nan, np.inf], ids=["nan", "inf"])
def test_nonfinite_rhs_is_reported(gpu, prec, bad):
    """A NaN or Inf in the rhs ends the solve with UAAMG_NONFINITE (the
    reference would run on with NaN scalars): the device CG scalars carry
    the non-finite value to the first test of p.Ap or r.r."""
    pb = CASES["3d_24x20x18_aniso_solids"]
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(1))
    m = np.flatnonzero(pb.fluid_interior_mask())
    rhs[m[len(m) // 3]] = bad
    c = _ctx(pb, precision=prec, tol_mode=U.TOL_REL, use_graph=1)
    p, st = c.solve(rhs, 100, 1e-6)
    c.close()
    assert st.status == U.NONFINITE, U.status_name(st.status)
    assert st.iters <= 1


@pytest.mark.parametrize("dim", [2, 3])
def test_all_solid_mask(gpu, dim):
    """Every interior cell solid: no fluid, so the projected rhs, r and p are
    all zero; the first p.Ap (= 0 < 1e-15) ends the solve as the reference's
    breakdown return (pcg.cpp:100) and p is left as it was."""
    n = (20, 16) if dim == 2 else (12, 10, 8)
    pb = ob.Problem(n)
    pb.interior(pb.solid)[...] = 1
    rhs = np.random.default_rng(2).standard_normal(pb.N)
    p0 = np.random.default_rng(3).standard_normal(pb.N)
    for prec in (U.FP64, U.FP32):
        c = _ctx(pb, precision=prec)
        p, st = c.solve(rhs, 50, 1e-8, p0=p0)
        c.close()
        assert st.status == U.BREAKDOWN and st.iters == 1, (prec, U.status_name(st.status), st.iters)
        np.testing.assert_array_equal(p, p0)


@pytest.mark.parametrize("prec", [U.FP64, U.FP32])
def test_mask_change_at_fixed_dims_with_graph(gpu, prec):
    """One context, the same dims, a different solid: the second solve (after
    a first one that captured its iterations into a hipGraph) uses the new
    mask, its coarse masks and Galerkin couplings, and equals the oracle's
    solve on the new mask (the reference wrapper kept the old one,
    cuda_pcg_solver.cu:45-49)."""
    base = CASES["3d_40x36x34_walls_solids"]
    pa = ob.Problem(base.n, base.L, base.solid, base.wall)
    pb = ob.Problem(base.n, base.L, cases.balls(base, [(0.3, 0.6, 0.5)], [0.15]), base.wall)
    assert not np.array_equal(pa.solid, pb.solid)
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(4))
    c = U.Context(precision=prec, tol_mode=U.TOL_REL, use_graph=1)
    c.set_grid(pa.n, pa.h, pa.solid, pa.wall)
    c.solve(1e3 * cases.random_rhs(pa, np.random.default_rng(5)), 500, 1e-6)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    p, st = c.solve(rhs, 500, 1e-8 if prec == U.FP64 else 1e-6)
    for lvl in range(len(c.hierarchy())):
        g, o = c.level_data(lvl), ob.level_data(pb, lvl)
        assert np.array_equal(g["cx"], o["cx"]) or prec != U.FP64, lvl
    c.close()
    po, res = ob.pcg(pb, rhs, 500, 1e-8, tol_mode=1)
    assert st.status == U.OK
    m = pb.fluid_interior_mask()
    a, b = p[m] - p[m].mean(), po[m] - po[m].mean()
    if prec == U.FP64:
        assert abs(st.iters - res.iters) <= 1, (st.iters, res.iters)
        assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()
    else:
        assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max()


def test_sizes_are_checked_before_the_call(gpu):
    """Arrays that do not match the grid raise instead of reading or writing
    past their end."""
    pb = CASES["3d_24x20x18_aniso_solids"]
    c = _ctx(pb, precision=U.FP32)
    with pytest.raises(ValueError):
        c.solve(np.zeros(pb.N - 1))
    with pytest.raises(ValueError):
        c.solve(np.zeros(pb.N), p0=np.zeros(pb.N + 5))
    with pytest.raises(ValueError):
        c.apply(np.zeros(pb.N // 2))
    nu, nv, nw = c.mac_sizes()
    with pytest.raises(ValueError):
        c.project(np.zeros(nu), np.zeros(nv - 1), np.zeros(nw), 1.0)
    with pytest.raises(ValueError):
        c.project(np.zeros(nu), np.zeros(nv), None, 1.0)
    # a [k][j][i] array held in Fortran memory order is read in the padded
    # layout's logical order (copied to C order), not as raw memory
    r = 1e3 * cases.random_rhs(pb, np.random.default_rng(6))
    p1, s1 = c.solve(r, 200, 1e-6)
    rf = np.asfortranarray(r.reshape(pb.shape[::-1]))
    assert not rf.flags.c_contiguous
    p2, s2 = c.solve(rf, 200, 1e-6)
    c.close()
    np.testing.assert_array_equal(p1, p2)
