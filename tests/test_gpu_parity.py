"""GPU parity: the HIP path through the C-ABI against the oracle on the same inputs."""
import numpy as np
import pytest

import oracle_bind as ob
import cases
import uaamg_amd as U

pytestmark = pytest.mark.gpu

CASES = cases.parity_cases()


def ctx_for(pb, **opts):
    c = U.Context(**opts)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    return c


@pytest.mark.parametrize("name,pb", CASES, ids=[c[0] for c in CASES])
def test_hierarchy_and_coefficients(gpu, name, pb):
    """Coarse masks and Galerkin couplings equal the oracle's, bit for bit."""
    c = ctx_for(pb)
    assert c.hierarchy() == ob.hierarchy_dims(pb)
    for lvl in range(len(c.hierarchy())):
        g = c.level_data(lvl)
        o = ob.level_data(pb, lvl)
        mask = np.zeros(g["solid"].size, bool)
        dims = o["dims"]
        sub = type("P", (), {})()
        # interior comparison of the mask; couplings everywhere (ghosts are zero on both sides)
        view = g["solid"].reshape([d + 2 for d in dims][::-1])
        oview = o["solid"].reshape([d + 2 for d in dims][::-1])
        sl = tuple([slice(1, -1)] * len(dims))
        assert np.array_equal(view[sl], oview[sl]), f"level {lvl} mask"
        for ch in ("cx", "cy", "diag", "excess") + (("cz",) if pb.dim == 3 else ()):
            assert np.array_equal(g[ch], o[ch]), f"level {lvl} {ch}"
    c.close()


@pytest.mark.parametrize("name,pb", CASES, ids=[c[0] for c in CASES])
def test_matvec_dot(gpu, name, pb):
    rng = np.random.default_rng(7)
    v = cases.random_rhs(pb, rng)
    c = ctx_for(pb)
    Av, vAv = c.matvec(v)
    Ao = ob.matvec(pb, v)
    m = pb.fluid_interior_mask()
    assert np.array_equal(Av[m], Ao[m])
    ref = float(np.sum(v[m] * Ao[m]))
    assert abs(vAv - ref) <= 1e-12 * abs(ref)
    c.close()


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("name,pb", CASES, ids=[c[0] for c in CASES])
def test_vcycle_fp64(gpu, name, pb, fused):
    """One V-cycle, FP64: the reference's GPU tests allow 1e-12*max(1,|z|)
    (test_cuda_uaamg.cu:56); the HIP path keeps the reference's operation order."""
    rng = np.random.default_rng(11)
    r = cases.random_rhs(pb, rng)
    c = ctx_for(pb, precision=U.FP64, fused=fused)
    z = c.apply(r)
    zo = ob.uaamg_apply(pb, r)
    m = np.zeros(pb.N, bool); pb.interior(m)[...] = True
    err = np.abs(z[m] - zo[m]).max()
    assert err <= 1e-12 * max(1.0, np.abs(zo[m]).max()), err
    c.close()


@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("name,pb", CASES, ids=[c[0] for c in CASES])
def test_vcycle_fp32(gpu, name, pb, fused):
    rng = np.random.default_rng(11)
    r = cases.random_rhs(pb, rng)
    c = ctx_for(pb, precision=U.FP32, fused=fused)
    z = c.apply(r)
    zo = ob.uaamg_apply(pb, r)
    m = np.zeros(pb.N, bool); pb.interior(m)[...] = True
    err = np.abs(z[m] - zo[m]).max() / np.abs(zo[m]).max()
    assert err <= 2e-5, err
    c.close()


def _demean(p, m):
    q = p[m]
    return q - q.mean()


# PCG needs a connected fluid (one null-space constant): the porous case has
# isolated fluid cells, so only its hierarchy, matvec and V-cycle are compared.
PCG_CASES = [c for c in CASES if "porous" not in c[0]]


@pytest.mark.parametrize("name,pb", PCG_CASES, ids=[c[0] for c in PCG_CASES])
def test_pcg_fp64(gpu, name, pb):
    """PCG FP64 vs the oracle at relres 1e-8: iterations within +-1 and
    pressure (mean removed) within 1e-9 of max|p|."""
    rng = np.random.default_rng(3)
    rhs = 1e3 * cases.random_rhs(pb, rng)   # div(u)/dt magnitudes: keeps pAp far above the 1e-15 guard
    po, res = ob.pcg(pb, rhs, 500, 1e-8, tol_mode=1)
    c = ctx_for(pb, precision=U.FP64, tol_mode=U.TOL_REL)
    p, st = c.solve(rhs, 500, 1e-8)
    assert res.status == 0 and st.status == 0
    assert abs(st.iters - res.iters) <= 1, (st.iters, res.iters)
    m = pb.fluid_interior_mask()
    a, b = _demean(p, m), _demean(po, m)
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()
    c.close()


@pytest.mark.parametrize("name,pb", PCG_CASES, ids=[c[0] for c in PCG_CASES])
def test_pcg_fp32(gpu, name, pb):
    """Plain FP32 form (FP32 vectors/levels, FP64 reductions) at relres 1e-6."""
    rng = np.random.default_rng(3)
    rhs = 1e3 * cases.random_rhs(pb, rng)
    po, res = ob.pcg(pb, rhs, 500, 1e-6, tol_mode=1)
    c = ctx_for(pb, precision=U.FP32, tol_mode=U.TOL_REL)
    p, st = c.solve(rhs, 500, 1e-6)
    assert st.status == 0
    # FP32 vectors/levels perturb the Krylov sequence: allow +-2 or 15 %
    print(f"{name}: fp32 {st.iters} iterations, oracle {res.iters}", flush=True)
    assert abs(st.iters - res.iters) <= max(2, int(0.15 * res.iters)), (st.iters, res.iters)
    m = pb.fluid_interior_mask()
    a, b = _demean(p, m), _demean(po, m)
    assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max()
    c.close()


@pytest.mark.parametrize("N", [32, 33, 64, 100])
@pytest.mark.parametrize("prec", [U.FP64, U.FP32])
def test_amgpcg(gpu, N, prec):
    rng = np.random.default_rng(N)
    b = np.zeros((N + 2) ** 2)
    b.reshape(N + 2, N + 2)[1:-1, 1:-1] = rng.standard_normal((N, N))
    xo, io = ob.amgpcg(N, b, 60, 1e-6)
    x, it = gpu.amgpcg(N, b, 60, 1e-6, prec)
    tol = 1e-10 if prec == U.FP64 else 1e-4
    assert abs(it - io) <= (0 if prec == U.FP64 else 1), (it, io)
    assert np.abs(x - xo).max() <= tol * np.abs(xo).max()


# ---- FP64 CG with FP32 coarse levels (CudaPCG3D::solve_mixed, cuda_pcg_3d_opt.cu:319-330) ----

@pytest.mark.parametrize("fused", [0, 1])
@pytest.mark.parametrize("name,pb", CASES, ids=[c[0] for c in CASES])
def test_vcycle_mixed_f64(gpu, name, pb, fused):
    """FP64 level-0 smoothing, FP32 coarse levels: the V-cycle is within the
    FP32 form's bar of the FP64 one."""
    rng = np.random.default_rng(11)
    r = cases.random_rhs(pb, rng)
    c = ctx_for(pb, precision=U.MIXED_F64, fused=fused)
    z = c.apply(r)
    zo = ob.uaamg_apply(pb, r)
    m = np.zeros(pb.N, bool); pb.interior(m)[...] = True
    err = np.abs(z[m] - zo[m]).max() / np.abs(zo[m]).max()
    assert err <= 2e-5, err
    c.close()


@pytest.mark.parametrize("name,pb", PCG_CASES, ids=[c[0] for c in PCG_CASES])
def test_pcg_mixed_f64(gpu, name, pb):
    """FP64 CG vectors and scalars with FP32 coarse levels: at relres 1e-8
    within 2 iterations (or 5 %) of the oracle (measured: +2 at most on these
    grids; the FP32 levels perturb the preconditioner, the FP64 Krylov
    vectors do not), pressure (mean removed) within 1e-7 of max|p|."""
    rng = np.random.default_rng(3)
    rhs = 1e3 * cases.random_rhs(pb, rng)
    po, res = ob.pcg(pb, rhs, 500, 1e-8, tol_mode=1)
    c = ctx_for(pb, precision=U.MIXED_F64, tol_mode=U.TOL_REL)
    p, st = c.solve(rhs, 500, 1e-8)
    assert res.status == 0 and st.status == 0, U.status_name(st.status)
    assert -1 <= st.iters - res.iters <= max(2, res.iters // 20), (st.iters, res.iters)
    m = pb.fluid_interior_mask()
    a, b = _demean(p, m), _demean(po, m)
    assert np.abs(a - b).max() <= 1e-7 * np.abs(b).max()
    c.close()


# ---- mixed form: BF16 level storage, FP32 CG, FP64 reductions ----

@pytest.mark.parametrize("prec", [U.MIXED_BF16, U.MIXED_BF16R], ids=["bf16", "bf16r"])
@pytest.mark.parametrize("name,pb", CASES, ids=[c[0] for c in CASES])
def test_vcycle_bf16_levels(gpu, name, pb, prec):
    """BF16 coarse levels (8-bit mantissa couplings, b, x; or the couplings
    only, with FP32 b and x) perturb the preconditioner at the 1e-2 level: it
    stays close to the FP64 V-cycle."""
    rng = np.random.default_rng(11)
    r = cases.random_rhs(pb, rng)
    c = ctx_for(pb, precision=prec)
    z = c.apply(r)
    zo = ob.uaamg_apply(pb, r)
    m = np.zeros(pb.N, bool); pb.interior(m)[...] = True
    err = np.abs(z[m] - zo[m]).max() / np.abs(zo[m]).max()
    assert err <= 5e-2, err
    c.close()


@pytest.mark.parametrize("prec,flexible", [(U.MIXED_BF16, 0), (U.MIXED_BF16, 1), (U.MIXED_BF16R, 0)],
                         ids=["bf16", "bf16_flexible", "bf16r"])
@pytest.mark.parametrize("name,pb", PCG_CASES, ids=[c[0] for c in PCG_CASES])
def test_pcg_bf16_levels(gpu, name, pb, prec, flexible):
    """The CG residual, not the preconditioner, sets the accuracy: at relres
    1e-6 the pressure is within 2e-5 of the FP64 oracle (mean removed), the
    bar of the plain FP32 form; the iteration count may grow by the
    preconditioner's perturbation."""
    rng = np.random.default_rng(3)
    rhs = 1e3 * cases.random_rhs(pb, rng)
    po, res = ob.pcg(pb, rhs, 500, 1e-10, tol_mode=1)
    c = ctx_for(pb, precision=prec, tol_mode=U.TOL_REL, flexible=flexible)
    p, st = c.solve(rhs, 500, 1e-6)
    assert st.status == 0, U.status_name(st.status)
    assert st.iters <= 1.5 * res.iters + 5, (st.iters, res.iters)
    m = pb.fluid_interior_mask()
    a, b = _demean(p, m), _demean(po, m)
    assert np.abs(a - b).max() <= 2e-5 * np.abs(b).max()
    c.close()


# ---- hipGraph replay of the CG iterations ----

@pytest.mark.parametrize("prec", [U.FP64, U.FP32, U.MIXED_BF16, U.MIXED_BF16R, U.MIXED_F64])
@pytest.mark.parametrize("name,pb", [c for c in PCG_CASES if c[0] in ("3d_24x20x18_aniso_solids", "2d_48x40_aniso_solids", "3d_40x36x34_walls_solids")],
                         ids=lambda v: v if isinstance(v, str) else "")
def test_graph_replay_matches_eager(gpu, name, pb, prec):
    """use_graph=1 replays captured 4-iteration chunks: every solve (several
    replays each, with eager work and synchronisation in between) gives the
    eager result bit for bit."""
    rng = np.random.default_rng(5)
    rhs = 1e3 * cases.random_rhs(pb, rng)
    eager = ctx_for(pb, precision=prec, tol_mode=U.TOL_REL, use_graph=0)
    p0, st0 = eager.solve(rhs, 400, 1e-6)
    graph = ctx_for(pb, precision=prec, tol_mode=U.TOL_REL, use_graph=1)
    for rep in range(3):
        p1, st1 = graph.solve(rhs, 400, 1e-6)
        assert st1.status == st0.status and st1.iters == st0.iters, (rep, st1.iters, st0.iters)
        assert np.array_equal(p1, p0), rep
        pe, _ = eager.solve(rhs, 400, 1e-6)       # eager work between the replays
        assert np.array_equal(pe, p0)
    eager.close(); graph.close()


# ---- factory key "cg": PCG with the identity preconditioner (factory.cpp:19, factory_3d.cpp:19) ----

@pytest.mark.parametrize("prec", [U.FP64, U.FP32], ids=["fp64", "fp32"])
@pytest.mark.parametrize("name,pb", PCG_CASES, ids=[c[0] for c in PCG_CASES])
def test_cg_identity_preconditioner(gpu, name, pb, prec):
    """Plain CG (z = r) against the oracle's identity-preconditioned PCG:
    FP64 within +-1 iteration of the oracle at relres 1e-8 and within 1e-9
    of its pressure at relres 1e-10 (at 1e-8 the hundreds of CG iterations,
    their dot products summed in different orders, left 2e-9 between the two
    stopping points); FP32 at relres 1e-6 within 2e-5."""
    rng = np.random.default_rng(3)
    rhs = 1e3 * cases.random_rhs(pb, rng)
    po8, res8 = ob.pcg(pb, rhs, 20000, 1e-8, tol_mode=1, precond=ob.PRECOND_IDENTITY)
    po, res = ob.pcg(pb, rhs, 20000, 1e-10, tol_mode=1, precond=ob.PRECOND_IDENTITY)
    assert res8.status == 0 and res.status == 0
    c = ctx_for(pb, precision=prec, tol_mode=U.TOL_REL, preconditioner=U.PRECOND_NONE)
    if prec == U.FP64:
        _, st8 = c.solve(rhs, 20000, 1e-8)
        assert abs(st8.iters - res8.iters) <= 1, (st8.iters, res8.iters)
    p, st = c.solve(rhs, 20000, 1e-10 if prec == U.FP64 else 1e-6)
    c.close()
    assert st.status == 0, U.status_name(st.status)
    assert not st.forms_used & (U.FORM_ROWS | U.FORM_TAIL | U.FORM_PMV)
    m = pb.fluid_interior_mask()
    a, b = _demean(p, m), _demean(po, m)
    assert np.abs(a - b).max() <= (1e-9 if prec == U.FP64 else 2e-5) * np.abs(b).max()


# ---- variable Laplacian (grid.h:21-23, pcg.cpp:36-46), 2D ----

VAR_CASES = [c for c in PCG_CASES if c[1].dim == 2]


@pytest.mark.parametrize("precond", [ob.PRECOND_UAAMG, ob.PRECOND_IDENTITY], ids=["uaamg", "cg"])
@pytest.mark.parametrize("name,pb", VAR_CASES, ids=[c[0] for c in VAR_CASES])
def test_variable_lap_pcg(gpu, name, pb, precond):
    """Symmetric variable coefficients (x couplings varying with y, y
    couplings with x): FP64 at relres 1e-10 within +-1 iteration and 1e-9
    of the oracle; the matvec alone equals the oracle's bit for bit."""
    lap = cases.separable_lap(pb)
    vpb = ob.Problem(pb.n, pb.L, pb.solid, pb.wall, lap=lap)
    rng = np.random.default_rng(3)
    rhs = 1e3 * cases.random_rhs(pb, rng)
    po, res = ob.pcg(vpb, rhs, 5000, 1e-10, tol_mode=1, precond=precond)
    assert res.status == 0
    c = ctx_for(pb, precision=U.FP64, tol_mode=U.TOL_REL, preconditioner=precond)
    c.set_variable_lap(lap)
    v = cases.random_rhs(pb, np.random.default_rng(7))
    Av, _ = c.matvec(v)
    m = pb.fluid_interior_mask()
    assert np.array_equal(Av[m], ob.matvec(vpb, v)[m])
    p, st = c.solve(rhs, 5000, 1e-10)
    assert st.status == 0
    assert abs(st.iters - res.iters) <= 1, (st.iters, res.iters)
    a, b = _demean(p, m), _demean(po, m)
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()
    # back to the uniform operator
    c.set_variable_lap(None)
    p2, st2 = c.solve(rhs, 5000, 1e-8)
    pu, ru = ob.pcg(pb, rhs, 5000, 1e-8, tol_mode=1, precond=precond)
    assert abs(st2.iters - ru.iters) <= 1
    c.close()


def test_variable_lap_smooth_cylinder_and_warm_start(gpu):
    """SmoothCylinder's coefficients (cylinder_model.cpp:50-71) make a
    non-symmetric operator on which the reference's PCG does not converge:
    ten iterations agree with the oracle's ten.  A warm start through the
    variable operator (r = b - A x0) on the symmetric case stays at the
    converged pressure."""
    from uaamg_amd import workloads as W
    n, L = (64, 48), (1.0, 0.75)
    solid, d, ox, oy = W.smooth_cylinder_lap(n, L, (0.5, 0.375), 0.3)
    vpb = ob.Problem(n, L, solid, lap=(d, ox, oy))
    rhs = 1e3 * cases.random_rhs(vpb, np.random.default_rng(9))
    po, res = ob.pcg(vpb, rhs, 10, 1e-14, tol_mode=1)
    c = U.Context(precision=U.FP64, tol_mode=U.TOL_REL)
    c.set_grid(n, vpb.h, solid)
    c.set_variable_lap((d, ox, oy))
    p, st = c.solve(rhs, 10, 1e-14)
    c.close()
    assert st.iters == res.iters == 10 and st.status == U.NOT_CONVERGED
    m = vpb.fluid_interior_mask()
    assert np.abs(p[m] - po[m]).max() <= 1e-9 * np.abs(po[m]).max()
    # warm start on the symmetric case: from the converged pressure nothing is left to do
    import torch
    name, pb = VAR_CASES[0]
    lap = cases.separable_lap(pb)
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(3))
    c = ctx_for(pb, precision=U.FP64, tol_mode=U.TOL_REL)
    c.set_variable_lap(lap)
    r = torch.from_numpy(rhs).to("cuda:0")
    pd = torch.zeros(pb.N, dtype=torch.float64, device="cuda:0")
    s1 = c.solve_device(r.data_ptr(), pd.data_ptr(), 0, 5000, 1e-6)
    first = pd.clone()
    s2 = c.solve_device(r.data_ptr(), pd.data_ptr(), 1, 5000, 1e-6)
    torch.cuda.synchronize()
    c.close()
    assert s1.status == 0 and s2.status == 0 and s2.iters < s1.iters, (s1.iters, s2.iters)
    vpb = ob.Problem(pb.n, pb.L, pb.solid, pb.wall, lap=lap)
    po, _ = ob.pcg(vpb, rhs, 5000, 1e-10, tol_mode=1)
    m = pb.fluid_interior_mask()
    for q in (first, pd):
        a = _demean(q.cpu().numpy(), m)
        assert np.abs(a - _demean(po, m)).max() <= 2e-5 * np.abs(po[m] - po[m].mean()).max()


def test_variable_lap_rejected_in_3d(gpu):
    name, pb = [c for c in PCG_CASES if c[1].dim == 3][0]
    c = ctx_for(pb)
    with pytest.raises(RuntimeError):
        c.set_variable_lap((np.ones(pb.N), np.zeros(pb.N), np.zeros(pb.N)))
    c.close()
