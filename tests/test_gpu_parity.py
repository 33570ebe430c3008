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

@pytest.mark.parametrize("prec", [U.FP64, U.FP32, U.MIXED_BF16, U.MIXED_BF16R])
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
