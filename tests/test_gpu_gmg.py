"""GPU parity of the factory's other PCG keys, 2D: "pcg" / "pcg_gmg"
(GMGPreconditioner) and "pcg_amg" (AMGPreconditioner), factory.cpp:20-22,
through the C-ABI against the oracle (pinned bit for bit against the
reference build and against a dense-matrix multigrid cycle)."""
import numpy as np
import pytest

import oracle_bind as ob
import cases
import uaamg_amd as U

pytestmark = pytest.mark.gpu

CASES2 = [c for c in cases.parity_cases() if c[1].dim == 2]
PCG2 = [c for c in CASES2 if "porous" not in c[0]]
# Neumann walls only: the reference's 2D grid has no other walls, and its
# GMG / AMG smoothers treat every wall as Neumann; on a Dirichlet grid the
# cycle is a weaker preconditioner of the operator (FP64 still within one
# iteration of the oracle, FP32 8 iterations more on 2d_40x30_dirichlet)
PCG2N = [c for c in PCG2 if not any(c[1].wall)]
KINDS = [(U.PRECOND_GMG, ob.PRECOND_GMG, "gmg"), (U.PRECOND_AMG, ob.PRECOND_AMG, "amg")]
KIDS = [k[2] for k in KINDS]


def _ctx(pb, **opts):
    c = U.Context(**opts)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    return c


def _demean(p, m):
    return p[m] - p[m].mean()


@pytest.mark.parametrize("kind", KINDS, ids=KIDS)
@pytest.mark.parametrize("prec", [U.FP64, U.FP32, U.MIXED_F64], ids=["fp64", "fp32", "mixed_f64"])
@pytest.mark.parametrize("name,pb", CASES2, ids=[c[0] for c in CASES2])
def test_gmg_amg_apply(gpu, name, pb, prec, kind):
    """One V(2,2) application z = M r.  FP64: each cell's update is the
    oracle's arithmetic in the oracle's order and the colours are updated as
    independent sets, so z equals the oracle's bit for bit; FP32 /
    FP32-level forms within 1e-4 (measured 1e-7 to 5e-7)."""
    upc, opc, tag = kind
    r = cases.random_rhs(pb, np.random.default_rng(11))
    c = _ctx(pb, precision=prec, preconditioner=upc)
    z = c.apply(r)
    c.close()
    zo = ob.precond_apply(pb, r, opc)
    m = np.zeros(pb.N, bool); pb.interior(m)[...] = True
    err = np.abs(z[m] - zo[m]).max() / np.abs(zo[m]).max()
    print(f"{name} {tag} prec={prec}: apply rel err {err:.3e}", flush=True)
    if prec == U.FP64:
        assert np.array_equal(z[m], zo[m]), err
    assert err <= 1e-4, err


@pytest.mark.parametrize("kind", KINDS, ids=KIDS)
@pytest.mark.parametrize("name,pb", PCG2, ids=[c[0] for c in PCG2])
def test_gmg_amg_pcg_fp64(gpu, name, pb, kind):
    """FP64 PCG: within +-1 iteration of the oracle at relres 1e-8, and the
    pressure (mean removed) within 1e-9 of max|p| at relres 1e-10."""
    upc, opc, tag = kind
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(3))
    po8, r8 = ob.pcg(pb, rhs, 500, 1e-8, tol_mode=1, precond=opc)
    po, res = ob.pcg(pb, rhs, 500, 1e-10, tol_mode=1, precond=opc)
    assert r8.status == 0 and res.status == 0
    c = _ctx(pb, precision=U.FP64, tol_mode=U.TOL_REL, preconditioner=upc)
    _, st8 = c.solve(rhs, 500, 1e-8)
    p, st = c.solve(rhs, 500, 1e-10)
    c.close()
    print(f"{name} {tag}: {st8.iters} iterations, oracle {r8.iters}", flush=True)
    assert st8.status == 0 and st.status == 0
    assert abs(st8.iters - r8.iters) <= 1, (st8.iters, r8.iters)
    assert not st.forms_used & (U.FORM_ROWS | U.FORM_TAIL | U.FORM_PMV)
    m = pb.fluid_interior_mask()
    a, b = _demean(p, m), _demean(po, m)
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


@pytest.mark.parametrize("kind", KINDS, ids=KIDS)
@pytest.mark.parametrize("prec", [U.FP32, U.MIXED_F64], ids=["fp32", "mixed_f64"])
@pytest.mark.parametrize("name,pb", PCG2N, ids=[c[0] for c in PCG2N])
def test_gmg_amg_pcg_reduced(gpu, name, pb, prec, kind):
    """FP32 vectors and levels at relres 1e-6: pressure within 2e-5 of the
    FP64 oracle, iterations within 2 or 15 %.  FP64 CG with FP32 levels at
    relres 1e-8: within 2 iterations (or 5 %) and 1e-7."""
    upc, opc, tag = kind
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(3))
    tol = 1e-6 if prec == U.FP32 else 1e-8
    po, res = ob.pcg(pb, rhs, 500, tol, tol_mode=1, precond=opc)
    pt, _ = ob.pcg(pb, rhs, 500, 1e-10, tol_mode=1, precond=opc)
    c = _ctx(pb, precision=prec, tol_mode=U.TOL_REL, preconditioner=upc)
    p, st = c.solve(rhs, 500, tol)
    c.close()
    print(f"{name} {tag} prec={prec}: {st.iters} iterations, oracle {res.iters}", flush=True)
    assert st.status == 0, U.status_name(st.status)
    slack = max(2, int(0.15 * res.iters)) if prec == U.FP32 else max(2, res.iters // 20)
    assert -slack <= st.iters - res.iters <= slack, (st.iters, res.iters)
    m = pb.fluid_interior_mask()
    a, b = _demean(p, m), _demean(pt, m)
    assert np.abs(a - b).max() <= (2e-5 if prec == U.FP32 else 1e-7) * np.abs(b).max()


@pytest.mark.parametrize("kind", KINDS, ids=KIDS)
def test_gmg_amg_variable_lap_and_graph(gpu, kind):
    """The variable-coefficient operator (pcg.cpp:36-46) under the GMG / AMG
    V-cycle (which keeps its own uniform stencil, as the reference's does):
    FP64 within +-1 iteration and 1e-9 of the oracle; the graph-replayed
    solve equals the eager one bit for bit, three times over."""
    upc, opc, tag = kind
    name, pb = PCG2[0]
    lap = cases.separable_lap(pb)
    vpb = ob.Problem(pb.n, pb.L, pb.solid, pb.wall, lap=lap)
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(4))
    po, res = ob.pcg(vpb, rhs, 2000, 1e-10, tol_mode=1, precond=opc)
    assert res.status == 0
    eager = _ctx(pb, precision=U.FP64, tol_mode=U.TOL_REL, preconditioner=upc, use_graph=0)
    eager.set_variable_lap(lap)
    p0, st0 = eager.solve(rhs, 2000, 1e-10)
    assert st0.status == 0 and abs(st0.iters - res.iters) <= 1, (st0.iters, res.iters)
    m = pb.fluid_interior_mask()
    assert np.abs(_demean(p0, m) - _demean(po, m)).max() <= 1e-9 * np.abs(_demean(po, m)).max()
    graph = _ctx(pb, precision=U.FP64, tol_mode=U.TOL_REL, preconditioner=upc, use_graph=1)
    graph.set_variable_lap(lap)
    for rep in range(3):
        p1, st1 = graph.solve(rhs, 2000, 1e-10)
        assert st1.iters == st0.iters and np.array_equal(p1, p0), rep
        pe, _ = eager.solve(rhs, 2000, 1e-10)
        assert np.array_equal(pe, p0)
    eager.close(); graph.close()


@pytest.mark.parametrize("upc", [U.PRECOND_GMG, U.PRECOND_AMG], ids=KIDS)
def test_gmg_amg_rejected_in_3d(gpu, upc):
    """The 3D factory has no GMG / AMG (factory_3d.cpp:21-23): a 3D grid
    on such a context is an argument error."""
    name, pb = [c for c in cases.parity_cases() if c[1].dim == 3][0]
    c = U.Context(preconditioner=upc)
    with pytest.raises(RuntimeError):
        c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    c.close()
