"""CPU tests of the oracle (the C restatement of the reference solve).

Pins, in order of strength:
  1. the golden fixtures written from the reference's own CPU solver
     (tests/golden/make_golden.py over oracle/_ref): bit-for-bit equality;
  2. the compiled reference itself (oracle/_ref), when present: bit-for-bit
     on fresh random cases;
  3. mathematics independent of the reference's code: graph-Laplacian
     brute force for the operator, A 1 = 0, dense R^T A P for the Galerkin
     coarse operator, symmetry of the V-cycle, a dense least-squares solve
     for PCG, and the reference's analytic sin*sin tests for amgpcg.
"""
import glob
import os

import numpy as np
import pytest

import oracle_bind as ob
import cases
from uaamg_amd import workloads as W

# the solver fixtures (the VTK frames, g7, are checked in test_vtk.py)
GOLDEN = sorted(p for p in glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz"))
                if not os.path.basename(p).startswith("g7_"))


def _problem(d):
    n = tuple(int(v) for v in d["n"])
    return ob.Problem(n, tuple(float(v) for v in d["L"]), d["solid"])


@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_oracle_matches_golden(path):
    d = np.load(path, allow_pickle=False)
    kind = str(d["kind"])
    if kind == "apply":
        pb = _problem(d)
        z = ob.uaamg_apply(pb, d["r"])
        m = np.zeros(pb.N, bool); pb.interior(m)[...] = True
        assert np.array_equal(z[m], d["z"][m])
    elif kind == "pcg":
        pb = _problem(d)
        p, res = ob.pcg(pb, d["rhs"], int(d["max_iter"]), float(d["tol"]))
        assert np.array_equal(p, d["p"])
        assert res.applies == int(d["applies"])
    elif kind == "pcg_ex":
        lap = (d["lap_diag"], d["lap_off_x"], d["lap_off_y"]) if "lap_diag" in d.files else None
        pb = ob.Problem(tuple(int(v) for v in d["n"]), tuple(float(v) for v in d["L"]), d["solid"], lap=lap)
        p, res = ob.pcg(pb, d["rhs"], int(d["max_iter"]), float(d["tol"]), precond=int(d["precond"]))
        assert res.status == (1 if "_40it" in path else 0)
        assert np.array_equal(p, d["p"])
    else:
        N = int(d["N"])
        b = d["b"] if "b" in d.files else None
        if b is None:
            b = np.zeros((N + 2) ** 2)
            h = 1.0 / (N + 1)
            ii = np.arange(1, N + 1)
            b.reshape(N + 2, N + 2)[1:-1, 1:-1] = 2 * np.pi ** 2 * np.outer(np.sin(np.pi * ii * h), np.sin(np.pi * ii * h))
        x, it = ob.amgpcg(N, b, int(d["max_iter"]), float(d["tol"]))
        assert it == int(d["iters"])
        if "x" in d.files:
            assert np.array_equal(x, d["x"])


@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (no reference tree here)")
@pytest.mark.parametrize("name,pb", cases.neumann_cases(), ids=[c[0] for c in cases.neumann_cases()])
def test_oracle_matches_reference_build(name, pb):
    rng = np.random.default_rng(99)
    r = cases.random_rhs(pb, rng)
    m = np.zeros(pb.N, bool); pb.interior(m)[...] = True
    assert np.array_equal(ob.uaamg_apply(pb, r)[m], ob.ref_uaamg_apply(pb, r)[m])
    p1, res = ob.pcg(pb, 1e3 * r, 300, 1e-3)
    p2, applies = ob.ref_pcg(pb, 1e3 * r, 300, 1e-3)
    assert np.array_equal(p1, p2) and res.applies == applies


# ---------------------------------------------------------------- operator

def graph_laplacian(pb, v):
    """(A v)_c = sum over fluid face-neighbours n of (v_c - v_n)/h_axis^2 on
    fluid cells (Neumann walls and solid faces carry no flux)."""
    fl = pb.fluid_interior_mask().reshape(pb.shape[::-1])
    V = v.reshape(pb.shape[::-1])
    out = np.zeros_like(V)
    for ax in range(pb.dim):
        h2 = pb.h[pb.dim - 1 - ax] ** 2
        for sgn in (1, -1):
            nb = np.roll(V, -sgn, axis=ax)
            fn = np.roll(fl, -sgn, axis=ax)
            out += np.where(fl & fn, (V - nb) / h2, 0.0)
    return np.where(fl, out, 0.0).reshape(-1)


@pytest.mark.parametrize("name,pb", cases.neumann_cases(), ids=[c[0] for c in cases.neumann_cases()])
def test_matvec_is_graph_laplacian(name, pb):
    rng = np.random.default_rng(5)
    v = cases.random_rhs(pb, rng)
    Av = ob.matvec(pb, v)
    ref = graph_laplacian(pb, v)
    m = pb.fluid_interior_mask()
    assert np.allclose(Av[m], ref[m], rtol=1e-12, atol=1e-9 * np.abs(ref).max())
    ones = np.zeros(pb.N); ones[m] = 1.0
    assert np.abs(ob.matvec(pb, ones)[m]).max() <= 1e-9 * max(1.0 / h ** 2 for h in pb.h)


def test_matvec_known_answers():
    # 1D-like profile on a 4x1 strip: constant -> 0; linear ramp -> interior 0, ends -+1/h^2
    pb = ob.Problem((4, 2), (4.0, 2.0))
    v = np.zeros(pb.N)
    V = pb.view(v)
    V[1:-1, 1:-1] = np.arange(1, 5)[None, :]
    Av = pb.view(ob.matvec(pb, v))
    assert np.allclose(Av[1, 1:-1], [-1.0, 0.0, 0.0, 1.0])


# ---------------------------------------------------------------- hierarchy

def dense_level0(pb):
    """Coefficient form of level 0 as a dense matrix over fluid cells."""
    idx = np.flatnonzero(pb.fluid_interior_mask())
    pos = {c: q for q, c in enumerate(idx)}
    n = len(idx)
    A = np.zeros((n, n))
    strides = [1, pb.shape[0]] + ([pb.shape[0] * pb.shape[1]] if pb.dim == 3 else [])
    for q, c in enumerate(idx):
        for ax in range(pb.dim):
            for sgn in (1, -1):
                cn = c + sgn * strides[ax]
                if cn in pos:
                    w = 1.0 / pb.h[ax] ** 2
                    A[q, pos[cn]] -= w
                    A[q, q] += w
    return A, idx


@pytest.mark.parametrize("n,L", [((8, 6), (1.0, 0.75)), ((9, 7), (1.0, 1.0)), ((8, 6, 6), (1.0, 0.75, 0.5)),
                                 ((9, 5, 6), (1.0, 1.0, 1.0))])
def test_galerkin_is_RtAP(n, L):
    """Without solids the reference's coarse stencil is exactly R A P with P
    the constant injection of fluid children (sinks at odd truncation)."""
    pb = ob.Problem(n, L)
    A, idx = dense_level0(pb)
    dims = ob.hierarchy_dims(pb)
    c = dims[1]
    cshape = tuple(v + 2 for v in c)
    cidx = []
    P = np.zeros((len(idx), int(np.prod(c))))
    pos = {cc: q for q, cc in enumerate(idx)}
    grid = np.stack(np.meshgrid(*[np.arange(1, v + 1) for v in c[::-1]], indexing="ij"), -1).reshape(-1, pb.dim)
    for q, cc in enumerate(grid):          # cc in (slow..fast) order
        ic = cc[::-1]
        cidx.append(int(sum(ic[a] * int(np.prod(cshape[:a])) for a in range(pb.dim))))
        for off in np.ndindex(*([2] * pb.dim)):
            f = [2 * ic[a] - 1 + off[a] for a in range(pb.dim)]
            fid = int(sum(f[a] * int(np.prod(pb.shape[:a])) for a in range(pb.dim)))
            if fid in pos:
                P[pos[fid], q] = 1.0
    Ac = P.T @ A @ P
    lv = ob.level_data(pb, 1)
    cidx = np.array(cidx)
    diag = lv["diag"][cidx]
    assert np.allclose(np.diag(Ac), diag, rtol=1e-13)
    cstr = [1, cshape[0]] + ([cshape[0] * cshape[1]] if pb.dim == 3 else [])
    where = {v: q for q, v in enumerate(cidx)}
    for q, cc in enumerate(cidx):
        for ax, ch in enumerate(["cx", "cy", "cz"][:pb.dim]):
            nb = cc + cstr[ax]
            if nb in where:
                assert np.isclose(-Ac[q, where[nb]], lv[ch][cc], rtol=1e-13)


def test_galerkin_closed_form_uniform():
    """Interior coarse couplings on a solid-free uniform grid grow by
    2^(d-1) per level (cuda_uaamg_preconditioner_3d.cu:358-363)."""
    for n, d in (((32, 32), 2), ((16, 16, 16), 3)):
        pb = ob.Problem(n)
        for lvl in range(1, len(ob.hierarchy_dims(pb))):
            lv = ob.level_data(pb, lvl)
            dims = lv["dims"]
            shape = tuple(v + 2 for v in dims)
            cx = lv["cx"].reshape(shape[::-1])
            inner = cx[(slice(2, -2),) * d] if d == 3 else cx[2:-2, 2:-3]
            expect = (1.0 / pb.h[0] ** 2) * (2 ** (d - 1)) ** lvl
            vals = inner[inner != 0]
            assert np.allclose(vals, expect, rtol=1e-14)


def test_coarse_solid_threshold():
    """restrictSolid: coarse solid iff >= 2 of 4 (2D) / >= 4 of 8 (3D) children solid."""
    pb = ob.Problem((4, 4))
    s = np.zeros(pb.shape[::-1], np.uint8)
    s[1, 1] = 1                      # 1 of 4 -> fluid
    s[1, 3] = s[2, 4] = 1            # 2 of 4 -> solid
    pb.solid = s.reshape(-1)
    lv = ob.level_data(pb, 1, coarsest_side=2)
    cs = lv["solid"].reshape(4, 4)
    assert cs[1, 1] == 0 and cs[1, 2] == 1
    pb3 = ob.Problem((4, 4, 4))
    s3 = np.zeros(pb3.shape[::-1], np.uint8)
    s3[1, 1, 1] = s3[1, 2, 1] = s3[2, 1, 1] = 1            # 3 of 8 -> fluid
    s3[1, 1, 3] = s3[1, 2, 3] = s3[2, 1, 3] = s3[2, 2, 4] = 1  # 4 of 8 -> solid
    pb3.solid = s3.reshape(-1)
    cs3 = ob.level_data(pb3, 1, coarsest_side=2)["solid"].reshape(4, 4, 4)
    assert cs3[1, 1, 1] == 0 and cs3[1, 1, 2] == 1


# ---------------------------------------------------------------- V-cycle and PCG

@pytest.mark.parametrize("name,pb", cases.parity_cases(), ids=[c[0] for c in cases.parity_cases()])
def test_vcycle_symmetric(name, pb):
    """Forward pre-smooth, reverse post-smooth, symmetric coarse operators:
    <M r1, r2> = <r1, M r2>."""
    rng = np.random.default_rng(8)
    m = pb.fluid_interior_mask()
    r1, r2 = cases.random_rhs(pb, rng), cases.random_rhs(pb, rng)
    r1[~m] = 0; r2[~m] = 0
    a = np.dot(ob.uaamg_apply(pb, r1)[m], r2[m])
    b = np.dot(r1[m], ob.uaamg_apply(pb, r2)[m])
    assert abs(a - b) <= 1e-11 * max(abs(a), abs(b))


@pytest.mark.parametrize("n,L,solid", [((12, 8), (1.0, 1.0), False), ((10, 9), (1.0, 0.6), True),
                                       ((6, 5, 4), (1.0, 1.0, 1.0), False), ((7, 6, 5), (1.0, 0.8, 0.9), True)])
def test_pcg_matches_dense_solve(n, L, solid):
    pb = ob.Problem(n, L)
    if solid:
        s = np.zeros(pb.shape[::-1], np.uint8)
        s[(slice(2, 4),) * pb.dim] = 1
        pb.solid = s.reshape(-1)
    rng = np.random.default_rng(2)
    rhs = 1e3 * cases.random_rhs(pb, rng)
    m = pb.fluid_interior_mask()
    rhs[m] -= rhs[m].mean()
    p, res = ob.pcg(pb, rhs, 200, 1e-10, tol_mode=1)
    assert res.status == 0
    A, idx = dense_level0(pb)
    x = np.linalg.lstsq(A, -rhs[idx], rcond=None)[0]   # lap p = rhs  <=>  (-lap) p = -rhs
    a = p[idx] - p[idx].mean()
    b = x - x.mean()
    assert np.abs(a - b).max() <= 1e-7 * np.abs(b).max()


def test_pcg_zero_rhs_returns_unchanged():
    """rhs = 0: pAp = 0 < 1e-15 at the first iteration, p untouched (pcg.cpp:100)."""
    pb = ob.Problem((8, 8))
    p0 = np.full(pb.N, 3.0)
    p, res = ob.pcg(pb, np.zeros(pb.N), 10, 1e-6, p0=p0)
    assert res.status == 2 and res.iters == 1 and np.array_equal(p, p0)


def test_amgpcg_analytic():
    """test_amgpcg.cpp:75-92: sin*sin at N=32, error < 1e-2 in <= 30 iterations."""
    N = 32
    h = 1.0 / (N + 1)
    ii = np.arange(1, N + 1)
    S = np.outer(np.sin(np.pi * ii * h), np.sin(np.pi * ii * h))
    b = np.zeros((N + 2, N + 2)); b[1:-1, 1:-1] = 2 * np.pi ** 2 * S
    x, it = ob.amgpcg(N, b.reshape(-1), 30, 1e-8)
    assert it <= 30
    assert np.abs(x.reshape(N + 2, N + 2)[1:-1, 1:-1] - S).max() < 1e-2


def test_dirichlet_extension_second_order():
    """Dirichlet walls (extension): lap p = 2 pi^2 sin sin on the unit square
    converges to -sin sin at second order."""
    errs = []
    for n in (16, 32, 64):
        pb = ob.Problem((n, n), wall=[ob.DIRICHLET] * 6)
        h = 1.0 / n
        c = (np.arange(n + 2) - 0.5) * h
        X, Y = np.meshgrid(c, c)
        rhs = 2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
        rhs[0, :] = rhs[-1, :] = rhs[:, 0] = rhs[:, -1] = 0
        # scaled by 1e4 so the absolute pAp < 1e-15 guard stays far away
        p, res = ob.pcg(pb, 1e4 * rhs.reshape(-1), 200, 1e-10, tol_mode=1)
        assert res.status == 0
        exact = -np.sin(np.pi * X) * np.sin(np.pi * Y)
        errs.append(np.abs(1e-4 * p.reshape(n + 2, n + 2)[1:-1, 1:-1] - exact[1:-1, 1:-1]).max())
    assert errs[1] < errs[0] / 3.2 and errs[2] < errs[1] / 3.2


# ---- pressure projection (pressure.cpp:5-32, pressure_3d.cpp:5-39) ----

PROJ_CASES = [c for c in cases.neumann_cases() if c[0] in
              ("2d_32x16", "2d_64x16_cylinder", "2d_33x17_odd", "3d_16_sphere", "3d_17x9x5_odd",
               "3d_24x20x18_aniso_solids")]


@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (no reference tree)")
@pytest.mark.parametrize("name,pb", PROJ_CASES, ids=[c[0] for c in PROJ_CASES])
def test_project_matches_reference(name, pb):
    """The oracle's projection equals the reference's PressureProjection[3D]
    ::project with the pcg_uaamg solver, bit for bit (u, v, w and p)."""
    rng = np.random.default_rng(5)
    u, v, w = cases.mac_fields(pb, rng, closed=False)
    dt = 0.01
    uo, vo, wo, po, res = ob.project(pb, u, v, w, dt, 200, 1e-10)
    ur, vr, wr, pr = ob.ref_project(pb, u, v, w, dt, 200, 1e-10)
    np.testing.assert_array_equal(po, pr)
    np.testing.assert_array_equal(uo, ur)
    np.testing.assert_array_equal(vo, vr)
    if pb.dim == 3:
        np.testing.assert_array_equal(wo, wr)


@pytest.mark.parametrize("name,pb", PROJ_CASES, ids=[c[0] for c in PROJ_CASES])
def test_project_divergence_free(name, pb):
    """After the projection of a closed field (no flux through walls or solid
    faces) the discrete divergence vanishes on fluid cells to the solver's
    tolerance: this fixes the rhs sign and scaling and the faces and sign of
    the correction (a wrong index or sign leaves O(1) divergence)."""
    rng = np.random.default_rng(6)
    u, v, w = cases.mac_fields(pb, rng, closed=True)
    dt = 0.02
    d0 = cases.mac_divergence(pb, u, v, w)
    fl = pb.fluid_interior_mask()
    uo, vo, wo, po, res = ob.project(pb, u, v, w, dt, 500, 1e-9 * np.abs(d0[fl]).max() / dt)
    # converged, or the reference's silent stop at p.Ap < 1e-15 (pcg.cpp:100),
    # which this absolute threshold reaches once r is at round-off
    assert res.status in (0, 2), res.status
    d1 = cases.mac_divergence(pb, uo, vo, wo)
    assert np.abs(d1[fl]).max() <= 1e-8 * np.abs(d0[fl]).max(), (np.abs(d1[fl]).max(), np.abs(d0[fl]).max())
    # faces touching a solid cell and the wall faces are left as they were
    nx = pb.n[0]
    U0 = u.reshape(-1, nx + 1)
    U1 = uo.reshape(-1, nx + 1)
    np.testing.assert_array_equal(U1[:, 0], U0[:, 0])
    np.testing.assert_array_equal(U1[:, nx], U0[:, nx])


def test_project_rhs_linear_field():
    """u = a x, v = b y, w = c z (closed form): div = a + b + c on every
    fluid cell, so rhs = (a + b + c)/dt there and 0 on solid cells."""
    pb = ob.Problem((12, 10, 8), (1.2, 0.5, 0.4))
    s = np.zeros(pb.shape[::-1], np.uint8)
    s[3:5, 4:6, 5:8] = 1
    pb.solid = s.reshape(-1)
    nx, ny, nz = pb.n
    a, b, c, dt = 2.0, -3.0, 0.5, 0.25
    xf = np.arange(nx + 1) * pb.h[0]
    yf = np.arange(ny + 1) * pb.h[1]
    zf = np.arange(nz + 1) * pb.h[2]
    u = np.broadcast_to(a * xf, (nz + 2, ny + 2, nx + 1)).reshape(-1)
    v = np.broadcast_to((b * yf)[None, :, None], (nz + 2, ny + 1, nx + 2)).reshape(-1)
    w = np.broadcast_to((c * zf)[:, None, None], (nz + 1, ny + 2, nx + 2)).reshape(-1)
    rhs = ob.project_rhs(pb, u, v, w, dt)
    fl = pb.fluid_interior_mask()
    np.testing.assert_allclose(rhs[fl], (a + b + c) / dt, rtol=1e-12)
    assert np.all(rhs[~fl] == 0)


# ---------------------------------------------------------------- "cg" key and the variable Laplacian

@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (no reference tree here)")
@pytest.mark.parametrize("name,pb", cases.neumann_cases(), ids=[c[0] for c in cases.neumann_cases()])
def test_oracle_cg_matches_reference_build(name, pb):
    """Key "cg": PCG with IdentityPreconditioner[3D] (factory.cpp:19,
    factory_3d.cpp:19), bit for bit against the reference build."""
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(5))
    p1, res = ob.pcg(pb, rhs, 300, 1e-3, precond=ob.PRECOND_IDENTITY)
    p2 = ob.ref_pcg_ex(pb, rhs, 300, 1e-3, precond=ob.PRECOND_IDENTITY)
    assert np.array_equal(p1, p2)


@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (no reference tree here)")
@pytest.mark.parametrize("precond", [ob.PRECOND_UAAMG, ob.PRECOND_IDENTITY], ids=["uaamg", "cg"])
@pytest.mark.parametrize("coef", ["separable", "smooth_cylinder"])
def test_oracle_variable_lap_matches_reference_build(precond, coef):
    """The variable-coefficient matvec (pcg.cpp:36-46) inside PCG, bit for
    bit against the reference build (the V-cycle stays the uniform one, as in
    the reference).  Symmetric coefficients (cases.separable_lap) converge;
    the coefficients SmoothCylinder writes make a non-symmetric operator whose
    range misses the zero-mean rhs, so the reference's PCG does not converge
    and 40 iterations are compared."""
    if coef == "separable":
        pb = ob.Problem((41, 30), (1.0, 0.8))
        s = pb.view(pb.solid); s[10:16, 12:20] = 1
        pb.lap = cases.separable_lap(pb)
        max_iter, rel = 600, 1e-6
    else:
        solid, d, ox, oy = W.smooth_cylinder_lap((40, 32), (1.0, 0.8), (0.5, 0.4), 0.33)
        pb = ob.Problem((40, 32), (1.0, 0.8), solid, lap=(d, ox, oy))
        max_iter, rel = 40, 1e-12
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(8))
    tol = rel * np.linalg.norm(rhs[pb.fluid_interior_mask()])
    p1, res = ob.pcg(pb, rhs, max_iter, tol, precond=precond)
    p2 = ob.ref_pcg_ex(pb, rhs, max_iter, tol, precond=precond)
    assert res.iters > 3 and res.status == (0 if coef == "separable" else 1)
    assert np.array_equal(p1, p2)


def test_smooth_cylinder_coefficients_match_reference():
    """workloads.smooth_cylinder_lap (the tests' generator of variable
    coefficients) equals SmoothCylinder::applyToLaplacian bit for bit."""
    if not ob.ref_available():
        pytest.skip("oracle/_ref not built (no reference tree here)")
    for n, L, c, R, tw in (((64, 48), (1.0, 0.75), (0.5, 0.37), 0.3, 0.0), ((33, 17), (2.0, 1.0), (1.1, 0.45), 0.4, 0.05)):
        a = W.smooth_cylinder_lap(n, L, c, R, tw)
        b = ob.ref_smooth_cylinder_lap(n, L, c, R, tw)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_variable_lap_with_uniform_coefficients_is_uniform():
    """Coefficients diag = 2(1/dx^2 + 1/dy^2), off = -1/h^2 turn the
    variable branch into the uniform operator exactly (a - b c == a + (-b) c
    in IEEE arithmetic): matvec and PCG agree bit for bit with the uniform
    form; a dropped or swapped coefficient would not."""
    for name, pb in cases.neumann_cases():
        if pb.dim != 2:
            continue
        ix, iy = 1 / pb.h[0] ** 2, 1 / pb.h[1] ** 2
        var = ob.Problem(pb.n, pb.L, pb.solid, pb.wall,
                         lap=(np.full(pb.N, 2 * (ix + iy)), np.full(pb.N, -ix), np.full(pb.N, -iy)))
        v = cases.random_rhs(pb, np.random.default_rng(2))
        assert np.array_equal(ob.matvec(var, v), ob.matvec(pb, v)), name
        # swapping the axes' coefficients changes the operator on anisotropic grids
        if ix != iy:
            swp = ob.Problem(pb.n, pb.L, pb.solid, pb.wall,
                             lap=(np.full(pb.N, 2 * (ix + iy)), np.full(pb.N, -iy), np.full(pb.N, -ix)))
            assert not np.array_equal(ob.matvec(swp, v), ob.matvec(pb, v)), name
        rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(3))
        p1, r1 = ob.pcg(var, rhs, 200, 1e-4)
        p2, r2 = ob.pcg(pb, rhs, 200, 1e-4)
        assert np.array_equal(p1, p2) and r1.iters == r2.iters, name


def test_cg_matches_dense_solve():
    """Plain CG (identity preconditioner) converges to the dense solution of
    the fluid-cell system (a least-squares solve of the singular Neumann
    matrix, mean removed), on a grid with a solid block."""
    pb = ob.Problem((9, 7), (1.0, 0.8))
    s = pb.view(pb.solid); s[3:5, 4:6] = 1
    m = pb.fluid_interior_mask()
    idx = np.flatnonzero(m)
    A = np.zeros((idx.size, idx.size))
    for c, k in enumerate(idx):
        e = np.zeros(pb.N); e[k] = 1.0
        A[:, c] = ob.matvec(pb, e)[idx]
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(4))
    rhs[m] -= rhs[m].mean()
    p, res = ob.pcg(pb, rhs, 500, 1e-8, precond=ob.PRECOND_IDENTITY)
    assert res.status == 0
    x = np.linalg.lstsq(A, -rhs[idx], rcond=None)[0]
    a, b = p[idx] - p[idx].mean(), x - x.mean()
    assert np.abs(a - b).max() <= 1e-9 * np.abs(b).max()


# ---------------------------------------------------------------- keys "pcg" / "pcg_gmg" / "pcg_amg" (2D)

GMG_AMG = [(ob.PRECOND_GMG, "gmg"), (ob.PRECOND_AMG, "amg")]


@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (no reference tree here)")
@pytest.mark.parametrize("precond,tag", GMG_AMG, ids=[t for _, t in GMG_AMG])
@pytest.mark.parametrize("name,pb", [c for c in cases.neumann_cases() if c[1].dim == 2],
                         ids=[c[0] for c in cases.neumann_cases() if c[1].dim == 2])
def test_oracle_gmg_amg_match_reference_build(name, pb, precond, tag):
    """PCG with GMGPreconditioner / AMGPreconditioner (factory.cpp:20-22),
    bit for bit against the reference build: iterates, residual path and
    iteration count."""
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(7))
    tol = 1e-8 * np.linalg.norm(rhs[pb.fluid_interior_mask()])
    p1, res = ob.pcg(pb, rhs, 300, tol, precond=precond)
    p2 = ob.ref_pcg_ex(pb, rhs, 300, tol, precond=precond)
    assert res.status == 0 and res.iters > 3
    assert np.array_equal(p1, p2)


def _coarse_solid(pb):
    """A coarse cell is solid when at least two of its 2x2 children are."""
    nx, ny = pb.n
    s = pb.view(pb.solid)[1:ny + 1, 1:nx + 1] != 0
    cx, cy = nx // 2, ny // 2
    cnt = s[:2 * cy, :2 * cx].reshape(cy, 2, cx, 2).sum(axis=(1, 3))
    out = np.zeros((cy + 2, cx + 2), np.uint8)
    out[1:-1, 1:-1] = cnt >= 2
    return out.reshape(-1)


def _dense_vcycle(pb, b, side=4, csweeps=10):
    """The GMG / AMG V(2,2) cycle as dense linear algebra over the fluid
    cells: red-black Gauss-Seidel on the graph Laplacian (odd colour i + j
    first forward, even first in reverse), the coarse rhs the mean of the
    fluid children's residuals b - A x, the coarse correction added to the
    fluid children of fluid coarse cells, the coarse operator rediscretised
    at twice the spacing."""
    A, idx = dense_level0(pb)
    sx = pb.n[0] + 2
    par = (idx % sx + idx // sx) & 1
    diag = np.diag(A).copy()

    def colour(x, bb, odd):
        for q in np.flatnonzero(par == odd):      # cells of one colour are decoupled
            if diag[q] > 0:
                x[q] += (bb[q] - A[q] @ x) / diag[q]

    def sweep(x, bb, reverse):
        colour(x, bb, 0 if reverse else 1)
        colour(x, bb, 1 if reverse else 0)

    bb = b[idx].copy()
    x = np.zeros(idx.size)
    if pb.n[0] <= side or pb.n[1] <= side or pb.n[0] // 2 < 2 or pb.n[1] // 2 < 2:
        for _ in range(csweeps):
            sweep(x, bb, False)
            sweep(x, bb, True)
        return idx, x
    for _ in range(2):
        sweep(x, bb, False)
    res = np.zeros(pb.N); res[idx] = bb - A @ x
    cn = (pb.n[0] // 2, pb.n[1] // 2)
    pc = ob.Problem(cn, (2 * pb.h[0] * cn[0], 2 * pb.h[1] * cn[1]), _coarse_solid(pb))
    fl = np.zeros(pb.N, bool); fl[idx] = True
    bc = np.zeros(pc.N)
    for jc in range(1, cn[1] + 1):
        for ic in range(1, cn[0] + 1):
            ch = [(2 * ic - 1 + di) + (2 * jc - 1 + dj) * sx for di in (0, 1) for dj in (0, 1)]
            ch = [c for c in ch if fl[c]]
            if ch:
                bc[ic + jc * (cn[0] + 2)] = np.mean(res[ch])
    cidx, xc = _dense_vcycle(pc, bc, side, csweeps)
    xcf = np.zeros(pc.N); xcf[cidx] = xc
    for q, c in enumerate(idx):
        i, j = c % sx, c // sx
        ic, jc = (i + 1) // 2, (j + 1) // 2
        if ic <= cn[0] and jc <= cn[1]:
            x[q] += xcf[ic + jc * (cn[0] + 2)]
    for _ in range(2):
        sweep(x, bb, True)
    return idx, x


@pytest.mark.parametrize("precond,tag", GMG_AMG, ids=[t for _, t in GMG_AMG])
@pytest.mark.parametrize("n,L,block", [((8, 6), (1.0, 0.75), None), ((16, 12), (1.0, 0.6), (4, 9, 3, 7)),
                                       ((13, 10), (1.3, 1.0), (2, 6, 5, 8)), ((20, 9), (2.0, 0.5), (7, 8, 1, 10))])
def test_gmg_amg_vcycle_is_dense_multigrid(precond, tag, n, L, block):
    """One GMG / AMG application equals the V(2,2) cycle written as dense
    linear algebra (_dense_vcycle): a wrong colour order, stencil term,
    restriction weight or prolongation target changes z.  The two classes
    differ only in the order of their floating-point operations."""
    pb = ob.Problem(n, L)
    if block:
        i0, i1, j0, j1 = block
        pb.view(pb.solid)[j0:j1, i0:i1] = 1
    r = cases.random_rhs(pb, np.random.default_rng(11))
    z = ob.precond_apply(pb, r, precond)
    idx, x = _dense_vcycle(pb, r)
    assert np.abs(z[idx] - x).max() <= 1e-11 * np.abs(x).max()
    # solid interior cells keep z = 0
    s = np.zeros(pb.N, bool); pb.interior(s)[...] = pb.interior(pb.solid) != 0
    assert not z[s].any()


def test_gmg_and_amg_agree_to_rounding():
    """GMG and AMG restate the same cycle (their residual and restriction
    formulas are equal in exact arithmetic): their PCG solves agree to
    rounding and take the same number of iterations within one."""
    for name, pb in cases.neumann_cases():
        if pb.dim != 2:
            continue
        rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(12))
        pg, rg = ob.pcg(pb, rhs, 300, 1e-10, tol_mode=1, precond=ob.PRECOND_GMG)
        pa, ra = ob.pcg(pb, rhs, 300, 1e-10, tol_mode=1, precond=ob.PRECOND_AMG)
        m = pb.fluid_interior_mask()
        a, b = pg[m] - pg[m].mean(), pa[m] - pa[m].mean()
        assert rg.status == 0 and ra.status == 0 and abs(rg.iters - ra.iters) <= 1, name
        assert np.abs(a - b).max() <= 1e-8 * np.abs(b).max(), name


# ---------------------------------------------------------------- hierarchy_mode 1 (keep fluid)

def _aggregation_P(pb, c):
    """P: fluid fine cell -> its coarse parent (constant injection over the
    2^d children), columns over every coarse interior cell."""
    A, idx = dense_level0(pb)
    cshape = tuple(v + 2 for v in c)
    pos = {cc: q for q, cc in enumerate(idx)}
    cidx = []
    P = np.zeros((len(idx), int(np.prod(c))))
    grid = np.stack(np.meshgrid(*[np.arange(1, v + 1) for v in c[::-1]], indexing="ij"), -1).reshape(-1, pb.dim)
    for q, cc in enumerate(grid):
        ic = cc[::-1]
        cidx.append(int(sum(ic[a] * int(np.prod(cshape[:a])) for a in range(pb.dim))))
        for off in np.ndindex(*([2] * pb.dim)):
            f = [2 * ic[a] - 1 + off[a] for a in range(pb.dim)]
            fid = int(sum(f[a] * int(np.prod(pb.shape[:a])) for a in range(pb.dim)))
            if fid in pos:
                P[pos[fid], q] = 1.0
    return A, P, np.array(cidx), cshape


@pytest.mark.parametrize("n,L,blocks", [((8, 6), (1.0, 0.75), [(2, 5, 1, 4)]), ((12, 10), (1.0, 1.0), [(3, 8, 2, 5), (9, 11, 6, 9)]),
                                        ((8, 6, 6), (1.0, 0.75, 0.5), [(2, 5, 1, 4, 2, 5)]),
                                        ((10, 8, 6), (1.0, 1.0, 0.8), [(3, 7, 2, 5, 1, 4)])])
def test_keep_fluid_galerkin_is_RtAP_with_solids(n, L, blocks):
    """hierarchy_mode 1: a coarse cell is solid only when all its children
    are, so every coarse coupling is the aggregation Galerkin product
    P^T A P, solids included (under the reference's threshold a coarse
    fluid cell keeps couplings into half-solid coarse cells, which no P^T A P
    has).  Blocks that cover some children of a coarse cell test it."""
    pb = ob.Problem(n, L)
    s = pb.view(pb.solid)
    for b in blocks:
        if pb.dim == 2:
            s[b[2]:b[3], b[0]:b[1]] = 1
        else:
            s[b[4]:b[5], b[2]:b[3], b[0]:b[1]] = 1
    A, P, cidx, cshape = _aggregation_P(pb, ob.hierarchy_dims(pb)[1])
    Ac = P.T @ A @ P
    lv = ob.level_data(pb, 1, hierarchy_mode=1)
    fluid_c = P.sum(axis=0) > 0
    assert np.array_equal(lv["solid"][cidx] == 0, fluid_c)          # solid iff no fluid child
    assert np.allclose(lv["diag"][cidx][fluid_c], np.diag(Ac)[fluid_c], rtol=1e-13)
    cstr = [1, cshape[0]] + ([cshape[0] * cshape[1]] if pb.dim == 3 else [])
    where = {v: q for q, v in enumerate(cidx)}
    checked = 0
    for q, cc in enumerate(cidx):
        if not fluid_c[q]:
            continue
        for ax, ch in enumerate(["cx", "cy", "cz"][:pb.dim]):
            nb = cc + cstr[ax]
            if nb in where:
                assert np.isclose(-Ac[q, where[nb]], lv[ch][cc], rtol=1e-13, atol=1e-9), (q, ch)
                checked += 1
    assert checked > 0
    # the reference's threshold does not give P^T A P here
    ref = ob.level_data(pb, 1)
    assert not np.array_equal(ref["solid"][cidx] == 0, fluid_c)


def test_keep_fluid_threshold():
    """Coarse solid only when every child is: 3 of 4 (2D) / 7 of 8 (3D)
    solid children leave the coarse cell fluid."""
    pb = ob.Problem((4, 4))
    s = np.zeros(pb.shape[::-1], np.uint8)
    s[1, 1] = s[1, 2] = s[2, 1] = 1              # 3 of 4 -> fluid
    s[1, 3] = s[1, 4] = s[2, 3] = s[2, 4] = 1    # 4 of 4 -> solid
    pb.solid = s.reshape(-1)
    cs = ob.level_data(pb, 1, coarsest_side=2, hierarchy_mode=1)["solid"].reshape(4, 4)
    assert cs[1, 1] == 0 and cs[1, 2] == 1
    pb3 = ob.Problem((4, 4, 4))
    s3 = np.zeros(pb3.shape[::-1], np.uint8)
    s3[1:3, 1:3, 1:3] = 1; s3[2, 2, 2] = 0        # 7 of 8 -> fluid
    s3[1:3, 1:3, 3:5] = 1                          # 8 of 8 -> solid
    pb3.solid = s3.reshape(-1)
    cs3 = ob.level_data(pb3, 1, coarsest_side=2, hierarchy_mode=1)["solid"].reshape(4, 4, 4)
    assert cs3[1, 1, 1] == 0 and cs3[1, 1, 2] == 1


@pytest.mark.parametrize("name,pb", [c for c in cases.parity_cases() if "porous" not in c[0]],
                         ids=[c[0] for c in cases.parity_cases() if "porous" not in c[0]])
def test_keep_fluid_vcycle_symmetric_and_not_slower(name, pb):
    """hierarchy_mode 1 keeps the V-cycle symmetric (<M r1, r2> = <r1, M r2>)
    and never needs more PCG iterations than the reference's hierarchy on
    the parity geometries (fewer where solids cut coarse cells)."""
    rng = np.random.default_rng(8)
    m = pb.fluid_interior_mask()
    r1, r2 = cases.random_rhs(pb, rng), cases.random_rhs(pb, rng)
    r1[~m] = 0; r2[~m] = 0
    a = np.dot(ob.uaamg_apply(pb, r1, hierarchy_mode=1)[m], r2[m])
    b = np.dot(r1[m], ob.uaamg_apply(pb, r2, hierarchy_mode=1)[m])
    assert abs(a - b) <= 1e-11 * max(abs(a), abs(b))
    rhs = 1e3 * cases.random_rhs(pb, np.random.default_rng(3))
    _, r0 = ob.pcg(pb, rhs, 2000, 1e-6, tol_mode=1)
    p1, r1_ = ob.pcg(pb, rhs, 2000, 1e-6, tol_mode=1, hierarchy_mode=1)
    assert r0.status == 0 and r1_.status == 0
    assert r1_.iters <= r0.iters, (r1_.iters, r0.iters)
    # and solves the same system: the two pressures agree to the tolerance's accuracy
    p0, _ = ob.pcg(pb, rhs, 2000, 1e-10, tol_mode=1)
    p1, _ = ob.pcg(pb, rhs, 2000, 1e-10, tol_mode=1, hierarchy_mode=1)
    x, y = p0[m] - p0[m].mean(), p1[m] - p1[m].mean()
    assert np.abs(x - y).max() <= 1e-7 * np.abs(x).max()


def test_keep_fluid_plume_iterations():
    """The 64^3 plume (8 spheres and a floor slab) at relres 1e-6: the
    reference's hierarchy needs 35 iterations, hierarchy_mode 1 needs 12."""
    wp = W.plume3d(seed=0, n=64)
    pb = ob.Problem(wp.n, wp.L, wp.solid, wp.wall)
    _, r0 = ob.pcg(pb, wp.rhs, 500, 1e-6, tol_mode=1)
    _, r1 = ob.pcg(pb, wp.rhs, 500, 1e-6, tol_mode=1, hierarchy_mode=1)
    assert (r0.iters, r1.iters) == (35, 12), (r0.iters, r1.iters)
