"""Write the golden fixtures under tests/golden/ from the reference's own CPU
solver (oracle/_ref/libref.so, built by `make -C oracle ref` from the
reference sources).  Run once in the build container; the fixtures are data
(inputs and the reference's outputs) and are committed, the script with them.

G1  2D UAAMG V-cycle z = M r   (uaamg_preconditioner.cpp:177-199)
G2  3D UAAMG V-cycle           (uaamg_preconditioner_3d.cpp:198-229)
G3  PCG p and preconditioner-application counts (pcg.cpp / pcg_3d.cpp)
G4  standalone amgpcg x and iteration counts (test/solver/test_amgpcg.cpp)
G7  VTK frames of VtkWriter / VtkWriter3D (src/io/vtk_writer.cpp, vtk_writer_3d.cpp)
G6  PCG with the GMG / AMG preconditioners (keys pcg, pcg_gmg, pcg_amg)
G5  PCG with the identity preconditioner (factory key "cg", factory.cpp:19)
    and with the variable Laplacian of SmoothCylinder::applyToLaplacian
    (cylinder_model.cpp:50-71) that PCG::solve's matvec reads (pcg.cpp:36-46)
Inputs follow the reference's GPU tests where they exist
(test_cuda_uaamg.cu:21-26, :70-75, :136-149; test_cuda_uaamg_3d.cu:15-25).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bind as ob  # noqa: E402
import cases as shared_cases  # noqa: E402


def grid_dims(pb):
    return list(pb.n) + ([1] if pb.dim == 2 else [])


def cyl_mask(pb, cx, cy, R):
    m = np.zeros(pb.shape[::-1], np.uint8)
    for j in range(1, pb.n[1] + 1):
        for i in range(1, pb.n[0] + 1):
            x, y = (i - 0.5) * pb.h[0], (j - 0.5) * pb.h[1]
            if np.sqrt((x - cx) ** 2 + (y - cy) ** 2) < R:
                m[j, i] = 1
    return m.reshape(-1)


def sphere_mask(pb, c, R):
    k, j, i = np.meshgrid(*[np.arange(s) for s in pb.shape[::-1]], indexing="ij")
    x, y, z = (i - 0.5) * pb.h[0], (j - 0.5) * pb.h[1], (k - 0.5) * pb.h[2]
    inside = ((x - c[0]) ** 2 + (y - c[1]) ** 2 + (z - c[2]) ** 2) < R * R
    inter = (i >= 1) & (i <= pb.n[0]) & (j >= 1) & (j <= pb.n[1]) & (k >= 1) & (k <= pb.n[2])
    return (inside & inter).astype(np.uint8).reshape(-1)


def gaussian(pb):
    out = np.zeros(pb.N)
    v = pb.view(out)
    if pb.dim == 2:
        for j in range(1, pb.n[1] + 1):
            for i in range(1, pb.n[0] + 1):
                d2 = (i - pb.n[0] / 2.0) ** 2 + (j - pb.n[1] / 2.0) ** 2
                v[j, i] = np.exp(-0.01 * d2)
    else:
        for k in range(1, pb.n[2] + 1):
            for j in range(1, pb.n[1] + 1):
                for i in range(1, pb.n[0] + 1):
                    d2 = (i - pb.n[0] / 2.0) ** 2 + (j - pb.n[1] / 2.0) ** 2 + (k - pb.n[2] / 2.0) ** 2
                    v[k, j, i] = np.exp(-0.01 * d2)
    return out


def ones_fluid(pb):
    out = np.zeros(pb.N)
    m = pb.fluid_interior_mask()
    out[m] = 1.0
    return out


def cases():
    rng = np.random.default_rng(20260522)
    out = []
    # G1: 2D V-cycles
    pb = ob.Problem((32, 16), (1.0, 1.0))
    out.append(("g1_2d_32x16_ones", "apply", pb, ones_fluid(pb)))
    pb = ob.Problem((64, 16), (4.0, 1.0)); pb.solid = cyl_mask(pb, 1.0, 0.5, 0.1)
    out.append(("g1_2d_64x16_cylinder", "apply", pb, ones_fluid(pb)))
    pb = ob.Problem((33, 17), (1.0, 0.7))
    r = np.zeros(pb.N); pb.interior(r)[...] = rng.standard_normal((17, 33))
    out.append(("g1_2d_33x17_odd_random", "apply", pb, r))
    pb = ob.Problem((40, 24), (1.0, 24.0 / 20.0)); pb.solid = cyl_mask(pb, 0.3, 0.6, 0.12)
    r = np.zeros(pb.N); pb.interior(r)[...] = rng.standard_normal((24, 40))
    out.append(("g1_2d_40x24_aniso_cylinder", "apply", pb, r))
    # G2: 3D V-cycles
    pb = ob.Problem((16, 8, 4), (1.0, 1.0, 1.0))
    out.append(("g2_3d_16x8x4_ones", "apply", pb, ones_fluid(pb)))
    pb = ob.Problem((16, 16, 16)); pb.solid = sphere_mask(pb, (0.5, 0.5, 0.5), 0.3)
    r = np.zeros(pb.N); pb.interior(r)[...] = rng.standard_normal((16, 16, 16))
    out.append(("g2_3d_16_sphere_random", "apply", pb, r))
    pb = ob.Problem((17, 9, 5), (1.0, 0.9, 0.8))
    r = np.zeros(pb.N); pb.interior(r)[...] = rng.standard_normal((5, 9, 17))
    out.append(("g2_3d_17x9x5_odd_random", "apply", pb, r))
    # G3: PCG
    pb = ob.Problem((32, 16), (1.0, 1.0))
    out.append(("g3_2d_32x16_gauss_tol1e-10", "pcg", pb, (gaussian(pb), 50, 1e-10)))
    pb = ob.Problem((64, 16), (4.0, 1.0)); pb.solid = cyl_mask(pb, 1.0, 0.5, 0.1)
    out.append(("g3_2d_64x16_cylinder_gauss", "pcg", pb, (1e3 * gaussian(pb), 200, 1e-6)))
    pb = ob.Problem((16, 8, 8))
    out.append(("g3_3d_16x8x8_gauss_tol1e-10", "pcg", pb, (gaussian(pb), 30, 1e-10)))
    pb = ob.Problem((16, 16, 16)); pb.solid = sphere_mask(pb, (0.5, 0.4, 0.5), 0.25)
    out.append(("g3_3d_16_sphere_gauss", "pcg", pb, (1e3 * gaussian(pb), 200, 1e-6)))
    pb = ob.Problem((17, 9, 5), (1.0, 0.9, 0.8))
    r = np.zeros(pb.N); pb.interior(r)[...] = 1e3 * rng.standard_normal((5, 9, 17))
    out.append(("g3_3d_17x9x5_odd_random", "pcg", pb, (r, 300, 1e-6)))
    return out


def g5_cases():
    out = []
    pb = ob.Problem((48, 40), (1.2, 1.0)); pb.solid = cyl_mask(pb, 0.4, 0.5, 0.15)
    out.append(("g5_2d_48x40_cylinder_cg", pb, 1e3 * gaussian(pb), 400, 1e-6, ob.PRECOND_IDENTITY))
    pb = ob.Problem((16, 8, 8), (1.0, 0.5, 0.5)); pb.solid = sphere_mask(pb, (0.6, 0.25, 0.25), 0.12)
    out.append(("g5_3d_16x8x8_sphere_cg", pb, 1e3 * gaussian(pb), 200, 1e-6, ob.PRECOND_IDENTITY))
    for prec, tag in ((ob.PRECOND_UAAMG, "uaamg"), (ob.PRECOND_IDENTITY, "cg")):
        # symmetric coefficients (x couplings vary with y only, y couplings with x only): converges
        pb = ob.Problem((48, 36), (1.2, 0.9)); pb.solid = cyl_mask(pb, 0.5, 0.45, 0.12)
        pb.lap = shared_cases.separable_lap(pb)
        out.append((f"g5_2d_48x36_separable_varlap_{tag}", pb, 1e3 * gaussian(pb), 600, 1e-6, prec))
        # SmoothCylinder's coefficients give a non-symmetric operator whose range misses the
        # zero-mean rhs: the reference's PCG does not converge; 40 iterations are the fixture
        n, L = (64, 48), (1.0, 0.75)
        solid, d, ox, oy = ob.ref_smooth_cylinder_lap(n, L, (0.5, 0.375), 0.3)
        pb = ob.Problem(n, L, solid, lap=(d, ox, oy))
        out.append((f"g5_2d_64x48_smoothcyl_varlap_{tag}_40it", pb, 1e3 * gaussian(pb), 40, 1e-12, prec))
    return out


def g6_cases():
    """Keys "pcg" / "pcg_gmg" (GMGPreconditioner) and "pcg_amg"
    (AMGPreconditioner), factory.cpp:20-22: 2D only."""
    out = []
    rng = np.random.default_rng(6)
    for prec, tag in ((ob.PRECOND_GMG, "gmg"), (ob.PRECOND_AMG, "amg")):
        pb = ob.Problem((48, 40), (1.2, 1.0)); pb.solid = cyl_mask(pb, 0.4, 0.5, 0.15)
        out.append((f"g6_2d_48x40_cylinder_{tag}", pb, 1e3 * gaussian(pb), 300, 1e-6, prec))
        pb = ob.Problem((33, 17), (1.0, 0.6))
        r = np.zeros(pb.N); pb.interior(r)[...] = 1e3 * rng.standard_normal((17, 33))
        out.append((f"g6_2d_33x17_odd_random_{tag}", pb, r, 300, 1e-8, prec))
        # x couplings varying with y, y couplings with x, under the GMG / AMG V-cycle
        pb = ob.Problem((48, 36), (1.2, 0.9)); pb.solid = cyl_mask(pb, 0.5, 0.45, 0.12)
        pb.lap = shared_cases.separable_lap(pb)
        out.append((f"g6_2d_48x36_separable_varlap_{tag}", pb, 1e3 * gaussian(pb), 600, 1e-6, prec))
    return out


def write_g5(groups=("g5",)):
    todo = (g5_cases() if "g5" in groups else []) + (g6_cases() if "g6" in groups else [])
    for name, pb, rhs, mi, rel, prec in todo:
        tol = rel * np.linalg.norm(rhs[pb.fluid_interior_mask()])    # absolute tolerance, relres rel
        p = ob.ref_pcg_ex(pb, rhs, mi, tol, precond=prec)
        d = dict(n=np.array(pb.n), L=np.array(pb.L), solid=pb.solid.copy(), rhs=rhs, p=p, max_iter=np.array(mi),
                 tol=np.array(tol), precond=np.array(prec))
        if pb.lap is not None:
            d.update(lap_diag=pb.lap[0], lap_off_x=pb.lap[1], lap_off_y=pb.lap[2])
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kind=np.array("pcg_ex"), **d)
        print("wrote", name)


def write_g7():
    """VtkWriter / VtkWriter3D frames (src/io/vtk_writer.cpp, vtk_writer_3d.cpp)
    of seeded MAC fields with a random solid mask; the fixture holds the
    inputs and the file the reference wrote, as bytes."""
    import tempfile
    for name, n, L, seed in (("g7_vtk_2d_6x5", (6, 5), (1.0, 0.8), 1), ("g7_vtk_3d_4x3x3", (4, 3, 3), (1.0, 0.75, 0.6), 2)):
        rng = np.random.default_rng(seed)
        shp_u = [n[0] + 1] + [v + 2 for v in n[1:]]
        shp_v = [n[0] + 2, n[1] + 1] + [v + 2 for v in n[2:]]
        u = rng.standard_normal(int(np.prod(shp_u)))
        v = rng.standard_normal(int(np.prod(shp_v)))
        w = rng.standard_normal((n[0] + 2) * (n[1] + 2) * (n[2] + 1)) if len(n) == 3 else None
        solid = (rng.random(int(np.prod([x + 2 for x in n]))) < 0.2).astype(np.uint8)
        with tempfile.TemporaryDirectory() as d:
            ob.ref_write_vtk(n, L, u, v, w, solid, 12, d)
            blob = open(os.path.join(d, "frame_00012.vtk"), "rb").read()
        extra = dict(w=w) if w is not None else {}
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kind=np.array("vtk"), n=np.array(n), L=np.array(L),
                            u=u, v=v, solid=solid, frame=np.array(12), vtk=np.frombuffer(blob, np.uint8), **extra)
        print("wrote", name, len(blob), "bytes")


def main(groups):
    assert ob.ref_available(), "build oracle/_ref first: make -C oracle ref"
    os.makedirs(HERE, exist_ok=True)
    if "g7" in groups:
        write_g7()
    if {"g5", "g6"} & set(groups):
        write_g5(groups)
    if not ({"g1", "g2", "g3", "g4"} & set(groups)):
        return
    index = []
    for name, kind, pb, arg in cases():
        d = dict(n=np.array(pb.n), L=np.array(pb.L), solid=pb.solid.copy())
        if kind == "apply":
            z = ob.ref_uaamg_apply(pb, arg)
            d.update(r=arg, z=z)
        else:
            rhs, mi, tol = arg
            # absolute tolerance scaled by |r0| (relres), as the reference PCG tests on |r|
            p, applies = ob.ref_pcg(pb, rhs, mi, tol * np.linalg.norm(rhs[pb.fluid_interior_mask()]) if tol > 1e-9 else tol)
            d.update(rhs=rhs, p=p, applies=np.array(applies), max_iter=np.array(mi),
                     tol=np.array(tol * np.linalg.norm(rhs[pb.fluid_interior_mask()]) if tol > 1e-9 else tol))
        np.savez_compressed(os.path.join(HERE, name + ".npz"), kind=np.array(kind), **d)
        index.append(name)
        print("wrote", name)
    # G4: standalone amgpcg (test_amgpcg.cpp:77-85 inputs, 2 pi^2 sin sin) and iteration counts
    for N in (32, 64, 128, 256):
        b = np.zeros((N + 2) ** 2)
        h = 1.0 / (N + 1)
        v = b.reshape(N + 2, N + 2)
        for i in range(1, N + 1):
            for j in range(1, N + 1):
                v[i, j] = 2 * np.pi ** 2 * np.sin(np.pi * i * h) * np.sin(np.pi * j * h)
        x, it = ob.ref_amgpcg(N, b, 30, 1e-8)
        d = dict(N=np.array(N), iters=np.array(it), max_iter=np.array(30), tol=np.array(1e-8))
        if N <= 64:
            d.update(b=b, x=x)
        np.savez_compressed(os.path.join(HERE, f"g4_amgpcg_sin_N{N}.npz"), kind=np.array("amgpcg"), **d)
        print("wrote amgpcg", N, it)


if __name__ == "__main__":
    # python make_golden.py [g1 g2 g3 g4 g5 g6 g7]  (default: all groups)
    main(sys.argv[1:] or ["g1", "g2", "g3", "g4", "g5", "g6", "g7"])
