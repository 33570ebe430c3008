"""Seeded problem generators shared by the parity tests (small sizes)."""
import numpy as np

import oracle_bind as ob


def coords(pb):
    """Cell-centre coordinates (padded arrays, [k][j][i] or [j][i] order)."""
    idx = np.meshgrid(*[np.arange(s) for s in pb.shape[::-1]], indexing="ij")
    return [(idx[len(idx) - 1 - d] - 0.5) * pb.h[d] for d in range(pb.dim)]


def balls(pb, centers, radii):
    X = coords(pb)
    m = np.zeros(pb.shape[::-1], bool)
    for c, r in zip(centers, radii):
        d2 = sum((X[q] - c[q]) ** 2 for q in range(pb.dim))
        m |= d2 < r * r
    return m.reshape(-1).astype(np.uint8)


def random_solids(pb, rng, frac=0.15, nobj=6):
    X = coords(pb)
    m = np.zeros(pb.shape[::-1], bool)
    for _ in range(nobj):
        if rng.random() < 0.5:
            c = [rng.uniform(0, pb.L[q]) for q in range(pb.dim)]
            r = rng.uniform(0.05, 0.2) * min(pb.L)
            m |= sum((X[q] - c[q]) ** 2 for q in range(pb.dim)) < r * r
        else:
            lo = [rng.uniform(0, pb.L[q]) for q in range(pb.dim)]
            ext = [rng.uniform(0.05, 0.3) * pb.L[q] for q in range(pb.dim)]
            box = np.ones_like(m)
            for q in range(pb.dim):
                box &= (X[q] >= lo[q]) & (X[q] <= lo[q] + ext[q])
            m |= box
        if m.mean() > frac:
            break
    return m.reshape(-1).astype(np.uint8)


def gaussian_rhs(pb):
    """exp(-0.01 |x - c|^2) in index units (test_cuda_uaamg.cu:70-75)."""
    idx = np.meshgrid(*[np.arange(s) for s in pb.shape[::-1]], indexing="ij")
    d2 = 0
    for d in range(pb.dim):
        ii = idx[len(idx) - 1 - d]
        d2 = d2 + (ii - pb.n[d] / 2.0) ** 2
    out = np.zeros(pb.N)
    pb.interior(out)[...] = pb.interior(np.exp(-0.01 * d2).reshape(-1))
    return out


def random_rhs(pb, rng):
    out = np.zeros(pb.N)
    pb.interior(out)[...] = rng.standard_normal(pb.interior(out).shape)
    return out


def parity_cases():
    """(name, Problem) pairs: the reference's own test geometries plus odd,
    anisotropic and solid-heavy grids."""
    rng = np.random.default_rng(1234)
    out = []
    out.append(("2d_32x16", ob.Problem((32, 16))))
    pb = ob.Problem((64, 16), (4.0, 1.0))
    pb.solid = balls(pb, [(1.0, 0.5)], [0.1])
    out.append(("2d_64x16_cylinder", pb))
    out.append(("2d_33x17_odd", ob.Problem((33, 17), (1.0, 0.7))))
    pb = ob.Problem((48, 40), (1.0, 40.0 / 30.0))
    pb.solid = random_solids(pb, rng)
    out.append(("2d_48x40_aniso_solids", pb))
    out.append(("3d_16x8x4", ob.Problem((16, 8, 4))))
    pb = ob.Problem((16, 16, 16))
    pb.solid = balls(pb, [(0.5, 0.5, 0.5)], [0.25])
    out.append(("3d_16_sphere", pb))
    out.append(("3d_17x9x5_odd", ob.Problem((17, 9, 5), (1.0, 0.9, 0.8))))
    pb = ob.Problem((24, 20, 18), (1.0, 1.1, 0.7))
    pb.solid = random_solids(pb, rng)
    out.append(("3d_24x20x18_aniso_solids", pb))
    # several workgroup tiles in y and z, mixed Dirichlet / Neumann walls
    pb = ob.Problem((40, 36, 34), (1.0, 0.9, 0.85),
                    wall=[ob.DIRICHLET, ob.NEUMANN, ob.NEUMANN, ob.DIRICHLET, ob.DIRICHLET, ob.NEUMANN])
    pb.solid = random_solids(pb, rng)
    out.append(("3d_40x36x34_walls_solids", pb))
    # x-rows longer than 128 cells (partly filled last lanes)
    pb = ob.Problem((130, 20, 12), (1.3, 0.2, 0.12))
    pb.solid = balls(pb, [(0.4, 0.1, 0.06)], [0.05])
    out.append(("3d_130x20x12", pb))
    out.append(("2d_40x30_dirichlet", ob.Problem((40, 30), (1.0, 0.75), wall=[ob.DIRICHLET] * 6)))
    # level 1 wider than 64 cells: the coarse level runs on the record row
    # kernels (several y tiles and z chunks, walls baked into the records)
    pb = ob.Problem((200, 70, 40), (1.0, 0.35, 0.2),
                    wall=[ob.NEUMANN, ob.DIRICHLET, ob.NEUMANN, ob.NEUMANN, ob.DIRICHLET, ob.NEUMANN])
    pb.solid = random_solids(pb, rng, frac=0.2, nobj=10)
    out.append(("3d_200x70x40_walls_solids", pb))
    # rows of exactly 256 cells (4 per lane, every lane inside the row): the
    # level-0 kernels without lane masks, with the Dirichlet excess on the
    # first and the last lane's cells
    pb = ob.Problem((256, 20, 12), (1.0, 0.1, 0.05),
                    wall=[ob.DIRICHLET, ob.DIRICHLET, ob.DIRICHLET, ob.NEUMANN, ob.NEUMANN, ob.DIRICHLET])
    pb.solid = balls(pb, [(0.3, 0.05, 0.025), (0.7, 0.04, 0.02)], [0.03, 0.02])
    out.append(("3d_256x20x12_walls", pb))
    # porous: 40 % random solid cells, many fluid cells with no fluid neighbour
    pb = ob.Problem((20, 18, 16))
    m = np.zeros(pb.shape[::-1], np.uint8)
    m[1:-1, 1:-1, 1:-1] = (rng.random((16, 18, 20)) < 0.4)
    pb.solid = m.reshape(-1)
    out.append(("3d_20x18x16_porous", pb))
    return out


def neumann_cases():
    """parity_cases() with all walls Neumann (the reference's own setting)."""
    return [(n, pb) for n, pb in parity_cases() if all(w == ob.NEUMANN for w in pb.wall)]


def mac_fields(pb, rng, closed=True):
    """Random MAC velocities u, v[, w] in the reference layouts (mesh.h:27-35,
    mesh_3d.h:29-37).  closed: zero normal velocity on the box walls and on
    faces touching a solid cell, so every fluid region has zero net flux and
    the projected field can be divergence-free."""
    nx, ny = pb.n[0], pb.n[1]
    nz = pb.n[2] if pb.dim == 3 else 1
    kk = nz + 2 if pb.dim == 3 else 1
    s = pb.solid.reshape(pb.shape[::-1]).astype(bool)          # [k][j][i] or [j][i]
    if pb.dim == 2:
        s = s[None]
    u = rng.standard_normal((kk, ny + 2, nx + 1))
    v = rng.standard_normal((kk, ny + 1, nx + 2))
    w = rng.standard_normal((nz + 1, ny + 2, nx + 2)) if pb.dim == 3 else None
    if closed:
        u[:, :, 0] = 0; u[:, :, nx] = 0
        v[:, 0, :] = 0; v[:, ny, :] = 0
        u[:, :, :] *= ~(s[:, :, :-1] | s[:, :, 1:])            # face i between cells i and i+1
        v[:, :, :] *= ~(s[:, :-1, :] | s[:, 1:, :])
        if w is not None:
            w[0] = 0; w[nz] = 0
            w[:, :, :] *= ~(s[:-1] | s[1:])
    return u.reshape(-1), v.reshape(-1), (w.reshape(-1) if w is not None else None)


def mac_divergence(pb, u, v, w):
    """fvc::divergence on every interior cell (operators.h:18-21, operators_3d.h:17-21), padded array."""
    nx, ny = pb.n[0], pb.n[1]
    nz = pb.n[2] if pb.dim == 3 else 1
    kk = nz + 2 if pb.dim == 3 else 1
    U = u.reshape(kk, ny + 2, nx + 1)
    V = v.reshape(kk, ny + 1, nx + 2)
    d = np.zeros((kk, ny + 2, nx + 2))
    ks = slice(1, -1) if pb.dim == 3 else slice(0, 1)
    d[ks, 1:-1, 1:-1] = (U[ks, 1:-1, 1:] - U[ks, 1:-1, :-1]) / pb.h[0] + (V[ks, 1:, 1:-1] - V[ks, :-1, 1:-1]) / pb.h[1]
    if pb.dim == 3:
        W = w.reshape(nz + 1, ny + 2, nx + 2)
        d[1:-1, 1:-1, 1:-1] += (W[1:, 1:-1, 1:-1] - W[:-1, 1:-1, 1:-1]) / pb.h[2]
    return d.reshape(-1)


def separable_lap(pb):
    """Variable Laplacian coefficients (Grid::lap_diag / lap_off_x /
    lap_off_y, grid.h:21-23) whose x couplings depend on y only and y
    couplings on x only: with the matvec's one off-diagonal per axis
    (pcg.cpp:44) the operator is symmetric with zero row sums
    (diag = -2 off_x - 2 off_y), so PCG converges on it."""
    import numpy as np
    j, i = np.meshgrid(np.arange(pb.n[1] + 2), np.arange(pb.n[0] + 2), indexing="ij")
    x, y = (i - 0.5) * pb.h[0], (j - 0.5) * pb.h[1]
    ox = -(1.0 + 0.5 * np.sin(2 * np.pi * y / pb.L[1])) / pb.h[0] ** 2
    oy = -(1.0 + 0.4 * np.cos(np.pi * x / pb.L[0])) / pb.h[1] ** 2
    ox, oy = ox.reshape(-1), oy.reshape(-1)
    return (-2 * ox - 2 * oy, ox, oy)
