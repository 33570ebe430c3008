"""Seeded synthetic pressure-solve inputs (host side, numpy).

Every generator returns a `Problem`: padded arrays in the reference layout
(include/core/mesh.h:32-35: idx = i + j*(nx+2) [+ k*(nx+2)*(ny+2)], i fastest),
spacing h = L/n (mesh.h:21-22) and per-face wall types.  The right-hand side
follows the reference's pressure projection, rhs = div(u)/dt on fluid cells
(src/pressure/pressure.cpp:10-14, include/ops/operators.h:18-21), made
zero-mean over fluid cells where the walls are all Neumann.

Workloads (sizes as named):
  pressure2d_mac   1024x512 MAC grid, dx=1/1024, dy=1/480, Neumann, solid disc
  plume3d          n^3 (n=256), Neumann, floor slab + 8 spheres, buoyant-plume rhs
  poisson2d_dirichlet  256^2 unit square, Dirichlet p=0, 2 pi^2 sin sin + noise
  batch_mixed      64 2D + 32 3D grids of uneven size, masks and walls
"""
import numpy as np

NEUMANN, DIRICHLET = 0, 1


class Problem:
    def __init__(self, n, L, solid, rhs, wall=None, name=""):
        self.n = tuple(int(v) for v in n)
        self.dim = len(self.n)
        self.L = tuple(float(v) for v in L)
        self.h = tuple(self.L[d] / self.n[d] for d in range(self.dim))
        self.shape = tuple(v + 2 for v in self.n)
        self.N = int(np.prod(self.shape))
        self.solid = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        self.rhs = np.ascontiguousarray(rhs, dtype=np.float64).reshape(-1)
        self.wall = list(wall) if wall is not None else [NEUMANN] * 6
        self.name = name
        assert self.solid.size == self.N and self.rhs.size == self.N

    @property
    def cells(self):
        return int(np.prod(self.n))

    def view(self, a):
        return a.reshape(self.shape[::-1])

    def fluid_mask(self):
        m = np.zeros(self.shape[::-1], bool)
        m[(slice(1, -1),) * self.dim] = True
        return m.reshape(-1) & (self.solid == 0)


def _centres(n, h):
    """Cell-centre coordinates of the padded grid along one axis: (i-0.5)*h, i=0..n+1."""
    return (np.arange(n + 2) - 0.5) * h


def _interior(shape_rev):
    return (slice(1, -1),) * len(shape_rev)


def _zero_mean(rhs, fluid):
    rhs[fluid] -= rhs[fluid].mean()
    rhs[~fluid] = 0.0
    return rhs


def pressure2d_mac(seed=0, nx=1024, ny=512, dt=1e-3, nvort=32):
    """MAC-grid pressure projection: Lx = 1, Ly = ny/480 so dx = 1/1024, dy = 1/480;
    Neumann walls; solid disc r < 0.05 at (0.2, 0.5); velocity from the stream
    function sum_v A exp(-|x-c|^2/s^2), A~U(-1,1), s~U(0.01,0.05), c~U(domain),
    sampled on faces; solid and wall faces zero; rhs = div(u)/dt, zero mean."""
    rng = np.random.default_rng(seed)
    Lx, Ly = 1.0, ny / 480.0
    dx, dy = Lx / nx, Ly / ny
    xc, yc = _centres(nx, dx), _centres(ny, dy)
    Y, X = np.meshgrid(yc, xc, indexing="ij")
    solid = ((X - 0.2) ** 2 + (Y - 0.5) ** 2 < 0.05 ** 2)
    A = rng.uniform(-1, 1, nvort)
    S = rng.uniform(0.01, 0.05, nvort)
    C = np.stack([rng.uniform(0, Lx, nvort), rng.uniform(0, Ly, nvort)], 1)
    # u at x-faces (i*dx, (j-0.5)dy), i = 0..nx ; v at y-faces ((i-0.5)dx, j*dy), j = 0..ny
    xf = np.arange(nx + 1) * dx
    yf = np.arange(ny + 1) * dy
    Yu, Xu = np.meshgrid(yc, xf, indexing="ij")     # (ny+2, nx+1)
    Yv, Xv = np.meshgrid(yf, xc, indexing="ij")     # (ny+1, nx+2)
    u = np.zeros_like(Xu)
    v = np.zeros_like(Xv)
    for a, s, c in zip(A, S, C):
        eu = np.exp(-((Xu - c[0]) ** 2 + (Yu - c[1]) ** 2) / s ** 2)
        u += a * (-2.0 * (Yu - c[1]) / s ** 2) * eu
        ev = np.exp(-((Xv - c[0]) ** 2 + (Yv - c[1]) ** 2) / s ** 2)
        v += a * (2.0 * (Xv - c[0]) / s ** 2) * ev
    u[:, 0] = 0.0; u[:, nx] = 0.0
    v[0, :] = 0.0; v[ny, :] = 0.0
    su = solid[:, :-1] | solid[:, 1:]          # face between cells i and i+1
    sv = solid[:-1, :] | solid[1:, :]
    u[su] = 0.0
    v[sv] = 0.0
    rhs = np.zeros((ny + 2, nx + 2))
    div = (u[1:-1, 1:] - u[1:-1, :-1]) / dx + (v[1:, 1:-1] - v[:-1, 1:-1]) / dy
    rhs[1:-1, 1:-1] = div / dt
    sol = solid.astype(np.uint8)
    pb = Problem((nx, ny), (Lx, Ly), sol.reshape(-1), rhs.reshape(-1), name="pressure2d_mac_1024x512")
    _zero_mean(pb.rhs, pb.fluid_mask())
    return pb


def _upsample_trilinear(coarse, n):
    """Trilinear upsampling of a (m+1)^3 node lattice on [0,1]^3 to n^3 cell centres."""
    m = coarse.shape[0] - 1
    t = (np.arange(n) + 0.5) / n * m
    i0 = np.clip(np.floor(t).astype(int), 0, m - 1)
    w = t - i0
    out = coarse
    for ax in range(3):
        a = np.take(out, i0, axis=ax)
        b = np.take(out, i0 + 1, axis=ax)
        shape = [1, 1, 1]
        shape[ax] = n
        ww = w.reshape(shape)
        out = a * (1 - ww) + b * ww
    return out


def plume3d(seed=0, n=256, nspheres=8, velocity=False):
    """n^3 cells, dx = 1/n, Neumann walls; solid floor slab y < 0.1 and
    `nspheres` spheres (r ~ U(0.03, 0.1), centres ~ U(0,1)^3); velocity
    T(x) curl(Psi), T a Gaussian temperature blob, Psi a 16^3 random vector
    field trilinearly upsampled; rhs = div(u) on fluid cells, zero mean.
    velocity: also keep the MAC face velocities as pb.u, pb.v, pb.w in the
    reference layouts (mesh_3d.h:29-37), ghost rows zero."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    c = _centres(n, h)
    # solid mask on the padded grid, [k][j][i] = (z, y, x)
    solid = np.zeros((n + 2, n + 2, n + 2), bool)
    solid |= (c[None, :, None] < 0.1)
    cen = rng.uniform(0, 1, (nspheres, 3))
    rad = rng.uniform(0.03, 0.1, nspheres)
    for (x0, y0, z0), r in zip(cen, rad):
        solid |= ((c[None, None, :] - x0) ** 2 + (c[None, :, None] - y0) ** 2 + (c[:, None, None] - z0) ** 2) < r * r
    psi = [_upsample_trilinear(rng.standard_normal((17, 17, 17)), n) for _ in range(3)]   # [z][y][x] order
    # curl on cell centres (central differences), psi = (px, py, pz)
    px, py, pz = psi[0], psi[1], psi[2]
    g = lambda f, ax: np.gradient(f, h, axis=ax)
    vx = g(pz, 1) - g(py, 0)
    vy = g(px, 0) - g(pz, 2)
    vz = g(py, 2) - g(px, 1)
    del psi, px, py, pz
    ci = c[1:-1]
    T = np.exp(-(((ci[None, None, :] - 0.5) ** 2 + (ci[None, :, None] - 0.3) ** 2 + (ci[:, None, None] - 0.5) ** 2) / 0.15 ** 2))
    vx *= T; vy *= T; vz *= T
    del T
    sol_in = solid[1:-1, 1:-1, 1:-1]
    # face velocities: average of neighbouring centres; wall and solid faces zero
    ux = np.zeros((n, n, n + 1)); ux[:, :, 1:-1] = 0.5 * (vx[:, :, 1:] + vx[:, :, :-1])
    ux[:, :, 1:-1][sol_in[:, :, 1:] | sol_in[:, :, :-1]] = 0.0
    div = (ux[:, :, 1:] - ux[:, :, :-1]) / h
    mac = {}
    if velocity:
        U = np.zeros((n + 2, n + 2, n + 1)); U[1:-1, 1:-1, :] = ux
        mac["u"] = U.reshape(-1)
    del ux, vx
    uy = np.zeros((n, n + 1, n)); uy[:, 1:-1, :] = 0.5 * (vy[:, 1:, :] + vy[:, :-1, :])
    uy[:, 1:-1, :][sol_in[:, 1:, :] | sol_in[:, :-1, :]] = 0.0
    div += (uy[:, 1:, :] - uy[:, :-1, :]) / h
    if velocity:
        V = np.zeros((n + 2, n + 1, n + 2)); V[1:-1, :, 1:-1] = uy
        mac["v"] = V.reshape(-1)
    del uy, vy
    uz = np.zeros((n + 1, n, n)); uz[1:-1, :, :] = 0.5 * (vz[1:, :, :] + vz[:-1, :, :])
    uz[1:-1, :, :][sol_in[1:, :, :] | sol_in[:-1, :, :]] = 0.0
    div += (uz[1:, :, :] - uz[:-1, :, :]) / h
    if velocity:
        W = np.zeros((n + 1, n + 2, n + 2)); W[:, 1:-1, 1:-1] = uz
        mac["w"] = W.reshape(-1)
    del uz, vz
    rhs = np.zeros((n + 2, n + 2, n + 2))
    rhs[1:-1, 1:-1, 1:-1] = div
    del div
    pb = Problem((n, n, n), (1.0, 1.0, 1.0), solid.astype(np.uint8).reshape(-1), rhs.reshape(-1),
                 name=f"plume3d_{n}^3")
    _zero_mean(pb.rhs, pb.fluid_mask())
    for k, a in mac.items():
        setattr(pb, k, a)
    return pb


def poisson2d_dirichlet(seed=0, n=256, noise=1e-3):
    """Unit square, n x n cells (h = 1/n), Dirichlet p = 0 on every wall,
    rhs = 2 pi^2 sin(pi x) sin(pi y) + noise * N(0,1); lap(p) = rhs has the
    analytic solution p = -sin(pi x) sin(pi y) up to the noise."""
    rng = np.random.default_rng(seed)
    h = 1.0 / n
    c = _centres(n, h)
    Y, X = np.meshgrid(c, c, indexing="ij")
    rhs = 2 * np.pi ** 2 * np.sin(np.pi * X) * np.sin(np.pi * Y)
    rhs[1:-1, 1:-1] += noise * rng.standard_normal((n, n))
    rhs[0, :] = rhs[-1, :] = 0.0
    rhs[:, 0] = rhs[:, -1] = 0.0
    exact = -np.sin(np.pi * X) * np.sin(np.pi * Y)
    pb = Problem((n, n), (1.0, 1.0), np.zeros((n + 2) ** 2, np.uint8), rhs.reshape(-1),
                 wall=[DIRICHLET] * 6, name=f"poisson2d_dirichlet_{n}^2")
    pb.exact = exact.reshape(-1)
    return pb


def _smoothstep(t):
    s = (t + 1.0) * 0.5
    return np.where(t <= -1.0, 0.0, np.where(t >= 1.0, 1.0, s * s * (3.0 - 2.0 * s)))


def smooth_cylinder_lap(n, L, centre, R, tw=0.0):
    """Variable Laplacian coefficients of a smooth cylinder, as
    SmoothCylinder::applyToLaplacian writes them into Grid::lap_diag,
    lap_off_x, lap_off_y (src/geometry/cylinder_model.cpp:50-71, the
    solid fraction of :40-43 and smoothstep of :5-10): cells with solid
    fraction > 0.9 become solid with diagonal 1e15, the others get face
    permeabilities 1 - (phi + phi_nb)/2 + 0.1.  Returns (solid, diag, off_x,
    off_y), padded 2D arrays (i fastest)."""
    nx, ny = int(n[0]), int(n[1])
    dx, dy = L[0] / nx, L[1] / ny
    tw = tw if tw > 0 else dx
    cx, cy = float(centre[0]), float(centre[1])
    eps, large = 0.1, 1e15

    def frac(x, y):
        return _smoothstep((np.sqrt((x - cx) * (x - cx) + (y - cy) * (y - cy)) - R) / tw)

    j, i = np.meshgrid(np.arange(1, ny + 1), np.arange(1, nx + 1), indexing="ij")
    xc, yc = (i - 0.5) * dx, (j - 0.5) * dy
    phi = frac(xc, yc)
    pL, pR, pB, pT = frac(xc - dx, yc), frac(xc + dx, yc), frac(xc, yc - dy), frac(xc, yc + dy)
    cL = 1.0 - 0.5 * (phi + pL) + eps
    cR = 1.0 - 0.5 * (phi + pR) + eps
    cB = 1.0 - 0.5 * (phi + pB) + eps
    cT = 1.0 - 0.5 * (phi + pT) + eps
    ix, iy = 1.0 / (dx * dx), 1.0 / (dy * dy)
    sol = phi > 0.9
    shape = (ny + 2, nx + 2)
    out = [np.zeros(shape, np.uint8), np.zeros(shape), np.zeros(shape), np.zeros(shape)]
    inner = (slice(1, -1), slice(1, -1))
    out[0][inner] = sol
    out[1][inner] = np.where(sol, large, (cL + cR) * ix + (cB + cT) * iy)
    out[2][inner] = np.where(sol, 0.0, -0.5 * (cL + cR) * ix)
    out[3][inner] = np.where(sol, 0.0, -0.5 * (cB + cT) * iy)
    return tuple(a.reshape(-1) for a in out)


def _random_mask(rng, n, frac_max=0.3):
    """Random boxes/spheres covering up to frac_max of the cells (padded, [k][j][i])."""
    dim = len(n)
    shape = tuple(v + 2 for v in n)[::-1]
    axes = [(np.arange(s) - 0.5) / (s - 2) for s in shape]
    grids = np.meshgrid(*axes, indexing="ij", sparse=True)   # broadcastable; grids[0] is the slowest axis
    m = np.zeros(shape, bool)
    target = rng.uniform(0.0, frac_max)
    for _ in range(64):
        if m[(slice(1, -1),) * dim].mean() >= target:
            break
        if rng.random() < 0.5:
            cc = rng.uniform(0, 1, dim)
            r = rng.uniform(0.03, 0.15)
            m |= sum((grids[q] - cc[q]) ** 2 for q in range(dim)) < r * r
        else:
            lo = rng.uniform(0, 1, dim)
            ext = rng.uniform(0.05, 0.25, dim)
            box = np.ones(shape, bool)
            for q in range(dim):
                box &= (grids[q] >= lo[q]) & (grids[q] <= lo[q] + ext[q])
            m |= box
    return m.astype(np.uint8).reshape(-1)


def _vortex_rhs(rng, n, solid, nvort=16):
    """Divergence of a random vortex field on the faces, zero on solid faces."""
    dim = len(n)
    shape = tuple(v + 2 for v in n)[::-1]
    s = solid.reshape(shape).astype(bool)
    axes = [(np.arange(sh) - 0.5) / (sh - 2) for sh in shape]
    grids = np.meshgrid(*axes, indexing="ij", sparse=True)
    vel = [np.zeros(shape) for _ in range(dim)]
    for _ in range(nvort):
        cc = rng.uniform(0, 1, dim)
        sg = rng.uniform(0.03, 0.15)
        a = rng.uniform(-1, 1)
        e = 1.0
        for q in range(dim):            # separable Gaussian: a product of 1D factors
            e = e * np.exp(-(grids[q] - cc[q]) ** 2 / sg ** 2)
        # rotation in the plane of the two fastest axes, plus a source term so div != 0
        vel[-1] += a * (grids[-2] - cc[-2]) * e
        vel[-2] -= a * (grids[-1] - cc[-1]) * e
        vel[-1] += 0.3 * a * (grids[-1] - cc[-1]) * e
    rhs = np.zeros(shape)
    inner = (slice(1, -1),) * dim
    for q in range(dim):
        v = vel[q]
        hq = 1.0 / n[dim - 1 - q]
        lo = [slice(1, -1)] * dim; hi = [slice(1, -1)] * dim
        hi[q] = slice(2, None); lo[q] = slice(0, -2)
        fp = 0.5 * (v[inner] + v[tuple(hi)])
        fm = 0.5 * (v[inner] + v[tuple(lo)])
        fp = np.where(s[inner] | s[tuple(hi)], 0.0, fp)
        fm = np.where(s[inner] | s[tuple(lo)], 0.0, fm)
        rhs[inner] += (fp - fm) / hq
    return rhs.reshape(-1)


def batch_mixed(seed=0, n2d=64, n3d=32, side2d=(64, 1024), side3d=(32, 160)):
    """Independent solves of uneven size: n2d square 2D grids with side
    log-uniform in side2d and n3d cubic 3D grids with side log-uniform in
    side3d; random solid masks (0-30 %), random per-face Neumann/Dirichlet
    walls, zero-mean random-vortex rhs."""
    rng = np.random.default_rng(seed)
    out = []
    specs = [(2, side2d)] * n2d + [(3, side3d)] * n3d
    for dim, (lo, hi) in specs:
        side = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        n = (side,) * dim
        solid = _random_mask(rng, n)
        wall = [int(rng.random() < 0.25) for _ in range(6)]
        rhs = _vortex_rhs(rng, n, solid)
        pb = Problem(n, (1.0,) * dim, solid, rhs, wall=wall, name=f"batch_{dim}d_{side}")
        _consistent_rhs(pb)
        out.append(pb)
    return out


def _consistent_rhs(pb):
    """Make the rhs solvable: zero on non-fluid cells and zero-mean on every
    connected fluid region that touches no Dirichlet wall (a random mask can
    cut off pockets of fluid; each such pocket carries its own constant)."""
    from scipy import ndimage
    fl = pb.fluid_mask()
    F = pb.view(fl)
    lab, nlab = ndimage.label(F)
    rhs = pb.rhs
    rhs[~fl] = 0.0
    if nlab == 0:
        return
    L = lab.reshape(-1)
    touches = np.zeros(nlab + 1, bool)
    V = lab
    for f in range(2 * pb.dim):
        if pb.wall[f] != DIRICHLET:
            continue
        ax = pb.dim - 1 - f // 2                     # face -x,+x,-y,+y,-z,+z -> array axis ([k][j][i])
        idx = [slice(None)] * pb.dim
        idx[ax] = 1 if f % 2 == 0 else V.shape[ax] - 2
        touches[np.unique(V[tuple(idx)])] = True
    cnt = np.bincount(L, minlength=nlab + 1).astype(float)
    sm = np.bincount(L, weights=rhs, minlength=nlab + 1)
    mean = np.where(cnt > 0, sm / np.maximum(cnt, 1), 0.0)
    mean[0] = 0.0
    mean[touches] = 0.0
    rhs -= mean[L]
    rhs[~fl] = 0.0


# ---- served stream: many concurrent simulations, one pressure solve per frame ----

def serve_sims(nsim, seed, dev):
    """Simulations of the served stream: half 2D (side log-uniform 64..2048),
    half 3D (side log-uniform 32..256), all walls Neumann, up to 3 disjoint
    solid balls away from the walls (the fluid stays connected), generated on
    the device."""
    import torch
    rng = np.random.default_rng(seed)
    sims = []
    for k in range(nsim):
        dim = 2 if k % 2 == 0 else 3
        lo, hi = (64, 2048) if dim == 2 else (32, 256)
        side = int(round(np.exp(rng.uniform(np.log(lo), np.log(hi)))))
        n = (side,) * dim
        shape = tuple(v + 2 for v in n)[::-1]
        ax = [((torch.arange(s, device=dev, dtype=torch.float32) - 0.5) / (s - 2)) for s in shape]
        X = torch.meshgrid(*ax, indexing="ij")
        solid = torch.zeros(shape, dtype=torch.bool, device=dev)
        balls = []
        for _ in range(int(rng.integers(0, 4))):
            for _try in range(20):
                r = rng.uniform(0.04, 0.1)
                c = rng.uniform(r + 0.05, 1 - r - 0.05, dim)
                if all(np.linalg.norm(c - c2) > r + r2 + 0.05 for c2, r2 in balls):
                    balls.append((c, r))
                    break
        for c, r in balls:
            d2 = sum((X[q] - float(c[dim - 1 - q])) ** 2 for q in range(dim))
            solid |= d2 < r * r
        inner = (slice(1, -1),) * dim
        fluid = torch.zeros(shape, dtype=torch.bool, device=dev)
        fluid[inner] = True
        fluid &= ~solid
        # smooth rhs: a few Gaussian sources and sinks, zero mean over the fluid
        rhs = torch.zeros(shape, dtype=torch.float64, device=dev)
        for _ in range(6):
            c = rng.uniform(0.1, 0.9, dim)
            s = rng.uniform(0.05, 0.2)
            rhs += rng.uniform(-1e3, 1e3) * torch.exp(-sum((X[q].double() - float(c[dim - 1 - q])) ** 2 for q in range(dim)) / s ** 2)
        rhs[~fluid] = 0
        rhs[fluid] -= rhs[fluid].mean()
        sims.append({"n": n, "h": tuple(1.0 / v for v in n), "solid": solid.reshape(-1).to(torch.uint8).cpu().numpy(),
                     "fluid": fluid.reshape(-1), "rhs": rhs.reshape(-1), "shape": shape, "dim": dim})
    return sims


def next_frame(sim, rng):
    """Frame t+1: frame t's rhs advected by a smooth (one-cell) shift plus 5 % new noise."""
    import torch
    r = sim["rhs"].reshape(sim["shape"])
    ax = int(rng.integers(0, sim["dim"]))
    r = torch.roll(r, shifts=int(rng.choice([-1, 1])), dims=ax)
    gen = torch.Generator(device=r.device)
    gen.manual_seed(int(rng.integers(0, 2 ** 62)))
    noise = torch.randn(r.shape, dtype=torch.float64, device=r.device, generator=gen)
    noise = noise * (0.05 * float(r.abs().max()))
    r = r + noise
    r = r.reshape(-1)
    f = sim["fluid"]
    r[~f] = 0
    r[f] -= r[f].mean()
    sim["rhs"] = r
    return r
