"""ctypes bindings to the CPU checker libraries (test infrastructure).

oracle/liboracle.so  -- the C restatement of the reference solve (the oracle).
oracle/_ref/libref.so -- the reference's own CPU solver compiled from its
sources (present only where the reference tree was available at build time).
"""
import ctypes as C
import os
import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libref.so")

NEUMANN, DIRICHLET = 0, 1


class OrcGrid(C.Structure):
    _fields_ = [("dim", C.c_int), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("dx", C.c_double), ("dy", C.c_double), ("dz", C.c_double),
                ("solid", C.c_void_p), ("wall", C.c_int * 6),
                ("lap_diag", C.c_void_p), ("lap_off_x", C.c_void_p), ("lap_off_y", C.c_void_p)]


class OrcOpts(C.Structure):
    _fields_ = [("nu_pre", C.c_int), ("nu_post", C.c_int), ("coarsest_side", C.c_int),
                ("coarsest_sweeps", C.c_int), ("tol_mode", C.c_int), ("project_mean", C.c_int),
                ("precond", C.c_int), ("hierarchy_mode", C.c_int)]


PRECOND_UAAMG, PRECOND_IDENTITY, PRECOND_GMG, PRECOND_AMG = 0, 1, 2, 3


class OrcResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("applies", C.c_int),
                ("rnorm", C.c_double), ("r0norm", C.c_double)]


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(ORACLE_SO)
        P = C.c_void_p
        _lib.orc_default_opts.argtypes = [P]
        _lib.orc_pcg.argtypes = [P, P, P, P, C.c_int, C.c_double, P]
        _lib.orc_uaamg_apply.argtypes = [P, P, P, P]
        _lib.orc_precond_apply.argtypes = [P, P, P, P]
        _lib.orc_hierarchy_dims.argtypes = [P, P, P, P, C.c_int]
        _lib.orc_level_data.argtypes = [P, P, C.c_int, P, P, P, P, P, P]
        _lib.orc_matvec.argtypes = [P, P, P]
        _lib.orc_amgpcg.argtypes = [C.c_int, P, P, C.c_int, C.c_double, P]
        _lib.orc_project.argtypes = [P, P, P, P, P, P, C.c_double, C.c_int, C.c_double, P]
        _lib.orc_project_rhs.argtypes = [P, P, P, P, C.c_double, P]
        _lib.orc_project_correct.argtypes = [P, P, P, P, P, C.c_double]
    return _lib


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        _ref = C.CDLL(REF_SO)
        P = C.c_void_p
        D, I = C.c_double, C.c_int
        _ref.ref_pcg.argtypes = [I, I, I, I, D, D, D, P, P, P, I, D, P]
        _ref.ref_uaamg_apply.argtypes = [I, I, I, I, D, D, D, P, P, P]
        _ref.ref_amgpcg.argtypes = [I, P, P, I, D, P]
        _ref.ref_spacing.argtypes = [D, I]
        _ref.ref_spacing.restype = D
        _ref.ref_project.argtypes = [I, I, I, I, D, D, D, P, P, P, P, P, D, I, D]
        _ref.ref_pcg_ex.argtypes = [I, I, I, I, D, D, D, P, P, P, I, D, I, P, P, P]
        _ref.ref_smooth_cylinder_lap.argtypes = [I, I, D, D, D, D, D, D, P, P, P, P]
        _ref.ref_write_vtk.argtypes = [I, I, I, I, D, D, D, P, P, P, P, I, C.c_char_p]
    return _ref


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Problem:
    """A padded grid problem: dims n, lengths L (dx = L/n), solid mask, walls."""

    def __init__(self, n, L=None, solid=None, wall=None, lap=None):
        """lap: None (uniform operator) or (diag, off_x, off_y) padded arrays, 2D
        (Grid::lap_diag / lap_off_x / lap_off_y, grid.h:21-23)."""
        self.lap = None if lap is None else tuple(np.ascontiguousarray(a, dtype=np.float64) for a in lap)
        self.dim = len(n)
        self.n = tuple(int(v) for v in n)
        self.L = tuple(float(v) for v in (L if L is not None else [1.0] * self.dim))
        self.h = tuple(self.L[d] / self.n[d] for d in range(self.dim))
        self.shape = tuple(v + 2 for v in self.n)          # (nx+2, ny+2[, nz+2]) i fastest
        self.N = int(np.prod(self.shape))
        if solid is None:
            solid = np.zeros(self.N, np.uint8)
        self.solid = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        self.wall = list(wall) if wall is not None else [NEUMANN] * 6

    def grid(self):
        g = OrcGrid()
        g.dim = self.dim
        g.nx, g.ny = self.n[0], self.n[1]
        g.nz = self.n[2] if self.dim == 3 else 1
        g.dx, g.dy = self.h[0], self.h[1]
        g.dz = self.h[2] if self.dim == 3 else 1.0
        g.solid = _p(self.solid)
        for f in range(6):
            g.wall[f] = self.wall[f]
        if self.lap is not None:
            g.lap_diag, g.lap_off_x, g.lap_off_y = (_p(a) for a in self.lap)
        self._keep = (self.solid, self.lap)
        return g

    def view(self, a):
        """Padded flat array -> array indexed [k][j][i] (or [j][i])."""
        return a.reshape(self.shape[::-1])

    def interior(self, a):
        v = self.view(a)
        if self.dim == 3:
            return v[1:-1, 1:-1, 1:-1]
        return v[1:-1, 1:-1]

    def mac_sizes(self):
        """Sizes of the MAC velocity arrays u, v[, w] (mesh.h:27-29, mesh_3d.h:29-31)."""
        nx, ny = self.n[0], self.n[1]
        if self.dim == 2:
            return ((nx + 1) * (ny + 2), (nx + 2) * (ny + 1), 1)
        nz = self.n[2]
        return ((nx + 1) * (ny + 2) * (nz + 2), (nx + 2) * (ny + 1) * (nz + 2), (nx + 2) * (ny + 2) * (nz + 1))

    def fluid_interior_mask(self):
        m = np.zeros(self.N, bool)
        self.interior(m)[...] = True
        return m & (self.solid == 0)


def default_opts(**kw):
    o = OrcOpts()
    lib().orc_default_opts(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def pcg(pb, rhs, max_iter, tol, p0=None, **opts):
    p = np.zeros(pb.N) if p0 is None else np.array(p0, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    g = pb.grid()
    o = default_opts(**opts)
    res = OrcResult()
    rc = lib().orc_pcg(C.byref(g), C.byref(o), _p(rhs), _p(p), max_iter, tol, C.byref(res))
    assert rc == 0, rc
    return p, res


def uaamg_apply(pb, r, **opts):
    z = np.zeros(pb.N)
    r = np.ascontiguousarray(r, dtype=np.float64)
    g = pb.grid()
    o = default_opts(**opts)
    rc = lib().orc_uaamg_apply(C.byref(g), C.byref(o), _p(r), _p(z))
    assert rc == 0, rc
    return z


def precond_apply(pb, r, precond, **opts):
    """One application of the preconditioner `precond` (PRECOND_*)."""
    z = np.zeros(pb.N)
    r = np.ascontiguousarray(r, dtype=np.float64)
    g = pb.grid()
    o = default_opts(precond=precond, **opts)
    rc = lib().orc_precond_apply(C.byref(g), C.byref(o), _p(r), _p(z))
    assert rc == 0, rc
    return z


def matvec(pb, v):
    Av = np.zeros(pb.N)
    v = np.ascontiguousarray(v, dtype=np.float64)
    g = pb.grid()
    assert lib().orc_matvec(C.byref(g), _p(v), _p(Av)) == 0
    return Av


def hierarchy_dims(pb, **opts):
    g = pb.grid()
    o = default_opts(**opts)
    nl = C.c_int(0)
    dims = np.zeros(3 * 64, np.int32)
    assert lib().orc_hierarchy_dims(C.byref(g), C.byref(o), C.byref(nl), _p(dims), 64) == 0
    return [tuple(dims[3 * l:3 * l + pb.dim]) for l in range(nl.value)]


def level_data(pb, level, **opts):
    dims = hierarchy_dims(pb, **opts)[level]
    N = int(np.prod([d + 2 for d in dims]))
    out = {k: np.zeros(N) for k in ("diag", "cx", "cy", "cz", "excess")}
    solid = np.zeros(N, np.uint8)
    g = pb.grid()
    o = default_opts(**opts)
    rc = lib().orc_level_data(C.byref(g), C.byref(o), level, _p(solid), _p(out["diag"]), _p(out["cx"]),
                              _p(out["cy"]), _p(out["cz"]), _p(out["excess"]))
    assert rc == 0, rc
    out["solid"] = solid
    out["dims"] = dims
    return out


def amgpcg(N, b, max_iter, tol):
    x = np.zeros((N + 2) * (N + 2))
    b = np.ascontiguousarray(b, dtype=np.float64)
    it = C.c_int(0)
    assert lib().orc_amgpcg(N, _p(b), _p(x), max_iter, tol, C.byref(it)) == 0
    return x, it.value


# ---- the reference's own solver (oracle/_ref), same calling conventions ----

def ref_pcg(pb, rhs, max_iter, tol, p0=None):
    p = np.zeros(pb.N) if p0 is None else np.array(p0, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    ap = C.c_int(0)
    nz = pb.n[2] if pb.dim == 3 else 1
    lz = pb.L[2] if pb.dim == 3 else 1.0
    ref().ref_pcg(pb.dim, pb.n[0], pb.n[1], nz, pb.L[0], pb.L[1], lz, _p(pb.solid), _p(rhs), _p(p),
                  max_iter, tol, C.byref(ap))
    return p, ap.value


def ref_pcg_ex(pb, rhs, max_iter, tol, precond=PRECOND_UAAMG, p0=None):
    """The reference's PCG with the preconditioner of a factory key and the
    problem's variable Laplacian (absolute tolerance, the reference's rule)."""
    p = np.zeros(pb.N) if p0 is None else np.array(p0, dtype=np.float64)
    rhs = np.ascontiguousarray(rhs, dtype=np.float64)
    nz = pb.n[2] if pb.dim == 3 else 1
    lz = pb.L[2] if pb.dim == 3 else 1.0
    lap = pb.lap if pb.lap is not None else (None, None, None)
    rc = ref().ref_pcg_ex(pb.dim, pb.n[0], pb.n[1], nz, pb.L[0], pb.L[1], lz, _p(pb.solid), _p(rhs), _p(p),
                          max_iter, tol, precond, *[(_p(a) if a is not None else None) for a in lap])
    assert rc == 0, rc
    return p


def ref_write_vtk(n, L, u, v, w, solid, frame, out_dir):
    """The reference's VtkWriter[3D]::write (vtk_writer.cpp, vtk_writer_3d.cpp)
    of a grid with these MAC arrays and padded solid mask into out_dir."""
    dim = len(n)
    nz = n[2] if dim == 3 else 1
    lz = L[2] if dim == 3 else 1.0
    u = np.ascontiguousarray(u, np.float64); v = np.ascontiguousarray(v, np.float64)
    w = np.ascontiguousarray(w if w is not None else np.zeros(1), np.float64)
    s = np.ascontiguousarray(solid, np.uint8)
    ref().ref_write_vtk(dim, n[0], n[1], nz, L[0], L[1], lz, _p(u), _p(v), _p(w), _p(s), frame, out_dir.encode())


def ref_smooth_cylinder_lap(n, L, centre, R, tw=0.0):
    """The reference's SmoothCylinder::applyToLaplacian (cylinder_model.cpp:50-71)
    on a fresh Grid: returns (solid, diag, off_x, off_y)."""
    N = (n[0] + 2) * (n[1] + 2)
    out = [np.zeros(N, np.uint8)] + [np.zeros(N) for _ in range(3)]
    ref().ref_smooth_cylinder_lap(n[0], n[1], L[0], L[1], centre[0], centre[1], R, tw, *[_p(a) for a in out])
    return tuple(out)


def ref_uaamg_apply(pb, r):
    z = np.zeros(pb.N)
    r = np.ascontiguousarray(r, dtype=np.float64)
    nz = pb.n[2] if pb.dim == 3 else 1
    lz = pb.L[2] if pb.dim == 3 else 1.0
    ref().ref_uaamg_apply(pb.dim, pb.n[0], pb.n[1], nz, pb.L[0], pb.L[1], lz, _p(pb.solid), _p(r), _p(z))
    return z


def ref_amgpcg(N, b, max_iter, tol):
    x = np.zeros((N + 2) * (N + 2))
    b = np.ascontiguousarray(b, dtype=np.float64)
    it = C.c_int(0)
    ref().ref_amgpcg(N, _p(b), _p(x), max_iter, tol, C.byref(it))
    return x, it.value


def project(pb, u, v, w, dt, max_iter, tol, **opts):
    """Pressure projection (oracle): returns (u, v, w, p, result)."""
    u = np.array(u, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    w = np.array(w if w is not None else np.zeros(1), dtype=np.float64)
    p = np.zeros(pb.N)
    g = pb.grid()
    o = default_opts(**opts)
    res = OrcResult()
    rc = lib().orc_project(C.byref(g), C.byref(o), _p(u), _p(v), _p(w), _p(p), dt, max_iter, tol, C.byref(res))
    assert rc == 0, rc
    return u, v, w, p, res


def project_rhs(pb, u, v, w, dt):
    u = np.ascontiguousarray(u, dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64)
    w = np.ascontiguousarray(w if w is not None else np.zeros(1), dtype=np.float64)
    rhs = np.zeros(pb.N)
    g = pb.grid()
    lib().orc_project_rhs(C.byref(g), _p(u), _p(v), _p(w), dt, _p(rhs))
    return rhs


def ref_project(pb, u, v, w, dt, max_iter, tol):
    """The reference's PressureProjection[3D]::project with the pcg_uaamg solver."""
    u = np.array(u, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    w = np.array(w if w is not None else np.zeros(1), dtype=np.float64)
    p = np.zeros(pb.N)
    nz = pb.n[2] if pb.dim == 3 else 1
    lz = pb.L[2] if pb.dim == 3 else 1.0
    ref().ref_project(pb.dim, pb.n[0], pb.n[1], nz, pb.L[0], pb.L[1], lz, _p(pb.solid), _p(u), _p(v), _p(w), _p(p),
                      dt, max_iter, tol)
    return u, v, w, p
