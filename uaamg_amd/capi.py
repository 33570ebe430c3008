"""ctypes view of the C-ABI in include/uaamg.h."""
import ctypes as C
import os
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

FP64, FP32, MIXED_BF16, MIXED_FP16, MIXED_BF16R, MIXED_F64 = 0, 1, 2, 3, 4, 5
TOL_ABS, TOL_REL = 0, 1
NEUMANN, DIRICHLET = 0, 1
OK, ERR_ARG, ERR_HIP, ERR_OOM, ERR_NODEVICE, ERR_ADMISSION = 0, 1, 2, 3, 4, 5
NOT_CONVERGED, BREAKDOWN, NONFINITE = 10, 11, 12
_NAMES = {0: "ok", 1: "bad-argument", 2: "hip-error", 3: "out-of-memory", 4: "no-device",
          5: "admission", 10: "not-converged", 11: "breakdown", 12: "nonfinite"}
K_MATVEC, K_UPDATE, K_DOWN0, K_UP0, K_PUPDATE, K_COARSE, K_INIT, K_OTHER = range(8)
FORM_ROWS, FORM_ROWS_COARSE, FORM_TAIL, FORM_PMV, FORM_XMASK = 1, 2, 4, 8, 16
FORM_NT_X, FORM_NT_P, FORM_NT_D, FORM_NT_U, FORM_SPLIT = 32, 64, 128, 256, 512
FORM_NT = FORM_NT_X | FORM_NT_P | FORM_NT_D | FORM_NT_U
PRECOND_UAAMG, PRECOND_NONE, PRECOND_GMG, PRECOND_AMG = 0, 1, 2, 3
HIERARCHY_REFERENCE, HIERARCHY_KEEP_FLUID = 0, 1
FORM_DEFAULT = FORM_ROWS | FORM_ROWS_COARSE | FORM_TAIL | FORM_PMV | FORM_NT


def status_name(rc):
    return _NAMES.get(rc, str(rc))


class GridDesc(C.Structure):
    _fields_ = [("dim", C.c_int), ("n", C.c_int * 3), ("h", C.c_double * 3),
                ("solid", C.c_void_p), ("wall", C.c_int * 6)]


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int), ("nu_pre", C.c_int), ("nu_post", C.c_int),
                ("coarsest_side", C.c_int), ("coarsest_sweeps", C.c_int), ("tol_mode", C.c_int),
                ("max_iter", C.c_int), ("tol", C.c_double), ("project_mean", C.c_int),
                ("fused", C.c_int), ("check_every", C.c_int), ("use_graph", C.c_int),
                ("timing", C.c_int), ("flexible", C.c_int), ("forms", C.c_int), ("xfork_level", C.c_int),
                ("xblocks", C.c_int), ("xthreads", C.c_int), ("rows_kc", C.c_int), ("march_kc_min", C.c_int),
                ("march_target", C.c_int), ("trace", C.c_int), ("preconditioner", C.c_int),
                ("hierarchy_mode", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("nlevels", C.c_int),
                ("rnorm", C.c_double), ("r0norm", C.c_double), ("ms_total", C.c_double),
                ("ms_setup", C.c_double), ("ms_kernel", C.c_double * 8), ("n_kernel", C.c_int * 8),
                ("alg_bytes", C.c_double), ("kernel_set", C.c_int), ("forms_used", C.c_int)]


EXPORTS = ["uaamg_default_opts", "uaamg_create", "uaamg_destroy", "uaamg_last_error", "uaamg_set_opts",
           "uaamg_set_grid", "uaamg_set_grid_device", "uaamg_set_variable_lap", "uaamg_solve", "uaamg_solve_device",
           "uaamg_apply_precond", "uaamg_matvec", "uaamg_hierarchy", "uaamg_level_data",
           "uaamg_amgpcg", "uaamg_amgpcg_noalloc", "uaamg_bytes_required", "uaamg_project", "uaamg_project_device",
           "uaamg_batch_create", "uaamg_batch_destroy", "uaamg_batch_last_error", "uaamg_batch_add",
           "uaamg_batch_add_device", "uaamg_batch_clear", "uaamg_batch_size", "uaamg_batch_bytes",
           "uaamg_batch_tuning",
           "uaamg_batch_solve",
           "uaamg_server_create", "uaamg_server_destroy", "uaamg_server_last_error", "uaamg_server_register",
           "uaamg_server_submit", "uaamg_server_step", "uaamg_server_request", "uaamg_server_status"]


def lib_path():
    return os.path.join(HERE, "lib", "libuaamg_hip.so")


class Lib:
    def __init__(self, path=None):
        path = path or lib_path()
        if not os.path.exists(path):
            raise RuntimeError(f"HIP library not built: {path} (run __graft_entry__.build())")
        self.path = path
        self.L = L = C.CDLL(path)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        L.uaamg_default_opts.argtypes = [P]
        L.uaamg_create.argtypes = [C.POINTER(C.c_void_p), P]
        L.uaamg_destroy.argtypes = [P]
        L.uaamg_destroy.restype = None
        L.uaamg_last_error.argtypes = [P]
        L.uaamg_last_error.restype = C.c_char_p
        L.uaamg_set_opts.argtypes = [P, P]
        L.uaamg_set_grid.argtypes = [P, P]
        L.uaamg_set_grid_device.argtypes = [P, P]
        L.uaamg_set_variable_lap.argtypes = [P, P, P, P]
        L.uaamg_solve.argtypes = [P, P, P, I, D, P]
        L.uaamg_solve_device.argtypes = [P, P, P, I, I, D, P]
        L.uaamg_apply_precond.argtypes = [P, P, P]
        L.uaamg_matvec.argtypes = [P, P, P, P]
        L.uaamg_hierarchy.argtypes = [P, P, P, I]
        L.uaamg_level_data.argtypes = [P, I, P, P, P, P, P, P]
        L.uaamg_amgpcg.argtypes = [I, P, P, I, D, P, I]
        L.uaamg_amgpcg_noalloc.argtypes = [I, P, P, I, D, P, I, P]
        L.uaamg_bytes_required.argtypes = [P, P]
        L.uaamg_bytes_required.restype = C.c_size_t
        L.uaamg_project.argtypes = [P, P, P, P, P, D, I, D, P]
        L.uaamg_project_device.argtypes = [P, P, P, P, P, D, I, D, P]
        L.uaamg_batch_create.argtypes = [C.POINTER(C.c_void_p), P]
        L.uaamg_batch_destroy.argtypes = [P]
        L.uaamg_batch_destroy.restype = None
        L.uaamg_batch_last_error.argtypes = [P]
        L.uaamg_batch_last_error.restype = C.c_char_p
        L.uaamg_batch_add.argtypes = [P, P, P]
        L.uaamg_batch_add_device.argtypes = [P, P, P]
        L.uaamg_batch_clear.argtypes = [P]
        L.uaamg_batch_size.argtypes = [P]
        L.uaamg_batch_bytes.argtypes = [P]
        L.uaamg_batch_bytes.restype = C.c_size_t
        L.uaamg_batch_tuning.argtypes = [P, I, P, P, P]
        L.uaamg_batch_solve.argtypes = [P, P, P, P, I, D, P]
        L.uaamg_server_create.argtypes = [C.POINTER(C.c_void_p), P, C.c_size_t, I]
        L.uaamg_server_destroy.argtypes = [P]
        L.uaamg_server_destroy.restype = None
        L.uaamg_server_last_error.argtypes = [P]
        L.uaamg_server_last_error.restype = C.c_char_p
        L.uaamg_server_register.argtypes = [P, I, P]
        L.uaamg_server_submit.argtypes = [P, I, P, P, I, I, D, P]
        L.uaamg_server_step.argtypes = [P, P, I, P]
        L.uaamg_server_request.argtypes = [P, I, P]
        L.uaamg_server_status.argtypes = [P, P, P, P, P]

    def default_opts(self, **kw):
        o = Opts()
        self.L.uaamg_default_opts(C.byref(o))
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def amgpcg(self, N, b, max_iter, tol, precision=FP64):
        x = np.zeros((N + 2) * (N + 2))
        b = np.ascontiguousarray(b, dtype=np.float64).reshape(-1)
        if b.size != x.size:
            raise ValueError(f"b: {b.size} values, N = {N} needs {x.size}")
        it = C.c_int(0)
        rc = self.L.uaamg_amgpcg(N, _p(b), _p(x), max_iter, tol, C.byref(it), precision)
        if rc != OK:
            raise RuntimeError(f"uaamg_amgpcg failed: {status_name(rc)}")
        return x, it.value


_LIB = None


def load(path=None):
    global _LIB
    if _LIB is None or path is not None:
        _LIB = Lib(path)
    return _LIB


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def make_grid(n, h, solid, wall=None):
    g = GridDesc()
    g.dim = len(n)
    for q in range(3):
        g.n[q] = int(n[q]) if q < len(n) else 1
        g.h[q] = float(h[q]) if q < len(h) else 1.0
    g.solid = _p(solid)
    w = wall if wall is not None else [NEUMANN] * 6
    for f in range(6):
        g.wall[f] = int(w[f])
    return g


class Context:
    """One solver context (device hierarchy + CG buffers) on its own HIP stream."""

    def __init__(self, lib=None, **opts):
        self.lib = lib or load()
        self.opts = self.lib.default_opts(**opts)
        h = C.c_void_p()
        rc = self.lib.L.uaamg_create(C.byref(h), C.byref(self.opts))
        if rc != OK:
            raise RuntimeError(f"uaamg_create failed: {status_name(rc)}")
        self.h = h
        self.n = None

    h = None

    def close(self):
        if self.h:
            self.lib.L.uaamg_destroy(self.h)
            self.h = None

    __del__ = close

    def _err(self, what, rc):
        msg = self.lib.L.uaamg_last_error(self.h).decode()
        raise RuntimeError(f"{what}: {status_name(rc)}: {msg}")

    def set_opts(self, **kw):
        for k, v in kw.items():
            setattr(self.opts, k, v)
        rc = self.lib.L.uaamg_set_opts(self.h, C.byref(self.opts))
        if rc != OK:
            self._err("set_opts", rc)

    def set_grid(self, n, h, solid, wall=None):
        self.n = tuple(int(v) for v in n)
        self.dim = len(n)
        self._solid = np.ascontiguousarray(solid, dtype=np.uint8).reshape(-1)
        self.N = int(np.prod([v + 2 for v in n]))
        if self._solid.size != self.N:
            raise ValueError(f"solid: {self._solid.size} values, the grid needs {self.N}")
        g = make_grid(n, h, self._solid, wall)
        rc = self.lib.L.uaamg_set_grid(self.h, C.byref(g))
        if rc != OK:
            self._err("set_grid", rc)

    def _cells(self, a, what, n=None):
        """C-order FP64 copy of a padded array, checked against the grid's size."""
        if self.n is None:
            raise RuntimeError("no grid set")
        n = self.N if n is None else n
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(-1)
        if a.size != n:
            raise ValueError(f"{what}: {a.size} values, the grid needs {n}")
        return a

    def mac_sizes(self):
        """Lengths of the MAC face arrays u, v, w (mesh.h:27-35, mesh_3d.h:29-37)."""
        nx, ny = self.n[0], self.n[1]
        if self.dim == 2:
            return (nx + 1) * (ny + 2), (nx + 2) * (ny + 1), 0
        nz = self.n[2]
        return (nx + 1) * (ny + 2) * (nz + 2), (nx + 2) * (ny + 1) * (nz + 2), (nx + 2) * (ny + 2) * (nz + 1)

    def set_variable_lap(self, lap):
        """lap = (diag, off_x, off_y) padded arrays (2D), or None for the uniform operator."""
        if lap is None:
            rc = self.lib.L.uaamg_set_variable_lap(self.h, None, None, None)
        else:
            self._lap = [self._cells(a, name) for a, name in zip(lap, ("lap_diag", "lap_off_x", "lap_off_y"))]
            rc = self.lib.L.uaamg_set_variable_lap(self.h, *[_p(a) for a in self._lap])
        if rc != OK:
            self._err("set_variable_lap", rc)

    def solve(self, rhs, max_iter=-1, tol=-1.0, p0=None):
        rhs = self._cells(rhs, "rhs")
        p = np.zeros(self.N) if p0 is None else self._cells(p0, "p0").copy()
        st = Stats()
        rc = self.lib.L.uaamg_solve(self.h, _p(rhs), _p(p), int(max_iter), float(tol), C.byref(st))
        if rc not in (OK, NOT_CONVERGED, BREAKDOWN, NONFINITE):
            self._err("solve", rc)
        return p, st

    def solve_device(self, rhs_ptr, p_ptr, warm=0, max_iter=-1, tol=-1.0):
        st = Stats()
        rc = self.lib.L.uaamg_solve_device(self.h, C.c_void_p(rhs_ptr), C.c_void_p(p_ptr), int(warm),
                                           int(max_iter), float(tol), C.byref(st))
        if rc not in (OK, NOT_CONVERGED, BREAKDOWN, NONFINITE):
            self._err("solve_device", rc)
        return st

    def project(self, u, v, w, dt, max_iter=-1, tol=-1.0):
        """Pressure projection on host MAC arrays (copies); returns (u, v, w, p, Stats)."""
        nu, nv, nw = self.mac_sizes()
        u = self._cells(u, "u", nu).copy()
        v = self._cells(v, "v", nv).copy()
        if self.dim == 3 and w is None:
            raise ValueError("w: a 3D projection needs w")
        w = self._cells(w, "w", nw).copy() if self.dim == 3 else None
        p = np.zeros(self.N)
        st = Stats()
        rc = self.lib.L.uaamg_project(self.h, _p(u), _p(v), _p(w) if w is not None else None, _p(p), float(dt),
                                      int(max_iter), float(tol), C.byref(st))
        if rc not in (OK, NOT_CONVERGED, BREAKDOWN, NONFINITE):
            self._err("project", rc)
        return u, v, w, p, st

    def project_device(self, u_ptr, v_ptr, w_ptr, p_ptr, dt, max_iter=-1, tol=-1.0):
        st = Stats()
        rc = self.lib.L.uaamg_project_device(self.h, C.c_void_p(u_ptr), C.c_void_p(v_ptr),
                                             C.c_void_p(w_ptr) if w_ptr else None, C.c_void_p(p_ptr), float(dt),
                                             int(max_iter), float(tol), C.byref(st))
        if rc not in (OK, NOT_CONVERGED, BREAKDOWN, NONFINITE):
            self._err("project_device", rc)
        return st

    def apply(self, r):
        r = self._cells(r, "r")
        z = np.zeros(self.N)
        rc = self.lib.L.uaamg_apply_precond(self.h, _p(r), _p(z))
        if rc != OK:
            self._err("apply", rc)
        return z

    def matvec(self, v):
        v = self._cells(v, "v")
        Av = np.zeros(self.N)
        d = C.c_double(0)
        rc = self.lib.L.uaamg_matvec(self.h, _p(v), _p(Av), C.byref(d))
        if rc != OK:
            self._err("matvec", rc)
        return Av, d.value

    def hierarchy(self):
        nl = C.c_int(0)
        dims = np.zeros(3 * 64, np.int32)
        rc = self.lib.L.uaamg_hierarchy(self.h, C.byref(nl), _p(dims), 64)
        if rc != OK:
            self._err("hierarchy", rc)
        return [tuple(dims[3 * l:3 * l + self.dim]) for l in range(nl.value)]

    def level_data(self, level):
        dims = self.hierarchy()[level]
        N = int(np.prod([d + 2 for d in dims]))
        out = {k: np.zeros(N) for k in ("diag", "cx", "cy", "cz", "excess")}
        solid = np.zeros(N, np.uint8)
        rc = self.lib.L.uaamg_level_data(self.h, level, _p(solid), _p(out["diag"]), _p(out["cx"]),
                                         _p(out["cy"]), _p(out["cz"]), _p(out["excess"]))
        if rc != OK:
            self._err("level_data", rc)
        out["solid"] = solid
        out["dims"] = dims
        return out


class Batch:
    """Many independent solves run as one set of launches (uaamg_batch_*)."""

    def __init__(self, lib=None, **opts):
        self.lib = lib or load()
        opts.setdefault("precision", FP32)
        self.opts = self.lib.default_opts(**opts)
        h = C.c_void_p()
        rc = self.lib.L.uaamg_batch_create(C.byref(h), C.byref(self.opts))
        if rc != OK:
            raise RuntimeError(f"uaamg_batch_create failed: {status_name(rc)}")
        self.h = h
        self.members = []

    h = None

    def close(self):
        if self.h:
            self.lib.L.uaamg_batch_destroy(self.h)
            self.h = None

    __del__ = close

    def _err(self, what, rc):
        msg = self.lib.L.uaamg_batch_last_error(self.h).decode()
        raise RuntimeError(f"{what}: {status_name(rc)}: {msg}")

    def add(self, n, h, solid, wall=None, device=False):
        """Add a member; solid is a host numpy mask or (device=True) a device pointer."""
        if device:
            g = make_grid(n, h, np.zeros(1, np.uint8), wall)
            g.solid = C.c_void_p(int(solid))
            fn = self.lib.L.uaamg_batch_add_device
        else:
            solid = np.ascontiguousarray(solid, dtype=np.uint8)
            g = make_grid(n, h, solid, wall)
            fn = self.lib.L.uaamg_batch_add
        i = C.c_int(-1)
        rc = fn(self.h, C.byref(g), C.byref(i))
        if rc != OK:
            self._err("uaamg_batch_add", rc)
        self.members.append(tuple(n))
        return i.value

    def clear(self):
        rc = self.lib.L.uaamg_batch_clear(self.h)
        if rc != OK:
            self._err("uaamg_batch_clear", rc)
        self.members = []

    def size(self):
        return self.lib.L.uaamg_batch_size(self.h)

    def bytes(self):
        return self.lib.L.uaamg_batch_bytes(self.h)

    def tuning(self, dim):
        """Launch shapes the autotuner chose for the dim-D members:
        {"march_target", "march_kc_min", "rows_kc"}, or None without such members."""
        t, k, r = C.c_int(0), C.c_int(0), C.c_int(0)
        if self.lib.L.uaamg_batch_tuning(self.h, int(dim), C.byref(t), C.byref(k), C.byref(r)) != OK:
            return None
        return {"march_target": t.value, "march_kc_min": k.value, "rows_kc": r.value}

    def solve_device(self, rhs_ptrs, p_ptrs, warm=None, max_iter=-1, tol=-1.0):
        """rhs_ptrs / p_ptrs: device pointers (ints) of every member, in order.
        Returns (status, [Stats])."""
        n = len(rhs_ptrs)
        assert n == len(p_ptrs) == self.size()
        R = (C.c_void_p * n)(*[C.c_void_p(int(x)) for x in rhs_ptrs])
        Pp = (C.c_void_p * n)(*[C.c_void_p(int(x)) for x in p_ptrs])
        W = (C.c_int * n)(*([int(w) for w in warm] if warm is not None else [0] * n))
        st = (Stats * n)()
        rc = self.lib.L.uaamg_batch_solve(self.h, R, Pp, W, max_iter, tol, st)
        if rc not in (OK, NOT_CONVERGED, BREAKDOWN, NONFINITE):
            self._err("uaamg_batch_solve", rc)
        return rc, list(st)


class Server:
    """Request loop over continuous batches (uaamg_server_*)."""

    def __init__(self, budget_bytes, max_active=64, lib=None, **opts):
        self.lib = lib or load()
        opts.setdefault("precision", FP32)
        self.opts = self.lib.default_opts(**opts)
        h = C.c_void_p()
        rc = self.lib.L.uaamg_server_create(C.byref(h), C.byref(self.opts), int(budget_bytes), int(max_active))
        if rc != OK:
            raise RuntimeError(f"uaamg_server_create failed: {status_name(rc)}")
        self.h = h
        self._done = (C.c_int * 4096)()

    h = None

    def close(self):
        if self.h:
            self.lib.L.uaamg_server_destroy(self.h)
            self.h = None

    __del__ = close

    def _err(self, what, rc):
        msg = self.lib.L.uaamg_server_last_error(self.h).decode()
        raise RuntimeError(f"{what}: {status_name(rc)}: {msg}")

    def register(self, sim, n, h, solid, wall=None):
        solid = np.ascontiguousarray(solid, dtype=np.uint8)
        g = make_grid(n, h, solid, wall)
        rc = self.lib.L.uaamg_server_register(self.h, int(sim), C.byref(g))
        if rc != OK:
            self._err("uaamg_server_register", rc)

    def submit(self, sim, rhs_ptr, p_ptr, warm=0, max_iter=-1, tol=-1.0):
        r = C.c_int(-1)
        rc = self.lib.L.uaamg_server_submit(self.h, int(sim), C.c_void_p(int(rhs_ptr)), C.c_void_p(int(p_ptr)),
                                            int(warm), int(max_iter), float(tol), C.byref(r))
        if rc != OK:
            self._err("uaamg_server_submit", rc)
        return r.value

    def step(self):
        n = C.c_int(0)
        rc = self.lib.L.uaamg_server_step(self.h, self._done, len(self._done), C.byref(n))
        if rc != OK:
            self._err("uaamg_server_step", rc)
        return [self._done[k] for k in range(n.value)]

    def request(self, req):
        st = Stats()
        rc = self.lib.L.uaamg_server_request(self.h, int(req), C.byref(st))
        return st if rc == OK else None

    def status(self):
        q, a, r = C.c_int(0), C.c_int(0), C.c_int(0)
        b = C.c_size_t(0)
        self.lib.L.uaamg_server_status(self.h, C.byref(q), C.byref(a), C.byref(r), C.byref(b))
        return {"queued": q.value, "active": a.value, "resident": r.value, "bytes": b.value}
