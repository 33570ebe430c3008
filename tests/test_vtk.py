"""CPU: ASCII VTK frames of the host mirror (include/uaamg_vtk.hpp,
VtkWriter / VtkWriter3D) against the reference's own writers
(src/io/vtk_writer.cpp, vtk_writer_3d.cpp, compiled into oracle/_ref) byte
for byte, against golden frames the reference wrote (tests/golden/g7_*), and
against the field definitions recomputed with numpy."""
import glob
import os
import subprocess

import numpy as np
import pytest

import oracle_bind as ob

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOST_TOOL = os.path.join(ROOT, "build", "host_solver_test")
GOLDEN = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "g7_*.npz")))

needs_tool = pytest.mark.skipif(not os.path.exists(HOST_TOOL), reason="host tool not built")


def _fields(n, seed, solid_frac=0.2):
    rng = np.random.default_rng(seed)
    if len(n) == 2:
        nx, ny = n
        u = rng.standard_normal((nx + 1) * (ny + 2))
        v = rng.standard_normal((nx + 2) * (ny + 1))
        w = None
        solid = (rng.random((nx + 2) * (ny + 2)) < solid_frac).astype(np.uint8)
    else:
        nx, ny, nz = n
        u = rng.standard_normal((nx + 1) * (ny + 2) * (nz + 2))
        v = rng.standard_normal((nx + 2) * (ny + 1) * (nz + 2))
        w = rng.standard_normal((nx + 2) * (ny + 2) * (nz + 1))
        solid = (rng.random((nx + 2) * (ny + 2) * (nz + 2)) < solid_frac).astype(np.uint8)
    return u, v, w, solid


def _ours(tmp, n, L, u, v, w, solid, frame):
    d = str(tmp)
    nz = n[2] if len(n) == 3 else 1
    lz = L[2] if len(n) == 3 else 1.0
    with open(os.path.join(d, "spec.txt"), "w") as f:
        f.write(f"{len(n)} {n[0]} {n[1]} {nz} {L[0]!r} {L[1]!r} {lz!r} {frame}\n")
    u.astype(np.float64).tofile(os.path.join(d, "u.f64"))
    v.astype(np.float64).tofile(os.path.join(d, "v.f64"))
    (w if w is not None else np.zeros(1)).astype(np.float64).tofile(os.path.join(d, "w.f64"))
    solid.astype(np.uint8).tofile(os.path.join(d, "solid.u8"))
    os.makedirs(os.path.join(d, "ours"), exist_ok=True)
    r = subprocess.run([HOST_TOOL, "vtk", d], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    with open(os.path.join(d, "ours", f"frame_{frame:05d}.vtk"), "rb") as f:
        return f.read()


CASES = [((6, 5), (1.0, 0.8)), ((17, 9), (2.0, 0.7)), ((4, 3, 3), (1.0, 0.75, 0.6)), ((9, 7, 5), (1.3, 1.0, 0.9))]


@needs_tool
@pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built (no reference tree here)")
@pytest.mark.parametrize("n,L", CASES, ids=["x".join(map(str, c[0])) for c in CASES])
def test_vtk_equals_reference_writer(tmp_path, n, L):
    u, v, w, solid = _fields(n, 5)
    frame = 37
    ours = _ours(tmp_path, n, L, u, v, w, solid, frame)
    refdir = tmp_path / "ref"
    refdir.mkdir()
    ob.ref_write_vtk(n, L, u, v, w, solid, frame, str(refdir))
    with open(refdir / f"frame_{frame:05d}.vtk", "rb") as f:
        ref = f.read()
    assert ours == ref


@needs_tool
@pytest.mark.parametrize("path", GOLDEN, ids=[os.path.basename(p)[:-4] for p in GOLDEN])
def test_vtk_equals_golden(tmp_path, path):
    d = np.load(path, allow_pickle=False)
    n = tuple(int(x) for x in d["n"])
    L = tuple(float(x) for x in d["L"])
    w = d["w"] if len(n) == 3 else None
    ours = _ours(tmp_path, n, L, d["u"], d["v"], w, d["solid"], int(d["frame"]))
    assert ours == d["vtk"].tobytes()


def _parse(blob):
    lines = blob.decode().splitlines()
    out, k = {}, 0
    while k < len(lines):
        ln = lines[k]
        if ln.startswith("VECTORS"):
            name = ln.split()[1]
            vals = []
            k += 1
            while k < len(lines) and not lines[k].startswith("SCALARS"):
                vals.append([float(x) for x in lines[k].split()])
                k += 1
            out[name] = np.array(vals)
            continue
        if ln.startswith("SCALARS"):
            name = ln.split()[1]
            k += 2
            vals = []
            while k < len(lines) and not lines[k].startswith(("SCALARS", "VECTORS")):
                vals.append(float(lines[k]))
                k += 1
            out[name] = np.array(vals)
            continue
        out.setdefault("header", []).append(ln)
        k += 1
    return out


@needs_tool
@pytest.mark.parametrize("n,L", CASES, ids=["x".join(map(str, c[0])) for c in CASES])
def test_vtk_fields_are_the_definitions(tmp_path, n, L):
    """Independent of the reference's code: the corner velocity is the mean
    of the in-domain faces that touch the corner, the divergence is the MAC
    divergence of the clamped cell, the solid marker is the clamped cell's
    (all to the 6 printed digits)."""
    u, v, w, solid = _fields(n, 9)
    f = _parse(_ours(tmp_path, n, L, u, v, w, solid, 3))
    dim = len(n)
    h = [L[a] / n[a] for a in range(dim)]
    assert f["header"][1] == "LFM %dD Fluid - Frame 3" % dim
    if dim == 2:
        nx, ny = n
        U = u.reshape(ny + 2, nx + 1); V = v.reshape(ny + 1, nx + 2); S = solid.reshape(ny + 2, nx + 2)
        vel = f["velocity"].reshape(ny + 1, nx + 1, 3)
        div = f["divergence"].reshape(ny + 1, nx + 1)
        sol = f["solid"].reshape(ny + 1, nx + 1)
        for j in range(ny + 1):
            for i in range(nx + 1):
                us = [U[j, i - 1]] * (i > 0 and 0 < j <= ny) + [U[j, i]] * (i < nx and 0 < j <= ny)
                vs = [V[j - 1, i]] * (j > 0 and 0 < i <= nx) + [V[j, i]] * (j < ny and 0 < i <= nx)
                assert np.isclose(vel[j, i, 0], np.mean(us) if us else 0.0, rtol=1e-5, atol=1e-6)
                assert np.isclose(vel[j, i, 1], np.mean(vs) if vs else 0.0, rtol=1e-5, atol=1e-6)
                ci, cj = min(max(i, 1), nx), min(max(j, 1), ny)
                dv = (U[cj, ci] - U[cj, ci - 1]) / h[0] + (V[cj, ci] - V[cj - 1, ci]) / h[1]
                assert np.isclose(div[j, i], dv, rtol=1e-5, atol=1e-5)
                assert sol[j, i] == (1.0 if S[cj, ci] else 0.0)
    else:
        nx, ny, nz = n
        U = u.reshape(nz + 2, ny + 2, nx + 1); V = v.reshape(nz + 2, ny + 1, nx + 2)
        W = w.reshape(nz + 1, ny + 2, nx + 2); S = solid.reshape(nz + 2, ny + 2, nx + 2)
        vel = f["velocity"].reshape(nz + 1, ny + 1, nx + 1, 3)
        div = f["divergence"].reshape(nz + 1, ny + 1, nx + 1)
        sol = f["solid"].reshape(nz + 1, ny + 1, nx + 1)
        for k in range(nz + 1):
            for j in range(ny + 1):
                for i in range(nx + 1):
                    iin, jin, kin = 0 < i <= nx, 0 < j <= ny, 0 < k <= nz
                    us = [U[k, j, i - 1]] * (i > 0 and jin and kin) + [U[k, j, i]] * (i < nx and jin and kin)
                    ws = [W[k - 1, j, i]] * (k > 0 and iin and jin) + [W[k, j, i]] * (k < nz and iin and jin)
                    assert np.isclose(vel[k, j, i, 0], np.mean(us) if us else 0.0, rtol=1e-5, atol=1e-6)
                    assert np.isclose(vel[k, j, i, 2], np.mean(ws) if ws else 0.0, rtol=1e-5, atol=1e-6)
                    ci, cj, ck = min(max(i, 1), nx), min(max(j, 1), ny), min(max(k, 1), nz)
                    dv = ((U[ck, cj, ci] - U[ck, cj, ci - 1]) / h[0] + (V[ck, cj, ci] - V[ck, cj - 1, ci]) / h[1]
                          + (W[ck, cj, ci] - W[ck - 1, cj, ci]) / h[2])
                    assert np.isclose(div[k, j, i], dv, rtol=1e-5, atol=1e-5)
                    assert sol[k, j, i] == (1.0 if S[ck, cj, ci] else 0.0)
