"""GPU: every benchmark configuration at its full size, pinned to the oracle.

- 2D MAC pressure solve, 1024 x 512, dx = 1/1024, dy = 1/480, solid disc:
  FP64 and FP32 against the oracle's PCG on the full grid.
- 3D plume: the 64^3 version against the oracle's PCG (FP64, FP32, FP64 CG
  with FP32 levels); at 256^3 the true residual |b - A p| / |b| of the FP64
  and FP32 solves through the oracle's operator.
- The batched mixed-size stream (64 2D + 32 3D grids of uneven size, masks
  and walls) in FP32 and with BF16 levels: every member against a converged
  FP64 solve, the eight smallest against the oracle, every member against its
  own single-grid solve.
- The served stream (the benchmark's simulation generator, 20 simulations,
  warm-started frames): sampled results against the oracle or a converged
  FP64 solve, and a request alone against the same request in the mix.
"""
import time

import numpy as np
import pytest

import oracle_bind as ob
import uaamg_amd as U
from uaamg_amd import workloads as W

pytestmark = pytest.mark.gpu


def _opb(pb):
    return ob.Problem(pb.n, pb.L, pb.solid, pb.wall)


def _err(pb, p, q, mask=None):
    """max |p - q| / max |q| over fluid cells, constants removed on pure-Neumann grids."""
    m = pb.fluid_mask() if mask is None else mask
    a, b = p[m], q[m]
    if all(w == U.NEUMANN for w in pb.wall[:2 * pb.dim]):
        a = a - a.mean()
        b = b - b.mean()
    return float(np.abs(a - b).max() / np.abs(b).max())


def _solve(pb, prec, tol, max_iter=2000, **opts):
    c = U.Context(precision=prec, tol_mode=U.TOL_REL, **opts)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    p, st = c.solve(pb.rhs, max_iter, tol)
    c.close()
    assert st.status == U.OK, (pb.name, prec, U.status_name(st.status), st.iters)
    return p, st


def _converged64(pb, rhs=None, tol=1e-10):
    """Converged FP64 solution (relres 1e-10).  The rhs is scaled by a power
    of two (exact in floating point) so that the absolute p.Ap < 1e-15 guard
    of the reference loop (pcg.cpp:100) cannot end the solve early; p is
    scaled back."""
    rhs = pb.rhs if rhs is None else rhs
    s = 2.0 ** np.round(np.log2(1e8 / max(np.linalg.norm(rhs), 1e-300)))
    c = U.Context(precision=U.FP64, tol_mode=U.TOL_REL)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    p, st = c.solve(s * rhs, 5000, tol)
    c.close()
    assert st.status == U.OK, (pb.name, U.status_name(st.status), st.rnorm / st.r0norm)
    return p / s, st


def _oracle(pb, rhs, tol, max_iter=5000, **kw):
    """The oracle's PCG on the power-of-two scaled rhs (as _converged64)."""
    s = 2.0 ** np.round(np.log2(1e8 / max(np.linalg.norm(rhs), 1e-300)))
    po, res = ob.pcg(_opb(pb), s * rhs, max_iter, tol, tol_mode=1, **kw)
    assert res.status == 0, (pb.name, res.status)
    return po / s, res


def _true_relres(pb, p):
    """|b - A x| / |b| with the oracle's operator (pcg.cpp:35-61), b the PCG's
    negated, mean-projected rhs (pcg.cpp:22-32), x = p (cold start)."""
    fl = pb.fluid_mask()
    b = np.zeros(pb.N)
    b[fl] = -pb.rhs[fl]
    if all(w == U.NEUMANN for w in pb.wall[:2 * pb.dim]):
        b[fl] -= b[fl].mean()
    x = np.where(fl, p, 0.0)
    r = b - ob.matvec(_opb(pb), x)
    return float(np.linalg.norm(r[fl]) / np.linalg.norm(b[fl]))


# ---------------------------------------------------------------- 2D MAC, full size

def test_pressure2d_full_size_vs_oracle(gpu):
    """1024 x 512 anisotropic MAC solve with a solid disc at relres 1e-6:
    FP64 within +-1 iteration of the oracle and within 1e-8 of its pressure
    (measured 3.6e-9: both stop at relres 1e-6 after 84 iterations whose dot
    products sum in different orders); FP32 within 2e-5 of it, with the
    measured iteration overhead."""
    pb = W.pressure2d_mac(seed=0)
    t = time.time()
    po, res = ob.pcg(_opb(pb), pb.rhs, 2000, 1e-6, tol_mode=1)
    print(f"[c2] oracle {res.iters} iterations in {time.time() - t:.1f} s", flush=True)
    assert res.status == 0
    p64, s64 = _solve(pb, U.FP64, 1e-6)
    assert abs(s64.iters - res.iters) <= 1, (s64.iters, res.iters)
    assert _err(pb, p64, po) <= 1e-8, _err(pb, p64, po)
    p32, s32 = _solve(pb, U.FP32, 1e-6)
    print(f"[c2] fp64 {s64.iters} fp32 {s32.iters} oracle {res.iters}; fp32 err {_err(pb, p32, po):.2e}", flush=True)
    assert s32.iters <= res.iters + max(2, (15 * res.iters) // 100), (s32.iters, res.iters)
    assert _err(pb, p32, po) <= 2e-5


# ---------------------------------------------------------------- 3D plume

def test_plume64_vs_oracle(gpu):
    """The benchmark's plume at 64^3 (floor slab + 8 spheres) against the
    oracle's PCG: at relres 1e-6 the FP64 form takes the oracle's iterations
    (+-1); at relres 1e-10 (rhs scaled by a power of two, see _converged64)
    the FP64 and FP64-CG/FP32-levels pressures are within 1e-9 / 1e-8 of the
    oracle's, and the FP32 form's at relres 1e-6 within 2e-5.  (At relres
    1e-8 the FP64 form was 4e-9 from the oracle: both stop on the same test
    after iterations whose dot products sum in different orders.)"""
    pb = W.plume3d(seed=0, n=64)
    t = time.time()
    po6, res6 = ob.pcg(_opb(pb), pb.rhs, 1000, 1e-6, tol_mode=1)
    po, res = _oracle(pb, pb.rhs, 1e-10)
    print(f"[c3-64] oracle {res6.iters} / {res.iters} iterations (relres 1e-6 / 1e-10) in {time.time() - t:.1f} s",
          flush=True)
    assert res6.status == 0
    _, s6 = _solve(pb, U.FP64, 1e-6)
    assert abs(s6.iters - res6.iters) <= 1, (s6.iters, res6.iters)
    _, sm6 = _solve(pb, U.MIXED_F64, 1e-6)
    p64, s64 = _converged64(pb)
    assert abs(s64.iters - res.iters) <= 1, (s64.iters, res.iters)
    s = 2.0 ** np.round(np.log2(1e8 / np.linalg.norm(pb.rhs)))
    c = U.Context(precision=U.MIXED_F64, tol_mode=U.TOL_REL)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    pm, sm = c.solve(s * pb.rhs, 5000, 1e-10)
    c.close()
    pm = pm / s
    p32, s32 = _solve(pb, U.FP32, 1e-6)
    print(f"[c3-64] relres 1e-6: fp64 {s6.iters} mixed64 {sm6.iters} fp32 {s32.iters} iterations (oracle "
          f"{res6.iters}); relres 1e-10: fp64 {s64.iters} mixed64 {sm.iters}; errors {_err(pb, p64, po):.1e} "
          f"{_err(pb, pm, po):.1e} {_err(pb, p32, po):.1e}", flush=True)
    assert sm.status == U.OK
    assert _err(pb, p64, po) <= 1e-9
    assert _err(pb, pm, po) <= 1e-8
    assert _err(pb, p32, po) <= 2e-5


def test_plume256_true_residual(gpu):
    """At 256^3 the oracle's PCG would take minutes per iteration; the
    solves are pinned by their true residual through the oracle's operator
    instead: the FP64 and FP64-CG forms' recursive relres 1e-6 is their true
    relres.  The FP32 form returns an FP32 iterate: rounding the FP64
    solution itself to FP32 already leaves a true relres near 1e-5 (the
    operator amplifies the cell-to-cell rounding noise by up to 12/h^2), and
    the FP32 iterate, rounded once per x += alpha p over ~100 iterations,
    sits at about ten times that floor (measured 8.98e-5 against 9.65e-6);
    its pressure is within 2e-5 of the FP64 one."""
    pb = W.plume3d(seed=0, n=256)
    out = {}
    for prec in (U.FP64, U.MIXED_F64, U.FP32):
        p, st = _solve(pb, prec, 1e-6, max_iter=400, check_every=4)
        rr = _true_relres(pb, p)
        out[prec] = (p, st.iters, rr)
        print(f"[c3-256] precision {prec}: {st.iters} iterations, recursive relres "
              f"{st.rnorm / st.r0norm:.3e}, true relres {rr:.3e}", flush=True)
    floor = _true_relres(pb, out[U.FP64][0].astype(np.float32).astype(np.float64))
    print(f"[c3-256] true relres of the FP64 solution rounded to FP32: {floor:.3e}", flush=True)
    assert out[U.FP64][2] <= 1.0e-6 and out[U.MIXED_F64][2] <= 1.0e-6
    assert out[U.FP32][2] <= 15 * max(floor, 1e-6), (out[U.FP32][2], floor)
    assert _err(pb, out[U.FP32][0], out[U.FP64][0]) <= 2e-5


# ---------------------------------------------------------------- batched mixed-size stream

@pytest.fixture(scope="module")
def batch_stream(gpu):
    """The 96 grids and their converged FP64 solutions (relres 1e-10)."""
    pbs = W.batch_mixed(seed=0)
    ref = [_converged64(pb)[0] for pb in pbs]
    return pbs, ref


def _batch(pbs, prec):
    import torch
    dev = torch.device("cuda:0")
    b = U.Batch(precision=prec, tol_mode=U.TOL_REL, use_graph=1, max_iter=2000)
    for pb in pbs:
        b.add(pb.n, pb.h, pb.solid, pb.wall)
    rhs = [torch.from_numpy(pb.rhs).to(dev) for pb in pbs]
    out = [torch.zeros(pb.N, dtype=torch.float64, device=dev) for pb in pbs]
    torch.cuda.synchronize()
    rc, st = b.solve_device([t.data_ptr() for t in rhs], [t.data_ptr() for t in out], None, 2000, 1e-6)
    torch.cuda.synchronize()
    b.close()
    return rc, st, [t.cpu().numpy() for t in out]


@pytest.mark.parametrize("prec", [U.FP32, U.MIXED_BF16], ids=["fp32", "bf16"])
def test_batch_stream_vs_fp64(batch_stream, prec):
    """Every member of the 96-grid stream at relres 1e-6 is within 1e-5 of the
    converged FP64 solution (relative to max |p|)."""
    pbs, ref = batch_stream
    rc, st, ps = _batch(pbs, prec)
    assert rc == U.OK, U.status_name(rc)
    errs = [_err(pb, p, q) for pb, p, q in zip(pbs, ps, ref)]
    worst = int(np.argmax(errs))
    print(f"[c4] precision {prec}: max error {errs[worst]:.2e} ({pbs[worst].name}), "
          f"iterations max {max(s.iters for s in st)}", flush=True)
    assert all(s.status == U.OK for s in st)
    assert max(errs) <= 1e-5, (pbs[worst].name, errs[worst])


def test_batch_stream_smallest_vs_oracle(batch_stream):
    """The eight smallest members against the oracle's PCG (relres 1e-10,
    their own walls and masks): FP64 solve within 1e-9, batched FP32 and BF16
    within 1e-5."""
    pbs, ref = batch_stream
    order = sorted(range(len(pbs)), key=lambda k: pbs[k].cells)[:8]
    b32 = _batch(pbs, U.FP32)[2]
    bbf = _batch(pbs, U.MIXED_BF16)[2]
    for k in order:
        pb = pbs[k]
        po, res = _oracle(pb, pb.rhs, 1e-10)
        assert _err(pb, ref[k], po) <= 1e-9, pb.name
        assert _err(pb, b32[k], po) <= 1e-5, pb.name
        assert _err(pb, bbf[k], po) <= 1e-5, pb.name


def test_batch_stream_members_equal_solo(batch_stream):
    """In the batch each member runs the single-grid kernels' arithmetic with
    its own reductions (summed over other partial groupings): iterations
    within 2 % (at least 2) and pressure within 1e-5 of its single-grid FP32
    solve."""
    pbs, _ = batch_stream
    rc, st, ps = _batch(pbs, U.FP32)
    worst_it, worst_e = 0, 0.0
    for pb, s, p in zip(pbs, st, ps):
        p1, s1 = _solve(pb, U.FP32, 1e-6)
        assert abs(s.iters - s1.iters) <= max(2, s1.iters // 50), (pb.name, s.iters, s1.iters)
        e = _err(pb, p, p1)
        assert e <= 1e-5, (pb.name, e)
        worst_it, worst_e = max(worst_it, abs(s.iters - s1.iters)), max(worst_e, e)
    print(f"[c4] batch vs solo: iterations differ by at most {worst_it}, pressure by {worst_e:.1e}", flush=True)


# ---------------------------------------------------------------- served stream

def test_serving_stream(gpu):
    """20 simulations of the served stream's generator (2D side 64-2048, 3D
    side 32-256, solid balls, frames advected by a one-cell shift plus 5 %
    noise), three frames each, the later ones warm-started from the previous
    pressure, relres 1e-5, BF16 levels, continuous admission into 8 slots:
    every request completes; sampled results are within 2e-5 of the oracle
    (small grids) or of a converged FP64 solve (all sampled); a request served
    alone equals the same request in the mix to 1e-5 (the same kernels; the
    batch's launch shapes and partial groupings differ)."""
    import torch
    dev = torch.device("cuda:0")
    sims = W.serve_sims(20, 7, dev)
    opts = dict(precision=U.MIXED_BF16, tol_mode=U.TOL_REL, use_graph=1, max_iter=500, flexible=1)
    srv = U.Server(8 << 30, max_active=8, **opts)
    for k, s in enumerate(sims):
        srv.register(k, s["n"], s["h"], s["solid"])
    rng = np.random.default_rng(11)
    p = [torch.zeros(int(np.prod(s["shape"])), dtype=torch.float64, device=dev) for s in sims]
    frames = [0] * len(sims)
    record = {}                 # req -> (sim, rhs, x0, warm)

    def submit(k):
        r = (W.next_frame(sims[k], rng) if frames[k] else sims[k]["rhs"]).clone()
        warm = 1 if frames[k] else 0
        x0 = p[k].clone()
        torch.cuda.current_stream().synchronize()
        req = srv.submit(k, r.data_ptr(), p[k].data_ptr(), warm, 500, 1e-5)
        record[req] = (k, r, x0, warm)
        frames[k] += 1
        return req

    live = {submit(k): k for k in range(len(sims))}
    results = {}
    t0 = time.time()
    while live:
        for d in srv.step():
            k = live.pop(d)
            torch.cuda.synchronize()
            results[d] = (srv.request(d), p[k].clone())
            if frames[k] < 3:
                live[submit(k)] = k
        assert time.time() - t0 < 300, "server did not finish"
    srv.close()
    assert len(results) == 3 * len(sims)
    assert all(st.status == U.OK for st, _ in results.values())
    assert all(st.ms_total > 0 for st, _ in results.values())

    def solve64(s, rhs):
        # power-of-two scaling of the rhs, as _converged64
        sc = 2.0 ** np.round(np.log2(1e8 / float(rhs.norm())))
        c = U.Context(precision=U.FP64, tol_mode=U.TOL_REL)
        c.set_grid(s["n"], s["h"], s["solid"])
        q = torch.zeros_like(rhs)
        rs = rhs * sc
        st = c.solve_device(rs.data_ptr(), q.data_ptr(), 0, 5000, 1e-10)
        torch.cuda.synchronize()
        c.close()
        assert st.status == U.OK, U.status_name(st.status)
        return q / sc

    def err(s, a, b):
        f = s["fluid"]
        a, b = a[f], b[f]
        a, b = a - a.mean(), b - b.mean()
        return float((a - b).abs().max() / b.abs().max())

    sampled = sorted(results)[::4]
    n_oracle = 0
    worst = 0.0
    for req in sampled:
        k, r, x0, warm = record[req]
        s = sims[k]
        served = results[req][1]
        e64 = err(s, served, solve64(s, r))
        worst = max(worst, e64)
        assert e64 <= 2e-5, (req, s["n"], e64)
        if int(np.prod(s["n"])) <= 1 << 17:
            opb = ob.Problem(s["n"], (1.0,) * s["dim"], s["solid"])
            rn = r.cpu().numpy()
            sc = 2.0 ** np.round(np.log2(1e8 / np.linalg.norm(rn)))
            po, res = ob.pcg(opb, sc * rn, 5000, 1e-10, tol_mode=1)
            assert res.status == 0
            po = po / sc
            eo = err(s, served, torch.from_numpy(po).to(dev))
            assert eo <= 2e-5, (req, s["n"], eo)
            n_oracle += 1
    print(f"[c5] {len(results)} requests, {len(sampled)} sampled (max error vs FP64 {worst:.2e}), "
          f"{n_oracle} against the oracle", flush=True)
    assert n_oracle >= 2

    # a request alone (same rhs, same warm-start pressure) equals the mix
    for req in sampled[:4]:
        k, r, x0, warm = record[req]
        s = sims[k]
        alone = U.Server(8 << 30, max_active=2, **opts)
        alone.register(k, s["n"], s["h"], s["solid"])
        q = x0.clone()
        torch.cuda.synchronize()
        a_req = alone.submit(k, r.data_ptr(), q.data_ptr(), warm, 500, 1e-5)
        done = []
        while a_req not in done:
            done += alone.step()
        torch.cuda.synchronize()
        st_alone = alone.request(a_req)
        alone.close()
        e = err(s, q, results[req][1])
        print(f"[c5] request {req} ({s['n']}): alone {st_alone.iters} iterations, in the mix "
              f"{results[req][0].iters}; difference {e:.1e}", flush=True)
        assert abs(st_alone.iters - results[req][0].iters) <= 2, (st_alone.iters, results[req][0].iters)
        assert e <= 1e-5, (req, e)
