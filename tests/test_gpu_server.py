"""GPU: the serving loop (uaamg_server_*) -- continuous admission of solves
from many simulations, resident hierarchies under an HBM budget, warm
starts -- against single-grid solves, and alone vs in the mix."""
import time

import numpy as np
import pytest

import cases
import uaamg_amd as U

pytestmark = pytest.mark.gpu

SIMS = [c for c in cases.parity_cases() if "porous" not in c[0] and c[0] != "3d_16x8x4"]   # serving needs >= 2 levels


def _rhs(pb, seed):
    return 1e3 * cases.random_rhs(pb, np.random.default_rng(seed))


def _err(pb, p, q):
    m = pb.fluid_interior_mask()
    a, b = p[m], q[m]
    if all(w == 0 for w in pb.wall[:2 * pb.dim]):
        a = a - a.mean()
        b = b - b.mean()
    return float(np.abs(a - b).max() / np.abs(b).max())


def _run(server, jobs):
    """jobs: list of (sim, rhs tensor, p tensor, warm); returns {req: stats}."""
    reqs = [server.submit(sim, r.data_ptr(), p.data_ptr(), warm, 500, 1e-6) for sim, r, p, warm in jobs]
    left = set(reqs)
    t0 = time.time()
    while left:
        for d in server.step():
            left.discard(d)
        assert time.time() - t0 < 60, "server did not finish"
    return {r: server.request(r) for r in reqs}


def _dev(a):
    import torch
    return torch.from_numpy(np.array(a, dtype=np.float64)).to("cuda:0")


def test_server_matches_single_solves_and_alone_equals_mix(gpu):
    import torch
    mix = U.Server(1 << 34, max_active=32, precision=U.FP32, tol_mode=U.TOL_REL, use_graph=1)
    for k, (name, pb) in enumerate(SIMS):
        mix.register(k, pb.n, pb.h, pb.solid, pb.wall)
    rh = [_dev(_rhs(pb, 3)) for _, pb in SIMS]
    ps = [_dev(np.zeros(pb.N)) for _, pb in SIMS]
    st = _run(mix, [(k, rh[k], ps[k], 0) for k in range(len(SIMS))])
    torch.cuda.synchronize()
    for k, (name, pb) in enumerate(SIMS):
        s = list(st.values())[k]
        assert s.status == 0, name
        c = U.Context(precision=U.FP32, tol_mode=U.TOL_REL)
        c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
        p1, s1 = c.solve(_rhs(pb, 3), 500, 1e-6)
        c.close()
        assert abs(s.iters - s1.iters) <= 2, (name, s.iters, s1.iters)
        assert _err(pb, ps[k].cpu().numpy(), p1) <= 1e-5, name
    # the same request alone
    for k in (0, 5, 9):
        name, pb = SIMS[k]
        alone = U.Server(1 << 34, max_active=4, precision=U.FP32, tol_mode=U.TOL_REL, use_graph=1)
        alone.register(k, pb.n, pb.h, pb.solid, pb.wall)
        p_alone = _dev(np.zeros(pb.N))
        _run(alone, [(k, rh[k], p_alone, 0)])
        torch.cuda.synchronize()
        assert _err(pb, p_alone.cpu().numpy(), ps[k].cpu().numpy()) <= 1e-6, name
        alone.close()
    mix.close()


def test_server_warm_start_frames(gpu):
    """Consecutive frames of one simulation: the warm start from the previous
    pressure needs fewer iterations than a cold start."""
    import torch
    name, pb = [c for c in SIMS if c[0] == "3d_24x20x18_aniso_solids"][0]
    srv = U.Server(1 << 34, max_active=4, precision=U.FP32, tol_mode=U.TOL_REL, use_graph=1)
    srv.register(0, pb.n, pb.h, pb.solid, pb.wall)
    r0 = _rhs(pb, 3)
    p = _dev(np.zeros(pb.N))
    s0 = list(_run(srv, [(0, _dev(r0), p, 0)]).values())[0]
    r1 = r0 + 0.02 * _rhs(pb, 4)
    s1 = list(_run(srv, [(0, _dev(r1), p, 1)]).values())[0]
    torch.cuda.synchronize()
    assert s0.status == 0 and s1.status == 0
    assert s1.iters < s0.iters, (s1.iters, s0.iters)
    c = U.Context(precision=U.FP64, tol_mode=U.TOL_REL)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    pr, _ = c.solve(r1, 500, 1e-10)
    c.close()
    assert _err(pb, p.cpu().numpy(), pr) <= 2e-5
    srv.close()


def test_server_admission_under_budget(gpu):
    """A budget that holds only a few hierarchies: every request still
    completes (idle simulations are evicted and rebuilt on demand)."""
    import torch
    small = [c for c in SIMS if c[1].dim == 3]
    import ctypes
    o32 = U.load().default_opts(precision=U.FP32)
    est = max(U.load().L.uaamg_bytes_required(ctypes.byref(U.capi.make_grid(pb.n, pb.h, pb.solid, pb.wall)),
                                              ctypes.byref(o32)) for _, pb in small)
    budget = int(1.05 * est)      # the largest hierarchy fits only alone
    srv = U.Server(budget, max_active=8, precision=U.FP32, tol_mode=U.TOL_REL, use_graph=1)
    for k, (name, pb) in enumerate(small):
        srv.register(k, pb.n, pb.h, pb.solid, pb.wall)
    jobs = []
    for rep in range(2):
        for k, (name, pb) in enumerate(small):
            jobs.append((k, _dev(_rhs(pb, 10 + k)), _dev(np.zeros(pb.N)), 0))
    st = _run(srv, jobs)
    torch.cuda.synchronize()
    assert all(s.status == 0 for s in st.values())
    status = srv.status()
    assert status["bytes"] <= budget
    assert status["resident"] < len(small)     # some hierarchies were evicted
    srv.close()


def test_server_step_returns_every_completion(gpu):
    """A step returns at most max_done finished requests; the rest come back
    from the following steps, each with its latency taken when it finished."""
    import ctypes
    import torch
    srv = U.Server(1 << 34, max_active=16, precision=U.FP32, tol_mode=U.TOL_REL, use_graph=1)
    for k, (name, pb) in enumerate(SIMS):
        srv.register(k, pb.n, pb.h, pb.solid, pb.wall)
    keep = [(_dev(_rhs(pb, 3)), _dev(np.zeros(pb.N))) for _, pb in SIMS]
    reqs = [srv.submit(k, r.data_ptr(), p.data_ptr(), 0, 500, 1e-6) for k, (r, p) in enumerate(keep)]
    done = (ctypes.c_int * 1)()
    n = ctypes.c_int(0)
    got = []
    t0 = time.time()
    while len(got) < len(reqs):
        assert srv.lib.L.uaamg_server_step(srv.h, done, 1, ctypes.byref(n)) == U.OK
        assert n.value <= 1
        got += [done[0]] if n.value else []
        assert time.time() - t0 < 60, "server did not finish"
    torch.cuda.synchronize()
    assert sorted(got) == sorted(reqs)
    for r in reqs:
        st = srv.request(r)
        assert st is not None and st.status == U.OK and st.ms_total > 0, r
    srv.close()
