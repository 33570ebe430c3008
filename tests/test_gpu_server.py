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
    # a batch member's device bytes (members leave out the single-grid form's
    # fused p-update buffers, UAAMG_FORM_PMV), as the server estimates them
    o32 = U.load().default_opts(precision=U.FP32, forms=U.FORM_DEFAULT & ~U.FORM_PMV)
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


def test_server_small_requests_do_not_wait_for_large_chunks(gpu):
    """Size classes: a 3D grid of 2^18 or more padded cells runs in a batch
    and on a stream of its own, so a small 2D simulation in a closed loop
    (resubmitted as soon as it finishes) completes frames at the pace of its
    own chunks while a 256^3 request runs.  Were both in one batch, every
    small chunk would wait for a large one, and the small simulation could
    finish at most (large chunks) / (chunks per small request) frames; the
    test asks for twice that."""
    import torch
    from uaamg_amd import workloads as W
    big = W.plume3d(seed=0, n=256)
    name, small = [c for c in SIMS if c[0] == "2d_48x40_aniso_solids"][0]
    srv = U.Server(1 << 34, max_active=8, precision=U.FP32, tol_mode=U.TOL_REL, use_graph=1)
    srv.register(0, big.n, big.h, big.solid, big.wall)
    srv.register(1, small.n, small.h, small.solid, small.wall)
    rb, pbig = _dev(big.rhs), _dev(np.zeros(big.N))
    rs = _dev(_rhs(small, 3))
    psmall = [_dev(np.zeros(small.N)) for _ in range(2)]
    req_big = srv.submit(0, rb.data_ptr(), pbig.data_ptr(), 0, 500, 1e-6)
    frames, cur, last, last_req = 0, 0, None, None
    pending = srv.submit(1, rs.data_ptr(), psmall[cur].data_ptr(), 0, 500, 1e-6)
    big_done = False
    t0 = time.time()
    while not big_done or pending is not None:
        for d in srv.step():
            if d == req_big:
                big_done = True
            elif d == pending:
                frames += 0 if big_done else 1
                last, last_req, pending = cur, pending, None
                if not big_done:   # the next frame of the small simulation, into the other buffer
                    cur ^= 1
                    psmall[cur].zero_()
                    # torch's stream only: a device-wide synchronize would wait for the large chunk
                    torch.cuda.current_stream().synchronize()
                    pending = srv.submit(1, rs.data_ptr(), psmall[cur].data_ptr(), 0, 500, 1e-6)
        assert time.time() - t0 < 60, "server did not finish"
    torch.cuda.synchronize()
    sb, ss = srv.request(req_big), srv.request(last_req)
    lockstep = -(-sb.iters // 4) / -(-ss.iters // 4)   # chunks of 4 iterations (check_every)
    print(f"large request {sb.iters} iterations, small {ss.iters}; small frames completed meanwhile: {frames} "
          f"(one batch would allow {lockstep:.1f})", flush=True)
    assert sb.status == 0 and ss.status == 0
    assert frames >= 2 * lockstep, (frames, sb.iters, ss.iters)
    c = U.Context(precision=U.FP32, tol_mode=U.TOL_REL)
    c.set_grid(big.n, big.h, big.solid, big.wall)
    p1, s1 = c.solve(big.rhs, 500, 1e-6)
    c.close()
    m = big.fluid_mask()
    a, b = pbig.cpu().numpy()[m], p1[m]
    a, b = a - a.mean(), b - b.mean()
    assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()
    c = U.Context(precision=U.FP32, tol_mode=U.TOL_REL)
    c.set_grid(small.n, small.h, small.solid, small.wall)
    p2, _ = c.solve(_rhs(small, 3), 500, 1e-6)
    c.close()
    assert _err(small, psmall[last].cpu().numpy(), p2) <= 1e-5
    srv.close()
