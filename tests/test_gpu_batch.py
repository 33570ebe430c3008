"""GPU: batched solves (uaamg_batch_*) -- many grids of uneven size, mask and
walls in one set of launches -- against the oracle, against single-grid
solves, alone vs in a mix, warm start, graph replay."""
import numpy as np
import pytest

import cases
import oracle_bind as ob
import uaamg_amd as U

pytestmark = pytest.mark.gpu

MEMBERS = [c for c in cases.parity_cases() if "porous" not in c[0]]


def _rhs(pb, seed):
    return 1e3 * cases.random_rhs(pb, np.random.default_rng(seed))


def _demean(p, m):
    q = p[m]
    return q - q.mean()


def _err(pb, p, po):
    m = pb.fluid_interior_mask()
    a, b = _demean(p, m), _demean(po, m)
    return float(np.abs(a - b).max() / np.abs(b).max())


def _solve_batch(members, prec, tol=1e-6, warm=None, p0=None, use_graph=1, rhs=None, seeds=None):
    import torch
    dev = torch.device("cuda:0")
    b = U.Batch(precision=prec, tol_mode=U.TOL_REL, use_graph=use_graph)
    rh, pd = [], []
    for k, (name, pb) in enumerate(members):
        b.add(pb.n, pb.h, pb.solid, pb.wall)
        r = rhs[k] if rhs is not None else _rhs(pb, seeds[k] if seeds else 3)
        rh.append(torch.from_numpy(r).to(dev))
        pd.append(torch.from_numpy(p0[k] if p0 is not None else np.zeros(pb.N)).to(dev))
    torch.cuda.synchronize()
    rc, st = b.solve_device([t.data_ptr() for t in rh], [t.data_ptr() for t in pd], warm, 500, tol)
    torch.cuda.synchronize()
    out = [t.cpu().numpy() for t in pd]
    # the autotuner's choice is one of its candidates (coordinate descent
    # over march density, march chunk and level-0 row tile)
    for d in {pb.dim for _, pb in members}:
        t = b.tuning(d)
        assert t is not None and t["march_target"] in (64, 256, 1024) and t["march_kc_min"] in (2, 4, 8), t
        assert t["rows_kc"] in (0, 8, 16), t
    b.close()
    return rc, st, out


@pytest.mark.parametrize("prec", [U.FP32, U.MIXED_BF16])
def test_batch_matches_oracle(gpu, prec):
    rc, st, ps = _solve_batch(MEMBERS, prec)
    for (name, pb), s, p in zip(MEMBERS, st, ps):
        po, res = ob.pcg(pb, _rhs(pb, 3), 500, 1e-10, tol_mode=1)
        assert s.status == 0, (name, U.status_name(s.status))
        assert _err(pb, p, po) <= 2e-5, name
    assert rc == 0


def test_batch_matches_single_grid_solves(gpu):
    """Same kernels, per-grid reductions: iterations and pressure as the
    single-grid FP32 solve of each member."""
    rc, st, ps = _solve_batch(MEMBERS, U.FP32)
    for (name, pb), s, p in zip(MEMBERS, st, ps):
        c = U.Context(precision=U.FP32, tol_mode=U.TOL_REL)
        c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
        p1, s1 = c.solve(_rhs(pb, 3), 500, 1e-6)
        c.close()
        assert abs(s.iters - s1.iters) <= 2, (name, s.iters, s1.iters)
        assert _err(pb, p, p1) <= 1e-5, name


def test_batch_member_alone_equals_in_mix(gpu):
    """A member's result does not depend on the other members: alone and in a
    mix of every other geometry (autotuned launch shapes may differ) the
    pressure agrees to 1e-6 of max|p| and the iteration count to +-1."""
    for k in (1, 5, 8):
        alone = _solve_batch([MEMBERS[k]], U.FP32)
        mix = _solve_batch(MEMBERS, U.FP32)
        pb = MEMBERS[k][1]
        assert abs(alone[1][0].iters - mix[1][k].iters) <= 1
        assert _err(pb, alone[2][0], mix[2][k]) <= 1e-6


def test_batch_graph_equals_eager(gpu):
    g = _solve_batch(MEMBERS, U.FP32, use_graph=1)
    e = _solve_batch(MEMBERS, U.FP32, use_graph=0)
    for k, (name, pb) in enumerate(MEMBERS):
        assert g[1][k].iters == e[1][k].iters, name
        assert np.array_equal(g[2][k], e[2][k]), name


def test_batch_warm_start(gpu):
    """warm_start = 1 from the converged pressure: every member is already
    converged or needs very few iterations, and p stays the solution."""
    rc, st, ps = _solve_batch(MEMBERS, U.FP32, tol=1e-6)
    rc2, st2, ps2 = _solve_batch(MEMBERS, U.FP32, tol=1e-6, warm=[1] * len(MEMBERS), p0=ps)
    for k, (name, pb) in enumerate(MEMBERS):
        assert st2[k].status == 0, name
        assert st2[k].iters <= 2, (name, st2[k].iters)
        assert _err(pb, ps2[k], ps[k]) <= 2e-5, name
