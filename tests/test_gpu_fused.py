"""GPU: the fused FP32 3D CG kernels (p update + operator, r update in the
level-0 down kernel, x update on a second stream) against the separate-pass
form of the same engine, including the stop paths (converged, max_iter,
breakdown on a zero rhs, a solve that stops inside a captured graph chunk)."""
import numpy as np
import pytest

import cases
import uaamg_amd as U

pytestmark = pytest.mark.gpu

NAMES = ("3d_24x20x18_aniso_solids", "3d_40x36x34_walls_solids", "3d_130x20x12", "3d_200x70x40_walls_solids")
FUSED = [c for c in cases.parity_cases() if c[0] in NAMES]


def _ctx(pb, pmv, **opts):
    forms = opts.pop("forms", U.FORM_DEFAULT)
    c = U.Context(forms=forms if pmv else forms & ~U.FORM_PMV, **opts)
    c.set_grid(pb.n, pb.h, pb.solid, pb.wall)
    return c


def _solve(pb, pmv, rhs, max_iter, tol, **opts):
    c = _ctx(pb, pmv, precision=U.FP32, tol_mode=U.TOL_REL, **opts)
    p, st = c.solve(rhs, max_iter, tol)
    c.close()
    return p, st


@pytest.mark.parametrize("use_graph", [0, 1])
@pytest.mark.parametrize("name,pb", FUSED, ids=[c[0] for c in FUSED])
def test_fused_matches_separate_passes(gpu, name, pb, use_graph):
    """Converged solves: the same iterations (+-1) and p within 2e-6 of
    max|p| (the fused updates round in FP32 FMAs)."""
    rhs = cases.gaussian_rhs(pb)
    p1, s1 = _solve(pb, True, rhs, 500, 1e-6, use_graph=use_graph, check_every=4)
    p0, s0 = _solve(pb, False, rhs, 500, 1e-6, use_graph=use_graph, check_every=4)
    assert s1.kernel_set == 2 and s0.kernel_set == 0
    assert s1.forms_used & U.FORM_SPLIT and not s0.forms_used & (U.FORM_PMV | U.FORM_SPLIT)
    assert s1.status == U.OK and s0.status == U.OK
    assert abs(s1.iters - s0.iters) <= 1, (s1.iters, s0.iters)
    m = pb.fluid_interior_mask()
    a, b = p1[m] - p1[m].mean(), p0[m] - p0[m].mean()
    assert np.abs(a - b).max() <= 2e-6 * np.abs(b).max()


# a grid whose level 1 runs on the record row kernels (64 < nx/2 <= 128), as
# at 256^3: the stop paths with those kernels in the V-cycle
L1REC = [k for k, c in enumerate(FUSED) if c[0] == "3d_200x70x40_walls_solids"][0]


@pytest.mark.parametrize("case", [0, L1REC])
@pytest.mark.parametrize("max_iter", [1, 5, 6])
def test_fused_max_iter_stop(gpu, max_iter, case):
    """Stopped by max_iter (inside and at the end of a 4-iteration graph
    chunk): the x of the last iteration is in p, as in the separate passes."""
    name, pb = FUSED[case]
    rhs = cases.gaussian_rhs(pb)
    p1, s1 = _solve(pb, True, rhs, max_iter, 1e-12, use_graph=1, check_every=4)
    p0, s0 = _solve(pb, False, rhs, max_iter, 1e-12, use_graph=1, check_every=4)
    assert s1.status == s0.status == U.NOT_CONVERGED
    assert s1.iters == s0.iters == max_iter
    m = pb.fluid_interior_mask()
    assert np.abs(p1[m] - p0[m]).max() <= 1e-5 * np.abs(p0[m]).max()


@pytest.mark.parametrize("case", [0, L1REC])
def test_fused_zero_rhs(gpu, case):
    """A zero rhs stops at the first operator (p.Ap = 0 < 1e-15, pcg.cpp:100):
    the same status and iteration count as the separate passes, p unchanged."""
    name, pb = FUSED[case]
    rhs = np.zeros(pb.N)
    p0 = np.random.default_rng(3).standard_normal(pb.N)
    for pmv in (True, False):
        c = _ctx(pb, pmv, precision=U.FP32, tol_mode=U.TOL_ABS)
        p, st = c.solve(rhs, 50, 1e-8, p0=p0)
        c.close()
        assert st.status in (U.OK, U.BREAKDOWN), U.status_name(st.status)
        np.testing.assert_array_equal(p, p0)
        if pmv:
            ref = (st.status, st.iters)
        else:
            assert (st.status, st.iters) == ref


@pytest.mark.parametrize("case", [1, L1REC])
def test_fused_warm_start_and_repeat(gpu, case):
    """Two solves in a row on one context (buffers alternate between solves)
    agree bit for bit, and a warm start from the converged pressure (relres
    measured against its own, already small, initial residual) stays at that
    pressure."""
    import torch
    name, pb = FUSED[case]
    rhs = cases.gaussian_rhs(pb)
    c = _ctx(pb, True, precision=U.FP32, tol_mode=U.TOL_REL, use_graph=1, check_every=4)
    dev = torch.device("cuda:0")
    r = torch.from_numpy(rhs).to(dev)
    pa = torch.zeros(pb.N, dtype=torch.float64, device=dev)
    pb2 = torch.zeros(pb.N, dtype=torch.float64, device=dev)
    s1 = c.solve_device(r.data_ptr(), pa.data_ptr(), 0, 500, 1e-6)
    first = pa.clone()
    s2 = c.solve_device(r.data_ptr(), pb2.data_ptr(), 0, 500, 1e-6)
    s3 = c.solve_device(r.data_ptr(), pa.data_ptr(), 1, 500, 1e-6)
    torch.cuda.synchronize()
    c.close()
    assert s1.status == s2.status == U.OK and s1.iters == s2.iters
    np.testing.assert_array_equal(first.cpu().numpy(), pb2.cpu().numpy())   # deterministic, buffers reset per solve
    assert s3.status == U.OK and s3.iters < s1.iters, (U.status_name(s3.status), s3.iters)
    m = pb.fluid_interior_mask()
    a, b = pa.cpu().numpy()[m], first.cpu().numpy()[m]
    assert np.abs(a - b).max() <= 1e-5 * np.abs(b).max()


@pytest.mark.parametrize("opts,used,unused",
                         [({"forms": U.FORM_DEFAULT | U.FORM_XMASK}, U.FORM_XMASK, U.FORM_NT_X),
                          ({"xfork_level": 0, "xblocks": 2048}, U.FORM_NT, 0),
                          ({"xfork_level": 99, "xthreads": 512}, U.FORM_NT, 0),
                          ({"forms": U.FORM_DEFAULT & ~U.FORM_NT}, 0, U.FORM_NT)],
                         ids=["mask", "fork0_2048", "fork_tail_512", "temporal"])
def test_fused_x_update_forms(gpu, opts, used, unused):
    """The side-stream x update (x += alpha p, elementwise) in its other forms
    (mask read, other fork points and launch shapes), and plain instead of
    nontemporal loads and stores in it and the level-0 kernels, give
    bit-identical pressure to the default form: only where and how the same
    FP64 updates run changes.  uaamg_stats.forms_used shows that the form
    under test is the one that ran."""
    name, pb = FUSED[L1REC]
    rhs = cases.gaussian_rhs(pb)
    p_def, s_def = _solve(pb, True, rhs, 500, 1e-6, use_graph=1, check_every=4)
    p_alt, s_alt = _solve(pb, True, rhs, 500, 1e-6, use_graph=1, check_every=4, **opts)
    assert s_def.kernel_set == s_alt.kernel_set == 2
    assert s_def.forms_used & U.FORM_NT == U.FORM_NT, hex(s_def.forms_used)   # the default runs every nontemporal form
    assert s_alt.forms_used & used == used and not s_alt.forms_used & unused, hex(s_alt.forms_used)
    assert s_def.status == s_alt.status == U.OK and s_def.iters == s_alt.iters
    np.testing.assert_array_equal(p_alt, p_def)
