#!/usr/bin/env python3
"""Benchmark: PCG solves/s and HBM GB/s at 256^3 (relres 1e-6), one MI355X.

One step = one complete cold-start pressure solve of the 256^3 plume
workload (8 seeded solid spheres + floor slab, Neumann walls, zero-mean
divergence rhs), including the per-solve hierarchy set-up from the solid mask
on device, in the plain FP32 form (FP32 CG vectors and levels, FP64
reductions/scalars), stopping at relative residual 1e-6.  Inputs (mask, rhs)
are resident in HBM before the timed region.

Prints one JSON line on stdout; diagnostics go to stderr.
"""
import argparse
import ctypes
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "PCG solves/s & HBM GB/s at 256^3 (relres 1e-6); served p50/p99 latency"
HBM_PEAK_GBS = 8000.0
KNAMES = ["matvec_dot", "update", "down0", "up0", "pupdate", "coarse", "init", "other"]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def per_cell_bytes(cls, sv, sl, dim, mode):
    """Algorithmic HBM bytes per interior cell for one launch of a level-0 kernel
    class (each array read or written once; nbr = 1-byte neighbour mask).
    mode "plain": operator, CG update, down, up, p update as five kernels;
    "pmv": the p update fused into the operator pass; "split": in addition the
    r update rides on the down kernel and the update kernel only does x."""
    rho = 1.0 / 8 if dim == 3 else 1.0 / 4
    table = {
        "matvec_dot": 2 * sv + 1,          # read p, nbr; write Ap
        "update": 6 * sv + 1,              # read x, p, r, Ap, nbr; write x, r
        "down0": sv + 1 + rho * sl,        # read r, nbr; write coarse b
        "up0": 2 * sv + 1 + rho * sl,      # read r, nbr, coarse x; write z
        "pupdate": 3 * sv + 1,             # read z, p, nbr; write p
    }
    if mode in ("pmv", "split"):
        table["matvec_dot"] = 4 * sv + 1   # read z, p_old, nbr; write p_new, Ap
    if mode == "split":
        table["update"] = 3 * sv           # read x, p; write x (p is 0 off the fluid cells: no mask read)
        table["down0"] = 3 * sv + 1 + rho * sl   # read r_old, Ap, nbr; write r_new, coarse b
    return table.get(cls)


def cpu_baseline(pb, k_iters, budget_s=30.0):
    """Time the reference CPU solver (oracle/_ref, compiled from the reference
    sources) or, when absent, the oracle restatement, single thread, on a
    bounded sample of the same workload: solves with max_iter = 1 and 3 give
    t_iter = (t3 - t1) / 2 and t_setup = t1 - t_iter; a full solve is
    extrapolated as t_setup + k * t_iter with k the reference's own iteration
    count for relres 1e-6 (the FP64 GPU solve's, which equals the oracle's
    within one)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_bind as ob
    kind = "reference" if ob.ref_available() else "port"
    opb = ob.Problem(pb.n, pb.L, pb.solid)
    tol = 1e-30   # never met: the sample runs exactly max_iter iterations

    def run(mi):
        t = time.perf_counter()
        if kind == "reference":
            ob.ref_pcg(opb, pb.rhs, mi, tol)
        else:
            ob.pcg(opb, pb.rhs, mi, tol)
        return time.perf_counter() - t

    t1 = run(1)
    t3 = run(3)
    t_iter = max((t3 - t1) / 2, 1e-9)
    t_setup = max(t1 - t_iter, 0.0)
    t_solve = t_setup + k_iters * t_iter
    cpu = platform.processor() or platform.machine()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    cpu = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {
        "value": 1.0 / t_solve,
        "unit": "solves/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"{'oracle/_ref (reference PCG3D+UAAMGPreconditioner3D, g++ -O2)' if kind == 'reference' else 'oracle restatement (C, gcc -O2)'}"
                   f" FP64 1 thread on {pb.name}: max_iter=1 took {t1:.2f}s, max_iter=3 took {t3:.2f}s;"
                   f" solve extrapolated as setup {t_setup:.2f}s + {k_iters} x {t_iter:.2f}s/iter"
                   f" ({k_iters} = the FP64 iteration count at relres 1e-6);"
                   f" host {cpu}, nproc {os.cpu_count()}"),
        "s_per_iter": t_iter,
    }


def pmc_summary(fetch_csv, write_csv, out=sys.stdout):
    """Per-dispatch and per-step HBM bytes from two rocprofv3 --pmc runs
    (FETCH_SIZE, then WRITE_SIZE, counter_collection.csv) of
    `bench.py --steps K --warmup W --copy-calibration ...`.  Both counters
    are scaled by what they read for the calibration's 256 MiB device copies
    (the largest __amd_rocclr_copyBuffer dispatches).  A step (set-up from
    the mask, zeroing of p, the solve) begins at a k_nbr0 dispatch; the
    per-step total sums every dispatch from the first k_nbr0 on, divided by
    the number of solves (k_ctrl_init dispatches)."""
    import csv

    def load(path):
        rows = list(csv.DictReader(open(path)))
        rows.sort(key=lambda r: int(r["Dispatch_Id"]))
        return rows

    fr, wr = load(fetch_csv), load(write_csv)
    if len(fr) != len(wr):
        raise SystemExit(f"dispatch counts differ: {len(fr)} vs {len(wr)}")
    calib = 268435456.0
    cf = [float(r["Counter_Value"]) for r in fr if r["Kernel_Name"] == "__amd_rocclr_copyBuffer"]
    cw = [float(r["Counter_Value"]) for r in wr if r["Kernel_Name"] == "__amd_rocclr_copyBuffer"]
    big_f, big_w = sorted(cf)[-3:], sorted(cw)[-3:]
    kf = calib / (np.mean(big_f) * 1024.0)
    kw = calib / (np.mean(big_w) * 1024.0)
    per = {}
    start = next(i for i, r in enumerate(fr) if "k_nbr0" in r["Kernel_Name"])
    nsolve = sum(1 for r in fr[start:] if "k_ctrl_init" in r["Kernel_Name"])
    tot_f = tot_w = 0.0
    for a, b in zip(fr[start:], wr[start:]):
        f = float(a["Counter_Value"]) * 1024.0 * kf
        w = float(b["Counter_Value"]) * 1024.0 * kw
        tot_f += f
        tot_w += w
        name = a["Kernel_Name"].split("(")[0]
        key = (name, int(a["Grid_Size"]))
        e = per.setdefault(key, [0, 0.0, 0.0])
        e[0] += 1; e[1] += f; e[2] += w
    print(f"calibration: the 3 copies of 268435456 bytes read FETCH_SIZE {np.mean(big_f):.1f} KiB and WRITE_SIZE "
          f"{np.mean(big_w):.1f} KiB per dispatch, so bytes = FETCH_SIZE x {kf * 1024:.1f} and "
          f"WRITE_SIZE x {kw * 1024:.1f} per KiB", file=out)
    print(f"steps counted: {nsolve} (dispatches {len(fr) - start})", file=out)
    print(f"per step: fetch {tot_f / nsolve / 1e6:.1f} MB, write {tot_w / nsolve / 1e6:.1f} MB, "
          f"total {(tot_f + tot_w) / nsolve / 1e6:.1f} MB", file=out)
    print("", file=out)
    print("| kernel | threads | dispatches per step | fetch MB / dispatch | write MB / dispatch | total MB per step |", file=out)
    print("|---|---:|---:|---:|---:|---:|", file=out)
    for (name, grid), (n, f, w) in sorted(per.items(), key=lambda kv: -(kv[1][1] + kv[1][2])):
        if (f + w) / nsolve < 1e6:
            continue
        print(f"| `{name[:90]}` | {grid} | {n / nsolve:.1f} | {f / n / 1e6:.1f} | {w / n / 1e6:.1f} | "
              f"{(f + w) / nsolve / 1e6:.1f} |", file=out)
    return (tot_f + tot_w) / nsolve


def trace_summary(stats_csv, trace_csv, top=20, out=sys.stdout):
    """Markdown tables of a rocprofv3 --kernel-trace --stats run: the top
    kernels by total time, and the average duration per kernel and grid size
    (one row per level for the kernels every level runs)."""
    import csv
    rows = list(csv.DictReader(open(stats_csv)))
    print("| kernel | calls | avg us | total ms | % |", file=out)
    print("|---|---:|---:|---:|---:|", file=out)
    for r in rows[:top]:
        print(f"| `{r['Name'].split('(')[0][:90]}` | {r['Calls']} | {float(r['AverageNs']) / 1e3:.2f} | "
              f"{float(r['TotalDurationNs']) / 1e6:.2f} | {float(r['Percentage']):.1f} |", file=out)
    print(f"\nSum of kernel time: {sum(float(r['TotalDurationNs']) for r in rows) / 1e6:.1f} ms.\n", file=out)
    per = {}
    for r in csv.DictReader(open(trace_csv)):
        grid = int(r["Grid_Size_X"]) * int(r["Grid_Size_Y"]) * int(r["Grid_Size_Z"])
        key = (r["Kernel_Name"].split("(")[0], grid, r["VGPR_Count"], r["LDS_Block_Size"])
        e = per.setdefault(key, [0, 0.0])
        e[0] += 1
        e[1] += (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) / 1e3
    print("| kernel | threads | VGPRs | LDS bytes | launches | avg us |", file=out)
    print("|---|---:|---:|---:|---:|---:|", file=out)
    for (name, grid, vg, lds), (n, t) in sorted(per.items(), key=lambda kv: -kv[1][1])[:top]:
        print(f"| `{name[:90]}` | {grid} | {vg} | {lds} | {n} | {t / n:.2f} |", file=out)


def load_traffic(workload):
    """HBM bytes per launch of the dominant kernel measured with rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration), if a
    measurement for this workload is committed under profiles/."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
        return d.get(workload)
    except (OSError, ValueError):
        return None


def free_parts(n, solid, wall):
    """Flat masks of the fluid regions on which a solution is fixed only up
    to a constant: the connected fluid components (face neighbours, as the
    stencil couples them) that touch no Dirichlet wall.  Two solutions of the
    same system may differ by a constant on each (enclosed pockets of a
    random mask, or the whole grid when every wall is Neumann); Krylov
    iterates take it from the preconditioner, so comparisons remove each
    part's mean."""
    from scipy import ndimage
    dim = len(n)
    shape = tuple(v + 2 for v in n)[::-1]
    fl = (np.asarray(solid).reshape(shape) == 0)
    inner = np.zeros(shape, bool)
    inner[(slice(1, -1),) * dim] = True
    lab, nc = ndimage.label(fl & inner)
    touch = np.zeros(nc + 1, bool)
    # wall order x-, x+, y-, y+, z-, z+; axis a of the padded array is dim-1-a
    for f in range(2 * dim):
        if wall[f] == 0:
            continue
        ax = dim - 1 - f // 2
        idx = [slice(None)] * dim
        idx[ax] = 1 if f % 2 == 0 else n[f // 2]
        touch[np.unique(lab[tuple(idx)])] = True
    return [(lab == c).reshape(-1) for c in range(1, nc + 1) if not touch[c]]


def rel_err(a, b, fluid, parts):
    """max |a - b| / max |b| over the fluid cells, each free part's mean removed."""
    import torch
    a, b = a.clone(), b.clone()
    for m in parts:
        a[m] -= a[m].mean()
        b[m] -= b[m].mean()
    return float((a[fluid] - b[fluid]).abs().max() / b[fluid].abs().max().clamp_min(1e-300))


def bench_batch(args):
    """Batched mixed-size stream: 64 2D grids (side log-uniform 64..1024) and
    32 3D grids (side 32..160), random masks and walls, relres 1e-6.  The
    plain form solves them one after another with single-grid FP32 contexts
    (eager and hipGraph); the batched form runs all 96 in one set of launches
    (FP32, and BF16 levels with FP32 CG and FP64 reductions; hipGraph replay,
    autotuned launch shapes).  Hierarchies are resident in every form (built
    before the timed region).  Errors are measured against the FP64 GPU path,
    which matches the FP64 oracle to 1e-9 of max|p| (parity tests)."""
    import torch
    import uaamg_amd as U
    from uaamg_amd import workloads as W
    t0 = time.perf_counter()
    pbs = W.batch_mixed(seed=0)
    log(f"[batch] generated {len(pbs)} grids ({sum(p.cells for p in pbs) / 1e6:.1f} M cells) in {time.perf_counter() - t0:.1f}s")
    dev = torch.device("cuda:0")
    rhs = [torch.from_numpy(p.rhs).to(dev) for p in pbs]
    tol = 1e-6

    def plain(prec, use_graph, reps):
        ctxs = []
        for p in pbs:
            c = U.Context(precision=prec, tol_mode=U.TOL_REL, use_graph=use_graph, max_iter=1000)
            c.set_grid(p.n, p.h, p.solid, p.wall)
            ctxs.append(c)
        outs = [torch.zeros(p.N, dtype=torch.float64, device=dev) for p in pbs]
        for k, c in enumerate(ctxs):      # warm-up (and graph capture)
            c.solve_device(rhs[k].data_ptr(), outs[k].data_ptr(), 0, 1000, tol)
        torch.cuda.synchronize()
        t = time.perf_counter()
        iters = []
        for _ in range(reps):
            for k, c in enumerate(ctxs):
                outs[k].zero_()
                st = c.solve_device(rhs[k].data_ptr(), outs[k].data_ptr(), 0, 1000, tol)
                iters.append(st.iters)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        for c in ctxs:
            c.close()
        return dt, outs, iters

    tuned = {}

    def batched(prec, reps, hierarchy_mode=U.HIERARCHY_REFERENCE):
        b = U.Batch(precision=prec, tol_mode=U.TOL_REL, use_graph=1, max_iter=1000, hierarchy_mode=hierarchy_mode)
        for p in pbs:
            b.add(p.n, p.h, p.solid, p.wall)
        outs = [torch.zeros(p.N, dtype=torch.float64, device=dev) for p in pbs]
        R = [t.data_ptr() for t in rhs]
        P = [t.data_ptr() for t in outs]
        b.solve_device(R, P, None, 1000, tol)      # warm-up: autotuner + graph capture
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(reps):
            for o in outs:
                o.zero_()
            rc, st = b.solve_device(R, P, None, 1000, tol)
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t) / reps
        if hierarchy_mode == U.HIERARCHY_REFERENCE:
            tuned[prec] = {f"{d}d": b.tuning(d) for d in (2, 3)}
        b.close()
        return dt, outs, [s.iters for s in st], rc

    reps = 2
    t_plain, out_plain, it_plain = plain(U.FP32, 0, reps)
    log(f"[batch] plain fp32 eager: {t_plain * 1e3:.1f} ms for {len(pbs)} solves")
    t_plain_g, _, _ = plain(U.FP32, 1, reps)
    log(f"[batch] plain fp32 graph: {t_plain_g * 1e3:.1f} ms")
    t_b32, out_b32, it_b32, rc32 = batched(U.FP32, reps)
    log(f"[batch] batched fp32: {t_b32 * 1e3:.1f} ms")
    t_bbf, out_bbf, it_bbf, rcbf = batched(U.MIXED_BF16, reps)
    log(f"[batch] batched bf16 levels: {t_bbf * 1e3:.1f} ms")
    t_bkf, out_bkf, it_bkf, rckf = batched(U.FP32, reps, U.HIERARCHY_KEEP_FLUID)
    log(f"[batch] batched fp32, keep-fluid hierarchy: {t_bkf * 1e3:.1f} ms")
    # converged FP64 references (relres 1e-10); the rhs is scaled by a power of
    # two (exact) so the absolute p.Ap < 1e-15 guard of the reference loop
    # cannot end a solve early
    out_64 = []
    for k, p in enumerate(pbs):
        sc = 2.0 ** float(np.round(np.log2(1e8 / max(float(rhs[k].norm()), 1e-300))))
        c = U.Context(precision=U.FP64, tol_mode=U.TOL_REL)
        c.set_grid(p.n, p.h, p.solid, p.wall)
        q = torch.zeros(p.N, dtype=torch.float64, device=dev)
        rs = rhs[k] * sc
        torch.cuda.synchronize()
        c.solve_device(rs.data_ptr(), q.data_ptr(), 0, 5000, 1e-10)
        torch.cuda.synchronize()
        c.close()
        out_64.append(q / sc)

    parts = [[torch.from_numpy(m).to(dev) for m in free_parts(p.n, p.solid, p.wall)] for p in pbs]
    fluids = [torch.from_numpy(p.fluid_mask()).to(dev) for p in pbs]

    def err(outs):
        return max(rel_err(a, b, f, q) for a, b, f, q in zip(outs, out_64, fluids, parts))

    n = len(pbs)
    best_plain = min(t_plain, t_plain_g)
    res = {
        "workload": "batch_mixed: 64 2D grids side logU[64,1024] + 32 3D grids side logU[32,160], random masks "
                    "(0-30 % boxes/spheres), random Neumann/Dirichlet walls, random-vortex rhs (seed 0), relres 1e-6",
        "members": n,
        "cells": int(sum(p.cells for p in pbs)),
        "plain_fp32_solves_per_s": round(n / t_plain, 2),
        "plain_fp32_graph_solves_per_s": round(n / t_plain_g, 2),
        "batched_fp32_solves_per_s": round(n / t_b32, 2),
        "batched_bf16_solves_per_s": round(n / t_bbf, 2),
        "speedup_bf16_batched_vs_best_plain": round(best_plain / t_bbf, 3),
        "max_rel_err_vs_fp64": {"plain_fp32": err(out_plain), "batched_fp32": err(out_b32),
                                "batched_bf16": err(out_bbf), "reference": "FP64 at relres 1e-10"},
        "iters_max": {"plain_fp32": int(max(it_plain)), "batched_fp32": int(max(it_b32)),
                      "batched_bf16": int(max(it_bbf))},
        "status": {"batched_fp32": U.status_name(rc32), "batched_bf16": U.status_name(rcbf)},
        "autotuned": {"fp32": tuned.get(U.FP32), "bf16": tuned.get(U.MIXED_BF16)},
        # the keep-fluid hierarchy (not the reference's algorithm): reported apart
        "keep_fluid_hierarchy": {"batched_fp32_solves_per_s": round(n / t_bkf, 2), "iters_max": int(max(it_bkf)),
                                 "max_rel_err_vs_fp64": err(out_bkf), "status": U.status_name(rckf)},
        "hierarchies": "resident (built before timing)",
    }
    log(f"[batch] {json.dumps(res)}")
    return res


def bench_keys(args):
    """The 1024 x 512 MAC pressure solve (solid disc, relres 1e-6) under each
    PCG key of the 2D factory (factory.cpp:19-23): pcg_uaamg, pcg_gmg (= pcg),
    pcg_amg and cg, FP64 and FP32.  Per key and precision: iterations and the
    median ms per solve over args.steps solves (inputs resident, grid set
    once, p zeroed before each solve)."""
    import torch
    import uaamg_amd as U
    from uaamg_amd import workloads as W
    pb = W.pressure2d_mac(seed=0)
    dev = torch.device("cuda:0")
    rhs_d = torch.from_numpy(pb.rhs).to(dev)
    p_d = torch.zeros(pb.N, dtype=torch.float64, device=dev)
    keys = [("pcg_uaamg", U.PRECOND_UAAMG), ("pcg_gmg", U.PRECOND_GMG), ("pcg_amg", U.PRECOND_AMG), ("cg", U.PRECOND_NONE)]
    rows = []
    for key, pc in keys:
        for pname, prec in (("fp64", U.FP64), ("fp32", U.FP32)):
            mi = 20000 if pc == U.PRECOND_NONE else 2000
            ctx = U.Context(precision=prec, tol_mode=U.TOL_REL, preconditioner=pc, use_graph=1)
            ctx.set_grid(pb.n, pb.h, pb.solid, pb.wall)
            times, st = [], None
            for k in range(args.warmup + args.steps):
                p_d.zero_()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                st = ctx.solve_device(rhs_d.data_ptr(), p_d.data_ptr(), 0, mi, 1e-6)
                torch.cuda.synchronize()
                if k >= args.warmup:
                    times.append(1e3 * (time.perf_counter() - t0))
            ctx.close()
            ms = float(np.median(times))
            row = {"key": key, "precision": pname, "iters": st.iters, "status": U.status_name(st.status),
                   "ms_per_solve": round(ms, 3), "us_per_iter": round(1e3 * ms / max(1, st.iters), 1)}
            log(f"[bench] keys {key:9s} {pname}: {st.iters:5d} iterations {ms:9.2f} ms/solve "
                f"{row['us_per_iter']:8.1f} us/iteration ({row['status']})")
            rows.append(row)
    return {"workload": f"{pb.name}: 2D MAC pressure solve 1024x512, solid disc, relres 1e-6, hipGraph chunks",
            "steps": args.steps, "rows": rows}


def bench_project(args):
    """Pressure projection around the solve (PressureProjection3D::project,
    pressure_3d.cpp:5-39) at 256^3: the plume's MAC velocities (the solve
    benchmark's rhs is their divergence), dt = 1, FP32 relres 1e-6, every
    array resident in HBM; each timed projection starts from the same
    velocities (restored outside the timed call)."""
    import torch
    import uaamg_amd as U
    from uaamg_amd import workloads as W
    pb = W.plume3d(seed=0, n=args.n, velocity=True)
    dev = torch.device("cuda:0")
    u0, v0, w0 = (torch.from_numpy(a).to(dev) for a in (pb.u, pb.v, pb.w))
    u, v, w = u0.clone(), v0.clone(), w0.clone()
    p = torch.zeros(pb.N, dtype=torch.float64, device=dev)
    ctx = U.Context(precision=U.FP32, tol_mode=U.TOL_REL, max_iter=1000, tol=1e-6, check_every=args.check_every)
    ctx.set_grid(pb.n, pb.h, pb.solid, pb.wall)

    def once():
        u.copy_(u0); v.copy_(v0); w.copy_(w0)
        torch.cuda.synchronize()
        t = time.perf_counter()
        st = ctx.project_device(u.data_ptr(), v.data_ptr(), w.data_ptr(), p.data_ptr(), 1.0, 1000, 1e-6)
        return time.perf_counter() - t, st

    for _ in range(max(1, args.warmup)):
        once()
    ts, its = [], []
    for _ in range(args.steps):
        dt_, st = once()
        if st.status != 0:
            raise RuntimeError(f"projection solve: {U.status_name(st.status)}")
        ts.append(dt_)
        its.append(st.iters)
    # divergence before / after (fluid cells), on the device
    n = pb.n[0]

    def div(uu, vv, ww):
        U3 = uu.view(n + 2, n + 2, n + 1)
        V3 = vv.view(n + 2, n + 1, n + 2)
        W3 = ww.view(n + 1, n + 2, n + 2)
        d = (U3[1:-1, 1:-1, 1:] - U3[1:-1, 1:-1, :-1]) + (V3[1:-1, 1:, 1:-1] - V3[1:-1, :-1, 1:-1]) \
            + (W3[1:, 1:-1, 1:-1] - W3[:-1, 1:-1, 1:-1])
        return d * n
    fl = torch.from_numpy(pb.solid.reshape(n + 2, n + 2, n + 2)[1:-1, 1:-1, 1:-1] == 0).to(dev)
    d0 = div(u0, v0, w0)[fl]
    d1 = div(u, v, w)[fl]
    r0 = float((d0 - d0.mean()).abs().max())
    r1 = float((d1 - d1.mean()).abs().max())
    ctx.close()
    t = float(np.median(ts))
    res = {
        "workload": f"{pb.name} MAC velocities u = T curl(Psi) (the solve benchmark's rhs is their divergence), "
                    "dt = 1, FP32 relres 1e-6: rhs = div(u)/dt, p = 0, solve, face correction",
        "projections_per_s": round(1.0 / t, 3),
        "ms_per_projection": round(1e3 * t, 3),
        "iters": int(np.median(its)),
        "max_div_fluid_before": r0,
        "max_div_fluid_after": r1,
    }
    log(f"[project] {json.dumps(res)}")
    return res


def bench_serve(args):
    """Served stream: 200 concurrent simulations submit one pressure solve per
    frame (relres 1e-5, warm start from their previous pressure).  Closed
    loop first (`slots` requests outstanding, each completion followed by a
    frame of a uniformly drawn simulation) to measure the capacity, then
    Poisson arrivals over uniformly drawn simulations at 75 % of it; latency
    = submit -> completion, also per size class of the server.  Errors of sampled requests against an FP64 GPU solve of the
    same rhs (the FP64 path matches the FP64 oracle to 1e-9)."""
    import torch
    import uaamg_amd as U
    from uaamg_amd import workloads as W
    dev = torch.device("cuda:0")
    t0 = time.perf_counter()
    nsim = args.serve_sims
    sims = W.serve_sims(nsim, 7, dev)
    cells = sum(int(np.prod(s["n"])) for s in sims)
    log(f"[serve] generated {nsim} simulations ({cells / 1e6:.1f} M cells) in {time.perf_counter() - t0:.1f}s")
    budget = int(args.serve_budget_gb * (1 << 30))
    srv = U.Server(budget, max_active=args.serve_slots, precision=U.MIXED_BF16, tol_mode=U.TOL_REL, use_graph=1,
                   max_iter=500, flexible=1)
    for k, s in enumerate(sims):
        srv.register(k, s["n"], s["h"], s["solid"])
    p = [torch.zeros(int(np.prod(s["shape"])), dtype=torch.float64, device=dev) for s in sims]
    rng = np.random.default_rng(11)
    tol = 1e-5
    frames = [0] * nsim

    def submit(k, keep=None):
        r = W.next_frame(sims[k], rng) if frames[k] else sims[k]["rhs"]
        r = r.clone()
        if keep is not None:
            keep.append(r)
        # the rhs is produced on torch's stream and read on the server's:
        # it must be complete before the server can admit the request
        torch.cuda.current_stream().synchronize()
        req = srv.submit(k, r.data_ptr(), p[k].data_ptr(), 1 if frames[k] else 0, 500, tol)
        frames[k] += 1
        return req, r

    # warm-up: the first (cold) frame of every simulation, drained
    live = {}
    for k in range(nsim):
        req, r = submit(k)
        live[req] = (k, r)
    while live:
        for d in srv.step():
            live.pop(d, None)
    torch.cuda.synchronize()
    # closed loop with the open loop's request mix: `slots` requests
    # outstanding, every completion followed by a frame of a uniformly drawn
    # simulation (a simulation drawn again while its frame runs queues behind
    # it), so the capacity is that of the mix the arrivals bring
    t_closed = time.perf_counter()
    done_closed = 0
    for _ in range(args.serve_slots):
        req, r = submit(int(rng.integers(0, nsim)))
        live[req] = r
    while time.perf_counter() - t_closed < args.serve_seconds:
        for d in srv.step():
            live.pop(d)
            done_closed += 1
            req, r = submit(int(rng.integers(0, nsim)))
            live[req] = r
    cap = done_closed / (time.perf_counter() - t_closed)
    while live:
        for d in srv.step():
            live.pop(d, None)
    log(f"[serve] closed-loop capacity {cap:.1f} solves/s ({done_closed} solves)")

    # open loop: Poisson arrivals at 75 % of capacity
    lam = 0.75 * cap
    nreq = max(50, int(lam * args.serve_seconds))
    arr = np.cumsum(rng.exponential(1.0 / lam, nreq))
    order = rng.integers(0, nsim, nreq)
    pending_sim = {k: [] for k in range(nsim)}
    t_start = time.perf_counter()
    nxt = 0
    live = {}
    lat = []
    iters = []
    lat_class = {}   # size class (the server's: 2D / 3D, fewer / at least 2^18 padded cells) -> latencies
    sample = {}
    while nxt < nreq or live:
        now = time.perf_counter() - t_start
        while nxt < nreq and arr[nxt] <= now:
            k = int(order[nxt])
            keep = [] if len(sample) < 8 and nxt % 7 == 0 else None
            req, r = submit(k, keep)
            live[req] = (k, time.perf_counter())
            if keep is not None:
                sample[req] = (k, r, frames[k])
            nxt += 1
        for d in srv.step():
            k, ts = live.pop(d)
            st = srv.request(d)
            lat.append(st.ms_total)
            iters.append(st.iters)
            cls = ("3d" if len(sims[k]["n"]) == 3 else "2d") + ("_large" if int(np.prod(sims[k]["shape"])) >= (1 << 18) else "_small")
            lat_class.setdefault(cls, []).append(st.ms_total)
            if d in sample:
                sample[d] = sample[d] + (p[k].clone(),)
                torch.cuda.current_stream().synchronize()   # copied before the next frame of k can write p[k]
        if not live and nxt < nreq:
            time.sleep(max(0.0, min(0.002, arr[nxt] - (time.perf_counter() - t_start))))
    wall = time.perf_counter() - t_start
    torch.cuda.synchronize()
    # errors of sampled requests vs an FP64 solve of the same rhs
    errs = []
    for req, item in sample.items():
        if len(item) < 4:
            continue
        k, r, fr, pserved = item
        s = sims[k]
        c = U.Context(precision=U.FP64, tol_mode=U.TOL_REL)
        c.set_grid(s["n"], s["h"], s["solid"])
        pref = torch.zeros_like(r)
        c.solve_device(r.data_ptr(), pref.data_ptr(), 0, 1000, 1e-10)
        c.close()
        f = s["fluid"]
        parts = [torch.from_numpy(m).to(f.device) for m in free_parts(s["n"], s["solid"], [0] * 6)]
        errs.append(rel_err(pserved, pref, f, parts))
    st = srv.status()
    srv.close()
    lat = np.array(lat)
    res = {
        "workload": f"serving: {nsim} concurrent simulations (2D side logU[64,2048], 3D side logU[32,256], "
                    "Neumann walls, 0-3 solid balls), frame rhs = previous rhs shifted one cell + 5 % noise, "
                    "relres 1e-5, warm start, BF16 levels; Poisson arrivals at 75 % of the closed-loop capacity",
        "capacity_solves_per_s": round(cap, 2),
        "arrival_rate_per_s": round(lam, 2),
        "requests": int(nreq),
        "solves_per_s": round(nreq / wall, 2),
        "p50_ms": round(float(np.percentile(lat, 50)), 3),
        "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_iters": round(float(np.mean(iters)), 2),
        "by_size_class": {c: {"requests": len(v), "p50_ms": round(float(np.percentile(v, 50)), 3),
                              "p99_ms": round(float(np.percentile(v, 99)), 3)} for c, v in sorted(lat_class.items())},
        "hbm_budget_gb": args.serve_budget_gb,
        "slots": args.serve_slots,
        "max_rel_err_vs_fp64_sampled": max(errs) if errs else None,
        "sampled": len(errs),
        "resident_at_end": st["resident"],
    }
    log(f"[serve] {json.dumps(res)}")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="grid side (the metric is quoted at 256)")
    ap.add_argument("--precision", default="fp32", choices=["fp32", "fp64", "bf16", "bf16r", "mixed64"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--check-every", type=int, default=4, help="CG iterations queued between host convergence checks")
    ap.add_argument("--profile-all", action="store_true", help="time every kernel class (stderr)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched mixed-size stream")
    ap.add_argument("--batch-only", action="store_true", help="only the batched mixed-size stream")
    ap.add_argument("--no-serve", action="store_true", help="skip the served stream")
    ap.add_argument("--serve-only", action="store_true", help="only the served stream")
    ap.add_argument("--serve-sims", type=int, default=200)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-budget-gb", type=float, default=8.0)
    ap.add_argument("--serve-slots", type=int, default=64)
    ap.add_argument("--keys-only", action="store_true",
                    help="only the 1024x512 2D solve under each PCG key (pcg_uaamg, pcg_gmg, pcg_amg, cg)")
    ap.add_argument("--no-project", action="store_true", help="skip the projection measurement")
    ap.add_argument("--no-keep-fluid", action="store_true", help="skip the keep-fluid hierarchy measurement")
    ap.add_argument("--project-only", action="store_true", help="only the projection measurement")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="summarize two rocprofv3 --pmc runs (FETCH_SIZE, WRITE_SIZE) of this benchmark and exit")
    ap.add_argument("--trace-summary", nargs=2, metavar=("KERNEL_STATS_CSV", "KERNEL_TRACE_CSV"),
                    help="tabulate a rocprofv3 --kernel-trace --stats run of this benchmark and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run a plain 256 MiB device copy (3x), the reference kernel for scaling "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE readings to bytes")
    args = ap.parse_args()

    if args.pmc_summary:
        pmc_summary(*args.pmc_summary)
        return
    if args.trace_summary:
        trace_summary(*args.trace_summary)
        return
    if args.gpus != 1:
        print(f"bench.py: the benchmark runs on one GPU only (asked for --gpus {args.gpus}); "
              "independent solves shard per GPU with no communication; a single solve would need z-slab halos.", file=sys.stderr)
        sys.exit(2)

    if args.batch_only:
        print(json.dumps({"batched": bench_batch(args)}), flush=True)
        return
    if args.serve_only:
        print(json.dumps({"serving": bench_serve(args)}), flush=True)
        return
    if args.project_only:
        print(json.dumps({"projection": bench_project(args)}), flush=True)
        return
    if args.keys_only:
        print(json.dumps({"keys_2d": bench_keys(args)}), flush=True)
        return

    import torch
    import uaamg_amd as U
    from uaamg_amd import workloads as W

    prec = {"fp32": U.FP32, "fp64": U.FP64, "bf16": U.MIXED_BF16, "bf16r": U.MIXED_BF16R, "mixed64": U.MIXED_F64}[args.precision]
    t0 = time.perf_counter()
    pb = W.plume3d(seed=0, n=args.n)
    log(f"[bench] generated {pb.name} in {time.perf_counter() - t0:.1f}s, solid fraction {pb.solid.mean():.3f}")

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        src = torch.ones(64 << 20, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] copy calibration: 3 x copy of {src.numel() * 4} bytes (read {src.numel() * 4}, write {src.numel() * 4})")
        del src, dst
    rhs_d = torch.from_numpy(pb.rhs).to(dev)
    solid_d = torch.from_numpy(pb.solid).to(dev)
    p_d = torch.zeros(pb.N, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    ctx = U.Context(precision=prec, tol_mode=U.TOL_REL, max_iter=1000, tol=1e-6, check_every=args.check_every)
    grid = U.capi.make_grid(pb.n, pb.h, np.zeros(1, np.uint8), pb.wall)
    grid.solid = ctypes.c_void_p(solid_d.data_ptr())
    ctx.n = pb.n; ctx.dim = pb.dim; ctx.N = pb.N

    def step():
        rc = ctx.lib.L.uaamg_set_grid_device(ctx.h, ctypes.byref(grid))
        if rc != 0:
            raise RuntimeError(f"set_grid_device: {U.status_name(rc)}")
        p_d.zero_()
        return ctx.solve_device(rhs_d.data_ptr(), p_d.data_ptr(), 0, 1000, 1e-6)

    for _ in range(args.warmup):
        st = step()
    torch.cuda.synchronize()

    # timed region: K complete solves
    ctx.set_opts(timing=0)
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    iters = []
    alg = 0.0
    for _ in range(args.steps):
        st = step()
        iters.append(st.iters)
        alg += st.alg_bytes
        if st.status != 0:
            raise RuntimeError(f"solve did not converge: {U.status_name(st.status)} after {st.iters} iterations")
    torch.cuda.synchronize()
    t_total = time.perf_counter() - t_start
    ms_per_step = 1e3 * t_total / args.steps
    solves_per_s = args.steps / t_total
    hbm_gbs = alg / t_total / 1e9

    # per-kernel device time with HIP events on the library's stream, over the
    # same K solves repeated with event timing on
    mask = 0xFF if args.profile_all else (1 << U.capi.K_MATVEC) | (1 << U.capi.K_UPDATE) | (1 << U.capi.K_DOWN0) | (1 << U.capi.K_UP0) | (1 << U.capi.K_PUPDATE)
    ctx.set_opts(timing=mask)
    ms_k = np.zeros(8); n_k = np.zeros(8)
    for _ in range(args.steps):
        st = step()
        ms_k += np.array(st.ms_kernel[:]); n_k += np.array(st.n_kernel[:])
    ctx.set_opts(timing=0)
    torch.cuda.synchronize()
    for c in range(8):
        if n_k[c]:
            log(f"[bench] {KNAMES[c]:>10s}: {ms_k[c] / n_k[c] * 1e3:8.1f} us/launch x {int(n_k[c] / args.steps)} per solve")
    sv = 8 if prec in (U.FP64, U.MIXED_F64) else 4
    sl = {U.FP64: 8, U.FP32: 4, U.MIXED_BF16R: 4, U.MIXED_F64: 4}.get(prec, 2)   # coarse-level vector element size
    mode = {0: "plain", 1: "pmv", 2: "split"}[st.kernel_set]   # the level-0 kernel set the solves ran
    # the dominant kernel on the solve's critical path: in the split form the
    # x update runs on a second stream beside the coarse levels, so it is
    # reported (stderr) but not a candidate
    cand = [c for c in range(5) if n_k[c] > 0 and not (mode == "split" and c == U.capi.K_UPDATE)]
    dom = max(cand, key=lambda c: ms_k[c])
    dname = KNAMES[dom]
    avg_s = ms_k[dom] / n_k[dom] / 1e3
    bytes_launch = per_cell_bytes(dname, sv, sl, pb.dim, mode) * pb.cells
    achieved = bytes_launch / avg_s / 1e9
    traffic = load_traffic(f"{pb.name}:{args.precision}:{mode}:{dname}")
    step_bytes = load_traffic(f"{pb.name}:{args.precision}:step")   # rocprof FETCH+WRITE of one whole step

    k_med = int(np.median(iters))
    out = {
        "metric": METRIC,
        "value": round(solves_per_s, 3),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp32": "fp32", "fp64": "fp64", "bf16": "fp32+bf16", "bf16r": "fp32+bf16", "mixed64": "fp64+fp32"}[args.precision],
        "data": "synthetic",
        "config": {
            "workload": f"{pb.name}: 256^3 MAC pressure solve, dx=1/256, Neumann walls, floor slab y<0.1 + 8 seeded spheres, "
                        "buoyant-plume zero-mean divergence rhs (seed 0); cold start, set-up included, relres 1e-6",
            "grid": list(pb.n),
            "relres": 1e-6,
            "precision": {"fp32": "FP32 vectors+levels, FP64 reductions", "fp64": "FP64",
                          "bf16": "FP32 CG, BF16 levels", "bf16r": "FP32 CG and level vectors, BF16 level records",
                          "mixed64": "FP64 CG and level-0 smoothing, FP32 coarse levels"}[args.precision],
            "iters": k_med,
            "hbm_gbs_solve": round(hbm_gbs, 1),
            "alg_bytes_per_solve": alg / args.steps,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": {"matvec_dot": "p update + operator (k_rows_pmv)" if mode != "plain" else "operator (k_matvec_dot)",
                       "update": "x update (k_xupdate, second stream)" if mode == "split" else "CG update (k_update)",
                       "down0": "level-0 down (k_rows_down0" + ("r)" if mode == "split" else ")"),
                       "up0": "level-0 up (k_rows_up0)", "pupdate": "p update (k_pupdate)"}[dname],
            "kernel_mode": mode,
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "bytes_per_launch": bytes_launch,
            "avg_launch_us": round(avg_s * 1e6, 2),
            # whole step (set-up, zeroing, solve) as rocprof counts it, over this run's step time
            "measured_bytes_per_step": step_bytes,
            "measured_gbs_solve": round(step_bytes / (ms_per_step * 1e-3) / 1e9, 1) if step_bytes else None,
        },
        "cpu_baseline": None,
    }
    if not args.no_keep_fluid:
        # the same workload with the keep-fluid hierarchy (a coarse cell is solid
        # only when all its children are; no counterpart in the reference):
        # reported beside the reference-compatible `value`, never as it
        p_ref = p_d.clone()
        ck = U.Context(precision=prec, tol_mode=U.TOL_REL, max_iter=1000, tol=1e-6, check_every=args.check_every,
                       hierarchy_mode=U.HIERARCHY_KEEP_FLUID)
        ck.n = pb.n; ck.dim = pb.dim; ck.N = pb.N

        def step_kf():
            rc = ck.lib.L.uaamg_set_grid_device(ck.h, ctypes.byref(grid))
            if rc != 0:
                raise RuntimeError(f"set_grid_device: {U.status_name(rc)}")
            p_d.zero_()
            return ck.solve_device(rhs_d.data_ptr(), p_d.data_ptr(), 0, 1000, 1e-6)

        for _ in range(max(1, args.warmup)):
            step_kf()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        kf_iters = []
        for _ in range(args.steps):
            stk = step_kf()
            kf_iters.append(stk.iters)
            if stk.status != 0:
                raise RuntimeError(f"keep-fluid solve did not converge: {U.status_name(stk.status)}")
        torch.cuda.synchronize()
        t_kf = (time.perf_counter() - t0) / args.steps
        ck.close()
        fl = torch.from_numpy(pb.fluid_mask()).to(dev)
        parts = [torch.from_numpy(m).to(dev) for m in free_parts(pb.n, pb.solid, pb.wall)]
        out["keep_fluid_hierarchy"] = {
            "hierarchy_mode": "keep_fluid (a coarse cell is solid only when all its children are; "
                              "not the reference's algorithm, so not `value`)",
            "solves_per_s": round(1.0 / t_kf, 3),
            "ms_per_step": round(1e3 * t_kf, 3),
            "iters": int(np.median(kf_iters)),
            "iters_reference_hierarchy": k_med,
            "max_rel_diff_vs_reference_hierarchy": rel_err(p_d, p_ref, fl, parts),
        }
        log(f"[bench] keep-fluid hierarchy: {out['keep_fluid_hierarchy']}")
    if not args.no_project or not args.no_batch:
        del rhs_d, solid_d, p_d
        ctx.close()
    if not args.no_project:
        out["projection"] = bench_project(args)
    if not args.no_batch:
        out["batched"] = bench_batch(args)
    if not args.no_serve:
        out["serving"] = bench_serve(args)
    if not args.no_cpu_baseline:
        try:
            # the reference's own iteration count: the FP64 form's (= the oracle's within one)
            c64 = U.Context(precision=U.FP64, tol_mode=U.TOL_REL)
            c64.set_grid(pb.n, pb.h, pb.solid, pb.wall)
            _, st64 = c64.solve(pb.rhs, 1000, 1e-6)
            c64.close()
            out["cpu_baseline"] = cpu_baseline(pb, st64.iters)
        except Exception as e:   # the baseline is reported, never required
            log(f"[bench] cpu baseline failed: {e}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
