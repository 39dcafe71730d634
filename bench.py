#!/usr/bin/env python3
"""PLBFS benchmark on one MI355X.

Workload (default): 88^3 hexes split into 6 Kuhn tetrahedra each (4,088,832
cells, 8,224,128 faces) over a 2:1:1 channel (176 x 88 x 88 lattice units),
interior vertices jittered by +-10% (seed 4), cell labels permuted (seed 5),
uniform inflow u = 0.05, Re = 100 on D = 0.15 H, with an immersed rigid sphere
of diameter 0.15 H (2,000 Fibonacci nodes, Peskin 4-point kernel) at
(0.5 H, 0.5 H, 0.5 H).  A "step" is one time step (IBM update + 4 RK stages)
over every cell; value = cell-updates/s with all data resident in HBM.
Prints one JSON line.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "3d-plfbs-gpu_amd"))

import numpy as np  # noqa: E402

from plbfs import _native as nat  # noqa: E402
from plbfs import case as pc  # noqa: E402
from plbfs import meshgen  # noqa: E402
from plbfs import objects  # noqa: E402
from plbfs.solver import GpuSolver, HostMesh  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "cell-updates/s + wall time to residual 1e-6 (warm-start/FP32 vs plain FP64)"


def workload(name):
    if name == "tet88":
        n = 88
        pm = meshgen.box_mesh(n, n, n, lengths=(2.0 * n, n, n), tets=True, jitter=0.10, jitter_seed=4,
                              permute_seed=5, patch_types=pc.CHANNEL_PATCHES)
        # fneq_min 0.55: the reference's conservative stability floor on tau
        # (src/unstructured_mesh.cpp:981-983); the sphere wake diverges without it
        case = pc.channel_case(re=100.0, length_scale=0.15 * n, fneq_min=0.55)
        case.lsq_exact_det = 1  # the reference's LSQ determinant diverges on jittered tets
        desc = "88^3x6 Kuhn tets (4,088,832 cells) 2:1:1 channel, jitter 10%, permuted, Re=100"
        ibm = dict(n_nodes=2000, centre=(0.5 * n, 0.5 * n, 0.5 * n), radius=0.075 * n, stiffness=5.0, h=1.0,
                   kernel=1)
    elif name == "hex32":
        pm = meshgen.box_mesh(32, 32, 32, patch_types=pc.CAVITY_PATCHES)
        case = pc.cavity_case(re=100.0)
        desc = "lid-driven cavity 32^3 hex (32,768 cells), Re=100"
        ibm = None
    elif name == "chan96":
        pm = meshgen.box_mesh(96, 48, 48, lengths=(96.0, 48.0, 48.0), jitter=0.15, jitter_seed=2,
                              permute_seed=3, patch_types=pc.CHANNEL_PATCHES)
        case = pc.channel_case(re=40.0, length_scale=12.0, fneq_min=0.55)
        desc = "channel 96x48x48 hex (221,184 cells), jitter 15%, permuted, Re=40"
        ibm = dict(n_nodes=450, centre=(24.0, 24.0, 24.0), radius=6.0, stiffness=5.0, h=1.0, kernel=1)
    else:
        raise SystemExit("unknown workload %s" % name)
    return pm, case, desc, ibm


def algorithmic_bytes(info, P=8, Gb=8):
    """Per-launch algorithmic bytes of the two stage kernels (two-pass dataflow,
    every array element moved once at storage precision):
      gradient kernel: N*(4P state + 12P gradient write) + S*(3Gb weights + 4 index)
      flux kernel:     N*(32P + 3Gb) + F*(8 + 10Gb)
    N cells, F faces, S stencil slots; their sum per cell is the per-stage figure
    48P + s(3Gb+4) + (F/N)(8+10Gb) + 3Gb."""
    N, F, S = info["cells"], info["faces"], info["stencil_slots"]
    grad = N * 16 * P + S * (3 * Gb + 4)
    flux = N * (32 * P + 3 * Gb) + F * (8 + 10 * Gb)
    return grad, flux


def measured_traffic(workload, precision, block_cells, suffix=""):
    """HBM bytes per launch of the flux kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a device copy), or
    None when no summary matches this workload, precision and block size."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    key = "%s/%s/b%d%s" % (workload, precision, block_cells, suffix)
    return table.get(key)


def cpu_baseline(pm, case, st, seconds_hint):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    o = pyoracle.Oracle(pm, pc.BoundCase(case, pm))
    o.set_state(*st)
    t = time.perf_counter()
    r = o.run(1)
    dt = time.perf_counter() - t
    o.close()
    assert r["rc"] == 0, r["rc"]
    return dict(value=pm.n_cells / dt, unit="cell-updates/s", cores=1, kind="port",
                sample="1 time step (4 RK stages) of the same %d-cell mesh by the single-threaded FP64 C oracle "
                       "(restatement of src/Solver.cpp), %.1f s" % (pm.n_cells, dt),
                host=host_cpu())


def host_cpu():
    """CPU model of the host the baseline ran on, its logical CPUs and the
    ones this process may use."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return dict(model=model, nproc=os.cpu_count(), usable=usable)


def run_batch(args):
    """B = 128 cases (seeded Re, sphere size and position) on one jittered
    64 x 32 x 32 hex channel, FP32 storage, autotuned tiling; value =
    case-cell-updates/s."""
    from plbfs import warmstart as ws
    from plbfs.solver import GpuBatch
    B = args.cases
    pm = meshgen.box_mesh(64, 32, 32, lengths=(64.0, 32.0, 32.0), jitter=0.15, jitter_seed=12, permute_seed=13,
                          patch_types=pc.CHANNEL_PATCHES)
    hm = HostMesh.from_polymesh(pm, 0.01, 1.0, 0.0)
    draws = ws.sample_cases(B, (64.0, 32.0, 32.0), seed=2024)
    cases = [ws.case_for(d, precision=1 if args.precision == "fp32" else 0, tol=0.0, check=max(args.steps, 1))
             for d in draws]
    bt = GpuBatch(hm, cases[0].params, cases[0].patches_for(pm.patch_names), [c.visc for c in cases])
    geo = hm.geometry()
    bt.set_state(-1, *ws.uniform_state(cases[0], geo, pm.n_cells))
    nodes = 0
    for b, (d, c) in enumerate(zip(draws, cases)):
        bt.set_ibm(b, **ws.ibm_args(d, c))
        nodes += d.n_nodes
    t = time.perf_counter()
    if args.tiling:
        bc, cp, sg = (int(x) for x in args.tiling.split(","))
        bt.set_tiling(bc, cp, sg)
        print("tiling: block_cells %d cases_per_block %d stage %d (given)" % (bc, cp, sg), file=sys.stderr,
              flush=True)
    else:
        bc, cp, sg, tune_ms = bt.autotune(3)
        print("autotune: block_cells %d cases_per_block %d stage %d (%.3f ms/step, %.1f s)" % (
            bc, cp, sg, tune_ms, time.perf_counter() - t), file=sys.stderr, flush=True)
    if args.warmup > 0:
        bt.run(args.warmup)
    rc, res = bt.run(args.steps)
    assert rc == 0 and all(r["steps"] == args.warmup + args.steps for r in res), (rc, res[:2])
    ms = res[0]["ms"]
    value = pm.n_cells * B * args.steps / (ms * 1e-3)
    P = 4 if args.precision == "fp32" else 8
    # contract figure per case-cell-update: 4 stages x (48 P per case + geometry shared by the B cases)
    # (stencil slots = cell-face incidences F + NI; geometry stored at the field precision)
    F, N, NI = pm.n_faces, pm.n_cells, pm.n_internal
    geo_stage = (F + NI) * (3 * P + 4) + F * (8 + 10 * P) + N * 3 * P
    b_update = 4 * (48 * P + geo_stage / (N * B))
    step_gbs = value * b_update / 1e9
    out = {
        "metric": METRIC, "value": value, "unit": "case-cell-updates/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms / args.steps, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32" if P == 4 else "f64",
        "data": "synthetic mesh (seeded), seeded case draws, uniform-inflow initial state",
        "config": {"workload": "%d cases on a jittered 64x32x32 hex channel (65,536 cells), Re~U[20,200], "
                               "sphere D~U[6,14] cells (%d Lagrangian nodes in total), Peskin-4" % (B, nodes),
                   "cases": B, "cells": N, "faces": F, "block_cells": bc, "cases_per_block": cp,
                   "staged": bool(sg)},
        "roofline": {"bound": "hbm", "achieved": step_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": step_gbs / HBM_PEAK_GBS, "traffic": None, "kernel": "whole step (graph replay)",
                     "alg_bytes_per_case_cell_update": b_update,
                     # rocprofv3 FETCH/WRITE bytes of one kb_flux launch at this tiling, when measured
                     "kb_flux_traffic": measured_traffic("batch128", args.precision, bc, "c%d" % cp)},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def run_batch_to_tolerance(args):
    """The 128 batched cases (the seeded draws of run_batch, FP32 storage)
    each run to max(u, v, w) residual args.tol on the shared 64 x 32 x 32
    jittered hex channel, in chunks of args.chunk steps; a case freezes when it
    converges.  Reports every case's draw, steps, status and the wall time of
    the chunk in which it froze, and the whole batch's wall time."""
    from plbfs import warmstart as ws
    from plbfs.solver import GpuBatch
    B = args.cases
    pm = meshgen.box_mesh(64, 32, 32, lengths=(64.0, 32.0, 32.0), jitter=0.15, jitter_seed=12, permute_seed=13,
                          patch_types=pc.CHANNEL_PATCHES)
    hm = HostMesh.from_polymesh(pm, 0.01, 1.0, 0.0)
    draws = ws.sample_cases(B, (64.0, 32.0, 32.0), seed=2024)
    cases = [ws.case_for(d, precision=1, tol=args.tol, check=100) for d in draws]
    bt = GpuBatch(hm, cases[0].params, cases[0].patches_for(pm.patch_names), [c.visc for c in cases])
    bt.set_state(-1, *ws.uniform_state(cases[0], hm.geometry(), pm.n_cells))
    for b, (d, c) in enumerate(zip(draws, cases)):
        bt.set_ibm(b, **ws.ibm_args(d, c))
    if args.tiling:
        bt.set_tiling(*(int(x) for x in args.tiling.split(",")))
    else:
        bt.autotune(3)
    info = bt.layout_info()
    froze = [None] * B
    t0 = time.perf_counter()
    done_steps = 0
    res = None
    while done_steps < args.max_steps:
        n = min(args.chunk, args.max_steps - done_steps)
        rc, res = bt.run(n)
        done_steps += n
        wall = time.perf_counter() - t0
        for b, r in enumerate(res):
            if froze[b] is None and (r["converged"] or r["status"] != 0):
                froze[b] = wall
        live = sum(1 for f in froze if f is None)
        print("  %d steps, %.1f s, %d of %d cases still running" % (done_steps, wall, live, B), file=sys.stderr,
              flush=True)
        if live == 0:
            break
    wall = time.perf_counter() - t0
    rows = [dict(case=b, re=d.re, diameter=d.diameter, nodes=d.n_nodes, steps=r["steps"], converged=r["converged"],
                 status=r["status"], residual=max(r["residual"][1:]), wall_s=froze[b])
            for b, (d, r) in enumerate(zip(draws, res))]
    conv = [r for r in rows if r["converged"]]
    out = {"workload": "%d cases on a jittered 64x32x32 hex channel, FP32, each to residual %g" % (B, args.tol),
           "tolerance": args.tol, "block_cells": info["block_cells"], "cases_per_block": info["cases_per_block"],
           "wall_s": wall, "steps_run": done_steps, "converged": len(conv), "cases": rows}
    if conv:
        st = sorted(r["steps"] for r in conv)
        out["steps_median"] = st[len(st) // 2]
        out["steps_max"] = st[-1]
    print(json.dumps(out))


def run_to_tolerance(args, emit=True):
    """Wall time to max(u, v, w) residual 1e-6 on the chosen workload: plain
    FP64 and FP32 storage from the uniform start, and FP32 from the learned
    warm start (trained parameters shipped in plbfs/data, or --params) with
    its probe-based fallback; velocity rel-L2 of each FP32 result against
    FP64 at convergence.  One JSON line."""
    from plbfs import warmstart as ws
    pm, case, desc, ibm = workload(args.workload)
    if args.re:
        case.reynolds_no = args.re
        desc += " (run at Re=%g)" % args.re
    if args.no_ibm:
        ibm = None
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    cells = hm.geometry()
    N = pm.n_cells
    st = case.initial_state(cells["cx"][:N], cells["cy"][:N], cells["cz"][:N])
    draw = None
    ibm_kw = None
    if ibm:
        draw = ws.Draw(case.reynolds_no, 2.0 * ibm["radius"], tuple(ibm["centre"]), nodes=ibm["n_nodes"])
        ibm_kw = ws.ibm_args(draw, case, stiffness=ibm["stiffness"], h=ibm["h"])
        ibm_kw["kernel"] = ibm["kernel"]
    out, states = {}, {}
    params = ws.load_params(args.params or None) if draw is not None else None
    legs = [("fp64", 0, False), ("fp32", 1, False)] + ([("fp32_warm", 1, True)] if params is not None else [])
    if getattr(args, "legs", ""):
        legs = [leg for leg in legs if leg[0] in args.legs.split(",")]
    for name, prec, warm in legs:
        case.tolerance = args.tol
        case.check_interval = 100
        case.precision = prec
        if warm:
            net = ws.network_for(hm, N, params)
            r, state, info = ws.warm_solve(pm, hm, case, ibm_kw, net, draw, args.max_steps, probe=args.probe,
                                           label=name)
            net.close()
        else:
            r, state = ws.solve_to_tolerance(pm, hm, case, st, ibm_kw, args.max_steps, label=name)
            info = None
        out[name] = dict(steps=r["steps"], converged=r["converged"], rc=r["rc"], seconds=r["seconds"],
                         device_ms=r["ms"], residual=r["residual"])
        if info:
            out[name]["start"] = info
        print("%s: %d steps, converged %s, %.1f s" % (name, r["steps"], r["converged"], r["seconds"]),
              file=sys.stderr, flush=True)
        states[name] = state
    if getattr(args, "save_velocity", "") and "fp64" in states:
        np.save(args.save_velocity, np.stack(states["fp64"][1:], 1).astype(np.float32))
    if getattr(args, "ref_velocity", ""):
        # the FP64 solution of a separate run, rounded to float32 (relative
        # rounding 6e-8, far below the 1e-4 compared against)
        uref = np.load(args.ref_velocity).astype(np.float64)
        if uref.shape != (N, 3):
            raise SystemExit("--ref-velocity: shape %s, expected (%d, 3)" % (uref.shape, N))
        for name in states:
            u = np.stack(states[name][1:], 1)
            out[name]["velocity_rel_l2_vs_fp64_file"] = float(np.linalg.norm(u - uref) / np.linalg.norm(uref))
    if "fp64" in states:
        u64 = np.stack(states["fp64"][1:], 1)
        for name in states:
            if name != "fp64":
                u = np.stack(states[name][1:], 1)
                out[name]["velocity_rel_l2_vs_fp64"] = float(np.linalg.norm(u - u64) / np.linalg.norm(u64))
        if "fp32_warm" in out and out["fp32_warm"]["converged"] and out["fp64"]["converged"]:
            out["speedup_fp64_rest_over_fp32_warm"] = out["fp64"]["seconds"] / out["fp32_warm"]["seconds"]
        if "fp32" in out and out["fp32"]["converged"] and out["fp64"]["converged"]:
            out["speedup_fp64_rest_over_fp32_rest"] = out["fp64"]["seconds"] / out["fp32"]["seconds"]
    out["workload"] = desc
    out["tolerance"] = args.tol
    if emit:
        print(json.dumps(out))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="tet88")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ibm", action="store_true", help="leave the immersed sphere out")
    ap.add_argument("--precision", choices=("fp64", "fp32"), default="fp64",
                    help="storage precision of fields, gradients and face data")
    ap.add_argument("--cases", type=int, default=128, help="cases of the batch128 workload")
    ap.add_argument("--tiling", default="", help="batch128: cells per block, cases per workgroup, staging "
                                                 "(e.g. 64,8,0) instead of the autotuner")
    ap.add_argument("--time-to-tol", action="store_true",
                    help="measure wall time to the residual tolerance (FP64 and FP32) instead of throughput")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--chunk", type=int, default=1000, help="batch128 --time-to-tol: steps between progress reads")
    ap.add_argument("--max-steps", type=int, default=60000)
    ap.add_argument("--params", default="", help="trained warm-start parameters (.npy); default plbfs/data")
    ap.add_argument("--probe", type=int, default=200, help="steps of the warm-start fallback probe")
    ap.add_argument("--re", type=float, default=0.0, help="--time-to-tol: Reynolds number other than the workload's")
    ap.add_argument("--legs", default="", help="--time-to-tol: subset of fp64,fp32,fp32_warm to run")
    ap.add_argument("--save-velocity", default="",
                    help="--time-to-tol: write the FP64 leg's converged (u, v, w) as an N x 3 float32 .npy")
    ap.add_argument("--ref-velocity", default="",
                    help="--time-to-tol: FP64 velocity written by --save-velocity in an earlier run of the same "
                         "workload; each leg's velocity rel-L2 against it is reported")
    ap.add_argument("--no-ttt", action="store_true",
                    help="throughput only: skip the time-to-1e-6 comparison on the 96x48x48 channel")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a device copy of 1 GiB first (PMC FETCH/WRITE calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only (the solver has no multi-GPU path); use --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    if nat.lib().plbfs_device_count() <= 0:
        print("bench.py: no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_calibration:
        # a 1 GiB device copy; with HSA_ENABLE_SDMA=0 the runtime runs it as
        # its blit kernel, whose dispatches the counters see
        import torch
        a = torch.empty(1 << 27, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        print("copy calibration: 1073741824 B read + written", file=sys.stderr, flush=True)

    if args.workload == "batch128":
        return run_batch_to_tolerance(args) if args.time_to_tol else run_batch(args)
    if args.time_to_tol:
        return run_to_tolerance(args)
    t0 = time.perf_counter()
    pm, case, desc, ibm = workload(args.workload)
    if args.no_ibm:
        ibm = None
    t_mesh = time.perf_counter() - t0
    print("mesh %s: %d cells %d faces (%.1f s)" % (args.workload, pm.n_cells, pm.n_faces, t_mesh), file=sys.stderr,
          flush=True)
    case.tolerance = 0.0
    case.check_interval = max(args.steps, 1)
    case.precision = 1 if args.precision == "fp32" else 0
    t0 = time.perf_counter()
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    t_setup = time.perf_counter() - t0
    info = s.layout_info()
    cells = hm.geometry()
    st = case.initial_state(cells["cx"], cells["cy"], cells["cz"])
    s.set_state(*st)
    if ibm:
        X, dA = objects.fibonacci_sphere(ibm["n_nodes"], ibm["centre"], ibm["radius"])
        s.set_ibm(X, X, stiffness=ibm["stiffness"], node_area=dA, dt=case.time_marching_step, h=ibm["h"],
                  kernel=ibm["kernel"])
    print("setup %.1f s, layout %s" % (t_setup, info), file=sys.stderr, flush=True)

    if args.warmup > 0:
        s.run(args.warmup)
    # timed region: K steps, graph replay, device clock from hipEvents
    t0 = time.perf_counter()
    r = s.run(args.steps)
    wall = time.perf_counter() - t0
    assert r["rc"] == 0 and r["steps"] == args.steps, r
    ms_step = r["ms"] / args.steps
    value = pm.n_cells * args.steps / (r["ms"] * 1e-3)
    print("timed %d steps: %.3f ms/step device, %.3f ms/step host wall, residual %s" %
          (args.steps, ms_step, wall * 1e3 / args.steps, r["residual"]), file=sys.stderr, flush=True)

    # per-kernel durations over the same number of steps with events on the solver stream
    s.enable_timing(True)
    s.run(args.steps)
    kt = {k: s.kernel_stats(i) for k, i in (("grad", 0), ("flux", 1), ("ibm", 2), ("reduce", 3))}
    s.enable_timing(False)
    P = 4 if args.precision == "fp32" else 8
    g_bytes, f_bytes = algorithmic_bytes(info, P=P, Gb=P)
    flux_ms = kt["flux"][0] / max(kt["flux"][1], 1)
    grad_ms = kt["grad"][0] / max(kt["grad"][1], 1)
    flux_gbs = f_bytes / (flux_ms * 1e-3) / 1e9
    grad_gbs = g_bytes / (grad_ms * 1e-3) / 1e9
    ibm_ms = kt["ibm"][0] / max(args.steps, 1)
    print("kernels: flux %.3f ms (%.0f GB/s alg), grad %.3f ms (%.0f GB/s alg), ibm %.3f ms/step, reduce %.3f ms" %
          (flux_ms, flux_gbs, grad_ms, grad_gbs, ibm_ms, kt["reduce"][0] / max(kt["reduce"][1], 1)),
          file=sys.stderr)
    block_cells = int(info["block_cells"])
    # the committed counter runs include the immersed sphere (its body force is read by the flux kernel)
    traffic = measured_traffic(args.workload, args.precision, block_cells) if ibm else None

    cpu = None
    if not args.no_cpu:
        print("cpu baseline ...", file=sys.stderr, flush=True)
        cpu = cpu_baseline(pm, case, st, 20)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic mesh (seeded), uniform-inflow initial state",
        "config": {"workload": desc, "cells": pm.n_cells, "faces": pm.n_faces, "precision": args.precision,
                   "ibm": ({"nodes": ibm["n_nodes"], "diameter": 2 * ibm["radius"], "kernel": "peskin4",
                            "ms_per_step": ibm_ms} if ibm else None),
                   "block_cells": block_cells,
                   "lsq_determinant": "exact" if case.lsq_exact_det else "reference"},
        "roofline": {"bound": "hbm", "achieved": flux_gbs, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": flux_gbs / HBM_PEAK_GBS, "traffic": traffic, "kernel": "k_flux_update",
                     "kernel_ms": flux_ms, "alg_bytes_per_launch": f_bytes,
                     "grad_kernel": {"achieved": grad_gbs, "kernel_ms": grad_ms, "alg_bytes_per_launch": g_bytes},
                     "step_alg_bytes_per_cell": 4 * (g_bytes + f_bytes) / pm.n_cells,
                     "step_achieved": 4 * (g_bytes + f_bytes) / (ms_step * 1e-3) / 1e9},
        "cpu_baseline": cpu,
        "wall_s_per_step_host": wall / args.steps,
    }
    if not args.no_ttt:
        # the second half of the metric: wall time to residual 1e-6, FP64 from
        # rest against FP32 from the learned warm start, on the 96x48x48 hex
        # channel (on the 4.1M-tet case the three solves take ~30 min together)
        s.close()
        t_args = argparse.Namespace(**vars(args))
        t_args.workload, t_args.re, t_args.no_ibm = "chan96", 0.0, False
        t_args.tol, t_args.max_steps = 1e-6, 200000
        print("time to tolerance on chan96 ...", file=sys.stderr, flush=True)
        tt = run_to_tolerance(t_args, emit=False)
        out["time_to_tolerance"] = {
            "workload": tt["workload"], "tolerance": tt["tolerance"],
            "fp64_rest_s": tt["fp64"]["seconds"], "fp64_rest_steps": tt["fp64"]["steps"],
            "fp32_rest_s": tt["fp32"]["seconds"], "fp32_rest_steps": tt["fp32"]["steps"],
            "fp32_warm_s": tt.get("fp32_warm", {}).get("seconds"),
            "fp32_warm_steps": tt.get("fp32_warm", {}).get("steps"),
            "converged": all(tt[k]["converged"] for k in ("fp64", "fp32", "fp32_warm") if k in tt),
            "speedup_fp64_rest_over_fp32_warm": tt.get("speedup_fp64_rest_over_fp32_warm"),
            "speedup_fp64_rest_over_fp32_rest": tt.get("speedup_fp64_rest_over_fp32_rest"),
            "fp32_warm_velocity_rel_l2_vs_fp64": tt.get("fp32_warm", {}).get("velocity_rel_l2_vs_fp64"),
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
