"""Learned warm start: converged fields from the batched solver -> network -> start state.

Pipeline (all solver and network work on the GPU through the C-ABI):
  sample_cases   seeded (Re, sphere diameter, centre) draws in the upstream third
                 of a channel, N_lag = round(pi D^2) Fibonacci nodes per sphere;
  generate       runs the draws as one batch (plbfs_batch_*) to the residual
                 tolerance and returns the converged fields;
  dataset        node inputs and normalised targets per field;
  train          Adam on the network (plbfs_mpnn_train);
  warm_state     predicted start state for a new draw, with the fallback to
                 the uniform start when its first-step residual is worse.
"""
from __future__ import annotations

import math
import sys
from dataclasses import dataclass

import numpy as np

from . import case as pc
from . import mpnn
from . import objects
from .solver import GpuBatch, GpuSolver, HostMesh


@dataclass
class Draw:
    re: float
    diameter: float
    centre: tuple

    @property
    def n_nodes(self):
        return objects.sphere_nodes_for_diameter(self.diameter)


def sample_cases(n, lengths, seed, re_range=(20.0, 200.0), d_range=(6.0, 14.0)):
    """Seeded draws: Re ~ U[re_range], D ~ U[d_range] cells, centre x uniform in
    the upstream third, y and z within 2 cells of the channel axis."""
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = lengths
    out = []
    for _ in range(n):
        re = float(rng.uniform(*re_range))
        d = float(rng.uniform(*d_range))
        x0 = float(rng.uniform(0.5 * d + 3.0, Lx / 3.0))
        y0 = float(0.5 * Ly + rng.uniform(-2.0, 2.0))
        z0 = float(0.5 * Lz + rng.uniform(-2.0, 2.0))
        out.append(Draw(re, d, (x0, y0, z0)))
    return out


def case_for(draw, precision=1, tol=1e-6, check=100):
    c = pc.channel_case(re=draw.re, length_scale=draw.diameter)
    c.precision = precision
    c.tolerance = tol
    c.check_interval = check
    return c


def ibm_args(draw, case, stiffness=5.0):
    X, dA = objects.fibonacci_sphere(draw.n_nodes, draw.centre, 0.5 * draw.diameter)
    return dict(x_ref=X, x=X, stiffness=stiffness, node_area=dA, dt=case.time_marching_step, h=1.0, kernel=1)


def uniform_state(case, geo, n):
    return case.initial_state(geo["cx"][:n], geo["cy"][:n], geo["cz"][:n])


def generate(pm, hm, draws, max_steps, tiling=None, precision=1):
    """Runs every draw in one batch to the tolerance; returns (fields, results)."""
    cases = [case_for(d, precision) for d in draws]
    bt = GpuBatch(hm, cases[0].params, cases[0].patches_for(pm.patch_names), [c.visc for c in cases])
    if tiling:
        bt.set_tiling(*tiling)
    geo = hm.geometry()
    N = pm.n_cells
    bt.set_state(-1, *uniform_state(cases[0], geo, N))
    for b, (d, c) in enumerate(zip(draws, cases)):
        bt.set_ibm(b, **ibm_args(d, c))
    done = 0
    while done < max_steps:
        n = min(5000, max_steps - done)
        rc, res = bt.run(n)
        done += n
        live = sum(1 for r in res if not r["converged"] and r["status"] == 0)
        print("  batch of %d: %d steps, %d cases still running" % (len(draws), done, live), file=sys.stderr,
              flush=True)
        if live == 0:
            break
    fields = [bt.state(b) for b in range(len(draws))]
    bt.close()
    return fields, res


def dataset(geo, n, draws, fields):
    X = np.stack([mpnn.node_inputs(geo, n, d.centre, d.diameter, d.re) for d in draws])
    U = pc.Case().mach_no
    T = np.stack([mpnn.targets(*f, U) for f in fields])
    return X, T


def predict(net, geo, n, draw):
    y = net.forward(mpnn.node_inputs(geo, n, draw.centre, draw.diameter, draw.re))
    return mpnn.fields(y, pc.Case().mach_no)


def first_step_residual(pm, hm, case, state, ibm):
    """max(u, v, w) residual of one step from the given state (the fallback test)."""
    s = GpuSolver(hm, dict(case.params, tolerance=0.0), case.patches_for(pm.patch_names))
    s.set_state(*state)
    s.set_ibm(**ibm)
    r = s.run(1)
    s.close()
    return max(r["residual"][1:])


def warm_state(pm, hm, net, draw, case):
    """Predicted state, or the uniform start when that is no better."""
    geo = hm.geometry()
    n = pm.n_cells
    ibm = ibm_args(draw, case)
    pred = predict(net, geo, n, draw)
    uni = uniform_state(case, geo, n)
    r_pred = first_step_residual(pm, hm, case, pred, ibm)
    r_uni = first_step_residual(pm, hm, case, uni, ibm)
    if not math.isfinite(r_pred) or r_pred >= r_uni:
        return uni, dict(used="uniform", r_pred=r_pred, r_uniform=r_uni)
    return pred, dict(used="prediction", r_pred=r_pred, r_uniform=r_uni)


def solve_to_tolerance(pm, hm, case, state, ibm, max_steps):
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    s.set_state(*state)
    s.set_ibm(**ibm)
    steps, ms = 0, 0.0
    while steps < max_steps:
        r = s.run(min(5000, max_steps - steps))
        steps += r["steps"]
        ms += r["ms"]
        if r["converged"] or r["rc"] != 0:
            break
    out = dict(rc=r["rc"], steps=steps, converged=r["converged"], ms=ms, residual=r["residual"])
    st = s.state()
    s.close()
    return out, st


def cfg4_mesh():
    """64 x 32 x 32 hex channel, interior vertices jittered by 15% (seed 12), cells permuted (seed 13)."""
    from . import meshgen
    return meshgen.box_mesh(64, 32, 32, lengths=(64.0, 32.0, 32.0), jitter=0.15, jitter_seed=12, permute_seed=13,
                            patch_types=pc.CHANNEL_PATCHES)


def main(argv=None):
    """Train on converged batch fields, then compare solves from the prediction
    and from the uniform start on held-out draws; prints one JSON report."""
    import argparse
    import json
    import sys
    import time
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=256)
    ap.add_argument("--test", type=int, default=8)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--max-steps", type=int, default=40000)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--save", default="")
    args = ap.parse_args(argv)
    pm = cfg4_mesh()
    hm = HostMesh.from_polymesh(pm, 0.01, 1.0, 0.0)
    geo = hm.geometry()
    N = pm.n_cells
    L = (64.0, 32.0, 32.0)
    draws = sample_cases(args.train + args.test, L, seed=args.seed)
    train_d, test_d = draws[: args.train], draws[args.train:]
    t0 = time.perf_counter()
    fields, conv = [], []
    for k in range(0, len(train_d), args.batch):
        f, res = generate(pm, hm, train_d[k:k + args.batch], args.max_steps)
        fields += f
        conv += res
        print("generated %d fields, %d converged, steps %d..%d (%.1f s)" % (
            len(fields), sum(r["converged"] for r in conv), min(r["steps"] for r in res),
            max(r["steps"] for r in res), time.perf_counter() - t0), file=sys.stderr, flush=True)
    X, T = dataset(geo, N, train_d, fields)
    src, dst, ef, _ = mpnn.graph_from_geometry(geo, N)
    net = mpnn.Mpnn(N, src, dst, ef, train=True)
    net.set_params(mpnn.init_params(args.seed))
    t1 = time.perf_counter()
    hist = net.train(X, T, mpnn.output_mask(N), epochs=args.epochs, lr=args.lr, seed=args.seed)
    t_train = time.perf_counter() - t1
    print("trained %d epochs in %.1f s, loss %.4g -> %.4g" % (args.epochs, t_train, hist[0], hist[-1]),
          file=sys.stderr, flush=True)
    params = net.params()
    if args.save:
        np.save(args.save, params)
    net.close()
    infer = mpnn.Mpnn(N, src, dst, ef, train=False)
    infer.set_params(params)
    rows = []
    for d in test_d:
        c32 = case_for(d, precision=1)
        c64 = case_for(d, precision=0)
        ibm = ibm_args(d, c32)
        st_w, info = warm_state(pm, hm, infer, d, c32)
        uni = uniform_state(c32, geo, N)
        r_w, s_w = solve_to_tolerance(pm, hm, c32, st_w, ibm, args.max_steps)
        r_u32, _ = solve_to_tolerance(pm, hm, c32, uni, ibm, args.max_steps)
        r_u64, s_64 = solve_to_tolerance(pm, hm, c64, uni, ibm, args.max_steps)
        uw = np.stack(s_w[1:], 1)
        u64 = np.stack(s_64[1:], 1)
        rows.append(dict(re=d.re, diameter=d.diameter, start=info["used"], r_pred=info["r_pred"],
                         r_uniform=info["r_uniform"], warm_fp32=r_w, rest_fp32=r_u32, rest_fp64=r_u64,
                         vel_rel_l2_vs_fp64=float(np.linalg.norm(uw - u64) / np.linalg.norm(u64))))
        print("test Re %.1f D %.2f: %s start; steps warm/fp32 %d, rest/fp32 %d, rest/fp64 %d; ms %.0f / %.0f / %.0f"
              % (d.re, d.diameter, info["used"], r_w["steps"], r_u32["steps"], r_u64["steps"], r_w["ms"],
                 r_u32["ms"], r_u64["ms"]), file=sys.stderr, flush=True)
    speed = [r["rest_fp64"]["ms"] / r["warm_fp32"]["ms"] for r in rows if r["warm_fp32"]["converged"]
             and r["rest_fp64"]["converged"]]
    report = dict(mesh="64x32x32 jittered hex channel (65,536 cells)", train_fields=len(fields),
                  train_converged=int(sum(r["converged"] for r in conv)), epochs=args.epochs,
                  train_seconds=t_train, loss_hist=[float(v) for v in hist], params=int(params.size),
                  tests=rows, speedup_fp64_rest_over_fp32_warm=speed)
    print(json.dumps(report))
    return report


if __name__ == "__main__":
    main()
