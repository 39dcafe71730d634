"""Learned warm start: converged fields from the batched solver -> network -> start state.

Pipeline (all solver and network work on the GPU through the C-ABI):
  sample_cases   seeded (Re, sphere diameter, centre) draws in the upstream third
                 of a channel, N_lag = round(pi D^2) Fibonacci nodes per sphere;
  generate       runs the draws as one batch (plbfs_batch_*) to the residual
                 tolerance and returns the converged fields;
  dataset        node inputs and normalised targets per converged field;
  train          Adam on the network (plbfs_mpnn_train) over the disjoint union
                 of the training meshes' cell graphs, so that one network sees
                 hex and tet cells alike;
  warm_solve     predicted start state for a new draw, with the fallback to the
                 uniform start when the prediction's residual after a short
                 probe is no better than the uniform start's, then the plain
                 solve to the tolerance.

The channel has a uniform inflow, far-field sides and a fixed-density outlet,
so the uniform free stream is the start from rest of this flow: only the
immersed sphere drives the transient.  A residual taken on the first step
compares the two starts unfairly (the free stream is an exact steady state
away from the sphere until its wake forms), so the fallback compares the two
residuals after `probe` steps.
"""
from __future__ import annotations

import math
import os
import sys
import time
from dataclasses import dataclass

import numpy as np

from . import case as pc
from . import meshgen
from . import mpnn
from . import objects
from .solver import GpuBatch, GpuSolver, HostMesh

# Spring stiffness of the tether between each Lagrangian node and its
# reference position (force = -k dA (X - X_ref)).
STIFFNESS = 5.0
# Floor of the non-equilibrium weight tau through the face streaming dt cap
# (src/unstructured_mesh.cpp:979-983, whose comment gives 0.55 as the
# conservative stability bound).  Without it these sphere flows diverge once
# the cell Reynolds number exceeds ~5.
FNEQ_MIN = 0.55


@dataclass
class Draw:
    re: float
    diameter: float
    centre: tuple
    nodes: int = 0

    @property
    def n_nodes(self):
        return self.nodes or objects.sphere_nodes_for_diameter(self.diameter)


def sample_cases(n, lengths, seed, re_range=(20.0, 200.0), d_range=(6.0, 14.0), d_frac=None):
    """Seeded draws: Re ~ U[re_range]; D ~ U[d_range] in mesh units, or, with
    d_frac = (lo, hi), D ~ U[lo, hi] x channel height; centre x uniform in the
    upstream third, y and z within 2/32 of the height of the channel axis."""
    rng = np.random.default_rng(seed)
    Lx, Ly, Lz = lengths
    out = []
    for _ in range(n):
        re = float(rng.uniform(*re_range))
        d = float(rng.uniform(*d_range)) if d_frac is None else float(rng.uniform(*d_frac)) * Ly
        x0 = float(rng.uniform(0.5 * d + 3.0 * Ly / 32.0, Lx / 3.0))
        y0 = float(0.5 * Ly + rng.uniform(-2.0, 2.0) * Ly / 32.0)
        z0 = float(0.5 * Lz + rng.uniform(-2.0, 2.0) * Lz / 32.0)
        out.append(Draw(re, d, (x0, y0, z0)))
    return out


def case_for(draw, precision=1, tol=1e-6, check=100, fneq_min=FNEQ_MIN):
    c = pc.channel_case(re=draw.re, length_scale=draw.diameter, fneq_min=fneq_min)
    c.precision = precision
    c.tolerance = tol
    c.check_interval = check
    return c


def ibm_args(draw, case, stiffness=None, h=1.0):
    X, dA = objects.fibonacci_sphere(draw.n_nodes, draw.centre, 0.5 * draw.diameter)
    return dict(x_ref=X, x=X, stiffness=STIFFNESS if stiffness is None else stiffness, node_area=dA,
                dt=case.time_marching_step, h=h, kernel=1)


def mesh_for(pm, case):
    """Host mesh whose face streaming dt carries this case's cap (a single-case
    solver takes the face dt from its mesh; a batch caps it per case itself)."""
    return HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)


def uniform_state(case, geo, n):
    return case.initial_state(geo["cx"][:n], geo["cy"][:n], geo["cz"][:n])


# --------------------------------------------------------------------------
# meshes

def channel_mesh(name):
    """(polyMesh, lengths) of the named channel mesh:
      cfg4    64 x 32 x 32 hex, interior vertices jittered 15% (seed 12), permuted (seed 13)
      tet32   32 x 16 x 16 hexes split into 6 tets each (49,152 cells) over 64 x 32 x 32,
              jitter 10% (seed 14), permuted (seed 15)
      chan96  96 x 48 x 48 hex, jitter 15% (seed 2), permuted (seed 3)
      tet88   88^3 hexes split into 6 tets (4,088,832 cells) over 176 x 88 x 88,
              jitter 10% (seed 4), permuted (seed 5)"""
    if name == "cfg4":
        L = (64.0, 32.0, 32.0)
        pm = meshgen.box_mesh(64, 32, 32, lengths=L, jitter=0.15, jitter_seed=12, permute_seed=13,
                              patch_types=pc.CHANNEL_PATCHES)
    elif name == "tet32":
        L = (64.0, 32.0, 32.0)
        pm = meshgen.box_mesh(32, 16, 16, lengths=L, tets=True, jitter=0.10, jitter_seed=14, permute_seed=15,
                              patch_types=pc.CHANNEL_PATCHES)
    elif name == "chan96":
        L = (96.0, 48.0, 48.0)
        pm = meshgen.box_mesh(96, 48, 48, lengths=L, jitter=0.15, jitter_seed=2, permute_seed=3,
                              patch_types=pc.CHANNEL_PATCHES)
    elif name == "tet88":
        L = (176.0, 88.0, 88.0)
        pm = meshgen.box_mesh(88, 88, 88, lengths=L, tets=True, jitter=0.10, jitter_seed=4, permute_seed=5,
                              patch_types=pc.CHANNEL_PATCHES)
    else:
        raise ValueError("unknown mesh %s" % name)
    return pm, L


def cfg4_mesh():
    return channel_mesh("cfg4")[0]


def lsq_exact(name):
    # the reference's LSQ determinant expression diverges on jittered tets
    return 1 if name.startswith("tet") else 0


# --------------------------------------------------------------------------
# data

def generate(pm, hm, draws, max_steps, tiling=None, precision=1, lsq_exact_det=0, stiffness=None, tol=1e-6):
    """Runs every draw in one batch to the tolerance; returns (fields, results)."""
    cases = [case_for(d, precision, tol=tol) for d in draws]
    for c in cases:
        c.lsq_exact_det = lsq_exact_det
    bt = GpuBatch(hm, cases[0].params, cases[0].patches_for(pm.patch_names), [c.visc for c in cases])
    geo = hm.geometry()
    N = pm.n_cells
    bt.set_state(-1, *uniform_state(cases[0], geo, N))
    for b, (d, c) in enumerate(zip(draws, cases)):
        bt.set_ibm(b, **ibm_args(d, c, stiffness))
    if tiling == "auto":
        bt.autotune(3)
    elif tiling:
        bt.set_tiling(*tiling)
    done = 0
    while done < max_steps:
        n = min(2500, max_steps - done)
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


def usable(result, tol=1e-3):
    """A field is a training target when its solve ended cleanly at a residual
    below tol (converged, or nearly steady when the step cap came first)."""
    return result["status"] == 0 and max(result["residual"][1:]) < tol


def dataset(geo, n, draws, fields):
    X = np.stack([mpnn.node_inputs(geo, n, d.centre, d.diameter, d.re) for d in draws])
    U = pc.Case().mach_no
    T = np.stack([mpnn.targets(*f, U) for f in fields])
    return X, T


class UnionGraph:
    """Disjoint union of the cell graphs of several meshes: node k of mesh m is
    node off[m] + k.  A training sample concatenates one field per mesh."""

    def __init__(self, parts):
        self.n = [int(n) for _, n in parts]
        self.off = np.concatenate([[0], np.cumsum(self.n)]).astype(np.int64)
        src, dst, ef = [], [], []
        for (geo, n), o in zip(parts, self.off):
            s, d, e, _ = mpnn.graph_from_geometry(geo, n)
            src.append(s + o)
            dst.append(d + o)
            ef.append(e)
        self.src = np.concatenate(src).astype(np.int32)
        self.dst = np.concatenate(dst).astype(np.int32)
        self.ef = np.concatenate(ef).astype(np.float32)
        self.N = int(self.off[-1])


def train_epochs(net, X, T, mask, epochs, lr, seed, save=""):
    """One plbfs_mpnn_train call per epoch (so that progress is printed), with a
    cosine step size from lr down to lr/20 over the epochs; the Adam moments
    persist in the network between calls.  With `save`, the parameters are
    written there after every epoch."""
    hist = []
    for ep in range(epochs):
        ph = ep / (epochs - 1) if epochs > 1 else 0.0
        lr_ep = lr / 20 + 0.5 * (lr - lr / 20) * (1 + math.cos(math.pi * ph))
        t = time.perf_counter()
        h = net.train(X, T, mask, epochs=1, lr=lr_ep, seed=seed + ep)
        hist.append(float(h[0]))
        if save:
            np.save(save, net.params())
        print("  epoch %d/%d: loss %.4g (lr %.2e, %.1f s)" % (ep + 1, epochs, hist[-1], lr_ep,
                                                            time.perf_counter() - t), file=sys.stderr, flush=True)
    return np.array(hist)


def predict(net, geo, n, draw):
    y = net.forward(mpnn.node_inputs(geo, n, draw.centre, draw.diameter, draw.re))
    return mpnn.fields(y, pc.Case().mach_no)


# --------------------------------------------------------------------------
# solves

def _solver(pm, hm, case, ibm):
    return GpuSolver(hm, case.params, case.patches_for(pm.patch_names))


def _start(s, state, ibm):
    s.set_state(*state)
    if ibm:
        s.set_ibm(**ibm)


def _run_to_tol(s, max_steps, label="", done=0):
    steps, ms = done, 0.0
    r = None
    while steps < max_steps:
        r = s.run(min(5000, max_steps - steps))
        steps += r["steps"]
        ms += r["ms"]
        if label:
            print("  %s: %d steps, residual %.3e" % (label, steps, max(r["residual"][1:])), file=sys.stderr,
                  flush=True)
        if r["converged"] or r["rc"] != 0:
            break
    return dict(rc=r["rc"], steps=steps, converged=r["converged"], ms=ms, residual=max(r["residual"][1:]))


def solve_to_tolerance(pm, hm, case, state, ibm, max_steps, label=""):
    """Plain solve from `state` on host mesh hm (built for this case, see
    mesh_for); returns (result, final state).  result["seconds"] is the host
    wall time of the solve loop."""
    s = _solver(pm, hm, case, ibm)
    t = time.perf_counter()
    _start(s, state, ibm)
    out = _run_to_tol(s, max_steps, label)
    out["seconds"] = time.perf_counter() - t
    st = s.state()
    s.close()
    return out, st


def warm_solve(pm, hm, case, ibm, net, draw, max_steps, probe=200, label=""):
    """Warm-started solve on host mesh hm (built for this case, see mesh_for):
    network inference, the probe of both starts, then the solve from the better one.  The wall time covers all three; the
    network's graph set-up belongs to mesh preprocessing and is not in it.
    Returns (result, final state, info)."""
    geo = hm.geometry()
    n = pm.n_cells
    s = _solver(pm, hm, case, ibm)
    t0 = time.perf_counter()
    pred = predict(net, geo, n, draw)
    t_inf = time.perf_counter() - t0
    uni = uniform_state(case, geo, n)
    probe_ms = 0.0
    r_uni = r_pred = float("nan")
    use_pred = True
    if probe > 0:
        _start(s, uni, ibm)
        ru = s.run(probe)
        r_uni = max(ru["residual"][1:])
        probe_ms += ru["ms"]
        _start(s, pred, ibm)
        rp = s.run(probe)
        r_pred = max(rp["residual"][1:])
        probe_ms += rp["ms"]
        use_pred = rp["rc"] == 0 and math.isfinite(r_pred) and r_pred < r_uni
        if use_pred and rp["converged"]:
            out = dict(rc=0, steps=rp["steps"], converged=True, ms=0.0, residual=r_pred)
        else:
            if not use_pred:
                _start(s, uni, ibm)
            out = _run_to_tol(s, max_steps, label, done=probe if use_pred else 0)
    else:
        _start(s, pred, ibm)
        out = _run_to_tol(s, max_steps, label)
    out["ms"] += probe_ms
    out["seconds"] = time.perf_counter() - t0
    st = s.state()
    s.close()
    info = dict(used="prediction" if use_pred else "uniform", r_pred=r_pred, r_uniform=r_uni, probe=probe,
                inference_s=t_inf)
    return out, st, info


def load_params(path=None):
    """Trained parameters shipped with the package (or from `path`)."""
    path = path or os.path.join(os.path.dirname(__file__), "data", "mpnn_channel.npy")
    if not os.path.exists(path):
        return None
    return np.load(path, allow_pickle=False)


def network_for(hm, n, params):
    src, dst, ef, _ = mpnn.graph_from_geometry(hm.geometry(), n)
    net = mpnn.Mpnn(n, src, dst, ef, train=False)
    net.set_params(params)
    return net


# --------------------------------------------------------------------------
# experiment driver

def train_main(args):
    parts, Xs, Ts, conv = [], [], [], {}
    t0 = time.perf_counter()
    for name, count in ((m, int(c)) for m, c in (s.split(":") for s in args.meshes.split(","))):
        pm, L = channel_mesh(name)
        hm = HostMesh.from_polymesh(pm, 0.01, 1.0, 0.0)
        geo = hm.geometry()
        N = pm.n_cells
        draws = sample_cases(count, L, seed=args.seed + len(parts), d_frac=(args.d_lo, args.d_hi))
        fields, keep = [], []
        for k in range(0, count, args.batch):
            f, res = generate(pm, hm, draws[k:k + args.batch], args.gen_steps, tiling="auto",
                              lsq_exact_det=lsq_exact(name), stiffness=args.stiffness)
            for d, fld, r in zip(draws[k:k + args.batch], f, res):
                if usable(r):
                    fields.append(fld)
                    keep.append(d)
            print("%s: %d usable fields of %d, steps %d..%d (%.1f s)" % (
                name, len(fields), min(k + args.batch, count), min(r["steps"] for r in res),
                max(r["steps"] for r in res), time.perf_counter() - t0), file=sys.stderr, flush=True)
        X, T = dataset(geo, N, keep, fields)
        conv[name] = dict(drawn=count, usable=len(fields))
        parts.append((geo, N))
        Xs.append(X)
        Ts.append(T)
        hm.close()
    g = UnionGraph(parts)
    S = max(len(x) for x in Xs)
    # one sample = one field of every mesh; shorter lists repeat
    X = np.concatenate([x[np.arange(S) % len(x)] for x in Xs], axis=1)
    T = np.concatenate([t[np.arange(S) % len(t)] for t in Ts], axis=1)
    del Xs, Ts
    net = mpnn.Mpnn(g.N, g.src, g.dst, g.ef, train=True)
    net.set_params(mpnn.init_params(args.seed))
    t1 = time.perf_counter()
    hist = train_epochs(net, X, T, mpnn.output_mask(g.N), args.epochs, args.lr, args.seed, save=args.save)
    t_train = time.perf_counter() - t1
    print("trained %d epochs on %d samples in %.1f s, loss %.4g -> %.4g" % (
        args.epochs, S, t_train, hist[0], hist[-1]), file=sys.stderr, flush=True)
    params = net.params()
    net.close()
    if args.save:
        np.save(args.save, params)
    return dict(meshes=conv, samples=S, epochs=args.epochs, train_seconds=t_train,
                loss_hist=[float(v) for v in hist], params=int(params.size)), params


def eval_main(args, params):
    pm, L = channel_mesh(args.eval)
    hm = HostMesh.from_polymesh(pm, 0.01, 1.0, 0.0)
    N = pm.n_cells
    net = network_for(hm, N, params)
    geo = hm.geometry()
    draws = sample_cases(args.test, L, seed=args.seed + 1000, d_frac=(args.d_lo, args.d_hi))
    rows = []
    for d in draws:
        c32 = case_for(d, precision=1)
        c64 = case_for(d, precision=0)
        for c in (c32, c64):
            c.lsq_exact_det = lsq_exact(args.eval)
        ibm = ibm_args(d, c32, args.stiffness)
        hc = mesh_for(pm, c32)
        r_w, s_w, info = warm_solve(pm, hc, c32, ibm, net, d, args.max_steps, probe=args.probe, label="warm32")
        uni = uniform_state(c32, geo, N)
        r_u32, _ = solve_to_tolerance(pm, hc, c32, uni, ibm, args.max_steps, label="rest32")
        r_u64, s_64 = solve_to_tolerance(pm, hc, c64, uni, ibm, args.max_steps, label="rest64")
        hc.close()
        uw = np.stack(s_w[1:], 1)
        u64 = np.stack(s_64[1:], 1)
        rows.append(dict(re=d.re, diameter=d.diameter, centre=d.centre, start=info["used"], r_pred=info["r_pred"],
                         r_uniform=info["r_uniform"], warm_fp32=r_w, rest_fp32=r_u32, rest_fp64=r_u64,
                         vel_rel_l2_vs_fp64=float(np.linalg.norm(uw - u64) / np.linalg.norm(u64))))
        print("%s Re %.1f D %.2f: %s start (probe %.2e vs %.2e); steps warm32 %d rest32 %d rest64 %d; "
              "s %.2f / %.2f / %.2f" % (args.eval, d.re, d.diameter, info["used"], info["r_pred"], info["r_uniform"],
                                        r_w["steps"], r_u32["steps"], r_u64["steps"], r_w["seconds"],
                                        r_u32["seconds"], r_u64["seconds"]), file=sys.stderr, flush=True)
    net.close()
    ok = [r for r in rows if r["warm_fp32"]["converged"] and r["rest_fp64"]["converged"]]
    speedup = [r["rest_fp64"]["seconds"] / r["warm_fp32"]["seconds"] for r in ok]
    err = [r["vel_rel_l2_vs_fp64"] for r in ok]
    return dict(mesh=args.eval, cells=N, tests=rows, converged=len(ok),
                speedup_fp64_rest_over_fp32_warm=speedup,
                speedup_median=float(np.median(speedup)) if ok else None,
                speedup_worst=float(np.min(speedup)) if ok else None,
                vel_rel_l2_worst=float(np.max(err)) if ok else None,
                steps_ratio_rest_over_warm_fp32=[r["rest_fp32"]["steps"] / max(r["warm_fp32"]["steps"], 1)
                                                 for r in ok if r["rest_fp32"]["converged"]])


def main(argv=None):
    """Train on converged batch fields (--meshes cfg4:384,tet32:256) and save the
    parameters, and/or compare solves from the prediction and from the uniform
    start on held-out draws of one mesh (--eval); prints one JSON report."""
    import argparse
    import json
    ap = argparse.ArgumentParser()
    ap.add_argument("--meshes", default="", help="training meshes and draw counts, e.g. cfg4:384,tet32:256")
    ap.add_argument("--eval", default="", help="mesh of the held-out comparison (cfg4, tet32, chan96, tet88)")
    ap.add_argument("--test", type=int, default=8)
    ap.add_argument("--batch", type=int, default=128)
    ap.add_argument("--max-steps", type=int, default=60000, help="step cap of the held-out solves")
    ap.add_argument("--gen-steps", type=int, default=20000, help="step cap of the training-field batches")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=7)
    ap.add_argument("--d-lo", type=float, default=0.14, help="smallest sphere diameter / channel height")
    ap.add_argument("--d-hi", type=float, default=0.44, help="largest sphere diameter / channel height")
    ap.add_argument("--stiffness", type=float, default=STIFFNESS)
    ap.add_argument("--probe", type=int, default=200)
    ap.add_argument("--save", default="")
    ap.add_argument("--params", default="")
    args = ap.parse_args(argv)
    report = {}
    params = None
    if args.meshes:
        report["train"], params = train_main(args)
    if args.eval:
        params = params if params is not None else load_params(args.params or None)
        if params is None:
            raise SystemExit("no trained parameters")
        report["eval"] = eval_main(args, params)
    print(json.dumps(report))
    return report


if __name__ == "__main__":
    main()
