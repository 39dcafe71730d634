"""Shared builders for tests: meshes, cases, oracle and GPU solver pairs."""
import numpy as np

from plbfs import case as pc
from plbfs import meshgen

import pyoracle


def cavity_mesh(n=6, tets=False, jitter=0.15, permute=True, seed=7):
    return meshgen.box_mesh(n, n, n, tets=tets, jitter=jitter, jitter_seed=seed,
                            permute_seed=seed + 1 if permute else None, patch_types=pc.CAVITY_PATCHES)


def channel_mesh(nx=12, ny=6, nz=6, tets=False, jitter=0.15, permute=True, seed=11, h=1.0):
    return meshgen.box_mesh(nx, ny, nz, lengths=(nx * h, ny * h, nz * h), tets=tets, jitter=jitter,
                            jitter_seed=seed, permute_seed=seed + 1 if permute else None,
                            patch_types=pc.CHANNEL_PATCHES)


def oracle_for(pm, case):
    return pyoracle.Oracle(pm, pc.BoundCase(case, pm))


def initial_state(case, oracle):
    cells = oracle.cells()
    return case.initial_state(cells["cx"], cells["cy"], cells["cz"])


def perturbed_state(case, oracle, amp=0.01, seed=0):
    """Initial state plus a smooth, seeded perturbation of every field."""
    rho, u, v, w = initial_state(case, oracle)
    cells = oracle.cells()
    rng = np.random.default_rng(seed)
    k = rng.uniform(1.0, 3.0, size=(4, 3))
    ph = rng.uniform(0, 2 * np.pi, size=4)
    x = np.stack([cells["cx"], cells["cy"], cells["cz"]], 1)
    L = np.maximum(x.max(0) - x.min(0), 1e-12)
    waves = [np.sin((x / L) @ k[q] + ph[q]) for q in range(4)]
    U = case.mach_no
    return (rho + amp * waves[0], u + amp * U * waves[1], v + amp * U * waves[2], w + amp * U * waves[3])


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


PERIODIC_Z_PATCHES = {"x0": "inlet", "x1": "outlet", "y0": "farfield", "y1": "farfield", "z0": "zlow", "z1": "zhigh"}


def periodic_z_case(lz, re=40.0, length_scale=4.0):
    """Channel whose z sides are periodic: the z0 ghosts take the cells
    next to z1 (translate +lz) and vice versa."""
    case = pc.channel_case(re=re, length_scale=length_scale)
    per = dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="periodic", rho_type="periodic")
    case.bcs["zlow"] = dict(per, per_translate=(0.0, 0.0, lz))
    case.bcs["zhigh"] = dict(per, per_translate=(0.0, 0.0, -lz))
    return case


def periodic_z_mesh(nx=10, ny=5, nz=6, jitter=0.15, seed=21, tets=False, permute=True):
    return meshgen.box_mesh(nx, ny, nz, lengths=(float(nx), float(ny), float(nz)), tets=tets, jitter=jitter,
                            jitter_seed=seed, permute_seed=seed + 1 if permute else None,
                            patch_types=PERIODIC_Z_PATCHES)
