"""Case description: the parameter derivation of the reference's XML
preprocessing, restated for programmatic cases.

Follows src/preprocessor.cpp:54-84 (velocities scaled by mach_no/scale),
:187-302 (boundary types by token), src/global_variables.cpp:20-56
(nu = mach_no*ref_length/Re, preconditioned gamma), src/Solver.cpp:213
(body force = -pressure_gradient) and src/Solution.cpp:94-189 (initial
conditions as a vector magnitude for rho and a component sum for u).
The XML reader in csrc/host/case.cpp applies the same rules to files.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

BC_CODES = {"dirichlet": 1, "neumann": 2, "periodic": 3, "parabolic-N": 4, "parabolic-W": 5,
            "inlet": 6, "wall": 7, "symmetry-x": 8, "parabolic": 9}
GRAD_CODES = {"LS": 0, "hybrid": 1}          # anything else: Green-Gauss (2)
DT_CODES = {"local": 0, "constant": 1, "min": 2, "talts": 3}


@dataclass
class Case:
    mach_no: float = 0.05
    scale: float = 1.0
    reynolds_no: float = 100.0
    ref_length: float = 1.0
    pre_condition_gamma: float = 1.0
    preconditioning: bool = False
    time_marching_step: float = 0.5
    fneq_min: float = 0.0
    tolerance: float = 1e-30
    gradient_calc: str = "LS"
    time_stepping: str = "local"
    pressure_gradient: float = 0.0
    average_rho: float = 1.0
    rho_origin_magnitude: tuple = (1.0, 0.0, 0.0)
    rho_gradient: tuple = (0.0, 0.0, 0.0)
    origin_location: tuple = (0.0, 0.0, 0.0)
    vel_origin_magnitude: tuple = (0.0, 0.0, 0.0)
    vel_gradient: tuple = (0.0, 0.0, 0.0)
    # patch type token -> dict(rho, u, v, w, vel_type, rho_type) with unscaled velocity
    bcs: dict = field(default_factory=dict)
    precision: int = 0
    check_interval: int = 10
    lsq_exact_det: int = 0
    # uns_boundary_conditions/periodic_translation: one vector for every periodic
    # patch (src/preprocessor.cpp:299-301); a bc dict may override it per patch
    periodic_translation: tuple = (0.0, 0.0, 0.0)
    # geometry x, y, z: the domain lengths that the parabolic inlet profiles
    # (types 4 and 5) scale by, after multiplication by `scale`
    # (src/preprocessor.cpp:48,467-469, src/Solution.cpp:804-813)
    domain: tuple = (1.0, 1.0, 1.0)

    @property
    def factor(self):
        return self.mach_no / self.scale

    @property
    def gamma(self):
        g = self.pre_condition_gamma
        if self.preconditioning and abs(g - 1) < 0.00001:
            g = self.mach_no * 2.0 * math.sqrt(3)
        return g

    @property
    def visc(self):
        return self.mach_no * self.ref_length / self.reynolds_no

    @property
    def params(self):
        return dict(visc=self.visc, gamma=self.gamma, fneq_min=self.fneq_min, cfl=self.time_marching_step,
                    max_velocity=self.mach_no, ref_rho=self.average_rho, force=-self.pressure_gradient,
                    tolerance=self.tolerance, gradient_type=GRAD_CODES.get(self.gradient_calc, 2),
                    time_stepping=DT_CODES[self.time_stepping], precision=self.precision,
                    check_interval=self.check_interval, lsq_exact_det=self.lsq_exact_det,
                    domain_x=self.domain[0] * self.scale, domain_y=self.domain[1] * self.scale)

    def patches_for(self, patch_names):
        """Patch table in the mesh's patch order, matched by type token."""
        out = []
        f = self.factor
        for name in patch_names:
            if name not in self.bcs:
                raise KeyError("no boundary condition for patch type %r" % name)
            b = self.bcs[name]
            out.append(dict(name=name, rho_type=BC_CODES[b["rho_type"]], vel_type=BC_CODES[b["vel_type"]],
                            rho=float(b.get("rho", 1.0)), u=float(b.get("u", 0.0)) * f,
                            v=float(b.get("v", 0.0)) * f, w=float(b.get("w", 0.0)) * f,
                            is_empty=int(name == "empty"),
                            per_translate=tuple(float(t) for t in b.get("per_translate", self.periodic_translation))))
        return out

    def initial_state(self, cx, cy, cz):
        """rho,u,v,w over the given centroids (all cells incl. ghosts)."""
        f = self.factor
        ox, oy, oz = self.origin_location
        dx, dy, dz = cx - ox, cy - oy, cz - oz
        rm, rg = self.rho_origin_magnitude, self.rho_gradient
        tx = rm[0] + rg[0] * dx
        ty = rm[1] + rg[1] * dy
        tz = rm[2] + rg[2] * dz
        rho = np.sqrt(tx ** 2 + ty ** 2 + tz ** 2)
        vm = [v * f for v in self.vel_origin_magnitude]
        vg = [v * f for v in self.vel_gradient]
        ux = vm[0] + vg[0] * dx
        uy = vm[1] + vg[1] * dy
        uz = vm[2] + vg[2] * dz
        u = ux + uy + uz
        return rho, u, np.zeros_like(rho), np.zeros_like(rho)


class BoundCase:
    """A Case bound to a mesh's patch order (what the oracle and solver consume)."""

    def __init__(self, case: Case, mesh):
        self.case = case
        self.params = case.params
        self.patches = case.patches_for(mesh.patch_names)


def cavity_case(re=100.0, lid_u=1.0, mach=0.05):
    """Lid-driven cavity: patch type 'lid' moves in +x, 'wall' is no-slip;
    velocity Dirichlet, density Neumann (the cavity row of the benchmark set)."""
    return Case(mach_no=mach, reynolds_no=re, ref_length=1.0,
                bcs={"lid": dict(rho=1.0, u=lid_u, v=0.0, w=0.0, vel_type="dirichlet", rho_type="neumann"),
                     "wall": dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="dirichlet", rho_type="neumann")})


CAVITY_PATCHES = {"x0": "wall", "x1": "wall", "y0": "wall", "y1": "lid", "z0": "wall", "z1": "wall"}
CHANNEL_PATCHES = {"x0": "inlet", "x1": "outlet", "y0": "farfield", "y1": "farfield", "z0": "farfield",
                   "z1": "farfield"}


def channel_case(re=40.0, length_scale=12.0, mach=0.05, sides="farfield", fneq_min=0.0):
    """Channel: uniform inflow (velocity Dirichlet, density Neumann), outlet with
    zero-gradient velocity and fixed density.  Lateral sides: ``farfield`` holds
    the free-stream velocity (Dirichlet, density Neumann); ``symmetry`` uses the
    reference's only slip type symmetry-x (mirrors v); ``neumann`` is zero
    gradient.  The reference's Neumann ghost gradient (grad * (1 - n), without
    abs) makes zero-gradient sides unstable on these channels, so far-field is
    the default.  ``length_scale`` is the Reynolds length (sphere diameter).
    ``fneq_min`` > 0 caps every face's streaming dt at 3 nu / (gamma fneq_min),
    which keeps the non-equilibrium weight tau >= fneq_min
    (src/unstructured_mesh.cpp:979-983); flows past an immersed sphere at cell
    Reynolds numbers above ~5 diverge without it."""
    side = {"farfield": dict(rho=1.0, u=1.0, v=0.0, w=0.0, vel_type="dirichlet", rho_type="neumann"),
            "symmetry": dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="symmetry-x", rho_type="symmetry-x"),
            "neumann": dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="neumann", rho_type="neumann")}[sides]
    return Case(mach_no=mach, reynolds_no=re, ref_length=length_scale, fneq_min=fneq_min,
                vel_origin_magnitude=(1.0, 0.0, 0.0),
                bcs={"inlet": dict(rho=1.0, u=1.0, v=0.0, w=0.0, vel_type="dirichlet", rho_type="neumann"),
                     "outlet": dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="neumann", rho_type="dirichlet"),
                     "farfield": side})
