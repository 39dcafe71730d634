"""GPU parity for the boundary types whose ghost the reference does not fully
write (src/Solution.cpp:759-833): parabolic-N (4) and parabolic-W (5) set u
and v from a profile at the centroid of cell b, b the boundary index, and
leave w; "parabolic" (9) and the velocity "wall" (7) write nothing; density
types 4, 5 and 9 write nothing.  What is not written keeps the ghost's
initial state (plbfs_set_ghost_state).  Their ghost gradients take the
reference's last branch, the gradient of the periodic node, which is cell 0
on a patch that is not periodic (src/gradients.cpp:434-440,488-502,
src/Boundary_Conditions.cpp:416).  Bars as the other boundary-type tests:
FP64 1e-11 one stage and 1e-10 after 100 steps (60 on tets), FP32 2e-5 one
stage and 1e-5 velocity, 1e-4 density deviation after the same steps.
"""
import numpy as np
import pytest

from helpers import oracle_for, perturbed_state, rel_err
from plbfs import case as pc
from plbfs import meshgen
from plbfs.solver import GpuBatch, GpuSolver, HostMesh

pytestmark = pytest.mark.gpu

KINDS = {
    "parabolic-N": dict(rho=1.0, u=1.0, v=0.2, w=0.3, vel_type="parabolic-N", rho_type="parabolic"),
    "parabolic-W": dict(rho=1.0, u=1.0, v=0.2, w=0.3, vel_type="parabolic-W", rho_type="dirichlet"),
    "parabolic": dict(rho=1.0, u=1.0, v=0.0, w=0.0, vel_type="parabolic", rho_type="neumann"),
    "wall": dict(rho=1.0, u=1.0, v=0.0, w=0.0, vel_type="wall", rho_type="parabolic-N"),
}


def para_case(kind):
    case = pc.channel_case(re=40.0, length_scale=4.0)
    case.bcs["inlet"] = dict(KINDS[kind])
    case.domain = (10.5, 6.0, 5.0)  # the profile's X and Y, not the box's
    return case


def para_mesh(tets):
    return meshgen.box_mesh(10, 5, 5, lengths=(10.0, 5.0, 5.0), tets=tets, jitter=0.15, jitter_seed=11,
                            permute_seed=12, patch_types=pc.CHANNEL_PATCHES)


def ghost_start(case, o, N):
    # a perturbed state over all cells; the ghosts' part is what the unwritten
    # ghost components keep, w made nonzero so that it shows
    st = list(perturbed_state(case, o, amp=0.01, seed=17))
    st[3] = st[3] + 0.005
    return st


def make_solver(pm, case, st):
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    s.set_state(*st)
    s.set_ghost_state(*(x[pm.n_cells:] for x in st))
    return s


@pytest.mark.parametrize("kind", sorted(KINDS))
@pytest.mark.parametrize("tets", [False, True])
def test_parabolic_and_wall_types_match_oracle(gpu_available, kind, tets):
    pm = para_mesh(tets)
    case = para_case(kind)
    case.lsq_exact_det = 1 if tets else 0
    o = oracle_for(pm, case)
    N = pm.n_cells
    st = ghost_start(case, o, N)
    print("case", kind, "tets", tets, "cells", N, flush=True)
    Ro, go = o.eval_residual(*st)
    s = make_solver(pm, case, st)
    R, g = s.eval_residual()
    for q in range(12):
        assert rel_err(g[:, q], go[:N, q]) < 1e-11, ("grad", q)
    for q in range(4):
        assert rel_err(R[:, q], Ro[:, q]) < 1e-11, ("R", q)
    o.set_state(*st)
    n = 60 if tets else 100
    assert o.run(n)["rc"] == 0
    assert s.run(n)["rc"] == 0
    errs = [rel_err(a, b) for a, b in zip(s.state(), o.state())]
    assert max(errs) < 1e-10, errs


@pytest.mark.parametrize("kind", sorted(KINDS))
def test_parabolic_and_wall_types_fp32(gpu_available, kind):
    pm = para_mesh(True)
    case = para_case(kind)
    case.lsq_exact_det = 1
    case.precision = 1
    o = oracle_for(pm, case)
    st = ghost_start(case, o, pm.n_cells)
    print("case", kind, flush=True)
    Ro, _ = o.eval_residual(*st)
    s = make_solver(pm, case, st)
    R, _ = s.eval_residual()
    for q in range(4):
        assert rel_err(R[:, q], Ro[:, q]) < 2e-5, ("R", q)
    o.set_state(*st)
    assert o.run(60)["rc"] == 0
    assert s.run(60)["rc"] == 0
    so, sg = o.state(), s.state()
    assert rel_err(sg[0] - 1.0, so[0] - 1.0) < 1e-4
    for q in (1, 2, 3):
        assert rel_err(sg[q], so[q]) < 1e-5, q


def test_ghost_state_is_required_where_a_type_keeps_it(gpu_available):
    pm = para_mesh(False)
    case = para_case("parabolic")
    o = oracle_for(pm, case)
    st = ghost_start(case, o, pm.n_cells)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    s.set_state(*st)
    with pytest.raises(RuntimeError, match="plbfs_set_ghost_state"):
        s.run(5)
    # a batch has no per-case ghost state: such patches are refused there
    with pytest.raises(RuntimeError):
        GpuBatch(hm, case.params, case.patches_for(pm.patch_names), [case.visc, case.visc])
