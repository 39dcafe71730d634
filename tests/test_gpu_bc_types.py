"""GPU parity for the boundary types the other parity tests do not reach:
symmetry-x (type 8) sides, and an inlet that holds velocity (type 6) and
density (types 6 and 7).  FP64: the bars of the other FP64 parity tests
(1e-11 one-stage, 1e-10 after 100 steps, 60 on tets).  FP32 storage: 2e-5
one-stage residual, 1e-5 velocity and 1e-4 density deviation after the same
steps.  Batched: each case of a two-viscosity batch equals its single-case
run bit for bit."""
import numpy as np
import pytest

from helpers import oracle_for, perturbed_state, rel_err
from plbfs import case as pc
from plbfs import meshgen
from plbfs.solver import GpuBatch, GpuSolver, HostMesh

pytestmark = pytest.mark.gpu

SYM_SIDES = {"x0": "inlet", "x1": "outlet", "y0": "sym", "y1": "sym", "z0": "farfield", "z1": "farfield"}


def bc_case(kind):
    case = pc.channel_case(re=40.0, length_scale=4.0)
    if kind == "symmetry":
        case.bcs["sym"] = dict(rho=1.0, u=0.5, v=0.0, w=0.0, vel_type="symmetry-x", rho_type="symmetry-x")
    elif kind == "inlet6":
        case.bcs["inlet"] = dict(rho=1.0, u=1.0, v=0.0, w=0.0, vel_type="inlet", rho_type="inlet")
    else:
        case.bcs["inlet"] = dict(rho=1.0, u=1.0, v=0.0, w=0.0, vel_type="inlet", rho_type="wall")
    return case


def bc_mesh(kind, tets):
    pt = SYM_SIDES if kind == "symmetry" else pc.CHANNEL_PATCHES
    return meshgen.box_mesh(10, 5, 5, lengths=(10.0, 5.0, 5.0), tets=tets, jitter=0.15, jitter_seed=11,
                            permute_seed=12, patch_types=pt)


@pytest.mark.parametrize("kind", ["symmetry", "inlet6", "inlet7"])
@pytest.mark.parametrize("tets", [False, True])
def test_boundary_types_match_oracle(gpu_available, kind, tets):
    pm = bc_mesh(kind, tets)
    case = bc_case(kind)
    case.lsq_exact_det = 1 if tets else 0
    o = oracle_for(pm, case)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    st = perturbed_state(case, o, amp=0.01, seed=17)
    print("case", kind, "tets", tets, "cells", pm.n_cells, flush=True)
    Ro, go = o.eval_residual(*st)
    s.set_state(*st)
    R, g = s.eval_residual()
    N = pm.n_cells
    for q in range(12):
        assert rel_err(g[:, q], go[:N, q]) < 1e-11, ("grad", q)
    for q in range(4):
        assert rel_err(R[:, q], Ro[:, q]) < 1e-11, ("R", q)
    o.set_state(*st)
    s.set_state(*st)
    n = 60 if tets else 100
    assert o.run(n)["rc"] == 0
    assert s.run(n)["rc"] == 0
    errs = [rel_err(a, b) for a, b in zip(s.state(), o.state())]
    assert max(errs) < 1e-10, errs


@pytest.mark.parametrize("kind", ["symmetry", "inlet6", "inlet7"])
@pytest.mark.parametrize("tets", [False, True])
def test_boundary_types_fp32_match_oracle(gpu_available, kind, tets):
    pm = bc_mesh(kind, tets)
    case = bc_case(kind)
    case.lsq_exact_det = 1
    case.precision = 1
    o = oracle_for(pm, case)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    st = perturbed_state(case, o, amp=0.01, seed=17)
    print("case", kind, "tets", tets, "cells", pm.n_cells, flush=True)
    Ro, _ = o.eval_residual(*st)
    s.set_state(*st)
    R, _ = s.eval_residual()
    for q in range(4):
        assert rel_err(R[:, q], Ro[:, q]) < 2e-5, ("R", q)
    o.set_state(*st)
    s.set_state(*st)
    n = 60 if tets else 100
    assert o.run(n)["rc"] == 0
    assert s.run(n)["rc"] == 0
    so, sg = o.state(), s.state()
    assert rel_err(sg[0] - 1.0, so[0] - 1.0) < 1e-4
    for q in (1, 2, 3):
        assert rel_err(sg[q], so[q]) < 1e-5, q


@pytest.mark.parametrize("kind", ["symmetry", "inlet7"])
@pytest.mark.parametrize("precision", [0, 1])
def test_batch_boundary_types_equal_single_runs(gpu_available, kind, precision):
    pm = bc_mesh(kind, False)
    cases = []
    for re in (30.0, 90.0):
        c = bc_case(kind)
        c.reynolds_no = re
        c.precision = precision
        cases.append(c)
    assert cases[0].visc != cases[1].visc
    o = oracle_for(pm, cases[0])
    st = perturbed_state(cases[0], o, amp=0.01, seed=4)
    print("case", kind, "precision", precision, flush=True)
    hm = HostMesh.from_polymesh(pm, cases[0].visc, cases[0].gamma, cases[0].fneq_min)
    bt = GpuBatch(hm, cases[0].params, cases[0].patches_for(pm.patch_names), [c.visc for c in cases])
    bt.set_state(-1, *st)
    rc, res = bt.run(40)
    assert rc == 0 and all(r["steps"] == 40 for r in res), res
    for b, c in enumerate(cases):
        h1 = HostMesh.from_polymesh(pm, c.visc, c.gamma, c.fneq_min)
        s = GpuSolver(h1, c.params, c.patches_for(pm.patch_names))
        s.set_state(*st)
        assert s.run(40)["rc"] == 0
        s1, sb = s.state(), bt.state(b)
        for q in range(4):
            assert np.array_equal(sb[q], s1[q]), (b, q, rel_err(sb[q], s1[q]))


def empty_case_and_mesh(kind, tets):
    # "thin": one cell thick in x with empty x sides, which hold boundary face 0
    # (computed, Solver.cpp:308 skips only face > n_internal); "slab": a
    # channel one cell thick in z with empty z sides
    wall = dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="dirichlet", rho_type="neumann")
    empty = dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="neumann", rho_type="neumann")
    if kind == "thin":
        pt = {"x0": "empty", "x1": "empty", "y0": "wall", "y1": "wall", "z0": "wall", "z1": "wall"}
        pm = meshgen.box_mesh(1, 6, 5, lengths=(1.0, 6.0, 5.0), tets=tets, jitter=0.15, jitter_seed=3,
                              permute_seed=4, patch_types=pt)
        case = pc.Case(mach_no=0.05, reynolds_no=40.0, ref_length=4.0, bcs={"empty": empty, "wall": wall})
    else:
        pt = {"x0": "inlet", "x1": "outlet", "y0": "wall", "y1": "wall", "z0": "empty", "z1": "empty"}
        pm = meshgen.box_mesh(12, 6, 1, lengths=(12.0, 6.0, 1.0), tets=tets, jitter=0.15, jitter_seed=5,
                              permute_seed=6, patch_types=pt)
        case = pc.channel_case(re=40.0, length_scale=4.0)
        case.bcs["wall"] = wall
        case.bcs["empty"] = empty
    case.lsq_exact_det = 1 if tets else 0
    return case, pm


@pytest.mark.parametrize("kind", ["thin", "slab"])
@pytest.mark.parametrize("tets", [False, True])
def test_empty_patches_match_oracle(gpu_available, kind, tets):
    case, pm = empty_case_and_mesh(kind, tets)
    o = oracle_for(pm, case)
    assert pm.patch_names[pm.bface_patch[0]] == ("empty" if kind == "thin" else "inlet")
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    st = perturbed_state(case, o, amp=0.01, seed=23)
    print("empty", kind, "tets", tets, "cells", pm.n_cells, flush=True)
    Ro, go = o.eval_residual(*st)
    s.set_state(*st)
    R, g = s.eval_residual()
    N = pm.n_cells
    for q in range(12):
        assert rel_err(g[:, q], go[:N, q]) < 1e-11, ("grad", q)
    for q in range(4):
        assert rel_err(R[:, q], Ro[:, q]) < 1e-11, ("R", q)
    o.set_state(*st)
    # the one-cell-thick meshes are violent (the thin box's unbalanced face 0
    # drives rho to 2.5; on tets both reach rho blow-up in the oracle within 30
    # steps): 20 steps there, 100 on the hex slab
    n = 100 if (kind == "slab" and not tets) else 20
    assert o.run(n)["rc"] == 0
    assert s.run(n)["rc"] == 0
    errs = [rel_err(a, b) for a, b in zip(s.state(), o.state())]
    assert max(errs) < 1e-10, errs
    # FP32 storage on the same case
    case.precision = 1
    s32 = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    s32.set_state(*st)
    R32, _ = s32.eval_residual()
    for q in range(4):
        assert rel_err(R32[:, q], Ro[:, q]) < 2e-5, ("R32", q)
