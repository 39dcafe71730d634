"""GPU immersed boundary (pair-list interpolation/spreading) against the oracle.

Bars: node positions, velocities and forces to 1e-11 relative, cell force
density and fields after 20 coupled steps to 1e-10 relative; force
conservation sum_c f_c V_c = sum_n F_n to 1e-12.
"""
import numpy as np
import pytest

from helpers import channel_mesh, oracle_for, perturbed_state, rel_err
from plbfs import case as pc
from plbfs import objects
from plbfs.solver import GpuSolver, HostMesh

import pyoracle

pytestmark = pytest.mark.gpu


def setup(tets, kernel, jitter=0.12, n_nodes=80, steps=20):
    pm = channel_mesh(16, 10, 10, tets=tets, jitter=jitter)
    case = pc.channel_case(re=40.0, length_scale=4.0)
    case.lsq_exact_det = 1 if tets else 0
    o = oracle_for(pm, case)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    st = perturbed_state(case, o, amp=0.01, seed=4)
    o.set_state(*st)
    s.set_state(*st)
    X, dA = objects.fibonacci_sphere(n_nodes, (6.0, 5.0, 5.0), 2.0)
    args = dict(stiffness=0.5, node_area=dA, dt=case.time_marching_step, h=1.0, kernel=kernel)
    pyoracle.set_ibm(o, X, X, **args)
    s.set_ibm(X, X, **args)
    return pm, o, s, X


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("tets", [False, True])
def test_ibm_coupled_steps_match_oracle(gpu_available, tets, kernel):
    pm, o, s, X = setup(tets, kernel)
    n = 20
    ro = o.run(n)
    rg = s.run(n)
    assert ro["rc"] == 0 and rg["rc"] == 0, (ro["rc"], rg["rc"])
    xo, uo, fo = pyoracle.ibm_state(o)
    xg, ug, fg = s.ibm_state()
    assert rel_err(xg - X, xo - X) < 1e-9
    assert rel_err(ug, uo) < 1e-9
    assert rel_err(fg, fo) < 1e-9
    fb_o = pyoracle.body_force(o)
    fb_g = s.body_force()
    assert rel_err(fb_g, fb_o) < 1e-9
    for a, b in zip(s.state(), o.state()):
        assert rel_err(a, b) < 1e-10


@pytest.mark.parametrize("kernel", [0, 1])
def test_ibm_force_conservation(gpu_available, kernel):
    pm, o, s, X = setup(True, kernel, steps=1)
    assert s.run(3)["rc"] == 0
    _, _, F = s.ibm_state()
    f = s.body_force()
    vol = s.mesh.geometry()["vol"]
    assert np.allclose((f * vol[:, None]).sum(0), F.sum(0), rtol=1e-12, atol=1e-16)


def test_ibm_rejects_nodes_without_cells(gpu_available):
    # candidate lists are built around the nodes' positions, which may lie
    # anywhere relative to their anchors; a node with no cell within its
    # kernel support plus margin is an error
    pm, o, s, X = setup(False, 1)
    from plbfs import _native as nat
    s.set_ibm(X, X + 2.0, stiffness=1.0, node_area=1.0, dt=0.5, h=1.0, kernel=1)
    with pytest.raises(nat.PlbfsError) as e:
        s.set_ibm(X, X + 50.0, stiffness=1.0, node_area=1.0, dt=0.5, h=1.0, kernel=1)
    assert e.value.rc == -7


def test_ibm_lists_follow_drifting_nodes(gpu_available):
    """Nodes that drift half their list margin (0.5 h) get new candidate lists
    built around their positions, so a sphere carried 3 h downstream runs
    without PLBFS_ERR_IBM; the run equals, bit for bit, one that stops after
    every step and calls plbfs_set_ibm afresh whenever a node has drifted that
    far (the FP64 local time step restarts from the same state bits)."""
    pm = channel_mesh(20, 8, 8, jitter=0.1)
    case = pc.channel_case(re=40.0, length_scale=3.0)
    case.check_interval = 25
    X, dA = objects.fibonacci_sphere(60, (4.0, 4.0, 4.0), 1.5)
    # a weak tether: the nodes are carried with the stream, and their small
    # spring force still reaches the fields through the spreading
    ibm = dict(stiffness=1e-4, node_area=dA, dt=case.time_marching_step, h=1.0, kernel=1)
    runs = []
    for stepwise in (False, True):
        hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
        s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
        cells = hm.geometry()
        s.set_state(*case.initial_state(cells["cx"], cells["cy"], cells["cz"]))
        s.set_ibm(X, X, **ibm)
        rebuilds = 0
        if not stepwise:
            r = s.run(200)
            assert r["rc"] == 0 and r["steps"] == 200, r
        else:
            centre = X.copy()
            for t in range(200):
                assert s.run(1)["rc"] == 0, t
                x, _, _ = s.ibm_state()
                if np.abs(x - centre).max() > 0.5:
                    s.set_ibm(X, x, **ibm)
                    centre = x.copy()
                    rebuilds += 1
        runs.append((s.state(), s.ibm_state(), rebuilds))
    (sa, ia, _), (sb, ib, nb) = runs
    assert nb >= 4, nb
    assert (ia[0] - X)[:, 0].min() > 3.0  # every node carried more than 3 h downstream
    for q in range(4):
        assert np.array_equal(sa[q], sb[q]), q
    for q in range(3):
        assert np.array_equal(ia[q], ib[q]), q


def test_sphere_drag_points_downstream(gpu_available):
    """A sphere held by springs in a uniform stream: the fluid pushes it in +x
    (minus the sum of the spring forces), and lift is small against drag."""
    from helpers import channel_mesh
    from plbfs import objects
    pm = channel_mesh(16, 8, 8, jitter=0.1)
    case = pc.channel_case(re=20.0, length_scale=3.0)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    cells = hm.geometry()
    s.set_state(*case.initial_state(cells["cx"], cells["cy"], cells["cz"]))
    X, dA = objects.fibonacci_sphere(100, (5.0, 4.0, 4.0), 1.5)
    s.set_ibm(X, X, stiffness=5.0, node_area=dA, dt=case.time_marching_step, h=1.0, kernel=1)
    assert s.run(300)["rc"] == 0
    F = s.ibm_body_force()
    _, _, fn = s.ibm_state()
    assert np.allclose(F, -fn.sum(0))
    assert F[0] > 0 and abs(F[1]) < 0.2 * F[0] and abs(F[2]) < 0.2 * F[0], F


@pytest.mark.parametrize("tets", [False, True])
def test_rigid_cylinder_coupled_steps_match_oracle(gpu_available, tets):
    # the reference's only rigid object (lagrangian_object.cpp:79-113) with its
    # own IBM parameters: 2-point hat of width delta_x = 2 dt (Solver_gpu.cu:517)
    # and node area 2 pi r / n * depth (immersed_boundary_method.cu:168)
    pm = channel_mesh(16, 10, 10, tets=tets, jitter=0.12)
    case = pc.channel_case(re=40.0, length_scale=4.0)
    case.lsq_exact_det = 1 if tets else 0
    o = oracle_for(pm, case)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    st = perturbed_state(case, o, amp=0.01, seed=6)
    o.set_state(*st)
    s.set_state(*st)
    X, dA = objects.rigid_cylinder(36, (6.0, 5.0, 5.0), 2.0, 6.0)
    assert X.shape == (36, 3)
    args = dict(stiffness=0.5, node_area=dA, dt=case.time_marching_step, h=2 * 0.5, kernel=0)
    pyoracle.set_ibm(o, X, X, **args)
    s.set_ibm(X, X, **args)
    ro, rg = o.run(20), s.run(20)
    assert ro["rc"] == 0 and rg["rc"] == 0, (ro["rc"], rg["rc"])
    xo, uo, fo = pyoracle.ibm_state(o)
    xg, ug, fg = s.ibm_state()
    assert rel_err(xg - X, xo - X) < 1e-9 and rel_err(ug, uo) < 1e-9 and rel_err(fg, fo) < 1e-9
    assert np.abs(xg - X).max() > 0  # the nodes moved
    for a, b in zip(s.state(), o.state()):
        assert rel_err(a, b) < 1e-10
