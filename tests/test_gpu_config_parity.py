"""FP64 parity at the workloads' own sizes (GPU through the C-ABI, oracle on the host).

  32^3 cavity      the uniform 32 x 32 x 32 hex lid-driven cavity (32,768 cells),
                   Re 100, lid u = 0.05, from rest: fields after 100 steps within
                   1e-10 relative of the oracle, residual history alike.
  96x48x48 channel the jittered (15 %, seed 2), permuted (seed 3) 221,184-cell
                   hex channel of the sphere workload with the immersed boundary
                   off, Re 40 on D = 12, fneq_min 0.55, from a smooth seeded
                   perturbation of the inflow state (the uniform inflow itself
                   is steady): fields after 100 steps within 1e-10.
  converged state  a jittered, permuted 10 x 5 x 5 channel run by both sides to
                   max(u, v, w) residual 1e-6: the same stop step and fields
                   within 1e-8.
The oracle takes ~17 s for the cavity's 100 steps and ~4 min for the channel's
on one host core, hence the per-test time limits.
"""
import numpy as np
import pytest

from helpers import channel_mesh, initial_state, oracle_for, perturbed_state, rel_err
from plbfs import case as pc
from plbfs import meshgen
from plbfs.solver import GpuSolver, HostMesh

pytestmark = pytest.mark.gpu


def run_both(pm, case, st, steps):
    o = oracle_for(pm, case)
    o.set_state(*st)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    s.set_state(*st)
    rg = s.run(steps)
    ro = o.run(steps)
    so, sg = o.state(), s.state()
    o.close()
    return ro, rg, so, sg


@pytest.mark.timeout(300)
def test_cavity_32_cubed_100_steps(gpu_available):
    pm = meshgen.box_mesh(32, 32, 32, patch_types=pc.CAVITY_PATCHES)
    assert (pm.n_cells, pm.n_faces) == (32768, 101376)
    case = pc.cavity_case(re=100.0)
    o = oracle_for(pm, case)
    st = initial_state(case, o)
    o.close()
    ro, rg, so, sg = run_both(pm, case, st, 100)
    assert ro["rc"] == 0 and rg["rc"] == 0 and rg["steps"] == 100
    errs = [rel_err(sg[q], so[q]) for q in range(4)]
    print("32^3 cavity, 100 steps: field errors vs oracle", errs)
    assert max(errs) < 1e-10, errs
    assert np.max(np.abs(rg["hist"][:, 1:] - ro["hist"][:, 1:]) / ro["hist"][:, 1:]) < 1e-8


@pytest.mark.timeout(900)
def test_channel_96x48x48_plain_100_steps(gpu_available):
    pm = meshgen.box_mesh(96, 48, 48, lengths=(96.0, 48.0, 48.0), jitter=0.15, jitter_seed=2, permute_seed=3,
                          patch_types=pc.CHANNEL_PATCHES)
    assert pm.n_cells == 221184
    case = pc.channel_case(re=40.0, length_scale=12.0, fneq_min=0.55)
    o = oracle_for(pm, case)
    st = perturbed_state(case, o, amp=0.01, seed=17)
    o.close()
    ro, rg, so, sg = run_both(pm, case, st, 100)
    assert ro["rc"] == 0 and rg["rc"] == 0 and rg["steps"] == 100
    errs = [rel_err(sg[q], so[q]) for q in range(4)]
    print("96x48x48 channel, 100 steps: field errors vs oracle", errs)
    assert max(errs) < 1e-10, errs
    # the perturbation has moved: the comparison is not of two unchanged states
    N = pm.n_cells
    assert max(rel_err(sg[q], st[q][:N]) for q in range(4)) > 1e-4


@pytest.mark.timeout(300)
def test_converged_state_matches_oracle(gpu_available):
    pm = channel_mesh(10, 5, 5)
    case = pc.channel_case(re=40.0, length_scale=4.0)
    case.tolerance = 1e-6
    case.check_interval = 50
    o = oracle_for(pm, case)
    st = perturbed_state(case, o, amp=0.01, seed=3)
    o.close()
    ro, rg, so, sg = run_both(pm, case, st, 40000)
    assert ro["converged"] and rg["converged"], (ro["steps"], rg["steps"])
    assert rg["steps"] == ro["steps"], (rg["steps"], ro["steps"])
    errs = [rel_err(sg[q], so[q]) for q in range(4)]
    print("10x5x5 channel converged at step %d: field errors vs oracle %s" % (rg["steps"], errs))
    assert max(errs) < 1e-8, errs
