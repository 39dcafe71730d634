"""plbfs_run_noalloc against plbfs_run (GPU, through the C-ABI).

The allocation-free time loop runs the same kernels as plbfs_run in the
library's fixed per-device scratch array, so after the same steps its fields
must equal plbfs_run's bit for bit, with the same step count, status and
residual history.  Two calls in a row share the array: the second runs a
smaller mesh (fewer flux blocks, fewer steps) after the first has left its
partials and history there.  A call that needs more scratch than the array
holds is refused and leaves the state alone.
"""
import numpy as np
import pytest

from helpers import channel_mesh, initial_state, oracle_for, perturbed_state
from plbfs import _native as nat
from plbfs import case as pc
from plbfs import objects
from plbfs.solver import GpuSolver, HostMesh

pytestmark = pytest.mark.gpu


def solver_pair(pm, case, st, ibm=None):
    out = []
    for _ in range(2):
        hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
        s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
        s.set_state(*st)
        if ibm:
            s.set_ibm(**ibm)
        out.append(s)
    return out


def assert_same_run(ra, rb, sa, sb):
    assert (rb["rc"], rb["status"], rb["steps"], rb["converged"], rb["total"]) == \
        (ra["rc"], ra["status"], ra["steps"], ra["converged"], ra["total"])
    assert rb["residual"] == ra["residual"]
    assert np.array_equal(rb["hist"], ra["hist"])
    for q, (a, b) in enumerate(zip(sa.state(), sb.state())):
        assert np.array_equal(a, b), q


def test_noalloc_run_equals_run_on_two_calls_in_a_row(gpu_available):
    # call 1: 16 x 8 x 8 jittered hex channel, FP64, immersed sphere, 60 steps
    pm = channel_mesh(16, 8, 8)
    case = pc.channel_case(re=40.0, length_scale=4.0)
    case.check_interval = 7
    st = initial_state(case, oracle_for(pm, case))
    X, dA = objects.fibonacci_sphere(80, (6.0, 4.0, 4.0), 1.5)
    ibm = dict(x_ref=X, x=X, stiffness=2.0, node_area=dA, dt=case.time_marching_step, h=1.0, kernel=1)
    a, b = solver_pair(pm, case, st, ibm)
    big_blocks = a.layout_info()["blocks"]
    ra, rb = a.run(60), b.run_noalloc(60)
    assert ra["rc"] == 0 and ra["steps"] == 60
    assert_same_run(ra, rb, a, b)

    # call 2: a smaller mesh (fewer blocks) with FP32 storage and the "min"
    # time step (its block minima also live in the scratch), 25 steps
    pm2 = channel_mesh(6, 4, 4, tets=True, seed=5)
    case2 = pc.channel_case(re=60.0, length_scale=2.0)
    case2.precision = 1
    case2.time_stepping = "min"
    case2.lsq_exact_det = 1
    o2 = oracle_for(pm2, case2)
    st2 = perturbed_state(case2, o2, amp=0.01, seed=4)
    a2, b2 = solver_pair(pm2, case2, st2)
    assert a2.layout_info()["blocks"] < big_blocks
    ra2, rb2 = a2.run(25), b2.run_noalloc(25)
    assert ra2["rc"] == 0 and ra2["steps"] == 25
    assert_same_run(ra2, rb2, a2, b2)

    # the first context continues where it stopped, in both variants
    ra3, rb3 = a.run(10), b.run_noalloc(10)
    assert ra3["total"] == 70
    assert_same_run(ra3, rb3, a, b)


def test_noalloc_run_stops_on_convergence_like_run(gpu_available):
    pm = channel_mesh(8, 4, 4)
    case = pc.channel_case(re=40.0, length_scale=2.0)
    case.tolerance = 1e-5
    case.check_interval = 25
    o = oracle_for(pm, case)
    st = perturbed_state(case, o, amp=0.01, seed=9)
    a, b = solver_pair(pm, case, st)
    ra, rb = a.run(20000), b.run_noalloc(20000)
    assert ra["converged"] and ra["steps"] < 20000
    assert_same_run(ra, rb, a, b)


def test_noalloc_run_refuses_more_than_its_scratch(gpu_available):
    pm = channel_mesh(8, 4, 4)
    case = pc.channel_case(re=40.0, length_scale=2.0)
    st = initial_state(case, oracle_for(pm, case))
    a, = solver_pair(pm, case, st)[:1]
    before = a.state()
    with pytest.raises(nat.PlbfsError) as e:
        a.run_noalloc(1 << 20)  # 32 MiB of residual history
    assert e.value.rc == -1 and "scratch" in str(e.value)
    for q, x in enumerate(a.state()):
        assert np.array_equal(x, before[q]), q
