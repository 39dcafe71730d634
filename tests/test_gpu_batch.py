"""Batched cases (plbfs_batch_*) against the oracle and the single-case solver.

Each case of a batch runs the same per-cell arithmetic as a single-case
context, so its fields must equal the single-case GPU run bit for bit, for
every tiling (cells per flux block x cases per workgroup); in FP64 they must
also match the oracle run with that case's viscosity to 1e-10 after the same
steps.  Converged cases freeze while the others continue.
"""
import numpy as np
import pytest

from helpers import channel_mesh, initial_state, oracle_for, perturbed_state, rel_err
from plbfs import case as pc
from plbfs import objects
from plbfs.solver import GpuBatch, GpuSolver, HostMesh

pytestmark = pytest.mark.gpu

RES = [30.0, 60.0, 120.0]


def cases_for(res, precision=0, tol=1e-30, check=10, fneq_min=0.0):
    out = []
    for re in res:
        c = pc.channel_case(re=re, length_scale=4.0, fneq_min=fneq_min)
        c.precision = precision
        c.tolerance = tol
        c.check_interval = check
        out.append(c)
    return out


def make_batch(pm, cases):
    hm = HostMesh.from_polymesh(pm, cases[0].visc, cases[0].gamma, cases[0].fneq_min)
    bt = GpuBatch(hm, cases[0].params, cases[0].patches_for(pm.patch_names), [c.visc for c in cases])
    return hm, bt


def single_run(pm, case, st, steps, ibm=None):
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    s = GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    s.set_state(*st)
    if ibm:
        s.set_ibm(**ibm)
    r = s.run(steps)
    return r, s.state()


@pytest.mark.parametrize("precision,fneq_min", [(0, 0.0), (1, 0.0), (0, 0.55), (1, 0.55)])
def test_batch_cases_equal_single_runs(gpu_available, precision, fneq_min):
    pm = channel_mesh(10, 5, 5)
    cases = cases_for(RES, precision, fneq_min=fneq_min)
    o = oracle_for(pm, cases[0])
    st = perturbed_state(cases[0], o, amp=0.01, seed=2)
    hm, bt = make_batch(pm, cases)
    bt.set_state(-1, *st)
    rc, res = bt.run(40)
    assert rc == 0 and all(r["steps"] == 40 for r in res), res
    for b, c in enumerate(cases):
        r1, s1 = single_run(pm, c, st, 40)
        sb = bt.state(b)
        for q in range(4):
            assert np.array_equal(sb[q], s1[q]), (b, q, rel_err(sb[q], s1[q]))
        assert np.allclose(res[b]["residual"], r1["residual"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("fneq_min", [0.0, 0.55])
def test_batch_fp64_matches_oracle_per_case(gpu_available, fneq_min):
    # with fneq_min > 0 each case's face streaming dt carries its own cap
    # 3 nu / (gamma fneq_min), which the batch applies on the device
    pm = channel_mesh(10, 5, 5)
    cases = cases_for(RES, fneq_min=fneq_min)
    o0 = oracle_for(pm, cases[0])
    st = perturbed_state(cases[0], o0, amp=0.01, seed=6)
    hm, bt = make_batch(pm, cases)
    bt.set_state(-1, *st)
    assert bt.run(100)[0] == 0
    for b, c in enumerate(cases):
        o = oracle_for(pm, c)
        o.set_state(*st)
        assert o.run(100)["rc"] == 0
        so, sb = o.state(), bt.state(b)
        errs = [rel_err(sb[q], so[q]) for q in range(4)]
        assert max(errs) < 1e-10, (b, errs)


@pytest.mark.parametrize("tiling", [(64, 4, 0), (128, 2, 1), (256, 1, 1), (512, 1, 0), (64, 8, 1), (128, 4, 1)])
def test_batch_tiling_does_not_change_fields(gpu_available, tiling):
    pm = channel_mesh(12, 6, 6, tets=True)
    res = [25.0, 50.0, 75.0, 100.0, 150.0, 200.0, 40.0, 80.0]
    cases = cases_for(res, precision=1)
    for c in cases:
        c.lsq_exact_det = 1
    o = oracle_for(pm, cases[0])
    st = perturbed_state(cases[0], o, amp=0.01, seed=8)
    outs = []
    for t in [(256, 1, 0), tiling]:
        hm, bt = make_batch(pm, cases)
        bt.set_tiling(*t)
        info = bt.layout_info()
        assert (info["block_cells"], info["cases_per_block"], info["stage"]) == t
        bt.set_state(-1, *st)
        assert bt.run(25)[0] == 0
        outs.append([bt.state(b) for b in range(len(cases))])
        bt.close()
    for b in range(len(cases)):
        for q in range(4):
            assert np.array_equal(outs[0][b][q], outs[1][b][q]), (b, q)


def test_converged_case_freezes(gpu_available):
    # the same Re twice, one case started from rest: different convergence steps
    pm = channel_mesh(12, 6, 6)
    cases = cases_for([40.0, 40.0], tol=1e-6, check=50)
    o = oracle_for(pm, cases[0])
    st0 = initial_state(cases[0], o)
    st1 = perturbed_state(cases[0], o, amp=0.02, seed=1)
    hm, bt = make_batch(pm, cases)
    bt.set_state(0, *st0)
    bt.set_state(1, *st1)
    rc, res = bt.run(30000)
    assert rc == 0 and all(r["converged"] for r in res), res
    assert res[0]["steps"] != res[1]["steps"]
    for b, st in enumerate((st0, st1)):
        r1, s1 = single_run(pm, cases[b], st, 30000)
        assert r1["converged"] and r1["steps"] == res[b]["steps"], (b, r1["steps"], res[b]["steps"])
        for q in range(4):
            assert np.array_equal(bt.state(b)[q], s1[q]), (b, q)


def test_batch_per_case_spheres_equal_single_runs(gpu_available):
    pm = channel_mesh(16, 8, 8)
    cases = cases_for([40.0, 80.0, 20.0])
    spheres = [((5.0, 4.0, 4.0), 1.5, 60), ((6.0, 4.2, 3.8), 2.0, 110), ((5.5, 3.9, 4.1), 1.2, 45)]
    o = oracle_for(pm, cases[0])
    st = initial_state(cases[0], o)
    hm, bt = make_batch(pm, cases)
    bt.set_state(-1, *st)
    ibms = []
    for b, (ctr, rad, n) in enumerate(spheres):
        X, dA = objects.fibonacci_sphere(n, ctr, rad)
        ibm = dict(x_ref=X, x=X, stiffness=2.0, node_area=dA, dt=cases[b].time_marching_step, h=1.0, kernel=1)
        bt.set_ibm(b, **ibm)
        ibms.append(ibm)
    rc, res = bt.run(30)
    assert rc == 0, res
    for b, c in enumerate(cases):
        r1, s1 = single_run(pm, c, st, 30, ibm=ibms[b])
        assert r1["rc"] == 0
        for q in range(4):
            assert np.array_equal(bt.state(b)[q], s1[q]), (b, q, rel_err(bt.state(b)[q], s1[q]))


@pytest.mark.parametrize("precision,tiling", [(1, (64, 4, 0)), (1, (64, 8, 0)), (1, (128, 2, 1)), (0, (64, 2, 0))])
def test_batch_border_pass_does_not_change_fields(gpu_available, monkeypatch, precision, tiling):
    # PLBFS_BATCH_SPLIT_BORDER=1: faces between two flux blocks are evaluated
    # once per case by kb_flux_border instead of once by each block; the same
    # face arithmetic, so every case's fields stay bit for bit the same
    pm = channel_mesh(12, 6, 6, tets=True)
    res = [25.0, 50.0, 75.0, 100.0, 150.0, 200.0, 40.0, 80.0]
    cases = cases_for(res, precision=precision, fneq_min=0.55)
    for c in cases:
        c.lsq_exact_det = 1
    o = oracle_for(pm, cases[0])
    st = perturbed_state(cases[0], o, amp=0.01, seed=9)
    outs = []
    for split in ("0", "1"):
        monkeypatch.setenv("PLBFS_BATCH_SPLIT_BORDER", split)
        hm, bt = make_batch(pm, cases)
        bt.set_tiling(*tiling)
        bt.set_state(-1, *st)
        rc, r = bt.run(25)
        assert rc == 0 and all(x["steps"] == 25 for x in r), r
        outs.append([bt.state(b) for b in range(len(cases))])
        bt.close()
    for b in range(len(cases)):
        for q in range(4):
            assert np.array_equal(outs[0][b][q], outs[1][b][q]), (b, q, rel_err(outs[1][b][q], outs[0][b][q]))


def test_autotune_keeps_results(gpu_available):
    pm = channel_mesh(12, 6, 6)
    cases = cases_for([30.0, 60.0, 90.0, 120.0], precision=1)
    o = oracle_for(pm, cases[0])
    st = perturbed_state(cases[0], o, amp=0.01, seed=5)
    hm, bt = make_batch(pm, cases)
    bt.set_state(-1, *st)
    bc, cp, sg, ms = bt.autotune(3)
    assert bc in (64, 128, 256, 512) and cp in (1, 2, 4) and sg in (0, 1) and ms > 0
    info = bt.layout_info()
    assert (info["block_cells"], info["cases_per_block"], info["stage"]) == (bc, cp, sg)
    assert bt.run(20)[0] == 0
    ref = GpuBatch(hm, cases[0].params, cases[0].patches_for(pm.patch_names), [c.visc for c in cases])
    ref.set_state(-1, *st)
    assert ref.run(20)[0] == 0
    for b in range(4):
        for q in range(4):
            assert np.array_equal(bt.state(b)[q], ref.state(b)[q]), (b, q)


def test_set_state_of_one_case_keeps_the_others(gpu_available):
    # re-seeding case 1 must leave case 0's state, step count and RK history alone
    pm = channel_mesh(10, 5, 5)
    cases = cases_for(RES)
    o = oracle_for(pm, cases[0])
    st = perturbed_state(cases[0], o, amp=0.01, seed=3)
    st1 = perturbed_state(cases[0], o, amp=0.02, seed=4)
    hm, bt = make_batch(pm, cases)
    bt.set_state(-1, *st)
    assert bt.run(10)[0] == 0
    before = bt.state(0)
    bt.set_state(1, *st1)
    after = bt.state(0)
    for q in range(4):
        assert np.array_equal(before[q], after[q]), q
    rc, res = bt.run(10)
    assert rc == 0 and [r["steps"] for r in res] == [20, 10, 20], res
    for b, (start, n) in enumerate([(st, 20), (st1, 10), (st, 20)]):
        r1, s1 = single_run(pm, cases[b], start, n)
        for q in range(4):
            assert np.array_equal(bt.state(b)[q], s1[q]), (b, q)


def test_refused_tiling_leaves_the_batch_runnable(gpu_available):
    # 1024 cells x 8 FP64 cases cannot fit the LDS: the call fails and the
    # batch keeps its tiling and gives the same fields as an untouched batch
    pm = channel_mesh(12, 6, 6)
    cases = cases_for(RES)
    o = oracle_for(pm, cases[0])
    st = perturbed_state(cases[0], o, amp=0.01, seed=9)
    hm, bt = make_batch(pm, cases)
    bt.set_tiling(128, 1, 0)
    info = bt.layout_info()
    with pytest.raises(RuntimeError):
        bt.set_tiling(1024, 8, 0)
    assert bt.layout_info() == info
    bt.set_state(-1, *st)
    assert bt.run(15)[0] == 0
    hm2, ref = make_batch(pm, cases)
    ref.set_tiling(128, 1, 0)
    ref.set_state(-1, *st)
    assert ref.run(15)[0] == 0
    for b in range(len(cases)):
        for q in range(4):
            assert np.array_equal(bt.state(b)[q], ref.state(b)[q]), (b, q)


def test_node_outside_its_support_stops_only_its_case(gpu_available):
    # a node 2.2 h outside the inlet: candidate cells exist (within support +
    # margin) but every Peskin weight is zero, so its case stops with
    # PLBFS_ERR_IBM (not NaN) and the other cases run on
    pm = channel_mesh(16, 8, 8)
    cases = cases_for([40.0, 80.0, 20.0])
    o = oracle_for(pm, cases[0])
    st = initial_state(cases[0], o)
    hm, bt = make_batch(pm, cases)
    bt.set_state(-1, *st)
    X, dA = objects.fibonacci_sphere(60, (6.0, 4.0, 4.0), 1.5)
    ibm = dict(x_ref=X, x=X, stiffness=2.0, node_area=dA, dt=cases[0].time_marching_step, h=1.0, kernel=1)
    Xs = np.vstack([X, [[-2.2, 4.0, 4.0]]])
    bt.set_ibm(0, **ibm)
    bt.set_ibm(1, **dict(ibm, x_ref=Xs, x=Xs))
    rc, res = bt.run(10)
    assert rc == -7 and res[1]["status"] == -7 and res[1]["steps"] == 1, res
    assert res[0]["status"] == 0 and res[2]["status"] == 0 and res[0]["steps"] == 10 == res[2]["steps"], res
    assert all(np.all(np.isfinite(a)) for a in bt.state(1))
    r0, s0 = single_run(pm, cases[0], st, 10, ibm=ibm)
    r2, s2 = single_run(pm, cases[2], st, 10)
    for q in range(4):
        assert np.array_equal(bt.state(0)[q], s0[q]), q
        assert np.array_equal(bt.state(2)[q], s2[q]), q
