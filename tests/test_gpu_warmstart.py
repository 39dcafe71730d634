"""What the trained warm-start network gains over the method without it.

A small network is trained on the GPU (plbfs_mpnn_train) on converged fields
that the batched solver produces for seeded (Re, sphere) draws on a small
jittered channel.  For a held-out draw, the plain FP32 solver started from the
network's prediction must reach the residual tolerance (1e-5 here, to keep the
test short) in fewer steps than from the uniform start, and land on the same
steady state.  An untrained
network's prediction must lose the probe, so that the solve falls back to the
uniform start and gives exactly the plain solve's fields.
"""
import numpy as np
import pytest

from plbfs import meshgen, mpnn
from plbfs import case as pc
from plbfs import warmstart as ws
from plbfs.solver import HostMesh

pytestmark = pytest.mark.gpu

L = (32.0, 16.0, 16.0)
TOL = 1e-5


@pytest.fixture(scope="module")
def channel():
    pm = meshgen.box_mesh(32, 16, 16, lengths=L, jitter=0.10, jitter_seed=21, permute_seed=22,
                          patch_types=pc.CHANNEL_PATCHES)
    hm = HostMesh.from_polymesh(pm, 0.01, 1.0, 0.0)
    return pm, hm


def draws(n, seed):
    return ws.sample_cases(n, L, seed=seed, re_range=(40.0, 120.0), d_range=(3.0, 5.0))


@pytest.fixture(scope="module")
def trained(channel, gpu_available):
    pm, hm = channel
    train = draws(24, seed=31)
    fields, res = ws.generate(pm, hm, train, 30000, tol=TOL)
    keep = [k for k, r in enumerate(res) if ws.usable(r)]
    assert len(keep) >= 20, [(r["status"], r["residual"]) for r in res]
    geo = hm.geometry()
    X, T = ws.dataset(geo, pm.n_cells, [train[k] for k in keep], [fields[k] for k in keep])
    src, dst, ef, _ = mpnn.graph_from_geometry(geo, pm.n_cells)
    net = mpnn.Mpnn(pm.n_cells, src, dst, ef, train=True)
    net.set_params(mpnn.init_params(3))
    hist = net.train(X, T, mpnn.output_mask(pm.n_cells), epochs=60, lr=2e-3, seed=5)
    params = net.params()
    net.close()
    assert hist[-1] < 0.2 * hist[0], hist
    return params


def held_out():
    d = draws(1, seed=99)[0]
    return ws.Draw(80.0, 4.0, d.centre)


def test_warm_start_cuts_steps(channel, trained):
    pm, hm = channel
    d = held_out()
    case = ws.case_for(d, precision=1, tol=TOL)
    ibm = ws.ibm_args(d, case)
    geo = hm.geometry()
    hc = ws.mesh_for(pm, case)
    rest, s_rest = ws.solve_to_tolerance(pm, hc, case, ws.uniform_state(case, geo, pm.n_cells), ibm, 80000)
    net = ws.network_for(hm, pm.n_cells, trained)
    warm, s_warm, info = ws.warm_solve(pm, hc, case, ibm, net, d, 80000, probe=100)
    net.close()
    assert rest["converged"] and warm["converged"], (rest, warm)
    assert info["used"] == "prediction", info
    print("steps from rest %d, from the prediction %d" % (rest["steps"], warm["steps"]))
    assert warm["steps"] < 0.8 * rest["steps"], (warm["steps"], rest["steps"])
    u_r = np.stack(s_rest[1:], 1)
    u_w = np.stack(s_warm[1:], 1)
    # both stop at a residual of 1e-5 on a slowly decaying tail, so the two
    # fields agree to a few 1e-3, not to round-off
    assert np.linalg.norm(u_w - u_r) / np.linalg.norm(u_r) < 1e-2


def test_untrained_network_falls_back(channel, gpu_available):
    pm, hm = channel
    d = held_out()
    case = ws.case_for(d, precision=1, tol=TOL)
    ibm = ws.ibm_args(d, case)
    geo = hm.geometry()
    hc = ws.mesh_for(pm, case)
    rest, s_rest = ws.solve_to_tolerance(pm, hc, case, ws.uniform_state(case, geo, pm.n_cells), ibm, 80000)
    net = ws.network_for(hm, pm.n_cells, mpnn.init_params(11) * 4.0)
    warm, s_warm, info = ws.warm_solve(pm, hc, case, ibm, net, d, 80000, probe=100)
    net.close()
    assert info["used"] == "uniform", info
    assert warm["steps"] == rest["steps"]
    for q in range(4):
        assert np.array_equal(s_warm[q], s_rest[q]), q
