"""Warm-start network on the GPU (BF16 MFMA kernels) against the NumPy reference.

The reference rounds the same GEMM operands to bfloat16 and accumulates in
float64; the GPU accumulates in FP32, so a few operands round to the
neighbouring bfloat16 value downstream.  Bars (relative L2 per tensor):
outputs 1e-2, every parameter gradient 3e-2.  The reference's own backward is
pinned against finite differences on the CPU (test_mpnn_ref.py).
"""
import numpy as np
import pytest

import mpnn_ref as M
from helpers import channel_mesh
from plbfs import case as pc
from plbfs import mpnn
from plbfs.solver import HostMesh

pytestmark = pytest.mark.gpu


def problem(seed=0):
    pm = channel_mesh(8, 4, 4)
    case = pc.channel_case(re=40.0, length_scale=3.0)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    geo = hm.geometry()
    N = pm.n_cells
    src, dst, ef, _ = mpnn.graph_from_geometry(geo, N)
    x = mpnn.node_inputs(geo, N, (4.0, 2.0, 2.0), 2.0, 80.0)
    rng = np.random.default_rng(seed)
    tgt = np.zeros((N, 16), np.float32)
    tgt[:, :4] = rng.normal(size=(N, 4))
    return N, src, dst, ef, x, tgt


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(np.linalg.norm(b), 1e-30))


def test_forward_matches_reference(gpu_available):
    N, src, dst, ef, x, _ = problem()
    P = M.init_params(1)
    net = mpnn.Mpnn(N, src, dst, ef, train=False)
    net.set_params(mpnn.flatten_params(P, M.mlp_shapes()))
    y = net.forward(x)
    G = M.Graph(src, dst, N)
    yr, _ = M.forward(P, G, x.astype(np.float64), ef[G.order].astype(np.float64), bf16_mode=True)
    assert rel(y[:, :4], yr[:, :4]) < 1e-2


def test_gradients_match_reference(gpu_available):
    N, src, dst, ef, x, tgt = problem()
    P = M.init_params(2)
    net = mpnn.Mpnn(N, src, dst, ef, train=True)
    shapes = M.mlp_shapes()
    net.set_params(mpnn.flatten_params(P, shapes))
    mask = mpnn.output_mask(N)
    loss, g = net.loss_grad(x, tgt, mask)
    G = M.Graph(src, dst, N)
    lr, _, gr = M.loss_and_grads(P, G, x.astype(np.float64), ef[G.order].astype(np.float64),
                                 tgt.astype(np.float64), mask.astype(np.float64), bf16_mode=True)
    assert abs(loss - lr) / lr < 1e-2
    gg = mpnn.unflatten(g, shapes)
    for name, _, _ in shapes:
        for key in ("W1", "b1", "W2", "b2"):
            assert rel(gg[name][key], gr[name][key]) < 3e-2, (name, key)


def test_gradients_are_deterministic(gpu_available):
    N, src, dst, ef, x, tgt = problem(3)
    net = mpnn.Mpnn(N, src, dst, ef, train=True)
    net.set_params(mpnn.flatten_params(M.init_params(4), M.mlp_shapes()))
    mask = mpnn.output_mask(N)
    _, g1 = net.loss_grad(x, tgt, mask)
    _, g2 = net.loss_grad(x, tgt, mask)
    assert np.array_equal(g1, g2)


def test_training_reduces_loss(gpu_available):
    N, src, dst, ef, x, _ = problem()
    net = mpnn.Mpnn(N, src, dst, ef, train=True)
    net.set_params(mpnn.flatten_params(M.init_params(5), M.mlp_shapes()))
    # a smooth target the inputs determine
    tgt = np.zeros((N, 16), np.float32)
    tgt[:, 0] = x[:, 0] * x[:, 1]
    tgt[:, 1] = np.tanh(3 * x[:, 3])
    tgt[:, 2] = x[:, 2]
    X = np.stack([x] * 4)
    T = np.stack([tgt] * 4)
    hist = net.train(X, T, mpnn.output_mask(N), epochs=40, lr=2e-3, seed=3)
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.2 * hist[0], hist
