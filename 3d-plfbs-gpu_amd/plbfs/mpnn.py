"""Warm-start network front end (include/plbfs_mpnn.h) and its features.

Node inputs (16 per cell, 9 used): centroid scaled to [-1, 1] per axis, signed
distance to the immersed sphere over its diameter, normalised Reynolds number,
the offset from the sphere centre over its diameter (clipped to +-8, over 4),
and the sphere diameter over the channel height (centred on 0.3, over 0.15).
The offset tells the network which side of the sphere a cell is on (upstream
or in the wake), which the distance alone does not.
Edge inputs (16 per directed interior face, 6 used): the face's n A over h^2
and the centroid offset x_src - x_dst over h, with h the mean cell size; the
reverse edge flips both.  Outputs (16, 4 used): (rho - 1) / RHO_SCALE and
(u, v, w) / U.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as nat

D = 16
RHO_SCALE = 0.01
RE_MID, RE_HALF = 110.0, 90.0


def fptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def graph_from_geometry(geo, n_cells):
    """Directed interior-face edges (both directions) and their features."""
    own = geo["owner"]
    nb = geo["nbr"]
    NI = int(np.sum(nb < n_cells) if nb.size else 0)
    f = np.arange(len(own))
    interior = (nb < n_cells) & (f < len(own))
    o, n = own[interior], nb[interior]
    nA = geo["fn"][interior] * geo["area"][interior, None]
    cx = np.stack([geo["cx"], geo["cy"], geo["cz"]], 1)
    h = float(np.mean(geo["vol"]) ** (1.0 / 3.0))
    d_on = (cx[o] - cx[n]) / h
    src = np.concatenate([o, n]).astype(np.int32)
    dst = np.concatenate([n, o]).astype(np.int32)
    ef = np.zeros((2 * len(o), D), np.float32)
    ef[: len(o), 0:3] = nA / (h * h)
    ef[len(o):, 0:3] = -nA / (h * h)
    ef[: len(o), 3:6] = d_on          # src = owner, dst = neighbour: x_src - x_dst
    ef[len(o):, 3:6] = -d_on
    return src, dst, ef, NI


def node_inputs(geo, n_cells, centre, diameter, re):
    cx = np.stack([geo["cx"][:n_cells], geo["cy"][:n_cells], geo["cz"][:n_cells]], 1)
    lo, hi = cx.min(0), cx.max(0)
    x = np.zeros((n_cells, D), np.float32)
    x[:, 0:3] = 2 * (cx - lo) / np.maximum(hi - lo, 1e-30) - 1
    x[:, 3] = (np.linalg.norm(cx - np.asarray(centre), axis=1) - 0.5 * diameter) / diameter
    x[:, 4] = (re - RE_MID) / RE_HALF
    x[:, 5:8] = np.clip((cx - np.asarray(centre)) / diameter, -8.0, 8.0) / 4.0
    x[:, 8] = (diameter / max(hi[1] - lo[1], 1e-30) - 0.3) / 0.15
    return x


def targets(rho, u, v, w, U):
    t = np.zeros((len(rho), D), np.float32)
    t[:, 0] = (rho - 1.0) / RHO_SCALE
    t[:, 1] = u / U
    t[:, 2] = v / U
    t[:, 3] = w / U
    return t


def fields(y, U):
    return 1.0 + RHO_SCALE * y[:, 0].astype(np.float64), U * y[:, 1], U * y[:, 2], U * y[:, 3]


def output_mask(n):
    m = np.zeros((n, D), np.float32)
    m[:, :4] = 1
    return m


class Mpnn:
    def __init__(self, n_nodes, src, dst, edge_feat, train=True):
        self.N = int(n_nodes)
        self._src = np.ascontiguousarray(src, dtype=np.int32)
        self._dst = np.ascontiguousarray(dst, dtype=np.int32)
        self._ef = np.ascontiguousarray(edge_feat, dtype=np.float32).reshape(-1)
        self.P = nat.lib().plbfs_mpnn_n_params()
        h = C.c_void_p()
        nat.check(nat.lib().plbfs_mpnn_create(self.N, len(self._src), nat.iptr(self._src), nat.iptr(self._dst),
                                              fptr(self._ef), 1 if train else 0, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            nat.lib().plbfs_mpnn_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.P
        nat.check(nat.lib().plbfs_mpnn_set_params(self.h, fptr(p)))

    def params(self):
        p = np.zeros(self.P, np.float32)
        nat.check(nat.lib().plbfs_mpnn_get_params(self.h, fptr(p)))
        return p

    def forward(self, x):
        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((self.N, D), np.float32)
        nat.check(nat.lib().plbfs_mpnn_forward(self.h, fptr(x), fptr(y)))
        return y

    def loss_grad(self, x, target, mask):
        x, t, m = (np.ascontiguousarray(a, dtype=np.float32) for a in (x, target, mask))
        g = np.zeros(self.P, np.float32)
        loss = C.c_double(0)
        nat.check(nat.lib().plbfs_mpnn_loss_grad(self.h, fptr(x), fptr(t), fptr(m), C.byref(loss), fptr(g)))
        return loss.value, g

    def adam_step(self, lr, b1=0.9, b2=0.999, eps=1e-8):
        nat.check(nat.lib().plbfs_mpnn_adam_step(self.h, lr, b1, b2, eps))

    def train(self, X, T, mask, epochs, lr, seed=1):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        m = np.ascontiguousarray(mask, dtype=np.float32)
        hist = np.zeros(epochs)
        nat.check(nat.lib().plbfs_mpnn_train(self.h, X.shape[0], fptr(X), fptr(T), fptr(m), int(epochs), lr, seed,
                                             nat.dptr(hist)))
        return hist


H = 64
N_LAYERS = 6


def mlp_shapes():
    """(name, k_in, k_out) of every MLP in the library's parameter order."""
    out = [("enc", D, H)]
    for l in range(N_LAYERS):
        out.append(("msg%d" % l, 2 * H + D, H))
        out.append(("upd%d" % l, 2 * H, H))
    out.append(("dec", H, D))
    return out


def init_params(seed=0):
    """He-normal first layers, 1/sqrt(64) second layers (halved in the residual
    update MLPs), zero biases; flat FP32 vector."""
    rng = np.random.default_rng(seed)
    P = {}
    for name, k, o in mlp_shapes():
        P[name] = dict(W1=rng.normal(0, np.sqrt(2.0 / k), (k, H)), b1=np.zeros(H),
                       W2=rng.normal(0, np.sqrt(1.0 / H), (H, o)) * (0.5 if name.startswith("upd") else 1.0),
                       b2=np.zeros(o))
    return flatten_params(P, mlp_shapes())


def flatten_params(P, shapes):
    """Flat FP32 vector in the library's order from {name: dict(W1, b1, W2, b2)}."""
    out = []
    for name, _, _ in shapes:
        for k in ("W1", "b1", "W2", "b2"):
            out.append(np.asarray(P[name][k], dtype=np.float64).reshape(-1))
    return np.concatenate(out).astype(np.float32)


def unflatten(vec, shapes, H=64):
    P, p = {}, 0
    for name, k, o in shapes:
        d = {}
        for key, shp in (("W1", (k, H)), ("b1", (H,)), ("W2", (H, o)), ("b2", (o,))):
            n = int(np.prod(shp))
            d[key] = np.asarray(vec[p:p + n], dtype=np.float64).reshape(shp)
            p += n
        P[name] = d
    return P
