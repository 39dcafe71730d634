"""NumPy reference of the warm-start message-passing network (test reference).

The same network as 3d-plfbs-gpu_amd/csrc/device/mpnn.hip, written plainly:
forward, and backward by the chain rule.  With ``bf16=True`` every GEMM
operand is rounded to bfloat16 (round to nearest even) exactly where the GPU
kernels round it, and products accumulate in float64, so GPU results agree to
FP32-accumulation noise.  With ``bf16=False`` it is a float64 network whose
backward is checked against central finite differences.

Network (6 layers, hidden 64, every MLP = Linear -> SiLU -> Linear):
  h  = enc(x)                                  x: [N, 16] (5 used)
  repeat 6 times:
    m_e = msg([h[dst_e], h[src_e], f_e])       f_e: [E, 16] (6 used); K = 144
    a_t = sum_{e: dst_e = t} m_e / max(deg_t, 1)
    h   = h + upd([h, a])                      K = 128
  y  = dec(h)                                  y: [N, 16] (4 used)
"""
import numpy as np

H = 64
D_IN = 16
D_EDGE = 16
D_OUT = 16
N_LAYERS = 6
K_MSG = 2 * H + D_EDGE   # 144
K_UPD = 2 * H            # 128


def bf16(a):
    """Round float32/float64 values to bfloat16 (nearest even), returned as float64."""
    a32 = np.asarray(a, dtype=np.float32)
    u = a32.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def silu(a):
    return a / (1.0 + np.exp(-a))


def dsilu(a):
    s = 1.0 / (1.0 + np.exp(-a))
    return s * (1.0 + a * (1.0 - s))


def mlp_shapes():
    """(name, K_in, K_out) of every MLP in parameter order."""
    out = [("enc", D_IN, H)]
    for l in range(N_LAYERS):
        out.append(("msg%d" % l, K_MSG, H))
        out.append(("upd%d" % l, K_UPD, H))
    out.append(("dec", H, D_OUT))
    return out


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    P = {}
    for name, k, o in mlp_shapes():
        P[name] = dict(W1=rng.normal(0, np.sqrt(2.0 / k), (k, H)), b1=np.zeros(H),
                       W2=rng.normal(0, np.sqrt(1.0 / H), (H, o)) * (0.5 if name.startswith("upd") else 1.0),
                       b2=np.zeros(o))
    return P


def n_params():
    return sum(k * H + H + H * o + o for _, k, o in mlp_shapes())


class Graph:
    """Directed edges sorted by destination; CSR of incoming edges per node."""

    def __init__(self, src, dst, n_nodes):
        order = np.lexsort((src, dst))
        self.src = np.asarray(src)[order].astype(np.int32)
        self.dst = np.asarray(dst)[order].astype(np.int32)
        self.N = n_nodes
        self.order = order
        deg = np.bincount(self.dst, minlength=n_nodes)
        self.off = np.concatenate([[0], np.cumsum(deg)]).astype(np.int32)
        self.inv_deg = 1.0 / np.maximum(deg, 1)


def _mlp_fwd(p, z, rnd):
    a1 = rnd(z) @ rnd(p["W1"]) + p["b1"]
    s1 = silu(a1)
    y = rnd(s1) @ rnd(p["W2"]) + p["b2"]
    return y, (z, a1, s1)


def _mlp_bwd(p, cache, dy, rnd):
    z, a1, s1 = cache
    g = {}
    g["W2"] = rnd(s1).T @ rnd(dy)
    g["b2"] = dy.sum(0)
    ds1 = rnd(dy) @ rnd(p["W2"]).T
    da1 = ds1 * dsilu(a1)
    g["W1"] = rnd(z).T @ rnd(da1)
    g["b1"] = da1.sum(0)
    dz = rnd(da1) @ rnd(p["W1"]).T
    return g, dz


def forward(P, G, x, ef, bf16_mode=True):
    rnd = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    caches = {}
    h, caches["enc"] = _mlp_fwd(P["enc"], x, rnd)
    hs = [h]
    for l in range(N_LAYERS):
        z = np.concatenate([h[G.dst], h[G.src], ef], 1)
        m, caches["msg%d" % l] = _mlp_fwd(P["msg%d" % l], z, rnd)
        agg = np.zeros_like(h)
        np.add.at(agg, G.dst, m)
        agg *= G.inv_deg[:, None]
        u = np.concatenate([h, agg], 1)
        dh, caches["upd%d" % l] = _mlp_fwd(P["upd%d" % l], u, rnd)
        h = h + dh
        hs.append(h)
    y, caches["dec"] = _mlp_fwd(P["dec"], h, rnd)
    return y, caches


def backward(P, G, caches, dy, bf16_mode=True):
    rnd = bf16 if bf16_mode else (lambda a: np.asarray(a, dtype=np.float64))
    grads = {}
    grads["dec"], dh = _mlp_bwd(P["dec"], caches["dec"], dy, rnd)
    for l in reversed(range(N_LAYERS)):
        g, du = _mlp_bwd(P["upd%d" % l], caches["upd%d" % l], dh, rnd)
        grads["upd%d" % l] = g
        dh = dh + du[:, :H]
        dagg = du[:, H:] * G.inv_deg[:, None]
        dm = dagg[G.dst]
        g, dz = _mlp_bwd(P["msg%d" % l], caches["msg%d" % l], dm, rnd)
        grads["msg%d" % l] = g
        np.add.at(dh, G.dst, dz[:, :H])
        np.add.at(dh, G.src, dz[:, H:2 * H])
    grads["enc"], _ = _mlp_bwd(P["enc"], caches["enc"], dh, rnd)
    return grads


def loss_and_grads(P, G, x, ef, target, mask, bf16_mode=True):
    """Mean squared error over the used output channels of all nodes."""
    y, caches = forward(P, G, x, ef, bf16_mode)
    diff = (y - target) * mask
    n = mask.sum()
    loss = float((diff ** 2).sum() / n)
    dy = 2.0 * diff / n
    return loss, y, backward(P, G, caches, dy, bf16_mode)
