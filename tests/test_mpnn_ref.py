"""CPU pins of the warm-start network reference (tests/mpnn_ref.py).

The float64 network's chain-rule backward equals central finite differences
(step 1e-6) on a small random graph, entry by entry on sampled weights of
every MLP; the bf16 rounding helper is exact on representable values and
rounds half to even.
"""
import numpy as np
import pytest

import mpnn_ref as M


def small_problem(seed=0, n=7, e=18):
    rng = np.random.default_rng(seed)
    src = rng.integers(0, n, e)
    dst = (src + rng.integers(1, n, e)) % n
    G = M.Graph(src, dst, n)
    x = np.zeros((n, M.D_IN))
    x[:, :5] = rng.normal(size=(n, 5))
    ef = np.zeros((e, M.D_EDGE))
    ef[:, :6] = rng.normal(size=(e, 6))[G.order]
    tgt = np.zeros((n, M.D_OUT))
    tgt[:, :4] = rng.normal(size=(n, 4))
    mask = np.zeros((n, M.D_OUT))
    mask[:, :4] = 1
    return G, x, ef, tgt, mask


def test_parameter_count_is_about_150k():
    assert 140_000 < M.n_params() < 175_000


def test_bf16_rounding():
    vals = np.array([1.0, -2.5, 3.0e-3, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -9])
    r = M.bf16(vals)
    assert r[0] == 1.0 and r[1] == -2.5
    assert r[3] == 1.0                                 # half-way, even mantissa stays
    assert r[4] == 1.0 + 2.0 ** -6                     # half-way, odd mantissa rounds up
    assert r[5] == 1.0
    assert abs(r[2] - 3.0e-3) / 3.0e-3 < 2 ** -8


@pytest.mark.parametrize("name", ["enc", "msg0", "upd2", "msg5", "dec"])
def test_backward_matches_finite_differences(name):
    G, x, ef, tgt, mask = small_problem()
    P = M.init_params(3)
    loss, _, grads = M.loss_and_grads(P, G, x, ef, tgt, mask, bf16_mode=False)
    rng = np.random.default_rng(1)
    for key in ("W1", "b1", "W2", "b2"):
        W = P[name][key]
        for _ in range(4):
            idx = tuple(rng.integers(0, s) for s in W.shape)
            old = W[idx]
            W[idx] = old + 1e-6
            lp = M.loss_and_grads(P, G, x, ef, tgt, mask, bf16_mode=False)[0]
            W[idx] = old - 1e-6
            lm = M.loss_and_grads(P, G, x, ef, tgt, mask, bf16_mode=False)[0]
            W[idx] = old
            fd = (lp - lm) / 2e-6
            an = grads[name][key][idx]
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (name, key, idx, fd, an)
