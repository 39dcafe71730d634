"""Pins of the oracle's immersed-boundary restatement (CPU only).

- 2-point hat on the reference's own setting (uniform, unpermuted unit hex grid,
  cell index i + j nx + k nx ny): interpolation and spreading equal the
  reference's structured formulas (src/immersed_boundary_method.cu:13-90,
  :299-375), restated here from that index arithmetic, not from the oracle.
- Both kernels reproduce linear velocity fields exactly at interior nodes (the
  first-moment condition of the hat and of Peskin's 4-point function).
- Spreading conserves the total force: sum_c f_c V_c = sum_n F_n.
- Interpolation and spreading are adjoint: sum_n F_n . U_n = sum_c V_c f_c . u_c
  (same node positions).
- Node update: X = X0 + U dt, F = -k dA (X - X_ref)  (:161-190, :244-295).
"""
import numpy as np
import pytest

from helpers import oracle_for
from plbfs import case as pc
from plbfs import meshgen

import pyoracle


def unit_grid(n=(10, 9, 8), jitter=0.0, permute=None, tets=False):
    return meshgen.box_mesh(*n, lengths=tuple(float(v) for v in n), tets=tets, jitter=jitter, jitter_seed=3,
                            permute_seed=permute, patch_types=pc.CHANNEL_PATCHES)


def ref_hat_interp(X, u, nx, ny):
    """The reference's structured 2-point interpolation, origin 0, delta_x = 1."""
    xr, yr, zr = X
    out = np.zeros(3)
    x0, y0, z0 = int(round(xr)) - 1, int(round(abs(yr))) - 1, int(round(abs(zr))) - 1
    for i in (x0, x0 + 1):
        for j in (y0, y0 + 1):
            for k in (z0, z0 + 1):
                idx = i + j * nx + k * ny * nx
                wgt = (1 - abs(i + 0.5 - xr)) * (1 - abs(j + 0.5 - yr)) * (1 - abs(k + 0.5 - zr))
                out += u[idx] * wgt
    return out


def nodes(rng, n, lo, hi):
    return rng.uniform(lo, hi, size=(n, 3))


def test_hat_matches_reference_structured_formula():
    pm = unit_grid()
    o = oracle_for(pm, pc.channel_case())
    rng = np.random.default_rng(0)
    X = nodes(rng, 20, 2.0, 6.0)
    T = o.T
    c = o.cells()
    uf = [np.sin(c["cx"]) * np.cos(c["cy"]), c["cz"] * 0.1, np.cos(c["cx"] + c["cz"])]
    pyoracle.set_ibm(o, X, X, stiffness=2.0, node_area=0.5, dt=0.0, h=1.0, kernel=0)
    pyoracle.ibm_step(o, *uf)
    x, U, F = pyoracle.ibm_state(o)
    for k in range(len(X)):
        ref = ref_hat_interp(X[k], np.stack(uf, 1), 10, 9)
        assert np.allclose(U[k], ref, rtol=0, atol=1e-14)
    # reference spreading: f_c = sum_n F_n w / dx^3
    f = pyoracle.body_force(o)
    fr = np.zeros((o.N, 3))
    for k in range(len(X)):
        xr, yr, zr = x[k]
        x0, y0, z0 = int(round(xr)) - 1, int(round(yr)) - 1, int(round(zr)) - 1
        for i in (x0, x0 + 1):
            for j in (y0, y0 + 1):
                for kk in (z0, z0 + 1):
                    idx = i + j * 10 + kk * 90
                    wgt = (1 - abs(i + 0.5 - xr)) * (1 - abs(j + 0.5 - yr)) * (1 - abs(kk + 0.5 - zr))
                    fr[idx] += F[k] * wgt
    assert np.allclose(f, fr, rtol=0, atol=1e-14)
    # forces: F = -k dA (X - X_ref), X = X0 + U dt with dt = 0
    assert np.allclose(F, 0.0)


@pytest.mark.parametrize("kernel", [0, 1])
@pytest.mark.parametrize("jitter", [0.0, 0.15])
def test_linear_field_reproduced(kernel, jitter):
    pm = unit_grid(n=(12, 12, 12), jitter=jitter, permute=5)
    o = oracle_for(pm, pc.channel_case())
    c = o.cells()
    a = np.array([[0.3, -0.1, 0.2], [0.05, 0.4, -0.3], [-0.2, 0.1, 0.15]])
    b = np.array([0.01, -0.02, 0.03])
    uf = [b[q] + a[q, 0] * c["cx"] + a[q, 1] * c["cy"] + a[q, 2] * c["cz"] for q in range(3)]
    rng = np.random.default_rng(1)
    X = nodes(rng, 15, 4.0, 8.0)
    pyoracle.set_ibm(o, X, X, stiffness=0.0, node_area=1.0, dt=0.0, h=1.0, kernel=kernel)
    pyoracle.ibm_step(o, *uf)
    _, U, _ = pyoracle.ibm_state(o)
    exp = b[None, :] + X @ a.T
    if jitter == 0.0:
        assert np.allclose(U, exp, atol=1e-13)
    else:
        # the per-node normalisation keeps constants exact; linear fields are
        # exact only to the jitter-induced first-moment error
        assert np.allclose(U, exp, atol=0.25 * np.abs(a).max())


@pytest.mark.parametrize("kernel", [0, 1])
def test_force_conservation_and_adjointness(kernel):
    pm = unit_grid(n=(12, 10, 10), jitter=0.12, permute=7, tets=True)
    o = oracle_for(pm, pc.channel_case())
    rng = np.random.default_rng(2)
    Xr = nodes(rng, 30, 3.5, 6.5)
    X = Xr + rng.uniform(-0.1, 0.1, size=Xr.shape)
    c = o.cells()
    uf = [np.sin(0.3 * c["cx"]), np.cos(0.2 * c["cy"]), np.sin(0.1 * c["cz"] + 1)]
    pyoracle.set_ibm(o, Xr, X, stiffness=3.0, node_area=0.7, dt=0.0, h=1.0, kernel=kernel)
    pyoracle.ibm_step(o, *uf)
    x, U, F = pyoracle.ibm_state(o)
    assert np.allclose(F, -3.0 * 0.7 * (X - Xr), atol=1e-15)
    f = pyoracle.body_force(o)
    vol = c["vol"][: o.N]
    assert np.allclose((f * vol[:, None]).sum(0), F.sum(0), rtol=1e-12, atol=1e-15)
    # adjointness at fixed positions (dt = 0: interpolation and spreading use the same X)
    lhs = np.sum(F * U)
    u3 = np.stack(uf, 1)[: o.N]
    rhs = np.sum(vol[:, None] * f * u3)
    assert abs(lhs - rhs) < 1e-12 * max(abs(lhs), 1e-30)


def test_node_update_moves_with_interpolated_velocity():
    pm = unit_grid()
    o = oracle_for(pm, pc.channel_case())
    X = np.array([[4.3, 4.6, 3.9], [5.1, 3.2, 4.4]])
    c = o.cells()
    uf = [np.full(o.T, 0.05), np.full(o.T, -0.01), np.full(o.T, 0.02)]
    pyoracle.set_ibm(o, X, X, stiffness=1.5, node_area=0.2, dt=0.5, h=1.0, kernel=1)
    pyoracle.ibm_step(o, *uf)
    x, U, F = pyoracle.ibm_state(o)
    assert np.allclose(U, [0.05, -0.01, 0.02], atol=1e-15)
    assert np.allclose(x, X + U * 0.5, atol=1e-15)
    assert np.allclose(F, -1.5 * 0.2 * (x - X), atol=1e-15)
