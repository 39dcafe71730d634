"""Pins of the oracle's face flux, LSQ determinant and ghost cells that the
uniform-state, rest-state and conservation pins leave open (CPU only).

Each check runs against the oracle library the tests use, and against a copy
of the oracle built with one planted mistake (OM_PLANT in plbfs_oracle.c);
a pin counts only if it accepts the first and rejects the second.

  shear closed form   On uniform hexes a linear shear u = a y (rho uniform,
                      v = w = 0) gets exact LSQ gradients, and the D3Q15
                      moments give, through every interior face of normal n,
                        mass   rho U n_x A
                        mom_x  rho A (n_x (1/3 + U^2/gamma + nu a^2 dt / gamma^2)
                                      - nu a n_y / gamma)
                        mom_y  rho A (n_y / 3 - nu a n_x / gamma)
                        mom_z  rho A n_z (1/3 + nu a^2 dt / gamma^2)
                      with U = a y_face.  This fixes tau' = 3 nu / (dt gamma),
                      the sign of the lattice offsets and the 1/gamma of feq.
  geometric flux      On jittered hexes and tets, with quadratic fields, the
                      flux of every interior face equals the LBFS flux written
                      from its definition: lattice point x_f - e dt takes the
                      state of the side of the face plane it lies on (the
                      owner when it lies on the plane), from that side's linear
                      reconstruction; the rest point takes the mean of the two
                      face values; f = feq(I) - tau' (feq(I) - feq(lattice)).
  ghost gradients     At inlet (type 6) faces the ghost-side points carry no
                      gradient offset; at symmetry-x (type 8) faces the density
                      ghost gradient is zero and the velocity one is
                      2 (u_b - u_o) n / |x_g - x_o|.
  LSQ determinant     On hexes sheared in x by y and z (LHS_yz = 0) the
                      reference expression (gradients.cpp:282-284) reduces to
                      zz (xx yy + xy^2) - xy xz - yy xz^2.
  parabolic ghosts    Types 4 and 5 (Solution.cpp:804-813) evaluate the profile
                      at the centroid of cell b, b the boundary index, whose
                      position on an unpermuted box is known in closed form;
                      type 5's u mixes x and y; w (and all of type 9) keeps the
                      initial state.
"""
import numpy as np
import pytest

from helpers import cavity_mesh, oracle_for
from plbfs import case as pc
from plbfs import meshgen

import pyoracle

# D3Q15 from its definition: rest, the 6 axis directions, the 8 body diagonals
E = np.array([[0, 0, 0]] + [list(v) for a in range(3) for v in (np.eye(3, dtype=int)[a], -np.eye(3, dtype=int)[a])]
             + [[x, y, z] for x in (1, -1) for y in (1, -1) for z in (1, -1)], dtype=float)
W = np.array([2 / 9] + [1 / 9] * 6 + [1 / 72] * 8)


def feq(rho, u, gamma):
    """Preconditioned D3Q15 equilibrium, rho [..], u [..,3] -> [..,15]."""
    eu = u @ E.T
    uu = np.sum(u * u, axis=-1)[..., None]
    return W * rho[..., None] * (1 + 3 * eu + 4.5 * eu * eu / gamma - 1.5 * uu / gamma)


def lbfs_flux(n, A, dt, xf, nu, gamma, side_state):
    """Flux [F,4] through faces from their definition.  side_state(k_mask, xp)
    returns the primitive state [F,4] at points xp [F,3] (for lattice points,
    xp = xf - e dt) taken from the owner where k_mask is True, else from the
    neighbour; the rest point is passed xp = xf and both masks."""
    F = n.shape[0]
    lat = np.zeros((F, 15, 4))
    lat[:, 0] = 0.5 * (side_state(np.ones(F, bool), xf) + side_state(np.zeros(F, bool), xf))
    for k in range(1, 15):
        xp = xf - E[k] * dt[:, None]
        owner = ((xp - xf) * n).sum(1) <= 0  # behind the face or on its plane
        lat[:, k] = side_state(owner, xp)
    f_lat = np.stack([feq(lat[:, k, 0], lat[:, k, 1:], gamma)[:, k] for k in range(15)], 1)
    rI = f_lat.sum(1)
    uI = (f_lat @ E) / rI[:, None]
    fI = feq(rI, uI, gamma)
    tau = 3 * nu / (dt * gamma)
    f = fI - tau[:, None] * (fI - f_lat)
    en = E @ n.T  # [15,F]
    mass = (f * en.T).sum(1)
    mom = np.einsum("fk,kf,ka->fa", f, en, E)
    return np.concatenate([mass[:, None], mom], 1) * A[:, None]


def _fail_with(plant, check):
    """A check must reject the oracle with planted mistake `plant`."""
    with pytest.raises(AssertionError):
        check(pyoracle.plant_path(plant))


# ---------------------------------------------------------------------------
# Linear shear on uniform hexes: closed form
# ---------------------------------------------------------------------------

SHEAR_PATCHES = {"x0": "side", "x1": "side", "y0": "bottom", "y1": "lid", "z0": "side", "z1": "side"}


def check_shear_closed_form(libpath=None, gamma=1.0):
    n_cells = 5
    pm = meshgen.box_mesh(n_cells, n_cells, n_cells, patch_types=SHEAR_PATCHES)
    a, rho0, mach = 0.05, 1.02, 0.05
    neu = dict(rho=rho0, u=0.0, v=0.0, w=0.0, vel_type="neumann", rho_type="neumann")
    case = pc.Case(mach_no=mach, reynolds_no=5.0, pre_condition_gamma=gamma,
                   bcs={"side": neu,
                        "bottom": dict(neu, vel_type="dirichlet"),
                        "lid": dict(neu, vel_type="dirichlet", u=a / mach)})
    o = pyoracle.Oracle(pm, pc.BoundCase(case, pm), libpath)
    c, f = o.cells(), o.faces()
    st = [np.full(o.T, rho0), a * c["cy"], np.zeros(o.T), np.zeros(o.T)]
    _, g = o.eval_residual(*st)
    # the ghosts mirror the field exactly, so every gradient is (0, a, 0) for u
    assert np.abs(g[: o.N, 4] - a).max() < 1e-12 and np.abs(np.delete(g[: o.N], 4, 1)).max() < 1e-12
    nu, gam = case.visc, case.gamma
    for fi in range(pm.n_internal):
        n = np.array([f["ni"][fi], f["nj"][fi], f["nk"][fi]])
        A, dt, U = f["area"][fi], f["dt"][fi], a * f["fy"][fi]
        q = nu * a * a * dt / gam ** 2
        exp = rho0 * A * np.array([U * n[0],
                                   n[0] * (1 / 3 + U * U / gam + q) - nu * a * n[1] / gam,
                                   n[1] / 3 - nu * a * n[0] / gam,
                                   n[2] * (1 / 3 + q)])
        F = o.face_flux(fi)
        # the viscous parts are ~1e-5 of the pressure part; hold them to 1e-9 of themselves
        assert np.allclose(F, exp, rtol=0, atol=1e-9 * rho0 * nu * a * A / gam), (fi, F, exp)


@pytest.mark.parametrize("gamma", [1.0, 0.3])
def test_shear_flux_closed_form(gamma):
    check_shear_closed_form(gamma=gamma)


@pytest.mark.parametrize("plant", [1, 2, 3])
def test_shear_pin_rejects_planted_mistakes(plant):
    # 1: tau' without 1/gamma (visible at gamma != 1), 2: tau' doubled, 3: offsets +e.grad dt
    _fail_with(plant, lambda lp: check_shear_closed_form(lp, gamma=0.3))


# ---------------------------------------------------------------------------
# Geometric LBFS flux on jittered meshes
# ---------------------------------------------------------------------------

def _quadratic_state(x, rng, rho0=1.0, U=0.05):
    """Smooth quadratic fields over points x [P,3] (seeded coefficients)."""
    cl = rng.uniform(-1, 1, size=(4, 3))
    cq = rng.uniform(-1, 1, size=(4, 3, 3))
    lin = np.einsum("qa,pa->pq", cl, x)
    quad = np.einsum("qab,pa,pb->pq", cq, x, x)
    amp = np.array([0.02, U, U, U])
    base = np.array([rho0, U, 0.0, 0.0])
    return base + amp * (lin + quad)


def _side_state_fn(o, pm, st, g, faces):
    """Linear reconstruction of the owner's or the neighbour's field at a point."""
    c = o.cells()
    xc = np.stack([c["cx"], c["cy"], c["cz"]], 1)
    own = pm.owner[faces]
    nbr = o.faces()["nbr"][faces]
    S = np.stack(st, 1)
    G = g.reshape(-1, 4, 3)

    def side_state(owner_mask, xp):
        cell = np.where(owner_mask, own, nbr)
        return S[cell] + np.einsum("fqa,fa->fq", G[cell], xp - xc[cell])
    return side_state


def check_geometric_flux(libpath=None, tets=False, gamma=1.0, jitter=0.15):
    pm = cavity_mesh(5, tets=tets, jitter=jitter, seed=3)
    case = pc.cavity_case(re=2.0)
    case.pre_condition_gamma = gamma
    case.lsq_exact_det = 1
    o = pyoracle.Oracle(pm, pc.BoundCase(case, pm), libpath)
    c = o.cells()
    x = np.stack([c["cx"], c["cy"], c["cz"]], 1)
    st = list(_quadratic_state(x, np.random.default_rng(5)).T)
    _, g = o.eval_residual(*st)
    f = o.faces()
    faces = np.arange(pm.n_internal)
    n = np.stack([f["ni"], f["nj"], f["nk"]], 1)
    if jitter > 0:
        # where a lattice direction lies in the face plane analytically, e.n is
        # rounding noise and so is the side it picks; those faces are left to
        # the unjittered mesh, where e.n = 0 exactly
        tie = np.abs(n[faces] @ E[1:].T).min(1) < 1e-10
        faces = faces[~tie]
        assert faces.size > 250
    n = n[faces]
    xf = np.stack([f["fx"], f["fy"], f["fz"]], 1)[faces]
    exp = lbfs_flux(n, f["area"][faces], f["dt"][faces], xf, case.visc, case.gamma,
                    _side_state_fn(o, pm, st, g, faces))
    got = np.array([o.face_flux(fi) for fi in faces])
    # the non-equilibrium part of the momentum flux is well above round-off
    # (the mass flux is rho_I u_I . n A whatever tau' is)
    visc = np.abs(exp - lbfs_flux(n, f["area"][faces], f["dt"][faces], xf, 0.0, case.gamma,
                                  _side_state_fn(o, pm, st, g, faces))).max(0)
    assert np.all(visc[1:] > 1e-2 * np.abs(exp).max(0)[1:])
    assert np.abs(got - exp).max(0).max() < 1e-12 * np.abs(exp).max(), np.abs(got - exp).max(0)


@pytest.mark.parametrize("tets", [False, True])
@pytest.mark.parametrize("gamma", [1.0, 0.3])
def test_flux_equals_its_definition_on_jittered_meshes(tets, gamma):
    check_geometric_flux(tets=tets, gamma=gamma)


def test_flux_on_face_planes_takes_the_owner_side():
    # unjittered hexes: every face contains lattice directions (e.n = 0 exactly),
    # and a bilinear field gives the two sides different in-plane gradients
    check_geometric_flux(jitter=0.0)


@pytest.mark.parametrize("plant", [1, 2, 3, 4, 5])
def test_geometric_flux_pin_rejects_planted_mistakes(plant):
    # 4: lattice side flipped, 5: rest point from the owner only (and 1-3 as above)
    _fail_with(plant, lambda lp: check_geometric_flux(lp, tets=True, gamma=0.3))


def test_face_plane_pin_rejects_planted_mistake():
    # 11: points on the face plane from the neighbour side
    _fail_with(11, lambda lp: check_geometric_flux(lp, jitter=0.0))


# ---------------------------------------------------------------------------
# Ghost gradients of inlet (6) and symmetry-x (8) faces
# ---------------------------------------------------------------------------

GHOST_PATCHES = {"x0": "inlet", "x1": "outlet", "y0": "sym", "y1": "far", "z0": "far", "z1": "far"}


def _ghost_case(inlet_type):
    far = dict(rho=1.0, u=1.0, v=0.0, w=0.0, vel_type="dirichlet", rho_type="neumann")
    return pc.Case(mach_no=0.05, reynolds_no=1.0, ref_length=4.0, lsq_exact_det=1,
                   bcs={"inlet": dict(rho=1.01, u=1.0, v=0.2, w=-0.1, vel_type=inlet_type, rho_type=inlet_type),
                        "outlet": dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="neumann", rho_type="dirichlet"),
                        "sym": dict(rho=1.0, u=0.3, v=0.0, w=0.1, vel_type="symmetry-x", rho_type="symmetry-x"),
                        "far": far})


def check_boundary_ghost_gradients(libpath=None, tets=False):
    pm = meshgen.box_mesh(6, 4, 4, lengths=(6.0, 4.0, 4.0), tets=tets, jitter=0.15, jitter_seed=8,
                          permute_seed=9, patch_types=GHOST_PATCHES)
    case = _ghost_case("inlet")
    bc = pc.BoundCase(case, pm)
    o = pyoracle.Oracle(pm, bc, libpath)
    c = o.cells()
    x = np.stack([c["cx"], c["cy"], c["cz"]], 1)
    st = list(_quadratic_state(x / 4.0, np.random.default_rng(2)).T)
    _, g = o.eval_residual(*st)
    f = o.faces()
    N, NI = pm.n_cells, pm.n_internal
    S = np.stack(st, 1)
    G = g.reshape(-1, 4, 3)
    names = [p["name"] for p in bc.patches]
    for name, ptype in (("inlet", 6), ("sym", 8)):
        P = bc.patches[names.index(name)]
        faces = NI + np.flatnonzero(pm.bface_patch == names.index(name))
        faces = faces[faces > NI]  # boundary face 0 takes the two-sided branch
        own = pm.owner[faces]
        gh = f["nbr"][faces]
        n = np.stack([f["ni"], f["nj"], f["nk"]], 1)[faces]
        xf = np.stack([f["fx"], f["fy"], f["fz"]], 1)[faces]
        Go = G[own]
        if ptype == 6:
            Gg = np.zeros_like(Go)
        else:
            d = np.linalg.norm(x[gh] - x[own], axis=1)
            ub = np.array([P["u"], P["v"], P["w"]])
            Gg = np.zeros_like(Go)
            Gg[:, 1:, :] = (2 * (ub[None, :] - S[own, 1:]) / d[:, None])[:, :, None] * n[:, None, :]
        # the face value: type 6 extrapolates the owner with the ghost gradient
        # (zero), type 8 keeps rho_o, u_o, w_o and sets v = 0
        fv = S[own].copy()
        if ptype == 8:
            fv[:, 2] = 0.0

        def side_state(owner_mask, xp):
            grad = np.where(owner_mask[:, None, None], Go, Gg)
            return fv + np.einsum("fqa,fa->fq", grad, xp - xf)
        exp = lbfs_flux(n, f["area"][faces], f["dt"][faces], xf, case.visc, case.gamma, side_state)
        got = np.array([o.face_flux(fi) for fi in faces])
        assert np.abs(got - exp).max() < 1e-12 * np.abs(exp).max(), (name, np.abs(got - exp).max())
        # and the ghost gradients the oracle used are these
        assert np.abs(G[gh] - Gg).max() < 1e-14 * max(np.abs(Gg).max(), 1.0), name


@pytest.mark.parametrize("tets", [False, True])
def test_inlet_and_symmetry_ghost_gradients(tets):
    check_boundary_ghost_gradients(tets=tets)


@pytest.mark.parametrize("plant", [7, 8])
def test_ghost_gradient_pin_rejects_planted_mistakes(plant):
    # 7: type-6 ghost gradients copy the owner's, 8: type-8 density ghost gradient copies the owner's
    _fail_with(plant, lambda lp: check_boundary_ghost_gradients(lp, tets=True))


# ---------------------------------------------------------------------------
# LSQ determinant: the "- yz - zx" part
# ---------------------------------------------------------------------------

def check_lsq_determinant_xz(libpath=None):
    pm = cavity_mesh(4, tets=False, jitter=0.0)
    pts = pm.points.astype(np.float64)
    pts[:, 0] += 0.5 * pts[:, 1] + 0.25 * pts[:, 2]  # exact in float32 on the 1/4 lattice
    pm.points = pts.astype(np.float32)
    case = pc.cavity_case()
    case.lsq_exact_det = 0
    o = pyoracle.Oracle(pm, pc.BoundCase(case, pm), libpath)
    c = o.cells()
    xc = np.stack([c["cx"], c["cy"], c["cz"]], 1)
    off, cell, _ = o.stencil()
    inv, rhs, _ = o.lsq()
    L = np.zeros((o.N, 3, 3))
    for i in range(o.N):
        for s, k in enumerate(range(off[i], off[i + 1])):
            L[i] += np.outer(rhs[i, s], xc[cell[k]] - xc[i])
    scale = np.abs(L).max()
    # interior cells (no ghost in the stencil): LHS_yz = 0, LHS_xy and LHS_xz not
    inner = [i for i in range(o.N) if np.all(cell[off[i]:off[i + 1]] < o.N)]
    assert len(inner) >= 8
    Li = L[inner]
    xx, yy, zz, xy, xz = Li[:, 0, 0], Li[:, 1, 1], Li[:, 2, 2], Li[:, 0, 1], Li[:, 0, 2]
    assert np.abs(Li[:, 1, 2]).max() < 1e-12 * scale
    assert np.abs(xy).min() > 1e-3 * scale and np.abs(xz).min() > 1e-3 * scale
    det_ref = zz * (xx * yy + xy * xy) - xy * xz - yy * xz * xz
    s = np.linalg.det(Li) / det_ref
    prod = inv.reshape(o.N, 3, 3)[inner] @ Li
    assert np.abs(prod - s[:, None, None] * np.eye(3)).max() < 1e-9


def test_lsq_reference_determinant_with_xz_coupling():
    check_lsq_determinant_xz()


def test_lsq_determinant_pin_rejects_planted_mistake():
    _fail_with(6, check_lsq_determinant_xz)


# ---------------------------------------------------------------------------
# Parabolic ghosts (types 4, 5, 9)
# ---------------------------------------------------------------------------

PARA_PATCHES = {"x0": "inlet", "x1": "outlet", "y0": "far", "y1": "far", "z0": "far", "z1": "far"}


def check_parabolic_ghosts(libpath=None, vel_type="parabolic-N"):
    nx, ny, nz, h = 6, 4, 3, 0.5
    pm = meshgen.box_mesh(nx, ny, nz, lengths=(nx * h, ny * h, nz * h), patch_types=PARA_PATCHES)
    X, Y = 2.5, 1.75  # domain lengths (geometry x, y), deliberately not the box's
    ub, vb = 0.8, 0.3
    case = pc.Case(mach_no=0.05, reynolds_no=10.0, domain=(X, Y, 1.0),
                   vel_origin_magnitude=(1.0, 0.0, 0.0), rho_gradient=(0.01, 0.02, 0.0),
                   bcs={"inlet": dict(rho=1.0, u=ub, v=vb, w=0.4, vel_type=vel_type, rho_type="parabolic"),
                        "outlet": dict(rho=1.0, vel_type="neumann", rho_type="dirichlet"),
                        "far": dict(rho=1.0, u=1.0, vel_type="dirichlet", rho_type="neumann")})
    bc = pc.BoundCase(case, pm)
    o = pyoracle.Oracle(pm, bc, libpath)
    c = o.cells()
    st = list(case.initial_state(c["cx"], c["cy"], c["cz"]))
    st[3] = st[3] + 0.01  # a nonzero initial w, which the profiles leave in place
    o.eval_residual(*st)
    rho_t, u_t, v_t, w_t = o.temp()
    N, NI = pm.n_cells, pm.n_internal
    fac = case.factor
    b = np.flatnonzero(pm.bface_patch == 0)  # the inlet's boundary faces
    # cell b of an unpermuted box: b = ix + nx (iy + ny iz), centre ((i + 1/2) h)
    ix, iy = b % nx, (b // nx) % ny
    cxb, cyb = (ix + 0.5) * h, (iy + 0.5) * h
    if vel_type == "parabolic-N":
        eu = 4 * ub * fac / Y ** 2 * cyb * (Y - cyb)
        ev = 4 * vb * fac / Y ** 2 * cyb * (Y - cyb)
    elif vel_type == "parabolic-W":
        eu = 4 * ub * fac / X ** 2 * cxb * (X - cyb)
        ev = 4 * vb * fac / X ** 2 * cxb * (X - cxb)
    else:  # "parabolic": no velocity update at all
        eu, ev = st[1][N + b], st[2][N + b]
    assert np.abs(u_t[N + b] - eu).max() < 1e-15 and np.abs(v_t[N + b] - ev).max() < 1e-15
    # w and the type-9 density keep the initial ghost state
    assert np.array_equal(w_t[N + b], st[3][N + b]) and np.array_equal(rho_t[N + b], st[0][N + b])
    assert np.abs(eu).max() > 0.1 * ub * fac


@pytest.mark.parametrize("vel_type", ["parabolic-N", "parabolic-W", "parabolic"])
def test_parabolic_ghost_states(vel_type):
    check_parabolic_ghosts(vel_type=vel_type)


@pytest.mark.parametrize("plant,vel_type", [(9, "parabolic-N"), (9, "parabolic-W"), (10, "parabolic-W")])
def test_parabolic_pin_rejects_planted_mistakes(plant, vel_type):
    # 9: the owner's centroid instead of cell b's, 10: type 5 u with (X - x)
    _fail_with(plant, lambda lp: check_parabolic_ghosts(lp, vel_type=vel_type))
