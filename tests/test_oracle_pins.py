"""Pins of the CPU oracle against closed forms and invariants (CPU only).

The reference ships no tests, fixtures or sample meshes, and its solver cannot
be built here (it needs Boost, TecIO and the CUDA headers), so each part of the
restatement is pinned by what the mathematics fixes:
  geometry  - cube volumes/areas/normals/centroids, ghost mirror, closure,
              total volume on jittered meshes, float32 points;
  LSQ       - LHS rebuilt from the oracle's own RHS and stencil offsets,
              inv(LHS) * LHS = I with the true determinant and = s*I with the
              reference's expression (s != 1 exactly when LHS_xy != 0);
              exact gradients of linear fields;
  flux      - uniform state gives the Euler flux rho(u.n)A, rho(n/3 + u(u.n)/gamma)A;
              rest state gives R = 0; global conservation sum R_rho V = -sum F_boundary;
  RK        - one step of a near-equilibrium single cell equals the reference
              scheme's amplification 1 + Z + 7/12 Z^2 + 5/24 Z^3 + 1/24 Z^4
              (Z = dt J), distinct from classical RK4;
  forcing   - a uniform body force on a resting uniform state adds dt*f per step;
  ghosts    - the free stream is steady in the channel; Neumann ghost gradients
              keep the sign of n; symmetry-x ghosts mirror the owner and drag
              the stream toward the patch velocity; a type-7 density ghost
              continues the owner's gradient; empty patches are skipped from
              the second boundary face on, and boundary face 0 takes the
              two-sided branch; on x-by-y sheared hexes the reference LSQ
              determinant flips the sign of its xy term.
"""
import numpy as np
import pytest

from helpers import cavity_mesh, initial_state, oracle_for, rel_err
from plbfs import case as pc
from plbfs import meshgen

import pyoracle


def allneumann_case():
    return pc.Case(mach_no=0.05, reynolds_no=100.0,
                   bcs={t: dict(rho=1.0, u=0, v=0, w=0, vel_type="neumann", rho_type="neumann")
                        for t in ("wall", "lid")})


def test_cube_geometry_closed_forms():
    h = 0.25
    pm = meshgen.box_mesh(4, 4, 4, patch_types=pc.CAVITY_PATCHES)
    o = oracle_for(pm, pc.cavity_case())
    c, f = o.cells(), o.faces()
    N = pm.n_cells
    assert np.allclose(c["vol"][:N], h ** 3, rtol=1e-12, atol=0)
    assert np.allclose(f["area"], h * h, rtol=0, atol=1e-17)
    n = np.stack([f["ni"], f["nj"], f["nk"]], 1)
    assert np.allclose(np.sort(np.abs(n), 1), [0, 0, 1], atol=1e-15)
    # centroids at cube centres
    ijk = np.stack([(c["cx"][:N] / h - 0.5), (c["cy"][:N] / h - 0.5), (c["cz"][:N] / h - 0.5)], 1)
    assert np.allclose(ijk, np.round(ijk), atol=1e-12)
    # face dt = 0.5 sqrt(A) when fneq_min = 0 (infinite cap)
    assert np.allclose(f["dt"], 0.5 * h, atol=1e-16)
    # ghost centroid = mirror of the owner across the boundary face
    NI = pm.n_internal
    ow = pm.owner[NI:]
    fc = np.stack([f["fx"][NI:], f["fy"][NI:], f["fz"][NI:]], 1)
    xo = np.stack([c["cx"][ow], c["cy"][ow], c["cz"][ow]], 1)
    xg = np.stack([c["cx"][N:], c["cy"][N:], c["cz"][N:]], 1)
    assert np.allclose(xg, 2 * fc - xo, atol=1e-15)


@pytest.mark.parametrize("tets", [False, True])
def test_jittered_geometry_invariants(tets):
    pm = cavity_mesh(5, tets=tets, jitter=0.15)
    o = oracle_for(pm, pc.cavity_case())
    c, f = o.cells(), o.faces()
    N = pm.n_cells
    assert abs(c["vol"][:N].sum() - 1.0) < 1e-12
    assert np.all(c["vol"][:N] > 0)
    nA = np.stack([f["ni"], f["nj"], f["nk"]], 1) * f["area"][:, None]
    cl = np.zeros((N, 3))
    np.add.at(cl, pm.owner, nA)
    np.add.at(cl, pm.neighbour, -nA[: pm.n_internal])
    if tets:
        # planar triangles: unit normal times area is the exact vector area
        assert np.abs(cl).max() < 1e-15
    else:
        # warped quads: the reference's unit normal of the summed fan cross
        # products times the summed triangle areas overstates the vector area,
        # so cells close only to second order in the warp
        assert 1e-6 < np.abs(cl).max() < 0.05 * f["area"].max()
    # normals point from owner towards the face (outward for the owner)
    fc = np.stack([f["fx"], f["fy"], f["fz"]], 1)
    xo = np.stack([c["cx"][pm.owner], c["cy"][pm.owner], c["cz"][pm.owner]], 1)
    assert np.all(np.einsum("ij,ij->i", fc - xo, nA) > 0)


def test_points_are_single_precision():
    pm = meshgen.box_mesh(1, 1, 1, lengths=(0.1, 0.3, 0.7), patch_types=pc.CAVITY_PATCHES)
    o = oracle_for(pm, pc.cavity_case())
    v = o.cells()["vol"][0]
    exact32 = float(np.float32(0.1)) * float(np.float32(0.3)) * float(np.float32(0.7))
    assert abs(v - exact32) < 1e-14 * exact32
    assert abs(v - 0.1 * 0.3 * 0.7) > 1e-10 * exact32


def _lhs_from_rhs(o):
    """sum_j RHS_j d_j^T: for pure least squares RHS_j = w_j d_j, so this is the LHS."""
    c = o.cells()
    x = np.stack([c["cx"], c["cy"], c["cz"]], 1)
    off, cell, face = o.stencil()
    inv, rhs, _ = o.lsq()
    L = np.zeros((o.N, 3, 3))
    for i in range(o.N):
        for s, k in enumerate(range(off[i], off[i + 1])):
            d = x[cell[k]] - x[i]
            L[i] += np.outer(rhs[i, s], d)
    return inv.reshape(o.N, 3, 3), L


@pytest.mark.parametrize("tets", [False, True])
def test_lsq_inverse_structure(tets):
    pm = cavity_mesh(4, tets=tets, jitter=0.15)
    case = pc.cavity_case()
    case.lsq_exact_det = 1
    inv, L = _lhs_from_rhs(oracle_for(pm, case))
    prod = inv @ L
    assert np.abs(prod - np.eye(3)).max() < 1e-9
    case.lsq_exact_det = 0
    inv_r, L_r = _lhs_from_rhs(oracle_for(pm, case))
    prod = inv_r @ L_r
    s = prod[:, 0, 0]
    # the reference expression only rescales the inverse
    assert np.abs(prod - s[:, None, None] * np.eye(3)).max() < 1e-9 * np.abs(s).max()
    assert np.abs(s - 1).max() > 1e-3  # and it does rescale on jittered cells


def test_lsq_reference_determinant_vanishes_on_axis_aligned_hexes():
    pm = meshgen.box_mesh(4, 4, 4, patch_types=pc.CAVITY_PATCHES)
    inv, L = _lhs_from_rhs(oracle_for(pm, pc.cavity_case()))
    assert np.abs(inv @ L - np.eye(3)).max() < 1e-12


@pytest.mark.parametrize("tets", [False, True])
def test_linear_field_gradient_exact(tets):
    pm = cavity_mesh(5, tets=tets, jitter=0.15)
    case = allneumann_case()
    case.lsq_exact_det = 1
    o = oracle_for(pm, case)
    c = o.cells()
    a = np.array([0.3, -0.2, 0.7])
    phi = 1.0 + a[0] * c["cx"] + a[1] * c["cy"] + a[2] * c["cz"]
    off, cell, face = o.stencil()
    inv, rhs, _ = o.lsq()
    for i in range(o.N):
        acc = np.zeros(3)
        for s, k in enumerate(range(off[i], off[i + 1])):
            acc += rhs[i, s] * (phi[cell[k]] - phi[i])
        g = inv[i].reshape(3, 3) @ acc
        assert np.allclose(g, a, atol=1e-10)


def _interior_faces_away_from_boundary(pm, o):
    off, cell, face = o.stencil()
    touch = np.zeros(o.N, bool)
    for i in range(o.N):
        touch[i] = np.any(cell[off[i]:off[i + 1]] >= o.N)
    # cells whose stencil has a ghost, and their neighbours, see non-uniform gradients
    near = touch.copy()
    for i in np.flatnonzero(touch):
        near[cell[off[i]:off[i + 1]][cell[off[i]:off[i + 1]] < o.N]] = True
    nb = pm.neighbour
    return [f for f in range(pm.n_internal) if not near[pm.owner[f]] and not near[nb[f]]]


@pytest.mark.parametrize("gamma", [1.0, 0.3])
def test_uniform_state_gives_euler_flux(gamma):
    pm = cavity_mesh(6, tets=True, jitter=0.1)
    case = pc.cavity_case()
    case.pre_condition_gamma = gamma
    o = oracle_for(pm, case)
    rho0, u0 = 1.02, np.array([0.03, -0.01, 0.02])
    st = [np.full(o.T, rho0), np.full(o.T, u0[0]), np.full(o.T, u0[1]), np.full(o.T, u0[2])]
    o.eval_residual(*st)
    f = o.faces()
    faces = _interior_faces_away_from_boundary(pm, o)
    assert len(faces) > 50
    for fi in faces:
        n = np.array([f["ni"][fi], f["nj"][fi], f["nk"][fi]])
        A = f["area"][fi]
        F = o.face_flux(fi)
        un = u0 @ n
        exp = np.concatenate([[rho0 * un], rho0 * (n / 3 + u0 * un / gamma)]) * A
        assert np.allclose(F, exp, rtol=0, atol=1e-15)


@pytest.mark.parametrize("tets,jitter", [(False, 0.0), (True, 0.15)])
def test_rest_state_has_zero_residual(tets, jitter):
    # meshes with planar faces, where the reference's face vector areas close
    pm = cavity_mesh(5, tets=tets, jitter=jitter)
    o = oracle_for(pm, allneumann_case())
    st = [np.full(o.T, 1.0), np.zeros(o.T), np.zeros(o.T), np.zeros(o.T)]
    R, _ = o.eval_residual(*st)
    assert np.abs(R).max() < 1e-13


@pytest.mark.parametrize("tets", [False, True])
def test_global_conservation(tets):
    pm = cavity_mesh(5, tets=tets, jitter=0.1)
    case = pc.cavity_case()
    o = oracle_for(pm, case)
    from helpers import perturbed_state
    st = perturbed_state(case, o, amp=0.02, seed=4)
    R, _ = o.eval_residual(*st)
    vol = o.cells()["vol"][: o.N]
    bsum = sum(o.face_flux(f) for f in range(pm.n_internal, pm.n_faces))
    total = (R * vol[:, None]).sum(0)
    assert np.allclose(total, -bsum, atol=1e-13)


def test_body_force_adds_dt_f_per_step():
    pm = cavity_mesh(3, tets=False, jitter=0.0)
    case = allneumann_case()
    case.pressure_gradient = -1e-5   # force = -pressure_gradient = +1e-5 in x
    o = oracle_for(pm, case)
    st = [np.full(o.T, 1.0), np.zeros(o.T), np.zeros(o.T), np.zeros(o.T)]
    o.set_state(*st)
    dt = o.dt_local()
    r = o.run(1)
    assert r["rc"] == 0
    rho, u, v, w = o.state()
    assert np.allclose(u, dt * 1e-5, rtol=1e-9, atol=0)
    assert np.allclose(rho, 1.0, atol=1e-15)


def test_local_dt_closed_form_with_x_slot_areas():
    """Uniform cubes of side h: the reference accumulates every face's
    (|n_x|+|n_y|+|n_z|) A / 2 into the x slot (Solver.cpp:1679-1681), so
    A_x = 3 h^2 and A_y = A_z = 0, and with u = (U, 0, 0)
      lambda = (U + c) 3h^2 / gamma + 4 (2 nu / (gamma rho V)) (3h^2)^2,
      dt = CFL V / lambda.
    Per-direction areas (h^2 each) would give (U + 3c) h^2 instead."""
    n = 4
    h = 1.0 / n
    pm = meshgen.box_mesh(n, n, n, patch_types=pc.CAVITY_PATCHES)
    case = pc.cavity_case()
    o = oracle_for(pm, case)
    U = 0.05
    st = [np.full(o.T, 1.0), np.full(o.T, U), np.zeros(o.T), np.zeros(o.T)]
    o.set_state(*st)
    dt = o.dt_local()
    gam, nu, cfl = case.gamma, case.visc, case.time_marching_step
    c = U * np.sqrt(1 - gam + (gam / np.sqrt(3) / U) ** 2)
    V, Ax = h ** 3, 3 * h * h
    lam = (U + c) * Ax / gam + 4 * (2 * nu / (gam * 1.0 * V)) * Ax * Ax
    assert np.allclose(dt, cfl * V / lam, rtol=1e-12, atol=0)
    lam_split = (U + 3 * c) * h * h / gam + 4 * (2 * nu / (gam * V)) * 3 * h ** 4
    assert not np.allclose(dt, cfl * V / lam_split, rtol=1e-3)


def test_rk_amplification_polynomial():
    """Single hex cell, Dirichlet no-slip walls: one step from a small perturbation
    of rest equals P(dt J) delta with the reference's stage coefficients."""
    pm = meshgen.box_mesh(1, 1, 1, patch_types=pc.CAVITY_PATCHES)
    case = pc.cavity_case(lid_u=0.0)
    o = oracle_for(pm, case)
    T = o.T
    base = np.array([1.0, 0.0, 0.0, 0.0])

    def R_of(x):
        st = [np.full(T, x[q]) for q in range(4)]
        R, _ = o.eval_residual(*st)
        return R[0]

    eps = 1e-7
    J = np.zeros((4, 4))
    for q in range(4):
        e = np.zeros(4)
        e[q] = eps
        J[:, q] = (R_of(base + e) - R_of(base - e)) / (2 * eps)
    d0 = np.array([2e-6, 3e-6, -1e-6, 2e-6])
    x0 = base + d0
    o.set_state(*[np.full(T, x0[q]) for q in range(4)])
    dt = o.dt_local()[0]
    r = o.run(1)
    assert r["rc"] == 0
    x1 = np.array([v[0] for v in o.state()])
    # conservative variables: (rho, rho u, ...) ~ primitives to O(d^2)
    Z = dt * J
    I = np.eye(4)
    P_ref = I + Z + 7 / 12 * Z @ Z + 5 / 24 * Z @ Z @ Z + 1 / 24 * Z @ Z @ Z @ Z
    P_rk4 = I + Z + Z @ Z / 2 + Z @ Z @ Z / 6 + Z @ Z @ Z @ Z / 24
    d_ref, d_rk4 = P_ref @ d0, P_rk4 @ d0
    err = np.abs((x1 - base) - d_ref).max()
    assert np.abs(d_ref - d_rk4).max() > 50 * err  # the test tells the two schemes apart
    assert err < 1e-3 * np.abs(d0).max()


def test_periodic_z_residual_is_z_independent():
    """Uniform hexes, z sides periodic, a field that varies in x and y only:
    each z column must have one residual (the periodic ghost is the next cell
    of the column, so boundary cells see what interior cells see), and every
    partner must be the cell at the other end of its column.  Cells stay
    in generation order so that every x and y face is oriented the same way in
    every layer (lattice points with e.n = 0 are taken from the owner's side)."""
    from helpers import periodic_z_case, periodic_z_mesh
    nx, ny, nz = 6, 4, 5
    pm = periodic_z_mesh(nx, ny, nz, jitter=0.0, permute=False)
    case = periodic_z_case(float(nz))
    o = oracle_for(pm, case)
    cells = o.cells()
    x, y, z = cells["cx"], cells["cy"], cells["cz"]
    rho = 1.0 + 0.01 * np.sin(0.7 * x + 0.3) * np.cos(0.9 * y)
    u = 0.05 * (1 + 0.3 * np.cos(0.5 * x) * np.sin(0.8 * y + 0.2))
    v = 0.01 * np.sin(0.6 * x + 0.9 * y)
    w = np.zeros_like(x)
    R, _ = o.eval_residual(rho, u, v, w)
    N = o.N
    key = np.round(x[:N], 6) * 1000 + np.round(y[:N], 6)
    # the owner of the first boundary face takes the interior branch with a
    # ghost (the reference's "neighbour > n_cells" test), so its column is left out
    skip = key[pm.owner[pm.n_internal]]
    checked = 0
    for k in np.unique(key):
        if k == skip:
            continue
        col = R[key == k]
        assert len(col) == nz
        checked += 1
        spread = np.abs(col - col.mean(0)).max()
        assert spread <= 1e-13 * max(1.0, np.abs(R).max()), (k, spread)
    assert checked == nx * ny - 1
    # every periodic ghost's partner is the cell at the other end of its column
    part = o.partner()
    NI = pm.n_internal
    for b in range(pm.n_faces - NI):
        own = pm.owner[NI + b]
        fz = o.faces()["fz"][NI + b]
        if abs(fz) < 1e-9 or abs(fz - nz) < 1e-9:
            p = part[b]
            assert abs(x[p] - x[own]) < 1e-9 and abs(y[p] - y[own]) < 1e-9
            assert abs(z[p] - (nz - z[own])) < 1e-9
    # and the field is not trivially at rest
    assert np.abs(R[:, 1]).max() > 1e-6


@pytest.mark.parametrize("fneq_min", [0.0, 0.55])
def test_free_stream_is_steady_in_the_channel(fneq_min):
    # uniform inflow state everywhere: the ghosts of the Dirichlet inlet and
    # sides (velocity 2U - u = U, density Neumann rho), of the outlet (velocity
    # Neumann u, density Dirichlet 2 rho_b - rho = rho_b) and their gradients
    # all reproduce the free stream, so every residual vanishes; a wrong sign
    # in any of these ghost formulas leaves a residual at that patch
    from helpers import channel_mesh
    pm = channel_mesh(10, 5, 5, tets=True)
    case = pc.channel_case(re=40.0, length_scale=4.0, fneq_min=fneq_min)
    case.lsq_exact_det = 1
    o = oracle_for(pm, case)
    cells = o.cells()
    st = case.initial_state(cells["cx"], cells["cy"], cells["cz"])
    assert np.allclose(st[1], case.mach_no) and np.allclose(st[0], 1.0)
    R, g = o.eval_residual(*st)
    assert np.abs(R).max() < 1e-14, np.abs(R).max()
    assert np.abs(g).max() < 1e-12
    # and the same state with the inflow speed doubled is not steady at the inlet
    st2 = [st[0], 2 * st[1], st[2], st[3]]
    R2, _ = o.eval_residual(*st2)
    assert np.abs(R2).max() > 1e-6


def test_neumann_ghost_gradient_keeps_the_sign_of_n():
    # Neumann ghost gradients are grad * (1 - n) without abs (gradients.cpp:592-648):
    # on the x = 0 wall (n = -x) the normal component doubles, on x = 1 (n = +x)
    # it vanishes.  A tangential velocity mirror-symmetric about x = 1/2
    # therefore gets a residual that is mirror-symmetric away from the x walls
    # but not in the cells on them; (1 - |n|) would keep it symmetric everywhere.
    pm = cavity_mesh(6, tets=False, jitter=0.0)
    o = oracle_for(pm, allneumann_case())
    c = o.cells()
    N = o.N
    x = c["cx"]
    # v (a y velocity) is unchanged by the mirror x -> 1 - x, so its residual
    # mirrors too wherever the boundary treatment is symmetric
    st = [np.full(o.T, 1.0), np.zeros(o.T), 0.01 + 0.4 * (x - 0.5) ** 2, np.zeros(o.T)]
    R, g = o.eval_residual(*st)
    xyz = np.stack([c["cx"][:N], c["cy"][:N], c["cz"][:N]], 1)
    mx = xyz.copy()
    mx[:, 0] = 1.0 - mx[:, 0]
    dist = np.linalg.norm(xyz[None, :, :] - mx[:, None, :], axis=2)
    mirror = dist.argmin(1)
    assert dist.min(1).max() < 1e-6  # float32 points: centroids mirror to ~1e-8
    h = 1.0 / 6
    wall = (xyz[:, 0] < h) | (xyz[:, 0] > 1 - h)
    d = np.abs(R[:, 2] - R[mirror, 2])
    scale = np.abs(R[:, 2]).max()
    assert d[~wall].max() < 1e-6 * scale  # float32 points: the mesh mirrors to ~1e-8
    assert d[wall].max() > 1e-3 * scale


@pytest.mark.parametrize("inlet_vel,inlet_rho", [("dirichlet", "neumann"), ("inlet", "inlet"), ("inlet", "wall")])
def test_ghost_states_take_the_patch_or_the_owner_value(inlet_vel, inlet_rho):
    # free stream U with density rho_i inside; the inlet patch holds U and rho = 1.
    # Velocity types 1 (2U_b - u) and 6 (U_b) both reproduce U; density type 2
    # copies the owner (steady for any rho_i), types 6 and 7 impose the patch
    # density (steady only for rho_i = 1)
    from helpers import channel_mesh
    pm = channel_mesh(10, 5, 5, tets=True)
    case = pc.channel_case(re=40.0, length_scale=4.0)
    case.bcs["inlet"] = dict(rho=1.0, u=1.0, v=0.0, w=0.0, vel_type=inlet_vel, rho_type=inlet_rho)
    case.lsq_exact_det = 1
    o = oracle_for(pm, case)
    cells = o.cells()
    st = case.initial_state(cells["cx"], cells["cy"], cells["cz"])
    R, _ = o.eval_residual(*st)
    assert np.abs(R).max() < 1e-14
    st_r = [np.full(o.T, 1.01), st[1], st[2], st[3]]
    R_r, _ = o.eval_residual(*st_r)
    near_outlet_or_inlet = np.abs(R_r).max()
    if inlet_rho == "neumann":
        # the outlet imposes rho = 1 (type 1): only its cells feel the jump
        x = cells["cx"][: o.N]
        assert np.abs(R_r[x < 0.5 * x.max()]).max() < 1e-14
    else:
        x = cells["cx"][: o.N]
        assert np.abs(R_r[x < 0.2 * x.max()]).max() > 1e-6
    assert near_outlet_or_inlet > 1e-6


SYM_Y0_PATCHES = {"x0": "wall", "x1": "wall", "y0": "sym", "y1": "wall", "z0": "wall", "z1": "wall"}


def test_symmetry_ghost_is_the_mirror_image():
    # symmetry-x (type 8) ghosts copy rho, u, w and negate v (Solution.cpp:759-833).
    # On axis-aligned hexes the ghost centroid is the owner's mirror image in the
    # y = 0 plane, so a field that is even in y for rho, u, w and odd for v is
    # sampled exactly by the ghosts and the least-squares gradient of the cells
    # on that plane is exact.  A v that is not odd about the plane is not.
    n = 6
    pm = meshgen.box_mesh(n, n, n, patch_types=SYM_Y0_PATCHES)
    neu = dict(rho=1.0, u=0, v=0, w=0, vel_type="neumann", rho_type="neumann")
    sym = dict(rho=1.0, u=0, v=0, w=0, vel_type="symmetry-x", rho_type="symmetry-x")
    case = pc.Case(mach_no=0.05, reynolds_no=100.0, bcs={"wall": neu, "sym": sym})
    o = oracle_for(pm, case)
    c = o.cells()
    N = o.N
    x, y, z = c["cx"], c["cy"], c["cz"]
    h = 1.0 / n
    on_plane = (y[:N] < h) & (x[:N] > h) & (x[:N] < 1 - h) & (z[:N] > h) & (z[:N] < 1 - h)
    assert on_plane.sum() == (n - 2) ** 2
    want = np.array([0.1, 0, 0.2, 0.3, 0, -0.1, 0, 0.4, 0, 0.2, 0, 0.05])
    st = [1 + 0.1 * x + 0.2 * z, 0.3 * x - 0.1 * z, 0.4 * y, 0.2 * x + 0.05 * z]
    _, g = o.eval_residual(*st)
    assert np.abs(g[:N][on_plane] - want).max() < 1e-12
    st[2] = 0.4 * y + 0.1
    _, g = o.eval_residual(*st)
    assert np.abs(g[:N][on_plane, 7] - 0.4).min() > 0.1


def test_symmetry_patch_velocity_drags_the_stream():
    # the symmetry-x ghost velocity gradient is the Dirichlet one,
    # 2 (U_patch - u) / |d| n (gradients.cpp:592-648 via the type-8 branches),
    # so a uniform stream U along y-normal symmetry sides is steady when the
    # patch holds U, and the patch value otherwise acts like a sliding wall:
    # the cells on those sides are slowed by a slower patch, sped up by a
    # faster one, in proportion to U_patch - U; every other cell stays steady
    pt = {"x0": "inlet", "x1": "outlet", "y0": "sym", "y1": "sym", "z0": "farfield", "z1": "farfield"}
    pm = meshgen.box_mesh(10, 5, 5, lengths=(10.0, 5.0, 5.0), tets=True, jitter=0.15, jitter_seed=11,
                          permute_seed=12, patch_types=pt)
    ru = {}
    for u_patch in (1.0, 0.0, 0.5, 2.0):
        case = pc.channel_case(re=40.0, length_scale=4.0)
        case.bcs["sym"] = dict(rho=1.0, u=u_patch, v=0.0, w=0.0, vel_type="symmetry-x", rho_type="symmetry-x")
        case.lsq_exact_det = 1
        o = oracle_for(pm, case)
        cells = o.cells()
        st = case.initial_state(cells["cx"], cells["cy"], cells["cz"])
        R, _ = o.eval_residual(*st)
        y = cells["cy"][: o.N]
        side = (y < 1.0) | (y > 4.0)
        assert np.abs(R[~side]).max() < 1e-14, u_patch
        ru[u_patch] = R[side, 1]
    assert np.abs(ru[1.0]).max() < 1e-14
    assert ru[0.0].max() < 1e-14 and ru[0.0].min() < -1e-4
    assert ru[2.0].min() > -1e-14 and ru[2.0].max() > 1e-4
    # half the velocity difference, half the drag
    assert abs(ru[0.5].min() / ru[0.0].min() - 0.5) < 0.01


def _x0_face_fluxes(rho_type_x0):
    # rho = a + b x with a uniform stream U x, on a 4-cell-long box whose x = 0
    # patch has the given density type (patch value = the field at the ghost
    # centroid x = -1/2, velocity Dirichlet U), and on an 8-cell box over
    # [-4, 4] where x = 0 is interior.  The sides are Neumann and the far ends
    # Dirichlet at the field's face value, so every ghost outside x = 0
    # reproduces the linear field and every LSQ gradient near x = 0 is exact.
    # Returns the fluxes of the x = 0 faces of both meshes, matched by
    # centroid and given the box's outward orientation.
    a, b = 1.0, 0.02
    pt = {"x0": "inl", "x1": "out", "y0": "side", "y1": "side", "z0": "side", "z1": "side"}
    side = dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="neumann", rho_type="neumann")
    out = dict(rho=a + 4 * b, u=1.0, v=0.0, w=0.0, vel_type="neumann", rho_type="dirichlet")
    meshes = [(meshgen.box_mesh(4, 3, 3, lengths=(4.0, 3.0, 3.0), patch_types=pt),
               {"inl": dict(rho=a - 0.5 * b, u=1.0, v=0.0, w=0.0, vel_type="dirichlet", rho_type=rho_type_x0),
                "out": out, "side": side}),
              (meshgen.box_mesh(8, 3, 3, lengths=(8.0, 3.0, 3.0), origin=(-4.0, 0.0, 0.0), patch_types=pt),
               {"inl": dict(out, rho=a - 4 * b), "out": out, "side": side})]
    fluxes = []
    for pm, bcs in meshes:
        case = pc.Case(mach_no=0.05, reynolds_no=40.0, ref_length=4.0, bcs=bcs)
        case.lsq_exact_det = 1
        o = oracle_for(pm, case)
        x = o.cells()["cx"]
        o.eval_residual(a + b * x, np.full(o.T, case.mach_no), np.zeros(o.T), np.zeros(o.T))
        f = o.faces()
        sel = np.where(np.abs(f["fx"]) < 1e-9)[0]
        assert len(sel) == 9
        order = np.lexsort((f["fz"][sel], f["fy"][sel]))
        fluxes.append(np.array([o.face_flux(int(k)) * -np.sign(f["ni"][k]) for k in sel[order]]))
    return fluxes


def test_wall_density_ghost_continues_the_owner_gradient():
    # density type 7 ("wall") holds the patch value in the ghost and copies the
    # owner's gradient to it (gradients.cpp:395-504): with the patch value on
    # the linear field, its boundary flux equals the flux of the same face
    # inside a longer box.  Type 6 ("inlet", same ghost value, zero ghost
    # gradient) does not, nor do 1 and 2.
    fa, fb = _x0_face_fluxes("wall")
    scale = np.abs(fb).max()
    assert np.abs(fa - fb).max() < 1e-13 * scale
    for other in ("inlet", "neumann", "dirichlet"):
        fa, fb = _x0_face_fluxes(other)
        assert np.abs(fa - fb).max() > 1e-3 * scale, other


def _thin_box_residual(x_sides):
    # one cell thick in x, a field of y and z only: the x faces carry fluxes
    # that cancel pairwise when both are computed (periodic x, where each
    # cell is its own partner) and are skipped on "empty" patches
    if x_sides == "empty":
        pt_x = {"x0": "empty", "x1": "empty"}
        bcs = {"empty": dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="neumann", rho_type="neumann")}
    else:
        pt_x = {"x0": "xlo", "x1": "xhi"}
        per = dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="periodic", rho_type="periodic")
        bcs = {"xlo": dict(per, per_translate=(1.0, 0.0, 0.0)), "xhi": dict(per, per_translate=(-1.0, 0.0, 0.0))}
    bcs["wall"] = dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="dirichlet", rho_type="neumann")
    pm = meshgen.box_mesh(1, 5, 4, lengths=(1.0, 5.0, 4.0), permute_seed=3,
                          patch_types=dict(pt_x, y0="wall", y1="wall", z0="wall", z1="wall"))
    o = oracle_for(pm, pc.Case(mach_no=0.05, reynolds_no=40.0, ref_length=4.0, bcs=bcs))
    c = o.cells()
    y, z = c["cy"], c["cz"]
    st = [1.0 + 0.01 * np.sin(y) * np.cos(z), 0.05 * np.cos(0.7 * y + 0.3 * z), 0.01 * np.sin(z), 0.01 * np.cos(y)]
    R, _ = o.eval_residual(*st)
    return R, pm


def test_empty_patch_skip_starts_after_the_first_boundary_face():
    # the reference skips an "empty" face only if face > n_internal
    # (Solver.cpp:308), so the first boundary face is still computed: on a
    # one-cell-thick box with empty x sides (x0 holds boundary face 0) every
    # residual equals the periodic-x one, except the owner of boundary face 0,
    # which keeps that face's flux (the x pressure flux rho/3 A)
    Re, pm = _thin_box_residual("empty")
    Rp, _ = _thin_box_residual("periodic")
    scale = np.abs(Rp).max()
    d = np.abs(Re - Rp).max(1)
    first = pm.owner[pm.n_internal]
    others = np.arange(pm.n_cells) != first
    assert d[others].max() < 1e-13 * scale
    assert d[first] > scale


def test_first_boundary_face_takes_the_interior_branch():
    # the reference treats a face as a boundary face only if its neighbour
    # index is > n_cells (Solver.cpp:861); the ghost of boundary face 0 has
    # index n_cells, so that face is reconstructed two-sidedly from the ghost
    # state and gradient instead of taking the Dirichlet value.  With y and z
    # periodic and a field of x alone, every x = 0 face sees the same cells;
    # on a quadratic density the two-sided value differs from the patch value,
    # so face 0's flux differs from the other x = 0 faces, which all agree
    a, b, q = 1.0, 0.02, 0.004
    pt = {"x0": "inl", "x1": "out", "y0": "ylo", "y1": "yhi", "z0": "zlo", "z1": "zhi"}
    pm = meshgen.box_mesh(6, 3, 3, lengths=(6.0, 3.0, 3.0), permute_seed=5, patch_types=pt)
    per = dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="periodic", rho_type="periodic")
    bcs = {"inl": dict(rho=a, u=1.0, v=0.0, w=0.0, vel_type="dirichlet", rho_type="dirichlet"),
           "out": dict(rho=1.0, u=0.0, v=0.0, w=0.0, vel_type="neumann", rho_type="neumann"),
           "ylo": dict(per, per_translate=(0.0, 3.0, 0.0)), "yhi": dict(per, per_translate=(0.0, -3.0, 0.0)),
           "zlo": dict(per, per_translate=(0.0, 0.0, 3.0)), "zhi": dict(per, per_translate=(0.0, 0.0, -3.0))}
    case = pc.Case(mach_no=0.05, reynolds_no=40.0, ref_length=4.0, bcs=bcs)
    o = oracle_for(pm, case)
    x = o.cells()["cx"]
    o.eval_residual(a + b * x + q * x * x, np.full(o.T, case.mach_no), np.zeros(o.T), np.zeros(o.T))
    f = o.faces()
    x0 = [k for k in range(pm.n_internal, pm.n_faces) if abs(f["fx"][k]) < 1e-9]
    assert x0[0] == pm.n_internal and len(x0) == 9
    F = np.array([o.face_flux(k) for k in x0])
    scale = np.abs(F).max()
    assert np.abs(F[1:] - F[1]).max() < 1e-13 * scale
    assert np.abs(F[0] - F[1]).max() > 1e-6 * scale


def test_lsq_reference_determinant_on_sheared_cells():
    # hexes sheared in x by y (x += 0.3 y) couple x and y in the LSQ matrix but
    # leave z uncoupled (LHS_xz = LHS_yz = 0).  There the reference expression
    # (gradients.cpp:282-284) reduces to zz (xx yy + xy^2) against the true
    # zz (xx yy - xy^2), so its inverse is the true one scaled by
    # s = (xx yy - xy^2) / (xx yy + xy^2) in every cell
    pm = cavity_mesh(4, tets=False, jitter=0.0)
    pts = pm.points.astype(np.float64)
    pts[:, 0] += 0.3 * pts[:, 1]
    pm.points = pts.astype(np.float32).astype(pm.points.dtype)
    case = pc.cavity_case()
    case.lsq_exact_det = 0
    inv, L = _lhs_from_rhs(oracle_for(pm, case))
    xx, yy, xy = L[:, 0, 0], L[:, 1, 1], L[:, 0, 1]
    assert np.abs(L[:, 0, 2]).max() < 1e-12 * np.abs(L).max() and np.abs(L[:, 1, 2]).max() < 1e-12 * np.abs(L).max()
    assert np.abs(xy).min() > 1e-3 * np.abs(L).max()
    prod = inv @ L
    s = (xx * yy - xy * xy) / (xx * yy + xy * xy)
    assert np.abs(prod - s[:, None, None] * np.eye(3)).max() < 1e-9
