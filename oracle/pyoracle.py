"""ctypes binding of the CPU oracle (plbfs_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the reference result or the timed CPU baseline,
never as part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")
_libs = {}

ERR = {0: "ok", -1: "bad input", -2: "nan", -3: "rho blow-up", -4: "unsupported bc type"}


class MeshIn(C.Structure):
    _fields_ = [("n_points", C.c_int), ("n_cells", C.c_int), ("n_faces", C.c_int), ("n_internal", C.c_int),
                ("px", C.POINTER(C.c_float)), ("py", C.POINTER(C.c_float)), ("pz", C.POINTER(C.c_float)),
                ("face_off", C.POINTER(C.c_int)), ("face_verts", C.POINTER(C.c_int)),
                ("owner", C.POINTER(C.c_int)), ("neighbour", C.POINTER(C.c_int)),
                ("bface_patch", C.POINTER(C.c_int))]


class Patch(C.Structure):
    _fields_ = [("rho_type", C.c_int), ("vel_type", C.c_int), ("rho", C.c_double), ("u", C.c_double),
                ("v", C.c_double), ("w", C.c_double), ("is_empty", C.c_int), ("per_translate", C.c_double * 3)]


class Params(C.Structure):
    _fields_ = [("visc", C.c_double), ("gamma", C.c_double), ("fneq_min", C.c_double), ("cfl", C.c_double),
                ("max_velocity", C.c_double), ("ref_rho", C.c_double), ("force", C.c_double),
                ("tolerance", C.c_double), ("gradient_type", C.c_int), ("time_stepping", C.c_int),
                ("lsq_exact_det", C.c_int), ("domain_x", C.c_double), ("domain_y", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def plant_path(n):
    """Library of the oracle with planted mistake n (OM_PLANT in plbfs_oracle.c),
    built on demand; used only to show that a pin rejects it."""
    path = os.path.join(_HERE, "_build", "liboracle_plant%d.so" % n)
    if not os.path.exists(path):
        subprocess.check_call(["make", "-s", "-C", _HERE, "plants"])
    return path


def lib(path=None):
    path = path or _LIB_PATH
    if path not in _libs:
        if path == _LIB_PATH and not os.path.exists(path):
            build()
        L = C.CDLL(path)
        vp = C.c_void_p
        dp = C.POINTER(C.c_double)
        ip = C.POINTER(C.c_int)
        L.om_create.restype = vp
        L.om_create.argtypes = [C.POINTER(MeshIn), C.POINTER(Params), C.POINTER(Patch), C.c_int, ip]
        L.om_destroy.argtypes = [vp]
        L.om_n_total.argtypes = [vp]
        L.om_stencil_size.argtypes = [vp]
        L.om_get_cells.argtypes = [vp, dp, dp, dp, dp]
        L.om_get_faces.argtypes = [vp, dp, dp, dp, dp, dp, dp, dp, dp, ip]
        L.om_get_lsq.argtypes = [vp, dp, dp, dp]
        L.om_get_stencil.argtypes = [vp, ip, ip, ip]
        L.om_get_partner.argtypes = [vp, ip]
        L.om_set_state.argtypes = [vp, dp, dp, dp, dp]
        L.om_get_state.argtypes = [vp, dp, dp, dp, dp]
        L.om_get_residual.argtypes = [vp, dp]
        L.om_set_body_force.argtypes = [vp, dp, dp, dp]
        L.om_eval_residual.argtypes = [vp, dp, dp, dp, dp, dp, dp]
        L.om_face_flux.argtypes = [vp, C.c_int, dp]
        L.om_get_dt.argtypes = [vp, dp]
        L.om_get_temp.argtypes = [vp, dp, dp, dp, dp]
        L.om_run.argtypes = [vp, C.c_int, dp, ip, ip]
        L.om_run.restype = C.c_int
        _libs[path] = L
    return _libs[path]


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    """One oracle solver instance on a PolyMesh with a case (see plbfs.case)."""

    def __init__(self, mesh, case, libpath=None):
        self.L = L = lib(libpath)
        self._keep = []

        def keep(a):
            self._keep.append(a)
            return a

        pts = np.ascontiguousarray(mesh.points, dtype=np.float32)
        px, py, pz = (keep(np.ascontiguousarray(pts[:, q])) for q in range(3))
        fo = keep(np.ascontiguousarray(mesh.face_off, dtype=np.int32))
        fv = keep(np.ascontiguousarray(mesh.face_verts, dtype=np.int32))
        ow = keep(np.ascontiguousarray(mesh.owner, dtype=np.int32))
        nb = keep(np.ascontiguousarray(mesh.neighbour, dtype=np.int32))
        bp = keep(np.ascontiguousarray(mesh.bface_patch, dtype=np.int32))
        self.mi = MeshIn(pts.shape[0], mesh.n_cells, mesh.n_faces, mesh.n_internal,
                         _p(px, C.c_float), _p(py, C.c_float), _p(pz, C.c_float),
                         _p(fo, C.c_int), _p(fv, C.c_int), _p(ow, C.c_int),
                         _p(nb, C.c_int) if nb.size else None, _p(bp, C.c_int))
        pr = case.params
        self.params = Params(pr["visc"], pr["gamma"], pr["fneq_min"], pr["cfl"], pr["max_velocity"],
                             pr["ref_rho"], pr["force"], pr["tolerance"], pr["gradient_type"],
                             pr["time_stepping"], int(pr.get("lsq_exact_det", 0)),
                             float(pr.get("domain_x", 1.0)), float(pr.get("domain_y", 1.0)))
        patches = (Patch * len(case.patches))()
        for q, P in enumerate(case.patches):
            patches[q] = Patch(P["rho_type"], P["vel_type"], P["rho"], P["u"], P["v"], P["w"],
                               int(P.get("is_empty", 0)), (C.c_double * 3)(*P.get("per_translate", (0, 0, 0))))
        self._patches = patches
        err = C.c_int(0)
        self.h = L.om_create(C.byref(self.mi), C.byref(self.params), patches, len(case.patches), C.byref(err))
        if not self.h:
            raise ValueError("oracle create failed: %s" % ERR.get(err.value, err.value))
        self.N = mesh.n_cells
        self.F = mesh.n_faces
        self.T = L.om_n_total(self.h)

    def close(self):
        if self.h:
            self.L.om_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def cells(self):
        a = [np.zeros(self.T) for _ in range(4)]
        self.L.om_get_cells(self.h, *(_p(x, C.c_double) for x in a))
        return dict(cx=a[0], cy=a[1], cz=a[2], vol=a[3])

    def faces(self):
        a = [np.zeros(self.F) for _ in range(8)]
        nbr = np.zeros(self.F, dtype=np.int32)
        self.L.om_get_faces(self.h, *(_p(x, C.c_double) for x in a), _p(nbr, C.c_int))
        return dict(fx=a[0], fy=a[1], fz=a[2], ni=a[3], nj=a[4], nk=a[5], area=a[6], dt=a[7], nbr=nbr)

    def lsq(self):
        inv = np.zeros((self.N, 9))
        rhs = np.zeros((self.N, 6, 3))
        ca = np.zeros(self.N)
        self.L.om_get_lsq(self.h, _p(inv, C.c_double), _p(rhs, C.c_double), _p(ca, C.c_double))
        return inv, rhs, ca

    def stencil(self):
        n = self.L.om_stencil_size(self.h)
        off = np.zeros(self.N + 1, dtype=np.int32)
        cell = np.zeros(n, dtype=np.int32)
        face = np.zeros(n, dtype=np.int32)
        self.L.om_get_stencil(self.h, _p(off, C.c_int), _p(cell, C.c_int), _p(face, C.c_int))
        return off, cell, face

    def partner(self):
        out = np.zeros(max(self.T - self.N, 1), dtype=np.int32)
        self.L.om_get_partner(self.h, _p(out, C.c_int))
        return out[: self.T - self.N]

    def set_state(self, rho, u, v, w):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, u, v, w)]
        for x in arrs:
            assert x.shape == (self.T,)
        self.L.om_set_state(self.h, *(_p(x, C.c_double) for x in arrs))

    def state(self):
        a = [np.zeros(self.N) for _ in range(4)]
        self.L.om_get_state(self.h, *(_p(x, C.c_double) for x in a))
        return a

    def set_body_force(self, fx=None, fy=None, fz=None):
        if fx is None:
            self.L.om_set_body_force(self.h, None, None, None)
            return
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (fx, fy, fz)]
        self._force = arrs
        self.L.om_set_body_force(self.h, *(_p(x, C.c_double) for x in arrs))

    def eval_residual(self, rho, u, v, w):
        arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (rho, u, v, w)]
        R = np.zeros((self.N, 4))
        g = np.zeros((self.T, 12))
        self.L.om_eval_residual(self.h, *(_p(x, C.c_double) for x in arrs), _p(R, C.c_double), _p(g, C.c_double))
        return R, g

    def temp(self):
        """Temp state over T cells after eval_residual (ghosts from the boundary update)."""
        a = [np.zeros(self.T) for _ in range(4)]
        self.L.om_get_temp(self.h, *(_p(x, C.c_double) for x in a))
        return a

    def face_flux(self, face):
        out = np.zeros(4)
        self.L.om_face_flux(self.h, int(face), _p(out, C.c_double))
        return out

    def dt_local(self):
        out = np.zeros(self.N)
        self.L.om_get_dt(self.h, _p(out, C.c_double))
        return out

    def run(self, nsteps):
        hist = np.zeros((max(nsteps, 1), 4))
        done = C.c_int(0)
        conv = C.c_int(0)
        rc = self.L.om_run(self.h, int(nsteps), _p(hist, C.c_double), C.byref(done), C.byref(conv))
        return dict(rc=rc, steps=done.value, converged=bool(conv.value), hist=hist[:done.value])


def _bind_ibm(L):
    dp = C.POINTER(C.c_double)
    L.om_set_ibm.argtypes = [C.c_void_p, C.c_int, dp, dp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]
    L.om_ibm_step.argtypes = [C.c_void_p, dp, dp, dp]
    L.om_get_ibm.argtypes = [C.c_void_p, dp, dp, dp]
    L.om_get_body_force.argtypes = [C.c_void_p, dp]


def set_ibm(o, x_ref, x, stiffness, node_area, dt, h, kernel):
    """Lagrangian nodes for an Oracle (kernel 0 = 2-point hat, 1 = 4-point Peskin)."""
    L = o.L
    _bind_ibm(L)
    xr = np.ascontiguousarray(x_ref, dtype=np.float64).reshape(-1)
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    L.om_set_ibm(o.h, xr.size // 3, _p(xr, C.c_double), _p(xx, C.c_double), stiffness, node_area, dt, h, kernel)
    o.n_nodes = xr.size // 3


def ibm_step(o, u, v, w):
    L = o.L
    _bind_ibm(L)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, v, w)]
    L.om_ibm_step(o.h, *(_p(a, C.c_double) for a in arrs))


def ibm_state(o):
    L = o.L
    _bind_ibm(L)
    n = o.n_nodes
    x, u, f = (np.zeros(3 * n) for _ in range(3))
    L.om_get_ibm(o.h, _p(x, C.c_double), _p(u, C.c_double), _p(f, C.c_double))
    return x.reshape(n, 3), u.reshape(n, 3), f.reshape(n, 3)


def body_force(o):
    L = o.L
    _bind_ibm(L)
    f = np.zeros(3 * o.N)
    L.om_get_body_force(o.h, _p(f, C.c_double))
    return f.reshape(o.N, 3)
