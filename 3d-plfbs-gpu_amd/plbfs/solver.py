"""Python front end over the C-ABI (plumbing for tests and benchmarks).

``HostMesh`` wraps the C++ mesh object (reader/geometry); ``GpuSolver`` wraps a
solver context.  All arrays are in the caller's (OpenFOAM file) cell order.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _native as nat


class HostMesh:
    """Host mesh built by libplbfs (OpenFOAM files or in-memory arrays)."""

    def __init__(self, handle, patch_names=None):
        self.h = handle
        L = nat.lib()
        info = np.zeros(6, dtype=np.int64)
        nat.check(L.plbfs_mesh_info(self.h, info.ctypes.data_as(C.POINTER(C.c_int64))))
        self.n_cells, self.n_faces, self.n_internal, self.n_points, self.n_patches, self.n_ghosts = (
            int(x) for x in info)
        if patch_names is None:
            patch_names = []
            buf = C.create_string_buffer(256)
            for p in range(self.n_patches):
                nat.check(L.plbfs_mesh_patch_type(self.h, p, buf, 256))
                patch_names.append(buf.value.decode())
        self.patch_names = list(patch_names)
        self.view = nat.MeshView()
        nat.check(L.plbfs_mesh_view_get(self.h, C.byref(self.view)))

    @classmethod
    def from_polymesh(cls, pm, visc, gamma=1.0, fneq_min=0.0):
        L = nat.lib()
        pts = np.ascontiguousarray(pm.points, dtype=np.float32)
        px, py, pz = (np.ascontiguousarray(pts[:, q]) for q in range(3))
        fo = np.ascontiguousarray(pm.face_off, dtype=np.int32)
        fv = np.ascontiguousarray(pm.face_verts, dtype=np.int32)
        ow = np.ascontiguousarray(pm.owner, dtype=np.int32)
        nb = np.ascontiguousarray(pm.neighbour, dtype=np.int32)
        names = (C.c_char_p * len(pm.patch_names))(*[n.encode() for n in pm.patch_names])
        sizes = np.ascontiguousarray(pm.patch_sizes, dtype=np.int32)
        h = C.c_void_p()
        nat.check(L.plbfs_mesh_from_arrays(pts.shape[0], nat.fptr(px), nat.fptr(py), nat.fptr(pz), pm.n_cells,
                                           pm.n_faces, pm.n_internal, nat.iptr(fo), nat.iptr(fv), nat.iptr(ow),
                                           nat.iptr(nb) if nb.size else None, len(pm.patch_names), names,
                                           nat.iptr(sizes), visc, gamma, fneq_min, C.byref(h)))
        return cls(h, pm.patch_names)

    @classmethod
    def read_openfoam(cls, prefix, visc, gamma=1.0, fneq_min=0.0):
        h = C.c_void_p()
        nat.check(nat.lib().plbfs_mesh_read_openfoam(prefix.encode(), visc, gamma, fneq_min, C.byref(h)))
        return cls(h)

    def close(self):
        if self.h:
            nat.lib().plbfs_mesh_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _arr(self, ptr, n, dtype=np.float64):
        return np.ctypeslib.as_array(ptr, shape=(n,)).astype(dtype, copy=True)

    def geometry(self):
        v = self.view
        T = self.n_cells + self.n_ghosts
        F = self.n_faces
        cc = self._arr(v.cell_centroid, 3 * T).reshape(T, 3)
        return dict(cx=cc[:, 0], cy=cc[:, 1], cz=cc[:, 2], vol=self._arr(v.cell_volume, self.n_cells),
                    fc=self._arr(v.face_centroid, 3 * F).reshape(F, 3),
                    fn=self._arr(v.face_normal, 3 * F).reshape(F, 3), area=self._arr(v.face_area, F),
                    dt=self._arr(v.face_dt, F), owner=self._arr(v.owner, F, np.int32),
                    nbr=self._arr(v.neighbour, F, np.int32))

    def bface_patch(self):
        p = nat.lib().plbfs_mesh_bface_patch(self.h)
        return np.ctypeslib.as_array(p, shape=(self.n_ghosts,)).copy() if self.n_ghosts else np.zeros(0, np.int32)

    def raw(self):
        L = nat.lib()
        P, F, NI = self.n_points, self.n_faces, self.n_internal
        px, py, pz = (np.zeros(P, np.float32) for _ in range(3))
        ow = np.zeros(F, np.int32)
        nb = np.zeros(max(NI, 1), np.int32)
        fo = np.zeros(F + 1, np.int32)
        fv = np.zeros(int(L.plbfs_mesh_n_face_verts(self.h)), np.int32)
        nat.check(L.plbfs_mesh_raw(self.h, nat.fptr(px), nat.fptr(py), nat.fptr(pz), nat.iptr(ow), nat.iptr(nb),
                                   nat.iptr(fo), nat.iptr(fv)))
        return dict(points=np.stack([px, py, pz], 1), owner=ow, neighbour=nb[:NI], face_off=fo, face_verts=fv)


def make_params(p):
    return nat.Params(p["visc"], p["gamma"], p["fneq_min"], p["cfl"], p["max_velocity"], p["ref_rho"], p["force"],
                      p["tolerance"], p["gradient_type"], p["time_stepping"], p.get("precision", 0),
                      p.get("check_interval", 10), int(p.get("lsq_exact_det", 0)), float(p.get("domain_x", 1.0)),
                      float(p.get("domain_y", 1.0)))


def make_patches(patches):
    arr = (nat.Patch * len(patches))()
    for q, P in enumerate(patches):
        arr[q] = nat.Patch(P["rho_type"], P["vel_type"], P["rho"], P["u"], P["v"], P["w"], int(P.get("is_empty", 0)),
                           (C.c_double * 3)(*P.get("per_translate", (0.0, 0.0, 0.0))))
    return arr


def state_arrays(arrs, n):
    """The first n values of rho, u, v, w as contiguous float64 arrays.  The
    C-ABI reads n values from each pointer, so a shorter input is an error."""
    out = []
    for name, x in zip(("rho", "u", "v", "w"), arrs):
        a = np.asarray(x, dtype=np.float64).reshape(-1)
        if a.size < n:
            raise ValueError("%s has %d values; the mesh has %d cells" % (name, a.size, n))
        out.append(np.ascontiguousarray(a[:n]))
    return out


def node_arrays(x_ref, x):
    """Anchor and current node positions as flat float64 [3n] arrays of one n."""
    xr = np.ascontiguousarray(x_ref, dtype=np.float64).reshape(-1)
    xx = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
    if xr.size % 3 or xx.size != xr.size:
        raise ValueError("node arrays must both hold 3n values (got %d and %d)" % (xr.size, xx.size))
    return xr, xx


class GpuSolver:
    """One solver context on the current GPU."""

    def __init__(self, hmesh: HostMesh, params: dict, patches: list):
        L = nat.lib()
        self.mesh = hmesh
        self.N = hmesh.n_cells
        self._params = make_params(params)
        self._patches = make_patches(patches)
        self._bp = np.ascontiguousarray(hmesh.bface_patch(), dtype=np.int32)
        h = C.c_void_p()
        nat.check(L.plbfs_create(C.byref(hmesh.view), C.byref(self._params), self._patches, len(patches),
                                 nat.iptr(self._bp) if self._bp.size else None, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            nat.lib().plbfs_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, rho, u, v, w):
        a = state_arrays((rho, u, v, w), self.N)
        nat.check(nat.lib().plbfs_set_state(self.h, *(nat.dptr(x) for x in a)))

    def state(self):
        a = [np.zeros(self.N) for _ in range(4)]
        nat.check(nat.lib().plbfs_get_state(self.h, *(nat.dptr(x) for x in a)))
        return a

    def set_ghost_state(self, rho, u, v, w):
        """Initial ghost state, one value per boundary face (plbfs_set_ghost_state)."""
        a = state_arrays((rho, u, v, w), self.mesh.n_faces - self.mesh.n_internal)
        nat.check(nat.lib().plbfs_set_ghost_state(self.h, *(nat.dptr(x) for x in a)))

    def run(self, nsteps, raise_on_error=False):
        r = nat.Result()
        rc = nat.lib().plbfs_run(self.h, int(nsteps), C.byref(r))
        if rc not in (0, -2, -3, -7) or (raise_on_error and rc != 0):
            nat.check(rc)
        hist = np.zeros((max(r.steps, 1), 4))
        nat.check(nat.lib().plbfs_get_residual_history(self.h, nat.dptr(hist), r.steps))
        return dict(rc=rc, status=r.status, steps=r.steps, converged=bool(r.converged), total=r.total_steps,
                    residual=list(r.residual), ms=r.solve_ms, hist=hist[: r.steps])

    def run_noalloc(self, nsteps):
        """plbfs_run_noalloc: the same steps as run(), working in the library's
        fixed per-device scratch array instead of allocated memory."""
        r = nat.Result()
        hist = np.zeros((max(int(nsteps), 1), 4))
        rc = nat.lib().plbfs_run_noalloc(self.h, int(nsteps), C.byref(r), nat.dptr(hist))
        if rc not in (0, -2, -3, -7):
            nat.check(rc)
        return dict(rc=rc, status=r.status, steps=r.steps, converged=bool(r.converged), total=r.total_steps,
                    residual=list(r.residual), ms=r.solve_ms, hist=hist[: r.steps])

    def eval_residual(self):
        g = np.zeros((self.N, 12))
        R = np.zeros((self.N, 4))
        nat.check(nat.lib().plbfs_eval_residual(self.h, nat.dptr(g), nat.dptr(R)))
        return R, g

    def dt_local(self):
        d = np.zeros(self.N)
        nat.check(nat.lib().plbfs_get_dt(self.h, nat.dptr(d)))
        return d

    def set_ibm(self, x_ref, x, stiffness, node_area, dt, h, kernel):
        xr, xx = node_arrays(x_ref, x)
        self._ibm_keep = (xr, xx)
        nodes = nat.IbmNodes(xr.size // 3, nat.dptr(xr), nat.dptr(xx), stiffness, node_area, dt, h, kernel)
        nat.check(nat.lib().plbfs_set_ibm(self.h, C.byref(nodes)))
        self.n_nodes = xr.size // 3

    def ibm_state(self):
        n = self.n_nodes
        x, v, f = (np.zeros(3 * n) for _ in range(3))
        nat.check(nat.lib().plbfs_get_ibm(self.h, nat.dptr(x), nat.dptr(v), nat.dptr(f)))
        return x.reshape(n, 3), v.reshape(n, 3), f.reshape(n, 3)

    def body_force(self):
        f = np.zeros(3 * self.N)
        nat.check(nat.lib().plbfs_get_body_force(self.h, nat.dptr(f)))
        return f.reshape(self.N, 3)

    def wall_forces(self, patch):
        """(drag, lift) on the boundary faces of patch index `patch` (reference sums)."""
        out = np.zeros(2)
        nat.check(nat.lib().plbfs_wall_forces(self.h, C.byref(self.mesh.view), nat.iptr(self._bp) if self._bp.size
                                              else None, int(patch), nat.dptr(out)))
        return out

    def ibm_body_force(self):
        out = np.zeros(3)
        nat.check(nat.lib().plbfs_ibm_body_force(self.h, nat.dptr(out)))
        return out

    def enable_timing(self, on=True):
        nat.check(nat.lib().plbfs_enable_timing(self.h, 1 if on else 0))

    def kernel_stats(self, which):
        ms = C.c_double(0)
        n = C.c_int32(0)
        nat.check(nat.lib().plbfs_kernel_stats(self.h, which, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def layout_info(self):
        info = np.zeros(10, dtype=np.int64)
        nat.check(nat.lib().plbfs_layout_info(self.h, info.ctypes.data_as(C.POINTER(C.c_int64))))
        keys = ["cells", "faces", "items", "stencil_slots", "ghosts", "blocks", "max_items", "gather",
                "block_cells", "lds_bytes"]
        return dict(zip(keys, (int(x) for x in info)))


class GpuBatch:
    """B cases on one mesh (plbfs_batch_*): per-case viscosity, state and IBM nodes."""

    def __init__(self, hmesh: HostMesh, params: dict, patches: list, case_visc):
        L = nat.lib()
        self.mesh = hmesh
        self.N = hmesh.n_cells
        self.B = len(case_visc)
        self._params = make_params(params)
        self._patches = make_patches(patches)
        self._bp = np.ascontiguousarray(hmesh.bface_patch(), dtype=np.int32)
        self._visc = np.ascontiguousarray(case_visc, dtype=np.float64)
        h = C.c_void_p()
        nat.check(L.plbfs_batch_create(C.byref(hmesh.view), C.byref(self._params), self._patches, len(patches),
                                       nat.iptr(self._bp) if self._bp.size else None, self.B, nat.dptr(self._visc),
                                       C.byref(h)))
        self.h = h
        self._ibm_keep = {}

    def close(self):
        if self.h:
            nat.lib().plbfs_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, case, rho, u, v, w):
        a = state_arrays((rho, u, v, w), self.N)
        nat.check(nat.lib().plbfs_batch_set_state(self.h, int(case), *(nat.dptr(x) for x in a)))

    def state(self, case):
        a = [np.zeros(self.N) for _ in range(4)]
        nat.check(nat.lib().plbfs_batch_get_state(self.h, int(case), *(nat.dptr(x) for x in a)))
        return a

    def set_ibm(self, case, x_ref, x, stiffness, node_area, dt, h, kernel):
        xr, xx = node_arrays(x_ref, x)
        self._ibm_keep[case] = (xr, xx)
        nodes = nat.IbmNodes(xr.size // 3, nat.dptr(xr), nat.dptr(xx), stiffness, node_area, dt, h, kernel)
        nat.check(nat.lib().plbfs_batch_set_ibm(self.h, int(case), C.byref(nodes)))

    def run(self, nsteps):
        res = (nat.Result * self.B)()
        rc = nat.lib().plbfs_batch_run(self.h, int(nsteps), res)
        if rc not in (0, -2, -3, -7):
            nat.check(rc)
        return rc, [dict(status=r.status, steps=r.steps, converged=bool(r.converged), residual=list(r.residual),
                         ms=r.solve_ms) for r in res]

    def set_tiling(self, block_cells, cases_per_block, stage=0):
        nat.check(nat.lib().plbfs_batch_set_tiling(self.h, int(block_cells), int(cases_per_block), int(stage)))

    def autotune(self, trial_steps=5):
        """(cells per block, cases per workgroup, staging, ms per step) of the fastest tiling."""
        bc = C.c_int32(0)
        cp = C.c_int32(0)
        sg = C.c_int32(0)
        ms = C.c_double(0)
        nat.check(nat.lib().plbfs_batch_autotune(self.h, int(trial_steps), C.byref(bc), C.byref(cp), C.byref(sg),
                                                 C.byref(ms)))
        return bc.value, cp.value, sg.value, ms.value

    def layout_info(self):
        info = np.zeros(5, dtype=np.int64)
        nat.check(nat.lib().plbfs_batch_layout_info(self.h, info.ctypes.data_as(C.POINTER(C.c_int64))))
        return dict(zip(["cases", "block_cells", "cases_per_block", "lds_bytes", "stage"], (int(x) for x in info)))
