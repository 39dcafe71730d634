"""The whole program on the GPU: the plbfs executable runs an XML case end to
end in the order of src/program.cpp:40-137 (case and mesh reading, patch table,
initial conditions, the solver hand-off, output), and its restart fields equal
a direct solver run of the same case bit for bit."""
import os
import subprocess

import numpy as np
import pytest

from helpers import channel_mesh
from plbfs import _native as nat
from plbfs import meshgen
from plbfs.solver import GpuSolver, HostMesh
from test_host import CASE_XML, load_case

import ctypes as C

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def read_field(path):
    vals = [line.split(",") for line in open(path).read().splitlines() if line.strip()]
    assert [int(v[0]) for v in vals] == list(range(len(vals)))
    return np.array([float(v[1]) for v in vals])


def test_plbfs_executable_runs_a_case(gpu_available, tmp_path):
    exe = os.path.join(ROOT, "3d-plfbs-gpu_amd", "build", "plbfs")
    pm = channel_mesh(10, 5, 5, jitter=0.1)
    prefix = meshgen.write_polymesh(pm, str(tmp_path / "mesh"))
    out = tmp_path / "out"
    xml = tmp_path / "case.xml"
    xml.write_text(CASE_XML.format(out=str(out), mesh=prefix, prec="false"))
    print("plbfs", xml, flush=True)
    r = subprocess.run([exe, str(xml)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    # the same case through the C-ABI directly
    L = nat.lib()
    h = load_case(str(xml))
    p = nat.Params()
    nat.check(L.plbfs_case_params(h, C.byref(p)))
    hm = HostMesh.read_openfoam(prefix, p.visc, p.gamma, p.fneq_min)
    NP = len(pm.patch_names)
    patches = (nat.Patch * NP)()
    nat.check(L.plbfs_case_patches(h, hm.h, patches, NP))
    T = hm.n_cells + hm.n_ghosts
    st = [np.zeros(T) for _ in range(4)]
    nat.check(L.plbfs_case_initial_state(h, hm.h, *(nat.dptr(a) for a in st)))
    L.plbfs_case_free(h)
    params = {f: getattr(p, f) for f, _ in nat.Params._fields_}
    pl = [dict(rho_type=q.rho_type, vel_type=q.vel_type, rho=q.rho, u=q.u, v=q.v, w=q.w, is_empty=q.is_empty,
               per_translate=tuple(q.per_translate)) for q in patches]
    s = GpuSolver(hm, params, pl)
    s.set_state(*st)
    s.set_ghost_state(*(a[hm.n_cells:] for a in st))
    rd = s.run(20)
    assert rd["rc"] == 0 and rd["steps"] == 20
    direct = s.state()
    for q, name in enumerate(("rho", "u", "v", "w")):
        assert np.array_equal(read_field(str(out / (name + ".txt"))), direct[q]), name
    err = open(out / "error.txt").read().splitlines()
    assert len(err) == 20
    last = [float(x) for x in err[-1].split(",")]
    assert last[0] == 19 and np.allclose(last[2:], rd["residual"], rtol=1e-15, atol=0)
    g = open(out / "globals.txt").read()
    assert "Steps:20" in g and "Reynolds: 40" in g
    assert any(n.endswith(".dat") for n in os.listdir(out / "plt"))
