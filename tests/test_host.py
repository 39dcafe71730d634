"""Host-side checks that need no GPU: library exports, mesh generator, OpenFOAM
reader round trip, product geometry against the oracle, XML case derivation,
field and Tecplot output, and clean errors without a device."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from helpers import cavity_mesh, channel_mesh, oracle_for
from plbfs import _native as nat
from plbfs import case as pc
from plbfs import meshgen
from plbfs.solver import GpuSolver, HostMesh

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(plbfs_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exports_every_declared_symbol():
    L = nat.lib()
    decl = declared_symbols()
    assert decl == set(nat.EXPORTS)
    for name in decl:
        assert hasattr(L, name), name


@pytest.mark.parametrize("tets", [False, True])
def test_meshgen_counts_and_topology(tets):
    n = (4, 3, 2)
    pm = meshgen.box_mesh(*n, tets=tets, jitter=0.1, jitter_seed=1, permute_seed=2, patch_types=pc.CHANNEL_PATCHES)
    H = n[0] * n[1] * n[2]
    C_ = 6 * H if tets else H
    bq = 2 * (n[0] * n[1] + n[1] * n[2] + n[0] * n[2])
    B = 2 * bq if tets else bq
    fpc = 4 if tets else 6
    assert pm.n_cells == C_
    assert pm.n_faces - pm.n_internal == B
    assert 2 * pm.n_internal + B == fpc * C_
    assert np.all(pm.owner[: pm.n_internal] < pm.neighbour)
    # internal faces in upper-triangular order
    key = pm.owner[: pm.n_internal].astype(np.int64) * pm.n_cells + pm.neighbour
    assert np.all(np.diff(key) > 0)
    assert sum(pm.patch_sizes) == B


def test_config3_face_counts_formula():
    # 88^3 x 6 tets: F = (4 * 4088832 + 92928) / 2
    n = 88
    cells = 6 * n ** 3
    bdry = 2 * 2 * 3 * n * n
    assert cells == 4088832 and bdry == 92928
    assert (4 * cells + bdry) // 2 == 8224128


@pytest.mark.parametrize("tets", [False, True])
def test_polymesh_roundtrip_through_cpp_reader(tmp_path, tets):
    pm = cavity_mesh(4, tets=tets, jitter=0.12)
    prefix = meshgen.write_polymesh(pm, str(tmp_path / "polyMesh"))
    lines = open(prefix + "owner").read().split("\n")
    assert lines[12].lstrip().startswith("note") and "nInternalFaces:%d" % pm.n_internal in lines[12]
    hm = HostMesh.read_openfoam(prefix, visc=0.0005)
    assert (hm.n_cells, hm.n_faces, hm.n_internal) == (pm.n_cells, pm.n_faces, pm.n_internal)
    raw = hm.raw()
    assert np.array_equal(raw["points"], pm.points)  # float32 values survive the text round trip
    assert np.array_equal(raw["owner"], pm.owner)
    assert np.array_equal(raw["neighbour"], pm.neighbour)
    assert np.array_equal(raw["face_off"], pm.face_off)
    assert np.array_equal(raw["face_verts"], pm.face_verts)
    assert hm.patch_names == pm.patch_names
    assert np.array_equal(hm.bface_patch(), pm.bface_patch)


def test_reader_errors_are_clean(tmp_path):
    with pytest.raises(nat.PlbfsError) as e:
        HostMesh.read_openfoam(str(tmp_path) + "/missing/", visc=0.001)
    assert e.value.rc == -1


@pytest.mark.parametrize("fname,pattern,repl", [
    ("owner", r"nFaces:\d+", "nFaces:-5"),
    ("owner", r"nFaces:\d+", "nFaces:99999999999"),
    ("owner", r"nPoints:\d+", "nPoints:2000000000"),
    ("owner", r"nInternalFaces:\d+", "nInternalFaces:100000"),
    ("boundary", r"nFaces\s+\d+;", "nFaces abc;"),
    ("boundary", r"nFaces\s+\d+;", "nFaces 99999999999999999999;"),
])
def test_reader_rejects_bad_counts(tmp_path, fname, pattern, repl):
    # counts that are negative, overflow int32, exceed what the file can hold
    # or are not numbers are input errors, not crashes or giant allocations
    import shutil
    pm = cavity_mesh(3)
    good = meshgen.write_polymesh(pm, str(tmp_path / "good"))
    HostMesh.read_openfoam(good, visc=0.001)
    bad = str(tmp_path / "bad")
    shutil.copytree(good, bad)
    path = os.path.join(bad, fname)
    text = open(path).read()
    new, n = re.subn(pattern, repl, text, count=1)
    assert n == 1
    open(path, "w").write(new)
    with pytest.raises(nat.PlbfsError) as e:
        HostMesh.read_openfoam(bad + "/", visc=0.001)
    assert e.value.rc == -1


@pytest.mark.parametrize("tets,jitter", [(False, 0.15), (True, 0.1)])
def test_product_geometry_equals_oracle_bitwise(tets, jitter):
    pm = cavity_mesh(5, tets=tets, jitter=jitter)
    case = pc.cavity_case()
    o = oracle_for(pm, case)
    hm = HostMesh.from_polymesh(pm, case.visc, case.gamma, case.fneq_min)
    g = hm.geometry()
    c, f = o.cells(), o.faces()
    for k in ("cx", "cy", "cz"):
        assert np.array_equal(g[k], c[k])
    assert np.array_equal(g["vol"], c["vol"][: pm.n_cells])
    assert np.array_equal(g["fc"], np.stack([f["fx"], f["fy"], f["fz"]], 1))
    assert np.array_equal(g["fn"], np.stack([f["ni"], f["nj"], f["nk"]], 1))
    assert np.array_equal(g["area"], f["area"])
    assert np.array_equal(g["dt"], f["dt"])
    assert np.array_equal(g["nbr"], f["nbr"])


CASE_XML = """<root>
<global_variables>
  <tolerance>1e-6</tolerance><pre_condition_gamma>1</pre_condition_gamma>
  <simulation_length>20</simulation_length><time_marching_step>0.5</time_marching_step>
  <reynolds_no>40</reynolds_no><mach_no>0.05</mach_no><simulation_name>chan</simulation_name>
  <output_directory>{out}</output_directory><max_multi_grid_levels>0</max_multi_grid_levels>
  <FMG_levels>0</FMG_levels><FMG_tolerance>0</FMG_tolerance><dissipation_kappa_2>0</dissipation_kappa_2>
  <dissipation_kappa_4>0</dissipation_kappa_4><martinelli_exponent>0</martinelli_exponent>
  <testcase>1</testcase><mesh_type>3</mesh_type><scale>1</scale><womersley_no>0</womersley_no>
  <output_step>10</output_step><import_file>{mesh}</import_file><import_format>openfoam</import_format>
  <gradient_calc>LS</gradient_calc><restart_analysis>false</restart_analysis>
  <time_stepping>local</time_stepping><preconditioning>{prec}</preconditioning><gpu_solver>1</gpu_solver>
  <fneq_min>0</fneq_min>
</global_variables>
<geometry><x>10</x><y>5</y><z>5</z><dx>1</dx><dy>1</dy><dz>1</dz><streaming_dt>0.5</streaming_dt>
  <origin_x>0</origin_x><origin_y>0</origin_y><origin_z>0</origin_z><ref_length>4</ref_length></geometry>
<initial_conditions><average_rho>1</average_rho><pressure_gradient>0</pressure_gradient>
  <rho><gradient><x>0</x><y>0</y><z>0</z></gradient><origin_magnitude><x>1</x><y>0</y><z>0</z></origin_magnitude></rho>
  <origin_location><x>0</x><y>0</y><z>0</z></origin_location>
  <vel><gradient><x>0</x><y>0</y><z>0</z></gradient><origin_magnitude><x>1</x><y>0</y><z>0</z></origin_magnitude></vel>
  <u>0</u><v>0</v><w>0</w></initial_conditions>
<boundary_conditions/>
<uns_boundary_conditions>
  <bc><name>inlet</name><rho>1</rho><u>1</u><v>0</v><w>0</w><vel_type>dirichlet</vel_type><rho_type>neumann</rho_type></bc>
  <bc><name>outlet</name><rho>1</rho><u>0</u><v>0</v><w>0</w><vel_type>neumann</vel_type><rho_type>dirichlet</rho_type></bc>
  <bc><name>farfield</name><rho>1</rho><u>1</u><v>0</v><w>0</w><vel_type>dirichlet</vel_type><rho_type>neumann</rho_type></bc>
  <periodic_translation><x>0</x><y>0</y><z>0</z></periodic_translation>
</uns_boundary_conditions>
</root>
"""


def load_case(path):
    L = nat.lib()
    L.plbfs_case_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    L.plbfs_case_params.argtypes = [C.c_void_p, C.POINTER(nat.Params)]
    L.plbfs_case_patches.argtypes = [C.c_void_p, C.c_void_p, C.POINTER(nat.Patch), C.c_int32]
    L.plbfs_case_initial_state.argtypes = [C.c_void_p, C.c_void_p] + [C.POINTER(C.c_double)] * 4
    L.plbfs_case_free.argtypes = [C.c_void_p]
    h = C.c_void_p()
    nat.check(L.plbfs_case_load(path.encode(), C.byref(h)))
    return h


@pytest.mark.parametrize("prec", ["false", "true"])
def test_xml_case_matches_python_derivation(tmp_path, prec):
    pm = channel_mesh(10, 5, 5, jitter=0.1)
    prefix = meshgen.write_polymesh(pm, str(tmp_path / "mesh"))
    xml = tmp_path / "case.xml"
    xml.write_text(CASE_XML.format(out=str(tmp_path / "out"), mesh=prefix, prec=prec))
    L = nat.lib()
    h = load_case(str(xml))
    p = nat.Params()
    nat.check(L.plbfs_case_params(h, C.byref(p)))
    ref = pc.channel_case(re=40.0, length_scale=4.0)
    ref.preconditioning = prec == "true"
    rp = ref.params
    assert p.visc == pytest.approx(rp["visc"], rel=1e-15)
    assert p.gamma == pytest.approx(rp["gamma"], rel=1e-15)
    assert p.cfl == 0.5 and p.max_velocity == 0.05 and p.tolerance == 1e-6
    hm = HostMesh.read_openfoam(prefix, p.visc, p.gamma, p.fneq_min)
    patches = (nat.Patch * hm.n_patches)()
    nat.check(L.plbfs_case_patches(h, hm.h, patches, hm.n_patches))
    exp = ref.patches_for(hm.patch_names)
    for q in range(hm.n_patches):
        assert (patches[q].rho_type, patches[q].vel_type) == (exp[q]["rho_type"], exp[q]["vel_type"])
        assert patches[q].u == pytest.approx(exp[q]["u"], rel=1e-15)
    T = hm.n_cells + hm.n_ghosts
    arrs = [np.zeros(T) for _ in range(4)]
    nat.check(L.plbfs_case_initial_state(h, hm.h, *(nat.dptr(a) for a in arrs)))
    g = hm.geometry()
    er = ref.initial_state(g["cx"], g["cy"], g["cz"])
    for a, b in zip(arrs, er):
        assert np.allclose(a, b, rtol=1e-15, atol=0)
    L.plbfs_case_free(h)


def test_unknown_patch_token_is_a_clean_error(tmp_path):
    pm = cavity_mesh(3)
    prefix = meshgen.write_polymesh(pm, str(tmp_path / "mesh"))
    xml = tmp_path / "case.xml"
    xml.write_text(CASE_XML.format(out=str(tmp_path / "out"), mesh=prefix, prec="false"))
    L = nat.lib()
    h = load_case(str(xml))
    hm = HostMesh.read_openfoam(prefix, 0.001)
    patches = (nat.Patch * hm.n_patches)()
    assert L.plbfs_case_patches(h, hm.h, patches, hm.n_patches) == -4   # 'wall'/'lid' have no <bc>
    L.plbfs_case_free(h)


@pytest.mark.parametrize("edit", ["re_text", "re_zero", "deep", "sphere_nodes"])
def test_bad_case_values_are_clean_errors(tmp_path, edit):
    pm = cavity_mesh(3)
    prefix = meshgen.write_polymesh(pm, str(tmp_path / "mesh"))
    text = CASE_XML.format(out=str(tmp_path / "out"), mesh=prefix, prec="false")
    if edit == "re_text":
        text = text.replace("<reynolds_no>40</reynolds_no>", "<reynolds_no>forty</reynolds_no>")
    elif edit == "re_zero":
        text = text.replace("<reynolds_no>40</reynolds_no>", "<reynolds_no>0</reynolds_no>")
    elif edit == "deep":
        text = text.replace("</root>", "<x>" * 100 + "</x>" * 100 + "</root>")
    else:
        text = text.replace("</root>", "<lagrangian_objects><object><name>s</name><num_nodes>1e20</num_nodes>"
                            "<radius>1</radius><depth>1</depth><stiffness>1</stiffness><centre_x>0</centre_x>"
                            "<centre_y>0</centre_y><centre_z>0</centre_z><type>rigid_sphere</type></object>"
                            "</lagrangian_objects></root>")
    xml = tmp_path / "case.xml"
    xml.write_text(text)
    L = nat.lib()
    L.plbfs_case_load.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    h = C.c_void_p()
    rc = L.plbfs_case_load(str(xml).encode(), C.byref(h))
    if edit != "sphere_nodes":
        assert rc == -1 and not h.value
        return
    assert rc == 0
    L.plbfs_case_object_nodes.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_int32,
                                          C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int32)]
    n = C.c_int32(0)
    assert L.plbfs_case_object_nodes(h, 0, None, 0, C.byref(n), None, None, None) == -1
    L.plbfs_case_free.argtypes = [C.c_void_p]
    L.plbfs_case_free(h)


def test_fields_roundtrip_full_precision(tmp_path):
    L = nat.lib()
    L.plbfs_write_fields.argtypes = [C.c_char_p, C.c_int32] + [C.POINTER(C.c_double)] * 4
    L.plbfs_read_fields.argtypes = [C.c_char_p, C.c_int32] + [C.POINTER(C.c_double)] * 4
    rng = np.random.default_rng(0)
    a = [rng.standard_normal(57) for _ in range(4)]
    nat.check(L.plbfs_write_fields(str(tmp_path).encode(), 57, *(nat.dptr(x) for x in a)))
    b = [np.zeros(57) for _ in range(4)]
    nat.check(L.plbfs_read_fields(str(tmp_path).encode(), 57, *(nat.dptr(x) for x in b)))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert open(tmp_path / "rho.txt").readline().startswith("0 ,")


def _tecplot_numbers(text, skip_lines):
    """Whitespace-separated numbers after the first skip_lines lines."""
    return np.array([float(t) for t in " ".join(text.splitlines()[skip_lines:]).split()])


def test_tecplot_grid_and_solution_files(tmp_path):
    """Grid file: node coordinates and polyhedral faces.  Solution file: the
    reference's 15 cell-centred variables (tecplot_output.cpp:347-348) in
    order, p = rho, velocities over max_velocity, centroids, zero vort/stf,
    the given residual and dt, and fneq = 3 nu / (dt_face gamma) of the last
    face a cell owns (Solver.cpp:680), "empty" faces past face NI skipped."""
    nu, gam, U = 0.004, 0.7, 0.05
    pm = meshgen.box_mesh(3, 3, 3, lengths=(3.0, 3.0, 1.5), permute_seed=2,
                          patch_types={"x0": "wall", "x1": "wall", "y0": "wall", "y1": "wall", "z0": "empty",
                                       "z1": "empty"})
    hm = HostMesh.from_polymesh(pm, nu, gam)
    N, F, NI = pm.n_cells, pm.n_faces, pm.n_internal
    L = nat.lib()
    L.plbfs_write_tecplot_grid.argtypes = [C.c_void_p, C.c_char_p]
    L.plbfs_write_tecplot.argtypes = ([C.c_void_p, C.c_char_p] + [C.POINTER(C.c_double)] * 6 + [C.c_double] * 4)
    grid = tmp_path / "grid.dat"
    nat.check(L.plbfs_write_tecplot_grid(hm.h, str(grid).encode()))
    gt = grid.read_text()
    assert "FILETYPE = GRID" in gt and "ELEMENTS=%d FACES=%d" % (N, F) in gt
    P = pm.points.shape[0]
    nums = _tecplot_numbers(gt, 4)
    assert np.array_equal(nums[:3 * P], np.asarray(pm.points, np.float32).T.reshape(-1).astype(np.float64))
    rng = np.random.default_rng(1)
    fields = [1 + 0.01 * rng.standard_normal(N)] + [U * rng.standard_normal(N) for _ in range(3)]
    res = rng.standard_normal(4 * N)
    dtl = rng.uniform(0.1, 0.5, N)
    sol = tmp_path / "7.dat"
    nat.check(L.plbfs_write_tecplot(hm.h, str(sol).encode(), *(nat.dptr(x) for x in fields), nat.dptr(res),
                                    nat.dptr(dtl), U, nu, gam, 7.0))
    st = sol.read_text()
    assert st.splitlines()[2] == ('VARIABLES = "p" "u" "v" "w" "x" "y" "z" "vort" "stf" "res_rho" "res_u" '
                                  '"res_v" "res_w" "dt" "fneq"')
    assert "FILETYPE = SOLUTION" in st and "VARLOCATION=([1-15]=CELLCENTERED)" in st and "SOLUTIONTIME=7" in st
    cols = _tecplot_numbers(st, 4).reshape(15, N)
    geo = hm.geometry()
    assert np.array_equal(cols[0], fields[0])
    for q in (1, 2, 3):
        assert np.array_equal(cols[q], fields[q] / U)
    assert np.array_equal(cols[4], geo["cx"][:N]) and np.array_equal(cols[6], geo["cz"][:N])
    assert not cols[7].any() and not cols[8].any()
    for q in range(4):
        assert np.array_equal(cols[9 + q], res.reshape(N, 4)[:, q])
    assert np.array_equal(cols[13], dtl)
    # faces of the 1 x 1 x 0.5 cells: x and y faces A = 0.5, z faces A = 1;
    # dt_face = 0.5 sqrt(A); the z sides are empty, so their faces past NI
    # do not set fneq
    empty = np.isin(pm.bface_patch, [pm.patch_names.index("empty")])
    want = np.zeros(N)
    for f in range(F):
        if f > NI and empty[f - NI]:
            continue
        want[pm.owner[f]] = nu * 3 / geo["dt"][f] / gam
    assert np.allclose(cols[14], want, rtol=1e-15, atol=0)
    # both face sizes occur, and cells whose owned faces are all empty keep 0
    assert set(np.round(want, 12)) == {round(nu * 3 / (0.5 * np.sqrt(0.5)) / gam, 12),
                                       round(nu * 3 / 0.5 / gam, 12), 0.0}


def test_tecplot_cylinder_zone(tmp_path):
    """A 3 x 2 rigid cylinder as the reference's FEPOLYGON zone
    (tecplot_output.cpp:966-1057): ring edges, axial edges, and the elements
    on their two sides; positions over the domain lengths."""
    L = nat.lib()
    L.plbfs_write_tecplot_nodes.argtypes = ([C.c_char_p, C.c_char_p, C.c_int32] + [C.POINTER(C.c_double)] * 3
                                            + [C.c_double] * 3 + [C.c_int32, C.c_int32, C.c_double])
    x = np.arange(18.0) + 1.0
    vel = -np.arange(18.0)
    path = tmp_path / "cyl.dat"
    nat.check(L.plbfs_write_tecplot_nodes(str(path).encode(), b"cyl", 6, nat.dptr(x), nat.dptr(vel), None,
                                          2.0, 4.0, 8.0, 3, 2, 1.0))
    lines = path.read_text().splitlines()
    assert "ZONETYPE=FEPOLYGON NODES=6 ELEMENTS=3 FACES=9" in lines[2]
    nums = _tecplot_numbers(chr(10).join(lines), 3)
    data = nums[:54].reshape(9, 6)
    X = x.reshape(6, 3)
    assert np.array_equal(data[0], X[:, 0] / 2.0) and np.array_equal(data[2], X[:, 2] / 8.0)
    assert np.array_equal(data[3], vel.reshape(6, 3)[:, 0]) and not data[6:].any()
    rest = nums[54:].astype(int)
    edges = rest[:18].reshape(9, 2).tolist()
    assert edges == [[1, 2], [2, 3], [3, 1], [4, 5], [5, 6], [6, 4], [1, 4], [2, 5], [3, 6]]
    left, right = rest[18:27].tolist(), rest[27:36].tolist()
    assert left == [0, 0, 0, 1, 2, 3, 1, 2, 3] and right == [1, 2, 3, 0, 0, 0, 3, 1, 2]
    # element e is the quad of ring nodes e, e+1 (mod 3) on both rings: every
    # edge lies on the elements it names
    quad = {e: {e, e % 3 + 1, e + 3, e % 3 + 4} for e in (1, 2, 3)}
    for (a, b), l_, r_ in zip(edges, left, right):
        for e in (l_, r_):
            if e:
                assert {a, b} <= quad[e], (a, b, e)


def test_create_without_device_is_a_clean_error():
    if nat.lib().plbfs_device_count() > 0:
        pytest.skip("a device is visible")
    pm = cavity_mesh(3)
    case = pc.cavity_case()
    hm = HostMesh.from_polymesh(pm, case.visc)
    with pytest.raises(nat.PlbfsError) as e:
        GpuSolver(hm, case.params, case.patches_for(pm.patch_names))
    assert e.value.rc == -6


def test_spring_network_object_nodes(tmp_path):
    """A spring_network object's nodes come from its mesh file as written
    (lagrangian_object.cpp:116-213); node area uses the reference's cylinder
    expression with the file's node count; bad connectivity is a clean error."""
    nodes = np.array([[1.0, 2.0, 3.0], [1.5, 2.0, 3.0], [1.0, 2.5, 3.0], [1.0, 2.0, 3.5]])
    mesh = tmp_path / "net.txt"
    lines = ["4 2 2"] + ["%.17g %.17g %.17g" % tuple(p) for p in nodes]
    lines += ["1 2 3", "1 3 4", "1 2 3 4", "2 3 4 1", "2 2 3", "2 1 3", "2 1 2", "1 1"]
    mesh.write_text("\n".join(lines) + "\n")
    obj = ("<lagrangian_objects><object><name>net</name><num_nodes>999</num_nodes><radius>0.5</radius>"
           "<depth>2</depth><stiffness>3</stiffness><centre_x>0</centre_x><centre_y>0</centre_y>"
           "<centre_z>0</centre_z><type>spring_network</type><mesh_file>%s</mesh_file></object>"
           "</lagrangian_objects>\n</root>" % mesh)
    xml = tmp_path / "case.xml"
    xml.write_text(CASE_XML.format(out=str(tmp_path / "out"), mesh="unused/", prec="false").replace("</root>", obj))
    L = nat.lib()
    h = load_case(str(xml))
    L.plbfs_case_object_nodes.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_int32,
                                          C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int32)]
    assert L.plbfs_case_n_objects(h) == 1
    x = np.zeros(12)
    n, k = C.c_int32(0), C.c_int32(0)
    st, area = C.c_double(0), C.c_double(0)
    nat.check(L.plbfs_case_object_nodes(h, 0, nat.dptr(x), 4, C.byref(n), C.byref(st), C.byref(area), C.byref(k)))
    assert n.value == 4 and st.value == 3.0
    assert np.array_equal(x.reshape(4, 3), nodes)
    assert area.value == pytest.approx(2 * np.pi * 0.5 / 4 * 2.0, rel=1e-15)
    L.plbfs_case_free(h)
    mesh.write_text("\n".join(lines).replace("1 2 3 4", "1 2 3 9") + "\n")
    h = load_case(str(xml))
    rc = L.plbfs_case_object_nodes(h, 0, nat.dptr(x), 4, C.byref(n), C.byref(st), C.byref(area), C.byref(k))
    assert rc == -1
    L.plbfs_case_free(h)


def test_warm_start_union_graph_and_draws():
    from plbfs import mpnn
    from plbfs import warmstart as ws
    parts = []
    for tets in (False, True):
        pm = meshgen.box_mesh(4, 3, 2, tets=tets, jitter=0.1, jitter_seed=1, permute_seed=2)
        hm = HostMesh.from_polymesh(pm, 0.01)
        parts.append((hm.geometry(), pm.n_cells))
    g = ws.UnionGraph(parts)
    single = [mpnn.graph_from_geometry(geo, n) for geo, n in parts]
    assert g.N == parts[0][1] + parts[1][1]
    assert len(g.src) == sum(len(s[0]) for s in single)
    e0 = len(single[0][0])
    assert np.array_equal(g.src[:e0], single[0][0]) and np.array_equal(g.src[e0:], single[1][0] + parts[0][1])
    assert np.array_equal(g.dst[e0:], single[1][1] + parts[0][1])
    assert g.src[:e0].max() < parts[0][1] <= g.src[e0:].min()
    assert np.array_equal(g.ef[e0:], single[1][2])
    # every directed edge has its reverse with negated features
    key = {(s, d): k for k, (s, d) in enumerate(zip(g.src, g.dst))}
    for k, (s, d) in enumerate(zip(g.src, g.dst)):
        assert np.array_equal(g.ef[key[(d, s)]], -g.ef[k])
    # sphere sizes scale with the channel height; centres in the upstream third
    for L in ((64.0, 32.0, 32.0), (176.0, 88.0, 88.0)):
        dr = ws.sample_cases(200, L, seed=3, d_frac=(0.14, 0.44))
        D = np.array([d.diameter for d in dr])
        assert D.min() >= 0.14 * L[1] and D.max() <= 0.44 * L[1]
        assert all(0.5 * d.diameter < d.centre[0] <= L[0] / 3 for d in dr)
        assert all(abs(d.centre[1] - 0.5 * L[1]) <= L[1] / 16 for d in dr)


def test_shipped_warm_start_parameters():
    from plbfs import warmstart as ws
    p = ws.load_params()
    assert p is not None, "plbfs/data/mpnn_channel.npy missing"
    assert p.dtype == np.float32 and p.shape == (nat.lib().plbfs_mpnn_n_params(),)
    assert np.all(np.isfinite(p)) and 0 < np.abs(p).max() < 100


def test_warm_start_training_schedule_and_field_filter():
    from plbfs import warmstart as ws

    class FakeNet:
        def __init__(self):
            self.calls = []

        def train(self, X, T, mask, epochs, lr, seed):
            self.calls.append((epochs, lr, seed))
            return np.array([1.0 / (len(self.calls))])

        def params(self):
            return np.zeros(3, np.float32)

    net = FakeNet()
    hist = ws.train_epochs(net, None, None, None, epochs=5, lr=1e-3, seed=7)
    lrs = [c[1] for c in net.calls]
    assert [c[0] for c in net.calls] == [1] * 5 and [c[2] for c in net.calls] == [7, 8, 9, 10, 11]
    assert lrs[0] == pytest.approx(1e-3) and lrs[-1] == pytest.approx(5e-5)
    assert all(a > b for a, b in zip(lrs, lrs[1:]))          # cosine decay, monotone
    assert lrs[2] == pytest.approx(0.5 * (1e-3 + 5e-5))     # half-way point of the cosine
    assert np.allclose(hist, [1, 1 / 2, 1 / 3, 1 / 4, 1 / 5])
    ok = dict(status=0, residual=[0, 2e-4, 1e-4, 1e-4], converged=False)
    assert ws.usable(ok)
    assert not ws.usable(dict(ok, status=-3))
    assert not ws.usable(dict(ok, residual=[0, 5e-3, 0, 0]))


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_program_reads_the_case_up_to_the_device(tmp_path):
    # the plbfs executable (src/main.cpp entry) runs the XML case through the
    # reader, the patch table and the initial conditions; without a device it
    # stops at plbfs_create with PLBFS_ERR_NODEVICE, and a bad case stops earlier
    if nat.lib().plbfs_device_count() > 0:
        pytest.skip("a device is visible")
    import subprocess
    exe = os.path.join(ROOT, "3d-plfbs-gpu_amd", "build", "plbfs")
    pm = channel_mesh(10, 5, 5, jitter=0.1)
    prefix = meshgen.write_polymesh(pm, str(tmp_path / "mesh"))
    xml = tmp_path / "case.xml"
    xml.write_text(CASE_XML.format(out=str(tmp_path / "out"), mesh=prefix, prec="false"))
    r = subprocess.run([exe, str(xml)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "error -6" in r.stderr, r.stderr
    bad = tmp_path / "bad.xml"
    bad.write_text(CASE_XML.format(out=str(tmp_path / "out"), mesh=str(tmp_path / "nomesh") + "/", prec="false"))
    r = subprocess.run([exe, str(bad)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "error -1" in r.stderr and "owner" in r.stderr, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr


def test_wrapper_rejects_short_state_and_node_arrays():
    # the C-ABI reads n values per pointer: the Python wrappers refuse shorter inputs
    from plbfs.solver import node_arrays, state_arrays
    full = [np.arange(12.0)] * 4
    out = state_arrays(full, 10)
    assert all(a.shape == (10,) and a.flags.c_contiguous for a in out)
    with pytest.raises(ValueError):
        state_arrays([np.arange(12.0), np.arange(9.0), np.arange(12.0), np.arange(12.0)], 10)
    xr, xx = node_arrays(np.zeros((4, 3)), np.ones((4, 3)))
    assert xr.shape == xx.shape == (12,)
    with pytest.raises(ValueError):
        node_arrays(np.zeros((4, 3)), np.zeros((3, 3)))
    with pytest.raises(ValueError):
        node_arrays(np.zeros(10), np.zeros(10))


C_CONSUMER = r"""
#include <stdio.h>
#include "plbfs_hip.h"
#include "plbfs_case.h"
#include "plbfs_mpnn.h"
int main(int argc, char **argv) {
    plbfs_mesh *m = NULL;
    int64_t info[6];
    plbfs_mesh_view v;
    if (argc < 2) return 2;
    if (plbfs_mesh_read_openfoam(argv[1], 0.001, 1.0, 0.0, &m) != PLBFS_OK) {
        printf("error %s\n", plbfs_last_error());
        return 1;
    }
    plbfs_mesh_info(m, info);
    plbfs_mesh_view_get(m, &v);
    printf("%lld %lld %d %d\n", (long long)info[0], (long long)v.n_faces, plbfs_mpnn_n_params(),
           plbfs_device_count() >= 0);
    plbfs_mesh_free(m);
    return 0;
}
"""


def test_headers_are_plain_c(tmp_path):
    # what a cgo / ctypes / JNI binding sees: the three headers compile as
    # pedantic C99 without warnings, and a C program links the library and
    # reads a mesh through it
    import subprocess
    src = tmp_path / "consumer.c"
    src.write_text(C_CONSUMER)
    exe = tmp_path / "consumer"
    build = os.path.join(ROOT, "3d-plfbs-gpu_amd", "build")
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
                        "-I" + os.path.join(ROOT, "include"), str(src), "-L" + build, "-lplbfs",
                        "-Wl,-rpath," + build, "-o", str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    pm = cavity_mesh(3)
    prefix = meshgen.write_polymesh(pm, str(tmp_path / "mesh"))
    r = subprocess.run([str(exe), prefix], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    cells, faces, n_params, ok = r.stdout.split()
    assert (int(cells), int(faces), int(ok)) == (pm.n_cells, pm.n_faces, 1)
    assert int(n_params) > 0


@pytest.mark.parametrize("num_nodes", [16, 20])
def test_rigid_cylinder_object_nodes(tmp_path, num_nodes):
    """A rigid_cylinder object (lagrangian_object.cpp:79-113): (int)sqrt(n)
    rings of (int)sqrt(n) nodes, ring j at z = cz + depth/2 - j depth/rings -
    depth/(2 rings), node i at angle 2 pi i/rings from +y towards +x (x = cx +
    r sin, y = cy + r cos); node area 2 pi r / n * depth with the file's n
    (immersed_boundary_method.cu:168).  With n = 20 the reference fills 16 of
    its 20 node slots; the loader gives the 16 it fills."""
    r, depth, cx, cy, cz = 2.0, 4.0, 5.0, 4.0, 3.0
    obj = ("<lagrangian_objects><object><name>cyl</name><num_nodes>%d</num_nodes><radius>%g</radius>"
           "<depth>%g</depth><stiffness>3</stiffness><centre_x>%g</centre_x><centre_y>%g</centre_y>"
           "<centre_z>%g</centre_z><type>rigid_cylinder</type></object>"
           "</lagrangian_objects>\n</root>" % (num_nodes, r, depth, cx, cy, cz))
    xml = tmp_path / "case.xml"
    xml.write_text(CASE_XML.format(out=str(tmp_path / "out"), mesh="unused/", prec="false").replace("</root>", obj))
    L = nat.lib()
    h = load_case(str(xml))
    L.plbfs_case_object_nodes.argtypes = [C.c_void_p, C.c_int32, C.POINTER(C.c_double), C.c_int32,
                                          C.POINTER(C.c_int32), C.POINTER(C.c_double), C.POINTER(C.c_double),
                                          C.POINTER(C.c_int32)]
    x = np.zeros(3 * num_nodes)
    n, k = C.c_int32(0), C.c_int32(0)
    st, area = C.c_double(0), C.c_double(0)
    nat.check(L.plbfs_case_object_nodes(h, 0, nat.dptr(x), num_nodes, C.byref(n), C.byref(st), C.byref(area),
                                        C.byref(k)))
    L.plbfs_case_free(h)
    assert n.value == 16 and k.value == 0  # 4 rings of 4, the 2-point hat kernel
    X = x[: 3 * 16].reshape(16, 3)
    # closed form: rings at z = 3 + 2 - j - 0.5, quarter turns from +y
    zs = [4.5, 3.5, 2.5, 1.5]
    quarter = [(0.0, 1.0), (1.0, 0.0), (0.0, -1.0), (-1.0, 0.0)]
    for j in range(4):
        for i in range(4):
            sx, sy = quarter[i]
            assert np.allclose(X[4 * j + i], (cx + r * sx, cy + r * sy, zs[j]), rtol=0, atol=1e-14), (j, i)
    assert area.value == pytest.approx(2 * np.pi * r / num_nodes * depth, rel=1e-15)
    from plbfs import objects
    Xp, dAp = objects.rigid_cylinder(num_nodes, (cx, cy, cz), r, depth)
    assert np.array_equal(Xp, X) and dAp == area.value
