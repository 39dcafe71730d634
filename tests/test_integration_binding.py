"""The reference-side binding printed in INTEGRATION.md compiles against the
reference's member names (CPU only).

The block is compiled with g++ -fsyntax-only against tests/ref_decls/Solver_gpu.h,
which declares only the reference members the binding touches.  Each of those
declarations cites the reference header line it restates; where the reference
tree is present, every declared name must appear on the cited line.  A binding
that uses a member the reference does not have (the accessor
get_pressure_gradient, say) must fail to compile.
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DECLS = os.path.join(ROOT, "tests", "ref_decls")
REF_SRC = "/root/reference/src"

needs_gxx = pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")


def binding_source():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    sec = text[text.index("## Reference-side binding (C++)"):]
    m = re.search(r"```cpp\n(.*?)```", sec, re.S)
    assert m, "no C++ block in the reference-side binding section"
    return m.group(1)


def compile_binding(src, tmp_path):
    path = tmp_path / "gpu_solver_plbfs.cpp"
    path.write_text(src)
    return subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", "-I", DECLS,
                           "-I", os.path.join(ROOT, "include"), str(path)], capture_output=True, text=True)


@needs_gxx
def test_binding_compiles_against_the_reference_members(tmp_path):
    r = compile_binding(binding_source(), tmp_path)
    assert r.returncode == 0, r.stderr


@needs_gxx
@pytest.mark.parametrize("good,bad", [
    ("-init_conds.pressure_gradient", "-source.get_pressure_gradient()"),
    ("time_mode(globals.time_stepping)", "time_mode(globals.gpu_time_stepping)"),
    ("globals.ref_rho", "soln.get_ref_rho()"),
])
def test_a_member_the_reference_lacks_does_not_compile(tmp_path, good, bad):
    src = binding_source()
    assert good in src
    r = compile_binding(src.replace(good, bad), tmp_path)
    assert r.returncode != 0


def test_binding_maps_what_the_case_loader_maps():
    src = binding_source()
    # time stepping from the string the CPU solver reads (Solver.cpp:1752-1773)
    assert 'if (ts == "local") return PLBFS_DT_LOCAL;' in src
    assert 'if (ts == "min") return PLBFS_DT_MIN;' in src
    assert 'if (ts == "talts") return PLBFS_DT_TALTS;' in src
    assert "return PLBFS_DT_CONSTANT;" in src
    assert "gpu_time_stepping" not in src
    # IBM kernel width delta_x = 2 dt (Solver_gpu.cu:517), force -pressure_gradient (Solver.cpp:213)
    assert "2 * domain.dt," in src
    assert "prm.force = -init_conds.pressure_gradient;" in src
    loader = open(os.path.join(ROOT, "3d-plfbs-gpu_amd", "csrc", "host", "case.cpp")).read()
    assert 'p.force = -c->n("initial_conditions/pressure_gradient", 0.0);' in loader
    assert 'ts == "local" ? PLBFS_DT_LOCAL' in loader and 'ts == "min" ? PLBFS_DT_MIN' in loader
    assert 'ts == "talts" ? PLBFS_DT_TALTS' in loader


KEYWORDS = {"double", "int", "void", "const", "std", "string", "vector", "class", "public", "return", "arg", "i"}


def declared_names(decl):
    decl = re.sub(r"=\s*[-0-9.]+", "", decl)
    if "(" in decl:
        return [re.findall(r"(\w+)\s*\(", decl)[0]]
    return [w for w in re.findall(r"[A-Za-z_]\w*", decl) if w not in KEYWORDS]


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="reference tree not present")
def test_each_declaration_is_on_the_reference_line_it_cites():
    checked = 0
    for line in open(os.path.join(DECLS, "Solver_gpu.h")):
        m = re.match(r"\s*(.*?;)\s*/\*\s*(\w+\.h):(\d+)(?:-(\d+))?\s*\*/", line)
        if not m:
            continue
        decl, hdr, a, b = m.group(1), m.group(2), int(m.group(3)), int(m.group(4) or m.group(3))
        ref = open(os.path.join(REF_SRC, hdr), errors="replace").read().splitlines()[a - 1:b]
        for name in declared_names(decl):
            assert any(re.search(r"\b%s\b" % re.escape(name), r) for r in ref), (name, hdr, a)
            checked += 1
    assert checked >= 60, checked
