"""Failure paths of the mesh front end, on the CPU (each in a child
process, since they abort as the reference's asserts do,
src/mesh/mesh.c:194-195,484,545,652-654): a negatively oriented element,
a face shared by three elements, an element naming a missing node and a
truncated file each stop with a message naming the cause."""
import os
import subprocess
import sys
import textwrap

import gdon3d as G

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def base_mesh(tmp_path):
    path = str(tmp_path / "k2.msh")
    assert G.load_library().mesh_kuhn_cube_write_msh22(path.encode(), 2, 0.0, 0) == 0
    with open(path) as f:
        return f.read().split("\n")


def elements_range(lines):
    i = lines.index("$Elements")
    n = int(lines[i + 1])
    return i + 2, i + 2 + n


def run_mesh(path, hmin=True):
    code = textwrap.dedent(f"""
        import sys, ctypes as C
        sys.path.insert(0, {os.path.join(ROOT, 'dg-cfl-less-maxwell-solver_amd')!r})
        import gdon3d as G
        lib = G.load_library()
        m = G.GdnMesh()
        lib.mesh_read_from_msh22_file(C.byref(m), {path!r}.encode())
        lib.mesh_build_connectivity(C.byref(m))
        if {hmin!r}:
            lib.mesh_get_t4_hmin(C.byref(m))
        print("returned")
    """)
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True,
                          timeout=120)


def write(tmp_path, name, lines):
    p = str(tmp_path / name)
    with open(p, "w") as f:
        f.write("\n".join(lines))
    return p


def test_valid_mesh_passes(tmp_path):
    r = run_mesh(write(tmp_path, "ok.msh", base_mesh(tmp_path)))
    assert r.returncode == 0 and "returned" in r.stdout


def test_negatively_oriented_element(tmp_path):
    lines = base_mesh(tmp_path)
    a, _ = elements_range(lines)
    f = lines[a + 3].split()  # "id 4 2 1 1 v0 v1 v2 v3": swap v2 and v3
    f[-1], f[-2] = f[-2], f[-1]
    lines[a + 3] = " ".join(f)
    r = run_mesh(write(tmp_path, "inv.msh", lines))
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "negatively" in r.stderr


def test_face_shared_by_three_elements(tmp_path):
    lines = base_mesh(tmp_path)
    a, b = elements_range(lines)
    n = b - a
    dup = lines[a + 5].split()
    dup[0] = str(n + 1)
    lines[a - 1] = str(n + 1)
    lines.insert(b, " ".join(dup))
    r = run_mesh(write(tmp_path, "dup.msh", lines), hmin=False)
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "more than two elements" in r.stderr


def test_missing_node(tmp_path):
    lines = base_mesh(tmp_path)
    a, _ = elements_range(lines)
    f = lines[a].split()
    f[-1] = "100000"
    lines[a] = " ".join(f)
    r = run_mesh(write(tmp_path, "miss.msh", lines), hmin=False)
    assert r.returncode != 0 and "missing node" in r.stderr


def test_truncated_file(tmp_path):
    lines = base_mesh(tmp_path)
    a, b = elements_range(lines)
    r = run_mesh(write(tmp_path, "trunc.msh", lines[: a + 3]), hmin=False)
    assert r.returncode != 0 and "returned" not in r.stdout
    assert "gdon3d mesh:" in r.stderr
