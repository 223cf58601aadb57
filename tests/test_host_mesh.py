"""CPU tests of the cloth mesh host side: include/mpm/LagrangianMesh.h (built with g++ here, no GPU
calls) and the Python OBJ reader/writer of the package, against each other and against fixed
meshes whose parse is known by hand.

  OBJ reader   LagrangianMesh::ObjMesh (LagrangianMesh.cpp:198-280): `v` and `f` lines, 1-based
               indices; extended here to `a/b/c` corner forms, negative indices and polygon fans
  writer       the mesh frame file of HybridSolver.cpp:1006-1030
"""
import os
import subprocess

import numpy as np
import pytest

from anisotropicelastoplasticity_amd import api, scenes

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# a quad (fan split into 2 triangles), a triangle with a/b/c corners, one with negative indices,
# comments, normals and texture lines that the reader skips
HAND_OBJ = """# hand-written
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
vn 0 0 1
vt 0.5 0.5
f 1 2 3 4
v 2 0 0.5
f 2/1/1 5/1/1 3//1
f -1 -4 -5
o name
"""
HAND_X = np.array([[0, 1, 1, 0, 2], [0, 0, 1, 1, 0], [0, 0, 0, 0, 0.5]], np.float32)
HAND_F = np.array([[0, 1, 2], [0, 2, 3], [1, 4, 2], [4, 1, 0]], np.int32)


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    d = tmp_path_factory.mktemp("meshio")
    exe = str(d / "mesh_io_check")
    res = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall", "-Werror", os.path.join(HERE, "host", "mesh_io_check.cpp"),
                          "-o", exe], capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    return exe, d


def _run(checker, path):
    exe, d = checker
    res = subprocess.run([exe, str(d), str(path)], capture_output=True, text=True, timeout=60)
    assert res.returncode == 0, res.stderr
    lines = res.stdout.split("\n")
    nv, nf = map(int, lines[0].split())
    x = np.array([list(map(float, l.split())) for l in lines[1:1 + nv]], np.float32).T
    f = np.array([list(map(int, l.split())) for l in lines[1 + nv:1 + nv + nf]], np.int32).reshape(-1, 3)
    return x, f


def test_cpp_reader_hand_mesh(checker, tmp_path):
    p = tmp_path / "hand.obj"
    p.write_text(HAND_OBJ)
    x, f = _run(checker, p)
    np.testing.assert_array_equal(x, HAND_X)
    np.testing.assert_array_equal(f, HAND_F)


def test_python_reader_hand_mesh(tmp_path):
    p = tmp_path / "hand.obj"
    p.write_text(HAND_OBJ)
    x, f = api.read_mesh_obj(str(p))
    np.testing.assert_array_equal(x, HAND_X)
    np.testing.assert_array_equal(f, HAND_F)


def test_writer_reader_round_trip_both_sides(checker, tmp_path):
    x, faces = scenes.grid_mesh(7, 5, 0.125, (0.25, 0.5), 0.75)  # exact binary fractions
    p = tmp_path / "grid.obj"
    api.write_mesh_obj(str(p), x, faces)
    lines = p.read_text().splitlines()
    assert len(lines) == 35 + faces.shape[0]
    assert lines[0].startswith("v ") and lines[35].startswith("f ")
    assert lines[35].split()[1:] == [str(int(k) + 1) for k in faces[0]]
    xp, fp = api.read_mesh_obj(str(p))
    np.testing.assert_array_equal(xp, x)
    np.testing.assert_array_equal(fp, faces)
    xc, fc = _run(checker, p)
    np.testing.assert_array_equal(xc, x)
    np.testing.assert_array_equal(fc, faces)


@pytest.mark.parametrize("text", ["v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 0\n", "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 4\n",
                                  "v 0 0 0\nv 1 0 0\nf 1 2\n"])
def test_python_reader_rejects_bad_faces(tmp_path, text):
    p = tmp_path / "bad.obj"
    p.write_text(text)
    with pytest.raises(api.MPMError):
        api.read_mesh_obj(str(p))
