"""OBJ / MTL loading (ilm_obj_*) against hand-derived expectations that
follow the reference loader's rules (Scene/WavefrontSceneLoader.cpp)."""
import os
import tempfile

import numpy as np
import pytest

OBJ = """# test scene
mtllib scene.mtl
o thing
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 2 2 2
vn 0 0 1
vn 0 1 0
vt 0 0
vt 1 0
usemtl white
f 1//1 2//1 3//1
f 1//1 3//1 4//1 5//1
usemtl lamp
f 1/1/2 2/2/2 3/1/2
f 1 2 3 4 5
f 1/2 2/2 3
usemtl nosuch
f 2 3 6
"""

MTL = """newmtl white
Kd 0.7 0.7 0.7
newmtl lamp
Kd 0.1 0.2 0.3
Ks 0.5 0.5 0.5
Ns 20
Ke 4 5 6
"""


@pytest.fixture()
def obj_file():
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "scene.obj"), "w") as f:
            f.write(OBJ)
        with open(os.path.join(d, "scene.mtl"), "w") as f:
            f.write(MTL)
        yield os.path.join(d, "scene.obj")


def test_obj_faces_vertices_and_materials(ilm, obj_file):
    pos, nrm, idx, tm, tl, mats, lights = ilm.load_obj(obj_file)
    # faces: tri (white), quad -> 2 tris (white), tri (lamp), the pentagon is
    # skipped, "p/t" corners read as position/normal plus a one-field corner
    # that keeps the previous normal (lamp), tri with an unknown material
    # (material 0)
    assert len(idx) == 6
    assert tm.tolist() == [0, 0, 0, 1, 1, 0]
    assert tl.tolist() == [-1, -1, -1, 0, 0, -1]
    # vertex records: (1,0,1) (2,0,1) (3,0,1) (4,0,1) (5,0,1) (1,1,2) (2,2,2)
    # (3,1,2) (1,0,2) (2,0,2) (3,0,2) (2,0,0) (3,0,0) (6,0,0)
    keys = [(1, 0, 1), (2, 0, 1), (3, 0, 1), (4, 0, 1), (5, 0, 1), (1, 1, 2), (2, 2, 2), (3, 1, 2),
            (1, 0, 2), (2, 0, 2), (3, 0, 2), (2, 0, 0), (3, 0, 0), (6, 0, 0)]
    assert len(pos) == len(keys)
    P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [2, 2, 2]], np.float32)
    assert np.array_equal(pos, P[[k[0] - 1 for k in keys]])
    assert idx.tolist() == [[0, 1, 2], [0, 2, 3], [0, 3, 4], [5, 6, 7], [8, 9, 10], [11, 12, 13]]
    # explicit normals are kept; vertices without one get the normal of their triangle
    assert np.array_equal(nrm[:5], np.tile([0, 0, 1], (5, 1)).astype(np.float32))
    assert np.array_equal(nrm[5:11], np.tile([0, 1, 0], (6, 1)).astype(np.float32))
    n = np.cross(P[2] - P[1], P[5] - P[1])
    n /= np.linalg.norm(n)
    assert np.allclose(nrm[11:], np.tile(n, (3, 1)), atol=1e-6)
    assert np.allclose(mats, [[0.7, 0.7, 0.7, 0, 0, 0, 0, 0], [0.1, 0.2, 0.3, 0.5, 0.5, 0.5, 20, 1]])
    assert np.allclose(lights, [[4, 5, 6, 0]])


def test_obj_errors(ilm, obj_file):
    with pytest.raises(ilm.IlmError):
        ilm.load_obj(obj_file + ".missing")
    bad = obj_file.replace("scene.obj", "bad.obj")
    with open(bad, "w") as f:
        f.write("v 0 0 0\nv 1 0 0\nf 1 2 7\n")
    with pytest.raises(ilm.IlmError):
        ilm.load_obj(bad)
