"""Scene scripts (.ilm) through the plug-in's EnvironmentLoader
(illuminaprt_amd/plugin/ilm_environment.cpp), run by the sandbox program.

The CPU tests load scripts on the host only and check the resolved job
against expectations worked out here: the box room of the plain
path-tracing workload written as a script (wall and light Quads with the
reference's tessellation, Quad.cpp:11-29; the two boxes as one Wavefront
model placed twice with Rotation / Scaling / Translation in the reference's
Matrix3x3 convention), and the lexer's and loader's rules
(System/Lexer.cpp:27-114, Scene/EnvironmentLoader.cpp:163-912). The GPU
tests render the script and compare it with the Python binding on the same
arrays (identical) and with the box-room workload (same expectation)."""
import os
import subprocess

import numpy as np
import pytest

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "illuminaprt_amd")
PLUGIN = os.path.join(PKG, "libilm_plugin.so")
SANDBOX = os.path.join(PKG, "ilm_plugin_sandbox")

CUBE_OBJ = """mtllib cube.mtl
v 0 0 0
v 1 0 0
v 1 1 0
v 0 1 0
v 0 0 1
v 1 0 1
v 1 1 1
v 0 1 1
vn -1 0 0
vn 1 0 0
vn 0 -1 0
vn 0 1 0
vn 0 0 -1
vn 0 0 1
usemtl white
f 1//1 5//1 8//1 4//1
f 2//2 3//2 7//2 6//2
f 1//3 2//3 6//3 5//3
f 4//4 8//4 7//4 3//4
f 1//5 4//5 3//5 2//5
f 5//6 6//6 7//6 8//6
"""

CUBE_MTL = """newmtl white
Kd 0.73 0.73 0.73
"""

# Matrix3x3 elements in storage order; Product gives out_i = sum_j m[3j+i] v_j,
# so this maps (x, y, z) to (z, y, -x): a quarter turn about y.
ROT_Y90 = [0, 0, -1, 0, 1, 0, 1, 0, 0]


def _run(*args, timeout=120, cwd=None):
    if not (os.path.exists(PLUGIN) and os.path.exists(SANDBOX)):
        pytest.fail("plug-in not built: run __graft_entry__.build()")
    r = subprocess.run([SANDBOX, PLUGIN, *args], capture_output=True, text=True, timeout=timeout, cwd=cwd)
    print(r.stdout, r.stderr, flush=True)
    return r


def _f(v):
    return repr(float(np.float32(v)))


def _vec(v):
    return "{" + ", ".join(_f(x) for x in v) + "}"


def room_script(d, width=64, height=64, samples=16, fmt="PFM", depth=5):
    """The box room as a script; returns its path. Walls and the light are
    Quads with Vertices (p0, p3, p2, p1) so that the reference's
    tessellation gives the workload's own triangles (p0, p2, p3), (p0, p1, p2)."""
    from illuminaprt_amd import workloads as w
    pos, nrm, idx, tm, tl, mats, lights = w.box_room()
    with open(os.path.join(d, "cube.obj"), "w") as f:
        f.write(CUBE_OBJ)
    with open(os.path.join(d, "cube.mtl"), "w") as f:
        f.write(CUBE_MTL)
    names = ["white", "red", "green", "emitter"]
    shapes, prims = [], []
    for q in range(len(idx) // 2):
        b = idx[2 * q][0]
        p = pos[b:b + 4]
        m, l = int(tm[2 * q]), int(tl[2 * q])
        if 8 <= q < 20:
            continue  # the boxes come from the Wavefront model
        shapes.append(f"  Shape {{\n    Id = quad{q}\n    Type = Quad\n"
                      f"    Vertices = {{{_vec(p[0])}, {_vec(p[3])}, {_vec(p[2])}, {_vec(p[1])}}}\n  }}\n")
        if l >= 0:
            prims.append(f"      Primitive {{\n        Id = prim{q}\n        Type = Emissive\n        Shape = quad{q}\n"
                         f"        Material = {names[m]}\n        Light = ceiling\n      }}\n")
        else:
            prims.append(f"      Primitive {{\n        Id = prim{q}\n        Type = Geometry\n        Shape = quad{q}\n"
                         f"        Material = {names[m]}\n      }}\n")
    box_prims = (
        "      Primitive {\n        Id = box_a\n        Type = Geometry\n        Shape = cube\n"
        "        Material = cube_materials\n"
        f"        Rotation = {_vec(ROT_Y90)}\n        Scaling = {{0.3, 0.3, 0.3}}\n"
        "        Translation = {0.13, 0, 0.45}\n      }\n"
        "      Primitive {\n        Id = box_b\n        Type = Geometry\n        Shape = cube\n"
        "        Material = cube_materials\n"
        "        Scaling = {0.3, 0.6, 0.3}\n        Translation = {0.55, 0, 0.45}\n      }\n")
    # the light quad last, after the boxes, as in the workload
    light_prim = prims.pop()
    script = (
        "# box room\n"
        "Cameras {\n  Camera {\n    Id = cam\n    Type = Perspective\n    Position = {0.5, 0.5, -1.4}\n"
        "    Target = {0.5, 0.5, 0.5}\n    Up = {0, 1, 0}\n    Fov = 40\n    Aspect = 1\n  }\n}\n"
        "Shapes {\n" + "".join(shapes) +
        "  Shape {\n    Id = cube\n    Type = WavefrontModel\n    Filename = cube.obj\n    Container = BVHMesh\n"
        "    MaterialGroupId = cube_materials\n  }\n}\n"
        "Materials {\n" + "".join(
            f"  Material {{\n    Id = {n}\n    Type = Matte\n    Reflectivity = {_vec(mats[k][:3])}\n  }}\n"
            for k, n in enumerate(names)) + "}\n"
        f"Lights {{\n  Light {{\n    Id = ceiling\n    Type = DiffuseArea\n    Power = {_vec(lights[0][:3])}\n"
        "    Shape = quad20\n  }\n}\n"
        "Filters {\n  Filter {\n    Id = box\n    Type = Box\n  }\n}\n"
        "Samplers {\n  Sampler {\n    Id = rnd\n    Type = Random\n  }\n}\n"
        f"Devices {{\n  Device {{\n    Id = img\n    Type = Image\n    Width = {width}\n    Height = {height}\n"
        f"    Format = {fmt}\n    Filename = room.{fmt.lower()}\n  }}\n}}\n"
        f"Integrators {{\n  Integrator {{\n    Id = pt\n    Type = PathTracing\n    RayDepth = {depth}\n"
        "    ShadowRays = 1\n    Epsilon = 0.0001\n  }\n}\n"
        f"Renderers {{\n  Renderer {{\n    Id = basic\n    Type = Basic\n    Samples = {samples}\n"
        "    Integrator = pt\n    Device = img\n    Filter = box\n  }\n}\n"
        "Environment {\n  Renderer = basic\n  Scene {\n    Sampler = rnd\n    Camera = cam\n    Lights = {ceiling}\n"
        "    Space {\n      Id = space\n      Type = Basic\n      Primitives {\n" + "".join(prims) + box_prims +
        light_prim + "      }\n    }\n  }\n}\n")
    path = os.path.join(d, "room.ilm")
    with open(path, "w") as f:
        f.write(script)
    return path


def read_job(path):
    with open(path, "rb") as f:
        data = f.read()
    o = 0

    def take(fmt, n):
        nonlocal o
        dt = np.dtype(fmt)
        a = np.frombuffer(data[o:o + dt.itemsize * int(n)], dtype=dt)
        o += dt.itemsize * int(n)
        return a

    nv, nt, nmat, nl = take("<i8", 4)
    W, H, spp, depth, builder = take("<i4", 5)
    eps = float(take("<f4", 1)[0])
    cam = take("<f4", 11)
    job = dict(W=int(W), H=int(H), spp=int(spp), depth=int(depth), builder=int(builder), eps=eps, cam=cam,
               pos=take("<f4", 3 * nv).reshape(-1, 3), idx=take("<i4", 3 * nt).reshape(-1, 3),
               nrm=take("<f4", 3 * nv).reshape(-1, 3), tri_mat=take("<i4", nt), tri_light=take("<i4", nt),
               mats=take("<f4", 8 * nmat).reshape(-1, 8), lights=take("<f4", 4 * nl).reshape(-1, 4))
    npr = int(take("<i8", 1)[0])
    job["first"] = take("<i8", npr + 1)
    assert o == len(data)
    return job


def load(tmp_path, script_text=None, script_path=None):
    if script_path is None:
        script_path = os.path.join(tmp_path, "s.ilm")
        with open(script_path, "w") as f:
            f.write(script_text)
    out = os.path.join(tmp_path, "job.bin")
    r = _run("ilm-load", script_path, out)
    return r, (read_job(out) if r.returncode == 0 else None)


def test_room_script_resolves_to_the_workload(tmp_path, ilm):
    from illuminaprt_amd import workloads as w
    r, job = load(tmp_path, script_path=room_script(str(tmp_path)))
    assert r.returncode == 0 and "warning" not in r.stdout
    pos, nrm, idx, tm, tl, mats, lights = w.box_room()
    cam = w.room_camera()
    assert (job["W"], job["H"], job["spp"], job["depth"]) == (64, 64, 16, 5)
    assert job["eps"] == np.float32(1e-4) and job["builder"] == ilm.ILM_BUILDER_SAH
    np.testing.assert_array_equal(job["cam"], np.array(cam.as_list(), np.float32))
    np.testing.assert_array_equal(job["lights"], lights)
    # primitives: 8 wall quads, the two boxes (12 triangles each), the light quad
    assert job["first"].tolist() == [2 * k for k in range(9)] + [28, 40, 42]
    P = job["pos"][job["idx"]]               # (nt, 3, 3) world triangles
    want = pos[idx]
    wall_q = list(range(8)) + [20]
    got_walls = np.concatenate([P[:16], P[40:]])
    want_walls = np.concatenate([want[[2 * q + s for q in wall_q for s in (0, 1)]]])
    # the reference's Quad tessellation of (p0, p3, p2, p1): (p0, p2, p3), (p0, p1, p2)
    want_swapped = want_walls.reshape(-1, 2, 3, 3)[:, ::-1].reshape(-1, 3, 3)
    np.testing.assert_array_equal(got_walls, want_swapped)
    # boxes: the cube placed by R (S p) + T covers the workload's box faces
    # (same planes, same areas, same area-weighted centroids per box)
    def area_centroid(tris):
        a = 0.5 * np.linalg.norm(np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0]), axis=1)
        return a.sum(), (a[:, None] * tris.mean(axis=1)).sum(axis=0) / a.sum()
    for k in range(2):
        ga, gc = area_centroid(P[16 + 12 * k:28 + 12 * k].astype(np.float64))
        wa, wc = area_centroid(want[16 + 12 * k:28 + 12 * k].astype(np.float64))
        assert abs(ga - wa) < 1e-6 and np.abs(gc - wc).max() < 1e-6
        lo, hi = want[16 + 12 * k:28 + 12 * k].reshape(-1, 3).min(0), want[16 + 12 * k:28 + 12 * k].reshape(-1, 3).max(0)
        assert np.abs(P[16 + 12 * k:28 + 12 * k].reshape(-1, 3).min(0) - lo).max() < 1e-6
        assert np.abs(P[16 + 12 * k:28 + 12 * k].reshape(-1, 3).max(0) - hi).max() < 1e-6
    # box normals: the OBJ's face normals turned by the rotation (outward)
    ctr = np.array([[0.28, 0.15, 0.30], [0.70, 0.30, 0.60]])
    for k in range(2):
        t = slice(16 + 12 * k, 28 + 12 * k)
        n = job["nrm"][job["idx"][t]].mean(axis=1)
        assert (np.einsum("ij,ij->i", n, P[t].mean(axis=1) - ctr[k]) > 0).all()
    # walls keep their workload normals' direction
    gn = job["nrm"][job["idx"][:16]].mean(axis=1)
    wn = nrm[idx[:16]].mean(axis=1)
    wn = wn.reshape(-1, 2, 3)[:, ::-1].reshape(-1, 3)
    np.testing.assert_allclose(gn, wn, atol=1e-6)
    # materials: first use order (white, red, green, the cube group, emitter)
    M = job["mats"][job["tri_mat"]][:, :3]
    np.testing.assert_allclose(M, np.concatenate([mats[tm[:16]], mats[tm[16:40]], mats[tm[40:]]])[:, :3], atol=1e-7)
    assert (job["tri_light"] == np.where(np.arange(42) >= 40, 0, -1)).all()


def _env(body, renderer="basic"):
    return (
        "Cameras {\n  Camera {\n    Id = cam\n    Type = Perspective\n  }\n}\n"
        "Shapes {\n  Shape {\n    Id = tri\n    Type = Triangle\n    P0 = {0, 0, 1}\n    P1 = {1, 0, 1}\n"
        "    P2 = {0, 1, 1}\n  }\n  Shape {\n    Id = ball\n    Type = Sphere\n    Centre = {0, 0, 0}\n"
        "    Radius = 1\n  }\n}\n"
        "Materials {\n  Material {\n    Id = m\n    Type = Matte\n    Reflectivity = {0.5, 0.5, 0.5}\n  }\n"
        "  Material {\n    Id = chrome\n    Type = Mirror\n    Reflectivity = {1, 1, 1}\n  }\n}\n"
        "Lights {\n  Light {\n    Id = l\n    Type = DiffuseArea\n    Power = {1, 2, 3}\n  }\n"
        "  Light {\n    Id = bulb\n    Type = Point\n  }\n}\n"
        "Samplers {\n  Sampler {\n    Id = s\n    Type = Jitter\n  }\n}\n"
        "Devices {\n  Device {\n    Id = d\n    Type = Image\n    Width = 8\n    Height = 4\n  }\n}\n"
        "Integrators {\n  Integrator {\n    Id = pt\n    Type = PathTracing\n  }\n"
        "  Integrator {\n    Id = ic\n    Type = IC\n  }\n}\n"
        "Renderers {\n  Renderer {\n    Id = basic\n    Type = Basic\n    Integrator = pt\n    Device = d\n  }\n"
        "  Renderer {\n    Id = icr\n    Type = Basic\n    Integrator = ic\n    Device = d\n  }\n}\n"
        f"Environment {{\n  Renderer = {renderer}\n  Scene {{\n    Sampler = s\n    Camera = cam\n" + body + "  }\n}\n")


SPACE_ONE = ("    Lights = {l}\n    Space {\n      Id = sp\n      Type = Basic\n      Primitives {\n"
             "        Primitive {\n          Id = p\n          Type = Emissive\n          Shape = tri\n"
             "          Material = m\n          Light = l\n        }\n      }\n    }\n")


def test_defaults_and_unused_unsupported_entries(tmp_path, ilm):
    r, job = load(tmp_path, _env(SPACE_ONE))
    assert r.returncode == 0, r.stderr
    # factory defaults: Perspective at (0,0,-1) looking at the origin, fov 60,
    # aspect 1; PathTracing depth 6, epsilon 1e-4; Basic 1 sample
    np.testing.assert_array_equal(job["cam"], np.array([0, 0, -1, 0, 0, 0, 0, 1, 0, 60, 1], np.float32))
    assert (job["W"], job["H"], job["spp"], job["depth"]) == (8, 4, 1, 6)
    assert job["builder"] == ilm.ILM_BUILDER_SAH  # a Triangle shape
    # a renderer without a filter binding is reported (EnvironmentLoader.cpp:685-690)
    assert [l for l in r.stdout.splitlines() if l.startswith("warning")] == [
        "warning: [Renderer] Unable to bind filter []; the box filter is used"]
    np.testing.assert_array_equal(job["lights"], [[1, 2, 3, 0]])
    assert job["tri_light"].tolist() == [0] and job["mats"].shape == (1, 8)
    # the unsupported Sphere, Mirror, Point light and IC integrator are unused: no error


@pytest.mark.parametrize("swap, message", [
    (("Shape = tri", "Shape = ball"), "spheres"),
    (("Material = m", "Material = chrome"), "Mirror"),
    (("Lights = {l}", "Lights = {l, bulb}"), "Point"),
    (("Shape = tri", "Shape = nosuch"), "Instance does not exist"),
])
def test_used_unsupported_entries_fail_loudly(tmp_path, swap, message):
    r, _ = load(tmp_path, _env(SPACE_ONE.replace(*swap)))
    assert r.returncode == 3 and message in r.stderr


def test_unsupported_integrator_fails(tmp_path):
    r, _ = load(tmp_path, _env(SPACE_ONE, renderer="icr"))
    assert r.returncode == 3 and "IC integrator" in r.stderr


def test_light_must_be_bound_to_its_emissive_primitive(tmp_path):
    geo = SPACE_ONE.replace("Type = Emissive", "Type = Geometry")
    r, _ = load(tmp_path, _env(geo))
    assert r.returncode == 3 and "not bound to an Emissive primitive" in r.stderr
    r, _ = load(tmp_path, _env(SPACE_ONE.replace("Lights = {l}", "Lights = {}")))
    assert r.returncode == 3 and "not in the Scene's Lights list" in r.stderr


def test_lexer_rules(tmp_path):
    # a name glued to '=' keeps the digits and swallows the next name as its
    # value (Lexer.cpp:71-85): Fov and Aspect keep their defaults
    cam = ("Cameras {\n  Camera {\n    Id = cam\n    Type = Perspective\n    Fov=30\n    Aspect = 2\n"
           "    Distance = 2\n  }\n}\n")
    text = _env(SPACE_ONE).replace("Cameras {\n  Camera {\n    Id = cam\n    Type = Perspective\n  }\n}\n", cam)
    r, job = load(tmp_path, text)
    assert r.returncode == 0, r.stderr
    # Distance 2 folds into the field of view: 2 atan(tan(30 deg) / 2)
    want = np.float32(2 * np.degrees(np.arctan(np.tan(np.radians(30.0)) / 2)))
    assert abs(job["cam"][9] - want) < 1e-4 and job["cam"][10] == 1.0
    # comment lines, blank lines, and a '#' inside a value (not a comment:
    # the value runs to the end of the line, and a string argument is its
    # first word, ArgumentMap.h:85-97), so the camera's Id is still "cam"
    text2 = _env(SPACE_ONE).replace("    Id = cam\n", "# the camera\n\n    Id = cam  # not a comment\n")
    r, _ = load(tmp_path, text2)
    assert r.returncode == 0, r.stderr
    # the stream ends inside a value: the reference throws
    r, _ = load(tmp_path, _env(SPACE_ONE) + "Trailing = 1")
    assert r.returncode == 3 and "Unexpected end of file" in r.stderr
    # a name at the end of the stream: no sections are read
    r, _ = load(tmp_path, _env(SPACE_ONE) + "Trailing\n")
    assert r.returncode == 3 and "Missing Environment block" in r.stderr
    # entries without Id or Type are skipped with a warning
    r, _ = load(tmp_path, _env(SPACE_ONE).replace("Samplers {\n", "Samplers {\n  Sampler {\n    Type = Random\n  }\n"))
    assert r.returncode == 0 and "without Id or Type" in r.stdout


def test_scene_script_errors(tmp_path):
    r = _run("ilm-load", os.path.join(str(tmp_path), "missing.ilm"), os.path.join(str(tmp_path), "o.bin"))
    assert r.returncode == 3 and "cannot open" in r.stderr
    r, _ = load(tmp_path, "Environment {\n}\n}\n")
    assert r.returncode == 3 and "unbalanced" in r.stderr


@pytest.mark.gpu
def test_room_script_renders_like_the_binding_and_the_workload(tmp_path, gpu_ctx, ilm):
    from illuminaprt_amd import workloads as w
    d = str(tmp_path)
    script = room_script(d, width=64, height=64, samples=16)
    r, job = load(tmp_path, script_path=script)
    assert r.returncode == 0
    out = os.path.join(d, "img.bin")
    r = _run("ilm-render", script, "99", out, cwd=d)
    assert r.returncode == 0, r.stderr
    img = np.fromfile(out, np.float32).reshape(64, 64, 3)
    # the device file (PFM) holds the same pixels
    with open(os.path.join(d, "room.pfm"), "rb") as f:
        head = f.read(len(b"PF\n64 64\n-1\n"))
        pfm = np.frombuffer(f.read(), "<f4").reshape(64, 64, 3)
    assert head.startswith(b"PF") and np.array_equal(pfm, img)
    # the Python binding on the job's arrays with the same seed: identical
    # (the device writes frame pixel (x, y) at (w-1-x, h-1-y))
    m = ilm.Mesh(gpu_ctx, job["pos"], job["idx"])
    m.build()
    sc = ilm.Scene(gpu_ctx, m, job["nrm"], job["tri_mat"], job["tri_light"], job["mats"], job["lights"])
    acc = gpu_ctx.buffer(64 * 64 * 24).zero()
    cam = ilm.Camera.make(job["cam"][0:3], job["cam"][3:6], job["cam"][6:9], float(job["cam"][9]),
                          float(job["cam"][10]))
    sc.render(cam, 64, 64, 16, acc, seed=99, max_depth=5, eps=job["eps"])
    frame = (acc.download(np.float64, 64 * 64 * 3).reshape(64, 64, 3) / 16).astype(np.float32)
    assert np.array_equal(img, frame[::-1, ::-1])
    # the workload's own box room: the same expectation
    pos, nrm, idx, tm, tl, mats, lights = w.box_room()
    m2 = ilm.Mesh(gpu_ctx, pos, idx)
    m2.build()
    sc2 = ilm.Scene(gpu_ctx, m2, nrm, tm, tl, mats, lights)

    def render(s, spp, seed):
        acc.zero()
        s.render(w.room_camera(), 64, 64, spp, acc, seed=seed, max_depth=5)
        return acc.download(np.float64, 64 * 64 * 3).reshape(64, 64, 3) / spp

    ref = render(sc2, 1024, 5)
    e_script = w.rel_rmse(render(sc, 64, 6), ref)
    e_work = w.rel_rmse(render(sc2, 64, 7), ref)
    print(f"[ilm room] relRMSE script {e_script:.4f} workload {e_work:.4f}", flush=True)
    assert e_script < 1.2 * e_work
    acc.free()
    for o in (sc, m, sc2, m2):
        o.close()
