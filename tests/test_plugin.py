"""The C++ plug-in layer (illuminaprt_amd/plugin/): loaded with dlopen by a
sandbox program, registered with an EngineKernel, its GpuBVHMesh and
GpuPathTracing created through the factory managers from argument strings,
as the reference's engine does with a plug-in (System/PlugInManager.cpp,
Scene/EnvironmentLoader.cpp:567-647).

The host test needs no GPU. The GPU tests run the sandbox as a child process
and check what it writes against the oracle (brute-force closest and any hit;
the path integrator restatement with the same seed) and against the Python
binding of the same C ABI.
"""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from tests import meshgen

PKG = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "illuminaprt_amd")
PLUGIN = os.path.join(PKG, "libilm_plugin.so")
SANDBOX = os.path.join(PKG, "ilm_plugin_sandbox")


def _run(*args, timeout=120):
    if not (os.path.exists(PLUGIN) and os.path.exists(SANDBOX)):
        pytest.fail("plug-in not built: run __graft_entry__.build()")
    r = subprocess.run([SANDBOX, PLUGIN, *args], capture_output=True, text=True, timeout=timeout)
    print(r.stdout, r.stderr, flush=True)
    return r


def test_plugin_host_checks():
    r = _run("host")
    assert r.returncode == 0 and "host ok" in r.stdout


def test_plugin_exports_entry_symbols():
    import ctypes
    lib = ctypes.CDLL(PLUGIN)
    # the six symbols of Core/System/LibrarySymbols.h:17-23
    for s in ("GetPlugInSingleton", "DisposePlugInSingleton", "CreatePlugInInstance", "GetPlugInInstance",
              "DisposePlugInInstance", "DisposePlugInInstances"):
        assert hasattr(lib, s), s


def test_plugin_multi_instance_symbols_host():
    # Without a device only the handle bookkeeping is checked; the context
    # release is checked by the gpu variant below.
    r = _run("multi")
    assert r.returncode == 0 and "multi ok" in r.stdout


@pytest.mark.gpu
def test_plugin_dispose_releases_device_context(gpu_ctx):
    # Initialise creates the shared context; the no-argument
    # DisposePlugInSingleton and the last DisposePlugInInstance release it.
    r = _run("multi")
    assert r.returncode == 0 and "multi ok" in r.stdout


def test_plugin_reports_errors_as_exceptions():
    # A failure inside the plug-in reaches the caller as the plug-in's
    # exception (exit status 3 with its message), before any device work.
    with tempfile.TemporaryDirectory() as d:
        r = _run("trace", os.path.join(d, "missing.bin"), os.path.join(d, "o.bin"), "Builder = SAH")
        assert r.returncode == 3 and "cannot open" in r.stderr


def _trace_case(tmp, pos, idx, rays, nper, args):
    inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
    with open(inp, "wb") as f:
        np.array([len(pos), len(idx), len(rays), nper], np.int64).tofile(f)
        np.ascontiguousarray(pos, np.float32).tofile(f)
        np.ascontiguousarray(idx, np.int32).tofile(f)
        np.ascontiguousarray(rays, np.float32).tofile(f)
    r = _run("trace", inp, out, args)
    assert r.returncode == 0, r.stderr
    raw = np.fromfile(out, np.uint8)
    n = len(rays)
    hit_dt = np.dtype([("t", "<f4"), ("prim", "<i4"), ("u", "<f4"), ("v", "<f4")])
    o = 0
    batch = raw[o:o + 16 * n].view(hit_dt)
    o += 16 * n
    occ = raw[o:o + n]
    o += n
    per = raw[o:o + 16 * nper].view(hit_dt)
    o += 16 * nper
    per_occ = raw[o:o + nper]
    return batch, occ, per, per_occ


@pytest.mark.gpu
@pytest.mark.parametrize("args", ["Builder = SAH",
                                  "Builder = Reference; MaximumLeafObjects = 10; MaximumTreeDepth = 35"])
def test_plugin_trace_matches_oracle(gpu_ctx, orc, args):
    pos, idx = meshgen.icosphere(3)
    rays = np.concatenate([meshgen.rays_toward(3000, pos, seed=5), meshgen.random_rays(1096, seed=6)])
    nper = 256
    with tempfile.TemporaryDirectory() as tmp:
        batch, occ, per, per_occ = _trace_case(tmp, pos, idx, rays, nper, args)
    om = orc.Mesh(pos, idx)
    pc, tc = om.trace(rays, mode=0, accel=0)
    pa, _ = om.trace(rays, mode=1, accel=0)
    same = (batch["prim"] == pc) | ((batch["prim"] >= 0) & (pc >= 0) & (np.abs(batch["t"] - tc) <= 1e-5))
    print(f"[{args}] closest agreement {same.mean():.6f}, hits {(pc >= 0).mean():.3f}", flush=True)
    assert same.mean() >= 0.9995
    assert ((occ != 0) == (pa >= 0)).mean() >= 0.9995
    # The per-ray IShape::Intersects path returns what the batch returns.
    assert np.array_equal(per["prim"], batch["prim"][:nper])
    assert np.array_equal(per["t"], batch["t"][:nper])
    assert np.array_equal(per_occ, occ[:nper])


@pytest.mark.gpu
def test_plugin_render_matches_binding_and_oracle(gpu_ctx, ilm, orc):
    from illuminaprt_amd import workloads as w
    from tests.test_gpu_pt import _same_sample_agreement

    pos, nrm, idx, tm, tl, mats, lights = w.box_room()
    cam = w.room_camera()
    W = H = 64
    spp, seed, depth = 16, 21, 5
    with tempfile.TemporaryDirectory() as tmp:
        inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(inp, "wb") as f:
            np.array([len(pos), len(idx), len(mats), len(lights)], np.int64).tofile(f)
            np.array([W, H, spp, seed], np.int32).tofile(f)
            np.array(cam.as_list(), np.float32).tofile(f)
            for a, dt in ((pos, np.float32), (idx, np.int32), (nrm, np.float32), (tm, np.int32),
                          (tl, np.int32), (mats, np.float32), (lights, np.float32)):
                np.ascontiguousarray(a, dt).tofile(f)
        r = _run("render", inp, out, f"RayDepth = {depth}; ShadowRays = 1; Epsilon = 0.0001")
        assert r.returncode == 0, r.stderr
        img = np.fromfile(out, np.float32).reshape(H, W, 3).astype(np.float64)
    # Same library through the Python binding: same paths, same sums.
    m = ilm.Mesh(gpu_ctx, pos, idx)
    m.build()
    sc = ilm.Scene(gpu_ctx, m, nrm, tm, tl, mats, lights)
    acc = gpu_ctx.buffer(W * H * 3 * 8).zero()
    sc.render(cam, W, H, spp, acc, max_depth=depth, rr_depth=4, seed=seed)
    ref = acc.download(np.float64, W * H * 3).reshape(H, W, 3) / spp
    acc.free()
    sc.close()
    m.close()
    np.testing.assert_allclose(img, ref, rtol=1e-6, atol=1e-7)
    c = orc.Scene(pos, nrm, idx, tm, tl, mats, lights).render(cam.as_list(), W, H, spp, max_depth=depth,
                                                              rr_depth=4, seed=seed)
    _same_sample_agreement(img * spp, c, spp, "plug-in room 64x64 16spp")


@pytest.mark.gpu
def test_plugin_space_matches_oracle(gpu_ctx, orc):
    # GpuSpace (flattened primitives, one BVH) against the BasicSpace
    # restatement (rays transformed into each primitive's object space).
    inst = meshgen.instances(seed=1)
    wp, _, _, _ = meshgen.flatten_instances(inst)
    rays = meshgen.rays_toward(8000, wp, seed=3, dist=6.0)
    n, nper = len(rays), 200
    with tempfile.TemporaryDirectory() as tmp:
        inp, out = os.path.join(tmp, "in.bin"), os.path.join(tmp, "out.bin")
        with open(inp, "wb") as f:
            np.array([len(inst), n, nper], np.int64).tofile(f)
            for pos, idx, rot, scl, trn in inst:
                np.array([len(pos), len(idx)], np.int64).tofile(f)
                xf = np.concatenate([np.eye(3, dtype=np.float32).reshape(-1) if rot is None else rot,
                                     np.ones(3, np.float32) if scl is None else scl,
                                     np.zeros(3, np.float32) if trn is None else trn]).astype(np.float32)
                xf.tofile(f)
                np.ascontiguousarray(pos, np.float32).tofile(f)
                np.ascontiguousarray(idx, np.int32).tofile(f)
            np.ascontiguousarray(rays, np.float32).tofile(f)
        r = _run("space", inp, out, "")
        assert r.returncode == 0, r.stderr
        raw = np.fromfile(out, np.uint8)
    dt = np.dtype([("t", "<f4"), ("prim", "<i4"), ("tri", "<i4")])
    batch = raw[:12 * n].view(dt)
    per = raw[12 * n:12 * (n + nper)].view(dt)
    occ = raw[12 * (n + nper):]
    space = orc.Space()
    keep = []
    for pos, idx, rot, scl, trn in inst:
        m = orc.Mesh(pos, idx)
        m.build(10, 35)
        keep.append(m)
        space.add(m, rot, scl, trn)
    p, tr, t = space.trace(rays, mode=0)
    same = (batch["prim"] == p) & (batch["tri"] == tr)
    near = (batch["prim"] >= 0) & (p >= 0) & (np.abs(batch["t"] - t) <= 1e-5)
    print(f"plug-in space agreement {(same | near).mean():.6f}", flush=True)
    assert (same | near).mean() >= 0.9995
    assert np.array_equal(per["prim"], batch["prim"][:nper]) and np.array_equal(per["tri"], batch["tri"][:nper])
    pa, _, _ = space.trace(rays[:nper], mode=1)
    assert ((occ != 0) == (pa >= 0)).mean() >= 0.99
