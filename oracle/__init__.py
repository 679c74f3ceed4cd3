"""CPU oracle (test infrastructure only).

ctypes bindings for ``oracle/_build/liboracle.so``, the C++ restatement of the
reference renderer's CPU hot path (see ``oracle/oracle.h`` for what each entry
restates and where). Only ``tests/``, ``__graft_entry__.smoke()`` and the
CPU-baseline leg of ``bench.py`` may import this module, and only as the
checker or as the timed CPU baseline; the product library never calls it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_SO = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
_u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")


def build() -> str:
    """Compile the oracle with its Makefile (g++ only)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _SO


class RenderParams(C.Structure):
    _fields_ = [
        ("cam", C.c_float * 11),
        ("width", C.c_int32), ("height", C.c_int32),
        ("spp", C.c_int32), ("spp_offset", C.c_int32),
        ("max_depth", C.c_int32), ("rr_depth", C.c_int32),
        ("eps", C.c_float), ("seed", C.c_uint32), ("rng_mode", C.c_int32),
        ("x0", C.c_int32), ("y0", C.c_int32), ("x1", C.c_int32), ("y1", C.c_int32),
        ("pix0", C.c_int64),
    ]


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_SO):
        raise RuntimeError(f"oracle library missing: {_SO} (run __graft_entry__.build())")
    L = C.CDLL(_SO)
    fp3 = C.POINTER(C.c_float)
    L.orc_tri_intersect.argtypes = [_f32p, _f32p, _f32p, _f32p, _f32p, C.c_float, fp3, fp3, fp3]
    L.orc_tri_intersect.restype = C.c_int
    L.orc_aabb_intersect.argtypes = [_f32p, _f32p, _f32p, _f32p, C.c_float, C.c_float, fp3, fp3]
    L.orc_aabb_intersect.restype = C.c_int
    L.orc_mt_words.argtypes = [C.c_uint32, _u32p, C.c_int]
    L.orc_mt_floats.argtypes = [C.c_uint32, _f32p, C.c_int]
    L.orc_philox4x64.argtypes = [_u64p, _u64p, _u64p]
    L.orc_cosine_hemisphere.argtypes = [C.c_float, C.c_float, _f32p]
    L.orc_uniform_triangle.argtypes = [C.c_float, C.c_float, fp3, fp3]
    L.orc_quad_sample.argtypes = [_f32p, C.c_float, C.c_float, _f32p, _f32p]
    L.orc_camera_ray.argtypes = [_f32p, C.c_int, C.c_int, C.c_int, C.c_int, C.c_float, C.c_float, _f32p, _f32p]
    L.orc_bsdf_eval.argtypes = [_f32p, C.c_int64, _f32p, _f32p, _f32p, _f32p]
    L.orc_bsdf_sample.argtypes = [_f32p, C.c_int64, _f32p, _f32p, _f32p, _f32p, _f32p]
    L.orc_ic_render.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, _f64p, _i64p, C.c_int]
    L.orc_ic_render.restype = C.c_int
    L.orc_mesh_create.argtypes = [_f32p, C.c_int64, _i32p, C.c_int64]
    L.orc_mesh_create.restype = C.c_void_p
    L.orc_mesh_destroy.argtypes = [C.c_void_p]
    L.orc_bvh_build.argtypes = [C.c_void_p, C.c_int, C.c_int]
    L.orc_bvh_build.restype = C.c_int64
    L.orc_bvh_stats.argtypes = [C.c_void_p, _i64p]
    L.orc_bvh_export.argtypes = [C.c_void_p, _f32p, _i32p, _i32p, _i32p, _i32p]
    L.orc_kd_build.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int]
    L.orc_kd_build.restype = C.c_int64
    L.orc_kd_stats.argtypes = [C.c_void_p, _i64p]
    L.orc_kd_export.argtypes = [C.c_void_p, _f32p, _i32p, _f32p, _i32p, _i32p, _i32p]
    L.orc_space_create.restype = C.c_void_p
    L.orc_space_destroy.argtypes = [C.c_void_p]
    L.orc_space_add.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
    L.orc_space_add.restype = C.c_int
    L.orc_space_trace.argtypes = [C.c_void_p, _f32p, C.c_int64, C.c_int, _i32p, _i32p, C.c_void_p, C.c_int]
    L.orc_space_trace.restype = C.c_int
    L.orc_trace.argtypes = [C.c_void_p, _f32p, C.c_int64, C.c_int, C.c_int, _i32p, C.c_void_p, C.c_void_p, C.c_int]
    L.orc_trace.restype = C.c_int
    L.orc_scene_create.argtypes = [_f32p, _f32p, C.c_int64, _i32p, _i32p, _i32p, C.c_int64, _f32p, C.c_int, _f32p, C.c_int]
    L.orc_scene_create.restype = C.c_void_p
    L.orc_scene_destroy.argtypes = [C.c_void_p]
    L.orc_render.argtypes = [C.c_void_p, C.POINTER(RenderParams), _f64p, _i64p, C.c_int]
    L.orc_render.restype = C.c_int
    _lib = L
    return L


def _f3(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.float32).reshape(3))


def tri_intersect(p0, p1, p2, o, d, ray_max=np.finfo(np.float32).max):
    t, b, g = C.c_float(), C.c_float(), C.c_float()
    hit = lib().orc_tri_intersect(_f3(p0), _f3(p1), _f3(p2), _f3(o), _f3(d), ray_max,
                                  C.byref(t), C.byref(b), C.byref(g))
    return bool(hit), t.value, b.value, g.value


def aabb_intersect(bmin, bmax, o, d, tmin=0.0, tmax=np.finfo(np.float32).max):
    a, b = C.c_float(), C.c_float()
    hit = lib().orc_aabb_intersect(_f3(bmin), _f3(bmax), _f3(o), _f3(d), tmin, tmax, C.byref(a), C.byref(b))
    return bool(hit), a.value, b.value


def mt_words(seed, n):
    out = np.zeros(n, np.uint32)
    lib().orc_mt_words(seed, out, n)
    return out


def mt_floats(seed, n):
    out = np.zeros(n, np.float32)
    lib().orc_mt_floats(seed, out, n)
    return out


def philox4x64(ctr, key):
    out = np.zeros(4, np.uint64)
    lib().orc_philox4x64(np.ascontiguousarray(ctr, np.uint64), np.ascontiguousarray(key, np.uint64), out)
    return out


def cosine_hemisphere(u, v):
    out = np.zeros(3, np.float32)
    lib().orc_cosine_hemisphere(u, v, out)
    return out


def uniform_triangle(u, v):
    b1, b2 = C.c_float(), C.c_float()
    lib().orc_uniform_triangle(u, v, C.byref(b1), C.byref(b2))
    return b1.value, b2.value


def quad_sample(v1, v2, v3, v4, u, v):
    """The reference's Quad::SamplePoint: (point, normalised winding normal)."""
    q = np.ascontiguousarray(np.concatenate([v1, v2, v3, v4]), np.float32)
    p, n = np.zeros(3, np.float32), np.zeros(3, np.float32)
    lib().orc_quad_sample(q, u, v, p, n)
    return p, n


def camera_ray(cam, w, h, px, py, u, v):
    o = np.zeros(3, np.float32)
    d = np.zeros(3, np.float32)
    lib().orc_camera_ray(np.ascontiguousarray(cam, np.float32), w, h, px, py, u, v, o, d)
    return o, d


def _mat8(kd, ks, exponent, kind):
    return np.array([*kd, *ks, exponent, kind], np.float32)


def bsdf_eval(kd, ks, exponent, kind, wo, wi):
    """f (n x 3) and SampleF's pdf (n) of a material in its local frame."""
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    wi = np.ascontiguousarray(wi, np.float32).reshape(-1, 3)
    n = len(wo)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().orc_bsdf_eval(_mat8(kd, ks, exponent, kind), n, wo, wi, f, pdf)
    return f, pdf


def bsdf_sample(kd, ks, exponent, kind, wo, uvw):
    """BSDF::SampleF for n outgoing directions: wi (n x 3), f (n x 3), pdf (n)."""
    wo = np.ascontiguousarray(wo, np.float32).reshape(-1, 3)
    uvw = np.ascontiguousarray(uvw, np.float32).reshape(-1, 3)
    n = len(wo)
    wi = np.zeros((n, 3), np.float32)
    f = np.zeros((n, 3), np.float32)
    pdf = np.zeros(n, np.float32)
    lib().orc_bsdf_sample(_mat8(kd, ks, exponent, kind), n, wo, uvw, wi, f, pdf)
    return wi, f, pdf


class Mesh:
    """Triangle mesh with the BVHMesh restatement (reference topology)."""

    def __init__(self, pos, idx):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
        self.nv, self.nt = len(pos), len(idx)
        self._h = lib().orc_mesh_create(pos, self.nv, idx, self.nt)
        self.nnodes = 0

    def build(self, max_leaf=10, max_depth=35):
        self.nnodes = lib().orc_bvh_build(self._h, max_leaf, max_depth)
        return self.nnodes

    def stats(self):
        out = np.zeros(6, np.int64)
        lib().orc_bvh_stats(self._h, out)
        return dict(zip(["internal", "leaves", "max_depth", "refs", "max_leaf", "empty"], out.tolist()))

    def export(self):
        n = self.nnodes
        box = np.zeros(6 * n, np.float32)
        child = np.zeros(2 * n, np.int32)
        lf = np.zeros(n, np.int32)
        lc = np.zeros(n, np.int32)
        refs = np.zeros(max(1, self.stats()["refs"]), np.int32)
        lib().orc_bvh_export(self._h, box, child, lf, lc, refs)
        return box.reshape(n, 6), child.reshape(n, 2), lf, lc, refs

    def kd_build(self, kind=1, max_leaf=10, max_depth=35):
        """kind 1: KDTreeMesh, kind 2: KDTreeMeshEx. Trace with accel=2."""
        n = lib().orc_kd_build(self._h, kind, max_leaf, max_depth)
        if n < 0:
            raise ValueError("bad kd-tree kind or depth")
        self.kd_nodes = n
        return n

    def kd_stats(self):
        out = np.zeros(4, np.int64)
        lib().orc_kd_stats(self._h, out)
        return dict(zip(["nodes", "leaves", "max_depth", "refs"], out.tolist()))

    def kd_export(self):
        n = self.kd_nodes
        box = np.zeros(6 * n, np.float32)
        child = np.zeros(2 * n, np.int32)
        split = np.zeros(2 * n, np.float32)
        lf = np.zeros(n, np.int32)
        lc = np.zeros(n, np.int32)
        refs = np.zeros(max(1, self.kd_stats()["refs"]), np.int32)
        lib().orc_kd_export(self._h, box, child, split, lf, lc, refs)
        return box.reshape(n, 6), child.reshape(n, 2), split.reshape(n, 2), lf, lc, refs

    def trace(self, rays, mode=0, accel=1, nthreads=0, want_bary=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        prim = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        bary = np.zeros(2 * n, np.float32) if want_bary else None
        used = lib().orc_trace(self._h, rays, n, mode, accel, prim, t.ctypes.data,
                               bary.ctypes.data if want_bary else None, nthreads)
        self.last_threads = used
        if want_bary:
            return prim, t, bary.reshape(n, 2)
        return prim, t

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_mesh_destroy(self._h)
            self._h = None


class Space:
    """BasicSpace restatement: primitives = (Mesh, rotation elements, scaling,
    translation) traced in object space after Ray::ApplyInverse."""

    def __init__(self):
        self._h = lib().orc_space_create()
        self._keep = []

    def add(self, mesh: "Mesh", rot=None, scl=None, trn=None) -> int:
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32).reshape(-1) for a in (rot, scl, trn)]
        self._keep.append((mesh, arrs))
        ptr = [None if a is None else a.ctypes.data for a in arrs]
        return lib().orc_space_add(self._h, mesh._h, *ptr)

    def trace(self, rays, mode=0, nthreads=0):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = len(rays)
        prim = np.zeros(n, np.int32)
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        lib().orc_space_trace(self._h, rays, n, mode, prim, tri, t.ctypes.data, nthreads)
        return prim, tri, t

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_space_destroy(self._h)
            self._h = None


class IcParams(C.Structure):
    _fields_ = [("cache_depth", C.c_int32), ("error_threshold", C.c_float), ("ambient_resolution", C.c_float),
                ("ambient_multiplier", C.c_float), ("azimuth_strata", C.c_int32), ("altitude_strata", C.c_int32),
                ("ray_depth", C.c_int32), ("eps", C.c_float), ("indirect_factor", C.c_float),
                ("miss_radiance", C.c_float * 3)]


class Scene:
    """Shaded scene for the PathIntegrator restatement."""

    def __init__(self, pos, nrm, idx, tri_mat, tri_light, mats, lights):
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
        tri_mat = np.ascontiguousarray(tri_mat, np.int32)
        tri_light = np.ascontiguousarray(tri_light, np.int32)
        mats = np.ascontiguousarray(mats, np.float32).reshape(-1, 8)
        lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 4)
        self._h = lib().orc_scene_create(pos, nrm, len(pos), idx, tri_mat, tri_light, len(idx),
                                         mats, len(mats), lights, len(lights))

    def render(self, cam, width, height, spp, max_depth=5, rr_depth=4, eps=1e-4, seed=1234,
               rng_mode=0, spp_offset=0, window=None, nthreads=0, accum=None, pix0=0):
        """pix0: offset of the pixel counter of the counter-based streams (the
        first pixel of this image in a batch of images)."""
        p = RenderParams()
        for i, v in enumerate(cam):
            p.cam[i] = float(v)
        p.width, p.height, p.spp, p.spp_offset = width, height, spp, spp_offset
        p.max_depth, p.rr_depth, p.eps, p.seed, p.rng_mode = max_depth, rr_depth, eps, seed, rng_mode
        p.pix0 = pix0
        if window is not None:
            p.x0, p.y0, p.x1, p.y1 = window
        if accum is None:
            accum = np.zeros((height, width, 3), np.float64)
        counts = np.zeros(2, np.int64)
        used = lib().orc_render(self._h, C.byref(p), accum.reshape(-1), counts, nthreads)
        self.last_threads = used
        self.last_rays = counts
        return accum

    def ic_render(self, cam, width, height, spp, seed=1234, window=None, nthreads=0, cache_depth=16,
                  error_threshold=1.0, ambient_resolution=1e-3, ambient_multiplier=1e3, azimuth_strata=12,
                  altitude_strata=4, ray_depth=6, eps=1e-4, indirect_factor=1.0 / np.pi,
                  miss_radiance=(20.0, 50.0, 100.0), spp_offset=0):
        """ICIntegrator restatement (ICIntegrator factory defaults): radiance
        sums per pixel (height, width, 3); self.last_counts = camera rays,
        gather path rays, shadow rays, records created."""
        p = RenderParams()
        for i, v in enumerate(cam):
            p.cam[i] = float(v)
        p.width, p.height, p.spp, p.spp_offset, p.seed = width, height, spp, spp_offset, seed
        if window is not None:
            p.x0, p.y0, p.x1, p.y1 = window
        q = IcParams(cache_depth, error_threshold, ambient_resolution, ambient_multiplier, azimuth_strata,
                     altitude_strata, ray_depth, eps, indirect_factor, (C.c_float * 3)(*miss_radiance))
        accum = np.zeros((height, width, 3), np.float64)
        counts = np.zeros(4, np.int64)
        self.last_threads = lib().orc_ic_render(self._h, C.byref(p), C.byref(q), accum.reshape(-1), counts, nthreads)
        self.last_counts = counts
        return accum

    def __del__(self):
        if getattr(self, "_h", None) and _lib is not None:
            _lib.orc_scene_destroy(self._h)
            self._h = None
