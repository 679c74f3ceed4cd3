"""MI355X ray traversal and path integration for IlluminaPRT scenes.

Python bindings over the C ABI of ``libilm_gpu.so`` (declared in
``include/ilm_gpu.h``). The library is built in-tree by
``illuminaprt_amd.build.build()``. Every compute call runs on the GPU; when
no HIP device is visible, ``Context`` raises ``IlmError`` (there is no CPU
fallback).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libilm_gpu.so")
_lib = None

ILM_BUILDER_SAH = 0
ILM_BUILDER_REFERENCE = 1
ILM_BUILDER_KDTREE = 2
ILM_BUILDER_KDTREE_EX = 3
ILM_BUILDER_LBVH = 4
ILM_NODES_FLOAT, ILM_NODES_QUANTIZED = 0, 1
ILM_IC_WARD = 0
ILM_IC_TABELION = 1
ILM_TONE_GLOBAL, ILM_TONE_DRAGO = 0, 1
ILM_IMAGE_PFM, ILM_IMAGE_PPM = 0, 1
ILM_SPAWN_SHADOW = 0
ILM_SPAWN_COSINE = 1
RAY_BYTES = 32
HIT_BYTES = 16
HIT_DTYPE = np.dtype([("t", np.float32), ("prim", np.int32), ("u", np.float32), ("v", np.float32)])


class IlmError(RuntimeError):
    pass


class BvhParams(C.Structure):
    _fields_ = [("builder", C.c_int32), ("width", C.c_int32), ("max_leaf", C.c_int32),
                ("max_depth", C.c_int32), ("threads", C.c_int32), ("node_format", C.c_int32)]


class BvhInfo(C.Structure):
    _fields_ = [("nodes", C.c_int64), ("node_bytes", C.c_int64), ("tri_bytes", C.c_int64),
                ("width", C.c_int32), ("depth", C.c_int32), ("max_leaf", C.c_int32),
                ("stack_need", C.c_int32), ("build_ms", C.c_double), ("sah_cost", C.c_double),
                ("node_format", C.c_int32), ("node_stride", C.c_int32)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class Camera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("target", C.c_float * 3), ("up", C.c_float * 3),
                ("fov_deg", C.c_float), ("aspect", C.c_float)]

    @classmethod
    def make(cls, pos, target, up=(0, 1, 0), fov_deg=40.0, aspect=1.0):
        c = cls()
        for i in range(3):
            c.pos[i], c.target[i], c.up[i] = pos[i], target[i], up[i]
        c.fov_deg, c.aspect = fov_deg, aspect
        return c

    def as_list(self):
        return [*self.pos, *self.target, *self.up, self.fov_deg, self.aspect]


class RenderParams(C.Structure):
    _fields_ = [("cam", Camera), ("width", C.c_int32), ("height", C.c_int32), ("spp", C.c_int32),
                ("spp_offset", C.c_int32), ("max_depth", C.c_int32), ("rr_depth", C.c_int32),
                ("eps", C.c_float), ("seed", C.c_uint32), ("batch_paths", C.c_int32)]


class BatchView(C.Structure):
    _fields_ = [("scene", C.c_int32), ("cam", Camera), ("width", C.c_int32), ("height", C.c_int32),
                ("spp", C.c_int32)]


class BatchTuning(C.Structure):
    _fields_ = [("node_format", C.c_int32), ("width", C.c_int32), ("batch_paths", C.c_int64), ("best_ms", C.c_double),
                ("candidates", C.c_int32), ("candidate_format", C.c_int32 * 16), ("candidate_width", C.c_int32 * 16),
                ("candidate_batch", C.c_int64 * 16), ("candidate_ms", C.c_double * 16)]

    def as_dict(self):
        n = self.candidates
        return {"node_format": self.node_format, "width": self.width, "batch_paths": self.batch_paths,
                "best_ms": self.best_ms,
                "table": [(self.candidate_format[i], self.candidate_width[i], self.candidate_batch[i],
                           self.candidate_ms[i]) for i in range(n)]}


class NrcParams(C.Structure):
    _fields_ = [("bounds_lo", C.c_float * 3), ("bounds_hi", C.c_float * 3), ("lr", C.c_float),
                ("beta1", C.c_float), ("beta2", C.c_float), ("eps", C.c_float), ("seed", C.c_uint32)]


class NrcRenderParams(C.Structure):
    _fields_ = [("prefix_depth", C.c_int32), ("suffix_depth", C.c_int32), ("train_fraction", C.c_float),
                ("train_batch", C.c_int32), ("train_steps", C.c_int32), ("accumulate", C.c_int32)]


class IcParams(C.Structure):
    _fields_ = [("cache_depth", C.c_int32), ("error_threshold", C.c_float), ("ambient_resolution", C.c_float),
                ("ambient_multiplier", C.c_float), ("azimuth_strata", C.c_int32), ("altitude_strata", C.c_int32),
                ("ray_depth", C.c_int32), ("eps", C.c_float), ("indirect_factor", C.c_float),
                ("miss_radiance", C.c_float * 3), ("placement_stride", C.c_int32), ("variant", C.c_int32)]


class IcStats(C.Structure):
    _fields_ = [("records_created", C.c_int64), ("records_total", C.c_int64), ("gather_rays", C.c_int64),
                ("lookups", C.c_int64), ("ms", C.c_double)]


class RenderStats(C.Structure):
    _fields_ = [("camera_rays", C.c_int64), ("extension_rays", C.c_int64), ("shadow_rays", C.c_int64),
                ("ms", C.c_double), ("train_records", C.c_int64)]


# name -> (restype, argtypes); every exported symbol of include/ilm_gpu.h
_vp, _i32, _i64, _u32, _f32, _sz = C.c_void_p, C.c_int32, C.c_int64, C.c_uint32, C.c_float, C.c_size_t
SIGNATURES = {
    "ilm_ctx_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "ilm_ctx_destroy": (C.c_int, [_vp]),
    "ilm_last_error": (C.c_char_p, []),
    "ilm_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ilm_device_info": (C.c_int, [_vp, C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
    "ilm_device_memory": (C.c_int, [_vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "ilm_ctx_set_trace_refill": (C.c_int, [_vp, C.c_int]),
    "ilm_alloc": (C.c_int, [_vp, _sz, C.POINTER(_vp)]),
    "ilm_free": (C.c_int, [_vp, _vp]),
    "ilm_upload": (C.c_int, [_vp, _vp, _vp, _sz]),
    "ilm_download": (C.c_int, [_vp, _vp, _vp, _sz]),
    "ilm_memset": (C.c_int, [_vp, _vp, C.c_int, _sz]),
    "ilm_sync": (C.c_int, [_vp]),
    "ilm_device_copy": (C.c_int, [_vp, _vp, _vp, _sz, _vp]),
    "ilm_read_probe": (C.c_int, [_vp, _vp, _sz, C.c_int, C.c_int, _vp, _vp]),
    "ilm_timer_mark": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "ilm_timer_elapsed": (C.c_int, [_vp, C.c_int, C.c_int, C.POINTER(C.c_float)]),
    "ilm_mesh_create": (C.c_int, [_vp, _vp, _i64, _vp, _i64, C.POINTER(_vp)]),
    "ilm_mesh_destroy": (C.c_int, [_vp]),
    "ilm_bvh_build": (C.c_int, [_vp, C.POINTER(BvhParams)]),
    "ilm_bvh_get_info": (C.c_int, [_vp, C.POINTER(BvhInfo)]),
    "ilm_bvh_export": (C.c_int, [_vp, _vp, _vp]),
    "ilm_trace_closest": (C.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "ilm_trace_any": (C.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "ilm_trace_closest_noalloc": (C.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "ilm_trace_stats": (C.c_int, [_vp, _vp, _i64, C.c_int, _vp, _vp]),
    "ilm_bvh_visit_counts": (C.c_int, [_vp, _vp, _i64, C.c_int, _vp, C.c_int]),
    "ilm_raygen_camera": (C.c_int, [_vp, C.POINTER(Camera), C.c_int, C.c_int, C.c_int, _u32, _u32, _vp, _vp]),
    "ilm_raygen_camera_list": (C.c_int, [_vp, C.POINTER(Camera), C.c_int, C.c_int, _vp, _i64, _i64, _u32, _vp, _vp]),
    "ilm_raygen_secondary": (C.c_int, [_vp, _vp, _vp, _vp, _i64, C.c_int, _vp, _f32, _u32, _vp, _vp]),
    "ilm_compact_hits": (C.c_int, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "ilm_scenegen_icosphere_size": (C.c_int, [C.c_int, C.POINTER(_i64), C.POINTER(_i64)]),
    "ilm_scenegen_icosphere": (C.c_int, [C.c_int, _f32, C.c_int, _u32, _vp, _vp]),
    "ilm_scenegen_boxes": (C.c_int, [C.c_int, _f32, _f32, _f32, _u32, _vp, _vp, _i64]),
    "ilm_scenegen_vertex_normals": (C.c_int, [_vp, _i64, _vp, _i64, _vp]),
    "ilm_instances_create": (C.c_int, [_vp, _vp, C.c_int, C.POINTER(_vp)]),
    "ilm_instances_destroy": (C.c_int, [_vp]),
    "ilm_instances_set": (C.c_int, [_vp, _vp, _i64]),
    "ilm_instances_trace_closest": (C.c_int, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "ilm_instances_trace_any": (C.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "ilm_instances_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64),
                                     C.POINTER(C.c_int32)]),
    "ilm_instances_check": (C.c_int, [_vp]),
    "ilm_space_create": (C.c_int, [_vp, C.POINTER(_vp)]),
    "ilm_space_destroy": (C.c_int, [_vp]),
    "ilm_space_add": (C.c_int, [_vp, _vp, _i64, _vp, _i64, _vp, _vp, _vp, _vp, C.c_int32, C.c_int32,
                                C.POINTER(C.c_int32)]),
    "ilm_space_build": (C.c_int, [_vp, C.POINTER(BvhParams)]),
    "ilm_space_mesh": (C.c_int, [_vp, C.POINTER(_vp)]),
    "ilm_space_size": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    "ilm_space_shading": (C.c_int, [_vp, _vp, _vp, _vp]),
    "ilm_space_resolve": (C.c_int, [_vp, _vp, _i64, _vp, _vp, _vp]),
    "ilm_space_geometry": (C.c_int, [_vp, _vp, _vp]),
    "ilm_obj_load": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "ilm_obj_counts": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]),
    "ilm_obj_arrays": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "ilm_obj_material_info": (C.c_int, [_vp, _vp, C.c_char_p, C.c_int]),
    "ilm_obj_destroy": (C.c_int, [_vp]),
    "ilm_ic_create": (C.c_int, [_vp, C.POINTER(IcParams), C.POINTER(_vp)]),
    "ilm_ic_destroy": (C.c_int, [_vp]),
    "ilm_ic_reset": (C.c_int, [_vp]),
    "ilm_ic_render": (C.c_int, [_vp, C.POINTER(RenderParams), _vp, C.POINTER(IcStats)]),
    "ilm_ic_insert": (C.c_int, [_vp, _vp, _i64]),
    "ilm_ic_info": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64), C.POINTER(C.c_float),
                              C.POINTER(C.c_float), _vp]),
    "ilm_ic_records": (C.c_int, [_vp, _vp]),
    "ilm_ic_node_records": (C.c_int, [_vp, _i64, C.POINTER(C.c_int32), C.POINTER(C.c_int32), _vp, C.c_int32]),
    "ilm_ic_lookup": (C.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "ilm_persistent_load": (C.c_int, [C.c_char_p, C.POINTER(_vp)]),
    "ilm_persistent_counts": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64), C.POINTER(_i64)]),
    "ilm_persistent_arrays": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "ilm_persistent_build": (C.c_int, [_vp, _vp, C.c_int]),
    "ilm_persistent_destroy": (C.c_int, [_vp]),
    "ilm_image_resolve": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_double, C.c_int, _vp, _vp]),
    "ilm_image_tonemap": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp]),
    "ilm_image_write": (C.c_int, [C.c_char_p, _vp, C.c_int, C.c_int, C.c_int]),
    "ilm_scene_create": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_int, _vp, C.c_int, C.POINTER(_vp)]),
    "ilm_scene_destroy": (C.c_int, [_vp]),
    "ilm_render": (C.c_int, [_vp, C.POINTER(RenderParams), _vp, C.POINTER(RenderStats)]),
    "ilm_nrc_create": (C.c_int, [_vp, C.POINTER(NrcParams), C.POINTER(_vp)]),
    "ilm_nrc_destroy": (C.c_int, [_vp]),
    "ilm_nrc_param_count": (C.c_int, [C.POINTER(C.c_int)]),
    "ilm_nrc_set_weights": (C.c_int, [_vp, _vp]),
    "ilm_nrc_get_weights": (C.c_int, [_vp, _vp]),
    "ilm_nrc_get_gradients": (C.c_int, [_vp, _vp]),
    "ilm_nrc_infer": (C.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "ilm_nrc_train_step": (C.c_int, [_vp, _vp, _vp, _i64, _vp, _vp]),
    "ilm_scene_set_graph": (C.c_int, [_vp, C.c_int]),
    "ilm_batch_create": (C.c_int, [_vp, C.POINTER(_vp), C.c_int, C.POINTER(_vp)]),
    "ilm_batch_destroy": (C.c_int, [_vp]),
    "ilm_render_batch": (C.c_int, [_vp, C.POINTER(BatchView), C.c_int, C.POINTER(RenderParams), _vp,
                                   C.POINTER(RenderStats)]),
    "ilm_batch_tree": (C.c_int, [_vp, C.POINTER(_vp), _vp]),
    "ilm_batch_set_graph": (C.c_int, [_vp, C.c_int]),
    "ilm_batch_configure": (C.c_int, [_vp, C.c_int, C.c_int, _i64]),
    "ilm_batch_autotune": (C.c_int, [_vp, C.POINTER(BatchView), C.c_int, C.POINTER(RenderParams),
                                     C.POINTER(BatchTuning)]),
    "ilm_render_nrc": (C.c_int, [_vp, _vp, C.POINTER(RenderParams), C.POINTER(NrcRenderParams), _vp,
                                 C.POINTER(RenderStats)]),
    "ilm_render_batch_nrc": (C.c_int, [_vp, _vp, C.POINTER(BatchView), C.c_int, C.POINTER(RenderParams),
                                       C.POINTER(NrcRenderParams), _vp, C.POINTER(RenderStats)]),
}


def load():
    """Loads libilm_gpu.so (built by illuminaprt_amd.build.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise IlmError(f"{LIB_PATH} is missing: run __graft_entry__.build() (hipcc, gfx950)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = load().ilm_last_error()
        raise IlmError(f"ilm error {rc}: {msg.decode() if msg else ''}")


def gpu_available() -> bool:
    n = C.c_int(0)
    check(load().ilm_device_count(C.byref(n)))
    return n.value > 0


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


# ---------------------------------------------------------------- device memory
class DeviceBuffer:
    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        check(load().ilm_alloc(ctx.h, self.nbytes, C.byref(p)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray, offset: int = 0):
        arr = np.ascontiguousarray(arr)
        assert offset + arr.nbytes <= self.nbytes
        check(load().ilm_upload(self.ctx.h, self.ptr + offset, _ptr(arr), arr.nbytes))
        return self

    def download(self, dtype, count=None, offset: int = 0):
        dtype = np.dtype(dtype)
        if count is None:
            count = (self.nbytes - offset) // dtype.itemsize
        out = np.empty(count, dtype)
        assert offset + out.nbytes <= self.nbytes
        check(load().ilm_download(self.ctx.h, _ptr(out), self.ptr + offset, out.nbytes))
        return out

    def zero(self):
        check(load().ilm_memset(self.ctx.h, self.ptr, 0, self.nbytes))
        return self

    def free(self):
        if self.ptr:
            check(load().ilm_free(self.ctx.h, self.ptr))
            self.ptr = None


class Context:
    """A device context (one stream, owned allocations and scratch)."""

    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(load().ilm_ctx_create(C.byref(h), device))
        self.h = h.value
        name = C.create_string_buffer(64)
        cu = C.c_int()
        check(load().ilm_device_info(self.h, name, 64, C.byref(cu)))
        self.name, self.cu_count = name.value.decode(), cu.value

    def close(self):
        if self.h:
            check(load().ilm_ctx_destroy(self.h))
            self.h = None

    def buffer(self, nbytes) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> DeviceBuffer:
        arr = np.ascontiguousarray(arr)
        return DeviceBuffer(self, max(arr.nbytes, 16)).upload(arr)

    def sync(self):
        check(load().ilm_sync(self.h))

    def set_trace_refill(self, idle_lanes: int):
        check(load().ilm_ctx_set_trace_refill(self.h, idle_lanes))

    def memory(self) -> tuple[int, int]:
        """(free, total) bytes of device memory."""
        fr, tot = C.c_uint64(), C.c_uint64()
        check(load().ilm_device_memory(self.h, C.byref(fr), C.byref(tot)))
        return fr.value, tot.value

    def copy(self, dst: DeviceBuffer, src: DeviceBuffer, nbytes: int):
        check(load().ilm_device_copy(self.h, dst.ptr, src.ptr, nbytes, None))

    def read_probe(self, src: DeviceBuffer, nbytes: int, passes: int, blocks_per_cu: int, sink: DeviceBuffer):
        """Launches the cache-bandwidth probe (reads cu_count * blocks_per_cu * passes * nbytes bytes)."""
        check(load().ilm_read_probe(self.h, src.ptr, nbytes, passes, blocks_per_cu, sink.ptr, None))

    def mark(self) -> int:
        s = C.c_int()
        check(load().ilm_timer_mark(self.h, C.byref(s)))
        return s.value

    def elapsed_ms(self, s0: int, s1: int) -> float:
        ms = C.c_float()
        check(load().ilm_timer_elapsed(self.h, s0, s1, C.byref(ms)))
        return ms.value

    # ---- ray sets
    def raygen_camera(self, cam: Camera, width, height, out: DeviceBuffer, jitter=False, seed=0, sample=0):
        check(load().ilm_raygen_camera(self.h, C.byref(cam), width, height, int(jitter), seed, sample, out.ptr, None))

    def raygen_camera_list(self, cam: Camera, width, height, pixels: DeviceBuffer, npix, n, out: DeviceBuffer,
                           seed=0):
        check(load().ilm_raygen_camera_list(self.h, C.byref(cam), width, height, pixels.ptr, npix, n, seed,
                                            out.ptr, None))

    def compact_hits(self, hits: DeviceBuffer, n, idx_out: DeviceBuffer) -> int:
        cnt = self.buffer(8)
        check(load().ilm_compact_hits(self.h, hits.ptr, n, idx_out.ptr, cnt.ptr, None))
        c = int(cnt.download(np.int64, 1)[0])
        cnt.free()
        return c


class Mesh:
    """A triangle mesh with a GPU BVH (replaces the reference's BVHMesh)."""

    def __init__(self, ctx: Context | None, pos: np.ndarray, idx: np.ndarray):
        self.ctx = ctx
        self.pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        self.idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
        h = C.c_void_p()
        check(load().ilm_mesh_create(ctx.h if ctx else None, _ptr(self.pos), len(self.pos), _ptr(self.idx),
                                     len(self.idx), C.byref(h)))
        self.h = h.value

    @property
    def ntri(self):
        return len(self.idx)

    def build(self, builder=ILM_BUILDER_SAH, width=4, max_leaf=0, max_depth=0, threads=0,
              node_format=ILM_NODES_FLOAT) -> dict:
        p = BvhParams(builder, width, max_leaf, max_depth, threads, node_format)
        check(load().ilm_bvh_build(self.h, C.byref(p)))
        return self.info()

    def info(self) -> dict:
        i = BvhInfo()
        check(load().ilm_bvh_get_info(self.h, C.byref(i)))
        return i.as_dict()

    def export(self):
        info = self.info()
        nodes = np.zeros(info["node_bytes"], np.uint8)
        order = np.zeros(info["tri_bytes"] // 48, np.int32)
        check(load().ilm_bvh_export(self.h, _ptr(nodes), _ptr(order)))
        return nodes, order

    def trace_closest(self, rays: DeviceBuffer, n: int, hits: DeviceBuffer):
        check(load().ilm_trace_closest(self.h, rays.ptr, n, hits.ptr, None))

    def trace_any(self, rays: DeviceBuffer, n: int, occ: DeviceBuffer):
        check(load().ilm_trace_any(self.h, rays.ptr, n, occ.ptr, None))

    def trace_closest_noalloc(self, rays: DeviceBuffer, n: int, hits: DeviceBuffer):
        """trace_closest without any device allocation (fixed per-GPU scratch)."""
        check(load().ilm_trace_closest_noalloc(self.h, rays.ptr, n, hits.ptr, None))

    def visit_counts(self, rays: np.ndarray, any_hit: bool = False, threads: int = 0) -> np.ndarray:
        """Node visits and triangle tests per ray (n x 2 int32) of the ordered,
        non-speculative traversal of this tree, counted on the host."""
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        out = np.zeros((len(rays), 2), np.int32)
        check(load().ilm_bvh_visit_counts(self.h, _ptr(rays), len(rays), int(any_hit), _ptr(out), threads))
        return out

    def trace_stats(self, rays: DeviceBuffer, n: int, any_hit: bool, visits: DeviceBuffer):
        check(load().ilm_trace_stats(self.h, rays.ptr, n, int(any_hit), visits.ptr, None))

    def spawn(self, rays: DeviceBuffer, hits: DeviceBuffer, n: int, kind: int, out: DeviceBuffer, light=None,
              eps=1e-4, seed=0, src: DeviceBuffer | None = None):
        lp = None
        if light is not None:
            light = np.ascontiguousarray(light, np.float32).reshape(9)
            lp = _ptr(light)
        check(load().ilm_raygen_secondary(self.h, rays.ptr, hits.ptr, src.ptr if src else None, n, kind, lp,
                                          eps, seed, out.ptr, None))

    def close(self):
        if self.h and not getattr(self, "borrowed", False):
            check(load().ilm_mesh_destroy(self.h))
        self.h = None


class Batch:
    """Several scenes rendered by one frame's launch sequence."""

    def __init__(self, ctx: Context, scenes):
        self.ctx = ctx
        arr = (C.c_void_p * len(scenes))(*[sc.h for sc in scenes])
        h = C.c_void_p()
        check(load().ilm_batch_create(ctx.h, arr, len(scenes), C.byref(h)))
        self.h = h.value
        self.nscenes = len(scenes)

    @staticmethod
    def _views(views):
        arr = (BatchView * len(views))()
        for i, (scene, cam, w, h, spp) in enumerate(views):
            arr[i].scene, arr[i].cam, arr[i].width, arr[i].height, arr[i].spp = scene, cam, w, h, spp
        return arr

    def render(self, views, accum: DeviceBuffer, max_depth=5, rr_depth=4, eps=1e-4, seed=1234, spp_offset=0,
               batch_paths=0) -> dict:
        """views: list of (scene index, Camera, width, height, spp)."""
        arr = self._views(views)
        p = RenderParams(Camera(), 0, 0, 0, spp_offset, max_depth, rr_depth, eps, seed, batch_paths)
        st = RenderStats()
        check(load().ilm_render_batch(self.h, arr, len(views), C.byref(p), accum.ptr, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def render_nrc(self, nrc: "Nrc", views, accum: DeviceBuffer | None, seed=1234, spp_offset=0, rr_depth=4,
                   eps=1e-4, prefix_depth=2, suffix_depth=8, train_fraction=1 / 16, train_batch=65536,
                   train_steps=1, batch_paths=0) -> dict:
        """The cache renderer over views of the batch (one network for all)."""
        arr = self._views(views)
        p = RenderParams(Camera(), 0, 0, 0, spp_offset, prefix_depth + suffix_depth, rr_depth, eps, seed,
                         batch_paths)
        q = NrcRenderParams(prefix_depth, suffix_depth, train_fraction, train_batch, train_steps,
                            1 if accum is not None else 0)
        st = RenderStats()
        check(load().ilm_render_batch_nrc(self.h, nrc.h, arr, len(views), C.byref(p), C.byref(q),
                                          accum.ptr if accum else None, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def autotune(self, views, max_depth=5, rr_depth=4, eps=1e-4, seed=1234) -> dict:
        arr = self._views(views)
        p = RenderParams(Camera(), 0, 0, 0, 0, max_depth, rr_depth, eps, seed, 0)
        t = BatchTuning()
        check(load().ilm_batch_autotune(self.h, arr, len(views), C.byref(p), C.byref(t)))
        return t.as_dict()

    def configure(self, node_format=ILM_NODES_FLOAT, batch_paths=0, width=4):
        """Node format and BVH width (2, 4, 8; 2 and 8 with float nodes) of
        later frames, and paths per wavefront batch (0 = default)."""
        check(load().ilm_batch_configure(self.h, node_format, width, batch_paths))

    def tree(self):
        """(packed mesh, roots): the batch's float BVH4 as a Mesh view (not
        owned; export() and info() work on it) and each scene's root node."""
        h = C.c_void_p()
        roots = np.zeros(self.nscenes, np.uint32)
        check(load().ilm_batch_tree(self.h, C.byref(h), _ptr(roots)))
        m = Mesh.__new__(Mesh)
        m.h, m.ctx, m.borrowed = h.value, self.ctx, True
        return m, roots

    def set_graph(self, enable: bool):
        check(load().ilm_batch_set_graph(self.h, int(enable)))

    def close(self):
        if self.h:
            load().ilm_batch_destroy(self.h)
            self.h = None


class IrradianceCache:
    """ICIntegrator's irradiance cache on the GPU (ilm_ic_*), over a Scene."""

    def __init__(self, scene: "Scene", cache_depth=0, error_threshold=0.0, ambient_resolution=0.0,
                 ambient_multiplier=0.0, azimuth_strata=0, altitude_strata=0, ray_depth=0, eps=0.0,
                 indirect_factor=0.0, miss_radiance=(20.0, 50.0, 100.0), placement_stride=0, variant=0):
        """variant: ILM_IC_WARD (ICIntegrator) or ILM_IC_TABELION (MLICIntegrator)."""
        p = IcParams(cache_depth, error_threshold, ambient_resolution, ambient_multiplier, azimuth_strata,
                     altitude_strata, ray_depth, eps, indirect_factor, (C.c_float * 3)(*miss_radiance),
                     placement_stride, variant)
        h = C.c_void_p()
        check(load().ilm_ic_create(scene.h, C.byref(p), C.byref(h)))
        self.h = h.value
        self.ctx = scene.ctx

    def render(self, cam: Camera, width, height, spp, accum: DeviceBuffer, seed=1234, spp_offset=0,
               batch_paths=0) -> dict:
        p = RenderParams(cam, width, height, spp, spp_offset, 1, 4, 1e-4, seed, batch_paths)
        st = IcStats()
        check(load().ilm_ic_render(self.h, C.byref(p), accum.ptr, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def insert(self, records):
        """records: n x 10 float32 (P, N, E, Ri)."""
        r = np.ascontiguousarray(records, np.float32).reshape(-1, 10)
        check(load().ilm_ic_insert(self.h, _ptr(r), len(r)))

    def info(self) -> dict:
        nr, nn, nf = _i64(), _i64(), _i64()
        rmin, rmax = C.c_float(), C.c_float()
        root = np.zeros(6, np.float32)
        check(load().ilm_ic_info(self.h, C.byref(nr), C.byref(nn), C.byref(nf), C.byref(rmin), C.byref(rmax),
                                 _ptr(root)))
        return {"records": nr.value, "nodes": nn.value, "refs": nf.value, "rmin": rmin.value, "rmax": rmax.value,
                "root": root}

    def records(self):
        n = self.info()["records"]
        out = np.zeros((n, 11), np.float32)
        if n:
            check(load().ilm_ic_records(self.h, _ptr(out)))
        return out

    def node(self, i):
        cnt, child = C.c_int32(), C.c_int32()
        check(load().ilm_ic_node_records(self.h, i, C.byref(cnt), C.byref(child), None, 0))
        recs = np.zeros(max(1, cnt.value), np.int32)
        check(load().ilm_ic_node_records(self.h, i, C.byref(cnt), C.byref(child), _ptr(recs), cnt.value))
        return child.value, recs[:cnt.value]

    def lookup(self, points):
        """points: n x 6 (P, N) -> n x 4 (weighted mean irradiance, weight sum)."""
        pts = np.ascontiguousarray(points, np.float32).reshape(-1, 6)
        n = len(pts)
        pd = self.ctx.to_device(pts)
        od = self.ctx.buffer(max(16, n * 16))
        check(load().ilm_ic_lookup(self.h, pd.ptr, n, od.ptr, None))
        self.ctx.sync()
        out = od.download(np.float32, 4 * n).reshape(n, 4)
        pd.free()
        od.free()
        return out

    def reset(self):
        check(load().ilm_ic_reset(self.h))

    def close(self):
        if self.h:
            load().ilm_ic_destroy(self.h)
            self.h = None


class Persistent:
    """PersistentMesh files (<trunk>.ver / .tri / .tre): a mesh and its prebuilt kd-tree."""

    def __init__(self, trunk: str):
        h = C.c_void_p()
        check(load().ilm_persistent_load(trunk.encode(), C.byref(h)))
        self.h = h.value
        nv, nt, nn, nr = _i64(), _i64(), _i64(), _i64()
        check(load().ilm_persistent_counts(self.h, C.byref(nv), C.byref(nt), C.byref(nn), C.byref(nr)))
        self.nv, self.nt, self.tree_nodes, self.tree_refs = nv.value, nt.value, nn.value, nr.value

    def arrays(self):
        """pos, nrm (nv x 3 float32), idx (nt x 3 int32), group (nt int32)."""
        pos = np.zeros((self.nv, 3), np.float32)
        nrm = np.zeros((self.nv, 3), np.float32)
        idx = np.zeros((self.nt, 3), np.int32)
        group = np.zeros(self.nt, np.int32)
        check(load().ilm_persistent_arrays(self.h, _ptr(pos), _ptr(nrm), _ptr(idx), _ptr(group)))
        return pos, nrm, idx, group

    def build(self, mesh: "Mesh", node_format=ILM_NODES_FLOAT) -> dict:
        """Installs the file's kd-tree as mesh's BVH (mesh made from arrays())."""
        check(load().ilm_persistent_build(self.h, mesh.h, node_format))
        return mesh.info()

    def close(self):
        if self.h:
            load().ilm_persistent_destroy(self.h)
            self.h = None


class Instance(C.Structure):
    _fields_ = [("mesh", C.c_int32), ("rot", C.c_float * 9), ("scl", C.c_float * 3), ("trn", C.c_float * 3)]


class Instances:
    """The flat instance table: a two-level BVH over shared meshes (each mesh
    stored once; instances are transforms under a top-level tree)."""

    def __init__(self, ctx: Context, meshes):
        self.ctx = ctx
        arr = (C.c_void_p * len(meshes))(*[m.h for m in meshes])
        h = C.c_void_p()
        check(load().ilm_instances_create(ctx.h, arr, len(meshes), C.byref(h)))
        self.h = h.value

    def set(self, instances):
        """instances: list of (mesh index, rot (9, or None), scl (3, or None), trn (3, or None))."""
        arr = (Instance * max(1, len(instances)))()
        for i, (mesh, rot, scl, trn) in enumerate(instances):
            arr[i].mesh = mesh
            r = np.eye(3, dtype=np.float32).reshape(-1) if rot is None else np.asarray(rot, np.float32).reshape(-1)
            sc = np.ones(3, np.float32) if scl is None else np.asarray(scl, np.float32).reshape(-1)
            t = np.zeros(3, np.float32) if trn is None else np.asarray(trn, np.float32).reshape(-1)
            for k in range(9):
                arr[i].rot[k] = float(r[k])
            for k in range(3):
                arr[i].scl[k], arr[i].trn[k] = float(sc[k]), float(t[k])
        check(load().ilm_instances_set(self.h, arr, len(instances)))

    def trace_closest(self, rays: DeviceBuffer, n: int, hits: DeviceBuffer, inst: DeviceBuffer):
        check(load().ilm_instances_trace_closest(self.h, rays.ptr, n, hits.ptr, inst.ptr, None))

    def trace_any(self, rays: DeviceBuffer, n: int, occ: DeviceBuffer):
        check(load().ilm_instances_trace_any(self.h, rays.ptr, n, occ.ptr, None))

    def info(self) -> dict:
        v = [C.c_int64() for _ in range(4)]
        need = C.c_int32()
        check(load().ilm_instances_info(self.h, *[C.byref(x) for x in v], C.byref(need)))
        return {"instances": v[0].value, "mesh_bytes": v[1].value, "table_bytes": v[2].value,
                "tlas_bytes": v[3].value, "stack_need": need.value}

    def check(self):
        check(load().ilm_instances_check(self.h))

    def close(self):
        if self.h:
            load().ilm_instances_destroy(self.h)
            self.h = None


class Space:
    """Primitives (mesh + the reference's Transformation + material + light)
    flattened into one world-space mesh (replaces BasicSpace)."""

    def __init__(self, ctx: Context | None):
        self.ctx = ctx
        h = C.c_void_p()
        check(load().ilm_space_create(ctx.h if ctx else None, C.byref(h)))
        self.h = h.value
        self.mesh = None

    def add(self, pos, idx, nrm=None, rot=None, scl=None, trn=None, material=0, light=-1) -> int:
        pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
        idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
        arrs = [None if a is None else np.ascontiguousarray(a, np.float32).reshape(-1) for a in (nrm, rot, scl, trn)]
        out = C.c_int32()
        check(load().ilm_space_add(self.h, _ptr(pos), len(pos), _ptr(idx), len(idx),
                                   *[None if a is None else _ptr(a) for a in arrs], material, light, C.byref(out)))
        return out.value

    def build(self, builder=ILM_BUILDER_SAH, width=4, max_leaf=0, max_depth=0, threads=0,
              node_format=ILM_NODES_FLOAT) -> dict:
        p = BvhParams(builder, width, max_leaf, max_depth, threads, node_format)
        check(load().ilm_space_build(self.h, C.byref(p)))
        h = C.c_void_p()
        check(load().ilm_space_mesh(self.h, C.byref(h)))
        m = Mesh.__new__(Mesh)       # a view of the space's mesh; the space owns it
        m.ctx, m.h, m.borrowed = self.ctx, h.value, True
        np_, nv, nt = C.c_int64(), C.c_int64(), C.c_int64()
        check(load().ilm_space_size(self.h, C.byref(np_), C.byref(nv), C.byref(nt)))
        self.nprims, self.nv, self.nt = np_.value, nv.value, nt.value
        m.pos, m.idx = None, np.zeros((self.nt, 3), np.int32)
        self.mesh = m
        return m.info()

    def shading(self):
        nrm = np.zeros((self.nv, 3), np.float32)
        tm = np.zeros(self.nt, np.int32)
        tl = np.zeros(self.nt, np.int32)
        check(load().ilm_space_shading(self.h, _ptr(nrm), _ptr(tm), _ptr(tl)))
        return nrm, tm, tl

    def geometry(self):
        pos = np.zeros((self.nv, 3), np.float32)
        idx = np.zeros((self.nt, 3), np.int32)
        check(load().ilm_space_geometry(self.h, _ptr(pos), _ptr(idx)))
        return pos, idx

    def resolve(self, hits: DeviceBuffer, n: int, prim: DeviceBuffer, local: DeviceBuffer):
        check(load().ilm_space_resolve(self.h, hits.ptr, n, prim.ptr, local.ptr, None))

    def close(self):
        if self.h:
            load().ilm_space_destroy(self.h)
            self.h = None
            self.mesh = None


class Scene:
    """Shading data over a Mesh, rendered by the wavefront path integrator
    (replaces the reference's PathIntegrator + BaseRenderer::RenderRegion)."""

    def __init__(self, ctx: Context, mesh: Mesh, nrm, tri_mat, tri_light, mats, lights):
        self.ctx, self.mesh = ctx, mesh
        nrm = np.ascontiguousarray(nrm, np.float32).reshape(-1, 3)
        tri_mat = np.ascontiguousarray(tri_mat, np.int32)
        tri_light = np.ascontiguousarray(tri_light, np.int32)
        mats = np.ascontiguousarray(mats, np.float32).reshape(-1, 8)
        lights = np.ascontiguousarray(lights, np.float32).reshape(-1, 4)
        h = C.c_void_p()
        check(load().ilm_scene_create(ctx.h, mesh.h, _ptr(nrm), _ptr(tri_mat), _ptr(tri_light), _ptr(mats),
                                      len(mats), _ptr(lights), len(lights), C.byref(h)))
        self.h = h.value

    def render(self, cam: Camera, width, height, spp, accum: DeviceBuffer, max_depth=5, rr_depth=4, eps=1e-4,
               seed=1234, spp_offset=0, batch_paths=0) -> dict:
        p = RenderParams(cam, width, height, spp, spp_offset, max_depth, rr_depth, eps, seed, batch_paths)
        st = RenderStats()
        check(load().ilm_render(self.h, C.byref(p), accum.ptr, C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def set_graph(self, enable: bool):
        check(load().ilm_scene_set_graph(self.h, int(enable)))

    def render_nrc(self, nrc: "Nrc", cam: Camera, width, height, spp, accum: DeviceBuffer | None, seed=1234,
                   spp_offset=0, rr_depth=4, eps=1e-4, prefix_depth=2, suffix_depth=8, train_fraction=1 / 16,
                   train_batch=65536, train_steps=1, batch_paths=0) -> dict:
        p = RenderParams(cam, width, height, spp, spp_offset, prefix_depth + suffix_depth, rr_depth, eps, seed,
                         batch_paths)
        q = NrcRenderParams(prefix_depth, suffix_depth, train_fraction, train_batch, train_steps,
                            1 if accum is not None else 0)
        st = RenderStats()
        check(load().ilm_render_nrc(self.h, nrc.h, C.byref(p), C.byref(q), accum.ptr if accum else None,
                                    C.byref(st)))
        return {k: getattr(st, k) for k, _ in st._fields_}

    def close(self):
        if self.h:
            check(load().ilm_scene_destroy(self.h))
            self.h = None


class Nrc:
    """Neural radiance cache (fused bf16 MFMA MLP, trained online)."""

    QUERY_FLOATS = 16

    def __init__(self, ctx: Context, bounds_lo, bounds_hi, lr=1e-3, beta1=0.9, beta2=0.99, eps=1e-8, seed=1):
        self.ctx = ctx
        p = NrcParams()
        for i in range(3):
            p.bounds_lo[i], p.bounds_hi[i] = bounds_lo[i], bounds_hi[i]
        p.lr, p.beta1, p.beta2, p.eps, p.seed = lr, beta1, beta2, eps, seed
        h = C.c_void_p()
        check(load().ilm_nrc_create(ctx.h, C.byref(p), C.byref(h)))
        self.h = h.value
        n = C.c_int()
        check(load().ilm_nrc_param_count(C.byref(n)))
        self.nparams = n.value

    def weights(self):
        w = np.zeros(self.nparams, np.float32)
        check(load().ilm_nrc_get_weights(self.h, _ptr(w)))
        return w

    def gradients(self):
        g = np.zeros(self.nparams, np.float32)
        check(load().ilm_nrc_get_gradients(self.h, _ptr(g)))
        return g

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        assert w.size == self.nparams
        check(load().ilm_nrc_set_weights(self.h, _ptr(w)))

    def infer(self, queries: DeviceBuffer, n: int, out: DeviceBuffer):
        check(load().ilm_nrc_infer(self.h, queries.ptr, n, out.ptr, None))

    def train_step(self, queries: DeviceBuffer, targets: DeviceBuffer, n: int, loss: DeviceBuffer | None = None):
        check(load().ilm_nrc_train_step(self.h, queries.ptr, targets.ptr, n, loss.ptr if loss else None, None))

    def close(self):
        if self.h:
            check(load().ilm_nrc_destroy(self.h))
            self.h = None


# ---------------------------------------------------------------- host-side helpers
def rays_to_device(ctx: Context, rays: np.ndarray) -> DeviceBuffer:
    rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
    return ctx.to_device(rays)


def load_obj(path: str):
    """Reads an OBJ (+ MTL) file: returns pos, nrm, idx, tri_mat, tri_light,
    mats, lights in the order Scene takes them."""
    L = load()
    h = C.c_void_p()
    check(L.ilm_obj_load(path.encode(), C.byref(h)))
    try:
        nv, nt, nm, nl = C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()
        check(L.ilm_obj_counts(h, C.byref(nv), C.byref(nt), C.byref(nm), C.byref(nl)))
        pos = np.zeros((nv.value, 3), np.float32)
        nrm = np.zeros((nv.value, 3), np.float32)
        idx = np.zeros((nt.value, 3), np.int32)
        tm = np.zeros(nt.value, np.int32)
        tl = np.zeros(nt.value, np.int32)
        mats = np.zeros((nm.value, 8), np.float32)
        lights = np.zeros((nl.value, 4), np.float32)
        check(L.ilm_obj_arrays(h, _ptr(pos), _ptr(nrm), _ptr(idx), _ptr(tm), _ptr(tl), _ptr(mats),
                               _ptr(lights) if nl.value else None))
    finally:
        L.ilm_obj_destroy(h)
    return pos, nrm, idx, tm, tl, mats, lights


def image_resolve(ctx: Context, accum: DeviceBuffer, width: int, height: int, samples: int, out: DeviceBuffer,
                  flip=False):
    check(load().ilm_image_resolve(ctx.h, accum.ptr, width, height, 1.0 / samples, int(flip), out.ptr, None))


def image_tonemap(ctx: Context, src: DeviceBuffer, width: int, height: int, op: int, out: DeviceBuffer):
    check(load().ilm_image_tonemap(ctx.h, src.ptr, width, height, op, out.ptr, None))


def image_write(path: str, rgb: np.ndarray, fmt=ILM_IMAGE_PFM):
    rgb = np.ascontiguousarray(rgb, np.float32)
    h, w = rgb.shape[:2]
    check(load().ilm_image_write(path.encode(), _ptr(rgb), w, h, fmt))


def icosphere(level: int, amplitude=0.0, octaves=0, seed=0):
    L = load()
    nv, nt = C.c_int64(), C.c_int64()
    check(L.ilm_scenegen_icosphere_size(level, C.byref(nv), C.byref(nt)))
    pos = np.zeros((nv.value, 3), np.float32)
    idx = np.zeros((nt.value, 3), np.int32)
    check(L.ilm_scenegen_icosphere(level, amplitude, octaves, seed, _ptr(pos), _ptr(idx)))
    return pos, idx


def boxes(n: int, emin=0.01, emax=0.1, radius=3.0, seed=11, vertex_offset=0):
    pos = np.zeros((24 * n, 3), np.float32)
    idx = np.zeros((12 * n, 3), np.int32)
    check(load().ilm_scenegen_boxes(n, emin, emax, radius, seed, _ptr(pos), _ptr(idx), vertex_offset))
    return pos, idx


def vertex_normals(pos, idx):
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    idx = np.ascontiguousarray(idx, np.int32).reshape(-1, 3)
    nrm = np.zeros_like(pos)
    check(load().ilm_scenegen_vertex_normals(_ptr(pos), len(pos), _ptr(idx), len(idx), _ptr(nrm)))
    return nrm


def decode_bvh4(nodes_bytes: np.ndarray, node_format: int = 0):
    """Splits an exported BVH4 node array into boxes (n,4,6) and child words (n,4).
    Quantised nodes are decoded to float boxes lo = p + qlo 2^e, hi = p + qhi 2^e
    (float64, as the traversal decodes them in exact arithmetic)."""
    if node_format == ILM_NODES_QUANTIZED:
        raw = nodes_bytes.view(np.uint8).reshape(-1, 64)
        p = raw[:, 0:12].copy().view(np.float32).reshape(-1, 3).astype(np.float64)
        ex = raw[:, 12:15].astype(np.int64) - 127                      # (n, 3)
        child = raw[:, 16:32].copy().view(np.uint32).reshape(-1, 4)
        q = raw[:, 32:56].reshape(-1, 6, 4).astype(np.float64)          # qlox,qhix,qloy,qhiy,qloz,qhiz x child
        step = np.ldexp(1.0, ex)                                       # (n, 3)
        lo = p[:, None, :] + np.stack([q[:, 0], q[:, 2], q[:, 4]], axis=-1) * step[:, None, :]
        hi = p[:, None, :] + np.stack([q[:, 1], q[:, 3], q[:, 5]], axis=-1) * step[:, None, :]
        return np.concatenate([lo, hi], axis=-1), child
    return decode_bvh(nodes_bytes, 4)


def decode_bvh(nodes_bytes: np.ndarray, width: int):
    """Float nodes of width 2, 4 or 8 (W lox, W hix, W loy, W hiy, W loz, W hiz,
    W child words; BVH2 nodes padded to 64 bytes) -> boxes (n,W,6), child words (n,W)."""
    stride = {2: 64, 4: 112, 8: 224}[width]
    raw = nodes_bytes.view(np.uint8).reshape(-1, stride)
    f = raw[:, :24 * width].copy().view(np.float32).reshape(-1, 6, width)
    child = raw[:, 24 * width:28 * width].copy().view(np.uint32).reshape(-1, width)
    boxes = np.stack([f[:, 0], f[:, 2], f[:, 4], f[:, 1], f[:, 3], f[:, 5]], axis=-1)  # (n,W,6) lo xyz, hi xyz
    return boxes, child
