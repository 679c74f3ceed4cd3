"""Intersection throughput benchmark (one MI355X).

Workload: 1,334,720 triangles (displaced level-8 icosphere + 2000 boxes) and
three 4,194,304-ray sets (primary, shadow, incoherent), built by
illuminaprt_amd.workloads. A step is one closest-hit launch over the
4,194,304 incoherent rays with rays resident in HBM. The primary and shadow
sets are timed the same way and reported beside it.

Prints one JSON line. `value` = incoherent Mrays/s. `roofline` prices the
traversal kernel at two levels of the memory hierarchy:
- HBM (`roofline.achieved` / `frac`): the bytes a launch must move through
  HBM -- 32 B read per ray, 16 B written per hit (1 B per occlusion flag)
  and the tree and triangle arrays read once -- per launch duration, against
  the 8 TB/s peak. `traffic` is the HBM traffic per launch measured by the
  FETCH_SIZE / WRITE_SIZE counters (profiles/traffic_trace_incoherent.json,
  used only when it was measured on the current traversal source);
- L2 (`roofline.l2`): bytes per ray = 32 + 16 + N_node * S_node + N_tri * 48
  (S_node = 64 bytes for the quantised BVH4 traced by default, 112 with
  --nodes float), with N_node and N_tri the mean node
  visits and triangle tests of a deterministic ordered traversal of the same
  flattened tree over the same rays on the host (ilm_bvh_visit_counts), per
  launch duration, against the L2 read bandwidth this run measures with the
  cache-bandwidth probe (ilm_read_probe over a 2 MiB buffer).
`cpu_baseline` times the oracle (the C++ restatement of the reference BVHMesh
build + recursive traversal) on a bounded sample of the same incoherent rays
on all host cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mrays/s incoherent @1M tris; time-to-0.05 relRMSE, neural cache vs plain PT"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_FLOPS = 157.3e12   # MI355X FP32 vector peak
TRI_BYTES, RAY_BYTES, HIT_BYTES = 48, 32, 16


def train_fraction(args, paths):
    """Fraction of a frame's paths extended as training paths: as many as
    the 65,536-sample training batch holds, at most --cache-train-frac."""
    return min(args.cache_train_frac, 65536 / max(1, paths))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_launches(ctx, launch, steps):
    ctx.sync()
    t0 = time.perf_counter()
    marks = [ctx.mark()]
    for _ in range(steps):
        launch()
        marks.append(ctx.mark())
    ctx.sync()
    wall = time.perf_counter() - t0
    per = [ctx.elapsed_ms(marks[i], marks[i + 1]) for i in range(steps)]
    return wall, per


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_trace_incoherent.json")
TRACE_SOURCES = [os.path.join(ROOT, "illuminaprt_amd", "csrc", f) for f in ("trace.hip", "geom.hpp")]


def trace_source_sha16():
    """Digest of the traversal kernel's sources: a traffic measurement counts
    for this run only if it was taken on the same sources."""
    import hashlib
    h = hashlib.sha256()
    for f in TRACE_SOURCES:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def node_format(ilm, args):
    """The BVH4 node format every single-scene leg builds (--nodes)."""
    return ilm.ILM_NODES_QUANTIZED if args.nodes == "quantized" else ilm.ILM_NODES_FLOAT


def trace_kernel_name(nodes):
    """The closest-hit traversal kernel that traces float or quantised BVH4 nodes."""
    return f"k_trace4<false, false, {1 if nodes == 'quantized' else 0}, false, false>"


def traffic_from_profiles(nodes):
    """(HBM bytes per incoherent launch, note): from the committed counter
    summary when it was measured on the current traversal sources and on the
    kernel of the node format traced."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, "no counter summary"
    with open(TRAFFIC_FILE) as f:
        d = json.load(f)
    rel = os.path.relpath(TRAFFIC_FILE, ROOT)
    if d.get("trace_source_sha16") != trace_source_sha16():
        return None, f"{rel} was measured on other traversal sources ({d.get('trace_source_sha16')})"
    if d.get("kernel") != "ilm::" + trace_kernel_name(nodes):
        return None, f"{rel} was measured on {d.get('kernel')}, not on this node format's kernel"
    return d.get("hbm_bytes_per_launch"), f"{rel} (FETCH_SIZE and WRITE_SIZE, traversal sources {d['trace_source_sha16']})"


def pmc_summary(dirs, copy_mib, out_path=TRAFFIC_FILE, nodes="quantized"):
    """Reads four rocprofv3 --pmc runs (CSV output directories: trace FETCH_SIZE,
    trace WRITE_SIZE, copy FETCH_SIZE, copy WRITE_SIZE), scales each trace
    counter by what it reads for the plain copy of copy_mib MiB each way,
    writes the traffic summary JSON and prints it."""
    import csv
    import glob

    def per_dispatch(d, counter):
        vals = {}
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                for row in csv.DictReader(fh):
                    if row.get("Counter_Name") != counter:
                        continue
                    key = (int(row["Dispatch_Id"]), row["Kernel_Name"])
                    vals[key] = vals.get(key, 0.0) + float(row["Counter_Value"])
        return [(k[0], k[1], v) for k, v in sorted(vals.items())]

    tf, tw, cf, cw = dirs
    kib = copy_mib * 1024.0
    copy_f = [v for _, k, v in per_dispatch(cf, "FETCH_SIZE") if "k_copy" in k]
    copy_w = [v for _, k, v in per_dispatch(cw, "WRITE_SIZE") if "k_copy" in k]
    fs, ws = kib / float(np.mean(copy_f)), kib / float(np.mean(copy_w))
    name = trace_kernel_name(nodes)
    tr_f = [v for _, k, v in per_dispatch(tf, "FETCH_SIZE") if name in k]
    tr_w = [v for _, k, v in per_dispatch(tw, "WRITE_SIZE") if name in k]
    # the last six dispatches trace the incoherent set (1 warm-up + 5 timed)
    f_kib, w_kib = float(np.mean(tr_f[-6:])), float(np.mean(tr_w[-6:]))
    out = {"kernel": "ilm::" + name, "workload": "incoherent rays, 1,334,720 triangles, 4,194,304 rays",
           "trace_source_sha16": trace_source_sha16(),
           "fetch_kib_per_dispatch": round(f_kib, 1), "write_kib_per_dispatch": round(w_kib, 1),
           "copy_fetch_kib": round(float(np.mean(copy_f)), 1), "copy_write_kib": round(float(np.mean(copy_w)), 1),
           "fetch_scale": round(fs, 4), "write_scale": round(ws, 4),
           "hbm_bytes_per_launch": int(round((f_kib * fs + w_kib * ws) * 1024.0))}
    with open(out_path, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print(json.dumps(out))


def l2_peak_gbs(ctx, steps=5):
    """Aggregate L2 read bandwidth measured by the cache-bandwidth probe: a
    2 MiB buffer (inside each XCD's 4 MiB L2, far beyond a CU's 32 KiB L1)
    read 8 times by 8 workgroups per CU."""
    nbytes, passes, per_cu = 2 << 20, 8, 8
    src, sink = ctx.buffer(nbytes).zero(), ctx.buffer(16)
    launch = lambda: ctx.read_probe(src, nbytes, passes, per_cu, sink)
    for _ in range(2):
        launch()
    _, per = time_launches(ctx, launch, steps)
    src.free()
    sink.free()
    ms = float(np.mean(per))
    return ctx.cu_count * per_cu * passes * nbytes / (ms * 1e-3) / 1e9, ms


def cpu_baseline(pos, idx, rays_host, budget_s=15.0):
    import oracle
    oracle.lib()
    t0 = time.perf_counter()
    om = oracle.Mesh(pos, idx)
    om.build(10, 35)
    build_s = time.perf_counter() - t0
    probe = rays_host[:20000]
    t0 = time.perf_counter()
    om.trace(probe, mode=0, accel=1)
    rate = len(probe) / (time.perf_counter() - t0)
    k = int(min(len(rays_host), max(20000, rate * budget_s)))
    sel = rays_host[:k]
    t0 = time.perf_counter()
    om.trace(sel, mode=0, accel=1)
    dt = time.perf_counter() - t0
    return {
        "value": round(k / dt / 1e6, 4),
        "unit": "Mrays/s",
        "cores": int(om.last_threads),
        "kind": "port",
        "sample": f"first {k} of the 4,194,304 incoherent rays, closest hit, BVHMesh restatement "
                  f"(reference topology, MaximumLeafObjects 10, depth 35; host build {build_s:.1f} s untimed)",
    }


def path_tracing_room(ctx, ilm, workloads, args):
    """Box room, 256 x 256, 16 spp, depth 5: GPU frame time and Mrays/s over
    camera + extension + shadow rays, beside the oracle's frame."""
    sa = workloads.box_room()
    pos, nrm, idx, tm, tl, mats, lights = sa
    mesh = ilm.Mesh(ctx, pos, idx)
    mesh.build(node_format=node_format(ilm, args))
    sc = ilm.Scene(ctx, mesh, nrm, tm, tl, mats, lights)
    cam = workloads.room_camera()
    acc = ctx.buffer(256 * 256 * 3 * 8).zero()
    for i in range(args.warmup):
        sc.render(cam, 256, 256, 16, acc, seed=100 + i)
    frames = []
    for i in range(max(3, args.steps // 4)):
        frames.append(sc.render(cam, 256, 256, 16, acc, seed=1000 + i))
    ms = float(np.mean([f["ms"] for f in frames]))
    rays = frames[0]["camera_rays"] + frames[0]["extension_rays"] + frames[0]["shadow_rays"]
    out = {"workload": "box room 40+2 tris, 256x256, 16 spp, depth 5", "frame_ms": round(ms, 3),
           "mrays_s": round(rays / ms / 1e3, 1), "rays_per_frame": int(rays)}
    if not args.no_cpu:
        import oracle
        s = oracle.Scene(*sa)
        t0 = time.perf_counter()
        s.render(cam.as_list(), 256, 256, 16, seed=1234)
        dt = time.perf_counter() - t0
        out["cpu_frame_ms"] = round(dt * 1e3, 1)
        out["cpu_mrays_s"] = round(float(s.last_rays.sum()) / dt / 1e6, 2)
        out["cpu_cores"] = int(s.last_threads)
    log(f"path tracing room: {out}")
    acc.free()
    sc.close()
    mesh.close()
    return out


def frames_leg(ctx, ilm, workloads, args):
    """A few plain and neural-cache frames of the radiance-cache workload
    (1920 x 1080, depth 10, 1 spp each): a short run for kernel-level
    profiles of the wavefront integrator."""
    W, H, depth = args.cache_width, args.cache_height, 10
    pos, nrm, idx, tm, tl, mats, lights = workloads.cache_scene()
    mesh = ilm.Mesh(ctx, pos, idx)
    mesh.build(node_format=node_format(ilm, args))
    sc = ilm.Scene(ctx, mesh, nrm, tm, tl, mats, lights)
    cam = workloads.cache_camera(W / H)
    acc = ctx.buffer(W * H * 24).zero()
    net = ilm.Nrc(ctx, *workloads.CACHE_BOUNDS, lr=args.cache_lr, seed=5)
    plain, cache = [], []
    for i in range(args.warmup + args.steps):
        st = sc.render(cam, W, H, 1, acc, max_depth=depth, seed=i)
        if i >= args.warmup:
            plain.append(st)
    for i in range(args.warmup + args.steps):
        st = sc.render_nrc(net, cam, W, H, 1, acc, seed=500 + i, prefix_depth=2, suffix_depth=depth - 2)
        if i >= args.warmup:
            cache.append(st)
    net.close()
    acc.free()
    sc.close()
    mesh.close()

    def summary(fr):
        ms = float(np.mean([f["ms"] for f in fr]))
        rays = fr[0]["camera_rays"] + fr[0]["extension_rays"] + fr[0]["shadow_rays"]
        return {"frame_ms": round(ms, 3), "rays_per_frame": int(rays), "mrays_s": round(rays / ms / 1e3, 1)}

    out = {"workload": f"radiance-cache scene {len(idx)} tris, {W}x{H}, 1 spp, depth {depth}",
           "plain": summary(plain), "neural_cache": summary(cache)}
    log(f"frames: {out}")
    return out


def ic_frames_leg(ctx, ilm, workloads, args):
    """Six irradiance-cache frames of the radiance-cache workload (1920 x 1080,
    1 spp each, the cache kept across frames, error threshold --ic-threshold,
    indirect factor 1, misses 0): a short run for kernel-level profiles of the
    irradiance cache's placement passes, lookups and gather paths."""
    W, H, depth = args.cache_width, args.cache_height, 10
    pos, nrm, idx, tm, tl, mats, lights = workloads.cache_scene()
    mesh = ilm.Mesh(ctx, pos, idx)
    mesh.build(node_format=node_format(ilm, args))
    sc = ilm.Scene(ctx, mesh, nrm, tm, tl, mats, lights)
    cam = workloads.cache_camera(W / H)
    acc = ctx.buffer(W * H * 24).zero()
    ic = ilm.IrradianceCache(sc, ray_depth=depth - 1, error_threshold=args.ic_threshold, indirect_factor=1.0,
                             miss_radiance=(0.0, 0.0, 0.0))
    frames = [ic.render(cam, W, H, 1, acc, seed=30001 + f, spp_offset=f) for f in range(6)]
    ic.close()
    acc.free()
    sc.close()
    mesh.close()
    out = {"workload": f"radiance-cache scene {len(idx)} tris, {W}x{H}, 1 spp, irradiance cache "
                       f"(threshold {args.ic_threshold})",
           "frame_ms": [round(f["ms"], 2) for f in frames], "records_total": [int(f["records_total"]) for f in frames],
           "gather_rays": [int(f["gather_rays"]) for f in frames]}
    log(f"ic frames: {out}")
    return out


def _trace_rate(ctx, mesh, buf, n, steps, warmup, hits):
    for _ in range(warmup):
        mesh.trace_closest(buf, n, hits)
    _, per = time_launches(ctx, lambda: mesh.trace_closest(buf, n, hits), steps)
    ms = float(np.mean(per))
    return round(n / ms / 1e3, 1), round(ms, 4)


def lbvh_leg(ctx, ilm, pos, idx, rs, sah_info, args):
    """The device builder on the same 1,334,720 triangles: build time against
    the host SAH build, and the incoherent trace rate on its tree."""
    m = ilm.Mesh(ctx, pos, idx)
    builds = []
    for _ in range(3):
        builds.append(m.build(builder=ilm.ILM_BUILDER_LBVH)["build_ms"])
    info = m.info()
    hits = ctx.buffer(rs.n * HIT_BYTES)
    rate, ms = _trace_rate(ctx, m, rs.incoherent, rs.n, max(3, args.steps // 4), args.warmup, hits)
    hits.free()
    m.close()
    out = {"device_lbvh_build_ms": round(min(builds), 2), "host_sah_build_ms": round(sah_info["build_ms"], 1),
           "lbvh_incoherent_mrays_s": rate, "lbvh_kernel_ms": ms, "lbvh_sah_cost": round(info["sah_cost"], 2),
           "sah_cost": round(sah_info["sah_cost"], 2), "lbvh_nodes": info["nodes"]}
    log(f"bvh build: {out}")
    return out


def large_scene_leg(ctx, ilm, workloads, args):
    """Two displaced level-9 icospheres and 50,000 boxes (11,085,760
    triangles), one 3840 x 2160 camera: build times (device LBVH, host SAH),
    HBM footprint and primary / incoherent trace rates on both trees."""
    t0 = time.perf_counter()
    pos, idx = workloads.large_scene()
    gen_s = time.perf_counter() - t0
    W, H = 3840, 2160
    n = W * H
    out = {"workload": f"{len(idx)} tris, {W}x{H} camera rays and their cosine bounces", "triangles": int(len(idx)),
           "rays_per_set": n, "generate_s": round(gen_s, 1), "nodes": args.nodes}
    light = np.array([-0.5, 4.0, -0.5, 1.0, 0.0, 0.0, 0.0, 0.0, 1.0], np.float32)
    for name, builder in (("lbvh", ilm.ILM_BUILDER_LBVH), ("sah", ilm.ILM_BUILDER_SAH)):
        m = ilm.Mesh(ctx, pos, idx)
        info = m.build(builder=builder, node_format=node_format(ilm, args))
        rs = workloads.intersection_rays(ctx, m, W, H, cam=workloads.large_camera(W / H), light=light)
        hits = ctx.buffer(n * HIT_BYTES)
        steps = max(3, args.steps // 4)
        p_rate, _ = _trace_rate(ctx, m, rs.primary, n, steps, args.warmup, hits)
        i_rate, i_ms = _trace_rate(ctx, m, rs.incoherent, n, steps, args.warmup, hits)
        out[name] = {"build_ms": round(info["build_ms"], 1), "nodes": info["nodes"],
                     "bvh_bytes": info["node_bytes"] + info["tri_bytes"], "sah_cost": round(info["sah_cost"], 2),
                     "primary_mrays_s": p_rate, "incoherent_mrays_s": i_rate, "incoherent_kernel_ms": i_ms}
        out["ray_bytes"] = n * (3 * RAY_BYTES + 2 * HIT_BYTES)
        hits.free()
        rs.free()
        m.close()
        log(f"large scene {name}: {out[name]}")
    return out


def instances_leg(ctx, ilm, workloads, args):
    """The flat instance table: 2,048 instances of three meshes under a
    top-level tree, against the same instances flattened into one mesh
    (device LBVH). 2048 x 2048 camera rays and their cosine bounces (made on
    the flattened mesh), closest hits by both; the table's device bytes
    against the flattened mesh's, and the time of ilm_instances_set (the
    top-level rebuild of an update)."""
    meshes, placements = workloads.instanced_scene()
    gm = []
    for pos, idx in meshes:
        m = ilm.Mesh(ctx, pos, idx)
        m.build()
        gm.append(m)
    table = ilm.Instances(ctx, gm)
    t0 = time.perf_counter()
    table.set(placements)
    set_ms = (time.perf_counter() - t0) * 1e3
    info = table.info()
    fpos, fidx = workloads.flatten_placements(meshes, placements)
    flat = ilm.Mesh(ctx, fpos, fidx)
    finfo = flat.build(builder=ilm.ILM_BUILDER_LBVH)
    W = H = 2048
    n = W * H
    rs = workloads.intersection_rays(ctx, flat, W, H, cam=workloads.instanced_camera(),
                                     light=np.array([-2.0, 12.0, -2.0, 4.0, 0, 0, 0, 0, 4.0], np.float32))
    hits, inst = ctx.buffer(n * HIT_BYTES), ctx.buffer(n * 4)
    steps = max(3, args.steps // 4)
    out = {"workload": f"{len(placements)} instances of {len(meshes)} meshes ({len(fidx)} triangles flattened), "
                       f"{W}x{H} camera rays and their cosine bounces",
           "instances": info["instances"], "table_bytes": info["table_bytes"], "mesh_bytes": info["mesh_bytes"],
           "tlas_bytes": info["tlas_bytes"], "flattened_bytes": finfo["node_bytes"] + finfo["tri_bytes"],
           "set_ms": round(set_ms, 2)}
    for name, buf in (("primary", rs.primary), ("incoherent", rs.incoherent)):
        for _ in range(args.warmup):
            table.trace_closest(buf, n, hits, inst)
        _, per = time_launches(ctx, lambda: table.trace_closest(buf, n, hits, inst), steps)
        out[f"{name}_mrays_s"] = round(n / float(np.mean(per)) / 1e3, 1)
        rate, _ = _trace_rate(ctx, flat, buf, n, steps, args.warmup, hits)
        out[f"{name}_flattened_mrays_s"] = rate
    table.check()
    for b in (hits, inst):
        b.free()
    rs.free()
    flat.close()
    table.close()
    for m in gm:
        m.close()
    log(f"instances: {out}")
    return out


def large_render_leg(ctx, ilm, workloads, args):
    """11,085,808 triangles (two displaced level-9 icospheres, 50,000 boxes and
    a 48-triangle room with three panel lights; 30% glossy), eight 3840 x 2160
    cameras, depth 12, rendered as one batch (one scene, eight views): plain
    frames and neural-cache frames (2-bounce prefix, 10-bounce training
    suffix), 1 spp per view per frame (a 64-spp frame is 64 of them), with
    --large-batch paths per wavefront batch (default 2^26: the frame's 66 M
    paths at once) and the --nodes tree format. The
    plain frames' per-view means check the cache frames' (after `skip`
    training frames). HBM footprint = device memory in use after the frames
    minus before the scene was created."""
    free0, total = ctx.memory()
    t0 = time.perf_counter()
    pos, nrm, idx, tm, tl, mats, lights = workloads.large_render_scene()
    gen_s = time.perf_counter() - t0
    m = ilm.Mesh(ctx, pos, idx)
    info = m.build(builder=ilm.ILM_BUILDER_LBVH)
    sc = ilm.Scene(ctx, m, nrm, tm, tl, mats, lights)
    b = ilm.Batch(ctx, [sc])
    b.configure(node_format=node_format(ilm, args), batch_paths=args.large_batch)
    sc.close()
    m.close()
    del pos, nrm, idx, tm, tl
    W, H, nv, depth, prefix = 3840, 2160, 8, 12, 2
    views = [(0, c, W, H, 1) for c in workloads.large_cameras(nv, W / H)]
    npix = nv * W * H
    acc = ctx.buffer(npix * 24).zero()

    def view_means():
        img = acc.download(np.float64, npix * 3).reshape(nv, -1)
        return img.mean(axis=1)

    b.render(views, acc, max_depth=depth, seed=300)   # allocates the wavefront state
    acc.zero()
    plain = [b.render(views, acc, max_depth=depth, seed=301 + i) for i in range(args.large_frames)]
    plain_means = view_means() / args.large_frames
    net = ilm.Nrc(ctx, *workloads.LARGE_BOUNDS, lr=args.large_lr, seed=5)
    acc.zero()
    skip, cache = args.large_skip, []
    for i in range(skip + args.large_frames):
        cache.append(b.render_nrc(net, views, acc if i >= skip else None, seed=400 + i, prefix_depth=prefix,
                                  suffix_depth=depth - prefix, train_fraction=train_fraction(args, npix),
                                  train_batch=65536, train_steps=args.large_train_steps))
    cache_means = view_means() / args.large_frames
    free1, _ = ctx.memory()
    net.close()
    acc.free()
    b.close()

    def summary(frames):
        ms = float(np.mean([f["ms"] for f in frames]))
        rays = frames[0]["camera_rays"] + frames[0]["extension_rays"] + frames[0]["shadow_rays"]
        return {"frame_ms": round(ms, 1), "rays_per_frame": int(rays), "mrays_s": round(rays / ms / 1e3, 1),
                "ms_64spp": round(64 * ms, 0)}

    out = {"workload": f"{info['triangles'] if 'triangles' in info else 11085808} tris, {nv} cameras "
                       f"{W}x{H}, depth {depth}, 1 spp per view per frame, one batch",
           "generate_s": round(gen_s, 1), "lbvh_build_ms": round(info["build_ms"], 1), "nodes": args.nodes,
           "batch_paths": args.large_batch or 2**22,
           "hbm_footprint_gib": round((free0 - free1) / 2**30, 2), "hbm_total_gib": round(total / 2**30, 1),
           "plain": summary(plain), "neural_cache": summary(cache[skip:]),
           "cache_training_frames": skip, "cache_train_steps_per_frame": args.large_train_steps,
           "cache_train_fraction": round(train_fraction(args, npix), 6),
           "cache_over_plain_view_means": [round(float(c / p), 4) for c, p in zip(cache_means, plain_means)]}
    log(f"large render: {out}")
    return out


def batch_leg(ctx, ilm, workloads, args):
    """64 box rooms, triangle counts log-uniform in [1k, 2M], square cameras of
    128..1024 pixels, depth 8, rendered by one frame's launch sequence (a
    batch): device BVH builds, autotuned node format and batch size, then
    timed frames at 2 spp (the 32-spp frame is 16 of them)."""
    rng = np.random.default_rng(2024)
    t0 = time.perf_counter()
    scenes, meshes, views, build_ms, tris, pix = [], [], [], 0.0, 0, 0
    for k in range(args.batch_scenes):
        (pos, nrm, idx, tm, tl, mats, lights), cam, side, _ = workloads.batch_scene(rng)
        m = ilm.Mesh(ctx, pos, idx)
        build_ms += m.build(builder=ilm.ILM_BUILDER_LBVH)["build_ms"]
        scenes.append(ilm.Scene(ctx, m, nrm, tm, tl, mats, lights))
        meshes.append(m)
        views.append((k, cam, side, side, args.batch_spp))
        tris += len(idx)
        pix += side * side
    b = ilm.Batch(ctx, scenes)
    for sc in scenes:
        sc.close()
    for m in meshes:
        m.close()
    setup_s = time.perf_counter() - t0
    tune = b.autotune(views, max_depth=8)
    log(f"batch: {len(views)} scenes, {tris} tris, {pix} pixels, setup {setup_s:.1f} s, autotune {tune}")
    acc = ctx.buffer(pix * 24).zero()
    frames = []
    for graph in (False, True):
        b.set_graph(graph)
        for i in range(2):   # the first graph frame captures
            b.render(views, acc, max_depth=8, seed=50 + i)
        for i in range(max(2, args.steps // 8)):
            frames.append((graph, b.render(views, acc, max_depth=8, seed=100 + i)))
    b.set_graph(False)
    acc.free()
    b.close()
    out = {"workload": f"{len(views)} box rooms, {tris} tris (1k-2M log-uniform per room), {pix} pixels, "
                       f"{args.batch_spp} spp per frame, depth 8",
           "triangles": tris, "pixels": pix, "device_build_ms": round(build_ms, 1),
           "autotune": {"node_format": tune["node_format"], "width": tune["width"], "batch_paths": tune["batch_paths"],
                        "table_ms": [[f, wd, bp, round(ms, 2)] for f, wd, bp, ms in tune["table"]]}}
    for graph in (False, True):
        fr = [f for g, f in frames if g == graph]
        ms = float(np.mean([f["ms"] for f in fr]))
        rays = fr[0]["camera_rays"] + fr[0]["extension_rays"] + fr[0]["shadow_rays"]
        out["graph" if graph else "eager"] = {"frame_ms": round(ms, 2), "rays_per_frame": int(rays),
                                               "mrays_s": round(rays / ms / 1e3, 1)}
    log(f"batch: {out}")
    return out


# the network's algorithmic FLOP per query: 83 inputs -> 5 x 64 (ReLU) -> 3
MLP_FLOP_PER_QUERY = 2 * (83 * 64 + 4 * 64 * 64 + 64 * 3)
# as executed: inputs padded to 96, outputs to 16
MLP_FLOP_EXECUTED = 2 * (96 * 64 + 4 * 64 * 64 + 64 * 16)
MFMA_BF16_PEAK_TFLOPS = 2500.0


def _ref_noise2(r1, r2):
    """Squared relRMSE of the mean of two independent renders R = (R1+R2)/2
    against the truth, estimated from their difference:
    E[(R1-R2)^2] = 4 Var(R). Subtracting it from a measured squared relRMSE
    removes the reference's own noise from the figure."""
    ref = 0.5 * (r1 + r2)
    return float(np.mean((r1 - r2) ** 2 / (ref ** 2 + 0.01)) / 4.0), ref


def _corrected(img, ref, noise2):
    e2 = float(np.mean((img - ref) ** 2 / (ref ** 2 + 0.01)))
    return float(np.sqrt(max(e2 - noise2, 0.0))), float(np.sqrt(e2))


class _TargetSearch:
    """Progressive frames checked against the error target at growing
    intervals: the host reads the image only at those checks, so the frames
    in between run back to back (a read after every frame idles the GPU
    while the host computes the error, and the next frame starts slower).
    Those frames are kept as device snapshots of the accumulated image, so a
    check that meets the target is traced back to the first frame since the
    previous check that met it, whose cumulative time is the renderer's time
    to the target (the same frame a check after every frame finds while the
    error falls from frame to frame)."""

    def __init__(self, ctx, buf, nbytes, err):
        self.ctx, self.buf, self.nbytes, self.err = ctx, buf, nbytes, err
        self.snaps, self.pending = [], []

    def keep(self, frame, ms, divisor):
        j = len(self.pending)
        if j == len(self.snaps):
            self.snaps.append(self.ctx.buffer(self.nbytes))
        self.ctx.copy(self.snaps[j], self.buf, self.nbytes)
        self.pending.append((frame, ms, divisor, j))

    def first_hit(self, target):
        """(frame, ms, error) of the first kept frame at the target, or None."""
        for frame, ms, divisor, j in self.pending:
            e = self.err(self.snaps[j], divisor)
            if e <= target:
                return frame, ms, e
        return None

    def clear(self):
        self.pending = []

    def free(self):
        for b in self.snaps:
            b.free()
        self.snaps = []


def cache_convergence(ctx, ilm, workloads, args):
    """Time to relRMSE 0.05 on the radiance-cache workload (displaced level-8
    sphere, 30% glossy, in a closed room with three panel lights), plain
    wavefront PT against the neural-cache renderer, both timed with events
    on the render stream from a cold start (the cache starts untrained).
    The reference is the mean of two independent plain renders; its own
    noise is estimated from their difference and removed (_ref_noise2)."""
    W, H = args.cache_width, args.cache_height
    depth, prefix = 10, 2
    pos, nrm, idx, tm, tl, mats, lights = workloads.cache_scene()
    mesh = ilm.Mesh(ctx, pos, idx)
    mesh.build(node_format=node_format(ilm, args))
    sc = ilm.Scene(ctx, mesh, nrm, tm, tl, mats, lights)
    cam = workloads.cache_camera(W / H)
    n3 = W * H * 3
    buf = ctx.buffer(n3 * 8)

    def plain(spp, seed, spp_offset=0):
        buf.zero()
        done, ms = 0, 0.0
        while done < spp:
            k = min(args.cache_ref_chunk, spp - done)
            ms += sc.render(cam, W, H, k, buf, max_depth=depth, seed=seed, spp_offset=spp_offset + done)["ms"]
            done += k
        return buf.download(np.float64, n3).reshape(H, W, 3) / spp, ms

    t0 = time.perf_counter()
    r1, ms_ref = plain(args.cache_ref_spp, seed=90001)
    r2, _ = plain(args.cache_ref_spp, seed=90002)
    noise2, ref = _ref_noise2(r1, r2)
    del r1, r2
    log(f"cache reference: 2 x {args.cache_ref_spp} spp in {time.perf_counter() - t0:.1f} s, "
        f"ref noise relRMSE {np.sqrt(noise2):.4f}, {ms_ref / args.cache_ref_spp:.2f} ms/spp")

    target = args.cache_target

    def err(b, divisor):
        return _corrected(b.download(np.float64, n3).reshape(H, W, 3) / divisor, ref, noise2)[0]

    search = _TargetSearch(ctx, buf, n3 * 8, err)
    # Plain PT: progressive 1-spp frames, the error checked at every frame up
    # to 8 and then every quarter more; the first frame at the target is
    # found among the snapshots since the previous check. The trace holds
    # the checks and that frame.
    buf.zero()
    plain_ms, spp, trace = 0.0, 0, []
    check = 1
    plain_hit = None
    plain_rays = 0
    while spp < args.cache_max_spp:
        st = sc.render(cam, W, H, 1, buf, max_depth=depth, seed=7001, spp_offset=spp)
        plain_ms += st["ms"]
        plain_rays += st["camera_rays"] + st["extension_rays"] + st["shadow_rays"]
        spp += 1
        if spp != check:
            search.keep(spp, plain_ms, spp)
            continue
        e = err(buf, spp)
        if e <= target:
            first = search.first_hit(target) or (spp, plain_ms, e)
            trace.append((first[0], round(first[1], 2), round(first[2], 5)))
            plain_hit = (first[0], first[1])
            break
        trace.append((spp, round(plain_ms, 2), round(e, 5)))
        search.clear()
        check = spp * 2 if spp < 8 else spp + max(1, spp // 4)
    search.clear()
    # 1/sqrt(spp) fit of the last point: spp needed for the target.
    s_last, ms_last, e_last = trace[-1]
    spp_fit = s_last * (e_last / target) ** 2
    plain_out = {"ms_per_spp": round(plain_ms / spp, 3), "spp_to_target": plain_hit[0] if plain_hit else None,
                 "ms_to_target": round(plain_hit[1], 1) if plain_hit else None,
                 "spp_fit": round(spp_fit, 1), "ms_fit": round(spp_fit * plain_ms / spp, 1),
                 "mrays_s": round(plain_rays / plain_ms / 1e3, 1), "trace": trace}
    log(f"cache plain: {plain_out}")

    # Neural cache: 1-spp frames, each followed by its Adam steps; the image is
    # the running mean of frames from `skip` on (earlier frames see a cold
    # cache). The configured run trains at Adam lr 1e-3 on 65,536-record
    # batches from 1/16 of the paths; the variant at a larger learning rate on
    # one batch filled from fewer paths.
    def nrc_run(lr, frac, steps, skip, seed):
        net = ilm.Nrc(ctx, *workloads.CACHE_BOUNDS, lr=lr, seed=5)
        buf.zero()
        nrc_ms, frames, acc_frames, ntrace = 0.0, 0, 0, []
        nrc_hit = None
        records = 0
        while frames < args.cache_max_frames:
            st = sc.render_nrc(net, cam, W, H, 1, buf if frames >= skip else None, seed=seed + frames,
                               prefix_depth=prefix, suffix_depth=depth - prefix, train_fraction=frac,
                               train_batch=65536, train_steps=steps)
            nrc_ms += st["ms"]
            records = st["train_records"]
            frames += 1
            if frames > skip:
                acc_frames += 1
                # checked like the plain frames, at 1, 2, 4 and every 4th frame
                if not (acc_frames in (1, 2, 4) or acc_frames % 4 == 0):
                    search.keep(frames, nrc_ms, acc_frames)
                    continue
                e = err(buf, acc_frames)
                if e <= target:
                    first = search.first_hit(target) or (frames, nrc_ms, e)
                    ntrace.append((first[0], round(first[1], 2), round(first[2], 5)))
                    nrc_hit = (first[0], first[1])
                    break
                ntrace.append((frames, round(nrc_ms, 2), round(e, 5)))
                search.clear()
        search.clear()
        out = {"adam_lr": lr, "train_fraction": round(frac, 6), "train_records_per_frame": int(records),
               "train_batch": 65536, "train_steps_per_frame": steps, "skip_frames": skip,
               "ms_per_frame": round(nrc_ms / frames, 3), "frames_to_target": nrc_hit[0] if nrc_hit else None,
               "ms_to_target": round(nrc_hit[1], 1) if nrc_hit else None, "trace": ntrace}
        return out, net

    nrc_out, net = nrc_run(args.cache_lr, args.cache_train_frac, args.cache_train_steps, args.cache_skip, 20001)
    log(f"cache nrc: {nrc_out}")
    variant = None
    if args.cache_variant:
        variant, vnet = nrc_run(args.cache_variant_lr, min(args.cache_train_frac, 65536 / (W * H)),
                                args.cache_variant_steps, args.cache_skip, 40001)
        vnet.close()
        log(f"cache nrc variant: {variant}")

    # Irradiance cache (the reference's comparator): 1-spp frames, the cache
    # kept across frames; indirect factor 1 so that it converges to what the
    # path integrator does. Stops at the target or at the plain renderer's
    # time to it.
    ic_out = None
    if not args.no_ic:
        ic = ilm.IrradianceCache(sc, ray_depth=depth - 1, error_threshold=args.ic_threshold, indirect_factor=1.0,
                                 miss_radiance=(0.0, 0.0, 0.0))
        buf.zero()
        ic_ms, f, itrace, ic_hit, best_ic = 0.0, 0, [], None, float("inf")
        budget = plain_out["ms_to_target"] or plain_out["ms_fit"]
        check = 1
        while ic_ms < budget and f < args.cache_max_frames:
            st = ic.render(cam, W, H, 1, buf, seed=30001 + f, spp_offset=f)
            ic_ms += st["ms"]
            f += 1
            if f != check:
                search.keep(f, ic_ms, f)
                continue
            e = err(buf, f)
            best_ic = min(best_ic, e)
            if e <= target:
                first = search.first_hit(target) or (f, ic_ms, e)
                itrace.append((first[0], round(first[1], 2), round(first[2], 5), st["records_total"]))
                ic_hit = (first[0], first[1])
                break
            itrace.append((f, round(ic_ms, 2), round(e, 5), st["records_total"]))
            search.clear()
            check = f * 2 if f < 8 else f + max(1, f // 4)
        if ic_hit is None:
            # the frames after the last check, for the best error reached
            for frame, ms, divisor, j in search.pending:
                best_ic = min(best_ic, err(search.snaps[j], divisor))
        search.clear()
        ic_out = {"error_threshold": args.ic_threshold, "ms_per_frame": round(ic_ms / f, 3),
                  "frames_to_target": ic_hit[0] if ic_hit else None,
                  "ms_to_target": round(ic_hit[1], 1) if ic_hit else None,
                  "best_rel_rmse": round(best_ic, 5), "records": ic.info()["records"], "trace": itrace}
        ic.close()
        log(f"cache ic: {ic_out}")

    # MLP inference throughput on one frame's worth of queries.
    nq = W * H
    rng = np.random.default_rng(3)
    q = np.zeros((nq, 16), np.float32)
    lo, hi = np.array(workloads.CACHE_BOUNDS[0]), np.array(workloads.CACHE_BOUNDS[1])
    q[:, :3] = lo + (hi - lo) * rng.random((nq, 3))
    q[:, 3:7] = rng.random((nq, 4))
    q[:, 7:14] = rng.random((nq, 7))
    qd = ctx.buffer(q.nbytes)
    qd.upload(q)
    od = ctx.buffer(nq * 16)
    for _ in range(2):
        net.infer(qd, nq, od)
    _, per = time_launches(ctx, lambda: net.infer(qd, nq, od), 10)
    inf_ms = float(np.mean(per))
    tflops = nq * MLP_FLOP_PER_QUERY / (inf_ms * 1e-3) / 1e12
    tflops_exec = nq * MLP_FLOP_EXECUTED / (inf_ms * 1e-3) / 1e12
    # One training step (fused forward + backward, weight-gradient reduction,
    # Adam and bf16 repack) on a frame's 65,536 samples: about 3x the forward
    # FLOP per sample
    nt = 65536
    t = np.concatenate([rng.random((nt, 3)), np.ones((nt, 1))], axis=1).astype(np.float32)
    td = ctx.buffer(t.nbytes)
    td.upload(t)
    for _ in range(2):
        net.train_step(qd, td, nt)
    _, per = time_launches(ctx, lambda: net.train_step(qd, td, nt), 10)
    tr_ms = float(np.mean(per))
    tr_tflops = nt * 3 * MLP_FLOP_PER_QUERY / (tr_ms * 1e-3) / 1e12
    mlp = {"queries": nq, "kernel_ms": round(inf_ms, 4), "tflops": round(tflops, 1),
           "flop_per_query": MLP_FLOP_PER_QUERY, "roofline_frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4),
           "executed_flop_per_query": MLP_FLOP_EXECUTED, "executed_tflops": round(tflops_exec, 1),
           "peak_tflops": MFMA_BF16_PEAK_TFLOPS, "kernel": "ilm::k_infer",
           "train_step": {"samples": nt, "ms": round(tr_ms, 4), "tflops": round(tr_tflops, 1),
                          "flop_per_sample": 3 * MLP_FLOP_PER_QUERY,
                          "roofline_frac": round(tr_tflops / MFMA_BF16_PEAK_TFLOPS, 4)}}
    log(f"cache mlp inference: {mlp}")
    search.free()
    for b in (qd, od, td, buf):
        b.free()
    net.close()
    sc.close()
    mesh.close()
    speedup = None
    p_ms = plain_out["ms_to_target"] or plain_out["ms_fit"]
    if nrc_out["ms_to_target"] and p_ms:
        speedup = round(p_ms / nrc_out["ms_to_target"], 3)
    # The CPU restatement of the reference path (BVHMesh topology, unordered
    # traversal, PathIntegrator) on a fixed window of 1/256 of the image at
    # 1 spp: a rate, beside the GPU's plain frames
    cpu = None
    if not args.no_cpu:
        import oracle
        s = oracle.Scene(pos, nrm, idx, tm, tl, mats, lights)
        ww, wh = W // 16, H // 16
        x0, y0 = (W - ww) // 2, (H - wh) // 2
        t0 = time.perf_counter()
        s.render(workloads.cache_camera(W / H).as_list(), W, H, 1, max_depth=depth, seed=7,
                 window=(x0, y0, x0 + ww, y0 + wh))
        dt = time.perf_counter() - t0
        rays = int(s.last_rays.sum())
        cpu = {"window": [x0, y0, x0 + ww, y0 + wh], "seconds": round(dt, 2), "rays": rays,
               "mrays_s": round(rays / dt / 1e6, 4), "cores": int(s.last_threads),
               "gpu_plain_mrays_s": round(plain_out.get("mrays_s", 0.0), 1)}
        # the irradiance-cache integrator (ICIntegrator restatement) on the same
        # window at 1 spp, the cache built from empty, the GPU leg's settings
        t0 = time.perf_counter()
        s.ic_render(workloads.cache_camera(W / H).as_list(), W, H, 1, seed=7, window=(x0, y0, x0 + ww, y0 + wh),
                    error_threshold=args.ic_threshold, ray_depth=depth - 1, indirect_factor=1.0,
                    miss_radiance=(0.0, 0.0, 0.0))
        dt = time.perf_counter() - t0
        c = s.last_counts
        cpu["irradiance_cache"] = {"seconds": round(dt, 2), "camera_rays": int(c[0]), "gather_rays": int(c[1]),
                                   "shadow_rays": int(c[2]), "records": int(c[3]),
                                   "mrays_s": round(float(c[0] + c[1] + c[2]) / dt / 1e6, 4),
                                   "cores": int(s.last_threads), "error_threshold": args.ic_threshold}
        del s
        log(f"cache cpu: {cpu}")
    vspeed = None
    if variant and variant["ms_to_target"] and p_ms:
        variant["speedup_time_to_target"] = round(p_ms / variant["ms_to_target"], 3)
    return {"workload": f"radiance-cache scene {len(idx)} tris, {W}x{H}, depth {depth} "
                        f"(cache after {prefix} bounces)",
            "target_rel_rmse": target, "reference_spp": 2 * args.cache_ref_spp,
            "reference_noise_rel_rmse": round(float(np.sqrt(noise2)), 5),
            "plain_pt": plain_out, "neural_cache": nrc_out, "irradiance_cache": ic_out,
            "speedup_time_to_target": speedup, "neural_cache_variant": variant, "mlp": mlp,
            "cpu_reference_path": cpu}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--only", choices=["incoherent", "primary", "shadow"], default=None,
                    help="time one ray set only (profiling runs)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run the plain copy kernel over MIB mebibytes (bandwidth / counter calibration)")
    ap.add_argument("--no-cache", action="store_true", help="skip the time-to-relRMSE leg")
    ap.add_argument("--no-large", action="store_true", help="skip the 11M-triangle leg")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-scene batched leg")
    ap.add_argument("--batch-only", action="store_true", help="run only the 64-scene batched leg")
    ap.add_argument("--no-instances", action="store_true", help="skip the instance-table leg")
    ap.add_argument("--instances-only", action="store_true", help="run only the instance-table leg")
    ap.add_argument("--no-large-render", action="store_true", help="skip the 8-camera 4K batched render leg")
    ap.add_argument("--large-render-only", action="store_true", help="run only the 8-camera 4K render leg")
    ap.add_argument("--large-frames", type=int, default=2, help="timed frames per renderer (large render leg)")
    ap.add_argument("--large-skip", type=int, default=8, help="cache training frames before timing (large render)")
    ap.add_argument("--large-train-steps", type=int, default=8, help="Adam steps per cache frame (large render)")
    ap.add_argument("--large-batch", type=int, default=1 << 26,
                    help="paths per wavefront batch of the large render (0: the library default, 2^22); the "
                         "default holds a frame's 66 M paths in one batch (30.6 GiB in all)")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-spp", type=int, default=2)
    ap.add_argument("--refill", type=int, default=0,
                    help="idle lanes that trigger a traversal wave's refill, for every ray set (0: 24, 56 for camera rays)")
    ap.add_argument("--nodes", choices=["float", "quantized"], default="quantized",
                    help="BVH4 node format of the traced trees")
    ap.add_argument("--width", type=int, choices=[2, 4, 8], default=4, help="BVH width (float nodes)")
    ap.add_argument("--cache-only", action="store_true", help="run only the time-to-relRMSE leg")
    ap.add_argument("--frames-only", action="store_true",
                    help="run only a few plain and cache frames of the cache workload (profiling)")
    ap.add_argument("--ic-frames-only", action="store_true",
                    help="run only six irradiance-cache frames of the cache workload (profiling)")
    ap.add_argument("--l2-probe-only", action="store_true", help="run only the L2 read-bandwidth probe")
    ap.add_argument("--pmc-summary", nargs=4, metavar="DIR",
                    help="summarise rocprofv3 --pmc CSV directories (trace FETCH_SIZE, trace WRITE_SIZE, copy "
                         "FETCH_SIZE, copy WRITE_SIZE) into the traffic summary and exit")
    ap.add_argument("--pmc-copy-mib", type=int, default=1024, help="size of the copy run of --pmc-summary")
    ap.add_argument("--pmc-out", default=None, help="where --pmc-summary writes (default: the committed summary)")
    ap.add_argument("--cache-width", type=int, default=1920)
    ap.add_argument("--cache-height", type=int, default=1080)
    ap.add_argument("--cache-ref-spp", type=int, default=2048, help="spp of each of the two reference renders")
    ap.add_argument("--cache-ref-chunk", type=int, default=64)
    ap.add_argument("--cache-target", type=float, default=0.05)
    ap.add_argument("--cache-max-spp", type=int, default=1024)
    ap.add_argument("--cache-max-frames", type=int, default=400)
    ap.add_argument("--cache-skip", type=int, default=0, help="neural-cache frames rendered before accumulating")
    ap.add_argument("--cache-lr", type=float, default=1e-3, help="Adam learning rate of the neural cache")
    ap.add_argument("--cache-train-frac", type=float, default=1 / 16,
                    help="fraction of paths extended as training paths (records in batches of 65,536)")
    ap.add_argument("--cache-train-steps", type=int, default=12, help="Adam steps per neural-cache frame")
    ap.add_argument("--cache-variant", type=int, default=1,
                    help="also run the neural cache at --cache-variant-lr on one batch filled from fewer paths")
    ap.add_argument("--cache-variant-lr", type=float, default=1e-2)
    ap.add_argument("--cache-variant-steps", type=int, default=3)
    ap.add_argument("--large-lr", type=float, default=1e-2, help="Adam learning rate of the large-scene cache")
    ap.add_argument("--no-ic", action="store_true", help="skip the irradiance-cache comparator")
    ap.add_argument("--ic-threshold", type=float, default=0.3, help="irradiance-cache error threshold")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(a multi-GPU run would shard the rays by ranges and replicate the BVH)", file=sys.stderr)
        sys.exit(2)

    if args.pmc_summary:
        pmc_summary(args.pmc_summary, args.pmc_copy_mib, args.pmc_out or TRAFFIC_FILE, args.nodes)
        return

    import illuminaprt_amd as ilm
    from illuminaprt_amd import workloads

    ilm.load()
    ctx = ilm.Context(0)
    if args.copy_probe:
        nbytes = args.copy_probe << 20
        src, dst = ctx.buffer(nbytes), ctx.buffer(nbytes)
        for _ in range(args.warmup):
            ctx.copy(dst, src, nbytes)
        _, per = time_launches(ctx, lambda: ctx.copy(dst, src, nbytes), args.steps)
        ms = float(np.mean(per))
        print(json.dumps({"copy_bytes": nbytes, "kernel_ms": ms, "gbs_read_plus_write": 2 * nbytes / ms / 1e6}))
        return
    if args.batch_only:
        print(json.dumps({"batched_scenes": batch_leg(ctx, ilm, workloads, args)}), flush=True)
        ctx.close()
        return
    if args.frames_only:
        print(json.dumps({"frames": frames_leg(ctx, ilm, workloads, args)}), flush=True)
        ctx.close()
        return
    if args.instances_only:
        print(json.dumps({"instances": instances_leg(ctx, ilm, workloads, args)}), flush=True)
        ctx.close()
        return
    if args.ic_frames_only:
        print(json.dumps({"ic_frames": ic_frames_leg(ctx, ilm, workloads, args)}), flush=True)
        ctx.close()
        return
    if args.l2_probe_only:
        gbs, ms = l2_peak_gbs(ctx, args.steps)
        print(json.dumps({"l2_read_gbs": round(gbs, 1), "kernel_ms": round(ms, 4)}), flush=True)
        ctx.close()
        return
    if args.large_render_only:
        print(json.dumps({"large_render": large_render_leg(ctx, ilm, workloads, args)}), flush=True)
        ctx.close()
        return
    if args.cache_only:
        print(json.dumps({"time_to_rel_rmse": cache_convergence(ctx, ilm, workloads, args)}), flush=True)
        ctx.close()
        return
    t0 = time.perf_counter()
    nodes_fmt = ilm.ILM_NODES_QUANTIZED if args.nodes == "quantized" else ilm.ILM_NODES_FLOAT
    pos, idx = workloads.intersection_scene()
    mesh = ilm.Mesh(ctx, pos, idx)
    info = mesh.build(node_format=nodes_fmt, width=args.width)
    rs = workloads.intersection_rays(ctx, mesh)
    n = rs.n
    log(f"setup {time.perf_counter() - t0:.1f} s: {len(idx)} tris, bvh {info}")

    hits = ctx.buffer(n * HIT_BYTES)
    occ = ctx.buffer(n)
    visits = ctx.buffer(n * 8)
    # idle-lane refill threshold per set: the library default (24) for the
    # incoherent and shadow sets, 56 for coherent camera rays, as the path
    # integrator traces them
    refill = {"incoherent": 24, "primary": 56, "shadow": 24}
    if args.refill:
        refill = {k: args.refill for k in refill}
    sets = {
        "incoherent": (lambda: mesh.trace_closest(rs.incoherent, n, hits), rs.incoherent, False),
        "primary": (lambda: mesh.trace_closest(rs.primary, n, hits), rs.primary, False),
        "shadow": (lambda: mesh.trace_any(rs.shadow, n, occ), rs.shadow, True),
    }
    if args.only:
        sets = {args.only: sets[args.only]}
    res = {}
    tree_bytes = info["node_bytes"] + info["tri_bytes"]
    for name, (launch, buf, anyhit) in sets.items():
        ctx.set_trace_refill(refill[name])
        # node visits and triangle tests: the ordered traversal of the same
        # tree on the host (the roofline's counts), and the speculative
        # kernel's own counters beside them
        t0 = time.perf_counter()
        c = mesh.visit_counts(buf.download(np.float32, 8 * n).reshape(n, 8), any_hit=anyhit)
        nn, nt = float(c[:, 0].mean()), float(c[:, 1].mean())
        count_s = time.perf_counter() - t0
        mesh.trace_stats(buf, n, anyhit, visits)
        ctx.sync()
        v = visits.download(np.int32, 2 * n).reshape(n, 2)
        sn, st_ = float(v[:, 0].mean()), float(v[:, 1].mean())
        steps = args.steps if name == "incoherent" else max(3, args.steps // 4)
        for _ in range(args.warmup):
            launch()
        wall, per = time_launches(ctx, launch, steps)
        kern_ms = float(np.mean(per))
        out_bytes = 1 if anyhit else HIT_BYTES
        res[name] = {
            "mrays_s": n / (kern_ms * 1e-3) / 1e6,
            "wall_mrays_s": n * steps / wall / 1e6,
            "kernel_ms": kern_ms,
            "wall_ms_per_step": wall * 1e3 / steps,
            "nodes_per_ray": nn,
            "tris_per_ray": nt,
            "nodes_per_ray_speculative": sn,
            "tris_per_ray_speculative": st_,
            "count_s": count_s,
            "bytes_per_ray": RAY_BYTES + out_bytes + nn * info["node_stride"] + nt * TRI_BYTES,
            "hbm_bytes": n * (RAY_BYTES + out_bytes) + tree_bytes,
            "steps": steps,
            "refill_idle_lanes": refill[name],
        }
        log(f"{name}: {res[name]}")
    ctx.set_trace_refill(24)
    ctx.sync()
    if "incoherent" not in res:
        print(json.dumps(res), flush=True)
        return

    inc = res["incoherent"]
    achieved = inc["hbm_bytes"] / (inc["kernel_ms"] * 1e-3) / 1e9
    l2_achieved = inc["bytes_per_ray"] * n / (inc["kernel_ms"] * 1e-3) / 1e9
    l2_peak, l2_probe_ms = l2_peak_gbs(ctx)
    log(f"L2 read probe: {l2_peak:.0f} GB/s ({l2_probe_ms:.3f} ms)")
    traffic, traffic_note = traffic_from_profiles(args.nodes)
    cpu = None
    if not args.no_cpu:
        rays_host = rs.incoherent.download(np.float32, 8 * n).reshape(n, 8)
        cpu = cpu_baseline(pos, idx, rays_host, args.cpu_budget)
        log(f"cpu baseline: {cpu}")
    build = None if args.only else lbvh_leg(ctx, ilm, pos, idx, rs, info, args)
    room = None if args.only else path_tracing_room(ctx, ilm, workloads, args)
    cache = None if (args.only or args.no_cache) else cache_convergence(ctx, ilm, workloads, args)
    large = None if (args.only or args.no_large) else large_scene_leg(ctx, ilm, workloads, args)
    batch = None if (args.only or args.no_batch) else batch_leg(ctx, ilm, workloads, args)
    lrender = None if (args.only or args.no_large_render) else large_render_leg(ctx, ilm, workloads, args)
    inst = None if (args.only or args.no_instances) else instances_leg(ctx, ilm, workloads, args)
    out = {
        "metric": METRIC,
        "value": round(inc["wall_mrays_s"], 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": inc["steps"],
        "warmup": args.warmup,
        "ms_per_step": round(inc["wall_ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "intersection-only: displaced level-8 icosphere + 2000 boxes, incoherent cosine rays",
            "triangles": int(len(idx)),
            "rays_per_step": n,
            "bvh": {k: info[k] for k in ("width", "nodes", "depth", "max_leaf", "node_bytes", "tri_bytes")},
            "device": ctx.name,
            "cus": ctx.cu_count,
        },
        # per ray class: the HBM fraction by the bytes a launch must move, the
        # L2 fraction by bytes per ray of the ordered traversal against the
        # measured L2 read bandwidth, and the VALU fraction of the FP32 vector
        # peak by the estimate 12 FP32 operations per box test (a node tests
        # one box per child slot) + 30 per triangle
        "rays": {k: {"mrays_s": round(v["mrays_s"], 2), "kernel_ms": round(v["kernel_ms"], 4),
                     "nodes_per_ray": round(v["nodes_per_ray"], 3), "tris_per_ray": round(v["tris_per_ray"], 3),
                     "nodes_per_ray_speculative": round(v["nodes_per_ray_speculative"], 3),
                     "tris_per_ray_speculative": round(v["tris_per_ray_speculative"], 3),
                     "host_count_s": round(v["count_s"], 2),
                     "bytes_per_ray": round(v["bytes_per_ray"], 1),
                     "hbm_frac": round(v["hbm_bytes"] / (v["kernel_ms"] * 1e-3) / (HBM_PEAK_GBS * 1e9), 4),
                     "l2_frac": round(v["bytes_per_ray"] * v["mrays_s"] * 1e6 / (l2_peak * 1e9), 4),
                     "valu_frac_est": round((12 * v["nodes_per_ray"] * info["width"] + 30 * v["tris_per_ray"])
                                            * v["mrays_s"] * 1e6 / FP32_PEAK_FLOPS, 4)}
                 for k, v in res.items()},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_note,
            "traffic_frac": round(traffic / (inc["kernel_ms"] * 1e-3) / (HBM_PEAK_GBS * 1e9), 4) if traffic else None,
            "kernel": f"ilm::k_trace4<false, false, {1 if args.nodes == 'quantized' else {4: 0, 2: 2, 8: 3}[args.width]}, false, false>",
            "algorithmic_bytes_per_launch": int(inc["hbm_bytes"]),
            "algorithmic_bytes": "32 B read per ray + 16 B written per hit + tree and triangle arrays once "
                                 f"({tree_bytes} B)",
            "l2": {"achieved": round(l2_achieved, 1), "peak": round(l2_peak, 1), "unit": "GB/s",
                   "frac": round(l2_achieved / l2_peak, 4), "bytes_per_ray": round(inc["bytes_per_ray"], 1),
                   "nodes_per_ray": round(inc["nodes_per_ray"], 3), "tris_per_ray": round(inc["tris_per_ray"], 3),
                   "node_bytes": info["node_stride"], "tri_bytes": TRI_BYTES,
                   "peak_source": "ilm_read_probe: 2 MiB buffer read 8 times by 8 workgroups per CU"},
        },
        "cpu_baseline": cpu,
        "path_tracing_room": room,
        "time_to_rel_rmse": cache,
        "bvh_build": build,
        "large_scene": large,
        "batched_scenes": batch,
        "large_render": lrender,
        "instances": inst,
    }
    print(json.dumps(out), flush=True)
    for b in (hits, occ, visits):
        b.free()
    rs.free()
    mesh.close()
    ctx.close()


if __name__ == "__main__":
    main()
