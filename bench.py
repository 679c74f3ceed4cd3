"""Intersection throughput benchmark (one MI355X).

Workload: 1,334,720 triangles (displaced level-8 icosphere + 2000 boxes) and
three 4,194,304-ray sets (primary, shadow, incoherent), built by
illuminaprt_amd.workloads. A step is one closest-hit launch over the
4,194,304 incoherent rays with rays resident in HBM. The primary and shadow
sets are timed the same way and reported beside it.

Prints one JSON line. `value` = incoherent Mrays/s. `roofline` prices the
traversal kernel against HBM bandwidth with algorithmic bytes per ray
= 32 (ray) + 16 (hit) + N_node * 112 + N_tri * 48, where N_node and N_tri are
the mean node visits and triangle tests of the same traversal over the same
rays (counted by the kernel's statistics variant). `cpu_baseline` times the
oracle (the C++ restatement of the reference BVHMesh build + recursive
traversal) on a bounded sample of the same incoherent rays on all host cores.
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
NODE_BYTES, TRI_BYTES, RAY_BYTES, HIT_BYTES = 112, 48, 32, 16


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


def traffic_from_profiles():
    p = os.path.join(ROOT, "profiles", "traffic_trace_incoherent.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        return json.load(f).get("hbm_bytes_per_launch")


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
    mesh.build()
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
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported "
              "(the traversal shards by ray ranges; see DESIGN.md)", file=sys.stderr)
        sys.exit(2)

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
    t0 = time.perf_counter()
    pos, idx = workloads.intersection_scene()
    mesh = ilm.Mesh(ctx, pos, idx)
    info = mesh.build()
    rs = workloads.intersection_rays(ctx, mesh)
    n = rs.n
    log(f"setup {time.perf_counter() - t0:.1f} s: {len(idx)} tris, bvh {info}")

    hits = ctx.buffer(n * HIT_BYTES)
    occ = ctx.buffer(n)
    visits = ctx.buffer(n * 8)
    sets = {
        "incoherent": (lambda: mesh.trace_closest(rs.incoherent, n, hits), rs.incoherent, False),
        "primary": (lambda: mesh.trace_closest(rs.primary, n, hits), rs.primary, False),
        "shadow": (lambda: mesh.trace_any(rs.shadow, n, occ), rs.shadow, True),
    }
    if args.only:
        sets = {args.only: sets[args.only]}
    res = {}
    for name, (launch, buf, anyhit) in sets.items():
        mesh.trace_stats(buf, n, anyhit, visits)
        ctx.sync()
        v = visits.download(np.int32, 2 * n).reshape(n, 2)
        nn, nt = float(v[:, 0].mean()), float(v[:, 1].mean())
        steps = args.steps if name == "incoherent" else max(3, args.steps // 4)
        for _ in range(args.warmup):
            launch()
        wall, per = time_launches(ctx, launch, steps)
        kern_ms = float(np.mean(per))
        res[name] = {
            "mrays_s": n / (kern_ms * 1e-3) / 1e6,
            "wall_mrays_s": n * steps / wall / 1e6,
            "kernel_ms": kern_ms,
            "wall_ms_per_step": wall * 1e3 / steps,
            "nodes_per_ray": nn,
            "tris_per_ray": nt,
            "bytes_per_ray": RAY_BYTES + (1 if anyhit else HIT_BYTES) + nn * NODE_BYTES + nt * TRI_BYTES,
            "steps": steps,
        }
        log(f"{name}: {res[name]}")
    ctx.sync()
    if "incoherent" not in res:
        print(json.dumps(res), flush=True)
        return

    inc = res["incoherent"]
    achieved = inc["bytes_per_ray"] * n / (inc["kernel_ms"] * 1e-3) / 1e9
    cpu = None
    if not args.no_cpu:
        rays_host = rs.incoherent.download(np.float32, 8 * n).reshape(n, 8)
        cpu = cpu_baseline(pos, idx, rays_host, args.cpu_budget)
        log(f"cpu baseline: {cpu}")
    room = None if args.only else path_tracing_room(ctx, ilm, workloads, args)
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
        "rays": {k: {"mrays_s": round(v["mrays_s"], 2), "kernel_ms": round(v["kernel_ms"], 4),
                     "nodes_per_ray": round(v["nodes_per_ray"], 3), "tris_per_ray": round(v["tris_per_ray"], 3)}
                 for k, v in res.items()},
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic_from_profiles(),
            "kernel": "ilm::k_trace4<false,false>",
            "bytes_per_ray": round(inc["bytes_per_ray"], 1),
        },
        "cpu_baseline": cpu,
        "path_tracing_room": room,
    }
    print(json.dumps(out), flush=True)
    for b in (hits, occ, visits):
        b.free()
    rs.free()
    mesh.close()
    ctx.close()


if __name__ == "__main__":
    main()
