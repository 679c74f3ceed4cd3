"""Writes the converged reference image of the box-room workload.

The oracle (the C++ restatement of the reference PathIntegrator) renders the
256 x 256 box room with 65,536 samples per pixel, depth 5, roulette from depth
4, accumulating in FP64, with a seed (99) independent of the seeds the GPU
tests use. Output: tests/golden/room_ref_256_64k.npz with the per-pixel mean
radiance (float32, 256 x 256 x 3) and the sample count.

Run: python tests/golden/make_room_reference.py [--spp 65536] [--threads 0]
"""
import argparse
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--spp", type=int, default=65536)
    ap.add_argument("--chunk", type=int, default=1024)
    ap.add_argument("--threads", type=int, default=0)
    ap.add_argument("--out", default=os.path.join(HERE, "room_ref_256_64k.npz"))
    args = ap.parse_args()
    import oracle
    from illuminaprt_amd import workloads as w
    oracle.build()
    pos, nrm, idx, tm, tl, mats, lights = w.box_room()
    s = oracle.Scene(pos, nrm, idx, tm, tl, mats, lights)
    cam = w.room_camera().as_list()
    acc = np.zeros((256, 256, 3), np.float64)
    t0 = time.time()
    for off in range(0, args.spp, args.chunk):
        s.render(cam, 256, 256, min(args.chunk, args.spp - off), max_depth=5, rr_depth=4, seed=99,
                 spp_offset=off, nthreads=args.threads, accum=acc)
        print(f"{off + args.chunk}/{args.spp} spp, {time.time() - t0:.0f} s", flush=True)
    np.savez_compressed(args.out, mean=(acc / args.spp).astype(np.float32), spp=args.spp)


if __name__ == "__main__":
    main()
