"""The traversal entry point that allocates nothing (ilm_trace_closest_noalloc)
against ilm_trace_closest, through the C ABI: two calls in a row, the second
with fewer rays, on a fresh context (so that the allocating entry point has
not grown the context's spill area yet). Results agree with the allocating
entry point's to the traversal's parity bar (index equal, or both hit within
1e-6 relative t), rays beyond n are left untouched, and the device's free
memory does not change across the calls."""
import numpy as np
import pytest

from tests import meshgen

pytestmark = pytest.mark.gpu


def _scene():
    sp, si = meshgen.icosphere(6)
    tp, ti = meshgen.random_triangles(20000, seed=9, extent=1.5)
    return meshgen.merge([(sp, si), (tp, ti)])


@pytest.mark.parametrize("builder", [0, 1])
def test_noalloc_matches_trace_closest(ilm, builder):
    if not ilm.gpu_available():
        pytest.fail("test marked gpu but no HIP device is visible")
    ctx = ilm.Context(0)
    try:
        pos, idx = _scene()
        rays = np.concatenate([meshgen.rays_toward(150000, pos, seed=21), meshgen.random_rays(50000, seed=22)])
        m = ilm.Mesh(ctx, pos, idx)
        info = m.build(builder=builder)
        print(f"[noalloc b{builder}] stack_need={info['stack_need']} nodes={info['nodes']}", flush=True)
        n = len(rays)
        rb = ilm.rays_to_device(ctx, rays)
        ha, hb = ctx.buffer(n * 16), ctx.buffer(n * 16)
        sentinel = np.zeros(n, ilm.HIT_DTYPE)
        sentinel["prim"] = -7
        ctx.sync()
        for k, cnt in enumerate((n, 1000)):
            hb.upload(sentinel)
            free1, _ = ctx.memory()
            lib = ilm.load()
            ilm.check(lib.ilm_trace_closest_noalloc(m.h, rb.ptr, cnt, hb.ptr, None))
            ctx.sync()
            free2, _ = ctx.memory()
            assert free2 == free1, (k, free1, free2)
            ilm.check(lib.ilm_trace_closest(m.h, rb.ptr, cnt, ha.ptr, None))
            ctx.sync()
            a, b = ha.download(ilm.HIT_DTYPE, n)[:cnt], hb.download(ilm.HIT_DTYPE, n)
            assert (b["prim"][cnt:] == -7).all()
            b = b[:cnt]
            same = a["prim"] == b["prim"]
            both = (a["prim"] >= 0) & (b["prim"] >= 0)
            near = both & (np.abs(a["t"].astype(np.float64) - b["t"]) <= 1e-6 * np.abs(a["t"]))
            print(f"[noalloc b{builder} call {k}] n={cnt} index-equal={same.mean():.7f} "
                  f"agree={np.mean(same | near):.7f} bit-equal={np.mean(a.view(np.uint32).reshape(-1, 4) == b.view(np.uint32).reshape(-1, 4)):.7f}",
                  flush=True)
            assert np.mean(same | near) == 1.0
            assert same.mean() >= 0.9999
        for x in (rb, ha, hb):
            x.free()
        m.close()
    finally:
        ctx.close()
