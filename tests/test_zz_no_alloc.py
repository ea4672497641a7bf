"""srt_trace_rays_noalloc: ray queries whose scratch (fetch counters and
traversal-stack spill) is a file-scope device array instead of context
allocations.  Its closest and any hits equal srt_trace_rays's bit for bit on
two calls in a row, the second with fewer rays, and the context's HBM use
does not change across the calls (no allocation)."""
import numpy as np
import pytest

from srt_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _rays(s, n, seed):
    rng = np.random.default_rng(seed)
    lo, hi = s.positions.min(0), s.positions.max(0)
    o = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    tmax = rng.uniform(0.05, float(np.linalg.norm(hi - lo)), size=n).astype(np.float32)
    return o, d, tmax


def test_noalloc_equals_trace_rays():
    ctx = abi.Context(0)
    try:
        s = scenes.spheres()
        h = ctx.upload(s)
        ctx.build_bvh(h)
        n_max = 300000
        bufs = {k: ctx.device_alloc(n_max * w) for k, w in (("o", 12), ("d", 12), ("tm", 4), ("t", 4), ("u", 4),
                                                               ("v", 4), ("p", 4))}
        for call, (n, seed) in enumerate(((n_max, 1), (77777, 2))):
            o, d, tmax = _rays(s, n, seed)
            ctx.upload_bytes(bufs["o"], o)
            ctx.upload_bytes(bufs["d"], d)
            ctx.upload_bytes(bufs["tm"], tmax)
            for any_hit in (False, True):
                used = ctx.lib.srt_hbm_used(ctx.h)
                ctx.trace_noalloc(h, n, bufs["o"], bufs["d"], bufs["tm"], bufs["t"], bufs["u"], bufs["v"], bufs["p"],
                                  any_hit=any_hit)
                assert ctx.lib.srt_hbm_used(ctx.h) == used
                t = np.zeros(n, np.float32); u = np.zeros(n, np.float32); v = np.zeros(n, np.float32)
                p = np.zeros(n, np.int32)
                ctx.download_bytes(p, bufs["p"])
                ref = ctx.trace(h, o, d, tmax=tmax, any_hit=any_hit)
                np.testing.assert_array_equal(p, ref[3])
                if not any_hit:
                    ctx.download_bytes(t, bufs["t"]); ctx.download_bytes(u, bufs["u"]); ctx.download_bytes(v, bufs["v"])
                    for x, y in ((t, ref[0]), (u, ref[1]), (v, ref[2])):
                        assert np.array_equal(x.view(np.uint32), y.view(np.uint32))
                    assert 0.05 < (p >= 0).mean() < 0.95
                else:
                    assert 0.02 < p.mean() < 0.98
        # camera rays of the Cornell box through the same scratch after the sphere scene
        c = scenes.cornell()
        hc = ctx.upload(c)
        ctx.build_bvh(hc)
        from oracle import pyoracle as O
        o, d = O.primary_rays(c.camera, 64, 48, 3)
        n = o.shape[0]
        tmax = np.full(n, 1e16, np.float32)
        ctx.upload_bytes(bufs["o"], o); ctx.upload_bytes(bufs["d"], d); ctx.upload_bytes(bufs["tm"], tmax)
        ctx.trace_noalloc(hc, n, bufs["o"], bufs["d"], bufs["tm"], bufs["t"], bufs["u"], bufs["v"], bufs["p"])
        p = np.zeros(n, np.int32)
        ctx.download_bytes(p, bufs["p"])
        np.testing.assert_array_equal(p, ctx.trace(hc, o, d)[3])
        for b in bufs.values():
            ctx.device_free(b)
    finally:
        ctx.close()


def test_noalloc_on_an_instanced_scene():
    """The same entry point on a two-level scene: closest and any hits equal
    srt_trace_rays_inst's bit for bit (the instance index is not returned
    there), with unchanged HBM use."""
    ctx = abi.Context(0)
    try:
        s = scenes.spheres_instanced(num_spheres=400, num_meshes=6, seed=9)
        h = ctx.upload(s)
        ctx.build_bvh(h)
        n = 100000
        rng = np.random.default_rng(4)
        c = s.inst_xf[:, :, 3].astype(np.float64)
        o = rng.uniform(c.min(0) - 1.0, c.max(0) + 1.0, size=(n, 3))
        d = rng.normal(size=(n, 3))
        aim = rng.uniform(size=n) < 0.5
        d[aim] = c[rng.integers(0, c.shape[0], aim.sum())] - o[aim]
        o = o.astype(np.float32)
        d = d.astype(np.float32)
        tmax = rng.uniform(1.0, 150.0, size=n).astype(np.float32)
        bufs = {k: ctx.device_alloc(n * w) for k, w in (("o", 12), ("d", 12), ("tm", 4), ("t", 4), ("u", 4),
                                                           ("v", 4), ("p", 4))}
        ctx.upload_bytes(bufs["o"], o)
        ctx.upload_bytes(bufs["d"], d)
        ctx.upload_bytes(bufs["tm"], tmax)
        for any_hit in (False, True):
            used = ctx.lib.srt_hbm_used(ctx.h)
            ctx.trace_noalloc(h, n, bufs["o"], bufs["d"], bufs["tm"], bufs["t"], bufs["u"], bufs["v"], bufs["p"],
                              any_hit=any_hit)
            assert ctx.lib.srt_hbm_used(ctx.h) == used
            p = np.zeros(n, np.int32)
            ctx.download_bytes(p, bufs["p"])
            ref = ctx.trace_inst(h, o, d, tmax=tmax, any_hit=any_hit)
            if any_hit:
                np.testing.assert_array_equal(p, ref)
                assert 0.02 < p.mean() < 0.98
            else:
                np.testing.assert_array_equal(p, ref[3])
                t = np.zeros(n, np.float32); u = np.zeros(n, np.float32); v = np.zeros(n, np.float32)
                ctx.download_bytes(t, bufs["t"]); ctx.download_bytes(u, bufs["u"]); ctx.download_bytes(v, bufs["v"])
                hit = p >= 0
                for x, y in ((t, ref[0]), (u, ref[1]), (v, ref[2])):
                    assert np.array_equal(x[hit].view(np.uint32), y[hit].view(np.uint32))
                assert 0.05 < hit.mean() < 0.95
        for b in bufs.values():
            ctx.device_free(b)
    finally:
        ctx.close()


def test_noalloc_rejects_bad_arguments():
    ctx = abi.Context(0)
    try:
        s = scenes.cornell()
        h = ctx.upload(s)
        with pytest.raises(abi.SrtError):          # no BVH yet
            ctx.trace_noalloc(h, 4, 1, 1, 1, 1, 1, 1, 1)
        ctx.build_bvh(h)
        assert ctx.lib.srt_trace_rays_noalloc(ctx.h, h, -1, None, None, 0.0, None, 0, None, None, None, None) == -1
    finally:
        ctx.close()
