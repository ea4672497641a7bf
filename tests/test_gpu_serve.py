"""Serving loop: requests of different scenes, resolutions, sample counts and
depths batched into shared path chunks render the same bits as each request
alone through srt_render; admission follows the HBM budget."""
import numpy as np
import pytest

from srt_amd import abi, scenes

pytestmark = pytest.mark.gpu


def _cam_variant(cam, dx):
    c = dict(cam)
    c["position"] = [cam["position"][0] + dx, cam["position"][1], cam["position"][2]]
    return c


@pytest.fixture(scope="module")
def served():
    ctx = abi.Context(0)
    cb = scenes.cornell()
    sp = scenes.spheres()
    hc = ctx.upload(cb)
    ctx.build_bvh(hc)
    hs = ctx.upload(sp)
    ctx.build_bvh(hs)
    ctx.mlp_create(seed=7)
    reqs = [
        (hc, cb.camera, 64, 48, 8, 5, 4, 1, 1, 11, 0),
        (hc, _cam_variant(cb.camera, 20.0), 32, 32, 16, 3, 2, 1, 0, 12, 0),
        (hs, sp.camera, 96, 54, 2, 8, 4, 1, 1, 13, 0),
        (hc, cb.camera, 40, 40, 4, 6, 4, 0, 0, 14, 0),
        (hs, _cam_variant(sp.camera, -30.0), 48, 27, 3, 4, 4, 1, 1, 15, 0),
        (hc, cb.camera, 32, 24, 4, 5, 4, 1, 1, 16, 1),      # radiance-MLP query at vertex 2
    ]
    alone = [ctx.render(h, cam, W, H, spp, d, rr, nee, mis, seed, mlp_mode=m)
             for (h, cam, W, H, spp, d, rr, nee, mis, seed, m) in reqs]
    yield ctx, reqs, alone
    ctx.close()


def _submit_all(ctx, reqs):
    out = []
    for (h, cam, W, H, spp, d, rr, nee, mis, seed, m) in reqs:
        out.append(ctx.submit(h, cam, W, H, spp, d, rr, nee, mis, seed, mlp_mode=m))
    return out


@pytest.mark.parametrize("pool", [3000, 20000, 0])
def test_mix_equals_alone(served, pool):
    ctx, reqs, alone = served
    ctx.serve_start(pool)
    try:
        tickets = _submit_all(ctx, reqs)
        for k, (t, img) in enumerate(tickets):
            res = ctx.wait(t)
            assert res["status"] == 2, (k, res)
            assert np.array_equal(img, alone[k]), f"request {k} differs in the mix (pool {pool})"
        st = ctx.serve_stats()
        assert st["served_paths"] == sum(W * H * spp for (_, _, W, H, spp, *_r) in reqs)
        if pool == 3000:
            # small chunks: requests on one scene share chunks
            assert max(ctx.poll(t)["chunks"] for t, _ in tickets) > 1
    finally:
        ctx.serve_stop()


def test_entry_points_refuse_while_serving(served):
    ctx, reqs, alone = served
    ctx.serve_start(0)
    try:
        with pytest.raises(abi.SrtError) as e:
            ctx.render(reqs[0][0], reqs[0][1], 8, 8, 1)
        assert e.value.code == -7
        t, img = ctx.submit(*reqs[0][:10])
        assert ctx.wait(t)["status"] == 2
        assert np.array_equal(img, alone[0])
    finally:
        ctx.serve_stop()
    # usable again after stop
    img = ctx.render(*reqs[0][:10])
    assert np.array_equal(img, alone[0])


def test_admission_by_budget(served):
    ctx, reqs, alone = served
    h, cam, W, H, spp, d, rr, nee, mis, seed, m = reqs[0]
    film = W * H * 28
    ctx.set_option("keep_build_scratch", 0)     # memory in use stays fixed while serving
    budget0 = ctx.lib.srt_hbm_used(ctx.h)
    ctx.set_option("hbm_budget", budget0 + film + film // 2)   # one request's film at a time
    ctx.serve_start(0)
    try:
        tickets = [ctx.submit(h, cam, W, H, spp, d, rr, nee, mis, seed) for _ in range(4)]
        big, _ = ctx.submit(h, cam, 1024, 1024, 1, d, rr, nee, mis, seed)   # 28 MiB film never fits
        res = [ctx.wait(t) for t, _ in tickets]
        rb = ctx.wait(big)
    finally:
        ctx.serve_stop()
        ctx.set_option("hbm_budget", 1 << 62)
        ctx.set_option("keep_build_scratch", 1)
    for (t, img), r in zip(tickets, res):
        assert r["status"] == 2
        assert np.array_equal(img, alone[0])
    # admitted one after another
    for a, b in zip(res, res[1:]):
        assert b["admit_ms"] >= a["done_ms"]
    assert rb["status"] == -4


def test_stop_without_drain_cancels_queued(served):
    ctx, reqs, alone = served
    h, cam, W, H, spp, d, rr, nee, mis, seed, m = reqs[0]
    ctx.serve_start(0)
    try:
        first, img = ctx.submit(h, cam, W, H, spp, d, rr, nee, mis, seed)
        rest = [ctx.submit(h, cam, 256, 256, 64, d, rr, nee, mis, seed + k)[0] for k in range(200)]
        assert ctx.wait(first)["status"] == 2
        assert np.array_equal(img, alone[0])
    finally:
        ctx.serve_stop(drain=False)
    # the context is usable again and holds no request films
    used = ctx.lib.srt_hbm_used(ctx.h)
    assert np.array_equal(ctx.render(*reqs[0][:10]), alone[0])
    assert ctx.lib.srt_hbm_used(ctx.h) == used


def test_pool_outside_32bit_slots_is_refused(served):
    """Path slots are 32-bit: a serving pool outside [64, 2^31 - 1] is an
    error at srt_serve_start (0 takes the context's pool), and the context
    stays idle and usable."""
    ctx, reqs, alone = served
    for bad in (1 << 31, (1 << 31) + 5, 63, -1):
        with pytest.raises(abi.SrtError):
            ctx.serve_start(bad)
        assert ctx.lib.srt_set_option(ctx.h, b"profile", 0) == 0      # not serving
    ctx.serve_start((1 << 31) - 1)
    try:
        t, img = ctx.submit(*reqs[0][:10])
        assert ctx.wait(t)["status"] == 2
        assert np.array_equal(img, alone[0])
    finally:
        ctx.serve_stop(drain=True)


def test_request_arriving_mid_chunk_joins_the_next_chunk(served):
    """A large request is cut into chunks; a small request submitted while it
    runs is admitted between two of its chunks, runs in one chunk ahead of the
    large one's remaining paths (its due time, submit time + remaining paths
    at 2 Gpaths/s, is earlier), finishes first, and both images equal their
    solo renders."""
    ctx, reqs, alone = served
    hs = reqs[2][0]                                  # the 1M-triangle scene
    cam = reqs[2][1]
    big = (hs, cam, 1024, 1024, 64, 4, 4, 1, 1, 21)   # 67.1 M paths
    small = reqs[0]
    big_alone = ctx.render(*big)
    pool = 1 << 22
    ctx.serve_start(pool)
    try:
        tb, img_b = ctx.submit(*big)
        # wait until the large request has run at least one chunk and is still running
        import time
        t_end = time.time() + 30
        while ctx.poll(tb)["chunks"] < 1 and time.time() < t_end:
            time.sleep(0.0002)
        assert ctx.poll(tb)["status"] == 1
        ts, img_s = ctx.submit(*small[:10])
        rs = ctx.wait(ts)
        rb = ctx.wait(tb)
    finally:
        ctx.serve_stop()
    assert rs["status"] == 2 and rb["status"] == 2
    assert rs["chunks"] == 1
    assert rb["chunks"] == (64 * 1024 * 1024 + pool - 1) // pool
    assert rs["done_ms"] < rb["done_ms"]
    assert rs["admit_ms"] > rb["admit_ms"]
    assert np.array_equal(img_s, alone[0])
    assert np.array_equal(img_b, big_alone)


def test_each_scene_queries_its_own_mlp(served):
    """Radiance MLPs bound to scenes (srt_mlp_select): a scene's mlp_mode-1
    render uses its own MLP, a scene without one the context's own, alone and
    in the serving mix; releasing a scene frees its MLP."""
    ctx, reqs, alone = served
    hc = reqs[0][0]
    cb = scenes.cornell()
    h2 = ctx.upload(cb)
    ctx.build_bvh(h2)
    try:
        ctx.mlp_select(h2)
        ctx.mlp_create(seed=99)
        own = ctx.render(h2, cb.camera, 32, 24, 4, 5, 4, 1, 1, 16, mlp_mode=1)
        ctx.mlp_select(0)
        shared = ctx.render(hc, cb.camera, 32, 24, 4, 5, 4, 1, 1, 16, mlp_mode=1)
        assert np.array_equal(shared, alone[5])         # the context's own MLP (seed 7)
        assert not np.array_equal(own, shared)          # same scene data, different MLP
        ctx.serve_start(0)
        try:
            t1, i1 = ctx.submit(h2, cb.camera, 32, 24, 4, 5, 4, 1, 1, 16, mlp_mode=1)
            t2, i2 = ctx.submit(hc, cb.camera, 32, 24, 4, 5, 4, 1, 1, 16, mlp_mode=1)
            assert ctx.wait(t1)["status"] == 2 and ctx.wait(t2)["status"] == 2
        finally:
            ctx.serve_stop()
        assert np.array_equal(i1, own) and np.array_equal(i2, shared)
        used = ctx.lib.srt_hbm_used(ctx.h)
    finally:
        ctx.release(h2)
    assert ctx.lib.srt_hbm_used(ctx.h) < used          # the scene and its MLP are freed
    assert np.array_equal(ctx.render(hc, cb.camera, 32, 24, 4, 5, 4, 1, 1, 16, mlp_mode=1), shared)
