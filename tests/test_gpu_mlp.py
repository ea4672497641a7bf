"""GPU radiance MLP (BF16 MFMA, FP32 accumulation) against the float64
reference of tests/mlp_ref.py with BF16 rounding emulated where the kernels
round: forward outputs, gradients of every parameter, one Adam step in closed
form (Kingma & Ba 2015), determinism, checkpoint round trip, and that online
training reduces the loss."""
import os

import numpy as np
import pytest

import mlp_ref as M

pytestmark = pytest.mark.gpu


@pytest.fixture
def mctx(ctx):
    ctx.mlp_create(seed=3)
    return ctx


def data(n, seed):
    rng = np.random.default_rng(seed)
    u = rng.uniform(size=(n, 9)).astype(np.float32)
    t = (0.3 + 0.25 * np.sin(2 * np.pi * u[:, :3]) * np.cos(np.pi * u[:, 3:6]) + 0.2 * u[:, 6:9]).astype(np.float32)
    return u, t


@pytest.mark.parametrize("n", [1, 15, 16, 1000, 4099])
def test_forward_matches_reference(mctx, n):
    p = M.init_params(5)
    mctx.mlp_set_params(p)
    assert np.array_equal(mctx.mlp_params(), p.astype(np.float32))
    u, _ = data(n, n)
    y = mctx.mlp_infer(u)
    ref, _ = M.forward(p, u, emulate=True)
    scale = np.abs(ref).max() + 1e-6
    # BF16 activations: a value at a rounding boundary can round the other way
    assert np.abs(y - ref).max() <= 2e-2 * scale
    assert np.abs(y - ref).mean() <= 1e-3 * scale


def test_gradients_match_reference(mctx):
    p = M.init_params(7)
    mctx.mlp_set_params(p)
    u, t = data(8192, 1)
    loss, g = mctx.mlp_train_step(u, t, apply=False, want_grads=True)
    gr = M.backward(p, u, t, emulate=True)
    lr_ = M.loss(p, u, t, emulate=True)
    assert abs(loss - lr_) / lr_ < 1e-3
    assert np.array_equal(mctx.mlp_params(), p.astype(np.float32))      # apply=0 leaves weights
    o = 0
    for (r, c) in M.SHAPES:
        for size in (r * c, r):
            a, b = g[o:o + size], gr[o:o + size]
            assert np.linalg.norm(a - b) <= 2e-2 * np.linalg.norm(b) + 1e-9, (r, c, size)
            o += size


def test_adam_first_step_closed_form(mctx):
    p = M.init_params(9).astype(np.float32)
    mctx.mlp_set_params(p)
    mctx.mlp_set_hparams(lr=1e-3, beta1=0.9, beta2=0.999, eps=1e-8)
    u, t = data(4096, 2)
    _, g = mctx.mlp_train_step(u, t, apply=False, want_grads=True)
    mctx.mlp_train_step(u, t, apply=True)
    p1 = mctx.mlp_params()
    # t = 1: m_hat = g, v_hat = g^2 -> step lr * g / (|g| + eps)
    g64 = g.astype(np.float64)
    expect = p.astype(np.float64) - 1e-3 * g64 / (np.abs(g64) + 1e-8)
    assert np.allclose(p1, expect, rtol=0, atol=2e-6)


def test_training_is_deterministic_and_learns(ctx, tmp_path):
    u, t = data(1 << 14, 3)
    runs = []
    for rep in range(2):
        ctx.mlp_create(seed=11, lr=3e-3)
        losses = [ctx.mlp_train_step(u, t) for _ in range(150)]
        runs.append((losses, ctx.mlp_params()))
    assert np.array_equal(runs[0][1], runs[1][1])
    assert runs[0][0] == runs[1][0]
    losses = runs[0][0]
    assert losses[-1] < 0.2 * losses[0], (losses[0], losses[-1])
    # checkpoint round trip: same weights and the same next step
    path = str(tmp_path / "mlp.bin")
    ctx.mlp_save(path)
    l_a = ctx.mlp_train_step(u, t)
    p_a = ctx.mlp_params()
    ctx.mlp_create(seed=99)
    ctx.mlp_load(path)
    assert np.array_equal(ctx.mlp_params(), runs[0][1])
    l_b = ctx.mlp_train_step(u, t)
    assert l_a == l_b and np.array_equal(ctx.mlp_params(), p_a)


def test_integrator_hooks_train_and_query(ctx):
    """Online training from the tracer's own vertex-2 samples, then paths that
    stop at vertex 2 and take the MLP's prediction: finite, deterministic, and
    close to the plain tracer's converged mean on the Cornell box."""
    from srt_amd import abi, scenes
    s = scenes.cornell()
    h = ctx.upload(s)
    ctx.mlp_create(seed=1, lr=1e-3)
    with pytest.raises(abi.SrtError):
        c2 = abi.Context(0)
        h2 = c2.upload(s)
        try:
            c2.render(h2, s.camera, 32, 32, 1, 5, 4, 1, 1, 0, mlp_mode=1)     # no MLP on this context
        finally:
            c2.close()
    losses = []
    for f in range(40):
        loss, n = ctx.mlp_train_frame(h, s.camera, 128, 128, 4, 5, 4, seed=1000 + f, max_samples=1 << 15, steps=4)
        assert 0 < n <= 1 << 15
        losses.append(loss)
    assert np.isfinite(losses).all()
    assert np.mean(losses[-5:]) < np.mean(losses[:5])
    p_a = ctx.mlp_params()
    img = ctx.render(h, s.camera, 64, 64, 64, 5, 4, 1, 1, 7, mlp_mode=1)
    assert np.isfinite(img).all()
    plain = ctx.render(h, s.camera, 64, 64, 256, 5, 4, 1, 1, 8)
    assert abs(img[:, :3].mean() / plain[:, :3].mean() - 1) < 0.15
    # same training again from the same seed: identical weights
    ctx.mlp_create(seed=1, lr=1e-3)
    for f in range(40):
        ctx.mlp_train_frame(h, s.camera, 128, 128, 4, 5, 4, seed=1000 + f, max_samples=1 << 15, steps=4)
    assert np.array_equal(ctx.mlp_params(), p_a)
    ctx.release(h)


def test_query_paths_independent_of_chunking(ctx):
    """Mode-1 paths queue their vertex-2 queries per queue partition and the
    queries are compacted before inference: the image is the same for one
    chunk and for chunks of 1000 paths (ragged partitions, other query order)."""
    from srt_amd import scenes
    s = scenes.cornell()
    h = ctx.upload(s)
    ctx.mlp_create(seed=3, lr=1e-3)
    ctx.mlp_train_frame(h, s.camera, 64, 64, 2, 5, 4, seed=5, max_samples=1 << 12, steps=2)
    a = ctx.render(h, s.camera, 48, 40, 8, 5, 4, 1, 1, 9, mlp_mode=1)
    ctx.set_option("path_pool", 1000)
    b = ctx.render(h, s.camera, 48, 40, 8, 5, 4, 1, 1, 9, mlp_mode=1)
    ctx.set_option("path_pool", 1 << 27)
    ctx.release(h)
    assert np.isfinite(a).all()
    assert np.array_equal(a, b)
