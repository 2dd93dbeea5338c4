"""Learned radiance cache on the GPU: the hash-grid encoding, the fused MFMA
MLP forward and backward and the Adam step against a float reference
(tests/nrc_reference.py), training convergence on a smooth function, and the
learned texel tracer's plumbing against the plain tracer."""

from __future__ import annotations

import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

from tests import nrc_reference as ref  # noqa: E402

BBOX_MIN = np.array([-1.0, 0.0, -1.0])
BBOX_MAX = np.array([1.0, 2.0, 1.0])


def _say(msg: str) -> None:
    print(f"[gpu-nrc] {msg}", flush=True)
    sys.stdout.flush()


def _cache(device, **kw):
    from nan_amd.radiance_cache import RadianceCache, RadianceCacheConfig

    return RadianceCache(device, BBOX_MIN, BBOX_MAX, RadianceCacheConfig(**kw))


def _inputs(n, seed, device):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(BBOX_MIN, BBOX_MAX, size=(n, 3))
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    alb = rng.uniform(0.05, 0.9, size=(n, 3))
    t = lambda a: torch.from_numpy(a.astype(np.float32)).to(device.torch_device)  # noqa: E731
    return (pos, nrm, alb), (t(pos), t(nrm), t(alb))


def _randomize(cache, seed, table_scale=1.0):
    rng = np.random.default_rng(seed)
    w, t = cache.get_params()
    w = rng.uniform(-0.3, 0.3, size=w.shape).astype(np.float32)
    t = rng.uniform(-table_scale, table_scale, size=t.shape).astype(np.float32)
    cache.set_params(w, t)
    return w, t


@pytest.mark.parametrize("n", [1, 63, 1000, 70_001])
def test_encoding_matches_reference(device, n):
    _say(f"encode n={n}")
    cache = _cache(device, log2_table_size=14, base_resolution=4.0, max_resolution=300.0)
    _, table = _randomize(cache, 1)
    (pos, nrm, alb), dev_in = _inputs(n, 2, device)
    got = cache.encode(*dev_in).cpu().numpy()
    want = ref.encode(pos, nrm, alb, table, BBOX_MIN, BBOX_MAX, 4.0, 300.0, 14)
    want16 = want.astype(np.float16).astype(np.float64)
    # both sides round to FP16 once; FP32 vs FP64 interpolation may flip that rounding
    np.testing.assert_allclose(got, want16, rtol=2e-3, atol=2e-3)
    assert np.all(got[:, 38:] == 0.0)
    # the levels span dense and hashed indexing
    _, dense = ref.level_grid(16, 14, 4.0, 300.0)
    assert any(dense) and not all(dense)


@pytest.mark.parametrize("n", [64, 1000, 65_536])
def test_mlp_forward_matches_reference(device, n):
    _say(f"forward n={n}")
    cache = _cache(device)
    w, _ = _randomize(cache, 3, table_scale=0.5)
    _, dev_in = _inputs(n, 4, device)
    enc = cache.encode(*dev_in)
    got = cache.forward_raw(*dev_in).cpu()
    want, _, _ = ref.mlp(enc.cpu(), w)
    want = want[:, :4].detach()
    scale = want.abs().max().item()
    assert scale > 1e-2
    err = (got - want).abs().max().item()
    assert err <= 2e-3 * scale + 1e-4, f"max |gpu - ref| = {err} (scale {scale})"


@pytest.mark.parametrize("n", [64, 5000, 65_536])
def test_mlp_backward_matches_autograd(device, n):
    _say(f"backward n={n}")
    cache = _cache(device)
    w, _ = _randomize(cache, 5, table_scale=0.5)
    _, dev_in = _inputs(n, 6, device)
    rng = np.random.default_rng(7)
    target = torch.from_numpy(rng.uniform(0.0, 2.0, size=(n, 3)).astype(np.float32)).to(device.torch_device)
    enc = cache.encode(*dev_in)
    wg, ig = cache.backward_raw(*dev_in, target)
    out, x_leaf, ws = ref.mlp(enc.cpu(), w)
    # the kernels round the unnormalised per-record gradients to FP16 and
    # divide by 3n at the end; so does the reference
    ref.relative_l2_sum(out, target.cpu()).backward()
    scale = 1.0 / (3.0 * n)
    want_w = torch.cat([x.grad.reshape(-1) for x in ws]).numpy() * scale
    got_w = wg.cpu().numpy()
    rel = np.linalg.norm(got_w - want_w) / np.linalg.norm(want_w)
    assert rel < 1e-2, f"weight gradient relative error {rel}"
    for k, (o, i) in enumerate(ref.LAYER_SHAPES):  # every layer, not just the largest
        off = sum(a * b for a, b in ref.LAYER_SHAPES[:k])
        g, h = got_w[off:off + o * i], want_w[off:off + o * i]
        assert np.linalg.norm(g - h) <= 2e-2 * np.linalg.norm(h) + 1e-9, f"layer {k}"
    got_i = ig.cpu().numpy()
    want_i = x_leaf.grad.numpy() * scale
    rel_i = np.linalg.norm(got_i - want_i) / np.linalg.norm(want_i)
    assert rel_i < 1e-2, f"input gradient relative error {rel_i}"


def test_backward_is_deterministic(device):
    _say("backward determinism")
    cache = _cache(device)
    _randomize(cache, 8, table_scale=0.5)
    _, dev_in = _inputs(20_000, 9, device)
    target = torch.rand((20_000, 3), device=device.torch_device)
    a, _ = cache.backward_raw(*dev_in, target)
    b, _ = cache.backward_raw(*dev_in, target)
    assert torch.equal(a, b)


def test_first_adam_step_moves_by_learning_rate(device):
    _say("adam step")
    lr = 1e-2
    cache = _cache(device, learning_rate=lr, epsilon=1e-15)
    w0, t0 = _randomize(cache, 10, table_scale=0.5)
    _, dev_in = _inputs(4096, 11, device)
    target = torch.rand((4096, 3), device=device.torch_device) * 2.0
    g, _ = cache.backward_raw(*dev_in, target)
    g = g.cpu().numpy()
    cache.train_step(*dev_in, target)
    w1, t1 = cache.get_params()
    big = np.abs(g) > 1e-7
    assert big.sum() > 1000
    # step 1 of Adam: m_hat = g, v_hat = g^2, so every weight moves by -lr * sign(g)
    np.testing.assert_allclose((w1 - w0)[big], -lr * np.sign(g[big]), rtol=0, atol=1e-6)
    # lazy table update: untouched entries keep their values, touched ones move by lr
    moved = t1 != t0
    assert moved.any() and (~moved).any()
    step = np.abs((t1 - t0)[moved])
    assert step.max() <= lr * 1.001 + 1e-6  # |m_hat| / (sqrt(v_hat) + eps) <= 1
    assert abs(np.median(step) - lr) < 1e-4 * lr


def test_training_fits_a_smooth_function(device):
    """Online training (fresh records every step, as the renderer trains)
    on a smooth function of position and albedo."""
    _say("training convergence")
    cache = _cache(device, max_resolution=128.0)
    n = 16_384

    def f(pos, alb):
        return np.stack([np.sin(3 * pos[:, 0]) + 1.2, np.cos(2 * pos[:, 1]) + 1.1, pos[:, 2] ** 2 + 0.5 * alb[:, 0]],
                        axis=1)

    losses = []
    for step in range(300):
        (pos, nrm, alb), dev_in = _inputs(n, 1000 + step, device)
        target = torch.from_numpy(f(pos, alb).astype(np.float32)).to(device.torch_device)
        losses.append(cache.train_step(*dev_in, target).item())
    (pos2, nrm2, alb2), dev_in2 = _inputs(4096, 13, device)
    f2 = f(pos2, alb2)
    pred = cache.infer(*dev_in2).cpu().numpy()
    rel = np.sqrt(np.mean((pred - f2) ** 2) / np.mean(f2 ** 2))
    _say(f"loss {losses[0]:.4f} -> {losses[-1]:.5f}, held-out relative RMSE {rel:.4f}")
    assert losses[-1] < 0.01 * losses[0]
    assert rel < 0.05


@pytest.mark.parametrize("spp", [1, 2])
def test_learned_tracer_with_zero_cache_is_the_one_bounce_plain_tracer(device, cornell, spp):
    """With all weights zero the cache predicts 0, and on the Cornell box (no
    sun, so a one-bounce suffix is empty) the learned form estimates what the
    plain tracer with max_bounces=1 does.  At 1 sample per texel the front
    slots see the same random numbers and must agree sample by sample (back
    slots face out of the box and stay black); with more samples the plain
    tracer's suffix draws shift the per-texel stream, so the estimates agree
    in the mean only."""
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say(f"learned vs plain one-bounce spp={spp}")
    node, layout, _ = cornell
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=spp, max_bounces=1, mode="megakernel")
    cache = RadianceCache.for_scene(device, sc)
    w, _ = cache.get_params()
    enc = device.create_command_encoder()
    slots = layout.total_payload_count
    plain = torch.zeros((slots, 4), dtype=torch.float32, device=device.torch_device)
    learned = torch.zeros_like(plain)
    counters = torch.zeros((8,), dtype=torch.int64, device=device.torch_device)
    for it in range(3):
        cache.set_params(np.zeros_like(w))
        tracer.execute(enc, plain, it, reset=(it == 0))
        # train_samples=0: no training, the zero cache stays zero
        tracer.execute_learned(enc, learned, it, cache, reset=(it == 0), train_samples=0, emitter_nee=False,
                               ramped_mean=False, counters=counters)
    p, q = plain.cpu().numpy(), learned.cpu().numpy()
    assert np.array_equal(p[:, 3].view(np.uint32), q[:, 3].view(np.uint32))
    assert (p[:, 3].view(np.uint32) == 3).all()
    if spp == 1:
        np.testing.assert_allclose(q[:, :3], p[:, :3], rtol=2e-6, atol=1e-7)
    else:
        mp, mq = p[:, :3].astype(np.float64).mean(), q[:, :3].astype(np.float64).mean()
        assert mp > 0.1 and abs(mq - mp) < 0.02 * mp
    c = counters.cpu().numpy()
    assert c[4] == 3 * slots  # slot updates
    assert c[0] == 3 * slots * spp  # one extend ray per texel sample


def test_learned_tracer_trains_and_converges(device, cornell):
    """After 64 frames the learned form (light sampling, ramped mean, four
    Adam steps per frame) is unbiased in the mean and within 1.5x of the plain
    tracer's RMSE against a 512-frame plain bake (measured: about 1.1x)."""
    from nan_amd.radiance_cache import RadianceCache, RadianceCacheConfig
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say("learned tracer convergence")
    node, layout, _ = cornell
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode="megakernel")
    enc = device.create_command_encoder()
    slots = layout.total_payload_count
    bake = torch.zeros((slots, 4), dtype=torch.float32, device=device.torch_device)
    for it in range(512):
        tracer.execute(enc, bake, 100_000 + it, reset=(it == 0))
    cache = RadianceCache.for_scene(device, sc, RadianceCacheConfig(minibatches=4))
    learned = torch.zeros_like(bake)
    plain = torch.zeros_like(bake)
    for it in range(64):
        tracer.execute_learned(enc, learned, it, cache, reset=(it == 0))
        tracer.execute(enc, plain, it, reset=(it == 0))
    b = bake[:, :3].cpu().numpy().astype(np.float64)
    lq = learned[:, :3].cpu().numpy().astype(np.float64)
    pq = plain[:, :3].cpu().numpy().astype(np.float64)
    assert np.isfinite(lq).all()
    assert (learned[:, 3].cpu().numpy().view(np.uint32) == 64).all()
    rmse_l = np.sqrt(np.mean((lq - b) ** 2))
    rmse_p = np.sqrt(np.mean((pq - b) ** 2))
    _say(f"64 frames: RMSE learned {rmse_l:.4f} plain {rmse_p:.4f}; means {lq.mean():.4f} {pq.mean():.4f} "
         f"{b.mean():.4f}")
    # the cache carries the bounces after the first: total energy must match
    assert abs(lq.mean() - b.mean()) < 0.05 * b.mean()
    assert rmse_l < 1.5 * rmse_p


def _small_sunlit():
    from nan_amd.mesh_colors import MeshColorsLayout
    from nan_amd.scenes import synthetic_scene

    node = synthetic_scene(2, terrain_quads=32, box_grid=4, spheres=10, extent=20.0)
    layout = MeshColorsLayout.build(node, texels_per_unit=4.0, min_resolution=1, max_resolution=8,
                                    max_total_texels=2_000_000)
    return node, layout


def test_texel_query_form_with_zero_cache_traces_no_bounce(device):
    """The texel-query form with an untrained zero cache is the sun term
    alone: no extend ray per texel, every slot updated once per frame, and
    every value within [0, sun illuminance]."""
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say("texel-query form, zero cache")
    node, layout = _small_sunlit()
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode="megakernel")
    cache = RadianceCache.for_scene(device, sc)
    w, _ = cache.get_params()
    cache.set_params(np.zeros_like(w))
    out = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    counters = torch.zeros((8,), dtype=torch.int64, device=device.torch_device)
    tracer.execute_learned(device.create_command_encoder(), out, 0, cache, reset=True, train_samples=0,
                           ramped_mean=False, texel_query=True, counters=counters)
    c = counters.cpu().numpy()
    assert c[0] == 0 and c[4] == layout.total_payload_count
    v = out.cpu().numpy()
    sun = node.lighting.pack()
    illum = sun[4:7] * sun[8:11]
    assert (v[:, :3] >= 0.0).all() and (v[:, :3] <= illum * (1 + 1e-6)).all()
    assert (v[:, :3].max(axis=1) > 0.0).mean() > 0.3  # the lit part of the scene


def test_learned_forms_refuse_russian_roulette_scenes(device):
    """The learned forms' suffix paths have no Russian roulette: a scene with
    it on is refused with an error, and leaves the output untouched."""
    from nan_amd import _native
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say("learned forms refuse roulette scenes")
    node, layout = _small_sunlit()
    node.russian_roulette = True
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode="megakernel")
    cache = RadianceCache.for_scene(device, sc)
    out = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
    for texel_query in (False, True):
        with pytest.raises(_native.NanError, match="Russian roulette"):
            tracer.execute_learned(device.create_command_encoder(), out, 0, cache, reset=True, train_samples=0,
                                   texel_query=texel_query)
    torch.cuda.synchronize()
    assert not out.any()


def test_texel_query_form_converges_faster_than_plain(device):
    """On a sunlit scene the texel-query form (cache queried at the texel, no
    bounce ray) stays unbiased in the mean (within 3% of a 256-frame plain
    bake) and after 32 frames is closer to the bake than the plain tracer
    after 4 frames."""
    from nan_amd.radiance_cache import RadianceCache, RadianceCacheConfig
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say("texel-query form convergence")
    node, layout = _small_sunlit()
    sc = Scene(device, node)
    tracer = TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3)
    enc = device.create_command_encoder()
    slots = layout.total_payload_count
    bake = torch.zeros((slots, 4), dtype=torch.float32, device=device.torch_device)
    for it in range(256):
        tracer.execute(enc, bake, 100_000 + it, reset=(it == 0))
    plain = torch.zeros_like(bake)
    for it in range(4):
        tracer.execute(enc, plain, it, reset=(it == 0))
    cache = RadianceCache.for_scene(device, sc, RadianceCacheConfig(minibatches=8))
    learned = torch.zeros_like(bake)
    for it in range(32):
        tracer.execute_learned(enc, learned, it, cache, reset=(it == 0), train_samples=131072, texel_query=True)
    b = bake[:, :3].cpu().numpy().astype(np.float64)
    lq = learned[:, :3].cpu().numpy().astype(np.float64)
    pq = plain[:, :3].cpu().numpy().astype(np.float64)
    assert np.isfinite(lq).all()
    rmse_l = np.sqrt(np.mean((lq - b) ** 2))
    rmse_p4 = np.sqrt(np.mean((pq - b) ** 2))
    _say(f"RMSE texel-query@32 {rmse_l:.4f}, plain@4 {rmse_p4:.4f}; means {lq.mean():.4f} {b.mean():.4f}")
    assert abs(lq.mean() - b.mean()) < 0.03 * b.mean()
    assert rmse_l < rmse_p4


@pytest.mark.parametrize("texel_query", [False, True])
def test_learned_forms_on_the_wavefront_match_the_per_texel_kernels(device, texel_query):
    """In wavefront mode the learned forms run on the tracer's queues; their
    training records and (with a fixed, untrained cache) their payloads are
    bit-identical to the per-texel kernels of megakernel mode."""
    from nan_amd.radiance_cache import RadianceCache
    from nan_amd.scene import Scene
    from nan_amd.texture_space_path_tracer import TextureSpacePathTracer

    _say(f"learned wavefront vs per-texel (texel_query={texel_query})")
    node, layout = _small_sunlit()
    sc = Scene(device, node)
    tracers = {m: TextureSpacePathTracer(device, sc, layout, samples_per_texel=1, max_bounces=3, mode=m)
               for m in ("megakernel", "wavefront")}
    n = min(60_000, layout.total_surface_texels)
    recs = {m: t.learned_training_records(5, n, texel_query=texel_query) for m, t in tracers.items()}
    for k in ("pos", "nrm", "alb", "target", "valid"):
        a, b = recs["megakernel"][k].cpu().numpy(), recs["wavefront"][k].cpu().numpy()
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), k
    valid = recs["wavefront"]["valid"].cpu().numpy()
    assert valid.mean() > (0.99 if texel_query else 0.3)
    cache = RadianceCache.for_scene(device, sc)
    w, t = cache.get_params()
    rng = np.random.default_rng(11)
    cache.set_params((rng.standard_normal(w.shape) * 0.1).astype(w.dtype),
                     (rng.standard_normal(t.shape) * 0.1).astype(t.dtype))
    outs = {}
    for m, tr in tracers.items():
        out = torch.zeros((layout.total_payload_count, 4), dtype=torch.float32, device=device.torch_device)
        enc = device.create_command_encoder()
        for it in range(2):
            tr.execute_learned(enc, out, it, cache, reset=(it == 0), train_samples=0, texel_query=texel_query)
        outs[m] = out.cpu().numpy()
    np.testing.assert_array_equal(outs["megakernel"].view(np.uint32), outs["wavefront"].view(np.uint32))
    assert outs["wavefront"][:, :3].max() > 0.0


@pytest.mark.parametrize("case", ["uniform_16k", "uniform_64k", "clustered_16k"])
def test_table_gradient_fp16_sums_against_float64(device, case):
    """The hash-table gradient as the Adam step receives it (one packed FP16
    atomic per sample corner) against the same per-sample gradients summed
    in float64.  Training records drawn over the scene share a coarse dense
    entry among tens to hundreds of corners: every level stays within 3e-3
    of the exact sum.  Records clustered in a 0.05-unit cube share a few
    entries among thousands: FP16 sums then lose up to a fifth of a hot
    entry's gradient (measured 0.17 at worst), the price of one atomic per
    corner at this GPU's 20.9 G atomics/s; the bound below states it."""
    n = 65536 if case == "uniform_64k" else 16384
    _say(f"table gradient {case}")
    cache = _cache(device)
    rng = np.random.default_rng(3)
    _randomize(cache, 4, table_scale=0.5)
    if case.startswith("uniform"):
        pos = rng.uniform(BBOX_MIN, BBOX_MAX, size=(n, 3))
    else:
        pos = np.clip(np.array([0.1, 0.9, -0.2]) + rng.uniform(-0.025, 0.025, size=(n, 3)), BBOX_MIN, BBOX_MAX)
    nrm = rng.normal(size=(n, 3))
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    alb = rng.uniform(0.05, 0.9, size=(n, 3))
    tgt = rng.uniform(0.0, 2.0, size=(n, 3))
    t = lambda a: torch.from_numpy(a.astype(np.float32)).to(device.torch_device)  # noqa: E731
    args = (t(pos), t(nrm), t(alb), t(tgt))
    _, ig = cache.backward_raw(*args)
    got = cache.table_grad_raw(*args).cpu().numpy().astype(np.float64)
    want = ref.table_gradient(pos, ig.cpu().numpy(), BBOX_MIN, BBOX_MAX, 16, got.shape[1], 16.0, 512.0, 19)
    rel = np.array([np.linalg.norm(got[l] - want[l]) / np.linalg.norm(want[l]) for l in range(16)])
    _say(f"{case}: per-level relative error max {rel.max():.2e}")
    assert rel.max() < (3e-3 if case.startswith("uniform") else 0.3), rel
    # no entry outside the touched ones receives anything
    assert np.all(got[want == 0.0] == 0.0)
