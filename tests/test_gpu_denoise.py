"""GPU parity of the learned denoiser: BF16 input features bit-exact with the
oracle, MFMA forward and weight gradients against the FP64 oracle, gradients
deterministic, training on tracer-generated pairs reduces the error, and the
trained network beats the raw 1-spp estimate (and, for the soft shadows, the
non-learned 5x5 box filter it adds its correction to)."""
import numpy as np
import pytest

from luz_amd import denoiser as dnm
from oracle import denoise as odn
from tests.helpers import denoise_rmse

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def frame(ctx):
    return dnm.make_frame(ctx, 3, crop=(96, 64), target_spp=64, focus="instance")


def _params(seed):
    rng = np.random.default_rng(seed)
    mats = []
    for shape in odn.SHAPES:
        fan = shape[1] if len(shape) == 2 else 64
        if len(shape) == 2:
            mats.append(rng.normal(0.0, 1.0 / np.sqrt(fan), shape))
        else:
            mats.append(rng.uniform(-0.1, 0.2, shape))
    return odn.pack(mats).astype(np.float32)


def _all_samples(fr):
    L = max(fr.num_lights, 1)
    px = np.repeat(fr.foreground.astype(np.int64), L)
    li = np.tile(np.arange(L), len(fr.foreground))
    return px, li


def test_features_bit_exact(ctx, frame):
    px, li = _all_samples(frame)
    print(f"{len(px)} samples of a {frame.width}x{frame.height} frame", flush=True)
    got = dnm.features(ctx, frame, (px | (li << 24)).astype(np.uint32))
    want = odn.features(frame.host(), px, li)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    bg = np.setdiff1d(np.arange(frame.width * frame.height), frame.foreground)[:5]
    if len(bg):
        assert not dnm.features(ctx, frame, bg.astype(np.uint32)).any()


def _check_forward(ctx, fr):
    """Inference on every pixel of `fr` against the FP64 oracle's forward pass
    on the same weights: visibility and AO within 2e-2 (max) and 3e-3 (RMS),
    background pixels 0."""
    dn = dnm.Denoiser(ctx, 0, seed=1)
    dn.params = _params(7)
    n, L = fr.width * fr.height, fr.num_lights
    ao_o, vis_o = ctx.buffer(n * 4), ctx.buffer(max(L, 1) * n * 4)
    dn.denoise(fr, ao_o, vis_o)
    ao = ao_o.download(np.float32, n).astype(np.float64)
    vis = vis_o.download(np.float32, (max(L, 1), n)).astype(np.float64)
    px, li = _all_samples(fr)
    host = fr.host()
    y, _, _ = odn.forward(odn.device_weights(dn.params), odn.features(host, px, li))
    y = y + odn.box_base(host, px, li)
    want_vis = np.clip(y[:, 1], 0, 1)
    want_ao = np.clip(y[:, 0].reshape(-1, max(L, 1)).mean(1), 0, 1)
    fg = fr.foreground
    err_a = ao[fg] - want_ao
    err_v = vis[li, px] - want_vis if L else np.zeros(1)
    print(f"{L} lights: y0 in [{y[:, 0].min():.3f}, {y[:, 0].max():.3f}], y1 in [{y[:, 1].min():.3f}, "
          f"{y[:, 1].max():.3f}]; vis err max {np.abs(err_v).max():.2e} rms {np.sqrt(np.mean(err_v ** 2)):.2e}; "
          f"ao err max {np.abs(err_a).max():.2e}", flush=True)
    assert np.abs(err_v).max() < 2e-2 and np.sqrt(np.mean(err_v ** 2)) < 3e-3
    assert np.abs(err_a).max() < 2e-2 and np.sqrt(np.mean(err_a ** 2)) < 3e-3
    bgm = np.ones(n, bool)
    bgm[fg] = False
    assert not ao[bgm].any()
    if L:
        assert not vis[:, bgm].any()
    dn.close()


def test_forward_matches_fp64_oracle(ctx, frame):
    _check_forward(ctx, frame)


@pytest.mark.parametrize("nl", [0, 1, 2, 4, 5])
def test_forward_any_light_count(ctx, frame, nl):
    # inference takes the lights in groups of three: the frame's first nl light
    # planes exercise a last group of one or two lights, and no lights at all
    n = frame.width * frame.height
    sh = frame.shadow.download(np.uint8, (frame.num_lights, n))
    planes = ctx.to_device(np.ascontiguousarray(sh[:nl])) if nl else frame.shadow
    sub = dnm.Frame(ctx, frame.width, frame.height, frame.position, frame.normal, frame.cam_pos, nl,
                    planes, frame.shadow_samples, frame.ao, frame.ao_samples)
    sub.foreground = frame.foreground
    _check_forward(ctx, sub)


def test_gradient_matches_fp64_oracle_and_is_deterministic(ctx, frame):
    N = 4096
    dn = dnm.Denoiser(ctx, N, seed=1)
    dn.params = _params(9)
    rng = np.random.default_rng(4)
    px = frame.foreground[rng.integers(0, len(frame.foreground), N)].astype(np.int64)
    li = rng.integers(0, frame.num_lights, N)
    buf = ctx.to_device((px | (li << 24)).astype(np.uint32))
    grads = []
    for _ in range(2):
        dn.accumulate(frame, buf, N, 1.0 / N)
        g, sse = dn.gradient(host=True, loss=True)
        grads.append(g)
    np.testing.assert_array_equal(grads[0].view(np.uint32), grads[1].view(np.uint32))
    w = odn.device_weights(dn.params)
    host = frame.host()
    X = odn.features(host, px, li)
    # the network's outputs are added to the box base: its targets are the residuals
    T = np.stack([frame.target_ao_host[px], frame.target_vis_host[li, px]], 1).astype(np.float64)
    T = T - odn.box_base(host, px, li)
    want = odn.backward(w, X, T, scale=1.0 / N)
    y, _, _ = odn.forward(w, X)
    want_sse = float(np.sum((y - T) ** 2))
    print(f"sse gpu {sse:.6f} oracle {want_sse:.6f}", flush=True)
    assert abs(sse - want_sse) <= 1e-2 * want_sse
    o = 0
    for k, shape in enumerate(odn.SHAPES):
        m = int(np.prod(shape))
        a, b = grads[0][o:o + m].astype(np.float64), want[o:o + m]
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        print(f"param block {k} {shape}: rel err {rel:.2e} cos {cos:.6f}", flush=True)
        assert rel < 2e-2 and cos > 0.9995
        o += m
    dn.close()


def test_training_on_tracer_pairs_reduces_error(ctx):
    frames = [dnm.make_frame(ctx, 200 + i, crop=(64, 64), target_spp=64, focus="instance") for i in range(16)]
    held = dnm.make_frame(ctx, 300, crop=(128, 128), target_spp=128, focus="instance")
    print(f"held-out frame: {len(held.hard)} noisy samples", flush=True)
    dn = dnm.Denoiser(ctx, 8 * 2048, seed=0)
    before = denoise_rmse(ctx, dn, held)
    hist = dnm.train(ctx, dn, frames, steps=300, frames_per_step=8, samples_per_frame=2048, lr=3e-3,
                     log_every=50, log=lambda m: print(m, flush=True))
    after = denoise_rmse(ctx, dn, held)
    print(f"held-out before {before}\nafter {after}", flush=True)
    assert hist[-1][1] < 0.5 * hist[0][1]
    assert after["vis"]["mlp"] < after["vis"]["raw"] and after["vis"]["mlp"] < 0.6 * before["vis"]["mlp"]
    assert after["ao"]["mlp"] < before["ao"]["mlp"]
    # a crop without occluders in AO range has an exact 1-spp AO (raw error 0)
    assert after["ao"]["mlp"] < (after["ao"]["raw"] if after["ao"]["raw"] > 0.005 else 0.05)
    dn.close()


@pytest.mark.skipif(not dnm.WEIGHTS.exists(), reason="no trained weights committed")
def test_trained_network_beats_unlearned_estimates(ctx):
    dn = dnm.Denoiser(ctx, 0)
    dn.load(dnm.WEIGHTS)
    for seed in (900001, 900002):
        fr = dnm.make_frame(ctx, seed, crop=(256, 256), target_spp=256, focus="instance")
        r = denoise_rmse(ctx, dn, fr)
        print(f"scene {seed}: {r}", flush=True)
        assert r["vis"]["mlp"] < r["vis"]["box"] and r["vis"]["mlp"] < r["vis"]["raw"]
        # AO: below the raw estimate and within the 0.01 target; the box filter it corrects can
        # be closer on a crop with little occlusion (on the whole 1080p frame the network wins)
        assert r["ao"]["mlp"] < r["ao"]["raw"] and r["ao"]["mlp"] < 0.01
    dn.close()


def test_shade_fractions_equals_shade_on_exact_fractions(ctx, frame):
    """With 1-spp counts the fractions are exactly 0 or 1, so lighting from
    fractions must reproduce lighting from counts bit for bit."""
    n, L = frame.width * frame.height, frame.num_lights
    g = frame.gbuffer.desc()
    sb = frame.scene_block
    for l in range(L):
        sb.lights[l].num_shadow_samples = 1
    sb.ao_num_samples = 1
    # the G-buffer's material planes were released: give it neutral ones
    planes = [ctx.to_device(np.full(n, v, np.uint32)) for v in (0xFF8040C0, 0xFFFF8060, 0xFF000000)]
    g.albedo, g.material, g.emission = (p.ptr for p in planes)
    out_c, out_f = ctx.buffer(n * 16), ctx.buffer(n * 16)
    ctx.shade(g, sb, frame.shadow, frame.ao, out_c)
    sh = frame.shadow.download(np.uint8, (L, n)).astype(np.float32)
    vis = ctx.to_device(1.0 - sh)
    aof = ctx.to_device(frame.ao.download(np.uint8, n).astype(np.float32))
    ctx.shade_fractions(g, sb, vis, aof, out_f)
    np.testing.assert_array_equal(out_f.download(np.uint32, (n, 4)), out_c.download(np.uint32, (n, 4)))
