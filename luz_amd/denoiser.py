"""Learned denoiser of 1-spp ambient occlusion and soft shadows.

Wraps the denoiser C ABI (luzrt_denoiser_*), produces training pairs with
the GPU tracer itself (1-spp inputs, 256-spp targets from two independent
128-spp passes), trains with the hand-written BF16 MFMA kernels, and
evaluates against the plain multi-spp tracer.

    python -m luz_amd.denoiser train --scenes 2000 --steps 4000 --out luz_amd/data/denoiser.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import math
import time
from pathlib import Path

import numpy as np

from . import api
from .api import lib, _P

N_PARAMS = 22274
WEIGHTS = Path(__file__).resolve().parent / "data" / "denoiser.npy"
# AO radius (aoMax) of the denoiser's frames.  The AO ray origin is lifted by
# 0.01 x the camera distance along the normal (light.frag:229-231); at the
# sphere fields' camera distances (100-250) that lift exceeds a radius of 1 and
# AO would be 1 on every pixel, so these frames look for occluders out to 16.
AO_RADIUS = 16.0


class DnFrame(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("position", C.c_void_p),
                ("normal", C.c_void_p), ("cam_pos", C.c_float * 3), ("num_lights", C.c_uint32),
                ("shadow_counts", C.c_void_p), ("shadow_samples", C.c_uint32),
                ("ao_counts", C.c_void_p), ("ao_samples", C.c_uint32)]


_sigs = {
    "luzrt_denoiser_create": (C.c_int, [_P, C.c_uint32, C.c_uint64, C.POINTER(_P)]),
    "luzrt_denoiser_destroy": (C.c_int, [_P, _P]),
    "luzrt_denoiser_get_params": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "luzrt_denoiser_set_params": (C.c_int, [_P, _P, _P, C.c_uint32]),
    "luzrt_denoise": (C.c_int, [_P, _P, C.POINTER(DnFrame), _P, _P]),
    "luzrt_denoiser_features": (C.c_int, [_P, C.POINTER(DnFrame), _P, C.c_uint32, _P]),
    "luzrt_denoiser_accumulate": (C.c_int, [_P, _P, C.POINTER(DnFrame), _P, _P, _P, C.c_uint32, C.c_float]),
    "luzrt_denoiser_gradient": (C.c_int, [_P, _P, _P, _P]),
    "luzrt_denoiser_adam": (C.c_int, [_P, _P, C.c_float, C.c_float, C.c_float, C.c_float]),
}
for _name, (_res, _args) in _sigs.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args
EXPORTED = tuple(_sigs)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


class Frame:
    """Device planes of one frame for the denoiser: G-buffer position/normal,
    noisy counts, and (for training) targets and the list of foreground pixels."""

    def __init__(self, ctx: api.Context, width: int, height: int, position, normal, cam_pos,
                 num_lights: int, shadow, shadow_samples: int, ao, ao_samples: int):
        self.ctx = ctx
        self.width, self.height, self.num_lights = width, height, num_lights
        self.position, self.normal, self.shadow, self.ao = position, normal, shadow, ao
        self.cam_pos = tuple(float(c) for c in cam_pos)
        self.shadow_samples, self.ao_samples = shadow_samples, ao_samples
        self.target_ao = self.target_vis = None
        self.foreground = None   # host array of non-background pixel indices

    def desc(self) -> DnFrame:
        f = DnFrame()
        f.width, f.height = self.width, self.height
        f.position, f.normal = self.position.ptr, self.normal.ptr
        for k in range(3):
            f.cam_pos[k] = self.cam_pos[k]
        f.num_lights = self.num_lights
        f.shadow_counts = self.shadow.ptr if self.num_lights else None
        f.shadow_samples = self.shadow_samples
        f.ao_counts = self.ao.ptr
        f.ao_samples = self.ao_samples
        return f

    def host(self) -> dict:
        """The frame as the oracle reads it."""
        n = self.width * self.height
        L = self.num_lights
        return {"width": self.width, "height": self.height,
                "position": self.position.download(np.float32, (n, 3)),
                "normal": self.normal.download(np.float32, (n, 3)),
                "cam": np.array(self.cam_pos, np.float32), "num_lights": L,
                "shadow": self.shadow.download(np.uint8, (L, n)) if L else np.zeros((0, n), np.uint8),
                "shadow_n": self.shadow_samples, "ao": self.ao.download(np.uint8, n),
                "ao_n": self.ao_samples}


class Denoiser:
    """luzrt_denoiser: BF16 MFMA MLP (150 -> 64 x4 -> 2) with FP32 master weights."""

    def __init__(self, ctx: api.Context, batch_capacity: int = 0, seed: int = 0):
        self.ctx = ctx
        h = _P()
        ctx._check(lib.luzrt_denoiser_create(ctx.handle, batch_capacity, seed, C.byref(h)))
        self.handle = h.value
        self.capacity = batch_capacity

    def close(self) -> None:
        if self.handle:
            lib.luzrt_denoiser_destroy(self.ctx.handle, self.handle)
            self.handle = None

    @property
    def params(self) -> np.ndarray:
        out = np.empty(N_PARAMS, np.float32)
        self.ctx._check(lib.luzrt_denoiser_get_params(self.ctx.handle, self.handle, _ptr(out), N_PARAMS))
        return out

    @params.setter
    def params(self, value) -> None:
        v = np.ascontiguousarray(value, np.float32)
        assert v.shape == (N_PARAMS,)
        self.ctx._check(lib.luzrt_denoiser_set_params(self.ctx.handle, self.handle, _ptr(v), N_PARAMS))

    def save(self, path) -> None:
        np.save(path, self.params)

    def load(self, path) -> None:
        self.params = np.load(path, allow_pickle=False)

    def denoise(self, frame: Frame, ao_out: api.DeviceBuffer, vis_out: api.DeviceBuffer) -> None:
        d = frame.desc()
        self.ctx._check(lib.luzrt_denoise(self.ctx.handle, self.handle, C.byref(d), ao_out.ptr,
                                          vis_out.ptr if frame.num_lights else None))

    def accumulate(self, frame: Frame, samples: api.DeviceBuffer, count: int, grad_scale: float) -> None:
        d = frame.desc()
        self.ctx._check(lib.luzrt_denoiser_accumulate(
            self.ctx.handle, self.handle, C.byref(d), frame.target_ao.ptr,
            frame.target_vis.ptr if frame.num_lights else None, samples.ptr, count, grad_scale))

    def gradient(self, host: bool = False, loss: bool = False):
        """(gradient or None, sum of squared errors or None); clears the batch.
        Either request synchronises."""
        g = np.empty(N_PARAMS, np.float32) if host else None
        sse = C.c_float()
        self.ctx._check(lib.luzrt_denoiser_gradient(self.ctx.handle, self.handle,
                                                    _ptr(g) if host else None,
                                                    C.byref(sse) if loss else None))
        return g, (float(sse.value) if loss else None)

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8) -> None:
        self.ctx._check(lib.luzrt_denoiser_adam(self.ctx.handle, self.handle, lr, beta1, beta2, eps))


def features(ctx: api.Context, frame: Frame, samples: np.ndarray) -> np.ndarray:
    """Device-computed input features of samples (pixel | light << 24) as FP32."""
    s = np.ascontiguousarray(samples, np.uint32)
    ds = ctx.to_device(s)
    out = ctx.buffer(len(s) * 150 * 2)
    d = frame.desc()
    ctx._check(lib.luzrt_denoiser_features(ctx.handle, C.byref(d), ds.ptr, len(s), out.ptr))
    bits = out.download(np.uint16, (len(s), 150))
    return (bits.astype(np.uint32) << 16).view(np.float32)


# ------------------------------------------------------------------ data from the tracer
def crop_camera(cam: api.Camera, W: int, H: int, x0: int, y0: int, w: int, h: int) -> api.Camera:
    """Camera whose w x h image is the window (x0, y0) of the W x H image of `cam`."""
    c = api.Camera()
    C.memmove(C.byref(c), C.byref(cam), C.sizeof(api.Camera))
    cx = (2.0 * x0 + w) / W - 1.0
    cy = (2.0 * y0 + h) / H - 1.0
    for k in range(3):
        c.forward[k] = cam.forward[k] + cx * cam.right[k] + cy * cam.up[k]
        c.right[k] = cam.right[k] * (w / W)
        c.up[k] = cam.up[k] * (h / H)
    return c


def render_counts(ctx: api.Context, gs: api.GPUScene, gd: api.GBufferDevice, sb: api.SceneBlock,
                  spp: int, frame: int):
    """Shadow (occluded) and AO (unoccluded) counts with `spp` samples per light and AO."""
    blk = api.SceneBlock()
    C.memmove(C.byref(blk), C.byref(sb), C.sizeof(api.SceneBlock))
    for l in range(blk.num_lights):
        blk.lights[l].num_shadow_samples = spp
    blk.ao_num_samples = spp
    blk.frame = frame
    n = gd.width * gd.height
    sh = ctx.buffer(max(blk.num_lights, 1) * n)
    ao = ctx.buffer(n)
    ctx.trace_visibility(gs.tlas, gd.desc(), blk, sh, ao)
    return sh, ao


def project(cam: api.Camera, W: int, H: int, point) -> tuple:
    """Pixel coordinates of a world point through the primary-ray camera."""
    e = np.array(cam.eye, np.float64)
    f = np.array(cam.forward, np.float64)
    r = np.array(cam.right, np.float64)
    u = np.array(cam.up, np.float64)
    d = np.asarray(point, np.float64) - e
    t = d @ f / (f @ f)
    if t <= 0:
        return None
    nx = d @ r / (r @ r) / t
    ny = d @ u / (u @ u) / t
    return ((nx + 1.0) * 0.5 * W, (ny + 1.0) * 0.5 * H)


def crop_origin(h: api.HostScene, cam: api.Camera, W: int, H: int, w: int, hh: int, rng,
                focus: str) -> tuple:
    """Crop origin: uniform, or centred on a random mesh instance in view
    (focus="instance"), where the shadow and AO noise concentrates."""
    if focus == "instance":
        o2w, _, _ = h.nodes()
        pts = [project(cam, W, H, o2w[i].reshape(3, 4)[:, 3]) for i in range(len(o2w))]
        pts = [p for p in pts if p is not None and 0 <= p[0] < W and 0 <= p[1] < H]
        if pts:
            cx, cy = pts[int(rng.integers(0, len(pts)))]
            x0 = int(np.clip(cx - w / 2, 0, W - w))
            y0 = int(np.clip(cy - hh / 2, 0, H - hh))
            return x0, y0
    return int(rng.integers(0, W - w + 1)), int(rng.integers(0, H - hh + 1))


def make_frame(ctx: api.Context, seed: int, crop=(256, 256), rng=None, target_spp: int = 256,
               kind: int = api.SCENE_SPHERES, focus: str = "uniform", ao_radius: float = None) -> Frame:
    """One training/evaluation frame of scene `seed`: a crop of its image (or the
    whole image when crop is None), 1-spp inputs and target_spp targets, AO out
    to `ao_radius` (AO_RADIUS by default)."""
    h = api.HostScene(kind, seed)
    gs = api.GPUScene(ctx, h)
    W, H, cam = h.camera()
    if crop is not None:
        w, hh = crop
        rng = rng or np.random.default_rng(seed)
        x0, y0 = crop_origin(h, cam, W, H, w, hh, rng, focus)
        cam = crop_camera(cam, W, H, x0, y0, w, hh)
        W, H = w, hh
    _, _, mats = h.nodes()
    gd = api.GBufferDevice(ctx, W, H)
    ctx.primary_gbuffer(gs.tlas, cam, mats, gd.desc())
    sb = h.pack()
    sb.ao_max = AO_RADIUS if ao_radius is None else ao_radius
    L = sb.num_lights
    sh1, ao1 = render_counts(ctx, gs, gd, sb, 1, 0)
    n = W * H
    half = target_spp // 2
    sha, aoa = render_counts(ctx, gs, gd, sb, half, 1)
    shb, aob = render_counts(ctx, gs, gd, sb, target_spp - half, 2)
    ao_t = (aoa.download(np.uint8, n).astype(np.float32) + aob.download(np.uint8, n)) / target_spp
    sh_t = (sha.download(np.uint8, (max(L, 1), n)).astype(np.float32) +
            shb.download(np.uint8, (max(L, 1), n))) / target_spp
    fr = Frame(ctx, W, H, gd.position, gd.normal, tuple(sb.cam_pos), L, sh1, 1, ao1, 1)
    fr.target_ao = ctx.to_device(ao_t.astype(np.float32))
    fr.target_vis = ctx.to_device((1.0 - sh_t[:L]).astype(np.float32) if L else np.zeros(1, np.float32))
    nrm = gd.normal.download(np.float32, (n, 3))
    fr.foreground = np.flatnonzero(np.abs(nrm).sum(1) != 0).astype(np.uint32)
    fr.gbuffer = gd
    fr.scene_block = sb
    fr.seed = seed
    fr.target_ao_host = ao_t
    fr.target_vis_host = (1.0 - sh_t[:L]) if L else np.zeros((0, n), np.float32)
    # (pixel, light) samples whose targets are fractional: where the noise is
    fg = np.zeros(n, bool)
    fg[fr.foreground] = True
    frac_ao = (ao_t > 0) & (ao_t < 1) & fg
    hard = []
    for l in range(max(L, 1)):
        m = frac_ao.copy()
        if L:
            v = fr.target_vis_host[l]
            m |= (v > 0) & (v < 1) & fg
        idx = np.flatnonzero(m).astype(np.uint32)
        hard.append(idx | np.uint32(l << 24))
    fr.hard = np.concatenate(hard) if hard else np.zeros(0, np.uint32)
    for p in ("albedo", "material", "emission"):
        getattr(gd, p).free()
    gs.close()
    h.close()
    return fr


# ------------------------------------------------------------------ training
def train(ctx: api.Context, dn: Denoiser, frames, steps: int, frames_per_step: int = 8,
          samples_per_frame: int = 8192, lr: float = 2e-3, lr_min: float = 2e-5, seed: int = 0,
          log_every: int = 250, log=print, hard_fraction: float = 0.5) -> list:
    """Adam on the mean squared error of random (foreground pixel, light)
    samples, a `hard_fraction` of them drawn from the samples with fractional
    targets (penumbrae and partial occlusion, where the 1-spp noise is)."""
    rng = np.random.default_rng(seed)
    n_batch = frames_per_step * samples_per_frame
    assert n_batch <= dn.capacity
    scale = 2.0 / n_batch
    sample_bufs = [ctx.buffer(samples_per_frame * 4) for _ in range(frames_per_step)]
    history = []
    t0 = time.perf_counter()
    for step in range(steps):
        picks = rng.choice(len(frames), frames_per_step, replace=len(frames) < frames_per_step)
        for buf, fi in zip(sample_bufs, picks):
            fr = frames[fi]
            # half uniform over (foreground pixel, light), half over the noisy samples
            nh = int(samples_per_frame * hard_fraction) if len(fr.hard) else 0
            nu = samples_per_frame - nh
            px = fr.foreground[rng.integers(0, len(fr.foreground), nu)]
            li = rng.integers(0, max(fr.num_lights, 1), nu).astype(np.uint32)
            smp = np.concatenate([(px | (li << 24)).astype(np.uint32),
                                  fr.hard[rng.integers(0, len(fr.hard), nh)] if nh else np.zeros(0, np.uint32)])
            buf.upload(smp)
            dn.accumulate(fr, buf, samples_per_frame, scale)
        want_loss = (step % log_every == 0) or step == steps - 1
        _, sse = dn.gradient(loss=want_loss)
        rate = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * step / max(steps - 1, 1)))
        dn.adam(rate)
        if want_loss:
            mse = sse / (2 * n_batch)
            history.append((step, mse))
            log(f"step {step}: mse {mse:.6f} lr {rate:.2e} ({time.perf_counter() - t0:.1f} s)")
    return history


def evaluate(ctx: api.Context, dn: Denoiser, fr: Frame) -> dict:
    """RMSE over foreground pixels against the frame's targets: raw 1-spp and denoised."""
    n = fr.width * fr.height
    L = fr.num_lights
    ao_o, vis_o = ctx.buffer(n * 4), ctx.buffer(max(L, 1) * n * 4)
    dn.denoise(fr, ao_o, vis_o)
    fg = fr.foreground
    ao_d = ao_o.download(np.float32, n)[fg].astype(np.float64)
    vis_d = vis_o.download(np.float32, (max(L, 1), n))[:L, fg].astype(np.float64)
    ao_r = fr.ao.download(np.uint8, n)[fg] / float(fr.ao_samples)
    vis_r = 1.0 - fr.shadow.download(np.uint8, (max(L, 1), n))[:L, fg] / float(max(fr.shadow_samples, 1))
    ta = fr.target_ao_host[fg].astype(np.float64)
    tv = fr.target_vis_host[:, fg].astype(np.float64)
    r = lambda a, b: float(np.sqrt(np.mean((a - b) ** 2))) if a.size else 0.0
    return {"ao_raw": r(ao_r, ta), "ao_mlp": r(ao_d, ta), "vis_raw": r(vis_r, tv), "vis_mlp": r(vis_d, tv)}


def main() -> int:
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--scenes", type=int, default=2000)
    t.add_argument("--crop", type=int, default=256)
    t.add_argument("--steps", type=int, default=4000)
    t.add_argument("--frames-per-step", type=int, default=8)
    t.add_argument("--samples-per-frame", type=int, default=8192)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--first-scene", type=int, default=100000, help="training scene seeds start here")
    t.add_argument("--out", default=str(WEIGHTS))
    t.add_argument("--ao-radius", type=float, default=AO_RADIUS)
    t.add_argument("--eval-scenes", type=int, default=4, help="held-out scenes evaluated after training")
    args = ap.parse_args()
    ctx = api.Context(0)
    t0 = time.perf_counter()
    rng = np.random.default_rng(args.seed)
    frames = []
    for i in range(args.scenes):
        focus = "instance" if i % 4 else "uniform"
        frames.append(make_frame(ctx, args.first_scene + i, (args.crop, args.crop), rng, focus=focus,
                                 ao_radius=args.ao_radius))
        if i % 100 == 0:
            print(f"data: {i + 1}/{args.scenes} frames ({time.perf_counter() - t0:.1f} s)", flush=True)
    print(f"data: {len(frames)} frames of {args.crop}x{args.crop} in {time.perf_counter() - t0:.1f} s", flush=True)
    dn = Denoiser(ctx, args.frames_per_step * args.samples_per_frame, args.seed)
    train(ctx, dn, frames, args.steps, args.frames_per_step, args.samples_per_frame, args.lr,
          log=lambda m: print(m, flush=True))
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    dn.save(args.out)
    print(f"saved {args.out} ({time.perf_counter() - t0:.1f} s total)", flush=True)
    del frames
    for i in range(args.eval_scenes):
        fr = make_frame(ctx, 900001 + i, (args.crop, args.crop), rng, focus="instance",
                        ao_radius=args.ao_radius)
        print(f"held-out scene {900001 + i}: " +
              " ".join(f"{k} {v:.5f}" for k, v in evaluate(ctx, dn, fr).items()), flush=True)
    dn.close()
    ctx.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
