"""CPU oracle of the learned denoiser (test infrastructure only).

Restates, in numpy, what luz_amd/csrc/denoise.hip computes, so that GPU
results can be checked:
  features()   the 150 BF16 input features of (pixel, light) samples, with the
               same FP32 operation order as the device (bit-exact);
  forward()    the MLP in FP64 on BF16-rounded weights and FP32 biases;
  box_base()   the base the network's outputs are added to (the 5x5 mean of
               the noisy AO and visibility over foreground pixels, box_filter());
  backward()   the gradient of  s/2 * sum (y - target)^2  in FP64, in the
               parameter order of the C ABI (W1 b1 W2 b2 W3 b3 W4 b4 W5 b5).
The denoiser has no counterpart in the reference (light.frag:111-146 uses
the noisy ray-query estimates directly); the channels it denoises are the
shadow fraction of light.frag:104-108 and the AO fraction of :130-134.

Parity of backward() is pinned by central finite differences of the FP64
loss (tests/test_oracle_denoise.py); features() by closed-form frames.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID, N_OUT = 150, 64, 2
SHAPES = [(N_HID, N_IN), (N_HID,), (N_HID, N_HID), (N_HID,), (N_HID, N_HID), (N_HID,),
          (N_HID, N_HID), (N_HID,), (N_OUT, N_HID), (N_OUT,)]
N_PARAMS = sum(int(np.prod(s)) for s in SHAPES)
assert N_PARAMS == 22274


def bf16_round(x) -> np.ndarray:
    """Round FP32 values to the nearest BF16 (ties to even), returned as FP32."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(params: np.ndarray):
    """[W1, b1, ..., W5, b5] views of a flat parameter vector."""
    out, o = [], 0
    for s in SHAPES:
        n = int(np.prod(s))
        out.append(params[o:o + n].reshape(s))
        o += n
    return out


def pack(mats) -> np.ndarray:
    return np.concatenate([np.asarray(m, np.float64).reshape(-1) for m in mats])


def _dist(pos: np.ndarray, cam) -> np.ndarray:
    c = np.asarray(cam, np.float32)
    d = (pos - c).astype(np.float32)
    return np.sqrt((d[..., 0] * d[..., 0] + d[..., 1] * d[..., 1]) + d[..., 2] * d[..., 2]).astype(np.float32)


def features(frame: dict, pixels: np.ndarray, lights: np.ndarray) -> np.ndarray:
    """(N, 150) FP32 features (BF16 values) of samples (pixels[i], lights[i]).

    frame: width, height, position (npix,3) f32, normal (npix,3) f32, cam (3,),
    num_lights, shadow (L,npix) u8, shadow_n, ao (npix,) u8, ao_n."""
    W, H = frame["width"], frame["height"]
    pos = np.asarray(frame["position"], np.float32).reshape(-1, 3)
    nrm = np.asarray(frame["normal"], np.float32).reshape(-1, 3)
    L = frame["num_lights"]
    ao_n = np.float32(frame["ao_n"])
    sh_n = np.float32(frame.get("shadow_n", 1))
    nn = nrm * nrm
    bg = np.sqrt((nn[:, 0] + nn[:, 1]) + nn[:, 2]) == 0
    dist = _dist(pos, frame["cam"])
    ao = np.asarray(frame["ao"], np.uint8).astype(np.float32) / ao_n
    pixels = np.asarray(pixels, np.int64)
    lights = np.asarray(lights, np.int64)
    x, y = pixels % W, pixels // W
    dc = dist[pixels]
    X = np.zeros((len(pixels), N_IN), np.float32)
    centre_fg = ~bg[pixels]     # background samples have all-zero features
    for n in range(25):
        dx, dy = n % 5 - 2, n // 5 - 2
        q = np.clip(y + dy, 0, H - 1) * W + np.clip(x + dx, 0, W - 1)
        c0 = ao[q]
        if L:
            sh = np.asarray(frame["shadow"], np.uint8).reshape(L, -1)
            c1 = np.float32(1.0) - sh[lights, q].astype(np.float32) / sh_n
        else:
            c1 = np.zeros(len(q), np.float32)
        with np.errstate(divide="ignore", invalid="ignore"):
            rel = np.where(dc > 0, ((dist[q] - dc) / dc) * np.float32(16.0), np.float32(0.0)).astype(np.float32)
        c2 = np.clip(rel, np.float32(-1.0), np.float32(1.0))
        b = bg[q]
        X[:, n * 6 + 0] = np.where(b, 0.0, c0)
        X[:, n * 6 + 1] = np.where(b, 0.0, c1)
        X[:, n * 6 + 2] = np.where(b, 1.0, c2)
        X[:, n * 6 + 3:n * 6 + 6] = np.where(b[:, None], 0.0, nrm[q])
    X[~centre_fg] = 0.0
    return bf16_round(X)


def device_weights(params: np.ndarray):
    """The weights as the device uses them: BF16 matrices, FP32 biases (FP64 arrays)."""
    m = unpack(np.asarray(params, np.float32))
    return [bf16_round(v).astype(np.float64) if i % 2 == 0 else v.astype(np.float64)
            for i, v in enumerate(m)]


def forward(w, X: np.ndarray):
    """FP64 forward; returns (y (N,2), pre-activations [z1..z4], inputs [x, h1..h4])."""
    a = np.asarray(X, np.float64)
    zs, hs = [], [a]
    for k in range(4):
        z = a @ w[2 * k].T + w[2 * k + 1]
        zs.append(z)
        a = np.maximum(z, 0.0)
        hs.append(a)
    y = a @ w[8].T + w[9]
    return y, zs, hs


def loss(w, X, T, scale: float = 1.0) -> float:
    y, _, _ = forward(w, X)
    return 0.5 * scale * float(np.sum((y - T) ** 2))


def backward(w, X, T, scale: float = 1.0) -> np.ndarray:
    """Gradient of scale/2 * sum (y - T)^2 over the samples, flat parameter order."""
    y, zs, hs = forward(w, X)
    dz = scale * (y - T)
    grads = [None] * 10
    grads[8] = dz.T @ hs[4]
    grads[9] = dz.sum(0)
    dh = dz @ w[8]
    for k in range(3, -1, -1):
        dz = dh * (zs[k] > 0)
        grads[2 * k] = dz.T @ hs[k]
        grads[2 * k + 1] = dz.sum(0)
        dh = dz @ w[2 * k]
    return pack(grads)


def box_filter(frame: dict, light: int, radius: int = 2):
    """Non-learned baseline: mean of the noisy AO and visibility over the
    (2r+1)^2 window of non-background neighbours (clamped to the image)."""
    W, H = frame["width"], frame["height"]
    nrm = np.asarray(frame["normal"], np.float32).reshape(H, W, 3)
    fg = np.abs(nrm).sum(-1) != 0
    ao = np.asarray(frame["ao"], np.float64).reshape(H, W) / frame["ao_n"]
    if frame["num_lights"]:
        vis = 1.0 - np.asarray(frame["shadow"], np.float64).reshape(-1, H, W)[light] / frame["shadow_n"]
    else:
        vis = np.zeros((H, W))
    acc_a = np.zeros((H, W)); acc_v = np.zeros((H, W)); cnt = np.zeros((H, W))
    for dy in range(-radius, radius + 1):
        for dx in range(-radius, radius + 1):
            ys = np.clip(np.arange(H) + dy, 0, H - 1)
            xs = np.clip(np.arange(W) + dx, 0, W - 1)
            m = fg[np.ix_(ys, xs)]
            acc_a += np.where(m, ao[np.ix_(ys, xs)], 0.0)
            acc_v += np.where(m, vis[np.ix_(ys, xs)], 0.0)
            cnt += m
    cnt = np.maximum(cnt, 1)
    return (acc_a / cnt).reshape(-1), (acc_v / cnt).reshape(-1)


def box_base(frame: dict, pixels: np.ndarray, lights: np.ndarray) -> np.ndarray:
    """(N, 2) FP64 base of samples (pixels[i], lights[i]): the device adds the
    network's (AO, visibility) outputs to box_filter()'s (AO, visibility)."""
    pixels = np.asarray(pixels, np.int64)
    lights = np.asarray(lights, np.int64)
    out = np.zeros((len(pixels), 2))
    for l in np.unique(lights):
        a, v = box_filter(frame, int(l))
        m = lights == l
        out[m, 0] = a[pixels[m]]
        out[m, 1] = v[pixels[m]]
    return out
