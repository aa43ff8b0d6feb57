"""The denoiser oracle pinned on the CPU: BF16 rounding known answers,
features of closed-form frames, FP64 gradients against central finite
differences, and the non-learned box-filter baseline."""
import numpy as np
import pytest

from oracle import denoise as odn


def test_bf16_round_known_answers():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -1.0 - 2.0 ** -8 - 2.0 ** -20,
                  0.0, -0.0, 2.0 ** -130], np.float32)
    got = odn.bf16_round(x)
    assert got[0] == 1.0 and got[1] == 1.0                 # tie to even (mantissa 0)
    assert got[2] == np.float32(1.0 + 2.0 ** -6)           # tie to even (mantissa 2)
    assert got[3] == np.float32(-1.0 - 2.0 ** -7)          # above the tie rounds away
    assert got[4] == 0.0 and np.signbit(got[5])
    assert got[6] == np.float32(2.0 ** -130)               # subnormal kept (7 bits suffice)
    assert not (got.view(np.uint32) & 0xFFFF).any()        # every result is a BF16 value


def _plane_frame(W=7, H=5, depth=8.0):
    """Camera at the origin looking down -z at a plane z = -depth; normals +z."""
    n = W * H
    pos = np.zeros((n, 3), np.float32)
    pos[:, 2] = -depth
    nrm = np.zeros((n, 3), np.float32)
    nrm[:, 2] = 1.0
    return {"width": W, "height": H, "position": pos, "normal": nrm, "cam": np.zeros(3, np.float32),
            "num_lights": 2, "shadow": np.zeros((2, n), np.uint8), "shadow_n": 4,
            "ao": np.full(n, 3, np.uint8), "ao_n": 4}


def test_features_of_a_flat_frame():
    f = _plane_frame()
    f["shadow"][1, :] = 1          # light 1: 1 of 4 samples occluded everywhere
    X = odn.features(f, np.array([17, 17, 0]), np.array([0, 1, 1]))
    for i, vis in ((0, 1.0), (1, 0.75), (2, 0.75)):
        x = X[i].reshape(25, 6)
        np.testing.assert_array_equal(x[:, 0], 0.75)       # 3 of 4 AO samples unoccluded
        np.testing.assert_array_equal(x[:, 1], vis)
        np.testing.assert_array_equal(x[:, 2], 0.0)        # same distance everywhere
        np.testing.assert_array_equal(x[:, 3:], [[0.0, 0.0, 1.0]] * 25)


def test_features_depth_clamp_edges_and_background():
    W, H = 7, 5
    f = _plane_frame(W, H)
    p = 2 * W + 3                                         # centre pixel (3, 2)
    f["position"][p + 1, 2] = -8.5                        # (4, 2): 16 * 0.5/8 = 1
    f["position"][p - 1, 2] = -7.75                       # (2, 2): 16 * -0.25/8 = -0.5
    f["position"][p + W, 2] = -20.0                       # (3, 3): clamps to 1
    f["normal"][p - W] = 0.0                              # (3, 1) background
    x = odn.features(f, np.array([p]), np.array([0]))[0].reshape(25, 6)
    assert x[13, 2] == 1.0 and x[11, 2] == -0.5 and x[17, 2] == 1.0
    np.testing.assert_array_equal(x[7], [0, 0, 1, 0, 0, 0])
    # corner pixel: the window clamps to the image, so (-2,-2) is the pixel itself
    xc = odn.features(f, np.array([0]), np.array([0]))[0].reshape(25, 6)
    np.testing.assert_array_equal(xc[0], xc[12])
    np.testing.assert_array_equal(xc[6], xc[12])
    # a background centre has no features
    assert not odn.features(f, np.array([p - W]), np.array([0])).any()


def _random_net(rng, scale=0.3):
    p = rng.normal(0.0, scale, odn.N_PARAMS).astype(np.float32)
    return odn.device_weights(p)


def test_backward_matches_central_differences():
    rng = np.random.default_rng(5)
    w = _random_net(rng)
    X = rng.uniform(-1, 1, (7, odn.N_IN))
    T = rng.uniform(0, 1, (7, 2))
    g = odn.backward(w, X, T, scale=0.7)
    flat = odn.pack(w)
    sizes = [m.size for m in w]
    starts = np.cumsum([0] + sizes[:-1])
    idx = np.concatenate([rng.choice(n, min(n, 40), replace=False) + s for s, n in zip(starts, sizes)])
    h = 1e-6
    for i in idx:
        wp, wm = flat.copy(), flat.copy()
        wp[i] += h
        wm[i] -= h
        up = [a.reshape(m.shape) for a, m in zip(np.split(wp, starts[1:]), w)]
        um = [a.reshape(m.shape) for a, m in zip(np.split(wm, starts[1:]), w)]
        fd = (odn.loss(up, X, T, 0.7) - odn.loss(um, X, T, 0.7)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])


def test_forward_closed_form():
    # zero weights: outputs are the output biases; one path through unit weights
    w = [np.zeros_like(m) for m in _random_net(np.random.default_rng(0))]
    w[9] = np.array([0.25, -0.5])
    X = np.random.default_rng(1).uniform(-1, 1, (3, odn.N_IN))
    y, _, _ = odn.forward(w, X)
    np.testing.assert_array_equal(y, [[0.25, -0.5]] * 3)
    w[0][0, 7] = 1.0                       # h1[0] = relu(x7)
    for k in (2, 4, 6):
        w[k][0, 0] = 1.0                   # carried through layers 2-4
    w[8][1, 0] = 2.0                       # y1 = -0.5 + 2 relu(x7)
    y, _, _ = odn.forward(w, X)
    np.testing.assert_allclose(y[:, 1], -0.5 + 2 * np.maximum(X[:, 7], 0), rtol=0, atol=1e-15)


def test_box_filter_averages_foreground_only():
    f = _plane_frame()
    f["ao"] = np.arange(35, dtype=np.uint8) % 5
    f["normal"][::2] = 0.0                               # every other pixel background
    a, v = odn.box_filter(f, 0, radius=0)
    fg = np.abs(f["normal"]).sum(1) != 0
    np.testing.assert_allclose(a[fg], f["ao"][fg] / 4.0)
    a2, _ = odn.box_filter(_plane_frame(), 1)
    np.testing.assert_allclose(a2, 0.75)
