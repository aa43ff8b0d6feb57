"""The TAA / compose oracle pinned on the CPU: a static scene is a fixed
point, history off screen passes the source through, out-of-box history is
clipped, a camera motion reprojects history onto the source (and the wrong
motion does not), the Mitchell reconstruction is normalised, the
anti-flicker blend of one pixel equals taa.comp's formula in FP64, and
ACES + gamma known answers."""
import numpy as np

from oracle import oracle as orc

W, H = 32, 16


def perspective(fovy=1.0, aspect=W / H, near=0.1, far=100.0):
    """glm::perspective, [0, 1] depth, with the reference's y flip (column-major)."""
    t = np.tan(fovy / 2)
    P = np.zeros((4, 4))
    P[0, 0] = 1 / (aspect * t)
    P[1, 1] = -1 / t
    P[2, 2] = far / (near - far)
    P[3, 2] = -1.0
    P[2, 3] = -(far * near) / (far - near)
    return P   # row-major maths matrix


def view_at(ex):
    V = np.eye(4)
    V[0, 3] = -ex   # camera at (ex, 0, 0) looking down -z
    return V


def flat(M):
    return np.asarray(M, np.float64).T.reshape(16).astype(np.float32)   # column-major


def plane(depth=10.0, ex=0.0):
    """World positions of the pixel centres on the plane z = -depth, camera at x = ex."""
    P = perspective()
    xs = (np.arange(W) + 0.5) / W * 2 - 1
    ys = (np.arange(H) + 0.5) / H * 2 - 1
    X, Y = np.meshgrid(xs, ys)
    vx = X * depth / P[0, 0]
    vy = Y * depth / P[1, 1]
    pos = np.stack([vx + ex, vy, np.full_like(vx, -depth)], -1).reshape(-1, 3).astype(np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (W * H, 1))
    return pos, nrm


def run(light, history, pos, nrm, vp, pvp, reconstruct=0):
    ivp = np.linalg.inv(vp.astype(np.float64).reshape(4, 4).T).T.reshape(16).astype(np.float32)
    return orc.taa(light, history, pos, nrm, W, H, vp, pvp, ivp, reconstruct=reconstruct)


def test_static_uniform_scene_is_a_fixed_point():
    pos, nrm = plane()
    vp = flat(perspective() @ view_at(0))
    c = np.tile(np.array([0.3, 0.6, 0.9, 1.0], np.float32), (W * H, 1))
    for rec in (0, 1):
        out = run(c, c, pos, nrm, vp, vp, rec)
        np.testing.assert_allclose(out, c, rtol=4e-7, atol=0)


def test_history_off_screen_passes_the_source_through():
    pos, nrm = plane()
    vp = flat(perspective() @ view_at(0))
    pvp = flat(perspective() @ view_at(1000.0))     # every point far off the previous screen
    rng = np.random.default_rng(0)
    src = rng.uniform(0, 2, (W * H, 4)).astype(np.float32)
    out = run(src, np.zeros_like(src), pos, nrm, vp, pvp)
    np.testing.assert_array_equal(out, src)


def test_out_of_box_history_is_clipped():
    pos, nrm = plane()
    vp = flat(perspective() @ view_at(0))
    c = np.tile(np.array([0.5, 0.25, 0.125, 1.0], np.float32), (W * H, 1))
    out = run(c, 100 * c, pos, nrm, vp, vp)
    np.testing.assert_allclose(out, c, rtol=1e-6)


def _smooth(rng):
    yy, xx = np.mgrid[0:H, 0:W]
    img = np.stack([np.sin(xx * 0.3) + 1.5, np.cos(yy * 0.4) + 1.5, 0.1 * xx + 0.2, np.ones_like(xx)], -1)
    return img.reshape(-1, 4).astype(np.float32)


def test_motion_reprojects_history_onto_the_source():
    depth = 10.0
    P = perspective()
    px = 2 * depth / P[0, 0] / W * 2            # world width of two pixels on the plane
    pos, nrm = plane(depth, ex=0.0)
    vp = flat(P @ view_at(0.0))
    pvp = flat(P @ view_at(-px))                 # previous camera two pixels to the left
    src = _smooth(np.random.default_rng(1))
    img = src.reshape(H, W, 4)
    # the previous frame saw world point x at a pixel two columns further right
    hist = np.roll(img, 2, axis=1).reshape(-1, 4)
    out = run(src, hist, pos, nrm, vp, pvp)
    inner = np.zeros((H, W), bool)
    inner[1:-1, 3:-3] = True
    inner = inner.reshape(-1)
    err = np.abs(out - src)[inner].max()
    wrong = run(src, hist, pos, nrm, pvp, vp)   # motion reversed
    err_wrong = np.abs(wrong - src)[inner].max()
    assert err < 1e-4 and err_wrong > 1e-2, (err, err_wrong)


def test_anti_flicker_blend_matches_the_formula():
    pos, nrm = plane()
    vp = flat(perspective() @ view_at(0))
    a = np.array([0.2, 0.4, 0.1, 1.0], np.float32)
    b = np.array([1.2, 0.9, 0.6, 1.0], np.float32)
    yy, xx = np.mgrid[0:H, 0:W]
    checker = ((xx + yy) % 2 == 0).reshape(-1)
    src = np.where(checker[:, None], a, b).astype(np.float32)
    hv = np.array([0.7, 0.6, 0.3, 1.0], np.float32)        # inside the neighbourhood box
    hist = np.tile(hv, (W * H, 1))
    out = run(src, hist, pos, nrm, vp, vp)
    p = 5 * W + 7                                            # an interior pixel
    s = src[p].astype(np.float64)
    h = hv.astype(np.float64)
    lum = lambda c: c[0] * 0.2127 + c[1] * 0.7152 + c[2] * 0.0722
    sw = 0.05 / (1 + lum(s[:3] / (s[:3].max() + 1)))
    hw = 0.95 / (1 + lum(h[:3] / (h[:3].max() + 1)))
    want = (s * sw + h * hw) / (sw + hw)
    np.testing.assert_allclose(out[p], want, rtol=2e-6)


def test_compose_known_answers():
    x = np.array([[0, 1, 1e4, 0.5]], np.float32)
    rgba8, rgb = orc.compose(x)
    f = lambda v: ((v * (2.51 * v + 0.03)) / (v * (2.43 * v + 0.59) + 0.14)) ** (1 / 2.2)
    np.testing.assert_allclose(rgb[0, :3], [0.0, f(1.0), f(1e4)], rtol=1e-6, atol=1e-7)
    assert rgb[0, 3] == 1.0
    assert rgba8[0] & 0xFF == 0 and (rgba8[0] >> 8) & 0xFF == round(f(1.0) * 255) and (rgba8[0] >> 16) & 0xFF == 255
    assert rgba8[0] >> 24 == 255
