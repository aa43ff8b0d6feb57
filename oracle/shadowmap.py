"""CPU oracle of the shadow-map path (test infrastructure only).

Restates luz_amd/csrc/shadowmap.hip with numpy FP32 arithmetic in the same
operation order, and the oracle's closest-hit traversal for the map
texels:
  map_rays()        one ray per texel (orthographic near-to-far ray through
                    the texel centre, or the light-to-texel ray of a cube face);
  build()           depth per texel: the hit parameter t (NDC z of an
                    orthographic map) or min(t |d| / z_far, 1), 1 without a hit;
  visibility()      light.frag:147-165 EvaluateShadow with shadowType == Map:
                    bilinear depth lookup (REPEAT for 2D maps, clamped within
                    a cube face) and the depth comparisons, 1 = shadowed.
Pinned by tests/test_oracle_shadowmap.py against ray-traced hard shadows of
the same scene (maps and rays must agree away from shadow edges).
"""
from __future__ import annotations

import numpy as np

F = np.float32


def _mat_vec(m, x, y, z, w):
    """Column-major 4x4 (float32[16]) times (x, y, z, w), summed left to right."""
    m = np.asarray(m, F).reshape(16)
    out = []
    for r in range(4):
        acc = m[r] * x
        acc = acc + m[4 + r] * y
        acc = acc + m[8 + r] * z
        acc = acc + m[12 + r] * w
        out.append(acc.astype(F))
    return out


def _vlen(x, y, z):
    return np.sqrt((x * x + y * y) + z * z).astype(F)


def _cube_dir(f, sc, tc):
    one = np.ones_like(sc)
    return {0: (one, -tc, -sc), 1: (-one, -tc, sc), 2: (sc, one, tc), 3: (sc, -one, -tc),
            4: (sc, -tc, one), 5: (-sc, -tc, -one)}[f]


def map_rays(m: dict) -> np.ndarray:
    """(faces * res * res, 8) float32 rays in texel order (face, row j, column i)."""
    res = m["resolution"]
    idx = np.arange(res, dtype=F)
    c = ((idx + F(0.5)) / F(res) * F(2.0) - F(1.0)).astype(F)
    X, Y = np.meshgrid(c, c)          # X[j, i] = c[i], Y[j, i] = c[j]
    X, Y = X.reshape(-1), Y.reshape(-1)
    rays = []
    if m["light_type"] == 0:
        lp = np.asarray(m["light_pos"], F)
        for f in range(6):
            dx, dy, dz = _cube_dir(f, X, Y)
            r = np.zeros((len(X), 8), F)
            r[:, 0:3] = lp
            r[:, 3] = 0.0
            r[:, 4], r[:, 5], r[:, 6] = dx, dy, dz
            r[:, 7] = m["z_far"]
            rays.append(r)
    else:
        z0, one = np.zeros_like(X), np.ones_like(X)
        a = _mat_vec(m["inv_view_proj"], X, Y, z0, one)
        b = _mat_vec(m["inv_view_proj"], X, Y, one, one)
        r = np.zeros((len(X), 8), F)
        r[:, 0], r[:, 1], r[:, 2] = a[0], a[1], a[2]
        r[:, 3] = 0.0
        r[:, 4], r[:, 5], r[:, 6] = (b[0] - a[0]), (b[1] - a[1]), (b[2] - a[2])
        r[:, 7] = 1.0
        rays.append(r)
    return np.concatenate(rays)


def build(scene, m: dict) -> np.ndarray:
    """Depth per texel (float32) from the oracle scene's closest hits."""
    rays = map_rays(m)
    t, _, _ = scene.closest_hit(rays)
    hit = np.isfinite(t)
    depth = np.ones(len(rays), F)
    if m["light_type"] == 0:
        dist = (t * _vlen(rays[:, 4], rays[:, 5], rays[:, 6])).astype(F)
        v = (dist / F(m["z_far"])).astype(F)
        depth[hit] = np.minimum(v[hit], F(1.0))
    else:
        depth[hit] = t[hit]
    return depth


def _sample(img, res, u, v, repeat):
    tx = (u * F(res) - F(0.5)).astype(F)
    ty = (v * F(res) - F(0.5)).astype(F)
    fx, fy = np.floor(tx), np.floor(ty)
    ax, ay = (tx - fx).astype(F), (ty - fy).astype(F)
    x0, y0 = fx.astype(np.int64), fy.astype(np.int64)
    if repeat:
        xa, xb, ya, yb = x0 % res, (x0 + 1) % res, y0 % res, (y0 + 1) % res
    else:
        xa, xb = np.clip(x0, 0, res - 1), np.clip(x0 + 1, 0, res - 1)
        ya, yb = np.clip(y0, 0, res - 1), np.clip(y0 + 1, 0, res - 1)
    d00, d10 = img[ya * res + xa], img[ya * res + xb]
    d01, d11 = img[yb * res + xa], img[yb * res + xb]
    top = (d00 * (F(1.0) - ax) + d10 * ax).astype(F)
    bot = (d01 * (F(1.0) - ax) + d11 * ax).astype(F)
    return (top * (F(1.0) - ay) + bot * ay).astype(F)


def visibility(g: dict, width: int, height: int, cam_pos, maps) -> np.ndarray:
    """(num maps, npix) u8: 1 where the light is shadowed; maps[l] = (desc, depth or None)."""
    P = np.asarray(g["position"], F).reshape(-1, 3)
    N = np.asarray(g["normal"], F).reshape(-1, 3)
    fg = _vlen(N[:, 0], N[:, 1], N[:, 2]) != 0
    c = np.asarray(cam_pos, F)
    lc = _vlen(P[:, 0] - c[0], P[:, 1] - c[1], P[:, 2] - c[2])
    bias = np.maximum((lc * F(0.01)).astype(F), F(0.05)).astype(F)
    so = (P + N * bias[:, None]).astype(F)
    out = np.zeros((len(maps), len(P)), np.uint8)
    for l, (m, depth) in enumerate(maps):
        if depth is None:
            out[l] = fg
            continue
        res = m["resolution"]
        if m["light_type"] == 0:
            lp = np.asarray(m["light_pos"], F)
            lx, ly, lz = P[:, 0] - lp[0], P[:, 1] - lp[1], P[:, 2] - lp[2]
            ax, ay, az = np.abs(lx), np.abs(ly), np.abs(lz)
            isx = (ax >= ay) & (ax >= az)
            isy = ~isx & (ay >= az)
            face = np.where(isx, np.where(lx >= 0, 0, 1), np.where(isy, np.where(ly >= 0, 2, 3), np.where(lz >= 0, 4, 5)))
            ma = np.where(isx, ax, np.where(isy, ay, az))
            sc = np.where(isx, np.where(lx >= 0, -lz, lz), np.where(isy, lx, np.where(lz >= 0, lx, -lx)))
            tc = np.where(isx, -ly, np.where(isy, np.where(ly >= 0, lz, -lz), -ly))
            with np.errstate(divide="ignore", invalid="ignore"):
                u = ((sc / ma + F(1.0)) * F(0.5)).astype(F)
                v = ((tc / ma + F(1.0)) * F(0.5)).astype(F)
            d = np.zeros(len(P), F)
            for f in range(6):
                sel = fg & (face == f)
                if sel.any():
                    d[sel] = _sample(depth[f * res * res:(f + 1) * res * res], res, u[sel], v[sel], False)
            s = ((_vlen(lx, ly, lz) - F(0.05)).astype(F) >= (d * F(m["z_far"])).astype(F))
        else:
            q = _mat_vec(m["view_proj"], so[:, 0], so[:, 1], so[:, 2], np.ones(len(P), F))
            u = (q[0] * F(0.5) + F(0.5)).astype(F)
            v = (q[1] * F(0.5) + F(0.5)).astype(F)
            d = _sample(depth, res, u, v, True)
            s = q[2] >= d
        out[l] = (s & fg).astype(np.uint8)
    return out
