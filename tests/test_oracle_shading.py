"""The oracle's GGX lighting (light.frag:17-49,171-235) against closed forms
(Walter et al. 2007 GGX D, Schlick-GGX G with k=(r+1)^2/8 (Karis 2013),
Schlick Fresnel) on single-pixel G-buffers."""
import math

import numpy as np
import pytest

from luz_amd import api
from oracle import oracle as orc


def _pixel(pos, nrm, albedo=(200, 100, 50), rough=128, metal=0, occl=255, emission=(0, 0, 0)):
    a = albedo[0] | albedo[1] << 8 | albedo[2] << 16 | 255 << 24
    m = rough | metal << 8 | occl << 16 | 255 << 24
    e = emission[0] | emission[1] << 8 | emission[2] << 16 | 255 << 24
    return {"position": np.array([pos], np.float32), "normal": np.array([nrm], np.float32),
            "albedo": np.array([a], np.uint32), "material": np.array([m], np.uint32),
            "emission": np.array([e], np.uint32)}


def _scene(cam, lights, ao_n=4, ambient=0.01):
    sb = api.SceneBlock()
    for i, (typ, pos, dirn, col, inten, ns) in enumerate(lights):
        L = sb.lights[i]
        L.type = typ
        for k in range(3):
            L.position[k] = pos[k]
            L.direction[k] = dirn[k]
            L.color[k] = col[k]
        L.intensity = inten
        L.num_shadow_samples = ns
        L.inner_angle, L.outer_angle = math.radians(60), math.radians(50)
    sb.num_lights = len(lights)
    for k in range(3):
        sb.ambient_color[k] = 1.0
        sb.cam_pos[k] = cam[k]
    sb.ambient_intensity = ambient
    sb.ao_num_samples = ao_n
    return sb


def _srgb(c):
    return (c / 255.0) ** 2.2


@pytest.mark.parametrize("rough,metal", [(128, 0), (40, 0), (255, 0), (200, 255), (90, 128)])
def test_normal_incidence_closed_form(rough, metal):
    # camera and point light straight above the surface: N = V = L = H
    g = _pixel((0, 0, 0), (0, 0, 1), rough=rough, metal=metal)
    d = 4.0
    sb = _scene((0, 0, 10), [(api.LIGHT_POINT, (0, 0, d), (0, -1, 0), (1.0, 0.5, 0.25), 20.0, 1)])
    shadow = np.zeros((1, 1), np.uint8)
    ao = np.array([4], np.uint8)
    r, m = rough / 255.0, metal / 255.0
    alb = np.array([_srgb(200), _srgb(100), _srgb(50)])
    a2 = (r * r) ** 2
    D = a2 / (math.pi * a2 * a2)          # NdotH = 1: D = 1/(pi a^2)
    G = 1.0                               # NdotV = NdotL = 1
    F0 = 0.04 * (1 - m) + alb * m
    F = F0                                # HdotV = 1
    spec = D * G * F / (4.0 + 1e-4)
    kD = (1 - F) * (1 - m)
    radiance = np.array([1.0, 0.5, 0.25]) * 20.0 / d**2
    amb = float(np.float32(0.01))  # the SceneBlock stores the ambient intensity in FP32
    want = (kD * alb / math.pi + spec) * radiance + amb * alb
    for f64 in (False, True):
        out = orc.shade(g, 1, 1, sb, shadow, ao, f64=f64)[0]
        # FP32 loses ~1e-4 relative to the cancellation in NdotH^2 (a^2 - 1) + 1 at low roughness
        np.testing.assert_allclose(out[:3], want, rtol=3e-4 if not f64 else 1e-9)
        assert out[3] == 1.0


def test_fresnel_grazing_and_masking():
    # light at 60 degrees from the normal, camera along the normal
    th = math.radians(60)
    Ld = np.array([math.sin(th), 0.0, math.cos(th)])
    g = _pixel((0, 0, 0), (0, 0, 1), rough=100, metal=0)
    sb = _scene((0, 0, 5), [(api.LIGHT_DIRECTIONAL, (0, 0, 0), tuple(-Ld), (1, 1, 1), 2.0, 1)])
    out = orc.shade(g, 1, 1, sb, np.zeros((1, 1), np.uint8), np.array([4], np.uint8), f64=True)[0]
    r = 100 / 255.0
    a2 = (r * r) ** 2
    Ld = -np.asarray(tuple(-Ld), np.float32).astype(np.float64)   # as stored in the LightBlock
    Ld /= np.linalg.norm(Ld)
    V = np.array([0, 0, 1.0])
    H = (V + Ld) / np.linalg.norm(V + Ld)
    nh, nv, nl, hv = H[2], 1.0, Ld[2], H @ V
    D = a2 / (math.pi * (nh * nh * (a2 - 1) + 1) ** 2)
    k = (r + 1) ** 2 / 8
    G = (nl / (nl * (1 - k) + k)) * (nv / (nv * (1 - k) + k))
    F = 0.04 + 0.96 * (1 - hv) ** 5
    alb = np.array([_srgb(200), _srgb(100), _srgb(50)])
    spec = D * G * F / (4 * nv * nl + 1e-4)
    want = ((1 - F) * alb / math.pi + spec) * 2.0 * nl + float(np.float32(0.01)) * alb
    np.testing.assert_allclose(out[:3], want, rtol=1e-9)


def test_shadow_ao_emission_background():
    g = _pixel((0, 0, 0), (0, 0, 1), emission=(255, 0, 128), occl=128)
    sb = _scene((0, 0, 10), [(api.LIGHT_POINT, (0, 0, 4), (0, -1, 0), (1, 1, 1), 20.0, 4)], ao_n=8)
    alb = np.array([_srgb(200), _srgb(100), _srgb(50)])
    emis = np.array([1.0, 0.0, 128 / 255.0])
    # fully shadowed light, 2 of 8 AO samples unoccluded
    out = orc.shade(g, 1, 1, sb, np.array([[4]], np.uint8), np.array([2], np.uint8), f64=True)[0]
    amb = float(np.float32(0.01))
    np.testing.assert_allclose(out[:3], amb * alb * (128 / 255.0) * (2 / 8) + emis, rtol=1e-12)
    # background pixel -> ambient colour * intensity
    bg = _pixel((0, 0, 0), (0, 0, 0))
    out = orc.shade(bg, 1, 1, sb, np.zeros((1, 1), np.uint8), np.zeros(1, np.uint8))[0]
    np.testing.assert_allclose(out, [0.01, 0.01, 0.01, 1.0], rtol=1e-6)
    # a light facing away contributes nothing
    sb2 = _scene((0, 0, 10), [(api.LIGHT_POINT, (1, 0, -4), (0, -1, 0), (1, 1, 1), 20.0, 1)], ao_n=0)
    out = orc.shade(g, 1, 1, sb2, np.zeros((1, 1), np.uint8), np.zeros(1, np.uint8), f64=True)[0]
    np.testing.assert_allclose(out[:3], amb * alb * (128 / 255.0) + emis, rtol=1e-12)


def test_spot_cone_uses_radian_angles():
    # light.frag:202-204 compares cos(theta) with angles stored in radians
    g = _pixel((0, 0, 0), (0, 0, 1))
    inner, outer = math.radians(60), math.radians(50)   # 1.047, 0.873 rad
    sb = _scene((0, 0, 10), [(api.LIGHT_SPOT, (0, 0, 4), (0, 0, -1), (1, 1, 1), 20.0, 1)], ao_n=0)
    out = orc.shade(g, 1, 1, sb, np.zeros((1, 1), np.uint8), np.zeros(1, np.uint8), f64=True)[0]
    theta = 1.0   # dot(L, -dir)
    cone = min(max((theta - outer) / (inner - outer), 0.0), 1.0)
    assert 0.0 < cone < 1.0
    r = 128 / 255.0
    alb = np.array([_srgb(200), _srgb(100), _srgb(50)])
    a2 = (r * r) ** 2
    F = 0.04
    spec = (1 / (math.pi * a2)) * F / (4 + 1e-4)
    inner32, outer32 = float(np.float32(inner)), float(np.float32(outer))
    cone = min(max((theta - outer32) / (inner32 - outer32), 0.0), 1.0)
    want = ((1 - F) * alb / math.pi + spec) * 20.0 / 16.0 * cone + float(np.float32(0.01)) * alb
    np.testing.assert_allclose(out[:3], want, rtol=1e-9)


def test_shadow_type_outcomes():
    """EvaluateShadow by shadow type (light.frag:141-168): ray-traced scenes
    use occluded/samples; ShadowDisabled, and a shadow-map scene whose light
    has no map (no map sample), shade every light as fully shadowed; a map
    decision given as one sample shades as that decision."""
    g = _pixel((0, 0, 0), (0, 0, 1), emission=(255, 0, 128), occl=128)
    lights = [(api.LIGHT_POINT, (0, 0, 4), (0, -1, 0), (1, 1, 1), 20.0, 4),
              (api.LIGHT_SPOT, (0.5, 0, 4), (0, 0, -1), (1, 0.5, 0.2), 30.0, 4)]
    alb = np.array([_srgb(200), _srgb(100), _srgb(50)])
    emis = np.array([1.0, 0.0, 128 / 255.0])
    dark = float(np.float32(0.01)) * alb * (128 / 255.0) * (3 / 8) + emis
    counts = np.array([[1], [2]], np.uint8)
    ao = np.array([3], np.uint8)
    sb = _scene((0, 0, 10), lights, ao_n=8)
    lit = orc.shade(g, 1, 1, sb, counts, ao, f64=True)[0]
    assert np.all(lit[:3] > dark + 1e-3)
    for kind in (api.SHADOW_DISABLED, api.SHADOW_MAP):
        sb.shadow_type = kind
        out = orc.shade(g, 1, 1, sb, counts, ao, f64=True)[0]
        if kind == api.SHADOW_MAP:   # counts are map decisions out of 4 "samples" here
            np.testing.assert_allclose(out, lit, rtol=1e-12)
        for l in range(2):
            sb.lights[l].num_shadow_samples = 0
        out = orc.shade(g, 1, 1, sb, counts, ao, f64=True)[0]
        np.testing.assert_allclose(out[:3], dark, rtol=1e-12)
        for l in range(2):
            sb.lights[l].num_shadow_samples = 4
    # a map decision of "shadowed" for light 0 and "lit" for light 1
    sb.shadow_type = api.SHADOW_MAP
    for l in range(2):
        sb.lights[l].num_shadow_samples = 1
    half = orc.shade(g, 1, 1, sb, np.array([[1], [0]], np.uint8), ao, f64=True)[0]
    sb.shadow_type = api.SHADOW_RAYTRACED
    rt = orc.shade(g, 1, 1, sb, np.array([[1], [0]], np.uint8), ao, f64=True)[0]
    np.testing.assert_allclose(half, rt, rtol=1e-12)
    assert np.all(half[:3] > dark + 1e-3)


def test_visibility_casts_shadow_rays_only_in_ray_traced_scenes():
    from tests.helpers import oracle_scene
    s = api.HostScene(api.SCENE_CUBE, 0)
    W, H, cam = s.camera()
    _, _, mats = s.nodes()
    o = oracle_scene(s)
    g = o.primary_gbuffer(cam, mats, 64, 64)
    sb = s.pack()
    active = int((np.abs(g["normal"]).sum(1) != 0).sum())
    sh_rt, ao_rt, st_rt = o.trace_visibility(g, 64, 64, sb)
    assert sh_rt.sum() > 0 and st_rt.rays == active * (1 + sb.ao_num_samples)
    for kind in (api.SHADOW_DISABLED, api.SHADOW_MAP):
        sb.shadow_type = kind
        sh, ao, st = o.trace_visibility(g, 64, 64, sb)
        assert sh.sum() == 0 and st.rays == active * sb.ao_num_samples
        np.testing.assert_array_equal(ao, ao_rt)
    s.close()
