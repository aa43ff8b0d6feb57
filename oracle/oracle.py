"""ctypes wrapper of the CPU oracle (oracle/_build/libluz_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, as the checker.  The product (luz_amd) never
imports it.  Build with `make -C oracle`.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "_build" / "libluz_oracle.so"

if not LIB_PATH.exists():
    raise ImportError(f"oracle library missing at {LIB_PATH}; build it with `make -C oracle`")

lib = C.CDLL(str(LIB_PATH))


class OrcLight(C.Structure):
    _fields_ = [("color", C.c_float * 3), ("intensity", C.c_float),
                ("position", C.c_float * 3), ("inner_angle", C.c_float),
                ("direction", C.c_float * 3), ("outer_angle", C.c_float),
                ("type", C.c_int32), ("num_shadow_samples", C.c_int32),
                ("radius", C.c_float), ("pad", C.c_int32)]


class OrcParams(C.Structure):
    _fields_ = [("lights", OrcLight * 64),
                ("ambient_color", C.c_float * 3), ("ambient_intensity", C.c_float),
                ("cam_pos", C.c_float * 3), ("num_lights", C.c_int32),
                ("ao_min", C.c_float), ("ao_max", C.c_float),
                ("ao_num_samples", C.c_int32), ("frame", C.c_int32),
                ("shadow_type", C.c_int32), ("pad", C.c_int32 * 3)]


class OrcGBuffer(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("position", C.c_void_p), ("normal", C.c_void_p), ("albedo", C.c_void_p),
                ("material", C.c_void_p), ("emission", C.c_void_p)]


class OrcCamera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("tmin", C.c_float), ("tmax", C.c_float)]


class OrcMaterial(C.Structure):
    _fields_ = [("albedo", C.c_uint32), ("material", C.c_uint32), ("emission", C.c_uint32),
                ("pad", C.c_uint32)]


class OrcStats(C.Structure):
    _fields_ = [("rays", C.c_uint64), ("node_visits", C.c_uint64), ("tri_tests", C.c_uint64),
                ("inst_visits", C.c_uint64), ("hits", C.c_uint64), ("wide_node_visits", C.c_uint64)]


class OrcTaaDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("light", C.c_void_p),
                ("history", C.c_void_p), ("position", C.c_void_p), ("normal", C.c_void_p),
                ("view_proj", C.c_float * 16), ("prev_view_proj", C.c_float * 16),
                ("inv_view_proj", C.c_float * 16), ("jitter", C.c_float * 2), ("prev_jitter", C.c_float * 2),
                ("reconstruct", C.c_int32)]


_P = C.c_void_p
_F = C.POINTER(C.c_float)
_sigs = {
    "orc_pcg": (C.c_uint32, [C.c_uint32]),
    "orc_pcg4d": (None, [_P, _P]),
    "orc_sample2": (None, [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "orc_sincos_turns": (None, [C.c_float, _F, _F]),
    "orc_morton3": (C.c_uint32, [C.c_uint32, C.c_uint32, C.c_uint32]),
    "orc_intersect_tri": (C.c_int, [_P, _P, C.c_float, C.c_float, _P, _P, _P, _F, _P]),
    "orc_box_test": (C.c_int, [_P, _P, C.c_float, C.c_float, _P, _P]),
    "orc_invert_3x4": (None, [_P, _P]),
    "orc_scene_create": (_P, []),
    "orc_scene_destroy": (None, [_P]),
    "orc_scene_add_mesh": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32]),
    "orc_scene_refit_mesh": (C.c_int, [_P, C.c_int, _P]),
    "orc_scene_set_instances": (C.c_int, [_P, _P, _P, _P, _P, C.c_uint32]),
    "orc_blas_arrays": (C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(C.c_uint32), C.POINTER(_P),
                                  C.POINTER(_P), C.POINTER(C.c_uint32)]),
    "orc_tlas_arrays": (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_uint32), C.POINTER(_P),
                                  C.POINTER(_P), C.POINTER(C.c_uint32)]),
    "orc_any_hit": (C.c_int, [_P, _P, C.c_uint32, _P, C.c_int, C.POINTER(OrcStats)]),
    "orc_closest_hit": (C.c_int, [_P, _P, C.c_uint32, _P, _P, _P, C.c_int]),
    "orc_primary_gbuffer": (C.c_int, [_P, C.POINTER(OrcCamera), _P, C.c_uint32, C.c_uint32, C.c_uint32,
                                      _P, _P, _P, _P, _P, C.c_int]),
    "orc_trace_visibility": (C.c_int, [_P, C.POINTER(OrcGBuffer), C.POINTER(OrcParams), _P, _P,
                                       C.POINTER(OrcStats), C.c_int, C.c_uint32]),
    "orc_trace_visibility_f64": (C.c_int, [_P, C.POINTER(OrcGBuffer), C.POINTER(OrcParams), _P, _P,
                                           C.c_int, C.c_uint32]),
    "orc_scene_build_wide": (C.c_int, [_P]),
    "orc_blas_wide_arrays": (C.c_int, [_P, C.c_int, C.POINTER(_P), C.POINTER(C.c_uint32), C.POINTER(_P),
                                       C.POINTER(_P), C.POINTER(C.c_uint32)]),
    "orc_any_hit_wide": (C.c_int, [_P, _P, C.c_uint32, _P, C.POINTER(OrcStats)]),
    "orc_any_hit_f64": (C.c_int, [_P, _P, C.c_uint32, _P]),
    "orc_depth_to_world": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, _P, _P, _P]),
    "orc_trace_visibility_wide": (C.c_int, [_P, C.POINTER(OrcGBuffer), C.POINTER(OrcParams), _P, _P,
                                            C.POINTER(OrcStats), C.c_int, C.c_uint32]),
    "orc_visibility_rays": (C.c_int64, [C.POINTER(OrcGBuffer), C.POINTER(OrcParams), _P, C.c_int64]),
    "orc_shade": (C.c_int, [C.POINTER(OrcGBuffer), C.POINTER(OrcParams), _P, _P, _P, C.c_int]),
    "orc_shade_f64": (C.c_int, [C.POINTER(OrcGBuffer), C.POINTER(OrcParams), _P, _P, _P, C.c_int]),
    "orc_taa": (C.c_int, [C.POINTER(OrcTaaDesc), _P]),
    "orc_compose": (C.c_int, [_P, C.c_uint32, _P, _P]),
}
for _n, (_r, _a) in _sigs.items():
    _f = getattr(lib, _n)
    _f.restype = _r
    _f.argtypes = _a


def _p(a: np.ndarray) -> int:
    return a.ctypes.data


def threads_default() -> int:
    return int(os.environ.get("LUZ_ORACLE_THREADS", "0")) or (os.cpu_count() or 1)


def params_from(scene_block) -> OrcParams:
    """Byte copy of a luz_amd.api.SceneBlock (identical layout)."""
    p = OrcParams()
    C.memmove(C.addressof(p), C.addressof(scene_block), C.sizeof(OrcParams))
    return p


def camera_from(cam) -> OrcCamera:
    c = OrcCamera()
    C.memmove(C.addressof(c), C.addressof(cam), C.sizeof(OrcCamera))
    return c


def pcg(v: int) -> int:
    return lib.orc_pcg(v)


def pcg4d(v):
    a = np.asarray(v, np.uint32)
    out = np.zeros(4, np.uint32)
    lib.orc_pcg4d(_p(a), _p(out))
    return out


def sample2(x, y, frame, slot):
    uv = np.zeros(2, np.float32)
    lib.orc_sample2(x, y, frame, slot, _p(uv))
    return uv


def sincos_turns(w: float):
    s, c = C.c_float(), C.c_float()
    lib.orc_sincos_turns(w, C.byref(s), C.byref(c))
    return s.value, c.value


def morton3(x, y, z) -> int:
    return lib.orc_morton3(x, y, z)


def intersect_tri(o, d, tmin, tmax, v0, v1, v2):
    arr = [np.asarray(x, np.float32) for x in (o, d, v0, v1, v2)]
    t = C.c_float()
    bary = np.zeros(3, np.float32)
    hit = lib.orc_intersect_tri(_p(arr[0]), _p(arr[1]), tmin, tmax, _p(arr[2]), _p(arr[3]), _p(arr[4]),
                                C.byref(t), _p(bary))
    return bool(hit), t.value, bary


def box_test(o, d, tmin, tmax, lo, hi) -> bool:
    arr = [np.asarray(x, np.float32) for x in (o, d, lo, hi)]
    return bool(lib.orc_box_test(_p(arr[0]), _p(arr[1]), tmin, tmax, _p(arr[2]), _p(arr[3])))


def depth_to_world(depth: np.ndarray, width: int, height: int, inverse_proj, inverse_view) -> np.ndarray:
    """World positions (float32 [H*W, 3]) of a depth image [H, pitch/4] by DepthToWorld."""
    d = np.ascontiguousarray(depth, np.float32)
    pitch = d.shape[-1] * 4 if d.ndim == 2 else width * 4
    ip = np.ascontiguousarray(inverse_proj, np.float32).reshape(16)
    iv = np.ascontiguousarray(inverse_view, np.float32).reshape(16)
    out = np.zeros((width * height, 3), np.float32)
    if lib.orc_depth_to_world(_p(d), width, height, pitch, _p(ip), _p(iv), _p(out)) != 0:
        raise ValueError("bad depth image")
    return out


def invert_3x4(m):
    a = np.asarray(m, np.float32).reshape(12)
    out = np.zeros(12, np.float32)
    lib.orc_invert_3x4(_p(a), _p(out))
    return out


class OracleScene:
    """Meshes + instances with the oracle's own BLAS/TLAS builds."""

    def __init__(self):
        self.h = lib.orc_scene_create()
        self._keep = []

    def close(self):
        if self.h:
            lib.orc_scene_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_mesh(self, vertices: np.ndarray, indices: np.ndarray, stride_bytes: int = 48) -> int:
        v = np.ascontiguousarray(vertices, np.float32)
        i = np.ascontiguousarray(indices, np.uint32)
        nv = v.nbytes // stride_bytes
        r = lib.orc_scene_add_mesh(self.h, _p(v), stride_bytes, nv, _p(i), i.size)
        if r < 0:
            raise ValueError("orc_scene_add_mesh failed")
        return r

    def refit_mesh(self, mesh: int, vertices: np.ndarray) -> None:
        v = np.ascontiguousarray(vertices, np.float32)
        if lib.orc_scene_refit_mesh(self.h, mesh, _p(v)) != 0:
            raise ValueError("refit failed")

    def set_instances(self, o2w: np.ndarray, mesh_ids, custom=None, mask=None) -> None:
        o = np.ascontiguousarray(o2w, np.float32).reshape(-1, 12)
        m = np.ascontiguousarray(mesh_ids, np.uint32)
        c = np.ascontiguousarray(custom if custom is not None else np.arange(len(m)), np.uint32)
        k = np.ascontiguousarray(mask if mask is not None else np.full(len(m), 0xFF), np.uint32)
        self._keep = [o, m, c, k]
        if lib.orc_scene_set_instances(self.h, _p(o), _p(m), _p(c), _p(k), len(m)) != 0:
            raise ValueError("orc_scene_set_instances failed")

    def blas_arrays(self, mesh: int):
        nodes, nn, tris, prims, nt = C.c_void_p(), C.c_uint32(), C.c_void_p(), C.c_void_p(), C.c_uint32()
        lib.orc_blas_arrays(self.h, mesh, C.byref(nodes), C.byref(nn), C.byref(tris), C.byref(prims), C.byref(nt))
        n = np.ctypeslib.as_array(C.cast(nodes.value, C.POINTER(C.c_uint8)), shape=(nn.value, 64)).copy()
        t = np.zeros((nt.value, 9), np.float32)
        p = np.zeros(nt.value, np.uint32)
        if nt.value:
            t = np.ctypeslib.as_array(C.cast(tris.value, C.POINTER(C.c_float)), shape=(nt.value, 9)).copy()
            p = np.ctypeslib.as_array(C.cast(prims.value, C.POINTER(C.c_uint32)), shape=(nt.value,)).copy()
        return n, t, p

    def tlas_arrays(self):
        nodes, nn, order, w2o, cnt = C.c_void_p(), C.c_uint32(), C.c_void_p(), C.c_void_p(), C.c_uint32()
        lib.orc_tlas_arrays(self.h, C.byref(nodes), C.byref(nn), C.byref(order), C.byref(w2o), C.byref(cnt))
        n = np.ctypeslib.as_array(C.cast(nodes.value, C.POINTER(C.c_uint8)), shape=(nn.value, 64)).copy()
        k = cnt.value
        o = np.zeros(k, np.uint32)
        w = np.zeros((k, 12), np.float32)
        if k:
            o = np.ctypeslib.as_array(C.cast(order.value, C.POINTER(C.c_uint32)), shape=(k,)).copy()
            w = np.ctypeslib.as_array(C.cast(w2o.value, C.POINTER(C.c_float)), shape=(k, 12)).copy()
        return n, w, o

    def build_wide(self) -> None:
        if lib.orc_scene_build_wide(self.h) != 0:
            raise ValueError("orc_scene_build_wide failed")

    def blas_wide_arrays(self, mesh: int):
        """(nodes u8[n,96], tris f32[t,9], prim ids u32[t]) of the 8-wide BLAS."""
        nodes, nn, tris, prims, nt = C.c_void_p(), C.c_uint32(), C.c_void_p(), C.c_void_p(), C.c_uint32()
        lib.orc_blas_wide_arrays(self.h, mesh, C.byref(nodes), C.byref(nn), C.byref(tris), C.byref(prims),
                                 C.byref(nt))
        n = np.ctypeslib.as_array(C.cast(nodes.value, C.POINTER(C.c_uint8)), shape=(nn.value, 96)).copy()
        t = np.zeros((nt.value, 9), np.float32)
        p = np.zeros(nt.value, np.uint32)
        if nt.value:
            t = np.ctypeslib.as_array(C.cast(tris.value, C.POINTER(C.c_float)), shape=(nt.value, 9)).copy()
            p = np.ctypeslib.as_array(C.cast(prims.value, C.POINTER(C.c_uint32)), shape=(nt.value,)).copy()
        return n, t, p

    def any_hit_wide(self, rays: np.ndarray):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros(len(r), np.uint8)
        st = OrcStats()
        lib.orc_any_hit_wide(self.h, _p(r), len(r), _p(hits), C.byref(st))
        return hits, st

    def any_hit_f64(self, rays: np.ndarray) -> np.ndarray:
        """FP64 any-hit of rays given as 8 doubles (origin, tmin, direction, tmax)."""
        r = np.ascontiguousarray(rays, np.float64).reshape(-1, 8)
        hits = np.zeros(len(r), np.uint8)
        lib.orc_any_hit_f64(self.h, _p(r), len(r), _p(hits))
        return hits

    def any_hit(self, rays: np.ndarray, brute: bool = False):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        hits = np.zeros(len(r), np.uint8)
        st = OrcStats()
        lib.orc_any_hit(self.h, _p(r), len(r), _p(hits), int(brute), C.byref(st))
        return hits, st

    def closest_hit(self, rays: np.ndarray, brute: bool = False):
        r = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        t = np.zeros(len(r), np.float32)
        inst = np.zeros(len(r), np.uint32)
        prim = np.zeros(len(r), np.uint32)
        lib.orc_closest_hit(self.h, _p(r), len(r), _p(t), _p(inst), _p(prim), int(brute))
        return t, inst, prim

    def primary_gbuffer(self, cam, materials, width: int, height: int, threads: int = 0) -> dict:
        n = width * height
        out = {"position": np.zeros((n, 3), np.float32), "normal": np.zeros((n, 3), np.float32),
               "albedo": np.zeros(n, np.uint32), "material": np.zeros(n, np.uint32),
               "emission": np.zeros(n, np.uint32)}
        mats = (OrcMaterial * max(len(materials), 1))()
        for i, m in enumerate(materials):
            C.memmove(C.addressof(mats[i]), C.addressof(m), 16)
        r = lib.orc_primary_gbuffer(self.h, C.byref(camera_from(cam)), C.cast(mats, C.c_void_p), len(materials),
                                    width, height, _p(out["position"]), _p(out["normal"]), _p(out["albedo"]),
                                    _p(out["material"]), _p(out["emission"]), threads or threads_default())
        if r != 0:
            raise ValueError("orc_primary_gbuffer failed")
        return out

    def trace_visibility(self, g: dict, width: int, height: int, params, threads: int = 0, f64: bool = False,
                         row_step: int = 1, wide: bool = False):
        """Shadow counts [lights][W*H], AO counts [W*H], stats.  With row_step > 1 only
        rows 0, row_step, ... are traced; the other rows stay zero."""
        gb, keep = gbuffer_struct(g, width, height)
        p = params if isinstance(params, OrcParams) else params_from(params)
        nl = max(p.num_lights, 1)
        shadow = np.zeros((nl, width * height), np.uint8)
        ao = np.zeros(width * height, np.uint8)
        st = OrcStats()
        if f64:
            r = lib.orc_trace_visibility_f64(self.h, C.byref(gb), C.byref(p), _p(shadow), _p(ao),
                                             threads or threads_default(), row_step)
        elif wide:
            r = lib.orc_trace_visibility_wide(self.h, C.byref(gb), C.byref(p), _p(shadow), _p(ao), C.byref(st),
                                              threads or threads_default(), row_step)
        else:
            r = lib.orc_trace_visibility(self.h, C.byref(gb), C.byref(p), _p(shadow), _p(ao), C.byref(st),
                                         threads or threads_default(), row_step)
        if r != 0:
            raise ValueError("orc_trace_visibility failed")
        return shadow[: p.num_lights], ao, st


def gbuffer_struct(g: dict, width: int, height: int):
    keep = {k: np.ascontiguousarray(g[k]) for k in ("position", "normal", "albedo", "material", "emission")}
    gb = OrcGBuffer(width, height, _p(keep["position"]), _p(keep["normal"]), _p(keep["albedo"]),
                    _p(keep["material"]), _p(keep["emission"]))
    return gb, keep


def visibility_rays(g: dict, width: int, height: int, params) -> np.ndarray:
    gb, keep = gbuffer_struct(g, width, height)
    p = params if isinstance(params, OrcParams) else params_from(params)
    n = lib.orc_visibility_rays(C.byref(gb), C.byref(p), None, 0)
    rays = np.zeros((n, 8), np.float32)
    lib.orc_visibility_rays(C.byref(gb), C.byref(p), _p(rays), n)
    return rays


def shade(g: dict, width: int, height: int, params, shadow, ao, f64: bool = False, threads: int = 0):
    gb, keep = gbuffer_struct(g, width, height)
    p = params if isinstance(params, OrcParams) else params_from(params)
    sh = np.ascontiguousarray(shadow, np.uint8)
    a = np.ascontiguousarray(ao, np.uint8)
    out = np.zeros((width * height, 4), np.float64 if f64 else np.float32)
    fn = lib.orc_shade_f64 if f64 else lib.orc_shade
    fn(C.byref(gb), C.byref(p), _p(sh), _p(a), _p(out), threads or threads_default())
    return out


def taa(light, history, position, normal, width, height, view_proj, prev_view_proj, inv_view_proj,
        jitter=(0.0, 0.0), prev_jitter=(0.0, 0.0), reconstruct=0) -> np.ndarray:
    """Temporal accumulation (taa.comp:main) of (H*W, 4) FP32 images; returns (H*W, 4)."""
    arrs = [np.ascontiguousarray(a, np.float32) for a in (light, history, position, normal)]
    d = OrcTaaDesc()
    d.width, d.height = width, height
    d.light, d.history, d.position, d.normal = (_p(a) for a in arrs)
    for name, m in (("view_proj", view_proj), ("prev_view_proj", prev_view_proj), ("inv_view_proj", inv_view_proj)):
        getattr(d, name)[:] = [float(v) for v in np.asarray(m, np.float32).reshape(16)]
    d.jitter[:] = [float(v) for v in np.asarray(jitter, np.float32)]
    d.prev_jitter[:] = [float(v) for v in np.asarray(prev_jitter, np.float32)]
    d.reconstruct = reconstruct
    out = np.empty((width * height, 4), np.float32)
    if lib.orc_taa(C.byref(d), _p(out)) != 0:
        raise ValueError("orc_taa failed")
    return out


def compose(light):
    """ACES + gamma of (N, 4) FP32: (RGBA8 u32 (N,), RGBA32F (N, 4))."""
    lt = np.ascontiguousarray(light, np.float32).reshape(-1, 4)
    n = len(lt)
    rgba8 = np.empty(n, np.uint32)
    rgb = np.empty((n, 4), np.float32)
    lib.orc_compose(_p(lt), n, _p(rgba8), _p(rgb))
    return rgba8, rgb
