"""ctypes bindings of the luzrt C ABI (include/luzrt.h, include/luzscene.h).

The native library luz_amd/_lib/libluzrt.so is required: importing this
module raises ImportError when it has not been built (there is no fallback).

Names follow the reference interface they replace:
  create_blas / cmd_build_blas  -> vkw::CreateBLAS + CmdBuildBLAS (VulkanWrapper.cpp:743-836, 1089-1104)
  create_tlas / cmd_build_tlas  -> vkw::CreateTLAS + CmdBuildTLAS (VulkanWrapper.cpp:657-741, 1106-1139)
  GPUScene.add_mesh / update_resources_gpu -> GPUScene::AddMesh / UpdateResourcesGPU (GPUScene.cpp:108-138, 348-366)
  light_pass -> DeferredRenderer::LightPass (DeferredRenderer.cpp:324-345)
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libluzrt.so"

if not LIB_PATH.exists():
    raise ImportError(f"luzrt native library missing at {LIB_PATH}; build it with `python -m luz_amd.build`")

lib = C.CDLL(str(LIB_PATH))

MAX_LIGHTS = 64
MAX_INSTANCES = 8192
LIGHT_POINT, LIGHT_SPOT, LIGHT_DIRECTIONAL = 0, 1, 2
SCENE_CUBE, SCENE_SPHERES, SCENE_TERRAIN, SCENE_RANDOM = 1, 2, 4, 5
RAY_ORDER_AUTO, RAY_ORDER_SLOT, RAY_ORDER_GROUPED = 0, 1, 2
PRECISION_PLAIN, PRECISION_COMPACT, PRECISION_WIDE = 0, 1, 2
SHADOW_DISABLED, SHADOW_RAYTRACED, SHADOW_MAP = 0, 1, 2

STATUS = {0: "ok", -1: "invalid argument", -2: "out of memory", -3: "device error", -4: "in use",
          -5: "limit exceeded", -6: "no device", -7: "unsupported"}


class LuzError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS.get(status, status)}: {message}")
        self.status = status


class Light(C.Structure):
    _fields_ = [("color", C.c_float * 3), ("intensity", C.c_float),
                ("position", C.c_float * 3), ("inner_angle", C.c_float),
                ("direction", C.c_float * 3), ("outer_angle", C.c_float),
                ("type", C.c_int32), ("num_shadow_samples", C.c_int32),
                ("radius", C.c_float), ("pad", C.c_int32)]


class SceneBlock(C.Structure):
    _fields_ = [("lights", Light * MAX_LIGHTS),
                ("ambient_color", C.c_float * 3), ("ambient_intensity", C.c_float),
                ("cam_pos", C.c_float * 3), ("num_lights", C.c_int32),
                ("ao_min", C.c_float), ("ao_max", C.c_float),
                ("ao_num_samples", C.c_int32), ("frame", C.c_int32),
                ("shadow_type", C.c_int32), ("pad", C.c_int32 * 3)]

    def __init__(self, *args, **kwargs):
        # a new block is ray-traced, as a new Luz scene is (SceneAsset::shadowType);
        # the zeroed C struct would mean ShadowDisabled
        super().__init__(*args, **kwargs)
        if "shadow_type" not in kwargs:
            self.shadow_type = SHADOW_RAYTRACED


class GBuffer(C.Structure):
    """luzrt_gbuffer: planes with optional strides/pitches (0 = packed) and the
    depth mode (position None, depth + inverse projection and view)."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("position", C.c_void_p), ("normal", C.c_void_p), ("albedo", C.c_void_p),
                ("material", C.c_void_p), ("emission", C.c_void_p),
                ("position_stride", C.c_uint32), ("normal_stride", C.c_uint32),
                ("position_pitch", C.c_uint32), ("normal_pitch", C.c_uint32),
                ("albedo_pitch", C.c_uint32), ("material_pitch", C.c_uint32),
                ("emission_pitch", C.c_uint32), ("depth_pitch", C.c_uint32),
                ("depth", C.c_void_p),
                ("inverse_proj", C.c_float * 16), ("inverse_view", C.c_float * 16)]


class GBufferOut(C.Structure):
    """luzrt_gbuffer_out: tightly packed planes written by the primary rays."""
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32),
                ("position", C.c_void_p), ("normal", C.c_void_p), ("albedo", C.c_void_p),
                ("material", C.c_void_p), ("emission", C.c_void_p)]


class Camera(C.Structure):
    _fields_ = [("eye", C.c_float * 3), ("forward", C.c_float * 3), ("right", C.c_float * 3),
                ("up", C.c_float * 3), ("tmin", C.c_float), ("tmax", C.c_float)]


class Material(C.Structure):
    _fields_ = [("albedo", C.c_uint32), ("material", C.c_uint32), ("emission", C.c_uint32),
                ("pad", C.c_uint32)]


class Instance(C.Structure):
    _fields_ = [("blas", C.c_void_p), ("object_to_world", (C.c_float * 4) * 3),
                ("custom_index", C.c_uint32), ("mask", C.c_uint32)]


class ShadowMap(C.Structure):
    _fields_ = [("light_type", C.c_int32), ("resolution", C.c_uint32), ("view_proj", C.c_float * 16),
                ("inv_view_proj", C.c_float * 16), ("light_pos", C.c_float * 3), ("z_far", C.c_float),
                ("depth", C.c_void_p)]

    def as_dict(self) -> dict:
        return {"light_type": self.light_type, "resolution": self.resolution,
                "view_proj": np.array(self.view_proj[:], np.float32),
                "inv_view_proj": np.array(self.inv_view_proj[:], np.float32),
                "light_pos": np.array(self.light_pos[:], np.float32), "z_far": self.z_far}

    @property
    def texels(self) -> int:
        return (6 if self.light_type == LIGHT_POINT else 1) * self.resolution ** 2


class TaaDesc(C.Structure):
    _fields_ = [("width", C.c_uint32), ("height", C.c_uint32), ("light", C.c_void_p),
                ("history", C.c_void_p), ("out", C.c_void_p), ("position", C.c_void_p),
                ("normal", C.c_void_p), ("view_proj", C.c_float * 16), ("prev_view_proj", C.c_float * 16),
                ("inv_view_proj", C.c_float * 16), ("jitter", C.c_float * 2), ("prev_jitter", C.c_float * 2),
                ("reconstruct", C.c_int32), ("pad", C.c_int32)]


class Timings(C.Structure):
    _fields_ = [("build_ms", C.c_float), ("compact_ms", C.c_float), ("raygen_ms", C.c_float),
                ("trace_ms", C.c_float), ("resolve_ms", C.c_float), ("shade_ms", C.c_float),
                ("primary_ms", C.c_float), ("denoise_ms", C.c_float), ("post_ms", C.c_float),
                ("rays", C.c_uint64),
                ("active_pixels", C.c_uint64),
                ("traced_rays", C.c_uint64)]


class BatchScene(C.Structure):
    _fields_ = [("tlas", C.c_void_p), ("gbuffer", GBuffer), ("scene", SceneBlock),
                ("camera", C.POINTER(Camera)), ("materials", C.POINTER(Material)),
                ("material_count", C.c_uint32), ("shadow_counts", C.c_void_p),
                ("ao_counts", C.c_void_p), ("rgba32f_out", C.c_void_p)]


assert C.sizeof(Light) == 64
assert C.sizeof(SceneBlock) == 64 * 64 + 64
assert C.sizeof(Instance) == 64

_P = C.c_void_p
_sigs = {
    "luzrt_create": (C.c_int, [C.POINTER(_P), C.c_int]),
    "luzrt_destroy": (C.c_int, [_P]),
    "luzrt_last_error": (C.c_char_p, [_P]),
    "luzrt_status_string": (C.c_char_p, [C.c_int]),
    "luzrt_sync": (C.c_int, [_P]),
    "luzrt_get_timings": (C.c_int, [_P, C.POINTER(Timings)]),
    "luzrt_device_info": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]),
    "luzrt_set_trace_config": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]),
    "luzrt_set_ray_order": (C.c_int, [_P, C.c_uint32]),
    "luzrt_set_group_culling": (C.c_int, [_P, C.c_int]),
    "luzrt_set_stage_timing": (C.c_int, [_P, C.c_int]),
    "luzrt_set_step_limit": (C.c_int, [_P, C.c_uint32]),
    "luzrt_trace_visibility_static": (C.c_int, [_P, _P, C.POINTER(GBuffer), C.POINTER(SceneBlock), _P, _P]),
    "luzrt_set_precision": (C.c_int, [_P, C.c_int]),
    "luzrt_blas_wide_arrays": (C.c_int, [_P, _P, C.POINTER(_P), C.POINTER(C.c_uint32), C.POINTER(_P),
                                         C.POINTER(_P), C.POINTER(C.c_uint32)]),
    "luzrt_malloc": (C.c_int, [_P, C.c_size_t, C.POINTER(_P)]),
    "luzrt_free": (C.c_int, [_P, _P]),
    "luzrt_upload": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "luzrt_download": (C.c_int, [_P, _P, _P, C.c_size_t]),
    "luzrt_blas_build": (C.c_int, [_P, _P, C.c_uint32, C.c_uint32, _P, C.c_uint32, C.POINTER(_P)]),
    "luzrt_blas_refit": (C.c_int, [_P, _P, _P]),
    "luzrt_blas_destroy": (C.c_int, [_P, _P]),
    "luzrt_blas_arrays": (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_uint32), C.POINTER(_P),
                                    C.POINTER(_P), C.POINTER(C.c_uint32)]),
    "luzrt_tlas_create": (C.c_int, [_P, C.c_uint32, C.POINTER(_P)]),
    "luzrt_tlas_build": (C.c_int, [_P, _P, C.POINTER(Instance), C.c_uint32]),
    "luzrt_tlas_destroy": (C.c_int, [_P, _P]),
    "luzrt_tlas_arrays": (C.c_int, [_P, C.POINTER(_P), C.POINTER(C.c_uint32), C.POINTER(_P),
                                    C.POINTER(_P), C.POINTER(C.c_uint32)]),
    "luzrt_trace_rays": (C.c_int, [_P, _P, _P, C.c_uint32, _P]),
    "luzrt_trace_visibility": (C.c_int, [_P, _P, C.POINTER(GBuffer), C.POINTER(SceneBlock), _P, _P]),
    "luzrt_shade": (C.c_int, [_P, C.POINTER(GBuffer), C.POINTER(SceneBlock), _P, _P, _P]),
    "luzrt_light_pass": (C.c_int, [_P, _P, C.POINTER(GBuffer), C.POINTER(SceneBlock), _P]),
    "luzrt_shade_fractions": (C.c_int, [_P, C.POINTER(GBuffer), C.POINTER(SceneBlock), _P, _P, _P]),
    "luzrt_taa": (C.c_int, [_P, C.POINTER(TaaDesc)]),
    "luzrt_shadow_map_build": (C.c_int, [_P, _P, C.POINTER(ShadowMap)]),
    "luzrt_shadow_map_visibility": (C.c_int, [_P, C.POINTER(GBuffer), C.POINTER(SceneBlock), _P, C.c_uint32, _P]),
    "luzrt_light_pass_shadow_maps": (C.c_int, [_P, _P, C.POINTER(GBuffer), C.POINTER(SceneBlock), _P, C.c_uint32,
                                               _P]),
    "luzrt_compose": (C.c_int, [_P, C.c_uint32, C.c_uint32, _P, _P, _P]),
    "luzrt_primary_gbuffer": (C.c_int, [_P, _P, C.POINTER(Camera), C.POINTER(Material), C.c_uint32,
                                        C.POINTER(GBuffer)]),
    "luzscene_generate": (_P, [C.c_int, C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32]),
    "luzscene_destroy": (None, [_P]),
    "luzscene_mesh_count": (C.c_uint32, [_P]),
    "luzscene_mesh": (C.c_int, [_P, C.c_uint32, C.POINTER(_P), C.POINTER(C.c_uint32), C.POINTER(_P),
                                C.POINTER(C.c_uint32)]),
    "luzscene_node_count": (C.c_uint32, [_P]),
    "luzscene_nodes": (C.c_int, [_P, _P, _P, _P]),
    "luzscene_pack": (C.c_int, [_P, C.POINTER(SceneBlock)]),
    "luzscene_camera": (C.c_int, [_P, C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(Camera)]),
    "luzscene_animate": (C.c_int, [_P, C.c_uint32]),
    "luzscene_view_proj": (C.c_int, [_P, _P, _P, _P]),
    "luzscene_camera_matrices": (C.c_int, [_P, _P, _P, _P, _P, _P]),
    "luzscene_jitter": (None, [C.c_uint32, C.c_uint32, C.c_uint32, _P]),
    "luzscene_shadow_map": (C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(ShadowMap)]),
    "luzscene_set_shadow_type": (C.c_int, [_P, C.c_int]),
    "luzscene_load": (_P, [C.c_char_p, C.c_char_p, C.POINTER(C.c_uint32), C.c_char_p, C.c_uint32]),
    "luzscene_save": (C.c_int, [_P, C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32]),
    "luzscene_set_camera": (C.c_int, [_P, _P, _P, C.c_float, C.c_uint32, C.c_uint32, C.c_float, C.c_float]),
    "luzscene_add_light": (C.c_int, [_P, C.c_int, _P, _P, _P, C.c_float, C.c_float]),
    "luzscene_set_samples": (C.c_int, [_P, C.c_int, C.c_int, C.c_float, C.c_float]),
    "luzscene_instanced_triangles": (C.c_uint64, [_P]),
    "luzgpu_create": (_P, [_P, _P]),
    "luzgpu_destroy": (None, [_P]),
    "luzgpu_update": (C.c_int, [_P, _P]),
    "luzgpu_tlas": (_P, [_P]),
    "luzgpu_blas": (_P, [_P, C.c_uint32]),
    "luzgpu_vertices": (_P, [_P, C.c_uint32]),
    "luzgpu_instances": (C.c_uint32, [_P, _P, _P]),
    "luzrt_batch_create": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(_P)]),
    "luzrt_batch_add_refit": (C.c_int, [_P, _P, _P, _P]),
    "luzrt_batch_set_instances": (C.c_int, [_P, _P, C.c_uint32, C.POINTER(Instance), C.c_uint32]),
    "luzrt_batch_frame": (C.c_int, [_P, _P, C.c_uint32, C.c_int]),
    "luzrt_batch_rays": (C.c_int, [_P, _P, C.POINTER(C.c_uint64)]),
    "luzrt_batch_destroy": (C.c_int, [_P, _P]),
}
for _name, (_res, _args) in _sigs.items():
    _f = getattr(lib, _name)
    _f.restype = _res
    _f.argtypes = _args

EXPORTED = tuple(_sigs)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ------------------------------------------------------------------ context
class DeviceBuffer:
    """A device allocation owned by a Context."""

    def __init__(self, ctx: "Context", nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = _P()
        ctx._check(lib.luzrt_malloc(ctx.handle, max(self.nbytes, 1), C.byref(p)))
        self.ptr = p.value

    def upload(self, arr: np.ndarray) -> "DeviceBuffer":
        arr = np.ascontiguousarray(arr)
        assert arr.nbytes <= self.nbytes, (arr.nbytes, self.nbytes)
        self.ctx._check(lib.luzrt_upload(self.ctx.handle, self.ptr, _ptr(arr), arr.nbytes))
        return self

    def download(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        assert out.nbytes <= self.nbytes
        self.ctx._check(lib.luzrt_download(self.ctx.handle, _ptr(out), self.ptr, out.nbytes))
        return out

    def free(self) -> None:
        if self.ptr:
            lib.luzrt_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.handle:
                self.free()
        except Exception:
            pass


class Context:
    """One device and one stream (luzrt_ctx)."""

    def __init__(self, device: int = 0):
        h = _P()
        s = lib.luzrt_create(C.byref(h), device)
        if s != 0:
            raise LuzError(s, f"luzrt_create(device={device}) failed")
        self.handle = h.value

    def _check(self, s: int) -> None:
        if s != 0:
            raise LuzError(s, lib.luzrt_last_error(self.handle).decode())

    def close(self) -> None:
        if self.handle:
            lib.luzrt_destroy(self.handle)
            self.handle = None

    def sync(self) -> None:
        self._check(lib.luzrt_sync(self.handle))

    def device_info(self) -> tuple:
        """(compute units, peak engine clock in kHz) of the context's device."""
        cu, clk = C.c_uint32(), C.c_uint32()
        self._check(lib.luzrt_device_info(self.handle, C.byref(cu), C.byref(clk)))
        return cu.value, clk.value

    def timings(self) -> Timings:
        t = Timings()
        self._check(lib.luzrt_get_timings(self.handle, C.byref(t)))
        return t

    def set_trace_config(self, groups_per_cu: int = 0, refill: int = 0, stack_entries: int = 0,
                         tile_width: int = 0) -> None:
        self._check(lib.luzrt_set_trace_config(self.handle, groups_per_cu, refill, stack_entries, tile_width))

    def set_ray_order(self, order: int) -> None:
        """RAY_ORDER_AUTO, RAY_ORDER_SLOT or RAY_ORDER_GROUPED (same results, different speed)."""
        self._check(lib.luzrt_set_ray_order(self.handle, order))

    def set_group_culling(self, enable: bool) -> None:
        """Ray-group culling of visibility calls on (default) or off (same results)."""
        self._check(lib.luzrt_set_group_culling(self.handle, 1 if enable else 0))

    def set_stage_timing(self, enable: bool) -> None:
        """Per-stage HIP events on (default) or off; off, timings() reports 0 per stage."""
        self._check(lib.luzrt_set_stage_timing(self.handle, 1 if enable else 0))

    def set_step_limit(self, max_steps: int) -> None:
        """Traversal steps per ray before it is abandoned (0: the default 2^24)."""
        self._check(lib.luzrt_set_step_limit(self.handle, max_steps))

    def set_precision(self, mode: int) -> None:
        """PRECISION_PLAIN, PRECISION_COMPACT (8-wide 8-bit BVH, packed-FP16 shading) or
        PRECISION_WIDE (the 8-wide BVH with FP32 shading: PLAIN's results)."""
        self._check(lib.luzrt_set_precision(self.handle, mode))

    def buffer(self, nbytes: int) -> DeviceBuffer:
        return DeviceBuffer(self, nbytes)

    def to_device(self, arr: np.ndarray) -> DeviceBuffer:
        arr = np.ascontiguousarray(arr)
        return DeviceBuffer(self, arr.nbytes).upload(arr)

    # -------------------------------------------------- acceleration structures
    def create_blas(self, vertices: DeviceBuffer, stride: int, vertex_count: int,
                    indices: DeviceBuffer, index_count: int) -> "Blas":
        h = _P()
        self._check(lib.luzrt_blas_build(self.handle, vertices.ptr, stride, vertex_count,
                                         indices.ptr, index_count, C.byref(h)))
        return Blas(self, h.value, index_count // 3)

    def create_tlas(self, max_instances: int) -> "Tlas":
        h = _P()
        self._check(lib.luzrt_tlas_create(self.handle, max_instances, C.byref(h)))
        return Tlas(self, h.value)

    # -------------------------------------------------- the lighting path
    def trace_rays(self, tlas: "Tlas", rays: DeviceBuffer, count: int, hits: DeviceBuffer) -> None:
        self._check(lib.luzrt_trace_rays(self.handle, tlas.handle, rays.ptr, count, hits.ptr))

    def trace_visibility(self, tlas: "Tlas", g: GBuffer, scene: SceneBlock, shadow: DeviceBuffer,
                         ao: DeviceBuffer) -> None:
        self._check(lib.luzrt_trace_visibility(self.handle, tlas.handle, C.byref(g), C.byref(scene),
                                               shadow.ptr, ao.ptr))

    def shade(self, g: GBuffer, scene: SceneBlock, shadow: DeviceBuffer, ao: DeviceBuffer,
              out: DeviceBuffer) -> None:
        self._check(lib.luzrt_shade(self.handle, C.byref(g), C.byref(scene), shadow.ptr, ao.ptr, out.ptr))

    def shade_fractions(self, g: GBuffer, scene: SceneBlock, visibility: DeviceBuffer, ao: DeviceBuffer,
                        out: DeviceBuffer) -> None:
        """Lighting from visibility/AO fractions (e.g. the denoiser's outputs)."""
        self._check(lib.luzrt_shade_fractions(self.handle, C.byref(g), C.byref(scene),
                                              visibility.ptr if scene.num_lights else None, ao.ptr, out.ptr))

    def shadow_map_build(self, tlas: "Tlas", m: ShadowMap) -> None:
        """Renders m.depth (device, m.texels floats) by closest-hit rays."""
        self._check(lib.luzrt_shadow_map_build(self.handle, tlas.handle, C.byref(m)))

    def shadow_map_visibility(self, g: GBuffer, scene: SceneBlock, maps, out: DeviceBuffer) -> None:
        arr = (ShadowMap * max(len(maps), 1))(*maps)
        self._check(lib.luzrt_shadow_map_visibility(self.handle, C.byref(g), C.byref(scene),
                                                    C.cast(arr, C.c_void_p), len(maps), out.ptr))

    def light_pass_shadow_maps(self, tlas: "Tlas", g: GBuffer, scene: SceneBlock, maps, out: DeviceBuffer) -> None:
        arr = (ShadowMap * max(len(maps), 1))(*maps)
        self._check(lib.luzrt_light_pass_shadow_maps(self.handle, tlas.handle, C.byref(g), C.byref(scene),
                                                     C.cast(arr, C.c_void_p), len(maps), out.ptr))

    def taa(self, desc: TaaDesc) -> None:
        """Temporal accumulation of desc.light with desc.history into desc.out."""
        self._check(lib.luzrt_taa(self.handle, C.byref(desc)))

    def compose(self, width: int, height: int, light: DeviceBuffer, rgba8: DeviceBuffer = None,
                rgba32f: DeviceBuffer = None) -> None:
        """ACES tonemap + gamma of a light buffer to RGBA8 and/or RGBA32F."""
        self._check(lib.luzrt_compose(self.handle, width, height, light.ptr, rgba8.ptr if rgba8 else None,
                                      rgba32f.ptr if rgba32f else None))

    def light_pass(self, tlas: "Tlas", g: GBuffer, scene: SceneBlock, out: DeviceBuffer) -> None:
        self._check(lib.luzrt_light_pass(self.handle, tlas.handle, C.byref(g), C.byref(scene), out.ptr))

    def primary_gbuffer(self, tlas: "Tlas", cam: Camera, materials, out: GBuffer) -> None:
        n = len(materials)
        arr = (Material * max(n, 1))(*materials)
        self._check(lib.luzrt_primary_gbuffer(self.handle, tlas.handle, C.byref(cam), arr, n, C.byref(out)))


class Blas:
    def __init__(self, ctx: Context, handle: int, tri_count: int):
        self.ctx, self.handle, self.tri_count = ctx, handle, tri_count

    def refit(self, vertices: DeviceBuffer) -> None:
        self.ctx._check(lib.luzrt_blas_refit(self.ctx.handle, self.handle, vertices.ptr))

    def arrays(self):
        """(nodes u8[n,64], tris f32[t,9], prim_ids u32[t]) downloaded from the device."""
        nodes, nn, tris, prims, nt = _P(), C.c_uint32(), _P(), _P(), C.c_uint32()
        self.ctx._check(lib.luzrt_blas_arrays(self.handle, C.byref(nodes), C.byref(nn), C.byref(tris),
                                              C.byref(prims), C.byref(nt)))
        return _download_arrays(self.ctx, nodes.value, nn.value, tris.value, prims.value, nt.value)

    def wide_arrays(self):
        """(nodes u8[n,96], tris f32[t,9], prim_ids u32[t]) of the 8-wide form."""
        nodes, nn, tris, prims, nt = _P(), C.c_uint32(), _P(), _P(), C.c_uint32()
        self.ctx._check(lib.luzrt_blas_wide_arrays(self.ctx.handle, self.handle, C.byref(nodes), C.byref(nn),
                                                   C.byref(tris), C.byref(prims), C.byref(nt)))
        n_out = np.empty((nn.value, 96), np.uint8)
        if nn.value:
            self.ctx._check(lib.luzrt_download(self.ctx.handle, _ptr(n_out), nodes.value, n_out.nbytes))
        t_out = np.empty((nt.value, 9), np.float32)
        p_out = np.empty(nt.value, np.uint32)
        if nt.value:
            self.ctx._check(lib.luzrt_download(self.ctx.handle, _ptr(t_out), tris.value, t_out.nbytes))
            self.ctx._check(lib.luzrt_download(self.ctx.handle, _ptr(p_out), prims.value, p_out.nbytes))
        return n_out, t_out, p_out

    def destroy(self) -> None:
        self.ctx._check(lib.luzrt_blas_destroy(self.ctx.handle, self.handle))
        self.handle = None


def _download_arrays(ctx, nodes, nn, tris, prims, nt):
    n_out = np.empty((nn, 64), np.uint8)
    ctx._check(lib.luzrt_download(ctx.handle, _ptr(n_out), nodes, n_out.nbytes))
    t_out = np.empty((nt, 9), np.float32)
    p_out = np.empty(nt, np.uint32)
    if nt:
        ctx._check(lib.luzrt_download(ctx.handle, _ptr(t_out), tris, t_out.nbytes))
        ctx._check(lib.luzrt_download(ctx.handle, _ptr(p_out), prims, p_out.nbytes))
    return n_out, t_out, p_out


class Tlas:
    def __init__(self, ctx: Context, handle: int):
        self.ctx, self.handle = ctx, handle

    def build(self, instances) -> None:
        n = len(instances)
        arr = (Instance * max(n, 1))(*instances)
        self.ctx._check(lib.luzrt_tlas_build(self.ctx.handle, self.handle, arr, n))

    def arrays(self):
        """(nodes u8[n,64], world_to_object f32[k,12], order u32[k])."""
        nodes, nn, w2o, order, cnt = _P(), C.c_uint32(), _P(), _P(), C.c_uint32()
        self.ctx._check(lib.luzrt_tlas_arrays(self.handle, C.byref(nodes), C.byref(nn), C.byref(w2o),
                                              C.byref(order), C.byref(cnt)))
        n_out = np.empty((nn.value, 64), np.uint8)
        self.ctx._check(lib.luzrt_download(self.ctx.handle, _ptr(n_out), nodes.value, n_out.nbytes))
        k = cnt.value
        w = np.empty((k, 12), np.float32)
        o = np.empty(k, np.uint32)
        if k:
            self.ctx._check(lib.luzrt_download(self.ctx.handle, _ptr(w), w2o.value, w.nbytes))
            self.ctx._check(lib.luzrt_download(self.ctx.handle, _ptr(o), order.value, o.nbytes))
        return n_out, w, o

    def destroy(self) -> None:
        self.ctx._check(lib.luzrt_tlas_destroy(self.ctx.handle, self.handle))
        self.handle = None


def make_instance(blas: Blas, object_to_world: np.ndarray, custom_index: int, mask: int = 0xFF) -> Instance:
    inst = Instance()
    inst.blas = blas.handle
    m = np.asarray(object_to_world, np.float32).reshape(3, 4)
    for r in range(3):
        for c in range(4):
            inst.object_to_world[r][c] = float(m[r, c])
    inst.custom_index = custom_index
    inst.mask = mask
    return inst


# ------------------------------------------------------------------ scene model
class HostScene:
    """Luz's SceneAsset with procedural content (luzscene)."""

    def __init__(self, kind: int, seed: int = 0, tri_budget: int = 0, width: int = 0, height: int = 0):
        self.missing_meshes = 0
        self.handle = lib.luzscene_generate(kind, seed, tri_budget, width, height) if kind else None
        if kind and not self.handle:
            raise ValueError(f"unknown scene kind {kind}")

    @classmethod
    def load(cls, path, bin_path=None) -> "HostScene":
        """Scene ingest: .luz (+ .luzbin), .gltf/.glb or .obj."""
        s = cls(0)
        err = C.create_string_buffer(512)
        missing = C.c_uint32()
        s.handle = lib.luzscene_load(str(path).encode(), str(bin_path).encode() if bin_path else None,
                                     C.byref(missing), err, 512)
        if not s.handle:
            raise LuzError(-1, err.value.decode())
        s.missing_meshes = missing.value
        return s

    def save(self, path, bin_path=None) -> None:
        err = C.create_string_buffer(512)
        if lib.luzscene_save(self.handle, str(path).encode(), str(bin_path).encode() if bin_path else None, err, 512):
            raise LuzError(-1, err.value.decode())

    def set_camera(self, eye, center, fov_degrees: float, width: int, height: int, near: float = 0.0,
                   far: float = 0.0) -> None:
        e, c = np.asarray(eye, np.float32), np.asarray(center, np.float32)
        if lib.luzscene_set_camera(self.handle, _ptr(e), _ptr(c), fov_degrees, width, height, near, far):
            raise LuzError(-1, "bad camera")

    def add_light(self, kind: int, position, direction=None, color=(1, 1, 1), intensity: float = 10.0,
                  radius: float = 0.0) -> None:
        p = np.asarray(position, np.float32)
        d = None if direction is None else np.asarray(direction, np.float32)
        c = np.asarray(color, np.float32)
        if lib.luzscene_add_light(self.handle, kind, _ptr(p), _ptr(d) if d is not None else None, _ptr(c),
                                  intensity, radius):
            raise LuzError(-5, "bad light or more than 64 lights")

    def shadow_map(self, light: int, resolution: int) -> ShadowMap:
        """The light's shadow-map descriptor (depth left unset)."""
        m = ShadowMap()
        if lib.luzscene_shadow_map(self.handle, light, resolution, C.byref(m)) != 0:
            raise LuzError(-1, f"no shadow map for light {light}")
        return m

    def set_shadow_type(self, kind: int) -> None:
        if lib.luzscene_set_shadow_type(self.handle, kind):
            raise LuzError(-1, "shadow type must be SHADOW_DISABLED, SHADOW_RAYTRACED or SHADOW_MAP")

    def set_samples(self, light_samples: int, ao_samples: int, ao_min: float = 1e-4, ao_max: float = 1.0) -> None:
        if lib.luzscene_set_samples(self.handle, light_samples, ao_samples, ao_min, ao_max):
            raise LuzError(-5, "bad sample counts")

    def close(self) -> None:
        if self.handle:
            lib.luzscene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def mesh_count(self) -> int:
        return lib.luzscene_mesh_count(self.handle)

    def mesh(self, i: int):
        v, nv, idx, ni = _P(), C.c_uint32(), _P(), C.c_uint32()
        if lib.luzscene_mesh(self.handle, i, C.byref(v), C.byref(nv), C.byref(idx), C.byref(ni)) != 0:
            raise IndexError(i)
        verts = np.ctypeslib.as_array(C.cast(v.value, C.POINTER(C.c_float)), shape=(nv.value, 12)).copy()
        inds = np.ctypeslib.as_array(C.cast(idx.value, C.POINTER(C.c_uint32)), shape=(ni.value,)).copy()
        return verts, inds

    def nodes(self):
        n = lib.luzscene_node_count(self.handle)
        o2w = np.zeros((n, 12), np.float32)
        mesh = np.zeros(n, np.uint32)
        mats = (Material * max(n, 1))()
        lib.luzscene_nodes(self.handle, _ptr(o2w), _ptr(mesh), C.cast(mats, C.c_void_p))
        return o2w, mesh, [mats[i] for i in range(n)]

    def pack(self) -> SceneBlock:
        s = SceneBlock()
        if lib.luzscene_pack(self.handle, C.byref(s)) != 0:
            raise LuzError(-5, "scene has more than 64 lights")
        return s

    def camera(self):
        w, h, cam = C.c_uint32(), C.c_uint32(), Camera()
        lib.luzscene_camera(self.handle, C.byref(w), C.byref(h), C.byref(cam))
        return w.value, h.value, cam

    def view_proj(self, jitter=None):
        """(view_proj, inverse) as column-major float32[16] arrays."""
        vp = np.zeros(16, np.float32)
        ivp = np.zeros(16, np.float32)
        j = None if jitter is None else np.asarray(jitter, np.float32)
        if lib.luzscene_view_proj(self.handle, _ptr(j) if j is not None else None, _ptr(vp), _ptr(ivp)) != 0:
            raise LuzError(-1, "singular view-projection")
        return vp, ivp

    def camera_matrices(self, jitter=None) -> dict:
        """view, proj and their inverses (column-major float32[16]), as SceneBlock carries them."""
        m = {k: np.zeros(16, np.float32) for k in ("view", "proj", "inverse_view", "inverse_proj")}
        j = None if jitter is None else np.asarray(jitter, np.float32)
        if lib.luzscene_camera_matrices(self.handle, _ptr(j) if j is not None else None, _ptr(m["view"]),
                                        _ptr(m["proj"]), _ptr(m["inverse_view"]), _ptr(m["inverse_proj"])) != 0:
            raise LuzError(-1, "singular camera matrix")
        return m

    def animate(self, frame: int) -> None:
        lib.luzscene_animate(self.handle, frame)

    @property
    def instanced_triangles(self) -> int:
        return lib.luzscene_instanced_triangles(self.handle)


class GPUScene:
    """GPUScene: one BLAS per mesh, TLAS over mesh nodes (luzgpu)."""

    def __init__(self, ctx: Context, scene: HostScene):
        self.ctx = ctx
        self.handle = lib.luzgpu_create(ctx.handle, scene.handle)
        if not self.handle:
            raise LuzError(-3, "GPUScene creation failed: " + lib.luzrt_last_error(ctx.handle).decode())
        self.tlas = Tlas(ctx, lib.luzgpu_tlas(self.handle))

    def update_resources_gpu(self, scene: HostScene) -> None:
        self.ctx._check(lib.luzgpu_update(self.handle, scene.handle))

    def blas(self, mesh: int) -> Blas:
        return Blas(self.ctx, lib.luzgpu_blas(self.handle, mesh), 0)

    def vertices(self, mesh: int) -> int:
        """Device address of the mesh's vertex buffer (MeshVertex array)."""
        return lib.luzgpu_vertices(self.handle, mesh)

    def instances(self, scene: HostScene):
        """The scene's TLAS instances (ctypes array) and their count."""
        n = lib.luzgpu_instances(self.handle, scene.handle, None)
        arr = (Instance * max(n, 1))()
        lib.luzgpu_instances(self.handle, scene.handle, C.cast(arr, C.c_void_p))
        return arr, n

    def close(self) -> None:
        if self.handle:
            lib.luzgpu_destroy(self.handle)
            self.handle = None


class GBufferDevice:
    """Device G-buffer planes for width x height pixels."""

    def __init__(self, ctx: Context, width: int, height: int):
        n = width * height
        self.width, self.height = width, height
        self.position = ctx.buffer(n * 12)
        self.normal = ctx.buffer(n * 12)
        self.albedo = ctx.buffer(n * 4)
        self.material = ctx.buffer(n * 4)
        self.emission = ctx.buffer(n * 4)

    def desc(self) -> GBuffer:
        return GBuffer(self.width, self.height, self.position.ptr, self.normal.ptr, self.albedo.ptr,
                       self.material.ptr, self.emission.ptr)

    def upload(self, planes: dict) -> None:
        for k in ("position", "normal", "albedo", "material", "emission"):
            getattr(self, k).upload(planes[k])

    def download(self) -> dict:
        n = self.width * self.height
        return {"position": self.position.download(np.float32, (n, 3)),
                "normal": self.normal.download(np.float32, (n, 3)),
                "albedo": self.albedo.download(np.uint32, n),
                "material": self.material.download(np.uint32, n),
                "emission": self.emission.download(np.uint32, n)}


# ------------------------------------------------------------------ batched frames
def jitter(index: int, width: int, height: int) -> np.ndarray:
    """CameraNode::NextJitter's Halton(2,3) NDC jitter of frame `index`."""
    j = np.zeros(2, np.float32)
    lib.luzscene_jitter(index, width, height, _ptr(j))
    return j


class Batch:
    """Many scenes rendered per frame (luzrt_batch): batched BLAS refit and
    TLAS rebuild, primary G-buffers, one persistent traversal launch over all
    scenes' shadow and AO rays, shading; captured into a hipGraph."""

    def __init__(self, ctx: Context, scenes):
        """scenes: list of dicts with keys tlas (Tlas), gbuffer (GBuffer),
        scene (SceneBlock), camera (Camera or None), materials (list),
        shadow, ao, out (DeviceBuffer)."""
        self.ctx = ctx
        n = len(scenes)
        self._keep = []
        arr = (BatchScene * n)()
        for i, d in enumerate(scenes):
            e = arr[i]
            e.tlas = d["tlas"].handle
            e.gbuffer = d["gbuffer"]
            e.scene = d["scene"]
            if d.get("camera") is not None:
                cam = Camera()
                C.memmove(C.byref(cam), C.byref(d["camera"]), C.sizeof(Camera))
                self._keep.append(cam)
                e.camera = C.pointer(cam)
            mats = d.get("materials") or []
            if mats:
                m = (Material * len(mats))(*mats)
                self._keep.append(m)
                e.materials = C.cast(m, C.POINTER(Material))
            e.material_count = len(mats)
            e.shadow_counts = d["shadow"].ptr
            e.ao_counts = d["ao"].ptr
            e.rgba32f_out = d["out"].ptr
        h = _P()
        ctx._check(lib.luzrt_batch_create(ctx.handle, C.cast(arr, C.c_void_p), n, C.byref(h)))
        self.handle = h.value

    def add_refit(self, blas: Blas, vertices: int) -> None:
        self.ctx._check(lib.luzrt_batch_add_refit(self.ctx.handle, self.handle, blas.handle, vertices))

    def set_instances(self, scene: int, instances, count: int) -> None:
        self.ctx._check(lib.luzrt_batch_set_instances(self.ctx.handle, self.handle, scene, instances, count))

    def frame(self, frame: int, use_graph: bool = True) -> None:
        self.ctx._check(lib.luzrt_batch_frame(self.ctx.handle, self.handle, frame, 1 if use_graph else 0))

    def rays(self) -> int:
        r = C.c_uint64()
        self.ctx._check(lib.luzrt_batch_rays(self.ctx.handle, self.handle, C.byref(r)))
        return r.value

    def close(self) -> None:
        if self.handle:
            lib.luzrt_batch_destroy(self.ctx.handle, self.handle)
            self.handle = None
