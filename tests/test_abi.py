"""The C-ABI library loads, exports every function the public headers declare,
and keeps the documented struct layouts and error behaviour (no GPU needed)."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import pytest

from luz_amd import api, denoiser

ROOT = Path(__file__).resolve().parents[1]
DECL = re.compile(r"^\s*(?:const\s+)?(?:\w+\s*\*?\s+)+\**\s*(luz(?:rt|scene|gpu)_\w+)\s*\(", re.M)


def _declared():
    names = set()
    for h in sorted((ROOT / "include").glob("*.h")):
        names.update(DECL.findall(h.read_text()))
    return names


def test_every_declared_function_is_exported():
    names = _declared()
    assert len(names) >= 35
    missing = [n for n in sorted(names) if not hasattr(api.lib, n)]
    assert not missing, missing
    # and the Python binding covers them all
    assert names <= set(api.EXPORTED) | set(denoiser.EXPORTED) | {"luzrt_status_string"}


def test_struct_layouts():
    assert C.sizeof(api.Light) == 64           # LightBlock subset, 16-byte rows
    assert C.sizeof(api.SceneBlock) == 4160
    assert C.sizeof(api.Instance) == 64
    assert C.sizeof(api.Material) == 16
    assert C.sizeof(api.Camera) == 56
    assert api.SceneBlock.num_lights.offset == 64 * 64 + 28


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_struct_layouts_match_the_c_header(tmp_path):
    """sizeof/offsetof of the structs passed by pointer, from the C compiler."""
    src = tmp_path / "layout.c"
    src.write_text("""
#include <stddef.h>
#include <stdio.h>
#include "luzrt.h"
int main(void) {
  printf("%zu %zu %zu %zu\\n", sizeof(luzrt_batch_scene), offsetof(luzrt_batch_scene, camera),
         offsetof(luzrt_batch_scene, rgba32f_out), sizeof(luzrt_dn_frame));
  printf("%zu %zu %zu %zu\\n", offsetof(luzrt_dn_frame, ao_samples), sizeof(luzrt_timings),
         offsetof(luzrt_timings, denoise_ms), offsetof(luzrt_timings, rays));
  printf("%zu %zu %zu %zu %zu\\n", sizeof(luzrt_gbuffer), offsetof(luzrt_gbuffer, normal_stride),
         offsetof(luzrt_gbuffer, depth), offsetof(luzrt_gbuffer, inverse_view), sizeof(luzrt_gbuffer_out));
  return 0;
}
""")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [C.sizeof(api.BatchScene), api.BatchScene.camera.offset, api.BatchScene.rgba32f_out.offset,
            C.sizeof(denoiser.DnFrame), denoiser.DnFrame.ao_samples.offset, C.sizeof(api.Timings),
            api.Timings.denoise_ms.offset, api.Timings.rays.offset,
            C.sizeof(api.GBuffer), api.GBuffer.normal_stride.offset, api.GBuffer.depth.offset,
            api.GBuffer.inverse_view.offset, C.sizeof(api.GBufferOut)]
    assert got == want


def test_status_strings_and_null_context():
    assert api.lib.luzrt_status_string(0) == b"ok"
    assert api.lib.luzrt_status_string(-4) == b"object in use"
    assert api.lib.luzrt_trace_visibility(None, None, None, None, None, None) == -1
    assert api.lib.luzrt_light_pass(None, None, None, None, None) == -1
    assert api.lib.luzrt_sync(None) == -1
    assert api.lib.luzrt_last_error(None) == b"null context"


def _has_gpu():
    try:
        import torch
        return torch.cuda.device_count() > 0
    except Exception:
        return False


@pytest.mark.skipif(_has_gpu(), reason="checks the no-device error path")
def test_create_without_device_reports_no_device():
    h = C.c_void_p()
    assert api.lib.luzrt_create(C.byref(h), 0) == -6
    with pytest.raises(api.LuzError):
        api.Context(0)
