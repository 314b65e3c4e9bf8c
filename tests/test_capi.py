"""C-ABI boundary checks that need no GPU: the library loads, exports every
function include/umbreon_gpu.h declares, the Python struct mirrors match the
header's layout, and failures come back as status codes with messages."""
from __future__ import annotations

import ctypes as C
import re
import subprocess
from pathlib import Path

import pytest

from umbreon_amd import capi

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "umbreon_gpu.h"


def _declared():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(ubg_[a-z_]+)\s*\(", txt)))


def test_header_declarations_are_mirrored():
    assert _declared() == sorted(capi.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = capi.load()
    for name in _declared():
        assert hasattr(lib, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", str(capi.GPU_LIB)], capture_output=True,
                         text=True, check=True).stdout
    exported = set(re.findall(r"\bT (ubg_\w+)", out))
    assert set(_declared()) <= exported


def _c_sizeof(structs):
    """Compile a tiny C program against the header and print sizeof/offsets."""
    src = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', 'int main(void){']
    for s in structs:
        src.append(f'printf("{s} %zu\\n", sizeof({s}));')
    src.append('printf("frame.stats %zu\\n", offsetof(ubg_frame, stats));')
    src.append('printf("frame.kernel_ms %zu\\n", offsetof(ubg_frame, kernel_ms));')
    src.append('printf("params.flags %zu\\n", offsetof(ubg_render_params, flags));')
    src.append('return 0;}')
    import tempfile
    with tempfile.TemporaryDirectory() as td:
        c = Path(td) / "s.c"
        c.write_text("\n".join(src))
        exe = Path(td) / "s"
        subprocess.run(["gcc", "-std=c99", str(c), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return dict(line.split() for line in out.strip().splitlines())


def test_struct_layouts_match_header():
    m = {"ubg_material": capi.Material, "ubg_scene_desc": capi.SceneDesc, "ubg_camera": capi.Camera,
         "ubg_light": capi.Light, "ubg_fog": capi.Fog, "ubg_render_params": capi.RenderParams,
         "ubg_ray_stats": capi.RayStats, "ubg_frame": capi.Frame, "ubg_scene_info": capi.SceneInfo,
         "ubg_ray": capi.Ray, "ubg_hit": capi.Hit}
    sizes = _c_sizeof(list(m))
    for k, v in m.items():
        assert int(sizes[k]) == C.sizeof(v), k
    assert int(sizes["frame.stats"]) == capi.Frame.stats.offset
    assert int(sizes["frame.kernel_ms"]) == capi.Frame.kernel_ms.offset
    assert int(sizes["params.flags"]) == capi.RenderParams.flags.offset
    assert C.sizeof(capi.Ray) == capi.ray_dtype().itemsize
    assert C.sizeof(capi.Hit) == capi.hit_dtype().itemsize


def test_status_strings_and_null_arguments():
    lib = capi.load()
    assert lib.ubg_abi_version() == capi.ABI_VERSION
    assert lib.ubg_status_string(capi.UBG_ERR_UNSUPPORTED) == b"unsupported"
    assert lib.ubg_context_create(0, None) == capi.UBG_ERR_INVALID_ARGUMENT
    assert b"null" in lib.ubg_last_error()
    info = capi.SceneInfo()
    assert lib.ubg_scene_get_info(None, C.byref(info)) == capi.UBG_ERR_INVALID_ARGUMENT
    assert lib.ubg_render(None, None, None, None) == capi.UBG_ERR_INVALID_ARGUMENT


def test_no_device_is_an_error_not_a_fallback():
    lib = capi.load()
    h = C.c_void_p()
    st = lib.ubg_context_create(0, C.byref(h))
    if st == capi.UBG_OK:
        lib.ubg_context_destroy(h)
        pytest.skip("a GPU is present; this checks the no-device path")
    assert st == capi.UBG_ERR_DEVICE
    assert lib.ubg_last_error()


def test_unsupported_options_raise():
    from umbreon_amd.scene import RenderOptions, Scene, params
    # GI combinations beyond the pt1 full-res gather are refused by the
    # library itself (UBG_ERR_UNSUPPORTED); the parameters pass through
    p, _ = params(Scene(), RenderOptions(gi=True, gi_integrator=1, pt1_denoise=False, pt1_spp=4))
    assert (p.gi, p.gi_integrator, p.pt1_spp, p.pt1_denoise) == (1, 1, 4, 0)
    with pytest.raises(capi.UmbreonError):
        params(Scene(), RenderOptions(denoiser=1))
    p, _ = params(Scene(), RenderOptions(aa_mode=1, aa_threshold=0.25, aa_depth=4))
    assert (p.aa_mode, p.aa_depth) == (1, 4) and abs(p.aa_threshold - 0.25) < 1e-7
