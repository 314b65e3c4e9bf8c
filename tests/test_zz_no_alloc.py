"""ubg_render_noalloc (frame scratch carved from the library's fixed per-GPU
array, no device allocation) against ubg_render on the same scenes and
device output buffers: bit-identical frames and ray counts, on two calls in a
row, the second with a smaller frame (the array keeps the first call's
contents), and the refusals (host outputs, kernel timing, a frame whose
scratch does not fit)."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

from umbreon_amd import capi
from umbreon_amd.scene import params
from umbreon_amd.scenes import ball_and_stick_scene, molecular_surface_scene

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def device():
    from umbreon_amd.render import Device
    d = Device(0)
    yield d
    d.close()


class _Hip:
    """hipMalloc / hipMemcpy / hipFree through the HIP runtime (the test's
    own device buffers; only the entry point under test must not allocate)."""

    def __init__(self):
        self.lib = C.CDLL("libamdhip64.so")

    def alloc(self, nbytes: int, fill: int = 0xFF) -> C.c_void_p:
        p = C.c_void_p()
        assert self.lib.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        assert self.lib.hipMemset(p, C.c_int(fill), C.c_size_t(nbytes)) == 0
        assert self.lib.hipDeviceSynchronize() == 0
        return p

    def to_host(self, p: C.c_void_p, shape, dtype) -> np.ndarray:
        out = np.empty(shape, dtype)
        assert self.lib.hipDeviceSynchronize() == 0
        assert self.lib.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, C.c_size_t(out.nbytes), 2) == 0  # D2H
        return out

    def free(self, p: C.c_void_p) -> None:
        self.lib.hipFree(p)


def _frame(entry, bs, scene, opt, flags=0, memory=1):
    """One frame through `entry` (a C-ABI function) into device buffers;
    returns (status, host copies of colour, depth, view z, ids, AO and the
    ray statistics)."""
    hip = _Hip()
    lib = capi.load()
    p, keep = params(scene, opt, flags)
    ss = max(1, int(opt.supersample))
    hw, hh = opt.width * ss, opt.height * ss
    shapes = {"color": ((opt.height, opt.width, 4), np.float32), "depth": ((hh, hw), np.float32),
              "view_z": ((hh, hw), np.float32), "hit_id": ((hh, hw), np.uint32), "ao_open": ((hh, hw), np.float32)}
    bufs = {k: hip.alloc(int(np.prod(sh)) * 4) for k, (sh, _) in shapes.items()}
    try:
        fr = capi.Frame()
        fr.color, fr.depth, fr.view_z = bufs["color"], bufs["depth"], bufs["view_z"]
        fr.hit_id, fr.ao_open = bufs["hit_id"], bufs["ao_open"]
        fr.memory = memory
        st = getattr(lib, entry)(bs.device.handle, bs.handle, C.byref(p), C.byref(fr))
        del keep
        if st != 0:
            return st, None
        host = {k: hip.to_host(bufs[k], sh, dt) for k, (sh, dt) in shapes.items()}
        host["stats"] = fr.stats.as_dict()
        host["layers"] = int(fr.layers)
        return 0, host
    finally:
        for v in bufs.values():
            hip.free(v)


def _same(a, b):
    for k in ("color", "depth", "view_z", "hit_id", "ao_open"):
        assert np.array_equal(a[k].view(np.uint32), b[k].view(np.uint32)), k
    assert a["stats"] == b["stats"] and a["layers"] == b["layers"]


def test_noalloc_matches_render_on_two_calls(device):
    from umbreon_amd.render import BuiltScene
    cases = [ball_and_stick_scene(n_atoms=4000, width=480, height=270, supersample=2),
             molecular_surface_scene(n_atoms=1500, width=200, height=120)]
    for scene, opt in cases:
        bs = BuiltScene(device, scene)
        try:
            st, ref = _frame("ubg_render", bs, scene, opt)
            assert st == 0
            st, got = _frame("ubg_render_noalloc", bs, scene, opt)
            assert st == 0, capi.load().ubg_last_error()
            print(f"{opt.width}x{opt.height} ss{opt.supersample}: {got['layers']} layers, stats {got['stats']}",
                  flush=True)
            _same(ref, got)
            assert (got["color"][..., 3] >= 0).all() and got["stats"]["primary_rays"] > 0
            # and the ordinary entry point still works after it (its own buffers came back)
            st, again = _frame("ubg_render", bs, scene, opt)
            assert st == 0
            _same(ref, again)
        finally:
            bs.close()


def test_noalloc_refusals(device):
    from umbreon_amd.render import BuiltScene
    scene, opt = ball_and_stick_scene(n_atoms=500, width=64, height=48)
    bs = BuiltScene(device, scene)
    try:
        # host outputs
        st, _ = _frame("ubg_render_noalloc", bs, scene, opt, memory=0)
        assert st == capi.UBG_ERR_INVALID_ARGUMENT
        # per-kernel timing creates events
        st, _ = _frame("ubg_render_noalloc", bs, scene, opt, flags=capi.FLAG_KERNEL_TIMES)
        assert st == capi.UBG_ERR_INVALID_ARGUMENT
        # a frame beyond the fixed scratch: 3840 x 2160 hi-res (rejected before any launch)
        _, big = ball_and_stick_scene(n_atoms=500, width=1920, height=1080, supersample=2)
        st, _ = _frame("ubg_render_noalloc", bs, scene, big)
        assert st == capi.UBG_ERR_INVALID_ARGUMENT
        assert b"MiB" in capi.load().ubg_last_error()
        # the entry point works again after the refusals
        st, ref = _frame("ubg_render", bs, scene, opt)
        st2, got = _frame("ubg_render_noalloc", bs, scene, opt)
        assert st == 0 and st2 == 0
        _same(ref, got)
    finally:
        bs.close()
