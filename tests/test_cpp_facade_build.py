"""The C++ facade builds and exports the reference's public entry points
(no GPU needed: link/symbol checks only)."""
from __future__ import annotations

import re
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "umbreon_amd" / "lib"


def _demangled_exports(lib: Path) -> str:
    return subprocess.run(["nm", "-DC", "--defined-only", str(lib)], capture_output=True, text=True,
                          check=True).stdout


def test_facade_exports_public_api():
    out = _demangled_exports(LIB / "libumbreon.so")
    for sym in [
        "umbreon::render(umbreon::Scene const&, umbreon::RenderOptions const&)",
        "umbreon::render(umbreon::Scene const&, umbreon::RenderOptions const&, umbreon::RenderProgress&)",
        "umbreon::renderAsync(umbreon::Scene, umbreon::RenderOptions)",
        "umbreon::RenderTask::get()",
        "umbreon::RenderTask::wait_for(",
        "umbreon::setLogSink(",
        "umbreon::logMessage(",
        "umbreon::boxDownsample(",
        "umbreon::applyAssumedGamma(",
        "umbreon::srgbEncode8(",
        "umbreon::srgbEncodeF(float)",
        "umbreon::srgbDecodeF(float)",
        "umbreon::toString(umbreon::RenderPhase)",
        "umbreon::RenderProgress::fraction() const",
    ]:
        assert sym in out, sym


def test_facade_links_only_the_c_abi():
    deps = subprocess.run(["readelf", "-d", str(LIB / "libumbreon.so")], capture_output=True, text=True,
                          check=True).stdout
    assert "libumbreon_gpu.so" in deps
    # the facade reaches the GPU only through the C-ABI symbols
    und = subprocess.run(["nm", "-D", "--undefined-only", str(LIB / "libumbreon.so")], capture_output=True,
                         text=True, check=True).stdout
    assert set(re.findall(r"\b(ubg_\w+)", und)) >= {"ubg_context_create", "ubg_scene_create_ex", "ubg_render"}
    assert "hipLaunchKernel" not in und


def test_examples_built():
    for exe in ("api_selftest", "molecule_async"):
        assert (LIB / exe).exists(), exe


REF_EXAMPLE = Path("/root/reference/examples/minimal_render.cpp")


@pytest.mark.skipif(not REF_EXAMPLE.exists(), reason="the reference tree is not present")
def test_reference_example_compiles_against_the_facade(tmp_path):
    """Source compatibility: the reference's own minimal consumer, compiled
    where it lies, against these public headers and linked with
    libumbreon.so (compiled and linked only; nothing from the reference is
    copied or run)."""
    exe = tmp_path / "ref_minimal_render"
    r = subprocess.run(["g++", "-std=c++17", "-O0", "-I", str(ROOT / "include"), str(REF_EXAMPLE), "-L", str(LIB),
                        "-lumbreon", "-lumbreon_gpu", f"-Wl,-rpath,{LIB}", "-pthread", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    assert exe.exists()
