"""The C++ drop-in facade (umbreon::render / renderAsync, libumbreon.so) on
the GPU: runs the compiled self-test and the minimal example."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "umbreon_amd" / "lib"

pytestmark = pytest.mark.gpu


def test_cpp_api_selftest(tmp_path):
    r = subprocess.run([str(LIB / "api_selftest")], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failure(s)" in r.stdout


def test_minimal_example_writes_image(tmp_path):
    r = subprocess.run([str(LIB / "minimal_render")], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    assert r.returncode == 0, r.stdout + r.stderr
    ppm = (tmp_path / "minimal.ppm").read_bytes()
    assert ppm.startswith(b"P6\n256 256\n255\n") and len(ppm) == len(b"P6\n256 256\n255\n") + 256 * 256 * 3
