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


def test_async_molecule_example_writes_image(tmp_path):
    r = subprocess.run([str(LIB / "molecule_async")], capture_output=True, text=True, timeout=120, cwd=tmp_path)
    print(r.stdout[-3000:], r.stderr[-2000:])
    assert r.returncode == 0, r.stdout + r.stderr
    ppm = (tmp_path / "benzene.ppm").read_bytes()
    head = b"P6\n320 240\n255\n"
    assert ppm.startswith(head) and len(ppm) == len(head) + 320 * 240 * 3
