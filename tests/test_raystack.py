"""The traversal stack's bookkeeping (lucida_amd/csrc/raystack.hpp: register
top, LDS ring, device-memory spill), compiled for the host with g++ and
driven against a plain array stack by tests/raystack_check.cpp."""
from __future__ import annotations

import shutil
import subprocess
import tempfile
from pathlib import Path

import pytest

from conftest import ROOT


def test_stack_matches_a_plain_stack():
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    with tempfile.TemporaryDirectory() as d:
        exe = Path(d) / "raystack_check"
        subprocess.run([gxx, "-std=c++17", "-O1", "-Wall", f"-I{ROOT / 'lucida_amd' / 'csrc'}",
                        str(ROOT / "tests" / "raystack_check.cpp"), "-o", str(exe)], check=True)
        out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok")
