"""The Rust binding's scene descriptor, exercised from C: tests/abi_desc_check.c
fills lucida_scene_desc field by field as INTEGRATION.md's scene_desc() does
(an instanced scene under a sky, and a sky-lit analytic scene) and requires
films bit-identical to the same scenes uploaded through the library's own
path. The program is built with g++ by __graft_entry__.build()."""
from __future__ import annotations

import subprocess
from pathlib import Path

import pytest

BIN = Path(__file__).resolve().parent / "_bin" / "abi_desc_check"


@pytest.mark.gpu
def test_c_client_fills_the_scene_descriptor_like_the_binding():
    if not BIN.exists():
        pytest.fail(f"{BIN} is missing: run __graft_entry__.build()")
    print("case: C client abi_desc_check (instance-forest, sunset)", flush=True)
    r = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("films bit-identical") == 2


def test_c_client_source_is_plain_c():
    """The check program is C (no C++), so the boundary it exercises is the C
    ABI itself: it compiles with the C front end."""
    src = BIN.parent.parent / "abi_desc_check.c"
    r = subprocess.run(["gcc", "-x", "c", "-std=c11", "-fsyntax-only", "-Wall", "-Werror",
                        f"-I{src.parent.parent / 'include'}", str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
