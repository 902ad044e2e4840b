"""The megakernel (pt_gpu::render_native_on, crates/gpu/src/lib.rs:391;
shaders/trace/megakernel.wgsl) against the wavefront and the oracle.

Both architectures run the same shading code (shade_bounce), the same walks
and the same per-pixel sample order, so their films must be bit-identical
(the reference holds its two to mean_rel ~8e-9, README.md:1364); the
megakernel is also checked against the oracle, and both architectures must
report the same convergence (crates/gpu/tests/convergence.rs:128)."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import lucida_amd as L

pytestmark = pytest.mark.gpu

CASES = [  # scene, width, height, spp, depth, samples per batch (0: default)
    ("box-c1", 64, 64, 16, 8, 0),
    ("cornell-box", 96, 64, 8, 5, 3),
    ("cornell-mesh", 64, 64, 8, 8, 0),
    ("glass-box", 64, 64, 8, 24, 0),
    ("sunset", 64, 64, 8, 6, 0),
    ("metal-sweep", 64, 48, 8, 10, 0),
    ("instances-small", 64, 48, 6, 12, 4),
    ("instance-forest", 64, 64, 8, 5, 0),
]


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("name,w,h,spp,depth,spb", CASES)
def test_megakernel_film_equals_wavefront_film(gpu, oracle, name, w, h, spp, depth, spb):
    say(f"megakernel case: {name} {w}x{h} x {spp} spp, {depth} bounces, batch {spb or 'default'}")
    blob = L.Blob(name)
    scene = L.Scene(gpu, blob, build_lbvh=True)
    u = blob.uniforms(w, h, samples=spp, max_depth=depth)
    cfg = L.WavefrontConfig(spb, 0, 0, 0, 1, 0, 0, 0)
    fa, fb = L.Film(gpu, w, h), L.Film(gpu, w, h)
    sa = L.render_with(gpu, scene, u, spp, fa, "megakernel", cfg)
    sb = L.render_with(gpu, scene, u, spp, fb, "wavefront", cfg)
    a, b = fa.download(), fb.download()
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    assert (sa.extension_rays, sa.shadow_rays) == (sb.extension_rays, sb.shadow_rays)
    # the walks find the same hits, but a lane's speculative visits (past a
    # postponed leaf) depend on which rays share its wave
    for f in ("node_visits", "triangle_tests"):
        assert abs(getattr(sa, f) - getattr(sb, f)) <= 0.05 * getattr(sb, f), f
    assert sa.stack_overflows == 0
    assert fa.convergence() == fb.convergence()
    d, keep = scene.host_desc(blob)
    ref, _, _ = oracle.render(d, u, background=list(u.background), guides=False)
    c = oracle.compare(L.film_mean(a, w, h), ref)
    say(f"  bit-identical; vs oracle mean_rel {c['mean_rel']:.2e}, energy {c['mean_a']:.5f}/{c['mean_b']:.5f}")
    assert c["mean_rel"] < 2e-3 and abs(c["mean_a"] / c["mean_b"] - 1) < 2e-3


def test_megakernel_progressive_and_rejections(gpu):
    blob = L.Blob("box-c1")
    scene = L.Scene(gpu, blob)
    one = L.Film(gpu, 32, 32)
    L.render_megakernel(gpu, scene, blob.uniforms(32, 32, samples=6), 6, one)
    two = L.Film(gpu, 32, 32)
    L.render_megakernel(gpu, scene, blob.uniforms(32, 32, samples=2), 2, two)
    L.render_megakernel(gpu, scene, blob.uniforms(32, 32, samples=4, sample_offset=2), 4, two)
    a, b = L.film_mean(one.download(), 32, 32), L.film_mean(two.download(), 32, 32)
    np.testing.assert_allclose(a, b, rtol=1e-5, atol=1e-6)
    with pytest.raises(L.GpuError) as e:
        L.render_megakernel(gpu, scene, blob.uniforms(16, 16, samples=1), 1, one)
    assert e.value.kind == "InvalidArg"
