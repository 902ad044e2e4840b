"""GPU wavefront tracer and a-trous denoiser parity against the CPU oracle.

Both sides draw the same counter-based streams (PCG-RXS-M-XS and Owen-scrambled
Sobol), so the images differ only by float reassociation, libm ulps and the
rare path that flips at a silhouette. Gates (from the reference's own
agreement tests, crates/gpu/tests/cpu_gpu_agreement.rs:83-87,
wavefront_agreement.rs, denoise_agreement.rs, tightened where the build states
a stricter bar): per-pixel mean |delta| <= 1e-3, mean_rel < 2e-3, energy ratio
within 1e-3; a-trous mean_rel < 4e-3.
"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import lucida_amd as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


_SCENES = {}


def scene_for(gpu, oracle, name, **kw):
    """Blob with the GPU LBVH installed (compacted), plus the device scene."""
    key = (name, tuple(sorted(kw.items())))
    if key not in _SCENES:
        blob = L.Blob(name, **kw)
        scene = L.Scene(gpu, blob, build_lbvh=True)
        nodes, tris = scene.download_bvh()
        blob.set_bvh(nodes, tris)
        _SCENES[key] = (blob, scene)
    return _SCENES[key]


def gpu_render(gpu, scene, u, samples, config=None):
    film = L.Film(gpu, u.width, u.height)
    st = L.render_wavefront(gpu, scene, u, samples, film, config)
    acc = film.download()
    return acc, st


def means(acc, u):
    s = np.maximum(acc["samples"], 1.0)[:, None]
    rgb = (acc["radiance"] / s).reshape(u.height, u.width, 3)
    alb = (acc["albedo"] / s).reshape(u.height, u.width, 3)
    nrm = (acc["normal"] / s).reshape(u.height, u.width, 3)
    dep = (acc["depth"] / s[:, 0]).reshape(u.height, u.width)
    return rgb, alb, nrm, dep


def check_agreement(oracle, gpu_rgb, cpu_rgb, label, mean_abs=1e-3, mean_rel=2e-3, energy=1e-3):
    d = oracle.compare(gpu_rgb, cpu_rgb)
    say(f"  {label}: mean|d|={d['mean_abs']:.3e} mean_rel={d['mean_rel']:.3e} "
        f"energy {d['mean_a']:.6f}/{d['mean_b']:.6f} max_abs={d['max_abs']:.3e}")
    assert d["mean_abs"] <= mean_abs, d
    assert d["mean_rel"] < mean_rel, d
    assert abs(d["mean_a"] / d["mean_b"] - 1.0) < energy, d
    return d


def test_ggx_table_matches_oracle(gpu, oracle):
    """The device-built energy table (bsdf.rs:1200-1227) equals the oracle's to
    float rounding; the oracle's is pinned to the reference's shipped table."""
    say("case: GGX energy table (device kernel vs oracle)")
    dev = L.energy_table(gpu)
    np.testing.assert_allclose(dev, oracle.ggx_table(), rtol=0, atol=2e-6)


@pytest.mark.parametrize("sampler", [L.SAMPLER_INDEPENDENT, L.SAMPLER_SOBOL])
def test_config1_box_matches_oracle(gpu, oracle, sampler):
    """Benchmark config 1 at its full size: 256^2, 64 spp, 8 bounces, MIS."""
    blob, scene = scene_for(gpu, oracle, "box-c1")
    u = blob.uniforms(256, 256, samples=64, max_depth=8, sampler=sampler)
    say(f"case: box-c1 256x256 64spp sampler={sampler}")
    acc, st = gpu_render(gpu, scene, u, 64)
    assert st.stack_overflows == 0
    rgb, alb, nrm, dep = means(acc, u)
    cpu, guides, counts = oracle.render(blob.desc(), u)
    check_agreement(oracle, rgb, cpu, "box-c1")
    # guides agree too (first-hit albedo, normal, depth)
    np.testing.assert_allclose(alb, guides[..., 0:3], atol=2e-3)
    np.testing.assert_allclose(dep, guides[..., 6], rtol=1e-3, atol=1e-3)


@pytest.mark.parametrize("mode", [L.SAMPLING_BSDF, L.SAMPLING_NEE, L.SAMPLING_MIS])
def test_sampling_modes_match_oracle(gpu, oracle, mode):
    blob, scene = scene_for(gpu, oracle, "box-c1")
    u = blob.uniforms(96, 96, samples=32, max_depth=8, sampling=mode, sampler=L.SAMPLER_SOBOL)
    say(f"case: box-c1 96x96 32spp mode={mode}")
    acc, _ = gpu_render(gpu, scene, u, 32)
    cpu, _, _ = oracle.render(blob.desc(), u, guides=False)
    check_agreement(oracle, means(acc, u)[0], cpu, f"mode {mode}")


@pytest.mark.parametrize("depth", [1, 2, 4, 16])
def test_depths_match_oracle(gpu, oracle, depth):
    blob, scene = scene_for(gpu, oracle, "cornell-box")
    u = blob.uniforms(96, 96, samples=16, max_depth=depth, sampler=L.SAMPLER_INDEPENDENT)
    say(f"case: cornell-box (analytic) 96x96 16spp depth={depth}")
    acc, _ = gpu_render(gpu, scene, u, 16)
    cpu, _, _ = oracle.render(blob.desc(), u, guides=False)
    check_agreement(oracle, means(acc, u)[0], cpu, f"depth {depth}")


def test_mesh_scene_matches_oracle(gpu, oracle):
    blob, scene = scene_for(gpu, oracle, "cornell-mesh")
    u = blob.uniforms(128, 128, samples=16, max_depth=8)
    say("case: cornell-mesh 128x128 16spp")
    acc, st = gpu_render(gpu, scene, u, 16)
    assert st.stack_overflows == 0
    cpu, _, _ = oracle.render(blob.desc(), u, guides=False)
    check_agreement(oracle, means(acc, u)[0], cpu, "cornell-mesh")


def test_config2_geometry_matches_oracle(gpu, oracle):
    """The 1.05M-triangle benchmark scene, rendered small (the bench runs it at 1024^2 x 16)."""
    blob, scene = scene_for(gpu, oracle, "sphere-c2")
    u = blob.uniforms(128, 128, samples=8, max_depth=8)
    say("case: sphere-c2 (1.05M tris) 128x128 8spp")
    cfg = L.WavefrontConfig(0, 0, 2, 0, 1)  # the binary walk: the oracle repeats its node-visit counts
    acc, st = gpu_render(gpu, scene, u, 8, cfg)
    assert st.stack_overflows == 0
    cpu, _, counts = oracle.render(blob.desc(), u, guides=False)
    check_agreement(oracle, means(acc, u)[0], cpu, "sphere-c2")
    say(f"node visits per extension ray: {st.extend_node_visits / st.extension_rays:.2f}")
    # the same rays were traced: counts agree to the handful of paths that
    # flip at silhouettes
    assert abs(st.extension_rays - counts.extension_rays) <= 2e-3 * counts.extension_rays
    assert abs(st.shadow_rays - counts.shadow_rays) <= 2e-3 * counts.shadow_rays
    # the oracle counts closest-hit traversal only (Scene::intersect_counting);
    # the GPU walk visits every node the reference's walk visits and, past a
    # postponed leaf, some that the walk would cull: at least the oracle's
    # counts (less the paths that flip at silhouettes), a few percent above
    say(f"closest-hit visits / oracle {st.extend_node_visits / counts.node_visits:.4f}, "
        f"tests / oracle {st.extend_triangle_tests / counts.triangle_tests:.4f}")
    assert (1 - 5e-3) * counts.node_visits <= st.extend_node_visits <= 1.05 * counts.node_visits
    assert (1 - 5e-3) * counts.triangle_tests <= st.extend_triangle_tests <= 1.05 * counts.triangle_tests


def test_progressive_equals_one_shot(gpu, oracle):
    blob, scene = scene_for(gpu, oracle, "box-c1")
    u = blob.uniforms(64, 64, samples=16, max_depth=8)
    say("case: progressive (2 x 8 spp) vs one-shot 16 spp")
    one, _ = gpu_render(gpu, scene, u, 16)
    film = L.Film(gpu, 64, 64)
    L.render_wavefront(gpu, scene, u, 8, film)
    u2 = blob.uniforms(64, 64, samples=8, max_depth=8, sample_offset=8)
    L.render_wavefront(gpu, scene, u2, 8, film)
    two = film.download()
    a, b = means(one, u)[0], means(two, u)[0]
    assert np.abs(a - b).max() < 1e-5


def test_batch_size_does_not_change_the_image(gpu, oracle):
    blob, scene = scene_for(gpu, oracle, "box-c1")
    u = blob.uniforms(64, 64, samples=12, max_depth=8)
    say("case: samples_per_batch 1 / 5 / 12")
    imgs = []
    for spb in (1, 5, 12):
        acc, _ = gpu_render(gpu, scene, u, 12, L.WavefrontConfig(spb, 0, 0, 0, 0))
        imgs.append(means(acc, u)[0])
    assert np.abs(imgs[0] - imgs[1]).max() < 1e-5
    assert np.abs(imgs[0] - imgs[2]).max() < 1e-5


@pytest.mark.parametrize("name,width", [("sphere-c2", 0), ("sphere-c2", 2), ("instances-c3", 0)])
def test_ray_order_does_not_change_the_image(gpu, oracle, name, width):
    """Sorting the extension queues only regroups rays into waves, so on these
    scenes films and ray counts are bit-identical (node visits and triangle
    tests differ a little: a lane's speculative visits past a postponed leaf
    depend on which rays share its wave; for the same reason a ray meeting two
    triangles at distances equal to within rounding could keep either)."""
    if name.startswith("instances"):
        blob = L.Blob(name)
        scene = L.Scene(gpu, blob)
    else:
        blob, scene = scene_for(gpu, oracle, name)
    u = blob.uniforms(96, 80, samples=6, max_depth=8)
    say(f"case: {name} 96x80 6spp bvh_width={width}, queue order kept vs sorted")
    runs = []
    for order in (L.WavefrontConfig.RAY_ORDER_NEVER, L.WavefrontConfig.RAY_ORDER_ALWAYS):
        cfg = L.WavefrontConfig(0, 0, width, 0, 1, 1, order)
        runs.append(gpu_render(gpu, scene, u, 6, cfg))
    (a, sa), (b, sb) = runs
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    for f in ("extension_rays", "shadow_rays", "stack_overflows"):
        assert getattr(sa, f) == getattr(sb, f), f
    for f in ("node_visits", "triangle_tests"):
        assert abs(getattr(sa, f) - getattr(sb, f)) <= 0.05 * getattr(sb, f), f
    assert sa.queue_sorts == 0 and sb.queue_sorts == 8  # one batch of 6 spp: the queues are sorted after every SHADE
    eff = lambda st: (st.extend_node_visits / st.extend_lane_slots,
                      (st.node_visits - st.extend_node_visits) / st.connect_lane_slots)
    say("SIMD efficiency (closest hit, occlusion): kept %.3f %.3f, sorted %.3f %.3f" % (*eff(sa), *eff(sb)))


def test_bvh_equals_brute_force(gpu, oracle):
    """GPU BVH traversal vs GPU brute force (num_bvh_nodes = 0), cpu_gpu_agreement.rs:267-302."""
    blob = L.Blob("icosphere-2")
    a = blob.arrays()
    brute = L.Scene(gpu, blob, build_lbvh=False)
    scene = L.Scene(gpu, blob, build_lbvh=True)
    u = blob.uniforms(64, 64, samples=4, max_depth=4)
    u_brute = blob.uniforms(64, 64, samples=4, max_depth=4)
    u_brute.num_bvh_nodes = 0
    say("case: icosphere-2 BVH vs brute force")
    acc_b, _ = gpu_render(gpu, brute, u_brute, 4)
    acc_t, _ = gpu_render(gpu, scene, u, 4)
    # background is black and there is no light: compare first-hit guides
    gb, gt = means(acc_b, u), means(acc_t, u)
    np.testing.assert_allclose(gt[3], gb[3], atol=1e-6)
    np.testing.assert_allclose(gt[1], gb[1], atol=1e-6)
    assert len(a["triangles"]) == 320


def test_atrous_matches_oracle(gpu, oracle):
    blob, scene = scene_for(gpu, oracle, "box-c1")
    u = blob.uniforms(128, 128, samples=4, max_depth=8)
    say("case: a-trous on a 128x128 4spp render")
    film = L.Film(gpu, 128, 128)
    L.render_wavefront(gpu, scene, u, 4, film)
    before = film.download()
    out = L.Film(gpu, 128, 128)
    p = L.DenoiseParams.default()
    L.Denoiser(gpu).run(film, out, p)
    den = out.download()
    # the accumulation is untouched (denoise_agreement.rs:70-92)
    assert before.tobytes() == film.download().tobytes()
    rgb, alb, nrm, dep = means(before, u)
    alb_f = np.maximum(alb, 1e-3)
    cpu = oracle.denoise(rgb, alb_f, nrm, dep, p)
    gpu_den = den["radiance"].reshape(128, 128, 3)
    d = oracle.compare(gpu_den, cpu)
    say(f"  a-trous mean_rel={d['mean_rel']:.3e} max_abs={d['max_abs']:.3e}")
    assert d["mean_rel"] < 4e-3
    assert np.all(den["samples"] == 1.0)


def test_atrous_zero_iterations_is_identity(gpu, oracle):
    blob, scene = scene_for(gpu, oracle, "box-c1")
    u = blob.uniforms(32, 32, samples=2, max_depth=4)
    film = L.Film(gpu, 32, 32)
    L.render_wavefront(gpu, scene, u, 2, film)
    out = L.Film(gpu, 32, 32)
    p = L.DenoiseParams.default()
    p.iterations = 0
    L.Denoiser(gpu).run(film, out, p)
    rgb = means(film.download(), u)[0].reshape(-1, 3)
    np.testing.assert_allclose(out.download()["radiance"], rgb, rtol=2e-6, atol=1e-7)


def test_render_rejects_mismatched_film(gpu, oracle):
    blob, scene = scene_for(gpu, oracle, "box-c1")
    u = blob.uniforms(32, 32, samples=1)
    film = L.Film(gpu, 16, 16)
    with pytest.raises(L.GpuError) as e:
        L.render_wavefront(gpu, scene, u, 1, film)
    assert e.value.kind == "InvalidArg"


def test_only_the_binary_walk_is_accepted(gpu):
    """bvh_width is 0 or 2 (the binary walk) and lucida_scene_cast takes only
    LUCIDA_CAST_ANY_HIT: anything else is an InvalidArg, not a silent fallback."""
    import ctypes as C
    say("case: bvh_width 4 and an unknown cast flag are rejected")
    blob = L.Blob("cornell-box")
    scene = L.Scene(gpu, blob)
    film = L.Film(gpu, 16, 16)
    u = blob.uniforms(16, 16, samples=1, max_depth=2)
    for render in (L.render_wavefront, L.render_megakernel):
        with pytest.raises(L.GpuError) as e:
            render(gpu, scene, u, 1, film, L.WavefrontConfig(0, 0, 4, 0, 0))
        assert e.value.kind == "InvalidArg"
    rays = np.zeros((1, 8), np.float32)
    rays[0, 6], rays[0, 7] = 1.0, 1e30
    hits = np.zeros((1, 4), np.float32)
    status = L.lib().lucida_scene_cast(gpu._h, scene._h, rays.ctypes.data, 1, 2, hits.ctypes.data, None)
    assert L.GpuError.KINDS.get(status) == "InvalidArg"
    L.render_wavefront(gpu, scene, u, 1, film, L.WavefrontConfig(0, 0, 2, 0, 0))  # the binary walk still renders
    gpu.check_errors()
