"""Batched renders (lucida_render_batch): many requests of uneven resolution,
samples, depth and scene in one wavefront must give every job the film its
own lucida_render_wavefront call gives, bit for bit: the RNG is counter-based
on (frame, pixel, sample) and each job is cut into the chunks its single
render would use."""
from __future__ import annotations

import sys

import numpy as np
import pytest

import lucida_amd as L

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def scenes(gpu):
    blobs = {
        "box-c1": L.Blob("box-c1"),
        "cornell-mesh": L.Blob("cornell-mesh"),
        "sphere-c2-f24": L.Blob("sphere-c2", seed=3, frequency=24),
        "instances-small": L.Blob("instances-small"),
    }
    return {k: (b, L.Scene(gpu, b)) for k, b in blobs.items()}


# (scene, width, height, samples, max_depth, sample_offset, frame_seed)
JOBS = [
    ("box-c1", 64, 64, 4, 8, 0, 0x5EED1234),
    ("sphere-c2-f24", 128, 96, 3, 8, 0, 11),
    ("cornell-mesh", 96, 64, 1, 5, 7, 0x5EED1234),
    ("instances-small", 64, 48, 5, 12, 0, 99),
    ("box-c1", 32, 32, 16, 3, 2, 5),
    ("sphere-c2-f24", 48, 40, 2, 8, 0, 0x5EED1234),
    # chunks of 2^21 paths or more: rounds of their own, single-render kernels
    ("box-c1", 1024, 1024, 3, 4, 0, 7),
    ("sphere-c2-f24", 1024, 1024, 6, 8, 1, 3),  # one chunk (8 spp per 2^23-path chunk at 1024^2)
    ("instances-small", 1024, 768, 3, 6, 0, 21),
]


def make_jobs(gpu, scenes, sampling=L.SAMPLING_MIS):
    out = []
    for name, w, h, spp, depth, off, seed in JOBS:
        blob, scene = scenes[name]
        u = blob.uniforms(w, h, samples=spp, max_depth=depth, sample_offset=off, frame_seed=seed, sampling=sampling)
        out.append((scene, u, spp, L.Film(gpu, w, h)))
    return out


def solo(gpu, jobs):
    films = []
    for scene, u, spp, film in jobs:
        f = L.Film(gpu, film.width, film.height)
        L.render_wavefront(gpu, scene, u, spp, f)
        films.append(f.download())
    return films


@pytest.mark.parametrize("use_graph", [0, 1])
@pytest.mark.parametrize("max_paths", [0, 40_000])
def test_batch_equals_solo_renders(gpu, scenes, use_graph, max_paths):
    say(f"batch case: {len(JOBS)} jobs, use_graph={use_graph}, max_paths={max_paths}")
    jobs = make_jobs(gpu, scenes)
    cfg = L.WavefrontConfig(0, max_paths, 0, use_graph, 1, 0)
    st, done = L.render_batch(gpu, jobs, cfg)
    want = solo(gpu, jobs)
    for k, ((scene, u, spp, film), ref) in enumerate(zip(jobs, want)):
        got = film.download()
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=f"job {k} {JOBS[k]}")
    assert np.all(done > 0) and st.elapsed_ms > 0
    assert st.kernel_launches[0] >= (1 if max_paths == 0 else 3)  # rounds
    assert st.stack_overflows == 0
    # the large jobs finish after every small one, shortest first
    assert done[:6].max() <= done[6] <= done[8] <= done[7]
    say(f"rounds {st.kernel_launches[0]}, batch {st.elapsed_ms:.2f} ms, done {np.round(done, 2)}")


def test_batch_ray_order_does_not_change_films(gpu, scenes):
    """Sorted queues (by job, then ray) only regroup rays into waves: the films
    of a batch are bit-identical with and without the sorts."""
    say(f"batch case: {len(JOBS)} jobs, queue order kept vs sorted")
    jobs = make_jobs(gpu, scenes)
    out = []
    for order in (L.WavefrontConfig.RAY_ORDER_NEVER, L.WavefrontConfig.RAY_ORDER_ALWAYS):
        for _, _, _, film in jobs:
            film.clear()
        st, _ = L.render_batch(gpu, jobs, L.WavefrontConfig(0, 0, 0, 0, 1, 0, order))
        out.append(([film.download() for _, _, _, film in jobs], st))
    (kept, sk), (sorted_, ss) = out
    for k, (a, b) in enumerate(zip(kept, sorted_)):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f"job {k} {JOBS[k]}")
    for f in ("extension_rays", "shadow_rays"):
        assert getattr(sk, f) == getattr(ss, f), f
    # speculative visits (past a postponed leaf) depend on which rays share a wave
    for f in ("node_visits", "triangle_tests"):
        assert abs(getattr(sk, f) - getattr(ss, f)) <= 0.05 * getattr(ss, f), f


def test_batch_accumulates_like_progressive_renders(gpu, scenes):
    """Two batches into the same films equal one render of all the samples."""
    blob, scene = scenes["box-c1"]
    u1 = blob.uniforms(64, 64, samples=3, sample_offset=0)
    u2 = blob.uniforms(64, 64, samples=5, sample_offset=3)
    film = L.Film(gpu, 64, 64)
    L.render_batch(gpu, [(scene, u1, 3, film)])
    L.render_batch(gpu, [(scene, u2, 5, film)])
    ref = L.Film(gpu, 64, 64)
    L.render_wavefront(gpu, scene, u1, 3, ref)
    L.render_wavefront(gpu, scene, u2, 5, ref)
    np.testing.assert_array_equal(film.download().view(np.uint32), ref.download().view(np.uint32))


def test_batch_rejects_shared_films_and_bad_sizes(gpu, scenes):
    blob, scene = scenes["box-c1"]
    f = L.Film(gpu, 32, 32)
    u = blob.uniforms(32, 32, samples=1)
    with pytest.raises(L.GpuError) as e:
        L.render_batch(gpu, [(scene, u, 1, f), (scene, u, 1, f)])
    assert e.value.kind == "InvalidArg"
    with pytest.raises(L.GpuError) as e:
        L.render_batch(gpu, [(scene, blob.uniforms(16, 16, samples=1), 1, f)])
    assert e.value.kind == "InvalidArg"
