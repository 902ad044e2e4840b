"""The benchmark's batched request stream (64 jobs, PCG64 seed 6: resolutions
128^2-1024^2, 1-64 spp, the Cornell box and four displaced-sphere rooms) run
through lucida_render_batch exactly as the benchmark runs it: with the
autotuner's launch shape, once replayed from a hipGraph and once on the
context's two pipelines. Every job's film must equal its own
lucida_render_wavefront bit for bit (jobs of 1024^2 x 16 or 64 spp span
several rounds, alternating pipelines); one job of a mixed round is also
rendered by the oracle and compared at the renderer's gates."""
from __future__ import annotations

import os
import sys

import numpy as np
import pytest

import lucida_amd as L
from lucida_amd import autotune
from lucida_amd import request_stream as RS

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def stream(gpu):
    blobs = RS.stream_blobs()
    scenes = [L.Scene(gpu, b) for b in blobs]
    reqs = RS.stream_requests()
    solo = []
    say(f"stream: {len(reqs)} jobs, {sum(r * r * s for _, r, s, _ in reqs)} paths; solo renders first")
    for scene, u, spp, film in RS.make_jobs(gpu, blobs, scenes, reqs):
        L.render_wavefront(gpu, scene, u, spp, film)
        solo.append(film.download())
        del film
    return blobs, scenes, reqs, solo


@pytest.mark.parametrize("use_graph", [1, 0])
def test_stream_batch_equals_solo_renders(gpu, stream, tmp_path, use_graph):
    blobs, scenes, reqs, solo = stream
    os.environ["LUCIDA_AUTOTUNE_CACHE"] = str(tmp_path / "autotune.json")
    jobs = RS.make_jobs(gpu, blobs, scenes, reqs)
    cfg = autotune.tune(gpu, jobs, repeats=1)
    cfg.use_graph = use_graph
    say(f"stream case: autotuned max_paths {cfg.max_paths}, use_graph {use_graph}")
    st, done = L.render_batch(gpu, jobs, cfg)
    assert st.stack_overflows == 0 and np.all(done > 0)
    multi_round = 0
    for k, ((scene, u, spp, film), ref) in enumerate(zip(jobs, solo)):
        got = film.download()
        np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=f"job {k} {reqs[k]}")
        multi_round += u.width * u.height * spp > (1 << 23)
    assert multi_round >= 4  # jobs spanning several rounds are covered
    say(f"  rounds {st.kernel_launches[0]}, {st.elapsed_ms:.1f} ms, {multi_round} jobs over several rounds")


def test_stream_job_matches_the_oracle(gpu, oracle, stream):
    """The cheapest 128^2 job of the stream, rendered inside the batch, against
    the oracle on the device's own BVH."""
    blobs, scenes, reqs, solo = stream
    k = min((i for i, r in enumerate(reqs) if r[1] == 128), key=lambda i: (reqs[i][2], reqs[i][0] == 0))
    si, res, spp, fs = reqs[k]
    say(f"stream case: job {k} (scene {si}, {res}^2, {spp} spp) against the oracle")
    u = blobs[si].uniforms(res, res, samples=spp, max_depth=RS.MAX_DEPTH, frame_seed=fs)
    d, keep = scenes[si].host_desc(blobs[si])
    ref, _, _ = oracle.render(d, u, background=list(u.background), guides=False)
    img = L.film_mean(solo[k], res, res)
    c = oracle.compare(img, ref)
    say(f"  mean_rel {c['mean_rel']:.2e} mean|d| {c['mean_abs']:.2e} energy {c['mean_a']:.5f}/{c['mean_b']:.5f}")
    assert c["mean_abs"] <= 1e-3 and c["mean_rel"] < 2e-3 and abs(c["mean_a"] / c["mean_b"] - 1) < 1e-3
