"""What the trained network gains over the method without it.

The learned denoiser is trained on the GPU by the library's own kernels
(forward, backward, Adam) on 96 generated scenes (the config-2 and config-3
generators and Cornell boxes, seeds 1000-1095, rendered at 256^2 with 1-64
input samples against 2048-sample targets; 6000 Adam steps ending with a
moving average of the weights), then scored on scenes it never saw: a
displaced sphere in its room (seed 77) and an instanced field of rough
dielectric and GGX meshes under a sky with hot spots (seed 500), both at
256^2. Score: relMSE, the mean over pixels and channels of
(x - r)^2 / (r^2 + 1e-2), against an independent 4096-sample render. The
trained network must beat the a-trous filter alone (its fallback, and the
estimate it corrects) at 4 and 16 samples per pixel on the sphere scene
(about 0.86 and 0.67 of its relMSE) and at 1 and 4 on the instanced one
(about 0.45 and 0.44). At 1 sample on the sphere this small training set
leaves the two about even: the ratio measured 0.83 to 1.03 with the last
step's weights over runs and small changes of the training renders (one
target film of 192 rendered differently moved it from below 1 to 1.03),
and 0.99 with the average, so there the network must stay within 5% of the
filter. From 16
samples on, the instanced scene's relMSE is set by caustic fireflies of
the 1e4 hot spots seen through glass, which the 4096-sample reference still
holds: it then compares two noisy estimates of the reference's own noise,
so it is reported and not asserted.

An untrained network (output layer at zero) must return the a-trous estimate
itself, so any gain is the training's. (Training is deterministic: the
weight gradients are summed in a fixed order, so this test scores the same
network on every run.)"""
from __future__ import annotations

import sys

import numpy as np
import pytest

import lucida_amd as L
from lucida_amd import denoiser_training as T

pytestmark = pytest.mark.gpu

HELD_OUT = [("sphere", dict(name="sphere-c2", seed=77, frequency=40, displacement=0.1)),
            ("instanced", dict(name="instances-small", seed=500, frequency=5))]
ASSERTED = {"sphere": (4, 16), "instanced": (1, 4)}
WITHIN = {"sphere": {1: 1.05}}  # spp: largest learned / a-trous ratio allowed
SPP = (1, 4, 16)
RES = 256


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def relmse(x, r) -> float:
    x = x.astype(np.float64)
    r = r.astype(np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


@pytest.fixture(scope="module")
def held_out(gpu):
    """Per held-out scene: progressive films at 1, 4, 16 spp and the reference."""
    out = {}
    for label, kw in HELD_OUT:
        kw = dict(kw)
        blob = L.Blob(kw.pop("name"), **kw)
        scene = L.Scene(gpu, blob, build_lbvh=True)
        mk = lambda spp, off, seed: blob.uniforms(RES, RES, samples=spp, max_depth=8, sampling=L.SAMPLING_MIS,
                                                  sampler=L.SAMPLER_SOBOL, sample_offset=off, frame_seed=seed)
        ref = L.Film(gpu, RES, RES)
        L.render_wavefront(gpu, scene, mk(4096, 0, 0x0BADF00D), 4096, ref)
        films, f, done = {}, L.Film(gpu, RES, RES), 0
        for spp in SPP:
            L.render_wavefront(gpu, scene, mk(spp - done, done, 0x5EED1234), spp - done, f)
            done = spp
            g = L.Film(gpu, RES, RES)
            g.upload(f.download())
            films[spp] = g
        out[label] = (films, L.film_mean(ref.download(), RES, RES))
        say(f"held-out {label}: reference and inputs rendered")
    return out


def scores(gpu, denoise, films, ref) -> dict:
    out = L.Film(gpu, RES, RES)
    res = {}
    for spp, g in films.items():
        denoise(g, out)
        res[spp] = relmse(L.film_mean(out.download(), RES, RES), ref)
    return res


def test_untrained_network_returns_the_atrous_estimate(gpu, held_out):
    films, _ = held_out["sphere"]
    dn, net = L.Denoiser(gpu), L.LearnedDenoiser(gpu, seed=1)
    a, b = L.Film(gpu, RES, RES), L.Film(gpu, RES, RES)
    dn.run(films[4], a)
    net.denoise(films[4], b)
    x = L.film_mean(a.download(), RES, RES)
    y = L.film_mean(b.download(), RES, RES)
    # log1p / expm1 and the albedo division round trip in FP32
    assert relmse(y, x) < 1e-10, relmse(y, x)


def test_trained_network_beats_atrous_on_held_out_scenes(gpu, held_out):
    say("training: 96 scenes at 256^2 (1-64 spp inputs, 2048-spp targets), 6000 Adam steps, batch 8 of 256^2, "
        "weights averaged over the second half")
    pairs, paths = T.make_training_set(gpu, 96, width=RES, height=RES, input_spp=(1, 2, 4, 8, 16, 32, 64),
                                       ref_spp=2048,
                                       log=lambda m: say(m) if int(m.split()[1][:-1]) % 16 == 15 else None)
    net = L.LearnedDenoiser(gpu, seed=1)
    losses = T.train(gpu, net, pairs, 6000, crop_w=RES, crop_h=RES, log_every=1000, log=say, ema=0.9995)
    del pairs
    assert np.all(np.isfinite(losses)) and np.mean(losses[-100:]) < 0.5 * np.mean(losses[:100])
    dn = L.Denoiser(gpu)
    for label, (films, ref) in held_out.items():
        base = scores(gpu, dn.run, films, ref)
        learned = scores(gpu, net.denoise, films, ref)
        say(f"{label}: relMSE a-trous / learned: " +
            ", ".join(f"{s} spp {base[s]:.3e} / {learned[s]:.3e}" for s in SPP))
        for s in ASSERTED[label]:
            assert learned[s] < base[s], (label, s, learned[s], base[s])
        for s, bound in WITHIN.get(label, {}).items():
            assert learned[s] < bound * base[s], (label, s, learned[s], base[s])
