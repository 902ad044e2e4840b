"""Training data and training loop for the learned denoiser.

The tracer renders procedurally varied scenes twice: a low-sample input film
(1 to 256 samples per pixel, with its albedo / normal / depth guides) and an
independent high-sample target film of the same view (another frame seed, so
no sample is shared). The network is trained on random crops of those pairs
by the library's own forward, backward and Adam kernels; everything stays
resident on the device.

    python -m lucida_amd.denoiser_training --scenes 32 --steps 3000 --out weights.f32

With --eval the trained network is scored on held-out scenes (held_out_curves)
and a JSON summary of the run is written to --json.
"""
from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

import lucida_amd as L


def _rotate_y(v, angle: float):
    c, s = math.cos(angle), math.sin(angle)
    return (c * v[0] + s * v[2], v[1], -s * v[0] + c * v[2])


def orbit_camera(u: L.Uniforms, angle: float, lift: float = 0.0) -> L.Uniforms:
    """The scene's camera turned by `angle` about the vertical axis through the
    origin and raised by `lift` (camera frame vectors rotate with it)."""
    for name in ("cam_origin", "cam_upper_left"):
        v = _rotate_y(tuple(getattr(u, name)), angle)
        getattr(u, name)[:] = (v[0], v[1] + lift, v[2])
    for name in ("cam_horizontal", "cam_vertical"):
        getattr(u, name)[:] = _rotate_y(tuple(getattr(u, name)), angle)
    return u


def scene_variant(k: int, rng: np.random.Generator):
    """Scene k of the training set (seeds 1000 + k, so the benchmark scene,
    seed 0, is never in it): mostly displaced spheres in the room with random
    noise amplitudes, tessellations and orbiting cameras (the config-2
    generator), some instanced fields of rough dielectric and GGX meshes under
    a log-normal sky with hot spots (the config-3 generator), some Cornell
    boxes."""
    kind = rng.random()
    if kind < 0.7:
        blob = L.Blob("sphere-c2", seed=1000 + k, frequency=int(rng.integers(24, 96)),
                      displacement=float(rng.uniform(0.01, 0.12)))
        return blob, float(rng.uniform(-0.9, 0.9)), float(rng.uniform(-0.2, 0.4))
    if kind < 0.85:
        return L.Blob("instances-small", seed=1000 + k, frequency=int(rng.integers(3, 8))), 0.0, 0.0
    return L.Blob("box-c1" if kind < 0.93 else "cornell-mesh"), 0.0, 0.0


# Input sample counts of the training pairs (one per scene, drawn uniformly):
# the range the denoiser is used over, 1 to 256 spp.
INPUT_SPP = (1, 2, 4, 8, 16, 32, 64, 128, 256)


def make_training_set(gpu: L.Gpu, scenes: int = 24, width: int = 256, height: int = 256,
                      input_spp=INPUT_SPP, ref_spp: int = 4096, depth: int = 8, seed: int = 7,
                      log=None):
    """Renders the (input, target) film pairs; returns them and the paths traced."""
    rng = np.random.default_rng(seed)
    pairs = []
    paths = 0
    for k in range(scenes):
        blob, angle, lift = scene_variant(k, rng)
        scene = L.Scene(gpu, blob, build_lbvh=True)
        spp = int(rng.choice(input_spp))
        mk = lambda s, fs: orbit_camera(blob.uniforms(width, height, samples=s, max_depth=depth,
                                                       sampling=L.SAMPLING_MIS, sampler=L.SAMPLER_SOBOL,
                                                       frame_seed=fs), angle, lift)
        film_in = L.Film(gpu, width, height)
        film_ref = L.Film(gpu, width, height)
        L.render_wavefront(gpu, scene, mk(spp, 0x5EED0000 + k), spp, film_in)
        L.render_wavefront(gpu, scene, mk(ref_spp, 0x7A7A0000 + k), ref_spp, film_ref)
        paths += width * height * (spp + ref_spp)
        pairs.append((film_in, film_ref))
        if log:
            log(f"scene {k}: {blob.desc().num_triangles} tris, input {spp} spp")
    gpu.check_errors()
    return pairs, paths


def train(gpu: L.Gpu, net: L.LearnedDenoiser, pairs, steps: int = 1500, batch: int = 8, crop_w: int = 128,
          crop_h: int = 128, lr: float = 1e-3, seed: int = 11, log_every: int = 0, log=None, ema: float = 0.0,
          ema_every: int = 10):
    """Adam on random crops of random pairs; returns the per-step losses.
    With ema > 0 the network ends with a moving average of its weights over
    the second half of training (decay `ema` per step, sampled every
    `ema_every` steps) instead of the last step's weights."""
    rng = np.random.default_rng(seed)
    losses = []
    avg = None
    beta = ema ** ema_every if ema > 0 else 0.0
    for step in range(steps):
        idx = rng.integers(0, len(pairs), size=batch)
        ins = [pairs[i][0] for i in idx]
        refs = [pairs[i][1] for i in idx]
        crops = [(int(rng.integers(0, ins[j].width - crop_w + 1)), int(rng.integers(0, ins[j].height - crop_h + 1)))
                 for j in range(batch)]
        # cosine decay to a tenth of the initial rate
        rate = lr * (0.1 + 0.9 * 0.5 * (1.0 + math.cos(math.pi * step / max(steps, 1))))
        losses.append(net.train_step(ins, refs, crop_w, crop_h, crops=crops, lr=rate))
        if ema > 0 and step >= steps // 2 and (step + 1) % ema_every == 0:
            p = net.params().astype(np.float64)
            avg = p if avg is None else beta * avg + (1.0 - beta) * p
        if log and log_every and (step + 1) % log_every == 0:
            log(f"step {step + 1}: loss {np.mean(losses[-log_every:]):.5f}")
    if avg is not None:
        net.set_params(avg.astype(np.float32))
    gpu.check_errors()
    return losses


def relmse(x: np.ndarray, r: np.ndarray) -> float:
    """Mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2)."""
    x = x.astype(np.float64)
    r = r.astype(np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


# Scenes no training pair comes from (training seeds are 1000 + k): the
# benchmark scene (seed 0) at its full size, and the two scenes the gain test
# scores.
HELD_OUT = (("benchmark sphere-c2 seed 0, 1024^2", dict(name="sphere-c2", frequency=229), 1024),
            ("sphere-c2 seed 77, 256^2", dict(name="sphere-c2", seed=77, frequency=40, displacement=0.1), 256),
            ("instances-small seed 500, 256^2", dict(name="instances-small", seed=500, frequency=5), 256))


def held_out_curves(gpu: L.Gpu, net: L.LearnedDenoiser, max_spp: int = 256, ref_spp: int = 4096,
                    depth: int = 8, log=None) -> list:
    """relMSE of the plain estimate, the a-trous filter and the network after
    1, 2, 4, ..., max_spp progressive samples, against an independent
    ref_spp render (another frame seed), per held-out scene."""
    out = []
    dn = L.Denoiser(gpu)
    for label, kw, res in HELD_OUT:
        kw = dict(kw)
        blob = L.Blob(kw.pop("name"), **kw)
        scene = L.Scene(gpu, blob, build_lbvh=True)
        mk = lambda spp, off, seed: blob.uniforms(res, res, samples=spp, max_depth=depth, sampling=L.SAMPLING_MIS,
                                                  sampler=L.SAMPLER_SOBOL, sample_offset=off, frame_seed=seed)
        ref_film = L.Film(gpu, res, res)
        L.render_wavefront(gpu, scene, mk(ref_spp, 0, 0x0BADF00D), ref_spp, ref_film)
        ref = L.film_mean(ref_film.download(), res, res)
        del ref_film
        film, den = L.Film(gpu, res, res), L.Film(gpu, res, res)
        rows, done, spp = [], 0, 1
        while spp <= max_spp:
            L.render_wavefront(gpu, scene, mk(spp - done, done, 0x5EED1234), spp - done, film)
            done = spp
            plain = relmse(L.film_mean(film.download(), res, res), ref)
            dn.run(film, den)
            atrous = relmse(L.film_mean(den.download(), res, res), ref)
            net.denoise(film, den)
            learned = relmse(L.film_mean(den.download(), res, res), ref)
            rows.append({"spp": spp, "plain": plain, "atrous": atrous, "learned": learned,
                         "learned_over_atrous": round(learned / atrous, 4)})
            spp *= 2
        gpu.check_errors()
        out.append({"scene": label, "reference_spp": ref_spp, "curve": rows})
        if log:
            log(f"held out {label}: learned / a-trous " +
                " ".join(f"{r['spp']}:{r['learned_over_atrous']:.3f}" for r in rows))
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--scenes", type=int, default=32)
    ap.add_argument("--ref-spp", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=128, help="training crop side")
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--ema", type=float, default=0.0,
                    help="end with a moving average of the weights over the second half (decay per step, e.g. 0.999)")
    ap.add_argument("--out", type=str, default="", help="write the trained parameters (f32) here")
    ap.add_argument("--eval", action="store_true", help="score the network on the held-out scenes")
    ap.add_argument("--json", type=str, default="", help="write a JSON summary of the run here")
    a = ap.parse_args()
    say = lambda m: print(m, flush=True)
    gpu = L.Gpu(0)
    t0 = time.perf_counter()
    pairs, paths = make_training_set(gpu, a.scenes, ref_spp=a.ref_spp, log=say)
    t1 = time.perf_counter()
    say(f"data: {len(pairs)} pairs, {paths / 1e9:.2f} Gpaths in {t1 - t0:.1f} s")
    net = L.LearnedDenoiser(gpu, seed=1)
    losses = train(gpu, net, pairs, a.steps, a.batch, crop_w=a.crop, crop_h=a.crop, lr=a.lr, log_every=500,
                   log=say, ema=a.ema)
    t2 = time.perf_counter()
    del pairs
    say(f"trained {a.steps} steps in {t2 - t1:.1f} s, final loss {np.mean(losses[-100:]):.5f}")
    if a.out:
        net.params().tofile(a.out)
    summary = {"scenes": a.scenes, "input_spp": list(INPUT_SPP), "target_spp": a.ref_spp, "crop": a.crop, "ema": a.ema,
               "batch": a.batch, "lr": a.lr, "steps": a.steps, "params": L.LearnedDenoiser.param_count(),
               "input_channels": 16, "data_paths": paths, "data_s": round(t1 - t0, 1), "train_s": round(t2 - t1, 1),
               "loss_first_500": float(np.mean(losses[:500])), "loss_last_500": float(np.mean(losses[-500:])),
               "loss_every_500": [round(float(np.mean(losses[k:k + 500])), 6) for k in range(0, len(losses), 500)]}
    if a.eval:
        summary["held_out"] = held_out_curves(gpu, net, log=say)
    if a.json:
        with open(a.json, "w") as f:
            json.dump(summary, f, indent=1)
    gpu.close()
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
