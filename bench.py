"""Benchmark: the wavefront path tracer on the 1.05M-triangle scene.

Workload (one "step"): render the noise-displaced geodesic sphere (1,048,820
triangles) in its 1,024-triangle room lit by 16 emissive quads, at 1024x1024,
16 samples per pixel, 8 bounces, NEE + MIS, Owen-scrambled Sobol, into a
cleared film. The scene, its GPU-built tree (the LBVH restructured by
treelets; --tree lbvh keeps the LBVH as built) and the film are resident in
HBM before timing starts. value = paths per second over the timed steps; the
same step over the LBVH as built is timed beside it (extra.lbvh_tree_render).

Prints one JSON line. Extra figures (Mrays/s, LBVH Mtris/s, a-trous time,
per-kernel times) ride in "extra". "time_to_relmse" holds the metric's second
half: device time until the plain and the denoised images of the same scene
reach relMSE 0.01 against an independent 4096-spp render.

    python bench.py [--gpus 1] [--steps 7] [--warmup 2] [--cpu-seconds 15] [--no-cpu] [--no-relmse]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mpaths/s @1M tris 8 bounces; time-to-relMSE 0.01: denoised vs plain"
HBM_PEAK_GBS = 8000.0
TRAFFIC_SUMMARY = ROOT / "profiles" / "wavefront_c2_1024x16spp_hbm_bytes.md"
SQ_SUMMARY = ROOT / "profiles" / "wavefront_c2_1024x16spp_sq_counters.md"
# MI355X: 256 CUs x 4 SIMDs, 2.4 GHz peak engine clock; a SIMD issues one
# wave64 VALU instruction per 4 clocks.
SIMDS, CLOCK_HZ, CLOCKS_PER_VALU = 1024, 2.4e9, 4
# Algorithmic bytes (the reference's struct sizes): a closest-hit ray reads
# its queue index (4) and origin + direction (32) and writes a hit (16); a
# node visit reads the two 32-byte children of a binary node; a triangle
# test reads its three 16-byte edge vectors. A shadow ray reads its 48-byte
# record and adds 24 bytes of radiance. The LBVH costs 436 bytes per
# triangle, the a-trous filter 644 per pixel.
RAY_BYTES, NODE_BYTES, TRI_BYTES, SHADOW_BYTES = 52, 64, 48, 72
LBVH_BYTES_PER_TRI, ATROUS_BYTES_PER_PX = 436, 644


def profile_figure(path: Path, key: str):
    """A `key: value` line of a committed rocprofv3 summary, or None."""
    try:
        for line in path.read_text().splitlines():
            if line.startswith(key + ":"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def lbvh_build_rate(L, gpu, blob, runs: int = 7, treelet: bool = False) -> dict:
    """Device build of the scene's tree (inputs already on the GPU): the LBVH,
    or the LBVH restructured by treelets."""
    a = blob.arrays()
    tris = np.ascontiguousarray(a["triangles"])
    pos = np.ascontiguousarray(a["positions"])
    n = len(tris)
    lib = L.lib()
    ptrs = {}
    sizes = {"tris": tris.nbytes, "pos": pos.nbytes, "nodes": (2 * n - 1) * 32, "prims": n * 4,
             "sorted": tris.nbytes}
    for k, b in sizes.items():
        p = C.c_void_p()
        L._check(lib.lucida_device_alloc(gpu._h, b, C.byref(p)))
        ptrs[k] = p
    L._check(lib.lucida_copy_to_device(gpu._h, ptrs["tris"], tris.ctypes.data, tris.nbytes))
    L._check(lib.lucida_copy_to_device(gpu._h, ptrs["pos"], pos.ctypes.data, pos.nbytes))
    times, stages = [], []
    count = C.c_size_t()
    for i in range(runs + 1):
        st = np.zeros(5, np.float32)
        gpu.check_errors()
        t0 = time.perf_counter()
        if treelet:
            L._check(lib.lucida_bvh_build_treelet_device(gpu._h, ptrs["tris"], n, ptrs["pos"], L.TREELET_ROUNDS,
                                                         ptrs["nodes"], C.byref(count), ptrs["prims"],
                                                         ptrs["sorted"], st.ctypes.data))
        else:
            L._check(lib.lucida_lbvh_build_device(gpu._h, ptrs["tris"], n, ptrs["pos"], ptrs["nodes"],
                                                  C.byref(count), ptrs["prims"], ptrs["sorted"], st.ctypes.data))
        gpu.check_errors()
        t1 = time.perf_counter()
        if i > 0:
            times.append(t1 - t0)
            stages.append(st.copy())
    for p in ptrs.values():
        lib.lucida_device_free(gpu._h, p)
    med = float(np.median(times))
    st = np.median(np.stack(stages), axis=0)
    device_us = float(st.sum())
    if treelet:
        return {"triangles": n, "nodes": int(count.value), "rounds": L.TREELET_ROUNDS, "wall_ms": round(med * 1e3, 4),
                "device_us": round(device_us, 1), "runs": runs, "statistic": "median",
                "stage_us": dict(zip(("lbvh_to_fitted_tree", "treelet_rounds", "layout"),
                                     [round(float(x), 1) for x in st[:3]])),
                "mtris_per_s": round(n / (device_us * 1e-6) / 1e6, 1), "mtris_per_s_wall": round(n / med / 1e6, 1)}
    gbs = LBVH_BYTES_PER_TRI * n / (device_us * 1e-6) / 1e9
    return {"triangles": n, "nodes": int(count.value), "wall_ms": round(med * 1e3, 4),
            "device_us": round(device_us, 1), "runs": runs, "statistic": "median",
            "stage_us": dict(zip(("prepare", "sort", "karras", "fit", "relayout"), [round(float(x), 1) for x in st])),
            "mtris_per_s": round(n / (device_us * 1e-6) / 1e6, 1), "mtris_per_s_wall": round(n / med / 1e6, 1),
            "roofline": {"bytes_per_tri": LBVH_BYTES_PER_TRI, "achieved_gbs": round(gbs, 1),
                         "frac_of_8tbs": round(gbs / HBM_PEAK_GBS, 4)}}


def cpu_baseline(blob, u, target_s: float, runs: int = 7) -> dict:
    """The CPU oracle (a C++ restatement of the reference tracer) on a sample
    of rows of the same workload, spread evenly over the image, with one
    thread per core this process may use (each row a work item, so every
    thread is busy): one warm-up, then the median of `runs` timed renders of
    the sample (crates/cli/src/bin/bench.rs:192-212), its size chosen so all
    runs take about target_s."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle
    threads = len(os.sched_getaffinity(0))
    threads = min(threads, int(os.environ.get("OMP_NUM_THREADS", threads)))
    desc = blob.desc()
    spread = lambda n: ((np.arange(n) + 0.5) * u.height / n).astype(np.uint32)
    n = threads
    t0 = time.perf_counter()
    pyoracle.render_rows(desc, u, spread(n), threads=threads)  # warm-up and sizing
    dt = time.perf_counter() - t0
    n = int(max(threads, min(u.height, n * target_s / runs / max(dt, 1e-3))))
    n = max(threads, n // threads * threads)  # whole rounds of rows per thread
    rows = spread(n)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        pyoracle.render_rows(desc, u, rows, threads=threads)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    paths = n * u.width * u.samples_per_launch
    return {"value": paths / dt / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"{n} rows spread evenly over the {u.width}x{u.height} image (every {u.height / n:.1f}th), "
                      f"{u.samples_per_launch} spp, {u.max_depth} bounces ({paths} paths), one thread per core: "
                      f"median of {runs} runs after a warm-up, {dt:.2f} s each"}


def other_configs(L, gpu, treelet: bool) -> dict:
    """Throughput of the other two render workloads at their full sizes: the
    Cornell box (256^2, 64 spp, 8 bounces, independent PCG samples) and the
    two-level instanced scene (4,096 instances of a 9,680-triangle mesh, a
    2048x1024 environment map, 1024^2, 32 spp, 12 bounces, Sobol), the latter
    over both trees when `treelet` (both levels restructured; and as built).
    Scenes are resident before timing; one warm-up render each."""
    out = {}
    runs = [("box_c1", "box-c1", 256, 256, 64, 8, L.SAMPLER_INDEPENDENT, treelet),
            ("instances_c3", "instances-c3", 1024, 1024, 32, 12, L.SAMPLER_SOBOL, treelet)]
    if treelet:
        runs.append(("instances_c3_lbvh_tree", "instances-c3", 1024, 1024, 32, 12, L.SAMPLER_SOBOL, False))
    for key, name, w, h, spp, depth, sampler, tl in runs:
        blob = L.Blob(name)
        t0 = time.perf_counter()
        scene = L.Scene(gpu, blob, treelet=tl)
        gpu.check_errors()
        build_ms = (time.perf_counter() - t0) * 1e3
        film = L.Film(gpu, w, h)
        u = blob.uniforms(w, h, samples=spp, max_depth=depth, sampling=L.SAMPLING_MIS, sampler=sampler)
        L.render_wavefront(gpu, scene, u, spp, film)
        gpu.check_errors()
        film.clear()
        t0 = time.perf_counter()
        st = L.render_wavefront(gpu, scene, u, spp, film)
        gpu.check_errors()
        sec = time.perf_counter() - t0
        counts = scene.counts()
        out[key] = {"mpaths_per_s": round(w * h * spp / sec / 1e6, 2), "ms": round(sec * 1e3, 2),
                    "device_ms": round(st.elapsed_ms, 2), "width": w, "height": h, "spp": spp, "max_depth": depth,
                    "extension_rays": st.extension_rays, "shadow_rays": st.shadow_rays,
                    "scene_upload_and_build_ms": round(build_ms, 1), "triangles": counts.get("triangles"),
                    "instances": counts.get("instances"), "tree": "treelet" if tl else "lbvh"}
        del film, scene
    return out


def architectures(L, gpu, runs: int = 3) -> dict:
    """The reference's architecture comparison (README.md:703-714: wavefront
    speed over megakernel speed at 512^2, 8 spp, depth 8): both renders of
    each scene timed, median of `runs` after a warm-up. The films are
    bit-identical (tests/test_megakernel_gpu.py)."""
    out = {}
    for name in ("furnace-test", "cornell-box", "cornell-mesh", "bvh-stress"):
        blob = L.Blob(name)
        scene = L.Scene(gpu, blob, build_lbvh=True)
        u = blob.uniforms(512, 512, samples=8, max_depth=8)
        film = L.Film(gpu, 512, 512)
        ms = {}
        for arch in L.ARCHITECTURES:
            L.render_with(gpu, scene, u, 8, film, arch)
            gpu.check_errors()
            t = []
            for _ in range(runs):
                film.clear()
                t0 = time.perf_counter()
                L.render_with(gpu, scene, u, 8, film, arch)
                t.append((time.perf_counter() - t0) * 1e3)
            ms[arch] = float(np.median(t))
        out[name] = {"megakernel_ms": round(ms["megakernel"], 3), "wavefront_ms": round(ms["wavefront"], 3),
                     "wavefront_speed_over_megakernel": round(ms["megakernel"] / ms["wavefront"], 3)}
        del film, scene
    return out


def relmse(x: np.ndarray, ref: np.ndarray) -> float:
    """Relative MSE, mean over pixels and channels of (x - r)^2 / (r^2 + 1e-2).
    The reference defines no relMSE; this is the definition used throughout."""
    x = x.astype(np.float64)
    r = ref.astype(np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + 1e-2)))


def time_to_relmse(L, gpu, scene, blob, width: int, height: int, depth: int, thresholds, ref_spp: int,
                   max_spp: int, net=None) -> dict:
    """Progressive render of the same scene, sample counts 1, 2, 4, ..., max_spp,
    scored after each doubling against an independent ref_spp render (another
    frame seed, so no sample is shared). Time = device render time of all
    samples so far (+ the denoiser's time for the denoised variants); the
    crossing of each threshold is interpolated log-log between the
    bracketing sample counts, with time linear in samples. The first
    threshold is the metric's (0.01); a tighter one (1e-3) separates the
    denoisers, which both reach 0.01 at one sample."""
    threshold = min(thresholds)
    mk = lambda spp, off, seed: blob.uniforms(width, height, samples=spp, max_depth=depth, sampling=L.SAMPLING_MIS,
                                              sampler=L.SAMPLER_SOBOL, sample_offset=off, frame_seed=seed)
    ref_film = L.Film(gpu, width, height)
    t0 = time.perf_counter()
    st = L.render_wavefront(gpu, scene, mk(ref_spp, 0, 0x0BADF00D), ref_spp, ref_film)
    gpu.check_errors()
    ref_s = time.perf_counter() - t0
    ref = L.film_mean(ref_film.download(), width, height)
    del ref_film
    film = L.Film(gpu, width, height)
    den = L.Film(gpu, width, height)
    dn = L.Denoiser(gpu)
    variants = {"plain": [], "atrous": []}
    if net is not None:
        variants["learned"] = []
    # One untimed call of each denoiser at this size first: the first call
    # allocates its planes (a one-time set-up, not a per-frame cost).
    dn.run(film, den)
    if net is not None:
        net.denoise(film, den)
    gpu.check_errors()
    render_ms = 0.0
    done = 0
    spp = 1
    while spp <= max_spp:
        add = spp - done
        st = L.render_wavefront(gpu, scene, mk(add, done, 0x5EED1234), add, film)
        render_ms += st.elapsed_ms
        done = spp
        acc = film.download()
        variants["plain"].append((spp, render_ms, relmse(L.film_mean(acc, width, height), ref)))
        for name in [k for k in variants if k != "plain"]:
            gpu.check_errors()
            ta = time.perf_counter()
            if name == "atrous":
                dn.run(film, den)
            else:
                net.denoise(film, den)
            gpu.check_errors()
            dms = (time.perf_counter() - ta) * 1e3
            variants[name].append((spp, render_ms + dms, relmse(L.film_mean(den.download(), width, height), ref)))
        if all(v[-1][2] <= threshold for v in variants.values()):
            break
        spp *= 2

    def crossing(pts, threshold):
        for i, (s, t, e) in enumerate(pts):
            if e <= threshold:
                if i == 0:
                    return {"spp": float(s), "ms": round(t, 3), "relmse": e}
                s0, t0_, e0 = pts[i - 1]
                f = (np.log(e0) - np.log(threshold)) / (np.log(e0) - np.log(e))
                s_x = float(np.exp(np.log(s0) + f * (np.log(s) - np.log(s0))))
                return {"spp": round(s_x, 2), "ms": round(t0_ + (t - t0_) * (s_x - s0) / (s - s0), 3),
                        "relmse_at_bracket": [e0, e]}
        return None

    out = {"threshold": thresholds[0], "thresholds": list(thresholds),
           "definition": "mean (x-r)^2/(r^2+1e-2) over pixels and channels",
           "resolution": [width, height], "reference_spp": ref_spp, "reference_render_s": round(ref_s, 2),
           "max_spp": max_spp}
    for name, pts in variants.items():
        out[name] = {"curve": [[s, round(t, 3), round(e, 7)] for s, t, e in pts]}
        for k, th in enumerate(thresholds):
            out[name]["reached" if k == 0 else f"reached_{th:g}"] = crossing(pts, th)
    for k, th in enumerate(thresholds):
        key = "reached" if k == 0 else f"reached_{th:g}"
        p = out["plain"][key]
        for name in [v for v in variants if v != "plain"]:
            r = out[name][key]
            out[name]["speedup_vs_plain" if k == 0 else f"speedup_vs_plain_{th:g}"] = \
                round(p["ms"] / r["ms"], 2) if p and r else None
    if "learned" in variants:
        out["learned_vs_atrous_relmse"] = {str(s): round(e / a, 4) for (s, _, e), (_, _, a)
                                           in zip(variants["learned"], variants["atrous"])}
    return out


def batch_c4(L, gpu, jobs_n: int = 64, seed: int = 6, treelet: bool = True) -> dict:
    """The batched request stream: `jobs_n` renders drawn by a PCG64 stream
    (seed 6) over resolution {128, 256, 512, 1024}^2 x spp {1, 4, 16, 64} x
    scene {the Cornell box, four displaced-sphere rooms (seeds 1-4, 81,920 +
    1,024 triangles)}, 8 bounces, all trees resident (LBVHs restructured by
    treelets when `treelet`, else as built). Run once job after job
    (each lucida_render_wavefront waits for the previous) and once as one
    lucida_render_batch with the autotuned launch shape; latency = time from
    the start until a job's film is complete, all jobs arriving at once."""
    from lucida_amd import autotune
    from lucida_amd import request_stream as RS
    blobs = RS.stream_blobs()
    scenes = [L.Scene(gpu, b, treelet=treelet) for b in blobs]
    reqs = RS.stream_requests(jobs_n, seed, len(blobs))
    paths = sum(r * r * spp for _, r, spp, _ in reqs)

    def jobs():
        return RS.make_jobs(gpu, blobs, scenes, reqs)

    seq_jobs = jobs()
    L.render_wavefront(gpu, seq_jobs[0][0], seq_jobs[0][1], 1, L.Film(gpu, seq_jobs[0][1].width,
                                                                       seq_jobs[0][1].height))  # warm-up
    gpu.check_errors()
    t0 = time.perf_counter()
    seq_done = []
    for scene, u, spp, film in seq_jobs:
        L.render_wavefront(gpu, scene, u, spp, film)
        gpu.check_errors()
        seq_done.append((time.perf_counter() - t0) * 1e3)
    seq_ms = seq_done[-1]
    del seq_jobs
    bjobs = jobs()
    t_tune = time.perf_counter()
    cfg = autotune.tune(gpu, bjobs)
    tune_s = time.perf_counter() - t_tune
    trials = autotune.last_trials(gpu, bjobs)
    for _, _, _, f in bjobs:
        f.clear()
    gpu.check_errors()
    t0 = time.perf_counter()
    st, done = L.render_batch(gpu, bjobs, cfg)
    gpu.check_errors()
    batch_ms = (time.perf_counter() - t0) * 1e3
    q = lambda a, p: round(float(np.percentile(a, p)), 2)
    return {"jobs": jobs_n, "paths": int(paths), "max_depth": 8, "tree": "treelet" if treelet else "lbvh",
            "sequential": {"ms": round(seq_ms, 2), "mpaths_per_s": round(paths / seq_ms / 1e3, 2),
                           "latency_p50_ms": q(seq_done, 50), "latency_p99_ms": q(seq_done, 99)},
            "batched": {"ms": round(batch_ms, 2), "device_ms": round(st.elapsed_ms, 2),
                        "mpaths_per_s": round(paths / batch_ms / 1e3, 2), "rounds": int(st.kernel_launches[0]),
                        "latency_p50_ms": q(done, 50), "latency_p99_ms": q(done, 99),
                        "config": {"max_paths": cfg.max_paths, "use_graph": cfg.use_graph},
                        "autotune_s": round(tune_s, 2),
                        "autotune_trials_ms": {f"max_paths {t['max_paths']} graph {t['use_graph']}": t["ms"]
                                               for t in trials}}}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=7)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--width", type=int, default=1024)
    ap.add_argument("--height", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--depth", type=int, default=8)
    ap.add_argument("--frequency", type=int, default=229)
    ap.add_argument("--samples-per-batch", type=int, default=0)
    ap.add_argument("--ray-order", type=int, default=0, choices=(0, 1, 2),
                    help="extension-queue order: 0 automatic (sorted for this scene), 1 kept, 2 always sorted")
    ap.add_argument("--tree", choices=("treelet", "lbvh"), default="treelet",
                    help="the scene's tree: the LBVH restructured by treelets, or the LBVH as built")
    ap.add_argument("--no-lbvh-compare", action="store_true",
                    help="skip timing the step over the LBVH as built (profiling runs)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--relmse-ref-spp", type=int, default=4096)
    ap.add_argument("--relmse-max-spp", type=int, default=1024)
    ap.add_argument("--no-relmse", action="store_true")
    ap.add_argument("--net-scenes", type=int, default=96, help="training scenes for the learned denoiser")
    ap.add_argument("--net-ref-spp", type=int, default=4096, help="target samples per pixel of the training pairs")
    ap.add_argument("--net-steps", type=int, default=4000, help="Adam steps (0: no learned denoiser)")
    ap.add_argument("--net-crop", type=int, default=256, help="training crop side (8 crops per step)")
    ap.add_argument("--net-ema", type=float, default=0.9995,
                    help="decay of the moving average of the weights over the second half of training (0: last weights)")
    ap.add_argument("--serialize", action="store_true",
                    help="time the steps serialized (one pipeline, every kernel alone on one stream), as the "
                         "rocprofv3 counter runs see them")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched request stream")
    ap.add_argument("--no-configs", action="store_true",
                    help="skip the Cornell-box and instanced-scene throughput figures")
    ap.add_argument("--denoiser-timing", action="store_true",
                    help="only time the two denoisers on a 1-spp render of the benchmark scene (untrained network "
                         "weights: inference cost does not depend on them) and exit")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the library's streaming copy of 256 MiB three times (the known-size kernel "
                         "the FETCH_SIZE / WRITE_SIZE counters are scaled by) and exit")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {a.gpus} is not supported "
              f"(the path would shard by sample range, merging the films with one sum all-reduce)", file=sys.stderr)
        return 2

    import lucida_amd as L

    gpu = L.Gpu(0)
    dev_name, cus = gpu.info()
    if a.copy_calibration:
        lib = L.lib()
        nbytes = 256 << 20
        bufs = []
        for _ in range(2):
            p = C.c_void_p()
            L._check(lib.lucida_device_alloc(gpu._h, nbytes, C.byref(p)))
            bufs.append(p)
        for _ in range(3):
            L._check(lib.lucida_copy_device(gpu._h, bufs[1], bufs[0], nbytes))
        for p in bufs:
            lib.lucida_device_free(gpu._h, p)
        print(json.dumps({"copy_calibration_bytes": nbytes, "launches": 3}), flush=True)
        gpu.close()
        return 0
    blob = L.Blob("sphere-c2", frequency=a.frequency)
    ntris = blob.desc().num_triangles
    if a.denoiser_timing:
        scene = L.Scene(gpu, blob, build_lbvh=True)
        film, den = L.Film(gpu, a.width, a.height), L.Film(gpu, a.width, a.height)
        u = blob.uniforms(a.width, a.height, samples=1, max_depth=a.depth, sampling=L.SAMPLING_MIS,
                          sampler=L.SAMPLER_SOBOL)
        st = L.render_wavefront(gpu, scene, u, 1, film)
        dn, net = L.Denoiser(gpu), L.LearnedDenoiser(gpu, seed=1)
        res = {"render_1spp_ms": round(st.elapsed_ms, 3)}
        for name, run in (("atrous_ms", lambda: dn.run(film, den)), ("learned_ms", lambda: net.denoise(film, den))):
            run()  # first call allocates
            gpu.check_errors()
            t0 = time.perf_counter()
            for _ in range(10):
                run()
            gpu.check_errors()
            res[name] = round((time.perf_counter() - t0) / 10 * 1e3, 3)
        print(json.dumps({"denoiser_timing": res, "width": a.width, "height": a.height}), flush=True)
        gpu.close()
        return 0
    lbvh = lbvh_build_rate(L, gpu, blob)
    treelet = a.tree == "treelet"
    treelet_build = lbvh_build_rate(L, gpu, blob, treelet=True) if treelet else None
    scene = L.Scene(gpu, blob, build_lbvh=True, treelet=treelet)
    nodes, tris_sorted = scene.download_bvh()
    blob.set_bvh(nodes, tris_sorted)  # the CPU baseline traverses the same tree
    film = L.Film(gpu, a.width, a.height)
    u = blob.uniforms(a.width, a.height, samples=a.spp, max_depth=a.depth, sampling=L.SAMPLING_MIS,
                      sampler=L.SAMPLER_SOBOL)
    # Timed renders run as a caller runs them: no traversal counters, no
    # per-kernel events. Kernel times come from untimed renders of the same
    # (deterministic) step: one as the bench runs it (kernels overlapping
    # across streams and pipelines) and one serialized (every kernel alone on
    # one stream), whose times are solo times; node visits and triangle tests
    # from one more, counted, render.
    mk_cfg = lambda count=0, profile=0, serialize=0: L.WavefrontConfig(
        a.samples_per_batch, 0, 0, 0, count, profile, a.ray_order, serialize)
    cfg = mk_cfg(serialize=1 if a.serialize else 0)

    for _ in range(a.warmup):
        film.clear()
        L.render_wavefront(gpu, scene, u, a.spp, film, cfg)
    gpu.check_errors()

    step_s = []
    t0 = time.perf_counter()
    for _ in range(a.steps):
        film.clear()
        ts = time.perf_counter()
        L.render_wavefront(gpu, scene, u, a.spp, film, cfg)  # returns when the render is complete
        step_s.append(time.perf_counter() - ts)
    gpu.check_errors()
    total_s = time.perf_counter() - t0
    sec = float(np.median(step_s))

    # The same step over the LBVH as built (the reference's tree), timed the
    # same way, beside the restructured tree's.
    lbvh_tree_render = None
    if treelet and not a.no_lbvh_compare:
        # a fresh blob: `blob` now carries the restructured nodes for the CPU baseline
        plain = L.Scene(gpu, L.Blob("sphere-c2", frequency=a.frequency), build_lbvh=True)
        film.clear()
        L.render_wavefront(gpu, plain, u, a.spp, film, cfg)
        plain_s = []
        for _ in range(3):
            film.clear()
            ts = time.perf_counter()
            L.render_wavefront(gpu, plain, u, a.spp, film, cfg)
            plain_s.append(time.perf_counter() - ts)
        gpu.check_errors()
        plain_ms = float(np.median(plain_s)) * 1e3
        lbvh_tree_render = {"ms_per_step": round(plain_ms, 3), "steps": 3, "statistic": "median",
                            "mpaths_per_s": round(a.width * a.height * a.spp / (plain_ms * 1e-3) / 1e6, 3),
                            "treelet_speedup": round(plain_ms / (sec * 1e3), 4)}
        del plain

    def kernel_times(serialize: int) -> dict:
        film.clear()
        st = L.render_wavefront(gpu, scene, u, a.spp, film, mk_cfg(profile=1, serialize=serialize))
        gpu.check_errors()
        ms = {k: float(st.kernel_ms[i]) for i, k in enumerate(L.RenderStats.KERNELS)}
        ms["queue_sort"] = float(st.queue_sort_ms)
        n = {k: int(st.kernel_launches[i]) for i, k in enumerate(L.RenderStats.KERNELS)}
        n["queue_sort"] = int(st.queue_sorts)
        return {"ms": ms, "launches": n, "device_ms": float(st.elapsed_ms)}

    overlapped = kernel_times(0)
    solo = kernel_times(1)
    film.clear()
    cst = L.render_wavefront(gpu, scene, u, a.spp, film, mk_cfg(count=1))
    gpu.check_errors()
    ext_rays, shadow_rays = cst.extension_rays, cst.shadow_rays
    ext_visits, ext_tests = cst.extend_node_visits, cst.extend_triangle_tests
    con_visits, con_tests = cst.node_visits - ext_visits, cst.triangle_tests - ext_tests
    ext_slots, con_slots = cst.extend_lane_slots, cst.connect_lane_slots
    paths = a.width * a.height * a.spp
    value = paths / sec / 1e6

    # Roofline of the dominant kernel, EXTEND (the closest-hit walk), from its
    # solo launch time: algorithmic bytes per launch / average solo launch.
    ext_launches = max(solo["launches"]["extend"], 1)
    ext_solo_ms = solo["ms"]["extend"] / ext_launches
    ext_bytes = (ext_rays * RAY_BYTES + NODE_BYTES * ext_visits + TRI_BYTES * ext_tests) / ext_launches
    achieved = ext_bytes / (ext_solo_ms * 1e-3) / 1e9
    con_launches = max(solo["launches"]["connect"], 1)
    con_solo_ms = solo["ms"]["connect"] / con_launches
    con_bytes = (shadow_rays * SHADOW_BYTES + NODE_BYTES * con_visits + TRI_BYTES * con_tests) / con_launches
    dominant = max((k for k in solo["ms"] if k != "queue_sort"), key=lambda k: solo["ms"][k])

    def measured(kernel: str, ms: float) -> dict:
        """HBM bytes and VALU wave instructions per launch from the committed
        rocprofv3 summaries (serialized launches of the same step), over this
        run's solo launch time."""
        traffic = profile_figure(TRAFFIC_SUMMARY, f"{kernel}_hbm_bytes_per_launch")
        valu = profile_figure(SQ_SUMMARY, f"{kernel}_valu_insts_per_launch")
        active = profile_figure(SQ_SUMMARY, f"{kernel}_active_lane_fraction")
        out = {"hbm_bytes_per_launch": traffic, "valu_insts_per_launch": valu, "active_lane_fraction": active}
        if traffic:
            out["hbm_gbs"] = round(traffic / (ms * 1e-3) / 1e9, 1)
            out["hbm_frac_of_8tbs"] = round(traffic / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        if valu:
            out["valu_issue_frac"] = round(valu / (SIMDS * CLOCK_HZ / CLOCKS_PER_VALU * ms * 1e-3), 4)
        return out

    # a-trous on the final film (1024^2): median of 7 after a warm-up
    den = L.Film(gpu, a.width, a.height)
    dn = L.Denoiser(gpu)
    dn.run(film, den)
    gpu.check_errors()
    at = []
    for _ in range(7):
        td = time.perf_counter()
        dn.run(film, den)
        at.append(time.perf_counter() - td)
    atrous_ms = float(np.median(at)) * 1e3
    atrous_gbs = ATROUS_BYTES_PER_PX * a.width * a.height / (atrous_ms * 1e-3) / 1e9

    r3 = lambda d: {k: round(v, 3) for k, v in d.items()}
    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpaths/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(sec * 1e3, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic (procedural scene, seeded)",
        "config": {"workload": f"sphere-c2: displaced geodesic sphere {ntris - 1024} tris + 1,024-tri room, "
                               f"16 emissive quads, {a.width}x{a.height}, {a.spp} spp, {a.depth} bounces, "
                               f"NEE+MIS, Sobol", "triangles": ntris, "width": a.width, "height": a.height,
                   "spp": a.spp, "max_depth": a.depth, "parallelism": "single GPU",
                   "tree": (f"LBVH restructured by treelets ({L.TREELET_ROUNDS} rounds), built on the GPU"
                            if treelet else "LBVH as built on the GPU")},
        "timing": {"statistic": "median of the timed steps' wall times (each step a complete render)",
                   "mean_ms_per_step": round(total_s / a.steps * 1e3, 3),
                   "min_ms": round(min(step_s) * 1e3, 3), "max_ms": round(max(step_s) * 1e3, 3),
                   "serialized": bool(a.serialize)},
        "roofline": {"bound": "hbm", "kernel": "extend (closest-hit traversal)", "achieved": round(achieved, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": profile_figure(TRAFFIC_SUMMARY, "extend_hbm_bytes_per_launch"),
                     "achieved_is": "algorithmic bytes (queue index, ray, hit, 64 B per node visit, 48 B per "
                                    "triangle test) per launch over the solo launch time; most node and triangle "
                                    "reads are served by L2 and the MALL, so this is not HBM traffic",
                     "limiter": "VALU issue over diverged lanes (see measured.valu_issue_frac and "
                                "active_lane_fraction), not HBM bandwidth (measured.hbm_frac_of_8tbs)",
                     "avg_launch_ms": round(ext_solo_ms, 4), "launch_time": "solo (serialized render)",
                     "bytes_per_launch": int(ext_bytes),
                     "measured": measured("extend", ext_solo_ms)},
        "extra": {
            "device": dev_name, "compute_units": cus,
            "mrays_per_s": round((ext_rays + shadow_rays) / sec / 1e6, 2),
            "extension_rays_per_step": ext_rays, "shadow_rays_per_step": shadow_rays,
            "bvh_walk": "binary",
            "walk_simd_efficiency": {"extend": round(ext_visits / max(ext_slots, 1), 3),
                                     "connect": round(con_visits / max(con_slots, 1), 3)},
            "node_visits_per_ray": round(cst.node_visits / max(ext_rays + shadow_rays, 1), 2),
            "triangle_tests_per_ray": round(cst.triangle_tests / max(ext_rays + shadow_rays, 1), 2),
            "kernel_ms_per_step_solo": r3(solo["ms"]),
            "kernel_ms_per_step_overlapped": r3(overlapped["ms"]),
            "kernel_launches_per_step": solo["launches"],
            "device_ms_per_step": {"as_timed": round(overlapped["device_ms"], 3),
                                   "serialized": round(solo["device_ms"], 3)},
            "dominant_kernel": dominant,
            "connect_roofline": {"achieved": round(con_bytes / (con_solo_ms * 1e-3) / 1e9, 1),
                                 "frac": round(con_bytes / (con_solo_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                                 "avg_launch_ms": round(con_solo_ms, 4), "bytes_per_launch": int(con_bytes),
                                 "measured": measured("connect", con_solo_ms)},
            "lbvh": lbvh,
            "treelet_build": treelet_build,
            "lbvh_tree_render": lbvh_tree_render,
            "atrous": {"ms": round(atrous_ms, 3), "bytes_per_px": ATROUS_BYTES_PER_PX,
                       "achieved_gbs": round(atrous_gbs, 1), "frac_of_8tbs": round(atrous_gbs / HBM_PEAK_GBS, 4)},
        },
    }
    out["extra"]["convergence_meter"] = round(film.convergence(), 6)
    if not a.no_configs:
        out["extra"]["other_configs"] = other_configs(L, gpu, treelet)
        out["extra"]["architectures_512x512_8spp"] = architectures(L, gpu)
    if not a.no_relmse:
        net = None
        if a.net_steps > 0:
            from lucida_amd import denoiser_training as T
            t0 = time.perf_counter()
            progress = lambda msg: print(f"denoiser training: {msg}", file=sys.stderr, flush=True)
            pairs, data_paths = T.make_training_set(
                gpu, a.net_scenes, input_spp=T.INPUT_SPP, ref_spp=a.net_ref_spp,
                log=lambda msg: progress(msg) if msg.startswith("scene ") and int(msg.split()[1][:-1]) % 8 == 7
                else None)
            t1 = time.perf_counter()
            net = L.LearnedDenoiser(gpu, seed=1)
            losses = T.train(gpu, net, pairs, a.net_steps, crop_w=a.net_crop, crop_h=a.net_crop, log_every=1000,
                             log=progress, ema=a.net_ema)
            t2 = time.perf_counter()
            del pairs
            out["learned_denoiser"] = {
                "params": L.LearnedDenoiser.param_count(),
                "training_scenes": a.net_scenes, "crops": f"8 x {a.net_crop}x{a.net_crop} per step",
                "steps": a.net_steps, "weight_average_decay": a.net_ema,
                "target_spp": a.net_ref_spp, "data_paths": data_paths, "data_s": round(t1 - t0, 2),
                "train_s": round(t2 - t1, 2), "ms_per_step": round((t2 - t1) / a.net_steps * 1e3, 3),
                "loss_first_100": float(np.mean(losses[:100])), "loss_last_100": float(np.mean(losses[-100:])),
                "input_spp": list(T.INPUT_SPP), "input_channels": 16,
                "held_out": "the benchmark scene (seed 0, frequency 229) is not in the training set"}
        out["time_to_relmse"] = time_to_relmse(L, gpu, scene, blob, a.width, a.height, a.depth, (0.01, 1e-3),
                                               a.relmse_ref_spp, a.relmse_max_spp, net)
    if not a.no_batch:
        out["batch"] = batch_c4(L, gpu, treelet=treelet)
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(blob, u, a.cpu_seconds)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    gpu.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
