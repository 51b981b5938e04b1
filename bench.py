"""Benchmark of the ray hot path on one MI355X: closest-hit Mrays/s on the 1.3M-triangle mesh.

Workload (1,310,720-triangle pure-traversal configuration): an icosphere of 8 subdivisions, radius 1,
5-octave seeded value-noise radial displacement of 10% amplitude, FP32 vertices, binned-SAH BVH with
12 bins (leaves of at most 4 triangles). Two ray batches of 4,194,304 rays each, resident in HBM:
  (a) coherent primary rays of a 2048x2048 pinhole through pixel centres, framed as the reference's
      bvh_stats does (FOV 40 degrees, elevation 20 degrees, tangent to the bounding sphere);
  (b) incoherent rays, origins uniform in the bounding sphere, directions uniform on the sphere.
tmin 1e-4, tmax 1000. One step = closest hit over both batches (8,388,608 rays), the two batches
traced at once on two streams (`--sequential`: one after the other on one stream).

Prints one JSON line. `roofline` describes the dominant kernel (closest hit): algorithmic bytes per
launch = sum over rays of 32 (ray) + 16 (hit) + node records fetched x their size (128 bytes on the
4-wide tree the timed kernels walk) + 48 x triangle records fetched, with the counts measured
exactly by the fetch-counting walk of the same tree on the same rays, divided by the time the
step's two launches occupy the GPU (HIP events: fork to join of the step on the context's stream;
`frac_alone` divides by the two launches' times each alone). `cpu_baseline` times the f64 CPU restatement of the
reference algorithm (the reference's own BVH parameters: 8 bins, t_trav = 0) over the same rays;
`parity` checks the timed launches' own hits against the FP32 and f64 restatements.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
SEED_MESH = 2024
SEED_RAYS = 7

# Memory-side bytes of one k_closest launch on each batch of this workload, from rocprofv3
# FETCH_SIZE and WRITE_SIZE passes scaled by a 1 GiB copy (FETCH x2, WRITE x1), measured on the
# bench's fixed launches; profiles/k_closest_icosphere1p3m_counters.md. Keyed by (batch, block, lds_stack,
# refill, tree width, packet, image width of the 8 x 16 tiles or 0 for batch order): reported only
# for the kernel variants that were measured.
_COUNTERS = "profiles/k_closest_icosphere1p3m_counters.md"
MEASURED_TRAFFIC = {
    ("coherent", 128, 12, 0, 4, 0, 2048): {"bytes": 254.70e6, "source": _COUNTERS},
    ("incoherent", 128, 8, 1, 4, 0, 0): {"bytes": 2280.90e6, "source": _COUNTERS},
}
LAUNCH_KEYS = ("block", "waves_per_cu", "lds_stack", "refill", "width", "packet")
LAUNCH_DEFAULTS = dict(block=128, waves_per_cu=16, lds_stack=16, refill=0, width=2, packet=0)
# The launches of the two batches, fixed so that runs are reproducible and match the measured
# counters. Tree width, block, stack depth and refill are the autotuner's picks on each batch alone
# (pbrt_mesh_autotune, which alone picks 64 and 24 waves per CU); the grids were then swept with the
# two batches running at once, where over-subscribed grids on both sides win (waves per CU 128 and
# 32: 7,342-7,492 Mrays/s against 6,860 with the picks made alone). --autotune searches again.
DEFAULT_BATCH_LAUNCHES = "128,128,12,0,4;128,32,8,1,4"


def rows(r):
    return np.ascontiguousarray(r).view(np.float32).reshape(-1, 8)


def make_workload(pa, side=2048, n_incoherent=1 << 22):
    v, f = pa.icosphere(8, radius=1.0, noise="fbm", amp=0.10, octaves=5, freq=1.5, seed=SEED_MESH)
    coherent = pa.orbit_rays(v, 0, side)
    lo, hi = v.min(0).astype(np.float64), v.max(0).astype(np.float64)
    center = (lo + hi) * 0.5
    radius = float(np.sqrt(((hi - lo) ** 2).sum()) * 0.5)
    incoherent = pa.random_rays(center, radius, n_incoherent, seed=SEED_RAYS)
    return v, f, coherent, incoherent


def host_info():
    """What the host offers this process: CPUs in its affinity mask, the machine's CPU count, the
    cgroup CPU quota (cpu.max: quota and period in microseconds, if readable) and the model name."""
    info = {"affinity_cpus": len(os.sched_getaffinity(0)), "cpu_count": os.cpu_count()}
    try:
        with open("/sys/fs/cgroup/cpu.max") as fh:
            q = fh.read().split()
        info["cgroup_cpu_max"] = " ".join(q)
        if q and q[0] != "max":
            info["cgroup_cpus"] = round(int(q[0]) / int(q[1]), 2)
    except (OSError, ValueError, IndexError):
        info["cgroup_cpu_max"] = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def _fastest(fn, passes=3):
    fn()  # warm-up
    best = None
    for _ in range(passes):
        t0 = time.perf_counter()
        out = fn()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best, out


def cpu_baseline(pa, v, f, coherent, incoherent, side, threads, device_hits, device_occ):
    """The CPU path timed on this host, and the checks of the device's results against it.

    Timing: Oracle A (f64 restatement of the reference traversal over the reference's BVH, 8 bins,
    t_trav = 0) over the whole step (both 4M-ray batches) on `threads` threads, one warm-up pass then
    the fastest of 3 (the bvh_stats protocol, bvh_stats.rs:162-185); Oracle B (FP32, the device's
    tree) the same way; Oracle A on one thread over every 16th image row of batch (a) and the first
    131,072 rays of batch (b), for comparison with the reference's single-thread figures.
    Parity (the timed kernels' own outputs, read back after the timed loop): closest hits against
    Oracle B field by field, any hit against Oracle B's, every triangle that differs from Oracle A
    classified, hit t against the f64 re-evaluation of the hit triangle."""
    from oracle import parity
    from oracle import pyoracle as po
    t0 = time.time()
    ma = parity.oracle_a(v, f)
    mb = parity.oracle_b(v, f, pa.DEVICE_BVH)
    build_s = time.time() - t0
    batches = {"coherent": rows(coherent), "incoherent": rows(incoherent)}
    sample = np.concatenate(list(batches.values()))
    ta, ra = _fastest(lambda: ma.trace(sample, mode=po.MODE_REFERENCE, threads=threads))
    tb, rb = _fastest(lambda: mb.trace(sample, mode=po.MODE_CANONICAL, threads=threads))
    small = np.concatenate([rows(coherent).reshape(side, side, 8)[::16].reshape(-1, 8), rows(incoherent)[:131072]])
    t1, _ = _fastest(lambda: ma.trace(small, mode=po.MODE_REFERENCE, threads=1), passes=1)
    parity_out = {}
    off = 0
    for k, r in batches.items():
        n = len(r)
        sl = slice(off, off + n)
        off += n
        a_k = {x: ra[x][sl] for x in ("prim", "t", "u", "v")}
        b_k = {x: rb[x][sl] for x in ("prim", "t", "u", "v")}
        h = device_hits[k]
        any_b = mb.trace(r, mode=po.MODE_ANY, threads=threads)["prim"] >= 0
        cls = parity.classify_vs_reference(a_k, h["prim"], h["t"], h["u"], h["v"], v, f, r)
        cls.pop("unexplained_rays")
        parity_out[k] = {
            "mismatch_vs_oracle_b": parity.compare_hits(h, b_k),
            "anyhit_mismatch_vs_oracle_b": int((device_occ[k].astype(bool) != any_b).sum()),
            "vs_oracle_a": cls,
            "t_vs_fp64": parity.t_error_by_incidence(v, f, r, h["prim"].astype(np.int64), h["t"]),
        }
    host = host_info()
    return dict(value=len(sample) / ta / 1e6, unit="Mrays/s", cores=threads, kind="port",
                sample=f"the full step ({len(sample)} rays, both batches) on {threads} threads (the CPUs this "
                       f"process may use, at most 16, the host share of one GPU); f64 C++ restatement of the "
                       f"reference traversal over the reference BVH (8 bins, t_trav=0); one warm-up pass, then "
                       f"the fastest of 3; BVH builds {build_s:.1f}s not timed",
                seconds=ta, oracle_b_mrays_s=len(sample) / tb / 1e6,
                one_thread_mrays_s=round(len(small) / t1 / 1e6, 3), one_thread_sample_rays=len(small),
                host=host), parity_out


def first_below(prog, bar):
    """Milliseconds and samples per pixel at the first pass whose RMSE is at or below `bar`."""
    hit = np.nonzero(prog["rmse"] <= bar)[0]
    if len(hit) == 0:
        return None, None
    k = int(hit[0])
    return float(prog["ms"][k]), int(prog["spp"][k])


NRC_VARIANTS = {
    # the headline race: the reference's estimator (next-event estimation, emission after specular
    # bounces) under the Sobol sampler, on both sides
    "sobol": dict(sampler="sobol", mis=False),
    "sobol_mis": dict(sampler="sobol", mis=True),
    "independent": dict(sampler="independent", mis=False),
    "independent_mis": dict(sampler="independent", mis=True),
}


# Pass sizes each side of a race is run at; each side is timed at the size that reaches the bar
# soonest. Larger passes trace faster per sample, but the bar is only seen at the end of a pass:
# on the cache workload, plain path tracing with the Sobol sampler took 504 ms with passes of 4
# samples per pixel, 489 with 16 and 477 with 64 (the bar, first crossed at 242 spp, seen at 256),
# while with MIS, whose bar is crossed at 160 spp, passes of 64 see it only at 192.
PLAIN_PASS_SIZES = (4, 16, 64)
CACHE_PASS_SIZES = (4, 8, 16)


def nrc_race(pa, ctx, scene, cam, ref, bar, frames, plain_spp, sampler, mis):
    """One race to RMSE `bar` against `ref`: plain path tracing against the radiance cache, both with
    the given sampler and estimator, each at its fastest pass size, and the cache-terminated render
    without the learned part."""
    W, H, D = cam["width"], cam["height"], cam["max_depth"]
    view = dict(fov=cam["fov"], mis=mis, sampler=sampler)
    E, L = cam["eye"], cam["look"]
    plain_by, best_p = {}, None
    for sp in PLAIN_PASS_SIZES:
        _, st_, pr_ = scene.render_ex(W, H, plain_spp, D, E, L, seed=8, reference=ref, samples_per_pass=sp, **view)
        t, n = first_below(pr_, bar)
        plain_by[sp] = {"ms": t, "spp": n, "ms_per_spp": float(pr_["ms"][-1] / pr_["spp"][-1])}
        if t is not None and (best_p is None or t < best_p[0]):
            best_p = (t, n, sp, st_, pr_)
    t_plain, spp_plain, plain_pass, plain_st, plain_pr = best_p
    cache_by, best_c = {}, None
    for fp in CACHE_PASS_SIZES:
        c_ = pa.RadianceCache(ctx, seed=4)
        _, st_, pr_ = scene.render_ex(W, H, frames, D, E, L, seed=8, cache=c_, reference=ref, samples_per_pass=fp,
                                      query_vertex=3, train_stride=33, train_batch=65536, train_steps=1, **view)
        t, n = first_below(pr_, bar)
        cache_by[fp] = {"ms": t, "frames": n, "ms_per_frame": float(pr_["ms"][-1] / frames)}
        if t is not None and (best_c is None or t < best_c[0]):
            best_c = (t, n, fp, st_, pr_, c_)
    t_nrc, f_nrc, cache_pass, nrc_st, nrc_pr, cache = best_c
    k256 = int(np.searchsorted(nrc_pr["spp"], 256))
    # the same cache-terminated render without the learned part (paths end at vertex 3 with nothing)
    zero = pa.RadianceCache(ctx, seed=4)
    zero.params = np.zeros(pa.NRC_PARAMS, np.float32)
    cut, _, cut_pr = scene.render_ex(W, H, frames, D, E, L, seed=8, cache=zero, reference=ref,
                                     samples_per_pass=cache_pass, train_stride=0, **view)
    print(f"[bench] nrc ({sampler}, mis {mis}): plain reaches RMSE {bar} at {spp_plain} spp / {t_plain} ms "
          f"({plain_pass} spp per pass); cache at {f_nrc} frames / {t_nrc} ms ({cache_pass} frames per pass); at 256 "
          f"frames cache RMSE {nrc_pr['rmse'][k256]:.4f}; after {frames} frames {nrc_pr['rmse'][-1]:.4f} "
          f"(cut {cut_pr['rmse'][-1]:.4f})", file=sys.stderr, flush=True)
    return cache, {
        "sampler": sampler, "mis": mis,
        "plain_ms": t_plain, "plain_spp": spp_plain, "plain_spp_per_pass": plain_pass,
        "cache_ms": t_nrc, "cache_frames": f_nrc, "cache_frames_per_pass": cache_pass,
        "time_ratio": (t_nrc / t_plain) if (t_nrc and t_plain) else None,
        "plain_by_pass_size": plain_by, "cache_by_pass_size": cache_by,
        "cache_rmse_256_frames": float(nrc_pr["rmse"][k256]),
        "plain_rmse_256_spp": float(plain_pr["rmse"][int(np.searchsorted(plain_pr["spp"], 256))]),
        "cache_rmse_final": float(nrc_pr["rmse"][-1]), "cut_rmse_final": float(cut_pr["rmse"][-1]),
        "cache_ms_per_frame": float(nrc_pr["ms"][-1] / frames),
        "plain_ms_per_spp": float(plain_pr["ms"][-1] / plain_pr["spp"][-1]),
        "plain_mrays_s": plain_st["rays"] / (plain_pr["ms"][-1] * 1e-3) / 1e6,
        "plain_samples_s": W * H * plain_pr["spp"][-1] / (plain_pr["ms"][-1] * 1e-3),
        "cache_samples_s": W * H * frames / (nrc_pr["ms"][-1] * 1e-3),
        "cache_queries_per_frame": nrc_st["cache_queries"] / frames,
    }


def nrc_measure(pa, ctx, bar=1e-2, frames=512, ref_spp=4096, plain_spp=1024,
                variants=("sobol", "sobol_mis")):
    """Time to RMSE `bar` against a `ref_spp` plain render, plain path tracing versus the radiance
    cache, on the cache workload (the room with the 1.3M-triangle mesh and GGX materials, 1280x720,
    max depth 8). RMSE is over pixels and channels of linear radiance, reduced in FP64 on the device
    after every pass; times are HIP events on the render stream from the first launch of the render.
    Plain path tracing runs passes of PLAIN_PASS_SIZES samples per pixel, consecutive passes side by
    side on two streams; the cache renders 1 sample per pixel per frame, CACHE_PASS_SIZES frames per
    pass, one training step of 2^16 samples per frame from the 3 % training paths, and ends the
    other paths into the cache at their third vertex; each side is timed at the pass size that
    reaches the bar soonest. Each variant
    (NRC_VARIANTS: sampler and estimator, the same on both sides) races against the one reference
    render (independent sampler, the reference's estimator); the first variant is the headline."""
    from pbrt_amd import scenes
    desc, cam = scenes.nrc_room()
    t0 = time.time()
    scene = pa.Scene(ctx, desc)
    build_s = time.time() - t0
    W, H, D = cam["width"], cam["height"], cam["max_depth"]
    ref, ref_st, ref_pr = scene.render_ex(W, H, ref_spp, D, cam["eye"], cam["look"], seed=7, fov=cam["fov"])
    print(f"[bench] nrc: reference {ref_spp} spp in {ref_pr['ms'][-1] / 1e3:.1f}s, mean radiance {ref.mean():.4f}",
          file=sys.stderr, flush=True)
    races = {}
    cache = None
    for v in variants:
        c, races[v] = nrc_race(pa, ctx, scene, cam, ref, bar, frames, plain_spp, **NRC_VARIANTS[v])
        cache = cache or c

    mlp = mlp_measure(pa, ctx, cache, W * H)
    head = races[variants[0]]
    return dict({
        "metric": f"time to RMSE {bar}, radiance cache vs plain path tracing",
        "workload": "room with the 1,310,720-triangle displaced icosphere and GGX materials (roughness "
                    "U[0.05,0.5]), 1280x720, 1 spp per frame, max depth 8; RMSE of linear radiance against a "
                    f"{ref_spp}-spp plain render (mean radiance {ref.mean():.4f})",
        "headline_variant": variants[0],
        "frames": frames, "plain_pass_sizes": list(PLAIN_PASS_SIZES), "cache_pass_sizes": list(CACHE_PASS_SIZES),
        "reference_s": float(ref_pr["ms"][-1] / 1e3), "scene_build_s": build_s,
        "variants": races,
        "mlp": mlp,
    }, **head)


def mlp_measure(pa, ctx, cache, n_inf=1280 * 720, n_tr=65536, reps=20):
    """The radiance cache's MLP alone: inference over `n_inf` encoded rows (one 1280x720 frame's
    worth of queries) and one training step (forward, backward, partial reduction, Adam, BF16
    copies) of `n_tr` rows, as HIP events over `reps` launches on the library's stream; TFLOP/s of
    the algorithmic products (2 x 17,664 per row forward, 3x that per training row)."""
    rng = np.random.default_rng(1)
    raw = rng.uniform(0.0, 1.0, (max(n_inf, n_tr), 12)).astype(np.float32)
    d_raw = ctx.upload(raw)
    d_x = ctx.alloc(len(raw) * pa.NRC_FEATURES * 2)
    d_y = ctx.alloc(n_inf * 12)
    d_t = ctx.upload(rng.uniform(0.0, 1.0, (len(raw), 3)).astype(np.float32))
    cache.encode(d_raw, len(raw), d_x)
    e0, e1 = ctx.event(), ctx.event()
    cache.infer(d_x, n_inf, d_y)
    e0.record(ctx.stream)
    for _ in range(reps):
        cache.infer(d_x, n_inf, d_y)
    e1.record(ctx.stream)
    infer_ms = e0.elapsed_ms(e1) / reps
    cache.train_step(d_x, d_t, n_tr)
    e0.record(ctx.stream)
    for _ in range(reps):
        cache.train_step(d_x, d_t, n_tr)
    e1.record(ctx.stream)
    train_ms = e0.elapsed_ms(e1) / reps
    flop_q = 2 * (81 * 64 + 3 * 64 * 64 + 64 * 3)
    infer_tf = n_inf * flop_q / (infer_ms * 1e-3) / 1e12
    train_tf = n_tr * 3 * flop_q / (train_ms * 1e-3) / 1e12
    return {"infer_ms": infer_ms, "infer_queries": n_inf, "infer_tflops": infer_tf,
            "train_step_ms": train_ms, "train_samples": n_tr, "train_tflops": train_tf,
            "peak_bf16_tflops": 2500.0, "infer_frac": infer_tf / 2500.0, "train_frac": train_tf / 2500.0,
            "flop_per_query": flop_q}


def jobs_measure(pa, ctx):
    """The batched-jobs workload: 256 mesh instances (1k-200k triangles each, 9.1M in all) and 32
    CSG objects; eight render jobs (256^2 to 1024^2, 16 or 64 spp, depth 8) rendered as one batch
    sharing the wavefront's queues with per-bounce graphs, against the same jobs rendered one after
    the other, both with the scene walk's launch settings the autotuner picks on short renders (4
    spp) of the view of the job with the most samples (the library's default settings timed beside
    them). GPU time from HIP events on the render stream around each call. Beside them, sixteen
    small jobs (64^2 to 256x144, 1-4 spp), none of which fills the GPU alone, batched and one after
    the other."""
    from pbrt_amd import scenes
    t0 = time.time()
    desc = scenes.batch_scene()
    gen_s = time.time() - t0
    t0 = time.time()
    scene = pa.Scene(ctx, desc)
    build_s = time.time() - t0
    # the same scene with its 256 meshes built in one batched device build
    pa.TriangleMesh(ctx, *pa.icosphere(2), device_build=True)  # loads the build kernels
    t0 = time.time()
    scene_dev = pa.Scene(ctx, desc, device_build=True)
    dev_build_s = time.time() - t0
    jobs = scenes.batch_jobs()
    e0, e1 = ctx.event(), ctx.event()

    def timed(fn):
        e0.record(ctx.stream)
        out = fn()
        e1.record(ctx.stream)
        ctx.synchronize()
        return out, e0.elapsed_ms(e1)

    scene.render_jobs(jobs[:1], use_graphs=True)  # warm-up
    (_, st_d), default_ms = timed(lambda: scene.render_jobs(jobs, use_graphs=True))
    jt = max(jobs, key=lambda j: j["width"] * j["height"] * j["spp"])
    scene_launch = scene.autotune_render(jt["width"], jt["height"], 4, jt["max_depth"], jt["eye"], jt["look"],
                                         fov=jt["fov"], reps=2)
    print(f"[bench] jobs scene launch tuned on the largest job's view: {scene_launch}; batch at the default launch "
          f"{default_ms:.1f} ms", file=sys.stderr, flush=True)
    (films, st), batch_ms = timed(lambda: scene.render_jobs(jobs, use_graphs=True))
    (_, st_eager), eager_ms = timed(lambda: scene.render_jobs(jobs, use_graphs=False))
    seq_ms, seq_rays = 0.0, 0
    for j in jobs:
        (_, sj), ms = timed(lambda j=j: scene.render(j["width"], j["height"], j["spp"], j["max_depth"], j["eye"],
                                                      j["look"], fov=j["fov"], seed=j["seed"], film=False))
        seq_ms += ms
        seq_rays += sj["rays"]
    paths = st["paths"]
    # sixteen small jobs, none of which fills the GPU alone: where sharing the queues pays
    small = scenes.small_jobs()
    scene.render_jobs(small, use_graphs=True)  # warm-up
    (_, st_small), small_batch_ms = timed(lambda: scene.render_jobs(small, use_graphs=True))
    (_, _), small_eager_ms = timed(lambda: scene.render_jobs(small, use_graphs=False))
    small_seq_ms = 0.0
    for j in small:
        (_, _), ms = timed(lambda j=j: scene.render(j["width"], j["height"], j["spp"], j["max_depth"], j["eye"],
                                                     j["look"], fov=j["fov"], seed=j["seed"], film=False))
        small_seq_ms += ms
    small_out = {
        "workload": "16 jobs " + ", ".join(f"{j['width']}x{j['height']}x{j['spp']}" for j in small) + ", max depth 8",
        "batch_ms": small_batch_ms, "batch_no_graphs_ms": small_eager_ms, "sequential_ms": small_seq_ms,
        "speedup_vs_sequential": small_seq_ms / small_batch_ms, "rays": st_small["rays"], "paths": st_small["paths"],
        "batch_mrays_s": st_small["rays"] / small_batch_ms / 1e3}
    print(f"[bench] small jobs: batch {small_batch_ms:.2f} ms, eager batch {small_eager_ms:.2f} ms, one after "
          f"another {small_seq_ms:.2f} ms", file=sys.stderr, flush=True)
    j0 = jobs[0]
    f_host, _ = scene.render(j0["width"], j0["height"], 4, j0["max_depth"], j0["eye"], j0["look"], fov=j0["fov"])
    f_dev, _ = scene_dev.render(j0["width"], j0["height"], 4, j0["max_depth"], j0["eye"], j0["look"], fov=j0["fov"])
    builds = {"scene_s": build_s, "mesh_bvh_s": scene.build_seconds(), "device_scene_s": dev_build_s,
              "device_mesh_bvh_s": scene_dev.build_seconds(), "same_film": bool(np.array_equal(f_host, f_dev))}
    del scene_dev
    print(f"[bench] jobs scene mesh BVHs: host (16 threads) {builds['mesh_bvh_s'] * 1e3:.0f} ms, device "
          f"{builds['device_mesh_bvh_s'] * 1e3:.1f} ms, same film {builds['same_film']}", file=sys.stderr, flush=True)
    print(f"[bench] jobs: batch {batch_ms:.1f} ms ({st['rays'] / batch_ms / 1e3:.0f} Mrays/s), eager batch "
          f"{eager_ms:.1f} ms, one after another {seq_ms:.1f} ms", file=sys.stderr, flush=True)
    return {
        "workload": f"{len(desc.meshes)} mesh instances ({sum(len(f) for _, f in desc.meshes)} triangles), "
                    f"{sum(1 for o in desc.objects if o[0] == 4)} CSG objects; 8 jobs "
                    + ", ".join(f"{j['width']}x{j['height']}x{j['spp']}" for j in jobs) + ", max depth 8",
        "scene_gen_s": gen_s, "scene_build_s": build_s, "scene_builds": builds,
        "tuned_scene_launch": scene_launch, "batch_default_launch_ms": default_ms,
        "batch_ms": batch_ms, "batch_mrays_s": st["rays"] / batch_ms / 1e3, "batch_samples_s": paths / batch_ms * 1e3,
        "batch_no_graphs_ms": eager_ms, "sequential_ms": seq_ms, "sequential_mrays_s": seq_rays / seq_ms / 1e3,
        "speedup_vs_sequential": seq_ms / batch_ms, "passes": st["passes"], "rays": st["rays"], "paths": paths,
        "small_jobs": small_out,
    }


def render_measure(pa, ctx):
    """Path-tracing throughput of the wavefront renderer on the two image configurations: the small
    scene (5,120-triangle icosphere, floor, 64 spheres, point light; 256^2, 16 spp, depth 8) and the
    room (walls, emissive ceiling panel as an area light, two boxes, an 81,920-triangle icosphere;
    512^2, 64 spp, depth 8), each with the reference's estimator, with MIS and with Russian
    roulette. GPU time from the render's own progress events (first launch to the last pass), the
    fastest of three renders."""
    from pbrt_amd import scenes
    out = {}
    for name, (desc, cam), (w, h, spp) in [("small_256x256_16spp", scenes.small(), (256, 256, 16)),
                                             ("room_512x512_64spp", scenes.room(), (512, 512, 64))]:
        scene = pa.Scene(ctx, desc)
        scene.render(64, 64, 1, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"])  # warm-up
        row = {}
        for key, opt in (("reference_estimator", {}), ("mis", dict(mis=True)), ("roulette", dict(roulette=True))):
            ms = None
            for _ in range(3):  # the fastest of three renders
                _, st, pr = scene.render_ex(w, h, spp, cam["max_depth"], cam["eye"], cam["look"], fov=cam["fov"],
                                            seed=1, **opt)
                ms = float(pr["ms"][-1]) if ms is None else min(ms, float(pr["ms"][-1]))
            row[key] = {
                "ms": ms, "mrays_s": st["rays"] / ms / 1e3, "samples_s": w * h * spp / ms * 1e3,
                "rays_per_sample": st["rays"] / (w * h * spp), "passes": st["passes"]}
        out[name] = row
        print(f"[bench] render {name}: {row['reference_estimator']['mrays_s']:.0f} Mrays/s, "
              f"{row['reference_estimator']['samples_s'] / 1e6:.1f} M samples/s (MIS "
              f"{row['mis']['mrays_s']:.0f} Mrays/s; roulette {row['roulette']['samples_s'] / 1e6:.1f} M samples/s)",
              file=sys.stderr, flush=True)
    return out


def build_measure(pa, ctx, v, f, host_mesh, host_s, runs=5):
    """The mesh BVH built on the device (pbrt_mesh_create_device) against the host build of the same
    mesh: wall time of each call (vertex and index upload included in both), and whether the device
    arrays equal the host's byte for byte."""
    pa.TriangleMesh(ctx, v, f, device_build=True, **pa.DEVICE_BVH)  # warm-up (code objects, allocator)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        dm = pa.TriangleMesh(ctx, v, f, device_build=True, **pa.DEVICE_BVH)
        times.append(time.perf_counter() - t0)
    hn, ht, hb, hr = host_mesh.layout()
    dn, dt, db, dr = dm.layout()
    same = bool(hr == dr and np.array_equal(hb, db) and np.array_equal(hn, dn) and np.array_equal(ht, dt))
    dev_ms = float(np.median(times)) * 1e3
    out = {"triangles": len(f), "host_ms": round(host_s * 1e3, 1), "device_ms": round(dev_ms, 2),
           "speedup": round(host_s * 1e3 / dev_ms, 1), "device_mtris_s": round(len(f) / dev_ms / 1e3, 1),
           "identical_arrays": same}
    print(f"[bench] BVH build: host {out['host_ms']} ms, device {out['device_ms']} ms ({out['speedup']}x), "
          f"identical arrays: {same}", file=sys.stderr, flush=True)
    return out


def render_profile(pa, ctx, spp=16, frames_per_pass=8):
    """A short run for kernel traces of the renderer on the radiance-cache workload: `spp` samples
    per pixel of plain path tracing (4 per pass) and `spp` frames of the cache render."""
    from pbrt_amd import scenes
    desc, cam = scenes.nrc_room()
    scene = pa.Scene(ctx, desc)
    W, H, D = cam["width"], cam["height"], cam["max_depth"]
    view = dict(fov=cam["fov"])
    _, st, pr = scene.render_ex(W, H, spp, D, cam["eye"], cam["look"], seed=1, samples_per_pass=4, **view)
    plain = dict(ms=float(pr["ms"][-1]), rays=st["rays"])
    cache = pa.RadianceCache(ctx, seed=4)
    _, st, pr = scene.render_ex(W, H, spp, D, cam["eye"], cam["look"], seed=1, cache=cache,
                                samples_per_pass=frames_per_pass, **view)
    return {"plain": plain, "cache": dict(ms=float(pr["ms"][-1]), rays=st["rays"], frames=spp,
                                          queries=st["cache_queries"])}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--waves", type=int, default=0)
    ap.add_argument("--lds-stack", type=int, default=0)
    ap.add_argument("--refill", type=int, default=0)
    ap.add_argument("--batch-launches", default="",
                    help="fixed launch per batch, 'block,waves,lds_stack,refill[,width[,packet]]' for the coherent then the "
                         "incoherent batch, separated by ';' (no autotuning)")
    ap.add_argument("--side", type=int, default=2048)
    ap.add_argument("--calibrate", action="store_true", help="also run a 1 GiB plain copy (counter calibration)")
    ap.add_argument("--no-autotune", action="store_true", help="keep the library's default launch shape")
    ap.add_argument("--sequential", action="store_true",
                    help="trace the step's two batches one after the other on one stream instead of at once")
    ap.add_argument("--no-tiles", action="store_true",
                    help="take the coherent (image) batch in row order instead of 8 x 16 tiles")
    ap.add_argument("--autotune", action="store_true", help="autotune each batch's launch instead of the fixed picks")
    ap.add_argument("--no-nrc", action="store_true", help="skip the radiance-cache time-to-RMSE leg")
    ap.add_argument("--nrc-only", action="store_true", help="run only the radiance-cache leg (prints its JSON)")
    ap.add_argument("--mlp-only", action="store_true",
                    help="run only the radiance cache's MLP throughput (inference and training step; prints its JSON)")
    ap.add_argument("--nrc-variants", default="sobol,sobol_mis",
                    help="radiance-cache races, comma-separated, the first the headline: " + ", ".join(NRC_VARIANTS))
    ap.add_argument("--no-jobs", action="store_true", help="skip the batched-jobs leg")
    ap.add_argument("--no-render", action="store_true", help="skip the path-tracing throughput leg")
    ap.add_argument("--render-only", action="store_true", help="run only the path-tracing throughput leg")
    ap.add_argument("--no-build", action="store_true", help="skip the device BVH build leg")
    ap.add_argument("--build-only", action="store_true", help="run only the device BVH build leg (prints its JSON)")
    ap.add_argument("--jobs-only", action="store_true", help="run only the batched-jobs leg (prints its JSON)")
    ap.add_argument("--render-profile", action="store_true",
                    help="only a short plain + cache render of the cache workload, for kernel traces")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: there is no multi-GPU path (rays and render jobs would shard without "
              "exchange; not built)", file=sys.stderr)
        sys.exit(2)

    import pbrt_amd as pa

    if args.nrc_only:
        print(json.dumps(nrc_measure(pa, pa.Context(0), variants=tuple(args.nrc_variants.split(",")))))
        return
    if args.mlp_only:
        ctx = pa.Context(0)
        print(json.dumps(mlp_measure(pa, ctx, pa.RadianceCache(ctx, seed=4))))
        return
    if args.jobs_only:
        print(json.dumps(jobs_measure(pa, pa.Context(0))))
        return
    if args.render_only:
        print(json.dumps(render_measure(pa, pa.Context(0))))
        return
    if args.render_profile:
        print(json.dumps(render_profile(pa, pa.Context(0))))
        return

    t0 = time.time()
    v, f, coherent, incoherent = make_workload(pa, side=args.side)
    gen_s = time.time() - t0
    print(f"[bench] workload generated in {gen_s:.1f}s: {len(f)} triangles, {len(coherent)}+{len(incoherent)} rays",
          file=sys.stderr, flush=True)

    ctx = pa.Context(0)
    t0 = time.time()
    mesh = pa.TriangleMesh(ctx, v, f, **pa.DEVICE_BVH)
    build_s = time.time() - t0
    if args.block or args.waves or args.lds_stack or args.refill:
        mesh.set_launch(args.block, args.waves, args.lds_stack, args.refill)
    bs = mesh.build_stats()
    node_bytes, tri_bytes = mesh.device_bytes()
    print(f"[bench] BVH built in {build_s:.1f}s: {bs}", file=sys.stderr, flush=True)
    if args.build_only:
        print(json.dumps(build_measure(pa, ctx, v, f, mesh, build_s)))
        return

    batches = {"coherent": coherent, "incoherent": incoherent}
    all_rays = ctx.upload(np.concatenate([coherent, incoherent]))
    all_hits = ctx.alloc(16 * (len(coherent) + len(incoherent)))
    dev = {}
    off = 0
    for k, r in batches.items():
        n = len(r)
        dev[k] = (all_rays.view(32 * off, 32 * n), all_hits.view(16 * off, 16 * n), n)
        off += n

    # the coherent batch is a side x side image: its kernels take it in 8 x 16 tiles
    # (pbrt_mesh_set_ray_tiles); the incoherent batch in batch order
    tiles = {"coherent": 0 if args.no_tiles or args.side % 16 else args.side, "incoherent": 0}

    # launch shape per batch: the autotuner's choice on that batch's own rays, or the fixed one
    tuned = None
    launches = {k: mesh.get_launch() for k in dev}
    fixed = args.batch_launches or (None if (args.autotune or args.no_autotune or args.block or args.waves
                                             or args.lds_stack or args.refill) else DEFAULT_BATCH_LAUNCHES)
    if fixed:
        for k, spec in zip(dev, fixed.split(";")):
            vals = [int(x) for x in spec.split(",")]
            launches[k] = dict(LAUNCH_DEFAULTS, **dict(zip(LAUNCH_KEYS, vals)))
    elif not (args.block or args.waves or args.lds_stack or args.refill or args.no_autotune):
        tuned = {}
        for k, (dr, dh, n) in dev.items():
            mesh.set_ray_tiles(tiles[k])
            tuned[k] = mesh.autotune(dr, n, dh, any_hit=False, reps=2)
            launches[k] = {x: tuned[k][x] for x in LAUNCH_KEYS}
        print(f"[bench] autotuned launches: {tuned}", file=sys.stderr, flush=True)

    # exact per-ray work on the same rays, untimed: the records the timed walk fetches (its tree
    # width, pbrt_mesh_trace_work) give the algorithmic bytes; the reference's instrumented counters
    # on the binary tree are reported beside them, and its hits are kept to check the timed launches
    work, binary_hits = {}, {}
    for k, (dr, dh, n) in dev.items():
        ds, dw = ctx.alloc(12 * n), ctx.alloc(8 * n)
        mesh.set_launch(block=128, waves_per_cu=16, lds_stack=16, width=2)
        mesh.set_ray_tiles(0)
        mesh.query_instrumented(dr, n, dh, ds)
        st = ds.read(np.uint32, 3 * n).reshape(n, 3).astype(np.int64)
        binary_hits[k] = dh.read(pa.HIT_DTYPE, n)
        interior = (st[:, 1] - 1) // 2
        bytes_binary = int(n * (32 + 16) + 64 * interior.sum() + 48 * st[:, 2].sum())
        mesh.set_launch(**launches[k])
        mesh.set_ray_tiles(tiles[k])
        mesh.query_work(dr, n, dh, dw)
        wk = dw.read(np.uint32, 2 * n).reshape(n, 2).astype(np.int64)
        node_bytes_rec = 128 if launches[k]["width"] == 4 else 64
        bytes_alg = int(n * (32 + 16) + node_bytes_rec * wk[:, 0].sum() + 48 * wk[:, 1].sum())
        work[k] = dict(n=n, hit_rate=float((binary_hits[k]["prim"] >= 0).mean()),
                       node_records_per_ray=float(wk[:, 0].mean()), node_record_bytes=node_bytes_rec,
                       tri_records_per_ray=float(wk[:, 1].mean()),
                       bytes_per_launch=bytes_alg, bytes_per_ray=bytes_alg / n,
                       binary_tree={"nodes_per_ray": float(st[:, 0].mean()),
                                    "box_tests_per_ray": float(st[:, 1].mean()),
                                    "tri_tests_per_ray": float(st[:, 2].mean()),
                                    "bytes_per_launch": bytes_binary})
        ds.free()
        dw.free()
    if mesh.status() != 0:
        raise RuntimeError("traversal kernel reported a stack overflow")

    stream = ctx.stream
    # The two batches are independent: a step traces them at once, the coherent batch on the
    # context's stream and the incoherent one on a second stream (the mesh gives each stream its own
    # spill area), forked from and joined back into the context's stream. --sequential traces them
    # one after the other on one stream. One event pair per launch and per step, so the timed loop
    # never waits on the host.
    side = ctx.new_stream()
    concurrent = not args.sequential
    on = {k: (side.h if concurrent and j == 1 else stream) for j, k in enumerate(dev)}
    evs = [[(ctx.event(), ctx.event()) for _ in dev] for _ in range(args.steps)]
    spans = [(ctx.event(), ctx.event(), ctx.event()) for _ in range(args.steps)]

    def step(s=None):
        fork, joined, end = spans[s] if s is not None else spans[0]
        fork.record(stream)
        if concurrent:
            fork.wait(ctx, side.h)
        for j, (k, (dr, dh, n)) in enumerate(dev.items()):
            mesh.set_launch(**launches[k])
            mesh.set_ray_tiles(tiles[k])
            if s is not None:
                evs[s][j][0].record(on[k])
            mesh.query(dr, n, dh, stream=on[k])
            if s is not None:
                evs[s][j][1].record(on[k])
        if concurrent:
            joined.record(side.h)
            joined.wait(ctx, stream)
        end.record(stream)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t0 = time.perf_counter()
    for s in range(args.steps):
        step(s)
    ctx.synchronize()
    wall = time.perf_counter() - t0
    per_batch_ms = {k: [evs[s][j][0].elapsed_ms(evs[s][j][1]) for s in range(args.steps)]
                    for j, k in enumerate(dev)}
    span_ms = float(np.mean([spans[s][0].elapsed_ms(spans[s][2]) for s in range(args.steps)]))
    # each batch alone on the context's stream, after the timed loop: its own kernel time
    alone_ms = {}
    for k, (dr, dh, n) in dev.items():
        mesh.set_launch(**launches[k])
        mesh.set_ray_tiles(tiles[k])
        e0, e1 = evs[0][0]
        ts = []
        for _ in range(5):
            e0.record(stream)
            mesh.query(dr, n, dh, stream=stream)
            e1.record(stream)
            ts.append(e0.elapsed_ms(e1))
        alone_ms[k] = float(np.mean(ts))
    # the timed launches' own results: no stack overflow, and the hits of the reference's
    # instrumented walk on the binary tree, byte for byte
    status_after = mesh.status()
    if status_after != 0:
        raise RuntimeError("traversal kernel reported a stack overflow in the timed loop")
    timed_hits = {k: dh.read(pa.HIT_DTYPE, n) for k, (dr, dh, n) in dev.items()}
    timed_vs_binary = {k: int((timed_hits[k].view(np.uint8).reshape(-1, 16)
                               != binary_hits[k].view(np.uint8).reshape(-1, 16)).any(1).sum()) for k in dev}

    total_rays = sum(n for (_, _, n) in dev.values())
    kern_ms = {k: float(np.mean(x)) for k, x in per_batch_ms.items()}
    # the time the step's traversal launches occupy the GPU: fork to join when they run at once,
    # the sum of the two launches when they run one after the other
    step_kernel_ms = span_ms if concurrent else sum(kern_ms.values())
    ms_per_step = wall * 1e3 / args.steps
    value = total_rays / (ms_per_step * 1e-3) / 1e6
    per_batch = {k: dict(mrays_s=work[k]["n"] / (alone_ms[k] * 1e-3) / 1e6, kernel_ms=alone_ms[k],
                         kernel_ms_in_step=kern_ms[k], stream="second" if on[k] != stream else "context", **work[k])
                 for k in dev}
    # roofline over both launches of the dominant kernel: algorithmic bytes / the time they occupy
    alg_bytes = sum(work[k]["bytes_per_launch"] for k in dev)
    achieved = alg_bytes / (step_kernel_ms * 1e-3) / 1e9
    measured = [MEASURED_TRAFFIC.get((k,) + tuple(launches[k][x] for x in ("block", "lds_stack", "refill", "width",
                                                                          "packet")) + (tiles[k],))
                for k in dev]
    measured = None if (args.side != 2048 or None in measured) else measured

    # any-hit throughput on the same batches (informational)
    anyhit, occlusion = {}, {}
    for k, (dr, dh, n) in dev.items():
        occ = ctx.alloc(n)
        mesh.set_launch(**launches[k])
        mesh.set_ray_tiles(tiles[k])
        mesh.intersect_p(dr, n, occ)
        e0, e1 = evs[0][0]
        e0.record(stream)
        for _ in range(5):
            mesh.intersect_p(dr, n, occ)
        e1.record(stream)
        anyhit[k] = n * 5 / (e0.elapsed_ms(e1) * 1e-3) / 1e6
        occlusion[k] = occ.read(np.uint8, n)
        occ.free()
    if mesh.status() != 0:
        raise RuntimeError("any-hit kernel reported a stack overflow")

    copy_gbs = None
    if args.calibrate:
        # a plain copy of 1 GiB: the known-size launch that hardware byte counters are scaled by
        nbytes = 1 << 30
        src, dst = ctx.alloc(nbytes), ctx.alloc(nbytes)
        src.fill(1)
        e0, e1 = evs[0][0]
        e0.record(stream)
        pa._native.call("pbrt_dev_copy", ctx.h, dst.ptr, src.ptr, nbytes, stream)
        e1.record(stream)
        copy_gbs = 2 * nbytes / (e0.elapsed_ms(e1) * 1e-3) / 1e9
        src.free()
        dst.free()

    nrc = None
    if not args.no_nrc:
        nrc = nrc_measure(pa, ctx, variants=tuple(args.nrc_variants.split(",")))
    jobs = None
    if not args.no_jobs:
        jobs = jobs_measure(pa, ctx)
    render = None
    if not args.no_render:
        render = render_measure(pa, ctx)

    build = None
    if not args.no_build:
        build = build_measure(pa, ctx, v, f, mesh, build_s)

    cpu, parity = None, None
    if not args.no_cpu:
        threads = min(16, len(os.sched_getaffinity(0)))
        cpu, parity = cpu_baseline(pa, v, f, coherent, incoherent, args.side, threads, timed_hits, occlusion)
        print(f"[bench] cpu baseline: {cpu['value']:.2f} Mrays/s (Oracle A), {cpu['oracle_b_mrays_s']:.2f} "
              f"(Oracle B) on {threads} threads; host {cpu['host']}", file=sys.stderr, flush=True)
        print(f"[bench] parity: {parity}", file=sys.stderr, flush=True)
    parity_line = {"timed_hits_vs_binary_instrumented_walk": timed_vs_binary, "status_after_timed_loop": status_after}
    if parity:
        parity_line.update({k: {"mismatch_vs_oracle_b": p["mismatch_vs_oracle_b"]["total"],
                                "anyhit_mismatch_vs_oracle_b": p["anyhit_mismatch_vs_oracle_b"],
                                "mismatch_rate_vs_oracle_a": p["vs_oracle_a"]["rate"],
                                "vs_oracle_a": p["vs_oracle_a"],
                                "max_rel_t_err_fp64": p["t_vs_fp64"]["max_rel"],
                                "max_rel_t_err_fp64_non_grazing": p["t_vs_fp64"]["max_rel_non_grazing"],
                                "t_err_above_1e-5": p["t_vs_fp64"]["above_1e-5"]} for k, p in parity.items()})

    out = {
        "metric": "closest-hit Mrays/s (1.3M tris); time-to-RMSE 1e-2, radiance cache vs plain PT",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {
            "workload": "closest-hit traversal, 1,310,720-triangle displaced icosphere (8 subdivisions, 5-octave "
                        "value noise, 10% amplitude), binned SAH 12 bins / leaves <= 4; batches of 4,194,304 "
                        "coherent (2048^2 pinhole) + 4,194,304 incoherent rays per step",
            "triangles": len(f),
            "rays_per_step": total_rays,
            "bvh": dict(bs, node_bytes=node_bytes, tri_bytes=tri_bytes),
            "parallelism": "single GPU; the step's two batches on two streams at once" if not args.sequential
                           else "single GPU; the step's two batches one after the other",
            "launch": launches,
            "autotuned": tuned is not None,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": sum(m["bytes"] for m in measured) / len(measured) if measured else None,
            "kernel": ", ".join(f"{'k_closest_packet' if launches[k]['packet'] else 'k_closest4' if launches[k]['width'] == 4 else 'k_closest'} ({k}{', 8x16 tiles' if tiles[k] else ''})" for k in dev),
            "kernel_time": ("fork-to-join span of the two launches running at once, HIP events on the context's "
                            "stream" if concurrent else "sum of the two launches' HIP event times"),
            "kernel_time_ms": round(step_kernel_ms, 4),
            "frac_alone": round(alg_bytes / (sum(alone_ms.values()) * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "mrays_s_kernels_one_after_another": round(total_rays / (sum(alone_ms.values()) * 1e-3) / 1e6, 1),
            "algorithmic_bytes_per_launch": alg_bytes / len(dev),
            "algorithmic_bytes_rule": "per ray 32 (ray) + 16 (hit) + node records fetched x record size (128 B "
                                      "4-wide, 64 B binary) + triangle records fetched x 48 B, counted by "
                                      "pbrt_mesh_trace_work on the same rays and tree",
            "frac_binary_tree_bytes": round(sum(work[k]["binary_tree"]["bytes_per_launch"] for k in dev)
                                            / (step_kernel_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
            "traffic_source": measured[0]["source"] if measured else None,
        },
        "parity": parity_line,
        "per_batch": per_batch,
        "anyhit_mrays_s": anyhit,
        "copy_gbs": copy_gbs,
        "radiance_cache": nrc,
        "render_jobs": jobs,
        "render": render,
        "bvh_build": build,
        "cpu_baseline": cpu,
        "gpu_vs_cpu": round(value / cpu["value"], 1) if cpu else None,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
