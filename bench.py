"""Benchmark: textured mesh, 1920x1080 RGBA8 4x MSAA + resolve, 522,242 indexed triangles,
2048^2 mip-chained noise texture with 16x anisotropic filtering, Lambert lighting, depth LESS,
camera at (0, 1.5, -1.5) looking at the origin (a secondary line repeats it with a close camera
that fills ~80% of the frame).

One step is one full frame through the renderer API on one MI355X: clear colour, clear
depth/stencil, draw_index, MSAA resolve, flush.  Inputs (vertex/index buffers, texture) are
resident in HBM before timing starts.  Prints one JSON line.

  python bench.py [--gpus 1] [--steps 50] [--warmup 5] [--no-cpu-baseline] [--cpu-seconds 15]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

from salviarenderer_amd import api as A  # noqa: E402
from salviarenderer_amd import hip_lib, scenes  # noqa: E402

W, H, SAMPLES = 1920, 1080, 4
HBM_PEAK_GBS = 8000.0  # MI355X HBM3E


def texture_chain_bytes(size=2048):
    total, s = 0, size
    while s >= 1:
        total += s * s * 4
        s //= 2
    return total


# Algorithmic bytes of one launch of the dominant kernel, vis_kernel<2,4,true> (visibility + shading of
# the draw, with the frame's clears fused into it): per sample the RG32F depth/stencil written once
# (8 B) and the RGBA8 colour written once (4 B); the 2048^2 texture chain read at most once.
RASTER_BYTES = W * H * SAMPLES * (8 + 4) + texture_chain_bytes(2048)
RASTER_KERNEL = "vis_kernel<2, 4, true>"
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_vis_kernel_textured_mesh_1080p_msaa4.json")
ISSUE_FILE = os.path.join(ROOT, "profiles", "pmc_vis_kernel_textured_mesh_1080p_msaa4.json")


def measured_traffic():
    """HBM bytes per vis_kernel launch from the committed FETCH_SIZE/WRITE_SIZE counter passes,
    scaled by the copy-kernel calibration recorded with them; None if absent."""
    if not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as fh:
        t = json.load(fh)
    return float(t["traffic_bytes"])


def measured_issue():
    """VALU and LDS issue of one vis_kernel launch from the committed SQ counter passes; None if absent."""
    if not os.path.exists(ISSUE_FILE):
        return None
    with open(ISSUE_FILE) as fh:
        t = json.load(fh)
    return {"limiter": t["limiter"], "valu_issue_frac": t["valu_issue_frac"], "wait_frac": t["wait_frac"],
            "salu_per_valu": t["salu_per_valu"],
            "lds_bank_conflict_cycles_per_lds_inst": t["lds_bank_conflict_cycles_per_lds_inst"],
            "source": os.path.relpath(ISSUE_FILE, ROOT)}


def copy_calibration(mib, reps):
    """Copies a mib-MiB RGBA32F surface reps times with the library's 1-sample resolve kernel (a
    plain copy: every byte read once and written once), for counter calibration."""
    lib = hip_lib()
    r = lib.renderer(A.RENDERER_BENCHMARK)
    side = int((mib * 1024 * 1024 // 16) ** 0.5)
    src = r.create_tex2d(side, side, 1, A.PF_RGBA32F)
    dst = r.create_tex2d(side, side, 1, A.PF_RGBA32F)
    s, d = src.surface(0), dst.surface(0)
    r.clear_color(s, (0.25, 0.5, 0.75, 1.0))
    for _ in range(reps):
        r.resolve(s, d)
    r.flush()
    print(json.dumps({"copy_bytes_read_per_launch": side * side * 16, "copy_bytes_written_per_launch": side * side * 16,
                      "launches": reps}), flush=True)


def run_frames(lib, steps, warmup, profile=False, camera="oblique", sasl=False):
    """Times `steps` frames with no query active (the renderer as an application runs it); with
    `profile`, the same number of frames is then rendered again under the pipeline-profiles query,
    whose per-stage HIP events (recorded on the draw's stream between its kernels) give the stage
    and kernel times without adding their own gaps to the timed frames."""
    r = lib.renderer(A.RENDERER_BENCHMARK)
    f = scenes.setup_mesh(r, W, H, SAMPLES, camera=camera, sasl=sasl)
    for _ in range(warmup):
        f.render()
        r.flush()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t = time.perf_counter()
        f.render()
        r.flush()
        lat.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    prof = None
    if profile:
        q = r.create_query(A.QUERY_PIPELINE_PROFILES)
        r.begin(q)
        for _ in range(steps):
            f.render()
            r.flush()
        r.end(q)
        st, data = r.get_data(q)
        prof = list(data)
    return total, np.array(lat), prof, f.n_tris


def multipass_frames(lib, steps, warmup):
    """Secondary workload (not the headline): shadow map + PCF-shaded, translucent and stencil-mirror
    passes over 200,000 random triangles at 1920x1080 4xMSAA (scenes.MultipassFrame)."""
    r = lib.renderer(A.RENDERER_BENCHMARK)
    f = scenes.MultipassFrame(r)
    for _ in range(warmup):
        f.render()
        r.flush()
    t0 = time.perf_counter()
    for _ in range(steps):
        f.render()
        r.flush()
    total = time.perf_counter() - t0
    q = r.create_query(A.QUERY_PIPELINE_PROFILES)  # stage times from a second, profiled pass
    r.begin(q)
    for _ in range(steps):
        f.render()
        r.flush()
    r.end(q)
    prof = list(r.get_data(q)[1])
    return {"workload": "multipass 1920x1080 4xMSAA: 2048^2 shadow map of all 200k triangles, 140k opaque "
                        "PCF-shaded + 60k translucent triangles blended back to front, stencil mirror mark and "
                        "mirrored redraw (5 draws)",
            "ms_per_frame": total / steps * 1e3, "frames_per_s": steps / total, "frames": steps,
            "stage_ms_per_frame": {"setup_clip": prof[2] / steps / 1e6, "bin_sort": prof[5] / steps / 1e6,
                                   "raster": prof[6] / steps / 1e6}}


def batched_jobs(lib, launches, plain_passes=1, autotune_reps=1, config=None):
    """Secondary workload: 256 independent jobs (scenes.batch_job_specs: 256^2..1920x1080, 1k-1M
    triangles, plain-scene shaders, resident noise textures), issued one by one through the
    renderer API and as one recorded batch (one launch per stage, HIP graph), both timed on the
    wall clock around complete passes."""
    r = lib.renderer(A.RENDERER_BENCHMARK)
    specs = scenes.batch_job_specs()
    t = time.perf_counter()
    jobs = scenes.BatchedJobs(r, specs)
    r.flush()
    setup_s = time.perf_counter() - t
    print(f"batched: {len(specs)} jobs resident ({setup_s:.1f} s)", file=sys.stderr, flush=True)
    jobs.render_all()  # warm-up
    r.flush()
    t0 = time.perf_counter()
    for _ in range(plain_passes):
        jobs.render_all()
        r.flush()
    plain_s = (time.perf_counter() - t0) / plain_passes
    print(f"batched: one-by-one pass {plain_s * 1e3:.1f} ms", file=sys.stderr, flush=True)
    b = jobs.record()
    if config:
        b.configure(*config)
        sweep = []
    else:
        sweep = b.autotune(reps=autotune_reps)
    info = b.info()
    b.launch()
    r.flush()
    t0 = time.perf_counter()
    for _ in range(launches):
        b.launch()
    r.flush()
    batch_s = (time.perf_counter() - t0) / launches
    status = b.status()
    print(f"batched: batch launch {batch_s * 1e3:.1f} ms, status {status}", file=sys.stderr, flush=True)
    n = len(specs)
    return {"workload": f"{n} independent jobs: 256^2/512^2/1024^2/1920x1080 RGBA8 + depth, 1k-1M triangles "
                        "(37.4M total), pass-through VS, Gouraud PS, depth LESS, 1 sample",
            "frames_per_s": n / batch_s, "batch_ms": batch_s * 1e3, "launches": launches,
            "one_by_one_frames_per_s": n / plain_s, "one_by_one_ms": plain_s * 1e3,
            "speedup": plain_s / batch_s, "status": status,
            "mtri_per_s": info.primitives / batch_s / 1e6,
            "tuning": {"tile": info.tile, "bin_items": info.bin_items, "vis_threads": info.vis_threads},
            "sweep_ms": [[tl, bi, th, round(ms, 3)] for tl, bi, th, ms in sweep],
            "device_bytes": info.device_bytes, "records": info.records}


def serving(lib, n_requests=10000, rates=(0.25, 0.5, 0.75, 1.0), seed=0x5A11A005):
    """Secondary workload: Poisson streams of render requests drawn from the batched jobs (resident
    meshes and textures) with 1x or 4x MSAA at random, through the request server (continuous
    batching); first all at once (capacity), then at fractions of that capacity."""
    from concurrent.futures import ThreadPoolExecutor

    r = lib.renderer(A.RENDERER_BENCHMARK)
    desc = A.default_server_desc(lib)
    desc.staging_bytes = 4 << 30
    srv = r.create_server(desc)
    specs = scenes.batch_job_specs()
    t = time.perf_counter()
    with ThreadPoolExecutor(8) as ex:
        for sp, geo, tex in zip(specs, ex.map(scenes.batch_triangles, specs),
                                ex.map(lambda sp: scenes.noise_texture(sp.tex_size, sp.seed), specs)):
            srv.put_mesh(sp.seed, geo)
            srv.put_texture(sp.seed, tex)
    print(f"serving: {len(specs)} meshes + textures resident ({time.perf_counter() - t:.1f} s)", file=sys.stderr, flush=True)
    rng = np.random.default_rng(seed)

    def run(rate):
        jobs = rng.integers(0, len(specs), n_requests)
        samples = rng.choice([1, 4], n_requests)
        gaps = rng.exponential(1e9 / rate, n_requests) if rate else np.zeros(n_requests)
        start = srv.now() + 300_000_000
        arrivals = start + np.cumsum(gaps).astype(np.int64)
        tickets = {}
        for j, s_, a in zip(jobs, samples, arrivals):
            sp = specs[int(j)]
            tickets[srv.submit(sp.seed, sp.n_tris, sp.width, sp.height, int(s_), texture_key=sp.seed,
                               arrival_ns=int(a))] = (int(j), int(s_))
        lat, sizes, done_last, fails = [], [], 0, 0
        while len(lat) < n_requests:
            for c in srv.poll(256, wait_ms=50, copy=False):
                lat.append((c["done_ns"] - c["arrival_ns"]) / 1e6)
                sizes.append(c["batch_jobs"])
                done_last = max(done_last, c["done_ns"])
                fails += c["status"] != 0
                srv.release(c["ticket"])
        span = (done_last - int(arrivals[0])) / 1e9
        lat = np.array(lat)
        return {"offered_per_s": None if not rate else rate, "frames_per_s": n_requests / span,
                "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_batch": float(np.mean(sizes)), "failed": int(fails)}

    cap = run(None)
    print(f"serving: capacity {cap['frames_per_s']:.0f} frames/s", file=sys.stderr, flush=True)
    sweep = []
    for f in rates:
        res = run(f * cap["frames_per_s"])
        res["load"] = f
        sweep.append(res)
        print(f"serving: load {f} -> {res['frames_per_s']:.0f} frames/s p50 {res['p50_ms']:.2f} ms "
              f"p99 {res['p99_ms']:.2f} ms", file=sys.stderr, flush=True)
    # spot check: a few served images against the same jobs rendered one by one through the API
    small = sorted(range(len(specs)), key=lambda j: specs[j].width * specs[j].height * specs[j].n_tris)[:4]
    checked = mismatches = 0
    for j in small:
        for smp in (1, 4):
            sp = specs[j]
            t = srv.submit(sp.seed, sp.n_tris, sp.width, sp.height, smp, texture_key=sp.seed)
            got = None
            while got is None:
                for c in srv.poll(16, wait_ms=50):
                    if c["ticket"] == t:
                        got = c["color"]
            ref = scenes.BatchedJobs(lib.renderer(), [scenes.JobSpec(sp.seed, sp.width, sp.height, sp.n_tris,
                                                                     sp.tex_size, smp)], with_textures=False)
            ref.render_all()
            ref.r.flush()
            checked += 1
            mismatches += int(not np.array_equal(got, ref.output(0).ravel()))
    st = srv.stats()
    return {"workload": f"{n_requests} requests per rate from the 256 batched jobs, 1x/4x MSAA at random, "
                        "Poisson arrivals; meshes and textures resident (LRU pool cap 200 GiB)",
            "capacity": cap, "sweep": sweep, "checked_against_plain": checked, "mismatches": mismatches,
            "batches": st.batches, "retries": st.retries,
            "arena_peak_bytes": st.arena_peak_bytes, "pool_bytes": st.pool_bytes}


def close_camera(lib, steps, warmup):
    """Secondary: the same frame with the close camera (grid over ~80% of the frame)."""
    total, lat, prof, n_tris = run_frames(lib, steps, warmup, profile=True, camera="close")
    draws = max(prof[7], 1)
    return {"workload": "textured mesh as the headline, camera (0, 0.8, -0.6) -> (0, 0.5, 0): grid over ~80% "
                        "of the frame", "ms_per_frame": total / steps * 1e3, "mpix_per_s": W * H * steps / total / 1e6,
            "p99_ms": float(np.percentile(lat, 99) * 1e3), "raster_kernel_ms": prof[6] / draws / 1e6,
            "bin_sort_ms": prof[5] / draws / 1e6, "setup_clip_ms": prof[2] / draws / 1e6}


def sasl_frames(lib, steps, warmup):
    """Secondary: the headline frame shaded by the SASL Lambert shader (set_pixel_shader_code), which
    runs the quad interpreter."""
    total, lat, prof, n_tris = run_frames(lib, steps, warmup, profile=True, sasl=True)
    draws = max(prof[7], 1)
    return {"workload": "textured mesh as the headline, shaded by a SASL pixel shader (tex2D, normalize, saturate, "
                        "dot)", "ms_per_frame": total / steps * 1e3, "mpix_per_s": W * H * steps / total / 1e6,
            "raster_kernel_ms": prof[6] / draws / 1e6, "shade_kernel_ms": (prof[9] if len(prof) > 9 else 0) / draws / 1e6,
            "bin_sort_ms": prof[5] / draws / 1e6, "setup_clip_ms": prof[2] / draws / 1e6}


def cpu_baseline(seconds):
    oracle = os.path.join(ROOT, "oracle", "libsrx_cpu.so")
    if not os.path.exists(oracle):
        return None
    lib = A.Lib(oracle)
    threads = int(lib.dll.srxo_thread_count())
    host_cores = os.cpu_count()
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else host_cores
    r = lib.renderer(A.RENDERER_BENCHMARK)
    f = scenes.setup_mesh(r, W, H, SAMPLES)
    f.render()  # warm-up frame
    n, t0 = 0, time.perf_counter()
    while True:
        f.render()
        r.flush()
        n += 1
        el = time.perf_counter() - t0
        if el >= seconds or n >= 50:
            break
    return {"value": W * H * n / el / 1e6, "unit": "Mpix/s", "cores": threads, "kind": "port",
            "host_cores": host_cores, "affinity_cores": usable,
            "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "sample": f"{n} full frames (clear, draw_index, resolve) of this workload on the CPU oracle, a "
                      f"scalar C++ restatement of the reference's pipeline (its thread-pool structure, no SSE), "
                      f"{el:.1f} s on {threads} threads (the pool is capped by OMP_NUM_THREADS, this host "
                      f"share's CPU count; the machine reports {host_cores} cores, {usable} in this "
                      f"process's affinity mask)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-multipass", action="store_true", help="skip the secondary multipass measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the secondary batched-jobs measurement")
    ap.add_argument("--batched-only", action="store_true", help="only the batched-jobs measurement")
    ap.add_argument("--no-serving", action="store_true", help="skip the secondary serving measurement")
    ap.add_argument("--headline-only", action="store_true",
                    help="only the headline frames (no CPU baseline and no secondary workloads), for profiling")
    ap.add_argument("--serving-only", action="store_true", help="only the serving measurement")
    ap.add_argument("--sasl-only", action="store_true",
                    help="only the textured mesh shaded by the SASL pixel shader (steps/warmup as given), for profiling")
    ap.add_argument("--serving-requests", type=int, default=10000)
    ap.add_argument("--batch-config", default="", metavar="TILE,BIN_ITEMS,THREADS",
                    help="configure the batch instead of autotuning it")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run the copy kernel used to calibrate HBM byte counters, then exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration, 5)
        return
    if args.gpus != 1:
        print(f"bench.py: this pipeline runs on one GPU; --gpus {args.gpus} is not supported "
              "(multi-GPU sharding is design only)", file=sys.stderr)
        sys.exit(2)

    lib = hip_lib()
    config = tuple(int(x) for x in args.batch_config.split(",")) if args.batch_config else None
    if config is not None and len(config) != 3:
        ap.error("--batch-config takes TILE,BIN_ITEMS,THREADS")
    if args.batched_only:
        print(json.dumps(batched_jobs(lib, 3, config=config)), flush=True)
        return
    if args.serving_only:
        print(json.dumps(serving(lib, args.serving_requests)), flush=True)
        return
    if args.sasl_only:
        print(json.dumps(sasl_frames(lib, args.steps, args.warmup)), flush=True)
        return
    total, lat, prof, n_tris = run_frames(lib, args.steps, args.warmup, profile=True)
    fps = args.steps / total
    draws = max(prof[7], 1)
    # raster stage = the fused kernel (the separate shading slot only brackets an empty interval)
    vis_ns = prof[6] / draws
    achieved = RASTER_BYTES / (vis_ns * 1e-9) / 1e9
    out = {
        "metric": "shaded Mpix/s + Mtri/s (1080p 4xMSAA); batched frames/s; p99 latency",
        "value": W * H * fps / 1e6,
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "textured mesh 1920x1080 RGBA8 4xMSAA + resolve, 512x512-vertex grid "
                               "(522,242 indexed tris), 2048^2 RGBA8 noise texture, 16x anisotropic, Lambert, "
                               "camera (0, 1.5, -1.5) -> origin",
                   "width": W, "height": H, "samples": SAMPLES, "triangles": n_tris},
        "mtri_per_s": n_tris * fps / 1e6,
        "frames_per_s": fps,
        "p50_ms": float(np.percentile(lat, 50) * 1e3),
        "p99_ms": float(np.percentile(lat, 99) * 1e3),
        "stage_ms": {"vertex": prof[1] / draws / 1e6, "setup_clip": prof[2] / draws / 1e6,
                     "scan": prof[3] / draws / 1e6, "bin_sort": prof[5] / draws / 1e6,
                     "raster": vis_ns / 1e6},
        # the path has no MFMA work, so HBM is the roofline priced here; the kernel's actual limiter
        # (instruction issue or latency) is reported beside it from the committed counter passes
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(),
                     "kernel": RASTER_KERNEL, "algorithmic_bytes": RASTER_BYTES, "kernel_ms": vis_ns / 1e6,
                     "issue": measured_issue()},
        "cpu_baseline": None if (args.no_cpu_baseline or args.headline_only) else cpu_baseline(args.cpu_seconds),
        "secondary": {},
    }
    if args.headline_only:
        print(json.dumps(out), flush=True)
        return
    out["secondary"]["close_camera"] = close_camera(lib, 20, 3)
    out["secondary"]["sasl"] = sasl_frames(lib, 10, 2)
    if not args.no_multipass:
        out["secondary"]["multipass"] = multipass_frames(lib, 20, 3)
    if not args.no_batched:
        out["secondary"]["batched"] = batched_jobs(lib, 3, config=config)
    if not args.no_serving:
        out["secondary"]["serving"] = serving(lib, args.serving_requests)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
