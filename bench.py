#!/usr/bin/env python3
"""Benchmark of the MI355X pathtracer path.

Workload: displaced icosphere (20 * 4^8 = 1,310,720 triangles) on a plane with
a box light and a sphere light, Lambert/Phong, thin lens; 1920x1080, 16 spp,
max depth 8, FP32 traversal of an SAH kd-tree. A step renders one full frame
(every sample of every pixel) into a device-resident image; the scene is
resident in HBM before timing starts. Prints one JSON line.

value = executed closest-hit queries (Scene::intersection-equivalent calls)
per second of the timed steps, in millions. The roofline object covers the
dominant kernel (the trace kernel); cpu_baseline times the reference renderer
(or the oracle port when the reference checker is not built) on the host.
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRACE_BYTES_PER_QUERY = 48      # read ray (2 x float4) + write hit (float4)
PIPE_BYTES_PER_QUERY = 184      # ray/hit/path-state streams of trace + shade + compaction
PIPE_BYTES_PER_SAMPLE = 152     # camera ray + state write, FP64 accumulate read + write


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--subdiv", type=int, default=8)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--rdepth", type=int, default=8)
    ap.add_argument("--accel", choices=["kdtree", "octree"], default="kdtree")
    ap.add_argument("--spp-pass", type=int, default=0, help="samples per wavefront pass (0: auto)")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=1,
                    help="CPU baseline threads (the reference's glibc rand() lock makes more threads slower)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tuning", default="", help="launch shape as JSON (xt_tuning fields); skips the autotuner")
    return ap.parse_args()


def cpu_baseline(scene_path, args, threads):
    """Reference renderer (oracle/_ref) on a tile subset of the same frame;
    falls back to the oracle port. Returns the cpu_baseline object."""
    from oracle import oracle as O
    W, H = args.width, args.height
    tiles = ((W + 31) // 32) * ((H + 31) // 32)
    if O.ref_available():
        L = O.ref()
        L.xr_render_tiles.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 7 + [ctypes.c_uint, ctypes.c_int,
                                                                               ctypes.c_void_p, ctypes.c_void_p]
        t0 = time.perf_counter()
        r = O.RefScene(scene_path)
        load_s = time.perf_counter() - t0
        log(f"cpu baseline: reference scene loaded in {load_s:.1f}s")
        sec, q = ctypes.c_double(), ctypes.c_uint64()
        # calibrate on 8 tiles per thread spread over the frame (the cost per
        # tile varies widely: sky tiles are nearly free), then size the sample
        stride = max(1, tiles // (8 * threads))
        L.xr_render_tiles(r._h, 3, W, H, args.spp, 1, args.rdepth, threads, 1, stride, ctypes.byref(sec),
                          ctypes.byref(q))
        log(f"cpu baseline: calibration {(tiles + stride - 1) // stride} tiles on {threads} threads in {sec.value:.2f}s")
        n_cal = (tiles + stride - 1) // stride
        per_tile = sec.value / n_cal
        want = int(args.cpu_seconds * threads / max(per_tile, 1e-4))
        if want >= 2 * n_cal:
            stride2 = max(2, tiles // want)
            L.xr_render_tiles(r._h, 3, W, H, args.spp, 1, args.rdepth, threads, 1, stride2, ctypes.byref(sec),
                              ctypes.byref(q))
            stride = stride2
            log(f"cpu baseline: {(tiles + stride - 1) // stride} tiles in {sec.value:.2f}s")
        n_tiles = (tiles + stride - 1) // stride
        return {"value": q.value / sec.value / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "reference",
                "sample": f"reference pathtracer (compiled -O3 from its sources, OpenMP dynamic tiles, {threads} "
                          f"thread(s)) on every {stride}th 32x32 tile ({n_tiles} of {tiles}) at full spp; "
                          f"{q.value} Scene::intersection calls in {sec.value:.2f}s; scene load + octree build "
                          f"{load_s:.1f}s not timed"}
    import xtracer_amd as xa
    from xtracer_amd import scenes
    s = xa.Scene(scene_path, build=True)
    orc = O.OracleScene(s)
    cam = scenes.camera_dict(**scenes.ICOSPHERE_CAMERA)
    stride = max(1, tiles // threads)
    t0 = time.perf_counter()
    _, nq = orc.render(3, W, H, cam, samples=args.spp, rdepth=args.rdepth, threads=threads, tile_stride=stride)
    dt = time.perf_counter() - t0
    return {"value": nq / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"FP64 oracle port on every {stride}th tile at full spp; {nq} queries in {dt:.2f}s"}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would render its own "
              "tiles or jobs with a replicated scene); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch

    import xtracer_amd as xa
    from xtracer_amd import scenes

    t0 = time.perf_counter()
    path = scenes.icosphere_scene(seed=0, subdiv=args.subdiv)
    gen_s = time.perf_counter() - t0
    log(f"scene generated in {gen_s:.1f}s: {path}")
    t0 = time.perf_counter()
    scene = xa.Scene(path)
    build_s = time.perf_counter() - t0
    info = scene.info()
    log(f"scene loaded + trees built in {build_s:.1f}s: {info.n_prims} prims, {info.n_kd_nodes} kd nodes")
    dev = xa.Device(0)
    sid = dev.upload(scene)
    log("scene uploaded")
    accel = xa.ACCEL_KDTREE if args.accel == "kdtree" else xa.ACCEL_OCTREE
    out = torch.empty((args.height, args.width, 4), dtype=torch.float32, device="cuda")
    cam = scene.camera()

    def job(k):
        return xa.make_job(sid, xa.PATHTRACER, args.width, args.height, cam, samples=args.spp, rdepth=args.rdepth,
                           accel=accel, seed=1000 + k, samples_per_pass=args.spp_pass)

    if args.tuning:
        dev.set_tuning(**json.loads(args.tuning))
    tuning = dev.tuning()
    if not args.no_autotune and not args.tuning:
        from xtracer_amd.autotune import autotune
        t0 = time.perf_counter()
        tuning, _ = autotune(dev, job(999), out.data_ptr(), wave_paths=(4 << 20, 16 << 20), log=log)
        log(f"autotune picked {tuning} in {time.perf_counter() - t0:.1f}s")
    for k in range(args.warmup):
        t0 = time.perf_counter()
        st = dev.render_device(job(k), out.data_ptr())
        log(f"warmup {k}: {time.perf_counter() - t0:.3f}s, {st.queries} queries, {st.retests} retests")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats.append(dev.render_device(job(args.warmup + k), out.data_ptr()))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {wall:.3f}s")
    queries = sum(s.queries for s in stats)
    samples = sum(s.samples for s in stats)
    trace_ms = sum(s.ms_trace for s in stats)
    launches = sum(s.trace_launches for s in stats)
    dev_ms = sum(s.ms_total for s in stats)
    value = queries / wall / 1e6
    achieved = queries * TRACE_BYTES_PER_QUERY / (trace_ms * 1e-3) / 1e9 if trace_ms > 0 else 0.0
    pipe_bytes = queries * PIPE_BYTES_PER_QUERY + samples * PIPE_BYTES_PER_SAMPLE
    pipe = pipe_bytes / (wall) / 1e9
    threads = args.cpu_threads if args.cpu_threads > 0 else min(16, os.cpu_count() or 1)
    cpu = None if args.no_cpu_baseline else cpu_baseline(path, args, threads)
    line = {
        "metric": "Mrays/s at 1080p 16spp (pathtracer, plain one-job-per-launch)",
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded displaced icosphere scene, generated in-process)",
        "config": {"workload": f"pathtracer icosphere {info.n_prims} prims {args.width}x{args.height} "
                               f"{args.spp}spp rdepth {args.rdepth} {args.accel} thin-lens",
                   "width": args.width, "height": args.height, "spp": args.spp, "rdepth": args.rdepth,
                   "prims": info.n_prims, "accel": args.accel},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "k_trace (kd-tree)", "bytes_per_query": TRACE_BYTES_PER_QUERY,
                     "avg_launch_ms": round(trace_ms / max(1, launches), 4), "launches": launches},
        "roofline_pipeline": {"achieved": round(pipe, 2), "frac": round(pipe / HBM_PEAK_GBS, 5),
                              "bytes": pipe_bytes, "formula": "queries*184 + samples*152 over the timed wall time"},
        "cpu_baseline": cpu,
        "extra": {"queries_per_frame": queries // max(1, args.steps), "retests": sum(s.retests for s in stats),
                  "device_ms_per_frame": round(dev_ms / args.steps, 3), "trace_ms_per_frame":
                  round(trace_ms / args.steps, 3), "scene_gen_s": round(gen_s, 2), "load_build_s": round(build_s, 2),
                  "kd_nodes": info.n_kd_nodes, "kd_refs": info.n_kd_refs, "kd_depth": info.kd_depth,
                  "passes": stats[0].passes if stats else 0, "tuning": tuning},
    }
    if cpu:
        line["extra"]["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
