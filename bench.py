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
import csv
import ctypes
import glob
import json
import os
import shutil
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
TRACE_BYTES_PER_QUERY = 48      # read ray (2 x float4) + write hit (float4)
PIPE_BYTES_PER_QUERY = 184      # ray/hit/path-state streams of trace + shade + compaction
PIPE_BYTES_PER_SAMPLE = 152     # camera ray + state write, FP64 accumulate read + write
TRACE_KERNEL = "k_trace"
COPY_KERNEL = "__amd_rocclr_copyBuffer"
COPY_PROBE_MIB = 1024


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["icosphere", "strands"], default="icosphere",
                    help="icosphere: displaced icosphere, 1920x1080 16 spp depth 8 (the headline); strands: 50,000 "
                         "random-walk strands (2.95 M thin triangles), 1024x1024 32 spp depth 6")
    ap.add_argument("--subdiv", type=int, default=8)
    ap.add_argument("--width", type=int, default=0, help="0: the workload's")
    ap.add_argument("--height", type=int, default=0, help="0: the workload's")
    ap.add_argument("--spp", type=int, default=0, help="0: the workload's")
    ap.add_argument("--rdepth", type=int, default=0, help="0: the workload's")
    ap.add_argument("--accel", choices=["kdtree", "octree"], default="kdtree")
    ap.add_argument("--spp-pass", type=int, default=0, help="samples per wavefront pass (0: auto)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="CPU baseline threads (0: the host's cores, at most 16, the GPU's share of the host)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--autotune", action="store_true",
                    help="search the launch shapes first (default: the library's defaults, which are the shapes "
                         "this search picked for this workload on MI355X)")
    ap.add_argument("--tuning", default="", help="launch shape as JSON (xt_tuning fields); skips the autotuner")
    ap.add_argument("--one-group-frames", type=int, default=3,
                    help="frames rendered after the timed steps with each pass as one group (frame_lanes 1), "
                         "reported as roofline.one_group (0: skip)")
    ap.add_argument("--serve-jobs", type=int, default=256, help="jobs of the mixed batch (0: skip serving)")
    ap.add_argument("--serve-requests", type=int, default=512, help="Poisson requests per arrival rate")
    ap.add_argument("--serve-checked", type=int, default=4,
                    help="served requests per arrival rate whose image is compared with the oracle's")
    ap.add_argument("--serve-rates", default="0.5,0.8,0.95", help="arrival rates as fractions of capacity")
    ap.add_argument("--serve-config", default="",
                    help="serving engine settings as JSON (xt_serve_config: regen, graph, lane_streams as booleans; "
                         "rounds_per_step, paths, sample_buffer_mb)")
    ap.add_argument("--serve-log", default="", help="file to append one JSON line per served or plain request to")
    ap.add_argument("--strands", type=int, default=50000, help="strands of the hair scene in the serving mix")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="after the timed steps, copy MIB MiB device-to-device once (a kernel of known traffic "
                         "that scales hardware byte counters)")
    ap.add_argument("--traffic", choices=["auto", "off"], default="auto",
                    help="auto: before touching the GPU, run two 2-frame copies of this workload under rocprofv3 "
                         "counter collection (FETCH_SIZE; WRITE_SIZE + TCC_HIT/TCC_MISS) and report the trace "
                         "kernel's measured traffic per query and L2 hit rate")
    args = ap.parse_args()
    dflt = {"icosphere": (1920, 1080, 16, 8), "strands": (1024, 1024, 32, 6)}[args.workload]
    args.width = args.width or dflt[0]
    args.height = args.height or dflt[1]
    args.spp = args.spp or dflt[2]
    args.rdepth = args.rdepth or dflt[3]
    return args


def _counter_rows(out_dir):
    """(kernel name, counter, value) rows of a rocprofv3 --pmc csv run."""
    rows = []
    for f in glob.glob(os.path.join(out_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(f, newline="") as fh:
            for r in csv.DictReader(fh):
                rows.append((r.get("Kernel_Name", ""), r.get("Counter_Name", ""), float(r.get("Counter_Value", 0))))
    return rows


def traffic_probe(args):
    """Runs a 2-frame copy of this benchmark (same scene and frame, plus a
    1 GiB device copy) twice under rocprofv3 counter collection, one counter
    group per run, and returns the trace kernel's HBM-side traffic per query
    (FETCH_SIZE and WRITE_SIZE, each scaled by what it reads for the copy of
    known size) and its L2 hit rate. Must run before this process touches the
    GPU: the children are separate processes."""
    exe = shutil.which("rocprofv3")
    if not exe:
        return {"error": "rocprofv3 not found"}
    child = [sys.executable, os.path.abspath(__file__), "--steps", "1", "--warmup", "1", "--no-cpu-baseline",
             "--serve-jobs", "0", "--traffic", "off", "--one-group-frames", "0", "--copy-probe", str(COPY_PROBE_MIB),
             "--workload", args.workload, "--strands", str(args.strands),
             "--subdiv", str(args.subdiv), "--width", str(args.width), "--height", str(args.height),
             "--spp", str(args.spp), "--rdepth", str(args.rdepth), "--accel", args.accel]
    if args.tuning:
        child += ["--tuning", args.tuning]
    passes = (["FETCH_SIZE"], ["WRITE_SIZE", "TCC_HIT_sum", "TCC_MISS_sum"])
    sums, copy, queries = {}, {}, []
    with tempfile.TemporaryDirectory(prefix="xt_pmc_") as tmp:
        for k, counters in enumerate(passes):
            out = os.path.join(tmp, f"pass{k}")
            cmd = ["timeout", "-s", "KILL", "180", exe, "--pmc", *counters, "--output-format", "csv", "-d", out,
                   "-o", "run", "--", *child]
            log(f"traffic probe pass {k}: {' '.join(counters)}")
            r = subprocess.run(cmd, cwd=tmp, capture_output=True, text=True)
            lines = [ln for ln in r.stdout.splitlines() if ln.startswith("{")]
            if not lines:
                return {"error": f"counter run {k} ended with status {r.returncode} and no result line",
                        "stderr_tail": r.stderr[-400:]}
            queries.append(json.loads(lines[-1])["extra"]["queries_per_frame"] * 2)
            for name, ctr, val in _counter_rows(out):
                if name.split("<")[0].split("(")[0].strip().endswith(TRACE_KERNEL):
                    sums[ctr] = sums.get(ctr, 0.0) + val
                elif COPY_KERNEL in name:
                    copy[ctr] = max(copy.get(ctr, 0.0), val)
    need = ("FETCH_SIZE", "WRITE_SIZE")
    if any(c not in sums or not copy.get(c) for c in need):
        return {"error": f"counters missing: trace {sorted(sums)}, copy {sorted(copy)}"}
    kib = COPY_PROBE_MIB * 1024.0
    f_fetch, f_write = kib / copy["FETCH_SIZE"], kib / copy["WRITE_SIZE"]
    fetch_b = sums["FETCH_SIZE"] * f_fetch * 1024.0 / queries[0]
    write_b = sums["WRITE_SIZE"] * f_write * 1024.0 / queries[1]
    hit, miss = sums.get("TCC_HIT_sum"), sums.get("TCC_MISS_sum")
    return {"bytes_per_query": round(fetch_b + write_b, 1), "fetch_bytes_per_query": round(fetch_b, 1),
            "write_bytes_per_query": round(write_b, 1), "fetch_scale": round(f_fetch, 4),
            "write_scale": round(f_write, 4),
            "l2_hit_rate": round(hit / (hit + miss), 4) if hit is not None and miss and hit + miss > 0 else None,
            "method": "rocprofv3 --pmc over 2 frames of this workload in child processes; FETCH_SIZE and "
                      "WRITE_SIZE summed over the k_trace dispatches, each scaled by its reading for a "
                      f"{COPY_PROBE_MIB} MiB device copy, divided by the frames' queries; "
                      "L2 hit rate = TCC_HIT / (TCC_HIT + TCC_MISS) over the same dispatches"}


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for ln in fh:
                if ln.startswith("model name"):
                    return ln.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_baseline(scene_path, args, threads):
    """Reference renderer (oracle/_ref) on a spread subset of the frame's
    32x32 tiles, on `threads` host threads and on 1 thread; falls back to the
    oracle port. A query's cost does not depend on how many samples a pixel
    takes, so the sample runs at 1 spp, 4 tiles per thread. Returns the
    cpu_baseline object (value: the run on `threads` threads)."""
    from oracle import oracle as O
    W, H = args.width, args.height
    tiles = ((W + 31) // 32) * ((H + 31) // 32)
    model = _cpu_model()
    if O.ref_available():
        L = O.ref()
        L.xr_render_tiles.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 7 + [ctypes.c_uint, ctypes.c_int,
                                                                               ctypes.c_void_p, ctypes.c_void_p]
        t0 = time.perf_counter()
        r = O.RefScene(scene_path)
        load_s = time.perf_counter() - t0
        log(f"cpu baseline: reference scene loaded in {load_s:.1f}s")

        def run(nt):
            # 4 tiles per thread spread evenly over the frame at 1 spp; a tile's
            # cost ranges from milliseconds (sky) to seconds (tiles over the
            # mesh's symmetry planes), so a sample is not extrapolated
            sec, q = ctypes.c_double(), ctypes.c_uint64()
            stride = max(1, tiles // (4 * nt))
            L.xr_render_tiles(r._h, 3, W, H, 1, 1, args.rdepth, nt, 1, stride, ctypes.byref(sec), ctypes.byref(q))
            log(f"cpu baseline: {(tiles + stride - 1) // stride} tiles at 1 spp on {nt} threads in {sec.value:.2f}s")
            return q.value, sec.value, stride, 1

        qa, sa, stride, spp = run(threads)
        q1, s1, stride1, spp1 = run(1) if threads > 1 else (qa, sa, stride, spp)
        n_tiles = (tiles + stride - 1) // stride
        return {"value": qa / sa / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "reference",
                "cpu_model": model, "host_threads": os.cpu_count(),
                "single_thread": {"value": q1 / s1 / 1e6, "cores": 1, "queries": q1, "seconds": round(s1, 2),
                                  "tiles": (tiles + stride1 - 1) // stride1, "spp": spp1},
                "sample": f"reference pathtracer (compiled -O3 from its sources, its OpenMP dynamic tile loop on "
                          f"{threads} thread(s)) on every {stride}th 32x32 tile ({n_tiles} of {tiles}) of the same "
                          f"frame at {spp} spp, rdepth {args.rdepth}; {qa} Scene::intersection calls in "
                          f"{sa:.2f}s; scene load + octree build {load_s:.1f}s not timed; the same on 1 thread "
                          f"in single_thread"}
    import xtracer_amd as xa
    from xtracer_amd import scenes
    s = xa.Scene(scene_path, build=True)
    orc = O.OracleScene(s)
    cam = scenes.camera_dict(**(scenes.ICOSPHERE_CAMERA if args.workload == "icosphere" else scenes.STRANDS_CAMERA))
    stride = max(1, tiles // (4 * threads))
    t0 = time.perf_counter()
    _, nq = orc.render(3, W, H, cam, samples=1, rdepth=args.rdepth, threads=threads, tile_stride=stride)
    dt = time.perf_counter() - t0
    return {"value": nq / dt / 1e6, "unit": "Mrays/s", "cores": threads, "kind": "port", "cpu_model": model,
            "sample": f"FP64 oracle port on every {stride}th tile at 1 spp on {threads} threads; {nq} queries in "
                      f"{dt:.2f}s"}


def _pct(xs, q):
    import numpy as np
    return float(np.percentile(np.asarray(xs, dtype=np.float64), q)) if len(xs) else None


def _orbit_pose(cam, rng):
    """A random pose on the sphere through the scene camera's position around
    its target (seeded), same lens, rounded to float like the scene loader."""
    import numpy as np
    p, t = np.asarray(cam["position"], float), np.asarray(cam["target"], float)
    r = float(np.linalg.norm(p - t))
    az, el = rng.uniform(0, 2 * np.pi), rng.uniform(0.05, 0.6)
    pos = t + r * np.array([np.cos(el) * np.sin(az), np.sin(el), np.cos(el) * np.cos(az)])
    f = lambda v: tuple(float(np.float32(c)) for c in v)  # noqa: E731
    return {"position": f(pos), "target": f(t), "fov": float(np.float32(cam.get("fov", 45.0))),
            "aperture": float(np.float32(cam.get("aperture", 0.0))),
            "flength": float(np.float32(cam.get("flength", 0.0)))}


def serving(dev, xa, scenes, args, ico_sid, ico_scene, frame_mrays):
    """Mixed-job batch (served vs one job per launch) and Poisson arrivals
    with p50/p99 latency. Scenes: 64 spheres (ao), displaced icosphere,
    strands and terrain (pathtracer); resolutions from {128^2 .. 1024^2} and
    1..64 spp. Per-request image error against the FP64 oracle (same seeded
    paths) on a sample of the small spheres / terrain requests."""
    import numpy as np
    import torch
    from oracle.oracle import OracleScene
    rng = np.random.default_rng(2024)
    pool = []
    t0 = time.perf_counter()
    for kind, fn, kw, integ, rdepth in (
            ("spheres", scenes.spheres, {}, xa.AO, 1),
            ("strands", scenes.strands_scene, {"strands": args.strands}, xa.PATHTRACER, 6),
            ("terrain", scenes.terrain_scene, {}, xa.PATHTRACER, 8)):
        sc = xa.Scene(fn(**kw))
        pool.append((kind, sc, dev.upload(sc), integ, rdepth, scenes.SCENES[kind][1]))
        log(f"serving scene {kind}: {sc.info().n_prims} prims")
    pool.insert(1, ("icosphere", ico_scene, ico_sid, xa.PATHTRACER, 8, scenes.ICOSPHERE_CAMERA))
    setup_s = time.perf_counter() - t0
    # kd-trees of the resident scenes rebuilt on the device (the host build's
    # tree node for node), against the host build's time for the same tree
    kd_build = []
    for kind, sc, sid, _, _, _ in pool:
        st = dev.build_kd(sid)
        i = sc.info()
        kd_build.append({"scene": kind, "prims": i.n_prims, "nodes": st["n_nodes"], "refs": st["n_refs"],
                         "device_ms": round(st["ms"], 2), "host_s": round(i.build_seconds_kd, 3)})
        log(f"kd-tree of {kind} built on the device in {st['ms']:.1f} ms (host build {i.build_seconds_kd:.2f} s)")

    def draw(n, key0, orbit):
        jobs = []
        for k in range(n):
            e = pool[int(rng.integers(len(pool)))]
            kind, sc, sid, integ, rdepth, camd = e
            res = int(rng.choice([128, 256, 512, 1024]))
            spp = int(rng.choice([1, 4, 16, 64]))
            pose = _orbit_pose(camd, rng) if orbit else dict(camd)
            cam = xa.make_camera(pose["position"], pose["target"], (0.0, 1.0, 0.0), pose.get("fov", 45.0),
                                 pose.get("aperture", 0.0), pose.get("flength", 0.0)) if orbit else sc.camera()
            jobs.append((e, pose, xa.make_job(sid, integ, res, res, cam, samples=spp, rdepth=rdepth, seed=7,
                                              job_key=key0 + k)))
        return jobs

    def plain(jobs, arrivals=None):
        out = torch.empty((1024, 1024, 4), dtype=torch.float32, device="cuda")
        lat, q = [], 0
        t0 = time.perf_counter()
        for k, (_, _, j) in enumerate(jobs):
            if arrivals is not None:
                wait = t0 + arrivals[k] - time.perf_counter()
                if wait > 0:
                    time.sleep(wait)
            st = dev.render_device(j, out.data_ptr())
            q += st.queries
            lat.append((time.perf_counter() - t0 - (arrivals[k] if arrivals is not None else 0.0)) * 1e3)
        return time.perf_counter() - t0, q, lat, {}, lat

    def served(jobs, arrivals=None, keep=()):
        info0 = dev.serve_info()
        t0 = time.perf_counter()
        pending, lat, k, imgs = {}, [], 0, {}
        by_job = [0.0] * len(jobs)
        while k < len(jobs) or pending:
            now = time.perf_counter() - t0
            while k < len(jobs) and (arrivals is None or arrivals[k] <= now):
                j = jobs[k][2]
                late = 0.0 if arrivals is None else (time.perf_counter() - t0 - arrivals[k]) * 1e3
                pending[dev.submit(j)] = (k, j.width, j.height, late)
                k += 1
            if not pending:
                time.sleep(max(0.0, t0 + arrivals[k] - time.perf_counter()))
                continue
            dev.serve_step()
            for t in list(pending):
                i, w, h, late = pending[t]
                done, img, st = dev.poll(t, w, h)
                if done:
                    lat.append(late + st.latency_ms)
                    by_job[i] = late + st.latency_ms
                    if i in keep:
                        imgs[i] = img
                    del pending[t]
        wall = time.perf_counter() - t0
        info1 = dev.serve_info()
        return wall, info1["queries"] - info0["queries"], lat, imgs, by_job

    def reqlog(part, load, policy, jobs, arrivals, by_job):
        """One JSON line per request (scene, size, samples, arrival, latency)
        into --serve-log."""
        if not args.serve_log:
            return
        with open(args.serve_log, "a") as f:
            for k, ((e, _, j), ms) in enumerate(zip(jobs, by_job)):
                f.write(json.dumps({"part": part, "load": load, "policy": policy, "k": k, "scene": e[0],
                                    "res": j.width, "spp": j.samples,
                                    "arrival_s": None if arrivals is None else round(float(arrivals[k]), 5),
                                    "latency_ms": round(ms, 2)}) + "\n")

    def summary(wall, q, lat, _imgs, _by_job=None):
        return {"wall_s": round(wall, 3), "Mrays_s": round(q / wall / 1e6, 1),
                "p50_ms": round(_pct(lat, 50), 1), "p99_ms": round(_pct(lat, 99), 1)}

    def request_errors(jobs, imgs):
        """Relative L1 difference of served images from the oracle's render of
        the same job (same seeded paths), and the share of pixels within 1e-3."""
        out = []
        for i, img in sorted(imgs.items()):
            e, pose, j = jobs[i]
            ref, _ = OracleScene(e[1]).render(j.integrator, j.width, j.height, scenes.camera_dict(**pose),
                                              samples=j.samples, rdepth=j.rdepth, seed=j.seed, job_key=j.job_key)
            a, b = img[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
            rel = float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-30))
            agree = float(np.isclose(a, b, rtol=1e-3, atol=1e-4).all(axis=2).mean())
            out.append({"scene": e[0], "res": j.width, "spp": j.samples, "rel_l1": round(rel, 6),
                        "pixels_within_1e-3": round(agree, 5)})
        return out

    def pick_checked(jobs, n):
        small = [i for i, (e, _, j) in enumerate(jobs)
                 if e[0] in ("spheres", "terrain") and j.width <= 256 and j.samples <= 16]
        return set(small[:n])

    policies = ("fcfs", "srpt", "hybrid")
    if args.serve_config:
        dev.serve_configure(**json.loads(args.serve_config))
    res = {"scenes": [p[0] for p in pool], "setup_s": round(setup_s, 1), "engine": dev.serve_config(),
           "policies": list(policies), "kd_build": kd_build,
           "kd_build_host_threads": os.cpu_count()}
    jobs = draw(args.serve_jobs, 10000, False)
    samples = sum(j.width * j.height * j.samples for _, _, j in jobs)
    served(jobs[:4])  # warm the server's buffers and graph
    pw, pq, plat, _, pby = plain(jobs)
    reqlog("batch", None, "plain", jobs, None, pby)
    res["batch"] = {"jobs": len(jobs), "samples": samples, "plain": summary(pw, pq, plat, None)}
    walls = {}
    for pol in policies:
        dev.set_serve_policy(pol)
        sw, sq, slat, _, sby = served(jobs)
        reqlog("batch", None, pol, jobs, None, sby)
        walls[pol] = sw
        res["batch"]["served_" + pol] = summary(sw, sq, slat, None)
        log(f"serving batch ({pol}): plain {pw:.2f}s served {sw:.2f}s")
    res["batch"]["served_over_plain"] = {pol: round(pw / walls[pol], 3) for pol in policies}
    # the served engine can at best run the mix at the single-frame rate:
    # served/plain <= frame Mrays/s / plain Mrays/s on this mix
    plain_mrays = pq / pw / 1e6
    res["batch"]["served_over_plain_bound"] = round(frame_mrays / plain_mrays, 3) if plain_mrays > 0 else None
    mu = len(jobs) / min(walls.values())
    res["capacity_jobs_s"] = round(mu, 2)
    res["poisson"] = []
    for rate in [float(x) for x in args.serve_rates.split(",") if x]:
        lam = rate * mu
        jobs = draw(args.serve_requests, 20000 + int(rate * 1000) * 100, True)
        arrivals = np.cumsum(rng.exponential(1.0 / lam, len(jobs)))
        keep = pick_checked(jobs, args.serve_checked)
        entry = {"load": rate, "lambda_jobs_s": round(lam, 2), "requests": len(jobs)}
        imgs = {}
        for pol in ("srpt", "hybrid"):
            dev.set_serve_policy(pol)
            wall, q, lat, got, by = served(jobs, arrivals, keep)
            reqlog("poisson", rate, pol, jobs, arrivals, by)
            entry["served_" + pol] = summary(wall, q, lat, None)
            imgs = imgs or got
        pr = plain(jobs, arrivals)
        reqlog("poisson", rate, "plain", jobs, arrivals, pr[4])
        entry["plain"] = summary(*pr)
        t0 = time.perf_counter()
        entry["request_error"] = request_errors(jobs, imgs)
        entry["request_error_s"] = round(time.perf_counter() - t0, 1)
        res["poisson"].append(entry)
        log(f"poisson load {rate}: p99 plain {entry['plain']['p99_ms']} ms, "
            + ", ".join(f"{pol} {entry['served_' + pol]['p99_ms']} ms" for pol in ("srpt", "hybrid")))
    dev.set_serve_policy("srpt")
    return res


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would render its own "
              "tiles or jobs with a replicated scene); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    # counter runs first: they are child processes, and this one has not
    # touched the GPU yet
    probe = traffic_probe(args) if args.traffic == "auto" else None
    if probe:
        log(f"traffic probe: {probe}")
    import torch

    import xtracer_amd as xa
    from xtracer_amd import scenes

    t0 = time.perf_counter()
    if args.workload == "icosphere":
        path = scenes.icosphere_scene(seed=0, subdiv=args.subdiv)
    else:
        path = scenes.strands_scene(strands=args.strands)
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
    tune_table = None
    if args.autotune and not args.tuning:
        from xtracer_amd.autotune import autotune
        t0 = time.perf_counter()
        tuning, table = autotune(dev, job(999), out.data_ptr(), wave_paths=(16 << 20, 64 << 20), log=log)
        tune_table = [dict(t, ms=round(ms, 3)) for t, ms in table]
        log(f"autotune picked {tuning} in {time.perf_counter() - t0:.1f}s")
    for k in range(args.warmup):
        t0 = time.perf_counter()
        st = dev.render_device(job(k), out.data_ptr())
        log(f"warmup {k}: {time.perf_counter() - t0:.3f}s, {st.queries} queries, {st.retests} retests, "
            f"{st.trace_fail} FP32 walks and {st.retest_fail} FP64 walks out of steps")
    torch.cuda.synchronize()
    stats = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        stats.append(dev.render_device(job(args.warmup + k), out.data_ptr()))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"timed {args.steps} steps in {wall:.3f}s")
    if args.copy_probe > 0:
        src = torch.empty(args.copy_probe << 18, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        del src, dst
    one_group = None
    if args.one_group_frames > 0 and tuning.get("frame_lanes", 1) != 1:
        # the same frame with each pass as one group: the trace launches then
        # run alone, so their summed time is the kernel's own
        dev.set_tuning(frame_lanes=1)
        dev.render_device(job(args.warmup + args.steps), out.data_ptr())
        og = [dev.render_device(job(args.warmup + args.steps + 1 + k), out.data_ptr())
              for k in range(args.one_group_frames)]
        dev.set_tuning(frame_lanes=tuning["frame_lanes"])
        oq, otr = sum(s.queries for s in og), sum(s.ms_trace for s in og)
        ol = sum(s.trace_launches for s in og)
        o_ach = oq * TRACE_BYTES_PER_QUERY / (otr * 1e-3) / 1e9 if otr > 0 else 0.0
        one_group = {"frames": len(og), "device_ms_per_frame": round(sum(s.ms_total for s in og) / len(og), 3),
                     "trace_ms_per_frame": round(otr / len(og), 3),
                     "avg_launch_ms": round(otr / max(1, ol), 4), "launches": ol,
                     "achieved": round(o_ach, 2), "frac": round(o_ach / HBM_PEAK_GBS, 5)}
        log(f"one group per pass: {one_group}")
    queries = sum(s.queries for s in stats)
    samples = sum(s.samples for s in stats)
    trace_ms = sum(s.ms_trace for s in stats)
    busy_ms = sum(s.ms_trace_busy for s in stats)
    launches = sum(s.trace_launches for s in stats)
    dev_ms = sum(s.ms_total for s in stats)
    value = queries / wall / 1e6
    # the trace launches of a frame's two pixel groups run at the same time:
    # the kernel's rate is its bytes over the time in which any launch runs
    # (the union of the launches' intervals; the sum of their times when
    # they do not overlap)
    achieved = queries * TRACE_BYTES_PER_QUERY / (busy_ms * 1e-3) / 1e9 if busy_ms > 0 else 0.0
    pipe_bytes = queries * PIPE_BYTES_PER_QUERY + samples * PIPE_BYTES_PER_SAMPLE
    pipe = pipe_bytes / (wall) / 1e9
    threads = args.cpu_threads if args.cpu_threads > 0 else min(16, os.cpu_count() or 1)
    serve = serving(dev, xa, scenes, args, sid, scene, value) if args.serve_jobs > 0 else None
    cpu = None if args.no_cpu_baseline else cpu_baseline(path, args, threads)
    line = {
        "metric": ("Mrays/s at 1080p 16spp (pathtracer, plain one-job-per-launch)" if args.workload == "icosphere"
                   else f"Mrays/s at {args.width}x{args.height} {args.spp}spp (strands pathtracer, plain)"),
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
        "data": f"synthetic (seeded {args.workload} scene, generated in-process)",
        "config": {"workload": (f"pathtracer icosphere {info.n_prims} prims {args.width}x{args.height} "
                                f"{args.spp}spp rdepth {args.rdepth} {args.accel} thin-lens"
                                if args.workload == "icosphere" else
                                f"pathtracer strands {info.n_prims} prims {args.width}x{args.height} "
                                f"{args.spp}spp rdepth {args.rdepth} {args.accel} blinn-phong"),
                   "width": args.width, "height": args.height, "spp": args.spp, "rdepth": args.rdepth,
                   "prims": info.n_prims, "accel": args.accel},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5),
                     "traffic": (int(probe["bytes_per_query"] * queries / max(1, launches))
                                 if probe and "bytes_per_query" in probe and launches else None),
                     "traffic_per_query": probe.get("bytes_per_query") if probe else None,
                     "l2_hit_rate": probe.get("l2_hit_rate") if probe else None,
                     "traffic_probe": probe,
                     "kernel": "k_trace (kd-tree)", "bytes_per_query": TRACE_BYTES_PER_QUERY,
                     "avg_launch_ms": round(trace_ms / max(1, launches), 4), "launches": launches,
                     "busy_ms_per_frame": round(busy_ms / max(1, args.steps), 3),
                     "concurrent_groups": tuning.get("frame_lanes", 1),
                     "achieved_formula": "queries * 48 B over the time in which at least one trace launch runs",
                     "one_group": one_group},
        "roofline_pipeline": {"achieved": round(pipe, 2), "frac": round(pipe / HBM_PEAK_GBS, 5),
                              "bytes": pipe_bytes, "formula": "queries*184 + samples*152 over the timed wall time"},
        "cpu_baseline": cpu,
        "extra": {"queries_per_frame": queries // max(1, args.steps), "retests": sum(s.retests for s in stats),
                  "device_ms_per_frame": round(dev_ms / args.steps, 3), "trace_ms_per_frame":
                  round(trace_ms / args.steps, 3), "scene_gen_s": round(gen_s, 2), "load_build_s": round(build_s, 2),
                  "kd_nodes": info.n_kd_nodes, "kd_refs": info.n_kd_refs, "kd_depth": info.kd_depth,
                  "passes": stats[0].passes if stats else 0, "tuning": tuning},
    }
    if tune_table:
        line["autotune"] = tune_table
    if serve:
        line["serving"] = serve
    if cpu:
        line["extra"]["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    dev.close()


if __name__ == "__main__":
    main()
