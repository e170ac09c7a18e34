"""Benchmark of pine's path-tracing hot path on one MI355X.

Headline workload (`value`, Mrays/s): closest-hit BVH traversal of a
1,048,576-triangle displaced geodesic icosphere by 16,777,216 rays (4,194,304
primary rays of a 2048x2048 pinhole camera + 12,582,912 cosine-distributed
rays leaving seeded surface points).  One step = one pg_trace_closest pass
over all rays, rays and hierarchy resident in HBM; by default the launch
first sorts the rays on the device by direction octant and origin cell
(PG_LAUNCH_REORDER: a key kernel and a radix sort, inside the step) and
writes every hit to its ray's slot.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--width 2] [--block 128]

Prints one JSON line.  The roofline object accounts the dominant kernel
(k_trace): algorithmic bytes per launch = 40 B compulsory per ray (32 B ray
read + 8 B hit written) + node bytes x nodes fetched + record bytes x prims
tested, the counts taken from the instrumented variant of the same kernel over
the same rays; duration = average of HIP events around each launch on the
launch stream.  `traffic` is measured in the same run: rocprofv3 --pmc
FETCH_SIZE and WRITE_SIZE over the same launches in two child runs, scaled by
a known-size copy.  cpu_baseline is pine's own BVH::Intersect (pine built from
its sources, oracle/_ref/ref_harness) over all rays on the host cores; the
oracle's CPU port of the traversal is reported beside it.

Secondary workloads, each next to pine itself on the host cores:
  path_tracing       the seeded 64-spheres scene, 512x512, 64 spp, maxDepth 8 (also
                     in spectral mode, four wavelengths per path; GPU only)
  mesh_path_tracing  the 2^20-triangle icosphere in a lit scene, 512x512, 64 spp
  volume             a 256^3 fBm density grid, 1024x1024, 256 spp, maxDepth 64,
                     16^3-voxel majorant bricks, HG g = 0.85, albedo 0.99, lit by a
                     directional sun and a constant environment (pine, which has no
                     constant environment light, renders it under its Sky)
and, GPU only:
  batch              64 scenes of uneven size (10k..2M triangles plus instanced
                     boxes, 256^2..1024^2, 16..128 spp) autotuned, then rendered one
                     by one and as one batch over concurrent streams with
                     graph-replayed bounce loops
  guided             the occluded-light box: plain vs learned-guided path tracing,
                     time to reach the plain tracer's 512-spp relMSE

Every child process (pine's runs, the traffic runs) is started and finished
before this process makes its first HIP call, and pine's timed regions run
one at a time with nothing else on the host.  Under a profiler that preloads
its library (rocprofv3) no child is started at all.
"""
import argparse
import csv
import glob
import json
import math
import os
import shutil
import signal
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
MFMA_BF16_PEAK_TFLOPS = 2500.0
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")

# Queue-record bytes of the fused path-tracing pipeline (the algorithmic
# traffic of its wavefront stages; scene data such as BVH nodes, triangles and
# materials is not counted):
#   per camera sample  108 B: raygen writes the ray (32) and the path state
#                      (L, beta, pdf, sampler index and dimension, depth,
#                      film position, queue slot: 56); the film pass reads L
#                      and the film position (20);
#   per extension ray  220 B: trace reads queue slot and ray (36) and writes
#                      hit and uv (16); shade reads slot, hit, uv, ray and
#                      path state (96) and writes the next ray, its queue slot
#                      and the updated state (72);
#   per shadow ray     128 B: shade writes ray, path id and contribution (48);
#                      the occlusion trace reads the ray and writes a flag
#                      (36); resolve reads flag, path id and contribution and
#                      updates L (44).
QUEUE_BYTES_SAMPLE, QUEUE_BYTES_EXT, QUEUE_BYTES_SHADOW = 108, 220, 128
# Guide network evaluation: 2 x 25,449 multiply-adds of the 225-64-64-64-41 MLP.
GUIDE_FLOP_PER_QUERY = 2 * 25449


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled():
    """True when this process runs under a profiler that preloads its library:
    then no child process may be started (the profiler has already set up the
    GPU in this process)."""
    return any(k.startswith("ROCPROF") for k in os.environ) or \
        "rocprof" in os.environ.get("LD_PRELOAD", "")


def cpu_share():
    """The CPUs this process may run on: its affinity mask, capped by a cgroup
    (v2) CPU quota when one is set.  Pine's ParallelFor uses
    hardware_concurrency(); threads beyond this share would only time-slice."""
    try:
        aff = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        aff = os.cpu_count() or 1
    quota = None
    try:
        q, period = open("/sys/fs/cgroup/cpu.max").read().split()
        if q != "max":
            quota = int(q) / int(period)
    except (OSError, ValueError):
        pass
    n = aff if quota is None else max(1, min(aff, int(math.floor(quota))))
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return n, {"threads": n, "host_cpus": os.cpu_count(), "affinity_cpus": aff,
               "cgroup_cpu_quota": quota, "cpu_model": model}


class PineRun:
    """pine itself (oracle/_ref/ref_harness: pine's library built from its
    sources) in a child process: `render` (pine's PathIntegrator on a scene
    file; samples/s excludes scene loading and BVH build, as pine's own
    [Rendering]Average line does) or `time` (pine's BVH::Intersect over a ray
    file).  With gate=True the child builds its scene and BVH, prints
    "ready" and waits for go() before its timed region, so several children
    can build at once and still be timed one at a time."""

    def __init__(self, args, scene_text=None, files=(), gate=False):
        self.dir = tempfile.TemporaryDirectory()
        for name, writer in files:
            writer(os.path.join(self.dir.name, name))
        if scene_text is not None:
            with open(os.path.join(self.dir.name, "scene.txt"), "w") as f:
                f.write(scene_text)
        env = dict(os.environ)
        if gate:
            env["PG_HARNESS_GATE"] = "1"
        self.t0 = time.time()
        self.p = subprocess.Popen([HARNESS, *args], cwd=self.dir.name, stdin=subprocess.PIPE,
                                  stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)

    def wait_ready(self):
        """Waits for the child's "ready" line (pine's own log lines come first)."""
        while True:
            line = self.p.stdout.readline()
            if not line:
                raise RuntimeError(f"pine run ended before getting ready: "
                                   f"{self.p.stderr.read()[-500:]}")
            if line.strip() == "ready":
                break
        self.build_s = time.time() - self.t0

    def go(self):
        self.p.stdin.write("go\n")
        self.p.stdin.flush()

    def result(self, timeout=900):
        out, err = self.p.communicate(timeout=timeout)
        wall = time.time() - self.t0
        self.dir.cleanup()
        line = [l for l in out.splitlines() if l.startswith("{")]
        if self.p.returncode != 0 or not line:
            log(f"[bench] pine run failed: {err[-500:]}")
            return None
        r = json.loads(line[-1])
        r["wall_s"] = wall
        return r


def place(src):
    """File writer for PineRun: hard link, else copy."""
    def w(dst):
        try:
            os.link(src, dst)
        except OSError:
            shutil.copyfile(src, dst)
    return w


def write_mesh_bin(path, v, i):
    """The harness's scene.bin with one triangle mesh."""
    with open(path, "wb") as f:
        f.write(np.array([0x31534750, 1, 0, len(v), len(i)], np.uint32).tobytes())
        f.write(np.ascontiguousarray(v, np.float32).tobytes())
        f.write(np.ascontiguousarray(i, np.uint32).tobytes())


def write_rays_bin(path, rays):
    with open(path, "wb") as f:
        f.write(np.array([len(rays)], np.uint64).tobytes())
        f.write(np.ascontiguousarray(rays, np.float32).tobytes())


def cpu_baseline_from(r, what, share):
    if r is None:
        return None
    n = r["w"] * r["h"] * r["spp"]
    return {"value": round(r["msamples_per_s"] * 1e6, 1), "unit": "samples/s",
            "cores": r["threads"], "kind": "reference", "host": share,
            "sample": f"{what}: pine renders {r['w']}x{r['h']} at {r['spp']} spp ({n} samples, "
                      f"{r['render_s']:.1f}s render, {r['wall_s']:.0f}s with scene load and BVH "
                      f"build), alone on the host before any GPU work"}


def pine_baselines(args, W, work, v, i, rays, threads, share):
    """pine on the host cores, every timed region alone and all of it before
    this process makes its first HIP call (no child process is started once
    the GPU is in use).  pine's two single-threaded BVH builds of the
    2^20-triangle mesh (about 1.5 minutes each) run side by side; their timed
    parts then run one at a time:
      mesh      the mesh scene at the GPU leg's size (512x512, 64 spp);
      trace     BVH::Intersect (bvh.cpp:669-709) over all 16,777,216 rays;
      spheres   the spheres scene at the GPU leg's size (512x512, 64 spp);
      volume    the 256^3 cloud at the GPU leg's film (1024x1024) with 8 spp
                (pine has no majorant bricks: its global majorant)."""
    t0 = time.time()
    T = str(threads)
    out = {}
    obj = os.path.join(work, "ico1m.obj")
    mesh_bin, rays_bin = os.path.join(work, "ico1m.bin"), os.path.join(work, "rays16m.bin")
    write_mesh_bin(mesh_bin, v, i)
    write_rays_bin(rays_bin, rays)
    runs = {}
    if not args.no_pt:
        runs["mesh"] = PineRun(["render", "./scene.txt", "film.bin", T],
                               W.mesh_scene_text("ico1m.obj", width=512, height=512, spp=64),
                               files=[("ico1m.obj", place(obj))], gate=True)
    runs["trace"] = PineRun(["time", "scene.bin", "rays.bin", T],
                            files=[("scene.bin", place(mesh_bin)), ("rays.bin", place(rays_bin))],
                            gate=True)
    for r in runs.values():
        r.wait_ready()
    log(f"[bench] pine: BVH builds done in {time.time() - t0:.0f} s")
    for k, r in runs.items():
        r.go()
        out[k] = r.result()
        if out[k]:
            out[k]["build_s"] = round(r.build_s, 1)
    os.remove(rays_bin)
    if not args.no_pt:
        text = W.spheres_scene_text(seed=0, width=512, height=512, spp=args.pt_spp, max_depth=8)
        out["spheres"] = PineRun(["render", "./scene.txt", "film.bin", T], text).result()
        vtext = W.cloud_scene_text("cloud256.vol", width=1024, height=1024, spp=8, max_depth=64)
        out["volume"] = PineRun(["render", "./scene.txt", "film.bin", T], vtext,
                                files=[("cloud256.vol", place(os.path.join(work, "cloud256.vol")))]
                                ).result()
    log(f"[bench] pine baselines done in {time.time() - t0:.0f} s on {threads} threads")
    return out


def traffic_probe(args):
    """HBM bytes per closest-hit launch, measured now: two child runs of this
    script's --traffic-probe mode under rocprofv3 --pmc (FETCH_SIZE, then
    WRITE_SIZE), made before this process touches the GPU.  Each counter is
    scaled by what it reads for a device-to-device copy of the 512 MiB ray
    buffer in the same run (the copy reads and writes exactly its size).
    Returns a dict, or None with the reason logged."""
    exe = shutil.which("rocprofv3")
    if exe is None:
        log("[bench] traffic: rocprofv3 not found")
        return None
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        d = tempfile.mkdtemp()
        cmd = [exe, "--pmc", counter, "-d", d, "-o", "run", "--output-format", "csv", "--",
               sys.executable, os.path.abspath(__file__), "--traffic-probe",
               "--width", str(args.width), "--block", str(args.block), "--flags", str(args.flags),
               "--leaf", str(args.leaf)]
        p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                             start_new_session=True)
        try:
            out, _ = p.communicate(timeout=180)
        except subprocess.TimeoutExpired:
            os.killpg(p.pid, signal.SIGKILL)
            p.communicate()
            log(f"[bench] traffic: {counter} run timed out")
            shutil.rmtree(d, ignore_errors=True)
            return None
        rows = []
        for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
            with open(f) as fh:
                rows += list(csv.DictReader(fh))
        shutil.rmtree(d, ignore_errors=True)
        rows = [r for r in rows if r.get("Counter_Name") == counter]
        rows.sort(key=lambda r: int(r.get("Dispatch_Id") or 0))
        trace = [float(r["Counter_Value"]) for r in rows if "k_trace" in r.get("Kernel_Name", "")]
        copy = [float(r["Counter_Value"]) for r in rows if "copyBuffer" in r.get("Kernel_Name", "")]
        if p.returncode != 0 or len(trace) < 2 or not copy:
            log(f"[bench] traffic: {counter} run gave rc {p.returncode}, {len(trace)} trace and "
                f"{len(copy)} copy rows: {out[-400:]}")
            return None
        # The copy kernel moves 512 MiB each way: scale = true KiB / counted KiB.
        scale = (512 * 1024) / copy[-1]
        # Dispatch 0 is the instrumented pass; the plain launches follow.
        kib = float(np.mean(trace[1:]))
        res[counter] = {"kib_per_dispatch": round(kib, 1), "copy_scale": round(scale, 4),
                        "bytes": kib * scale * 1024}
    out = {"bytes": int(res["FETCH_SIZE"]["bytes"] + res["WRITE_SIZE"]["bytes"]),
           "fetch": res["FETCH_SIZE"], "write": res["WRITE_SIZE"],
           "method": "rocprofv3 --pmc FETCH_SIZE, then WRITE_SIZE, on this script's "
                     "--traffic-probe (one instrumented + 3 plain launches and a 512 MiB copy), "
                     "each counter scaled by its reading of the copy"}
    out["valu"] = valu_probe(args, exe)
    return out


VALU_COUNTERS = ("SQ_WAVES", "SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "GRBM_GUI_ACTIVE")


def valu_probe(args, exe):
    """VALU utilisation of the plain traversal launches, measured now in a third
    child run under rocprofv3 --pmc: busy = 4 x SQ_ACTIVE_INST_VALU (quad-cycles)
    over 1,024 SIMDs x GRBM_GUI_ACTIVE / 8 (the GUI-active cycles of the eight
    XCDs are summed).  Returns a dict, or None with the reason logged."""
    d = tempfile.mkdtemp()
    cmd = [exe, "--pmc", *VALU_COUNTERS, "-d", d, "-o", "run", "--output-format", "csv", "--",
           sys.executable, os.path.abspath(__file__), "--traffic-probe",
           "--width", str(args.width), "--block", str(args.block), "--flags", str(args.flags),
           "--leaf", str(args.leaf)]
    p = subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                         start_new_session=True)
    try:
        out, _ = p.communicate(timeout=180)
    except subprocess.TimeoutExpired:
        os.killpg(p.pid, signal.SIGKILL)
        p.communicate()
        log("[bench] valu: counter run timed out")
        shutil.rmtree(d, ignore_errors=True)
        return None
    per = {}
    for f in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(f) as fh:
            for r in csv.DictReader(fh):
                if "k_trace" in r.get("Kernel_Name", "") and r.get("Counter_Name") in VALU_COUNTERS:
                    key = int(r.get("Dispatch_Id") or 0)
                    per.setdefault(key, {}).setdefault(r["Counter_Name"], 0.0)
                    per[key][r["Counter_Name"]] += float(r["Counter_Value"])
    shutil.rmtree(d, ignore_errors=True)
    launches = [per[k] for k in sorted(per)][1:]  # dispatch 0: the instrumented pass
    if p.returncode != 0 or not launches or any(len(c) != len(VALU_COUNTERS) for c in launches):
        log(f"[bench] valu: counter run gave rc {p.returncode}, {len(launches)} launches: {out[-400:]}")
        return None
    busy = [4.0 * c["SQ_ACTIVE_INST_VALU"] / (1024.0 * c["GRBM_GUI_ACTIVE"] / 8.0) for c in launches]
    per_wave = [c["SQ_INSTS_VALU"] / c["SQ_WAVES"] for c in launches]
    return {"busy": round(float(np.mean(busy)), 4), "insts_per_wave": round(float(np.mean(per_wave)), 1),
            "method": "rocprofv3 --pmc " + " ".join(VALU_COUNTERS) + " on the same --traffic-probe "
                      "launches; busy = 4 x SQ_ACTIVE_INST_VALU / (1024 SIMDs x GRBM_GUI_ACTIVE / 8)"}


def traffic_probe_main(args):
    """--traffic-probe: the traversal workload's launches and the calibration
    copy, nothing else (run under rocprofv3 by traffic_probe)."""
    import pine_amd as pg
    from pine_amd import hip
    from pine_amd import workloads as W
    v, i = W.make_mesh(W.ICOSPHERE_1M)
    rays = W.make_rays(W.RAYS_16M, v, i)
    acc = pg.Accel()
    acc.add_mesh(v, i)
    acc.build(width=args.width, leaf_max=args.leaf)
    acc.set_launch(args.block, args.flags)
    d_rays = hip.DeviceArray.from_numpy(rays)
    d_hits = hip.DeviceArray((d_rays.shape[0], 2), np.int32)
    acc.trace_counts(d_rays)
    for _ in range(3):
        acc.trace_closest(d_rays, d_hits)
    dst = hip.DeviceArray(d_rays.shape, np.float32)
    hip.check(hip.rt().hipMemcpy(dst.ptr, d_rays.ptr, d_rays.nbytes, hip.hipMemcpyDeviceToDevice))
    hip.synchronize()
    print("traffic probe done", flush=True)


def traversal(args, pg, hip, W, v, i, rays, traffic):

    acc = pg.Accel()
    acc.add_mesh(v, i)
    acc.build(width=args.width, leaf_max=args.leaf)
    acc.set_launch(args.block, args.flags)
    st = acc.stats()
    launch_flags = args.flags if args.flags >= 0 else pg.LAUNCH_DEFAULT
    log(f"[bench] accel built in {st['build_ms']:.0f} ms: {st['n_nodes']} nodes, depth {st['max_depth']}")

    d_rays = hip.DeviceArray.from_numpy(rays)
    n = d_rays.shape[0]
    d_hits = hip.DeviceArray((n, 2), np.int32)

    # Instrumented pass: work per ray over the same rays and tree.
    ch, cnt = acc.trace_counts(d_rays)
    hip.synchronize()
    c = cnt.numpy()
    nodes_fetched = int(c[:, 0].sum(dtype=np.int64))
    prims_tested = int(c[:, 1].sum(dtype=np.int64))
    del c, cnt, ch
    node_bytes = 64 if st["width"] == 2 else 128
    compulsory = 40 * n
    traversal_bytes = nodes_fetched * node_bytes + prims_tested * 48
    algo_bytes = compulsory + traversal_bytes

    # Launches go to the default stream; events are recorded on that stream.
    for _ in range(args.warmup):
        acc.trace_closest(d_rays, d_hits)
    hip.synchronize()
    starts = [hip.Event() for _ in range(args.steps)]
    ends = [hip.Event() for _ in range(args.steps)]
    hip.synchronize()
    w0 = time.perf_counter()
    for k in range(args.steps):
        starts[k].record()
        acc.trace_closest(d_rays, d_hits)
        ends[k].record()
    hip.synchronize()
    wall = time.perf_counter() - w0
    step_ms = float(np.mean([s.elapsed_ms(e) for s, e in zip(starts, ends)]))
    ms_per_step = wall * 1e3 / args.steps
    mrays = n / (ms_per_step * 1e-3) * 1e-6
    hits = d_hits.numpy()
    hit_frac = float((hits[:, 0] >= 0).mean())
    # The traversal kernel alone (a reordered launch also runs the key kernel
    # and the sort): HIP events the library records around that kernel on its
    # stream (PG_LAUNCH_TIMED), one launch at a time.
    acc.set_launch(args.block, launch_flags | pg.LAUNCH_TIMED)
    trace_ms = []
    for _ in range(args.steps):
        acc.trace_closest(d_rays, d_hits)
        trace_ms.append(acc.stats()["last_trace_ms"])
    acc.set_launch(args.block, launch_flags)
    kern_ms = float(np.mean(trace_ms))
    same_timed = bool(np.array_equal(d_hits.numpy(), hits))
    # The same rays by the stackless traversal (per-ray trail, key table,
    # wave-wide node fetches; binary trees): same hits bit for bit, its rate.
    variants = {}
    if st["width"] == 2 and not (launch_flags & pg.LAUNCH_STACKLESS):
        acc.set_launch(args.block, launch_flags | pg.LAUNCH_STACKLESS)
        acc.trace_closest(d_rays, d_hits)
        ev = [hip.Event() for _ in range(10)]
        for k in range(5):
            ev[2 * k].record()
            acc.trace_closest(d_rays, d_hits)
            ev[2 * k + 1].record()
        hip.synchronize()
        sl_ms = float(np.mean([ev[2 * k].elapsed_ms(ev[2 * k + 1]) for k in range(5)]))
        same = bool(np.array_equal(d_hits.numpy(), hits))
        variants["stackless"] = {"step_ms": round(sl_ms, 3), "launch_flags": launch_flags | pg.LAUNCH_STACKLESS,
                                 "mrays_per_s": round(n / (sl_ms * 1e-3) * 1e-6, 1),
                                 "hits_identical": same}
        log(f"[bench] stackless traversal: step {sl_ms:.3f} ms, same hits: {same}")
        acc.set_launch(args.block, launch_flags)
    if args.copy_calib:
        dst = hip.DeviceArray(d_rays.shape, np.float32)
        hip.check(hip.rt().hipMemcpy(dst.ptr, d_rays.ptr, d_rays.nbytes, hip.hipMemcpyDeviceToDevice))
        hip.synchronize()
        dst.free()
    # The same rays in the callers' order, no sort: the reordering's gain.
    unordered = None
    if launch_flags & pg.LAUNCH_REORDER:
        acc.set_launch(args.block, launch_flags & ~pg.LAUNCH_REORDER)
        acc.trace_closest(d_rays, d_hits)
        ev = [hip.Event() for _ in range(10)]
        for k in range(5):
            ev[2 * k].record()
            acc.trace_closest(d_rays, d_hits)
            ev[2 * k + 1].record()
        hip.synchronize()
        un_ms = float(np.mean([ev[2 * k].elapsed_ms(ev[2 * k + 1]) for k in range(5)]))
        unordered = {"step_ms": round(un_ms, 3), "mrays_per_s": round(n / (un_ms * 1e-3) * 1e-6, 1),
                     "hits_identical": bool(np.array_equal(d_hits.numpy(), hits))}
        acc.set_launch(args.block, launch_flags)
        log(f"[bench] rays in the callers' order: step {un_ms:.3f} ms, same hits: "
            f"{unordered['hits_identical']}")
    log(f"[bench] {ms_per_step:.3f} ms/step (events {step_ms:.3f} ms), traversal kernel {kern_ms:.3f} ms, "
        f"{mrays:.1f} Mrays/s, hit fraction {hit_frac:.3f}")
    d_rays.free()
    d_hits.free()
    acc.release()

    achieved = algo_bytes / (kern_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic["bytes"] if traffic else None,
        "kernel": f"k_trace<W={st['width']}>",
        "bytes_per_launch": {"compulsory": compulsory, "traversal": traversal_bytes,
                             "node_fetch": nodes_fetched * node_bytes,
                             "counters": traffic["bytes"] if traffic else None},
        # The algorithmic figure counts every node visit at its full size,
        # L2 and MALL hits included; the counters see only L2 misses.
        "frac_counters": round(traffic["bytes"] / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)
        if traffic else None,
        "compulsory_frac": round(compulsory / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "limiter": (f"VALU issue ({traffic['valu']['busy']:.0%} busy, measured in this run), "
                    if traffic and traffic.get("valu") else "VALU issue (78% busy in "
                    "profiles/k_trace2_reordered_icosphere1m_rays16m.md), ") +
                   "not HBM bandwidth; the algorithmic figure counts every node visit at 64 B, "
                   "cache hits included, so it exceeds the HBM peak",
        "valu": traffic.get("valu") if traffic else None,
        "kernel_ms": round(kern_ms, 4),
        "per_ray": {"nodes": round(nodes_fetched / n, 3), "prims": round(prims_tested / n, 3)},
        "traffic_method": traffic["method"] if traffic else "not measured in this run",
    }
    config = {"workload": "bvh-traversal: icosphere 1,048,576 tris, 16,777,216 rays",
              "triangles": int(len(i)), "rays": int(n), "node_width": st["width"],
              "leaf_max": args.leaf, "block": args.block, "launch_flags": launch_flags,
              "hit_fraction": round(hit_frac, 4), "build_ms": round(st["build_ms"], 1),
              "step_ms_events": round(step_ms, 4), "trace_kernel_ms": round(kern_ms, 4),
              "reorder_ms": round(step_ms - kern_ms, 4) if launch_flags & pg.LAUNCH_REORDER else 0.0,
              "timed_hits_identical": same_timed, "unordered": unordered, "variants": variants}
    return {"mrays": mrays, "ms_per_step": ms_per_step, "roofline": roofline, "config": config,
            "rays": rays, "hits": hits}


def traversal_cpu(trav, oracle, v, i, threads, share):
    """The oracle's CPU port of pine's traversal (pine-order stack traversal
    of a binned-SAH binary BVH) in this process, on the same rays: a second
    CPU figure beside pine's own BVH::Intersect."""
    rays, hits = trav["rays"], trav["hits"]
    tri = oracle.tri9(v, i)
    tb = time.time()
    cb = oracle.CpuBVH(tri)
    build_s = time.time() - tb
    prim, t, secs = cb.trace(rays, threads=threads)
    cpu = {"value": round(len(rays) / secs * 1e-6, 3), "unit": "Mrays/s", "cores": threads,
           "kind": "port", "host": share,
           "sample": f"all {len(rays)} rays of the stream, same mesh; pine-order BVH2 "
                     f"traversal, tree built in {build_s:.1f}s (not timed)",
           "agrees_with_gpu_t": round(float((t.view(np.int32) == hits[:, 1]).mean()), 6)}
    log(f"[bench] cpu port {cpu['value']} Mrays/s on {threads} threads ({secs:.2f}s)")
    return cpu


def queue_roofline(inf, samples, ms):
    """Whole-pipeline roofline of a fused path-tracing render from its
    queue-record bytes (QUEUE_BYTES_*) over the render's time on its stream."""
    ext = int(inf["last_extension_rays"])
    shadow = int(inf["last_rays"]) - ext
    b = samples * QUEUE_BYTES_SAMPLE + ext * QUEUE_BYTES_EXT + shadow * QUEUE_BYTES_SHADOW
    achieved = b / (ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes": int(b), "extension_rays": ext, "shadow_rays": shadow,
            "basis": f"queue records: {QUEUE_BYTES_SAMPLE} B per camera sample, "
                     f"{QUEUE_BYTES_EXT} B per extension ray, {QUEUE_BYTES_SHADOW} B per shadow "
                     f"ray, over the whole render time (all kernels)"}


def render_gpu(pg, text, directory):
    scene = pg.Scene(text=text, directory=directory)
    scene.render()  # warm, untimed: path-state slot sized, bounce loop recorded for the batch size
    inf = scene.render()
    scene.release()
    return inf


def batch_leg(args, pg, W, work):
    """The batched workload: 64 scenes of uneven size (W.batch_workload)
    rendered one by one (pg_render, graph replay), as one batch over concurrent
    streams (pg_render_batch), and as a batch with the launch-by-launch loop;
    each scene first autotuned (pg_scene_autotune)."""
    t0 = time.time()
    items = W.batch_workload(args.batch_scenes)
    W.write_box_obj(os.path.join(work, W.BOX_OBJ))
    scenes = []
    for it in items:
        v, i = W.make_mesh(it["mesh"])
        name = f"batch{it['index']}.obj"
        pg.write_obj_fixed(os.path.join(work, name), v, i)
        scenes.append(pg.Scene(text=W.batch_scene_text(it, name), directory=work))
        os.remove(os.path.join(work, name))
    setup_s = time.time() - t0
    log(f"[bench] batch: {len(scenes)} scenes, {sum(it['triangles'] for it in items)} triangles "
        f"loaded and built in {setup_s:.1f} s")
    t0 = time.time()
    tuned = [s.autotune(probe_spp=1, try_width=True) for s in scenes] if not args.no_tune else []
    tune_s = time.time() - t0
    samples = sum(it["width"] * it["height"] * it["spp"] for it in items)
    # Warm passes, untimed: size the path-state slots (slot 0 for renders one
    # by one, one per stream for the batch) and record every scene's bounce
    # loop for its batch size on each.
    for s in scenes:
        s.render()
    pg.hip.synchronize()
    t0 = time.perf_counter()
    for s in scenes:
        s.render()
    seq_s = time.perf_counter() - t0
    seq_rays = sum(s.info()["last_rays"] for s in scenes)
    pg.render_batch(scenes, streams=args.batch_streams)
    res = pg.render_batch(scenes, streams=args.batch_streams)
    ext = sum(int(s.info()["last_extension_rays"]) for s in scenes)
    roof = queue_roofline({"last_extension_rays": ext, "last_rays": res["rays"]}, samples, res["ms"])
    res_ng = pg.render_batch(scenes, streams=args.batch_streams, graph=False)
    for s in scenes:
        s.release()
    out = {
        "workload": f"{len(items)} scenes: displaced icosphere of log-uniform 10k..2M triangles "
                    f"and 4..24 instanced boxes in a seeded 16-sphere scene, 256^2..1024^2, "
                    f"16..128 spp, maxDepth 8",
        "triangles": int(sum(it["triangles"] for it in items)), "camera_samples": int(samples),
        "samples_per_s": round(samples / (res["ms"] * 1e-3), 1),
        "mrays_per_s": round(res["rays"] / (res["ms"] * 1e-3) * 1e-6, 2),
        "batch_ms": round(res["ms"], 2), "streams": args.batch_streams,
        "sequential_ms": round(seq_s * 1e3, 2),
        "sequential_samples_per_s": round(samples / seq_s, 1),
        "batch_launch_loop_ms": round(res_ng["ms"], 2),
        "speedup_vs_sequential": round(seq_s * 1e3 / res["ms"], 3),
        "graph_speedup": round(res_ng["ms"] / res["ms"], 3),
        "rays_match_sequential": bool(res["rays"] == seq_rays),
        "roofline": roof,
        "setup_s": round(setup_s, 1), "autotune_s": round(tune_s, 1),
        "tuned": {"trace_block": sorted({t["trace_block"] for t in tuned}),
                  "bvh_width": sorted({t["bvh_width"] for t in tuned}),
                  "postpone_leaves": sum(1 for t in tuned if t["trace_flags"] & 1),
                  "paths_in_flight": sorted({t["paths_in_flight"] for t in tuned}),
                  "probe_gain": round(float(np.mean([t["default_ms"] / t["best_ms"] for t in tuned])), 3)
                  if tuned else None},
    }
    log(f"[bench] batch: {out['samples_per_s'] / 1e6:.1f} M samples/s in {res['ms']:.0f} ms "
        f"({out['speedup_vs_sequential']}x one by one, {out['graph_speedup']}x the launch loop)")
    return out


GUIDED_RR_DEPTH = 4      # Russian roulette from depth 4 in the guided renders (pine: 2)
# Training schedule of the guiding leg: 6 passes of 2 spp, the last 3 kept in the image
# (time to the plain tracer's 512-spp error, 16-range reference: 8 x 4 spp kept 4: 330 ms;
# 8 x 3: 292; 6 x 3: 273; 8 x 2: 280; 4 x 4: 283; 6 x 2: 261).
GUIDED_TRAIN_DEFAULT = dict(train_passes=6, train_spp=2, keep_passes=3)


def guided_leg(args, pg, W, work):
    """Learned path guiding on the occluded-light box at equal time.  Three
    tracers, each at five sample counts, relMSE against a reference of the
    plain tracer (pine_amd.metrics: mean of 1024-spp ranges without singular
    ones): the plain tracer (pine's Russian roulette from depth 2), the plain
    tracer with roulette from depth 4, and the guided tracer (6 training
    passes of 2 spp by default, the last 3 kept in the image, roulette from
    depth 4) with the same samples per pixel in the image as the plain one.  Reported: time
    for each to reach the plain tracer's error at 512 spp, and the split of
    the gain into the later roulette (plain rr4 vs plain) and guiding itself
    (guided vs plain rr4)."""
    from pine_amd import metrics as M
    train_cfg = dict(train_passes=args.guide_passes, train_spp=args.guide_train_spp,
                        keep_passes=args.guide_keep)
    kept_spp = train_cfg["train_spp"] * train_cfg["keep_passes"]
    res = args.guide_res
    v, i = W.occluded_box_clutter()
    W.write_obj(os.path.join(work, "clutter.obj"), v, i)
    s = pg.Scene(text=W.occluded_box_scene_text("clutter.obj", width=res, height=res, spp=64,
                                                 max_depth=8), directory=work)
    t0 = time.time()
    ref = M.robust_reference(s, 1024, args.guide_ref_chunks,
                             progress=lambda c: log(f"[bench] guided: reference {c} x 1024 spp"))
    ref_s = time.time() - t0
    spps = (32, 64, 128, 256, 512, 1024)

    untrimmed = {}

    def score(key, err):
        img = M.linear(s.film())
        err.append(M.relmse(img, ref))
        untrimmed.setdefault(key, []).append(M.relmse_untrimmed(img, ref))

    def plain(rr):
        ms, err = [], []
        for spp in spps:
            s.render(spp=spp, sample_offset=1 << 20, rr_depth=rr)
            ms.append(s.info()["last_render_ms"])
            score(f"plain_rr{rr}", err)
        return ms, err

    pt, pe = plain(0)
    qt, qe = plain(GUIDED_RR_DEPTH)
    gt, ge, train_ms, step_ms = [], [], [], []
    for spp in spps:
        g = pg.Guide(seed=3, learning_rate=3e-3)
        inf = s.render_guided(g, final_spp=spp - kept_spp, sample_offset=1 << 21,
                              rr_depth=GUIDED_RR_DEPTH, steps_per_pass=args.guide_steps, **train_cfg)
        gt.append(inf["last_guided_ms"])
        train_ms.append(inf["last_guided_ms"] - inf["last_render_ms"])
        step_ms.append(inf["guide_step_ms"])
        score("guided", ge)
        g.release()
    # Guide inference on the matrix cores: one more guided render of the same
    # size drawn launch by launch, its inference kernels timed by HIP events.
    g = pg.Guide(seed=3, learning_rate=3e-3)
    inf = s.render_guided(g, final_spp=64 - kept_spp, sample_offset=1 << 21,
                          rr_depth=GUIDED_RR_DEPTH, graph=False, steps_per_pass=args.guide_steps,
                          **train_cfg)
    g.release()
    iq, ims = int(inf["guide_infer_queries"]), float(inf["guide_infer_ms"])
    tflops = iq * GUIDE_FLOP_PER_QUERY / (ims * 1e-3) / 1e12 if ims > 0 else None
    mfma = {"bound": "mfma", "achieved": round(tflops, 1) if tflops else None,
            "peak": MFMA_BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tflops / MFMA_BF16_PEAK_TFLOPS, 4) if tflops else None, "traffic": None,
            "kernel": "k_guide_infer", "queries": iq, "kernel_ms": round(ims, 2),
            "basis": f"{GUIDE_FLOP_PER_QUERY} FLOP per network evaluation (2 x 25,449 MACs) over "
                     f"the inference kernels' time in a 64-spp guided render"}
    s.release()
    target = pe[spps.index(512)]
    tp, tq, tg = (M.time_to_error(t, e, target) for t, e in ((pt, pe), (qt, qe), (gt, ge)))
    # A series already below the target at its cheapest point reports that
    # point's time: its speedup is then a lower bound.
    bracketed = all(e[0] > target for e in (pe, qe, ge))

    def series(ms, err, key):
        u = untrimmed[key]
        return {"ms": [round(x, 2) for x in ms], "relmse": [float(f"{x:.4g}") for x in err],
                "relmse_untrimmed": [float(f"{x:.4g}") for x, _ in u],
                "nonfinite_pixels": [n for _, n in u]}

    def ratio(a, b):
        return round(a / b, 3) if a and b else None

    out = {
        "workload": f"occluded-light box (512,000 clutter triangles), {res}x{res}, maxDepth 8; "
                    f"guide trained {train_cfg['train_passes']} passes x "
                    f"{train_cfg['train_spp']} spp, the last {train_cfg['keep_passes']} "
                    f"kept in the image; Russian roulette from depth {GUIDED_RR_DEPTH} when guided",
        "reference": f"plain tracer, {args.guide_ref_chunks} x 1024 spp ({ref_s:.0f} s), "
                     f"per-pixel mean of the ranges without singular ones",
        "metric": "relMSE = mean (img - ref)^2 / (ref^2 + 0.01) over pixels, the 0.1% largest "
                  "dropped (relmse_untrimmed: none dropped, non-finite pixels counted apart)",
        "spp_in_image": list(spps),
        "plain": series(pt, pe, "plain_rr0"),
        f"plain_rr{GUIDED_RR_DEPTH}": series(qt, qe, f"plain_rr{GUIDED_RR_DEPTH}"),
        "guided": dict(series(gt, ge, "guided"), final_spp=[x - kept_spp for x in spps],
                       training_ms=round(float(np.mean(train_ms)), 2),
                       training_adam_steps_ms=round(float(np.mean(step_ms)), 2)),
        "target_relmse": float(f"{target:.4g}"),
        "plain_ms_to_target": round(tp, 2) if tp else None,
        f"plain_rr{GUIDED_RR_DEPTH}_ms_to_target": round(tq, 2) if tq else None,
        "guided_ms_to_target": round(tg, 2) if tg else None,
        "equal_time_speedup": ratio(tp, tg),
        "speedups_are": "interpolated" if bracketed else "lower bounds (a series starts below the target)",
        "equal_time_speedup_from_roulette": ratio(tp, tq),
        "equal_time_speedup_from_guiding": ratio(tq, tg),
        "equal_spp_error_ratio": [round(a / b, 3) for a, b in zip(pe, ge)],
        # Untrimmed errors are listed per series but give no time-to-error:
        # single finite paths of ~1e9 radiance in the plain images put their
        # untrimmed relMSE at ~1e17 at random sample counts.
        "roofline": mfma,
    }
    log(f"[bench] guided: plain {tp:.1f} ms, plain rr{GUIDED_RR_DEPTH} {tq:.1f} ms, guided {tg:.1f} ms "
        f"to relMSE {target:.3g}" if tp and tq and tg else f"[bench] guided: target {target:.3g} not reached")
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=2)
    ap.add_argument("--block", type=int, default=128)
    ap.add_argument("--flags", type=int, default=8,
                    help="pg_accel_set_launch flags (1 postpone leaves, 2 XCD remap, 4 stackless, "
                         "8 rays reordered on the device); default 8; -1 = the library's default")
    ap.add_argument("--leaf", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-traffic", action="store_true",
                    help="skip the rocprofv3 traffic measurement of the traversal kernel")
    ap.add_argument("--no-pt", action="store_true", help="skip the path-tracing workloads")
    ap.add_argument("--pt-spp", type=int, default=64)
    ap.add_argument("--volume-spp", type=int, default=256)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched multi-scene leg")
    ap.add_argument("--batch-scenes", type=int, default=64)
    ap.add_argument("--batch-streams", type=int, default=4)
    ap.add_argument("--no-tune", action="store_true", help="batch leg without the autotuner")
    ap.add_argument("--no-guide", action="store_true", help="skip the path-guiding leg")
    ap.add_argument("--guide-res", type=int, default=1024,
                    help="occluded-box resolution of the guiding leg")
    ap.add_argument("--guide-ref-chunks", type=int, default=64,
                    help="1024-spp ranges in the guiding leg's reference")
    ap.add_argument("--guide-passes", type=int, default=GUIDED_TRAIN_DEFAULT["train_passes"],
                    help="training passes of the guiding leg")
    ap.add_argument("--guide-train-spp", type=int, default=GUIDED_TRAIN_DEFAULT["train_spp"],
                    help="samples per pixel of each training pass")
    ap.add_argument("--guide-keep", type=int, default=GUIDED_TRAIN_DEFAULT["keep_passes"],
                    help="last training passes kept in the image")
    ap.add_argument("--guide-steps", type=int, default=0,
                    help="Adam steps per training pass of the guiding leg (0: one epoch of its "
                         "records in at most 32 steps)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one device-to-device copy of the ray buffer (512 MiB), the "
                         "known-size kernel that scales rocprof's FETCH_SIZE/WRITE_SIZE")
    ap.add_argument("--traffic-probe", action="store_true", help=argparse.SUPPRESS)
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the traversal "
              "shards by rays with the hierarchy replicated, see DESIGN.md.", file=sys.stderr)
        sys.exit(2)
    if args.traffic_probe:
        traffic_probe_main(args)
        return

    # Host-side preparation, then every child process (pine's CPU baselines,
    # the rocprofv3 traffic runs), all before this process makes a HIP call.
    import pine_amd as pg
    from pine_amd import workloads as W
    threads, share = cpu_share()
    no_children = profiled()
    if no_children:
        log("[bench] under a profiler: no child processes (CPU baselines and traffic skipped)")
    cpu_legs = not args.no_cpu and os.path.exists(HARNESS) and not no_children
    work = tempfile.TemporaryDirectory()
    t0 = time.time()
    v, i = W.make_mesh(W.ICOSPHERE_1M)
    rays = W.make_rays(W.RAYS_16M, v, i)
    log(f"[bench] mesh and rays in {time.time() - t0:.2f}s: {len(i)} triangles, {len(rays)} rays")
    if not args.no_pt:
        W.write_obj(os.path.join(work.name, "ico1m.obj"), v, i)
        pg.write_volume(os.path.join(work.name, "cloud256.vol"), pg.gen_density_grid(W.CLOUD_GRID))
    pine = pine_baselines(args, W, work.name, v, i, rays, threads, share) if cpu_legs else {}
    traffic = None
    if not args.no_traffic and not no_children:
        t0 = time.time()
        traffic = traffic_probe(args)
        log(f"[bench] traffic measured in {time.time() - t0:.0f} s: "
            f"{traffic['bytes'] / 1e9:.2f} GB per launch, VALU busy "
            f"{traffic['valu']['busy'] if traffic['valu'] else None}" if traffic else "[bench] traffic: none")

    from pine_amd import hip
    if hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    trav = traversal(args, pg, hip, W, v, i, rays, traffic)

    pt = mesh_pt = vol = None
    if not args.no_pt:
        # Spheres scene.
        text = W.spheres_scene_text(seed=0, width=512, height=512, spp=args.pt_spp, max_depth=8)
        inf = render_gpu(pg, text, work.name)
        ms = inf["last_render_ms"]
        samples = 512 * 512 * args.pt_spp
        pt = {"workload": f"spheres scene 512x512, {args.pt_spp} spp, maxDepth 8",
              "samples_per_s": round(samples / (ms * 1e-3), 1),
              "render_ms": round(ms, 2),
              "mrays_per_s": round(inf["last_rays"] / (ms * 1e-3) * 1e-6, 2),
              "roofline": queue_roofline(inf, samples, ms)}
        log(f"[bench] spheres: {pt['samples_per_s'] / 1e6:.1f} M samples/s")
        # The same scene with four wavelengths per path (spectral mode).
        sc = pg.Scene(text=text, directory=work.name)
        sc.render(spectral=True)  # warm, untimed
        sc.render(spectral=True)
        sms = sc.info()["last_render_ms"]
        sc.release()
        pt["spectral"] = {"samples_per_s": round(samples / (sms * 1e-3), 1),
                          "render_ms": round(sms, 2), "wavelengths_per_path": 4}
        log(f"[bench] spheres, spectral: {pt['spectral']['samples_per_s'] / 1e6:.1f} M samples/s")
        # 2^20-triangle mesh scene.
        mesh_spp = 64
        text = W.mesh_scene_text("ico1m.obj", width=512, height=512, spp=mesh_spp)
        inf = render_gpu(pg, text, work.name)
        ms = inf["last_render_ms"]
        samples = 512 * 512 * mesh_spp
        mesh_pt = {"workload": f"2^20-triangle icosphere scene 512x512, {mesh_spp} spp, maxDepth 8",
                   "samples_per_s": round(samples / (ms * 1e-3), 1),
                   "render_ms": round(ms, 2),
                   "mrays_per_s": round(inf["last_rays"] / (ms * 1e-3) * 1e-6, 2),
                   "roofline": queue_roofline(inf, samples, ms)}
        log(f"[bench] mesh: {mesh_pt['samples_per_s'] / 1e6:.1f} M samples/s")
        # Volume: 256^3 grid with majorant bricks, lit by a directional sun and
        # a constant environment; the same scene under pine's Sky (sun and sky
        # gradient) is the one pine itself can render.
        text = W.cloud_scene_text("cloud256.vol", width=1024, height=1024, spp=args.volume_spp,
                                  max_depth=64, bricks=W.CLOUD_BRICK, lighting="sun+constant")
        inf = render_gpu(pg, text, work.name)
        ms = inf["last_render_ms"]
        lookups = inf["medium_steps"]
        achieved = lookups * (8 * 4 + 4) / (ms * 1e-3) / 1e9
        sky = render_gpu(pg, W.cloud_scene_text("cloud256.vol", width=1024, height=1024,
                                                spp=args.volume_spp, max_depth=64,
                                                bricks=W.CLOUD_BRICK), work.name)
        vol = {"workload": f"256^3 fBm grid, 1024x1024, {args.volume_spp} spp, maxDepth 64, "
                           f"{W.CLOUD_BRICK}^3 majorant bricks, directional sun + constant "
                           f"environment",
               "sky_variant_samples_per_s": round(1024 * 1024 * args.volume_spp /
                                                  (sky["last_render_ms"] * 1e-3), 1),
               "samples_per_s": round(1024 * 1024 * args.volume_spp / (ms * 1e-3), 1),
               "render_ms": round(ms, 2), "density_lookups": int(lookups),
               "mrays_per_s": round(inf["last_rays"] / (ms * 1e-3) * 1e-6, 2),
               "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "traffic": None,
                            "basis": "36 B per density lookup (8 taps + majorant) over the "
                                     "whole render time"}}
        log(f"[bench] volume: {vol['samples_per_s'] / 1e6:.1f} M samples/s, {lookups} lookups")

    batch = guided = None
    if not args.no_batch:
        batch = batch_leg(args, pg, W, work.name)
    if not args.no_guide:
        guided = guided_leg(args, pg, W, work.name)

    cpu = None
    if pine.get("trace"):
        r = pine["trace"]
        cpu = {"value": round(r["mrays_per_s"], 3), "unit": "Mrays/s", "cores": r["threads"],
               "kind": "reference", "host": share,
               "sample": f"pine's own BVH::Intersect (bvh.cpp:669-709) over all {r['rays']} rays "
                         f"of the stream on pine's own BVH of the same mesh ({r['trace_s']:.2f} s "
                         f"timed; its single-threaded build took {r['build_ms'] / 1e3:.0f} s, "
                         f"not timed), alone on the host before any GPU work"}
        log(f"[bench] pine BVH::Intersect: {cpu['value']} Mrays/s on {r['threads']} threads")
    for res, key, what, gpu_key in (
            (pt, "spheres", "same scene file", "samples_per_s"),
            (mesh_pt, "mesh", "same scene file", "samples_per_s"),
            (vol, "volume", "same grid, medium and film lit by pine's Sky (pine has no constant "
                            "environment light), pine's global majorant (no bricks), 8 spp; the "
                            "speedup is against the GPU's render of this Sky scene",
             "sky_variant_samples_per_s")):
        if res is not None and pine.get(key):
            res["cpu_baseline"] = cpu_baseline_from(pine[key], what, share)
            res["speedup_vs_cpu"] = round(res[gpu_key] / res["cpu_baseline"]["value"], 1)
            log(f"[bench] {res['workload'][:32]}: {res['speedup_vs_cpu']}x pine on "
                f"{res['cpu_baseline']['cores']} cores")
    port = None
    if not args.no_cpu:
        import oracle
        port = traversal_cpu(trav, oracle, v, i, threads, share)
        if cpu is None:
            cpu = port
    work.cleanup()

    out = {
        "metric": "Mrays/s at 1 GPU (closest hit)", "value": round(trav["mrays"], 2),
        "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(trav["ms_per_step"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": trav["config"],
        "roofline": trav["roofline"],
        "cpu_baseline": cpu,
        "cpu_port": port if port is not cpu else None,
        "path_tracing": pt,
        "mesh_path_tracing": mesh_pt,
        "volume": vol,
        "batch": batch,
        "guided": guided,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
