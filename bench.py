"""Benchmark: closest-hit BVH traversal on one MI355X.

Workload (the traversal microbenchmark): a 1,048,576-triangle displaced
geodesic icosphere and 16,777,216 rays (4,194,304 primary rays of a 2048x2048
pinhole camera + 12,582,912 cosine-distributed rays leaving seeded surface
points).  One step = one pg_trace_closest pass over all rays, rays and
hierarchy resident in HBM.  value = Mrays/s.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--width 2] [--block 128]

Prints one JSON line.  The roofline object accounts the dominant kernel
(k_trace*): algorithmic bytes per launch = 40 B compulsory per ray (32 B ray
read + 8 B hit written) + node bytes x nodes fetched + record bytes x prims
tested, the counts taken from the instrumented variant of the same kernel
over the same rays; duration = average of HIP events around each launch on the
launch stream.  cpu_baseline times the oracle's CPU port of pine's traversal
(pine-order stack traversal of a binned-SAH binary BVH) on all host cores.
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=2)
    ap.add_argument("--block", type=int, default=128)
    ap.add_argument("--flags", type=int, default=-1,
                    help="pg_accel_set_launch flags (1 postpone leaves, 2 XCD remap); -1 = default")
    ap.add_argument("--leaf", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline legs")
    ap.add_argument("--no-pt", action="store_true", help="skip the path-tracing measurement")
    ap.add_argument("--pt-spp", type=int, default=64)
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run one device-to-device copy of the ray buffer (512 MiB), the "
                         "known-size kernel that scales rocprof's FETCH_SIZE/WRITE_SIZE")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the traversal "
              "shards by rays with the hierarchy replicated, see DESIGN.md.", file=sys.stderr)
        sys.exit(2)

    import pine_amd as pg
    from pine_amd import hip
    from pine_amd import workloads as W

    if hip.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)

    t0 = time.time()
    v, i = W.make_mesh(W.ICOSPHERE_1M)
    rays = W.make_rays(W.RAYS_16M, v, i)
    log(f"[bench] workload generated in {time.time() - t0:.2f}s: {len(i)} triangles, {len(rays)} rays")

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
    prim_bytes = 48
    compulsory = 40 * n
    traversal = nodes_fetched * node_bytes + prims_tested * prim_bytes
    algo_bytes = compulsory + traversal

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
    kern_ms = float(np.mean([s.elapsed_ms(e) for s, e in zip(starts, ends)]))
    ms_per_step = wall * 1e3 / args.steps
    mrays = n / (ms_per_step * 1e-3) * 1e-6
    hits = d_hits.numpy()
    hit_frac = float((hits[:, 0] >= 0).mean())
    if args.copy_calib:
        dst = hip.DeviceArray(d_rays.shape, np.float32)
        hip.check(hip.rt().hipMemcpy(dst.ptr, d_rays.ptr, d_rays.nbytes, hip.hipMemcpyDeviceToDevice))
        hip.synchronize()
        dst.free()
    log(f"[bench] {ms_per_step:.3f} ms/step, kernel {kern_ms:.3f} ms, {mrays:.1f} Mrays/s, "
        f"hit fraction {hit_frac:.3f}")

    achieved = algo_bytes / (kern_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(st["width"], args.leaf, args.block, launch_flags),
        "kernel": f"k_trace{st['width']}",
        "bytes_per_launch": {"compulsory": compulsory, "traversal": traversal},
        "compulsory_frac": round(compulsory / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "per_ray": {"nodes": round(nodes_fetched / n, 3), "prims": round(prims_tested / n, 3)},
    }

    cpu = None
    if not args.no_cpu:
        import oracle
        threads = min(os.cpu_count() or 1, 16)
        tri = oracle.tri9(v, i)
        tb = time.time()
        cb = oracle.CpuBVH(tri)
        build_s = time.time() - tb
        sel = np.arange(n)  # the whole stream
        prim, t, secs = cb.trace(rays[sel], threads=threads)
        cpu = {"value": round(len(sel) / secs * 1e-6, 3), "unit": "Mrays/s", "cores": threads,
               "kind": "port",
               "sample": f"all {len(sel)} rays of the stream, same mesh; pine-order BVH2 "
                         f"traversal, tree built in {build_s:.1f}s (not timed)"}
        # Sanity: same closest distances as the GPU on the sample.
        same_t = float((t.view(np.int32) == hits[sel, 1]).mean())
        cpu["agrees_with_gpu_t"] = round(same_t, 6)
        log(f"[bench] cpu baseline {cpu['value']} Mrays/s on {threads} threads ({secs:.2f}s)")

    pt = None if args.no_pt else path_tracing(args, pg, W)

    out = {
        "metric": "Mrays/s at 1 GPU (closest hit)", "value": round(mrays, 2), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "bvh-traversal: icosphere 1,048,576 tris, 16,777,216 rays",
                   "triangles": int(len(i)), "rays": int(n), "node_width": st["width"],
                   "leaf_max": args.leaf, "block": args.block, "launch_flags": launch_flags, "hit_fraction": round(hit_frac, 4),
                   "build_ms": round(st["build_ms"], 1)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "path_tracing": pt,
    }
    print(json.dumps(out), flush=True)


def measured_traffic(width, leaf, block, flags):
    """HBM bytes per launch of the trace kernel for this launch configuration,
    as measured by rocprofv3 FETCH_SIZE + WRITE_SIZE (scaled by a known-size
    copy) and recorded in the kernel's profile summary; None when that
    configuration has no recorded measurement."""
    import re
    path = os.path.join(ROOT, "profiles", "k_trace2_icosphere1m_rays16m.md")
    if not os.path.exists(path):
        return None
    key = f"width={width} leaf={leaf} block={block} flags={flags}"
    for line in open(path):
        m = re.match(r"\s*traffic_bytes_per_launch k_trace\S* (.*): (\d+)\s*$", line)
        if m and m.group(1).strip() == key:
            return int(m.group(2))
    return None


def path_tracing(args, pg, W):
    """Secondary measurement: samples/s of pine's Path integrator on the seeded
    spheres scene (64 spheres, one area light, 512x512, maxDepth 8), beside pine
    itself (oracle/_ref/ref_harness render mode) on the host cores."""
    import subprocess
    import tempfile
    text = W.spheres_scene_text(seed=0, width=512, height=512, spp=args.pt_spp, max_depth=8)
    scene = pg.Scene(text=text)
    scene.render(spp=4)  # warm-up: allocations and first launches
    inf = scene.render()
    ms = inf["last_render_ms"]
    samples = 512 * 512 * args.pt_spp
    res = {"workload": f"spheres scene 512x512, {args.pt_spp} spp, maxDepth 8",
           "samples_per_s": round(samples / (ms * 1e-3), 1), "render_ms": round(ms, 2),
           "mrays_per_s": round(inf["last_rays"] / (ms * 1e-3) * 1e-6, 2)}
    scene.release()
    log(f"[bench] path tracing {res['samples_per_s'] / 1e6:.2f} M samples/s "
        f"({res['mrays_per_s']} Mrays/s incl. shadow rays)")
    harness = os.path.join(ROOT, "oracle", "_ref", "ref_harness")
    if not args.no_cpu and os.path.exists(harness):
        threads = min(os.cpu_count() or 1, 16)
        with tempfile.TemporaryDirectory() as d:
            with open(os.path.join(d, "scene.txt"), "w") as f:
                f.write(text)
            p = subprocess.run([harness, "render", "./scene.txt", os.path.join(d, "film.bin"),
                                str(threads)], capture_output=True, text=True, cwd=d, timeout=300)
        line = [l for l in p.stdout.splitlines() if l.startswith("{")]
        if p.returncode == 0 and line:
            ref = json.loads(line[-1])
            cps = ref["msamples_per_s"] * 1e6
            res["cpu_baseline"] = {"value": round(cps, 1), "unit": "samples/s", "cores": threads,
                                   "kind": "reference",
                                   "sample": "same scene file rendered by pine's PathIntegrator "
                                             f"({ref['render_s']:.1f}s)"}
            res["speedup_vs_cpu"] = round(res["samples_per_s"] / cps, 1)
            log(f"[bench] pine on {threads} cores: {cps / 1e6:.3f} M samples/s")
    return res


if __name__ == "__main__":
    main()
