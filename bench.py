"""Benchmark of pine's path-tracing hot path on one MI355X.

Headline workload (`value`, Mrays/s): closest-hit BVH traversal of a
1,048,576-triangle displaced geodesic icosphere by 16,777,216 rays (4,194,304
primary rays of a 2048x2048 pinhole camera + 12,582,912 cosine-distributed
rays leaving seeded surface points).  One step = one pg_trace_closest pass
over all rays, rays and hierarchy resident in HBM.

    python bench.py [--gpus 1] [--steps 20] [--warmup 3] [--width 2] [--block 128]

Prints one JSON line.  The roofline object accounts the dominant kernel
(k_trace): algorithmic bytes per launch = 40 B compulsory per ray (32 B ray
read + 8 B hit written) + node bytes x nodes fetched + record bytes x prims
tested, the counts taken from the instrumented variant of the same kernel over
the same rays; duration = average of HIP events around each launch on the
launch stream.  cpu_baseline times the oracle's CPU port of pine's traversal
(pine-order stack traversal of a binned-SAH binary BVH) on the host cores.

Secondary workloads, each next to pine itself (oracle/_ref/ref_harness, pine's
own PathIntegrator built from its sources) on the host cores, at matched
relMSE (the GPU reproduces pine's per-pixel sample sequence):
  path_tracing       the seeded 64-spheres scene, 512x512, 64 spp, maxDepth 8
  mesh_path_tracing  the 2^20-triangle icosphere in a lit scene, 512x512
  volume             a 256^3 fBm density grid, 1024x1024, 256 spp, maxDepth 64,
                     16^3-voxel majorant bricks, HG g = 0.85, albedo 0.99
"""
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
HARNESS = os.path.join(ROOT, "oracle", "_ref", "ref_harness")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_threads():
    return min(os.cpu_count() or 1, 16)


class PineRun:
    """pine's PathIntegrator on a scene file, in a child process (the harness's
    render mode); its samples/s excludes scene loading and BVH build, as pine's
    own [Rendering]Average line does."""

    def __init__(self, scene_text, files=(), threads=None):
        self.dir = tempfile.TemporaryDirectory()
        for name, writer in files:
            writer(os.path.join(self.dir.name, name))
        with open(os.path.join(self.dir.name, "scene.txt"), "w") as f:
            f.write(scene_text)
        self.threads = threads or cpu_threads()
        self.t0 = time.time()
        self.p = subprocess.Popen([HARNESS, "render", "./scene.txt", "film.bin", str(self.threads)],
                                  cwd=self.dir.name, stdout=subprocess.PIPE,
                                  stderr=subprocess.PIPE, text=True)

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
            import shutil
            shutil.copyfile(src, dst)
    return w


def cpu_baseline_from(r, what):
    if r is None:
        return None
    n = r["w"] * r["h"] * r["spp"]
    return {"value": round(r["msamples_per_s"] * 1e6, 1), "unit": "samples/s",
            "cores": r["threads"], "kind": "reference",
            "sample": f"{what}: pine renders {r['w']}x{r['h']} at {r['spp']} spp ({n} samples, "
                      f"{r['render_s']:.1f}s render, {r['wall_s']:.0f}s with scene load and BVH build)"}


def traversal(args, pg, hip, W, v, i):
    t0 = time.time()
    rays = W.make_rays(W.RAYS_16M, v, i)
    log(f"[bench] rays generated in {time.time() - t0:.2f}s: {len(i)} triangles, {len(rays)} rays")

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
    d_rays.free()
    d_hits.free()
    acc.release()

    achieved = algo_bytes / (kern_ms * 1e-3) / 1e9
    roofline = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": measured_traffic(st["width"], args.leaf, args.block, launch_flags),
        "kernel": f"k_trace<W={st['width']}>",
        "bytes_per_launch": {"compulsory": compulsory, "traversal": traversal_bytes},
        "compulsory_frac": round(compulsory / (kern_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 5),
        "per_ray": {"nodes": round(nodes_fetched / n, 3), "prims": round(prims_tested / n, 3)},
    }
    config = {"workload": "bvh-traversal: icosphere 1,048,576 tris, 16,777,216 rays",
              "triangles": int(len(i)), "rays": int(n), "node_width": st["width"],
              "leaf_max": args.leaf, "block": args.block, "launch_flags": launch_flags,
              "hit_fraction": round(hit_frac, 4), "build_ms": round(st["build_ms"], 1)}
    return {"mrays": mrays, "ms_per_step": ms_per_step, "roofline": roofline, "config": config,
            "rays": rays, "hits": hits}


def traversal_cpu(trav, oracle, v, i):
    rays, hits = trav["rays"], trav["hits"]
    threads = cpu_threads()
    tri = oracle.tri9(v, i)
    tb = time.time()
    cb = oracle.CpuBVH(tri)
    build_s = time.time() - tb
    prim, t, secs = cb.trace(rays, threads=threads)
    cpu = {"value": round(len(rays) / secs * 1e-6, 3), "unit": "Mrays/s", "cores": threads,
           "kind": "port",
           "sample": f"all {len(rays)} rays of the stream, same mesh; pine-order BVH2 "
                     f"traversal, tree built in {build_s:.1f}s (not timed)",
           "agrees_with_gpu_t": round(float((t.view(np.int32) == hits[:, 1]).mean()), 6)}
    log(f"[bench] cpu baseline {cpu['value']} Mrays/s on {threads} threads ({secs:.2f}s)")
    return cpu


def render_gpu(pg, text, directory, warm_spp=2):
    scene = pg.Scene(text=text, directory=directory)
    scene.render(spp=warm_spp)  # allocations and first launches
    inf = scene.render()
    scene.release()
    return inf


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
    ap.add_argument("--no-pt", action="store_true", help="skip the path-tracing workloads")
    ap.add_argument("--pt-spp", type=int, default=64)
    ap.add_argument("--volume-spp", type=int, default=256)
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

    cpu_legs = not args.no_cpu and os.path.exists(HARNESS)
    work = tempfile.TemporaryDirectory()
    v, i = W.make_mesh(W.ICOSPHERE_1M)
    # pine's mesh scene spends about two minutes in its single-threaded BVH
    # build; it runs beside the GPU-only measurements, and the CPU-heavy legs
    # wait for it.
    obj = os.path.join(work.name, "ico1m.obj")
    if not args.no_pt:
        W.write_obj(obj, v, i)
    pine_mesh = None
    if cpu_legs and not args.no_pt:
        pine_mesh = PineRun(W.mesh_scene_text("ico1m.obj", width=128, height=128, spp=64),
                            files=[("ico1m.obj", place(obj))])

    trav = traversal(args, pg, hip, W, v, i)

    pt = mesh_pt = vol = None
    if not args.no_pt:
        # Spheres scene.
        text = W.spheres_scene_text(seed=0, width=512, height=512, spp=args.pt_spp, max_depth=8)
        inf = render_gpu(pg, text, work.name)
        ms = inf["last_render_ms"]
        pt = {"workload": f"spheres scene 512x512, {args.pt_spp} spp, maxDepth 8",
              "samples_per_s": round(512 * 512 * args.pt_spp / (ms * 1e-3), 1),
              "render_ms": round(ms, 2),
              "mrays_per_s": round(inf["last_rays"] / (ms * 1e-3) * 1e-6, 2)}
        log(f"[bench] spheres: {pt['samples_per_s'] / 1e6:.1f} M samples/s")
        # 2^20-triangle mesh scene.
        mesh_spp = 64
        text = W.mesh_scene_text("ico1m.obj", width=512, height=512, spp=mesh_spp)
        inf = render_gpu(pg, text, work.name)
        ms = inf["last_render_ms"]
        mesh_pt = {"workload": f"2^20-triangle icosphere scene 512x512, {mesh_spp} spp, maxDepth 8",
                   "samples_per_s": round(512 * 512 * mesh_spp / (ms * 1e-3), 1),
                   "render_ms": round(ms, 2),
                   "mrays_per_s": round(inf["last_rays"] / (ms * 1e-3) * 1e-6, 2)}
        log(f"[bench] mesh: {mesh_pt['samples_per_s'] / 1e6:.1f} M samples/s")
        # Volume: 256^3 grid with majorant bricks.
        grid = pg.gen_density_grid(W.CLOUD_GRID)
        pg.write_volume(os.path.join(work.name, "cloud256.vol"), grid)
        text = W.cloud_scene_text("cloud256.vol", width=1024, height=1024, spp=args.volume_spp,
                                  max_depth=64, bricks=W.CLOUD_BRICK)
        inf = render_gpu(pg, text, work.name)
        ms = inf["last_render_ms"]
        lookups = inf["medium_steps"]
        achieved = lookups * (8 * 4 + 4) / (ms * 1e-3) / 1e9
        vol = {"workload": f"256^3 fBm grid, 1024x1024, {args.volume_spp} spp, maxDepth 64, "
                           f"{W.CLOUD_BRICK}^3 majorant bricks",
               "samples_per_s": round(1024 * 1024 * args.volume_spp / (ms * 1e-3), 1),
               "render_ms": round(ms, 2), "density_lookups": int(lookups),
               "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                            "basis": "36 B per density lookup (8 taps + majorant) over the "
                                     "whole render time"}}
        log(f"[bench] volume: {vol['samples_per_s'] / 1e6:.1f} M samples/s, {lookups} lookups")

    cpu = None
    if not args.no_cpu:
        import oracle
        if pine_mesh is not None:
            mesh_pt["cpu_baseline"] = cpu_baseline_from(pine_mesh.result(),
                                                        "same mesh scene at lower resolution")
        if pt is not None and cpu_legs:
            text = W.spheres_scene_text(seed=0, width=512, height=512, spp=args.pt_spp, max_depth=8)
            pt["cpu_baseline"] = cpu_baseline_from(PineRun(text).result(), "same scene file")
            vtext = W.cloud_scene_text("cloud256.vol", width=256, height=256, spp=16, max_depth=64)
            vol_file = os.path.join(work.name, "cloud256.vol")
            r = PineRun(vtext, files=[("cloud256.vol", place(vol_file))]).result()
            vol["cpu_baseline"] = cpu_baseline_from(r, "same grid and medium with pine's global "
                                                       "majorant, lower resolution and spp")
        for res in (pt, mesh_pt, vol):
            if res and res.get("cpu_baseline"):
                res["speedup_vs_cpu"] = round(res["samples_per_s"] / res["cpu_baseline"]["value"], 1)
                log(f"[bench] {res['workload'][:32]}: {res['speedup_vs_cpu']}x pine on "
                    f"{res['cpu_baseline']['cores']} cores")
        cpu = traversal_cpu(trav, oracle, v, i)
    work.cleanup()

    out = {
        "metric": "Mrays/s at 1 GPU (closest hit)", "value": round(trav["mrays"], 2),
        "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(trav["ms_per_step"], 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": trav["config"],
        "roofline": trav["roofline"],
        "cpu_baseline": cpu,
        "path_tracing": pt,
        "mesh_path_tracing": mesh_pt,
        "volume": vol,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
