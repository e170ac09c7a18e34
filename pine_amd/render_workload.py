"""Renders one benchmark workload in a process of its own and prints its
rate: for profiling a workload's kernels without the rest of the benchmark.
The workload renders twice, a warm render and the timed one, both at full
size.

    python -m pine_amd.render_workload {spheres,mesh,cloud,box} [--spp N] [--res R]

spheres  the seeded 64-spheres scene (default 512x512, 64 spp, maxDepth 8)
mesh     the 2^20-triangle displaced icosphere in a lit scene (512x512, 64 spp)
cloud    the 256^3 fBm density grid with 16^3-voxel majorant bricks, directional
         sun and constant environment (1024x1024, 256 spp, maxDepth 64)
box      the occluded-light box with 512,000 clutter triangles (1024x1024, 64 spp)
"""
import argparse
import os
import tempfile


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("workload", choices=["spheres", "mesh", "cloud", "box"])
    ap.add_argument("--spp", type=int, default=0)
    ap.add_argument("--res", type=int, default=0)
    args = ap.parse_args(argv)

    import pine_amd as pg
    from pine_amd import workloads as W

    work = tempfile.mkdtemp()
    if args.workload == "spheres":
        res, spp = args.res or 512, args.spp or 64
        text = W.spheres_scene_text(seed=0, width=res, height=res, spp=spp, max_depth=8)
    elif args.workload == "mesh":
        res, spp = args.res or 512, args.spp or 64
        v, i = W.make_mesh(W.ICOSPHERE_1M)
        W.write_obj(os.path.join(work, "ico1m.obj"), v, i)
        text = W.mesh_scene_text("ico1m.obj", width=res, height=res, spp=spp)
    elif args.workload == "cloud":
        res, spp = args.res or 1024, args.spp or 256
        pg.write_volume(os.path.join(work, "cloud256.vol"), pg.gen_density_grid(W.CLOUD_GRID))
        text = W.cloud_scene_text("cloud256.vol", width=res, height=res, spp=spp, max_depth=64,
                                  bricks=W.CLOUD_BRICK, lighting="sun+constant")
    else:
        res, spp = args.res or 1024, args.spp or 64
        v, i = W.occluded_box_clutter()
        W.write_obj(os.path.join(work, "clutter.obj"), v, i)
        text = W.occluded_box_scene_text("clutter.obj", width=res, height=res, spp=spp, max_depth=8)
    s = pg.Scene(text=text, directory=work)
    s.render()  # warm: sizes the path state and records the bounce loop for the batch size
    inf = s.render()
    ms = inf["last_render_ms"]
    print(f"{args.workload} {res}x{res}, {spp} spp: {ms:.1f} ms, "
          f"{res * res * spp / (ms * 1e-3) * 1e-6:.1f} M samples/s, {inf['last_rays']} rays, "
          f"{inf['medium_steps']} density lookups", flush=True)
    s.release()


if __name__ == "__main__":
    main()
