"""Plain and learned-guided path tracing of the occluded-light box, one after
the other, with their times: a small program to try guiding on a scene whose
light arrives through a gap, or to profile the guided kernels on their own.

    python -m pine_amd.guided [--res 1024] [--spp 64]

The guide trains for 8 passes of 4 spp (kept in the image), then renders
`spp` more samples; the plain render takes `spp` samples.  Both images are
unbiased estimates of the same scene.
"""
import argparse
import os
import tempfile


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--res", type=int, default=1024)
    ap.add_argument("--spp", type=int, default=64)
    args = ap.parse_args(argv)

    import pine_amd as pg
    from pine_amd import workloads as W

    work = tempfile.mkdtemp()
    v, i = W.occluded_box_clutter()
    W.write_obj(os.path.join(work, "clutter.obj"), v, i)
    s = pg.Scene(text=W.occluded_box_scene_text("clutter.obj", width=args.res, height=args.res,
                                                 spp=args.spp, max_depth=8), directory=work)
    s.render(spp=args.spp, sample_offset=1 << 20)
    print(f"plain: {args.spp} spp in {s.info()['last_render_ms']:.1f} ms", flush=True)
    g = pg.Guide(seed=3, learning_rate=3e-3)
    inf = s.render_guided(g, train_passes=8, train_spp=4, final_spp=args.spp, sample_offset=1 << 21,
                          keep_passes=8)
    print(f"guided: 32 training + {args.spp} spp in {inf['last_guided_ms']:.1f} ms "
          f"(final render {inf['last_render_ms']:.1f} ms)", flush=True)
    g.release()
    s.release()


if __name__ == "__main__":
    main()
