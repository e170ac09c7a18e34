"""Benchmark workloads, named by what they are.

``ICOSPHERE_1M``: geodesic icosphere of frequency 229 (20*229^2 = 1,048,820
triangles) with its last 244 triangles dropped, giving exactly 2^20 =
1,048,576 triangles, displaced radially by seeded 5-octave value-noise fBm of
amplitude 0.1.  (20*n^2 never equals 2^20, so a truncated frequency-229 sphere
is the closest geodesic construction.)

``RAYS_16M``: 4,194,304 primary rays of a 2048x2048 pinhole camera at
(0, 0.6, 3) looking at the origin with a 45 degree vertical field of view,
followed by 12,582,912 cosine-distributed rays leaving seeded points of the
mesh, origins offset by pine's OffsetRayOrigin.
"""
from dataclasses import dataclass

import numpy as np


@dataclass(frozen=True)
class MeshSpec:
    freq: int
    amp: float
    seed: int
    max_triangles: int


@dataclass(frozen=True)
class RaySpec:
    res: int
    eye: tuple
    fov_deg: float
    n_bounce: int
    seed: int

    @property
    def n(self):
        return self.res * self.res + self.n_bounce


ICOSPHERE_1M = MeshSpec(freq=229, amp=0.1, seed=1, max_triangles=1 << 20)
RAYS_16M = RaySpec(res=2048, eye=(0.0, 0.6, 3.0), fov_deg=45.0, n_bounce=12 << 20, seed=7)

# Smaller cases of the same construction for parity tests.
ICOSPHERE_5K = MeshSpec(freq=16, amp=0.1, seed=1, max_triangles=0)
ICOSPHERE_82K = MeshSpec(freq=64, amp=0.1, seed=1, max_triangles=0)
RAYS_SMALL = RaySpec(res=96, eye=(0.0, 0.6, 3.0), fov_deg=45.0, n_bounce=8192, seed=7)


def make_mesh(spec):
    from . import gen_icosphere
    return gen_icosphere(spec.freq, spec.amp, spec.seed, spec.max_triangles)


def make_rays(spec, vertices, indices):
    from . import gen_rays
    return gen_rays(vertices, indices, spec.res, spec.eye, spec.fov_deg, spec.n_bounce, spec.seed)


def spheres_scene(seed=0, n_spheres=64):
    """Seeded spheres scene: n_spheres spheres with radius U(0.2, 1.0) inside a
    20^3 box standing on a 100x100 ground rect, one 6x6 rect area light of
    emission 10 above them.  Materials: 50% diffuse (albedo U(0.1, 0.9)), 30%
    dielectric (IOR 1.5, roughness 0), 20% conductor (roughness U(0.05, 0.5),
    albedo U(0.5, 0.95)).  Returns a dict of plain lists (scene-file terms)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    spheres = []
    for k in range(n_spheres):
        r = float(rng.uniform(0.2, 1.0))
        c = (float(rng.uniform(-10, 10)), float(r + rng.uniform(0, 20 - 2 * r)),
             float(rng.uniform(-10, 10)))
        u = float(rng.uniform())
        if u < 0.5:
            mat = {"type": "Diffuse", "albedo": float(rng.uniform(0.1, 0.9))}
        elif u < 0.8:
            mat = {"type": "Dielectric", "eta": 1.5, "roughness": 0.0}
        else:
            mat = {"type": "Conductor", "roughness": float(rng.uniform(0.05, 0.5)),
                   "albedo": float(rng.uniform(0.5, 0.95))}
        spheres.append({"c": c, "r": r, "material": mat})
    return {
        "spheres": spheres,
        # Rect{position, ex, ey}: normal = normalize(cross(ex, ey)) = +y
        "ground": {"position": (0.0, 0.0, 0.0), "ex": (0.0, 0.0, 100.0), "ey": (100.0, 0.0, 0.0),
                   "material": {"type": "Diffuse", "albedo": 0.5}},
        # Light faces down: cross(ex, ey) = -y
        "light": {"position": (0.0, 24.0, 0.0), "ex": (6.0, 0.0, 0.0), "ey": (0.0, 0.0, 6.0),
                  "emission": 10.0},
        "camera": {"from": (0.0, 12.0, 38.0), "to": (0.0, 6.0, 0.0), "fov": 0.9},
    }


def _fmt(x):
    # Plain decimals only: pine's stof (pstd/string.cpp:50-68) has no exponents.
    s = f"{float(x):.6f}".rstrip("0").rstrip(".")
    return s if s not in ("", "-0") else "0"


def _v3(v):
    return " ".join(_fmt(x) for x in v)


def spheres_scene_text(seed=0, width=512, height=512, spp=64, max_depth=8, n_spheres=64,
                       output="spheres.bmp", sampler="Halton", randomize=None, environment=None,
                       area_light=True):
    """The seeded spheres scene as pine scene-file text (parser.cpp grammar);
    `sampler` names pine's sampler type, `randomize` its randomizeStrategy;
    `environment` (a dict: "type" plus parameters) adds an Environment light,
    `area_light` False drops the emissive rect."""
    sc = spheres_scene(seed, n_spheres)
    cam = sc["camera"]
    out = [
        "Integrator: Path{",
        f"    maxDepth: {max_depth}",
        f"    sampler: {sampler}{{",
        f"        samplesPerPixel: {spp}",
    ] + ([f"        randomizeStrategy: {randomize}"] if randomize else []) + [
        "    }",
        "}",
        "Camera: ThinLen{",
        "    film{",
        f"        outputFileName: {output}",
        f"        size: {width} {height}",
        "        filter: LanczosSinc{",
        "            radius: 0.5",
        "        }",
        "    }",
        f"    from: {_v3(cam['from'])}",
        f"    to: {_v3(cam['to'])}",
        f"    fov: {_fmt(cam['fov'])}",
        "}",
        "Material light: Emissive{",
        f"    color: {_fmt(sc['light']['emission'])}",
        "}",
        "Material ground: Layered{",
        "    layer0: Diffuse{",
        f"        albedo: {_fmt(sc['ground']['material']['albedo'])}",
        "    }",
        "}",
    ]
    for k, s in enumerate(sc["spheres"]):
        m = s["material"]
        out.append(f"Material m{k}: Layered{{")
        if m["type"] == "Diffuse":
            out += ["    layer0: Diffuse{", f"        albedo: {_fmt(m['albedo'])}", "    }"]
        elif m["type"] == "Dielectric":
            out += ["    layer0: Dielectric{", f"        eta: {_fmt(m['eta'])}",
                    f"        roughness: {_fmt(m['roughness'])}", "    }"]
        else:
            out += ["    layer0: Conductor{", f"        albedo: {_fmt(m['albedo'])}",
                    f"        roughness: {_fmt(m['roughness'])}", "    }"]
        out.append("}")
    g, li = sc["ground"], sc["light"]
    out += ["Shape: Rect{", f"    position: {_v3(g['position'])}", f"    ex: {_v3(g['ex'])}",
            f"    ey: {_v3(g['ey'])}", "    material: ground", "}"]
    if area_light:
        out += ["Shape: Rect{", f"    position: {_v3(li['position'])}", f"    ex: {_v3(li['ex'])}",
                f"    ey: {_v3(li['ey'])}", "    material: light", "}"]
    if environment:
        out += ["Light: Environment{", f"    environment: {environment['type']}{{"]
        out += [f"        {k}: {v}" for k, v in environment.items() if k != "type"]
        out += ["    }", "}"]
    for k, s in enumerate(sc["spheres"]):
        out += ["Shape: Sphere{", f"    position: {_v3(s['c'])}", f"    radius: {_fmt(s['r'])}",
                f"    material: m{k}", "}"]
    return "\n".join(out) + "\n"


BATCH_SCENES = 64


def batch_workload(n=BATCH_SCENES, seed=7, min_triangles=10_000, max_triangles=2_000_000):
    """The batched-rendering workload: n scenes of uneven size.  Scene k is the
    seeded spheres scene (materials and light as the spheres workload, 16
    spheres) with a displaced icosphere at its centre whose triangle count is
    log-uniform in [min_triangles, max_triangles], plus 4..24 instances of one
    box mesh (BOX_OBJ, 12 triangles) standing on the ground, each scaled
    U(0.3, 2.5) per axis and given one of the scene's sphere materials;
    square resolution uniform in {256, 384, ..., 1024}, spp uniform in
    {16, 24, ..., 128}, maxDepth 8.  Returns one dict per scene;
    batch_scene_text writes its scene file."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n):
        tris = int(round(float(np.exp(rng.uniform(np.log(min_triangles), np.log(max_triangles))))))
        res = int(rng.choice(np.arange(256, 1025, 128)))
        spp = int(rng.choice(np.arange(16, 129, 8)))
        freq = int(np.ceil(np.sqrt(tris / 20.0)))
        brng = np.random.default_rng(seed * 1000 + k + 500_000)
        boxes = []
        for _ in range(int(brng.integers(4, 25))):
            sc = brng.uniform(0.3, 2.5, 3)
            pos = (float(brng.uniform(-12, 12)), float(sc[1] / 2), float(brng.uniform(-12, 12)))
            boxes.append(dict(scale=tuple(float(x) for x in sc), position=pos,
                              material=int(brng.integers(0, 16))))
        out.append(dict(index=k, triangles=tris, mesh=MeshSpec(freq, 0.1, seed * 1000 + k, tris),
                        width=res, height=res, spp=spp, scene_seed=seed * 1000 + k, boxes=boxes))
    return out


def reduced_batch_item(item, res=64, spp=32, max_triangles=40_000):
    """A batch_workload item at a smaller film, spp and mesh (the icosphere
    truncated to max_triangles), same scene otherwise: for parity renders
    that pine itself finishes in seconds."""
    it = dict(item)
    tris = min(item["triangles"], max_triangles)
    m = item["mesh"]
    it.update(width=res, height=res, spp=spp, triangles=tris,
              mesh=MeshSpec(m.freq, m.amp, m.seed, tris))
    return it


BOX_OBJ = "box.obj"


def write_box_obj(path):
    """The unit cube [-0.5, 0.5]^3 as 8 vertices and 12 outward triangles."""
    v = [(x, y, z) for x in (-0.5, 0.5) for y in (-0.5, 0.5) for z in (-0.5, 0.5)]
    f = [(1, 2, 4), (1, 4, 3), (5, 7, 8), (5, 8, 6), (1, 5, 6), (1, 6, 2),
         (3, 4, 8), (3, 8, 7), (1, 3, 7), (1, 7, 5), (2, 6, 8), (2, 8, 4)]
    with open(path, "w") as fh:
        for p in v:
            fh.write("v " + " ".join(_fmt(x) for x in p) + "\n")
        for t in f:
            fh.write("f %d %d %d\n" % t)


def batch_scene_text(item, obj_file, box_file=BOX_OBJ):
    """Scene file of one batch_workload item (its icosphere in obj_file, the
    instanced box in box_file)."""
    text = spheres_scene_text(seed=item["scene_seed"], width=item["width"], height=item["height"],
                              spp=item["spp"], max_depth=8, n_spheres=16,
                              output=f"batch{item['index']}.bmp")
    out = ["Material batch_mesh: Layered{", "    layer0: Diffuse{", "        albedo: 0.6 0.55 0.5",
           "    }", "}", "Shape: TriangleMesh{", f"    file: {obj_file}", "    scale: 3 3 3",
           "    position: 0 3 0", "    material: batch_mesh", "}"]
    for b in item.get("boxes", ()):
        out += ["Shape: TriangleMesh{", f"    file: {box_file}", f"    scale: {_v3(b['scale'])}",
                f"    position: {_v3(b['position'])}", f"    material: m{b['material']}", "}"]
    return text + "\n".join(out) + "\n"


def write_obj(path, vertices, indices):
    """Writes a mesh as OBJ text with plain decimals (pine's stof has no exponent)."""
    with open(path, "w") as f:
        for v in vertices:
            f.write("v " + " ".join(_fmt(x) for x in v) + "\n")
        for t in indices:
            f.write("f %d %d %d\n" % (t[0] + 1, t[1] + 1, t[2] + 1))


def mixed_scene_text(obj_file, width=96, height=72, spp=64, max_depth=6):
    """A small scene exercising what the spheres scene does not: point and
    directional lights next to an area light (power-based light selection),
    two-layer materials, a rough dielectric, an OBJ triangle mesh, a plane and
    a scaled/positioned mesh, a Box filter and a thin-lens camera."""
    return f"""Integrator: Path{{
    maxDepth: {max_depth}
    sampler: Halton{{
        samplesPerPixel: {spp}
    }}
}}
Camera: ThinLen{{
    film{{
        outputFileName: mixed.bmp
        size: {width} {height}
        filter: Box{{
            radius: 0.5
        }}
    }}
    from: 0.3 2.2 6.5
    to: 0 0.6 0
    fov: 0.8
    lensRadius: 0.02
    focusDistance: 6
}}
Material floor: Layered{{
    layer0: Diffuse{{
        albedo: 0.6 0.55 0.5
    }}
}}
Material coated: Layered{{
    layer0: Diffuse{{
        albedo: 0.2 0.4 0.8
    }}
    layer1: Conductor{{
        albedo: 0.9
        roughness: 0.3
    }}
}}
Material frosted: Layered{{
    layer0: Dielectric{{
        eta: 1.33
        roughness: 0.25
    }}
}}
Material gold: Layered{{
    layer0: Conductor{{
        albedo: 0.95 0.7 0.3
        roughness: 0.15
    }}
}}
Material lamp: Emissive{{
    color: 4 3.5 3
}}
Light: Point{{
    position: -2 3 2
    color: 6 6 8
}}
Light: Directional{{
    direction: 0.3 1 0.5
    color: 0.8 0.8 0.7
}}
Shape: Plane{{
    position: 0 0 0
    normal: 0 1 0
    material: floor
}}
Shape: Sphere{{
    position: -1.2 0.7 0
    radius: 0.7
    material: coated
}}
Shape: Sphere{{
    position: 1.2 0.6 0.4
    radius: 0.6
    material: frosted
}}
Shape: TriangleMesh{{
    file: {obj_file}
    scale: 0.5 0.5 0.5
    position: 0 0.5 -1.4
    material: gold
}}
Shape: Rect{{
    position: 0 3 0
    ex: 1.2 0 0
    ey: 0 0 1.2
    material: lamp
}}
"""


# --- heterogeneous medium ------------------------------------------------------

# The volume benchmark: a 256^3 fBm density grid, max(0, fbm6 - 0.3) * 40, with
# 16^3-voxel majorant bricks, HG g = 0.85, single-scattering albedo 0.99.
CLOUD_GRID = 256
CLOUD_BRICK = 16


def cloud_scene_text(volume_file, width=1024, height=1024, spp=256, max_depth=64, bricks=0,
                     sigma_a=0.01, sigma_s=0.99, g=0.85, scale=2.0, passthrough=False,
                     output="cloud.bmp", lighting="sky"):
    """A density grid seen from outside, the camera inside the medium as in
    pine's scenes/cloud.txt (the grid bounds the density), a diffuse ground and
    pine's Sky environment light (its sun is sampled as a delta light).
    `bricks` > 0 adds the brick-majorant extension parameter.  `passthrough`
    adds a material-less rect in front of the grid (surfaces that rays and
    shadow rays cross, path.cpp:42-46 and integrator.cpp:80-84).
    lighting="sun+constant" replaces the Sky by a Directional sun (colour 3,
    from the Sky's sun direction) and a Constant environment of radiance 0.3
    (this library's extension: pine has no constant environment light)."""
    out = [
        "Integrator: Path{",
        f"    maxDepth: {max_depth}",
        "    sampler: Halton{",
        f"        samplesPerPixel: {spp}",
        "    }",
        "}",
        "Camera: ThinLen{",
        "    film{",
        f"        outputFileName: {output}",
        f"        size: {width} {height}",
        "        filter: LanczosSinc{",
        "            radius: 0.5",
        "        }",
        "    }",
        "    from: 0.3 0.4 3.6",
        "    to: 0 0 0",
        "    fov: 0.7",
        "    medium: cloud",
        "}",
        "Material ground: Layered{",
        "    layer0: Diffuse{",
        "        albedo: 0.35 0.3 0.25",
        "    }",
        "}",
    ] + ([
        "Light: Environment{",
        "    environment: Sky{",
        "        sunDirection: -1 6 1",
        "        sunColor: 1 1 1",
        "    }",
        "}",
    ] if lighting == "sky" else [
        "Light: Directional{",
        "    direction: -1 6 1",
        "    color: 3 3 3",
        "}",
        "Light: Environment{",
        "    environment: Constant{",
        "        color: 0.3 0.3 0.3",
        "    }",
        "}",
    ]) + [
        "Shape: Rect{",
        "    position: 0 -1.2 0",
        "    ex: 0 0 100",
        "    ey: 100 0 0",
        "    material: ground",
        "}",
        "Medium cloud: Grid{",
        f"    file: {volume_file}",
        f"    sigma_a: {_fmt(sigma_a)} {_fmt(sigma_a)} {_fmt(sigma_a)}",
        f"    sigma_s: {_fmt(sigma_s)} {_fmt(sigma_s)} {_fmt(sigma_s)}",
        f"    g: {_fmt(g)}",
        "    position: 0 0 0",
        f"    scale: {_fmt(scale)}",
    ]
    if bricks:
        out.append(f"    majorantBricks: {bricks}")
    out.append("}")
    if passthrough:
        out += ["Shape: Rect{", "    position: 0.2 0.1 1.6", "    ex: 0.8 0 0", "    ey: 0 0.6 0.1",
                "}"]
    return "\n".join(out) + "\n"


def slab_scene_text(volume_file, width=48, height=48, spp=256, sigma=1.0, albedo=0.0,
                    view="through", bricks=0):
    """Closed-form transmittance scenes on a grid of constant density, scale 1
    (the grid is the cube [-0.5, 0.5]^3, model distance = world distance):
      view="through": the camera looks along -z through the cube at an emissive
        wall (Le = 1); with albedo 0 a pixel's expectation is exp(-sigma * d * len)
        (delta tracking, GridMedium::Sample);
      view="shadow": a directional light straight down casts the cube's
        shadow on a white ground seen from the side; direct light only
        (maxDepth 2), so shadowed / lit radiance = exp(-sigma * d * 1)
        (ratio tracking, GridMedium::Tr)."""
    head = ["Integrator: Path{", "    maxDepth: " + ("2" if view == "shadow" else "8"),
            "    sampler: Halton{", f"        samplesPerPixel: {spp}", "    }", "}"]
    med = ["Medium fog: Grid{", f"    file: {volume_file}",
           f"    sigma_a: {_fmt(sigma * (1 - albedo))} {_fmt(sigma * (1 - albedo))} {_fmt(sigma * (1 - albedo))}",
           f"    sigma_s: {_fmt(sigma * albedo)} {_fmt(sigma * albedo)} {_fmt(sigma * albedo)}",
           "    g: 0", "    position: 0 0 0", "    scale: 1"]
    if bricks:
        med.append(f"    majorantBricks: {bricks}")
    med.append("}")
    film = ["    film{", "        outputFileName: slab.bmp", f"        size: {width} {height}",
            "        filter: Box{", "            radius: 0.5", "        }", "    }"]
    if view == "through":
        cam = ["Camera: ThinLen{", *film, "    from: 0 0 3", "    to: 0 0 0", "    fov: 0.25",
               "    medium: fog", "}"]
        scene = ["Material wall: Emissive{", "    color: 1 1 1", "}",
                 "Shape: Rect{", "    position: 0 0 -2", "    ex: 20 0 0", "    ey: 0 20 0",
                 "    material: wall", "}"]
    else:
        # The camera stays below the cube, so only shadow rays cross it.
        cam = ["Camera: ThinLen{", *film, "    from: 0 -0.7 3", "    to: 0 -1 0",
               "    fov: 0.9", "    medium: fog", "}"]
        scene = ["Material white: Layered{", "    layer0: Diffuse{", "        albedo: 1 1 1", "    }",
                 "}", "Light: Directional{", "    direction: 0 1 0", "    color: 1 1 1", "}",
                 "Shape: Rect{", "    position: 0 -1 0", "    ex: 0 0 40", "    ey: 40 0 0",
                 "    material: white", "}"]
    return "\n".join(head + cam + scene + med) + "\n"


def mesh_scene_text(obj_file, width=512, height=512, spp=64, max_depth=8, output="mesh.bmp"):
    """The path-tracing scene around the 2^20-triangle traversal mesh: the
    displaced icosphere (TriangleMesh from an OBJ file) with a coated diffuse
    material on a diffuse ground, lit by one rectangular area light."""
    return f"""Integrator: Path{{
    maxDepth: {max_depth}
    sampler: Halton{{
        samplesPerPixel: {spp}
    }}
}}
Camera: ThinLen{{
    film{{
        outputFileName: {output}
        size: {width} {height}
        filter: LanczosSinc{{
            radius: 0.5
        }}
    }}
    from: 0 0.8 3.4
    to: 0 0 0
    fov: 0.8
}}
Material ground: Layered{{
    layer0: Diffuse{{
        albedo: 0.5 0.5 0.5
    }}
}}
Material body: Layered{{
    layer0: Diffuse{{
        albedo: 0.7 0.45 0.3
    }}
    layer1: Conductor{{
        albedo: 0.9
        roughness: 0.35
    }}
}}
Material lamp: Emissive{{
    color: 12 11 10
}}
Shape: TriangleMesh{{
    file: {obj_file}
    material: body
}}
Shape: Rect{{
    position: 0 -1.15 0
    ex: 0 0 60
    ey: 60 0 0
    material: ground
}}
Shape: Rect{{
    position: 1.5 3 1
    ex: 1.5 0 0
    ey: 0 0 1.5
    material: lamp
}}
"""


# --- occluded-light box (learned path guiding) ----------------------------------

def occluded_box_clutter(n_objects=100, freq=16, seed=21):
    """Seeded clutter for the guiding scene: n_objects displaced icospheres
    (20 freq^2 triangles each; 100 x 5,120 = 512,000 by default) resting on the
    floor of the main room.  Returns (vertices, indices) of one mesh."""
    rng = np.random.default_rng(seed)
    verts, idx, base = [], [], 0
    for k in range(n_objects):
        v, i = make_mesh(MeshSpec(freq, 0.2, seed + k, 0))
        r = rng.uniform(0.04, 0.13)
        c = np.array([rng.uniform(-0.85, 0.85), r, rng.uniform(-0.25, 0.85)])
        verts.append((v * r + c).astype(np.float32))
        idx.append(i + base)
        base += len(v)
    return np.concatenate(verts), np.concatenate(idx).astype(np.uint32)


def occluded_box_scene_text(obj_file, width=1024, height=1024, spp=64, max_depth=8,
                            output="box.bmp"):
    """A closed five-wall box (open towards the camera) split by a partition
    with a 0.3 x 0.3 gap; the only light hangs behind the partition, so most
    radiance reaches the room indirectly through the gap."""
    def rect(pos, ex, ey, mat=None):
        out = ["Shape: Rect{", f"    position: {_v3(pos)}", f"    ex: {_v3(ex)}",
               f"    ey: {_v3(ey)}"]
        if mat:
            out.append(f"    material: {mat}")
        return out + ["}"]
    lines = [
        "Integrator: Path{", f"    maxDepth: {max_depth}", "    sampler: Halton{",
        f"        samplesPerPixel: {spp}", "    }", "}",
        "Camera: ThinLen{", "    film{", f"        outputFileName: {output}",
        f"        size: {width} {height}", "        filter: LanczosSinc{", "            radius: 0.5",
        "        }", "    }", "    from: 0 1 3.3", "    to: 0 0.9 0", "    fov: 0.75", "}",
        "Material white: Layered{", "    layer0: Diffuse{", "        albedo: 0.7 0.7 0.7", "    }", "}",
        "Material red: Layered{", "    layer0: Diffuse{", "        albedo: 0.7 0.15 0.1", "    }", "}",
        "Material green: Layered{", "    layer0: Diffuse{", "        albedo: 0.15 0.6 0.15", "    }", "}",
        "Material clutter: Layered{", "    layer0: Diffuse{", "        albedo: 0.55 0.5 0.4", "    }", "}",
        "Material lamp: Emissive{", "    color: 30 28 25", "}",
    ]
    # Walls: floor, ceiling, back, left, right (centre positions, full extents).
    lines += rect((0, 0, 0), (0, 0, 2), (2, 0, 0), "white")
    lines += rect((0, 2, 0), (2, 0, 0), (0, 0, 2), "white")
    lines += rect((0, 1, -1), (2, 0, 0), (0, 2, 0), "white")
    lines += rect((-1, 1, 0), (0, 2, 0), (0, 0, 2), "red")
    lines += rect((1, 1, 0), (0, 0, 2), (0, 2, 0), "green")
    # Partition at z = -0.4 with a 0.3 x 0.3 gap centred at (0, 1.0).
    z, gy = -0.4, 1.0
    lines += rect((-0.575, 1, z), (0.85, 0, 0), (0, 2, 0), "white")
    lines += rect((0.575, 1, z), (0.85, 0, 0), (0, 2, 0), "white")
    lines += rect((0, (gy - 0.15) / 2, z), (0.3, 0, 0), (0, gy - 0.15, 0), "white")
    lines += rect((0, (gy + 0.15 + 2) / 2, z), (0.3, 0, 0), (0, 2 - gy - 0.15, 0), "white")
    # The lamp under the ceiling of the compartment behind the partition,
    # facing down: no point of the room sees it, light reaches the room only
    # after bouncing off the compartment, through the gap.
    lines += rect((0, 1.95, -0.72), (0.5, 0, 0), (0, 0, 0.4), "lamp")
    lines += ["Shape: TriangleMesh{", f"    file: {obj_file}", "    material: clutter", "}"]
    return "\n".join(lines) + "\n"


def nodes_scene_text(width=160, height=90, spp=64, max_depth=6, output="nodes.bmp"):
    """A scene whose materials are pine node graphs (node.h): a plane with a
    checkerboard albedo over its texture coordinates, a sphere with a noise bump
    map, a conductor with a Noise3D albedo and a checkerboard roughness, and
    spheres whose inputs combine the arithmetic nodes (Composite, Decompose,
    Sin, Cos, Tan, Pow, Sqr, Sqrt, Length, Fract, Invert, Add, Substract,
    Divide, MultiplyAdd, Position, Normal, TexCoord).  One emissive rect."""
    lines = [
        "Integrator: Path{", f"    maxDepth: {max_depth}", "    sampler: Halton{",
        f"        samplesPerPixel: {spp}", "    }", "}",
        "Camera: ThinLen{", "    film{", f"        outputFileName: {output}",
        f"        size: {width} {height}", "        filter: LanczosSinc{", "            radius: 0.5",
        "        }", "    }", "    from: 0 3.5 13", "    to: 0 0.4 0", "    fov: 0.42", "}",
        "Shape: Rect{", "    position: 6 12 5", "    ex: 3 0 0", "    ey: 0 0 3", "    material: lamp", "}",
        "Shape: Plane{", "    position: 0 0 0", "    normal: 0 1 0", "    material: floor", "}",
    ]
    for k, (x, mat) in enumerate([(-3.0, "bumpy"), (-1.5, "rusty"), (0.0, "arith"), (1.5, "glassy"),
                                  (3.0, "trig")]):
        lines += ["Shape: Sphere{", f"    position: {x} 0.6 {0.4 * (k % 2)}", "    radius: 0.6",
                  f"    material: {mat}", "}"]
    lines += [
        "Material lamp: Emissive{", "    color: 60", "}",
        "Material floor: Layered{", "    layer0: Diffuse{", "        albedo: MultiplyAdd{",
        "            input: Checkerboard{", "                input: TexCoord", "                frequency: 3.0",
        "            }", "            mulFactor: 0.8", "            addFactor: 0.1", "        }", "    }", "}",
        "Material bumpy: Layered{", "    layer0: Diffuse{", "        albedo: 0.8 0.75 0.7", "    }",
        "    bumpMap: Multiply{", "        input: Noise{", "            input: Position",
        "            frequency: 12", "            octaves: 3.0", "        }", "        factor: 0.04", "    }", "}",
        "Material rusty: Layered{", "    layer0: Conductor{", "        albedo: Noise3D{",
        "            input: Position", "            frequency: 6", "            octaves: 3", "        }",
        "        roughness: MultiplyAdd{", "            input: Checkerboard{", "                input: TexCoord",
        "                frequency: 2.0", "            }", "            mulFactor: 0.3",
        "            addFactor: 0.15", "        }", "    }", "}",
        "Material arith: Layered{", "    layer0: Diffuse{", "        albedo: Composite{",
        "            inputX: Fract{", "                input: Decompose{", "                    input: Position",
        "                    dimension: 0", "                }", "            }",
        "            inputY: Sqr{", "                input: Decompose{", "                    input: Normal",
        "                    dimension: 1", "                }", "            }",
        "            inputZ: Invert{", "                input: Divide{", "                    input: Length{",
        "                        input: Position", "                    }", "                    factor: 4.0",
        "                }", "            }", "        }", "    }", "}",
        "Material glassy: Layered{", "    layer0: Diffuse{", "        albedo: 0.2 0.5 0.3", "    }",
        "    layer1: Dielectric{", "        roughness: Sqrt{", "            input: Substract{",
        "                input: 0.2", "                factor: Multiply{",
        "                    input: 0.15", "                    factor: Checkerboard{",
        "                        input: Position", "                        frequency: 4",
        "                    }", "                }", "            }", "        }", "        eta: Add{",
        "            input: 1.3", "            factor: 0.2", "        }", "    }", "}",
        "Material trig: Layered{", "    layer0: Diffuse{", "        albedo: Pow{", "            input: Composite{",
        "                inputX: Sin{", "                    input: Decompose{", "                        input: TexCoord",
        "                        dimension: 0", "                    }", "                }",
        "                inputY: Cos{", "                    input: Decompose{",
        "                        input: TexCoord", "                        dimension: 1",
        "                    }", "                }", "                inputZ: Tan{", "                    input: 0.5",
        "                }", "            }", "            exp: 2.0", "        }", "    }", "}",
    ]
    return "\n".join(lines) + "\n"
