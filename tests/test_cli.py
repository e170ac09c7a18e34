"""The `pine scene.txt` command line (pine_amd/pine, pine_cli.cpp): pine's
src/cli/main.cpp over the C ABI, with pine's Film::Finalize output.

CPU: `pine --finalize` of a film must write the BMP pine's own Film::Finalize
writes for the same accumulators (tests/golden/bmp.npz, made by pine's Film
and SaveImage through the oracle harness), byte for byte, with the filmic tone
mapping on and off.  GPU: `pine scene.txt` renders the golden scene; its PFM
and film match pine's film and its BMP is the finalisation of its own film."""
import os
import subprocess

import numpy as np
import pytest

from conftest import golden

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "pine_amd", "pine")


def cli():
    if not os.path.exists(CLI):
        pytest.skip("pine_amd/pine not built")
    return CLI


def write_film(path, film):
    h, w, _ = film.shape
    with open(path, "wb") as f:
        f.write(np.array([w, h], np.int32).tobytes())
        f.write(np.ascontiguousarray(film, np.float32).tobytes())


def read_film(path):
    a = np.fromfile(path, np.float32)
    w, h = a[:2].view(np.int32)
    return a[2:].reshape(h, w, 4)


def finalize(tmp_path, film, tone_map, name="out.bmp"):
    fp, bp = str(tmp_path / "film.bin"), str(tmp_path / name)
    write_film(fp, film)
    r = subprocess.run([cli(), "--finalize", fp, bp] + ([] if tone_map else ["--no-tone-mapping"]),
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    return np.fromfile(bp, np.uint8)


@pytest.mark.parametrize("tone_map", [1, 0])
@pytest.mark.parametrize("name", ["spheres", "mixed", "nodes", "synthetic"])
def test_finalize_writes_pines_bmp(tmp_path, name, tone_map):
    fx = golden("bmp")
    film = fx["synthetic_film"] if name == "synthetic" else golden("render_" + name)["film"]
    ours = finalize(tmp_path, film, tone_map)
    ref = fx[f"{name}_tm{tone_map}"]
    assert ours.size == ref.size
    d = np.flatnonzero(ours != ref)
    assert len(d) == 0, f"{len(d)} bytes differ from pine's BMP, first at {d[:8]}"


def test_bmp_layout(tmp_path):
    """A 4x2 film with one coloured pixel: BGR order, bottom row first,
    pine's header fields."""
    film = np.zeros((2, 4, 4), np.float32)
    film[..., 3] = 1.0
    film[0, 1, :3] = (1.0, 0.25, 0.0)   # top row, x = 1
    b = finalize(tmp_path, film, tone_map=0)
    assert b[:2].tobytes() == b"BM"
    assert int.from_bytes(b[2:6].tobytes(), "little") == 54 + 3 * 4 * 2
    assert int.from_bytes(b[18:22].tobytes(), "little") == 4
    assert int.from_bytes(b[22:26].tobytes(), "little") == 2
    px = b[54:].reshape(2, 4, 3)           # bottom row first
    assert (px[0] == 0).all()
    r, g = 255, int(np.float32(0.25) ** np.float32(1 / 2.2) * 256)
    assert abs(int(px[1, 1, 2]) - r) == 0 and abs(int(px[1, 1, 1]) - g) <= 1 and px[1, 1, 0] == 0


def test_usage_without_arguments():
    r = subprocess.run([cli()], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "Usage: pine" in r.stdout


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"PF"
        w, h = (int(x) for x in f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(), np.float32 if scale < 0 else ">f4")
    return data.reshape(h, w, 3)[::-1]


@pytest.mark.gpu
@pytest.mark.parametrize("tone_map", [True, False])
def test_cli_renders_the_golden_scene(pg, cuda, tmp_path, tone_map):
    """`pine scene.txt` on the golden spheres scene: the film it writes matches
    pine's film as the render test's bar, the PFM is that film's rgb / weight,
    and the BMP is `pine --finalize` of the same film, so it differs from
    pine's own BMP only where the films differ."""
    from test_render_gpu import check_film, linear
    fx = golden("render_spheres")
    text = fx["scene"].item().decode()
    assert "outputFileName: spheres.bmp" in text
    if not tone_map:
        text = text.replace("        outputFileName: spheres.bmp",
                            "        outputFileName: spheres.bmp\n        applyToneMapping: false")
    (tmp_path / "scene.txt").write_text(text)
    r = subprocess.run([cli(), "scene.txt", "--film", "film.bin"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stderr
    film = read_film(str(tmp_path / "film.bin"))
    check_film(film, fx["film"], 2.5e-4, 0.35)
    pfm = read_pfm(str(tmp_path / "spheres.pfm"))
    assert np.array_equal(pfm, linear(film).astype(np.float32))
    bmp = np.fromfile(str(tmp_path / "spheres.bmp"), np.uint8)
    own = finalize(tmp_path, film, tone_map, "own.bmp")
    assert np.array_equal(bmp, own)
    pine_bmp = golden("bmp")[f"spheres_tm{int(tone_map)}"]
    diff = np.abs(bmp.astype(int) - pine_bmp)
    print(f"BMP vs pine's: {(diff > 0).mean():.4f} of bytes differ, max {diff.max()}")
    assert (diff > 0).mean() < 0.02


@pytest.mark.gpu
def test_cli_writes_no_bmp_for_other_extensions(pg, cuda, tmp_path):
    """pine's SaveImage writes BMP files only; for outputFileName result.png it
    warns and writes nothing (the PFM is this command's own addition)."""
    fx = golden("render_spheres")
    text = fx["scene"].item().decode().replace("outputFileName: spheres.bmp",
                                               "outputFileName: result.png")
    (tmp_path / "scene.txt").write_text(text)
    r = subprocess.run([cli(), "scene.txt", "--spp", "4"], cwd=str(tmp_path),
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "unsupported image file extension" in r.stdout
    assert not (tmp_path / "result.png").exists() and not (tmp_path / "result.bmp").exists()
    assert (tmp_path / "result.pfm").exists()
