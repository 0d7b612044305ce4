"""GPU: scenes loaded by the asset front end (a .glb with PNG textures, an
emitter and a Radiance .hdr sky) render on the HIP path exactly as the CPU
oracle renders the same ingested arrays, and pt-render writes the image."""
import io
import math
import os
import subprocess

import numpy as np
import pytest

import binding as oracle
import ptcore
from glb_util import GlbBuilder, hdr, png
from test_gpu_parity import assert_parity

try:
    from PIL import Image
except ImportError:  # the lossless/lossy comparison needs an independent decoder
    Image = None

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def sphere(nlat, nlon, r, c):
    P, N, UV, I = [], [], [], []
    for i in range(nlat + 1):
        th = math.pi * i / nlat
        for j in range(nlon + 1):
            ph = 2 * math.pi * j / nlon
            n = [math.sin(th) * math.cos(ph), math.cos(th), -math.sin(th) * math.sin(ph)]
            P.append([c[k] + r * n[k] for k in range(3)])
            N.append(n)
            UV.append([j / nlon, i / nlat])
    for i in range(nlat):
        for j in range(nlon):
            a, b = i * (nlon + 1) + j, (i + 1) * (nlon + 1) + j
            I += [a, a + 1, b, a + 1, b + 1, b]
    return (np.array(P, np.float32), np.array(N, np.float32), np.array(UV, np.float32), np.array(I, np.uint32))


def scene_files():
    rng = np.random.default_rng(11)
    g = GlbBuilder()
    checker = ((np.indices((64, 64)).sum(0) // 8) % 2 * 180 + 40).astype(np.uint8)
    t_base = g.image(png(np.stack([checker, checker // 2, 255 - checker], -1)))
    t_mr = g.image(png(np.stack([np.zeros((32, 32), np.uint8), rng.integers(0, 256, (32, 32), dtype=np.uint8),
                                 rng.integers(0, 256, (32, 32), dtype=np.uint8)], -1)))
    g.doc["materials"] = [
        {"pbrMetallicRoughness": {"baseColorTexture": {"index": t_base}, "metallicRoughnessTexture": {"index": t_mr},
                                  "metallicFactor": 0.6, "roughnessFactor": 0.8},
         "extras": {"OpenPBR_specular_roughness": 0.2}},
        {"pbrMetallicRoughness": {"baseColorFactor": [0.7, 0.7, 0.7, 1.0], "metallicFactor": 0.0}},
        {"emissiveFactor": [1.0, 0.9, 0.7], "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 30.0}},
         "pbrMetallicRoughness": {"metallicFactor": 0.0}}]
    P, N, UV, I = sphere(24, 48, 0.6, (0.0, 0.6, 0.0))
    sph = {"attributes": {"POSITION": g.accessor(P), "NORMAL": g.accessor(N), "TEXCOORD_0": g.accessor(UV)},
           "indices": g.accessor(I), "material": 0}
    fl = np.array([[-3, 0, -3], [3, 0, -3], [3, 0, 3], [-3, 0, 3]], np.float32)
    floor = {"attributes": {"POSITION": g.accessor(fl), "NORMAL": g.accessor(np.tile([0, 1, 0], (4, 1)).astype(np.float32)),
                            "TEXCOORD_0": g.accessor(np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32))},
             "indices": g.accessor(np.array([0, 2, 1, 0, 3, 2], np.uint32)), "material": 1}
    lp = np.array([[-0.4, 2.5, -0.4], [0.4, 2.5, -0.4], [0.4, 2.5, 0.4], [-0.4, 2.5, 0.4]], np.float32)
    light = {"attributes": {"POSITION": g.accessor(lp)}, "indices": g.accessor(np.array([0, 1, 2, 0, 2, 3], np.uint32)),
             "material": 2}
    g.doc["meshes"] = [{"primitives": [sph]}, {"primitives": [floor, light]}]
    g.doc["nodes"] = [{"mesh": 0, "rotation": [0.0, math.sin(0.3), 0.0, math.cos(0.3)], "translation": [0.2, 0.0, 0.0]},
                      {"mesh": 1}]
    g.doc["scenes"] = [{"nodes": [0, 1]}]
    g.doc["scene"] = 0
    sky = np.zeros((32, 64, 4), np.uint8)
    sky[..., :3] = rng.integers(100, 256, (32, 64, 3))
    sky[..., 3] = 128 + rng.integers(0, 3, (32, 64))
    return g.build(), hdr(sky)


@pytest.fixture(scope="module")
def ingested(ctx):
    glb, sky = scene_files()
    a = ptcore.Assets(glb, sky)
    w = a.world(skybox_luminance=2.0, camera_position=(0.0, 1.2, 3.5), pitch=-12.0, fov=50.0, aspect=1.0)
    s = ptcore.Scene(ctx, w.desc)
    yield w, s
    s.close()


def test_ingested_primary_hits_bitexact(ctx, ingested):
    w, s = ingested
    t = ptcore.Tracer(ctx, s, 160, 160)
    tri, tt = t.primary_hits(3)
    ref_tri, ref_t = oracle.primary_hits(w.desc, 160, 160, 3)
    assert (tri >= 0).mean() > 0.5
    assert np.array_equal(tri, ref_tri) and np.array_equal(tt[tri >= 0], ref_t[ref_tri >= 0])
    t.close()


@pytest.mark.parametrize("schedule", [ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT])
def test_ingested_radiance_matches_oracle(ctx, ingested, schedule):
    w, s = ingested
    t = ptcore.Tracer(ctx, s, 128, 128, schedule)
    t.set_max_bounces(6)
    t.trace(8)
    ref, _, _ = oracle.render_f32(w.desc, 128, 128, 6, 8)
    assert_parity(t.accumulation(), ref, f"ingested glb schedule={schedule}")
    t.close()


def test_pt_render_writes_the_image(tmp_path):
    glb, sky = scene_files()
    (tmp_path / "scene.glb").write_bytes(glb)
    (tmp_path / "sky.hdr").write_bytes(sky)
    decoded = {}
    for ext in ["png", "webp", "jpg"]:
        out = tmp_path / f"render.{ext}"
        r = subprocess.run([os.path.join(ROOT, "path-tracer_amd", "pt-render"), "render", "--scene",
                            str(tmp_path / "scene.glb"), "--skybox", str(tmp_path / "sky.hdr"), "--width", "96",
                            "--height", "64", "--samples", "4", "--tile-size", "48", "--output", str(out)],
                           capture_output=True, text=True, timeout=120)
        print(ext, r.stdout, r.stderr)
        assert r.returncode == 0
        data = out.read_bytes()
        assert data[:4] in (b"\x89PNG", b"RIFF", b"\xff\xd8\xff\xe0")
        if Image is not None:
            decoded[ext] = np.asarray(Image.open(io.BytesIO(data)).convert("RGB")).astype(np.int32)
    if decoded:
        # the same deterministic render saved three ways: lossless files agree exactly, JPEG closely
        assert decoded["png"].shape == (64, 96, 3)
        assert np.array_equal(decoded["png"], decoded["webp"])
        assert np.abs(decoded["jpg"] - decoded["png"]).mean() < 4.0


def render_tile(ctx, scene_desc, samples=4):
    s = ptcore.Scene(ctx, scene_desc)
    t = ptcore.Tracer(ctx, s, 64, 64)
    t.set_tile((1920, 1080), (800, 500), (64, 64))
    t.set_max_bounces(6)
    t.trace(samples)
    out = t.accumulation().copy()
    t.close()
    s.close()
    return out


def test_exported_field_scene_and_its_ptscene_render_as_the_original(ctx, field, tmp_path):
    """The 262k-triangle field scene written as .glb + FLOAT EXR sky (loaded back with
    its render flags), and its built arrays as .ptscene, render bit for bit as the world
    they came from."""
    want = render_tile(ctx, field.desc)
    assert np.isfinite(want).all() and want[..., :3].sum() > 0
    ptcore.save_scene(field.desc, tmp_path / "field.ptscene")
    f = ptcore.SceneFile(tmp_path / "field.ptscene")
    got = render_tile(ctx, f.desc)
    print("ptscene: bit-exact", np.array_equal(got, want), flush=True)
    assert np.array_equal(got, want)
    f.close()
    g = ptcore.Generated(ptcore.SCENE_FIELD, 2, 1920, 1080)
    assert ptcore.export_glb(g.desc, tmp_path / "field.glb", tmp_path / "field_sky.exr") == 0
    flags = g.render_flags()
    opts = dict(zip(flags[::2], flags[1::2]))
    a = ptcore.Assets(tmp_path / "field.glb", tmp_path / "field_sky.exr")
    w = a.world(skybox_luminance=float(opts["--skybox-luminance"]),
                camera_position=[float(c) for c in opts["--camera-position"].split(",")],
                yaw=float(opts["--camera-yaw"]), pitch=float(opts["--camera-pitch"]),
                fov=float(opts["--camera-fov"]), aspect=1920 / 1080)
    got = render_tile(ctx, w.desc)
    print("glb: bit-exact", np.array_equal(got, want), flush=True)
    assert np.array_equal(got, want)
