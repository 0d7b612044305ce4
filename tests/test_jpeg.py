"""JPEG decoding for glTF textures and sky images (csrc/jpeg_decode.cpp). The
reference decodes them with the `image` crate (world.rs:887-889, 946-969,
1872-1873). CPU only.

Pinning:
- An independent decoder (Pillow over libjpeg-turbo) decodes every file. Ours
  must agree within 3 levels, mean below 0.1, over baseline and progressive
  files; 4:4:4, 4:2:2 and 4:2:0 sampling; odd sizes down to 1x1; restart
  intervals; optimised Huffman tables. The residual is libjpeg-turbo's integer
  IDCT and fixed-point colour conversion against our exact ones. Grey images
  agree exactly.
- A baseline and a progressive file of the same image and quality carry the
  same quantised coefficients. They must decode to identical pixels, which
  pins spectral selection and successive approximation exactly.
- Byte equality with the `image` crate's own decoder is unpinned: the crate
  is not in the reference tree.
"""
import io

import numpy as np
import pytest

import ptcore
from glb_util import GlbBuilder
from test_assets import QUAD_I, QUAD_P, QUAD_UV, quad_glb

Image = pytest.importorskip("PIL.Image")


def smooth(h, w):
    yy, xx = np.mgrid[0:h, 0:w]
    return np.stack([xx * 255 / max(w, 1), yy * 255 / max(h, 1), 128 + 100 * np.sin(xx / 5.0) * np.cos(yy / 7.0)],
                    -1).clip(0, 255).astype(np.uint8)


def jpeg(img, **kw):
    buf = io.BytesIO()
    Image.fromarray(img).save(buf, "JPEG", **kw)
    return buf.getvalue()


def ours(data):
    """Decode through the sky loader: into_rgba32f of 8-bit samples is sample / 255."""
    a = ptcore.Assets(quad_glb(), data)
    return np.round(a.arrays()["skybox"][..., :3].astype(np.float64) * 255).astype(np.int32)


def pillow(data):
    return np.asarray(Image.open(io.BytesIO(data)).convert("RGB")).astype(np.int32)


@pytest.mark.parametrize("subsampling", [0, 1, 2])
@pytest.mark.parametrize("progressive", [False, True])
def test_decode_matches_libjpeg(subsampling, progressive):
    rng = np.random.default_rng(subsampling + 3 * progressive)
    for (h, w) in [(37, 23), (64, 64), (101, 77), (17, 300), (1, 1), (9, 2)]:
        img = smooth(h, w) if h * w > 4 else rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        for q in (50, 95):
            data = jpeg(img, quality=q, subsampling=subsampling, progressive=progressive)
            d = np.abs(ours(data) - pillow(data))
            assert d.max() <= 3 and (h * w < 256 or d.mean() < 0.1), (h, w, q, d.max(), d.mean())


def test_restart_intervals_optimised_tables_and_noise():
    noise = np.random.default_rng(5).integers(0, 256, (48, 40, 3), dtype=np.uint8)
    for kw in [dict(quality=90, subsampling=2, restart_marker_blocks=3),
               dict(quality=90, subsampling=0, restart_marker_rows=1),
               dict(quality=75, subsampling=2, optimize=True),
               dict(quality=90, subsampling=2, progressive=True, restart_marker_blocks=5)]:
        data = jpeg(noise, **kw)
        assert b"\xff\xdd" in data or "restart" not in str(kw)  # a DRI segment was written
        d = np.abs(ours(data) - pillow(data))
        assert d.max() <= 3 and d.mean() < 0.1, (kw, d.max(), d.mean())


def test_progressive_equals_baseline_exactly():
    for sub in (0, 2):
        img = smooth(70, 90)
        base = ours(jpeg(img, quality=80, subsampling=sub))
        prog = ours(jpeg(img, quality=80, subsampling=sub, progressive=True))
        assert np.array_equal(base, prog)


def test_grey_is_exact_and_expands_to_rgb():
    g = smooth(50, 60)[..., 0]
    for prog in (False, True):
        buf = io.BytesIO()
        Image.fromarray(g, "L").save(buf, "JPEG", quality=85, progressive=prog)
        data = buf.getvalue()
        got = ours(data)
        want = np.asarray(Image.open(io.BytesIO(data))).astype(np.int32)
        assert np.array_equal(got[..., 0], want) and np.array_equal(got[..., 1], want) and np.array_equal(got[..., 2], want)


def test_our_encoder_round_trips_through_our_decoder():
    img = smooth(40, 56)
    data = ptcore.encode_image(img, "jpeg", 90)
    d = np.abs(ours(data) - pillow(data))
    assert d.max() <= 3 and d.mean() < 0.1


def test_jpeg_texture_in_a_glb():
    img = smooth(24, 40)
    data = jpeg(img, quality=90, subsampling=2)
    g = GlbBuilder()
    tex = g.image(data)
    g.doc["images"][-1]["mimeType"] = "image/jpeg"
    attrs = {"POSITION": g.accessor(QUAD_P), "TEXCOORD_0": g.accessor(QUAD_UV)}
    g.doc["materials"] = [{"pbrMetallicRoughness": {"baseColorTexture": {"index": tex}}}]
    g.doc["meshes"] = [{"primitives": [{"attributes": attrs, "indices": g.accessor(QUAD_I), "material": 0}]}]
    g.doc["nodes"] = [{"mesh": 0}]
    a = ptcore.Assets(g.build(), jpeg(smooth(4, 8)))
    arr = a.arrays()
    assert a.info()["srgb_atlas_count"] == 1
    m = arr["materials"][0]
    S = 8192
    x, y = int(round(m["base_color_uv"][0] * S)), int(round(m["base_color_uv"][1] * S))
    texels = arr["srgb_atlases"].reshape(S, S, 4)[y:y + 24, x:x + 40]
    assert np.all(texels[..., 3] == 255)
    d = np.abs(texels[..., :3].astype(np.int32) - pillow(data))
    assert d.max() <= 3 and d.mean() < 0.1


def test_unsupported_and_corrupt_files_are_rejected():
    cmyk = io.BytesIO()
    Image.new("CMYK", (16, 16), (10, 20, 30, 40)).save(cmyk, "JPEG")
    with pytest.raises(ptcore.PtError, match="4-component"):
        ptcore.Assets(quad_glb(), cmyk.getvalue())
    good = jpeg(smooth(32, 32), quality=90)
    with pytest.raises(ptcore.PtError, match="JPEG"):
        ptcore.Assets(quad_glb(), good[:20])  # header cut short
    g = GlbBuilder()
    g.image(ptcore.encode_image(smooth(8, 8), "webp"))
    g.doc["meshes"], g.doc["nodes"] = [], []
    with pytest.raises(ptcore.PtError, match="WebP"):
        ptcore.Assets(g.build(), good)
    sof = good.index(b"\xff\xc0")
    bad = bytearray(good)
    bad[sof + 4] = 12  # 12-bit samples
    with pytest.raises(ptcore.PtError, match="8-bit"):
        ptcore.Assets(quad_glb(), bytes(bad))
