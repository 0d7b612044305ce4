"""The asset front end (pt_assets.h): glTF scenes (Scene::new, world.rs:887-1423),
PNG textures and Radiance .hdr skies. CPU only.

Pinning: expected values are computed independently in numpy from the glTF
specification and the reference's conversion rules (node matrices as
glam-ordered float32 products, normals by the inverse transpose, RGBA8
conversion rules, RGBE = c * 2^(e - 136)), and tangents by closed forms (a
planar quad's tangent is the world direction of +u) and invariants (unit
length, orthogonal to the normal, handedness). Parity with the bevy_mikktspace
crate itself is unpinned: the crate is not in the reference tree and no
reference test fixes its output.
"""
import math
import os

import numpy as np
import pytest

import ptcore
from glb_util import GlbBuilder, exr, hdr, png, rgbe_to_float

REF = "/root/reference/assets"
SKY = hdr(np.array([[[128, 128, 128, 129], [64, 32, 16, 130]]], np.uint8), rle=False)


def f32(x):
    return np.float32(x)


def quat_mat(q):
    x, y, z, w = (f32(v) for v in q)
    x2, y2, z2 = x + x, y + y, z + z
    xx, xy, xz, yy, yz, zz = x * x2, x * y2, x * z2, y * y2, y * z2, z * z2
    wx, wy, wz = w * x2, w * y2, w * z2
    m = np.eye(4, dtype=np.float32)  # m[col][row]
    m[0, :3] = (f32(1) - (yy + zz), xy + wz, xz - wy)
    m[1, :3] = (xy - wz, f32(1) - (xx + zz), yz + wx)
    m[2, :3] = (xz + wy, yz - wx, f32(1) - (xx + yy))
    return m


def mat_mul(a, b):  # column-major, glam order ((x + y) + z) + w
    o = np.zeros((4, 4), np.float32)
    for j in range(4):
        for r in range(4):
            o[j, r] = ((a[0, r] * b[j, 0] + a[1, r] * b[j, 1]) + a[2, r] * b[j, 2]) + a[3, r] * b[j, 3]
    return o


def trs(t=(0, 0, 0), q=(0, 0, 0, 1), s=(1, 1, 1)):
    T = np.eye(4, dtype=np.float32)
    T[3, :3] = t
    S = np.diag(np.array([*s, 1], np.float32))
    return mat_mul(mat_mul(T, quat_mat(q)), S)


def point(m, p):
    return np.array([((m[0, r] * f32(p[0]) + m[1, r] * f32(p[1])) + m[2, r] * f32(p[2])) + m[3, r] for r in range(3)],
                    np.float32)


QUAD_P = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0]], np.float32)
QUAD_UV = np.array([[0, 0], [1, 0], [1, 1], [0, 1]], np.float32)
QUAD_I = np.array([0, 1, 2, 0, 2, 3], np.uint32)


def quad_glb(node=None, normals=True, uvs=True, tangents=None, material=None, extra_doc=None):
    g = GlbBuilder()
    attrs = {"POSITION": g.accessor(QUAD_P)}
    if normals:
        attrs["NORMAL"] = g.accessor(np.tile(np.array([0, 0, 1], np.float32), (4, 1)))
    if uvs:
        attrs["TEXCOORD_0"] = g.accessor(QUAD_UV)
    if tangents is not None:
        attrs["TANGENT"] = g.accessor(np.asarray(tangents, np.float32))
    prim = {"attributes": attrs, "indices": g.accessor(QUAD_I)}
    if material is not None:
        prim["material"] = 0
        g.doc["materials"] = [material]
    g.doc["meshes"] = [{"primitives": [prim]}]
    g.doc["nodes"] = [dict(mesh=0, **(node or {}))]
    g.doc["scenes"] = [{"nodes": [0]}]
    g.doc["scene"] = 0
    if extra_doc:
        g.doc.update(extra_doc)
    return g.build()


def load(glb, sky=SKY):
    a = ptcore.Assets(glb, sky)
    return a, a.arrays()


def test_node_transform_normals_and_mikktspace_quad():
    q = (0.0, math.sin(math.pi / 8), 0.0, math.cos(math.pi / 8))  # 45 degrees about +y
    node = {"translation": [1.5, -2.0, 0.25], "rotation": list(q), "scale": [2.0, 2.0, 2.0]}
    a, arr = load(quad_glb(node))
    assert a.info()["triangle_count"] == 2
    m = trs(node["translation"], q, node["scale"])
    want = np.array([[point(m, QUAD_P[i]) for i in QUAD_I[3 * t:3 * t + 3]] for t in range(2)])
    got = arr["geometries"]["v"].reshape(2, 3, 3)
    assert np.array_equal(got, want)
    # normals: inverse transpose of a rotation with uniform scale = the rotation, normalised
    n_want = np.array([math.sin(math.pi / 4), 0.0, math.cos(math.pi / 4)])
    n = arr["attributes"]["n"].reshape(-1, 3)
    assert np.allclose(n, n_want, atol=1e-6)
    # MikkTSpace on a planar quad: tangent = world direction of +u; uv area > 0 so the
    # frame preserves orientation (+1), which the reference stores negated (w = -1)
    t_want = np.array([math.cos(math.pi / 4), 0.0, -math.sin(math.pi / 4)])
    for k in ("t0", "t1", "t2"):
        t = arr["attributes"][k]
        assert np.allclose(t[:, :3], t_want, atol=1e-6)
        assert np.all(t[:, 3] == -1.0)


def test_given_tangents_are_transformed_and_placeholders_ignored():
    t = np.tile(np.array([0.0, 1.0, 0.0, -1.0], np.float32), (4, 1))
    q = (0.0, 0.0, math.sin(math.pi / 4), math.cos(math.pi / 4))  # 90 degrees about +z
    _, arr = load(quad_glb({"rotation": list(q)}, tangents=t))
    got = arr["attributes"]["t0"]
    assert np.allclose(got[:, :3], [-1.0, 0.0, 0.0], atol=1e-6) and np.all(got[:, 3] == -1.0)
    # the (1,0,0,1) placeholder switches to computed tangents (+u direction, negated handedness)
    ph = np.tile(np.array([1.0, 0.0, 0.0, 1.0], np.float32), (4, 1))
    _, arr = load(quad_glb(tangents=ph))
    assert np.allclose(arr["attributes"]["t1"][:, :3], [1.0, 0.0, 0.0]) and np.all(arr["attributes"]["t1"][:, 3] == -1)
    # w varying inside a triangle also switches
    var = t.copy()
    var[1, 3] = 1.0
    _, arr = load(quad_glb(tangents=var))
    assert np.allclose(arr["attributes"]["t2"][:, :3], [1.0, 0.0, 0.0])


def test_flat_normals_and_dummy_uvs_when_missing():
    _, arr = load(quad_glb(normals=False, uvs=False))
    assert np.array_equal(arr["attributes"]["n"].reshape(-1, 3), np.tile([0.0, 0.0, 1.0], (6, 1)).astype(np.float32))
    dummy = np.array([[0, 0], [1, 0], [0, 1]], np.float32)
    uv = arr["attributes"]["uv"].reshape(2, 3, 2)
    for t in range(2):
        for c in range(3):
            assert np.array_equal(uv[t, c], dummy[QUAD_I[3 * t + c] % 3])


def test_materials_defaults_extensions_and_extras():
    mat = {"pbrMetallicRoughness": {"baseColorFactor": [0.5, 0.25, 0.125, 0.75], "metallicFactor": 0.2,
                                    "roughnessFactor": 0.7},
           "emissiveFactor": [1.0, 0.5, 0.0], "alphaMode": "MASK", "alphaCutoff": 0.3,
           "extensions": {"KHR_materials_emissive_strength": {"emissiveStrength": 40.0},
                          "KHR_materials_specular": {"specularFactor": 0.4, "specularColorFactor": [0.9, 0.8, 0.7]},
                          "KHR_materials_ior": {"ior": 1.33}},
           "extras": {"OpenPBR_specular_roughness": 0.05, "OpenPBR_base_weight": 0.9, "OpenPBR_specular_haze": 2.0,
                      "unknown_key": 1}}
    a, arr = load(quad_glb(material=mat))
    m = arr["materials"][0]
    assert tuple(m["base_color"]) == tuple(np.float32([0.5, 0.25, 0.125, 0.75]))
    assert m["base_metalness"] == f32(0.2) and m["base_diffuse_roughness"] == f32(0.7)
    assert tuple(m["emission_color_and_luminance"]) == (1.0, 0.5, 0.0, 40.0)
    assert tuple(m["specular_color_and_weight"]) == tuple(np.float32([0.9, 0.8, 0.7, 0.4]))
    assert m["specular_ior"] == f32(1.33) and m["alpha_mode"] == 1 and m["alpha_cutoff"] == f32(0.3)
    assert m["specular_roughness"] == f32(0.05) and m["base_weight"] == f32(0.9)
    assert m["specular_haze"] == 0.0  # out of [0, 1]: ignored with a warning
    assert m["base_color_tex_layer"] == -1 and m["emission_weight"] == 1.0 and m["specular_haze_spread"] == f32(0.3)
    w = a.warnings()
    assert any("unknown_key" in x for x in w) and any("OpenPBR_specular_haze" in x for x in w)
    # no material: the glTF default (white, metallic 1, roughness 1) and reference defaults
    _, arr = load(quad_glb())
    m = arr["materials"][0]
    assert tuple(m["base_color"]) == (1, 1, 1, 1) and m["base_metalness"] == 1 and m["base_diffuse_roughness"] == 1
    assert tuple(m["specular_color_and_weight"]) == (1, 1, 1, 1) and m["specular_ior"] == f32(1.5)
    assert m["specular_roughness"] == f32(0.3) and m["alpha_mode"] == 0 and m["alpha_cutoff"] == f32(0.5)


def test_hierarchy_order_material_sharing_and_default_scene():
    g = GlbBuilder()
    pos = g.accessor(QUAD_P)
    idx = g.accessor(QUAD_I)
    g.doc["materials"] = [{"pbrMetallicRoughness": {"baseColorFactor": [1, 0, 0, 1]}},
                          {"pbrMetallicRoughness": {"baseColorFactor": [0, 1, 0, 1]}}]
    g.doc["meshes"] = [{"primitives": [{"attributes": {"POSITION": pos}, "indices": idx, "material": 1},
                                       {"attributes": {"POSITION": pos}, "indices": idx, "material": 0},
                                       {"attributes": {"POSITION": pos}, "indices": idx, "mode": 1}]}]
    g.doc["nodes"] = [{"children": [1, 2], "translation": [10, 0, 0]},
                      {"mesh": 0, "translation": [0, 5, 0]},
                      {"mesh": 0, "scale": [1, 1, 3], "translation": [0, 0, 7]}]
    g.doc["scenes"] = [{"nodes": [0]}]
    g.doc["scene"] = 0
    a, arr = load(g.build())
    # LIFO traversal: node 0, then its last child (2), then child 1; each mesh: primitive order,
    # the non-triangle primitive skipped with a warning
    assert a.info()["triangle_count"] == 8
    v = arr["geometries"]["v"].reshape(8, 3, 3)
    assert np.all(v[:4, :, 2] == 7.0) and np.all(v[:4, :, 0] >= 10.0)  # node 2 first
    assert np.all(v[4:, :, 1] >= 5.0)
    # materials created at first use and shared: node 2 uses material 1 then 0
    assert arr["materials"].shape[0] == 2
    assert list(arr["attributes"]["material"]) == [0, 0, 1, 1, 0, 0, 1, 1]
    assert tuple(arr["materials"][0]["base_color"]) == (0, 1, 0, 1)
    assert any("not triangles" in x for x in a.warnings())
    # without a default scene every node is a root: the children are flattened twice more
    del g.doc["scene"]
    a2, _ = load(g.build())
    assert a2.info()["triangle_count"] == 8 + 8


def test_png_textures_into_srgb_and_linear_atlases():
    rgb = np.arange(4 * 3 * 3, dtype=np.uint8).reshape(4, 3, 3) * 7
    la = np.stack([np.arange(6, dtype=np.uint8).reshape(2, 3) * 40, np.full((2, 3), 200, np.uint8)], -1)
    wide = (np.arange(2 * 2 * 4, dtype=np.uint16).reshape(2, 2, 4) * 4099)
    pal_idx = np.array([[0, 1, 2], [2, 1, 0]], np.uint8)
    palette = [[255, 0, 0], [0, 255, 0], [0, 0, 255]]
    g = GlbBuilder()
    t_rgb = g.image(png(rgb))
    t_la = g.image(png(la))
    t_wide = g.image(png(wide))
    t_pal = g.image(png(pal_idx, palette=palette, trns=[128]))
    attrs = {"POSITION": g.accessor(QUAD_P), "TEXCOORD_0": g.accessor(QUAD_UV)}
    g.doc["materials"] = [{"pbrMetallicRoughness": {"baseColorTexture": {"index": t_rgb},
                                                    "metallicRoughnessTexture": {"index": t_la}},
                           "normalTexture": {"index": t_wide, "scale": 0.5},
                           "emissiveTexture": {"index": t_pal}}]
    g.doc["meshes"] = [{"primitives": [{"attributes": attrs, "indices": g.accessor(QUAD_I), "material": 0}]}]
    g.doc["nodes"] = [{"mesh": 0}]
    a, arr = load(g.build())
    info = a.info()
    assert info["srgb_atlas_count"] == 1 and info["linear_atlas_count"] == 1 and info["atlas_size"] == 8192
    m = arr["materials"][0]
    assert m["normal_scale"] == 0.5
    S = 8192
    srgb = arr["srgb_atlases"].reshape(S, S, 4)
    lin = arr["linear_atlases"].reshape(S, S, 4)

    def block(atlas, uv, h, w):
        x, y = int(round(uv[0] * S)), int(round(uv[1] * S))
        assert uv[2] == np.float32(w) / np.float32(S) and uv[3] == np.float32(h) / np.float32(S)
        return atlas[y:y + h, x:x + w]

    assert m["base_color_tex_layer"] == 0 and m["emission_tex_layer"] == 0
    assert np.array_equal(block(srgb, m["base_color_uv"], 4, 3)[..., :3], rgb)
    assert np.all(block(srgb, m["base_color_uv"], 4, 3)[..., 3] == 255)
    pal = np.array(palette, np.uint8)[pal_idx]
    got = block(srgb, m["emission_uv"], 2, 3)
    assert np.array_equal(got[..., :3], pal) and np.array_equal(got[..., 3], np.where(pal_idx == 0, 128, 255))
    # RG8 (grey + alpha) -> (l, a, 0, 255); 16-bit -> high bytes
    got = block(lin, m["metallic_roughness_uv"], 2, 3)
    assert np.array_equal(got[..., 0], la[..., 0]) and np.array_equal(got[..., 1], la[..., 1])
    assert np.all(got[..., 2] == 0) and np.all(got[..., 3] == 255)
    assert np.array_equal(block(lin, m["normal_uv"], 2, 2), (wide >> 8).astype(np.uint8))


def test_hdr_sky_rle_and_flat():
    rng = np.random.default_rng(3)
    rgbe = rng.integers(0, 256, (5, 40, 4), dtype=np.uint8)
    rgbe[:, 10:30, :] = rgbe[:, 10:11, :]  # long runs
    rgbe[2, :, 3] = 0  # e = 0 is black
    want = rgbe_to_float(rgbe)
    for rle in (True, False):
        a = ptcore.Assets()
        a.load_skybox(hdr(rgbe, rle))
        info = a.info()
        assert (info["skybox_width"], info["skybox_height"]) == (40, 5)
    a = ptcore.Assets(quad_glb(), hdr(rgbe, True))
    sky = a.arrays()["skybox"]
    assert np.array_equal(sky[..., :3], want) and np.all(sky[..., 3] == 1.0)
    # 8-bit PNG skies decode to sample / 255 without gamma
    px = rng.integers(0, 256, (2, 4, 3), dtype=np.uint8)
    a = ptcore.Assets(quad_glb(), png(px))
    sky = a.arrays()["skybox"]
    assert np.array_equal(sky[..., :3], px.astype(np.float32) / np.float32(255.0))


def test_bad_inputs_are_rejected():
    a = ptcore.Assets()
    with pytest.raises(ptcore.PtError):
        a.load_scene(b"glTF\x02\x00\x00\x00\xff\xff\x00\x00")
    with pytest.raises(ptcore.PtError):
        a.load_scene(b'{"asset": {"version": "2.0"}, "nodes": [{"mesh": 3}], "meshes": []}')
    with pytest.raises(ptcore.PtError):
        a.load_skybox(b"\x76\x2f\x31\x01 not really exr")
    with pytest.raises(ptcore.PtError):
        a.world_desc()  # nothing loaded
    bad = GlbBuilder()
    bad.doc["meshes"] = [{"primitives": [{"attributes": {"POSITION": bad.accessor(QUAD_P)},
                                          "indices": bad.accessor(np.array([0, 1, 9], np.uint32))}]}]
    bad.doc["nodes"] = [{"mesh": 0}]
    with pytest.raises(ptcore.PtError) as e:
        a.load_scene(bad.build())
    assert "out of range" in str(e.value)


def test_mikktspace_frames_on_a_uv_sphere_are_orthonormal():
    # latitude-longitude sphere with smooth normals: tangents follow +u (east), unit, orthogonal to n
    nlat, nlon = 12, 24
    P, N, UV = [], [], []
    for i in range(nlat + 1):
        th = math.pi * i / nlat
        for j in range(nlon + 1):
            ph = 2 * math.pi * j / nlon
            n = [math.sin(th) * math.cos(ph), math.cos(th), -math.sin(th) * math.sin(ph)]
            P.append(n)
            N.append(n)
            UV.append([j / nlon, i / nlat])
    I = []
    for i in range(nlat):
        for j in range(nlon):
            a, b = i * (nlon + 1) + j, (i + 1) * (nlon + 1) + j
            I += [a, b, a + 1, a + 1, b, b + 1]
    g = GlbBuilder()
    attrs = {"POSITION": g.accessor(np.array(P, np.float32)), "NORMAL": g.accessor(np.array(N, np.float32)),
             "TEXCOORD_0": g.accessor(np.array(UV, np.float32))}
    g.doc["meshes"] = [{"primitives": [{"attributes": attrs, "indices": g.accessor(np.array(I, np.uint32))}]}]
    g.doc["nodes"] = [{"mesh": 0}]
    _, arr = load(g.build())
    at = arr["attributes"]
    for c, k in enumerate(("t0", "t1", "t2")):
        t = at[k][:, :3].astype(np.float64)
        n = at["n"][:, 3 * c:3 * c + 3].astype(np.float64)
        ok = np.abs(n[:, 1]) < 0.99  # away from the poles
        assert np.allclose(np.linalg.norm(t[ok], axis=1), 1.0, atol=1e-5)
        assert np.abs((t[ok] * n[ok]).sum(1)).max() < 1e-4
        east = np.stack([n[:, 2], np.zeros(len(n)), -n[:, 0]], 1)  # d(position)/du
        east /= np.linalg.norm(east, axis=1, keepdims=True) + 1e-30
        assert (t[ok] * east[ok]).sum(1).min() > 0.95
        # this winding has negative uv area: orientation-reversing frames, stored negated (+1)
        assert np.all(at[k][:, 3] == 1.0)


def awkward_grid(n, seed):
    """An n x n heightfield with a mirrored-uv half, a collapsed-uv band, some
    reversed windings, a NaN normal and degenerate triangles. Its last rows use
    duplicates of their vertices, which must weld with the originals."""
    rng = np.random.default_rng(seed)
    ys, xs = np.mgrid[0:n, 0:n].astype(np.float32)
    P = np.stack([xs / n, ys / n, rng.normal(0, 0.05, (n, n)).astype(np.float32)], -1).reshape(-1, 3)
    UV = np.stack([xs / n, ys / n], -1).reshape(-1, 2)
    UV[xs.reshape(-1) > n / 2, 0] = 1.0 - UV[xs.reshape(-1) > n / 2, 0]
    UV[ys.reshape(-1) < 0.15 * n] = 0.25
    N = np.tile(np.float32([0, 0, 1]), (n * n, 1)) + rng.normal(0, 0.1, (n * n, 3)).astype(np.float32)
    N /= np.linalg.norm(N, axis=1, keepdims=True)
    N[n + 1] = np.nan
    x, y = np.meshgrid(np.arange(n - 1), np.arange(n - 1))
    a = (y * n + x).reshape(-1)
    quads = np.stack([a, a + 1, a + n + 1, a, a + n + 1, a + n], 1).reshape(-1, 3)
    flip = rng.random(len(quads)) < 0.1
    quads[flip] = quads[flip][:, ::-1]
    dup_from = n * (n - 8)
    P = np.concatenate([P, P[dup_from:]])
    N = np.concatenate([N, N[dup_from:]])
    UV = np.concatenate([UV, UV[dup_from:]])
    tail = quads.min(1) >= dup_from
    quads[tail] += n * n - dup_from
    I = np.concatenate([quads.reshape(-1), [0, 0, 1, 5, 6, 5]]).astype(np.uint32)
    return P.astype(np.float32), N.astype(np.float32), UV.astype(np.float32), I


def test_mikktspace_weld_and_thread_count():
    """The weld and the derivative pass run on several threads above 65536
    vertices: the frames must be the same on one thread, and corners of
    duplicated vertices must carry the frames of the originals they weld with."""
    n = 300
    P, N, UV, I = awkward_grid(n, 3)
    g = GlbBuilder()
    attrs = {"POSITION": g.accessor(P), "NORMAL": g.accessor(N), "TEXCOORD_0": g.accessor(UV)}
    for k in ("min", "max"):  # optional for normals, and JSON has no NaN
        g.doc["accessors"][attrs["NORMAL"]].pop(k)
    g.doc["meshes"] = [{"primitives": [{"attributes": attrs, "indices": g.accessor(I)}]}]
    g.doc["nodes"] = [{"mesh": 0}]
    glb = g.build()
    out = {}
    saved = os.environ.get("PT_HOST_THREADS")
    try:
        for t in ("1", "8"):
            os.environ["PT_HOST_THREADS"] = t
            _, arr = load(glb)
            out[t] = arr["attributes"].copy()
    finally:
        if saved is None:
            os.environ.pop("PT_HOST_THREADS", None)
        else:
            os.environ["PT_HOST_THREADS"] = saved
    assert out["1"].tobytes() == out["8"].tobytes()
    # the same mesh without duplicates: every corner's frame is unchanged
    dup_from = n * (n - 8)
    I2 = I.copy()
    I2[I2 >= n * n] -= n * n - dup_from
    g2 = GlbBuilder()
    attrs2 = {"POSITION": g2.accessor(P[:n * n]), "NORMAL": g2.accessor(N[:n * n]),
              "TEXCOORD_0": g2.accessor(UV[:n * n])}
    for k in ("min", "max"):
        g2.doc["accessors"][attrs2["NORMAL"]].pop(k)
    g2.doc["meshes"] = [{"primitives": [{"attributes": attrs2, "indices": g2.accessor(I2)}]}]
    g2.doc["nodes"] = [{"mesh": 0}]
    _, arr2 = load(g2.build())
    for k in ("t0", "t1", "t2"):
        assert np.array_equal(arr2["attributes"][k], out["8"][k], equal_nan=True)


@pytest.mark.skipif(not os.path.exists(REF), reason="the reference's asset files are not present")
def test_reference_default_assets_load():
    a = ptcore.Assets(os.path.join(REF, "glb", "box.glb"), os.path.join(REF, "skybox", "kloppenheim_06_puresky_2k.hdr"))
    info = a.info()
    assert info["material_count"] == 8 and info["srgb_atlas_count"] == 0
    assert (info["skybox_width"], info["skybox_height"]) == (2048, 1024)
    arr = a.arrays()
    assert np.isfinite(arr["skybox"]).all() and (arr["skybox"] >= 0).all()
    t = arr["attributes"]["t0"]
    assert np.allclose(np.linalg.norm(t[:, :3], axis=1), 1.0, atol=1e-4)
    w = a.world(skybox_luminance=8000.0, aspect=16 / 9)
    assert w.desc.triangle_count == info["triangle_count"]
    assert w.desc.skybox.average_luminance == 8000.0


@pytest.mark.parametrize("compression", [0, 2, 3])
def test_exr_sky_channels_and_compression(compression):
    rng = np.random.default_rng(compression + 1)
    h, w = 37, 23
    r = rng.uniform(0, 40, (h, w)).astype(np.float32)
    g = rng.uniform(0, 1, (h, w)).astype(np.float16)
    b = np.round(rng.uniform(0, 3, (h, w))).astype(np.float32)  # runs compress under ZIP
    a = ptcore.Assets(quad_glb(), exr({"R": r, "G": g, "B": b}, compression))
    sky = a.arrays()["skybox"]
    assert sky.shape == (h, w, 4)
    assert np.array_equal(sky[..., 0], r) and np.array_equal(sky[..., 1], g.astype(np.float32))
    assert np.array_equal(sky[..., 2], b) and np.all(sky[..., 3] == 1.0)  # no A channel: opaque
    a2 = ptcore.Assets(quad_glb(), exr({"R": r, "G": r, "B": r, "A": np.zeros((h, w), np.float32)}, compression))
    assert np.all(a2.arrays()["skybox"][..., 3] == 0.0)


@pytest.mark.skipif(not os.path.exists(REF), reason="the reference's asset files are not present")
def test_reference_4k_exr_sky_matches_its_2k_hdr():
    # the reference's 4k sky is PIZ-compressed FLOAT RGBA; its 2k .hdr is the same sky, RGBE-quantised:
    # the 4k image averaged over 2x2 pixels must agree with it to the RGBE quantisation level
    a = ptcore.Assets(quad_glb(), os.path.join(REF, "skybox", "kloppenheim_06_puresky_4k.exr"))
    ex = a.arrays()["skybox"].copy()
    b = ptcore.Assets(quad_glb(), os.path.join(REF, "skybox", "kloppenheim_06_puresky_2k.hdr"))
    hd = b.arrays()["skybox"]
    assert ex.shape == (2048, 4096, 4) and np.isfinite(ex).all() and np.all(ex[..., 3] == 1.0)
    lum = lambda x: (x[..., :3].astype(np.float64) * [0.2126, 0.7152, 0.0722]).sum(-1)
    le, lh = lum(ex.reshape(1024, 2, 2048, 2, 4).mean((1, 3))), lum(hd)
    rel = np.abs(le - lh) / np.maximum(lh, 1e-6)
    assert np.median(rel) < 0.01 and np.percentile(rel, 99) < 0.03
    assert np.corrcoef(le.ravel(), lh.ravel())[0, 1] > 0.9999


def test_huge_declared_sizes_are_rejected_before_allocation():
    # a small file whose header declares a giant image must fail at once, not allocate or decode it
    import struct
    import time
    import zlib

    small = png(np.zeros((4, 4, 3), np.uint8))
    big_png = bytearray(small)
    ihdr = small.index(b"IHDR")
    big_png[ihdr + 4:ihdr + 12] = struct.pack(">II", 20000, 20000)
    big_png[ihdr + 17:ihdr + 21] = struct.pack(">I", zlib.crc32(bytes(big_png[ihdr:ihdr + 17])) & 0xFFFFFFFF)
    big_hdr = b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 20000 +X 20000\n" + bytes(64)
    small_exr = exr({"R": np.ones((2, 2), np.float32), "G": np.ones((2, 2), np.float32), "B": np.ones((2, 2), np.float32)})
    dw = small_exr.index(b"dataWindow\0box2i\0") + len(b"dataWindow\0box2i\0") + 4
    big_exr = small_exr[:dw] + struct.pack("<iiii", 0, 0, 19999, 19999) + small_exr[dw + 16:]
    files = [bytes(big_png), big_hdr, big_exr]
    try:
        import io
        from PIL import Image
        buf = io.BytesIO()
        Image.fromarray(np.zeros((16, 16, 3), np.uint8)).save(buf, "JPEG")
        jpg = bytearray(buf.getvalue())
        sof = jpg.index(b"\xff\xc0")
        jpg[sof + 5:sof + 9] = struct.pack(">HH", 65535, 65407)  # height, width
        files.append(bytes(jpg))
    except ImportError:
        pass
    for data in files:
        t0 = time.perf_counter()
        with pytest.raises(ptcore.PtError, match="exceeds the image size limit"):
            ptcore.Assets(quad_glb(), data)
        assert time.perf_counter() - t0 < 1.0


def test_error_messages_are_valid_text():
    # the message quotes the file's bytes: invalid UTF-8 becomes '?', valid UTF-8 is kept
    bad = b"#?RADIANCE\nFORMAT=\xf9\xfe\xc3(\xc3\xa9\n\n-Y 1 +X 1\n" + bytes(4)
    with pytest.raises(ptcore.PtError) as e:
        ptcore.Assets(quad_glb(), bad)
    msg = str(e.value)
    assert "unsupported HDR pixel format ???(é" in msg
