"""World editing and take_updates (world.rs:589-884) and pick_material
(world.rs:1595-1688). CPU only: host-side producers.

Pinning: an edit followed by take_updates must leave the world exactly as a
fresh pt_world_build of the edited input would (camera, emissive alias table,
fluxes and light probabilities), so the update path cannot drift from the
builder that test_world.py pins against its Python restatement. pick_material
is pinned against a float64 brute-force ray cast over every triangle.
"""
import ctypes as C
import math

import numpy as np
import pytest

import ptcore

CORNELL_LIGHT = 7  # the emitter's material in the Cornell generator


def fresh(kind=ptcore.SCENE_CORNELL, seed=1, edit=None):
    g = ptcore.Generated(kind, seed, 256, 256)
    if edit:
        edit(g)
    return ptcore.World(g.desc, keep=g), g


def scene_arrays(world):
    d = world.desc
    return (bytes(d.camera), bytes(d.scene_uniforms), bytes(d.skybox),
            ptcore._view(d.scene_alias, ptcore.SCENE_ALIAS_DTYPE, d.scene_alias_count).tobytes())


def rads(deg):
    # the builder's degrees -> radians: deg * (pi_f32 / 180) in float32
    return float(np.float32(deg) * (np.float32(math.pi) / np.float32(180.0)))


def test_no_pending_edits_reports_nothing():
    w, _ = fresh()
    u = w.take_updates()
    assert not any(u[k] for k in ("camera_changed", "scene_uniforms_changed", "scene_alias_changed",
                                  "skybox_uniforms_changed", "reset_accumulation"))
    assert len(u["material_indices"]) == 0
    # re-submitting the current camera is not a change either
    w.edit_camera()
    assert not w.take_updates()["camera_changed"]


def test_camera_edit_equals_fresh_build():
    a, _ = fresh()
    a.edit_camera(yaw=rads(12.0), pitch=rads(-7.0), fov=rads(60.0))
    u = a.take_updates()
    assert u["camera_changed"] and u["reset_accumulation"]
    assert not u["scene_uniforms_changed"] and not u["scene_alias_changed"]

    def edit(g):
        g.desc.camera_yaw_degrees, g.desc.camera_pitch_degrees, g.desc.camera_fov_degrees = 12.0, -7.0, 60.0
    b, _ = fresh(edit=edit)
    assert bytes(u["camera"]) == bytes(b.desc.camera)
    assert a.camera() == b.camera()
    # the initial camera is the scene-bounds camera, untouched by edits
    assert a.camera(initial=True) == b.camera(initial=True)
    assert a.camera(initial=True)["yaw"] == 0.0 and a.camera(initial=True)["fov"] == pytest.approx(math.pi / 2)


def test_emission_edit_equals_fresh_build():
    a, _ = fresh()
    cur, ini = a.material(CORNELL_LIGHT)
    m = cur.copy()
    m["emission_color_and_luminance"][3] = 37.5
    m["emission_color_and_luminance"][0] = 0.25
    assert a.edit_material(CORNELL_LIGHT, m)
    u = a.take_updates()
    assert u["scene_alias_changed"] and u["scene_uniforms_changed"] and u["reset_accumulation"]
    assert list(u["material_indices"]) == [CORNELL_LIGHT]
    assert u["materials"][0].tobytes() == m.tobytes()

    def edit(g):
        mats = ptcore._view(g.desc.materials, ptcore.MATERIAL_DTYPE, g.desc.material_count)
        mats[CORNELL_LIGHT] = m
    b, _ = fresh(edit=edit)
    assert scene_arrays(a) == scene_arrays(b)
    assert bytes(u["scene_uniforms"]) == bytes(b.desc.scene_uniforms)
    assert u["scene_alias"].tobytes() == scene_arrays(b)[3]
    # the initial material is kept for emission scaling
    assert a.material(CORNELL_LIGHT)[1].tobytes() == ini.tobytes()


def test_non_emission_edit_keeps_lights():
    a, _ = fresh()
    m = a.material(2)[0].copy()
    m["base_color"][:3] = (0.9, 0.1, 0.2)
    m["alpha_mode"] = 2  # not editable: stays opaque
    assert a.edit_material(2, m)
    u = a.take_updates()
    assert u["reset_accumulation"] and not u["scene_alias_changed"] and not u["scene_uniforms_changed"]
    got = a.material(2)[0]
    assert tuple(got["base_color"][:3]) == pytest.approx((0.9, 0.1, 0.2))
    assert got["alpha_mode"] == 0


def test_edit_material_rejects_non_finite_and_unknown():
    a, _ = fresh()
    m = a.material(1)[0].copy()
    m["specular_ior"] = np.inf
    assert not a.edit_material(1, m)
    assert not a.edit_material(10_000, a.material(1)[0])
    assert not a.take_updates()["reset_accumulation"]


def test_emission_scale():
    a, _ = fresh()
    base = a.material(CORNELL_LIGHT)[1]["emission_color_and_luminance"][3]
    a.edit_emission_scale(2.5)
    assert a.emission_scale() == 2.5
    u = a.take_updates()
    assert list(u["material_indices"]) == [CORNELL_LIGHT]  # only emissive materials are scaled
    assert a.material(CORNELL_LIGHT)[0]["emission_color_and_luminance"][3] == np.float32(base) * np.float32(2.5)
    a.edit_emission_scale(1.0)
    a.edit_emission_scale(2.5)  # back to the current state: nothing pending
    assert len(a.take_updates()["material_indices"]) == 0
    a.edit_emission_scale(-3.0)
    assert a.emission_scale() == 0.0


def test_skybox_luminance_edit_equals_fresh_build():
    a, _ = fresh(ptcore.SCENE_FIELD, 2)
    a.edit_skybox_luminance(3.25)
    u = a.take_updates()
    assert u["skybox_uniforms_changed"] and u["scene_uniforms_changed"] and u["reset_accumulation"]
    assert not u["scene_alias_changed"]

    def edit(g):
        g.desc.skybox_luminance = 3.25
    b, _ = fresh(ptcore.SCENE_FIELD, 2, edit)
    assert scene_arrays(a)[1:3] == scene_arrays(b)[1:3]
    assert bytes(u["skybox_uniforms"]) == bytes(b.desc.skybox)
    a.edit_skybox_luminance(3.25)
    assert not a.take_updates()["skybox_uniforms_changed"]


def forward(yaw, pitch):
    # glam Mat3::from_rotation_y(yaw) * Mat3::from_rotation_x(pitch) * -Z
    return np.array([-math.sin(yaw) * math.cos(pitch), math.sin(pitch), -math.cos(yaw) * math.cos(pitch)])


def brute_pick(arrays, pos, d):
    v = arrays["geometries"]["v"].astype(np.float64).reshape(-1, 3, 3)
    e1, e2 = v[:, 1] - v[:, 0], v[:, 2] - v[:, 0]
    p = np.cross(d, e2)
    det = (e1 * p).sum(1)
    ok = np.abs(det) > 1e-12
    inv = np.where(ok, 1.0 / np.where(ok, det, 1.0), 0.0)
    s = pos - v[:, 0]
    uu = (s * p).sum(1) * inv
    q = np.cross(s, e1)
    vv = (q * d).sum(1) * inv
    t = (q * e2).sum(1) * inv
    hit = ok & (uu >= 0) & (vv >= 0) & (uu + vv <= 1) & (t > 0)
    if not hit.any():
        return -1, None
    t = np.where(hit, t, np.inf)
    i = int(np.argmin(t))
    return int(arrays["attributes"]["material"][i]), np.sort(t[hit])


def test_pick_material_matches_brute_force():
    for kind, seed in ((ptcore.SCENE_CORNELL, 1), (ptcore.SCENE_FIELD, 2)):
        w, _ = fresh(kind, seed)
        arr = w.arrays()
        cam = w.camera()
        rng = np.random.default_rng(5)
        checked = 0
        for k in range(60):
            pos = np.array(cam["position"], np.float32) + rng.normal(0, 0.2, 3).astype(np.float32)
            yaw, pitch = float(rng.uniform(-1.2, 1.2)), float(rng.uniform(-1.2, 0.6))
            got = w.pick_material(pos, yaw, pitch)
            want, ts = brute_pick(arr, pos.astype(np.float64), forward(yaw, pitch))
            if ts is not None and len(ts) > 1 and ts[1] - ts[0] < 1e-4 * ts[0]:
                continue  # near-tie between two surfaces: either answer is right
            assert got == want, (kind, k, got, want)
            checked += 1
        assert checked >= 40
    # looking away from everything
    w, _ = fresh()
    assert w.pick_material((0.0, 0.0, 50.0), math.pi, 0.0) == -1
    # the selection reported by take_updates follows the camera
    w.edit_camera(yaw=0.05)
    u = w.take_updates()
    assert u["selected_material"] == w.pick_material(w.camera()["position"], 0.05, w.camera()["pitch"])
