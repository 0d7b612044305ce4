"""GPU: editing a world and applying its updates to a resident scene
(PathTracer::apply_updates, path_tracer.rs:272-282) renders exactly what a
scene built fresh from the edited input renders, and resets accumulation only
when the updates ask for it."""
import math

import numpy as np
import pytest

import ptcore

pytestmark = pytest.mark.gpu

LIGHT = 7


def render(ctx, scene, spp=4, w=96, h=96, schedule=ptcore.SCHEDULE_WAVEFRONT, tracer=None):
    t = tracer or ptcore.Tracer(ctx, scene, w, h, schedule)
    t.set_max_bounces(6)
    t.trace(spp)
    out = t.accumulation()
    if tracer is None:
        t.close()
    return out


@pytest.mark.parametrize("schedule", [ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT])
def test_applied_edits_equal_fresh_scene(ctx, schedule):
    g = ptcore.Generated(ptcore.SCENE_CORNELL, 1, 96, 96)
    a = ptcore.World(g.desc, keep=g)
    sa = ptcore.Scene(ctx, a.desc)
    t = ptcore.Tracer(ctx, sa, 96, 96, schedule)
    t.set_max_bounces(6)
    t.trace(2)
    assert t.sample_count == 2
    # edits: emitter brighter and tinted, a wall recoloured, camera turned, sky dimmer
    m = a.material(LIGHT)[0].copy()
    m["emission_color_and_luminance"] = (1.0, 0.6, 0.3, 40.0)
    assert a.edit_material(LIGHT, m)
    wall = a.material(3)[0].copy()
    wall["base_color"][:3] = (0.2, 0.7, 0.3)
    assert a.edit_material(3, wall)
    yaw = float(np.float32(5.0) * (np.float32(math.pi) / np.float32(180.0)))
    a.edit_camera(yaw=yaw)
    a.edit_skybox_luminance(0.25)
    u = a.take_updates()
    assert u["reset_accumulation"]
    sa.apply_updates(u)
    t.trace(4)
    assert t.sample_count == 4  # the update reset the accumulation
    got = t.accumulation()
    t.close()

    def edit(gen):
        mats = ptcore._view(gen.desc.materials, ptcore.MATERIAL_DTYPE, gen.desc.material_count)
        mats[LIGHT] = m
        mats[3] = wall
        gen.desc.camera_yaw_degrees = 5.0
        gen.desc.skybox_luminance = 0.25
    g2 = ptcore.Generated(ptcore.SCENE_CORNELL, 1, 96, 96)
    edit(g2)
    b = ptcore.World(g2.desc, keep=g2)
    sb = ptcore.Scene(ctx, b.desc)
    want = render(ctx, sb, 4, schedule=schedule)
    sb.close()
    sa.close()
    assert np.array_equal(got, want)


def test_updates_without_reset_keep_accumulating(ctx):
    g = ptcore.Generated(ptcore.SCENE_CORNELL, 1, 64, 64)
    a = ptcore.World(g.desc, keep=g)
    s = ptcore.Scene(ctx, a.desc)
    t = ptcore.Tracer(ctx, s, 64, 64)
    t.trace(3)
    a.edit_camera()  # same camera: nothing changes
    u = a.take_updates()
    assert not u["reset_accumulation"]
    s.apply_updates(u)
    t.trace(2)
    assert t.sample_count == 5
    t.close()
    s.close()
