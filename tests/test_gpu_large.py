"""GPU parity on the 4.2M-triangle textured scene at its benchmark size:
3840x2160, depth 8, seed 3. It has a 1202^2 heightfield, 4096 icospheres,
8 Perlin 2048^2 textures (sRGB albedo, linear roughness) and an RGBE sky.

The bar is the one in test_gpu_parity: hit ids and distances bit-exact, and
radiance within the stated tolerance, which the images also meet bit for bit.
The checks run against the CPU oracle on the identical BVH:
- the full 4K frame of primary hits;
- tiles of radiance under both schedules.
"""
import numpy as np
import pytest

import binding as oracle
import ptcore
from test_gpu_parity import assert_parity

pytestmark = pytest.mark.gpu

W, H, BOUNCES = 3840, 2160, 8


@pytest.fixture(scope="module")
def large():
    print("generating the 4.2M-triangle scene", flush=True)
    return ptcore.World.generate(ptcore.SCENE_LARGE, 3, W, H)


@pytest.fixture(scope="module")
def large_scene(ctx, large):
    s = ptcore.Scene(ctx, large.desc)
    yield s
    s.close()


def test_primary_hits_large_full_frame_bitexact(ctx, large, large_scene):
    t = ptcore.Tracer(ctx, large_scene, W, H)
    tri, tt = t.primary_hits(5)
    ref_tri, ref_t = oracle.primary_hits(large.desc, W, H, 5)
    mism = int((tri != ref_tri).sum())
    print(f"large primary hits: {tri.size} pixels, {mism} mismatches, {int((tri >= 0).sum())} hits", flush=True)
    assert mism == 0
    assert np.array_equal(tt[tri >= 0], ref_t[ref_tri >= 0])
    t.close()


@pytest.mark.parametrize("schedule", [ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT])
@pytest.mark.parametrize("origin", [(1856, 1040), (200, 1500), (3500, 600)])
def test_radiance_large_tiles(ctx, large, large_scene, schedule, origin):
    tw, th, spp = 96, 64, 2
    t = ptcore.Tracer(ctx, large_scene, tw, th, schedule)
    t.set_tile((W, H), origin, (tw, th))
    t.set_max_bounces(BOUNCES)
    t.trace(spp)
    gpu = t.accumulation()
    ref, _, _ = oracle.render_f32(large.desc, W, H, BOUNCES, spp, tile=(tw, th), origin=origin)
    assert_parity(gpu, ref, f"large tile {origin} schedule={schedule}")
    t.close()
