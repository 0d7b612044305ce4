"""GPU parity: HIP kernels through the C ABI against the CPU oracle.

Bar (stated by the spec of this path): primary-ray hit triangle ids are
bit-exact on the identical BVH; per-pixel radiance under the reference's RNG
seeding agrees with the FP32 oracle to mean-image relative L1 <= 1e-4 with
<= 0.1% of pixels diverging (a pixel diverges when its relative difference
exceeds 1e-3). With the deterministic math layer and the shader's evaluation
order on both sides, both round identically, so the tests also require the
images to be bit-equal (checked beside the tolerance, which they report).
"""
import numpy as np
import pytest

import binding as oracle
import ptcore

pytestmark = pytest.mark.gpu

REL_L1_TOL = 1e-4
DIVERGE_FRAC_TOL = 1e-3
PIXEL_REL_TOL = 1e-3


def compare(gpu, ref):
    g, r = gpu[..., :3].astype(np.float64), ref[..., :3].astype(np.float64)
    rel_l1 = np.abs(g - r).sum() / max(np.abs(r).sum(), 1e-300)
    scale = np.maximum(np.abs(r).max(axis=-1), 1e-30)
    pix_rel = np.abs(g - r).max(axis=-1) / scale
    diverging = float((pix_rel > PIXEL_REL_TOL).mean())
    exact = float(np.all(gpu[..., :3] == ref[..., :3], axis=-1).mean())
    return rel_l1, diverging, exact


def assert_parity(gpu, ref, label, bit_exact=True):
    rel_l1, diverging, exact = compare(gpu, ref)
    print(f"{label}: rel L1 {rel_l1:.3g}, diverging {diverging:.3%}, bit-exact pixels {exact:.3%}", flush=True)
    assert np.isfinite(gpu).all()
    assert rel_l1 <= REL_L1_TOL, f"{label}: rel L1 {rel_l1}"
    assert diverging <= DIVERGE_FRAC_TOL, f"{label}: diverging fraction {diverging}"
    if bit_exact:
        assert exact == 1.0, f"{label}: {1 - exact:.4%} of pixels differ in some bit"


@pytest.fixture(scope="module")
def cornell_scene(ctx, cornell):
    s = ptcore.Scene(ctx, cornell.desc)
    yield s
    s.close()


@pytest.fixture(scope="module")
def field_scene(ctx, field):
    s = ptcore.Scene(ctx, field.desc)
    yield s
    s.close()


@pytest.mark.parametrize("sample", [1, 2, 7])
def test_primary_hits_cornell_bitexact(ctx, cornell, cornell_scene, sample):
    t = ptcore.Tracer(ctx, cornell_scene, 256, 256)
    tri, tt = t.primary_hits(sample)
    ref_tri, ref_t = oracle.primary_hits(cornell.desc, 256, 256, sample)
    assert np.array_equal(tri, ref_tri)
    assert np.array_equal(tt[tri >= 0], ref_t[ref_tri >= 0])
    t.close()


def test_primary_hits_field_full_frame_bitexact(ctx, field, field_scene):
    t = ptcore.Tracer(ctx, field_scene, 1920, 1080)
    tri, tt = t.primary_hits(1)
    ref_tri, ref_t = oracle.primary_hits(field.desc, 1920, 1080, 1)
    mism = int((tri != ref_tri).sum())
    print(f"field primary hits: {tri.size} pixels, {mism} mismatches, {int((tri >= 0).sum())} hits", flush=True)
    assert mism == 0
    assert np.array_equal(tt[tri >= 0], ref_t[ref_tri >= 0])
    t.close()


@pytest.mark.parametrize("schedule", [ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT])
def test_radiance_cornell(ctx, cornell, cornell_scene, schedule):
    w = h = 256
    spp, bounces = 16, 8
    t = ptcore.Tracer(ctx, cornell_scene, w, h, schedule)
    t.set_max_bounces(bounces)
    done = 0
    while done < spp:
        done += t.trace(spp - done)
    assert t.sample_count == spp
    gpu = t.accumulation()
    storage = t.storage()
    ref, ref_storage, _ = oracle.render_f32(cornell.desc, w, h, bounces, spp)
    assert_parity(gpu, ref, f"cornell schedule={schedule}")
    assert np.allclose(storage[..., :3], ref_storage[..., :3], rtol=1e-5, atol=0)
    t.close()


@pytest.mark.parametrize("schedule", [ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT])
def test_radiance_field_tile(ctx, field, field_scene, schedule):
    # a 256x192 tile of the 1080p frame, 2 samples, depth 6
    iw, ih, tw, th, ox, oy = 1920, 1080, 256, 192, 832, 444
    spp, bounces = 2, 6
    t = ptcore.Tracer(ctx, field_scene, tw, th, schedule)
    t.set_tile((iw, ih), (ox, oy), (tw, th))
    t.set_max_bounces(bounces)
    t.trace(spp)
    gpu = t.accumulation()
    ref, _, _ = oracle.render_f32(field.desc, iw, ih, bounces, spp, tile=(tw, th), origin=(ox, oy))
    assert_parity(gpu, ref, f"field tile schedule={schedule}")
    t.close()


def test_wavefront_equals_megakernel_bitwise(ctx, field_scene):
    out = []
    for schedule in (ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT):
        t = ptcore.Tracer(ctx, field_scene, 480, 270, schedule)
        t.set_tile((1920, 1080), (720, 405), (480, 270))
        t.set_max_bounces(6)
        t.trace(3)
        out.append(t.accumulation())
        t.close()
    assert np.array_equal(out[0], out[1])


def test_intersect_queries_match_oracle(ctx, field, field_scene):
    rng = np.random.default_rng(7)
    n = 200_000
    o = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
    o[:, 1] = rng.uniform(-0.3, 0.8, n)
    d = rng.normal(size=(n, 3)).astype(np.float32)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([o, d], axis=1).astype(np.float32)
    hit, t, v, w, pairs, tris = field_scene.intersect(rays)
    rh, rt, rv, rw, rp, rtr = oracle.intersect(field.desc, rays)
    assert np.array_equal(hit, rh)
    m = hit >= 0
    assert np.array_equal(t[m], rt[m]) and np.array_equal(v[m], rv[m]) and np.array_equal(w[m], rw[m])
    assert np.array_equal(pairs, rp) and np.array_equal(tris, rtr)
    tmax = np.where(m, rt * 0.999, np.float32(3.0e38)).astype(np.float32)
    occ = field_scene.occluded(rays, tmax)
    rocc = oracle.occluded(field.desc, rays, tmax)
    assert np.array_equal(occ, rocc)


ALPHA_MASK, ALPHA_BLEND = 1, 2


@pytest.fixture(scope="module")
def alpha_cornell(ctx):
    # the Cornell box with alpha-tested materials: a stochastic (blend) tall box, a cut-out (mask,
    # alpha < cutoff) back wall that shows the sky, and a mask short box that passes the cutoff
    w = ptcore.World.generate(ptcore.SCENE_CORNELL, 1, 256, 256)
    m = w.arrays()["materials"]
    for k, mode, alpha in ((5, ALPHA_BLEND, 0.55), (2, ALPHA_MASK, 0.3), (6, ALPHA_MASK, 0.8)):
        m["alpha_mode"][k] = mode
        m["base_color"][k, 3] = alpha
    assert (m["alpha_mode"][[2, 5, 6]] != 0).all()
    s = ptcore.Scene(ctx, w.desc)
    yield w, s
    s.close()


@pytest.mark.parametrize("sample", [1, 5])
def test_alpha_primary_hits_bitexact(ctx, alpha_cornell, sample):
    w, s = alpha_cornell
    t = ptcore.Tracer(ctx, s, 256, 256)
    tri, tt = t.primary_hits(sample)
    ref_tri, ref_t = oracle.primary_hits(w.desc, 256, 256, sample)
    assert np.array_equal(tri, ref_tri)
    assert np.array_equal(tt[tri >= 0], ref_t[ref_tri >= 0])
    t.close()


@pytest.mark.parametrize("schedule", [ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT])
def test_alpha_radiance_cornell(ctx, alpha_cornell, schedule):
    w, s = alpha_cornell
    spp, bounces = 8, 8
    t = ptcore.Tracer(ctx, s, 256, 256, schedule)
    t.set_max_bounces(bounces)
    t.trace(spp)
    gpu = t.accumulation()
    ref, _, _ = oracle.render_f32(w.desc, 256, 256, bounces, spp)
    assert_parity(gpu, ref, f"alpha cornell schedule={schedule}")
    t.close()


@pytest.mark.parametrize("schedule", [ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT])
def test_tile_past_the_image_edge_keeps_outside_pixels(ctx, field, field_scene, schedule):
    """A 37x23 tile (not a multiple of the 8x8 blocks) moved from inside the
    image to its corner, where 20x10 of its pixels remain in the image. Only
    those are traced (path_tracer.rs:334-338, trace.slang:731-733); the others
    keep what the previous tile position accumulated, as the reference
    leaves them. Compared with the oracle continuing from the same buffer."""
    iw, ih, tw, th, bounces = 1920, 1080, 37, 23, 6
    t = ptcore.Tracer(ctx, field_scene, tw, th, schedule)
    t.set_max_bounces(bounces)
    t.set_tile((iw, ih), (700, 500), (tw, th))
    t.trace(2)
    first = t.accumulation()
    ref_first, _, _ = oracle.render_f32(field.desc, iw, ih, bounces, 2, tile=(tw, th), origin=(700, 500))
    assert_parity(first, ref_first, f"inner tile schedule={schedule}")
    t.set_tile((iw, ih), (iw - 20, ih - 10), (tw, th))
    t.trace(3)
    second = t.accumulation()
    t.close()
    ref, _, _ = oracle.render_f32(field.desc, iw, ih, bounces, 3, tile=(tw, th), origin=(iw - 20, ih - 10),
                                  accum=ref_first.copy())
    assert_parity(second[:10, :20], ref[:10, :20], f"edge tile schedule={schedule}")
    assert np.array_equal(second, ref), "pixels outside the image must keep their previous values"


@pytest.mark.parametrize("bounces", [0, 1, 2])
def test_shallow_depths_match_oracle(ctx, cornell, cornell_scene, bounces):
    """max_bounces 0 (camera-visible emission and direct light only), 1 and 2:
    the terminal-bounce rules (no BSDF sample at the last surface, NEE without
    MIS weight there) under both schedules."""
    w = h = 128
    ref, _, _ = oracle.render_f32(cornell.desc, 256, 256, bounces, 4, tile=(w, h), origin=(64, 64))
    for schedule in (ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT):
        t = ptcore.Tracer(ctx, cornell_scene, w, h, schedule)
        t.set_tile((256, 256), (64, 64), (w, h))
        t.set_max_bounces(bounces)
        t.trace(4)
        assert_parity(t.accumulation(), ref, f"depth {bounces} schedule={schedule}")
        t.close()


def test_profiling_levels(ctx, cornell, cornell_scene):
    """Level 1 times every kernel class, level 3 only the traversal launches,
    level 2 adds the device counters; the image is the same at every level."""
    images = []
    for level in (0, 1, 2, 3):
        t = ptcore.Tracer(ctx, cornell_scene, 64, 64)
        t.set_max_bounces(4)
        t.set_profiling(level)
        t.trace(2)
        st = t.stats()
        images.append(t.accumulation())
        if level in (1, 2):
            assert st["extend_ms"] > 0 and st["shade_ms"] > 0 and st["total_ms"] > 0
        if level == 3:
            assert st["extend_ms"] > 0 and st["extend_launches"] == 6
            assert st["shade_ms"] == 0 and st["total_ms"] == 0
        assert (st["pair_tests"] > 0) == (level == 2)
        t.close()
    for img in images[1:]:
        assert np.array_equal(img, images[0])
    with pytest.raises(ptcore.PtError):
        t = ptcore.Tracer(ctx, cornell_scene, 8, 8)
        try:
            t.set_profiling(4)
        finally:
            t.close()


@pytest.mark.parametrize("atlas_size", [12, 16])
def test_textured_materials_match_oracle(ctx, atlas_size):
    """Textured materials at random uvs, inside and outside [0, 1] (the fetch
    clamps), with a side that is a multiple of 8 and one that is not: every
    material samples a random base-colour layer and the metallic-roughness
    layer, and both schedules reproduce the oracle's image."""
    rng = np.random.default_rng(atlas_size)
    g = ptcore.Generated(ptcore.SCENE_CORNELL, 1, 128, 128)
    a = g.arrays()
    geo, att, mat = a["geometries"].copy(), a["attributes"].copy(), a["materials"].copy()
    att["uv"] = rng.uniform(-0.2, 1.2, att["uv"].shape).astype(np.float32)
    mat["base_color_tex_layer"] = rng.integers(0, 2, len(mat))
    mat["metallic_roughness_tex_layer"] = 0
    mat["base_color_uv"] = [0.0, 0.0, 1.0, 1.0]
    mat["metallic_roughness_uv"] = [0.0, 0.0, 1.0, 1.0]
    srgb = rng.integers(0, 256, (2, atlas_size, atlas_size, 4), dtype=np.uint8)
    lin = rng.integers(0, 256, (1, atlas_size, atlas_size, 4), dtype=np.uint8)
    inp = ptcore.WorldInput(geo, att, mat, a["skybox"], srgb=srgb, linear=lin, atlas_size=atlas_size)
    world = inp.build()
    scene = ptcore.Scene(ctx, world.desc)
    w = h = 64
    ref, _, _ = oracle.render_f32(world.desc, 128, 128, 4, 2, tile=(w, h), origin=(32, 32))
    tri_ref, _ = oracle.primary_hits(world.desc, 128, 128, 1)
    for schedule in (ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT):
        t = ptcore.Tracer(ctx, scene, w, h, schedule)
        t.set_tile((128, 128), (32, 32), (w, h))
        t.set_max_bounces(4)
        t.trace(2)
        assert_parity(t.accumulation(), ref, f"atlas {atlas_size} schedule={schedule}")
        t.close()
    scene.close()
    assert (tri_ref >= 0).any()


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_openpbr_materials_match_oracle(ctx, seed):
    """Materials drawn across the OpenPBR parameters the shading reads:
    anisotropic GGX, partial metalness and specular weight, coloured specular,
    IOR 1.1-2.5, and tangent-space normal maps at several scales with random
    tangents of either handedness. Both schedules reproduce the oracle's image."""
    from materials import random_openpbr_world
    world = random_openpbr_world(seed)
    scene = ptcore.Scene(ctx, world.desc)
    w = h = 64
    ref, _, _ = oracle.render_f32(world.desc, 128, 128, 6, 4, tile=(w, h), origin=(32, 32))
    for schedule in (ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT):
        t = ptcore.Tracer(ctx, scene, w, h, schedule)
        t.set_tile((128, 128), (32, 32), (w, h))
        t.set_max_bounces(6)
        t.trace(4)
        assert_parity(t.accumulation(), ref, f"materials seed {seed} schedule={schedule}")
        t.close()
    scene.close()


def test_huge_roughness_takes_the_general_estimate(ctx):
    """Surfaces with GGX alpha beyond 2^56 run the E_dielectric estimate in its general form (full
    normalisation of each sampled half vector); the image still equals the oracle's under both
    schedules, as do the other surfaces, which take the restricted-domain form."""
    from materials import random_openpbr_world
    world = random_openpbr_world(5, huge_roughness=True)
    scene = ptcore.Scene(ctx, world.desc)
    w = h = 64
    ref, _, _ = oracle.render_f32(world.desc, 128, 128, 6, 4, tile=(w, h), origin=(32, 32))
    assert np.isfinite(ref).all()
    for schedule in (ptcore.SCHEDULE_MEGAKERNEL, ptcore.SCHEDULE_WAVEFRONT):
        t = ptcore.Tracer(ctx, scene, w, h, schedule)
        t.set_tile((128, 128), (32, 32), (w, h))
        t.set_max_bounces(6)
        t.trace(4)
        assert_parity(t.accumulation(), ref, f"huge roughness schedule={schedule}")
        t.close()
    scene.close()


def test_context_info(ctx):
    """pt_context_get_info reports the device the kernels were built for."""
    i = ctx.info()
    print(i, flush=True)
    assert i["arch"].startswith("gfx950") and i["cu_count"] > 0 and i["clock_khz"] > 0
    assert i["hbm_bytes"] > (1 << 30) and i["l2_bytes"] > 0


def test_measure_gather(ctx):
    """pt_context_measure_gather gives a positive lane-access rate, a higher one when four lanes share each
    64-byte segment than when every lane has its own, and refuses other lane counts."""
    own = ctx.measure_gather(1)
    shared = ctx.measure_gather(4)
    print(f"gather: {own:.3f} lane accesses per CU per cycle (own segment), {shared:.3f} (four lanes)", flush=True)
    assert 0.05 < own < 64.0 and shared > 1.5 * own
    with pytest.raises(ptcore.PtError):
        ctx.measure_gather(2)


def test_fast_math_matches_correct_rounding(ctx):
    """The restricted-domain square root equals the device's correctly rounded sqrtf on every float32 in its
    domain (+-0, [2^-96, +inf], NaN) and on the negative normals, -inf and negative NaNs (all NaN); the
    restricted-domain division equals a / b on 2^28 operand pairs of its domain for each of two seeds."""
    assert ctx.check_fast_math(0x00000000, 0x00000000) == (0, 0)  # +0
    assert ctx.check_fast_math(0x80000000, 0x80000000) == (0, 0)  # -0
    assert ctx.check_fast_math(0x0F800000, 0x7FFFFFFF) == (0, 0)  # 2^-96 .. +inf, positive NaNs
    assert ctx.check_fast_math(0x80800000, 0xFFFFFFFF) == (0, 0)  # negative normals, -inf, negative NaNs
    # outside the domain the hardware root alone does not give sqrtf's result (the compiler's sequence
    # rescales these arguments first); no caller passes them: each caller's argument is 0, NaN, or at
    # least 2^-24 in magnitude
    tiny, _ = ctx.check_fast_math(0x00000001, 0x0F7FFFFF)
    neg_sub, _ = ctx.check_fast_math(0x80000001, 0x807FFFFF)
    print(f"sqrt outside the domain: (0, 2^-96) {tiny} of {0x0F7FFFFF} differ, negative subnormals "
          f"{neg_sub} of {0x7FFFFF}", flush=True)
    for seed in (1, 2):
        assert ctx.check_fast_math(0, 0, 1 << 28, seed) == (0, 0), f"division seed {seed}"
    with pytest.raises(ptcore.PtError):
        ctx.check_fast_math(5, 4)


def test_axis_aligned_rays_match_oracle(ctx, field, field_scene):
    """Rays with zero direction components, whose reciprocals are infinite: the slab test then
    meets 0 * inf = NaN where an origin lies on a box plane, and the min/max rules decide. Along the
    six axes (signed zeros included), in the coordinate planes, and from origins placed on the
    Cornell box's axis-aligned walls and the field scene's node bounds: hit ids, distances,
    barycentrics, per-ray counters and occlusion equal the oracle's."""
    rng = np.random.default_rng(11)
    cornell = ptcore.World.generate(ptcore.SCENE_CORNELL, 1, 64, 64)
    cornell_scene = ptcore.Scene(ctx, cornell.desc)
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                     [-0.0, 0, 1], [0, -0.0, -1], [1, -0.0, -0.0]], dtype=np.float32)
    try:
        for world, scene, lo, hi in ((field, field_scene, -1.2, 1.2), (cornell, cornell_scene, -1.0, 1.0)):
            n = 60_000
            o = rng.uniform(lo, hi, (n, 3)).astype(np.float32)
            d = np.empty((n, 3), np.float32)
            k = n // 3
            d[:k] = axes[rng.integers(0, len(axes), k)]
            plane = rng.normal(size=(k, 3)).astype(np.float32)
            plane[np.arange(k), rng.integers(0, 3, k)] = 0.0
            d[k:2 * k] = plane / np.linalg.norm(plane, axis=1, keepdims=True)
            # nearly axis-aligned: tiny components, whose reciprocals are huge but finite
            near = axes[rng.integers(0, 6, n - 2 * k)].copy()
            near += (rng.choice([-1.0, 1.0], (n - 2 * k, 3)) * 10.0 ** rng.uniform(-30, -8, (n - 2 * k, 3))).astype(np.float32)
            d[2 * k:] = near
            # a third of the origins on planes the boxes use: the scene's extreme coordinates
            bv = world.arrays()["bvh_nodes"]
            planes = np.concatenate([bv["aabb_min"].ravel(), bv["aabb_max"].ravel()]).astype(np.float32)
            on = rng.uniform(size=n) < 0.33
            axis = rng.integers(0, 3, n)
            o[on, axis[on]] = rng.choice(planes, on.sum())
            # and, for half the rays with zero direction components, a zero origin coordinate there:
            # -o * (1/d) is then NaN too, the axis drops out of every slab test, and the walk goes deep
            zero = (d == 0.0) & (rng.uniform(size=(n, 1)) < 0.5)
            o[zero] = 0.0
            rays = np.concatenate([o, d], axis=1).astype(np.float32)
            hit, t, v, w, pairs, tris = scene.intersect(rays)
            rh, rt, rv, rw, rp, rtr = oracle.intersect(world.desc, rays)
            assert np.array_equal(hit, rh)
            m = hit >= 0
            # exactly axis- or plane-parallel rays miss unless their origin's coordinate on that axis
            # is 0: a box plane's fma(lo, 1/d, -o * (1/d)) is then either inf - inf = NaN, which the min/max
            # rules drop, or an infinity that empties the interval (the reference's own behaviour)
            assert m[:2 * k].any() and m[2 * k:].any() and (pairs[:2 * k] > 1).any()
            assert np.array_equal(t[m], rt[m]) and np.array_equal(v[m], rv[m]) and np.array_equal(w[m], rw[m])
            assert np.array_equal(pairs, rp) and np.array_equal(tris, rtr)
            tmax = np.where(m, rt * 0.999, np.float32(3.0e38)).astype(np.float32)
            assert np.array_equal(scene.occluded(rays, tmax), oracle.occluded(world.desc, rays, tmax))
    finally:
        cornell_scene.close()
