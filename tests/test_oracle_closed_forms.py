"""Pinning the FP32 oracle with what the mathematics fixes. CPU only.

Each check is chosen so that a plausible slip in the restatement (a dropped
term, a wrong sign or index, a transposed operand) fails it:
  * Fresnel: normal incidence ((n-1)/(n+1))^2, total internal reflection,
    the textbook s/p equations at random angles (float64 numpy);
  * MIS: power heuristic weights of two strategies sum to one;
  * modulated IOR: weight 0 gives 1, weight 1 gives the IOR;
  * EON diffuse: white furnace (albedo of a rho = 1 lobe integrates to ~1);
  * BSDF sampling vs its pdf: histogram of sampled directions against the
    integral of pdf_total over the same bins, and the rejected fraction;
  * light sampling: E[1/pdf] equals the solid angle of a triangle light
    (Van Oosterom-Strackee) and 4*pi for the sky; the sky pdf integrates to 1;
  * direct lighting: a Lambertian floor under a parallel square emitter
    converges to rho * L * F with the closed-form point-to-rectangle factor;
  * float vs double oracle on the same random stream;
  * BVH traversal: closest t equals brute force over all triangles, and rays
    aimed at shared edges and vertices of a closed mesh never miss.
"""
import ctypes as C
import math

import numpy as np
import pytest

import binding as oracle
import ptcore


def fv(*x):
    return np.ascontiguousarray(np.array(x, np.float32))


def fresnel_ref(n, cos_i):
    sin_t2 = (1 - cos_i ** 2) / n ** 2
    if sin_t2 >= 1:
        return 1.0
    cos_t = math.sqrt(1 - sin_t2)
    rs = (cos_i - n * cos_t) / (cos_i + n * cos_t)
    rp = (cos_t - n * cos_i) / (cos_t + n * cos_i)
    return 0.5 * (rs * rs + rp * rp)


def test_fresnel_closed_forms():
    L = oracle.lib()
    for n in (1.01, 1.33, 1.5, 2.4, 3.0):
        assert abs(L.oracle_fresnel_exact(n, 1.0) - ((n - 1) / (n + 1)) ** 2) < 1e-6
    assert L.oracle_fresnel_exact(0.7, 0.3) == 1.0  # sin_i = 0.95 > 0.7: TIR
    rng = np.random.default_rng(0)
    for _ in range(2000):
        n, c = rng.uniform(0.5, 3.0), rng.uniform(0.0, 1.0)
        assert abs(L.oracle_fresnel_exact(n, c) - fresnel_ref(np.float32(n), np.float32(c))) < 2e-5


def test_power_heuristic_partition_of_unity():
    L = oracle.lib()
    rng = np.random.default_rng(1)
    for _ in range(2000):
        a, b = np.exp(rng.uniform(-30, 30, 2))
        assert abs(L.oracle_power_heuristic(a, b) + L.oracle_power_heuristic(b, a) - 1.0) < 1e-6
    assert L.oracle_power_heuristic(0.0, 0.0) == 0.0
    assert L.oracle_power_heuristic(1e30, 1e30) == pytest.approx(0.5)


def test_specular_ior_modulated_limits():
    L = oracle.lib()
    for n in (1.0, 1.2, 1.5, 2.5):
        assert L.oracle_specular_ior_modulated(n, 0.0) == 1.0
        assert L.oracle_specular_ior_modulated(n, 1.0) == pytest.approx(n, rel=1e-5)
    assert L.oracle_specular_ior_modulated(0.8, 1.0) == pytest.approx(0.8, rel=1e-5)


def lobes(params, wo, wi, e=(0.0, 0.0, 0.0)):
    f = np.zeros(3, np.float32)
    pdf = C.c_float()
    oracle.lib().oracle_evaluate_lobes(oracle._ptr(params), oracle._ptr(wo), oracle._ptr(wi), oracle._ptr(fv(*e)),
                                       oracle._ptr(f), C.byref(pdf))
    return f, pdf.value


def hemisphere_grid(n_theta=96, n_phi=192):
    # midpoint grid in (cos theta, phi): equal-solid-angle cells
    mu = (np.arange(n_theta) + 0.5) / n_theta
    phi = (np.arange(n_phi) + 0.5) / n_phi * 2 * np.pi
    cell = (1.0 / n_theta) * (2 * np.pi / n_phi)
    return mu, phi, cell


def material(base=(1, 1, 1), weight=1.0, metal=0.0, rough_d=0.5, spec=(1, 1, 1), spec_w=0.0, ior=1.5, rough_s=0.3,
             aniso=0.0):
    return fv(*base, weight, metal, rough_d, *spec, spec_w, ior, rough_s, aniso)


@pytest.mark.parametrize("rough", [0.0, 0.3, 0.7, 1.0])
@pytest.mark.parametrize("mu_o", [0.2, 0.6, 0.95])
def test_eon_white_furnace(rough, mu_o):
    prm = material(rough_d=rough)
    wo = fv(math.sqrt(1 - mu_o ** 2), 0.0, mu_o)
    mu, phi, cell = hemisphere_grid(64, 96)
    total = 0.0
    for m in mu:
        s = math.sqrt(1 - m * m)
        for p in phi:
            f, _ = lobes(prm, wo, fv(s * math.cos(p), s * math.sin(p), m))
            total += f[0] * m * cell
    # the reference notes its white furnace "fails slightly" (trace.slang:1132)
    assert abs(total - 1.0) < 0.03, total


@pytest.mark.parametrize("case", [
    dict(m=material(base=(0.6, 0.5, 0.4), rough_d=0.8, spec_w=1.0, rough_s=0.4), wo=(0.5, 0.1, 0.86)),
    dict(m=material(base=(0.9, 0.8, 0.7), metal=1.0, spec_w=1.0, rough_s=0.25), wo=(0.3, -0.4, 0.866)),
    dict(m=material(base=(0.3, 0.3, 0.3), rough_d=0.3, spec_w=1.0, rough_s=0.8), wo=(0.8, 0.0, 0.6)),
    # anisotropic GGX (trace.slang:369-375): alpha_x != alpha_y stretches the lobe along the tangent
    dict(m=material(base=(0.5, 0.6, 0.7), metal=0.6, rough_d=0.5, spec_w=1.0, rough_s=0.5, aniso=0.7),
         wo=(0.4, 0.5, 0.77)),
    dict(m=material(base=(0.9, 0.7, 0.5), metal=1.0, spec_w=1.0, rough_s=0.7, aniso=0.4), wo=(-0.3, 0.6, 0.74)),
])
def test_bsdf_sampling_matches_pdf(case):
    prm, wo = case["m"], fv(*case["wo"])
    wo = fv(*(wo / np.linalg.norm(wo)))
    e = fv(0.04, 0.04, 0.04)
    rng = np.random.default_rng(4)
    n = 120_000
    r = rng.random((n, 3)).astype(np.float32)
    wi = np.zeros(3, np.float32)
    L = oracle.lib()
    nt, nphi = 8, 16
    hist = np.zeros((nt, nphi))
    rejected = 0
    for k in range(n):
        ok = L.oracle_sample_bsdf(oracle._ptr(prm), oracle._ptr(wo), oracle._ptr(e), float(r[k, 0]), float(r[k, 1]),
                                  float(r[k, 2]), oracle._ptr(wi))
        if not ok:
            rejected += 1
            continue
        it = min(int(wi[2] * nt), nt - 1)
        ip = int((math.atan2(wi[1], wi[0]) % (2 * math.pi)) / (2 * math.pi) * nphi) % nphi
        hist[it, ip] += 1
    # expected: integral of pdf_total over each bin (10x10 sub-cells)
    expect = np.zeros((nt, nphi))
    sub = 10
    for it in range(nt):
        for ip in range(nphi):
            acc = 0.0
            for a in range(sub):
                m = (it + (a + 0.5) / sub) / nt
                s = math.sqrt(1 - m * m)
                for b in range(sub):
                    p = (ip + (b + 0.5) / sub) / nphi * 2 * math.pi
                    _, pdf = lobes(prm, wo, fv(s * math.cos(p), s * math.sin(p), m), e)
                    acc += pdf
            expect[it, ip] = acc / (sub * sub) * (1.0 / nt) * (2 * math.pi / nphi) * n
    # Poisson z-scores of the bins: chi^2 per bin near 1 and no bin beyond 4.5 sigma
    mask = expect > 200
    z = (hist[mask] - expect[mask]) / np.sqrt(expect[mask])
    assert np.mean(z * z) < 1.5 and np.abs(z).max() < 4.5, (np.mean(z * z), np.abs(z).max())
    assert abs((n - rejected) - expect.sum()) / n < 0.02


def light_test_scene(sky_black=True):
    # floor + one emissive triangle above it
    floor = [[-5, 0, -5], [5, 0, -5], [5, 0, 5]], [[-5, 0, -5], [5, 0, 5], [-5, 0, 5]]
    tri = [[-0.4, 1.0, -0.3], [0.5, 1.2, -0.2], [0.1, 0.9, 0.6]]
    geo, attr = ptcore.triangles_from_vertices([*floor, tri], [0, 0, 1])
    mats = np.array([ptcore.default_material(specular_color_and_weight=(1, 1, 1, 0)),
                     ptcore.default_material(emission_color_and_luminance=(1, 1, 1, 10))])
    sky = np.zeros((8, 16, 4), np.float32) if sky_black else np.full((8, 16, 4), 0.5, np.float32)
    return ptcore.WorldInput(geo, attr, mats, sky, skybox_luminance=1.0).build(), np.array(tri, np.float64)


def solid_angle_triangle(p, tri):
    a, b, c = (np.asarray(v, np.float64) - p for v in tri)
    la, lb, lc = (np.linalg.norm(v) for v in (a, b, c))
    num = abs(np.dot(a, np.cross(b, c)))
    den = la * lb * lc + np.dot(a, b) * lc + np.dot(a, c) * lb + np.dot(b, c) * la
    return 2 * math.atan2(num, den)


def sample_light_many(desc, origin, n, seed=9):
    L = oracle.lib()
    st = C.c_uint32(oracle.lib().oracle_seed(seed, 1, 1))
    d, em = np.zeros(3, np.float32), np.zeros(3, np.float32)
    dist, pdf = C.c_float(), C.c_float()
    o = fv(*origin)
    out = np.zeros(n)
    tri = np.zeros(n, np.int32)
    for k in range(n):
        tri[k] = L.oracle_sample_light(C.byref(desc), oracle._ptr(o), C.byref(st), oracle._ptr(d), C.byref(dist),
                                       oracle._ptr(em), C.byref(pdf))
        out[k] = pdf.value
    return out, tri


def test_triangle_light_pdf_solid_angle():
    world, tri = light_test_scene(sky_black=True)
    u = world.desc.scene_uniforms
    assert u.emissive_triangles_count == 1 and list(u.light_probabilities) == [1.0, 0.0]
    a = world.arrays()
    # the emitter's BVH-order position
    emitter = [tuple(g["v"].reshape(3, 3).tolist()) for g in a["geometries"]].index(tuple(tri.astype(np.float32).tolist()))
    p = np.array([0.05, 0.0, 0.1])
    pdfs, tris = sample_light_many(world.desc, p, 60_000)
    assert (tris == emitter).all()
    est = np.mean(1.0 / pdfs)
    want = solid_angle_triangle(p, tri)
    assert abs(est - want) / want < 0.01, (est, want)


def test_sky_pdf_normalised_and_sampled(field):
    desc = field.desc
    # skybox_light_pdf integrates to 1 over the sphere
    nt, nphi = 256, 512
    mu = (np.arange(nt) + 0.5) / nt * 2 - 1
    phi = (np.arange(nphi) + 0.5) / nphi * 2 * np.pi
    total = 0.0
    d = np.zeros(3, np.float32)
    L = oracle.lib()
    for m in mu[::1]:
        s = math.sqrt(1 - m * m)
        for p in phi[::4]:
            d[:] = (s * math.cos(p), m, s * math.sin(p))
            total += L.oracle_skybox_light_pdf(C.byref(desc), oracle._ptr(d), 1.0)
    total *= (2.0 / nt) * (2 * np.pi / (nphi / 4))
    assert abs(total - 1.0) < 0.01, total
    # with sky-only light selection E[1/pdf] = 4 pi
    d2 = ptcore.SceneDesc.from_buffer_copy(desc)
    d2.scene_uniforms.light_probabilities[0] = 0.0
    d2.scene_uniforms.light_probabilities[1] = 1.0
    pdfs, tris = sample_light_many(d2, (0.0, 0.5, 0.0), 60_000)
    assert (tris == -1).all()
    est = np.mean(1.0 / pdfs)
    assert abs(est - 4 * np.pi) / (4 * np.pi) < 0.02, est


def rect_factor_corner(a, b, c):
    # differential area to a parallel rectangle a x b whose corner is above it at distance c
    A, B = a / c, b / c
    return (A / math.sqrt(1 + A * A) * math.atan(B / math.sqrt(1 + A * A)) +
            B / math.sqrt(1 + B * B) * math.atan(A / math.sqrt(1 + B * B))) / (2 * math.pi)


def test_direct_lighting_closed_form():
    rho, Le, h, s = 0.6, 5.0, 1.0, 0.8
    floor = [[-5, 0, -5], [5, 0, 5], [5, 0, -5]], [[-5, 0, -5], [-5, 0, 5], [5, 0, 5]]
    half = s / 2
    light = [[-half, h, -half], [half, h, -half], [half, h, half]], [[-half, h, -half], [half, h, half], [-half, h, half]]
    geo, attr = ptcore.triangles_from_vertices([*floor, *light], [0, 0, 1, 1])
    mats = np.array([ptcore.default_material(base_color=(rho, rho, rho, 1), base_diffuse_roughness=0.0,
                                             specular_color_and_weight=(1, 1, 1, 0)),
                     ptcore.default_material(emission_color_and_luminance=(1, 1, 1, Le),
                                             specular_color_and_weight=(1, 1, 1, 0))])
    sky = np.zeros((4, 8, 4), np.float32)
    world = ptcore.WorldInput(geo, attr, mats, sky, skybox_luminance=0.0, camera_position=(0.0, 0.5, 0.0),
                              yaw=0.0, pitch=-90.0, fov=0.5, aspect=1.0).build()
    acc = oracle.render_f32(world.desc, 1, 1, 0, 40_000)[0]
    est = acc[0, 0, 0] / 40_000
    want = rho * Le * 4 * rect_factor_corner(half, half, h)
    assert abs(est - want) / want < 0.01, (est, want)


def test_f32_and_f64_oracles_agree(cornell):
    f32 = oracle.render_f32(cornell.desc, 48, 48, 8, 16)[0][..., :3].astype(np.float64)
    f64 = oracle.render_f64(cornell.desc, 48, 48, 8, 16)[..., :3]
    rel = np.abs(f32 - f64).sum() / np.abs(f64).sum()
    assert rel < 2e-3, rel


def random_rays(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    o = rng.uniform(lo, hi, (n, 3))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return np.concatenate([o, d], 1).astype(np.float32)


@pytest.mark.parametrize("kind,tris,n", [(ptcore.SCENE_CORNELL, 0, 20000), (ptcore.SCENE_SIZED, 3000, 4000)])
def test_traversal_equals_brute_force(kind, tris, n):
    world = ptcore.World.generate(kind, 5, 64, 64, tris)
    rays = random_rays(n, -1.2, 1.2, 3)
    hit, t, *_ = oracle.intersect(world.desc, rays)
    bh, bt = oracle.intersect_brute(world.desc, rays)
    assert np.array_equal(hit >= 0, bh >= 0)
    m = hit >= 0
    assert np.array_equal(t[m], bt[m])
    # ids equal except where two triangles give the very same t (shared edges)
    diff = m & (hit != bh)
    assert diff.mean() < 1e-3


def test_watertight_shared_edges_and_vertices():
    world = ptcore.World.generate(ptcore.SCENE_SIZED, 8, 64, 64, 4000)
    a = world.arrays()
    v = a["geometries"]["v"].reshape(-1, 3, 3).astype(np.float64)
    rng = np.random.default_rng(2)
    sel = rng.choice(len(v), 3000, replace=False)
    # targets: vertices and edge midpoints of interior ground triangles seen
    # from above (on the outer rim a rounded direction may pass outside)
    targets = []
    for i in sel:
        tv = v[i]
        if tv[:, 1].max() > 0.35 or np.abs(tv[:, [0, 2]]).max() > 0.9:
            continue
        targets += [tv[0], (tv[0] + tv[1]) / 2]
    targets = np.array(targets)
    origin = targets + np.array([0.013, 2.0, -0.021])
    d = targets - origin
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.concatenate([origin, d], 1).astype(np.float32)
    # The reference's edge functions use fma (trace.slang:819-821), so the
    # two triangles of a shared edge do not compute exact negations of each
    # other: the test is watertight only up to rare misses.
    bh, _ = oracle.intersect_brute(world.desc, rays)
    assert (bh < 0).mean() <= 1e-3, f"{int((bh < 0).sum())} of {len(bh)} rays slipped through"
    # Its slab test (trace.slang:597-609) is not conservative either: a ray
    # aimed exactly at a box corner can round to an empty interval and skip
    # the leaf. The BVH traversal keeps that behaviour (hit ids must match
    # the reference's), so corner-aimed rays miss more often than generic ones.
    hit, *_ = oracle.intersect(world.desc, rays)
    print(f"corner-aimed rays: brute-force misses {(bh < 0).mean():.3%}, BVH misses {(hit < 0).mean():.3%}")
    assert ((hit >= 0) <= (bh >= 0)).all(), "the BVH found a hit brute force did not"
    m = hit >= 0
    _, bt = oracle.intersect_brute(world.desc, rays)
    t = oracle.intersect(world.desc, rays)[1]
    assert (t[m] >= bt[m]).all()


@pytest.mark.parametrize("mode,alpha,cutoff,expect_pass", [(2, 0.3, 0.5, 0.7), (2, 0.85, 0.5, 0.15),
                                                            (1, 0.3, 0.5, 1.0), (1, 0.7, 0.5, 0.0)])
def test_alpha_coverage_closed_form(mode, alpha, cutoff, expect_pass):
    # a black alpha-tested quad filling the view in front of a constant sky of luminance L:
    # blend lets a ray through with probability 1 - alpha (trace.slang:426-454), mask iff alpha < cutoff,
    # so the pixel's mean radiance at depth 0 is (pass probability) * L
    L = 2.0
    quad = [[-5, 0, -5], [5, 0, 5], [5, 0, -5]], [[-5, 0, -5], [-5, 0, 5], [5, 0, 5]]
    geo, attr = ptcore.triangles_from_vertices([*quad], [0, 0])
    mats = np.array([ptcore.default_material(base_color=(0, 0, 0, alpha), alpha_mode=mode, alpha_cutoff=cutoff,
                                             specular_color_and_weight=(1, 1, 1, 0))])
    sky = np.ones((4, 8, 4), np.float32)
    world = ptcore.WorldInput(geo, attr, mats, sky, skybox_luminance=L, camera_position=(0.0, 0.5, 0.0),
                              yaw=0.0, pitch=-90.0, fov=0.5, aspect=1.0).build()
    n = 40_000
    acc = oracle.render_f32(world.desc, 1, 1, 0, n)[0]
    est = acc[0, 0, :3] / n
    want = expect_pass * L
    assert np.allclose(est, est[0])
    assert abs(est[0] - want) <= 0.015 * L, (est, want)
