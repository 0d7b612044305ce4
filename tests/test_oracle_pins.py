"""CPU: pins of the oracle functions that the radiance parity tests cannot
see, each against what the mathematics fixes (no reference run is possible).

- GGX distribution (trace.slang:249-250): the projected microfacet area
  integrates to one, int D(h) cos(theta_h) dw = 1, for anisotropic alpha.
- Smith Lambda (trace.slang:347-352): by definition of the Smith masking
  function for D, the visible-normal area satisfies
  int <wo, h>+ D(h) dw = cos(theta_o) (1 + Lambda(wo)) (Heitz 2014, eq. 2);
  Lambda >= 0, so the height-correlated G2 never exceeds G1.
- The GGX lobe values of evaluate_lobes (trace.slang:229-262): a white
  conductor's directional albedo by quadrature of f cos equals the VNDF
  expectation of G2 / G1 (the sampler's own estimator), and the conductor is
  reciprocal; the F82-tint Fresnel takes its defining values (F0 at normal
  incidence, specular tint times Schlick at cos = 1/7, one at grazing); the
  dielectric lobe is the exact Fresnel factor times the same microfacet term.
- estimate_e_dielectric (trace.slang:396-411): its mean over many seeds is
  the float64 quadrature of F G2 / G1 over the visible normals, for IOR
  below and above one, isotropic and anisotropic alpha; a scripted random
  stream pins the draw order (rand1, then 2 pi rand) and the loop.
- Sky fetch (trace.slang:473-475, world_resources.rs:451-456): on a sky
  linear in texel coordinates bilinear filtering with clamp-to-edge is
  exact, so evaluate_skybox equals the linear function.
- sRGB decode table and the display channel: the IEC 61966-2-1 EOTF and the
  ACES fit + sRGB encode of display.wgsl:19-46 in float64.
"""
import ctypes as C
import math

import numpy as np
import pytest

import binding as oracle
import ptcore

L = oracle.lib()


def f3(v):
    return (C.c_float * 3)(*[float(x) for x in v])


def ggx_d(h, ax, ay):
    return L.oracle_ggx_d(f3(h), ax, ay)


def lam(v, ax, ay):
    return L.oracle_ggx_lambda(f3(v), ax, ay)


def vndf(wo, ax, ay, r, two_pi_r):
    out = (C.c_float * 3)()
    L.oracle_sample_vndf(f3(wo), ax, ay, r, two_pi_r, out)
    return np.array(out[:], np.float64)


def lobes(prm, wo, wi, e=(0.0, 0.0, 0.0)):
    f = (C.c_float * 3)()
    pdf = C.c_float()
    L.oracle_evaluate_lobes((C.c_float * 13)(*prm), f3(wo), f3(wi), f3(e), f, C.byref(pdf))
    return np.array(f[:], np.float64), pdf.value


def dir_at(theta, phi=0.0):
    return np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])


def hemisphere_quad(fn, scale, n_s=160, n_phi=192):
    """int over the upper hemisphere of fn(h) dw_h for an integrand peaked at
    the pole with angular width about `scale`: tan(theta) = scale s / (1 - s),
    Gauss-Legendre in s, uniform (periodic trapezoid) in phi."""
    xs, ws = np.polynomial.legendre.leggauss(n_s)
    s = 0.5 * (xs + 1.0)
    ws = 0.5 * ws
    total = 0.0
    for phi in (np.arange(n_phi) + 0.5) * (2 * math.pi / n_phi):
        for sk, wk in zip(s, ws):
            t = scale * sk / (1.0 - sk)
            theta = math.atan(t)
            dtheta_ds = scale / (1.0 - sk) ** 2 / (1.0 + t * t)
            total += fn(dir_at(theta, phi)) * math.sin(theta) * dtheta_ds * wk * (2 * math.pi / n_phi)
    return total


ALPHAS = [(0.05, 0.05), (0.3, 0.3), (1.0, 1.0), (0.08, 0.4), (0.6, 0.05)]


@pytest.mark.parametrize("ax,ay", ALPHAS)
def test_ggx_distribution_is_normalised(ax, ay):
    got = hemisphere_quad(lambda h: ggx_d(h, ax, ay) * h[2], min(ax, ay))
    assert abs(got - 1.0) < 2e-3, got
    assert ggx_d([0.0, 0.0, 1.0], ax, ay) == pytest.approx(1.0 / (math.pi * ax * ay), rel=1e-6)


@pytest.mark.parametrize("ax,ay", ALPHAS)
@pytest.mark.parametrize("theta,phi", [(0.3, 0.0), (1.0, 0.7), (1.35, 2.0), (1.5, 4.0)])
def test_smith_lambda_is_the_visible_area_of_d(ax, ay, theta, phi):
    wo = dir_at(theta, phi)
    area = hemisphere_quad(lambda h: max(float(np.dot(wo, h)), 0.0) * ggx_d(h, ax, ay), min(ax, ay), n_s=200, n_phi=256)
    want = wo[2] * (1.0 + lam(wo, ax, ay))
    assert area == pytest.approx(want, rel=3e-3), (area, want)


def test_lambda_is_nonnegative_so_g2_never_exceeds_g1():
    rng = np.random.default_rng(1)
    v = rng.normal(size=(4000, 3))
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    v[:, 2] = np.abs(v[:, 2]) + 1e-3
    for (ax, ay), x in zip(ALPHAS * 800, v):
        assert lam(x, ax, ay) >= 0.0
    assert lam([0.0, 0.0, 1.0], 0.7, 0.2) == 0.0


# white conductor: base (1,1,1), base weight 1, metal 1, diffuse roughness 0,
# specular colour (1,1,1), specular weight 1, IOR 1.5, roughness r, anisotropy a
def conductor(r, a, base=(1.0, 1.0, 1.0), spec=(1.0, 1.0, 1.0), base_weight=1.0):
    return [*base, base_weight, 1.0, 0.0, *spec, 1.0, 1.5, r, a]


@pytest.mark.parametrize("rough,aniso", [(0.2, 0.0), (0.5, 0.0), (0.9, 0.0), (0.4, 0.7)])
@pytest.mark.parametrize("theta", [0.2, 0.9, 1.3])
def test_conductor_albedo_quadrature_equals_vndf_expectation(rough, aniso, theta):
    ax, ay = C.c_float(), C.c_float()
    L.oracle_alpha_anisotropic(rough, aniso, C.byref(ax), C.byref(ay))
    ax, ay = ax.value, ay.value
    wo = dir_at(theta, 0.4)
    prm = conductor(rough, aniso)

    def integrand(h):  # f cos(theta_i) over w_i, written over h: dw_i = 4 <wo, h> dw_h
        c = float(np.dot(wo, h))
        if c <= 0.0:
            return 0.0
        wi = 2.0 * c * h - wo
        if wi[2] <= 1e-6:
            return 0.0
        f, _ = lobes(prm, wo, wi)
        return f[0] * wi[2] * 4.0 * c

    albedo = hemisphere_quad(integrand, min(ax, ay), n_s=160, n_phi=160)
    # the sampler's estimator: E over h ~ VNDF(wo) of G2 / G1 on w_i in the upper hemisphere
    rng = np.random.default_rng(7)
    u = rng.random((60000, 2)).astype(np.float32)
    opl = 1.0 + lam(wo, ax, ay)
    acc = []
    for r1, r2 in u:
        h = vndf(wo, ax, ay, float(r1), float(2 * math.pi * r2))
        wi = 2.0 * float(np.dot(wo, h)) * h - wo
        acc.append(opl / (opl + lam(wi, ax, ay)) if wi[2] > 0 else 0.0)
    acc = np.array(acc)
    mc, se = acc.mean(), acc.std() / math.sqrt(len(acc))
    print(f"alpha ({ax:.3g},{ay:.3g}) theta {theta}: quadrature {albedo:.5f}, VNDF estimate {mc:.5f} +- {se:.1e}")
    assert albedo <= 1.0 + 1e-3
    assert abs(albedo - mc) < max(4 * se, 0.01 * mc)


def test_conductor_is_reciprocal():
    rng = np.random.default_rng(3)
    prm = conductor(0.45, 0.5, base=(0.9, 0.6, 0.3), spec=(0.7, 0.9, 1.0), base_weight=0.8)
    for _ in range(200):
        a, b = rng.normal(size=3), rng.normal(size=3)
        a[2], b[2] = abs(a[2]) + 0.05, abs(b[2]) + 0.05
        a, b = a / np.linalg.norm(a), b / np.linalg.norm(b)
        fab, _ = lobes(prm, a, b)
        fba, _ = lobes(prm, b, a)
        assert np.allclose(fab, fba, rtol=2e-5, atol=1e-12)


def mirror_pair(mu):
    """w_o and w_i symmetric about the normal, so H = n and H . w_o = mu."""
    s = math.sqrt(1.0 - mu * mu)
    return np.array([s, 0.0, mu]), np.array([-s, 0.0, mu])


def test_f82_tint_defining_values():
    F0 = np.array([0.9, 0.6, 0.3])
    tint = np.array([0.5, 0.8, 1.0])

    def fresnel(mu, spec, f0=F0, rough=0.5):
        wo, wi = mirror_pair(mu)
        f, _ = lobes(conductor(rough, 0.0, base=tuple(f0), spec=tuple(spec)), wo, wi)
        g, _ = lobes(conductor(rough, 0.0), wo, wi)  # F = 1: the microfacet term alone
        return f / g

    schlick = lambda mu: F0 + (1.0 - F0) * (1.0 - mu) ** 5
    assert np.allclose(fresnel(1.0, tint), F0, rtol=1e-5)                 # normal incidence: F0, whatever the tint
    assert np.allclose(fresnel(1.0 / 7.0, tint), tint * schlick(1.0 / 7.0), rtol=1e-5)  # the F82 point
    assert np.allclose(fresnel(1.0 / 7.0, (1, 1, 1)), schlick(1.0 / 7.0), rtol=1e-5)
    for mu in (0.05, 0.3, 0.7):
        assert np.allclose(fresnel(mu, (1, 1, 1)), schlick(mu), rtol=1e-5)  # no tint: Schlick
    # grazing: one (F falls off linearly with slope 5 (1 - F0) + the tint term's coefficient, ~8.5 here)
    assert np.allclose(fresnel(1e-5, tint), 1.0, atol=2e-4)


@pytest.mark.parametrize("ior", [1.5, 2.2, 1.1])
def test_dielectric_lobe_is_fresnel_times_the_microfacet_term(ior):
    for mu in (0.95, 0.6, 0.25, 0.08):
        wo, wi = mirror_pair(mu)
        f, _ = lobes([0.5, 0.5, 0.5, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 1.0, ior, 0.5, 0.0], wo, wi)
        g, _ = lobes(conductor(0.5, 0.0), wo, wi)
        # textbook unpolarised Fresnel at IOR `ior` (specular weight 1 leaves it unmodulated)
        st = math.sqrt(1 - mu * mu) / ior
        ct = math.sqrt(1 - st * st)
        rs = (mu - ior * ct) / (mu + ior * ct)
        rp = (ct - ior * mu) / (ct + ior * mu)
        assert f / g == pytest.approx([0.5 * (rs * rs + rp * rp)] * 3, rel=2e-5)


# ---------------------------------------------------------------- E_dielectric
def d64(h, ax, ay):
    k = h[..., 0] ** 2 / ax ** 2 + h[..., 1] ** 2 / ay ** 2 + h[..., 2] ** 2
    return 1.0 / (math.pi * ax * ay * k * k)


def lam64(v, ax, ay):
    return 0.5 * np.sqrt(1.0 + (ax * ax * v[..., 0] ** 2 + ay * ay * v[..., 1] ** 2) / v[..., 2] ** 2) - 0.5


def fresnel64(ior, c):
    s2 = (1.0 - c * c) / (ior * ior)
    ct = np.sqrt(np.maximum(1.0 - s2, 0.0))
    with np.errstate(invalid="ignore", divide="ignore"):  # the total-reflection branch is taken there
        rs = (c - ior * ct) / (c + ior * ct)
        rp = (ct - ior * c) / (ct + ior * c)
    return np.where(s2 >= 1.0, 1.0, 0.5 * (rs * rs + rp * rp))


def e_dielectric_quadrature(wo, ax, ay, ior, n_s=400, n_phi=512):
    """int over h of F(h . wo) G2(wo, wi) / G1(wo) D_wo(h) with wi = reflect(-wo, h) above the
    surface, D_wo(h) = G1(wo) <wo, h>+ D(h) / cos(theta_o) the visible-normal density."""
    xs, ws = np.polynomial.legendre.leggauss(n_s)
    s = 0.5 * (xs + 1.0)
    ws = 0.5 * ws
    scale = min(ax, ay)
    t = scale * s / (1.0 - s)
    theta = np.arctan(t)
    dth = scale / (1.0 - s) ** 2 / (1.0 + t * t)
    phi = (np.arange(n_phi) + 0.5) * (2 * math.pi / n_phi)
    TH, PH = np.meshgrid(theta, phi, indexing="ij")
    h = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1)
    c = h @ wo
    wi = 2.0 * c[..., None] * h - wo
    ok = (c > 0) & (wi[..., 2] > 0)
    wiz = np.where(ok, wi[..., 2], 1.0)
    wi = np.where(ok[..., None], wi, np.array([0.0, 0.0, 1.0]))
    opl = 1.0 + lam64(wo, ax, ay)
    g2_over_g1 = opl / (opl + lam64(wi, ax, ay))
    dvis = np.maximum(c, 0.0) * d64(h, ax, ay) / (wo[2] * opl)
    val = np.where(ok, fresnel64(ior, np.clip(c, 0, 1)) * g2_over_g1 * dvis, 0.0)
    del wiz
    weight = (np.sin(TH) * dth[:, None] * ws[:, None]) * (2 * math.pi / n_phi)
    return float((val * weight).sum())


@pytest.mark.parametrize("rough,aniso,ior,theta", [
    (0.3, 0.0, 1.5, 0.4), (0.3, 0.0, 1.5, 1.3), (0.7, 0.0, 1.5, 0.9), (0.5, 0.6, 1.5, 1.0),
    (0.5, 0.0, 0.7, 0.5), (0.5, 0.0, 0.7, 1.2), (0.85, 0.3, 2.4, 0.8)])
def test_e_dielectric_mean_is_the_albedo_it_estimates(rough, aniso, ior, theta):
    ax, ay = C.c_float(), C.c_float()
    L.oracle_alpha_anisotropic(rough, aniso, C.byref(ax), C.byref(ay))
    ax, ay = ax.value, ay.value
    wo = dir_at(theta, 0.3)
    spec = (1.0, 0.5, 0.25)
    n = 8192
    est = np.empty(n)
    out = (C.c_float * 3)()
    for k in range(n):
        st = C.c_uint32(L.oracle_seed(k % 97, k // 97, 1 + k % 13))
        L.oracle_estimate_e_dielectric(f3(wo), ax, ay, ior, f3(spec), C.byref(st), out)
        assert out[1] == pytest.approx(0.5 * out[0], rel=1e-6) and out[2] == pytest.approx(0.25 * out[0], rel=1e-6)
        est[k] = out[0]
    mean, se = est.mean(), est.std(ddof=1) / math.sqrt(n)
    quad = e_dielectric_quadrature(wo, ax, ay, ior)
    print(f"alpha ({ax:.3g},{ay:.3g}) ior {ior} theta {theta}: mean {mean:.5f} +- {se:.1e}, quadrature {quad:.5f}")
    assert abs(mean - quad) < 4 * se + 2e-4


def test_e_dielectric_draw_order():
    """With the oracle's own stream, the estimate equals the loop rebuilt from
    the pinned pieces with rand1 drawn first and 2 pi rand second; the
    swapped order gives other values."""
    wo = dir_at(0.8, 0.5)
    ax, ay, ior = 0.3, 0.12, 1.45
    opl = 1.0 + lam(wo, ax, ay)
    swapped_differs = 0
    for seed in range(50):
        s0 = L.oracle_seed(seed, 3, 1)
        st = C.c_uint32(s0)
        out = (C.c_float * 3)()
        L.oracle_estimate_e_dielectric(f3(wo), ax, ay, ior, f3((1, 1, 1)), C.byref(st), out)
        st2 = C.c_uint32(s0)
        draws = [L.oracle_rand(C.byref(st2)) for _ in range(32)]
        assert st.value == st2.value  # 32 draws, no more
        res = []
        for order in ((0, 1), (1, 0)):
            total = np.float32(0)
            for i in range(16):
                r1, r2 = draws[2 * i + order[0]], draws[2 * i + order[1]]
                h = vndf(wo, ax, ay, r1, float(np.float32(2.0) * np.float32(3.14159265359) * np.float32(r2)))
                wi = 2.0 * float(np.dot(wo, h)) * h - wo
                if wi[2] <= 0:
                    continue
                fr = L.oracle_fresnel_exact(ior, float(np.dot(h, wo)))
                total += np.float32(fr * opl / (opl + lam(wi, ax, ay)))
            res.append(float(total) / 16.0)
        assert out[0] == pytest.approx(res[0], rel=2e-5, abs=1e-7)
        swapped_differs += abs(res[1] - out[0]) > 1e-3 * max(out[0], 1e-3)
    assert swapped_differs >= 40


# ---------------------------------------------------------------- sky, sRGB, display
def sky_desc(rgba, lum):
    h, w, _ = rgba.shape
    d = ptcore.SceneDesc()
    d.skybox_rgba = rgba.ctypes.data_as(C.c_void_p)
    d.skybox.width, d.skybox.height = w, h
    d.skybox.average_luminance = lum
    return d


@pytest.mark.parametrize("w,h", [(64, 32), (1024, 512), (7, 5)])
def test_sky_fetch_is_exact_on_a_linear_sky(w, h):
    X, Y = np.meshgrid(np.arange(w, dtype=np.float64), np.arange(h, dtype=np.float64))
    lin = lambda x, y: np.stack([1.0 + 0.01 * x + 0.02 * y, 2.0 + 0.03 * x, 3.0 + 0.05 * y], -1)
    rgba = np.concatenate([lin(X, Y), np.ones((h, w, 1))], -1).astype(np.float32)
    lum = 1.75
    d = sky_desc(rgba, lum)
    rng = np.random.default_rng(w)
    dirs = rng.normal(size=(3000, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    out = (C.c_float * 3)()
    worst = 0.0
    for v in dirs:
        L.oracle_evaluate_skybox(C.byref(d), f3(v), out)
        u = (math.atan2(v[2], v[0]) / (2 * math.pi) + 0.5) % 1.0
        vv = math.atan2(math.hypot(v[0], v[2]), v[1]) / math.pi
        x = min(max(u * w - 0.5, 0.0), w - 1.0)
        y = min(max(vv * h - 0.5, 0.0), h - 1.0)
        want = lin(np.array(x), np.array(y)) * lum
        worst = max(worst, float(np.max(np.abs(np.array(out[:]) - want) / want)))
    print(f"sky {w}x{h}: worst relative error {worst:.2e}")
    assert worst < 2e-5


def test_srgb_decode_table_is_the_iec_eotf():
    for i in range(256):
        c = i / 255.0
        want = c / 12.92 if c <= 0.04045 else ((c + 0.055) / 1.055) ** 2.4
        assert L.oracle_srgb_decode(i) == np.float32(want)


def test_display_channel_is_aces_then_srgb():
    xs = np.concatenate([np.linspace(-1, 12, 6001), np.geomspace(1e-6, 1e6, 3001)]).astype(np.float32)
    for c in xs:
        c64 = float(c)
        m = min(max((c64 * (2.51 * c64 + 0.03)) / (c64 * (2.43 * c64 + 0.59) + 0.14), 0.0), 1.0)
        want = 1.055 * m ** (1 / 2.4) - 0.055 if m > 0.0031308 else 12.92 * m
        assert L.oracle_display_channel(c64) == pytest.approx(want, abs=3e-7, rel=3e-6)


# ---------------------------------------------------------------- shading frame, textures, camera
def test_alpha_anisotropic_keeps_the_rms_roughness():
    """OpenPBR eq. 21: alpha_b = alpha_t (1 - anisotropy), and the RMS of the
    two alphas is roughness^2 (until the 1e-4 floor)."""
    for r in (0.05, 0.3, 0.7, 1.0):
        for a in (0.0, 0.25, 0.5, 0.9):
            ax, ay = C.c_float(), C.c_float()
            L.oracle_alpha_anisotropic(r, a, C.byref(ax), C.byref(ay))
            assert math.sqrt((ax.value ** 2 + ay.value ** 2) / 2) == pytest.approx(r * r, rel=1e-5)
            assert ay.value == pytest.approx(max(ax.value * (1 - a), 1e-4), rel=1e-5)


def test_orthonormal_tangent():
    """Gram-Schmidt of the tangent against the unit normal; for a tangent along
    the normal (or zero), the Duff et al. 2017 basis vector, which is unit and
    perpendicular to the normal too."""
    rng = np.random.default_rng(11)
    out = (C.c_float * 3)()
    for k in range(2000):
        n = rng.normal(size=3)
        n /= np.linalg.norm(n)
        t = rng.normal(size=3) if k % 4 else (n * rng.uniform(-2, 2) if k % 8 else np.zeros(3))
        L.oracle_orthonormal_tangent(f3(t), f3(n), out)
        v = np.array(out[:], np.float64)
        assert abs(np.linalg.norm(v) - 1) < 1e-5 and abs(np.dot(v, n)) < 1e-5
        p = t - np.dot(n, t) * n
        if np.dot(p, p) > 1e-6:
            assert np.allclose(v, p / np.linalg.norm(p), atol=1e-5)  # the projection, normalised


@pytest.mark.parametrize("srgb", [0, 1])
def test_atlas_fetch_is_nearest_with_clamp(srgb):
    size, layers = 16, 3
    rng = np.random.default_rng(5)
    atl = rng.integers(0, 256, (layers, size, size, 4), dtype=np.uint8)
    ptr = atl.ctypes.data_as(C.c_void_p)
    out = (C.c_float * 4)()
    for _ in range(3000):
        layer = int(rng.integers(0, layers))
        u, v = rng.uniform(-0.3, 1.3, 2).astype(np.float32)
        L.oracle_sample_atlas(ptr, size, layer, float(u), float(v), srgb, out)
        ix = min(max(int(math.floor(float(u) * size)), 0), size - 1)
        iy = min(max(int(math.floor(float(v) * size)), 0), size - 1)
        px = atl[layer, iy, ix]
        want = [L.oracle_srgb_decode(int(c)) if srgb else np.float32(c) / np.float32(255) for c in px[:3]]
        assert np.array_equal(np.float32(out[:3]), np.float32(want))
        assert np.float32(out[3]) == np.float32(px[3]) / np.float32(255)


def test_primary_rays_follow_the_pinhole_camera():
    """Camera::gpu_camera (world.rs:2159-2173) and the primary ray
    (trace.slang:738-748): the image plane sits at unit distance along the
    forward axis R_y(yaw) R_x(pitch) (-Z), spans 2 tan(fov / 2) vertically
    (image rows going down) and aspect times that horizontally; a pixel's ray
    goes through (pixel + jitter) / size with the jitter drawn x first."""
    yaw, pitch, fov, aspect, W, H = 30.0, 10.0, 70.0, 1.6, 160, 100
    geo, attr = ptcore.triangles_from_vertices([[[0, 0, -5], [1, 0, -5], [0, 1, -5]]], [0])
    world = ptcore.WorldInput(geo, attr, np.array([ptcore.default_material()]), np.ones((4, 8, 4), np.float32),
                              camera_position=(1.0, 2.0, 3.0), yaw=yaw, pitch=pitch, fov=fov, aspect=aspect).build()
    cam = world.desc.camera
    y, p, f = math.radians(yaw), math.radians(pitch), math.radians(fov)
    fwd = np.array([-math.sin(y) * math.cos(p), math.sin(p), -math.cos(y) * math.cos(p)])
    right = np.array([math.cos(y), 0.0, -math.sin(y)])
    down = -np.cross(right, fwd)
    hv = np.array(cam.horizontal[:3], np.float64)
    vv = np.array(cam.vertical[:3], np.float64)
    llc = np.array(cam.lower_left_corner_minus_position[:3], np.float64)
    assert np.allclose(cam.position[:3], [1, 2, 3])
    assert np.allclose(vv, down * 2 * math.tan(f / 2), atol=1e-6)
    assert np.allclose(hv, right * aspect * 2 * math.tan(f / 2), atol=1e-6)
    assert np.allclose(llc + 0.5 * hv + 0.5 * vv, fwd, atol=1e-6)
    out = (C.c_float * 3)()
    for px, py in [(0, 0), (W - 1, H - 1), (80, 50), (17, 93)]:
        st = C.c_uint32(L.oracle_seed(px, py, 3))
        st2 = C.c_uint32(st.value)
        jx, jy = L.oracle_rand(C.byref(st2)), L.oracle_rand(C.byref(st2))
        L.oracle_primary_dir(C.byref(world.desc), px, py, W, H, C.byref(st), out)
        assert st.value == st2.value
        d = llc + (px + jx) / W * hv + (py + jy) / H * vv
        assert np.allclose(out[:], d / np.linalg.norm(d), atol=2e-6)


# ---------------------------------------------------------------- the whole path
def closed_box(rho, le):
    """A closed 2x2x2 box of 12 triangles, every one Lambertian with albedo rho
    and emitting le (specular weight 0, diffuse roughness 0), camera at the centre."""
    c = [[x, y, z] for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]
    faces = [(0, 1, 3, 2), (4, 6, 7, 5), (0, 4, 5, 1), (2, 3, 7, 6), (0, 2, 6, 4), (1, 5, 7, 3)]
    tris = []
    for a, b, cc, d in faces:
        tris += [[c[a], c[b], c[cc]], [c[a], c[cc], c[d]]]
    geo, attr = ptcore.triangles_from_vertices(tris, [0] * 12)
    mat = ptcore.default_material(base_color=(rho, rho, rho, 1), base_diffuse_roughness=0.0,
                                  specular_color_and_weight=(1, 1, 1, 0), emission_color_and_luminance=(1, 1, 1, le))
    return ptcore.WorldInput(geo, attr, np.array([mat]), np.zeros((4, 8, 4), np.float32), skybox_luminance=0.0,
                             camera_position=(0.0, 0.0, 0.0), yaw=20.0, pitch=10.0, fov=100.0, aspect=1.0).build()


@pytest.mark.parametrize("rho,depth", [(0.5, 2), (0.5, 6), (0.8, 8), (0.3, 0)])
@pytest.mark.parametrize("precision", ["f32", "f64"])
def test_path_furnace_closed_form(rho, depth, precision):
    """Inside a closed box whose walls all emit le and reflect with albedo
    rho, a path of max_bounces D carries le * sum_{k=0}^{D+1} rho^k in
    expectation: the emission seen at surfaces 0..D (direct at the first,
    then split between BSDF sampling and light sampling by the MIS weights)
    plus the light sampled from the last surface. This pins the assembly of
    the bounce loop (throughput, both MIS weights, NEE, Russian roulette past
    bounce 3) by the closed form."""
    le, n, spp = 2.0, 24, 256
    world = closed_box(rho, le)
    if precision == "f32":
        acc = oracle.render_f32(world.desc, n, n, depth, spp)[0][..., :3].astype(np.float64)
    else:
        acc = oracle.render_f64(world.desc, n, n, depth, spp)[..., :3]
    px = acc[..., 0].reshape(-1) / spp
    want = le * sum(rho ** k for k in range(depth + 2))
    mean, se = px.mean(), px.std(ddof=1) / math.sqrt(px.size)
    print(f"rho {rho} depth {depth} {precision}: mean {mean:.5f} +- {se:.1e}, closed form {want:.5f}")
    assert abs(mean - want) < 4 * se + 1e-4 * want
    assert np.allclose(acc[..., 0], acc[..., 1]) and np.allclose(acc[..., 0], acc[..., 2])
