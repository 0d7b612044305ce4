// trace_oracle.cpp - CPU oracle: scalar restatement of the reference's trace
// kernel, src/render/trace.slang. TEST INFRASTRUCTURE ONLY: it is the checker
// that the HIP kernels are compared against and the CPU baseline that the
// benchmark times beside them; no product code links it.
//
// Templated on the arithmetic type T. T = float is the FP32 parity oracle:
// it evaluates every expression in the order the shader writes it, with fma
// exactly where the shader writes fma, and takes its transcendentals from
// pt_math.h (the same deterministic libm the kernels use). T = double is the
// FP64 convergence oracle (std:: transcendentals, same estimator and constants).
//
// Build: -O3 -ffp-contract=off (no contraction, IEEE division and sqrt).
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../path-tracer_amd/csrc/pt_math.h"

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

#define F(x) T(x##f)

// ------------------------------------------------------------------ math policy
template <class T>
struct Mth;
template <>
struct Mth<float> {
    static float sqrt(float x) { return sqrtf(x); }
    static float fma(float a, float b, float c) { return fmaf(a, b, c); }
    static float sin(float x) { return ptm::sin_(x); }
    static float cos(float x) { return ptm::cos_(x); }
    static float acos(float x) { return ptm::acos_(x); }
    static float atan2(float y, float x) { return ptm::atan2_(y, x); }
    static float pow(float x, float y) { return ptm::pow_(x, y); }
    static float floor(float x) { return floorf(x); }
    static float min(float a, float b) { return fminf(a, b); }
    static float max(float a, float b) { return fmaxf(a, b); }
};
template <>
struct Mth<double> {
    static double sqrt(double x) { return std::sqrt(x); }
    static double fma(double a, double b, double c) { return std::fma(a, b, c); }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double acos(double x) { return std::acos(x); }
    static double atan2(double y, double x) { return std::atan2(y, x); }
    static double pow(double x, double y) { return std::pow(x, y); }
    static double floor(double x) { return std::floor(x); }
    static double min(double a, double b) { return std::fmin(a, b); }
    static double max(double a, double b) { return std::fmax(a, b); }
};

// ------------------------------------------------------------------ vectors
template <class T>
struct V3 {
    T x, y, z;
    T operator[](int k) const { return k == 0 ? x : (k == 1 ? y : z); }
};
template <class T> V3<T> operator+(V3<T> a, V3<T> b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
template <class T> V3<T> operator-(V3<T> a, V3<T> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> V3<T> operator-(V3<T> a) { return {-a.x, -a.y, -a.z}; }
template <class T> V3<T> operator*(V3<T> a, V3<T> b) { return {a.x * b.x, a.y * b.y, a.z * b.z}; }
template <class T> V3<T> operator*(V3<T> a, T s) { return {a.x * s, a.y * s, a.z * s}; }
template <class T> V3<T> operator*(T s, V3<T> a) { return {s * a.x, s * a.y, s * a.z}; }
template <class T> V3<T> operator/(V3<T> a, T s) { return {a.x / s, a.y / s, a.z / s}; }
template <class T> V3<T> operator/(V3<T> a, V3<T> b) { return {a.x / b.x, a.y / b.y, a.z / b.z}; }
template <class T> V3<T> sub_s(V3<T> a, T s) { return {a.x - s, a.y - s, a.z - s}; }
template <class T> V3<T> add_s(V3<T> a, T s) { return {a.x + s, a.y + s, a.z + s}; }
template <class T> V3<T> s_sub(T s, V3<T> a) { return {s - a.x, s - a.y, s - a.z}; }
template <class T> T dot(V3<T> a, V3<T> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class T> V3<T> cross(V3<T> a, V3<T> b) {
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class T> T length(V3<T> a) { return Mth<T>::sqrt(dot(a, a)); }
template <class T> V3<T> normalize(V3<T> a) { return a * (T(1) / Mth<T>::sqrt(dot(a, a))); }
template <class T> T rsqrt(T x) { return T(1) / Mth<T>::sqrt(x); }
template <class T> T saturate(T x) { return Mth<T>::min(Mth<T>::max(x, T(0)), T(1)); }
template <class T> T clampv(T x, T lo, T hi) { return Mth<T>::min(Mth<T>::max(x, lo), hi); }
template <class T> V3<T> saturate3(V3<T> a) { return {saturate(a.x), saturate(a.y), saturate(a.z)}; }
template <class T> T lerp(T a, T b, T t) { return a * (T(1) - t) + b * t; }
template <class T> V3<T> lerp3(V3<T> a, V3<T> b, T t) { return a * (T(1) - t) + b * t; }
template <class T> V3<T> fma3(V3<T> a, V3<T> b, V3<T> c) {
    return {Mth<T>::fma(a.x, b.x, c.x), Mth<T>::fma(a.y, b.y, c.y), Mth<T>::fma(a.z, b.z, c.z)};
}
template <class T> V3<T> reflect(V3<T> i, V3<T> n) { return i - n * (T(2) * dot(n, i)); }
template <class T> T frac(T x) { return x - Mth<T>::floor(x); }
template <class T> bool any_pos(V3<T> a) { return a.x > T(0) || a.y > T(0) || a.z > T(0); }
template <class T> T fifth_power(T x) { return x * x * x * x * x; }  // trace.slang:220-222

// ------------------------------------------------------------------ RNG (trace.slang:195-202)
struct Rng {
    uint32_t state;
    float next() {
        state = state * 747796405u + 2891336453u;
        uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
        uint32_t bits = 0x3f800000u | (((word >> 22u) ^ word) >> 9u);
        float f;
        std::memcpy(&f, &bits, 4);
        return f - 1.0f;
    }
};
uint32_t seed_of(uint32_t px, uint32_t py, uint32_t sample_count) {  // trace.slang:736-737
    uint32_t s = (px * 0x9E3779B9u) ^ (py * 0x85EBCA6Bu) ^ (0u * 0xC2B2AE35u) ^
                 (sample_count * 0x27D4EB2Du) ^ 0xA511E9B3u;
    return s | 1u;
}

// ------------------------------------------------------------------ scene access
struct Counters {
    uint64_t closest = 0, shadow = 0, pairs = 0, tris = 0;
};

float srgb_lut[256];
struct LutInit {
    LutInit() {
        for (int i = 0; i < 256; ++i) {
            double c = i / 255.0;
            double l = c <= 0.04045 ? c / 12.92 : std::pow((c + 0.055) / 1.055, 2.4);
            srgb_lut[i] = (float)l;
        }
    }
} lut_init;

template <class T>
V3<T> vtx(const PtTriangleGeometry& g, int k) {
    return {T(g.v[3 * k]), T(g.v[3 * k + 1]), T(g.v[3 * k + 2])};
}

// Atlas fetch: nearest filter, clamp-to-edge, LOD 0 (world_resources.rs:447-450).
template <class T>
void sample_atlas(const uint8_t* atlases, uint32_t size, int layer, T u, T v, bool srgb, T out[4]) {
    T fx = Mth<T>::floor(u * T(size));
    T fy = Mth<T>::floor(v * T(size));
    int64_t ix = fx < T(0) ? 0 : (fx > T(size - 1) ? (int64_t)size - 1 : (int64_t)fx);
    int64_t iy = fy < T(0) ? 0 : (fy > T(size - 1) ? (int64_t)size - 1 : (int64_t)fy);
    const uint8_t* p = atlases + (((uint64_t)layer * size + (uint64_t)iy) * size + (uint64_t)ix) * 4;
    for (int c = 0; c < 3; ++c) out[c] = srgb ? T(srgb_lut[p[c]]) : T((float)p[c] / 255.0f);
    out[3] = T((float)p[3] / 255.0f);
}

template <class T>
struct Kernel {
    using M = Mth<T>;
    const pt_scene_desc& s;
    Counters& cnt;
    explicit Kernel(const pt_scene_desc& sc, Counters& c) : s(sc), cnt(c) {}

    // Gram-Schmidt with Duff et al. fallback (trace.slang:207-218)
    V3<T> orthonormal_tangent(V3<T> tangent, V3<T> normal) {
        V3<T> projected = tangent - dot(normal, tangent) * normal;
        T length_squared = dot(projected, projected);
        if (length_squared > F(1e-12)) return projected * rsqrt(length_squared);
        T sg = normal.z >= T(0) ? T(1) : T(-1);
        T a = T(-1) / (sg + normal.z);
        return {M::fma(sg * normal.x * normal.x, a, T(1)), sg * normal.x * normal.y * a, -sg * normal.x};
    }

    // trace.slang:330-341
    T fresnel_exact(T ior, T cos_i) {
        T sin_t_div_ior_squared = M::fma(-cos_i, cos_i, T(1)) / (ior * ior);
        if (sin_t_div_ior_squared >= T(1)) return T(1);
        T cos_t = M::sqrt(T(1) - sin_t_div_ior_squared);
        T Rs = M::fma(-ior, cos_t, cos_i) / M::fma(ior, cos_t, cos_i);
        T Rp = M::fma(-ior, cos_i, cos_t) / M::fma(ior, cos_i, cos_t);
        return F(0.5) * (Rs * Rs + Rp * Rp);
    }
    // trace.slang:347-352
    T ggx_lambda(V3<T> v, T ax, T ay) {
        return M::fma(F(0.5), M::sqrt(T(1) + (ax * ax * v.x * v.x + ay * ay * v.y * v.y) / (v.z * v.z)), F(-0.5));
    }
    // trace.slang:356-364
    V3<T> sample_vndf(V3<T> wo, T ax, T ay, T r, T two_pi_r) {
        V3<T> wo_warped = normalize(V3<T>{wo.x * ax, wo.y * ay, wo.z});
        T z = M::fma(T(1) - r, T(1) + wo_warped.z, -wo_warped.z);
        T sin_theta = M::sqrt(saturate(M::fma(-z, z, T(1))));
        T x = sin_theta * M::cos(two_pi_r);
        T y = sin_theta * M::sin(two_pi_r);
        V3<T> h = V3<T>{x, y, z} + wo_warped;
        return normalize(V3<T>{h.x * ax, h.y * ay, h.z});
    }
    // trace.slang:369-375
    void alpha_anisotropic(T roughness, T anisotropy, T& ax, T& ay) {
        T isotropy = T(1) - anisotropy;
        T alpha_t = roughness * roughness * F(1.41421356237) * rsqrt(M::fma(isotropy, isotropy, T(1)));
        ax = M::max(alpha_t, F(1e-4));
        ay = M::max(alpha_t * isotropy, F(1e-4));
    }
    // trace.slang:379-391
    T specular_ior_modulated(T ior, T weight) {
        T num = ior - T(1);
        T den = ior + T(1);
        T f0 = num * num / (den * den);
        T sgn = (ior - T(1)) > T(0) ? T(1) : ((ior - T(1)) < T(0) ? T(-1) : T(0));
        T f0_sqrt = sgn * M::sqrt(M::min(weight * f0, T(ptm::bits_to_float(0x3f800000u - 1u))));
        return (T(1) + f0_sqrt) / (T(1) - f0_sqrt);
    }
    // trace.slang:318-323
    T e_fon_exact(T mu, T r) {
        T AF = T(1) / M::fma(F(0.28779340921), r, T(1));
        T Si = M::sqrt(M::fma(-mu, mu, T(1)));
        T G = Si * (M::acos(mu) - Si * mu) +
              T(ptm::bits_to_float(0x3f2aaaabu)) * ((Si / mu) * (T(1) - (Si * Si * Si)) - Si);
        return AF + r * AF * F(0.31830988618) * G;
    }
    // trace.slang:416-424
    T power_heuristic(T pdf, T other) {
        T scale = M::max(pdf, other);
        if (scale <= T(0)) return T(0);
        T p = pdf / scale;
        T q = other / scale;
        return p * p / (p * p + q * q);
    }

    struct Shading {
        V3<T> base_color;
        T base_weight, base_metalness, base_diffuse_roughness;
        V3<T> specular_color;
        T specular_weight, sim, ax, ay, one_plus_lambda_wo;
        V3<T> F0_conductor, E_dielectric;
        T ltc[4];
        V3<T> bx, by, bz;  // rows of the LTC orthonormal basis
        T P_u, P_c, p_ggx, p_diffuse;
    };

    // GGX distribution of evaluate_lobes, trace.slang:249-250
    T ggx_d(V3<T> H, T ax, T ay) {
        T D_helper = (H.x * H.x) / (ax * ax) + (H.y * H.y) / (ay * ay) + H.z * H.z;
        return F(0.31830988618) / (ax * ay * D_helper * D_helper);
    }
    // trace.slang:229-315
    void evaluate_lobes(const Shading& sh, V3<T> wo, V3<T> wi, V3<T>& f_total, T& pdf_total) {
        V3<T> H = normalize(wo + wi);
        T HdotW = dot(H, wo);
        T G1 = T(1) / sh.one_plus_lambda_wo;
        T G2 = T(1) / (sh.one_plus_lambda_wo + ggx_lambda(wi, sh.ax, sh.ay));
        T D = ggx_d(H, sh.ax, sh.ay);
        T ts = D * G2 / (F(4.0) * wi.z * wo.z);
        T omh = T(1) - HdotW;
        T k = F(0.00105024005) * HdotW * omh;
        V3<T> fa = s_sub(T(1), sh.specular_color) * k;
        V3<T> fb = add_s(sh.F0_conductor * F(9031.0), F(7776.0));
        V3<T> fc = sub_s(sh.F0_conductor, T(1));
        V3<T> F82 = sh.F0_conductor - fifth_power(omh) * fma3(fa, fb, fc);
        V3<T> f_conductor = saturate3(sh.specular_weight * F82) * ts;
        V3<T> f_dielectric = sh.specular_color * fresnel_exact(sh.sim, HdotW) * ts;
        V3<T> rho = sh.base_color;
        T r = sh.base_diffuse_roughness;
        T s_f = wi.x * wo.x + wi.y * wo.y;
        T sovertF = s_f > T(0) ? s_f / M::max(wo.z, wi.z) : s_f;
        T AF = T(1) / M::fma(F(0.28779340921), r, T(1));
        V3<T> f_ss = rho * F(0.31830988618) * AF * M::fma(r, sovertF, T(1));
        T EFo = e_fon_exact(wo.z, r);
        T EFi = e_fon_exact(wi.z, r);
        T avgEF = AF * M::fma(F(0.07248821245), r, T(1));
        V3<T> rho_ms = rho * rho * avgEF / s_sub(T(1), rho * (T(1) - avgEF));
        V3<T> f_ms = (rho_ms * F(0.31830988618)) * M::max(F(1.0e-7), T(1) - EFo) *
                     M::max(F(1.0e-7), T(1) - EFi) / M::max(F(1.0e-7), T(1) - avgEF);
        V3<T> f_diffuse = sh.base_weight * (f_ss + f_ms);
        V3<T> f_glossy_diffuse = f_dielectric + s_sub(T(1), sh.E_dielectric) * f_diffuse;
        f_total = lerp3(f_glossy_diffuse, f_conductor, sh.base_metalness);
        T pdf_ggx = D * G1 / (F(4.0) * wo.z);
        V3<T> wl = {dot(sh.bx, wi), dot(sh.by, wi), dot(sh.bz, wi)};
        T a = sh.ltc[0], b = sh.ltc[1], c = sh.ltc[2], d = sh.ltc[3];
        T detM = c * M::fma(-b, d, a);
        V3<T> wh = {c * (wl.x - b * wl.z), (a - b * d) * wl.y, -c * (d * wl.x - a * wl.z)};
        T lenSqr = dot(wh, wh);
        T s_pdf = M::fma(F(0.5), rsqrt(M::fma(d, d, T(1))), F(0.5));
        T pdf_c = (detM * detM) / (lenSqr * lenSqr) * M::max(wh.z, T(0)) / (F(3.14159265359) * s_pdf);
        T pdf_eon = sh.P_u * F(0.15915494309) + sh.P_c * pdf_c;
        pdf_total = sh.p_ggx * pdf_ggx + sh.p_diffuse * pdf_eon;
    }

    // trace.slang:396-411
    V3<T> estimate_e_dielectric(V3<T> wo, T opl, T ax, T ay, T sim, V3<T> spec_color, Rng& rng) {
        T sum_W = T(0);
        for (uint32_t i = 0; i < 16u; ++i) {
            T rand1 = T(rng.next());
            T two_pi_rand = F(2.0) * F(3.14159265359) * T(rng.next());
            V3<T> H = sample_vndf(wo, ax, ay, rand1, two_pi_rand);
            V3<T> wi = reflect(-wo, H);
            if (wi.z <= T(0)) continue;
            T Fr = fresnel_exact(sim, dot(H, wo));
            sum_W += Fr * opl / (opl + ggx_lambda(wi, ax, ay));
        }
        return spec_color * sum_W / F(16.0);
    }

    // trace.slang:426-454
    T coverage(const PtMaterial& mat, T u, T v) {
        if (mat.alpha_mode == PT_ALPHA_MODE_OPAQUE) return T(1);
        T alpha = T(mat.base_color[3]);
        if (mat.base_color_tex_layer >= 0) {
            T tex[4];
            sample_atlas<T>(s.srgb_atlases, s.atlas_size, mat.base_color_tex_layer,
                            M::fma(u, T(mat.base_color_uv[2]), T(mat.base_color_uv[0])),
                            M::fma(v, T(mat.base_color_uv[3]), T(mat.base_color_uv[1])), true, tex);
            alpha *= tex[3];
        }
        if (mat.alpha_mode == PT_ALPHA_MODE_MASK) return alpha >= T(mat.alpha_cutoff) ? T(1) : T(0);
        return alpha;
    }
    void bary_uv(const PtTriangleAttribute& at, T u, T v, T w, T& ou, T& ov) {
        ou = frac(u * T(at.uv[0]) + v * T(at.uv[2]) + w * T(at.uv[4]));
        ov = frac(u * T(at.uv[1]) + v * T(at.uv[3]) + w * T(at.uv[5]));
    }
    bool passes_through(uint32_t tri, T v, T w, Rng& rng) {
        const PtTriangleAttribute& at = s.attributes[tri];
        const PtMaterial& mat = s.materials[at.material];
        if (mat.alpha_mode == PT_ALPHA_MODE_OPAQUE) return false;
        T uu, vv;
        bary_uv(at, T(1) - v - w, v, w, uu, vv);
        return coverage(mat, uu, vv) <= T(rng.next());
    }
    V3<T> evaluate_emission(const PtMaterial& mat, T u, T v) {
        V3<T> color = {T(mat.emission_color_and_luminance[0]), T(mat.emission_color_and_luminance[1]),
                       T(mat.emission_color_and_luminance[2])};
        if (mat.emission_tex_layer >= 0) {
            T tex[4];
            sample_atlas<T>(s.srgb_atlases, s.atlas_size, mat.emission_tex_layer,
                            M::fma(u, T(mat.emission_uv[2]), T(mat.emission_uv[0])),
                            M::fma(v, T(mat.emission_uv[3]), T(mat.emission_uv[1])), true, tex);
            color = color * V3<T>{tex[0], tex[1], tex[2]};
        }
        return T(mat.emission_weight) * T(mat.emission_color_and_luminance[3]) * color;
    }
    // trace.slang:467-472
    void skybox_uv(V3<T> d, T& u, T& v) {
        u = frac(M::fma(M::atan2(d.z, d.x), F(0.15915494309189535), F(0.5)));
        v = M::atan2(M::sqrt(d.x * d.x + d.z * d.z), d.y) * F(0.31830988618);
    }
    // bilinear, clamp-to-edge sky fetch (world_resources.rs:451-456)
    V3<T> sky_bilinear(T u, T v) {
        const uint32_t W = s.skybox.width, Hh = s.skybox.height;
        T x = u * T(W) - F(0.5);
        T y = v * T(Hh) - F(0.5);
        T x0 = M::floor(x), y0 = M::floor(y);
        T fx = x - x0, fy = y - y0;
        auto cl = [](int64_t i, uint32_t n) -> uint32_t {
            return (uint32_t)(i < 0 ? 0 : (i >= (int64_t)n ? (int64_t)n - 1 : i));
        };
        int64_t ix = (int64_t)x0, iy = (int64_t)y0;
        uint32_t xa = cl(ix, W), xb = cl(ix + 1, W), ya = cl(iy, Hh), yb = cl(iy + 1, Hh);
        auto px = [&](uint32_t xx, uint32_t yy) {
            const float* p = s.skybox_rgba + ((uint64_t)yy * W + xx) * 4;
            return V3<T>{T(p[0]), T(p[1]), T(p[2])};
        };
        V3<T> top = lerp3(px(xa, ya), px(xb, ya), fx);
        V3<T> bot = lerp3(px(xa, yb), px(xb, yb), fx);
        return lerp3(top, bot, fy);
    }
    V3<T> evaluate_skybox(V3<T> d) {
        T u, v;
        skybox_uv(d, u, v);
        return sky_bilinear(u, v) * T(s.skybox.average_luminance);
    }
    // trace.slang:476-481
    T skybox_texel_solid_angle(uint32_t row) {
        T step = F(3.14159265359) / T(s.skybox.height);
        uint32_t pole_row = std::min(row, s.skybox.height - 1u - row);
        return (F(2.0) * F(3.14159265359) / T(s.skybox.width)) * F(2.0) * M::sin((T(pole_row) + F(0.5)) * step) *
               M::sin(F(0.5) * step);
    }
    // trace.slang:482-489
    T skybox_light_pdf(V3<T> d, T category_probability) {
        if (category_probability <= T(0)) return T(0);
        T u, v;
        skybox_uv(d, u, v);
        uint32_t px = (uint32_t)(u * T(s.skybox.width));
        uint32_t py = (uint32_t)(v * T(s.skybox.height));
        px = std::min(px, s.skybox.width - 1u);
        py = std::min(py, s.skybox.height - 1u);
        return category_probability * T(s.skybox_alias[py * s.skybox.width + px].selection_probability) /
               skybox_texel_solid_angle(py);
    }
    // trace.slang:490-507
    T triangle_selection_probability(uint32_t tri) {
        uint32_t first = 0, end = s.scene_uniforms.emissive_triangles_count;
        while (first < end) {
            uint32_t middle = first + (end - first) / 2u;
            const PtSceneAliasEntry& e = s.scene_alias[middle];
            if (e.triangle_index < tri) first = middle + 1u;
            else if (e.triangle_index > tri) end = middle;
            else return T(e.selection_probability);
        }
        return T(0);
    }
    // trace.slang:508-523
    T triangle_light_pdf(uint32_t tri, V3<T> dir, T d2, T probability) {
        if (probability <= T(0)) return T(0);
        const PtTriangleGeometry& g = s.geometries[tri];
        V3<T> v0 = vtx<T>(g, 0), v1 = vtx<T>(g, 1), v2 = vtx<T>(g, 2);
        T projected_area = F(0.5) * std::fabs(dot(cross(v1 - v0, v2 - v0), dir));
        if (projected_area <= T(0)) return T(0);
        return probability * d2 / projected_area;
    }

    struct LightSample {
        V3<T> direction{T(0), T(0), T(0)};
        T distance = T(0);
        V3<T> emission{T(0), T(0), T(0)};
        T pdf = T(0);
        int32_t triangle_index = -1;
    };
    // trace.slang:531-592
    LightSample sample_light(V3<T> origin, Rng& rng) {
        LightSample r;
        T px = T(s.scene_uniforms.light_probabilities[0]);
        T py = T(s.scene_uniforms.light_probabilities[1]);
        if (px + py <= T(0)) return r;
        if (T(rng.next()) < px) {
            uint32_t count = s.scene_uniforms.emissive_triangles_count;
            uint32_t slot = std::min((uint32_t)(T(rng.next()) * T(count)), count - 1u);
            const PtSceneAliasEntry& cand = s.scene_alias[slot];
            const PtSceneAliasEntry& entry = s.scene_alias[T(rng.next()) < T(cand.threshold) ? slot : cand.alias_index];
            const PtTriangleGeometry& g = s.geometries[entry.triangle_index];
            V3<T> v0 = vtx<T>(g, 0), v1 = vtx<T>(g, 1), v2 = vtx<T>(g, 2);
            T root = M::sqrt(T(rng.next()));
            T u = T(1) - root;
            T v = root * (T(1) - T(rng.next()));
            T w = T(1) - u - v;
            V3<T> delta = u * v0 + v * v1 + w * v2 - origin;
            T d2 = dot(delta, delta);
            if (d2 <= T(0)) return r;
            r.distance = M::sqrt(d2);
            r.direction = delta / r.distance;
            r.triangle_index = (int32_t)entry.triangle_index;
            r.pdf = triangle_light_pdf(entry.triangle_index, r.direction, d2, px * T(entry.selection_probability));
            const PtTriangleAttribute& at = s.attributes[entry.triangle_index];
            T uu, vv;
            bary_uv(at, u, v, w, uu, vv);
            const PtMaterial& mat = s.materials[at.material];
            r.emission = coverage(mat, uu, vv) * evaluate_emission(mat, uu, vv);
        } else {
            uint32_t W = s.skybox.width, Hh = s.skybox.height;
            uint32_t count = W * Hh;
            uint32_t slot = std::min((uint32_t)(T(rng.next()) * T(count)), count - 1u);
            const PtSkyboxAliasEntry& cand = s.skybox_alias[slot];
            uint32_t index = T(rng.next()) < T(cand.threshold) ? slot : cand.alias_index;
            uint32_t row = index / W, column = index % W;
            T phi = ((T(column) + T(rng.next())) / T(W) - F(0.5)) * (F(2.0) * F(3.14159265359));
            bool north = row < Hh / 2u;
            uint32_t pole_row = north ? row : Hh - 1u - row;
            T half_step = F(0.5) * F(3.14159265359) / T(Hh);
            T a = M::sin(T(pole_row) * half_step);
            T b = M::sin(T(pole_row + 1u) * half_step);
            T shs = lerp(a * a, b * b, T(rng.next()));
            T cos_theta = T(1) - F(2.0) * shs;
            T sin_theta = F(2.0) * M::sqrt(M::max(T(0), shs * (T(1) - shs)));
            r.direction = {sin_theta * M::cos(phi), north ? cos_theta : -cos_theta, sin_theta * M::sin(phi)};
            r.distance = T(ptm::FLT_MAX_);
            r.pdf = skybox_light_pdf(r.direction, py);
            r.emission = evaluate_skybox(r.direction);
        }
        return r;
    }

    // trace.slang:597-609
    T intersect_aabb(const PtBvhNode& n, V3<T> neg_o_rcp, V3<T> rcp, T hit_t) {
        V3<T> mn = {T(n.aabb_min[0]), T(n.aabb_min[1]), T(n.aabb_min[2])};
        V3<T> mx = {T(n.aabb_max[0]), T(n.aabb_max[1]), T(n.aabb_max[2])};
        V3<T> t0 = fma3(mn, rcp, neg_o_rcp);
        V3<T> t1 = fma3(mx, rcp, neg_o_rcp);
        V3<T> tmin = {M::min(t0.x, t1.x), M::min(t0.y, t1.y), M::min(t0.z, t1.z)};
        V3<T> tmax = {M::max(t0.x, t1.x), M::max(t0.y, t1.y), M::max(t0.z, t1.z)};
        T t_near = M::max(M::max(tmin.x, tmin.y), tmin.z);
        T t_far = M::min(M::min(tmax.x, tmax.y), tmax.z);
        return (t_near <= t_far && t_far > T(0) && t_near < hit_t) ? M::max(T(0), t_near) : T(-1);
    }

    struct Watertight {
        int kx, ky, kz;
        T Sx, Sy, Sz;
    };
    // trace.slang:769-783
    Watertight setup(V3<T> d, V3<T> rcp) {
        V3<T> ad = {std::fabs(d.x), std::fabs(d.y), std::fabs(d.z)};
        int kxy = ad.x > ad.y ? 0 : 1;
        int kz = ad[kxy] > ad.z ? kxy : 2;
        int kx_val = kz == 2 ? 0 : kz + 1;
        int ky_val = kx_val == 2 ? 0 : kx_val + 1;
        bool is_neg = d[kz] < T(0);
        Watertight w;
        w.kx = is_neg ? ky_val : kx_val;
        w.ky = is_neg ? kx_val : ky_val;
        w.kz = kz;
        w.Sz = rcp[kz];
        w.Sx = -d[w.kx] * w.Sz;
        w.Sy = -d[w.ky] * w.Sz;
        return w;
    }
    // trace.slang:805-835; returns true and T/v/w if the triangle is hit in (0, hit_t)
    bool tri_test(uint32_t tri, V3<T> o, const Watertight& wt, T hit_t, T& tt, T& bv, T& bw) {
        const PtTriangleGeometry& g = s.geometries[tri];
        V3<T> A = vtx<T>(g, 0) - o, B = vtx<T>(g, 1) - o, C = vtx<T>(g, 2) - o;
        T Ax = M::fma(wt.Sx, A[wt.kz], A[wt.kx]);
        T Ay = M::fma(wt.Sy, A[wt.kz], A[wt.ky]);
        T Bx = M::fma(wt.Sx, B[wt.kz], B[wt.kx]);
        T By = M::fma(wt.Sy, B[wt.kz], B[wt.ky]);
        T Cx = M::fma(wt.Sx, C[wt.kz], C[wt.kx]);
        T Cy = M::fma(wt.Sy, C[wt.kz], C[wt.ky]);
        T U = M::fma(Cx, By, -Cy * Bx);
        T V = M::fma(Ax, Cy, -Ay * Cx);
        T W = M::fma(Bx, Ay, -By * Ax);
        if (M::min(U, M::min(V, W)) < T(0) && M::max(U, M::max(V, W)) > T(0)) return false;
        T det = U + V + W;
        if (det == T(0)) return false;
        T inv_det = T(1) / det;
        T t = wt.Sz * M::fma(U, A[wt.kz], M::fma(V, B[wt.kz], W * C[wt.kz])) * inv_det;
        if (!(t > T(0) && t < hit_t)) return false;
        tt = t;
        bv = V * inv_det;
        bw = W * inv_det;
        return true;
    }

    // closest hit, trace.slang:785-884
    int32_t closest(V3<T> o, V3<T> d, V3<T> rcp, Rng* rng, T& hit_t, T& hit_v, T& hit_w,
                    uint32_t* pairs = nullptr, uint32_t* tris = nullptr) {
        Watertight wt = setup(d, rcp);
        hit_t = T(ptm::FLT_MAX_);
        hit_v = hit_w = T(0);
        int32_t hit = -1;
        uint32_t stack[128];
        uint32_t node_index = 0, sp = 0, np = 0, nt = 0;
        V3<T> neg = (-o) * rcp;
        for (;;) {
            const PtBvhNode& node = s.bvh_nodes[node_index];
            if (node.tri_count > 0u) {
                for (uint32_t i = 0; i < node.tri_count; ++i) {
                    uint32_t ti = node.left_first + i;
                    ++nt;
                    T tt, bv, bw;
                    if (tri_test(ti, o, wt, hit_t, tt, bv, bw)) {
                        if (rng && passes_through(ti, bv, bw, *rng)) continue;
                        hit_t = tt;
                        hit = (int32_t)ti;
                        hit_v = bv;
                        hit_w = bw;
                    }
                }
            } else {
                uint32_t li = node.left_first, ri = node.left_first + 1u;
                ++np;
                T dl = intersect_aabb(s.bvh_nodes[li], neg, rcp, hit_t);
                T dr = intersect_aabb(s.bvh_nodes[ri], neg, rcp, hit_t);
                bool hl = dl >= T(0), hr = dr >= T(0);
                if (hl && hr) {
                    bool left_closer = dl < dr;
                    stack[sp++] = left_closer ? ri : li;
                    node_index = left_closer ? li : ri;
                    continue;
                } else if (hl) {
                    node_index = li;
                    continue;
                } else if (hr) {
                    node_index = ri;
                    continue;
                }
            }
            if (sp == 0u) break;
            node_index = stack[--sp];
        }
        if (pairs) *pairs = np;
        if (tris) *tris = nt;
        cnt.pairs += np;
        cnt.tris += nt;
        return hit;
    }

    // any hit, trace.slang:614-717
    bool occluded(V3<T> o, V3<T> d, T hit_t, int32_t light_tri, Rng& rng) {
        V3<T> rcp = V3<T>{T(1), T(1), T(1)} / d;
        Watertight wt = setup(d, rcp);
        uint32_t stack[128];
        uint32_t node_index = 0, sp = 0;
        V3<T> neg = (-o) * rcp;
        for (;;) {
            const PtBvhNode& node = s.bvh_nodes[node_index];
            if (node.tri_count > 0u) {
                for (uint32_t i = 0; i < node.tri_count; ++i) {
                    uint32_t ti = node.left_first + i;
                    if ((int32_t)ti == light_tri) continue;
                    ++cnt.tris;
                    T tt, bv, bw;
                    if (tri_test(ti, o, wt, hit_t, tt, bv, bw) && !passes_through(ti, bv, bw, rng)) return true;
                }
            } else {
                uint32_t li = node.left_first, ri = node.left_first + 1u;
                ++cnt.pairs;
                T dl = intersect_aabb(s.bvh_nodes[li], neg, rcp, hit_t);
                T dr = intersect_aabb(s.bvh_nodes[ri], neg, rcp, hit_t);
                bool hl = dl >= T(0), hr = dr >= T(0);
                if (hl && hr) {
                    bool left_closer = dl < dr;
                    stack[sp++] = left_closer ? ri : li;
                    node_index = left_closer ? li : ri;
                    continue;
                } else if (hl) {
                    node_index = li;
                    continue;
                } else if (hr) {
                    node_index = ri;
                    continue;
                }
            }
            if (sp == 0u) break;
            node_index = stack[--sp];
        }
        return false;
    }

    V3<T> primary_dir(uint32_t px, uint32_t py, uint32_t iw, uint32_t ih, Rng& rng) {
        T r0 = T(rng.next());
        T r1 = T(rng.next());
        T u = (T(px) + r0) / T(iw);
        T v = (T(py) + r1) / T(ih);
        const PtCamera& c = s.camera;
        V3<T> llc = {T(c.lower_left_corner_minus_position[0]), T(c.lower_left_corner_minus_position[1]),
                     T(c.lower_left_corner_minus_position[2])};
        V3<T> h = {T(c.horizontal[0]), T(c.horizontal[1]), T(c.horizontal[2])};
        V3<T> vv = {T(c.vertical[0]), T(c.vertical[1]), T(c.vertical[2])};
        return normalize(llc + u * h + v * vv);
    }

    void texture_rgba(const uint8_t* atl, int layer, const float* uvo, T u, T v, bool srgb, T out[4]) {
        sample_atlas<T>(atl, s.atlas_size, layer, M::fma(u, T(uvo[2]), T(uvo[0])), M::fma(v, T(uvo[3]), T(uvo[1])),
                        srgb, out);
    }

    // one path sample, trace.slang:719-1297
    V3<T> trace_path(uint32_t px, uint32_t py, uint32_t iw, uint32_t ih, uint32_t max_bounces, uint32_t sample_count) {
        Rng rng{seed_of(px, py, sample_count)};
        const PtCamera& cam = s.camera;
        V3<T> ray_origin = {T(cam.position[0]), T(cam.position[1]), T(cam.position[2])};
        V3<T> ray_direction = primary_dir(px, py, iw, ih, rng);
        V3<T> rcp = V3<T>{T(1), T(1), T(1)} / ray_direction;
        V3<T> throughput = {T(1), T(1), T(1)};
        V3<T> final_color = {T(0), T(0), T(0)};
        T previous_bsdf_pdf = T(0);
        const T prob_geo = T(s.scene_uniforms.light_probabilities[0]);
        const T prob_sky = T(s.scene_uniforms.light_probabilities[1]);

        for (uint32_t bounce = 0; bounce <= max_bounces; ++bounce) {
            T hit_t, hit_v, hit_w;
            ++cnt.closest;
            int32_t hit_index = closest(ray_origin, ray_direction, rcp, &rng, hit_t, hit_v, hit_w);
            if (hit_index == -1) {
                T mis = T(1);
                if (bounce > 0u) mis = power_heuristic(previous_bsdf_pdf, skybox_light_pdf(ray_direction, prob_sky));
                final_color = final_color + throughput * evaluate_skybox(ray_direction) * mis;
                break;
            }
            const PtTriangleAttribute& at = s.attributes[hit_index];
            T hit_u = T(1) - hit_v - hit_w;
            T iu, iv;
            bary_uv(at, hit_u, hit_v, hit_w, iu, iv);
            const PtMaterial& mat = s.materials[at.material];
            Shading sh;
            sh.base_color = {T(mat.base_color[0]), T(mat.base_color[1]), T(mat.base_color[2])};
            sh.base_weight = T(mat.base_weight);
            sh.base_metalness = T(mat.base_metalness);
            sh.base_diffuse_roughness = T(mat.base_diffuse_roughness);
            sh.specular_color = {T(mat.specular_color_and_weight[0]), T(mat.specular_color_and_weight[1]),
                                 T(mat.specular_color_and_weight[2])};
            sh.specular_weight = T(mat.specular_color_and_weight[3]);
            T specular_roughness = T(mat.specular_roughness);
            T anisotropy = T(mat.specular_roughness_anisotropy);
            T specular_ior = T(mat.specular_ior);
            T tex[4];
            if (mat.base_color_tex_layer >= 0) {
                texture_rgba(s.srgb_atlases, mat.base_color_tex_layer, mat.base_color_uv, iu, iv, true, tex);
                sh.base_color = sh.base_color * V3<T>{tex[0], tex[1], tex[2]};
            }
            if (mat.metallic_roughness_tex_layer >= 0) {
                texture_rgba(s.linear_atlases, mat.metallic_roughness_tex_layer, mat.metallic_roughness_uv, iu, iv,
                             false, tex);
                sh.base_metalness *= tex[2];
                sh.base_diffuse_roughness *= tex[1];
            }
            if (mat.specular_color_tex_layer >= 0) {
                texture_rgba(s.srgb_atlases, mat.specular_color_tex_layer, mat.specular_color_uv, iu, iv, true, tex);
                sh.specular_color = sh.specular_color * V3<T>{tex[0], tex[1], tex[2]};
            }
            if (mat.specular_weight_tex_layer >= 0) {
                texture_rgba(s.linear_atlases, mat.specular_weight_tex_layer, mat.specular_weight_uv, iu, iv, false,
                             tex);
                sh.specular_weight *= tex[3];
            }
            V3<T> emission = evaluate_emission(mat, iu, iv);
            T emission_mis = T(1);
            if (bounce > 0u && prob_geo > T(0) && any_pos(emission)) {
                T light_pdf = triangle_light_pdf((uint32_t)hit_index, ray_direction, hit_t * hit_t,
                                                 prob_geo * triangle_selection_probability((uint32_t)hit_index));
                emission_mis = power_heuristic(previous_bsdf_pdf, light_pdf);
            }
            final_color = final_color + throughput * emission * emission_mis;

            V3<T> hit_point = ray_direction * hit_t + ray_origin;
            V3<T> Vv = -ray_direction;
            const PtTriangleGeometry& g = s.geometries[hit_index];
            V3<T> v0 = vtx<T>(g, 0), v1 = vtx<T>(g, 1), v2 = vtx<T>(g, 2);
            V3<T> flat_N_raw = normalize(cross(v1 - v0, v2 - v0));
            T bfm = dot(Vv, flat_N_raw) > T(0) ? T(1) : T(-1);
            V3<T> flat_N = flat_N_raw * bfm;
            V3<T> n0 = {T(at.n[0]), T(at.n[1]), T(at.n[2])};
            V3<T> n1 = {T(at.n[3]), T(at.n[4]), T(at.n[5])};
            V3<T> n2 = {T(at.n[6]), T(at.n[7]), T(at.n[8])};
            V3<T> N = normalize(hit_u * n0 + hit_v * n1 + hit_w * n2) * bfm;
            V3<T> t0 = {T(at.t0[0]), T(at.t0[1]), T(at.t0[2])};
            V3<T> t1 = {T(at.t1[0]), T(at.t1[1]), T(at.t1[2])};
            V3<T> t2 = {T(at.t2[0]), T(at.t2[1]), T(at.t2[2])};
            T t0w = T(at.t0[3]);
            V3<T> Tt = orthonormal_tangent(hit_u * t0 + hit_v * t1 + hit_w * t2, N) * bfm;
            V3<T> Bt = cross(N, Tt) * t0w * bfm;
            bool N_changed = false;
            if (mat.normal_tex_layer >= 0) {
                texture_rgba(s.linear_atlases, mat.normal_tex_layer, mat.normal_uv, iu, iv, false, tex);
                V3<T> local_n = sub_s(V3<T>{tex[0], tex[1], tex[2]} * F(2.0), T(1));
                T ns = T(mat.normal_scale);
                local_n = normalize(V3<T>{local_n.x * ns, local_n.y * ns, local_n.z});
                N = normalize(local_n.x * Tt + local_n.y * Bt + local_n.z * N);
                N_changed = true;
            }
            T flat_NdotV = dot(flat_N, Vv);
            T threshold = F(0.5) * flat_NdotV;
            T NdotV = dot(N, Vv);
            if (NdotV < threshold) {
                N = normalize(lerp3(N, flat_N, (threshold - NdotV) / (flat_NdotV - NdotV)));
                NdotV = dot(N, Vv);
                N_changed = true;
            }
            if (N_changed) {
                Tt = orthonormal_tangent(Tt, N);
                Bt = cross(N, Tt) * t0w * bfm;
            }
            alpha_anisotropic(specular_roughness, anisotropy, sh.ax, sh.ay);
            sh.sim = specular_ior_modulated(specular_ior, sh.specular_weight);
            sh.F0_conductor = sh.base_weight * sh.base_color;
            V3<T> wo = {dot(Tt, Vv), dot(Bt, Vv), NdotV};
            sh.one_plus_lambda_wo = T(1) + ggx_lambda(wo, sh.ax, sh.ay);
            const T r = sh.base_diffuse_roughness, z = wo.z;
            sh.ltc[0] = T(1) + r * (F(0.303392) + (F(-0.518982) + F(0.111709) * z) * z +
                                    (F(-0.276266) + F(0.335918) * z) * r);
            sh.ltc[1] = r * (F(-1.16407) + F(1.15859) * z + (F(0.150815) - F(0.150105) * z) * r) /
                        (z * z * z - F(1.43545));
            sh.ltc[2] = T(1) + r * (F(0.20013) + (F(-0.506373) + F(0.261777) * z) * z);
            sh.ltc[3] = r * (F(0.540852) + (F(-1.01625) + F(0.475392) * z) * z) / (F(-1.0743) + (F(0.0725628) + z) * z);
            T lsq = wo.x * wo.x + wo.y * wo.y;
            V3<T> X = {T(1), T(0), T(0)};
            if (lsq > T(0)) X = V3<T>{wo.x, wo.y, T(0)} * rsqrt(lsq);
            sh.bx = X;
            sh.by = {-X.y, X.x, T(0)};
            sh.bz = {T(0), T(0), T(1)};
            sh.E_dielectric = estimate_e_dielectric(wo, sh.one_plus_lambda_wo, sh.ax, sh.ay, sh.sim,
                                                    sh.specular_color, rng);
            const V3<T> LUMA = {F(0.2126), F(0.7152), F(0.0722)};
            T albedo_gloss = dot(sh.E_dielectric, LUMA);
            T albedo_diffuse = (T(1) - albedo_gloss) * sh.base_weight * dot(sh.base_color, LUMA);
            V3<T> fsm = sh.F0_conductor + s_sub(T(1), sh.F0_conductor) * fifth_power(T(1) - wo.z);
            T albedo_metal = dot(fsm, LUMA);
            T w_ggx = lerp(albedo_gloss, albedo_metal, sh.base_metalness);
            T w_diffuse = (T(1) - sh.base_metalness) * albedo_diffuse;
            sh.p_ggx = clampv(w_ggx / M::max(w_ggx + w_diffuse, F(1e-7)), F(0.1), F(0.9));
            sh.p_diffuse = T(1) - sh.p_ggx;
            sh.P_u = M::pow(r, F(0.1)) * (F(0.162925) + (F(-0.372058) + (F(0.538233) - F(0.290822) * z) * z) * z);
            sh.P_c = T(1) - sh.P_u;

            V3<T> next_origin = hit_point + flat_N * M::max(F(1e-4), F(1e-4) * length(hit_point));
            LightSample light = sample_light(next_origin, rng);
            V3<T> lwi = {dot(Tt, light.direction), dot(Bt, light.direction), dot(N, light.direction)};
            if (light.pdf > T(0) && lwi.z > T(0) && any_pos(light.emission)) {
                V3<T> f;
                T pdf;
                evaluate_lobes(sh, wo, lwi, f, pdf);
                ++cnt.shadow;
                if (!occluded(next_origin, light.direction, light.distance, light.triangle_index, rng)) {
                    T weight = bounce < max_bounces ? power_heuristic(light.pdf, pdf) : T(1);
                    final_color = final_color + throughput * f * lwi.z * light.emission * (weight / light.pdf);
                }
            }
            if (bounce == max_bounces) break;

            T two_pi_rand = F(2.0) * F(3.14159265359) * T(rng.next());
            T rand1 = T(rng.next());
            V3<T> wi;
            if (T(rng.next()) < sh.p_ggx) {
                wi = reflect(-wo, sample_vndf(wo, sh.ax, sh.ay, rand1, two_pi_rand));
            } else if (rand1 < sh.P_u) {
                T u1 = rand1 / sh.P_u;
                T sinTheta = M::sqrt(T(1) - u1 * u1);
                wi = {sinTheta * M::cos(two_pi_rand), sinTheta * M::sin(two_pi_rand), u1};
            } else {
                T u1 = (rand1 - sh.P_u) / sh.P_c;
                T a = sh.ltc[0], b = sh.ltc[1], c = sh.ltc[2], d = sh.ltc[3];
                T R = M::sqrt(u1);
                T y = R * M::sin(two_pi_rand);
                T sp = M::fma(F(0.5), rsqrt(M::fma(d, d, T(1))), F(0.5));
                T x = -lerp(M::sqrt(M::fma(-y, y, T(1))), R * M::cos(two_pi_rand), sp);
                V3<T> wh = {x, y, M::sqrt(M::max(T(1) - (x * x + y * y), T(0)))};
                V3<T> wu = {a * wh.x + b * wh.z, c * wh.y, d * wh.x + wh.z};
                wi = normalize(wu.x * sh.bx + wu.y * sh.by + wu.z * sh.bz);
            }
            if (wi.z <= T(0)) break;
            V3<T> f;
            T pdf;
            evaluate_lobes(sh, wo, wi, f, pdf);
            if (pdf <= T(0)) break;
            throughput = throughput * (f * wi.z / pdf);
            previous_bsdf_pdf = pdf;
            if (bounce > 3u) {
                T p_survive = clampv(M::max(throughput.x, M::max(throughput.y, throughput.z)), F(0.05), F(0.95));
                if (T(rng.next()) > p_survive) break;
                throughput = throughput / p_survive;
            }
            ray_origin = next_origin;
            ray_direction = wi.x * Tt + wi.y * Bt + wi.z * N;
            rcp = V3<T>{T(1), T(1), T(1)} / ray_direction;
        }
        return final_color;
    }
};

template <class T, class A>
int render_rows(const pt_scene_desc* s, const PtTraceParameters* p, uint32_t tile_w, uint32_t tile_h,
                uint32_t row_begin, uint32_t row_end, uint32_t samples, A* accum, float* storage, int threads,
                oracle_stats* stats) {
    if (!s || !p || !accum || samples == 0) return 1;
    Counters total;
    const uint32_t iw = p->image_size[0], ih = p->image_size[1];
    const uint32_t ox = p->tile_origin[0], oy = p->tile_origin[1];
#ifdef _OPENMP
    int nt = threads > 0 ? threads : omp_get_max_threads();
#else
    int nt = 1;
    (void)threads;
#endif
    std::vector<Counters> per(nt);
    const int64_t n = (int64_t)tile_w * (row_end - row_begin);
#pragma omp parallel for schedule(dynamic, 64) num_threads(nt)
    for (int64_t i = 0; i < n; ++i) {
#ifdef _OPENMP
        Counters& c = per[omp_get_thread_num()];
#else
        Counters& c = per[0];
#endif
        uint32_t gx = (uint32_t)(i % tile_w), gy = row_begin + (uint32_t)(i / tile_w);
        uint32_t px = gx + ox, py = gy + oy;
        if (px >= iw || py >= ih) continue;
        Kernel<T> k(*s, c);
        uint64_t idx = ((uint64_t)gy * tile_w + gx) * 4;
        for (uint32_t si = 0; si < samples; ++si) {
            uint32_t sc = p->sample_count + si;
            V3<T> L = k.trace_path(px, py, iw, ih, p->max_bounces, sc);
            T ar = sc > 1u ? T(accum[idx + 0]) : T(0);
            T ag = sc > 1u ? T(accum[idx + 1]) : T(0);
            T ab = sc > 1u ? T(accum[idx + 2]) : T(0);
            accum[idx + 0] = A(ar + L.x);
            accum[idx + 1] = A(ag + L.y);
            accum[idx + 2] = A(ab + L.z);
            accum[idx + 3] = A(1);
        }
        if (storage) {
            uint32_t sc = p->sample_count + samples - 1u;
            for (int ch = 0; ch < 3; ++ch)
                storage[idx + ch] = (float)(T(PT_EXPOSURE) * T(accum[idx + ch]) / T(sc));
            storage[idx + 3] = 1.0f;
        }
    }
    for (auto& c : per) {
        total.closest += c.closest;
        total.shadow += c.shadow;
        total.pairs += c.pairs;
        total.tris += c.tris;
    }
    if (stats) {
        stats->closest_rays = total.closest;
        stats->shadow_rays = total.shadow;
        stats->pair_tests = total.pairs;
        stats->tri_tests = total.tris;
    }
    return 0;
}

Counters g_dummy;

}  // namespace

extern "C" {

int oracle_render_f32(const pt_scene_desc* scene, const PtTraceParameters* params, uint32_t tile_w,
                      uint32_t tile_h, uint32_t samples, float* accum, float* storage, int threads,
                      oracle_stats* stats) {
    return render_rows<float, float>(scene, params, tile_w, tile_h, 0, tile_h, samples, accum, storage, threads,
                                     stats);
}
int oracle_render_f64(const pt_scene_desc* scene, const PtTraceParameters* params, uint32_t tile_w,
                      uint32_t tile_h, uint32_t samples, double* accum, int threads, oracle_stats* stats) {
    return render_rows<double, double>(scene, params, tile_w, tile_h, 0, tile_h, samples, accum, nullptr, threads,
                                       stats);
}
int oracle_render_rows_f32(const pt_scene_desc* scene, const PtTraceParameters* params, uint32_t tile_w,
                           uint32_t tile_h, uint32_t row_begin, uint32_t row_end, uint32_t samples, float* accum,
                           int threads, oracle_stats* stats) {
    if (row_end > tile_h || row_begin > row_end) return 1;
    return render_rows<float, float>(scene, params, tile_w, tile_h, row_begin, row_end, samples, accum, nullptr,
                                     threads, stats);
}

int oracle_primary_hits(const pt_scene_desc* s, const PtTraceParameters* p, uint32_t tile_w, uint32_t tile_h,
                        int32_t* tri, float* t) {
    const uint32_t iw = p->image_size[0], ih = p->image_size[1];
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)tile_w * tile_h; ++i) {
        Counters c;
        Kernel<float> k(*s, c);
        uint32_t px = (uint32_t)(i % tile_w) + p->tile_origin[0], py = (uint32_t)(i / tile_w) + p->tile_origin[1];
        if (px >= iw || py >= ih) {
            tri[i] = -1;
            t[i] = 0.0f;
            continue;
        }
        Rng rng{seed_of(px, py, p->sample_count)};
        V3<float> d = k.primary_dir(px, py, iw, ih, rng);
        V3<float> o = {s->camera.position[0], s->camera.position[1], s->camera.position[2]};
        V3<float> rcp = V3<float>{1.0f, 1.0f, 1.0f} / d;
        float ht, hv, hw;
        tri[i] = k.closest(o, d, rcp, &rng, ht, hv, hw);
        t[i] = ht;
    }
    return 0;
}

int oracle_intersect(const pt_scene_desc* s, const float* rays, uint32_t count, int32_t* hit, float* t, float* v,
                     float* w, uint32_t* pair_tests, uint32_t* tri_tests) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        Counters c;
        Kernel<float> k(*s, c);
        const float* r = rays + i * 6;
        V3<float> o = {r[0], r[1], r[2]}, d = {r[3], r[4], r[5]};
        V3<float> rcp = V3<float>{1.0f, 1.0f, 1.0f} / d;
        float ht, hv, hw;
        uint32_t np = 0, nt = 0;
        hit[i] = k.closest(o, d, rcp, nullptr, ht, hv, hw, &np, &nt);
        if (t) t[i] = ht;
        if (v) v[i] = hv;
        if (w) w[i] = hw;
        if (pair_tests) pair_tests[i] = np;
        if (tri_tests) tri_tests[i] = nt;
    }
    return 0;
}

int oracle_intersect_brute(const pt_scene_desc* s, const float* rays, uint32_t count, int32_t* hit, float* t) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        Counters c;
        Kernel<float> k(*s, c);
        const float* r = rays + i * 6;
        V3<float> o = {r[0], r[1], r[2]}, d = {r[3], r[4], r[5]};
        V3<float> rcp = V3<float>{1.0f, 1.0f, 1.0f} / d;
        auto wt = k.setup(d, rcp);
        float best = ptm::FLT_MAX_;
        int32_t h = -1;
        for (uint32_t ti = 0; ti < s->triangle_count; ++ti) {
            float tt, bv, bw;
            if (k.tri_test(ti, o, wt, best, tt, bv, bw)) {
                best = tt;
                h = (int32_t)ti;
            }
        }
        hit[i] = h;
        t[i] = best;
    }
    return 0;
}

int oracle_occluded(const pt_scene_desc* s, const float* rays, const float* tmax, const int32_t* skip, uint32_t count,
                    uint8_t* occluded) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int64_t i = 0; i < (int64_t)count; ++i) {
        Counters c;
        Kernel<float> k(*s, c);
        Rng rng{1u};
        const float* r = rays + i * 6;
        occluded[i] = k.occluded({r[0], r[1], r[2]}, {r[3], r[4], r[5]}, tmax[i], skip ? skip[i] : -1, rng) ? 1 : 0;
    }
    return 0;
}

static pt_scene_desc g_empty_scene;

float oracle_fresnel_exact(float ior, float cos_i) { return Kernel<float>(g_empty_scene, g_dummy).fresnel_exact(ior, cos_i); }
float oracle_ggx_lambda(const float v[3], float ax, float ay) {
    return Kernel<float>(g_empty_scene, g_dummy).ggx_lambda({v[0], v[1], v[2]}, ax, ay);
}
void oracle_sample_vndf(const float wo[3], float ax, float ay, float r, float two_pi_r, float h_out[3]) {
    V3<float> h = Kernel<float>(g_empty_scene, g_dummy).sample_vndf({wo[0], wo[1], wo[2]}, ax, ay, r, two_pi_r);
    h_out[0] = h.x;
    h_out[1] = h.y;
    h_out[2] = h.z;
}
float oracle_ggx_d(const float h[3], float ax, float ay) {
    return Kernel<float>(g_empty_scene, g_dummy).ggx_d({h[0], h[1], h[2]}, ax, ay);
}
void oracle_estimate_e_dielectric(const float wo[3], float ax, float ay, float sim, const float spec[3], uint32_t* rng,
                                  float out[3]) {
    Rng r{*rng};
    V3<float> e = Kernel<float>(g_empty_scene, g_dummy)
                      .estimate_e_dielectric({wo[0], wo[1], wo[2]}, 1.0f + Kernel<float>(g_empty_scene, g_dummy).ggx_lambda(
                                                                               {wo[0], wo[1], wo[2]}, ax, ay),
                                             ax, ay, sim, {spec[0], spec[1], spec[2]}, r);
    *rng = r.state;
    out[0] = e.x;
    out[1] = e.y;
    out[2] = e.z;
}
void oracle_alpha_anisotropic(float roughness, float anisotropy, float* ax, float* ay) {
    Kernel<float>(g_empty_scene, g_dummy).alpha_anisotropic(roughness, anisotropy, *ax, *ay);
}
void oracle_evaluate_skybox(const pt_scene_desc* s, const float dir[3], float out[3]) {
    Counters c;
    V3<float> v = Kernel<float>(*s, c).evaluate_skybox({dir[0], dir[1], dir[2]});
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}
float oracle_srgb_decode(uint32_t byte) { return srgb_lut[byte & 255u]; }
void oracle_orthonormal_tangent(const float t[3], const float n[3], float out[3]) {
    V3<float> v = Kernel<float>(g_empty_scene, g_dummy).orthonormal_tangent({t[0], t[1], t[2]}, {n[0], n[1], n[2]});
    out[0] = v.x;
    out[1] = v.y;
    out[2] = v.z;
}
void oracle_sample_atlas(const uint8_t* atlases, uint32_t size, int layer, float u, float v, int srgb, float out[4]) {
    sample_atlas<float>(atlases, size, layer, u, v, srgb != 0, out);
}
void oracle_primary_dir(const pt_scene_desc* s, uint32_t px, uint32_t py, uint32_t iw, uint32_t ih, uint32_t* rng,
                        float out[3]) {
    Counters c;
    Rng r{*rng};
    V3<float> d = Kernel<float>(*s, c).primary_dir(px, py, iw, ih, r);
    *rng = r.state;
    out[0] = d.x;
    out[1] = d.y;
    out[2] = d.z;
}
float oracle_e_fon_exact(float mu, float r) { return Kernel<float>(g_empty_scene, g_dummy).e_fon_exact(mu, r); }
float oracle_specular_ior_modulated(float ior, float weight) {
    return Kernel<float>(g_empty_scene, g_dummy).specular_ior_modulated(ior, weight);
}
float oracle_power_heuristic(float pdf, float other) {
    return Kernel<float>(g_empty_scene, g_dummy).power_heuristic(pdf, other);
}

// params: base rgb(3), base_weight, metalness, diffuse_roughness, spec rgb(3), spec_weight, spec_ior,
// spec_roughness, spec_roughness_anisotropy
static Kernel<float>::Shading make_shading(const float prm[13], V3<float> wo, const float e[3]) {
    Kernel<float> k(g_empty_scene, g_dummy);
    Kernel<float>::Shading sh;
    sh.base_color = {prm[0], prm[1], prm[2]};
    sh.base_weight = prm[3];
    sh.base_metalness = prm[4];
    sh.base_diffuse_roughness = prm[5];
    sh.specular_color = {prm[6], prm[7], prm[8]};
    sh.specular_weight = prm[9];
    k.alpha_anisotropic(prm[11], prm[12], sh.ax, sh.ay);
    sh.sim = k.specular_ior_modulated(prm[10], sh.specular_weight);
    sh.F0_conductor = sh.base_weight * sh.base_color;
    sh.one_plus_lambda_wo = 1.0f + k.ggx_lambda(wo, sh.ax, sh.ay);
    const float r = sh.base_diffuse_roughness, z = wo.z;
    sh.ltc[0] = 1.0f + r * (0.303392f + (-0.518982f + 0.111709f * z) * z + (-0.276266f + 0.335918f * z) * r);
    sh.ltc[1] = r * (-1.16407f + 1.15859f * z + (0.150815f - 0.150105f * z) * r) / (z * z * z - 1.43545f);
    sh.ltc[2] = 1.0f + r * (0.20013f + (-0.506373f + 0.261777f * z) * z);
    sh.ltc[3] = r * (0.540852f + (-1.01625f + 0.475392f * z) * z) / (-1.0743f + (0.0725628f + z) * z);
    float lsq = wo.x * wo.x + wo.y * wo.y;
    V3<float> X = {1.0f, 0.0f, 0.0f};
    if (lsq > 0.0f) X = V3<float>{wo.x, wo.y, 0.0f} * (1.0f / sqrtf(lsq));
    sh.bx = X;
    sh.by = {-X.y, X.x, 0.0f};
    sh.bz = {0.0f, 0.0f, 1.0f};
    sh.E_dielectric = {e[0], e[1], e[2]};
    const V3<float> LUMA = {0.2126f, 0.7152f, 0.0722f};
    float albedo_gloss = dot(sh.E_dielectric, LUMA);
    float albedo_diffuse = (1.0f - albedo_gloss) * sh.base_weight * dot(sh.base_color, LUMA);
    V3<float> fsm = sh.F0_conductor + s_sub(1.0f, sh.F0_conductor) * fifth_power(1.0f - wo.z);
    float albedo_metal = dot(fsm, LUMA);
    float w_ggx = lerp(albedo_gloss, albedo_metal, sh.base_metalness);
    float w_diffuse = (1.0f - sh.base_metalness) * albedo_diffuse;
    sh.p_ggx = clampv(w_ggx / fmaxf(w_ggx + w_diffuse, 1e-7f), 0.1f, 0.9f);
    sh.p_diffuse = 1.0f - sh.p_ggx;
    sh.P_u = ptm::pow_(r, 0.1f) * (0.162925f + (-0.372058f + (0.538233f - 0.290822f * z) * z) * z);
    sh.P_c = 1.0f - sh.P_u;
    return sh;
}

void oracle_evaluate_lobes(const float prm[13], const float wo[3], const float wi[3], const float e[3], float f_out[3],
                           float* pdf_out) {
    Kernel<float> k(g_empty_scene, g_dummy);
    V3<float> w = {wo[0], wo[1], wo[2]};
    auto sh = make_shading(prm, w, e);
    V3<float> f;
    float pdf;
    k.evaluate_lobes(sh, w, {wi[0], wi[1], wi[2]}, f, pdf);
    f_out[0] = f.x;
    f_out[1] = f.y;
    f_out[2] = f.z;
    *pdf_out = pdf;
}

int oracle_sample_bsdf(const float prm[13], const float wo[3], const float e[3], float r0, float r1, float r2,
                       float wi_out[3]) {
    Kernel<float> k(g_empty_scene, g_dummy);
    V3<float> w = {wo[0], wo[1], wo[2]};
    auto sh = make_shading(prm, w, e);
    float two_pi_rand = 2.0f * 3.14159265359f * r0;
    V3<float> wi;
    if (r2 < sh.p_ggx) {
        wi = reflect(-w, k.sample_vndf(w, sh.ax, sh.ay, r1, two_pi_rand));
    } else if (r1 < sh.P_u) {
        float u1 = r1 / sh.P_u;
        float sinTheta = sqrtf(1.0f - u1 * u1);
        wi = {sinTheta * ptm::cos_(two_pi_rand), sinTheta * ptm::sin_(two_pi_rand), u1};
    } else {
        float u1 = (r1 - sh.P_u) / sh.P_c;
        float a = sh.ltc[0], b = sh.ltc[1], c = sh.ltc[2], d = sh.ltc[3];
        float R = sqrtf(u1);
        float y = R * ptm::sin_(two_pi_rand);
        float sp = fmaf(0.5f, 1.0f / sqrtf(fmaf(d, d, 1.0f)), 0.5f);
        float x = -lerp(sqrtf(fmaf(-y, y, 1.0f)), R * ptm::cos_(two_pi_rand), sp);
        V3<float> wh = {x, y, sqrtf(fmaxf(1.0f - (x * x + y * y), 0.0f))};
        V3<float> wu = {a * wh.x + b * wh.z, c * wh.y, d * wh.x + wh.z};
        wi = normalize(wu.x * sh.bx + wu.y * sh.by + wu.z * sh.bz);
    }
    wi_out[0] = wi.x;
    wi_out[1] = wi.y;
    wi_out[2] = wi.z;
    return wi.z > 0.0f ? 1 : 0;
}

int oracle_sample_light(const pt_scene_desc* s, const float origin[3], uint32_t* rng, float dir_out[3],
                        float* distance, float emission_out[3], float* pdf) {
    Counters c;
    Kernel<float> k(*s, c);
    Rng r{*rng};
    auto ls = k.sample_light({origin[0], origin[1], origin[2]}, r);
    *rng = r.state;
    dir_out[0] = ls.direction.x;
    dir_out[1] = ls.direction.y;
    dir_out[2] = ls.direction.z;
    *distance = ls.distance;
    emission_out[0] = ls.emission.x;
    emission_out[1] = ls.emission.y;
    emission_out[2] = ls.emission.z;
    *pdf = ls.pdf;
    return ls.triangle_index;
}
float oracle_skybox_light_pdf(const pt_scene_desc* s, const float dir[3], float p) {
    Counters c;
    return Kernel<float>(*s, c).skybox_light_pdf({dir[0], dir[1], dir[2]}, p);
}
float oracle_triangle_light_pdf(const pt_scene_desc* s, uint32_t tri, const float dir[3], float d2, float prob) {
    Counters c;
    return Kernel<float>(*s, c).triangle_light_pdf(tri, {dir[0], dir[1], dir[2]}, d2, prob);
}
float oracle_rand(uint32_t* state) {
    Rng r{*state};
    float f = r.next();
    *state = r.state;
    return f;
}
uint32_t oracle_seed(uint32_t px, uint32_t py, uint32_t sample_count) { return seed_of(px, py, sample_count); }

// Evaluate one deterministic-libm function over arrays (fn: 0 sin, 1 cos,
// 2 acos, 3 atan2(x, y), 4 pow(x, y), 5 log2, 6 exp2, 7 and 8: the sine and
// the cosine from the kernels' one-reduction sincos_).
int oracle_math(int fn, const float* x, const float* y, float* out, uint32_t n) {
    for (uint32_t i = 0; i < n; ++i) {
        float sn, cs;
        switch (fn) {
            case 0: out[i] = ptm::sin_(x[i]); break;
            case 1: out[i] = ptm::cos_(x[i]); break;
            case 2: out[i] = ptm::acos_(x[i]); break;
            case 3: out[i] = ptm::atan2_(x[i], y[i]); break;
            case 4: out[i] = ptm::pow_(x[i], y[i]); break;
            case 5: out[i] = ptm::log2_(x[i]); break;
            case 6: out[i] = ptm::exp2_(x[i]); break;
            case 7: ptm::sincos_(x[i], sn, cs); out[i] = sn; break;
            case 8: ptm::sincos_(x[i], sn, cs); out[i] = cs; break;
            default: return 1;
        }
    }
    return 0;
}

// display.wgsl:19-46: ACES filmic fit, then the sRGB OETF
float oracle_display_channel(float c) {
    float a = 2.51f, b = 0.03f, cc = 2.43f, d = 0.59f, e = 0.14f;
    float m = (c * (a * c + b)) / (c * (cc * c + d) + e);
    m = fminf(fmaxf(m, 0.0f), 1.0f);
    return m > 0.0031308f ? ptm::pow_(m, 1.0f / 2.4f) * 1.055f - 0.055f : m * 12.92f;
}

}  // extern "C"
