// Device math for the SRT path-tracing kernels (gfx950).
//
// Every function here evaluates IEEE single-precision operations in a fixed
// order; the library is compiled with -ffp-contract=off and correctly rounded
// division and square root, so the kernels reproduce a host execution of the
// same formulas bit for bit.  Formulas follow the reference renderer's device
// headers (paths relative to the reference repository):
//   src/device_include/math/basic.h:26-258   sampling and Disney helpers
//   src/device_include/math/onb.h:5-40       orthonormal basis
//   src/device_include/scene/bxdf.h:59-387   BxDF
//   src/device_include/scene/gmaterial.h     material frame handling
// with these deliberate differences:
//   - normalize(v) = v * (1 / sqrt(dot(v, v))) instead of rsqrtf;
//   - sin/cos/log/exp/pow are the portable implementations below (Cody-Waite
//     reduction + Taylor polynomials; atanh series for log), not libm/ocml;
//   - RNG is PCG32 (O'Neill 2014) with one stream per (pixel, sample).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SRT_D __device__ __forceinline__

namespace srt {

// ---------------------------------------------------------------------------
// vectors (component-wise operators; lerp = a + t(b - a);
// clamp = max(a, min(f, b)); reflect(i, n) = i - (2n) dot(n, i))
// ---------------------------------------------------------------------------
struct f3 { float x, y, z; };
SRT_D f3 mk(float x, float y, float z) { f3 r; r.x = x; r.y = y; r.z = z; return r; }
SRT_D f3 mk(float a) { return mk(a, a, a); }
SRT_D f3 operator+(f3 a, f3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
SRT_D f3 operator-(f3 a, f3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
SRT_D f3 operator*(f3 a, f3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
SRT_D f3 operator/(f3 a, f3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
SRT_D f3 operator*(f3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
SRT_D f3 operator*(float s, f3 a) { return mk(s * a.x, s * a.y, s * a.z); }
SRT_D f3 operator/(f3 a, float s) { return mk(a.x / s, a.y / s, a.z / s); }
SRT_D f3 operator-(f3 a) { return mk(-a.x, -a.y, -a.z); }
SRT_D float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
SRT_D f3 cross(f3 a, f3 b) { return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
SRT_D float length(f3 v) { return __builtin_sqrtf(dot(v, v)); }
SRT_D f3 normalize(f3 v) { float inv = 1.0f / __builtin_sqrtf(dot(v, v)); return v * inv; }
SRT_D f3 lerp3(f3 a, f3 b, float t) { return a + t * (b - a); }
SRT_D float lerp1(float a, float b, float t) { return a + t * (b - a); }
SRT_D float clampf(float f, float a, float b) { return fmaxf(a, fminf(f, b)); }
SRT_D f3 reflect(f3 i, f3 n) { return i - (2.0f * n) * dot(n, i); }
SRT_D f3 refract(f3 i, f3 n, float eta) {
    float n_i = dot(n, -i);
    float k = 1.0f - eta * eta * (1.0f - n_i * n_i);
    if (k < 0.0f) return mk(0.0f);
    return eta * i + (eta * n_i - __builtin_sqrtf(k)) * n;
}
SRT_D f3 ld3(const float* p) { return mk(p[0], p[1], p[2]); }

constexpr float kPi = 3.14159265358979323846f;
constexpr float kInvPi = 0.318309886183790671538f;

// ---------------------------------------------------------------------------
// portable transcendentals
//   sincos: Cody-Waite reduction by pi/2 (three-part constant), Taylor
//           polynomials of degree 9 (sin) and 10 (cos) on [-pi/4, pi/4]
//   log:    m in [sqrt(1/2), sqrt(2)), log m = 2 atanh((m-1)/(m+1))
//   exp:    x = k ln2 + r, Taylor to r^8, exponent-field scaling
// ---------------------------------------------------------------------------
SRT_D void pt_sincosf(float x, float* s, float* c) {
    const float two_over_pi = 0.636619772367581343f;
    const float P1 = 1.5703125f;
    const float P2 = 4.837512969970703125e-4f;
    const float P3 = 7.54978995489188216e-8f;
    float kf = __builtin_rintf(x * two_over_pi);
    float r = x - kf * P1;
    r = r - kf * P2;
    r = r - kf * P3;
    float r2 = r * r;
    float sp = r2 * 2.75573192239858907e-6f;
    sp = sp + -1.98412698412698413e-4f;
    sp = sp * r2;
    sp = sp + 8.33333333333333333e-3f;
    sp = sp * r2;
    sp = sp + -1.66666666666666667e-1f;
    sp = sp * r2;
    sp = sp * r;
    float sv = r + sp;
    float cp = r2 * -2.75573192239858907e-7f;
    cp = cp + 2.48015873015873016e-5f;
    cp = cp * r2;
    cp = cp + -1.38888888888888889e-3f;
    cp = cp * r2;
    cp = cp + 4.16666666666666667e-2f;
    cp = cp * r2;
    cp = cp + -0.5f;
    cp = cp * r2;
    float cv = 1.0f + cp;
    int q = ((int)kf) & 3;
    float so, co;
    if (q == 0) { so = sv; co = cv; }
    else if (q == 1) { so = cv; co = -sv; }
    else if (q == 2) { so = -sv; co = -cv; }
    else { so = -cv; co = sv; }
    *s = so; *c = co;
}
SRT_D float pt_sinf(float x) { float s, c; pt_sincosf(x, &s, &c); return s; }
SRT_D float pt_cosf(float x) { float s, c; pt_sincosf(x, &s, &c); return c; }

SRT_D float pt_logf(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -__builtin_huge_valf() : __builtin_nanf("");
    if (x == __builtin_huge_valf()) return x;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 16777216.0f; e = -24; }
    uint32_t u = __float_as_uint(x);
    e += (int)((u >> 23) & 0xffu) - 127;
    float m = __uint_as_float((u & 0x007fffffu) | 0x3f800000u);
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float f = m - 1.0f;
    float s = f / (m + 1.0f);
    float z = s * s;
    float p = z * 2.22222222222222222e-1f;
    p = p + 2.85714285714285714e-1f;
    p = p * z;
    p = p + 4.0e-1f;
    p = p * z;
    p = p + 6.66666666666666667e-1f;
    p = p * z;
    p = p * s;
    float lm = (s + s) + p;
    float ef = (float)e;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.42860676533018704e-6f;
    return ef * ln2_hi + (lm + ef * ln2_lo);
}

SRT_D float pt_expf(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return __builtin_huge_valf();
    if (x < -103.972084f) return 0.0f;
    const float log2e = 1.44269504088896341f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.42860676533018704e-6f;
    float kf = __builtin_rintf(x * log2e);
    float r = x - kf * ln2_hi;
    r = r - kf * ln2_lo;
    float p = r * 2.48015873015873016e-5f;
    p = p + 1.98412698412698413e-4f;
    p = p * r;
    p = p + 1.38888888888888889e-3f;
    p = p * r;
    p = p + 8.33333333333333333e-3f;
    p = p * r;
    p = p + 4.16666666666666667e-2f;
    p = p * r;
    p = p + 1.66666666666666667e-1f;
    p = p * r;
    p = p + 0.5f;
    p = p * r;
    p = p * r;
    float er = (1.0f + r) + p;
    int k = (int)kf;
    if (k > 127) { er = er * 1.70141183e38f; k -= 127; }
    if (k < -126) { er = er * 1.17549435e-38f; k += 126; }
    return er * __uint_as_float((uint32_t)(k + 127) << 23);
}

SRT_D float pt_powf(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : __builtin_huge_valf();
    if (x < 0.0f) return __builtin_nanf("");
    if (x == 1.0f) return 1.0f;
    return pt_expf(y * pt_logf(x));
}

// ---------------------------------------------------------------------------
// RNG: PCG32 (XSH-RR), stream per (pixel, sample); PCG-hash seed mixing
// (Jarzynski & Olano 2020).
// ---------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state, inc;
    SRT_D uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    SRT_D float next_float() { return (float)(next() >> 8) * 5.9604644775390625e-8f; }
};
SRT_D uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
SRT_D Pcg32 path_rng(uint32_t pixel, uint32_t seed, uint32_t sample) {
    Pcg32 r;
    r.state = 0u;
    r.inc = ((uint64_t)sample << 1u) | 1u;
    r.next();
    r.state += (uint64_t)pcg_hash(pixel + pcg_hash(seed));
    r.next();
    return r;
}
// the stream increment is implied by the sample index, so a path carries only
// its 64-bit state between kernels
SRT_D Pcg32 rng_resume(uint64_t state, uint32_t sample) {
    Pcg32 r; r.state = state; r.inc = ((uint64_t)sample << 1u) | 1u; return r;
}

// ---------------------------------------------------------------------------
// sampling helpers (basic.h)
// ---------------------------------------------------------------------------
struct f2 { float x, y; };
// basic.h:26-38 (cdf in global memory)
SRT_D int binary_search(const float* cdf, int num, float x) {
    int lf = 0, rg = num - 1;
    while (lf <= rg) {
        int mid = (lf + rg) / 2;
        if (cdf[mid] <= x) lf = mid + 1; else rg = mid - 1;
    }
    return min(rg + 1, num - 1);
}
// basic.h:103-120
SRT_D f2 concentric_sample_disk(f2 s) {
    float ox = 2.0f * s.x - 1.0f, oy = 2.0f * s.y - 1.0f;
    f2 r0; r0.x = 0.0f; r0.y = 0.0f;
    if (ox == 0.0f && oy == 0.0f) return r0;
    float theta, r;
    const float pi_4 = 0.785398163397448309616f, pi_2 = 1.57079632679489661923f;
    if (fabsf(ox) > fabsf(oy)) { r = ox; theta = pi_4 * (oy / ox); }
    else { r = oy; theta = pi_2 - pi_4 * (ox / oy); }
    float sn, cs;
    pt_sincosf(theta, &sn, &cs);
    f2 o; o.x = r * cs; o.y = r * sn;
    return o;
}
// basic.h:127-137
SRT_D f3 cosine_sample_hemisphere(f2 s) {
    f2 d = concentric_sample_disk(s);
    float z = __builtin_sqrtf(fmaxf(0.0f, 1.0f - d.x * d.x - d.y * d.y));
    return mk(d.x, d.y, z);
}
SRT_D float cosine_hemisphere_pdf(float c) { return c * kInvPi; }
// basic.h:83-89
SRT_D f3 uniform_sample_sphere(f2 s) {
    float z = 1.0f - 2.0f * s.x;
    float r = __builtin_sqrtf(1.0f - z * z);
    float phi = 2.0f * kPi * s.y;
    float sp, cp;
    pt_sincosf(phi, &sp, &cp);
    return mk(r * cp, r * sp, z);
}
// basic.h:139-143
SRT_D f2 uniform_sample_triangle(f2 s) {
    float sx = __builtin_sqrtf(s.x);
    f2 o; o.x = 1.0f - sx; o.y = s.y * sx;
    return o;
}
// basic.h:154-165
SRT_D float schlick_fresnel(float cos_theta, float eta) {
    float r0 = (eta - 1.0f) / (eta + 1.0f);
    r0 = r0 * r0;
    float m = clampf(1.0f - cos_theta, 0.0f, 1.0f);
    return r0 + (1.0f - r0) * m * m * m * m * m;
}
SRT_D float schlick_fresnel(float cos_theta) { return schlick_fresnel(cos_theta, 1.0f); }
SRT_D float fresnel_mix(float metallic, float eta, float c) {
    return lerp1(schlick_fresnel(c, eta), schlick_fresnel(c), metallic);
}
SRT_D float luminance(f3 c) { return dot(c, mk(0.2126f, 0.7152f, 0.0722f)); }
// basic.h:172-201
SRT_D float GTR1(float n_h, float a) {
    if (a >= 1.0f) return kInvPi;
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * n_h * n_h;
    return (a2 - 1.0f) * kInvPi / (pt_logf(a2) * t);
}
SRT_D float GTR2(float n_h, float a) {
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * n_h * n_h;
    return a2 * kInvPi / (t * t);
}
SRT_D float GTR2_aniso(float n_h, float h_x, float h_y, float ax, float ay) {
    float a = h_x / ax, b = h_y / ay;
    float t = a * a + b * b + n_h * n_h;
    return kInvPi / (ax * ay * t * t);
}
SRT_D float smithG_GGX(float n_v, float alphaG) {
    float a = alphaG * alphaG;
    float b = n_v * n_v;
    return (2.0f * n_v) / (n_v + __builtin_sqrtf(a + b - a * b));
}
// basic.h:203-212
SRT_D f3 sample_GTR1(float a, f2 s) {
    if (a >= 1.0f) return cosine_sample_hemisphere(s);
    float a2 = a * a;
    float phi = 2.0f * kPi * s.x;
    float cos_theta = __builtin_sqrtf((1.0f - pt_powf(a2, 1.0f - s.y)) / (1.0f - a2));
    float sin_theta = clampf(__builtin_sqrtf(1.0f - (cos_theta * cos_theta)), 0.0f, 1.0f);
    float sn, cs;
    pt_sincosf(phi, &sn, &cs);
    return mk(cs * sin_theta, sn * sin_theta, cos_theta);
}
// basic.h:223-230
SRT_D f3 sample_GTR2_aniso(float ax, float ay, f2 s) {
    float phi = 2.0f * kPi * s.x;
    float sn, cs;
    pt_sincosf(phi, &sn, &cs);
    float sin_phi = ay * sn;
    float cos_phi = ax * cs;
    float tan_theta = __builtin_sqrtf(s.y / (1.0f - s.y));
    return normalize(mk(tan_theta * cos_phi, tan_theta * sin_phi, 1.0f));
}
// basic.h:232-250
SRT_D f3 sample_GGXVNDF(f3 V, float a, f2 s) {
    f3 Vh = normalize(mk(a * V.x, a * V.y, V.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    f3 T1 = (lensq > 0.0f) ? mk(-Vh.y, Vh.x, 0.0f) * (1.0f / __builtin_sqrtf(lensq)) : mk(1.0f, 0.0f, 0.0f);
    f3 T2 = cross(Vh, T1);
    float r = __builtin_sqrtf(s.x);
    float phi = 2.0f * kPi * s.y;
    float sn, cs;
    pt_sincosf(phi, &sn, &cs);
    float t1 = r * cs;
    float t2 = r * sn;
    float ss = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - ss) * __builtin_sqrtf(1.0f - t1 * t1) + ss * t2;
    f3 Nh = t1 * T1 + t2 * T2 + __builtin_sqrtf(fmaxf(0.0f, 1.0f - t1 * t1 - t2 * t2)) * Vh;
    return normalize(mk(a * Nh.x, a * Nh.y, fmaxf(0.0f, Nh.z)));
}

// ---------------------------------------------------------------------------
// Onb (onb.h:22-26 + to_world / to_local)
// ---------------------------------------------------------------------------
struct Onb {
    f3 z, x, y;
};
SRT_D Onb make_onb(f3 n, f3 t) {
    Onb o; o.z = n; o.x = normalize(t - n * dot(t, n)); o.y = cross(o.z, o.x); return o;
}
SRT_D f3 to_world(const Onb& o, f3 p) { return p.x * o.x + p.y * o.y + p.z * o.z; }
SRT_D f3 to_local(const Onb& o, f3 p) { return mk(dot(p, o.x), dot(p, o.y), dot(p, o.z)); }

// ---------------------------------------------------------------------------
// Materials: 80-byte record (type, 12 Disney parameters, base colour,
// emission colour, intensity).
// ---------------------------------------------------------------------------
struct MatRec {
    int32_t type;
    float ior, metallic, subsurface, roughness, specular, specularTint, anisotropic, sheen, sheenTint,
        clearcoat, clearcoatGloss, specTrans;
    float base_color[3];
    float emission_color[3];
    float intensity;
};
static_assert(sizeof(MatRec) == 80, "material record is 80 bytes");
enum { T_DIFFUSE = 0, T_DIFFTRANS = 1, T_DIELECTRIC = 2, T_DISNEY = 3 };

// gmaterial.h:20-31
SRT_D bool mat_is_emissive(const MatRec& m) { return m.intensity > 0.0f; }
SRT_D bool mat_is_specular(const MatRec& m) {
    return (m.type == T_DIELECTRIC) || (m.type == T_DISNEY && m.metallic >= 0.99f && m.roughness <= 0.01f) ||
           (m.type == T_DISNEY && m.specTrans >= 0.99f && m.roughness <= 0.01f && m.clearcoatGloss <= 0.01f);
}
SRT_D bool mat_is_transmissive(const MatRec& m) {
    return (m.type == T_DIELECTRIC) || (m.type == T_DIFFTRANS) || (m.type == T_DISNEY && m.metallic < 1.0f && m.specTrans > 0.0f);
}

struct BxSample { f3 f; f3 wi; float pdf; bool delta; };

// Disney quantities that depend only on the material and its colour
// (bxdf.h:93-99, :233-240); computed once per shading point.
struct DisneyK {
    float lum;
    f3 ctint, spec_color, sheen_color;
    float ax, ay;
};
SRT_D DisneyK disney_consts(const MatRec& m, f3 color) {
    DisneyK k;
    k.lum = luminance(color);
    k.ctint = (k.lum > 0.0f) ? color / k.lum : mk(1.0f);
    k.spec_color = lerp3((m.specular * 0.08f) * lerp3(mk(1.0f), k.ctint, m.specularTint), color, m.metallic);
    k.sheen_color = lerp3(mk(1.0f), k.ctint, m.sheenTint);
    float aspect = __builtin_sqrtf(1.0f - m.anisotropic * 0.9f);
    float a2 = m.roughness * m.roughness;
    k.ax = fmaxf(0.001f, a2 / aspect);
    k.ay = fmaxf(0.001f, a2 * aspect);
    return k;
}

// bxdf.h:93-174 (Disney evaluation)
SRT_D f3 disney_eval(const MatRec& m, const DisneyK& K, f3 wi, f3 wo, f3 color, float eta) {
    {
        f3 V = wo, L = wi;
        f3 H = (L.z > 0.0f) ? normalize(L + V) : normalize(L + V * eta);
        if (H.z <= 0.0f) H = -H;
        float V_H = dot(V, H), L_H = dot(L, H);
        const float lum = K.lum;
        const f3 spec_color = K.spec_color;
        const f3 sheen_color = K.sheen_color;
        float FM = fresnel_mix(m.metallic, eta, V_H);
        float diffuse_w = lum * (1.0f - m.metallic) * (1.0f - m.specTrans);
        float spec_reflect_w = luminance(lerp3(spec_color, mk(1.0f), FM));
        float spec_refract_w = (1.0f - FM) * (1.0f - m.metallic) * m.specTrans * lum;
        float clearcoat_w = m.clearcoat * (1.0f - m.metallic);

        f3 f_diffuse = mk(0.0f);
        if (diffuse_w > 0.0f && L.z > 0.0f) {
            float FL = schlick_fresnel(L.z), FV = schlick_fresnel(V.z);
            float FH = schlick_fresnel(L_H);
            float Fd90 = 0.5f + 2.0f * L_H * L_H * m.roughness;
            float Fd = lerp1(1.0f, Fd90, FL) * lerp1(1.0f, Fd90, FV);
            float Fss90 = L_H * L_H * m.roughness;
            float Fss = lerp1(1.0f, Fss90, FL) * lerp1(1.0f, Fss90, FV);
            float ss = 1.25f * (Fss * (1.0f / (L.z + V.z) - 0.5f) + 0.5f);
            f3 Fsheen = (FH * m.sheen) * sheen_color;
            f_diffuse = ((kInvPi * lerp1(Fd, ss, m.subsurface)) * color + Fsheen) * (1.0f - m.metallic) * (1.0f - m.specTrans);
        }
        f3 f_spec_reflect = mk(0.0f);
        if (spec_reflect_w > 0.0f && L.z > 0.0f && V.z > 0.0f) {
            float _FM = fresnel_mix(m.metallic, eta, L_H);
            f3 F = lerp3(spec_color, mk(1.0f), _FM);
            float D = GTR2_aniso(H.z, H.x, H.y, K.ax, K.ay);
            float G = smithG_GGX(fabsf(L.z), m.roughness) * smithG_GGX(fabsf(V.z), m.roughness);
            f_spec_reflect = F * D * G / (4.0f * L.z * V.z);
        }
        f3 f_spec_refract = mk(0.0f);
        if (spec_refract_w > 0.0f && L.z < 0.0f) {
            float F = schlick_fresnel(fabsf(V_H), eta);
            float D = GTR2(H.z, m.roughness);
            float denom = (L_H + V_H * eta) * (L_H + V_H * eta);
            float G = smithG_GGX(fabsf(L.z), m.roughness) * smithG_GGX(fabsf(V.z), m.roughness) * fabsf(L_H) * fabsf(V_H) * eta * eta / denom;
            f3 refract_color = mk(__builtin_sqrtf(color.x), __builtin_sqrtf(color.y), __builtin_sqrtf(color.z));
            f_spec_refract = refract_color * (1.0f - m.metallic) * m.specTrans * (1.0f - F) * D * G / (fabsf(L.z) * fabsf(V.z));
        }
        f3 f_clearcoat = mk(0.0f);
        if (clearcoat_w > 0.0f && L.z > 0.0f && V.z > 0.0f) {
            float FH = schlick_fresnel(L_H, 1.0f / 1.5f);
            float F = lerp1(0.04f, 1.0f, FH);
            float D = GTR1(H.z, m.clearcoatGloss);
            float G = smithG_GGX(L.z, 0.25f) * smithG_GGX(V.z, 0.25f);
            f_clearcoat = mk(0.25f) * m.clearcoat * F * D * G / (4.0f * L.z * V.z);
        }
        return f_diffuse + f_spec_reflect + f_spec_refract + f_clearcoat;
    }
}

// bxdf.h:79-180
SRT_D f3 bxdf_eval(const MatRec& m, f3 wi, f3 wo, f3 color, bool inner) {
    switch (m.type) {
    case T_DIFFUSE: return (wi.z <= 0.0f) ? mk(0.0f) : color * kInvPi;
    case T_DIFFTRANS: return color * kInvPi;
    case T_DIELECTRIC: return mk(0.0f);
    case T_DISNEY: return disney_eval(m, disney_consts(m, color), wi, wo, color, inner ? m.ior : 1.0f / m.ior);
    default: return mk(0.0f);
    }
}

struct Lobes { float dw, srw, stw, cw, ax, ay; };
// bxdf.h:233-252
SRT_D Lobes disney_lobes(const MatRec& m, const DisneyK& K, f3 V, float eta) {
    Lobes w;
    const float lum = K.lum;
    const f3 spec_color = K.spec_color;
    w.ax = K.ax;
    w.ay = K.ay;
    float FM = fresnel_mix(m.metallic, eta, V.z);
    w.dw = lum * (1.0f - m.metallic) * (1.0f - m.specTrans);
    w.srw = luminance(lerp3(spec_color, mk(1.0f), FM));
    w.stw = (1.0f - FM) * (1.0f - m.metallic) * m.specTrans * lum;
    w.cw = m.clearcoat * (1.0f - m.metallic);
    float total = w.dw + w.srw + w.stw + w.cw;
    w.dw /= total; w.srw /= total; w.stw /= total; w.cw /= total;
    return w;
}
// bxdf.h:289-310 / 365-380
SRT_D float disney_pdf(const MatRec& m, f3 L, f3 V, float eta, const Lobes& w) {
    f3 H_reflect = normalize(L + V);
    if (H_reflect.z < 0.0f) H_reflect = -H_reflect;
    f3 H_refract = normalize(L + V * eta);
    if (H_refract.z < 0.0f) H_refract = -H_refract;
    float pdf = 0.0f;
    if (w.dw > 0.0f && L.z > 0.0f) pdf += w.dw * cosine_hemisphere_pdf(L.z);
    if (w.srw > 0.0f && dot(V, H_reflect) > 0.0f)
        pdf += w.srw * GTR2_aniso(H_reflect.z, H_reflect.x, H_reflect.y, w.ax, w.ay) * H_reflect.z / (4.0f * dot(V, H_reflect));
    if (w.stw > 0.0f && dot(V, H_refract) > 0.0f && dot(L, H_refract) < 0.0f) {
        float V_H = dot(V, H_refract), L_H = dot(L, H_refract);
        float denom = (L_H + V_H * eta);
        pdf += w.stw * GTR2(H_refract.z, m.roughness) * smithG_GGX(V.z, m.roughness) * fmaxf(dot(V, H_refract), 0.0f) / V.z *
               fabsf(L_H) / (denom * denom);
    }
    if (w.cw > 0.0f && dot(V, H_reflect) > 0.0f)
        pdf += w.cw * GTR1(H_reflect.z, m.clearcoatGloss) * H_reflect.z / (4.0f * dot(V, H_reflect));
    return pdf;
}

SRT_D BxSample disney_sample(const MatRec& m, const DisneyK& K, const Lobes& w, f3 wo, f2 rnd, f3 color, float eta);

// bxdf.h:182-319
SRT_D BxSample bxdf_sample(const MatRec& m, f3 wo, f2 rnd, f3 color, bool inner) {
    BxSample none; none.f = mk(0.0f); none.wi = mk(0.0f); none.pdf = 0.0f; none.delta = false;
    switch (m.type) {
    case T_DIFFUSE: {
        f3 wi = cosine_sample_hemisphere(rnd);
        float pdf = cosine_hemisphere_pdf(wi.z);
        BxSample r; r.f = bxdf_eval(m, wi, wo, color, inner); r.wi = wi; r.pdf = pdf; r.delta = false; return r;
    }
    case T_DIFFTRANS: {
        if (rnd.x < 0.5f) {
            rnd.x *= 2.0f;
            f3 wi = cosine_sample_hemisphere(rnd);
            float pdf = cosine_hemisphere_pdf(wi.z);
            BxSample r; r.f = bxdf_eval(m, wi, wo, color, inner); r.wi = wi; r.pdf = 0.5f * pdf; r.delta = false; return r;
        } else {
            rnd.x = (rnd.x - 0.5f) * 2.0f;
            f3 wi = cosine_sample_hemisphere(rnd);
            float pdf = cosine_hemisphere_pdf(wi.z);
            BxSample r; r.f = bxdf_eval(m, -wi, wo, color, inner); r.wi = -wi; r.pdf = 0.5f * pdf; r.delta = false; return r;
        }
    }
    case T_DIELECTRIC: {
        float eta = inner ? m.ior : 1.0f / m.ior;
        f3 n = mk(0.0f, 0.0f, 1.0f);
        f3 wi = refract(-wo, n, eta);
        float reflect_ratio = (dot(wi, wi) == 0.0f) ? 1.0f : schlick_fresnel(wo.z, eta);
        if (rnd.x <= reflect_ratio) wi = reflect(-wo, n);
        wi = normalize(wi);
        BxSample r; r.f = color; r.wi = wi; r.pdf = 1.0f; r.delta = true; return r;
    }
    case T_DISNEY: {
        float eta = inner ? m.ior : 1.0f / m.ior;
        const DisneyK K = disney_consts(m, color);
        return disney_sample(m, K, disney_lobes(m, K, wo, eta), wo, rnd, color, eta);
    }
    default: return none;
    }
}

// bxdf.h:228-313 (Disney sampling) with the shading point's constants and lobe weights
SRT_D BxSample disney_sample(const MatRec& m, const DisneyK& K, const Lobes& w, f3 wo, f2 rnd, f3 color, float eta) {
    BxSample none; none.f = mk(0.0f); none.wi = mk(0.0f); none.pdf = 0.0f; none.delta = false;
    {
        f3 V = wo;
        f3 L = mk(0.0f);
        if (rnd.x < w.dw) {
            rnd.x /= w.dw;
            L = cosine_sample_hemisphere(rnd);
        } else if (rnd.x < w.dw + w.srw) {
            rnd.x = (rnd.x - w.dw) / w.srw;
            f3 H = sample_GTR2_aniso(w.ax, w.ay, rnd);
            if (dot(V, H) <= 0.0f) return none;
            L = normalize(reflect(-V, H));
        } else if (rnd.x < w.dw + w.srw + w.stw) {
            rnd.x = (rnd.x - w.dw - w.srw) / w.stw;
            f3 H = sample_GGXVNDF(V, m.roughness, rnd);
            L = refract(-V, H, eta);
            if (dot(V, H) <= 0.0f || dot(L, L) == 0.0f) return none;
            L = normalize(L);
        } else {
            rnd.x = (rnd.x - w.dw - w.srw - w.stw) / w.cw;
            f3 H = sample_GTR1(m.clearcoatGloss, rnd);
            if (dot(V, H) <= 0.0f) return none;
            L = normalize(reflect(-V, H));
        }
        BxSample r;
        r.pdf = disney_pdf(m, L, V, eta, w);
        r.f = disney_eval(m, K, L, wo, color, eta);
        r.wi = L;
        r.delta = false;
        return r;
    }
}

// bxdf.h:321-387
SRT_D float bxdf_pdf(const MatRec& m, f3 wi, f3 wo, f3 color, bool inner) {
    switch (m.type) {
    case T_DIFFUSE: return (wi.z <= 0.0f) ? 0.0f : cosine_hemisphere_pdf(wi.z);
    case T_DIFFTRANS: return 0.5f * cosine_hemisphere_pdf(fabsf(wi.z));
    case T_DIELECTRIC: return 0.0f;
    case T_DISNEY: {
        float eta = inner ? m.ior : 1.0f / m.ior;
        return disney_pdf(m, wi, wo, eta, disney_lobes(m, disney_consts(m, color), wo, eta));
    }
    default: return 0.0f;
    }
}

// gmaterial.h:59-104
SRT_D f3 mat_eval(const MatRec& m, f3 wi, f3 wo, const Onb& onb, f3 color) {
    f3 wil = to_local(onb, wi), wol = to_local(onb, wo);
    bool inner = false;
    if (wol.z <= 0.0f) { wol = -wol; wil = -wil; inner = true; }
    return bxdf_eval(m, wil, wol, color, inner);
}
SRT_D BxSample mat_sample(const MatRec& m, f3 wo, f2 rnd, const Onb& onb, f3 color) {
    f3 wol = to_local(onb, wo);
    bool inner = false;
    if (wol.z <= 0.0f) { wol = -wol; inner = true; }
    BxSample s = bxdf_sample(m, wol, rnd, color, inner);
    if (inner) s.wi = -s.wi;
    s.wi = to_world(onb, s.wi);
    return s;
}
SRT_D float mat_pdf(const MatRec& m, f3 wi, f3 wo, const Onb& onb, f3 color) {
    f3 wil = to_local(onb, wi), wol = to_local(onb, wo);
    bool inner = false;
    if (wol.z <= 0.0f) { wol = -wol; wil = -wil; inner = true; }
    return bxdf_pdf(m, wil, wol, color, inner);
}

// One shading point: the outgoing direction in the local frame (flipped to
// the upper hemisphere, gmaterial.h:59-104) and, for Disney materials, the
// constants and lobe weights that every evaluation at this point shares.
// sp_eval / sp_pdf / sp_sample give the same values as mat_eval / mat_pdf /
// mat_sample with the same arguments.
struct ShadePoint {
    Onb onb;
    f3 wol;
    bool inner;
    float eta;
    DisneyK K;
    Lobes w;
};
SRT_D ShadePoint shade_point(const MatRec& m, f3 wo, const Onb& onb, f3 color) {
    ShadePoint p;
    p.onb = onb;
    p.wol = to_local(onb, wo);
    p.inner = false;
    if (p.wol.z <= 0.0f) { p.wol = -p.wol; p.inner = true; }
    p.eta = p.inner ? m.ior : 1.0f / m.ior;
    if (m.type == T_DISNEY) {
        p.K = disney_consts(m, color);
        p.w = disney_lobes(m, p.K, p.wol, p.eta);
    }
    return p;
}
SRT_D f3 sp_eval(const MatRec& m, const ShadePoint& p, f3 wi, f3 color) {
    f3 wil = to_local(p.onb, wi);
    if (p.inner) wil = -wil;
    if (m.type == T_DISNEY) return disney_eval(m, p.K, wil, p.wol, color, p.eta);
    return bxdf_eval(m, wil, p.wol, color, p.inner);
}
SRT_D float sp_pdf(const MatRec& m, const ShadePoint& p, f3 wi, f3 color) {
    f3 wil = to_local(p.onb, wi);
    if (p.inner) wil = -wil;
    if (m.type == T_DISNEY) return disney_pdf(m, wil, p.wol, p.eta, p.w);
    return bxdf_pdf(m, wil, p.wol, color, p.inner);
}
SRT_D BxSample sp_sample(const MatRec& m, const ShadePoint& p, f2 rnd, f3 color) {
    BxSample s = (m.type == T_DISNEY) ? disney_sample(m, p.K, p.w, p.wol, rnd, color, p.eta)
                                      : bxdf_sample(m, p.wol, rnd, color, p.inner);
    if (p.inner) s.wi = -s.wi;
    s.wi = to_world(p.onb, s.wi);
    return s;
}

}  // namespace srt
