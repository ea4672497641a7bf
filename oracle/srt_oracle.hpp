// SRT path-tracing hot path: CPU oracle (test infrastructure, not product code).
//
// This file is a CPU restatement of the reference renderer's megakernel path
// tracer and of the data it reads.  It exists only to check the HIP product
// path: tests, the smoke check and the CPU-baseline leg of the benchmark link
// it; nothing in the product loads it.
//
// Everything is templated on Real:
//   Real = float   -> the host execution of the reference integrator; it uses
//                     the same IEEE operations in the same order as the GPU
//                     kernels (no FMA contraction, correctly rounded / and
//                     sqrt, the portable transcendentals below), so per-pixel
//                     results can be compared at 1e-4.
//   Real = double  -> the FP64 oracle used for RMSE of converged images; it
//                     uses libm's double functions.
//
// Reference files followed (paths relative to the reference repository):
//   src/shader/integrator/path_tracer.cu:50-169      integrator loop
//   src/device_include/params/hit_info.h:22-65        hit reconstruction
//   src/device_include/scene/bxdf.h:59-387            BxDF (Disney etc.)
//   src/device_include/scene/gmaterial.h:6-104        material frame handling
//   src/device_include/scene/light.h:48-84            light selection
//   src/device_include/scene/gmesh.h:45-85            triangle sampling
//   src/device_include/scene/camera.h:27-32,107-117,165-170  pinhole camera
//   src/device_include/math/transform.h:136-145       LookAt
//   src/device_include/math/basic.h:26-258            sampling / Disney helpers
//   src/device_include/math/onb.h:5-40                orthonormal basis
//   src/scene/scene.cu:240-287                        light CDF
//   src/scene/scene.cu:336-395                        linear blend skinning
//   src/scene/mesh.cpp:3-18                           per-mesh area CDF
//   src/passes/postprocess/accumulate.cu:22-26        running mean
//   src/passes/postprocess/tonemapping.cu:14-67       tone operators
// Deliberate deviations (identical in the product kernels):
//   D1 RNG: PCG32 (O'Neill 2014); stream per (pixel, sample): initstate =
//      hash(pixel, seed) (PCG-hash, Jarzynski & Olano 2020), initseq = sample.
//      Consumption order inside one sample follows path_tracer.cu:66-159.
//   D2 CDF remap uses the float value of cdf[idx-1] (light.h:65 and
//      gmesh.h:49 truncate it to int).
//   D3 closest hit = lexicographic minimum of (t, primitive id) over all
//      triangles with t in [tmin, tmax], watertight test of Woop, Benthin and
//      Wald (JCGT 2013); shadow rays stop at the first hit in range.
//   D4 no fast math: normalize(v) = v * (1 / sqrt(dot(v, v))).
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <limits>

namespace orc {

// ---------------------------------------------------------------------------
// Portable float transcendentals (bit-identical wherever IEEE single precision
// with round-to-nearest and no contraction is used).
//   sin/cos: Cody & Waite (1980) additive range reduction by pi/2 in three
//            parts, then Taylor polynomials on [-pi/4, pi/4] (truncation error
//            < 2e-9, far below float rounding).
//   log:     exponent split, m in [sqrt(1/2), sqrt(2)), log m = 2 atanh(s),
//            s = (m-1)/(m+1), odd series to s^9 (truncation < 1e-9).
//   exp:     x = k ln2 + r with ln2 split in two, Taylor to r^8 on |r|<=ln2/2,
//            scale by 2^k through the exponent field.
//   pow:     exp(y log x) for x > 0.
// ---------------------------------------------------------------------------
inline uint32_t f2u(float f) { uint32_t u; std::memcpy(&u, &f, 4); return u; }
inline float u2f(uint32_t u) { float f; std::memcpy(&f, &u, 4); return f; }

inline void pt_sincosf(float x, float* s, float* c) {
    const float two_over_pi = 0.636619772367581343f;
    const float P1 = 1.5703125f;                      // 201/128 exactly
    const float P2 = 4.837512969970703125e-4f;        // next 12 bits of pi/2
    const float P3 = 7.54978995489188216e-8f;         // remainder of pi/2
    float kf = std::rint(x * two_over_pi);
    float r = x - kf * P1;
    r = r - kf * P2;
    r = r - kf * P3;
    float r2 = r * r;
    // sin r = r + r^3 (-1/6 + r^2 (1/120 + r^2 (-1/5040 + r^2 / 362880)))
    float sp = r2 * 2.75573192239858907e-6f;
    sp = sp + -1.98412698412698413e-4f;
    sp = sp * r2;
    sp = sp + 8.33333333333333333e-3f;
    sp = sp * r2;
    sp = sp + -1.66666666666666667e-1f;
    sp = sp * r2;
    sp = sp * r;
    float sv = r + sp;
    // cos r = 1 + r^2 (-1/2 + r^2 (1/24 + r^2 (-1/720 + r^2 (1/40320 - r^2/3628800))))
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
    if (q == 0) { *s = sv; *c = cv; }
    else if (q == 1) { *s = cv; *c = -sv; }
    else if (q == 2) { *s = -sv; *c = -cv; }
    else { *s = -cv; *c = sv; }
}
inline float pt_sinf(float x) { float s, c; pt_sincosf(x, &s, &c); return s; }
inline float pt_cosf(float x) { float s, c; pt_sincosf(x, &s, &c); return c; }

inline float pt_logf(float x) {
    if (!(x > 0.0f)) return (x == 0.0f) ? -INFINITY : NAN;
    if (x == INFINITY) return INFINITY;
    int e = 0;
    if (x < 1.17549435e-38f) { x = x * 16777216.0f; e = -24; }   // denormal input
    uint32_t u = f2u(x);
    e += (int)((u >> 23) & 0xff) - 127;
    float m = u2f((u & 0x007fffffu) | 0x3f800000u);              // [1, 2)
    if (m > 1.41421356f) { m = m * 0.5f; e = e + 1; }
    float f = m - 1.0f;
    float s = f / (m + 1.0f);
    float z = s * s;
    // log m = 2s + s*z*(2/3 + z*(2/5 + z*(2/7 + z*2/9)))
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
    const float ln2_hi = 0.693145751953125f;            // 11 significant bits
    const float ln2_lo = 1.42860676533018704e-6f;
    return ef * ln2_hi + (lm + ef * ln2_lo);
}

inline float pt_expf(float x) {
    if (x != x) return x;
    if (x > 88.7228394f) return INFINITY;
    if (x < -103.972084f) return 0.0f;
    const float log2e = 1.44269504088896341f;
    const float ln2_hi = 0.693145751953125f;
    const float ln2_lo = 1.42860676533018704e-6f;
    float kf = std::rint(x * log2e);
    float r = x - kf * ln2_hi;
    r = r - kf * ln2_lo;
    // e^r = 1 + r + r^2/2 + ... + r^8/40320
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
    // scale by 2^k in at most two exact steps
    if (k > 127) { er = er * 1.70141183e38f; k -= 127; }      // 2^127
    if (k < -126) { er = er * 1.17549435e-38f; k += 126; }    // 2^-126
    return er * u2f((uint32_t)(k + 127) << 23);
}

inline float pt_powf(float x, float y) {
    if (y == 0.0f) return 1.0f;
    if (x == 0.0f) return (y > 0.0f) ? 0.0f : INFINITY;
    if (x < 0.0f) return NAN;
    if (x == 1.0f) return 1.0f;
    return pt_expf(y * pt_logf(x));
}

// Math traits: float path uses the portable functions, double path libm.
template <class R> struct M;
template <> struct M<float> {
    static float sqrt(float x) { return std::sqrt(x); }
    static float sin(float x) { return pt_sinf(x); }
    static float cos(float x) { return pt_cosf(x); }
    static float log(float x) { return pt_logf(x); }
    static float pow(float x, float y) { return pt_powf(x, y); }
    static float abs(float x) { return std::fabs(x); }
    static float fmax(float a, float b) { return std::fmax(a, b); }
    static float fmin(float a, float b) { return std::fmin(a, b); }
};
template <> struct M<double> {
    static double sqrt(double x) { return std::sqrt(x); }
    static double sin(double x) { return std::sin(x); }
    static double cos(double x) { return std::cos(x); }
    static double log(double x) { return std::log(x); }
    static double pow(double x, double y) { return std::pow(x, y); }
    static double abs(double x) { return std::fabs(x); }
    static double fmax(double a, double b) { return std::fmax(a, b); }
    static double fmin(double a, double b) { return std::fmin(a, b); }
};

// ---------------------------------------------------------------------------
// Vector math (semantics of the helper functions used by the reference:
// component-wise operators, lerp = a + t(b-a), clamp = max(a, min(f, b)),
// reflect(i, n) = i - 2 n dot(n, i)).
// ---------------------------------------------------------------------------
template <class R> struct V3 {
    R x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(R a, R b, R c) : x(a), y(b), z(c) {}
    explicit V3(R a) : x(a), y(a), z(a) {}
};
template <class R> inline V3<R> operator+(V3<R> a, V3<R> b) { return V3<R>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class R> inline V3<R> operator-(V3<R> a, V3<R> b) { return V3<R>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class R> inline V3<R> operator*(V3<R> a, V3<R> b) { return V3<R>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class R> inline V3<R> operator/(V3<R> a, V3<R> b) { return V3<R>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class R> inline V3<R> operator*(V3<R> a, R s) { return V3<R>(a.x * s, a.y * s, a.z * s); }
template <class R> inline V3<R> operator*(R s, V3<R> a) { return V3<R>(s * a.x, s * a.y, s * a.z); }
template <class R> inline V3<R> operator/(V3<R> a, R s) { return V3<R>(a.x / s, a.y / s, a.z / s); }
template <class R> inline V3<R> operator-(V3<R> a) { return V3<R>(-a.x, -a.y, -a.z); }
template <class R> inline R dot(V3<R> a, V3<R> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class R> inline V3<R> cross(V3<R> a, V3<R> b) {
    return V3<R>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class R> inline R length(V3<R> v) { return M<R>::sqrt(dot(v, v)); }
template <class R> inline V3<R> normalize(V3<R> v) { R inv = R(1) / M<R>::sqrt(dot(v, v)); return v * inv; }
template <class R> inline V3<R> lerp(V3<R> a, V3<R> b, R t) { return a + t * (b - a); }
template <class R> inline R lerp(R a, R b, R t) { return a + t * (b - a); }
template <class R> inline R clampf(R f, R a, R b) { return M<R>::fmax(a, M<R>::fmin(f, b)); }
template <class R> inline V3<R> reflect(V3<R> i, V3<R> n) { return i - (R(2) * n) * dot(n, i); }
template <class R> inline V3<R> refract(V3<R> i, V3<R> n, R eta) {
    R n_i = dot(n, -i);
    R k = R(1) - eta * eta * (R(1) - n_i * n_i);
    if (k < R(0)) return V3<R>(R(0));
    return eta * i + (eta * n_i - M<R>::sqrt(k)) * n;
}
template <class R> inline R max_comp(V3<R> v) { return M<R>::fmax(v.x, M<R>::fmax(v.y, v.z)); }
template <class R> inline V3<R> cvt(const float* p) { return V3<R>((R)p[0], (R)p[1], (R)p[2]); }

template <class R> struct C {
    static R pi() { return (R)3.14159265358979323846; }
    static R inv_pi() { return (R)0.318309886183790671538; }
};

// ---------------------------------------------------------------------------
// RNG (D1): PCG32, XSH-RR output (O'Neill 2014), float = top 24 bits * 2^-24.
// ---------------------------------------------------------------------------
struct Pcg32 {
    uint64_t state = 0, inc = 1;
    uint32_t next() {
        uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((-rot) & 31u));
    }
    void seed(uint64_t initstate, uint64_t initseq) {
        state = 0u;
        inc = (initseq << 1u) | 1u;
        next();
        state += initstate;
        next();
    }
    float next_float() { return (float)(next() >> 8) * 5.9604644775390625e-8f; }
};
// PCG-hash (Jarzynski & Olano 2020, "Hash Functions for GPU Rendering").
inline uint32_t pcg_hash(uint32_t v) {
    uint32_t state = v * 747796405u + 2891336453u;
    uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
    return (word >> 22u) ^ word;
}
inline uint32_t seed_hash(uint32_t pixel, uint32_t seed) { return pcg_hash(pixel + pcg_hash(seed)); }
inline void path_rng(Pcg32& r, uint32_t pixel, uint32_t seed, uint32_t sample) {
    r.seed((uint64_t)seed_hash(pixel, seed), (uint64_t)sample);
}

// ---------------------------------------------------------------------------
// Sampling helpers (basic.h:26-258)
// ---------------------------------------------------------------------------
// basic.h:26-38
inline int binary_search(const float* cdf, int num, float x) {
    int lf = 0, rg = num - 1;
    while (lf <= rg) {
        int mid = (lf + rg) / 2;
        if (cdf[mid] <= x) lf = mid + 1; else rg = mid - 1;
    }
    return std::min(rg + 1, num - 1);
}
template <class R> struct V2 { R x, y; };

// basic.h:96-101
template <class R> inline V2<R> uniform_sample_disk(V2<R> s) {
    R r = M<R>::sqrt(s.x);
    R th = R(2) * C<R>::pi() * s.y;
    return V2<R>{r * M<R>::cos(th), r * M<R>::sin(th)};
}
// basic.h:103-120
template <class R> inline V2<R> concentric_sample_disk(V2<R> s) {
    R ox = R(2) * s.x - R(1), oy = R(2) * s.y - R(1);
    if (ox == R(0) && oy == R(0)) return V2<R>{R(0), R(0)};
    R theta, r;
    const R pi_4 = (R)0.785398163397448309616, pi_2 = (R)1.57079632679489661923;
    if (M<R>::abs(ox) > M<R>::abs(oy)) { r = ox; theta = pi_4 * (oy / ox); }
    else { r = oy; theta = pi_2 - pi_4 * (ox / oy); }
    return V2<R>{r * M<R>::cos(theta), r * M<R>::sin(theta)};
}
// basic.h:127-137
template <class R> inline V3<R> cosine_sample_hemisphere(V2<R> s) {
    V2<R> d = concentric_sample_disk(s);
    R z = M<R>::sqrt(M<R>::fmax(R(0), R(1) - d.x * d.x - d.y * d.y));
    return V3<R>(d.x, d.y, z);
}
template <class R> inline R cosine_hemisphere_pdf(R c) { return c * C<R>::inv_pi(); }
// basic.h:83-89
template <class R> inline V3<R> uniform_sample_sphere(V2<R> s) {
    R z = R(1) - R(2) * s.x;
    R r = M<R>::sqrt(R(1) - z * z);
    R phi = R(2) * C<R>::pi() * s.y;
    return V3<R>(r * M<R>::cos(phi), r * M<R>::sin(phi), z);
}
// basic.h:139-143
template <class R> inline V2<R> uniform_sample_triangle(V2<R> s) {
    R sx = M<R>::sqrt(s.x);
    return V2<R>{R(1) - sx, s.y * sx};
}
// basic.h:154-160
template <class R> inline R schlick_fresnel(R cos_theta, R eta = R(1)) {
    R r0 = (eta - R(1)) / (eta + R(1));
    r0 = r0 * r0;
    R m = clampf(R(1) - cos_theta, R(0), R(1));
    return r0 + (R(1) - r0) * m * m * m * m * m;
}
template <class R> inline R fresnel_mix(R metallic, R eta, R c) {
    return lerp(schlick_fresnel(c, eta), schlick_fresnel(c), metallic);
}
template <class R> inline R luminance(V3<R> c) { return dot(c, V3<R>((R)0.2126f, (R)0.7152f, (R)0.0722f)); }
// basic.h:172-179
template <class R> inline R GTR1(R n_h, R a) {
    if (a >= R(1)) return C<R>::inv_pi();
    R a2 = a * a;
    R t = R(1) + (a2 - R(1)) * n_h * n_h;
    return (a2 - R(1)) * C<R>::inv_pi() / (M<R>::log(a2) * t);
}
// basic.h:181-186
template <class R> inline R GTR2(R n_h, R a) {
    R a2 = a * a;
    R t = R(1) + (a2 - R(1)) * n_h * n_h;
    return a2 * C<R>::inv_pi() / (t * t);
}
// basic.h:188-194
template <class R> inline R GTR2_aniso(R n_h, R h_x, R h_y, R ax, R ay) {
    R a = h_x / ax, b = h_y / ay;
    R t = a * a + b * b + n_h * n_h;
    return C<R>::inv_pi() / (ax * ay * t * t);
}
// basic.h:196-201
template <class R> inline R smithG_GGX(R n_v, R alphaG) {
    R a = alphaG * alphaG;
    R b = n_v * n_v;
    return (R(2) * n_v) / (n_v + M<R>::sqrt(a + b - a * b));
}
// basic.h:203-212
template <class R> inline V3<R> sample_GTR1(R a, V2<R> s) {
    if (a >= R(1)) return cosine_sample_hemisphere(s);
    R a2 = a * a;
    R phi = R(2) * C<R>::pi() * s.x;
    R cos_theta = M<R>::sqrt((R(1) - M<R>::pow(a2, R(1) - s.y)) / (R(1) - a2));
    R sin_theta = clampf(M<R>::sqrt(R(1) - (cos_theta * cos_theta)), R(0), R(1));
    return V3<R>(M<R>::cos(phi) * sin_theta, M<R>::sin(phi) * sin_theta, cos_theta);
}
// basic.h:223-230
template <class R> inline V3<R> sample_GTR2_aniso(R ax, R ay, V2<R> s) {
    R phi = R(2) * C<R>::pi() * s.x;
    R sin_phi = ay * M<R>::sin(phi);
    R cos_phi = ax * M<R>::cos(phi);
    R tan_theta = M<R>::sqrt(s.y / (R(1) - s.y));
    return normalize(V3<R>(tan_theta * cos_phi, tan_theta * sin_phi, R(1)));
}
// basic.h:232-250 (rsqrtf replaced by 1/sqrt, D4)
template <class R> inline V3<R> sample_GGXVNDF(V3<R> V, R a, V2<R> s) {
    V3<R> Vh = normalize(V3<R>(a * V.x, a * V.y, V.z));
    R lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    V3<R> T1 = (lensq > R(0)) ? V3<R>(-Vh.y, Vh.x, R(0)) * (R(1) / M<R>::sqrt(lensq)) : V3<R>(R(1), R(0), R(0));
    V3<R> T2 = cross(Vh, T1);
    R r = M<R>::sqrt(s.x);
    R phi = R(2) * C<R>::pi() * s.y;
    R t1 = r * M<R>::cos(phi);
    R t2 = r * M<R>::sin(phi);
    R ss = R(0.5) * (R(1) + Vh.z);
    t2 = (R(1) - ss) * M<R>::sqrt(R(1) - t1 * t1) + ss * t2;
    V3<R> Nh = t1 * T1 + t2 * T2 + M<R>::sqrt(M<R>::fmax(R(0), R(1) - t1 * t1 - t2 * t2)) * Vh;
    return normalize(V3<R>(a * Nh.x, a * Nh.y, M<R>::fmax(R(0), Nh.z)));
}

// ---------------------------------------------------------------------------
// Onb (onb.h:5-40)
// ---------------------------------------------------------------------------
template <class R> struct Onb {
    V3<R> z, x, y;
    Onb() {}
    Onb(V3<R> n, V3<R> t) : z(n) { x = normalize(t - z * dot(t, z)); y = cross(z, x); }
    // onb.h:31: from three given axes (Transform::apply_onb, transform.h:72-75)
    Onb(V3<R> n, V3<R> t, V3<R> b) : z(n), x(t), y(b) {}
    V3<R> to_world(V3<R> p) const { return p.x * x + p.y * y + p.z * z; }
    V3<R> to_local(V3<R> p) const { return V3<R>(dot(p, x), dot(p, y), dot(p, z)); }
};

// ---------------------------------------------------------------------------
// Material: 80-byte record shared with the test harness.
// ---------------------------------------------------------------------------
struct MaterialRec {
    int32_t type;   // 0 Diffuse, 1 DiffuseTransmission, 2 Dielectric, 3 Disney
    float ior, metallic, subsurface, roughness, specular, specularTint, anisotropic,
        sheen, sheenTint, clearcoat, clearcoatGloss, specTrans;
    float base_color[3];
    float emission_color[3];
    float intensity;
};
static_assert(sizeof(MaterialRec) == 80, "material record is 80 bytes");

enum { T_DIFFUSE = 0, T_DIFFTRANS = 1, T_DIELECTRIC = 2, T_DISNEY = 3 };

template <class R> struct BxDFSample { V3<R> f; V3<R> wi; R pdf = R(0); bool delta = false; };

// bxdf.h:59-387
template <class R> struct BxDF {
    int type;
    R ior, metallic, subsurface, roughness, specular, specularTint, anisotropic, sheen, sheenTint,
        clearcoat, clearcoatGloss, specTrans;
    explicit BxDF(const MaterialRec& m)
        : type(m.type), ior(m.ior), metallic(m.metallic), subsurface(m.subsurface), roughness(m.roughness),
          specular(m.specular), specularTint(m.specularTint), anisotropic(m.anisotropic), sheen(m.sheen),
          sheenTint(m.sheenTint), clearcoat(m.clearcoat), clearcoatGloss(m.clearcoatGloss), specTrans(m.specTrans) {}

    // bxdf.h:79-180
    V3<R> eval(V3<R> wi, V3<R> wo, V3<R> color, bool inner) const {
        switch (type) {
        case T_DIFFUSE: return (wi.z <= R(0)) ? V3<R>(R(0)) : color * C<R>::inv_pi();
        case T_DIFFTRANS: return color * C<R>::inv_pi();
        case T_DIELECTRIC: return V3<R>(R(0));
        case T_DISNEY: {
            R eta = inner ? ior : R(1) / ior;
            V3<R> V = wo, L = wi;
            V3<R> H = (L.z > R(0)) ? normalize(L + V) : normalize(L + V * eta);
            if (H.z <= R(0)) H = -H;
            R V_H = dot(V, H), L_H = dot(L, H);
            R lum = luminance(color);
            V3<R> ctint = (lum > R(0)) ? color / lum : V3<R>(R(1));
            V3<R> spec_color = lerp((specular * (R)0.08f) * lerp(V3<R>(R(1)), ctint, specularTint), color, metallic);
            V3<R> sheen_color = lerp(V3<R>(R(1)), ctint, sheenTint);
            R FM = fresnel_mix(metallic, eta, V_H);
            R diffuse_w = lum * (R(1) - metallic) * (R(1) - specTrans);
            R spec_reflect_w = luminance(lerp(spec_color, V3<R>(R(1)), FM));
            R spec_refract_w = (R(1) - FM) * (R(1) - metallic) * specTrans * lum;
            R clearcoat_w = clearcoat * (R(1) - metallic);

            V3<R> f_diffuse(R(0));
            if (diffuse_w > R(0) && L.z > R(0)) {
                R FL = schlick_fresnel(L.z), FV = schlick_fresnel(V.z);
                R FH = schlick_fresnel(L_H);
                R Fd90 = R(0.5) + R(2) * L_H * L_H * roughness;
                R Fd = lerp(R(1), Fd90, FL) * lerp(R(1), Fd90, FV);
                R Fss90 = L_H * L_H * roughness;
                R Fss = lerp(R(1), Fss90, FL) * lerp(R(1), Fss90, FV);
                R ss = (R)1.25f * (Fss * (R(1) / (L.z + V.z) - R(0.5)) + R(0.5));
                V3<R> Fsheen = (FH * sheen) * sheen_color;
                f_diffuse = ((C<R>::inv_pi() * lerp(Fd, ss, subsurface)) * color + Fsheen) * (R(1) - metallic) * (R(1) - specTrans);
            }
            V3<R> f_spec_reflect(R(0));
            if (spec_reflect_w > R(0) && L.z > R(0) && V.z > R(0)) {
                R _FM = fresnel_mix(metallic, eta, L_H);
                V3<R> F = lerp(spec_color, V3<R>(R(1)), _FM);
                R aspect = M<R>::sqrt(R(1) - anisotropic * (R)0.9f);
                R a2 = roughness * roughness;
                R ax = M<R>::fmax((R)0.001f, a2 / aspect), ay = M<R>::fmax((R)0.001f, a2 * aspect);
                R D = GTR2_aniso(H.z, H.x, H.y, ax, ay);
                R G = smithG_GGX(M<R>::abs(L.z), roughness) * smithG_GGX(M<R>::abs(V.z), roughness);
                f_spec_reflect = F * D * G / (R(4) * L.z * V.z);
            }
            V3<R> f_spec_refract(R(0));
            if (spec_refract_w > R(0) && L.z < R(0)) {
                R F = schlick_fresnel(M<R>::abs(V_H), eta);
                R D = GTR2(H.z, roughness);
                R denom = (L_H + V_H * eta) * (L_H + V_H * eta);
                R G = smithG_GGX(M<R>::abs(L.z), roughness) * smithG_GGX(M<R>::abs(V.z), roughness) *
                      M<R>::abs(L_H) * M<R>::abs(V_H) * eta * eta / denom;
                V3<R> refract_color(M<R>::sqrt(color.x), M<R>::sqrt(color.y), M<R>::sqrt(color.z));
                f_spec_refract = refract_color * (R(1) - metallic) * specTrans * (R(1) - F) * D * G /
                                 (M<R>::abs(L.z) * M<R>::abs(V.z));
            }
            V3<R> f_clearcoat(R(0));
            if (clearcoat_w > R(0) && L.z > R(0) && V.z > R(0)) {
                R FH = schlick_fresnel(L_H, R(1) / (R)1.5f);
                R F = lerp((R)0.04f, R(1), FH);
                R D = GTR1(H.z, clearcoatGloss);
                R G = smithG_GGX(L.z, (R)0.25f) * smithG_GGX(V.z, (R)0.25f);
                f_clearcoat = V3<R>((R)0.25f) * clearcoat * F * D * G / (R(4) * L.z * V.z);
            }
            return f_diffuse + f_spec_reflect + f_spec_refract + f_clearcoat;
        }
        default: return V3<R>(R(0));
        }
    }

    // Disney lobe weights (bxdf.h:242-252 / 353-363), shared by sample and pdf.
    void lobe_weights(V3<R> V, V3<R> color, R eta, R& dw, R& srw, R& stw, R& cw, R& ax, R& ay) const {
        R lum = luminance(color);
        V3<R> ctint = (lum > R(0)) ? color / lum : V3<R>(R(1));
        V3<R> spec_color = lerp((specular * (R)0.08f) * lerp(V3<R>(R(1)), ctint, specularTint), color, metallic);
        R aspect = M<R>::sqrt(R(1) - anisotropic * (R)0.9f);
        R a2 = roughness * roughness;
        ax = M<R>::fmax((R)0.001f, a2 / aspect);
        ay = M<R>::fmax((R)0.001f, a2 * aspect);
        R FM = fresnel_mix(metallic, eta, V.z);
        dw = lum * (R(1) - metallic) * (R(1) - specTrans);
        srw = luminance(lerp(spec_color, V3<R>(R(1)), FM));
        stw = (R(1) - FM) * (R(1) - metallic) * specTrans * lum;
        cw = clearcoat * (R(1) - metallic);
        R total = dw + srw + stw + cw;
        dw /= total; srw /= total; stw /= total; cw /= total;
    }

    // mixture pdf of the Disney lobes (bxdf.h:289-310 and 365-380)
    R disney_pdf(V3<R> L, V3<R> V, R eta, R dw, R srw, R stw, R cw, R ax, R ay) const {
        V3<R> H_reflect = normalize(L + V);
        if (H_reflect.z < R(0)) H_reflect = -H_reflect;
        V3<R> H_refract = normalize(L + V * eta);
        if (H_refract.z < R(0)) H_refract = -H_refract;
        R pdf = R(0);
        if (dw > R(0) && L.z > R(0)) pdf += dw * cosine_hemisphere_pdf(L.z);
        if (srw > R(0) && dot(V, H_reflect) > R(0))
            pdf += srw * GTR2_aniso(H_reflect.z, H_reflect.x, H_reflect.y, ax, ay) * H_reflect.z / (R(4) * dot(V, H_reflect));
        if (stw > R(0) && dot(V, H_refract) > R(0) && dot(L, H_refract) < R(0)) {
            R V_H = dot(V, H_refract), L_H = dot(L, H_refract);
            R denom = (L_H + V_H * eta);
            pdf += stw * GTR2(H_refract.z, roughness) * smithG_GGX(V.z, roughness) * M<R>::fmax(dot(V, H_refract), R(0)) /
                   V.z * M<R>::abs(L_H) / (denom * denom);
        }
        if (cw > R(0) && dot(V, H_reflect) > R(0))
            pdf += cw * GTR1(H_reflect.z, clearcoatGloss) * H_reflect.z / (R(4) * dot(V, H_reflect));
        return pdf;
    }

    // bxdf.h:182-319
    BxDFSample<R> sample(V3<R> wo, V2<R> rnd, V3<R> color, bool inner) const {
        switch (type) {
        case T_DIFFUSE: {
            V3<R> wi = cosine_sample_hemisphere(rnd);
            R pdf = cosine_hemisphere_pdf(wi.z);
            return {eval(wi, wo, color, inner), wi, pdf, false};
        }
        case T_DIFFTRANS: {
            if (rnd.x < R(0.5)) {
                rnd.x *= R(2);
                V3<R> wi = cosine_sample_hemisphere(rnd);
                R pdf = cosine_hemisphere_pdf(wi.z);
                return {eval(wi, wo, color, inner), wi, R(0.5) * pdf, false};
            } else {
                rnd.x = (rnd.x - R(0.5)) * R(2);
                V3<R> wi = cosine_sample_hemisphere(rnd);
                R pdf = cosine_hemisphere_pdf(wi.z);
                return {eval(-wi, wo, color, inner), -wi, R(0.5) * pdf, false};
            }
        }
        case T_DIELECTRIC: {
            R eta = inner ? ior : R(1) / ior;
            V3<R> n(R(0), R(0), R(1));
            V3<R> wi = refract(-wo, n, eta);
            R reflect_ratio = (dot(wi, wi) == R(0)) ? R(1) : schlick_fresnel(wo.z, eta);
            if (rnd.x <= reflect_ratio) wi = reflect(-wo, n);
            wi = normalize(wi);
            return {color, wi, R(1), true};
        }
        case T_DISNEY: {
            R eta = inner ? ior : R(1) / ior;
            V3<R> V = wo;
            R dw, srw, stw, cw, ax, ay;
            lobe_weights(V, color, eta, dw, srw, stw, cw, ax, ay);
            V3<R> L(R(0));
            if (rnd.x < dw) {
                rnd.x /= dw;
                L = cosine_sample_hemisphere(rnd);
            } else if (rnd.x < dw + srw) {
                rnd.x = (rnd.x - dw) / srw;
                V3<R> H = sample_GTR2_aniso(ax, ay, rnd);
                if (dot(V, H) <= R(0)) return BxDFSample<R>();
                L = normalize(reflect(-V, H));
            } else if (rnd.x < dw + srw + stw) {
                rnd.x = (rnd.x - dw - srw) / stw;
                V3<R> H = sample_GGXVNDF(V, roughness, rnd);
                L = refract(-V, H, eta);
                if (dot(V, H) <= R(0) || dot(L, L) == R(0)) return BxDFSample<R>();
                L = normalize(L);
            } else {
                rnd.x = (rnd.x - dw - srw - stw) / cw;
                V3<R> H = sample_GTR1(clearcoatGloss, rnd);
                if (dot(V, H) <= R(0)) return BxDFSample<R>();
                L = normalize(reflect(-V, H));
            }
            R pdf = disney_pdf(L, V, eta, dw, srw, stw, cw, ax, ay);
            return {eval(L, wo, color, inner), L, pdf, false};
        }
        default: return BxDFSample<R>();
        }
    }

    // bxdf.h:321-387
    R sample_pdf(V3<R> wi, V3<R> wo, V3<R> color, bool inner) const {
        switch (type) {
        case T_DIFFUSE: return (wi.z <= R(0)) ? R(0) : cosine_hemisphere_pdf(wi.z);
        case T_DIFFTRANS: return R(0.5) * cosine_hemisphere_pdf(M<R>::abs(wi.z));
        case T_DIELECTRIC: return R(0);
        case T_DISNEY: {
            R eta = inner ? ior : R(1) / ior;
            R dw, srw, stw, cw, ax, ay;
            lobe_weights(wo, color, eta, dw, srw, stw, cw, ax, ay);
            return disney_pdf(wi, wo, eta, dw, srw, stw, cw, ax, ay);
        }
        default: return R(0);
        }
    }
};

// gmaterial.h:20-31
inline bool mat_is_emissive(const MaterialRec& m) { return m.intensity > 0.0f; }
inline bool mat_is_specular(const MaterialRec& m) {
    return (m.type == T_DIELECTRIC) ||
           (m.type == T_DISNEY && m.metallic >= 0.99f && m.roughness <= 0.01f) ||
           (m.type == T_DISNEY && m.specTrans >= 0.99f && m.roughness <= 0.01f && m.clearcoatGloss <= 0.01f);
}
inline bool mat_is_transmissive(const MaterialRec& m) {
    return (m.type == T_DIELECTRIC) || (m.type == T_DIFFTRANS) ||
           (m.type == T_DISNEY && m.metallic < 1.0f && m.specTrans > 0.0f);
}

// gmaterial.h:59-104: local frame, inner flag when wo is below the shading normal
template <class R> inline V3<R> mat_eval(const BxDF<R>& b, V3<R> wi, V3<R> wo, const Onb<R>& onb, V3<R> color) {
    V3<R> wil = onb.to_local(wi), wol = onb.to_local(wo);
    bool inner = false;
    if (wol.z <= R(0)) { wol = -wol; wil = -wil; inner = true; }
    return b.eval(wil, wol, color, inner);
}
template <class R> inline BxDFSample<R> mat_sample(const BxDF<R>& b, V3<R> wo, V2<R> rnd, const Onb<R>& onb, V3<R> color) {
    V3<R> wol = onb.to_local(wo);
    bool inner = false;
    if (wol.z <= R(0)) { wol = -wol; inner = true; }
    BxDFSample<R> s = b.sample(wol, rnd, color, inner);
    if (inner) s.wi = -s.wi;
    s.wi = onb.to_world(s.wi);
    return s;
}
template <class R> inline R mat_pdf(const BxDF<R>& b, V3<R> wi, V3<R> wo, const Onb<R>& onb, V3<R> color) {
    V3<R> wil = onb.to_local(wi), wol = onb.to_local(wo);
    bool inner = false;
    if (wol.z <= R(0)) { wol = -wol; wil = -wil; inner = true; }
    return b.sample_pdf(wil, wol, color, inner);
}

// ---------------------------------------------------------------------------
// Scene: triangle meshes.  Without instances every mesh is one identity
// instance in world space.  With instances (scene.cu:225-288, the IAS of
// optix_ray_tracer.cu:337-359) the meshes are in object space and instance i
// places mesh inst_mesh[i] with the 3x4 row-major transform inst_xf[12 i..]
// (Transform::apply_point / apply_vector, transform.h:37-52); its inverse
// inst_inv is computed in double from the 3x4 (cofactors / determinant,
// translation -A^-1 t) and rounded to float, as the product library does.
// ---------------------------------------------------------------------------
struct Scene {
    int nv = 0, nf = 0, nmesh = 0, nmat = 0;
    std::vector<float> pos, nrm, tan;          // 3 * nv (nrm/tan may be empty)
    std::vector<uint32_t> idx;                 // 3 * nf
    std::vector<int> mesh_first;               // nmesh + 1
    std::vector<int> mesh_mat;                 // nmesh
    std::vector<int> tri_mesh;                 // nf
    std::vector<MaterialRec> mats;
    float env[3] = {0, 0, 0};
    // lights (scene.cu:240-287, mesh.cpp:3-18)
    std::vector<float> mesh_area;              // nmesh
    std::vector<float> area_cdf;               // nf (per mesh segment)
    std::vector<int> mesh_light;               // nmesh, -1 if not emissive
    std::vector<int> light_mesh;               // nlights
    std::vector<float> light_cdf;              // nlights
    float light_weight_sum = 0.0f;
    // instances (empty: one identity instance per mesh, world-space meshes)
    std::vector<int> inst_mesh;                // ninst
    std::vector<float> inst_xf, inst_inv;      // 12 per instance
    std::vector<int> light_inst;               // nlights: the light's instance (instanced scenes)
    int nlights() const { return (int)light_mesh.size(); }
    bool instanced() const { return !inst_mesh.empty(); }
    int ninst() const { return (int)inst_mesh.size(); }
};

// 3x4 affine inverse: double cofactors / determinant, translation -A^-1 t,
// rounded to float.  Returns false for a singular transform.
inline bool affine_inverse(const float m[12], float out[12]) {
    const double a = m[0], b = m[1], c = m[2], d = m[4], e = m[5], f = m[6], g = m[8], h = m[9], k = m[10];
    const double A = e * k - f * h, B = f * g - d * k, Cc = d * h - e * g;
    const double det = a * A + b * B + c * Cc;
    if (!(det != 0.0) || !std::isfinite(det)) return false;
    const double inv[9] = {A / det, (c * h - b * k) / det, (b * f - c * e) / det,
                           B / det, (a * k - c * g) / det, (c * d - a * f) / det,
                           Cc / det, (b * g - a * h) / det, (a * e - b * d) / det};
    const double t[3] = {m[3], m[7], m[11]};
    for (int r = 0; r < 3; r++) {
        out[4 * r + 0] = (float)inv[3 * r + 0];
        out[4 * r + 1] = (float)inv[3 * r + 1];
        out[4 * r + 2] = (float)inv[3 * r + 2];
        out[4 * r + 3] = (float)(-(inv[3 * r + 0] * t[0] + inv[3 * r + 1] * t[1] + inv[3 * r + 2] * t[2]));
    }
    return true;
}
// transform.h:37-52, in float in the reference's operation order
inline void xf_point(const float* m, const float p[3], float out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[4 * r] * p[0] + m[4 * r + 1] * p[1] + m[4 * r + 2] * p[2] + m[4 * r + 3];
}
inline void xf_vector(const float* m, const float v[3], float out[3]) {
    for (int r = 0; r < 3; r++) out[r] = m[4 * r] * v[0] + m[4 * r + 1] * v[1] + m[4 * r + 2] * v[2];
}
template <class R> inline V3<R> xf_point(const float* m, V3<R> p) {
    return V3<R>((R)m[0] * p.x + (R)m[1] * p.y + (R)m[2] * p.z + (R)m[3],
                 (R)m[4] * p.x + (R)m[5] * p.y + (R)m[6] * p.z + (R)m[7],
                 (R)m[8] * p.x + (R)m[9] * p.y + (R)m[10] * p.z + (R)m[11]);
}
template <class R> inline V3<R> xf_vector(const float* m, V3<R> v) {
    return V3<R>((R)m[0] * v.x + (R)m[1] * v.y + (R)m[2] * v.z, (R)m[4] * v.x + (R)m[5] * v.y + (R)m[6] * v.z,
                 (R)m[8] * v.x + (R)m[9] * v.y + (R)m[10] * v.z);
}

// mesh.cpp:3-18 and scene.cu:240-264 (light weight = intensity * mesh area)
inline void scene_finalize(Scene& s) {
    s.mesh_area.assign(s.nmesh, 0.0f);
    s.area_cdf.assign(s.nf, 0.0f);
    s.tri_mesh.assign(s.nf, 0);
    for (int m = 0; m < s.nmesh; m++) {
        float area = 0.0f;
        for (int f = s.mesh_first[m]; f < s.mesh_first[m + 1]; f++) {
            s.tri_mesh[f] = m;
            V3<float> v0 = cvt<float>(&s.pos[3 * s.idx[3 * f + 0]]);
            V3<float> v1 = cvt<float>(&s.pos[3 * s.idx[3 * f + 1]]);
            V3<float> v2 = cvt<float>(&s.pos[3 * s.idx[3 * f + 2]]);
            float A = 0.5f * length(cross(v1 - v0, v2 - v0));
            area += A;
            s.area_cdf[f] = area;
        }
        for (int f = s.mesh_first[m]; f < s.mesh_first[m + 1]; f++) s.area_cdf[f] /= area;
        s.mesh_area[m] = area;
    }
    s.mesh_light.assign(s.nmesh, -1);
    s.light_mesh.clear();
    s.light_inst.clear();
    s.light_cdf.clear();
    float wsum = 0.0f;
    // lights: emissive instances (scene.cu:240-264), weight = intensity x the
    // mesh's object-space area; without instances, emissive meshes
    const int nunits = s.instanced() ? s.ninst() : s.nmesh;
    for (int u = 0; u < nunits; u++) {
        const int m = s.instanced() ? s.inst_mesh[u] : u;
        const MaterialRec& mat = s.mats[s.mesh_mat[m]];
        if (mat.intensity <= 0.0f) continue;
        if (!s.instanced()) s.mesh_light[m] = (int)s.light_mesh.size();
        s.light_mesh.push_back(m);
        if (s.instanced()) s.light_inst.push_back(u);
        float w = mat.intensity * s.mesh_area[m];
        wsum += w;
        s.light_cdf.push_back(wsum);
    }
    for (auto& c : s.light_cdf) c /= wsum;
    s.light_weight_sum = wsum;
}

// ---------------------------------------------------------------------------
// Camera (camera.h:107-117, 165-170; transform.h:136-145; renderer.cu:125-140).
// The basis is computed once in float on the host, as the reference does.
// ---------------------------------------------------------------------------
struct Camera {
    float pos[3], X[3], Y[3], Z[3];
    float nw, nh;
};
inline Camera make_camera(const float pos[3], const float target[3], const float up[3], float fov_deg, int W, int H) {
    using F = float;
    V3<F> p = cvt<F>(pos), t = cvt<F>(target), u = cvt<F>(up);
    // LookAt columns
    V3<F> z = normalize(t - p);
    V3<F> x = normalize(cross(u, z));
    V3<F> y = cross(z, x);
    // CameraController(Transform, r) re-normalises each column (camera.h:45-48)
    x = normalize(x); y = normalize(y); z = normalize(z);
    Camera c;
    c.pos[0] = p.x; c.pos[1] = p.y; c.pos[2] = p.z;
    c.X[0] = x.x; c.X[1] = x.y; c.X[2] = x.z;
    c.Y[0] = y.x; c.Y[1] = y.y; c.Y[2] = y.z;
    c.Z[0] = z.x; c.Z[1] = z.y; c.Z[2] = z.z;
    float aspect = (float)W / (float)H;
    float rad = (fov_deg * 0.5f) * 3.14159265358979323846f / 180.0f;   // Radians (basic.h:16-19)
    c.nh = 2.0f * std::tan(rad);
    c.nw = c.nh * aspect;
    return c;
}

// ---------------------------------------------------------------------------
// Ray / triangle intersection (D3): watertight test, Woop, Benthin and Wald,
// "Watertight Ray/Triangle Intersection", JCGT 2(1), 2013.
// ---------------------------------------------------------------------------
struct RayPre {
    int kx, ky, kz;
    float Sx, Sy, Sz;
    float o[3];
};
inline RayPre ray_pre(const float o[3], const float d[3]) {
    RayPre r;
    float ax = std::fabs(d[0]), ay = std::fabs(d[1]), az = std::fabs(d[2]);
    int kz = (ax > ay) ? ((ax > az) ? 0 : 2) : ((ay > az) ? 1 : 2);
    int kx = kz + 1; if (kx == 3) kx = 0;
    int ky = kx + 1; if (ky == 3) ky = 0;
    if (d[kz] < 0.0f) std::swap(kx, ky);
    r.kx = kx; r.ky = ky; r.kz = kz;
    r.Sx = d[kx] / d[kz];
    r.Sy = d[ky] / d[kz];
    r.Sz = 1.0f / d[kz];
    r.o[0] = o[0]; r.o[1] = o[1]; r.o[2] = o[2];
    return r;
}
// returns true with t, u (weight of v1), v (weight of v2) if t in [tmin, tmax]
inline bool intersect_tri(const RayPre& r, const float* v0, const float* v1, const float* v2, float tmin, float tmax,
                          float* t_out, float* u_out, float* v_out) {
    float A[3] = {v0[0] - r.o[0], v0[1] - r.o[1], v0[2] - r.o[2]};
    float B[3] = {v1[0] - r.o[0], v1[1] - r.o[1], v1[2] - r.o[2]};
    float Cc[3] = {v2[0] - r.o[0], v2[1] - r.o[1], v2[2] - r.o[2]};
    float Ax = A[r.kx] - r.Sx * A[r.kz];
    float Ay = A[r.ky] - r.Sy * A[r.kz];
    float Bx = B[r.kx] - r.Sx * B[r.kz];
    float By = B[r.ky] - r.Sy * B[r.kz];
    float Cx = Cc[r.kx] - r.Sx * Cc[r.kz];
    float Cy = Cc[r.ky] - r.Sy * Cc[r.kz];
    float U = Cx * By - Cy * Bx;
    float V = Ax * Cy - Ay * Cx;
    float W = Bx * Ay - By * Ax;
    if (U == 0.0f || V == 0.0f || W == 0.0f) {
        double CxBy = (double)Cx * (double)By, CyBx = (double)Cy * (double)Bx;
        U = (float)(CxBy - CyBx);
        double AxCy = (double)Ax * (double)Cy, AyCx = (double)Ay * (double)Cx;
        V = (float)(AxCy - AyCx);
        double BxAy = (double)Bx * (double)Ay, ByAx = (double)By * (double)Ax;
        W = (float)(BxAy - ByAx);
    }
    if ((U < 0.0f || V < 0.0f || W < 0.0f) && (U > 0.0f || V > 0.0f || W > 0.0f)) return false;
    float det = U + V + W;
    if (det == 0.0f) return false;
    float Az = r.Sz * A[r.kz];
    float Bz = r.Sz * B[r.kz];
    float Cz = r.Sz * Cc[r.kz];
    float T = U * Az + V * Bz + W * Cz;
    float rcp = 1.0f / det;
    float t = T * rcp;
    if (!(t >= tmin && t <= tmax)) return false;
    *t_out = t;
    *u_out = V * rcp;
    *v_out = W * rcp;
    return true;
}

// ---------------------------------------------------------------------------
// BVH: LBVH (Karras 2012) over 63-bit Morton codes, bottom-up bounds, SAH
// collapse to leaves of at most max_leaf triangles, and a binned-SAH rebuild
// of the top of the tree over a frontier of subtrees ("clusters").  Every
// choice is made with integer or exact min/max arithmetic, or with float
// formulas evaluated in a fixed order, so a GPU build that follows the same
// steps produces the same bytes.
// ---------------------------------------------------------------------------
struct BvhNode {             // 64 bytes
    float lx[2], ly[2];      // left child box: lo.x hi.x lo.y hi.y
    float rx[2], ry[2];      // right child box
    float lz[2], rz[2];      // left lo.z hi.z, right lo.z hi.z
    int32_t left, right;     // >= 0 internal node index; < 0 leaf reference
    int32_t pad0, pad1;
};
static_assert(sizeof(BvhNode) == 64, "node is 64 bytes");

inline int32_t leaf_ref(uint32_t first, uint32_t count) { return (int32_t)(0x80000000u | (first << 3) | (count - 1u)); }
inline uint32_t leaf_first(int32_t ref) { return ((uint32_t)ref & 0x7fffffffu) >> 3; }
inline uint32_t leaf_count(int32_t ref) { return ((uint32_t)ref & 7u) + 1u; }

// Total order on floats (-0 < +0) used for every min/max of the build.
inline uint32_t fkey(float f) { uint32_t u = f2u(f); return (u & 0x80000000u) ? ~u : (u | 0x80000000u); }
inline float kmin(float a, float b) { return (fkey(b) < fkey(a)) ? b : a; }
inline float kmax(float a, float b) { return (fkey(b) > fkey(a)) ? b : a; }

struct Box {
    float lo[3], hi[3];
    Box() { for (int i = 0; i < 3; i++) { lo[i] = INFINITY; hi[i] = -INFINITY; } }
    void grow(const Box& b) { for (int i = 0; i < 3; i++) { lo[i] = kmin(lo[i], b.lo[i]); hi[i] = kmax(hi[i], b.hi[i]); } }
    void grow(const float* p) { for (int i = 0; i < 3; i++) { lo[i] = kmin(lo[i], p[i]); hi[i] = kmax(hi[i], p[i]); } }
};
// surface area 2(dx dy + dy dz + dz dx), 0 for an empty box
inline float box_area(const Box& b) {
    float dx = b.hi[0] - b.lo[0], dy = b.hi[1] - b.lo[1], dz = b.hi[2] - b.lo[2];
    if (!(dx >= 0.0f)) return 0.0f;
    return 2.0f * ((dx * dy + dy * dz) + dz * dx);
}

struct BvhParams {
    int max_leaf = 4;          // <= 8
    int cluster_size = 0;      // 0: automatic = max(64, n / 1024)
    float c_trav = 1.0f, c_isect = 1.0f;
    int sah_bins = 16;
};

struct Bvh {
    std::vector<BvhNode> nodes;
    std::vector<uint32_t> order;        // sorted position -> primitive id
    std::vector<float> tris;            // 12 floats per sorted triangle (v0,prim | v1,mesh | v2,0)
    int32_t root = 0;                   // node index or leaf ref
    int max_depth = 0;
    int num_clusters = 0;
    int num_top_nodes = 0;
    int cluster_size = 0;
    std::vector<int32_t> karras_left, karras_right;   // refs: >=0 internal, <0 ~leaf
    std::vector<uint64_t> codes;        // sorted codes
};

inline uint64_t expand_bits21(uint32_t v) {
    uint64_t x = v & 0x1fffffu;
    x = (x | (x << 32)) & 0x1f00000000ffffULL;
    x = (x | (x << 16)) & 0x1f0000ff0000ffULL;
    x = (x | (x << 8)) & 0x100f00f00f00f00fULL;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ULL;
    x = (x | (x << 2)) & 0x1249249249249249ULL;
    return x;
}
// centroid -> 21-bit grid coordinate per axis
inline uint32_t morton_axis(float c, float lo, float ext) {
    float t = (ext > 0.0f) ? (c - lo) / ext : 0.0f;
    float q = t * 2097152.0f;
    uint32_t qi = (q >= 2097151.0f) ? 2097151u : (q > 0.0f ? (uint32_t)q : 0u);
    return qi;
}
inline uint64_t morton3(uint32_t x, uint32_t y, uint32_t z) {
    return (expand_bits21(x) << 2) | (expand_bits21(y) << 1) | expand_bits21(z);
}

inline int clz64(uint64_t v) { return v ? __builtin_clzll(v) : 64; }
inline int clz32(uint32_t v) { return v ? __builtin_clz(v) : 32; }
// Karras delta with index tie-break
inline int kdelta(const std::vector<uint64_t>& c, int n, int i, int j) {
    if (j < 0 || j >= n) return -1;
    uint64_t a = c[i], b = c[j];
    if (a != b) return clz64(a ^ b);
    return 64 + clz32((uint32_t)i ^ (uint32_t)j);
}

void bvh_build(const Scene& s, const BvhParams& prm, Bvh& out);
// instanced scenes: one BVH per mesh over the mesh's object-space triangles
// (sorted triangles carry the scene's primitive ids)
void blas_build(const Scene& s, const BvhParams& prm, std::vector<Bvh>& out);
void bvh_refit(const Scene& s, Bvh& bvh);   // same topology, boxes from current vertex positions

// closest hit over the BVH: returns sorted position (or -1); inst: the
// instance hit (instanced scenes; -1 otherwise)
struct Hit { float t, u, v; int32_t prim; int32_t pos; int32_t inst = -1; };
bool trace_closest(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax, Hit* hit, int* nodes_visited = nullptr, int* tris_tested = nullptr);
bool trace_any(const Bvh& bvh, const float o[3], const float d[3], float tmin, float tmax);
// brute force over the scene (reference semantics of D3)
bool brute_closest(const Scene& s, const float o[3], const float d[3], float tmin, float tmax, Hit* hit);
bool brute_any(const Scene& s, const float o[3], const float d[3], float tmin, float tmax);

// integrator parameters (path_tracer.h:34-38 defaults)
struct PtParams {
    int spp = 1, max_depth = 16, rr_depth = 4, use_nee = 1, use_mis = 1;
    uint32_t seed = 0;
};

}  // namespace orc
