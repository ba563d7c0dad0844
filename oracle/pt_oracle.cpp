// CPU oracle — TEST INFRASTRUCTURE (the parity reference and CPU baseline), never product code.
//
// Restatement of the reference GLSL path, function by function, in the same
// arithmetic order (left-to-right as the GLSL source groups it), templated on
// the scalar T. Citations are file:line in knightcrawler25/GLSL-PathTracer.
//
// Conventions where the GLSL spec leaves the choice to the GL driver (these are
// "parity unpinned at the GL-driver boundary"): normalize(v) = v / sqrt(dot(v,v));
// min(x,y) = y < x ? y : x, max(x,y) = x < y ? y : x (GLSL spec definitions);
// mat * v = sum of columns scaled by components, in column order; inverse() by
// cofactor expansion; texture() = GL_LINEAR bilinear with GL_REPEAT per the GL
// spec formula, full-precision weights; fragment centre = pixel + 0.5; the
// triangle test's uvt.xyz / det = uvt.xyz * (1 / det), the reciprocal correctly
// rounded (GLSL allows 2.5 ulp for a division; this is within 1.5).
//
// What pins each part (tests/test_oracle_*.py):
//   pcg4d / rand / seeds          numpy uint32 re-implementation, known answers
//   SphereIntersect, RectIntersect, AABBIntersect   closed forms
//   ClosestHit / AnyHit           FP64 brute force over every triangle and light
//   Disney lobes, sampling        closed forms (Lambert, GGX conductor, Walter dielectric),
//                                 sample-vs-pdf histogram agreement, pdf integral,
//                                 reciprocity of reflection lobes
//   Light sampling                rect/sphere points on the light from the independent RNG,
//                                 E[1/pdf] = solid angle (rect) or hemisphere quadrature
//                                 (sphere); distant light draws nothing
//   PathTrace                     analytic single-bounce irradiance (diffuse plane under a
//                                 rect light), white furnace, FP64 instantiation agreement
//   EnvMap CDF / search / eval    texel = linear scan of the CDF for the same draw, corner
//                                 colour = mean of four texels, pdf integrates to 1,
//                                 rotation = azimuth shift, sample pdf = eval pdf
//   Raygen (CameraRay)            tent jitter (moments, KS), aperture disc, focal-plane meet
//   Media                         absorb / emissive closed forms across a slab, free-flight
//                                 fraction exp(-density * chord)
//   Alpha-tested AnyHit           float64 brute force with a block mask; blend comparison
//   (test_oracle_sampling.py pins the last six rows)
//   sin / cos (float)             double sin/cos within 1.6 ulp over [-64, 64], symmetry,
//                                 exact zero
// Parity unpinned: texture filtering weights, GLSL inverse(), the ULPs of the
// other transcendental functions (pow, exp, log, acos, atan2, tan).
#include "pt_oracle.h"

#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// ------------------------------------------------------------ vector types
template <class T> struct V2 { T x, y; };
template <class T> struct V3 {
    T x, y, z;
    V3() : x(0), y(0), z(0) {}
    V3(T a, T b, T c) : x(a), y(b), z(c) {}
    explicit V3(T a) : x(a), y(a), z(a) {}
};
template <class T> struct V4 { T x, y, z, w; };

template <class T> inline V3<T> operator+(V3<T> a, V3<T> b) { return V3<T>(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> inline V3<T> operator-(V3<T> a, V3<T> b) { return V3<T>(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> inline V3<T> operator*(V3<T> a, V3<T> b) { return V3<T>(a.x * b.x, a.y * b.y, a.z * b.z); }
template <class T> inline V3<T> operator/(V3<T> a, V3<T> b) { return V3<T>(a.x / b.x, a.y / b.y, a.z / b.z); }
template <class T> inline V3<T> operator*(V3<T> a, T s) { return V3<T>(a.x * s, a.y * s, a.z * s); }
template <class T> inline V3<T> operator*(T s, V3<T> a) { return V3<T>(s * a.x, s * a.y, s * a.z); }
template <class T> inline V3<T> operator/(V3<T> a, T s) { return V3<T>(a.x / s, a.y / s, a.z / s); }
template <class T> inline V3<T> operator-(V3<T> a) { return V3<T>(-a.x, -a.y, -a.z); }
template <class T> inline V2<T> operator*(V2<T> a, T s) { return V2<T>{a.x * s, a.y * s}; }
template <class T> inline V2<T> operator+(V2<T> a, V2<T> b) { return V2<T>{a.x + b.x, a.y + b.y}; }
template <class T> inline V2<T> operator-(V2<T> a, V2<T> b) { return V2<T>{a.x - b.x, a.y - b.y}; }

template <class T> inline T gmin(T a, T b) { return b < a ? b : a; }
template <class T> inline T gmax(T a, T b) { return a < b ? b : a; }
template <class T> inline T gclamp(T x, T lo, T hi) { return gmin(gmax(x, lo), hi); }
template <class T> inline T gmix(T x, T y, T a) { return x * (T(1) - a) + y * a; }
template <class T> inline V3<T> gmix(V3<T> x, V3<T> y, T a) { return x * (T(1) - a) + y * a; }
template <class T> inline V3<T> vmin(V3<T> a, V3<T> b) { return V3<T>(gmin(a.x, b.x), gmin(a.y, b.y), gmin(a.z, b.z)); }
template <class T> inline V3<T> vmax(V3<T> a, V3<T> b) { return V3<T>(gmax(a.x, b.x), gmax(a.y, b.y), gmax(a.z, b.z)); }
template <class T> inline T dot(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> inline V3<T> cross(V3<T> a, V3<T> b) {
    return V3<T>(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
template <class T> inline T length(V3<T> a) { return std::sqrt(dot(a, a)); }
template <class T> inline V3<T> normalize(V3<T> a) { return a / length(a); }
template <class T> inline V3<T> reflect(V3<T> I, V3<T> N) { return I - N * (T(2) * dot(N, I)); }
template <class T> inline V3<T> refract(V3<T> I, V3<T> N, T eta) {
    T d = dot(N, I);
    T k = T(1) - eta * eta * (T(1) - d * d);
    if (k < T(0)) return V3<T>(T(0));
    return eta * I - (eta * d + std::sqrt(k)) * N;
}
template <class T> inline V3<T> vexp(V3<T> a) { return V3<T>(std::exp(a.x), std::exp(a.y), std::exp(a.z)); }
template <class T> inline V3<T> vpow(V3<T> a, T p) { return V3<T>(std::pow(a.x, p), std::pow(a.y, p), std::pow(a.z, p)); }
// pow of the texture decode (GetMaterial, texture ^ 2.2): GLSL leaves pow's
// precision to the driver; the float path takes the value rounded once from
// double (correctly rounded but for vanishingly rare double roundings).
template <class T> inline V3<T> vpow_tex(V3<T> a, T p) { return vpow(a, p); }
template <> inline V3<float> vpow_tex<float>(V3<float> a, float p) {
    auto f = [p](float x) { return static_cast<float>(std::pow(static_cast<double>(x), static_cast<double>(p))); };
    return V3<float>(f(a.x), f(a.y), f(a.z));
}

// sin and cos of the path's angles (sampling azimuths, the aperture angle, the
// environment map's theta and phi; all within [-2pi, 2pi]). GLSL leaves their
// precision to the driver, so the float path fixes one sequence of float
// operations: k = floor(x * 2/pi + 0.5); r = x - k * pi/2 with pi/2 split in
// four floats (8, 12, 12 and 24 significant bits; the first three products
// exact for |k| < 2^12); Taylor series in r^2 to degree 11 (sin) and 12 (cos);
// quadrant k mod 4. The double path calls std::sin / std::cos.
// Pinned by tests/test_oracle_sampling.py against double sin/cos (<= 1.6 ulp).
inline void sincos_f(float x, float& sn, float& cs) {
    const float k = std::floor(x * 0x1.45f306p-1f + 0.5f);
    const float p1 = 0x1.92p+0f, p2 = 0x1.fb6p-12f, p3 = -0x1.778p-25f, p4 = 0x1.68c234p-39f;
    const float r = (((x - k * p1) - k * p2) - k * p3) - k * p4;
    const float t = r * r;
    // 1/n! rounded to float
    const float s3 = 0x1.555556p-3f, s5 = 0x1.111112p-7f, s7 = 0x1.a01a02p-13f, s9 = 0x1.71de3ap-19f,
                s11 = 0x1.ae6456p-26f;
    const float c4 = 0x1.555556p-5f, c6 = 0x1.6c16c2p-10f, c8 = 0x1.a01a02p-16f, c10 = 0x1.27e4fcp-22f,
                c12 = 0x1.1eed8ep-29f;
    const float ps = r + r * t * (-s3 + t * (s5 + t * (-s7 + t * (s9 - t * s11))));
    const float pc = (1.0f - 0.5f * t) + t * t * (c4 + t * (-c6 + t * (c8 + t * (-c10 + t * c12))));
    switch (static_cast<int>(k) & 3) {
        case 0: sn = ps; cs = pc; break;
        case 1: sn = pc; cs = -ps; break;
        case 2: sn = -ps; cs = -pc; break;
        default: sn = -pc; cs = ps; break;
    }
}
template <class T> inline T tsin(T x) { return std::sin(x); }
template <class T> inline T tcos(T x) { return std::cos(x); }
template <> inline float tsin<float>(float x) { float s, c; sincos_f(x, s, c); return s; }
template <> inline float tcos<float>(float x) { float s, c; sincos_f(x, s, c); return c; }

// 4x4 and 3x3 matrices stored as columns (GLSL / reference Mat4 convention).
template <class T> struct M4 { T m[4][4]; };
template <class T> struct M3 { T m[3][3]; };

template <class T> inline V3<T> mul_point(const M4<T>& a, V3<T> v, T w) {
    V3<T> c0(a.m[0][0], a.m[0][1], a.m[0][2]), c1(a.m[1][0], a.m[1][1], a.m[1][2]);
    V3<T> c2(a.m[2][0], a.m[2][1], a.m[2][2]), c3(a.m[3][0], a.m[3][1], a.m[3][2]);
    return ((c0 * v.x + c1 * v.y) + c2 * v.z) + c3 * w;
}
template <class T> inline V3<T> mul3(const M3<T>& a, V3<T> v) {
    V3<T> c0(a.m[0][0], a.m[0][1], a.m[0][2]), c1(a.m[1][0], a.m[1][1], a.m[1][2]);
    V3<T> c2(a.m[2][0], a.m[2][1], a.m[2][2]);
    return (c0 * v.x + c1 * v.y) + c2 * v.z;
}
template <class T> M3<T> transpose3(const M3<T>& a) {
    M3<T> r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = a.m[j][i];
    return r;
}
template <class T> M3<T> inverse3(const M3<T>& a) {
    const T (*m)[3] = a.m;
    T c00 = m[1][1] * m[2][2] - m[2][1] * m[1][2];
    T c01 = m[0][1] * m[2][2] - m[2][1] * m[0][2];
    T c02 = m[0][1] * m[1][2] - m[1][1] * m[0][2];
    T det = m[0][0] * c00 - m[1][0] * c01 + m[2][0] * c02;
    T inv = T(1) / det;
    M3<T> r;
    r.m[0][0] = c00 * inv;
    r.m[1][0] = -(m[1][0] * m[2][2] - m[2][0] * m[1][2]) * inv;
    r.m[2][0] = (m[1][0] * m[2][1] - m[2][0] * m[1][1]) * inv;
    r.m[0][1] = -c01 * inv;
    r.m[1][1] = (m[0][0] * m[2][2] - m[2][0] * m[0][2]) * inv;
    r.m[2][1] = -(m[0][0] * m[2][1] - m[2][0] * m[0][1]) * inv;
    r.m[0][2] = c02 * inv;
    r.m[1][2] = -(m[0][0] * m[1][2] - m[1][0] * m[0][2]) * inv;
    r.m[2][2] = (m[0][0] * m[1][1] - m[1][0] * m[0][1]) * inv;
    return r;
}
template <class T> M4<T> inverse4(const M4<T>& a) {
    // Cofactor expansion with 2x2 sub-determinants.
    const T (*m)[4] = a.m;
    T s0 = m[0][0] * m[1][1] - m[1][0] * m[0][1];
    T s1 = m[0][0] * m[1][2] - m[1][0] * m[0][2];
    T s2 = m[0][0] * m[1][3] - m[1][0] * m[0][3];
    T s3 = m[0][1] * m[1][2] - m[1][1] * m[0][2];
    T s4 = m[0][1] * m[1][3] - m[1][1] * m[0][3];
    T s5 = m[0][2] * m[1][3] - m[1][2] * m[0][3];
    T c5 = m[2][2] * m[3][3] - m[3][2] * m[2][3];
    T c4 = m[2][1] * m[3][3] - m[3][1] * m[2][3];
    T c3 = m[2][1] * m[3][2] - m[3][1] * m[2][2];
    T c2 = m[2][0] * m[3][3] - m[3][0] * m[2][3];
    T c1 = m[2][0] * m[3][2] - m[3][0] * m[2][2];
    T c0 = m[2][0] * m[3][1] - m[3][0] * m[2][1];
    T det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    T id = T(1) / det;
    M4<T> r;
    r.m[0][0] = (m[1][1] * c5 - m[1][2] * c4 + m[1][3] * c3) * id;
    r.m[0][1] = (-m[0][1] * c5 + m[0][2] * c4 - m[0][3] * c3) * id;
    r.m[0][2] = (m[3][1] * s5 - m[3][2] * s4 + m[3][3] * s3) * id;
    r.m[0][3] = (-m[2][1] * s5 + m[2][2] * s4 - m[2][3] * s3) * id;
    r.m[1][0] = (-m[1][0] * c5 + m[1][2] * c2 - m[1][3] * c1) * id;
    r.m[1][1] = (m[0][0] * c5 - m[0][2] * c2 + m[0][3] * c1) * id;
    r.m[1][2] = (-m[3][0] * s5 + m[3][2] * s2 - m[3][3] * s1) * id;
    r.m[1][3] = (m[2][0] * s5 - m[2][2] * s2 + m[2][3] * s1) * id;
    r.m[2][0] = (m[1][0] * c4 - m[1][1] * c2 + m[1][3] * c0) * id;
    r.m[2][1] = (-m[0][0] * c4 + m[0][1] * c2 - m[0][3] * c0) * id;
    r.m[2][2] = (m[3][0] * s4 - m[3][1] * s2 + m[3][3] * s0) * id;
    r.m[2][3] = (-m[2][0] * s4 + m[2][1] * s2 - m[2][3] * s0) * id;
    r.m[3][0] = (-m[1][0] * c3 + m[1][1] * c1 - m[1][2] * c0) * id;
    r.m[3][1] = (m[0][0] * c3 - m[0][1] * c1 + m[0][2] * c0) * id;
    r.m[3][2] = (-m[3][0] * s3 + m[3][1] * s1 - m[3][2] * s0) * id;
    r.m[3][3] = (m[2][0] * s3 - m[2][1] * s1 + m[2][2] * s0) * id;
    return r;
}

// ------------------------------------------------------------ constants
// globals.glsl:25-44; literals are float in GLSL, so the float instantiation
// uses the float-rounded values and the double one the double values.
template <class T> struct K {
    static constexpr T PI = T(3.14159265358979323);
    static constexpr T INV_PI = T(0.31830988618379067);
    static constexpr T TWO_PI = T(6.28318530717958648);
    static constexpr T INV_TWO_PI = T(0.15915494309189533);
    static constexpr T INV_4_PI = T(0.07957747154594766);
    static constexpr T EPS = T(0.0003);
    static constexpr T INF = T(1000000.0);
};
template <> struct K<float> {
    static constexpr float PI = 3.14159265358979323f;
    static constexpr float INV_PI = 0.31830988618379067f;
    static constexpr float TWO_PI = 6.28318530717958648f;
    static constexpr float INV_TWO_PI = 0.15915494309189533f;
    static constexpr float INV_4_PI = 0.07957747154594766f;
    static constexpr float EPS = 0.0003f;
    static constexpr float INF = 1000000.0f;
};

enum { QUAD_LIGHT = 0, SPHERE_LIGHT = 1, DISTANT_LIGHT = 2 };
enum { ALPHA_MODE_OPAQUE = 0, ALPHA_MODE_BLEND = 1, ALPHA_MODE_MASK = 2 };
enum { MEDIUM_NONE = 0, MEDIUM_ABSORB = 1, MEDIUM_SCATTER = 2, MEDIUM_EMISSIVE = 3 };

// ------------------------------------------------------------ RNG (globals.glsl:143-166)
inline void pcg4d(uint32_t v[4]) {
    for (int i = 0; i < 4; ++i) v[i] = v[i] * 1664525u + 1013904223u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
    for (int i = 0; i < 4; ++i) v[i] ^= v[i] >> 16u;
    v[0] += v[1] * v[3]; v[1] += v[2] * v[0]; v[2] += v[0] * v[1]; v[3] += v[1] * v[2];
}
// float(seed.x) / float(0xffffffffu): the divisor rounds to 2^32, result in [0, 1].
// The draw is always float, also in the double instantiation, so both follow
// the same random decisions.
inline float rand_f(uint32_t s[4]) {
    pcg4d(s);
    return static_cast<float>(s[0]) / static_cast<float>(0xffffffffu);
}

template <class T> inline T Luminance(V3<T> c) {
    return T(0.212671f) * c.x + T(0.715160f) * c.y + T(0.072169f) * c.z;
}
template <> inline double Luminance<double>(V3<double> c) {
    return 0.212671 * c.x + 0.715160 * c.y + 0.072169 * c.z;
}

// ------------------------------------------------------------ structs (globals.glsl:46-139)
template <class T> struct Ray { V3<T> origin, direction; };
template <class T> struct Medium { int type = 0; T density = 0; V3<T> color; T anisotropy = 0; };
template <class T> struct Material {
    V3<T> baseColor; T opacity = 0; int alphaMode = 0; T alphaCutoff = 0;
    V3<T> emission; T anisotropic = 0, metallic = 0, roughness = 0, subsurface = 0, specularTint = 0;
    T sheen = 0, sheenTint = 0, clearcoat = 0, clearcoatRoughness = 0, specTrans = 0, ior = 0, ax = 0, ay = 0;
    Medium<T> medium;
};
template <class T> struct Light { V3<T> position, emission, u, v; T radius, area, type; };
template <class T> struct State {
    int depth = 0; T eta = 0; T hitDist = 0;
    V3<T> fhp, normal, ffnormal, tangent, bitangent;
    bool isEmitter = false;
    V2<T> texCoord{0, 0};
    int matID = 0;
    Material<T> mat;
    Medium<T> medium;
    // frozen-sampling mode only: the material evaluated for f and the medium
    // whose factors weight the path (see Ctx::eval_mats)
    Material<T> mat_e;
    T eta_e = 0;
    Medium<T> medium_e;
    // instrumentation (not in the reference): which primitive was hit
    int hitInstance = -1, hitTriRow = -1, hitLight = -1;
    T baryU = 0, baryV = 0;
};
template <class T> struct ScatterSampleRec { V3<T> L, f; T pdf = 0; };
template <class T> struct LightSampleRec { V3<T> normal, emission, direction; T dist = 0, pdf = 0; };

struct Counters { int64_t rays_closest = 0, rays_shadow = 0, nodes = 0, tris = 0, instances = 0, lights = 0; };

// ------------------------------------------------------------ scene access
struct SceneData {
    const oracle_scene* s;
    const oracle_options* o;
    std::vector<float> tex_f;  // unused placeholder for alignment of future caches
};

template <class T> struct Ctx {
    const oracle_scene* s;
    const oracle_options* o;
    uint32_t seed[4];
    Counters* cnt;
    bool stack_overflow = false;
    // Frozen-sampling mode (test instrument for the adjoint, not in the
    // reference): directions, lobe choices, pdfs, MIS weights and Russian
    // roulette come from the scene's materials; the BSDF values f that weight
    // the path come from this second material table, and so do the medium
    // factors: absorption and emission with the second table's colour and
    // density, and in a scattering medium, whose free-flight distances s are
    // still drawn with the scene's density, the weight of a scatter event
    // colour_e * density_e exp(-density_e s) / (density exp(-density s)) and of
    // a segment of length d crossed without scattering
    // exp(-density_e d) / exp(-density d) (1 and colour when both tables agree).
    // A central difference of the image over eval_mats is then the derivative
    // the path-replay adjoint estimates, sample by sample.
    const float* eval_mats = nullptr;
    T rand() { return T(rand_f(seed)); }
};

template <class T> inline V3<T> fetch3(const float* p) { return V3<T>(T(p[0]), T(p[1]), T(p[2])); }

template <class T> Light<T> fetch_light(const oracle_scene* s, int i) {
    const float* p = s->lights + 15 * i;
    return Light<T>{fetch3<T>(p), fetch3<T>(p + 3), fetch3<T>(p + 6), fetch3<T>(p + 9), T(p[12]), T(p[13]), T(p[14])};
}

// GL_LINEAR + GL_REPEAT per the GL specification (8.14.2): texel centres at
// (i + 0.5) / size, weights from the fractional part.
template <class T> V4<T> bilinear(const float* img, int ch, int W, int H, T u, T v, bool rgba8, const uint8_t* img8) {
    T x = u * T(W) - T(0.5), y = v * T(H) - T(0.5);
    T fx = std::floor(x), fy = std::floor(y);
    T ax = x - fx, ay = y - fy;
    long long i0 = static_cast<long long>(fx), j0 = static_cast<long long>(fy);
    auto wrap = [](long long i, int n) { long long r = i % n; return static_cast<int>(r < 0 ? r + n : r); };
    int xs[2] = {wrap(i0, W), wrap(i0 + 1, W)}, ys[2] = {wrap(j0, H), wrap(j0 + 1, H)};
    T w[4] = {(T(1) - ax) * (T(1) - ay), ax * (T(1) - ay), (T(1) - ax) * ay, ax * ay};
    T acc[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k) {
        int xx = xs[k & 1], yy = ys[k >> 1];
        size_t base = (static_cast<size_t>(yy) * W + xx) * ch;
        for (int c = 0; c < ch; ++c) {
            T val = rgba8 ? T(img8[base + c]) / T(255) : T(img[base + c]);
            acc[c] = acc[c] + w[k] * val;
        }
    }
    if (ch < 4) acc[3] = T(1);
    return V4<T>{acc[0], acc[1], acc[2], acc[3]};
}

template <class T> V4<T> tex_array(const oracle_scene* s, V2<T> uv, T layer) {
    int l = static_cast<int>(std::floor(layer + T(0.5)));
    l = std::max(0, std::min(s->num_textures - 1, l));
    const uint8_t* base = s->textures + static_cast<size_t>(l) * 4 * s->tex_width * s->tex_height;
    return bilinear<T>(nullptr, 4, s->tex_width, s->tex_height, uv.x, uv.y, true, base);
}

// ------------------------------------------------------------ intersection.glsl:25-82
template <class T> T SphereIntersect(T rad, V3<T> pos, const Ray<T>& r) {
    V3<T> op = pos - r.origin;
    T eps = T(0.001);
    T b = dot(op, r.direction);
    T det = b * b - dot(op, op) + rad * rad;
    if (det < T(0)) return K<T>::INF;
    det = std::sqrt(det);
    T t1 = b - det;
    if (t1 > eps) return t1;
    T t2 = b + det;
    if (t2 > eps) return t2;
    return K<T>::INF;
}

template <class T> T RectIntersect(V3<T> pos, V3<T> u, V3<T> v, V4<T> plane, const Ray<T>& r) {
    V3<T> n(plane.x, plane.y, plane.z);
    T dt = dot(r.direction, n);
    T t = (plane.w - dot(n, r.origin)) / dt;
    if (t > K<T>::EPS) {
        V3<T> p = r.origin + r.direction * t;
        V3<T> vi = p - pos;
        T a1 = dot(u, vi);
        if (a1 >= T(0) && a1 <= T(1)) {
            T a2 = dot(v, vi);
            if (a2 >= T(0) && a2 <= T(1)) return t;
        }
    }
    return K<T>::INF;
}

template <class T> T AABBIntersect(V3<T> minCorner, V3<T> maxCorner, const Ray<T>& r) {
    V3<T> invDir = V3<T>(T(1)) / r.direction;
    V3<T> f = (maxCorner - r.origin) * invDir;
    V3<T> n = (minCorner - r.origin) * invDir;
    V3<T> tmax = vmax(f, n);
    V3<T> tmin = vmin(f, n);
    T t1 = gmin(tmax.x, gmin(tmax.y, tmax.z));
    T t0 = gmax(tmin.x, gmax(tmin.y, tmin.z));
    return (t1 >= t0) ? (t0 > T(0) ? t0 : t1) : T(-1);
}

// ------------------------------------------------------------ sampling.glsl:25-277
template <class T> T GTR1(T NDotH, T a) {
    if (a >= T(1)) return K<T>::INV_PI;
    T a2 = a * a;
    T t = T(1) + (a2 - T(1)) * NDotH * NDotH;
    return (a2 - T(1)) / (K<T>::PI * std::log(a2) * t);
}
// clamp(sqrt(1 - c * c), 0, 1) of SampleGTR1 and SampleHG (sampling.glsl):
// where rounding leaves |c| a hair above 1 GLSL's sqrt is undefined; a GPU's
// max/min return the non-NaN operand, so the reference's clamp gives 0, and
// so does this. Elsewhere the clamped root itself.
template <class T> inline T sin_from_cos(T c) {
    const T s2 = T(1) - (c * c);
    return s2 > T(0) ? gmin(std::sqrt(s2), T(1)) : T(0);
}
template <class T> V3<T> SampleGTR1(T rgh, T r1, T r2) {
    T a = gmax(T(0.001), rgh);
    T a2 = a * a;
    T phi = r1 * K<T>::TWO_PI;
    T cosTheta = std::sqrt((T(1) - std::pow(a2, T(1) - r2)) / (T(1) - a2));
    T sinTheta = sin_from_cos(cosTheta);
    T sinPhi = tsin(phi), cosPhi = tcos(phi);
    return V3<T>(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}
template <class T> V3<T> SampleGGXVNDF(V3<T> V, T ax, T ay, T r1, T r2) {
    V3<T> Vh = normalize(V3<T>(ax * V.x, ay * V.y, V.z));
    T lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    V3<T> T1 = lensq > T(0) ? V3<T>(-Vh.y, Vh.x, T(0)) * (T(1) / std::sqrt(lensq)) : V3<T>(T(1), T(0), T(0));
    V3<T> T2 = cross(Vh, T1);
    T r = std::sqrt(r1);
    T phi = T(2) * K<T>::PI * r2;
    T t1 = r * tcos(phi);
    T t2 = r * tsin(phi);
    T s = T(0.5) * (T(1) + Vh.z);
    t2 = (T(1) - s) * std::sqrt(T(1) - t1 * t1) + s * t2;
    V3<T> Nh = t1 * T1 + t2 * T2 + std::sqrt(gmax(T(0), T(1) - t1 * t1 - t2 * t2)) * Vh;
    return normalize(V3<T>(ax * Nh.x, ay * Nh.y, gmax(T(0), Nh.z)));
}
template <class T> T GTR2Aniso(T NDotH, T HDotX, T HDotY, T ax, T ay) {
    T a = HDotX / ax;
    T b = HDotY / ay;
    T c = a * a + b * b + NDotH * NDotH;
    return T(1) / (K<T>::PI * ax * ay * c * c);
}
template <class T> T SmithG(T NDotV, T alphaG) {
    T a = alphaG * alphaG;
    T b = NDotV * NDotV;
    return (T(2) * NDotV) / (NDotV + std::sqrt(a + b - a * b));
}
template <class T> T SmithGAniso(T NDotV, T VDotX, T VDotY, T ax, T ay) {
    T a = VDotX * ax;
    T b = VDotY * ay;
    T c = NDotV;
    return (T(2) * NDotV) / (NDotV + std::sqrt(a * a + b * b + c * c));
}
template <class T> T SchlickWeight(T u) {
    T m = gclamp(T(1) - u, T(0), T(1));
    T m2 = m * m;
    return m2 * m2 * m;
}
template <class T> T DielectricFresnel(T cosThetaI, T eta) {
    T sinThetaTSq = eta * eta * (T(1) - cosThetaI * cosThetaI);
    if (sinThetaTSq > T(1)) return T(1);
    T cosThetaT = std::sqrt(gmax(T(1) - sinThetaTSq, T(0)));
    T rs = (eta * cosThetaT - cosThetaI) / (eta * cosThetaT + cosThetaI);
    T rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    return T(0.5) * (rs * rs + rp * rp);
}
template <class T> V3<T> CosineSampleHemisphere(T r1, T r2) {
    V3<T> dir;
    T r = std::sqrt(r1);
    T phi = K<T>::TWO_PI * r2;
    dir.x = r * tcos(phi);
    dir.y = r * tsin(phi);
    dir.z = std::sqrt(gmax(T(0), T(1) - dir.x * dir.x - dir.y * dir.y));
    return dir;
}
template <class T> V3<T> UniformSampleHemisphere(T r1, T r2) {
    T r = std::sqrt(gmax(T(0), T(1) - r1 * r1));
    T phi = K<T>::TWO_PI * r2;
    return V3<T>(r * tcos(phi), r * tsin(phi), r1);
}
template <class T> T PowerHeuristic(T a, T b) {
    T t = a * a;
    return t / (b * b + t);
}
template <class T> void Onb(V3<T> N, V3<T>& Tt, V3<T>& B) {
    V3<T> up = std::fabs(N.z) < T(0.9999999f) ? V3<T>(T(0), T(0), T(1)) : V3<T>(T(1), T(0), T(0));
    Tt = normalize(cross(up, N));
    B = cross(N, Tt);
}
template <> void Onb<double>(V3<double> N, V3<double>& Tt, V3<double>& B) {
    V3<double> up = std::fabs(N.z) < 0.9999999 ? V3<double>(0, 0, 1) : V3<double>(1, 0, 0);
    Tt = normalize(cross(up, N));
    B = cross(N, Tt);
}
template <class T> void SampleSphereLight(Ctx<T>& c, const Light<T>& light, V3<T> scatterPos, LightSampleRec<T>& ls) {
    T r1 = c.rand();
    T r2 = c.rand();
    V3<T> sc = scatterPos - light.position;
    T distToSphereCenter = length(sc);
    sc = sc / distToSphereCenter;
    V3<T> sampledDir = UniformSampleHemisphere(r1, r2);
    V3<T> Tt, B;
    Onb(sc, Tt, B);
    sampledDir = Tt * sampledDir.x + B * sampledDir.y + sc * sampledDir.z;
    V3<T> lightSurfacePos = light.position + sampledDir * light.radius;
    ls.direction = lightSurfacePos - scatterPos;
    ls.dist = length(ls.direction);
    T distSq = ls.dist * ls.dist;
    ls.direction = ls.direction / ls.dist;
    ls.normal = normalize(lightSurfacePos - light.position);
    ls.emission = light.emission * T(c.s->num_lights);
    ls.pdf = distSq / (light.area * T(0.5) * std::fabs(dot(ls.normal, ls.direction)));
}
template <class T> void SampleRectLight(Ctx<T>& c, const Light<T>& light, V3<T> scatterPos, LightSampleRec<T>& ls) {
    T r1 = c.rand();
    T r2 = c.rand();
    V3<T> lightSurfacePos = light.position + light.u * r1 + light.v * r2;
    ls.direction = lightSurfacePos - scatterPos;
    ls.dist = length(ls.direction);
    T distSq = ls.dist * ls.dist;
    ls.direction = ls.direction / ls.dist;
    ls.normal = normalize(cross(light.u, light.v));
    ls.emission = light.emission * T(c.s->num_lights);
    ls.pdf = distSq / (light.area * std::fabs(dot(ls.normal, ls.direction)));
}
template <class T> void SampleDistantLight(Ctx<T>& c, const Light<T>& light, V3<T> scatterPos, LightSampleRec<T>& ls) {
    ls.direction = normalize(light.position - V3<T>(T(0)));
    ls.normal = normalize(scatterPos - light.position);
    ls.emission = light.emission * T(c.s->num_lights);
    ls.dist = K<T>::INF;
    ls.pdf = T(1);
}
template <class T> void SampleOneLight(Ctx<T>& c, const Light<T>& light, V3<T> scatterPos, LightSampleRec<T>& ls) {
    int type = static_cast<int>(light.type);
    if (type == QUAD_LIGHT) SampleRectLight(c, light, scatterPos, ls);
    else if (type == SPHERE_LIGHT) SampleSphereLight(c, light, scatterPos, ls);
    else SampleDistantLight(c, light, scatterPos, ls);
}
template <class T> V3<T> SampleHG(V3<T> V, T g, T r1, T r2) {
    T cosTheta;
    if (std::fabs(g) < T(0.001)) cosTheta = T(1) - T(2) * r2;
    else {
        T sqrTerm = (T(1) - g * g) / (T(1) + g - T(2) * g * r2);
        cosTheta = -(T(1) + g * g - sqrTerm * sqrTerm) / (T(2) * g);
    }
    T phi = r1 * K<T>::TWO_PI;
    T sinTheta = sin_from_cos(cosTheta);
    T sinPhi = tsin(phi), cosPhi = tcos(phi);
    V3<T> v1, v2;
    Onb(V, v1, v2);
    return sinTheta * cosPhi * v1 + sinTheta * sinPhi * v2 + cosTheta * V;
}
template <class T> T PhaseHG(T cosTheta, T g) {
    T denom = T(1) + g * g + T(2) * g * cosTheta;
    return K<T>::INV_4_PI * (T(1) - g * g) / (denom * std::sqrt(denom));
}

// ------------------------------------------------------------ envmap.glsl:28-83
template <class T> V2<T> BinarySearch(const oracle_scene* s, T value) {
    int W = s->env_width, H = s->env_height;
    int lower = 0, upper = H - 1;
    while (lower < upper) {
        int mid = (lower + upper) >> 1;
        if (value < T(s->env_cdf[mid * W + (W - 1)])) upper = mid;
        else lower = mid + 1;
    }
    int y = std::max(0, std::min(H - 1, lower));
    lower = 0; upper = W - 1;
    while (lower < upper) {
        int mid = (lower + upper) >> 1;
        if (value < T(s->env_cdf[y * W + mid])) upper = mid;
        else lower = mid + 1;
    }
    int x = std::max(0, std::min(W - 1, lower));
    return V2<T>{T(x) / T(W), T(y) / T(H)};
}
template <class T> V4<T> envtex(const oracle_scene* s, V2<T> uv) {
    return bilinear<T>(s->env_img, 3, s->env_width, s->env_height, uv.x, uv.y, false, nullptr);
}
template <class T> V4<T> EvalEnvMap(const oracle_scene* s, const oracle_options* o, const Ray<T>& r) {
    T theta = std::acos(gclamp(r.direction.y, T(-1), T(1)));
    V2<T> uv{(K<T>::PI + std::atan2(r.direction.z, r.direction.x)) * K<T>::INV_TWO_PI, theta * K<T>::INV_PI};
    uv.x = uv.x + T(o->envmap_rot);
    V4<T> c4 = envtex<T>(s, uv);
    V3<T> color(c4.x, c4.y, c4.z);
    T pdf = Luminance(color) / T(s->env_total_sum);
    return V4<T>{color.x, color.y, color.z,
                 (pdf * T(s->env_width) * T(s->env_height)) / (K<T>::TWO_PI * K<T>::PI * tsin(theta))};
}
template <class T> V4<T> SampleEnvMap(Ctx<T>& c, V3<T>& color) {
    const oracle_scene* s = c.s;
    V2<T> uv = BinarySearch<T>(s, c.rand() * T(s->env_total_sum));
    V4<T> c4 = envtex<T>(s, uv);
    color = V3<T>(c4.x, c4.y, c4.z);
    T pdf = Luminance(color) / T(s->env_total_sum);
    uv.x = uv.x - T(c.o->envmap_rot);
    T phi = uv.x * K<T>::TWO_PI;
    T theta = uv.y * K<T>::PI;
    if (tsin(theta) == T(0)) pdf = T(0);
    return V4<T>{-tsin(theta) * tcos(phi), tcos(theta), -tsin(theta) * tsin(phi),
                 (pdf * T(s->env_width) * T(s->env_height)) / (K<T>::TWO_PI * K<T>::PI * tsin(theta))};
}

// ------------------------------------------------------------ disney.glsl:37-351
template <class T> V3<T> ToWorld(V3<T> X, V3<T> Y, V3<T> Z, V3<T> V) { return V.x * X + V.y * Y + V.z * Z; }
template <class T> V3<T> ToLocal(V3<T> X, V3<T> Y, V3<T> Z, V3<T> V) { return V3<T>(dot(V, X), dot(V, Y), dot(V, Z)); }

template <class T> void TintColors(const Material<T>& mat, T eta, T& F0, V3<T>& Csheen, V3<T>& Cspec0) {
    T lum = Luminance(mat.baseColor);
    V3<T> ctint = lum > T(0) ? mat.baseColor / lum : V3<T>(T(1));
    F0 = (T(1) - eta) / (T(1) + eta);
    F0 = F0 * F0;
    Cspec0 = F0 * gmix(V3<T>(T(1)), ctint, mat.specularTint);
    Csheen = gmix(V3<T>(T(1)), ctint, mat.sheenTint);
}
template <class T> V3<T> EvalDisneyDiffuse(const Material<T>& mat, V3<T> Csheen, V3<T> V, V3<T> L, V3<T> H, T& pdf) {
    pdf = T(0);
    if (L.z <= T(0)) return V3<T>(T(0));
    T LDotH = dot(L, H);
    T Rr = T(2) * mat.roughness * LDotH * LDotH;
    T FL = SchlickWeight(L.z);
    T FV = SchlickWeight(V.z);
    T Fretro = Rr * (FL + FV + FL * FV * (Rr - T(1)));
    T Fd = (T(1) - T(0.5) * FL) * (T(1) - T(0.5) * FV);
    T Fss90 = T(0.5) * Rr;
    T Fss = gmix(T(1), Fss90, FL) * gmix(T(1), Fss90, FV);
    T ss = T(1.25) * (Fss * (T(1) / (L.z + V.z) - T(0.5)) + T(0.5));
    T FH = SchlickWeight(LDotH);
    V3<T> Fsheen = FH * mat.sheen * Csheen;
    pdf = L.z * K<T>::INV_PI;
    return K<T>::INV_PI * mat.baseColor * gmix(Fd + Fretro, ss, mat.subsurface) + Fsheen;
}
template <class T> V3<T> EvalMicrofacetReflection(const Material<T>& mat, V3<T> V, V3<T> L, V3<T> H, V3<T> F, T& pdf) {
    pdf = T(0);
    if (L.z <= T(0)) return V3<T>(T(0));
    T D = GTR2Aniso(H.z, H.x, H.y, mat.ax, mat.ay);
    T G1 = SmithGAniso(std::fabs(V.z), V.x, V.y, mat.ax, mat.ay);
    T G2 = G1 * SmithGAniso(std::fabs(L.z), L.x, L.y, mat.ax, mat.ay);
    pdf = G1 * D / (T(4) * V.z);
    return F * D * G2 / (T(4) * L.z * V.z);
}
template <class T> V3<T> EvalMicrofacetRefraction(const Material<T>& mat, T eta, V3<T> V, V3<T> L, V3<T> H, V3<T> F, T& pdf) {
    pdf = T(0);
    if (L.z >= T(0)) return V3<T>(T(0));
    T LDotH = dot(L, H);
    T VDotH = dot(V, H);
    T D = GTR2Aniso(H.z, H.x, H.y, mat.ax, mat.ay);
    T G1 = SmithGAniso(std::fabs(V.z), V.x, V.y, mat.ax, mat.ay);
    T G2 = G1 * SmithGAniso(std::fabs(L.z), L.x, L.y, mat.ax, mat.ay);
    T denom = LDotH + VDotH * eta;
    denom = denom * denom;
    T eta2 = eta * eta;
    T jacobian = std::fabs(LDotH) / denom;
    pdf = G1 * gmax(T(0), VDotH) * D * jacobian / V.z;
    return vpow(mat.baseColor, T(0.5)) * (V3<T>(T(1)) - F) * D * G2 * std::fabs(VDotH) * jacobian * eta2 /
           std::fabs(L.z * V.z);
}
template <class T> V3<T> EvalClearcoat(const Material<T>& mat, V3<T> V, V3<T> L, V3<T> H, T& pdf) {
    pdf = T(0);
    if (L.z <= T(0)) return V3<T>(T(0));
    T VDotH = dot(V, H);
    T F = gmix(T(0.04), T(1), SchlickWeight(VDotH));
    T D = GTR1(H.z, mat.clearcoatRoughness);
    T G = SmithG(L.z, T(0.25)) * SmithG(V.z, T(0.25));
    T jacobian = T(1) / (T(4) * VDotH);
    pdf = D * H.z * jacobian;
    return V3<T>(F) * D * G;
}

template <class T> struct LobePr { T diffPr, dielectricPr, metalPr, glassPr, clearCtPr, dielectricWt, metalWt, glassWt; };
template <class T> LobePr<T> Lobes(const Material<T>& mat, V3<T> Cspec0, T Vz) {
    LobePr<T> p;
    p.dielectricWt = (T(1) - mat.metallic) * (T(1) - mat.specTrans);
    p.metalWt = mat.metallic;
    p.glassWt = (T(1) - mat.metallic) * mat.specTrans;
    T schlickWt = SchlickWeight(Vz);
    p.diffPr = p.dielectricWt * Luminance(mat.baseColor);
    p.dielectricPr = p.dielectricWt * Luminance(gmix(Cspec0, V3<T>(T(1)), schlickWt));
    p.metalPr = p.metalWt * Luminance(gmix(mat.baseColor, V3<T>(T(1)), schlickWt));
    p.glassPr = p.glassWt;
    p.clearCtPr = T(0.25) * mat.clearcoat;
    T invTotalWt = T(1) / (p.diffPr + p.dielectricPr + p.metalPr + p.glassPr + p.clearCtPr);
    p.diffPr = p.diffPr * invTotalWt;
    p.dielectricPr = p.dielectricPr * invTotalWt;
    p.metalPr = p.metalPr * invTotalWt;
    p.glassPr = p.glassPr * invTotalWt;
    p.clearCtPr = p.clearCtPr * invTotalWt;
    return p;
}

template <class T> V3<T> DisneyEval(const State<T>& state, V3<T> V, V3<T> N, V3<T> L, T& pdf) {
    pdf = T(0);
    V3<T> f(T(0));
    V3<T> Tt, B;
    Onb(N, Tt, B);
    V = ToLocal(Tt, B, N, V);
    L = ToLocal(Tt, B, N, L);
    V3<T> H;
    if (L.z > T(0)) H = normalize(L + V);
    else H = normalize(L + V * state.eta);
    if (H.z < T(0)) H = -H;
    V3<T> Csheen, Cspec0;
    T F0;
    TintColors(state.mat, state.eta, F0, Csheen, Cspec0);
    LobePr<T> p = Lobes(state.mat, Cspec0, V.z);
    bool reflect = L.z * V.z > T(0);
    T tmpPdf = T(0);
    T VDotH = std::fabs(dot(V, H));
    if (p.diffPr > T(0) && reflect) {
        f = f + EvalDisneyDiffuse(state.mat, Csheen, V, L, H, tmpPdf) * p.dielectricWt;
        pdf = pdf + tmpPdf * p.diffPr;
    }
    if (p.dielectricPr > T(0) && reflect) {
        T F = (DielectricFresnel(VDotH, T(1) / state.mat.ior) - F0) / (T(1) - F0);
        f = f + EvalMicrofacetReflection(state.mat, V, L, H, gmix(Cspec0, V3<T>(T(1)), F), tmpPdf) * p.dielectricWt;
        pdf = pdf + tmpPdf * p.dielectricPr;
    }
    if (p.metalPr > T(0) && reflect) {
        V3<T> F = gmix(state.mat.baseColor, V3<T>(T(1)), SchlickWeight(VDotH));
        f = f + EvalMicrofacetReflection(state.mat, V, L, H, F, tmpPdf) * p.metalWt;
        pdf = pdf + tmpPdf * p.metalPr;
    }
    if (p.glassPr > T(0)) {
        T F = DielectricFresnel(VDotH, state.eta);
        if (reflect) {
            f = f + EvalMicrofacetReflection(state.mat, V, L, H, V3<T>(F), tmpPdf) * p.glassWt;
            pdf = pdf + tmpPdf * p.glassPr * F;
        } else {
            f = f + EvalMicrofacetRefraction(state.mat, state.eta, V, L, H, V3<T>(F), tmpPdf) * p.glassWt;
            pdf = pdf + tmpPdf * p.glassPr * (T(1) - F);
        }
    }
    if (p.clearCtPr > T(0) && reflect) {
        f = f + EvalClearcoat(state.mat, V, L, H, tmpPdf) * T(0.25) * state.mat.clearcoat;
        pdf = pdf + tmpPdf * p.clearCtPr;
    }
    return f * std::fabs(L.z);
}

template <class T> V3<T> DisneySample(Ctx<T>& c, const State<T>& state, V3<T> V, V3<T> N, V3<T>& L, T& pdf) {
    pdf = T(0);
    T r1 = c.rand();
    T r2 = c.rand();
    V3<T> Tt, B;
    Onb(N, Tt, B);
    V = ToLocal(Tt, B, N, V);
    V3<T> Csheen, Cspec0;
    T F0;
    TintColors(state.mat, state.eta, F0, Csheen, Cspec0);
    LobePr<T> p = Lobes(state.mat, Cspec0, V.z);
    T cdf[5];
    cdf[0] = p.diffPr;
    cdf[1] = cdf[0] + p.dielectricPr;
    cdf[2] = cdf[1] + p.metalPr;
    cdf[3] = cdf[2] + p.glassPr;
    cdf[4] = cdf[3] + p.clearCtPr;
    T r3 = c.rand();
    if (r3 < cdf[0]) {
        L = CosineSampleHemisphere(r1, r2);
    } else if (r3 < cdf[2]) {
        V3<T> H = SampleGGXVNDF(V, state.mat.ax, state.mat.ay, r1, r2);
        if (H.z < T(0)) H = -H;
        L = normalize(reflect(-V, H));
    } else if (r3 < cdf[3]) {
        V3<T> H = SampleGGXVNDF(V, state.mat.ax, state.mat.ay, r1, r2);
        T F = DielectricFresnel(std::fabs(dot(V, H)), state.eta);
        if (H.z < T(0)) H = -H;
        r3 = (r3 - cdf[2]) / (cdf[3] - cdf[2]);
        if (r3 < F) L = normalize(reflect(-V, H));
        else L = normalize(refract(-V, H, state.eta));
    } else {
        V3<T> H = SampleGTR1(state.mat.clearcoatRoughness, r1, r2);
        if (H.z < T(0)) H = -H;
        L = normalize(reflect(-V, H));
    }
    L = ToWorld(Tt, B, N, L);
    V = ToWorld(Tt, B, N, V);
    return DisneyEval(state, V, N, L, pdf);
}

// ------------------------------------------------------------ closest_hit.glsl:25-266
// Instance inverses (closest_hit.glsl:163-164 calls GLSL inverse() per TLAS
// leaf, and :246 transpose(inverse(mat3(transform))) for the normal; their
// precision is left to the GL driver). 1 (default) = both inverses computed in
// double and rounded to the scalar type, i.e. correctly rounded; 0 = cofactor
// inverses evaluated in the scalar type on every visit.
// The two differ only on rays that tie between coplanar faces or pass through
// a shared edge (tests/test_oracle_geometry.py bounds this). Set before use.
int g_inverse_mode = 1;
template <class T> M4<T> instance_inverse(const M4<T>& m) {
    if (g_inverse_mode == 0) return inverse4(m);
    M4<double> md;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) md.m[i][j] = static_cast<double>(m.m[i][j]);
    M4<double> id = inverse4(md);
    M4<T> r;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) r.m[i][j] = static_cast<T>(id.m[i][j]);
    return r;
}
template <class T> M3<T> normal_inverse(const M3<T>& m) {
    if (g_inverse_mode == 0) return inverse3(m);
    M3<double> md;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) md.m[i][j] = static_cast<double>(m.m[i][j]);
    M3<double> id = inverse3(md);
    M3<T> r;
    for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) r.m[i][j] = static_cast<T>(id.m[i][j]);
    return r;
}

template <class T> M4<T> fetch_transform(const oracle_scene* s, int inst) {
    M4<T> m;
    const float* p = s->transforms + 16 * inst;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) m.m[i][j] = T(p[4 * i + j]);
    return m;
}

template <class T> bool ClosestHit(Ctx<T>& c, const Ray<T>& r, State<T>& state, LightSampleRec<T>& lightSample) {
    const oracle_scene* s = c.s;
    const oracle_options* o = c.o;
    T t = K<T>::INF;
    T d;
    state.hitLight = -1; state.hitInstance = -1; state.hitTriRow = -1;
    if (c.cnt) c.cnt->rays_closest++;

    if (o->opt_lights && (!o->opt_hide_emitters || state.depth > 0)) {
        for (int i = 0; i < s->num_lights; i++) {
            if (c.cnt) c.cnt->lights++;
            Light<T> L = fetch_light<T>(s, i);
            V3<T> u = L.u, v = L.v;
            if (L.type == T(QUAD_LIGHT)) {
                V3<T> normal = normalize(cross(u, v));
                if (dot(normal, r.direction) > T(0)) continue;
                V4<T> plane{normal.x, normal.y, normal.z, dot(normal, L.position)};
                u = u * (T(1) / dot(u, u));
                v = v * (T(1) / dot(v, v));
                d = RectIntersect(L.position, u, v, plane, r);
                if (d < T(0)) d = K<T>::INF;
                if (d < t) {
                    t = d;
                    T cosTheta = dot(-r.direction, normal);
                    lightSample.pdf = (t * t) / (L.area * cosTheta);
                    lightSample.emission = L.emission;
                    state.isEmitter = true;
                    state.hitLight = i;
                }
            }
            if (L.type == T(SPHERE_LIGHT)) {
                d = SphereIntersect(L.radius, L.position, r);
                if (d < T(0)) d = K<T>::INF;
                if (d < t) {
                    t = d;
                    V3<T> hitPt = r.origin + t * r.direction;
                    T cosTheta = dot(-r.direction, normalize(hitPt - L.position));
                    lightSample.pdf = (t * t) / (L.area * cosTheta * T(0.5));
                    lightSample.emission = L.emission;
                    state.isEmitter = true;
                    state.hitLight = i;
                }
            }
        }
    }

    int stack[64];
    int ptr = 0;
    stack[ptr++] = -1;
    int index = s->top_level_index;
    T leftHit = T(0), rightHit = T(0);
    int currMatID = 0, currInst = -1;
    bool BLAS = false;
    int triID[3] = {-1, -1, -1};
    int triRow = -1, triInst = -1;
    M4<T> transMat, transform;
    for (int i = 0; i < 4; ++i) for (int j = 0; j < 4; ++j) { transMat.m[i][j] = T(i == j); transform.m[i][j] = T(i == j); }
    V3<T> bary;
    V4<T> vert0{}, vert1{}, vert2{};
    Ray<T> rTrans = r;

    while (index != -1) {
        const float* nd = s->nodes + 9 * static_cast<size_t>(index);
        int leftIndex = static_cast<int>(nd[6]);
        int rightIndex = static_cast<int>(nd[7]);
        int leaf = static_cast<int>(nd[8]);
        if (leaf > 0) {
            for (int i = 0; i < rightIndex; i++) {
                if (c.cnt) c.cnt->tris++;
                const int32_t* vi = s->vert_indices + 3 * static_cast<size_t>(leftIndex + i);
                const float* p0 = s->vertices_uvx + 4 * static_cast<size_t>(vi[0]);
                const float* p1 = s->vertices_uvx + 4 * static_cast<size_t>(vi[1]);
                const float* p2 = s->vertices_uvx + 4 * static_cast<size_t>(vi[2]);
                V4<T> v0{T(p0[0]), T(p0[1]), T(p0[2]), T(p0[3])};
                V4<T> v1{T(p1[0]), T(p1[1]), T(p1[2]), T(p1[3])};
                V4<T> v2{T(p2[0]), T(p2[1]), T(p2[2]), T(p2[3])};
                V3<T> e0 = V3<T>(v1.x, v1.y, v1.z) - V3<T>(v0.x, v0.y, v0.z);
                V3<T> e1 = V3<T>(v2.x, v2.y, v2.z) - V3<T>(v0.x, v0.y, v0.z);
                V3<T> pv = cross(rTrans.direction, e1);
                T det = dot(e0, pv);
                V3<T> tv = rTrans.origin - V3<T>(v0.x, v0.y, v0.z);
                V3<T> qv = cross(tv, e0);
                V4<T> uvt;
                uvt.x = dot(tv, pv);
                uvt.y = dot(rTrans.direction, qv);
                uvt.z = dot(e1, qv);
                const T rdet = T(1) / det;  // uvt.xyz / det as one reciprocal and three products
                uvt.x = uvt.x * rdet; uvt.y = uvt.y * rdet; uvt.z = uvt.z * rdet;
                uvt.w = T(1) - uvt.x - uvt.y;
                if (uvt.x >= T(0) && uvt.y >= T(0) && uvt.z >= T(0) && uvt.w >= T(0) && uvt.z < t) {
                    t = uvt.z;
                    triID[0] = vi[0]; triID[1] = vi[1]; triID[2] = vi[2];
                    triRow = leftIndex + i;
                    triInst = currInst;
                    state.matID = currMatID;
                    bary = V3<T>(uvt.w, uvt.x, uvt.y);
                    vert0 = v0; vert1 = v1; vert2 = v2;
                    transform = transMat;
                }
            }
        } else if (leaf < 0) {
            if (c.cnt) c.cnt->instances++;
            currInst = -leaf - 1;
            transMat = fetch_transform<T>(s, currInst);
            M4<T> inv = instance_inverse(transMat);
            rTrans.origin = mul_point(inv, r.origin, T(1));
            rTrans.direction = mul_point(inv, r.direction, T(0));
            if (ptr >= 64) { c.stack_overflow = true; break; }
            stack[ptr++] = -1;
            index = leftIndex;
            BLAS = true;
            currMatID = rightIndex;
            continue;
        } else {
            if (c.cnt) c.cnt->nodes++;
            const float* ln = s->nodes + 9 * static_cast<size_t>(leftIndex);
            const float* rn = s->nodes + 9 * static_cast<size_t>(rightIndex);
            leftHit = AABBIntersect(fetch3<T>(ln), fetch3<T>(ln + 3), rTrans);
            rightHit = AABBIntersect(fetch3<T>(rn), fetch3<T>(rn + 3), rTrans);
            if (leftHit > T(0) && rightHit > T(0)) {
                int deferred = -1;
                if (leftHit > rightHit) { index = rightIndex; deferred = leftIndex; }
                else { index = leftIndex; deferred = rightIndex; }
                if (ptr >= 64) { c.stack_overflow = true; break; }
                stack[ptr++] = deferred;
                continue;
            } else if (leftHit > T(0)) {
                index = leftIndex;
                continue;
            } else if (rightHit > T(0)) {
                index = rightIndex;
                continue;
            }
        }
        index = stack[--ptr];
        if (BLAS && index == -1) {
            BLAS = false;
            index = stack[--ptr];
            rTrans.origin = r.origin;
            rTrans.direction = r.direction;
        }
    }

    if (t == K<T>::INF) return false;
    state.hitDist = t;
    state.fhp = r.origin + r.direction * t;

    if (triID[0] != -1) {
        state.isEmitter = false;
        state.hitLight = -1;
        state.hitInstance = triInst;
        state.hitTriRow = triRow;
        state.baryU = bary.y; state.baryV = bary.z;
        const float* q0 = s->normals_uvy + 4 * static_cast<size_t>(triID[0]);
        const float* q1 = s->normals_uvy + 4 * static_cast<size_t>(triID[1]);
        const float* q2 = s->normals_uvy + 4 * static_cast<size_t>(triID[2]);
        V4<T> n0{T(q0[0]), T(q0[1]), T(q0[2]), T(q0[3])};
        V4<T> n1{T(q1[0]), T(q1[1]), T(q1[2]), T(q1[3])};
        V4<T> n2{T(q2[0]), T(q2[1]), T(q2[2]), T(q2[3])};
        V2<T> t0{vert0.w, n0.w}, t1{vert1.w, n1.w}, t2{vert2.w, n2.w};
        state.texCoord = t0 * bary.x + t1 * bary.y + t2 * bary.z;
        V3<T> normal = normalize(V3<T>(n0.x, n0.y, n0.z) * bary.x + V3<T>(n1.x, n1.y, n1.z) * bary.y +
                                 V3<T>(n2.x, n2.y, n2.z) * bary.z);
        M3<T> m3;
        for (int i = 0; i < 3; ++i) for (int j = 0; j < 3; ++j) m3.m[i][j] = transform.m[i][j];
        state.normal = normalize(mul3(transpose3(normal_inverse(m3)), normal));
        state.ffnormal = dot(state.normal, r.direction) <= T(0) ? state.normal : -state.normal;
        V3<T> deltaPos1 = V3<T>(vert1.x, vert1.y, vert1.z) - V3<T>(vert0.x, vert0.y, vert0.z);
        V3<T> deltaPos2 = V3<T>(vert2.x, vert2.y, vert2.z) - V3<T>(vert0.x, vert0.y, vert0.z);
        V2<T> deltaUV1 = t1 - t0;
        V2<T> deltaUV2 = t2 - t0;
        T invdet = T(1) / (deltaUV1.x * deltaUV2.y - deltaUV1.y * deltaUV2.x);
        state.tangent = (deltaPos1 * deltaUV2.y - deltaPos2 * deltaUV1.y) * invdet;
        state.bitangent = (deltaPos2 * deltaUV1.x - deltaPos1 * deltaUV2.x) * invdet;
        state.tangent = normalize(mul3(m3, state.tangent));
        state.bitangent = normalize(mul3(m3, state.bitangent));
    }
    return true;
}

// ------------------------------------------------------------ anyhit.glsl:25-215
template <class T> bool AnyHit(Ctx<T>& c, const Ray<T>& r, T maxDist) {
    const oracle_scene* s = c.s;
    const oracle_options* o = c.o;
    if (c.cnt) c.cnt->rays_shadow++;
    if (o->opt_lights) {
        for (int i = 0; i < s->num_lights; i++) {
            if (c.cnt) c.cnt->lights++;
            Light<T> L = fetch_light<T>(s, i);
            V3<T> u = L.u, v = L.v;
            if (L.type == T(QUAD_LIGHT)) {
                V3<T> normal = normalize(cross(u, v));
                V4<T> plane{normal.x, normal.y, normal.z, dot(normal, L.position)};
                u = u * (T(1) / dot(u, u));
                v = v * (T(1) / dot(v, v));
                T d = RectIntersect(L.position, u, v, plane, r);
                if (d > T(0) && d < maxDist) return true;
            }
            if (L.type == T(SPHERE_LIGHT)) {
                T d = SphereIntersect(L.radius, L.position, r);
                if (d > T(0) && d < maxDist) return true;
            }
        }
    }
    const bool alpha = o->opt_alpha_test && !o->opt_medium;
    int stack[64];
    int ptr = 0;
    stack[ptr++] = -1;
    int index = s->top_level_index;
    int currMatID = 0;
    bool BLAS = false;
    Ray<T> rTrans = r;
    while (index != -1) {
        const float* nd = s->nodes + 9 * static_cast<size_t>(index);
        int leftIndex = static_cast<int>(nd[6]);
        int rightIndex = static_cast<int>(nd[7]);
        int leaf = static_cast<int>(nd[8]);
        if (leaf > 0) {
            for (int i = 0; i < rightIndex; i++) {
                if (c.cnt) c.cnt->tris++;
                const int32_t* vi = s->vert_indices + 3 * static_cast<size_t>(leftIndex + i);
                const float* p0 = s->vertices_uvx + 4 * static_cast<size_t>(vi[0]);
                const float* p1 = s->vertices_uvx + 4 * static_cast<size_t>(vi[1]);
                const float* p2 = s->vertices_uvx + 4 * static_cast<size_t>(vi[2]);
                V3<T> a0 = fetch3<T>(p0), a1 = fetch3<T>(p1), a2 = fetch3<T>(p2);
                V3<T> e0 = a1 - a0;
                V3<T> e1 = a2 - a0;
                V3<T> pv = cross(rTrans.direction, e1);
                T det = dot(e0, pv);
                V3<T> tv = rTrans.origin - a0;
                V3<T> qv = cross(tv, e0);
                V4<T> uvt;
                uvt.x = dot(tv, pv);
                uvt.y = dot(rTrans.direction, qv);
                uvt.z = dot(e1, qv);
                const T rdet = T(1) / det;  // uvt.xyz / det as one reciprocal and three products
                uvt.x = uvt.x * rdet; uvt.y = uvt.y * rdet; uvt.z = uvt.z * rdet;
                uvt.w = T(1) - uvt.x - uvt.y;
                if (uvt.x >= T(0) && uvt.y >= T(0) && uvt.z >= T(0) && uvt.w >= T(0) && uvt.z < maxDist) {
                    if (alpha) {
                        const float* q0 = s->normals_uvy + 4 * static_cast<size_t>(vi[0]);
                        const float* q1 = s->normals_uvy + 4 * static_cast<size_t>(vi[1]);
                        const float* q2 = s->normals_uvy + 4 * static_cast<size_t>(vi[2]);
                        V2<T> t0{T(p0[3]), T(q0[3])}, t1{T(p1[3]), T(q1[3])}, t2{T(p2[3]), T(q2[3])};
                        V2<T> texCoord = t0 * uvt.w + t1 * uvt.x + t2 * uvt.y;
                        const float* mp = s->materials + 32 * static_cast<size_t>(currMatID);
                        T texID = T(mp[24]);
                        T alphaV = (s->num_textures > 0) ? tex_array<T>(s, texCoord, texID).w : T(1);
                        T opacity = T(mp[28]);
                        int alphaMode = static_cast<int>(mp[29]);
                        T alphaCutoff = T(mp[30]);
                        opacity = opacity * alphaV;
                        bool skip = (alphaMode == ALPHA_MODE_MASK && opacity < alphaCutoff) ||
                                    (alphaMode == ALPHA_MODE_BLEND && c.rand() > opacity);
                        if (!skip) return true;
                    } else {
                        return true;
                    }
                }
            }
        } else if (leaf < 0) {
            if (c.cnt) c.cnt->instances++;
            M4<T> tr = fetch_transform<T>(s, -leaf - 1);
            M4<T> inv = instance_inverse(tr);
            rTrans.origin = mul_point(inv, r.origin, T(1));
            rTrans.direction = mul_point(inv, r.direction, T(0));
            if (ptr >= 64) { c.stack_overflow = true; return false; }
            stack[ptr++] = -1;
            index = leftIndex;
            BLAS = true;
            currMatID = rightIndex;
            continue;
        } else {
            if (c.cnt) c.cnt->nodes++;
            const float* ln = s->nodes + 9 * static_cast<size_t>(leftIndex);
            const float* rn = s->nodes + 9 * static_cast<size_t>(rightIndex);
            T leftHit = AABBIntersect(fetch3<T>(ln), fetch3<T>(ln + 3), rTrans);
            T rightHit = AABBIntersect(fetch3<T>(rn), fetch3<T>(rn + 3), rTrans);
            if (leftHit > T(0) && rightHit > T(0)) {
                int deferred = -1;
                if (leftHit > rightHit) { index = rightIndex; deferred = leftIndex; }
                else { index = leftIndex; deferred = rightIndex; }
                if (ptr >= 64) { c.stack_overflow = true; return false; }
                stack[ptr++] = deferred;
                continue;
            } else if (leftHit > T(0)) { index = leftIndex; continue; }
            else if (rightHit > T(0)) { index = rightIndex; continue; }
        }
        index = stack[--ptr];
        if (BLAS && index == -1) {
            BLAS = false;
            index = stack[--ptr];
            rTrans.origin = r.origin;
            rTrans.direction = r.direction;
        }
    }
    return false;
}

// ------------------------------------------------------------ pathtrace.glsl:25-476
template <class T> void GetMaterialFrom(Ctx<T>& c, State<T>& state, const Ray<T>& r, const float* table) {
    const oracle_scene* s = c.s;
    const oracle_options* o = c.o;
    const float* p = table + 32 * static_cast<size_t>(state.matID);
    Material<T> mat;
    mat.baseColor = fetch3<T>(p);
    mat.anisotropic = T(p[3]);
    mat.emission = fetch3<T>(p + 4);
    mat.metallic = T(p[8]);
    mat.roughness = gmax(T(p[9]), T(0.001));
    mat.subsurface = T(p[10]);
    mat.specularTint = T(p[11]);
    mat.sheen = T(p[12]);
    mat.sheenTint = T(p[13]);
    mat.clearcoat = T(p[14]);
    mat.clearcoatRoughness = gmix(T(0.1), T(0.001), T(p[15]));
    mat.specTrans = T(p[16]);
    mat.ior = T(p[17]);
    mat.medium.type = static_cast<int>(p[18]);
    mat.medium.density = T(p[19]);
    mat.medium.color = fetch3<T>(p + 20);
    mat.medium.anisotropy = gclamp(T(p[23]), T(-0.9), T(0.9));
    int texIDs[4] = {static_cast<int>(p[24]), static_cast<int>(p[25]), static_cast<int>(p[26]), static_cast<int>(p[27])};
    mat.opacity = T(p[28]);
    mat.alphaMode = static_cast<int>(p[29]);
    mat.alphaCutoff = T(p[30]);

    if (texIDs[0] >= 0 && s->num_textures > 0) {
        V4<T> col = tex_array<T>(s, state.texCoord, T(texIDs[0]));
        mat.baseColor = mat.baseColor * vpow_tex(V3<T>(col.x, col.y, col.z), T(2.2f));
        mat.opacity = mat.opacity * col.w;
    }
    if (texIDs[1] >= 0 && s->num_textures > 0) {
        V4<T> t4 = tex_array<T>(s, state.texCoord, T(texIDs[1]));
        mat.metallic = t4.z;
        mat.roughness = gmax(t4.y * t4.y, T(0.001));
    }
    if (texIDs[2] >= 0 && s->num_textures > 0) {
        V4<T> t4 = tex_array<T>(s, state.texCoord, T(texIDs[2]));
        V3<T> texNormal(t4.x, t4.y, t4.z);
        if (o->opt_opengl_normalmap) texNormal.y = T(1) - texNormal.y;
        texNormal = normalize(texNormal * T(2) - V3<T>(T(1)));
        V3<T> origNormal = state.normal;
        state.normal = normalize(state.tangent * texNormal.x + state.bitangent * texNormal.y + state.normal * texNormal.z);
        state.ffnormal = dot(origNormal, r.direction) <= T(0) ? state.normal : -state.normal;
    }
    if (o->opt_roughness_mollification && state.depth > 0)
        mat.roughness = gmax(gmix(T(0), state.mat.roughness, T(o->roughness_mollification_amt)), mat.roughness);
    if (texIDs[3] >= 0 && s->num_textures > 0) {
        V4<T> t4 = tex_array<T>(s, state.texCoord, T(texIDs[3]));
        mat.emission = vpow_tex(V3<T>(t4.x, t4.y, t4.z), T(2.2f));
    }
    T aspect = std::sqrt(T(1) - mat.anisotropic * T(0.9));
    mat.ax = gmax(T(0.001), mat.roughness / aspect);
    mat.ay = gmax(T(0.001), mat.roughness * aspect);
    state.mat = mat;
    state.eta = dot(r.direction, state.normal) < T(0) ? (T(1) / mat.ior) : mat.ior;
}

template <class T> void GetMaterial(Ctx<T>& c, State<T>& state, const Ray<T>& r) {
    if (c.eval_mats) {
        State<T> e = state;  // shading-normal edits of this call are discarded
        e.mat = state.mat_e;
        GetMaterialFrom(c, e, r, c.eval_mats);
        state.mat_e = e.mat;
        state.eta_e = e.eta;
    }
    GetMaterialFrom(c, state, r, c.s->materials);
}

// f of the BSDF that weights the path: the scene's material, or in
// frozen-sampling mode the eval table's (pdf always from the scene's).
template <class T> V3<T> EvalF(const Ctx<T>& c, const State<T>& state, V3<T> V, V3<T> N, V3<T> L, T& pdf) {
    V3<T> f = DisneyEval(state, V, N, L, pdf);
    if (c.eval_mats) {
        State<T> e = state;
        e.mat = state.mat_e;
        e.eta = state.eta_e;
        T unused;
        f = DisneyEval(e, V, N, L, unused);
    }
    return f;
}

template <class T> V3<T> EvalTransmittance(Ctx<T>& c, Ray<T> r) {
    LightSampleRec<T> lightSample;
    State<T> state;
    V3<T> transmittance(T(1));
    for (int depth = 0; depth < c.o->max_depth; depth++) {
        bool hit = ClosestHit(c, r, state, lightSample);
        if (!hit || state.isEmitter) break;
        GetMaterial(c, state, r);
        bool alphatest = (state.mat.alphaMode == ALPHA_MODE_MASK && state.mat.opacity < state.mat.alphaCutoff) ||
                         (state.mat.alphaMode == ALPHA_MODE_BLEND && c.rand() > state.mat.opacity);
        bool refractive = (T(1) - state.mat.metallic) * state.mat.specTrans > T(0);
        if (hit && !(alphatest || refractive)) return V3<T>(T(0));
        if (dot(r.direction, state.normal) > T(0) && state.mat.medium.type != MEDIUM_NONE) {
            V3<T> color = state.mat.medium.type == MEDIUM_ABSORB ? V3<T>(T(1)) - state.mat.medium.color : V3<T>(T(1));
            transmittance = transmittance * vexp(-color * state.mat.medium.density * state.hitDist);
        }
        r.origin = state.fhp + r.direction * K<T>::EPS;
    }
    return transmittance;
}

template <class T> V3<T> DirectLight(Ctx<T>& c, const Ray<T>& r, const State<T>& state, bool isSurface) {
    const oracle_scene* s = c.s;
    const oracle_options* o = c.o;
    V3<T> Ld(T(0));
    V3<T> Li(T(0));
    V3<T> scatterPos = state.fhp + state.normal * K<T>::EPS;
    ScatterSampleRec<T> scatterSample;
    const bool volmis = o->opt_medium && o->opt_vol_mis;

    if (o->opt_envmap && !o->opt_uniform_light) {
        V4<T> dirPdf = SampleEnvMap(c, Li);
        V3<T> lightDir(dirPdf.x, dirPdf.y, dirPdf.z);
        T lightPdf = dirPdf.w;
        Ray<T> shadowRay{scatterPos, lightDir};
        if (volmis) {
            Li = Li * EvalTransmittance(c, shadowRay);
            if (isSurface) scatterSample.f = EvalF(c, state, -r.direction, state.ffnormal, lightDir, scatterSample.pdf);
            else {
                T p = PhaseHG(dot(-r.direction, lightDir), state.medium.anisotropy);
                scatterSample.f = V3<T>(p);
                scatterSample.pdf = p;
            }
            if (scatterSample.pdf > T(0)) {
                T misWeight = PowerHeuristic(lightPdf, scatterSample.pdf);
                if (misWeight > T(0)) Ld = Ld + misWeight * Li * scatterSample.f * T(o->envmap_intensity) / lightPdf;
            }
        } else {
            bool inShadow = AnyHit(c, shadowRay, K<T>::INF - K<T>::EPS);
            if (!inShadow) {
                scatterSample.f = EvalF(c, state, -r.direction, state.ffnormal, lightDir, scatterSample.pdf);
                if (scatterSample.pdf > T(0)) {
                    T misWeight = PowerHeuristic(lightPdf, scatterSample.pdf);
                    if (misWeight > T(0)) Ld = Ld + misWeight * Li * scatterSample.f * T(o->envmap_intensity) / lightPdf;
                }
            }
        }
    }

    if (o->opt_lights) {
        LightSampleRec<T> lightSample;
        // int(rand() * N); rand() == 1.0 would index one past the end: clamped.
        int index = static_cast<int>(c.rand() * T(s->num_lights));
        if (index > s->num_lights - 1) index = s->num_lights - 1;
        Light<T> light = fetch_light<T>(s, index);
        SampleOneLight(c, light, scatterPos, lightSample);
        Li = lightSample.emission;
        if (dot(lightSample.direction, lightSample.normal) < T(0)) {
            Ray<T> shadowRay{scatterPos, lightSample.direction};
            if (volmis) {
                Li = Li * EvalTransmittance(c, shadowRay);
                if (isSurface) scatterSample.f = EvalF(c, state, -r.direction, state.ffnormal, lightSample.direction, scatterSample.pdf);
                else {
                    T p = PhaseHG(dot(-r.direction, lightSample.direction), state.medium.anisotropy);
                    scatterSample.f = V3<T>(p);
                    scatterSample.pdf = p;
                }
                T misWeight = T(1);
                if (light.area > T(0)) misWeight = PowerHeuristic(lightSample.pdf, scatterSample.pdf);
                if (scatterSample.pdf > T(0)) Ld = Ld + misWeight * scatterSample.f * Li / lightSample.pdf;
            } else {
                bool inShadow = AnyHit(c, shadowRay, lightSample.dist - K<T>::EPS);
                if (!inShadow) {
                    scatterSample.f = EvalF(c, state, -r.direction, state.ffnormal, lightSample.direction, scatterSample.pdf);
                    T misWeight = T(1);
                    if (light.area > T(0)) misWeight = PowerHeuristic(lightSample.pdf, scatterSample.pdf);
                    if (scatterSample.pdf > T(0)) Ld = Ld + misWeight * Li * scatterSample.f / lightSample.pdf;
                }
            }
        }
    }
    return Ld;
}

template <class T> V4<T> PathTrace(Ctx<T>& c, Ray<T> r) {
    const oracle_options* o = c.o;
    V3<T> radiance(T(0));
    V3<T> throughput(T(1));
    V3<T> thr_s(T(1));  // throughput of the sampling materials (== throughput unless frozen-sampling)
    State<T> state;
    LightSampleRec<T> lightSample;
    ScatterSampleRec<T> scatterSample;
    T alpha = T(1);
    bool inMedium = false, mediumSampled = false, surfaceScatter = false;
    int passThrough = 0;  // guard for the alpha pass-through depth-- loop

    for (state.depth = 0;; state.depth++) {
        bool hit = ClosestHit(c, r, state, lightSample);
        if (!hit) {
            if ((o->opt_background || o->opt_transparent_background) && state.depth == 0) alpha = T(0);
            if (!o->opt_hide_emitters || state.depth > 0) {
                if (o->opt_uniform_light) {
                    radiance = radiance + V3<T>(T(o->uniform_light_col[0]), T(o->uniform_light_col[1]), T(o->uniform_light_col[2])) * throughput;
                } else if (o->opt_envmap) {
                    V4<T> envMapColPdf = EvalEnvMap(c.s, o, r);
                    T misWeight = T(1);
                    if (state.depth > 0) misWeight = PowerHeuristic(scatterSample.pdf, envMapColPdf.w);
                    if (o->opt_medium && !o->opt_vol_mis && !surfaceScatter) misWeight = T(1);
                    if (misWeight > T(0))
                        radiance = radiance + misWeight * V3<T>(envMapColPdf.x, envMapColPdf.y, envMapColPdf.z) * throughput * T(o->envmap_intensity);
                }
            }
            break;
        }

        GetMaterial(c, state, r);
        radiance = radiance + state.mat.emission * throughput;

        if (o->opt_lights && state.isEmitter) {
            T misWeight = T(1);
            if (state.depth > 0) misWeight = PowerHeuristic(scatterSample.pdf, lightSample.pdf);
            if (o->opt_medium && !o->opt_vol_mis && !surfaceScatter) misWeight = T(1);
            radiance = radiance + misWeight * lightSample.emission * throughput;
            break;
        }
        if (state.depth == o->max_depth) break;

        if (o->opt_medium) {
            mediumSampled = false;
            surfaceScatter = false;
            if (inMedium) {
                const Medium<T>& me = c.eval_mats ? state.medium_e : state.medium;
                if (state.medium.type == MEDIUM_ABSORB) {
                    throughput = throughput * vexp(-(V3<T>(T(1)) - me.color) * state.hitDist * me.density);
                    thr_s = thr_s * vexp(-(V3<T>(T(1)) - state.medium.color) * state.hitDist * state.medium.density);
                } else if (state.medium.type == MEDIUM_EMISSIVE) {
                    radiance = radiance + me.color * state.hitDist * me.density * throughput;
                } else {
                    T scatterDist = gmin(-std::log(c.rand()) / state.medium.density, state.hitDist);
                    mediumSampled = scatterDist < state.hitDist;
                    if (c.eval_mats && !mediumSampled)
                        throughput = throughput * (std::exp(-me.density * state.hitDist) /
                                                   std::exp(-state.medium.density * state.hitDist));
                    if (mediumSampled) {
                        if (c.eval_mats)
                            throughput = throughput * (me.color * ((me.density * std::exp(-me.density * scatterDist)) /
                                                                   (state.medium.density *
                                                                    std::exp(-state.medium.density * scatterDist))));
                        else
                            throughput = throughput * state.medium.color;
                        thr_s = thr_s * state.medium.color;
                        r.origin = r.origin + r.direction * scatterDist;
                        state.fhp = r.origin;
                        radiance = radiance + DirectLight(c, r, state, false) * throughput;
                        T ra = c.rand();
                        T rb = c.rand();
                        V3<T> scatterDir = SampleHG(-r.direction, state.medium.anisotropy, ra, rb);
                        scatterSample.pdf = PhaseHG(dot(-r.direction, scatterDir), state.medium.anisotropy);
                        r.direction = scatterDir;
                    }
                }
            }
        }

        if (!o->opt_medium || !mediumSampled) {
            bool passed = false;
            if (o->opt_alpha_test) {
                passed = (state.mat.alphaMode == ALPHA_MODE_MASK && state.mat.opacity < state.mat.alphaCutoff) ||
                         (state.mat.alphaMode == ALPHA_MODE_BLEND && c.rand() > state.mat.opacity);
            }
            if (passed) {
                scatterSample.L = r.direction;
                state.depth--;
                if (++passThrough > 1024) break;
            } else {
                surfaceScatter = true;
                radiance = radiance + DirectLight(c, r, state, true) * throughput;
                scatterSample.f = DisneySample(c, state, -r.direction, state.ffnormal, scatterSample.L, scatterSample.pdf);
                if (scatterSample.pdf > T(0)) {
                    thr_s = thr_s * (scatterSample.f / scatterSample.pdf);
                    if (c.eval_mats) {
                        T unused;
                        scatterSample.f = EvalF(c, state, -r.direction, state.ffnormal, scatterSample.L, unused);
                    }
                    throughput = throughput * (scatterSample.f / scatterSample.pdf);
                } else {
                    break;
                }
            }
            r.direction = scatterSample.L;
            r.origin = state.fhp + r.direction * K<T>::EPS;
            if (o->opt_medium) {
                if (dot(r.direction, state.normal) < T(0) && state.mat.medium.type != MEDIUM_NONE) {
                    inMedium = true;
                    state.medium = state.mat.medium;
                    if (c.eval_mats) state.medium_e = state.mat_e.medium;
                } else if (state.mat.medium.type != MEDIUM_NONE) {
                    inMedium = false;
                }
            }
        }

        if (o->opt_rr && state.depth >= o->rr_depth) {
            T q = gmin(gmax(thr_s.x, gmax(thr_s.y, thr_s.z)) + T(0.001), T(0.95));
            if (c.rand() > q) break;
            throughput = throughput / q;
            thr_s = thr_s / q;
        }
    }
    return V4<T>{radiance.x, radiance.y, radiance.z, alpha};
}

// ------------------------------------------------------------ tile.glsl:41-75
// tile.glsl:41-71: seed the pixel's RNG and generate its camera ray.
template <class T> Ray<T> CameraRay(Ctx<T>& c, const oracle_camera* cam, int px, int py, int sample) {
    const oracle_options* o = c.o;
    // InitRNG(gl_FragCoord.xy, frameNum): seed = (x, y, frame, x + y); with one tile the
    // first accumulated frame is frameNum 2.
    c.seed[0] = static_cast<uint32_t>(px);
    c.seed[1] = static_cast<uint32_t>(py);
    c.seed[2] = static_cast<uint32_t>(sample + 2);
    c.seed[3] = static_cast<uint32_t>(px) + static_cast<uint32_t>(py);
    V2<T> res{T(o->width), T(o->height)};
    V2<T> coords{(T(px) + T(0.5)) / res.x, (T(py) + T(0.5)) / res.y};
    T r1 = T(2) * c.rand();
    T r2 = T(2) * c.rand();
    V2<T> jitter;
    jitter.x = r1 < T(1) ? std::sqrt(r1) - T(1) : T(1) - std::sqrt(T(2) - r1);
    jitter.y = r2 < T(1) ? std::sqrt(r2) - T(1) : T(1) - std::sqrt(T(2) - r2);
    jitter.x = jitter.x / (res.x * T(0.5));
    jitter.y = jitter.y / (res.y * T(0.5));
    V2<T> d{(coords.x * T(2) - T(1)) + jitter.x, (coords.y * T(2) - T(1)) + jitter.y};
    T scale = std::tan(T(cam->fov) * T(0.5));
    d.y = d.y * (res.y / res.x * scale);
    d.x = d.x * scale;
    V3<T> right = fetch3<T>(cam->right), up = fetch3<T>(cam->up), fwd = fetch3<T>(cam->forward);
    V3<T> rayDir = normalize(d.x * right + d.y * up + fwd);
    V3<T> focalPoint = T(cam->focalDist) * rayDir;
    T cam_r1 = c.rand() * K<T>::TWO_PI;
    T cam_r2 = c.rand() * T(cam->aperture);
    V3<T> randomAperturePos = (tcos(cam_r1) * right + tsin(cam_r1) * up) * std::sqrt(cam_r2);
    V3<T> finalRayDir = normalize(focalPoint - randomAperturePos);
    return Ray<T>{fetch3<T>(cam->position) + randomAperturePos, finalRayDir};
}

template <class T> V4<T> TileSample(Ctx<T>& c, const oracle_camera* cam, int px, int py, int sample) {
    Ray<T> ray = CameraRay(c, cam, px, py, sample);
    return PathTrace(c, ray);
}

template <class T, class Acc>
int RenderT(const oracle_scene* s, const oracle_camera* cam, const oracle_options* o, int first, int n,
            int row_begin, int row_end, int row_step, Acc* accum, int nthreads, oracle_counters* out,
            const float* eval_mats = nullptr) {
    const int W = o->width;
    const int rows = row_step > 0 ? (row_end - row_begin + row_step - 1) / row_step : 0;
    const int tiles_x = (W + 7) / 8, tiles_y = (rows + 7) / 8;
    const int ntiles = tiles_x * tiles_y;
    Counters total;
    int overflow = 0;
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
    {
        Counters local;
        Ctx<T> c{s, o, {0, 0, 0, 0}, out ? &local : nullptr};
        c.eval_mats = eval_mats;
#pragma omp for schedule(dynamic)
        for (int tile = 0; tile < ntiles; ++tile) {
            const int tx = tile % tiles_x, ty = tile / tiles_x;
            for (int jr = ty * 8; jr < std::min(rows, ty * 8 + 8); ++jr) {
                const int py = row_begin + jr * row_step;
                for (int px = tx * 8; px < std::min(W, tx * 8 + 8); ++px) {
                    for (int smp = first; smp < first + n; ++smp) {
                        V4<T> col = TileSample(c, cam, px, py, smp);
                        Acc* a = accum + 4 * (static_cast<size_t>(py) * W + px);
                        a[0] += static_cast<Acc>(col.x);
                        a[1] += static_cast<Acc>(col.y);
                        a[2] += static_cast<Acc>(col.z);
                        a[3] += static_cast<Acc>(col.w);
                    }
                }
            }
        }
#pragma omp critical
        {
            total.rays_closest += local.rays_closest; total.rays_shadow += local.rays_shadow;
            total.nodes += local.nodes; total.tris += local.tris; total.instances += local.instances;
            total.lights += local.lights;
            if (c.stack_overflow) overflow = 1;
        }
    }
    if (out) {
        out->rays_closest = total.rays_closest; out->rays_shadow = total.rays_shadow; out->nodes = total.nodes;
        out->tris = total.tris; out->instances = total.instances; out->lights = total.lights;
    }
    return overflow ? -4 : 0;
}

Material<float> MaterialFromRecord(const float* p) {
    // GetMaterial without textures, depth 0 (pathtrace.glsl:25-115).
    oracle_scene s{};
    oracle_options o{};
    s.materials = p;
    s.num_materials = 1;
    Ctx<float> c{&s, &o, {0, 0, 0, 0}, nullptr};
    State<float> st;
    st.matID = 0;
    st.normal = V3<float>(0, 0, 1);
    Ray<float> r{V3<float>(0, 0, 1), V3<float>(0, 0, -1)};
    GetMaterial(c, st, r);
    return st.mat;
}

}  // namespace

extern "C" {

void oracle_set_inverse_mode(int32_t mode) { g_inverse_mode = mode; }

int oracle_eval_fn(int32_t fn, int32_t n, const float* in, float* out) {
    static const int IN[] = {2, 5, 2, 5, 2, 1, 2, 7, 3, 2, 2, 6, 2, 10, 19, 12, 1};
    static const int OUT[] = {1, 1, 1, 1, 1, 1, 1, 3, 3, 3, 3, 3, 1, 1, 1, 1, 2};
    if (fn < 0 || fn > 16) return -1;
    const int ni = IN[fn], no = OUT[fn];
    for (int i = 0; i < n; ++i) {
        const float* a = in + static_cast<size_t>(i) * ni;
        float* o = out + static_cast<size_t>(i) * no;
        V3<float> v;
        switch (fn) {
            case 0: o[0] = GTR1(a[0], a[1]); break;
            case 1: o[0] = GTR2Aniso(a[0], a[1], a[2], a[3], a[4]); break;
            case 2: o[0] = SmithG(a[0], a[1]); break;
            case 3: o[0] = SmithGAniso(a[0], a[1], a[2], a[3], a[4]); break;
            case 4: o[0] = DielectricFresnel(a[0], a[1]); break;
            case 5: o[0] = SchlickWeight(a[0]); break;
            case 6: o[0] = PowerHeuristic(a[0], a[1]); break;
            case 7: v = SampleGGXVNDF(fetch3<float>(a), a[3], a[4], a[5], a[6]); break;
            case 8: v = SampleGTR1(a[0], a[1], a[2]); break;
            case 9: v = CosineSampleHemisphere(a[0], a[1]); break;
            case 10: v = UniformSampleHemisphere(a[0], a[1]); break;
            case 11: v = SampleHG(fetch3<float>(a), a[3], a[4], a[5]); break;
            case 12: o[0] = PhaseHG(a[0], a[1]); break;
            case 13: {
                Ray<float> r{fetch3<float>(a + 4), fetch3<float>(a + 7)};
                o[0] = SphereIntersect(a[0], fetch3<float>(a + 1), r);
                break;
            }
            case 14: {
                Ray<float> r{fetch3<float>(a + 13), fetch3<float>(a + 16)};
                V4<float> pl{a[9], a[10], a[11], a[12]};
                o[0] = RectIntersect(fetch3<float>(a), fetch3<float>(a + 3), fetch3<float>(a + 6), pl, r);
                break;
            }
            case 15: {
                Ray<float> r{fetch3<float>(a + 6), fetch3<float>(a + 9)};
                o[0] = AABBIntersect(fetch3<float>(a), fetch3<float>(a + 3), r);
                break;
            }
            case 16: o[0] = tsin(a[0]); o[1] = tcos(a[0]); break;
        }
        if (no == 3) { o[0] = v.x; o[1] = v.y; o[2] = v.z; }
    }
    return 0;
}

int oracle_render(int32_t precision, const oracle_scene* s, const oracle_camera* cam, const oracle_options* o,
                  int32_t first_sample, int32_t num_samples, int32_t row_begin, int32_t row_end, int32_t row_step,
                  float* accum_f32, double* accum_f64, int32_t num_threads, oracle_counters* counters) {
    if (!s || !cam || !o) return -1;
    if (precision == 64) {
        if (!accum_f64) return -1;
        return RenderT<double, double>(s, cam, o, first_sample, num_samples, row_begin, row_end, row_step, accum_f64,
                                       num_threads, counters);
    }
    if (!accum_f32) return -1;
    return RenderT<float, float>(s, cam, o, first_sample, num_samples, row_begin, row_end, row_step, accum_f32,
                                 num_threads, counters);
}

int oracle_trace_samples(const oracle_scene* s, const oracle_camera* cam, const oracle_options* o, int32_t n,
                         const int32_t* px, const int32_t* py, const int32_t* sample, float* rgba) {
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i) {
        Ctx<float> c{s, o, {0, 0, 0, 0}, nullptr};
        V4<float> col = TileSample(c, cam, px[i], py[i], sample[i]);
        rgba[4 * i] = col.x; rgba[4 * i + 1] = col.y; rgba[4 * i + 2] = col.z; rgba[4 * i + 3] = col.w;
    }
    return 0;
}

int oracle_camera_rays(const oracle_camera* cam, const oracle_options* o, int32_t n, const int32_t* px,
                       const int32_t* py, const int32_t* sample, float* orig, float* dir) {
    if (!cam || !o || n < 0 || !px || !py || !sample || !orig || !dir) return -1;
    oracle_scene s{};
    for (int i = 0; i < n; ++i) {
        Ctx<float> c{&s, o, {0, 0, 0, 0}, nullptr};
        Ray<float> r = CameraRay(c, cam, px[i], py[i], sample[i]);
        orig[3 * i] = r.origin.x; orig[3 * i + 1] = r.origin.y; orig[3 * i + 2] = r.origin.z;
        dir[3 * i] = r.direction.x; dir[3 * i + 1] = r.direction.y; dir[3 * i + 2] = r.direction.z;
    }
    return 0;
}

int oracle_closest_hit(const oracle_scene* s, const oracle_options* o, const float* orig, const float* dir, int32_t n,
                       float* t, int32_t* instance, int32_t* triangle, int32_t* light, float* uv,
                       oracle_counters* counters) {
    Counters total;
#pragma omp parallel
    {
        Counters local;
        Ctx<float> c{s, o, {0, 0, 0, 0}, counters ? &local : nullptr};
#pragma omp for schedule(dynamic, 256)
        for (int i = 0; i < n; ++i) {
            Ray<float> r{V3<float>(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]),
                         V3<float>(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2])};
            State<float> st;
            st.depth = 1;
            LightSampleRec<float> ls;
            bool hit = ClosestHit(c, r, st, ls);
            t[i] = hit ? st.hitDist : K<float>::INF;
            instance[i] = hit ? st.hitInstance : -1;
            triangle[i] = hit ? st.hitTriRow : -1;
            light[i] = hit ? st.hitLight : -1;
            uv[2 * i] = hit && st.hitTriRow >= 0 ? st.baryU : 0.f;
            uv[2 * i + 1] = hit && st.hitTriRow >= 0 ? st.baryV : 0.f;
        }
#pragma omp critical
        {
            total.rays_closest += local.rays_closest; total.nodes += local.nodes; total.tris += local.tris;
            total.instances += local.instances; total.lights += local.lights;
        }
    }
    if (counters) {
        counters->rays_closest = total.rays_closest; counters->rays_shadow = 0; counters->nodes = total.nodes;
        counters->tris = total.tris; counters->instances = total.instances; counters->lights = total.lights;
    }
    return 0;
}

int oracle_any_hit(const oracle_scene* s, const oracle_options* o, const float* orig, const float* dir,
                   const float* max_dist, int32_t n, int32_t* occluded) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
        Ctx<float> c{s, o, {1, 2, 3, 4}, nullptr};
        Ray<float> r{V3<float>(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]),
                     V3<float>(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2])};
        occluded[i] = AnyHit(c, r, max_dist[i]) ? 1 : 0;
    }
    return 0;
}

void oracle_pcg4d(const uint32_t seed[4], int32_t n, uint32_t* states_out, float* rand_out) {
    uint32_t v[4] = {seed[0], seed[1], seed[2], seed[3]};
    for (int i = 0; i < n; ++i) {
        float r = rand_f(v);
        if (states_out) std::memcpy(states_out + 4 * i, v, 16);
        if (rand_out) rand_out[i] = r;
    }
}

int oracle_disney_eval(const float* mat, float eta, const float* V, const float* N, const float* L, float* f_out,
                       float* pdf_out) {
    State<float> st;
    st.mat = MaterialFromRecord(mat);
    V3<float> v(V[0], V[1], V[2]), n(N[0], N[1], N[2]), l(L[0], L[1], L[2]);
    st.eta = eta >= 0.f ? eta : (dot(-v, n) < 0.f ? 1.f / st.mat.ior : st.mat.ior);
    float pdf;
    V3<float> f = DisneyEval(st, v, n, l, pdf);
    f_out[0] = f.x; f_out[1] = f.y; f_out[2] = f.z;
    *pdf_out = pdf;
    return 0;
}

int oracle_disney_sample(const float* mat, float eta, const float* V, const float* N, uint32_t seed[4], float* L_out,
                         float* f_out, float* pdf_out) {
    oracle_scene s{};
    oracle_options o{};
    Ctx<float> c{&s, &o, {seed[0], seed[1], seed[2], seed[3]}, nullptr};
    State<float> st;
    st.mat = MaterialFromRecord(mat);
    V3<float> v(V[0], V[1], V[2]), n(N[0], N[1], N[2]), l;
    st.eta = eta >= 0.f ? eta : (dot(-v, n) < 0.f ? 1.f / st.mat.ior : st.mat.ior);
    float pdf;
    V3<float> f = DisneySample(c, st, v, n, l, pdf);
    L_out[0] = l.x; L_out[1] = l.y; L_out[2] = l.z;
    f_out[0] = f.x; f_out[1] = f.y; f_out[2] = f.z;
    *pdf_out = pdf;
    std::memcpy(seed, c.seed, 16);
    return 0;
}

int oracle_sample_light(const float* light, int32_t num_lights, const float* scatter_pos, uint32_t seed[4],
                        float* rec) {
    oracle_scene s{};
    s.num_lights = num_lights;
    oracle_options o{};
    Ctx<float> c{&s, &o, {seed[0], seed[1], seed[2], seed[3]}, nullptr};
    Light<float> L{fetch3<float>(light), fetch3<float>(light + 3), fetch3<float>(light + 6), fetch3<float>(light + 9),
                   light[12], light[13], light[14]};
    LightSampleRec<float> ls;
    SampleOneLight(c, L, fetch3<float>(scatter_pos), ls);
    float v[13] = {ls.direction.x, ls.direction.y, ls.direction.z, ls.normal.x, ls.normal.y, ls.normal.z,
                   ls.emission.x, ls.emission.y, ls.emission.z, ls.dist, ls.pdf, 0, 0};
    std::memcpy(rec, v, sizeof(v));
    std::memcpy(seed, c.seed, 16);
    return 0;
}

int oracle_eval_envmap(const oracle_scene* s, const oracle_options* o, const float* dir, float* rgb_pdf) {
    Ray<float> r{V3<float>(0, 0, 0), V3<float>(dir[0], dir[1], dir[2])};
    V4<float> v = EvalEnvMap(s, o, r);
    rgb_pdf[0] = v.x; rgb_pdf[1] = v.y; rgb_pdf[2] = v.z; rgb_pdf[3] = v.w;
    return 0;
}

int oracle_sample_envmap(const oracle_scene* s, const oracle_options* o, uint32_t seed[4], float* dir_pdf, float* rgb) {
    Ctx<float> c{s, o, {seed[0], seed[1], seed[2], seed[3]}, nullptr};
    V3<float> col;
    V4<float> v = SampleEnvMap(c, col);
    dir_pdf[0] = v.x; dir_pdf[1] = v.y; dir_pdf[2] = v.z; dir_pdf[3] = v.w;
    rgb[0] = col.x; rgb[1] = col.y; rgb[2] = col.z;
    std::memcpy(seed, c.seed, 16);
    return 0;
}

int oracle_render_frozen(const oracle_scene* s, const float* eval_materials, const oracle_camera* cam,
                         const oracle_options* o, int32_t first_sample, int32_t num_samples, float* accum,
                         int32_t num_threads) {
    if (!s || !eval_materials || !cam || !o || !accum) return -1;
    return RenderT<float, float>(s, cam, o, first_sample, num_samples, 0, o->height, 1, accum, num_threads, nullptr,
                                 eval_materials);
}

// tonemap.glsl:59-133 for one RGBA32F accumulation image (rows bottom-up),
// then the float -> unsigned byte conversion of glGetTexImage
// (Renderer.cpp:619-634): clamp to [0, 1], round to nearest.
int oracle_tonemap(const float* accum, int32_t w, int32_t h, float inv_spp, int32_t enable_tonemap,
                   int32_t enable_aces, int32_t simple_aces, int32_t bg_mode, const float* bg, uint8_t* out) {
    if (!accum || !out || w <= 0 || h <= 0) return -1;
    auto clamp01 = [](float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); };
    auto dot3 = [](const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    // mat3 constructor columns; color * M = (dot(color, c0), dot(color, c1), dot(color, c2))
    static const float IN[3][3] = {{0.59719f, 0.35458f, 0.04823f}, {0.07600f, 0.90834f, 0.01566f},
                                   {0.02840f, 0.13383f, 0.83777f}};
    static const float OUT[3][3] = {{1.60475f, -0.53108f, -0.07367f}, {-0.10208f, 1.10813f, -0.00605f},
                                    {-0.00327f, -0.07276f, 1.07602f}};
    for (int y = 0; y < h; ++y) {
        for (int x = 0; x < w; ++x) {
            const float* a = accum + 4 * (static_cast<size_t>(y) * w + x);
            float c[3] = {a[0] * inv_spp, a[1] * inv_spp, a[2] * inv_spp};
            const float alpha = a[3] * inv_spp;
            if (enable_tonemap) {
                if (enable_aces) {
                    if (simple_aces) {
                        for (int k = 0; k < 3; ++k)
                            c[k] = clamp01((c[k] * (2.51f * c[k] + 0.03f)) / (c[k] * (2.43f * c[k] + 0.59f) + 0.14f));
                    } else {
                        float t[3] = {dot3(c, IN[0]), dot3(c, IN[1]), dot3(c, IN[2])};
                        for (int k = 0; k < 3; ++k) {
                            float av = t[k] * (t[k] + 0.0245786f) - 0.000090537f;
                            float bv = t[k] * (0.983729f * t[k] + 0.4329510f) + 0.238081f;
                            t[k] = av / bv;
                        }
                        c[0] = clamp01(dot3(t, OUT[0]));
                        c[1] = clamp01(dot3(t, OUT[1]));
                        c[2] = clamp01(dot3(t, OUT[2]));
                    }
                } else {
                    const float lum = 0.212671f * c[0] + 0.715160f * c[1] + 0.072169f * c[2];
                    const float s = 1.0f / (1.0f + lum / 1.5f);
                    for (int k = 0; k < 3; ++k) c[k] = c[k] * s;
                }
            }
            for (int k = 0; k < 3; ++k) c[k] = std::pow(c[k], 1.0f / 2.2f);
            float out_alpha = 1.0f;
            float bgc[3] = {bg[0], bg[1], bg[2]};
            if (bg_mode == 2) {
                out_alpha = alpha;
                const float fx = std::floor((static_cast<float>(x) + 0.5f) / 10.0f);
                const float fy = std::floor((static_cast<float>(y) + 0.5f) / 10.0f);
                const float m = std::fmod(fx + fy, 2.0f);
                const float res = std::max(m > 0.0f ? 1.0f : (m < 0.0f ? -1.0f : 0.0f), 0.0f);
                for (int k = 0; k < 3; ++k) bgc[k] = 0.1f * (1.0f - res) + 0.2f * res;
            }
            if (bg_mode != 0)
                for (int k = 0; k < 3; ++k) c[k] = bgc[k] * (1.0f - alpha) + c[k] * alpha;
            uint8_t* o = out + 4 * (static_cast<size_t>(y) * w + x);
            for (int k = 0; k < 3; ++k) o[k] = static_cast<uint8_t>(clamp01(c[k]) * 255.0f + 0.5f);
            o[3] = static_cast<uint8_t>(clamp01(out_alpha) * 255.0f + 0.5f);
        }
    }
    return 0;
}

}  // extern "C"
