// Device side of the path: scene layout in HBM, RNG, Disney BSDF, light and
// environment sampling, two-level BVH traversal. gfx950 only.
//
// Arithmetic follows the reference shaders expression by expression
// (src/shaders/common/*.glsl), compiled with FP contraction off, so results
// match the host transcription up to the last-ulp differences of the
// transcendental functions. Citations are file:line in the reference.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace ptd {

// ------------------------------------------------------------------ layout
// BVH2 node, 64 bytes = 4 x 16-byte loads: both children's boxes plus their
// references. A reference >= 0 is an internal node; < 0 is a leaf:
//   TLAS leaf   -1 - instance
//   BLAS leaf   -1 - (first_row << 2 | count)     (count 1..3)
struct alignas(16) GNode {
    float4 a;  // lmin.x lmin.y lmin.z lmax.x
    float4 b;  // lmax.y lmax.z rmin.x rmin.y
    float4 c;  // rmin.z rmax.x rmax.y rmax.z
    int4 d;    // left ref, right ref, unused, unused
};

// Triangle record in leaf order (vertIndices row order), 48 bytes:
// v0, e0 = v1 - v0, e1 = v2 - v0 (the subtractions the shader does per test).
struct alignas(16) GTri {
    float4 v0;
    float4 e0;
    float4 e1;
};

// Four-wide node (the reference's BVH2 collapsed: each node holds up to four
// of its descendants), for every BLAS and the TLAS: the children's boxes as
// six rows of four (min x, max x, min y, max y, min z, max z of children
// 0..3) and their references (internal: index of a GNode4; leaf: as in GNode,
// -1 - (row << 2 | count) in a BLAS and -1 - instance in the TLAS; EMPTY4: no
// child). Two 64-byte lines.
struct alignas(16) GNode4 {
    float4 lox, hix, loy, hiy, loz, hiz;
    int4 ref;
    int4 pad;
};
constexpr int EMPTY4 = static_cast<int>(0x80000000u);

// Instance: inverse affine as 4 columns (world -> object), forward linear
// part as 3 columns, BLAS root reference and material id (from the TLAS leaf).
struct alignas(16) GInst {
    float4 i0;  // ic0.x ic0.y ic0.z ic1.x
    float4 i1;  // ic1.y ic1.z ic2.x ic2.y
    float4 i2;  // ic2.z ic3.x ic3.y ic3.z
    float4 f0;  // fc0.x fc0.y fc0.z fc1.x
    float4 f1;  // fc1.y fc1.z fc2.x fc2.y
    float4 f2;  // fc2.z, blas_root (int bits), material (int bits), four-wide blas_root (int bits)
};

// Light record (src/core/Scene.h:50-59) plus the per-light constants the
// intersection tests derive (closest_hit.glsl:50-54).
struct alignas(16) GLight {
    float4 p;   // position, radius
    float4 e;   // emission, area
    float4 u;   // u, type
    float4 v;   // v, unused
    float4 n;   // normalize(cross(u, v)), dot(n, position)
    float4 uu;  // u * (1 / dot(u, u))
    float4 vv;  // v * (1 / dot(v, v))
    float4 pad;
};

struct DevScene {
    const GNode* __restrict__ nodes;
    const GNode4* __restrict__ nodes4;  // four-wide TLAS and BLAS nodes; null: BVH2 form
    const GTri* __restrict__ tris;
    const int4* __restrict__ vidx;      // vertIndices rows (x, y, z, unused)
    const float4* __restrict__ verts;   // verticesUVX
    const float4* __restrict__ norms;   // normalsUVY
    const float4* __restrict__ mats;    // 8 per material
    const GInst* __restrict__ inst;
    const GLight* __restrict__ lights;
    const uint8_t* __restrict__ tex;    // RGBA8 layers
    const float* __restrict__ env_img;  // RGB32F
    const float* __restrict__ env_cdf;
    int root;                           // TLAS root reference (in nodes4 when that is set)
    int root_bin;                       // EMPTY4, or the BVH2 TLAS root that any-hit traversals drawing
                                        // random numbers walk (nodes4 set and a material blends)
    int num_lights;
    int tex_w, tex_h, num_tex;
    int env_w, env_h;
    float env_total;
};

enum : uint32_t {
    OPT_LIGHTS = 1u << 0, OPT_RR = 1u << 1, OPT_UNIFORM_LIGHT = 1u << 2, OPT_ENVMAP = 1u << 3,
    OPT_HIDE_EMITTERS = 1u << 4, OPT_BACKGROUND = 1u << 5, OPT_TRANSPARENT_BACKGROUND = 1u << 6,
    OPT_ALPHA_TEST = 1u << 7, OPT_MEDIUM = 1u << 8, OPT_VOL_MIS = 1u << 9,
    OPT_ROUGHNESS_MOLLIFICATION = 1u << 10, OPT_OPENGL_NORMALMAP = 1u << 11,
};

struct DevOpts {
    uint32_t flags;
    int width, height, max_depth, rr_depth;
    float ucol_x, ucol_y, ucol_z;
    float env_intensity, env_rot;       // env_rot = degrees / 360
    float moll_amt;
    float cam_pos[3], cam_up[3], cam_right[3], cam_fwd[3];
    float cam_fov, cam_focal, cam_aperture;
};

// ------------------------------------------------------------------ vec3
struct vec3 { float x, y, z; };
__device__ __forceinline__ vec3 V(float x, float y, float z) { return vec3{x, y, z}; }
__device__ __forceinline__ vec3 V1(float a) { return vec3{a, a, a}; }
__device__ __forceinline__ vec3 operator+(vec3 a, vec3 b) { return V(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ vec3 operator-(vec3 a, vec3 b) { return V(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ vec3 operator*(vec3 a, vec3 b) { return V(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ vec3 operator/(vec3 a, vec3 b) { return V(a.x / b.x, a.y / b.y, a.z / b.z); }
__device__ __forceinline__ vec3 operator*(vec3 a, float s) { return V(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ vec3 operator*(float s, vec3 a) { return V(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ vec3 operator/(vec3 a, float s) { return V(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ vec3 operator-(vec3 a) { return V(-a.x, -a.y, -a.z); }
// GLSL spec min/max (y < x ? y : x); NaN behaviour equals the host transcription.
__device__ __forceinline__ float gmin(float a, float b) { return b < a ? b : a; }
__device__ __forceinline__ float gmax(float a, float b) { return a < b ? b : a; }
__device__ __forceinline__ float gclamp(float x, float lo, float hi) { return gmin(gmax(x, lo), hi); }
__device__ __forceinline__ float gmix(float x, float y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ vec3 gmix(vec3 x, vec3 y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ float dot(vec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ vec3 cross(vec3 a, vec3 b) {
    return V(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(vec3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ vec3 normalize(vec3 a) { return a / length(a); }
__device__ __forceinline__ vec3 reflect(vec3 I, vec3 N) { return I - N * (2.0f * dot(N, I)); }
__device__ __forceinline__ vec3 refract(vec3 I, vec3 N, float eta) {
    float d = dot(N, I);
    float k = 1.0f - eta * eta * (1.0f - d * d);
    if (k < 0.0f) return V1(0.0f);
    return eta * I - (eta * d + sqrtf(k)) * N;
}
__device__ __forceinline__ vec3 vexp(vec3 a) { return V(expf(a.x), expf(a.y), expf(a.z)); }
__device__ __forceinline__ vec3 vpow(vec3 a, float p) { return V(powf(a.x, p), powf(a.y, p), powf(a.z, p)); }
// pow of the texture decode (texture value ^ 2.2 in GetMaterial): evaluated in
// double and rounded once, i.e. the correctly rounded float in all but
// vanishingly rare double-rounding cases, so host and device agree bit for bit
// (GLSL leaves pow's precision to the driver).
__device__ __forceinline__ float pow_cr(float x, float p) {
    return static_cast<float>(pow(static_cast<double>(x), static_cast<double>(p)));
}
__device__ __forceinline__ vec3 vpow_cr(vec3 a, float p) { return V(pow_cr(a.x, p), pow_cr(a.y, p), pow_cr(a.z, p)); }
// sin and cos of the path code's angles (all within [-2pi, 2pi]; the
// reduction holds for |x| < 2^12). GLSL leaves their precision to the driver,
// and device and host math libraries round some arguments differently, which
// near-specular lobes turn into visible differences, so both sides use this
// one sequence of float operations: k = nearest integer to x * 2/pi;
// r = x - k * pi/2 by a four-part Cody-Waite subtraction (the first three
// parts hold 8, 12 and 12 significant bits, so k times each is exact); the
// Taylor series of sin r to r^11 and of cos r to r^12 in Horner form in r^2
// (truncation below 1e-11 for |r| <= pi/4); the quadrant k mod 4 selects
// +-sin r or +-cos r. Within 1.6 ulp of the exact values.
__host__ __device__ __forceinline__ void sincos_pt(float x, float& sn, float& cs) {
    // 2/pi; pi/2 = 0x1.92p+0 + 0x1.fb6p-12 - 0x1.778p-25 + 0x1.68c234p-39;
    // 1/3!, 1/5!, ..., 1/11! and 1/4!, 1/6!, ..., 1/12!, each rounded to float
    const float k = floorf(x * 0x1.45f306p-1f + 0.5f);
    const float r = (((x - k * 0x1.92p+0f) - k * 0x1.fb6p-12f) - k * -0x1.778p-25f) - k * 0x1.68c234p-39f;
    const float t = r * r;
    const float sp = r + r * t * (-0x1.555556p-3f + t * (0x1.111112p-7f + t * (-0x1.a01a02p-13f +
                                 t * (0x1.71de3ap-19f - t * 0x1.ae6456p-26f))));
    const float cp = (1.0f - 0.5f * t) + t * t * (0x1.555556p-5f + t * (-0x1.6c16c2p-10f + t * (0x1.a01a02p-16f +
                                                  t * (-0x1.27e4fcp-22f + t * 0x1.1eed8ep-29f))));
    const int q = static_cast<int>(k) & 3;
    sn = q == 0 ? sp : q == 1 ? cp : q == 2 ? -sp : -cp;
    cs = q == 0 ? cp : q == 1 ? -sp : q == 2 ? -cp : sp;
}
__host__ __device__ __forceinline__ float sin_pt(float x) { float s, c; sincos_pt(x, s, c); return s; }
__host__ __device__ __forceinline__ float cos_pt(float x) { float s, c; sincos_pt(x, s, c); return c; }
__device__ __forceinline__ vec3 xyz(float4 a) { return V(a.x, a.y, a.z); }
__device__ __forceinline__ float4 F4(vec3 a, float w) { return make_float4(a.x, a.y, a.z, w); }

// ------------------------------------------------------------------ constants (globals.glsl:25-44)
constexpr float PI = 3.14159265358979323f;
constexpr float INV_PI = 0.31830988618379067f;
constexpr float TWO_PI = 6.28318530717958648f;
constexpr float INV_TWO_PI = 0.15915494309189533f;
constexpr float INV_4_PI = 0.07957747154594766f;
constexpr float EPS = 0.0003f;
constexpr float INF = 1000000.0f;
enum { QUAD_LIGHT = 0, SPHERE_LIGHT = 1, DISTANT_LIGHT = 2 };
enum { ALPHA_MODE_OPAQUE = 0, ALPHA_MODE_BLEND = 1, ALPHA_MODE_MASK = 2 };
enum { MEDIUM_NONE = 0, MEDIUM_ABSORB = 1, MEDIUM_SCATTER = 2, MEDIUM_EMISSIVE = 3 };

// ------------------------------------------------------------------ RNG (globals.glsl:149-166)
struct Rng { uint32_t x, y, z, w; };
__device__ __forceinline__ void pcg4d(Rng& v) {
    v.x = v.x * 1664525u + 1013904223u; v.y = v.y * 1664525u + 1013904223u;
    v.z = v.z * 1664525u + 1013904223u; v.w = v.w * 1664525u + 1013904223u;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
    v.x ^= v.x >> 16u; v.y ^= v.y >> 16u; v.z ^= v.z >> 16u; v.w ^= v.w >> 16u;
    v.x += v.y * v.w; v.y += v.z * v.x; v.z += v.x * v.y; v.w += v.y * v.z;
}
// float(seed.x) / float(0xffffffff); the divisor is 2^32 in float, so [0, 1].
__device__ __forceinline__ float rnd(Rng& r) {
    pcg4d(r);
    return static_cast<float>(r.x) / 4294967296.0f;
}
__device__ __forceinline__ Rng init_rng(uint32_t px, uint32_t py, uint32_t frame) {
    return Rng{px, py, frame, px + py};
}

__device__ __forceinline__ float Luminance(vec3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// ------------------------------------------------------------------ structs (globals.glsl:52-139)
struct Medium { int type; float density; vec3 color; float anisotropy; };
struct Material {
    vec3 baseColor; float opacity; int alphaMode; float alphaCutoff;
    vec3 emission; float anisotropic, metallic, roughness, subsurface, specularTint;
    float sheen, sheenTint, clearcoat, clearcoatRoughness, specTrans, ior, ax, ay;
    Medium medium;
};
struct SurfState {
    vec3 fhp, normal, ffnormal, tangent, bitangent;
    float tu, tv;  // texCoord
    int matID;
};
struct LightSample { vec3 normal, emission, direction; float dist, pdf; };

// ------------------------------------------------------------------ textures
// GL_LINEAR + GL_REPEAT per the GL specification (8.14.2), full-precision weights.
__device__ __forceinline__ int wrapi(int i, int n) { int r = i % n; return r < 0 ? r + n : r; }
__device__ float4 tex_layer(const DevScene& sc, float u, float v, float layer) {
    int l = static_cast<int>(floorf(layer + 0.5f));
    l = l < 0 ? 0 : (l > sc.num_tex - 1 ? sc.num_tex - 1 : l);
    const int W = sc.tex_w, H = sc.tex_h;
    const uint8_t* base = sc.tex + static_cast<size_t>(l) * 4 * W * H;
    float x = u * static_cast<float>(W) - 0.5f, y = v * static_cast<float>(H) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
    int xs0 = wrapi(i0, W), xs1 = wrapi(i0 + 1, W), ys0 = wrapi(j0, H), ys1 = wrapi(j0 + 1, H);
    float w[4] = {(1.0f - ax) * (1.0f - ay), ax * (1.0f - ay), (1.0f - ax) * ay, ax * ay};
    int xx[4] = {xs0, xs1, xs0, xs1}, yy[4] = {ys0, ys0, ys1, ys1};
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
    for (int k = 0; k < 4; ++k) {
        const uint8_t* t = base + (static_cast<size_t>(yy[k]) * W + xx[k]) * 4;
        for (int c = 0; c < 4; ++c) acc[c] = acc[c] + w[k] * (static_cast<float>(t[c]) / 255.0f);
    }
    return make_float4(acc[0], acc[1], acc[2], acc[3]);
}
__device__ vec3 env_lookup(const DevScene& sc, float u, float v) {
    const int W = sc.env_w, H = sc.env_h;
    float x = u * static_cast<float>(W) - 0.5f, y = v * static_cast<float>(H) - 0.5f;
    float fx = floorf(x), fy = floorf(y);
    float ax = x - fx, ay = y - fy;
    int i0 = static_cast<int>(fx), j0 = static_cast<int>(fy);
    int xs0 = wrapi(i0, W), xs1 = wrapi(i0 + 1, W), ys0 = wrapi(j0, H), ys1 = wrapi(j0 + 1, H);
    float w[4] = {(1.0f - ax) * (1.0f - ay), ax * (1.0f - ay), (1.0f - ax) * ay, ax * ay};
    int xx[4] = {xs0, xs1, xs0, xs1}, yy[4] = {ys0, ys0, ys1, ys1};
    float acc[3] = {0.f, 0.f, 0.f};
    for (int k = 0; k < 4; ++k) {
        const float* t = sc.env_img + (static_cast<size_t>(yy[k]) * W + xx[k]) * 3;
        for (int c = 0; c < 3; ++c) acc[c] = acc[c] + w[k] * t[c];
    }
    return V(acc[0], acc[1], acc[2]);
}

// ------------------------------------------------------------------ sampling.glsl:25-277
__device__ float GTR1(float NDotH, float a) {
    if (a >= 1.0f) return INV_PI;
    float a2 = a * a;
    float t = 1.0f + (a2 - 1.0f) * NDotH * NDotH;
    return (a2 - 1.0f) / (PI * logf(a2) * t);
}
// clamp(sqrt(1 - c * c), 0, 1) of SampleGTR1 and SampleHG (sampling.glsl). When
// rounding leaves |c| a hair above 1, GLSL's sqrt is undefined and a GPU's
// max/min return the non-NaN operand, so the reference's clamp gives 0; with
// the spec's max (x < y ? y : x) the NaN would reach the pixel. Both sides
// take 0 there; elsewhere the value is the clamped root bit for bit.
__device__ __forceinline__ float sin_from_cos(float c) {
    const float s2 = 1.0f - (c * c);
    return s2 > 0.0f ? gmin(sqrtf(s2), 1.0f) : 0.0f;
}
__device__ vec3 SampleGTR1(float rgh, float r1, float r2) {
    float a = gmax(0.001f, rgh);
    float a2 = a * a;
    float phi = r1 * TWO_PI;
    float cosTheta = sqrtf((1.0f - powf(a2, 1.0f - r2)) / (1.0f - a2));
    float sinTheta = sin_from_cos(cosTheta);
    float sinPhi = sin_pt(phi), cosPhi = cos_pt(phi);
    return V(sinTheta * cosPhi, sinTheta * sinPhi, cosTheta);
}
__device__ vec3 SampleGGXVNDF(vec3 Vv, float ax, float ay, float r1, float r2) {
    vec3 Vh = normalize(V(ax * Vv.x, ay * Vv.y, Vv.z));
    float lensq = Vh.x * Vh.x + Vh.y * Vh.y;
    vec3 T1 = lensq > 0.0f ? V(-Vh.y, Vh.x, 0.0f) * (1.0f / sqrtf(lensq)) : V(1.0f, 0.0f, 0.0f);
    vec3 T2 = cross(Vh, T1);
    float r = sqrtf(r1);
    float phi = 2.0f * PI * r2;
    float t1 = r * cos_pt(phi);
    float t2 = r * sin_pt(phi);
    float s = 0.5f * (1.0f + Vh.z);
    t2 = (1.0f - s) * sqrtf(1.0f - t1 * t1) + s * t2;
    vec3 Nh = t1 * T1 + t2 * T2 + sqrtf(gmax(0.0f, 1.0f - t1 * t1 - t2 * t2)) * Vh;
    return normalize(V(ax * Nh.x, ay * Nh.y, gmax(0.0f, Nh.z)));
}
__device__ __forceinline__ float GTR2Aniso(float NDotH, float HDotX, float HDotY, float ax, float ay) {
    float a = HDotX / ax;
    float b = HDotY / ay;
    float c = a * a + b * b + NDotH * NDotH;
    return 1.0f / (PI * ax * ay * c * c);
}
__device__ __forceinline__ float SmithG(float NDotV, float alphaG) {
    float a = alphaG * alphaG;
    float b = NDotV * NDotV;
    return (2.0f * NDotV) / (NDotV + sqrtf(a + b - a * b));
}
__device__ __forceinline__ float SmithGAniso(float NDotV, float VDotX, float VDotY, float ax, float ay) {
    float a = VDotX * ax;
    float b = VDotY * ay;
    float c = NDotV;
    return (2.0f * NDotV) / (NDotV + sqrtf(a * a + b * b + c * c));
}
__device__ __forceinline__ float SchlickWeight(float u) {
    float m = gclamp(1.0f - u, 0.0f, 1.0f);
    float m2 = m * m;
    return m2 * m2 * m;
}
__device__ float DielectricFresnel(float cosThetaI, float eta) {
    float sinThetaTSq = eta * eta * (1.0f - cosThetaI * cosThetaI);
    if (sinThetaTSq > 1.0f) return 1.0f;
    float cosThetaT = sqrtf(gmax(1.0f - sinThetaTSq, 0.0f));
    float rs = (eta * cosThetaT - cosThetaI) / (eta * cosThetaT + cosThetaI);
    float rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    return 0.5f * (rs * rs + rp * rp);
}
__device__ vec3 CosineSampleHemisphere(float r1, float r2) {
    vec3 dir;
    float r = sqrtf(r1);
    float phi = TWO_PI * r2;
    dir.x = r * cos_pt(phi);
    dir.y = r * sin_pt(phi);
    dir.z = sqrtf(gmax(0.0f, 1.0f - dir.x * dir.x - dir.y * dir.y));
    return dir;
}
__device__ vec3 UniformSampleHemisphere(float r1, float r2) {
    float r = sqrtf(gmax(0.0f, 1.0f - r1 * r1));
    float phi = TWO_PI * r2;
    return V(r * cos_pt(phi), r * sin_pt(phi), r1);
}
__device__ __forceinline__ float PowerHeuristic(float a, float b) {
    float t = a * a;
    return t / (b * b + t);
}
__device__ __forceinline__ void Onb(vec3 N, vec3& T, vec3& B) {
    vec3 up = fabsf(N.z) < 0.9999999f ? V(0.0f, 0.0f, 1.0f) : V(1.0f, 0.0f, 0.0f);
    T = normalize(cross(up, N));
    B = cross(N, T);
}
__device__ vec3 SampleHG(vec3 Vv, float g, float r1, float r2) {
    float cosTheta;
    if (fabsf(g) < 0.001f) cosTheta = 1.0f - 2.0f * r2;
    else {
        float sqrTerm = (1.0f - g * g) / (1.0f + g - 2.0f * g * r2);
        cosTheta = -(1.0f + g * g - sqrTerm * sqrTerm) / (2.0f * g);
    }
    float phi = r1 * TWO_PI;
    float sinTheta = sin_from_cos(cosTheta);
    float sinPhi = sin_pt(phi), cosPhi = cos_pt(phi);
    vec3 v1, v2;
    Onb(Vv, v1, v2);
    return sinTheta * cosPhi * v1 + sinTheta * sinPhi * v2 + cosTheta * Vv;
}
__device__ __forceinline__ float PhaseHG(float cosTheta, float g) {
    float denom = 1.0f + g * g + 2.0f * g * cosTheta;
    return INV_4_PI * (1.0f - g * g) / (denom * sqrtf(denom));
}

// Light sampling (sampling.glsl:186-248)
__device__ void SampleOneLight(const GLight& L, vec3 scatterPos, int numLights, Rng& rng, LightSample& ls) {
    int type = static_cast<int>(L.u.w);
    vec3 pos = xyz(L.p);
    if (type == QUAD_LIGHT) {
        float r1 = rnd(rng);
        float r2 = rnd(rng);
        vec3 lightSurfacePos = pos + xyz(L.u) * r1 + xyz(L.v) * r2;
        ls.direction = lightSurfacePos - scatterPos;
        ls.dist = length(ls.direction);
        float distSq = ls.dist * ls.dist;
        ls.direction = ls.direction / ls.dist;
        ls.normal = normalize(cross(xyz(L.u), xyz(L.v)));
        ls.emission = xyz(L.e) * static_cast<float>(numLights);
        ls.pdf = distSq / (L.e.w * fabsf(dot(ls.normal, ls.direction)));
    } else if (type == SPHERE_LIGHT) {
        float r1 = rnd(rng);
        float r2 = rnd(rng);
        vec3 sc = scatterPos - pos;
        float distToSphereCenter = length(sc);
        sc = sc / distToSphereCenter;
        vec3 sampledDir = UniformSampleHemisphere(r1, r2);
        vec3 T, B;
        Onb(sc, T, B);
        sampledDir = T * sampledDir.x + B * sampledDir.y + sc * sampledDir.z;
        vec3 lightSurfacePos = pos + sampledDir * L.p.w;
        ls.direction = lightSurfacePos - scatterPos;
        ls.dist = length(ls.direction);
        float distSq = ls.dist * ls.dist;
        ls.direction = ls.direction / ls.dist;
        ls.normal = normalize(lightSurfacePos - pos);
        ls.emission = xyz(L.e) * static_cast<float>(numLights);
        ls.pdf = distSq / (L.e.w * 0.5f * fabsf(dot(ls.normal, ls.direction)));
    } else {
        ls.direction = normalize(pos - V1(0.0f));
        ls.normal = normalize(scatterPos - pos);
        ls.emission = xyz(L.e) * static_cast<float>(numLights);
        ls.dist = INF;
        ls.pdf = 1.0f;
    }
}

// ------------------------------------------------------------------ envmap.glsl:28-83
__device__ void BinarySearch(const DevScene& sc, float value, float& u, float& v) {
    const int W = sc.env_w, H = sc.env_h;
    int lower = 0, upper = H - 1;
    while (lower < upper) {
        int mid = (lower + upper) >> 1;
        if (value < sc.env_cdf[mid * W + (W - 1)]) upper = mid;
        else lower = mid + 1;
    }
    int y = lower < 0 ? 0 : (lower > H - 1 ? H - 1 : lower);
    lower = 0; upper = W - 1;
    while (lower < upper) {
        int mid = (lower + upper) >> 1;
        if (value < sc.env_cdf[y * W + mid]) upper = mid;
        else lower = mid + 1;
    }
    int x = lower < 0 ? 0 : (lower > W - 1 ? W - 1 : lower);
    u = static_cast<float>(x) / static_cast<float>(W);
    v = static_cast<float>(y) / static_cast<float>(H);
}
__device__ float4 EvalEnvMap(const DevScene& sc, const DevOpts& op, vec3 dir) {
    float theta = acosf(gclamp(dir.y, -1.0f, 1.0f));
    float u = (PI + atan2f(dir.z, dir.x)) * INV_TWO_PI + op.env_rot;
    float v = theta * INV_PI;
    vec3 color = env_lookup(sc, u, v);
    float pdf = Luminance(color) / sc.env_total;
    return F4(color, (pdf * static_cast<float>(sc.env_w) * static_cast<float>(sc.env_h)) / (TWO_PI * PI * sin_pt(theta)));
}
__device__ float4 SampleEnvMap(const DevScene& sc, const DevOpts& op, Rng& rng, vec3& color) {
    float u, v;
    BinarySearch(sc, rnd(rng) * sc.env_total, u, v);
    color = env_lookup(sc, u, v);
    float pdf = Luminance(color) / sc.env_total;
    u = u - op.env_rot;
    float phi = u * TWO_PI;
    float theta = v * PI;
    if (sin_pt(theta) == 0.0f) pdf = 0.0f;
    return make_float4(-sin_pt(theta) * cos_pt(phi), cos_pt(theta), -sin_pt(theta) * sin_pt(phi),
                       (pdf * static_cast<float>(sc.env_w) * static_cast<float>(sc.env_h)) / (TWO_PI * PI * sin_pt(theta)));
}

// ------------------------------------------------------------------ disney.glsl:37-351
__device__ __forceinline__ vec3 ToWorld(vec3 X, vec3 Y, vec3 Z, vec3 v) { return v.x * X + v.y * Y + v.z * Z; }
__device__ __forceinline__ vec3 ToLocal(vec3 X, vec3 Y, vec3 Z, vec3 v) { return V(dot(v, X), dot(v, Y), dot(v, Z)); }

struct Lobes { float diffPr, dielectricPr, metalPr, glassPr, clearCtPr, dielectricWt, metalWt, glassWt, F0; vec3 Csheen, Cspec0; };

__device__ __forceinline__ Lobes LobeWeights(const Material& m, float eta, float Vz) {
    Lobes p;
    float lum = Luminance(m.baseColor);
    vec3 ctint = lum > 0.0f ? m.baseColor / lum : V1(1.0f);
    p.F0 = (1.0f - eta) / (1.0f + eta);
    p.F0 = p.F0 * p.F0;
    p.Cspec0 = p.F0 * gmix(V1(1.0f), ctint, m.specularTint);
    p.Csheen = gmix(V1(1.0f), ctint, m.sheenTint);
    p.dielectricWt = (1.0f - m.metallic) * (1.0f - m.specTrans);
    p.metalWt = m.metallic;
    p.glassWt = (1.0f - m.metallic) * m.specTrans;
    float schlickWt = SchlickWeight(Vz);
    p.diffPr = p.dielectricWt * Luminance(m.baseColor);
    p.dielectricPr = p.dielectricWt * Luminance(gmix(p.Cspec0, V1(1.0f), schlickWt));
    p.metalPr = p.metalWt * Luminance(gmix(m.baseColor, V1(1.0f), schlickWt));
    p.glassPr = p.glassWt;
    p.clearCtPr = 0.25f * m.clearcoat;
    float invTotalWt = 1.0f / (p.diffPr + p.dielectricPr + p.metalPr + p.glassPr + p.clearCtPr);
    p.diffPr = p.diffPr * invTotalWt;
    p.dielectricPr = p.dielectricPr * invTotalWt;
    p.metalPr = p.metalPr * invTotalWt;
    p.glassPr = p.glassPr * invTotalWt;
    p.clearCtPr = p.clearCtPr * invTotalWt;
    return p;
}

__device__ vec3 EvalMicrofacetReflection(const Material& m, vec3 Vv, vec3 L, vec3 H, vec3 F, float& pdf) {
    pdf = 0.0f;
    if (L.z <= 0.0f) return V1(0.0f);
    float D = GTR2Aniso(H.z, H.x, H.y, m.ax, m.ay);
    float G1 = SmithGAniso(fabsf(Vv.z), Vv.x, Vv.y, m.ax, m.ay);
    float G2 = G1 * SmithGAniso(fabsf(L.z), L.x, L.y, m.ax, m.ay);
    pdf = G1 * D / (4.0f * Vv.z);
    return F * D * G2 / (4.0f * L.z * Vv.z);
}

__device__ vec3 DisneyEval(const Material& m, float eta, vec3 Vv, vec3 N, vec3 L, float& pdf) {
    pdf = 0.0f;
    vec3 f = V1(0.0f);
    vec3 T, B;
    Onb(N, T, B);
    Vv = ToLocal(T, B, N, Vv);
    L = ToLocal(T, B, N, L);
    vec3 H;
    if (L.z > 0.0f) H = normalize(L + Vv);
    else H = normalize(L + Vv * eta);
    if (H.z < 0.0f) H = -H;
    Lobes p = LobeWeights(m, eta, Vv.z);
    bool refl = L.z * Vv.z > 0.0f;
    float tmpPdf = 0.0f;
    float VDotH = fabsf(dot(Vv, H));
    if (p.diffPr > 0.0f && refl) {
        // EvalDisneyDiffuse (disney.glsl:61-88); L.z > 0 here since refl and V.z > 0 or both < 0
        vec3 fd = V1(0.0f);
        tmpPdf = 0.0f;
        if (L.z > 0.0f) {
            float LDotH = dot(L, H);
            float Rr = 2.0f * m.roughness * LDotH * LDotH;
            float FL = SchlickWeight(L.z);
            float FV = SchlickWeight(Vv.z);
            float Fretro = Rr * (FL + FV + FL * FV * (Rr - 1.0f));
            float Fd = (1.0f - 0.5f * FL) * (1.0f - 0.5f * FV);
            float Fss90 = 0.5f * Rr;
            float Fss = gmix(1.0f, Fss90, FL) * gmix(1.0f, Fss90, FV);
            float ss = 1.25f * (Fss * (1.0f / (L.z + Vv.z) - 0.5f) + 0.5f);
            float FH = SchlickWeight(LDotH);
            vec3 Fsheen = FH * m.sheen * p.Csheen;
            tmpPdf = L.z * INV_PI;
            fd = INV_PI * m.baseColor * gmix(Fd + Fretro, ss, m.subsurface) + Fsheen;
        }
        f = f + fd * p.dielectricWt;
        pdf = pdf + tmpPdf * p.diffPr;
    }
    if (p.dielectricPr > 0.0f && refl) {
        float F = (DielectricFresnel(VDotH, 1.0f / m.ior) - p.F0) / (1.0f - p.F0);
        f = f + EvalMicrofacetReflection(m, Vv, L, H, gmix(p.Cspec0, V1(1.0f), F), tmpPdf) * p.dielectricWt;
        pdf = pdf + tmpPdf * p.dielectricPr;
    }
    if (p.metalPr > 0.0f && refl) {
        vec3 F = gmix(m.baseColor, V1(1.0f), SchlickWeight(VDotH));
        f = f + EvalMicrofacetReflection(m, Vv, L, H, F, tmpPdf) * p.metalWt;
        pdf = pdf + tmpPdf * p.metalPr;
    }
    if (p.glassPr > 0.0f) {
        float F = DielectricFresnel(VDotH, eta);
        if (refl) {
            f = f + EvalMicrofacetReflection(m, Vv, L, H, V1(F), tmpPdf) * p.glassWt;
            pdf = pdf + tmpPdf * p.glassPr * F;
        } else {
            // EvalMicrofacetRefraction (disney.glsl:104-123)
            vec3 fr = V1(0.0f);
            tmpPdf = 0.0f;
            if (!(L.z >= 0.0f)) {
                float LDotH = dot(L, H);
                float VDotH2 = dot(Vv, H);
                float D = GTR2Aniso(H.z, H.x, H.y, m.ax, m.ay);
                float G1 = SmithGAniso(fabsf(Vv.z), Vv.x, Vv.y, m.ax, m.ay);
                float G2 = G1 * SmithGAniso(fabsf(L.z), L.x, L.y, m.ax, m.ay);
                float denom = LDotH + VDotH2 * eta;
                denom = denom * denom;
                float eta2 = eta * eta;
                float jacobian = fabsf(LDotH) / denom;
                tmpPdf = G1 * gmax(0.0f, VDotH2) * D * jacobian / Vv.z;
                fr = vpow(m.baseColor, 0.5f) * (V1(1.0f) - V1(F)) * D * G2 * fabsf(VDotH2) * jacobian * eta2 /
                     fabsf(L.z * Vv.z);
            }
            f = f + fr * p.glassWt;
            pdf = pdf + tmpPdf * p.glassPr * (1.0f - F);
        }
    }
    if (p.clearCtPr > 0.0f && refl) {
        // EvalClearcoat (disney.glsl:125-140)
        vec3 fc = V1(0.0f);
        tmpPdf = 0.0f;
        if (L.z > 0.0f) {
            float VDotH2 = dot(Vv, H);
            float F = gmix(0.04f, 1.0f, SchlickWeight(VDotH2));
            float D = GTR1(H.z, m.clearcoatRoughness);
            float G = SmithG(L.z, 0.25f) * SmithG(Vv.z, 0.25f);
            float jacobian = 1.0f / (4.0f * VDotH2);
            tmpPdf = D * H.z * jacobian;
            fc = V1(F) * D * G;
        }
        f = f + fc * 0.25f * m.clearcoat;
        pdf = pdf + tmpPdf * p.clearCtPr;
    }
    return f * fabsf(L.z);
}

__device__ vec3 DisneySample(const Material& m, float eta, vec3 Vv, vec3 N, Rng& rng, vec3& L, float& pdf) {
    pdf = 0.0f;
    float r1 = rnd(rng);
    float r2 = rnd(rng);
    vec3 T, B;
    Onb(N, T, B);
    Vv = ToLocal(T, B, N, Vv);
    Lobes p = LobeWeights(m, eta, Vv.z);
    float cdf0 = p.diffPr;
    float cdf1 = cdf0 + p.dielectricPr;
    float cdf2 = cdf1 + p.metalPr;
    float cdf3 = cdf2 + p.glassPr;
    float r3 = rnd(rng);
    if (r3 < cdf0) {
        L = CosineSampleHemisphere(r1, r2);
    } else if (r3 < cdf2) {
        vec3 H = SampleGGXVNDF(Vv, m.ax, m.ay, r1, r2);
        if (H.z < 0.0f) H = -H;
        L = normalize(reflect(-Vv, H));
    } else if (r3 < cdf3) {
        vec3 H = SampleGGXVNDF(Vv, m.ax, m.ay, r1, r2);
        float F = DielectricFresnel(fabsf(dot(Vv, H)), eta);
        if (H.z < 0.0f) H = -H;
        r3 = (r3 - cdf2) / (cdf3 - cdf2);
        if (r3 < F) L = normalize(reflect(-Vv, H));
        else L = normalize(refract(-Vv, H, eta));
    } else {
        vec3 H = SampleGTR1(m.clearcoatRoughness, r1, r2);
        if (H.z < 0.0f) H = -H;
        L = normalize(reflect(-Vv, H));
    }
    L = ToWorld(T, B, N, L);
    Vv = ToWorld(T, B, N, Vv);
    return DisneyEval(m, eta, Vv, N, L, pdf);
}

// ------------------------------------------------------------------ intersection.glsl:25-66
__device__ __forceinline__ float SphereIntersect(float rad, vec3 pos, vec3 ro, vec3 rd) {
    vec3 op = pos - ro;
    float eps = 0.001f;
    float b = dot(op, rd);
    float det = b * b - dot(op, op) + rad * rad;
    if (det < 0.0f) return INF;
    det = sqrtf(det);
    float t1 = b - det;
    if (t1 > eps) return t1;
    float t2 = b + det;
    if (t2 > eps) return t2;
    return INF;
}
__device__ __forceinline__ float RectIntersect(vec3 pos, vec3 u, vec3 v, float4 plane, vec3 ro, vec3 rd) {
    vec3 n = xyz(plane);
    float dt = dot(rd, n);
    float t = (plane.w - dot(n, ro)) / dt;
    if (t > EPS) {
        vec3 p = ro + rd * t;
        vec3 vi = p - pos;
        float a1 = dot(u, vi);
        if (a1 >= 0.0f && a1 <= 1.0f) {
            float a2 = dot(v, vi);
            if (a2 >= 0.0f && a2 <= 1.0f) return t;
        }
    }
    return INF;
}

// ------------------------------------------------------------------ material (pathtrace.glsl:25-115)
__device__ Material GetMaterial(const DevScene& sc, const DevOpts& op, SurfState& st, vec3 rd, int depth,
                                float prev_roughness, float& eta) {
    const float4* p = sc.mats + 8 * st.matID;
    float4 p1 = p[0], p2 = p[1], p3 = p[2], p4 = p[3], p5 = p[4], p6 = p[5], p7 = p[6], p8 = p[7];
    Material m;
    m.baseColor = xyz(p1);
    m.anisotropic = p1.w;
    m.emission = xyz(p2);
    m.metallic = p3.x;
    m.roughness = gmax(p3.y, 0.001f);
    m.subsurface = p3.z;
    m.specularTint = p3.w;
    m.sheen = p4.x;
    m.sheenTint = p4.y;
    m.clearcoat = p4.z;
    m.clearcoatRoughness = gmix(0.1f, 0.001f, p4.w);
    m.specTrans = p5.x;
    m.ior = p5.y;
    m.medium.type = static_cast<int>(p5.z);
    m.medium.density = p5.w;
    m.medium.color = xyz(p6);
    m.medium.anisotropy = gclamp(p6.w, -0.9f, 0.9f);
    int texX = static_cast<int>(p7.x), texY = static_cast<int>(p7.y), texZ = static_cast<int>(p7.z),
        texW = static_cast<int>(p7.w);
    m.opacity = p8.x;
    m.alphaMode = static_cast<int>(p8.y);
    m.alphaCutoff = p8.z;
    if (sc.num_tex > 0) {
        if (texX >= 0) {
            float4 col = tex_layer(sc, st.tu, st.tv, static_cast<float>(texX));
            m.baseColor = m.baseColor * vpow_cr(V(col.x, col.y, col.z), 2.2f);
            m.opacity = m.opacity * col.w;
        }
        if (texY >= 0) {
            float4 t4 = tex_layer(sc, st.tu, st.tv, static_cast<float>(texY));
            m.metallic = t4.z;
            m.roughness = gmax(t4.y * t4.y, 0.001f);
        }
        if (texZ >= 0) {
            float4 t4 = tex_layer(sc, st.tu, st.tv, static_cast<float>(texZ));
            vec3 tn = V(t4.x, t4.y, t4.z);
            if (op.flags & OPT_OPENGL_NORMALMAP) tn.y = 1.0f - tn.y;
            tn = normalize(tn * 2.0f - V1(1.0f));
            vec3 orig = st.normal;
            st.normal = normalize(st.tangent * tn.x + st.bitangent * tn.y + st.normal * tn.z);
            st.ffnormal = dot(orig, rd) <= 0.0f ? st.normal : -st.normal;
        }
    }
    if ((op.flags & OPT_ROUGHNESS_MOLLIFICATION) && depth > 0)
        m.roughness = gmax(gmix(0.0f, prev_roughness, op.moll_amt), m.roughness);
    if (sc.num_tex > 0 && texW >= 0) {
        float4 t4 = tex_layer(sc, st.tu, st.tv, static_cast<float>(texW));
        m.emission = vpow_cr(V(t4.x, t4.y, t4.z), 2.2f);
    }
    float aspect = sqrtf(1.0f - m.anisotropic * 0.9f);
    m.ax = gmax(0.001f, m.roughness / aspect);
    m.ay = gmax(0.001f, m.roughness * aspect);
    eta = dot(rd, st.normal) < 0.0f ? (1.0f / m.ior) : m.ior;
    return m;
}

// ------------------------------------------------------------------ traversal
struct Hit {
    float t, u, v;
    int tri;    // vertIndices row, -1 if none
    int inst;   // instance of the triangle
    int light;  // light index when a light is the closest hit, else -1
};

struct TravCount { uint32_t nodes, tris, insts, lights, spills, rays; };

// Traversal stack: entries [0, S) in LDS (entry i of lane tid at lds[i*bs + tid],
// conflict-free), deeper entries in a global backing store indexed by the
// thread's slot. MAXD bounds the total depth; deeper trees set `overflow`.
// The reference's binary traversal keeps 64 entries (closest_hit.glsl:89);
// four-wide nodes push up to three entries per level over half as many
// levels, so this stack has half as much room again.
constexpr int MAXD = 96;

struct Stack {
    int* lds;          // this block's LDS array
    int* spill;        // global backing store (MAXD - S entries per thread slot)
    int lane;          // threadIdx.x
    int bs;            // blockDim.x
    int S;             // LDS entries
    size_t gslot;      // this thread's slot in the backing store
    size_t gstride;    // number of thread slots
    __device__ __forceinline__ void push(int sp, int v) {
        if (sp < S) lds[sp * bs + lane] = v;
        else spill[static_cast<size_t>(sp - S) * gstride + gslot] = v;
    }
    // The backing-store read is a relaxed atomic load so that the compiler
    // keeps the two reads apart: merged, they become one generic (flat) load,
    // which puts the LDS pop on the vector-memory path and its wait counter.
    __device__ __forceinline__ int pop(int sp) {
        if (sp < S) return lds[sp * bs + lane];
        return __hip_atomic_load(spill + static_cast<size_t>(sp - S) * gstride + gslot, __ATOMIC_RELAXED,
                                 __HIP_MEMORY_SCOPE_WAVEFRONT);
    }
};

__device__ __forceinline__ void inst_to_object(const GInst& I, vec3 o, vec3 d, vec3& oo, vec3& od) {
    vec3 c0 = V(I.i0.x, I.i0.y, I.i0.z), c1 = V(I.i0.w, I.i1.x, I.i1.y);
    vec3 c2 = V(I.i1.z, I.i1.w, I.i2.x), c3 = V(I.i2.y, I.i2.z, I.i2.w);
    oo = ((c0 * o.x + c1 * o.y) + c2 * o.z) + c3;
    od = (c0 * d.x + c1 * d.y) + c2 * d.z;
}

// Reference slab test (intersection.glsl:68-82) returning the ordering key
// (t0 if t0 > 0 else t1; -1 on a miss) and the entry distance t0.
__device__ __forceinline__ float slab(float mnx, float mny, float mnz, float mxx, float mxy, float mxz, vec3 o,
                                      vec3 inv, float& t0) {
    float fx = (mxx - o.x) * inv.x, fy = (mxy - o.y) * inv.y, fz = (mxz - o.z) * inv.z;
    float nx = (mnx - o.x) * inv.x, ny = (mny - o.y) * inv.y, nz = (mnz - o.z) * inv.z;
    float t1 = fminf(fmaxf(fx, nx), fminf(fmaxf(fy, ny), fmaxf(fz, nz)));
    t0 = fmaxf(fminf(fx, nx), fmaxf(fminf(fy, ny), fminf(fz, nz)));
    return (t1 >= t0) ? (t0 > 0.0f ? t0 : t1) : -1.0f;
}

// ClosestHit (closest_hit.glsl:25-218) / AnyHit (anyhit.glsl:25-213) over the
// BVH2 re-layout. Visit order equals the reference (nearer child first by the
// same key, farther pushed, LIFO). Subtrees whose entry distance exceeds the
// current best are skipped; that never changes which primitive wins except for
// exact ties decided by rounding. ANY returns at the first accepted triangle
// with t < tmax.
// Alpha test of an any-hit candidate (anyhit.glsl:118-140): the instance
// material's opacity times the base-colour texture's alpha at the hit
// (texture layer texIDs.x, clamped like a GL array texture; alpha 1 without
// textures). MASK: transparent below the cutoff; BLEND: transparent with
// probability 1 - opacity, drawing from the path's RNG (null: opaque).
__device__ __forceinline__ bool alpha_transparent(const DevScene& sc, int row, int mat, float u, float v, float w,
                                                  Rng* rng) {
    const int4 vi = sc.vidx[row];
    const float4 p0 = sc.verts[vi.x], p1 = sc.verts[vi.y], p2 = sc.verts[vi.z];
    const float q0 = sc.norms[vi.x].w, q1 = sc.norms[vi.y].w, q2 = sc.norms[vi.z].w;
    const float tu = p0.w * w + p1.w * u + p2.w * v;
    const float tv = q0 * w + q1 * u + q2 * v;
    const float4 texIDs = sc.mats[8 * mat + 6];
    const float4 ap = sc.mats[8 * mat + 7];
    const float a = sc.num_tex > 0 ? tex_layer(sc, tu, tv, texIDs.x).w : 1.0f;
    const float opacity = ap.x * a;
    const int mode = static_cast<int>(ap.y);
    return (mode == ALPHA_MODE_MASK && opacity < ap.z) || (mode == ALPHA_MODE_BLEND && rng && rnd(*rng) > opacity);
}

// Position of a closest-hit traversal between two launches: the node or leaf
// it processes next (ref), the stack depth (entries [0, sp) are on the
// Stack), the stack depth at which the current instance's BLAS began (-1 in
// the TLAS) and that instance. {root, 0, -1, -1} starts a traversal.
struct TravState { int ref, sp, blas_base, cur_inst; };

// Two-level traversal. ANY: return at the first accepted hit closer than
// tmax; with `alpha` set, candidates are alpha-tested first (AnyHit with
// OPT_ALPHA_TEST and no media).
// RESUME (closest hit only): start from *ts (with the stack entries and h
// holding the best hit so far, tmax = h.t) and stop after `budget` loop
// iterations (budget < 0: no limit); returns true when the traversal ended
// and false when it stopped, its position saved in *ts. The iterations and
// their arithmetic are the same however the traversal is split: the object-
// space ray of the current instance is recomputed from the same record.
// UNORDERED (any hit without random draws): the children a node's boxes hit
// are visited in child order instead of nearest first. Whether anything is
// hit does not depend on the order; what would is a stochastic alpha test,
// whose draws follow the visit order, so callers that pass an RNG keep the
// reference's order.
template <bool ANY, bool COUNT, bool RESUME = false, bool UNORDERED = false>
__device__ bool traverse(const DevScene& sc, vec3 O, vec3 D, float tmax, Stack& stk, Hit& h, TravCount& cnt,
                         bool& overflow, bool alpha = false, Rng* rng = nullptr, TravState* ts = nullptr,
                         int budget = -1) {
    static_assert(!(ANY && RESUME), "only closest-hit traversals are resumed");
    static_assert(!UNORDERED || ANY, "only any-hit traversals ignore the visit order");
    if (COUNT) cnt.rays++;
    const vec3 invD = V(1.0f / D.x, 1.0f / D.y, 1.0f / D.z);
    vec3 o = O, d = D, inv = invD;
    float best = tmax;
    int sp = 0, blas_base = -1, cur_inst = -1, cur_mat = 0;
    int ref = sc.root;
    if (RESUME) {
        ref = ts->ref; sp = ts->sp; blas_base = ts->blas_base; cur_inst = ts->cur_inst;
        if (blas_base >= 0) {
            inst_to_object(sc.inst[cur_inst], O, D, o, d);
            inv = V(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
        }
    }
    for (;;) {
        if (RESUME && budget-- == 0) {
            ts->ref = ref; ts->sp = sp; ts->blas_base = blas_base; ts->cur_inst = cur_inst;
            return false;
        }
        bool do_pop = false;
        if (ref >= 0 && sc.nodes4) {
            // Four-wide node (TLAS or BLAS): the hit children in order of the
            // reference's key (nearest first; equal keys in child order), the
            // nearest visited next, the others pushed farthest first.
            if (COUNT) cnt.nodes += 2;
            const GNode4& n = sc.nodes4[ref];
            const float4 lx = n.lox, hx = n.hix, ly = n.loy, hy = n.hiy, lz = n.loz, hz = n.hiz;
            const int4 rf = n.ref;
            float t0a, t0b, t0c, t0d;
            float ka = slab(lx.x, ly.x, lz.x, hx.x, hy.x, hz.x, o, inv, t0a);
            float kb = slab(lx.y, ly.y, lz.y, hx.y, hy.y, hz.y, o, inv, t0b);
            float kc = slab(lx.z, ly.z, lz.z, hx.z, hy.z, hz.z, o, inv, t0c);
            float kd = slab(lx.w, ly.w, lz.w, hx.w, hy.w, hz.w, o, inv, t0d);
            const bool ha = rf.x != EMPTY4 && ka > 0.0f && !(t0a > best);
            const bool hb = rf.y != EMPTY4 && kb > 0.0f && !(t0b > best);
            const bool hc = rf.z != EMPTY4 && kc > 0.0f && !(t0c > best);
            const bool hd = rf.w != EMPTY4 && kd > 0.0f && !(t0d > best);
            const int c = int(ha) + int(hb) + int(hc) + int(hd);
            if (c == 0) {
                do_pop = true;
            } else if (UNORDERED) {
                // the first hit child next, the others pushed
                if (sp + c - 1 > MAXD) { overflow = true; break; }
                int first = EMPTY4;
                if (ha) first = rf.x;
                if (hb) { if (first == EMPTY4) first = rf.y; else stk.push(sp++, rf.y); }
                if (hc) { if (first == EMPTY4) first = rf.z; else stk.push(sp++, rf.z); }
                if (hd) { if (first == EMPTY4) first = rf.w; else stk.push(sp++, rf.w); }
                ref = first;
            } else {
                const float big = __builtin_inff();
                float q0 = ha ? ka : big, q1 = hb ? kb : big, q2 = hc ? kc : big, q3 = hd ? kd : big;
                int r0 = rf.x, r1 = rf.y, r2 = rf.z, r3 = rf.w;
                // sorting network on (key, ref); a pair swaps only on a strictly smaller key
#define PT_CSWAP(qa, ra, qb, rb)                                      \
    if (qb < qa) {                                                    \
        const float tq = qa; qa = qb; qb = tq;                        \
        const int tr = ra; ra = rb; rb = tr;                          \
    }
                PT_CSWAP(q0, r0, q1, r1)
                PT_CSWAP(q2, r2, q3, r3)
                PT_CSWAP(q0, r0, q2, r2)
                PT_CSWAP(q1, r1, q3, r3)
                PT_CSWAP(q1, r1, q2, r2)
#undef PT_CSWAP
                if (sp + c - 1 > MAXD) { overflow = true; break; }
                if (COUNT) {
                    const int lo = sp > stk.S ? sp : stk.S, hi = sp + c - 1;
                    if (hi > lo) cnt.spills += hi - lo;
                }
                if (c > 3) stk.push(sp++, r3);
                if (c > 2) stk.push(sp++, r2);
                if (c > 1) stk.push(sp++, r1);
                ref = r0;
            }
        } else if (ref >= 0) {
            if (COUNT) cnt.nodes++;
            const GNode n = sc.nodes[ref];
            float tl0, tr0;
            float kl = slab(n.a.x, n.a.y, n.a.z, n.a.w, n.b.x, n.b.y, o, inv, tl0);
            float kr = slab(n.b.z, n.b.w, n.c.x, n.c.y, n.c.z, n.c.w, o, inv, tr0);
            bool hl = kl > 0.0f && !(tl0 > best);
            bool hr = kr > 0.0f && !(tr0 > best);
            if (hl && hr) {
                int nearer = (!UNORDERED && kl > kr) ? n.d.y : n.d.x;
                int farther = (!UNORDERED && kl > kr) ? n.d.x : n.d.y;
                if (sp >= MAXD) { overflow = true; break; }
                if (COUNT && sp >= stk.S) cnt.spills++;
                stk.push(sp++, farther);
                ref = nearer;
            } else if (hl) {
                ref = n.d.x;
            } else if (hr) {
                ref = n.d.y;
            } else {
                do_pop = true;
            }
        } else if (blas_base < 0) {
            // TLAS leaf: enter the instance's BLAS with the ray in object space.
            if (COUNT) cnt.insts++;
            cur_inst = -1 - ref;
            const GInst I = sc.inst[cur_inst];
            if (ANY && alpha) cur_mat = __float_as_int(I.f2.z);
            inst_to_object(I, O, D, o, d);
            inv = V(1.0f / d.x, 1.0f / d.y, 1.0f / d.z);
            blas_base = sp;
            ref = __float_as_int(sc.nodes4 ? I.f2.w : I.f2.y);
        } else {
            const int code = -1 - ref;
            const int first = code >> 2, count = code & 3;
            for (int i = 0; i < count; ++i) {
                if (COUNT) cnt.tris++;
                const GTri T = sc.tris[first + i];
                vec3 v0 = xyz(T.v0), e0 = xyz(T.e0), e1 = xyz(T.e1);
                vec3 pv = cross(d, e1);
                float det = dot(e0, pv);
                vec3 tv = o - v0;
                vec3 qv = cross(tv, e0);
                // uvt.xyz / det (closest_hit.glsl:140) as one correctly rounded
                // reciprocal and three products: within GLSL's 2.5 ulp for a
                // division, the oracle's arithmetic, two division sequences
                // fewer per triangle
                const float rdet = 1.0f / det;
                float uu = dot(tv, pv) * rdet;
                float vv = dot(d, qv) * rdet;
                float tt = dot(e1, qv) * rdet;
                float ww = 1.0f - uu - vv;
                if (uu >= 0.0f && vv >= 0.0f && tt >= 0.0f && ww >= 0.0f && tt < best) {
                    if (ANY && !(alpha && alpha_transparent(sc, first + i, cur_mat, uu, vv, ww, rng))) return true;
                    if (ANY) continue;
                    best = tt;
                    h.t = tt; h.u = uu; h.v = vv; h.tri = first + i; h.inst = cur_inst; h.light = -1;
                }
            }
            do_pop = true;
        }
        if (do_pop) {
            if (blas_base >= 0 && sp == blas_base) {
                blas_base = -1;
                o = O; d = D; inv = invD;
            }
            if (sp == 0) break;
            ref = stk.pop(--sp);
        }
    }
    return RESUME;
}

// Light loop of ClosestHit (closest_hit.glsl:35-85): closest analytic light.
template <bool COUNT>
__device__ __forceinline__ void hit_lights(const DevScene& sc, vec3 ro, vec3 rd, Hit& h, TravCount& cnt) {
    for (int i = 0; i < sc.num_lights; i++) {
        if (COUNT) cnt.lights++;
        const GLight& L = sc.lights[i];
        float type = L.u.w;
        float d;
        if (type == static_cast<float>(QUAD_LIGHT)) {
            vec3 normal = xyz(L.n);
            if (dot(normal, rd) > 0.0f) continue;
            d = RectIntersect(xyz(L.p), xyz(L.uu), xyz(L.vv), L.n, ro, rd);
            if (d < 0.0f) d = INF;
            if (d < h.t) { h.t = d; h.light = i; }
        }
        if (type == static_cast<float>(SPHERE_LIGHT)) {
            d = SphereIntersect(L.p.w, xyz(L.p), ro, rd);
            if (d < 0.0f) d = INF;
            if (d < h.t) { h.t = d; h.light = i; }
        }
    }
}

// Light loop of AnyHit (anyhit.glsl:28-63).
template <bool COUNT>
__device__ __forceinline__ bool any_light(const DevScene& sc, vec3 ro, vec3 rd, float maxDist, TravCount& cnt) {
    for (int i = 0; i < sc.num_lights; i++) {
        if (COUNT) cnt.lights++;
        const GLight& L = sc.lights[i];
        float type = L.u.w;
        if (type == static_cast<float>(QUAD_LIGHT)) {
            float d = RectIntersect(xyz(L.p), xyz(L.uu), xyz(L.vv), L.n, ro, rd);
            if (d > 0.0f && d < maxDist) return true;
        }
        if (type == static_cast<float>(SPHERE_LIGHT)) {
            float d = SphereIntersect(L.p.w, xyz(L.p), ro, rd);
            if (d > 0.0f && d < maxDist) return true;
        }
    }
    return false;
}

// Emitter record of a light hit: lightSample.pdf and emission as ClosestHit
// sets them (closest_hit.glsl:59-82).
__device__ __forceinline__ float light_hit_pdf(const DevScene& sc, int li, vec3 ro, vec3 rd, float t) {
    const GLight& L = sc.lights[li];
    if (L.u.w == static_cast<float>(QUAD_LIGHT)) {
        float cosTheta = dot(-rd, xyz(L.n));
        return (t * t) / (L.e.w * cosTheta);
    }
    vec3 hitPt = ro + t * rd;
    float cosTheta = dot(-rd, normalize(hitPt - xyz(L.p)));
    return (t * t) / (L.e.w * cosTheta * 0.5f);
}

// Hit attributes of a triangle hit (closest_hit.glsl:220-263).
__device__ void surface_attributes(const DevScene& sc, const Hit& h, vec3 ro, vec3 rd, SurfState& st) {
    st.fhp = ro + rd * h.t;
    const int4 vi = sc.vidx[h.tri];
    const float4 a0 = sc.verts[vi.x], a1 = sc.verts[vi.y], a2 = sc.verts[vi.z];
    const float4 n0 = sc.norms[vi.x], n1 = sc.norms[vi.y], n2 = sc.norms[vi.z];
    const GInst I = sc.inst[h.inst];
    float bx = 1.0f - h.u - h.v, by = h.u, bz = h.v;
    st.tu = (a0.w * bx + a1.w * by) + a2.w * bz;
    st.tv = (n0.w * bx + n1.w * by) + n2.w * bz;
    vec3 normal = normalize(xyz(n0) * bx + xyz(n1) * by + xyz(n2) * bz);
    vec3 ic0 = V(I.i0.x, I.i0.y, I.i0.z), ic1 = V(I.i0.w, I.i1.x, I.i1.y), ic2 = V(I.i1.z, I.i1.w, I.i2.x);
    st.normal = normalize(V(dot(ic0, normal), dot(ic1, normal), dot(ic2, normal)));
    st.ffnormal = dot(st.normal, rd) <= 0.0f ? st.normal : -st.normal;
    vec3 dp1 = xyz(a1) - xyz(a0), dp2 = xyz(a2) - xyz(a0);
    float du1x = a1.w - a0.w, du1y = n1.w - n0.w, du2x = a2.w - a0.w, du2y = n2.w - n0.w;
    float invdet = 1.0f / (du1x * du2y - du1y * du2x);
    vec3 tg = (dp1 * du2y - dp2 * du1y) * invdet;
    vec3 bt = (dp2 * du1x - dp1 * du2x) * invdet;
    vec3 fc0 = V(I.f0.x, I.f0.y, I.f0.z), fc1 = V(I.f0.w, I.f1.x, I.f1.y), fc2 = V(I.f1.z, I.f1.w, I.f2.x);
    st.tangent = normalize((fc0 * tg.x + fc1 * tg.y) + fc2 * tg.z);
    st.bitangent = normalize((fc0 * bt.x + fc1 * bt.y) + fc2 * bt.z);
    st.matID = __float_as_int(I.f2.z);
}

}  // namespace ptd
