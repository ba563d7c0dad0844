// Derivatives of the Disney BSDF value with respect to the optimisable
// material parameters, for the path-replay adjoint.
//
// The adjoint differentiates f only: sampled directions, lobe choices, MIS
// weights, pdfs and Russian-roulette probabilities are held fixed (detached),
// which keeps the gradient estimator unbiased (E_p[df/p] = d/dθ ∫f). The
// value part of every expression follows DisneyEval (disney.glsl:37-351) in
// the same order; derivatives use forward-mode dual numbers with one
// component per parameter.
#pragma once

#include "pt_device.h"

namespace ptd {

// Parameter order of the gradient record of one material (the C ABI's
// PT_ADJOINT_PARAMS layout): material-record fields baseColor.rgb (floats 0-2),
// roughness (9), metallic (8), ior (17), clearcoat (14).
constexpr int NPARAM = 7;
// Medium parameters differentiated per material (PT_ADJOINT_MEDIUM_PARAMS):
// density, colour r, g, b (record floats 19, 20-22).
constexpr int NMEDIUM = 4;
enum { P_BASE_R = 0, P_BASE_G = 1, P_BASE_B = 2, P_ROUGHNESS = 3, P_METALLIC = 4, P_IOR = 5, P_CLEARCOAT = 6 };

struct Dual {
    float v;
    float d[NPARAM];
};
struct dvec3 { Dual x, y, z; };

__device__ __forceinline__ Dual dc(float v) {
    Dual r;
    r.v = v;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = 0.0f;
    return r;
}
__device__ __forceinline__ Dual dvar(float v, int k, float dk) {
    Dual r = dc(v);
    r.d[k] = dk;
    return r;
}
__device__ __forceinline__ Dual operator+(Dual a, Dual b) {
    Dual r; r.v = a.v + b.v;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
__device__ __forceinline__ Dual operator-(Dual a, Dual b) {
    Dual r; r.v = a.v - b.v;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
__device__ __forceinline__ Dual operator-(Dual a) {
    Dual r; r.v = -a.v;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = -a.d[k];
    return r;
}
__device__ __forceinline__ Dual operator*(Dual a, Dual b) {
    Dual r; r.v = a.v * b.v;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    Dual r; r.v = a.v / b.v;
    const float ib = 1.0f / b.v;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
__device__ __forceinline__ Dual operator+(Dual a, float b) { a.v = a.v + b; return a; }
__device__ __forceinline__ Dual operator+(float a, Dual b) { b.v = a + b.v; return b; }
__device__ __forceinline__ Dual operator-(Dual a, float b) { a.v = a.v - b; return a; }
__device__ __forceinline__ Dual operator-(float a, Dual b) {
    Dual r = -b; r.v = a - b.v; return r;
}
__device__ __forceinline__ Dual operator*(Dual a, float s) {
    Dual r; r.v = a.v * s;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = a.d[k] * s;
    return r;
}
__device__ __forceinline__ Dual operator*(float s, Dual a) {
    Dual r; r.v = s * a.v;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = s * a.d[k];
    return r;
}
__device__ __forceinline__ Dual operator/(Dual a, float s) {
    Dual r; r.v = a.v / s;
    const float is = 1.0f / s;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = a.d[k] * is;
    return r;
}
__device__ __forceinline__ Dual operator/(float a, Dual b) {
    Dual r; r.v = a / b.v;
    const float s = -r.v / b.v;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = s * b.d[k];
    return r;
}
__device__ __forceinline__ Dual dsqrt(Dual a) {
    Dual r; r.v = sqrtf(a.v);
    const float s = r.v > 0.0f ? 0.5f / r.v : 0.0f;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = a.d[k] * s;
    return r;
}
// a^p for a constant exponent
__device__ __forceinline__ Dual dpow(Dual a, float p) {
    Dual r; r.v = powf(a.v, p);
    const float s = a.v > 0.0f ? p * r.v / a.v : 0.0f;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) r.d[k] = a.d[k] * s;
    return r;
}
__device__ __forceinline__ Dual dabs(Dual a) { return a.v < 0.0f ? -a : a; }
// GLSL max/min/clamp on the value, derivative of the selected operand
__device__ __forceinline__ Dual dmax(Dual a, float b) { return a.v < b ? dc(b) : a; }
__device__ __forceinline__ Dual dmax(float a, Dual b) { return a < b.v ? b : dc(a); }
__device__ __forceinline__ Dual dclamp01(Dual x) { return x.v < 0.0f ? dc(0.0f) : (x.v > 1.0f ? dc(1.0f) : x); }
__device__ __forceinline__ Dual dmix(Dual x, Dual y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ Dual dmix(float x, Dual y, float a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ Dual dmix(Dual x, float y, Dual a) { return x * (1.0f - a) + y * a; }
__device__ __forceinline__ Dual dmix(float x, float y, Dual a) { return x * (1.0f - a) + y * a; }

__device__ __forceinline__ dvec3 D3(Dual x, Dual y, Dual z) { return dvec3{x, y, z}; }
__device__ __forceinline__ dvec3 D3c(vec3 a) { return dvec3{dc(a.x), dc(a.y), dc(a.z)}; }
__device__ __forceinline__ dvec3 D3s(Dual a) { return dvec3{a, a, a}; }
__device__ __forceinline__ dvec3 operator+(dvec3 a, dvec3 b) { return D3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ dvec3 operator*(dvec3 a, dvec3 b) { return D3(a.x * b.x, a.y * b.y, a.z * b.z); }
__device__ __forceinline__ dvec3 operator*(dvec3 a, Dual s) { return D3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ dvec3 operator*(Dual s, dvec3 a) { return D3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ dvec3 operator*(dvec3 a, float s) { return D3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ dvec3 operator*(float s, dvec3 a) { return D3(s * a.x, s * a.y, s * a.z); }
__device__ __forceinline__ dvec3 operator/(dvec3 a, Dual s) { return D3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ dvec3 operator-(dvec3 a) { return D3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ dvec3 dmix(dvec3 x, dvec3 y, float a) { return D3(dmix(x.x, y.x, a), dmix(x.y, y.y, a), dmix(x.z, y.z, a)); }
__device__ __forceinline__ dvec3 dmix(dvec3 x, float y, Dual a) { return D3(dmix(x.x, y, a), dmix(x.y, y, a), dmix(x.z, y, a)); }
__device__ __forceinline__ dvec3 dmix1(float x, dvec3 y, float a) { return D3(dmix(x, y.x, a), dmix(x, y.y, a), dmix(x, y.z, a)); }
__device__ __forceinline__ Dual ddot(dvec3 a, vec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ Dual ddot(vec3 a, dvec3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ dvec3 dnormalize(dvec3 a) {
    Dual len = dsqrt(a.x * a.x + a.y * a.y + a.z * a.z);
    return a / len;
}
__device__ __forceinline__ Dual dLuminance(dvec3 c) { return 0.212671f * c.x + 0.715160f * c.y + 0.072169f * c.z; }

// The differentiable fields of a Material, seeded with d(field)/d(param).
struct MatD {
    dvec3 baseColor;
    Dual metallic, roughness, clearcoat, ior, ax, ay, eta;
};

// Seeds of the raw material-record parameters through GetMaterial
// (pathtrace.glsl:25-115): textures replace metallic/roughness (derivative 0)
// and scale baseColor; roughness = max(r, 0.001). `m` is the processed
// material of the hit, so every value equals the forward pass's.
__device__ __forceinline__ MatD seed_material(const DevScene& sc, const SurfState& st, const Material& m, vec3 rd) {
    const float4* p = sc.mats + 8 * st.matID;
    const float4 p3 = p[2], p7 = p[6];
    vec3 bscale = V1(1.0f);
    bool tex_mr = false;
    if (sc.num_tex > 0) {
        if (static_cast<int>(p7.x) >= 0) {
            float4 col = tex_layer(sc, st.tu, st.tv, p7.x);
            bscale = vpow_cr(V(col.x, col.y, col.z), 2.2f);
        }
        tex_mr = static_cast<int>(p7.y) >= 0;
    }
    MatD d;
    d.baseColor = D3(dvar(m.baseColor.x, P_BASE_R, bscale.x), dvar(m.baseColor.y, P_BASE_G, bscale.y),
                     dvar(m.baseColor.z, P_BASE_B, bscale.z));
    d.metallic = dvar(m.metallic, P_METALLIC, tex_mr ? 0.0f : 1.0f);
    d.roughness = dvar(m.roughness, P_ROUGHNESS, (tex_mr || p3.y < 0.001f) ? 0.0f : 1.0f);
    d.clearcoat = dvar(m.clearcoat, P_CLEARCOAT, 1.0f);
    d.ior = dvar(m.ior, P_IOR, 1.0f);
    const float aspect = sqrtf(1.0f - m.anisotropic * 0.9f);
    d.ax = dmax(0.001f, d.roughness / aspect);
    d.ay = dmax(0.001f, d.roughness * aspect);
    d.eta = dot(rd, st.normal) < 0.0f ? (1.0f / d.ior) : d.ior;
    return d;
}

__device__ __forceinline__ Dual dSchlickWeight(Dual u) {
    Dual m = dclamp01(1.0f - u);
    Dual m2 = m * m;
    return m2 * m2 * m;
}
__device__ __forceinline__ Dual dDielectricFresnel(Dual cosThetaI, Dual eta) {
    Dual sinThetaTSq = eta * eta * (1.0f - cosThetaI * cosThetaI);
    if (sinThetaTSq.v > 1.0f) return dc(1.0f);
    Dual cosThetaT = dsqrt(dmax(1.0f - sinThetaTSq, 0.0f));
    Dual rs = (eta * cosThetaT - cosThetaI) / (eta * cosThetaT + cosThetaI);
    Dual rp = (eta * cosThetaI - cosThetaT) / (eta * cosThetaI + cosThetaT);
    return 0.5f * (rs * rs + rp * rp);
}
__device__ __forceinline__ Dual dGTR2Aniso(Dual NDotH, Dual HDotX, Dual HDotY, Dual ax, Dual ay) {
    Dual a = HDotX / ax;
    Dual b = HDotY / ay;
    Dual c = a * a + b * b + NDotH * NDotH;
    return 1.0f / (PI * ax * ay * c * c);
}
__device__ __forceinline__ Dual dSmithGAniso(float NDotV, float VDotX, float VDotY, Dual ax, Dual ay) {
    Dual a = VDotX * ax;
    Dual b = VDotY * ay;
    float c = NDotV;
    return (2.0f * NDotV) / (NDotV + dsqrt(a * a + b * b + c * c));
}
__device__ __forceinline__ Dual dGTR1(Dual NDotH, float a) {
    if (a >= 1.0f) return dc(INV_PI);
    float a2 = a * a;
    Dual t = 1.0f + (a2 - 1.0f) * NDotH * NDotH;
    return (a2 - 1.0f) / (PI * logf(a2) * t);
}

__device__ __forceinline__ dvec3 dEvalMicrofacetReflection(const MatD& d, vec3 Vv, vec3 L, dvec3 H, dvec3 F) {
    if (L.z <= 0.0f) return D3s(dc(0.0f));
    Dual D = dGTR2Aniso(H.z, H.x, H.y, d.ax, d.ay);
    Dual G1 = dSmithGAniso(fabsf(Vv.z), Vv.x, Vv.y, d.ax, d.ay);
    Dual G2 = G1 * dSmithGAniso(fabsf(L.z), L.x, L.y, d.ax, d.ay);
    return F * D * G2 / dc(4.0f * L.z * Vv.z);
}

// f of DisneyEval and its derivatives. Lobe gating uses the forward pass's
// float lobe probabilities, so the same lobes are evaluated.
__device__ dvec3 DisneyEvalD(const Material& m, const MatD& d, vec3 Vw, vec3 N, vec3 Lw) {
    dvec3 f = D3s(dc(0.0f));
    vec3 T, B;
    Onb(N, T, B);
    const vec3 Vv = ToLocal(T, B, N, Vw);
    const vec3 L = ToLocal(T, B, N, Lw);
    dvec3 H;
    if (L.z > 0.0f) {
        H = D3c(normalize(L + Vv));
    } else {
        dvec3 s = D3(L.x + Vv.x * d.eta, L.y + Vv.y * d.eta, L.z + Vv.z * d.eta);
        H = dnormalize(s);
    }
    if (H.z.v < 0.0f) H = -H;
    const Lobes p = LobeWeights(m, d.eta.v, Vv.z);
    // LobeWeights (disney.glsl:142-176), the parts f depends on
    Dual lum = dLuminance(d.baseColor);
    dvec3 ctint = lum.v > 0.0f ? d.baseColor / lum : D3s(dc(1.0f));
    Dual F0 = (1.0f - d.eta) / (1.0f + d.eta);
    F0 = F0 * F0;
    dvec3 Cspec0 = F0 * dmix1(1.0f, ctint, m.specularTint);
    dvec3 Csheen = dmix1(1.0f, ctint, m.sheenTint);
    Dual dielectricWt = (1.0f - d.metallic) * (1.0f - m.specTrans);
    Dual metalWt = d.metallic;
    Dual glassWt = (1.0f - d.metallic) * m.specTrans;

    const bool refl = L.z * Vv.z > 0.0f;
    Dual VDotH = dabs(ddot(Vv, H));
    if (p.diffPr > 0.0f && refl) {
        dvec3 fd = D3s(dc(0.0f));
        if (L.z > 0.0f) {
            Dual LDotH = ddot(L, H);
            Dual Rr = 2.0f * d.roughness * LDotH * LDotH;
            float FL = SchlickWeight(L.z);
            float FV = SchlickWeight(Vv.z);
            Dual Fretro = Rr * (FL + FV + FL * FV * (Rr - 1.0f));
            float Fd = (1.0f - 0.5f * FL) * (1.0f - 0.5f * FV);
            Dual Fss90 = 0.5f * Rr;
            Dual Fss = dmix(1.0f, Fss90, FL) * dmix(1.0f, Fss90, FV);
            Dual ss = 1.25f * (Fss * (1.0f / (L.z + Vv.z) - 0.5f) + 0.5f);
            Dual FH = dSchlickWeight(LDotH);
            dvec3 Fsheen = FH * m.sheen * Csheen;
            Dual mixv = dmix(Fd + Fretro, ss, m.subsurface);
            fd = INV_PI * d.baseColor * mixv + Fsheen;
        }
        f = f + fd * dielectricWt;
    }
    if (p.dielectricPr > 0.0f && refl) {
        Dual F = (dDielectricFresnel(VDotH, 1.0f / d.ior) - F0) / (1.0f - F0);
        f = f + dEvalMicrofacetReflection(d, Vv, L, H, dmix(Cspec0, 1.0f, F)) * dielectricWt;
    }
    if (p.metalPr > 0.0f && refl) {
        dvec3 F = dmix(d.baseColor, 1.0f, dSchlickWeight(VDotH));
        f = f + dEvalMicrofacetReflection(d, Vv, L, H, F) * metalWt;
    }
    if (p.glassPr > 0.0f) {
        Dual F = dDielectricFresnel(VDotH, d.eta);
        if (refl) {
            f = f + dEvalMicrofacetReflection(d, Vv, L, H, D3s(F)) * glassWt;
        } else {
            dvec3 fr = D3s(dc(0.0f));
            if (!(L.z >= 0.0f)) {
                Dual LDotH = ddot(L, H);
                Dual VDotH2 = ddot(Vv, H);
                Dual D = dGTR2Aniso(H.z, H.x, H.y, d.ax, d.ay);
                Dual G1 = dSmithGAniso(fabsf(Vv.z), Vv.x, Vv.y, d.ax, d.ay);
                Dual G2 = G1 * dSmithGAniso(fabsf(L.z), L.x, L.y, d.ax, d.ay);
                Dual denom = LDotH + VDotH2 * d.eta;
                denom = denom * denom;
                Dual eta2 = d.eta * d.eta;
                Dual jacobian = dabs(LDotH) / denom;
                dvec3 sq = D3(dsqrt(d.baseColor.x), dsqrt(d.baseColor.y), dsqrt(d.baseColor.z));
                Dual s = (1.0f - F) * D * G2 * dabs(VDotH2) * jacobian * eta2 / dc(fabsf(L.z * Vv.z));
                fr = sq * s;
            }
            f = f + fr * glassWt;
        }
    }
    if (p.clearCtPr > 0.0f && refl) {
        dvec3 fc = D3s(dc(0.0f));
        if (L.z > 0.0f) {
            Dual VDotH2 = ddot(Vv, H);
            Dual F = dmix(0.04f, 1.0f, dSchlickWeight(VDotH2));
            Dual D = dGTR1(H.z, m.clearcoatRoughness);
            float G = SmithG(L.z, 0.25f) * SmithG(Vv.z, 0.25f);
            fc = D3s(F * D * G);
        }
        f = f + fc * (0.25f * d.clearcoat);
    }
    return f * fabsf(L.z);
}

// sum over colour channels of w_c * df_c/dparam_k
__device__ __forceinline__ void grad_dot(vec3 w, const dvec3& df, float g[NPARAM]) {
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) g[k] = w.x * df.x.d[k] + w.y * df.y.d[k] + w.z * df.z.d[k];
}

}  // namespace ptd
