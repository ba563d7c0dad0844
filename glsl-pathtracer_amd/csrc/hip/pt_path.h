// Per-path device functions of the bounce loop: ray generation, the two
// ray queries as the path uses them, DirectLight, and the PathTrace body split
// at next-event estimation so the plain kernel and the wavefront stages share
// it. Kernel-independent: nothing here uses wave intrinsics, LDS or grids.
//
// Reference behaviour (file:line in knightcrawler25/GLSL-PathTracer):
//   ray generation      src/shaders/tile.glsl:41-75
//   bounce loop         src/shaders/common/pathtrace.glsl:285-476
//   DirectLight (NEE)   src/shaders/common/pathtrace.glsl:158-283
#pragma once

#include "pt_device.h"

namespace ptd {
// Camera ray of tile.glsl:41-69 for pixel (px, py).
__device__ void camera_ray(const DevOpts& op, int px, int py, Rng& rng, vec3& ro, vec3& rd) {
    const float resx = static_cast<float>(op.width), resy = static_cast<float>(op.height);
    float cx = (static_cast<float>(px) + 0.5f) / resx;
    float cy = (static_cast<float>(py) + 0.5f) / resy;
    float r1 = 2.0f * rnd(rng);
    float r2 = 2.0f * rnd(rng);
    float jx = r1 < 1.0f ? sqrtf(r1) - 1.0f : 1.0f - sqrtf(2.0f - r1);
    float jy = r2 < 1.0f ? sqrtf(r2) - 1.0f : 1.0f - sqrtf(2.0f - r2);
    jx = jx / (resx * 0.5f);
    jy = jy / (resy * 0.5f);
    float dx = (cx * 2.0f - 1.0f) + jx;
    float dy = (cy * 2.0f - 1.0f) + jy;
    float scale = tanf(op.cam_fov * 0.5f);
    dy = dy * (resy / resx * scale);
    dx = dx * scale;
    vec3 right = V(op.cam_right[0], op.cam_right[1], op.cam_right[2]);
    vec3 up = V(op.cam_up[0], op.cam_up[1], op.cam_up[2]);
    vec3 fwd = V(op.cam_fwd[0], op.cam_fwd[1], op.cam_fwd[2]);
    vec3 rayDir = normalize(dx * right + dy * up + fwd);
    vec3 focalPoint = op.cam_focal * rayDir;
    float cam_r1 = rnd(rng) * TWO_PI;
    float cam_r2 = rnd(rng) * op.cam_aperture;
    vec3 ap = (cos_pt(cam_r1) * right + sin_pt(cam_r1) * up) * sqrtf(cam_r2);
    rd = normalize(focalPoint - ap);
    ro = V(op.cam_pos[0], op.cam_pos[1], op.cam_pos[2]) + ap;
}

// AnyHit (anyhit.glsl:25-215): lights, then the BVH; with OPT_ALPHA_TEST and
// no media, alpha-tested candidates are skipped (BLEND draws from `rng`).
// UNORDERED: the caller passes no RNG (no stochastic alpha in this query),
// so children are visited in child order (see traverse).
template <bool COUNT, bool UNORDERED = false>
__device__ __forceinline__ bool occluded(const DevScene& sc, const DevOpts& op, vec3 ro, vec3 rd, float maxDist,
                                         Stack& stk, TravCount& cnt, bool& ovf, Rng* rng = nullptr) {
    if ((op.flags & OPT_LIGHTS) && any_light<COUNT>(sc, ro, rd, maxDist, cnt)) return true;
    Hit h;
    const bool alpha = (op.flags & OPT_ALPHA_TEST) && !(op.flags & OPT_MEDIUM);
    if (UNORDERED) return traverse<true, COUNT, false, true>(sc, ro, rd, maxDist, stk, h, cnt, ovf, alpha, nullptr);
    // A stochastic alpha test draws in visit order, and four-wide nodes visit
    // children nearest first among up to four, not in the binary tree's order
    // (anyhit.glsl:118-141 over the reference's BVH2): with blending materials
    // this traversal walks the binary form.
    DevScene s = sc;
    if (rng && alpha && sc.root_bin != EMPTY4) {
        s.nodes4 = nullptr;
        s.root = sc.root_bin;
    }
    return traverse<true, COUNT>(s, ro, rd, maxDist, stk, h, cnt, ovf, alpha, rng);
}

template <bool COUNT>
__device__ __forceinline__ Hit closest(const DevScene& sc, const DevOpts& op, vec3 ro, vec3 rd, int depth, Stack& stk,
                                       TravCount& cnt, bool& ovf) {
    Hit h;
    h.t = INF; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1; h.light = -1;
    if ((op.flags & OPT_LIGHTS) && (!(op.flags & OPT_HIDE_EMITTERS) || depth > 0)) hit_lights<COUNT>(sc, ro, rd, h, cnt);
    traverse<false, COUNT>(sc, ro, rd, h.t, stk, h, cnt, ovf);
    return h;
}

// closest() in pieces: the analytic lights, then at most `budget` traversal
// iterations (budget < 0: all). Returns false when stopped, with the position
// in ts and the best hit so far in h; traverse<false, false, true> resumes it.
__device__ __forceinline__ bool closest_begin(const DevScene& sc, const DevOpts& op, vec3 ro, vec3 rd, int depth,
                                              Stack& stk, bool& ovf, Hit& h, TravState& ts, int budget) {
    TravCount cnt{0, 0, 0, 0, 0, 0};
    h.t = INF; h.u = 0.f; h.v = 0.f; h.tri = -1; h.inst = -1; h.light = -1;
    if ((op.flags & OPT_LIGHTS) && (!(op.flags & OPT_HIDE_EMITTERS) || depth > 0)) hit_lights<false>(sc, ro, rd, h, cnt);
    ts = TravState{sc.root, 0, -1, -1};
    return traverse<false, false, true>(sc, ro, rd, h.t, stk, h, cnt, ovf, false, nullptr, &ts, budget);
}

// EvalTransmittance (pathtrace.glsl:118-156): product of medium
// transmittances along a shadow ray, zero when an opaque surface blocks it.
template <bool COUNT>
__device__ vec3 eval_transmittance(const DevScene& sc, const DevOpts& op, vec3 ro, vec3 rd, Rng& rng, Stack& stk,
                                   TravCount& cnt, bool& ovf) {
    vec3 tr = V1(1.0f);
    SurfState st;
    st.matID = 0; st.tu = 0.f; st.tv = 0.f; st.normal = V1(0.f); st.ffnormal = V1(0.f);
    st.tangent = V1(0.f); st.bitangent = V1(0.f); st.fhp = V1(0.f);
    float prev_rough = 0.f;
    for (int depth = 0; depth < op.max_depth; depth++) {
        Hit h = closest<COUNT>(sc, op, ro, rd, 0, stk, cnt, ovf);
        bool hit = h.t != INF;
        if (!hit || h.tri < 0) break;
        surface_attributes(sc, h, ro, rd, st);
        float eta;
        Material m = GetMaterial(sc, op, st, rd, 0, prev_rough, eta);
        prev_rough = m.roughness;
        bool alphatest = (m.alphaMode == ALPHA_MODE_MASK && m.opacity < m.alphaCutoff) ||
                         (m.alphaMode == ALPHA_MODE_BLEND && rnd(rng) > m.opacity);
        bool refractive = (1.0f - m.metallic) * m.specTrans > 0.0f;
        if (hit && !(alphatest || refractive)) return V1(0.0f);
        if (dot(rd, st.normal) > 0.0f && m.medium.type != MEDIUM_NONE) {
            vec3 color = m.medium.type == MEDIUM_ABSORB ? V1(1.0f) - m.medium.color : V1(1.0f);
            tr = tr * vexp(-color * m.medium.density * h.t);
        }
        ro = st.fhp + rd * EPS;
    }
    return tr;
}

// Per-path state carried between bounces (the locals of PathTrace).
struct PathRegs {
    vec3 ro, rd, thr, rad;
    float alpha, prev_pdf, prev_rough;
    int depth, pass, prev_mat;
    bool in_medium, surface_scatter;
    float tu, tv;
    Medium med;
    int med_mat;  // material whose medium the path entered last (-1: none)
    Rng rng;
};

// Result of DirectLight's sampling step; the visibility test is separate so
// the wavefront form can defer it to the shadow stage.
struct NeeRay { vec3 o, d, contrib, w; float maxd; bool valid, tested; };

// The sampling and BSDF part of DirectLight for one light kind. kind 0 =
// environment (SampleEnvMap), 1 = analytic light. RNG draws happen in the
// reference order; contrib excludes the final `* throughput`; w is contrib
// without the BSDF value (the factor the adjoint multiplies df/dparam by).
// `tested`: the reference runs AnyHit for this ray (always for the
// environment, for a light when it faces the point) even where the BSDF
// gives no contribution, which matters when AnyHit draws random numbers.
__device__ NeeRay nee_setup(const DevScene& sc, const DevOpts& op, int kind, vec3 rd_in, const SurfState& st,
                            const Material& m, float eta, Rng& rng) {
    NeeRay r;
    r.valid = false;
    r.tested = kind == 0;
    r.contrib = V1(0.0f);
    r.w = V1(0.0f);
    vec3 scatterPos = st.fhp + st.normal * EPS;
    r.o = scatterPos;
    if (kind == 0) {
        vec3 Li;
        float4 dp = SampleEnvMap(sc, op, rng, Li);
        vec3 lightDir = xyz(dp);
        float lightPdf = dp.w;
        r.d = lightDir;
        r.maxd = INF - EPS;
        float pdf;
        vec3 f = DisneyEval(m, eta, -rd_in, st.ffnormal, lightDir, pdf);
        if (pdf > 0.0f) {
            float mis = PowerHeuristic(lightPdf, pdf);
            if (mis > 0.0f) {
                r.contrib = mis * Li * f * op.env_intensity / lightPdf;
                r.w = mis * Li * op.env_intensity / lightPdf;
                r.valid = true;
            }
        }
        return r;
    }
    int index = static_cast<int>(rnd(rng) * static_cast<float>(sc.num_lights));
    if (index > sc.num_lights - 1) index = sc.num_lights - 1;
    const GLight L = sc.lights[index];
    LightSample ls;
    SampleOneLight(L, scatterPos, sc.num_lights, rng, ls);
    vec3 Li = ls.emission;
    r.d = ls.direction;
    r.maxd = ls.dist - EPS;
    if (dot(ls.direction, ls.normal) < 0.0f) {
        r.tested = true;
        float pdf;
        vec3 f = DisneyEval(m, eta, -rd_in, st.ffnormal, ls.direction, pdf);
        float mis = 1.0f;
        if (L.e.w > 0.0f) mis = PowerHeuristic(ls.pdf, pdf);
        if (pdf > 0.0f) {
            r.contrib = mis * Li * f / ls.pdf;
            r.w = mis * Li / ls.pdf;
            r.valid = true;
        }
    }
    return r;
}

// DirectLight with the volume-MIS variant (transmittance instead of a binary
// test); used by the plain kernel only.
template <bool COUNT>
__device__ vec3 direct_light_volmis(const DevScene& sc, const DevOpts& op, vec3 rd_in, const SurfState& st,
                                    const Material& m, float eta, const Medium& med, bool isSurface, Rng& rng,
                                    Stack& stk, TravCount& cnt, bool& ovf) {
    vec3 Ld = V1(0.0f);
    vec3 scatterPos = st.fhp + st.normal * EPS;
    if ((op.flags & OPT_ENVMAP) && !(op.flags & OPT_UNIFORM_LIGHT)) {
        vec3 Li;
        float4 dp = SampleEnvMap(sc, op, rng, Li);
        vec3 lightDir = xyz(dp);
        float lightPdf = dp.w;
        Li = Li * eval_transmittance<COUNT>(sc, op, scatterPos, lightDir, rng, stk, cnt, ovf);
        vec3 f;
        float pdf;
        if (isSurface) f = DisneyEval(m, eta, -rd_in, st.ffnormal, lightDir, pdf);
        else { float p = PhaseHG(dot(-rd_in, lightDir), med.anisotropy); f = V1(p); pdf = p; }
        if (pdf > 0.0f) {
            float mis = PowerHeuristic(lightPdf, pdf);
            if (mis > 0.0f) Ld = Ld + mis * Li * f * op.env_intensity / lightPdf;
        }
    }
    if (op.flags & OPT_LIGHTS) {
        int index = static_cast<int>(rnd(rng) * static_cast<float>(sc.num_lights));
        if (index > sc.num_lights - 1) index = sc.num_lights - 1;
        const GLight L = sc.lights[index];
        LightSample ls;
        SampleOneLight(L, scatterPos, sc.num_lights, rng, ls);
        vec3 Li = ls.emission;
        if (dot(ls.direction, ls.normal) < 0.0f) {
            Li = Li * eval_transmittance<COUNT>(sc, op, scatterPos, ls.direction, rng, stk, cnt, ovf);
            vec3 f;
            float pdf;
            if (isSurface) f = DisneyEval(m, eta, -rd_in, st.ffnormal, ls.direction, pdf);
            else { float p = PhaseHG(dot(-rd_in, ls.direction), med.anisotropy); f = V1(p); pdf = p; }
            float mis = 1.0f;
            if (L.e.w > 0.0f) mis = PowerHeuristic(ls.pdf, pdf);
            if (pdf > 0.0f) Ld = Ld + mis * f * Li / ls.pdf;
        }
    }
    return Ld;
}

// DirectLight with binary visibility (pathtrace.glsl:158-283, non-VOL_MIS).
// With stochastic alpha in AnyHit every ray the reference tests is traced,
// so that the path's random-number stream advances as the reference's does;
// otherwise only rays that can contribute are.
template <bool COUNT>
__device__ vec3 direct_light(const DevScene& sc, const DevOpts& op, vec3 rd_in, const SurfState& st,
                             const Material& m, float eta, Rng& rng, Stack& stk, TravCount& cnt, bool& ovf) {
    vec3 Ld = V1(0.0f);
    const bool rng_in_anyhit = (op.flags & OPT_ALPHA_TEST) && !(op.flags & OPT_MEDIUM);
    for (int kind = 0; kind < 2; ++kind) {
        if (kind == 0 && !((op.flags & OPT_ENVMAP) && !(op.flags & OPT_UNIFORM_LIGHT))) continue;
        if (kind == 1 && !(op.flags & OPT_LIGHTS)) continue;
        NeeRay r = nee_setup(sc, op, kind, rd_in, st, m, eta, rng);
        if (r.valid || (rng_in_anyhit && r.tested)) {
            const bool occ = occluded<COUNT>(sc, op, r.o, r.d, r.maxd, stk, cnt, ovf, &rng);
            if (r.valid && !occ) Ld = Ld + r.contrib;
        }
    }
    return Ld;
}

// Event of one bounce after the closest hit.
enum BounceEvent { EV_END = 0, EV_SURFACE = 1, EV_PASS = 2, EV_MEDIUM = 3 };

// Miss branch of PathTrace (pathtrace.glsl:305-339).
__device__ void on_miss(const DevScene& sc, const DevOpts& op, PathRegs& p) {
    if ((op.flags & (OPT_BACKGROUND | OPT_TRANSPARENT_BACKGROUND)) && p.depth == 0) p.alpha = 0.0f;
    if (!(op.flags & OPT_HIDE_EMITTERS) || p.depth > 0) {
        if (op.flags & OPT_UNIFORM_LIGHT) {
            p.rad = p.rad + V(op.ucol_x, op.ucol_y, op.ucol_z) * p.thr;
        } else if (op.flags & OPT_ENVMAP) {
            float4 e = EvalEnvMap(sc, op, p.rd);
            float mis = 1.0f;
            if (p.depth > 0) mis = PowerHeuristic(p.prev_pdf, e.w);
            if ((op.flags & OPT_MEDIUM) && !(op.flags & OPT_VOL_MIS) && !p.surface_scatter) mis = 1.0f;
            if (mis > 0.0f) p.rad = p.rad + mis * xyz(e) * p.thr * op.env_intensity;
        }
    }
}

// The medium segment a bounce ended (for the adjoint): the path crossed d = the
// hit distance inside the medium of material `mat` (-1: not in a medium) and,
// in a scattering medium, scattered at distance s when `scattered`; thr and
// rad are the throughput and radiance before the medium's factor was applied.
struct MedRec {
    int mat, type;
    float d, s, density;
    vec3 color, thr, rad;
    bool scattered;
};

// Bounce body from the hit up to next-event estimation (pathtrace.glsl:341-425):
// material, emission, emitter MIS, depth limit, medium events, alpha test.
// `st` keeps the previous bounce's attributes when a light is hit, as the
// shader's State does. Returns the event that decides what follows; with `mr`,
// records the medium segment for the adjoint.
__device__ int bounce_pre(const DevScene& sc, const DevOpts& op, const Hit& h, PathRegs& p, SurfState& st,
                          Material& m, float& eta, MedRec* mr = nullptr) {
    if (mr) mr->mat = -1;
    const bool is_emitter = h.tri < 0;
    float light_pdf = 0.f;
    vec3 light_em = V1(0.f);
    if (is_emitter) {
        light_pdf = light_hit_pdf(sc, h.light, p.ro, p.rd, h.t);
        light_em = xyz(sc.lights[h.light].e);
        st.fhp = p.ro + p.rd * h.t;
        st.matID = p.prev_mat;
        st.tu = p.tu; st.tv = p.tv;
    } else {
        surface_attributes(sc, h, p.ro, p.rd, st);
    }
    m = GetMaterial(sc, op, st, p.rd, p.depth, p.prev_rough, eta);
    p.prev_rough = m.roughness;
    p.prev_mat = st.matID;
    p.tu = st.tu; p.tv = st.tv;
    p.rad = p.rad + m.emission * p.thr;
    if ((op.flags & OPT_LIGHTS) && is_emitter) {
        float mis = 1.0f;
        if (p.depth > 0) mis = PowerHeuristic(p.prev_pdf, light_pdf);
        if ((op.flags & OPT_MEDIUM) && !(op.flags & OPT_VOL_MIS) && !p.surface_scatter) mis = 1.0f;
        p.rad = p.rad + mis * light_em * p.thr;
        return EV_END;
    }
    if (p.depth == op.max_depth) return EV_END;
    if (op.flags & OPT_MEDIUM) {
        p.surface_scatter = false;
        if (p.in_medium) {
            if (mr) {
                mr->mat = p.med_mat; mr->type = p.med.type; mr->d = h.t; mr->s = 0.0f;
                mr->density = p.med.density; mr->color = p.med.color; mr->thr = p.thr; mr->rad = p.rad;
                mr->scattered = false;
            }
            if (p.med.type == MEDIUM_ABSORB) {
                p.thr = p.thr * vexp(-(V1(1.0f) - p.med.color) * h.t * p.med.density);
            } else if (p.med.type == MEDIUM_EMISSIVE) {
                p.rad = p.rad + p.med.color * h.t * p.med.density * p.thr;
            } else {
                float scatterDist = gmin(-logf(rnd(p.rng)) / p.med.density, h.t);
                if (scatterDist < h.t) {
                    if (mr) { mr->scattered = true; mr->s = scatterDist; }
                    p.thr = p.thr * p.med.color;
                    p.ro = p.ro + p.rd * scatterDist;
                    st.fhp = p.ro;
                    return EV_MEDIUM;
                }
            }
        }
    }
    if (op.flags & OPT_ALPHA_TEST) {
        bool passed = (m.alphaMode == ALPHA_MODE_MASK && m.opacity < m.alphaCutoff) ||
                      (m.alphaMode == ALPHA_MODE_BLEND && rnd(p.rng) > m.opacity);
        if (passed) {
            p.depth--;
            p.pass++;
            return EV_PASS;
        }
    }
    p.surface_scatter = true;
    return EV_SURFACE;
}

// BSDF sample of a SURFACE event, reported to the adjoint.
struct BsdfOut { vec3 f, L; float pdf; };

// Bounce body after NEE: BSDF or phase sampling, medium bookkeeping, Russian
// roulette (pathtrace.glsl:392-471). Returns false when the path ends.
__device__ bool bounce_post(const DevOpts& op, PathRegs& p, const SurfState& st, const Material& m, float eta, int ev,
                            BsdfOut* bo = nullptr) {
    if (ev == EV_MEDIUM) {
        float ra = rnd(p.rng);
        float rb = rnd(p.rng);
        vec3 scatterDir = SampleHG(-p.rd, p.med.anisotropy, ra, rb);
        p.prev_pdf = PhaseHG(dot(-p.rd, scatterDir), p.med.anisotropy);
        p.rd = scatterDir;
    } else {
        vec3 L;
        if (ev == EV_PASS) {
            if (p.pass > 1024) return false;
            L = p.rd;
        } else {
            float pdf;
            vec3 f = DisneySample(m, eta, -p.rd, st.ffnormal, p.rng, L, pdf);
            if (bo) { bo->f = f; bo->L = L; bo->pdf = pdf; }
            p.prev_pdf = pdf;
            if (pdf > 0.0f) p.thr = p.thr * (f / pdf);
            else return false;
        }
        p.rd = L;
        p.ro = st.fhp + p.rd * EPS;
        if (op.flags & OPT_MEDIUM) {
            if (dot(p.rd, st.normal) < 0.0f && m.medium.type != MEDIUM_NONE) {
                p.in_medium = true;
                p.med = m.medium;
                p.med_mat = st.matID;
            } else if (m.medium.type != MEDIUM_NONE) {
                p.in_medium = false;
            }
        }
    }
    if ((op.flags & OPT_RR) && p.depth >= op.rr_depth) {
        float q = gmin(gmax(p.thr.x, gmax(p.thr.y, p.thr.z)) + 0.001f, 0.95f);
        if (rnd(p.rng) > q) return false;
        p.thr = p.thr / q;
    }
    return true;
}

__device__ __forceinline__ void init_path(PathRegs& p) {
    p.thr = V1(1.0f); p.rad = V1(0.0f);
    p.alpha = 1.0f; p.prev_pdf = 0.0f; p.prev_rough = 0.0f;
    p.depth = 0; p.pass = 0; p.prev_mat = 0;
    p.in_medium = false; p.surface_scatter = false;
    p.tu = 0.f; p.tv = 0.f;
    p.med.type = 0; p.med.density = 0.f; p.med.color = V1(0.f); p.med.anisotropy = 0.f;
    p.med_mat = -1;
}

__device__ __forceinline__ void init_surf(SurfState& st) {
    st.fhp = V1(0.f); st.normal = V1(0.f); st.ffnormal = V1(0.f); st.tangent = V1(0.f); st.bitangent = V1(0.f);
    st.tu = 0.f; st.tv = 0.f; st.matID = 0;
}

// NEE rays of one SURFACE or MEDIUM event in the reference's draw order
// (environment first, then one analytic light).
__device__ __forceinline__ void nee_rays(const DevScene& sc, const DevOpts& op, const PathRegs& p0, Rng& rng,
                                         const SurfState& st, const Material& m, float eta, NeeRay nee[2]) {
    nee[0].valid = nee[1].valid = false;
    if ((op.flags & OPT_ENVMAP) && !(op.flags & OPT_UNIFORM_LIGHT)) nee[0] = nee_setup(sc, op, 0, p0.rd, st, m, eta, rng);
    if (op.flags & OPT_LIGHTS) nee[1] = nee_setup(sc, op, 1, p0.rd, st, m, eta, rng);
}

// Full PathTrace for one sample (megakernel form). Inlined into each caller:
// called out of line it spills the path's registers to scratch.
template <bool COUNT>
__device__ __forceinline__ float4 path_trace(const DevScene& sc, const DevOpts& op, PathRegs& p, Stack& stk, TravCount& cnt,
                             bool& ovf) {
    SurfState st;
    init_surf(st);
    for (;; p.depth++) {
        Hit h = closest<COUNT>(sc, op, p.ro, p.rd, p.depth, stk, cnt, ovf);
        if (h.t == INF) { on_miss(sc, op, p); break; }
        Material m;
        float eta;
        const int ev = bounce_pre(sc, op, h, p, st, m, eta);
        if (ev == EV_END) break;
        if (ev != EV_PASS) {
            vec3 Ld = V1(0.0f);
            if ((op.flags & OPT_MEDIUM) && (op.flags & OPT_VOL_MIS)) {
                Ld = direct_light_volmis<COUNT>(sc, op, p.rd, st, m, eta, p.med, ev == EV_SURFACE, p.rng, stk, cnt, ovf);
            } else {
                Ld = direct_light<COUNT>(sc, op, p.rd, st, m, eta, p.rng, stk, cnt, ovf);
            }
            p.rad = p.rad + Ld * p.thr;
        }
        if (!bounce_post(op, p, st, m, eta, ev)) break;
    }
    return make_float4(p.rad.x, p.rad.y, p.rad.z, p.alpha);
}

}  // namespace ptd
