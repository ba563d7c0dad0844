// Path-replay adjoint, per path vertex: the replay of one bounce with the
// NEE visibility the primal pass recorded, and the gradient of that vertex's
// contribution with respect to its material's parameters. Kernel-independent
// (used by the replay kernel and by host builds of the device path).
//
// The forward model is the bounce loop of pathtrace.glsl:285-476 with
// DirectLight (:158-283) and DisneyEval (disney.glsl:244-351), as pt_path.h
// implements it. The replay repeats the primal pass's decisions exactly (same
// seeds, same draws); only the BSDF values are differentiated, sampling,
// MIS weights and Russian roulette are held fixed (path-replay
// backpropagation with detached sampling).
#pragma once

#include "pt_adjoint.h"
#include "pt_path.h"

namespace ptd {

// Gradient of one SURFACE vertex:
//   sum_c delta_c * (lcont_c / f_c) * df_c/dtheta            (BSDF-sampled continuation)
// + sum_{k visible} sum_c delta_c * thr_c * w_k,c * df_k,c/dtheta   (NEE rays)
// lcont is the radiance the path gathers after this vertex's NEE (the
// continuation through the BSDF sample, including its Russian-roulette
// weight); f its BSDF value (with |cos|), w_k the NEE weight without the BSDF
// value. A channel with f_c == 0 or lcont_c == 0 contributes nothing.
// Returns false (g left undefined) when a term is not finite.
__device__ __forceinline__ bool vertex_gradient(const DevScene& sc, const SurfState& st, const Material& m,
                                                vec3 rd_in, vec3 delta, vec3 lcont, const BsdfOut& bo,
                                                const NeeRay nee[2], int vis, vec3 thr_nee, float g[NPARAM]) {
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) g[k] = 0.0f;
    const MatD md = seed_material(sc, st, m, rd_in);
    const vec3 Vw = -rd_in;
    if (bo.pdf > 0.0f) {
        const float qx = (bo.f.x != 0.0f && lcont.x != 0.0f) ? delta.x * (lcont.x / bo.f.x) : 0.0f;
        const float qy = (bo.f.y != 0.0f && lcont.y != 0.0f) ? delta.y * (lcont.y / bo.f.y) : 0.0f;
        const float qz = (bo.f.z != 0.0f && lcont.z != 0.0f) ? delta.z * (lcont.z / bo.f.z) : 0.0f;
        if (qx != 0.0f || qy != 0.0f || qz != 0.0f) {
            const dvec3 df = DisneyEvalD(m, md, Vw, st.ffnormal, bo.L);
            grad_dot(V(qx, qy, qz), df, g);
        }
    }
    for (int k = 0; k < 2; ++k) {
        if (!nee[k].valid || !((vis >> k) & 1)) continue;
        const dvec3 dl = DisneyEvalD(m, md, Vw, st.ffnormal, nee[k].d);
        float t[NPARAM];
        grad_dot(delta * thr_nee * nee[k].w, dl, t);
#pragma unroll
        for (int j = 0; j < NPARAM; ++j) g[j] = g[j] + t[j];
    }
    bool finite = true;
#pragma unroll
    for (int k = 0; k < NPARAM; ++k) finite = finite && (g[k] - g[k] == 0.0f);
    return finite;
}

// Gradient of the medium segment a bounce ended (pathtrace.glsl:370-410),
// with the path's later radiance lafter (everything gathered after the
// medium's factor was applied) and delta = dLoss/dradiance:
//   absorb     factor exp(-(1 - c) d sigma) per channel:
//              d/dsigma = sum_c delta_c lafter_c (-(1 - c_c) d), d/dc_c = delta_c lafter_c d sigma
//   emissive   radiance += c d sigma thr:
//              d/dsigma = sum_c delta_c c_c d thr_c, d/dc_c = delta_c d sigma thr_c
//   scatter    distances drawn from sigma exp(-sigma s) (detached), so the path's
//              factor is its score: a scatter at s weighs colour * sigma
//              exp(-sigma s) / pdf, d log/dsigma = 1/sigma - s, d log/dc_c = 1/c_c
//              (a channel with c_c == 0 or lafter_c == 0 contributes nothing);
//              a crossing of length d weighs exp(-sigma d) / pdf, d log/dsigma = -d.
// Returns false when a term is not finite.
__device__ __forceinline__ bool medium_gradient(const MedRec& mr, vec3 delta, vec3 lafter, float g[NMEDIUM]) {
    const float c[3] = {mr.color.x, mr.color.y, mr.color.z};
    const float w[3] = {delta.x * lafter.x, delta.y * lafter.y, delta.z * lafter.z};
    const float t[3] = {mr.thr.x, mr.thr.y, mr.thr.z};
    const float dl[3] = {delta.x, delta.y, delta.z};
    const float l[3] = {lafter.x, lafter.y, lafter.z};
    g[0] = 0.0f;
    for (int k = 0; k < 3; ++k) {
        float gs = 0.0f, gc = 0.0f;
        if (mr.type == MEDIUM_ABSORB) {
            gs = w[k] * (-(1.0f - c[k]) * mr.d);
            gc = w[k] * (mr.d * mr.density);
        } else if (mr.type == MEDIUM_EMISSIVE) {
            gs = dl[k] * c[k] * mr.d * t[k];
            gc = dl[k] * mr.d * mr.density * t[k];
        } else if (mr.scattered) {
            gs = w[k] * (1.0f / mr.density - mr.s);
            gc = (c[k] != 0.0f && l[k] != 0.0f) ? w[k] / c[k] : 0.0f;
        } else {
            gs = w[k] * -mr.d;
        }
        g[0] = g[0] + gs;
        g[1 + k] = gc;
    }
    bool finite = true;
    for (int k = 0; k < NMEDIUM; ++k) finite = finite && (g[k] - g[k] == 0.0f);
    return finite;
}

// Result of one replayed vertex.
struct ReplayVertex {
    bool alive;    // the path continues (the primal queued it for the next bounce)
    int mat;       // material of a SURFACE vertex whose gradient is in g, else -1
    bool finite;   // false: the vertex's gradient had a non-finite term and is dropped
    int med_mat;   // material of the medium segment whose gradient is in mg, else -1
    bool mfinite;  // false: the segment's gradient had a non-finite term and is dropped
};

// One bounce of the replay after its closest hit `h` (the shade and shadow
// stages of the primal in one step): material and emission, the medium and
// alpha events, NEE with the primal's visibility bits `vis` (bit k: ray k was
// visible), the BSDF or phase sample and Russian roulette. p advances exactly
// as in the primal; `ltot` is the primal path's final radiance and `delta`
// dLoss/dradiance of its pixel. The radiance update repeats the shadow stage's
// arithmetic (visible contributions summed in ray order, times the throughput
// at NEE time), so p.rad stays bit-identical to the primal's.
__device__ __forceinline__ ReplayVertex replay_vertex(const DevScene& sc, const DevOpts& op, const Hit& h,
                                                      PathRegs& p, int vis, vec3 delta, vec3 ltot,
                                                      float g[NPARAM], float mg[NMEDIUM] = nullptr) {
    ReplayVertex out{false, -1, true, -1, true};
    SurfState st;
    init_surf(st);
    Material m;
    float eta;
    MedRec mr;
    const int ev = bounce_pre(sc, op, h, p, st, m, eta, mg ? &mr : nullptr);
    if (mg && mr.mat >= 0) {
        // everything after the medium's factor: the primal's final radiance
        // less what the path held before it (emission of this hit included)
        out.med_mat = mr.mat;
        out.mfinite = medium_gradient(mr, delta, ltot - mr.rad, mg);
    }
    if (ev == EV_END) return out;
    NeeRay nee[2];
    nee[0].valid = nee[1].valid = false;
    vec3 thr_nee = V1(0.0f);
    if (ev != EV_PASS) {
        nee_rays(sc, op, p, p.rng, st, m, eta, nee);
        thr_nee = p.thr;
        vec3 Ld = V1(0.0f);
        bool lit = false;
        for (int k = 0; k < 2; ++k) {
            if (!nee[k].valid || !((vis >> k) & 1)) continue;
            Ld = Ld + nee[k].contrib;
            lit = true;
        }
        if (lit) p.rad = p.rad + Ld * thr_nee;
    }
    const vec3 rd_in = p.rd;
    const vec3 rad_vertex = p.rad;
    BsdfOut bo;
    bo.pdf = 0.0f;
    out.alive = bounce_post(op, p, st, m, eta, ev, &bo);
    if (ev == EV_SURFACE) {
        out.mat = st.matID;
        out.finite = vertex_gradient(sc, st, m, rd_in, delta, ltot - rad_vertex, bo, nee, vis, thr_nee, g);
    }
    if (out.alive) p.depth++;
    return out;
}

}  // namespace ptd
