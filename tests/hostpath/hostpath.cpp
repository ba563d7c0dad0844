// CPU build of the renderer's per-path device code, for the host-path
// tests: the GPU scene layout (pt_layout.h) over the same pt_path.h /
// pt_replay.h functions the HIP kernels call, run per pixel on the host.
// host_render follows the plain kernel (one PathTrace per sample);
// host_adjoint follows the wavefront stage order of the adjoint's primal
// pass (closest hit, shade, shadow per bounce, recording NEE visibility) and
// then its replay. Test infrastructure: the product never links this.
#include <omp.h>

#include <cstdio>
#include <cstring>
#include <vector>
#include "pt_path.h"
#include "pt_replay.h"
#include "pt_layout.h"

using namespace ptd;


struct HostScene {
    ptl::Relayout R;
    std::vector<GTri> tris;
    std::vector<int4> vidx;
    std::vector<GInst> insts;
    std::vector<GLight> lights;
    DevScene sc{};
    DevOpts op{};
    HostScene(const pt_scene_view* v, const pt_options* o, const pt_camera* cam) {
        R = ptl::relayout(v->nodes, v->num_nodes, v->top_level_index, v->num_tris, v->num_instances);
        ptl::make_tris(*v, tris, vidx);
        insts.resize(v->num_instances);
        for (int i = 0; i < v->num_instances; ++i)
            insts[i] = ptl::make_inst(v->transforms + 16 * (size_t)i, R.inst_blas_ref[i], R.inst_material[i],
                                      R.inst_blas4_ref[i]);
        lights.resize(std::max(1, v->num_lights));
        for (int i = 0; i < v->num_lights; ++i) lights[i] = ptl::make_light(v->lights[i]);
        bool any_alpha = false, any_medium = false;
        for (int i = 0; i < v->num_materials; ++i) {
            if ((int)v->materials[i].alphaMode != 0) any_alpha = true;
            if ((int)v->materials[i].mediumType != 0) any_medium = true;
        }
        bool has_env = v->env_width > 0 && v->env_height > 0 && v->env_img && v->env_cdf;
        sc.nodes = R.nodes.data(); sc.tris = tris.data(); sc.vidx = vidx.data();
        sc.nodes4 = o->bvh_width == 4 ? R.wide.nodes.data() : nullptr;
        sc.verts = reinterpret_cast<const float4*>(v->vertices_uvx);
        sc.norms = reinterpret_cast<const float4*>(v->normals_uvy);
        sc.mats = reinterpret_cast<const float4*>(v->materials);
        sc.inst = insts.data(); sc.lights = lights.data();
        sc.tex = (v->num_textures > 0 && v->texture_maps) ? v->texture_maps : nullptr;
        sc.env_img = has_env ? v->env_img : nullptr;
        sc.env_cdf = has_env ? v->env_cdf : nullptr;
        sc.root = sc.nodes4 ? R.root4 : R.root;
        bool any_blend = false;
        for (int i = 0; i < v->num_materials; ++i)
            if ((int)v->materials[i].alphaMode == ALPHA_MODE_BLEND) any_blend = true;
        sc.root_bin = (sc.nodes4 && any_blend) ? R.root : EMPTY4;
        sc.num_lights = v->num_lights;
        sc.tex_w = v->tex_width; sc.tex_h = v->tex_height;
        sc.num_tex = sc.tex ? v->num_textures : 0;
        sc.env_w = has_env ? v->env_width : 0; sc.env_h = has_env ? v->env_height : 0;
        sc.env_total = has_env ? v->env_total_sum : 0.f;
        op = ptl::make_dev_opts(*o, *cam, v->num_lights, has_env, any_alpha, any_medium);
    }
};

// Wavefront stage order per path: closest hit, shade (bounce_pre, NEE rays,
// bounce_post), shadow (visibility, radiance). Records visibility bits.
static float4 primal_path(const DevScene& sc, const DevOpts& op, PathRegs& p, Stack& stk, std::vector<int>& vis) {
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    const int D = op.max_depth + 1 + ((op.flags & OPT_ALPHA_TEST) ? 64 : 0);
    vis.assign(D, 0);
    for (int b = 0; b < D; ++b) {
        Hit h = closest<false>(sc, op, p.ro, p.rd, p.depth, stk, cnt, ovf);
        if (h.t == INF) { on_miss(sc, op, p); break; }
        SurfState st; init_surf(st);
        Material m; float eta;
        const int ev = bounce_pre(sc, op, h, p, st, m, eta);
        if (ev == EV_END) break;
        NeeRay nee[2]; nee[0].valid = nee[1].valid = false;
        vec3 thr_nee = V1(0.f);
        if (ev != EV_PASS) {
            nee_rays(sc, op, p, p.rng, st, m, eta, nee);
            thr_nee = p.thr;
        }
        const bool alive = bounce_post(op, p, st, m, eta, ev);
        vec3 Ld = V1(0.f); bool lit = false;
        for (int k = 0; k < 2; ++k) {
            if (!nee[k].valid) continue;
            if (occluded<false>(sc, op, nee[k].o, nee[k].d, nee[k].maxd, stk, cnt, ovf)) continue;
            Ld = Ld + nee[k].contrib; lit = true; vis[b] |= 1 << k;
        }
        if (lit) p.rad = p.rad + Ld * thr_nee;
        if (!alive) break;
        p.depth++;
    }
    return make_float4(p.rad.x, p.rad.y, p.rad.z, p.alpha);
}

// Gradient sums (double) per material parameter; dropped: vertices with a
// non-finite gradient term; mismatch: paths whose replayed radiance differs
// from the primal's in any bit.
// grad: num_materials * NPARAM sums; mgrad (null: not computed): the medium
// gradient, num_materials * NMEDIUM sums.
extern "C" int host_adjoint(const pt_scene_view* v, const pt_options* o, const pt_camera* cam, int first, int n,
                            const float* dimg, double* grad, double* mgrad, int* dropped, int* mismatch_out) {
    try {
        HostScene H(v, o, cam);
        const DevScene& sc = H.sc;
        const DevOpts op = H.op;
        const int W = op.width, Hh = op.height, nm = v->num_materials;
        const float inv_n = 1.0f / (float)n;
        int drop = 0, mismatch = 0;
#pragma omp parallel reduction(+ : drop, mismatch)
        {
            std::vector<int> lds(MAXD + 1), spill(MAXD + 1), vis;
            Stack stk{lds.data(), spill.data(), 0, 1, MAXD, 0, 1};
            TravCount cnt{0, 0, 0, 0, 0, 0};
            bool ovf = false;
            std::vector<double> gl(nm * NPARAM, 0.0), ml(nm * NMEDIUM, 0.0);
#pragma omp for schedule(dynamic, 4)
            for (int py = 0; py < Hh; ++py)
                for (int px = 0; px < W; ++px)
                    for (int s = first; s < first + n; ++s) {
                        PathRegs p;
                        init_path(p);
                        p.rng = init_rng((uint32_t)px, (uint32_t)py, (uint32_t)(s + 2));
                        camera_ray(op, px, py, p.rng, p.ro, p.rd);
                        PathRegs p0 = p;
                        float4 lt = primal_path(sc, op, p, stk, vis);
                        const float* dp = dimg + 3 * ((size_t)py * W + px);
                        const vec3 delta = V(dp[0], dp[1], dp[2]) * inv_n;
                        p = p0;
                        const int D = (int)vis.size();
                        for (int b = 0; b < D; ++b) {
                            Hit h = closest<false>(sc, op, p.ro, p.rd, p.depth, stk, cnt, ovf);
                            if (h.t == INF) { on_miss(sc, op, p); break; }
                            float g[NPARAM], mg[NMEDIUM];
                            ReplayVertex rv = replay_vertex(sc, op, h, p, vis[b], delta, V(lt.x, lt.y, lt.z), g,
                                                            mgrad ? mg : nullptr);
                            if (rv.med_mat >= 0) {
                                if (!rv.mfinite) drop++;
                                else for (int k = 0; k < NMEDIUM; ++k) ml[rv.med_mat * NMEDIUM + k] += mg[k];
                            }
                            if (rv.mat >= 0) {
                                if (!rv.finite) drop++;
                                else for (int k = 0; k < NPARAM; ++k) gl[rv.mat * NPARAM + k] += g[k];
                            }
                            if (!rv.alive) break;
                        }
                        if (p.rad.x != lt.x || p.rad.y != lt.y || p.rad.z != lt.z) mismatch++;
                    }
#pragma omp critical
            {
                for (int i = 0; i < nm * NPARAM; ++i) grad[i] += gl[i];
                if (mgrad)
                    for (int i = 0; i < nm * NMEDIUM; ++i) mgrad[i] += ml[i];
            }
        }
        *dropped = drop;
        *mismatch_out = mismatch;
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "host_adjoint: %s\n", e.what());
        return -1;
    }
}

extern "C" int host_render(const pt_scene_view* v, const pt_options* o, const pt_camera* cam, int first, int n,
                           float* accum) {
    try {
        HostScene H(v, o, cam);
        const DevScene& sc = H.sc;
        const DevOpts op = H.op;
        const int W = op.width, Hh = op.height;
#pragma omp parallel
        {
            std::vector<int> lds(MAXD + 1), spill(MAXD + 1);
            Stack stk{lds.data(), spill.data(), 0, 1, MAXD, 0, 1};
            TravCount cnt{0, 0, 0, 0, 0, 0};
            bool ovf = false;
#pragma omp for schedule(dynamic, 4)
            for (int py = 0; py < Hh; ++py)
                for (int px = 0; px < W; ++px)
                    for (int s = first; s < first + n; ++s) {
                        PathRegs p;
                        init_path(p);
                        p.rng = init_rng((uint32_t)px, (uint32_t)py, (uint32_t)(s + 2));
                        camera_ray(op, px, py, p.rng, p.ro, p.rd);
                        float4 c = path_trace<false>(sc, op, p, stk, cnt, ovf);
                        float* a = accum + 4 * ((size_t)py * W + px);
                        a[0] += c.x; a[1] += c.y; a[2] += c.z; a[3] += c.w;
                    }
        }
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "host_render: %s\n", e.what());
        return -1;
    }
}

// Closest hits of n rays traced whole and traced in pieces of `budget`
// traversal iterations, the stack moved to a fresh stack between pieces (as
// the capped extend stage and its resume stage do, with 16 entries in "LDS"
// and the rest in a backing store). Writes the number of rays whose hit
// differs in any bit and the number of pieces traced.
extern "C" int host_closest_split(const pt_scene_view* v, const pt_options* o, const pt_camera* cam, int n,
                                  const float* orig, const float* dir, int budget, int* differ, long long* pieces) {
    try {
        HostScene H(v, o, cam);
        const DevScene& sc = H.sc;
        const DevOpts op = H.op;
        int bad = 0;
        long long np = 0;
#pragma omp parallel reduction(+ : bad, np)
        {
            const int S = 16;
            std::vector<int> l0(S), s0(MAXD + 1), l1(S), s1(MAXD + 1), l2(MAXD + 1), s2(MAXD + 1);
            Stack a{l0.data(), s0.data(), 0, 1, S, 0, 1}, b{l1.data(), s1.data(), 0, 1, S, 0, 1};
            Stack whole{l2.data(), s2.data(), 0, 1, MAXD, 0, 1};
            TravCount cnt{0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 256)
            for (int i = 0; i < n; ++i) {
                const vec3 ro = V(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]);
                const vec3 rd = V(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
                bool ovf0 = false, ovf1 = false;
                const Hit ref = closest<false>(sc, op, ro, rd, 1, whole, cnt, ovf0);
                Hit h;
                TravState ts;
                Stack* cur = &a;
                Stack* nxt = &b;
                bool done = closest_begin(sc, op, ro, rd, 1, *cur, ovf1, h, ts, budget);
                np++;
                while (!done) {
                    for (int k = 0; k < ts.sp; ++k) nxt->push(k, cur->pop(k));
                    std::swap(cur, nxt);
                    done = traverse<false, false, true>(sc, ro, rd, h.t, *cur, h, cnt, ovf1, false, nullptr, &ts,
                                                        budget);
                    np++;
                }
                const bool same = memcmp(&ref.t, &h.t, 4) == 0 && memcmp(&ref.u, &h.u, 4) == 0 &&
                                  memcmp(&ref.v, &h.v, 4) == 0 && ref.tri == h.tri && ref.inst == h.inst &&
                                  ref.light == h.light && ovf0 == ovf1;
                if (!same) bad++;
            }
        }
        *differ = bad;
        *pieces = np;
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "host_closest_split: %s\n", e.what());
        return -1;
    }
}

// Any-hit answers of n rays (segments of length maxd) with children visited
// nearest first and in child order: the number of rays whose answers differ.
extern "C" int host_any_orders(const pt_scene_view* v, const pt_options* o, const pt_camera* cam, int n,
                               const float* orig, const float* dir, const float* maxd, int* differ, int* hits) {
    try {
        HostScene H(v, o, cam);
        const DevScene& sc = H.sc;
        const DevOpts op = H.op;
        int bad = 0, nh = 0;
#pragma omp parallel reduction(+ : bad, nh)
        {
            std::vector<int> l0(16), s0(MAXD + 1);
            Stack stk{l0.data(), s0.data(), 0, 1, 16, 0, 1};
            TravCount cnt{0, 0, 0, 0, 0, 0};
#pragma omp for schedule(dynamic, 256)
            for (int i = 0; i < n; ++i) {
                const vec3 ro = V(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]);
                const vec3 rd = V(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
                bool ovf = false;
                const bool a = occluded<false>(sc, op, ro, rd, maxd[i], stk, cnt, ovf);
                const bool b = occluded<false, true>(sc, op, ro, rd, maxd[i], stk, cnt, ovf);
                if (a != b) bad++;
                if (a) nh++;
            }
        }
        *differ = bad;
        *hits = nh;
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "host_any_orders: %s\n", e.what());
        return -1;
    }
}

// The four-wide TLAS of the host re-layout (what pt_upload_scene uploads):
// room entries of 32 floats written to out (when cap >= room), the room and
// the root reference.
extern "C" int host_tlas4(const pt_scene_view* v, float* out, int cap, int* room, int* root) {
    try {
        const ptl::Relayout R = ptl::relayout(v->nodes, v->num_nodes, v->top_level_index, v->num_tris, v->num_instances);
        const int m = static_cast<int>(R.wide.nodes.size()) - R.tlas4_base;
        *room = m;
        *root = R.root4;
        if (out && cap >= m) memcpy(out, R.wide.nodes.data() + R.tlas4_base, sizeof(GNode4) * static_cast<size_t>(m));
        return 0;
    } catch (const std::exception& e) {
        fprintf(stderr, "host_tlas4: %s\n", e.what());
        return -1;
    }
}
