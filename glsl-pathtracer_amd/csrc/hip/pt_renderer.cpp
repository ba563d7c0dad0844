// Host side of the HIP renderer and its C ABI (pt_* in include/glslpt.h).
//
// Mirrors GLSLPT::Renderer (src/core/Renderer.h:104-185): the scene arrays of
// InitGPUDataBuffers (src/core/Renderer.cpp:135-249) are uploaded once and
// re-laid out for the GPU; the OPT_* switches of InitShaders (:400-459)
// become a runtime flag word; Render adds samples to an RGBA32F accumulation
// buffer (:546-590); instance edits refresh transforms, materials and the TLAS
// node range (:649-665).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/glslpt.h"
#include "pt_kernels.h"
#include "pt_layout.h"
#include "pt_tlas.h"

using ptd::DevOpts;
using ptd::DevScene;
using ptd::GInst;
using ptd::GLight;
using ptd::GNode;
using ptd::GTri;
using ptl::ArgError;
using ptl::make_inst;
using ptl::make_light;
using ptl::relayout;
using ptl::relayout_tlas;
using ptl::Relayout;
using ptl::TlasRelayout;

namespace {

struct HipError : std::exception {
    std::string msg;
    explicit HipError(std::string m) : msg(std::move(m)) {}
    const char* what() const noexcept override { return msg.c_str(); }
};
#define HIPCHK(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) throw HipError(std::string(#x) + ": " + hipGetErrorString(e_)); \
    } while (0)

template <class T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    void free() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    void ensure(size_t count) {
        if (count <= n && p) return;
        free();
        if (count == 0) return;
        HIPCHK(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
        n = count;
    }
    void upload(const T* src, size_t count, hipStream_t s) {
        ensure(count);
        if (count) HIPCHK(hipMemcpyAsync(p, src, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
};

struct StageEvent { int stage; hipEvent_t a, b; };

}  // namespace

struct pt_renderer_s {
    int device = 0;
    hipStream_t stream = nullptr;
    int num_cu = 0;
    int lds_bytes = 65536;  // LDS per workgroup (hipDeviceProp_t::sharedMemPerBlock)
    std::string err;

    // scene (GPU layout)
    DevBuf<GNode> nodes;
    DevBuf<ptd::GNode4> nodes4;            // four-wide BLAS nodes, then the four-wide TLAS
    std::unordered_map<int, int> root4;    // BLAS root: BVH2 reference -> four-wide reference
    int root2 = 0, root4_tlas = 0;         // TLAS root in the BVH2 and in the four-wide form
    int t4_base = 0, t4_room = 0;          // four-wide TLAS entries [t4_base, t4_base + t4_room) of nodes4
    DevBuf<GTri> tris;
    DevBuf<int4> vidx;
    DevBuf<float4> verts, norms, mats;
    DevBuf<GInst> inst;
    DevBuf<GLight> lights;
    DevBuf<uint8_t> tex;
    DevBuf<float> env_img, env_cdf, env_total;
    DevBuf<int> env_ok;
    DevBuf<double> env_scratch;  // chunk sums, grids, functions and states of the CDF scan
    DevScene sc{};
    bool has_scene = false;
    int num_tris = 0, num_instances = 0, num_materials = 0;
    bool any_alpha = false, any_medium = false, has_env = false;
    bool any_blend = false;  // a material with alphaMode BLEND (stochastic alpha)
    // host copy of the flat node array (for TLAS refresh) and node re-layout map
    std::vector<pt_bvh_node> flat;
    int top_level_index = 0;
    // per-instance BLAS root reference and material id (TLAS leaves)
    std::vector<int> inst_blas, inst_blas4, inst_mat;
    std::unordered_map<int, int> root_ref;  // BLAS roots of the last full re-layout
    // device TLAS rebuild: inputs, scratch, and the TLAS placement
    bool tlas_rebuildable = false;
    bool tlas_on_device = false;  // the device TLAS was rebuilt after the last host node upload
    int tlas_base = 0;
    DevBuf<float> t_xf, t_local, t_box, t_cen;
    DevBuf<int> t_blas, t_blas4, t_mat, t_ord, t_pre, t_pos, t_pseg;
    DevBuf<int> t_seg;
    DevBuf<unsigned long long> t_keys;

    // render state
    pt_options opts{};
    pt_camera cam{};
    bool has_opts = false, has_cam = false;
    DevBuf<float4> accum;
    int accum_w = 0, accum_h = 0;

    // wavefront buffers
    DevBuf<float4> p_ro, p_rd, p_thr, p_rad, p_aux, p_med, p_hit;
    DevBuf<uint4> p_rng;
    DevBuf<int4> p_meta;
    DevBuf<int2> p_hit2;
    DevBuf<int2> s_idx;
    DevBuf<float4> s_o, s_t, s_d, s_c;
    DevBuf<float4> a_rad;         // adjoint: replay radiance
    DevBuf<unsigned> a_vis;       // adjoint: primal NEE visibility bits
    DevBuf<float> a_dimg, a_grad; // adjoint: dLoss/dimage, gradient sums
    DevBuf<int> a_flags;          // adjoint: [0] material index error, [1] dropped vertices
    // batched render: job / plane / segment / pass tables, pass state, images
    DevBuf<ptd::JobDev> b_jobs;
    DevBuf<ptd::PlaneDev> b_planes;
    DevBuf<ptd::SegDev> b_segs;
    DevBuf<ptd::PassDev> b_passes;
    DevBuf<int> b_cur;
    DevBuf<float4> b_accum;
    DevBuf<int> qa, qb, hq, counters;
    // capped extend stage: stopped traversals (ptd::ContBuf)
    DevBuf<int> c_slot, c_stack;
    DevBuf<int4> c_st, c_hi;
    DevBuf<float4> c_hf;
    DevBuf<int> spill;
    DevBuf<unsigned long long> work_counts;  // [0,6) closest-hit work, [8,14) shadow work: nodes, tris, insts, lights, spills, rays
    DevBuf<int> ovf;
    DevBuf<uchar4> ldr;

    // stats
    pt_stats stats{};
    std::vector<StageEvent> pending;
    std::vector<hipEvent_t> free_events;
    hipEvent_t ev_r0 = nullptr, ev_r1 = nullptr;
    bool render_pending = false;

    ~pt_renderer_s() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& e : pending) { (void)hipEventDestroy(e.a); (void)hipEventDestroy(e.b); }
        for (auto e : free_events) (void)hipEventDestroy(e);
        if (ev_r0) (void)hipEventDestroy(ev_r0);
        if (ev_r1) (void)hipEventDestroy(ev_r1);
        DevBuf<int>* ib[] = {&qa, &qb, &hq, &counters, &spill, &ovf};
        for (auto* b : ib) b->free();
        nodes.free(); nodes4.free(); tris.free(); vidx.free(); verts.free(); norms.free(); mats.free(); inst.free();
        lights.free(); tex.free(); env_img.free(); env_cdf.free(); env_total.free(); env_ok.free(); env_scratch.free(); accum.free();
        p_ro.free(); p_rd.free(); p_thr.free(); p_rad.free(); p_aux.free(); p_med.free(); p_hit.free();
        p_rng.free(); p_meta.free(); p_hit2.free(); s_idx.free(); s_o.free(); s_t.free(); s_d.free(); s_c.free();
        a_rad.free(); a_vis.free(); a_dimg.free(); a_grad.free(); a_flags.free();
        c_slot.free(); c_stack.free(); c_st.free(); c_hi.free(); c_hf.free();
        b_jobs.free(); b_planes.free(); b_segs.free(); b_passes.free(); b_cur.free(); b_accum.free();
        work_counts.free(); ldr.free();
        t_xf.free(); t_local.free(); t_box.free(); t_cen.free(); t_blas.free(); t_blas4.free(); t_mat.free(); t_ord.free();
        t_pre.free(); t_pos.free(); t_pseg.free(); t_seg.free(); t_keys.free();
        if (stream) (void)hipStreamDestroy(stream);
    }

    hipEvent_t get_event() {
        if (!free_events.empty()) {
            hipEvent_t e = free_events.back();
            free_events.pop_back();
            return e;
        }
        hipEvent_t e;
        HIPCHK(hipEventCreate(&e));
        return e;
    }
    void stage_begin(int stage, hipEvent_t& a) {
        (void)stage;
        a = get_event();
        HIPCHK(hipEventRecord(a, stream));
    }
    void stage_end(int stage, hipEvent_t a) {
        hipEvent_t b = get_event();
        HIPCHK(hipEventRecord(b, stream));
        pending.push_back({stage, a, b});
        stats.launches_stage[stage]++;
    }
    void resolve_events() {
        if (pending.empty() && !render_pending) return;
        HIPCHK(hipStreamSynchronize(stream));
        for (auto& e : pending) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, e.a, e.b));
            stats.ms_stage[e.stage] += ms;
            free_events.push_back(e.a);
            free_events.push_back(e.b);
        }
        pending.clear();
        if (render_pending) {
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, ev_r0, ev_r1));
            stats.ms_last_render = ms;
            render_pending = false;
        }
        if (opts.count_work && work_counts.p) {
            unsigned long long c[16];
            HIPCHK(hipMemcpy(c, work_counts.p, sizeof(c), hipMemcpyDeviceToHost));
            stats.nodes_visited = static_cast<int64_t>(c[0]);
            stats.tris_tested = static_cast<int64_t>(c[1]);
            stats.instances_entered = static_cast<int64_t>(c[2]);
            stats.lights_tested = static_cast<int64_t>(c[3]);
            stats.rays_closest = static_cast<int64_t>(c[5]);
            stats.shadow_nodes = static_cast<int64_t>(c[8]);
            stats.shadow_tris = static_cast<int64_t>(c[9]);
            stats.shadow_instances = static_cast<int64_t>(c[10]);
            stats.shadow_lights = static_cast<int64_t>(c[11]);
            stats.rays_shadow = static_cast<int64_t>(c[13]);
            stats.stack_spills = static_cast<int64_t>(c[4] + c[12]);
        }
        if (ovf.p) {
            int f = 0;
            HIPCHK(hipMemcpy(&f, ovf.p, sizeof(int), hipMemcpyDeviceToHost));
            stats.stack_overflow = f;
        }
    }

    // The BLAS form the kernels traverse follows pt_options.bvh_width.
    void select_bvh() {
        sc.nodes4 = (opts.bvh_width == 4 && nodes4.p) ? nodes4.p : nullptr;
        sc.root = sc.nodes4 ? root4_tlas : root2;
        sc.root_bin = (sc.nodes4 && any_blend) ? root2 : ptd::EMPTY4;
    }
    void set_tlas_roots(const Relayout& R) {
        root2 = R.root;
        root4_tlas = R.root4;
        t4_base = R.tlas4_base;
        t4_room = static_cast<int>(R.wide.nodes.size()) - R.tlas4_base;
    }
    // Four-wide BLAS roots of BVH2 BLAS root references.
    std::vector<int> blas4_of(const std::vector<int>& blas2) const {
        std::vector<int> out(blas2.size());
        for (size_t i = 0; i < blas2.size(); ++i) {
            auto it = root4.find(blas2[i]);
            if (it == root4.end()) throw ArgError(PT_ERR_STATE, "instance names a BLAS root without four-wide nodes");
            out[i] = it->second;
        }
        return out;
    }

    DevOpts dev_opts() const {
        return ptl::make_dev_opts(opts, cam, sc.num_lights, has_env, any_alpha, any_medium);
    }

    int block_size() const { return opts.block_size > 0 ? opts.block_size : 256; }
    int stack_S() const { return std::max(1, std::min(ptd::MAXD, opts.stack_depth > 0 ? opts.stack_depth : 24)); }

    void ensure_spill(size_t threads) {
        spill.ensure(threads * static_cast<size_t>(ptd::MAXD - stack_S() + 1));
    }
    void ensure_misc() {
        work_counts.ensure(16);
        ovf.ensure(4);
    }
    // Continuation buffer for passes of up to max_paths paths: room for a
    // quarter of them (later bounces hold only the paths still alive, and of
    // those the capped stage stops the slow tail; a traversal that finds no
    // room finishes in the capped stage).
    ptd::ContBuf cont_buf(size_t max_paths) {
        const size_t cap = std::min<size_t>(std::max<size_t>(max_paths / 4, 64), INT32_MAX / ptd::CONT_STACK);
        c_slot.ensure(cap); c_st.ensure(cap); c_hf.ensure(cap); c_hi.ensure(cap);
        c_stack.ensure(cap * ptd::CONT_STACK);
        return ptd::ContBuf{c_slot.p, c_st.p, c_hf.p, c_hi.p, c_stack.p, static_cast<int>(cap)};
    }
};

namespace {

int Fail(pt_handle h, int code, const std::string& m) {
    if (h) h->err = m;
    return code;
}

template <class F>
int Guard(pt_handle h, F&& f) {
    if (!h) return PT_ERR_INVALID_ARG;
    try {
        return f();
    } catch (const HipError& e) {
        return Fail(h, PT_ERR_HIP, e.msg);
    } catch (const ArgError& e) {
        return Fail(h, e.code, e.msg);
    } catch (const std::bad_alloc&) {
        return Fail(h, PT_ERR_OVERFLOW, "out of host memory");
    } catch (const std::exception& e) {
        return Fail(h, PT_ERR_STATE, e.what());
    }
}

ptd::PathBuf path_buf(pt_renderer_s* h) {
    ptd::PathBuf pb;
    pb.ro = h->p_ro.p; pb.rd = h->p_rd.p; pb.thr = h->p_thr.p; pb.rad = h->p_rad.p; pb.rng = h->p_rng.p;
    pb.meta = h->p_meta.p; pb.aux = h->p_aux.p; pb.med = h->any_medium ? h->p_med.p : nullptr;
    pb.hit = h->p_hit.p; pb.hit2 = h->p_hit2.p;
    return pb;
}

void check_ready(pt_renderer_s* h) {
    if (!h->has_scene) throw ArgError(PT_ERR_STATE, "no scene uploaded (pt_upload_scene)");
    if (!h->has_opts) throw ArgError(PT_ERR_STATE, "no options set (pt_set_options)");
    if (!h->has_cam) throw ArgError(PT_ERR_STATE, "no camera set (pt_set_camera)");
}

// Counter layout for one wavefront pass of D bounces (all zeroed per pass):
//   [0, D+1)            paths entering extend at bounce b
//   [D+1, 2D+2)         hits entering shade at bounce b
//   [2D+2, 3D+3)        shadow rays of bounce b
//   [3D+3, 6D+6)        work counters (extend, shade, shadow) of bounce b
// Per pass: extend-queue counts [0, D], hit counts [D+1, 2D+1], shadow counts
// [2D+2, 3D+2], three work counters per bounce [3D+3, 6D+3), then per bounce
// the continuation count and the resume stage's work counter [6D+6, 8D+6).
int counter_cells(int D) { return 8 * D + 6; }
int* cont_cells(int* cnt, int D, int b) { return cnt + 6 * D + 6 + 2 * b; }

// NEE visibility is resolved inside the shade stage (shadow records are not
// used) when AnyHit draws from the path's random numbers between NEE sampling
// and BSDF sampling (stochastic BLEND alpha without media, anyhit.glsl:118-141)
// or when it is a transmittance (volume MIS with media, pathtrace.glsl:118-156).
// The scene needs NEE inside the shade stage: stochastic alpha draws from the
// path's RNG in shadow rays, volume MIS attenuates them through media.
bool needs_inline_nee(const pt_renderer_s* h, const DevOpts& op) {
    if ((op.flags & ptd::OPT_MEDIUM) && (op.flags & ptd::OPT_VOL_MIS)) return true;
    return (op.flags & ptd::OPT_ALPHA_TEST) && !(op.flags & ptd::OPT_MEDIUM) && h->any_blend;
}

// NEE in the shade stage: where the scene needs it, or by pt_options.nee_inline.
bool inline_nee(const pt_renderer_s* h, const DevOpts& op) {
    return needs_inline_nee(h, op) || h->opts.nee_inline == 1;
}

void ensure_shadow(pt_renderer_s* h, size_t max_paths) {
    h->s_idx.ensure(max_paths); h->s_o.ensure(max_paths); h->s_t.ensure(max_paths);
    h->s_d.ensure(2 * max_paths); h->s_c.ensure(2 * max_paths);
}

ptd::ShadowBuf shadow_buf(pt_renderer_s* h) { return ptd::ShadowBuf{h->s_idx.p, h->s_o.p, h->s_t.p, h->s_d.p, h->s_c.p}; }

// Launch shape and buffers of the wavefront form for passes of up to
// spp_pass samples of every pixel.
struct Wavefront {
    int npix, spp_pass, D, bs, S, b_ext, b_sha, b_sdw;
    bool cw, inline_nee;
    size_t max_paths;
    int cap;            // extend_cap of later bounces (0: off)
    ptd::ContBuf cb;    // continuation buffer when cap > 0
};

// adjoint: NEE always in the shadow stage, which records the visibility bits
// the replay applies.
Wavefront wavefront_setup(pt_renderer_s* h, const DevOpts& op, int num_samples, bool adjoint = false) {
    Wavefront w;
    const int W = op.width, H = op.height;
    const int tiles = ((W + 7) / 8) * ((H + 7) / 8);
    w.npix = tiles * 64;
    w.spp_pass = std::max(1, std::min(h->opts.wave_spp > 0 ? h->opts.wave_spp : 1, num_samples));
    const size_t max_paths = static_cast<size_t>(w.npix) * w.spp_pass;
    if (max_paths > static_cast<size_t>(INT32_MAX / 2)) throw ArgError(PT_ERR_OVERFLOW, "too many paths per pass");
    w.max_paths = max_paths;
    h->p_ro.ensure(max_paths); h->p_rd.ensure(max_paths); h->p_thr.ensure(max_paths); h->p_rad.ensure(max_paths);
    h->p_rng.ensure(max_paths); h->p_meta.ensure(max_paths); h->p_aux.ensure(max_paths);
    if (h->any_medium) h->p_med.ensure(max_paths);
    h->p_hit.ensure(max_paths); h->p_hit2.ensure(max_paths);
    w.inline_nee = adjoint ? false : inline_nee(h, op);
    if (!w.inline_nee) ensure_shadow(h, max_paths);
    h->qa.ensure(max_paths); h->qb.ensure(max_paths); h->hq.ensure(max_paths);
    // Bounce iterations: maxDepth + 1 closest hits; alpha pass-through repeats a depth.
    w.D = op.max_depth + 1 + ((op.flags & ptd::OPT_ALPHA_TEST) ? 64 : 0);
    h->counters.ensure(counter_cells(w.D));
    h->ensure_misc();
    w.bs = h->block_size();
    w.S = h->stack_S();
    w.cw = h->opts.count_work != 0;
    w.b_ext = ptk::persistent_blocks(ptk::KIND_EXTEND, w.bs, w.S, w.cw, h->num_cu);
    // instrumented runs and the adjoint's passes traverse uncapped
    w.cap = (adjoint || w.cw) ? 0 : h->opts.extend_cap;
    w.cb = w.cap > 0 ? h->cont_buf(max_paths) : ptd::ContBuf{};
    w.b_sha = ptk::persistent_blocks(w.inline_nee ? ptk::KIND_SHADE_INLINE : ptk::KIND_SHADE, w.bs, w.S, w.cw, h->num_cu);
    w.b_sdw = ptk::persistent_blocks(ptk::KIND_SHADOW, w.bs, w.S, w.cw, h->num_cu);
    // every traversing grid (extend, shadow, an inline-NEE shade) indexes the
    // stack's backing store by its own thread slot
    h->ensure_spill(static_cast<size_t>(std::max({w.b_ext, w.b_sdw, w.inline_nee ? w.b_sha : 0})) * w.bs);
    return w;
}

// Every bounce of one pass (paths [0, npaths), samples fs...): extend, shade,
// shadow per bounce. vis (the adjoint's primal pass, else null): visibility
// words of the pass, ceil(D / 16) arrays of max_paths words.
void run_bounces(pt_renderer_s* h, const DevOpts& op, const Wavefront& w, const ptd::PathBuf& pb,
                 const ptd::ShadowBuf& sb, int npaths, int fs, unsigned* vis) {
    int* cnt = h->counters.p;
    hipStream_t s = h->stream;
    const int D = w.D;
    hipEvent_t ev;
    ptk::launch_fill(s, cnt, counter_cells(D), 0);
    int* qin = h->qa.p;
    int* qout = h->qb.p;
    int* c_ext = cnt;  // count of the current extend queue
    const int cadence = std::max(1, h->opts.compact_every);
    for (int b = 0; b < D; ++b) {
        int* c_hit = cnt + (D + 1) + b;
        int* c_sh = cnt + 2 * (D + 1) + b;
        int* wk = cnt + 3 * (D + 1) + 3 * b;
        const bool primary = b == 0;
        // bounce 0 always compacts (its input is the implicit slot range);
        // later bounces compact every `cadence`-th time, otherwise the next
        // extend re-reads this queue and skips ended paths.
        const bool compact = primary || (b % cadence) == 0;
        int* c_next = compact ? cnt + b + 1 : c_ext;
        int* c_cont = cont_cells(cnt, D, b);
        const bool capped = w.cap > 0 && !primary;
        h->stage_begin(PT_STAGE_EXTEND, ev);
        ptk::launch_extend(s, w.b_ext, w.bs, h->sc, op, pb, primary, qin, c_ext, npaths, w.npix, fs, h->hq.p, c_hit,
                           wk + 0, w.S, h->spill.p, h->work_counts.p, h->ovf.p, w.cw, capped ? w.cap : 0, &w.cb,
                           c_cont);
        if (capped)
            ptk::launch_extend_resume(s, w.b_ext, w.bs, h->sc, op, pb, w.cb, c_cont, h->hq.p, c_cont + 1, w.S,
                                      h->spill.p, h->ovf.p);
        h->stage_end(PT_STAGE_EXTEND, ev);
        h->stage_begin(PT_STAGE_SHADE, ev);
        ptk::launch_shade(s, w.b_sha, w.bs, h->sc, op, pb, sb, nullptr, primary, w.inline_nee, h->hq.p, c_hit, w.npix,
                          fs, qout, c_next, c_sh, wk + 1, compact, w.S, h->spill.p, h->work_counts.p + 8, h->ovf.p);
        h->stage_end(PT_STAGE_SHADE, ev);
        if (!w.inline_nee) {
            unsigned* vb = vis ? vis + static_cast<size_t>(b / ptd::VIS_BOUNCES_PER_WORD) * w.max_paths : nullptr;
            h->stage_begin(PT_STAGE_SHADOW, ev);
            ptk::launch_shadow(s, w.b_sdw, w.bs, h->sc, op, pb, sb, c_sh, wk + 2, w.S, h->spill.p,
                               h->work_counts.p + 8, h->ovf.p, w.cw, vb, 2 * (b % ptd::VIS_BOUNCES_PER_WORD));
            h->stage_end(PT_STAGE_SHADOW, ev);
        }
        if (compact) {
            std::swap(qin, qout);
            c_ext = c_next;
        }
    }
}

void render_wavefront(pt_renderer_s* h, int first_sample, int num_samples) {
    const DevOpts op = h->dev_opts();
    const Wavefront w = wavefront_setup(h, op, num_samples);
    const ptd::PathBuf pb = path_buf(h);
    const ptd::ShadowBuf sb = shadow_buf(h);
    for (int s0 = 0; s0 < num_samples; s0 += w.spp_pass) {
        const int ns = std::min(w.spp_pass, num_samples - s0);
        run_bounces(h, op, w, pb, sb, w.npix * ns, first_sample + s0, nullptr);
        hipEvent_t ev;
        h->stage_begin(PT_STAGE_ACCUM, ev);
        ptk::launch_accum(h->stream, op, pb, h->accum.p, w.npix, ns);
        h->stage_end(PT_STAGE_ACCUM, ev);
        HIPCHK(hipGetLastError());
        if (w.cw) {  // instrumented runs: hits per bounce of this pass (the counters are reused per pass)
            std::vector<int> hits(w.D);
            HIPCHK(hipMemcpyAsync(hits.data(), h->counters.p + (w.D + 1), w.D * sizeof(int), hipMemcpyDeviceToHost,
                                  h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
            for (int b = 0; b < w.D; ++b) h->stats.bounce_hits[std::min(b, 15)] += hits[b];
        }
    }
}

// Path-replay adjoint (pt_replay.h). Per pass of samples: the primal pass is
// the forward wavefront pass, which leaves every path's total radiance in the
// radiance buffer and records which NEE rays were visible; the replay pass
// re-traces the same paths (closest hits again, same seeds, same decisions)
// on a second radiance buffer, applies the recorded visibility instead of
// shadow rays, and accumulates each surface vertex's parameter gradient (and,
// with `medium`, each medium segment's, after them in the same buffer).
void adjoint_wavefront(pt_renderer_s* h, int first_sample, int num_samples, float inv_n, bool medium) {
    const DevOpts op = h->dev_opts();
    if (op.flags & ptd::OPT_ROUGHNESS_MOLLIFICATION)
        throw ArgError(PT_ERR_UNSUPPORTED, "the adjoint does not differentiate roughness mollification");
    if (needs_inline_nee(h, op))
        throw ArgError(PT_ERR_UNSUPPORTED, "the adjoint replays recorded NEE visibility: stochastic (blend) alpha "
                                           "in shadow rays and volume MIS are not supported");
    const Wavefront w = wavefront_setup(h, op, num_samples, true);
    const size_t grad_bytes =
        static_cast<size_t>(h->num_materials) * (ptd::NPARAM + (medium ? ptd::NMEDIUM : 0)) * sizeof(float);
    const bool lds = grad_bytes <= 32768;
    const int b_rep = ptk::persistent_blocks(ptk::KIND_REPLAY, w.bs, w.S, false, h->num_cu, lds ? grad_bytes : 0);
    const int nwords = (w.D + ptd::VIS_BOUNCES_PER_WORD - 1) / ptd::VIS_BOUNCES_PER_WORD;
    const size_t vis_words = static_cast<size_t>(nwords) * w.max_paths;
    if (vis_words > static_cast<size_t>(INT32_MAX)) throw ArgError(PT_ERR_OVERFLOW, "too many paths per pass");
    h->a_rad.ensure(w.max_paths);
    h->a_vis.ensure(vis_words);
    const ptd::PathBuf pb = path_buf(h);
    ptd::PathBuf pb2 = pb;
    pb2.rad = h->a_rad.p;
    const ptd::ShadowBuf sb = shadow_buf(h);
    ptd::ReplayBuf rb{};
    rb.ltot = pb.rad;
    rb.dimg = h->a_dimg.p;
    rb.inv_n = inv_n;
    rb.grad = h->a_grad.p;
    rb.mgrad = medium ? h->a_grad.p + static_cast<size_t>(h->num_materials) * ptd::NPARAM : nullptr;
    rb.err = h->a_flags.p;
    rb.dropped = h->a_flags.p + 1;
    rb.num_mats = h->num_materials;
    rb.lds = lds ? 1 : 0;
    hipStream_t s = h->stream;
    int* cnt = h->counters.p;
    const int D = w.D;
    const int cadence = std::max(1, h->opts.compact_every);
    for (int s0 = 0; s0 < num_samples; s0 += w.spp_pass) {
        const int ns = std::min(w.spp_pass, num_samples - s0);
        const int npaths = w.npix * ns, fs = first_sample + s0;
        ptk::launch_fill(s, reinterpret_cast<int*>(h->a_vis.p), static_cast<int>(vis_words), 0);
        run_bounces(h, op, w, pb, sb, npaths, fs, h->a_vis.p);
        // replay: the extend stage of the forward renderer, then the replay stage
        ptk::launch_fill(s, cnt, counter_cells(D), 0);
        int* qin = h->qa.p;
        int* qout = h->qb.p;
        int* c_ext = cnt;
        hipEvent_t ev;
        for (int b = 0; b < D; ++b) {
            int* c_hit = cnt + (D + 1) + b;
            int* wk = cnt + 3 * (D + 1) + 3 * b;
            const bool primary = b == 0;
            const bool compact = primary || (b % cadence) == 0;
            int* c_next = compact ? cnt + b + 1 : c_ext;
            h->stage_begin(PT_STAGE_EXTEND, ev);
            ptk::launch_extend(s, w.b_ext, w.bs, h->sc, op, pb2, primary, qin, c_ext, npaths, w.npix, fs, h->hq.p,
                               c_hit, wk + 0, w.S, h->spill.p, h->work_counts.p, h->ovf.p, false);
            h->stage_end(PT_STAGE_EXTEND, ev);
            rb.vis = h->a_vis.p + static_cast<size_t>(b / ptd::VIS_BOUNCES_PER_WORD) * w.max_paths;
            h->stage_begin(PT_STAGE_ADJOINT, ev);
            ptk::launch_replay_shade(s, b_rep, w.bs, h->sc, op, pb2, primary, h->hq.p, c_hit, w.npix, fs, qout, c_next,
                                     wk + 1, compact, rb, 2 * (b % ptd::VIS_BOUNCES_PER_WORD));
            h->stage_end(PT_STAGE_ADJOINT, ev);
            if (compact) {
                std::swap(qin, qout);
                c_ext = c_next;
            }
        }
        HIPCHK(hipGetLastError());
    }
}

// ---- batched render of uneven jobs
// Plan: every (job, sample) is a plane of npix_job path slots; planes are
// taken job by job, sample by sample, and packed into passes of at most
// `budget` slots (a plane never splits). Each pass runs every bounce of the
// deepest job; shallower paths end at their own depth.
struct BatchPlan {
    std::vector<ptd::JobDev> jobs;
    std::vector<ptd::PlaneDev> planes;
    std::vector<ptd::SegDev> segs;
    std::vector<ptd::PassDev> passes;
    size_t pixels = 0;    // sum of W*H over jobs
    long long items = 0;  // sum of npix over jobs (pixel items of the one-launch plain form)
    int max_paths = 0;    // largest pass
    int max_depth = 0;
};

BatchPlan plan_batch(const pt_job* jobs, int n, long long budget) {
    BatchPlan P;
    long long max_npix = 0;
    for (int j = 0; j < n; ++j) {
        const pt_job& J = jobs[j];
        if (J.width <= 0 || J.height <= 0 || J.spp <= 0 || J.max_depth < 0 || !J.accum_out)
            throw ArgError(PT_ERR_INVALID_ARG, "job " + std::to_string(j) + ": invalid size, spp, depth or output");
        if (static_cast<long long>(J.width) * J.height > (1LL << 26))
            throw ArgError(PT_ERR_INVALID_ARG, "job " + std::to_string(j) + ": resolution too large");
        ptd::JobDev d{};
        for (int i = 0; i < 3; ++i) {
            d.cam_pos[i] = J.camera.position[i];
            d.cam_up[i] = J.camera.up[i];
            d.cam_right[i] = J.camera.right[i];
            d.cam_fwd[i] = J.camera.forward[i];
        }
        d.cam_fov = J.camera.fov;
        d.cam_focal = J.camera.focalDist;
        d.cam_aperture = J.camera.aperture;
        d.width = J.width;
        d.height = J.height;
        d.max_depth = J.max_depth;
        d.npix = ((J.width + 7) / 8) * ((J.height + 7) / 8) * 64;
        d.spp = J.spp;
        d.item0 = static_cast<int>(P.items);
        d.accum_off = static_cast<long long>(P.pixels);
        P.items += d.npix;
        P.pixels += static_cast<size_t>(J.width) * J.height;
        P.jobs.push_back(d);
        max_npix = std::max<long long>(max_npix, d.npix);
        P.max_depth = std::max(P.max_depth, J.max_depth);
    }
    if (P.items > INT32_MAX / 2) throw ArgError(PT_ERR_OVERFLOW, "too many pixels in the batch");
    budget = std::max(budget, max_npix);
    if (budget > INT32_MAX / 2) throw ArgError(PT_ERR_OVERFLOW, "too many paths per pass");
    ptd::PassDev cur{};
    long long used = 0;
    auto close_pass = [&]() {
        if (cur.nplanes == 0) return;
        cur.npaths = static_cast<int>(used);
        cur.seg0 = static_cast<int>(P.segs.size());
        int thread0 = 0;
        for (int q = cur.plane0; q < cur.plane0 + cur.nplanes; ++q) {
            const ptd::PlaneDev& pl = P.planes[q];
            if (P.segs.size() > static_cast<size_t>(cur.seg0) && P.segs.back().job == pl.job) {
                P.segs.back().nplanes++;
            } else {
                P.segs.push_back({pl.job, q, 1, thread0});
                thread0 += P.jobs[pl.job].npix;
            }
        }
        cur.nsegs = static_cast<int>(P.segs.size()) - cur.seg0;
        cur.nthreads = thread0;
        P.max_paths = std::max(P.max_paths, cur.npaths);
        P.passes.push_back(cur);
        cur = ptd::PassDev{};
        cur.plane0 = static_cast<int>(P.planes.size());
        used = 0;
    };
    cur.plane0 = 0;
    for (int j = 0; j < n; ++j) {
        for (int smp = 0; smp < jobs[j].spp; ++smp) {
            const int np = P.jobs[j].npix;
            if (used + np > budget) close_pass();
            P.planes.push_back({j, smp, static_cast<int>(used), 0});
            cur.nplanes++;
            used += np;
        }
    }
    close_pass();
    return P;
}

// Plain form of pt_batch_render: every job in one persistent launch of the
// plain kernel (k_plain_batch), each thread one pixel, its samples in order.
void batch_render_plain(pt_renderer_s* h, const pt_job* jobs, int n) {
    const DevOpts op = h->dev_opts();
    const BatchPlan P = plan_batch(jobs, n, 0);
    hipStream_t s = h->stream;
    h->b_jobs.upload(P.jobs.data(), P.jobs.size(), s);
    h->b_accum.ensure(P.pixels);
    h->counters.ensure(std::max<size_t>(h->counters.n, 8));
    h->ensure_misc();
    const int bs = h->block_size(), S = h->stack_S();
    const int blocks = ptk::persistent_blocks(ptk::KIND_PLAIN_BATCH, bs, S, false, h->num_cu);
    h->ensure_spill(static_cast<size_t>(blocks) * bs);
    ptk::launch_fill(s, h->counters.p, 1, 0);
    hipEvent_t ev;
    h->stage_begin(PT_STAGE_BATCH, ev);
    ptk::launch_plain_batch(s, blocks, bs, h->sc, op, h->b_jobs.p, n, static_cast<int>(P.items), h->b_accum.p,
                            h->counters.p, S, h->spill.p, h->ovf.p);
    h->stage_end(PT_STAGE_BATCH, ev);
    HIPCHK(hipGetLastError());
    for (int j = 0; j < n; ++j) {
        const ptd::JobDev& J = P.jobs[j];
        HIPCHK(hipMemcpyAsync(jobs[j].accum_out, h->b_accum.p + J.accum_off,
                              static_cast<size_t>(J.width) * J.height * sizeof(float4), hipMemcpyDeviceToHost, s));
    }
}

void batch_render(pt_renderer_s* h, const pt_job* jobs, int n) {
    if (h->opts.mode == PT_MODE_PLAIN) {
        batch_render_plain(h, jobs, n);
        return;
    }
    DevOpts op = h->dev_opts();
    const bool inl = inline_nee(h, op);
    // pass budget: wave_spp samples of the largest job
    long long max_npix = 0;
    for (int j = 0; j < n; ++j)
        max_npix = std::max<long long>(max_npix, static_cast<long long>((jobs[j].width + 7) / 8) * ((jobs[j].height + 7) / 8) * 64);
    const BatchPlan P = plan_batch(jobs, n, std::max(1, h->opts.wave_spp) * max_npix);
    const size_t max_paths = static_cast<size_t>(P.max_paths);
    h->p_ro.ensure(max_paths); h->p_rd.ensure(max_paths); h->p_thr.ensure(max_paths); h->p_rad.ensure(max_paths);
    h->p_rng.ensure(max_paths); h->p_meta.ensure(max_paths); h->p_aux.ensure(max_paths);
    if (h->any_medium) h->p_med.ensure(max_paths);
    h->p_hit.ensure(max_paths); h->p_hit2.ensure(max_paths);
    if (!inl) ensure_shadow(h, max_paths);
    h->qa.ensure(max_paths); h->qb.ensure(max_paths); h->hq.ensure(max_paths);
    const int D = P.max_depth + 1 + ((op.flags & ptd::OPT_ALPHA_TEST) ? 64 : 0);
    const int ncells = counter_cells(D);
    h->counters.ensure(ncells);
    h->ensure_misc();
    const int bs = h->block_size(), S = h->stack_S();
    const int b_ext = ptk::persistent_blocks(ptk::KIND_EXTEND, bs, S, false, h->num_cu);
    const int cap = h->opts.extend_cap;
    const ptd::ContBuf cb = cap > 0 ? h->cont_buf(max_paths) : ptd::ContBuf{};
    const int b_sha = ptk::persistent_blocks(inl ? ptk::KIND_SHADE_BATCH_INLINE : ptk::KIND_SHADE_BATCH, bs, S, false,
                                             h->num_cu);
    const int b_sdw = ptk::persistent_blocks(ptk::KIND_SHADOW, bs, S, false, h->num_cu);
    const int b_acc = 4 * h->num_cu;
    h->ensure_spill(static_cast<size_t>(std::max({b_ext, b_sdw, inl ? b_sha : 0})) * bs);
    hipStream_t s = h->stream;
    h->b_jobs.upload(P.jobs.data(), P.jobs.size(), s);
    h->b_planes.upload(P.planes.data(), P.planes.size(), s);
    h->b_segs.upload(P.segs.data(), P.segs.size(), s);
    h->b_passes.upload(P.passes.data(), P.passes.size(), s);
    h->b_cur.ensure(16);
    h->b_accum.ensure(P.pixels);
    HIPCHK(hipMemsetAsync(h->b_cur.p, 0, 16 * sizeof(int), s));
    HIPCHK(hipMemsetAsync(h->b_accum.p, 0, P.pixels * sizeof(float4), s));
    const ptd::BatchBuf bb{h->b_jobs.p, h->b_planes.p, h->b_segs.p, h->b_passes.p, h->b_cur.p, h->b_accum.p};
    const ptd::PathBuf pb = path_buf(h);
    const ptd::ShadowBuf sb = shadow_buf(h);
    int* cnt = h->counters.p;
    const int cadence = std::max(1, h->opts.compact_every);
    const bool graph = h->opts.use_graph != 0;

    // the stage launches of one pass (the same for every pass: pass-specific
    // values are read from the device-side pass state)
    auto one_pass = [&](bool timed) {
        hipEvent_t ev;
        auto stage = [&](int id, auto&& launch) {
            if (timed) h->stage_begin(id, ev);
            launch();
            if (timed) h->stage_end(id, ev);
        };
        ptk::launch_pass_begin(s, bb, cnt, ncells);
        int* qin = h->qa.p;
        int* qout = h->qb.p;
        int* c_ext = cnt;
        for (int b = 0; b < D; ++b) {
            int* c_hit = cnt + (D + 1) + b;
            int* c_sh = cnt + 2 * (D + 1) + b;
            int* wk = cnt + 3 * (D + 1) + 3 * b;
            const bool primary = b == 0;
            const bool compact = primary || (b % cadence) == 0;
            int* c_next = compact ? cnt + b + 1 : c_ext;
            int* c_cont = cont_cells(cnt, D, b);
            stage(PT_STAGE_EXTEND, [&] {
                if (primary) {
                    ptk::launch_extend_batch(s, b_ext, bs, h->sc, op, pb, bb, h->hq.p, c_hit, wk, S, h->spill.p, h->ovf.p);
                } else {
                    ptk::launch_extend(s, b_ext, bs, h->sc, op, pb, false, qin, c_ext, 0, 0, 0, h->hq.p, c_hit, wk, S,
                                       h->spill.p, h->work_counts.p, h->ovf.p, false, cap, &cb, c_cont);
                    if (cap > 0)
                        ptk::launch_extend_resume(s, b_ext, bs, h->sc, op, pb, cb, c_cont, h->hq.p, c_cont + 1,
                                                  S, h->spill.p, h->ovf.p);
                }
            });
            stage(PT_STAGE_SHADE, [&] {
                ptk::launch_shade(s, b_sha, bs, h->sc, op, pb, sb, &bb, primary, inl, h->hq.p, c_hit, 0, 0, qout,
                                  c_next, c_sh, wk + 1, compact, S, h->spill.p, h->work_counts.p + 8, h->ovf.p);
            });
            if (!inl)
                stage(PT_STAGE_SHADOW, [&] {
                    ptk::launch_shadow(s, b_sdw, bs, h->sc, op, pb, sb, c_sh, wk + 2, S, h->spill.p,
                                       h->work_counts.p + 8, h->ovf.p, false);
                });
            if (compact) {
                std::swap(qin, qout);
                c_ext = c_next;
            }
        }
        stage(PT_STAGE_ACCUM, [&] { ptk::launch_accum_batch(s, b_acc, pb, bb); });
        ptk::launch_pass_end(s, bb);
    };

    const int npasses = static_cast<int>(P.passes.size());
    if (graph) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HIPCHK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
        one_pass(false);
        hipError_t ec = hipStreamEndCapture(s, &g);
        if (ec != hipSuccess) throw HipError(std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
        ec = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (ec != hipSuccess) {
            (void)hipGraphDestroy(g);
            throw HipError(std::string("hipGraphInstantiate: ") + hipGetErrorString(ec));
        }
        hipEvent_t ev;
        h->stage_begin(PT_STAGE_BATCH, ev);
        for (int p = 0; p < npasses; ++p) ec = ec == hipSuccess ? hipGraphLaunch(ge, s) : ec;
        h->stage_end(PT_STAGE_BATCH, ev);
        HIPCHK(hipStreamSynchronize(s));
        (void)hipGraphExecDestroy(ge);
        (void)hipGraphDestroy(g);
        if (ec != hipSuccess) throw HipError(std::string("hipGraphLaunch: ") + hipGetErrorString(ec));
    } else {
        hipEvent_t ev;
        h->stage_begin(PT_STAGE_BATCH, ev);
        for (int p = 0; p < npasses; ++p) one_pass(true);
        h->stage_end(PT_STAGE_BATCH, ev);
    }
    HIPCHK(hipGetLastError());
    for (int j = 0; j < n; ++j) {
        const ptd::JobDev& J = P.jobs[j];
        HIPCHK(hipMemcpyAsync(jobs[j].accum_out, h->b_accum.p + J.accum_off,
                              static_cast<size_t>(J.width) * J.height * sizeof(float4), hipMemcpyDeviceToHost, s));
    }
}

void render_plain(pt_renderer_s* h, int first_sample, int num_samples) {
    const DevOpts op = h->dev_opts();
    const int W = op.width, H = op.height;
    const int items = ((W + 7) / 8) * ((H + 7) / 8) * 64;
    h->counters.ensure(std::max<size_t>(h->counters.n, 8));
    h->ensure_misc();
    const int bs = h->block_size(), S = h->stack_S();
    const bool cw = h->opts.count_work != 0;
    const int blocks = ptk::persistent_blocks(ptk::KIND_PLAIN, bs, S, cw, h->num_cu);
    h->ensure_spill(static_cast<size_t>(blocks) * bs);
    for (int smp = first_sample; smp < first_sample + num_samples; ++smp) {
        hipEvent_t ev;
        ptk::launch_fill(h->stream, h->counters.p, 1, 0);
        h->stage_begin(PT_STAGE_PLAIN, ev);
        ptk::launch_plain(h->stream, blocks, bs, h->sc, op, h->accum.p, smp, items, h->counters.p, S, h->spill.p,
                          h->work_counts.p, h->ovf.p, cw);
        h->stage_end(PT_STAGE_PLAIN, ev);
        HIPCHK(hipGetLastError());
    }
}

}  // namespace

extern "C" {

void pt_default_options(pt_options* o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->width = 1280; o->height = 720;                 // RenderOptions defaults (Renderer.h:41-72)
    o->max_depth = 2;
    o->enable_rr = 1; o->rr_depth = 2;
    o->uniform_light_col[0] = o->uniform_light_col[1] = o->uniform_light_col[2] = 0.3f;
    o->envmap_intensity = 1.0f;
    o->opengl_normal_map = 1;
    o->mode = PT_MODE_WAVEFRONT;
    o->block_size = 256;
    o->stack_depth = 24;
    o->wave_spp = 1;
    o->compact_every = 1;
    o->bvh_width = 2;
}

int pt_device_count(int32_t* n) {
    if (!n) return PT_ERR_INVALID_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *n = c;
    return PT_OK;
}

// Message of the last failure that had no handle to hold it (pt_create,
// pt_env_cdf_noalloc), in a fixed per-thread buffer.
static thread_local char g_handleless_err[512] = "no error";

__attribute__((visibility("hidden"))) void pt_set_handleless_error(const char* msg) {
    std::snprintf(g_handleless_err, sizeof(g_handleless_err), "%s", msg);
}

int pt_create(int32_t device, pt_handle* out) {
    if (!out) {
        pt_set_handleless_error("pt_create: null output pointer");
        return PT_ERR_INVALID_ARG;
    }
    *out = nullptr;
    {
        const ptk::LayoutSizes a = ptk::layout_sizes(), b = ptk::layout_sizes_here();
        const ptk::TlasLayout c = ptk::tlas_layout(), d = ptk::tlas_layout_here();
        if (std::memcmp(&a, &b, sizeof(a)) != 0 || std::memcmp(&c, &d, sizeof(c)) != 0) {
            pt_set_handleless_error("pt_create: the library's kernels and host code were built from different "
                                    "headers (record layouts differ); rebuild it");
            return PT_ERR_STATE;
        }
    }
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
        pt_set_handleless_error("pt_create: no HIP device visible");
        return PT_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= n) {
        pt_set_handleless_error(("pt_create: device " + std::to_string(device) + " out of range (" + std::to_string(n) +
                                 " visible)").c_str());
        return PT_ERR_INVALID_ARG;
    }
    pt_renderer_s* h = new (std::nothrow) pt_renderer_s;
    if (!h) {
        pt_set_handleless_error("pt_create: out of host memory");
        return PT_ERR_OVERFLOW;
    }
    try {
        h->device = device;
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
        hipDeviceProp_t prop;
        HIPCHK(hipGetDeviceProperties(&prop, device));
        h->num_cu = prop.multiProcessorCount;
        h->lds_bytes = static_cast<int>(prop.sharedMemPerBlock);
        HIPCHK(hipEventCreate(&h->ev_r0));
        HIPCHK(hipEventCreate(&h->ev_r1));
        pt_default_options(&h->opts);
    } catch (const std::exception& e) {
        pt_set_handleless_error((std::string("pt_create: ") + e.what()).c_str());
        delete h;
        return PT_ERR_HIP;
    }
    *out = h;
    return PT_OK;
}

int pt_destroy(pt_handle h) {
    if (!h) return PT_OK;
    (void)hipSetDevice(h->device);
    delete h;
    return PT_OK;
}

// A null handle reports the last failure of this thread that had no handle
// (pt_create, pt_env_cdf_noalloc).
const char* pt_last_error(pt_handle h) { return h ? h->err.c_str() : g_handleless_err; }

// Uploads an RGB32F environment image and builds its CDF on the device;
// returns the total luminance (the CDF's last entry).
static float build_env(pt_renderer_s* h, const float* rgb, int width, int height) {
    const size_t n = static_cast<size_t>(width) * height;
    if (n > (1u << 28)) throw ArgError(PT_ERR_INVALID_ARG, "environment map too large");
    h->env_img.upload(rgb, 3 * n, h->stream);
    h->env_cdf.ensure(n);
    h->env_total.ensure(1);
    h->env_ok.ensure(1);
    h->env_scratch.ensure((ptk::env_cdf_scratch_bytes(static_cast<int>(n)) + 7) / 8);
    ptk::launch_env_cdf(h->stream, h->env_img.p, static_cast<int>(n), h->env_cdf.p, h->env_total.p, h->env_ok.p,
                        h->env_scratch.p);
    HIPCHK(hipGetLastError());
    float total = 0.f;
    HIPCHK(hipMemcpyAsync(&total, h->env_total.p, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return total;
}

int pt_upload_scene(pt_handle h, const pt_scene_view* v) {
    return Guard(h, [&] {
        if (!v || !v->nodes || v->num_nodes <= 0 || !v->vert_indices || v->num_tris <= 0 || !v->vertices_uvx ||
            !v->normals_uvy || v->num_vertices <= 0 || !v->materials || v->num_materials <= 0 || !v->transforms ||
            v->num_instances <= 0 || (v->num_lights > 0 && !v->lights))
            return Fail(h, PT_ERR_INVALID_ARG, "incomplete scene view");
        HIPCHK(hipSetDevice(h->device));
        HIPCHK(hipStreamSynchronize(h->stream));
        for (int i = 0; i < 3 * v->num_tris; ++i)
            if (v->vert_indices[i] < 0 || v->vert_indices[i] >= v->num_vertices)
                return Fail(h, PT_ERR_INVALID_ARG, "vert_indices entry out of range");
        Relayout R = relayout(v->nodes, v->num_nodes, v->top_level_index, v->num_tris, v->num_instances);
        for (int i = 0; i < v->num_instances; ++i)
            if (R.inst_material[i] < 0 || R.inst_material[i] >= v->num_materials)
                return Fail(h, PT_ERR_INVALID_ARG, "TLAS leaf material out of range");
        std::vector<GTri> tris;
        std::vector<int4> vidx;
        ptl::make_tris(*v, tris, vidx);
        std::vector<GInst> insts(v->num_instances);
        for (int i = 0; i < v->num_instances; ++i)
            insts[i] = make_inst(v->transforms + 16 * static_cast<size_t>(i), R.inst_blas_ref[i], R.inst_material[i],
                                 R.inst_blas4_ref[i]);
        std::vector<GLight> lights(std::max(1, v->num_lights));
        for (int i = 0; i < v->num_lights; ++i) lights[i] = make_light(v->lights[i]);
        hipStream_t s = h->stream;
        h->nodes.upload(R.nodes.data(), R.nodes.size(), s);
        h->nodes4.upload(R.wide.nodes.data(), R.wide.nodes.size(), s);
        h->tris.upload(tris.data(), tris.size(), s);
        h->vidx.upload(vidx.data(), vidx.size(), s);
        h->verts.upload(reinterpret_cast<const float4*>(v->vertices_uvx), v->num_vertices, s);
        h->norms.upload(reinterpret_cast<const float4*>(v->normals_uvy), v->num_vertices, s);
        h->mats.upload(reinterpret_cast<const float4*>(v->materials), 8 * static_cast<size_t>(v->num_materials), s);
        h->inst.upload(insts.data(), insts.size(), s);
        h->lights.upload(lights.data(), lights.size(), s);
        h->any_alpha = h->any_medium = h->any_blend = false;
        for (int i = 0; i < v->num_materials; ++i) {
            if (static_cast<int>(v->materials[i].alphaMode) != 0) h->any_alpha = true;
            if (static_cast<int>(v->materials[i].alphaMode) == ptd::ALPHA_MODE_BLEND) h->any_blend = true;
            if (static_cast<int>(v->materials[i].mediumType) != 0) h->any_medium = true;
        }
        if (v->num_textures > 0 && v->texture_maps) {
            h->tex.upload(v->texture_maps, 4 * static_cast<size_t>(v->tex_width) * v->tex_height * v->num_textures, s);
        }
        h->has_env = v->env_width > 0 && v->env_height > 0 && v->env_img;
        float env_total = v->env_total_sum;
        if (h->has_env) {
            const size_t n = static_cast<size_t>(v->env_width) * v->env_height;
            if (v->env_cdf) {
                h->env_img.upload(v->env_img, 3 * n, s);
                h->env_cdf.upload(v->env_cdf, n, s);
            } else {
                env_total = build_env(h, v->env_img, v->env_width, v->env_height);
            }
        }
        DevScene& sc = h->sc;
        sc.nodes = h->nodes.p; sc.tris = h->tris.p; sc.vidx = h->vidx.p; sc.verts = h->verts.p; sc.norms = h->norms.p;
        sc.mats = h->mats.p; sc.inst = h->inst.p; sc.lights = h->lights.p;
        sc.tex = (v->num_textures > 0 && v->texture_maps) ? h->tex.p : nullptr;
        sc.env_img = h->has_env ? h->env_img.p : nullptr;
        sc.env_cdf = h->has_env ? h->env_cdf.p : nullptr;
        h->set_tlas_roots(R);
        h->select_bvh();
        sc.num_lights = v->num_lights;
        sc.tex_w = v->tex_width; sc.tex_h = v->tex_height;
        sc.num_tex = sc.tex ? v->num_textures : 0;
        sc.env_w = h->has_env ? v->env_width : 0;
        sc.env_h = h->has_env ? v->env_height : 0;
        sc.env_total = h->has_env ? env_total : 0.f;
        h->flat.assign(v->nodes, v->nodes + v->num_nodes);
        h->top_level_index = v->top_level_index;
        h->inst_blas = R.inst_blas_ref;
        h->inst_blas4 = R.inst_blas4_ref;
        h->inst_mat = R.inst_material;
        h->root_ref = std::move(R.root_ref);
        h->root4 = std::move(R.wide.root4);
        h->tlas_rebuildable = R.tlas_rebuildable;
        h->tlas_on_device = false;
        h->tlas_base = R.tlas_base;
        h->t_local.upload(R.inst_local.data(), R.inst_local.size(), s);
        h->t_blas.upload(R.inst_blas_ref.data(), R.inst_blas_ref.size(), s);
        h->t_blas4.upload(R.inst_blas4_ref.data(), R.inst_blas4_ref.size(), s);
        h->t_mat.upload(R.inst_material.data(), R.inst_material.size(), s);
        h->num_tris = v->num_tris;
        h->num_instances = v->num_instances;
        h->num_materials = v->num_materials;
        HIPCHK(hipStreamSynchronize(s));
        h->has_scene = true;
        return PT_OK;
    });
}

int pt_update_instances(pt_handle h, const float* transforms, int32_t num_instances, const pt_material* materials,
                        int32_t num_materials, const pt_bvh_node* nodes, int32_t first_node, int32_t num_nodes) {
    return Guard(h, [&] {
        if (!h->has_scene) return Fail(h, PT_ERR_STATE, "no scene uploaded");
        if (num_instances != h->num_instances || !transforms) return Fail(h, PT_ERR_INVALID_ARG, "instance count mismatch");
        if (materials && num_materials != h->num_materials) return Fail(h, PT_ERR_INVALID_ARG, "material count mismatch");
        const int nflat = static_cast<int>(h->flat.size());
        if (nodes && (first_node < 0 || num_nodes < 0 || first_node + num_nodes > nflat))
            return Fail(h, PT_ERR_INVALID_ARG, "node range out of bounds");
        if (nodes && h->tlas_on_device && !(first_node <= h->top_level_index && first_node + num_nodes == nflat))
            return Fail(h, PT_ERR_STATE, "the TLAS was rebuilt on the device: pass the whole TLAS node range");
        HIPCHK(hipSetDevice(h->device));
        HIPCHK(hipStreamSynchronize(h->stream));
        std::vector<GInst> insts(num_instances);
        TlasRelayout T;
        const bool tlas_only = nodes && h->tlas_rebuildable && first_node == h->top_level_index &&
                               first_node + num_nodes == nflat &&
                               relayout_tlas(nodes, first_node, nflat, num_instances, h->tlas_base, h->root_ref, T);
        if (tlas_only) {
            // The TLAS range only: n - 1 nodes and the instance records cross the bus.
            for (int i = 0; i < num_instances; ++i)
                if (T.mat[i] < 0 || T.mat[i] >= h->num_materials)
                    return Fail(h, PT_ERR_INVALID_ARG, "TLAS leaf material out of range");
            const std::vector<int> blas4 = h->blas4_of(T.blas);
            for (int i = 0; i < num_instances; ++i)
                insts[i] = make_inst(transforms + 16 * static_cast<size_t>(i), T.blas[i], T.mat[i], blas4[i]);
            std::vector<float> local(6 * static_cast<size_t>(num_instances));
            for (int i = 0; i < num_instances; ++i)
                for (int j = 0; j < 3; ++j) {
                    local[6 * i + j] = h->flat[T.roots[i]].bboxmin[j];
                    local[6 * i + 3 + j] = h->flat[T.roots[i]].bboxmax[j];
                }
            if (num_instances > 1) {
                HIPCHK(hipMemcpyAsync(h->nodes.p + h->tlas_base, T.nodes.data(), T.nodes.size() * sizeof(GNode),
                                      hipMemcpyHostToDevice, h->stream));
                // the four-wide TLAS of the new nodes, in place (its root stays at t4_base)
                std::vector<ptd::GNode4> t4;
                ptl::collapse4_tlas(T.nodes.data(), h->tlas_base, h->tlas_base, h->t4_base, h->t4_room, t4);
                HIPCHK(hipMemcpyAsync(h->nodes4.p + h->t4_base, t4.data(), t4.size() * sizeof(ptd::GNode4),
                                      hipMemcpyHostToDevice, h->stream));
                HIPCHK(hipStreamSynchronize(h->stream));
            }
            h->t_local.upload(local.data(), local.size(), h->stream);
            h->t_blas.upload(T.blas.data(), T.blas.size(), h->stream);
            h->t_blas4.upload(blas4.data(), blas4.size(), h->stream);
            h->t_mat.upload(T.mat.data(), T.mat.size(), h->stream);
            std::copy(nodes, nodes + num_nodes, h->flat.begin() + first_node);
            h->inst_blas = T.blas;
            h->inst_blas4 = blas4;
            h->inst_mat = T.mat;
            h->tlas_on_device = false;
        } else if (nodes) {
            // Everything is checked before any state changes.
            std::vector<pt_bvh_node> flat = h->flat;
            std::copy(nodes, nodes + num_nodes, flat.begin() + first_node);
            Relayout R = relayout(flat.data(), nflat, h->top_level_index, h->num_tris, h->num_instances);
            for (int i = 0; i < num_instances; ++i)
                if (R.inst_material[i] < 0 || R.inst_material[i] >= h->num_materials)
                    return Fail(h, PT_ERR_INVALID_ARG, "TLAS leaf material out of range");
            for (int i = 0; i < num_instances; ++i)
                insts[i] = make_inst(transforms + 16 * static_cast<size_t>(i), R.inst_blas_ref[i], R.inst_material[i],
                                     R.inst_blas4_ref[i]);
            h->nodes.upload(R.nodes.data(), R.nodes.size(), h->stream);
            h->nodes4.upload(R.wide.nodes.data(), R.wide.nodes.size(), h->stream);
            h->t_local.upload(R.inst_local.data(), R.inst_local.size(), h->stream);
            h->t_blas.upload(R.inst_blas_ref.data(), R.inst_blas_ref.size(), h->stream);
            h->t_blas4.upload(R.inst_blas4_ref.data(), R.inst_blas4_ref.size(), h->stream);
            h->t_mat.upload(R.inst_material.data(), R.inst_material.size(), h->stream);
            h->flat.swap(flat);
            h->inst_blas = R.inst_blas_ref;
            h->inst_blas4 = R.inst_blas4_ref;
            h->inst_mat = R.inst_material;
            h->root_ref = std::move(R.root_ref);
            h->root4 = std::move(R.wide.root4);
            h->tlas_rebuildable = R.tlas_rebuildable;
            h->tlas_base = R.tlas_base;
            h->tlas_on_device = false;
            h->sc.nodes = h->nodes.p;
            h->set_tlas_roots(R);
            h->select_bvh();
        } else {
            // TLAS unchanged: only the instance records.
            for (int i = 0; i < num_instances; ++i)
                insts[i] = make_inst(transforms + 16 * static_cast<size_t>(i), h->inst_blas[i], h->inst_mat[i],
                                     h->inst_blas4[i]);
        }
        h->inst.upload(insts.data(), insts.size(), h->stream);
        if (materials) h->mats.upload(reinterpret_cast<const float4*>(materials), 8 * static_cast<size_t>(num_materials), h->stream);
        h->sc.inst = h->inst.p;
        h->sc.mats = h->mats.p;
        HIPCHK(hipStreamSynchronize(h->stream));
        return PT_OK;
    });
}

int pt_rebuild_instances(pt_handle h, const float* transforms, int32_t num_instances, const int32_t* material_ids) {
    return Guard(h, [&] {
        if (!h->has_scene) return Fail(h, PT_ERR_STATE, "no scene uploaded (pt_upload_scene)");
        if (!transforms || num_instances != h->num_instances) return Fail(h, PT_ERR_INVALID_ARG, "instance count mismatch");
        if (!h->tlas_rebuildable)
            return Fail(h, PT_ERR_UNSUPPORTED, "the uploaded TLAS is not a one-instance-per-leaf preorder tree at the end of the node array");
        const int n = num_instances;
        for (int i = 0; i < n; ++i) {
            GInst g;
            if (!ptd::inst_record(transforms + 16 * static_cast<size_t>(i), 0, 0, &g))
                return Fail(h, PT_ERR_INVALID_ARG, "singular instance transform");
        }
        if (material_ids)
            for (int i = 0; i < n; ++i)
                if (material_ids[i] < 0 || material_ids[i] >= h->num_materials)
                    return Fail(h, PT_ERR_INVALID_ARG, "instance material out of range");
        HIPCHK(hipSetDevice(h->device));
        hipStream_t s = h->stream;
        const size_t nn = static_cast<size_t>(n), half = nn / 2 + 1;
        h->t_xf.upload(transforms, 16 * nn, s);
        if (material_ids) {
            h->t_mat.upload(material_ids, nn, s);
            h->inst_mat.assign(material_ids, material_ids + n);
        }
        h->t_box.ensure(6 * nn);
        h->t_cen.ensure(3 * nn);
        h->t_ord.ensure(2 * nn);
        h->t_pre.ensure(nn + 2);
        h->t_pos.ensure(2 * nn);
        h->t_pseg.ensure(nn);
        h->t_seg.ensure(ptd::kTlasSegWords * half);
        h->t_keys.ensure(24 * half);
        ptd::TlasBuild a{};
        a.n = n;
        a.base = h->tlas_base;
        a.xf = h->t_xf.p;
        a.local = h->t_local.p;
        a.blas_ref = h->t_blas.p;
        a.blas4_ref = h->t_blas4.p;
        a.mat = h->t_mat.p;
        a.inst = h->inst.p;
        a.nodes = h->nodes.p;
        a.box = h->t_box.p;
        a.cen = h->t_cen.p;
        a.ord0 = h->t_ord.p;
        a.ord1 = h->t_ord.p + nn;
        a.pre = h->t_pre.p;
        a.apos = h->t_pos.p;
        a.bpos = h->t_pos.p + nn;
        a.pseg = h->t_pseg.p;
        a.seg = h->t_seg.p;
        a.keys = h->t_keys.p;
        a.negz = h->t_pre.p + nn + 1;
        a.lds_bytes = h->lds_bytes;
        HIPCHK(hipMemsetAsync(a.negz, 0, sizeof(int), s));
        ptk::launch_tlas_rebuild(s, a);
        HIPCHK(hipGetLastError());
        if (n >= 2) {
            // the four-wide TLAS of the rebuilt nodes (the order arrays are free again)
            ptk::launch_tlas4(s, h->nodes.p, h->tlas_base, h->nodes4.p + h->t4_base, h->t4_base, n - 1, h->t_ord.p,
                              h->t_ord.p + nn);
            HIPCHK(hipGetLastError());
        }
        HIPCHK(hipStreamSynchronize(s));
        h->tlas_on_device = true;
        return PT_OK;
    });
}

int pt_read_tlas(pt_handle h, float* nodes, int32_t num_nodes, float* instances, int32_t num_instances, int32_t* root) {
    return Guard(h, [&] {
        if (!h->has_scene) return Fail(h, PT_ERR_STATE, "no scene uploaded (pt_upload_scene)");
        const int n = h->num_instances;
        if (nodes && !h->tlas_rebuildable)
            return Fail(h, PT_ERR_UNSUPPORTED, "the uploaded TLAS is not a one-instance-per-leaf preorder tree at the end of the node array");
        if (nodes && num_nodes != n - 1) return Fail(h, PT_ERR_INVALID_ARG, "nodes must hold instances - 1 records");
        if (instances && num_instances != n) return Fail(h, PT_ERR_INVALID_ARG, "instances must hold one record per instance");
        HIPCHK(hipSetDevice(h->device));
        if (nodes && n > 1)
            HIPCHK(hipMemcpyAsync(nodes, h->nodes.p + h->tlas_base, static_cast<size_t>(n - 1) * sizeof(GNode),
                                  hipMemcpyDeviceToHost, h->stream));
        if (instances)
            HIPCHK(hipMemcpyAsync(instances, h->inst.p, static_cast<size_t>(n) * sizeof(GInst), hipMemcpyDeviceToHost,
                                  h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (root) *root = h->sc.root;
        return PT_OK;
    });
}

int pt_read_tlas4(pt_handle h, float* nodes, int32_t num_nodes, int32_t* room, int32_t* root) {
    return Guard(h, [&] {
        if (!h->has_scene) return Fail(h, PT_ERR_STATE, "no scene uploaded (pt_upload_scene)");
        if (room) *room = h->t4_room;
        if (nodes) {
            if (num_nodes != h->t4_room) return Fail(h, PT_ERR_INVALID_ARG, "nodes must hold the four-wide TLAS room");
            HIPCHK(hipSetDevice(h->device));
            HIPCHK(hipMemcpyAsync(nodes, h->nodes4.p + h->t4_base, static_cast<size_t>(h->t4_room) * sizeof(ptd::GNode4),
                                  hipMemcpyDeviceToHost, h->stream));
            HIPCHK(hipStreamSynchronize(h->stream));
        }
        if (root) *root = h->root4_tlas;
        return PT_OK;
    });
}

int pt_set_env_map(pt_handle h, const float* rgb, int32_t width, int32_t height) {
    return Guard(h, [&] {
        if (!h->has_scene) return Fail(h, PT_ERR_STATE, "no scene uploaded (pt_upload_scene)");
        if (!rgb || width <= 0 || height <= 0) return Fail(h, PT_ERR_INVALID_ARG, "invalid environment map");
        HIPCHK(hipSetDevice(h->device));
        HIPCHK(hipStreamSynchronize(h->stream));
        const float total = build_env(h, rgb, width, height);
        h->has_env = true;
        h->sc.env_img = h->env_img.p;
        h->sc.env_cdf = h->env_cdf.p;
        h->sc.env_w = width;
        h->sc.env_h = height;
        h->sc.env_total = total;
        return PT_OK;
    });
}

int pt_read_env_cdf(pt_handle h, float* cdf, size_t count, float* total) {
    return Guard(h, [&] {
        if (!h->has_env) return Fail(h, PT_ERR_STATE, "no environment map");
        const size_t n = static_cast<size_t>(h->sc.env_w) * h->sc.env_h;
        if (!cdf || count != n) return Fail(h, PT_ERR_INVALID_ARG, "cdf buffer must hold width*height floats");
        HIPCHK(hipSetDevice(h->device));
        HIPCHK(hipMemcpyAsync(cdf, h->env_cdf.p, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if (total) *total = h->sc.env_total;
        return PT_OK;
    });
}

int pt_set_options(pt_handle h, const pt_options* o) {
    return Guard(h, [&] {
        if (!o || o->width <= 0 || o->height <= 0 || o->max_depth < 0) return Fail(h, PT_ERR_INVALID_ARG, "invalid options");
        if (o->block_size != 0 && o->block_size != 64 && o->block_size != 128 && o->block_size != 256)
            return Fail(h, PT_ERR_INVALID_ARG, "block_size must be 64, 128 or 256");
        if (o->stack_depth < 0 || o->stack_depth > ptd::MAXD) return Fail(h, PT_ERR_INVALID_ARG, "stack_depth out of range");
        if (o->mode != PT_MODE_PLAIN && o->mode != PT_MODE_WAVEFRONT) return Fail(h, PT_ERR_INVALID_ARG, "unknown mode");
        if (o->compact_every < 0 || o->compact_every > 64) return Fail(h, PT_ERR_INVALID_ARG, "compact_every out of range");
        if (static_cast<long long>(o->width) * o->height > (1LL << 28)) return Fail(h, PT_ERR_INVALID_ARG, "resolution too large");
        if (o->bvh_width != 0 && o->bvh_width != 2 && o->bvh_width != 4)
            return Fail(h, PT_ERR_INVALID_ARG, "bvh_width must be 2 or 4");
        if (o->nee_inline != 0 && o->nee_inline != 1) return Fail(h, PT_ERR_INVALID_ARG, "nee_inline must be 0 or 1");
        if (o->extend_cap < 0 || o->extend_cap > 4096) return Fail(h, PT_ERR_INVALID_ARG, "extend_cap must be in [0, 4096]");
        HIPCHK(hipSetDevice(h->device));
        h->opts = *o;
        h->has_opts = true;
        h->select_bvh();
        if (o->width != h->accum_w || o->height != h->accum_h) {
            HIPCHK(hipStreamSynchronize(h->stream));
            h->accum.ensure(static_cast<size_t>(o->width) * o->height);
            h->accum_w = o->width;
            h->accum_h = o->height;
            HIPCHK(hipMemsetAsync(h->accum.p, 0, sizeof(float4) * h->accum.n, h->stream));
        }
        return PT_OK;
    });
}

int pt_set_camera(pt_handle h, const pt_camera* c) {
    return Guard(h, [&] {
        if (!c) return Fail(h, PT_ERR_INVALID_ARG, "null camera");
        h->cam = *c;
        h->has_cam = true;
        return PT_OK;
    });
}

int pt_clear_accum(pt_handle h) {
    return Guard(h, [&] {
        if (!h->accum.p) return Fail(h, PT_ERR_STATE, "no accumulation buffer (pt_set_options)");
        HIPCHK(hipSetDevice(h->device));
        HIPCHK(hipMemsetAsync(h->accum.p, 0, sizeof(float4) * static_cast<size_t>(h->accum_w) * h->accum_h, h->stream));
        return PT_OK;
    });
}

int pt_render(pt_handle h, int32_t first_sample, int32_t num_samples) {
    return Guard(h, [&] {
        check_ready(h);
        if (first_sample < 0 || num_samples < 0) return Fail(h, PT_ERR_INVALID_ARG, "negative sample range");
        if (num_samples == 0) return PT_OK;
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        if (h->opts.count_work) {
            h->ensure_misc();
            HIPCHK(hipMemsetAsync(h->work_counts.p, 0, 16 * sizeof(unsigned long long), h->stream));
        }
        h->ensure_misc();
        HIPCHK(hipMemsetAsync(h->ovf.p, 0, sizeof(int), h->stream));
        HIPCHK(hipEventRecord(h->ev_r0, h->stream));
        if (h->opts.mode == PT_MODE_PLAIN) render_plain(h, first_sample, num_samples);
        else render_wavefront(h, first_sample, num_samples);
        HIPCHK(hipEventRecord(h->ev_r1, h->stream));
        h->render_pending = true;
        return PT_OK;
    });
}

int pt_synchronize(pt_handle h) {
    return Guard(h, [&] {
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        HIPCHK(hipStreamSynchronize(h->stream));
        return h->stats.stack_overflow ? Fail(h, PT_ERR_OVERFLOW, "BVH traversal exceeded the maximum stack depth") : PT_OK;
    });
}

int pt_read_accum(pt_handle h, float* rgba, size_t num_floats) {
    return Guard(h, [&] {
        if (!h->accum.p) return Fail(h, PT_ERR_STATE, "no accumulation buffer");
        const size_t need = 4 * static_cast<size_t>(h->accum_w) * h->accum_h;
        if (!rgba || num_floats < need) return Fail(h, PT_ERR_INVALID_ARG, "output buffer too small");
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        HIPCHK(hipMemcpyAsync(rgba, h->accum.p, need * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return PT_OK;
    });
}

int pt_read_ldr(pt_handle h, int32_t num_samples, int32_t enable_tonemap, int32_t enable_aces, int32_t simple_aces_fit,
                const float background[3], uint8_t* rgba, size_t num_bytes) {
    return Guard(h, [&] {
        if (!h->accum.p) return Fail(h, PT_ERR_STATE, "no accumulation buffer");
        const size_t npx = static_cast<size_t>(h->accum_w) * h->accum_h;
        if (!rgba || num_bytes < 4 * npx || num_samples <= 0) return Fail(h, PT_ERR_INVALID_ARG, "invalid LDR request");
        HIPCHK(hipSetDevice(h->device));
        h->ldr.ensure(npx);
        const float bg0[3] = {1.f, 1.f, 1.f};
        int bg_mode = h->opts.transparent_background ? 2 : (h->opts.enable_background ? 1 : 0);
        ptk::launch_tonemap(h->stream, h->accum.p, h->accum_w, h->accum_h, 1.0f / static_cast<float>(num_samples),
                            enable_tonemap, enable_aces, simple_aces_fit, bg_mode, background ? background : bg0, h->ldr.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(rgba, h->ldr.p, 4 * npx, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return PT_OK;
    });
}

int pt_get_stats(pt_handle h, pt_stats* out) {
    return Guard(h, [&] {
        if (!out) return Fail(h, PT_ERR_INVALID_ARG, "null stats");
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        *out = h->stats;
        return PT_OK;
    });
}

int pt_reset_stats(pt_handle h) {
    return Guard(h, [&] {
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        std::memset(&h->stats, 0, sizeof(h->stats));
        return PT_OK;
    });
}

int pt_closest_hit(pt_handle h, const float* origins, const float* dirs, int32_t n, float* t, int32_t* instance,
                   int32_t* triangle, int32_t* light, float* uv) {
    return Guard(h, [&] {
        if (!h->has_scene) return Fail(h, PT_ERR_STATE, "no scene uploaded");
        if (n < 0 || (n > 0 && (!origins || !dirs || !t || !instance || !triangle || !light || !uv)))
            return Fail(h, PT_ERR_INVALID_ARG, "invalid query arrays");
        if (n == 0) return PT_OK;
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        DevBuf<float> o, d, tt, uvb;
        DevBuf<int> ib, tb, lb;
        o.upload(origins, 3 * static_cast<size_t>(n), h->stream);
        d.upload(dirs, 3 * static_cast<size_t>(n), h->stream);
        tt.ensure(n); uvb.ensure(2 * static_cast<size_t>(n)); ib.ensure(n); tb.ensure(n); lb.ensure(n);
        h->ensure_misc();
        const int S = h->has_opts ? h->stack_S() : 24;
        const size_t grid = ((static_cast<size_t>(n) + 255) / 256) * 256;
        DevBuf<int> spill;
        spill.ensure(grid * static_cast<size_t>(ptd::MAXD - S + 1));
        HIPCHK(hipMemsetAsync(h->ovf.p, 0, sizeof(int), h->stream));
        DevOpts op = h->has_opts ? h->dev_opts() : DevOpts{};
        if (!h->has_opts && h->sc.num_lights > 0) op.flags |= ptd::OPT_LIGHTS;
        ptk::launch_query_closest(h->stream, h->sc, op, o.p, d.p, n, tt.p, ib.p, tb.p, lb.p, uvb.p, S, spill.p, h->ovf.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(t, tt.p, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(instance, ib.p, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(triangle, tb.p, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(light, lb.p, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(uv, uvb.p, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, h->stream));
        int f = 0;
        HIPCHK(hipMemcpyAsync(&f, h->ovf.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        o.free(); d.free(); tt.free(); uvb.free(); ib.free(); tb.free(); lb.free(); spill.free();
        return f ? Fail(h, PT_ERR_OVERFLOW, "BVH traversal exceeded the maximum stack depth") : PT_OK;
    });
}

int pt_any_hit(pt_handle h, const float* origins, const float* dirs, const float* max_dist, int32_t n, int32_t* occluded) {
    return Guard(h, [&] {
        if (!h->has_scene) return Fail(h, PT_ERR_STATE, "no scene uploaded");
        if (n < 0 || (n > 0 && (!origins || !dirs || !max_dist || !occluded))) return Fail(h, PT_ERR_INVALID_ARG, "invalid query arrays");
        if (n == 0) return PT_OK;
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        DevBuf<float> o, d, m;
        DevBuf<int> oc;
        o.upload(origins, 3 * static_cast<size_t>(n), h->stream);
        d.upload(dirs, 3 * static_cast<size_t>(n), h->stream);
        m.upload(max_dist, n, h->stream);
        oc.ensure(n);
        h->ensure_misc();
        const int S = h->has_opts ? h->stack_S() : 24;
        const size_t grid = ((static_cast<size_t>(n) + 255) / 256) * 256;
        DevBuf<int> spill;
        spill.ensure(grid * static_cast<size_t>(ptd::MAXD - S + 1));
        HIPCHK(hipMemsetAsync(h->ovf.p, 0, sizeof(int), h->stream));
        DevOpts op = h->has_opts ? h->dev_opts() : DevOpts{};
        if (!h->has_opts && h->sc.num_lights > 0) op.flags |= ptd::OPT_LIGHTS;
        ptk::launch_query_any(h->stream, h->sc, op, o.p, d.p, m.p, n, oc.p, S, spill.p, h->ovf.p);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(occluded, oc.p, sizeof(int) * n, hipMemcpyDeviceToHost, h->stream));
        int f = 0;
        HIPCHK(hipMemcpyAsync(&f, h->ovf.p, sizeof(int), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        o.free(); d.free(); m.free(); oc.free(); spill.free();
        return f ? Fail(h, PT_ERR_OVERFLOW, "BVH traversal exceeded the maximum stack depth") : PT_OK;
    });
}

int pt_batch_render(pt_handle h, const pt_job* jobs, int32_t num_jobs) {
    return Guard(h, [&] {
        if (!h->has_scene) return Fail(h, PT_ERR_STATE, "no scene uploaded (pt_upload_scene)");
        if (!h->has_opts) return Fail(h, PT_ERR_STATE, "no options set (pt_set_options)");
        if (num_jobs < 0 || (num_jobs > 0 && !jobs)) return Fail(h, PT_ERR_INVALID_ARG, "invalid job list");
        if (num_jobs == 0) return PT_OK;
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        h->ensure_misc();
        HIPCHK(hipMemsetAsync(h->ovf.p, 0, sizeof(int), h->stream));
        HIPCHK(hipEventRecord(h->ev_r0, h->stream));
        batch_render(h, jobs, num_jobs);
        HIPCHK(hipEventRecord(h->ev_r1, h->stream));
        h->render_pending = true;
        h->resolve_events();
        return h->stats.stack_overflow ? Fail(h, PT_ERR_OVERFLOW, "BVH traversal exceeded the maximum stack depth") : PT_OK;
    });
}

// pt_adjoint (medium false) and pt_adjoint_medium (medium true).
static int adjoint_call(pt_handle h, int32_t first_sample, int32_t num_samples, const float* dL_dimage,
                        float* d_params, float* d_medium, bool medium) {
    return Guard(h, [&] {
        check_ready(h);
        if (first_sample < 0 || num_samples <= 0) return Fail(h, PT_ERR_INVALID_ARG, "invalid sample range");
        if (!dL_dimage || !d_params || (medium && !d_medium))
            return Fail(h, PT_ERR_INVALID_ARG, "null gradient buffers");
        if (!medium) d_medium = nullptr;
        if (h->num_materials <= 0) return Fail(h, PT_ERR_STATE, "scene has no materials");
        HIPCHK(hipSetDevice(h->device));
        h->resolve_events();
        const size_t npx = static_cast<size_t>(h->opts.width) * h->opts.height;
        const size_t np = static_cast<size_t>(h->num_materials) * ptd::NPARAM;
        const size_t ng = np + (d_medium ? static_cast<size_t>(h->num_materials) * ptd::NMEDIUM : 0);
        h->a_dimg.upload(dL_dimage, 3 * npx, h->stream);
        h->a_grad.ensure(ng);
        h->a_flags.ensure(2);
        ptk::launch_fill(h->stream, reinterpret_cast<int*>(h->a_grad.p), static_cast<int>(ng), 0);
        ptk::launch_fill(h->stream, h->a_flags.p, 2, 0);
        h->ensure_misc();
        HIPCHK(hipMemsetAsync(h->ovf.p, 0, sizeof(int), h->stream));
        HIPCHK(hipEventRecord(h->ev_r0, h->stream));
        adjoint_wavefront(h, first_sample, num_samples, 1.0f / static_cast<float>(num_samples), d_medium != nullptr);
        HIPCHK(hipEventRecord(h->ev_r1, h->stream));
        h->render_pending = true;
        std::vector<float> g(ng);
        int flags[2] = {0, 0};
        HIPCHK(hipMemcpyAsync(g.data(), h->a_grad.p, ng * sizeof(float), hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipMemcpyAsync(flags, h->a_flags.p, sizeof(flags), hipMemcpyDeviceToHost, h->stream));
        h->resolve_events();
        h->stats.adjoint_dropped = flags[1];
        if (flags[0]) return Fail(h, PT_ERR_STATE, "adjoint: a path vertex named a material outside the scene's table");
        for (size_t i = 0; i < np; ++i) d_params[i] += g[i];
        for (size_t i = np; i < ng; ++i) d_medium[i - np] += g[i];
        return h->stats.stack_overflow ? Fail(h, PT_ERR_OVERFLOW, "BVH traversal exceeded the maximum stack depth") : PT_OK;
    });
}

int pt_adjoint(pt_handle h, int32_t first_sample, int32_t num_samples, const float* dL_dimage, float* d_params) {
    return adjoint_call(h, first_sample, num_samples, dL_dimage, d_params, nullptr, false);
}

int pt_adjoint_medium(pt_handle h, int32_t first_sample, int32_t num_samples, const float* dL_dimage,
                      float* d_params, float* d_medium) {
    return adjoint_call(h, first_sample, num_samples, dL_dimage, d_params, d_medium, true);
}

}  // extern "C"
