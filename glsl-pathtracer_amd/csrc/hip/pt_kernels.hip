// Kernels of the path: plain one-launch-per-frame megakernel, wavefront
// stages, ray queries, accumulation and tonemapping. gfx950, wave64. The
// per-path functions they share are in pt_path.h.
//
// Reference behaviour (file:line in knightcrawler25/GLSL-PathTracer):
//   bounce loop         src/shaders/common/pathtrace.glsl:285-476
//   tonemap             src/shaders/tonemap.glsl:59-133
#include "pt_kernels.h"
#include "pt_path.h"
#include "pt_replay.h"

namespace ptd {

// ------------------------------------------------------------------ helpers
__device__ __forceinline__ int lane_id() { return __lane_id(); }

// Work distribution of the persistent kernels. Every wave's first batch of 64
// items is static (global wave index * 64); later batches are claimed in runs
// of RUN batches with one atomic per wave on `counter`, offset past the
// static range. All waves of a grid share that one counter, and a wave waits
// for the contended atomic's return before its next batch: claimed one batch
// at a time by the camera-ray stage at 1080p it took half of the stage's time.
// Runs trade that against the tail a long run left at the end of the queue
// holds. Measured at 1080p (ms per frame): camera rays in runs of 4, 20.9 ->
// 19.4 in the closest-hit stage against a static assignment (which gives a
// wave tiles far apart); shade, whose items cost about the same, 8.2 -> 5.6
// from runs of 1 to 16; shadow and later closest-hit bounces, whose items vary
// with the traversal, best one batch at a time (shadow 6.5 -> 9.2 at 16).
constexpr int RUN_CAMERA = 4, RUN_TRAVERSE = 1, RUN_SHADE = 16;
__device__ __forceinline__ int static_total() { return static_cast<int>(gridDim.x * (blockDim.x >> 6)) * 64; }
__device__ __forceinline__ int first_batch() {
    return static_cast<int>(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64;
}
template <int RUN>
struct WorkCursor {
    int base, end;  // current batch; end of the claimed run
    __device__ __forceinline__ WorkCursor() : base(first_batch()), end(first_batch() + 64) {}
    __device__ __forceinline__ void next(int* counter) {
        base += 64;
        if (base < end) return;
        int b = 0;
        if (lane_id() == 0) b = atomicAdd(counter, 64 * RUN);
        base = __builtin_amdgcn_readfirstlane(b) + static_total();
        end = base + 64 * RUN;
    }
};

// Wave-level stream compaction: one atomic per wave, lanes write at their
// prefix position. Must be called by every lane of the wave.
__device__ __forceinline__ int wave_append(bool keep, int* counter) {
    const uint64_t mask = __ballot(keep);
    const int n = __popcll(mask);
    const int prefix = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mask >> 32),
                                                 __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mask), 0u));
    int base = 0;
    if (lane_id() == 0 && n > 0) base = atomicAdd(counter, n);
    base = __builtin_amdgcn_readfirstlane(base);
    return keep ? base + prefix : -1;
}

template <bool COUNT>
__device__ __forceinline__ void flush_counts(const TravCount& c, unsigned long long* out) {
    if (!COUNT) return;
    atomicAdd(out + 0, static_cast<unsigned long long>(c.nodes));
    atomicAdd(out + 1, static_cast<unsigned long long>(c.tris));
    atomicAdd(out + 2, static_cast<unsigned long long>(c.insts));
    atomicAdd(out + 3, static_cast<unsigned long long>(c.lights));
    atomicAdd(out + 4, static_cast<unsigned long long>(c.spills));
    atomicAdd(out + 5, static_cast<unsigned long long>(c.rays));
}

// Pixel of work item `i`: 8x8 pixel tiles, one tile per wave, rows bottom-up.
__device__ __forceinline__ void item_pixel(int i, int width, int& px, int& py) {
    const int tiles_x = (width + 7) >> 3;
    const int tile = i >> 6, l = i & 63;
    px = (tile % tiles_x) * 8 + (l & 7);
    py = (tile / tiles_x) * 8 + (l >> 3);
}

// ------------------------------------------------------------------ plain kernel
template <bool COUNT>
__global__ void __launch_bounds__(256) k_plain(DevScene sc, DevOpts op, float4* __restrict__ accum, int sample,
                                               int num_items, int* work, int S, int* spill,
                                               unsigned long long* counters, int* ovf_flag) {
    extern __shared__ int lds[];
    Stack stk{lds, spill, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x), S,
              blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x,
              static_cast<size_t>(gridDim.x) * blockDim.x};
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    for (WorkCursor<RUN_TRAVERSE> wc; wc.base < num_items; wc.next(work)) {
        const int i = wc.base + lane_id();
        int px, py;
        item_pixel(i, op.width, px, py);
        if (i < num_items && px < op.width && py < op.height) {
            PathRegs p;
            init_path(p);
            p.rng = init_rng(static_cast<uint32_t>(px), static_cast<uint32_t>(py), static_cast<uint32_t>(sample + 2));
            camera_ray(op, px, py, p.rng, p.ro, p.rd);
            float4 c = path_trace<COUNT>(sc, op, p, stk, cnt, ovf);
            float4* a = accum + static_cast<size_t>(py) * op.width + px;
            float4 acc = *a;
            acc.x += c.x; acc.y += c.y; acc.z += c.z; acc.w += c.w;
            *a = acc;
        }
    }
    flush_counts<COUNT>(cnt, counters);
    if (ovf) atomicOr(ovf_flag, 1);
}

// ------------------------------------------------------------------ wavefront
// Path slot -> pixel: slot = sample_in_pass * npix + item.
__device__ __forceinline__ bool slot_pixel(const DevOpts& op, int slot, int npix, int& s, int& px, int& py) {
    s = slot / npix;
    item_pixel(slot - s * npix, op.width, px, py);
    return px < op.width && py < op.height;
}

// meta.y: pass-through count (bits 0-15), in-medium (16), surface-scatter (17),
// ended (18, PATH_DEAD), medium type (20-21).
constexpr int MEDIUM_TYPE_SHIFT = 20;

__device__ __forceinline__ void load_path(const PathBuf& pb, int slot, PathRegs& p) {
    float4 o4 = pb.ro[slot], d4 = pb.rd[slot], t4 = pb.thr[slot], r4 = pb.rad[slot], a4 = pb.aux[slot];
    uint4 g = pb.rng[slot];
    int4 mt = pb.meta[slot];
    p.ro = xyz(o4); p.prev_pdf = o4.w;
    p.rd = xyz(d4); p.prev_rough = d4.w;
    p.thr = xyz(t4);
    p.rad = xyz(r4); p.alpha = r4.w;
    p.rng = Rng{g.x, g.y, g.z, g.w};
    p.depth = mt.x; p.pass = mt.y & 0xffff;
    p.in_medium = (mt.y >> 16) & 1; p.surface_scatter = (mt.y >> 17) & 1;
    p.med.type = (mt.y >> MEDIUM_TYPE_SHIFT) & 3;
    p.prev_mat = mt.z; p.med_mat = mt.w;
    p.tu = a4.x; p.tv = a4.y; p.med.anisotropy = a4.z;
    if (pb.med) { float4 m4 = pb.med[slot]; p.med.color = xyz(m4); p.med.density = m4.w; }
    else { p.med.color = V1(0.f); p.med.density = 0.f; }
}

__device__ __forceinline__ void store_path(const PathBuf& pb, int slot, const PathRegs& p) {
    pb.ro[slot] = F4(p.ro, p.prev_pdf);
    pb.rd[slot] = F4(p.rd, p.prev_rough);
    pb.thr[slot] = F4(p.thr, 0.f);
    pb.rad[slot] = F4(p.rad, p.alpha);
    pb.rng[slot] = make_uint4(p.rng.x, p.rng.y, p.rng.z, p.rng.w);
    pb.meta[slot] = make_int4(p.depth,
                              (p.pass & 0xffff) | (p.in_medium ? 1 << 16 : 0) | (p.surface_scatter ? 1 << 17 : 0) |
                                  (p.med.type << MEDIUM_TYPE_SHIFT),
                              p.prev_mat, p.med_mat);
    pb.aux[slot] = make_float4(p.tu, p.tv, p.med.anisotropy, 0.f);
    if (pb.med) pb.med[slot] = F4(p.med.color, p.med.density);
}

// meta.y flag of a path that ended in a bounce whose queue was not compacted.
constexpr int PATH_DEAD = 1 << 18;

// Camera ray and initial state of a primary path (tile.glsl:41-69); the
// bounce-0 stages recompute it instead of storing it.
__device__ __forceinline__ void primary_path(const DevOpts& op, int px, int py, int sample, PathRegs& p) {
    init_path(p);
    p.rng = init_rng(static_cast<uint32_t>(px), static_cast<uint32_t>(py), static_cast<uint32_t>(sample + 2));
    camera_ray(op, px, py, p.rng, p.ro, p.rd);
}

// ---- batched render: slot -> (job, sample, pixel item) of the current pass
__device__ __forceinline__ void batch_decode(const BatchBuf& bb, int slot, int& job, int& sample, int& item) {
    const int p0 = bb.cur[1], np = bb.cur[2];
    int lo = p0, hi = p0 + np - 1;  // last plane with slot0 <= slot
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (bb.planes[mid].slot0 <= slot) lo = mid;
        else hi = mid - 1;
    }
    const PlaneDev pl = bb.planes[lo];
    job = pl.job;
    sample = pl.sample;
    item = slot - pl.slot0;
}

// Options of one job: shared switches, the job's camera, size and depth.
__device__ __forceinline__ DevOpts job_opts(const DevOpts& op, const JobDev& J) {
    DevOpts o = op;
    o.width = J.width;
    o.height = J.height;
    o.max_depth = J.max_depth;
    for (int i = 0; i < 3; ++i) {
        o.cam_pos[i] = J.cam_pos[i];
        o.cam_up[i] = J.cam_up[i];
        o.cam_right[i] = J.cam_right[i];
        o.cam_fwd[i] = J.cam_fwd[i];
    }
    o.cam_fov = J.cam_fov;
    o.cam_focal = J.cam_focal;
    o.cam_aperture = J.cam_aperture;
    return o;
}

// End of a path's closest-hit step: a miss finishes the path here (environment
// or uniform light), a hit is recorded for the shade stage (returns true).
template <bool PRIMARY>
__device__ __forceinline__ bool finish_extend(const DevScene& sc, const DevOpts& op, const PathBuf& pb, int slot,
                                              PathRegs& p, const Hit& h) {
    if (h.t == INF) {
        if (!PRIMARY) {
            const float4 o4 = pb.ro[slot], t4 = pb.thr[slot], r4 = pb.rad[slot];
            p.prev_pdf = o4.w; p.thr = xyz(t4); p.rad = xyz(r4); p.alpha = r4.w;
            p.surface_scatter = (pb.meta[slot].y >> 17) & 1;
        }
        on_miss(sc, op, p);
        pb.rad[slot] = F4(p.rad, p.alpha);
        if (!PRIMARY) pb.meta[slot].y = PATH_DEAD;
        return false;
    }
    pb.hit[slot] = make_float4(h.t, h.u, h.v, __int_as_float(h.tri));
    pb.hit2[slot] = make_int2(h.inst, h.light);
    return true;
}

// Record r of the continuation buffer: a stopped traversal, its stack and
// the hit-queue position kept for it (pos).
__device__ __forceinline__ void save_cont(const ContBuf& cb, int r, int slot, const TravState& ts, const Hit& h,
                                          Stack& stk, int pos) {
    cb.slot[r] = slot;
    cb.st[r] = make_int4(ts.ref, ts.sp, ts.blas_base, ts.cur_inst);
    cb.hf[r] = make_float4(h.t, h.u, h.v, 0.f);
    cb.hi[r] = make_int4(h.tri, h.inst, h.light, pos);
    for (int k = 0; k < ts.sp; ++k) cb.stack[static_cast<size_t>(k) * cb.cap + r] = stk.pop(k);
}

// Extend: closest hit of every queued path. Misses finish here (environment
// or uniform light, pathtrace.glsl:305-339) and write only their radiance;
// hits are compacted into the shade queue. At bounce 0 (PRIMARY) the input is
// every slot of the pass and the camera ray is generated in place; BATCH
// (bounce 0 of a batched pass) takes the slot count and each slot's job,
// sample and pixel from the device-side pass state. Seven waves per SIMD (72
// registers): at eight (64) more values spill and the traversal loop reloads
// them around its stack pushes, which costs issue slots in a stage bound by
// VALU issue (1080p: 19.6 -> 17.9 ms per frame; 4K heavy scene: 630 -> 569 ms;
// six waves, 80 registers, is slower on both).
// CAPPED (later bounces): a traversal runs at most `budget` loop iterations
// here; a wave does not wait on its longest rays. Stopped traversals go to the
// continuation buffer (while it has room and their stack is at most
// CONT_STACK deep; otherwise they finish here) and k_extend_resume finishes
// them, densely packed. The hits are the same either way. A stopped traversal
// keeps its place in the hit queue beside the batch's finished hits, so the
// shade stage reads the paths of a batch together (k_extend_resume writes the
// slot there, or -1 for a miss, which the shade stage skips).
template <bool COUNT, bool PRIMARY, bool BATCH, bool CAPPED = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 7))) k_extend(DevScene sc, DevOpts op, PathBuf pb, const int* __restrict__ qin,
                                                const int* count_in, int npaths, int npix, int first_sample,
                                                int* hitq, int* hit_count, int* work, int S, int* spill,
                                                unsigned long long* counters, int* ovf_flag, BatchBuf bb,
                                                int budget, ContBuf cb, int* cont_count) {
    static_assert(!CAPPED || (!COUNT && !PRIMARY), "only uninstrumented later bounces are capped");
    extern __shared__ int lds[];
    Stack stk{lds, spill, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x), S,
              blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x,
              static_cast<size_t>(gridDim.x) * blockDim.x};
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    const int count = PRIMARY ? (BATCH ? bb.cur[3] : npaths) : *count_in;
    for (WorkCursor<PRIMARY ? RUN_CAMERA : RUN_TRAVERSE> wc; wc.base < count; wc.next(work)) {
        const int i = wc.base + lane_id();
        bool keep = false, valid = false, done = false;
        int slot = -1;
        PathRegs p;
        Hit h;
        TravState ts;
        if (i < count) {
            slot = PRIMARY ? i : qin[i];
            valid = true;
            if (PRIMARY && BATCH) {
                int job, smp, item, px, py;
                batch_decode(bb, slot, job, smp, item);
                const DevOpts oj = job_opts(op, bb.jobs[job]);
                item_pixel(item, oj.width, px, py);
                valid = px < oj.width && py < oj.height;
                if (valid) primary_path(oj, px, py, smp, p);
            } else if (PRIMARY) {
                int s, px, py;
                valid = slot_pixel(op, slot, npix, s, px, py);
                if (valid) primary_path(op, px, py, first_sample + s, p);
            } else {
                const int2 mt = *reinterpret_cast<const int2*>(&pb.meta[slot]);
                valid = !(mt.y & PATH_DEAD);  // ended path left in an uncompacted queue
                if (valid) {
                    const float4 o4 = pb.ro[slot], d4 = pb.rd[slot];
                    p.ro = xyz(o4); p.rd = xyz(d4);
                    p.depth = mt.x;
                }
            }
            if (valid) {
                if (CAPPED) {
                    done = closest_begin(sc, op, p.ro, p.rd, p.depth, stk, ovf, h, ts, budget);
                } else {
                    h = closest<COUNT>(sc, op, p.ro, p.rd, p.depth, stk, cnt, ovf);
                    done = true;
                }
            }
        }
        if (CAPPED) {
            bool susp = valid && !done && ts.sp <= CONT_STACK;
            const int r = wave_append(susp, cont_count);
            if (susp && r >= cb.cap) susp = false;  // no room: finish here
            if (!susp && valid && !done) {
                traverse<false, false, true>(sc, p.ro, p.rd, h.t, stk, h, cnt, ovf, false, nullptr, &ts, -1);
                done = true;
            }
            if (valid && done) keep = finish_extend<PRIMARY>(sc, op, pb, slot, p, h);
            const int pos = wave_append(keep || susp, hit_count);
            if (keep) hitq[pos] = slot;
            if (susp) save_cont(cb, r, slot, ts, h, stk, pos);
        } else {
            if (valid && done) keep = finish_extend<PRIMARY>(sc, op, pb, slot, p, h);
            const int pos = wave_append(keep, hit_count);
            if (pos >= 0) hitq[pos] = slot;
        }
    }
    flush_counts<COUNT>(cnt, counters);
    if (ovf) atomicOr(ovf_flag, 1);
}

// The traversals a capped extend stage stopped, packed densely: each record's
// stack back on this thread's stack, the traversal run to its end, then the
// same finish as the extend stage (records [0, min(*cont_count, cap))); the
// slot, or -1 for a miss, goes to the hit-queue position the record kept.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7, 7)))
k_extend_resume(DevScene sc, DevOpts op, PathBuf pb, ContBuf cb, const int* cont_count, int* hitq,
                int* work, int S, int* spill, int* ovf_flag) {
    extern __shared__ int lds[];
    Stack stk{lds, spill, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x), S,
              blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x,
              static_cast<size_t>(gridDim.x) * blockDim.x};
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    const int count = min(*cont_count, cb.cap);
    for (WorkCursor<RUN_TRAVERSE> wc; wc.base < count; wc.next(work)) {
        const int r = wc.base + lane_id();
        bool keep = false;
        int slot = -1;
        if (r < count) {
            slot = cb.slot[r];
            const int4 st = cb.st[r];
            TravState ts{st.x, st.y, st.z, st.w};
            const float4 hf = cb.hf[r];
            const int4 hi = cb.hi[r];
            Hit h;
            h.t = hf.x; h.u = hf.y; h.v = hf.z; h.tri = hi.x; h.inst = hi.y; h.light = hi.z;
            for (int k = 0; k < ts.sp; ++k) stk.push(k, cb.stack[static_cast<size_t>(k) * cb.cap + r]);
            PathRegs p;
            const float4 o4 = pb.ro[slot], d4 = pb.rd[slot];
            p.ro = xyz(o4); p.rd = xyz(d4);
            p.depth = pb.meta[slot].x;
            traverse<false, false, true>(sc, p.ro, p.rd, h.t, stk, h, cnt, ovf, false, nullptr, &ts, -1);
            keep = finish_extend<false>(sc, op, pb, slot, p, h);
            hitq[hi.w] = keep ? slot : -1;
        }
    }
    if (ovf) atomicOr(ovf_flag, 1);
}

// Shade: one bounce of every path that hit something. The NEE rays of the
// vertex go to the shadow queue as one record; continuing paths are
// compacted into the next extend queue; ending paths write only radiance.
// INLINE_NEE: the vertex's NEE visibility is resolved here instead, with
// DirectLight's own traversals in the reference's order (pathtrace.glsl:158-283).
// Scenes need it when AnyHit draws from the path's random numbers between
// NEE sampling and BSDF sampling (stochastic BLEND alpha, anyhit.glsl:118-141)
// or when visibility is a transmittance (volume MIS, pathtrace.glsl:118-156);
// no shadow records are written then. The LDS holds the traversal stack.
template <bool PRIMARY, bool BATCH, bool INLINE_NEE>
__global__ void __launch_bounds__(256) k_shade(DevScene sc, DevOpts op0, PathBuf pb, ShadowBuf sb,
                                               const int* __restrict__ hitq, const int* hit_count, int npix,
                                               int first_sample, int* qout, int* count_out, int* shadow_count,
                                               int* work, int compact, BatchBuf bb, int S, int* spill,
                                               unsigned long long* counters, int* ovf_flag) {
    extern __shared__ int lds[];
    Stack stk{lds, spill, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x), S,
              blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x,
              static_cast<size_t>(gridDim.x) * blockDim.x};
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    const int count = *hit_count;
    for (WorkCursor<RUN_SHADE> wc; wc.base < count; wc.next(work)) {
        const int i = wc.base + lane_id();
        int slot = -1;
        bool keep = false;
        NeeRay nee[2];
        nee[0].valid = nee[1].valid = false;
        vec3 thr_nee = V1(0.f);
        if (i < count) slot = hitq[i];  // -1: a place a resumed traversal left (a miss)
        if (slot >= 0) {
            PathRegs p;
            DevOpts op = op0;
            if (BATCH) {
                int job, smp, item;
                batch_decode(bb, slot, job, smp, item);
                op = job_opts(op0, bb.jobs[job]);
                if (PRIMARY) {
                    int px, py;
                    item_pixel(item, op.width, px, py);
                    primary_path(op, px, py, smp, p);
                }
            } else if (PRIMARY) {
                int s, px, py;
                slot_pixel(op, slot, npix, s, px, py);
                primary_path(op, px, py, first_sample + s, p);
            }
            if (!PRIMARY) load_path(pb, slot, p);
            const float4 h4 = pb.hit[slot];
            const int2 h2 = pb.hit2[slot];
            Hit h;
            h.t = h4.x; h.u = h4.y; h.v = h4.z; h.tri = __float_as_int(h4.w); h.inst = h2.x; h.light = h2.y;
            SurfState st;
            init_surf(st);
            Material m;
            float eta;
            bool alive = false;
            const int ev = bounce_pre(sc, op, h, p, st, m, eta);
            if (ev != EV_END) {
                if (ev != EV_PASS) {
                    if (INLINE_NEE) {
                        vec3 Ld;
                        if ((op.flags & OPT_MEDIUM) && (op.flags & OPT_VOL_MIS))
                            Ld = direct_light_volmis<false>(sc, op, p.rd, st, m, eta, p.med, ev == EV_SURFACE, p.rng,
                                                            stk, cnt, ovf);
                        else
                            Ld = direct_light<false>(sc, op, p.rd, st, m, eta, p.rng, stk, cnt, ovf);
                        p.rad = p.rad + Ld * p.thr;
                    } else {
                        nee_rays(sc, op, p, p.rng, st, m, eta, nee);
                        thr_nee = p.thr;
                    }
                }
                alive = bounce_post(op, p, st, m, eta, ev);
                if (alive) p.depth++;
            }
            if (alive) {
                store_path(pb, slot, p);
            } else {
                pb.rad[slot] = F4(p.rad, p.alpha);
                // without compaction the path stays in the queue: mark it ended
                if (!compact) pb.meta[slot].y = PATH_DEAD;
            }
            keep = alive;
        }
        if (!INLINE_NEE) {
            // NEE rays of this vertex (environment, light): one shadow record
            // per path holding both, so one thread adds their visible
            // contributions in DirectLight's order (no two threads update a
            // path's radiance).
            const int mask = (nee[0].valid ? 1 : 0) | (nee[1].valid ? 2 : 0);
            const int pos = wave_append(mask != 0, shadow_count);
            if (pos >= 0) {
                sb.idx[pos] = make_int2(slot, mask);
                sb.o[pos] = F4(nee[(mask & 1) ? 0 : 1].o, 0.f);  // both rays start at the scatter point
                sb.t[pos] = F4(thr_nee, 0.f);
                for (int k = 0; k < 2; ++k) {
                    if (!((mask >> k) & 1)) continue;
                    sb.d[2 * pos + k] = F4(nee[k].d, nee[k].maxd);
                    sb.c[2 * pos + k] = F4(nee[k].contrib, 0.f);
                }
            }
        }
        if (compact) {
            const int pos = wave_append(keep, count_out);
            if (pos >= 0) qout[pos] = slot;
        }
    }
    if (INLINE_NEE) {
        flush_counts<false>(cnt, counters);
        if (ovf) atomicOr(ovf_flag, 1);
    }
}

// Shadow: any-hit tests of the queued NEE rays of a path; the visible
// contributions are summed as DirectLight does and added to the path's
// radiance times the throughput at NEE time. `vis` (the adjoint's primal
// pass, else null): the word array of this bounce's visibility bits, where
// bits vshift + k are set for the visible rays k of the record's path slot.
// Compiled for 8 waves per SIMD (64 VGPRs, a few spilled): measured 9% faster
// than the unconstrained 5 waves on the 1080p workload.
template <bool COUNT>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(8, 8))) k_shadow(DevScene sc, DevOpts op, PathBuf pb, ShadowBuf sb, const int* count_in,
                                                int* work, int S, int* spill, unsigned long long* counters,
                                                int* ovf_flag, unsigned* vis, int vshift) {
    extern __shared__ int lds[];
    Stack stk{lds, spill, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x), S,
              blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x,
              static_cast<size_t>(gridDim.x) * blockDim.x};
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    const int count = *count_in;
    for (WorkCursor<RUN_TRAVERSE> wc; wc.base < count; wc.next(work)) {
        const int i = wc.base + lane_id();
        if (i < count) {
            const int2 id = sb.idx[i];
            const float4 o4 = sb.o[i];
            vec3 Ld = V1(0.0f);
            int lit = 0;
            for (int k = 0; k < 2; ++k) {
                if (!((id.y >> k) & 1)) continue;
                const float4 d4 = sb.d[2 * i + k];
                if (occluded<COUNT, true>(sc, op, xyz(o4), xyz(d4), d4.w, stk, cnt, ovf)) continue;
                Ld = Ld + xyz(sb.c[2 * i + k]);
                lit |= 1 << k;
            }
            if (lit) {
                const float4 r = pb.rad[id.x];
                pb.rad[id.x] = F4(xyz(r) + Ld * xyz(sb.t[i]), r.w);
                if (vis) vis[id.x] |= static_cast<unsigned>(lit) << vshift;
            }
        }
    }
    flush_counts<COUNT>(cnt, counters);
    if (ovf) atomicOr(ovf_flag, 1);
}

// Replay stage of the path-replay adjoint: one bounce of every path that hit
// something, with the NEE visibility of the primal pass (ReplayBuf.vis bits
// at vshift), so the path advances exactly as in the primal (pt_replay.h);
// each SURFACE vertex adds its material-parameter gradient. Per-block sums in
// LDS (num_mats * NPARAM floats when ReplayBuf.lds), flushed with one global
// atomic per non-zero entry; otherwise global atomics. Material ids outside
// [0, num_mats) set ReplayBuf.err; vertices with a non-finite gradient term
// are dropped and counted in ReplayBuf.dropped.
template <bool PRIMARY>
__global__ void __launch_bounds__(256) k_replay_shade(DevScene sc, DevOpts op, PathBuf pb,
                                                      const int* __restrict__ hitq, const int* hit_count, int npix,
                                                      int first_sample, int* qout, int* count_out, int* work,
                                                      int compact, ReplayBuf rb, int vshift) {
    extern __shared__ float gacc[];
    const int ng0 = rb.num_mats * NPARAM;
    const int ng = ng0 + (rb.mgrad ? rb.num_mats * NMEDIUM : 0);
    if (rb.lds) {
        for (int j = threadIdx.x; j < ng; j += blockDim.x) gacc[j] = 0.0f;
        __syncthreads();
    }
    const int count = *hit_count;
    for (WorkCursor<RUN_SHADE> wc; wc.base < count; wc.next(work)) {
        const int i = wc.base + lane_id();
        int slot = -1;
        bool keep = false;
        if (i < count) {
            slot = hitq[i];
            int s, px, py;
            slot_pixel(op, slot, npix, s, px, py);
            PathRegs p;
            if (PRIMARY) primary_path(op, px, py, first_sample + s, p);
            else load_path(pb, slot, p);
            const float4 h4 = pb.hit[slot];
            const int2 h2 = pb.hit2[slot];
            Hit h;
            h.t = h4.x; h.u = h4.y; h.v = h4.z; h.tri = __float_as_int(h4.w); h.inst = h2.x; h.light = h2.y;
            const float* dp = rb.dimg + 3 * (static_cast<size_t>(py) * op.width + px);
            const vec3 delta = V(dp[0], dp[1], dp[2]) * rb.inv_n;
            const int vis = static_cast<int>((rb.vis[slot] >> vshift) & 3u);
            float g[NPARAM], mg[NMEDIUM];
            const ReplayVertex rv = replay_vertex(sc, op, h, p, vis, delta, xyz(rb.ltot[slot]), g,
                                                  rb.mgrad ? mg : nullptr);
            if (rv.med_mat >= 0) {
                if (rv.med_mat >= rb.num_mats) {
                    atomicOr(rb.err, 1);
                } else if (!rv.mfinite) {
                    atomicAdd(rb.dropped, 1);
                } else {
                    const int g0 = rv.med_mat * NMEDIUM;
#pragma unroll
                    for (int k = 0; k < NMEDIUM; ++k) {
                        if (mg[k] == 0.0f) continue;
                        if (rb.lds)
                            __hip_atomic_fetch_add(&gacc[ng0 + g0 + k], mg[k], __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_WORKGROUP);
                        else atomicAdd(rb.mgrad + g0 + k, mg[k]);
                    }
                }
            }
            if (rv.mat >= 0) {
                if (rv.mat >= rb.num_mats) {
                    atomicOr(rb.err, 1);
                } else if (!rv.finite) {
                    atomicAdd(rb.dropped, 1);
                } else {
                    const int g0 = rv.mat * NPARAM;
#pragma unroll
                    for (int k = 0; k < NPARAM; ++k) {
                        if (g[k] == 0.0f) continue;
                        if (rb.lds) __hip_atomic_fetch_add(&gacc[g0 + k], g[k], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                        else atomicAdd(rb.grad + g0 + k, g[k]);
                    }
                }
            }
            if (rv.alive) {
                store_path(pb, slot, p);
            } else {
                pb.rad[slot] = F4(p.rad, p.alpha);
                if (!compact) pb.meta[slot].y = PATH_DEAD;
            }
            keep = rv.alive;
        }
        if (compact) {
            const int pos = wave_append(keep, count_out);
            if (pos >= 0) qout[pos] = slot;
        }
    }
    if (rb.lds) {
        __syncthreads();
        for (int j = threadIdx.x; j < ng; j += blockDim.x)
            if (gacc[j] != 0.0f) atomicAdd(j < ng0 ? rb.grad + j : rb.mgrad + (j - ng0), gacc[j]);
    }
}

// Batched render in one persistent launch of the plain form. A work item is
// one pixel of one job (8x8 tiles, so a wave's batch of 64 lies in one job);
// its thread traces the pixel's samples 0..spp-1 in order and stores their
// sum, the float sequence pt_render adds into its accumulation buffer, so
// every job's image equals the job rendered alone. Paths stay in registers
// from camera ray to termination: no path state is written between bounces.
// Two waves per SIMD (the occupancy of k_plain): left to itself the allocator
// takes 4 registers more than that allows.
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2)))
k_plain_batch(DevScene sc, DevOpts op0, const JobDev* jobs, int njobs, int items, float4* __restrict__ accum,
              int* work, int S, int* spill, int* ovf_flag) {
    extern __shared__ int lds[];
    Stack stk{lds, spill, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x), S,
              blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x,
              static_cast<size_t>(gridDim.x) * blockDim.x};
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    for (WorkCursor<RUN_TRAVERSE> wc; wc.base < items; wc.next(work)) {
        int lo = 0, hi = njobs - 1;  // the last job whose first item is <= the batch's
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (jobs[mid].item0 <= wc.base) lo = mid;
            else hi = mid - 1;
        }
        const JobDev* Jp = jobs + lo;
        const int width = Jp->width, height = Jp->height, spp = Jp->spp;
        const int item = wc.base - Jp->item0 + lane_id();
        int px, py;
        item_pixel(item, width, px, py);
        if (item < Jp->npix && px < width && py < height) {
            float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
            for (int smp = 0; smp < spp; ++smp) {
                // The camera is read again for each sample, after the previous
                // trace (jobs is not restrict, so the trace's stores keep the
                // loads here): only max_depth stays live through the trace.
                PathRegs p;
                primary_path(job_opts(op0, *Jp), px, py, smp, p);
                DevOpts op = op0;
                op.max_depth = Jp->max_depth;
                const float4 c = path_trace<false>(sc, op, p, stk, cnt, ovf);
                acc.x += c.x; acc.y += c.y; acc.z += c.z; acc.w += c.w;
            }
            accum[Jp->accum_off + static_cast<long long>(py) * width + px] = acc;
        }
    }
    if (ovf) atomicOr(ovf_flag, 1);
}

// Accumulate: accum[pixel] += radiance of samples s = 0..nsamp-1 in order.
__global__ void __launch_bounds__(256) k_accum(DevOpts op, PathBuf pb, float4* __restrict__ accum, int npix, int nsamp) {
    const int item = blockIdx.x * blockDim.x + threadIdx.x;
    if (item >= npix) return;
    int px, py;
    item_pixel(item, op.width, px, py);
    if (px >= op.width || py >= op.height) return;
    float4* a = accum + static_cast<size_t>(py) * op.width + px;
    float4 acc = *a;
    for (int s = 0; s < nsamp; ++s) {
        const float4 r = pb.rad[static_cast<size_t>(s) * npix + item];
        acc.x += r.x; acc.y += r.y; acc.z += r.z; acc.w += r.w;
    }
    *a = acc;
}

// ---- batched render: pass bookkeeping and accumulation
// First node of a batched pass: make pass cur[0] current (copy its PassDev to
// cur[1..8)) and zero the pass's queue and work counters.
__global__ void __launch_bounds__(256) k_pass_begin(BatchBuf bb, int* counters, int ncounters) {
    for (int i = threadIdx.x; i < ncounters; i += blockDim.x) counters[i] = 0;
    if (threadIdx.x < 8) {
        const int* src = reinterpret_cast<const int*>(bb.passes + bb.cur[0]);
        bb.cur[1 + threadIdx.x] = threadIdx.x < 7 ? src[threadIdx.x] : 0;
    }
}

// Last node of a batched pass: advance to the next pass.
__global__ void k_pass_end(BatchBuf bb) {
    if (threadIdx.x == 0) bb.cur[0] = bb.cur[0] + 1;
}

// Accumulation of a batched pass: one thread per (job, pixel item) of the
// pass; each adds its pixel's samples of this pass in sample order, so every
// job's image is bit-identical to rendering that job alone.
__global__ void __launch_bounds__(256) k_accum_batch(PathBuf pb, BatchBuf bb) {
    const int seg0 = bb.cur[4], nsegs = bb.cur[5], nthreads = bb.cur[6];
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nthreads; t += gridDim.x * blockDim.x) {
        int k = seg0;
        while (k + 1 < seg0 + nsegs && bb.segs[k + 1].thread0 <= t) ++k;
        const SegDev sg = bb.segs[k];
        const JobDev J = bb.jobs[sg.job];
        const int item = t - sg.thread0;
        int px, py;
        item_pixel(item, J.width, px, py);
        if (px >= J.width || py >= J.height) continue;
        float4* a = bb.accum + J.accum_off + static_cast<size_t>(py) * J.width + px;
        float4 acc = *a;
        for (int q = 0; q < sg.nplanes; ++q) {
            const float4 r = pb.rad[bb.planes[sg.plane0 + q].slot0 + item];
            acc.x += r.x; acc.y += r.y; acc.z += r.z; acc.w += r.w;
        }
        *a = acc;
    }
}

// Environment-map CDF (EnvironmentMap::BuildCDF, src/core/EnvironmentMap.cpp:39-61):
// the luminance of every texel, then one running sum over the row-major image
// in the same order and float arithmetic as the host build.
//
// The luminance is elementwise and runs over the whole grid, written into the
// CDF buffer. The running sum is a chain of dependent fp32 adds, which a
// different association would round differently, so one lane performs it: one
// workgroup, lane 0 of wave 0 adding chunks of ENV_CHUNK texels held in LDS
// (16 per step, the next 16 already loaded), while waves 1..3 store the chunk
// before and load the chunk after. Adding texel 0 to -0.0f returns it exactly.
constexpr int ENV_CHUNK = 4096;
constexpr int ENV_SCAN_THREADS = 256;

// Luminance of every texel; clears *ok when one is negative, -0 or not finite
// (the binade scan below needs finite values >= +0).
__global__ void k_env_lum(const float* __restrict__ img, int n, float* __restrict__ w, int* __restrict__ ok) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float y = 0.212671f * img[3 * static_cast<size_t>(i)] + 0.715160f * img[3 * static_cast<size_t>(i) + 1] +
                    0.072169f * img[3 * static_cast<size_t>(i) + 2];
    w[i] = y;
    if (!(y >= 0.0f) || (__float_as_uint(y) >> 31) || !(y <= 3.40282347e38f)) *ok = 0;
}

// Exact parallel form of the same running sum, for finite luminances >= +0.
// Between two crossings of a power of two by the running sum r, every add
// rounds on one grid of spacing u (the ulp of r), so with r = R*u (R < 2^24)
// each add increases R by an integer: floor(w/u), plus one when the fraction
// of w/u is above one half, or, at exactly one half, when R + floor(w/u) is
// odd (ties to even). An add is a function of the parity of R only, so a scan
// of (increment, parity) pairs over a chunk gives every R at once. The first
// add whose exact sum leaves the grid (R + floor(w/u) >= 2^24) is done as a
// float add, and the scan restarts on the new grid from the element after it;
// an add that rounds up onto 2^24 ends the grid too. The values equal the
// one-lane chain bit for bit.
constexpr int ZS_THREADS = 1024, ZS_PER = 16, ZS_CHUNK = ZS_THREADS * ZS_PER;
constexpr int ZS_TOP = 1 << 24, ZS_SAT = 1 << 26;
constexpr int ZL_POS = 0, ZL_K = 1, ZL_R = 2, ZL_EXIT = 3, ZL_INF = 4, ZL_CUR = 5, ZL_TOT = 8, ZL_PRE = 8 + 4 * 16;
constexpr int ZL_WORDS = ZL_PRE + 4 * 17;

struct ZFun {  // increments of R and output parities for input parity 0 and 1
    int a, b, pa, pb;
};
__device__ __forceinline__ ZFun zf_identity() { return ZFun{0, 0, 0, 1}; }
__device__ __forceinline__ ZFun zf_then(ZFun f, ZFun g) {  // f, then g
    ZFun h;
    h.a = min(f.a + (f.pa ? g.b : g.a), ZS_SAT);
    h.pa = f.pa ? g.pb : g.pa;
    h.b = min(f.b + (f.pb ? g.b : g.a), ZS_SAT);
    h.pb = f.pb ? g.pb : g.pa;
    return h;
}
__device__ __forceinline__ int zf_inc(int F, int cls, int parity) {
    return F + (cls == 2 ? 1 : (cls == 1 ? ((parity + F) & 1) : 0));
}
__device__ __forceinline__ ZFun zf_elem(int F, int cls) {
    const int a = zf_inc(F, cls, 0), b = zf_inc(F, cls, 1);
    return ZFun{a, b, a & 1, (1 + b) & 1};
}
// Grid of a value r >= +0: spacing 2^k and r = R * 2^k.
__device__ __forceinline__ void zone_of(float r, int& k, int& R) {
    const unsigned bits = __float_as_uint(r);
    const int e = static_cast<int>((bits >> 23) & 0xffu);
    k = e == 0 ? -149 : e - 150;
    R = static_cast<int>(e == 0 ? (bits & 0x7fffffu) : ((bits & 0x7fffffu) | 0x800000u));
}

// Texels [base, base + ZS_PER) below `end` on the grid 2^k: the input values,
// their integer parts F and fraction classes, and their composed function.
__device__ __forceinline__ ZFun zs_load(const float* cdf, int base, int end, int k, float (&w)[ZS_PER],
                                        int (&F)[ZS_PER], int (&cls)[ZS_PER]) {
    ZFun L = zf_identity();
#pragma unroll
    for (int j = 0; j < ZS_PER; ++j) {
        w[j] = 0.0f;
        F[j] = 0;
        cls[j] = 0;
        if (base + j < end) {
            w[j] = cdf[base + j];
            const float q = ldexpf(w[j], -k);
            if (!(q < 16777216.0f)) {
                F[j] = ZS_TOP;
            } else {
                const float fq = floorf(q);
                const float d = q - fq;
                F[j] = static_cast<int>(fq);
                cls[j] = d < 0.5f ? 0 : (d == 0.5f ? 1 : 2);
            }
            L = zf_then(L, zf_elem(F[j], cls[j]));
        }
    }
    return L;
}

// Block scan of the threads' functions (every thread calls it): returns the
// composition of all earlier threads' functions; `tot` gets the whole block's.
__device__ __forceinline__ ZFun zs_block_scan(ZFun L, int* lds, ZFun& tot) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int nwaves = ZS_THREADS / 64;
    ZFun I = L;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        ZFun o;
        o.a = __shfl_up(I.a, d);
        o.b = __shfl_up(I.b, d);
        o.pa = __shfl_up(I.pa, d);
        o.pb = __shfl_up(I.pb, d);
        if (lane >= d) I = zf_then(o, I);
    }
    if (lane == 63) {
        lds[ZL_TOT + 4 * wave] = I.a;
        lds[ZL_TOT + 4 * wave + 1] = I.b;
        lds[ZL_TOT + 4 * wave + 2] = I.pa;
        lds[ZL_TOT + 4 * wave + 3] = I.pb;
    }
    __syncthreads();
    if (wave == 0) {  // exclusive prefix of the wave totals, one lane per wave
        ZFun t = zf_identity();
        if (lane < nwaves)
            t = ZFun{lds[ZL_TOT + 4 * lane], lds[ZL_TOT + 4 * lane + 1], lds[ZL_TOT + 4 * lane + 2],
                     lds[ZL_TOT + 4 * lane + 3]};
#pragma unroll
        for (int d = 1; d < nwaves; d <<= 1) {
            ZFun o;
            o.a = __shfl_up(t.a, d);
            o.b = __shfl_up(t.b, d);
            o.pa = __shfl_up(t.pa, d);
            o.pb = __shfl_up(t.pb, d);
            if (lane >= d) t = zf_then(o, t);
        }
        ZFun e;
        e.a = __shfl_up(t.a, 1);
        e.b = __shfl_up(t.b, 1);
        e.pa = __shfl_up(t.pa, 1);
        e.pb = __shfl_up(t.pb, 1);
        if (lane == 0) e = zf_identity();
        if (lane < nwaves) {
            lds[ZL_PRE + 4 * lane] = e.a;
            lds[ZL_PRE + 4 * lane + 1] = e.b;
            lds[ZL_PRE + 4 * lane + 2] = e.pa;
            lds[ZL_PRE + 4 * lane + 3] = e.pb;
        }
        if (lane == nwaves - 1) {
            lds[ZL_PRE + 4 * nwaves] = t.a;
            lds[ZL_PRE + 4 * nwaves + 1] = t.b;
            lds[ZL_PRE + 4 * nwaves + 2] = t.pa;
            lds[ZL_PRE + 4 * nwaves + 3] = t.pb;
        }
    }
    __syncthreads();
    ZFun X;
    X.a = __shfl_up(I.a, 1);
    X.b = __shfl_up(I.b, 1);
    X.pa = __shfl_up(I.pa, 1);
    X.pb = __shfl_up(I.pb, 1);
    if (lane == 0) X = zf_identity();
    X = zf_then(ZFun{lds[ZL_PRE + 4 * wave], lds[ZL_PRE + 4 * wave + 1], lds[ZL_PRE + 4 * wave + 2],
                     lds[ZL_PRE + 4 * wave + 3]}, X);
    tot = ZFun{lds[ZL_PRE + 4 * nwaves], lds[ZL_PRE + 4 * nwaves + 1], lds[ZL_PRE + 4 * nwaves + 2],
               lds[ZL_PRE + 4 * nwaves + 3]};
    return X;
}

// The exact running sum of texels [pos, end) from the state in LDS (ZL_POS,
// ZL_K, ZL_R, ZL_INF), leaving the state after texel end - 1 (every thread of
// the block calls it).
__device__ void zone_walk(float* cdf, int end, int* lds) {
    const int tid = threadIdx.x;
    for (;;) {
        const int pos = lds[ZL_POS];
        if (pos >= end) break;
        if (lds[ZL_INF]) {
            for (int i = pos + tid; i < end; i += ZS_THREADS) cdf[i] = __uint_as_float(0x7f800000u);
            __syncthreads();
            if (tid == 0) lds[ZL_POS] = end;
            __syncthreads();
            break;
        }
        const int k = lds[ZL_K], R0 = lds[ZL_R];
        const int base = pos + tid * ZS_PER;
        float w[ZS_PER];
        int F[ZS_PER], cls[ZS_PER];
        const ZFun L = zs_load(cdf, base, end, k, w, F, cls);
        if (tid == 0) lds[ZL_EXIT] = 0x7fffffff;
        ZFun tot;
        const ZFun X = zs_block_scan(L, lds, tot);
        const int Rstart = min(R0 + ((R0 & 1) ? X.b : X.a), ZS_SAT);
        // exact walk over this thread's elements
        int Rv[ZS_PER];
        int R = Rstart, ex = -1, exA = 0;
#pragma unroll
        for (int j = 0; j < ZS_PER; ++j) {
            Rv[j] = R;
            if (ex < 0 && base + j < end) {
                if (R + F[j] >= ZS_TOP) {
                    ex = j;
                    exA = 1;
                } else {
                    R += zf_inc(F[j], cls[j], R & 1);
                    Rv[j] = R;
                    if (R >= ZS_TOP) ex = j;
                }
            }
        }
        if (ex >= 0) atomicMin(&lds[ZL_EXIT], base + ex);
        __syncthreads();
        const int estar = lds[ZL_EXIT];
#pragma unroll
        for (int j = 0; j < ZS_PER; ++j)
            if (base + j < end && base + j < estar) cdf[base + j] = ldexpf(static_cast<float>(Rv[j]), k);
        if (estar != 0x7fffffff) {
            if (estar >= base && estar < base + ZS_PER) {
                float rv = 0.0f;
#pragma unroll
                for (int j = 0; j < ZS_PER; ++j) {
                    if (base + j != estar) continue;
                    if (exA) {
                        const int Rp = j == 0 ? Rstart : Rv[j > 0 ? j - 1 : 0];
                        rv = ldexpf(static_cast<float>(Rp), k) + w[j];
                    } else {
                        rv = ldexpf(static_cast<float>(Rv[j]), k);
                    }
                }
                cdf[estar] = rv;
                int nk, nR;
                zone_of(rv, nk, nR);
                lds[ZL_POS] = estar + 1;
                lds[ZL_K] = nk;
                lds[ZL_R] = nR;
                lds[ZL_INF] = (__float_as_uint(rv) & 0x7fffffffu) == 0x7f800000u;
            }
        } else {
            const int last = min(pos + ZS_CHUNK, end) - 1;
            if (last >= base && last < base + ZS_PER) {
#pragma unroll
                for (int j = 0; j < ZS_PER; ++j)
                    if (base + j == last) lds[ZL_R] = Rv[j];
                lds[ZL_POS] = last + 1;
            }
        }
        __syncthreads();
    }
}

// The same sum spread over the grid. Texels 1..n-1 form chunks of ZS_CHUNK.
// (1) Each chunk's sum in double gives an estimate of the running sum before
// it, hence a guess of its grid. (2) Each chunk's function on its guessed grid.
// (3) One workgroup carries the exact state from chunk to chunk: a chunk whose
// grid was guessed right and that stays on it is applied as a whole; any other
// chunk is summed exactly by zone_walk. (4) The chunks applied as a whole write
// their values from the state carried into them.
// Scratch per chunk: double sum, guessed grid, function (4 ints), carried R and
// a flag (the chunk was applied as a whole).
__global__ void __launch_bounds__(256) k_env_chunk_sum(const float* __restrict__ cdf, int n, double* __restrict__ csum,
                                                       const int* __restrict__ ok) {
    if (*ok == 0) return;
    extern __shared__ int lds[];
    double* part = reinterpret_cast<double*>(lds);
    const int c = blockIdx.x;
    const int lo = 1 + c * ZS_CHUNK, hi = min(lo + ZS_CHUNK, n);
    double acc = 0.0;
    for (int i = lo + static_cast<int>(threadIdx.x); i < hi; i += 256) acc += static_cast<double>(cdf[i]);
    part[threadIdx.x] = acc;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (static_cast<int>(threadIdx.x) < h) part[threadIdx.x] += part[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) csum[c] = part[0];
}

__global__ void k_env_guess(const float* __restrict__ cdf, int nchunks, const double* __restrict__ csum,
                            int* __restrict__ kguess, const int* __restrict__ ok) {
    if (*ok == 0 || threadIdx.x != 0) return;
    double run = static_cast<double>(cdf[0]);
    for (int c = 0; c < nchunks; ++c) {
        int k, R;
        zone_of(static_cast<float>(run), k, R);
        kguess[c] = k;
        run += csum[c];
    }
}

__global__ void __launch_bounds__(ZS_THREADS) k_env_chunk_fun(const float* __restrict__ cdf, int n,
                                                              const int* __restrict__ kguess, int4* __restrict__ fun,
                                                              const int* __restrict__ ok) {
    if (*ok == 0) return;
    extern __shared__ int lds[];
    const int c = blockIdx.x;
    const int lo = 1 + c * ZS_CHUNK, hi = min(lo + ZS_CHUNK, n);
    float w[ZS_PER];
    int F[ZS_PER], cls[ZS_PER];
    const ZFun L = zs_load(cdf, lo + static_cast<int>(threadIdx.x) * ZS_PER, hi, kguess[c], w, F, cls);
    ZFun tot;
    (void)zs_block_scan(L, lds, tot);
    if (threadIdx.x == 0) fun[c] = make_int4(tot.a, tot.b, tot.pa, tot.pb);
}

__global__ void __launch_bounds__(ZS_THREADS) k_env_carry(float* cdf, int n, int nchunks, const int* __restrict__ kguess,
                                                          const int4* __restrict__ fun, int2* __restrict__ cstate,
                                                          float* __restrict__ total, const int* __restrict__ ok) {
    if (*ok == 0) return;
    extern __shared__ int lds[];
    const int tid = threadIdx.x;
    if (tid == 0) {
        const float r0 = n > 0 ? cdf[0] : 0.0f;  // cdf[0] = w[0]
        int k, R;
        zone_of(r0, k, R);
        lds[ZL_POS] = 1;
        lds[ZL_K] = k;
        lds[ZL_R] = R;
        lds[ZL_INF] = (__float_as_uint(r0) & 0x7fffffffu) == 0x7f800000u;
        lds[ZL_CUR] = 0;
    }
    __syncthreads();
    for (;;) {
        if (tid == 0) {  // chunks that stay on their guessed grid: O(1) each
            int c = lds[ZL_CUR], R = lds[ZL_R], pos = lds[ZL_POS];
            const int k = lds[ZL_K];
            while (c < nchunks && !lds[ZL_INF]) {
                const int lo = 1 + c * ZS_CHUNK;
                if (pos != lo || k != kguess[c]) break;
                const int4 f = fun[c];
                const int inc = (R & 1) ? f.y : f.x;
                if (R + inc >= ZS_TOP) break;
                cstate[c] = make_int2(R, 1);
                R += inc;
                pos = min(lo + ZS_CHUNK, n);
                ++c;
            }
            lds[ZL_CUR] = c;
            lds[ZL_R] = R;
            lds[ZL_POS] = pos;
        }
        __syncthreads();
        const int c = lds[ZL_CUR];
        if (c >= nchunks) break;
        // the next chunk leaves its grid, or the grid was guessed wrong: exactly
        if (tid == 0) cstate[c] = make_int2(0, 0);
        zone_walk(cdf, min(1 + (c + 1) * ZS_CHUNK, n), lds);
        if (tid == 0) lds[ZL_CUR] = c + 1;
        __syncthreads();
    }
    if (tid == 0)
        *total = n <= 0 ? 0.0f : (lds[ZL_INF] ? __uint_as_float(0x7f800000u) : ldexpf(static_cast<float>(lds[ZL_R]), lds[ZL_K]));
}

__global__ void __launch_bounds__(ZS_THREADS) k_env_chunk_write(float* cdf, int n, const int* __restrict__ kguess,
                                                                const int2* __restrict__ cstate,
                                                                const int* __restrict__ ok) {
    if (*ok == 0) return;
    const int c = blockIdx.x;
    const int2 cs = cstate[c];
    if (cs.y == 0) return;  // summed exactly by k_env_carry
    extern __shared__ int lds[];
    const int k = kguess[c];
    const int lo = 1 + c * ZS_CHUNK, hi = min(lo + ZS_CHUNK, n);
    const int base = lo + static_cast<int>(threadIdx.x) * ZS_PER;
    float w[ZS_PER];
    int F[ZS_PER], cls[ZS_PER];
    const ZFun L = zs_load(cdf, base, hi, k, w, F, cls);
    ZFun tot;
    const ZFun X = zs_block_scan(L, lds, tot);
    int R = cs.x + ((cs.x & 1) ? X.b : X.a);
    float out[ZS_PER];
#pragma unroll
    for (int j = 0; j < ZS_PER; ++j) {
        R += zf_inc(F[j], cls[j], R & 1);
        out[j] = ldexpf(static_cast<float>(R), k);
    }
    __syncthreads();  // every thread has read its inputs before any writes
#pragma unroll
    for (int j = 0; j < ZS_PER; ++j)
        if (base + j < hi) cdf[base + j] = out[j];
}

__global__ void __launch_bounds__(ENV_SCAN_THREADS) k_env_scan(float* cdf, int n, float* __restrict__ total,
                                                               const int* __restrict__ ok) {
    if (*ok != 0) return;
    extern __shared__ int lds[];
    float* buf = reinterpret_cast<float*>(lds);  // two chunks
    const int tid = threadIdx.x;
    const int nch = (n + ENV_CHUNK - 1) / ENV_CHUNK;
    for (int j = tid; j < min(ENV_CHUNK, n); j += ENV_SCAN_THREADS) buf[j] = cdf[j];
    __syncthreads();
    float run = -0.0f;
    for (int c = 0; c < nch; ++c) {
        float* cur = buf + (c & 1) * ENV_CHUNK;
        float* oth = buf + ((c + 1) & 1) * ENV_CHUNK;
        const int cnt = min(ENV_CHUNK, n - c * ENV_CHUNK);
        if (tid == 0) {
            int j = 0;
            if (cnt >= 16) {
                float4 q0 = reinterpret_cast<const float4*>(cur)[0], q1 = reinterpret_cast<const float4*>(cur)[1];
                float4 q2 = reinterpret_cast<const float4*>(cur)[2], q3 = reinterpret_cast<const float4*>(cur)[3];
                for (; j + 16 <= cnt; j += 16) {
                    float4 v[4] = {q0, q1, q2, q3};
                    if (j + 32 <= cnt) {
                        const float4* nx = reinterpret_cast<const float4*>(cur + j + 16);
                        q0 = nx[0]; q1 = nx[1]; q2 = nx[2]; q3 = nx[3];
                    }
                    for (int k = 0; k < 4; ++k) {
                        run = run + v[k].x; v[k].x = run;
                        run = run + v[k].y; v[k].y = run;
                        run = run + v[k].z; v[k].z = run;
                        run = run + v[k].w; v[k].w = run;
                    }
                    float4* o = reinterpret_cast<float4*>(cur + j);
                    o[0] = v[0]; o[1] = v[1]; o[2] = v[2]; o[3] = v[3];
                }
            }
            for (; j < cnt; ++j) {
                run = run + cur[j];
                cur[j] = run;
            }
        } else if (tid >= 64) {
            // Whole float4s, every load of a thread issued before its LDS writes.
            constexpr int LT = ENV_SCAN_THREADS - 64, Q = ENV_CHUNK / 4, PER = (Q + LT - 1) / LT;
            const int t = tid - 64;
            if (c > 0) {
                float4* dst = reinterpret_cast<float4*>(cdf + static_cast<size_t>(c - 1) * ENV_CHUNK);
                const float4* src = reinterpret_cast<const float4*>(oth);
#pragma unroll
                for (int k = 0; k < PER; ++k)
                    if (t + k * LT < Q) dst[t + k * LT] = src[t + k * LT];
            }
            if (c + 1 < nch) {
                const size_t nb = static_cast<size_t>(c + 1) * ENV_CHUNK;
                const int ncnt = min(ENV_CHUNK, n - (c + 1) * ENV_CHUNK), nq = ncnt / 4;
                const float4* src = reinterpret_cast<const float4*>(cdf + nb);
                float4 r[PER];
#pragma unroll
                for (int k = 0; k < PER; ++k)
                    if (t + k * LT < nq) r[k] = src[t + k * LT];
                float4* dst = reinterpret_cast<float4*>(oth);
#pragma unroll
                for (int k = 0; k < PER; ++k)
                    if (t + k * LT < nq) dst[t + k * LT] = r[k];
                // The last partial float4 goes to the thread that stored that
                // float4 of the previous chunk above, after its store.
                if (t == nq % LT)
                    for (int j = 4 * nq; j < ncnt; ++j) oth[j] = cdf[nb + j];
            }
        }
        __syncthreads();
    }
    if (nch > 0) {
        const int last = nch - 1;
        const float* cur = buf + (last & 1) * ENV_CHUNK;
        for (int j = tid; j < n - last * ENV_CHUNK; j += ENV_SCAN_THREADS) cdf[last * ENV_CHUNK + j] = cur[j];
    }
    if (tid == 0) *total = n > 0 ? run : 0.0f;
}

__global__ void k_fill_int(int* p, int n, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// ------------------------------------------------------------------ queries
__global__ void __launch_bounds__(256) k_query_closest(DevScene sc, DevOpts op, const float* __restrict__ orig,
                                                       const float* __restrict__ dir, int n, float* t, int* inst,
                                                       int* tri, int* light, float* uv, int S, int* spill,
                                                       int* ovf_flag) {
    extern __shared__ int lds[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    Stack stk{lds, spill, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x), S,
              static_cast<size_t>(i), static_cast<size_t>(gridDim.x) * blockDim.x};
    if (i >= n) return;
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    vec3 o = V(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]);
    vec3 d = V(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    Hit h = closest<false>(sc, op, o, d, 1, stk, cnt, ovf);
    t[i] = h.t;
    inst[i] = h.tri >= 0 ? h.inst : -1;
    tri[i] = h.tri;
    light[i] = h.light;
    uv[2 * i] = h.tri >= 0 ? h.u : 0.f;
    uv[2 * i + 1] = h.tri >= 0 ? h.v : 0.f;
    if (ovf) atomicOr(ovf_flag, 1);
}

__global__ void __launch_bounds__(256) k_query_any(DevScene sc, DevOpts op, const float* __restrict__ orig,
                                                   const float* __restrict__ dir, const float* __restrict__ maxd,
                                                   int n, int* occ, int S, int* spill, int* ovf_flag) {
    extern __shared__ int lds[];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    Stack stk{lds, spill, static_cast<int>(threadIdx.x), static_cast<int>(blockDim.x), S,
              static_cast<size_t>(i), static_cast<size_t>(gridDim.x) * blockDim.x};
    if (i >= n) return;
    TravCount cnt{0, 0, 0, 0, 0, 0};
    bool ovf = false;
    vec3 o = V(orig[3 * i], orig[3 * i + 1], orig[3 * i + 2]);
    vec3 d = V(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]);
    occ[i] = occluded<false, true>(sc, op, o, d, maxd[i], stk, cnt, ovf) ? 1 : 0;
    if (ovf) atomicOr(ovf_flag, 1);
}

// ------------------------------------------------------------------ tonemap (tonemap.glsl:59-133)
__global__ void k_tonemap(const float4* __restrict__ accum, int w, int h, float inv_spp, int enable_tonemap,
                          int enable_aces, int simple_aces, int bg_mode, float bgx, float bgy, float bgz,
                          uchar4* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const float4 a = accum[i];
    vec3 color = V(a.x * inv_spp, a.y * inv_spp, a.z * inv_spp);
    const float alpha = a.w * inv_spp;
    if (enable_tonemap) {
        if (enable_aces) {
            if (simple_aces) {
                const float A = 2.51f, B = 0.03f, Y = 2.43f, D = 0.59f, E = 0.14f;
                vec3 num = color * (A * color + V1(B));
                vec3 den = color * (Y * color + V1(D)) + V1(E);
                vec3 r = num / den;
                color = V(gclamp(r.x, 0.f, 1.f), gclamp(r.y, 0.f, 1.f), gclamp(r.z, 0.f, 1.f));
            } else {
                // ACES fitted: input matrix, RRT+ODT fit, output matrix. GLSL's
                // `color * mat3(c0, c1, c2)` is (dot(color, c0), dot(color, c1), dot(color, c2)).
                const vec3 i0 = V(0.59719f, 0.35458f, 0.04823f), i1 = V(0.07600f, 0.90834f, 0.01566f),
                           i2 = V(0.02840f, 0.13383f, 0.83777f);
                const vec3 o0 = V(1.60475f, -0.53108f, -0.07367f), o1 = V(-0.10208f, 1.10813f, -0.00605f),
                           o2 = V(-0.00327f, -0.07276f, 1.07602f);
                vec3 c = V(dot(color, i0), dot(color, i1), dot(color, i2));
                vec3 aa = c * (c + V1(0.0245786f)) - V1(0.000090537f);
                vec3 bb = c * (0.983729f * c + V1(0.4329510f)) + V1(0.238081f);
                c = aa / bb;
                c = V(dot(c, o0), dot(c, o1), dot(c, o2));
                color = V(gclamp(c.x, 0.f, 1.f), gclamp(c.y, 0.f, 1.f), gclamp(c.z, 0.f, 1.f));
            }
        } else {
            color = color * 1.0f / (1.0f + Luminance(color) / 1.5f);
        }
    }
    color = vpow(color, 1.0f / 2.2f);
    float out_alpha = 1.0f;
    vec3 bg = V(bgx, bgy, bgz);
    if (bg_mode == 2) {  // transparent background: checkerboard, alpha from the path
        out_alpha = alpha;
        const int px = i % w, py = i / w;
        float fx = floorf((static_cast<float>(px) + 0.5f) / 10.0f), fy = floorf((static_cast<float>(py) + 0.5f) / 10.0f);
        float m = fmodf(fx + fy, 2.0f);
        float res = gmax(m > 0.f ? 1.f : (m < 0.f ? -1.f : 0.f), 0.0f);
        bg = gmix(V1(0.1f), V1(0.2f), res);
    }
    if (bg_mode != 0) color = gmix(bg, color, alpha);
    auto q = [](float v) {
        v = v < 0.f ? 0.f : (v > 1.f ? 1.f : v);
        return static_cast<unsigned char>(v * 255.0f + 0.5f);
    };
    out[i] = make_uchar4(q(color.x), q(color.y), q(color.z), q(out_alpha));
}

}  // namespace ptd

// ------------------------------------------------------------------ launchers
namespace ptk {
using namespace ptd;

template <class K>
static int max_blocks(K kernel, int bs, size_t lds) {
    int n = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, bs, lds) != hipSuccess) n = 1;
    return n < 1 ? 1 : n;
}

int persistent_blocks(int kind, int bs, int S, bool count, int num_cu, size_t extra_lds) {
    const size_t lds = static_cast<size_t>(S) * bs * sizeof(int);
    int per_cu = 1;
    switch (kind) {
        case KIND_PLAIN: per_cu = count ? max_blocks(k_plain<true>, bs, lds) : max_blocks(k_plain<false>, bs, lds); break;
        case KIND_PLAIN_BATCH: per_cu = max_blocks(k_plain_batch, bs, lds); break;
        case KIND_EXTEND: per_cu = count ? max_blocks(k_extend<true, false, false>, bs, lds) : max_blocks(k_extend<false, false, false>, bs, lds); break;
        case KIND_SHADOW: per_cu = count ? max_blocks(k_shadow<true>, bs, lds) : max_blocks(k_shadow<false>, bs, lds); break;
        case KIND_SHADE: per_cu = max_blocks(k_shade<false, false, false>, bs, 0); break;
        case KIND_SHADE_BATCH: per_cu = max_blocks(k_shade<false, true, false>, bs, 0); break;
        case KIND_SHADE_INLINE: per_cu = max_blocks(k_shade<false, false, true>, bs, lds); break;
        case KIND_SHADE_BATCH_INLINE: per_cu = max_blocks(k_shade<false, true, true>, bs, lds); break;
        case KIND_REPLAY: per_cu = max_blocks(k_replay_shade<false>, bs, extra_lds); break;
        default: break;
    }
    return per_cu * num_cu;
}

void launch_plain(hipStream_t s, int blocks, int bs, const DevScene& sc, const DevOpts& op, float4* accum, int sample,
                  int num_items, int* work, int S, int* spill, unsigned long long* counters, int* ovf, bool count) {
    const size_t lds = static_cast<size_t>(S) * bs * sizeof(int);
    if (count) hipLaunchKernelGGL(k_plain<true>, dim3(blocks), dim3(bs), lds, s, sc, op, accum, sample, num_items, work, S, spill, counters, ovf);
    else hipLaunchKernelGGL(k_plain<false>, dim3(blocks), dim3(bs), lds, s, sc, op, accum, sample, num_items, work, S, spill, counters, ovf);
}

void launch_extend(hipStream_t s, int blocks, int bs, const DevScene& sc, const DevOpts& op, const PathBuf& pb,
                   bool primary, const int* qin, const int* count_in, int npaths, int npix, int first_sample,
                   int* hitq, int* hit_count, int* work, int S, int* spill, unsigned long long* counters, int* ovf,
                   bool count, int budget, const ContBuf* cb, int* cont_count) {
    const size_t lds = static_cast<size_t>(S) * bs * sizeof(int);
    const BatchBuf nb{};
    const ContBuf nc{};
    if (budget > 0 && !primary && !count && cb && cont_count) {
        hipLaunchKernelGGL((k_extend<false, false, false, true>), dim3(blocks), dim3(bs), lds, s, sc, op, pb, qin, count_in, npaths, npix, first_sample, hitq, hit_count, work, S, spill, counters, ovf, nb, budget, *cb, cont_count);
        return;
    }
#define PT_EXT(C, P) hipLaunchKernelGGL((k_extend<C, P, false>), dim3(blocks), dim3(bs), lds, s, sc, op, pb, qin, count_in, npaths, npix, first_sample, hitq, hit_count, work, S, spill, counters, ovf, nb, 0, nc, nullptr)
    if (count) { if (primary) PT_EXT(true, true); else PT_EXT(true, false); }
    else { if (primary) PT_EXT(false, true); else PT_EXT(false, false); }
#undef PT_EXT
}

void launch_extend_resume(hipStream_t s, int blocks, int bs, const DevScene& sc, const DevOpts& op, const PathBuf& pb,
                          const ContBuf& cb, const int* cont_count, int* hitq, int* work, int S, int* spill,
                          int* ovf) {
    const size_t lds = static_cast<size_t>(S) * bs * sizeof(int);
    hipLaunchKernelGGL(k_extend_resume, dim3(blocks), dim3(bs), lds, s, sc, op, pb, cb, cont_count, hitq, work, S, spill, ovf);
}

void launch_shade(hipStream_t s, int blocks, int bs, const DevScene& sc, const DevOpts& op, const PathBuf& pb,
                  const ShadowBuf& sb, const BatchBuf* bb, bool primary, bool inline_nee, const int* hitq,
                  const int* hit_count, int npix, int first_sample, int* qout, int* count_out, int* shadow_count,
                  int* work, bool compact, int S, int* spill, unsigned long long* counters, int* ovf) {
    const int c = compact ? 1 : 0;
    const BatchBuf b = bb ? *bb : BatchBuf{};
    const size_t lds = inline_nee ? static_cast<size_t>(S) * bs * sizeof(int) : 0;
#define PT_SHADE(P, B, I) hipLaunchKernelGGL((k_shade<P, B, I>), dim3(blocks), dim3(bs), lds, s, sc, op, pb, sb, hitq, hit_count, npix, first_sample, qout, count_out, shadow_count, work, c, b, S, spill, counters, ovf)
    if (bb) {
        if (inline_nee) { if (primary) PT_SHADE(true, true, true); else PT_SHADE(false, true, true); }
        else { if (primary) PT_SHADE(true, true, false); else PT_SHADE(false, true, false); }
    } else {
        if (inline_nee) { if (primary) PT_SHADE(true, false, true); else PT_SHADE(false, false, true); }
        else { if (primary) PT_SHADE(true, false, false); else PT_SHADE(false, false, false); }
    }
#undef PT_SHADE
}

void launch_replay_shade(hipStream_t s, int blocks, int bs, const DevScene& sc, const DevOpts& op, const PathBuf& pb,
                         bool primary, const int* hitq, const int* hit_count, int npix, int first_sample, int* qout,
                         int* count_out, int* work, bool compact, const ReplayBuf& rb, int vshift) {
    const int c = compact ? 1 : 0;
    const size_t lds = rb.lds ? static_cast<size_t>(rb.num_mats) * (NPARAM + (rb.mgrad ? NMEDIUM : 0)) * sizeof(float)
                              : 0;
    if (primary) hipLaunchKernelGGL((k_replay_shade<true>), dim3(blocks), dim3(bs), lds, s, sc, op, pb, hitq, hit_count, npix, first_sample, qout, count_out, work, c, rb, vshift);
    else hipLaunchKernelGGL((k_replay_shade<false>), dim3(blocks), dim3(bs), lds, s, sc, op, pb, hitq, hit_count, npix, first_sample, qout, count_out, work, c, rb, vshift);
}

void launch_pass_begin(hipStream_t s, const BatchBuf& bb, int* counters, int ncounters) {
    hipLaunchKernelGGL(k_pass_begin, dim3(1), dim3(256), 0, s, bb, counters, ncounters);
}

void launch_pass_end(hipStream_t s, const BatchBuf& bb) {
    hipLaunchKernelGGL(k_pass_end, dim3(1), dim3(64), 0, s, bb);
}

void launch_extend_batch(hipStream_t s, int blocks, int bs, const DevScene& sc, const DevOpts& op, const PathBuf& pb,
                         const BatchBuf& bb, int* hitq, int* hit_count, int* work, int S, int* spill, int* ovf) {
    const size_t lds = static_cast<size_t>(S) * bs * sizeof(int);
    hipLaunchKernelGGL((k_extend<false, true, true>), dim3(blocks), dim3(bs), lds, s, sc, op, pb, nullptr, nullptr, 0, 0, 0, hitq, hit_count, work, S, spill, nullptr, ovf, bb, 0, ContBuf{}, nullptr);
}

void launch_accum_batch(hipStream_t s, int blocks, const PathBuf& pb, const BatchBuf& bb) {
    hipLaunchKernelGGL(k_accum_batch, dim3(blocks), dim3(256), 0, s, pb, bb);
}

void launch_plain_batch(hipStream_t s, int blocks, int bs, const DevScene& sc, const DevOpts& op, const JobDev* jobs,
                        int njobs, int items, float4* accum, int* work, int S, int* spill, int* ovf) {
    const size_t lds = static_cast<size_t>(S) * bs * sizeof(int);
    hipLaunchKernelGGL(k_plain_batch, dim3(blocks), dim3(bs), lds, s, sc, op, jobs, njobs, items, accum, work, S, spill,
                       ovf);
}

void launch_shadow(hipStream_t s, int blocks, int bs, const DevScene& sc, const DevOpts& op, const PathBuf& pb,
                   const ShadowBuf& sb, const int* count_in, int* work, int S, int* spill,
                   unsigned long long* counters, int* ovf, bool count, unsigned* vis, int vshift) {
    const size_t lds = static_cast<size_t>(S) * bs * sizeof(int);
    if (count) hipLaunchKernelGGL((k_shadow<true>), dim3(blocks), dim3(bs), lds, s, sc, op, pb, sb, count_in, work, S, spill, counters, ovf, vis, vshift);
    else hipLaunchKernelGGL((k_shadow<false>), dim3(blocks), dim3(bs), lds, s, sc, op, pb, sb, count_in, work, S, spill, counters, ovf, vis, vshift);
}

void launch_accum(hipStream_t s, const DevOpts& op, const PathBuf& pb, float4* accum, int npix, int nsamp) {
    hipLaunchKernelGGL(k_accum, dim3((npix + 255) / 256), dim3(256), 0, s, op, pb, accum, npix, nsamp);
}

static_assert(ZS_CHUNK == ENV_CDF_CHUNK, "chunk size of the CDF scan");
static_assert(sizeof(double) + sizeof(int) + sizeof(int4) + sizeof(int2) == 36, "CDF scan scratch per chunk");
size_t env_cdf_scratch_bytes(int n) { return env_cdf_scratch_size(n); }

LayoutSizes layout_sizes() { return layout_sizes_here(); }

void launch_env_cdf(hipStream_t s, const float* img, int n, float* cdf, float* total, int* ok, void* scratch) {
    launch_fill(s, ok, 1, 1);
    if (n > 0) hipLaunchKernelGGL(k_env_lum, dim3((n + 255) / 256), dim3(256), 0, s, img, n, cdf, ok);
    const int nchunks = n > 1 ? (n - 2) / ZS_CHUNK + 1 : 0;
    // scratch: int4 first (16-byte aligned), then double, int2, int
    int4* fun = static_cast<int4*>(scratch);
    double* csum = reinterpret_cast<double*>(fun + nchunks);
    int2* cstate = reinterpret_cast<int2*>(csum + nchunks);
    int* kguess = reinterpret_cast<int*>(cstate + nchunks);
    // exactly one of the two paths does the work, as *ok says
    if (nchunks > 0) {
        hipLaunchKernelGGL(k_env_chunk_sum, dim3(nchunks), dim3(256), 256 * sizeof(double), s, cdf, n, csum, ok);
        hipLaunchKernelGGL(k_env_guess, dim3(1), dim3(64), 0, s, cdf, nchunks, csum, kguess, ok);
        hipLaunchKernelGGL(k_env_chunk_fun, dim3(nchunks), dim3(ZS_THREADS), ZL_WORDS * sizeof(int), s, cdf, n, kguess,
                           fun, ok);
    }
    hipLaunchKernelGGL(k_env_carry, dim3(1), dim3(ZS_THREADS), ZL_WORDS * sizeof(int), s, cdf, n, nchunks, kguess, fun,
                       cstate, total, ok);
    if (nchunks > 0)
        hipLaunchKernelGGL(k_env_chunk_write, dim3(nchunks), dim3(ZS_THREADS), ZL_WORDS * sizeof(int), s, cdf, n,
                           kguess, cstate, ok);
    hipLaunchKernelGGL(k_env_scan, dim3(1), dim3(ENV_SCAN_THREADS), 2 * ENV_CHUNK * sizeof(float), s, cdf, n, total, ok);
}

void launch_fill(hipStream_t s, int* p, int n, int v) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((n + 255) / 256), dim3(256), 0, s, p, n, v);
}

void launch_query_closest(hipStream_t s, const DevScene& sc, const DevOpts& op, const float* o, const float* d, int n,
                          float* t, int* inst, int* tri, int* light, float* uv, int S, int* spill, int* ovf) {
    const int bs = 256;
    hipLaunchKernelGGL(k_query_closest, dim3((n + bs - 1) / bs), dim3(bs), static_cast<size_t>(S) * bs * 4, s, sc, op, o, d, n, t, inst, tri, light, uv, S, spill, ovf);
}

void launch_query_any(hipStream_t s, const DevScene& sc, const DevOpts& op, const float* o, const float* d,
                      const float* maxd, int n, int* occ, int S, int* spill, int* ovf) {
    const int bs = 256;
    hipLaunchKernelGGL(k_query_any, dim3((n + bs - 1) / bs), dim3(bs), static_cast<size_t>(S) * bs * 4, s, sc, op, o, d, maxd, n, occ, S, spill, ovf);
}

void launch_tonemap(hipStream_t s, const float4* accum, int w, int h, float inv_spp, int tm, int aces, int simple,
                    int bg_mode, const float* bg, uchar4* out) {
    hipLaunchKernelGGL(k_tonemap, dim3((w * h + 255) / 256), dim3(256), 0, s, accum, w, h, inv_spp, tm, aces, simple, bg_mode, bg[0], bg[1], bg[2], out);
}

}  // namespace ptk
