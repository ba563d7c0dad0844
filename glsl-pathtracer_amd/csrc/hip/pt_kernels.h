// Launch interface between the renderer (host) and the kernels.
#pragma once

#include "pt_adjoint.h"

namespace ptd {

// Wavefront path state, structure of arrays indexed by path slot
// (slot = sample_in_pass * pixels + pixel item).
struct PathBuf {
    float4* ro;    // origin, previous BSDF/phase pdf
    float4* rd;    // direction, previous material roughness
    float4* thr;   // throughput
    float4* rad;   // radiance, alpha
    uint4* rng;    // pcg4d state
    int4* meta;    // depth, pass-through count | flags | medium type << 20, previous material, medium's material
    float4* aux;   // previous texCoord (u, v), medium anisotropy
    float4* med;   // medium colour, density (only with OPT_MEDIUM, else null)
    float4* hit;   // closest hit: t, u, v, triangle row (int bits)
    int2* hit2;    // instance, light
};

// Closest-hit traversals the capped extend stage stopped (pt_options.
// extend_cap), resumed by a second, compacted launch. Record r: the path
// slot, the traversal position, the best hit so far and the stack entries
// (entry k at stack[k * cap + r]).
constexpr int CONT_STACK = 24;  // a traversal deeper than this is not stopped
struct ContBuf {
    int* slot;
    int4* st;     // ref, sp, blas_base, cur_inst
    float4* hf;   // best hit t, u, v
    int4* hi;     // best hit triangle row, instance, light
    int* stack;   // [CONT_STACK * cap]
    int cap;      // records
};

// Shadow (NEE visibility) queue: one record per path with up to two rays
// (k = 0 environment, 1 analytic light) from the same scatter point.
struct ShadowBuf {
    int2* idx;  // [n]    path slot, mask of valid rays (bit k: ray k)
    float4* o;  // [n]    origin (w unused)
    float4* t;  // [n]    throughput at NEE time (w unused)
    float4* d;  // [2n]   ray k: direction, max distance
    float4* c;  // [2n]   ray k: contribution without the throughput
};

// Visibility record of the adjoint's primal pass: per path slot and bounce,
// bit k of the bounce's two bits says NEE ray k (0 environment, 1 analytic
// light) was valid and unoccluded. 16 bounces per 32-bit word; word array w
// (max_paths words) covers bounces [16 w, 16 w + 16), at bit 2 (b mod 16).
constexpr int VIS_BOUNCES_PER_WORD = 16;

// Replay pass of the adjoint: inputs, the primal's records and the outputs.
struct ReplayBuf {
    const float4* ltot;   // primal radiance of every path slot of the pass
    const unsigned* vis;  // primal NEE visibility bits of the current bounce's word (pt_replay.h)
    const float* dimg;    // dLoss / d(mean image), RGB per pixel, rows bottom-up
    float inv_n;          // 1 / number of samples in the mean image
    float* grad;          // [num_mats * NPARAM] gradient sums
    float* mgrad;         // [num_mats * NMEDIUM] medium gradient sums (null: not computed)
    int* err;             // set to 1 when a vertex names a material outside [0, num_mats)
    int* dropped;         // vertices and medium segments whose gradient had a non-finite term
    int num_mats;
    int lds;              // 1: accumulate per block in LDS (grad, then mgrad)
};

// ---- batched render of many jobs (one scene, own camera / size / depth each)
// One job of a batch.
struct JobDev {
    float cam_pos[3], cam_up[3], cam_right[3], cam_fwd[3];
    float cam_fov, cam_focal, cam_aperture;
    int width, height, max_depth, npix;  // npix: pixel items (8x8 tiles * 64)
    int spp, item0;                      // samples; first pixel item in the one-launch plain form
    long long accum_off;                 // float4 offset of the job's image in the batch accumulation
};
// One sample of one job: npix consecutive path slots of a pass starting at slot0.
struct PlaneDev { int job, sample, slot0, pad; };
// The planes of one job inside one pass (for the accumulation kernel).
struct SegDev { int job, plane0, nplanes, thread0; };
// One pass: planes [plane0, plane0 + nplanes), npaths slots, segments [seg0, seg0 + nsegs).
struct PassDev { int plane0, nplanes, npaths, seg0, nsegs, nthreads, pad0, pad1; };
// Pass state read by the kernels: cur[0] = index of the pass being run,
// cur[1..8) = a copy of that pass's PassDev (written by k_pass_begin).
struct BatchBuf {
    const JobDev* jobs;
    const PlaneDev* planes;
    const SegDev* segs;
    const PassDev* passes;
    int* cur;
    float4* accum;
};

}  // namespace ptd

namespace ptk {
enum { KIND_PLAIN = 0, KIND_EXTEND = 1, KIND_SHADE = 2, KIND_SHADOW = 3, KIND_SHADE_BATCH = 4,
       KIND_SHADE_INLINE = 5, KIND_SHADE_BATCH_INLINE = 6, KIND_REPLAY = 7, KIND_PLAIN_BATCH = 8 };

// Resident blocks for a persistent launch (occupancy x CUs); extra_lds is the
// replay stage's per-block gradient table.
int persistent_blocks(int kind, int bs, int S, bool count, int num_cu, size_t extra_lds = 0);
void launch_plain(hipStream_t s, int blocks, int bs, const ptd::DevScene& sc, const ptd::DevOpts& op, float4* accum,
                  int sample, int num_items, int* work, int S, int* spill, unsigned long long* counters, int* ovf,
                  bool count);
void launch_extend(hipStream_t s, int blocks, int bs, const ptd::DevScene& sc, const ptd::DevOpts& op,
                   const ptd::PathBuf& pb, bool primary, const int* qin, const int* count_in, int npaths, int npix,
                   int first_sample, int* hitq, int* hit_count, int* work, int S, int* spill,
                   unsigned long long* counters, int* ovf, bool count, int budget = 0,
                   const ptd::ContBuf* cb = nullptr, int* cont_count = nullptr);
// A later bounce with budget > 0 (and not counting): traversals stop after
// `budget` iterations into cb, each keeping a place in the hit queue;
// launch_extend_resume then finishes them and fills those places (the slot,
// or -1 for a miss).
void launch_extend_resume(hipStream_t s, int blocks, int bs, const ptd::DevScene& sc, const ptd::DevOpts& op,
                          const ptd::PathBuf& pb, const ptd::ContBuf& cb, const int* cont_count, int* hitq,
                          int* work, int S, int* spill, int* ovf);
// Shade stage; bb: the batched form's pass state (null: one frame); inline_nee:
// NEE visibility resolved in the stage (stochastic alpha, volume MIS), else
// queued as shadow records. S and spill: the traversal stack (inline_nee).
void launch_shade(hipStream_t s, int blocks, int bs, const ptd::DevScene& sc, const ptd::DevOpts& op,
                  const ptd::PathBuf& pb, const ptd::ShadowBuf& sb, const ptd::BatchBuf* bb, bool primary,
                  bool inline_nee, const int* hitq, const int* hit_count, int npix, int first_sample, int* qout,
                  int* count_out, int* shadow_count, int* work, bool compact, int S, int* spill,
                  unsigned long long* counters, int* ovf);
// Shadow stage; vis (the adjoint's primal pass, else null): the bounce's word
// array of visibility bits, set at bit vshift + k for visible ray k.
void launch_shadow(hipStream_t s, int blocks, int bs, const ptd::DevScene& sc, const ptd::DevOpts& op,
                   const ptd::PathBuf& pb, const ptd::ShadowBuf& sb, const int* count_in, int* work, int S,
                   int* spill, unsigned long long* counters, int* ovf, bool count, unsigned* vis = nullptr,
                   int vshift = 0);
// Replay stage of the adjoint (extend is the forward renderer's).
void launch_replay_shade(hipStream_t s, int blocks, int bs, const ptd::DevScene& sc, const ptd::DevOpts& op,
                         const ptd::PathBuf& pb, bool primary, const int* hitq, const int* hit_count, int npix,
                         int first_sample, int* qout, int* count_out, int* work, bool compact,
                         const ptd::ReplayBuf& rb, int vshift);
void launch_accum(hipStream_t s, const ptd::DevOpts& op, const ptd::PathBuf& pb, float4* accum, int npix, int nsamp);
// batched render: the stage kernels of one pass, parameterised by device-side pass state
void launch_pass_begin(hipStream_t s, const ptd::BatchBuf& bb, int* counters, int ncounters);
void launch_pass_end(hipStream_t s, const ptd::BatchBuf& bb);
void launch_extend_batch(hipStream_t s, int blocks, int bs, const ptd::DevScene& sc, const ptd::DevOpts& op,
                         const ptd::PathBuf& pb, const ptd::BatchBuf& bb, int* hitq, int* hit_count, int* work, int S,
                         int* spill, int* ovf);
void launch_accum_batch(hipStream_t s, int blocks, const ptd::PathBuf& pb, const ptd::BatchBuf& bb);
// Batched render in one persistent launch of the plain form (jobs sorted by
// item0, `items` = the sum of their npix).
void launch_plain_batch(hipStream_t s, int blocks, int bs, const ptd::DevScene& sc, const ptd::DevOpts& op,
                        const ptd::JobDev* jobs, int njobs, int items, float4* accum, int* work, int S, int* spill,
                        int* ovf);
// Environment-map CDF and total luminance of an RGB32F image of n texels
// (ok: one int of scratch; scratch: env_cdf_scratch_bytes(n) bytes, 8-byte aligned).
// The parallel scan works in chunks of ENV_CDF_CHUNK texels (texels 1..n-1)
// and keeps 36 bytes per chunk (sum, grid, function, carried state) plus 64.
constexpr int ENV_CDF_CHUNK = 16384;
constexpr size_t env_cdf_scratch_size(long long n) {
    return (n > 1 ? static_cast<size_t>((n - 2) / ENV_CDF_CHUNK + 1) : 0) * 36 + 64;
}
size_t env_cdf_scratch_bytes(int n);
// Sizes of the records the launchers take from the host code, as this
// translation unit compiled them; the host side compares them with its own
// at pt_create, so a library linked from objects built against different
// headers refuses to run instead of handing the kernels misread pointers.
struct LayoutSizes { size_t scene, opts, path, shadow, batch, replay, inst, node, job, cont; };
LayoutSizes layout_sizes();
inline LayoutSizes layout_sizes_here() {
    return LayoutSizes{sizeof(ptd::DevScene), sizeof(ptd::DevOpts), sizeof(ptd::PathBuf), sizeof(ptd::ShadowBuf),
                       sizeof(ptd::BatchBuf), sizeof(ptd::ReplayBuf), sizeof(ptd::GInst), sizeof(ptd::GNode),
                       sizeof(ptd::JobDev), sizeof(ptd::ContBuf)};
}
void launch_env_cdf(hipStream_t s, const float* img, int n, float* cdf, float* total, int* ok, void* scratch);
void launch_fill(hipStream_t s, int* p, int n, int v);
void launch_query_closest(hipStream_t s, const ptd::DevScene& sc, const ptd::DevOpts& op, const float* o,
                          const float* d, int n, float* t, int* inst, int* tri, int* light, float* uv, int S,
                          int* spill, int* ovf);
void launch_query_any(hipStream_t s, const ptd::DevScene& sc, const ptd::DevOpts& op, const float* o, const float* d,
                      const float* maxd, int n, int* occ, int S, int* spill, int* ovf);
void launch_tonemap(hipStream_t s, const float4* accum, int w, int h, float inv_spp, int tm, int aces, int simple,
                    int bg_mode, const float* bg, uchar4* out);
}  // namespace ptk
