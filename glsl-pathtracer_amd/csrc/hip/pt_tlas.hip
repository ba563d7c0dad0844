// Device rebuild of the instance TLAS.
//
// Scene::RebuildInstances (src/core/Scene.cpp:200-214) recomputes the world
// box of every instance from its BLAS bounds and transform (:150-184), builds
// the top-level tree with RadeonRays::Bvh(10, 64, false) — split at the centre
// of the centroid bounds on their widest axis, one instance per leaf
// (thirdparty/RadeonRays/bvh.cpp:68-243, :359-490) — and BvhTranslator::
// UpdateTLAS writes it in preorder, left child first (bvh_translator.cpp:52-80,
// :108-114). These kernels produce the same tree and write it straight into the
// BVH2 node array, together with the instance records.
//
// The reference splits one node at a time with an in-place two-pointer
// partition. Its result has a closed form in the elements' side flags, which
// lets one workgroup split every node of a tree level at once:
//   * with nL elements going left, the split is start + nL;
//   * the k-th right-going element from the left among positions below the
//     split trades places with the k-th left-going element from the right
//     among positions at or above it; every other element stays.
// The child boxes are unions, which any order reproduces except for the sign
// of a zero coordinate: the reference keeps the first element grown at the
// extreme value. Each element therefore carries its grow-order rank:
//   * left child: position - start (the partition grows it left to right);
//   * right child: the positions b_1 > b_2 > ... > b_m that received a swapped
//     element, and the split, cut the right part into runs [b_k, b_(k-1)),
//     with b_0 = end and the last run starting at the split. A run is grown at
//     its lowest position first and then downward from its top, the highest run
//     first. So rank(p) = end - p, except at a run's lowest position, whose rank
//     is end - (top of the run).
// Box reductions use (value, rank) keys: the smallest value wins, and among
// equal values the earliest rank. When every element lands on one side, the
// reference falls back to the middle split and grows the halves in position
// order into the boxes the partition already grew, so the side that received
// everything keeps the box of the whole node; with a zero centroid extent it
// does not partition and grows both halves in position order.
#include "pt_tlas.h"

namespace ptk {
namespace {

using ptd::GInst;
using ptd::Child4;
using ptd::GNode;
using ptd::GNode4;
using ptd::TlasBuild;
using ptd::expand4;
using ptd::make_node4;
using ptd::kTlasSegWords;

constexpr int kBuildThreads = 1024;
// LDS words of the build kernel.
constexpr int L_CARRY = 0, L_NSEG = 1, L_NEXT = 2, L_WTOT = 4;  // L_WTOT: one word per wave
constexpr int kBuildLdsWords = L_WTOT + kBuildThreads / 64;

// std::min / std::max of the reference's Vec3::Min / Max (Vec3.h:119-135):
// the first argument wins ties.
__device__ __forceinline__ float min_ref(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float max_ref(float a, float b) { return (a < b) ? b : a; }

// Reduction keys. High word: the value in an order-preserving encoding with
// -0 and +0 equal. Low word: the grow-order rank (lowest wins among equal
// values) above the sign bit of the value, which a zero result takes.
__device__ __forceinline__ unsigned ord_bits(float v) {
    const unsigned u = (v == 0.0f) ? 0u : __float_as_uint(v);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ unsigned long long key_min(float v, unsigned rank) {
    return (static_cast<unsigned long long>(ord_bits(v)) << 32) | (rank << 1) | (__float_as_uint(v) >> 31);
}
__device__ __forceinline__ unsigned long long key_max(float v, unsigned rank) {
    return (static_cast<unsigned long long>(ord_bits(v)) << 32) | ((0x7FFFFFFFu - rank) << 1) |
           (__float_as_uint(v) >> 31);
}
__device__ __forceinline__ float key_value(unsigned long long k) {
    const unsigned o = static_cast<unsigned>(k >> 32);
    unsigned u = (o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o;
    if (u == 0u) u = static_cast<unsigned>(k & 1ull) << 31;
    return __uint_as_float(u);
}

// The rank only decides the sign of a zero extreme. Without any -0.0 among
// the world boxes and centres every zero is +0.0, and the value word alone
// (32-bit keys) gives the same bounds.
template <typename K> struct Key;
template <> struct Key<unsigned long long> {
    static __device__ __forceinline__ unsigned long long lo(float v, unsigned r) { return key_min(v, r); }
    static __device__ __forceinline__ unsigned long long hi(float v, unsigned r) { return key_max(v, r); }
    static __device__ __forceinline__ float value(unsigned long long k) { return key_value(k); }
};
template <> struct Key<unsigned> {
    static __device__ __forceinline__ unsigned lo(float v, unsigned) { return ord_bits(v); }
    static __device__ __forceinline__ unsigned hi(float v, unsigned) { return ord_bits(v); }
    static __device__ __forceinline__ float value(unsigned o) {
        return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
    }
};

}  // namespace

// Per instance: world box (Scene.cpp:155-178), its centre (bbox.cpp:28) and
// the instance record.
__global__ void k_tlas_prep(TlasBuild a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float* m = a.xf + 16 * static_cast<size_t>(i);
    const float* lb = a.local + 6 * static_cast<size_t>(i);
    float* box = a.box + 6 * static_cast<size_t>(i);
    float* cen = a.cen + 3 * static_cast<size_t>(i);
    bool negz = false;
    for (int r = 0; r < 3; ++r) {
        const float xa = m[r] * lb[0], xb = m[r] * lb[3];
        const float ya = m[4 + r] * lb[1], yb = m[4 + r] * lb[4];
        const float za = m[8 + r] * lb[2], zb = m[8 + r] * lb[5];
        const float mn = min_ref(xa, xb) + min_ref(ya, yb) + min_ref(za, zb) + m[12 + r];
        const float mx = max_ref(xa, xb) + max_ref(ya, yb) + max_ref(za, zb) + m[12 + r];
        box[r] = mn;
        box[3 + r] = mx;
        cen[r] = (mx + mn) * 0.5f;
        negz = negz || __float_as_uint(mn) == 0x80000000u || __float_as_uint(mx) == 0x80000000u ||
               __float_as_uint(cen[r]) == 0x80000000u;
    }
    if (negz) a.negz[0] = 1;
    GInst g;
    // The host validated every transform; a singular one cannot reach here.
    (void)ptd::inst_record(m, a.blas_ref[i], a.mat[i], &g, a.blas4_ref[i]);
    a.inst[i] = g;
}

// Level-synchronous top-down build in one workgroup (n >= 2). K: key type;
// ArrLds, KeyLds: the arrays, and the keys, in LDS (fixed per instantiation, so
// their accesses compile to LDS instructions).
template <typename K, bool ArrLds, bool KeyLds>
__device__ __forceinline__ void tlas_build(const TlasBuild& a, int* lds) {
    constexpr K kNone = static_cast<K>(~static_cast<K>(0));
    const int n = a.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6;
    constexpr bool in_lds = ArrLds;
    int* const arr = lds + kBuildLdsWords;
    int* ord = in_lds ? arr : a.ord0;
    int* ord2 = in_lds ? arr + n : a.ord1;
    int* const pseg = in_lds ? arr + 2 * n : a.pseg;
    int* const apos = in_lds ? arr + 3 * n : a.apos;
    int* const bpos = in_lds ? arr + 4 * n : a.bpos;
    int* const pre = in_lds ? arr + 5 * n : a.pre;  // n + 1 words
    // Segment table (layout in pt_tlas.h), after the per-position arrays.
    const int hs = n / 2 + 1;
    int* const sw = in_lds ? arr + 6 * n + 1 : a.seg;
    int* const sstart = sw;  // start, count, node: [2 hs], current level at lv * hs
    int* const scount = sw + 2 * hs;
    int* const snode = sw + 4 * hs;
    float* const scb = reinterpret_cast<float*>(sw + 6 * hs);
    float* const sborder = reinterpret_cast<float*>(sw + 12 * hs);
    int* const sinfo = sw + 13 * hs;
    int* const ssplit = sw + 14 * hs;
    int* const scsplit = sw + 15 * hs;
    int* const sm = sw + 16 * hs;
    int* const schild = sw + 17 * hs;  // child c of segment s at c * hs + s
    int lv = 0;
    // 32-bit keys after the segment table when the launcher reserved the LDS.
    K* const keys = KeyLds ? reinterpret_cast<K*>(sw + kTlasSegWords * hs) : reinterpret_cast<K*>(a.keys);
    // Key j of level segment s at keys[j * ks + s]: the runs of one wave reduce
    // into neighbouring slots. In HBM, the keys of one node lie in 24 cache lines.
    const size_t ks = static_cast<size_t>(n / 2 + 1);

    // Root: all elements in index order; centroid bounds grown in index order
    // (bvh.cpp:370-376).
    if (tid == 0) {
        for (int j = 0; j < 3; ++j) { keys[j * ks] = kNone; keys[(3 + j) * ks] = 0; }
        lds[L_NSEG] = 1;
        lds[L_NEXT] = 0;
    }
    for (int p = tid; p < n; p += nthr) { ord[p] = p; pseg[p] = 0; }
    __syncthreads();
    {
        // Reduced per thread, then over the wave; one atomic per wave and key.
        K q[6] = {kNone, kNone, kNone, 0, 0, 0};
        for (int p = tid; p < n; p += nthr) {
            const float* c = a.cen + 3 * static_cast<size_t>(p);
            for (int j = 0; j < 3; ++j) {
                const K kn = Key<K>::lo(c[j], p), kx = Key<K>::hi(c[j], p);
                q[j] = kn < q[j] ? kn : q[j];
                q[3 + j] = kx > q[3 + j] ? kx : q[3 + j];
            }
        }
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const K kn = __shfl_xor(q[j], d), kx = __shfl_xor(q[3 + j], d);
                q[j] = kn < q[j] ? kn : q[j];
                q[3 + j] = kx > q[3 + j] ? kx : q[3 + j];
            }
        }
        if (lane == 0) {
            for (int j = 0; j < 3; ++j) {
                atomicMin(keys + j * ks, q[j]);
                atomicMax(keys + (3 + j) * ks, q[3 + j]);
            }
        }
    }
    __syncthreads();
    if (tid == 0) {
        sstart[0] = 0;
        scount[0] = n;
        snode[0] = a.base;
        for (int j = 0; j < 6; ++j) scb[j] = Key<K>::value(keys[j * ks]);
    }
    __syncthreads();

    for (;;) {
        const int nseg = lds[L_NSEG];
        if (nseg == 0) break;

        // Split plane of every node of the level (bvh.cpp:94-96, :125-129).
        for (int s = tid; s < nseg; s += nthr) {
            const float* cb = scb + 6 * s;
            const float ex = cb[3] - cb[0], ey = cb[4] - cb[1], ez = cb[5] - cb[2];
            int axis = 0;  // bbox::maxdim (bbox.h:70-83)
            if (ex >= ey && ex >= ez) axis = 0;
            else if (ey >= ex && ey >= ez) axis = 1;
            else if (ez >= ex && ez >= ey) axis = 2;
            const float ext = axis == 0 ? ex : (axis == 1 ? ey : ez);
            sborder[s] = (cb[3 + axis] + cb[axis]) * 0.5f;
            sinfo[s] = (ext > 0.f ? 3 : 0) | (axis << 2) | (((scount[lv * hs + s] + sstart[lv * hs + s]) & 1) << 4);
            for (int j = 0; j < 24; ++j) {
                const int r = j % 12;  // 0-2, 6-8 minima; 3-5, 9-11 maxima
                keys[j * ks + s] = (r < 3 || (r >= 6 && r < 9)) ? kNone : 0;
            }
        }
        if (tid == 0) lds[L_CARRY] = 0;
        __syncthreads();

        // Side flags (near2far: centroid < border goes left; otherwise
        // centroid >= border goes left) and their exclusive prefix count.
        for (int base = 0; base < n; base += nthr) {
            const int p = base + tid;
            int f = 0;
            if (p < n) {
                const int s = pseg[p];
                const int in = s >= 0 ? sinfo[s] : 0;
                if ((in & 3) == 3) {
                    const float c = a.cen[3 * static_cast<size_t>(ord[p]) + ((in >> 2) & 3)];
                    f = (in & 16) ? (c < sborder[s]) : (c >= sborder[s]);
                }
            }
            const unsigned long long mask = __ballot(f);
            const int below = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                static_cast<unsigned>(mask >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u)));
            if (lane == 0) lds[L_WTOT + wave] = __popcll(mask);
            __syncthreads();
            // Every wave adds up the counts of the waves before it.
            int off = lds[L_CARRY], tot = 0;
#pragma unroll
            for (int w = 0; w < kBuildThreads / 64; ++w) {
                const int t = lds[L_WTOT + w];
                off += w < wave ? t : 0;
                tot += t;
            }
            if (p < n) pre[p] = off + below;
            __syncthreads();
            if (tid == 0) lds[L_CARRY] += tot;
        }
        if (tid == 0) pre[n] = lds[L_CARRY];
        __syncthreads();

        // Split of every node; fallback to the middle (bvh.cpp:198-212).
        for (int s = tid; s < nseg; s += nthr) {
            const int st = sstart[lv * hs + s], cnt = scount[lv * hs + s], e = st + cnt;
            int csplit = st + (cnt >> 1), mm = 0, in = sinfo[s];
            if ((in & 3) == 3) {
                const int split = st + (pre[e] - pre[st]);
                ssplit[s] = split;
                if (split == st) in = (in & ~3) | 1;
                else if (split == e) in = (in & ~3) | 2;
                else {
                    csplit = split;
                    mm = (split - st) - (pre[split] - pre[st]);
                }
            }
            sinfo[s] = in;
            sm[s] = mm;
            scsplit[s] = csplit;
        }
        __syncthreads();

        // Positions of the swapped pairs, by rank: a = right-going below the
        // split (from the left), b = left-going at or above it (from the right).
        for (int p = tid; p < n; p += nthr) {
            const int s = pseg[p];
            if (s < 0 || (sinfo[s] & 3) != 3) continue;
            const int st = sstart[lv * hs + s], sp = ssplit[s];
            const int f = pre[p + 1] - pre[p];
            if (p < sp) {
                if (!f) apos[st + (p - st) - (pre[p] - pre[st])] = p;
            } else if (f) {
                bpos[st + sm[s] - 1 - (pre[p] - pre[sp])] = p;
            }
        }
        __syncthreads();

        // Move the elements and reduce the child boxes: within a wave first, then
        // one store, or one atomic when the child spans several waves, per key
        // and run of lanes.
        for (int base = 0; base < n; base += nthr) {
            const int p = base + tid;
            const int s = p < n ? pseg[p] : -1;
            int rk[2] = {-1, -1};  // grow-order rank of this lane's element in each child, -1: none
            int el = 0, st = 0, e = 0, cs = 0, mode = 0;
            if (p < n && s < 0) ord2[p] = ord[p];
            if (s >= 0) {
                st = sstart[lv * hs + s];
                e = st + scount[lv * hs + s];
                cs = scsplit[s];
                mode = sinfo[s] & 3;
                int src = p;
                if (mode == 0) {
                    if (p < cs) rk[0] = p - st;
                    else rk[1] = p - cs;
                } else if (mode == 1) {
                    rk[1] = (p == st) ? 0 : e - p;
                    if (p < cs) rk[0] = p - st;
                } else if (mode == 2) {
                    rk[0] = p - st;
                    if (p >= cs) rk[1] = p - cs;
                } else {
                    const int sp = ssplit[s], mm = sm[s];
                    const int f = pre[p + 1] - pre[p];
                    if (p < sp) {
                        if (!f) src = bpos[st + (p - st) - (pre[p] - pre[st])];
                        rk[0] = p - st;
                    } else if (f) {
                        const int r = mm - 1 - (pre[p] - pre[sp]);
                        src = apos[st + r];
                        rk[1] = e - (r == 0 ? e : bpos[st + r - 1]);
                    } else if (p == sp) {
                        rk[1] = e - (mm == 0 ? e : bpos[st + mm - 1]);
                    } else {
                        rk[1] = e - p;
                    }
                }
                el = ord[src];
                ord2[p] = el;
            }
            if (__ballot(s >= 0) == 0ull) continue;
            float v[9];  // box min xyz, box max xyz, centre xyz
            if (s >= 0) {
                const float* b = a.box + 6 * static_cast<size_t>(el);
                const float* ce = a.cen + 3 * static_cast<size_t>(el);
                for (int j = 0; j < 6; ++j) v[j] = b[j];
                for (int j = 0; j < 3; ++j) v[6 + j] = ce[j];
            }
            // A position adds to one child, except in the fallback splits, where
            // the side that received everything spans the whole node and the
            // other half adds to the other child as well. The first pass takes
            // the spanning child (else the only one), the second the other.
            const int cp = mode == 1 ? 1 : (mode == 2 ? 0 : (rk[0] >= 0 ? 0 : 1));
            for (int pass = 0; pass < 2; ++pass) {
                const int c = pass == 0 ? cp : 1 - cp;
                const bool mine = s >= 0 && rk[c] >= 0;
                if (__ballot(mine) == 0ull) continue;
                // The child's positions [lo, hi) are consecutive lanes, so a
                // segmented inclusive scan over the wave leaves the reduction of
                // the run in its last lane; the scan stops at the longest run.
                const int lo = (c == 1 && mode != 1) ? cs : st;
                const int hi = (c == 0 && mode != 2) ? cs : e;
                // keys 0-2 box min, 3-5 box max, 6-8 centre min, 9-11 centre max
                K q[12];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    q[j] = mine ? Key<K>::lo(v[j], rk[c]) : kNone;
                    q[3 + j] = mine ? Key<K>::hi(v[3 + j], rk[c]) : 0;
                    q[6 + j] = mine ? Key<K>::lo(v[6 + j], rk[c]) : kNone;
                    q[9 + j] = mine ? Key<K>::hi(v[6 + j], rk[c]) : 0;
                }
                for (int d = 1; d < 64; d <<= 1) {
                    const bool take = mine && lane >= d && p - d >= lo;
                    if (__ballot(take) == 0ull) break;
                    K t[12];
#pragma unroll
                    for (int j = 0; j < 12; ++j) t[j] = __shfl_up(q[j], d);
                    if (take) {
#pragma unroll
                        for (int j = 0; j < 3; ++j) {
                            q[j] = t[j] < q[j] ? t[j] : q[j];
                            q[3 + j] = t[3 + j] > q[3 + j] ? t[3 + j] : q[3 + j];
                            q[6 + j] = t[6 + j] < q[6 + j] ? t[6 + j] : q[6 + j];
                            q[9 + j] = t[9 + j] > q[9 + j] ? t[9 + j] : q[9 + j];
                        }
                    }
                }
                if (mine && (lane == 63 || p == hi - 1)) {
                    // When the whole child lies in this wave the run is the whole
                    // reduction and is stored as is.
                    K* k = keys + static_cast<size_t>(12 * c) * ks + s;
                    if (lo >= p - lane && p == hi - 1) {
#pragma unroll
                        for (int j = 0; j < 12; ++j) k[j * ks] = q[j];
                    } else {
#pragma unroll
                        for (int j = 0; j < 3; ++j) {
                            atomicMin(k + j * ks, q[j]);
                            atomicMax(k + (3 + j) * ks, q[3 + j]);
                            atomicMin(k + (6 + j) * ks, q[6 + j]);
                            atomicMax(k + (9 + j) * ks, q[9 + j]);
                        }
                    }
                }
            }
        }
        __syncthreads();

        // Node records (bvh_translator.cpp:52-80) and the next level. The
        // children that are inner nodes are numbered in position order (a
        // prefix count over the level's segments, themselves in position
        // order), so the runs of one wave reduce into neighbouring key slots.
        for (int sb = 0; sb < nseg; sb += nthr) {
            const int s = sb + tid;
            const bool live = s < nseg;
            int st = 0, cs = 0, gn = 0, nl = 0, nr = 0;
            float v[24];
            if (live) {
                for (int j = 0; j < 24; ++j) v[j] = Key<K>::value(keys[j * ks + s]);
                st = sstart[lv * hs + s];
                cs = scsplit[s];
                gn = snode[lv * hs + s];
                nl = cs - st;
                nr = st + scount[lv * hs + s] - cs;
                GNode node;
                node.a = make_float4(v[0], v[1], v[2], v[3]);
                node.b = make_float4(v[4], v[5], v[12], v[13]);
                node.c = make_float4(v[14], v[15], v[16], v[17]);
                node.d = make_int4(nl == 1 ? -1 - ord2[st] : gn + 1, nr == 1 ? -1 - ord2[cs] : gn + nl, 0, 0);
                a.nodes[gn] = node;
            }
            const unsigned long long ml = __ballot(live && nl >= 2), mr = __ballot(live && nr >= 2);
            const int below = static_cast<int>(
                __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(ml >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(ml), 0u)) +
                __builtin_amdgcn_mbcnt_hi(static_cast<unsigned>(mr >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mr), 0u)));
            if (lane == 0) lds[L_WTOT + wave] = __popcll(ml) + __popcll(mr);
            __syncthreads();
            int id = lds[L_NEXT], tot = 0;
#pragma unroll
            for (int w = 0; w < kBuildThreads / 64; ++w) {
                const int t = lds[L_WTOT + w];
                id += w < wave ? t : 0;
                tot += t;
            }
            id += below;
            if (live) {
                const int nx = (lv ^ 1) * hs;
                for (int c = 0; c < 2; ++c) {
                    const int cnt = c == 0 ? nl : nr;
                    int cid = -1;
                    if (cnt >= 2) {
                        cid = id++;
                        sstart[nx + cid] = c == 0 ? st : cs;
                        scount[nx + cid] = cnt;
                        snode[nx + cid] = c == 0 ? gn + 1 : gn + nl;
                        // Only the start of a level reads the centroid bounds.
                        for (int j = 0; j < 6; ++j) scb[6 * cid + j] = v[12 * c + 6 + j];
                    }
                    schild[c * hs + s] = cid;
                }
            }
            __syncthreads();
            if (tid == 0) lds[L_NEXT] += tot;
        }
        __syncthreads();

        for (int p = tid; p < n; p += nthr) {
            const int s = pseg[p];
            if (s < 0) continue;
            pseg[p] = p < scsplit[s] ? schild[s] : schild[hs + s];
        }
        {
            int* t = ord; ord = ord2; ord2 = t;
            lv ^= 1;
        }
        __syncthreads();
        if (tid == 0) {
            lds[L_NSEG] = lds[L_NEXT];
            lds[L_NEXT] = 0;
        }
        __syncthreads();
    }
}

__global__ __launch_bounds__(kBuildThreads) void k_tlas_build(TlasBuild a) {
    extern __shared__ int lds[];
    using u64 = unsigned long long;
    if (a.negz[0]) {
        // The LDS reserved for 32-bit keys does not hold 64-bit ones.
        if (a.arr_lds) tlas_build<u64, true, false>(a, lds);
        else tlas_build<u64, false, false>(a, lds);
    } else if (a.key_lds) {
        tlas_build<unsigned, true, true>(a, lds);
    } else if (a.arr_lds) {
        tlas_build<unsigned, true, false>(a, lds);
    } else {
        tlas_build<unsigned, false, false>(a, lds);
    }
}

// Four-wide TLAS of a rebuilt BVH2 TLAS (one workgroup), the array the host's
// collapse4_tlas writes: breadth first, level by level. Each thread converts
// one node of the level (expand4); the places of its internal children in the
// next level come from a workgroup prefix sum over the threads' counts, so
// the numbering is the host's first-in-first-out order. out holds `room`
// entries; the tail after the last node is zeroed. qa and qb hold the BVH2
// references of the current and the next level (room entries each).
constexpr int kTlas4Threads = 256;
__global__ __launch_bounds__(kTlas4Threads) void k_tlas4(const GNode* __restrict__ nodes, int root2,
                                                         GNode4* __restrict__ out, int base, int room, int* qa,
                                                         int* qb) {
    constexpr int kWaves = kTlas4Threads / 64;
    __shared__ int wsum[kWaves];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    if (tid == 0) qa[0] = root2;
    __syncthreads();
    int cur = 1, start = 0;  // nodes of this level, index of its first node
    // a tree of at most room internal nodes never fails the bound (it guards the loop)
    while (cur > 0 && start + cur <= room) {
        int nxt = 0;
        for (int c0 = 0; c0 < cur; c0 += kTlas4Threads) {
            const int i = c0 + tid;
            Child4 ch[4];
            int k = 0, inner = 0;
            if (i < cur) {
                k = expand4(nodes, 0, qa[i], ch);
                for (int j = 0; j < k; ++j) inner += ch[j].ref >= 0 ? 1 : 0;
            }
            int incl = inner;
            for (int d = 1; d < 64; d <<= 1) {
                const int t = __shfl_up(incl, d, 64);
                if (lane >= d) incl += t;
            }
            if (lane == 63) wsum[wv] = incl;
            __syncthreads();
            int woff = 0, total = 0;
            for (int w = 0; w < kWaves; ++w) {
                if (w < wv) woff += wsum[w];
                total += wsum[w];
            }
            if (i < cur) {
                int p = nxt + woff + incl - inner;
                int ref[4];
                for (int j = 0; j < k; ++j) {
                    ref[j] = ch[j].ref;
                    if (ref[j] >= 0) {
                        qb[p] = ch[j].ref;
                        ref[j] = base + start + cur + p;
                        ++p;
                    }
                }
                out[start + i] = make_node4(ch, k, ref);
            }
            nxt += total;
            __syncthreads();
        }
        start += cur;
        cur = nxt;
        int* t = qa;
        qa = qb;
        qb = t;
    }
    for (int i = start + tid; i < room; i += kTlas4Threads) out[i] = GNode4{};
}

void launch_tlas_rebuild(hipStream_t s, const TlasBuild& a) {
    hipLaunchKernelGGL(k_tlas_prep, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    if (a.n >= 2) {
        // The per-position arrays (two orders, segment of each position, the two
        // swap-position tables, the side-flag prefix count: 6n + 1 words) and the
        // segment table live in LDS after its first kBuildLdsWords words where the
        // workgroup's LDS holds them, and the 32-bit keys (24 per segment) too where it holds
        // those as well: on gfx950 (160 KB) up to 2640 and 1487 instances.
        const size_t cap = static_cast<size_t>(a.lds_bytes) / sizeof(int);
        const size_t hs = static_cast<size_t>(a.n) / 2 + 1;
        const size_t arr = 6 * static_cast<size_t>(a.n) + 1 + kTlasSegWords * hs;
        TlasBuild b = a;
        b.arr_lds = b.key_lds = 0;
        size_t words = kBuildLdsWords;
        if (words + arr <= cap) {
            words += arr;
            b.arr_lds = 1;
            if (words + 24 * hs <= cap) {
                words += 24 * hs;
                b.key_lds = 1;
            }
        }
        hipLaunchKernelGGL(k_tlas_build, dim3(1), dim3(kBuildThreads), words * sizeof(int), s, b);
    }
}

void launch_tlas4(hipStream_t s, const GNode* nodes, int root2, GNode4* out, int base, int room, int* qa, int* qb) {
    hipLaunchKernelGGL(k_tlas4, dim3(1), dim3(kTlas4Threads), 0, s, nodes, root2, out, base, room, qa, qb);
}

TlasLayout tlas_layout() { return tlas_layout_here(); }

}  // namespace ptk
