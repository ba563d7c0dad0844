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
using ptd::GNode;
using ptd::TlasBuild;
using ptd::TlasSeg;

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
    for (int r = 0; r < 3; ++r) {
        const float xa = m[r] * lb[0], xb = m[r] * lb[3];
        const float ya = m[4 + r] * lb[1], yb = m[4 + r] * lb[4];
        const float za = m[8 + r] * lb[2], zb = m[8 + r] * lb[5];
        const float mn = min_ref(xa, xb) + min_ref(ya, yb) + min_ref(za, zb) + m[12 + r];
        const float mx = max_ref(xa, xb) + max_ref(ya, yb) + max_ref(za, zb) + m[12 + r];
        box[r] = mn;
        box[3 + r] = mx;
        cen[r] = (mx + mn) * 0.5f;
    }
    GInst g;
    // The host validated every transform; a singular one cannot reach here.
    (void)ptd::inst_record(m, a.blas_ref[i], a.mat[i], &g);
    a.inst[i] = g;
}

// Level-synchronous top-down build in one workgroup (n >= 2).
__global__ __launch_bounds__(kBuildThreads) void k_tlas_build(TlasBuild a) {
    extern __shared__ int lds[];
    const int n = a.n;
    const int tid = threadIdx.x, nthr = blockDim.x;
    const int lane = tid & 63, wave = tid >> 6, nwaves = (nthr + 63) >> 6;
    int* ord = a.ord0;
    int* ord2 = a.ord1;
    TlasSeg* cur = a.seg0;
    TlasSeg* nxt = a.seg1;
    unsigned long long* keys = a.keys;
    // Key j of level segment s at keys[j * ks + s]: the 24 keys of one node lie
    // in 24 cache lines (atomics to one line are serialised in L2).
    const size_t ks = static_cast<size_t>(n / 2 + 1);

    // Root: all elements in index order; centroid bounds grown in index order
    // (bvh.cpp:370-376).
    if (tid == 0) {
        for (int j = 0; j < 3; ++j) { keys[j * ks] = ~0ull; keys[(3 + j) * ks] = 0ull; }
        lds[L_NSEG] = 1;
        lds[L_NEXT] = 0;
    }
    for (int p = tid; p < n; p += nthr) { ord[p] = p; a.pseg[p] = 0; }
    __syncthreads();
    for (int p = tid; p < n; p += nthr) {
        const float* c = a.cen + 3 * static_cast<size_t>(p);
        for (int j = 0; j < 3; ++j) {
            atomicMin(keys + j * ks, key_min(c[j], p));
            atomicMax(keys + (3 + j) * ks, key_max(c[j], p));
        }
    }
    __syncthreads();
    if (tid == 0) {
        TlasSeg& g = cur[0];
        g.start = 0;
        g.count = n;
        g.node = a.base;
        for (int j = 0; j < 6; ++j) g.cb[j] = key_value(keys[j * ks]);
    }
    __syncthreads();

    for (;;) {
        const int nseg = lds[L_NSEG];
        if (nseg == 0) break;

        // Split plane of every node of the level (bvh.cpp:94-96, :125-129).
        for (int s = tid; s < nseg; s += nthr) {
            TlasSeg& g = cur[s];
            const float ex = g.cb[3] - g.cb[0], ey = g.cb[4] - g.cb[1], ez = g.cb[5] - g.cb[2];
            int axis = 0;  // bbox::maxdim (bbox.h:70-83)
            if (ex >= ey && ex >= ez) axis = 0;
            else if (ey >= ex && ey >= ez) axis = 1;
            else if (ez >= ex && ez >= ey) axis = 2;
            const float ext = axis == 0 ? ex : (axis == 1 ? ey : ez);
            g.axis = axis;
            g.border = (g.cb[3 + axis] + g.cb[axis]) * 0.5f;
            g.mode = ext > 0.f ? 3 : 0;
            g.near2far = (g.count + g.start) & 1;
            for (int j = 0; j < 24; ++j) {
                const int r = j % 12;  // 0-2, 6-8 minima; 3-5, 9-11 maxima
                keys[j * ks + s] = (r < 3 || (r >= 6 && r < 9)) ? ~0ull : 0ull;
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
                const int s = a.pseg[p];
                if (s >= 0 && cur[s].mode == 3) {
                    const TlasSeg& g = cur[s];
                    const float c = a.cen[3 * static_cast<size_t>(ord[p]) + g.axis];
                    f = g.near2far ? (c < g.border) : (c >= g.border);
                }
            }
            const unsigned long long mask = __ballot(f);
            const int below = static_cast<int>(__builtin_amdgcn_mbcnt_hi(
                static_cast<unsigned>(mask >> 32), __builtin_amdgcn_mbcnt_lo(static_cast<unsigned>(mask), 0u)));
            if (lane == 0) lds[L_WTOT + wave] = __popcll(mask);
            __syncthreads();
            if (tid == 0) {
                int c = lds[L_CARRY];
                for (int w = 0; w < nwaves; ++w) {
                    const int t = lds[L_WTOT + w];
                    lds[L_WTOT + w] = c;
                    c += t;
                }
                lds[L_CARRY] = c;
            }
            __syncthreads();
            if (p < n) a.pre[p] = lds[L_WTOT + wave] + below;
            __syncthreads();
        }
        if (tid == 0) a.pre[n] = lds[L_CARRY];
        __syncthreads();

        // Split of every node; fallback to the middle (bvh.cpp:198-212).
        for (int s = tid; s < nseg; s += nthr) {
            TlasSeg& g = cur[s];
            const int st = g.start, e = st + g.count;
            int csplit = st + (g.count >> 1);
            g.m = 0;
            if (g.mode == 3) {
                const int split = st + (a.pre[e] - a.pre[st]);
                g.split = split;
                if (split == st) g.mode = 1;
                else if (split == e) g.mode = 2;
                else {
                    csplit = split;
                    g.m = (split - st) - (a.pre[split] - a.pre[st]);
                }
            }
            g.csplit = csplit;
        }
        __syncthreads();

        // Positions of the swapped pairs, by rank: a = right-going below the
        // split (from the left), b = left-going at or above it (from the right).
        for (int p = tid; p < n; p += nthr) {
            const int s = a.pseg[p];
            if (s < 0 || cur[s].mode != 3) continue;
            const TlasSeg& g = cur[s];
            const int f = a.pre[p + 1] - a.pre[p];
            if (p < g.split) {
                if (!f) a.apos[g.start + (p - g.start) - (a.pre[p] - a.pre[g.start])] = p;
            } else if (f) {
                a.bpos[g.start + g.m - 1 - (a.pre[p] - a.pre[g.split])] = p;
            }
        }
        __syncthreads();

        // Move the elements and reduce the child boxes: within a wave first, then
        // one store, or one atomic when the child spans several waves, per key
        // and run of lanes.
        for (int base = 0; base < n; base += nthr) {
            const int p = base + tid;
            const int s = p < n ? a.pseg[p] : -1;
            int rk[2] = {-1, -1};  // grow-order rank of this lane's element in each child, -1: none
            int el = 0;
            if (p < n && s < 0) ord2[p] = ord[p];
            if (s >= 0) {
                const TlasSeg& g = cur[s];
                const int st = g.start, e = st + g.count;
                int src = p;
                if (g.mode == 0) {
                    if (p < g.csplit) rk[0] = p - st;
                    else rk[1] = p - g.csplit;
                } else if (g.mode == 1) {
                    rk[1] = (p == st) ? 0 : e - p;
                    if (p < g.csplit) rk[0] = p - st;
                } else if (g.mode == 2) {
                    rk[0] = p - st;
                    if (p >= g.csplit) rk[1] = p - g.csplit;
                } else {
                    const int f = a.pre[p + 1] - a.pre[p];
                    if (p < g.split) {
                        if (!f) src = a.bpos[st + (p - st) - (a.pre[p] - a.pre[st])];
                        rk[0] = p - st;
                    } else if (f) {
                        const int r = g.m - 1 - (a.pre[p] - a.pre[g.split]);
                        src = a.apos[st + r];
                        rk[1] = e - (r == 0 ? e : a.bpos[st + r - 1]);
                    } else if (p == g.split) {
                        rk[1] = e - (g.m == 0 ? e : a.bpos[st + g.m - 1]);
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
            for (int c = 0; c < 2; ++c) {
                // The lanes adding to one child of one node are consecutive
                // positions, so a segmented inclusive scan over the wave leaves
                // each run's reduction in its last lane, which sends the atomics.
                const bool mine = rk[c] >= 0;
                if (__ballot(mine) == 0ull) continue;
                const int grp = mine ? s : -1;
                // keys 0-2 box min, 3-5 box max, 6-8 centre min, 9-11 centre max
                unsigned long long q[12];
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    q[j] = mine ? key_min(v[j], rk[c]) : ~0ull;
                    q[3 + j] = mine ? key_max(v[3 + j], rk[c]) : 0ull;
                    q[6 + j] = mine ? key_min(v[6 + j], rk[c]) : ~0ull;
                    q[9 + j] = mine ? key_max(v[6 + j], rk[c]) : 0ull;
                }
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int gs = __shfl_up(grp, d);
                    const bool take = lane >= d && gs == grp;
                    unsigned long long t[12];
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
                const int gn = __shfl_down(grp, 1);
                if (mine && (lane == 63 || gn != grp)) {
                    // The child's positions [lo, hi): when they all lie in this
                    // wave the run is the whole reduction and is stored as is.
                    const TlasSeg& g = cur[s];
                    const int e = g.start + g.count;
                    const int lo = (c == 1 && g.mode != 1) ? g.csplit : g.start;
                    const int hi = (c == 0 && g.mode != 2) ? g.csplit : e;
                    unsigned long long* k = keys + static_cast<size_t>(12 * c) * ks + s;
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

        // Node records (bvh_translator.cpp:52-80) and the next level.
        for (int s = tid; s < nseg; s += nthr) {
            TlasSeg& g = cur[s];
            float v[24];
            for (int j = 0; j < 24; ++j) v[j] = key_value(keys[j * ks + s]);
            const int st = g.start, e = st + g.count, cs = g.csplit;
            const int nl = cs - st, nr = e - cs;
            GNode node;
            node.a = make_float4(v[0], v[1], v[2], v[3]);
            node.b = make_float4(v[4], v[5], v[12], v[13]);
            node.c = make_float4(v[14], v[15], v[16], v[17]);
            node.d = make_int4(nl == 1 ? -1 - ord2[st] : g.node + 1, nr == 1 ? -1 - ord2[cs] : g.node + nl, 0, 0);
            a.nodes[g.node] = node;
            g.child[0] = g.child[1] = -1;
            for (int c = 0; c < 2; ++c) {
                const int cnt = c == 0 ? nl : nr;
                if (cnt < 2) continue;
                const int id = atomicAdd(&lds[L_NEXT], 1);
                TlasSeg& ch = nxt[id];
                ch.start = c == 0 ? st : cs;
                ch.count = cnt;
                ch.node = c == 0 ? g.node + 1 : g.node + nl;
                for (int j = 0; j < 6; ++j) ch.cb[j] = v[12 * c + 6 + j];
                g.child[c] = id;
            }
        }
        __syncthreads();

        for (int p = tid; p < n; p += nthr) {
            const int s = a.pseg[p];
            if (s < 0) continue;
            const TlasSeg& g = cur[s];
            a.pseg[p] = p < g.csplit ? g.child[0] : g.child[1];
        }
        {
            int* t = ord; ord = ord2; ord2 = t;
            TlasSeg* u = cur; cur = nxt; nxt = u;
        }
        __syncthreads();
        if (tid == 0) {
            lds[L_NSEG] = lds[L_NEXT];
            lds[L_NEXT] = 0;
        }
        __syncthreads();
    }
}

void launch_tlas_rebuild(hipStream_t s, const TlasBuild& a) {
    hipLaunchKernelGGL(k_tlas_prep, dim3((a.n + 255) / 256), dim3(256), 0, s, a);
    if (a.n >= 2)
        hipLaunchKernelGGL(k_tlas_build, dim3(1), dim3(kBuildThreads), kBuildLdsWords * sizeof(int), s, a);
}

}  // namespace ptk
