// Instance records and the device rebuild of the instance TLAS: the record
// math shared by the host upload and the device rebuild, the rebuild's
// buffers, and its launcher.
#pragma once

#include "pt_device.h"

namespace ptd {

// Instance record from a column-major Mat4 (16 floats): inverse of the
// affine part in double, rounded to float, plus the forward linear part, the
// BLAS root reference and the material id. The same code runs on the host
// (scene upload) and on the device (instance rebuild), so both give the same
// bits. Returns false for a singular or non-finite linear part.
__host__ __device__ __forceinline__ bool inst_record(const float* m, int blas_ref, int material, GInst* g,
                                                     int blas4_ref = 0) {
    double a[3][3];  // a[col][row]
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) a[c][r] = m[4 * c + r];
    const double det = a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2]) -
                       a[1][0] * (a[0][1] * a[2][2] - a[2][1] * a[0][2]) +
                       a[2][0] * (a[0][1] * a[1][2] - a[1][1] * a[0][2]);
    if (det == 0.0 || !__builtin_isfinite(det)) return false;
    const double id = 1.0 / det;
    double b[3][3];  // b[col][row] = inverse
    b[0][0] = (a[1][1] * a[2][2] - a[2][1] * a[1][2]) * id;
    b[1][0] = -(a[1][0] * a[2][2] - a[2][0] * a[1][2]) * id;
    b[2][0] = (a[1][0] * a[2][1] - a[2][0] * a[1][1]) * id;
    b[0][1] = -(a[0][1] * a[2][2] - a[2][1] * a[0][2]) * id;
    b[1][1] = (a[0][0] * a[2][2] - a[2][0] * a[0][2]) * id;
    b[2][1] = -(a[0][0] * a[2][1] - a[2][0] * a[0][1]) * id;
    b[0][2] = (a[0][1] * a[1][2] - a[1][1] * a[0][2]) * id;
    b[1][2] = -(a[0][0] * a[1][2] - a[1][0] * a[0][2]) * id;
    b[2][2] = (a[0][0] * a[1][1] - a[1][0] * a[0][1]) * id;
    const double t[3] = {m[12], m[13], m[14]};
    float ic[4][3];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) ic[c][r] = static_cast<float>(b[c][r]);
    for (int r = 0; r < 3; ++r) ic[3][r] = static_cast<float>(-(b[0][r] * t[0] + b[1][r] * t[1] + b[2][r] * t[2]));
    g->i0 = make_float4(ic[0][0], ic[0][1], ic[0][2], ic[1][0]);
    g->i1 = make_float4(ic[1][1], ic[1][2], ic[2][0], ic[2][1]);
    g->i2 = make_float4(ic[2][2], ic[3][0], ic[3][1], ic[3][2]);
    g->f0 = make_float4(m[0], m[1], m[2], m[4]);
    g->f1 = make_float4(m[5], m[6], m[8], m[9]);
    g->f2 = make_float4(m[10], __builtin_bit_cast(float, blas_ref), __builtin_bit_cast(float, material),
                        __builtin_bit_cast(float, blas4_ref));
    return true;
}

// One four-wide node from BVH2 node r (n2[r - off] holds it): starting from
// its two children, while fewer than four are held, the internal one of
// largest surface area (the first on ties; area in double, no contraction) is
// replaced in place by its two children, so the children keep the BVH2's
// left-to-right order. Returns the number of children (2..4); internal
// children keep their BVH2 references, leaves their leaf references. Shared
// by the host collapse of every BLAS and TLAS and the device TLAS collapse.
struct Child4 { float lo[3], hi[3]; int ref; };
__host__ __device__ __forceinline__ int expand4(const GNode* n2, int off, int r, Child4 ch[4]) {
    auto kids = [&](int q, Child4& l, Child4& rr) {
        const GNode g = n2[q - off];
        l.lo[0] = g.a.x; l.lo[1] = g.a.y; l.lo[2] = g.a.z; l.hi[0] = g.a.w; l.hi[1] = g.b.x; l.hi[2] = g.b.y;
        l.ref = g.d.x;
        rr.lo[0] = g.b.z; rr.lo[1] = g.b.w; rr.lo[2] = g.c.x; rr.hi[0] = g.c.y; rr.hi[1] = g.c.z; rr.hi[2] = g.c.w;
        rr.ref = g.d.y;
    };
    auto area = [](const Child4& c) {
        const double dx = double(c.hi[0]) - c.lo[0], dy = double(c.hi[1]) - c.lo[1], dz = double(c.hi[2]) - c.lo[2];
        return dx * dy + dy * dz + dz * dx;
    };
    kids(r, ch[0], ch[1]);
    int k = 2;
    while (k < 4) {
        int best = -1;
        double ba = -1.0;
        for (int j = 0; j < k; ++j)
            if (ch[j].ref >= 0 && area(ch[j]) > ba) { ba = area(ch[j]); best = j; }
        if (best < 0) break;
        Child4 l, rr;
        kids(ch[best].ref, l, rr);
        for (int j = k; j > best + 1; --j) ch[j] = ch[j - 1];
        ch[best] = l;
        ch[best + 1] = rr;
        ++k;
    }
    return k;
}

// GNode4 of up to four children; refs[] are the references written (internal
// children already renumbered), EMPTY4 past count.
__host__ __device__ __forceinline__ GNode4 make_node4(const Child4 ch[4], int count, const int refs[4]) {
    float lo[3][4], hi[3][4];
    int rf[4];
    for (int k = 0; k < 4; ++k) {
        for (int a = 0; a < 3; ++a) {
            lo[a][k] = k < count ? ch[k].lo[a] : 0.f;
            hi[a][k] = k < count ? ch[k].hi[a] : 0.f;
        }
        rf[k] = k < count ? refs[k] : EMPTY4;
    }
    GNode4 g;
    g.lox = make_float4(lo[0][0], lo[0][1], lo[0][2], lo[0][3]);
    g.hix = make_float4(hi[0][0], hi[0][1], hi[0][2], hi[0][3]);
    g.loy = make_float4(lo[1][0], lo[1][1], lo[1][2], lo[1][3]);
    g.hiy = make_float4(hi[1][0], hi[1][1], hi[1][2], hi[1][3]);
    g.loz = make_float4(lo[2][0], lo[2][1], lo[2][2], lo[2][3]);
    g.hiz = make_float4(hi[2][0], hi[2][1], hi[2][2], hi[2][3]);
    g.ref = make_int4(rf[0], rf[1], rf[2], rf[3]);
    g.pad = make_int4(0, 0, 0, 0);
    return g;
}

// Segment table of the build: one segment per node of the level under
// construction (elements [start, start + count) of the order array), as word
// arrays of h = n / 2 + 1 entries each, in this order:
//   start[2], count[2], node[2]  (two buffers: current and next level)
//   cb[6 h]     centroid bounds, min xyz, max xyz per segment (read when the
//               level starts, written for the next level's segments later)
//   border      split plane (float)
//   info        mode | axis << 2 | near2far << 4; mode 0 middle split without
//               partition, 1 all right, 2 all left, 3 partitioned
//   split       partition split (modes 1..3)
//   csplit      first element of the right child
//   m           swapped pairs (mode 3)
//   child[2]    next-level segment of each child, -1 for a leaf
constexpr int kTlasSegWords = 19;

// Buffers of one rebuild over n instances (device pointers).
struct TlasBuild {
    int n;                  // instances
    int base;               // node index of the TLAS root (n >= 2)
    const float* xf;        // [n * 16] instance transforms, column-major
    const float* local;     // [n * 6] object-space box of each instance's BLAS: min xyz, max xyz
    const int* blas_ref;    // [n] BLAS root reference
    const int* blas4_ref;   // [n] four-wide BLAS root reference
    const int* mat;         // [n] material id
    GInst* inst;            // [n] out: instance records
    GNode* nodes;           // out: nodes [base, base + n - 1)
    float* box;             // [n * 6] world boxes
    float* cen;             // [n * 3] box centres
    int* ord0;              // [n] element order (two buffers, swapped per level)
    int* ord1;
    int* pre;               // [n + 1] exclusive prefix count of left-going flags
    int* apos;              // [n] positions of the swapped pairs, per node
    int* bpos;
    int* pseg;              // [n] level segment of each position, -1 when in a leaf
    int* seg;               // [(n / 2 + 1) * kTlasSegWords] segment table
    unsigned long long* keys;  // [(n / 2 + 1) * 24] child box reduction keys per segment
    int* negz;              // [1] set when a world box or centre coordinate is -0.0
    int lds_bytes;          // LDS one workgroup may use on this device
    int arr_lds;            // set by the launcher: order, prefix, pair and segment arrays in LDS
    int key_lds;            // set by the launcher: 32-bit keys kept in LDS too
};

}  // namespace ptd

namespace ptk {
// Instance records, world boxes and the TLAS nodes of a rebuild (a.n >= 1).
void launch_tlas_rebuild(hipStream_t s, const ptd::TlasBuild& a);
// Four-wide TLAS (room entries at out, references counted from base) of the
// BVH2 TLAS rooted at root2 >= 0 with at most room internal nodes; qa and qb
// are room-entry scratch arrays.
void launch_tlas4(hipStream_t s, const ptd::GNode* nodes, int root2, ptd::GNode4* out, int base, int room, int* qa,
                  int* qb);
// Record sizes as the rebuild's unit compiled them (compared with the host
// code's at pt_create).
struct TlasLayout { size_t build, inst, node; };
TlasLayout tlas_layout();
inline TlasLayout tlas_layout_here() { return TlasLayout{sizeof(ptd::TlasBuild), sizeof(ptd::GInst), sizeof(ptd::GNode)}; }
}  // namespace ptk
