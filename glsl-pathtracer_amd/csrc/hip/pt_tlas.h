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
// Record sizes as the rebuild's unit compiled them (compared with the host
// code's at pt_create).
struct TlasLayout { size_t build, inst, node; };
TlasLayout tlas_layout();
inline TlasLayout tlas_layout_here() { return TlasLayout{sizeof(ptd::TlasBuild), sizeof(ptd::GInst), sizeof(ptd::GNode)}; }
}  // namespace ptk
