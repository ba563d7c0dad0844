// Host-side conversion of the reference's flat scene arrays into the GPU
// layout (BVH2 nodes, instance and light records) and of the RenderOptions
// subset into the kernels' option word. Used by the renderer's uploads and by
// host builds of the device path.
//
// Reference layouts (file:line in knightcrawler25/GLSL-PathTracer):
//   flat node array     thirdparty/RadeonRays/bvh_translator.cpp:32-137
//   Light record        src/core/Scene.h:50-59
//   OPT_* switches      src/core/Renderer.cpp:400-459
#pragma once

#include <algorithm>
#include <cmath>
#include <exception>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../../include/glslpt.h"
#include "pt_device.h"
#include "pt_tlas.h"

namespace ptl {

using ptd::DevOpts;
using ptd::GInst;
using ptd::GLight;
using ptd::GNode;

// Invalid or unsupported input: a pt_* status code and a message.
struct ArgError : std::exception {
    std::string msg;
    int code;
    ArgError(int c, std::string m) : msg(std::move(m)), code(c) {}
    const char* what() const noexcept override { return msg.c_str(); }
};

inline GInst make_inst(const float* m, int blas_ref, int material, int blas4_ref) {
    GInst g;
    if (!ptd::inst_record(m, blas_ref, material, &g, blas4_ref))
        throw ArgError(PT_ERR_INVALID_ARG, "singular instance transform");
    return g;
}

// Four-wide BLAS nodes from the BVH2 re-layout, each from its BVH2 node by
// ptd::expand4. A node's internal children get consecutive indices when the
// node is converted (depth first). root4 maps each BLAS root's BVH2 reference
// to its four-wide one.
struct Wide4 {
    std::vector<ptd::GNode4> nodes;
    std::unordered_map<int, int> root4;
};
inline void collapse4_root(const std::vector<GNode>& n2, int root, Wide4& W) {
    if (W.root4.count(root)) return;
    if (root < 0) {  // a BLAS that is one leaf
        W.root4[root] = root;
        return;
    }
    struct Job { int ref2, slot; };  // BVH2 node to convert, its GNode4 index
    std::vector<Job> todo;
    auto alloc = [&](int ref2) {
        const int idx = static_cast<int>(W.nodes.size());
        W.nodes.emplace_back();
        todo.push_back({ref2, idx});
        return idx;
    };
    W.root4[root] = alloc(root);
    while (!todo.empty()) {
        const Job jb = todo.back();
        todo.pop_back();
        ptd::Child4 ch[4];
        const int k = ptd::expand4(n2.data(), 0, jb.ref2, ch);
        int ref[4];
        for (int j = 0; j < k; ++j) ref[j] = ch[j].ref;
        // the first child is converted next (depth first)
        for (int j = k - 1; j >= 0; --j)
            if (ref[j] >= 0) ref[j] = alloc(ref[j]);
        W.nodes[jb.slot] = ptd::make_node4(ch, k, ref);
    }
}

// Four-wide TLAS from the BVH2 TLAS (n2[r - off] holds node r, root2 the
// root), appended to out in breadth-first order (the root first, then each
// level's internal children in order), references counted from base (the
// index of the first appended entry), then zero entries up to `room` in all.
// The device collapse after a TLAS rebuild (k_tlas4) writes the same array.
// Returns the four-wide root reference (root2 itself when it is a leaf).
inline int collapse4_tlas(const GNode* n2, int off, int root2, int base, size_t room, std::vector<ptd::GNode4>& out) {
    const size_t first = out.size();
    int root4 = root2;
    if (root2 >= 0) {
        root4 = base;
        std::vector<int> fifo{root2};
        for (size_t q = 0; q < fifo.size(); ++q) {
            ptd::Child4 ch[4];
            const int k = ptd::expand4(n2, off, fifo[q], ch);
            int ref[4];
            for (int j = 0; j < k; ++j) {
                ref[j] = ch[j].ref;
                if (ref[j] >= 0) {
                    ref[j] = base + static_cast<int>(fifo.size());
                    fifo.push_back(ch[j].ref);
                }
            }
            out.push_back(ptd::make_node4(ch, k, ref));
        }
    }
    if (out.size() - first < room) out.resize(first + room, ptd::GNode4{});
    return root4;
}

// Re-layout of the reference's flat node array (bvh_translator.cpp:32-137)
// into BVH2 nodes. Only nodes reachable from the TLAS root are converted;
// GNode order follows the flat (preorder) order, so subtrees stay contiguous.
struct Relayout {
    std::vector<GNode> nodes;
    int root = 0;
    std::vector<int> inst_blas_ref, inst_material;
    std::vector<float> inst_local;  // box of each instance's BLAS root: min xyz, max xyz
    std::unordered_map<int, int> root_ref;  // flat index of each BLAS root a TLAS leaf names -> its reference
    Wide4 wide;                             // four-wide BLAS nodes and their roots, then the four-wide TLAS
    std::vector<int> inst_blas4_ref;        // each instance's four-wide BLAS root reference
    int tlas4_base = 0;                     // first four-wide TLAS entry in wide.nodes (at least max(n - 1, 1) follow)
    int root4 = 0;                          // four-wide TLAS root reference
    // The TLAS as the device rebuild writes it: one instance per leaf, every
    // instance once, internal nodes numbered in preorder from tlas_base to the
    // end of the node array.
    bool tlas_rebuildable = false;
    int tlas_base = 0;
};

inline Relayout relayout(const pt_bvh_node* flat, int num_nodes, int top, int num_tris, int num_instances) {
    Relayout R;
    R.inst_blas_ref.assign(num_instances, 0);
    R.inst_material.assign(num_instances, 0);
    R.inst_local.assign(6 * static_cast<size_t>(num_instances), 0.f);
    auto kind = [&](int i) {
        float z = flat[i].LRLeaf[2];
        return z > 0.f ? 1 : (z < 0.f ? -1 : 0);
    };
    auto as_int = [](float f) { return static_cast<int>(f); };
    std::vector<int> map(num_nodes, -1);
    std::vector<char> reach(num_nodes, 0);
    // mark reachable internal nodes; TLAS first, then each BLAS root
    std::vector<int> st;
    if (top < 0 || top >= num_nodes) throw ArgError(PT_ERR_INVALID_ARG, "top_level_index out of range");
    st.push_back(top);
    std::vector<char> blas_seen(num_nodes, 0);
    while (!st.empty()) {
        int i = st.back();
        st.pop_back();
        if (i < 0 || i >= num_nodes) throw ArgError(PT_ERR_INVALID_ARG, "node child index out of range");
        if (reach[i]) throw ArgError(PT_ERR_INVALID_ARG, "node reachable twice (not a tree)");
        reach[i] = 1;
        int k = kind(i);
        if (k == 0) {
            st.push_back(as_int(flat[i].LRLeaf[0]));
            st.push_back(as_int(flat[i].LRLeaf[1]));
        } else if (k < 0) {
            int inst = -as_int(flat[i].LRLeaf[2]) - 1;
            if (inst < 0 || inst >= num_instances) throw ArgError(PT_ERR_INVALID_ARG, "TLAS leaf instance out of range");
            int root = as_int(flat[i].LRLeaf[0]);
            if (root < 0 || root >= num_nodes) throw ArgError(PT_ERR_INVALID_ARG, "BLAS root out of range");
            if (!blas_seen[root]) {
                blas_seen[root] = 1;
                if (reach[root]) throw ArgError(PT_ERR_INVALID_ARG, "BLAS root inside the TLAS");
                st.push_back(root);
            }
        } else {
            int first = as_int(flat[i].LRLeaf[0]), count = as_int(flat[i].LRLeaf[1]);
            if (count < 1 || count > 3) throw ArgError(PT_ERR_UNSUPPORTED, "BLAS leaves must hold 1..3 triangles");
            if (first < 0 || first + count > num_tris) throw ArgError(PT_ERR_INVALID_ARG, "BLAS leaf triangle range out of bounds");
            if (first >= (1 << 28)) throw ArgError(PT_ERR_UNSUPPORTED, "too many triangles for the leaf encoding");
        }
    }
    int n = 0;
    for (int i = 0; i < num_nodes; ++i)
        if (reach[i] && kind(i) == 0) map[i] = n++;
    R.nodes.resize(std::max(n, 1));
    auto ref_of = [&](int i) -> int {
        int k = kind(i);
        if (k == 0) return map[i];
        if (k < 0) return -1 - (-as_int(flat[i].LRLeaf[2]) - 1);
        int first = as_int(flat[i].LRLeaf[0]), count = as_int(flat[i].LRLeaf[1]);
        return -1 - ((first << 2) | count);
    };
    for (int i = 0; i < num_nodes; ++i) {
        if (!reach[i]) continue;
        int k = kind(i);
        if (k == 0) {
            int l = as_int(flat[i].LRLeaf[0]), r = as_int(flat[i].LRLeaf[1]);
            const float* lmn = flat[l].bboxmin; const float* lmx = flat[l].bboxmax;
            const float* rmn = flat[r].bboxmin; const float* rmx = flat[r].bboxmax;
            GNode& g = R.nodes[map[i]];
            g.a = make_float4(lmn[0], lmn[1], lmn[2], lmx[0]);
            g.b = make_float4(lmx[1], lmx[2], rmn[0], rmn[1]);
            g.c = make_float4(rmn[2], rmx[0], rmx[1], rmx[2]);
            g.d = make_int4(ref_of(l), ref_of(r), 0, 0);
        } else if (k < 0) {
            int inst = -as_int(flat[i].LRLeaf[2]) - 1;
            const int root = as_int(flat[i].LRLeaf[0]);
            R.inst_blas_ref[inst] = ref_of(root);
            R.root_ref[root] = R.inst_blas_ref[inst];
            R.inst_material[inst] = as_int(flat[i].LRLeaf[1]);
            for (int j = 0; j < 3; ++j) {
                R.inst_local[6 * inst + j] = flat[root].bboxmin[j];
                R.inst_local[6 * inst + 3 + j] = flat[root].bboxmax[j];
            }
        }
    }
    R.root = ref_of(top);
    {
        bool ok = true;
        int internal = 0, leaves = 0;
        const int base = kind(top) == 0 ? map[top] : n;
        std::vector<char> inst_seen(num_instances, 0);
        st.assign(1, top);
        while (!st.empty() && ok) {
            const int i = st.back();
            st.pop_back();
            const int k = kind(i);
            if (k == 0) {
                ok = map[i] == base + internal;
                ++internal;
                st.push_back(as_int(flat[i].LRLeaf[1]));
                st.push_back(as_int(flat[i].LRLeaf[0]));
            } else if (k < 0) {
                const int inst = -as_int(flat[i].LRLeaf[2]) - 1;
                ok = !inst_seen[inst];
                inst_seen[inst] = 1;
                ++leaves;
            } else {
                ok = false;
            }
        }
        R.tlas_rebuildable = ok && leaves == num_instances && internal == num_instances - 1 && base + internal == n;
        R.tlas_base = base;
    }
    R.inst_blas4_ref.assign(num_instances, 0);
    for (int i = 0; i < num_instances; ++i) {
        collapse4_root(R.nodes, R.inst_blas_ref[i], R.wide);
        R.inst_blas4_ref[i] = R.wide.root4[R.inst_blas_ref[i]];
    }
    // The four-wide TLAS after the BLASes, room for n - 1 nodes (what any
    // TLAS over n instances collapses to at most; a device rebuild may change
    // the shape).
    R.tlas4_base = static_cast<int>(R.wide.nodes.size());
    R.root4 = collapse4_tlas(R.nodes.data(), 0, R.root, R.tlas4_base, std::max(num_instances - 1, 1), R.wide.nodes);
    return R;
}

// TLAS-only re-layout for an update of the TLAS node range [top, nflat) with
// the BLASes unchanged (the reference uploads exactly that range,
// src/core/Renderer.cpp:659-663): the n - 1 internal nodes, numbered in
// preorder from base, plus each instance's BLAS root reference and material.
// Returns false when the new range is not a one-instance-per-leaf preorder
// tree or names a BLAS root that the last full re-layout did not convert;
// the caller then re-lays the whole array out.
struct TlasRelayout {
    std::vector<GNode> nodes;
    std::vector<int> blas, mat, roots;
};
inline bool relayout_tlas(const pt_bvh_node* tl, int top, int nflat, int n, int base,
                   const std::unordered_map<int, int>& root_ref, TlasRelayout& T) {
    auto at = [&](int i) -> const pt_bvh_node& { return tl[i - top]; };
    auto kind = [&](int i) { const float z = at(i).LRLeaf[2]; return z > 0.f ? 1 : (z < 0.f ? -1 : 0); };
    auto as_int = [](float f) { return static_cast<int>(f); };
    T.nodes.assign(std::max(n - 1, 0), GNode{});
    T.blas.assign(n, 0);
    T.mat.assign(n, 0);
    T.roots.assign(n, -1);
    std::vector<char> seen(n, 0);
    struct Item { int i, parent, side; };
    std::vector<Item> st{{top, -1, 0}};
    int internal = 0, leaves = 0;
    while (!st.empty()) {
        const Item it = st.back();
        st.pop_back();
        if (it.i < top || it.i >= nflat) return false;
        int ref;
        const int k = kind(it.i);
        if (k == 0) {
            if (internal >= n - 1) return false;
            const int idx = internal++;
            ref = base + idx;
            const int l = as_int(at(it.i).LRLeaf[0]), r = as_int(at(it.i).LRLeaf[1]);
            if (l < top || l >= nflat || r < top || r >= nflat) return false;
            const float* lmn = at(l).bboxmin; const float* lmx = at(l).bboxmax;
            const float* rmn = at(r).bboxmin; const float* rmx = at(r).bboxmax;
            GNode& g = T.nodes[idx];
            g.a = make_float4(lmn[0], lmn[1], lmn[2], lmx[0]);
            g.b = make_float4(lmx[1], lmx[2], rmn[0], rmn[1]);
            g.c = make_float4(rmn[2], rmx[0], rmx[1], rmx[2]);
            g.d = make_int4(0, 0, 0, 0);
            st.push_back({r, idx, 1});
            st.push_back({l, idx, 0});
        } else if (k < 0) {
            const int inst = -as_int(at(it.i).LRLeaf[2]) - 1;
            if (inst < 0 || inst >= n || seen[inst]) return false;
            seen[inst] = 1;
            ++leaves;
            const int root = as_int(at(it.i).LRLeaf[0]);
            const auto f = root_ref.find(root);
            if (f == root_ref.end()) return false;
            T.blas[inst] = f->second;
            T.mat[inst] = as_int(at(it.i).LRLeaf[1]);
            T.roots[inst] = root;
            ref = -1 - inst;
        } else {
            return false;
        }
        if (it.parent >= 0) {
            if (it.side == 0) T.nodes[it.parent].d.x = ref;
            else T.nodes[it.parent].d.y = ref;
        }
    }
    return internal == n - 1 && leaves == n;
}

inline GLight make_light(const pt_light& l) {
    GLight g{};
    g.p = make_float4(l.position[0], l.position[1], l.position[2], l.radius);
    g.e = make_float4(l.emission[0], l.emission[1], l.emission[2], l.area);
    g.u = make_float4(l.u[0], l.u[1], l.u[2], l.type);
    g.v = make_float4(l.v[0], l.v[1], l.v[2], 0.f);
    // normalize(cross(u, v)), plane w, u / |u|^2, v / |v|^2 (closest_hit.glsl:50-54),
    // in the same float operations the shader performs.
    const float* u = l.u; const float* v = l.v;
    float cx = u[1] * v[2] - u[2] * v[1], cy = u[2] * v[0] - u[0] * v[2], cz = u[0] * v[1] - u[1] * v[0];
    float len = std::sqrt(cx * cx + cy * cy + cz * cz);
    float nx = cx / len, ny = cy / len, nz = cz / len;
    float w = nx * l.position[0] + ny * l.position[1] + nz * l.position[2];
    g.n = make_float4(nx, ny, nz, w);
    float iu = 1.0f / (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    float iv = 1.0f / (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    g.uu = make_float4(u[0] * iu, u[1] * iu, u[2] * iu, 0.f);
    g.vv = make_float4(v[0] * iv, v[1] * iv, v[2] * iv, 0.f);
    return g;
}

// Triangle records in vertIndices row order (= BLAS leaf order): v0 and the
// two edges the shader's intersection test forms (closest_hit.glsl:124-129),
// plus the rows themselves for the hit attributes. Indices are checked by
// the caller.
inline void make_tris(const pt_scene_view& v, std::vector<ptd::GTri>& tris, std::vector<int4>& vidx) {
    tris.resize(v.num_tris);
    vidx.resize(v.num_tris);
    const float* P = v.vertices_uvx;
    for (int r = 0; r < v.num_tris; ++r) {
        const int* vi = v.vert_indices + 3 * r;
        const float* a = P + 4 * static_cast<size_t>(vi[0]);
        const float* b = P + 4 * static_cast<size_t>(vi[1]);
        const float* c = P + 4 * static_cast<size_t>(vi[2]);
        tris[r].v0 = make_float4(a[0], a[1], a[2], 0.f);
        tris[r].e0 = make_float4(b[0] - a[0], b[1] - a[1], b[2] - a[2], 0.f);
        tris[r].e1 = make_float4(c[0] - a[0], c[1] - a[1], c[2] - a[2], 0.f);
        vidx[r] = make_int4(vi[0], vi[1], vi[2], 0);
    }
}

// The kernels' option word and uniforms (InitShaders' OPT_* defines and the
// uniforms of Renderer::Update, src/core/Renderer.cpp:400-459, :764-811).
inline DevOpts make_dev_opts(const pt_options& opts, const pt_camera& cam, int num_lights, bool has_env,
                             bool any_alpha, bool any_medium) {
    DevOpts o{};
    uint32_t f = 0;
    if (num_lights > 0) f |= ptd::OPT_LIGHTS;
    if (opts.enable_rr) f |= ptd::OPT_RR;
    if (opts.enable_uniform_light) f |= ptd::OPT_UNIFORM_LIGHT;
    if (opts.enable_envmap && has_env) f |= ptd::OPT_ENVMAP;
    if (opts.hide_emitters) f |= ptd::OPT_HIDE_EMITTERS;
    if (opts.enable_background) f |= ptd::OPT_BACKGROUND;
    if (opts.transparent_background) f |= ptd::OPT_TRANSPARENT_BACKGROUND;
    if (any_alpha) f |= ptd::OPT_ALPHA_TEST;
    if (any_medium) f |= ptd::OPT_MEDIUM;
    if (opts.enable_volume_mis) f |= ptd::OPT_VOL_MIS;
    if (opts.enable_roughness_mollification) f |= ptd::OPT_ROUGHNESS_MOLLIFICATION;
    if (opts.opengl_normal_map) f |= ptd::OPT_OPENGL_NORMALMAP;
    o.flags = f;
    o.width = opts.width;
    o.height = opts.height;
    o.max_depth = opts.max_depth;
    o.rr_depth = opts.rr_depth;
    o.ucol_x = opts.uniform_light_col[0];
    o.ucol_y = opts.uniform_light_col[1];
    o.ucol_z = opts.uniform_light_col[2];
    o.env_intensity = opts.envmap_intensity;
    o.env_rot = opts.envmap_rot / 360.0f;
    o.moll_amt = opts.roughness_mollification_amt;
    for (int i = 0; i < 3; ++i) {
        o.cam_pos[i] = cam.position[i];
        o.cam_up[i] = cam.up[i];
        o.cam_right[i] = cam.right[i];
        o.cam_fwd[i] = cam.forward[i];
    }
    o.cam_fov = cam.fov;
    o.cam_focal = cam.focalDist;
    o.cam_aperture = cam.aperture;
    return o;
}

}  // namespace ptl
