// Scene surface, BLAS/TLAS orchestration and flattening.
// Reference behaviour followed (file:line):
//   Scene::Add*                src/core/Scene.cpp:56-146
//   instance boxes (createTLAS) src/core/Scene.cpp:148-187
//   ProcessScene flattening    src/core/Scene.cpp:216-297
//   BVH flattening             thirdparty/RadeonRays/bvh_translator.cpp:32-137
//   Camera basis               src/core/Camera.cpp:120-188
//   OBJ triangle soup          src/core/Mesh.cpp:44-119
//   env-map CDF                src/core/EnvironmentMap.cpp:39-61
#include "scene.h"

#include "image_io.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>

namespace glslpt {

namespace {
inline float Degrees(float r) { return r * (180.f / kPi); }
inline float Radians(float d) { return d * (kPi / 180.f); }

BBox TriBox(const Vec4& a, const Vec4& b, const Vec4& c) {
    BBox bb;
    bb.grow(Vec3(a.x, a.y, a.z));
    bb.grow(Vec3(b.x, b.y, b.z));
    bb.grow(Vec3(c.x, c.y, c.z));
    return bb;
}
}  // namespace

// ---------------------------------------------------------------- Mesh

bool Mesh::LoadFromArrays(const float* pos, const float* nrm, const float* uvs, int num_tris) {
    if (!pos || num_tris <= 0) return false;
    verticesUVX.resize(3 * static_cast<size_t>(num_tris));
    normalsUVY.resize(3 * static_cast<size_t>(num_tris));
    for (int t = 0; t < num_tris; ++t) {
        Vec3 p[3];
        for (int k = 0; k < 3; ++k) p[k] = Vec3(pos[9 * t + 3 * k], pos[9 * t + 3 * k + 1], pos[9 * t + 3 * k + 2]);
        Vec3 fn;
        if (!nrm) fn = Vec3::Normalize(Vec3::Cross(p[1] - p[0], p[2] - p[0]));
        for (int k = 0; k < 3; ++k) {
            float u, v;
            if (uvs) { u = uvs[6 * t + 2 * k]; v = uvs[6 * t + 2 * k + 1]; }
            else { u = (k == 0) ? 0.f : (k == 1 ? 0.f : 1.f); v = (k == 0) ? 0.f : 1.f; }
            Vec3 n = nrm ? Vec3(nrm[9 * t + 3 * k], nrm[9 * t + 3 * k + 1], nrm[9 * t + 3 * k + 2]) : fn;
            verticesUVX[3 * t + k] = Vec4(p[k].x, p[k].y, p[k].z, u);
            normalsUVY[3 * t + k] = Vec4(n.x, n.y, n.z, v);
        }
    }
    return true;
}

bool Mesh::LoadFromFile(const std::string& filename) {
    std::ifstream in(filename);
    if (!in) return false;
    name = filename;
    std::vector<Vec3> P, N;
    std::vector<float> T;  // pairs
    struct Corner { int p, t, n; };
    std::vector<Corner> tris;
    auto fix = [](int idx, int count) { return idx < 0 ? count + idx : idx - 1; };
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ss(line);
        std::string tag;
        if (!(ss >> tag)) continue;
        if (tag == "v") { Vec3 v; ss >> v.x >> v.y >> v.z; P.push_back(v); }
        else if (tag == "vn") { Vec3 v; ss >> v.x >> v.y >> v.z; N.push_back(v); }
        else if (tag == "vt") { float a = 0, b = 0; ss >> a >> b; T.push_back(a); T.push_back(b); }
        else if (tag == "f") {
            std::vector<Corner> poly;
            std::string tok;
            while (ss >> tok) {
                Corner c{-1, -1, -1};
                int vals[3] = {0, 0, 0};
                bool have[3] = {false, false, false};
                size_t s = 0;
                for (int k = 0; k < 3 && s <= tok.size(); ++k) {
                    size_t e = tok.find('/', s);
                    std::string part = tok.substr(s, e == std::string::npos ? std::string::npos : e - s);
                    if (!part.empty()) { vals[k] = std::stoi(part); have[k] = true; }
                    if (e == std::string::npos) break;
                    s = e + 1;
                }
                if (!have[0]) return false;
                c.p = fix(vals[0], static_cast<int>(P.size()));
                c.t = have[1] ? fix(vals[1], static_cast<int>(T.size() / 2)) : -1;
                c.n = have[2] ? fix(vals[2], static_cast<int>(N.size())) : -1;
                poly.push_back(c);
            }
            for (size_t k = 2; k < poly.size(); ++k) {  // fan triangulation
                tris.push_back(poly[0]); tris.push_back(poly[k - 1]); tris.push_back(poly[k]);
            }
        }
    }
    if (tris.empty()) return false;
    const size_t nt = tris.size() / 3;
    verticesUVX.resize(tris.size());
    normalsUVY.resize(tris.size());
    for (size_t t = 0; t < nt; ++t) {
        Vec3 p[3];
        for (int k = 0; k < 3; ++k) {
            const Corner& c = tris[3 * t + k];
            if (c.p < 0 || c.p >= static_cast<int>(P.size())) return false;
            p[k] = P[c.p];
        }
        Vec3 fn = Vec3::Normalize(Vec3::Cross(p[1] - p[0], p[2] - p[0]));
        for (int k = 0; k < 3; ++k) {
            const Corner& c = tris[3 * t + k];
            float tx, ty;
            if (!T.empty() && c.t >= 0 && 2 * c.t + 1 < static_cast<int>(T.size())) {
                tx = T[2 * c.t]; ty = 1.0f - T[2 * c.t + 1];
            } else {
                if (k == 0) tx = ty = 0; else if (k == 1) tx = 0, ty = 1; else tx = ty = 1;
            }
            Vec3 n = (c.n >= 0 && c.n < static_cast<int>(N.size())) ? N[c.n] : fn;
            verticesUVX[3 * t + k] = Vec4(p[k].x, p[k].y, p[k].z, tx);
            normalsUVY[3 * t + k] = Vec4(n.x, n.y, n.z, ty);
        }
    }
    return true;
}

void Mesh::BuildBVH() {
    const int n = static_cast<int>(verticesUVX.size() / 3);
    std::vector<BBox> boxes(n);
    for (int i = 0; i < n; ++i)
        boxes[i] = TriBox(verticesUVX[3 * i], verticesUVX[3 * i + 1], verticesUVX[3 * i + 2]);
    bvh = BuildSahBvh(boxes.data(), n, 2.0f, 64);
}

// ---------------------------------------------------------------- Camera
// Trigonometry goes through the C library's double-precision functions with
// float arguments, matching the unqualified calls at src/core/Camera.cpp:124-186.

Camera::Camera(Vec3 eye, Vec3 lookat, float fov_degrees) {
    position = eye;
    pivot = lookat;
    Vec3 dir = Vec3::Normalize(pivot - position);
    pitch = Degrees(static_cast<float>(::asin(static_cast<double>(dir.y))));
    yaw = Degrees(static_cast<float>(::atan2(static_cast<double>(dir.z), static_cast<double>(dir.x))));
    radius = Vec3::Distance(eye, lookat);
    fov = Radians(fov_degrees);
    focalDist = 0.1f;
    aperture = 0.f;
    UpdateCamera();
}

void Camera::UpdateCamera() {
    Vec3 f;
    const double ry = Radians(yaw), rp = Radians(pitch);
    f.x = static_cast<float>(::cos(ry) * ::cos(rp));
    f.y = static_cast<float>(::sin(rp));
    f.z = static_cast<float>(::sin(ry) * ::cos(rp));
    forward = Vec3::Normalize(f);
    position = pivot + (forward * -1.0f) * radius;
    right = Vec3::Normalize(Vec3::Cross(forward, worldUp));
    up = Vec3::Normalize(Vec3::Cross(right, forward));
}

void Camera::OffsetOrientation(float dx, float dy) { pitch -= dy; yaw += dx; UpdateCamera(); }
void Camera::Strafe(float dx, float dy) { pivot = pivot + (right * -dx + up * dy); UpdateCamera(); }
void Camera::SetRadius(float dr) { radius += dr; UpdateCamera(); }
void Camera::SetFov(float degrees) { fov = Radians(degrees); }

// ---------------------------------------------------------------- Scene

Scene::~Scene() {
    for (Mesh* m : meshes) delete m;
    delete camera;
}

int Scene::AddMesh(const std::string& filename) {
    for (size_t i = 0; i < meshes.size(); ++i)
        if (meshes[i]->name == filename) return static_cast<int>(i);
    Mesh* m = new Mesh;
    if (!m->LoadFromFile(filename)) {
        error = "unable to load model " + filename;
        delete m;
        return -1;
    }
    meshes.push_back(m);
    return static_cast<int>(meshes.size() - 1);
}

int Scene::AddMesh(const float* positions, const float* normals, const float* uvs, int num_tris,
                   const std::string& name) {
    if (!name.empty())
        for (size_t i = 0; i < meshes.size(); ++i)
            if (meshes[i]->name == name) return static_cast<int>(i);
    Mesh* m = new Mesh;
    m->name = name;
    if (!m->LoadFromArrays(positions, normals, uvs, num_tris)) {
        error = "invalid mesh arrays";
        delete m;
        return -1;
    }
    meshes.push_back(m);
    return static_cast<int>(meshes.size() - 1);
}

int Scene::AddMaterial(const Material& material) {
    materials.push_back(material);
    return static_cast<int>(materials.size() - 1);
}

int Scene::AddMeshInstance(const MeshInstance& instance) {
    meshInstances.push_back(instance);
    return static_cast<int>(meshInstances.size() - 1);
}

int Scene::AddLight(const Light& light) {
    lights.push_back(light);
    return static_cast<int>(lights.size() - 1);
}

int Scene::AddTexture(const uint8_t* rgba, int width, int height, const std::string& name) {
    for (size_t i = 0; i < textures.size(); ++i)
        if (!name.empty() && textures[i].name == name) return static_cast<int>(i);
    if (!rgba || width <= 0 || height <= 0) { error = "invalid texture"; return -1; }
    Texture t{name, width, height, std::vector<uint8_t>(rgba, rgba + 4 * static_cast<size_t>(width) * height)};
    textures.push_back(std::move(t));
    return static_cast<int>(textures.size() - 1);
}

int Scene::AddTexture(const std::string& filename) {
    for (size_t i = 0; i < textures.size(); ++i)
        if (textures[i].name == filename) return static_cast<int>(i);
    int w = 0, h = 0;
    std::vector<uint8_t> rgba;
    std::string err;
    if (!LoadPNG(filename, w, h, rgba, err)) {
        error = "unable to load texture " + filename + ": " + err;
        return -1;
    }
    textures.push_back(Texture{filename, w, h, std::move(rgba)});
    return static_cast<int>(textures.size() - 1);
}

void Scene::AddEnvMap(const std::string& filename) {
    int w = 0, h = 0;
    std::vector<float> rgb;
    std::string err;
    if (LoadHDR(filename, w, h, rgb, err)) {
        AddEnvMap(rgb.data(), w, h);
        return;
    }
    error = "unable to load HDR " + filename + ": " + err;
    envImg.clear();
    envCdf.clear();
    envWidth = envHeight = 0;
    envTotalSum = 0.f;
    envMapModified = true;
    dirty = true;
}

void Scene::AddCamera(Vec3 eye, Vec3 lookat, float fov_degrees) {
    delete camera;
    camera = new Camera(eye, lookat, fov_degrees);
}

void Scene::AddEnvMap(const float* rgb, int width, int height) {
    envWidth = width; envHeight = height;
    envImg.assign(rgb, rgb + 3 * static_cast<size_t>(width) * height);
    envCdf.resize(static_cast<size_t>(width) * height);
    // Single running sum over the row-major image.
    float run = 0.f;
    for (size_t i = 0; i < envCdf.size(); ++i) {
        float w = 0.212671f * envImg[3 * i] + 0.715160f * envImg[3 * i + 1] + 0.072169f * envImg[3 * i + 2];
        run = (i == 0) ? w : run + w;
        envCdf[i] = run;
    }
    envTotalSum = envCdf.empty() ? 0.f : envCdf.back();
    envMapModified = true;
    dirty = true;
}

void Scene::BuildTLAS() {
    std::vector<BBox> boxes(meshInstances.size());
    for (size_t i = 0; i < meshInstances.size(); ++i) {
        const BBox& b = meshes[meshInstances[i].meshID]->bvh.bounds;
        const Mat4& m = meshInstances[i].transform;
        Vec3 right(m[0][0], m[0][1], m[0][2]), up(m[1][0], m[1][1], m[1][2]);
        Vec3 fwd(m[2][0], m[2][1], m[2][2]), tr(m[3][0], m[3][1], m[3][2]);
        Vec3 xa = right * b.pmin.x, xb = right * b.pmax.x;
        Vec3 ya = up * b.pmin.y, yb = up * b.pmax.y;
        Vec3 za = fwd * b.pmin.z, zb = fwd * b.pmax.z;
        boxes[i].pmin = Vec3::Min(xa, xb) + Vec3::Min(ya, yb) + Vec3::Min(za, zb) + tr;
        boxes[i].pmax = Vec3::Max(xa, xb) + Vec3::Max(ya, yb) + Vec3::Max(za, zb) + tr;
    }
    tlas = BuildMedianBvh(boxes.data(), static_cast<int>(boxes.size()));
    sceneBounds = tlas.bounds;
}

namespace {
// Preorder, left child first: a node's left child is the next index, its right
// child follows the whole left subtree.
template <class LeafFn>
void FlattenTree(const BvhTree& t, int base, std::vector<FlatNode>& out, LeafFn leaf_fn) {
    struct Item { int node; int parent; bool left; };
    std::vector<Item> st;
    st.push_back({0, -1, false});
    int cur = base;
    while (!st.empty()) {
        Item it = st.back();
        st.pop_back();
        const int idx = cur++;
        const BvhBuildNode& n = t.nodes[it.node];
        if (it.parent >= 0) {
            if (it.left) out[it.parent].LRLeaf.x = static_cast<float>(idx);
            else out[it.parent].LRLeaf.y = static_cast<float>(idx);
        }
        out[idx].bboxmin = n.bounds.pmin;
        out[idx].bboxmax = n.bounds.pmax;
        out[idx].LRLeaf.z = 0.f;
        if (n.leaf) {
            leaf_fn(n, out[idx]);
        } else {
            st.push_back({n.rc, idx, false});
            st.push_back({n.lc, idx, true});
        }
    }
}
}  // namespace

void Scene::FlattenTLAS() {
    FlattenTree(tlas, topLevelIndex, nodes, [&](const BvhBuildNode& n, FlatNode& f) {
        const int inst = tlas.packed[n.start];
        f.LRLeaf.x = static_cast<float>(blasRootIndex[meshInstances[inst].meshID]);
        f.LRLeaf.y = static_cast<float>(meshInstances[inst].materialID);
        f.LRLeaf.z = static_cast<float>(-inst - 1);
    });
}

bool Scene::ProcessScene() {
    error.clear();
    if (meshes.empty() || meshInstances.empty()) { error = "scene needs at least one mesh and one instance"; return false; }
    for (const MeshInstance& mi : meshInstances) {
        if (mi.meshID < 0 || mi.meshID >= static_cast<int>(meshes.size())) { error = "instance references a missing mesh"; return false; }
        if (mi.materialID < 0 || mi.materialID >= static_cast<int>(materials.size())) { error = "instance references a missing material"; return false; }
    }
    for (Mesh* m : meshes) m->BuildBVH();
    BuildTLAS();

    // Node array: every BLAS (in mesh order), then the TLAS with room for
    // 2 * instances nodes.
    int count = 0;
    blasRootIndex.clear();
    for (Mesh* m : meshes) { blasRootIndex.push_back(count); count += static_cast<int>(m->bvh.nodes.size()); }
    topLevelIndex = count;
    nodes.assign(count + 2 * meshInstances.size(), FlatNode{});
    int tri_base = 0;
    for (size_t i = 0; i < meshes.size(); ++i) {
        FlattenTree(meshes[i]->bvh, blasRootIndex[i], nodes, [&](const BvhBuildNode& n, FlatNode& f) {
            f.LRLeaf.x = static_cast<float>(tri_base + n.start);
            f.LRLeaf.y = static_cast<float>(n.count);
            f.LRLeaf.z = 1.f;
        });
        tri_base += static_cast<int>(meshes[i]->bvh.packed.size());
    }
    FlattenTLAS();

    vertIndices.clear(); verticesUVX.clear(); normalsUVY.clear();
    int vbase = 0;
    for (Mesh* m : meshes) {
        for (int p : m->bvh.packed) vertIndices.push_back(Indices{3 * p + vbase, 3 * p + 1 + vbase, 3 * p + 2 + vbase});
        verticesUVX.insert(verticesUVX.end(), m->verticesUVX.begin(), m->verticesUVX.end());
        normalsUVY.insert(normalsUVY.end(), m->normalsUVY.begin(), m->normalsUVY.end());
        vbase += static_cast<int>(m->verticesUVX.size());
    }
    transforms.resize(meshInstances.size());
    for (size_t i = 0; i < meshInstances.size(); ++i) transforms[i] = meshInstances[i].transform;

    // Texture array: each texture resampled (nearest) to the array size.
    const size_t tex_bytes = 4 * static_cast<size_t>(texArrayWidth) * texArrayHeight;
    textureMapsArray.assign(tex_bytes * textures.size(), 0);
    for (size_t i = 0; i < textures.size(); ++i) {
        const Texture& t = textures[i];
        uint8_t* dst = &textureMapsArray[i * tex_bytes];
        for (int y = 0; y < texArrayHeight; ++y)
            for (int x = 0; x < texArrayWidth; ++x) {
                int sx = static_cast<int>((static_cast<long long>(x) * t.width) / texArrayWidth);
                int sy = static_cast<int>((static_cast<long long>(y) * t.height) / texArrayHeight);
                std::memcpy(dst + 4 * (static_cast<size_t>(y) * texArrayWidth + x),
                            &t.rgba[4 * (static_cast<size_t>(sy) * t.width + sx)], 4);
            }
    }

    if (!camera) {
        Vec3 ext = sceneBounds.extents(), c = sceneBounds.center();
        AddCamera(Vec3(c.x, c.y, c.z + Vec3::Length(ext) * 2.0f), c, 45.0f);
    }
    initialized = true;
    return true;
}

bool Scene::RebuildInstances() {
    if (!initialized) { error = "ProcessScene has not run"; return false; }
    BuildTLAS();
    FlattenTLAS();
    for (size_t i = 0; i < meshInstances.size(); ++i) transforms[i] = meshInstances[i].transform;
    instancesModified = true;
    dirty = true;
    return true;
}

}  // namespace glslpt
