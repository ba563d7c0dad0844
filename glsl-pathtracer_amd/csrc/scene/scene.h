// Host scene surface: the reference's Scene / Mesh / MeshInstance / Camera
// (src/core/Scene.h:74-125, src/core/Mesh.h:30-70, src/core/Camera.h:30-65),
// re-implemented for a HIP backend. It keeps the reference's method names and
// the flat array layouts that the GPU side consumes
// (src/core/Renderer.cpp:135-249), so a caller written against the reference
// Scene builds the same arrays here.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "bvh_build.h"
#include "pt_math.h"

namespace glslpt {

// 32 floats, the GPU record of one material (src/core/Material.h:47-120).
struct Material {
    Vec3 baseColor{1.f, 1.f, 1.f}; float anisotropic = 0.f;
    Vec3 emission{0.f, 0.f, 0.f};  float padding1 = 0.f;
    float metallic = 0.f, roughness = 0.5f, subsurface = 0.f, specularTint = 0.f;
    float sheen = 0.f, sheenTint = 0.f, clearcoat = 0.f, clearcoatGloss = 0.f;
    float specTrans = 0.f, ior = 1.5f, mediumType = 0.f, mediumDensity = 0.f;
    Vec3 mediumColor{1.f, 1.f, 1.f}; float mediumAnisotropy = 0.f;
    float baseColorTexId = -1.f, metallicRoughnessTexID = -1.f, normalmapTexID = -1.f,
          emissionmapTexID = -1.f;
    float opacity = 1.f, alphaMode = 0.f, alphaCutoff = 0.f, padding2 = 0.f;
};
static_assert(sizeof(Material) == 128, "material record is 8 x vec4");

enum LightType { RectLight = 0, SphereLight = 1, DistantLight = 2 };

// 15 floats (src/core/Scene.h:50-59).
struct Light {
    Vec3 position, emission, u, v;
    float radius = 0.f, area = 0.f, type = 0.f;
};
static_assert(sizeof(Light) == 60, "light record is 5 x vec3");

struct Indices { int x, y, z; };

// Flattened BVH node (thirdparty/RadeonRays/bvh_translator.h:45-50): child or
// leaf data are stored as floats, exact below 2^24.
struct FlatNode { Vec3 bboxmin, bboxmax, LRLeaf; };
static_assert(sizeof(FlatNode) == 36, "flat node is 3 x vec3");

class Mesh {
public:
    std::vector<Vec4> verticesUVX;  // position + u, three per triangle (not shared)
    std::vector<Vec4> normalsUVY;   // normal + v
    BvhTree bvh;
    std::string name;

    // Triangle soup from arrays: positions[9*n], normals[9*n] (may be null:
    // face normals), uvs[6*n] (may be null: (0,0),(0,1),(1,1) per triangle as
    // src/core/Mesh.cpp:86-94).
    bool LoadFromArrays(const float* positions, const float* normals, const float* uvs, int num_tris);
    // Wavefront OBJ (v / vn / vt / f, polygons fan-triangulated), with the
    // reference's v -> 1 - v flip (src/core/Mesh.cpp:44-119).
    bool LoadFromFile(const std::string& filename);
    void BuildBVH();  // per-triangle boxes -> SAH BLAS (src/core/Mesh.cpp:121-139)
};

struct MeshInstance {
    std::string name;
    int meshID = 0;
    Mat4 transform;
    int materialID = 0;
};

class Camera {
public:
    Camera(Vec3 eye, Vec3 lookat, float fov_degrees);
    Vec3 position, up, right, forward;
    float focalDist = 0.1f, aperture = 0.f, fov = 0.f;  // fov in radians
    void OffsetOrientation(float dx, float dy);
    void Strafe(float dx, float dy);
    void SetRadius(float dr);
    void SetFov(float degrees);
private:
    void UpdateCamera();
    Vec3 worldUp{0.f, 1.f, 0.f}, pivot;
    float pitch = 0.f, radius = 0.f, yaw = 0.f;
};

class Scene {
public:
    Scene() = default;
    ~Scene();
    Scene(const Scene&) = delete;
    Scene& operator=(const Scene&) = delete;

    int AddMesh(const std::string& filename);
    int AddMesh(const float* positions, const float* normals, const float* uvs, int num_tris,
                const std::string& name);
    int AddMaterial(const Material& material);
    int AddMeshInstance(const MeshInstance& instance);
    int AddLight(const Light& light);
    int AddTexture(const uint8_t* rgba, int width, int height, const std::string& name);
    // Texture from a PNG file, named by its path (src/core/Scene.cpp:85-107);
    // -1 if it cannot be read.
    int AddTexture(const std::string& filename);
    void AddCamera(Vec3 eye, Vec3 lookat, float fov_degrees);
    void AddEnvMap(const float* rgb, int width, int height);
    // Environment map from a Radiance HDR file (src/core/Scene.cpp:116-132);
    // a file that cannot be read leaves the scene without an environment map.
    void AddEnvMap(const std::string& filename);

    // Builds every BLAS, the TLAS and the flat arrays (src/core/Scene.cpp:216-297).
    // Returns false and sets `error` on invalid input.
    bool ProcessScene();
    // Rebuilds only the TLAS after instance edits (src/core/Scene.cpp:200-214).
    bool RebuildInstances();

    std::vector<Mesh*> meshes;
    std::vector<Material> materials;
    std::vector<MeshInstance> meshInstances;
    std::vector<Light> lights;
    Camera* camera = nullptr;

    // Flat arrays (what the GPU backend uploads).
    std::vector<FlatNode> nodes;
    int topLevelIndex = 0;
    std::vector<Indices> vertIndices;
    std::vector<Vec4> verticesUVX, normalsUVY;
    std::vector<Mat4> transforms;
    BBox sceneBounds;

    // Textures, resized to texArrayWidth x texArrayHeight RGBA8 layers.
    struct Texture { std::string name; int width, height; std::vector<uint8_t> rgba; };
    std::vector<Texture> textures;
    std::vector<uint8_t> textureMapsArray;
    int texArrayWidth = 2048, texArrayHeight = 2048;

    // Environment map (RGB32F) and its running-sum luminance CDF
    // (src/core/EnvironmentMap.cpp:39-61).
    std::vector<float> envImg, envCdf;
    int envWidth = 0, envHeight = 0;
    float envTotalSum = 0.f;

    bool initialized = false, dirty = true, instancesModified = false, envMapModified = false;
    std::string error;

private:
    BvhTree tlas;
    std::vector<int> blasRootIndex;
    void BuildTLAS();
    void FlattenTLAS();
};

}  // namespace glslpt
