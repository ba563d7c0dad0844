// Test-infrastructure harness around the reference's own host C++ (Scene,
// Mesh, Camera, RadeonRays Bvh/SplitBvh/BvhTranslator), compiled from the
// reference sources by oracle/Makefile into oracle/_ref/libref_scene.so.
// It serves the tests only: they call it to obtain the reference's flattened
// arrays for given triangle soups and instances, and compare the product's
// scene builder against them bit for bit. Nothing in the product links it.
//
// Reference entry points exercised:
//   Scene::AddMaterial/AddMeshInstance/AddCamera/ProcessScene  src/core/Scene.cpp:56-297
//   Scene::RebuildInstances                                    src/core/Scene.cpp:200-214
//   Mesh::LoadFromFile (tinyobj)                               src/core/Mesh.cpp:44-119
//   Camera basis                                               src/core/Camera.cpp:120-188
//   LoadSceneFromFile (.scene)                                 src/loaders/Loader.cpp:36-517
#include <cstring>
#include <vector>

#include "Camera.h"
#include "Loader.h"
#include "Mesh.h"
#include "Scene.h"

using namespace GLSLPT;

namespace {
struct Built {
    Scene* scene = nullptr;
};
Built g_last;
}  // namespace

extern "C" {

// Loads an OBJ with the reference loader and returns its triangle soup.
// Call with out arrays NULL to get the vertex count, then again to fill.
int ref_load_obj(const char* filename, int* num_vertices, float* verts_uvx, float* normals_uvy) {
    Mesh mesh;
    if (!mesh.LoadFromFile(filename)) return -1;
    *num_vertices = static_cast<int>(mesh.verticesUVX.size());
    if (verts_uvx) std::memcpy(verts_uvx, mesh.verticesUVX.data(), mesh.verticesUVX.size() * sizeof(Vec4));
    if (normals_uvy) std::memcpy(normals_uvy, mesh.normalsUVY.data(), mesh.normalsUVY.size() * sizeof(Vec4));
    return 0;
}

// Builds a scene with the reference code. Meshes are given as concatenated
// Vec4 arrays (tri_counts[i] triangles each); instances by mesh id, material
// id and a 16-float Mat4. Keeps the scene for ref_get_* until the next call.
int ref_build_scene(int num_meshes, const int* tri_counts, const float* verts_uvx, const float* normals_uvy,
                    int num_materials, int num_instances, const int* mesh_ids, const int* mat_ids,
                    const float* transforms, const float* camera_eye_lookat_fov) {
    delete g_last.scene;
    g_last.scene = new Scene();
    Scene* s = g_last.scene;
    size_t off = 0;
    for (int i = 0; i < num_meshes; ++i) {
        Mesh* m = new Mesh;
        m->name = "mesh" + std::to_string(i);
        const size_t nv = 3 * static_cast<size_t>(tri_counts[i]);
        m->verticesUVX.resize(nv);
        m->normalsUVY.resize(nv);
        std::memcpy(m->verticesUVX.data(), verts_uvx + 4 * off, nv * sizeof(Vec4));
        std::memcpy(m->normalsUVY.data(), normals_uvy + 4 * off, nv * sizeof(Vec4));
        off += nv;
        s->meshes.push_back(m);
    }
    for (int i = 0; i < num_materials; ++i) s->AddMaterial(Material());
    for (int i = 0; i < num_instances; ++i) {
        Mat4 xf;
        std::memcpy(xf.data, transforms + 16 * i, sizeof(xf.data));
        s->AddMeshInstance(MeshInstance("inst", mesh_ids[i], xf, mat_ids[i]));
    }
    if (camera_eye_lookat_fov) {
        const float* c = camera_eye_lookat_fov;
        s->AddCamera(Vec3(c[0], c[1], c[2]), Vec3(c[3], c[4], c[5]), c[6]);
    }
    s->ProcessScene();
    return 0;
}

// Sizes of the last build: nodes, topLevelIndex, vertIndices rows, vertices.
int ref_get_sizes(int* num_nodes, int* top_level_index, int* num_tris, int* num_vertices) {
    Scene* s = g_last.scene;
    if (!s) return -1;
    *num_nodes = static_cast<int>(s->bvhTranslator.nodes.size());
    *top_level_index = s->bvhTranslator.topLevelIndex;
    *num_tris = static_cast<int>(s->vertIndices.size());
    *num_vertices = static_cast<int>(s->verticesUVX.size());
    return 0;
}

// Copies the flattened arrays and the camera basis of the last build.
// camera: position, up, right, forward (12 floats) + fov, focalDist, aperture.
int ref_get_arrays(float* nodes, int* vert_indices, float* verts_uvx, float* normals_uvy, float* camera15) {
    Scene* s = g_last.scene;
    if (!s) return -1;
    std::memcpy(nodes, s->bvhTranslator.nodes.data(), s->bvhTranslator.nodes.size() * 36);
    std::memcpy(vert_indices, s->vertIndices.data(), s->vertIndices.size() * 12);
    std::memcpy(verts_uvx, s->verticesUVX.data(), s->verticesUVX.size() * 16);
    std::memcpy(normals_uvy, s->normalsUVY.data(), s->normalsUVY.size() * 16);
    if (camera15 && s->camera) {
        const Camera* c = s->camera;
        const Vec3* v[4] = {&c->position, &c->up, &c->right, &c->forward};
        for (int i = 0; i < 4; ++i) { camera15[3 * i] = v[i]->x; camera15[3 * i + 1] = v[i]->y; camera15[3 * i + 2] = v[i]->z; }
        camera15[12] = c->fov; camera15[13] = c->focalDist; camera15[14] = c->aperture;
    }
    return 0;
}

// Replaces every instance transform of the last scene (16 floats each, Mat4
// data order) and runs the reference's Scene::RebuildInstances
// (src/core/Scene.cpp:200-214); ref_get_arrays then returns the rebuilt nodes.
int ref_rebuild_instances(const float* transforms) {
    Scene* s = g_last.scene;
    if (!s) return -1;
    for (size_t i = 0; i < s->meshInstances.size(); ++i)
        std::memcpy(s->meshInstances[i].transform.data, transforms + 16 * i, sizeof(s->meshInstances[i].transform.data));
    s->RebuildInstances();
    return 0;
}

// Loads a .scene file with the reference loader into a fresh scene, copies
// its RenderOptions into the scene as the application does
// (src/Main.cpp:121-154, without the default-HDR fallback) and processes it.
// options35: the 35 RenderOptions values in pts_render_options order, as floats.
// Returns 0, or -1 when the loader fails.
int ref_load_scene_file(const char* filename, float* options35) {
    delete g_last.scene;
    g_last.scene = new Scene();
    RenderOptions ro;
    if (!LoadSceneFromFile(filename, g_last.scene, ro)) return -1;
    g_last.scene->renderOptions = ro;
    g_last.scene->ProcessScene();
    const float o[] = {
        float(ro.renderResolution.x), float(ro.renderResolution.y), float(ro.windowResolution.x),
        float(ro.windowResolution.y), ro.uniformLightCol.x, ro.uniformLightCol.y, ro.uniformLightCol.z,
        ro.backgroundCol.x, ro.backgroundCol.y, ro.backgroundCol.z, float(ro.tileWidth), float(ro.tileHeight),
        float(ro.maxDepth), float(ro.maxSpp), float(ro.RRDepth), float(ro.texArrayWidth), float(ro.texArrayHeight),
        float(ro.denoiserFrameCnt), float(ro.enableRR), float(ro.enableDenoiser), float(ro.enableTonemap),
        float(ro.enableAces), 0.f /* simpleAcesFit: not initialised by RenderOptions() */, float(ro.openglNormalMap),
        float(ro.enableEnvMap), float(ro.enableUniformLight), float(ro.hideEmitters), float(ro.enableBackground),
        float(ro.transparentBackground), float(ro.independentRenderSize), float(ro.enableRoughnessMollification),
        float(ro.enableVolumeMIS), ro.envMapIntensity, ro.envMapRot, ro.roughnessMollificationAmt};
    std::memcpy(options35, o, sizeof(o));
    return 0;
}

// Counts of the last scene's materials, lights, instances, textures (and the
// texture-array layer size) and whether an environment map was loaded.
int ref_get_counts(int* counts7) {
    Scene* s = g_last.scene;
    if (!s) return -1;
    counts7[0] = static_cast<int>(s->materials.size());
    counts7[1] = static_cast<int>(s->lights.size());
    counts7[2] = static_cast<int>(s->transforms.size());
    counts7[3] = static_cast<int>(s->textures.size());
    counts7[4] = s->renderOptions.texArrayWidth;
    counts7[5] = s->renderOptions.texArrayHeight;
    counts7[6] = s->envMap ? 1 : 0;
    return 0;
}

// Copies materials (32 floats each), lights (15) and transforms (16) of the
// last scene.
int ref_get_records(float* materials, float* lights, float* transforms) {
    Scene* s = g_last.scene;
    if (!s) return -1;
    std::memcpy(materials, s->materials.data(), s->materials.size() * sizeof(Material));
    std::memcpy(lights, s->lights.data(), s->lights.size() * sizeof(Light));
    std::memcpy(transforms, s->transforms.data(), s->transforms.size() * sizeof(Mat4));
    return 0;
}

// Decoded pixels of texture i of the last scene (before the resize into the
// texture array): call with rgba NULL for the size.
int ref_get_texture(int i, int* w, int* h, unsigned char* rgba) {
    Scene* s = g_last.scene;
    if (!s || i < 0 || i >= static_cast<int>(s->textures.size())) return -1;
    const Texture* t = s->textures[i];
    *w = t->width;
    *h = t->height;
    if (rgba) std::memcpy(rgba, t->texData.data(), t->texData.size());
    return 0;
}

// Environment map of the last scene as EnvironmentMap holds it: size, RGB
// image, running-sum CDF (BuildCDF, src/core/EnvironmentMap.cpp:39-61) and
// totalSum. Call with img NULL for the size.
int ref_get_env(int* w, int* h, float* img, float* cdf, float* total) {
    Scene* s = g_last.scene;
    if (!s || !s->envMap) return -1;
    const EnvironmentMap* e = s->envMap;
    *w = e->width;
    *h = e->height;
    const size_t n = static_cast<size_t>(e->width) * e->height;
    if (img) std::memcpy(img, e->img, 3 * n * sizeof(float));
    if (cdf) std::memcpy(cdf, e->cdf, n * sizeof(float));
    if (total) *total = e->totalSum;
    return 0;
}

void ref_free(void) {
    delete g_last.scene;
    g_last.scene = nullptr;
}

}  // extern "C"
