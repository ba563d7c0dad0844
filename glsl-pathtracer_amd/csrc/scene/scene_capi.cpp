// extern "C" surface of the scene builder (pts_* in include/glslpt.h).
#include <cstring>
#include <exception>
#include <new>

#include "../../../include/glslpt.h"
#include "scene.h"

struct pts_scene_s {
    glslpt::Scene scene;
    std::string err;
};

namespace {
static_assert(sizeof(pt_material) == sizeof(glslpt::Material), "material layout");
static_assert(sizeof(pt_light) == sizeof(glslpt::Light), "light layout");
static_assert(sizeof(pt_bvh_node) == sizeof(glslpt::FlatNode), "node layout");
static_assert(sizeof(glslpt::Mat4) == 64, "mat4 layout");

int Fail(pts_scene s, int code, const std::string& msg) {
    if (s) s->err = msg;
    return code;
}

template <class F>
int Guard(pts_scene s, F&& f) {
    if (!s) return PT_ERR_INVALID_ARG;
    try {
        return f();
    } catch (const std::bad_alloc&) {
        return Fail(s, PT_ERR_OVERFLOW, "out of host memory");
    } catch (const std::exception& e) {
        return Fail(s, PT_ERR_STATE, e.what());
    }
}

glslpt::Mat4 ToMat4(const float* t) {
    glslpt::Mat4 m;
    std::memcpy(m.data, t, sizeof(m.data));
    return m;
}
}  // namespace

extern "C" {

int pts_create(pts_scene* out) {
    if (!out) return PT_ERR_INVALID_ARG;
    *out = new (std::nothrow) pts_scene_s;
    return *out ? PT_OK : PT_ERR_OVERFLOW;
}

int pts_destroy(pts_scene s) {
    delete s;
    return PT_OK;
}

const char* pts_last_error(pts_scene s) { return s ? s->err.c_str() : "null scene handle"; }

int pts_add_mesh_obj(pts_scene s, const char* filename, int32_t* id) {
    return Guard(s, [&] {
        if (!filename || !id) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        *id = s->scene.AddMesh(std::string(filename));
        return *id < 0 ? Fail(s, PT_ERR_IO, s->scene.error) : PT_OK;
    });
}

int pts_add_mesh(pts_scene s, const float* positions, const float* normals, const float* uvs,
                 int32_t num_tris, const char* name, int32_t* id) {
    return Guard(s, [&] {
        if (!positions || num_tris <= 0 || !id) return Fail(s, PT_ERR_INVALID_ARG, "mesh needs positions and num_tris > 0");
        *id = s->scene.AddMesh(positions, normals, uvs, num_tris, name ? name : "");
        return *id < 0 ? Fail(s, PT_ERR_INVALID_ARG, s->scene.error) : PT_OK;
    });
}

int pts_add_material(pts_scene s, const pt_material* m, int32_t* id) {
    return Guard(s, [&] {
        if (!m || !id) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        glslpt::Material mat;
        std::memcpy(static_cast<void*>(&mat), m, sizeof(mat));
        *id = s->scene.AddMaterial(mat);
        return PT_OK;
    });
}

int pts_add_mesh_instance(pts_scene s, int32_t mesh_id, const float* transform, int32_t material_id,
                          int32_t* id) {
    return Guard(s, [&] {
        if (!transform || !id) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        glslpt::MeshInstance mi;
        mi.meshID = mesh_id;
        mi.materialID = material_id;
        mi.transform = ToMat4(transform);
        *id = s->scene.AddMeshInstance(mi);
        return PT_OK;
    });
}

int pts_add_light(pts_scene s, const pt_light* l, int32_t* id) {
    return Guard(s, [&] {
        if (!l || !id) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        glslpt::Light light;
        std::memcpy(static_cast<void*>(&light), l, sizeof(light));
        *id = s->scene.AddLight(light);
        return PT_OK;
    });
}

int pts_add_texture(pts_scene s, const uint8_t* rgba, int32_t width, int32_t height, const char* name,
                    int32_t* id) {
    return Guard(s, [&] {
        if (!id) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        *id = s->scene.AddTexture(rgba, width, height, name ? name : "");
        return *id < 0 ? Fail(s, PT_ERR_INVALID_ARG, s->scene.error) : PT_OK;
    });
}

int pts_add_camera(pts_scene s, const float eye[3], const float lookat[3], float fov_degrees) {
    return Guard(s, [&] {
        if (!eye || !lookat) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        s->scene.AddCamera(glslpt::Vec3(eye[0], eye[1], eye[2]), glslpt::Vec3(lookat[0], lookat[1], lookat[2]),
                           fov_degrees);
        return PT_OK;
    });
}

int pts_add_env_map(pts_scene s, const float* rgb, int32_t width, int32_t height) {
    return Guard(s, [&] {
        if (!rgb || width <= 0 || height <= 0) return Fail(s, PT_ERR_INVALID_ARG, "invalid environment map");
        s->scene.AddEnvMap(rgb, width, height);
        return PT_OK;
    });
}

int pts_set_texture_array_size(pts_scene s, int32_t width, int32_t height) {
    return Guard(s, [&] {
        if (width <= 0 || height <= 0) return Fail(s, PT_ERR_INVALID_ARG, "invalid texture array size");
        s->scene.texArrayWidth = width;
        s->scene.texArrayHeight = height;
        return PT_OK;
    });
}

int pts_process(pts_scene s) {
    return Guard(s, [&] {
        return s->scene.ProcessScene() ? PT_OK : Fail(s, PT_ERR_INVALID_ARG, s->scene.error);
    });
}

int pts_rebuild_instances(pts_scene s) {
    return Guard(s, [&] {
        return s->scene.RebuildInstances() ? PT_OK : Fail(s, PT_ERR_STATE, s->scene.error);
    });
}

int pts_set_instance_transform(pts_scene s, int32_t instance, const float* transform) {
    return Guard(s, [&] {
        if (!transform || instance < 0 || instance >= static_cast<int32_t>(s->scene.meshInstances.size()))
            return Fail(s, PT_ERR_INVALID_ARG, "instance index out of range");
        s->scene.meshInstances[instance].transform = ToMat4(transform);
        return PT_OK;
    });
}

int pts_get_view(pts_scene s, pt_scene_view* v) {
    return Guard(s, [&] {
        if (!v) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        const glslpt::Scene& sc = s->scene;
        if (!sc.initialized) return Fail(s, PT_ERR_STATE, "ProcessScene has not run");
        std::memset(v, 0, sizeof(*v));
        v->nodes = reinterpret_cast<const pt_bvh_node*>(sc.nodes.data());
        v->num_nodes = static_cast<int32_t>(sc.nodes.size());
        v->top_level_index = sc.topLevelIndex;
        v->vert_indices = reinterpret_cast<const int32_t*>(sc.vertIndices.data());
        v->num_tris = static_cast<int32_t>(sc.vertIndices.size());
        v->vertices_uvx = reinterpret_cast<const float*>(sc.verticesUVX.data());
        v->normals_uvy = reinterpret_cast<const float*>(sc.normalsUVY.data());
        v->num_vertices = static_cast<int32_t>(sc.verticesUVX.size());
        v->materials = reinterpret_cast<const pt_material*>(sc.materials.data());
        v->num_materials = static_cast<int32_t>(sc.materials.size());
        v->transforms = reinterpret_cast<const float*>(sc.transforms.data());
        v->num_instances = static_cast<int32_t>(sc.transforms.size());
        v->lights = reinterpret_cast<const pt_light*>(sc.lights.data());
        v->num_lights = static_cast<int32_t>(sc.lights.size());
        v->texture_maps = sc.textureMapsArray.empty() ? nullptr : sc.textureMapsArray.data();
        v->tex_width = sc.texArrayWidth;
        v->tex_height = sc.texArrayHeight;
        v->num_textures = static_cast<int32_t>(sc.textures.size());
        v->env_img = sc.envImg.empty() ? nullptr : sc.envImg.data();
        v->env_cdf = sc.envCdf.empty() ? nullptr : sc.envCdf.data();
        v->env_width = sc.envWidth;
        v->env_height = sc.envHeight;
        v->env_total_sum = sc.envTotalSum;
        return PT_OK;
    });
}

int pts_get_camera(pts_scene s, pt_camera* c) {
    return Guard(s, [&] {
        if (!c) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        const glslpt::Camera* cam = s->scene.camera;
        if (!cam) return Fail(s, PT_ERR_STATE, "scene has no camera");
        auto put = [](float* d, const glslpt::Vec3& v) { d[0] = v.x; d[1] = v.y; d[2] = v.z; };
        put(c->position, cam->position);
        put(c->up, cam->up);
        put(c->right, cam->right);
        put(c->forward, cam->forward);
        c->fov = cam->fov;
        c->focalDist = cam->focalDist;
        c->aperture = cam->aperture;
        return PT_OK;
    });
}

int pts_camera_offset_orientation(pts_scene s, float dx, float dy) {
    return Guard(s, [&] {
        if (!s->scene.camera) return Fail(s, PT_ERR_STATE, "scene has no camera");
        s->scene.camera->OffsetOrientation(dx, dy);
        return PT_OK;
    });
}

int pts_get_blas_root(pts_scene s, int32_t mesh_id, int32_t* root) {
    return Guard(s, [&] {
        if (!root || mesh_id < 0 || mesh_id >= static_cast<int32_t>(s->scene.meshes.size()))
            return Fail(s, PT_ERR_INVALID_ARG, "mesh index out of range");
        int32_t r = 0;
        for (int32_t i = 0; i < mesh_id; ++i) r += static_cast<int32_t>(s->scene.meshes[i]->bvh.nodes.size());
        *root = r;
        return PT_OK;
    });
}

}  // extern "C"
