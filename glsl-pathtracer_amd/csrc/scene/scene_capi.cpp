// extern "C" surface of the scene builder (pts_* in include/glslpt.h).
#include <cstring>
#include <exception>
#include <new>

#include "../../../include/glslpt.h"
#include "scene.h"
#include "scene_loader.h"

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

namespace {
void ToC(const glslpt::RenderOptions& r, pts_render_options* o) {
    o->render_resolution[0] = r.renderResolution[0]; o->render_resolution[1] = r.renderResolution[1];
    o->window_resolution[0] = r.windowResolution[0]; o->window_resolution[1] = r.windowResolution[1];
    o->uniform_light_col[0] = r.uniformLightCol.x; o->uniform_light_col[1] = r.uniformLightCol.y;
    o->uniform_light_col[2] = r.uniformLightCol.z;
    o->background_col[0] = r.backgroundCol.x; o->background_col[1] = r.backgroundCol.y;
    o->background_col[2] = r.backgroundCol.z;
    o->tile_width = r.tileWidth; o->tile_height = r.tileHeight; o->max_depth = r.maxDepth; o->max_spp = r.maxSpp;
    o->rr_depth = r.RRDepth; o->tex_array_width = r.texArrayWidth; o->tex_array_height = r.texArrayHeight;
    o->denoiser_frame_cnt = r.denoiserFrameCnt;
    o->enable_rr = r.enableRR; o->enable_denoiser = r.enableDenoiser; o->enable_tonemap = r.enableTonemap;
    o->enable_aces = r.enableAces; o->simple_aces_fit = r.simpleAcesFit; o->opengl_normal_map = r.openglNormalMap;
    o->enable_envmap = r.enableEnvMap; o->enable_uniform_light = r.enableUniformLight;
    o->hide_emitters = r.hideEmitters; o->enable_background = r.enableBackground;
    o->transparent_background = r.transparentBackground; o->independent_render_size = r.independentRenderSize;
    o->enable_roughness_mollification = r.enableRoughnessMollification;
    o->enable_volume_mis = r.enableVolumeMIS;
    o->envmap_intensity = r.envMapIntensity; o->envmap_rot = r.envMapRot;
    o->roughness_mollification_amt = r.roughnessMollificationAmt;
}

glslpt::RenderOptions FromC(const pts_render_options* o) {
    glslpt::RenderOptions r;
    r.renderResolution[0] = o->render_resolution[0]; r.renderResolution[1] = o->render_resolution[1];
    r.windowResolution[0] = o->window_resolution[0]; r.windowResolution[1] = o->window_resolution[1];
    r.uniformLightCol = glslpt::Vec3(o->uniform_light_col[0], o->uniform_light_col[1], o->uniform_light_col[2]);
    r.backgroundCol = glslpt::Vec3(o->background_col[0], o->background_col[1], o->background_col[2]);
    r.tileWidth = o->tile_width; r.tileHeight = o->tile_height; r.maxDepth = o->max_depth; r.maxSpp = o->max_spp;
    r.RRDepth = o->rr_depth; r.texArrayWidth = o->tex_array_width; r.texArrayHeight = o->tex_array_height;
    r.denoiserFrameCnt = o->denoiser_frame_cnt;
    r.enableRR = o->enable_rr; r.enableDenoiser = o->enable_denoiser; r.enableTonemap = o->enable_tonemap;
    r.enableAces = o->enable_aces; r.simpleAcesFit = o->simple_aces_fit; r.openglNormalMap = o->opengl_normal_map;
    r.enableEnvMap = o->enable_envmap; r.enableUniformLight = o->enable_uniform_light;
    r.hideEmitters = o->hide_emitters; r.enableBackground = o->enable_background;
    r.transparentBackground = o->transparent_background; r.independentRenderSize = o->independent_render_size;
    r.enableRoughnessMollification = o->enable_roughness_mollification;
    r.enableVolumeMIS = o->enable_volume_mis;
    r.envMapIntensity = o->envmap_intensity; r.envMapRot = o->envmap_rot;
    r.roughnessMollificationAmt = o->roughness_mollification_amt;
    return r;
}
}  // namespace

int pts_get_texture(pts_scene s, int32_t id, int32_t* width, int32_t* height, uint8_t* rgba) {
    return Guard(s, [&] {
        if (!width || !height) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        if (id < 0 || id >= static_cast<int32_t>(s->scene.textures.size())) return Fail(s, PT_ERR_INVALID_ARG, "no such texture");
        const auto& t = s->scene.textures[id];
        *width = t.width;
        *height = t.height;
        if (rgba) std::memcpy(rgba, t.rgba.data(), t.rgba.size());
        return PT_OK;
    });
}

void pts_default_render_options(pts_render_options* o) {
    if (o) ToC(glslpt::RenderOptions(), o);
}

int pts_load_scene_file(pts_scene s, const char* filename, pts_render_options* options) {
    return Guard(s, [&] {
        if (!filename || !options) return Fail(s, PT_ERR_INVALID_ARG, "null argument");
        glslpt::RenderOptions r = FromC(options);
        std::string log;
        const std::string fn(filename);
        const bool ok = glslpt::LoadSceneFromFile(fn, s->scene, r, log);
        ToC(r, options);
        if (!ok) {
            const bool gltf = log.find("glTF") != std::string::npos;
            return Fail(s, gltf ? PT_ERR_UNSUPPORTED : PT_ERR_IO, log);
        }
        s->err = log;
        return PT_OK;
    });
}

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
