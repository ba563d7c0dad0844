// HipRenderer: C++ mirror of GLSLPT::Renderer (src/core/Renderer.h:104-185)
// over the glslpt C ABI. Header-only; link libglslpt_scene.so and
// libglslpt_hip.so.
//
// The reference renders one tile of one sample per Render() call and shows a
// low-resolution preview while the scene is dirty (src/core/Renderer.cpp:
// 546-590, 641-812). Here a Render() call adds one full-frame sample on the
// GPU (the tile is the whole image), Present() has no GL work to do, and
// GetOutputBuffer() returns the tonemapped RGBA8 image allocated with new[]
// for the caller to delete[], as the reference does (Renderer.cpp:619-634).
#pragma once

#include <stdexcept>
#include <string>

#include "glslpt.h"

namespace glslpt {

class HipRenderer {
public:
    // Renderer(Scene*, shadersDirectory): the scene must have been processed
    // (pts_process); options play the role of Scene::renderOptions.
    HipRenderer(pts_scene scene, const pt_options& options, int device = 0) : scene_(scene), opts_(options) {
        check(pt_create(device, &h_), "pt_create");
        Upload();
    }
    ~HipRenderer() { pt_destroy(h_); }
    HipRenderer(const HipRenderer&) = delete;
    HipRenderer& operator=(const HipRenderer&) = delete;

    pt_options& Options() { return opts_; }

    // Re-reads the resolution and options (Renderer.cpp:251-279).
    void ResizeRenderer() { ApplyOptions(); }
    // No shaders to compile: option switches are re-applied (Renderer.cpp:380-390).
    void ReloadShaders() { ApplyOptions(); }

    // Adds the next sample of every pixel unless maxSpp is reached (Renderer.cpp:546-551).
    void Render() {
        if (max_spp_ > 0 && samples_ >= max_spp_) return;
        check(pt_render(h_, samples_, 1), "pt_render");
        ++samples_;
    }
    void Present() {}

    // Refreshes camera and instances; a dirty scene restarts accumulation
    // (Renderer.cpp:641-665, 733-744).
    void Update(float /*secondsElapsed*/) {
        pt_camera cam;
        check(pts_get_camera(scene_, &cam), "pts_get_camera");
        check(pt_set_camera(h_, &cam), "pt_set_camera");
        if (dirty_) {
            check(pt_clear_accum(h_), "pt_clear_accum");
            samples_ = 0;
            dirty_ = false;
        }
    }

    // After Scene::RebuildInstances: new transforms, materials and TLAS nodes.
    void InstancesModified() {
        pt_scene_view v;
        check(pts_get_view(scene_, &v), "pts_get_view");
        check(pt_update_instances(h_, v.transforms, v.num_instances, v.materials, v.num_materials,
                                  v.nodes + v.top_level_index, v.top_level_index,
                                  v.num_nodes - v.top_level_index),
              "pt_update_instances");
        dirty_ = true;
    }
    void CameraMoved() { dirty_ = true; }

    float GetProgress() const { return max_spp_ <= 0 ? 0.0f : samples_ * 100.0f / max_spp_; }
    int GetSampleCount() const { return samples_; }
    void SetMaxSpp(int spp) { max_spp_ = spp; }

    // Tonemapped RGBA8, rows bottom-up like glGetTexImage; caller delete[]s.
    void GetOutputBuffer(unsigned char** data, int& w, int& h, bool tonemap = true, bool aces = false) {
        w = opts_.width;
        h = opts_.height;
        *data = new unsigned char[static_cast<size_t>(w) * h * 4];
        const float bg[3] = {1.f, 1.f, 1.f};
        int rc = pt_read_ldr(h_, samples_ > 0 ? samples_ : 1, tonemap, aces, 0, bg, *data, static_cast<size_t>(w) * h * 4);
        if (rc != PT_OK) {
            delete[] *data;
            *data = nullptr;
            check(rc, "pt_read_ldr");
        }
    }
    // Raw RGBA32F accumulation (sum over samples).
    void GetAccumulation(float* rgba) {
        check(pt_read_accum(h_, rgba, static_cast<size_t>(opts_.width) * opts_.height * 4), "pt_read_accum");
    }

    pt_handle handle() const { return h_; }

private:
    void check(int rc, const char* what) {
        if (rc != PT_OK) throw std::runtime_error(std::string(what) + ": " + pt_last_error(h_));
    }
    void ApplyOptions() {
        check(pt_set_options(h_, &opts_), "pt_set_options");
        dirty_ = true;
    }
    void Upload() {
        pt_scene_view v;
        if (pts_get_view(scene_, &v) != PT_OK) throw std::runtime_error(pts_last_error(scene_));
        check(pt_upload_scene(h_, &v), "pt_upload_scene");
        ApplyOptions();
        Update(0.f);
    }

    pts_scene scene_;
    pt_options opts_;
    pt_handle h_ = nullptr;
    int samples_ = 0;
    int max_spp_ = -1;
    bool dirty_ = true;
};

}  // namespace glslpt
