// Loader for the reference's .scene files (src/loaders/Loader.cpp:36-517)
// into a Scene plus the RenderOptions the file sets.
#pragma once

#include <string>

#include "scene.h"

namespace glslpt {

// RenderOptions with the reference's defaults (src/core/Renderer.h:37-100).
struct RenderOptions {
    int renderResolution[2] = {1280, 720};
    int windowResolution[2] = {1280, 720};
    Vec3 uniformLightCol{0.3f, 0.3f, 0.3f};
    Vec3 backgroundCol{1.0f, 1.0f, 1.0f};
    int tileWidth = 100, tileHeight = 100;
    int maxDepth = 2, maxSpp = -1, RRDepth = 2;
    int texArrayWidth = 2048, texArrayHeight = 2048;
    int denoiserFrameCnt = 20;
    bool enableRR = true, enableDenoiser = false, enableTonemap = true, enableAces = false;
    bool simpleAcesFit = false, openglNormalMap = true, enableEnvMap = false, enableUniformLight = false;
    bool hideEmitters = false, enableBackground = false, transparentBackground = false;
    bool independentRenderSize = false, enableRoughnessMollification = false, enableVolumeMIS = false;
    float envMapIntensity = 1.0f, envMapRot = 0.0f, roughnessMollificationAmt = 0.0f;
};

// Parses `filename` into `scene` (materials, lights, camera, meshes and
// instances, textures, environment map) and `options`. Returns false when the
// file cannot be opened or names a glTF file; problems that the reference
// only reports (a missing mesh, texture or environment map, an unknown
// material name) are appended to `log` and the element is skipped.
bool LoadSceneFromFile(const std::string& filename, Scene& scene, RenderOptions& options, std::string& log);

}  // namespace glslpt
