// The .scene text format (src/loaders/Loader.cpp:36-517): blocks
//   material <name> { ... }   light { ... }   camera { ... }
//   renderer { ... }          mesh { ... }    gltf { ... }
// read line by line. Keys, their sscanf patterns and the way a block is
// recognised (a line containing the block's keyword) are the format, so they
// are kept exactly; so are its corner cases: a blank line inside a camera or
// mesh block counts as a `matrix` line (sscanf returns EOF there), the first
// definition of a material name wins, an unknown material name leaves the
// mesh's material unchanged, and relative file names are taken from the
// scene file's directory (the whole file name plus '/' when it has none).
// glTF blocks are rejected: this build has no glTF importer.
#include <cstdio>
#include <cstring>
#include <map>
#include <string>

#include "image_io.h"
#include "scene.h"
#include "scene_loader.h"

namespace glslpt {

namespace {

constexpr int kLine = 2048;

// "true" / "false" switch of the renderer block; anything else keeps the value.
void Switch(const char* v, bool& out) {
    if (std::strcmp(v, "false") == 0) out = false;
    else if (std::strcmp(v, "true") == 0) out = true;
}

bool ReadLine(FILE* f, char* line) { return std::fgets(line, kLine, f) != nullptr; }

// sscanf of a 16-float `matrix` line into data[col][row] order; true unless
// the pattern fails before its first conversion with input left (sscanf == 0).
bool MatrixLine(const char* line, Mat4& m) {
    return std::sscanf(line, " matrix %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f %f",
                       &m[0][0], &m[1][0], &m[2][0], &m[3][0], &m[0][1], &m[1][1], &m[2][1], &m[3][1],
                       &m[0][2], &m[1][2], &m[2][2], &m[3][2], &m[0][3], &m[1][3], &m[2][3], &m[3][3]) != 0;
}

void ReadMaterial(FILE* f, char* line, Scene& scene, const std::string& dir, Material& mat) {
    char albedo[100] = "none", mr[100] = "none", normal[100] = "none", emission[100] = "none";
    char alphaMode[20] = "none", mediumType[20] = "none";
    while (ReadLine(f, line)) {
        if (std::strchr(line, '}')) break;
        std::sscanf(line, " color %f %f %f", &mat.baseColor.x, &mat.baseColor.y, &mat.baseColor.z);
        std::sscanf(line, " opacity %f", &mat.opacity);
        std::sscanf(line, " alphamode %s", alphaMode);
        std::sscanf(line, " alphacutoff %f", &mat.alphaCutoff);
        std::sscanf(line, " emission %f %f %f", &mat.emission.x, &mat.emission.y, &mat.emission.z);
        std::sscanf(line, " metallic %f", &mat.metallic);
        std::sscanf(line, " roughness %f", &mat.roughness);
        std::sscanf(line, " subsurface %f", &mat.subsurface);
        std::sscanf(line, " speculartint %f", &mat.specularTint);
        std::sscanf(line, " anisotropic %f", &mat.anisotropic);
        std::sscanf(line, " sheen %f", &mat.sheen);
        std::sscanf(line, " sheentint %f", &mat.sheenTint);
        std::sscanf(line, " clearcoat %f", &mat.clearcoat);
        std::sscanf(line, " clearcoatgloss %f", &mat.clearcoatGloss);
        std::sscanf(line, " spectrans %f", &mat.specTrans);
        std::sscanf(line, " ior %f", &mat.ior);
        std::sscanf(line, " albedotexture %99s", albedo);
        std::sscanf(line, " metallicroughnesstexture %99s", mr);
        std::sscanf(line, " normaltexture %99s", normal);
        std::sscanf(line, " emissiontexture %99s", emission);
        std::sscanf(line, " mediumtype %19s", mediumType);
        std::sscanf(line, " mediumdensity %f", &mat.mediumDensity);
        std::sscanf(line, " mediumcolor %f %f %f", &mat.mediumColor.x, &mat.mediumColor.y, &mat.mediumColor.z);
        std::sscanf(line, " mediumanisotropy %f", &mat.mediumAnisotropy);
    }
    if (std::strcmp(albedo, "none") != 0) mat.baseColorTexId = static_cast<float>(scene.AddTexture(dir + albedo));
    if (std::strcmp(mr, "none") != 0) mat.metallicRoughnessTexID = static_cast<float>(scene.AddTexture(dir + mr));
    if (std::strcmp(normal, "none") != 0) mat.normalmapTexID = static_cast<float>(scene.AddTexture(dir + normal));
    if (std::strcmp(emission, "none") != 0) mat.emissionmapTexID = static_cast<float>(scene.AddTexture(dir + emission));
    if (std::strcmp(alphaMode, "opaque") == 0) mat.alphaMode = 0.f;
    else if (std::strcmp(alphaMode, "blend") == 0) mat.alphaMode = 1.f;
    else if (std::strcmp(alphaMode, "mask") == 0) mat.alphaMode = 2.f;
    if (std::strcmp(mediumType, "absorb") == 0) mat.mediumType = 1.f;
    else if (std::strcmp(mediumType, "scatter") == 0) mat.mediumType = 2.f;
    else if (std::strcmp(mediumType, "emissive") == 0) mat.mediumType = 3.f;
}

void ReadLight(FILE* f, char* line, Scene& scene) {
    Light light;
    Vec3 v1, v2;
    char type[20] = "none";
    while (ReadLine(f, line)) {
        if (std::strchr(line, '}')) break;
        std::sscanf(line, " position %f %f %f", &light.position.x, &light.position.y, &light.position.z);
        std::sscanf(line, " emission %f %f %f", &light.emission.x, &light.emission.y, &light.emission.z);
        std::sscanf(line, " radius %f", &light.radius);
        std::sscanf(line, " v1 %f %f %f", &v1.x, &v1.y, &v1.z);
        std::sscanf(line, " v2 %f %f %f", &v2.x, &v2.y, &v2.z);
        std::sscanf(line, " type %19s", type);
    }
    if (std::strcmp(type, "quad") == 0) {
        light.type = static_cast<float>(RectLight);
        light.u = v1 - light.position;
        light.v = v2 - light.position;
        light.area = Vec3::Length(Vec3::Cross(light.u, light.v));
    } else if (std::strcmp(type, "sphere") == 0) {
        light.type = static_cast<float>(SphereLight);
        light.area = 4.0f * kPi * light.radius * light.radius;
    } else if (std::strcmp(type, "distant") == 0) {
        light.type = static_cast<float>(DistantLight);
        light.area = 0.0f;
    }
    scene.AddLight(light);
}

void ReadCamera(FILE* f, char* line, Scene& scene) {
    Mat4 xform;
    Vec3 position, lookAt;
    float fov = 45.f, aperture = 0.f, focalDist = 1.f;
    bool matrix = false;
    while (ReadLine(f, line)) {
        if (std::strchr(line, '}')) break;
        std::sscanf(line, " position %f %f %f", &position.x, &position.y, &position.z);
        std::sscanf(line, " lookat %f %f %f", &lookAt.x, &lookAt.y, &lookAt.z);
        std::sscanf(line, " aperture %f ", &aperture);
        std::sscanf(line, " focaldist %f", &focalDist);
        std::sscanf(line, " fov %f", &fov);
        if (MatrixLine(line, xform)) matrix = true;
    }
    if (matrix) {
        const Vec3 forward(xform[2][0], xform[2][1], xform[2][2]);
        position = Vec3(xform[3][0], xform[3][1], xform[3][2]);
        lookAt = position + forward;
    }
    scene.AddCamera(position, lookAt, fov);
    scene.camera->aperture = aperture;
    scene.camera->focalDist = focalDist;
}

void ReadRenderer(FILE* f, char* line, Scene& scene, const std::string& dir, RenderOptions& o) {
    char envMap[200] = "none", rr[10] = "none", aces[10] = "none", nmap[10] = "none", hide[10] = "none";
    char transparent[10] = "none", background[10] = "none", independent[10] = "none", tonemap[10] = "none";
    char moll[10] = "none", volmis[10] = "none", uniform[10] = "none";
    while (ReadLine(f, line)) {
        if (std::strchr(line, '}')) break;
        std::sscanf(line, " envmapfile %199s", envMap);
        std::sscanf(line, " resolution %d %d", &o.renderResolution[0], &o.renderResolution[1]);
        std::sscanf(line, " windowresolution %d %d", &o.windowResolution[0], &o.windowResolution[1]);
        std::sscanf(line, " envmapintensity %f", &o.envMapIntensity);
        std::sscanf(line, " maxdepth %i", &o.maxDepth);
        std::sscanf(line, " maxspp %i", &o.maxSpp);
        std::sscanf(line, " tilewidth %i", &o.tileWidth);
        std::sscanf(line, " tileheight %i", &o.tileHeight);
        std::sscanf(line, " enablerr %9s", rr);
        std::sscanf(line, " rrdepth %i", &o.RRDepth);
        std::sscanf(line, " enabletonemap %9s", tonemap);
        std::sscanf(line, " enableaces %9s", aces);
        std::sscanf(line, " texarraywidth %i", &o.texArrayWidth);
        std::sscanf(line, " texarrayheight %i", &o.texArrayHeight);
        std::sscanf(line, " openglnormalmap %9s", nmap);
        std::sscanf(line, " hideemitters %9s", hide);
        std::sscanf(line, " enablebackground %9s", background);
        std::sscanf(line, " transparentbackground %9s", transparent);
        std::sscanf(line, " backgroundcolor %f %f %f", &o.backgroundCol.x, &o.backgroundCol.y, &o.backgroundCol.z);
        std::sscanf(line, " independentrendersize %9s", independent);
        std::sscanf(line, " envmaprotation %f", &o.envMapRot);
        std::sscanf(line, " enableroughnessmollification %9s", moll);
        std::sscanf(line, " roughnessmollificationamt %f", &o.roughnessMollificationAmt);
        std::sscanf(line, " enablevolumemis %9s", volmis);
        std::sscanf(line, " enableuniformlight %9s", uniform);
        std::sscanf(line, " uniformlightcolor %f %f %f", &o.uniformLightCol.x, &o.uniformLightCol.y,
                    &o.uniformLightCol.z);
    }
    if (std::strcmp(envMap, "none") != 0) {
        scene.AddEnvMap(dir + envMap);
        o.enableEnvMap = true;
    } else {
        o.enableEnvMap = false;
    }
    Switch(aces, o.enableAces);
    Switch(rr, o.enableRR);
    Switch(nmap, o.openglNormalMap);
    Switch(hide, o.hideEmitters);
    Switch(background, o.enableBackground);
    Switch(transparent, o.transparentBackground);
    Switch(independent, o.independentRenderSize);
    Switch(tonemap, o.enableTonemap);
    Switch(moll, o.enableRoughnessMollification);
    Switch(volmis, o.enableVolumeMIS);
    Switch(uniform, o.enableUniformLight);
    if (!o.independentRenderSize) {
        o.windowResolution[0] = o.renderResolution[0];
        o.windowResolution[1] = o.renderResolution[1];
    }
}

void ReadMesh(FILE* f, char* line, Scene& scene, const std::string& dir,
              const std::map<std::string, int>& materials, std::string& log) {
    std::string filename;
    float q[4] = {0.f, 0.f, 0.f, 0.f};
    Mat4 xform, translate, rot, scale;
    int material = 0;
    char meshName[200] = "none";
    bool matrix = false;
    while (ReadLine(f, line)) {
        if (std::strchr(line, '}')) break;
        char file[kLine];
        char matName[100];
        std::sscanf(line, " name %199[^\t\n]s", meshName);
        if (std::sscanf(line, " file %2047s", file) == 1) filename = dir + file;
        if (std::sscanf(line, " material %99s", matName) == 1) {
            auto it = materials.find(matName);
            if (it != materials.end()) material = it->second;
            else log += std::string("could not find material ") + matName + "\n";
        }
        if (MatrixLine(line, xform)) matrix = true;
        std::sscanf(line, " position %f %f %f", &translate[3][0], &translate[3][1], &translate[3][2]);
        std::sscanf(line, " scale %f %f %f", &scale[0][0], &scale[1][1], &scale[2][2]);
        if (std::sscanf(line, " rotation %f %f %f %f", &q[0], &q[1], &q[2], &q[3]) != 0)
            rot = Mat4::QuatToMatrix(q[0], q[1], q[2], q[3]);
    }
    if (filename.empty()) return;
    const int mesh = scene.AddMesh(filename);
    if (mesh == -1) {
        log += "unable to load model " + filename + "\n";
        return;
    }
    MeshInstance inst;
    if (std::strcmp(meshName, "none") != 0) inst.name = meshName;
    else inst.name = filename.substr(filename.find_last_of("/\\") + 1);
    inst.meshID = mesh;
    inst.transform = matrix ? xform : scale * rot * translate;
    inst.materialID = material;
    scene.AddMeshInstance(inst);
}

}  // namespace

bool LoadSceneFromFile(const std::string& filename, Scene& scene, RenderOptions& options, std::string& log) {
    FILE* f = std::fopen(filename.c_str(), "r");
    if (!f) {
        log += "couldn't open " + filename + " for reading\n";
        return false;
    }
    const std::string dir = filename.substr(0, filename.find_last_of("/\\")) + "/";
    std::map<std::string, int> materials;
    scene.AddMaterial(Material());  // default material, id 0
    char line[kLine];
    bool ok = true;
    while (ReadLine(f, line)) {
        if (line[0] == '#') continue;
        char name[kLine] = {0};
        if (std::sscanf(line, " material %2047s", name) == 1) {
            Material mat;
            ReadMaterial(f, line, scene, dir, mat);
            if (materials.find(name) == materials.end()) materials[name] = scene.AddMaterial(mat);
        }
        if (std::strstr(line, "light")) ReadLight(f, line, scene);
        if (std::strstr(line, "camera")) ReadCamera(f, line, scene);
        if (std::strstr(line, "renderer")) ReadRenderer(f, line, scene, dir, options);
        if (std::strstr(line, "mesh")) ReadMesh(f, line, scene, dir, materials, log);
        if (std::strstr(line, "gltf")) {
            std::string gltf;
            while (ReadLine(f, line)) {
                if (std::strchr(line, '}')) break;
                char file[kLine];
                if (std::sscanf(line, " file %2047s", file) == 1) gltf = dir + file;
            }
            if (!gltf.empty()) {
                log += "glTF scenes are not supported: " + gltf + "\n";
                ok = false;
                break;
            }
        }
    }
    std::fclose(f);
    scene.texArrayWidth = options.texArrayWidth;
    scene.texArrayHeight = options.texArrayHeight;
    return ok;
}

}  // namespace glslpt
