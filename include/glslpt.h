/*
 * glslpt — C ABI of the MI355X path-tracing backend.
 *
 * Two libraries sit behind this header:
 *   libglslpt_scene.so  host scene surface + BVH build + flattening (no GPU)
 *   libglslpt_hip.so    HIP renderer for gfx950 (wavefront and plain kernels)
 *
 * Conventions: every entry point returns PT_OK (0) or a negative PT_ERR_*;
 * no C++ exception crosses the ABI; the last error message of a handle is
 * returned by pt_last_error / pts_last_error. Arrays are plain host pointers
 * with explicit counts; the caller keeps ownership of everything it passes.
 * One host thread per handle; each renderer handle owns one HIP stream.
 *
 * Reference interfaces replaced (file:line into knightcrawler25/GLSL-PathTracer):
 *   pts_*   GLSLPT::Scene        src/core/Scene.h:74-125, src/core/Scene.cpp:56-297
 *   pt_*    GLSLPT::Renderer     src/core/Renderer.h:104-185, src/core/Renderer.cpp:135-812
 */
#ifndef GLSLPT_H
#define GLSLPT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PT_OK 0
#define PT_ERR_INVALID_ARG (-1)
#define PT_ERR_HIP (-2)
#define PT_ERR_STATE (-3)
#define PT_ERR_OVERFLOW (-4)
#define PT_ERR_IO (-5)
#define PT_ERR_UNSUPPORTED (-6)
#define PT_ERR_NO_DEVICE (-7)

/* ------------------------------------------------------------------ records */

/* 32 floats = 8 x vec4, the GPU material record (src/core/Material.h:47-120). */
typedef struct pt_material {
    float baseColor[3]; float anisotropic;
    float emission[3]; float padding1;
    float metallic, roughness, subsurface, specularTint;
    float sheen, sheenTint, clearcoat, clearcoatGloss;
    float specTrans, ior, mediumType, mediumDensity;
    float mediumColor[3]; float mediumAnisotropy;
    float baseColorTexId, metallicRoughnessTexID, normalmapTexID, emissionmapTexID;
    float opacity, alphaMode, alphaCutoff, padding2;
} pt_material;

/* 15 floats = 5 x vec3 (src/core/Scene.h:50-59). type: 0 quad, 1 sphere, 2 distant. */
typedef struct pt_light {
    float position[3]; float emission[3]; float u[3]; float v[3];
    float radius, area, type;
} pt_light;

/* Camera uniforms read by the ray generator (src/shaders/common/globals.glsl:83-92);
 * fov is horizontal, in radians. */
typedef struct pt_camera {
    float position[3]; float up[3]; float right[3]; float forward[3];
    float fov, focalDist, aperture;
} pt_camera;

/* Flattened BVH node (thirdparty/RadeonRays/bvh_translator.h:45-50).
 * LRLeaf: internal (left, right, 0); BLAS leaf (first vertIndices row, count, 1);
 * TLAS leaf (BLAS root node, material id, -instance-1). */
typedef struct pt_bvh_node { float bboxmin[3]; float bboxmax[3]; float LRLeaf[3]; } pt_bvh_node;

/* Exactly the arrays Renderer::InitGPUDataBuffers uploads
 * (src/core/Renderer.cpp:135-249) plus the uniforms derived from the scene. */
typedef struct pt_scene_view {
    const pt_bvh_node* nodes; int32_t num_nodes; int32_t top_level_index;
    const int32_t* vert_indices; int32_t num_tris;           /* 3 ints per row */
    const float* vertices_uvx; const float* normals_uvy;       /* 4 floats per vertex */
    int32_t num_vertices;
    const pt_material* materials; int32_t num_materials;
    const float* transforms; int32_t num_instances;            /* Mat4 data[4][4], column per data[i] */
    const pt_light* lights; int32_t num_lights;
    const uint8_t* texture_maps; int32_t tex_width, tex_height, num_textures; /* RGBA8 layers */
    const float* env_img; const float* env_cdf;                /* RGB32F image, R32F running-sum CDF
                                                                  (NULL: built on the device) */
    int32_t env_width, env_height; float env_total_sum;
} pt_scene_view;

/* ------------------------------------------------------------ scene builder */

typedef struct pts_scene_s* pts_scene;

int pts_create(pts_scene* out);
int pts_destroy(pts_scene s);
const char* pts_last_error(pts_scene s);
/* Scene::AddMesh(filename) — OBJ triangle soup (src/core/Scene.cpp:62-83). */
int pts_add_mesh_obj(pts_scene s, const char* filename, int32_t* id);
/* Scene::AddMesh from arrays: positions[9n], normals[9n] or NULL, uvs[6n] or NULL. */
int pts_add_mesh(pts_scene s, const float* positions, const float* normals, const float* uvs,
                 int32_t num_tris, const char* name, int32_t* id);
int pts_add_material(pts_scene s, const pt_material* m, int32_t* id);      /* Scene.cpp:109-114 */
/* transform: 16 floats, Mat4 data[4][4] (column per 4 floats) (Scene.cpp:134-139). */
int pts_add_mesh_instance(pts_scene s, int32_t mesh_id, const float* transform, int32_t material_id,
                          int32_t* id);
int pts_add_light(pts_scene s, const pt_light* l, int32_t* id);            /* Scene.cpp:141-146 */
int pts_add_texture(pts_scene s, const uint8_t* rgba, int32_t width, int32_t height, const char* name,
                    int32_t* id);                                          /* Scene.cpp:85-107 */
int pts_add_camera(pts_scene s, const float eye[3], const float lookat[3], float fov_degrees); /* :56-60 */
int pts_add_env_map(pts_scene s, const float* rgb, int32_t width, int32_t height);            /* :116-132 */
int pts_set_texture_array_size(pts_scene s, int32_t width, int32_t height);
int pts_process(pts_scene s);                                              /* Scene.cpp:216-297 */
int pts_rebuild_instances(pts_scene s);                                    /* Scene.cpp:200-214 */
int pts_set_instance_transform(pts_scene s, int32_t instance, const float* transform);
int pts_get_view(pts_scene s, pt_scene_view* out);  /* pointers valid until the next mutation */
int pts_get_camera(pts_scene s, pt_camera* out);
int pts_camera_offset_orientation(pts_scene s, float dx, float dy);       /* Camera.cpp:152-157 */
int pts_get_blas_root(pts_scene s, int32_t mesh_id, int32_t* root);
/* Texture `id` as added (before the resize into the texture array): size, and
 * its RGBA8 pixels when rgba is not NULL (width*height*4 bytes). */
int pts_get_texture(pts_scene s, int32_t id, int32_t* width, int32_t* height, uint8_t* rgba);

/* RenderOptions (src/core/Renderer.h:37-100) as a .scene file sets them. */
typedef struct pts_render_options {
    int32_t render_resolution[2], window_resolution[2];
    float uniform_light_col[3], background_col[3];
    int32_t tile_width, tile_height, max_depth, max_spp, rr_depth;
    int32_t tex_array_width, tex_array_height, denoiser_frame_cnt;
    int32_t enable_rr, enable_denoiser, enable_tonemap, enable_aces, simple_aces_fit, opengl_normal_map;
    int32_t enable_envmap, enable_uniform_light, hide_emitters, enable_background, transparent_background;
    int32_t independent_render_size, enable_roughness_mollification, enable_volume_mis;
    float envmap_intensity, envmap_rot, roughness_mollification_amt;
} pts_render_options;
/* RenderOptions() defaults (src/core/Renderer.h:39-72). */
void pts_default_render_options(pts_render_options* o);
/* LoadSceneFromFile (src/loaders/Loader.cpp:36-517): parses a .scene file into
 * the scene (materials, lights, camera, OBJ meshes and instances, PNG
 * textures, HDR environment map) and *options (in/out: start from
 * pts_default_render_options). The texture-array size of the scene is taken
 * from the options. Missing meshes, textures or maps are skipped as in the
 * reference (reported by pts_last_error after PT_OK); glTF blocks return
 * PT_ERR_UNSUPPORTED. Call pts_process afterwards. */
int pts_load_scene_file(pts_scene s, const char* filename, pts_render_options* options);

/* ------------------------------------------------------------------ renderer */

#define PT_MODE_WAVEFRONT 0
#define PT_MODE_PLAIN 1

/* RenderOptions subset read by the path (src/core/Renderer.h:37-100) plus the
 * backend's launch configuration. */
typedef struct pt_options {
    int32_t width, height;                 /* renderResolution */
    int32_t max_depth;                     /* maxDepth */
    int32_t enable_rr, rr_depth;           /* enableRR, RRDepth */
    int32_t enable_uniform_light; float uniform_light_col[3];
    int32_t enable_envmap; float envmap_intensity; float envmap_rot;   /* rot in degrees */
    int32_t hide_emitters;
    int32_t enable_background, transparent_background;
    int32_t enable_roughness_mollification; float roughness_mollification_amt;
    int32_t enable_volume_mis;
    int32_t opengl_normal_map;
    /* backend launch configuration (no reference counterpart) */
    int32_t mode;            /* PT_MODE_WAVEFRONT or PT_MODE_PLAIN */
    int32_t block_size;      /* 64, 128 or 256 */
    int32_t stack_depth;     /* LDS traversal-stack entries per thread: 16, 24 or 32 */
    int32_t wave_spp;        /* samples per wavefront pass */
    int32_t use_graph;       /* capture the bounce stages of a pass in a HIP graph */
    int32_t count_work;      /* instrumented kernels: count nodes / triangles / instances */
    int32_t compact_every;   /* compact the path queue every k-th bounce (1 = every bounce) */
    int32_t bvh_width;       /* 2: BLAS traversed as the reference's BVH2; 4: as four-wide nodes
                                collapsed from it (same hits but for exact ties) */
    int32_t nee_inline;      /* wavefront and batched forms: 1 = the shade stage traces its NEE rays
                                itself (no shadow stage); 0 = a shadow stage does, except where the
                                scene needs them inline (stochastic alpha, volume MIS). Same image
                                either way; pt_adjoint always uses the shadow stage */
    int32_t extend_cap;      /* wavefront and batched forms, bounces after the first: a closest-hit
                                traversal stops after this many steps and a second, densely packed
                                launch finishes the stopped ones, so waves do not wait on their
                                slowest rays (0 = off). Same hits, same image */
} pt_options;

typedef struct pt_stats {
    double ms_last_render;           /* device time of the last pt_render / pt_batch_render */
    double ms_stage[8];              /* per-stage device time accumulated since pt_reset_stats */
    int64_t launches_stage[8];
    /* count_work only: traversal work of closest-hit queries (extend stage, or
     * everything in the plain kernel) and of shadow queries (shadow stage) */
    int64_t rays_closest, nodes_visited, tris_tested, instances_entered, lights_tested;
    int64_t rays_shadow, shadow_nodes, shadow_tris, shadow_instances, shadow_lights;
    int64_t stack_spills;            /* traversal-stack entries that went to the global backing store */
    int32_t stack_overflow;          /* 1 if a traversal exceeded the maximum stack depth */
    int32_t adjoint_dropped;         /* last pt_adjoint: path vertices whose gradient had a non-finite
                                        term and was left out */
    /* count_work, wavefront form only: closest hits that entered the shade
     * stage at bounce b (b < 15; entry 15 sums bounces 15 and later), over
     * every pass since pt_reset_stats. hits[0] / samples is the fraction of
     * camera rays that hit; the sum / samples the mean path length in vertices. */
    int64_t bounce_hits[16];
} pt_stats;

/* stage indices of pt_stats.ms_stage */
#define PT_STAGE_RAYGEN 0
#define PT_STAGE_EXTEND 1
#define PT_STAGE_SHADE 2
#define PT_STAGE_SHADOW 3
#define PT_STAGE_ACCUM 4
#define PT_STAGE_PLAIN 5
#define PT_STAGE_ADJOINT 6   /* replay shade stage of pt_adjoint */
#define PT_STAGE_BATCH 7     /* whole pass sequence of pt_batch_render */

typedef struct pt_renderer_s* pt_handle;

void pt_default_options(pt_options* o);
int pt_device_count(int32_t* n);
int pt_create(int32_t device, pt_handle* out);                 /* Renderer::Renderer (Renderer.cpp:41-84) */
int pt_destroy(pt_handle h);                                   /* Renderer::~Renderer (:86-133) */
const char* pt_last_error(pt_handle h);
int pt_upload_scene(pt_handle h, const pt_scene_view* v);      /* InitGPUDataBuffers (:135-249) */
/* Instance edit refresh: transforms, materials and the TLAS node range
 * [first_node, first_node + num_nodes) (Renderer.cpp:649-665). */
int pt_update_instances(pt_handle h, const float* transforms, int32_t num_instances,
                        const pt_material* materials, int32_t num_materials,
                        const pt_bvh_node* nodes, int32_t first_node, int32_t num_nodes);
/* Instance edit with the TLAS rebuilt on the device: Scene::RebuildInstances
 * (src/core/Scene.cpp:200-214; world boxes :150-184; Bvh(10, 64, false) median
 * build, thirdparty/RadeonRays/bvh.cpp:68-243; UpdateTLAS,
 * bvh_translator.cpp:108-114) followed by the instancesModified refresh
 * (src/core/Renderer.cpp:649-665). Only the transforms (16 floats per instance,
 * column-major Mat4) and, if not NULL, the instance material ids cross to the
 * device; the TLAS nodes and instance records are written there, identical to
 * pts_rebuild_instances followed by pt_update_instances with the TLAS range.
 * Requires the uploaded TLAS to hold one instance per leaf, as the reference
 * builds it (else PT_ERR_UNSUPPORTED). Synchronous. */
int pt_rebuild_instances(pt_handle h, const float* transforms, int32_t num_instances,
                         const int32_t* material_ids);
/* Copies the device TLAS: its num_instances - 1 internal nodes (16 floats each:
 * left box min xyz, max xyz, right box min xyz, max xyz, left and right
 * reference as int bits, 2 unused), the instance records (24 floats each) and
 * the root reference. Any output may be NULL. */
int pt_read_tlas(pt_handle h, float* nodes, int32_t num_nodes, float* instances, int32_t num_instances,
                 int32_t* root);
/* Copies the four-wide TLAS that pt_options.bvh_width = 4 traverses: `room`
 * entries of 32 floats (children's boxes as min x, max x, min y, max y, min z,
 * max z rows of four, then four references as int bits: internal ones index
 * the renderer's four-wide node array, in which the first entry copied is the
 * root, leaves are -1 - instance, empty INT32_MIN; 4 unused), breadth first
 * from the root, unused entries zero; room is at least num_instances - 1.
 * root gets the root's index (or -1 - instance when the TLAS is one
 * instance). nodes may be NULL to query room. No counterpart in the reference
 * (its TLAS is binary): the inspection hook of this form, beside pt_read_tlas. */
int pt_read_tlas4(pt_handle h, float* nodes, int32_t num_nodes, int32_t* room, int32_t* root);
int pt_set_options(pt_handle h, const pt_options* o);          /* InitShaders defines + uniforms */
/* Replaces the environment map of the uploaded scene: uploads the RGB32F image
 * (width*height*3 floats, rows top-down as EnvironmentMap::LoadMap reads them)
 * and builds its luminance CDF and total on the device, bit-identical to
 * EnvironmentMap::BuildCDF (src/core/EnvironmentMap.cpp:39-61) and to
 * pts_add_env_map; envMapModified refresh (src/core/Renderer.cpp:666-678). */
int pt_set_env_map(pt_handle h, const float* rgb, int32_t width, int32_t height);
/* The device CDF of the current environment map (count = width*height floats)
 * and its total luminance. */
int pt_read_env_cdf(pt_handle h, float* cdf, size_t count, float* total);
/* pt_set_env_map's CDF build for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference. rgb: width*height*3 floats in
 * device memory (rows top-down); cdf: width*height floats and total: one float,
 * both in device memory, receive the CDF and its total luminance, bit-identical
 * to pt_set_env_map's. The work is enqueued on `stream` (a hipStream_t; NULL =
 * the null stream) on the current device and nothing is allocated: the scan's
 * scratch is a fixed device array of the library (~74 KB, one per GPU, held
 * until the process ends), shared by every call on that GPU, so calls that
 * overlap on one GPU race. At most PT_ENV_NOALLOC_MAX_TEXELS texels
 * (PT_ERR_OVERFLOW beyond). Errors are reported by pt_last_error(NULL). */
#define PT_ENV_NOALLOC_MAX_TEXELS (1 << 25)
int pt_env_cdf_noalloc(const float* rgb, int32_t width, int32_t height, float* cdf, float* total, void* stream);
int pt_set_camera(pt_handle h, const pt_camera* c);            /* camera uniforms (:764-771) */
int pt_clear_accum(pt_handle h);                               /* accumFBO clear (:733-744) */
/* Adds samples [first_sample, first_sample + num_samples) of every pixel to the
 * RGBA32F accumulation buffer; sample s uses RNG seed (x, y, s + 2, x + y). */
int pt_render(pt_handle h, int32_t first_sample, int32_t num_samples);
int pt_synchronize(pt_handle h);
int pt_read_accum(pt_handle h, float* rgba, size_t num_floats); /* GetOutputBuffer analogue */
/* Tonemapped, gamma-corrected RGBA8, rows bottom-up like the GL texture
 * (tonemap.glsl:97-133, Renderer.cpp:619-634). */
int pt_read_ldr(pt_handle h, int32_t num_samples, int32_t enable_tonemap, int32_t enable_aces,
                int32_t simple_aces_fit, const float background[3], uint8_t* rgba, size_t num_bytes);
int pt_get_stats(pt_handle h, pt_stats* out);
int pt_reset_stats(pt_handle h);

/* Ray queries against the uploaded scene (ClosestHit closest_hit.glsl:25-266,
 * AnyHit anyhit.glsl:25-215), host arrays in and out.
 * closest: t (INF on miss), instance (-1 if none), triangle = vertIndices row
 * (-1 if none), light index (-1 if none), barycentrics (u, v). */
int pt_closest_hit(pt_handle h, const float* origins, const float* dirs, int32_t n,
                   float* t, int32_t* instance, int32_t* triangle, int32_t* light, float* uv);
int pt_any_hit(pt_handle h, const float* origins, const float* dirs, const float* max_dist,
               int32_t n, int32_t* occluded);

/* ---- batched uneven jobs: one launch sequence over many render jobs ---- */
/* Renders every job (own camera, resolution, samples [0, spp) and max depth;
 * scene and the other options shared). Wavefront mode: passes that mix jobs;
 * the path slots of a pass are (job, sample, pixel) planes, packed up to
 * wave_spp samples worth of slots of the largest job, and with
 * pt_options.use_graph the stage launches of one pass are captured once in a
 * HIP graph launched once per pass. Plain mode: one persistent launch of the
 * plain kernel over every job's pixels, each thread tracing one pixel's
 * samples in order. Either way each job's image is identical to rendering it
 * alone with pt_render (same seeds, same order of accumulation). No reference
 * counterpart: the reference renders one scene view at a time
 * (Renderer::Render, src/core/Renderer.cpp:546-590). */
typedef struct pt_job {
    pt_camera camera;
    int32_t width, height, spp, max_depth;
    float* accum_out;        /* host RGBA32F buffer, width*height*4 floats: the sum over the job's
                                samples, written (not added), rows bottom-up */
} pt_job;
int pt_batch_render(pt_handle h, const pt_job* jobs, int32_t num_jobs);

/* ---- adjoint: gradient of an image loss w.r.t. material parameters ---- */
/* Path-replay backpropagation over samples [first_sample, first_sample +
 * num_samples) with the current scene, options and camera (no reference
 * counterpart; the forward path it differentiates is PathTrace,
 * pathtrace.glsl:285-476). Parameters per material, in this order:
 * baseColor.r, .g, .b (record floats 0-2), roughness (9), metallic (8),
 * ior (17), clearcoat (14). dL_dimage is dLoss/d(mean image) where the mean
 * image is the accumulation of those samples divided by num_samples: RGB
 * float per pixel (width*height*3, rows bottom-up like pt_read_accum).
 * d_params receives num_materials*7 floats, added to (not overwritten).
 * Sampling, MIS weights and Russian roulette are detached: the estimate is
 * unbiased for the gradient of the expected image. Runs the wavefront form
 * whatever pt_options.mode says: a primal pass (the forward render, which
 * also records which NEE rays were visible) and a replay pass that re-traces
 * the same paths and accumulates the gradients. Returns PT_ERR_UNSUPPORTED
 * with roughness mollification, stochastic (blend) alpha in shadow rays or
 * volume MIS enabled; PT_ERR_STATE if a path vertex names a material outside
 * the table. Vertices whose gradient has a non-finite term are left out and
 * counted in pt_stats.adjoint_dropped. */
#define PT_ADJOINT_PARAMS 7
int pt_adjoint(pt_handle h, int32_t first_sample, int32_t num_samples, const float* dL_dimage,
               float* d_params);

/* As pt_adjoint, plus the gradient with respect to each material's medium
 * (no reference counterpart; the forward path it differentiates is the
 * medium branch of PathTrace, pathtrace.glsl:370-410, and the medium a path
 * enters at :457-465): per material, in this order, medium density (record
 * float 19) and medium colour r, g, b (20-22), added to d_medium
 * (num_materials*4 floats). Free-flight distances of scattering media are
 * drawn with the current density and held fixed; the estimate includes their
 * score (1/density - s for a scatter at distance s, -d for a crossing of
 * length d), so it is unbiased for the gradient of the expected image. A
 * medium entry is credited to the material whose surface the path crossed to
 * enter it. Same errors as pt_adjoint; medium segments with a non-finite term
 * are left out and counted in pt_stats.adjoint_dropped. */
#define PT_ADJOINT_MEDIUM_PARAMS 4
int pt_adjoint_medium(pt_handle h, int32_t first_sample, int32_t num_samples, const float* dL_dimage,
                      float* d_params, float* d_medium);

#ifdef __cplusplus
}
#endif

#endif /* GLSLPT_H */
