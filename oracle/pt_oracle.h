/*
 * CPU oracle — TEST INFRASTRUCTURE, not product code.
 *
 * A C++ restatement of the reference's GLSL path (src/shaders/tile.glsl and
 * src/shaders/common/{globals,intersection,sampling,envmap,closest_hit,anyhit,
 * disney,pathtrace}.glsl) over the reference's flat scene arrays. tests/ and
 * __graft_entry__.smoke() compare the HIP renderer's results with it, and
 * bench.py times it as the CPU baseline ("port"). Templated on the scalar
 * (float for parity, double for the high-spp FP64 reference image).
 *
 * Pinning: see the header of pt_oracle.cpp for which check pins which function.
 */
#ifndef PT_ORACLE_H
#define PT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct oracle_scene {
    const float* nodes; int32_t num_nodes; int32_t top_level_index;  /* 9 floats per node */
    const int32_t* vert_indices; int32_t num_tris;
    const float* vertices_uvx; const float* normals_uvy; int32_t num_vertices;
    const float* materials; int32_t num_materials;                  /* 32 floats each */
    const float* transforms; int32_t num_instances;                 /* 16 floats each */
    const float* lights; int32_t num_lights;                        /* 15 floats each */
    const uint8_t* textures; int32_t tex_width, tex_height, num_textures;
    const float* env_img; const float* env_cdf; int32_t env_width, env_height; float env_total_sum;
} oracle_scene;

typedef struct oracle_camera {
    float position[3], up[3], right[3], forward[3];
    float fov, focalDist, aperture;
} oracle_camera;

/* OPT_* compile-time switches of the reference (src/core/Renderer.cpp:400-459)
 * and the uniforms that go with them (:498-543, :764-811). */
typedef struct oracle_options {
    int32_t width, height, max_depth;
    int32_t opt_lights, opt_rr, rr_depth;
    int32_t opt_uniform_light; float uniform_light_col[3];
    int32_t opt_envmap; float envmap_intensity; float envmap_rot;  /* envmap_rot = degrees / 360 */
    int32_t opt_hide_emitters, opt_background, opt_transparent_background;
    int32_t opt_alpha_test, opt_medium, opt_vol_mis;
    int32_t opt_roughness_mollification; float roughness_mollification_amt;
    int32_t opt_opengl_normalmap;
} oracle_options;

typedef struct oracle_counters {
    int64_t rays_closest, rays_shadow, nodes, tris, instances, lights;
} oracle_counters;

/* Adds samples [first_sample, first_sample + num_samples) of rows
 * row_begin, row_begin + row_step, ... < row_end into accum (RGBA, width*height*4,
 * rows bottom-up). precision 32 -> float path, accum_f32; 64 -> double path,
 * accum_f64. num_threads <= 0 uses all OpenMP threads. counters may be NULL. */
int oracle_render(int32_t precision, const oracle_scene* s, const oracle_camera* cam,
                  const oracle_options* o, int32_t first_sample, int32_t num_samples,
                  int32_t row_begin, int32_t row_end, int32_t row_step,
                  float* accum_f32, double* accum_f64, int32_t num_threads, oracle_counters* counters);

/* Frozen-sampling render (float path, all rows): samples, pdfs, MIS weights
 * and Russian roulette from s->materials, BSDF values f from eval_materials
 * (same layout). Used to check the path-replay adjoint: d(image)/d(eval
 * parameters) at eval == s->materials is what the adjoint estimates. */
int oracle_render_frozen(const oracle_scene* s, const float* eval_materials, const oracle_camera* cam,
                         const oracle_options* o, int32_t first_sample, int32_t num_samples, float* accum,
                         int32_t num_threads);

/* tonemap.glsl (gamma 2.2, luminance / ACES / simple ACES fit, background
 * and transparent-checker modes 0/1/2) plus glGetTexImage's float -> byte
 * conversion: accum RGBA32F (w*h*4) -> out RGBA8 (w*h*4). */
int oracle_tonemap(const float* accum, int32_t w, int32_t h, float inv_spp, int32_t enable_tonemap,
                   int32_t enable_aces, int32_t simple_aces, int32_t bg_mode, const float* bg, uint8_t* out);

/* Per-sample radiance (no accumulation): rgba[4 * (i)] for pixel list px/py,
 * sample index s[i]. */
int oracle_trace_samples(const oracle_scene* s, const oracle_camera* cam, const oracle_options* o,
                         int32_t n, const int32_t* px, const int32_t* py, const int32_t* sample, float* rgba);

/* Camera rays of tile.glsl:41-71 (float path): origin and direction of
 * sample[i] of pixel (px[i], py[i]), 3 floats each. */
int oracle_camera_rays(const oracle_camera* cam, const oracle_options* o, int32_t n, const int32_t* px,
                       const int32_t* py, const int32_t* sample, float* orig, float* dir);

/* ClosestHit / AnyHit queries (float path). Outputs as the GPU query. */
int oracle_closest_hit(const oracle_scene* s, const oracle_options* o, const float* orig, const float* dir,
                       int32_t n, float* t, int32_t* instance, int32_t* triangle, int32_t* light, float* uv,
                       oracle_counters* counters);
int oracle_any_hit(const oracle_scene* s, const oracle_options* o, const float* orig, const float* dir,
                   const float* max_dist, int32_t n, int32_t* occluded);

/* Instance inverse used by the traversal: 0 = per-visit inverse in the path's
 * scalar type (the shader's inverse(mat4)), 1 = double inverse rounded. */
void oracle_set_inverse_mode(int32_t mode);

/* Sampling and intersection primitives, n calls with packed arguments:
 *  0 GTR1(NdotH, a)  1 GTR2Aniso(NdotH, HdotX, HdotY, ax, ay)  2 SmithG(NdotV, a)
 *  3 SmithGAniso(NdotV, VdotX, VdotY, ax, ay)  4 DielectricFresnel(cos, eta)
 *  5 SchlickWeight(u)  6 PowerHeuristic(a, b)  7 SampleGGXVNDF(V3, ax, ay, r1, r2)->3
 *  8 SampleGTR1(rgh, r1, r2)->3  9 CosineSampleHemisphere(r1, r2)->3
 * 10 UniformSampleHemisphere(r1, r2)->3  11 SampleHG(V3, g, r1, r2)->3  12 PhaseHG(cos, g)
 * 13 SphereIntersect(rad, pos3, o3, d3)  14 RectIntersect(pos3, u3, v3, plane4, o3, d3)
 * 15 AABBIntersect(min3, max3, o3, d3)  16 sincos(x)->2 (the path's float sin, cos) */
int oracle_eval_fn(int32_t fn, int32_t n, const float* in, float* out);

/* RNG known-answer access: n successive pcg4d states from seed. */
void oracle_pcg4d(const uint32_t seed[4], int32_t n, uint32_t* states_out, float* rand_out);

/* BSDF unit access (float path). mat = 32-float material record, processed as
 * GetMaterial does without textures; eta < 0 means "derive from V and N". */
int oracle_disney_eval(const float* mat, float eta, const float* V, const float* N, const float* L,
                       float* f_out, float* pdf_out);
int oracle_disney_sample(const float* mat, float eta, const float* V, const float* N, uint32_t seed[4],
                         float* L_out, float* f_out, float* pdf_out);
/* Light sampling unit access: light = 15 floats; returns direction, normal,
 * emission, dist, pdf (13 floats) and advances seed. */
int oracle_sample_light(const float* light, int32_t num_lights, const float* scatter_pos, uint32_t seed[4],
                        float* rec_out);
/* Environment map lookups: EvalEnvMap (dir -> rgb, pdf) and SampleEnvMap. */
int oracle_eval_envmap(const oracle_scene* s, const oracle_options* o, const float* dir, float* rgb_pdf);
int oracle_sample_envmap(const oracle_scene* s, const oracle_options* o, uint32_t seed[4], float* dir_pdf,
                         float* rgb);

#ifdef __cplusplus
}
#endif

#endif
