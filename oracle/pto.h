/*
 * ORACLE — test infrastructure only.
 *
 * Plain-C restatement of vkPathTracer's scalar CPU path tracer (the reference's
 * correctness path). Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it, and only as the checker or the timed CPU baseline;
 * the product never links it.
 *
 * Built three ways by oracle/Makefile:
 *   liboracle_f32.so       fp32, portable transcendentals (pto_math.h)  — parity anchor for the GPU
 *   liboracle_f32_libm.so  fp32, glibc sinf/cosf/powf, i.e. the reference's own math
 *   liboracle_f64.so       the same estimator in fp64 (libm double)    — convergence scoring
 *
 * The record layouts below are byte-identical to include/vkpt_hip.h (checked by
 * tests/test_capi_layout.py) so one set of host arrays feeds both.
 */
#ifndef PTO_H
#define PTO_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* RTMaterial (reference src/pathtracer/PathTracer.h:242-261), 20 x 4 bytes. */
typedef struct {
  float albedo[3];
  float emissive[3];
  float roughness, metallic, ior, transmission, clearcoat, sheen, anisotropy, alpha;
  uint32_t material_model, material_effect, material_flags;
  uint32_t base_color_texture_index, normal_texture_index, pad0;
} PtoMaterial;

/* RTHitLight (PathTracer.h:341-349), 16 x 4 bytes. */
typedef struct {
  float position[3];
  float color[3];
  float intensity, radius;
  float direction[3];
  float spot_inner_cos, spot_outer_cos;
  float pad0, pad1, pad2;
} PtoLight;

/* RTInstance (PathTracer.h:277-296), 24 x 4 bytes. */
typedef struct {
  uint64_t entity_id;
  uint32_t geometry_id, first_triangle, triangle_count, material_index, flags, transform_revision;
  uint32_t local_first_vertex, local_vertex_count, local_first_index, local_index_count;
  float translation[3];
  float rotation[4]; /* quaternion x, y, z, w */
  float scale[3];
  uint32_t pad0, pad1;
} PtoInstance;

typedef struct {
  float position[3], target[3], up[3];
  float fov_deg, aperture_radius, focus_distance;
  uint32_t iris_blade_count;
  float iris_rotation_degrees, iris_roundness, anamorphic_squeeze;
} PtoCamera;

/* Flattened view of PathTracerSceneSnapshot (PathTracer.h:358-409). */
typedef struct {
  const float* vertices; /* xyz */
  const uint32_t* indices;
  const float* local_vertices;
  const uint32_t* local_indices;
  const PtoInstance* instances;
  const PtoMaterial* materials;
  const PtoLight* lights;
  const float* environment_map; /* rgb, width*height */
  uint32_t vertex_count, index_count, local_vertex_count, local_index_count;
  uint32_t instance_count, material_count, light_count;
  uint32_t environment_map_width, environment_map_height;
  float environment_color[3];
  float environment_map_scale[3];
  PtoCamera camera;
} PtoScene;

/* RenderSettings subset that the hot path reads (PathTracer.h:134-178). */
typedef struct {
  uint64_t seed;
  uint32_t width, height, spp, max_depth;
  uint32_t enable_nee, enable_mis, deterministic;
  uint32_t rr_start_depth;
  float rr_min_survival, rr_max_survival;
  float camera_aperture_radius, camera_focus_distance;
} PtoSettings;

/* SampleCounters (PathTracer.h:218-230), same field order. */
typedef struct {
  uint64_t samples, rays, triangle_tests, sdf_tests, sdf_steps, triangle_hits;
  uint64_t sdf_hits, sdf_misses, bvh_node_visits, bvh_leaf_visits, shadow_tests;
} PtoCounters;

/* Texture record (include/vkpt_hip.h PtTextureDesc): width x height rgb texels from texel `offset`. */
typedef struct {
  uint32_t offset, width, height, pad;
} PtoTextureDesc;

enum { PTO_BUILDER_MIDPOINT = 0, PTO_BUILDER_SAH = 1 };
enum { PTO_SHADOW_CLOSEST = 0, PTO_SHADOW_ANYHIT = 1 };

/* SDF primitive (include/vkpt_hip.h PtSdfPrimitive; RTSdfPrimitive, PathTracer.h:330-339). */
typedef struct {
  uint32_t shape;
  float position[3], scale[3], rotation[3];
  uint32_t material_index;
  float radius, param_a, param_b;
} PtoSdf;

typedef struct pto_accel pto_accel;

uint32_t pto_real_bytes(void);
uint32_t pto_uses_libm(void);
uint32_t pto_sizeof(int which); /* 0 material, 1 light, 2 instance, 3 camera, 4 scene, 5 settings, 6 counters,
                                   7 texture */

/* Math entry points (portable or libm, per build). */
float pto_sinf(float x);
float pto_cosf(float x);
float pto_powf(float x, float y);
float pto_atan2f(float y, float x);
float pto_acosf(float x);

/* splitmix64 and the Rng of ScalarCpuPathTracer.cpp:36-41, :383-393. */
uint64_t pto_splitmix64(uint64_t x);
uint64_t pto_rng_seed(uint64_t seed, uint64_t frame, uint64_t sample, uint64_t path_id,
                      uint64_t dimension, uint64_t depth, uint64_t pixel);
void pto_rng_draws(uint64_t state, uint32_t n, float* out, uint64_t* out_state);

/* Acceleration structure over the flattened instance triangles. */
pto_accel* pto_accel_build(const PtoScene* scene, int builder, uint32_t leaf_size, uint32_t bucket_count);
void pto_accel_free(pto_accel* a);
/* Albedo textures (no counterpart in the reference's CPU path): texcoords are 2 floats per scene
 * vertex; a material whose base_color_texture_index is below texture_count has its albedo
 * multiplied by the bilinear tap at the hit. Copied. Returns 0, or -1 for a texture outside the
 * texels. */
int pto_accel_set_textures(pto_accel* a, const float* texcoords, uint32_t texcoord_count, const float* texels,
                           uint64_t texel_count, const PtoTextureDesc* textures, uint32_t texture_count);
/* The bilinear tap of texture `tex` at barycentrics (bu, bv) of flattened primitive `prim`: colour,
 * the four texel indices and their weights. Returns -1 without textures or for bad indices. */
int pto_texture_sample(const pto_accel* a, uint32_t tex, uint32_t prim, float bu, float bv, double* out_rgb,
                       uint32_t* out_idx, double* out_w);
uint32_t pto_accel_node_count(const pto_accel* a);
uint32_t pto_accel_prim_count(const pto_accel* a);
uint32_t pto_accel_max_depth(const pto_accel* a);
/* bounds: 6 reals per node (min xyz, max xyz, padded); meta: left, right, first, count. */
int pto_accel_export(const pto_accel* a, void* bounds, int32_t* meta, uint32_t* prim_order,
                     void* prim_v0e1e2n /* 12 reals per primitive */, uint32_t* prim_material);

/* rays: 6 floats each (origin, direction). tmax may be NULL (infinity). */
int pto_trace_rays(const pto_accel* a, const float* rays, uint32_t n, int any_hit, const float* tmax,
                   float* out_t, int32_t* out_prim, PtoCounters* counters);
int pto_brute_force_rays(const pto_accel* a, const float* rays, uint32_t n, float* out_t, int32_t* out_prim);

/* One sample for each listed pixel (shade_pixel, ScalarCpuPathTracer.Render.cpp:228-284).
 * out_rgb: 3 reals per pixel; out_valid: 1 if finite. */
int pto_render_pixels(const PtoScene* scene, const pto_accel* a, const PtoSettings* s,
                      const uint32_t* pixels, uint32_t n, uint32_t sample_index, uint64_t frame_index,
                      int threads, int shadow_mode, void* out_rgb, uint8_t* out_valid, PtoCounters* counters);

/* Full tile loop over samples [sample_begin, sample_begin+sample_count): the film after
 * render_tile() for each sample in order. film_sum: 3 reals per pixel; film_count: u32. */
int pto_render_tile(const PtoScene* scene, const pto_accel* a, const PtoSettings* s,
                    uint32_t x0, uint32_t y0, uint32_t w, uint32_t h,
                    uint32_t sample_begin, uint32_t sample_count, uint64_t frame_index,
                    int threads, int shadow_mode, void* film_sum, uint32_t* film_count, PtoCounters* counters);

/* Path-replay adjoint of one sample per listed pixel (no reference counterpart; see pto.c).
 * delta: dLoss/d(path radiance), 3 floats per pixel. Gradients accumulate (+=) in double. */
int pto_adjoint_pixels(const PtoScene* scene, const pto_accel* a, const PtoSettings* s,
                       const uint32_t* pixels, uint32_t n, uint32_t sample_index, uint64_t frame_index,
                       int threads, int shadow_mode, const float* delta, double* g_albedo, double* g_emission,
                       double* g_light, double* g_rough, double* g_tex, void* out_rgb);

/* Glossy-branch direction samples and their d log(density)/d exponent (see pto.c). */
int pto_phong_samples(float exponent, const float refl[3], const float n[3], float ps, uint64_t state,
                      uint32_t count, float* out_dirs, double* out_score);

/* pto_trace_rays' traversal, counting the triangle tests the |det| < 1e-4 cut-off rejects:
 * out[0] = triangle tests, out[1] = rejections. */
int pto_det_rejections(const pto_accel* a, const float* rays, uint32_t n, int any_hit, uint64_t out[2]);
/* render_tile with TiledCpuPathTracer's band jobs (see pto.c); the CPU baseline's schedule. */
int pto_render_bands(const PtoScene* scene, const pto_accel* a, const PtoSettings* s, uint32_t band_rows,
                     uint32_t band_first, uint32_t band_stride, uint32_t sample_begin, uint32_t sample_count,
                     uint64_t frame_index, int threads, int shadow_mode, void* film_sum, uint32_t* film_count,
                     PtoCounters* counters);

/* Probes for closed-form checks (see pto.c): the primary rays and pixel jitter shade_pixel uses, the
 * pinhole direction at a given jitter, cosine-hemisphere samples, and light_pdf_for_direction. */
int pto_camera_rays(const PtoScene* scene, const PtoSettings* s, const uint32_t* pixels, uint32_t n,
                    uint32_t sample, uint64_t frame, float* out_rays, float* out_jitter);
int pto_camera_dir(const PtoScene* scene, const PtoSettings* s, uint32_t x, uint32_t y, float u, float v,
                   float out_dir[3]);
int pto_hemisphere_samples(const float n[3], uint64_t state, uint32_t count, float* out_dirs);
double pto_light_pdf(const PtoScene* scene, const float pos[3], const float dir[3]);

/* The scene's SDF primitives (intersect_scene tests them after the triangles; Unknown shapes dropped). */
int pto_accel_set_sdfs(pto_accel* a, const PtoSdf* sdfs, uint32_t count);
/* intersect_scene (ScalarCpuPathTracer.Intersect.cpp:28-132) for rays (o xyz, d xyz): t (inf: miss),
 * normal, material index, SDF flag; counters accumulate. */
int pto_intersect_scene_rays(const PtoScene* sc, const pto_accel* a, const float* rays, uint32_t n, float* out_t,
                             float* out_normal, uint32_t* out_material, uint8_t* out_sdf, PtoCounters* c);

/* CpuBvhAccelerator's dynamic-instance refit (CpuBvhAccelerator.cpp:310-358, :527-603): scene holds the
 * updated instance transforms; primitives are recomputed and node boxes refit bottom-up, the topology
 * kept. Returns -2 when the primitive count changed (a rebuild is needed). */
int pto_accel_refit(pto_accel* a, const PtoScene* scene);

/* FilmResolveSettings (PathTracer.h:67-74) after the host-side conversions: white-balance scale
 * (WhiteBalanceScale) and inverse gamma 1 / max(0.01, gamma). tone_map: 0 linear, 1 reinhard,
 * 2 filmic_approx, 3 aces_approx (ToneMapMode); output_transform: 0 gamma, 1 linear. */
typedef struct PtoResolve {
  float exposure;
  float white_balance[3];
  float inv_gamma;
  uint32_t tone_map, output_transform, clamp_output;
} PtoResolve;

/* ColorTemperatureToRgb / WhiteBalanceScale (FilmBuffer.cpp:30-49, :216-224), libm. */
void pto_color_temperature_rgb(float kelvin, float out[3]);
void pto_white_balance_scale(float kelvin, float out[3]);
/* FilmBuffer::resolve_hdr (FilmBuffer.cpp:160-214): sum3 holds 3 floats per pixel. */
void pto_resolve_hdr(const float* sum3, const uint32_t* count, uint64_t pixels, float* rgb);
/* ApplyFilmResolve (FilmBuffer.cpp:255-332) of resolve_hdr's image: 4 bytes (rgba) per pixel. */
void pto_resolve_ldr(const float* sum3, const uint32_t* count, uint64_t pixels, const PtoResolve* rs,
                     uint8_t* rgba);

#ifdef __cplusplus
}
#endif

#endif /* PTO_H */
