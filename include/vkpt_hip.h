/*
 * vkpt_hip.h — C ABI of the MI355X-native path-tracing hot path.
 *
 * This is the drop-in boundary for vkPathTracer's CPU path tracer
 * (reference: src/pathtracer/PathTracer.h). Every entry point replaces one
 * IPathTracer / IRayAccelerator method, cited next to it. Plain C types only:
 * pointers, sizes and POD records; no exceptions cross this boundary and every
 * call returns a pt_status mirroring vkpt::core::StatusCode
 * (src/core/contracts/Result.h:13-23). pt_hip_last_error() returns the message
 * of the last failed call on that handle (IPathTracer::status().last_error,
 * PathTracer.h:854-865).
 *
 * Threading: one caller thread per handle (the reference's contract,
 * RenderCoordinator.cpp:1353-1391). Each handle owns one HIP stream at default
 * priority and all device memory it allocates. Calls are not re-entrant.
 *
 * Record layouts are fixed (4-byte fields, sizes asserted below) so that the
 * same host arrays can feed this library and the CPU oracle.
 */
#ifndef VKPT_HIP_H
#define VKPT_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* vkpt::core::StatusCode, src/core/contracts/Result.h:13-23. */
typedef enum pt_status {
  PT_OK = 0,
  PT_UNSUPPORTED = 1,
  PT_INVALID_ARGUMENT = 2,
  PT_NOT_READY = 3,
  PT_BUSY = 4,
  PT_ALLOC_FAILED = 5,
  PT_TIMEOUT = 6,
  PT_CANCELLED = 7,
  PT_INTERNAL_ERROR = 8
} pt_status;

/* PathTracerLifecycle (PathTracer.h:1052-1067 state machine). */
typedef enum pt_lifecycle {
  PT_LIFECYCLE_UNINITIALIZED = 0,
  PT_LIFECYCLE_CONFIGURED = 1,
  PT_LIFECYCLE_SCENE_LOADED = 2,
  PT_LIFECYCLE_READY = 3
} pt_lifecycle;

/* RTMaterial, PathTracer.h:242-261. 80 bytes. */
typedef struct PtMaterial {
  float albedo[3];
  float emissive[3];
  float roughness, metallic, ior, transmission, clearcoat, sheen, anisotropy, alpha;
  uint32_t material_model, material_effect, material_flags;
  uint32_t base_color_texture_index, normal_texture_index, pad0;
} PtMaterial;

/* RTHitLight, PathTracer.h:341-349. 64 bytes. */
typedef struct PtLight {
  float position[3];
  float color[3];
  float intensity, radius;
  float direction[3];
  float spot_inner_cos, spot_outer_cos;
  float pad0, pad1, pad2;
} PtLight;

/* RTInstance, PathTracer.h:277-296. 96 bytes. flags bit 0 = kRTInstanceFlagDynamicTransform. */
typedef struct PtInstance {
  uint64_t entity_id;
  uint32_t geometry_id, first_triangle, triangle_count, material_index, flags, transform_revision;
  uint32_t local_first_vertex, local_vertex_count, local_first_index, local_index_count;
  float translation[3];
  float rotation[4]; /* quaternion x, y, z, w */
  float scale[3];
  uint32_t pad0, pad1;
} PtInstance;

/* Camera fields of PathTracerSceneSnapshot, PathTracer.h:392-408. 64 bytes. */
typedef struct PtCamera {
  float position[3], target[3], up[3];
  float fov_deg, aperture_radius, focus_distance;
  uint32_t iris_blade_count;
  float iris_rotation_degrees, iris_roundness, anamorphic_squeeze;
} PtCamera;

/* Flattened PathTracerSceneSnapshot (PathTracer.h:358-409). Host pointers; the
 * library copies what it needs during pt_hip_load_scene and keeps no pointer. */
typedef struct PtSceneView {
  const float* vertices; /* xyz per vertex */
  const uint32_t* indices;
  const float* local_vertices;
  const uint32_t* local_indices;
  const PtInstance* instances;
  const PtMaterial* materials;
  const PtLight* lights;
  const float* environment_map; /* rgb, width*height (unsupported: must be NULL or 0x0) */
  uint32_t vertex_count, index_count, local_vertex_count, local_index_count;
  uint32_t instance_count, material_count, light_count;
  uint32_t environment_map_width, environment_map_height;
  float environment_color[3];
  float environment_map_scale[3];
  PtCamera camera;
} PtSceneView;

/* RenderSettings fields the hot path reads (PathTracer.h:134-178). 56 bytes. */
typedef struct PtRenderSettings {
  uint64_t seed;
  uint32_t width, height, spp, max_depth;
  uint32_t enable_nee, enable_mis, deterministic;
  uint32_t rr_start_depth;
  float rr_min_survival, rr_max_survival;
  float camera_aperture_radius, camera_focus_distance;
} PtRenderSettings;

/* FilmResolveSettings, PathTracer.h:67-74. 24 bytes. tone_map: ToneMapMode (0 linear, 1 reinhard,
 * 2 filmic_approx, 3 aces_approx); output_transform: OutputTransformMode (0 gamma, 1 linear). */
typedef struct PtFilmResolve {
  float exposure, white_balance_kelvin;
  uint32_t tone_map, output_transform;
  float gamma;
  uint32_t clamp_output;
} PtFilmResolve;

/* FilmReadbackToken, PathTracer.h:894-901 (only the latest request's token is valid), and
 * FilmReadbackState, PathTracer.h:903-909. */
typedef struct PtFilmReadbackToken {
  uint64_t id, fence_value;
  uint32_t width, height;
} PtFilmReadbackToken;
enum {
  PT_READBACK_INVALID = 0,
  PT_READBACK_PENDING = 1,
  PT_READBACK_READY = 2,
  PT_READBACK_FAILED = 3,
  PT_READBACK_UNSUPPORTED = 4
};

/* RTSdfPrimitive, PathTracer.h:330-339. 56 bytes. shape: SdfShape (0 sphere, 1 box, 2 rounded box,
 * 3 plane, 4 torus, 5 capsule, 6 unknown); rotation: Euler angles in radians (x, y, z). */
typedef struct PtSdfPrimitive {
  uint32_t shape;
  float position[3], scale[3], rotation[3];
  uint32_t material_index;
  float radius, param_a, param_b;
} PtSdfPrimitive;

/* RTInstanceTransformUpdate, PathTracer.h:298-306. 64 bytes. An instance_index out of range (e.g.
 * 0xFFFFFFFF) selects the instance whose entity_id matches. */
typedef struct PtInstanceTransformUpdate {
  uint64_t entity_id;
  uint32_t instance_index, flags, transform_revision;
  float translation[3];
  float rotation[4]; /* quaternion x, y, z, w */
  float scale[3];
  uint32_t pad0;
} PtInstanceTransformUpdate;

/* InstanceTransformPlan / InstanceTransformUpdateResult, PathTracer.h:565-615 (status values in
 * InstanceTransformUpdateStatus order). 24 bytes. */
typedef struct PtInstanceTransformResult {
  uint32_t status, requested_count, matched_count, applied_count;
  double refit_ms;
} PtInstanceTransformResult;
enum {
  PT_XFORM_APPLIED_METADATA_ONLY = 0,
  PT_XFORM_APPLIED_INSTANCE_BUFFER_ONLY = 1,
  PT_XFORM_APPLIED_DYNAMIC_ACCEL_UPDATE = 2,
  PT_XFORM_APPLIED_FULL_STATIC_ACCEL_REBUILD = 3,
  PT_XFORM_APPLIED_FULL_SCENE_RELOAD = 4,
  PT_XFORM_BLOCKED_NEEDS_FULL_STATIC_ACCEL_REBUILD = 5,
  PT_XFORM_BLOCKED_NEEDS_FULL_SCENE_RELOAD = 6,
  PT_XFORM_UNSUPPORTED = 7,
  PT_XFORM_FAILED = 8
};

/* RenderTile, PathTracer.h:919-927. One tile adds one sample (sample_index) to each pixel. */
typedef struct PtRenderTile {
  uint32_t x, y, width, height, sample_index, tile_id, gpu_id;
} PtRenderTile;

/* SampleCounters, PathTracer.h:218-230 (same field order). */
typedef struct PtSampleCounters {
  uint64_t samples, rays, triangle_tests, sdf_tests, sdf_steps, triangle_hits;
  uint64_t sdf_hits, sdf_misses, bvh_node_visits, bvh_leaf_visits, shadow_tests;
} PtSampleCounters;

/* RayQueryHit, PathTracer.h:976-983, as CpuBvhAccelerator::intersect fills it
 * (CpuBvhAccelerator.cpp:446-452): on a hit, position = origin + direction * t, normal = the
 * primitive's face normal normalize(e1 x e2), material_index = its instance's material index (not
 * clamped to the material list) and primitive_index = its index in instance-flattening order; on a miss
 * every field is zero (`out = {}`). hit is the bool as a 32-bit word. 40 bytes. */
typedef struct PtRayQueryHit {
  uint32_t hit;
  float t;
  float position[3];
  float normal[3];
  uint32_t material_index;
  uint32_t primitive_index;
} PtRayQueryHit;

/* BVH build parameters. builder 0 = bucketed SAH (BvhBuildConfig,
 * src/gpu/D3D12GpuPathTracer.SceneBvh.h:15-19; defaults leaf 4, 8 buckets),
 * builder 1 = midpoint split (ParallelBvhBuilder, src/cpu/ParallelBvhBuilder.cpp:61-163).
 * The SAH tree is built on the GPU (the host builder's tree, node for node) unless flags has
 * PT_BVH_FLAG_HOST_BUILD or a vertex is not finite; the midpoint tree is built on the host. */
typedef struct PtBvhParams {
  uint32_t builder;
  uint32_t leaf_size;
  uint32_t bucket_count;
  uint32_t flags;
} PtBvhParams;

enum { PT_BVH_SAH = 0, PT_BVH_MIDPOINT = 1 };
enum { PT_BVH_FLAG_HOST_BUILD = 1 };

/* RayAcceleratorBuildInfo, PathTracer.h:992-999, plus depth, device bytes and whether every tree of
 * the build was built on the GPU (device_built; a tree built on the host clears it). */
typedef struct PtBuildInfo {
  double build_ms;
  uint64_t device_bytes;
  uint32_t built, primitive_count, node_count, leaf_count, max_depth, deterministic;
  uint32_t device_built, pad0;
} PtBuildInfo;

/* PathTracerStatus subset (PathTracer.h:854-865). */
typedef struct PtTracerStatus {
  uint64_t total_samples;
  uint64_t accumulation_gen;
  uint32_t lifecycle, scene_loaded, accel_valid, ready_to_render, current_sample, pad0;
} PtTracerStatus;

/* Per-kernel HIP-event timings collected while profiling is enabled. */
enum {
  PT_KERNEL_RAYGEN = 0,
  PT_KERNEL_EXTEND = 1,
  PT_KERNEL_SHADE = 2,
  PT_KERNEL_SHADOW = 3,
  PT_KERNEL_FILM = 4,
  PT_KERNEL_COUNT = 5
};
typedef struct PtKernelStats {
  double total_ms[PT_KERNEL_COUNT];
  uint64_t launches[PT_KERNEL_COUNT];
  uint64_t items[PT_KERNEL_COUNT]; /* paths / rays handed to each kernel */
  uint64_t inner_node_fetches;    /* 64-byte BVH node loads by extend + shadow */
  uint64_t shadow_node_visits, shadow_triangle_tests;
  uint64_t shadow_inner_node_fetches, shadow_leaf_visits, shadow_rays;
} PtKernelStats;

typedef struct pt_tracer pt_tracer;

/* Handle lifetime: the reference constructs ScalarCpuPathTracer directly
 * (BenchmarkRuntime.cpp:656-664). device is the HIP device ordinal. */
int pt_hip_create(int device, pt_tracer** out);
void pt_hip_destroy(pt_tracer* tracer);

/* IPathTracer::configure, PathTracer.h:1078 (ScalarCpuPathTracer.cpp:415-456). */
int pt_hip_configure(pt_tracer* tracer, const PtRenderSettings* settings);
/* IPathTracer::load_scene_snapshot, PathTracer.h:1091 (ScalarCpuPathTracer.cpp:458-477). */
int pt_hip_load_scene(pt_tracer* tracer, const PtSceneView* scene);
/* IPathTracer::build_or_update_acceleration, PathTracer.h:1095 (ScalarCpuPathTracer.cpp:479-520)
 * and IRayAccelerator::build / build_info, PathTracer.h:1023, :1034. params may be NULL. */
int pt_hip_build_accel(pt_tracer* tracer, const PtBvhParams* params, PtBuildInfo* info);
/* IPathTracer::update_render_settings, PathTracer.h:1083-1087: new integrator settings (seed, spp,
 * depth, NEE / MIS, roulette, lens defaults) without touching the scene or BVH. A different resolution
 * returns PT_UNSUPPORTED: the caller falls back to configure / load / build, as the contract says. */
int pt_hip_update_render_settings(pt_tracer* tracer, const PtRenderSettings* settings);
/* IPathTracer::update_camera / update_camera_state, PathTracer.h:1100-1106 (ScalarCpuPathTracer.cpp:522-546):
 * the scene's camera changes without re-uploading geometry or rebuilding the BVH. The accumulated film
 * is kept (the caller resets it, as with the reference). */
int pt_hip_update_camera(pt_tracer* tracer, const PtCamera* camera);
/* IPathTracer / IRayAccelerator plan_instance_transform_update and apply_instance_transform_update,
 * PathTracer.h:1009-1032, :1115-1137 (CpuBvhAccelerator.cpp:265-358): the plan is a pure query; apply
 * recomputes the updated dynamic instances' triangles from their local geometry on the GPU and refits
 * every node box bottom-up (the topology stays), then renders see the moved geometry. Instances
 * without dynamic local geometry block with PT_XFORM_BLOCKED_NEEDS_FULL_STATIC_ACCEL_REBUILD (the
 * caller rebuilds). The call returns PT_OK with the outcome in the result's status. */
int pt_hip_plan_instance_transform_update(pt_tracer* tracer, const PtInstanceTransformUpdate* updates,
                                          uint32_t count, PtInstanceTransformResult* plan);
int pt_hip_apply_instance_transform_update(pt_tracer* tracer, const PtInstanceTransformUpdate* updates,
                                           uint32_t count, PtInstanceTransformResult* result);
/* IPathTracer::reset_accumulation, PathTracer.h:1099 (ScalarCpuPathTracer.cpp:669-688). */
int pt_hip_reset_accumulation(pt_tracer* tracer);
/* IPathTracer::render_tile, PathTracer.h:1141 (ScalarCpuPathTracer.Render.cpp:140-226). */
int pt_hip_render_tile(pt_tracer* tracer, const PtRenderTile* tile, uint64_t frame_index);
/* IPathTracer::render_tiles, PathTracer.h:1157-1168: all tiles in one wavefront launch sequence.
 * Tiles are accumulated into the film in array order. */
int pt_hip_render_tiles(pt_tracer* tracer, const PtRenderTile* tiles, uint32_t count, uint64_t frame_index);
/* Film readback: FilmBuffer::raw / sample_counts (PathTracer.h:955-958). Host pointers:
 * rgb_sum holds 3 floats per pixel, sample_count one u32 per pixel (either may be NULL). */
int pt_hip_read_film(pt_tracer* tracer, float* rgb_sum, uint32_t* sample_count);
/* PathTracerSceneSnapshot::sdf_primitives (PathTracer.h:358-409), after pt_hip_load_scene and before
 * pt_hip_build_accel: primitives intersect_scene tests after the triangles
 * (ScalarCpuPathTracer.Intersect.cpp:91-129; intersectors ScalarCpuPathTracer.cpp:809-962), on
 * extension and shadow rays, counted in the sdf_* counters. Unknown shapes are dropped (never hit). */
int pt_hip_load_sdf_primitives(pt_tracer* tracer, const PtSdfPrimitive* prims, uint32_t count);
/* IPathTracer::request_film_readback / poll_film, PathTracer.h:1172-1179: the request enqueues the GPU
 * resolve of the film as it stands after all work queued so far (settings as pt_hip_resolve_ldr) and its
 * copy into pinned host memory on a second stream, and returns at once; rendering queued afterwards
 * overlaps the copy. poll_film sets *state to PT_READBACK_PENDING until the copy has landed, then to
 * PT_READBACK_READY and copies the image (4 bytes per pixel) into rgba8 (may be NULL); a token other
 * than the latest is PT_READBACK_INVALID. */
int pt_hip_request_film_readback(pt_tracer* tracer, const PtFilmResolve* settings, PtFilmReadbackToken* token);
int pt_hip_poll_film(pt_tracer* tracer, const PtFilmReadbackToken* token, uint8_t* rgba8, uint32_t* state);
/* IPathTracer::replace_film_history, PathTracer.h:1096-1098: replaces the accumulated film (host
 * pointers, 3 floats and one count per pixel) without touching scene or acceleration data. */
int pt_hip_replace_film(pt_tracer* tracer, const float* rgb_sum, const uint32_t* sample_count);
/* IPathTracer::resolve_hdr, PathTracer.h:1171 (FilmBuffer::resolve_hdr, FilmBuffer.cpp:160-214), evaluated
 * on the GPU: 3 floats per pixel (mean of the pixel's samples). rgb may be host or device memory. */
int pt_hip_resolve_hdr(pt_tracer* tracer, float* rgb);
/* IPathTracer::resolve_ldr, PathTracer.h:1170 (FilmBuffer::resolve_ldr -> ApplyFilmResolve,
 * FilmBuffer.cpp:156-158, :255-332), evaluated on the GPU: exposure, white balance, tone map, gamma,
 * clamp, 4 bytes (r, g, b, 255) per pixel. settings NULL = FilmResolveSettings{}; a caller applying the
 * camera's exposure controls passes CameraAdjustedFilmResolveSettings' result (FilmBuffer.cpp:226-253).
 * rgba8 may be host or device memory. */
int pt_hip_resolve_ldr(pt_tracer* tracer, const PtFilmResolve* settings, uint8_t* rgba8);
/* IPathTracer::read_counters, PathTracer.h:1180. */
int pt_hip_read_counters(pt_tracer* tracer, PtSampleCounters* counters);
/* IPathTracer::status, PathTracer.h:1181. */
int pt_hip_status(pt_tracer* tracer, PtTracerStatus* status);
/* IPathTracer::shutdown, PathTracer.h:1182 (ScalarCpuPathTracer.Render.cpp:463-481). */
int pt_hip_shutdown(pt_tracer* tracer);
/* Message of the last failing call on this handle ("" if none). */
const char* pt_hip_last_error(const pt_tracer* tracer);

/* IRayAccelerator::intersect, PathTracer.h:1032 (CpuBvhAccelerator.cpp:360-458), batched.
 * Device pointers: rays = 6 floats per ray (origin, direction); tmax = per-ray t limit or NULL
 * (infinity); out_t / out_prim = closest t and RayQueryHit::primitive_index (-1 on miss).
 * mode 0 = closest hit, mode 1 = any hit with t < tmax (out_t is the first accepted t).
 * counters (host, may be NULL) receive bvh_node_visits / bvh_leaf_visits / triangle_tests /
 * triangle_hits for this call. */
int pt_hip_trace_rays(pt_tracer* tracer, int mode, const float* rays, const float* tmax, uint32_t count,
                      float* out_t, int32_t* out_prim, PtSampleCounters* counters);

/* pt_hip_trace_rays without any allocation (no counterpart in the reference: for callers that cannot
 * allocate device memory — code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). Same arguments, results and counters as pt_hip_trace_rays; rays, tmax, out_t and out_prim
 * are device pointers. The traversal's scratch (counter stripes, queue cursor, error flag and the stack
 * spill area of a 512-workgroup grid) lives in one static 26 MiB device array per GPU, set up when
 * the first tracer on that GPU is created and held until the process ends. Trees deeper than
 * PT_NOALLOC_MAX_TREE_DEPTH levels are rejected with PT_UNSUPPORTED. Every call on one GPU shares that
 * GPU's array: calls that overlap on one GPU (two streams, two threads) race and must be serialised by
 * the caller. */
#define PT_NOALLOC_MAX_TREE_DEPTH 32
int pt_hip_trace_rays_noalloc(pt_tracer* tracer, int mode, const float* rays, const float* tmax, uint32_t count,
                              float* out_t, int32_t* out_prim, PtSampleCounters* counters);

/* IRayAccelerator::intersect, PathTracer.h:1032 (CpuBvhAccelerator.cpp:360-458), with the whole
 * RayQueryHit per ray (PtRayQueryHit). rays (6 floats each), tmax (per-ray limit or NULL = infinity)
 * and hits may be host or device memory; they are staged through device buffers the tracer owns.
 * mode 0 = closest hit (the reference's query), mode 1 = the first leaf batch with an accepted hit
 * below tmax. stats (host, may be NULL) receive RayQueryStats' four counters (PathTracer.h:985-990)
 * summed over the call, in the triangle_tests / triangle_hits / bvh_node_visits / bvh_leaf_visits
 * fields. */
int pt_hip_intersect(pt_tracer* tracer, int mode, const float* rays, const float* tmax, uint32_t count,
                     PtRayQueryHit* hits, PtSampleCounters* stats);

/* IPathTracer::update_scene_delta (PathTracer.h:1139; ApplySceneDeltaUpdate): replace the material
 * and / or light records without rebuilding geometry. Counts must equal the loaded scene's. Either
 * pointer may be NULL. Host pointers. */
int pt_hip_update_scene_delta(pt_tracer* tracer, const PtMaterial* materials, uint32_t material_count,
                              const PtLight* lights, uint32_t light_count);

/* RTSceneDeltaUpdate, PathTracer.h:433-461: per-index material and light records, per-instance material
 * indices (an instance named by index, or by entity_id when the index is out of range) and the
 * environment colour. Host pointers; a count of 0 with a NULL pointer is an empty list. */
typedef struct PtMaterialUpdate {
  uint32_t material_index, pad;
  PtMaterial material;
} PtMaterialUpdate;
typedef struct PtLightUpdate {
  uint32_t light_index;
  PtLight light;
} PtLightUpdate;
typedef struct PtInstanceMaterialUpdate {
  uint64_t entity_id;
  uint32_t instance_index; /* 0xFFFFFFFF (kInvalidRTInstanceIndex): look the instance up by entity_id */
  uint32_t material_index;
} PtInstanceMaterialUpdate;
typedef struct PtSceneDeltaUpdate {
  const PtMaterialUpdate* materials;
  const PtLightUpdate* lights;
  const PtInstanceMaterialUpdate* instance_materials;
  uint32_t material_count, light_count, instance_material_count;
  uint32_t environment_color_changed;
  float environment_color[3];
  uint32_t pad;
} PtSceneDeltaUpdate;

/* IPathTracer::update_scene_delta, PathTracer.h:1139, as ScalarCpuPathTracer::update_scene_delta
 * (ScalarCpuPathTracer.cpp:701-720) applies ApplySceneDeltaUpdate (PathTracer.cpp:427-482): every index
 * is checked first (PT_INVALID_ARGUMENT, nothing changed), then the records are replaced. Materials,
 * lights and the environment colour reach the next render; instance materials are baked into the
 * acceleration structure's primitives, as the reference's accelerator bakes them at build
 * (CpuBvhAccelerator.cpp:232-238), so they reach renders after the next pt_hip_build_accel. An update
 * that changes nothing (no records, and instance materials equal to the current ones) returns
 * PT_INVALID_ARGUMENT with "scene delta did not apply", as the reference returns false. */
int pt_hip_apply_scene_delta(pt_tracer* tracer, const PtSceneDeltaUpdate* update);

/* Path-replay adjoint of render_tiles (no counterpart in the reference, which has no
 * differentiation). For the same tiles and frame, replays every path with the same random
 * numbers and accumulates, with fp32 atomics, the gradient of
 *   loss = sum over pixels p, channels c of film_grad[p][c] * (film sum contributed by these tiles)[p][c]
 * into grad_albedo and grad_emission (3 floats per material), grad_light (3 floats per light
 * colour), grad_roughness (1 float per material: the Blinn-Phong NEE term plus the score of the
 * glossy lobe's sampling density; may be NULL) and grad_texels (3 floats per texel of the loaded
 * textures, in pt_hip_load_textures order; may be NULL). A textured material's albedo gradient is
 * taken with the texture held fixed and each texel's with the albedo held fixed. All are device
 * pointers; gradients are added to the buffers' contents. The film and the render counters are not
 * modified. Requires rr_start_depth <= 4. Right after render_tiles of the same tiles and frame (no
 * scene change between; other adjoints of that render may come between), the forward pass is not
 * replayed: each path's radiance from that render is reused. */
int pt_hip_render_adjoint(pt_tracer* tracer, const PtRenderTile* tiles, uint32_t count, uint64_t frame_index,
                          const float* film_grad, float* grad_albedo, float* grad_emission, float* grad_light,
                          float* grad_roughness, float* grad_texels);

/* Albedo textures (no counterpart in the reference's CPU path, whose hits carry no texture
 * coordinates; SceneSnapshot.h:118 holds per-vertex texcoords that it never reads). A texture is
 * width x height rgb fp32 texels, row-major with rows in increasing v, starting at texel `offset`
 * of the texel array (3 floats per texel). A material whose base_color_texture_index names one of
 * the scene's textures has base colour albedo * texture(u, v): bilinear, repeat addressing, texel
 * centres at ((i + 0.5) / width, (j + 0.5) / height), uv interpolated from the hit triangle's
 * vertex texcoords with the hit barycentrics. 16 bytes. */
typedef struct PtTextureDesc {
  uint32_t offset, width, height, pad;
} PtTextureDesc;
/* Texture coordinates (2 floats per vertex of PtSceneView.vertices; instanced geometry has none)
 * and textures of the loaded scene. Host pointers, copied. Call after pt_hip_load_scene and before
 * pt_hip_build_accel (a later call needs another build). texcoord_count must equal the scene's
 * vertex count. */
int pt_hip_load_textures(pt_tracer* tracer, const float* texcoords, uint32_t texcoord_count, const float* texels,
                         uint64_t texel_count, const PtTextureDesc* textures, uint32_t texture_count);
/* Replace every texel of the built scene (same count) without a rebuild. texels is a host or
 * device pointer. */
int pt_hip_update_texels(pt_tracer* tracer, const float* texels, uint64_t texel_count);

/* Block until all work on the handle's stream is done. */
int pt_hip_synchronize(pt_tracer* tracer);
/* Maximum paths in flight per wavefront pass (default 8M). */
int pt_hip_set_wavefront_capacity(pt_tracer* tracer, uint32_t paths);
/* Launch-shape tuning (no counterpart in the reference; chosen by vkpathtracer_amd.autotune).
 * In pt_hip_set_tuning a zero wavefront_capacity or refill_min keeps the current value; a block
 * count caps each kernel's occupancy-derived persistent grid, and zero restores that grid.
 * pt_hip_get_tuning reports the extend and shade grids. */
typedef struct PtTuning {
  uint32_t wavefront_capacity; /* paths in flight per wavefront pass */
  uint32_t refill_min;         /* 1..64: idle lanes a traversal wave gathers before it takes new rays */
  uint32_t traversal_blocks;   /* persistent workgroups of the extend / shadow / ray-query kernels */
  uint32_t shade_blocks;       /* workgroups of the grid-stride shade kernel */
} PtTuning;
int pt_hip_set_tuning(pt_tracer* tracer, const PtTuning* tuning);
int pt_hip_get_tuning(pt_tracer* tracer, PtTuning* tuning);
/* HIP-event timing per kernel (off by default). */
int pt_hip_set_profiling(pt_tracer* tracer, int enabled);
int pt_hip_read_kernel_stats(pt_tracer* tracer, PtKernelStats* stats);
int pt_hip_reset_kernel_stats(pt_tracer* tracer);

/* Batched rendering (no counterpart in the reference, whose tracers hold one scene): many scenes,
 * each with its own image size, spp and BVH parameters, packed into one device scene pool and
 * rendered by one wavefront launch sequence per frame (paths of every scene share the ray queues,
 * so small scenes fill the GPU together). Integrator settings (seed, depth, NEE, MIS, roulette,
 * lens defaults) are shared by the batch; each scene's film is its own. */
typedef struct pt_batch pt_batch;
int pt_batch_create(int device, const PtRenderSettings* settings, pt_batch** out);
void pt_batch_destroy(pt_batch* batch);
const char* pt_batch_last_error(const pt_batch* batch);
/* Copies a scene (as pt_hip_load_scene does); index receives its position in the batch. params may be NULL. */
int pt_batch_add_scene(pt_batch* batch, const PtSceneView* scene, uint32_t width, uint32_t height, uint32_t spp,
                       const PtBvhParams* params, uint32_t* index);
/* Another camera on scene `scene` (an index from pt_batch_add_scene): shares that scene's geometry,
 * materials, lights and textures in the pool, renders its own image and film. index receives its
 * position in the batch. */
int pt_batch_add_view(pt_batch* batch, uint32_t scene, const PtCamera* camera, uint32_t width, uint32_t height,
                      uint32_t spp, uint32_t* index);
/* Textures of scene `scene`, as pt_hip_load_textures (before pt_batch_build). */
int pt_batch_load_textures(pt_batch* batch, uint32_t scene, const float* texcoords, uint32_t texcoord_count,
                           const float* texels, uint64_t texel_count, const PtTextureDesc* textures,
                           uint32_t texture_count);
/* Builds every scene's tree on host threads, uploads the pool, allocates and clears the films.
 * info (may be NULL) receives pool totals. */
int pt_batch_build(pt_batch* batch, PtBuildInfo* info);
int pt_batch_reset_accumulation(pt_batch* batch);
/* Adds samples 0..spp-1 of every scene's full frame to its film. use_graph != 0 replays the frame's
 * launch sequence from a HIP graph captured on first use (re-captured after a tuning change). */
int pt_batch_render(pt_batch* batch, uint64_t frame_index, int use_graph);
/* update_scene_delta over the whole pool: materials (lights) of every scene in the order the scenes
 * were added, counts equal to the pool's. Host pointers. */
int pt_batch_update_scene_delta(pt_batch* batch, const PtMaterial* materials, uint32_t material_count,
                                const PtLight* lights, uint32_t light_count);
/* Replace every texel of the pool (every scene's textures in order; same count). Host or device pointer. */
int pt_batch_update_texels(pt_batch* batch, const float* texels, uint64_t texel_count);
/* Path-replay adjoint of pt_batch_render (every scene's and view's full frame at its spp), as
 * pt_hip_render_adjoint. film_grad holds 3 floats per pixel of every film, films in batch order;
 * gradient buffers hold every scene's materials (lights, texels) in the order the scenes were added
 * (views add none). Device pointers. */
int pt_batch_render_adjoint(pt_batch* batch, uint64_t frame_index, const float* film_grad, float* grad_albedo,
                            float* grad_emission, float* grad_light, float* grad_roughness, float* grad_texels);
/* Film of one scene, as pt_hip_read_film. */
int pt_batch_read_film(pt_batch* batch, uint32_t scene, float* rgb_sum, uint32_t* sample_count);
/* SDF primitives of one scene (not a view), as pt_hip_load_sdf_primitives; before pt_batch_build. */
int pt_batch_load_sdf_primitives(pt_batch* batch, uint32_t scene, const PtSdfPrimitive* prims, uint32_t count);
/* Camera of one scene or view, as pt_hip_update_camera (graph replays see the new camera). */
int pt_batch_update_camera(pt_batch* batch, uint32_t scene, const PtCamera* camera);
/* Resolved image of one scene, as pt_hip_resolve_hdr / pt_hip_resolve_ldr. */
int pt_batch_resolve_hdr(pt_batch* batch, uint32_t scene, float* rgb);
int pt_batch_resolve_ldr(pt_batch* batch, uint32_t scene, const PtFilmResolve* settings, uint8_t* rgba8);
/* Counters summed over the batch's scenes. */
int pt_batch_read_counters(pt_batch* batch, PtSampleCounters* counters);
int pt_batch_set_tuning(pt_batch* batch, const PtTuning* tuning);
int pt_batch_get_tuning(pt_batch* batch, PtTuning* tuning);
int pt_batch_set_profiling(pt_batch* batch, int enabled);
int pt_batch_read_kernel_stats(pt_batch* batch, PtKernelStats* stats);
int pt_batch_reset_kernel_stats(pt_batch* batch);
int pt_batch_synchronize(pt_batch* batch);

/* Host-only BVH construction (no device needed): the flattening and split decisions of
 * CpuBvhAccelerator::build (CpuBvhAccelerator.cpp:161-263) with the chosen builder. Call once with
 * NULL outputs to get the counts, then again with buffers of that size:
 *   bounds     6 floats per node (padded min xyz, max xyz), node 0 = root, pre-order numbering
 *   meta       4 int32 per node: left, right, first, count (leaves: left = right = -1)
 *   prim_order 1 u32 per primitive: leaf order -> flattened primitive index
 *   prim_geo   12 floats per primitive: v0, e1, e2, face normal */
int pt_bvh_build_host(const PtSceneView* scene, const PtBvhParams* params, uint32_t* node_count,
                      uint32_t* prim_count, uint32_t* max_depth, float* bounds, int32_t* meta,
                      uint32_t* prim_order, float* prim_geo);

/* The same export for the tree pt_hip_build_accel builds: the SAH tree built on the current HIP device
 * (built_on_device, may be NULL, is 1; 0 when a non-finite vertex sent the build to the host, or for
 * the midpoint builder). The outputs equal pt_bvh_build_host's for the same scene and parameters.
 * No counterpart in the reference, whose builders run on the CPU (D3D12GpuPathTracer.SceneBvh.cpp:73-238
 * for the bucketed SAH). */
int pt_bvh_build_device(const PtSceneView* scene, const PtBvhParams* params, uint32_t* node_count,
                        uint32_t* prim_count, uint32_t* max_depth, float* bounds, int32_t* meta,
                        uint32_t* prim_order, float* prim_geo, uint32_t* built_on_device);

/* ABI sizes, for bindings to check their mirrors of the records above. */
uint32_t pt_hip_sizeof(int which); /* 0 material, 1 light, 2 instance, 3 camera, 4 scene, 5 settings, 6 counters,
                                      7 tile, 8 tuning, 9 texture, 10 film resolve,
                                      11 instance transform update, 12 SDF primitive, 13 ray query hit,
                                      14 scene delta update, 15 material update, 16 light update,
                                      17 instance material update, 18 build info */
const char* pt_hip_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* VKPT_HIP_H */
