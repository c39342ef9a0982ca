// pt_device.h — device records shared by the wavefront kernels and the C ABI layer.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "vkpt_hip.h"

namespace ptd {

constexpr int kBlock = 256;        // threads per workgroup (4 waves of 64)
constexpr int kWavesPerBlock = kBlock / 64;
#ifndef PT_STACK_LDS
#define PT_STACK_LDS 8
#endif
constexpr int kStackLds = PT_STACK_LDS;  // newest traversal stack entries per lane held in LDS (a ring)
#ifndef PT_HOT_NODES
#define PT_HOT_NODES 64
#endif
// Inner nodes a traversal workgroup keeps in LDS: the pool's hot copies (DevScene::hot_base). With
// the stack ring this is 20 KiB per workgroup, seven workgroups per CU.
constexpr int kHotNodes = PT_HOT_NODES;
constexpr int kCounterSlots = 16;  // device u64 counters (see CounterSlot)

// Every same-address device atomic is serialized (about 11 ns each on MI355X), so counters and
// queue cursors are striped: counter slot k of stripe s lives at counters[s * kCounterSlots + k]
// (one 128-byte line per stripe) and a wave adds into stripe (global wave id % kStripes). Readers
// sum the stripes.
constexpr int kStripes = 64;
constexpr int kCounterWords = kStripes * kCounterSlots;

enum CounterSlot : int {
  kSamples = 0,
  kRays = 1,
  kTriTests = 2,
  kTriHits = 3,
  kNodeVisits = 4,
  kLeafVisits = 5,
  kShadowTests = 6,
  kInnerFetches = 7,
  kShadowNodeVisits = 8,
  kShadowTriTests = 9,
  kShadowRays = 10,
  kShadowInnerFetches = 11,
  kShadowLeafVisits = 12,
  kSdfTests = 13,  // SampleCounters::sdf_tests / sdf_steps / sdf_hits (sdf_misses = tests - hits)
  kSdfSteps = 14,
  kSdfHits = 15,
};

// Ray queues are split into kSegments segments of seg_cap entries. Queue entry j of a bounce
// (0 <= j < total) lives in the segment s with prefix[s] <= j < prefix[s + 1], at position
// s * seg_cap + (j - prefix[s]). Producers append a 64-entry input chunk c's outputs to segment
// c % kSegments, so the append cursors are spread over kSegments cache lines; a segment never
// receives more than seg_cap = 64 * ceil(ceil(capacity / 64) / kSegments) entries.
constexpr int kSegments = 64;
constexpr int kSegLine = 32;  // u32 per cursor line (128 bytes)
// u32 layout of the queue block
constexpr int kQueueInCount = 0;                                  // kSegments counts of this bounce's input
constexpr int kQueueOutCursor = kSegments;                        // kSegments cursors, kSegLine apart
constexpr int kQueueShadowCursor = kQueueOutCursor + kSegments * kSegLine;
// Traversal scheduling: the next chunk of queries to hand out, and the waves of the running launch
// that have finished; the launch's last wave sets both back to zero.
// Traversal chunks are handed out from one cursor per region of the queue (kRegions regions, one per
// XCD, each cursor on its own line).
constexpr int kRegions = 8;
constexpr int kQueueChunkCursor = kQueueShadowCursor + kSegments * kSegLine;
constexpr int kQueueWavesDone = kQueueChunkCursor + kRegions * kSegLine;
// The finished waves are counted in kDoneStripes stripes (wave w in stripe w % kDoneStripes, one line
// each); kQueueWavesDone counts the completed stripes. One counter for all waves would serialise every
// wave of a launch on one address (tens of microseconds for a resident grid).
constexpr int kDoneStripes = 32;
constexpr int kQueueDoneStripe = kQueueWavesDone + kSegLine;
constexpr int kQueueWords = kQueueDoneStripe + kDoneStripes * kSegLine;

__host__ __device__ constexpr uint32_t segment_capacity(uint32_t capacity) {
  return 64u * ((((capacity + 63u) / 64u) + kSegments - 1u) / kSegments);
}

struct DevCamera {
  float pos[3], fwd[3], right[3], up[3];
  float aspect, tan_half_fov, aperture, focus;
  float iris_roundness, iris_rotation, anamorphic;
  uint32_t iris_blades;
};

// One scene of the pool: where its tree, lights, environment and film live. A single-scene
// tracer has one record; a batch has one per scene, and every path carries its scene's index.
// RTSdfPrimitive (PathTracer.h:330-339) as the kernels use it: the pool material index, and the sines
// and cosines of -rotation (rotate_euler_inv) and +rotation (rotate_euler) per axis, formed once on the
// host with the same sin / cos the kernels use (pure functions, so the values are those a per-step
// evaluation would give).
struct DevSdf {
  uint32_t shape, material;
  float pos[3], scale[3];
  float radius, pa, pb;
  float inv_s[3], inv_c[3];
  float fwd_s[3], fwd_c[3];
  float pad;
};

struct SceneRec {
  int32_t root_ref;       // pool-rebased root reference
  uint32_t prim_count;
  float root_box[6];
  uint32_t light_base, light_count, has_transmissive, env_base;  // env_base: float offset in the env pool
  uint32_t env_w, env_h;  // 0 x 0: constant environment colour
  uint32_t width, height, film_base;  // film pixels [film_base, film_base + width * height)
  uint32_t sdf_base, sdf_count;       // the scene's SDF primitives in DevScene::sdfs
  float env[3], env_scale[3];
  DevCamera cam;
};

// One texture of the pool: width x height rgb fp32 texels starting at texel `offset` of
// DevScene::texels (3 floats per texel), rows in increasing v.
struct TexRec {
  uint32_t offset, width, height, pad;
};

// The scene pool as kernels see it: all scenes' nodes, triangles and primitives (node and
// primitive references rebased into the pool, material indices clamped per scene and rebased),
// materials and lights. Inside shading a lane works on a copy whose lights, light_count,
// has_transmissive and env fields are its own scene's (scene_view).
struct DevScene {
  const float4* nodes;   // 4 x float4 per inner node
  const float4* tris;    // 3 x float4 per leaf-ordered triangle
  const float4* prims;   // normal.xyz + pool material index, by pool primitive index
  const PtMaterial* materials;
  const PtLight* lights;
  uint32_t material_count, light_count, has_transmissive;
  uint32_t general_boxes;  // some node box coordinate is not finite: every ray takes the select form of slab
  float env[3];
  const float* env_map;  // rgb, env_w * env_h (nullptr: constant colour)
  uint32_t env_w, env_h;
  float env_scale[3];
  const SceneRec* recs;
  uint32_t scene_count;
  const float* texels;     // texture pool (rgb)
  const TexRec* textures;  // material base_color_texture_index values are pool texture indices
  uint32_t texture_count;
  uint32_t textured;       // some material has a texture
  const float* prim_uv;    // 6 floats per pool primitive: texture coordinates of v0, v1, v2
  const uint32_t* prim_tri;  // per pool primitive: its triangle record in `tris`
  const DevSdf* sdfs;      // SDF primitives of all scenes (SceneRec::sdf_base / sdf_count)
  uint32_t has_sdf;        // some scene has SDF primitives: extension rays take k_sdf_extend, shadow
                           // rays the closest-hit loop with SDF tests
  // Copies of the most visited inner nodes at node indices [hot_base, hot_base + hot_count) (after
  // the pool's own nodes; hot_count <= kHotNodes): a hot copy's children that are hot are hot
  // references, and a scene whose root is hot starts there. Traversal kernels read them from LDS.
  uint32_t hot_base, hot_count;
};

// Integrator settings shared by every scene of a launch (image sizes are per scene: SceneRec).
struct DevSettings {
  uint64_t seed;
  uint64_t frame;
  const uint64_t* frame_dev;  // when set, raygen reads the frame index here (graph replays)
  uint32_t max_depth, enable_nee, enable_mis, rr_start;
  float rr_min, rr_max;
};

// Tile of one wavefront pass: w*h paths render sample `sample` of this rectangle of scene `scene`'s
// image. Consecutive tiles of a pass with the same scene and rectangle form a group of G tiles
// (G = group >> 16; 1 for a lone tile) whose paths interleave: pixel i of the group's k-th tile
// (k = group & 0xFFFF) is path path_begin + i * G + k. A pixel's samples are neighbours in the ray
// queues, so a wave traces up to 64 samples of one pixel: nearly identical camera rays.
struct DevTile {
  uint32_t x, y, w, h, sample, path_begin, scene, group;
};

__host__ __device__ inline uint32_t tile_slot(const DevTile& t, uint32_t local) {
  return t.path_begin + local * (t.group >> 16) + (t.group & 0xFFFFu);
}

// Film accumulation of one scene in one pass: the bounding rectangle of the scene's tiles in the
// pass and their range in the pass's tile table (in render order). The film kernel gives each job
// ceil(w*h / 64) waves, numbered from wave_begin across the pass's jobs. group = G when the job's
// tiles are exactly one interleaved group of G tiles (DevTile::group), else 1.
struct FilmJob {
  uint32_t scene, x0, y0, w, h, first_tile, tile_count, wave_begin, group, pad[3];
};

// One ray queue: origin xyz, direction xyz (SoA), the path slot each entry serves and its scene.
// Forward passes keep a live path's throughput, radiance and RNG state with its queue entry (tp, rad,
// rng in queue order), so shading reads and writes them with the entry's coalesced loads and stores;
// a path's radiance reaches DevPaths::rad (indexed by path slot) once, when the path ends. Adjoint
// replays keep all three in DevPaths, by path slot, beside the rest of their per-path state.
struct DevRays {
  float* c[6];
  uint32_t* qpath;
  uint32_t* qscene;
  float* tp[3];
  float* rad[3];
  uint64_t* rng;
};

// Where a forward shadow record's light contribution goes (DevPaths::spath): the radiance of the
// path's queued continuation ray (kRadQueued | its entry in the output queue of the bounce), or, when
// the path ends at this bounce, DevPaths::rad at its path slot.
constexpr uint32_t kRadQueued = 0x80000000u;
// Adjoint replays: the path ended at this bounce and its replayed radiance is no longer needed.
constexpr uint32_t kRadNone = 0xFFFFFFFFu;
struct RadTarget {
  float* rad[3];  // the output queue's radiance arrays
};


// Per-path wavefront state. Ray and hit buffers, and a live path's state, are in queue order
// (double-buffered rays); indexed by the path's slot in the pass is only what a forward path leaves
// when it ends: its radiance and the number of extension rays it traced (for k_film's ray count of
// dropped samples).
struct DevPaths {
  DevRays rays[2];
  float* hit_t;
  int32_t* hit_prim;
  float* rad[3];          // radiance of ended paths
  uint32_t* ended_rays;   // extension rays of ended paths
  float* sray[6];        // shadow queue: origin xyz, direction xyz
  float* stmax;
  float* scontrib[3];
  uint32_t* spath;
  uint32_t* sscene;
  uint32_t* queue;               // kQueueWords: segment counts and append cursors
  unsigned long long* counters;  // kCounterWords: striped CounterSlot counters
  uint32_t* error;               // nonzero: a kernel refused work (stack overflow area too small)
  int32_t* stack_ref;            // traversal stack overflow beyond kStackLds
  float* stack_t;
  uint32_t stack_overflow_depth; // entries per lane in the overflow area
  uint32_t stack_threads;        // lanes the overflow area is sized for (persistent grid * kBlock)
  uint32_t capacity;             // paths per pass (path-state arrays)
  uint32_t seg_cap;              // entries per queue segment (queue arrays hold kSegments * seg_cap)
  uint32_t refill_min;           // idle lanes a traversal wave gathers before it takes new queries
};

struct DevFilm {
  float4* sum;       // rgb sum (w unused)
  uint32_t* count;
};

// FilmResolveSettings (PathTracer.h:67-74) after the host-side conversions: the white-balance scale
// (WhiteBalanceScale, FilmBuffer.cpp:216-224) and 1 / max(0.01, gamma).
struct DevResolve {
  float exposure;
  float wb[3];
  float inv_gamma;
  uint32_t tone_map;   // ToneMapMode: 0 linear, 1 reinhard, 2 filmic_approx, 3 aces_approx
  uint32_t gamma_out;  // OutputTransformMode::Gamma
  uint32_t clamp_out;
};

// Path-replay adjoint state (no counterpart in the reference). Per queue entry (items): forward
// radiance, the loss adjoint of that radiance and the bounce awaiting its remainder radiance; per path
// slot: the history of vertices before the roulette start depth (see k_shade). Per shadow record: the
// gradient pieces that apply if the light turns out visible.
constexpr int kAdjHist = 4;          // max roulette start depth supported by the adjoint
constexpr int kAdjLdsFloats = 6144;  // block-staged gradient slots, all copies (24 KB of LDS)
// A path's per-bounce adjoint state travels with its queue entry, double-buffered like the ray queues
// (items[b] belongs to DevPaths::rays[b]): `stride` entries apart, the float arrays forward radiance
// ltot xyz (kItemLtot), loss adjoint delta xyz (kItemDelta), the pending bounce's d log(bounce weight)
// / d albedo xyz (kItemPdx), its roughness score (kItemScore) and texture coordinate (kItemUv), and the
// u32 arrays pending material (kItemMat) and flags (kItemFlags: bit0 pending, bit1 active, bits2-3 c*,
// bits4-6 nhist, bit7 prev active, bits8-9 prev c*, bits10-12 prev nhist).
enum : uint32_t { kItemLtot = 0, kItemDelta = 3, kItemPdx = 6, kItemScore = 9, kItemUv = 10, kItemFloats = 12 };
enum : uint32_t { kItemMat = 0, kItemFlags = 1, kItemWords = 2 };
struct AdjItems {
  float* f;
  uint32_t* u;
};
struct DevAdjoint {
  AdjItems items[2];
  float* h_dx[kAdjHist][3];   // h_dx[h][c] = h_dx[0][0] + (3h + c) * stride
  uint32_t* h_mat[kAdjHist];  // h_mat[h] = h_mat[0] + h * stride
  float* s_galb[3];
  uint32_t* s_mat;
  float* s_glight[3];
  uint32_t* s_light;
  float* s_grough;  // per shadow record: d contribution / d roughness, weighted by the loss adjoint
  // texture coordinates of the history's and the shadow records' vertices (read for textured materials
  // only); h_uv[h][k] = h_uv[0][0] + (2h + k) * stride
  float* h_uv[kAdjHist][2];
  float* s_uv[2];
  float* g_albedo;    // 3 per material
  float* g_emission;  // 3 per material
  float* g_light;     // 3 per light (colour)
  float* g_rough;     // 1 per material (nullptr: not requested)
  float* g_tex;       // 3 per texel of the texture pool (nullptr: not requested)
  const float* film_grad;  // 3 per film pixel (all scenes' films in pool order): dLoss / d(film sum)
  uint32_t nmat, nlight, use_lds;
  uint32_t stride;  // entries per adjoint state array (history arrays by path slot, items by queue entry)
  uint32_t lds_copies, lds_stride;  // LDS staging: copies (a power of two) of the slots, stride apart
  // LDS staging: block b of every adjoint kernel adds its slots into row b of `partial` (plain
  // read-modify-write: kernels of one stream never overlap, and rows are private to block indices);
  // launch_grad_reduce adds the rows into the gradient buffers once at the end.
  float* partial;
  uint32_t partial_rows;
};

// Launchers (pt_kernels.hip). All launch on `stream`; none synchronizes.
void launch_init_queue(hipStream_t stream, const DevPaths& paths, uint32_t live);
void launch_raygen(hipStream_t stream, const DevPaths& paths, const DevScene& scene, const DevSettings& st,
                   const DevTile* tiles, uint32_t tile_count, uint32_t path_count);
void launch_extend(hipStream_t stream, const DevPaths& paths, const DevScene& scene, int in_buf, uint32_t grid);
void launch_shade(hipStream_t stream, const DevPaths& paths, const DevScene& scene, const DevSettings& st,
                  int in_buf, uint32_t depth, uint32_t grid);
// in_buf: the bounce's shade input; forward contributions of paths that continue go to the radiance of
// their entries in the other buffer.
void launch_shadow(hipStream_t stream, const DevPaths& paths, const DevScene& scene, int in_buf, uint32_t grid);
void launch_advance(hipStream_t stream, const DevPaths& paths);
// Adds the radiance of every tile of the pass into the films, pixel by pixel in tile order.
void launch_film(hipStream_t stream, const DevPaths& paths, const DevScene& scene, const DevFilm& film,
                 const DevTile* tiles, const FilmJob* jobs, uint32_t job_count, uint32_t job_waves);
void launch_trace_rays(hipStream_t stream, const DevPaths& paths, const DevScene& scene, int any_hit,
                       const float* rays, const float* tmax, uint32_t count, float* out_t, int32_t* out_prim,
                       uint32_t grid);
// The same queries, each storing a whole RayQueryHit (inst_mat: the instance material of each primitive).
void launch_intersect(hipStream_t stream, const DevPaths& paths, const DevScene& scene, int any_hit, const float* rays,
                      const float* tmax, uint32_t count, PtRayQueryHit* hits, const uint32_t* inst_mat, uint32_t grid);
// Adjoint variants: raygen also gathers each path's loss adjoint and the forward radiance the pass left
// in the path slots; shade and shadow accumulate gradients.
void launch_raygen_adjoint(hipStream_t stream, const DevPaths& paths, const DevAdjoint& adj, const DevScene& scene,
                           const DevSettings& st, const DevTile* tiles, uint32_t tile_count, uint32_t path_count);
void launch_shade_adjoint(hipStream_t stream, const DevPaths& paths, const DevAdjoint& adj, const DevScene& scene,
                          const DevSettings& st, int in_buf, uint32_t depth, uint32_t grid);
void launch_shadow_adjoint(hipStream_t stream, const DevPaths& paths, const DevAdjoint& adj, const DevScene& scene,
                           int in_buf, uint32_t grid);
void launch_grad_clear(hipStream_t stream, const DevAdjoint& adj);   // zero the partial rows
void launch_grad_reduce(hipStream_t stream, const DevAdjoint& adj);  // rows -> gradient buffers
constexpr uint32_t kGradRows = 1024;  // partial rows: an upper bound on every adjoint kernel's grid
void launch_zero_counters(hipStream_t stream, unsigned long long* counters);  // kCounterWords
void launch_clear_film(hipStream_t stream, const DevFilm& film, uint32_t pixels);
// Dynamic-instance refit (CpuBvhAccelerator.cpp:310-358, :527-603). One record per updated instance:
// its primitive range, local geometry and new TRS; tri_begin is the prefix of triangle counts.
struct RefitInst {
  uint32_t prim_first, tri_count, local_first_vertex, local_vertex_count;
  uint32_t local_first_index, tri_begin, pad0, pad1;
  float t[3], q[4], s[3], pad2[2];
};
// Recompute the updated instances' triangles (leaf-order records, primitive normals, padded boxes).
void launch_refit_tris(hipStream_t stream, const RefitInst* insts, uint32_t inst_count, uint32_t tri_count,
                       const float* local_v, const uint32_t* local_i, const uint32_t* leaf_pos, float4* tris,
                       float4* prims, float* tri_box);
// Refit inner nodes order[begin, begin + count) (their children already refit); flags non-finite boxes.
void launch_refit_nodes(hipStream_t stream, float4* nodes, const uint32_t* order, uint32_t begin, uint32_t count,
                        const float* tri_box, uint32_t* nonfinite);
// The scene record's root box from the refit root (inner node or leaf).
void launch_refit_root(hipStream_t stream, const float4* nodes, int32_t root_ref, const float* tri_box,
                       SceneRec* rec);
// Copies the refit boxes of the hot nodes' originals (src[h], h < count) into their hot copies at
// node indices hot_base + h (child references stay).
void launch_hot_refresh(hipStream_t stream, float4* nodes, const uint32_t* src, uint32_t hot_base, uint32_t count);
// resolve_hdr (rs == nullptr: 3 floats per pixel into out) or resolve_ldr (rgba8 per pixel into out).
void launch_resolve(hipStream_t stream, const DevFilm& film, uint32_t pixels, const DevResolve* rs, void* out);

// Workgroups of each kernel that are resident on the device at once (its persistent grid): a
// persistent kernel launched with more would leave the extra waves' chunks to a late tail.
struct Grids {
  uint32_t extend, shadow, shadow_adj, query, shade, shade_adj;
};
Grids resident_grids(int device);

// Allocation-free ray queries (pt_hip_trace_rays_noalloc): their scratch is one static device array per
// GPU. noalloc_setup (once per device, with that device current) looks its address up; noalloc_scratch
// returns the pieces (all null before a successful setup on that device).
constexpr uint32_t kNoAllocBlocks = 512;  // largest persistent grid the spill area is sized for
// spilled stack entries per lane: a tree of PT_NOALLOC_MAX_TREE_DEPTH levels needs that + 2 entries
constexpr uint32_t kNoAllocSpill = PT_NOALLOC_MAX_TREE_DEPTH + 2u - static_cast<uint32_t>(kStackLds);
struct NoAllocScratch {
  unsigned long long* counters;  // kCounterWords
  uint32_t* queue;               // kQueueWords
  uint32_t* error;
  int32_t* stack_ref;
  float* stack_t;
};
hipError_t noalloc_setup(int device);
NoAllocScratch noalloc_scratch(int device);
// Clears the scratch's counters, queue block and error flag, then runs the queries (launch_trace_rays)
// on a grid of min(grid, kNoAllocBlocks) workgroups over the scratch.
void launch_trace_rays_noalloc(hipStream_t stream, const NoAllocScratch& sc, uint32_t refill_min, const DevScene& scene,
                               int any_hit, const float* rays, const float* tmax, uint32_t count, float* out_t,
                               int32_t* out_prim, uint32_t grid);

}  // namespace ptd
