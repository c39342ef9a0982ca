// pt_capi.cpp — the C ABI (include/vkpt_hip.h): IPathTracer state machine, the scene pool in
// device memory, and the host loop that drives the wavefront kernels. A pt_tracer renders one
// scene; a pt_batch renders many scenes through the same pool and the same launches.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <new>
#include <queue>
#include <string>
#include <thread>
#include <vector>

#include "pt_build.h"
#include "pt_bvh.h"
#include "pt_device.h"
#include "pt_math.h"
#include "vkpt_hip.h"

static_assert(sizeof(PtFilmResolve) == 24, "PtFilmResolve layout");
static_assert(sizeof(PtSdfPrimitive) == 56, "PtSdfPrimitive layout");
static_assert(sizeof(ptd::DevSdf) == 96, "DevSdf layout");
static_assert(sizeof(PtInstanceTransformUpdate) == 64, "PtInstanceTransformUpdate layout");
static_assert(sizeof(PtInstanceTransformResult) == 24, "PtInstanceTransformResult layout");

namespace {

constexpr uint32_t kDefaultCapacity = 128u * 1024u * 1024u;  // paths per pass: 64 spp of 1080p in one pass (~20 GB)
constexpr uint32_t kDefaultRefillMin = 8u;

template <typename T>
struct DevArray {
  T* ptr = nullptr;
  size_t n = 0;
  void release() {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
    n = 0;
  }
  hipError_t ensure(size_t count) {
    if (count <= n && ptr) return hipSuccess;
    release();
    const hipError_t e = hipMalloc(reinterpret_cast<void**>(&ptr), std::max<size_t>(count, 1) * sizeof(T));
    if (e == hipSuccess) n = count;
    else ptr = nullptr;
    return e;
  }
};

struct V3h {
  float x, y, z;
};
V3h sub3(V3h a, V3h b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
float dot3(V3h a, V3h b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
V3h cross3(V3h a, V3h b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
V3h normalize3(V3h a) {  // ScalarCpuPathTracer.cpp:104-110
  const float l = std::sqrt(dot3(a, a));
  if (l <= 1e-4f) return {0.0f, 1.0f, 0.0f};
  return {a.x / l, a.y / l, a.z / l};
}

struct ProfiledLaunch {
  int kernel;
  hipEvent_t start, stop;
};

// One scene: the host copy of its snapshot, its image, its tree.
struct SceneSlot {
  std::vector<float> vertices, local_vertices, environment_map;
  std::vector<uint32_t> indices, local_indices;
  std::vector<PtInstance> instances;
  std::vector<PtMaterial> materials;
  std::vector<PtLight> lights;
  PtSceneView view{};
  std::vector<float> texcoords;  // 2 per vertex (may be empty)
  std::vector<float> texels;     // rgb
  std::vector<PtTextureDesc> textures;
  std::vector<PtSdfPrimitive> sdfs;  // shapes the reference intersects (Unknown dropped)
  int32_t view_of = -1;          // >= 0: another camera on that scene's geometry (batches)
  uint32_t width = 0, height = 0, spp = 0;  // spp: samples per batch frame
  PtBvhParams bvh_params{PT_BVH_SAH, 4u, 8u, 0u};
  ptb::Bvh bvh;
  bool built_on_device = false;  // the last build ran on the GPU
  ptb::DeviceLayout layout;      // the host layout of a tree built on the host (empty for device trees)
  // the tree as the pool sees it: inner nodes, root reference (scene-local) and root box; during a
  // build, where the device kept the tree (forest, scene; forest -1: the host layout holds it)
  uint32_t inner_count = 0;
  int32_t root_ref = 0;
  float root_box[6] = {0, 0, 0, 0, 0, 0};
  int forest = -1;
  uint32_t forest_scene = 0;
  std::vector<uint32_t> inst_first;  // per instance its first primitive, then the primitive count
  PtBuildInfo info{};
  std::string error;
};

// One wavefront pass of a render: tiles [first, first + count) of the plan (grouped by scene,
// render order kept within a scene), `paths` paths, and its film jobs.
struct Pass {
  size_t first = 0, count = 0;
  uint32_t paths = 0;
  size_t job_first = 0, job_count = 0;
  uint32_t job_waves = 0;
};

}  // namespace

struct pt_tracer {
  int device = 0;
  hipStream_t stream = nullptr;
  bool configured = false, scene_loaded = false, accel_valid = false;
  bool batch = false;
  PtRenderSettings settings{};
  std::string last_error;
  uint32_t current_sample = 0;
  uint64_t accumulation_gen = 0;

  std::vector<SceneSlot> scenes;  // a tracer uses scenes[0]
  PtBuildInfo info{};             // the pool's build (a tracer: its scene's)
  uint32_t film_pixels = 0;

  DevArray<float4> d_nodes, d_tris, d_prims, d_film_sum;
  DevArray<PtMaterial> d_materials;
  DevArray<PtLight> d_lights;
  DevArray<ptd::SceneRec> d_recs;
  DevArray<ptd::TexRec> d_texrecs;
  DevArray<ptd::DevSdf> d_sdfs;
  DevArray<float> d_texels, d_prim_uv;
  DevArray<uint32_t> d_prim_tri;
  DevArray<uint32_t> d_hot_src;    // the pool node each hot copy duplicates (refreshed after a refit)
  DevArray<float> d_grad_partial;  // adjoint: per-block gradient rows
  size_t texel_count = 0;
  DevArray<uint32_t> d_film_count, d_queue, d_error;
  DevArray<uint32_t> d_resolve;  // resolve_hdr / resolve_ldr output before its copy to the caller
  DevArray<unsigned long long> d_counters, d_query_counters;
  // pt_hip_intersect: staging for rays, limits and RayQueryHit records, and the instance material of
  // every primitive as the last build baked it (rebuilt after each build)
  DevArray<uint8_t> d_iq;
  hipError_t noalloc_status = hipErrorNotInitialized;  // pt_hip_trace_rays_noalloc's scratch lookup
  DevArray<uint32_t> d_inst_mat;
  bool inst_mat_ready = false;
  DevArray<float> d_env;
  DevArray<float> d_path_f;
  DevArray<uint32_t> d_path_u;
  DevArray<uint64_t> d_rng;
  DevArray<int32_t> d_stack_ref;
  DevArray<float> d_stack_t;
  DevArray<float> d_adj_f;
  DevArray<uint32_t> d_adj_u;
  DevArray<uint8_t> d_plan;     // tile tables and film jobs of the current render
  DevArray<uint64_t> d_frame;   // frame index read by raygen in graph replays
  uint32_t adj_capacity_alloc = 0;
  uint32_t path_capacity_alloc = 0;
  uint32_t capacity = kDefaultCapacity;
  uint32_t refill_min = kDefaultRefillMin;

  // dynamic-instance refit: padded triangle boxes in leaf order, local geometry, inner nodes ordered by
  // height (refit_level[h] = first node of height h + 1), updated instance records, non-finite flag
  bool refit_ready = false;
  DevArray<float> d_tri_box, d_local_v;
  DevArray<uint32_t> d_local_i, d_leaf_pos, d_refit_order, d_refit_flag;
  DevArray<ptd::RefitInst> d_refit_insts;
  std::vector<uint32_t> refit_level;

  // split film readback: the resolved image is copied on copy_stream into pinned host memory while
  // the main stream goes on rendering; rb_copied marks the copy's completion
  hipStream_t copy_stream = nullptr;
  hipEvent_t rb_resolved = nullptr, rb_copied = nullptr;
  DevArray<uint32_t> d_readback;
  uint8_t* rb_host = nullptr;
  size_t rb_host_bytes = 0;
  bool rb_inflight = false;
  uint64_t rb_next_id = 1;
  PtFilmReadbackToken rb_latest{};

  // pinned staging for host->device uploads that must follow the stream's earlier work
  void* stage = nullptr;
  size_t stage_bytes = 0;
  hipEvent_t stage_done = nullptr;
  bool stage_pending = false;

  // Per-path radiance of the last forward render, still in the path arrays when that render was
  // a single pass: an adjoint of the same tiles, frame and scene reuses it instead of replaying
  // the forward pass. Any change of scene, settings or path arrays clears it.
  bool fwd_rad_valid = false;
  std::vector<ptd::DevTile> fwd_work;
  uint64_t fwd_frame = 0;
  uint64_t scene_version = 0, fwd_scene_version = 0;

  // batch frame captured as a HIP graph (launch sequence of every pass)
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;

  ptd::DevScene dscene{};
  ptd::DevPaths paths{};
  ptd::Grids grids{}, grids_max{};  // launch grids in use, occupancy-derived maxima

  bool profiling = false;
  std::vector<ProfiledLaunch> pending;
  std::vector<hipEvent_t> event_pool;
  PtKernelStats kstats{};
};

struct pt_batch {
  pt_tracer t;
};

namespace {

int fail(pt_tracer* t, int code, const std::string& msg) {
  if (t) t->last_error = msg;
  return code;
}

int hip_fail(pt_tracer* t, hipError_t e, const char* what) {
  const int code = (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) ? PT_ALLOC_FAILED : PT_INTERNAL_ERROR;
  return fail(t, code, std::string(what) + ": " + hipGetErrorString(e));
}

#define PT_HIP(t, expr, what)                   \
  do {                                          \
    const hipError_t e_ = (expr);               \
    if (e_ != hipSuccess) return hip_fail((t), e_, (what)); \
  } while (0)

hipEvent_t take_event(pt_tracer* t) {
  if (!t->event_pool.empty()) {
    hipEvent_t e = t->event_pool.back();
    t->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Launch wrapper: brackets the launch with events while profiling.
template <typename F>
void timed(pt_tracer* t, int kernel, uint64_t items, F&& launch) {
  if (!t->profiling) {
    launch();
    return;
  }
  ProfiledLaunch pl{kernel, take_event(t), take_event(t)};
  if (pl.start) (void)hipEventRecord(pl.start, t->stream);
  launch();
  if (pl.stop) (void)hipEventRecord(pl.stop, t->stream);
  t->pending.push_back(pl);
  t->kstats.launches[kernel] += 1;
  t->kstats.items[kernel] += items;
}

void collect_profile(pt_tracer* t) {
  for (const ProfiledLaunch& pl : t->pending) {
    if (pl.start && pl.stop) {
      float ms = 0.0f;
      if (hipEventSynchronize(pl.stop) == hipSuccess && hipEventElapsedTime(&ms, pl.start, pl.stop) == hipSuccess)
        t->kstats.total_ms[pl.kernel] += ms;
    }
    if (pl.start) t->event_pool.push_back(pl.start);
    if (pl.stop) t->event_pool.push_back(pl.stop);
  }
  t->pending.clear();
}

int check_device_error(pt_tracer* t) {
  if (!t->d_error.ptr) return PT_OK;
  uint32_t flag = 0;
  PT_HIP(t, hipMemcpyAsync(&flag, t->d_error.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, t->stream), "read error flag");
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (flag) {
    (void)hipMemsetAsync(t->d_error.ptr, 0, sizeof(uint32_t), t->stream);
    return fail(t, PT_INTERNAL_ERROR, "traversal stack overflow area exhausted (BVH deeper than the build reported)");
  }
  return PT_OK;
}

int sync_all(pt_tracer* t) {
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (t->profiling) collect_profile(t);
  return check_device_error(t);
}

// Copies bytes to device memory in stream order through the pinned staging buffer (the copy runs
// after the stream's earlier kernels, which may still read dst's previous contents).
int stage_upload(pt_tracer* t, void* dst, const void* src, size_t bytes) {
  if (bytes == 0) return PT_OK;
  if (t->stage_pending) {
    PT_HIP(t, hipEventSynchronize(t->stage_done), "wait for staging buffer");
    t->stage_pending = false;
  }
  if (bytes > t->stage_bytes) {
    if (t->stage) (void)hipHostFree(t->stage);
    t->stage = nullptr;
    t->stage_bytes = 0;
    PT_HIP(t, hipHostMalloc(&t->stage, bytes, hipHostMallocDefault), "allocate staging buffer");
    t->stage_bytes = bytes;
  }
  std::memcpy(t->stage, src, bytes);
  PT_HIP(t, hipMemcpyAsync(dst, t->stage, bytes, hipMemcpyHostToDevice, t->stream), "upload");
  PT_HIP(t, hipEventRecord(t->stage_done, t->stream), "record staging event");
  t->stage_pending = true;
  return PT_OK;
}

void drop_graph(pt_tracer* t) {
  if (t->graph_exec) (void)hipGraphExecDestroy(t->graph_exec);
  if (t->graph) (void)hipGraphDestroy(t->graph);
  t->graph_exec = nullptr;
  t->graph = nullptr;
}

// Sums the striped device counters (ptd::kStripes stripes of ptd::kCounterSlots) into out.
int read_counter_block(pt_tracer* t, const unsigned long long* dev, unsigned long long out[ptd::kCounterSlots]) {
  std::vector<unsigned long long> raw(ptd::kCounterWords);
  PT_HIP(t, hipMemcpyAsync(raw.data(), dev, raw.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost, t->stream),
         "read counters");
  const int rc = sync_all(t);
  if (rc != PT_OK) return rc;
  for (int k = 0; k < ptd::kCounterSlots; ++k) {
    unsigned long long v = 0;
    for (int s = 0; s < ptd::kStripes; ++s) v += raw[static_cast<size_t>(s) * ptd::kCounterSlots + k];
    out[k] = v;
  }
  return PT_OK;
}

void free_scene_device(pt_tracer* t) {
  t->refit_ready = false;
  t->d_sdfs.release();
  t->d_tri_box.release();
  t->d_local_v.release();
  t->d_local_i.release();
  t->d_leaf_pos.release();
  t->d_refit_order.release();
  t->d_refit_flag.release();
  t->d_refit_insts.release();
  drop_graph(t);
  t->d_nodes.release();
  t->d_tris.release();
  t->d_prims.release();
  t->d_materials.release();
  t->d_lights.release();
  t->d_env.release();
  t->d_recs.release();
  t->d_texrecs.release();
  t->d_texels.release();
  t->d_prim_uv.release();
  t->d_prim_tri.release();
  t->d_hot_src.release();
}

void free_paths(pt_tracer* t) {
  drop_graph(t);
  t->fwd_rad_valid = false;
  t->d_path_f.release();
  t->d_path_u.release();
  t->d_rng.release();
  t->path_capacity_alloc = 0;
  t->d_adj_f.release();
  t->d_adj_u.release();
  t->d_grad_partial.release();
  t->adj_capacity_alloc = 0;
}

// Adjoint state: 33 + 2 * kAdjHist float arrays and 4 + kAdjHist u32 arrays of max(cap, kSegments *
// seg_cap) entries.
// Adjoint state: two buffers of queue-ordered items, then the shadow-record and history arrays.
constexpr size_t kAdjFloats = 2 * ptd::kItemFloats + 3 + 3 + 3 * ptd::kAdjHist + 1 + 2 + 2 * ptd::kAdjHist;
constexpr size_t kAdjWords = 2 * ptd::kItemWords + 2 + ptd::kAdjHist;
int ensure_adjoint(pt_tracer* t, uint32_t cap, ptd::DevAdjoint& A) {
  if (cap > t->adj_capacity_alloc) {
    t->d_adj_f.release();
    t->d_adj_u.release();
    t->adj_capacity_alloc = 0;
    const size_t m = std::max<size_t>(cap, static_cast<size_t>(ptd::kSegments) * ptd::segment_capacity(cap));
    PT_HIP(t, t->d_adj_f.ensure(kAdjFloats * m), "allocate adjoint state");
    PT_HIP(t, t->d_adj_u.ensure(kAdjWords * m), "allocate adjoint state");
    t->adj_capacity_alloc = cap;
  }
  const size_t n = std::max<size_t>(t->adj_capacity_alloc,
                                    static_cast<size_t>(ptd::kSegments) * ptd::segment_capacity(t->adj_capacity_alloc));
  float* f = t->d_adj_f.ptr;
  uint32_t* u = t->d_adj_u.ptr;
  for (int b = 0; b < 2; ++b) {
    A.items[b].f = f + static_cast<size_t>(b) * ptd::kItemFloats * n;
    A.items[b].u = u + static_cast<size_t>(b) * ptd::kItemWords * n;
  }
  float* fs = f + 2 * ptd::kItemFloats * n;  // shadow-record and history floats
  for (int c = 0; c < 3; ++c) {
    A.s_galb[c] = fs + (0 + c) * n;
    A.s_glight[c] = fs + (3 + c) * n;
    for (int h = 0; h < ptd::kAdjHist; ++h) A.h_dx[h][c] = fs + (6 + 3 * h + c) * n;
  }
  A.s_grough = fs + (6 + 3 * ptd::kAdjHist) * n;
  for (int k = 0; k < 2; ++k) {
    A.s_uv[k] = fs + (7 + 3 * ptd::kAdjHist + k) * n;
    for (int h = 0; h < ptd::kAdjHist; ++h) A.h_uv[h][k] = fs + (9 + 3 * ptd::kAdjHist + 2 * h + k) * n;
  }
  A.stride = static_cast<uint32_t>(n);
  uint32_t* us = u + 2 * ptd::kItemWords * n;
  A.s_mat = us;
  A.s_light = us + n;
  for (int h = 0; h < ptd::kAdjHist; ++h) A.h_mat[h] = us + (2 + h) * n;
  return PT_OK;
}

// Path-state and queue arrays for `cap` paths: 38 float arrays, 8 u32 arrays, 2 u64 arrays, each
// of max(cap, kSegments * seg_cap) entries (queue arrays are indexed by segment position).
int ensure_paths(pt_tracer* t, uint32_t cap) {
  if (cap <= t->path_capacity_alloc) return PT_OK;
  free_paths(t);
  constexpr size_t kF = 38, kU = 8;
  const uint32_t seg_cap = ptd::segment_capacity(cap);
  const size_t n = std::max<size_t>(cap, static_cast<size_t>(ptd::kSegments) * seg_cap);
  // shadow records name a queue entry with the top bit set (ptd::kRadQueued)
  if (n >= static_cast<size_t>(ptd::kRadQueued)) return fail(t, PT_INVALID_ARGUMENT, "wavefront capacity too large");
  PT_HIP(t, t->d_path_f.ensure(kF * n), "allocate path state");
  PT_HIP(t, t->d_path_u.ensure(kU * n), "allocate path state");
  PT_HIP(t, t->d_rng.ensure(2 * n), "allocate path state");
  float* f = t->d_path_f.ptr;
  uint32_t* u = t->d_path_u.ptr;
  ptd::DevPaths& P = t->paths;
  for (int b = 0; b < 2; ++b) {
    for (int k = 0; k < 6; ++k) P.rays[b].c[k] = f + (size_t)(6 * b + k) * n;
    P.rays[b].qpath = u + (size_t)b * n;
    P.rays[b].qscene = u + (size_t)(4 + b) * n;
    for (int k = 0; k < 3; ++k) {
      P.rays[b].tp[k] = f + (size_t)(26 + 6 * b + k) * n;
      P.rays[b].rad[k] = f + (size_t)(29 + 6 * b + k) * n;
    }
    P.rays[b].rng = t->d_rng.ptr + (size_t)b * n;
  }
  P.hit_t = f + 12 * n;
  P.hit_prim = reinterpret_cast<int32_t*>(u + 2 * n);
  for (int k = 0; k < 3; ++k) {
    P.rad[k] = f + (size_t)(13 + k) * n;
    P.scontrib[k] = f + (size_t)(23 + k) * n;
  }
  for (int k = 0; k < 6; ++k) P.sray[k] = f + (size_t)(16 + k) * n;
  P.stmax = f + 22 * n;
  P.spath = u + 3 * n;
  P.sscene = u + (size_t)6 * n;
  P.ended_rays = u + (size_t)7 * n;
  P.capacity = cap;
  P.seg_cap = seg_cap;
  t->path_capacity_alloc = cap;
  return PT_OK;
}

bool same_work(const std::vector<ptd::DevTile>& a, const std::vector<ptd::DevTile>& b) {
  if (a.size() != b.size()) return false;
  for (size_t i = 0; i < a.size(); ++i)
    if (a[i].x != b[i].x || a[i].y != b[i].y || a[i].w != b[i].w || a[i].h != b[i].h || a[i].sample != b[i].sample ||
        a[i].path_begin != b[i].path_begin || a[i].scene != b[i].scene || a[i].group != b[i].group)
      return false;
  return true;
}

// ------------------------------------------------------------------ scenes
// PathTracerSceneSnapshot copy (ScalarCpuPathTracer.cpp:458-477): an empty material list gets one
// default material (:469-471); the environment map is used only when it holds width * height
// texels (PathTracer.cpp:216-219).
void copy_scene(SceneSlot& s, const PtSceneView* sc) {
  s.sdfs.clear();
  auto copy = [](auto& dst, const auto* src, size_t n) {
    if (src && n) dst.assign(src, src + n);
    else dst.clear();
  };
  copy(s.vertices, sc->vertices, static_cast<size_t>(sc->vertex_count) * 3u);
  copy(s.indices, sc->indices, sc->index_count);
  copy(s.local_vertices, sc->local_vertices, static_cast<size_t>(sc->local_vertex_count) * 3u);
  copy(s.local_indices, sc->local_indices, sc->local_index_count);
  copy(s.instances, sc->instances, sc->instance_count);
  copy(s.materials, sc->materials, sc->material_count);
  copy(s.lights, sc->lights, sc->light_count);
  const size_t env_texels = static_cast<size_t>(sc->environment_map_width) * sc->environment_map_height;
  copy(s.environment_map, sc->environment_map, env_texels * 3u);
  if (s.materials.empty()) {
    PtMaterial m{};
    m.albedo[0] = m.albedo[1] = m.albedo[2] = 1.0f;
    m.roughness = 1.0f;
    m.ior = 1.5f;
    m.alpha = 1.0f;
    m.base_color_texture_index = m.normal_texture_index = 0xFFFFFFFFu;
    s.materials.push_back(m);
  }
  s.view = *sc;
  s.view.vertices = s.vertices.data();
  s.view.indices = s.indices.data();
  s.view.local_vertices = s.local_vertices.data();
  s.view.local_indices = s.local_indices.data();
  s.view.instances = s.instances.data();
  s.view.materials = s.materials.data();
  s.view.lights = s.lights.data();
  s.view.environment_map = nullptr;
  s.view.vertex_count = static_cast<uint32_t>(s.vertices.size() / 3u);
  s.view.index_count = static_cast<uint32_t>(s.indices.size());
  s.view.local_vertex_count = static_cast<uint32_t>(s.local_vertices.size() / 3u);
  s.view.local_index_count = static_cast<uint32_t>(s.local_indices.size());
  s.view.instance_count = static_cast<uint32_t>(s.instances.size());
  s.view.material_count = static_cast<uint32_t>(s.materials.size());
  s.view.light_count = static_cast<uint32_t>(s.lights.size());
  s.layout = ptb::DeviceLayout{};
  s.bvh = ptb::Bvh{};
  s.inner_count = 0;
  s.root_ref = 0;
  s.forest = -1;
  s.texcoords.clear();
  s.texels.clear();
  s.textures.clear();
  s.view_of = -1;
}

// Copies a scene's texture coordinates and textures after checking that every texture lies in
// the texel array and the coordinates cover the scene's vertices.
int copy_textures(pt_tracer* t, SceneSlot& s, const float* texcoords, uint32_t texcoord_count, const float* texels,
                  uint64_t texel_count, const PtTextureDesc* textures, uint32_t texture_count) {
  if (texcoord_count && texcoord_count != s.vertices.size() / 3u)
    return fail(t, PT_INVALID_ARGUMENT, "load_textures: texcoord_count differs from the vertex count");
  if ((texcoord_count && !texcoords) || (texel_count && !texels) || (texture_count && !textures))
    return fail(t, PT_INVALID_ARGUMENT, "load_textures: null array");
  for (uint32_t i = 0; i < texture_count; ++i) {
    const PtTextureDesc& d = textures[i];
    if (d.width == 0 || d.height == 0 || d.width > 65536u || d.height > 65536u)
      return fail(t, PT_INVALID_ARGUMENT, "load_textures: texture " + std::to_string(i) + " has a bad size");
    if (static_cast<uint64_t>(d.offset) + static_cast<uint64_t>(d.width) * d.height > texel_count)
      return fail(t, PT_INVALID_ARGUMENT, "load_textures: texture " + std::to_string(i) + " exceeds the texels");
  }
  s.texcoords.assign(texcoords, texcoords + 2u * static_cast<size_t>(texcoord_count));
  s.texels.assign(texels, texels + 3u * static_cast<size_t>(texel_count));
  s.textures.assign(textures, textures + texture_count);
  return PT_OK;
}

int check_bvh_params(pt_tracer* t, const PtBvhParams* params, PtBvhParams& p) {
  p = PtBvhParams{PT_BVH_SAH, 4u, 8u, 0u};
  if (params) p = *params;
  if (p.leaf_size == 0) p.leaf_size = 4u;
  if (p.bucket_count == 0) p.bucket_count = 8u;
  if (p.leaf_size > ptb::kMaxLeafSize) return fail(t, PT_INVALID_ARGUMENT, "leaf_size exceeds 31");
  if (p.builder != PT_BVH_SAH && p.builder != PT_BVH_MIDPOINT) return fail(t, PT_INVALID_ARGUMENT, "unknown builder");
  return PT_OK;
}

// A scene's build in three steps (each thread-safe per slot): flatten its triangles; its tree, the SAH
// tree on the GPU (the host builder's tree node for node; several scenes at once in one device build,
// kept on the device for the pool) unless the parameters ask for the host or a vertex is not finite, the
// midpoint tree on the host; then, for a tree built on the host, the device layout, packed by `threads`
// host threads.
struct FlatScene {
  ptb::HostVec<float> raw;
  ptb::HostVec<uint32_t> material;
  bool bounded = false;
};

bool flatten_scene(SceneSlot& s, FlatScene& f, unsigned threads) {
  s.error.clear();
  return ptb::flatten(s.view, s.bvh.prims, f.raw, s.error, s.texcoords.empty() ? nullptr : s.texcoords.data(),
                      &s.inst_first, threads, &f.material, &f.bounded);
}

bool wants_device_build(const SceneSlot& s) {
  return s.bvh_params.builder == PT_BVH_SAH && !(s.bvh_params.flags & PT_BVH_FLAG_HOST_BUILD);
}

ptg::DeviceBuild device_job(SceneSlot& s, const FlatScene& f) {
  ptg::DeviceBuild d{&s.bvh, &f.raw, s.bvh_params.leaf_size, s.bvh_params.bucket_count, 1};
  d.material = &f.material;
  d.bounded = f.bounded ? 1 : 0;
  return d;
}

// `kept`: the device built the tree and kept it (its forest scene); otherwise it is built here.
void finish_scene(SceneSlot& s, const FlatScene& f, const ptg::ForestScene* kept, unsigned threads, double shared_ms,
                  std::chrono::steady_clock::duration own) {
  const auto t0 = std::chrono::steady_clock::now();
  s.built_on_device = kept != nullptr;
  uint32_t node_count = 0;
  if (kept) {
    s.layout = ptb::DeviceLayout{};
    node_count = kept->node_count;
    s.inner_count = (node_count - 1u) / 2u;  // every inner node has two children
    const ptb::Node& r = kept->root;
    s.root_ref = r.count > 0 ? ~static_cast<int32_t>(static_cast<uint32_t>(r.count)) : 0;  // leaf: first 0
    for (int k = 0; k < 3; ++k) {
      s.root_box[k] = r.mn[k];
      s.root_box[3 + k] = r.mx[k];
    }
  } else {
    s.forest = -1;
    if (s.bvh_params.builder == PT_BVH_SAH) ptb::build_sah(s.bvh, f.raw, s.bvh_params.leaf_size, s.bvh_params.bucket_count);
    else ptb::build_midpoint(s.bvh, f.raw, s.bvh_params.leaf_size);
    ptb::pack_device_layout(s.bvh, s.layout, true, threads);
    node_count = static_cast<uint32_t>(s.bvh.nodes.size());
    s.inner_count = s.layout.inner_count;
    s.root_ref = s.layout.root_ref;
    std::memcpy(s.root_box, s.layout.root_box, sizeof(s.root_box));
  }
  s.info = PtBuildInfo{};
  s.info.build_ms = shared_ms + std::chrono::duration<double, std::milli>(own + (std::chrono::steady_clock::now() - t0))
                                    .count();
  s.info.built = 1;
  s.info.primitive_count = static_cast<uint32_t>(s.bvh.prims.size());
  s.info.node_count = node_count;
  s.info.leaf_count = s.bvh.leaf_count;
  s.info.max_depth = s.bvh.max_depth;
  s.info.deterministic = 1;
  s.info.device_built = s.built_on_device ? 1u : 0u;
}

// Builds jobs[i] for scenes[i] (those that want the device) into `forests`; records where each landed.
bool build_on_device(hipStream_t stream, std::vector<SceneSlot*>& scenes, std::vector<ptg::DeviceBuild>& jobs,
                     std::vector<ptg::DeviceForest>& forests, std::string& err) {
  if (jobs.empty()) return true;
  if (ptg::build_sah_device_many(stream, jobs, err, &forests) < 0) return false;
  for (size_t j = 0; j < jobs.size(); ++j) {
    scenes[j]->forest = jobs[j].rc == 0 ? jobs[j].forest : -1;
    scenes[j]->forest_scene = jobs[j].scene;
  }
  return true;
}

const ptg::ForestScene* kept_tree(const SceneSlot& s, const std::vector<ptg::DeviceForest>& forests) {
  return s.forest >= 0 ? &forests[static_cast<size_t>(s.forest)].scenes[s.forest_scene] : nullptr;
}

bool build_scene(SceneSlot& s, hipStream_t stream, unsigned threads, std::vector<ptg::DeviceForest>& forests) {
  const auto t0 = std::chrono::steady_clock::now();
  FlatScene f;
  s.forest = -1;
  if (!flatten_scene(s, f, threads)) return false;
  if (wants_device_build(s)) {
    std::vector<SceneSlot*> one{&s};
    std::vector<ptg::DeviceBuild> jobs{device_job(s, f)};
    if (!build_on_device(stream, one, jobs, forests, s.error)) return false;
  }
  finish_scene(s, f, kept_tree(s, forests), threads, 0.0, std::chrono::steady_clock::now() - t0);
  return true;
}

// compute_camera_basis, ScalarCpuPathTracer.cpp:139-152; camera_rays constants, Render.cpp:74-83
ptd::DevCamera make_camera(const PtCamera& c, uint32_t width, uint32_t height, float aperture_default,
                           float focus_default) {
  const V3h pos{c.position[0], c.position[1], c.position[2]};
  const V3h fwd = normalize3(sub3(V3h{c.target[0], c.target[1], c.target[2]}, pos));
  V3h right = normalize3(cross3(fwd, V3h{c.up[0], c.up[1], c.up[2]}));
  if (dot3(right, right) <= 1e-4f * 1e-4f) right = {1.0f, 0.0f, 0.0f};
  V3h up = normalize3(cross3(right, fwd));
  if (dot3(up, up) <= 1e-4f * 1e-4f) up = {0.0f, 1.0f, 0.0f};
  ptd::DevCamera C{};
  const V3h* src[4] = {&pos, &fwd, &right, &up};
  float* dst[4] = {C.pos, C.fwd, C.right, C.up};
  for (int i = 0; i < 4; ++i) {
    dst[i][0] = src[i]->x;
    dst[i][1] = src[i]->y;
    dst[i][2] = src[i]->z;
  }
  const float kPi = 3.14159265358979323846f;
  C.aspect = static_cast<float>(width) / std::max(1.0f, static_cast<float>(height));
  C.tan_half_fov = std::tan(0.5f * (c.fov_deg * (kPi / 180.0f)));
  C.aperture = c.aperture_radius > 0.0f ? c.aperture_radius : aperture_default;
  C.focus = c.focus_distance > 0.0f ? c.focus_distance : focus_default;
  C.iris_blades = std::min(c.iris_blade_count, 64u);
  C.iris_roundness = std::isfinite(c.iris_roundness) ? std::min(1.0f, std::max(0.0f, c.iris_roundness)) : 0.0f;
  C.iris_rotation = c.iris_rotation_degrees * (kPi / 180.0f);
  C.anamorphic = std::isfinite(c.anamorphic_squeeze) ? std::max(0.01f, c.anamorphic_squeeze) : 1.0f;
  return C;
}

bool transmissive(const PtMaterial& m) { return m.material_model == 5u || m.transmission > 0.05f; }

int32_t rebase_ref(int32_t ref, uint32_t node_base, uint32_t tri_base) {
  if (ref >= 0) return ref + static_cast<int32_t>(node_base);
  const uint32_t code = ~static_cast<uint32_t>(ref);
  const uint32_t first = (code >> ptb::kLeafCountBits) + tri_base, count = code & ptb::kMaxLeafSize;
  return ~static_cast<int32_t>((first << ptb::kLeafCountBits) | count);
}

// A scene's materials as the device sees them: texture indices rebased into the texture pool. An
// index beyond the scene's textures, or a scene without texture coordinates, means untextured.
std::vector<PtMaterial> pool_materials(const SceneSlot& s, uint32_t tex_base) {
  std::vector<PtMaterial> out = s.materials;
  for (PtMaterial& m : out)
    m.base_color_texture_index = m.base_color_texture_index < s.textures.size() && !s.texcoords.empty()
                                     ? m.base_color_texture_index + tex_base
                                     : 0xFFFFFFFFu;
  return out;
}

// Packs every scene's tree, materials, lights, textures and environment into the device pool,
// rebasing node, triangle, primitive, material and texture references, and writes the scene
// records. Films are laid out scene after scene (film_base); a tracer's single scene starts at
// pixel 0. A view shares its scene's pool entries and has its own camera and film.
int upload_pool(pt_tracer* t, const std::vector<ptg::DeviceForest>& forests) {
  t->inst_mat_ready = false;
  t->refit_ready = false;
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  free_scene_device(t);
  size_t n_nodes = 0, n_tris = 0, n_prims = 0, n_mats = 0, n_lights = 0, n_env = 0, n_tex = 0, n_texels = 0;
  uint32_t max_depth = 0;
  bool any_uv = false;
  for (const SceneSlot& s : t->scenes) {
    if (s.view_of >= 0) continue;
    n_nodes += s.inner_count;
    n_tris += s.info.primitive_count;
    n_prims += s.info.primitive_count;
    n_mats += s.materials.size();
    n_lights += s.lights.size();
    n_env += s.environment_map.size();
    n_tex += s.textures.size();
    n_texels += s.texels.size() / 3u;
    max_depth = std::max(max_depth, s.bvh.max_depth);
    any_uv = any_uv || !s.texcoords.empty();
  }
  if (n_tris >= ptb::kMaxPrimitives || n_nodes >= 0x7fffffffu)
    return fail(t, PT_UNSUPPORTED, "scene pool exceeds 2^26 triangles");
  if (n_texels >= (1ull << 32)) return fail(t, PT_UNSUPPORTED, "texture pool exceeds 2^32 texels");
  // The trees' arrays are written in place: trees the device build kept are packed on the device,
  // host-built trees uploaded from their host layouts; the hot copies go after the nodes.
  PT_HIP(t, t->d_nodes.ensure((n_nodes + ptd::kHotNodes) * 4u), "allocate BVH nodes");
  PT_HIP(t, t->d_tris.ensure(n_tris * 3u), "allocate triangles");
  PT_HIP(t, t->d_prims.ensure(n_prims), "allocate primitives");
  PT_HIP(t, t->d_hot_src.ensure(std::max<size_t>(ptd::kHotNodes, 1)), "allocate hot node table");
  if (any_uv) {  // scenes without texture coordinates keep zeros here
    PT_HIP(t, t->d_prim_uv.ensure(6u * n_prims), "allocate texture coordinates");
    PT_HIP(t, t->d_prim_tri.ensure(n_prims), "allocate triangle map");
    PT_HIP(t, hipMemsetAsync(t->d_prim_uv.ptr, 0, 6u * n_prims * sizeof(float), t->stream), "clear");
    PT_HIP(t, hipMemsetAsync(t->d_prim_tri.ptr, 0, n_prims * sizeof(uint32_t), t->stream), "clear");
    PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");  // before the host uploads below
  }
  std::vector<float> env(n_env), texels;
  std::vector<ptd::TexRec> texrecs;
  std::vector<PtMaterial> mats;
  std::vector<PtLight> lights;
  std::vector<ptd::SceneRec> recs(t->scenes.size());
  std::vector<ptd::DevSdf> sdfs;
  std::vector<uint32_t> slot_mat_base(t->scenes.size(), 0u);
  std::vector<std::vector<ptg::PoolDest>> dests(forests.size());
  for (size_t f = 0; f < forests.size(); ++f) dests[f].resize(forests[f].scenes.size());
  mats.reserve(n_mats);
  lights.reserve(n_lights);
  texels.reserve(3u * n_texels);
  texrecs.reserve(n_tex);
  uint32_t node_base = 0, tri_base = 0, prim_base = 0, env_base = 0, film_base = 0;
  bool any_transmissive = false, textured = false;
  std::atomic<bool> general_boxes{false};
  const unsigned host_threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (size_t si = 0; si < t->scenes.size(); ++si) {
    const SceneSlot& s = t->scenes[si];
    ptd::SceneRec& R = recs[si];
    if (s.view_of >= 0) {  // the scene's record, with this view's image, film and camera
      R = recs[static_cast<size_t>(s.view_of)];
    } else {
      const uint32_t nn = s.inner_count, np = s.info.primitive_count, nt = np;
      const uint32_t mat_base = static_cast<uint32_t>(mats.size()), light_base = static_cast<uint32_t>(lights.size());
      slot_mat_base[si] = mat_base;
      const uint32_t tex_base = static_cast<uint32_t>(texrecs.size());
      const uint32_t last_mat = static_cast<uint32_t>(s.materials.size()) - 1u;
      const bool with_uv = any_uv && !s.texcoords.empty();
      if (s.forest >= 0) {  // packed on the device below; texture coordinates from the flattening
        dests[static_cast<size_t>(s.forest)][s.forest_scene] =
            ptg::PoolDest{node_base, tri_base, prim_base, mat_base, last_mat, with_uv ? 1u : 0u};
        if (with_uv) {
          ptb::HostVec<float> uv(6u * np);
          ptb::parallel_for(np, host_threads, [&](size_t a, size_t b) {
            for (size_t i = a; i < b; ++i) std::memcpy(&uv[6u * i], s.bvh.prims[i].uv, 6 * sizeof(float));
          });
          PT_HIP(t, hipMemcpy(t->d_prim_uv.ptr + 6u * static_cast<size_t>(prim_base), uv.data(),
                              uv.size() * sizeof(float), hipMemcpyHostToDevice), "upload texture coordinates");
        }
      } else {  // the host layout, rebased into the pool
        const ptb::DeviceLayout& L = s.layout;
        ptb::HostVec<float> nodes(16u * nn), tris(12u * nt), prims(4u * np);  // every element written
        ptb::parallel_for(nn, host_threads, [&, node_base, tri_base](size_t a, size_t b) {
          bool general = false;
          for (size_t i = a; i < b; ++i) {
            const float* src = &L.nodes[16u * i];
            float* dst = &nodes[16u * i];
            std::memcpy(dst, src, 12 * sizeof(float));
            for (int k = 0; k < 12; ++k)
              if (!std::isfinite(src[k])) general = true;
            int32_t r[2];
            std::memcpy(r, src + 12, sizeof(r));
            for (int k = 0; k < 2; ++k) r[k] = rebase_ref(r[k], node_base, tri_base);
            std::memcpy(dst + 12, r, sizeof(r));
            std::memset(dst + 14, 0, 2 * sizeof(float));
          }
          if (general) general_boxes.store(true, std::memory_order_relaxed);
        });
        ptb::parallel_for(nt, host_threads, [&, prim_base](size_t a, size_t b) {
          for (size_t i = a; i < b; ++i) {
            const float* src = &L.tris[12u * i];
            float* dst = &tris[12u * i];
            std::memcpy(dst, src, 12 * sizeof(float));
            uint32_t pid;
            std::memcpy(&pid, src + 3, 4);
            pid += prim_base;
            std::memcpy(dst + 3, &pid, 4);
          }
        });
        ptb::parallel_for(np, host_threads, [&, mat_base, last_mat](size_t a, size_t b) {
          for (size_t i = a; i < b; ++i) {
            const float* src = &L.prims[4u * i];
            float* dst = &prims[4u * i];
            std::memcpy(dst, src, 3 * sizeof(float));
            uint32_t m;
            std::memcpy(&m, src + 3, 4);
            m = std::min(m, last_mat) + mat_base;  // the shading clamp of a scene-local index, then the pool base
            std::memcpy(dst + 3, &m, 4);
          }
        });
        PT_HIP(t, hipMemcpy(t->d_nodes.ptr + 4u * static_cast<size_t>(node_base), nodes.data(),
                            nodes.size() * sizeof(float), hipMemcpyHostToDevice), "upload BVH nodes");
        PT_HIP(t, hipMemcpy(t->d_tris.ptr + 3u * static_cast<size_t>(tri_base), tris.data(), tris.size() * sizeof(float),
                            hipMemcpyHostToDevice), "upload triangles");
        PT_HIP(t, hipMemcpy(t->d_prims.ptr + prim_base, prims.data(), prims.size() * sizeof(float),
                            hipMemcpyHostToDevice), "upload primitives");
        if (with_uv) {
          std::vector<uint32_t> prim_tri(np);
          for (uint32_t i = 0; i < np; ++i) prim_tri[i] = tri_base + L.prim_tri[i];
          PT_HIP(t, hipMemcpy(t->d_prim_uv.ptr + 6u * static_cast<size_t>(prim_base), L.prim_uv.data(),
                              L.prim_uv.size() * sizeof(float), hipMemcpyHostToDevice), "upload texture coordinates");
          PT_HIP(t, hipMemcpy(t->d_prim_tri.ptr + prim_base, prim_tri.data(), np * sizeof(uint32_t),
                              hipMemcpyHostToDevice), "upload triangle map");
        }
      }
      const std::vector<PtMaterial> pm = pool_materials(s, tex_base);
      for (const PtMaterial& m : pm)
        if (m.base_color_texture_index != 0xFFFFFFFFu) textured = true;
      mats.insert(mats.end(), pm.begin(), pm.end());
      lights.insert(lights.end(), s.lights.begin(), s.lights.end());
      std::copy(s.environment_map.begin(), s.environment_map.end(), env.begin() + env_base);
      const uint32_t texel_base = static_cast<uint32_t>(texels.size() / 3u);
      for (const PtTextureDesc& d : s.textures)
        texrecs.push_back(ptd::TexRec{texel_base + d.offset, d.width, d.height, 0u});
      texels.insert(texels.end(), s.texels.begin(), s.texels.end());

      R = ptd::SceneRec{};
      R.root_ref = np ? rebase_ref(s.root_ref, node_base, tri_base) : 0;
      R.prim_count = np;
      for (int k = 0; k < 6; ++k) R.root_box[k] = s.root_box[k];
      R.light_base = light_base;
      R.light_count = static_cast<uint32_t>(s.lights.size());
      R.has_transmissive = 0;
      for (const PtMaterial& m : s.materials)
        if (transmissive(m)) R.has_transmissive = 1;
      any_transmissive = any_transmissive || R.has_transmissive;
      R.env_base = env_base;
      const bool has_map = !s.environment_map.empty();
      R.env_w = has_map ? s.view.environment_map_width : 0u;
      R.env_h = has_map ? s.view.environment_map_height : 0u;
      for (int k = 0; k < 3; ++k) {
        R.env[k] = s.view.environment_color[k];
        R.env_scale[k] = s.view.environment_map_scale[k];
      }
      node_base += nn;
      tri_base += nt;
      prim_base += np;
      env_base += static_cast<uint32_t>(s.environment_map.size());
    }
    if (s.view_of < 0) {  // SDF primitives: pool material index (scene clamp, then base), trig of the Euler angles
      R.sdf_base = static_cast<uint32_t>(sdfs.size());
      R.sdf_count = static_cast<uint32_t>(s.sdfs.size());
      const uint32_t last = static_cast<uint32_t>(s.materials.size()) - 1u;
      for (const PtSdfPrimitive& q : s.sdfs) {
        ptd::DevSdf d{};
        d.shape = q.shape;
        d.material = std::min(q.material_index, last) + slot_mat_base[si];
        for (int k = 0; k < 3; ++k) {
          d.pos[k] = q.position[k];
          d.scale[k] = q.scale[k];
          ptm::sincosf_p(-q.rotation[k], d.inv_s[k], d.inv_c[k]);
          ptm::sincosf_p(q.rotation[k], d.fwd_s[k], d.fwd_c[k]);
        }
        d.radius = q.radius;
        d.pa = q.param_a;
        d.pb = q.param_b;
        sdfs.push_back(d);
      }
    }
    R.width = s.width;
    R.height = s.height;
    R.film_base = film_base;
    R.cam = make_camera(s.view.camera, s.width, s.height, t->settings.camera_aperture_radius,
                        t->settings.camera_focus_distance);
    film_base += s.width * s.height;
  }

  for (size_t f = 0; f < forests.size(); ++f) {
    std::string err;
    if (ptg::pack_pool(t->stream, forests[f], dests[f],
                       ptg::PoolArrays{t->d_nodes.ptr, t->d_tris.ptr, t->d_prims.ptr, any_uv ? t->d_prim_tri.ptr : nullptr},
                       err) < 0)
      return fail(t, PT_INTERNAL_ERROR, err);
  }

  // Hot copies for the traversal kernels' LDS: in each tree (the hot budget split evenly between the
  // pool's trees, in pool order) the inner nodes of largest box surface area, taken from the root
  // down (a child's box lies in its parent's, so the set is a top subtree), copied after the pool's
  // nodes with their hot children's references pointing at the copies. Scene records whose root is
  // hot start at its copy. Traversal order and results do not depend on where a node is stored.
  std::vector<uint32_t> hot_src;
  {
    std::vector<int32_t> roots;
    std::vector<float> boxes;
    for (size_t si = 0; si < t->scenes.size(); ++si)
      if (t->scenes[si].view_of < 0 && recs[si].prim_count && recs[si].root_ref >= 0) {
        roots.push_back(recs[si].root_ref);
        boxes.insert(boxes.end(), recs[si].root_box, recs[si].root_box + 6);
      }
    const uint32_t per = roots.empty() ? 0u : std::max<uint32_t>(1u, ptd::kHotNodes / static_cast<uint32_t>(roots.size()));
    std::string err;
    if (ptg::select_hot_nodes(t->stream, t->d_nodes.ptr, static_cast<uint32_t>(n_nodes), roots, boxes, per,
                              ptd::kHotNodes, t->d_hot_src.ptr, hot_src, err) < 0)
      return fail(t, PT_INTERNAL_ERROR, err);
  }
  const uint32_t hot_base = static_cast<uint32_t>(n_nodes), hot_count = static_cast<uint32_t>(hot_src.size());
  for (ptd::SceneRec& R : recs)
    if (R.prim_count && R.root_ref >= 0)
      for (uint32_t h = 0; h < hot_count; ++h)
        if (hot_src[h] == static_cast<uint32_t>(R.root_ref)) {
          R.root_ref = static_cast<int32_t>(hot_base + h);
          break;
        }

  uint64_t bytes = (n_nodes + hot_count) * 64u + hot_count * 4u + n_tris * 48u + n_prims * 16u +
                   (any_uv ? n_prims * 28u : 0u);
  auto upload = [&](auto& arr, const void* src, size_t count, size_t elem) -> hipError_t {
    hipError_t e = arr.ensure(count);
    if (e != hipSuccess) return e;
    bytes += count * elem;
    if (count) e = hipMemcpy(arr.ptr, src, count * elem, hipMemcpyHostToDevice);
    return e;
  };
  PT_HIP(t, upload(t->d_materials, mats.data(), mats.size(), sizeof(PtMaterial)), "upload materials");
  PT_HIP(t, upload(t->d_lights, lights.data(), lights.size(), sizeof(PtLight)), "upload lights");
  PT_HIP(t, upload(t->d_env, env.data(), env.size(), sizeof(float)), "upload environment");
  PT_HIP(t, upload(t->d_recs, recs.data(), recs.size(), sizeof(ptd::SceneRec)), "upload scene records");
  PT_HIP(t, upload(t->d_texels, texels.data(), texels.size(), sizeof(float)), "upload textures");
  PT_HIP(t, upload(t->d_texrecs, texrecs.data(), texrecs.size(), sizeof(ptd::TexRec)), "upload texture table");
  PT_HIP(t, upload(t->d_sdfs, sdfs.data(), sdfs.size(), sizeof(ptd::DevSdf)), "upload SDF primitives");
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");

  // Traversal stack spill area: a lane's stack never holds more than max_depth entries (one per
  // level of the current path); entries older than the LDS ring's kStackLds go here.
  const uint32_t need = max_depth + 2u;
  const uint32_t threads =
      std::max({t->grids_max.extend, t->grids_max.shadow, t->grids_max.shadow_adj, t->grids_max.query}) *
      static_cast<uint32_t>(ptd::kBlock);
  if (static_cast<uint64_t>(need) * threads >= (1ull << 32))  // the kernels index the area with 32-bit offsets
    return fail(t, PT_UNSUPPORTED, "traversal stack area exceeds 2^32 entries (tree too deep)");
  PT_HIP(t, t->d_stack_ref.ensure(static_cast<size_t>(need) * threads), "allocate traversal stack");
  PT_HIP(t, t->d_stack_t.ensure(static_cast<size_t>(need) * threads), "allocate traversal stack");
  bytes += static_cast<uint64_t>(need) * threads * 8u;
  t->paths.stack_ref = t->d_stack_ref.ptr;
  t->paths.stack_t = t->d_stack_t.ptr;
  t->paths.stack_overflow_depth = need;
  t->paths.stack_threads = threads;

  ptd::DevScene& S = t->dscene;
  S = ptd::DevScene{};
  S.nodes = t->d_nodes.ptr;
  S.hot_base = hot_base;
  S.hot_count = hot_count;
  S.tris = t->d_tris.ptr;
  S.prims = t->d_prims.ptr;
  S.materials = t->d_materials.ptr;
  S.lights = t->d_lights.ptr;
  S.material_count = static_cast<uint32_t>(mats.size());
  S.light_count = static_cast<uint32_t>(lights.size());
  S.has_transmissive = any_transmissive ? 1u : 0u;
  S.general_boxes = general_boxes.load() ? 1u : 0u;
  S.env_map = t->d_env.ptr;
  S.recs = t->d_recs.ptr;
  S.scene_count = static_cast<uint32_t>(recs.size());
  S.texels = t->d_texels.ptr;
  S.textures = t->d_texrecs.ptr;
  S.texture_count = static_cast<uint32_t>(texrecs.size());
  S.textured = textured ? 1u : 0u;
  S.prim_uv = any_uv ? t->d_prim_uv.ptr : nullptr;
  S.prim_tri = any_uv ? t->d_prim_tri.ptr : nullptr;
  S.sdfs = sdfs.empty() ? nullptr : t->d_sdfs.ptr;
  S.has_sdf = sdfs.empty() ? 0u : 1u;
  t->texel_count = texels.size() / 3u;

  t->info = PtBuildInfo{};
  t->info.device_bytes = bytes;
  t->info.built = 1;
  t->info.deterministic = 1;
  t->info.device_built = 1;
  for (const SceneSlot& s : t->scenes) {
    if (s.view_of >= 0) continue;
    t->info.device_built &= s.info.device_built;
    t->info.build_ms += s.info.build_ms;
    t->info.primitive_count += s.info.primitive_count;
    t->info.node_count += s.info.node_count;
    t->info.leaf_count += s.info.leaf_count;
    t->info.max_depth = std::max(t->info.max_depth, s.info.max_depth);
  }
  t->film_pixels = film_base;
  return PT_OK;
}

int alloc_film(pt_tracer* t, size_t pixels) {
  PT_HIP(t, t->d_film_sum.ensure(pixels), "allocate film");
  PT_HIP(t, t->d_film_count.ensure(pixels), "allocate film");
  ptd::launch_clear_film(t->stream, ptd::DevFilm{t->d_film_sum.ptr, t->d_film_count.ptr}, static_cast<uint32_t>(pixels));
  return PT_OK;
}

// ------------------------------------------------------------------ rendering
// Clamp tiles like render_tile_cancellable (ScalarCpuPathTracer.Render.cpp:172-180) and split
// any tile larger than one pass into row bands.
int build_work(pt_tracer* t, uint32_t scene, const PtRenderTile* tiles, uint32_t count, std::vector<ptd::DevTile>& work) {
  const uint32_t W = t->scenes[scene].width, H = t->scenes[scene].height;
  const uint32_t cap = std::max<uint32_t>(t->capacity, 1024u);
  for (uint32_t i = 0; i < count; ++i) {
    const PtRenderTile& tl = tiles[i];
    const uint32_t x0 = std::min(tl.x, W), y0 = std::min(tl.y, H);
    const uint32_t x1 = static_cast<uint32_t>(std::min<uint64_t>(W, static_cast<uint64_t>(x0) + tl.width));
    const uint32_t y1 = static_cast<uint32_t>(std::min<uint64_t>(H, static_cast<uint64_t>(y0) + tl.height));
    if (x0 >= x1 || y0 >= y1) continue;
    const uint32_t w = x1 - x0;
    if (w > cap) return fail(t, PT_INVALID_ARGUMENT, "tile wider than the wavefront capacity");
    const uint32_t rows_per = std::max(1u, cap / w);
    for (uint32_t y = y0; y < y1; y += rows_per) {
      const uint32_t h = std::min(rows_per, y1 - y);
      work.push_back(ptd::DevTile{x0, y, w, h, tl.sample_index, 0u, scene, 0u});
    }
  }
  return PT_OK;
}

ptd::DevSettings dev_settings(const pt_tracer* t, uint64_t frame_index) {
  ptd::DevSettings st{};
  st.seed = t->settings.seed;
  st.frame = frame_index;
  st.frame_dev = nullptr;
  st.max_depth = t->settings.max_depth;
  st.enable_nee = t->settings.enable_nee;
  st.enable_mis = t->settings.enable_mis;
  st.rr_start = t->settings.rr_start_depth;
  st.rr_min = t->settings.rr_min_survival;
  st.rr_max = t->settings.rr_max_survival;
  return st;
}

// Splits the work into passes of at most `capacity` paths (tiles in work order; inside a pass
// the tiles are grouped by scene, keeping their order within a scene), numbers each pass's paths,
// builds the film jobs and uploads tiles and jobs.
int plan_passes(pt_tracer* t, std::vector<ptd::DevTile>& work, std::vector<Pass>& passes,
                std::vector<ptd::FilmJob>& jobs) {
  const uint32_t cap = std::max<uint32_t>(t->capacity, 1024u);
  passes.clear();
  jobs.clear();
  size_t next = 0;
  while (next < work.size()) {
    Pass ps;
    ps.first = next;
    uint64_t paths = 0;
    size_t end = next;
    while (end < work.size()) {
      const uint64_t n = static_cast<uint64_t>(work[end].w) * work[end].h;
      if (end > next && paths + n > cap) break;
      paths += n;
      ++end;
    }
    std::stable_sort(work.begin() + static_cast<std::ptrdiff_t>(next), work.begin() + static_cast<std::ptrdiff_t>(end),
                     [](const ptd::DevTile& a, const ptd::DevTile& b) { return a.scene < b.scene; });
    // path numbering: runs of tiles with the same scene and rectangle interleave (DevTile::group)
    uint32_t pb = 0;
    for (size_t i = next; i < end;) {
      size_t j = i + 1;
      while (j < end && j - i < 0xFFFFu && work[j].scene == work[i].scene && work[j].x == work[i].x &&
             work[j].y == work[i].y && work[j].w == work[i].w && work[j].h == work[i].h)
        ++j;
      const uint32_t G = static_cast<uint32_t>(j - i);
      for (uint32_t k = 0; k < G; ++k) {
        work[i + k].path_begin = pb;
        work[i + k].group = (G << 16) | k;
      }
      pb += work[i].w * work[i].h * G;
      i = j;
    }
    ps.count = end - next;
    ps.paths = pb;
    ps.job_first = jobs.size();
    uint32_t waves = 0;
    for (size_t i = next; i < end;) {
      size_t j = i;
      ptd::FilmJob J{};
      J.scene = work[i].scene;
      uint32_t x0 = work[i].x, y0 = work[i].y, x1 = work[i].x + work[i].w, y1 = work[i].y + work[i].h;
      while (j < end && work[j].scene == J.scene) {
        x0 = std::min(x0, work[j].x);
        y0 = std::min(y0, work[j].y);
        x1 = std::max(x1, work[j].x + work[j].w);
        y1 = std::max(y1, work[j].y + work[j].h);
        ++j;
      }
      J.x0 = x0;
      J.y0 = y0;
      J.w = x1 - x0;
      J.h = y1 - y0;
      J.first_tile = static_cast<uint32_t>(i - next);
      J.tile_count = static_cast<uint32_t>(j - i);
      J.wave_begin = waves;
      waves += (J.w * J.h + 63u) / 64u;
      J.group = (work[i].group >> 16) == J.tile_count ? J.tile_count : 1u;
      jobs.push_back(J);
      i = j;
    }
    ps.job_count = jobs.size() - ps.job_first;
    ps.job_waves = waves;
    passes.push_back(ps);
    next = end;
  }
  const size_t tile_bytes = work.size() * sizeof(ptd::DevTile), job_bytes = jobs.size() * sizeof(ptd::FilmJob);
  std::vector<uint8_t> blob(tile_bytes + job_bytes);
  if (tile_bytes) std::memcpy(blob.data(), work.data(), tile_bytes);
  if (job_bytes) std::memcpy(blob.data() + tile_bytes, jobs.data(), job_bytes);
  if (blob.size() > t->d_plan.n) {
    PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");  // the old table may still be read
    PT_HIP(t, t->d_plan.ensure(blob.size() + blob.size() / 2), "allocate tile table");
  }
  return stage_upload(t, t->d_plan.ptr, blob.data(), blob.size());
}

// The bounce loop of one wavefront pass (extend -> shade -> shadow -> advance per depth) of `paths`
// paths. No queue of the pass holds more than `paths` entries, so no launch needs more than one
// workgroup per 256 of them (a wave takes 64 entries at a time): small passes launch small grids.
void bounce_loop(pt_tracer* t, const ptd::DevPaths& P, const ptd::DevSettings& st, const ptd::DevAdjoint* adj,
                 uint32_t paths) {
  const uint32_t need = std::max<uint32_t>(1u, (paths + 255u) / 256u);
  const auto grid = [need](uint32_t full) { return std::min(full, need); };
  int buf = 0;
  for (uint32_t depth = 0; depth < st.max_depth; ++depth) {
    timed(t, PT_KERNEL_EXTEND, 0, [&] { ptd::launch_extend(t->stream, P, t->dscene, buf, grid(t->grids.extend)); });
    timed(t, PT_KERNEL_SHADE, 0, [&] {
      if (adj) ptd::launch_shade_adjoint(t->stream, P, *adj, t->dscene, st, buf, depth, grid(t->grids.shade_adj));
      else ptd::launch_shade(t->stream, P, t->dscene, st, buf, depth, grid(t->grids.shade));
    });
    if (st.enable_nee && depth + 1u < st.max_depth)
      timed(t, PT_KERNEL_SHADOW, 0, [&] {
        if (adj) ptd::launch_shadow_adjoint(t->stream, P, *adj, t->dscene, buf, grid(t->grids.shadow_adj));
        else ptd::launch_shadow(t->stream, P, t->dscene, buf, grid(t->grids.shadow));
      });
    ptd::launch_advance(t->stream, P);
    buf ^= 1;
  }
}

// Allocates what the passes need (outside any graph capture).
int prepare_passes(pt_tracer* t, const std::vector<Pass>& passes, ptd::DevAdjoint* adj) {
  const uint32_t cap = std::max<uint32_t>(t->capacity, 1024u);
  uint32_t most = 0;
  for (const Pass& p : passes) most = std::max(most, p.paths);
  int rc = ensure_paths(t, std::max(most, std::min(cap, 1u << 20)));
  if (rc != PT_OK) return rc;
  if (adj) {
    rc = ensure_adjoint(t, t->path_capacity_alloc, *adj);
    if (rc != PT_OK) return rc;
  }
  return PT_OK;
}

// Launches every pass of a planned render (prepare_passes done).
// adj == nullptr: forward passes whose radiance is accumulated into the films in tile order.
// adj != nullptr: per pass, a forward replay (radiance kept per path) then the adjoint replay;
//                 the film and the render counters are left untouched. The forward replay is
//                 skipped when the previous call was the single-pass forward render of the same
//                 tiles, frame and scene (the autograd forward-then-backward pattern).
int launch_passes(pt_tracer* t, const std::vector<ptd::DevTile>& work, const std::vector<Pass>& passes,
                  const ptd::DevSettings& st, ptd::DevAdjoint* adj) {
  const ptd::DevFilm film{t->d_film_sum.ptr, t->d_film_count.ptr};
  const ptd::DevTile* dtiles = reinterpret_cast<const ptd::DevTile*>(t->d_plan.ptr);
  const ptd::FilmJob* djobs = reinterpret_cast<const ptd::FilmJob*>(t->d_plan.ptr + work.size() * sizeof(ptd::DevTile));
  for (const Pass& ps : passes) {
    const bool single_pass = passes.size() == 1;
    const bool reuse = adj && single_pass && t->fwd_rad_valid && st.frame == t->fwd_frame &&
                       t->scene_version == t->fwd_scene_version && same_work(work, t->fwd_work);
    ptd::DevPaths P = t->paths;
    if (adj) P.counters = t->d_query_counters.ptr;
    const ptd::DevTile* tl = dtiles + ps.first;
    const uint32_t nt = static_cast<uint32_t>(ps.count);
    if (!reuse) {
      ptd::launch_init_queue(t->stream, P, ps.paths);
      timed(t, PT_KERNEL_RAYGEN, ps.paths, [&] { ptd::launch_raygen(t->stream, P, t->dscene, st, tl, nt, ps.paths); });
      bounce_loop(t, P, st, nullptr, ps.paths);
    }
    if (adj) {
      // The adjoint replay leaves the path slots' radiance alone; a forward replay above has replaced it
      // with this pass's, which belongs to no recorded forward render.
      if (!reuse) t->fwd_rad_valid = false;
      ptd::launch_init_queue(t->stream, P, ps.paths);
      timed(t, PT_KERNEL_RAYGEN, ps.paths,
            [&] { ptd::launch_raygen_adjoint(t->stream, P, *adj, t->dscene, st, tl, nt, ps.paths); });
      bounce_loop(t, P, st, adj, ps.paths);
    } else {
      timed(t, PT_KERNEL_FILM, ps.paths, [&] {
        ptd::launch_film(t->stream, P, t->dscene, film, tl, djobs + ps.job_first, static_cast<uint32_t>(ps.job_count),
                         ps.job_waves);
      });
      t->fwd_rad_valid = single_pass;
      if (single_pass) {
        t->fwd_work = work;
        t->fwd_frame = st.frame;
        t->fwd_scene_version = t->scene_version;
      }
    }
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(t, le, "kernel launch");
  }
  return PT_OK;
}

int run_passes(pt_tracer* t, std::vector<ptd::DevTile>& work, const ptd::DevSettings& st, ptd::DevAdjoint* adj) {
  std::vector<Pass> passes;
  std::vector<ptd::FilmJob> jobs;
  int rc = plan_passes(t, work, passes, jobs);
  if (rc != PT_OK) return rc;
  rc = prepare_passes(t, passes, adj);
  if (rc != PT_OK) return rc;
  return launch_passes(t, work, passes, st, adj);
}

int init_tracer(pt_tracer* t, int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return PT_UNSUPPORTED;
  if (device < 0 || device >= count) return PT_INVALID_ARGUMENT;
  if (hipSetDevice(device) != hipSuccess) return PT_INTERNAL_ERROR;
  t->device = device;
  if (hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess) return PT_INTERNAL_ERROR;
  if (hipEventCreateWithFlags(&t->stage_done, hipEventDisableTiming) != hipSuccess) return PT_INTERNAL_ERROR;
  if (t->d_counters.ensure(ptd::kCounterWords) != hipSuccess || t->d_query_counters.ensure(ptd::kCounterWords) != hipSuccess ||
      t->d_queue.ensure(ptd::kQueueWords) != hipSuccess || t->d_error.ensure(1) != hipSuccess ||
      t->d_frame.ensure(1) != hipSuccess)
    return PT_ALLOC_FAILED;
  (void)hipMemsetAsync(t->d_counters.ptr, 0, sizeof(unsigned long long) * ptd::kCounterWords, t->stream);
  (void)hipMemsetAsync(t->d_queue.ptr, 0, sizeof(uint32_t) * ptd::kQueueWords, t->stream);
  (void)hipMemsetAsync(t->d_error.ptr, 0, sizeof(uint32_t), t->stream);
  t->paths.queue = t->d_queue.ptr;
  t->paths.counters = t->d_counters.ptr;
  t->paths.error = t->d_error.ptr;
  t->paths.refill_min = t->refill_min;
  t->grids = t->grids_max = ptd::resident_grids(device);
  // the allocation-free ray queries' static scratch on this device (looked up once per device)
  t->noalloc_status = ptd::noalloc_setup(device);
  // error messages of the allocation-free entry point fit without growing the string
  t->last_error.reserve(256);
  return PT_OK;
}

void destroy_tracer(pt_tracer* t) {
  (void)hipSetDevice(t->device);
  if (t->stream) (void)hipStreamSynchronize(t->stream);
  collect_profile(t);
  for (hipEvent_t e : t->event_pool) (void)hipEventDestroy(e);
  free_scene_device(t);
  free_paths(t);
  t->d_film_sum.release();
  t->d_film_count.release();
  t->d_resolve.release();
  t->d_queue.release();
  t->d_error.release();
  t->d_counters.release();
  t->d_query_counters.release();
  t->d_stack_ref.release();
  t->d_stack_t.release();
  t->d_plan.release();
  t->d_frame.release();
  if (t->copy_stream) (void)hipStreamSynchronize(t->copy_stream);
  t->d_readback.release();
  if (t->rb_host) (void)hipHostFree(t->rb_host);
  if (t->rb_resolved) (void)hipEventDestroy(t->rb_resolved);
  if (t->rb_copied) (void)hipEventDestroy(t->rb_copied);
  if (t->copy_stream) (void)hipStreamDestroy(t->copy_stream);
  if (t->stage) (void)hipHostFree(t->stage);
  if (t->stage_done) (void)hipEventDestroy(t->stage_done);
  if (t->stream) (void)hipStreamDestroy(t->stream);
}

int read_film_range(pt_tracer* t, size_t base, size_t pixels, float* rgb_sum, uint32_t* sample_count) {
  std::vector<float4> sum(rgb_sum ? pixels : 0);
  if (rgb_sum && pixels)
    PT_HIP(t, hipMemcpyAsync(sum.data(), t->d_film_sum.ptr + base, pixels * sizeof(float4), hipMemcpyDeviceToHost,
                             t->stream),
           "read film");
  if (sample_count && pixels)
    PT_HIP(t, hipMemcpyAsync(sample_count, t->d_film_count.ptr + base, pixels * sizeof(uint32_t),
                             hipMemcpyDeviceToHost, t->stream),
           "read film");
  const int rc = sync_all(t);
  if (rc != PT_OK) return rc;
  if (rgb_sum)
    for (size_t i = 0; i < pixels; ++i) {
      rgb_sum[3 * i] = sum[i].x;
      rgb_sum[3 * i + 1] = sum[i].y;
      rgb_sum[3 * i + 2] = sum[i].z;
    }
  return PT_OK;
}

void to_counters(const unsigned long long* c, PtSampleCounters* out) {
  *out = PtSampleCounters{};
  out->samples = c[ptd::kSamples];
  out->rays = c[ptd::kRays];
  out->triangle_tests = c[ptd::kTriTests];
  out->triangle_hits = c[ptd::kTriHits];
  out->bvh_node_visits = c[ptd::kNodeVisits];
  out->bvh_leaf_visits = c[ptd::kLeafVisits];
  out->shadow_tests = c[ptd::kShadowTests];
  out->sdf_tests = c[ptd::kSdfTests];
  out->sdf_steps = c[ptd::kSdfSteps];
  out->sdf_hits = c[ptd::kSdfHits];
  out->sdf_misses = c[ptd::kSdfTests] - c[ptd::kSdfHits];  // every test ends in a hit or a miss
}

// ColorTemperatureToRgb (FilmBuffer.cpp:30-49): the Tanner Helland fit, with the host's libm as in the
// reference.
void color_temperature_rgb(float kelvin, float out[3]) {
  const float t = std::clamp(kelvin, 1000.0f, 40000.0f) / 100.0f;
  float r = 1.0f, g = 1.0f, b = 1.0f;
  if (t <= 66.0f) {
    g = std::clamp(0.39008158f * std::log(std::max(1.0f, t)) - 0.63184144f, 0.0f, 1.0f);
    b = t <= 19.0f ? 0.0f : std::clamp(0.5432068f * std::log(std::max(1.0f, t - 10.0f)) - 1.1962541f, 0.0f, 1.0f);
  } else {
    r = std::clamp(1.2929362f * std::pow(t - 60.0f, -0.13320476f), 0.0f, 1.0f);
    g = std::clamp(1.1298909f * std::pow(t - 60.0f, -0.075514846f), 0.0f, 1.0f);
  }
  out[0] = r;
  out[1] = g;
  out[2] = b;
}

// FilmResolveSettings -> the kernel's record: WhiteBalanceScale (FilmBuffer.cpp:216-224) and the
// inverse gamma of ApplyFilmResolve (FilmBuffer.cpp:276).
int make_resolve(pt_tracer* t, const PtFilmResolve* in, ptd::DevResolve& rs) {
  const PtFilmResolve def{1.0f, 6500.0f, 0u, 0u, 2.2f, 1u};
  const PtFilmResolve& s = in ? *in : def;
  if (s.tone_map > 3u) return fail(t, PT_INVALID_ARGUMENT, "resolve_ldr: tone_map must be 0..3");
  if (s.output_transform > 1u) return fail(t, PT_INVALID_ARGUMENT, "resolve_ldr: output_transform must be 0 or 1");
  float d65[3], target[3];
  color_temperature_rgb(6500.0f, d65);
  color_temperature_rgb(s.white_balance_kelvin, target);
  rs.exposure = s.exposure;
  for (int k = 0; k < 3; ++k) rs.wb[k] = d65[k] / std::max(1.0e-3f, target[k]);
  rs.inv_gamma = 1.0f / std::max(0.01f, s.gamma);
  rs.tone_map = s.tone_map;
  rs.gamma_out = s.output_transform == 0u ? 1u : 0u;
  rs.clamp_out = s.clamp_output ? 1u : 0u;
  return PT_OK;
}

// Resolves film pixels [base, base + pixels) on the device and copies the image to out (host or device
// memory): 3 floats per pixel (rs == nullptr) or 4 bytes.
int resolve_range(pt_tracer* t, size_t base, size_t pixels, const ptd::DevResolve* rs, void* out) {
  if (pixels == 0) return PT_OK;
  if (pixels > 0xFFFFFFFFull) return fail(t, PT_INVALID_ARGUMENT, "resolve: image too large");
  const size_t words = rs ? pixels : 3u * pixels;
  PT_HIP(t, t->d_resolve.ensure(words), "allocate resolve buffer");
  const ptd::DevFilm film{t->d_film_sum.ptr + base, t->d_film_count.ptr + base};
  ptd::launch_resolve(t->stream, film, static_cast<uint32_t>(pixels), rs, t->d_resolve.ptr);
  PT_HIP(t, hipGetLastError(), "launch resolve");
  PT_HIP(t, hipMemcpyAsync(out, t->d_resolve.ptr, words * sizeof(uint32_t), hipMemcpyDefault, t->stream),
         "copy resolved image");
  return sync_all(t);
}

// The film range of one batch scene: its first pixel in the pooled film and its pixel count.
int batch_film_range(pt_batch* b, uint32_t scene, const char* what, size_t& base, size_t& pixels) {
  pt_tracer* t = &b->t;
  if (!t->accel_valid) return fail(t, PT_NOT_READY, std::string(what) + " requires pt_batch_build");
  if (scene >= t->scenes.size()) return fail(t, PT_INVALID_ARGUMENT, std::string(what) + ": no such scene");
  base = 0;
  for (uint32_t i = 0; i < scene; ++i) base += static_cast<size_t>(t->scenes[i].width) * t->scenes[i].height;
  pixels = static_cast<size_t>(t->scenes[scene].width) * t->scenes[scene].height;
  return PT_OK;
}

// IPathTracer::update_camera_state for scene slot `idx`: the slot's camera and, once the pool is on the
// device, its scene record's camera (the geometry and BVH stay).
int update_slot_camera(pt_tracer* t, uint32_t idx, const PtCamera& cam) {
  SceneSlot& s = t->scenes[idx];
  s.view.camera = cam;
  if (!t->accel_valid) return PT_OK;
  const ptd::DevCamera dc = make_camera(cam, s.width, s.height, t->settings.camera_aperture_radius,
                                        t->settings.camera_focus_distance);
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  PT_HIP(t, hipMemcpy(&t->d_recs.ptr[idx].cam, &dc, sizeof(dc), hipMemcpyHostToDevice), "upload camera");
  return PT_OK;
}

// resolve_instance_index (CpuBvhAccelerator.cpp:507-525): the update's instance, by index or entity id.
bool resolve_instance(const SceneSlot& s, const PtInstanceTransformUpdate& u, uint32_t& out) {
  uint32_t i = u.instance_index;
  if (i >= s.instances.size() && u.entity_id != 0u)
    for (size_t k = 0; k < s.instances.size(); ++k)
      if (s.instances[k].entity_id == u.entity_id) {
        i = static_cast<uint32_t>(k);
        break;
      }
  if (i >= s.instances.size()) return false;
  out = i;
  return true;
}

// plan_instance_transform_update (CpuBvhAccelerator.cpp:265-306).
PtInstanceTransformResult plan_refit(const pt_tracer* t, const PtInstanceTransformUpdate* ups, uint32_t n,
                                     const char*& message) {
  PtInstanceTransformResult r{PT_XFORM_FAILED, n, 0u, 0u, 0.0};
  if (!t->accel_valid || t->batch || t->scenes.size() != 1u || t->scenes[0].info.primitive_count == 0u) {
    message = "the BVH is not built";
    return r;
  }
  const SceneSlot& s = t->scenes[0];
  for (uint32_t k = 0; k < n; ++k) {
    uint32_t i = 0;
    if (!resolve_instance(s, ups[k], i)) {
      message = "transform update references an unknown instance";
      return r;
    }
    const PtInstance& inst = s.instances[i];
    if ((inst.flags & 1u) == 0u || inst.local_vertex_count == 0u || inst.local_index_count == 0u ||
        i + 1u >= s.inst_first.size()) {
      r.status = PT_XFORM_BLOCKED_NEEDS_FULL_STATIC_ACCEL_REBUILD;
      message = "transform update needs dynamic local geometry metadata";
      return r;
    }
    ++r.matched_count;
  }
  r.status = PT_XFORM_APPLIED_DYNAMIC_ACCEL_UPDATE;
  message = "the BVH can refit dynamic instance transforms";
  return r;
}

// One-time set-up of the refit data: padded boxes of every triangle in leaf order (from a flattening of
// the scene as it now stands, which is what the device holds), the local geometry, and the inner nodes
// grouped by height so that each launch only reads children already refit.
int prepare_refit(pt_tracer* t) {
  SceneSlot& s = t->scenes[0];
  ptb::HostVec<ptb::Prim> prims;
  ptb::HostVec<float> raw;
  std::string err;
  if (!ptb::flatten(s.view, prims, raw, err, nullptr, nullptr)) return fail(t, PT_INTERNAL_ERROR, "refit: " + err);
  const size_t n = prims.size();
  if (n != s.info.primitive_count) return fail(t, PT_INTERNAL_ERROR, "refit: primitive count changed");
  // each primitive's leaf position and the inner nodes' references: from the host layout of a tree
  // built on the host, else from the pool (one scene: its indices are the scene's own)
  const uint32_t inner = s.inner_count;
  std::vector<uint32_t> leaf_pos(s.layout.prim_tri.begin(), s.layout.prim_tri.end());
  std::vector<int32_t> refs(2u * inner);
  PT_HIP(t, t->d_leaf_pos.ensure(n), "allocate refit map");
  if (s.layout.prim_tri.size() != n) {
    ptg::launch_leaf_positions(t->stream, t->d_tris.ptr, static_cast<uint32_t>(n), t->d_leaf_pos.ptr);
    leaf_pos.resize(n);
    std::vector<float4> pool_nodes(4u * static_cast<size_t>(inner));
    PT_HIP(t, hipMemcpyAsync(leaf_pos.data(), t->d_leaf_pos.ptr, n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                             t->stream), "read refit map");
    PT_HIP(t, hipMemcpyAsync(pool_nodes.data(), t->d_nodes.ptr, pool_nodes.size() * sizeof(float4),
                             hipMemcpyDeviceToHost, t->stream), "read BVH nodes");
    PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
    for (uint32_t i = 0; i < inner; ++i) std::memcpy(&refs[2u * i], &pool_nodes[4u * i + 3u], 2 * sizeof(int32_t));
  } else {
    for (uint32_t i = 0; i < inner; ++i) std::memcpy(&refs[2u * i], &s.layout.nodes[16u * i + 12u], 2 * sizeof(int32_t));
  }
  std::vector<float> box(6 * n);
  auto mn3 = [](float a, float b, float c) { float m = a; if (b < m) m = b; if (c < m) m = c; return m; };
  auto mx3 = [](float a, float b, float c) { float m = a; if (m < b) m = b; if (m < c) m = c; return m; };
  for (size_t pid = 0; pid < n; ++pid) {
    const float* v = &raw[9 * pid];
    float* b = &box[6u * leaf_pos[pid]];
    for (int k = 0; k < 3; ++k) {
      b[k] = mn3(v[k], v[3 + k], v[6 + k]) - 1e-4f;
      b[3 + k] = mx3(v[k], v[3 + k], v[6 + k]) + 1e-4f;
    }
  }
  // node heights: children before parents (inner nodes are numbered in pre-order, so a reverse sweep
  // sees every child first)
  std::vector<uint32_t> height(inner, 0);
  uint32_t max_h = 0;
  for (uint32_t i = inner; i-- > 0;) {
    const int32_t* ref = &refs[2u * i];
    uint32_t h = 1;
    for (int c = 0; c < 2; ++c)
      if (ref[c] >= 0) h = std::max(h, height[static_cast<uint32_t>(ref[c])] + 1u);
    height[i] = h;
    max_h = std::max(max_h, h);
  }
  t->refit_level.assign(max_h + 1u, 0u);
  for (uint32_t i = 0; i < inner; ++i) ++t->refit_level[height[i]];
  uint32_t acc = 0;  // refit_level[h - 1] = first slot of height h after the prefix sum
  for (uint32_t h = 1; h <= max_h; ++h) {
    const uint32_t c = t->refit_level[h];
    t->refit_level[h - 1] = acc;
    acc += c;
  }
  t->refit_level[max_h] = acc;
  std::vector<uint32_t> order(inner), fill(t->refit_level.begin(), t->refit_level.end());
  for (uint32_t i = 0; i < inner; ++i) order[fill[height[i] - 1u]++] = i;
  PT_HIP(t, t->d_tri_box.ensure(box.size()), "allocate refit boxes");
  PT_HIP(t, hipMemcpy(t->d_tri_box.ptr, box.data(), box.size() * sizeof(float), hipMemcpyHostToDevice), "upload boxes");
  PT_HIP(t, hipMemcpy(t->d_leaf_pos.ptr, leaf_pos.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice),
         "upload refit map");
  PT_HIP(t, t->d_refit_order.ensure(order.size()), "allocate refit order");
  if (!order.empty())
    PT_HIP(t, hipMemcpy(t->d_refit_order.ptr, order.data(), order.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
           "upload refit order");
  PT_HIP(t, t->d_local_v.ensure(s.local_vertices.size()), "allocate local vertices");
  if (!s.local_vertices.empty())
    PT_HIP(t, hipMemcpy(t->d_local_v.ptr, s.local_vertices.data(), s.local_vertices.size() * sizeof(float),
                        hipMemcpyHostToDevice), "upload local vertices");
  PT_HIP(t, t->d_local_i.ensure(s.local_indices.size()), "allocate local indices");
  if (!s.local_indices.empty())
    PT_HIP(t, hipMemcpy(t->d_local_i.ptr, s.local_indices.data(), s.local_indices.size() * sizeof(uint32_t),
                        hipMemcpyHostToDevice), "upload local indices");
  PT_HIP(t, t->d_refit_flag.ensure(1), "allocate refit flag");
  t->refit_ready = true;
  return PT_OK;
}

// The SDF list of one slot (Unknown shapes dropped: intersect_scene never tests them).
int set_sdfs(pt_tracer* t, SceneSlot& s, const PtSdfPrimitive* prims, uint32_t count) {
  if (count && !prims) return fail(t, PT_INVALID_ARGUMENT, "load_sdf_primitives: null primitives");
  s.sdfs.clear();
  for (uint32_t i = 0; i < count; ++i)
    if (prims[i].shape <= 5u) s.sdfs.push_back(prims[i]);
  return PT_OK;
}

}  // namespace

extern "C" {

const char* pt_hip_version(void) { return "vkpathtracer_amd 0.2 (gfx950 wavefront, scene pools)"; }

uint32_t pt_hip_sizeof(int which) {
  switch (which) {
    case 0: return sizeof(PtMaterial);
    case 1: return sizeof(PtLight);
    case 2: return sizeof(PtInstance);
    case 3: return sizeof(PtCamera);
    case 4: return sizeof(PtSceneView);
    case 5: return sizeof(PtRenderSettings);
    case 6: return sizeof(PtSampleCounters);
    case 7: return sizeof(PtRenderTile);
    case 8: return sizeof(PtTuning);
    case 9: return sizeof(PtTextureDesc);
    case 10: return sizeof(PtFilmResolve);
    case 11: return sizeof(PtInstanceTransformUpdate);
    case 12: return sizeof(PtSdfPrimitive);
    case 13: return sizeof(PtRayQueryHit);
    case 14: return sizeof(PtSceneDeltaUpdate);
    case 15: return sizeof(PtMaterialUpdate);
    case 16: return sizeof(PtLightUpdate);
    case 17: return sizeof(PtInstanceMaterialUpdate);
    case 18: return sizeof(PtBuildInfo);
    default: return 0;
  }
}

int pt_hip_create(int device, pt_tracer** out) {
  if (!out) return PT_INVALID_ARGUMENT;
  *out = nullptr;
  pt_tracer* t = new (std::nothrow) pt_tracer();
  if (!t) return PT_ALLOC_FAILED;
  const int rc = init_tracer(t, device);
  if (rc != PT_OK) {
    destroy_tracer(t);
    delete t;
    return rc;
  }
  *out = t;
  return PT_OK;
}

void pt_hip_destroy(pt_tracer* t) {
  if (!t) return;
  destroy_tracer(t);
  delete t;
}

const char* pt_hip_last_error(const pt_tracer* t) { return t ? t->last_error.c_str() : "null tracer"; }

int pt_hip_configure(pt_tracer* t, const PtRenderSettings* s) {
  if (!t || !s) return fail(t, PT_INVALID_ARGUMENT, "configure: null argument");
  if (t->batch) return fail(t, PT_INVALID_ARGUMENT, "configure: use the pt_batch entry points on a batch");
  (void)hipSetDevice(t->device);
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (!(s->rr_min_survival == s->rr_min_survival) || !(s->rr_max_survival == s->rr_max_survival))
    return fail(t, PT_INVALID_ARGUMENT, "configure: russian roulette bounds are NaN");
  t->settings = *s;
  t->settings.max_depth = std::max(1u, s->max_depth);  // MakePathTraceSettings, PathTracer.cpp:283
  t->settings.camera_aperture_radius = std::max(0.0f, s->camera_aperture_radius);
  t->settings.camera_focus_distance = std::max(0.0f, s->camera_focus_distance);
  const int rc = alloc_film(t, static_cast<size_t>(s->width) * s->height);
  if (rc != PT_OK) return rc;
  ptd::launch_zero_counters(t->stream, t->d_counters.ptr);
  t->configured = true;
  ++t->scene_version;
  t->scene_loaded = false;
  t->accel_valid = false;
  t->scenes.clear();
  t->current_sample = 0;
  t->accumulation_gen = 0;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_load_scene(pt_tracer* t, const PtSceneView* sc) {
  if (!t || !sc) return fail(t, PT_INVALID_ARGUMENT, "load_scene: null argument");
  if (!t->configured) return fail(t, PT_NOT_READY, "load_scene_snapshot called before configure");
  t->scenes.resize(1);
  SceneSlot& s = t->scenes[0];
  copy_scene(s, sc);
  s.width = t->settings.width;
  s.height = t->settings.height;
  t->scene_loaded = true;
  t->accel_valid = false;
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_load_textures(pt_tracer* t, const float* texcoords, uint32_t texcoord_count, const float* texels,
                         uint64_t texel_count, const PtTextureDesc* textures, uint32_t texture_count) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (t->batch) return fail(t, PT_INVALID_ARGUMENT, "load_textures: use pt_batch_load_textures on a batch");
  if (!t->scene_loaded) return fail(t, PT_NOT_READY, "load_textures called before load_scene_snapshot");
  const int rc = copy_textures(t, t->scenes[0], texcoords, texcoord_count, texels, texel_count, textures, texture_count);
  if (rc != PT_OK) return rc;
  t->accel_valid = false;
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_load_sdf_primitives(pt_tracer* t, const PtSdfPrimitive* prims, uint32_t count) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (!t->scene_loaded || t->scenes.size() != 1u)
    return fail(t, PT_NOT_READY, "load_sdf_primitives called before load_scene_snapshot");
  const int rc = set_sdfs(t, t->scenes[0], prims, count);
  if (rc != PT_OK) return rc;
  t->accel_valid = false;  // the next build uploads them
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_update_texels(pt_tracer* t, const float* texels, uint64_t texel_count) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "update_texels requires an established acceleration");
  if (texel_count != t->texel_count) return fail(t, PT_INVALID_ARGUMENT, "update_texels: texel count differs");
  if (texel_count && !texels) return fail(t, PT_INVALID_ARGUMENT, "update_texels: null texels");
  (void)hipSetDevice(t->device);
  if (texel_count)
    PT_HIP(t,
           hipMemcpyAsync(t->d_texels.ptr, texels, 3u * sizeof(float) * texel_count, hipMemcpyDefault, t->stream),
           "upload texels");
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_build_accel(pt_tracer* t, const PtBvhParams* params, PtBuildInfo* info) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (!t->scene_loaded) return fail(t, PT_NOT_READY, "build_or_update_acceleration called before load_scene_snapshot");
  (void)hipSetDevice(t->device);
  PtBvhParams p;
  int rc = check_bvh_params(t, params, p);
  if (rc != PT_OK) return rc;
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  t->accel_valid = false;
  SceneSlot& s = t->scenes[0];
  s.bvh_params = p;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<ptg::DeviceForest> forests;
  if (!build_scene(s, t->stream, std::max(1u, std::min(16u, std::thread::hardware_concurrency())), forests))
    return fail(t, PT_INVALID_ARGUMENT, s.error);
  rc = upload_pool(t, forests);
  if (rc != PT_OK) return rc;
  t->info.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (info) *info = t->info;
  t->accel_valid = true;
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_reset_accumulation(pt_tracer* t) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (!t->configured) return fail(t, PT_NOT_READY, "reset_accumulation called before configure");
  (void)hipSetDevice(t->device);
  const uint32_t pixels = t->batch ? t->film_pixels : t->settings.width * t->settings.height;
  ptd::launch_clear_film(t->stream, ptd::DevFilm{t->d_film_sum.ptr, t->d_film_count.ptr}, pixels);
  ptd::launch_zero_counters(t->stream, t->d_counters.ptr);
  t->current_sample = 0;
  ++t->accumulation_gen;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_render_tiles(pt_tracer* t, const PtRenderTile* tiles, uint32_t count, uint64_t frame_index) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (t->batch) return fail(t, PT_INVALID_ARGUMENT, "render_tiles: use pt_batch_render on a batch");
  if (!t->configured || !t->scene_loaded)
    return fail(t, PT_NOT_READY, "render_tile requires a configured tracer with a scene");
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "render_tile requires build_or_update_acceleration");
  if (count && !tiles) return fail(t, PT_INVALID_ARGUMENT, "render_tiles: null tiles");
  (void)hipSetDevice(t->device);
  std::vector<ptd::DevTile> work;
  int rc = build_work(t, 0u, tiles, count, work);
  if (rc != PT_OK) return rc;
  rc = run_passes(t, work, dev_settings(t, frame_index), nullptr);
  if (rc != PT_OK) return rc;
  for (uint32_t i = 0; i < count; ++i) t->current_sample = std::max(t->current_sample, tiles[i].sample_index + 1u);
  if (t->profiling) {
    const int rc = sync_all(t);
    if (rc != PT_OK) return rc;
  }
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_render_tile(pt_tracer* t, const PtRenderTile* tile, uint64_t frame_index) {
  if (!tile) return fail(t, PT_INVALID_ARGUMENT, "render_tile: null tile");
  return pt_hip_render_tiles(t, tile, 1u, frame_index);
}

int pt_hip_synchronize(pt_tracer* t) {
  if (!t) return PT_INVALID_ARGUMENT;
  (void)hipSetDevice(t->device);
  return sync_all(t);
}

int pt_hip_read_film(pt_tracer* t, float* rgb_sum, uint32_t* sample_count) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (!t->configured) return fail(t, PT_NOT_READY, "read_film called before configure");
  (void)hipSetDevice(t->device);
  return read_film_range(t, 0, static_cast<size_t>(t->settings.width) * t->settings.height, rgb_sum, sample_count);
}

int pt_hip_replace_film(pt_tracer* t, const float* rgb_sum, const uint32_t* sample_count) {
  if (!t || !rgb_sum || !sample_count) return fail(t, PT_INVALID_ARGUMENT, "replace_film_history: null argument");
  if (!t->configured) return fail(t, PT_NOT_READY, "replace_film_history called before configure");
  (void)hipSetDevice(t->device);
  const size_t pixels = static_cast<size_t>(t->settings.width) * t->settings.height;
  std::vector<float4> sum(pixels);
  for (size_t i = 0; i < pixels; ++i)
    sum[i] = make_float4(rgb_sum[3 * i], rgb_sum[3 * i + 1], rgb_sum[3 * i + 2], 0.0f);
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (pixels) {
    PT_HIP(t, hipMemcpyAsync(t->d_film_sum.ptr, sum.data(), pixels * sizeof(float4), hipMemcpyHostToDevice, t->stream),
           "upload film");
    PT_HIP(t, hipMemcpyAsync(t->d_film_count.ptr, sample_count, pixels * sizeof(uint32_t), hipMemcpyHostToDevice,
                             t->stream),
           "upload film");
  }
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_resolve_hdr(pt_tracer* t, float* rgb) {
  if (!t || !rgb) return fail(t, PT_INVALID_ARGUMENT, "resolve_hdr: null argument");
  if (!t->configured) return fail(t, PT_NOT_READY, "resolve_hdr called before configure");
  (void)hipSetDevice(t->device);
  return resolve_range(t, 0, static_cast<size_t>(t->settings.width) * t->settings.height, nullptr, rgb);
}

int pt_hip_resolve_ldr(pt_tracer* t, const PtFilmResolve* settings, uint8_t* rgba8) {
  if (!t || !rgba8) return fail(t, PT_INVALID_ARGUMENT, "resolve_ldr: null argument");
  if (!t->configured) return fail(t, PT_NOT_READY, "resolve_ldr called before configure");
  ptd::DevResolve rs{};
  const int rc = make_resolve(t, settings, rs);
  if (rc != PT_OK) return rc;
  (void)hipSetDevice(t->device);
  return resolve_range(t, 0, static_cast<size_t>(t->settings.width) * t->settings.height, &rs, rgba8);
}

int pt_hip_read_counters(pt_tracer* t, PtSampleCounters* out) {
  if (!t || !out) return fail(t, PT_INVALID_ARGUMENT, "read_counters: null argument");
  (void)hipSetDevice(t->device);
  unsigned long long c[ptd::kCounterSlots];
  const int rc = read_counter_block(t, t->d_counters.ptr, c);
  if (rc != PT_OK) return rc;
  to_counters(c, out);
  return PT_OK;
}

int pt_hip_read_kernel_stats(pt_tracer* t, PtKernelStats* out) {
  if (!t || !out) return fail(t, PT_INVALID_ARGUMENT, "read_kernel_stats: null argument");
  (void)hipSetDevice(t->device);
  unsigned long long c[ptd::kCounterSlots];
  const int rc = read_counter_block(t, t->d_counters.ptr, c);
  if (rc != PT_OK) return rc;
  *out = t->kstats;
  out->inner_node_fetches = c[ptd::kInnerFetches];
  out->shadow_node_visits = c[ptd::kShadowNodeVisits];
  out->shadow_triangle_tests = c[ptd::kShadowTriTests];
  out->shadow_inner_node_fetches = c[ptd::kShadowInnerFetches];
  out->shadow_leaf_visits = c[ptd::kShadowLeafVisits];
  out->shadow_rays = c[ptd::kShadowRays];
  out->items[PT_KERNEL_EXTEND] = c[ptd::kRays];
  out->items[PT_KERNEL_SHADOW] = c[ptd::kShadowRays];
  out->items[PT_KERNEL_SHADE] = c[ptd::kRays];
  return PT_OK;
}

int pt_hip_reset_kernel_stats(pt_tracer* t) {
  if (!t) return PT_INVALID_ARGUMENT;
  const int rc = pt_hip_synchronize(t);
  t->kstats = PtKernelStats{};
  return rc;
}

int pt_hip_set_profiling(pt_tracer* t, int enabled) {
  if (!t) return PT_INVALID_ARGUMENT;
  const int rc = pt_hip_synchronize(t);
  t->profiling = enabled != 0;
  return rc;
}

int pt_hip_set_wavefront_capacity(pt_tracer* t, uint32_t paths) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (paths < 1024u) return fail(t, PT_INVALID_ARGUMENT, "wavefront capacity must be at least 1024 paths");
  const int rc = pt_hip_synchronize(t);
  if (rc != PT_OK) return rc;
  t->capacity = paths;
  drop_graph(t);
  if (t->path_capacity_alloc > paths) free_paths(t);
  return PT_OK;
}

int pt_hip_set_tuning(pt_tracer* t, const PtTuning* tu) {
  if (!t || !tu) return fail(t, PT_INVALID_ARGUMENT, "set_tuning: null argument");
  if (tu->wavefront_capacity && tu->wavefront_capacity < 1024u)
    return fail(t, PT_INVALID_ARGUMENT, "wavefront capacity must be at least 1024 paths");
  if (tu->refill_min > 64u) return fail(t, PT_INVALID_ARGUMENT, "refill_min must be in 1..64");
  (void)hipSetDevice(t->device);
  const int rc = pt_hip_synchronize(t);
  if (rc != PT_OK) return rc;
  drop_graph(t);
  if (tu->wavefront_capacity) {
    t->capacity = tu->wavefront_capacity;
    if (t->path_capacity_alloc > t->capacity) free_paths(t);
  }
  if (tu->refill_min) t->refill_min = t->paths.refill_min = tu->refill_min;
  auto pick = [](uint32_t want, uint32_t max) { return want ? std::min(want, max) : max; };
  t->grids.extend = pick(tu->traversal_blocks, t->grids_max.extend);
  t->grids.shadow = pick(tu->traversal_blocks, t->grids_max.shadow);
  t->grids.shadow_adj = pick(tu->traversal_blocks, t->grids_max.shadow_adj);
  t->grids.query = pick(tu->traversal_blocks, t->grids_max.query);
  t->grids.shade = pick(tu->shade_blocks, t->grids_max.shade);
  t->grids.shade_adj = pick(tu->shade_blocks, t->grids_max.shade_adj);
  return PT_OK;
}

int pt_hip_get_tuning(pt_tracer* t, PtTuning* tu) {
  if (!t || !tu) return fail(t, PT_INVALID_ARGUMENT, "get_tuning: null argument");
  tu->wavefront_capacity = t->capacity;
  tu->refill_min = t->refill_min;
  tu->traversal_blocks = t->grids.extend;
  tu->shade_blocks = t->grids.shade;
  return PT_OK;
}

int pt_hip_status(pt_tracer* t, PtTracerStatus* out) {
  if (!t || !out) return PT_INVALID_ARGUMENT;
  *out = PtTracerStatus{};
  out->lifecycle = !t->configured ? PT_LIFECYCLE_UNINITIALIZED
                   : !t->scene_loaded ? PT_LIFECYCLE_CONFIGURED
                   : !t->accel_valid ? PT_LIFECYCLE_SCENE_LOADED
                                     : PT_LIFECYCLE_READY;
  out->scene_loaded = t->scene_loaded;
  out->accel_valid = t->accel_valid;
  out->ready_to_render = t->configured && t->scene_loaded && t->accel_valid;
  out->current_sample = t->current_sample;
  out->accumulation_gen = t->accumulation_gen;
  if (t->configured) {
    PtSampleCounters c;
    if (pt_hip_read_counters(t, &c) == PT_OK) out->total_samples = c.samples;
  }
  return PT_OK;
}

int pt_hip_shutdown(pt_tracer* t) {
  if (!t) return PT_INVALID_ARGUMENT;
  (void)hipSetDevice(t->device);
  (void)hipStreamSynchronize(t->stream);
  free_scene_device(t);
  free_paths(t);
  t->d_film_sum.release();
  t->d_film_count.release();
  t->d_resolve.release();
  t->scenes.clear();
  t->configured = t->scene_loaded = t->accel_valid = false;
  t->current_sample = 0;
  t->accumulation_gen = 0;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_trace_rays(pt_tracer* t, int mode, const float* rays, const float* tmax, uint32_t count, float* out_t,
                      int32_t* out_prim, PtSampleCounters* counters) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (t->batch) return fail(t, PT_INVALID_ARGUMENT, "trace_rays: not available on a batch");
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "trace_rays requires build_or_update_acceleration");
  if (count && (!rays || !out_t || !out_prim)) return fail(t, PT_INVALID_ARGUMENT, "trace_rays: null buffer");
  if (mode != 0 && mode != 1) return fail(t, PT_INVALID_ARGUMENT, "trace_rays: mode must be 0 or 1");
  (void)hipSetDevice(t->device);
  ptd::DevPaths q = t->paths;
  q.counters = t->d_query_counters.ptr;
  ptd::launch_zero_counters(t->stream, q.counters);
  if (count)
    ptd::launch_trace_rays(t->stream, q, t->dscene, mode, rays, tmax, count, out_t, out_prim, t->grids.query);
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) return hip_fail(t, le, "kernel launch");
  unsigned long long c[ptd::kCounterSlots];
  const int rc = read_counter_block(t, q.counters, c);
  if (rc != PT_OK) return rc;
  if (counters) to_counters(c, counters);
  return PT_OK;
}

int pt_hip_trace_rays_noalloc(pt_tracer* t, int mode, const float* rays, const float* tmax, uint32_t count,
                              float* out_t, int32_t* out_prim, PtSampleCounters* counters) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (t->batch) return fail(t, PT_INVALID_ARGUMENT, "trace_rays_noalloc: not available on a batch");
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "trace_rays_noalloc requires build_or_update_acceleration");
  if (count && (!rays || !out_t || !out_prim)) return fail(t, PT_INVALID_ARGUMENT, "trace_rays_noalloc: null buffer");
  if (mode != 0 && mode != 1) return fail(t, PT_INVALID_ARGUMENT, "trace_rays_noalloc: mode must be 0 or 1");
  if (t->info.max_depth > PT_NOALLOC_MAX_TREE_DEPTH)
    return fail(t, PT_UNSUPPORTED, "trace_rays_noalloc: tree deeper than PT_NOALLOC_MAX_TREE_DEPTH levels");
  if (t->noalloc_status != hipSuccess) return hip_fail(t, t->noalloc_status, "trace_rays_noalloc: scratch lookup");
  (void)hipSetDevice(t->device);
  const ptd::NoAllocScratch sc = ptd::noalloc_scratch(t->device);
  if (!sc.counters) return fail(t, PT_INTERNAL_ERROR, "trace_rays_noalloc: no scratch on this device");
  ptd::launch_trace_rays_noalloc(t->stream, sc, t->paths.refill_min, t->dscene, mode, rays, tmax, count, out_t,
                                 out_prim, t->grids.query);
  const hipError_t le = hipGetLastError();
  if (le != hipSuccess) return hip_fail(t, le, "kernel launch");
  unsigned long long raw[ptd::kCounterWords];
  uint32_t error = 0;
  PT_HIP(t, hipMemcpyAsync(raw, sc.counters, sizeof(raw), hipMemcpyDeviceToHost, t->stream), "read counters");
  PT_HIP(t, hipMemcpyAsync(&error, sc.error, sizeof(error), hipMemcpyDeviceToHost, t->stream), "read error flag");
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (error) return fail(t, PT_INTERNAL_ERROR, "trace_rays_noalloc: traversal stack exceeded its spill area");
  if (counters) {
    unsigned long long c[ptd::kCounterSlots];
    for (int k = 0; k < ptd::kCounterSlots; ++k) {
      unsigned long long v = 0;
      for (int st = 0; st < ptd::kStripes; ++st) v += raw[static_cast<size_t>(st) * ptd::kCounterSlots + k];
      c[k] = v;
    }
    to_counters(c, counters);
  }
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_intersect(pt_tracer* t, int mode, const float* rays, const float* tmax, uint32_t count, PtRayQueryHit* hits,
                     PtSampleCounters* stats) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (t->batch) return fail(t, PT_INVALID_ARGUMENT, "intersect: not available on a batch");
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "intersect requires build_or_update_acceleration");
  if (count && (!rays || !hits)) return fail(t, PT_INVALID_ARGUMENT, "intersect: null buffer");
  if (mode != 0 && mode != 1) return fail(t, PT_INVALID_ARGUMENT, "intersect: mode must be 0 or 1");
  (void)hipSetDevice(t->device);
  if (!t->inst_mat_ready) {  // RayQueryHit::material_index: the instance material baked at build
    const ptb::HostVec<ptb::Prim>& prims = t->scenes[0].bvh.prims;
    std::vector<uint32_t> m(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) m[i] = prims[i].material;
    PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
    PT_HIP(t, t->d_inst_mat.ensure(m.size()), "allocate instance materials");
    if (!m.empty())
      PT_HIP(t, hipMemcpy(t->d_inst_mat.ptr, m.data(), m.size() * sizeof(uint32_t), hipMemcpyHostToDevice),
             "upload instance materials");
    t->inst_mat_ready = true;
  }
  const size_t ray_bytes = 24u * static_cast<size_t>(count), tmax_bytes = tmax ? 4u * static_cast<size_t>(count) : 0u;
  const size_t hit_off = (ray_bytes + tmax_bytes + 255u) & ~static_cast<size_t>(255u);
  const size_t total = hit_off + sizeof(PtRayQueryHit) * static_cast<size_t>(count);
  if (count) {
    // the previous call's copies read d_iq on this stream: the stream's order keeps them first
    PT_HIP(t, t->d_iq.ensure(total), "allocate ray query staging");
    PT_HIP(t, hipMemcpyAsync(t->d_iq.ptr, rays, ray_bytes, hipMemcpyDefault, t->stream), "stage rays");
    if (tmax)
      PT_HIP(t, hipMemcpyAsync(t->d_iq.ptr + ray_bytes, tmax, tmax_bytes, hipMemcpyDefault, t->stream), "stage limits");
  }
  ptd::DevPaths q = t->paths;
  q.counters = t->d_query_counters.ptr;
  ptd::launch_zero_counters(t->stream, q.counters);
  if (count) {
    const float* drays = reinterpret_cast<const float*>(t->d_iq.ptr);
    const float* dtmax = tmax ? reinterpret_cast<const float*>(t->d_iq.ptr + ray_bytes) : nullptr;
    PtRayQueryHit* dhits = reinterpret_cast<PtRayQueryHit*>(t->d_iq.ptr + hit_off);
    ptd::launch_intersect(t->stream, q, t->dscene, mode, drays, dtmax, count, dhits, t->d_inst_mat.ptr,
                          t->grids.query);
    const hipError_t le = hipGetLastError();
    if (le != hipSuccess) return hip_fail(t, le, "kernel launch");
    PT_HIP(t, hipMemcpyAsync(hits, dhits, sizeof(PtRayQueryHit) * static_cast<size_t>(count), hipMemcpyDefault,
                             t->stream),
           "copy ray query hits");
  }
  unsigned long long c[ptd::kCounterSlots];
  const int rc = read_counter_block(t, q.counters, c);
  if (rc != PT_OK) return rc;
  if (stats) {
    *stats = PtSampleCounters{};
    stats->triangle_tests = c[ptd::kTriTests];
    stats->triangle_hits = c[ptd::kTriHits];
    stats->bvh_node_visits = c[ptd::kNodeVisits];
    stats->bvh_leaf_visits = c[ptd::kLeafVisits];
  }
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_update_camera(pt_tracer* t, const PtCamera* camera) {
  if (!t || !camera) return fail(t, PT_INVALID_ARGUMENT, "update_camera: null argument");
  if (!t->configured || !t->scene_loaded || t->scenes.size() != 1u)
    return fail(t, PT_NOT_READY, "update_camera_state requires a configured tracer with a scene");
  (void)hipSetDevice(t->device);
  const int rc = update_slot_camera(t, 0u, *camera);
  if (rc != PT_OK) return rc;
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_update_render_settings(pt_tracer* t, const PtRenderSettings* s) {
  if (!t || !s) return fail(t, PT_INVALID_ARGUMENT, "update_render_settings: null argument");
  if (!t->configured || t->batch) return fail(t, PT_NOT_READY, "update_render_settings requires a configured tracer");
  if (s->width != t->settings.width || s->height != t->settings.height)
    return fail(t, PT_UNSUPPORTED, "update_render_settings: a new resolution needs configure");
  if (!(s->rr_min_survival == s->rr_min_survival) || !(s->rr_max_survival == s->rr_max_survival))
    return fail(t, PT_INVALID_ARGUMENT, "update_render_settings: russian roulette bounds are NaN");
  (void)hipSetDevice(t->device);
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  const bool lens = s->camera_aperture_radius != t->settings.camera_aperture_radius ||
                    s->camera_focus_distance != t->settings.camera_focus_distance;
  t->settings = *s;
  t->settings.max_depth = std::max(1u, s->max_depth);
  t->settings.camera_aperture_radius = std::max(0.0f, s->camera_aperture_radius);
  t->settings.camera_focus_distance = std::max(0.0f, s->camera_focus_distance);
  if (lens && t->scene_loaded) {
    const int rc = update_slot_camera(t, 0u, t->scenes[0].view.camera);
    if (rc != PT_OK) return rc;
  }
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_request_film_readback(pt_tracer* t, const PtFilmResolve* settings, PtFilmReadbackToken* token) {
  if (!t || !token) return fail(t, PT_INVALID_ARGUMENT, "request_film_readback: null argument");
  *token = PtFilmReadbackToken{};
  if (!t->configured) return fail(t, PT_NOT_READY, "request_film_readback called before configure");
  ptd::DevResolve rs{};
  int rc = make_resolve(t, settings, rs);
  if (rc != PT_OK) return rc;
  (void)hipSetDevice(t->device);
  const size_t pixels = static_cast<size_t>(t->settings.width) * t->settings.height;
  if (pixels == 0 || pixels > 0xFFFFFFFFull) return fail(t, PT_INVALID_ARGUMENT, "request_film_readback: image size");
  if (!t->copy_stream) {
    PT_HIP(t, hipStreamCreateWithFlags(&t->copy_stream, hipStreamNonBlocking), "create copy stream");
    PT_HIP(t, hipEventCreateWithFlags(&t->rb_resolved, hipEventDisableTiming), "create event");
    PT_HIP(t, hipEventCreateWithFlags(&t->rb_copied, hipEventDisableTiming), "create event");
  }
  if (t->rb_host_bytes < 4 * pixels) {
    PT_HIP(t, hipStreamSynchronize(t->copy_stream), "synchronize copy stream");
    if (t->rb_host) (void)hipHostFree(t->rb_host);
    t->rb_host = nullptr;
    t->rb_host_bytes = 0;
    PT_HIP(t, hipHostMalloc(reinterpret_cast<void**>(&t->rb_host), 4 * pixels, hipHostMallocDefault),
           "allocate pinned readback buffer");
    t->rb_host_bytes = 4 * pixels;
  }
  PT_HIP(t, t->d_readback.ensure(pixels), "allocate readback buffer");
  // the previous request's copy reads d_readback: the new resolve waits for it
  if (t->rb_inflight) PT_HIP(t, hipStreamWaitEvent(t->stream, t->rb_copied, 0), "order readback");
  const ptd::DevFilm film{t->d_film_sum.ptr, t->d_film_count.ptr};
  ptd::launch_resolve(t->stream, film, static_cast<uint32_t>(pixels), &rs, t->d_readback.ptr);
  PT_HIP(t, hipGetLastError(), "launch resolve");
  PT_HIP(t, hipEventRecord(t->rb_resolved, t->stream), "record event");
  PT_HIP(t, hipStreamWaitEvent(t->copy_stream, t->rb_resolved, 0), "order readback");
  PT_HIP(t, hipMemcpyAsync(t->rb_host, t->d_readback.ptr, 4 * pixels, hipMemcpyDeviceToHost, t->copy_stream),
         "copy readback");
  PT_HIP(t, hipEventRecord(t->rb_copied, t->copy_stream), "record event");
  t->rb_inflight = true;
  token->id = t->rb_next_id++;
  token->fence_value = token->id;
  token->width = t->settings.width;
  token->height = t->settings.height;
  t->rb_latest = *token;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_poll_film(pt_tracer* t, const PtFilmReadbackToken* token, uint8_t* rgba8, uint32_t* state) {
  if (!t || !token || !state) return fail(t, PT_INVALID_ARGUMENT, "poll_film: null argument");
  *state = PT_READBACK_INVALID;
  if (token->id == 0 || token->id != t->rb_latest.id || token->width != t->rb_latest.width ||
      token->height != t->rb_latest.height)
    return PT_OK;
  (void)hipSetDevice(t->device);
  const hipError_t q = hipEventQuery(t->rb_copied);
  if (q == hipErrorNotReady) {
    *state = PT_READBACK_PENDING;
    return PT_OK;
  }
  if (q != hipSuccess) {
    *state = PT_READBACK_FAILED;
    return fail(t, PT_INTERNAL_ERROR, std::string("poll_film: ") + hipGetErrorString(q));
  }
  t->rb_inflight = false;
  if (rgba8) std::memcpy(rgba8, t->rb_host, 4ull * token->width * token->height);
  *state = PT_READBACK_READY;
  return PT_OK;  // no wait on the main stream: rendering queued after the request keeps running
}

int pt_hip_plan_instance_transform_update(pt_tracer* t, const PtInstanceTransformUpdate* updates, uint32_t count,
                                          PtInstanceTransformResult* plan) {
  if (!t || !plan || (count && !updates)) return fail(t, PT_INVALID_ARGUMENT, "plan_instance_transform_update: null argument");
  const char* message = "";
  *plan = plan_refit(t, updates, count, message);
  if (plan->status != PT_XFORM_APPLIED_DYNAMIC_ACCEL_UPDATE) t->last_error = message;
  return PT_OK;
}

int pt_hip_apply_instance_transform_update(pt_tracer* t, const PtInstanceTransformUpdate* updates, uint32_t count,
                                           PtInstanceTransformResult* result) {
  if (!t || !result || (count && !updates))
    return fail(t, PT_INVALID_ARGUMENT, "apply_instance_transform_update: null argument");
  const char* message = "";
  *result = plan_refit(t, updates, count, message);
  if (result->status != PT_XFORM_APPLIED_DYNAMIC_ACCEL_UPDATE) {
    t->last_error = message;
    return PT_OK;
  }
  const auto t0 = std::chrono::steady_clock::now();
  (void)hipSetDevice(t->device);
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (!t->refit_ready) {
    const int rc = prepare_refit(t);
    if (rc != PT_OK) {
      result->status = PT_XFORM_FAILED;
      return rc;
    }
  }
  SceneSlot& s = t->scenes[0];
  // the instances touched, each once; refit_instance_primitives' checks (CpuBvhAccelerator.cpp:531-566)
  // run on all of them before anything changes
  std::vector<int32_t> slot_of(s.instances.size(), -1);
  std::vector<uint32_t> inst_of;
  for (uint32_t k = 0; k < count; ++k) {
    uint32_t i = 0;
    (void)resolve_instance(s, updates[k], i);
    if (slot_of[i] < 0) {
      slot_of[i] = static_cast<int32_t>(inst_of.size());
      inst_of.push_back(i);
    }
  }
  for (uint32_t i : inst_of) {
    const PtInstance& inst = s.instances[i];
    bool ok = s.inst_first[i + 1u] - s.inst_first[i] == inst.triangle_count;
    const uint64_t local_end = static_cast<uint64_t>(inst.local_first_vertex) + inst.local_vertex_count;
    for (uint32_t tri = 0; ok && tri < inst.triangle_count; ++tri) {
      const uint64_t base = static_cast<uint64_t>(inst.local_first_index) + 3u * tri;
      if (base + 2u >= s.local_indices.size()) {
        ok = false;
        break;
      }
      for (int c = 0; c < 3; ++c) {
        const uint64_t v = static_cast<uint64_t>(inst.local_first_vertex) + s.local_indices[base + c];
        if (v >= local_end || 3u * v + 2u >= s.local_vertices.size()) ok = false;
      }
    }
    if (!ok) {
      result->status = PT_XFORM_FAILED;
      t->last_error = "failed while refitting instance primitives";
      return PT_OK;
    }
  }
  // ApplyInstanceTransformUpdates on the slot's records, in update order (a later update of the same
  // instance wins, as sequential refits leave it); the next full build sees them too
  for (uint32_t k = 0; k < count; ++k) {
    uint32_t i = 0;
    (void)resolve_instance(s, updates[k], i);
    PtInstance& inst = s.instances[i];
    std::memcpy(inst.translation, updates[k].translation, sizeof(inst.translation));
    std::memcpy(inst.rotation, updates[k].rotation, sizeof(inst.rotation));
    std::memcpy(inst.scale, updates[k].scale, sizeof(inst.scale));
    inst.flags |= updates[k].flags;
    if (updates[k].transform_revision != 0u) inst.transform_revision = updates[k].transform_revision;
  }
  std::vector<ptd::RefitInst> recs(inst_of.size());
  uint32_t tri_total = 0;
  for (size_t r = 0; r < recs.size(); ++r) {
    const PtInstance& inst = s.instances[inst_of[r]];
    ptd::RefitInst& R = recs[r];
    R.prim_first = s.inst_first[inst_of[r]];
    R.tri_count = inst.triangle_count;
    R.local_first_vertex = inst.local_first_vertex;
    R.local_vertex_count = inst.local_vertex_count;
    R.local_first_index = inst.local_first_index;
    R.tri_begin = tri_total;
    std::memcpy(R.t, inst.translation, sizeof(R.t));
    std::memcpy(R.q, inst.rotation, sizeof(R.q));
    std::memcpy(R.s, inst.scale, sizeof(R.s));
    tri_total += inst.triangle_count;
  }
  PT_HIP(t, t->d_refit_insts.ensure(recs.size()), "allocate refit records");
  if (!recs.empty())
    PT_HIP(t, hipMemcpy(t->d_refit_insts.ptr, recs.data(), recs.size() * sizeof(ptd::RefitInst), hipMemcpyHostToDevice),
           "upload refit records");
  PT_HIP(t, hipMemsetAsync(t->d_refit_flag.ptr, 0, sizeof(uint32_t), t->stream), "clear refit flag");
  ptd::launch_refit_tris(t->stream, t->d_refit_insts.ptr, static_cast<uint32_t>(recs.size()), tri_total,
                         t->d_local_v.ptr, t->d_local_i.ptr, t->d_leaf_pos.ptr, t->d_tris.ptr, t->d_prims.ptr,
                         t->d_tri_box.ptr);
  for (size_t h = 0; h + 1 < t->refit_level.size(); ++h)
    ptd::launch_refit_nodes(t->stream, t->d_nodes.ptr, t->d_refit_order.ptr, t->refit_level[h],
                            t->refit_level[h + 1] - t->refit_level[h], t->d_tri_box.ptr, t->d_refit_flag.ptr);
  ptd::launch_refit_root(t->stream, t->d_nodes.ptr, s.root_ref, t->d_tri_box.ptr, t->d_recs.ptr);
  ptd::launch_hot_refresh(t->stream, t->d_nodes.ptr, t->d_hot_src.ptr, t->dscene.hot_base, t->dscene.hot_count);
  PT_HIP(t, hipGetLastError(), "launch refit");
  uint32_t nonfinite = 0;
  PT_HIP(t, hipMemcpyAsync(&nonfinite, t->d_refit_flag.ptr, sizeof(uint32_t), hipMemcpyDeviceToHost, t->stream),
         "read refit flag");
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (nonfinite) t->dscene.general_boxes = 1u;
  t->fwd_rad_valid = false;
  ++t->scene_version;
  result->applied_count = count;
  result->refit_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_update_scene_delta(pt_tracer* t, const PtMaterial* materials, uint32_t material_count,
                              const PtLight* lights, uint32_t light_count) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (!t->accel_valid || t->scenes.size() != 1u)
    return fail(t, PT_NOT_READY, "update_scene_delta requires an established acceleration");
  (void)hipSetDevice(t->device);
  SceneSlot& s = t->scenes[0];
  if (materials && material_count != s.materials.size())
    return fail(t, PT_INVALID_ARGUMENT, "update_scene_delta: material count differs from the loaded scene");
  if (lights && light_count != s.lights.size())
    return fail(t, PT_INVALID_ARGUMENT, "update_scene_delta: light count differs from the loaded scene");
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (materials) {
    s.materials.assign(materials, materials + material_count);
    const std::vector<PtMaterial> pm = pool_materials(s, 0u);
    PT_HIP(t, hipMemcpy(t->d_materials.ptr, pm.data(), sizeof(PtMaterial) * material_count, hipMemcpyHostToDevice),
           "upload materials");
    uint32_t any = 0, textured = 0;
    for (const PtMaterial& m : s.materials)
      if (transmissive(m)) any = 1;
    for (const PtMaterial& m : pm)
      if (m.base_color_texture_index != 0xFFFFFFFFu) textured = 1;
    t->dscene.textured = textured;
    t->dscene.has_transmissive = any;
    PT_HIP(t, hipMemcpy(&t->d_recs.ptr[0].has_transmissive, &any, sizeof(uint32_t), hipMemcpyHostToDevice),
           "upload scene record");
  }
  if (lights && light_count) {
    s.lights.assign(lights, lights + light_count);
    PT_HIP(t, hipMemcpy(t->d_lights.ptr, lights, sizeof(PtLight) * light_count, hipMemcpyHostToDevice),
           "upload lights");
  }
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_apply_scene_delta(pt_tracer* t, const PtSceneDeltaUpdate* u) {
  if (!t || !u) return fail(t, PT_INVALID_ARGUMENT, "update_scene_delta: null argument");
  if (!t->configured || !t->scene_loaded || t->scenes.size() != 1u)
    return fail(t, PT_NOT_READY, "update_scene_delta requires a configured tracer with a scene");
  if ((u->material_count && !u->materials) || (u->light_count && !u->lights) ||
      (u->instance_material_count && !u->instance_materials))
    return fail(t, PT_INVALID_ARGUMENT, "update_scene_delta: null list");
  SceneSlot& s = t->scenes[0];
  // ApplySceneDeltaUpdate (PathTracer.cpp:427-482): check every index before changing anything
  auto instance_of = [&](const PtInstanceMaterialUpdate& im) -> uint32_t {
    uint32_t idx = im.instance_index;
    if (idx >= s.instances.size() && im.entity_id != 0u)
      for (size_t k = 0; k < s.instances.size(); ++k)
        if (s.instances[k].entity_id == im.entity_id) {
          idx = static_cast<uint32_t>(k);
          break;
        }
    return idx;
  };
  for (uint32_t i = 0; i < u->material_count; ++i)
    if (u->materials[i].material_index >= s.materials.size())
      return fail(t, PT_INVALID_ARGUMENT, "scene delta did not apply: material index out of range");
  for (uint32_t i = 0; i < u->light_count; ++i)
    if (u->lights[i].light_index >= s.lights.size())
      return fail(t, PT_INVALID_ARGUMENT, "scene delta did not apply: light index out of range");
  for (uint32_t i = 0; i < u->instance_material_count; ++i)
    if (instance_of(u->instance_materials[i]) >= s.instances.size() ||
        u->instance_materials[i].material_index >= s.materials.size())
      return fail(t, PT_INVALID_ARGUMENT, "scene delta did not apply: instance or material index out of range");
  bool changed = false;
  for (uint32_t i = 0; i < u->material_count; ++i) {
    s.materials[u->materials[i].material_index] = u->materials[i].material;
    changed = true;
  }
  for (uint32_t i = 0; i < u->light_count; ++i) {
    s.lights[u->lights[i].light_index] = u->lights[i].light;
    changed = true;
  }
  for (uint32_t i = 0; i < u->instance_material_count; ++i) {
    PtInstance& inst = s.instances[instance_of(u->instance_materials[i])];
    if (inst.material_index != u->instance_materials[i].material_index) {
      inst.material_index = u->instance_materials[i].material_index;  // baked at the next build
      changed = true;
    }
  }
  if (u->environment_color_changed) {
    for (int k = 0; k < 3; ++k) s.view.environment_color[k] = u->environment_color[k];
    changed = true;
  }
  if (!changed) return fail(t, PT_INVALID_ARGUMENT, "scene delta did not apply");
  if (!t->accel_valid) {  // nothing on the device yet: the next build uploads the snapshot as it now is
    ++t->scene_version;
    t->last_error.clear();
    return PT_OK;
  }
  (void)hipSetDevice(t->device);
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (u->material_count) {
    const std::vector<PtMaterial> pm = pool_materials(s, 0u);
    PT_HIP(t, hipMemcpy(t->d_materials.ptr, pm.data(), sizeof(PtMaterial) * pm.size(), hipMemcpyHostToDevice),
           "upload materials");
    uint32_t any = 0, textured = 0;
    for (const PtMaterial& m : s.materials)
      if (transmissive(m)) any = 1;
    for (const PtMaterial& m : pm)
      if (m.base_color_texture_index != 0xFFFFFFFFu) textured = 1;
    t->dscene.textured = textured;
    t->dscene.has_transmissive = any;
    PT_HIP(t, hipMemcpy(&t->d_recs.ptr[0].has_transmissive, &any, sizeof(uint32_t), hipMemcpyHostToDevice),
           "upload scene record");
  }
  if (u->light_count)
    PT_HIP(t, hipMemcpy(t->d_lights.ptr, s.lights.data(), sizeof(PtLight) * s.lights.size(), hipMemcpyHostToDevice),
           "upload lights");
  if (u->environment_color_changed)
    PT_HIP(t, hipMemcpy(t->d_recs.ptr[0].env, s.view.environment_color, 3 * sizeof(float), hipMemcpyHostToDevice),
           "upload environment colour");
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

// Path-replay adjoint of `work` (one tracer's tiles or a batch frame). Gradient buffers are indexed
// by pool material, light and texel; film_grad by pool film pixel.
static int run_adjoint(pt_tracer* t, std::vector<ptd::DevTile>& work, uint64_t frame_index, const float* film_grad,
                float* grad_albedo, float* grad_emission, float* grad_light, float* grad_roughness,
                float* grad_texels) {
  if (!film_grad || !grad_albedo || !grad_emission || (!grad_light && t->dscene.light_count))
    return fail(t, PT_INVALID_ARGUMENT, "render_adjoint: null gradient buffer");
  if (t->settings.rr_start_depth > static_cast<uint32_t>(ptd::kAdjHist))
    return fail(t, PT_UNSUPPORTED, "render_adjoint supports russian roulette start depths up to 4");
  ptd::DevAdjoint A{};
  A.g_albedo = grad_albedo;
  A.g_emission = grad_emission;
  A.g_light = grad_light;
  A.g_rough = grad_roughness;
  A.g_tex = t->texel_count ? grad_texels : nullptr;
  A.film_grad = film_grad;
  A.nmat = t->dscene.material_count;
  A.nlight = t->dscene.light_count;
  const uint32_t slots = 7u * A.nmat + 3u * A.nlight;
  A.use_lds = slots <= static_cast<uint32_t>(ptd::kAdjLdsFloats) ? 1u : 0u;
  A.lds_stride = slots | 1u;  // odd: the copies of a slot sit in different banks
  A.lds_copies = 1;
  while (A.lds_copies < 16u && 2u * A.lds_copies * A.lds_stride <= static_cast<uint32_t>(ptd::kAdjLdsFloats))
    A.lds_copies *= 2u;
  if (A.use_lds) {
    PT_HIP(t, t->d_grad_partial.ensure(static_cast<size_t>(ptd::kGradRows) * slots), "allocate gradient rows");
    A.partial = t->d_grad_partial.ptr;
    A.partial_rows = ptd::kGradRows;
  }
  ptd::launch_grad_clear(t->stream, A);
  int rc = run_passes(t, work, dev_settings(t, frame_index), &A);
  if (rc != PT_OK) return rc;
  ptd::launch_grad_reduce(t->stream, A);
  if (t->profiling) {
    rc = sync_all(t);
    if (rc != PT_OK) return rc;
  }
  t->last_error.clear();
  return PT_OK;
}

int pt_hip_render_adjoint(pt_tracer* t, const PtRenderTile* tiles, uint32_t count, uint64_t frame_index,
                          const float* film_grad, float* grad_albedo, float* grad_emission, float* grad_light,
                          float* grad_roughness, float* grad_texels) {
  if (!t) return PT_INVALID_ARGUMENT;
  if (t->batch) return fail(t, PT_INVALID_ARGUMENT, "render_adjoint: use pt_batch_render_adjoint on a batch");
  if (!t->configured || !t->scene_loaded || !t->accel_valid)
    return fail(t, PT_NOT_READY, "render_adjoint requires a built acceleration structure");
  if (count && !tiles) return fail(t, PT_INVALID_ARGUMENT, "render_adjoint: null tiles");
  (void)hipSetDevice(t->device);
  std::vector<ptd::DevTile> work;
  const int rc = build_work(t, 0u, tiles, count, work);
  if (rc != PT_OK) return rc;
  return run_adjoint(t, work, frame_index, film_grad, grad_albedo, grad_emission, grad_light, grad_roughness,
                     grad_texels);
}

// ------------------------------------------------------------------ batches
int pt_batch_create(int device, const PtRenderSettings* settings, pt_batch** out) {
  if (!out || !settings) return PT_INVALID_ARGUMENT;
  *out = nullptr;
  if (!(settings->rr_min_survival == settings->rr_min_survival) ||
      !(settings->rr_max_survival == settings->rr_max_survival))
    return PT_INVALID_ARGUMENT;
  pt_batch* b = new (std::nothrow) pt_batch();
  if (!b) return PT_ALLOC_FAILED;
  pt_tracer* t = &b->t;
  const int rc = init_tracer(t, device);
  if (rc != PT_OK) {
    destroy_tracer(t);
    delete b;
    return rc;
  }
  t->batch = true;
  t->settings = *settings;
  t->settings.max_depth = std::max(1u, settings->max_depth);
  t->settings.camera_aperture_radius = std::max(0.0f, settings->camera_aperture_radius);
  t->settings.camera_focus_distance = std::max(0.0f, settings->camera_focus_distance);
  t->configured = true;
  *out = b;
  return PT_OK;
}

void pt_batch_destroy(pt_batch* b) {
  if (!b) return;
  destroy_tracer(&b->t);
  delete b;
}

const char* pt_batch_last_error(const pt_batch* b) { return b ? b->t.last_error.c_str() : "null batch"; }

int pt_batch_add_scene(pt_batch* b, const PtSceneView* scene, uint32_t width, uint32_t height, uint32_t spp,
                       const PtBvhParams* params, uint32_t* index) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (!scene) return fail(t, PT_INVALID_ARGUMENT, "add_scene: null scene");
  if (width == 0 || height == 0 || spp == 0) return fail(t, PT_INVALID_ARGUMENT, "add_scene: empty image or spp");
  PtBvhParams p;
  const int rc = check_bvh_params(t, params, p);
  if (rc != PT_OK) return rc;
  t->scenes.emplace_back();
  SceneSlot& s = t->scenes.back();
  copy_scene(s, scene);
  s.width = width;
  s.height = height;
  s.spp = spp;
  s.bvh_params = p;
  t->scene_loaded = true;
  t->accel_valid = false;
  if (index) *index = static_cast<uint32_t>(t->scenes.size() - 1u);
  t->last_error.clear();
  return PT_OK;
}

int pt_batch_add_view(pt_batch* b, uint32_t scene, const PtCamera* camera, uint32_t width, uint32_t height,
                      uint32_t spp, uint32_t* index) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (!camera) return fail(t, PT_INVALID_ARGUMENT, "add_view: null camera");
  if (scene >= t->scenes.size() || t->scenes[scene].view_of >= 0)
    return fail(t, PT_INVALID_ARGUMENT, "add_view: not a scene of the batch");
  if (width == 0 || height == 0 || spp == 0) return fail(t, PT_INVALID_ARGUMENT, "add_view: empty image or spp");
  t->scenes.emplace_back();
  SceneSlot& s = t->scenes.back();
  const SceneSlot& base = t->scenes[scene];
  s.view = base.view;
  s.view.camera = *camera;
  s.view_of = static_cast<int32_t>(scene);
  s.width = width;
  s.height = height;
  s.spp = spp;
  t->accel_valid = false;
  if (index) *index = static_cast<uint32_t>(t->scenes.size() - 1u);
  t->last_error.clear();
  return PT_OK;
}

int pt_batch_load_textures(pt_batch* b, uint32_t scene, const float* texcoords, uint32_t texcoord_count,
                           const float* texels, uint64_t texel_count, const PtTextureDesc* textures,
                           uint32_t texture_count) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (scene >= t->scenes.size() || t->scenes[scene].view_of >= 0)
    return fail(t, PT_INVALID_ARGUMENT, "load_textures: not a scene of the batch");
  const int rc =
      copy_textures(t, t->scenes[scene], texcoords, texcoord_count, texels, texel_count, textures, texture_count);
  if (rc != PT_OK) return rc;
  t->accel_valid = false;
  t->last_error.clear();
  return PT_OK;
}

int pt_batch_build(pt_batch* b, PtBuildInfo* info) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (t->scenes.empty()) return fail(t, PT_NOT_READY, "batch_build: no scenes");
  (void)hipSetDevice(t->device);
  const auto t0 = std::chrono::steady_clock::now();
  // the scenes are flattened in parallel host threads (largest first), their SAH trees built together
  // in one device build, then the host trees and device layouts again in parallel threads
  std::vector<size_t> order;
  for (size_t i = 0; i < t->scenes.size(); ++i)
    if (t->scenes[i].view_of < 0) order.push_back(i);
  std::sort(order.begin(), order.end(),
            [&](size_t a, size_t c) { return t->scenes[a].indices.size() > t->scenes[c].indices.size(); });
  const unsigned nthreads = std::max(1u, std::min<unsigned>(16u, std::thread::hardware_concurrency()));
  std::vector<char> ok(t->scenes.size(), 1);  // one byte per scene: threads write different elements
  std::vector<FlatScene> flat(t->scenes.size());
  std::vector<std::chrono::steady_clock::duration> own(t->scenes.size());
  auto each_scene = [&](auto fn) {
    std::vector<std::thread> pool;
    std::atomic<size_t> cursor{0};
    for (unsigned w = 0; w < std::min<size_t>(nthreads, order.size()); ++w)
      pool.emplace_back([&] {
        for (size_t k = cursor++; k < order.size(); k = cursor++) {
          const auto s0 = std::chrono::steady_clock::now();
          fn(order[k]);
          own[order[k]] += std::chrono::steady_clock::now() - s0;
        }
      });
    for (std::thread& th : pool) th.join();
  };
  for (size_t i : order) t->scenes[i].forest = -1;
  each_scene([&](size_t i) { ok[i] = flatten_scene(t->scenes[i], flat[i], 1u); });
  for (size_t i = 0; i < ok.size(); ++i)
    if (!ok[i]) return fail(t, PT_INVALID_ARGUMENT, "scene " + std::to_string(i) + ": " + t->scenes[i].error);
  std::vector<ptg::DeviceBuild> jobs;
  std::vector<SceneSlot*> job_scene;
  for (size_t i : order) {
    SceneSlot& s = t->scenes[i];
    if (!wants_device_build(s)) continue;
    jobs.push_back(device_job(s, flat[i]));
    job_scene.push_back(&s);
  }
  std::vector<ptg::DeviceForest> forests;
  const auto d0 = std::chrono::steady_clock::now();
  {
    std::string err;
    if (!build_on_device(t->stream, job_scene, jobs, forests, err)) return fail(t, PT_INTERNAL_ERROR, err);
  }
  const double device_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - d0).count();
  each_scene([&](size_t i) {
    finish_scene(t->scenes[i], flat[i], kept_tree(t->scenes[i], forests), 1u, device_ms, own[i]);
    flat[i] = FlatScene{};
  });
  int rc = upload_pool(t, forests);
  if (rc != PT_OK) return rc;
  rc = alloc_film(t, t->film_pixels);
  if (rc != PT_OK) return rc;
  ptd::launch_zero_counters(t->stream, t->d_counters.ptr);
  t->info.build_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (info) *info = t->info;
  t->accel_valid = true;
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_batch_reset_accumulation(pt_batch* b) { return b ? pt_hip_reset_accumulation(&b->t) : PT_INVALID_ARGUMENT; }

int pt_batch_update_scene_delta(pt_batch* b, const PtMaterial* materials, uint32_t material_count,
                                const PtLight* lights, uint32_t light_count) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "batch_update_scene_delta requires pt_batch_build");
  if (materials && material_count != t->dscene.material_count)
    return fail(t, PT_INVALID_ARGUMENT, "batch_update_scene_delta: material count differs from the pool's");
  if (lights && light_count != t->dscene.light_count)
    return fail(t, PT_INVALID_ARGUMENT, "batch_update_scene_delta: light count differs from the pool's");
  (void)hipSetDevice(t->device);
  PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");
  if (materials) {
    std::vector<PtMaterial> pool;
    pool.reserve(material_count);
    std::vector<ptd::SceneRec> recs(t->scenes.size());
    PT_HIP(t, hipMemcpy(recs.data(), t->d_recs.ptr, recs.size() * sizeof(ptd::SceneRec), hipMemcpyDeviceToHost),
           "read scene records");
    size_t at = 0;
    uint32_t tex_base = 0, any = 0, textured = 0;
    for (size_t si = 0; si < t->scenes.size(); ++si) {
      SceneSlot& sl = t->scenes[si];
      if (sl.view_of >= 0) {
        recs[si].has_transmissive = recs[static_cast<size_t>(sl.view_of)].has_transmissive;
        continue;
      }
      sl.materials.assign(materials + at, materials + at + sl.materials.size());
      at += sl.materials.size();
      const std::vector<PtMaterial> pm = pool_materials(sl, tex_base);
      tex_base += static_cast<uint32_t>(sl.textures.size());
      pool.insert(pool.end(), pm.begin(), pm.end());
      recs[si].has_transmissive = 0;
      for (const PtMaterial& m : sl.materials)
        if (transmissive(m)) recs[si].has_transmissive = 1;
      any |= recs[si].has_transmissive;
      for (const PtMaterial& m : pm)
        if (m.base_color_texture_index != 0xFFFFFFFFu) textured = 1;
    }
    PT_HIP(t, hipMemcpy(t->d_materials.ptr, pool.data(), sizeof(PtMaterial) * pool.size(), hipMemcpyHostToDevice),
           "upload materials");
    PT_HIP(t, hipMemcpy(t->d_recs.ptr, recs.data(), recs.size() * sizeof(ptd::SceneRec), hipMemcpyHostToDevice),
           "upload scene records");
    t->dscene.has_transmissive = any;
    t->dscene.textured = textured;
    drop_graph(t);  // the captured launches hold the scene flags by value
  }
  if (lights && light_count) {
    size_t at = 0;
    for (SceneSlot& sl : t->scenes) {
      if (sl.view_of >= 0) continue;
      sl.lights.assign(lights + at, lights + at + sl.lights.size());
      at += sl.lights.size();
    }
    PT_HIP(t, hipMemcpy(t->d_lights.ptr, lights, sizeof(PtLight) * light_count, hipMemcpyHostToDevice),
           "upload lights");
  }
  ++t->scene_version;
  t->last_error.clear();
  return PT_OK;
}

int pt_batch_update_texels(pt_batch* b, const float* texels, uint64_t texel_count) {
  return b ? pt_hip_update_texels(&b->t, texels, texel_count) : PT_INVALID_ARGUMENT;
}

// A batch frame: samples 0..spp-1 of every scene's (and view's) full image, scene after scene.
static int batch_frame_work(pt_tracer* t, std::vector<ptd::DevTile>& work) {
  for (uint32_t si = 0; si < t->scenes.size(); ++si) {
    const SceneSlot& s = t->scenes[si];
    std::vector<PtRenderTile> tiles;
    for (uint32_t k = 0; k < s.spp; ++k) tiles.push_back(PtRenderTile{0u, 0u, s.width, s.height, k, 0u, 0u});
    const int rc = build_work(t, si, tiles.data(), static_cast<uint32_t>(tiles.size()), work);
    if (rc != PT_OK) return rc;
  }
  return PT_OK;
}

int pt_batch_render_adjoint(pt_batch* b, uint64_t frame_index, const float* film_grad, float* grad_albedo,
                            float* grad_emission, float* grad_light, float* grad_roughness, float* grad_texels) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "batch_render_adjoint requires pt_batch_build");
  (void)hipSetDevice(t->device);
  std::vector<ptd::DevTile> work;
  const int rc = batch_frame_work(t, work);
  if (rc != PT_OK) return rc;
  return run_adjoint(t, work, frame_index, film_grad, grad_albedo, grad_emission, grad_light, grad_roughness,
                     grad_texels);
}

int pt_batch_render(pt_batch* b, uint64_t frame_index, int use_graph) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "batch_render requires pt_batch_build");
  (void)hipSetDevice(t->device);
  std::vector<ptd::DevTile> work;
  int wrc = batch_frame_work(t, work);
  if (wrc != PT_OK) return wrc;
  ptd::DevSettings st = dev_settings(t, frame_index);
  if (!use_graph) {
    const int rc = run_passes(t, work, st, nullptr);
    if (rc != PT_OK) return rc;
  } else {
    if (!t->graph_exec) {
      std::vector<Pass> passes;
      std::vector<ptd::FilmJob> jobs;
      int rc = plan_passes(t, work, passes, jobs);
      if (rc == PT_OK) rc = prepare_passes(t, passes, nullptr);
      if (rc != PT_OK) return rc;
      PT_HIP(t, hipStreamSynchronize(t->stream), "synchronize");  // staging copies done before capture
      t->stage_pending = false;
      st.frame_dev = t->d_frame.ptr;
      const bool prof = t->profiling;
      t->profiling = false;
      PT_HIP(t, hipStreamBeginCapture(t->stream, hipStreamCaptureModeThreadLocal), "begin graph capture");
      rc = launch_passes(t, work, passes, st, nullptr);
      hipGraph_t g = nullptr;
      const hipError_t ec = hipStreamEndCapture(t->stream, &g);
      t->profiling = prof;
      t->fwd_rad_valid = false;
      if (rc != PT_OK) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
      }
      if (ec != hipSuccess) return hip_fail(t, ec, "end graph capture");
      t->graph = g;
      PT_HIP(t, hipGraphInstantiate(&t->graph_exec, t->graph, nullptr, nullptr, 0), "instantiate graph");
    }
    const int rc = stage_upload(t, t->d_frame.ptr, &frame_index, sizeof(uint64_t));
    if (rc != PT_OK) return rc;
    // the replay rewrites every path's radiance: an adjoint must not take it for an earlier eager
    // render's
    t->fwd_rad_valid = false;
    PT_HIP(t, hipGraphLaunch(t->graph_exec, t->stream), "launch graph");
  }
  if (t->profiling) {
    const int rc = sync_all(t);
    if (rc != PT_OK) return rc;
  }
  t->last_error.clear();
  return PT_OK;
}

int pt_batch_read_film(pt_batch* b, uint32_t scene, float* rgb_sum, uint32_t* sample_count) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (!t->accel_valid) return fail(t, PT_NOT_READY, "batch_read_film requires pt_batch_build");
  if (scene >= t->scenes.size()) return fail(t, PT_INVALID_ARGUMENT, "batch_read_film: no such scene");
  (void)hipSetDevice(t->device);
  size_t base = 0;
  for (uint32_t i = 0; i < scene; ++i) base += static_cast<size_t>(t->scenes[i].width) * t->scenes[i].height;
  const SceneSlot& s = t->scenes[scene];
  return read_film_range(t, base, static_cast<size_t>(s.width) * s.height, rgb_sum, sample_count);
}

int pt_batch_load_sdf_primitives(pt_batch* b, uint32_t scene, const PtSdfPrimitive* prims, uint32_t count) {
  if (!b) return PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (scene >= t->scenes.size() || t->scenes[scene].view_of >= 0)
    return fail(t, PT_INVALID_ARGUMENT, "batch_load_sdf_primitives: no such scene (views share their scene's)");
  const int rc = set_sdfs(t, t->scenes[scene], prims, count);
  if (rc != PT_OK) return rc;
  t->accel_valid = false;
  t->last_error.clear();
  return PT_OK;
}

int pt_batch_update_camera(pt_batch* b, uint32_t scene, const PtCamera* camera) {
  if (!b || !camera) return b ? fail(&b->t, PT_INVALID_ARGUMENT, "batch_update_camera: null argument") : PT_INVALID_ARGUMENT;
  pt_tracer* t = &b->t;
  if (scene >= t->scenes.size()) return fail(t, PT_INVALID_ARGUMENT, "batch_update_camera: no such scene");
  (void)hipSetDevice(t->device);
  const int rc = update_slot_camera(t, scene, *camera);
  if (rc != PT_OK) return rc;
  t->last_error.clear();
  return PT_OK;
}

int pt_batch_resolve_hdr(pt_batch* b, uint32_t scene, float* rgb) {
  if (!b || !rgb) return b ? fail(&b->t, PT_INVALID_ARGUMENT, "batch_resolve_hdr: null argument") : PT_INVALID_ARGUMENT;
  size_t base = 0, pixels = 0;
  const int rc = batch_film_range(b, scene, "batch_resolve_hdr", base, pixels);
  if (rc != PT_OK) return rc;
  (void)hipSetDevice(b->t.device);
  return resolve_range(&b->t, base, pixels, nullptr, rgb);
}

int pt_batch_resolve_ldr(pt_batch* b, uint32_t scene, const PtFilmResolve* settings, uint8_t* rgba8) {
  if (!b || !rgba8) return b ? fail(&b->t, PT_INVALID_ARGUMENT, "batch_resolve_ldr: null argument") : PT_INVALID_ARGUMENT;
  size_t base = 0, pixels = 0;
  int rc = batch_film_range(b, scene, "batch_resolve_ldr", base, pixels);
  if (rc != PT_OK) return rc;
  ptd::DevResolve rs{};
  rc = make_resolve(&b->t, settings, rs);
  if (rc != PT_OK) return rc;
  (void)hipSetDevice(b->t.device);
  return resolve_range(&b->t, base, pixels, &rs, rgba8);
}

int pt_batch_read_counters(pt_batch* b, PtSampleCounters* out) {
  return b ? pt_hip_read_counters(&b->t, out) : PT_INVALID_ARGUMENT;
}
int pt_batch_set_tuning(pt_batch* b, const PtTuning* tu) { return b ? pt_hip_set_tuning(&b->t, tu) : PT_INVALID_ARGUMENT; }
int pt_batch_get_tuning(pt_batch* b, PtTuning* tu) { return b ? pt_hip_get_tuning(&b->t, tu) : PT_INVALID_ARGUMENT; }
int pt_batch_set_profiling(pt_batch* b, int enabled) {
  return b ? pt_hip_set_profiling(&b->t, enabled) : PT_INVALID_ARGUMENT;
}
int pt_batch_read_kernel_stats(pt_batch* b, PtKernelStats* out) {
  return b ? pt_hip_read_kernel_stats(&b->t, out) : PT_INVALID_ARGUMENT;
}
int pt_batch_reset_kernel_stats(pt_batch* b) { return b ? pt_hip_reset_kernel_stats(&b->t) : PT_INVALID_ARGUMENT; }
int pt_batch_synchronize(pt_batch* b) { return b ? pt_hip_synchronize(&b->t) : PT_INVALID_ARGUMENT; }

}  // extern "C"

namespace {
// The tree pt_hip_build_accel builds for `scene`, exported: on the host, or (device) the SAH tree on
// the current HIP device through a stream of its own.
int bvh_build_export(const PtSceneView* scene, const PtBvhParams* params, bool device, uint32_t* node_count,
                     uint32_t* prim_count, uint32_t* max_depth, float* bounds, int32_t* meta, uint32_t* prim_order,
                     float* prim_geo, uint32_t* built_on_device) {
  if (!scene) return PT_INVALID_ARGUMENT;
  PtBvhParams p{PT_BVH_SAH, 4u, 8u, 0u};
  if (params) p = *params;
  if (p.leaf_size == 0) p.leaf_size = 4u;
  if (p.bucket_count == 0) p.bucket_count = 8u;
  if (p.leaf_size > ptb::kMaxLeafSize || (p.builder != PT_BVH_SAH && p.builder != PT_BVH_MIDPOINT))
    return PT_INVALID_ARGUMENT;
  ptb::Bvh bvh;
  ptb::HostVec<float> raw;
  std::string err;
  const unsigned threads = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  if (!ptb::flatten(*scene, bvh.prims, raw, err, nullptr, nullptr, threads)) return PT_INVALID_ARGUMENT;
  int rc = 1;
  if (device && p.builder == PT_BVH_SAH) {
    hipStream_t stream = nullptr;
    if (hipStreamCreate(&stream) != hipSuccess) return PT_INTERNAL_ERROR;
    rc = ptg::build_sah_device(stream, bvh, raw, p.leaf_size, p.bucket_count, err);
    (void)hipStreamDestroy(stream);
    if (rc < 0) return PT_INTERNAL_ERROR;
  }
  if (built_on_device) *built_on_device = rc == 0 ? 1u : 0u;
  if (rc > 0) {
    if (p.builder == PT_BVH_SAH) ptb::build_sah(bvh, raw, p.leaf_size, p.bucket_count);
    else ptb::build_midpoint(bvh, raw, p.leaf_size);
  }
  if (node_count) *node_count = static_cast<uint32_t>(bvh.nodes.size());
  if (prim_count) *prim_count = static_cast<uint32_t>(bvh.prims.size());
  if (max_depth) *max_depth = bvh.max_depth;
  for (size_t i = 0; i < bvh.nodes.size(); ++i) {
    const ptb::Node& n = bvh.nodes[i];
    if (bounds)
      for (int k = 0; k < 3; ++k) {
        bounds[6 * i + k] = n.mn[k];
        bounds[6 * i + 3 + k] = n.mx[k];
      }
    if (meta) {
      meta[4 * i] = n.left;
      meta[4 * i + 1] = n.right;
      meta[4 * i + 2] = n.first;
      meta[4 * i + 3] = n.count;
    }
  }
  if (prim_order) std::copy(bvh.order.begin(), bvh.order.end(), prim_order);
  if (prim_geo)
    for (size_t i = 0; i < bvh.prims.size(); ++i) {
      const ptb::Prim& q = bvh.prims[i];
      float* g = prim_geo + 12 * i;
      for (int k = 0; k < 3; ++k) {
        g[k] = q.v0[k];
        g[3 + k] = q.e1[k];
        g[6 + k] = q.e2[k];
        g[9 + k] = q.n[k];
      }
    }
  return PT_OK;
}
}  // namespace

extern "C" int pt_bvh_build_host(const PtSceneView* scene, const PtBvhParams* params, uint32_t* node_count,
                                 uint32_t* prim_count, uint32_t* max_depth, float* bounds, int32_t* meta,
                                 uint32_t* prim_order, float* prim_geo) {
  return bvh_build_export(scene, params, false, node_count, prim_count, max_depth, bounds, meta, prim_order, prim_geo,
                          nullptr);
}

extern "C" int pt_bvh_build_device(const PtSceneView* scene, const PtBvhParams* params, uint32_t* node_count,
                                   uint32_t* prim_count, uint32_t* max_depth, float* bounds, int32_t* meta,
                                   uint32_t* prim_order, float* prim_geo, uint32_t* built_on_device) {
  return bvh_build_export(scene, params, true, node_count, prim_count, max_depth, bounds, meta, prim_order, prim_geo,
                          built_on_device);
}
