// hip_path_tracer.hpp — C++23 host classes over the C ABI (include/vkpt_hip.h).
//
// HipPathTracer has the shape of the reference's IPathTracer (src/pathtracer/PathTracer.h:1069-1184):
// configure / load_scene_snapshot / build_or_update_acceleration return a Status, reset_accumulation /
// render_tile / render_tiles / update_scene_delta return bool with the message kept for status(), and
// render_tiles takes a std::stop_token checked before the launch. A subclass of the reference's
// IPathTracer forwards each override to the member of the same name (INTEGRATION.md shows the
// adapter). HipRayAccelerator mirrors IRayAccelerator (PathTracer.h:1001-1036). Neither class owns
// any GPU code: every call goes through libvkpt_hip.so.
#pragma once

#include <cstddef>
#include <cstdint>
#include <span>
#include <stop_token>
#include <string>
#include <string_view>
#include <vector>

#include "vkpt_hip.h"

namespace vkpt_hip {

// vkpt::core::Status (src/core/contracts/Result.h:13-23): a code and a message; true when ok.
struct Status {
  int code = PT_OK;
  std::string message;
  bool ok() const { return code == PT_OK; }
  explicit operator bool() const { return ok(); }
  std::string_view code_name() const;  // StatusCodeName, Result.h:69-91
};

// PathTracerSceneSnapshot (PathTracer.h:358-409) as owned arrays, plus the albedo textures of the
// texture extension (texcoords: 2 floats per vertex; texels: 3 floats each).
struct SceneSnapshot {
  std::vector<float> vertices, local_vertices;
  std::vector<uint32_t> indices, local_indices;
  std::vector<PtInstance> instances;
  std::vector<PtMaterial> materials;
  std::vector<PtLight> lights;
  std::vector<float> environment_map;  // rgb, environment_map_width * environment_map_height
  uint32_t environment_map_width = 0, environment_map_height = 0;
  float environment_color[3] = {0.0f, 0.0f, 0.0f};
  float environment_map_scale[3] = {0.0f, 0.0f, 0.0f};
  PtCamera camera{};
  std::vector<float> texcoords, texels;
  std::vector<PtTextureDesc> textures;
  std::vector<PtSdfPrimitive> sdf_primitives;  // loaded after the triangles (pt_hip_load_sdf_primitives)

  // The C view of the arrays (valid while the snapshot is alive and unchanged).
  PtSceneView view() const;
  uint64_t triangle_count() const;
};

// Binary snapshot files ("VKPTSNAP", little-endian): a header of counts, then every array in
// SceneSnapshot's order (version 1); version 2 appends a u32 count and the SDF primitive records.
// Files without SDF primitives are written as version 1. vkpathtracer_amd.scene.Scene.save_snapshot
// writes the same format.
bool load_snapshot_file(const std::string& path, SceneSnapshot& out, std::string& error);
bool save_snapshot_file(const std::string& path, const SceneSnapshot& snapshot, std::string& error);

struct Film {  // FilmBuffer::raw / sample_counts (PathTracer.h:955-958)
  uint32_t width = 0, height = 0;
  std::vector<float> rgb_sum;     // 3 per pixel
  std::vector<uint32_t> counts;   // 1 per pixel
};

struct FilmHdr {  // FilmBuffer::resolve_hdr (FilmBuffer.cpp:160-214)
  uint32_t width = 0, height = 0;
  std::vector<float> rgb;  // 3 per pixel
};

struct FilmLdr {  // FilmLdr / ApplyFilmResolve (FilmBuffer.cpp:255-332)
  uint32_t width = 0, height = 0;
  std::vector<uint8_t> rgba8;  // 4 per pixel
};

// RTSceneDeltaUpdate (PathTracer.h:455-461): per-index material and light records, per-instance material
// indices and the environment colour.
struct SceneDeltaUpdate {
  std::vector<PtMaterialUpdate> materials;
  std::vector<PtLightUpdate> lights;
  std::vector<PtInstanceMaterialUpdate> instance_materials;
  bool environment_color_changed = false;
  float environment_color[3] = {0.0f, 0.0f, 0.0f};
};

struct TracerStatus {  // PathTracerStatus (PathTracer.h:830-851) and the last error message
  PtTracerStatus state{};
  std::string last_error;
};

class HipPathTracer {
 public:
  explicit HipPathTracer(int device = 0);
  ~HipPathTracer();
  HipPathTracer(const HipPathTracer&) = delete;
  HipPathTracer& operator=(const HipPathTracer&) = delete;

  // false when no HIP device could be opened (status().last_error says why)
  bool valid() const { return handle_ != nullptr; }

  Status configure(const PtRenderSettings& settings);
  Status load_scene_snapshot(const SceneSnapshot& scene);
  Status build_or_update_acceleration(const PtBvhParams* params = nullptr);
  bool reset_accumulation();
  bool render_tile(const PtRenderTile& tile, uint32_t frame_index);
  bool render_tile_cancellable(const PtRenderTile& tile, uint32_t frame_index, std::stop_token stop);
  // One wavefront launch sequence for all tiles; a stop requested before the launch returns false.
  bool render_tiles(std::span<const PtRenderTile> tiles, uint32_t frame_index, std::stop_token stop = {});
  bool supports_tile_rendering() const { return true; }
  bool supports_tile_batching() const { return true; }
  bool update_scene_delta(std::span<const PtMaterial> materials, std::span<const PtLight> lights);
  // IPathTracer::update_scene_delta(const RTSceneDeltaUpdate&) (PathTracer.h:1139): false when an index is
  // out of range or nothing changed, as ScalarCpuPathTracer::update_scene_delta.
  bool update_scene_delta(const SceneDeltaUpdate& update);
  FilmHdr resolve_hdr() const;
  // IPathTracer::resolve_ldr with the given FilmResolveSettings (nullptr: the defaults).
  FilmLdr resolve_ldr(const PtFilmResolve* settings = nullptr) const;
  // IPathTracer::replace_film_history (PathTracer.h:1096-1098).
  bool replace_film_history(const Film& film);
  Film film() const;
  PtSampleCounters read_counters() const;
  TracerStatus status() const;
  void shutdown();

  const PtBuildInfo& build_info() const { return build_info_; }
  bool set_tuning(const PtTuning& tuning);
  PtTuning tuning() const;
  pt_tracer* handle() const { return handle_; }

 private:
  bool fail_bool(int rc) const;
  Status to_status(int rc) const;

  pt_tracer* handle_ = nullptr;
  PtRenderSettings settings_{};
  PtBuildInfo build_info_{};
  mutable std::string error_;
};

// Ray, RayQueryHit, RayQueryStats and RayAcceleratorBuildInfo (PathTracer.h:30-53, :976-999), with the
// reference's field names so that an IRayAccelerator adapter copies them member for member.
struct Vec3 {
  float x = 0.0f, y = 0.0f, z = 0.0f;
};
struct Ray {
  Vec3 origin{};
  Vec3 direction{};
};
struct RayQueryHit {
  bool hit = false;
  float t = 0.0f;
  Vec3 position{};
  Vec3 normal{};
  uint32_t material_index = 0;
  uint32_t primitive_index = 0;
};
struct RayQueryStats {
  uint64_t triangle_tests = 0;
  uint64_t triangle_hits = 0;
  uint64_t bvh_node_visits = 0;
  uint64_t bvh_leaf_visits = 0;
};
struct RayAcceleratorBuildInfo {
  bool built = false;
  std::size_t primitive_count = 0;
  std::size_t node_count = 0;
  std::size_t leaf_count = 0;
  double build_ms = 0.0;
  bool deterministic = false;
};

// IRayAccelerator (PathTracer.h:1001-1036) over the GPU traversal. build() takes CpuBvhAccelerator's
// tree (midpoint splits, leaves of 4: ParallelBvhBuilder, CpuBvhAccelerator.cpp:252-256) unless a
// PtBvhParams names another; intersect() answers one query as the reference does (RayQueryHit filled,
// stats added to) and the span overload answers a batch in one launch; trace_rays() is the device-buffer
// form (t and primitive only). The builds are deterministic whatever the flag says; build_info()
// reports the flag as given, as the reference does.
class HipRayAccelerator {
 public:
  explicit HipRayAccelerator(int device = 0) : tracer_(device) {}
  bool build(const SceneSnapshot& scene, bool deterministic);
  bool build(const SceneSnapshot& scene, bool deterministic, const PtBvhParams& params);
  bool intersect(const Ray& ray, RayQueryHit& out, RayQueryStats* stats = nullptr) const;
  // One launch for all rays; out.size() must equal rays.size(). False (and every hit cleared) on an error.
  bool intersect(std::span<const Ray> rays, std::span<RayQueryHit> out, RayQueryStats* stats = nullptr) const;
  // rays: 6 floats per ray (origin, direction); tmax: per-ray limit or nullptr; all device pointers.
  Status trace_rays(bool any_hit, const float* rays, const float* tmax, uint32_t count, float* out_t,
                    int32_t* out_prim, PtSampleCounters* counters = nullptr);
  RayAcceleratorBuildInfo build_info() const { return info_; }
  void reset();
  const std::string& last_error() const { return error_; }

 private:
  HipPathTracer tracer_;
  RayAcceleratorBuildInfo info_{};
  mutable std::string error_;
};

}  // namespace vkpt_hip
