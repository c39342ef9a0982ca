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

// IRayAccelerator (PathTracer.h:1001-1036): build over a snapshot, then batched closest-hit or
// any-hit queries on device buffers (pt_hip_trace_rays).
class HipRayAccelerator {
 public:
  explicit HipRayAccelerator(int device = 0) : tracer_(device) {}
  Status build(const SceneSnapshot& scene, const PtBvhParams* params = nullptr);
  // rays: 6 floats per ray (origin, direction); tmax: per-ray limit or nullptr; all device pointers.
  Status intersect(bool any_hit, const float* rays, const float* tmax, uint32_t count, float* out_t,
                   int32_t* out_prim, PtSampleCounters* counters = nullptr);
  const PtBuildInfo& build_info() const { return tracer_.build_info(); }
  void reset() { tracer_.shutdown(); }

 private:
  HipPathTracer tracer_;
};

}  // namespace vkpt_hip
