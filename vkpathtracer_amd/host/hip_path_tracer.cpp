// hip_path_tracer.cpp — the C++ host classes over libvkpt_hip.so.
#include "hip_path_tracer.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <limits>

namespace vkpt_hip {

std::string_view Status::code_name() const {
  static constexpr std::string_view kNames[] = {"ok",    "unsupported", "invalid_argument", "not_ready",     "busy",
                                                "alloc_failed", "timeout",     "cancelled",        "internal_error"};
  return code >= 0 && code < static_cast<int>(std::size(kNames)) ? kNames[code] : "unknown";
}

// ------------------------------------------------------------------ snapshots
PtSceneView SceneSnapshot::view() const {
  PtSceneView v{};
  v.vertices = vertices.empty() ? nullptr : vertices.data();
  v.indices = indices.empty() ? nullptr : indices.data();
  v.local_vertices = local_vertices.empty() ? nullptr : local_vertices.data();
  v.local_indices = local_indices.empty() ? nullptr : local_indices.data();
  v.instances = instances.empty() ? nullptr : instances.data();
  v.materials = materials.empty() ? nullptr : materials.data();
  v.lights = lights.empty() ? nullptr : lights.data();
  v.environment_map = environment_map.empty() ? nullptr : environment_map.data();
  v.vertex_count = static_cast<uint32_t>(vertices.size() / 3u);
  v.index_count = static_cast<uint32_t>(indices.size());
  v.local_vertex_count = static_cast<uint32_t>(local_vertices.size() / 3u);
  v.local_index_count = static_cast<uint32_t>(local_indices.size());
  v.instance_count = static_cast<uint32_t>(instances.size());
  v.material_count = static_cast<uint32_t>(materials.size());
  v.light_count = static_cast<uint32_t>(lights.size());
  v.environment_map_width = environment_map.empty() ? 0u : environment_map_width;
  v.environment_map_height = environment_map.empty() ? 0u : environment_map_height;
  for (int k = 0; k < 3; ++k) {
    v.environment_color[k] = environment_color[k];
    v.environment_map_scale[k] = environment_map_scale[k];
  }
  v.camera = camera;
  return v;
}

uint64_t SceneSnapshot::triangle_count() const {
  uint64_t n = 0;
  for (const PtInstance& i : instances) n += i.triangle_count;
  return n;
}

namespace {

constexpr char kMagic[8] = {'V', 'K', 'P', 'T', 'S', 'N', 'A', 'P'};
constexpr uint32_t kVersion = 1, kVersionSdf = 2;

// Header after the magic: version, then the element counts of each array.
struct SnapHeader {
  uint32_t version;
  uint32_t vertex_count, index_count, local_vertex_count, local_index_count;
  uint32_t instance_count, material_count, light_count;
  uint32_t environment_map_width, environment_map_height;
  uint32_t texcoord_count, texture_count;
  uint64_t texel_count;
  float environment_color[3], environment_map_scale[3];
  PtCamera camera;
};

template <typename T>
bool read_array(std::ifstream& f, std::vector<T>& v, uint64_t n) {
  if (n > (1ull << 34) / sizeof(T)) return false;  // 16 GiB per array at most
  v.resize(static_cast<size_t>(n));
  if (n) f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
  return static_cast<bool>(f);
}

template <typename T>
void write_array(std::ofstream& f, const std::vector<T>& v) {
  if (!v.empty()) f.write(reinterpret_cast<const char*>(v.data()), static_cast<std::streamsize>(v.size() * sizeof(T)));
}

}  // namespace

bool load_snapshot_file(const std::string& path, SceneSnapshot& out, std::string& error) {
  std::ifstream f(path, std::ios::binary);
  if (!f) {
    error = "cannot open " + path;
    return false;
  }
  char magic[8];
  SnapHeader h{};
  f.read(magic, sizeof(magic));
  f.read(reinterpret_cast<char*>(&h), sizeof(h));
  if (!f || std::memcmp(magic, kMagic, sizeof(kMagic)) != 0 || (h.version != kVersion && h.version != kVersionSdf)) {
    error = path + ": not a version-1 or version-2 scene snapshot";
    return false;
  }
  SceneSnapshot s;
  bool ok = read_array(f, s.vertices, 3ull * h.vertex_count) && read_array(f, s.indices, h.index_count) &&
                  read_array(f, s.local_vertices, 3ull * h.local_vertex_count) &&
                  read_array(f, s.local_indices, h.local_index_count) &&
                  read_array(f, s.instances, h.instance_count) && read_array(f, s.materials, h.material_count) &&
                  read_array(f, s.lights, h.light_count) &&
                  read_array(f, s.environment_map,
                             3ull * h.environment_map_width * static_cast<uint64_t>(h.environment_map_height)) &&
                  read_array(f, s.texcoords, 2ull * h.texcoord_count) && read_array(f, s.texels, 3ull * h.texel_count) &&
                  read_array(f, s.textures, h.texture_count);
  uint32_t sdf_count = 0;
  if (ok && h.version == kVersionSdf) {
    f.read(reinterpret_cast<char*>(&sdf_count), sizeof(sdf_count));
    ok = static_cast<bool>(f) && read_array(f, s.sdf_primitives, sdf_count);
  }
  if (!ok) {
    error = path + ": truncated scene snapshot";
    return false;
  }
  s.environment_map_width = h.environment_map_width;
  s.environment_map_height = h.environment_map_height;
  for (int k = 0; k < 3; ++k) {
    s.environment_color[k] = h.environment_color[k];
    s.environment_map_scale[k] = h.environment_map_scale[k];
  }
  s.camera = h.camera;
  out = std::move(s);
  return true;
}

bool save_snapshot_file(const std::string& path, const SceneSnapshot& s, std::string& error) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    error = "cannot write " + path;
    return false;
  }
  SnapHeader h{};
  h.version = s.sdf_primitives.empty() ? kVersion : kVersionSdf;
  h.vertex_count = static_cast<uint32_t>(s.vertices.size() / 3u);
  h.index_count = static_cast<uint32_t>(s.indices.size());
  h.local_vertex_count = static_cast<uint32_t>(s.local_vertices.size() / 3u);
  h.local_index_count = static_cast<uint32_t>(s.local_indices.size());
  h.instance_count = static_cast<uint32_t>(s.instances.size());
  h.material_count = static_cast<uint32_t>(s.materials.size());
  h.light_count = static_cast<uint32_t>(s.lights.size());
  h.environment_map_width = s.environment_map.empty() ? 0u : s.environment_map_width;
  h.environment_map_height = s.environment_map.empty() ? 0u : s.environment_map_height;
  h.texcoord_count = static_cast<uint32_t>(s.texcoords.size() / 2u);
  h.texture_count = static_cast<uint32_t>(s.textures.size());
  h.texel_count = s.texels.size() / 3u;
  for (int k = 0; k < 3; ++k) {
    h.environment_color[k] = s.environment_color[k];
    h.environment_map_scale[k] = s.environment_map_scale[k];
  }
  h.camera = s.camera;
  f.write(kMagic, sizeof(kMagic));
  f.write(reinterpret_cast<const char*>(&h), sizeof(h));
  write_array(f, s.vertices);
  write_array(f, s.indices);
  write_array(f, s.local_vertices);
  write_array(f, s.local_indices);
  write_array(f, s.instances);
  write_array(f, s.materials);
  write_array(f, s.lights);
  write_array(f, s.environment_map);
  write_array(f, s.texcoords);
  write_array(f, s.texels);
  write_array(f, s.textures);
  if (h.version == kVersionSdf) {
    const uint32_t sdf_count = static_cast<uint32_t>(s.sdf_primitives.size());
    f.write(reinterpret_cast<const char*>(&sdf_count), sizeof(sdf_count));
    write_array(f, s.sdf_primitives);
  }
  if (!f) {
    error = "write failed: " + path;
    return false;
  }
  return true;
}

// ------------------------------------------------------------------ HipPathTracer
HipPathTracer::HipPathTracer(int device) {
  const int rc = pt_hip_create(device, &handle_);
  if (rc != PT_OK) {
    handle_ = nullptr;
    error_ = "pt_hip_create failed (no HIP device " + std::to_string(device) + ")";
  }
}

HipPathTracer::~HipPathTracer() {
  if (handle_) pt_hip_destroy(handle_);
}

Status HipPathTracer::to_status(int rc) const {
  Status s;
  s.code = rc;
  if (rc != PT_OK) s.message = handle_ ? pt_hip_last_error(handle_) : error_;
  return s;
}

bool HipPathTracer::fail_bool(int rc) const {
  if (rc == PT_OK) return true;
  if (handle_) error_ = pt_hip_last_error(handle_);
  return false;
}

Status HipPathTracer::configure(const PtRenderSettings& settings) {
  if (!handle_) return Status{PT_NOT_READY, error_};
  settings_ = settings;
  return to_status(pt_hip_configure(handle_, &settings));
}

Status HipPathTracer::load_scene_snapshot(const SceneSnapshot& scene) {
  if (!handle_) return Status{PT_NOT_READY, error_};
  const PtSceneView v = scene.view();
  int rc = pt_hip_load_scene(handle_, &v);
  if (rc == PT_OK && (!scene.texcoords.empty() || !scene.textures.empty()))
    rc = pt_hip_load_textures(handle_, scene.texcoords.empty() ? nullptr : scene.texcoords.data(),
                              static_cast<uint32_t>(scene.texcoords.size() / 2u),
                              scene.texels.empty() ? nullptr : scene.texels.data(), scene.texels.size() / 3u,
                              scene.textures.empty() ? nullptr : scene.textures.data(),
                              static_cast<uint32_t>(scene.textures.size()));
  if (rc == PT_OK && !scene.sdf_primitives.empty())
    rc = pt_hip_load_sdf_primitives(handle_, scene.sdf_primitives.data(),
                                    static_cast<uint32_t>(scene.sdf_primitives.size()));
  return to_status(rc);
}

Status HipPathTracer::build_or_update_acceleration(const PtBvhParams* params) {
  if (!handle_) return Status{PT_NOT_READY, error_};
  return to_status(pt_hip_build_accel(handle_, params, &build_info_));
}

bool HipPathTracer::reset_accumulation() { return handle_ && fail_bool(pt_hip_reset_accumulation(handle_)); }

bool HipPathTracer::render_tile(const PtRenderTile& tile, uint32_t frame_index) {
  return handle_ && fail_bool(pt_hip_render_tile(handle_, &tile, frame_index));
}

bool HipPathTracer::render_tile_cancellable(const PtRenderTile& tile, uint32_t frame_index, std::stop_token stop) {
  if (stop.stop_requested()) return false;
  return render_tile(tile, frame_index);
}

bool HipPathTracer::render_tiles(std::span<const PtRenderTile> tiles, uint32_t frame_index, std::stop_token stop) {
  if (!handle_ || stop.stop_requested()) return false;
  if (tiles.size() > std::numeric_limits<uint32_t>::max()) return false;
  return fail_bool(pt_hip_render_tiles(handle_, tiles.data(), static_cast<uint32_t>(tiles.size()), frame_index));
}

bool HipPathTracer::update_scene_delta(std::span<const PtMaterial> materials, std::span<const PtLight> lights) {
  if (!handle_) return false;
  return fail_bool(pt_hip_update_scene_delta(handle_, materials.empty() ? nullptr : materials.data(),
                                             static_cast<uint32_t>(materials.size()),
                                             lights.empty() ? nullptr : lights.data(),
                                             static_cast<uint32_t>(lights.size())));
}

bool HipPathTracer::update_scene_delta(const SceneDeltaUpdate& update) {
  if (!handle_) return false;
  PtSceneDeltaUpdate u{};
  u.materials = update.materials.empty() ? nullptr : update.materials.data();
  u.lights = update.lights.empty() ? nullptr : update.lights.data();
  u.instance_materials = update.instance_materials.empty() ? nullptr : update.instance_materials.data();
  u.material_count = static_cast<uint32_t>(update.materials.size());
  u.light_count = static_cast<uint32_t>(update.lights.size());
  u.instance_material_count = static_cast<uint32_t>(update.instance_materials.size());
  u.environment_color_changed = update.environment_color_changed ? 1u : 0u;
  for (int k = 0; k < 3; ++k) u.environment_color[k] = update.environment_color[k];
  return fail_bool(pt_hip_apply_scene_delta(handle_, &u));
}

FilmHdr HipPathTracer::resolve_hdr() const {
  FilmHdr out;
  if (!handle_) return out;
  out.width = settings_.width;
  out.height = settings_.height;
  out.rgb.resize(3ull * out.width * out.height);
  if (!fail_bool(pt_hip_resolve_hdr(handle_, out.rgb.data()))) out = FilmHdr{};
  return out;
}

FilmLdr HipPathTracer::resolve_ldr(const PtFilmResolve* settings) const {
  FilmLdr out;
  if (!handle_) return out;
  out.width = settings_.width;
  out.height = settings_.height;
  out.rgba8.resize(4ull * out.width * out.height);
  if (!fail_bool(pt_hip_resolve_ldr(handle_, settings, out.rgba8.data()))) out = FilmLdr{};
  return out;
}

bool HipPathTracer::replace_film_history(const Film& film) {
  if (!handle_) return false;
  if (film.width != settings_.width || film.height != settings_.height ||
      film.rgb_sum.size() != 3ull * film.width * film.height ||
      film.counts.size() != static_cast<size_t>(film.width) * film.height)
    return false;
  return fail_bool(pt_hip_replace_film(handle_, film.rgb_sum.data(), film.counts.data()));
}

Film HipPathTracer::film() const {
  Film out;
  if (!handle_) return out;
  out.width = settings_.width;
  out.height = settings_.height;
  out.rgb_sum.resize(3ull * out.width * out.height);
  out.counts.resize(static_cast<size_t>(out.width) * out.height);
  if (!fail_bool(pt_hip_read_film(handle_, out.rgb_sum.data(), out.counts.data()))) out = Film{};
  return out;
}

PtSampleCounters HipPathTracer::read_counters() const {
  PtSampleCounters c{};
  if (handle_) fail_bool(pt_hip_read_counters(handle_, &c));
  return c;
}

TracerStatus HipPathTracer::status() const {
  TracerStatus s;
  if (handle_) {
    pt_hip_status(handle_, &s.state);
    s.last_error = pt_hip_last_error(handle_);
  } else {
    s.last_error = error_;
  }
  return s;
}

void HipPathTracer::shutdown() {
  if (handle_) pt_hip_shutdown(handle_);
}

bool HipPathTracer::set_tuning(const PtTuning& tuning) { return handle_ && fail_bool(pt_hip_set_tuning(handle_, &tuning)); }

PtTuning HipPathTracer::tuning() const {
  PtTuning t{};
  if (handle_) fail_bool(pt_hip_get_tuning(handle_, &t));
  return t;
}

// ------------------------------------------------------------------ HipRayAccelerator
bool HipRayAccelerator::build(const SceneSnapshot& scene, bool deterministic) {
  return build(scene, deterministic, PtBvhParams{PT_BVH_MIDPOINT, 4u, 8u, 0u});
}

bool HipRayAccelerator::build(const SceneSnapshot& scene, bool deterministic, const PtBvhParams& params) {
  reset();
  info_.deterministic = deterministic;
  PtRenderSettings st{};
  st.width = 1;
  st.height = 1;
  st.spp = 1;
  st.max_depth = 1;
  st.rr_min_survival = 0.1f;
  st.rr_max_survival = 0.99f;
  Status s = tracer_.configure(st);
  if (s) s = tracer_.load_scene_snapshot(scene);
  if (s) s = tracer_.build_or_update_acceleration(&params);
  if (!s) {
    error_ = s.message;
    return false;
  }
  const PtBuildInfo& b = tracer_.build_info();
  info_.built = b.built != 0u;
  info_.primitive_count = b.primitive_count;
  info_.node_count = b.node_count;
  info_.leaf_count = b.leaf_count;
  info_.build_ms = b.build_ms;
  error_.clear();
  return info_.built && (b.node_count > 0u || b.primitive_count == 0u);
}

bool HipRayAccelerator::intersect(const Ray& ray, RayQueryHit& out, RayQueryStats* stats) const {
  return intersect(std::span<const Ray>(&ray, 1), std::span<RayQueryHit>(&out, 1), stats) && out.hit;
}

bool HipRayAccelerator::intersect(std::span<const Ray> rays, std::span<RayQueryHit> out, RayQueryStats* stats) const {
  for (RayQueryHit& h : out) h = RayQueryHit{};
  if (!tracer_.handle() || !info_.built || out.size() != rays.size()) {
    error_ = out.size() != rays.size() ? "intersect: output and ray counts differ" : "intersect before build";
    return false;
  }
  if (rays.empty()) return true;
  static_assert(sizeof(Ray) == 24 && sizeof(Vec3) == 12, "Ray must be 6 packed floats");
  std::vector<PtRayQueryHit> hits(rays.size());
  PtSampleCounters c{};
  const int rc = pt_hip_intersect(tracer_.handle(), 0, reinterpret_cast<const float*>(rays.data()), nullptr,
                                  static_cast<uint32_t>(rays.size()), hits.data(), &c);
  if (rc != PT_OK) {
    error_ = pt_hip_last_error(tracer_.handle());
    return false;
  }
  for (std::size_t i = 0; i < hits.size(); ++i) {
    const PtRayQueryHit& h = hits[i];
    RayQueryHit& o = out[i];
    o.hit = h.hit != 0u;
    o.t = h.t;
    o.position = Vec3{h.position[0], h.position[1], h.position[2]};
    o.normal = Vec3{h.normal[0], h.normal[1], h.normal[2]};
    o.material_index = h.material_index;
    o.primitive_index = h.primitive_index;
  }
  if (stats) {  // the reference's intersect adds to the caller's stats
    stats->triangle_tests += c.triangle_tests;
    stats->triangle_hits += c.triangle_hits;
    stats->bvh_node_visits += c.bvh_node_visits;
    stats->bvh_leaf_visits += c.bvh_leaf_visits;
  }
  return true;
}

Status HipRayAccelerator::trace_rays(bool any_hit, const float* rays, const float* tmax, uint32_t count, float* out_t,
                                     int32_t* out_prim, PtSampleCounters* counters) {
  if (!tracer_.handle()) return Status{PT_NOT_READY, tracer_.status().last_error};
  Status s;
  s.code = pt_hip_trace_rays(tracer_.handle(), any_hit ? 1 : 0, rays, tmax, count, out_t, out_prim, counters);
  if (!s) s.message = pt_hip_last_error(tracer_.handle());
  return s;
}

void HipRayAccelerator::reset() {
  tracer_.shutdown();
  info_ = RayAcceleratorBuildInfo{};
}

}  // namespace vkpt_hip
