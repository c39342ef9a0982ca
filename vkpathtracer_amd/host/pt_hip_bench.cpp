// pt_hip_bench — a ptbench-style command line over HipPathTracer (C++ host, C ABI underneath).
//
// Follows ptbench's render protocol (src/benchmark/BenchmarkRuntime.cpp:580-695): the scene
// conversion and BVH build are timed as build_ms, then reset_accumulation and one full-frame tile
// per sample are timed as render_ms; results.json carries BenchmarkResult's fields
// (BenchmarkSchema.h:86-119) with throughput.samples_per_sec = W*H*spp / render time and
// throughput.paths_per_sec = extension rays / render time (BenchmarkRuntime.cpp:755-757), plus
// the counters and Mpaths/s, Mrays/s. The output directory holds the artifact set of ptbench's
// contract (DefaultBenchmarkArtifactContract, BenchmarkSchema.cpp:758-775; WriteRunArtifacts,
// BenchmarkRunArtifacts.cpp:107-190): results.json / .csv, metadata.json, scene_snapshot.json,
// shader_manifest.json, asset_manifest.json, beauty.png (the GPU's LDR resolve), beauty.exr (the GPU's
// HDR resolve as OpenEXR), logs.jsonl and profiler_trace.json (host phases and per-kernel GPU time),
// plus film.bin (rgb sums and counts) and beauty.pfm.
//
//   pt_hip_bench --scene FILE.ptsnap [--width W] [--height H] [--spp N] [--max-depth D] [--seed S]
//                [--rr-start K] [--no-nee] [--no-mis] [--builder sah|midpoint] [--leaf L]
//                [--warmup F] [--frames F] [--per-sample-launches] [--device I] [--out DIR]
//                [--exposure E] [--white-balance K] [--tone-map linear|reinhard|filmic_approx|aces_approx]
//                [--output-transform gamma|linear] [--gamma G] [--no-clamp]
//   pt_hip_bench --check-snapshot FILE     (no device: print the snapshot's counts as JSON)
//   pt_hip_bench --validate-artifacts DIR  (no device: ValidateBenchmarkArtifactsOnDisk's report)
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iterator>
#include <span>
#include <string>
#include <vector>

#include "artifacts.hpp"
#include "hip_path_tracer.hpp"

namespace {

struct Options {
  std::string scene, out = "pt_hip_bench_out", check, validate, ray_queries;
  uint32_t single = 256;  // --ray-queries: rays also answered one intersect() call each
  PtFilmResolve resolve{1.0f, 6500.0f, 0u, 0u, 2.2f, 1u};  // FilmResolveSettings{}
  uint32_t width = 256, height = 256, spp = 16, max_depth = 8, rr_start = 3, leaf = 4, warmup = 1, frames = 1;
  uint64_t seed = 0xBADC0FFEEull;
  int device = 0;
  bool nee = true, mis = true, per_sample = false, midpoint = false;
};

int usage() {
  std::fprintf(stderr,
               "usage: pt_hip_bench --scene FILE.ptsnap [--width W] [--height H] [--spp N] [--max-depth D]\n"
               "                    [--seed S] [--rr-start K] [--no-nee] [--no-mis] [--builder sah|midpoint]\n"
               "                    [--leaf L] [--warmup F] [--frames F] [--per-sample-launches] [--device I]\n"
               "                    [--out DIR] [--exposure E] [--white-balance K] [--tone-map NAME]\n"
               "                    [--output-transform gamma|linear] [--gamma G] [--no-clamp]\n"
               "       pt_hip_bench --scene FILE.ptsnap --ray-queries RAYS.f32 [--single K] [--builder B]\n"
               "                    [--leaf L] [--out DIR]\n"
               "       pt_hip_bench --check-snapshot FILE\n"
               "       pt_hip_bench --validate-artifacts DIR\n");
  return 2;
}

bool parse(int argc, char** argv, Options& o) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    auto next = [&](const char* name) -> const char* {
      if (i + 1 >= argc) {
        std::fprintf(stderr, "%s needs a value\n", name);
        return nullptr;
      }
      return argv[++i];
    };
    auto num = [&](const char* name, auto& dst) {
      const char* v = next(name);
      if (!v) return false;
      char* end = nullptr;
      const unsigned long long x = std::strtoull(v, &end, 0);
      if (!end || *end) {
        std::fprintf(stderr, "%s: not a number: %s\n", name, v);
        return false;
      }
      dst = static_cast<std::remove_reference_t<decltype(dst)>>(x);
      return true;
    };
    auto fnum = [&](const char* name, float& dst) {
      const char* v = next(name);
      if (!v) return false;
      char* end = nullptr;
      const double x = std::strtod(v, &end);
      if (!end || *end || !std::isfinite(x)) {
        std::fprintf(stderr, "%s: not a number: %s\n", name, v);
        return false;
      }
      dst = static_cast<float>(x);
      return true;
    };
    auto choice = [&](const char* name, std::initializer_list<const char*> names, uint32_t& dst) {
      const char* v = next(name);
      if (!v) return false;
      uint32_t i = 0;
      for (const char* n : names) {
        if (std::strcmp(v, n) == 0) {
          dst = i;
          return true;
        }
        ++i;
      }
      std::fprintf(stderr, "%s: unknown value %s\n", name, v);
      return false;
    };
    bool ok = true;
    if (a == "--scene") { const char* v = next("--scene"); ok = v; if (v) o.scene = v; }
    else if (a == "--check-snapshot") { const char* v = next("--check-snapshot"); ok = v; if (v) o.check = v; }
    else if (a == "--out") { const char* v = next("--out"); ok = v; if (v) o.out = v; }
    else if (a == "--width") ok = num("--width", o.width);
    else if (a == "--height") ok = num("--height", o.height);
    else if (a == "--spp") ok = num("--spp", o.spp);
    else if (a == "--max-depth") ok = num("--max-depth", o.max_depth);
    else if (a == "--seed") ok = num("--seed", o.seed);
    else if (a == "--rr-start") ok = num("--rr-start", o.rr_start);
    else if (a == "--leaf") ok = num("--leaf", o.leaf);
    else if (a == "--warmup") ok = num("--warmup", o.warmup);
    else if (a == "--frames") ok = num("--frames", o.frames);
    else if (a == "--device") ok = num("--device", o.device);
    else if (a == "--ray-queries") { const char* v = next("--ray-queries"); ok = v; if (v) o.ray_queries = v; }
    else if (a == "--single") ok = num("--single", o.single);
    else if (a == "--no-nee") o.nee = false;
    else if (a == "--no-mis") o.mis = false;
    else if (a == "--per-sample-launches") o.per_sample = true;
    else if (a == "--validate-artifacts") { const char* v = next("--validate-artifacts"); ok = v; if (v) o.validate = v; }
    else if (a == "--exposure") ok = fnum("--exposure", o.resolve.exposure);
    else if (a == "--white-balance") ok = fnum("--white-balance", o.resolve.white_balance_kelvin);
    else if (a == "--gamma") ok = fnum("--gamma", o.resolve.gamma);
    else if (a == "--tone-map")
      ok = choice("--tone-map", {"linear", "reinhard", "filmic_approx", "aces_approx"}, o.resolve.tone_map);
    else if (a == "--output-transform") ok = choice("--output-transform", {"gamma", "linear"}, o.resolve.output_transform);
    else if (a == "--no-clamp") o.resolve.clamp_output = 0u;
    else if (a == "--builder") {
      const char* v = next("--builder");
      ok = v && (std::strcmp(v, "sah") == 0 || std::strcmp(v, "midpoint") == 0);
      if (ok) o.midpoint = std::strcmp(v, "midpoint") == 0;
    } else if (a == "--help" || a == "-h") return false;
    else {
      std::fprintf(stderr, "unknown argument %s\n", a.c_str());
      return false;
    }
    if (!ok) return false;
  }
  return !o.scene.empty() || !o.check.empty() || !o.validate.empty();
}

std::string json_escape(const std::string& s) {
  std::string r;
  for (char c : s) {
    if (c == '"' || c == '\\') r += '\\';
    if (static_cast<unsigned char>(c) < 0x20) continue;
    r += c;
  }
  return r;
}

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int check_snapshot(const std::string& path) {
  vkpt_hip::SceneSnapshot s;
  std::string err;
  if (!vkpt_hip::load_snapshot_file(path, s, err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 1;
  }
  double vsum = 0.0;
  for (float v : s.vertices) vsum += v;
  double tsum = 0.0;
  for (float v : s.texels) tsum += v;
  double ssum = 0.0;  // every float field of the SDF records, plus shape and material index
  for (const PtSdfPrimitive& p : s.sdf_primitives) {
    ssum += p.shape + p.material_index + p.radius + p.param_a + p.param_b;
    for (int k = 0; k < 3; ++k) ssum += p.position[k] + p.scale[k] + p.rotation[k];
  }
  std::printf(
      "{\"vertices\": %zu, \"indices\": %zu, \"local_vertices\": %zu, \"local_indices\": %zu, \"instances\": %zu, "
      "\"materials\": %zu, \"lights\": %zu, \"environment_map\": [%u, %u], \"texcoords\": %zu, \"texels\": %zu, "
      "\"textures\": %zu, \"triangles\": %llu, \"vertex_sum\": %.9g, \"texel_sum\": %.9g, \"fov_deg\": %.9g, "
      "\"sdf_primitives\": %zu, \"sdf_sum\": %.9g}\n",
      s.vertices.size() / 3, s.indices.size(), s.local_vertices.size() / 3, s.local_indices.size(), s.instances.size(),
      s.materials.size(), s.lights.size(), s.environment_map_width, s.environment_map_height, s.texcoords.size() / 2,
      s.texels.size() / 3, s.textures.size(), static_cast<unsigned long long>(s.triangle_count()), vsum, tsum,
      static_cast<double>(s.camera.fov_deg), s.sdf_primitives.size(), ssum);
  return 0;
}

bool write_film(const std::string& dir, const vkpt_hip::Film& film, const vkpt_hip::FilmHdr& hdr) {
  {
    std::ofstream f(dir + "/film.bin", std::ios::binary | std::ios::trunc);
    const uint32_t wh[2] = {film.width, film.height};
    f.write(reinterpret_cast<const char*>(wh), sizeof(wh));
    f.write(reinterpret_cast<const char*>(film.rgb_sum.data()),
            static_cast<std::streamsize>(film.rgb_sum.size() * sizeof(float)));
    f.write(reinterpret_cast<const char*>(film.counts.data()),
            static_cast<std::streamsize>(film.counts.size() * sizeof(uint32_t)));
    if (!f) return false;
  }
  // Portable float map: little-endian (negative scale), rows bottom to top.
  std::ofstream f(dir + "/beauty.pfm", std::ios::binary | std::ios::trunc);
  const std::string header = "PF\n" + std::to_string(hdr.width) + " " + std::to_string(hdr.height) + "\n-1.0\n";
  f.write(header.data(), static_cast<std::streamsize>(header.size()));
  for (uint32_t r = 0; r < hdr.height; ++r) {
    const float* row = hdr.rgb.data() + 3ull * hdr.width * (hdr.height - 1u - r);
    f.write(reinterpret_cast<const char*>(row), static_cast<std::streamsize>(3ull * hdr.width * sizeof(float)));
  }
  return static_cast<bool>(f);
}

// IRayAccelerator queries (PathTracer.h:1001-1036) through HipRayAccelerator: the rays of a raw file
// (6 little-endian floats each) answered in one batched intersect, the first `single` of them again by
// one intersect(Ray, RayQueryHit&, RayQueryStats*) call each. Writes DIR/hits.bin (PtRayQueryHit records
// of the batch) and DIR/ray_queries.json (build info, summed stats of both forms, and how many of the
// single-ray answers differ from the batch's).
int run_ray_queries(const Options& o, const vkpt_hip::SceneSnapshot& scene) {
  std::ifstream f(o.ray_queries, std::ios::binary);
  std::vector<char> raw((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
  if (!f.good() && !f.eof()) {
    std::fprintf(stderr, "cannot read %s\n", o.ray_queries.c_str());
    return 2;
  }
  if (raw.size() % sizeof(vkpt_hip::Ray) != 0) {
    std::fprintf(stderr, "%s: size is not a multiple of 24 bytes\n", o.ray_queries.c_str());
    return 2;
  }
  std::vector<vkpt_hip::Ray> rays(raw.size() / sizeof(vkpt_hip::Ray));
  if (!rays.empty()) std::memcpy(rays.data(), raw.data(), raw.size());
  vkpt_hip::HipRayAccelerator accel(o.device);
  const PtBvhParams bp{o.midpoint ? PT_BVH_MIDPOINT : PT_BVH_SAH, o.leaf, 8u, 0u};
  if (!accel.build(scene, true, bp)) {
    std::fprintf(stderr, "build failed: %s\n", accel.last_error().c_str());
    return 2;
  }
  std::vector<vkpt_hip::RayQueryHit> hits(rays.size());
  vkpt_hip::RayQueryStats batch_stats{}, single_stats{};
  if (!accel.intersect(std::span<const vkpt_hip::Ray>(rays), std::span<vkpt_hip::RayQueryHit>(hits), &batch_stats)) {
    std::fprintf(stderr, "intersect failed: %s\n", accel.last_error().c_str());
    return 2;
  }
  const size_t single = std::min<size_t>(o.single, rays.size());
  size_t mismatched = 0, single_hits = 0;
  for (size_t i = 0; i < single; ++i) {
    vkpt_hip::RayQueryHit h;
    const bool hit = accel.intersect(rays[i], h, &single_stats);
    single_hits += hit ? 1u : 0u;
    const vkpt_hip::RayQueryHit& b = hits[i];
    if (hit != b.hit || std::memcmp(&h.t, &b.t, sizeof(float)) != 0 || h.primitive_index != b.primitive_index ||
        h.material_index != b.material_index || std::memcmp(&h.position, &b.position, sizeof(h.position)) != 0 ||
        std::memcmp(&h.normal, &b.normal, sizeof(h.normal)) != 0)
      ++mismatched;
  }
  std::filesystem::create_directories(o.out);
  std::vector<PtRayQueryHit> recs(hits.size());
  for (size_t i = 0; i < hits.size(); ++i) {
    const vkpt_hip::RayQueryHit& h = hits[i];
    recs[i] = PtRayQueryHit{h.hit ? 1u : 0u, h.t, {h.position.x, h.position.y, h.position.z},
                            {h.normal.x, h.normal.y, h.normal.z}, h.material_index, h.primitive_index};
  }
  std::ofstream hb(std::filesystem::path(o.out) / "hits.bin", std::ios::binary);
  hb.write(reinterpret_cast<const char*>(recs.data()), static_cast<std::streamsize>(recs.size() * sizeof(PtRayQueryHit)));
  const vkpt_hip::RayAcceleratorBuildInfo bi = accel.build_info();
  std::ofstream js(std::filesystem::path(o.out) / "ray_queries.json");
  auto stats_json = [](const vkpt_hip::RayQueryStats& s) {
    return "{\"triangle_tests\": " + std::to_string(s.triangle_tests) + ", \"triangle_hits\": " +
           std::to_string(s.triangle_hits) + ", \"bvh_node_visits\": " + std::to_string(s.bvh_node_visits) +
           ", \"bvh_leaf_visits\": " + std::to_string(s.bvh_leaf_visits) + "}";
  };
  js << "{\"rays\": " << rays.size() << ", \"single\": " << single << ", \"single_hits\": " << single_hits
     << ", \"single_mismatched\": " << mismatched << ", \"build_info\": {\"built\": " << (bi.built ? "true" : "false")
     << ", \"primitive_count\": " << bi.primitive_count << ", \"node_count\": " << bi.node_count
     << ", \"leaf_count\": " << bi.leaf_count << ", \"deterministic\": " << (bi.deterministic ? "true" : "false")
     << "}, \"batch_stats\": " << stats_json(batch_stats) << ", \"single_stats\": " << stats_json(single_stats)
     << "}\n";
  if (!hb || !js) {
    std::fprintf(stderr, "cannot write %s\n", o.out.c_str());
    return 2;
  }
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  Options o;
  if (!parse(argc, argv, o)) return usage();
  if (!o.check.empty()) return check_snapshot(o.check);
  if (!o.validate.empty()) {
    const vkpt_hip::ArtifactValidation v = vkpt_hip::validate_artifacts(o.validate);
    std::printf("%s\n", v.to_json().c_str());
    return v.ok ? 0 : 1;
  }

  const auto t_total = std::chrono::steady_clock::now();
  vkpt_hip::SceneSnapshot scene;
  std::string err;
  if (!vkpt_hip::load_snapshot_file(o.scene, scene, err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 2;
  }
  if (!o.ray_queries.empty()) return run_ray_queries(o, scene);
  vkpt_hip::HipPathTracer tracer(o.device);
  if (!tracer.valid()) {
    std::fprintf(stderr, "%s\n", tracer.status().last_error.c_str());
    return 2;
  }
  PtRenderSettings st{};
  st.seed = o.seed;
  st.width = o.width;
  st.height = o.height;
  st.spp = o.spp;
  st.max_depth = o.max_depth;
  st.enable_nee = o.nee;
  st.enable_mis = o.mis;
  st.deterministic = 1;
  st.rr_start_depth = o.rr_start;
  st.rr_min_survival = 0.1f;
  st.rr_max_survival = 0.99f;
  const auto t_build = std::chrono::steady_clock::now();
  PtBvhParams bp{o.midpoint ? PT_BVH_MIDPOINT : PT_BVH_SAH, o.leaf, 8u, 0u};
  vkpt_hip::Status s = tracer.configure(st);
  if (s) s = tracer.load_scene_snapshot(scene);
  if (s) s = tracer.build_or_update_acceleration(&bp);
  if (!s) {
    std::fprintf(stderr, "%s: %s\n", std::string(s.code_name()).c_str(), s.message.c_str());
    return 2;
  }
  const double build_ms = ms_since(t_build);

  std::vector<PtRenderTile> tiles;
  for (uint32_t k = 0; k < o.spp; ++k) tiles.push_back(PtRenderTile{0u, 0u, o.width, o.height, k, 0u, 0u});
  auto frame = [&](uint32_t f) {
    if (!tracer.reset_accumulation()) return false;
    if (o.per_sample) {
      for (const PtRenderTile& t : tiles)
        if (!tracer.render_tile(t, f)) return false;
      return true;
    }
    return tracer.render_tiles(tiles, f);
  };
  for (uint32_t w = 0; w < o.warmup; ++w)
    if (!frame(0) || pt_hip_synchronize(tracer.handle()) != PT_OK) {
      std::fprintf(stderr, "render failed: %s\n", tracer.status().last_error.c_str());
      return 2;
    }
  const auto t_render = std::chrono::steady_clock::now();
  for (uint32_t f = 0; f < o.frames; ++f)
    if (!frame(0)) {
      std::fprintf(stderr, "render failed: %s\n", tracer.status().last_error.c_str());
      return 2;
    }
  if (pt_hip_synchronize(tracer.handle()) != PT_OK) {
    std::fprintf(stderr, "render failed: %s\n", tracer.status().last_error.c_str());
    return 2;
  }
  const double render_ms = ms_since(t_render) / o.frames;  // per frame
  const PtSampleCounters c = tracer.read_counters();       // the last frame's (reset clears them)

  // One more frame under per-kernel event timing (outside the timed region; it renders the same frame,
  // so the film resolved below is the timed frames' film): profiler_trace.json's GPU events.
  PtKernelStats ks{};
  bool have_ks = pt_hip_set_profiling(tracer.handle(), 1) == PT_OK && pt_hip_reset_kernel_stats(tracer.handle()) == PT_OK &&
                 frame(0) && pt_hip_synchronize(tracer.handle()) == PT_OK &&
                 pt_hip_read_kernel_stats(tracer.handle(), &ks) == PT_OK;
  (void)pt_hip_set_profiling(tracer.handle(), 0);

  const auto t_resolve = std::chrono::steady_clock::now();
  std::filesystem::create_directories(o.out);
  const vkpt_hip::Film film = tracer.film();
  const vkpt_hip::FilmHdr hdr = tracer.resolve_hdr();
  const vkpt_hip::FilmLdr ldr = tracer.resolve_ldr(&o.resolve);
  if (hdr.rgb.empty() || ldr.rgba8.empty()) {
    std::fprintf(stderr, "resolve failed: %s\n", tracer.status().last_error.c_str());
    return 2;
  }
  const double resolve_ms = ms_since(t_resolve);
  std::string werr;
  const std::string png = o.out + "/beauty.png", exr = o.out + "/beauty.exr";
  if (!write_film(o.out, film, hdr) || !vkpt_hip::save_png(png, ldr, werr) || !vkpt_hip::save_exr(exr, hdr, werr)) {
    std::fprintf(stderr, "cannot write the images to %s %s\n", o.out.c_str(), werr.c_str());
    return 2;
  }
  std::string png_bytes;
  {
    std::ifstream in(png, std::ios::binary);
    png_bytes.assign(std::istreambuf_iterator<char>(in), std::istreambuf_iterator<char>());
  }
  const std::string image_hash = vkpt_hip::hex64(vkpt_hip::fnv1a64(png_bytes.data(), png_bytes.size()));
  const double total_ms = ms_since(t_total);
  const double samples = static_cast<double>(o.width) * o.height * o.spp;
  const double sps = render_ms > 0.0 ? 1000.0 * samples / render_ms : 0.0;
  const double pps = render_ms > 0.0 ? 1000.0 * static_cast<double>(c.rays) / render_ms : 0.0;
  const double rays_all = static_cast<double>(c.rays) + static_cast<double>(c.shadow_tests);
  const std::string run_id = "pt_hip_bench-" + image_hash.substr(0, 8);
  const std::string scene_name = std::filesystem::path(o.scene).stem().string();
  const std::string dir = json_escape(o.out);

  std::ofstream j(o.out + "/results.json", std::ios::trunc);
  char buf[8192];
  std::snprintf(
      buf, sizeof(buf),
      "{\n"
      " \"run_id\": \"%s\",\n \"scene\": \"%s\",\n \"backend\": \"hip\",\n \"renderer_path\": \"hip-wavefront\",\n"
      " \"cpu_simd_mode\": \"none\",\n \"resolution\": {\"width\": %u, \"height\": %u},\n \"spp\": %u,\n"
      " \"seed\": %llu,\n \"max_depth\": %u,\n"
      " \"device_info\": {\"backend\": \"hip\", \"renderer_path\": \"hip-wavefront\", \"gpu_name\": \"%s\"},\n"
      " \"timing\": {\"total_ms\": %.3f, \"build_ms\": %.3f, \"render_ms\": %.3f, \"cpu_ms\": %.3f},\n"
      " \"timing_breakdown\": [{\"name\": \"scene_build\", \"category\": \"scene\", \"ms\": %.3f},"
      " {\"name\": \"render_samples\", \"category\": \"render\", \"ms\": %.3f},"
      " {\"name\": \"resolve\", \"category\": \"resolve\", \"ms\": %.3f}],\n"
      " \"throughput\": {\"paths_per_sec\": %.1f, \"samples_per_sec\": %.1f},\n"
      " \"mpaths_per_s\": %.3f,\n \"mrays_per_s\": %.3f,\n"
      " \"counters\": {\"samples\": %llu, \"rays\": %llu, \"triangle_tests\": %llu, \"triangle_hits\": %llu,"
      " \"bvh_node_visits\": %llu, \"bvh_leaf_visits\": %llu, \"shadow_tests\": %llu},\n"
      " \"bvh\": {\"builder\": \"%s\", \"leaf_size\": %u, \"nodes\": %u, \"max_depth\": %u, \"device_bytes\": %llu},\n"
      " \"resolve\": %s,\n"
      " \"frames\": %u,\n \"launches\": \"%s\",\n \"image_hash\": \"%s\",\n"
      " \"beauty_png\": \"%s/beauty.png\",\n \"beauty_exr\": \"%s/beauty.exr\",\n \"beauty_hdr\": \"%s/beauty.pfm\",\n"
      " \"logs_jsonl\": \"%s/logs.jsonl\",\n \"profiler_trace_json\": \"%s/profiler_trace.json\",\n"
      " \"film\": \"%s/film.bin\"\n}\n",
      run_id.c_str(), json_escape(o.scene).c_str(), o.width, o.height, o.spp, static_cast<unsigned long long>(o.seed),
      o.max_depth, json_escape(pt_hip_version()).c_str(), total_ms, build_ms, render_ms, render_ms, build_ms, render_ms,
      resolve_ms, pps, sps, sps / 1e6, render_ms > 0.0 ? rays_all / render_ms / 1e3 : 0.0,
      static_cast<unsigned long long>(c.samples), static_cast<unsigned long long>(c.rays),
      static_cast<unsigned long long>(c.triangle_tests), static_cast<unsigned long long>(c.triangle_hits),
      static_cast<unsigned long long>(c.bvh_node_visits), static_cast<unsigned long long>(c.bvh_leaf_visits),
      static_cast<unsigned long long>(c.shadow_tests), o.midpoint ? "midpoint" : "sah", o.leaf,
      tracer.build_info().node_count, tracer.build_info().max_depth,
      static_cast<unsigned long long>(tracer.build_info().device_bytes),
      vkpt_hip::resolve_settings_json(o.resolve).c_str(), o.frames,
      o.per_sample ? "one render_tile per sample" : "one render_tiles per frame", image_hash.c_str(), dir.c_str(),
      dir.c_str(), dir.c_str(), dir.c_str(), dir.c_str(), dir.c_str());
  j << buf;
  j.close();
  std::printf("%s", buf);

  // results.csv: WriteRunArtifacts' columns (BenchmarkRunArtifacts.cpp:135-148).
  std::ofstream csv(o.out + "/results.csv", std::ios::trunc);
  csv << "run_id,scene,backend,renderer_path,resolution_width,resolution_height,spp,max_depth,total_ms,build_ms,"
         "render_ms,cpu_ms,paths_per_sec,samples_per_sec,samples_per_sec_per_thread,paths_per_sec_per_thread,"
         "samples_per_sec_per_megapixel,normalized_score,reference_error,image_hash\n";
  char row[1024];
  const double mp = static_cast<double>(o.width) * o.height / 1e6;
  std::snprintf(row, sizeof(row), "%s,%s,hip,hip-wavefront,%u,%u,%u,%u,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%.6f,%s\n",
                run_id.c_str(), o.scene.c_str(), o.width, o.height, o.spp, o.max_depth, total_ms, build_ms, render_ms,
                render_ms, pps, sps, sps, pps, mp > 0 ? sps / mp : 0.0, 0.0, 0.0, image_hash.c_str());
  csv << row;
  csv.close();

  // metadata.json (SerializeMetadata, BenchmarkRunArtifacts.cpp:85-105): what this renderer can do.
  std::ofstream meta(o.out + "/metadata.json", std::ios::trunc);
  meta << "{\"run_id\":\"" << run_id << "\",\"command\":\"pt_hip_bench\",\"scene_path\":\"" << json_escape(o.scene)
       << "\",\"backend\":\"hip\",\"renderer_path\":\"hip-wavefront\",\"scene_name\":\"" << json_escape(scene_name)
       << "\",\"schema\":\"VKPTSNAP " << (scene.sdf_primitives.empty() ? 1 : 2)
       << "\",\"material_count\":" << scene.materials.size()
       << ",\"instance_count\":" << scene.instances.size() << ",\"triangle_count\":" << scene.triangle_count()
       << ",\"benchmark_capabilities\":{\"backends\":[\"hip\"],\"renderer_paths\":[\"hip-wavefront\"],"
          "\"tile_rendering\":true,\"tile_batching\":true,\"deterministic\":true,\"film_resolve\":\"gpu\","
          "\"artifacts\":[\"results.json\",\"results.csv\",\"beauty.png\",\"beauty.exr\",\"profiler_trace.json\"]},"
          "\"profiler_capabilities\":{\"host_phases\":true,\"gpu_kernel_events\":true}}\n";
  meta.close();

  // scene_snapshot.json (SerializeSceneSnapshot, BenchmarkRunArtifacts.cpp:18-83), from the snapshot file.
  std::ofstream snap(o.out + "/scene_snapshot.json", std::ios::trunc);
  snap << "{\"entity_ids\":[";
  for (size_t i = 0; i < scene.instances.size(); ++i) snap << (i ? "," : "") << scene.instances[i].entity_id;
  snap << "],\"renderables\":[";
  for (size_t i = 0; i < scene.instances.size(); ++i) {
    const PtInstance& in = scene.instances[i];
    snap << (i ? "," : "") << "{\"entity_id\":" << in.entity_id << ",\"mesh_id\":" << in.geometry_id
         << ",\"material_id\":" << in.material_index << ",\"transform\":[" << in.translation[0] << ","
         << in.translation[1] << "," << in.translation[2] << "," << in.rotation[0] << "," << in.rotation[1] << ","
         << in.rotation[2] << "," << in.rotation[3] << "," << in.scale[0] << "," << in.scale[1] << "," << in.scale[2]
         << "]}";
  }
  snap << "],\"lights\":[";
  for (size_t i = 0; i < scene.lights.size(); ++i) {
    const PtLight& l = scene.lights[i];
    snap << (i ? "," : "") << "{\"position\":[" << l.position[0] << "," << l.position[1] << "," << l.position[2]
         << "],\"color\":[" << l.color[0] << "," << l.color[1] << "," << l.color[2] << "],\"intensity\":"
         << l.intensity << ",\"radius\":" << l.radius << "}";
  }
  snap << "],\"materials\":[";
  for (size_t i = 0; i < scene.materials.size(); ++i) {
    const PtMaterial& m = scene.materials[i];
    snap << (i ? "," : "") << "{\"id\":" << i << ",\"albedo\":[" << m.albedo[0] << "," << m.albedo[1] << ","
         << m.albedo[2] << "],\"roughness\":" << m.roughness << ",\"emission\":[" << m.emissive[0] << ","
         << m.emissive[1] << "," << m.emissive[2] << "],\"material_model\":" << m.material_model << "}";
  }
  snap << "],\"scene_hash\":\"" << vkpt_hip::hex64(vkpt_hip::fnv1a64(scene.vertices.data(), scene.vertices.size() * 4))
       << "\"}\n";
  snap.close();

  // shader_manifest.json: the device record layouts and kernels this renderer runs (where the reference
  // lists its GPU packing strides, PathTracer.h:778-788).
  std::ofstream sm(o.out + "/shader_manifest.json", std::ios::trunc);
  sm << "{\"kernels\":[\"k_raygen\",\"k_extend\",\"k_shade\",\"k_shadow\",\"k_film\",\"k_resolve\"],"
        "\"records\":{\"material\":" << pt_hip_sizeof(0) << ",\"light\":" << pt_hip_sizeof(1) << ",\"instance\":"
     << pt_hip_sizeof(2) << ",\"camera\":" << pt_hip_sizeof(3) << ",\"bvh_node\":64,\"triangle\":48},"
        "\"version\":\"" << json_escape(pt_hip_version()) << "\"}\n";
  sm.close();
  std::ofstream am(o.out + "/asset_manifest.json", std::ios::trunc);
  am << "{\"asset_refs\":[\"" << json_escape(o.scene) << "\"]}\n";
  am.close();

  // logs.jsonl (BenchmarkRuntime.cpp:873-889) and profiler_trace.json (SerializeProfilerTrace,
  // BenchmarkSchemaCapabilities.cpp:199-218).
  std::ofstream lg(o.out + "/logs.jsonl", std::ios::trunc);
  lg << "{\"severity\":\"info\",\"message\":\"run complete\",\"fields\":[{\"k\":\"scene\",\"v\":\""
     << json_escape(o.scene) << "\"},{\"k\":\"backend\",\"v\":\"hip\"},{\"k\":\"renderer\",\"v\":\"hip-wavefront\"},"
     << "{\"k\":\"total_ms\",\"v\":\"" << total_ms << "\"}]}\n";
  lg.close();
  std::ofstream tr(o.out + "/profiler_trace.json", std::ios::trunc);
  tr << "{\"events\":[{\"kind\":\"cpu\",\"name\":\"scene_build\",\"category\":\"scene\",\"thread_id\":0,\"start_ms\":0,"
     << "\"duration_ms\":" << build_ms << "},{\"kind\":\"cpu\",\"name\":\"render_samples\",\"category\":\"render\","
     << "\"thread_id\":0,\"start_ms\":" << build_ms << ",\"duration_ms\":" << render_ms * o.frames << "},"
     << "{\"kind\":\"cpu\",\"name\":\"resolve\",\"category\":\"resolve\",\"thread_id\":0,\"start_ms\":"
     << build_ms + render_ms * o.frames << ",\"duration_ms\":" << resolve_ms << "}";
  if (have_ks) {
    static const char* names[PT_KERNEL_COUNT] = {"raygen", "extend", "shade", "shadow", "film"};
    double t = 0.0;
    for (int k = 0; k < PT_KERNEL_COUNT; ++k) {
      tr << ",{\"kind\":\"gpu\",\"name\":\"" << names[k] << "\",\"category\":\"gpu_kernel\",\"thread_id\":1,"
         << "\"start_ms\":" << t << ",\"duration_ms\":" << ks.total_ms[k] << ",\"launches\":" << ks.launches[k] << "}";
      t += ks.total_ms[k];
    }
  }
  tr << "]}\n";
  tr.close();
  const vkpt_hip::ArtifactValidation v = vkpt_hip::validate_artifacts(o.out);
  if (!v.ok) {
    std::fprintf(stderr, "artifact contract not met: %s\n", v.to_json().c_str());
    return 2;
  }
  return 0;
}
