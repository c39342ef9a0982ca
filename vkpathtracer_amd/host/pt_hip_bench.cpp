// pt_hip_bench — a ptbench-style command line over HipPathTracer (C++ host, C ABI underneath).
//
// Follows ptbench's render protocol (src/benchmark/BenchmarkRuntime.cpp:580-695): the scene
// conversion and BVH build are timed as build_ms, then reset_accumulation and one full-frame tile
// per sample are timed as render_ms; results.json carries BenchmarkResult's fields
// (BenchmarkSchema.h:86-119) with throughput.samples_per_sec = W*H*spp / render time and
// throughput.paths_per_sec = extension rays / render time (BenchmarkRuntime.cpp:755-757), plus
// the counters and Mpaths/s, Mrays/s. The film is written as film.bin (rgb sums and counts) and the
// resolved HDR image as beauty.pfm.
//
//   pt_hip_bench --scene FILE.ptsnap [--width W] [--height H] [--spp N] [--max-depth D] [--seed S]
//                [--rr-start K] [--no-nee] [--no-mis] [--builder sah|midpoint] [--leaf L]
//                [--warmup F] [--frames F] [--per-sample-launches] [--device I] [--out DIR]
//   pt_hip_bench --check-snapshot FILE     (no device: print the snapshot's counts as JSON)
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <string>
#include <vector>

#include "hip_path_tracer.hpp"

namespace {

struct Options {
  std::string scene, out = "pt_hip_bench_out", check;
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
               "                    [--out DIR]\n"
               "       pt_hip_bench --check-snapshot FILE\n");
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
    else if (a == "--no-nee") o.nee = false;
    else if (a == "--no-mis") o.mis = false;
    else if (a == "--per-sample-launches") o.per_sample = true;
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
  return !o.scene.empty() || !o.check.empty();
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
  std::printf(
      "{\"vertices\": %zu, \"indices\": %zu, \"local_vertices\": %zu, \"local_indices\": %zu, \"instances\": %zu, "
      "\"materials\": %zu, \"lights\": %zu, \"environment_map\": [%u, %u], \"texcoords\": %zu, \"texels\": %zu, "
      "\"textures\": %zu, \"triangles\": %llu, \"vertex_sum\": %.9g, \"texel_sum\": %.9g, \"fov_deg\": %.9g}\n",
      s.vertices.size() / 3, s.indices.size(), s.local_vertices.size() / 3, s.local_indices.size(), s.instances.size(),
      s.materials.size(), s.lights.size(), s.environment_map_width, s.environment_map_height, s.texcoords.size() / 2,
      s.texels.size() / 3, s.textures.size(), static_cast<unsigned long long>(s.triangle_count()), vsum, tsum,
      static_cast<double>(s.camera.fov_deg));
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

}  // namespace

int main(int argc, char** argv) {
  Options o;
  if (!parse(argc, argv, o)) return usage();
  if (!o.check.empty()) return check_snapshot(o.check);

  const auto t_total = std::chrono::steady_clock::now();
  vkpt_hip::SceneSnapshot scene;
  std::string err;
  if (!vkpt_hip::load_snapshot_file(o.scene, scene, err)) {
    std::fprintf(stderr, "%s\n", err.c_str());
    return 2;
  }
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

  const auto t_resolve = std::chrono::steady_clock::now();
  std::filesystem::create_directories(o.out);
  const vkpt_hip::Film film = tracer.film();
  const vkpt_hip::FilmHdr hdr = tracer.resolve_hdr();
  if (!write_film(o.out, film, hdr)) {
    std::fprintf(stderr, "cannot write the film to %s\n", o.out.c_str());
    return 2;
  }
  const double resolve_ms = ms_since(t_resolve);
  const double total_ms = ms_since(t_total);
  const double samples = static_cast<double>(o.width) * o.height * o.spp;
  const double sps = render_ms > 0.0 ? 1000.0 * samples / render_ms : 0.0;
  const double pps = render_ms > 0.0 ? 1000.0 * static_cast<double>(c.rays) / render_ms : 0.0;
  const double rays_all = static_cast<double>(c.rays) + static_cast<double>(c.shadow_tests);

  std::ofstream j(o.out + "/results.json", std::ios::trunc);
  char buf[4096];
  std::snprintf(
      buf, sizeof(buf),
      "{\n"
      " \"scene\": \"%s\",\n \"backend\": \"hip\",\n \"renderer_path\": \"hip-wavefront\",\n"
      " \"cpu_simd_mode\": \"none\",\n \"resolution\": {\"width\": %u, \"height\": %u},\n \"spp\": %u,\n"
      " \"seed\": %llu,\n \"max_depth\": %u,\n"
      " \"device_info\": {\"backend\": \"hip\", \"renderer_path\": \"hip-wavefront\", \"gpu_name\": \"%s\"},\n"
      " \"timing\": {\"total_ms\": %.3f, \"build_ms\": %.3f, \"render_ms\": %.3f, \"cpu_ms\": %.3f},\n"
      " \"timing_breakdown\": [{\"name\": \"scene_build\", \"category\": \"scene\", \"ms\": %.3f},"
      " {\"name\": \"render_samples\", \"category\": \"render\", \"ms\": %.3f},"
      " {\"name\": \"resolve_and_write\", \"category\": \"io\", \"ms\": %.3f}],\n"
      " \"throughput\": {\"paths_per_sec\": %.1f, \"samples_per_sec\": %.1f},\n"
      " \"mpaths_per_s\": %.3f,\n \"mrays_per_s\": %.3f,\n"
      " \"counters\": {\"samples\": %llu, \"rays\": %llu, \"triangle_tests\": %llu, \"triangle_hits\": %llu,"
      " \"bvh_node_visits\": %llu, \"bvh_leaf_visits\": %llu, \"shadow_tests\": %llu},\n"
      " \"bvh\": {\"builder\": \"%s\", \"leaf_size\": %u, \"nodes\": %u, \"max_depth\": %u, \"device_bytes\": %llu},\n"
      " \"frames\": %u,\n \"launches\": \"%s\",\n \"beauty_hdr\": \"%s/beauty.pfm\",\n \"film\": \"%s/film.bin\"\n}\n",
      json_escape(o.scene).c_str(), o.width, o.height, o.spp, static_cast<unsigned long long>(o.seed), o.max_depth,
      json_escape(pt_hip_version()).c_str(), total_ms, build_ms, render_ms, render_ms, build_ms, render_ms, resolve_ms,
      pps, sps, sps / 1e6, render_ms > 0.0 ? rays_all / render_ms / 1e3 : 0.0,
      static_cast<unsigned long long>(c.samples), static_cast<unsigned long long>(c.rays),
      static_cast<unsigned long long>(c.triangle_tests), static_cast<unsigned long long>(c.triangle_hits),
      static_cast<unsigned long long>(c.bvh_node_visits), static_cast<unsigned long long>(c.bvh_leaf_visits),
      static_cast<unsigned long long>(c.shadow_tests), o.midpoint ? "midpoint" : "sah", o.leaf,
      tracer.build_info().node_count, tracer.build_info().max_depth,
      static_cast<unsigned long long>(tracer.build_info().device_bytes), o.frames,
      o.per_sample ? "one render_tile per sample" : "one render_tiles per frame", json_escape(o.out).c_str(),
      json_escape(o.out).c_str());
  j << buf;
  std::printf("%s", buf);
  return j ? 0 : 2;
}
