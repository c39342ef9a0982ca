// artifacts.cpp — PNG / OpenEXR writers and the benchmark artifact contract (see artifacts.hpp).
#include "artifacts.hpp"

#include <array>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <sstream>

namespace vkpt_hip {
namespace {

namespace fs = std::filesystem;

// CRC-32 (ISO 3309, reflected polynomial 0xEDB88320) as PNG chunks use it.
uint32_t crc32(uint32_t crc, const uint8_t* p, size_t n) {
  static const std::array<uint32_t, 256> table = [] {
    std::array<uint32_t, 256> t{};
    for (uint32_t i = 0; i < 256; ++i) {
      uint32_t c = i;
      for (int k = 0; k < 8; ++k) c = (c & 1u) ? 0xEDB88320u ^ (c >> 1) : c >> 1;
      t[i] = c;
    }
    return t;
  }();
  for (size_t i = 0; i < n; ++i) crc = table[(crc ^ p[i]) & 0xFFu] ^ (crc >> 8);
  return crc;
}

// Adler-32 (RFC 1950) with the modulo deferred over 5552-byte runs.
uint32_t adler32(const std::vector<uint8_t>& d) {
  uint32_t a = 1, b = 0;
  for (size_t i = 0; i < d.size();) {
    const size_t end = std::min(d.size(), i + 5552);
    for (; i < end; ++i) {
      a += d[i];
      b += a;
    }
    a %= 65521u;
    b %= 65521u;
  }
  return (b << 16) | a;
}

void put_be32(std::vector<uint8_t>& o, uint32_t v) {
  for (int s = 24; s >= 0; s -= 8) o.push_back(static_cast<uint8_t>(v >> s));
}

void put_le(std::vector<uint8_t>& o, uint64_t v, int bytes) {
  for (int i = 0; i < bytes; ++i) o.push_back(static_cast<uint8_t>(v >> (8 * i)));
}

void png_chunk(std::vector<uint8_t>& o, const char type[4], const std::vector<uint8_t>& data) {
  put_be32(o, static_cast<uint32_t>(data.size()));
  const size_t start = o.size();
  o.insert(o.end(), type, type + 4);
  o.insert(o.end(), data.begin(), data.end());
  put_be32(o, crc32(0xFFFFFFFFu, o.data() + start, o.size() - start) ^ 0xFFFFFFFFu);
}

bool write_bytes(const std::string& path, const std::vector<uint8_t>& bytes, std::string& error) {
  std::ofstream f(path, std::ios::binary | std::ios::trunc);
  if (!f) {
    error = "failed to open " + path;
    return false;
  }
  f.write(reinterpret_cast<const char*>(bytes.data()), static_cast<std::streamsize>(bytes.size()));
  if (!f) error = "failed to write " + path;
  return static_cast<bool>(f);
}

void exr_attr(std::vector<uint8_t>& o, const char* name, const char* type, const std::vector<uint8_t>& value) {
  o.insert(o.end(), name, name + std::strlen(name) + 1);
  o.insert(o.end(), type, type + std::strlen(type) + 1);
  put_le(o, value.size(), 4);
  o.insert(o.end(), value.begin(), value.end());
}

std::vector<uint8_t> le_ints(std::initializer_list<int32_t> v) {
  std::vector<uint8_t> o;
  for (int32_t x : v) put_le(o, static_cast<uint32_t>(x), 4);
  return o;
}

std::vector<uint8_t> le_floats(std::initializer_list<float> v) {
  std::vector<uint8_t> o;
  for (float x : v) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    put_le(o, u, 4);
  }
  return o;
}

// The value of "key": "..." in a flat JSON text (the fields results.json names are plain strings).
std::string json_string_field(const std::string& text, const std::string& key) {
  const std::string pat = "\"" + key + "\"";
  size_t p = text.find(pat);
  if (p == std::string::npos) return {};
  p = text.find(':', p + pat.size());
  if (p == std::string::npos) return {};
  p = text.find('"', p + 1);
  if (p == std::string::npos) return {};
  std::string out;
  for (size_t i = p + 1; i < text.size(); ++i) {
    if (text[i] == '\\' && i + 1 < text.size()) {
      out += text[++i];
      continue;
    }
    if (text[i] == '"') return out;
    out += text[i];
  }
  return {};
}

std::string read_text(const fs::path& p) {
  std::ifstream f(p, std::ios::binary);
  std::ostringstream s;
  s << f.rdbuf();
  return s.str();
}

// resolve_artifact_path (BenchmarkSchema.cpp:273-286).
fs::path resolve_artifact_path(const fs::path& dir, const std::string& value) {
  if (value.empty()) return {};
  const fs::path p(value);
  if (p.is_absolute()) return p;
  std::error_code ec;
  if (fs::exists(p, ec) && !ec) return p;
  return dir / p.filename();
}

std::string json_array(const std::vector<std::string>& v) {
  std::string o = "[";
  for (size_t i = 0; i < v.size(); ++i) {
    if (i) o += ",";
    o += "\"";
    for (char c : v[i]) {
      if (c == '"' || c == '\\') o += '\\';
      o += c;
    }
    o += "\"";
  }
  return o + "]";
}

}  // namespace

uint64_t fnv1a64(const void* data, size_t size) {
  uint64_t h = 1469598103934665603ull;
  const uint8_t* p = static_cast<const uint8_t*>(data);
  for (size_t i = 0; i < size; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

std::string hex64(uint64_t v) {
  char buf[17];
  std::snprintf(buf, sizeof(buf), "%016llx", static_cast<unsigned long long>(v));
  return buf;
}

bool save_png(const std::string& path, const FilmLdr& img, std::string& error) {
  if (img.width == 0 || img.height == 0 || img.rgba8.size() < 4ull * img.width * img.height) {
    error = "empty image";
    return false;
  }
  const size_t row = 4ull * img.width;
  // Scanlines with filter type 0, then a zlib stream of stored deflate blocks (RFC 1950 / 1951).
  std::vector<uint8_t> raw;
  raw.reserve((row + 1) * img.height);
  for (uint32_t y = 0; y < img.height; ++y) {
    raw.push_back(0);
    raw.insert(raw.end(), img.rgba8.begin() + static_cast<ptrdiff_t>(row * y),
               img.rgba8.begin() + static_cast<ptrdiff_t>(row * (y + 1)));
  }
  std::vector<uint8_t> z = {0x78, 0x01};
  for (size_t off = 0; off < raw.size();) {
    const size_t len = std::min<size_t>(65535, raw.size() - off);
    z.push_back(off + len >= raw.size() ? 1 : 0);
    put_le(z, len, 2);
    put_le(z, ~len & 0xFFFFu, 2);
    z.insert(z.end(), raw.begin() + static_cast<ptrdiff_t>(off), raw.begin() + static_cast<ptrdiff_t>(off + len));
    off += len;
  }
  put_be32(z, adler32(raw));
  std::vector<uint8_t> out = {0x89, 'P', 'N', 'G', '\r', '\n', 0x1A, '\n'};
  std::vector<uint8_t> ihdr;
  put_be32(ihdr, img.width);
  put_be32(ihdr, img.height);
  ihdr.insert(ihdr.end(), {8, 6, 0, 0, 0});  // 8-bit RGBA, deflate, adaptive filters, no interlace
  png_chunk(out, "IHDR", ihdr);
  png_chunk(out, "IDAT", z);
  png_chunk(out, "IEND", {});
  return write_bytes(path, out, error);
}

bool save_exr(const std::string& path, const FilmHdr& img, std::string& error) {
  if (img.width == 0 || img.height == 0 || img.rgb.size() < 3ull * img.width * img.height) {
    error = "empty image";
    return false;
  }
  const int32_t xmax = static_cast<int32_t>(img.width) - 1, ymax = static_cast<int32_t>(img.height) - 1;
  std::vector<uint8_t> o = {0x76, 0x2F, 0x31, 0x01, 2, 0, 0, 0};  // magic 20000630, version 2, scanline
  std::vector<uint8_t> ch;
  for (const char* name : {"B", "G", "R"}) {  // alphabetical, as the format requires
    ch.insert(ch.end(), name, name + 2);
    put_le(ch, 2, 4);            // FLOAT
    ch.insert(ch.end(), {0, 0, 0, 0});  // pLinear, reserved
    put_le(ch, 1, 4);            // x sampling
    put_le(ch, 1, 4);            // y sampling
  }
  ch.push_back(0);
  exr_attr(o, "channels", "chlist", ch);
  exr_attr(o, "compression", "compression", {0});
  exr_attr(o, "dataWindow", "box2i", le_ints({0, 0, xmax, ymax}));
  exr_attr(o, "displayWindow", "box2i", le_ints({0, 0, xmax, ymax}));
  exr_attr(o, "lineOrder", "lineOrder", {0});
  exr_attr(o, "pixelAspectRatio", "float", le_floats({1.0f}));
  exr_attr(o, "screenWindowCenter", "v2f", le_floats({0.0f, 0.0f}));
  exr_attr(o, "screenWindowWidth", "float", le_floats({1.0f}));
  o.push_back(0);  // end of header
  const uint64_t block = 8ull + 12ull * img.width;  // y, size, then 3 channels of width floats
  const uint64_t first = o.size() + 8ull * img.height;
  for (uint32_t y = 0; y < img.height; ++y) put_le(o, first + block * y, 8);
  for (uint32_t y = 0; y < img.height; ++y) {
    put_le(o, y, 4);
    put_le(o, 12ull * img.width, 4);
    for (int c = 2; c >= 0; --c)  // B, G, R
      for (uint32_t x = 0; x < img.width; ++x) {
        uint32_t u;
        std::memcpy(&u, &img.rgb[3ull * (static_cast<size_t>(y) * img.width + x) + c], 4);
        put_le(o, u, 4);
      }
  }
  return write_bytes(path, o, error);
}

ArtifactContract default_artifact_contract() {
  return {{"results.json", "results.csv", "metadata.json", "scene_snapshot.json", "shader_manifest.json",
           "asset_manifest.json", "beauty.png", "beauty.exr", "logs.jsonl", "profiler_trace.json"},
          {"reference.exr", "diff_heatmap.png"}};
}

ArtifactValidation validate_artifacts(const std::string& directory) {
  ArtifactValidation v;
  v.directory = directory;
  const fs::path dir(directory);
  std::error_code ec;
  if (directory.empty() || !fs::is_directory(dir, ec) || ec) {
    v.missing.push_back("artifact_directory");
    return v;
  }
  const ArtifactContract c = default_artifact_contract();
  auto check = [&](const std::string& name, bool required) {
    const fs::path p = dir / name;
    std::error_code e;
    if (!fs::exists(p, e) || e) {
      if (required) v.missing.push_back(name);
      return;
    }
    if (!fs::is_regular_file(p, e) || e) {
      v.invalid.push_back(name + ": not a regular file");
      return;
    }
    v.present.push_back(name);
    if (fs::file_size(p, e) == 0 && !e) v.empty.push_back(name);
  };
  for (const auto& n : c.required) check(n, true);
  for (const auto& n : c.optional) check(n, false);
  const std::string results = read_text(dir / "results.json");
  if (results.find('{') == std::string::npos) {
    v.invalid.push_back("results.json: failed to parse benchmark result schema");
  } else {
    for (const char* key : {"beauty_png", "beauty_exr", "logs_jsonl", "profiler_trace_json"}) {
      const std::string value = json_string_field(results, key);
      if (value.empty()) {
        v.invalid.push_back(std::string("results.json: missing ") + key);
        continue;
      }
      const fs::path r = resolve_artifact_path(dir, value);
      if (r.empty() || !fs::exists(r, ec) || ec)
        v.invalid.push_back(std::string("results.json: ") + key + " does not resolve to an artifact file");
    }
  }
  if (read_text(dir / "metadata.json").find("\"benchmark_capabilities\"") == std::string::npos)
    v.invalid.push_back("metadata.json: missing benchmark_capabilities");
  if (read_text(dir / "profiler_trace.json").find("\"events\"") == std::string::npos)
    v.invalid.push_back("profiler_trace.json: missing events array");
  v.ok = v.missing.empty() && v.empty.empty() && v.invalid.empty();
  return v;
}

std::string ArtifactValidation::to_json() const {
  std::string dir_escaped;
  for (char ch : directory) {
    if (ch == '"' || ch == '\\') dir_escaped += '\\';
    dir_escaped += ch;
  }
  return std::string("{\"ok\":") + (ok ? "true" : "false") + ",\"artifact_directory\":\"" + dir_escaped +
         "\",\"present_files\":" + json_array(present) + ",\"missing_files\":" + json_array(missing) +
         ",\"empty_files\":" + json_array(empty) + ",\"invalid_files\":" + json_array(invalid) + "}";
}

std::string resolve_settings_json(const PtFilmResolve& s) {
  static const char* tone[] = {"linear", "reinhard", "filmic_approx", "aces_approx"};
  std::ostringstream o;
  o << "{\"exposure\":" << s.exposure << ",\"white_balance_kelvin\":" << s.white_balance_kelvin
    << ",\"tone_map\":\"" << (s.tone_map < 4 ? tone[s.tone_map] : "linear") << "\",\"output_transform\":\""
    << (s.output_transform == 1 ? "linear" : "gamma") << "\",\"gamma\":" << s.gamma
    << ",\"clamp_output\":" << (s.clamp_output ? "true" : "false") << "}";
  return o.str();
}

}  // namespace vkpt_hip
