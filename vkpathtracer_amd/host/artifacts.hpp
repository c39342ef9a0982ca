// artifacts.hpp — image files and the benchmark artifact contract of the C++ host.
//
// save_png mirrors SavePngCompat (ImageIo.cpp:127-189): an RGBA8 PNG whose zlib stream uses stored
// (uncompressed) deflate blocks, so the bytes depend only on the pixels. save_exr writes the HDR image
// as a real OpenEXR 2 file (single-part scanline, uncompressed, 32-bit float B, G, R channels) where the
// reference writes a text placeholder under that name (SaveExrCompat, ImageIo.cpp:191-217).
// validate_artifacts restates ValidateBenchmarkArtifactsOnDisk (BenchmarkSchema.cpp:780-846) over the
// contract of DefaultBenchmarkArtifactContract (BenchmarkSchema.cpp:758-775).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "hip_path_tracer.hpp"

namespace vkpt_hip {

// Fnv1a64 / Hex64 (MaterialDescriptors.cpp:50-65): ptbench's image_hash is this hash of beauty.png.
uint64_t fnv1a64(const void* data, size_t size);
std::string hex64(uint64_t value);

bool save_png(const std::string& path, const FilmLdr& image, std::string& error);
bool save_exr(const std::string& path, const FilmHdr& image, std::string& error);

struct ArtifactContract {
  std::vector<std::string> required, optional;
};
ArtifactContract default_artifact_contract();

struct ArtifactValidation {
  bool ok = false;
  std::string directory;
  std::vector<std::string> present, missing, empty, invalid;
  std::string to_json() const;  // SerializeBenchmarkArtifactValidation's fields
};
ArtifactValidation validate_artifacts(const std::string& directory);

// SerializeFilmResolveSettings (FilmBuffer.cpp:334-358).
std::string resolve_settings_json(const PtFilmResolve& settings);

}  // namespace vkpt_hip
