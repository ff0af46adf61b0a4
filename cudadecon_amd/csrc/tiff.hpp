// Minimal float32 TIFF reader/writer for OTF files.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cdn {

struct TiffImage {
  int width = 0, height = 0, depth = 0;  // depth = number of pages
  std::vector<float> data;               // width * height * depth, x fastest
};

TiffImage tiff_read(const std::string& path);
void tiff_write_f32(const std::string& path, const float* data, int width, int height, int depth);

}  // namespace cdn
