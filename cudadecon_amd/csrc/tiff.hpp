// Baseline TIFF I/O for OTF files and image stacks (no libtiff dependency).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace cdn {

struct TiffImage {
  int width = 0, height = 0, depth = 0;  // depth = number of pages
  int bits = 0;                          // bits per sample of the file (8, 16, 32)
  int format = 1;                        // SampleFormat: 1 unsigned, 2 signed, 3 IEEE float
  std::vector<float> data;               // width * height * depth, x fastest
};

// Reads every page of an uncompressed, single-sample, classic TIFF.  Throws
// std::runtime_error on anything else, and on malformed files: truncated
// fields or strips, empty fields, IFD chains that loop, pages of different
// sizes, and image data that would exceed the file's own size.
TiffImage tiff_read(const std::string& path);

// Optional metadata of written stacks: an ImageDescription string (ImageJ
// reads "spacing=", "unit=", "slices=" from it) and the x/y pixel size as
// XResolution / YResolution (pixels per unit).
struct TiffMeta {
  std::string description;
  float pixel_x = 0.f, pixel_y = 0.f;  // 0: no resolution tags
};

// Little-endian classic TIFF, one page per z-plane (x fastest).  Throws if
// the file would exceed the 4 GiB a classic TIFF can address.
void tiff_write_f32(const std::string& path, const float* data, int width, int height, int depth,
                    const TiffMeta* meta = nullptr);
void tiff_write_u16(const std::string& path, const uint16_t* data, int width, int height, int depth,
                    const TiffMeta* meta = nullptr);

}  // namespace cdn
