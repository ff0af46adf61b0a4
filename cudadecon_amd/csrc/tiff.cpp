// Baseline TIFF I/O (no libtiff dependency).
//
// Reads uncompressed, single-sample, chunky images with 8/16/32-bit unsigned,
// signed or IEEE-float samples in strips, either byte order, one page per
// z-plane; returns float32 data with x fastest.  This covers the OTF files
// the reference reads through CImg (src/RL-Biggs-Andrews.cpp:567-573): RA
// OTFs are width 2*nzotf x height nr x 1 page, 3-D OTFs (nx/2+1)*2 x ny x nz
// (src/radialft-nonSIM.cpp:182-188); and the 16-bit stacks the command-line
// caller reads (src/linearDecon.cpp:19-35).  Writes little-endian float32 or
// uint16 pages with an optional ImageJ-style description
// (src/linearDecon.cpp:41-51).
//
// Every value read from the file is checked before it is used: field counts
// against the file size (no allocation larger than the file), zero counts,
// IFD offsets already visited (a loop), and the total decoded image size
// against the file size.
#include "tiff.hpp"

#include <algorithm>
#include <cstring>
#include <fstream>
#include <limits>
#include <set>
#include <stdexcept>

namespace cdn {

namespace {

[[noreturn]] void bad(const std::string& what) { throw std::runtime_error("TIFF: " + what); }

struct Reader {
  std::vector<uint8_t> d;
  bool le = true;
  void need(size_t o, size_t n) const {
    if (o > d.size() || n > d.size() - o) bad("truncated file");
  }
  uint16_t u16(size_t o) const {
    need(o, 2);
    return le ? (uint16_t)(d[o] | (d[o + 1] << 8)) : (uint16_t)((d[o] << 8) | d[o + 1]);
  }
  uint32_t u32(size_t o) const {
    need(o, 4);
    return le ? (uint32_t)d[o] | ((uint32_t)d[o + 1] << 8) | ((uint32_t)d[o + 2] << 16) | ((uint32_t)d[o + 3] << 24)
              : ((uint32_t)d[o] << 24) | ((uint32_t)d[o + 1] << 16) | ((uint32_t)d[o + 2] << 8) | (uint32_t)d[o + 3];
  }
  // location and element size of an IFD entry's values, checked to lie in
  // the file; count >= 1
  void field(size_t entry, size_t* off, size_t* sz, uint32_t* count) const {
    const uint16_t type = u16(entry + 2);
    *count = u32(entry + 4);
    *sz = (type == 3) ? 2 : (type == 4) ? 4 : (type == 1) ? 1 : 0;
    if (*sz == 0) bad("unsupported field type " + std::to_string(type));
    if (*count == 0) bad("empty field (tag " + std::to_string(u16(entry)) + ")");
    const size_t bytes = *sz * (size_t)*count;
    *off = bytes > 4 ? (size_t)u32(entry + 8) : entry + 8;
    need(*off, bytes);
  }
  uint32_t at(size_t off, size_t sz, size_t i) const {
    const size_t o = off + i * sz;
    return (sz == 2) ? u16(o) : (sz == 4) ? u32(o) : d[o];
  }
  // the first value of a scalar tag
  uint32_t first(size_t entry) const {
    size_t off, sz;
    uint32_t count;
    field(entry, &off, &sz, &count);
    return at(off, sz, 0);
  }
  // every value of an array tag (its size is bounded by the file's)
  std::vector<uint32_t> values(size_t entry) const {
    size_t off, sz;
    uint32_t count;
    field(entry, &off, &sz, &count);
    std::vector<uint32_t> v(count);
    for (uint32_t i = 0; i < count; ++i) v[i] = at(off, sz, i);
    return v;
  }
};

}  // namespace

TiffImage tiff_read(const std::string& path) {
  Reader R;
  {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("cannot open TIFF file '" + path + "'");
    R.d.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  }
  if (R.d.size() < 8) bad("file too short");
  if (R.d[0] == 'I' && R.d[1] == 'I')
    R.le = true;
  else if (R.d[0] == 'M' && R.d[1] == 'M')
    R.le = false;
  else
    bad("bad byte-order mark");
  if (R.u16(2) != 42) bad("not a classic TIFF");
  TiffImage img;
  uint32_t ifd = R.u32(4);
  std::set<uint32_t> seen;
  size_t decoded = 0;  // image bytes of the pages read so far
  int pages = 0;
  while (ifd != 0) {
    if (!seen.insert(ifd).second) bad("IFD chain loops back to offset " + std::to_string(ifd));
    const uint16_t n = R.u16(ifd);
    R.need((size_t)ifd + 2, 12 * (size_t)n + 4);
    uint32_t w = 0, h = 0, bps = 1, comp = 1, spp = 1, fmt = 1, planar = 1;
    std::vector<uint32_t> offs, counts;
    for (uint16_t e = 0; e < n; ++e) {
      const size_t ent = ifd + 2 + 12 * (size_t)e;
      switch (R.u16(ent)) {
        case 256: w = R.first(ent); break;
        case 257: h = R.first(ent); break;
        case 258: bps = R.first(ent); break;
        case 259: comp = R.first(ent); break;
        case 273: offs = R.values(ent); break;
        case 277: spp = R.first(ent); break;
        case 279: counts = R.values(ent); break;
        case 284: planar = R.first(ent); break;
        case 339: fmt = R.first(ent); break;
        case 322: bad("tiled images are not supported");
        default: break;
      }
    }
    if (comp != 1) bad("compressed images are not supported");
    if (spp != 1 || planar != 1) bad("only single-sample images are supported");
    if (!(bps == 8 || bps == 16 || bps == 32)) bad("unsupported bits per sample");
    if (!(fmt == 1 || fmt == 2 || fmt == 3) || (fmt == 3 && bps != 32)) bad("unsupported sample format");
    if (w == 0 || h == 0) bad("empty page");
    if (offs.empty()) bad("page without strip offsets");
    const size_t bytes = (size_t)bps / 8;
    const size_t npix = (size_t)w * h;
    // an uncompressed page stores all its samples; the pages together
    // cannot hold more image bytes than the file (this also bounds pages
    // that point at the same strips)
    if (npix > R.d.size() / bytes || decoded + npix * bytes > R.d.size()) bad("image data larger than the file");
    decoded += npix * bytes;
    if (pages == 0) {
      img.width = (int)w;
      img.height = (int)h;
      img.bits = (int)bps;
      img.format = (int)fmt;
    } else if ((int)w != img.width || (int)h != img.height || (int)bps != img.bits || (int)fmt != img.format) {
      bad("pages of different sizes or sample types");
    }
    const size_t base = img.data.size();
    img.data.resize(base + npix);
    size_t pix = 0;
    for (size_t s = 0; s < offs.size() && pix < npix; ++s) {
      const size_t cnt = s < counts.size() ? counts[s] : npix * bytes;
      const size_t n_s = std::min(cnt / bytes, npix - pix);
      size_t o = offs[s];
      R.need(o, n_s * bytes);
      for (size_t i = 0; i < n_s; ++i, o += bytes) {
        float v;
        if (bps == 32) {
          const uint32_t u = R.u32(o);
          if (fmt == 3) {
            std::memcpy(&v, &u, 4);
          } else if (fmt == 2) {
            v = (float)(int32_t)u;
          } else {
            v = (float)u;
          }
        } else if (bps == 16) {
          const uint16_t u = R.u16(o);
          v = (fmt == 2) ? (float)(int16_t)u : (float)u;
        } else {
          v = (fmt == 2) ? (float)(int8_t)R.d[o] : (float)R.d[o];
        }
        img.data[base + pix++] = v;
      }
    }
    if (pix != npix) bad("missing image data");
    ++pages;
    ifd = R.u32(ifd + 2 + 12 * (size_t)n);
  }
  img.depth = pages;
  if (pages == 0) bad("no image");
  return img;
}

namespace {

void write_pages(const std::string& path, const void* data, int bits, int fmt, int width, int height, int depth,
                 const TiffMeta* meta) {
  if (width <= 0 || height <= 0 || depth <= 0) throw std::runtime_error("TIFF: non-positive image size");
  const uint64_t page_bytes = (uint64_t)width * height * (bits / 8);
  const bool res = meta && meta->pixel_x > 0.f && meta->pixel_y > 0.f;
  const std::string desc = meta ? meta->description : std::string();
  const int NE = 10 + (desc.empty() ? 0 : 1) + (res ? 3 : 0);
  const uint64_t ifd_bytes = 2 + 12 * (uint64_t)NE + 4;
  // out-of-line values after each IFD: description (+NUL, padded to even), 2 rationals
  const uint64_t desc_bytes = desc.empty() ? 0 : (desc.size() + 2) / 2 * 2;
  const uint64_t extra = desc_bytes + (res ? 16 : 0);
  const uint64_t total = 8 + (uint64_t)depth * (ifd_bytes + extra + page_bytes);
  if (total > std::numeric_limits<uint32_t>::max())
    throw std::runtime_error("TIFF: '" + path + "' would exceed the 4 GiB of a classic TIFF");
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write TIFF file '" + path + "'");
  auto put16 = [&](uint16_t v) { f.write((const char*)&v, 2); };
  auto put32 = [&](uint32_t v) { f.write((const char*)&v, 4); };
  f.write("II", 2);
  put16(42);
  uint32_t pos = 8;
  put32(pos);
  // pixels per unit as a rational: 1e6 / (size * 1e6)
  auto rational = [](float size) { return (uint32_t)std::max(1.0, (double)size * 1e6 + 0.5); };
  for (int z = 0; z < depth; ++z) {
    const uint32_t extra_off = pos + (uint32_t)ifd_bytes;
    const uint32_t data_off = extra_off + (uint32_t)extra;
    const uint32_t next = (z + 1 < depth) ? data_off + (uint32_t)page_bytes : 0;
    put16((uint16_t)NE);
    auto ent = [&](uint16_t tag, uint16_t type, uint32_t count, uint32_t val) {
      put16(tag);
      put16(type);
      put32(count);
      if (type == 3 && count == 1) {
        put16((uint16_t)val);
        put16(0);
      } else {
        put32(val);
      }
    };
    // tags in ascending order
    ent(256, 4, 1, (uint32_t)width);
    ent(257, 4, 1, (uint32_t)height);
    ent(258, 3, 1, (uint32_t)bits);
    ent(259, 3, 1, 1);
    ent(262, 3, 1, 1);
    if (!desc.empty()) ent(270, 2, (uint32_t)desc.size() + 1, extra_off);
    ent(273, 4, 1, data_off);
    ent(277, 3, 1, 1);
    ent(278, 4, 1, (uint32_t)height);
    ent(279, 4, 1, (uint32_t)page_bytes);
    if (res) {
      ent(282, 5, 1, extra_off + (uint32_t)desc_bytes);
      ent(283, 5, 1, extra_off + (uint32_t)desc_bytes + 8);
      ent(296, 3, 1, 1);
    }
    ent(339, 3, 1, (uint32_t)fmt);
    put32(next);
    if (!desc.empty()) {
      f.write(desc.c_str(), (std::streamsize)desc.size() + 1);
      if ((desc.size() + 1) % 2) f.put('\0');
    }
    if (res) {
      put32(1000000);
      put32(rational(meta->pixel_x));
      put32(1000000);
      put32(rational(meta->pixel_y));
    }
    f.write((const char*)data + (size_t)z * page_bytes, (std::streamsize)page_bytes);
    pos = data_off + (uint32_t)page_bytes;
  }
  if (!f) throw std::runtime_error("TIFF: write failed for '" + path + "'");
}

}  // namespace

void tiff_write_f32(const std::string& path, const float* data, int width, int height, int depth,
                    const TiffMeta* meta) {
  write_pages(path, data, 32, 3, width, height, depth, meta);
}

void tiff_write_u16(const std::string& path, const uint16_t* data, int width, int height, int depth,
                    const TiffMeta* meta) {
  write_pages(path, data, 16, 1, width, height, depth, meta);
}

}  // namespace cdn
