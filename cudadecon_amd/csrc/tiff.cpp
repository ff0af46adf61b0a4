// Minimal baseline TIFF I/O for OTF files (no libtiff dependency).
//
// Reads uncompressed, single-sample, chunky images with 8/16/32-bit unsigned,
// signed or IEEE-float samples in strips, either byte order, one page per
// z-plane; returns float32 data with x fastest.  This covers the OTF files
// the reference reads through CImg (src/RL-Biggs-Andrews.cpp:567-573): RA
// OTFs are width 2*nzotf x height nr x 1 page, 3-D OTFs (nx/2+1)*2 x ny x nz
// (src/radialft-nonSIM.cpp:182-188).  Writes little-endian float32 pages.
#include "tiff.hpp"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <stdexcept>
#include <vector>

namespace cdn {

namespace {

struct Reader {
  std::vector<uint8_t> d;
  bool le = true;
  uint16_t u16(size_t o) const {
    if (o + 2 > d.size()) throw std::runtime_error("TIFF: truncated");
    return le ? (uint16_t)(d[o] | (d[o + 1] << 8)) : (uint16_t)((d[o] << 8) | d[o + 1]);
  }
  uint32_t u32(size_t o) const {
    if (o + 4 > d.size()) throw std::runtime_error("TIFF: truncated");
    return le ? (uint32_t)d[o] | ((uint32_t)d[o + 1] << 8) | ((uint32_t)d[o + 2] << 16) | ((uint32_t)d[o + 3] << 24)
              : ((uint32_t)d[o] << 24) | ((uint32_t)d[o + 1] << 16) | ((uint32_t)d[o + 2] << 8) | (uint32_t)d[o + 3];
  }
  // value(s) of an IFD entry as unsigned integers
  std::vector<uint32_t> values(size_t entry) const {
    const uint16_t type = u16(entry + 2);
    const uint32_t count = u32(entry + 4);
    const size_t sz = (type == 3) ? 2 : (type == 4) ? 4 : (type == 1) ? 1 : 0;
    if (sz == 0) throw std::runtime_error("TIFF: unsupported field type");
    size_t off = entry + 8;
    if (sz * count > 4) off = u32(entry + 8);
    std::vector<uint32_t> v(count);
    for (uint32_t i = 0; i < count; ++i) {
      const size_t o = off + i * sz;
      v[i] = (sz == 2) ? u16(o) : (sz == 4) ? u32(o) : d.at(o);
    }
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
  if (R.d.size() < 8) throw std::runtime_error("TIFF: file too short");
  if (R.d[0] == 'I' && R.d[1] == 'I')
    R.le = true;
  else if (R.d[0] == 'M' && R.d[1] == 'M')
    R.le = false;
  else
    throw std::runtime_error("TIFF: bad byte-order mark");
  if (R.u16(2) != 42) throw std::runtime_error("TIFF: not a classic TIFF");
  TiffImage img;
  uint32_t ifd = R.u32(4);
  int pages = 0;
  while (ifd != 0) {
    const uint16_t n = R.u16(ifd);
    uint32_t w = 0, h = 0, bps = 1, comp = 1, spp = 1, fmt = 1, planar = 1;
    std::vector<uint32_t> offs, counts;
    for (uint16_t e = 0; e < n; ++e) {
      const size_t ent = ifd + 2 + 12 * (size_t)e;
      const uint16_t tag = R.u16(ent);
      switch (tag) {
        case 256: w = R.values(ent)[0]; break;
        case 257: h = R.values(ent)[0]; break;
        case 258: bps = R.values(ent)[0]; break;
        case 259: comp = R.values(ent)[0]; break;
        case 273: offs = R.values(ent); break;
        case 277: spp = R.values(ent)[0]; break;
        case 279: counts = R.values(ent); break;
        case 284: planar = R.values(ent)[0]; break;
        case 339: fmt = R.values(ent)[0]; break;
        case 322: throw std::runtime_error("TIFF: tiled images are not supported");
        default: break;
      }
    }
    if (comp != 1) throw std::runtime_error("TIFF: compressed images are not supported");
    if (spp != 1 || planar != 1) throw std::runtime_error("TIFF: only single-sample images are supported");
    if (!(bps == 8 || bps == 16 || bps == 32)) throw std::runtime_error("TIFF: unsupported bits per sample");
    if (pages == 0) {
      img.width = (int)w;
      img.height = (int)h;
    } else if ((int)w != img.width || (int)h != img.height) {
      throw std::runtime_error("TIFF: pages of different sizes");
    }
    const size_t bytes = (size_t)bps / 8;
    const size_t npix = (size_t)w * h;
    const size_t base = img.data.size();
    img.data.resize(base + npix);
    size_t pix = 0;
    for (size_t s = 0; s < offs.size() && pix < npix; ++s) {
      const size_t cnt = s < counts.size() ? counts[s] : npix * bytes;
      const size_t n_s = std::min(cnt / bytes, npix - pix);
      size_t o = offs[s];
      if (o + n_s * bytes > R.d.size()) throw std::runtime_error("TIFF: strip outside file");
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
    if (pix != npix) throw std::runtime_error("TIFF: missing image data");
    ++pages;
    ifd = R.u32(ifd + 2 + 12 * (size_t)n);
  }
  img.depth = pages;
  if (pages == 0) throw std::runtime_error("TIFF: no image");
  return img;
}

void tiff_write_f32(const std::string& path, const float* data, int width, int height, int depth) {
  std::ofstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot write TIFF file '" + path + "'");
  auto put16 = [&](uint16_t v) { f.write((const char*)&v, 2); };
  auto put32 = [&](uint32_t v) { f.write((const char*)&v, 4); };
  f.write("II", 2);
  put16(42);
  const uint32_t page_bytes = (uint32_t)width * height * 4;
  const int NE = 10;
  const uint32_t ifd_bytes = 2 + 12 * NE + 4;
  uint32_t pos = 8;
  put32(pos);
  for (int z = 0; z < depth; ++z) {
    const uint32_t data_off = pos + ifd_bytes;
    const uint32_t next = (z + 1 < depth) ? data_off + page_bytes : 0;
    put16(NE);
    auto ent = [&](uint16_t tag, uint16_t type, uint32_t val) {
      put16(tag);
      put16(type);
      put32(1);
      if (type == 3) {
        put16((uint16_t)val);
        put16(0);
      } else {
        put32(val);
      }
    };
    ent(256, 4, (uint32_t)width);
    ent(257, 4, (uint32_t)height);
    ent(258, 3, 32);
    ent(259, 3, 1);
    ent(262, 3, 1);
    ent(273, 4, data_off);
    ent(277, 3, 1);
    ent(278, 4, (uint32_t)height);
    ent(279, 4, page_bytes);
    ent(339, 3, 3);
    put32(next);
    f.write((const char*)(data + (size_t)z * width * height), page_bytes);
    pos = data_off + page_bytes;
  }
  if (!f) throw std::runtime_error("TIFF: write failed for '" + path + "'");
}

}  // namespace cdn
