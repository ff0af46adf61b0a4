// Host-side scalar logic of the RL path (declarations and reference lines in
// geometry.hpp).  The float/double rounding points follow the reference.
#include "geometry.hpp"

#include "host_common.hpp"
#include "tiff.hpp"

#include <cmath>

namespace cdn {

// widths beyond this are rejected (far above any FFT length the library runs)
static constexpr double kMaxWidth = 1u << 30;

bool not_good_dimension(unsigned n) {
  while (true) {
    if (n <= 1) return false;
    if (n == 2 || n == 3 || n == 5 || n == 7) return false;
    if (n % 2 == 0) { n /= 2; continue; }
    if (n % 3 == 0) { n /= 3; continue; }
    if (n % 5 == 0) { n /= 5; continue; }
    if (n % 7 == 0) { n /= 7; continue; }
    return true;
  }
}

unsigned find_optimal_dimension(unsigned n, int step) {
  if (n <= 1) return n;
  unsigned o = n;
  while (not_good_dimension(o)) o += step;
  return o;
}

void OtfTable::set_from_image(const float* img, int width, int height, int depth, float dr_psf, float dz_psf) {
  if (!img) fail("OTF image is null");
  if (width <= 0 || height <= 0 || depth <= 0) fail("OTF image has a non-positive size");
  if (depth > 1) {
    nxotf = width / 2;
    nyotf = height;
    nzotf = depth;
  } else {
    nxotf = height;
    nyotf = 1;
    nzotf = width / 2;
  }
  if (nxotf < 2 || nzotf < 1) fail("OTF table is too small");
  const size_t n = (size_t)2 * nxotf * nyotf * nzotf;
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(img[i])) fail("OTF table holds a NaN or Inf value (entry " + std::to_string(i) + ")");
  dkx = 1.0f / ((float)((nxotf - 1) * 2) * dr_psf);
  dky = (nyotf > 1) ? 1.0f / ((float)nyotf * dr_psf) : dkx;
  dkz = 1.0f / ((float)nzotf * dz_psf);
  data.assign(img, img + n);
}

OtfTable load_otf_file(const std::string& path, float dr_psf, float dz_psf) {
  const TiffImage img = tiff_read(path);
  OtfTable t;
  t.set_from_image(img.data.data(), img.width, img.height, img.depth, dr_psf, dz_psf);
  return t;
}

SessionGeometry session_geometry(int nx, int ny, int nz, float dr, float dz, float deskewAngle, float rotationAngle,
                                 int outputWidth, bool bSkewedDecon) {
  if (nx <= 0 || ny <= 0 || nz <= 0) fail("non-positive dimensions");
  if (!std::isfinite(dr) || !std::isfinite(dz) || !std::isfinite(deskewAngle) || !std::isfinite(rotationAngle))
    fail("non-finite pixel size or angle");
  if (!(dr > 0.f) || !(dz > 0.f)) fail("pixel sizes must be positive");
  if (outputWidth < 0) fail("negative output width");
  SessionGeometry S;
  S.output_ny = find_optimal_dimension(ny);
  S.output_nz = find_optimal_dimension(nz);
  const bool has_deskew = std::fabs(deskewAngle) > 0.0;
  // the CLI also optimises the width for skewed decon
  // (src/linearDecon.cpp:643); the library path leaves it unoptimised and
  // then plans an FFT of the deskewed width over the raw width
  S.output_nx = (has_deskew && !bSkewedDecon) ? (unsigned)nx : find_optimal_dimension(nx);
  S.crop = (S.output_nx != (unsigned)nx) || (S.output_ny != (unsigned)ny) || (S.output_nz != (unsigned)nz);
  S.deskewFactor = 0.;
  S.deskewedXdim = S.output_nx;
  if (has_deskew) {
    if (deskewAngle < 0) deskewAngle += 180.;
    S.deskewFactor = std::cos(deskewAngle * M_PI / 180.) * dz / dr;
    if (outputWidth == 0) {
      const double w = S.deskewedXdim +
                       std::floor(S.output_nz * dz * std::fabs(std::cos(deskewAngle * M_PI / 180.)) / dr);
      if (!(w < kMaxWidth)) fail("deskewed width out of range");
      S.deskewedXdim = (unsigned)w;
    } else {
      S.deskewedXdim = (unsigned)outputWidth;
    }
    S.deskewedXdim = find_optimal_dimension(S.deskewedXdim);
    dz *= std::fabs(std::sin(deskewAngle * M_PI / 180.));
  }
  if (std::fabs(rotationAngle) > 0.0) {
    rotationAngle *= M_PI / 180;
    const float stretch = dr / dz;
    S.rot_on = true;
    S.rot[0] = std::cos(rotationAngle) * stretch;
    S.rot[1] = std::sin(rotationAngle) * stretch;
    S.rot[2] = -std::sin(rotationAngle);
    S.rot[3] = std::cos(rotationAngle);
    if (S.rot[1] == 0.f || S.rot[0] == 0.f) fail("degenerate rotation angle");
    int nxo, nzo;
    rotated_dims(S.rot, S.deskewedXdim, S.output_nz, &nxo, &nzo);
    if (nxo <= 0 || nzo <= 0) fail("rotation gives an empty result");
  }
  S.dr = dr;
  S.dz = dz;
  S.fft_nx = bSkewedDecon ? S.output_nx : S.deskewedXdim;
  return S;
}

void rotated_dims(const float* p, unsigned nx, unsigned nz, int* nx_o, int* nz_o) {
  // sizes outside int (a near-degenerate matrix) come back as 0: no result
  auto to_int = [](float v) { return (v >= 0.f && v < 2147483520.f) ? (int)v : 0; };
  *nz_o = to_int((float)nz * p[3] / p[0]);
  *nx_o = to_int((float)nx * p[3] + (float)nz * p[2] * p[2] / std::fabs(p[1]));
}

void session_kscales(const SessionGeometry& g, const OtfTable& t, unsigned fft_nx, float ks[3]) {
  const float dkx = (float)(1.0 / (g.dr * (float)fft_nx));
  const float dky = (float)(1.0 / (g.dr * (float)g.output_ny));
  const float dkz = (float)(1.0 / (g.dz * (float)g.output_nz));
  ks[0] = dkx / t.dkx;
  ks[1] = dky / t.dky;
  ks[2] = dkz / t.dkz;
}

unsigned deskew_interface_width(int nx, float dz, float dr, float deskewAngle, int outputWidth, unsigned output_nz) {
  if (nx <= 0 || outputWidth < 0) return 0;
  unsigned xdim = (unsigned)nx;
  if (std::fabs(deskewAngle) > 0.0) {
    if (deskewAngle < 0) deskewAngle += 180.;
    if (outputWidth == 0) {
      const double w = xdim + std::floor(output_nz * dz * std::fabs(std::cos(deskewAngle * M_PI / 180.)) / dr) / 4.;
      xdim = (w >= 0.0 && w < kMaxWidth) ? (unsigned)w : 0;  // 0: no valid width (NaN, infinite, negative)
    } else {
      xdim = (unsigned)outputWidth;
    }
  }
  return xdim;
}

}  // namespace cdn
