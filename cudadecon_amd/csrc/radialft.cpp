// libradialft.so: makeOTF, the OTF of a measured PSF stack, computed on the
// host (the steps of src/radialft_interface.cpp:95-518, restated here).
//
//   locate_bead          (:199-255)  brightest voxel refined by a parabola
//                                    along each axis, +0.6 voxel in z; the
//                                    background is the mean of the in-focus
//                                    plane's rows more than 20 above the bead
//   half_spectrum                    R2C of PSF - background by a mixed-radix
//                                    DFT in double (any length; FFTW FP32 in
//                                    the reference), kept as complex<float>
//   move_to_origin       (:286-319)  phase ramp that puts the bead at (0,0,0)
//   ring_average         (:321-381)  mean over rings of rint(kr) per kz, then
//                                    Hermitian along kz
//   mask_support         (:383-411)  zero outside the objective's support
//   extrapolate_origin   (:413-443)  kr = 0 from a straight-line fit over
//                                    kr = 1..n (3-D form :445-486: the two
//                                    diagonals of every kz plane)
//   normalize            (:488-518)  max |OTF| = 1
//
// Each rounding point (float or double) is the reference's, so that OTF
// files agree with its output to FP32 precision.
#include "../../include/radialft.h"

#include "tiff.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

using zd = std::complex<double>;
using zf = std::complex<float>;

[[noreturn]] void refuse(const std::string& why) { throw std::runtime_error(why); }

// index i of an n-point transform as a signed frequency
int signed_freq(int i, int n) { return i > n / 2 ? i - n : i; }

// ---------------------------------------------------------------------------
// DFT of any length in double: split by the smallest prime factor p of n
// into p interleaved subsequences of length n / p, transform those, and
// combine X[k + q m] = sum_r w^{r (k + q m)} Y_r[k], w = e^{-2 pi i / n}.
// ---------------------------------------------------------------------------
class Dft {
 public:
  static void forward(zd* a, int n) {
    if (n < 2) return;
    const int p = smallest_factor(n), m = n / p;
    std::vector<zd> parts((size_t)n);
    for (int j = 0; j < n; ++j) parts[(size_t)(j % p) * m + j / p] = a[j];
    for (int r = 0; r < p; ++r) forward(&parts[(size_t)r * m], m);
    const std::vector<zd>& w = roots(n);
    for (int k = 0; k < n; ++k) {
      const int kk = k % m;
      zd acc = parts[(size_t)kk];
      for (int r = 1; r < p; ++r) acc += parts[(size_t)r * m + kk] * w[(size_t)(((long)r * k) % n)];
      a[k] = acc;
    }
  }

 private:
  static int smallest_factor(int n) {
    for (int d = 2; (long)d * d <= n; ++d)
      if (n % d == 0) return d;
    return n;
  }
  static const std::vector<zd>& roots(int n) {
    static std::map<int, std::vector<zd>> table;
    auto it = table.find(n);
    if (it != table.end()) return it->second;
    std::vector<zd> w((size_t)n);
    for (int t = 0; t < n; ++t) w[(size_t)t] = std::polar(1.0, -2.0 * M_PI * (double)t / (double)n);
    return table.emplace(n, std::move(w)).first->second;
  }
};

// ---------------------------------------------------------------------------
// The PSF and its bead
// ---------------------------------------------------------------------------
struct Psf {
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> v;  // (nz, ny, nx), x fastest
  float operator()(int z, int y, int x) const { return v[((size_t)z * ny + y) * nx + x]; }
};

struct Bead {
  float x = 0, y = 0, z = 0;  // sub-voxel position
  float background = 0;
};

// vertex of the parabola through (-1, lo), (0, mid), (1, hi): 0 when the
// three are collinear or the vertex lies more than 1.5 voxels away
float vertex(float lo, float mid, float hi) {
  const float half_slope = 0.5 * (hi - lo);
  const float bend = (hi + lo) - 2 * mid;
  if (bend == 0) return 0.f;
  const float t = -half_slope / bend;
  return (t > 1.5 || t < -1.5) ? 0.f : t;
}

int wrap(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

Bead locate_bead(const Psf& p) {
  // the first brightest voxel in z, y, x order
  float top = 0.f;
  int bz = 0, by = 0, bx = 0;
  for (int z = 0; z < p.nz; ++z)
    for (int y = 0; y < p.ny; ++y)
      for (int x = 0; x < p.nx; ++x)
        if (p(z, y, x) > top) {
          top = p(z, y, x);
          bz = z;
          by = y;
          bx = x;
        }
  Bead b;
  b.z = bz + vertex(p(wrap(bz - 1, p.nz), by, bx), top, p(wrap(bz + 1, p.nz), by, bx));
  b.z += 0.6;  // the reference's fixed axial offset
  b.y = by + vertex(p(bz, wrap(by - 1, p.ny), bx), top, p(bz, wrap(by + 1, p.ny), bx));
  b.x = bx + vertex(p(bz, by, wrap(bx - 1, p.nx)), top, p(bz, by, wrap(bx + 1, p.nx)));
  const int plane = (int)std::floor(b.z);
  double acc = 0;
  for (int y = 0; y < b.y - 20; ++y)
    for (int x = 0; x < p.nx; ++x) acc += p(plane, y, x);
  b.background = acc / ((b.y - 20) * p.nx);
  return b;
}

// ---------------------------------------------------------------------------
// Half spectrum (nz, ny, nx/2+1)
// ---------------------------------------------------------------------------
struct Spectrum {
  int nx = 0, nh = 0, ny = 0, nz = 0;
  std::vector<zf> c;
  zf& at(int z, int y, int x) { return c[((size_t)z * ny + y) * nh + x]; }
  const zf& at(int z, int y, int x) const { return c[((size_t)z * ny + y) * nh + x]; }
};

Spectrum half_spectrum(const Psf& p, float background) {
  const int nx = p.nx, ny = p.ny, nz = p.nz;
  std::vector<zd> w(p.v.size());
  for (size_t i = 0; i < w.size(); ++i) w[i] = zd(p.v[i] - background, 0.0);
  auto cell = [&](int z, int y, int x) -> zd& { return w[((size_t)z * ny + y) * nx + x]; };
  for (long row = 0; row < (long)ny * nz; ++row) Dft::forward(&w[(size_t)row * nx], nx);
  std::vector<zd> line((size_t)std::max(ny, nz));
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x) {
      for (int y = 0; y < ny; ++y) line[(size_t)y] = cell(z, y, x);
      Dft::forward(line.data(), ny);
      for (int y = 0; y < ny; ++y) cell(z, y, x) = line[(size_t)y];
    }
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      for (int z = 0; z < nz; ++z) line[(size_t)z] = cell(z, y, x);
      Dft::forward(line.data(), nz);
      for (int z = 0; z < nz; ++z) cell(z, y, x) = line[(size_t)z];
    }
  Spectrum s;
  s.nx = nx;
  s.nh = nx / 2 + 1;
  s.ny = ny;
  s.nz = nz;
  s.c.resize((size_t)s.nh * ny * nz);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < s.nh; ++x) {
        const zd& d = cell(z, y, x);
        s.at(z, y, x) = zf((float)d.real(), (float)d.imag());
      }
  return s;
}

void move_to_origin(Spectrum& s, const Bead& b) {
  const float wz = 2 * M_PI * b.z / s.nz, wy = 2 * M_PI * b.y / s.ny, wx = 2 * M_PI * b.x / s.nx;
  for (int z = 0; z < s.nz; ++z) {
    const float pz = wz * signed_freq(z, s.nz);
    for (int y = 0; y < s.ny; ++y) {
      const float py = wy * signed_freq(y, s.ny);
      for (int x = 0; x < s.nh; ++x) {
        const float ph = pz + py + wx * x;
        s.at(z, y, x) *= zf(std::cos(ph), std::sin(ph));
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Rotationally averaged table (nr = nx/2+1 radii, nz), kz fastest
// ---------------------------------------------------------------------------
struct RadialOtf {
  int nr = 0, nz = 0;
  std::vector<zf> c;
  zf& at(int r, int z) { return c[(size_t)r * nz + z]; }
};

RadialOtf ring_average(const Spectrum& s) {
  RadialOtf o;
  o.nr = s.nh;
  o.nz = s.nz;
  o.c.assign((size_t)o.nr * o.nz, zf(0.f, 0.f));
  const int slots = o.nr * o.nz;
  std::vector<int> hits((size_t)slots, 0);
  for (int z = 0; z < s.nz; ++z)
    for (int y = 0; y < s.ny; ++y) {
      const int fy = signed_freq(y, s.ny);
      for (int x = 0; x < s.nh; ++x) {
        const float r = std::sqrt((double)(x * x + fy * fy));
        if (!(r < o.nr)) continue;
        const int slot = (int)std::rint(r) * o.nz + z;
        if (slot >= slots) continue;
        o.c[(size_t)slot] += s.at(z, y, x);
        ++hits[(size_t)slot];
      }
    }
  for (int i = 0; i < slots; ++i)
    if (hits[(size_t)i] > 0) o.c[(size_t)i] /= (float)hits[(size_t)i];
  // Hermitian along kz: OTF(kr, -kz) = conj OTF(kr, kz)
  for (int r = 0; r < o.nr; ++r) {
    o.at(r, 0) = zf(o.at(r, 0).real(), 0.f);
    for (int kz = 1; kz <= o.nz / 2; ++kz) {
      const zf both = (o.at(r, kz) + std::conj(o.at(r, o.nz - kz))) / 2.f;
      o.at(r, kz) = both;
      o.at(r, o.nz - kz) = std::conj(both);
    }
  }
  return o;
}

// Zero every (kr, kz) the objective cannot pass: kr beyond 2 NA / lambda (or
// the caller's radius in pixels, when smaller), and per kr the kz band
// between the two edges of the support.
void mask_support(RadialOtf& o, float dkr, float dkz, int lambda_nm, int kr_cap, float na, float n_imm) {
  if (!(na > 0.f && na < n_imm)) refuse("makeOTF: cleanup needs 0 < NA < the immersion index");
  const float wl = lambda_nm * 0.001;  // micrometres
  const float cos_max = std::cos(std::asin(na / n_imm));
  float kr_edge = 2 * na / wl;
  if (kr_cap != 0 && kr_cap * dkr < kr_edge) kr_edge = kr_cap * dkr;
  for (int r = 0; r < o.nr; ++r) {
    const float kr = r * dkr;
    if (kr > kr_edge) {
      for (int z = 0; z < o.nz; ++z) o.at(r, z) = 0;
      continue;
    }
    const float tilt = std::asin((na - kr * wl) / n_imm);
    const float kz_edge = (n_imm / wl) * (std::cos(tilt) - cos_max);
    const int first = std::max(0, (int)std::rint((kz_edge / dkz) + 1));
    for (int z = first; z <= o.nz - first; ++z) o.at(r, z) = 0;
  }
}

// least-squares line through (k, v[k]), k = 1..n, evaluated at k = 0
zd line_at_zero(const std::vector<zd>& v, int n) {
  const double kbar = (n + 1) / 2.;
  zd vbar = 0;
  for (int k = 1; k <= n; ++k) vbar += v[(size_t)k];
  vbar /= (double)n;
  zd cov = 0, var = 0;
  for (int k = 1; k <= n; ++k) {
    cov += (k - kbar) * (v[(size_t)k] - vbar);
    var += (k - kbar) * (k - kbar);
  }
  return vbar - (cov / var) * kbar;
}

zf to_float(const zd& d) { return zf((float)d.real(), (float)d.imag()); }

// kr = 0 of every kz from kr = 1..n; false (nothing usable) when the
// extrapolated kz = 0 value is not above its kr = 1 neighbour
bool extrapolate_origin(RadialOtf& o, int n) {
  std::vector<zd> col((size_t)n + 1);
  for (int z = 0; z < o.nz; ++z) {
    for (int k = 1; k <= n; ++k) col[(size_t)k] = zd(o.at(k, z));
    o.at(0, z) = to_float(line_at_zero(col, n));
    if (z == 0 && std::abs(o.at(0, 0)) <= std::abs(o.at(1, 0))) return false;
  }
  return true;
}

// 3-D form: the (0, 0, 0) value only, from the sums over kz of the +45 and
// -45 degree diagonals (ky = kx and ky = -kx) at kx = 1..n.  The diagonals
// stop at ky = ny - 1, where the reference would read past the plane of a
// stack wider than tall.
void extrapolate_origin_3d(Spectrum& s, int n) {
  std::vector<zf> diag((size_t)s.nh, zf(0.f, 0.f));
  for (int z = 0; z < s.nz; ++z)
    for (int k = 1; k < s.nh && k < s.ny; ++k) {
      diag[(size_t)k] += s.at(z, k, k);
      diag[(size_t)k] += s.at(z, s.ny - k, k);
    }
  std::vector<zd> col((size_t)n + 1);
  for (int k = 1; k <= n; ++k) col[(size_t)k] = zd(diag[(size_t)k]);
  s.at(0, 0, 0) = to_float(line_at_zero(col, n));
}

// max |OTF| = 1: the averaged table is multiplied by 1 / max, the 3-D one
// divided by max (the two rounding forms of the reference)
void normalize(std::vector<zf>& c, bool divide) {
  float top = 0;
  for (const zf& v : c) top = std::max(top, std::abs(v));
  if (divide) {
    for (zf& v : c) v = zf(v.real() / top, v.imag() / top);
  } else {
    const float inv = 1 / top;
    for (zf& v : c) v *= inv;
  }
}

}  // namespace

extern "C" {

const char* radialft_last_error(void) { return g_err.c_str(); }

int makeOTF(const char* const ifiles, const char* const ofiles, int lambdanm, float dz, int interpkr,
            bool bUserBackground, float background, float NA, float NIMM, float dr, int krmax, bool bDoCleanup,
            bool b3Dout) {
  try {
    g_err.clear();
    if (!ifiles || !ofiles) refuse("makeOTF: null file name");
    cdn::TiffImage t = cdn::tiff_read(ifiles);
    Psf psf;
    psf.nx = t.width;
    psf.ny = t.height;
    psf.nz = t.depth;
    psf.v = std::move(t.data);
    if (psf.nx < 2 || psf.ny < 2 || psf.nz < 2) refuse("makeOTF: the PSF must be at least 2x2x2");
    const Bead bead = locate_bead(psf);
    Spectrum s = half_spectrum(psf, bUserBackground ? background : bead.background);
    move_to_origin(s, bead);
    if (interpkr >= s.nh) refuse("makeOTF: interpkr must be below nx/2+1");
    if (b3Dout) {
      if (interpkr > 1) extrapolate_origin_3d(s, interpkr);
      normalize(s.c, true);
      cdn::tiff_write_f32(ofiles, reinterpret_cast<const float*>(s.c.data()), 2 * s.nh, s.ny, s.nz);
      return 0;
    }
    RadialOtf o = ring_average(s);
    if (bDoCleanup) mask_support(o, 1 / (psf.ny * dr), 1 / (psf.nz * dz), lambdanm, krmax, NA, NIMM);
    if (interpkr > 0)
      while (!extrapolate_origin(o, interpkr))
        if (--interpkr < 4) refuse("makeOTF: fewer than 4 kr samples left for the kr = 0 extrapolation (interpkr)");
    normalize(o.c, false);
    cdn::tiff_write_f32(ofiles, reinterpret_cast<const float*>(o.c.data()), 2 * o.nz, o.nr, 1);
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "unknown error";
  }
  return 1;
}

}  // extern "C"
