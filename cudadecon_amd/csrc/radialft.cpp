// libradialft.so: makeOTF, OTF generation from a measured PSF stack
// (src/radialft_interface.cpp:95-518), on the host.
//
// Steps, each following the reference function named:
//   determine_center_and_background (:199-255)  peak by parabolic fits, z + 0.6,
//                                               background from rows y < yc - 20
//   R2C of PSF - background                     own mixed-radix FFT in double
//                                               (any length; FFTW FP32 in the
//                                               reference), stored as complex<float>
//   shift_center (:286-319)                     phase ramp moving the bead to the origin
//   radialft (:321-381)                         average over rint(kr) bins, kz Hermitian
//   cleanup (:383-411)                          zero outside the OTF support (optional)
//   fixorigin (:413-443, 3-D :445-486)          kr = 0 by linear extrapolation
//   rescale (:488-518)                          max |OTF| = 1
#include "../../include/radialft.h"

#include "tiff.hpp"

#include <cmath>
#include <complex>
#include <cstdio>
#include <algorithm>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

typedef std::complex<double> cd;
typedef std::complex<float> cf;

// e^{-2 pi i t / n}, t < n, one table per length
const std::vector<cd>& twiddles(int n) {
  static std::map<int, std::vector<cd>> cache;
  auto it = cache.find(n);
  if (it != cache.end()) return it->second;
  std::vector<cd> w((size_t)n);
  for (int t = 0; t < n; ++t) {
    const double a = -2.0 * M_PI * (double)t / (double)n;
    w[(size_t)t] = cd(std::cos(a), std::sin(a));
  }
  return cache.emplace(n, std::move(w)).first->second;
}

// In-place forward DFT of any length: decimation in time over the smallest
// prime factor p, X[k + m q] = sum_r W_n^{r (k + m q)} Y_r[k] (m = n / p).
void fft_line(cd* x, int n) {
  if (n <= 1) return;
  int p = 2;
  while ((long)p * p <= n && n % p) ++p;
  if (n % p) p = n;
  const int m = n / p;
  std::vector<cd> sub((size_t)n);
  for (int r = 0; r < p; ++r)
    for (int j = 0; j < m; ++j) sub[(size_t)r * m + j] = x[(size_t)j * p + r];
  for (int r = 0; r < p; ++r) fft_line(&sub[(size_t)r * m], m);
  const std::vector<cd>& w = twiddles(n);
  for (int q = 0; q < p; ++q)
    for (int k = 0; k < m; ++k) {
      const long kk = k + (long)m * q;
      cd s = sub[(size_t)k];
      for (int r = 1; r < p; ++r) s += sub[(size_t)r * m + k] * w[(size_t)((r * kk) % n)];
      x[kk] = s;
    }
}

// R2C of a (nz, ny, nx) real volume: (nz, ny, nx/2+1) complex<float>
std::vector<cf> rfft3(const std::vector<float>& v, int nx, int ny, int nz) {
  std::vector<cd> a((size_t)nx * ny * nz);
  for (size_t i = 0; i < a.size(); ++i) a[i] = cd(v[i], 0.0);
  std::vector<cd> line;
  for (long r = 0; r < (long)ny * nz; ++r) fft_line(&a[(size_t)r * nx], nx);
  line.resize((size_t)ny);
  for (int z = 0; z < nz; ++z)
    for (int x = 0; x < nx; ++x) {
      for (int y = 0; y < ny; ++y) line[(size_t)y] = a[((size_t)z * ny + y) * nx + x];
      fft_line(line.data(), ny);
      for (int y = 0; y < ny; ++y) a[((size_t)z * ny + y) * nx + x] = line[(size_t)y];
    }
  line.resize((size_t)nz);
  for (int y = 0; y < ny; ++y)
    for (int x = 0; x < nx; ++x) {
      for (int z = 0; z < nz; ++z) line[(size_t)z] = a[((size_t)z * ny + y) * nx + x];
      fft_line(line.data(), nz);
      for (int z = 0; z < nz; ++z) a[((size_t)z * ny + y) * nx + x] = line[(size_t)z];
    }
  const int nh = nx / 2 + 1;
  std::vector<cf> out((size_t)nh * ny * nz);
  for (long r = 0; r < (long)ny * nz; ++r)
    for (int x = 0; x < nh; ++x) out[(size_t)r * nh + x] = cf((float)a[(size_t)r * nx + x].real(),
                                                                (float)a[(size_t)r * nx + x].imag());
  return out;
}

float fitparabola(float a1, float a2, float a3) {
  const float slope = 0.5 * (a3 - a1);
  const float curve = (a3 + a1) - 2 * a2;
  if (curve == 0) return 0.f;
  const float peak = -slope / curve;
  if (peak > 1.5 || peak < -1.5) return 0.f;
  return peak;
}

void center_and_background(const std::vector<float>& s, int nx, int ny, int nz, float* xc, float* yc, float* zc,
                           float* bg) {
  float maxval = 0.f;
  int mi = 0, mj = 0, mk = 0;
  for (int k = 0; k < nz; ++k)
    for (int i = 0; i < ny; ++i)
      for (int j = 0; j < nx; ++j) {
        const float v = s[((size_t)k * ny + i) * nx + j];
        if (v > maxval) {
          maxval = v;
          mi = i;
          mj = j;
          mk = k;
        }
      }
  auto at = [&](int k, int i, int j) { return s[((size_t)k * ny + i) * nx + j]; };
  const int im = mi - 1 < 0 ? mi - 1 + ny : mi - 1, ip = mi + 1 >= ny ? mi + 1 - ny : mi + 1;
  const int jm = mj - 1 < 0 ? mj - 1 + nx : mj - 1, jp = mj + 1 >= nx ? mj + 1 - nx : mj + 1;
  const int km = mk - 1 < 0 ? mk - 1 + nz : mk - 1, kp = mk + 1 >= nz ? mk + 1 - nz : mk + 1;
  *zc = mk + fitparabola(at(km, mi, mj), maxval, at(kp, mi, mj));
  *zc += 0.6;
  *yc = mi + fitparabola(at(mk, im, mj), maxval, at(mk, ip, mj));
  *xc = mj + fitparabola(at(mk, mi, jm), maxval, at(mk, mi, jp));
  double sum = 0;
  const int infocus = (int)std::floor(*zc);
  for (int i = 0; i < *yc - 20; ++i)
    for (int j = 0; j < nx; ++j) sum += at(infocus, i, j);
  *bg = sum / ((*yc - 20) * nx);
}

void shift_center(cf* bands, int nx, int ny, int nz, float xc, float yc, float zc) {
  const int nh = nx / 2 + 1;
  const float dphiz = 2 * M_PI * zc / nz, dphiy = 2 * M_PI * yc / ny, dphix = 2 * M_PI * xc / nx;
  for (int kin = 0; kin < nz; ++kin) {
    const int kz = kin > nz / 2 ? kin - nz : kin;
    const float phi1 = dphiz * kz;
    for (int iin = 0; iin < ny; ++iin) {
      const int ky = iin > ny / 2 ? iin - ny : iin;
      const float phi2 = dphiy * ky;
      for (int kx = 0; kx < nh; ++kx) {
        const float phi = phi1 + phi2 + dphix * kx;
        cf& b = bands[((size_t)kin * ny + iin) * nh + kx];
        b = b * cf(std::cos(phi), std::sin(phi));
      }
    }
  }
}

void radialft(const cf* band, int nx, int ny, int nz, cf* avg) {
  const int nh = nx / 2 + 1, nxz = nh * nz;
  std::vector<int> count((size_t)nxz, 0);
  for (int kin = 0; kin < nz; ++kin)
    for (int iin = 0; iin < ny; ++iin) {
      const int ky = iin > ny / 2 ? iin - ny : iin;
      for (int kx = 0; kx < nh; ++kx) {
        const float rdist = std::sqrt((double)(kx * kx + ky * ky));
        if (rdist < nh) {
          const int out = (int)std::rint(rdist) * nz + kin;
          if (out < nxz) {
            avg[out] += band[((size_t)kin * ny + iin) * nh + kx];
            count[(size_t)out]++;
          }
        }
      }
    }
  for (int i = 0; i < nxz; ++i)
    if (count[(size_t)i] > 0) avg[i] /= (float)count[(size_t)i];
  for (int kx = 0; kx < nh; ++kx) {
    avg[kx * nz] = cf(avg[kx * nz].real(), 0.f);
    for (int kz = 1; kz <= nz / 2; ++kz) {
      const int i = kx * nz + kz, ic = kx * nz + (nz - kz);
      avg[i] = (avg[i] + std::conj(avg[ic])) / 2.f;
      avg[ic] = std::conj(avg[i]);
    }
  }
}

void cleanup(cf* otf, int nx, int nz, float dkr, float dkz, int lamdanm, int krmax_user, float NA, float NIMM) {
  const int ncl = nx / 2 + 1;
  const float lamda = lamdanm * 0.001;
  const float sinalpha = NA / NIMM;
  const float cosalpha = std::cos(std::asin(sinalpha));
  float krmax = 2 * NA / lamda;
  if (krmax_user * dkr < krmax && krmax_user != 0) krmax = krmax_user * dkr;
  for (int ix = 0; ix < ncl; ++ix) {
    const float kr = ix * dkr;
    if (kr <= krmax) {
      const float beta = std::asin((NA - kr * lamda) / NIMM);
      const float kzedge = (NIMM / lamda) * (std::cos(beta) - cosalpha);
      const int kzstart = (int)std::rint((kzedge / dkz) + 1);
      const int kzend = nz - kzstart;
      for (int iz = kzstart; iz <= kzend; ++iz) otf[ix * nz + iz] = 0;
    } else {
      for (int iz = 0; iz < nz; ++iz) otf[ix * nz + iz] = 0;
    }
  }
}

// least-squares intercept at x = 0 of vals[1..kx2] (double, as the reference)
cd intercept(const cd* vals, int kx2) {
  const double mean_kx = (kx2 + 1) / 2.;
  cd mean_val = 0, num = 0, den = 0;
  for (int x = 1; x <= kx2; ++x) mean_val += vals[x];
  mean_val /= (double)kx2;
  for (int x = 1; x <= kx2; ++x) {
    num += (x - mean_kx) * (vals[x] - mean_val);
    den += (x - mean_kx) * (x - mean_kx);
  }
  return mean_val - (num / den) * mean_kx;
}

bool fixorigin(cf* otf, int nx, int nz, int kx2) {
  std::vector<cd> col((size_t)kx2 + 1);
  for (int z = 0; z < nz; ++z) {
    for (int x = 1; x <= kx2; ++x) col[(size_t)x] = cd(otf[x * nz + z]);
    const cd c = intercept(col.data(), kx2);
    otf[z] = cf((float)c.real(), (float)c.imag());
    if (z == 0 && std::abs(otf[z]) <= std::abs(otf[nz + z])) return false;
  }
  return true;
}

// 3-D variant: project the +45 and -45 degree radial lines of every kz plane
// onto kx and extrapolate the (0, 0, 0) value only (lines stop at ky = ny - 1,
// where the reference would read past the plane of a stack wider than tall)
void fixorigin3d(cf* otf, int nh, int ny, int nz, int kx2) {
  std::vector<cf> proj((size_t)nh, cf(0.f, 0.f));
  for (int z = 0; z < nz; ++z)
    for (int x = 1; x < nh && x < ny; ++x) {
      proj[(size_t)x] += otf[((size_t)z * ny + x) * nh + x];
      proj[(size_t)x] += otf[((size_t)z * ny + (ny - x)) * nh + x];
    }
  std::vector<cd> col((size_t)kx2 + 1);
  for (int x = 1; x <= kx2; ++x) col[(size_t)x] = cd(proj[(size_t)x]);
  const cd c = intercept(col.data(), kx2);
  otf[0] = cf((float)c.real(), (float)c.imag());
}

// RA table: multiply by 1 / max (:488-504); 3-D: divide by max (:506-518)
void rescale(cf* otf, size_t n, bool divide) {
  float valmax = 0;
  for (size_t i = 0; i < n; ++i) valmax = std::max(valmax, std::abs(otf[i]));
  if (divide) {
    for (size_t i = 0; i < n; ++i) otf[i] = cf(otf[i].real() / valmax, otf[i].imag() / valmax);
  } else {
    const float s = 1 / valmax;
    for (size_t i = 0; i < n; ++i) otf[i] *= s;
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
    if (!ifiles || !ofiles) throw std::runtime_error("makeOTF: null file name");
    cdn::TiffImage t = cdn::tiff_read(ifiles);
    const int nx = t.width, ny = t.height, nz = t.depth;
    if (nx < 2 || ny < 2 || nz < 2) throw std::runtime_error("makeOTF: the PSF must be at least 2x2x2");
    const float dkr = 1 / (ny * dr);
    const float dkz = 1 / (nz * dz);
    float xc, yc, zc, est;
    center_and_background(t.data, nx, ny, nz, &xc, &yc, &zc, &est);
    if (!bUserBackground) background = est;
    for (float& v : t.data) v -= background;
    std::vector<cf> bands = rfft3(t.data, nx, ny, nz);
    shift_center(bands.data(), nx, ny, nz, xc, yc, zc);
    const int nh = nx / 2 + 1;
    if (b3Dout) {
      if (interpkr > 1) {
        if (interpkr >= nh) throw std::runtime_error("makeOTF: interpkr must be below nx/2+1");
        fixorigin3d(bands.data(), nh, ny, nz, interpkr);
      }
      rescale(bands.data(), bands.size(), true);
      cdn::tiff_write_f32(ofiles, reinterpret_cast<const float*>(bands.data()), 2 * nh, ny, nz);
    } else {
      std::vector<cf> avg((size_t)nh * nz, cf(0.f, 0.f));
      radialft(bands.data(), nx, ny, nz, avg.data());
      if (bDoCleanup) cleanup(avg.data(), nx, nz, dkr, dkz, lambdanm, krmax, NA, NIMM);
      if (interpkr > 0) {
        if (interpkr >= nh) throw std::runtime_error("makeOTF: interpkr must be below nx/2+1");
        while (!fixorigin(avg.data(), nx, nz, interpkr)) {
          --interpkr;
          if (interpkr < 4) throw std::runtime_error("#pixels < 4 used in kr=0 extrapolation");
        }
      }
      rescale(avg.data(), avg.size(), false);
      cdn::tiff_write_f32(ofiles, reinterpret_cast<const float*>(avg.data()), 2 * nz, nh, 1);
    }
    return 0;
  } catch (const std::exception& e) {
    g_err = e.what();
  } catch (...) {
    g_err = "unknown error";
  }
  return 1;
}

}  // extern "C"
