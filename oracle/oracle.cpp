// CPU oracle for the Richardson-Lucy / Biggs-Andrews deconvolution path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (libcudaDecon.so, the
// cudadecon_amd package) links, loads or calls this file.  It is used by
// tests/, by __graft_entry__.smoke() as the checker, and by bench.py's
// cpu_baseline leg (the FP32 flavour, timed on the host cores).
//
// Every function restates the reference algorithm in plain C++, templated on
// the arithmetic type: double is the correctness oracle, float (OpenMP) is
// the timed CPU baseline.  The FFT is this file's own mixed-radix Stockham
// transform (any length; primes > 7 by direct DFT), pinned by numpy/pocketfft
// golden vectors and brute-force DFTs in tests/.
//
// Reference files restated (paths relative to the reference repository):
//   src/RLgpuImpl.cu:124-132          background subtraction
//   src/RLgpuImpl.cu:156-225          filterGPU (scale 1/N, R2C, x OTF, C2R)
//   src/RLgpuImpl.cu:228-381          OTF interpolation (RA bilinear / 3D trilinear)
//   src/RLgpuImpl.cu:416-434          LRcore
//   src/RLgpuImpl.cu:462-508          current estimate, difference
//   src/RLgpuImpl.cu:510-633          acceleration factor (two dot products)
//   src/RLgpuImpl.cu:658-666          prediction update
//   src/RLgpuImpl.cu:668-827          mean above background (two reductions)
//   src/RLgpuImpl.cu:829-894          rescale, apodize
//   src/RLgpuImpl.cu:908-925          z blend (intended per-slice taper)
//   src/RL-Biggs-Andrews.cpp:18-43    findOptimalDimension
//   src/RL-Biggs-Andrews.cpp:273-367  RL-BA iteration
//   src/geometryTransform.cu:4-31     deskew
//   src/geometryTransform.cu:50-88    rotate
//   src/geometryTransform.cu:183-253  affine (texture trilinear, border 0)
//   src/radialft_interface.cpp:95-518 radially averaged OTF generation

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>
#include <omp.h>

namespace {

template <class T> using cx = std::complex<T>;

// ---------------------------------------------------------------------------
// 1-D mixed-radix Stockham FFT, any length.  Forward: X[k] = sum x[n] e^{-2 pi i nk/N};
// inverse is unnormalised.  Radices 4, 2, 3, 5, 7 have direct butterflies;
// any other prime factor uses a generic R-point DFT.
// ---------------------------------------------------------------------------
template <class T> struct Fft1 {
  int n = 0;
  std::vector<int> rad;
  std::vector<cx<T>> w;  // w[m] = exp(-2 pi i m / n), m in [0, n)

  explicit Fft1(int n_) : n(n_) {
    int m = n;
    for (int r : {4, 2, 3, 5, 7}) {
      while (m % r == 0) { rad.push_back(r); m /= r; }
    }
    for (int p = 11; m > 1; p += 2) {
      while (m % p == 0) { rad.push_back(p); m /= p; }
    }
    w.resize(n);
    for (int i = 0; i < n; ++i) {
      // angle reduced to [0, 2pi) exactly in integers before the double math
      long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)i / (long double)n;
      w[i] = cx<T>((T)std::cos(a), (T)std::sin(a));
    }
  }

  // In-place transform of a (length n) using work (length n).
  void run(cx<T>* a, cx<T>* work, bool inverse) const {
    if (n == 1) return;
    cx<T>* src = a;
    cx<T>* dst = work;
    int Lp = 1;
    for (int R : rad) {
      const int S = n / R;
      const int stepN = n / (Lp * R);  // twiddle index multiplier: W_{Lp R}^{tk} = w[t k stepN]
      cx<T> xt[64];
      for (int j = 0; j < S; ++j) {
        const int k = j % Lp;
        for (int t = 0; t < R; ++t) {
          cx<T> v = src[j + t * S];
          if (k) {
            cx<T> tw = w[(long)t * k * stepN % n];
            if (inverse) tw = std::conj(tw);
            v *= tw;
          }
          xt[t] = v;
        }
        const int base = (j - k) * R + k;
        butterfly(xt, R, inverse);
        for (int q = 0; q < R; ++q) dst[base + q * Lp] = xt[q];
      }
      std::swap(src, dst);
      Lp *= R;
    }
    if (src != a) std::memcpy(a, src, sizeof(cx<T>) * n);
  }

  void butterfly(cx<T>* x, int R, bool inverse) const {
    const T sg = inverse ? T(1) : T(-1);  // sign of the exponent
    if (R == 2) {
      cx<T> a = x[0], b = x[1];
      x[0] = a + b; x[1] = a - b;
    } else if (R == 4) {
      cx<T> a = x[0] + x[2], b = x[0] - x[2], c = x[1] + x[3], d = x[1] - x[3];
      cx<T> jd(-sg * d.imag(), sg * d.real());  // (sg i) * d
      x[0] = a + c; x[2] = a - c; x[1] = b + jd; x[3] = b - jd;
    } else {
      // generic odd R: X_q = x0 + sum_t x_t W_R^{tq}, using exact W_n powers
      cx<T> y[64];
      const int st = n / R;
      for (int q = 0; q < R; ++q) {
        cx<T> s = x[0];
        for (int t = 1; t < R; ++t) {
          cx<T> tw = w[(long)((t * q) % R) * st];
          if (sg > 0) tw = std::conj(tw);
          s += x[t] * tw;
        }
        y[q] = s;
      }
      for (int q = 0; q < R; ++q) x[q] = y[q];
    }
  }
};

// ---------------------------------------------------------------------------
// 3-D real <-> half-complex transforms, layout (nz, ny, nx) x fastest; the
// half spectrum is (nz, ny, nx/2+1).  Semantics follow numpy.fft.rfftn /
// irfftn: C2R runs complex inverses over z and y, then a real inverse along x
// that ignores the imaginary parts of the kx = 0 and kx = nx/2 bins.
// Unnormalised in both directions (as cuFFT at src/RLgpuImpl.cu:183,218).
// ---------------------------------------------------------------------------
template <class T> struct Fft3 {
  int nx, ny, nz, nxh;
  Fft1<T> fx, fy, fz;
  Fft3(int nx_, int ny_, int nz_)
      : nx(nx_), ny(ny_), nz(nz_), nxh(nx_ / 2 + 1), fx(nx_), fy(ny_), fz(nz_) {}

  // Column pass over axis with `len` points at element stride `stride` (in
  // complex units), for `ncols` contiguous columns per block and `nblk` blocks.
  void columns(cx<T>* S, const Fft1<T>& f, int len, long stride, long blk_stride, int nblk,
               int ncols, bool inverse) const {
    const int CB = 16;
    const int nb = (ncols + CB - 1) / CB;
#pragma omp parallel
    {
      std::vector<cx<T>> line((size_t)CB * len), work(len);
#pragma omp for schedule(static)
      for (long t = 0; t < (long)nblk * nb; ++t) {
        const long b = t / nb;
        const int c0 = (int)(t % nb) * CB;
        const int cn = std::min(CB, ncols - c0);
        cx<T>* base = S + b * blk_stride + c0;
        for (int i = 0; i < len; ++i)
          for (int c = 0; c < cn; ++c) line[(size_t)c * len + i] = base[i * stride + c];
        for (int c = 0; c < cn; ++c) f.run(&line[(size_t)c * len], work.data(), inverse);
        for (int i = 0; i < len; ++i)
          for (int c = 0; c < cn; ++c) base[i * stride + c] = line[(size_t)c * len + i];
      }
    }
  }

  // R2C: real rows are paired as z = a + i b, transformed once, and split.
  void r2c(const T* in, cx<T>* out) const {
    const long rows = (long)ny * nz;
    const long pairs = (rows + 1) / 2;
#pragma omp parallel
    {
      std::vector<cx<T>> z(nx), work(nx);
#pragma omp for schedule(static)
      for (long p = 0; p < pairs; ++p) {
        const long r0 = 2 * p, r1 = 2 * p + 1;
        const T* a = in + r0 * nx;
        const T* b = r1 < rows ? in + r1 * nx : nullptr;
        for (int i = 0; i < nx; ++i) z[i] = cx<T>(a[i], b ? b[i] : T(0));
        fx.run(z.data(), work.data(), false);
        for (int k = 0; k < nxh; ++k) {
          cx<T> zk = z[k], zn = std::conj(z[(nx - k) % nx]);
          out[r0 * nxh + k] = (zk + zn) * T(0.5);
          if (b) {
            cx<T> d = (zk - zn) * T(0.5);
            out[r1 * nxh + k] = cx<T>(d.imag(), -d.real());  // d / i
          }
        }
      }
    }
    columns(out, fy, ny, nxh, (long)ny * nxh, nz, nxh, false);
    columns(out, fz, nz, (long)ny * nxh, 0, 1, ny * nxh, false);
  }

  // C2R; `spec` is consumed (overwritten).
  void c2r(cx<T>* spec, T* out) const {
    columns(spec, fz, nz, (long)ny * nxh, 0, 1, ny * nxh, true);
    columns(spec, fy, ny, nxh, (long)ny * nxh, nz, nxh, true);
    const long rows = (long)ny * nz;
    const long pairs = (rows + 1) / 2;
#pragma omp parallel
    {
      std::vector<cx<T>> z(nx), work(nx);
#pragma omp for schedule(static)
      for (long p = 0; p < pairs; ++p) {
        const long r0 = 2 * p, r1 = 2 * p + 1;
        const bool hasb = r1 < rows;
        for (int k = 0; k < nxh; ++k) {
          cx<T> A = spec[r0 * nxh + k];
          cx<T> B = hasb ? spec[r1 * nxh + k] : cx<T>(0);
          if (k == 0 || 2 * k == nx) { A.imag(0); B.imag(0); }
          // Z[k] = A[k] + i B[k];  Z[n-k] = conj(A[k]) + i conj(B[k])
          z[k] = cx<T>(A.real() - B.imag(), A.imag() + B.real());
          if (k > 0 && 2 * k != nx)
            z[nx - k] = cx<T>(A.real() + B.imag(), -A.imag() + B.real());
        }
        fx.run(z.data(), work.data(), true);
        for (int i = 0; i < nx; ++i) {
          out[r0 * nx + i] = z[i].real();
          if (hasb) out[r1 * nx + i] = z[i].imag();
        }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// OTF interpolation onto the (nz, ny, nx/2+1) half grid.
// Restates dev_otfinterpolate + makeOTFarray_kernel (src/RLgpuImpl.cu:228-366):
// signed ky, kz (k > n>>1 -> k - n), scaled by the k-scales; RA (nyotf == 1):
// bilinear in (kr, kz), zero unless kr < nxotf-1, kz wraps at iz = nzotf-1;
// 3-D: trilinear, ky and kz wrap, kx does not.
// ---------------------------------------------------------------------------
template <class T>
void otf_interp(const float* raw /*complex64 interleaved*/, int nxotf, int nyotf, int nzotf,
                int nx, int ny, int nz, T kxscale, T kyscale, T kzscale, cx<T>* out) {
  const int nxh = nx / 2 + 1;
  auto at = [&](long idx) { return cx<T>(raw[2 * idx], raw[2 * idx + 1]); };
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; ++z) {
    const int kzi = z > (nz >> 1) ? z - nz : z;
    for (int y = 0; y < ny; ++y) {
      const int kyi = y > (ny >> 1) ? y - ny : y;
      for (int x = 0; x < nxh; ++x) {
        const T kx = x * kxscale, ky = kyi * kyscale, kz = kzi * kzscale;
        const T kzindex = kz < 0 ? kz + nzotf : kz;
        cx<T> v(0, 0);
        if (nyotf == 1) {
          const T kr = std::sqrt(kx * kx + ky * ky);
          if (kr < nxotf - 1 && kzindex < nzotf) {
            const int ir = (int)std::floor(kr), iz = (int)std::floor(kzindex);
            const T ar = kr - ir, az = kzindex - iz;
            const int iz1 = (iz == nzotf - 1) ? 0 : iz + 1;
            v = (T(1) - ar) * (at((long)ir * nzotf + iz) * (T(1) - az) + at((long)ir * nzotf + iz1) * az) +
                ar * (at((long)(ir + 1) * nzotf + iz) * (T(1) - az) + at((long)(ir + 1) * nzotf + iz1) * az);
          }
        } else {
          const T kxindex = kx;
          const T kyindex = ky < 0 ? ky + nyotf : ky;
          if (kxindex < nxotf - 1 && kyindex < nyotf && kzindex < nzotf) {
            const int ix = (int)std::floor(kxindex), iy = (int)std::floor(kyindex), iz = (int)std::floor(kzindex);
            const int iy1 = (iy + 1) % nyotf, iz1 = (iz + 1) % nzotf;
            const T ax = kxindex - ix, ay = kyindex - iy, az = kzindex - iz;
            const long nxy = (long)nxotf * nyotf;
            auto g = [&](int zz, int yy, int xx) { return at(zz * nxy + (long)yy * nxotf + xx); };
            v = (T(1) - az) * (g(iz, iy, ix) * (T(1) - ay) * (T(1) - ax) + g(iz, iy, ix + 1) * (T(1) - ay) * ax +
                               g(iz, iy1, ix) * ay * (T(1) - ax) + g(iz, iy1, ix + 1) * ay * ax) +
                az * (g(iz1, iy, ix) * (T(1) - ay) * (T(1) - ax) + g(iz1, iy, ix + 1) * (T(1) - ay) * ax +
                      g(iz1, iy1, ix) * ay * (T(1) - ax) + g(iz1, iy1, ix + 1) * ay * ax);
          }
        }
        out[((long)z * ny + y) * nxh + x] = v;
      }
    }
  }
}

// conv(v, H) = C2R(H . R2C(v / N))  (src/RLgpuImpl.cu:156-225)
template <class T>
void filter(T* v, const Fft3<T>& F, const cx<T>* H, bool conjH, cx<T>* S) {
  const long N = (long)F.nx * F.ny * F.nz;
  const long M = (long)F.nz * F.ny * F.nxh;
  const T inv = T(1) / T(N);
  F.r2c(v, S);
#pragma omp parallel for schedule(static)
  for (long i = 0; i < M; ++i) S[i] *= (conjH ? std::conj(H[i]) : H[i]) * inv;
  F.c2r(S, v);
}

// RL-Biggs-Andrews iteration (src/RL-Biggs-Andrews.cpp:273-367).
// raw: preprocessed stack (also X_0).  Y0: optional seed for Y at k = 0
// (the FlatStart hook at :305-316).  lambdas: optional, nIter entries.
template <class T>
void rl_ba(const T* raw, int nx, int ny, int nz, const cx<T>* H, int nIter, const T* Y0, T* X,
           double* lambdas, double* iter_seconds = nullptr) {
  const long N = (long)nx * ny * nz;
  const long M = (long)nz * ny * (nx / 2 + 1);
  Fft3<T> F(nx, ny, nz);
  std::vector<T> Xp(N), Y(N), C(N), G1(N), G2(N);
  std::vector<cx<T>> S(M);
  const T eps = std::numeric_limits<float>::epsilon();
  std::memcpy(X, raw, sizeof(T) * N);
  for (int k = 0; k < nIter; ++k) {
    const auto t0 = std::chrono::steady_clock::now();
    double lambda = 0;
    if (k > 1) {
      double num = 0, den = 0;
#pragma omp parallel for reduction(+ : num, den) schedule(static)
      for (long i = 0; i < N; ++i) {
        num += (double)G1[i] * (double)G2[i];
        den += (double)G2[i] * (double)G2[i];
      }
      lambda = num / (den + (double)std::numeric_limits<float>::epsilon());
      lambda = std::max(std::min(lambda, 1.), 0.);
      const T lam = (T)lambda;
#pragma omp parallel for schedule(static)
      for (long i = 0; i < N; ++i) {
        T y = X[i] + lam * (X[i] - Xp[i]);
        Y[i] = y > 0 ? y : T(0);
      }
    } else if (k == 0 && Y0) {
      std::memcpy(Y.data(), Y0, sizeof(T) * N);
    } else {
      std::memcpy(Y.data(), X, sizeof(T) * N);
    }
    if (lambdas) lambdas[k] = lambda;
    std::memcpy(Xp.data(), X, sizeof(T) * N);
    if (k > 0) std::memcpy(G2.data(), G1.data(), sizeof(T) * N);
    std::memcpy(C.data(), Y.data(), sizeof(T) * N);
    filter(C.data(), F, H, false, S.data());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      T c = std::fabs(C[i]) > 0 ? C[i] : eps;
      c = raw[i] / c + eps;
      if (c > T(10)) c = T(10);
      if (c < T(-10)) c = T(-10);
      C[i] = c;
    }
    filter(C.data(), F, H, true, S.data());
#pragma omp parallel for schedule(static)
    for (long i = 0; i < N; ++i) {
      T x = C[i] * Y[i];
      X[i] = x > 0 ? x : T(0);
      G1[i] = X[i] - Y[i];
    }
    if (iter_seconds)
      iter_seconds[k] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
}

}  // namespace

extern "C" {

int orc_max_threads() { return omp_get_max_threads(); }
void orc_set_threads(int n) { omp_set_num_threads(n); }

// findOptimalDimension (src/RL-Biggs-Andrews.cpp:18-43): largest n' <= n
// (step -1) or smallest n' >= n (step +1) that is 2,3,5,7-smooth; the primes
// 2,3,5,7 themselves count as good.  n <= 1 is returned unchanged (the
// reference recurses forever at 1).
static int orc_not_good(unsigned n) {
  if (n <= 1) return 0;
  if (n == 2 || n == 3 || n == 5 || n == 7) return 0;
  if (n % 2 == 0) return orc_not_good(n / 2);
  if (n % 3 == 0) return orc_not_good(n / 3);
  if (n % 5 == 0) return orc_not_good(n / 5);
  if (n % 7 == 0) return orc_not_good(n / 7);
  return 1;
}
unsigned orc_find_optimal_dimension(unsigned n, int step) {
  if (n <= 1) return n;
  unsigned o = n;
  while (orc_not_good(o)) o += step;
  return o;
}

// 1-D complex FFT of `howmany` contiguous lines (interleaved complex128).
void orc_fft1d_c2c_f64(double* data, int n, int howmany, int inverse) {
  Fft1<double> f(n);
  std::vector<cx<double>> work(n);
  for (int h = 0; h < howmany; ++h) f.run((cx<double>*)data + (long)h * n, work.data(), inverse != 0);
}

void orc_rfftn_f64(const double* in, int nx, int ny, int nz, double* out) {
  Fft3<double> F(nx, ny, nz);
  F.r2c(in, (cx<double>*)out);
}
void orc_irfftn_f64(const double* spec, int nx, int ny, int nz, double* out) {
  Fft3<double> F(nx, ny, nz);
  const long M = (long)nz * ny * (nx / 2 + 1);
  std::vector<cx<double>> S((const cx<double>*)spec, (const cx<double>*)spec + M);
  F.c2r(S.data(), out);
}
void orc_rfftn_f32(const float* in, int nx, int ny, int nz, float* out) {
  Fft3<float> F(nx, ny, nz);
  F.r2c(in, (cx<float>*)out);
}
void orc_irfftn_f32(const float* spec, int nx, int ny, int nz, float* out) {
  Fft3<float> F(nx, ny, nz);
  const long M = (long)nz * ny * (nx / 2 + 1);
  std::vector<cx<float>> S((const cx<float>*)spec, (const cx<float>*)spec + M);
  F.c2r(S.data(), out);
}

void orc_otf_interp_f64(const float* raw, int nxotf, int nyotf, int nzotf, int nx, int ny, int nz,
                        double kxscale, double kyscale, double kzscale, double* out) {
  otf_interp<double>(raw, nxotf, nyotf, nzotf, nx, ny, nz, kxscale, kyscale, kzscale, (cx<double>*)out);
}
void orc_otf_interp_f32(const float* raw, int nxotf, int nyotf, int nzotf, int nx, int ny, int nz,
                        float kxscale, float kyscale, float kzscale, float* out) {
  otf_interp<float>(raw, nxotf, nyotf, nzotf, nx, ny, nz, kxscale, kyscale, kzscale, (cx<float>*)out);
}

// conv(v, H) in place (one filterGPU call); conj selects conj(H).
void orc_filter_f64(double* v, int nx, int ny, int nz, const double* H, int conj) {
  Fft3<double> F(nx, ny, nz);
  std::vector<cx<double>> S((long)nz * ny * (nx / 2 + 1));
  filter<double>(v, F, (const cx<double>*)H, conj != 0, S.data());
}

void orc_rl_ba_f64(const double* raw, int nx, int ny, int nz, const double* H, int nIter,
                   const double* Y0, double* X, double* lambdas) {
  rl_ba<double>(raw, nx, ny, nz, (const cx<double>*)H, nIter, Y0, X, lambdas);
}
void orc_rl_ba_f32(const float* raw, int nx, int ny, int nz, const float* H, int nIter,
                   const float* Y0, float* X, double* lambdas) {
  rl_ba<float>(raw, nx, ny, nz, (const cx<float>*)H, nIter, Y0, X, lambdas);
}
// the same with the wall time of every iteration (the timed CPU baseline)
void orc_rl_ba_f32_timed(const float* raw, int nx, int ny, int nz, const float* H, int nIter, float* X,
                         double* lambdas, double* iter_seconds) {
  rl_ba<float>(raw, nx, ny, nz, (const cx<float>*)H, nIter, nullptr, X, lambdas, iter_seconds);
}

// ---------------------------------------------------------------------------
// Preprocessing and geometry, FP64 arithmetic unless the reference rounds to
// float at a named point (then the same rounding is applied here).
// ---------------------------------------------------------------------------

// uint16 -> float with corner crop (src/RL-Biggs-Andrews.cpp:674-677)
void orc_crop_u16(const uint16_t* raw, int nx, int ny, int nz, int onx, int ony, int onz, double* out) {
  for (int z = 0; z < onz; ++z)
    for (int y = 0; y < ony; ++y)
      for (int x = 0; x < onx; ++x)
        out[((long)z * ony + y) * onx + x] = raw[((long)z * ny + y) * nx + x];
}

// apodize_x_kernel then apodize_y_kernel (src/RLgpuImpl.cu:864-894)
void orc_apodize(double* img, int nx, int ny, int nz, int napodize) {
  if (napodize <= 0) return;
  const double PI = 3.14159265358979323846;
  for (int z = 0; z < nz; ++z) {
    double* s = img + (long)z * nx * ny;
    for (int k = 0; k < nx; ++k) {
      const float diff = (float)(((float)s[(long)(ny - 1) * nx + k] - (float)s[k]) / 2.0);
      for (int l = 0; l < napodize; ++l) {
        const float fact = (float)(diff * (1.0 - std::sin((((float)l + 0.5) / (float)napodize) * PI * 0.5)));
        s[(long)l * nx + k] = (float)((float)s[(long)l * nx + k] + fact);
        s[(long)(ny - 1 - l) * nx + k] = (float)((float)s[(long)(ny - 1 - l) * nx + k] - fact);
      }
    }
    for (int l = 0; l < ny; ++l) {
      double* r = s + (long)l * nx;
      const float diff = (float)(((float)r[nx - 1] - (float)r[0]) / 2.0);
      for (int k = 0; k < napodize; ++k) {
        const float fact = (float)(diff * (1.0 - std::sin(((k + 0.5) / (float)napodize) * PI * 0.5)));
        r[k] = (float)((float)r[k] + fact);
        r[nx - 1 - k] = (float)((float)r[nx - 1 - k] - fact);
      }
    }
  }
}

// bgsubtr_kernel (src/RLgpuImpl.cu:124-132)
void orc_bgsub(double* img, long n, double bg) {
  for (long i = 0; i < n; ++i) {
    double v = img[i] - bg;
    img[i] = v > 0 ? v : 0;
  }
}

// meanAboveBackground_GPU (src/RLgpuImpl.cu:668-718): mean (rounded to float),
// then sum and count of voxels strictly above it; returns sum / count.
double orc_mean_above(const double* img, long n) {
  double sum = 0;
  for (long i = 0; i < n; ++i) sum += img[i];
  const float mean = (float)(sum / (double)n);
  double s = 0;
  long c = 0;
  for (long i = 0; i < n; ++i)
    if ((float)img[i] > mean) { s += img[i]; ++c; }
  return s / (double)c;
}

// rescale_GPU (src/RLgpuImpl.cu:829-846): factor passed as float
void orc_rescale(double* img, long n, float scale) {
  for (long i = 0; i < n; ++i) img[i] *= (double)scale;
}

// Intended z-edge taper of zBlend_kernel (src/RLgpuImpl.cu:908-925) without
// its race: the first/last-slice difference is taken once, then slice zi and
// nz-1-zi receive +/- diff * (1 - sin((zi+0.5)/nZblend * pi/2)).
void orc_zblend(double* img, int nx, int ny, int nz, int nZblend) {
  if (nZblend <= 0) return;
  const double PI = 3.14159265358979323846;
  const long nxy = (long)nx * ny;
  for (long p = 0; p < nxy; ++p) {
    const float diff = (float)img[(long)(nz - 1) * nxy + p] - (float)img[p];
    for (int zi = 0; zi < nZblend; ++zi) {
      const float fact = (float)(diff * (1.0 - std::sin(((zi + 0.5) / (float)nZblend) * PI * 0.5)));
      img[(long)zi * nxy + p] = (float)((float)img[(long)zi * nxy + p] + fact);
      img[(long)(nz - zi - 1) * nxy + p] = (float)((float)img[(long)(nz - zi - 1) * nxy + p] - fact);
    }
  }
}

// deskew_kernel (src/geometryTransform.cu:4-31).  xin is evaluated in double
// and rounded to float, as in the reference; valid iff 0 <= xin < nx-1.
void orc_deskew(const double* in, int nx, int ny, int nz, double deskewFactor, double* out, int nxOut,
                int extraShift, double padVal) {
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int xo = 0; xo < nxOut; ++xo) {
        const float xin = (float)(((double)xo - nxOut / 2. + extraShift) - deskewFactor * ((double)z - nz / 2.) + nx / 2.);
        const long io = ((long)z * ny + y) * nxOut + xo;
        if (xin >= 0 && xin < nx - 1) {
          const float fl = std::floor(xin);
          const long ii = ((long)z * ny + y) * nx + (long)fl;
          const double off = (double)(xin - fl);
          out[io] = (1 - off) * in[ii] + off * in[ii + 1];
        } else {
          out[io] = padVal;
        }
      }
}

// rotate_kernel (src/geometryTransform.cu:50-88): float rounding of the
// centred coordinates and of xin/zin as in the reference; z-flipped output.
void orc_rotate(const double* in, int nx_in, int ny, int nz_in, const float* rot, double* out, int nx_out,
                int nz_out) {
  for (int zo = 0; zo < nz_out; ++zo)
    for (int y = 0; y < ny; ++y)
      for (int xo = 0; xo < nx_out; ++xo) {
        const float xc = (float)(xo - nx_out / 2.);
        const float zc = (float)(zo - nz_out / 2.);
        const float zin = (float)((double)(rot[0] * zc + rot[1] * xc) + nz_in / 2.);
        const float xin = (float)((double)(rot[2] * zc + rot[3] * xc) + nx_in / 2.);
        const long io = ((long)(nz_out - 1 - zo) * ny + y) * nx_out + xo;
        if (xin >= 0 && xin < nx_in - 1 && zin >= 0 && zin < nz_in - 1) {
          const float fx = std::floor(xin), fz = std::floor(zin);
          const long i00 = ((long)fz * ny + y) * nx_in + (long)fx;
          const long nxy = (long)nx_in * ny;
          const double xo_ = xin - fx, zo_ = zin - fz;
          out[io] = (1 - zo_) * ((1 - xo_) * in[i00] + xo_ * in[i00 + 1]) +
                    zo_ * ((1 - xo_) * in[i00 + nxy] + xo_ * in[i00 + nxy + 1]);
        } else {
          out[io] = 0;
        }
      }
}

// transformKernel / transformKernelRA + tex3D linear filtering with border
// addressing (src/geometryTransform.cu:183-253).  Texture sampling at
// unnormalised (tu, tv, tw) reads texels around (tu-0.5, tv-0.5, tw-0.5);
// texels outside the volume contribute 0.  The reference's 8-bit fixed-point
// texture weights are not reproduced: weights are exact here.
static double tex_sample(const double* in, int nx, int ny, int nz, double x, double y, double z) {
  const double xb = x - 0.5, yb = y - 0.5, zb = z - 0.5;
  const double fx = std::floor(xb), fy = std::floor(yb), fz = std::floor(zb);
  const double ax = xb - fx, ay = yb - fy, az = zb - fz;
  auto tex = [&](double i, double j, double k) -> double {
    if (i < 0 || j < 0 || k < 0 || i > nx - 1 || j > ny - 1 || k > nz - 1) return 0.0;
    return in[((long)k * ny + (long)j) * nx + (long)i];
  };
  return (1 - az) * ((1 - ay) * ((1 - ax) * tex(fx, fy, fz) + ax * tex(fx + 1, fy, fz)) +
                     ay * ((1 - ax) * tex(fx, fy + 1, fz) + ax * tex(fx + 1, fy + 1, fz))) +
         az * ((1 - ay) * ((1 - ax) * tex(fx, fy, fz + 1) + ax * tex(fx + 1, fy, fz + 1)) +
               ay * ((1 - ax) * tex(fx, fy + 1, fz + 1) + ax * tex(fx + 1, fy + 1, fz + 1)));
}

void orc_affine(const double* in, int nx, int ny, int nz, const float* m, double* out) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        const float u = x, v = y, w = z;
        const float tu = m[0] * u + m[1] * v + m[2] * w + m[3] + 0.5f;
        const float tv = m[4] * u + m[5] * v + m[6] * w + m[7] + 0.5f;
        const float tw = m[8] * u + m[9] * v + m[10] * w + m[11] + 0.5f;
        out[((long)z * ny + y) * nx + x] = tex_sample(in, nx, ny, nz, tu, tv, tw);
      }
}

void orc_affine_ra(const double* in, int nx, int ny, int nz, float dx, float dy, float dz, const float* m,
                   double* out) {
#pragma omp parallel for schedule(static)
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        float u = x, v = y, w = z;
        u = (float)(0.5 + (u - 0.5) * dx);
        v = (float)(0.5 + (v - 0.5) * dy);
        w = (float)(0.5 + (w - 0.5) * dz);
        float tu = m[0] * u + m[1] * v + m[2] * w + m[3];
        float tv = m[4] * u + m[5] * v + m[6] * w + m[7];
        float tw = m[8] * u + m[9] * v + m[10] * w + m[11];
        tu = (float)(0.5 + (tu - 0.5) / dx);
        tv = (float)(0.5 + (tv - 0.5) / dy);
        tw = (float)(0.5 + (tw - 0.5) / dz);
        out[((long)z * ny + y) * nx + x] = tex_sample(in, nx, ny, nz, tu, tv, tw);
      }
}

}  // extern "C"
