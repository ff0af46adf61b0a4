// Autotuner of the FFT pass launch shapes.
//
// Every length n of the given grids has up to three roles: x passes (n = nx),
// y column transforms (n = ny) and the z filter (n = nz).  For each length
// the tuner times, on the largest grid that uses it in each role:
//   x passes   x-R2C + LRcore + x-update with 64, 128 (default) or 256
//              threads per workgroup;
//   columns    the column FFT (y) and / or the z filter with half, the
//              default and twice the default columns per tile (within 1024
//              threads and 160 KB of LDS), and twice the columns at half the
//              threads per line; each with cached or non-temporal loads and
//              with the built-in or the other workgroup-to-XCD renumbering
//              (consecutive tiles of one line on one XCD share its L2 and its
//              translation caches); then, for a z length, the chosen shape's
//              z filter built for 6 waves per SIMD and with 4 or 6 OTF table
//              rows staged in LDS per tile;
// and installs the fastest combination (fft_set_variant).  Column variants
// only change how lines are grouped into workgroups and threads, so their
// results are bitwise those of the defaults; an x-pass variant also regroups
// the dot products' per-block partial sums (FP64, another summation order).  Candidates are compiled at run time (hipRTC),
// several at once; decisions are cached in <cache>/tune_gfx950_v2.txt keyed
// by length, role and the grid timed.
#include "tune.hpp"

#include "common.hpp"
#include "decon.hpp"
#include "fft_kernels.hpp"
#include "fft_kernels_dev.hpp"
#include "jit.hpp"
#include "rl_kernels.hpp"

#include <algorithm>
#include <array>
#include <cfloat>
#include <cstdio>
#include <fstream>
#include <map>
#include <memory>
#include <sstream>
#include <thread>

namespace cdn {
namespace {

struct Roles {
  bool x = false, y = false, z = false;
  std::array<int, 3> gx{}, gy{}, gz{};  // representative grids {nz, ny, nx}
};

long vox(const std::array<int, 3>& g) { return (long)g[0] * g[1] * g[2]; }

// LDS bytes of a column / z-filter workgroup of length n, tile width L and
// TPT threads per line (half when the exchange goes in halves)
long col_lds_bytes(int n, int L, int tpt) {
  const bool half = CDN_SPLIT_LDS != 0 && plan_exact(n, tpt);
  return (long)n * L * (half ? 4 : 8) + 8L * zf_rows(n, ZF_RA) * n;  // + the z filter's staged OTF rows
}

std::vector<FftVariant> col_candidates(int n) {
  const CShape d = c_shape(n);
  std::vector<FftVariant> out;
  // {threads per line, columns}: the built-in threads with half, the same
  // and twice the columns; half the threads (twice the elements each, up
  // to 32) with twice the columns, for a wider tile within 1024 threads
  std::vector<std::pair<int, int>> shapes{{d.TPT, d.L / 2}, {d.TPT, d.L}, {d.TPT, d.L * 2}};
  if (d.TPT % 2 == 0 && n / (d.TPT / 2) <= 32) shapes.push_back({d.TPT / 2, d.L * 2});
  for (auto [tpt, L] : shapes) {
    if (L < 1 || L * tpt > 1024 || col_lds_bytes(n, L, tpt) > 163840) continue;
    for (int nt : {0, 1})
      for (int remap : {-1, L * 8 >= 128 ? 2 : 0}) {  // the built-in renumbering and the other one
        FftVariant v;
        if (L != d.L) v.cols = L;
        if (tpt != d.TPT) v.col_tpt = tpt;
        if (nt != 0) v.col_nt_load = nt;
        v.xcd_remap = remap;
        out.push_back(v);
      }
  }
  return out;
}

std::vector<FftVariant> x_candidates(int n) {
  std::vector<FftVariant> out;
  for (int t : {64, 128, 256}) {
    const XShape xs = x_shape(n, t);
    if (xs.THREADS > 1024 || xs.THREADS < 64) continue;
    FftVariant v;
    if (t != 128) v.xthreads = t;
    // distinct when any x pass (update, LRcore, R2C) gets another workgroup
    auto threads = [&](int xt) {
      return std::array<int, 3>{x_shape(n, xt, XK_PLAIN).THREADS, x_shape(n, xt, XK_LR).THREADS,
                                x_shape(n, xt, XK_R2C).THREADS};
    };
    bool dup = false;
    for (const FftVariant& o : out) dup |= threads(o.xthreads) == threads(v.xthreads);
    if (!dup) out.push_back(v);
  }
  return out;
}

// Device buffers and plans of one grid, for timing
struct Bench {
  Fft3 F;
  OtfTable T;
  OtfDev D;
  DevBuf vol, spec, aux, part;
  hipStream_t st = nullptr;
  void setup(const std::array<int, 3>& g, hipStream_t s) {
    st = s;
    F.setup(g[2], g[1], g[0], st);
    // a rotationally averaged table of 33 radii x 17 kz samples, scaled so
    // that its support ends at 70% of the x and y Nyquist frequencies: the
    // z filter interpolates in about 40% of the columns, as with a real OTF
    const int nr = 33, nzo = 17;
    const std::vector<float> tab((size_t)2 * nzo * nr, 0.5f);
    T.set_from_image(tab.data(), 2 * nzo, nr, 1, 1.f, 1.f);
    const float ks[3] = {(nr - 1) / (0.35f * g[2]), (nr - 1) / (0.35f * g[1]), (nzo - 1) / (0.35f * g[0])};
    D.setup(T, ks, g[0], st);
    vol.ensure(sizeof(float) * F.nvox);
    aux.ensure(sizeof(float) * F.nvox);
    spec.ensure(sizeof(float2) * F.nspec);
    part.ensure(sizeof(double) * 2 * (size_t)x_update_blocks_max((long)F.ny * F.nz));
    fill(st, vol.as<float>(), F.nvox, 1.f);
    fill(st, aux.as<float>(), F.nvox, 1.f);
    fill(st, reinterpret_cast<float*>(spec.p), 2 * F.nspec, 0.f);
  }
  void run(int role) {
    const long rows = (long)F.ny * F.nz;
    float2* S = spec.as<float2>();
    if (role == 0) {
      launch_x_r2c(F.nx, st, vol.as<float>(), nullptr, nullptr, nullptr, XIN_PLAIN, S, rows, F.pitch,
                   F.twx.as<float2>());
      launch_x_lrcore(F.nx, st, S, vol.as<float>(), rows, F.pitch, F.twx.as<float2>(), FLT_EPSILON);
      launch_x_update(F.nx, st, S, vol.as<float>(), aux.as<float>(), aux.as<float>(), part.as<double>(), 1, rows,
                      F.pitch, F.twx.as<float2>());
    } else if (role == 1) {
      F.col(st, S, 1, false);
      F.col(st, S, 1, true);
    } else {
      launch_z_filter(F.nz, st, S, F.nxh, (long)F.ny * F.pitch, F.pitch, F.ny, 0, D.params(T, 1.0 / F.nvox),
                      F.twz.as<float2>());
    }
  }
};

// median of 3 trials of `reps` repetitions (ms per repetition)
double time_role(Bench& b, int role, int reps) {
  hipEvent_t e0, e1;
  CDN_HIP(hipEventCreate(&e0));
  CDN_HIP(hipEventCreate(&e1));
  b.run(role);  // warm-up (loads the module)
  std::vector<double> t;
  for (int k = 0; k < 3; ++k) {
    CDN_HIP(hipEventRecord(e0, b.st));
    for (int r = 0; r < reps; ++r) b.run(role);
    CDN_HIP(hipEventRecord(e1, b.st));
    CDN_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CDN_HIP(hipEventElapsedTime(&ms, e0, e1));
    t.push_back(ms / reps);
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::sort(t.begin(), t.end());
  return t[1];
}

std::string vstr(const FftVariant& v) {
  std::ostringstream o;
  o << "{\"cols\": " << v.cols << ", \"xthreads\": " << v.xthreads << ", \"col_nt_load\": " << v.col_nt_load
    << ", \"xcd_remap\": " << v.xcd_remap << ", \"col_tpt\": " << v.col_tpt << ", \"zf_waves\": " << v.zf_waves
    << ", \"zf_rows\": " << v.zf_rows << "}";
  return o.str();
}

std::string gkey(int n, const char* role, const std::array<int, 3>& g) {
  std::ostringstream o;
  o << n << ' ' << role << ' ' << g[0] << 'x' << g[1] << 'x' << g[2];
  return o.str();
}

std::map<std::string, FftVariant> load_cache(const std::string& path) {
  std::map<std::string, FftVariant> m;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    int n;
    std::string role, grid;
    FftVariant v;
    if (is >> n >> role >> grid >> v.cols >> v.xthreads >> v.col_nt_load >> v.xcd_remap >> v.col_tpt >> v.zf_waves >>
        v.zf_rows)
      m[std::to_string(n) + ' ' + role + ' ' + grid] = v;  // the layout gkey writes
  }
  return m;
}

void append_cache(const std::string& path, const std::string& key, const FftVariant& v) {
  std::ofstream f(path, std::ios::app);
  if (f)
    f << key << ' ' << v.cols << ' ' << v.xthreads << ' ' << v.col_nt_load << ' ' << v.xcd_remap << ' ' << v.col_tpt
      << ' ' << v.zf_waves << ' ' << v.zf_rows << '\n';
}

// compile (hipRTC, on several threads) every variant that is not built in
void prepare_parallel(const std::vector<std::pair<int, FftVariant>>& todo) {
  std::vector<std::pair<int, FftVariant>> jit;
  for (auto& t : todo)
    if (!t.second.is_default() || !fft_length_precompiled(t.first)) jit.push_back(t);
  const unsigned nth = std::max(1u, std::min(8u, std::thread::hardware_concurrency()));
  std::vector<std::thread> pool;
  std::vector<std::string> errs(nth);
  for (unsigned w = 0; w < nth; ++w)
    pool.emplace_back([&, w]() {
      try {
        for (size_t i = w; i < jit.size(); i += nth) jit_prepare(jit[i].first, jit[i].second);
      } catch (const std::exception& e) {
        errs[w] = e.what();
      }
    });
  for (auto& t : pool) t.join();
  for (auto& e : errs)
    if (!e.empty()) fail("autotune: " + e);
}

}  // namespace

std::string autotune(const std::vector<std::array<int, 3>>& grids, int reps, bool use_cache) {
  if (grids.empty()) fail("autotune: no grids");
  reps = std::max(1, reps);
  std::map<int, Roles> lengths;
  for (const auto& g : grids) {
    for (int k = 0; k < 3; ++k)
      if (!fft_length_supported(g[k])) fail("autotune: FFT length " + std::to_string(g[k]) + " is not supported");
    Roles& rx = lengths[g[2]];
    if (!rx.x || vox(g) > vox(rx.gx)) rx.gx = g;
    rx.x = true;
    Roles& ry = lengths[g[1]];
    if (!ry.y || vox(g) > vox(ry.gy)) ry.gy = g;
    ry.y = true;
    Roles& rz = lengths[g[0]];
    if (!rz.z || vox(g) > vox(rz.gz)) rz.gz = g;
    rz.z = true;
  }
  const std::string dir = jit_cache_dir();
  const std::string cpath = dir.empty() ? "" : dir + "/tune_gfx950_v2.txt";
  std::map<std::string, FftVariant> cache;
  if (use_cache && !cpath.empty()) cache = load_cache(cpath);

  // every candidate of every length, compiled up front on several threads
  std::vector<std::pair<int, FftVariant>> todo;
  for (auto& [n, r] : lengths) {
    if (r.x)
      for (const FftVariant& v : x_candidates(n)) todo.push_back({n, v});
    if (r.y || r.z)
      for (const FftVariant& v : col_candidates(n)) todo.push_back({n, v});
  }
  prepare_parallel(todo);

  hipStream_t st;
  CDN_HIP(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
  std::ostringstream rep;
  rep << "{\"lengths\": [";
  bool first = true;
  try {
    std::map<std::array<int, 3>, std::unique_ptr<Bench>> benches;
    auto bench = [&](const std::array<int, 3>& g) -> Bench& {
      auto it = benches.find(g);
      if (it == benches.end()) {
        benches.clear();  // one grid's buffers at a time
        it = benches.emplace(g, std::make_unique<Bench>()).first;
        it->second->setup(g, st);
      }
      return *it->second;
    };
    for (auto& [n, r] : lengths) {
      FftVariant best;
      rep << (first ? "" : ", ") << "{\"n\": " << n;
      first = false;
      if (r.x) {
        const std::string key = gkey(n, "x", r.gx);
        FftVariant bx;
        if (cache.count(key)) {
          bx = cache[key];
          rep << ", \"x\": {\"cached\": " << vstr(bx) << "}";
        } else {
          double tbest = 1e30;
          rep << ", \"x\": {\"grid\": [" << r.gx[0] << ", " << r.gx[1] << ", " << r.gx[2] << "], \"ms\": [";
          bool f2 = true;
          for (const FftVariant& v : x_candidates(n)) {
            fft_set_variant(n, v);
            const double t = time_role(bench(r.gx), 0, reps);
            rep << (f2 ? "" : ", ") << "[" << vstr(v) << ", " << t << "]";
            f2 = false;
            if (t < tbest * 0.99 || (v.is_default() && t <= tbest * 1.01)) {
              tbest = std::min(tbest, t);
              bx = v;
            }
          }
          rep << "]}";
          if (!cpath.empty()) append_cache(cpath, key, bx);
        }
        best.xthreads = bx.xthreads;
      }
      if (r.y || r.z) {
        const std::string key = gkey(n, r.y && r.z ? "yz" : (r.y ? "y" : "z"), r.y ? r.gy : r.gz);
        FftVariant bc;
        if (cache.count(key)) {
          bc = cache[key];
          rep << ", \"columns\": {\"cached\": " << vstr(bc) << "}";
        } else {
          double tbest = 1e30;
          rep << ", \"columns\": {\"ms\": [";
          bool f2 = true;
          for (const FftVariant& v : col_candidates(n)) {
            FftVariant w = v;
            w.xthreads = 0;
            fft_set_variant(n, w);
            double t = 0;
            if (r.y) t += time_role(bench(r.gy), 1, reps);
            if (r.z) t += time_role(bench(r.gz), 2, reps);
            rep << (f2 ? "" : ", ") << "[" << vstr(w) << ", " << t << "]";
            f2 = false;
            if (t < tbest * 0.99 || (w.is_default() && t <= tbest * 1.01)) {
              tbest = std::min(tbest, t);
              bc = w;
            }
          }
          rep << "]";
          // then the z filter of the chosen shape with a tighter register
          // budget (6 waves per SIMD: 3 workgroups of 512 threads per CU)
          // and with 4 or 6 OTF table rows staged in LDS per tile
          if (r.z) {
            const CShape cz = c_shape(n, bc.cols > 0 ? bc.cols : -1, bc.col_tpt > 0 ? bc.col_tpt : -1);
            std::vector<std::pair<int, FftVariant>> zc;
            if (bc.zf_waves == 0 && zf_waves(n, ZF_RA) < 6 && cz.THREADS == 512 && cz.E == 16) {
              FftVariant w = bc;
              w.zf_waves = 6;
              zc.push_back({n, w});
            }
            if (bc.zf_rows == 0 && zf_rows(n, ZF_RA) == 0)
              for (int rows : {4, 6}) {
                if (col_lds_bytes(n, cz.L, cz.TPT) + 8L * rows * n > 163840) continue;
                FftVariant w = bc;
                w.zf_rows = rows;
                zc.push_back({n, w});
              }
            prepare_parallel(zc);
            fft_set_variant(n, bc);
            double tb = time_role(bench(r.gz), 2, reps);
            rep << ", \"z_filter\": [[" << vstr(bc) << ", " << tb << "]";
            FftVariant b2 = bc;
            for (auto& [len, w] : zc) {
              fft_set_variant(n, w);
              const double t = time_role(bench(r.gz), 2, reps);
              rep << ", [" << vstr(w) << ", " << t << "]";
              if (t < tb * 0.99) {
                tb = t;
                b2 = w;
              }
            }
            rep << "]";
            bc = b2;
          }
          rep << "}";
          if (!cpath.empty()) append_cache(cpath, key, bc);
        }
        best.cols = bc.cols;
        best.col_nt_load = bc.col_nt_load;
        best.xcd_remap = bc.xcd_remap;
        best.col_tpt = bc.col_tpt;
        best.zf_waves = bc.zf_waves;
        best.zf_rows = bc.zf_rows;
      }
      fft_set_variant(n, best);
      if (!best.is_default()) jit_prepare(n, best);
      rep << ", \"chosen\": " << vstr(best) << "}";
    }
    benches.clear();
    CDN_HIP(hipStreamSynchronize(st));
  } catch (...) {
    (void)hipStreamDestroy(st);
    throw;
  }
  (void)hipStreamDestroy(st);
  rep << "]}";
  return rep.str();
}

}  // namespace cdn
