// hipRTC compilation and launch of FFT pass kernels for lengths that are not
// precompiled.  Code objects are kept in memory per length and, when the
// cache directory ($XDG_CACHE_HOME or $HOME/.cache, /cudadecon_amd) is
// writable, on disk keyed by a hash of the kernel sources, options and length.
#include "jit.hpp"

#include "common.hpp"
#include "fft_core.hpp"

#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "jit_sources.inc"  // kJitHeaderNames, kJitHeaderText, kJitHeaderCount (generated)

namespace cdn {
namespace {

constexpr int kMaxJitLength = 16384;

const char* const kExpr[JK_COUNT] = {"cdn::k_x_r2c<%d, false>",   "cdn::k_x_c2r<%d>",
                                     "cdn::k_x_lrcore<%d>",       "cdn::k_x_update<%d>",
                                     "cdn::k_col<%d, false>",     "cdn::k_col<%d, true>",
                                     "cdn::k_z_filter<%d>",       "cdn::k_x_r2c<%d, true>",
                                     "cdn::k_col_ragged<%d, false, false>", "cdn::k_col_ragged<%d, true, false>",
                                     "cdn::k_z_ragged<%d, false>", "cdn::k_col_ragged<%d, false, true>",
                                     "cdn::k_col_ragged<%d, true, true>", "cdn::k_z_ragged<%d, true>",
                                     "cdn::k_z_filter<%d, 1>",    "cdn::k_z_ragged<%d, false, 1>",
                                     "cdn::k_z_ragged<%d, true, 1>"};
const char* const kOpts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature",
                             "-Xclang", "-packed-fp32-ops"};

struct Compiled {
  std::vector<char> code;
  std::string names[JK_COUNT];
};
struct Loaded {
  hipModule_t mod = nullptr;
  hipFunction_t fn[JK_COUNT] = {};
};

std::mutex g_mu;
std::map<std::pair<int, std::string>, Compiled> g_code;                 // (length, variant defines)
std::map<std::pair<std::pair<int, int>, std::string>, Loaded> g_loaded;  // ((device, length), defines)

std::string expr(int k, int n) {
  char b[64];
  std::snprintf(b, sizeof b, kExpr[k], n);
  return b;
}

uint64_t fnv1a(const std::string& s, uint64_t h = 1469598103934665603ULL) {
  for (unsigned char c : s) {
    h ^= c;
    h *= 1099511628211ULL;
  }
  return h;
}

std::string cache_dir() {
  const char* x = std::getenv("XDG_CACHE_HOME");
  std::string base;
  if (x && *x)
    base = x;
  else if (const char* h = std::getenv("HOME"))
    base = std::string(h) + "/.cache";
  else
    return "";
  const std::string d = base + "/cudadecon_amd";
  ::mkdir(base.c_str(), 0755);
  ::mkdir(d.c_str(), 0755);
  return d;
}

bool load_cached(const std::string& path, Compiled& c) {
  std::ifstream f(path, std::ios::binary);
  if (!f) return false;
  std::string line;
  for (int k = 0; k < JK_COUNT; ++k)
    if (!std::getline(f, c.names[k]) || c.names[k].empty()) return false;
  c.code.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  return !c.code.empty();
}

void store_cached(const std::string& path, const Compiled& c) {
  const std::string tmp = path + ".tmp" + std::to_string(::getpid());
  {
    std::ofstream f(tmp, std::ios::binary);
    if (!f) return;
    for (int k = 0; k < JK_COUNT; ++k) f << c.names[k] << '\n';
    f.write(c.code.data(), (std::streamsize)c.code.size());
    if (!f) return;
  }
  std::rename(tmp.c_str(), path.c_str());
}

#define CDN_RTC(call)                                                                     \
  do {                                                                                    \
    hiprtcResult r_ = (call);                                                             \
    if (r_ != HIPRTC_SUCCESS) fail(std::string(#call) + ": " + hiprtcGetErrorString(r_)); \
  } while (0)

std::vector<std::string> split_opts(const std::string& s) {
  std::vector<std::string> out;
  std::istringstream is(s);
  std::string w;
  while (is >> w) out.push_back(w);
  return out;
}

Compiled compile(int n, const std::string& defines) {
  const std::string src = "#include \"fft_kernels_dev.hpp\"\n";
  std::string key = src + std::to_string(n) + defines;
  for (const char* o : kOpts) key += o;
  for (const char* e : kExpr) key += e;
  for (int i = 0; i < kJitHeaderCount; ++i) key += kJitHeaderText[i];
  char hex[32];
  std::snprintf(hex, sizeof hex, "%016llx", (unsigned long long)fnv1a(key));
  const std::string dir = cache_dir();
  const std::string path = dir.empty() ? "" : dir + "/fft_" + std::to_string(n) + "_" + hex + ".hsaco";
  Compiled c;
  if (!path.empty() && load_cached(path, c)) return c;

  hiprtcProgram prog;
  CDN_RTC(hiprtcCreateProgram(&prog, src.c_str(), "cdn_fft_jit.hip", kJitHeaderCount, kJitHeaderText,
                              kJitHeaderNames));
  std::string exprs[JK_COUNT];
  for (int k = 0; k < JK_COUNT; ++k) {
    exprs[k] = expr(k, n);
    CDN_RTC(hiprtcAddNameExpression(prog, exprs[k].c_str()));
  }
  const std::vector<std::string> extra = split_opts(defines);
  std::vector<const char*> opts(std::begin(kOpts), std::end(kOpts));
  for (const std::string& e : extra) opts.push_back(e.c_str());
  const hiprtcResult rc = hiprtcCompileProgram(prog, (int)opts.size(), opts.data());
  if (rc != HIPRTC_SUCCESS) {
    size_t ls = 0;
    hiprtcGetProgramLogSize(prog, &ls);
    std::string log(ls, '\0');
    if (ls) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    fail("run-time compilation of the FFT kernels for length " + std::to_string(n) + " failed: " + log);
  }
  for (int k = 0; k < JK_COUNT; ++k) {
    const char* lowered = nullptr;
    CDN_RTC(hiprtcGetLoweredName(prog, exprs[k].c_str(), &lowered));
    c.names[k] = lowered;
  }
  size_t cs = 0;
  CDN_RTC(hiprtcGetCodeSize(prog, &cs));
  c.code.resize(cs);
  CDN_RTC(hiprtcGetCode(prog, c.code.data()));
  hiprtcDestroyProgram(&prog);
  if (!path.empty()) store_cached(path, c);
  return c;
}

}  // namespace

std::string jit_cache_dir() { return cache_dir(); }

bool jit_length_ok(int n) {
  if (n < 1 || n > kMaxJitLength) return false;
  return make_plan(n).ok;
}

std::string FftVariant::defines(int n) const {
  std::string d;
  if (cols > 0) d += " -DCDN_COLS_OVERRIDE=" + std::to_string(n) + "," + std::to_string(cols);
  if (xthreads > 0) d += " -DCDN_XTHREADS_OVERRIDE=" + std::to_string(n) + "," + std::to_string(xthreads);
  if (col_nt_load >= 0) d += " -DCDN_NT_COL_LOAD=" + std::to_string(col_nt_load);
  if (xcd_remap >= 0) d += " -DCDN_XCD_REMAP=" + std::to_string(xcd_remap);
  if (col_tpt > 0) d += " -DCDN_CTPT_OVERRIDE=" + std::to_string(n) + "," + std::to_string(col_tpt);
  if (zf_waves > 0) d += " -DCDN_ZF_WAVES=" + std::to_string(zf_waves);
  if (zf_rows != 0) d += " -DCDN_ZF_ROWS_OVERRIDE=" + std::to_string(n) + "," + std::to_string(zf_rows);
  return d;
}

void jit_prepare(int n, const FftVariant& v) {
  const std::string d = v.defines(n);
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if (g_code.count({n, d})) return;
  }
  if (!jit_length_ok(n)) fail("FFT length " + std::to_string(n) + " is not 2,3,5,7-smooth or exceeds 16384");
  // compile outside the lock, so that several variants can compile at once
  Compiled c = compile(n, d);
  std::lock_guard<std::mutex> lk(g_mu);
  g_code.emplace(std::make_pair(n, d), std::move(c));
}

void jit_launch(int n, JitKernel k, unsigned grid, unsigned block, hipStream_t st, void** args,
                const FftVariant& v) {
  jit_prepare(n, v);
  const std::string d = v.defines(n);
  int dev = 0;
  CDN_HIP(hipGetDevice(&dev));
  hipFunction_t fn;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_loaded.find({{dev, n}, d});
    if (it == g_loaded.end()) {
      const Compiled& c = g_code.at({n, d});
      Loaded L;
      CDN_HIP(hipModuleLoadData(&L.mod, c.code.data()));
      for (int j = 0; j < JK_COUNT; ++j) CDN_HIP(hipModuleGetFunction(&L.fn[j], L.mod, c.names[j].c_str()));
      it = g_loaded.emplace(std::make_pair(std::make_pair(dev, n), d), L).first;
    }
    fn = it->second.fn[k];
  }
  if (grid == 0) return;
  CDN_HIP(hipModuleLaunchKernel(fn, grid, 1, 1, block, 1, 1, 0, st, args, nullptr));
}

}  // namespace cdn
