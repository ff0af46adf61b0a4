// cudaDecon: the command-line caller of the RL path (src/linearDecon.cpp:170-1431,
// src/boostfs.cpp:15-153), driving libcudaDecon.so through its C ABI.
//
//   cudaDecon [options] input-dir filename-pattern otf-file
//
// For every TIFF stack in input-dir whose path matches the pattern (sorted by
// name; --skip N leaves out the first N; --no_overwrite those already in
// GPUdecon/), the stack is cropped to 2,3,5,7-smooth sizes about its centre,
// handed to RL_interface (preprocessing, deskew, RL-BA iterations, rotation on
// the GPU), and written to input-dir/GPUdecon/<name>_decon.tif with an ImageJ
// description; optional max-intensity projections (GPUdecon/MIPs), the
// deskewed raw stack (Deskewed/), a final crop, uint16 output, Y tiling with
// stitching, a flat start at the stack's median, bleach correction across
// files, or the one-step Wiener filter instead of RL.  The next stack is read
// while the current one is processed, and results are written by a thread of
// their own while the next is processed; ">>>file_finished" is printed after
// each file (LLSpy waits for it).
#include "../../include/cudaDecon.h"
#include "geometry.hpp"
#include "tiff.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <filesystem>
#include <future>
#include <iostream>
#include <map>
#include <regex>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

namespace fs = std::filesystem;
using cdn::TiffImage;

namespace {

struct Options {
  float drdata = 0.104f, dzdata = 0.25f, drpsf = 0.104f, dzpsf = 0.1f, wavelength = 0.525f, wiener = -1.f;
  float background = 90.f, NA = 1.2f, deskew = 0.f, rotate = 0.f, padval = 0.f;
  int napodize = 15, nzblend = 0, iters = 15, width = 0, shift = 0, skip = 0, tile = 0, tile_overlap = 30;
  bool dup_rev = false, skewed = false, save_deskewed = false, uint16 = false, no_adjust = false;
  bool flat_start = false, bleach = false, no_overwrite = false, blend_tiles = false;
  std::vector<int> crop;       // x1 x2 y1 y2 z1 z2 (inclusive)
  std::vector<int> mip, rmip;  // 3 flags each
  std::string input_dir, pattern, otf_file;
};

void usage() {
  std::cout
      << "cudaDecon " << get_version() << " (MI355X build)\n"
      << "usage: cudaDecon [options] input-dir filename-pattern otf-file\n"
      << "  --drdata v            image x-y pixel size (um), 0.104\n"
      << "  -z, --dzdata v        image z step (um), 0.25\n"
      << "  --drpsf v             PSF x-y pixel size (um), 0.104\n"
      << "  -Z, --dzpsf v         PSF z step (um), 0.1\n"
      << "  -l, --wavelength v    emission wavelength (um), 0.525\n"
      << "  -W, --wiener v        Wiener constant; > 0: one-step Wiener filter instead of RL\n"
      << "  -b, --background v    background to subtract, 90\n"
      << "  -e, --napodize n      edge pixels to soften, 15\n"
      << "  -E, --nzblend n       top and bottom sections to blend, 0\n"
      << "  -d, --dupRevStack     duplicate the reversed stack before decon\n"
      << "  -n, --NA v            numerical aperture, 1.2\n"
      << "  -i, --RL n            Richardson-Lucy iterations, 15\n"
      << "  -D, --deskew v        deskew angle (deg), 0\n"
      << "  --dcbds               deskew after decon (sample-scan PSF, 3-D OTF)\n"
      << "  --padval v            value of the deskew's padding, 0\n"
      << "  -w, --width n         width of the deskewed image, 0 (automatic)\n"
      << "  -x, --shift n         extra x shift of the deskewed image, 0\n"
      << "  -R, --rotate v        rotation about y after decon (deg), 0\n"
      << "  -S, --saveDeskewedRaw save the deskewed raw stack\n"
      << "  -C, --crop x1 x2 y1 y2 z1 z2   crop the result (inclusive bounds)\n"
      << "  -M, --MIP a b c       max-intensity projections of the result along x y z (0/1 each)\n"
      << "  -m, --rMIP a b c      the same for the deskewed raw stack\n"
      << "  -u, --uint16          save results as uint16\n"
      << "  -a, --DoNotAdjustResForFFT   keep the data size (it must then be 2,3,5,7-smooth)\n"
      << "  --FlatStart           start RL from a flat image at the stack's median\n"
      << "  -p, --bleachCorrection  rescale every stack to the first one's mean above background\n"
      << "  --skip n              skip the first n files\n"
      << "  --no_overwrite        skip files already in GPUdecon/\n"
      << "  --tile n              Y tiles of n rows (0: none; < 0: 128)\n"
      << "  --TileOverlap n       overlap between tiles, 30\n"
      << "  --BlendTileOverlap    blend the overlaps\n"
      << "  -Q, --DevQuery        list the GPUs and quit\n"
      << "  -v, --version         print the version and quit\n"
      << "  -h, --help            this message\n";
}

[[noreturn]] void bad_arg(const std::string& m) { throw std::runtime_error(m); }

float to_f(const std::string& s, const std::string& opt) {
  try {
    size_t n = 0;
    const float v = std::stof(s, &n);
    if (n != s.size()) throw std::invalid_argument(s);
    return v;
  } catch (const std::exception&) {
    bad_arg("option " + opt + " needs a number, got '" + s + "'");
  }
}
int to_i(const std::string& s, const std::string& opt) {
  try {
    size_t n = 0;
    const int v = std::stoi(s, &n);
    if (n != s.size()) throw std::invalid_argument(s);
    return v;
  } catch (const std::exception&) {
    bad_arg("option " + opt + " needs an integer, got '" + s + "'");
  }
}

// returns 0 to continue, 1 when the program should stop (help, version, query)
int parse(int argc, char** argv, Options& o) {
  std::map<std::string, float*> fl = {
      {"drdata", &o.drdata}, {"dzdata", &o.dzdata}, {"z", &o.dzdata},       {"drpsf", &o.drpsf},
      {"dzpsf", &o.dzpsf},   {"Z", &o.dzpsf},       {"wavelength", &o.wavelength}, {"l", &o.wavelength},
      {"wiener", &o.wiener}, {"W", &o.wiener},      {"background", &o.background}, {"b", &o.background},
      {"NA", &o.NA},         {"n", &o.NA},          {"deskew", &o.deskew},   {"D", &o.deskew},
      {"rotate", &o.rotate}, {"R", &o.rotate},      {"padval", &o.padval}};
  std::map<std::string, int*> in = {{"napodize", &o.napodize}, {"e", &o.napodize}, {"nzblend", &o.nzblend},
                                    {"E", &o.nzblend},         {"RL", &o.iters},    {"i", &o.iters},
                                    {"width", &o.width},       {"w", &o.width},     {"shift", &o.shift},
                                    {"x", &o.shift},           {"skip", &o.skip},   {"tile", &o.tile},
                                    {"TileOverlap", &o.tile_overlap}};
  std::map<std::string, bool*> sw = {{"dupRevStack", &o.dup_rev},
                                     {"d", &o.dup_rev},
                                     {"dcbds", &o.skewed},
                                     {"saveDeskewedRaw", &o.save_deskewed},
                                     {"S", &o.save_deskewed},
                                     {"uint16", &o.uint16},
                                     {"u", &o.uint16},
                                     {"DoNotAdjustResForFFT", &o.no_adjust},
                                     {"a", &o.no_adjust},
                                     {"FlatStart", &o.flat_start},
                                     {"bleachCorrection", &o.bleach},
                                     {"p", &o.bleach},
                                     {"no_overwrite", &o.no_overwrite},
                                     {"BlendTileOverlap", &o.blend_tiles}};
  std::map<std::string, std::pair<std::vector<int>*, int>> multi = {
      {"crop", {&o.crop, 6}}, {"C", {&o.crop, 6}}, {"MIP", {&o.mip, 3}},
      {"M", {&o.mip, 3}},     {"rMIP", {&o.rmip, 3}}, {"m", {&o.rmip, 3}}};
  std::vector<std::string> pos;
  if (argc <= 1) {
    usage();
    return 1;
  }
  for (int a = 1; a < argc; ++a) {
    std::string t = argv[a];
    if (t.size() < 2 || t[0] != '-' || (t.size() > 1 && (std::isdigit((unsigned char)t[1]) || t[1] == '.'))) {
      pos.push_back(t);
      continue;
    }
    std::string name = t[1] == '-' ? t.substr(2) : t.substr(1), val;
    const size_t eq = name.find('=');
    bool has_val = eq != std::string::npos;
    if (has_val) {
      val = name.substr(eq + 1);
      name = name.substr(0, eq);
    }
    auto next = [&]() -> std::string {
      if (has_val) {
        has_val = false;
        return val;
      }
      if (a + 1 >= argc) bad_arg("option --" + name + " needs a value");
      return argv[++a];
    };
    if (name == "h" || name == "help") {
      usage();
      return 1;
    }
    if (name == "v" || name == "version") {
      std::cout << get_version() << "\n";
      return 1;
    }
    if (name == "Q" || name == "DevQuery") {
      int n = 0;
      if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
      std::cout << "Detected " << n << " GPU(s)\n";
      for (int d = 0; d < n; ++d) {
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, d) != hipSuccess) continue;
        std::cout << "\nDevice " << d << ": \"" << p.name << "\" (" << p.gcnArchName << ")\n"
                  << "  Global memory: " << (p.totalGlobalMem >> 20) << " MB\n"
                  << "  Compute units: " << p.multiProcessorCount << "\n";
      }
      return 1;
    }
    if (fl.count(name)) {
      *fl[name] = to_f(next(), name);
    } else if (in.count(name)) {
      *in[name] = to_i(next(), name);
    } else if (sw.count(name)) {
      *sw[name] = true;
    } else if (multi.count(name)) {
      auto& m = multi[name];
      m.first->clear();
      for (int k = 0; k < m.second; ++k) m.first->push_back(to_i(next(), name));
    } else if (name == "LSC" || name == "Pad" || name == "lzw") {
      bad_arg("option --" + name + " is not supported by this build");
    } else {
      bad_arg("unknown option '" + t + "'");
    }
  }
  if (pos.size() != 3) bad_arg("expected input-dir, filename-pattern and otf-file");
  o.input_dir = pos[0];
  o.pattern = pos[1];
  o.otf_file = pos[2];
  if (o.skip < 0) bad_arg("--skip must not be negative");
  return 0;
}

// gatherMatchingFiles (src/boostfs.cpp:15-127): the pattern becomes the
// regex ".*<pattern>" (".*<pattern>.*\.[tT][iI][fF]" unless it names .tif);
// with no_overwrite a file is left out when a file in GPUdecon/ matching the
// same regex has a name that starts with the input's stem.
std::vector<std::string> gather(const std::string& dir, std::string pattern, bool no_overwrite) {
  if (pattern.rfind(".tif") == std::string::npos && pattern.rfind(".TIF") == std::string::npos)
    pattern.append(".*\\.[tT][iI][fF]");
  pattern.insert(0, ".*");
  std::cout << "file regex pattern: " << pattern << std::endl;
  const std::regex re(pattern);
  std::vector<std::string> done;
  const fs::path out = fs::path(dir) / "GPUdecon";
  if (no_overwrite && fs::exists(out))
    for (const auto& e : fs::directory_iterator(out)) {
      const std::string s = e.path().string();
      if (e.is_regular_file() && std::regex_match(s, re)) done.push_back(e.path().stem().string());
    }
  std::vector<std::string> files;
  for (const auto& e : fs::directory_iterator(dir)) {
    const std::string s = e.path().string();
    if (!e.is_regular_file() || !std::regex_match(s, re)) continue;
    if (no_overwrite) {
      const std::string stem = e.path().stem().string();
      if (std::any_of(done.begin(), done.end(), [&](const std::string& d) { return d.rfind(stem, 0) == 0; }))
        continue;
    }
    files.push_back(s);
  }
  std::sort(files.begin(), files.end());
  return files;
}

// output path: dir/subdir/<name with `insert` before its last 4 characters>
// (makeOutputFilePath, src/boostfs.cpp:140-153, for names ending in .tif)
std::string out_path(const std::string& dir, const std::string& subdir, const std::string& in,
                     const std::string& insert) {
  std::string base = fs::path(in).filename().string();
  base.insert(base.size() >= 4 ? base.size() - 4 : 0, insert);
  return (fs::path(dir) / subdir / base).string();
}

// ImageJ description (make_Image_Description, src/linearDecon.cpp:41-51)
std::string imagej(int nz, float dz) {
  std::ostringstream s;
  s << "ImageJ=1.52o\nspacing=" << std::to_string(dz) << "\nunit=micron\nimages=" << nz << "\nslices=" << nz
    << "\n";
  return s.str();
}

struct Stack {
  int nx = 0, ny = 0, nz = 0;
  std::vector<float> v;  // x fastest
  float& at(int x, int y, int z) { return v[((size_t)z * ny + y) * nx + x]; }
  float at(int x, int y, int z) const { return v[((size_t)z * ny + y) * nx + x]; }
};

Stack load_stack(const std::string& path) {
  TiffImage t = cdn::tiff_read(path);
  if (t.format == 3 || t.bits > 16)
    throw std::runtime_error(path + ": raw stacks must hold 8- or 16-bit integers (" + std::to_string(t.bits) +
                             "-bit " + (t.format == 3 ? "float" : "integer") + " found)");
  Stack s;
  s.nx = t.width;
  s.ny = t.height;
  s.nz = t.depth;
  s.v = std::move(t.data);
  return s;
}

// region [x0, x0 + nx) x [y0, y0 + ny) x [z0, z0 + nz); rows beyond the stack
// repeat its edge (CImg get_crop with boundary conditions, src/linearDecon.cpp:536-544)
Stack crop(const Stack& s, int x0, int y0, int z0, int nx, int ny, int nz) {
  Stack o;
  o.nx = nx;
  o.ny = ny;
  o.nz = nz;
  o.v.resize((size_t)nx * ny * nz);
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x)
        o.at(x, y, z) = s.at(std::clamp(x + x0, 0, s.nx - 1), std::clamp(y + y0, 0, s.ny - 1),
                             std::clamp(z + z0, 0, s.nz - 1));
  return o;
}

// max-intensity projection (MaxIntProj, src/linearDecon.cpp:1434-1473):
// axis 0 -> (ny wide, nz high), 1 -> (nx, nz), 2 -> (nx, ny)
Stack mip(const Stack& s, int axis) {
  Stack o;
  o.nz = 1;
  o.nx = axis == 0 ? s.ny : s.nx;
  o.ny = axis == 2 ? s.ny : s.nz;
  o.v.assign((size_t)o.nx * o.ny, -1e10f);
  for (int z = 0; z < s.nz; ++z)
    for (int y = 0; y < s.ny; ++y)
      for (int x = 0; x < s.nx; ++x) {
        const float v = s.at(x, y, z);
        float& m = axis == 0 ? o.at(y, z, 0) : axis == 1 ? o.at(x, z, 0) : o.at(x, y, 0);
        if (v > m) m = v;
      }
  return o;
}

// median of the stack's values (CImg median(): the mean of the two middle
// values for an even count), from a histogram of the uint16 values
float median_u16(const std::vector<uint16_t>& v) {
  std::vector<size_t> h(65536, 0);
  for (uint16_t u : v) ++h[u];
  auto kth = [&](size_t k) {
    size_t c = 0;
    for (int i = 0; i < 65536; ++i) {
      c += h[i];
      if (c > k) return (float)i;
    }
    return 65535.f;
  };
  const size_t s = v.size();
  const float a = kth(s >> 1);
  return (s % 2) ? a : (float)((a + kth((s >> 1) - 1)) / 2);
}

void save(const std::string& path, const Stack& s, bool as_u16, float dr, float dz) {
  cdn::TiffMeta m;
  m.description = imagej(s.nz, dz);
  m.pixel_x = dr;
  m.pixel_y = dr;
  if (as_u16) {
    std::vector<uint16_t> u(s.v.size());
    for (size_t i = 0; i < u.size(); ++i) u[i] = (uint16_t)std::clamp(s.v[i], 0.f, 65535.f);
    cdn::tiff_write_u16(path, u.data(), s.nx, s.ny, s.nz, &m);
  } else {
    cdn::tiff_write_f32(path, s.v.data(), s.nx, s.ny, s.nz, &m);
  }
  std::cout << "Output: " << path << "\n";
}

void check(int rc, const char* what) {
  if (!rc) throw std::runtime_error(std::string(what) + " failed: " + cudaDecon_last_error());
}

// the CLI's data size (src/linearDecon.cpp:619-666): each axis cropped to
// the nearest 2,3,5,7-smooth length below it (an odd width above 1000 one
// step further); the width only when no deskew precedes the decon
void cli_sizes(const Options& o, int nx, int ny, int nz, int* cx, int* cy, int* cz) {
  *cx = nx;
  *cy = ny;
  *cz = nz;
  if (o.iters > 0 && !o.no_adjust) {
    *cy = (int)cdn::find_optimal_dimension(ny);
    *cz = (int)cdn::find_optimal_dimension(nz);
    if (!(std::fabs(o.deskew) > 0.0) || o.skewed) {
      unsigned w = cdn::find_optimal_dimension(nx);
      if (w % 2 && w > 1000) w = cdn::find_optimal_dimension(w - 1);
      *cx = (int)w;
    }
  }
}

struct Result {
  Stack decon, deskewed;
  bool have_decon = false, have_deskewed = false;
};

// one tile (or the whole stack) through the library
Result process(const Options& o, const Stack& raw, bool reinit) {
  int cx, cy, cz;
  cli_sizes(o, raw.nx, raw.ny, raw.nz, &cx, &cy, &cz);
  // centred crop (src/linearDecon.cpp:1094-1113)
  const Stack c = crop(raw, std::max((raw.nx - cx) / 2, 0), std::max((raw.ny - cy) / 2, 0),
                       std::max((raw.nz - cz) / 2, 0), cx, cy, cz);
  if (reinit) {
    std::vector<char> otf(o.otf_file.begin(), o.otf_file.end());
    otf.push_back(0);
    check(RL_interface_init(cx, cy, cz, o.drdata, o.dzdata, o.drpsf, o.dzpsf, o.deskew, o.rotate, o.width, o.skewed,
                            otf.data()),
          "RL_interface_init");
    const unsigned ox = cudaDecon_get_session_dims(2), oy = get_output_ny(), oz = cudaDecon_get_session_dims(0);
    if ((int)ox != cx || (int)oy != cy || (int)oz != cz)
      throw std::runtime_error("stack size " + std::to_string(cx) + " x " + std::to_string(cy) + " x " +
                               std::to_string(cz) + " is not 2,3,5,7-smooth (remove -a)");
  }
  std::vector<uint16_t> u(c.v.size());
  for (size_t i = 0; i < u.size(); ++i) u[i] = (uint16_t)c.v[i];
  Result r;
  if (o.wiener > 0) {
    if (std::fabs(o.deskew) > 0.0 || std::fabs(o.rotate) > 0.0)
      throw std::runtime_error("the Wiener filter runs without deskew or rotation");
    const TiffImage otf = cdn::tiff_read(o.otf_file);
    if (otf.depth != 1) throw std::runtime_error("the Wiener filter needs a rotationally averaged OTF");
    float ks[3];
    check(cudaDecon_get_session_kscale(ks), "cudaDecon_get_session_kscale");
    r.decon = c;
    check(cudaDecon_wiener(r.decon.v.data(), cx, cy, cz, otf.data.data(), otf.width, otf.height, ks[0], ks[1], ks[2],
                           1.f / (o.drdata * cx), 1.f / (o.drdata * cy), 2 * o.NA / o.wavelength, o.wiener,
                           o.background),
          "cudaDecon_wiener");
    r.have_decon = true;
    return r;
  }
  const bool desk = std::fabs(o.deskew) > 0.0;
  if (o.iters > 0 || (o.save_deskewed && desk) || std::fabs(o.rotate) > 0.0) {
    if (o.flat_start) {
      check(cudaDecon_set_flat_start(1, median_u16(u)), "cudaDecon_set_flat_start");
    } else {
      check(cudaDecon_set_flat_start(0, 0.f), "cudaDecon_set_flat_start");
    }
    r.decon.nx = (int)get_output_nx();
    r.decon.ny = (int)get_output_ny();
    r.decon.nz = (int)get_output_nz();
    r.decon.v.resize((size_t)r.decon.nx * r.decon.ny * r.decon.nz);
    float* dptr = nullptr;
    if (o.save_deskewed && desk) {
      r.deskewed.nx = (int)cudaDecon_get_session_dims(1);
      r.deskewed.ny = cy;
      r.deskewed.nz = cz;
      r.deskewed.v.resize((size_t)r.deskewed.nx * cy * cz);
      dptr = r.deskewed.v.data();
      r.have_deskewed = true;
    }
    check(RL_interface(u.data(), cx, cy, cz, r.decon.v.data(), dptr, o.background, o.bleach, o.save_deskewed && desk,
                       o.iters, o.shift, o.napodize, o.nzblend, o.padval, o.dup_rev, o.skewed),
          "RL_interface");
    r.have_decon = true;
  } else {
    std::cerr << "Nothing is performed\n";
  }
  return r;
}

// Y-tile stitching (src/linearDecon.cpp:1177-1245)
void stitch(Stack& dst, const Stack& t, int tile_index, int ntiles, int overlap, int y_off, bool blend_tiles,
            std::vector<uint8_t>& weight) {
  const int no_zone = (int)std::ceil((overlap - 4) / 2.0);
  const int blend_zone = overlap - no_zone - no_zone;
  const bool first = tile_index == 0, last = tile_index + 1 >= ntiles;
  for (int z = 0; z < t.nz; ++z)
    for (int y = 0; y < t.ny; ++y) {
      const int yd = y + y_off;
      if (yd >= dst.ny) continue;
      float blend;
      if (first && y <= overlap)
        blend = 1;
      else if (last && y >= t.ny - overlap)
        blend = 1;
      else if (y <= no_zone)
        blend = 0;
      else if (y <= no_zone + blend_zone)
        blend = (float)((double)(y - no_zone) / blend_zone);
      else if (y >= t.ny - no_zone)
        blend = 0;
      else
        blend = 1;
      for (int x = 0; x < t.nx; ++x) {
        float& d = dst.at(x, yd, z);
        uint8_t& w = weight[((size_t)z * dst.ny + yd) * dst.nx + x];
        const float v = t.at(x, y, z);
        if (blend_tiles) {
          if (blend > 0 && blend < 1 && w > 0)
            d = v * blend + d * (1.0f - blend);
          else if (blend > 0 && w == 0)
            d = v;
          else if (blend == 1)
            d = v;
          if (blend > 0) w = 1;
        } else {
          const bool past_front = y >= std::floor((float)overlap / 2.0f) || first;
          const bool before_end = y < std::floor(t.ny - (float)overlap / 2.0f) || last;
          if (past_front && before_end) d = v;
        }
      }
    }
}

// The thread that writes one file's results while the next file is
// processed.  An error while writing is reported by the next wait(); the
// thread is joined on every way out of run(), errors included.
struct Saver {
  std::thread t;
  std::exception_ptr err;
  template <class F> void start(F&& f) {
    wait();
    t = std::thread([this, f = std::forward<F>(f)]() mutable {
      try {
        f();
      } catch (...) {
        err = std::current_exception();
      }
    });
  }
  void wait() {
    if (t.joinable()) t.join();
    if (err) {
      std::exception_ptr e = err;
      err = nullptr;
      std::rethrow_exception(e);
    }
  }
  ~Saver() {
    if (t.joinable()) t.join();
  }
};

int run(int argc, char** argv) {
  Options o;
  if (parse(argc, argv, o)) return 0;
  std::cout << "Looking for files to process... ";
  const std::vector<std::string> files = gather(o.input_dir, o.pattern, o.no_overwrite);
  std::cout << "Found " << files.size() << " file(s)." << std::endl;
  if (files.size() <= (size_t)o.skip) {
    std::cout << "\nNo files need processing!" << std::endl;
    return 0;
  }
  const bool desk = std::fabs(o.deskew) > 0.0, rot = std::fabs(o.rotate) > 0.0;
  const fs::path dir(o.input_dir);
  if (o.iters > 0 || rot || o.wiener > 0) fs::create_directories(dir / "GPUdecon");
  if (o.save_deskewed && desk) fs::create_directories(dir / "Deskewed");
  const bool mips = o.mip.size() == 3 && (o.mip[0] || o.mip[1] || o.mip[2]) && o.iters > 0;
  const bool rmips = o.rmip.size() == 3 && (o.rmip[0] || o.rmip[1] || o.rmip[2]) && o.save_deskewed && desk;
  if (mips) fs::create_directories(dir / "GPUdecon" / "MIPs");
  if (rmips) fs::create_directories(dir / "Deskewed" / "MIPs");

  const auto t_start = std::chrono::steady_clock::now();
  std::future<Stack> next = std::async(std::launch::async, load_stack, files[(size_t)o.skip]);
  Saver saver;
  int last_nx = 0, last_ny = 0, last_nz = 0;
  for (size_t f = (size_t)o.skip; f < files.size(); ++f) {
    std::cout << "\nLoading raw_image: " << f + 1 << " out of " << files.size() << ".\n" << files[f] << std::endl;
    Stack file = next.get();
    if (f + 1 < files.size()) next = std::async(std::launch::async, load_stack, files[f + 1]);
    // Y tiles (src/linearDecon.cpp:500-566)
    int ntiles = 1, tile_size = file.ny, overlap = 0;
    if (o.tile != 0) {
      tile_size = std::min(o.tile < 0 ? 128 : o.tile, file.ny);
      if (tile_size <= o.tile_overlap) throw std::runtime_error("--tile must exceed --TileOverlap");
      ntiles = (int)std::ceil((double)file.ny / ((double)tile_size - o.tile_overlap));
      if (ntiles > 1) overlap = (int)std::floor(((double)tile_size * ntiles - file.ny) / (ntiles - 1));
    }
    std::cout << "# of tiles: " << ntiles << ". Tile size: " << tile_size << " Y pix. Overlap: " << overlap
              << " Y pix." << std::endl;
    Result out;
    Stack stitched;
    std::vector<uint8_t> weight;
    for (int t = 0; t < ntiles; ++t) {
      const int y_off = (int)std::floor((double)t * (tile_size - overlap));
      const Stack tile = ntiles > 1 ? crop(file, 0, y_off, 0, file.nx, tile_size, file.nz) : Stack();
      const Stack& raw = ntiles > 1 ? tile : file;
      const bool reinit = raw.nx != last_nx || raw.ny != last_ny || raw.nz != last_nz;
      last_nx = raw.nx;
      last_ny = raw.ny;
      last_nz = raw.nz;
      Result r = process(o, raw, reinit);
      if (ntiles == 1) {
        out = std::move(r);
        break;
      }
      if (t == 0) {
        stitched.nx = r.decon.nx;
        stitched.ny = file.ny;
        stitched.nz = r.decon.nz;
        stitched.v.assign((size_t)stitched.nx * stitched.ny * stitched.nz, 0.f);
        weight.assign(stitched.v.size(), 0);
      }
      stitch(stitched, r.decon, t, ntiles, overlap, y_off, o.blend_tiles, weight);
      out.have_decon = r.have_decon;
    }
    if (ntiles > 1) out.decon = std::move(stitched);
    if (o.crop.size() == 6) {
      auto cr = [&](Stack& s) {
        const int x1 = std::clamp(o.crop[0], 0, s.nx - 1), x2 = std::clamp(o.crop[1], x1, s.nx - 1);
        const int y1 = std::clamp(o.crop[2], 0, s.ny - 1), y2 = std::clamp(o.crop[3], y1, s.ny - 1);
        const int z1 = std::clamp(o.crop[4], 0, s.nz - 1), z2 = std::clamp(o.crop[5], z1, s.nz - 1);
        s = crop(s, x1, y1, z1, x2 - x1 + 1, y2 - y1 + 1, z2 - z1 + 1);
      };
      if (out.have_decon) cr(out.decon);
      if (out.have_deskewed) cr(out.deskewed);
    }
    saver.wait();  // the previous file's writes
    // voxel sizes (src/linearDecon.cpp:1014-1022): dz of the deskewed grid,
    // dr after a rotation
    const float dz_desk = desk ? o.dzdata * std::fabs(std::sin(o.deskew * (float)M_PI / 180.f)) : o.dzdata;
    const float dz_out = rot ? o.drdata : dz_desk;
    const std::string in = files[f];
    saver.start([o, in, out = std::move(out), dz_desk, dz_out, mips, rmips, dir]() {
      const char* ax[3] = {"_MIP_x", "_MIP_y", "_MIP_z"};
      if (out.have_deskewed) {
        save(out_path(dir.string(), "Deskewed", in, "_deskewed"), out.deskewed, o.uint16, o.drdata, dz_desk);
        if (rmips)
          for (int a = 0; a < 3; ++a)
            if (o.rmip[a])
              save(out_path(dir.string(), "Deskewed/MIPs", in, ax[a]), mip(out.deskewed, a), false, o.drdata,
                   dz_desk);
      }
      if (out.have_decon && mips)
        for (int a = 0; a < 3; ++a)
          if (o.mip[a])
            save(out_path(dir.string(), "GPUdecon/MIPs", in, ax[a]), mip(out.decon, a), o.uint16, o.drdata, dz_out);
      if (out.have_decon && (o.iters > 0 || std::fabs(o.rotate) > 0.0 || o.wiener > 0))
        save(out_path(dir.string(), "GPUdecon", in, "_decon"), out.decon, o.uint16, o.drdata, dz_out);
    });
    std::printf(">>>file_finished\n");
    std::fflush(stdout);
  }
  saver.wait();
  RL_cleanup();
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
  const size_t n = files.size() - (size_t)o.skip;
  std::cout << "*** Finished! Elapsed " << s << " seconds. ";
  if (o.skip) std::cout << "Skipped " << o.skip << " images. ";
  std::cout << "Processed " << n << " images. " << s / (double)n << " seconds per image. ***" << std::endl;
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception& e) {
    std::cerr << "\n!!Error occurred: " << e.what() << std::endl;
    return 1;
  }
}
