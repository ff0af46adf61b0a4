// Host code that parses user input, run under AddressSanitizer and
// UndefinedBehaviorSanitizer: the TIFF reader, the OTF table loader, the
// session geometry of RL_interface_init, the Deskew_interface width rule
// and makeOTF (libradialft), on every file named on the command line and on
// a grid of edge-case parameters.  Each file prints one line
// "<name>: read WxHxD" or "<name>: rejected: <reason>"; the process exits 0
// unless a sanitizer stops it.
#include "geometry.hpp"
#include "tiff.hpp"

#include <cmath>
#include <cstdio>
#include <exception>
#include <limits>
#include <string>

extern "C" int makeOTF(const char* const ifiles, const char* const ofiles, int lambdanm, float dz, int interpkr,
                       bool bUserBackground, float background, float NA, float NIMM, float dr, int krmax,
                       bool bDoCleanup, bool b3Dout);

static std::string base(const std::string& p) {
  const size_t s = p.find_last_of('/');
  return s == std::string::npos ? p : p.substr(s + 1);
}

int main(int argc, char** argv) {
  const char* out_otf = argc > 1 ? argv[1] : "/dev/null";
  for (int i = 2; i < argc; ++i) {
    const std::string path = argv[i];
    try {
      const cdn::TiffImage img = cdn::tiff_read(path);
      double sum = 0;
      for (float v : img.data) sum += v;  // touch every decoded sample
      std::printf("%s: read %dx%dx%d sum %.6g\n", base(path).c_str(), img.width, img.height, img.depth, sum);
    } catch (const std::exception& e) {
      std::printf("%s: rejected: %s\n", base(path).c_str(), e.what());
    }
    try {
      const cdn::OtfTable t = cdn::load_otf_file(path, 0.104f, 0.2f);
      std::printf("%s: otf %dx%dx%d\n", base(path).c_str(), t.nxotf, t.nyotf, t.nzotf);
    } catch (const std::exception& e) {
      std::printf("%s: otf rejected: %s\n", base(path).c_str(), e.what());
    }
    const int rc = makeOTF(path.c_str(), out_otf, 520, 0.2f, 10, false, 90.f, 1.25f, 1.3f, 0.104f, 0, true, false);
    std::printf("%s: makeOTF %d\n", base(path).c_str(), rc);
  }

  // session geometry over edge-case arguments: either a result or an error
  const float nan = std::numeric_limits<float>::quiet_NaN(), inf = std::numeric_limits<float>::infinity();
  const int dims[] = {-5, 0, 1, 2, 7, 11, 127, 826, 8191, 20000};
  const float sizes[] = {-0.1f, 0.f, 1e-30f, 0.104f, 0.3f, 1e30f, nan, inf};
  const float angles[] = {0.f, 31.5f, -31.5f, 90.f, 180.f, 1e-6f, 1e30f, nan, -inf};
  const int widths[] = {-3, 0, 1, 810, 1 << 30};
  long ok = 0, rejected = 0;
  for (int n : dims)
    for (float dz : sizes)
      for (float ang : angles)
        for (float rot : angles)
          for (int w : widths)
            for (int skew = 0; skew < 2; ++skew) {
              try {
                const cdn::SessionGeometry g = cdn::session_geometry(n, n + 1, n + 2, 0.104f, dz, ang, rot, w, skew);
                int nxo = 0, nzo = 0;
                if (g.rot_on) cdn::rotated_dims(g.rot, g.deskewedXdim, g.output_nz, &nxo, &nzo);
                ++ok;
              } catch (const std::exception&) {
                ++rejected;
              }
              (void)cdn::deskew_interface_width(n, dz, 0.104f, ang, w, (unsigned)std::max(n, 0));
            }
  for (unsigned n = 0; n < 20000; ++n) (void)cdn::find_optimal_dimension(n);
  std::printf("geometry: %ld ok, %ld rejected\n", ok, rejected);
  return 0;
}
