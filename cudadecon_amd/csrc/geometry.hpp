// Host-side scalar logic of the RL path: FFT-friendly sizes, the session
// geometry RL_interface_init derives, output sizes, OTF k-scales and the raw
// OTF table.  Plain C++ (no HIP), so the command-line caller and the CPU
// sanitizer build use the same code as the library.
#pragma once

#include <string>
#include <vector>

namespace cdn {

// notGoodDimension / findOptimalDimension (src/RL-Biggs-Andrews.cpp:18-43):
// the nearest 2,3,5,7-smooth length stepping by `step` from n.  n <= 1 is
// returned as is (the reference recurses forever at 1).
bool not_good_dimension(unsigned n);
unsigned find_optimal_dimension(unsigned n, int step = -1);

// Raw OTF as loaded (complex64 interleaved) and determine_OTF_dimensions
// (src/RL-Biggs-Andrews.cpp:506-529).
struct OtfTable {
  std::vector<float> data;
  int nxotf = 0, nyotf = 0, nzotf = 0;
  float dkx = 0, dky = 0, dkz = 0;
  // throws if the image is too small or holds NaN / Inf values (a single
  // non-finite OTF value makes every voxel of every filtered volume NaN)
  void set_from_image(const float* img, int width, int height, int depth, float dr_psf, float dz_psf);
};

// OTF TIFF file -> table (RA: width 2 nzotf x height nr; 3-D: width
// 2 nxotf x height nyotf x nzotf pages).
OtfTable load_otf_file(const std::string& path, float dr_psf, float dz_psf);

// Sizes and matrices RL_interface_init derives from its arguments
// (src/RL-Biggs-Andrews.cpp:531-655).
struct SessionGeometry {
  unsigned output_nx = 0, output_ny = 0, output_nz = 0;  // after the corner crop
  bool crop = false;
  double deskewFactor = 0.0;
  unsigned deskewedXdim = 0;
  bool rot_on = false;
  float rot[4] = {0, 0, 0, 0};
  float dr = 0, dz = 0;  // dz after the |sin| factor when deskewing
  unsigned fft_nx = 0;   // x length of the iteration grid
};
SessionGeometry session_geometry(int nx, int ny, int nz, float dr, float dz, float deskewAngle, float rotationAngle,
                                 int outputWidth, bool bSkewedDecon);

// rotated sizes (src/RL-Biggs-Andrews.cpp:410-411), float arithmetic, with
// fabs on p[1] in both (the reference's get_output_nx omits it, :479)
void rotated_dims(const float* p, unsigned nx, unsigned nz, int* nx_o, int* nz_o);

// OTF k-scales of an iteration grid of x length fft_nx
// (src/RL-Biggs-Andrews.cpp:634-642)
void session_kscales(const SessionGeometry& g, const OtfTable& t, unsigned fft_nx, float ks[3]);

// Output width of Deskew_interface (src/RL-Biggs-Andrews.cpp:736-744): the
// caller's outputWidth, or nx + floor(output_nz dz |cos| / dr) / 4 from the
// session's output_nz when outputWidth is 0; nx with a zero angle.
unsigned deskew_interface_width(int nx, float dz, float dr, float deskewAngle, int outputWidth, unsigned output_nz);

}  // namespace cdn
