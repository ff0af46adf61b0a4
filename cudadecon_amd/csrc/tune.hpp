// Autotuner of the FFT pass launch shapes (tune.cpp).
#pragma once

#include <array>
#include <string>
#include <vector>

namespace cdn {

// Tune the lengths used by the given grids ({nz, ny, nx} each) and install
// the fastest variant of each (fft_set_variant).  Returns a JSON report.
std::string autotune(const std::vector<std::array<int, 3>>& grids, int reps, bool use_cache);

}  // namespace cdn
