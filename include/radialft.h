/*
 * libradialft.so - OTF generation from a measured PSF, the C entry point of
 * the reference's libradialft (src/radialft_interface.cpp:52-56, built as
 * libradialft by src/CMakeLists.txt:186-200).  Host code only (the PSF is
 * small; the reference runs it on the CPU with FFTW).
 */
#ifndef RADIALFT_AMD_H
#define RADIALFT_AMD_H

#ifdef __cplusplus
#define RFT_DEFAULT(v) = v
extern "C" {
#else
#include <stdbool.h>
#define RFT_DEFAULT(v)
#endif

/* replaces makeOTF (src/radialft_interface.cpp:52-56, body :95-196).
 * Reads the PSF stack `ifiles` (TIFF), writes the OTF TIFF `ofiles`:
 *  - b3Dout = false: rotationally averaged, width 2*nz floats x height
 *    nx/2+1 (kz fastest), optionally cleaned outside the support (NA, NIMM,
 *    wavelength, krmax) and with kr = 0 re-estimated from kr = 1..interpkr;
 *  - b3Dout = true: the full half spectrum, width (nx/2+1)*2 x ny x nz.
 * Both normalised to max |OTF| = 1.  Returns 0 on success and 1 on failure
 * (the reference's convention); radialft_last_error() gives the reason. */
int makeOTF(const char* const ifiles, const char* const ofiles, int lambdanm RFT_DEFAULT(520),
            float dz RFT_DEFAULT(0.102f), int interpkr RFT_DEFAULT(10), bool bUserBackground RFT_DEFAULT(false),
            float background RFT_DEFAULT(90), float NA RFT_DEFAULT(1.25f), float NIMM RFT_DEFAULT(1.3f),
            float dr RFT_DEFAULT(0.102f), int krmax RFT_DEFAULT(0), bool bDoCleanup RFT_DEFAULT(false),
            bool b3Dout RFT_DEFAULT(false));

/* extension: message of the last failed makeOTF call ("" if none) */
const char* radialft_last_error(void);

#ifdef __cplusplus
}
#endif
#undef RFT_DEFAULT

#endif /* RADIALFT_AMD_H */
