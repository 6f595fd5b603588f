// hit_spectrum.hpp -- the seeded spectrum of the decaying-turbulence initial
// condition, shared by the host generator (initial_hit) and the device one
// (nuc_init_hit_device), so that both synthesise the same modes.
#ifndef NUC3D_HIT_SPECTRUM_HPP
#define NUC3D_HIT_SPECTRUM_HPP
#include "nuc3d_gpu.h"

#include <complex>
#include <vector>

namespace nuc3d {

struct HitSpectrum {
    int n[3], K[3], W[3];                       // grid, highest mode number, 2K+1
    std::vector<std::complex<double>> spec[3];  // velocity component c at [(kz*W1 + ky)*W0 + kx]
    std::vector<std::complex<double>> tw[3];    // e^{+i 2 pi k x / n} at [(k+K)*n + x]
};

// Random-phase solenoidal modes with E(k) ~ (k/k0)^4 exp(-2 (k/k0)^2)
// (src/block.cpp:397-551), phases from mt19937_64(seed)
void hit_spectrum(const nuc_config &c, double k0, unsigned long long seed, HitSpectrum &S);

}  // namespace nuc3d
#endif
