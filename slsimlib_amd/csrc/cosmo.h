// cosmo.h -- flat LCDM distances usable on host and device.
//
// Plays the role of CosmoLib's COSMOLOGY::coorDist (used at
// MultiPlane/lens.cpp:776, 2489, 2661, 2679), which is not part of the
// reference tree.  D_C(z1, z2) = (c/H0) * int_{z1}^{z2} dz / E(z),
// E(z) = sqrt(Om (1+z)^3 + OL), evaluated by composite 8-point
// Gauss-Legendre on panels of width <= 0.125 in z.
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace slsim {

struct FlatLCDM {
  double Om, OL, h;

  __host__ __device__ inline double invE(double z) const {
    const double a = 1.0 + z;
    return 1.0 / sqrt(Om * a * a * a + OL);
  }

  __host__ __device__ inline double integral(double z1, double z2) const {
    // 8-point Gauss-Legendre nodes and weights on [-1, 1]
    const double xg[4] = {0.1834346424956498049394761, 0.5255324099163289858177390,
                          0.7966664774136267395915539, 0.9602898564975362316835609};
    const double wg[4] = {0.3626837833783619829651504, 0.3137066458778872873379622,
                          0.2223810344533744705443560, 0.1012285362903762591525314};
    const double span = z2 - z1;
    int np = (int)ceil(fabs(span) / 0.125);
    if (np < 1) np = 1;
    const double w = span / np;
    double sum = 0.0;
    for (int p = 0; p < np; ++p) {
      const double c = z1 + (p + 0.5) * w;
      const double hw = 0.5 * w;
      double s = 0.0;
      for (int k = 0; k < 4; ++k) s += wg[k] * (invE(c - hw * xg[k]) + invE(c + hw * xg[k]));
      sum += s * hw;
    }
    return sum;
  }

  // comoving distance (Mpc)
  __host__ __device__ inline double coorDist(double z1, double z2) const {
    if (z1 == z2) return 0.0;
    return (2997.92458 / h) * integral(z1, z2);
  }
};

}  // namespace slsim
