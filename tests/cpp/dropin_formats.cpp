// Reference-style user program against the drop-in header for particle
// files: MakeParticleLenses reads the csv4 table argv[1] (sizes from the
// device k-NN, cached as glmb), CreateHalos makes the halos, insertMainHalo
// puts them in a Lens, a GridMap is shot and its kappa map is written with
// GridMap::writeFits to argv[2].  Prints the particle and halo counts.
#include <cstdio>

#include "glamer_amd.h"

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s particles.csv kappa.fits\n", argv[0]);
    return 2;
  }
  COSMOLOGY cosmo(0.3, 0.7, 0.7);
  MakeParticleLenses maker(argv[1], SimFileFormat::csv4, 32, false);
  maker.CreateHalos(cosmo, 0.5, 0.0);
  long seed = -1;
  Lens lens(&seed, 2.0, cosmo);
  for (auto* h : maker.halos) lens.insertMainHalo(*h, true);
  const double center[2] = {0.0, 0.0};
  GridMap grid(&lens, 128, center, 60 * arcsecTOradians);
  grid.writeFits<double>(LensingVariable::KAPPA, argv[2]);
  std::printf("dropin_formats: %zu particles, %zu halos, %d planes\n", maker.data.size(),
              maker.halos.size(), lens.getNplanes());
  return 0;
}
