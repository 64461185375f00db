// Reference-style user program against the drop-in header: builds a
// projected-NFW particle halo, moves it into a Lens, shoots a GridMap and a
// list of RAYs with per-ray source redshifts, and writes the results as raw
// float64 to the file named by argv[1] (read back by the parity test); with
// argv[2], a rectangular GridMap's rays there.
#include <cstdio>
#include <vector>

#include "glamer_amd.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 2;
  }
  const size_t n = 16384;
  COSMOLOGY cosmo(0.3, 0.7, 0.7);
  std::vector<double> x(n), y(n);
  if (slsim_gen_nfw(1001, n, 1e13, 8.0, 0.5, 0.3, 0.7, 0.7, 0.0, 0.0, x.data(), y.data()))
    return 1;
  std::vector<ParticleType<double>> particles(n);
  for (size_t i = 0; i < n; ++i) {
    particles[i].x[0] = x[i];
    particles[i].x[1] = y[i];
    particles[i].x[2] = 0.0;
    particles[i].Mass = (float)(1e13 / n);
    particles[i].Size = 1e-4f;
    particles[i].type = 0;
  }
  long seed = -1;
  Lens lens(&seed, 2.0, cosmo);
  std::vector<ParticleType<double>> particles2 = particles;  // the halo takes its vector
  LensHaloParticles<ParticleType<double>> halo(particles, 0.5, cosmo, Point_2d(0, 0), false, 0);
  lens.moveinMainHalo(halo, true);

  const double center[2] = {0.0, 0.0};
  const unsigned long N = 64;
  GridMap grid(&lens, N, center, 30 * arcsecTOradians);
  std::vector<double> out;
  for (size_t i = 0; i < grid.getNumberOfPoints(); ++i) {
    RAY r = grid.ray(i);
    out.insert(out.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
  }
  PixelMap<double> kap = grid.writePixelMap<double>(LensingVariable::KAPPA);
  std::vector<RAY> rays(200);
  for (size_t i = 0; i < rays.size(); ++i) {
    rays[i].x[0] = (i % 20 - 9.5) * 1e-6;
    rays[i].x[1] = (i / 20 - 4.5) * 2e-6;
    rays[i].z = 1.0 + 0.01 * i;
  }
  lens.rayshooterInternal(rays.size(), rays.data());
  for (auto& r : rays)
    out.insert(out.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
  out.push_back(kap[0]);
  if (argc > 2) {
    // the rectangular constructor (gridmap.cpp:26-91): 40 x 24 points
    const double c2[2] = {1e-6, -2e-6};
    GridMap rect(&lens, 40, c2, 30 * arcsecTOradians, 18.5 * arcsecTOradians);
    std::vector<double> w = {(double)rect.getInitNgrid(), (double)rect.getNumberOfPoints()};
    for (size_t i = 0; i < rect.getNumberOfPoints(); ++i) {
      RAY r = rect.ray(i);
      w.insert(w.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
    }
    w.push_back(rect.EinsteinArea());
    // its magnification contours at invmag = 0.5 (count, then each as n, points)
    std::vector<std::vector<Point_2d>> cv;
    std::vector<bool> hb;
    rect.find_magnification_contour(cv, hb, 0.5);
    w.push_back((double)cv.size());
    for (auto& c : cv) {
      w.push_back((double)c.size());
      for (const Point_2d& p : c) w.insert(w.end(), {p[0], p[1]});
    }
    // deLens: no pixel left with det A < 0, y = x, and no contour below 0.5
    rect.deLens();
    std::vector<std::vector<Point_2d>> cv2;
    std::vector<bool> hb2;
    rect.find_magnification_contour(cv2, hb2, 0.5);
    w.push_back(rect.EinsteinArea());
    w.push_back((double)cv2.size());
    w.push_back(rect.ray(100).y[0] == rect.ray(100).x[0] ? 1.0 : 0.0);
    FILE* g = std::fopen(argv[2], "wb");
    if (!g) return 1;
    std::fwrite(w.data(), sizeof(double), w.size(), g);
    std::fclose(g);
  }
  // Lens accessors, then clearMainHalos: a lens with no plane shoots y = x
  if (lens.getNMainHalos() != 1 || lens.getZlens() != 0.5f ||
      lens.getAngDistLens() != cosmo.angDist((double)0.5f))
    return 3;
  lens.clearMainHalos();
  if (lens.getNMainHalos() != 0 || lens.getNplanes() != 0) return 4;
  // addplanes = false on a lens with no plane: one plane; replaceMainHalos
  // leaves one halo on one plane; clearMainHalos empties the lens again
  LensHaloParticles<ParticleType<double>> halo2(particles2, 0.5, cosmo, Point_2d(0, 0), false, 0);
  lens.insertMainHalo(halo2, false);
  if (lens.getNMainHalos() != 1 || lens.getNplanes() != 1) return 5;
  std::vector<LensHaloParticles<ParticleType<double>>> more;
  more.push_back(halo2);
  lens.replaceMainHalos(more, false);
  if (lens.getNMainHalos() != 1 || lens.getNplanes() != 1) return 6;
  lens.clearMainHalos();
  FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 1;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("dropin_gridmap: %zu rays, %d planes, kappa[0]=%.6e\n", out.size() / 8,
              lens.getNplanes(), kap[0]);
  return 0;
}
