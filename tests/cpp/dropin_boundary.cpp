// Reference-style user program for the rest of the drop-in boundary:
// LensHaloParticles built on a borrowed particle array (with recentring and
// rotation), GridMap's unlensed constructor, ReInitialize, move and copy,
// and the uniform pixel maps and FITS files.  (As in the reference, the
// pointer constructor is called with MinPSize given: with seven arguments
// it and the protected constructor MakeParticleLenses uses are ambiguous.)
// Checks that need no oracle are made here (exit codes 3..); the rest is
// written as raw float64 to argv[1] for the parity test:
//   [n_particles, rotated particles (x, y, z) ..., centre of mass (3),
//    densest point (3), GridMap rays of the rotated halo (8 per ray),
//    ReInitialize'd rays (8 per ray), uniform alpha1 map (Nx Ny values)]
// The uniform map is also written as FITS to argv[2] + ".alpha1.fits".
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "glamer_amd.h"

static bool same_rays(GridMap& a, GridMap& b) {
  if (a.getNumberOfPoints() != b.getNumberOfPoints()) return false;
  for (size_t i = 0; i < a.getNumberOfPoints(); ++i) {
    RAY r = a.ray(i), s = b.ray(i);
    if (r.x[0] != s.x[0] || r.x[1] != s.x[1] || r.y[0] != s.y[0] || r.y[1] != s.y[1]) return false;
    for (int k = 0; k < 4; ++k)
      if (r.A[k] != s.A[k]) return false;
  }
  return true;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s out.bin fits_base\n", argv[0]);
    return 2;
  }
  const size_t n = 8192;
  COSMOLOGY cosmo(0.3, 0.7, 0.7);
  std::vector<double> x(n), y(n);
  if (slsim_gen_nfw(1002, n, 1e13, 8.0, 0.5, 0.3, 0.7, 0.7, 0.0, 0.0, x.data(), y.data()))
    return 1;
  std::vector<ParticleType<double>> parts(n);
  for (size_t i = 0; i < n; ++i) {
    parts[i].x[0] = x[i] + 1e-3;
    parts[i].x[1] = y[i] - 2e-3;
    parts[i].x[2] = 1e-4 * std::sin((double)i);
    parts[i].Mass = (float)(1e13 / n);
    parts[i].Size = (float)(1e-4 * (1.0 + (i % 7) / 7.0));
    parts[i].type = 0;
  }
  std::vector<double> out;
  const double center[2] = {0.0, 0.0};
  const double range = 30 * arcsecTOradians;
  const unsigned long N = 48;
  long seed = -1;

  // 1. borrowed particles equal owned particles bit for bit
  {
    std::vector<ParticleType<double>> own = parts, borrowed = parts;
    LensHaloParticles<ParticleType<double>> h1(own, 0.5, cosmo, Point_2d(0, 0), false, 0);
    LensHaloParticles<ParticleType<double>> h2(borrowed.data(), n, 0.5, cosmo, Point_2d(0, 0),
                                               false, 0, 0.0f, -1.0);
    if (!own.empty() || borrowed.size() != n) return 3;   // the vector one takes its vector
    if (h2.getN() != n || &h2[0] != borrowed.data()) return 4;   // the pointer one borrows
    Lens l1(&seed, 2.0, cosmo), l2(&seed, 2.0, cosmo);
    l1.moveinMainHalo(h1, true);
    l2.moveinMainHalo(h2, true);
    GridMap g1(&l1, N, center, range), g2(&l2, N, center, range);
    if (!same_rays(g1, g2)) return 5;
  }

  // 2. recentring a borrowed array moves only its first particle (reference)
  {
    std::vector<ParticleType<double>> b = parts;
    LensHaloParticles<ParticleType<double>> h(b.data(), n, 0.5, cosmo, Point_2d(0, 0), true, 0,
                                              0.0f);
    if (b[0].x[0] == parts[0].x[0] || b[1].x[0] != parts[1].x[0]) return 6;
    const Point_3d<> c = h.CenterOfMass();
    if (c[0] != 0.0 || c[1] != 0.0 || c[2] != 0.0) return 7;
  }

  // 3. rotation of a borrowed array; its GridMap; centre and densest point
  std::vector<ParticleType<double>> rot = parts;
  LensHaloParticles<ParticleType<double>> hr(rot.data(), n, 0.5, cosmo, Point_2d(0.3, -0.2),
                                             false, 0, 0.0f);
  out.push_back((double)n);
  for (size_t i = 0; i < n; ++i) out.insert(out.end(), {rot[i].x[0], rot[i].x[1], rot[i].x[2]});
  const Point_3d<> cm = hr.CenterOfMass(), dp = hr.DensestPoint();
  out.insert(out.end(), {cm[0], cm[1], cm[2], dp[0], dp[1], dp[2]});
  Lens lens(&seed, 2.0, cosmo);
  lens.moveinMainHalo(hr, true);
  GridMap g(&lens, N, center, range);
  for (size_t i = 0; i < g.getNumberOfPoints(); ++i) {
    RAY r = g.ray(i);
    out.insert(out.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
  }

  // 4. moves and copies keep the rays; a copy outlives its original
  GridMap moved(std::move(g));
  GridMap assigned(N, center, range);
  assigned = std::move(moved);
  if (assigned.getNumberOfPoints() != (size_t)N * N) return 8;
  GridMap* copy = new GridMap(assigned);
  if (!same_rays(*copy, assigned)) return 9;
  {
    GridMap tmp(std::move(assigned));
  }
  std::vector<Point_2d> ys(1, copy->source_point(100));
  std::vector<int> mult;
  std::list<RAY> imgs = copy->find_images(ys, mult);
  if (mult.size() != 1 || mult[0] < 1) return 10;

  // 5. ReInitialize after the source moves: the rectangular constructor's grid
  lens.ResetSourcePlane(1.5);
  GridMap re = copy->ReInitialize(&lens);
  GridMap rect(&lens, N, center, range, range);
  if (!same_rays(re, rect) || re.getNumberOfPoints() != (size_t)N * N) return 11;
  for (size_t i = 0; i < re.getNumberOfPoints(); ++i) {
    RAY r = re.ray(i);
    out.insert(out.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
  }
  delete copy;

  // 6. the unlensed grid: y = x at PositionFromIndex, A = I, no contour
  GridMap un(N, center, range);
  for (size_t i = 0; i < un.getNumberOfPoints(); ++i) {
    RAY r = un.ray(i);
    const double ex = center[0] + range * (1.0 * (i % N) / (N - 1) - 0.5);
    const double ey = center[1] + range * (1.0 * (i / N) / (N - 1) - 0.5);
    if (r.x[0] != ex || r.x[1] != ey || r.y[0] != ex || r.y[1] != ey) return 12;
    if (r.A[0] != 1 || r.A[1] != 0 || r.A[2] != 0 || r.A[3] != 1) return 13;
  }
  if (un.EinsteinArea() != 0) return 14;
  std::vector<std::vector<Point_2d>> cv;
  std::vector<bool> hb;
  un.find_magnification_contour(cv, hb, 0.5);
  if (!cv.empty()) return 15;

  // 7. uniform maps: on the grid's own geometry the uniform map is
  // writePixelMap; a coarser, shifted map goes to the parity test and FITS
  {
    PixelMap<double> a = rect.writePixelMap<double>(LensingVariable::KAPPA);
    PixelMap<double> b =
        rect.writePixelMapUniform<double>(center, N, N, LensingVariable::KAPPA);
    for (size_t i = 0; i < a.size(); ++i)
      if (a[i] != b[i]) return 16;
  }
  const double c2[2] = {2e-6, -3e-6};
  const size_t Nx = 20, Ny = 13;
  PixelMap<double> u = rect.writePixelMapUniform<double>(c2, Nx, Ny, LensingVariable::ALPHA1);
  if (u.getNx() != Nx || u.getNy() != Ny) return 17;
  for (size_t i = 0; i < u.size(); ++i) out.push_back(u[i]);
  rect.writeFitsUniform<double>(c2, Nx, Ny, LensingVariable::ALPHA1, std::string(argv[2]));
  rect.writeFitsUniform<float>(LensingVariable::INVMAG, std::string(argv[2]) + ".invmag_grid.fits");

  FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 1;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("dropin_boundary: %zu values, %zu images of one source\n", out.size(), imgs.size());
  return 0;
}
