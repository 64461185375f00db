// Reference-style user program against the drop-in header for a light-cone
// that mixes a particle halo (z = 0.5) and a LensHaloHalos of 200 NFW halos
// placed with setTheta (z = 0.3): shoots a 64 x 64 GridMap and writes x, y
// and A of every ray as raw float64 to argv[1]; with argv[2], the images
// (GridMap::find_images) of four grid rays' source positions there; with
// argv[3], the critical curves (GridMap::find_crit) and the contours of
// GridMap::find_magnification_contour at invmag = 0.5, each curve as
// (type, hits_boundary, n, n points); with argv[4], a lens of one screened
// LensPlaneTree field plane of the same halos and 30 LensHaloPowerLaw halos
// at z = 0.9: its force at one position (5 values) and a 32 x 32 GridMap's
// x, y, A.
#include <cmath>
#include <cstdio>
#include <vector>

#include "glamer_amd.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 2;
  }
  COSMOLOGY cosmo(0.3, 0.7, 0.7);
  const size_t n = 16384;
  std::vector<double> x(n), y(n);
  if (slsim_gen_nfw(2024, n, 5e12, 8.0, 0.5, 0.3, 0.7, 0.7, 0.0, 0.0, x.data(), y.data())) return 1;
  std::vector<ParticleType<double>> particles(n);
  for (size_t i = 0; i < n; ++i) {
    particles[i].x[0] = x[i];
    particles[i].x[1] = y[i];
    particles[i].x[2] = 0.0;
    particles[i].Mass = (float)(5e12 / n);
    particles[i].Size = 2e-4f;
    particles[i].type = 0;
  }
  std::vector<LensHaloNFW> nfw;
  for (int i = 0; i < 200; ++i) {
    const float mass = (float)std::pow(10.0, 10.5 + 2.0 * ((i * 37) % 100) / 100.0);
    const float rsize = (float)(0.01 + 0.002 * (i % 17));
    const float conc = (float)(4.0 + (i % 9));
    nfw.emplace_back(mass, rsize, 0.3, conc, 1.0f, 0.0f, cosmo);
    nfw.back().setTheta(((i % 20) - 9.5) * 3e-6, ((i / 20) - 4.5) * 5e-6);
  }
  long seed = -1;
  Lens lens(&seed, 2.0, cosmo);
  LensHaloParticles<ParticleType<double>> halo(particles, 0.5, cosmo, Point_2d(0, 0), false, 0);
  lens.moveinMainHalo(halo, true);
  std::vector<LensHaloNFW> nfw2 = nfw;   // for the field plane below
  LensHaloHalos<LensHaloNFW> field(nfw, 0.3f, cosmo, 0.0f);
  lens.moveinMainHalo(field, true);
  const double center[2] = {0.0, 0.0};
  GridMap grid(&lens, 64, center, 60 * arcsecTOradians);
  std::vector<double> out;
  for (size_t i = 0; i < grid.getNumberOfPoints(); ++i) {
    RAY r = grid.ray(i);
    out.insert(out.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
  }
  FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 1;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  if (argc > 2) {
    // GridMap::find_images for the source positions of a few grid rays
    std::vector<Point_2d> ys;
    for (size_t i : {size_t(1000), size_t(2080), size_t(2112), size_t(3000)})
      ys.push_back(grid.source_point(i));
    std::vector<int> mult;
    std::list<RAY> imgs = grid.find_images(ys, mult);
    std::vector<double> w;
    for (int m : mult) w.push_back(m);
    for (const RAY& r : imgs)
      w.insert(w.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
    FILE* g = std::fopen(argv[2], "wb");
    if (!g) return 1;
    std::fwrite(w.data(), sizeof(double), w.size(), g);
    std::fclose(g);
  }
  if (argc > 3) {
    std::vector<std::vector<Point_2d>> curves, contours;
    std::vector<bool> hits, chits;
    std::vector<CritType> types;
    grid.find_crit(curves, hits, types);
    grid.find_magnification_contour(contours, chits, 0.5);
    std::vector<double> w = {(double)curves.size(), (double)contours.size()};
    for (size_t c = 0; c < curves.size(); ++c) {
      w.insert(w.end(), {(double)(int)types[c], (double)hits[c], (double)curves[c].size()});
      for (const Point_2d& p : curves[c]) w.insert(w.end(), {p[0], p[1]});
    }
    for (size_t c = 0; c < contours.size(); ++c) {
      w.insert(w.end(), {0.0, (double)chits[c], (double)contours[c].size()});
      for (const Point_2d& p : contours[c]) w.insert(w.end(), {p[0], p[1]});
    }
    FILE* g = std::fopen(argv[3], "wb");
    if (!g) return 1;
    std::fwrite(w.data(), sizeof(double), w.size(), g);
    std::fclose(g);
  }
  if (argc > 4) {
    // the NFW halos and 30 power-law halos (LensHaloPowerLaw) on one field plane
    std::vector<LensHaloPowerLaw> pls;
    for (int i = 0; i < 30; ++i) {
      pls.emplace_back((float)std::pow(10.0, 10.0 + 0.05 * i), (float)(0.01 + 0.001 * i), 0.3,
                       -0.4 + 0.07 * i, 1.0f, 0.0f, cosmo);
      pls.back().setTheta(((i % 6) - 2.5) * 7e-6, ((i / 6) - 2.0) * 6e-6);
    }
    std::vector<LensHalo*> ptrs;
    for (LensHaloNFW& h : nfw2) ptrs.push_back(&h);
    for (LensHaloPowerLaw& h : pls) ptrs.push_back(&h);
    LensPlaneTree fp(0.9f, ptrs.data(), ptrs.size(), 100.0, 2.0);
    double xx[2] = {1e-3, -2e-3}, a[2], k, g[3], ph;
    fp.force(a, &k, g, &ph, xx);
    std::vector<double> w = {a[0], a[1], k, g[0], g[1]};
    Lens lens2(&seed, 2.0, cosmo);
    lens2.insertFieldPlane(fp, 0.9);
    GridMap grid2(&lens2, 32, center, 60 * arcsecTOradians);
    for (size_t i = 0; i < grid2.getNumberOfPoints(); ++i) {
      RAY r = grid2.ray(i);
      w.insert(w.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
    }
    FILE* g4 = std::fopen(argv[4], "wb");
    if (!g4) return 1;
    std::fwrite(w.data(), sizeof(double), w.size(), g4);
    std::fclose(g4);
  }
  std::printf("dropin_halos: %zu rays, %d planes\n", out.size() / 8, lens.getNplanes());
  return 0;
}
