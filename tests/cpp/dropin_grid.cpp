// Reference-style user program for the GridMap and Grid callers of the
// ray shooter: ImageFinding::find_crit and find_magnification_contour over a
// GridMap, and an adaptive Grid refined along the critical curve
// (RefineLeaves, RefineLeaf, EinsteinArea, ReInitialize).  Checks that need
// no oracle are made here (exit codes 3..); the rest is written as raw
// float64 to argv[1] for the parity test:
//   [N, GridMap rays (8 per ray),
//    n_curves, per curve: type, touches_edge, n, critical_center (2),
//      critical_area, caustic_center (2), caustic_area, n_outline,
//      critcurve x (2 n), caustic points (2 n), outline (2 n_outline),
//    n_contours, per contour: hits, n, rays (x, y: 4 n),
//    Grid: n_refined, refined ids ..., n_points, points (x, y, A, gridsize:
//      9 per point), EinsteinArea]
#include <cmath>
#include <cstdio>
#include <vector>

#include "glamer_amd.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s out.bin\n", argv[0]);
    return 2;
  }
  const size_t n = 8192;
  COSMOLOGY cosmo(0.3, 0.7, 0.7);
  std::vector<double> x(n), y(n);
  if (slsim_gen_nfw(1003, n, 1e13, 8.0, 0.5, 0.3, 0.7, 0.7, 0.0, 0.0, x.data(), y.data()))
    return 1;
  std::vector<ParticleType<double>> parts(n);
  for (size_t i = 0; i < n; ++i) {
    parts[i].x[0] = x[i];
    parts[i].x[1] = y[i];
    parts[i].x[2] = 0;
    parts[i].Mass = (float)(1e13 / n);
    parts[i].Size = 1e-4f;
    parts[i].type = 0;
  }
  long seed = -1;
  Lens lens(&seed, 2.0, cosmo);
  LensHaloParticles<ParticleType<double>> halo(parts, 0.5, cosmo, Point_2d(0, 0), false, 0);
  lens.moveinMainHalo(halo, true);
  std::vector<double> out;
  const double center[2] = {0.0, 0.0};
  const double range = 30 * arcsecTOradians;

  // 1. ImageFinding over a GridMap
  const unsigned long N = 64;
  GridMap gm(&lens, N, center, range);
  out.push_back((double)N);
  for (size_t i = 0; i < N * N; ++i) {
    const RAY r = gm.ray(i);
    out.insert(out.end(), {r.x[0], r.x[1], r.y[0], r.y[1], r.A[0], r.A[1], r.A[2], r.A[3]});
  }
  std::vector<CriticalCurve> crit;
  ImageFinding::find_crit(lens, gm, crit);
  if (crit.empty()) return 3;
  out.push_back((double)crit.size());
  for (CriticalCurve& c : crit) {
    const size_t m = c.critcurve.size();
    if (c.caustic_curve_intersecting.size() != m || c.z_source != 2.0) return 4;
    if (c.caustic_intersections != -1) return 5;
    out.insert(out.end(), {(double)(int)c.type, (double)c.touches_edge, (double)m,
                           c.critical_center[0], c.critical_center[1], c.critical_area,
                           c.caustic_center[0], c.caustic_center[1], c.caustic_area,
                           (double)c.caustic_curve_outline.size()});
    for (const RAY& r : c.critcurve) out.insert(out.end(), {r.x[0], r.x[1]});
    for (const Point_2d& p : c.caustic_curve_intersecting) out.insert(out.end(), {p[0], p[1]});
    for (const Point_2d& p : c.caustic_curve_outline) out.insert(out.end(), {p[0], p[1]});
  }
  std::vector<std::vector<RAY>> contour;
  std::vector<bool> hits;
  ImageFinding::find_magnification_contour(lens, gm, 0.0, contour, hits);
  out.push_back((double)contour.size());
  for (size_t i = 0; i < contour.size(); ++i) {
    out.push_back((double)hits[i]);
    out.push_back((double)contour[i].size());
    for (const RAY& r : contour[i]) out.insert(out.end(), {r.x[0], r.x[1], r.y[0], r.y[1]});
  }

  // 2. an adaptive Grid refined along the critical curve
  Grid grid(&lens, 25, center, range);
  if (grid.getNumberOfPoints() != 625 || grid.getInitNgrid() != 25 || grid.getNgrid_block() != 3)
    return 7;
  std::vector<Point*> cells;
  std::vector<double> refined;
  for (unsigned long i = 0; i < grid.getNumberOfPoints(); ++i) {
    Point* p = grid.point(i);
    if (std::fabs(p->invmag()) < 0.3) {
      cells.push_back(p);
      refined.push_back((double)i);
    }
  }
  const double g0 = cells.empty() ? 0 : cells[0]->gridsize;
  Point* added = grid.RefineLeaves(&lens, cells);
  if (!added || added[0].head + 625 != grid.getNumberOfPoints()) return 8;
  if (cells[0]->gridsize != g0 / 3) return 9;
  // each new point's source point carries its ray
  if (added[0].image->image != &added[0] || added[0].A[0] != added[0].image->A[0]) return 10;
  Point* corner = grid.RefineLeaf(&lens, grid.point(0));
  if (!corner || corner[0].head != 3) return 11;
  out.push_back((double)refined.size());
  out.insert(out.end(), refined.begin(), refined.end());
  out.push_back((double)grid.getNumberOfPoints());
  for (unsigned long i = 0; i < grid.getNumberOfPoints(); ++i) {
    Point* p = grid.point(i);
    out.insert(out.end(), {p->x[0], p->x[1], p->image->x[0], p->image->x[1], p->A[0], p->A[1],
                           p->A[2], p->A[3], p->gridsize});
  }
  out.push_back(grid.EinsteinArea());
  Grid fresh = grid.ReInitialize(&lens);
  if (fresh.getNumberOfPoints() != 625 || fresh.getInitRange() != grid.getInitRange()) return 12;

  FILE* f = std::fopen(argv[1], "wb");
  if (!f) return 1;
  std::fwrite(out.data(), sizeof(double), out.size(), f);
  std::fclose(f);
  std::printf("dropin_grid: %zu values, %zu critical curves, %lu grid points\n", out.size(),
              crit.size(), grid.getNumberOfPoints());
  return 0;
}
