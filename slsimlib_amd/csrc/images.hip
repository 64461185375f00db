// images.hip -- GridMap::find_images on the device.
//
// GridMap::find_images (TreeCode_link/gridmap.cpp:927-1032, 1326-1348) finds
// the images of source positions ys on a shot grid: every grid cell is split
// into the triangles (k, k+N+1, k+1) and (k, k+N+1, k+N), and a triangle whose
// source-plane image contains ys (sign tests of three cross products, edges
// and a vertex included by the reference's rules) yields an image at its
// centroid (edge midpoint, vertex) with A the mean of its three A matrices.
// The device tests every (cell, source) pair: one thread per cell keeps its
// four source-plane corners in registers and runs through the sources staged
// in LDS; hits are appended per source, then ordered on the host by (cell,
// triangle), which is the reference's order (rows, then cells, first
// triangle first), and turned into rays.  The cross products are computed
// without FMA contraction, as the reference's, so every sign decision and
// every image is the reference's to the bit.

#include <algorithm>
#include <cmath>
#include <vector>

#include "slsim_internal.h"

namespace slsim {

namespace {

// hit: cell k, code = triangle (bit 0) | case << 1 with case 0 inside,
// 1..3 on the triangle's first/second/third edge, 4 on its first vertex
struct ImageHit {
  uint32_t cell;
  uint32_t code;
};

__device__ inline int sgn(double v) { return (0.0 < v) - (v < 0.0); }

// (y - a) ^ (b - a): Point_2d's cross product (include/point.h:129-131)
__device__ inline int side(double y0, double y1, double a0, double a1, double b0, double b1) {
#pragma clang fp contract(off)
  const double u0 = y0 - a0, u1 = y1 - a1;
  const double v0 = b0 - a0, v1 = b1 - a1;
  return sgn(u0 * v1 - u1 * v0);
}

// Triangle (a, b, c) against source q, as one triangle of find_images:
// -1 no image, else 0 inside, 1..3 on the first/second/third edge, 4 on the
// first vertex (the first triangle of a cell only)
__device__ inline int tri_code(double q0, double q1, double a0, double a1, double b0, double b1,
                               double c0, double c1, bool first) {
  const int s1 = side(q0, q1, a0, a1, b0, b1);
  const int s2 = side(q0, q1, b0, b1, c0, c1);
  const int s3 = side(q0, q1, c0, c1, a0, a1);
  const int sum = s1 + s2 + s3;
  if (sum == 3 || sum == -3) return 0;
  if (sum == 2) return s1 == 0 ? 1 : (s2 == 0 ? 2 : 3);
  if (first && sum != -2 && s1 == 0 && s3 == 0) return 4;
  return -1;
}

// Images of every grid position as a source (find_boundaries_of_caustics'
// find_images over i_points, include/gridmap.h:284-288), counted per source.
// Each triangle tests only the grid positions in its source-plane bounding
// box widened by one column and row each side; positions outside it are not
// inside, on an edge or on a vertex of the triangle.  A triangle with a
// non-finite corner, whose sign tests give zeros, tests every position.
// fractional grid column / row of a position (inverse of geom_position, to
// within rounding)
__device__ inline double geom_col(const GridGeom& g, double x) {
  if (!g.rect) return (x - g.c0) * ((double)(g.nx - 1) / g.rx) + 0.5 * (double)(g.nx - 1);
  return (x - (g.c0 - 0.5 * g.rx)) * ((double)(g.nx - 1) / g.rx);
}
__device__ inline double geom_row(const GridGeom& g, double y) {
  if (!g.rect) return (y - g.c1) * ((double)(g.nx - 1) / g.rx) + 0.5 * (double)(g.nx - 1);
  return (y - (g.c1 - 0.5 * g.ry)) * ((double)(g.nx - 1) / g.rx);
}

__global__ __launch_bounds__(256) void k_count_grid_images(const double* __restrict__ y,
                                                           GridGeom g,
                                                           int32_t* __restrict__ mult) {
#pragma clang fp contract(off)
  const int64_t nx = g.nx, ny = g.ny;
  const int64_t ncell = (nx - 1) * (ny - 1);
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int64_t k = (c / (nx - 1)) * nx + c % (nx - 1);
  const double p0x = y[2 * k], p0y = y[2 * k + 1];
  const double p1x = y[2 * (k + nx + 1)], p1y = y[2 * (k + nx + 1) + 1];
  for (int t = 0; t < 2; ++t) {
    const int64_t k2 = t == 0 ? k + 1 : k + nx;
    const double p2x = y[2 * k2], p2y = y[2 * k2 + 1];
    int64_t ilo = 0, ihi = nx - 1, jlo = 0, jhi = ny - 1;
    const bool finite = isfinite(p0x) && isfinite(p0y) && isfinite(p1x) && isfinite(p1y) &&
                        isfinite(p2x) && isfinite(p2y);
    if (finite) {
      const double fx0 = geom_col(g, fmin(fmin(p0x, p1x), p2x));
      const double fx1 = geom_col(g, fmax(fmax(p0x, p1x), p2x));
      const double fy0 = geom_row(g, fmin(fmin(p0y, p1y), p2y));
      const double fy1 = geom_row(g, fmax(fmax(p0y, p1y), p2y));
      const double topx = (double)(nx - 1), topy = (double)(ny - 1);
      if (fx1 < -1.0 || fy1 < -1.0 || fx0 > topx + 1.0 || fy0 > topy + 1.0) continue;
      ilo = (int64_t)fmax(floor(fx0) - 1.0, 0.0);
      ihi = (int64_t)fmin(ceil(fx1) + 1.0, topx);
      jlo = (int64_t)fmax(floor(fy0) - 1.0, 0.0);
      jhi = (int64_t)fmin(ceil(fy1) + 1.0, topy);
    }
    for (int64_t j = jlo; j <= jhi; ++j)
      for (int64_t i = ilo; i <= ihi; ++i) {
        const int64_t s = j * nx + i;
        double q0, q1;
        geom_position(g, s, &q0, &q1);
        if (tri_code(q0, q1, p0x, p0y, p1x, p1y, p2x, p2y, t == 0) >= 0) atomicAdd(&mult[s], 1);
      }
  }
}

// Sources binned on a G x G grid over their extent (CSR by bin, row-major):
// bin (bx, by) holds sorted positions sy[start[by G + bx] .. start[by G + bx + 1])
// of sources src[...]; sources with a non-finite coordinate are listed apart.
struct SourceBins {
  const double2* sy;
  const int32_t* src;
  const int32_t* start;
  const int32_t* odd;      // non-finite sources
  int32_t nodd;
  int32_t G;
  double x0, y0, inv_w;    // inverse bin width (same both axes)
};

__device__ inline void add_hit(uint32_t* count, ImageHit* hits, uint32_t cap, int32_t s,
                               uint32_t cell, uint32_t code) {
  const uint32_t at = atomicAdd(&count[s], 1u);
  if (at < cap) hits[(int64_t)s * cap + at] = {cell, code};
}

// One thread per cell; each of its two triangles tests the sources in the
// bins its source-plane bounding box covers (widened by a bin each side); a
// source outside that box fails the sign tests unless the triangle is
// degenerate to 1e-16 of its size.  A triangle with a non-finite corner,
// whose sign tests give zeros, tests every source, and every triangle tests
// the non-finite sources, as the reference's all-pairs search does.
__global__ __launch_bounds__(256) void k_find_images(const double* __restrict__ y, int64_t nx,
                                                     int64_t ny, const double* __restrict__ ys,
                                                     int nsrc, SourceBins b, uint32_t cap,
                                                     uint32_t* __restrict__ count,
                                                     ImageHit* __restrict__ hits) {
#pragma clang fp contract(off)
  const int64_t ncell = (nx - 1) * (ny - 1);
  const int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (c >= ncell) return;
  const int64_t k = (c / (nx - 1)) * nx + c % (nx - 1);
  const double p0x = y[2 * k], p0y = y[2 * k + 1];
  const double p1x = y[2 * (k + nx + 1)], p1y = y[2 * (k + nx + 1) + 1];
  for (int t = 0; t < 2; ++t) {
    const int64_t k2 = t == 0 ? k + 1 : k + nx;
    const double p2x = y[2 * k2], p2y = y[2 * k2 + 1];
    const uint32_t code0 = (uint32_t)t;
    const bool first = t == 0;
    const bool finite = isfinite(p0x) && isfinite(p0y) && isfinite(p1x) && isfinite(p1y) &&
                        isfinite(p2x) && isfinite(p2y);
    if (!finite) {
      for (int s = 0; s < nsrc; ++s) {
        const int cd = tri_code(ys[2 * s], ys[2 * s + 1], p0x, p0y, p1x, p1y, p2x, p2y, first);
        if (cd >= 0) add_hit(count, hits, cap, s, (uint32_t)k, ((uint32_t)cd << 1) | code0);
      }
      continue;
    }
    const double top = (double)(b.G - 1);
    const double fx0 = (fmin(fmin(p0x, p1x), p2x) - b.x0) * b.inv_w;
    const double fx1 = (fmax(fmax(p0x, p1x), p2x) - b.x0) * b.inv_w;
    const double fy0 = (fmin(fmin(p0y, p1y), p2y) - b.y0) * b.inv_w;
    const double fy1 = (fmax(fmax(p0y, p1y), p2y) - b.y0) * b.inv_w;
    if (b.G > 0 && !(fx1 < -1.0 || fy1 < -1.0 || fx0 > top + 1.0 || fy0 > top + 1.0)) {
      const int bx0 = (int)fmax(floor(fx0) - 1.0, 0.0), bx1 = (int)fmin(floor(fx1) + 1.0, top);
      const int by0 = (int)fmax(floor(fy0) - 1.0, 0.0), by1 = (int)fmin(floor(fy1) + 1.0, top);
      for (int by = by0; by <= by1; ++by) {
        const int32_t e = b.start[by * b.G + bx1 + 1];
        for (int32_t q = b.start[by * b.G + bx0]; q < e; ++q) {
          const double2 sq = b.sy[q];
          const int cd = tri_code(sq.x, sq.y, p0x, p0y, p1x, p1y, p2x, p2y, first);
          if (cd >= 0) add_hit(count, hits, cap, b.src[q], (uint32_t)k, ((uint32_t)cd << 1) | code0);
        }
      }
    }
    for (int32_t q = 0; q < b.nodd; ++q) {
      const int32_t s = b.odd[q];
      const int cd = tri_code(ys[2 * s], ys[2 * s + 1], p0x, p0y, p1x, p1y, p2x, p2y, first);
      if (cd >= 0) add_hit(count, hits, cap, s, (uint32_t)k, ((uint32_t)cd << 1) | code0);
    }
  }
}

// The rays of the listed hits: x from the triangle's image-plane corners, y =
// the source, A = ((A_a + A_b) + A_c) / 3.  list[h] = (source, hit slot).
__global__ void k_image_rays(const ImageHit* __restrict__ hits, const uint32_t* __restrict__ list,
                             int64_t nhit, uint32_t cap, const double* __restrict__ A, GridGeom g,
                             const double* __restrict__ ys, double* __restrict__ rays,
                             int64_t* __restrict__ tri) {
#pragma clang fp contract(off)
  const int64_t h = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (h >= nhit) return;
  const uint32_t src = list[2 * h], slot = list[2 * h + 1];
  const ImageHit e = hits[(int64_t)src * cap + slot];
  const int64_t k = e.cell;
  const int t = e.code & 1, cs = (int)(e.code >> 1);
  const int64_t v[3] = {k, k + g.nx + 1, t == 0 ? k + 1 : k + g.nx};
  double x[3][2];
  for (int q = 0; q < 3; ++q) geom_position(g, v[q], &x[q][0], &x[q][1]);
  double px, py;
  if (cs == 0) {
    px = (x[0][0] + x[1][0] + x[2][0]) / 3;
    py = (x[0][1] + x[1][1] + x[2][1]) / 3;
  } else if (cs == 4) {
    px = x[0][0];
    py = x[0][1];
  } else {
    const int a = cs == 1 ? 0 : (cs == 2 ? 1 : 2), b = cs == 1 ? 1 : (cs == 2 ? 2 : 0);
    px = (x[a][0] + x[b][0]) / 2;
    py = (x[a][1] + x[b][1]) / 2;
  }
  double* r = rays + 8 * h;
  r[0] = px;
  r[1] = py;
  r[2] = ys[2 * src];
  r[3] = ys[2 * src + 1];
  for (int m = 0; m < 4; ++m) {
    double s = 0;
    for (int q = 0; q < 3; ++q) s = s + A[4 * v[q] + m];
    r[4 + m] = s / 3;
  }
  for (int q = 0; q < 3; ++q) tri[3 * h + q] = v[q];
}

}  // namespace

int count_grid_images(slsim_ctx c, const GridGeom& g, const double* y_dev, int32_t* mult_dev) {
  hipStream_t st = c->stream;
  SLSIM_HIP_CHECK(hipMemsetAsync(mult_dev, 0, 4 * (size_t)(g.nx * g.ny), st));
  const int64_t ncell = g.nx >= 2 && g.ny >= 2 ? (g.nx - 1) * (g.ny - 1) : 0;
  if (ncell > 0) {
    k_count_grid_images<<<(unsigned)((ncell + 255) / 256), 256, 0, st>>>(y_dev, g, mult_dev);
    SLSIM_HIP_CHECK(hipGetLastError());
  }
  return SLSIM_OK;
}

int find_images(slsim_ctx c, const GridGeom& g, const double* y_dev, const double* A_dev,
                size_t nsrc, const double* ys, size_t cap, int32_t* multiplicity, double* rays,
                int64_t* triangles, size_t* nimages) {
  *nimages = 0;
  for (size_t s = 0; s < nsrc; ++s) multiplicity[s] = 0;
  if (g.nx < 2 || g.ny < 2 || nsrc == 0) return SLSIM_OK;
  hipStream_t st = c->stream;
  Scratch sc(c);
  double* dys;
  uint32_t* dcount;
  ImageHit* dhits;
  int rc;
  if ((rc = sc.get(&dys, 2 * nsrc)) || (rc = sc.get(&dcount, nsrc)) ||
      (rc = sc.get(&dhits, nsrc * cap)))
    return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dys, ys, 16 * nsrc, hipMemcpyHostToDevice, st));
  launch_zero_u32(st, dcount, (int)nsrc);
  // bin the finite sources: about one a bin, at most 1024 x 1024 bins
  SourceBins bins{};
  std::vector<int32_t> odd, order;
  double bx0 = HUGE_VAL, bx1 = -HUGE_VAL, by0 = HUGE_VAL, by1 = -HUGE_VAL;
  for (size_t q = 0; q < nsrc; ++q) {
    const double a = ys[2 * q], b = ys[2 * q + 1];
    if (!std::isfinite(a) || !std::isfinite(b)) {
      odd.push_back((int32_t)q);
      continue;
    }
    order.push_back((int32_t)q);
    bx0 = std::min(bx0, a); bx1 = std::max(bx1, a);
    by0 = std::min(by0, b); by1 = std::max(by1, b);
  }
  const int64_t nfin = (int64_t)order.size();
  int G = 0;
  double inv_w = 0.0;
  if (nfin > 0) {
    G = (int)std::min<double>(1024.0, std::max(1.0, std::ceil(std::sqrt((double)nfin))));
    const double ext = std::max(bx1 - bx0, by1 - by0);
    inv_w = ext > 0 && std::isfinite(ext) ? (double)G / ext : 0.0;
    if (!(inv_w > 0 && std::isfinite(inv_w))) {
      G = 1;
      inv_w = 0.0;
    }
  }
  const int64_t nbins = (int64_t)G * G;
  std::vector<int32_t> bstart((size_t)nbins + 1, 0), bin((size_t)nfin);
  for (int64_t q = 0; q < nfin; ++q) {
    const int32_t s = order[q];
    const double fx = (ys[2 * s] - bx0) * inv_w, fy = (ys[2 * s + 1] - by0) * inv_w;
    const int ix = (int)std::min<double>(G - 1, std::max(0.0, std::floor(fx)));
    const int iy = (int)std::min<double>(G - 1, std::max(0.0, std::floor(fy)));
    bin[q] = iy * G + ix;
    ++bstart[bin[q] + 1];
  }
  for (int64_t q = 0; q < nbins; ++q) bstart[q + 1] += bstart[q];
  std::vector<int32_t> fill(bstart.begin(), bstart.end() - 1), bsrc((size_t)nfin);
  std::vector<double> bsy(2 * (size_t)nfin);
  for (int64_t q = 0; q < nfin; ++q) {
    const int32_t at = fill[bin[q]]++;
    bsrc[at] = order[q];
    bsy[2 * at] = ys[2 * order[q]];
    bsy[2 * at + 1] = ys[2 * order[q] + 1];
  }
  int32_t *dsrc, *dstart, *dodd;
  double* dbsy;
  if ((rc = sc.get(&dsrc, (size_t)nfin)) || (rc = sc.get(&dstart, (size_t)nbins + 1)) ||
      (rc = sc.get(&dodd, odd.size())) || (rc = sc.get(&dbsy, 2 * (size_t)nfin)))
    return rc;
  if (nfin > 0) {
    SLSIM_HIP_CHECK(hipMemcpyAsync(dsrc, bsrc.data(), 4 * nfin, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(dbsy, bsy.data(), 16 * nfin, hipMemcpyHostToDevice, st));
  }
  SLSIM_HIP_CHECK(hipMemcpyAsync(dstart, bstart.data(), 4 * (nbins + 1), hipMemcpyHostToDevice,
                                 st));
  if (!odd.empty())
    SLSIM_HIP_CHECK(hipMemcpyAsync(dodd, odd.data(), 4 * odd.size(), hipMemcpyHostToDevice, st));
  bins.sy = (const double2*)dbsy;
  bins.src = dsrc;
  bins.start = dstart;
  bins.odd = dodd;
  bins.nodd = (int32_t)odd.size();
  bins.G = G;
  bins.x0 = bx0;
  bins.y0 = by0;
  bins.inv_w = inv_w;
  const int64_t ncell = (g.nx - 1) * (g.ny - 1);
  k_find_images<<<(unsigned)((ncell + 255) / 256), 256, 0, st>>>(y_dev, g.nx, g.ny, dys,
                                                                  (int)nsrc, bins, (uint32_t)cap,
                                                                  dcount, dhits);
  SLSIM_HIP_CHECK(hipGetLastError());
  std::vector<uint32_t> cnt(nsrc);
  SLSIM_HIP_CHECK(hipMemcpyAsync(cnt.data(), dcount, 4 * nsrc, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  for (size_t s = 0; s < nsrc; ++s)
    if (cnt[s] > cap) {
      set_error("find_images: source " + std::to_string(s) + " has " + std::to_string(cnt[s]) +
                " images, more than max_per_source = " + std::to_string(cap));
      return SLSIM_ERR_ARG;
    }
  std::vector<ImageHit> hits(nsrc * cap);
  SLSIM_HIP_CHECK(hipMemcpyAsync(hits.data(), dhits, sizeof(ImageHit) * nsrc * cap,
                                 hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  // the reference's order: per source, by cell, first triangle first
  std::vector<uint32_t> list;
  for (size_t s = 0; s < nsrc; ++s) {
    std::vector<uint32_t> ord(cnt[s]);
    for (uint32_t q = 0; q < cnt[s]; ++q) ord[q] = q;
    const ImageHit* hs = hits.data() + s * cap;
    std::sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) {
      const uint64_t ka = ((uint64_t)hs[a].cell << 1) | (hs[a].code & 1);
      const uint64_t kb = ((uint64_t)hs[b].cell << 1) | (hs[b].code & 1);
      return ka < kb;
    });
    for (uint32_t q : ord) {
      list.push_back((uint32_t)s);
      list.push_back(q);
    }
    multiplicity[s] = (int32_t)cnt[s];
  }
  const int64_t nhit = (int64_t)list.size() / 2;
  *nimages = (size_t)nhit;
  if (nhit == 0) return SLSIM_OK;
  uint32_t* dlist;
  double* drays;
  int64_t* dtri;
  if ((rc = sc.get(&dlist, list.size())) || (rc = sc.get(&drays, 8 * nhit)) ||
      (rc = sc.get(&dtri, 3 * nhit)))
    return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dlist, list.data(), 4 * list.size(), hipMemcpyHostToDevice, st));
  k_image_rays<<<(unsigned)((nhit + 255) / 256), 256, 0, st>>>(dhits, dlist, nhit, (uint32_t)cap,
                                                               A_dev, g, dys, drays, dtri);
  SLSIM_HIP_CHECK(hipGetLastError());
  SLSIM_HIP_CHECK(hipMemcpyAsync(rays, drays, 64 * nhit, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(triangles, dtri, 24 * nhit, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  return SLSIM_OK;
}

}  // namespace slsim
