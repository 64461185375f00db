// crit.hip -- GridMap's critical curves and magnification contours.
//
// GridMap::find_crit (TreeCode_link/gridmap.cpp:525-647),
// find_magnification_contour (gridmap.cpp:493-522) and
// find_boundaries_of_caustics (include/gridmap.h:278-296) threshold a
// per-pixel quantity of the shot grid (the eigenvalues of A, det A, or the
// number of images of the grid positions as sources) and trace the
// boundaries of the resulting bitmaps with Utilities::find_boundaries
// (include/concave_hull.h:200-494), a marching-squares walk over the cells
// between four pixels.
//
// The device classifies every pixel in one pass over A (32 bytes a pixel,
// written back as one byte), and a second pass marks, one bit a cell packed
// by wavefront ballots, the cells whose four pixels are not all equal: the
// only cells a walk can start in or pass through.  The host walks only those
// cells, in index order, so its work is the length of the boundaries, not the
// size of the grid.  The walk is table driven: for a cell's four pixel bits
// and the direction it was entered, the direction it leaves in, which is the
// reference's branch for that cell type.

#include <algorithm>
#include <cmath>
#include <iterator>
#include <string>
#include <thread>
#include <vector>

#include "slsim_internal.h"

namespace slsim {

namespace {

enum { kByEigen = 0, kByInvmag = 1, kByMult = 2 };

// Bit 0: lambda_1 < 0 (tangential), det A < invmag, or more than one image;
// bit 1: lambda_0 < 0 and lambda_1 < 0 (radial).  Matrix2x2::eigenv
// (include/point.h:404-411), det (point.h:315-317), no contraction.  counts
// take the bits before the outermost pixels are cleared (the reference counts
// before find_boundaries pads the bitmap).  A grid-stride loop over a fixed
// number of blocks, counts summed per wavefront by ballots and per block in
// LDS: one atomic per block (one per wavefront of a full-grid launch puts
// 1e5 atomics on one address, which costs more than reading A).
constexpr int kPixBlocks = 2048;
__global__ __launch_bounds__(256) void k_crit_pixels(const double2* __restrict__ A,
                                                     const int32_t* __restrict__ mult, int64_t nx,
                                                     int64_t ny,
                                                     int mode, double invmag,
                                                     uint8_t* __restrict__ px,
                                                     unsigned long long* __restrict__ counts) {
#pragma clang fp contract(off)
  __shared__ unsigned long long part[2][4];
  const int64_t n = nx * ny;
  unsigned long long c0 = 0, c1 = 0;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < n;
       base += (int64_t)gridDim.x * blockDim.x) {
    const int64_t k = base + threadIdx.x;
    uint32_t bits = 0;
    if (k < n) {
      if (mode == kByMult) {
        bits = mult[k] > 1;
      } else {
        const double2 lo = A[2 * k], hi = A[2 * k + 1];
        const double det = lo.x * hi.y - lo.y * hi.x;
        if (mode == kByInvmag) {
          bits = det < invmag;
        } else {
          const double m = (lo.x + hi.y) / 2;
          const double s = sqrt(fabs(m * m - det));
          const double l0 = m + s, l1 = m - s;
          bits = (l1 < 0 ? 1u : 0u) | (l0 < 0 && l1 < 0 ? 2u : 0u);
        }
      }
    }
    c0 += __popcll(__ballot(bits & 1));
    c1 += __popcll(__ballot((bits >> 1) & 1));
    if (k < n) {
      const int64_t i = k % nx, j = k / nx;
      if (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) bits = 0;
      px[k] = (uint8_t)bits;
    }
  }
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    part[0][w] = c0;
    part[1][w] = c1;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long s0 = 0, s1 = 0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      s0 += part[0][q];
      s1 += part[1][q];
    }
    if (s0) atomicAdd(&counts[0], s0);
    if (s1) atomicAdd(&counts[1], s1);
  }
}

// Cells k < n - N with k % N != N - 1 (N points a row, n in all) whose four
// pixels (k, k+1, k+N, k+N+1) differ in bit b: bit k % 64 of word k / 64 of
// edge[b].
__global__ __launch_bounds__(256) void k_crit_edges(const uint8_t* __restrict__ px, int64_t N,
                                                    int64_t n,
                                                    int nbits, unsigned long long* __restrict__ edge,
                                                    int64_t nwords) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  uint32_t q = 0;
  if (k < n - N && k % N != N - 1) {
    const uint32_t a = px[k], b = px[k + 1], c = px[k + N], d = px[k + N + 1];
    q = (a | b | c | d) & ~(a & b & c & d);
  }
  for (int bit = 0; bit < nbits; ++bit) {
    const unsigned long long w = __ballot((q >> bit) & 1);
    if ((threadIdx.x & 63) == 0 && k / 64 < nwords) edge[bit * nwords + k / 64] = w;
  }
}

__global__ void k_gather_points(const double* __restrict__ y, const int64_t* __restrict__ idx,
                                int64_t n, double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t k = idx[i];
  out[2 * i] = y[2 * k];
  out[2 * i + 1] = y[2 * k + 1];
}

// ---------------------------------------------------------------- host walk
// Directions: 0 = +x (leaves through the right side), 1 = -y (bottom),
// 2 = -x (left), 3 = +y (top).  kLeave[cell bits][direction entered], the
// cell bits being pixel k = 1, k+1 = 2, k+N = 4, k+N+1 = 8 (the reference's
// types 1, 10, 100, 1000).  Cells that are all in or all out have no entry.
constexpr int8_t kLeave[16][4] = {
    {-1, -1, -1, -1}, {1, 2, 2, 2}, {0, 0, 1, 0}, {0, 2, 2, 2},   // 0, 1, 2, 3
    {3, 2, 2, 2},     {3, 1, 3, 3}, {1, 0, 3, 2}, {3, 0, 3, 3},   // 4, 5, 6, 7
    {3, 0, 3, 3},     {3, 2, 1, 0}, {3, 1, 3, 3}, {3, 2, 2, 2},   // 8, 9, 10, 11
    {0, 2, 2, 2},     {0, 0, 1, 0}, {1, 2, 2, 2}, {-1, -1, -1, -1}};
// the crossing point in the cell for each leaving direction (pixel units)
constexpr double kCrossX[4] = {1.0, 0.5, 0.0, 0.5};
constexpr double kCrossY[4] = {0.5, 0.0, 0.5, 1.0};

struct Walked {
  std::vector<int64_t> off{0};   // points of curve c: pts[off[c] .. off[c + 1])
  std::vector<double> pts;       // pixel units
  std::vector<uint8_t> hits;
  std::vector<std::vector<int64_t>> inside;   // per curve, the island fill's seed pixels
};

// Every boundary of bit `bit` of the nx x ny pixels px, curves in the order
// of their first cell.
int walk_all(const uint8_t* px, int64_t nx, int64_t ny, int bit, const unsigned long long* edge,
             int64_t nwords, bool seeds, Walked& w) {
  const int64_t n = nx * ny, N = nx;
  std::vector<uint64_t> used((size_t)nwords, 0);
  auto cell_bits = [&](int64_t k) {
    return ((px[k] >> bit) & 1) | (((px[k + 1] >> bit) & 1) << 1) |
           (((px[k + N] >> bit) & 1) << 2) | (((px[k + N + 1] >> bit) & 1) << 3);
  };
  for (int64_t wd = 0; wd < nwords; ++wd) {
    uint64_t todo = edge[wd] & ~used[wd];
    while (todo) {
      const int64_t start = wd * 64 + __builtin_ctzll(todo);
      todo &= todo - 1;
      if ((used[wd] >> (start & 63)) & 1) continue;
      bool hit = false;
      std::vector<int64_t> seed;
      int64_t k = start, steps = 0;
      int dir = 0;
      do {
        if (steps >= n) {
          set_error("find_boundaries: caught in loop");
          return SLSIM_ERR_ARG;
        }
        if (k < 0 || k >= n - N) {
          set_error("find_boundaries: walked off the grid");
          return SLSIM_ERR_ARG;
        }
        const int64_t i = k % N, j = k / N;
        hit |= i == 0 || i == N - 2 || j == 0 || j == ny - 2;
        used[k >> 6] |= 1ull << (k & 63);
        ++steps;
        const int c = cell_bits(k);
        const int out = kLeave[c][dir];
        if (out < 0) {
          set_error("find_boundaries: off edge");
          return SLSIM_ERR_ARG;
        }
        // find_islands keeps a cell's inside pixels, except in its three-in
        // cell missing pixel k+N+1 (curve_routines.cpp:2808-2813 tests the
        // type 1011 there)
        if (seeds && c != 7) {
          if (c & 1) seed.push_back(k);
          if (c & 2) seed.push_back(k + 1);
          if (c & 4) seed.push_back(k + N);
          if (c & 8) seed.push_back(k + N + 1);
        }
        w.pts.push_back((double)i + kCrossX[out]);
        w.pts.push_back((double)j + kCrossY[out]);
        k += out == 0 ? 1 : out == 1 ? -N : out == 2 ? -1 : N;
        dir = out;
      } while (k != start);
      w.off.push_back((int64_t)w.pts.size() / 2);
      w.hits.push_back(hit);
      if (seeds) w.inside.push_back(std::move(seed));
    }
  }
  return SLSIM_OK;
}

// find_islands' fill: the runs of inside pixels that start at the seeds
std::vector<int64_t> fill_island(const uint8_t* px, std::vector<int64_t> seed) {
  std::sort(seed.begin(), seed.end());
  seed.erase(std::unique(seed.begin(), seed.end()), seed.end());
  std::vector<int64_t> out;
  int64_t k = 0;
  for (int64_t s : seed) {
    if (k < s) {
      k = s;
      while (px[k] & 1) out.push_back(k++);
    }
  }
  return out;
}

// Utilities::inCurve (include/concave_hull.h:1022-1063): winding count of the
// polygon h (n points) about (x0, x1), horizontal sides skipped, vertices on
// the ray counted once where the polygon crosses it
bool in_polygon(double x0, double x1, const double* h, int64_t n) {
#pragma clang fp contract(off)
  if (n < 3) return false;
  long w = 0;
  for (int64_t i = 0; i < n; ++i) {
    const int64_t j = (i + 1) % n;
    const double hiy = h[2 * i + 1], hjy = h[2 * j + 1];
    if (hjy == hiy) continue;
    const double d1 = hiy - x1, d2 = hjy - x1;
    const int up = d1 >= 0 ? 1 : -1;
    if (d1 * d2 < 0) {
      const double e1 = h[2 * i] - x0, e2 = h[2 * j] - x0;
      if (e1 >= 0 && e2 >= 0) {
        w += up;
      } else if (e1 > 0 || e2 > 0) {
        if (e1 - (e2 - e1) * d1 / (d2 - d1) >= 0) w += up;
      }
    } else if (d2 == 0 && h[2 * j] - x0 >= 0) {
      int64_t k = (j + 1) % n;
      while (h[2 * k + 1] - x1 == 0 && k != i) k = (k + 1) % n;
      if (k != i && d1 * (h[2 * k + 1] - x1) <= 0) w += up;
    }
  }
  return w != 0;
}

double turn(const double* o, const double* a, const double* b) {
#pragma clang fp contract(off)
  return (a[0] - o[0]) * (b[1] - o[1]) - (a[1] - o[1]) * (b[0] - o[0]);
}

// Utilities::convex_hull(P, hull) (include/concave_hull.h:560-609): the
// monotone chain over the points sorted by x then y, turns <= 0 dropped;
// positions into p (2 n doubles), counter-clockwise.  The sort is std::sort
// with the reference's comparison on the same initial order, so equal points
// land where the reference's do.
std::vector<int64_t> hull_of(const double* p, int64_t n) {
  std::vector<int64_t> hull;
  if (n <= 3) {
    for (int64_t i = 0; i < n; ++i) hull.push_back(i);
    return hull;
  }
  struct Pt {
    double x, y;
    int64_t i;
  };
  std::vector<Pt> s((size_t)n);
  for (int64_t i = 0; i < n; ++i) s[i] = {p[2 * i], p[2 * i + 1], i};
  std::sort(s.begin(), s.end(), [](const Pt& a, const Pt& b) {
    if (a.x == b.x) return a.y < b.y;
    return a.x < b.x;
  });
  hull.resize(2 * (size_t)n + 1);
  int64_t k = 0;
  auto P = [&](int64_t i) { return p + 2 * i; };
  for (int64_t i = 0; i < n; ++i) {
    while (k >= 2 && turn(P(hull[k - 2]), P(hull[k - 1]), P(s[i].i)) <= 0) --k;
    hull[k++] = s[i].i;
  }
  const int64_t t = k + 1;
  for (int64_t i = n - 2; i >= 0; --i) {
    while (k >= t && turn(P(hull[k - 2]), P(hull[k - 1]), P(s[i].i)) <= 0) --k;
    hull[k++] = s[i].i;
  }
  hull.resize(k - 1);
  return hull;
}

struct Curve {
  std::vector<double> pts;
  int type;
  bool hit;
};

void to_result(const std::vector<Curve>& cs, bool replaced, slsim_curves_s* r) {
  r->off.assign(1, 0);
  r->pts.clear();
  r->types.clear();
  r->hits.clear();
  for (const Curve& c : cs) {
    r->pts.insert(r->pts.end(), c.pts.begin(), c.pts.end());
    r->off.push_back((int64_t)r->pts.size() / 2);
    r->types.push_back(c.type);
    r->hits.push_back(c.hit);
  }
  r->hits_replaced = replaced;
}

std::vector<Curve> curves_of(const Walked& w, int type) {
  std::vector<Curve> cs(w.hits.size());
  for (size_t c = 0; c < cs.size(); ++c) {
    cs[c].pts.assign(w.pts.begin() + 2 * w.off[c], w.pts.begin() + 2 * w.off[c + 1]);
    cs[c].type = type;
    cs[c].hit = w.hits[c];
  }
  return cs;
}

// the classification and the boundary cells of the grid, copied to the host
struct Classified {
  std::vector<uint8_t> px;
  std::vector<unsigned long long> edge;
  int64_t nwords = 0;
  unsigned long long count[2] = {0, 0};
};

int classify(slsim_ctx c, Scratch& sc, int64_t N, int64_t ny, int mode, const double* A_dev,
             const int32_t* mult_dev, double invmag, int nbits, Classified& out) {
  hipStream_t st = c->stream;
  const int64_t n = N * ny;
  out.nwords = (n + 63) / 64;
  uint8_t* dpx;
  unsigned long long *dedge, *dcnt;
  int rc;
  if ((rc = sc.get(&dpx, (size_t)n)) || (rc = sc.get(&dedge, (size_t)(nbits * out.nwords))) ||
      (rc = sc.get(&dcnt, 2)))
    return rc;
  launch_zero_u32(st, (uint32_t*)dcnt, 4);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  k_crit_pixels<<<(unsigned)std::min<int64_t>(blocks, kPixBlocks), 256, 0, st>>>(
      (const double2*)A_dev, mult_dev, N, ny, mode, invmag, dpx, dcnt);
  SLSIM_HIP_CHECK(hipGetLastError());
  k_crit_edges<<<blocks, 256, 0, st>>>(dpx, N, n, nbits, dedge, out.nwords);
  SLSIM_HIP_CHECK(hipGetLastError());
  out.px.resize((size_t)n);
  out.edge.resize((size_t)(nbits * out.nwords));
  SLSIM_HIP_CHECK(hipMemcpyAsync(out.px.data(), dpx, (size_t)n, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(out.edge.data(), dedge, 8 * out.edge.size(),
                                 hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(out.count, dcnt, 16, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  return SLSIM_OK;
}

// pixel units -> p * resolution + x of pixel 0 (gridmap.cpp:516-520, 591-595;
// GridMap::getResolution = x range / (Ngrid_init - 1))
void rescale(std::vector<Curve>& cs, const GridGeom& g) {
#pragma clang fp contract(off)
  const double res = g.rx / (double)(g.nx - 1);
  double x0, x1;
  geom_position(g, 0, &x0, &x1);
  for (Curve& c : cs)
    for (size_t q = 0; q < c.pts.size(); q += 2) {
      c.pts[q] = c.pts[q] * res + x0;
      c.pts[q + 1] = c.pts[q + 1] * res + x1;
    }
}

}  // namespace

int find_crit(slsim_ctx c, const GridGeom& g, const double* y_dev, const double* A_dev,
              slsim_curves_s* out) {
  const int64_t N = g.nx, ny = g.ny;
  Scratch sc(c);
  Classified cl;
  int rc;
  if ((rc = classify(c, sc, N, ny, kByEigen, A_dev, nullptr, 0.0, 2, cl))) return rc;
  std::vector<Curve> cs;
  std::vector<std::vector<int64_t>> seeds;
  bool replaced = false;
  // the two bitmaps are walked at once, the radial one on a second thread
  Walked wr;
  int rc_rad = SLSIM_OK;
  std::string err_rad;
  std::thread radial;
  if (cl.count[1] > 0)
    radial = std::thread([&] {
      rc_rad = walk_all(cl.px.data(), N, ny, 1, cl.edge.data() + cl.nwords, cl.nwords, false, wr);
      if (rc_rad) err_rad = slsim_last_error();
    });
  if (cl.count[0] > 0) {
    Walked w;
    rc = walk_all(cl.px.data(), N, ny, 0, cl.edge.data(), cl.nwords, true, w);
    if (rc == SLSIM_OK) {
      cs = curves_of(w, SLSIM_CRIT_TANGENTIAL);
      seeds = std::move(w.inside);
      replaced = true;
    }
  }
  if (radial.joinable()) radial.join();
  if (rc) return rc;
  if (rc_rad) {
    set_error(err_rad);
    return rc_rad;
  }
  const size_t ntan = cs.size();
  if (cl.count[1] > 0) {
    std::vector<Curve> rad = curves_of(wr, SLSIM_CRIT_RADIAL);
    cs.insert(cs.end(), std::make_move_iterator(rad.begin()), std::make_move_iterator(rad.end()));
  }
  // Each radial curve, in order, moves to just after the first tangential
  // curve (in list order) that holds its first point (gridmap.cpp:578-597).
  // The tangential curves keep their order and all precede the radial curve
  // being placed, so its place is after the first tangential curve in their
  // own order that holds the point, ahead of the radial curves placed there
  // before it; curves held by none stay at the end in order.  In pixel units
  // the winding test's products are exact and a point outside a polygon's
  // extent crosses it as often upward as downward, so only the tangential
  // curves whose extent holds the point are tested, found through buckets of
  // 16 x 16 pixels listing them in order.
  if (cs.size() > ntan) {
    struct Box {
      double x0, x1, y0, y1;
    };
    std::vector<Box> box(ntan);
    for (size_t i = 0; i < ntan; ++i) {
      Box b = {HUGE_VAL, -HUGE_VAL, HUGE_VAL, -HUGE_VAL};
      for (size_t q = 0; q < cs[i].pts.size(); q += 2) {
        b.x0 = std::min(b.x0, cs[i].pts[q]);
        b.x1 = std::max(b.x1, cs[i].pts[q]);
        b.y0 = std::min(b.y0, cs[i].pts[q + 1]);
        b.y1 = std::max(b.y1, cs[i].pts[q + 1]);
      }
      box[i] = b;
    }
    constexpr double kBucket = 16.0;
    const int64_t nb =
        std::max<int64_t>(1, (int64_t)std::ceil((double)std::max(N, ny) / kBucket) + 1);
    auto bucket = [&](double v) {
      return std::min<int64_t>(nb - 1, std::max<int64_t>(0, (int64_t)(v / kBucket)));
    };
    std::vector<std::vector<uint32_t>> grid((size_t)(nb * nb));
    for (size_t i = 0; i < ntan; ++i)
      for (int64_t by = bucket(box[i].y0); by <= bucket(box[i].y1); ++by)
        for (int64_t bx = bucket(box[i].x0); bx <= bucket(box[i].x1); ++bx)
          grid[by * nb + bx].push_back((uint32_t)i);
    std::vector<std::vector<size_t>> held(ntan);
    std::vector<size_t> loose;
    for (size_t j = ntan; j < cs.size(); ++j) {
      const double px = cs[j].pts[0], py = cs[j].pts[1];
      bool placed = false;
      for (uint32_t i : grid[bucket(py) * nb + bucket(px)]) {
        const Box& b = box[i];
        if (px < b.x0 || px > b.x1 || py < b.y0 || py > b.y1) continue;
        if (in_polygon(px, py, cs[i].pts.data(), (int64_t)cs[i].pts.size() / 2)) {
          held[i].push_back(j);
          placed = true;
          break;
        }
      }
      if (!placed) loose.push_back(j);
    }
    std::vector<Curve> ordered;
    ordered.reserve(cs.size());
    for (size_t i = 0; i < ntan; ++i) {
      ordered.push_back(std::move(cs[i]));
      for (auto r = held[i].rbegin(); r != held[i].rend(); ++r) ordered.push_back(std::move(cs[*r]));
    }
    for (size_t j : loose) ordered.push_back(std::move(cs[j]));
    cs = std::move(ordered);
  }
  rescale(cs, g);
  // a tangential curve with no radial curve after it gets a pseudo curve:
  // the image positions of the convex hull of its island's source positions
  // (gridmap.cpp:605-645).  Whether a curve needs one depends only on the
  // curve after it, so the islands are gathered from the device together.
  std::vector<size_t> need;                 // positions in cs
  std::vector<std::vector<int64_t>> islands;
  std::vector<int64_t> flat;
  std::vector<int64_t> start{0};
  for (size_t j = 0, itan = 0; j < cs.size(); ++j) {
    if (cs[j].type != SLSIM_CRIT_TANGENTIAL) continue;
    if (j + 1 == cs.size() || cs[j + 1].type == SLSIM_CRIT_TANGENTIAL) {
      need.push_back(j);
      islands.push_back(fill_island(cl.px.data(), seeds[itan]));
      flat.insert(flat.end(), islands.back().begin(), islands.back().end());
      start.push_back((int64_t)flat.size());
    }
    ++itan;
  }
  const int64_t m = (int64_t)flat.size();
  std::vector<double> src(2 * (size_t)m);
  if (m > 0) {
    int64_t* didx;
    double* dsrc;
    if ((rc = sc.get(&didx, (size_t)m)) || (rc = sc.get(&dsrc, 2 * (size_t)m))) return rc;
    hipStream_t st = c->stream;
    SLSIM_HIP_CHECK(hipMemcpyAsync(didx, flat.data(), 8 * m, hipMemcpyHostToDevice, st));
    k_gather_points<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(y_dev, didx, m, dsrc);
    SLSIM_HIP_CHECK(hipGetLastError());
    SLSIM_HIP_CHECK(hipMemcpyAsync(src.data(), dsrc, 16 * m, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  }
  std::vector<Curve> full;
  full.reserve(cs.size() + need.size());
  for (size_t j = 0, q = 0; j < cs.size(); ++j) {
    full.push_back(std::move(cs[j]));
    if (q < need.size() && need[q] == j) {
      const std::vector<int64_t>& isl = islands[q];
      const std::vector<int64_t> h =
          hull_of(src.data() + 2 * start[q], (int64_t)isl.size());
      Curve pc;
      pc.type = SLSIM_CRIT_PSEUDO;
      pc.hit = false;
      pc.pts.resize(2 * h.size());
      for (size_t v = 0; v < h.size(); ++v)
        geom_position(g, isl[h[v]], &pc.pts[2 * v], &pc.pts[2 * v + 1]);
      full.push_back(std::move(pc));
      ++q;
    }
  }
  cs = std::move(full);
  to_result(cs, replaced, out);
  return SLSIM_OK;
}

int find_magnification_contour(slsim_ctx c, const GridGeom& g, const double* A_dev,
                               double invmag, slsim_curves_s* out) {
  const int64_t N = g.nx, ny = g.ny;
  Scratch sc(c);
  Classified cl;
  int rc;
  if ((rc = classify(c, sc, N, ny, kByInvmag, A_dev, nullptr, invmag, 1, cl))) return rc;
  std::vector<Curve> cs;
  bool replaced = false;
  if (cl.count[0] > 0) {
    Walked w;
    if ((rc = walk_all(cl.px.data(), N, ny, 0, cl.edge.data(), cl.nwords, false, w))) return rc;
    cs = curves_of(w, 0);
    replaced = true;
  }
  rescale(cs, g);
  to_result(cs, replaced, out);
  return SLSIM_OK;
}

int find_boundaries_of_caustics(slsim_ctx c, const GridGeom& g, const double* y_dev,
                                int32_t* mult_host, slsim_curves_s* out) {
  const int64_t N = g.nx, ny = g.ny;
  Scratch sc(c);
  int32_t* dmult;
  int rc;
  if ((rc = sc.get(&dmult, (size_t)(N * ny)))) return rc;
  if ((rc = count_grid_images(c, g, y_dev, dmult))) return rc;
  if (mult_host)
    SLSIM_HIP_CHECK(hipMemcpyAsync(mult_host, dmult, 4 * (size_t)(N * ny), hipMemcpyDeviceToHost,
                                   c->stream));
  Classified cl;
  if ((rc = classify(c, sc, N, ny, kByMult, nullptr, dmult, 0.0, 1, cl))) return rc;
  Walked w;
  if ((rc = walk_all(cl.px.data(), N, ny, 0, cl.edge.data(), cl.nwords, false, w))) return rc;
  to_result(curves_of(w, 0), true, out);
  return SLSIM_OK;
}

int find_boundaries_host(const uint8_t* bitmap, size_t nx_, size_t ny_, slsim_curves_s* out) {
  const int64_t nx = (int64_t)nx_, ny = (int64_t)ny_, n = nx * ny;
  std::vector<uint8_t> px(bitmap, bitmap + n);
  for (int64_t k = 0; k < n; ++k) {
    const int64_t i = k % nx, j = k / nx;
    px[k] = (i == 0 || i == nx - 1 || j == 0 || j == ny - 1) ? 0 : (px[k] != 0);
  }
  const int64_t nwords = (n + 63) / 64;
  std::vector<unsigned long long> edge((size_t)nwords, 0);
  for (int64_t k = 0; k < n - nx; ++k) {
    if (k % nx == nx - 1) continue;
    const int s = px[k] + px[k + 1] + px[k + nx] + px[k + nx + 1];
    if (s != 0 && s != 4) edge[k >> 6] |= 1ull << (k & 63);
  }
  Walked w;
  int rc;
  if ((rc = walk_all(px.data(), nx, ny, 0, edge.data(), nwords, false, w))) return rc;
  to_result(curves_of(w, 0), true, out);
  return SLSIM_OK;
}

std::vector<int64_t> convex_hull_host(const double* p, size_t n) { return hull_of(p, (int64_t)n); }

bool in_curve_host(const double* poly, size_t n, double x0, double x1) {
  return in_polygon(x0, x1, poly, (int64_t)n);
}

// ------------------------------------------------------------ tightest hull
// Utilities::TightestHull (TreeCode_link/curve_routines.cpp:3519-3592): the
// pixels (1/100 of the curve's narrower extent, 2 R margins) that lie within
// R = 1.58 pixels of a segment of the closed curve, the outer boundary of
// that bitmap (Utilities::find_boundaries with outer_only), and each of its
// points moved to the nearest point of the curve.  Both distance passes are
// pixels (or boundary points) times segments: one thread a pixel (boundary
// point), the segments read through the cache by every thread in step.

namespace {

// distance_to_segment (include/concave_hull.h:967-1001) with the closest
// point, Point_2d's arithmetic in the reference's order, no contraction
__device__ inline double seg_dist(double px, double py, double ax, double ay, double bx,
                                  double by, double* cx, double* cy) {
#pragma clang fp contract(off)
  const double Dx = bx - ax, Dy = by - ay;
  const double s = ((px - ax) * Dx + (py - ay) * Dy) / (Dx * Dx + Dy * Dy);
  double ox, oy;
  if (isnan(s) || s <= 0) {
    ox = ax;
    oy = ay;
  } else if (s >= 1) {
    ox = bx;
    oy = by;
  } else {
    ox = ax + Dx * s;
    oy = ay + Dy * s;
  }
  *cx = ox;
  *cy = oy;
  const double ex = ox - px, ey = oy - py;
  return sqrt(ex * ex + ey * ey);
}

// pixel (i, j) at (xs[i], ys[j]) is set when some segment (v_k, v_k+1 mod n)
// lies within R of it (the first such segment ends the scan)
__global__ __launch_bounds__(256) void k_hull_bitmap(const double2* __restrict__ v, int64_t nv,
                                                     const double* __restrict__ xs,
                                                     const double* __restrict__ ys, int64_t nx,
                                                     int64_t ny, double R,
                                                     uint8_t* __restrict__ bitmap) {
  const int64_t m = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (m >= nx * ny) return;
  const double px = xs[m % nx], py = ys[m / nx];
  uint8_t in = 0;
  for (int64_t k = 0; k < nv; ++k) {
    const double2 a = v[k], b = v[k + 1 < nv ? k + 1 : 0];
    double cx, cy;
    if (R > seg_dist(px, py, a.x, a.y, b.x, b.y, &cx, &cy)) {
      in = 1;
      break;
    }
  }
  bitmap[m] = in;
}

// each boundary point to the closest point of the curve (the first segment
// at the smallest distance)
__global__ __launch_bounds__(256) void k_hull_project(const double2* __restrict__ v, int64_t nv,
                                                      double2* __restrict__ pts, int64_t np) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double2 p = pts[i];
  double bx, by;
  double best = seg_dist(p.x, p.y, v[0].x, v[0].y, v[1].x, v[1].y, &bx, &by);
  for (int64_t k = 1; k < nv; ++k) {
    const double2 b = v[k + 1 < nv ? k + 1 : 0];
    double cx, cy;
    const double d = seg_dist(p.x, p.y, v[k].x, v[k].y, b.x, b.y, &cx, &cy);
    if (d < best) {
      best = d;
      bx = cx;
      by = cy;
    }
  }
  pts[i] = make_double2(bx, by);
}

}  // namespace

int tightest_hull(slsim_ctx c, const double* v, size_t n_, std::vector<double>& out) {
#pragma clang fp contract(off)
  out.clear();
  const int64_t n = (int64_t)n_;
  if (n <= 3) {
    out.assign(v, v + 2 * n);
    return SLSIM_OK;
  }
  double ll[2] = {v[0], v[1]}, ur[2] = {v[0], v[1]};
  for (int64_t k = 0; k < n; ++k) {
    ll[0] = std::min(v[2 * k], ll[0]);
    ll[1] = std::min(v[2 * k + 1], ll[1]);
    ur[0] = std::max(v[2 * k], ur[0]);
    ur[1] = std::max(v[2 * k + 1], ur[1]);
  }
  const double resolution = std::min(ur[0] - ll[0], ur[1] - ll[1]) / 100;
  if (resolution == 0) {
    out = {ll[0], ll[1], ur[0], ur[1]};
    return SLSIM_OK;
  }
  const double R = resolution * std::sqrt(2.5);
  ll[0] -= 2 * R;
  ll[1] -= 2 * R;
  ur[0] += 2 * R;
  ur[1] += 2 * R;
  const int64_t nx = (int64_t)((ur[0] - ll[0]) / resolution);
  const int64_t ny = (int64_t)((ur[1] - ll[1]) / resolution);
  if (nx < 2 || ny < 2 || nx * ny > ((int64_t)1 << 31)) {
    set_error("tightest_hull: bitmap size out of range");
    return SLSIM_ERR_ARG;
  }
  // the pixel positions as the reference steps them (p += resolution)
  std::vector<double> xs((size_t)nx), ys((size_t)ny);
  double p = ll[0];
  for (int64_t i = 0; i < nx; ++i, p += resolution) xs[(size_t)i] = p;
  p = ll[1];
  for (int64_t j = 0; j < ny; ++j, p += resolution) ys[(size_t)j] = p;
  hipStream_t st = c->stream;
  Scratch sc(c);
  double2* dv;
  double *dxs, *dys;
  uint8_t* dbm;
  int rc;
  if ((rc = sc.get(&dv, (size_t)n)) || (rc = sc.get(&dxs, (size_t)nx)) ||
      (rc = sc.get(&dys, (size_t)ny)) || (rc = sc.get(&dbm, (size_t)(nx * ny))))
    return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dv, v, 16 * (size_t)n, hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dxs, xs.data(), 8 * (size_t)nx, hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dys, ys.data(), 8 * (size_t)ny, hipMemcpyHostToDevice, st));
  k_hull_bitmap<<<(unsigned)((nx * ny + 255) / 256), 256, 0, st>>>(dv, n, dxs, dys, nx, ny, R, dbm);
  SLSIM_HIP_CHECK(hipGetLastError());
  std::vector<uint8_t> bm((size_t)(nx * ny));
  SLSIM_HIP_CHECK(hipMemcpyAsync(bm.data(), dbm, bm.size(), hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  bool any = false;
  for (uint8_t b : bm) any |= b != 0;
  if (!any) return SLSIM_OK;
  // the outer boundary: the walk's first curve
  slsim_curves_s env;
  if ((rc = find_boundaries_host(bm.data(), (size_t)nx, (size_t)ny, &env))) return rc;
  const int64_t np = env.off.size() > 1 ? env.off[1] : 0;
  if (np == 0) return SLSIM_OK;
  std::vector<double> pts((size_t)(2 * np));
  for (int64_t i = 0; i < np; ++i) {
    pts[(size_t)(2 * i)] = ll[0] + env.pts[(size_t)(2 * i)] * resolution;
    pts[(size_t)(2 * i + 1)] = ll[1] + env.pts[(size_t)(2 * i + 1)] * resolution;
  }
  double2* dp;
  if ((rc = sc.get(&dp, (size_t)np))) return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dp, pts.data(), 16 * (size_t)np, hipMemcpyHostToDevice, st));
  k_hull_project<<<(unsigned)((np + 255) / 256), 256, 0, st>>>(dv, n, dp, np);
  SLSIM_HIP_CHECK(hipGetLastError());
  SLSIM_HIP_CHECK(hipMemcpyAsync(pts.data(), dp, 16 * (size_t)np, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  out.swap(pts);
  return SLSIM_OK;
}

// Utilities::windings (TreeCode_link/curve_routines.cpp:1709-1736): the
// winding number of the closed curve about x (isLeft's value taken as a
// float, include/Tree.h:865-867) and the absolute shoelace area
int windings_host(const double* x, const double* pt, size_t n, double* area) {
#pragma clang fp contract(off)
  int wn = 0;
  *area = 0.0;
  if (n < 3) return 0;
  for (size_t i = 0; i < n; ++i) {
    const size_t k = i < n - 1 ? i + 1 : 0;
    *area += (pt[2 * i] + pt[2 * k]) * (pt[2 * i + 1] - pt[2 * k + 1]);
    const float left = (float)((pt[2 * k] - pt[2 * i]) * (x[1] - pt[2 * i + 1]) -
                               (x[0] - pt[2 * i]) * (pt[2 * k + 1] - pt[2 * i + 1]));
    if (pt[2 * i + 1] <= x[1]) {
      if (pt[2 * k + 1] > x[1] && left > 0) ++wn;
    } else {
      if (pt[2 * k + 1] <= x[1] && left < 0) --wn;
    }
  }
  *area = std::fabs(*area) * 0.5;
  return wn;
}

}  // namespace slsim
