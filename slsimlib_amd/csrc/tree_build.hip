// tree_build.hip -- device build of the particle quadtree of one halo.
//
// Reproduces the cell set of TreeQuadParticles::BuildQTreeNB/_BuildQTreeNB
// (include/quadTree.h:295-565): square root box (shorter side extended from
// p1), four children per split at the geometric centre with `<=` going low
// (Utilities::quickPartition, include/Tree.h:396-411), child order
// 0=(x>,y>) 1=(x<=,y>) 2=(x>,y<=) 3=(x<=,y<=), a leaf when a cell holds at
// most `bucket` particles, all four children always created.
//
// MI355X design: instead of the reference's recursive in-place partitions,
// every particle computes its full 32-level path key with exactly the
// reference's centre arithmetic, keys are radix sorted, and cells are
// emitted level by level as contiguous key ranges (binary searches on the
// sorted keys).  Moments are segmented reductions, one wavefront per cell.
// Everything is FP64; masses and sizes are float as in ParticleType.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "slsim_internal.h"

namespace slsim {

namespace {

constexpr int kBlock = 256;
constexpr int32_t kBigCell = 4096;   // cells above this are reduced by a block

template <typename T>
__global__ void k_bbox(const T* __restrict__ x, const T* __restrict__ y, size_t n,
                       double* __restrict__ part) {
  __shared__ double s[4][kBlock];
  double mnx = 1e300, mny = 1e300, mxx = -1e300, mxy = -1e300;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    double a = (double)x[i], b = (double)y[i];
    mnx = fmin(mnx, a); mxx = fmax(mxx, a);
    mny = fmin(mny, b); mxy = fmax(mxy, b);
  }
  s[0][threadIdx.x] = mnx; s[1][threadIdx.x] = mny;
  s[2][threadIdx.x] = mxx; s[3][threadIdx.x] = mxy;
  __syncthreads();
  for (int o = kBlock / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + o]);
      s[1][threadIdx.x] = fmin(s[1][threadIdx.x], s[1][threadIdx.x + o]);
      s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + o]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int k = 0; k < 4; ++k) part[4 * blockIdx.x + k] = s[k][0];
}

// Path key: 2 bits per level, most significant first; digit = 2*(y > yc) + (x > xc),
// so ascending keys list child3, child2, child1, child0 of every cell, the
// layout the reference's partitions leave in its index array.
template <typename T>
__global__ void k_keys(const T* __restrict__ x, const T* __restrict__ y, size_t n, double p1x,
                       double p1y, double p2x, double p2y, uint64_t* __restrict__ keys,
                       uint32_t* __restrict__ idx) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double xi = (double)x[i], yi = (double)y[i];
  double ax = p1x, ay = p1y, bx = p2x, by = p2y;
  uint64_t key = 0;
  for (int l = 0; l < kMaxLevels; ++l) {
    const double cx = (ax + bx) / 2;
    const double cy = (ay + by) / 2;
    const bool hx = xi > cx, hy = yi > cy;
    key = (key << 2) | (uint64_t)(2 * (int)hy + (int)hx);
    if (hx) ax = cx; else bx = cx;
    if (hy) ay = cy; else by = cy;
  }
  keys[i] = key;
  idx[i] = (uint32_t)i;
}

template <typename T>
__global__ void k_gather(const uint32_t* __restrict__ perm, const T* __restrict__ x,
                         const T* __restrict__ y, const float* __restrict__ m,
                         const float* __restrict__ h, size_t n, double* __restrict__ xs,
                         double* __restrict__ ys, float* __restrict__ ms,
                         float* __restrict__ hs) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint32_t p = perm[i];
  xs[i] = (double)x[p];
  ys[i] = (double)y[p];
  ms[i] = m[p];
  hs[i] = h[p];
}

// A cell splits when it holds more than `bucket` particles and, with a
// maximum range, unless it lies so far beyond it that theta exceeds
// 1.732 box / max(|centre| - max_range, 1.732 box) (include/quadTree.h:370-380;
// products and sums rounded one by one, as on the host).
__global__ void k_split_flags(int nlev, const int32_t* __restrict__ cnt, int bucket, int level,
                              const double* __restrict__ p1, const double* __restrict__ p2,
                              double theta, double max_range, int32_t* __restrict__ flag,
                              int32_t* __restrict__ too_deep) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlev) return;
  bool split = cnt[i] > bucket;
  if (split && max_range > 0) {
    const double cx = (p1[2 * i] + p2[2 * i]) / 2, cy = (p1[2 * i + 1] + p2[2 * i + 1]) / 2;
    const double boxsize = __dmul_rn(1.732, p2[2 * i] - p1[2 * i]);
    const double d = sqrt(__dadd_rn(__dmul_rn(cx, cx), __dmul_rn(cy, cy))) - max_range;
    const double theta_range = boxsize / (d > boxsize ? d : boxsize);
    split = !(theta > theta_range);
  }
  if (split && level >= kMaxLevels) {
    *too_deep = 1;
    split = false;
  }
  flag[i] = split ? 1 : 0;
}

__device__ inline int32_t lower_digit(const uint64_t* __restrict__ keys, int32_t lo, int32_t hi,
                                      int shift, uint32_t d) {
  // first index in [lo, hi) whose digit at `shift` is >= d (digits ascend)
  while (lo < hi) {
    int32_t mid = lo + ((hi - lo) >> 1);
    uint32_t dm = (uint32_t)((keys[mid] >> shift) & 3u);
    if (dm < d) lo = mid + 1; else hi = mid;
  }
  return lo;
}

__global__ void k_children(int nlev, int level, const uint64_t* __restrict__ keys,
                           const int32_t* __restrict__ st, const int32_t* __restrict__ cnt,
                           const double* __restrict__ p1, const double* __restrict__ p2,
                           const int32_t* __restrict__ flag, const int32_t* __restrict__ offs,
                           int32_t* __restrict__ child, int32_t* __restrict__ nst,
                           int32_t* __restrict__ ncnt, double* __restrict__ np1,
                           double* __restrict__ np2) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nlev) return;
  if (!flag[i]) { child[i] = -1; return; }
  const int32_t s = st[i], e = st[i] + cnt[i];
  const int shift = 2 * (kMaxLevels - 1 - level);
  int32_t b[5];
  b[0] = s;
  b[1] = lower_digit(keys, s, e, shift, 1);
  b[2] = lower_digit(keys, b[1], e, shift, 2);
  b[3] = lower_digit(keys, b[2], e, shift, 3);
  b[4] = e;
  const double ax = p1[2 * i], ay = p1[2 * i + 1], bx = p2[2 * i], by = p2[2 * i + 1];
  const double cx = (ax + bx) / 2, cy = (ay + by) / 2;
  const int32_t base = 4 * offs[i];
  child[i] = base;
  // child k holds digit 3 - k
  for (int k = 0; k < 4; ++k) {
    const int d = 3 - k;
    nst[base + k] = b[d];
    ncnt[base + k] = b[d + 1] - b[d];
  }
  // child0 (x>, y>)
  np1[2 * (base + 0)] = cx; np1[2 * (base + 0) + 1] = cy;
  np2[2 * (base + 0)] = bx; np2[2 * (base + 0) + 1] = by;
  // child1 (x<=, y>)
  np1[2 * (base + 1)] = ax; np1[2 * (base + 1) + 1] = cy;
  np2[2 * (base + 1)] = cx; np2[2 * (base + 1) + 1] = by;
  // child2 (x>, y<=)
  np1[2 * (base + 2)] = cx; np1[2 * (base + 2) + 1] = ay;
  np2[2 * (base + 2)] = bx; np2[2 * (base + 2) + 1] = cy;
  // child3 (x<=, y<=)
  np1[2 * (base + 3)] = ax; np1[2 * (base + 3) + 1] = ay;
  np2[2 * (base + 3)] = cx; np2[2 * (base + 3) + 1] = cy;
}

__device__ inline double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// A cell's moments and its critical radius (CalcMoments, include/quadTree.h:
// 600-628): the largest corner distance from the centre of mass, over theta.
__device__ inline void store_moments(int64_t node, double sx, double sy, double m, double q0,
                                     double q1, double q2, const double* __restrict__ p1,
                                     const double* __restrict__ p2, double theta,
                                     double* __restrict__ com, double* __restrict__ mass,
                                     double* __restrict__ quad, double* __restrict__ r2crit) {
  com[2 * node] = sx;
  com[2 * node + 1] = sy;
  mass[node] = m;
  quad[3 * node] = q0;
  quad[3 * node + 1] = q1;
  quad[3 * node + 2] = q2;
  double rcom = 0;
  const double c2[2] = {sx, sy};
  for (int d = 0; d < 2; ++d) {
    double a = (c2[d] - p1[2 * node + d]) * (c2[d] - p1[2 * node + d]);
    double b = (c2[d] - p2[2 * node + d]) * (c2[d] - p2[2 * node + d]);
    rcom += a > b ? a : b;
  }
  rcom = sqrt(rcom);
  if (theta > 0.0) {
    double r = 1.15470 * rcom / theta;
    r2crit[node] = r * r;
  } else {
    r2crit[node] = 1.0e100;
  }
}

// CalcMoments (include/quadTree.h:567-628), one wavefront per cell.
__global__ void k_moments(int64_t nnodes, const int32_t* __restrict__ st,
                          const int32_t* __restrict__ cnt, const double* __restrict__ p1,
                          const double* __restrict__ p2, const double* __restrict__ xs,
                          const double* __restrict__ ys, const float* __restrict__ ms,
                          double theta, double* __restrict__ com, double* __restrict__ mass,
                          double* __restrict__ quad, double* __restrict__ r2crit) {
  const int lane = threadIdx.x & 63;
  const int64_t node = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (node >= nnodes) return;
  const int32_t s = st[node], c = cnt[node];
  if (c > kBigCell) return;   // k_moments_partial / k_moments_finish
  double m = 0, am = 0;
  for (int32_t k = lane; k < c; k += 64) {
    double w = ms[s + k];
    m += w;
    am += fabs(w);
  }
  m = wave_sum(m);
  am = wave_sum(am);
  double sx = 0, sy = 0;
  for (int32_t k = lane; k < c; k += 64) {
    double w = fabs((double)ms[s + k]);
    sx += w * xs[s + k] / am;
    sy += w * ys[s + k] / am;
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  if (c == 0) { sx = 0; sy = 0; }
  double q0 = 0, q1 = 0, q2 = 0;
  for (int32_t k = lane; k < c; k += 64) {
    double dx = xs[s + k] - sx, dy = ys[s + k] - sy;
    double r2 = dx * dx + dy * dy;
    double w = ms[s + k];
    q0 += (r2 - 2 * dx * dx) * w;
    q1 += (r2 - 2 * dy * dy) * w;
    q2 += -2 * dx * dy * w;
  }
  q0 = wave_sum(q0);
  q1 = wave_sum(q1);
  q2 = wave_sum(q2);
  if (lane == 0)
    store_moments(node, sx, sy, m, q0, q1, q2, p1, p2, theta, com, mass, quad, r2crit);
}

__device__ inline double block_sum(double v, double* sh) {
  v = wave_sum(v);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double t = 0;
  const int nw = blockDim.x >> 6;
  for (int k = 0; k < nw; ++k) t += sh[k];
  return t;
}

// Cells of more than kBigCell particles are cut into chunks of kChunk
// particles, one 1024-thread block per chunk, each pass writing its partial
// sums per chunk; a finish kernel adds a cell's chunks in order (one thread
// per cell).  The passes follow k_moments: PASS 0 the mass and the absolute
// mass, PASS 1 the |m|-weighted centre (needs the absolute mass), PASS 2 the
// quadrupole (needs the centre).  (A block per cell left the root, with
// every particle of the tree, to one CU.)
constexpr int32_t kChunk = 16384;

// The chunk table of the big cells: the cells, their first chunk, their chunk
// count (nbig each), then per chunk its cell, its first particle (offset in
// the cell) and its cell's index in the table (nchunks each).
struct ChunkTable {
  const int32_t *cell, *first, *nch;
  const int32_t *ccell, *ck0, *cbig;
  int nbig;
  uint32_t nchunks;
};

__device__ inline void block_sums(double* v, int k, double* sh) {
  for (int j = 0; j < k; ++j) v[j] = block_sum(v[j], sh);
}

template <int PASS>
__global__ __launch_bounds__(1024) void k_moments_partial(
    ChunkTable ct, const int32_t* __restrict__ st, const int32_t* __restrict__ cnt,
    const double* __restrict__ xs, const double* __restrict__ ys, const float* __restrict__ ms,
    const double* __restrict__ amb, const double* __restrict__ com, double* __restrict__ part) {
  __shared__ double sh[16];
  const int64_t node = ct.ccell[blockIdx.x];
  const int32_t s = st[node] + ct.ck0[blockIdx.x];
  const int32_t c = min(kChunk, cnt[node] - ct.ck0[blockIdx.x]);
  double v[3] = {0, 0, 0};
  if (PASS == 0) {
    for (int32_t k = threadIdx.x; k < c; k += blockDim.x) {
      double w = ms[s + k];
      v[0] += w;
      v[1] += fabs(w);
    }
    block_sums(v, 2, sh);
  } else if (PASS == 1) {
    const double am = amb[ct.cbig[blockIdx.x]];
    for (int32_t k = threadIdx.x; k < c; k += blockDim.x) {
      double w = fabs((double)ms[s + k]);
      v[0] += w * xs[s + k] / am;
      v[1] += w * ys[s + k] / am;
    }
    block_sums(v, 2, sh);
  } else {
    const double sx = com[2 * node], sy = com[2 * node + 1];
    for (int32_t k = threadIdx.x; k < c; k += blockDim.x) {
      double dx = xs[s + k] - sx, dy = ys[s + k] - sy;
      double r2 = dx * dx + dy * dy;
      double w = ms[s + k];
      v[0] += (r2 - 2 * dx * dx) * w;
      v[1] += (r2 - 2 * dy * dy) * w;
      v[2] += -2 * dx * dy * w;
    }
    block_sums(v, 3, sh);
  }
  if (threadIdx.x < 3) part[3 * (size_t)blockIdx.x + threadIdx.x] = v[threadIdx.x];
}

template <int PASS>
__global__ void k_moments_finish(ChunkTable ct, const double* __restrict__ part,
                                 const double* __restrict__ p1, const double* __restrict__ p2,
                                 double theta, double* __restrict__ amb, double* __restrict__ com,
                                 double* __restrict__ mass, double* __restrict__ quad,
                                 double* __restrict__ r2crit) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ct.nbig) return;
  const int64_t node = ct.cell[i];
  double v[3] = {0, 0, 0};
  for (int j = ct.first[i]; j < ct.first[i] + ct.nch[i]; ++j)
    for (int k = 0; k < 3; ++k) v[k] += part[3 * (size_t)j + k];
  if (PASS == 0) {
    mass[node] = v[0];
    amb[i] = v[1];
  } else if (PASS == 1) {
    com[2 * node] = v[0];
    com[2 * node + 1] = v[1];
  } else {
    store_moments(node, com[2 * node], com[2 * node + 1], mass[node], v[0], v[1], v[2], p1, p2,
                  theta, com, mass, quad, r2crit);
  }
}

// Flip error of a cell: how much a ray's deflection changes when the cell's
// opening test is decided the other way, at the distance where that test
// flips (r^2 = max(r2crit, 5.83 box^2), the opening test of
// include/quadTree.h:1010): the cell's monopole + quadrupole term (the
// reference's far-cell formula, :1089-1123, with its factor 4) against the
// exact sum over its particles, point masses and, where the point is within
// 2h of a particle, B-spline corrected (quadTree.h:147-190; as a bound, for
// every particle), at eight points around the centre of mass.  The largest
// difference (walk units, alpha before the halo's sign flip) is kept.
constexpr int kFlipDirs = 8;

__device__ inline void flip_point(int k, double cx, double cy, double R, double* px, double* py) {
  const double phi = (k + 0.5) * (kPI / 4);
  *px = cx + R * cos(phi);
  *py = cy + R * sin(phi);
}

// partial sums over particles [s + k0, s + c) with stride `step` of the
// deflection at (qx, qy): point masses (pm) and B-spline corrected (bs)
__device__ inline void flip_sums(int32_t s, int32_t c, int32_t k0, int32_t step,
                                 const double* __restrict__ xs, const double* __restrict__ ys,
                                 const float* __restrict__ ms, const float* __restrict__ hs,
                                 double qx, double qy, double* pm, double* bs) {
  pm[0] = pm[1] = bs[0] = bs[1] = 0;
  for (int32_t k = k0; k < c; k += step) {
    const double x = xs[s + k], y = ys[s + k], m = ms[s + k], h = hs[s + k];
    const double d0 = x - qx, d1 = y - qy;
    double rr = d0 * d0 + d1 * d1;
    if (rr < 1e-20) rr = 1e-20;
    const double pref = m / (kPI * rr);
    pm[0] -= pref * d0;
    pm[1] -= pref * d1;
    double b0 = -pref * d0, b1 = -pref * d1;
    if (rr < 4 * h * h) {
      const double r = sqrt(rr), q = r / h, q2 = q * q, q3 = q2 * q;
      const double M = q > 1 ? (-1 + 20 * q2 * (1 - q + 3 * q2 / 8 - q3 / 20)) / 7
                             : 10 * q2 * (1 - 3 * q / 4 + 3 * q3 / 10) / 7;
      const double ar = (M - 1) / (kPI * r);
      b0 -= m * ar * d0 / r;
      b1 -= m * ar * d1 / r;
    }
    bs[0] += b0;
    bs[1] += b1;
  }
}

// the far-cell deflection at (qx, qy) against the sums: the larger difference
__device__ inline double flip_diff(double cx, double cy, double qx, double qy, double mass,
                                   double q0, double q1, double q2, const double* pm,
                                   const double* bs) {
  const double dx = cx - qx, dy = cy - qy;
  const double r2 = dx * dx + dy * dy, inv = 1 / r2;
  const double mi = mass * inv / kPI;
  double fx = -mi * dx, fy = -mi * dy;
  const double i2p = inv * inv / kPI;
  fx -= (q0 * dx + q2 * dy) * i2p;
  fy -= (q1 * dy + q2 * dx) * i2p;
  const double t = 4 * (q0 * dx * dx + q1 * dy * dy + 2 * q2 * dx * dy) * inv * i2p;
  fx += t * dx;
  fy += t * dy;
  return fmax(hypot(fx - pm[0], fy - pm[1]), hypot(fx - bs[0], fy - bs[1]));
}

__device__ inline double flip_radius(double r2crit, double box) {
  return sqrt(fmax(r2crit, 5.83 * (box * box)));
}

// one wavefront per cell of at most kBigCell particles
__global__ void k_flip_error(int64_t nnodes, const int32_t* __restrict__ st,
                             const int32_t* __restrict__ cnt, const double* __restrict__ p1,
                             const double* __restrict__ p2, const double* __restrict__ xs,
                             const double* __restrict__ ys, const float* __restrict__ ms,
                             const float* __restrict__ hs, const double* __restrict__ com,
                             const double* __restrict__ mass, const double* __restrict__ quad,
                             const double* __restrict__ r2crit, double* __restrict__ err) {
  const int lane = threadIdx.x & 63;
  const int64_t node = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (node >= nnodes) return;
  const int32_t s = st[node], c = cnt[node];
  if (c > kBigCell) return;   // k_flip_partial / k_flip_finish
  if (c == 0 || r2crit[node] >= 1e99) {
    if (lane == 0) err[node] = 0;
    return;
  }
  const double cx = com[2 * node], cy = com[2 * node + 1];
  const double R = flip_radius(r2crit[node], p2[2 * node] - p1[2 * node]);
  double e = 0;
  for (int d = 0; d < kFlipDirs; ++d) {
    double qx, qy, pm[2], bs[2];
    flip_point(d, cx, cy, R, &qx, &qy);
    flip_sums(s, c, lane, 64, xs, ys, ms, hs, qx, qy, pm, bs);
    for (int k = 0; k < 2; ++k) {
      pm[k] = wave_sum(pm[k]);
      bs[k] = wave_sum(bs[k]);
    }
    e = fmax(e, flip_diff(cx, cy, qx, qy, mass[node], quad[3 * node], quad[3 * node + 1],
                          quad[3 * node + 2], pm, bs));
  }
  if (lane == 0) err[node] = e;
}

// Cells of more than kBigCell particles, in the chunks of the moments: one
// 1024-thread block per chunk writing its partial sums per direction (pm, bs:
// 4 doubles) to part[chunk][direction]; k_flip_finish adds a cell's chunks in
// order and takes the largest difference.

__device__ inline void flip_geometry(int64_t node, const double* __restrict__ p1,
                                     const double* __restrict__ p2,
                                     const double* __restrict__ com,
                                     const double* __restrict__ r2crit, double* cx, double* cy,
                                     double* R) {
  *cx = com[2 * node];
  *cy = com[2 * node + 1];
  *R = flip_radius(r2crit[node], p2[2 * node] - p1[2 * node]);
}

__global__ __launch_bounds__(1024) void k_flip_partial(
    ChunkTable ct, const int32_t* __restrict__ st, const int32_t* __restrict__ cnt,
    const double* __restrict__ p1, const double* __restrict__ p2, const double* __restrict__ xs,
    const double* __restrict__ ys, const float* __restrict__ ms, const float* __restrict__ hs,
    const double* __restrict__ com, const double* __restrict__ r2crit,
    double* __restrict__ part) {
  __shared__ double sh[16];
  const int64_t node = ct.ccell[blockIdx.x];
  const int32_t k0 = ct.ck0[blockIdx.x];
  const int32_t len = min(kChunk, cnt[node] - k0);
  double* out = part + (size_t)blockIdx.x * kFlipDirs * 4;
  if (r2crit[node] >= 1e99) {
    if (threadIdx.x < kFlipDirs * 4) out[threadIdx.x] = 0;
    return;
  }
  double cx, cy, R;
  flip_geometry(node, p1, p2, com, r2crit, &cx, &cy, &R);
  for (int d = 0; d < kFlipDirs; ++d) {
    double qx, qy, pm[2], bs[2];
    flip_point(d, cx, cy, R, &qx, &qy);
    flip_sums(st[node] + k0, len, threadIdx.x, blockDim.x, xs, ys, ms, hs, qx, qy, pm, bs);
    for (int k = 0; k < 2; ++k) {
      pm[k] = block_sum(pm[k], sh);
      bs[k] = block_sum(bs[k], sh);
    }
    if (threadIdx.x == 0) {
      out[4 * d] = pm[0];
      out[4 * d + 1] = pm[1];
      out[4 * d + 2] = bs[0];
      out[4 * d + 3] = bs[1];
    }
  }
}

// one thread per big cell: its chunks' sums [first, first + n) in order
__global__ void k_flip_finish(ChunkTable ct, const double* __restrict__ p1, const double* __restrict__ p2,
                              const double* __restrict__ com, const double* __restrict__ mass,
                              const double* __restrict__ quad, const double* __restrict__ r2crit,
                              const double* __restrict__ part, double* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ct.nbig) return;
  const int64_t node = ct.cell[i];
  if (r2crit[node] >= 1e99) {
    err[node] = 0;
    return;
  }
  double cx, cy, R;
  flip_geometry(node, p1, p2, com, r2crit, &cx, &cy, &R);
  double e = 0;
  for (int d = 0; d < kFlipDirs; ++d) {
    double qx, qy, pm[2] = {0, 0}, bs[2] = {0, 0};
    flip_point(d, cx, cy, R, &qx, &qy);
    for (int j = ct.first[i]; j < ct.first[i] + ct.nch[i]; ++j) {
      const double* q = part + ((size_t)j * kFlipDirs + d) * 4;
      pm[0] += q[0];
      pm[1] += q[1];
      bs[0] += q[2];
      bs[1] += q[3];
    }
    e = fmax(e, flip_diff(cx, cy, qx, qy, mass[node], quad[3 * node], quad[3 * node + 1],
                          quad[3 * node + 2], pm, bs));
  }
  err[node] = e;
}

// Mass-sheet correction of a tree whose particle masses have both signs.
// The walk adds inv_area M_c (c - x) per far cell (c its |m|-weighted centre
// of mass) and inv_area m (x_p - x) per leaf particle
// (include/quadTree.h:861-888, 1035-1044).  Summed over any partition of the
// particles this is inv_area (S - M x) + inv_area sum over the far cells of
// (M_c c - D_c), with D_c = sum of m x_p over the cell and S that sum over the
// tree; the second term vanishes when all masses share a sign.  Per cell:
// pi inv_area (M_c c - D_c) as the FP32 walk adds it (it sums pi alpha).
__global__ void k_sheet_delta(int64_t nnodes, const int32_t* __restrict__ st,
                              const int32_t* __restrict__ cnt, const double* __restrict__ xs,
                              const double* __restrict__ ys, const float* __restrict__ ms,
                              const double* __restrict__ com, const double* __restrict__ mass,
                              double inv_area, float2* __restrict__ delta,
                              double* __restrict__ dip_root) {
  const int lane = threadIdx.x & 63;
  const int64_t node = (blockIdx.x * (int64_t)blockDim.x + threadIdx.x) >> 6;
  if (node >= nnodes) return;
  const int32_t s = st[node], c = cnt[node];
  double dx = 0, dy = 0;
  for (int32_t k = lane; k < c; k += 64) {
    const double w = ms[s + k];
    dx += w * xs[s + k];
    dy += w * ys[s + k];
  }
  dx = wave_sum(dx);
  dy = wave_sum(dy);
  if (lane == 0) {
    const double f = kPI * inv_area;
    delta[node] = make_float2((float)(f * (mass[node] * com[2 * node] - dx)),
                              (float)(f * (mass[node] * com[2 * node + 1] - dy)));
    if (node == 0) {
      dip_root[0] = dx;
      dip_root[1] = dy;
    }
  }
}

// Per particle: its leaf, and whether it joins an internal cell's big list
// (root: h > box/2; other internal cells: box/2 < h <= box;
// include/quadTree.h:403-431) or its leaf's (h < box, :379-396).
__global__ void k_register(size_t n, const uint64_t* __restrict__ keys,
                           const float* __restrict__ hs, const int32_t* __restrict__ child,
                           const double* __restrict__ p1, const double* __restrict__ p2,
                           int32_t* __restrict__ leaf_of, int32_t* __restrict__ nbig_leaf,
                           uint64_t* __restrict__ pairs, uint32_t* __restrict__ npairs,
                           uint32_t pair_cap) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h = hs[i];
  const uint64_t key = keys[i];
  int32_t node = 0;
  bool registered = false;
  for (int l = 0;; ++l) {
    const double box = p2[2 * node] - p1[2 * node];
    const int32_t ch = child[node];
    if (ch < 0) {
      leaf_of[i] = node;
      if (!registered && h < box) atomicAdd(&nbig_leaf[node], 1);
      break;
    }
    if (!registered) {
      const bool band = (node == 0) ? (h > box / 2) : (h > box / 2 && h <= box);
      if (band) {
        uint32_t slot = atomicAdd(npairs, 1u);
        if (slot < pair_cap) pairs[slot] = ((uint64_t)(uint32_t)node << 32) | (uint64_t)i;
        registered = true;
      }
    }
    const uint32_t d = (uint32_t)((key >> (2 * (kMaxLevels - 1 - l))) & 3u);
    node = ch + (int32_t)(3 - d);
  }
}

__global__ void k_pack_nodes(int64_t nnodes, const double* __restrict__ com,
                             const double* __restrict__ mass, const double* __restrict__ quad,
                             const double* __restrict__ r2crit, const double* __restrict__ p1,
                             const double* __restrict__ p2, const int32_t* __restrict__ child,
                             const int32_t* __restrict__ st, const int32_t* __restrict__ cnt,
                             const int32_t* __restrict__ bst, const int32_t* __restrict__ bcnt,
                             const double* __restrict__ flip, NodeD* __restrict__ nd,
                             NodeF* __restrict__ nf) {
  int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nnodes) return;
  NodeD a;
  a.cx = com[2 * i];
  a.cy = com[2 * i + 1];
  a.mass = mass[i];
  a.q0 = quad[3 * i];
  a.q1 = quad[3 * i + 1];
  a.q2 = quad[3 * i + 2];
  a.r2crit = r2crit[i];
  const double box = p2[2 * i] - p1[2 * i];
  a.box = box;
  a.near2 = 5.83 * (box * box);
  a.child = child[i];
  int cm = 0;
  if (a.child >= 0) {
    for (int k = 0; k < 4; ++k)
      if (cnt[a.child + k] > 0) cm |= 1 << k;
    a.start = bst[i];
    a.count = bcnt[i];
  } else {
    a.start = st[i];
    a.count = cnt[i];
  }
  a.cmask = cm;
  a.T = fmax(a.r2crit, a.near2);
  nd[i] = a;
  NodeF f;
  f.cx = a.cx;
  f.cy = a.cy;
  f.near2 = a.near2;
  f.T = fmax(a.r2crit, a.near2);
  f.mass = (float)a.mass;
  // 4 q: the FP32 walk's quadrupole term needs 4 (q . e) and (q . e) / 4
  f.q0 = 4.f * (float)a.q0;
  f.q1 = 4.f * (float)a.q1;
  f.q2 = 4.f * (float)a.q2;
  f.flip = (float)flip[i];
  f.child = a.child;
  f.start = a.start;
  f.count_cmask = a.count | (cm << 28);
  nf[i] = f;
}

__global__ void k_pack_parts(size_t n, const double* __restrict__ xs, const double* __restrict__ ys,
                             const float* __restrict__ ms, const float* __restrict__ hs,
                             const int32_t* __restrict__ leaf_of, const double* __restrict__ p1,
                             const double* __restrict__ p2, const double* __restrict__ com,
                             PartD* __restrict__ pd, PartF* __restrict__ pf) {
  size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int32_t l = leaf_of[i];
  const double box = p2[2 * l] - p1[2 * l];
  const bool listed = (double)hs[i] < box;
  PartD p;
  p.x = xs[i];
  p.y = ys[i];
  p.m = ms[i];
  // the sign bit of h carries the leaf's big-particle flag, h < box
  p.h = listed ? hs[i] : -hs[i];
  pd[i] = p;
  // the FP32 walk's copy: offset from the leaf's centre of mass, 4 h^2
  PartF f;
  f.lx = (float)(p.x - com[2 * l]);
  f.ly = (float)(p.y - com[2 * l + 1]);
  f.m = ms[i];
  // (with a margin of 2^-10 for the FP32 walk's offsets: the walk takes
  // the particles below it to the FP64 offsets and the exact 4 h^2 test;
  // and at least FLT_MIN, so that a particle on the ray takes that branch)
  const float hh4 = fmaxf(4.f * hs[i] * hs[i] * (1.f + 1.f / 1024.f), 1.17549435e-38f);
  f.hh4 = listed ? hh4 : -hh4;
  pf[i] = f;
}

__global__ void k_pack_big(uint32_t nb, const uint64_t* __restrict__ pairs,
                           const double* __restrict__ xs, const double* __restrict__ ys,
                           const float* __restrict__ ms, const float* __restrict__ hs,
                           PartD* __restrict__ bd) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const uint32_t i = (uint32_t)(pairs[k] & 0xffffffffu);
  PartD p;
  p.x = xs[i];
  p.y = ys[i];
  p.m = ms[i];
  p.h = hs[i];
  bd[k] = p;
}

__global__ void k_big_ranges(uint32_t nb, const uint64_t* __restrict__ pairs,
                             int32_t* __restrict__ bst, int32_t* __restrict__ bcnt) {
  uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nb) return;
  const int32_t node = (int32_t)(pairs[k] >> 32);
  if (k == 0 || (int32_t)(pairs[k - 1] >> 32) != node) bst[node] = (int32_t)k;
  atomicAdd(&bcnt[node], 1);
}

inline unsigned nblocks(size_t n, int b = kBlock) { return (unsigned)((n + b - 1) / b); }

// RAII list of temporaries owned by the context.
struct Temps {
  slsim_ctx ctx;
  std::vector<void*> ptrs;
  explicit Temps(slsim_ctx c) : ctx(c) {}
  ~Temps() {
    for (void* p : ptrs) ctx->release(p);
  }
  template <typename T>
  int get(T** p, size_t count) {
    void* v = nullptr;
    int rc = ctx->alloc(&v, std::max<size_t>(count, 1) * sizeof(T));
    if (rc) return rc;
    ptrs.push_back(v);
    *p = static_cast<T*>(v);
    return SLSIM_OK;
  }
};

}  // namespace

// Host orchestration of the build.  Returns a status code.
template <typename T>
int build_plane(slsim_ctx ctx, const T* x_host, const T* y_host, const float* m_host,
                const float* h_host, size_t n, slsim_plane pl) {
  hipStream_t st = ctx->stream;
  {
    bool pos = false, neg = false;
    for (size_t i = 0; i < n; ++i) {
      pos |= m_host[i] > 0;
      neg |= m_host[i] < 0;
    }
    pl->mixed_sign = pos && neg;
  }
  hipEvent_t ev0, ev1;
  SLSIM_HIP_CHECK(hipEventCreate(&ev0));
  SLSIM_HIP_CHECK(hipEventCreate(&ev1));
  Temps tmp(ctx);
  int rc;
  T *dx, *dy;
  float *dm, *dh;
  if ((rc = tmp.get(&dx, n)) || (rc = tmp.get(&dy, n)) || (rc = tmp.get(&dm, n)) ||
      (rc = tmp.get(&dh, n)))
    return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dx, x_host, n * sizeof(T), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dy, y_host, n * sizeof(T), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dm, m_host, n * sizeof(float), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dh, h_host, n * sizeof(float), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipEventRecord(ev0, st));

  // ---- root box (BuildQTreeNB, include/quadTree.h:302-333)
  const unsigned bb_blocks = std::min<unsigned>(1024, nblocks(n));
  double* dpart;
  if ((rc = tmp.get(&dpart, 4 * (size_t)bb_blocks))) return rc;
  k_bbox<T><<<bb_blocks, kBlock, 0, st>>>(dx, dy, n, dpart);
  std::vector<double> hpart(4 * (size_t)bb_blocks);
  SLSIM_HIP_CHECK(hipMemcpyAsync(hpart.data(), dpart, hpart.size() * sizeof(double),
                                 hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  double p1[2] = {1e300, 1e300}, p2[2] = {-1e300, -1e300};
  for (unsigned b = 0; b < bb_blocks; ++b) {
    p1[0] = std::min(p1[0], hpart[4 * b]);
    p1[1] = std::min(p1[1], hpart[4 * b + 1]);
    p2[0] = std::max(p2[0], hpart[4 * b + 2]);
    p2[1] = std::max(p2[1], hpart[4 * b + 3]);
  }
  if (n == 1) { p1[0] = -0.25; p1[1] = -0.25; p2[0] = 0.25; p2[1] = 0.25; }
  double lengths[2] = {p2[0] - p1[0], p2[1] - p1[1]};
  if (lengths[0] == 0 || lengths[1] == 0) {
    set_error("particles in same place.");
    return SLSIM_ERR_SAME_PLACE;
  }
  {
    int j = lengths[0] > lengths[1] ? 1 : 0;
    p2[j] = p1[j] + lengths[!j];
  }
  pl->p1[0] = p1[0]; pl->p1[1] = p1[1];
  pl->p2[0] = p2[0]; pl->p2[1] = p2[1];

  // ---- path keys and radix sort
  uint64_t *keys, *keys_s;
  uint32_t *idx, *perm;
  if ((rc = tmp.get(&keys, n)) || (rc = tmp.get(&keys_s, n)) || (rc = tmp.get(&idx, n)) ||
      (rc = tmp.get(&perm, n)))
    return rc;
  k_keys<T><<<nblocks(n), kBlock, 0, st>>>(dx, dy, n, p1[0], p1[1], p2[0], p2[1], keys, idx);
  size_t sort_bytes = 0;
  SLSIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, keys_s, idx,
                                                     perm, (int)n, 0, 64, st));
  void* sort_tmp;
  if ((rc = tmp.get(reinterpret_cast<char**>(&sort_tmp), sort_bytes))) return rc;
  SLSIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, keys, keys_s, idx,
                                                     perm, (int)n, 0, 64, st));
  double *xs, *ys;
  float *ms, *hs;
  if ((rc = tmp.get(&xs, n)) || (rc = tmp.get(&ys, n)) || (rc = tmp.get(&ms, n)) ||
      (rc = tmp.get(&hs, n)))
    return rc;
  k_gather<T><<<nblocks(n), kBlock, 0, st>>>(perm, dx, dy, dm, dh, n, xs, ys, ms, hs);

  // ---- cells, level by level
  struct Level {
    int32_t nlev;
    int32_t *st, *cnt, *child;
    double *p1, *p2;
  };
  std::vector<Level> levels;
  {
    Level L0;
    L0.nlev = 1;
    if ((rc = tmp.get(&L0.st, 1)) || (rc = tmp.get(&L0.cnt, 1)) || (rc = tmp.get(&L0.child, 1)) ||
        (rc = tmp.get(&L0.p1, 2)) || (rc = tmp.get(&L0.p2, 2)))
      return rc;
    int32_t hst = 0, hcnt = (int32_t)n;
    SLSIM_HIP_CHECK(hipMemcpyAsync(L0.st, &hst, 4, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(L0.cnt, &hcnt, 4, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(L0.p1, p1, 16, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(L0.p2, p2, 16, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
    levels.push_back(L0);
  }
  int32_t* too_deep;
  if ((rc = tmp.get(&too_deep, 1))) return rc;
  SLSIM_HIP_CHECK(hipMemsetAsync(too_deep, 0, 4, st));
  for (int level = 0;; ++level) {
    Level& L = levels.back();
    int32_t *flag, *offs;
    if ((rc = tmp.get(&flag, L.nlev)) || (rc = tmp.get(&offs, L.nlev))) return rc;
    k_split_flags<<<nblocks(L.nlev), kBlock, 0, st>>>(L.nlev, L.cnt, pl->bucket, level, L.p1,
                                                      L.p2, pl->theta, pl->max_range, flag,
                                                      too_deep);
    size_t scan_bytes = 0;
    SLSIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, flag, offs, L.nlev, st));
    char* scan_tmp;
    if ((rc = tmp.get(&scan_tmp, scan_bytes))) return rc;
    SLSIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, flag, offs, L.nlev, st));
    int32_t last[2];
    SLSIM_HIP_CHECK(hipMemcpyAsync(&last[0], offs + L.nlev - 1, 4, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(&last[1], flag + L.nlev - 1, 4, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
    const int32_t nsplit = last[0] + last[1];
    Level N;
    N.nlev = 4 * nsplit;
    if (nsplit > 0) {
      if ((rc = tmp.get(&N.st, N.nlev)) || (rc = tmp.get(&N.cnt, N.nlev)) ||
          (rc = tmp.get(&N.child, N.nlev)) || (rc = tmp.get(&N.p1, 2 * (size_t)N.nlev)) ||
          (rc = tmp.get(&N.p2, 2 * (size_t)N.nlev)))
        return rc;
    } else {
      N.st = N.cnt = N.child = nullptr;
      N.p1 = N.p2 = nullptr;
    }
    k_children<<<nblocks(L.nlev), kBlock, 0, st>>>(L.nlev, level, keys_s, L.st, L.cnt, L.p1, L.p2,
                                                   flag, offs, L.child, N.st, N.cnt, N.p1, N.p2);
    if (nsplit == 0) break;
    levels.push_back(N);
  }
  int32_t h_too_deep = 0;
  SLSIM_HIP_CHECK(hipMemcpyAsync(&h_too_deep, too_deep, 4, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  if (h_too_deep) {
    set_error("a cell at level 32 still holds more than bucket particles");
    return SLSIM_ERR_TOO_DEEP;
  }

  // ---- concatenate levels into global BFS arrays
  pl->level_offset.clear();
  int64_t nnodes = 0;
  for (auto& L : levels) {
    pl->level_offset.push_back(nnodes);
    nnodes += L.nlev;
  }
  pl->level_offset.push_back(nnodes);
  pl->nnodes = nnodes;
  pl->depth = (int32_t)levels.size() - 1;
  int32_t *gst, *gcnt, *gchild;
  double *gp1, *gp2;
  if ((rc = tmp.get(&gst, nnodes)) || (rc = tmp.get(&gcnt, nnodes)) || (rc = tmp.get(&gchild, nnodes)))
    return rc;
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&gp1), 16 * nnodes))) return rc;
  pl->box_p1 = gp1;
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&gp2), 16 * nnodes))) return rc;
  pl->box_p2 = gp2;
  std::vector<int32_t> hchild_fix;
  for (size_t l = 0; l < levels.size(); ++l) {
    const Level& L = levels[l];
    const int64_t o = pl->level_offset[l];
    SLSIM_HIP_CHECK(hipMemcpyAsync(gst + o, L.st, 4 * L.nlev, hipMemcpyDeviceToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(gcnt + o, L.cnt, 4 * L.nlev, hipMemcpyDeviceToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(gp1 + 2 * o, L.p1, 16 * L.nlev, hipMemcpyDeviceToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(gp2 + 2 * o, L.p2, 16 * L.nlev, hipMemcpyDeviceToDevice, st));
  }
  // child indices are relative to the next level's start: make them global on the host
  {
    std::vector<int32_t> hc(nnodes);
    for (size_t l = 0; l < levels.size(); ++l) {
      const Level& L = levels[l];
      SLSIM_HIP_CHECK(hipMemcpyAsync(hc.data() + pl->level_offset[l], L.child, 4 * L.nlev,
                                     hipMemcpyDeviceToHost, st));
    }
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
    int64_t nleaves = 0;
    for (size_t l = 0; l < levels.size(); ++l) {
      const int64_t o = pl->level_offset[l], e = pl->level_offset[l + 1];
      for (int64_t i = o; i < e; ++i) {
        if (hc[i] >= 0) hc[i] += (int32_t)pl->level_offset[l + 1];
        else ++nleaves;
      }
    }
    pl->nleaves = nleaves;
    SLSIM_HIP_CHECK(hipMemcpyAsync(gchild, hc.data(), 4 * nnodes, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  }

  // ---- moments
  double *com, *mass, *quad, *r2c;
  if ((rc = tmp.get(&com, 2 * nnodes)) || (rc = tmp.get(&mass, nnodes)) ||
      (rc = tmp.get(&quad, 3 * nnodes)) || (rc = tmp.get(&r2c, nnodes)))
    return rc;
  // the chunk table of the cells of more than kBigCell particles
  ChunkTable ct{};
  {
    std::vector<int32_t> hcnt(nnodes);
    SLSIM_HIP_CHECK(hipMemcpyAsync(hcnt.data(), gcnt, 4 * nnodes, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
    std::vector<int32_t> cell, first, nch, ccell, ck0, cbig;
    for (int64_t i = 0; i < nnodes; ++i) {
      if (hcnt[i] <= kBigCell) continue;
      first.push_back((int32_t)ccell.size());
      for (int32_t k0 = 0; k0 < hcnt[i]; k0 += kChunk) {
        ccell.push_back((int32_t)i);
        ck0.push_back(k0);
        cbig.push_back((int32_t)cell.size());
      }
      nch.push_back((int32_t)ccell.size() - first.back());
      cell.push_back((int32_t)i);
    }
    ct.nbig = (int)cell.size();
    ct.nchunks = (uint32_t)ccell.size();
    if (ct.nbig > 0) {
      std::vector<int32_t> tab;
      for (const auto* v : {&cell, &first, &nch, &ccell, &ck0, &cbig})
        tab.insert(tab.end(), v->begin(), v->end());
      int32_t* dtab;
      double *part, *amb;
      if ((rc = tmp.get(&dtab, tab.size())) || (rc = tmp.get(&part, 3 * (size_t)ct.nchunks)) ||
          (rc = tmp.get(&amb, ct.nbig)))
        return rc;
      SLSIM_HIP_CHECK(hipMemcpyAsync(dtab, tab.data(), 4 * tab.size(), hipMemcpyHostToDevice, st));
      const size_t nb = ct.nbig, nc = ct.nchunks;
      ct.cell = dtab;
      ct.first = dtab + nb;
      ct.nch = dtab + 2 * nb;
      ct.ccell = dtab + 3 * nb;
      ct.ck0 = dtab + 3 * nb + nc;
      ct.cbig = dtab + 3 * nb + 2 * nc;
      const unsigned fb = (unsigned)((nb + 63) / 64);
      k_moments_partial<0><<<ct.nchunks, 1024, 0, st>>>(ct, gst, gcnt, xs, ys, ms, amb, com, part);
      k_moments_finish<0><<<fb, 64, 0, st>>>(ct, part, gp1, gp2, pl->theta, amb, com, mass, quad,
                                             r2c);
      k_moments_partial<1><<<ct.nchunks, 1024, 0, st>>>(ct, gst, gcnt, xs, ys, ms, amb, com, part);
      k_moments_finish<1><<<fb, 64, 0, st>>>(ct, part, gp1, gp2, pl->theta, amb, com, mass, quad,
                                             r2c);
      k_moments_partial<2><<<ct.nchunks, 1024, 0, st>>>(ct, gst, gcnt, xs, ys, ms, amb, com, part);
      k_moments_finish<2><<<fb, 64, 0, st>>>(ct, part, gp1, gp2, pl->theta, amb, com, mass, quad,
                                             r2c);
      SLSIM_HIP_CHECK(hipStreamSynchronize(st));   // `tab` is pageable host memory
    }
  }
  k_moments<<<nblocks(64 * (size_t)nnodes), kBlock, 0, st>>>(nnodes, gst, gcnt, gp1, gp2, xs, ys,
                                                             ms, pl->theta, com, mass, quad, r2c);

  // ---- big-particle registration
  int32_t *leaf_of, *nbig;
  uint32_t* npairs;
  if ((rc = tmp.get(&leaf_of, n)) || (rc = tmp.get(&npairs, 1))) return rc;
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&nbig), 4 * nnodes))) return rc;
  pl->nbig = nbig;
  SLSIM_HIP_CHECK(hipMemsetAsync(nbig, 0, 4 * nnodes, st));
  SLSIM_HIP_CHECK(hipMemsetAsync(npairs, 0, 4, st));
  const uint32_t pair_cap = (uint32_t)std::min<size_t>(n, 1u << 30);
  uint64_t *pairs, *pairs_s;
  if ((rc = tmp.get(&pairs, pair_cap)) || (rc = tmp.get(&pairs_s, pair_cap))) return rc;
  k_register<<<nblocks(n), kBlock, 0, st>>>(n, keys_s, hs, gchild, gp1, gp2, leaf_of, nbig, pairs,
                                            npairs, pair_cap);
  uint32_t hpairs = 0;
  SLSIM_HIP_CHECK(hipMemcpyAsync(&hpairs, npairs, 4, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  pl->nbig_internal = hpairs;
  int32_t *bst, *bcnt;
  if ((rc = tmp.get(&bst, nnodes)) || (rc = tmp.get(&bcnt, nnodes))) return rc;
  SLSIM_HIP_CHECK(hipMemsetAsync(bst, 0, 4 * nnodes, st));
  SLSIM_HIP_CHECK(hipMemsetAsync(bcnt, 0, 4 * nnodes, st));
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&pl->bigD), sizeof(PartD) * std::max<uint32_t>(hpairs, 1))))
    return rc;
  if (hpairs > 0) {
    size_t sb = 0;
    SLSIM_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(nullptr, sb, pairs, pairs_s, (int)hpairs, 0, 64, st));
    char* stmp;
    if ((rc = tmp.get(&stmp, sb))) return rc;
    SLSIM_HIP_CHECK(hipcub::DeviceRadixSort::SortKeys(stmp, sb, pairs, pairs_s, (int)hpairs, 0, 64, st));
    k_big_ranges<<<nblocks(hpairs), kBlock, 0, st>>>(hpairs, pairs_s, bst, bcnt);
    k_pack_big<<<nblocks(hpairs), kBlock, 0, st>>>(hpairs, pairs_s, xs, ys, ms, hs, pl->bigD);
    // the reference's big-list length of internal cells, for export
    std::vector<int32_t> hb(nnodes), hnb(nnodes);
    SLSIM_HIP_CHECK(hipMemcpyAsync(hb.data(), bcnt, 4 * nnodes, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(hnb.data(), nbig, 4 * nnodes, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
    for (int64_t i = 0; i < nnodes; ++i) hnb[i] += hb[i];
    SLSIM_HIP_CHECK(hipMemcpyAsync(nbig, hnb.data(), 4 * nnodes, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  }

  // ---- pack walk records
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&pl->nodesD), sizeof(NodeD) * nnodes))) return rc;
  // three zero cells after the last: a group of four child slots read from
  // any cell stays inside the array (the FP32 walk starts a tree as a virtual
  // entry whose first slot is the root)
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&pl->nodesF), sizeof(NodeF) * (nnodes + 3))))
    return rc;
  SLSIM_HIP_CHECK(hipMemsetAsync(pl->nodesF + nnodes, 0, 3 * sizeof(NodeF), st));
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&pl->partsD), sizeof(PartD) * n))) return rc;
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&pl->partsF), sizeof(PartF) * n))) return rc;
  if ((rc = ctx->alloc(reinterpret_cast<void**>(&pl->npart), 4 * nnodes))) return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(pl->npart, gcnt, 4 * nnodes, hipMemcpyDeviceToDevice, st));
  double* flip;
  if ((rc = tmp.get(&flip, nnodes))) return rc;
  k_flip_error<<<nblocks(64 * (size_t)nnodes), kBlock, 0, st>>>(
      nnodes, gst, gcnt, gp1, gp2, xs, ys, ms, hs, com, mass, quad, r2c, flip);
  if (ct.nbig > 0) {
    double* part;
    if ((rc = tmp.get(&part, (size_t)ct.nchunks * kFlipDirs * 4))) return rc;
    k_flip_partial<<<ct.nchunks, 1024, 0, st>>>(ct, gst, gcnt, gp1, gp2, xs, ys, ms, hs, com, r2c,
                                                part);
    k_flip_finish<<<(unsigned)((ct.nbig + 63) / 64), 64, 0, st>>>(ct, gp1, gp2, com, mass, quad,
                                                                  r2c, part, flip);
  }
  k_pack_nodes<<<nblocks(nnodes), kBlock, 0, st>>>(nnodes, com, mass, quad, r2c, gp1, gp2, gchild,
                                                   gst, gcnt, bst, bcnt, flip, pl->nodesD,
                                                   pl->nodesF);
  k_pack_parts<<<nblocks(n), kBlock, 0, st>>>(n, xs, ys, ms, hs, leaf_of, gp1, gp2, com,
                                               pl->partsD, pl->partsF);
  // masses of both signs: the sheet's partition-dependent term per cell
  double* dip = nullptr;
  if (pl->mixed_sign && pl->inv_area != 0.0) {
    if ((rc = ctx->alloc(reinterpret_cast<void**>(&pl->sheetd), sizeof(float2) * (nnodes + 3))))
      return rc;
    SLSIM_HIP_CHECK(hipMemsetAsync(pl->sheetd + nnodes, 0, 3 * sizeof(float2), st));
    if ((rc = tmp.get(&dip, 2))) return rc;
    k_sheet_delta<<<nblocks(64 * (size_t)nnodes), kBlock, 0, st>>>(
        nnodes, gst, gcnt, xs, ys, ms, com, mass, pl->inv_area, pl->sheetd, dip);
  }
  SLSIM_HIP_CHECK(hipEventRecord(ev1, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  SLSIM_HIP_CHECK(hipGetLastError());
  {
    NodeD root;
    SLSIM_HIP_CHECK(hipMemcpy(&root, pl->nodesD, sizeof(NodeD), hipMemcpyDeviceToHost));
    pl->mass = root.mass;
    pl->com[0] = root.cx;
    pl->com[1] = root.cy;
    pl->dip[0] = root.mass * root.cx;
    pl->dip[1] = root.mass * root.cy;
    if (dip) SLSIM_HIP_CHECK(hipMemcpy(pl->dip, dip, 16, hipMemcpyDeviceToHost));
  }
  float ms_build = 0;
  SLSIM_HIP_CHECK(hipEventElapsedTime(&ms_build, ev0, ev1));
  pl->build_ms = ms_build;
  (void)hipEventDestroy(ev0);
  (void)hipEventDestroy(ev1);
  pl->bytes = nnodes * (sizeof(NodeD) + sizeof(NodeF) + 16 + 16 + 4 + 4) + 3 * sizeof(NodeF) +
              n * (sizeof(PartD) + sizeof(PartF)) + hpairs * sizeof(PartD);
  return SLSIM_OK;
}

template int build_plane<double>(slsim_ctx, const double*, const double*, const float*,
                                 const float*, size_t, slsim_plane);
template int build_plane<float>(slsim_ctx, const float*, const float*, const float*,
                                const float*, size_t, slsim_plane);

// Node export in depth-first pre-order (walkTree_recur's order).
int export_nodes(slsim_plane pl, double* dbl, int64_t* ints) {
  slsim_ctx ctx = pl->ctx;
  hipStream_t st = ctx->stream;
  const int64_t nn = pl->nnodes;
  std::vector<NodeD> nd(nn);
  std::vector<NodeF> nf(nn);
  std::vector<double> p1(2 * nn), p2(2 * nn);
  std::vector<int32_t> nb(nn), np(nn);
  SLSIM_HIP_CHECK(hipMemcpyAsync(nd.data(), pl->nodesD, sizeof(NodeD) * nn, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(nf.data(), pl->nodesF, sizeof(NodeF) * nn, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(p1.data(), pl->box_p1, 16 * nn, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(p2.data(), pl->box_p2, 16 * nn, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(nb.data(), pl->nbig, 4 * nn, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(np.data(), pl->npart, 4 * nn, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  std::vector<std::pair<int32_t, int32_t>> stack;   // (node, level)
  stack.push_back({0, 0});
  int64_t k = 0;
  while (!stack.empty()) {
    auto [i, lev] = stack.back();
    stack.pop_back();
    const NodeD& a = nd[i];
    double* d = dbl + 13 * k;
    d[0] = p1[2 * i]; d[1] = p1[2 * i + 1]; d[2] = p2[2 * i]; d[3] = p2[2 * i + 1];
    d[4] = a.cx; d[5] = a.cy; d[6] = a.mass; d[7] = a.q0; d[8] = a.q1; d[9] = a.q2;
    d[10] = a.r2crit; d[11] = nf[i].flip; d[12] = 0;
    int64_t* q = ints + 6 * k;
    q[0] = lev; q[1] = np[i]; q[2] = nb[i]; q[3] = a.child < 0 ? 1 : 0; q[4] = 0; q[5] = 0;
    ++k;
    if (a.child >= 0)
      for (int c = 3; c >= 0; --c) stack.push_back({a.child + c, lev + 1});
  }
  return SLSIM_OK;
}

}  // namespace slsim
