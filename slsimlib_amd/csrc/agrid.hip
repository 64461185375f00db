// agrid.hip -- the adaptive image-plane grid (Grid, TreeCode_link/
// grid_maintenance.cpp:22-134, 543-982, image_finder.cpp:790-831).
//
// The grid's points live in HBM as flat arrays in the reference's order
// (image position, source position, magnification matrix, cell size), the
// initial grid first and each refinement's new points appended after it.
// A refinement of many cells (RefineLeaves) is one device pass: every parent
// cell's eight sub-cell points are formed from its position and size, those
// outside the initial image box (or too small for double precision) are
// dropped by a stable compaction that keeps the reference's order, the rest
// are shot through the lens in one device call, and those whose source
// position falls outside the initial source box are dropped the same way.
// The parents' cell sizes are divided by three, and restored as the
// reference restores them (multiplied by three) when none of their points
// was kept.  RefineLeaf keeps the reference's swap order for its single cell.
// A grid given a server (slsim_agrid_set_server) sends each refinement's rays
// as one ray-list request, so that many grids refining at once share device
// launches through the server's continuous batching.
//
// The reference's kd-trees of image and source points (TreeStruct), used by
// its Grid-based image and critical-curve finders, are not built: the grid
// keeps the two root boxes that the refinement tests against.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

#include "slsim_internal.h"

namespace slsim {
namespace {

constexpr int kBlockSize = 256;
inline unsigned nblk(int64_t n) { return (unsigned)((n + kBlockSize - 1) / kBlockSize); }

// Initial grid positions.  square: Grid::xygridpoints (image_finder.cpp:
// 790-824) through Utilities::PositionFromIndex (utilities.cpp:149-158);
// otherwise the rectangular constructor's loop (grid_maintenance.cpp:
// 93-106).  No contraction: the reference's roundings.
__global__ void k_agrid_init(int64_t n, int64_t N, int square, double c0, double c1,
                             double rx, double ry, double* __restrict__ x,
                             double* __restrict__ gs) {
#pragma clang fp contract(off)
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t ii = i % N, jj = i / N;
  double x0, x1;
  if (square) {
    if (N == 1) {
      x0 = c0;
      x1 = c1;
    } else {
      x0 = c0 + rx * (1.0 * (double)ii / (double)(N - 1) - 0.5);
      x1 = c1 + rx * (1.0 * (double)jj / (double)(N - 1) - 0.5);
    }
  } else {
    x0 = c0 + rx * (double)ii / (double)(N - 1) - 0.5 * rx;
    x1 = c1 + rx * (double)jj / (double)(N - 1) - 0.5 * ry;
  }
  x[2 * i] = x0;
  x[2 * i + 1] = x1;
  gs[i] = rx / (double)(N - 1);
}

// per-block minima and maxima of the (n, 2) points
__global__ void k_agrid_minmax(int64_t n, const double* __restrict__ p,
                               double* __restrict__ part) {
  __shared__ double s[4][kBlockSize];
  double lo0 = HUGE_VAL, lo1 = HUGE_VAL, hi0 = -HUGE_VAL, hi1 = -HUGE_VAL;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    lo0 = fmin(lo0, p[2 * i]);
    hi0 = fmax(hi0, p[2 * i]);
    lo1 = fmin(lo1, p[2 * i + 1]);
    hi1 = fmax(hi1, p[2 * i + 1]);
  }
  s[0][threadIdx.x] = lo0;
  s[1][threadIdx.x] = lo1;
  s[2][threadIdx.x] = hi0;
  s[3][threadIdx.x] = hi1;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + w]);
      s[1][threadIdx.x] = fmin(s[1][threadIdx.x], s[1][threadIdx.x + w]);
      s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + w]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) part[4 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ inline bool inbox(double x0, double x1, const double* p1, const double* p2) {
  return x0 >= p1[0] && x0 <= p2[0] && x1 >= p1[1] && x1 <= p2[1];
}

// Box corners in constant-size kernel arguments.
struct Box {
  double p1[2], p2[2];
};

// The eight sub-cell points of each parent (Grid::xygridpoints with the
// centre left out, range = gridsize (3 - 1) / 3, PositionFromIndex over a
// 3 x 3 grid) and whether RefineLeaves keeps them (grid_maintenance.cpp:
// 722-732): inside the initial image box and not below 1e-10 of the
// position's size.
__global__ void k_agrid_children(int64_t np, const int64_t* __restrict__ parent,
                                 const double* __restrict__ x, const double* __restrict__ gs,
                                 Box ib, double* __restrict__ cx, double* __restrict__ cgs,
                                 uint32_t* __restrict__ keep) {
#pragma clang fp contract(off)
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= 8 * np) return;
  const int64_t q = parent[t / 8];
  const int c = (int)(t % 8);
  const int i = c < 4 ? c : c + 1;   // the 3 x 3 index, centre (4) left out
  const double range = gs[q] * (3 - 1) / 3;
  const double x0 = x[2 * q] + range * (1.0 * (double)(i % 3) / (double)(3 - 1) - 0.5);
  const double x1 = x[2 * q + 1] + range * (1.0 * (double)(i / 3) / (double)(3 - 1) - 0.5);
  const double g = range / (double)(3 - 1);
  cx[2 * t] = x0;
  cx[2 * t + 1] = x1;
  cgs[t] = g;
  keep[t] = inbox(x0, x1, ib.p1, ib.p2) && !(g < 1.0e-10 * fmax(fabs(x0), fabs(x1)));
}

// stable compaction: item t goes to slot pos[t] when keep[t]
__global__ void k_agrid_scatter(int64_t n, const uint32_t* __restrict__ keep,
                                const uint32_t* __restrict__ pos, const double* __restrict__ sx,
                                const double* __restrict__ sy, const double* __restrict__ sA,
                                const double* __restrict__ sg, int64_t* __restrict__ sp,
                                double* __restrict__ dx, double* __restrict__ dy,
                                double* __restrict__ dA, double* __restrict__ dg,
                                int64_t* __restrict__ dp) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n || !keep[t]) return;
  const int64_t k = pos[t];
  dx[2 * k] = sx[2 * t];
  dx[2 * k + 1] = sx[2 * t + 1];
  if (sy) {
    dy[2 * k] = sy[2 * t];
    dy[2 * k + 1] = sy[2 * t + 1];
    for (int c = 0; c < 4; ++c) dA[4 * k + c] = sA[4 * t + c];
  }
  dg[k] = sg[t];
  if (sp) dp[k] = sp[t];
}

__global__ void k_agrid_source_keep(int64_t n, const double* __restrict__ y, Box sb,
                                    uint32_t* __restrict__ keep) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n) return;
  keep[t] = inbox(y[2 * t], y[2 * t + 1], sb.p1, sb.p2);
}

// the parents' cell sizes: / 3, or (/ 3) * 3 when none of their points stayed
// (grid_maintenance.cpp:706-708, 736-740, 848-852)
__global__ void k_agrid_parents(int64_t np, const int64_t* __restrict__ parent,
                                const uint32_t* __restrict__ keep, double* __restrict__ gs) {
#pragma clang fp contract(off)
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= np) return;
  uint32_t any = 0;
  for (int c = 0; c < 8; ++c) any |= keep[8 * t + c];
  const int64_t q = parent[t];
  double g = gs[q] / 3;
  if (!any) g *= 3;
  gs[q] = g;
}

// per-block sums of gridsize^2 over the points with det A < 0
// (Grid::EinsteinArea, grid_maintenance.cpp:368-378)
__global__ void k_agrid_einstein(int64_t n, const double* __restrict__ A,
                                 const double* __restrict__ gs, double* __restrict__ part) {
#pragma clang fp contract(off)
  __shared__ double s[kBlockSize];
  double acc = 0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    const double det = A[4 * i] * A[4 * i + 3] - A[4 * i + 1] * A[4 * i + 2];
    if (det < 0) acc += gs[i] * gs[i];
  }
  s[threadIdx.x] = acc;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

constexpr unsigned kReduceBlocks = 512;

int arg_error(const char* msg) {
  set_error(msg);
  return SLSIM_ERR_ARG;
}

}  // namespace
}  // namespace slsim

using namespace slsim;

struct slsim_agrid_s {
  slsim_lens lens = nullptr;
  slsim_ctx ctx = nullptr;
  int precision = SLSIM_FP64;
  int64_t Ngrid_init = 0, Ngrid_init2 = 0;
  int square = 1;
  Box ib{}, sb{};            // the image and source trees' root boxes
  int64_t n = 0, cap = 0;
  double *x = nullptr, *y = nullptr, *A = nullptr, *gs = nullptr;
  slsim_server server = nullptr;   // refinements as ray-list requests of a server
  ~slsim_agrid_s() { release(); }
  int shoot_host(int64_t count, const double* hx, double* hy, double* hA);
  int shoot_dev(int64_t count, const double* dx, double* dy, double* dA);
  void release() {
    for (double* p : {x, y, A, gs})
      if (p) ctx->release(p);
    x = y = A = gs = nullptr;
    cap = 0;
  }
  int reserve(int64_t want);
  int minmax(const double* p, int64_t count, double lo[2], double hi[2]);
};

int slsim_agrid_s::reserve(int64_t want) {
  if (want <= cap) return SLSIM_OK;
  const int64_t nc = std::max<int64_t>(want, 2 * cap);
  double *nx = nullptr, *ny = nullptr, *nA = nullptr, *ng = nullptr;
  int rc;
  if ((rc = ctx->alloc((void**)&nx, 16 * (size_t)nc)) ||
      (rc = ctx->alloc((void**)&ny, 16 * (size_t)nc)) ||
      (rc = ctx->alloc((void**)&nA, 32 * (size_t)nc)) ||
      (rc = ctx->alloc((void**)&ng, 8 * (size_t)nc))) {
    for (double* p : {nx, ny, nA, ng})
      if (p) ctx->release(p);
    return rc;
  }
  hipStream_t st = ctx->stream;
  if (n > 0) {
    SLSIM_HIP_CHECK(hipMemcpyAsync(nx, x, 16 * (size_t)n, hipMemcpyDeviceToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(ny, y, 16 * (size_t)n, hipMemcpyDeviceToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(nA, A, 32 * (size_t)n, hipMemcpyDeviceToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(ng, gs, 8 * (size_t)n, hipMemcpyDeviceToDevice, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  }
  const int64_t keep_n = n;
  release();
  x = nx;
  y = ny;
  A = nA;
  gs = ng;
  cap = nc;
  n = keep_n;
  return SLSIM_OK;
}

int slsim_agrid_s::minmax(const double* p, int64_t count, double lo[2], double hi[2]) {
  Scratch sc(ctx);
  double* part;
  int rc;
  if ((rc = sc.get(&part, 4 * kReduceBlocks))) return rc;
  k_agrid_minmax<<<kReduceBlocks, kBlockSize, 0, ctx->stream>>>(count, p, part);
  SLSIM_HIP_CHECK(hipGetLastError());
  std::vector<double> h(4 * kReduceBlocks);
  SLSIM_HIP_CHECK(hipMemcpyAsync(h.data(), part, 8 * h.size(), hipMemcpyDeviceToHost,
                                 ctx->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  lo[0] = lo[1] = HUGE_VAL;
  hi[0] = hi[1] = -HUGE_VAL;
  for (unsigned b = 0; b < kReduceBlocks; ++b) {
    lo[0] = std::min(lo[0], h[4 * b]);
    lo[1] = std::min(lo[1], h[4 * b + 1]);
    hi[0] = std::max(hi[0], h[4 * b + 2]);
    hi[1] = std::max(hi[1], h[4 * b + 3]);
  }
  return SLSIM_OK;
}

// A refinement's rays: one device shoot, or one ray-list request of the
// grid's server (continuous batching with other clients' requests; the same
// results bit for bit).
int slsim_agrid_s::shoot_host(int64_t count, const double* hx, double* hy, double* hA) {
  if (count == 0) return SLSIM_OK;
  uint64_t id = 0;
  int rc;
  if ((rc = slsim_server_submit_rays(server, lens, (size_t)count, hx, nullptr, precision, &id)))
    return rc;
  slsim_request_info info;
  if ((rc = slsim_server_wait(server, id, hy, hA, &info))) return rc;
  return info.status;
}

int slsim_agrid_s::shoot_dev(int64_t count, const double* dx, double* dy, double* dA) {
  if (!server) return slsim_shoot_device(lens, (size_t)count, dx, nullptr, dy, dA, precision,
                                         nullptr);
  std::vector<double> hx(2 * (size_t)count), hy(2 * (size_t)count), hA(4 * (size_t)count);
  hipStream_t st = ctx->stream;
  SLSIM_HIP_CHECK(hipMemcpyAsync(hx.data(), dx, 8 * hx.size(), hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  int rc;
  if ((rc = shoot_host(count, hx.data(), hy.data(), hA.data()))) return rc;
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dy, hy.data(), 8 * hy.size(), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dA, hA.data(), 8 * hA.size(), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  return SLSIM_OK;
}

extern "C" {

int slsim_agrid_set_server(slsim_agrid g, slsim_server server) {
  if (!g) return arg_error("null argument");
  g->server = server;
  return SLSIM_OK;
}

int slsim_agrid_create(slsim_lens lens, size_t Nx, const double* center, double range_x,
                       double range_y, int square, int precision, slsim_agrid* out) {
  if (!lens || !center || !out) return arg_error("null argument");
  *out = nullptr;
  if (Nx < 1) return arg_error("Grid size is < 1");
  if (!(range_x > 0) || !(range_y > 0)) return arg_error("Grid range is <= 0");
  if (precision != SLSIM_FP64 && precision != SLSIM_FP32C && precision != SLSIM_DIRECT64)
    return arg_error("unknown precision");
  std::unique_ptr<slsim_agrid_s> g(new slsim_agrid_s);
  g->lens = lens;
  g->ctx = lens->ctx;
  g->precision = precision;
  g->square = square ? 1 : 0;
  g->Ngrid_init = (int64_t)Nx;
  // Ngrid_init2 = (int)(Ngrid_init * axisratio) (grid_maintenance.cpp:89)
  g->Ngrid_init2 = g->square ? (int64_t)Nx : (int64_t)(int)((double)(int)Nx * (range_y / range_x));
  if (g->Ngrid_init2 < 1) return arg_error("Grid y range holds no row of points");
  SLSIM_HIP_CHECK(hipSetDevice(g->ctx->device));
  const int64_t n = g->Ngrid_init * g->Ngrid_init2;
  int rc;
  if ((rc = g->reserve(n))) return rc;
  g->n = n;
  k_agrid_init<<<nblk(n), kBlockSize, 0, g->ctx->stream>>>(n, g->Ngrid_init, g->square,
                                                            center[0], center[1], range_x,
                                                            range_y, g->x, g->gs);
  SLSIM_HIP_CHECK(hipGetLastError());
  if ((rc = slsim_shoot_device(lens, (size_t)n, g->x, nullptr, g->y, g->A, precision, nullptr)))
    return rc;
  // TreeStruct roots: the image points' bounding box; the source points'
  // with a buffer of the range (grid_maintenance.cpp:49-50, 127-128;
  // tree_maintenance.cpp:96-124)
  if ((rc = g->minmax(g->x, n, g->ib.p1, g->ib.p2))) return rc;
  double lo[2], hi[2];
  if ((rc = g->minmax(g->y, n, lo, hi))) return rc;
  const double buffer = g->square ? range_x : std::max(range_x, range_y);
  g->sb.p1[0] = lo[0] - buffer;
  g->sb.p1[1] = lo[1] - buffer;
  g->sb.p2[0] = hi[0] + buffer;
  g->sb.p2[1] = hi[1] + buffer;
  *out = g.release();
  return SLSIM_OK;
}

int slsim_agrid_destroy(slsim_agrid g) {
  if (!g) return SLSIM_OK;
  (void)hipSetDevice(g->ctx->device);
  (void)hipStreamSynchronize(g->ctx->stream);
  delete g;
  return SLSIM_OK;
}

int slsim_agrid_get_info(slsim_agrid g, slsim_agrid_info* info) {
  if (!g || !info) return arg_error("null argument");
  info->npoints = g->n;
  info->ngrid_init = g->Ngrid_init;
  info->ngrid_init2 = g->Ngrid_init2;
  info->ngrid_block = 3;
  for (int k = 0; k < 2; ++k) {
    info->image_p1[k] = g->ib.p1[k];
    info->image_p2[k] = g->ib.p2[k];
    info->source_p1[k] = g->sb.p1[k];
    info->source_p2[k] = g->sb.p2[k];
  }
  return SLSIM_OK;
}

int slsim_agrid_device(slsim_agrid g, double** x_dev, double** y_dev, double** A_dev,
                       double** gridsize_dev) {
  if (!g) return arg_error("null argument");
  if (x_dev) *x_dev = g->x;
  if (y_dev) *y_dev = g->y;
  if (A_dev) *A_dev = g->A;
  if (gridsize_dev) *gridsize_dev = g->gs;
  return SLSIM_OK;
}

int slsim_agrid_read(slsim_agrid g, int64_t first, int64_t count, double* x, double* y,
                     double* A, double* gridsize) {
  if (!g) return arg_error("null argument");
  if (first < 0 || count < 0 || first + count > g->n) return arg_error("points out of range");
  if (count == 0) return SLSIM_OK;
  SLSIM_HIP_CHECK(hipSetDevice(g->ctx->device));
  hipStream_t st = g->ctx->stream;
  const size_t f = (size_t)first, c = (size_t)count;
  if (x) SLSIM_HIP_CHECK(hipMemcpyAsync(x, g->x + 2 * f, 16 * c, hipMemcpyDeviceToHost, st));
  if (y) SLSIM_HIP_CHECK(hipMemcpyAsync(y, g->y + 2 * f, 16 * c, hipMemcpyDeviceToHost, st));
  if (A) SLSIM_HIP_CHECK(hipMemcpyAsync(A, g->A + 4 * f, 32 * c, hipMemcpyDeviceToHost, st));
  if (gridsize)
    SLSIM_HIP_CHECK(hipMemcpyAsync(gridsize, g->gs + f, 8 * c, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  return SLSIM_OK;
}

int slsim_agrid_refine(slsim_agrid g, size_t nparents, const int64_t* ids, int64_t* first,
                       int64_t* count) {
  if (!g || !first || !count || (nparents && !ids)) return arg_error("null argument");
  *first = g->n;
  *count = 0;
  if (nparents == 0) return SLSIM_OK;
  // distinct points of the grid
  std::vector<int64_t> sorted(ids, ids + nparents);
  std::sort(sorted.begin(), sorted.end());
  if (sorted.front() < 0 || sorted.back() >= g->n) return arg_error("point id out of range");
  if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
    return arg_error("a point is listed twice");
  SLSIM_HIP_CHECK(hipSetDevice(g->ctx->device));
  hipStream_t st = g->ctx->stream;
  const int64_t np = (int64_t)nparents, nc = 8 * np;
  // "Cell size is too small for double precision" (grid_maintenance.cpp:
  // 695-699): checked on the host copies of the parents, before any change
  {
    std::vector<double> px(2), pg(1);
    for (size_t k = 0; k < nparents; ++k) {
      SLSIM_HIP_CHECK(hipMemcpyAsync(px.data(), g->x + 2 * ids[k], 16, hipMemcpyDeviceToHost, st));
      SLSIM_HIP_CHECK(hipMemcpyAsync(pg.data(), g->gs + ids[k], 8, hipMemcpyDeviceToHost, st));
      SLSIM_HIP_CHECK(hipStreamSynchronize(st));
      if (pg[0] < std::pow(10., -15) * std::max(std::fabs(px[0]), std::fabs(px[1])))
        return arg_error("Cell size is too small for double precision");
    }
  }
  int rc;
  if ((rc = g->reserve(g->n + nc))) return rc;
  Scratch sc(g->ctx);
  int64_t* dpar;
  double *cx, *cg, *cy, *cA, *kx, *kg;
  uint32_t *keep, *pos, *keep2, *pos2;
  int64_t* kpar;
  if ((rc = sc.get(&dpar, (size_t)np)) || (rc = sc.get(&cx, 2 * (size_t)nc)) ||
      (rc = sc.get(&cg, (size_t)nc)) || (rc = sc.get(&keep, (size_t)nc + 1)) ||
      (rc = sc.get(&pos, (size_t)nc + 1)) || (rc = sc.get(&kx, 2 * (size_t)nc)) ||
      (rc = sc.get(&kg, (size_t)nc)) || (rc = sc.get(&kpar, (size_t)nc)) ||
      (rc = sc.get(&cy, 2 * (size_t)nc)) || (rc = sc.get(&cA, 4 * (size_t)nc)) ||
      (rc = sc.get(&keep2, (size_t)nc + 1)) || (rc = sc.get(&pos2, (size_t)nc + 1)))
    return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dpar, ids, 8 * nparents, hipMemcpyHostToDevice, st));
  k_agrid_children<<<nblk(nc), kBlockSize, 0, st>>>(np, dpar, g->x, g->gs, g->ib, cx, cg, keep);
  SLSIM_HIP_CHECK(hipGetLastError());
  // the candidates' parent (index into the list), for the source pass
  std::vector<int64_t> hpar((size_t)nc);
  for (int64_t t = 0; t < nc; ++t) hpar[(size_t)t] = t / 8;
  int64_t* cpar;
  if ((rc = sc.get(&cpar, (size_t)nc))) return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(cpar, hpar.data(), 8 * (size_t)nc, hipMemcpyHostToDevice, st));
  // image-box pass: stable compaction
  size_t tmp_bytes = 0;
  SLSIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, keep, pos, (int)nc + 1, st));
  char* tmp;
  if ((rc = sc.get(&tmp, tmp_bytes))) return rc;
  SLSIM_HIP_CHECK(hipMemsetAsync(keep + nc, 0, 4, st));
  SLSIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, keep, pos, (int)nc + 1, st));
  uint32_t nadd = 0;
  SLSIM_HIP_CHECK(hipMemcpyAsync(&nadd, pos + nc, 4, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  k_agrid_scatter<<<nblk(nc), kBlockSize, 0, st>>>(nc, keep, pos, cx, nullptr, nullptr, cg, cpar,
                                                   kx, nullptr, nullptr, kg, kpar);
  SLSIM_HIP_CHECK(hipGetLastError());
  if (nadd > 0) {
    // one shoot of every kept point
    if ((rc = g->shoot_dev(nadd, kx, cy, cA))) return rc;
    // source-box pass
    k_agrid_source_keep<<<nblk(nadd), kBlockSize, 0, st>>>(nadd, cy, g->sb, keep2);
    SLSIM_HIP_CHECK(hipGetLastError());
    SLSIM_HIP_CHECK(hipMemsetAsync(keep2 + nadd, 0, 4, st));
    size_t tb2 = 0;
    SLSIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, tb2, keep2, pos2, (int)nadd + 1, st));
    char* tmp2;
    if ((rc = sc.get(&tmp2, tb2))) return rc;
    SLSIM_HIP_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp2, tb2, keep2, pos2, (int)nadd + 1, st));
    uint32_t nfin = 0;
    SLSIM_HIP_CHECK(hipMemcpyAsync(&nfin, pos2 + nadd, 4, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
    const int64_t base = g->n;
    k_agrid_scatter<<<nblk(nadd), kBlockSize, 0, st>>>(
        nadd, keep2, pos2, kx, cy, cA, kg, nullptr, g->x + 2 * base, g->y + 2 * base,
        g->A + 4 * base, g->gs + base, nullptr);
    SLSIM_HIP_CHECK(hipGetLastError());
    // which candidates survived both passes, per parent: a candidate kept
    // by the image pass at slot k survives when keep2[k]
    std::vector<uint32_t> k1((size_t)nc), p1((size_t)nc), k2((size_t)nadd);
    SLSIM_HIP_CHECK(hipMemcpyAsync(k1.data(), keep, 4 * (size_t)nc, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(p1.data(), pos, 4 * (size_t)nc, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(k2.data(), keep2, 4 * (size_t)nadd, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
    for (int64_t t = 0; t < nc; ++t)
      if (k1[(size_t)t]) k1[(size_t)t] = k2[p1[(size_t)t]];
    SLSIM_HIP_CHECK(hipMemcpyAsync(keep, k1.data(), 4 * (size_t)nc, hipMemcpyHostToDevice, st));
    g->n = base + nfin;
    *count = nfin;
  }
  k_agrid_parents<<<nblk(np), kBlockSize, 0, st>>>(np, dpar, keep, g->gs);
  SLSIM_HIP_CHECK(hipGetLastError());
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  return SLSIM_OK;
}

int slsim_agrid_refine_one(slsim_agrid g, int64_t id, int64_t* first, int64_t* count) {
  // Grid::RefineLeaf (grid_maintenance.cpp:543-679): one cell, its points
  // dropped by swaps with the last (the reference's order), the image-box
  // test only for a point on the box's edge
  if (!g || !first || !count) return arg_error("null argument");
  *first = g->n;
  *count = 0;
  if (id < 0 || id >= g->n) return arg_error("point id out of range");
  SLSIM_HIP_CHECK(hipSetDevice(g->ctx->device));
  hipStream_t st = g->ctx->stream;
  double px[2], pg;
  SLSIM_HIP_CHECK(hipMemcpyAsync(px, g->x + 2 * id, 16, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(&pg, g->gs + id, 8, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  if (!(pg > std::pow(10., -15))) return arg_error("cell too small to refine");
  double cx[16], cgs[8];
  {
#pragma clang fp contract(off)
    const double range = pg * (3 - 1) / 3;
    for (int c = 0; c < 8; ++c) {
      const int i = c < 4 ? c : c + 1;
      cx[2 * c] = px[0] + range * (1.0 * (double)(i % 3) / (double)(3 - 1) - 0.5);
      cx[2 * c + 1] = px[1] + range * (1.0 * (double)(i / 3) / (double)(3 - 1) - 0.5);
      cgs[c] = range / (double)(3 - 1);
    }
  }
  auto swap_pt = [](double* a, int w, int i, int j) {
    for (int k = 0; k < w; ++k) std::swap(a[w * i + k], a[w * j + k]);
  };
  auto in_box = [](const double* p, const Box& b) {
    return p[0] >= b.p1[0] && p[0] <= b.p2[0] && p[1] >= b.p1[1] && p[1] <= b.p2[1];
  };
  int nout = 0;
  if (px[0] == g->ib.p1[0] || px[0] == g->ib.p2[0] || px[1] == g->ib.p1[1] ||
      px[1] == g->ib.p2[1]) {
    for (int kk = 0; kk < 8; ++kk) {
      if (!in_box(cx + 2 * (kk - nout), g->ib)) {
        swap_pt(cx, 2, kk - nout, 7 - nout);
        swap_pt(cgs, 1, kk - nout, 7 - nout);
        ++nout;
      }
    }
  }
  const int ntemp = 8 - nout;
  int rc;
  if ((rc = g->reserve(g->n + 8))) return rc;
  double cy[16], cA[32];
  if (ntemp > 0) {
    Scratch sc(g->ctx);
    double *dx, *dy, *dA;
    if ((rc = sc.get(&dx, 16)) || (rc = sc.get(&dy, 16)) || (rc = sc.get(&dA, 32))) return rc;
    SLSIM_HIP_CHECK(hipMemcpyAsync(dx, cx, 16 * (size_t)ntemp, hipMemcpyHostToDevice, st));
    if ((rc = g->shoot_dev(ntemp, dx, dy, dA))) return rc;
    SLSIM_HIP_CHECK(hipMemcpyAsync(cy, dy, 16 * (size_t)ntemp, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(cA, dA, 32 * (size_t)ntemp, hipMemcpyDeviceToHost, st));
    SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  }
  int nout2 = 0;
  for (int kk = 0; kk < ntemp; ++kk) {
    if (!in_box(cy + 2 * (kk - nout2), g->sb)) {
      const int a = kk - nout2, b = ntemp - 1 - nout2;
      swap_pt(cx, 2, a, b);
      swap_pt(cgs, 1, a, b);
      swap_pt(cy, 2, a, b);
      swap_pt(cA, 4, a, b);
      ++nout2;
    }
  }
  const int nadd = ntemp - nout2;
  double ng;
  {
#pragma clang fp contract(off)
    ng = pg / 3;
    if (nadd <= 0) ng *= 3;
  }
  SLSIM_HIP_CHECK(hipMemcpyAsync(g->gs + id, &ng, 8, hipMemcpyHostToDevice, st));
  if (nadd > 0) {
    const size_t b = (size_t)g->n;
    SLSIM_HIP_CHECK(hipMemcpyAsync(g->x + 2 * b, cx, 16 * (size_t)nadd, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(g->y + 2 * b, cy, 16 * (size_t)nadd, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(g->A + 4 * b, cA, 32 * (size_t)nadd, hipMemcpyHostToDevice, st));
    SLSIM_HIP_CHECK(hipMemcpyAsync(g->gs + b, cgs, 8 * (size_t)nadd, hipMemcpyHostToDevice, st));
    g->n += nadd;
    *count = nadd;
  }
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  return SLSIM_OK;
}

int slsim_agrid_einstein_area(slsim_agrid g, double* area) {
  if (!g || !area) return arg_error("null argument");
  SLSIM_HIP_CHECK(hipSetDevice(g->ctx->device));
  Scratch sc(g->ctx);
  double* part;
  int rc;
  if ((rc = sc.get(&part, kReduceBlocks))) return rc;
  k_agrid_einstein<<<kReduceBlocks, kBlockSize, 0, g->ctx->stream>>>(g->n, g->A, g->gs, part);
  SLSIM_HIP_CHECK(hipGetLastError());
  std::vector<double> h(kReduceBlocks);
  SLSIM_HIP_CHECK(hipMemcpyAsync(h.data(), part, 8 * h.size(), hipMemcpyDeviceToHost,
                                 g->ctx->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(g->ctx->stream));
  double s = 0;
  for (double v : h) s += v;
  *area = s;
  return SLSIM_OK;
}

}  // extern "C"
