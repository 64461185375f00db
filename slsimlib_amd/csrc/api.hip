// api.hip -- C ABI of the ray shooter: contexts, planes, lenses, shooting.
//
// The lens bookkeeping restates the reference's plane handling:
//   Lens::addMainHaloToPlane   MultiPlane/lens.cpp:772-815
//   Lens::combinePlanes        MultiPlane/lens.cpp:2397-2515
//   Lens::FindSourcePlane      MultiPlane/lens.cpp:2659-2681
//   Lens::ResetSourcePlane     MultiPlane/lens.cpp:2609-2657 (nearest = false)
//   Lens::rayshooterInternal   FullRange/internal_rayshooter_multi.cpp:53-251
// Shooting runs one kernel launch per lens plane over all 64-ray bundles; the
// per-ray FP64 state lives in HBM between planes.

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "cosmo.h"
#include "slsim_internal.h"

namespace slsim {

static thread_local std::string g_last_error;
void set_error(const std::string& msg) { g_last_error = msg; }

namespace {

int arg_error(const char* msg) {
  set_error(msg);
  return SLSIM_ERR_ARG;
}




int alloc_state(Scratch& s, int64_t nslots, RayState* st) {
  char* d;
  int rc = s.get(&d, kStateBytesPerSlot * (size_t)nslots);
  if (rc) return rc;
  *st = state_view(reinterpret_cast<double*>(d), nslots);
  return SLSIM_OK;
}

double now_ms() {
  return std::chrono::duration<double, std::milli>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

}  // namespace
}  // namespace slsim

using namespace slsim;

// ------------------------------------------------- no-allocation entry point
// slsim_shoot_device_noalloc takes every device buffer of a call from one
// fixed array per device, so that callers that cannot allocate (serving
// loops and harnesses that forbid hipMalloc) can shoot.  Its address on each
// device is looked up once, when the first context on that device is
// created; host code uses that address, never the array's name.
//
// Size of one call of n <= 2^17 rays (all of it from the array): 128 B of ray
// state per slot over n + 63 slots (16.8 MB), Morton keys and indices 24 n B
// and the radix sort's temporary (about 12 n B); for the precision guard its
// lists (28 n B) and outputs (48 n B), and its own FP64 shoot of at most n
// rays over at most max(n + 63, 64 x 4096) slots (33.6 MB) with 36 n B of
// keys and sort space: under 80 MB in all, 128 MiB held per device.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
constexpr size_t kNoAllocMaxRays = (size_t)1 << 17;
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_setup_mu;
char* g_noalloc_ptr[kNoAllocMaxDevices];            // the array's address per device
std::mutex g_noalloc_call_mu[kNoAllocMaxDevices];   // calls on one device in turn

// With `device` current: look the array's address up once for the device.
int noalloc_setup(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return SLSIM_OK;  // the entry point refuses
  std::lock_guard<std::mutex> lk(g_noalloc_setup_mu);
  if (g_noalloc_ptr[device]) return SLSIM_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) {
    set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    return SLSIM_ERR_HIP;
  }
  g_noalloc_ptr[device] = static_cast<char*>(p);
  return SLSIM_OK;
}

// Bump allocation from the fixed array (256-byte aligned pieces); running
// out is an error, never an allocation.
struct FixedScratch {
  static constexpr bool kMayAllocate = false;
  char* base;
  size_t cap, off = 0;
  FixedScratch(char* b, size_t c) : base(b), cap(c) {}
  template <typename T>
  int get(T** p, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    const size_t at = (off + 255) & ~(size_t)255;
    if (at > cap || bytes > cap - at) {
      set_error("slsim_shoot_device_noalloc: the call needs more than its fixed scratch array");
      return SLSIM_ERR_BUDGET;
    }
    *p = reinterpret_cast<T*>(base + at);
    off = at + bytes;
    return SLSIM_OK;
  }
};

}  // namespace

// ------------------------------------------------------------------ context
int slsim_ctx_s::alloc(void** p, size_t bytes) {
  std::lock_guard<std::mutex> g(mu);
  if (used + bytes > budget) {
    set_error("HBM budget exceeded: need " + std::to_string(bytes) + " B, " +
              std::to_string(budget - used) + " B left");
    return SLSIM_ERR_BUDGET;
  }
  hipError_t e = hipMalloc(p, std::max<size_t>(bytes, 16));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return SLSIM_ERR_HIP;
  }
  used += bytes;
  allocs.push_back({*p, bytes});
  return SLSIM_OK;
}

int slsim_ctx_s::release(void* p) {
  if (!p) return SLSIM_OK;
  std::lock_guard<std::mutex> g(mu);
  for (size_t i = 0; i < allocs.size(); ++i) {
    if (allocs[i].first == p) {
      used -= allocs[i].second;
      allocs[i] = allocs.back();
      allocs.pop_back();
      (void)hipFree(p);
      return SLSIM_OK;
    }
  }
  set_error("release of an unknown device pointer");
  return SLSIM_ERR_ARG;
}

// Scratch blocks: best fit among cached blocks no more than twice the
// request, else a fresh allocation (dropping the cache if the budget is short).
int slsim_ctx_s::scratch_get(void** p, size_t bytes) {
  {
    std::lock_guard<std::mutex> g(mu);
    size_t best = (size_t)-1, best_sz = (size_t)-1;
    for (size_t i = 0; i < cache.size(); ++i) {
      const size_t sz = cache[i].second;
      if (sz >= bytes && sz <= 2 * bytes + 4096 && sz < best_sz) {
        best = i;
        best_sz = sz;
      }
    }
    if (best != (size_t)-1) {
      *p = cache[best].first;
      cache[best] = cache.back();
      cache.pop_back();
      return SLSIM_OK;
    }
  }
  int rc = alloc(p, bytes);
  if (rc == SLSIM_ERR_BUDGET && !cache.empty()) {
    trim_cache();
    rc = alloc(p, bytes);
  }
  return rc;
}

void slsim_ctx_s::scratch_put(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> g(mu);
  for (auto& a : allocs)
    if (a.first == p) {
      cache.push_back(a);
      return;
    }
}

void slsim_ctx_s::trim_cache() {
  std::vector<std::pair<void*, size_t>> c;
  {
    std::lock_guard<std::mutex> g(mu);
    c.swap(cache);
  }
  for (auto& a : c) release(a.first);
}

TreeDesc slsim_plane_s::desc() const {
  TreeDesc t;
  t.nodesD = nodesD;
  t.nodesF = nodesF;
  t.partsD = partsD;
  t.bigD = bigD;
  t.partsF = partsF;
  t.inv_area = inv_area;
  t.mass = mass;
  t.comx = com[0];
  t.comy = com[1];
  t.dipx = dip[0];
  t.dipy = dip[1];
  t.sheetd = sheetd;
  t.nnodes = (int32_t)nnodes;
  t.nparts = (int32_t)n;
  t.kind = kind;
  t.nhalos = kind == kTreeHalos ? (int32_t)n : 0;
  t.hnodes = hnodes;
  t.hbig = hbig;
  t.halos = halos;
  t.nfw = nfw;
  t.hs2 = hs2;
  t.hxl = hxl;
  t.hyl = hyl;
  t.hperiodic = hperiodic;
  t.pad2 = 0;
  return t;
}

extern "C" {

const char* slsim_last_error(void) { return g_last_error.c_str(); }

int slsim_ctx_create(int device, size_t hbm_budget_bytes, slsim_ctx* out) {
  if (!out) return arg_error("null output handle");
  int ndev = 0;
  SLSIM_HIP_CHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return arg_error("no such device");
  SLSIM_HIP_CHECK(hipSetDevice(device));
  int rc = noalloc_setup(device);
  if (rc) return rc;
  auto c = std::make_unique<slsim_ctx_s>();
  c->device = device;
  size_t fr = 0, tot = 0;
  SLSIM_HIP_CHECK(hipMemGetInfo(&fr, &tot));
  c->budget = hbm_budget_bytes ? hbm_budget_bytes : (size_t)(0.9 * (double)fr);
  SLSIM_HIP_CHECK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  *out = c.release();
  return SLSIM_OK;
}

int slsim_ctx_destroy(slsim_ctx ctx) {
  if (!ctx) return SLSIM_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (auto& a : ctx->allocs) (void)hipFree(a.first);
  ctx->allocs.clear();
  ctx->cache.clear();
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return SLSIM_OK;
}

int slsim_ctx_memory(slsim_ctx ctx, size_t* used, size_t* budget) {
  if (!ctx) return arg_error("null context");
  (void)hipStreamSynchronize(ctx->stream);
  ctx->trim_cache();
  if (used) *used = ctx->used;
  if (budget) *budget = ctx->budget;
  return SLSIM_OK;
}

int slsim_ctx_synchronize(slsim_ctx ctx) {
  if (!ctx) return arg_error("null context");
  SLSIM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return SLSIM_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ planes
static int plane_free_device(slsim_plane p) {
  slsim_ctx c = p->ctx;
  void* ptrs[] = {p->nodesD, p->nodesF, p->partsD, p->bigD,  p->partsF, p->box_p1, p->box_p2,
                  p->nbig,   p->npart,  p->hnodes, p->hbig, p->halos,  p->nfw,    p->sheetd};
  for (void* q : ptrs)
    if (q) c->release(q);
  return SLSIM_OK;
}

template <typename T>
static int plane_create(slsim_ctx ctx, const T* x, const T* y, const float* m, const float* h,
                        size_t n, double theta, int bucket, double inv_area, slsim_plane* out,
                        double max_range = -1) {
  if (!ctx || !out) return arg_error("null handle");
  if (n == 0 || !x || !y || !m || !h) return arg_error("empty particle set");
  if (n >= (size_t)1 << 31) return arg_error("more than 2^31 particles in one plane");
  if (bucket < 1) return arg_error("bucket must be >= 1");
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  auto p = std::make_unique<slsim_plane_s>();
  p->ctx = ctx;
  p->n = (int64_t)n;
  p->theta = theta;
  p->bucket = bucket;
  p->inv_area = inv_area;
  p->max_range = max_range;
  int rc = build_plane<T>(ctx, x, y, m, h, n, p.get());
  if (rc) {
    plane_free_device(p.get());
    return rc;
  }
  *out = p.release();
  return SLSIM_OK;
}

extern "C" {

int slsim_plane_create_f64(slsim_ctx ctx, const double* x, const double* y, const float* m,
                           const float* h, size_t n, double theta, int bucket, double inv_area,
                           slsim_plane* out) {
  return plane_create<double>(ctx, x, y, m, h, n, theta, bucket, inv_area, out);
}

int slsim_plane_create_f32(slsim_ctx ctx, const float* x, const float* y, const float* m,
                           const float* h, size_t n, double theta, int bucket, double inv_area,
                           slsim_plane* out) {
  return plane_create<float>(ctx, x, y, m, h, n, theta, bucket, inv_area, out);
}

int slsim_plane_create_range_f64(slsim_ctx ctx, const double* x, const double* y, const float* m,
                                 const float* h, size_t n, double theta, int bucket,
                                 double inv_area, double max_range, slsim_plane* out) {
  return plane_create<double>(ctx, x, y, m, h, n, theta, bucket, inv_area, out, max_range);
}

int slsim_plane_create_range_f32(slsim_ctx ctx, const float* x, const float* y, const float* m,
                                 const float* h, size_t n, double theta, int bucket,
                                 double inv_area, double max_range, slsim_plane* out) {
  return plane_create<float>(ctx, x, y, m, h, n, theta, bucket, inv_area, out, max_range);
}

int slsim_plane_tree_create_nfw(slsim_ctx ctx, const double* x, const double* y,
                                const float* mass, const float* rsize,
                                const float* concentration, size_t n, double sigma_background,
                                double inv_screening_scale, slsim_plane* out) {
  if (!ctx || !out) return arg_error("null handle");
  if (n == 0 || !x || !y || !mass || !rsize || !concentration) return arg_error("no halos");
  if (n >= (size_t)1 << 22) return arg_error("more than 2^22 halos in one plane");
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  auto p = std::make_unique<slsim_plane_s>();
  p->ctx = ctx;
  // planes.cpp:13-19: screened trees take theta 0.1 and the periodic buffer,
  // unscreened ones TreeQuadHalos' defaults (theta 0.5, no buffer)
  const bool scr = inv_screening_scale != 0;
  const std::vector<double> param(concentration, concentration + n);
  const int rc = build_halo_plane(p.get(), x, y, mass, rsize, nullptr, param.data(), n,
                                  sigma_background, scr ? 0.1 : 0.5);
  if (rc) {
    plane_free_device(p.get());
    return rc;
  }
  p->hs2 = inv_screening_scale * inv_screening_scale;
  p->hperiodic = scr;
  *out = p.release();
  return SLSIM_OK;
}

int slsim_halos_create(slsim_ctx ctx, const double* x, const double* y, const float* mass,
                       const float* rsize, const int32_t* kind, const double* param, size_t n,
                       double inv_area, double theta_force, double inv_screening_scale,
                       int periodic, slsim_plane* out) {
  if (!ctx || !out) return arg_error("null handle");
  if (n == 0 || !x || !y || !mass || !rsize || !kind || !param) return arg_error("no halos");
  if (n >= (size_t)1 << 22) return arg_error("more than 2^22 halos in one plane");
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  auto p = std::make_unique<slsim_plane_s>();
  p->ctx = ctx;
  const int rc = build_halo_plane(p.get(), x, y, mass, rsize, kind, param, n, inv_area,
                                  theta_force);
  if (rc) {
    plane_free_device(p.get());
    return rc;
  }
  p->hs2 = inv_screening_scale * inv_screening_scale;
  p->hperiodic = periodic != 0;
  *out = p.release();
  return SLSIM_OK;
}

int slsim_halos_create_nfw(slsim_ctx ctx, const double* x, const double* y, const float* mass,
                           const float* rsize, const float* concentration, size_t n,
                           float inv_area, float theta_force, slsim_plane* out) {
  if (!ctx || !out) return arg_error("null handle");
  if (n == 0 || !x || !y || !mass || !rsize || !concentration) return arg_error("no halos");
  if (n >= (size_t)1 << 22) return arg_error("more than 2^22 halos in one plane");
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  auto p = std::make_unique<slsim_plane_s>();
  p->ctx = ctx;
  const std::vector<double> param(concentration, concentration + n);
  const int rc = build_halo_plane(p.get(), x, y, mass, rsize, nullptr, param.data(), n, inv_area,
                                  theta_force);
  if (rc) {
    plane_free_device(p.get());
    return rc;
  }
  *out = p.release();
  return SLSIM_OK;
}

// the grid of a slsim_grid, checked
static int grid_geom(const slsim_grid* gd, GridGeom* g) {
  if (!gd) return arg_error("null grid");
  if (gd->layout != SLSIM_GRID_SQUARE && gd->layout != SLSIM_GRID_RECT)
    return arg_error("grid layout must be SLSIM_GRID_SQUARE or SLSIM_GRID_RECT");
  if (gd->layout == SLSIM_GRID_SQUARE && gd->nx != gd->ny)
    return arg_error("a square grid has nx == ny");
  if (gd->nx >= 65536 || gd->ny >= 65536) return arg_error("grid sides must be below 65536");
  if (!(gd->range_x > 0) || (gd->layout == SLSIM_GRID_RECT && !(gd->range_y > 0)))
    return arg_error("grid ranges must be positive");
  *g = GridGeom{(int64_t)gd->nx, (int64_t)gd->ny, gd->center[0], gd->center[1], gd->range_x,
                gd->layout == SLSIM_GRID_RECT ? gd->range_y : gd->range_x,
                gd->layout == SLSIM_GRID_RECT, 0};
  return SLSIM_OK;
}

static slsim_grid square_grid(size_t N, double c0, double c1, double range) {
  slsim_grid g;
  g.nx = g.ny = N;
  g.center[0] = c0;
  g.center[1] = c1;
  g.range_x = g.range_y = range;
  g.layout = SLSIM_GRID_SQUARE;
  g.pad = 0;
  return g;
}

int slsim_grid_find_images(slsim_ctx ctx, const slsim_grid* grid, const double* y_dev,
                           const double* A_dev, size_t nsrc, const double* ys,
                           size_t max_per_source, int32_t* multiplicity, double* rays,
                           int64_t* triangles, size_t* nimages) {
  if (!ctx || !nimages || (nsrc && (!ys || !multiplicity || !rays || !triangles)))
    return arg_error("null argument");
  GridGeom g;
  int rc;
  if ((rc = grid_geom(grid, &g))) return rc;
  if (g.nx * g.ny >= 1 && (!y_dev || !A_dev)) return arg_error("null grid outputs");
  if (max_per_source == 0 || max_per_source > 4096)
    return arg_error("find_images: max_per_source must be in [1, 4096]");
  if (nsrc > ((size_t)1 << 31) / max_per_source)
    return arg_error("find_images: nsrc * max_per_source must stay below 2^31");
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  return find_images(ctx, g, y_dev, A_dev, nsrc, ys, max_per_source, multiplicity, rays,
                     triangles, nimages);
}

int slsim_gridmap_find_images(slsim_ctx ctx, size_t N, double c0, double c1, double range,
                              const double* y_dev, const double* A_dev, size_t nsrc,
                              const double* ys, size_t max_per_source, int32_t* multiplicity,
                              double* rays, int64_t* triangles, size_t* nimages) {
  const slsim_grid g = square_grid(N, c0, c1, range);
  return slsim_grid_find_images(ctx, &g, y_dev, A_dev, nsrc, ys, max_per_source, multiplicity,
                                rays, triangles, nimages);
}

static int curves_call(slsim_ctx ctx, const slsim_grid* grid, slsim_curves* out,
                       const std::function<int(const GridGeom&, slsim_curves_s*)>& body) {
  if (!ctx || !out) return arg_error("null argument");
  *out = nullptr;
  GridGeom g;
  int rc;
  if ((rc = grid_geom(grid, &g))) return rc;
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  std::unique_ptr<slsim_curves_s> r(new slsim_curves_s);
  rc = body(g, r.get());
  if (rc == SLSIM_OK) *out = r.release();
  return rc;
}

int slsim_grid_find_crit(slsim_ctx ctx, const slsim_grid* grid, const double* y_dev,
                         const double* A_dev, slsim_curves* out) {
  if (!y_dev || !A_dev) return arg_error("null grid outputs");
  return curves_call(ctx, grid, out, [&](const GridGeom& g, slsim_curves_s* r) {
    return find_crit(ctx, g, y_dev, A_dev, r);
  });
}

int slsim_grid_find_magnification_contour(slsim_ctx ctx, const slsim_grid* grid,
                                          const double* A_dev, double invmag,
                                          slsim_curves* out) {
  if (!A_dev) return arg_error("null grid outputs");
  return curves_call(ctx, grid, out, [&](const GridGeom& g, slsim_curves_s* r) {
    return find_magnification_contour(ctx, g, A_dev, invmag, r);
  });
}

int slsim_grid_find_boundaries_of_caustics(slsim_ctx ctx, const slsim_grid* grid,
                                           const double* y_dev, int32_t* multiplicity,
                                           slsim_curves* out) {
  if (!y_dev) return arg_error("null grid outputs");
  return curves_call(ctx, grid, out, [&](const GridGeom& g, slsim_curves_s* r) {
    return find_boundaries_of_caustics(ctx, g, y_dev, multiplicity, r);
  });
}

int slsim_gridmap_find_crit(slsim_ctx ctx, size_t N, double c0, double c1, double range,
                            const double* y_dev, const double* A_dev, slsim_curves* out) {
  const slsim_grid g = square_grid(N, c0, c1, range);
  return slsim_grid_find_crit(ctx, &g, y_dev, A_dev, out);
}

int slsim_gridmap_find_magnification_contour(slsim_ctx ctx, size_t N, double c0, double c1,
                                             double range, const double* A_dev, double invmag,
                                             slsim_curves* out) {
  const slsim_grid g = square_grid(N, c0, c1, range);
  return slsim_grid_find_magnification_contour(ctx, &g, A_dev, invmag, out);
}

int slsim_gridmap_find_boundaries_of_caustics(slsim_ctx ctx, size_t N, double c0, double c1,
                                              double range, const double* y_dev,
                                              int32_t* multiplicity, slsim_curves* out) {
  const slsim_grid g = square_grid(N, c0, c1, range);
  return slsim_grid_find_boundaries_of_caustics(ctx, &g, y_dev, multiplicity, out);
}

int slsim_find_boundaries(const uint8_t* bitmap, size_t nx, size_t ny, slsim_curves* out) {
  if (!out || (nx * ny && !bitmap)) return arg_error("null argument");
  *out = nullptr;
  std::unique_ptr<slsim_curves_s> r(new slsim_curves_s);
  const int rc = find_boundaries_host(bitmap, nx, ny, r.get());
  if (rc == SLSIM_OK) *out = r.release();
  return rc;
}

int slsim_convex_hull(const double* points, size_t n, int64_t* hull, size_t* nhull) {
  if (!nhull || (n && (!points || !hull))) return arg_error("null argument");
  const std::vector<int64_t> h = convex_hull_host(points, n);
  std::copy(h.begin(), h.end(), hull);
  *nhull = h.size();
  return SLSIM_OK;
}

int slsim_in_curve(const double* poly, size_t n, const double* points, size_t m, int32_t* inside) {
  if ((n && !poly) || (m && (!points || !inside))) return arg_error("null argument");
  for (size_t i = 0; i < m; ++i) inside[i] = in_curve_host(poly, n, points[2 * i], points[2 * i + 1]);
  return SLSIM_OK;
}

int slsim_tightest_hull(slsim_ctx ctx, const double* points, size_t n, slsim_curves* out) {
  if (!ctx || !out || (n && !points)) return arg_error("null argument");
  *out = nullptr;
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  std::unique_ptr<slsim_curves_s> r(new slsim_curves_s);
  std::vector<double> h;
  const int rc = tightest_hull(ctx, points, n, h);
  if (rc) return rc;
  r->pts.swap(h);
  r->off = {0, (int64_t)r->pts.size() / 2};
  r->types = {0};
  r->hits = {0};
  r->hits_replaced = 1;
  *out = r.release();
  return SLSIM_OK;
}

int slsim_windings(const double* x, const double* points, size_t n, int32_t* winding,
                   double* area) {
  if (!x || !winding || !area || (n && !points)) return arg_error("null argument");
  *winding = windings_host(x, points, n, area);
  return SLSIM_OK;
}

int slsim_curves_get_info(slsim_curves c, slsim_curves_info* info) {
  if (!c || !info) return arg_error("null argument");
  info->n_curves = c->types.size();
  info->n_points = c->pts.size() / 2;
  info->hits_replaced = c->hits_replaced;
  return SLSIM_OK;
}

int slsim_curves_read(slsim_curves c, int64_t* offsets, double* points, int32_t* types,
                      int32_t* hits) {
  if (!c || !offsets || (!points && !c->pts.empty()) || (!c->types.empty() && (!types || !hits)))
    return arg_error("null argument");
  std::copy(c->off.begin(), c->off.end(), offsets);
  std::copy(c->pts.begin(), c->pts.end(), points);
  std::copy(c->types.begin(), c->types.end(), types);
  std::copy(c->hits.begin(), c->hits.end(), hits);
  return SLSIM_OK;
}

int slsim_curves_destroy(slsim_curves c) {
  delete c;
  return SLSIM_OK;
}

int slsim_plane_destroy(slsim_plane plane) {
  if (!plane) return SLSIM_OK;
  (void)hipStreamSynchronize(plane->ctx->stream);
  plane_free_device(plane);
  delete plane;
  return SLSIM_OK;
}

int slsim_plane_get_info(slsim_plane p, slsim_plane_info* info) {
  if (!p || !info) return arg_error("null handle");
  std::memset(info, 0, sizeof(*info));
  info->n_particles = p->n;
  info->n_nodes = p->nnodes;
  info->n_leaves = p->nleaves;
  info->n_big_internal = p->nbig_internal;
  info->depth = p->depth;
  info->device_bytes = (int64_t)p->bytes;
  info->build_ms = p->build_ms;
  info->root_p1[0] = p->p1[0]; info->root_p1[1] = p->p1[1];
  info->root_p2[0] = p->p2[0]; info->root_p2[1] = p->p2[1];
  return SLSIM_OK;
}

int slsim_plane_export_nodes(slsim_plane p, double* dbl, int64_t* ints) {
  if (!p || !dbl || !ints) return arg_error("null argument");
  if (p->kind != kTreeParticles) return arg_error("export_nodes: not a particle plane");
  SLSIM_HIP_CHECK(hipSetDevice(p->ctx->device));
  return export_nodes(p, dbl, ints);
}

int slsim_plane_force(slsim_plane p, size_t n, const double* rays, double* out, int precision) {
  if (!p || (n && (!rays || !out))) return arg_error("null argument");
  if (n == 0) return SLSIM_OK;
  slsim_ctx c = p->ctx;
  SLSIM_HIP_CHECK(hipSetDevice(c->device));
  Scratch s(c);
  TreeDesc* dt;
  double *dr, *dout;
  uint32_t* err;
  int rc;
  if ((rc = s.get(&dt, 1)) || (rc = s.get(&dr, 2 * n)) || (rc = s.get(&dout, 5 * n)) ||
      (rc = s.get(&err, 1)))
    return rc;
  TreeDesc t = p->desc();
  SLSIM_HIP_CHECK(hipMemcpyAsync(dt, &t, sizeof(t), hipMemcpyHostToDevice, c->stream));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dr, rays, 16 * n, hipMemcpyHostToDevice, c->stream));
  SLSIM_HIP_CHECK(hipMemsetAsync(err, 0, 4, c->stream));
  if (p->kind == kTreeHalos)
    launch_halo_force(c->stream, dt, dr, (int64_t)n, 0, dout, err);
  else
    launch_force(c->stream, precision, dt, dr, (int64_t)n, dout, err);
  SLSIM_HIP_CHECK(hipGetLastError());
  uint32_t herr = 0;
  SLSIM_HIP_CHECK(hipMemcpyAsync(out, dout, 40 * n, hipMemcpyDeviceToHost, c->stream));
  SLSIM_HIP_CHECK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (herr) { set_error("traversal stack overflow"); return SLSIM_ERR_TOO_DEEP; }
  return SLSIM_OK;
}

int slsim_plane_direct_sum(slsim_plane p, size_t n, const double* rays, double* out) {
  if (!p || (n && (!rays || !out))) return arg_error("null argument");
  if (n == 0) return SLSIM_OK;
  slsim_ctx c = p->ctx;
  SLSIM_HIP_CHECK(hipSetDevice(c->device));
  Scratch s(c);
  double *dr, *dout;
  int rc;
  if ((rc = s.get(&dr, 2 * n)) || (rc = s.get(&dout, 5 * n))) return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dr, rays, 16 * n, hipMemcpyHostToDevice, c->stream));
  if (p->kind == kTreeHalos) {
    TreeDesc* dt;
    if ((rc = s.get(&dt, 1))) return rc;
    TreeDesc t = p->desc();
    SLSIM_HIP_CHECK(hipMemcpyAsync(dt, &t, sizeof(t), hipMemcpyHostToDevice, c->stream));
    launch_halo_force(c->stream, dt, dr, (int64_t)n, 1, dout, nullptr);
  } else {
    launch_direct(c->stream, p->partsD, p->n, p->inv_area, dr, (int64_t)n, dout);
  }
  SLSIM_HIP_CHECK(hipGetLastError());
  SLSIM_HIP_CHECK(hipMemcpyAsync(out, dout, 40 * n, hipMemcpyDeviceToHost, c->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SLSIM_OK;
}

}  // extern "C"

// -------------------------------------------------------------------- lens
double slsim_lens_s::coorDist(double z1, double z2) const {
  FlatLCDM c{Om, OL, h};
  return c.coorDist(z1, z2);
}

// MultiPlane/lens.cpp:2397-2515: the field planes' distances then the main
// planes', sorted together (a field plane ahead of a main plane at the same
// distance)
void slsim_lens_s::combinePlanes() {
  plane_redshifts.clear();
  combined.clear();
  Dl = field_Dl;
  Dl.insert(Dl.end(), main_Dl.begin(), main_Dl.end());
  dDl.clear();
  std::vector<int> order(Dl.size());
  for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return Dl[a] < Dl[b]; });
  std::sort(Dl.begin(), Dl.end());
  for (size_t i = 1; i < Dl.size(); ++i)
    if ((Dl[i] - Dl[i - 1]) < kMinPlaneDist) Dl[i] = Dl[i - 1] + kMinPlaneDist;
  const int nf = (int)field_Dl.size();
  for (int i : order) {
    if (i < nf) {
      plane_redshifts.push_back(field_z[i]);
      combined.push_back({field_planes[i]});
    } else {
      plane_redshifts.push_back(main_z[i - nf]);
      combined.push_back(main_planes[i - nf]);
    }
  }
  if (!plane_redshifts.empty() && zsource <= plane_redshifts.back())
    zsource = plane_redshifts.back() + 0.1;
  plane_redshifts.push_back(zsource);
  Dl.push_back(coorDist(0, zsource));
  dDl.push_back(Dl[0]);
  for (size_t i = 1; i < Dl.size(); ++i) dDl.push_back(Dl[i] - Dl[i - 1]);
  trees_dirty = true;
}

void slsim_lens_s::FindSourcePlane(double zs, long& jmax, double& Dls, double& Ds) const {
  Ds = coorDist(0, zs);
  if (zs <= plane_redshifts[0]) { jmax = 0; Dls = Ds; return; }
  jmax = (long)plane_redshifts.size() - 1;
  if (jmax <= 0) { Dls = Ds; return; }
  while (plane_redshifts[jmax - 1] > zs) --jmax;
  Dls = coorDist(plane_redshifts[jmax - 1], zs);
}

int slsim_lens_s::upload_trees() {
  if (!trees_dirty) return SLSIM_OK;
  std::vector<TreeDesc> all;
  tree_off.clear();
  tree_cnt.clear();
  tree_hcnt.clear();
  tree_hkind.clear();
  for (const std::vector<slsim_plane>& trees : combined) {
    // particle trees first, then the halo-of-halos trees of the plane
    tree_off.push_back((int)all.size());
    int np = 0, nh = 0;
    for (slsim_plane p : trees)
      if (p->kind == kTreeParticles) { all.push_back(p->desc()); ++np; }
    int hk = 0;
    for (slsim_plane p : trees)
      if (p->kind == kTreeHalos) {
        all.push_back(p->desc());
        ++nh;
        hk = std::max(hk, p->hpowerlaw ? 2 : 1);
      }
    tree_hkind.push_back(hk);
    tree_cnt.push_back(np);
    tree_hcnt.push_back(nh);
  }
  if (trees_dev) ctx->release(trees_dev);
  trees_dev = nullptr;
  if (!all.empty()) {
    int rc = ctx->alloc(reinterpret_cast<void**>(&trees_dev), sizeof(TreeDesc) * all.size());
    if (rc) return rc;
    SLSIM_HIP_CHECK(hipMemcpyAsync(trees_dev, all.data(), sizeof(TreeDesc) * all.size(),
                                   hipMemcpyHostToDevice, ctx->stream));
    SLSIM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  }
  trees_dirty = false;
  return SLSIM_OK;
}

namespace slsim {

// Source selection of compute_points_parallel for a uniform source redshift
// (include/lens.h:940-953): jmax and the last plane's (aa, bb).
void uniform_source(const slsim_lens_s* L, double zs, int32_t* jmax, double* aal, double* bbl) {
  double source_z = zs;
  if (!(zs > 0)) source_z = L->toggle_source_plane ? L->zs_implant : L->plane_redshifts.back();
  const int nplanes = (int)L->plane_redshifts.size() - 1;
  if (source_z == L->plane_redshifts.back()) {
    *jmax = nplanes;
    *aal = L->dDl.back() / L->Dl.back();
    *bbl = L->Dl[L->Dl.size() - 2] / L->Dl.back();
  } else {
    long jm;
    double Dls, Ds;
    L->FindSourcePlane(source_z, jm, Dls, Ds);
    *jmax = (int32_t)jm;
    *aal = Dls / Ds;
    *bbl = jm > 0 ? L->Dl[jm - 1] / Ds : 0.0;
  }
}

// Per-plane step descriptors for one ray job (one per lens plane).
void make_steps_into(const slsim_lens_s* L, const RayState& st, int64_t nslots,
                     uint32_t* counters, StepJob* steps) {
  const int nplanes = (int)L->plane_redshifts.size() - 1;
  for (int j = 0; j < nplanes; ++j) {
    StepJob& s = steps[j];
    std::memset(&s, 0, sizeof(s));
    s.sbase = st.th0;
    s.nslots = (int32_t)nslots;
    s.j = j;
    s.ntrees = L->tree_cnt[j];
    s.nhtrees = L->tree_hcnt[j];
    s.hkind = L->tree_hkind[j];
    s.hpad = 0;
    s.trees = L->trees_dev + L->tree_off[j];
    s.Dphys = L->Dl[j] / (1 + L->plane_redshifts[j]);
    s.Gfac = L->charge * L->Dl[j] / (1 + L->plane_redshifts[j]);
    s.aa = L->dDl[j + 1] / L->Dl[j + 1];
    s.bb = L->Dl[j] / L->Dl[j + 1];
    s.charge = L->charge;
    s.counters = counters;
    double ascale = 0;
    for (slsim_plane p : L->combined[j])
      if (p->kind == kTreeParticles && p->p2[0] > p->p1[0])
        ascale += std::fabs(p->mass) / (kPI * 0.5 * (p->p2[0] - p->p1[0]));
    s.ascale = ascale;
    s.eps = kGuardEps;
  }
}

void make_steps(const slsim_lens_s* L, const RayState& st, int64_t nslots, uint32_t* counters,
                std::vector<StepJob>& steps) {
  steps.resize(L->plane_redshifts.size() - 1);
  make_steps_into(L, st, nslots, counters, steps.data());
}

}  // namespace slsim

namespace {

// Runs the plane steps [0, jmax) of one job on the context stream.
int run_steps(slsim_ctx c, const StepJob* djobs, const StepJob* steps, int jmax,
              uint32_t nbundles, int precision, bool count, uint32_t* err,
              slsim_shoot_stats* stats, uint32_t lone_rays = 0) {
  const bool timed = stats != nullptr;
  std::vector<hipEvent_t> ev(timed ? 2 * (size_t)jmax : 0);
  if (timed)
    for (auto& e : ev) SLSIM_HIP_CHECK(hipEventCreate(&e));
  for (int j = 0; j < jmax; ++j) {
    if (timed) SLSIM_HIP_CHECK(hipEventRecord(ev[2 * j], c->stream));
    if (lone_rays > 0)
      launch_plane_step_lone(c->stream, djobs + j, lone_rays, err);
    else
      launch_plane_step(precision, count, nbundles, c->stream, djobs + j, nullptr, err,
                        steps[j].nhtrees > 0 ? steps[j].hkind : 0);
    if (timed) SLSIM_HIP_CHECK(hipEventRecord(ev[2 * j + 1], c->stream));
  }
  SLSIM_HIP_CHECK(hipGetLastError());
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (timed) {
    double tot = 0, mx = 0;
    for (int j = 0; j < jmax; ++j) {
      float ms = 0;
      SLSIM_HIP_CHECK(hipEventElapsedTime(&ms, ev[2 * j], ev[2 * j + 1]));
      tot += ms;
      mx = std::max(mx, (double)ms);
    }
    stats->walk_ms = tot;
    stats->walk_ms_max_plane = mx;
    stats->n_walk_launches = jmax;
    for (auto& e : ev) (void)hipEventDestroy(e);
  }
  return SLSIM_OK;
}

int check_err(slsim_ctx c, uint32_t* err) {
  uint32_t herr = 0;
  SLSIM_HIP_CHECK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, c->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (herr) {
    set_error("traversal stack overflow");
    return SLSIM_ERR_TOO_DEEP;
  }
  return SLSIM_OK;
}

template <typename ScratchT>
int guard_refine(slsim_lens L, const RayState& st, int64_t nslots, const GuardMap& m,
                 int64_t nrays, const double* zs_dev, double zs_uniform, double* y_dev,
                 double* A_dev, ScratchT& s, int32_t* nguard);

// rayshooterInternal on device buffers (FullRange/internal_rayshooter_multi.cpp:
// 53-251): rays Morton-ordered into 64-ray bundles, one launch per plane, the
// results scattered back to input order.  Device scratch comes from `s`
// (the context's cache, or the fixed array of the no-allocation entry
// point); the lens's trees must be resident.  Host buffers are fixed-size
// arrays, so a call allocates nothing beyond `s`.
template <typename ScratchT>
int shoot_array(slsim_lens L, size_t n, const double* x_dev, const double* zs_dev,
                double* y_dev, double* A_dev, int precision, slsim_shoot_stats* stats,
                ScratchT& s, int lanes = kWave, double zs_uniform = 0.0) {
  const double t0 = now_ms();
  if (precision < SLSIM_FP64 || precision > SLSIM_DIRECT64) return arg_error("bad precision");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  if (n == 0) return SLSIM_OK;
  if (!x_dev || !y_dev || !A_dev) return arg_error("null buffer");
  if (n >= (size_t)1 << 32) return arg_error("too many rays in one call");
  if (L->trees_dirty) return arg_error("lens trees are not on the device");
  slsim_ctx c = L->ctx;
  const int nplanes = (int)L->plane_redshifts.size() - 1;
  if (nplanes > kMaxPlanesFixed && !ScratchT::kMayAllocate)
    return arg_error("too many lens planes for the no-allocation entry point");
  // lanes == 0: one wavefront per ray (the lone-ray FP64 walk) where every
  // plane has particle trees only, else full bundles; rays dense in the state
  const bool lone_asked = lanes == 0;
  if (lone_asked) lanes = kWave;
  if (lanes < 1 || lanes > kWave) return arg_error("bad rays per bundle");
  const int64_t nbundles = ((int64_t)n + lanes - 1) / lanes;
  const int64_t nslots = nbundles * kWave;
  int rc;
  RayState st;
  {
    char* d;
    if ((rc = s.get(&d, kStateBytesPerSlot * (size_t)nslots))) return rc;
    st = state_view(reinterpret_cast<double*>(d), nslots);
  }
  uint32_t* err;
  if ((rc = s.get(&err, 1))) return rc;
  launch_zero_u32(c->stream, err, 1);
  // Morton order of the rays: each wavefront gets a compact bundle
  const int nblk = (int)std::min<int64_t>(kMinmaxBlocks, (n + 255) / 256);
  double* part;
  if ((rc = s.get(&part, 4 * (size_t)nblk))) return rc;
  launch_minmax_rays(c->stream, x_dev, (int64_t)n, part, nblk);
  double hp[4 * kMinmaxBlocks];
  SLSIM_HIP_CHECK(hipMemcpyAsync(hp, part, 32 * (size_t)nblk, hipMemcpyDeviceToHost, c->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
  for (int b = 0; b < nblk; ++b) {
    lo[0] = std::min(lo[0], hp[4 * b]); lo[1] = std::min(lo[1], hp[4 * b + 1]);
    hi[0] = std::max(hi[0], hp[4 * b + 2]); hi[1] = std::max(hi[1], hp[4 * b + 3]);
  }
  const double w = std::max(std::max(hi[0] - lo[0], hi[1] - lo[1]), 1e-300);
  uint64_t *code, *code_s;
  uint32_t *idx, *perm;
  if ((rc = s.get(&code, n)) || (rc = s.get(&code_s, n)) || (rc = s.get(&idx, n)) ||
      (rc = s.get(&perm, n)))
    return rc;
  launch_ray_morton(c->stream, x_dev, (int64_t)n, lo[0], lo[1], 1.0 / w, code, idx);
  size_t sb = 0;
  SLSIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sb, code, code_s, idx, perm, (int)n,
                                                     0, 30, c->stream));
  char* stmp;
  if ((rc = s.get(&stmp, sb))) return rc;
  SLSIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(stmp, sb, code, code_s, idx, perm, (int)n, 0,
                                                     30, c->stream));
  int32_t jmax = 0;
  double aal = 0, bbl = 0;
  if (nplanes > 0 && !zs_dev) uniform_source(L, zs_uniform, &jmax, &aal, &bbl);
  double *zpl, *dl;
  if ((rc = s.get(&zpl, L->plane_redshifts.size())) || (rc = s.get(&dl, L->Dl.size()))) return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(zpl, L->plane_redshifts.data(), 8 * L->plane_redshifts.size(),
                                 hipMemcpyHostToDevice, c->stream));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dl, L->Dl.data(), 8 * L->Dl.size(), hipMemcpyHostToDevice, c->stream));
  FlatLCDM cosmo{L->Om, L->OL, L->h};
  launch_init_array(c->stream, st, nslots, (int64_t)n, perm, x_dev, nplanes > 0 ? zs_dev : nullptr,
                    jmax, aal, bbl, zpl, dl, nplanes, cosmo, lanes);
  const int steps_to_run = zs_dev ? nplanes : jmax;
  if (steps_to_run > 0) {
    StepJob steps_fixed[kMaxPlanesFixed];
    std::vector<StepJob> steps_more;
    StepJob* steps = steps_fixed;
    if (nplanes > kMaxPlanesFixed) {
      steps_more.resize(nplanes);
      steps = steps_more.data();
    }
    make_steps_into(L, st, nslots, nullptr, steps);
    bool lone = lone_asked && precision == SLSIM_FP64;
    for (int j = 0; j < steps_to_run && lone; ++j) lone = steps[j].nhtrees == 0;
    StepJob* djobs;
    if ((rc = s.get(&djobs, (size_t)nplanes))) return rc;
    SLSIM_HIP_CHECK(hipMemcpyAsync(djobs, steps, sizeof(StepJob) * (size_t)nplanes,
                                   hipMemcpyHostToDevice, c->stream));
    rc = run_steps(c, djobs, steps, steps_to_run, (uint32_t)nbundles, precision, false, err,
                   stats, lone ? (uint32_t)n : 0u);
    if (rc) return rc;
  }
  launch_finalize_array(c->stream, st, nslots, (int64_t)n, perm, y_dev, A_dev, lanes);
  SLSIM_HIP_CHECK(hipGetLastError());
  if ((rc = check_err(c, err))) return rc;
  if (precision == SLSIM_FP32C && steps_to_run > 0 && lanes == kWave && !lone_asked) {
    GuardMap m{};
    m.perm = perm;
    m.x = x_dev;
    m.zs = zs_dev;
    m.n = (int64_t)n;
    int32_t ng = 0;
    if ((rc = guard_refine(L, st, nslots, m, (int64_t)n, zs_dev, zs_uniform, y_dev, A_dev, s,
                           &ng)))
      return rc;
    if (stats) stats->n_guarded = ng;
  }
  if (stats) stats->total_ms = now_ms() - t0;
  return SLSIM_OK;
}

// The precision form's guard (SLSIM_FP32C): after the FP32 walk's shoot,
// the rays the walk flagged and the rays with |alpha| below kGuardSmallAlpha
// of the call's rms are shot again in FP64 (a ray's result does not depend
// on its bundle mates, so these equal the FP64 path's) and their outputs
// replaced.
template <typename ScratchT>
int guard_refine(slsim_lens L, const RayState& st, int64_t nslots, const GuardMap& m,
                 int64_t nrays, const double* zs_dev, double zs_uniform, double* y_dev,
                 double* A_dev, ScratchT& s, int32_t* nguard) {
  slsim_ctx c = L->ctx;
  int rc;
  double *acc, *part;
  uint32_t* count;
  if ((rc = s.get(&acc, 1)) || (rc = s.get(&count, 1)) || (rc = s.get(&part, kGuardBlocks)))
    return rc;
  launch_zero_u32(c->stream, reinterpret_cast<uint32_t*>(acc), 2);
  launch_zero_u32(c->stream, count, 1);
  launch_guard_sumsq(c->stream, st, nslots, m, part, acc);
  const uint32_t cap = (uint32_t)nrays;
  uint32_t* idx;
  double *xl, *zl = nullptr;
  if ((rc = s.get(&idx, cap)) || (rc = s.get(&xl, 2 * (size_t)cap))) return rc;
  if (zs_dev && (rc = s.get(&zl, cap))) return rc;
  launch_guard_flag(c->stream, st, nslots, m, nrays, acc, count, cap, idx, xl, zl);
  uint32_t h = 0;
  SLSIM_HIP_CHECK(hipMemcpyAsync(&h, count, 4, hipMemcpyDeviceToHost, c->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (nguard) *nguard = (int32_t)h;
  if (h == 0) return SLSIM_OK;
  double *yt, *At;
  if ((rc = s.get(&yt, 2 * (size_t)h)) || (rc = s.get(&At, 4 * (size_t)h))) return rc;
  // few scattered rays: one wavefront per ray (the lone-ray walk), else
  // fewer rays per bundle, so that each wavefront walks what its few rays
  // need (a full bundle of scattered rays walks their union)
  int lanes = 1;
  while (lanes < kWave && (int64_t)h > (int64_t)lanes * kGuardWaves) lanes *= 2;
  if ((int64_t)h <= kGuardLoneRays) lanes = 0;
  if ((rc = shoot_array(L, h, xl, zl, yt, At, SLSIM_FP64, nullptr, s, lanes, zs_uniform)))
    return rc;
  launch_guard_scatter(c->stream, h, idx, yt, At, y_dev, A_dev);
  SLSIM_HIP_CHECK(hipGetLastError());
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SLSIM_OK;
}

int gridmap_run(slsim_lens L, size_t N, double c0, double c1, double range, double zs,
                double* y_dev, double* A_dev, int precision, uint64_t* cnt_host,
                slsim_shoot_stats* stats, uint64_t* bundle_stats = nullptr) {
  const double t0 = now_ms();
  if (!L) return arg_error("null lens");
  if (N < 1) return arg_error("GridMap size is < 1!");
  if (!(range > 0)) return arg_error("GridMap range is <= 0!");
  slsim_ctx c = L->ctx;
  SLSIM_HIP_CHECK(hipSetDevice(c->device));
  int rc = L->upload_trees();
  if (rc) return rc;
  const int64_t tiles_x = ((int64_t)N + 7) / 8;
  const int64_t nbundles = tiles_x * tiles_x;
  const int64_t nslots = nbundles * kWave;
  if (nbundles > 0xffffffffll) return arg_error("grid too large");
  Scratch s(c);
  RayState st;
  if ((rc = alloc_state(s, nslots, &st))) return rc;
  uint32_t* err;
  if ((rc = s.get(&err, 1))) return rc;
  SLSIM_HIP_CHECK(hipMemsetAsync(err, 0, 4, c->stream));
  uint32_t* counters = nullptr;
  if (cnt_host) {
    if ((rc = s.get(&counters, kCountStride * (size_t)nslots))) return rc;
    SLSIM_HIP_CHECK(hipMemsetAsync(counters, 0, 4 * kCountStride * (size_t)nslots, c->stream));
  }
  int32_t jmax = 0;
  double aal = 0, bbl = 0;
  const int nplanes = (int)L->plane_redshifts.size() - 1;
  if (nplanes > 0) uniform_source(L, zs, &jmax, &aal, &bbl);
  launch_init_grid(c->stream, st, nslots, (int64_t)N, tiles_x, 3, c0, c1, range, jmax, aal, bbl);
  if (jmax > 0) {
    std::vector<StepJob> steps;
    make_steps(L, st, nslots, counters, steps);
    StepJob* djobs;
    if ((rc = s.get(&djobs, steps.size()))) return rc;
    SLSIM_HIP_CHECK(hipMemcpyAsync(djobs, steps.data(), sizeof(StepJob) * steps.size(),
                                   hipMemcpyHostToDevice, c->stream));
    rc = run_steps(c, djobs, steps.data(), jmax, (uint32_t)nbundles,
                   cnt_host ? SLSIM_FP64 : precision,
                   cnt_host != nullptr, err, stats);
    if (rc) return rc;
  }
  if (y_dev && A_dev)
    launch_finalize_grid(c->stream, st, nslots, (int64_t)N, tiles_x, 3, y_dev, A_dev);
  if (y_dev && A_dev && precision == SLSIM_FP32C && !cnt_host && jmax > 0) {
    if ((rc = check_err(c, err))) return rc;
    GuardMap m{};
    m.N = (int64_t)N;
    m.tiles_x = tiles_x;
    m.tw = 3;
    m.c0 = c0;
    m.c1 = c1;
    m.range = range;
    int32_t ng = 0;
    if ((rc = guard_refine(L, st, nslots, m, (int64_t)N * (int64_t)N, nullptr, zs, y_dev, A_dev, s,
                           &ng)))
      return rc;
    if (stats) stats->n_guarded = ng;
  }
  if (cnt_host) {
    uint64_t* dcnt;
    if ((rc = s.get(&dcnt, 5 * N * N))) return rc;
    launch_counts_grid(c->stream, counters, nslots, (int64_t)N, tiles_x, 3, dcnt);
    SLSIM_HIP_CHECK(hipMemcpyAsync(cnt_host, dcnt, 40 * N * N, hipMemcpyDeviceToHost, c->stream));
    if (bundle_stats) {
      std::vector<uint32_t> hc(kCountStride * (size_t)nslots);
      SLSIM_HIP_CHECK(hipMemcpyAsync(hc.data(), counters, 4 * hc.size(), hipMemcpyDeviceToHost,
                                     c->stream));
      SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
      uint64_t t[4] = {0, 0, 0, 0};
      for (int64_t k = 0; k < nslots; ++k) {
        const uint32_t* q = hc.data() + kCountStride * (size_t)k;
        t[0] += q[0];
        t[1] += q[5];
        t[2] += q[2];
        t[3] += q[6];
      }
      for (int k = 0; k < 4; ++k) bundle_stats[k] = t[k];
    }
  }
  SLSIM_HIP_CHECK(hipGetLastError());
  if ((rc = check_err(c, err))) return rc;
  if (stats) stats->total_ms = now_ms() - t0;
  return SLSIM_OK;
}

}  // namespace

namespace slsim {
int guard_refine_ctx(slsim_lens L, const RayState& st, int64_t nslots, const GuardMap& m,
                     int64_t nrays, const double* zs_dev, double zs_uniform, double* y_dev,
                     double* A_dev, int32_t* nguard) {
  Scratch s(L->ctx);
  return guard_refine(L, st, nslots, m, nrays, zs_dev, zs_uniform, y_dev, A_dev, s, nguard);
}
}  // namespace slsim

extern "C" {

int slsim_lens_create(slsim_ctx ctx, double zsource, double Om, double OL, double h,
                      slsim_lens* out) {
  if (!ctx || !out) return arg_error("null handle");
  if ((Om + OL) != 1.0) {
    set_error("Lens can only handle flat universes at present.");
    return SLSIM_ERR_NOT_FLAT;
  }
  if (!(h > 0) || !(zsource > 0)) return arg_error("h and zsource must be positive");
  auto L = std::make_unique<slsim_lens_s>();
  L->ctx = ctx;
  L->Om = Om;
  L->OL = OL;
  L->h = h;
  L->zsource = zsource;
  L->charge = 4 * kPI * kGrav;
  L->combinePlanes();
  *out = L.release();
  return SLSIM_OK;
}

int slsim_lens_destroy(slsim_lens L) {
  if (!L) return SLSIM_OK;
  (void)hipStreamSynchronize(L->ctx->stream);
  if (L->trees_dev) L->ctx->release(L->trees_dev);
  for (slsim_plane p : L->owned) slsim_plane_destroy(p);
  delete L;
  return SLSIM_OK;
}

int slsim_lens_add_halo(slsim_lens L, slsim_plane plane, float z) {
  if (!L || !plane) return arg_error("null handle");
  if (plane->ctx != L->ctx) return arg_error("plane and lens belong to different contexts");
  for (slsim_plane p : L->owned)
    if (p == plane) return arg_error("halo already in the lens");
  L->owned.push_back(plane);
  const double halo_z = z;
  const double halo_Dl = L->coorDist(0, halo_z);
  auto& mD = L->main_Dl;
  const size_t i = std::upper_bound(mD.begin(), mD.end(), halo_Dl) - mD.begin();
  if (i > 0 && (halo_Dl - mD[i - 1]) < kMinPlaneDist) {
    L->main_planes[i - 1].push_back(plane);
  } else if (i == mD.size()) {
    L->main_planes.push_back({plane});
    L->main_z.push_back(halo_z);
    mD.push_back(halo_Dl);
  } else if ((mD[i] - halo_Dl) < kMinPlaneDist) {
    L->main_planes[i].push_back(plane);
  } else {
    L->main_planes.insert(L->main_planes.begin() + i, {plane});
    L->main_z.insert(L->main_z.begin() + i, halo_z);
    mD.insert(mD.begin() + i, halo_Dl);
  }
  L->combinePlanes();
  return SLSIM_OK;
}

int slsim_lens_clear_main_halos(slsim_lens L) {
  if (!L) return arg_error("null lens");
  (void)hipStreamSynchronize(L->ctx->stream);
  std::vector<slsim_plane> keep;
  for (slsim_plane p : L->owned) {
    bool main = false;
    for (const auto& pl : L->main_planes)
      for (slsim_plane q : pl) main = main || q == p;
    if (main) slsim_plane_destroy(p);
    else keep.push_back(p);
  }
  L->owned.swap(keep);
  L->main_planes.clear();
  L->main_z.clear();
  L->main_Dl.clear();
  L->combinePlanes();
  return SLSIM_OK;
}

int slsim_lens_add_field_plane(slsim_lens L, slsim_plane plane, double z) {
  if (!L || !plane) return arg_error("null handle");
  if (plane->ctx != L->ctx) return arg_error("plane and lens belong to different contexts");
  if (plane->kind != kTreeHalos) return arg_error("a field plane is a tree of halos");
  for (slsim_plane p : L->owned)
    if (p == plane) return arg_error("halo already in the lens");
  L->owned.push_back(plane);
  L->field_planes.push_back(plane);
  L->field_z.push_back(z);
  L->field_Dl.push_back(L->coorDist(0, z));
  L->combinePlanes();
  return SLSIM_OK;
}

int slsim_lens_add_halo_nearest(slsim_lens L, slsim_plane plane, float z) {
  if (!L || !plane) return arg_error("null handle");
  if (plane->ctx != L->ctx) return arg_error("plane and lens belong to different contexts");
  for (slsim_plane p : L->owned)
    if (p == plane) return arg_error("halo already in the lens");
  L->owned.push_back(plane);
  const double halo_z = z;
  const double halo_Dl = L->coorDist(0, halo_z);
  auto& mD = L->main_Dl;
  if (mD.empty()) {
    // a plane of the halo; the reference then adds the halo to the closest
    // plane too, so it is on this plane twice (MultiPlane/lens.cpp:828-837)
    L->main_planes.push_back({plane});
    L->main_z.push_back(halo_z);
    mD.push_back(halo_Dl);
  }
  // the closest main plane in distance (the first of equals)
  size_t i = 0;
  for (size_t k = 1; k < mD.size(); ++k)
    if (std::fabs(mD[k] - halo_Dl) < std::fabs(mD[i] - halo_Dl)) i = k;
  L->main_planes[i].push_back(plane);
  L->combinePlanes();
  return SLSIM_OK;
}

int slsim_lens_tables(slsim_lens L, int* nplanes, double* z, double* Dl, double* dDl,
                      double* charge) {
  if (!L) return arg_error("null lens");
  if (nplanes) *nplanes = (int)L->plane_redshifts.size() - 1;
  for (size_t i = 0; i < L->plane_redshifts.size(); ++i) {
    if (z) z[i] = L->plane_redshifts[i];
    if (Dl) Dl[i] = L->Dl[i];
    if (dDl) dDl[i] = L->dDl[i];
  }
  if (charge) *charge = L->charge;
  return SLSIM_OK;
}

int slsim_lens_reset_source(slsim_lens L, double z) {
  if (!L) return arg_error("null lens");
  if (z <= 0) {
    L->toggle_source_plane = false;
    return SLSIM_OK;
  }
  L->toggle_source_plane = true;
  L->zs_implant = z;
  return SLSIM_OK;
}

// Numerical Recipes' locate (Press et al., Numerical Recipes in C, 2nd ed.,
// sec. 3.4; GLAMER's locateD is its double version): for the 1-based table
// xx[1..n], the jl with xx[jl] <= x < xx[jl+1] by bisection, 1 when x ==
// xx[1] and n - 1 when x == xx[n].
static unsigned long nr_locate(const double* xx, unsigned long n, double x) {
  unsigned long jl = 0, ju = n + 1;
  const bool ascnd = xx[n] >= xx[1];
  while (ju - jl > 1) {
    const unsigned long jm = (ju + jl) >> 1;
    if ((x >= xx[jm]) == ascnd) jl = jm;
    else ju = jm;
  }
  if (x == xx[1]) return 1;
  if (x == xx[n]) return n - 1;
  return jl;
}

int slsim_lens_reset_source_plane(slsim_lens L, double z, int nearest, int32_t* index,
                                  double* zs_implant) {
  if (!L) return arg_error("null lens");
  L->toggle_source_plane = true;
  unsigned long j = 0;
  if (z > 0) {
    const double Ds = L->coorDist(0, z);
    const unsigned long n = L->plane_redshifts.size() - 1;   // lensing planes
    if (L->Dl.size() > 1) j = nr_locate(L->Dl.data() - 1, n, Ds);
    if (j > 0) {
      if (L->Dl[j - 1] == Ds) {
        --j;
      } else if (nearest) {
        const double z1 = slsim_inv_coordist(L->Om, L->OL, L->h, L->Dl[j] - 0.5 * L->dDl[j]);
        if (z < z1) --j;
      }
    }
    L->zs_implant = (nearest && j < n) ? L->plane_redshifts[j] : z;
  }
  if (index) *index = (int32_t)j;
  if (zs_implant) *zs_implant = L->zs_implant;
  return SLSIM_OK;
}

int slsim_gridmap_shoot(slsim_lens L, size_t N, double c0, double c1, double range, double zs,
                        double* y_dev, double* A_dev, int precision, slsim_shoot_stats* stats) {
  if (!y_dev || !A_dev) return arg_error("null output buffer");
  if (precision < SLSIM_FP64 || precision > SLSIM_DIRECT64) return arg_error("bad precision");
  if (stats) std::memset(stats, 0, sizeof(*stats));
  return gridmap_run(L, N, c0, c1, range, zs, y_dev, A_dev, precision, nullptr, stats);
}

int slsim_gridmap_count(slsim_lens L, size_t N, double c0, double c1, double range, double zs,
                        uint64_t* cnt_host) {
  if (!cnt_host) return arg_error("null counter buffer");
  return gridmap_run(L, N, c0, c1, range, zs, nullptr, nullptr, SLSIM_FP64, cnt_host, nullptr);
}

int slsim_gridmap_bundle_stats(slsim_lens L, size_t N, double c0, double c1, double range,
                               double zs, uint64_t* out4) {
  if (!out4) return arg_error("null output");
  std::vector<uint64_t> cnt(5 * N * N);
  return gridmap_run(L, N, c0, c1, range, zs, nullptr, nullptr, SLSIM_FP64, cnt.data(), nullptr,
                     out4);
}

int slsim_shoot_device(slsim_lens L, size_t n, const double* x_dev, const double* zs_dev,
                       double* y_dev, double* A_dev, int precision, slsim_shoot_stats* stats) {
  if (!L) return arg_error("null lens");
  SLSIM_HIP_CHECK(hipSetDevice(L->ctx->device));
  int rc = L->upload_trees();
  if (rc) return rc;
  Scratch s(L->ctx);
  return shoot_array(L, n, x_dev, zs_dev, y_dev, A_dev, precision, stats, s);
}

int slsim_shoot_device_noalloc(slsim_lens L, size_t n, const double* x_dev, const double* zs_dev,
                               double* y_dev, double* A_dev, int precision,
                               slsim_shoot_stats* stats) {
  if (!L) return arg_error("null lens");
  if (n > kNoAllocMaxRays)
    return arg_error("slsim_shoot_device_noalloc: more than 131072 rays in one call");
  const int dev = L->ctx->device;
  if (dev >= kNoAllocMaxDevices || !g_noalloc_ptr[dev])
    return arg_error("slsim_shoot_device_noalloc: no scratch array on this device");
  if (L->trees_dirty)
    return arg_error("slsim_shoot_device_noalloc: the lens's planes are not on the device "
                     "(shoot once with an allocating entry point after changing the lens)");
  SLSIM_HIP_CHECK(hipSetDevice(dev));
  std::lock_guard<std::mutex> lk(g_noalloc_call_mu[dev]);
  FixedScratch s(g_noalloc_ptr[dev], kNoAllocScratchBytes);
  return shoot_array(L, n, x_dev, zs_dev, y_dev, A_dev, precision, stats, s);
}

int slsim_shoot(slsim_lens L, size_t n, const double* x, const double* zs, double* y, double* A,
                int precision, slsim_shoot_stats* stats) {
  if (!L) return arg_error("null lens");
  if (n == 0) return SLSIM_OK;
  if (!x || !y || !A) return arg_error("null buffer");
  slsim_ctx c = L->ctx;
  SLSIM_HIP_CHECK(hipSetDevice(c->device));
  const double t0 = now_ms();
  Scratch s(c);
  double *dx, *dzs = nullptr, *dy, *dA;
  int rc;
  if ((rc = s.get(&dx, 2 * n)) || (rc = s.get(&dy, 2 * n)) || (rc = s.get(&dA, 4 * n))) return rc;
  if (zs && (rc = s.get(&dzs, n))) return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dx, x, 16 * n, hipMemcpyHostToDevice, c->stream));
  if (zs) SLSIM_HIP_CHECK(hipMemcpyAsync(dzs, zs, 8 * n, hipMemcpyHostToDevice, c->stream));
  rc = slsim_shoot_device(L, n, dx, dzs, dy, dA, precision, stats);
  if (rc) return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(y, dy, 16 * n, hipMemcpyDeviceToHost, c->stream));
  SLSIM_HIP_CHECK(hipMemcpyAsync(A, dA, 32 * n, hipMemcpyDeviceToHost, c->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  if (stats) stats->total_ms = now_ms() - t0;
  return SLSIM_OK;
}

// ------------------------------------------------------------ device buffers
int slsim_device_alloc(slsim_ctx ctx, size_t bytes, void** dev) {
  if (!ctx || !dev) return arg_error("null argument");
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  return ctx->alloc(dev, bytes);
}

int slsim_device_free(slsim_ctx ctx, void* dev) {
  if (!ctx) return arg_error("null context");
  (void)hipStreamSynchronize(ctx->stream);
  return ctx->release(dev);
}

int slsim_memcpy_d2h(slsim_ctx ctx, void* host, const void* dev, size_t bytes) {
  if (!ctx) return arg_error("null context");
  SLSIM_HIP_CHECK(hipMemcpyAsync(host, dev, bytes, hipMemcpyDeviceToHost, ctx->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return SLSIM_OK;
}

int slsim_memcpy_h2d(slsim_ctx ctx, void* dev, const void* host, size_t bytes) {
  if (!ctx) return arg_error("null context");
  SLSIM_HIP_CHECK(hipMemcpyAsync(dev, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return SLSIM_OK;
}

int slsim_memcpy_d2d(slsim_ctx ctx, void* dst, const void* src, size_t bytes) {
  if (!ctx) return arg_error("null context");
  SLSIM_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return SLSIM_OK;
}

// --------------------------------------------------------------- cosmology
double slsim_coordist(double Om, double OL, double h, double z1, double z2) {
  FlatLCDM c{Om, OL, h};
  return c.coorDist(z1, z2);
}

double slsim_inv_coordist(double Om, double OL, double h, double D) {
  FlatLCDM c{Om, OL, h};
  double lo = 0.0, hi = 1.0;
  while (c.coorDist(0, hi) < D && hi < 1e4) hi *= 2;
  for (int it = 0; it < 200; ++it) {
    const double mid = 0.5 * (lo + hi);
    if (mid == lo || mid == hi) break;
    if (c.coorDist(0, mid) < D) lo = mid; else hi = mid;
  }
  return 0.5 * (lo + hi);
}

}  // extern "C"
