// smoothing.hip -- particle smoothing sizes on the device: for every particle,
// the distance to its (Nsmooth+1)-th nearest particle in 3-D, itself
// included (LensHaloParticles::calculate_smoothing,
// include/particle_halo.h:683-730, through TreeSimple::NNDistance,
// include/simpleTree.h:483-620).  The value is defined by the particle set
// alone, so it is computed here with a structure of our own:
//
//   1. 63-bit 3-D Morton keys over the bounding box, radix sort.
//   2. Leaves of 32 consecutive particles in key order, and a complete
//      binary tree of bounding boxes over the leaves (heap order: node 1 is
//      the root, the children of b are 2b and 2b+1, leaves at [L, 2L)).
//   3. One wavefront per 64 consecutive particles.  Each lane keeps a max-heap
//      of the K = Nsmooth+1 smallest squared distances in LDS.  An initial
//      bound comes from the 2K particles around it in key order; the tree is
//      then walked with one shared stack and per-lane masks, a box being taken
//      by the lanes whose bound reaches it.
//
// Squared distances use the reference's arithmetic: per dimension the
// difference in the coordinate type T, squared in FP64, summed x, y, z
// (simpleTree.h:518-536), so the K-th smallest is the same number.

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "slsim_internal.h"

namespace slsim {
namespace {

constexpr int kBlk = 256;
constexpr int kLeaf = 32;           // particles per leaf
constexpr int kKnnStack = 64;       // traversal stack entries per wavefront
constexpr int kMaxK = 120;          // K = Nsmooth + 1 limit: the lane heaps stay under 64 KiB of LDS

inline unsigned nblk(size_t n, int b = kBlk) { return (unsigned)((n + b - 1) / b); }

template <typename T>
__global__ void k_bbox3(const T* __restrict__ x, const T* __restrict__ y,
                        const T* __restrict__ z, size_t n, double* __restrict__ part) {
  __shared__ double s[6][kBlk];
  double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (size_t i = blockIdx.x * (size_t)kBlk + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlk) {
    const double p[3] = {(double)x[i], (double)y[i], (double)z[i]};
    for (int j = 0; j < 3; ++j) {
      lo[j] = fmin(lo[j], p[j]);
      hi[j] = fmax(hi[j], p[j]);
    }
  }
  for (int j = 0; j < 3; ++j) {
    s[j][threadIdx.x] = lo[j];
    s[3 + j][threadIdx.x] = hi[j];
  }
  __syncthreads();
  for (int o = kBlk / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o)
      for (int j = 0; j < 3; ++j) {
        s[j][threadIdx.x] = fmin(s[j][threadIdx.x], s[j][threadIdx.x + o]);
        s[3 + j][threadIdx.x] = fmax(s[3 + j][threadIdx.x], s[3 + j][threadIdx.x + o]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * blockIdx.x + threadIdx.x] = s[threadIdx.x][0];
}

__device__ inline uint64_t spread3(uint64_t v) {   // 21 bits -> every third bit
  v &= 0x1fffffull;
  v = (v | (v << 32)) & 0x1f00000000ffffull;
  v = (v | (v << 16)) & 0x1f0000ff0000ffull;
  v = (v | (v << 8)) & 0x100f00f00f00f00full;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
  v = (v | (v << 2)) & 0x1249249249249249ull;
  return v;
}

template <typename T>
__global__ void k_keys3(const T* __restrict__ x, const T* __restrict__ y, const T* __restrict__ z,
                        size_t n, double lx, double ly, double lz, double sx, double sy,
                        double sz, uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
  const size_t i = blockIdx.x * (size_t)kBlk + threadIdx.x;
  if (i >= n) return;
  const double top = 2097151.0;   // 2^21 - 1
  const uint64_t qx = (uint64_t)fmin(top, fmax(0.0, ((double)x[i] - lx) * sx));
  const uint64_t qy = (uint64_t)fmin(top, fmax(0.0, ((double)y[i] - ly) * sy));
  const uint64_t qz = (uint64_t)fmin(top, fmax(0.0, ((double)z[i] - lz) * sz));
  keys[i] = spread3(qx) | (spread3(qy) << 1) | (spread3(qz) << 2);
  idx[i] = (uint32_t)i;
}

template <typename T>
__global__ void k_gather3(const uint32_t* __restrict__ perm, const T* __restrict__ x,
                          const T* __restrict__ y, const T* __restrict__ z, size_t n,
                          T* __restrict__ xs, T* __restrict__ ys, T* __restrict__ zs) {
  const size_t i = blockIdx.x * (size_t)kBlk + threadIdx.x;
  if (i >= n) return;
  const uint32_t p = perm[i];
  xs[i] = x[p];
  ys[i] = y[p];
  zs[i] = z[p];
}

struct Box {
  double lo[3], hi[3];
};

// Leaf boxes; slots past the last leaf are empty (lo > hi: never reached).
template <typename T>
__global__ void k_leaf_boxes(const T* __restrict__ xs, const T* __restrict__ ys,
                             const T* __restrict__ zs, size_t n, uint32_t nleaf, uint32_t L,
                             Box* __restrict__ box) {
  const uint32_t l = blockIdx.x * kBlk + threadIdx.x;
  if (l >= L) return;
  Box b;
  for (int j = 0; j < 3; ++j) {
    b.lo[j] = INFINITY;
    b.hi[j] = -INFINITY;
  }
  if (l < nleaf) {
    const size_t s = (size_t)l * kLeaf, e = std::min(n, s + kLeaf);
    for (size_t i = s; i < e; ++i) {
      const double p[3] = {(double)xs[i], (double)ys[i], (double)zs[i]};
      for (int j = 0; j < 3; ++j) {
        b.lo[j] = fmin(b.lo[j], p[j]);
        b.hi[j] = fmax(b.hi[j], p[j]);
      }
    }
  }
  box[L + l] = b;
}

__global__ void k_merge_boxes(uint32_t first, uint32_t count, Box* __restrict__ box) {
  const uint32_t t = blockIdx.x * kBlk + threadIdx.x;
  if (t >= count) return;
  const uint32_t b = first + t;
  const Box c0 = box[2 * b], c1 = box[2 * b + 1];
  Box m;
  for (int j = 0; j < 3; ++j) {
    m.lo[j] = fmin(c0.lo[j], c1.lo[j]);
    m.hi[j] = fmax(c0.hi[j], c1.hi[j]);
  }
  box[b] = m;
}

// Squared distance with the reference's arithmetic.
template <typename T>
__device__ inline double r2_ref(T px, T py, T pz, T qx, T qy, T qz) {
#pragma clang fp contract(off)
  const double dx = (double)(T)(px - qx), dy = (double)(T)(py - qy), dz = (double)(T)(pz - qz);
  double r2 = dx * dx;
  r2 += dy * dy;
  r2 += dz * dz;
  return r2;
}

// Max-heap of this lane's K smallest values, interleaved over the wave's
// lanes in LDS: element j of lane l at h[64 j + l].
struct LaneHeap {
  double* h;
  int k, n;
  __device__ inline double top() const { return h[0]; }
  __device__ inline void push(double v) {   // n < k
    int i = n++;
    while (i > 0) {
      const int p = (i - 1) >> 1;
      const double pv = h[64 * p];
      if (!(pv < v)) break;
      h[64 * i] = pv;
      i = p;
    }
    h[64 * i] = v;
  }
  __device__ inline void replace_top(double v) {   // n == k, v < top
    int i = 0;
    for (;;) {
      const int c = 2 * i + 1;
      if (c >= n) break;
      double cv = h[64 * c];
      int m = c;
      if (c + 1 < n) {
        const double c2 = h[64 * (c + 1)];
        if (c2 > cv) {
          cv = c2;
          m = c + 1;
        }
      }
      if (!(cv > v)) break;
      h[64 * i] = cv;
      i = m;
    }
    h[64 * i] = v;
  }
  // take v into the K smallest (bound: inclusive limit while filling)
  __device__ inline void offer(double v, double bound) {
    if (n < k) {
      if (v <= bound) push(v);
    } else if (v < top()) {
      replace_top(v);
    }
  }
  __device__ inline double limit(double bound) const { return n < k ? bound : top(); }
};

__device__ inline double box_d2(const Box& b, double qx, double qy, double qz) {
  const double q[3] = {qx, qy, qz};
  double d2 = 0;
  for (int j = 0; j < 3; ++j) {
    const double t = fmax(fmax(b.lo[j] - q[j], q[j] - b.hi[j]), 0.0);
    d2 += t * t;
  }
  return d2;
}

template <typename T>
__global__ __launch_bounds__(64) void k_knn(const T* __restrict__ xs, const T* __restrict__ ys,
                                            const T* __restrict__ zs, const uint32_t* __restrict__ perm,
                                            size_t n, int K, uint32_t L, const Box* __restrict__ box,
                                            double slack, float* __restrict__ out,
                                            uint32_t* __restrict__ err) {
  extern __shared__ double s_heap[];               // 64 K doubles
  __shared__ uint32_t s_node[kKnnStack];
  __shared__ uint64_t s_mask[kKnnStack];
  const int lane = (int)threadIdx.x;
  const size_t i = blockIdx.x * (size_t)64 + lane;
  const bool valid = i < n;
  const size_t iq = valid ? i : n - 1;
  const T qx = xs[iq], qy = ys[iq], qz = zs[iq];
  LaneHeap hp{s_heap + lane, K, 0};
  // initial bound: the K smallest over the 2K particles around it in key order
  size_t w0 = iq > (size_t)K ? iq - K : 0;
  size_t w1 = std::min(n, w0 + 2 * (size_t)K);
  w0 = w1 > 2 * (size_t)K ? w1 - 2 * (size_t)K : 0;
  for (size_t j = w0; j < w1; ++j) hp.offer(r2_ref(xs[j], ys[j], zs[j], qx, qy, qz), INFINITY);
  const double bound = hp.top();
  hp.n = 0;
  // tree walk from the root
  const double qdx = (double)qx, qdy = (double)qy, qdz = (double)qz;
  const uint64_t vmask = __builtin_amdgcn_ballot_w64(valid);
  int sp = 0;
  uint32_t node = 1;
  uint64_t live = vmask;
  for (;;) {
    const Box b = box[node];
    const bool take = ((live >> lane) & 1ull) &&
                      box_d2(b, qdx, qdy, qdz) <= hp.limit(bound) * slack;
    const uint64_t tm = __builtin_amdgcn_ballot_w64(take);
    if (tm) {
      if (node >= L) {
        const size_t s = (size_t)(node - L) * kLeaf, e = std::min(n, s + kLeaf);
        for (size_t j = s; j < e; ++j) {
          const T px = xs[j], py = ys[j], pz = zs[j];
          if (take) hp.offer(r2_ref(px, py, pz, qx, qy, qz), bound);
        }
      } else {
        if (sp + 2 > kKnnStack) {
          if (lane == 0) atomicOr(err, 1u);
          break;
        }
        if (lane == 0) {
          s_node[sp] = 2 * node + 1;
          s_mask[sp] = tm;
          s_node[sp + 1] = 2 * node;
          s_mask[sp + 1] = tm;
        }
        sp += 2;
      }
    }
    if (sp == 0) break;
    --sp;
    __builtin_amdgcn_wave_barrier();
    node = (uint32_t)__builtin_amdgcn_readfirstlane((int)s_node[sp]);
    const uint64_t m = s_mask[sp];
    live = ((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(m >> 32)) << 32) |
           (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)m);
    __builtin_amdgcn_wave_barrier();
  }
  if (valid) out[perm[i]] = (float)sqrt(hp.top());
}

}  // namespace

template <typename T>
int smoothing_sizes(slsim_ctx ctx, const T* x, const T* y, const T* z, size_t n, int nsmooth,
                    float* out_host) {
  const int K = nsmooth + 1;
  hipStream_t st = ctx->stream;
  std::vector<void*> held;
  auto get = [&](void** p, size_t bytes) -> int {
    int rc = ctx->alloc(p, std::max<size_t>(bytes, 16));
    if (!rc) held.push_back(*p);
    return rc;
  };
  struct Release {
    slsim_ctx c;
    std::vector<void*>& v;
    ~Release() {
      (void)hipStreamSynchronize(c->stream);
      for (void* p : v) c->release(p);
    }
  } rel{ctx, held};
  int rc;
  T *dx, *dy, *dz, *sx, *sy, *sz;
  uint64_t *keys, *keys_s;
  uint32_t *idx, *perm, *err;
  double* part;
  float* dout;
  if ((rc = get((void**)&dx, n * sizeof(T))) || (rc = get((void**)&dy, n * sizeof(T))) ||
      (rc = get((void**)&dz, n * sizeof(T))) || (rc = get((void**)&sx, n * sizeof(T))) ||
      (rc = get((void**)&sy, n * sizeof(T))) || (rc = get((void**)&sz, n * sizeof(T))) ||
      (rc = get((void**)&keys, n * 8)) || (rc = get((void**)&keys_s, n * 8)) ||
      (rc = get((void**)&idx, n * 4)) || (rc = get((void**)&perm, n * 4)) ||
      (rc = get((void**)&err, 4)) || (rc = get((void**)&dout, n * 4)))
    return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(dx, x, n * sizeof(T), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dy, y, n * sizeof(T), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(dz, z, n * sizeof(T), hipMemcpyHostToDevice, st));
  // bounding box
  const unsigned bb = std::min<unsigned>(1024, nblk(n));
  if ((rc = get((void**)&part, 6 * bb * sizeof(double)))) return rc;
  k_bbox3<T><<<bb, kBlk, 0, st>>>(dx, dy, dz, n, part);
  std::vector<double> hpart(6 * bb);
  SLSIM_HIP_CHECK(hipMemcpyAsync(hpart.data(), part, hpart.size() * 8, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  double lo[3], hi[3];
  for (int j = 0; j < 3; ++j) {
    lo[j] = INFINITY;
    hi[j] = -INFINITY;
    for (unsigned b = 0; b < bb; ++b) {
      lo[j] = std::min(lo[j], hpart[6 * b + j]);
      hi[j] = std::max(hi[j], hpart[6 * b + 3 + j]);
    }
  }
  double sc[3];
  for (int j = 0; j < 3; ++j) sc[j] = hi[j] > lo[j] ? 2097152.0 / (hi[j] - lo[j]) : 0.0;
  // keys, sort, gather
  k_keys3<T><<<nblk(n), kBlk, 0, st>>>(dx, dy, dz, n, lo[0], lo[1], lo[2], sc[0], sc[1], sc[2],
                                       keys, idx);
  size_t sort_bytes = 0;
  SLSIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sort_bytes, keys, keys_s, idx, perm,
                                                      (int)n, 0, 63, st));
  void* sort_tmp;
  if ((rc = get(&sort_tmp, sort_bytes))) return rc;
  SLSIM_HIP_CHECK(hipcub::DeviceRadixSort::SortPairs(sort_tmp, sort_bytes, keys, keys_s, idx, perm,
                                                      (int)n, 0, 63, st));
  k_gather3<T><<<nblk(n), kBlk, 0, st>>>(perm, dx, dy, dz, n, sx, sy, sz);
  // leaf boxes and the tree above them
  const uint32_t nleaf = (uint32_t)((n + kLeaf - 1) / kLeaf);
  uint32_t L = 1;
  while (L < nleaf) L <<= 1;
  Box* box;
  if ((rc = get((void**)&box, 2 * (size_t)L * sizeof(Box)))) return rc;
  k_leaf_boxes<T><<<nblk(L), kBlk, 0, st>>>(sx, sy, sz, n, nleaf, L, box);
  for (uint32_t first = L / 2; first >= 1; first /= 2)
    k_merge_boxes<<<nblk(first), kBlk, 0, st>>>(first, first, box);
  // queries
  SLSIM_HIP_CHECK(hipMemsetAsync(err, 0, 4, st));
  const size_t lds = 64 * (size_t)K * sizeof(double);
  const double slack = sizeof(T) == 4 ? 1.0 + 1e-6 : 1.0 + 1e-12;
  k_knn<T><<<nblk(n, 64), 64, lds, st>>>(sx, sy, sz, perm, n, K, L, box, slack, dout, err);
  SLSIM_HIP_CHECK(hipGetLastError());
  uint32_t herr = 0;
  SLSIM_HIP_CHECK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipMemcpyAsync(out_host, dout, n * 4, hipMemcpyDeviceToHost, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  if (herr) {
    set_error("smoothing: traversal stack overflow");
    return SLSIM_ERR_TOO_DEEP;
  }
  return SLSIM_OK;
}

}  // namespace slsim

extern "C" {

static int smoothing_check(slsim_ctx ctx, const void* x, const void* y, const void* z, size_t n,
                           int nsmooth, const float* out) {
  if (!ctx || !x || !y || !z || !out) {
    slsim::set_error("null argument");
    return SLSIM_ERR_ARG;
  }
  if (nsmooth < 1 || nsmooth + 1 > slsim::kMaxK) {
    slsim::set_error("smoothing: Nsmooth must be in [1, 119]");
    return SLSIM_ERR_ARG;
  }
  if (n <= (size_t)nsmooth + 1) {
    slsim::set_error("number of neighbors > total number of particles");
    return SLSIM_ERR_ARG;
  }
  if (n >= 0x7fffffffull) {
    slsim::set_error("smoothing: too many particles");
    return SLSIM_ERR_ARG;
  }
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  return SLSIM_OK;
}

int slsim_smoothing_sizes_f64(slsim_ctx ctx, const double* x, const double* y, const double* z,
                              size_t n, int nsmooth, float* size_out) {
  int rc = smoothing_check(ctx, x, y, z, n, nsmooth, size_out);
  return rc ? rc : slsim::smoothing_sizes<double>(ctx, x, y, z, n, nsmooth, size_out);
}

int slsim_smoothing_sizes_f32(slsim_ctx ctx, const float* x, const float* y, const float* z,
                              size_t n, int nsmooth, float* size_out) {
  int rc = smoothing_check(ctx, x, y, z, n, nsmooth, size_out);
  return rc ? rc : slsim::smoothing_sizes<float>(ctx, x, y, z, n, nsmooth, size_out);
}

}  // extern "C"
