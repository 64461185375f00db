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
//   3. One wavefront per query particle at a time (consecutive in key order
//      along a wavefront's run).  A bound on the K-th squared distance comes
//      from the previous query of the run (its K nearest lie within
//      sqrt(K-th) + the step, taken when the step is at most half that), else
//      from the 2K particles around the query in key order (their K-th
//      smallest).  The lanes then test up to 64 boxes of the tree at once,
//      pushing the children of the taken internal boxes and queueing the
//      taken leaves, and scan the queued leaves two at a time (32 lanes a
//      leaf), appending
//      every squared distance within the bound to a buffer in LDS.  The K-th
//      smallest of the buffer is found by bisection on the FP64 bit pattern
//      (non-negative doubles order as their bits), counted with ballots.
//      A full buffer is cut to its K smallest (the bound becomes the K-th).
//      A query whose box stack would overflow LDS is listed and done again
//      by a kernel whose stack (in global memory) holds every cell.
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
constexpr int kKnnWaves = 4;        // wavefronts per block of the query kernel
constexpr int kKnnStack = 512;      // box stack entries per wavefront
constexpr int kKnnLeafQ = 128;      // queued leaves per wavefront (<= 64 added a step, < 2 left)
constexpr int kKnnBuf = 512;        // buffered squared distances per wavefront (8 slots of 64)
constexpr int kKnnRun = 8;          // queries per wavefront
constexpr int kKnnRetryWaves = 8;   // wavefronts of the global-stack retry kernel
constexpr int kMaxK = 120;          // K = Nsmooth + 1 limit (2K window: 4 slots of 64)

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

__device__ inline double box_d2(const Box& b, double qx, double qy, double qz) {
  const double q[3] = {qx, qy, qz};
  double d2 = 0;
  for (int j = 0; j < 3; ++j) {
    const double t = fmax(fmax(b.lo[j] - q[j], q[j] - b.hi[j]), 0.0);
    d2 += t * t;
  }
  return d2;
}

__device__ inline uint64_t ballot64(bool p) { return __builtin_amdgcn_ballot_w64(p); }
__device__ inline int lanes_below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                        __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

// The K-th smallest (1-based) of up to 64 S values held one per lane per
// slot (bit patterns of non-negative doubles; absent entries ~0): the largest
// pattern t with fewer than K values below it, built from the top bit down.
template <int S>
__device__ inline uint64_t kth_bits(const uint64_t (&v)[S], int slots, int K) {
  uint64_t res = 0;
  for (int b = 62; b >= 0; --b) {
    const uint64_t t = res | (1ull << b);
    int c = 0;
#pragma unroll
    for (int s = 0; s < S; ++s)
      if (s < slots) c += __builtin_popcountll(ballot64(v[s] < t));
    if (c < K) res = t;
  }
  return res;
}

template <int S>
__device__ inline void load_slots(const double* buf, int m, int lane, uint64_t (&v)[S]) {
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int j = 64 * s + lane;
    v[s] = j < m ? (uint64_t)__double_as_longlong(buf[j]) : ~0ull;
  }
}

// One query, by a whole wavefront: the K-th smallest squared distance from
// particle iq, the tree walked with the box stack `stk` of `cap` entries
// (pk >= 0: the previous query's K-th at (px, py, pz) may give the bound).
// Returns false, having written nothing, when the stack would overflow.
template <typename T>
__device__ bool knn_query(const T* __restrict__ xs, const T* __restrict__ ys,
                          const T* __restrict__ zs, size_t n, int K, uint32_t L,
                          const Box* __restrict__ box, double slack, size_t iq, double pk,
                          double px, double py, double pz, uint32_t* stk, int cap,
                          uint32_t* s_leaf, double* s_buf, double* kth_out) {
  const int lane = (int)(threadIdx.x & 63);
  const T qx = xs[iq], qy = ys[iq], qz = zs[iq];
  const double qdx = (double)qx, qdy = (double)qy, qdz = (double)qz;
  double bound;
  const double ddx = qdx - px, ddy = qdy - py, ddz = qdz - pz;
  const double dd = ddx * ddx + ddy * ddy + ddz * ddz;
  if (pk >= 0 && 4 * dd <= pk) {
    // the previous query's K nearest lie within sqrt(pk) + |q - p| of this
    // one (taken when the step is at most half that radius)
    const double r = sqrt(pk) + sqrt(dd);
    bound = r * r * slack * slack;
  } else {
    // the K-th smallest over the 2K particles around the query in key order
    size_t w0 = iq > (size_t)K ? iq - K : 0;
    const size_t w1 = std::min(n, w0 + 2 * (size_t)K);
    w0 = w1 > 2 * (size_t)K ? w1 - 2 * (size_t)K : 0;
    uint64_t v[4];
#pragma unroll
    for (int s = 0; s < 4; ++s) {
      const size_t j = w0 + 64 * s + lane;
      const size_t jj = j < w1 ? j : w0;
      const double r2 = r2_ref(xs[jj], ys[jj], zs[jj], qx, qy, qz);
      v[s] = j < w1 ? (uint64_t)__double_as_longlong(r2) : ~0ull;
    }
    bound = __longlong_as_double((long long)kth_bits(v, (int)((w1 - w0 + 63) / 64), K));
  }
  // the tree: boxes within the bound, leaves scanned two at a time
  bool strict = false;       // after a cut the buffer holds K values <= bound
  int sp = 1, nleaf = 0, m = 0;
  if (lane == 0) stk[0] = 1;
  while (sp > 0 || nleaf > 0) {
    if (sp > 0) {
      const int nn = sp < 64 ? sp : 64;
      __builtin_amdgcn_wave_barrier();
      const uint32_t node = lane < nn ? stk[sp - nn + lane] : 0u;
      __builtin_amdgcn_wave_barrier();
      sp -= nn;
      const bool take = lane < nn && box_d2(box[node], qdx, qdy, qdz) <= bound * slack;
      const bool inner = take && node < L, leaf = take && node >= L;
      const uint64_t bi = ballot64(inner), bl = ballot64(leaf);
      const int ci = __builtin_popcountll(bi), cl = __builtin_popcountll(bl);
      if (sp + 2 * ci > cap) return false;
      if (inner) {
        const int o = sp + 2 * lanes_below(bi);
        stk[o] = 2 * node + 1;
        stk[o + 1] = 2 * node;
      }
      if (leaf) s_leaf[nleaf + lanes_below(bl)] = node - L;
      sp += 2 * ci;
      nleaf += cl;
    }
    while (nleaf >= 2 || (sp == 0 && nleaf > 0)) {
      const int take2 = nleaf >= 2 ? 2 : 1;
      __builtin_amdgcn_wave_barrier();
      const int half = lane >> 5;
      const bool has = half < take2;
      const uint32_t lf = has ? s_leaf[nleaf - 1 - half] : 0u;
      __builtin_amdgcn_wave_barrier();
      nleaf -= take2;
      const size_t j = (size_t)lf * kLeaf + (lane & 31);
      const bool valid = has && j < n;
      const size_t jj = valid ? j : iq;
      const double r2 = r2_ref(xs[jj], ys[jj], zs[jj], qx, qy, qz);
      bool acc = valid && (strict ? r2 < bound : r2 <= bound);
      uint64_t ba = ballot64(acc);
      int ca = __builtin_popcountll(ba);
      if (m + ca > kKnnBuf) {
        // cut the buffer to its K smallest: those below the K-th, then
        // copies of the K-th up to K values
        uint64_t v[kKnnBuf / 64];
        __builtin_amdgcn_wave_barrier();
        load_slots(s_buf, m, lane, v);
        __builtin_amdgcn_wave_barrier();
        const uint64_t kb = kth_bits(v, (m + 63) / 64, K);
        int base = 0;
#pragma unroll
        for (int s = 0; s < kKnnBuf / 64; ++s) {
          const uint64_t b = ballot64(v[s] < kb);
          if (v[s] < kb) s_buf[base + lanes_below(b)] = __longlong_as_double((long long)v[s]);
          base += __builtin_popcountll(b);
        }
        bound = __longlong_as_double((long long)kb);
        for (int j2 = base + lane; j2 < K; j2 += 64) s_buf[j2] = bound;
        m = K;
        strict = true;
        acc = valid && r2 < bound;
        ba = ballot64(acc);
        ca = __builtin_popcountll(ba);
      }
      if (acc) s_buf[m + lanes_below(ba)] = r2;
      m += ca;
    }
  }
  uint64_t v[kKnnBuf / 64];
  __builtin_amdgcn_wave_barrier();
  load_slots(s_buf, m, lane, v);
  __builtin_amdgcn_wave_barrier();
  *kth_out = __longlong_as_double((long long)kth_bits(v, (m + 63) / 64, K));
  return true;
}

// Runs of kKnnRun consecutive queries per wavefront, the box stack in LDS; a
// query whose stack would overflow is listed for k_knn_retry.
template <typename T>
__global__ __launch_bounds__(64 * kKnnWaves) void k_knn(
    const T* __restrict__ xs, const T* __restrict__ ys, const T* __restrict__ zs,
    const uint32_t* __restrict__ perm, size_t n, int K, uint32_t L, const Box* __restrict__ box,
    double slack, float* __restrict__ out, uint32_t* __restrict__ nretry,
    uint32_t* __restrict__ retry) {
  __shared__ uint32_t s_stack_all[kKnnWaves][kKnnStack];
  __shared__ uint32_t s_leaf_all[kKnnWaves][kKnnLeafQ];
  __shared__ double s_buf_all[kKnnWaves][kKnnBuf];
  const int w = (int)(threadIdx.x >> 6), lane = (int)(threadIdx.x & 63);
  const size_t q0 = ((size_t)blockIdx.x * kKnnWaves + w) * kKnnRun;
  double pk = -1, px = 0, py = 0, pz = 0;   // the previous query's K-th and position
  for (size_t iq = q0; iq < q0 + kKnnRun && iq < n; ++iq) {
    double kth;
    if (!knn_query(xs, ys, zs, n, K, L, box, slack, iq, pk, px, py, pz, s_stack_all[w],
                   kKnnStack, s_leaf_all[w], s_buf_all[w], &kth)) {
      if (lane == 0) retry[atomicAdd(nretry, 1u)] = (uint32_t)iq;
      pk = -1;
      continue;
    }
    if (lane == 0) out[perm[iq]] = (float)sqrt(kth);
    pk = kth;
    px = (double)xs[iq];
    py = (double)ys[iq];
    pz = (double)zs[iq];
  }
}

// The listed queries again, one wavefront each in turn over a fixed grid,
// with a box stack in global memory of `cap` entries per wavefront: every
// cell of the tree at most once, so it cannot overflow.
template <typename T>
__global__ __launch_bounds__(64) void k_knn_retry(
    const T* __restrict__ xs, const T* __restrict__ ys, const T* __restrict__ zs,
    const uint32_t* __restrict__ perm, size_t n, int K, uint32_t L, const Box* __restrict__ box,
    double slack, float* __restrict__ out, const uint32_t* __restrict__ nretry,
    const uint32_t* __restrict__ retry, uint32_t* __restrict__ gstack, int cap,
    uint32_t* __restrict__ err) {
  __shared__ uint32_t s_leaf[kKnnLeafQ];
  __shared__ double s_buf[kKnnBuf];
  uint32_t* stk = gstack + (size_t)blockIdx.x * cap;
  const uint32_t nr = *nretry;
  for (uint32_t i = blockIdx.x; i < nr; i += gridDim.x) {
    const size_t iq = retry[i];
    double kth;
    if (!knn_query(xs, ys, zs, n, K, L, box, slack, iq, -1.0, 0.0, 0.0, 0.0, stk, cap, s_leaf,
                   s_buf, &kth)) {
      if (threadIdx.x == 0) atomicOr(err, 1u);
      continue;
    }
    if (threadIdx.x == 0) out[perm[iq]] = (float)sqrt(kth);
  }
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
      (rc = get((void**)&keys, std::max<size_t>(n, 64) * 8)) ||
      (rc = get((void**)&keys_s, std::max<size_t>(n, 64) * 8)) ||
      (rc = get((void**)&idx, n * 4)) || (rc = get((void**)&perm, n * 4)) ||
      (rc = get((void**)&err, 8)) || (rc = get((void**)&dout, n * 4)))
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
  SLSIM_HIP_CHECK(hipMemsetAsync(err, 0, 8, st));   // the error flag, the retry count
  const double slack = sizeof(T) == 4 ? 1.0 + 1e-6 : 1.0 + 1e-12;
  // the retry list and the retry kernel's stacks reuse the key arrays (dead
  // once the particles are in key order): n and 2 kKnnRetryWaves L <= n + 32
  // entries of 2 max(n, 64)
  uint32_t* nretry = err + 1;
  uint32_t* retry = reinterpret_cast<uint32_t*>(keys);
  uint32_t* gstack = reinterpret_cast<uint32_t*>(keys_s);
  k_knn<T><<<nblk(n, kKnnWaves * kKnnRun), 64 * kKnnWaves, 0, st>>>(
      sx, sy, sz, perm, n, K, L, box, slack, dout, nretry, retry);
  SLSIM_HIP_CHECK(hipGetLastError());
  // the queries whose LDS stack would overflow (usually none: the kernel
  // reads the count and exits), with a stack of every cell
  k_knn_retry<T><<<kKnnRetryWaves, 64, 0, st>>>(sx, sy, sz, perm, n, K, L, box, slack, dout,
                                               nretry, retry, gstack, (int)(2 * L), err);
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
