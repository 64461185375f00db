// shoot.hip -- multi-plane ray shooting kernels.
//
// One wavefront carries a bundle of 64 spatially coherent rays through one
// lens plane: the bundle walks the plane's quadtree depth first with a
// shared LDS stack whose entries hold a 64-bit live mask.  Every lane
// applies the reference's own per-ray opening criterion
// (TreeQuadParticles::walkTree_recur, include/quadTree.h:990-1128): a lane
// for which a cell is far takes its multipole and drops out of that
// subtree, a lane that opens it takes the leaf's particles and the
// big-particle B-spline corrections.  The bundle descends while any lane
// is live, so each ray sees exactly the cells it would see alone, in the
// same depth-first order (results do not depend on bundle mates).
// Cell records and particles are wave-uniform and are read with scalar
// loads.  The plane-to-plane recursion (Lens::compute_points_parallel,
// include/lens.h:913-1103) runs in FP64 after each plane.

#include <cstdint>

#include "cosmo.h"
#include "slsim_internal.h"

namespace slsim {

namespace {

__device__ inline int lane_id() { return (int)__lane_id(); }

__device__ inline uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// Lane masks kept in scalar registers: a compare's mask without a vector
// round trip, and a mask back to a per-lane condition at no cost.
__device__ inline uint64_t lanes_lt(double a, double b) {
  return __builtin_amdgcn_fcmp(a, b, 4 /* ordered less-than */);
}
__device__ inline uint64_t lanes_ltf(float a, float b) {
  return __builtin_amdgcn_fcmpf(a, b, 4 /* ordered less-than */);
}
__device__ inline bool lane_in(uint64_t mask) { return __builtin_amdgcn_inverse_ballot_w64(mask); }

__device__ inline int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

__device__ inline uint64_t uni64(uint64_t v) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// 1/x in FP64: hardware estimate refined by two Newton steps.
__device__ inline double rcp64(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = fma(-x, r, 1.0);
  r = fma(r, e, r);
  e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

__device__ inline float rcp32(float x) { return __builtin_amdgcn_rcpf(x); }

// Wave-uniform reads (cell records, particles of a cell, job descriptors) go
// through the constant address space so that they compile to scalar loads.
// (The host pass of this single-source file has no constant address space.)
#if defined(__HIP_DEVICE_COMPILE__)
template <typename T>
using cptr = const __attribute__((address_space(4))) T*;
#else
template <typename T>
using cptr = const T*;
#endif
template <typename T>
__device__ inline cptr<T> as_const(const T* p) {
  return (cptr<T>)(uintptr_t)p;
}

constexpr double kInvPI = 1.0 / kPI;
constexpr float kInvPIf = (float)(1.0 / kPI);

// XCD-aware block order: hardware dispatches block b to XCD b % 8; give each
// XCD a contiguous range of bundles so neighbouring tiles share its L2.
__device__ inline uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t xcd = b & 7u, idx = b >> 3;
  const uint32_t per = nb >> 3, rem = nb & 7u;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

struct Counts {
  uint32_t v = 0, c = 0, p = 0, b = 0, s = 0;   // this ray's work (reference criterion)
  uint32_t wv = 0, wp = 0;                       // the wavefront's cell evaluations and
                                                 // leaf-particle iterations
};

// b_spline_profile (include/quadTree.h:147-190), FP64, phi omitted.
__device__ inline void bspline64(double d0, double d1, double r, double Mass, double size,
                                 double& ax, double& ay, double& kap, double& g0, double& g1) {
  const double q = r / size;
  double M, sigma;
  if (q > 2) {
    sigma = 0;
    M = 1;
  } else {
    const double q2 = q * q, q3 = q2 * q;
    sigma = (8 - 12 * q + 6 * q2 - q3) / 4;
    if (q > 1) {
      sigma *= 10 / size / size / 7 * kInvPI;
      M = (-1 + 20 * q2 * (1 - q + 3 * q2 / 8 - q3 / 20)) / 7;
    } else {
      sigma = 10 * (sigma - 1 + 3 * q - 3 * q2 + q3) / size / size / 7 * kInvPI;
      M = 10 * q2 * (1 - 3 * q / 4 + 3 * q3 / 10) / 7;
    }
  }
  const double alpha_r = (M - 1) * kInvPI / r;
  const double gt = alpha_r / r - sigma;
  ax -= Mass * alpha_r * d0 / r;
  ay -= Mass * alpha_r * d1 / r;
  g0 -= gt * Mass * (d0 * d0 - d1 * d1) / r / r;
  g1 -= gt * Mass * 2 * d0 * d1 / r / r;
  kap += Mass * sigma;
}

struct Acc64 {
  double ax = 0, ay = 0, k = 0, g0 = 0, g1 = 0;
};

// Traversal shared by both walks.  The stack holds opened internal cells as
// (first child | non-empty-children mask << 28, live mask of the lanes that
// opened it).  Popping an entry evaluates its four contiguous children for
// those lanes: a lane for which a child is far takes its multipole, a lane
// that opens a leaf takes its particles, and internal children opened by
// any lane are pushed (child3 first, so child0's subtree is walked first).
// Each lane therefore sees the cells of its own per-ray criterion, in an
// order fixed by its own open set.
__device__ inline void push_entry(int32_t* s_node, uint64_t* s_mask, int& sp, int lane,
                                  int32_t entry, uint64_t mask, uint32_t* err) {
  if (sp >= kStackDepth) {
    if (lane == 0) atomicOr(err, 1u);
    return;
  }
  if (lane == 0) {
    s_node[sp] = entry;
    s_mask[sp] = mask;
  }
  ++sp;
}

// ---------------------------------------------------------------- FP64 walk
// Evaluates one cell for the lanes with lv set (include/quadTree.h:990-1128):
// returns the mask of lanes that opened it when it is internal, 0 otherwise.
// The open test is r^2 < T = max(r2crit, near2) (r^2 < r2crit || r^2 <
// near2); the near test alone gates the big particles and is made only where
// they are taken.  The mass sheet's terms (inv_area) are skipped when the
// tree has none (adding zero).
template <bool COUNT>
__device__ __forceinline__ uint64_t eval64(const NodeD& nd, uint64_t lvm, double rx, double ry,
                                           double ia, cptr<PartD> parts, cptr<PartD> bigs,
                                           Acc64& a, Counts& cnt) {
  const double dx = nd.cx - rx, dy = nd.cy - ry;
  const double r2 = dx * dx + dy * dy;
  const uint64_t omask = lanes_lt(r2, nd.T) & lvm;
  const bool lv = lane_in(lvm);
  const bool open = lane_in(omask);
  const bool sheet = ia != 0.0;
  if (COUNT) cnt.wv++;
  if (COUNT && lv) cnt.v++;
  if (omask != lvm && lv && !open) {
    // far cell: monopole + quadrupole (include/quadTree.h:1089-1123)
    if (COUNT) cnt.c++;
    const double inv = rcp64(r2);
    const double mi = nd.mass * inv * kInvPI;
    double tmp = sheet ? nd.mass * ia - mi : -mi;
    a.ax += tmp * dx;
    a.ay += tmp * dy;
    tmp = -2.0 * mi * inv;
    a.g0 += 0.5 * (dx * dx - dy * dy) * tmp;
    a.g1 += dx * dy * tmp;
    const double i2p = inv * inv * kInvPI;
    a.ax -= (nd.q0 * dx + nd.q2 * dy) * i2p;
    a.ay -= (nd.q1 * dy + nd.q2 * dx) * i2p;
    tmp = 4 * (nd.q0 * dx * dx + nd.q1 * dy * dy + 2 * nd.q2 * dx * dy) * inv * i2p;
    a.ax += tmp * dx;
    a.ay += tmp * dy;
    if (sheet) a.k -= ia * nd.mass;
  }
  const uint64_t om = omask;
  if (om == 0) return 0;
  if (nd.child < 0) {
    // leaf: every particle as a point mass (include/quadTree.h:1014-1048)
    for (int k = 0; k < nd.count; ++k) {
      const PartD p = parts[nd.start + k];
      if (open) {
        const double d0 = p.x - rx, d1 = p.y - ry;
        double rr = d0 * d0 + d1 * d1;
        if (rr < 1e-20) rr = 1e-20;
        const double m = p.m;
        const double inv = rcp64(rr);
        const double pref = m * inv * kInvPI;
        const double s = sheet ? ia * m - pref : -pref;
        a.ax += s * d0;
        a.ay += s * d1;
        const double tmp = -2.0 * pref * inv;
        a.g0 += 0.5 * (d0 * d0 - d1 * d1) * tmp;
        a.g1 += d0 * d1 * tmp;
        if (sheet) a.k -= m * ia;
      }
    }
    if (COUNT && open) cnt.p += nd.count;
    if (COUNT) cnt.wp += nd.count;
    const uint64_t nm = lanes_lt(r2, nd.near2) & lvm;
    if (nm) {
      const bool nearc = lane_in(nm);
      // the leaf's big particles: h < box (include/quadTree.h:379-396, 1051-1077)
      for (int k = 0; k < nd.count; ++k) {
        const PartD p = parts[nd.start + k];
        if (nearc && !signbit(p.h)) {
          if (COUNT) cnt.b++;
          const double d0 = p.x - rx, d1 = p.y - ry;
          double rr = d0 * d0 + d1 * d1;
          if (rr < 1e-20) rr = 1e-20;
          const double size = p.h;
          if (rr < 4 * size * size) {
            if (COUNT) cnt.s++;
            bspline64(d0, d1, sqrt(rr), (double)p.m, size, a.ax, a.ay, a.k, a.g0, a.g1);
          }
        }
      }
    }
    return 0;
  }
  if (nd.count > 0) {
    const uint64_t nm = lanes_lt(r2, nd.near2) & lvm;
    if (nm) {
      const bool nearc = lane_in(nm);
      for (int k = 0; k < nd.count; ++k) {
        const PartD p = bigs[nd.start + k];
        if (nearc) {
          if (COUNT) cnt.b++;
          const double d0 = p.x - rx, d1 = p.y - ry;
          double rr = d0 * d0 + d1 * d1;
          if (rr < 1e-20) rr = 1e-20;
          const double size = p.h;
          if (rr < 4 * size * size) {
            if (COUNT) cnt.s++;
            bspline64(d0, d1, sqrt(rr), (double)p.m, size, a.ax, a.ay, a.k, a.g0, a.g1);
          }
        }
      }
    }
  }
  return om;
}

// Walk of one tree, FP64 (plain method).  Must be called by all 64 lanes.
// A popped entry's children are taken last to first, each opened internal
// child pushed at once, so child 0's subtree is walked next.
template <bool COUNT>
__device__ void walk64(const TreeDesc& t, double rx, double ry, bool active, Acc64& a,
                       Counts& cnt, int32_t* s_node, uint64_t* s_mask, uint32_t* err) {
  const int lane = lane_id();
  if (ballot(active) == 0) return;
  const cptr<NodeD> nodes = as_const(t.nodesD);
  const cptr<PartD> parts = as_const(t.partsD);
  const cptr<PartD> bigs = as_const(t.bigD);
  const double ia = t.inv_area;
  int sp = 0;
  {
    const NodeD root = nodes[0];
    const uint64_t om = eval64<COUNT>(root, ballot(active), rx, ry, ia, parts, bigs, a, cnt);
    if (om) push_entry(s_node, s_mask, sp, lane, root.child | (root.cmask << 28), om, err);
  }
  while (sp > 0) {
    --sp;
    __builtin_amdgcn_wave_barrier();
    const int entry = uni(s_node[sp]);
    const uint64_t live = uni64(s_mask[sp]);
    __builtin_amdgcn_wave_barrier();
    const int base = entry & 0x0fffffff;
    const int cm = (entry >> 28) & 0xf;
#pragma unroll
    for (int k = 3; k >= 0; --k) {
      if (!((cm >> k) & 1)) continue;
      const NodeD nd = nodes[base + k];
      const uint64_t pm = eval64<COUNT>(nd, live, rx, ry, ia, parts, bigs, a, cnt);
      if (pm) push_entry(s_node, s_mask, sp, lane, nd.child | (nd.cmask << 28), pm, err);
    }
  }
}

// ------------------------------------------------------ halo-of-halos walk
// LensHaloNFW's profile functions (AnalyticNSIE/nfw_lens.cpp:82-125), which
// InterpolateFromTable evaluates directly in the first table interval
// (MultiPlane/lens_halos.cpp:566-581).
__device__ inline double nfw_g(double x) {
#pragma clang fp contract(off)
  if (x < 1e-5) x = 1e-5;
  const double ans = log(x / 2);
  if (x == 1.0) return ans + 1.0;
  if (x > 1.0) return ans + 2 * atan(sqrt((x - 1) / (x + 1))) / sqrt(x * x - 1);
  return ans + 2 * atanh(sqrt((1 - x) / (x + 1))) / sqrt(1 - x * x);
}
__device__ inline double nfw_f(double x) {
#pragma clang fp contract(off)
  if (x == 0) x = 1e-5;
  if (x == 1.0) return 1.0 / 3.0;
  if (x > 1.0) return (1 - 2 * atan(sqrt((x - 1) / (x + 1))) / sqrt(x * x - 1)) / (x * x - 1);
  return (1 - 2 * atanh(sqrt((1 - x) / (x + 1))) / sqrt(1 - x * x)) / (x * x - 1);
}
__device__ inline double nfw_g2(double x) {
#pragma clang fp contract(off)
  if (x == 0) x = 1e-5;
  if (x == 1) return 10 / 3. + 4 * log(0.5);
  const double y = x * x - 1;
  const double ans = 4 * log(x / 2) / x / x - 2 / y;
  if (x < 1.0) return ans + 4 * (2 / x / x + 1 / y) * atanh(sqrt((1 - x) / (x + 1))) / sqrt(-y);
  return ans + 4 * (2 / x / x + 1 / y) * atan(sqrt((x - 1) / (x + 1))) / sqrt(y);
}

// The closed forms below the first table interval (rays within 0.01 rscale
// of a centre), kept out of line: their log/atan/atanh would otherwise set
// the register count of every kernel that walks halo-of-halos trees.
__device__ __attribute__((noinline)) void nfw_closed(double y, double* g, double* f, double* g2) {
  *g = nfw_g(y);
  *f = nfw_f(y);
  *g2 = nfw_g2(y);
}

// InterpolateFromTable for the three tables at once: linear between the
// entries j and j+1, j = (int)(y / maxrm * NTABLE) (kept in range; halos are
// accepted only with Rsize / rscale below the table's end).
__device__ inline void nfw_interp(cptr<NfwEntry> T, double y, double& g, double& f, double& g2) {
#pragma clang fp contract(off)
  int j = (int)(y / kNfwMaxrm * kNfwTable);
  if (j == 0) {
    nfw_closed(y, &g, &f, &g2);
    return;
  }
  j = j < 0 ? 0 : (j > kNfwTable - 2 ? kNfwTable - 2 : j);
  const NfwEntry a = T[j];
  const double d = y - a.x;
  g = a.sg * d + a.g;
  f = a.sf * d + a.f;
  g2 = a.sg2 * d + a.g2;
}

// LensHalo::force_halo of a symmetric NFW without point subtraction
// (MultiPlane/lens_halos.cpp:1528-1672): force_point beyond Rmax (a point
// mass), force_halo_sym inside with alpha_h, kappa_h, gamma_h of
// include/lens_halos.h:842-853.  x0, x1: position relative to the halo.
// Divisions by r^2, pi, rscale and gmax are multiplications by reciprocals
// (a last-bit difference from the reference's quotients).
// LensHaloPowerLaw's force_halo_sym inside Rmax (rscale 1, xmax = Rsize,
// include/lens_halos.h:979-997): alpha_h clamps x below 1e-6 xmax, kappa_h
// and gamma_h only x = 0 (unreachable after the r^2 clamp).  Out of line, the
// five terms returned by value: NFW-only trees keep their registers.
struct Terms5 {
  double ax, ay, k, g0, g1;
};
__device__ __noinline__ Terms5 powerlaw_inside(double xmax, double beta, double r2, double inv,
                                               double prefac, double x0, double x1) {
#pragma clang fp contract(off)
  const double x = sqrt(r2);
  const double xa = x < xmax * 1.0e-6 ? 1e-6 * xmax : x;
  const double e = -beta + 2;
  const double pa = pow(xa / xmax, e);
  const double pk = xa == x ? pa : pow(x / xmax, e);
  Terms5 t;
  double tmp = -1.0 * pa * prefac;
  t.ax = tmp * x0;
  t.ay = tmp * x1;
  t.k = 0.5 * e * pk * prefac;
  tmp = -beta * pk * prefac * inv;
  t.g0 = 0.5 * (x0 * x0 - x1 * x1) * tmp;
  t.g1 = x0 * x1 * tmp;
  return t;
}

template <bool PL>
__device__ inline void nfw_force(const HaloNFW& h, cptr<NfwEntry> T, double x0, double x1,
                                 Acc64& a) {
#pragma clang fp contract(off)
  double r2 = x0 * x0 + x1 * x1;
  const double R2 = (double)(h.rmax * h.rmax);   // float Rmax squared in float
  if (r2 < R2) {
    if (r2 < 1e-20) r2 = 1e-20;
    const double inv = rcp64(r2);
    const double prefac = h.mass_pi * inv;
    if (r2 < R2) {
      if (PL && h.pad != 0.0f) {   // a power-law halo: gmax holds beta
        const Terms5 t = powerlaw_inside((double)h.rmax, h.gmax, r2, inv, prefac, x0, x1);
        a.ax += t.ax;
        a.ay += t.ay;
        a.k += t.k;
        a.g0 += t.g0;
        a.g1 += t.g1;
        return;
      }
      const double x = sqrt(r2) * h.inv_rscale;
      double g, f, g2;
      nfw_interp(T, x, g, f, g2);
      double tmp = -g * h.inv_gmax * prefac;
      a.ax += tmp * x0;
      a.ay += tmp * x1;
      a.k += 0.5 * x * x * f * h.inv_gmax * prefac;
      tmp = -0.5 * x * x * g2 * h.inv_gmax * prefac * inv;
      a.g0 += 0.5 * (x0 * x0 - x1 * x1) * tmp;
      a.g1 += x0 * x1 * tmp;
      return;
    }
    a.ax += -1.0 * prefac * x0;
    a.ay += -1.0 * prefac * x1;
    const double tmp = -2.0 * prefac * inv;
    a.g0 += 0.5 * (x0 * x0 - x1 * x1) * tmp;
    a.g1 += x0 * x1 * tmp;
    return;
  }
  const double inv = rcp64(r2);
  const double prefac = h.mass_pi * inv;
  a.ax += -1.0 * prefac * x0;
  a.ay += -1.0 * prefac * x1;
  const double tmp = -2.0 * prefac * inv;
  a.g0 += 0.5 * (x0 * x0 - x1 * x1) * tmp;
  a.g1 += x0 * x1 * tmp;
}

// A halo of the tree with the uniform background the walk removes with it
// (include/quadTreeHalos.h:668-693): the halo's force at halo - ray plus
// inv_area m (halo - ray) and -inv_area m.
template <bool PL>
__device__ inline void halo_term(const HaloNFW& h, cptr<NfwEntry> T, double rx, double ry,
                                 double ia, Acc64& a) {
#pragma clang fp contract(off)
  const double x0 = h.x - rx, x1 = h.y - ry;
  nfw_force<PL>(h, T, x0, x1, a);
  const double m = h.mass;
  const double prefac = m * ia;
  a.ax += prefac * x0;
  a.ay += prefac * x1;
  a.k -= m * ia;
}

// exp(-r2 s2), out of line: only screened field planes call it, and inlined
// it would set the halo kernels' register count
__device__ __noinline__ double screening(double r2, double s2) { return exp(-r2 * s2); }

// One cell of TreeQuadHalos::walkTree_iter (include/quadTreeHalos.h:
// 656-735) for the lanes in lvm: its big halos, then, for an internal cell,
// the monopole + quadrupole where the cell is far.  Returns the lanes that
// descend into its children.
template <bool PL>
__device__ __forceinline__ uint64_t evalh64(const HNode& nd, uint64_t lvm, double rx, double ry,
                                            double ia, const TreeDesc& t, Acc64& a) {
#pragma clang fp contract(off)
  const bool lv = lane_in(lvm);
  const cptr<HaloNFW> halos = as_const(t.halos);
  const cptr<int32_t> big = as_const(t.hbig);
  const cptr<NfwEntry> T = as_const(t.nfw);
  for (int k = 0; k < nd.nbig; ++k) {
    const HaloNFW h = halos[big[nd.big_off + k]];
    if (lv) halo_term<PL>(h, T, rx, ry, ia, a);
  }
  if (nd.child < 0) return 0;
  const double dx = nd.cx - rx, dy = nd.cy - ry;
  const double r2 = dx * dx + dy * dy;
  const uint64_t farm = lanes_lt(nd.r2crit, r2) & lvm;
  if (lane_in(farm)) {
    const double inv = rcp64(r2);
    // LensPlaneTree's screening exp(-r^2 s^2) of the far terms (:679-733);
    // exactly 1 for other trees
    const double scr = t.hs2 != 0.0 ? screening(r2, t.hs2) : 1.0;
    const double i2p = inv * inv * kInvPI * scr;
    double tmp = nd.mass * (ia - inv * kInvPI * scr);
    a.ax += tmp * dx;
    a.ay += tmp * dy;
    tmp = -2.0 * nd.mass * i2p;
    a.g0 += 0.5 * (dx * dx - dy * dy) * tmp;
    a.g1 += dx * dy * tmp;
    a.ax -= (nd.q0 * dx + nd.q2 * dy) * i2p;
    a.ay -= (nd.q1 * dy + nd.q2 * dx) * i2p;
    tmp = 4 * (nd.q0 * dx * dx + nd.q1 * dy * dy + 2 * nd.q2 * dx * dy) * inv * i2p;
    a.ax += tmp * dx;
    a.ay += tmp * dy;
    a.k -= nd.mass * ia;
  }
  return lvm & ~farm;
}

// Walk of one halo-of-halos tree for the whole wavefront.  A stack entry is
// a block of four contiguous cells (first cell | cells-with-halos mask << 28)
// with the lanes that reach it; the root is the block {cell 0}.  The cells
// of a block are evaluated last to first and each one a lane descends into
// is pushed at once, so the first cell's subtree is walked next (one copy of
// the cell evaluation keeps the kernel's registers down).
template <bool PL>
__device__ void walk_halos64(const TreeDesc& t, double rx, double ry, bool active, Acc64& a,
                             int32_t* s_node, uint64_t* s_mask, uint32_t* err) {
  const int lane = lane_id();
  const uint64_t am = ballot(active);
  if (am == 0) return;
  const cptr<HNode> nodes = as_const(t.hnodes);
  const double ia = t.inv_area;
  int sp = 0;
  push_entry(s_node, s_mask, sp, lane, 1 << 28, am, err);
  while (sp > 0) {
    --sp;
    __builtin_amdgcn_wave_barrier();
    const int entry = uni(s_node[sp]);
    const uint64_t live = uni64(s_mask[sp]);
    __builtin_amdgcn_wave_barrier();
    const int base = entry & 0x0fffffff;
    const int cm = (entry >> 28) & 0xf;
#pragma unroll 1
    for (int k = 3; k >= 0; --k) {
      if (!((cm >> k) & 1)) continue;
      const HNode nd = nodes[base + k];
      const uint64_t om = evalh64<PL>(nd, live, rx, ry, ia, t, a);
      if (om) push_entry(s_node, s_mask, sp, lane, nd.child | (nd.cmask << 28), om, err);
    }
  }
}

// Every halo of the plane with its background term, no tree (the error
// column of halo-of-halos planes).
template <bool PL>
__device__ void halos_direct64(const TreeDesc& t, double rx, double ry, bool active, Acc64& a) {
  if (!active) return;
  const cptr<HaloNFW> halos = as_const(t.halos);
  const cptr<NfwEntry> T = as_const(t.nfw);
  for (int i = 0; i < t.nhalos; ++i) halo_term<PL>(halos[i], T, rx, ry, t.inv_area, a);
}

// TreeQuadHalos::force2D (include/quadTreeHalos.h:557-592): the walk (or the
// sum over every halo), with the periodic buffer nine of them at the ray
// shifted by the halos' extents, x shift outer.
template <bool DIRECT, bool PL>
__device__ void halo_tree_force64(const TreeDesc& t, double rx, double ry, bool active, Acc64& a,
                                  int32_t* s_node, uint64_t* s_mask, uint32_t* err) {
#pragma clang fp contract(off)
  const int nshift = t.hperiodic ? 9 : 1;
#pragma unroll 1
  for (int q = 0; q < nshift; ++q) {
    double sx = rx, sy = ry;
    if (nshift > 1) {
      sx = rx + (q / 3 - 1) * t.hxl;
      sy = ry + (q % 3 - 1) * t.hyl;
    }
    if (DIRECT) halos_direct64<PL>(t, sx, sy, active, a);
    else walk_halos64<PL>(t, sx, sy, active, a, s_node, s_mask, err);
  }
}

// ---------------------------------------------------------------- FP32 walk
// The FP32 walk sums pi times the reference's terms, with 1/pi, the signs
// and the factor 2 of gamma1 taken out of the per-term arithmetic:
//   AX, AY = pi sum alpha,  G0 = -pi sum gamma0,  G1 = -(pi/2) sum gamma1,
// restored once per walk in FP64 (finish32).  Kappa has no point-mass term
// besides the sheet: its B-spline terms (pi sigma m, rare) go straight to an
// FP64 sum.  Cell quadrupoles are stored as 4 q.  The mass sheet (inv_area: quadTree.h:861-888, 1035-1044) adds
// inv_area m (c - x) to alpha and -inv_area m to kappa for every cell
// monopole and point mass; every particle enters exactly once as one or the
// other, so the sheet's total is inv_area M (c - x) and -inv_area M with M, c
// the tree's mass and centre of mass, added in FP64 at the end.

struct CSum {   // Kahan-compensated FP32 sum
  float s = 0.f, c = 0.f;   // c: the running compensation (lost low part, negated)
  __device__ inline void add(float x) {
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ inline double value() const { return (double)s - (double)c; }
};

// Near-field interactions deferred to FP64 (per lane, in LDS): the tree
// index in the high word, the particle index in the low word with its top
// bit set for an internal big-particle list entry.
constexpr int kNearQueue = 7;   // 7: the block's LDS stays under 20 KiB (8 blocks per CU)
constexpr uint32_t kBigListBit = 0x80000000u;

struct Acc32 {
  CSum ax, ay, g0, g1;
  // plain partial sums of one popped stack entry (four cells), then one
  // compensated add per entry
  float pax = 0.f, pay = 0.f, pg0 = 0.f, pg1 = 0.f;
  uint64_t* q = nullptr;   // the wave's queues: lane l's at q[l], q[64 + l], ...
  int nq = 0;
  uint32_t tree = 0;       // index of the tree being walked (uniform)
  // FP64 sums of the queued terms and of the B-spline kappa terms (the only
  // kappa terms besides the mass sheet), in the reference's scale and signs
  double fx = 0, fy = 0, fk = 0, fg0 = 0, fg1 = 0;
  __device__ inline void flush() {
    ax.add(pax); ay.add(pay); g0.add(pg0); g1.add(pg1);
    pax = pay = pg0 = pg1 = 0.f;
  }
  __device__ inline void kappa(float sigma_pi_m) { fk += kInvPI * (double)sigma_pi_m; }
  __device__ inline bool push(uint32_t idx) {
    if (nq >= kNearQueue) return false;
    q[64 * nq + lane_id()] = ((uint64_t)tree << 32) | idx;
    ++nq;
    return true;
  }
};

// M(q) / r^2 and pi sigma(q) of the reference's cubic B-spline
// (include/quadTree.h:147-190, M in the reference's form), written so that
// nothing cancels at small q.  inv = 1 / r^2.
__device__ inline void bspline_fused32(float r2, float inv, float h, float& Mr2, float& sigpi) {
  const float r = sqrtf(r2);
  const float ih = rcp32(h);
  const float q = r * ih;
  const float ih2 = ih * ih;
  if (q > 2.f) {
    Mr2 = inv;
    sigpi = 0.f;
  } else if (q > 1.f) {
    const float q2 = q * q, q3 = q2 * q;
    const float M = (-1.f + 20.f * q2 * (1.f - q + 0.375f * q2 - 0.05f * q3)) * (1.f / 7.f);
    Mr2 = M * inv;
    const float tq = 2.f - q;
    sigpi = (10.f / 28.f) * tq * tq * tq * ih2;
  } else {
    const float q2 = q * q, q3 = q2 * q;
    Mr2 = (10.f / 7.f) * (1.f - 0.75f * q + 0.3f * q3) * ih2;
    sigpi = (10.f / 7.f) * (1.f - 1.5f * q2 + 0.75f * q3) * ih2;
  }
}

// The reference's B-spline correction (bspline64's formulas) with the
// divisions taken as reciprocals: FP64-accurate, and light on registers
// after the FP32 walk.
__device__ inline void bspline64r(double d0, double d1, double rr, double Mass, double size,
                                  Acc32& a) {
  const double is = rcp64(size), is2 = is * is;
  const double r = sqrt(rr);
  const double ir = rcp64(r);
  const double q = r * is;
  double M, sigma;
  if (q > 2) {
    sigma = 0;
    M = 1;
  } else {
    const double q2 = q * q, q3 = q2 * q;
    sigma = (8 - 12 * q + 6 * q2 - q3) * 0.25;
    if (q > 1) {
      sigma *= (10.0 / 7.0) * kInvPI * is2;
      M = (-1 + 20 * q2 * (1 - q + 0.375 * q2 - 0.05 * q3)) * (1.0 / 7.0);
    } else {
      sigma = (10.0 / 7.0) * kInvPI * is2 * (sigma - 1 + 3 * q - 3 * q2 + q3);
      M = (10.0 / 7.0) * q2 * (1 - 0.75 * q + 0.3 * q3);
    }
  }
  const double alpha_r = (M - 1) * kInvPI * ir;
  const double gt = alpha_r * ir - sigma;
  const double ma = Mass * alpha_r * ir;
  a.fx -= ma * d0;
  a.fy -= ma * d1;
  const double mg = gt * Mass * ir * ir;
  a.fg0 -= mg * (d0 * d0 - d1 * d1);
  a.fg1 -= mg * 2 * d0 * d1;
  a.fk += Mass * sigma;
}

// The queued near-field terms, after the walk (its registers are free then),
// in FP64 with the reference's arithmetic: a bare point mass within 2h of the
// ray (its cell is not near, or does not list it as big, so the reference
// applies no B-spline) and the B-spline corrections of internal-cell lists
// (which cancel a point mass taken at the particle's leaf).  These terms grow
// as 1/r^2 near a particle; FP64 keeps them and their cancellation exact.
// The mass sheet of the point masses is in finish32.
__device__ void near_terms64(const TreeDesc* __restrict__ trees, double rx, double ry,
                             Acc32& a) {
  for (int i = 0; i < a.nq; ++i) {
    const uint64_t e = a.q[64 * i + lane_id()];
    const uint32_t idx = (uint32_t)e;
    const TreeDesc* t = trees + (uint32_t)(e >> 32);
    const bool big = (idx & kBigListBit) != 0;
    const PartD p = big ? t->bigD[idx & ~kBigListBit] : t->partsD[idx];
    const double d0 = p.x - rx, d1 = p.y - ry;
    double rr = d0 * d0 + d1 * d1;
    if (rr < 1e-20) rr = 1e-20;
    const double m = p.m;
    if (!big) {
      const double inv = rcp64(rr);
      const double pref = m * inv * kInvPI;
      const double tmp = -2.0 * pref * inv;
      a.fx -= pref * d0;
      a.fy -= pref * d1;
      a.fg0 += 0.5 * (d0 * d0 - d1 * d1) * tmp;
      a.fg1 += d0 * d1 * tmp;
    } else {
      const double size = p.h;
      if (rr < 4 * size * size) bspline64r(d0, d1, rr, m, size, a);
    }
  }
  a.nq = 0;
}

// The guard of the precision form: the FP32 ray can be up to dx (physical,
// this ray's bound) from where the FP64 ray is, so an opening test whose r^2
// lies within 2 sqrt(T) dx of its threshold T (dx << sqrt(T)) may go the
// other way in FP64 and change the deflection by the cell's flip error.
// With D = (2 dx)^2 such tests are (r^2 - T)^2 < T D: the lanes that pass add
// the cell's flip error to their risk.  A near test (r^2 < 5.83 box^2,
// gating big particles' B-spline corrections) that is marginal where a big
// particle lies within 2h of the ray adds an infinite risk (near_marginal).
__device__ __forceinline__ void guard_open(cptr<NodeF> nd, double r2, double T, uint64_t lk,
                                           double D, float& risk) {
  const double d = r2 - T;
  const uint64_t mk = lanes_lt(d * d, T * D) & lk;
  if (mk) risk += lane_in(mk) ? fabsf(nd->flip) : 0.f;
}

// |r^2 - near2| within the band sqrt(near2) 2 dx, from the FP32 r^2 (its
// rounding allowed for).  sD: 2 dx (1 + 2^-20), the margin covering the
// hardware square root's error.
__device__ inline bool near_marginal(float r2f, float nef, float sD) {
  return fabsf(r2f - nef) < fmaf(__builtin_amdgcn_sqrtf(nef), sD, nef * (1.0f / 4194304.0f));
}

// The FP32 sums are updated with the same instructions on every path: a
// lane that takes no part in a term gets zero coefficients (selected, never
// formed from an infinite 1/r^2), and the rare near-field branches only
// change coefficients; divergent branches around the sums themselves would
// cost register copies.

// Far cell: monopole + quadrupole (stored as 4 q) for the lanes with far
// set; the others get 1/r^2 = 0, which zeroes every term.
__device__ __forceinline__ void far32(const NodeF& nd, bool far, float ex, float ey, float r2f,
                                      Acc32& a) {
  const float inv = far ? rcp32(r2f) : 0.f;
  const float s = nd.mass * inv;
  const float u = inv * inv;
  const float qx0 = fmaf(nd.q0, ex, nd.q2 * ey);
  const float qx1 = fmaf(nd.q1, ey, nd.q2 * ex);
  const float qq = fmaf(ex, qx0, ey * qx1);
  const float kk = fmaf(qq * inv, u, -s);
  const float u4 = 0.25f * u;
  const float w = s * inv;
  a.pax = fmaf(kk, ex, fmaf(-qx0, u4, a.pax));
  a.pay = fmaf(kk, ey, fmaf(-qx1, u4, a.pay));
  a.pg0 = fmaf(w, fmaf(ex, ex, -ey * ey), a.pg0);
  a.pg1 = fmaf(w * ex, ey, a.pg1);
}

// Particles of an opened leaf.  particle - ray = (particle - leaf centre,
// FP32 from the build) + (centre - ray as the FP32 pair ex + exl of the
// walk's FP64 offset), so that no FP64 arithmetic runs per particle; the
// offsets are good to about 2^-24 of the leaf size.  A particle that may lie
// within 2h (the near field, whose B-spline terms are sensitive to the
// offset) takes its offset again from the FP64 positions, as the FP64 walk
// forms it, before the near-field rules apply.

// One leaf particle's offset and point-mass coefficients for a lane.
struct LeafTerm {
  float d0, d1, d1s, s, w;
};

// The near-field rules for a lane whose particle may lie within 2h: the
// offset again from the FP64 positions, then the B-spline correction, the
// FP64 queue or the bare point mass.
template <bool GUARD>
__device__ __forceinline__ void leaf_near32(int idx, float m, bool nearc, double rx, double ry,
                                            cptr<PartD> partsd, Acc32& a, float r2f, float nef,
                                            float sD, float& risk, LeafTerm& t) {
  const PartD q = partsd[idx];
  t.d0 = (float)(q.x - rx);
  t.d1 = (float)(q.y - ry);
  t.d1s = t.d1 * t.d1;
  const float rr = fmaxf(fmaf(t.d0, t.d0, t.d1s), 1e-20f);
  const float inv = rcp32(rr);
  t.s = m * inv;
  t.w = t.s * inv;
  const float hs = fabsf(q.h);
  if (rr < 4.f * hs * hs) {
    // a big particle within 2h: its correction hangs on the near test
    const bool listed = !signbit(q.h);
    if (GUARD && listed && near_marginal(r2f, nef, sD)) risk = __builtin_inff();
    if (nearc && listed) {
      // smoothed: the B-spline correction folded into the point mass
      float Mr2, sigpi;
      bspline_fused32(rr, inv, hs, Mr2, sigpi);
      t.s = m * Mr2;
      const float sm = m * sigpi;
      t.w = (t.s - sm) * inv;
      a.kappa(sm);
    } else if (a.push((uint32_t)idx)) {
      t.s = 0.f;    // deferred to FP64
      t.w = 0.f;
    }
  }
}

// The point-mass term of a leaf particle; rr: the squared offset.
__device__ __forceinline__ LeafTerm leaf_point32(const PartF& p, bool open, float ex, float ey,
                                                 float exl, float eyl, float& rr) {
  LeafTerm t;
  t.d0 = (p.lx + ex) + exl;
  t.d1 = (p.ly + ey) + eyl;
  t.d1s = t.d1 * t.d1;
  // (no clamp: |hh4| >= FLT_MIN takes r^2 = 0 to the near-field branch)
  rr = fmaf(t.d0, t.d0, t.d1s);
  const float inv = open ? rcp32(rr) : 0.f;       // 0: no term for this lane
  t.s = p.m * inv;
  t.w = t.s * inv;
  return t;
}

__device__ __forceinline__ void leaf_add32(const LeafTerm& t, Acc32& a) {
  a.pax = fmaf(-t.s, t.d0, a.pax);
  a.pay = fmaf(-t.s, t.d1, a.pay);
  a.pg0 = fmaf(t.w, fmaf(t.d0, t.d0, -t.d1s), a.pg0);
  a.pg1 = fmaf(t.w * t.d0, t.d1, a.pg1);
}

// Particles of an opened leaf (om: the lanes that opened it), kLeafStep at a
// time (one scalar load of their records): the lanes whose particle may lie
// within 2h (4h^2 carries a margin) form masks in scalar registers, and only
// a non-empty mask branches (uniformly, rarely) to the near-field rules, so
// the common path carries no exec-mask changes and a quarter of the loop's
// scalar work (10-plane 2^20-particle field: 22.5 -> 22.3 ms per plane).
constexpr int kLeafStep = 4;
template <bool GUARD>
__device__ __forceinline__ void leaf32(int start, int count, uint64_t om, bool nearc, float ex,
                                       float ey, float exl, float eyl, double rx, double ry,
                                       cptr<PartF> parts, cptr<PartD> partsd, Acc32& a,
                                       float r2f, float nef, float sD, float& risk) {
  const bool open = lane_in(om);
  int k = 0;
  for (; k + kLeafStep <= count; k += kLeafStep) {
    LeafTerm t[kLeafStep];
    uint64_t n[kLeafStep], any = 0;
#pragma unroll
    for (int j = 0; j < kLeafStep; ++j) {
      const PartF p = parts[start + k + j];
      float rr;
      t[j] = leaf_point32(p, open, ex, ey, exl, eyl, rr);
      n[j] = lanes_ltf(rr, fabsf(p.hh4)) & om;
      any |= n[j];
    }
    if (any) {
#pragma unroll
      for (int j = 0; j < kLeafStep; ++j)
        if (lane_in(n[j]))
          leaf_near32<GUARD>(start + k + j, parts[start + k + j].m, nearc, rx, ry, partsd, a, r2f,
                             nef, sD, risk, t[j]);
    }
#pragma unroll
    for (int j = 0; j < kLeafStep; ++j) leaf_add32(t[j], a);
  }
  for (; k < count; ++k) {
    const PartF p = parts[start + k];
    float rr;
    LeafTerm t = leaf_point32(p, open, ex, ey, exl, eyl, rr);
    const uint64_t n = lanes_ltf(rr, fabsf(p.hh4)) & om;
    if (n && lane_in(n))
      leaf_near32<GUARD>(start + k, p.m, nearc, rx, ry, partsd, a, r2f, nef, sD, risk, t);
    leaf_add32(t, a);
  }
}

// An internal cell's own big particles: B-spline corrections (near field,
// queued for FP64; FP32 here only when the lane's queue is full).
__device__ __forceinline__ void bigs32(int start, int count, bool nearc, double rx, double ry,
                                       cptr<PartD> bigs, Acc32& a) {
  for (int k = 0; k < count; ++k) {
    const PartD p = bigs[start + k];
    const float d0 = (float)(p.x - rx), d1 = (float)(p.y - ry);
    const float rr = fmaxf(fmaf(d0, d0, d1 * d1), 1e-20f);
    float am = 0.f, w = 0.f;
    if (nearc && rr < 4.0001f * p.h * p.h && !a.push((uint32_t)(start + k) | kBigListBit) &&
        rr < 4.f * p.h * p.h) {
      const float inv = rcp32(rr);
      float Mr2, sigpi;
      bspline_fused32(rr, inv, p.h, Mr2, sigpi);
      am = p.m * (Mr2 - inv);            // m (M - 1) / r^2
      const float sk = p.m * sigpi;
      w = (am - sk) * inv;
      a.kappa(sk);
    }
    a.pax = fmaf(-am, d0, a.pax);
    a.pay = fmaf(-am, d1, a.pay);
    a.pg0 = fmaf(w, fmaf(d0, d0, -d1 * d1), a.pg0);
    a.pg1 = fmaf(w * d0, d1, a.pg1);
  }
}

// Walk of the trees of one plane, FP32 arithmetic with compensated
// accumulation.  One loop serves every tree: when the stack is empty the next
// tree starts with its root as the only child of a virtual entry.  The
// children of a popped entry are taken last to first, each in full (far
// term, leaf particles or big list, push when opened), so that child 0's
// subtree is walked next; empty children cost a scalar branch.
template <bool GUARD>
__device__ void walk32(const TreeDesc* __restrict__ trees, int ntrees, double rx, double ry,
                       bool active, Acc32& a, int32_t* s_node, uint64_t* s_mask,
                       uint64_t* s_near, uint32_t* err, double D, float sD, float& risk) {
  const int lane = lane_id();
  a.q = s_near;
  a.nq = 0;
  const uint64_t amask = ballot(active);
  if (amask == 0) return;
  int h = 0, sp = 0;
  cptr<NodeF> nodes = as_const(as_const(trees)->nodesF);
  cptr<PartF> parts = as_const(as_const(trees)->partsF);
  cptr<PartD> partsd = as_const(as_const(trees)->partsD);
  cptr<PartD> bigs = as_const(as_const(trees)->bigD);
  cptr<float2> sheetd = as_const(as_const(trees)->sheetd);
  a.tree = 0;
  int entry = 1 << 28;            // base 0, child mask {0}: the root
  uint64_t live = amask;
  while (true) {
    const bool lv = lane_in(live);
    const int base = entry & 0x0fffffff;
    const int cm = (entry >> 28) & 0xf;
    int lists = 0;                // opened internal children with a big-particle list
#pragma unroll
    for (int k = 3; k >= 0; --k) {
      if (!((cm >> k) & 1)) continue;
      const cptr<NodeF> nd = nodes + base + k;
      const double dx = nd->cx - rx, dy = nd->cy - ry;
      const double r2 = dx * dx + dy * dy;
      const double T = nd->T;
      const uint64_t om = lanes_lt(r2, T) & live;   // r^2 < r2crit or r^2 < near2
      if (GUARD) guard_open(nd, r2, T, live, D, risk);
      const float ex = (float)dx, ey = (float)dy;
      const uint64_t fm = live & ~om;
      if (fm) {
        far32(*nd, lane_in(fm), ex, ey, (float)r2, a);
        if (sheetd) {
          // masses of both signs: the far cell's share of the mass sheet
          const float2 dl = sheetd[base + k];
          const float sel = lane_in(fm) ? 1.f : 0.f;
          a.pax = fmaf(sel, dl.x, a.pax);
          a.pay = fmaf(sel, dl.y, a.pay);
        }
      }
      if (om) {
        const int32_t child = nd->child;
        const int32_t cc = nd->count_cmask;
        const int cnt = cc & 0x0fffffff;
        if (child < 0) {
          const uint64_t nm = lanes_lt(r2, nd->near2) & om;
          const float exl = (float)(dx - (double)ex), eyl = (float)(dy - (double)ey);
          leaf32<GUARD>(nd->start, cnt, om, lane_in(nm), ex, ey, exl, eyl, rx, ry,
                        parts, partsd, a, (float)r2, (float)nd->near2, sD, risk);
        } else {
          if (cnt > 0) {
            // a big list: its corrections hang on the near test
            if (GUARD && lane_in(live) && near_marginal((float)r2, (float)nd->near2, sD))
              risk = __builtin_inff();
            lists |= 1 << k;
          }
          push_entry(s_node, s_mask, sp, lane, child | (cc & (int32_t)0xf0000000), om, err);
        }
      }
    }
    // big-particle lists of the opened internal children, for the lanes
    // within their near radius (r^2 < near2 <= T: such lanes opened the cell)
#pragma nounroll
    for (int k = 0; k < 4; ++k) {
      if (!((lists >> k) & 1)) continue;
      const cptr<NodeF> nd = nodes + base + k;
      const double dx = nd->cx - rx, dy = nd->cy - ry;
      const uint64_t nm = lanes_lt(dx * dx + dy * dy, nd->near2) & live;
      if (nm) bigs32(nd->start, nd->count_cmask & 0x0fffffff, lane_in(nm), rx, ry, bigs, a);
    }
    // one compensated add per entry, by the lanes the entry holds: a lane's
    // sums do not depend on which cells its bundle mates open
    if (lv) a.flush();
    if (sp > 0) {
      --sp;
      __builtin_amdgcn_wave_barrier();
      entry = uni(s_node[sp]);
      live = uni64(s_mask[sp]);
      __builtin_amdgcn_wave_barrier();
    } else {
      if (++h >= ntrees) break;
      const cptr<TreeDesc> td = as_const(trees) + h;
      nodes = as_const(td->nodesF);
      parts = as_const(td->partsF);
      partsd = as_const(td->partsD);
      bigs = as_const(td->bigD);
      sheetd = as_const(td->sheetd);
      a.tree = (uint32_t)h;
      entry = 1 << 28;
      live = amask;
    }
  }
  near_terms64(trees, rx, ry, a);
}

// The FP32 walk's sums in the reference's scale and signs, plus the FP64
// near-field terms: out = alpha0, alpha1, kappa, gamma0, gamma1.
__device__ inline void finish32(const Acc32& a, double* out) {
  out[0] = kInvPI * a.ax.value() + a.fx;
  out[1] = kInvPI * a.ay.value() + a.fy;
  out[2] = a.fk;
  out[3] = -kInvPI * a.g0.value() + a.fg0;
  out[4] = -2.0 * kInvPI * a.g1.value() + a.fg1;
}

// A tree's mass sheet for the FP32 walk, added to its outputs: inv_area
// (S - M x) with S the sum of m x over the particles, which is M c when the
// masses share a sign (the walk adds the far cells' partition terms of a
// tree with masses of both signs).
__device__ inline void sheet32(const TreeDesc& t, double rx, double ry, double* out) {
  if (t.inv_area != 0.0) {
    const double sheet = t.inv_area * t.mass;
    if (t.sheetd) {
      out[0] += t.inv_area * t.dipx - sheet * rx;
      out[1] += t.inv_area * t.dipy - sheet * ry;
    } else {
      out[0] += sheet * (t.comx - rx);
      out[1] += sheet * (t.comy - ry);
    }
    out[2] -= sheet;
  }
}

// FP64 direct sum of the smoothed-particle kernel over every particle of a
// tree (no cells): the error-column oracle.  Particles are wave-uniform reads.
__device__ void direct64(const TreeDesc& t, double rx, double ry, bool active, Acc64& a) {
  if (ballot(active) == 0) return;
  const cptr<PartD> parts = as_const(t.partsD);
  const double ia = t.inv_area;
  const int np = t.nparts;
  for (int k = 0; k < np; ++k) {
    const PartD p = parts[k];
    if (!active) continue;
    const double d0 = p.x - rx, d1 = p.y - ry;
    double rr = d0 * d0 + d1 * d1;
    if (rr < 1e-20) rr = 1e-20;
    const double m = p.m;
    const double inv = rcp64(rr);
    const double pref = m * inv * kInvPI;
    a.ax += (ia * m - pref) * d0;
    a.ay += (ia * m - pref) * d1;
    const double tmp = -2.0 * pref * inv;
    a.g0 += 0.5 * (d0 * d0 - d1 * d1) * tmp;
    a.g1 += d0 * d1 * tmp;
    a.k -= m * ia;
    const double size = fabs((double)p.h);
    if (rr < 4 * size * size) bspline64(d0, d1, sqrt(rr), m, size, a.ax, a.ay, a.k, a.g0, a.g1);
  }
}

// The precision form's bounds, carried with the recursion: the FP32 walk's
// deflection may be off by eps max(|alpha|, ascale) plus the flip errors of
// its marginal tests (risk, walk units); a position error eth moves the
// plane's deflection by at most |G| eth (G the step's Jacobian term); both
// enter SumPrevAlphas and the next position as the recursion combines them.
// The risk is summed over the planes in radians (|charge| times the walk's).
__device__ inline void guard_step(const StepJob& jb, int slot, double alpha0, double alpha1,
                                  double kappa, double gamma0, double gamma1, float risk) {
  const RayState st = state_view(jb.sbase, jb.nslots);
  const double amag = sqrt(alpha0 * alpha0 + alpha1 * alpha1);
  const double ea = jb.eps * fmax(amag, jb.ascale) + (double)risk;
  st.risk[slot] += (float)(fabs(jb.charge) * (double)risk);
  const double gn = fabs(jb.Gfac) * (fabs(kappa) + sqrt(gamma0 * gamma0 + gamma1 * gamma1));
  const double eth = st.eth[slot];
  const double eS = (double)st.eS[slot] + gn * eth + fabs(jb.charge) * ea;
  double aa, bb;
  if (jb.j < st.jmax[slot] - 1) {
    aa = jb.aa;
    bb = jb.bb;
  } else {
    aa = st.aal[slot];
    bb = st.bbl[slot];
  }
  st.eS[slot] = (float)eS;
  st.eth[slot] = (float)(fabs(bb) * eth + fabs(aa) * eS);
}

// FP64 plane recursion step (include/lens.h:1003-1066), without FMA
// contraction so that it performs the reference's roundings.
__device__ inline void recursion_step(const StepJob& jb, int slot, double alpha0, double alpha1,
                                      double kappa, double gamma0, double gamma1) {
#pragma clang fp contract(off)
  const RayState st = state_view(jb.sbase, jb.nslots);
  double G[4] = {kappa + gamma0, gamma1, gamma1, kappa - gamma0};
  for (int k = 0; k < 4; ++k) G[k] *= jb.Gfac;
  double sa0 = st.sa0[slot], sa1 = st.sa1[slot];
  sa0 -= jb.charge * alpha0;
  sa1 -= jb.charge * alpha1;
  double aa, bb;
  if (jb.j < st.jmax[slot] - 1) {
    aa = jb.aa;
    bb = jb.bb;
  } else {
    aa = st.aal[slot];
    bb = st.bbl[slot];
  }
  const double th0 = bb * st.th0[slot] + aa * sa0;
  const double th1 = bb * st.th1[slot] + aa * sa1;
  const double A[4] = {st.A0[slot], st.A1[slot], st.A2[slot], st.A3[slot]};
  double S[4] = {st.G0[slot], st.G1[slot], st.G2[slot], st.G3[slot]};
  const double GA[4] = {G[0] * A[0] + G[1] * A[2], G[0] * A[1] + G[1] * A[3],
                        G[2] * A[0] + G[3] * A[2], G[2] * A[1] + G[3] * A[3]};
  for (int k = 0; k < 4; ++k) S[k] -= GA[k];
  double An[4];
  for (int k = 0; k < 4; ++k) {
    const double t1 = A[k] * bb;
    const double t2 = S[k] * aa;
    An[k] = t1 + t2;
  }
  st.sa0[slot] = sa0; st.sa1[slot] = sa1;
  st.th0[slot] = th0; st.th1[slot] = th1;
  st.A0[slot] = An[0]; st.A1[slot] = An[1]; st.A2[slot] = An[2]; st.A3[slot] = An[3];
  st.G0[slot] = S[0]; st.G1[slot] = S[1]; st.G2[slot] = S[2]; st.G3[slot] = S[3];
}

__device__ inline double phys_coord(double th, double Dphys) {
#pragma clang fp contract(off)
  return th * Dphys;
}

}  // namespace

// One plane step of one 64-ray bundle (one work item); called by a whole
// wavefront.
// HK: 0 no halo-of-halos trees, 1 NFW members only, 2 power-law members too
template <int MODE, bool COUNT, int HK>
__device__ __forceinline__ void process_item(const StepJob* __restrict__ jobs, WorkItem it,
                                             int32_t* s_node, uint64_t* s_mask, uint64_t* s_near,
                                             uint32_t* __restrict__ err) {
  const int lane = lane_id();
  const StepJob jb = as_const(jobs)[it.job];
  const int slot = (int)it.bundle * kWave + lane;
  const bool valid = slot < jb.nslots;
  const RayState st = state_view(jb.sbase, jb.nslots);
  const int jm = valid ? st.jmax[slot] : 0;
  const bool active = jb.j < jm;
  if (ballot(active) == 0) return;
  double th0 = 0, th1 = 0;
  if (active) {
    th0 = st.th0[slot];
    th1 = st.th1[slot];
  }
  const double rx = phys_coord(th0, jb.Dphys), ry = phys_coord(th1, jb.Dphys);
  double alpha0 = 0, alpha1 = 0, kappa = 0, gamma0 = 0, gamma1 = 0;
  float plane_risk = 0.f;
  Counts cnt;
  const int ntrees = uni(jb.ntrees);
  if (MODE == SLSIM_FP32C) {
    // one set of FP32 sums over all halos of the plane, so that no FP64
    // totals are live during the walks
    Acc32 a;
    // (2 dx)^2, dx this ray's position bound, physical units
    const double dx2 = active ? 2.0 * (double)st.eth[slot] * jb.Dphys : 0.0;
    const double D = dx2 * dx2;
    float risk = 0.f;
    // (a plane may hold halo-of-halos trees only; walk32 starts at trees[0])
    if (HK == 0 || ntrees > 0)
      walk32<true>(jb.trees, ntrees, rx, ry, active, a, s_node, s_mask, s_near, err, D,
                   (float)dx2 * (1.f + 1.f / 1048576.f), risk);
    plane_risk = risk;
    double f[5];
    finish32(a, f);
    for (int h = 0; h < ntrees; ++h) sheet32(as_const(jb.trees)[h], rx, ry, f);
    alpha0 = f[0]; alpha1 = f[1]; kappa = f[2]; gamma0 = f[3]; gamma1 = f[4];
  }
  // halo-of-halos trees: FP64 in every mode (LensHaloHalos::force_halo and
  // the plane each flip alpha, so the plane's alpha is the tree's raw alpha);
  // compiled only into the kernels launched for planes that have them
  const int nhtrees = HK > 0 ? uni(jb.nhtrees) : 0;
  for (int h = 0; h < nhtrees; ++h) {
    const TreeDesc t = as_const(jb.trees)[ntrees + h];
    Acc64 a;
    halo_tree_force64<MODE == SLSIM_DIRECT64, HK == 2>(t, rx, ry, active, a, s_node, s_mask,
                                                      err);
    alpha0 += a.ax; alpha1 += a.ay; kappa += a.k; gamma0 += a.g0; gamma1 += a.g1;
  }
  for (int h = 0; h < (MODE == SLSIM_FP32C ? 0 : ntrees); ++h) {
    const TreeDesc t = as_const(jb.trees)[h];
    // LensPlaneSingular::force sums the halos; force_halo and the plane each
    // flip alpha, so the plane's alpha is the tree's raw alpha
    // (MultiPlane/planes.cpp:151-152, include/particle_halo.h:432-433).
    if (MODE == SLSIM_DIRECT64) {
      Acc64 a;
      direct64(t, rx, ry, active, a);
      alpha0 += a.ax; alpha1 += a.ay; kappa += a.k; gamma0 += a.g0; gamma1 += a.g1;
    } else {
      Acc64 a;
      walk64<COUNT>(t, rx, ry, active, a, cnt, s_node, s_mask, err);
      alpha0 += a.ax; alpha1 += a.ay; kappa += a.k; gamma0 += a.g0; gamma1 += a.g1;
    }
  }
  if (COUNT && jb.counters && lane == 0) {
    uint32_t* c = jb.counters + kCountStride * (size_t)slot;
    c[5] += cnt.wv; c[6] += cnt.wp;
  }
  if (!active) return;
  if (COUNT && jb.counters) {
    uint32_t* c = jb.counters + kCountStride * (size_t)slot;
    c[0] += cnt.v; c[1] += cnt.c; c[2] += cnt.p; c[3] += cnt.b; c[4] += cnt.s;
  }
  if (MODE == SLSIM_FP32C) guard_step(jb, slot, alpha0, alpha1, kappa, gamma0, gamma1, plane_risk);
  recursion_step(jb, slot, alpha0, alpha1, kappa, gamma0, gamma1);
}

// One plane step for every bundle of the listed jobs.  Grid: one wavefront
// per work item, 4 per block (items == null: job 0, bundle = item index).
template <int MODE, bool COUNT, int HK>
__device__ __forceinline__ void plane_step_grid(const StepJob* __restrict__ jobs,
                                                const WorkItem* __restrict__ items,
                                                uint32_t nitems, uint32_t* __restrict__ err) {
  __shared__ int32_t s_node[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_mask[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_near[kWavesPerBlock][kNearQueue * kWave];
  const int w = uni((int)(threadIdx.x >> 6));
  const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
  const uint32_t wi = blk * kWavesPerBlock + (uint32_t)w;
  if (wi >= nitems) return;
  WorkItem it{0u, wi};
  if (items) {
    it.job = as_const(items)[wi].job;
    it.bundle = as_const(items)[wi].bundle;
  }
  process_item<MODE, COUNT, HK>(jobs, it, s_node[w], s_mask[w], s_near[w], err);
}

// Persistent variant: a fixed grid whose wavefronts take work items from a
// queue counter until the list is exhausted (every wave reaches the exit).
template <int MODE, int HK>
__device__ __forceinline__ void plane_step_queue(const StepJob* __restrict__ jobs,
                                                 const WorkItem* __restrict__ items,
                                                 uint32_t nitems, uint32_t* __restrict__ counter,
                                                 uint32_t* __restrict__ err) {
  __shared__ int32_t s_node[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_mask[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_near[kWavesPerBlock][kNearQueue * kWave];
  const int w = uni((int)(threadIdx.x >> 6));
  const int lane = lane_id();
  for (;;) {
    uint32_t idx = 0;
    if (lane == 0) idx = atomicAdd(counter, 1u);
    idx = (uint32_t)uni((int)idx);
    if (idx >= nitems) break;
    WorkItem it;
    it.job = as_const(items)[idx].job;
    it.bundle = as_const(items)[idx].bundle;
    process_item<MODE, false, HK>(jobs, it, s_node[w], s_mask[w], s_near[w], err);
  }
}

// Fused steps: each wavefront takes a bundle through all of its plane steps
// (the state it writes at one step is its own input at the next), so that
// bundles of different requests and planes overlap without a launch
// boundary between steps.  counter == null: one wavefront per item.
template <int MODE, int HK>
__device__ __forceinline__ void plane_steps_fused(const StepJob* __restrict__ jobs,
                                                  const FusedItem* __restrict__ items,
                                                  uint32_t nitems, uint32_t* __restrict__ counter,
                                                  uint32_t* __restrict__ err) {
  __shared__ int32_t s_node[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_mask[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_near[kWavesPerBlock][kNearQueue * kWave];
  const int w = uni((int)(threadIdx.x >> 6));
  const int lane = lane_id();
  uint32_t idx = xcd_block(blockIdx.x, gridDim.x) * kWavesPerBlock + (uint32_t)w;
  for (;;) {
    if (counter) {
      uint32_t k = 0;
      if (lane == 0) k = atomicAdd(counter, 1u);
      idx = (uint32_t)uni((int)k);
    }
    if (idx >= nitems) break;
    const uint32_t job0 = as_const(items)[idx].job0;
    const uint32_t bundle = as_const(items)[idx].bundle;
    const uint32_t nj = as_const(items)[idx].njobs;
    for (uint32_t j = 0; j < nj; ++j)
      process_item<MODE, false, HK>(jobs, WorkItem{job0 + j, bundle}, s_node[w], s_mask[w],
                                    s_near[w], err);
    if (!counter) break;
  }
}

template <int MODE, int HK>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_plane_steps_fused(
    const StepJob* __restrict__ jobs, const FusedItem* __restrict__ items, uint32_t nitems,
    uint32_t* __restrict__ counter, uint32_t* __restrict__ err) {
  plane_steps_fused<MODE, HK>(jobs, items, nitems, counter, err);
}

// the FP32C walk at 8 waves per SIMD, as k_plane_step32
__global__ __launch_bounds__(kWave * kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(8)))
void k_plane_steps_fused32(const StepJob* __restrict__ jobs, const FusedItem* __restrict__ items,
                           uint32_t nitems, uint32_t* __restrict__ counter,
                           uint32_t* __restrict__ err) {
  plane_steps_fused<SLSIM_FP32C, 0>(jobs, items, nitems, counter, err);
}

template <int MODE, bool COUNT, int HK>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_plane_step(
    const StepJob* __restrict__ jobs, const WorkItem* __restrict__ items, uint32_t nitems,
    uint32_t* __restrict__ err) {
  plane_step_grid<MODE, COUNT, HK>(jobs, items, nitems, err);
}

template <int MODE, int HK>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_plane_step_queue(
    const StepJob* __restrict__ jobs, const WorkItem* __restrict__ items, uint32_t nitems,
    uint32_t* __restrict__ counter, uint32_t* __restrict__ err) {
  plane_step_queue<MODE, HK>(jobs, items, nitems, counter, err);
}

// The FP32C walk at 8 waves per SIMD (64 VGPRs): measured faster than 7
// waves without the few spills the cap costs.  Planes with halo-of-halos
// trees use k_plane_step<SLSIM_FP32C, false, 1 or 2> (no cap: the FP64 halo
// walk needs more registers).
__global__ __launch_bounds__(kWave * kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(8)))
void k_plane_step32(const StepJob* __restrict__ jobs, const WorkItem* __restrict__ items,
                    uint32_t nitems, uint32_t* __restrict__ err) {
  plane_step_grid<SLSIM_FP32C, false, false>(jobs, items, nitems, err);
}

__global__ __launch_bounds__(kWave * kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(8)))
void k_plane_step_queue32(const StepJob* __restrict__ jobs, const WorkItem* __restrict__ items,
                          uint32_t nitems, uint32_t* __restrict__ counter,
                          uint32_t* __restrict__ err) {
  plane_step_queue<SLSIM_FP32C, false>(jobs, items, nitems, counter, err);
}

// The FP64 walk of particle trees at 8 waves per SIMD: the 64-VGPR cap spills
// a few FP64 sums of the per-tree loop to scratch (once per tree), and the
// extra waves hide more of the walk's latency (10-plane 2^20-particle field,
// 2048^2 rays: 38.7 -> 37.3 ms per plane).  Halo-of-halos planes, and the
// work-queue and fused schedules, keep the uncapped kernels.
__global__ __launch_bounds__(kWave * kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(8)))
void k_plane_step64(const StepJob* __restrict__ jobs, const WorkItem* __restrict__ items,
                    uint32_t nitems, uint32_t* __restrict__ err) {
  plane_step_grid<SLSIM_FP64, false, false>(jobs, items, nitems, err);
}

// Planes with NFW-only halo-of-halos trees at 8 waves per SIMD as well: the
// cap spills some of the halo walk's registers to scratch (95 -> 64 VGPRs)
// and still gains (10,000-halo plane, 2048^2 FP64 grid: 76.7 -> 71.6 ms; the
// screened 1,000-halo field plane: 53.0 -> 48.2 ms).  Trees with power-law
// members (HK 2, about 120 VGPRs), the FP32C mode and the work-queue and fused
// schedules keep the uncapped kernels.
template <int MODE>
__global__ __launch_bounds__(kWave * kWavesPerBlock) __attribute__((amdgpu_waves_per_eu(8)))
void k_plane_step_nfw(const StepJob* __restrict__ jobs, const WorkItem* __restrict__ items,
                      uint32_t nitems, uint32_t* __restrict__ err) {
  plane_step_grid<MODE, false, 1>(jobs, items, nitems, err);
}

// ------------------------------------------------------- lone-ray FP64 walk
// One plane step of one ray per wavefront, for a few scattered rays (the
// precision form's guard), where a bundle walk would leave 63 lanes idle and
// wait on one lane's stack: the lanes take the ray's cells 64 at a time from
// an LDS stack (the last pushed first), a lane that opens an internal cell
// pushes its children, and the particles of the opened leaves and of the big
// lists the ray is near are then taken 64 at a time.  The cells, the opening
// test and the terms are the FP64 walk's (include/quadTree.h:990-1128), so
// the result is the bundle walk's up to the order of the FP64 sums; the
// lanes' sums are added in a fixed order.
namespace {

// a pop of up to 64 cells pushes at most 256: 64 + 192 per level of depth
constexpr int kLoneStack = 64 + 192 * 34;

__device__ inline int wave_prefix(int v, int& total) {
  const int lane = lane_id();
  int x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const int y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}

__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ void lone_walk64(const TreeDesc& t, double rx, double ry, Acc64& a, int32_t* s_cell,
                            int32_t* s_pst, int32_t* s_pcnt, int32_t* s_pkind, int32_t* s_pre,
                            uint32_t* __restrict__ err) {
  const int lane = lane_id();
  const NodeD* __restrict__ nodes = t.nodesD;
  const PartD* __restrict__ parts = t.partsD;
  const PartD* __restrict__ bigs = t.bigD;
  const double ia = t.inv_area;
  if (lane == 0) s_cell[0] = 0;
  int sp = 1;
  __syncthreads();
  while (sp > 0) {
    const int c = sp < kWave ? sp : kWave;
    sp -= c;
    const int cell = lane < c ? s_cell[sp + lane] : -1;
    __syncthreads();
    int child = 0, cmask = 0, pst = 0, pcnt = 0, pkind = -1;
    if (cell >= 0) {
      const NodeD nd = nodes[cell];
      const double dx = nd.cx - rx, dy = nd.cy - ry;
      const double r2 = dx * dx + dy * dy;
      const bool nearc = r2 < nd.near2;
      if (!(r2 < nd.r2crit || nearc)) {
        // far cell: monopole + quadrupole (include/quadTree.h:1089-1123)
        const double inv = rcp64(r2);
        const double mi = nd.mass * inv * kInvPI;
        double tmp = nd.mass * ia - mi;
        a.ax += tmp * dx;
        a.ay += tmp * dy;
        tmp = -2.0 * mi * inv;
        a.g0 += 0.5 * (dx * dx - dy * dy) * tmp;
        a.g1 += dx * dy * tmp;
        const double i2p = inv * inv * kInvPI;
        a.ax -= (nd.q0 * dx + nd.q2 * dy) * i2p;
        a.ay -= (nd.q1 * dy + nd.q2 * dx) * i2p;
        tmp = 4 * (nd.q0 * dx * dx + nd.q1 * dy * dy + 2 * nd.q2 * dx * dy) * inv * i2p;
        a.ax += tmp * dx;
        a.ay += tmp * dy;
        a.k -= ia * nd.mass;
      } else if (nd.child < 0) {
        pst = nd.start;
        pcnt = nd.count;
        pkind = nearc ? 1 : 0;
      } else {
        child = nd.child;
        cmask = nd.cmask;
        if (nd.count > 0 && nearc) {
          pst = nd.start;
          pcnt = nd.count;
          pkind = 2;
        }
      }
    }
    // children onto the stack, in lane order
    int nch_total = 0;
    int at = sp + wave_prefix(__builtin_popcount(cmask), nch_total);
    if (sp + nch_total > kLoneStack) {
      if (lane == 0) atomicOr(err, 1u);
      return;
    }
    for (int k = 0; k < 4; ++k)
      if ((cmask >> k) & 1) s_cell[at++] = child + k;
    sp += nch_total;
    // particle lists: the opened leaves and the near big lists of this pop
    int nref = 0, np = 0;
    const bool has = pkind >= 0 && pcnt > 0;
    const int ri = wave_prefix(has ? 1 : 0, nref);
    const int pre = wave_prefix(has ? pcnt : 0, np);
    if (has) {
      s_pst[ri] = pst;
      s_pcnt[ri] = pcnt;
      s_pkind[ri] = pkind;
      s_pre[ri] = pre;
    }
    __syncthreads();
    for (int q = lane; q < np; q += kWave) {
      int lo = 0, hi = nref - 1;   // the last list whose prefix is <= q
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (s_pre[mid] <= q) lo = mid;
        else hi = mid - 1;
      }
      const int idx = s_pst[lo] + (q - s_pre[lo]);
      const int kind = s_pkind[lo];
      if (kind == 2) {
        // an internal cell's big particle: its B-spline correction
        const PartD p = bigs[idx];
        const double d0 = p.x - rx, d1 = p.y - ry;
        double rr = d0 * d0 + d1 * d1;
        if (rr < 1e-20) rr = 1e-20;
        const double size = p.h;
        if (rr < 4 * size * size)
          bspline64(d0, d1, sqrt(rr), (double)p.m, size, a.ax, a.ay, a.k, a.g0, a.g1);
      } else {
        // a leaf particle as a point mass (include/quadTree.h:1014-1048),
        // and its B-spline correction when it is big and the leaf near
        const PartD p = parts[idx];
        const double d0 = p.x - rx, d1 = p.y - ry;
        double rr = d0 * d0 + d1 * d1;
        if (rr < 1e-20) rr = 1e-20;
        const double m = p.m;
        const double inv = rcp64(rr);
        const double pref = m * inv * kInvPI;
        a.ax += (ia * m - pref) * d0;
        a.ay += (ia * m - pref) * d1;
        const double tmp = -2.0 * pref * inv;
        a.g0 += 0.5 * (d0 * d0 - d1 * d1) * tmp;
        a.g1 += d0 * d1 * tmp;
        a.k -= m * ia;
        if (kind == 1 && !signbit(p.h)) {
          const double size = p.h;
          if (rr < 4 * size * size)
            bspline64(d0, d1, sqrt(rr), (double)p.m, size, a.ax, a.ay, a.k, a.g0, a.g1);
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

// jobs[0]'s plane step for the rays at slots 0 .. nrays-1 (dense state), one
// wavefront (block) per ray; planes of particle trees only.
__global__ __launch_bounds__(kWave) void k_plane_step_lone(const StepJob* __restrict__ jobs,
                                                          uint32_t nrays,
                                                          uint32_t* __restrict__ err) {
  __shared__ int32_t s_cell[kLoneStack];
  __shared__ int32_t s_pst[kWave], s_pcnt[kWave], s_pkind[kWave], s_pre[kWave];
  const uint32_t r = xcd_block(blockIdx.x, gridDim.x);
  if (r >= nrays) return;
  const StepJob jb = as_const(jobs)[0];
  const int slot = (int)r;
  const RayState st = state_view(jb.sbase, jb.nslots);
  if (jb.j >= st.jmax[slot]) return;
  const double rx = phys_coord(st.th0[slot], jb.Dphys), ry = phys_coord(st.th1[slot], jb.Dphys);
  double alpha0 = 0, alpha1 = 0, kappa = 0, gamma0 = 0, gamma1 = 0;
  const int ntrees = uni(jb.ntrees);
  for (int h = 0; h < ntrees; ++h) {
    const TreeDesc t = as_const(jb.trees)[h];
    Acc64 a;
    lone_walk64(t, rx, ry, a, s_cell, s_pst, s_pcnt, s_pkind, s_pre, err);
    alpha0 += wave_sum(a.ax);
    alpha1 += wave_sum(a.ay);
    kappa += wave_sum(a.k);
    gamma0 += wave_sum(a.g0);
    gamma1 += wave_sum(a.g1);
  }
  if (lane_id() == 0) recursion_step(jb, slot, alpha0, alpha1, kappa, gamma0, gamma1);
}

__global__ void k_zero_u32(uint32_t* __restrict__ p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

// GridMap ray layout (TreeCode_link/gridmap.cpp:293-321 via
// Utilities::PositionFromIndex, TreeCode_link/utilities.cpp:149-158) in
// tiles of 2^tw x 2^(6-tw) rays: bundle b covers tile (b % tiles_x, b / tiles_x).
__device__ inline bool grid_slot(int64_t slot, int64_t N, int64_t tiles_x, int tw, int64_t* out) {
  const int64_t b = slot >> 6;
  const int lane = (int)(slot & 63);
  const int64_t ix = (b % tiles_x) * (1 << tw) + (lane & ((1 << tw) - 1));
  const int64_t iy = (b / tiles_x) * (1 << (6 - tw)) + (lane >> tw);
  if (ix >= N || iy >= N) return false;
  *out = ix + N * iy;
  return true;
}

__device__ inline void init_slot(const RayState& st, int64_t slot, bool valid, double x0,
                                 double x1, int32_t jmax, double aal, double bbl) {
  st.th0[slot] = x0; st.th1[slot] = x1;
  st.sa0[slot] = x0; st.sa1[slot] = x1;
  st.A0[slot] = 1; st.A1[slot] = 0; st.A2[slot] = 0; st.A3[slot] = 1;
  st.G0[slot] = 1; st.G1[slot] = 0; st.G2[slot] = 0; st.G3[slot] = 1;
  st.jmax[slot] = valid ? jmax : 0;
  st.aal[slot] = aal;
  st.bbl[slot] = bbl;
  st.eth[slot] = 0.f;
  st.eS[slot] = 0.f;
  st.risk[slot] = 0.f;
}


__device__ inline void finalize_slot(const RayState& st, int64_t slot, int64_t i,
                                     double* __restrict__ y, double* __restrict__ A) {
  y[2 * i] = st.th0[slot];
  y[2 * i + 1] = st.th1[slot];
  A[4 * i] = st.A0[slot];
  A[4 * i + 1] = st.A1[slot];
  A[4 * i + 2] = st.A2[slot];
  A[4 * i + 3] = st.A3[slot];
}

__global__ void k_init_grid(RayState st, int64_t nslots, int64_t N, int64_t tiles_x, int tw,
                            double c0, double c1, double range, int32_t jmax, double aal,
                            double bbl) {
  const int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  int64_t i = 0;
  const bool valid = grid_slot(slot, N, tiles_x, tw, &i);
  double x0, x1;
  grid_position(i, N, c0, c1, range, &x0, &x1);
  init_slot(st, slot, valid, x0, x1, jmax, aal, bbl);
}

__global__ void k_finalize_grid(RayState st, int64_t nslots, int64_t N, int64_t tiles_x, int tw,
                                double* __restrict__ y, double* __restrict__ A) {
  const int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  int64_t i;
  if (!grid_slot(slot, N, tiles_x, tw, &i)) return;
  finalize_slot(st, slot, i, y, A);
}

__global__ void k_counts_grid(const uint32_t* __restrict__ c, int64_t nslots, int64_t N,
                              int64_t tiles_x, int tw, uint64_t* __restrict__ out) {
  const int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  int64_t i;
  if (!grid_slot(slot, N, tiles_x, tw, &i)) return;
  for (int k = 0; k < 5; ++k) out[5 * i + k] = c[kCountStride * slot + k];
}

// Many GridMap requests at once: thread per slot of the concatenated slot
// range, the request found by binary search over the slot offsets.
__device__ inline int find_req(const GridReq* __restrict__ reqs, int nreq, int64_t g) {
  int lo = 0, hi = nreq - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (reqs[mid].slot_off <= g) lo = mid; else hi = mid - 1;
  }
  return lo;
}

__global__ void k_init_batch(const GridReq* __restrict__ reqs, int nreq, int64_t total) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const GridReq& r = reqs[find_req(reqs, nreq, g)];
  const int64_t slot = g - r.slot_off;
  int64_t i = 0;
  const bool valid = grid_slot(slot, r.N, r.tiles_x, r.tw, &i);
  double x0, x1;
  grid_position(i, r.N, r.c0, r.c1, r.range, &x0, &x1);
  init_slot(r.st, slot, valid, x0, x1, r.jmax, r.aal, r.bbl);
}

__global__ void k_finalize_batch(const GridReq* __restrict__ reqs, int nreq, int64_t total) {
  const int64_t g = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (g >= total) return;
  const GridReq& r = reqs[find_req(reqs, nreq, g)];
  const int64_t slot = g - r.slot_off;
  int64_t i;
  if (!grid_slot(slot, r.N, r.tiles_x, r.tw, &i)) return;
  finalize_slot(r.st, slot, i, r.y, r.A);
}

// Array rays: slot s < n holds ray perm[s] (Morton order).  Per-ray source
// redshifts resolve jmax, Dls and Ds as Lens::FindSourcePlane does
// (MultiPlane/lens.cpp:2659-2681).
// `lanes` rays per 64-slot bundle (sparse bundles for a few scattered rays):
// sorted ray k in slot (k / lanes) 64 + k % lanes.
__device__ inline int64_t array_ray_index(int64_t slot, int lanes) {
  const int64_t l = slot & 63;
  return l < lanes ? (slot >> 6) * lanes + l : -1;
}

__global__ void k_init_array(RayState st, int64_t nslots, int64_t n, const uint32_t* __restrict__ perm,
                             const double* __restrict__ x, const double* __restrict__ zs,
                             int32_t jmax_u, double aal_u, double bbl_u,
                             const double* __restrict__ zpl, const double* __restrict__ Dl,
                             int32_t nplanes, FlatLCDM cosmo, int lanes) {
  const int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const int64_t k = array_ray_index(slot, lanes);
  const bool valid = k >= 0 && k < n;
  const int64_t i = valid ? (perm ? (int64_t)perm[k] : k) : 0;   // no perm: ray k in place
  const double x0 = valid ? x[2 * i] : 0.0, x1 = valid ? x[2 * i + 1] : 0.0;
  st.th0[slot] = x0; st.th1[slot] = x1;
  st.sa0[slot] = x0; st.sa1[slot] = x1;
  st.A0[slot] = 1; st.A1[slot] = 0; st.A2[slot] = 0; st.A3[slot] = 1;
  st.G0[slot] = 1; st.G1[slot] = 0; st.G2[slot] = 0; st.G3[slot] = 1;
  int32_t jm = jmax_u;
  double aal = aal_u, bbl = bbl_u;
  if (valid && zs) {
    const double z = zs[i];
    const double Ds = cosmo.coorDist(0.0, z);
    double Dls;
    if (z <= zpl[0]) {
      jm = 0;
      Dls = Ds;
    } else {
      jm = nplanes;
      while (jm > 0 && zpl[jm - 1] > z) --jm;
      Dls = cosmo.coorDist(zpl[jm - 1], z);
    }
    aal = Dls / Ds;
    bbl = jm > 0 ? Dl[jm - 1] / Ds : 0.0;
  }
  st.jmax[slot] = valid ? jm : 0;
  st.aal[slot] = aal;
  st.bbl[slot] = bbl;
  st.eth[slot] = 0.f;
  st.eS[slot] = 0.f;
  st.risk[slot] = 0.f;
}

__global__ void k_finalize_array(RayState st, int64_t nslots, int64_t n,
                                 const uint32_t* __restrict__ perm, double* __restrict__ y,
                                 double* __restrict__ A, int lanes) {
  const int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  const int64_t k = array_ray_index(slot, lanes);
  if (k < 0 || k >= n) return;
  const int64_t i = perm ? (int64_t)perm[k] : k;
  y[2 * i] = st.th0[slot];
  y[2 * i + 1] = st.th1[slot];
  A[4 * i] = st.A0[slot];
  A[4 * i + 1] = st.A1[slot];
  A[4 * i + 2] = st.A2[slot];
  A[4 * i + 3] = st.A3[slot];
}

// ----------------------------------------------- precision-form guard
// A call's slots: ray i and its image position.  Grid: the tiles of a
// GridMap; array: slot s < n holds ray perm[s] of x.
__device__ inline bool slot_ray(const GuardMap& m, int64_t slot, int64_t* i, double* x0,
                                double* x1) {
  if (m.x == nullptr) {
    if (!grid_slot(slot, m.N, m.tiles_x, m.tw, i)) return false;
    grid_position(*i, m.N, m.c0, m.c1, m.range, x0, x1);
    return true;
  }
  if (slot >= m.n) return false;
  *i = m.perm ? (int64_t)m.perm[slot] : slot;
  *x0 = m.x[2 * *i];
  *x1 = m.x[2 * *i + 1];
  return true;
}

// sum of |alpha|^2 = |x - y|^2 over the call's rays (y on the source plane
// is the state's theta after the last step): per-block partial sums in a
// fixed order, then k_guard_total adds them in block order (repeatable)
__global__ void k_guard_sumsq(RayState st, int64_t nslots, GuardMap m, double* __restrict__ part) {
  __shared__ double sh[256];
  double v = 0;
  for (int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; slot < nslots;
       slot += (int64_t)gridDim.x * blockDim.x) {
    int64_t i;
    double x0, x1;
    if (!slot_ray(m, slot, &i, &x0, &x1)) continue;
    const double a0 = x0 - st.th0[slot], a1 = x1 - st.th1[slot];
    v += a0 * a0 + a1 * a1;
  }
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_guard_total(const double* __restrict__ part, int nblk, double* __restrict__ acc) {
  if (threadIdx.x != 0) return;
  double t = 0;
  for (int b = 0; b < nblk; ++b) t += part[b];
  *acc = t;
}

// rays to shoot again in FP64: flagged by the walk's guard, or with |alpha|
// below kGuardSmallAlpha of the rms (acc: the sum of |alpha|^2 over n rays)
__global__ void k_guard_flag(RayState st, int64_t nslots, GuardMap m, int64_t nrays,
                             const double* __restrict__ acc, uint32_t* __restrict__ count,
                             uint32_t cap, uint32_t* __restrict__ idx, double* __restrict__ xl,
                             double* __restrict__ zl) {
  const int64_t slot = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (slot >= nslots) return;
  int64_t i;
  double x0, x1;
  if (!slot_ray(m, slot, &i, &x0, &x1)) return;
  const double a0 = x0 - st.th0[slot], a1 = x1 - st.th1[slot];
  const double ms = *acc / (double)nrays;   // mean |alpha|^2
  const double a2 = a0 * a0 + a1 * a1;
  const double floor2 = 1e-6 * ms;
  const double lim = kGuardTol / kGuardSafety * sqrt(fmax(a2, floor2));
  if ((double)st.risk[slot] <= lim && a2 >= kGuardSmallAlpha * kGuardSmallAlpha * ms) return;
  const uint32_t k = atomicAdd(count, 1u);
  if (k >= cap) return;
  idx[k] = (uint32_t)i;
  xl[2 * k] = x0;
  xl[2 * k + 1] = x1;
  if (zl) zl[k] = m.zs[i];
}

__global__ void k_guard_scatter(uint32_t n, const uint32_t* __restrict__ idx,
                                const double* __restrict__ yt, const double* __restrict__ At,
                                double* __restrict__ y, double* __restrict__ A) {
  const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int64_t i = idx[k];
  y[2 * i] = yt[2 * k];
  y[2 * i + 1] = yt[2 * k + 1];
  for (int c = 0; c < 4; ++c) A[4 * i + c] = At[4 * k + c];
}

// 30-bit Morton code of each ray over the rays' bounding box.
__global__ void k_ray_morton(const double* __restrict__ x, int64_t n, double x0, double y0,
                             double inv_w, uint64_t* __restrict__ code, uint32_t* __restrict__ idx) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  double fx = (x[2 * i] - x0) * inv_w, fy = (x[2 * i + 1] - y0) * inv_w;
  uint32_t ux = (uint32_t)fmin(fmax(fx * 32767.0, 0.0), 32767.0);
  uint32_t uy = (uint32_t)fmin(fmax(fy * 32767.0, 0.0), 32767.0);
  uint64_t c = 0;
  for (int b = 14; b >= 0; --b) c = (c << 2) | (((uy >> b) & 1u) << 1) | ((ux >> b) & 1u);
  code[i] = c;
  idx[i] = (uint32_t)i;
}

__global__ void k_minmax_rays(const double* __restrict__ x, int64_t n, double* __restrict__ part) {
  __shared__ double s[4][256];
  double a0 = 1e300, a1 = 1e300, b0 = -1e300, b1 = -1e300;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x) {
    a0 = fmin(a0, x[2 * i]); b0 = fmax(b0, x[2 * i]);
    a1 = fmin(a1, x[2 * i + 1]); b1 = fmax(b1, x[2 * i + 1]);
  }
  s[0][threadIdx.x] = a0; s[1][threadIdx.x] = a1; s[2][threadIdx.x] = b0; s[3][threadIdx.x] = b1;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
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

// Tree walk alone at physical positions (one plane, no recursion): the
// force2D_recur outputs in the tree's sign convention, 5 per ray.
template <bool F32>
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_force(
    const TreeDesc* __restrict__ tree, const double* __restrict__ rays, int64_t n,
    double* __restrict__ out, uint32_t* __restrict__ err) {
  __shared__ int32_t s_node[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_mask[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_near[kWavesPerBlock][kNearQueue * kWave];
  const int w = uni((int)(threadIdx.x >> 6));
  const int64_t i = (blockIdx.x * (int64_t)kWavesPerBlock + w) * kWave + lane_id();
  const bool active = i < n;
  const double rx = active ? rays[2 * i] : 0.0, ry = active ? rays[2 * i + 1] : 0.0;
  const TreeDesc t = *as_const(tree);
  double r[5];
  if (F32) {
    Acc32 a;
    float unused = 0.f;   // no recursion here: no guard
    walk32<false>(tree, 1, rx, ry, active, a, s_node[w], s_mask[w], s_near[w], err, 0.0, 0.f,
                  unused);
    finish32(a, r);
    sheet32(t, rx, ry, r);
  } else {
    Acc64 a;
    Counts c;
    walk64<false>(t, rx, ry, active, a, c, s_node[w], s_mask[w], err);
    r[0] = a.ax; r[1] = a.ay; r[2] = a.k; r[3] = a.g0; r[4] = a.g1;
  }
  if (active)
    for (int k = 0; k < 5; ++k) out[5 * i + k] = r[k];
}

template __global__ void k_force<false>(const TreeDesc*, const double*, int64_t, double*, uint32_t*);
template __global__ void k_force<true>(const TreeDesc*, const double*, int64_t, double*, uint32_t*);

// Halo-of-halos tree walk alone at physical positions (TreeQuadHalos::force2D
// outputs, 5 per ray); direct != 0 sums every halo instead.
__global__ __launch_bounds__(kWave * kWavesPerBlock) void k_halo_force(
    const TreeDesc* __restrict__ tree, const double* __restrict__ rays, int64_t n, int direct,
    double* __restrict__ out, uint32_t* __restrict__ err) {
  __shared__ int32_t s_node[kWavesPerBlock][kStackDepth];
  __shared__ uint64_t s_mask[kWavesPerBlock][kStackDepth];
  const int w = uni((int)(threadIdx.x >> 6));
  const int64_t i = (blockIdx.x * (int64_t)kWavesPerBlock + w) * kWave + lane_id();
  const bool active = i < n;
  const double rx = active ? rays[2 * i] : 0.0, ry = active ? rays[2 * i + 1] : 0.0;
  const TreeDesc t = *as_const(tree);
  Acc64 a;
  if (direct) halo_tree_force64<true, true>(t, rx, ry, active, a, s_node[w], s_mask[w], err);
  else halo_tree_force64<false, true>(t, rx, ry, active, a, s_node[w], s_mask[w], err);
  if (active) {
    out[5 * i] = a.ax; out[5 * i + 1] = a.ay; out[5 * i + 2] = a.k;
    out[5 * i + 3] = a.g0; out[5 * i + 4] = a.g1;
  }
}

// FP64 direct sum of the smoothed-particle kernel over all particles
// (no tree); particles staged through LDS in tiles of 256.
__global__ __launch_bounds__(256) void k_direct(const PartD* __restrict__ parts, int64_t np,
                                                double ia, const double* __restrict__ rays,
                                                int64_t n, double* __restrict__ out) {
  __shared__ PartD tile[256];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool active = i < n;
  const double rx = active ? rays[2 * i] : 0.0, ry = active ? rays[2 * i + 1] : 0.0;
  double ax = 0, ay = 0, k = 0, g0 = 0, g1 = 0;
  for (int64_t base = 0; base < np; base += 256) {
    __syncthreads();
    if (base + threadIdx.x < np) tile[threadIdx.x] = parts[base + threadIdx.x];
    __syncthreads();
    const int cnt = (int)((np - base) < 256 ? (np - base) : 256);
    for (int q = 0; q < cnt; ++q) {
      const PartD p = tile[q];
      const double d0 = p.x - rx, d1 = p.y - ry;
      double rr = d0 * d0 + d1 * d1;
      if (rr < 1e-20) rr = 1e-20;
      const double m = p.m;
      const double inv = rcp64(rr);
      const double pref = m * inv * kInvPI;
      ax += (ia * m - pref) * d0;
      ay += (ia * m - pref) * d1;
      const double tmp = -2.0 * pref * inv;
      g0 += 0.5 * (d0 * d0 - d1 * d1) * tmp;
      g1 += d0 * d1 * tmp;
      k -= m * ia;
      const double size = fabs((double)p.h);
      if (rr < 4 * size * size) bspline64(d0, d1, sqrt(rr), m, size, ax, ay, k, g0, g1);
    }
  }
  if (active) {
    out[5 * i] = ax; out[5 * i + 1] = ay; out[5 * i + 2] = k;
    out[5 * i + 3] = g0; out[5 * i + 4] = g1;
  }
}

// ------------------------------------------------------------ host launchers
namespace {
inline unsigned blocks_for(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }
}

template <int HK>
void launch_step_hk(int precision, bool count, unsigned nb, hipStream_t st, const StepJob* jobs,
                    const WorkItem* items, uint32_t nitems, uint32_t* err) {
  const dim3 blk(kWave * kWavesPerBlock);
  if (count)
    k_plane_step<SLSIM_FP64, true, HK><<<nb, blk, 0, st>>>(jobs, items, nitems, err);
  else if (HK == 1 && precision == SLSIM_FP64)
    k_plane_step_nfw<SLSIM_FP64><<<nb, blk, 0, st>>>(jobs, items, nitems, err);
  else if (precision == SLSIM_FP32C && HK == 0)
    k_plane_step32<<<nb, blk, 0, st>>>(jobs, items, nitems, err);
  else if (precision == SLSIM_FP32C)
    k_plane_step<SLSIM_FP32C, false, HK><<<nb, blk, 0, st>>>(jobs, items, nitems, err);
  else if (precision == SLSIM_DIRECT64)
    k_plane_step<SLSIM_DIRECT64, false, HK><<<nb, blk, 0, st>>>(jobs, items, nitems, err);
  else if (HK == 0)
    k_plane_step64<<<nb, blk, 0, st>>>(jobs, items, nitems, err);
  else
    k_plane_step<SLSIM_FP64, false, HK><<<nb, blk, 0, st>>>(jobs, items, nitems, err);
}

void launch_plane_step(int precision, bool count, uint32_t nitems, hipStream_t st,
                       const StepJob* jobs, const WorkItem* items, uint32_t* err, int halos) {
  const unsigned nb = blocks_for(nitems, kWavesPerBlock);
  if (nb == 0) return;
  if (halos == 2) launch_step_hk<2>(precision, count, nb, st, jobs, items, nitems, err);
  else if (halos == 1) launch_step_hk<1>(precision, count, nb, st, jobs, items, nitems, err);
  else launch_step_hk<0>(precision, count, nb, st, jobs, items, nitems, err);
}

template <int HK>
void launch_fused_hk(int precision, unsigned nb, uint32_t nitems, hipStream_t st,
                     const StepJob* jobs, const FusedItem* items, uint32_t* counter,
                     uint32_t* err) {
  const dim3 blk(kWave * kWavesPerBlock);
  if (precision == SLSIM_FP32C && HK == 0)
    k_plane_steps_fused32<<<nb, blk, 0, st>>>(jobs, items, nitems, counter, err);
  else if (precision == SLSIM_FP32C)
    k_plane_steps_fused<SLSIM_FP32C, HK><<<nb, blk, 0, st>>>(jobs, items, nitems, counter, err);
  else if (precision == SLSIM_DIRECT64)
    k_plane_steps_fused<SLSIM_DIRECT64, HK><<<nb, blk, 0, st>>>(jobs, items, nitems, counter,
                                                                  err);
  else
    k_plane_steps_fused<SLSIM_FP64, HK><<<nb, blk, 0, st>>>(jobs, items, nitems, counter, err);
}

void launch_plane_steps_fused(int precision, unsigned nblocks, uint32_t nitems, hipStream_t st,
                              const StepJob* jobs, const FusedItem* items, uint32_t* counter,
                              uint32_t* err, int halos) {
  if (nitems == 0) return;
  unsigned nb = nblocks;
  if (nb == 0) {
    nb = blocks_for(nitems, kWavesPerBlock);
    counter = nullptr;
  } else {
    launch_zero_u32(st, counter, 1);
  }
  if (halos == 2) launch_fused_hk<2>(precision, nb, nitems, st, jobs, items, counter, err);
  else if (halos == 1) launch_fused_hk<1>(precision, nb, nitems, st, jobs, items, counter, err);
  else launch_fused_hk<0>(precision, nb, nitems, st, jobs, items, counter, err);
}

void launch_plane_step_lone(hipStream_t st, const StepJob* job, uint32_t nrays, uint32_t* err) {
  if (nrays > 0) k_plane_step_lone<<<nrays, kWave, 0, st>>>(job, nrays, err);
}

void launch_zero_u32(hipStream_t st, uint32_t* p, int n) {
  k_zero_u32<<<(n + 255) / 256, 256, 0, st>>>(p, n);
}

template <int HK>
void launch_queue_hk(int precision, unsigned nblocks, uint32_t nitems, hipStream_t st,
                     const StepJob* jobs, const WorkItem* items, uint32_t* counter,
                     uint32_t* err) {
  const dim3 blk(kWave * kWavesPerBlock);
  if (precision == SLSIM_FP32C && HK == 0)
    k_plane_step_queue32<<<nblocks, blk, 0, st>>>(jobs, items, nitems, counter, err);
  else if (precision == SLSIM_FP32C)
    k_plane_step_queue<SLSIM_FP32C, HK><<<nblocks, blk, 0, st>>>(jobs, items, nitems, counter,
                                                                 err);
  else if (precision == SLSIM_DIRECT64)
    k_plane_step_queue<SLSIM_DIRECT64, HK><<<nblocks, blk, 0, st>>>(jobs, items, nitems,
                                                                    counter, err);
  else
    k_plane_step_queue<SLSIM_FP64, HK><<<nblocks, blk, 0, st>>>(jobs, items, nitems, counter,
                                                                err);
}

void launch_plane_step_queue(int precision, unsigned nblocks, uint32_t nitems, hipStream_t st,
                             const StepJob* jobs, const WorkItem* items, uint32_t* counter,
                             uint32_t* err, int halos) {
  launch_zero_u32(st, counter, 1);
  if (halos == 2) launch_queue_hk<2>(precision, nblocks, nitems, st, jobs, items, counter, err);
  else if (halos == 1) launch_queue_hk<1>(precision, nblocks, nitems, st, jobs, items, counter, err);
  else launch_queue_hk<0>(precision, nblocks, nitems, st, jobs, items, counter, err);
}

void launch_init_grid(hipStream_t st, const RayState& s, int64_t nslots, int64_t N,
                      int64_t tiles_x, int tw, double c0, double c1, double range, int32_t jmax,
                      double aal, double bbl) {
  k_init_grid<<<blocks_for(nslots, 256), 256, 0, st>>>(s, nslots, N, tiles_x, tw, c0, c1, range,
                                                       jmax, aal, bbl);
}

void launch_finalize_grid(hipStream_t st, const RayState& s, int64_t nslots, int64_t N,
                          int64_t tiles_x, int tw, double* y, double* A) {
  k_finalize_grid<<<blocks_for(nslots, 256), 256, 0, st>>>(s, nslots, N, tiles_x, tw, y, A);
}

void launch_counts_grid(hipStream_t st, const uint32_t* c, int64_t nslots, int64_t N,
                        int64_t tiles_x, int tw, uint64_t* out) {
  k_counts_grid<<<blocks_for(nslots, 256), 256, 0, st>>>(c, nslots, N, tiles_x, tw, out);
}

void launch_init_batch(hipStream_t st, const GridReq* reqs, int nreq, int64_t total) {
  if (total > 0) k_init_batch<<<blocks_for(total, 256), 256, 0, st>>>(reqs, nreq, total);
}

void launch_finalize_batch(hipStream_t st, const GridReq* reqs, int nreq, int64_t total) {
  if (total > 0) k_finalize_batch<<<blocks_for(total, 256), 256, 0, st>>>(reqs, nreq, total);
}

void launch_init_array(hipStream_t st, const RayState& s, int64_t nslots, int64_t n,
                       const uint32_t* perm, const double* x, const double* zs, int32_t jmax,
                       double aal, double bbl, const double* zpl, const double* Dl,
                       int32_t nplanes, const FlatLCDM& cosmo, int lanes) {
  k_init_array<<<blocks_for(nslots, 256), 256, 0, st>>>(s, nslots, n, perm, x, zs, jmax, aal, bbl,
                                                        zpl, Dl, nplanes, cosmo, lanes);
}

void launch_finalize_array(hipStream_t st, const RayState& s, int64_t nslots, int64_t n,
                           const uint32_t* perm, double* y, double* A, int lanes) {
  k_finalize_array<<<blocks_for(nslots, 256), 256, 0, st>>>(s, nslots, n, perm, y, A, lanes);
}

void launch_guard_sumsq(hipStream_t st, const RayState& s, int64_t nslots, const GuardMap& m,
                        double* part, double* acc) {
  const unsigned nb = (unsigned)std::min<int64_t>(kGuardBlocks, (nslots + 255) / 256);
  if (nb == 0) return;
  k_guard_sumsq<<<nb, 256, 0, st>>>(s, nslots, m, part);
  k_guard_total<<<1, 64, 0, st>>>(part, (int)nb, acc);
}

void launch_guard_flag(hipStream_t st, const RayState& s, int64_t nslots, const GuardMap& m,
                       int64_t nrays, const double* acc, uint32_t* count, uint32_t cap,
                       uint32_t* idx, double* xl, double* zl) {
  k_guard_flag<<<blocks_for(nslots, 256), 256, 0, st>>>(s, nslots, m, nrays, acc, count, cap, idx,
                                                        xl, zl);
}

void launch_guard_scatter(hipStream_t st, uint32_t n, const uint32_t* idx, const double* yt,
                          const double* At, double* y, double* A) {
  k_guard_scatter<<<blocks_for(n, 256), 256, 0, st>>>(n, idx, yt, At, y, A);
}

void launch_ray_morton(hipStream_t st, const double* x, int64_t n, double x0, double y0,
                       double inv_w, uint64_t* code, uint32_t* idx) {
  k_ray_morton<<<blocks_for(n, 256), 256, 0, st>>>(x, n, x0, y0, inv_w, code, idx);
}

void launch_minmax_rays(hipStream_t st, const double* x, int64_t n, double* part, int nblk) {
  k_minmax_rays<<<nblk, 256, 0, st>>>(x, n, part);
}

void launch_force(hipStream_t st, int precision, const TreeDesc* tree, const double* rays,
                  int64_t n, double* out, uint32_t* err) {
  const unsigned nb = blocks_for(n, kWave * kWavesPerBlock);
  if (nb == 0) return;
  if (precision == SLSIM_FP32C)
    k_force<true><<<nb, kWave * kWavesPerBlock, 0, st>>>(tree, rays, n, out, err);
  else
    k_force<false><<<nb, kWave * kWavesPerBlock, 0, st>>>(tree, rays, n, out, err);
}

void launch_halo_force(hipStream_t st, const TreeDesc* tree, const double* rays, int64_t n,
                       int direct, double* out, uint32_t* err) {
  const unsigned nb = blocks_for(n, kWave * kWavesPerBlock);
  if (nb == 0) return;
  k_halo_force<<<nb, kWave * kWavesPerBlock, 0, st>>>(tree, rays, n, direct, out, err);
}

void launch_direct(hipStream_t st, const PartD* parts, int64_t np, double ia, const double* rays,
                   int64_t n, double* out) {
  k_direct<<<blocks_for(n, 256), 256, 0, st>>>(parts, np, ia, rays, n, out);
}

}  // namespace slsim
