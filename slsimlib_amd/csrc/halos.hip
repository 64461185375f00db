// halos.hip -- halo-of-halos planes: LensHaloHalos<LensHaloNFW>.
//
// A LensHaloHalos (include/particle_halo.h:1092-1227) is a main halo made of
// many analytic halos on a TreeQuadHalos quadtree (include/quadTreeHalos.h);
// with it a light-cone mixes particle halos and NFW halos.  Halo counts are
// small next to particle counts (1e2-1e5 per plane), so the tree is built on
// the host exactly as the reference builds it (the same partitions in the
// same order, hence the same cells, big-halo lists and moments) and copied to
// HBM; the walk (walk_halos64) and LensHaloNFW's force run on the device in
// FP64.  LensHaloNFW's interpolation tables (MultiPlane/lens_halos.cpp:
// 352-353, 482-540, the profile functions of AnalyticNSIE/nfw_lens.cpp:
// 82-125) are made once per process on the host and copied per plane.

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "slsim_internal.h"

namespace slsim {

namespace {

// ------------------------------------------------------------ NFW tables
double g_fn(double x) {
  if (x < 1e-5) x = 1e-5;
  const double ans = std::log(x / 2);
  if (x == 1.0) return ans + 1.0;
  if (x > 1.0) return ans + 2 * std::atan(std::sqrt((x - 1) / (x + 1))) / std::sqrt(x * x - 1);
  return ans + 2 * std::atanh(std::sqrt((1 - x) / (x + 1))) / std::sqrt(1 - x * x);
}
double f_fn(double x) {
  if (x == 0) x = 1e-5;
  if (x == 1.0) return 1.0 / 3.0;
  if (x > 1.0)
    return (1 - 2 * std::atan(std::sqrt((x - 1) / (x + 1))) / std::sqrt(x * x - 1)) / (x * x - 1);
  return (1 - 2 * std::atanh(std::sqrt((1 - x) / (x + 1))) / std::sqrt(1 - x * x)) / (x * x - 1);
}
double g2_fn(double x) {
  if (x == 0) x = 1e-5;
  if (x == 1) return 10 / 3. + 4 * std::log(0.5);
  const double y = x * x - 1;
  const double ans = 4 * std::log(x / 2) / x / x - 2 / y;
  if (x < 1.0) return ans + 4 * (2 / x / x + 1 / y) * std::atanh(std::sqrt((1 - x) / (x + 1))) / std::sqrt(-y);
  return ans + 4 * (2 / x / x + 1 / y) * std::atan(std::sqrt((x - 1) / (x + 1))) / std::sqrt(y);
}

const std::vector<NfwEntry>& nfw_tables() {
  static std::vector<NfwEntry> t;
  static std::once_flag once;
  std::call_once(once, [] {
    t.resize(kNfwTable);
    const double dx = kNfwMaxrm / (double)kNfwTable;
    for (int i = 0; i < kNfwTable; ++i) {
      const double x = i * dx;
      t[i] = {x, g_fn(x), f_fn(x), g2_fn(x), 0, 0, 0, 0};
    }
    for (int i = 0; i + 1 < kNfwTable; ++i) {
      const double w = t[i + 1].x - t[i].x;
      t[i].sg = (t[i + 1].g - t[i].g) / w;
      t[i].sf = (t[i + 1].f - t[i].f) / w;
      t[i].sg2 = (t[i + 1].g2 - t[i].g2) / w;
    }
  });
  return t;
}

// InterpolateFromTable(gtable, y) (MultiPlane/lens_halos.cpp:566-581)
double interp_g(double y) {
  const int j = (int)(y / kNfwMaxrm * kNfwTable);
  if (j == 0) return g_fn(y);
  const auto& T = nfw_tables();
  return (T[j + 1].g - T[j].g) / (T[j + 1].x - T[j].x) * (y - T[j].x) + T[j].g;
}

// --------------------------------------------------------------- the tree
// TreeQuadHalos' BuildQTreeNB / _BuildQTreeNB / CalcMoments (include/
// quadTreeHalos.h:167-486), the reference's partitions in its order.
struct HBuild {
  const double* x;
  const double* y;
  const float* rmax;
  int bucket;
  double theta;
  std::vector<unsigned long> index;
  std::vector<double> work;
  struct Cell {
    double p1[2], p2[2];
    size_t off, n;
    int32_t child = -1;
    std::vector<int32_t> big;
  };
  std::vector<Cell> cells;
  int err = SLSIM_OK;
};

// Utilities::quickPartition (include/Tree.h:408-424): values <= pivot first,
// in the reference's swap order
void quick_partition(double pivot, size_t* cut, unsigned long* particles, double* arr, size_t n) {
  size_t p = 0;
  for (size_t i = 0; i < n; ++i) {
    if (arr[i] <= pivot) {
      std::swap(arr[p], arr[i]);
      std::swap(particles[p], particles[i]);
      ++p;
    }
  }
  *cut = p;
}

void build_cell(HBuild& B, size_t ci, bool top, int level) {
  if (B.err) return;
  if (level > 60) {
    set_error("halo tree deeper than 60 levels: more than `bucket` halos in one place");
    B.err = SLSIM_ERR_TOO_DEEP;
    return;
  }
  const double p1[2] = {B.cells[ci].p1[0], B.cells[ci].p1[1]};
  const double p2[2] = {B.cells[ci].p2[0], B.cells[ci].p2[1]};
  const size_t off = B.cells[ci].off, n = B.cells[ci].n;
  unsigned long* particles = B.index.data() + off;
  const double center[2] = {(p1[0] + p2[0]) / 2, (p1[1] + p2[1]) / 2};
  if (n <= (size_t)B.bucket) {
    // leaf: every halo is a big halo
    B.cells[ci].big.assign(particles, particles + n);
    return;
  }
  double* x = B.work.data();
  // halos too big for the children: Rmax in (box/2, box], all above box/2 at the top
  for (size_t i = 0; i < n; ++i) x[i] = B.rmax[particles[i]];
  const double half_box = (p2[0] - p1[0]) / 2;
  size_t cut, cut2;
  quick_partition(half_box, &cut, particles, x, n);
  if (cut < n) {
    if (top) {
      cut2 = n - cut;
    } else {
      quick_partition(2 * half_box, &cut2, &particles[cut], &x[cut], n - cut);
    }
    B.cells[ci].big.assign(particles + cut, particles + cut + cut2);
  }
  // all halos go on to the children (the reference keeps the big ones too)
  const size_t first = B.cells.size();
  B.cells.resize(first + 4);
  HBuild::Cell* ch = &B.cells[first];
  ch[0].p1[0] = center[0]; ch[0].p1[1] = center[1]; ch[0].p2[0] = p2[0]; ch[0].p2[1] = p2[1];
  ch[1].p1[0] = p1[0]; ch[1].p1[1] = center[1]; ch[1].p2[0] = center[0]; ch[1].p2[1] = p2[1];
  ch[2].p1[0] = center[0]; ch[2].p1[1] = p1[1]; ch[2].p2[0] = p2[0]; ch[2].p2[1] = center[1];
  ch[3].p1[0] = p1[0]; ch[3].p1[1] = p1[1]; ch[3].p2[0] = center[0]; ch[3].p2[1] = center[1];
  B.cells[ci].child = (int32_t)first;
  const double xcut = center[0], ycut = center[1];
  size_t cutx, cuty;
  for (size_t i = 0; i < n; ++i) x[i] = B.y[particles[i]];
  quick_partition(ycut, &cuty, particles, x, n);
  ch = &B.cells[first];
  if (cuty > 0) {
    for (size_t i = 0; i < cuty; ++i) x[i] = B.x[particles[i]];
    quick_partition(xcut, &cutx, particles, x, cuty);
    ch[3].off = off; ch[3].n = cutx;
    ch[2].off = off + cutx; ch[2].n = cuty - cutx;
  } else {
    ch[3].off = off; ch[3].n = 0;
    ch[2].off = off; ch[2].n = 0;
  }
  if (cuty < n) {
    for (size_t i = cuty; i < n; ++i) x[i - cuty] = B.x[particles[i]];
    quick_partition(xcut, &cutx, &particles[cuty], x, n - cuty);
    ch[1].off = off + cuty; ch[1].n = cutx;
    ch[0].off = off + cuty + cutx; ch[0].n = n - cuty - cutx;
  } else {
    ch[1].off = off + n; ch[1].n = 0;
    ch[0].off = off + n; ch[0].n = 0;
  }
  for (int k = 0; k < 4; ++k) build_cell(B, first + k, false, level + 1);
}

}  // namespace

// Builds the plane's halo tree and copies it to HBM.
int build_halo_plane(slsim_plane p, const double* x, const double* y, const float* mass,
                     const float* rsize, const int32_t* kind, const double* param, size_t n,
                     double inv_area, double theta_force) {
  // LensHaloNFW(mass, Rsize, zlens, concentration, fratio 1, ...)
  // (MultiPlane/lens_halos.cpp:389-440): rscale = Rsize / c and xmax =
  // Rsize / rscale in float, gmax = g(xmax) from the table, Rmax = Rsize.
  // LensHaloPowerLaw(mass, Rsize, zlens, beta, fratio 1, ...)
  // (lens_halos.cpp:805-866): beta < 2, rscale 1, xmax = Rmax = Rsize.
  std::vector<HaloNFW> h(n);
  std::vector<float> rmax(n);
  for (size_t i = 0; i < n; ++i) {
    const int k = kind ? kind[i] : 0;
    if (!std::isfinite(x[i]) || !std::isfinite(y[i]) || !(mass[i] >= 0) || (k != 0 && k != 1)) {
      set_error("halo " + std::to_string(i) + ": need a finite position, mass >= 0 and a kind "
                "of 0 (NFW) or 1 (power law)");
      return SLSIM_ERR_ARG;
    }
    h[i].x = x[i];
    h[i].y = y[i];
    h[i].mass = mass[i];
    h[i].rmax = rsize[i];
    h[i].mass_pi = (double)mass[i] / kPI;
    rmax[i] = rsize[i];
    if (k == 1) {
      const double beta = param[i];
      if (!(beta < 2.0) || !(rsize[i] > 0)) {
        set_error("halo " + std::to_string(i) + ": a power law needs beta < 2 (\"bad gamma\") "
                  "and Rsize > 0");
        return SLSIM_ERR_ARG;
      }
      h[i].rscale = 1.0f;
      h[i].gmax = beta;
      h[i].inv_gmax = 0;
      h[i].inv_rscale = 1.0;
      h[i].pad = 1.0f;
      p->hpowerlaw = 1;
      continue;
    }
    const float conc = (float)param[i];
    const float rs = rsize[i] / conc;
    const float xm = rsize[i] / rs;
    if (!(rsize[i] > 0) || !(conc > 0) || !(rs > 0) || !(xm < 99.98f)) {
      set_error("halo " + std::to_string(i) +
                ": need Rsize > 0, concentration > 0 with Rsize/rscale below 99.98 (the NFW "
                "tables end at x = 100), mass >= 0 and a finite position");
      return SLSIM_ERR_ARG;
    }
    h[i].rscale = rs;
    h[i].gmax = interp_g((double)xm);
    h[i].inv_gmax = 1.0 / h[i].gmax;
    h[i].inv_rscale = 1.0 / (double)rs;
    h[i].pad = 0;
  }
  HBuild B;
  B.x = x;
  B.y = y;
  B.rmax = rmax.data();
  B.bucket = 5;
  B.theta = theta_force;
  B.index.resize(n);
  for (size_t i = 0; i < n; ++i) B.index[i] = i;
  B.work.resize(n);
  double p1[2] = {x[0], y[0]}, p2[2] = {x[0], y[0]};
  for (size_t i = 0; i < n; ++i) {
    const double v[2] = {x[i], y[i]};
    for (int j = 0; j < 2; ++j) {
      if (v[j] < p1[j]) p1[j] = v[j];
      if (v[j] > p2[j]) p2[j] = v[j];
    }
  }
  if (n == 1) {
    p1[0] = p1[1] = -0.25;
    p2[0] = p2[1] = 0.25;
  }
  const double lengths[2] = {p2[0] - p1[0], p2[1] - p1[1]};
  p->hxl = lengths[0];   // original_xl, original_yl: the periodic buffer's shifts
  p->hyl = lengths[1];
  if (lengths[0] == 0 || lengths[1] == 0) {
    set_error("particles in same place.");
    return SLSIM_ERR_SAME_PLACE;
  }
  const int j = lengths[0] > lengths[1] ? 1 : 0;
  p2[j] = p1[j] + lengths[!j];
  B.cells.resize(1);
  B.cells[0].p1[0] = p1[0]; B.cells[0].p1[1] = p1[1];
  B.cells[0].p2[0] = p2[0]; B.cells[0].p2[1] = p2[1];
  B.cells[0].off = 0;
  B.cells[0].n = n;
  build_cell(B, 0, true, 0);
  if (B.err) return B.err;
  // CalcMoments over the final index order (include/quadTreeHalos.h:425-486)
  const size_t nc = B.cells.size();
  if (nc >= ((size_t)1 << 28)) {
    set_error("halo tree has more than 2^28 cells");
    return SLSIM_ERR_TOO_DEEP;
  }
  std::vector<HNode> nodes(nc);
  std::vector<int32_t> big;
  double theta = theta_force;
  for (size_t c = 0; c < nc; ++c) {
    const HBuild::Cell& cl = B.cells[c];
    HNode& nd = nodes[c];
    std::memset(&nd, 0, sizeof nd);
    const unsigned long* pt = B.index.data() + cl.off;
    double m = 0;
    for (size_t i = 0; i < cl.n; ++i) m += (double)mass[pt[i]];
    double cx = 0, cy = 0;
    for (size_t i = 0; i < cl.n; ++i) {
      const double tmp = mass[pt[i]];
      cx += tmp * x[pt[i]] / m;
      cy += tmp * y[pt[i]] / m;
    }
    double q0 = 0, q1 = 0, q2 = 0;
    for (size_t i = 0; i < cl.n; ++i) {
      const double x0 = x[pt[i]] - cx, x1 = y[pt[i]] - cy;
      const double xcut = x0 * x0 + x1 * x1;
      const double tmp = mass[pt[i]];
      q0 += (xcut - 2 * x0 * x0) * tmp;
      q1 += (xcut - 2 * x1 * x1) * tmp;
      q2 += -2 * x0 * x1 * tmp;
    }
    double rcom = 0;
    const double c2[2] = {cx, cy};
    for (int i = 0; i < 2; ++i) {
      const double a = (c2[i] - cl.p1[i]) * (c2[i] - cl.p1[i]);
      const double b = (c2[i] - cl.p2[i]) * (c2[i] - cl.p2[i]);
      rcom += a > b ? a : b;
    }
    rcom = std::sqrt(rcom);
    double r2crit;
    if (theta > 0.0) {
      r2crit = 1.15470 * rcom / theta;
      r2crit *= r2crit;
    } else {
      r2crit = 1.0e100;
    }
    nd.cx = cx;
    nd.cy = cy;
    nd.mass = m;
    nd.q0 = q0;
    nd.q1 = q1;
    nd.q2 = q2;
    nd.r2crit = r2crit;
    nd.child = cl.child;
    nd.cmask = 0;
    if (cl.child >= 0)
      for (int k = 0; k < 4; ++k)
        if (B.cells[cl.child + k].n > 0) nd.cmask |= 1 << k;
    nd.big_off = (int32_t)big.size();
    nd.nbig = (int32_t)cl.big.size();
    big.insert(big.end(), cl.big.begin(), cl.big.end());
  }
  if (big.empty()) big.push_back(0);
  // device copies
  slsim_ctx c = p->ctx;
  const auto& T = nfw_tables();
  int rc;
  if ((rc = c->alloc(reinterpret_cast<void**>(&p->hnodes), sizeof(HNode) * nc)) ||
      (rc = c->alloc(reinterpret_cast<void**>(&p->hbig), sizeof(int32_t) * big.size())) ||
      (rc = c->alloc(reinterpret_cast<void**>(&p->halos), sizeof(HaloNFW) * n)) ||
      (rc = c->alloc(reinterpret_cast<void**>(&p->nfw), sizeof(NfwEntry) * T.size())))
    return rc;
  SLSIM_HIP_CHECK(hipMemcpyAsync(p->hnodes, nodes.data(), sizeof(HNode) * nc,
                                 hipMemcpyHostToDevice, c->stream));
  SLSIM_HIP_CHECK(hipMemcpyAsync(p->hbig, big.data(), sizeof(int32_t) * big.size(),
                                 hipMemcpyHostToDevice, c->stream));
  SLSIM_HIP_CHECK(hipMemcpyAsync(p->halos, h.data(), sizeof(HaloNFW) * n, hipMemcpyHostToDevice,
                                 c->stream));
  SLSIM_HIP_CHECK(hipMemcpyAsync(p->nfw, T.data(), sizeof(NfwEntry) * T.size(),
                                 hipMemcpyHostToDevice, c->stream));
  SLSIM_HIP_CHECK(hipStreamSynchronize(c->stream));
  p->kind = kTreeHalos;
  p->n = (int64_t)n;
  p->nnodes = (int64_t)nc;
  p->inv_area = inv_area;
  p->theta = theta;
  p->bucket = 5;
  p->p1[0] = p1[0]; p->p1[1] = p1[1];
  p->p2[0] = p2[0]; p->p2[1] = p2[1];
  p->mass = nodes[0].mass;
  p->com[0] = nodes[0].cx;
  p->com[1] = nodes[0].cy;
  int64_t leaves = 0;
  for (const auto& nd : nodes) leaves += nd.child < 0;
  p->nleaves = leaves;
  p->bytes = sizeof(HNode) * nc + sizeof(int32_t) * big.size() + sizeof(HaloNFW) * n +
             sizeof(NfwEntry) * T.size();
  p->halos_host = std::move(h);
  return SLSIM_OK;
}

}  // namespace slsim
