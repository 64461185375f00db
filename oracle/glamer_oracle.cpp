// glamer_oracle.cpp -- CPU restatement of GLAMER's particle ray-shooting path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the timed
// CPU baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load it; the product library never links or calls it.
//
// It restates, in FP64 with float32 particle masses and sizes, the reference
// functions listed below (paths relative to the reference root):
//   TreeQuadParticles ctor / BuildQTreeNB / _BuildQTreeNB / CalcMoments
//                                        include/quadTree.h:247-628
//   TreeQuadParticles::b_spline_profile  include/quadTree.h:147-190
//   TreeQuadParticles::force2D_recur     include/quadTree.h:932-987
//   TreeQuadParticles::walkTree_recur    include/quadTree.h:990-1128
//   Utilities::quickPartition            include/Tree.h:396-411
//   LensHaloParticles::force_halo        include/particle_halo.h:426-434
//   LensPlaneSingular::force             MultiPlane/planes.cpp:114-161
//   Lens::addMainHaloToPlane             MultiPlane/lens.cpp:772-815
//   Lens::combinePlanes                  MultiPlane/lens.cpp:2397-2515
//   Lens::FindSourcePlane                MultiPlane/lens.cpp:2659-2681
//   Lens::compute_points_parallel        include/lens.h:913-1103
//   Lens::rayshooterInternal (threads)   FullRange/internal_rayshooter_multi.cpp:53-118
//   Utilities::PositionFromIndex         TreeCode_link/utilities.cpp:149-158
//   constants PI, Grav, arcsecTOradians  include/standard.h:38-51
//   LensHaloParticles::calculate_smoothing include/particle_halo.h:683-730
//   TreeSimple ctor / BuildTreeNB / _BuildTreeNB (median cut)
//                                        include/simpleTree.h:264-278, 842-960
//   TreeSimple::NNDistance / _findleaf   include/simpleTree.h:483-640
//   LensHaloNFW tables, functions, force_halo  MultiPlane/lens_halos.cpp:352-581,
//                                        1528-1672; AnalyticNSIE/nfw_lens.cpp:82-125
//   TreeQuadHalos build / CalcMoments / walkTree_iter  include/quadTreeHalos.h:127-735
//   LensHaloHalos::force_halo            include/particle_halo.h:1199-1209
//
// Cosmological distances come from CosmoLib in the reference, which is not
// part of the reference tree.  Here they are the flat-LCDM comoving distance
// D_C = (c/H0) * int dz / sqrt(Om (1+z)^3 + OL), integrated by adaptive
// Gauss-Kronrod (7/15) to 1e-14 relative.
//
// A second oracle, orc_direct_plane, sums the same smoothed-particle kernel
// (point mass plus the reference's B-spline correction for r < 2h) over all
// particles with no tree.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <queue>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace orc {

static const double PI = M_PI;               // standard.h:38-40 (M_PI)
static const double Grav = 4.7788e-20;       // standard.h:43, G/c^2 in Mpc/Msun
static const double MIN_PLANE_DIST = 1e-6;   // MultiPlane/lens.cpp:15

// Interaction counters of one walk (per ray per plane).
struct Counters {
  uint64_t visits = 0;      // nodes with nparticles > 0 given the open test
  uint64_t far = 0;         // far-cell (multipole) evaluations
  uint64_t leafp = 0;       // leaf particle point-mass interactions
  uint64_t bigtest = 0;     // big-particle candidate tests
  uint64_t bspline = 0;     // B-spline corrections applied
  void add(const Counters& o) {
    visits += o.visits; far += o.far; leafp += o.leafp;
    bigtest += o.bigtest; bspline += o.bspline;
  }
};

// ---------------------------------------------------------------- cosmology
struct Cosmo {
  double Om, OL, h;
  double DH() const { return 2997.92458 / h; }   // c/H0 in Mpc
  double invE(double z) const {
    double a = 1 + z;
    return 1.0 / std::sqrt(Om * a * a * a + OL);
  }
  double gk15(double a, double b, double* err) const {
    static const double xgk[8] = {0.991455371120812639206854697526329,
                                  0.949107912342758524526189684047851,
                                  0.864864423359769072789712788640926,
                                  0.741531185599394439863864773280788,
                                  0.586087235467691130294144845693013,
                                  0.405845151377397166906606412076961,
                                  0.207784955007898467600689403773245,
                                  0.000000000000000000000000000000000};
    static const double wgk[8] = {0.022935322010529224963732008058970,
                                  0.063092092629978553290700663189204,
                                  0.104790010322250183839876322541518,
                                  0.140653259715525918745189590510238,
                                  0.169004726639267902826583426598550,
                                  0.190350578064785409913256402421014,
                                  0.204432940075298892414161999234649,
                                  0.209482141084727828012999174891714};
    static const double wg[4] = {0.129484966168869693270611432679082,
                                 0.279705391489276667901467771423780,
                                 0.381830050505118944950369775488975,
                                 0.417959183673469387755102040816327};
    double c = 0.5 * (a + b), hl = 0.5 * (b - a);
    double fc = invE(c);
    double rk = fc * wgk[7], rg = fc * wg[3];
    for (int j = 0; j < 7; ++j) {
      double f1 = invE(c - hl * xgk[j]), f2 = invE(c + hl * xgk[j]);
      rk += wgk[j] * (f1 + f2);
      if (j % 2 == 1) rg += wg[j / 2] * (f1 + f2);
    }
    *err = std::fabs((rk - rg) * hl);
    return rk * hl;
  }
  double integ(double a, double b, int depth = 0) const {
    double err;
    double r = gk15(a, b, &err);
    if (depth > 40 || err <= 1e-15 * std::fabs(r) + 1e-300) return r;
    double m = 0.5 * (a + b);
    return integ(a, m, depth + 1) + integ(m, b, depth + 1);
  }
  // comoving distance between z1 and z2 (Mpc)
  double coorDist(double z1, double z2) const {
    if (z2 == z1) return 0.0;
    if (z2 < z1) return -coorDist(z2, z1);
    return DH() * integ(z1, z2);
  }
  double angDist(double z) const { return coorDist(0, z) / (1 + z); }
};

// ---------------------------------------------------------------- quadtree
struct Node {
  double p1[2], p2[2];
  double center[2];
  double mass = 0, quad[3] = {0, 0, 0}, r2crit = 0;
  size_t* particles = nullptr;   // points into TreeQuad::index
  size_t nparticles = 0;
  std::vector<size_t> big;      // big_particles
  int child[4] = {-1, -1, -1, -1};
  int level = 0;
  bool leaf() const {
    return child[0] < 0 && child[1] < 0 && child[2] < 0 && child[3] < 0;
  }
};

struct TreeQuad {
  // particles (positions promoted to double exactly as PType::operator[] -> PosType)
  std::vector<double> px, py;
  std::vector<float> pm, ph;
  size_t N = 0;
  double inv_area = 0;
  int Nbucket = 20;
  double force_theta = 0.1;
  double max_range = -1;   // LensHaloParticles' maximum range (particle_halo.h:416)
  std::vector<size_t> index;
  std::vector<Node> nodes;   // nodes[0] = root
  std::vector<double> workspace;

  // Utilities::quickPartition, include/Tree.h:396-411 (Lomuto, x <= pivot first)
  static void quickPartition(double pivot, size_t* pivotindex, size_t* particles,
                             double* arr, size_t n) {
    *pivotindex = 0;
    for (size_t i = 0; i < n; ++i) {
      if (arr[i] <= pivot) {
        std::swap(arr[*pivotindex], arr[i]);
        std::swap(particles[*pivotindex], particles[i]);
        ++(*pivotindex);
      }
    }
  }

  // BuildQTreeNB, include/quadTree.h:295-350
  void build() {
    double p1[2] = {px[0], py[0]}, p2[2] = {px[0], py[0]};
    for (size_t i = 0; i < N; ++i) {
      if (px[i] < p1[0]) p1[0] = px[i];
      if (px[i] > p2[0]) p2[0] = px[i];
      if (py[i] < p1[1]) p1[1] = py[i];
      if (py[i] > p2[1]) p2[1] = py[i];
    }
    if (N == 1) { p1[0] = -0.25; p1[1] = -0.25; p2[0] = 0.25; p2[1] = 0.25; }
    double lengths[2] = {p2[0] - p1[0], p2[1] - p1[1]};
    if (lengths[0] == 0 || lengths[1] == 0)
      throw std::invalid_argument("particles in same place.");
    int j = lengths[0] > lengths[1] ? 1 : 0;
    p2[j] = p1[j] + lengths[!j];

    nodes.clear();
    nodes.reserve(N / 4 + 16);
    Node root;
    root.p1[0] = p1[0]; root.p1[1] = p1[1];
    root.p2[0] = p2[0]; root.p2[1] = p2[1];
    root.particles = index.data();
    root.nparticles = N;
    root.level = 0;
    nodes.push_back(std::move(root));
    workspace.resize(N);
    build_rec(0);
    workspace.clear();
    workspace.shrink_to_fit();
  }

  // _BuildQTreeNB, include/quadTree.h:358-565 (MultiRadius = MultiMass = true)
  void build_rec(int ni) {
    Node* cb = &nodes[ni];
    cb->center[0] = (cb->p1[0] + cb->p2[0]) / 2;
    cb->center[1] = (cb->p1[1] + cb->p2[1]) / 2;
    cb->quad[0] = cb->quad[1] = cb->quad[2] = 0;
    cb->mass = 0.0;
    double theta_range = 2 * force_theta;
    if (max_range > 0) {
      // quadTree.h:370-376: cells far beyond max_range stay unsplit
      const double boxsize = 1.732 * (cb->p2[0] - cb->p1[0]);
      const double d =
          std::sqrt(cb->center[0] * cb->center[0] + cb->center[1] * cb->center[1]) - max_range;
      theta_range = boxsize / (d > boxsize ? d : boxsize);
    }
    const double box = cb->p2[0] - cb->p1[0];
    size_t* particles = cb->particles;
    const size_t n = cb->nparticles;

    if (n <= (size_t)Nbucket || force_theta > theta_range) {
      cb->big.clear();
      for (size_t i = 0; i < n; ++i) {
        double r = ph[particles[i]];
        if (r < box) cb->big.push_back(particles[i]);
      }
      return;
    }

    double* x = workspace.data();
    size_t cut, cut2;
    for (size_t i = 0; i < n; ++i) x[i] = ph[particles[i]];
    double maxsize = box / 2;
    quickPartition(maxsize, &cut, particles, x, n);
    if (cut < n) {
      if (ni == 0) {
        cut2 = n - cut;
      } else {
        quickPartition(2 * maxsize, &cut2, &particles[cut], &x[cut], n - cut);
      }
      cb->big.assign(particles + cut, particles + cut + cut2);
    }

    const size_t NpInChildren = n;
    if (NpInChildren == 0) return;

    // children: 0 = (x>xc, y>yc), 1 = (x<=xc, y>yc), 2 = (x>xc, y<=yc), 3 = (x<=xc, y<=yc)
    int first = (int)nodes.size();
    for (int c = 0; c < 4; ++c) {
      Node ch;
      ch.level = nodes[ni].level + 1;
      nodes.push_back(std::move(ch));
    }
    cb = &nodes[ni];   // re-fetch after possible reallocation
    for (int c = 0; c < 4; ++c) cb->child[c] = first + c;
    Node* c0 = &nodes[first + 0];
    Node* c1 = &nodes[first + 1];
    Node* c2 = &nodes[first + 2];
    Node* c3 = &nodes[first + 3];
    c0->p1[0] = cb->center[0]; c0->p1[1] = cb->center[1];
    c0->p2[0] = cb->p2[0];     c0->p2[1] = cb->p2[1];
    c1->p1[0] = cb->p1[0];     c1->p1[1] = cb->center[1];
    c1->p2[0] = cb->center[0]; c1->p2[1] = cb->p2[1];
    c2->p1[0] = cb->center[0]; c2->p1[1] = cb->p1[1];
    c2->p2[0] = cb->p2[0];     c2->p2[1] = cb->center[1];
    c3->p1[0] = cb->p1[0];     c3->p1[1] = cb->p1[1];
    c3->p2[0] = cb->center[0]; c3->p2[1] = cb->center[1];

    const double xcut = cb->center[0], ycut = cb->center[1];
    size_t cutx, cuty;
    for (size_t i = 0; i < NpInChildren; ++i) x[i] = py[particles[i]];
    quickPartition(ycut, &cuty, particles, x, NpInChildren);
    if (cuty > 0) {
      for (size_t i = 0; i < cuty; ++i) x[i] = px[particles[i]];
      quickPartition(xcut, &cutx, particles, x, cuty);
      c3->nparticles = cutx;
      c3->particles = cutx > 0 ? particles : nullptr;
      c2->nparticles = cuty - cutx;
      c2->particles = (cuty - cutx) > 0 ? &particles[cutx] : nullptr;
    } else {
      c3->nparticles = 0; c3->particles = nullptr;
      c2->nparticles = 0; c2->particles = nullptr;
    }
    if (cuty < NpInChildren) {
      for (size_t i = cuty; i < NpInChildren; ++i) x[i - cuty] = px[particles[i]];
      quickPartition(xcut, &cutx, &particles[cuty], x, NpInChildren - cuty);
      c1->nparticles = cutx;
      c1->particles = cutx > 0 ? &particles[cuty] : nullptr;
      c0->nparticles = NpInChildren - cuty - cutx;
      c0->particles = c0->nparticles > 0 ? &particles[cuty + cutx] : nullptr;
    } else {
      c1->nparticles = 0; c1->particles = nullptr;
      c0->nparticles = 0; c0->particles = nullptr;
    }
    for (int c = 0; c < 4; ++c) build_rec(first + c);
  }

  // CalcMoments, include/quadTree.h:567-628
  void calc_moments() {
    for (auto& b : nodes) {
      double absmass = 0;
      b.mass = 0;
      for (size_t i = 0; i < b.nparticles; ++i) {
        b.mass += (double)pm[b.particles[i]];
        absmass += std::fabs((double)pm[b.particles[i]]);
      }
      b.center[0] = b.center[1] = 0;
      for (size_t i = 0; i < b.nparticles; ++i) {
        double tmp = std::fabs((double)pm[b.particles[i]]);
        b.center[0] += tmp * px[b.particles[i]] / absmass;
        b.center[1] += tmp * py[b.particles[i]] / absmass;
      }
      b.quad[0] = b.quad[1] = b.quad[2] = 0;
      for (size_t i = 0; i < b.nparticles; ++i) {
        double xcm0 = px[b.particles[i]] - b.center[0];
        double xcm1 = py[b.particles[i]] - b.center[1];
        double xcut = xcm0 * xcm0 + xcm1 * xcm1;
        double tmp = pm[b.particles[i]];
        b.quad[0] += (xcut - 2 * xcm0 * xcm0) * tmp;
        b.quad[1] += (xcut - 2 * xcm1 * xcm1) * tmp;
        b.quad[2] += -2 * xcm0 * xcm1 * tmp;
      }
      double rcom = 0.0;
      for (int i = 0; i < 2; ++i) {
        double a = (b.center[i] - b.p1[i]) * (b.center[i] - b.p1[i]);
        double c = (b.center[i] - b.p2[i]) * (b.center[i] - b.p2[i]);
        rcom += a > c ? a : c;
      }
      rcom = std::sqrt(rcom);
      if (force_theta > 0.0) {
        b.r2crit = 1.15470 * rcom / force_theta;
        b.r2crit *= b.r2crit;
      } else {
        b.r2crit = 1.0e100;
      }
    }
  }

  // b_spline_profile, include/quadTree.h:147-190 (phi omitted: dead downstream)
  static void b_spline(const double* xcm, double r, double Mass, double size,
                       double* alpha, double* kappa, double* gamma) {
    double q = r / size;
    double M, sigma;
    if (q > 2) {
      sigma = 0;
      M = 1;
    } else {
      double q2 = q * q, q3 = q2 * q;
      sigma = (8 - 12 * q + 6 * q2 - q3) / 4;
      if (q > 1) {
        sigma *= 10 / size / size / 7 / PI;
        M = (-1 + 20 * q2 * (1 - q + 3 * q2 / 8 - q3 / 20)) / 7;
      } else {
        sigma = 10 * (sigma - 1 + 3 * q - 3 * q2 + q3) / size / size / 7 / PI;
        M = 10 * q2 * (1 - 3 * q / 4 + 3 * q3 / 10) / 7;
      }
    }
    double alpha_r = (M - 1) / PI / r;
    double gt = alpha_r / r - sigma;
    alpha[0] -= Mass * alpha_r * xcm[0] / r;
    alpha[1] -= Mass * alpha_r * xcm[1] / r;
    gamma[0] -= gt * Mass * (xcm[0] * xcm[0] - xcm[1] * xcm[1]) / r / r;
    gamma[1] -= gt * Mass * 2 * xcm[0] * xcm[1] / r / r;
    *kappa += Mass * sigma;
  }

  // walkTree_recur, include/quadTree.h:990-1128 (screening = 0, phi omitted)
  void walk(int ni, const double* ray, double* alpha, double* kappa, double* gamma,
            Counters* cnt) const {
    const Node& b = nodes[ni];
    if (b.nparticles == 0) return;
    double xcm[2] = {b.center[0] - ray[0], b.center[1] - ray[1]};
    double rcm2cell = xcm[0] * xcm[0] + xcm[1] * xcm[1];
    double boxsize2 = (b.p2[0] - b.p1[0]) * (b.p2[0] - b.p1[0]);
    if (cnt) cnt->visits++;
    if (rcm2cell < b.r2crit || rcm2cell < 5.83 * boxsize2) {
      if (b.leaf()) {
        for (size_t i = 0; i < b.nparticles; ++i) {
          size_t p = b.particles[i];
          double d[2] = {px[p] - ray[0], py[p] - ray[1]};
          double rcm2 = d[0] * d[0] + d[1] * d[1];
          if (rcm2 < 1e-20) rcm2 = 1e-20;
          double mass = pm[p];
          double prefac = mass / rcm2 / PI;
          alpha[0] += (inv_area * mass - prefac) * d[0];
          alpha[1] += (inv_area * mass - prefac) * d[1];
          double tmp = -2.0 * prefac / rcm2;
          gamma[0] += 0.5 * (d[0] * d[0] - d[1] * d[1]) * tmp;
          gamma[1] += d[0] * d[1] * tmp;
          *kappa -= mass * inv_area;
        }
        if (cnt) cnt->leafp += b.nparticles;
      }
      if (rcm2cell < 5.83 * boxsize2) {
        for (size_t p : b.big) {
          double d[2] = {px[p] - ray[0], py[p] - ray[1]};
          double rcm2 = d[0] * d[0] + d[1] * d[1];
          if (rcm2 < 1e-20) rcm2 = 1e-20;
          double size = ph[p];
          if (cnt) cnt->bigtest++;
          if (rcm2 < 4 * size * size) {
            b_spline(d, std::sqrt(rcm2), (double)pm[p], size, alpha, kappa, gamma);
            if (cnt) cnt->bspline++;
          }
        }
      }
      for (int c = 0; c < 4; ++c)
        if (b.child[c] >= 0) walk(b.child[c], ray, alpha, kappa, gamma, cnt);
    } else {
      if (cnt) cnt->far++;
      double tmp = b.mass * inv_area - b.mass / rcm2cell / PI;
      alpha[0] += tmp * xcm[0];
      alpha[1] += tmp * xcm[1];
      tmp = -2.0 * b.mass / PI / rcm2cell / rcm2cell;
      gamma[0] += 0.5 * (xcm[0] * xcm[0] - xcm[1] * xcm[1]) * tmp;
      gamma[1] += xcm[0] * xcm[1] * tmp;
      alpha[0] -= (b.quad[0] * xcm[0] + b.quad[2] * xcm[1]) / (rcm2cell * rcm2cell) / PI;
      alpha[1] -= (b.quad[1] * xcm[1] + b.quad[2] * xcm[0]) / (rcm2cell * rcm2cell) / PI;
      tmp = 4 * (b.quad[0] * xcm[0] * xcm[0] + b.quad[1] * xcm[1] * xcm[1] +
                 2 * b.quad[2] * xcm[0] * xcm[1]) /
            (rcm2cell * rcm2cell * rcm2cell) / PI;
      alpha[0] += tmp * xcm[0];
      alpha[1] += tmp * xcm[1];
      *kappa -= inv_area * b.mass;
    }
  }

  // force2D_recur, include/quadTree.h:932-987 (periodic_buffer = false)
  void force2D(const double* ray, double* alpha, double* kappa, double* gamma,
               Counters* cnt) const {
    alpha[0] = alpha[1] = gamma[0] = gamma[1] = gamma[2] = 0.0;
    *kappa = 0.0;
    if (N == 0) return;
    walk(0, ray, alpha, kappa, gamma, cnt);
  }
};

// ------------------------------------------------------------ LensHaloNFW
// Tables and profile functions (MultiPlane/lens_halos.cpp:352-353, 482-540,
// 566-581; AnalyticNSIE/nfw_lens.cpp:82-125).
struct NfwProfile {
  static constexpr long NTABLE = 10000;
  static constexpr double maxrm = 100.0;
  std::vector<double> xtable, ftable, gtable, g2table;
  static double gfunction(double x) {
    double ans;
    if (x < 1e-5) x = 1e-5;
    ans = log(x / 2);
    if (x == 1.0) { ans += 1.0; return ans; }
    if (x > 1.0) { ans += 2 * atan(sqrt((x - 1) / (x + 1))) / sqrt(x * x - 1); return ans; }
    ans += 2 * atanh(sqrt((1 - x) / (x + 1))) / sqrt(1 - x * x);
    return ans;
  }
  static double ffunction(double x) {
    if (x == 0) x = 1e-5;
    if (x == 1.0) return 1.0 / 3.0;
    if (x > 1.0) return (1 - 2 * atan(sqrt((x - 1) / (x + 1))) / sqrt(x * x - 1)) / (x * x - 1);
    return (1 - 2 * atanh(sqrt((1 - x) / (x + 1))) / sqrt(1 - x * x)) / (x * x - 1);
  }
  static double g2function(double x) {
    double ans, y;
    if (x == 0) x = 1e-5;
    if (x == 1) return 10 / 3. + 4 * log(0.5);
    y = x * x - 1;
    ans = 4 * log(x / 2) / x / x - 2 / y;
    if (x < 1.0) { ans += 4 * (2 / x / x + 1 / y) * atanh(sqrt((1 - x) / (x + 1))) / sqrt(-y); return ans; }
    ans += 4 * (2 / x / x + 1 / y) * atan(sqrt((x - 1) / (x + 1))) / sqrt(y);
    return ans;
  }
  NfwProfile() {
    const double dx = maxrm / (double)NTABLE;
    xtable.resize(NTABLE); ftable.resize(NTABLE); gtable.resize(NTABLE); g2table.resize(NTABLE);
    for (long i = 0; i < NTABLE; ++i) {
      const double x = i * dx;
      xtable[i] = x;
      ftable[i] = ffunction(x);
      gtable[i] = gfunction(x);
      g2table[i] = g2function(x);
    }
  }
  // which: 0 g, 1 f, 2 g2
  double interpolate(int which, double y) const {
    const int j = (int)(y / maxrm * NTABLE);
    if (j == 0) return which == 0 ? gfunction(y) : which == 1 ? ffunction(y) : g2function(y);
    if (j + 1 >= NTABLE) throw std::out_of_range("NFW table ends at maxrm");
    const std::vector<double>& t = which == 0 ? gtable : which == 1 ? ftable : g2table;
    return (t[j + 1] - t[j]) / (xtable[j + 1] - xtable[j]) * (y - xtable[j]) + t[j];
  }
  static const NfwProfile& get() {
    static NfwProfile p;
    return p;
  }
};

// One LensHaloNFW(mass, Rsize, zlens, concentration, fratio = 1) at a
// physical position (MultiPlane/lens_halos.cpp:389-440) and its force_halo
// (LensHalo::force_halo -> force_point / force_halo_sym, :1528-1672).
// A symmetric halo of a halo-of-halos or field tree: LensHaloNFW
// (lens_halos.cpp:389-440; alpha_h, kappa_h, gamma_h from its tables,
// include/lens_halos.h) or LensHaloPowerLaw (lens_halos.cpp:805-866,
// lens_halos.h:979-997: rscale 1, xmax = Rmax = Rsize, kappa ~ r^-beta).
struct NfwHalo {
  double pos[2];
  float mass, Rsize, rscale, Rmax;
  double xmax, gmax;
  bool power_law = false;
  double beta = 0;
  NfwHalo(double x, double y, float m, float rsize, float conc) {
    pos[0] = x;
    pos[1] = y;
    mass = m;
    Rsize = rsize;
    rscale = Rsize / conc;
    xmax = Rsize / rscale;
    gmax = NfwProfile::get().interpolate(0, xmax);
    Rmax = Rsize;
  }
  static NfwHalo power(double x, double y, float m, float rsize, double my_beta) {
    if (my_beta >= 2.0) throw std::invalid_argument("bad gamma");
    if (rsize <= 0.0f || std::isnan(rsize)) throw std::runtime_error("Bad halo");
    NfwHalo h(x, y, m, rsize, 1.0f);
    h.power_law = true;
    h.beta = my_beta;
    h.rscale = 1;
    h.xmax = h.Rsize / h.rscale;
    h.gmax = 0;
    h.Rmax = h.Rsize;
    return h;
  }
  double alpha_h(double x) const {
    if (power_law) {
      if (x < xmax * 1.0e-6) x = 1e-6 * xmax;
      return -1.0 * pow(x / xmax, -beta + 2);
    }
    return -1.0 * NfwProfile::get().interpolate(0, x) / gmax;
  }
  double kappa_h(double x) const {
    if (power_law) {
      if (x == 0) x = 1e-6 * xmax;
      const double tmp = x / xmax;
      return 0.5 * (-beta + 2) * pow(tmp, 2 - beta);
    }
    return 0.5 * x * x * NfwProfile::get().interpolate(1, x) / gmax;
  }
  double gamma_h(double x) const {
    if (power_law) {
      if (x == 0) x = 1e-6 * xmax;
      return -beta * pow(x / xmax, -beta + 2);
    }
    return -0.5 * x * x * NfwProfile::get().interpolate(2, x) / gmax;
  }
  void force_halo(double* alpha, double* kappa, double* gamma, const double* xcm) const {
    double rcm2 = xcm[0] * xcm[0] + xcm[1] * xcm[1];
    // force_point without point subtraction
    if (!(rcm2 < Rmax * Rmax)) {
      const double prefac = mass / rcm2 / PI;
      alpha[0] += -1.0 * prefac * xcm[0];
      alpha[1] += -1.0 * prefac * xcm[1];
      const double tmp = -2.0 * prefac / rcm2;
      gamma[0] += 0.5 * (xcm[0] * xcm[0] - xcm[1] * xcm[1]) * tmp;
      gamma[1] += xcm[0] * xcm[1] * tmp;
      return;
    }
    // force_halo_sym
    if (rcm2 < 1e-20) rcm2 = 1e-20;
    if (rcm2 < Rmax * Rmax) {
      const double prefac = mass / rcm2 / PI;
      const double x = sqrt(rcm2) / rscale;
      double tmp = (alpha_h(x) + 1.0 * false) * prefac;
      alpha[0] += tmp * xcm[0];
      alpha[1] += tmp * xcm[1];
      *kappa += kappa_h(x) * prefac;
      tmp = (gamma_h(x) + 2.0 * false) * prefac / rcm2;
      gamma[0] += 0.5 * (xcm[0] * xcm[0] - xcm[1] * xcm[1]) * tmp;
      gamma[1] += xcm[0] * xcm[1] * tmp;
    } else {
      const double prefac = mass / rcm2 / PI;
      alpha[0] += -1.0 * prefac * xcm[0];
      alpha[1] += -1.0 * prefac * xcm[1];
      const double tmp = -2.0 * prefac / rcm2;
      gamma[0] += 0.5 * (xcm[0] * xcm[0] - xcm[1] * xcm[1]) * tmp;
      gamma[1] += xcm[0] * xcm[1] * tmp;
    }
  }
};

// ----------------------------------------------------------- TreeQuadHalos
// include/quadTreeHalos.h:127-735 with bucket 5 and no maximum range:
// pointer cells, Utilities::quickPartition, the build and moments in the
// reference's order, walkTree_iter's depth-first walk with the screening
// factor exp(-r^2 s^2) on the cells' far terms (quadTreeHalos.h:679-733), and
// force2D's periodic buffer (nine walks at the ray shifted by the halos'
// extents, :566-580).  LensHaloHalos uses no screening and no buffer;
// LensPlaneTree (MultiPlane/planes.cpp:13-19) uses both, or neither with
// theta 0.5.
struct HaloTree {
  struct Branch {
    unsigned long* particles = nullptr;
    unsigned long nparticles = 0;
    std::vector<unsigned long> big;
    double center[2] = {0, 0};
    double mass = 0;
    double quad[3] = {0, 0, 0};
    double p1[2] = {0, 0}, p2[2] = {0, 0};
    double r2crit = 0;
    std::unique_ptr<Branch> child[4];
    bool leaf() const { return !child[0]; }
  };
  std::vector<NfwHalo> halos;
  std::vector<unsigned long> index;
  std::vector<double> work;
  double inv_area, theta;
  double s2 = 0;               // inv_screening_scale2
  bool periodic = false;
  double xl = 0, yl = 0;       // original_xl, original_yl
  int bucket = 5;
  std::unique_ptr<Branch> top;

  static void quickPartition(double pivot, unsigned long* pivotindex, unsigned long* particles,
                             double* arr, unsigned long N) {
    *pivotindex = 0;
    for (unsigned long i = 0; i < N; ++i) {
      if (arr[i] <= pivot) {
        std::swap(arr[*pivotindex], arr[i]);
        std::swap(particles[*pivotindex], particles[i]);
        ++(*pivotindex);
      }
    }
  }

  HaloTree(std::vector<NfwHalo> h, double my_inv_area, double my_theta, double inv_screening = 0,
           bool my_periodic = false)
      : halos(std::move(h)), inv_area(my_inv_area), theta(my_theta),
        s2(inv_screening * inv_screening), periodic(my_periodic) {
    const unsigned long N = halos.size();
    index.resize(N);
    for (unsigned long i = 0; i < N; ++i) index[i] = i;
    if (N == 0) return;
    double p1[2], p2[2];
    for (int j = 0; j < 2; ++j) p1[j] = p2[j] = halos[0].pos[j];
    for (unsigned long i = 0; i < N; ++i)
      for (int j = 0; j < 2; ++j) {
        if (halos[i].pos[j] < p1[j]) p1[j] = halos[i].pos[j];
        if (halos[i].pos[j] > p2[j]) p2[j] = halos[i].pos[j];
      }
    if (N == 1) { p1[0] = -0.25; p1[1] = -0.25; p2[0] = 0.25; p2[1] = 0.25; }
    const double lengths[2] = {p2[0] - p1[0], p2[1] - p1[1]};
    xl = lengths[0];
    yl = lengths[1];
    if (lengths[0] == 0 || lengths[1] == 0) throw std::invalid_argument("particles in same place.");
    const int j = lengths[0] > lengths[1] ? 1 : 0;
    p2[j] = p1[j] + lengths[!j];
    top.reset(new Branch);
    top->particles = index.data();
    top->nparticles = N;
    for (int k = 0; k < 2; ++k) { top->p1[k] = p1[k]; top->p2[k] = p2[k]; }
    work.resize(N);
    build(top.get(), true, 0);
    moments(top.get());
  }

  void build(Branch* b, bool at_top, int level) {
    if (level > 200) throw std::runtime_error("halo tree too deep");
    b->center[0] = (b->p1[0] + b->p2[0]) / 2;
    b->center[1] = (b->p1[1] + b->p2[1]) / 2;
    if (b->nparticles <= (unsigned long)bucket) {
      b->big.assign(b->particles, b->particles + b->nparticles);
      return;
    }
    double* x = work.data();
    unsigned long* particles = b->particles;
    const unsigned long n = b->nparticles;
    for (unsigned long i = 0; i < n; ++i) x[i] = halos[particles[i]].Rmax;
    const double half_box = (b->p2[0] - b->p1[0]) / 2;
    unsigned long cut, cut2;
    quickPartition(half_box, &cut, particles, x, n);
    if (cut < n) {
      if (at_top) cut2 = n - cut;
      else quickPartition(2 * half_box, &cut2, &particles[cut], &x[cut], n - cut);
      b->big.assign(particles + cut, particles + cut + cut2);
    }
    for (int k = 0; k < 4; ++k) b->child[k].reset(new Branch);
    Branch* c0 = b->child[0].get(); Branch* c1 = b->child[1].get();
    Branch* c2 = b->child[2].get(); Branch* c3 = b->child[3].get();
    c0->p1[0] = b->center[0]; c0->p1[1] = b->center[1]; c0->p2[0] = b->p2[0]; c0->p2[1] = b->p2[1];
    c1->p1[0] = b->p1[0]; c1->p1[1] = b->center[1]; c1->p2[0] = b->center[0]; c1->p2[1] = b->p2[1];
    c2->p1[0] = b->center[0]; c2->p1[1] = b->p1[1]; c2->p2[0] = b->p2[0]; c2->p2[1] = b->center[1];
    c3->p1[0] = b->p1[0]; c3->p1[1] = b->p1[1]; c3->p2[0] = b->center[0]; c3->p2[1] = b->center[1];
    const double xcut = b->center[0], ycut = b->center[1];
    unsigned long cutx, cuty;
    for (unsigned long i = 0; i < n; ++i) x[i] = halos[particles[i]].pos[1];
    quickPartition(ycut, &cuty, particles, x, n);
    if (cuty > 0) {
      for (unsigned long i = 0; i < cuty; ++i) x[i] = halos[particles[i]].pos[0];
      quickPartition(xcut, &cutx, particles, x, cuty);
      c3->nparticles = cutx; c3->particles = particles;
      c2->nparticles = cuty - cutx; c2->particles = &particles[cutx];
    }
    if (cuty < n) {
      for (unsigned long i = cuty; i < n; ++i) x[i - cuty] = halos[particles[i]].pos[0];
      quickPartition(xcut, &cutx, &particles[cuty], x, n - cuty);
      c1->nparticles = cutx; c1->particles = &particles[cuty];
      c0->nparticles = n - cuty - cutx; c0->particles = &particles[cuty + cutx];
    }
    for (int k = 0; k < 4; ++k) build(b->child[k].get(), false, level + 1);
  }

  void moments(Branch* b) {
    b->mass = 0;
    for (unsigned long i = 0; i < b->nparticles; ++i) b->mass += halos[b->particles[i]].mass;
    b->center[0] = b->center[1] = 0;
    for (unsigned long i = 0; i < b->nparticles; ++i) {
      const double tmp = halos[b->particles[i]].mass;
      b->center[0] += tmp * halos[b->particles[i]].pos[0] / b->mass;
      b->center[1] += tmp * halos[b->particles[i]].pos[1] / b->mass;
    }
    b->quad[0] = b->quad[1] = b->quad[2] = 0;
    for (unsigned long i = 0; i < b->nparticles; ++i) {
      const double xcm[2] = {halos[b->particles[i]].pos[0] - b->center[0],
                             halos[b->particles[i]].pos[1] - b->center[1]};
      const double xcut = pow(xcm[0], 2) + pow(xcm[1], 2);
      const double tmp = halos[b->particles[i]].mass;
      b->quad[0] += (xcut - 2 * xcm[0] * xcm[0]) * tmp;
      b->quad[1] += (xcut - 2 * xcm[1] * xcm[1]) * tmp;
      b->quad[2] += -2 * xcm[0] * xcm[1] * tmp;
    }
    double rcom = 0;
    for (int i = 0; i < 2; ++i)
      rcom += (pow(b->center[i] - b->p1[i], 2) > pow(b->center[i] - b->p2[i], 2))
                  ? pow(b->center[i] - b->p1[i], 2) : pow(b->center[i] - b->p2[i], 2);
    rcom = sqrt(rcom);
    if (theta > 0.0) {
      b->r2crit = 1.15470 * rcom / theta;
      b->r2crit *= b->r2crit;
    } else {
      b->r2crit = 1.0e100;
    }
    if (!b->leaf())
      for (int k = 0; k < 4; ++k) moments(b->child[k].get());
  }

  void walk(const Branch* b, const double* ray, double* alpha, double* kappa, double* gamma) const {
    for (unsigned long t : b->big) {
      const NfwHalo& h = halos[t];
      const double xcm[2] = {h.pos[0] - ray[0], h.pos[1] - ray[1]};
      // LensHalo::force_halo(..., subtract_point = false, screening) leaves
      // the screening unused for a symmetric halo (lens_halos.cpp:1555-1600)
      h.force_halo(alpha, kappa, gamma, xcm);
      const double mass = h.mass;
      const double prefac = mass * inv_area;
      alpha[0] += prefac * xcm[0];
      alpha[1] += prefac * xcm[1];
      *kappa -= mass * inv_area;
    }
    if (b->leaf()) return;
    const double xcm[2] = {b->center[0] - ray[0], b->center[1] - ray[1]};
    const double rcm2cell = xcm[0] * xcm[0] + xcm[1] * xcm[1];
    if (rcm2cell > b->r2crit) {
      const double screening = exp(-rcm2cell * s2);
      double tmp = b->mass * (inv_area - screening / rcm2cell / PI);
      alpha[0] += tmp * xcm[0];
      alpha[1] += tmp * xcm[1];
      tmp = -2.0 * b->mass / PI / rcm2cell / rcm2cell * screening;
      gamma[0] += 0.5 * (xcm[0] * xcm[0] - xcm[1] * xcm[1]) * tmp;
      gamma[1] += xcm[0] * xcm[1] * tmp;
      alpha[0] -= (b->quad[0] * xcm[0] + b->quad[2] * xcm[1]) / (rcm2cell * rcm2cell) / PI * screening;
      alpha[1] -= (b->quad[1] * xcm[1] + b->quad[2] * xcm[0]) / (rcm2cell * rcm2cell) / PI * screening;
      tmp = 4 * (b->quad[0] * xcm[0] * xcm[0] + b->quad[1] * xcm[1] * xcm[1] +
                 2 * b->quad[2] * xcm[0] * xcm[1]) / (rcm2cell * rcm2cell * rcm2cell) / PI * screening;
      alpha[0] += tmp * xcm[0];
      alpha[1] += tmp * xcm[1];
      *kappa -= b->mass * inv_area;
      return;
    }
    for (int k = 0; k < 4; ++k) walk(b->child[k].get(), ray, alpha, kappa, gamma);
  }

  // force2D (include/quadTreeHalos.h:557-592)
  void force2D(const double* ray, double* alpha, double* kappa, double* gamma) const {
    alpha[0] = alpha[1] = gamma[0] = gamma[1] = gamma[2] = 0.0;
    *kappa = 0.0;
    if (halos.empty()) return;
    if (periodic) {
      for (int i = -1; i < 2; ++i)
        for (int j = -1; j < 2; ++j) {
          const double r[2] = {ray[0] + i * xl, ray[1] + j * yl};
          walk(top.get(), r, alpha, kappa, gamma);
        }
    } else {
      walk(top.get(), ray, alpha, kappa, gamma);
    }
  }

  // every halo with its background term, no tree (nine shifted copies of the
  // ray with the periodic buffer)
  void direct(const double* ray, double* alpha, double* kappa, double* gamma) const {
    alpha[0] = alpha[1] = gamma[0] = gamma[1] = gamma[2] = 0.0;
    *kappa = 0.0;
    for (int i = -1; i < 2; ++i)
      for (int j = -1; j < 2; ++j) {
        if (!periodic && (i != 0 || j != 0)) continue;
        const double r[2] = {periodic ? ray[0] + i * xl : ray[0],
                             periodic ? ray[1] + j * yl : ray[1]};
        for (const NfwHalo& h : halos) {
          const double xcm[2] = {h.pos[0] - r[0], h.pos[1] - r[1]};
          h.force_halo(alpha, kappa, gamma, xcm);
          alpha[0] += h.mass * inv_area * xcm[0];
          alpha[1] += h.mass * inv_area * xcm[1];
          *kappa -= h.mass * inv_area;
        }
      }
  }
};

// LensHaloParticles or LensHaloHalos with posHalo = (0,0): force_halo negates
// the tree's alpha (include/particle_halo.h:426-434, 1199-1209).
struct Halo {
  std::unique_ptr<TreeQuad> tree;
  std::unique_ptr<HaloTree> htree;
  float z;
  void force_halo(const double* x, double* alpha, double* kappa, double* gamma,
                  Counters* cnt) const {
    if (htree) htree->force2D(x, alpha, kappa, gamma);
    else tree->force2D(x, alpha, kappa, gamma, cnt);
    alpha[0] *= -1;
    alpha[1] *= -1;
  }
};

// LensPlaneSingular::force, MultiPlane/planes.cpp:114-161
struct Plane {
  std::vector<Halo*> halos;
  void force(double* alpha, double* kappa, double* gamma, const double* xx,
             Counters* cnt) const {
    alpha[0] = alpha[1] = 0.0;
    *kappa = 0.0;
    gamma[0] = gamma[1] = gamma[2] = 0.0;
    for (const Halo* h : halos) {
      double at[2] = {0, 0}, kt = 0, gt[3] = {0, 0, 0};
      double xt[2] = {xx[0] - 0.0, xx[1] - 0.0};   // getX() = (0,0) for particles
      h->force_halo(xt, at, &kt, gt, cnt);
      alpha[0] -= at[0];
      alpha[1] -= at[1];
      *kappa += kt;
      gamma[0] += gt[0];
      gamma[1] += gt[1];
      gamma[2] += gt[2];
    }
  }
};

struct Lens {
  Cosmo cosmo;
  double zsource;
  double charge;
  std::vector<std::unique_ptr<Halo>> halos;
  std::vector<std::unique_ptr<Plane>> main_planes;
  std::vector<double> main_Dl, main_z;
  // field planes (LensPlaneTree), redshifts as doubles (lens.cpp:910-918)
  std::vector<std::unique_ptr<Plane>> field_planes;
  std::vector<double> field_Dl, field_z;
  // combined
  std::vector<Plane*> planes;
  std::vector<double> plane_redshifts, Dl, dDl;

  Lens(double zs, double Om, double OL, double h) : zsource(zs) {
    cosmo = Cosmo{Om, OL, h};
    charge = 4 * PI * Grav;   // MultiPlane/lens.cpp:93
    combinePlanes();
  }

  // addMainHaloToPlane, MultiPlane/lens.cpp:772-815
  void addMainHaloToPlane(Halo* halo) {
    double halo_z = halo->z;
    double halo_Dl = cosmo.coorDist(0, halo_z);
    size_t i = std::upper_bound(main_Dl.begin(), main_Dl.end(), halo_Dl) - main_Dl.begin();
    if (i > 0 && (halo_Dl - main_Dl[i - 1]) < MIN_PLANE_DIST) {
      main_planes[i - 1]->halos.push_back(halo);
    } else if (i == main_Dl.size()) {
      auto p = std::make_unique<Plane>();
      p->halos.push_back(halo);
      main_planes.push_back(std::move(p));
      main_z.push_back(halo_z);
      main_Dl.push_back(halo_Dl);
    } else if ((main_Dl[i] - halo_Dl) < MIN_PLANE_DIST) {
      main_planes[i]->halos.push_back(halo);
    } else {
      auto p = std::make_unique<Plane>();
      p->halos.push_back(halo);
      main_planes.insert(main_planes.begin() + i, std::move(p));
      main_z.insert(main_z.begin() + i, halo_z);
      main_Dl.insert(main_Dl.begin() + i, halo_Dl);
    }
  }

  // combinePlanes, MultiPlane/lens.cpp:2397-2515: field then main distances,
  // sorted together (ties keep field planes first)
  void combinePlanes() {
    planes.clear();
    plane_redshifts.clear();
    Dl = field_Dl;
    Dl.insert(Dl.end(), main_Dl.begin(), main_Dl.end());
    dDl.clear();
    std::vector<size_t> idx(Dl.size());
    for (size_t i = 0; i < idx.size(); ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return Dl[a] < Dl[b]; });
    std::sort(Dl.begin(), Dl.end());
    for (size_t i = 1; i < Dl.size(); ++i)
      if ((Dl[i] - Dl[i - 1]) < MIN_PLANE_DIST) Dl[i] = Dl[i - 1] + MIN_PLANE_DIST;
    const size_t nf = field_Dl.size();
    for (size_t i : idx) {
      if (i < nf) {
        plane_redshifts.push_back(field_z[i]);
        planes.push_back(field_planes[i].get());
      } else {
        plane_redshifts.push_back(main_z[i - nf]);
        planes.push_back(main_planes[i - nf].get());
      }
    }
    if (!plane_redshifts.empty() && zsource <= plane_redshifts.back())
      zsource = plane_redshifts.back() + 0.1;
    plane_redshifts.push_back(zsource);
    Dl.push_back(cosmo.coorDist(0, zsource));
    dDl.push_back(Dl[0]);
    for (size_t i = 1; i < Dl.size(); ++i) dDl.push_back(Dl[i] - Dl[i - 1]);
  }

  // a LensPlaneTree field plane at redshift z: its force is the tree's
  // force2D (planes.cpp:65-68), i.e. a plane of one halo-of-halos at (0,0)
  void addFieldPlane(std::unique_ptr<Halo> h, double z) {
    Halo* raw = h.get();
    halos.push_back(std::move(h));
    auto p = std::make_unique<Plane>();
    p->halos.push_back(raw);
    field_planes.push_back(std::move(p));
    field_z.push_back(z);
    field_Dl.push_back(cosmo.coorDist(0, z));
    combinePlanes();
  }

  // moveinMainHalo(addplanes = true), include/lens.h:172-185
  void moveinMainHalo(std::unique_ptr<Halo> h) {
    Halo* raw = h.get();
    halos.push_back(std::move(h));
    addMainHaloToPlane(raw);
    combinePlanes();
  }

  // addMainHaloToNearestPlane, MultiPlane/lens.cpp:822-840: the halo goes to
  // the main plane closest in distance; with no main plane it makes one and
  // is then added to it a second time.  Utilities::closest is not in the
  // reference tree: the first plane at the least |Dl - halo_Dl| is taken.
  void moveinMainHaloNearest(std::unique_ptr<Halo> h) {
    Halo* raw = h.get();
    halos.push_back(std::move(h));
    const double halo_Dl = cosmo.coorDist(0, raw->z);
    if (main_Dl.empty()) {
      auto p = std::make_unique<Plane>();
      p->halos.push_back(raw);
      main_planes.push_back(std::move(p));
      main_z.push_back(raw->z);
      main_Dl.push_back(halo_Dl);
    }
    size_t i = 0;
    for (size_t k = 1; k < main_Dl.size(); ++k)
      if (std::fabs(main_Dl[k] - halo_Dl) < std::fabs(main_Dl[i] - halo_Dl)) i = k;
    main_planes[i]->halos.push_back(raw);
    combinePlanes();
  }

  // FindSourcePlane, MultiPlane/lens.cpp:2659-2681
  void FindSourcePlane(double zs, long& jmax, double& Dls, double& Ds) const {
    Ds = cosmo.coorDist(0, zs);
    if (zs <= plane_redshifts[0]) { jmax = 0; Dls = Ds; return; }
    jmax = (long)plane_redshifts.size() - 1;
    if (jmax <= 0) { Dls = Ds; return; }
    while (plane_redshifts[jmax - 1] > zs) --jmax;
    Dls = cosmo.coorDist(plane_redshifts[jmax - 1], zs);
  }

  // compute_points_parallel, include/lens.h:913-1103, on SoA rays.
  // x: 2n (x0,x1 interleaved), y: 2n out, A: 4n out (a0..a3), zs: per-ray or null.
  void compute_points(size_t start, size_t end, const double* x, const double* zs_ray,
                      double source_z, double* y, double* A, Counters* cnt_ray) const {
    long jmax = (long)planes.size();
    double Dls_Ds = 0, D_Ds = 0;
    if (!zs_ray) {
      if (source_z == plane_redshifts.back()) {
        Dls_Ds = dDl.back() / Dl.back();
        D_Ds = Dl[Dl.size() - 2] / Dl.back();
      } else {
        double Dls, Ds;
        FindSourcePlane(source_z, jmax, Dls, Ds);
        Dls_Ds = Dls / Ds;
        if (jmax > 0) D_Ds = Dl[jmax - 1] / Ds;
      }
    }
    for (size_t i = start; i < end; ++i) {
      double theta[2] = {x[2 * i], x[2 * i + 1]};
      double SPA[2] = {theta[0], theta[1]};
      double SAG[4] = {1, 0, 0, 1};
      double Ai[4] = {1, 0, 0, 1};
      Counters c;
      if (zs_ray) {
        double Dls, Ds;
        FindSourcePlane(zs_ray[i], jmax, Dls, Ds);
        Dls_Ds = Dls / Ds;
        if (jmax > 0) D_Ds = Dl[jmax - 1] / Ds;
      }
      for (long j = 0; j < jmax; ++j) {
        double Dphysical = Dl[j] / (1 + plane_redshifts[j]);
        double xx[2] = {theta[0] * Dphysical, theta[1] * Dphysical};
        double alpha[2], kappa, gamma[3];
        planes[j]->force(alpha, &kappa, gamma, xx, cnt_ray ? &c : nullptr);
        double G[4] = {kappa + gamma[0], gamma[1], gamma[1], kappa - gamma[0]};
        double f = charge * Dl[j] / (1 + plane_redshifts[j]);
        for (int k = 0; k < 4; ++k) G[k] *= f;
        SPA[0] -= charge * alpha[0];
        SPA[1] -= charge * alpha[1];
        double aa, bb;
        if (j < jmax - 1) {
          aa = dDl[j + 1] / Dl[j + 1];
          bb = Dl[j] / Dl[j + 1];
        } else {
          aa = Dls_Ds;
          bb = D_Ds;
        }
        theta[0] = bb * theta[0] + aa * SPA[0];
        theta[1] = bb * theta[1] + aa * SPA[1];
        // SumPrevAG -= G * A  (Matrix2x2::operator*, include/point.h:258-268)
        double GA[4] = {G[0] * Ai[0] + G[1] * Ai[2], G[0] * Ai[1] + G[1] * Ai[3],
                        G[2] * Ai[0] + G[3] * Ai[2], G[2] * Ai[1] + G[3] * Ai[3]};
        for (int k = 0; k < 4; ++k) SAG[k] -= GA[k];
        // A = A * bb + SumPrevAG * aa
        for (int k = 0; k < 4; ++k) {
          double t1 = Ai[k] * bb;
          double t2 = SAG[k] * aa;
          Ai[k] = t1 + t2;
        }
      }
      y[2 * i] = theta[0];
      y[2 * i + 1] = theta[1];
      for (int k = 0; k < 4; ++k) A[4 * i + k] = Ai[k];
      if (cnt_ray) cnt_ray[i] = c;
    }
  }

  // rayshooterInternal, FullRange/internal_rayshooter_multi.cpp:53-118
  void shoot(size_t n, const double* x, const double* zs_ray, double* y, double* A,
             int nthreads, Counters* cnt_ray) const {
    if (plane_redshifts.size() == 1) {
      for (size_t i = 0; i < n; ++i) {
        y[2 * i] = x[2 * i];
        y[2 * i + 1] = x[2 * i + 1];
        A[4 * i] = 1; A[4 * i + 1] = 0; A[4 * i + 2] = 0; A[4 * i + 3] = 1;
      }
      return;
    }
    if (n == 0) return;
    double source_z = plane_redshifts.back();
    if (nthreads < 1) nthreads = 1;
    size_t chunk;
    do {
      chunk = n / (size_t)nthreads;
      if (chunk == 0) nthreads /= 2;
    } while (chunk == 0);
    std::vector<std::thread> thr;
    for (int t = 0; t < nthreads; ++t) {
      size_t size = chunk;
      if (t == nthreads - 1) size = n - (size_t)(nthreads - 1) * chunk;
      size_t start = (size_t)t * chunk;
      thr.emplace_back([=]() {
        compute_points(start, start + size, x, zs_ray, source_z, y, A, cnt_ray);
      });
    }
    for (auto& t : thr) t.join();
  }
};

// ------------------------------------------------- smoothing sizes (k-NN)
// LensHaloParticles::calculate_smoothing (include/particle_halo.h:683-730):
// Size of particle i = TreeSimple::NNDistance(x_i, Nsmooth + 1)
// (include/simpleTree.h:483-620) over a 3-D median-cut tree with buckets of
// 2 Nsmooth (simpleTree.h:264-278, 842-960): the distance to the
// (Nsmooth+1)-th nearest particle, the particle itself included.  Squared
// distances are summed per dimension from pow(x_p[j] - x_i[j], 2) with the
// difference in the coordinate type T (simpleTree.h:518-536).
template <typename T>
struct KdTree {
  struct Node {
    double p1[3], p2[3];
    int level, first, n;      // particles idx[first, first + n)
    int child1 = -1, child2 = -1, parent = -1;
  };
  const T* x[3];
  std::vector<int> idx;
  std::vector<Node> nodes;
  int bucket;

  KdTree(const T* xs, const T* ys, const T* zs, size_t n, int bucket_) : bucket(bucket_) {
    x[0] = xs; x[1] = ys; x[2] = zs;
    idx.resize(n);
    for (size_t i = 0; i < n; ++i) idx[i] = (int)i;
    Node root;
    for (int j = 0; j < 3; ++j) {
      root.p1[j] = root.p2[j] = x[j][0];
      for (size_t i = 0; i < n; ++i) {
        root.p1[j] = std::min(root.p1[j], (double)x[j][i]);
        root.p2[j] = std::max(root.p2[j], (double)x[j][i]);
      }
    }
    root.level = 0; root.first = 0; root.n = (int)n;
    nodes.push_back(root);
    build(0);
  }
  // _BuildTreeNB with median_cut: sort by the cut dimension, cut at n/2
  // (simpleTree.h:882-960)
  void build(int b) {
    if (nodes[b].n <= bucket) return;
    const int d = nodes[b].level % 3;
    const int first = nodes[b].first, n = nodes[b].n;
    std::stable_sort(idx.begin() + first, idx.begin() + first + n,
                     [&](int a, int c) { return x[d][a] < x[d][c]; });
    const int cut = n / 2;
    const double xc = x[d][idx[first + cut]];
    Node c1 = nodes[b], c2 = nodes[b];
    c1.p2[d] = xc; c2.p1[d] = xc;
    c1.level = c2.level = nodes[b].level + 1;
    c1.first = first; c1.n = cut;
    c2.first = first + cut; c2.n = n - cut;
    c1.parent = c2.parent = b;
    if (c1.n > 0) { nodes[b].child1 = (int)nodes.size(); nodes.push_back(c1); build(nodes[b].child1); }
    if (c2.n > 0) { nodes[b].child2 = (int)nodes.size(); nodes.push_back(c2); build(nodes[b].child2); }
  }
  double r2of(int p, const T* q) const {
    double r2 = 0;
    for (int j = 0; j < 3; ++j) r2 += std::pow(x[j][p] - q[j], 2);
    return r2;
  }
  // NNDistance (simpleTree.h:483-620)
  double nn_distance(const T* q, int k) const {
    int it = 0;                                   // _findleaf (simpleTree.h:625-640)
    while (nodes[it].child1 >= 0 || nodes[it].child2 >= 0) {
      const int d = nodes[it].level % 3;
      const int c1 = nodes[it].child1, c2 = nodes[it].child2;
      if (c1 >= 0 && nodes[c1].p2[d] > (double)q[d]) it = c1;
      else if (c2 >= 0) it = c2;
      else it = c1;
    }
    while (nodes[it].n < k) it = nodes[it].parent;
    const int pass = it;
    std::priority_queue<double> heap;
    const Node& P = nodes[pass];
    for (int i = 0; i < P.n; ++i) {
      const double r2 = r2of(idx[P.first + i], q);
      if (i < k) heap.push(r2);
      else if (r2 < heap.top()) { heap.push(r2); heap.pop(); }
    }
    double bestr = std::sqrt(heap.top());
    // walk from the top, skipping the pass subtree; a box is entered if it
    // overlaps [q - bestr, q + bestr] along the dimension it was cut in
    std::vector<int> stack;
    stack.push_back(0);
    while (!stack.empty()) {
      const int b = stack.back();
      stack.pop_back();
      if (b == pass) continue;
      const Node& B = nodes[b];
      if (b != 0) {
        const int d = (B.level - 1) % 3;
        if (!(((double)q[d] + bestr) > B.p1[d] && ((double)q[d] - bestr) < B.p2[d])) continue;
      }
      if (B.child1 < 0 && B.child2 < 0) {
        for (int i = 0; i < B.n; ++i) {
          const double r2 = r2of(idx[B.first + i], q);
          if (r2 < bestr * bestr) {
            heap.push(r2);
            heap.pop();
            bestr = std::sqrt(heap.top());
          }
        }
      } else {
        if (B.child2 >= 0) stack.push_back(B.child2);
        if (B.child1 >= 0) stack.push_back(B.child1);
      }
    }
    return bestr;
  }
};

// The definition itself: the k-th smallest squared distance over all
// particles, by selection.
template <typename T>
double nn_distance_brute(const T* const* x, size_t n, const T* q, int k,
                         std::vector<double>& buf) {
  buf.resize(n);
  for (size_t p = 0; p < n; ++p) {
    double r2 = 0;
    for (int j = 0; j < 3; ++j) r2 += std::pow(x[j][p] - q[j], 2);
    buf[p] = r2;
  }
  std::nth_element(buf.begin(), buf.begin() + (k - 1), buf.end());
  return std::sqrt(buf[k - 1]);
}

template <typename T>
int smoothing(const T* xs, const T* ys, const T* zs, size_t n, int nsmooth, int brute,
              float* out, int nthreads, const int64_t* which, size_t nwhich) {
  const int k = nsmooth + 1;
  if (n <= (size_t)k) return 1;           // "number of neighbors > total number of particles"
  std::unique_ptr<KdTree<T>> tree;
  if (!brute) tree.reset(new KdTree<T>(xs, ys, zs, n, 2 * nsmooth));
  const T* x[3] = {xs, ys, zs};
  const size_t nq = which ? nwhich : n;
  std::vector<std::thread> th;
  std::atomic<size_t> next(0);
  for (int t = 0; t < std::max(1, nthreads); ++t) {
    th.emplace_back([&]() {
      std::vector<double> buf;
      for (;;) {
        const size_t i0 = next.fetch_add(256);
        if (i0 >= nq) break;
        for (size_t a = i0; a < std::min(nq, i0 + 256); ++a) {
          const size_t i = which ? (size_t)which[a] : a;
          const T q[3] = {xs[i], ys[i], zs[i]};
          out[a] = (float)(brute ? nn_distance_brute(x, n, q, k, buf)
                                 : tree->nn_distance(q, k));
        }
      }
    });
  }
  for (auto& t : th) t.join();
  return 0;
}

}  // namespace orc

// ====================================================================== C ABI
extern "C" {

struct orc_counters {
  uint64_t visits, far, leafp, bigtest, bspline;
};

double orc_coordist(double Om, double OL, double h, double z1, double z2) {
  orc::Cosmo c{Om, OL, h};
  return c.coorDist(z1, z2);
}

// Builds a TreeQuadParticles over n particles. Positions are given as double
// (float positions promote exactly). Returns null and sets *err on failure.
void* orc_tree_build_range(const double* x, const double* y, const float* m, const float* h,
                           size_t n, double inv_area, int bucket, double theta, double max_range,
                           int* err);
void* orc_tree_build(const double* x, const double* y, const float* m, const float* h,
                     size_t n, double inv_area, int bucket, double theta, int* err) {
  return orc_tree_build_range(x, y, m, h, n, inv_area, bucket, theta, -1.0, err);
}

// the same with TreeQuadParticles' maximum_range (quadTree.h:266, 370-376)
void* orc_tree_build_range(const double* x, const double* y, const float* m, const float* h,
                           size_t n, double inv_area, int bucket, double theta, double max_range,
                           int* err) {
  auto t = new orc::TreeQuad();
  t->max_range = max_range;
  t->px.assign(x, x + n);
  t->py.assign(y, y + n);
  t->pm.assign(m, m + n);
  t->ph.assign(h, h + n);
  t->N = n;
  t->inv_area = inv_area;
  t->Nbucket = bucket;
  t->force_theta = theta;
  t->index.resize(n);
  for (size_t i = 0; i < n; ++i) t->index[i] = i;
  if (err) *err = 0;
  try {
    if (n > 0) {
      t->build();
      t->calc_moments();
    }
  } catch (const std::exception&) {
    if (err) *err = 1;
    delete t;
    return nullptr;
  }
  return t;
}

void orc_tree_free(void* t) { delete static_cast<orc::TreeQuad*>(t); }

size_t orc_tree_nnodes(void* tp) { return static_cast<orc::TreeQuad*>(tp)->nodes.size(); }

// Node export in depth-first pre-order. Per node:
//   dbl[13*i + ...] = p1x, p1y, p2x, p2y, cx, cy, mass, q0, q1, q2, r2crit, 0, 0
//   ints[6*i + ...] = level, nparticles, nbig, is_leaf, first_child, parent-less flag
void orc_tree_export(void* tp, double* dbl, int64_t* ints) {
  auto* t = static_cast<orc::TreeQuad*>(tp);
  // depth-first pre-order, child0 first (walkTree_recur's visiting order)
  std::vector<int> order, stack{0};
  while (!stack.empty()) {
    int k = stack.back();
    stack.pop_back();
    order.push_back(k);
    const auto& b = t->nodes[k];
    for (int c = 3; c >= 0; --c)
      if (b.child[c] >= 0) stack.push_back(b.child[c]);
  }
  for (size_t i = 0; i < order.size(); ++i) {
    const auto& b = t->nodes[order[i]];
    double* d = dbl + 13 * i;
    d[0] = b.p1[0]; d[1] = b.p1[1]; d[2] = b.p2[0]; d[3] = b.p2[1];
    d[4] = b.center[0]; d[5] = b.center[1]; d[6] = b.mass;
    d[7] = b.quad[0]; d[8] = b.quad[1]; d[9] = b.quad[2]; d[10] = b.r2crit;
    d[11] = 0; d[12] = 0;
    int64_t* q = ints + 6 * i;
    q[0] = b.level; q[1] = (int64_t)b.nparticles; q[2] = (int64_t)b.big.size();
    q[3] = b.leaf() ? 1 : 0; q[4] = b.child[0]; q[5] = 0;
  }
}

// Raw force2D_recur outputs (tree sign convention) at n rays (x interleaved).
// out: 5 per ray = alpha0, alpha1, kappa, gamma0, gamma1.
void orc_tree_force(void* tp, const double* rays, size_t n, double* out,
                    orc_counters* cnt, int nthreads) {
  auto* t = static_cast<orc::TreeQuad*>(tp);
  auto work = [&](size_t s, size_t e) {
    for (size_t i = s; i < e; ++i) {
      double a[2], k, g[3];
      orc::Counters c;
      t->force2D(rays + 2 * i, a, &k, g, cnt ? &c : nullptr);
      out[5 * i] = a[0]; out[5 * i + 1] = a[1]; out[5 * i + 2] = k;
      out[5 * i + 3] = g[0]; out[5 * i + 4] = g[1];
      if (cnt) {
        cnt[i].visits = c.visits; cnt[i].far = c.far; cnt[i].leafp = c.leafp;
        cnt[i].bigtest = c.bigtest; cnt[i].bspline = c.bspline;
      }
    }
  };
  if (nthreads <= 1 || n < 64) { work(0, n); return; }
  std::vector<std::thread> th;
  size_t chunk = (n + nthreads - 1) / nthreads;
  for (int k = 0; k < nthreads; ++k) {
    size_t s = k * chunk, e = std::min(n, s + chunk);
    if (s >= e) break;
    th.emplace_back(work, s, e);
  }
  for (auto& x : th) x.join();
}

// The same smoothed-particle kernel summed over every particle (no tree), in
// the tree's sign convention. out: 5 per ray.
void orc_direct_plane(const double* px, const double* py, const float* pm, const float* ph,
                      size_t np, double inv_area, const double* rays, size_t n, double* out,
                      int nthreads) {
  auto work = [&](size_t s, size_t e) {
    for (size_t i = s; i < e; ++i) {
      double alpha[2] = {0, 0}, kappa = 0, gamma[2] = {0, 0};
      const double* ray = rays + 2 * i;
      for (size_t p = 0; p < np; ++p) {
        double d[2] = {px[p] - ray[0], py[p] - ray[1]};
        double rcm2 = d[0] * d[0] + d[1] * d[1];
        if (rcm2 < 1e-20) rcm2 = 1e-20;
        double mass = pm[p];
        double prefac = mass / rcm2 / orc::PI;
        alpha[0] += (inv_area * mass - prefac) * d[0];
        alpha[1] += (inv_area * mass - prefac) * d[1];
        double tmp = -2.0 * prefac / rcm2;
        gamma[0] += 0.5 * (d[0] * d[0] - d[1] * d[1]) * tmp;
        gamma[1] += d[0] * d[1] * tmp;
        kappa -= mass * inv_area;
        double size = ph[p];
        if (rcm2 < 4 * size * size)
          orc::TreeQuad::b_spline(d, std::sqrt(rcm2), mass, size, alpha, &kappa, gamma);
      }
      out[5 * i] = alpha[0]; out[5 * i + 1] = alpha[1]; out[5 * i + 2] = kappa;
      out[5 * i + 3] = gamma[0]; out[5 * i + 4] = gamma[1];
    }
  };
  if (nthreads <= 1 || n < 16) { work(0, n); return; }
  std::vector<std::thread> th;
  size_t chunk = (n + nthreads - 1) / nthreads;
  for (int k = 0; k < nthreads; ++k) {
    size_t s = k * chunk, e = std::min(n, s + chunk);
    if (s >= e) break;
    th.emplace_back(work, s, e);
  }
  for (auto& x : th) x.join();
}

// B-spline correction alone (for closed-form checks). out: alpha0, alpha1, kappa, gamma0, gamma1
void orc_bspline(const double* xcm, double r, double mass, double size, double* out) {
  double a[2] = {0, 0}, k = 0, g[2] = {0, 0};
  orc::TreeQuad::b_spline(xcm, r, mass, size, a, &k, g);
  out[0] = a[0]; out[1] = a[1]; out[2] = k; out[3] = g[0]; out[4] = g[1];
}

void* orc_lens_create(double zsource, double Om, double OL, double h) {
  return new orc::Lens(zsource, Om, OL, h);
}
void orc_lens_free(void* l) { delete static_cast<orc::Lens*>(l); }

// Moves a built tree into the lens as one LensHaloParticles at float redshift z.
void orc_lens_add_halo(void* lp, void* tree, float z) {
  auto* L = static_cast<orc::Lens*>(lp);
  auto h = std::make_unique<orc::Halo>();
  h->tree.reset(static_cast<orc::TreeQuad*>(tree));
  h->z = z;
  L->moveinMainHalo(std::move(h));
}

// LensHaloHalos<LensHaloNFW>: a halo tree of n NFW halos (float mass, Rsize,
// concentration; physical positions), inv_area and theta_force as floats.
void* orc_halotree_build(const double* x, const double* y, const float* m, const float* rsize,
                         const float* conc, size_t n, float inv_area, float theta, int* status) {
  try {
    std::vector<orc::NfwHalo> h;
    for (size_t i = 0; i < n; ++i) h.emplace_back(x[i], y[i], m[i], rsize[i], conc[i]);
    *status = 0;
    return new orc::HaloTree(std::move(h), inv_area, theta);
  } catch (const std::exception&) {
    *status = 1;
    return nullptr;
  }
}
// A tree of mixed halos: type[i] 0 = NFW (param = concentration), 1 = power
// law (param = beta); inv_area, theta, screening and the buffer as given.
void* orc_halotree_build_mixed(const double* x, const double* y, const float* m,
                               const float* rsize, const int* type, const double* param, size_t n,
                               double inv_area, double theta, double inv_screening, int periodic,
                               int* status) {
  try {
    std::vector<orc::NfwHalo> h;
    for (size_t i = 0; i < n; ++i) {
      if (type[i] == 1) h.push_back(orc::NfwHalo::power(x[i], y[i], m[i], rsize[i], param[i]));
      else h.emplace_back(x[i], y[i], m[i], rsize[i], (float)param[i]);
    }
    *status = 0;
    return new orc::HaloTree(std::move(h), inv_area, theta, inv_screening, periodic != 0);
  } catch (const std::exception&) {
    *status = 1;
    return nullptr;
  }
}
// one power-law halo's force_halo at positions relative to it (5 per point)
void orc_powerlaw_force(float m, float rsize, double beta, const double* xcm, size_t n,
                        double* out) {
  const orc::NfwHalo h = orc::NfwHalo::power(0, 0, m, rsize, beta);
  for (size_t i = 0; i < n; ++i) {
    double a[2] = {0, 0}, k = 0, g[3] = {0, 0, 0};
    h.force_halo(a, &k, g, xcm + 2 * i);
    out[5 * i] = a[0]; out[5 * i + 1] = a[1]; out[5 * i + 2] = k;
    out[5 * i + 3] = g[0]; out[5 * i + 4] = g[1];
  }
}

// LensPlaneTree's tree: inv_area (sigma_background), theta, screening and
// the periodic buffer as doubles
void* orc_halotree_build_field(const double* x, const double* y, const float* m,
                               const float* rsize, const float* conc, size_t n, double inv_area,
                               double theta, double inv_screening, int periodic, int* status) {
  try {
    std::vector<orc::NfwHalo> h;
    for (size_t i = 0; i < n; ++i) h.emplace_back(x[i], y[i], m[i], rsize[i], conc[i]);
    *status = 0;
    return new orc::HaloTree(std::move(h), inv_area, theta, inv_screening, periodic != 0);
  } catch (const std::exception&) {
    *status = 1;
    return nullptr;
  }
}
void orc_halotree_free(void* t) { delete static_cast<orc::HaloTree*>(t); }
// force2D (direct = 0) or the sum over every halo (direct = 1), 5 per ray
void orc_halotree_force(void* tp, const double* rays, size_t n, int direct, double* out) {
  const auto* t = static_cast<orc::HaloTree*>(tp);
  for (size_t i = 0; i < n; ++i) {
    double a[2], k, g[3];
    if (direct) t->direct(rays + 2 * i, a, &k, g);
    else t->force2D(rays + 2 * i, a, &k, g);
    out[5 * i] = a[0]; out[5 * i + 1] = a[1]; out[5 * i + 2] = k;
    out[5 * i + 3] = g[0]; out[5 * i + 4] = g[1];
  }
}
// one NFW halo's force_halo at positions relative to it (5 per point)
void orc_nfw_force(float m, float rsize, float conc, const double* xcm, size_t n, double* out) {
  orc::NfwHalo h(0, 0, m, rsize, conc);
  for (size_t i = 0; i < n; ++i) {
    double a[2] = {0, 0}, k = 0, g[3] = {0, 0, 0};
    h.force_halo(a, &k, g, xcm + 2 * i);
    out[5 * i] = a[0]; out[5 * i + 1] = a[1]; out[5 * i + 2] = k;
    out[5 * i + 3] = g[0]; out[5 * i + 4] = g[1];
  }
}
// the profile functions and their table interpolation (which: 0 g, 1 f, 2 g2)
double orc_nfw_function(int which, double x) {
  return which == 0 ? orc::NfwProfile::gfunction(x)
         : which == 1 ? orc::NfwProfile::ffunction(x) : orc::NfwProfile::g2function(x);
}
double orc_nfw_interpolate(int which, double x) { return orc::NfwProfile::get().interpolate(which, x); }
// Moves a halo tree into the lens as one LensHaloHalos at float redshift z.
void orc_lens_add_halotree(void* lp, void* tree, float z) {
  auto* L = static_cast<orc::Lens*>(lp);
  auto h = std::make_unique<orc::Halo>();
  h->htree.reset(static_cast<orc::HaloTree*>(tree));
  h->z = z;
  L->moveinMainHalo(std::move(h));
}

// moveinMainHalo(halo, addplanes=false): particle tree (kind 0) or halo tree
void orc_lens_add_nearest(void* lp, void* tree, float z, int halotree) {
  auto* L = static_cast<orc::Lens*>(lp);
  auto h = std::make_unique<orc::Halo>();
  if (halotree) h->htree.reset(static_cast<orc::HaloTree*>(tree));
  else h->tree.reset(static_cast<orc::TreeQuad*>(tree));
  h->z = z;
  L->moveinMainHaloNearest(std::move(h));
}

// a LensPlaneTree field plane (the lens takes the tree) at redshift z
void orc_lens_add_field_plane(void* lp, void* tree, double z) {
  auto* L = static_cast<orc::Lens*>(lp);
  auto h = std::make_unique<orc::Halo>();
  h->htree.reset(static_cast<orc::HaloTree*>(tree));
  h->z = (float)z;
  L->addFieldPlane(std::move(h), z);
}

int orc_lens_nplanes(void* lp) { return (int)static_cast<orc::Lens*>(lp)->planes.size(); }

// Copies plane_redshifts, Dl and dDl (each nplanes + 1 long).
void orc_lens_tables(void* lp, double* z, double* Dl, double* dDl, double* charge) {
  auto* L = static_cast<orc::Lens*>(lp);
  for (size_t i = 0; i < L->plane_redshifts.size(); ++i) {
    z[i] = L->plane_redshifts[i];
    Dl[i] = L->Dl[i];
    dDl[i] = L->dDl[i];
  }
  *charge = L->charge;
}

// rayshooterInternal over n rays: x (2n) -> y (2n), A (4n). zs: per-ray source
// redshifts (LinkedPoint/RAY overloads) or null (Point overload).
void orc_lens_shoot(void* lp, size_t n, const double* x, const double* zs, double* y,
                    double* A, int nthreads, orc_counters* cnt) {
  auto* L = static_cast<orc::Lens*>(lp);
  std::vector<orc::Counters> c;
  if (cnt) c.resize(n);
  L->shoot(n, x, zs, y, A, nthreads, cnt ? c.data() : nullptr);
  if (cnt)
    for (size_t i = 0; i < n; ++i) {
      cnt[i].visits = c[i].visits; cnt[i].far = c[i].far; cnt[i].leafp = c[i].leafp;
      cnt[i].bigtest = c[i].bigtest; cnt[i].bspline = c[i].bspline;
    }
}

// GridMap square-grid ray positions (TreeCode_link/gridmap.cpp:293-321 via
// Utilities::PositionFromIndex, TreeCode_link/utilities.cpp:149-158).
void orc_gridmap_positions(size_t N, double c0, double c1, double range, double* x) {
  const double center[2] = {c0, c1};
  for (size_t i = 0; i < N * N; ++i) {
    if (N == 1) { x[2 * i] = center[0]; x[2 * i + 1] = center[1]; continue; }
    x[2 * i] = center[0] + range * (1.0 * (i % N) / (N - 1) - 0.5);
    x[2 * i + 1] = center[1] + range * (1.0 * (i / N) / (N - 1) - 0.5);
  }
}

// Smoothing sizes: out[a] for particle which[a] (all particles when which is
// null).  brute = 1 uses the selection definition instead of the reference's
// tree search.  Returns 1 when n <= nsmooth + 1 (the reference exits).
int orc_smoothing_f64(const double* x, const double* y, const double* z, size_t n, int nsmooth,
                      int brute, float* out, int nthreads, const int64_t* which, size_t nwhich) {
  return orc::smoothing<double>(x, y, z, n, nsmooth, brute, out, nthreads, which, nwhich);
}
int orc_smoothing_f32(const float* x, const float* y, const float* z, size_t n, int nsmooth,
                      int brute, float* out, int nthreads, const int64_t* which, size_t nwhich) {
  return orc::smoothing<float>(x, y, z, n, nsmooth, brute, out, nthreads, which, nwhich);
}

}  // extern "C"
