// slsim_internal.h -- device data layout and host-side objects of the ray shooter.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/slsim_amd.h"

namespace slsim {

// Tree levels encoded in the 64-bit Morton path key (2 bits per level).
constexpr int kMaxLevels = 32;
// Per-wavefront traversal stack entries: depth-first with four children per
// opened cell needs at most 3 * (kMaxLevels + 1) + 1.
constexpr int kStackDepth = 104;
constexpr int kWave = 64;
constexpr int kWavesPerBlock = 4;
// Counters per slot in counting runs: 5 per-ray work counts, then the
// bundle's wave-level cell evaluations and leaf-particle iterations (lane 0).
constexpr int kCountStride = 7;

constexpr double kPI = 3.14159265358979323846;   // M_PI (include/standard.h:38-40)
constexpr double kGrav = 4.7788e-20;              // G/c^2 in Mpc/Msun (standard.h:43)
constexpr double kMinPlaneDist = 1e-6;            // MultiPlane/lens.cpp:15

// Cell record of the FP64 walk (96 B, read with scalar loads).
//  leaf:     start/count = particle range in the tree-ordered particle array
//  internal: start/count = range of the cell's internal big-particle list,
//            child = index of the first of four contiguous children
//            (child0..child3 in the reference's order), cmask = non-empty children
struct alignas(16) NodeD {
  double cx, cy;        // centre of mass
  double mass;
  double q0, q1, q2;    // quadrupole about the centre of mass
  double r2crit;        // (1.15470 rcom / theta)^2
  double near2;         // 5.83 * box^2
  double box;           // cell width
  int32_t child;        // -1 for a leaf
  int32_t start;
  int32_t count;
  int32_t cmask;
  double T;             // max(r2crit, near2): r^2 < T opens the cell
};
static_assert(sizeof(NodeD) == 96, "NodeD layout");

// Cell record of the FP32 walk (64 B).  The opening test stays FP64 so that
// both walks open exactly the same cells for a ray: r^2 < T with T =
// max(r2crit, near2) (r^2 < r2crit or r^2 < near2), near2 gating the big
// particles.
struct alignas(16) NodeF {
  double cx, cy;
  double T, near2;
  float mass, q0, q1, q2;   // q0..q2: 4 times the quadrupole
  float flip;               // the cell's flip error (walk units), tree_build.hip
  int32_t child;
  int32_t start;
  int32_t count_cmask;  // count | cmask << 28
};
static_assert(sizeof(NodeF) == 64, "NodeF layout");

// Particle in tree order, read by both walks.  In the leaf particle array
// the sign bit of h carries the leaf's big-particle flag (h < box,
// include/quadTree.h:379-396): set = not listed; |h| is the size.  Internal
// big-particle lists store h as is.
struct PartD {
  double x, y;
  float m, h;
};
static_assert(sizeof(PartD) == 24, "PartD layout");

// Particle of a leaf for the FP32 walk: its offset from the leaf's centre of
// mass (FP32; the walk adds the FP32 offset of the centre from the ray), its
// mass, and 4 h^2 (1 + 2^-10) with the leaf's big-particle flag in the sign
// bit as PartD carries it (set = not listed).  Same order and indices as
// PartD.
struct PartF {
  float lx, ly;
  float m;
  float hh4;
};
static_assert(sizeof(PartF) == 16, "PartF layout");

// Halo-of-halos planes (LensHaloHalos<LensHaloNFW>, include/particle_halo.h:
// 1092-1227): cells of the TreeQuadHalos quadtree (include/quadTreeHalos.h),
// children of a cell contiguous (child0..child3 in the reference's order).
struct alignas(16) HNode {
  double cx, cy;        // centre of mass
  double mass;
  double q0, q1, q2;    // quadrupole about the centre of mass
  double r2crit;        // (1.15470 rcom / theta)^2
  int32_t child;        // first child, -1 for a leaf
  int32_t cmask;        // children with halos
  int32_t big_off;      // the cell's big-halo list in the tree's index array
  int32_t nbig;
  double pad;
};
static_assert(sizeof(HNode) == 80, "HNode layout");

// One symmetric LensHaloNFW (fratio 1) as the reference stores it: float
// mass, Rmax and rscale, gmax = g(Rsize/rscale) from the tables; with the
// reciprocals the walk multiplies by.
// A symmetric halo of a halo tree.  NFW: gmax = g(xmax) of the tables.  A
// power law (kind 1, LensHaloPowerLaw): gmax holds beta, rscale = 1, xmax =
// rmax = Rsize.
struct HaloNFW {
  double x, y;          // physical position (LensHalo::getX)
  double gmax;
  double inv_gmax;      // 1 / gmax
  double inv_rscale;    // 1 / rscale
  double mass_pi;       // mass / pi
  float mass, rmax, rscale, pad;   // pad: the kind, 0 NFW, 1 power law
};
static_assert(sizeof(HaloNFW) == 64, "HaloNFW layout");

// LensHaloNFW's tables (MultiPlane/lens_halos.cpp:352-353, 482-540):
// NTABLE entries over [0, maxrm), entry i = (x_i, g(x_i), f(x_i), g2(x_i))
// and the slopes (t_{i+1} - t_i) / (x_{i+1} - x_i) of InterpolateFromTable.
constexpr int kNfwTable = 10000;
constexpr double kNfwMaxrm = 100.0;
struct NfwEntry {
  double x, g, f, g2;
  double sg, sf, sg2, pad;
};

constexpr int32_t kTreeParticles = 0;
constexpr int32_t kTreeHalos = 1;

struct TreeDesc {
  const NodeD* nodesD;
  const NodeF* nodesF;
  const PartD* partsD;
  const PartD* bigD;
  const PartF* partsF;
  double inv_area;
  double mass, comx, comy;   // total mass and centre of mass (the mass sheet's terms)
  double dipx, dipy;         // sum of m x over the particles
  const float2* sheetd;      // masses of both signs (else null): per cell
                             // pi inv_area (M_c c - D_c), see tree_build.hip
  int32_t nnodes;
  int32_t nparts;
  // halo-of-halos trees (kind == kTreeHalos)
  int32_t kind;
  int32_t nhalos;
  const HNode* hnodes;
  const int32_t* hbig;       // big-halo lists (halo indices)
  const HaloNFW* halos;
  const NfwEntry* nfw;       // kNfwTable entries
  // LensPlaneTree's walk options: screening exp(-r^2 hs2) on far cells, and
  // the periodic buffer (nine walks at the ray shifted by the halos' extents)
  double hs2, hxl, hyl;
  int32_t hperiodic;
  int32_t pad2;
};

// Per-slot ray state, structure of arrays (slot = bundle * 64 + lane).
struct RayState {
  double* th0; double* th1;     // theta (radians) on the current plane
  double* sa0; double* sa1;     // SumPrevAlphas
  double* A0; double* A1; double* A2; double* A3;
  double* G0; double* G1; double* G2; double* G3;   // SumPrevAG
  double* aal; double* bbl;     // (aa, bb) of the last plane before the source
  int32_t* jmax;                // planes in front of the source; 0 = inactive slot
  // precision form only: bounds on how far the FP32-walk ray can be from the
  // FP64 ray (eth: theta on the current plane, eS: SumPrevAlphas, radians),
  // and the risk: the summed flip errors (radians at the source) of the
  // opening tests that fell within that distance of their thresholds
  float* eth; float* eS; float* risk;
};
constexpr size_t kStateBytesPerSlot = 14 * sizeof(double) + sizeof(int32_t) + 3 * sizeof(float);

// The 15 arrays of a ray state laid out back to back from `d` (n slots each).
__host__ __device__ inline RayState state_view(double* d, int64_t n) {
  RayState s;
  s.th0 = d;         s.th1 = d + n;
  s.sa0 = d + 2 * n; s.sa1 = d + 3 * n;
  s.A0 = d + 4 * n;  s.A1 = d + 5 * n;  s.A2 = d + 6 * n;  s.A3 = d + 7 * n;
  s.G0 = d + 8 * n;  s.G1 = d + 9 * n;  s.G2 = d + 10 * n; s.G3 = d + 11 * n;
  s.aal = d + 12 * n; s.bbl = d + 13 * n;
  s.jmax = reinterpret_cast<int32_t*>(d + 14 * n);
  s.eth = reinterpret_cast<float*>(s.jmax + n);
  s.eS = s.eth + n;
  s.risk = s.eS + n;
  return s;
}

// One plane step of one ray job.
struct StepJob {
  double* sbase;                // ray state, state_view(sbase, nslots)
  int32_t nslots;
  int32_t j;                    // plane index of this step
  int32_t ntrees;               // particle trees, trees[0, ntrees)
  int32_t nhtrees;              // halo-of-halos trees, trees[ntrees, ntrees + nhtrees)
  int32_t hkind;                // 0 no halo trees, 1 NFW members only, 2 power laws too
  int32_t hpad;
  const TreeDesc* trees;        // the plane's halos (device array)
  double Dphys;                 // Dl[j] / (1 + z_j)
  double Gfac;                  // charge * Dl[j] / (1 + z_j)
  double aa, bb;                // dDl[j+1]/Dl[j+1], Dl[j]/Dl[j+1]
  double charge;
  uint32_t* counters;           // kCountStride per slot or null
  // precision form's guard: the plane's deflection scale (walk units,
  // sum over its particle trees of M / (pi R), R half the root box) and the
  // FP32 walk's error per unit of deflection
  double ascale;
  double eps;
};

// The precision form's guard.  The FP32 walk's deflection is taken to be
// off by at most kGuardEps times the larger of the ray's deflection and the
// plane's deflection scale (measured: below 1e-7 of the rms per plane).  A
// ray is shot again in FP64 when its risk, times kGuardSafety, exceeds
// kGuardTol of max(|alpha|, 1e-3 rms) (the error measure of the precision
// form's bound), or when |alpha| is below kGuardSmallAlpha of the call's
// rms: the FP32 walk's absolute error is a fraction of the field's
// deflections, so its relative error grows where alpha vanishes.
constexpr double kGuardEps = 16.0 / 16777216.0;
constexpr double kGuardTol = 1e-5;
constexpr double kGuardSafety = 2.0;
constexpr double kGuardSmallAlpha = 0.01;
// the guard's FP64 shoot puts fewer than 64 rays in a bundle while that
// still leaves at most this many bundles (about the waves the GPU holds)
constexpr int64_t kGuardWaves = 8192;
// up to this many re-shot rays: one wavefront per ray
constexpr int64_t kGuardLoneRays = 8192;

// A call's slots for the guard kernels: x == null: the tiles of an N x N
// GridMap; else slot s < n holds ray perm[s] (perm == null: ray s) of x,
// with per-ray source redshifts zs (or null).
struct GuardMap {
  int64_t N, tiles_x;
  int32_t tw, pad;
  double c0, c1, range;
  const uint32_t* perm;
  const double* x;
  const double* zs;
  int64_t n;
};

struct WorkItem {
  uint32_t job;
  uint32_t bundle;
};

// A bundle taken through all of its plane steps by one wavefront: jobs
// job0 .. job0 + njobs - 1 in order (a batch's fused schedule).
struct FusedItem {
  uint32_t job0;
  uint32_t bundle;
  uint32_t njobs;
  uint32_t pad;
};

// One GridMap request of a batch: its grid, source and state.
struct GridReq {
  int64_t N, tiles_x, slot_off, nslots;
  double c0, c1, range;
  double aal, bbl;
  int32_t jmax, tw;             // tw: log2 of the bundle tile width (tile 2^tw x 2^(6-tw))
  double* y;
  double* A;
  RayState st;
};

// ------------------------------------------------------------- host objects
struct Plane;

// Error helpers
void set_error(const std::string& msg);
#define SLSIM_HIP_CHECK(expr)                                                   \
  do {                                                                          \
    hipError_t e_ = (expr);                                                     \
    if (e_ != hipSuccess) {                                                     \
      ::slsim::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));   \
      return SLSIM_ERR_HIP;                                                     \
    }                                                                           \
  } while (0)



// GridMap ray i of an N x N grid (TreeCode_link/gridmap.cpp:293-321,
// utilities.cpp:149-158), computed without contraction as the reference.
__host__ __device__ inline void grid_position(int64_t i, int64_t N, double c0, double c1,
                                              double range, double* x0, double* x1) {
#pragma clang fp contract(off)
  *x0 = c0;
  *x1 = c1;
  if (N > 1) {
    *x0 = c0 + range * (1.0 * (double)(i % N) / (double)(N - 1) - 0.5);
    *x1 = c1 + range * (1.0 * (double)(i / N) / (double)(N - 1) - 0.5);
  }
}

// A GridMap's points: the N x N grid of GridMap(lens, N, center, range)
// (grid_position) or the Nx x Ny grid of GridMap(lens, Nx, center, rangeX,
// rangeY) (TreeCode_link/gridmap.cpp:26-91: point ii + jj Nx at the lower-left
// corner plus (ii, jj) rangeX / (Nx - 1)).
struct GridGeom {
  int64_t nx, ny;
  double c0, c1, rx, ry;
  int32_t rect;
  int32_t pad;
};
inline GridGeom square_geom(int64_t N, double c0, double c1, double range) {
  return GridGeom{N, N, c0, c1, range, range, 0, 0};
}
__host__ __device__ inline void geom_position(const GridGeom& g, int64_t k, double* x0,
                                              double* x1) {
#pragma clang fp contract(off)
  if (!g.rect) {
    grid_position(k, g.nx, g.c0, g.c1, g.rx, x0, x1);
    return;
  }
  const double xo = g.c0 - 0.5 * g.rx, yo = g.c1 - 0.5 * g.ry;
  const double res = g.rx / (double)(g.nx - 1);
  *x0 = xo + (double)(k % g.nx) * res;
  *x1 = yo + (double)(k / g.nx) * res;
}

}  // namespace slsim

struct slsim_ctx_s {
  int device = 0;
  size_t budget = 0;
  size_t used = 0;
  hipStream_t stream = nullptr;
  std::mutex mu;
  std::vector<std::pair<void*, size_t>> allocs;
  // freed scratch blocks kept for reuse (still counted in `used`)
  std::vector<std::pair<void*, size_t>> cache;
  int alloc(void** p, size_t bytes);
  int release(void* p);
  int scratch_get(void** p, size_t bytes);
  void scratch_put(void* p);
  void trim_cache();
};

struct slsim_plane_s {
  slsim_ctx ctx = nullptr;
  int64_t n = 0, nnodes = 0, nleaves = 0, nbig_internal = 0;
  int32_t depth = 0;
  double p1[2] = {0, 0}, p2[2] = {0, 0};
  double theta = 0.1, inv_area = 0.0;
  double mass = 0, com[2] = {0, 0};   // total mass and centre of mass (root cell)
  double dip[2] = {0, 0};             // sum of m x over the particles (signed masses)
  bool mixed_sign = false;            // masses of both signs
  float2* sheetd = nullptr;           // per cell: the sheet's partition term (mixed signs)
  int bucket = 20;
  double max_range = -1;   // > 0: cells far beyond it are not split
  double build_ms = 0;
  size_t bytes = 0;
  std::vector<int64_t> level_offset;   // BFS level starts (host copy)
  slsim::NodeD* nodesD = nullptr;
  slsim::NodeF* nodesF = nullptr;
  slsim::PartD* partsD = nullptr;
  slsim::PartD* bigD = nullptr;
  slsim::PartF* partsF = nullptr;
  // halo-of-halos plane (kind == kTreeHalos)
  int32_t kind = 0;
  slsim::HNode* hnodes = nullptr;
  int32_t* hbig = nullptr;
  slsim::HaloNFW* halos = nullptr;
  slsim::NfwEntry* nfw = nullptr;
  std::vector<slsim::HaloNFW> halos_host;
  double hs2 = 0, hxl = 0, hyl = 0;   // screening and the buffer's shifts (field planes)
  int hperiodic = 0;
  int hpowerlaw = 0;                  // some member is a power law
  // build products kept for export and direct sums
  double* box_p1 = nullptr;           // 2 per node (p1x, p1y)
  double* box_p2 = nullptr;           // 2 per node
  int32_t* nbig = nullptr;            // reference big-list length per node
  int32_t* npart = nullptr;           // particles per node
  slsim::TreeDesc desc() const;
};

struct slsim_lens_s {
  slsim_ctx ctx = nullptr;
  double Om = 0.3, OL = 0.7, h = 0.7;
  double zsource = 0;
  double charge = 0;
  bool toggle_source_plane = false;
  double zs_implant = 0;
  std::vector<std::vector<slsim_plane>> main_planes;   // halos per main plane
  std::vector<double> main_Dl, main_z;
  std::vector<slsim_plane> owned;
  // field planes (LensPlaneTree): one halo-of-halos tree each, double redshifts
  std::vector<slsim_plane> field_planes;
  std::vector<double> field_z, field_Dl;
  // combined (combinePlanes): the trees of plane j
  std::vector<std::vector<slsim_plane>> combined;
  std::vector<double> plane_redshifts, Dl, dDl;
  slsim::TreeDesc* trees_dev = nullptr;                // per plane: halos' descs
  std::vector<int> tree_off, tree_cnt, tree_hcnt;   // per plane: first desc, particle / halo trees
  std::vector<int> tree_hkind;                     // per plane: StepJob::hkind
  bool trees_dirty = true;
  double coorDist(double z1, double z2) const;
  void combinePlanes();
  void FindSourcePlane(double zs, long& jmax, double& Dls, double& Ds) const;
  int upload_trees();
};

// curves found on a grid: points of curve c are pts[2 off[c] .. 2 off[c + 1])
struct slsim_curves_s {
  std::vector<int64_t> off{0};
  std::vector<double> pts;
  std::vector<int32_t> types;
  std::vector<int32_t> hits;
  int hits_replaced = 0;
};

namespace slsim {

// Host-side fixed sizes of a shoot call: partial blocks of the ray bounding
// box, and plane steps held in a stack array (more planes use the heap,
// except in the no-allocation entry point).
constexpr int kMinmaxBlocks = 1024;
constexpr int kMaxPlanesFixed = 256;

// Device scratch of one call, returned to the context's cache on scope exit.
struct Scratch {
  static constexpr bool kMayAllocate = true;
  slsim_ctx ctx;
  std::vector<void*> ptrs;
  explicit Scratch(slsim_ctx c) : ctx(c) {}
  ~Scratch() {
    for (void* p : ptrs) ctx->scratch_put(p);
  }
  template <typename T>
  int get(T** p, size_t count) {
    void* v = nullptr;
    int rc = ctx->scratch_get(&v, std::max<size_t>(count, 1) * sizeof(T));
    if (rc) return rc;
    ptrs.push_back(v);
    *p = static_cast<T*>(v);
    return SLSIM_OK;
  }
};
struct FlatLCDM;
// kernel launchers (shoot.hip)
// halos: some job of the launch has halo-of-halos trees (selects the kernels
// that walk them)
void launch_plane_step(int precision, bool count, uint32_t nitems, hipStream_t st,
                       const StepJob* jobs, const WorkItem* items, uint32_t* err, int halos);
// one plane step (the job at `job`) of the rays at slots 0 .. nrays-1, one
// wavefront per ray (FP64, planes of particle trees only)
void launch_plane_step_lone(hipStream_t st, const StepJob* job, uint32_t nrays, uint32_t* err);
// every plane step of the listed bundles in one launch, each wavefront taking
// its bundles through their steps (nblocks == 0: one wavefront per item,
// else a persistent grid of nblocks blocks on a work queue at `counter`)
void launch_plane_steps_fused(int precision, unsigned nblocks, uint32_t nitems, hipStream_t st,
                              const StepJob* jobs, const FusedItem* items, uint32_t* counter,
                              uint32_t* err, int halos);
void launch_zero_u32(hipStream_t st, uint32_t* p, int n);
void launch_plane_step_queue(int precision, unsigned nblocks, uint32_t nitems, hipStream_t st,
                             const StepJob* jobs, const WorkItem* items, uint32_t* counter,
                             uint32_t* err, int halos);
void launch_init_grid(hipStream_t st, const RayState& s, int64_t nslots, int64_t N,
                      int64_t tiles_x, int tw, double c0, double c1, double range, int32_t jmax,
                      double aal, double bbl);
void launch_finalize_grid(hipStream_t st, const RayState& s, int64_t nslots, int64_t N,
                          int64_t tiles_x, int tw, double* y, double* A);
void launch_counts_grid(hipStream_t st, const uint32_t* c, int64_t nslots, int64_t N,
                        int64_t tiles_x, int tw, uint64_t* out);
void launch_init_batch(hipStream_t st, const GridReq* reqs, int nreq, int64_t total);
void launch_finalize_batch(hipStream_t st, const GridReq* reqs, int nreq, int64_t total);
// array rays: `lanes` rays per 64-slot bundle (64: dense)
void launch_init_array(hipStream_t st, const RayState& s, int64_t nslots, int64_t n,
                       const uint32_t* perm, const double* x, const double* zs, int32_t jmax,
                       double aal, double bbl, const double* zpl, const double* Dl,
                       int32_t nplanes, const FlatLCDM& cosmo, int lanes = 64);
void launch_finalize_array(hipStream_t st, const RayState& s, int64_t nslots, int64_t n,
                           const uint32_t* perm, double* y, double* A, int lanes = 64);
void launch_ray_morton(hipStream_t st, const double* x, int64_t n, double x0, double y0,
                       double inv_w, uint64_t* code, uint32_t* idx);
constexpr int kGuardBlocks = 1024;   // partial sums of the guard's reduction
void launch_guard_sumsq(hipStream_t st, const RayState& s, int64_t nslots, const GuardMap& m,
                        double* part, double* acc);
void launch_guard_flag(hipStream_t st, const RayState& s, int64_t nslots, const GuardMap& m,
                       int64_t nrays, const double* acc, uint32_t* count, uint32_t cap,
                       uint32_t* idx, double* xl, double* zl);
void launch_guard_scatter(hipStream_t st, uint32_t n, const uint32_t* idx, const double* yt,
                          const double* At, double* y, double* A);
void launch_minmax_rays(hipStream_t st, const double* x, int64_t n, double* part, int nblk);
// kind[i]: 0 NFW (param = concentration, a float), 1 power law (param = beta);
// kind == nullptr: all NFW
int build_halo_plane(slsim_plane p, const double* x, const double* y, const float* mass,
                     const float* rsize, const int32_t* kind, const double* param, size_t n,
                     double inv_area, double theta_force);
void launch_halo_force(hipStream_t st, const TreeDesc* tree, const double* rays, int64_t n,
                       int direct, double* out, uint32_t* err);
void launch_force(hipStream_t st, int precision, const TreeDesc* tree, const double* rays,
                  int64_t n, double* out, uint32_t* err);
void launch_direct(hipStream_t st, const PartD* parts, int64_t np, double ia, const double* rays,
                   int64_t n, double* out);
// tree build (tree_build.hip)
template <typename T>
int build_plane(slsim_ctx ctx, const T* x, const T* y, const float* m, const float* h, size_t n,
                slsim_plane pl);
int export_nodes(slsim_plane pl, double* dbl, int64_t* ints);
// the precision form's guard over one finished shoot of L (api.hip): flagged
// and small-|alpha| rays shot again in FP64 on the context stream, to the
// per-ray source redshifts zs_dev or else to zs_uniform (<= 0: the lens's
// source)
int guard_refine_ctx(slsim_lens L, const RayState& st, int64_t nslots, const GuardMap& m,
                     int64_t nrays, const double* zs_dev, double zs_uniform, double* y_dev,
                     double* A_dev, int32_t* nguard);
// lens helpers (api.hip)
void uniform_source(const slsim_lens_s* L, double zs, int32_t* jmax, double* aal, double* bbl);
void make_steps(const slsim_lens_s* L, const RayState& st, int64_t nslots, uint32_t* counters,
                std::vector<StepJob>& steps);
// GridMap::find_images over a shot grid (images.hip)
int find_images(slsim_ctx c, const GridGeom& g, const double* y_dev, const double* A_dev,
                size_t nsrc, const double* ys, size_t cap, int32_t* multiplicity, double* rays,
                int64_t* triangles, size_t* nimages);
// images of every grid position as a source, per position (images.hip)
int count_grid_images(slsim_ctx c, const GridGeom& g, const double* y_dev, int32_t* mult_dev);
// critical curves and contours of a shot grid (crit.hip)
int find_crit(slsim_ctx c, const GridGeom& g, const double* y_dev, const double* A_dev,
              slsim_curves_s* out);
int find_magnification_contour(slsim_ctx c, const GridGeom& g, const double* A_dev,
                               double invmag, slsim_curves_s* out);
int find_boundaries_of_caustics(slsim_ctx c, const GridGeom& g, const double* y_dev,
                                int32_t* mult_host, slsim_curves_s* out);
// Utilities::find_boundaries, convex_hull and inCurve on host data (crit.hip)
int find_boundaries_host(const uint8_t* bitmap, size_t nx, size_t ny, slsim_curves_s* out);
std::vector<int64_t> convex_hull_host(const double* p, size_t n);
bool in_curve_host(const double* poly, size_t n, double x0, double x1);
int tightest_hull(slsim_ctx c, const double* v, size_t n, std::vector<double>& out);
int windings_host(const double* x, const double* pt, size_t n, double* area);

}  // namespace slsim
