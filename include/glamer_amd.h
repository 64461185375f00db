// glamer_amd.h -- drop-in C++ front end of the MI355X particle ray shooter.
//
// Header-only layer over the C ABI in slsim_amd.h that keeps the names,
// signatures, argument meaning and error behaviour of GLAMER's particle
// ray-shooting API, so that user code written against
//   ParticleType<T>            include/particle_types.h:46-58
//   Point_2d, Matrix2x2, Point, LinkedPoint, RAY   include/point.h:48-621
//   COSMOLOGY (flat LCDM part)  CosmoLib, used at MultiPlane/lens.cpp:776,2489
//   LensHaloParticles<PType>   include/particle_halo.h:58-113, 341-434
//   Lens                       include/lens.h:75-349, MultiPlane/lens.cpp
//   GridMap                    include/gridmap.h:31-180, TreeCode_link/gridmap.cpp
//   PixelMap<T>::printFITS     ImageProcessing/pixelmap.cpp:1318-1388
//   MakeParticleLenses, SimFileFormat  include/particle_halo.h:964-1080,
//                              include/particle_types.h:32-43
//   LensHaloNFW (symmetric), LensHaloHalos<LensHaloNFW>
//                              include/lens_halos.h:745-865, particle_halo.h:1092-1227
// compiles against this header and runs the rays on the GPU.  Link with
// libslsim_amd.so.  Reference behaviours kept on purpose are marked
// "(reference)".  Deliberate differences: a non-flat cosmology throws
// std::invalid_argument instead of exit(1); LensHaloParticles takes an
// optional trailing opening angle (the reference fixes 0.1); a copied
// GridMap links its own points (the reference's copy links into the
// original's).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <list>
#include <memory>
#include <type_traits>
#include <mutex>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "slsim_amd.h"

typedef double PosType;
typedef double KappaType;
#ifndef IndexType_declare
#define IndexType_declare
typedef size_t IndexType;   // include/standard.h:110-113 (reference)
#endif

#ifndef arcsecTOradians
#define arcsecTOradians 0.000004848136811   // include/standard.h:51 (reference)
#endif

namespace slsim_cpp {

inline void check(int rc) {
  if (rc == SLSIM_OK) return;
  std::string msg = slsim_last_error();
  if (rc == SLSIM_ERR_ARG || rc == SLSIM_ERR_SAME_PLACE || rc == SLSIM_ERR_NOT_FLAT)
    throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// One library context per process, on the device chosen before first use.
inline int& device_index() {
  static int d = 0;
  return d;
}
inline slsim_ctx context() {
  static slsim_ctx ctx = [] {
    slsim_ctx c = nullptr;
    check(slsim_ctx_create(device_index(), 0, &c));
    return c;
  }();
  return ctx;
}

// Precision of the tree walk used by Lens::rayshooterInternal and GridMap.
inline int& precision() {
  static int p = SLSIM_FP64;
  return p;
}

}  // namespace slsim_cpp

// ------------------------------------------------------------------- points
struct Point_2d {
  Point_2d() { x[0] = x[1] = 0.0; }
  Point_2d(double x1, double x2) { x[0] = x1; x[1] = x2; }
  Point_2d(const PosType* p) { x[0] = p[0]; x[1] = p[1]; }
  double& operator[](int i) { return x[i]; }
  double operator[](int i) const { return x[i]; }
  Point_2d operator-(const Point_2d& p) const { return Point_2d(x[0] - p.x[0], x[1] - p.x[1]); }
  Point_2d operator+(const Point_2d& p) const { return Point_2d(x[0] + p.x[0], x[1] + p.x[1]); }
  Point_2d& operator+=(const Point_2d& p) { x[0] += p.x[0]; x[1] += p.x[1]; return *this; }
  Point_2d& operator/=(PosType v) { x[0] /= v; x[1] /= v; return *this; }
  Point_2d operator*(PosType v) const { return Point_2d(x[0] * v, x[1] * v); }
  double length() const { return std::sqrt(x[0] * x[0] + x[1] * x[1]); }
  double x[2];
};

// include/point.h:35
enum class CritType { ND = 0, radial = 2, tangential = 1, pseudo = 3 };

namespace slsim_cpp {
// Reads and releases a curve list of the library.
template <typename P>
void take_curves(slsim_curves h, std::vector<std::vector<P>>& curves, size_t offset,
                 std::vector<int32_t>& types, std::vector<int32_t>& hits, int& replaced) {
  slsim_curves_info info;
  int rc = slsim_curves_get_info(h, &info);
  std::vector<int64_t> off(info.n_curves + 1);
  std::vector<double> pts(2 * info.n_points + 2);
  types.assign(info.n_curves + 1, 0);
  hits.assign(info.n_curves + 1, 0);
  if (rc == SLSIM_OK) rc = slsim_curves_read(h, off.data(), pts.data(), types.data(), hits.data());
  slsim_curves_destroy(h);
  check(rc);
  types.resize(info.n_curves);
  hits.resize(info.n_curves);
  replaced = info.hits_replaced;
  curves.resize(offset + info.n_curves);
  for (size_t c = 0; c < info.n_curves; ++c) {
    std::vector<P>& v = curves[offset + c];
    v.resize((size_t)(off[c + 1] - off[c]));
    for (size_t q = 0; q < v.size(); ++q) {
      v[q][0] = pts[2 * (off[c] + q)];
      v[q][1] = pts[2 * (off[c] + q) + 1];
    }
  }
}
}  // namespace slsim_cpp

namespace Utilities {
// include/concave_hull.h:200-494: the boundaries of the true pixels in pixel
// units, pixel (0,0) at (0,0); the edge pixels of bitmap are set false first,
// as the reference does to the caller's bitmap.
template <typename P>
void find_boundaries(std::vector<bool>& bitmap, long nx, std::vector<std::vector<P>>& points,
                     std::vector<bool>& hits_edge, bool add_to_vector = false,
                     bool outer_only = false) {
  const size_t n = bitmap.size();
  const long ny = nx > 0 ? (long)(n / nx) : 0;
  if (nx <= 0 || n != (size_t)(nx * ny)) throw std::invalid_argument("invalid size");
  std::vector<uint8_t> bm(n);
  for (size_t i = 0; i < n; ++i) bm[i] = bitmap[i];
  slsim_curves h = nullptr;
  slsim_cpp::check(slsim_find_boundaries(bm.data(), (size_t)nx, (size_t)ny, &h));
  for (long i = 0; i < nx; ++i) bitmap[i] = bitmap[i + nx * (ny - 1)] = false;
  for (long j = 0; j < ny; ++j) bitmap[j * nx] = bitmap[nx - 1 + j * nx] = false;
  const size_t offset = add_to_vector ? points.size() : 0;
  if (!add_to_vector) hits_edge.resize(0);
  std::vector<int32_t> types, hits;
  int replaced;
  std::vector<std::vector<P>> found;
  slsim_cpp::take_curves(h, found, 0, types, hits, replaced);
  if (outer_only && found.size() > 1) {
    found.resize(1);
    hits.resize(1);
  }
  points.resize(offset + found.size());
  for (size_t c = 0; c < found.size(); ++c) points[offset + c] = found[c];
  for (int32_t b : hits) hits_edge.push_back(b != 0);
}

// include/concave_hull.h:560-609: indices of the convex hull, counter-clockwise
template <typename T>
void convex_hull(const std::vector<T>& P, std::vector<size_t>& hull) {
  std::vector<double> p(2 * P.size());
  for (size_t i = 0; i < P.size(); ++i) {
    p[2 * i] = P[i][0];
    p[2 * i + 1] = P[i][1];
  }
  std::vector<int64_t> h(2 * P.size() + 1);
  size_t nh = 0;
  slsim_cpp::check(slsim_convex_hull(p.data(), P.size(), h.data(), &nh));
  hull.assign(h.begin(), h.begin() + nh);
}
// include/concave_hull.h:513-556: the hull's points
template <typename T>
std::vector<T> convex_hull(const std::vector<T>& PP) {
  if (PP.size() <= 3) return PP;
  std::vector<size_t> h;
  convex_hull(PP, h);
  std::vector<T> out;
  for (size_t i : h) out.push_back(PP[i]);
  return out;
}

// include/concave_hull.h:1022-1063
template <typename Ptype>
bool inCurve(const Ptype& x, const std::vector<Ptype>& H) {
  std::vector<double> h(2 * H.size());
  for (size_t i = 0; i < H.size(); ++i) {
    h[2 * i] = H[i][0];
    h[2 * i + 1] = H[i][1];
  }
  const double q[2] = {x[0], x[1]};
  int32_t in = 0;
  slsim_cpp::check(slsim_in_curve(h.data(), H.size(), q, 1, &in));
  return in != 0;
}
}  // namespace Utilities

template <typename T = double>
struct Matrix2x2 {
  T a[4];
  T& operator[](int i) { return a[i]; }
  T operator[](int i) const { return a[i]; }
  void setToI() { a[0] = a[3] = 1; a[1] = a[2] = 0; }
  static Matrix2x2<T> I() { Matrix2x2<T> m; m.setToI(); return m; }
  T det() const { return a[0] * a[3] - a[1] * a[2]; }
  // point.h:377-381 (reference, "GLAMER II" convention)
  T kappa() const { return 1 - (a[0] + a[3]) / 2; }
  T gamma1() const { return (a[0] - a[3]) / 2; }
  T gamma2() const { return (a[1] + a[2]) / 2; }
  T gamma3() const { return (a[2] - a[1]) / 2; }
  Matrix2x2<T> operator*(const Matrix2x2<T>& F) const {
    Matrix2x2<T> m;
    m.a[0] = a[0] * F.a[0] + a[1] * F.a[2];
    m.a[1] = a[0] * F.a[1] + a[1] * F.a[3];
    m.a[2] = a[2] * F.a[0] + a[3] * F.a[2];
    m.a[3] = a[2] * F.a[1] + a[3] * F.a[3];
    return m;
  }
  Matrix2x2<T> operator*(T f) const {
    Matrix2x2<T> m = *this;
    for (auto& v : m.a) v *= f;
    return m;
  }
};

enum Boo { NO, YES, MAYBE };

struct Point : public Point_2d {
  Point() { A.setToI(); }
  Point(PosType x1, PosType x2) : Point_2d(x1, x2) { A.setToI(); }
  Point* next = nullptr;
  Point* prev = nullptr;
  Point* image = nullptr;
  unsigned long id = 0;
  unsigned long head = 0;
  Boo in_image = NO;
  double dt = 0;
  Matrix2x2<KappaType> A;
  double gridsize = 0;
  float surface_brightness = 0;
  bool flag = false;
  PosType* ptr_y() { return image->x; }
  KappaType invmag() const { return A.det(); }
  KappaType kappa() const { return A.kappa(); }
  KappaType gamma1() const { return A.gamma1(); }
  KappaType gamma2() const { return A.gamma2(); }
  KappaType gamma3() const { return A.gamma3(); }
};

struct LinkedPoint : public Point {
  LinkedPoint() { image = &im; im.image = this; }
  LinkedPoint(const LinkedPoint& o) : Point(o), im(o.im) { image = &im; im.image = this; }
  Point_2d& y() { return im; }
 private:
  Point im;
};

struct RAY {
  RAY() { A.setToI(); }
  RAY(const Point& p, double zs = -1) : x(p.x[0], p.x[1]), y(p.image->x[0], p.image->x[1]),
                                         A(p.A), dt(p.dt), z(zs) {}
  Point_2d x;
  Point_2d y;
  Matrix2x2<KappaType> A;
  KappaType dt = 0;
  KappaType z = -1;
  PosType* ptr_y() { return y.x; }
  KappaType invmag() const { return A.det(); }
  KappaType kappa() const { return A.kappa(); }
  KappaType gamma1() const { return A.gamma1(); }
  KappaType gamma2() const { return A.gamma2(); }
  KappaType gamma3() const { return A.gamma3(); }
  Point_2d alpha() const { return x - y; }
};

template <typename T = float>
struct ParticleType {
  T& operator[](int i) { return x[i]; }
  T* operator*() { return x; }
  T x[3];
  float Mass;
  float Size;
  int type;
  float size() { return Size; }
  float mass() { return Mass; }
};

static_assert(sizeof(ParticleType<float>) == sizeof(slsim_particle_f32),
              "ParticleType<float> is the glmb record");

// include/particle_types.h:32-43
enum class SimFileFormat { glmb, csv3, csv4, csv5, csv6, gadget2, ascii };

template <typename T = double>
struct Point_3d {
  Point_3d() { x[0] = x[1] = x[2] = 0.0; }
  Point_3d(T xx, T yy, T zz) { x[0] = xx; x[1] = yy; x[2] = zz; }
  T& operator[](int i) { return x[i]; }
  T operator[](int i) const { return x[i]; }
  T x[3];
};

enum class LensingVariable { DELAYT, ALPHA, ALPHA1, ALPHA2, KAPPA, GAMMA, GAMMA1, GAMMA2,
                             GAMMA3, INVMAG, PHI, SurfBrightness };

// --------------------------------------------------------------- cosmology
// The flat-LCDM distances the particle path uses (CosmoLib's COSMOLOGY).
class COSMOLOGY {
 public:
  COSMOLOGY(double Omega_matter = 0.3, double Omega_lambda = 0.7, double h = 0.7)
      : Om(Omega_matter), OL(Omega_lambda), hh(h) {}
  double getOmega_matter() const { return Om; }
  double getOmega_lambda() const { return OL; }
  double gethubble() const { return hh; }
  double coorDist(double z1, double z2) const { return slsim_coordist(Om, OL, hh, z1, z2); }
  double coorDist(double z) const { return coorDist(0, z); }
  double angDist(double z) const { return coorDist(0, z) / (1 + z); }
  double invCoorDist(double D) const { return slsim_inv_coordist(Om, OL, hh, D); }
 private:
  double Om, OL, hh;
};

// ------------------------------------------------------ particle lens halo
class MakeParticleLenses;

template <typename PType>
class LensHaloParticles {
 public:
  // particle_halo.h:58-75: takes possession of the particles and builds the
  // device quadtree (bucket 20, opening angle theta_force).
  LensHaloParticles(std::vector<PType>& pvector, float redshift, const COSMOLOGY& cosmo,
                    Point_2d theta_rotate, bool recenter, double my_inv_area,
                    float MinPSize = 0, double max_range = -1, bool verbose = false,
                    double theta_force = 0.1)
      : zlens(redshift), inv_area(my_inv_area), min_size(MinPSize), max_range_(max_range) {
    (void)cosmo; (void)verbose;
    std::swap(pvector, trash_collector);
    pp = trash_collector.data();
    Npoints = trash_collector.size();
    set_up(theta_rotate, recenter, theta_force);
  }
  // particle_halo.h:78-93: does not take possession of the particles.  As in
  // the reference, recentring and rotation change the caller's array; the
  // device tree is built from it here, so the caller may free it afterwards
  // unless it reads particles back through operator[].
  LensHaloParticles(PType* begin, size_t n, float redshift, const COSMOLOGY& cosmo,
                    Point_2d theta_rotate, bool recenter, double my_inv_area,
                    float MinPSize = 0, double max_range = -1, bool verbose = false,
                    double theta_force = 0.1)
      : zlens(redshift), inv_area(my_inv_area), min_size(MinPSize), max_range_(max_range),
        pp(begin), Npoints(n) {
    (void)cosmo; (void)verbose;
    set_up(theta_rotate, recenter, theta_force);
  }
  // particle_halo.h:226-340: particles from an ASCII file ("# nparticles N"
  // header; x y z [mass] rows, comoving Mpc), sizes from
  // "<file>.<Nsmooth>sizes" or computed on the device and written there.
  LensHaloParticles(const std::string& simulation_filename, SimFileFormat format, PosType redshift,
                    int Nsmooth, const COSMOLOGY& cosmo, Point_2d theta_rotate, bool recenter,
                    bool my_multimass, double my_inv_area, PosType MinPSize = 0,
                    PosType rescale_mass = 1.0, bool verbose = false)
      : zlens((float)redshift), inv_area(my_inv_area), min_size((float)MinPSize) {
    (void)cosmo; (void)verbose;
    if (format != SimFileFormat::ascii) throw std::invalid_argument("bad format");
    size_t n = 0;
    slsim_cpp::check(slsim_ascii_halo_count(simulation_filename.c_str(), my_multimass, &n));
    using T = typename std::remove_reference<decltype(std::declval<PType&>()[0])>::type;
    std::vector<T> xyz(3 * n);
    std::vector<float> m(n), h(n);
    slsim_ascii_halo_info info;
    const double th[2] = {theta_rotate[0], theta_rotate[1]};
    if constexpr (sizeof(T) == sizeof(float))
      slsim_cpp::check(slsim_ascii_halo_load_f32(slsim_cpp::context(), simulation_filename.c_str(),
                                                 redshift, Nsmooth, th, recenter, my_multimass,
                                                 MinPSize, rescale_mass, n, (float*)xyz.data(),
                                                 m.data(), h.data(), &info));
    else
      slsim_cpp::check(slsim_ascii_halo_load_f64(slsim_cpp::context(), simulation_filename.c_str(),
                                                 redshift, Nsmooth, th, recenter, my_multimass,
                                                 MinPSize, rescale_mass, n, (double*)xyz.data(),
                                                 m.data(), h.data(), &info));
    trash_collector.resize(n);
    std::vector<T> x(n), y(n);
    for (size_t i = 0; i < n; ++i) {
      for (int k = 0; k < 3; ++k) trash_collector[i][k] = xyz[3 * i + k];
      trash_collector[i].Mass = m[i];
      trash_collector[i].Size = h[i];
      trash_collector[i].type = 0;
      x[i] = xyz[3 * i];
      y[i] = xyz[3 * i + 1];
    }
    pp = trash_collector.data();
    Npoints = n;
    mass = info.mass;
    for (int k = 0; k < 3; ++k) mcenter[k] = info.mcenter[k];
    build<T>(x, y, m, h, theta_force_);
  }
  LensHaloParticles(LensHaloParticles&& h) noexcept
      : zlens(h.zlens), inv_area(h.inv_area), min_size(h.min_size), mass(h.mass),
        theta_force_(h.theta_force_), max_range_(h.max_range_),
        trash_collector(std::move(h.trash_collector)),
        pp(h.owns() ? trash_collector.data() : h.pp), Npoints(h.Npoints), plane(h.plane) {
    for (int i = 0; i < 3; ++i) mcenter[i] = h.mcenter[i];
    for (int i = 0; i < 3; ++i) densest_point[i] = h.densest_point[i];
    h.pp = nullptr;
    h.Npoints = 0;
    h.plane = nullptr;
  }
  // a copy holds its own device tree, built from the same particles (its own
  // copy of them when the original owns them, else the same borrowed array)
  LensHaloParticles(const LensHaloParticles& h)
      : zlens(h.zlens), inv_area(h.inv_area), min_size(h.min_size), mass(h.mass),
        theta_force_(h.theta_force_), max_range_(h.max_range_),
        trash_collector(h.trash_collector),
        pp(h.owns() ? trash_collector.data() : h.pp), Npoints(h.Npoints) {
    for (int i = 0; i < 3; ++i) mcenter[i] = h.mcenter[i];
    for (int i = 0; i < 3; ++i) densest_point[i] = h.densest_point[i];
    rebuild();
  }
  ~LensHaloParticles() {
    if (plane) slsim_plane_destroy(plane);
  }
  size_t getN() const { return Npoints; }
  float getZlens() const { return zlens; }
  double get_mass() const { return mass; }
  PType& operator[](size_t i) { return pp[i]; }
  // particle_halo.h:107-110: the centre of mass and the smallest particle's
  // position in input coordinates (after recentring and rotation)
  Point_3d<> CenterOfMass() { return Point_3d<>(mcenter[0], mcenter[1], mcenter[2]); }
  Point_3d<> DensestPoint() {
    return Point_3d<>(densest_point[0], densest_point[1], densest_point[2]);
  }
  // particle_halo.h:436-443: rotates the particles (the caller's array when
  // borrowed) and rebuilds the tree with the reference's defaults there
  // (opening angle 0.1, no maximum range)
  void rotate(Point_2d theta) {
    rotate_particles(theta[0], theta[1]);
    if (plane) slsim_plane_destroy(plane);
    plane = nullptr;
    theta_force_ = 0.1;
    max_range_ = -1;
    rebuild();
  }
  slsim_plane release_plane() {
    slsim_plane p = plane;
    plane = nullptr;
    return p;
  }

  // particle_halo.h:683-730: pp[i].Size = distance to the (Nsmooth+1)-th
  // nearest particle in 3-D, itself included, computed on the device.
  static void calculate_smoothing(int Nsmooth, PType* pp, size_t Npoints, bool verbose = false) {
    (void)verbose;
    using T = typename std::remove_reference<decltype(pp[0][0])>::type;
    std::vector<T> x(Npoints), y(Npoints), z(Npoints);
    for (size_t i = 0; i < Npoints; ++i) {
      x[i] = pp[i][0];
      y[i] = pp[i][1];
      z[i] = pp[i][2];
    }
    std::vector<float> size(Npoints);
    if constexpr (sizeof(T) == sizeof(float))
      slsim_cpp::check(slsim_smoothing_sizes_f32(slsim_cpp::context(), (const float*)x.data(),
                                                 (const float*)y.data(), (const float*)z.data(),
                                                 Npoints, Nsmooth, size.data()));
    else
      slsim_cpp::check(slsim_smoothing_sizes_f64(slsim_cpp::context(), (const double*)x.data(),
                                                 (const double*)y.data(), (const double*)z.data(),
                                                 Npoints, Nsmooth, size.data()));
    for (size_t i = 0; i < Npoints; ++i) pp[i].Size = size[i];
  }

 protected:
  friend class MakeParticleLenses;
  // particle_halo.h:194-207, the constructor MakeParticleLenses uses: the
  // particles are borrowed, as in the reference (set_up changes them only
  // when recentring or rotating, which CreateHalos does not ask for).
  LensHaloParticles(PType* pdata, size_t Nparticles, float redshift, const COSMOLOGY& cosmo,
                    Point_2d theta_rotate, bool recenter, double my_inv_area, bool verbose = false)
      : zlens(redshift), inv_area(my_inv_area), min_size(0), pp(pdata), Npoints(Nparticles) {
    (void)cosmo; (void)verbose;
    set_up(theta_rotate, recenter, 0.1);
  }

 private:
  bool owns() const { return !trash_collector.empty(); }
  void rebuild() {
    const size_t n = Npoints;
    using T = typename std::remove_reference<decltype(std::declval<PType&>()[0])>::type;
    std::vector<T> x(n), y(n);
    std::vector<float> m(n), h(n);
    for (size_t i = 0; i < n; ++i) {
      x[i] = pp[i][0];
      y[i] = pp[i][1];
      m[i] = pp[i].mass();
      h[i] = pp[i].size();
    }
    build<T>(x, y, m, h, theta_force_);
  }
  // set_up, particle_halo.h:341-417
  void set_up(Point_2d theta_rotate, bool recenter, double theta_force) {
    theta_force_ = theta_force;
    const size_t n = Npoints;
    size_t smallest = 0;
    for (size_t i = 0; i < n; ++i)
      if (pp[i].Size < pp[smallest].Size) smallest = i;
    for (int j = 0; j < 3; ++j) densest_point[j] = n ? pp[smallest][j] : 0.0;
    mcenter[0] = mcenter[1] = mcenter[2] = 0;
    mass = 0;
    for (size_t i = 0; i < n; ++i) {
      for (int j = 0; j < 3; ++j) mcenter[j] += pp[i][j] * pp[i].mass();
      mass += pp[i].mass();
    }
    for (int j = 0; j < 3; ++j) mcenter[j] /= mass;
    if (recenter) {
      // (reference) the centre is zeroed inside the loop, so only the first
      // particle moves (particle_halo.h:395-411)
      for (size_t i = 0; i < n; ++i) {
        for (int j = 0; j < 3; ++j) pp[i][j] -= mcenter[j];
        for (int j = 0; j < 3; ++j) densest_point[j] -= mcenter[j];
        mcenter[0] = mcenter[1] = mcenter[2] = 0;
      }
    }
    rotate_particles(theta_rotate[0], theta_rotate[1]);
    rebuild();
  }
  template <typename T>
  void build(const std::vector<T>& x, const std::vector<T>& y, const std::vector<float>& m,
             const std::vector<float>& h, double theta_force) {
    if constexpr (sizeof(T) == sizeof(float))
      slsim_cpp::check(slsim_plane_create_range_f32(
          slsim_cpp::context(), (const float*)x.data(), (const float*)y.data(), m.data(),
          h.data(), x.size(), theta_force, 20, inv_area, max_range_, &plane));
    else
      slsim_cpp::check(slsim_plane_create_range_f64(
          slsim_cpp::context(), (const double*)x.data(), (const double*)y.data(), m.data(),
          h.data(), x.size(), theta_force, 20, inv_area, max_range_, &plane));
  }
  // rotate_particles, particle_halo.h:634-680
  void rotate_particles(PosType tx, PosType ty) {
    if (tx == 0.0 && ty == 0.0) return;
    const double cx = std::cos(tx), sx = std::sin(tx), cy = std::cos(ty), sy = std::sin(ty);
    const double c[3][3] = {{cy, 0, -sy}, {-sy * sx, cx, -cy * sx}, {cx, sx, cy * cx}};
    for (size_t i = 0; i < Npoints; ++i) {
      PType& p = pp[i];
      double t[3] = {0, 0, 0};
      for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r) t[r] += c[r][j] * p[j];
      for (int j = 0; j < 3; ++j) p[j] = t[j];
    }
    for (double* v : {mcenter, densest_point}) {
      double t[3] = {0, 0, 0};
      for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r) t[r] += c[r][j] * v[j];
      for (int j = 0; j < 3; ++j) v[j] = t[j];
    }
  }

  float zlens;
  double inv_area;
  float min_size;
  double mass = 0;
  double theta_force_ = 0.1;
  double max_range_ = -1;   // the tree's maximum range (particle_halo.h:416)
  double mcenter[3] = {0, 0, 0};
  double densest_point[3] = {0, 0, 0};
  std::vector<PType> trash_collector;   // the particles when the halo owns them
  PType* pp = nullptr;                  // the particles (owned or borrowed)
  size_t Npoints = 0;
  slsim_plane plane = nullptr;
};

// ------------------------------------------------------- NFW halo planes
enum class EllipMethod { Fourier, Pseudo, Schramm, Keeton };

// LensHaloNFW (include/lens_halos.h:745-865, MultiPlane/lens_halos.cpp:
// 389-440) as a member of a LensHaloHalos: symmetric profiles only (fratio
// 1; others throw std::invalid_argument).  The position is set in angle
// with setTheta and is theta * D_A(zlens) in physical Mpc, as the reference's.
// include/lens_halos.h: what a halo tree reads of a LensHalo (position,
// mass, Rsize) and the kind and parameter of its profile.
class LensHalo {
 public:
  virtual ~LensHalo() = default;
  void setTheta(PosType PosX, PosType PosY) { posHalo[0] = PosX; posHalo[1] = PosY; }
  void getX(PosType* MyPosHalo) const {
    MyPosHalo[0] = posHalo[0] * Dist;
    MyPosHalo[1] = posHalo[1] * Dist;
  }
  PosType operator[](int i) const { return posHalo[i] * Dist; }
  float get_mass() const { return mass; }
  float getRsize() const { return Rsize; }
  float get_Rmax() const { return Rsize; }
  PosType getZlens() const { return zlens; }
  // SLSIM_HALO_* and its parameter (concentration or beta)
  virtual int32_t profile_kind() const = 0;
  virtual double profile_param() const = 0;

 protected:
  LensHalo(float my_mass, float my_Rsize, PosType my_zlens, const COSMOLOGY& cosmo)
      : mass(my_mass), Rsize(my_Rsize), zlens(my_zlens) {
    Dist = cosmo.angDist(my_zlens);
  }
  float mass, Rsize;
  PosType zlens;
  PosType Dist = -1;
  PosType posHalo[2] = {0, 0};
};

// MultiPlane/lens_halos.cpp:389-440 (symmetric: fratio = 1)
class LensHaloNFW : public LensHalo {
 public:
  LensHaloNFW(float my_mass, float my_Rsize, PosType my_zlens, float my_concentration,
              float my_fratio, float my_pa, const COSMOLOGY& cosmo,
              EllipMethod my_ellip_method = EllipMethod::Pseudo)
      : LensHalo(my_mass, my_Rsize, my_zlens, cosmo), concentration(my_concentration) {
    (void)my_pa; (void)my_ellip_method;
    if (my_fratio != 1)
      throw std::invalid_argument("LensHaloNFW: only fratio = 1 (symmetric) halos are built");
  }
  float getConcentration() const { return concentration; }
  int32_t profile_kind() const override { return SLSIM_HALO_NFW; }
  double profile_param() const override { return concentration; }

 private:
  float concentration;
};

// MultiPlane/lens_halos.cpp:805-866 (symmetric: fratio = 1): kappa ~ r^-beta
// inside Rsize, beta < 2
class LensHaloPowerLaw : public LensHalo {
 public:
  LensHaloPowerLaw(float my_mass, float my_Rsize, PosType my_zlens, PosType my_beta,
                   float my_fratio, float my_pa, const COSMOLOGY& cosmo,
                   EllipMethod my_ellip_method = EllipMethod::Pseudo)
      : LensHalo(my_mass, my_Rsize, my_zlens, cosmo), beta(my_beta) {
    (void)my_pa; (void)my_ellip_method;
    if (my_beta >= 2.0) throw std::invalid_argument("bad gamma");
    if (my_Rsize <= 0.0 || std::isnan(my_Rsize)) throw std::runtime_error("Bad halo");
    if (my_fratio != 1)
      throw std::invalid_argument("LensHaloPowerLaw: only fratio = 1 (symmetric) halos are built");
  }
  int32_t profile_kind() const override { return SLSIM_HALO_POWERLAW; }
  double profile_param() const override { return beta; }

 private:
  PosType beta;
};

// LensHaloHalos<HType> (include/particle_halo.h:1092-1227): a main halo made
// of halos on a TreeQuadHalos; the tree is built and walked by the library.
template <typename HType>
class LensHaloHalos {
  static_assert(std::is_same<HType, LensHaloNFW>::value ||
                    std::is_same<HType, LensHaloPowerLaw>::value,
                "LensHaloHalos is built for LensHaloNFW and LensHaloPowerLaw members");

 public:
  LensHaloHalos(std::vector<HType>& pvector, float redshift, const COSMOLOGY& cosmo,
                float inv_area, float my_theta_force = 0.1, bool mu_periodic_buffer = false,
                bool verbose = false)
      : zlens(redshift) {
    (void)cosmo; (void)verbose;
    if (mu_periodic_buffer)
      throw std::invalid_argument("LensHaloHalos: the periodic buffer is not built");
    std::swap(pvector, trash_collector);
    const size_t n = trash_collector.size();
    std::vector<double> x(n), y(n), prm(n);
    std::vector<float> m(n), r(n);
    std::vector<int32_t> kind(n);
    mass = 0;
    for (size_t i = 0; i < n; ++i) {
      double p[2];
      trash_collector[i].getX(p);
      x[i] = p[0];
      y[i] = p[1];
      m[i] = trash_collector[i].get_mass();
      r[i] = trash_collector[i].getRsize();
      kind[i] = trash_collector[i].profile_kind();
      prm[i] = trash_collector[i].profile_param();
      mass += m[i];
    }
    slsim_cpp::check(slsim_halos_create(slsim_cpp::context(), x.data(), y.data(), m.data(),
                                        r.data(), kind.data(), prm.data(), n, inv_area,
                                        my_theta_force, 0.0, 0, &plane));
  }
  LensHaloHalos(LensHaloHalos&& h) noexcept
      : zlens(h.zlens), mass(h.mass), trash_collector(std::move(h.trash_collector)),
        plane(h.plane) {
    h.plane = nullptr;
  }
  LensHaloHalos(const LensHaloHalos&) = delete;
  ~LensHaloHalos() {
    if (plane) slsim_plane_destroy(plane);
  }
  size_t getN() const { return trash_collector.size(); }
  float getZlens() const { return zlens; }
  double get_mass() const { return mass; }
  slsim_plane release_plane() {
    slsim_plane p = plane;
    plane = nullptr;
    return p;
  }

 private:
  float zlens;
  double mass = 0;
  std::vector<HType> trash_collector;
  slsim_plane plane = nullptr;
};

// MultiPlane/planes.cpp:13-68, include/planes.h:35-63: a field plane of
// halos on a TreeQuadHalos (sigma_background as its uniform background; a
// nonzero inv_screening_scale screens its far cells and adds the periodic
// buffer), built and walked by the library.  force() is TreeQuadHalos::force2D
// at a physical position; phi and gamma[2] are 0.
class LensPlaneTree {
 public:
  LensPlaneTree(float z_, LensHalo** my_halos, IndexType Nhalos, PosType my_sigma_background,
                PosType my_inv_screening_scale = 0)
      : z(z_), halos(my_halos, my_halos + Nhalos) {
    std::vector<double> x(Nhalos), y(Nhalos), prm(Nhalos);
    std::vector<float> m(Nhalos), r(Nhalos);
    std::vector<int32_t> kind(Nhalos);
    for (IndexType i = 0; i < Nhalos; ++i) {
      double p[2];
      my_halos[i]->getX(p);
      x[i] = p[0];
      y[i] = p[1];
      m[i] = my_halos[i]->get_mass();
      r[i] = my_halos[i]->getRsize();
      kind[i] = my_halos[i]->profile_kind();
      prm[i] = my_halos[i]->profile_param();
    }
    // planes.cpp:13-19: screened trees take theta 0.1 and the periodic buffer
    const bool scr = my_inv_screening_scale != 0;
    slsim_cpp::check(slsim_halos_create(slsim_cpp::context(), x.data(), y.data(), m.data(),
                                        r.data(), kind.data(), prm.data(), Nhalos,
                                        my_sigma_background, scr ? 0.1 : 0.5,
                                        my_inv_screening_scale, scr, &plane));
  }
  LensPlaneTree(LensPlaneTree&& p) noexcept : z(p.z), halos(std::move(p.halos)), plane(p.plane) {
    p.plane = nullptr;
  }
  LensPlaneTree(const LensPlaneTree&) = delete;
  ~LensPlaneTree() {
    if (plane) slsim_plane_destroy(plane);
  }
  void force(PosType* alpha, KappaType* kappa, KappaType* gamma, KappaType* phi, PosType* xx) {
    double out[5];
    slsim_cpp::check(slsim_plane_force(plane, 1, xx, out, SLSIM_FP64));
    alpha[0] = out[0];
    alpha[1] = out[1];
    *kappa = out[2];
    gamma[0] = out[3];
    gamma[1] = out[4];
    gamma[2] = 0;
    *phi = 0;
  }
  std::vector<LensHalo*> getHalos() { return halos; }
  slsim_plane release_plane() {
    slsim_plane p = plane;
    plane = nullptr;
    return p;
  }
  float z;

 private:
  std::vector<LensHalo*> halos;
  slsim_plane plane = nullptr;
};

// --------------------------------------------------------------------- lens
class Lens {
 public:
  Lens(long* my_seed, PosType z_source, const COSMOLOGY& cosmo, bool verbose = false)
      : seed(my_seed), cosmo_(cosmo) {
    (void)verbose;
    slsim_cpp::check(slsim_lens_create(slsim_cpp::context(), z_source, cosmo.getOmega_matter(),
                                       cosmo.getOmega_lambda(), cosmo.gethubble(), &lens));
  }
  ~Lens() {
    if (lens) slsim_lens_destroy(lens);
  }
  Lens(const Lens&) = delete;

  // include/lens.h:172-185: the lens takes possession of the halo.
  // addplanes = false: the nearest main plane (MultiPlane/lens.cpp:822-840)
  template <typename T>
  void moveinMainHalo(T& halo_in, bool addplanes, bool verbose = false) {
    (void)verbose;
    T* halo = new T(std::move(halo_in));
    owned.emplace_back(halo, [](void* p) { delete static_cast<T*>(p); });
    main_zlens.push_back(halo->getZlens());
    if (addplanes)
      slsim_cpp::check(slsim_lens_add_halo(lens, halo->release_plane(), halo->getZlens()));
    else
      slsim_cpp::check(
          slsim_lens_add_halo_nearest(lens, halo->release_plane(), halo->getZlens()));
  }

  // include/lens.h:220-236: copies of the halos added to the existing ones
  template <typename T>
  void insertMainHalos(std::vector<T>& my_halos, bool addplanes, bool verbose = false) {
    for (T& h : my_halos) insertMainHalo(h, addplanes, verbose);
  }
  // include/lens.h:194-207, 244-263: the halos replace the main halos, one
  // plane per redshift
  template <typename T>
  void replaceMainHalo(const T& halo_in, bool addplanes, bool verbose = false) {
    (void)addplanes;
    clearMainHalos(verbose);
    insertMainHalo(halo_in, true, verbose);
  }
  template <typename T>
  void replaceMainHalos(std::vector<T>& my_halos, bool verbose) {
    clearMainHalos(verbose);
    insertMainHalos(my_halos, true, verbose);
  }

  // MultiPlane/lens.cpp:1025-1036
  void clearMainHalos(bool verbose = false) {
    (void)verbose;
    slsim_cpp::check(slsim_lens_clear_main_halos(lens));
    owned.clear();
    main_zlens.clear();
  }
  std::size_t getNMainHalos() const { return main_zlens.size(); }
  // include/lens.h:102-117: the first main halo's redshift and angular
  // size distance
  PosType getZlens() const {
    if (main_zlens.empty()) throw std::runtime_error("error, no main lens present");
    return main_zlens[0];
  }
  PosType getAngDistLens() const { return cosmo_.angDist(getZlens()); }
  PosType getfov() const { return fieldofview; }
  void setfov(PosType fov) { fieldofview = fov; }

  // A field plane at redshift z (the reference builds its field planes in
  // Lens::createFieldPlanes, MultiPlane/lens.cpp:349-455, with double
  // redshifts, and combinePlanes sorts them with the main planes); the lens
  // takes the plane.
  void insertFieldPlane(LensPlaneTree& plane, PosType z) {
    slsim_cpp::check(slsim_lens_add_field_plane(lens, plane.release_plane(), z));
  }

  // include/lens.h:149-170: the lens keeps a copy of the halo.
  template <typename T>
  void insertMainHalo(const T& halo_in, bool addplanes, bool verbose = false) {
    T copy(halo_in);
    moveinMainHalo(copy, addplanes, verbose);
  }

  int getNplanes() const {
    int n = 0;
    slsim_cpp::check(slsim_lens_tables(lens, &n, nullptr, nullptr, nullptr, nullptr));
    return n;
  }
  PosType getSourceZ() const {
    int n = getNplanes();
    std::vector<double> z(n + 1);
    slsim_cpp::check(slsim_lens_tables(lens, &n, z.data(), nullptr, nullptr, nullptr));
    return toggle ? zs_implant : z.back();
  }
  // MultiPlane/lens.cpp:2609-2657
  short ResetSourcePlane(PosType z, bool nearest = false, bool verbose = false) {
    (void)verbose;
    int32_t j = 0;
    double zi = 0;
    slsim_cpp::check(slsim_lens_reset_source_plane(lens, z, nearest, &j, &zi));
    toggle = true;
    zs_implant = zi;
    return (short)j;
  }
  void RevertSourcePlane() {
    slsim_cpp::check(slsim_lens_reset_source(lens, -1));
    toggle = false;
  }

  // FullRange/internal_rayshooter_multi.cpp:53-118
  void rayshooterInternal(unsigned long Npoints, Point* i_points, bool RSIverbose = false) {
    (void)RSIverbose;
    std::vector<size_t> idx;
    std::vector<double> x;
    for (unsigned long i = 0; i < Npoints; ++i) {
      if (i_points[i].in_image == MAYBE) continue;
      idx.push_back(i);
      x.push_back(i_points[i].x[0]);
      x.push_back(i_points[i].x[1]);
    }
    std::vector<double> y(x.size()), A(2 * x.size());
    shoot(idx.size(), x.data(), nullptr, y.data(), A.data());
    for (size_t k = 0; k < idx.size(); ++k) {
      Point& p = i_points[idx[k]];
      p.image->x[0] = y[2 * k];
      p.image->x[1] = y[2 * k + 1];
      for (int c = 0; c < 4; ++c) p.A.a[c] = A[4 * k + c];
      p.dt = 0;
      p.image->A = p.A;
      p.image->dt = 0;
    }
  }
  // internal_rayshooter_multi.cpp:120-182
  void rayshooterInternal(unsigned long Npoints, LinkedPoint* i_points,
                          std::vector<double>& source_zs, bool RSIverbose = false) {
    (void)RSIverbose;
    if (source_zs.size() != Npoints) throw std::invalid_argument("missing redshifts");
    std::vector<double> x(2 * Npoints), y(2 * Npoints), A(4 * Npoints);
    for (unsigned long i = 0; i < Npoints; ++i) {
      x[2 * i] = i_points[i].x[0];
      x[2 * i + 1] = i_points[i].x[1];
    }
    shoot(Npoints, x.data(), source_zs.data(), y.data(), A.data());
    for (unsigned long i = 0; i < Npoints; ++i) {
      LinkedPoint& p = i_points[i];
      p.image->x[0] = y[2 * i];
      p.image->x[1] = y[2 * i + 1];
      for (int c = 0; c < 4; ++c) p.A.a[c] = A[4 * i + c];
      p.dt = 0;
      p.image->A = p.A;
      p.image->dt = 0;
    }
  }
  // internal_rayshooter_multi.cpp:184-235: each ray carries its source z
  void rayshooterInternal(unsigned long Npoints, RAY* rays) {
    std::vector<double> x(2 * Npoints), zs(Npoints), y(2 * Npoints), A(4 * Npoints);
    for (unsigned long i = 0; i < Npoints; ++i) {
      x[2 * i] = rays[i].x[0];
      x[2 * i + 1] = rays[i].x[1];
      zs[i] = rays[i].z;
    }
    shoot(Npoints, x.data(), zs.data(), y.data(), A.data());
    for (unsigned long i = 0; i < Npoints; ++i) {
      rays[i].y[0] = y[2 * i];
      rays[i].y[1] = y[2 * i + 1];
      for (int c = 0; c < 4; ++c) rays[i].A.a[c] = A[4 * i + c];
      rays[i].dt = 0;
    }
  }
  void rayshooterInternal(RAY& ray) { rayshooterInternal(1, &ray); }
  // internal_rayshooter_multi.cpp:40-50: a single ray to the lens's source
  void rayshooter(RAY& ray) {
    LinkedPoint p;
    p.x[0] = ray.x[0];
    p.x[1] = ray.x[1];
    rayshooterInternal(1, static_cast<Point*>(&p));
    ray = RAY(p, -1);
  }

  slsim_lens handle() const { return lens; }
  const COSMOLOGY& getCosmo() const { return cosmo_; }

 private:
  void shoot(size_t n, const double* x, const double* zs, double* y, double* A) {
    if (n == 0) return;
    slsim_cpp::check(slsim_shoot(lens, n, x, zs, y, A, slsim_cpp::precision(), nullptr));
  }
  long* seed;
  COSMOLOGY cosmo_;
  slsim_lens lens = nullptr;
  bool toggle = false;
  double zs_implant = 0;
  PosType fieldofview = 0;
  std::vector<PosType> main_zlens;   // redshifts of the main halos, in order
  std::vector<std::unique_ptr<void, void (*)(void*)>> owned;
};
typedef Lens* LensHndl;

// ------------------------------------------------------- particle files
// MakeParticleLenses (include/particle_halo.h:964-1080, MultiPlane/
// particle_lens.cpp): reads csv3/csv4/csv5, Gadget-2 or glmb files (smoothing
// sizes on the device, cached in "<filename>_S<Nsmooth>.glmb") and makes one
// LensHaloParticles per particle type.  `data` is a std::vector-like view of
// the particles, which the library holds.
class MakeParticleLenses {
 public:
  struct Particles {
    slsim_mpl h = nullptr;
    ParticleType<float>* data() const {
      slsim_particle_f32* p = nullptr;
      size_t n = 0;
      slsim_cpp::check(slsim_mpl_data(h, &p, &n));
      return reinterpret_cast<ParticleType<float>*>(p);
    }
    size_t size() const {
      slsim_particle_f32* p = nullptr;
      size_t n = 0;
      slsim_cpp::check(slsim_mpl_data(h, &p, &n));
      return n;
    }
    ParticleType<float>& operator[](size_t i) const { return data()[i]; }
    ParticleType<float>* begin() const { return data(); }
    ParticleType<float>* end() const { return data() + size(); }
  };

  std::vector<LensHaloParticles<ParticleType<float>>*> halos;
  Particles data;

  MakeParticleLenses(const std::string& filename, SimFileFormat format, int Nsmooth, bool recenter,
                     bool ignore_type_in_smoothing = false) {
    slsim_cpp::check(slsim_mpl_create(slsim_cpp::context(), filename.c_str(), (int)format, Nsmooth,
                                      recenter, ignore_type_in_smoothing, &data.h));
  }
  MakeParticleLenses(const std::string& filename, bool recenter) {
    slsim_cpp::check(slsim_mpl_create_glmb(slsim_cpp::context(), filename.c_str(), recenter,
                                           &data.h));
  }
  MakeParticleLenses(const MakeParticleLenses&) = delete;
  ~MakeParticleLenses() {
    for (auto p : halos) delete p;
    if (data.h) slsim_mpl_destroy(data.h);
  }

  std::vector<size_t> getnp() const {
    slsim_mpl_info i = info();
    return std::vector<size_t>(i.nparticles, i.nparticles + 6);
  }
  std::vector<float> getmp() const {
    slsim_mpl_info i = info();
    return std::vector<float>(i.masses, i.masses + 6);
  }
  double sim_redshift() const { return info().z_original; }
  double getZoriginal() const { return info().z_original; }
  Point_3d<> getCenterOfMass() const {
    slsim_mpl_info i = info();
    return Point_3d<>(i.cm[0], i.cm[1], i.cm[2]);
  }
  void getBoundingBox(Point_3d<>& Xmin, Point_3d<>& Xmax) const {
    slsim_mpl_info i = info();
    Xmin = Point_3d<>(i.bbox_ll[0], i.bbox_ll[1], i.bbox_ll[2]);
    Xmax = Point_3d<>(i.bbox_ur[0], i.bbox_ur[1], i.bbox_ur[2]);
  }
  // particle_lens.cpp:139-151: halos made before are deleted
  void Recenter(Point_3d<> x) {
    slsim_cpp::check(slsim_mpl_recenter(data.h, x.x));
    for (auto p : halos) delete p;
    halos.clear();
  }
  void radialCut(Point_3d<> center, double radius) {
    slsim_cpp::check(slsim_mpl_radial_cut(data.h, center.x, radius));
  }
  void cylindricalCut(Point_2d center, double radius) {
    slsim_cpp::check(slsim_mpl_cylindrical_cut(data.h, center.x, radius));
  }
  Point_3d<> densest_particle() {
    Point_3d<> x;
    slsim_cpp::check(slsim_mpl_densest_particle(data.h, x.x));
    return x;
  }
  // particle_lens.cpp:153-209
  void CreateHalos(const COSMOLOGY& cosmo, double redshift, double inv_area) {
    slsim_cpp::check(slsim_mpl_physical_units(data.h, (float)cosmo.gethubble()));
    for (auto p : halos) delete p;
    halos.clear();
    std::vector<size_t> np = getnp();
    ParticleType<float>* pp = data.data();
    size_t skip = 0;
    for (int i = 0; i < 6; ++i) {
      if (np[i] > 0)
        halos.push_back(new LensHaloParticles<ParticleType<float>>(
            pp + skip, np[i], (float)redshift, cosmo, Point_2d(), false, inv_area, false));
      skip += np[i];
    }
  }

 private:
  slsim_mpl_info info() const {
    slsim_mpl_info i;
    slsim_cpp::check(slsim_mpl_get_info(data.h, &i));
    return i;
  }
};

// ---------------------------------------------------------------- pixel map
template <typename T>
class PixelMap {
 public:
  PixelMap() : Nx_(0), Ny_(0), res(0) { c[0] = c[1] = 0; }
  // pixelmap.cpp:111-130: square pixels of `resolution`, the map's lower-left
  // boundary at center - (N resolution) / 2
  PixelMap(const double center[2], size_t Nx, size_t Ny, double resolution)
      : Nx_(Nx), Ny_(Ny), res(resolution), map(Nx * Ny) {
    c[0] = center[0];
    c[1] = center[1];
  }
  T& operator[](size_t i) { return map[i]; }
  const T& operator[](size_t i) const { return map[i]; }
  // pixelmap.cpp: every pixel set to zero
  void Clean() { std::fill(map.begin(), map.end(), T(0)); }
  // pixelmap.cpp:2227-2242, 2266-2269: the pixel holding x, -1 outside
  long find_index(const double x[]) const {
    const long ix = (long)((x[0] - (c[0] - (Nx_ * res) / 2.)) / res);
    const long iy = (long)((x[1] - (c[1] - (Ny_ * res) / 2.)) / res);
    if (ix < 0 || ix >= (long)Nx_) return -1;
    if (iy < 0 || iy >= (long)Ny_) return -1;
    return ix + (long)Nx_ * iy;
  }
  Point_2d getCenter() const { return Point_2d(c[0], c[1]); }
  size_t size() const { return map.size(); }
  size_t getNx() const { return Nx_; }
  size_t getNy() const { return Ny_; }
  double getResolution() const { return res; }
  const std::vector<T>& data() const { return map; }
  void setRAandDec(double RAin, double DECin) {
    RA = RAin;
    DEC = DECin;
  }
  // pixelmap.cpp:1318-1388: a FITS primary image with the reference's keys
  // (written from the FITS standard; the reference uses cfitsio)
  void printFITS(std::string filename, bool flipX = false, bool ctype = false,
                 bool verbose = false) {
    (void)verbose;
    static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value,
                  "printFITS writes double or float maps");
    if (filename.empty())
      throw std::invalid_argument("Please enter a valid filename for the FITS file output");
    slsim_cpp::check(slsim_fits_write_pixelmap(filename.c_str(), map.data(),
                                               std::is_same<T, double>::value, Nx_, Ny_, res, c,
                                               RA, DEC, units, flipX, ctype));
  }
 private:
  size_t Nx_, Ny_;
  double res;
  double c[2];
  double RA = 0, DEC = 0;
  int units = 0;   // PixelMapUnits::ndef
  std::vector<T> map;
};

// ------------------------------------------------------------------ gridmap
struct GridMap {
  // TreeCode_link/gridmap.cpp:93-120: N1d x N1d rays over `range`, x fastest;
  // the grid is generated and shot on the device.
  GridMap(LensHndl lens, unsigned long N1d, const double center_[2], double range)
      : Ngrid_init(N1d), Ngrid_init2(N1d), x_range(range) {
    if (N1d < 1) throw std::runtime_error("GridMap size is < 1!");
    if (range <= 0) throw std::runtime_error("GridMap range is <= 0!");
    center[0] = center_[0];
    center[1] = center_[1];
    const size_t n = (size_t)N1d * N1d;
    i_points.resize(n);
    s_points.resize(n);
    i_points[0].head = n;
    s_points[0].head = n;
    for (size_t i = 0; i < n; ++i) {
      i_points[i].id = i;
      i_points[i].gridsize = N1d > 1 ? range / (N1d - 1) : range;
      if (N1d == 1) {
        i_points[i].x[0] = center[0];
        i_points[i].x[1] = center[1];
      } else {
        i_points[i].x[0] = center[0] + range * (1.0 * (i % N1d) / (N1d - 1) - 0.5);
        i_points[i].x[1] = center[1] + range * (1.0 * (i / N1d) / (N1d - 1) - 0.5);
      }
      s_points[i].id = i;
      i_points[i].image = &s_points[i];
      s_points[i].image = &i_points[i];
      s_points[i].gridsize = i_points[i].gridsize;
    }
    std::lock_guard<std::mutex> hold(grid_mutex());
    slsim_ctx ctx = slsim_cpp::context();
    // the outputs stay in HBM for find_images
    slsim_cpp::check(slsim_device_alloc(ctx, 16 * n, &yd));
    int rc = slsim_device_alloc(ctx, 32 * n, &Ad);
    if (rc == SLSIM_OK)
      rc = slsim_gridmap_shoot(lens->handle(), N1d, center[0], center[1], range, 0.0,
                               (double*)yd, (double*)Ad, slsim_cpp::precision(), nullptr);
    std::vector<double> y(2 * n), A(4 * n);
    if (rc == SLSIM_OK) rc = slsim_memcpy_d2h(ctx, y.data(), yd, 16 * n);
    if (rc == SLSIM_OK) rc = slsim_memcpy_d2h(ctx, A.data(), Ad, 32 * n);
    if (rc != SLSIM_OK) release_device();
    slsim_cpp::check(rc);
    for (size_t i = 0; i < n; ++i) {
      s_points[i].x[0] = y[2 * i];
      s_points[i].x[1] = y[2 * i + 1];
      for (int c = 0; c < 4; ++c) i_points[i].A.a[c] = A[4 * i + c];
      s_points[i].A = i_points[i].A;
    }
  }

  // gridmap.cpp:26-91: Nx x Ny points, Ny = (int)(Nx rangeY / rangeX), square
  // cells of rangeX / (Nx - 1), point ii + jj Nx at (xo + ii res, yo + jj res)
  // from the lower-left corner (xo, yo); shot as a ray list, outputs kept in
  // HBM for find_images and the critical curves.
  GridMap(LensHndl lens, unsigned long Nx, const PosType center_[2], PosType rangeX,
          PosType rangeY)
      : Ngrid_init(Nx), x_range(rangeX), rect(true) {
    if (Nx <= 0) throw std::runtime_error("cannot make GridMap with no points");
    if (rangeX <= 0 || rangeY <= 0) throw std::runtime_error("cannot make GridMap with no range");
    center[0] = center_[0];
    center[1] = center_[1];
    y_range = rangeY;
    axisratio = rangeY / rangeX;
    Ngrid_init2 = (int)(Ngrid_init * axisratio);
    const size_t n = (size_t)Ngrid_init * Ngrid_init2;
    i_points.resize(n);
    s_points.resize(n);
    if (n == 0) return;
    i_points[0].head = n;
    s_points[0].head = n;
    const double xo = center[0] - 0.5 * rangeX, yo = center[1] - 0.5 * rangeY;
    const double res = rangeX / (Ngrid_init - 1);
    std::vector<double> x(2 * n), y(2 * n), A(4 * n);
    for (long ii = 0; ii < Ngrid_init; ++ii) {
      const double xx = xo + ii * res;
      for (long jj = 0; jj < Ngrid_init2; ++jj) {
        const size_t i = (size_t)(ii + jj * Ngrid_init);
        i_points[i].id = i;
        i_points[i].x[0] = xx;
        i_points[i].x[1] = yo + jj * res;
        i_points[i].gridsize = res;
        x[2 * i] = i_points[i].x[0];
        x[2 * i + 1] = i_points[i].x[1];
      }
    }
    {
      std::lock_guard<std::mutex> hold(grid_mutex());
      slsim_ctx ctx = slsim_cpp::context();
      int rc = slsim_shoot(lens->handle(), n, x.data(), nullptr, y.data(), A.data(),
                           slsim_cpp::precision(), nullptr);
      if (rc == SLSIM_OK) rc = slsim_device_alloc(ctx, 16 * n, &yd);
      if (rc == SLSIM_OK) rc = slsim_device_alloc(ctx, 32 * n, &Ad);
      if (rc == SLSIM_OK) rc = slsim_memcpy_h2d(ctx, yd, y.data(), 16 * n);
      if (rc == SLSIM_OK) rc = slsim_memcpy_h2d(ctx, Ad, A.data(), 32 * n);
      if (rc != SLSIM_OK) release_device();
      slsim_cpp::check(rc);
    }
    for (size_t i = 0; i < n; ++i) {
      s_points[i].id = i;
      s_points[i].gridsize = res;
      i_points[i].image = &s_points[i];
      s_points[i].image = &i_points[i];
      s_points[i].x[0] = y[2 * i];
      s_points[i].x[1] = y[2 * i + 1];
      for (int c = 0; c < 4; ++c) i_points[i].A.a[c] = A[4 * i + c];
      s_points[i].A = i_points[i].A;
    }
  }

  // gridmap.cpp:122-144: a grid without lensing (y = x, A = I); its outputs
  // are put in HBM too, so that the device consumers accept it
  GridMap(unsigned long N1d, const double center_[2], double range)
      : Ngrid_init(N1d), Ngrid_init2(N1d), x_range(range) {
    if (N1d < 1) throw std::runtime_error("GridMap size is < 1!");
    if (range <= 0) throw std::runtime_error("GridMap range is <= 0!");
    center[0] = center_[0];
    center[1] = center_[1];
    const size_t n = (size_t)N1d * N1d;
    i_points.resize(n);
    s_points.resize(n);
    i_points[0].head = n;
    s_points[0].head = n;
    for (size_t i = 0; i < n; ++i) {
      for (Point* p : {&i_points[i], &s_points[i]}) {
        p->id = i;
        p->gridsize = range / (N1d - 1);
        position_from_index(i, p->x);
      }
      i_points[i].image = &s_points[i];
      s_points[i].image = &i_points[i];
    }
    std::lock_guard<std::mutex> hold(grid_mutex());
    upload();
  }

  // gridmap.h:139-180: moving keeps the points (and their links) and the
  // device outputs; a copy relinks its own points and puts its own outputs
  // in HBM (the reference's copy keeps links into the original's points)
  GridMap(GridMap&& grid) noexcept
      : Ngrid_init(grid.Ngrid_init), Ngrid_init2(grid.Ngrid_init2), x_range(grid.x_range),
        y_range(grid.y_range), axisratio(grid.axisratio), rect(grid.rect),
        i_points(std::move(grid.i_points)), s_points(std::move(grid.s_points)), yd(grid.yd),
        Ad(grid.Ad) {
    center[0] = grid.center[0];
    center[1] = grid.center[1];
    grid.yd = grid.Ad = nullptr;
  }
  GridMap(GridMap& grid)
      : Ngrid_init(grid.Ngrid_init), Ngrid_init2(grid.Ngrid_init2), x_range(grid.x_range),
        y_range(grid.y_range), axisratio(grid.axisratio), rect(grid.rect),
        i_points(grid.i_points), s_points(grid.s_points) {
    center[0] = grid.center[0];
    center[1] = grid.center[1];
    for (size_t i = 0; i < i_points.size(); ++i) {
      i_points[i].image = &s_points[i];
      s_points[i].image = &i_points[i];
    }
    std::lock_guard<std::mutex> hold(grid_mutex());
    upload();
  }
  GridMap& operator=(GridMap&& grid) noexcept {
    if (this == &grid) return *this;
    release_device();
    Ngrid_init = grid.Ngrid_init;
    Ngrid_init2 = grid.Ngrid_init2;
    x_range = grid.x_range;
    y_range = grid.y_range;
    axisratio = grid.axisratio;
    rect = grid.rect;
    center[0] = grid.center[0];
    center[1] = grid.center[1];
    std::swap(i_points, grid.i_points);
    std::swap(s_points, grid.s_points);
    std::swap(yd, grid.yd);
    std::swap(Ad, grid.Ad);
    return *this;
  }
  GridMap& operator=(const GridMap&) = delete;
  ~GridMap() { release_device(); }

  // gridmap.cpp:148-160: the rays shot again through `lens` (for example
  // after its source plane changed), as the reference does it: a new grid
  // made by the rectangular constructor with this grid's size and ranges
  GridMap ReInitialize(LensHndl lens) {
    return GridMap(lens, Ngrid_init, center, x_range, getYRange());
  }

  // include/gridmap.h:182-205
  struct Triangle {
    Triangle(size_t i, size_t j, size_t k) { index[0] = i; index[1] = j; index[2] = k; }
    Triangle() { index[0] = index[1] = index[2] = 0; }
    size_t operator[](int i) const { return index[i]; }
    size_t index[3];
  };

  // gridmap.cpp:927-963, 1326-1348: the images of each source position, as
  // rays (x the image, y the source, A the mean over the image's triangle)
  std::list<RAY> find_images(std::vector<Point_2d>& ys, std::vector<int>& multiplicity) const {
    std::vector<double> r;
    std::vector<int64_t> t;
    images(ys, multiplicity, r, t);
    std::list<RAY> rays;
    for (size_t i = 0; i < r.size() / 8; ++i) {
      rays.emplace_back();
      RAY& ray = rays.back();
      ray.x = Point_2d(r[8 * i], r[8 * i + 1]);
      ray.y = Point_2d(r[8 * i + 2], r[8 * i + 3]);
      for (int k = 0; k < 4; ++k) ray.A[k] = r[8 * i + 4 + k];
    }
    return rays;
  }
  // gridmap.cpp:1034-1098
  void find_images(Point_2d y, std::vector<Point_2d>& image_points,
                   std::vector<Triangle>& triangles) const {
    std::vector<Point_2d> ys(1, y);
    std::vector<int> mult;
    std::vector<double> r;
    std::vector<int64_t> t;
    images(ys, mult, r, t);
    image_points.clear();
    triangles.clear();
    for (size_t i = 0; i < r.size() / 8; ++i) {
      image_points.emplace_back(r[8 * i], r[8 * i + 1]);
      triangles.emplace_back((size_t)t[3 * i], (size_t)t[3 * i + 1], (size_t)t[3 * i + 2]);
    }
  }

  // gridmap.cpp:525-647: tangential curves (A's smaller eigenvalue < 0),
  // radial curves (both < 0) after the tangential curve holding them, and a
  // pseudo curve after a tangential curve with none; the classification on
  // the device, the boundary walk on the host
  void find_crit(std::vector<std::vector<Point_2d>>& points, std::vector<bool>& hits_boundary,
                 std::vector<CritType>& crit_type) {
    slsim_curves h = nullptr;
    const slsim_grid gd = grid_desc();
    slsim_cpp::check(slsim_grid_find_crit(slsim_cpp::context(), &gd, (const double*)yd,
                                          (const double*)Ad, &h));
    std::vector<int32_t> t, hb;
    int replaced;
    slsim_cpp::take_curves(h, points, 0, t, hb, replaced);
    crit_type.resize(t.size());
    for (size_t i = 0; i < t.size(); ++i) crit_type[i] = (CritType)t[i];
    if (replaced) hits_boundary.assign(hb.begin(), hb.end());
  }
  // gridmap.cpp:493-522
  void find_magnification_contour(std::vector<std::vector<Point_2d>>& curves,
                                  std::vector<bool>& hits_boundary, double invmag) {
    slsim_curves h = nullptr;
    const slsim_grid gd = grid_desc();
    slsim_cpp::check(slsim_grid_find_magnification_contour(slsim_cpp::context(), &gd,
                                                           (const double*)Ad, invmag, &h));
    std::vector<int32_t> t, hb;
    int replaced;
    slsim_cpp::take_curves(h, curves, 0, t, hb, replaced);
    if (replaced) hits_boundary.assign(hb.begin(), hb.end());
  }
  // gridmap.h:278-296: boundaries (pixel units) of the grid positions with
  // more than one image as sources
  void find_boundaries_of_caustics(std::vector<std::vector<Point_2d>>& boundaries,
                                   std::vector<bool>& hits_edge) {
    slsim_curves h = nullptr;
    const slsim_grid gd = grid_desc();
    slsim_cpp::check(slsim_grid_find_boundaries_of_caustics(slsim_cpp::context(), &gd,
                                                            (const double*)yd, nullptr, &h));
    std::vector<int32_t> t, hb;
    int replaced;
    slsim_cpp::take_curves(h, boundaries, 0, t, hb, replaced);
    hits_edge.assign(hb.begin(), hb.end());
  }

  // gridmap.cpp:323-331: the area of the pixels with det A < 0
  PosType EinsteinArea() const {
    size_t count = 0;
    const size_t N = (size_t)Ngrid_init * Ngrid_init2;
    for (size_t i = 0; i < N; ++i)
      if (i_points[i].invmag() < 0) ++count;
    return count * x_range * x_range / Ngrid_init / Ngrid_init;
  }
  // gridmap.cpp:148-163: every ray unlensed (A = I, y = x), also in HBM
  void deLens() {
    const size_t n = (size_t)Ngrid_init * Ngrid_init2;
    for (size_t i = 0; i < n; ++i) {
      i_points[i].A.setToI();
      s_points[i].A.setToI();
      i_points[i].dt = 0;
      s_points[i].dt = 0;
      s_points[i].x[0] = i_points[i].x[0];
      s_points[i].x[1] = i_points[i].x[1];
    }
    upload();
  }

  size_t getNumberOfPoints() const { return (size_t)Ngrid_init * Ngrid_init2; }
  int getInitNgrid() const { return Ngrid_init; }
  double getXRange() const { return x_range; }
  double getYRange() const { return x_range * axisratio; }
  double getResolution() const { return x_range / (Ngrid_init - 1); }
  Point* operator[](size_t i) { return i_points.data() + i; }
  Point_2d image_point(size_t i) { return Point_2d(i_points[i].x[0], i_points[i].x[1]); }
  Point_2d source_point(size_t i) { return Point_2d(s_points[i].x[0], s_points[i].x[1]); }
  RAY ray(size_t i) { return RAY(i_points[i]); }
  Point_2d getCenter() const { return Point_2d(center[0], center[1]); }

  // gridmap.h:533-603
  template <typename T>
  PixelMap<T> writePixelMap(LensingVariable lensvar) {
    PixelMap<T> map(center, Ngrid_init, Ngrid_init2, x_range / (Ngrid_init - 1));
    const size_t N = map.size();
    for (size_t i = 0; i < N; ++i)
      map[i] = point_value(i_points[i], lensvar, "GridMap::writePixelMapUniform()");
    return map;
  }

  // gridmap.h:524-531
  template <typename T>
  void writeFits(LensingVariable lensvar, std::string filename) {
    PixelMap<T> map = writePixelMap<T>(lensvar);
    map.printFITS(filename);
  }

  // gridmap.h:98-106, 506-524: an Nx x Ny map of pixels x_range / (Nx - 1)
  // wide around `center`, each grid point's value put in the pixel holding
  // the point (points outside the map dropped; where two points share a
  // pixel, the later in index order is kept -- the reference's threads keep
  // either)
  template <typename T>
  PixelMap<T> writePixelMapUniform(const PosType center_[], size_t Nx, size_t Ny,
                                   LensingVariable lensvar) {
    if (getNumberOfPoints() == 0) return PixelMap<T>();
    PixelMap<T> map(center_, Nx, Ny, x_range / (Nx - 1));
    map.Clean();
    writePixelMapUniform(map, lensvar);
    return map;
  }
  // gridmap.h:605-628
  template <typename T>
  void writePixelMapUniform(PixelMap<T>& map, LensingVariable lensvar) {
    if (getNumberOfPoints() == 0) return;
    map.Clean();
    const size_t n = getNumberOfPoints();
    for (size_t i = 0; i < n; ++i) {
      const double v = point_value(i_points[i], lensvar, "PixelMap<T>::AddGrid()");
      const long index = map.find_index(i_points[i].x);
      if (index != -1) map[index] = v;
    }
  }
  // gridmap.h:691-740: the uniform map written to filename + a tag naming
  // the quantity (".kappa.fits", ...)
  template <typename T>
  void writeFitsUniform(const PosType center_[], size_t Nx, size_t Ny, LensingVariable lensvar,
                        std::string filename) {
    std::string tag;
    switch (lensvar) {
      case LensingVariable::DELAYT: tag = ".dt.fits"; break;
      case LensingVariable::ALPHA1: tag = ".alpha1.fits"; break;
      case LensingVariable::ALPHA2: tag = ".alpha2.fits"; break;
      case LensingVariable::ALPHA: tag = ".alpha.fits"; break;
      case LensingVariable::KAPPA: tag = ".kappa.fits"; break;
      case LensingVariable::GAMMA1: tag = ".gamma1.fits"; break;
      case LensingVariable::GAMMA2: tag = ".gamma2.fits"; break;
      case LensingVariable::GAMMA3: tag = ".gamma3.fits"; break;
      case LensingVariable::GAMMA: tag = ".gamma.fits"; break;
      case LensingVariable::INVMAG: tag = ".invmag.fits"; break;
      case LensingVariable::SurfBrightness: tag = ".surfbright.fits"; break;
      default: break;
    }
    PixelMap<T> map = writePixelMapUniform<T>(center_, Nx, Ny, lensvar);
    map.printFITS(filename + tag);
  }
  // gridmap.h:108-116: the grid as it is (writePixelMap), as FITS
  template <typename T>
  void writeFitsUniform(LensingVariable lensvar, std::string filename) {
    PixelMap<T> map = writePixelMap<T>(lensvar);
    map.printFITS(filename);
  }

 private:
  // gridmap.cpp:15: one process-wide lock around the shoot (reference)
  static std::mutex& grid_mutex() {
    static std::mutex m;
    return m;
  }
  void release_device() {
    slsim_ctx ctx = slsim_cpp::context();
    if (yd) slsim_device_free(ctx, yd);
    if (Ad) slsim_device_free(ctx, Ad);
    yd = Ad = nullptr;
  }
  // the points' y and A into HBM (allocated on first use)
  void upload() {
    const size_t n = (size_t)Ngrid_init * Ngrid_init2;
    if (n == 0) return;
    std::vector<double> y(2 * n), A(4 * n);
    for (size_t i = 0; i < n; ++i) {
      y[2 * i] = s_points[i].x[0];
      y[2 * i + 1] = s_points[i].x[1];
      for (int c = 0; c < 4; ++c) A[4 * i + c] = i_points[i].A.a[c];
    }
    slsim_ctx ctx = slsim_cpp::context();
    int rc = SLSIM_OK;
    if (!yd) rc = slsim_device_alloc(ctx, 16 * n, &yd);
    if (rc == SLSIM_OK && !Ad) rc = slsim_device_alloc(ctx, 32 * n, &Ad);
    if (rc == SLSIM_OK) rc = slsim_memcpy_h2d(ctx, yd, y.data(), 16 * n);
    if (rc == SLSIM_OK) rc = slsim_memcpy_h2d(ctx, Ad, A.data(), 32 * n);
    if (rc != SLSIM_OK) release_device();
    slsim_cpp::check(rc);
  }
  // Utilities::PositionFromIndex (utilities.cpp:149-158) on this grid
  void position_from_index(size_t i, double* x) const {
    if (Ngrid_init == 1) {
      x[0] = center[0];
      x[1] = center[1];
      return;
    }
    x[0] = center[0] + x_range * (1.0 * (i % Ngrid_init) / (Ngrid_init - 1) - 0.5);
    x[1] = center[1] + x_range * (1.0 * (i / Ngrid_init) / (Ngrid_init - 1) - 0.5);
  }
  // one point's lensing quantity (gridmap.h:542-600, 640-681)
  static double point_value(const Point& p, LensingVariable lensvar, const char* who) {
    const double a0 = p.x[0] - p.image->x[0], a1 = p.x[1] - p.image->x[1];
    switch (lensvar) {
      case LensingVariable::ALPHA: return std::sqrt(a0 * a0 + a1 * a1);
      case LensingVariable::ALPHA1: return a0;
      case LensingVariable::ALPHA2: return a1;
      case LensingVariable::KAPPA: return p.kappa();
      case LensingVariable::GAMMA:
        return std::sqrt(p.gamma1() * p.gamma1() + p.gamma2() * p.gamma2());
      case LensingVariable::GAMMA1: return p.gamma1();
      case LensingVariable::GAMMA2: return p.gamma2();
      case LensingVariable::GAMMA3: return p.gamma3();
      case LensingVariable::INVMAG: return p.invmag();
      case LensingVariable::DELAYT: return p.dt;
      case LensingVariable::SurfBrightness: return p.surface_brightness;
      default:
        throw std::runtime_error(std::string(who) +
                                 " does not work for the input LensingVariable");
    }
  }
  void images(std::vector<Point_2d>& ys, std::vector<int>& multiplicity, std::vector<double>& r,
              std::vector<int64_t>& t) const {
    const size_t ns = ys.size();
    std::vector<double> yv(2 * ns);
    for (size_t s = 0; s < ns; ++s) { yv[2 * s] = ys[s][0]; yv[2 * s + 1] = ys[s][1]; }
    std::vector<int32_t> m(ns);
    for (size_t cap : {size_t(64), size_t(4096)}) {
      r.assign(8 * std::max<size_t>(ns * cap, 1), 0.0);
      t.assign(3 * std::max<size_t>(ns * cap, 1), 0);
      size_t n = 0;
      const slsim_grid gd = grid_desc();
      const int rc = slsim_grid_find_images(slsim_cpp::context(), &gd, (const double*)yd,
                                            (const double*)Ad, ns, yv.data(), cap, m.data(),
                                            r.data(), t.data(), &n);
      if (rc == SLSIM_ERR_ARG && cap < 4096) continue;
      slsim_cpp::check(rc);
      r.resize(8 * n);
      t.resize(3 * n);
      break;
    }
    multiplicity.assign(m.begin(), m.end());
  }
  // the grid for the library's slsim_grid_* entry points
  slsim_grid grid_desc() const {
    slsim_grid g;
    g.nx = (size_t)Ngrid_init;
    g.ny = (size_t)Ngrid_init2;
    g.center[0] = center[0];
    g.center[1] = center[1];
    g.range_x = x_range;
    g.range_y = rect ? y_range : x_range;
    g.layout = rect ? SLSIM_GRID_RECT : SLSIM_GRID_SQUARE;
    g.pad = 0;
    return g;
  }
  int Ngrid_init, Ngrid_init2;
  double x_range;
  double y_range = 0;
  double axisratio = 1.0;
  bool rect = false;
  double center[2];
  std::vector<Point> i_points;
  std::vector<Point> s_points;
  void* yd = nullptr;
  void* Ad = nullptr;
};

// --------------------------------------------------------- adaptive grid
// Grid (TreeCode_link/grid_maintenance.cpp, include/grid_maintenance.h:24-240):
// the points stay in device memory (slsim_agrid); this object keeps host
// Point copies, image points linked to their source points, in blocks: the
// initial grid, then one block per refinement, as the reference's point
// arrays are.  Refinements shoot with the lens the grid was made with.  The
// reference's kd-trees (i_tree, s_tree) and the finders that walk them are
// not built.
struct Grid {
  Grid(LensHndl lens, unsigned long N1d, const double center[2], double range)
      : lens_(lens) {
    if (N1d < 1) throw std::runtime_error("GridMap size is < 1!");
    if (range <= 0) throw std::runtime_error("GridMap range is <= 0!");
    slsim_cpp::check(slsim_agrid_create(lens->handle(), N1d, center, range, range, 1,
                                        slsim_cpp::precision(), &g_));
    take_initial();
  }
  Grid(LensHndl lens, unsigned long Nx, const PosType center[2], PosType rangeX, PosType rangeY)
      : lens_(lens) {
    if (Nx <= 0 || rangeX <= 0 || rangeY <= 0) throw std::runtime_error("");
    slsim_cpp::check(slsim_agrid_create(lens->handle(), Nx, center, rangeX, rangeY, 0,
                                        slsim_cpp::precision(), &g_));
    take_initial();
  }
  Grid(Grid&& o) noexcept
      : lens_(o.lens_), g_(o.g_), blocks_(std::move(o.blocks_)), by_id_(std::move(o.by_id_)) {
    o.g_ = nullptr;
  }
  Grid(const Grid&) = delete;
  Grid& operator=(const Grid&) = delete;
  ~Grid() {
    if (g_) slsim_agrid_destroy(g_);
  }

  // grid_maintenance.cpp:148-158: a new unrefined grid over the same box
  Grid ReInitialize(LensHndl lens) {
    const slsim_agrid_info i = info();
    const Point_2d c = getInitCenter();
    return Grid(lens, (unsigned long)i.ngrid_init, c.x, i.image_p2[0] - i.image_p1[0],
                i.image_p2[1] - i.image_p1[1]);
  }
  // grid_maintenance.cpp:543-679
  Point* RefineLeaf(LensHndl lens, Point* point) {
    check_lens(lens);
    int64_t first = 0, count = 0;
    slsim_cpp::check(slsim_agrid_refine_one(g_, (int64_t)point->id, &first, &count));
    refresh({(int64_t)point->id});
    return take(first, count);
  }
  // grid_maintenance.cpp:681-982: all the cells' new points shot together
  Point* RefineLeaves(LensHndl lens, std::vector<Point*>& points) {
    check_lens(lens);
    if (points.empty()) return nullptr;
    std::vector<int64_t> ids(points.size());
    for (size_t k = 0; k < points.size(); ++k) ids[k] = (int64_t)points[k]->id;
    int64_t first = 0, count = 0;
    slsim_cpp::check(slsim_agrid_refine(g_, ids.size(), ids.data(), &first, &count));
    refresh(ids);
    return take(first, count);
  }
  unsigned long getNumberOfPoints() const { return (unsigned long)info().npoints; }
  // grid_maintenance.cpp:368-378
  PosType EinsteinArea() const {
    double a = 0;
    slsim_cpp::check(slsim_agrid_einstein_area(g_, &a));
    return a;
  }
  int getInitNgrid() { return (int)info().ngrid_init; }
  int getNgrid_block() { return (int)info().ngrid_block; }
  double getInitRange() {
    const slsim_agrid_info i = info();
    return i.image_p2[0] - i.image_p1[0];
  }
  // image_finder.cpp:826-831
  Point_2d getInitCenter() {
    const slsim_agrid_info i = info();
    return Point_2d((i.image_p2[0] + i.image_p1[0]) / 2, (i.image_p2[1] + i.image_p1[1]) / 2);
  }
  // grid_maintenance.cpp:984-1005
  void ClearAllMarks() {
    for (Point* p : by_id_) {
      p->in_image = p->image->in_image = NO;
      p->surface_brightness = p->image->surface_brightness = 0;
    }
  }
  // the image point of id i (initial grid first, then the refinements' points)
  Point* point(unsigned long i) { return by_id_.at(i); }
  slsim_agrid handle() const { return g_; }

 private:
  struct Block {
    std::unique_ptr<Point[]> i, s;
  };
  slsim_agrid_info info() const {
    slsim_agrid_info i;
    slsim_cpp::check(slsim_agrid_get_info(g_, &i));
    return i;
  }
  void check_lens(LensHndl lens) const {
    if (lens != lens_)
      throw std::invalid_argument("Grid: refinements shoot with the lens the grid was made with");
  }
  void take_initial() {
    const slsim_agrid_info i = info();
    take(0, i.npoints);
  }
  // host copies of points first .. first+count-1 as a new block
  Point* take(int64_t first, int64_t count) {
    if (count <= 0) return nullptr;
    const size_t n = (size_t)count;
    std::vector<double> x(2 * n), y(2 * n), A(4 * n), gs(n);
    slsim_cpp::check(slsim_agrid_read(g_, first, count, x.data(), y.data(), A.data(), gs.data()));
    Block b{std::unique_ptr<Point[]>(new Point[n]), std::unique_ptr<Point[]>(new Point[n])};
    for (size_t k = 0; k < n; ++k) {
      Point& p = b.i[k];
      Point& q = b.s[k];
      p.x[0] = x[2 * k];
      p.x[1] = x[2 * k + 1];
      q.x[0] = y[2 * k];
      q.x[1] = y[2 * k + 1];
      for (int c = 0; c < 4; ++c) p.A.a[c] = q.A.a[c] = A[4 * k + c];
      p.gridsize = q.gridsize = gs[k];
      p.id = q.id = (unsigned long)(first + (int64_t)k);
      p.image = &q;
      q.image = &p;
      p.head = q.head = n;
    }
    Point* out = b.i.get();
    for (size_t k = 0; k < n; ++k) by_id_.push_back(&b.i[k]);
    blocks_.push_back(std::move(b));
    return out;
  }
  // the refined cells' sizes, as the device left them
  void refresh(const std::vector<int64_t>& ids) {
    for (int64_t id : ids) {
      double g = 0;
      slsim_cpp::check(slsim_agrid_read(g_, id, 1, nullptr, nullptr, nullptr, &g));
      by_id_.at((size_t)id)->gridsize = by_id_.at((size_t)id)->image->gridsize = g;
    }
  }
  LensHndl lens_;
  slsim_agrid g_ = nullptr;
  std::vector<Block> blocks_;
  std::vector<Point*> by_id_;
};
typedef Grid* GridHndl;

// ------------------------------------------------------- curves of a GridMap
namespace Utilities {
// TreeCode_link/curve_routines.cpp:3519-3592: the outer boundary of the
// pixels near the closed curve v, moved onto the curve (device passes)
inline std::vector<Point_2d> TightestHull(const std::vector<Point_2d>& v) {
  std::vector<double> p(2 * v.size());
  for (size_t i = 0; i < v.size(); ++i) {
    p[2 * i] = v[i][0];
    p[2 * i + 1] = v[i][1];
  }
  slsim_curves h = nullptr;
  slsim_cpp::check(slsim_tightest_hull(slsim_cpp::context(), p.data(), v.size(), &h));
  std::vector<std::vector<Point_2d>> out;
  std::vector<int32_t> t, hb;
  int replaced;
  slsim_cpp::take_curves(h, out, 0, t, hb, replaced);
  return out.empty() ? std::vector<Point_2d>() : out[0];
}
// TreeCode_link/curve_routines.cpp:1709-1736: winding number about x and
// the absolute area of the closed curve
inline int windings(const Point_2d& x, const std::vector<Point_2d>& point, PosType* area) {
  std::vector<double> p(2 * point.size());
  for (size_t i = 0; i < point.size(); ++i) {
    p[2 * i] = point[i][0];
    p[2 * i + 1] = point[i][1];
  }
  int32_t wn = 0;
  slsim_cpp::check(slsim_windings(x.x, p.data(), point.size(), &wn, area));
  return wn;
}
// curve_routines.cpp:1738-1765: the same over the image positions of rays
inline int windings(Point_2d& x, std::vector<RAY>& point, PosType* area) {
  std::vector<Point_2d> p(point.size());
  for (size_t i = 0; i < point.size(); ++i) p[i] = point[i].x;
  return windings(x, p, area);
}
}  // namespace Utilities

// include/grid_maintenance.h:245-330 (the data members)
struct CriticalCurve {
  std::vector<RAY> critcurve;
  std::vector<Point_2d> caustic_curve_outline;
  std::vector<Point_2d> caustic_curve_intersecting;
  std::vector<Point_2d> ellipse_curve;
  PosType z_source = 0.0;
  CritType type = CritType::ND;
  int caustic_intersections = -1;
  Point_2d critical_center;
  Point_2d caustic_center;
  PosType critical_area = 0.0;
  PosType caustic_area = 0.0;
  PosType contour_ell = 0.0;
  PosType ellipse_area = 0.0;
  bool touches_edge = false;
  bool inCausticCurve(Point_2d& x) { return Utilities::inCurve(x, caustic_curve_outline); }
  bool inCriticalCurve(Point_2d& x) {
    std::vector<Point_2d> c(critcurve.size());
    for (size_t i = 0; i < c.size(); ++i) c[i] = critcurve[i].x;
    return Utilities::inCurve(x, c);
  }
};

namespace ImageFinding {
// TreeCode_link/find_crit.cpp:1036-1096: the GridMap's critical curves as
// rays to the lens's source, their centres and areas, their caustics (the
// critical points shot through the lens in one batch per curve) and the
// caustics' outlines (TightestHull) and areas
inline void find_crit(Lens& lens, GridMap& gridmap, std::vector<CriticalCurve>& critcurves,
                      bool verbose = false) {
  (void)verbose;
  const double z_source = lens.getSourceZ();
  std::vector<std::vector<Point_2d>> crits;
  std::vector<bool> hits_boundary;
  std::vector<CritType> crit_type;
  gridmap.find_crit(crits, hits_boundary, crit_type);
  const size_t Ncrit = crits.size();
  critcurves.resize(Ncrit);
  for (size_t i = 0; i < Ncrit; ++i) {
    CriticalCurve& cc = critcurves[i];
    cc.type = crit_type[i];
    cc.touches_edge = i < hits_boundary.size() ? (bool)hits_boundary[i] : false;
    cc.z_source = z_source;
    const size_t Npoints = crits[i].size();
    cc.critcurve.resize(Npoints);
    cc.critical_center = Point_2d(0, 0);
    for (size_t j = 0; j < Npoints; ++j) {
      cc.critcurve[j].x = crits[i][j];
      cc.critical_center += crits[i][j];
      cc.critcurve[j].z = z_source;
    }
    cc.critical_center /= (PosType)Npoints;
    Utilities::windings(cc.critical_center, cc.critcurve, &cc.critical_area);
    lens.rayshooterInternal(cc.critcurve.size(), cc.critcurve.data());
    cc.caustic_center = Point_2d(0, 0);
    cc.caustic_curve_intersecting.resize(Npoints);
    for (size_t j = 0; j < Npoints; ++j) {
      cc.caustic_curve_intersecting[j] = cc.critcurve[j].y;
      cc.caustic_center += cc.critcurve[j].y;
    }
    cc.caustic_center /= (PosType)Npoints;
    cc.caustic_curve_outline = Utilities::TightestHull(cc.caustic_curve_intersecting);
    Utilities::windings(cc.caustic_center, cc.caustic_curve_outline, &cc.caustic_area);
    cc.caustic_intersections = -1;
  }
}

// find_crit.cpp:1098-1119: the GridMap's magnification contours as rays
// shot to the lens's source, one batch per contour
inline void find_magnification_contour(Lens& lens, GridMap& gridmap, double invmag,
                                       std::vector<std::vector<RAY>>& contour,
                                       std::vector<bool>& hits_boundary) {
  std::vector<std::vector<Point_2d>> image_plane_contour;
  gridmap.find_magnification_contour(image_plane_contour, hits_boundary, invmag);
  const size_t n = image_plane_contour.size();
  contour.resize(n);
  for (size_t i = 0; i < n; ++i) {
    const size_t m = image_plane_contour[i].size();
    contour[i].resize(m);
    for (size_t j = 0; j < m; ++j) {
      contour[i][j].x = image_plane_contour[i][j];
      contour[i][j].z = lens.getSourceZ();
    }
    lens.rayshooterInternal(m, contour[i].data());
  }
}
}  // namespace ImageFinding
