// crit_oracle.cpp -- CPU restatement of GridMap's critical-curve routines.
//
// TEST INFRASTRUCTURE ONLY (the checker); the product library never links or
// calls it.
//
// Restates, serially and in the reference's order:
//   Matrix2x2::eigenv and det            include/point.h:315-317, 404-411
//   Utilities::find_boundaries           include/concave_hull.h:200-494
//   Utilities::find_islands              TreeCode_link/curve_routines.cpp:2670-2920
//   Utilities::convex_hull (index form)  include/concave_hull.h:560-609
//   Utilities::inCurve                   include/concave_hull.h:1022-1063
//   GridMap::find_magnification_contour  TreeCode_link/gridmap.cpp:493-522
//   GridMap::find_crit                   TreeCode_link/gridmap.cpp:525-647
//   GridMap::find_boundaries_of_caustics include/gridmap.h:278-296
//   distance_to_segment (both overloads) include/concave_hull.h:967-1001
//   Utilities::TightestHull              TreeCode_link/curve_routines.cpp:3519-3592
//   Utilities::windings (Point_2d)       TreeCode_link/curve_routines.cpp:1709-1736
//   isLeft (float result)                include/Tree.h:865-867
// The marching-squares walk is kept as the reference writes it, one branch per
// cell type (pixel k = 1, k+1 = 10, k+nx = 100, k+nx+1 = 1000), so that it
// checks the product's table-driven walk independently.

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <list>
#include <set>
#include <stdexcept>
#include <vector>

namespace {

struct P2 {
  double v[2];
  double operator[](int i) const { return v[i]; }
};

int cell_type(const std::vector<bool>& bm, size_t k, size_t nx) {
  int t = 0;
  if (bm[k]) t += 1;
  if (bm[k + 1]) t += 10;
  if (bm[k + nx]) t += 100;
  if (bm[k + nx + 1]) t += 1000;
  return t;
}

void pad(std::vector<bool>& bm, size_t nx, size_t ny) {
  for (size_t i = 0; i < nx; ++i) bm[i] = false;
  for (size_t i = 0; i < nx; ++i) bm[i + nx * (ny - 1)] = false;
  for (size_t i = 0; i < ny; ++i) bm[i * nx] = false;
  for (size_t i = 0; i < ny; ++i) bm[nx - 1 + i * nx] = false;
}

// the next not-yet-used boundary cell after `from` (n - nx when none)
size_t next_start(const std::vector<bool>& bm, const std::vector<bool>& not_used, long from,
                  size_t n, size_t nx) {
  size_t k;
  for (k = (size_t)(from + 1); k < n - nx; ++k) {
    if (k % nx == nx - 1) continue;
    const int t = cell_type(bm, k, nx);
    if (t > 0 && t != 1111 && not_used[k]) break;
  }
  return k;
}

// One step of the walk: the crossing point (pixel units) and the move.
// face: 0 came moving +x, 1 moving -y, 2 moving -x, 3 moving +y.
void step(int type, int& face, size_t& k, size_t nx, P2& p) {
  const double i = (double)(k % nx), j = (double)(k / nx);
  auto down = [&] { p = {{i + 0.5, j}}; face = 1; k -= nx; };
  auto left = [&] { p = {{i, j + 0.5}}; face = 2; k -= 1; };
  auto right = [&] { p = {{i + 1, j + 0.5}}; face = 0; k += 1; };
  auto up = [&] { p = {{i + 0.5, j + 1}}; face = 3; k += nx; };
  if (type == 0 || type == 1111) {
    throw std::runtime_error("off edge!!");
  } else if (type == 1 || type == 1110) {
    if (face == 0) down(); else left();
  } else if (type == 10 || type == 1101) {
    if (face == 2) down(); else right();
  } else if (type == 100 || type == 1011) {
    if (face == 0) up(); else left();
  } else if (type == 1000 || type == 111) {
    if (face == 1) right(); else up();
  } else if (type == 11 || type == 1100) {
    if (face == 0) right(); else left();
  } else if (type == 1010 || type == 101) {
    if (face == 1) down(); else up();
  } else if (type == 1001) {
    if (face == 0) up(); else if (face == 1) left(); else if (face == 2) down(); else right();
  } else if (type == 110) {
    if (face == 0) down(); else if (face == 1) right(); else if (face == 2) up(); else left();
  }
}

// Utilities::find_boundaries (outer_only = false)
void find_boundaries(std::vector<bool>& bm, size_t nx, std::vector<std::vector<P2>>& curves,
                     std::vector<bool>& hits, bool add) {
  const size_t n = bm.size(), ny = n / nx;
  std::vector<bool> not_used(n, true);
  pad(bm, nx, ny);
  std::list<std::list<P2>> contours;
  if (!add) hits.resize(0);
  long kfirst = -1;
  for (;;) {
    size_t k = next_start(bm, not_used, kfirst, n, nx);
    kfirst = (long)k;
    if (k == n - nx) break;
    contours.emplace_back();
    hits.push_back(false);
    int face = 0;
    size_t n_edge = 0;
    while (k != (size_t)kfirst || n_edge == 0) {
      if (n_edge >= n) throw std::runtime_error("caught in loop.");
      if (k % nx == 0 || k % nx == nx - 2) hits.back() = true;
      if (k / nx == 0 || k / nx == ny - 2) hits.back() = true;
      not_used[k] = false;
      ++n_edge;
      P2 p;
      step(cell_type(bm, k, nx), face, k, nx, p);
      contours.back().push_back(p);
    }
  }
  const size_t off = add ? curves.size() : 0;
  curves.resize(off + contours.size());
  size_t c = off;
  for (auto& l : contours) curves[c++].assign(l.begin(), l.end());
}

// Utilities::find_islands: the same walk, collecting the inside pixels of
// each boundary cell (none for type 111, whose branch tests for 1011), then
// the runs of inside pixels that start at those pixels
void find_islands(std::vector<bool>& bm, size_t nx, std::vector<std::vector<long>>& idx,
                  std::vector<bool>& hits) {
  const size_t n = bm.size(), ny = n / nx;
  std::vector<bool> not_used(n, true);
  pad(bm, nx, ny);
  hits.resize(0);
  long kfirst = -1;
  for (;;) {
    size_t k = next_start(bm, not_used, kfirst, n, nx);
    kfirst = (long)k;
    if (k == n - nx) break;
    std::set<long> contour;
    hits.push_back(false);
    int face = 0;
    size_t n_edge = 0;
    while (k != (size_t)kfirst || n_edge == 0) {
      if (k % nx == 0 || k % nx == nx - 2) hits.back() = true;
      if (k / nx == 0 || k / nx == ny - 2) hits.back() = true;
      not_used[k] = false;
      ++n_edge;
      const long k0 = (long)k, k1 = k0 + 1, k2 = k0 + (long)nx, k3 = k2 + 1;
      const int type = cell_type(bm, k, nx);
      switch (type) {
        case 1: contour.insert(k0); break;
        case 1110: contour.insert(k1); contour.insert(k2); contour.insert(k3); break;
        case 10: contour.insert(k1); break;
        case 1101: contour.insert(k0); contour.insert(k2); contour.insert(k3); break;
        case 100: contour.insert(k2); break;
        case 1011: contour.insert(k0); contour.insert(k1); contour.insert(k3); break;
        case 1000: contour.insert(k3); break;
        case 111: break;
        case 11: contour.insert(k0); contour.insert(k1); break;
        case 1100: contour.insert(k2); contour.insert(k3); break;
        case 1010: contour.insert(k1); contour.insert(k3); break;
        case 101: contour.insert(k0); contour.insert(k2); break;
        case 1001: contour.insert(k0); contour.insert(k3); break;
        case 110: contour.insert(k1); contour.insert(k2); break;
        default: break;
      }
      P2 p;
      step(type, face, k, nx, p);
    }
    idx.emplace_back();
    std::vector<long>& island = idx.back();
    long q = 0;
    for (long c : contour) {
      if (q < c) {
        q = c;
        while (bm[q]) island.push_back(q++);
      }
    }
  }
}

double crossD(const P2& O, const P2& A, const P2& B) {
  return (A[0] - O[0]) * (B[1] - O[1]) - (A[1] - O[1]) * (B[0] - O[0]);
}

// Utilities::convex_hull(P, hull): Andrew's monotone chain over indices
void convex_hull(const std::vector<P2>& P, std::vector<size_t>& hull) {
  const size_t n = P.size();
  if (n <= 3) {
    hull.resize(n);
    for (size_t i = 0; i < n; ++i) hull[i] = i;
    return;
  }
  std::vector<size_t> order(n);
  for (size_t i = 0; i < n; ++i) order[i] = i;
  std::sort(order.begin(), order.end(), [&P](size_t a, size_t b) {
    if (P[a][0] == P[b][0]) return P[a][1] < P[b][1];
    return P[a][0] < P[b][0];
  });
  hull.assign(2 * n + 1, 0);
  size_t k = 0;
  for (size_t i = 0; i < n; ++i) {
    while (k >= 2 && crossD(P[hull[k - 2]], P[hull[k - 1]], P[order[i]]) <= 0) --k;
    hull[k++] = order[i];
  }
  const size_t t = k + 1;
  for (long i = (long)n - 2; i >= 0; --i) {
    while (k >= t && crossD(P[hull[k - 2]], P[hull[k - 1]], P[order[i]]) <= 0) --k;
    hull[k++] = order[i];
  }
  hull.resize(k);
  hull.pop_back();
}

// Utilities::inCurve: winding count of the polygon H about x
bool in_curve(const P2& x, const std::vector<P2>& H) {
  if (H.size() < 3) return false;
  long w = 0;
  const size_t n = H.size();
  for (size_t i = 0; i < n; ++i) {
    const size_t j = (i + 1) % n;
    if (H[j][1] == H[i][1]) continue;
    double p1y = H[i][1] - x[1], p2y = H[j][1] - x[1];
    if (p1y * p2y < 0) {
      const double p1x = H[i][0] - x[0], p2x = H[j][0] - x[0];
      if (p1x >= 0 && p2x >= 0) {
        w += 2 * (p1y >= 0) - 1;
      } else if (p1x > 0 || p2x > 0) {
        const double s = p1x - (p2x - p1x) * p1y / (p2y - p1y);
        if (s >= 0) w += 2 * (p1y >= 0) - 1;
      }
    } else if (p2y == 0) {
      const double p2x = H[j][0] - x[0];
      if (p2x >= 0) {
        size_t k = (j + 1) % n;
        while (H[k][1] - x[1] == 0 && k != i) k = (k + 1) % n;
        if (k != i && p1y * (H[k][1] - x[1]) <= 0) w += 2 * (p1y >= 0) - 1;
      }
    }
  }
  return w != 0;
}

void eigenv(const double* a, double lam[2]) {
  const double m = (a[0] + a[3]) / 2;
  const double det = a[0] * a[3] - a[1] * a[2];
  lam[0] = m + std::sqrt(std::fabs(m * m - det));
  lam[1] = m - std::sqrt(std::fabs(m * m - det));
}

enum { kTangential = 1, kRadial = 2, kPseudo = 3 };

struct Out {
  std::vector<std::vector<P2>> curves;
  std::vector<bool> hits;
  std::vector<int> types;
  bool hits_replaced = false;
};

void rescale(Out& o, size_t nx, double range, const double* x) {
  const double res = range / (double)(nx - 1);
  for (auto& c : o.curves)
    for (P2& p : c) p = {{p[0] * res + x[0], p[1] * res + x[1]}};
}

// GridMap::find_crit on an N x ny grid (ny rows of N points)
void find_crit(size_t N, size_t ny, const double* x, const double* y, const double* A,
               double range, Out& o) {
  const size_t n = N * ny;
  std::vector<bool> bm(n);
  size_t count = 0;
  double lam[2];
  for (size_t k = 0; k < n; ++k) {
    eigenv(A + 4 * k, lam);
    bm[k] = lam[1] < 0;
    count += bm[k];
  }
  std::vector<std::vector<long>> idx;
  if (count > 0) {
    find_boundaries(bm, N, o.curves, o.hits, false);
    find_islands(bm, N, idx, o.hits);
    o.hits_replaced = true;
  }
  o.types.assign(o.curves.size(), kTangential);
  const size_t ntan = o.curves.size();
  count = 0;
  for (size_t k = 0; k < n; ++k) {
    eigenv(A + 4 * k, lam);
    bm[k] = lam[0] < 0 && lam[1] < 0;
    count += bm[k];
  }
  if (count > 0) find_boundaries(bm, N, o.curves, o.hits, true);
  const size_t m = o.curves.size();
  o.types.resize(m);
  for (size_t i = ntan; i < m; ++i) o.types[i] = kRadial;
  // a radial curve moves to just after the first tangential curve holding
  // its first point
  for (size_t j = ntan; j < m; ++j) {
    for (size_t i = 0; i < j; ++i) {
      if (o.types[i] == kTangential && in_curve(o.curves[j][0], o.curves[i])) {
        for (size_t k = j; k > i + 1; --k) {
          std::swap(o.curves[k], o.curves[k - 1]);
          std::swap(o.types[k], o.types[k - 1]);
          const bool t = o.hits[k];
          o.hits[k] = o.hits[k - 1];
          o.hits[k - 1] = t;
        }
        break;
      }
    }
  }
  rescale(o, N, range, x);
  // a tangential curve with no radial one after it gets a pseudo curve: the
  // image positions of the convex hull of its island's source positions
  size_t itan = 0;
  for (size_t j = 0; j < o.curves.size(); ++j) {
    if (o.types[j] != kTangential) continue;
    if (j == o.curves.size() - 1 || o.types[j + 1] == kTangential) {
      const std::vector<long>& isl = idx[itan];
      std::vector<P2> src(isl.size());
      for (size_t i = 0; i < isl.size(); ++i) src[i] = {{y[2 * isl[i]], y[2 * isl[i] + 1]}};
      std::vector<size_t> h;
      convex_hull(src, h);
      std::vector<P2> v(h.size());
      for (size_t i = 0; i < h.size(); ++i) v[i] = {{x[2 * isl[h[i]]], x[2 * isl[h[i]] + 1]}};
      o.curves.push_back(v);
      o.types.push_back(kPseudo);
      o.hits.push_back(false);
      for (size_t k = o.curves.size() - 1; k > j + 1; --k) {
        std::swap(o.curves[k], o.curves[k - 1]);
        std::swap(o.types[k], o.types[k - 1]);
        const bool t = o.hits[k];
        o.hits[k] = o.hits[k - 1];
        o.hits[k - 1] = t;
      }
    }
    ++itan;
  }
}

size_t emit(const Out& o, size_t cap_curves, size_t cap_points, int64_t* off, double* pts,
            int32_t* types, int32_t* hits, size_t* npoints, int* replaced) {
  size_t np = 0;
  for (auto& c : o.curves) np += c.size();
  *npoints = np;
  *replaced = o.hits_replaced;
  if (o.curves.size() > cap_curves || np > cap_points || o.hits.size() > cap_curves)
    return o.curves.size();
  size_t q = 0;
  off[0] = 0;
  for (size_t i = 0; i < o.curves.size(); ++i) {
    for (const P2& p : o.curves[i]) {
      pts[2 * q] = p[0];
      pts[2 * q + 1] = p[1];
      ++q;
    }
    off[i + 1] = (int64_t)q;
    if (types) types[i] = i < o.types.size() ? o.types[i] : 0;
  }
  for (size_t i = 0; i < o.hits.size(); ++i) hits[i] = o.hits[i];
  return o.curves.size();
}

// ------------------------------------------------------------ tightest hull
P2 sub(const P2& a, const P2& b) { return {{a[0] - b[0], a[1] - b[1]}}; }
P2 add(const P2& a, const P2& b) { return {{a[0] + b[0], a[1] + b[1]}}; }
P2 scale(const P2& a, double s) { return {{a[0] * s, a[1] * s}}; }
double dot(const P2& a, const P2& b) { return a[0] * b[0] + a[1] * b[1]; }
double length(const P2& a) { return std::sqrt(a[0] * a[0] + a[1] * a[1]); }
double length_sqr(const P2& a) { return a[0] * a[0] + a[1] * a[1]; }

double distance_to_segment(const P2& P, const P2& S1, const P2& S2) {
  const P2 D = sub(S2, S1);
  const double s = dot(sub(P, S1), D) / length_sqr(D);
  if (std::isnan(s) || s <= 0) {
    return length(sub(P, S1));
  } else if (s >= 1) {
    return length(sub(P, S2));
  } else {
    return length(sub(add(S1, scale(D, s)), P));
  }
}

double distance_to_segment(const P2& P, const P2& S1, const P2& S2, P2& closest_point) {
  const P2 D = sub(S2, S1);
  const double s = dot(sub(P, S1), D) / length_sqr(D);
  if (std::isnan(s) || s <= 0) {
    closest_point = S1;
  } else if (s >= 1) {
    closest_point = S2;
  } else {
    closest_point = add(S1, scale(D, s));
  }
  return length(sub(closest_point, P));
}

std::vector<P2> tightest_hull(const std::vector<P2>& v) {
  if (v.size() <= 3) return v;
  P2 ll = v[0], ur = v[0];
  for (const P2& p : v) {
    ll.v[0] = p[0] < ll[0] ? p[0] : ll[0];
    ll.v[1] = p[1] < ll[1] ? p[1] : ll[1];
    ur.v[0] = p[0] > ur[0] ? p[0] : ur[0];
    ur.v[1] = p[1] > ur[1] ? p[1] : ur[1];
  }
  std::vector<P2> output;
  const double w = ur[0] - ll[0], h = ur[1] - ll[1];
  const double resolution = (w < h ? w : h) / 100;
  if (resolution == 0) {
    output.push_back(ll);
    output.push_back(ur);
    return output;
  }
  const double R = resolution * std::sqrt(2.5);
  ll.v[0] -= 2 * R;
  ll.v[1] -= 2 * R;
  ur.v[0] += 2 * R;
  ur.v[1] += 2 * R;
  const long nx = (ur[0] - ll[0]) / resolution;
  const long ny = (ur[1] - ll[1]) / resolution;
  const long nv = (long)v.size();
  long count = 0;
  P2 p = ll;
  std::vector<bool> bitmap(nx * ny, false);
  for (long j = 0; j < ny; ++j, p.v[1] += resolution) {
    p.v[0] = ll[0];
    for (long i = 0; i < nx; ++i, p.v[0] += resolution) {
      const long m = i + nx * j;
      for (long k = 0; k < nv; ++k) {
        if (R > distance_to_segment(p, v[k], v[(k + 1) % nv])) {
          bitmap[m] = true;
          ++count;
          break;
        }
      }
    }
  }
  if (count > 0) {
    std::vector<std::vector<P2>> envelopes;
    std::vector<bool> hits_edge;
    find_boundaries(bitmap, nx, envelopes, hits_edge, false);
    for (P2& q : envelopes[0]) q = add(ll, scale(q, resolution));
    const size_t n = envelopes[0].size();
    output.resize(n);
    P2 q, closest_point;
    for (size_t i = 0; i < n; ++i) {
      double min_d = distance_to_segment(envelopes[0][i], v[0], v[1], closest_point);
      for (long k = 1; k < nv; ++k) {
        const double tmp_d = distance_to_segment(envelopes[0][i], v[k], v[(k + 1) % nv], q);
        if (tmp_d < min_d) {
          min_d = tmp_d;
          closest_point = q;
        }
      }
      output[i] = closest_point;
    }
  }
  return output;
}

float is_left(const P2& p0, const P2& p1, const P2& x) {
  return (p1[0] - p0[0]) * (x[1] - p0[1]) - (x[0] - p0[0]) * (p1[1] - p0[1]);
}

int windings(const P2& x, const std::vector<P2>& point, double* area) {
  int wn = 0;
  const size_t Npoints = point.size();
  *area = 0.0;
  if (Npoints < 3) return 0;
  for (size_t i = 0; i < Npoints; ++i) {
    const size_t k = i < Npoints - 1 ? i + 1 : 0;
    *area += (point[i][0] + point[k][0]) * (point[i][1] - point[k][1]);
    if (point[i][1] <= x[1]) {
      if (point[k][1] > x[1])
        if (is_left(point[i], point[k], x) > 0) ++wn;
    } else {
      if (point[k][1] <= x[1])
        if (is_left(point[i], point[k], x) < 0) --wn;
    }
  }
  *area = std::fabs(*area) * 0.5;
  return wn;
}

}  // namespace

extern "C" {

// Utilities::TightestHull of the closed curve p (2 n doubles): the hull's
// point count (written to out only when it fits cap)
size_t orc_tightest_hull(const double* p, size_t n, size_t cap, double* out) {
  std::vector<P2> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = {{p[2 * i], p[2 * i + 1]}};
  const std::vector<P2> h = tightest_hull(v);
  if (h.size() <= cap)
    for (size_t i = 0; i < h.size(); ++i) {
      out[2 * i] = h[i][0];
      out[2 * i + 1] = h[i][1];
    }
  return h.size();
}

// Utilities::windings(x, points, &area): the winding number
int orc_windings(const double* x, const double* p, size_t n, double* area) {
  std::vector<P2> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = {{p[2 * i], p[2 * i + 1]}};
  return windings({{x[0], x[1]}}, v, area);
}

// mode 0: find_crit; 1: find_magnification_contour(invmag); 2:
// find_boundaries_of_caustics from the multiplicities `mult` of the grid
// positions as sources.  x, y: grid positions and source positions (2 N^2),
// A (4 N^2).  Returns the number of curves (nothing written when a cap is
// short); hits holds *nhits flags (find_crit keeps the caller's flags when it
// finds no tangential pixel: *replaced = 0).
size_t orc_find_crit2(int mode, size_t N, size_t ny, const double* x, const double* y,
                      const double* A, const int32_t* mult, double range, double invmag,
                      size_t cap_curves, size_t cap_points, int64_t* off, double* pts,
                      int32_t* types, int32_t* hits, size_t* nhits, size_t* npoints,
                      int* replaced) {
  Out o;
  const size_t n = N * ny;
  if (mode == 0) {
    find_crit(N, ny, x, y, A, range, o);
  } else if (mode == 1) {
    std::vector<bool> bm(n);
    size_t count = 0;
    for (size_t k = 0; k < n; ++k) {
      const double* a = A + 4 * k;
      bm[k] = a[0] * a[3] - a[1] * a[2] < invmag;
      count += bm[k];
    }
    if (count > 0) {
      find_boundaries(bm, N, o.curves, o.hits, false);
      o.hits_replaced = true;
    }
    rescale(o, N, range, x);
  } else {
    std::vector<bool> bm(n);
    for (size_t k = 0; k < n; ++k) bm[k] = mult[k] > 1;
    find_boundaries(bm, N, o.curves, o.hits, false);
    o.hits_replaced = true;
  }
  *nhits = o.hits.size();
  return emit(o, cap_curves, cap_points, off, pts, types, hits, npoints, replaced);
}

size_t orc_find_crit(int mode, size_t N, const double* x, const double* y, const double* A,
                     const int32_t* mult, double range, double invmag, size_t cap_curves,
                     size_t cap_points, int64_t* off, double* pts, int32_t* types, int32_t* hits,
                     size_t* nhits, size_t* npoints, int* replaced) {
  return orc_find_crit2(mode, N, N, x, y, A, mult, range, invmag, cap_curves, cap_points, off,
                        pts, types, hits, nhits, npoints, replaced);
}

// Utilities::find_boundaries on a bitmap (pixel units), for tests of the walk
size_t orc_find_boundaries(const uint8_t* bitmap, size_t n, size_t nx, size_t cap_curves,
                           size_t cap_points, int64_t* off, double* pts, int32_t* hits,
                           size_t* npoints) {
  std::vector<bool> bm(n);
  for (size_t k = 0; k < n; ++k) bm[k] = bitmap[k] != 0;
  Out o;
  find_boundaries(bm, nx, o.curves, o.hits, false);
  int r;
  return emit(o, cap_curves, cap_points, off, pts, nullptr, hits, npoints, &r);
}

// Utilities::find_islands on a bitmap: island pixel lists, concatenated
size_t orc_find_islands(const uint8_t* bitmap, size_t n, size_t nx, size_t cap_islands,
                        size_t cap_pixels, int64_t* off, int64_t* pix, size_t* npix) {
  std::vector<bool> bm(n);
  for (size_t k = 0; k < n; ++k) bm[k] = bitmap[k] != 0;
  std::vector<std::vector<long>> idx;
  std::vector<bool> hits;
  find_islands(bm, nx, idx, hits);
  size_t tot = 0;
  for (auto& v : idx) tot += v.size();
  *npix = tot;
  if (idx.size() > cap_islands || tot > cap_pixels) return idx.size();
  size_t q = 0;
  off[0] = 0;
  for (size_t i = 0; i < idx.size(); ++i) {
    for (long p : idx[i]) pix[q++] = p;
    off[i + 1] = (int64_t)q;
  }
  return idx.size();
}

// Utilities::convex_hull(P, hull): hull indices into P (2 n doubles)
size_t orc_convex_hull(const double* p, size_t n, int64_t* hull) {
  std::vector<P2> P(n);
  for (size_t i = 0; i < n; ++i) P[i] = {{p[2 * i], p[2 * i + 1]}};
  std::vector<size_t> h;
  convex_hull(P, h);
  for (size_t i = 0; i < h.size(); ++i) hull[i] = (int64_t)h[i];
  return h.size();
}

// Utilities::inCurve for each of the m points q against the polygon h (n)
void orc_in_curve(const double* h, size_t n, const double* q, size_t m, int32_t* out) {
  std::vector<P2> H(n);
  for (size_t i = 0; i < n; ++i) H[i] = {{h[2 * i], h[2 * i + 1]}};
  for (size_t i = 0; i < m; ++i) out[i] = in_curve({{q[2 * i], q[2 * i + 1]}}, H);
}

}  // extern "C"
