// images_oracle.cpp -- CPU restatement of GridMap::find_images.
//
// TEST INFRASTRUCTURE ONLY (the checker); the product library never links or
// calls it.
//
// Restates GridMap::_find_images2_, find_images and _find_images_
// (TreeCode_link/gridmap.cpp:965-1032, 1034-1098, 1326-1348) serially, with
// Point_2d's cross product (include/point.h:129-131) and sign
// (include/point.h:42), over the grid's image positions x, source positions y
// and matrices A.

#include <cstddef>
#include <cstdint>
#include <vector>

namespace {

int sign(double v) { return (0.0 < v) - (v < 0.0); }

struct P2 {
  double a, b;
};
P2 sub(P2 u, P2 v) { return {u.a - v.a, u.b - v.b}; }
double cross(P2 u, P2 v) { return u.a * v.b - u.b * v.a; }

struct Hit {
  size_t t[3];
  P2 x;
};

// find_images(Point_2d y, image_points, triangles) for one source
// (rows j < Ngrid_init2 - 1, cells i < Ngrid_init - 1, row stride Ngrid_init)
void one_source(size_t nx, size_t ny, const double* xg, const double* yg, P2 y,
                std::vector<Hit>& out) {
  auto X = [&](size_t k) { return P2{xg[2 * k], xg[2 * k + 1]}; };
  auto S = [&](size_t k) { return P2{yg[2 * k], yg[2 * k + 1]}; };
  for (size_t j = 0; j + 1 < ny; ++j) {
    size_t k = j * nx, k1 = k + nx + 1, k2 = k + 1, k3 = k + nx;
    for (size_t i = 0; i + 1 < nx; ++i) {
      const int sig1 = sign(cross(sub(y, S(k)), sub(S(k1), S(k))));
      int sig2 = sign(cross(sub(y, S(k1)), sub(S(k2), S(k1))));
      int sig3 = sign(cross(sub(y, S(k2)), sub(S(k), S(k2))));
      int sum = sig1 + sig2 + sig3;
      auto mid = [&](size_t a, size_t b) { P2 s{X(a).a + X(b).a, X(a).b + X(b).b}; return P2{s.a / 2, s.b / 2}; };
      if (sum == 3 || sum == -3) {
        P2 s{X(k).a + X(k1).a, X(k).b + X(k1).b};
        s = {s.a + X(k2).a, s.b + X(k2).b};
        out.push_back({{k, k1, k2}, {s.a / 3, s.b / 3}});
      } else if (sum == 2 || sum == -2) {
        if (sum > 0) {
          if (sig1 == 0) out.push_back({{k, k1, k2}, mid(k, k1)});
          else if (sig2 == 0) out.push_back({{k, k1, k2}, mid(k1, k2)});
          else out.push_back({{k, k1, k2}, mid(k2, k)});
        }
      } else if (sig1 == 0 && sig3 == 0) {
        out.push_back({{k, k1, k2}, X(k)});
      }
      sig2 = sign(cross(sub(y, S(k1)), sub(S(k3), S(k1))));
      sig3 = sign(cross(sub(y, S(k3)), sub(S(k), S(k3))));
      sum = sig1 + sig2 + sig3;
      if (sum == 3 || sum == -3) {
        P2 s{X(k).a + X(k1).a, X(k).b + X(k1).b};
        s = {s.a + X(k3).a, s.b + X(k3).b};
        out.push_back({{k, k1, k3}, {s.a / 3, s.b / 3}});
      } else if (sum == 2 || sum == -2) {
        if (sum > 0) {
          if (sig1 == 0) out.push_back({{k, k1, k3}, mid(k, k1)});
          else if (sig2 == 0) out.push_back({{k, k1, k3}, mid(k1, k3)});
          else out.push_back({{k, k1, k3}, mid(k3, k)});
        }
      }
      ++k; ++k1; ++k2; ++k3;
    }
  }
}

}  // namespace

extern "C" {

// xg, yg: 2 per grid point; A: 4 per grid point; ys: 2 per source.  Writes
// up to cap rays (8 doubles) and triangles (3 indices); returns the number of
// images (all of them, also beyond cap); multiplicity per source.
// An nx x ny grid (ny rows of nx points); orc_find_images is the N x N case.
size_t orc_find_images2(size_t nx, size_t ny, const double* xg, const double* yg,
                        const double* A, const double* ys, size_t nsrc, int32_t* multiplicity,
                        double* rays, int64_t* tris, size_t cap) {
  size_t n = 0;
  for (size_t s = 0; s < nsrc; ++s) {
    std::vector<Hit> h;
    if (nx >= 2 && ny >= 2) one_source(nx, ny, xg, yg, P2{ys[2 * s], ys[2 * s + 1]}, h);
    multiplicity[s] = (int32_t)h.size();
    for (const Hit& e : h) {
      if (n < cap) {
        double* r = rays + 8 * n;
        r[0] = e.x.a; r[1] = e.x.b; r[2] = ys[2 * s]; r[3] = ys[2 * s + 1];
        for (int m = 0; m < 4; ++m) {
          double acc = 0 * 1.0;
          for (int q = 0; q < 3; ++q) acc = acc + A[4 * e.t[q] + m];
          r[4 + m] = acc / 3;
        }
        for (int q = 0; q < 3; ++q) tris[3 * n + q] = (int64_t)e.t[q];
      }
      ++n;
    }
  }
  return n;
}

size_t orc_find_images(size_t N, const double* xg, const double* yg, const double* A,
                       const double* ys, size_t nsrc, int32_t* multiplicity, double* rays,
                       int64_t* tris, size_t cap) {
  return orc_find_images2(N, N, xg, yg, A, ys, nsrc, multiplicity, rays, tris, cap);
}

}  // extern "C"
