// generator.cpp -- seeded synthetic particle planes (projected NFW halos).
//
// Workload generator for the particle ray shooter, in the spirit of
// LensHaloParticles::makeSIE (include/particle_halo.h:130-156), which makes
// particle halos for testing.  Deterministic: std::mt19937_64 with this
// file's own uniform and normal transforms (no implementation-defined
// distributions), so every host draws the same particles: halo parameters
// come from one stream, each halo's particles from streams derived from the
// seed and the halo index, independent of the thread count.

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <random>
#include <thread>
#include <vector>

#include "../../include/slsim_amd.h"

namespace {

constexpr double kPI = 3.14159265358979323846;
constexpr double kArcsec = 0.000004848136811;   // arcsecTOradians (include/standard.h:51)

struct Rng {
  std::mt19937_64 g;
  explicit Rng(uint64_t s) : g(s) {}
  double uniform() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); }   // [0, 1)
  double normal() {
    double u1 = uniform(), u2 = uniform();
    if (u1 <= 0) u1 = 1e-300;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2 * kPI * u2);
  }
};

double Ez2(double z, double Om, double OL) {
  const double a = 1 + z;
  return Om * a * a * a + OL;
}

// m(x) = ln(1+x) - x/(1+x): NFW mass inside x = r/r_s
double nfw_m(double x) { return std::log1p(x) - x / (1 + x); }

// x in [0, c] with m(x) = target: Newton on m'(x) = x/(1+x)^2, kept inside a
// shrinking bracket (bisection when a step leaves it).
double nfw_invert(double target, double c) {
  double lo = 0, hi = c;
  double x = std::min(std::sqrt(2 * target), 0.5 * c);
  for (int it = 0; it < 100; ++it) {
    const double f = nfw_m(x) - target;
    if (f < 0) lo = x; else hi = x;
    const double d = x / ((1 + x) * (1 + x));
    double xn = x - f / d;
    if (!(xn > lo && xn < hi)) xn = 0.5 * (lo + hi);
    if (std::fabs(xn - x) <= 1e-15 * xn || hi - lo <= 1e-15 * hi) return xn;
    x = xn;
  }
  return x;
}

uint64_t splitmix(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// One chunk of up to 65536 particles of one halo, drawn from its own stream.
struct Chunk {
  uint64_t seed;
  size_t off, n;
  double rs, c, xc, yc;
};

void run_chunk(const Chunk& ch, double* x, double* y) {
  Rng rng(ch.seed);
  const double mc = nfw_m(ch.c);
  for (size_t i = ch.off; i < ch.off + ch.n; ++i) {
    const double u = rng.uniform();
    const double r = ch.rs * nfw_invert(u * mc, ch.c);
    const double ct = 2 * rng.uniform() - 1;
    const double ph = 2 * kPI * rng.uniform();
    const double st = std::sqrt(std::max(0.0, 1 - ct * ct));
    x[i] = ch.xc + r * st * std::cos(ph);
    y[i] = ch.yc + r * st * std::sin(ph);
  }
}

// Halo of n particles at offset off, split into chunks; chunk k of a halo
// uses the stream splitmix(seed + k).
void add_halo_chunks(std::vector<Chunk>& out, uint64_t seed, size_t off, size_t n, double rs,
                     double c, double xc, double yc) {
  const size_t chunk = 65536;
  for (size_t k = 0; k * chunk < n; ++k)
    out.push_back({splitmix(seed + k), off + k * chunk, std::min(chunk, n - k * chunk), rs, c,
                   xc, yc});
}

// Runs the chunks on up to 16 threads (results do not depend on the count).
void run_chunks(const std::vector<Chunk>& chunks, double* x, double* y) {
  const size_t nt = std::min<size_t>(std::max(1u, std::thread::hardware_concurrency()),
                                     std::min<size_t>(16, chunks.size()));
  if (nt <= 1) {
    for (const auto& ch : chunks) run_chunk(ch, x, y);
    return;
  }
  std::atomic<size_t> next{0};
  std::vector<std::thread> th;
  for (size_t t = 0; t < nt; ++t)
    th.emplace_back([&]() {
      for (size_t k = next++; k < chunks.size(); k = next++) run_chunk(chunks[k], x, y);
    });
  for (auto& t : th) t.join();
}

}  // namespace

extern "C" {

double slsim_nfw_r200(double mass, double z, double Om, double OL, double h) {
  // rho_crit,0 = 2.77536627e11 h^2 Msun / Mpc^3
  const double rho_c = 2.77536627e11 * h * h * Ez2(z, Om, OL);
  return std::cbrt(3.0 * mass / (800.0 * kPI * rho_c));
}

int slsim_gen_nfw(uint64_t seed, size_t n, double mass, double conc, double z, double Om,
                  double OL, double h, double xc, double yc, double* x, double* y) {
  if (!x || !y || !(conc > 0) || !(mass > 0)) return SLSIM_ERR_ARG;
  const double r200 = slsim_nfw_r200(mass, z, Om, OL, h);
  std::vector<Chunk> chunks;
  add_halo_chunks(chunks, splitmix(seed), 0, n, r200 / conc, conc, xc, yc);
  run_chunks(chunks, x, y);
  return SLSIM_OK;
}

int slsim_gen_halo_field(uint64_t seed, size_t n, int n_halos, double Mmin, double Mmax,
                         double field_arcsec, double z, double Om, double OL, double h,
                         float softening, double* x, double* y, float* m, float* s) {
  if (!x || !y || !m || !s || n_halos < 1 || !(Mmin > 0) || !(Mmax >= Mmin))
    return SLSIM_ERR_ARG;
  Rng rng(seed);
  const double DA = slsim_coordist(Om, OL, h, 0.0, z) / (1 + z);
  const double half = 0.5 * field_arcsec * kArcsec * DA;
  std::vector<double> M(n_halos), c(n_halos), cx(n_halos), cy(n_halos);
  for (int k = 0; k < n_halos; ++k) {
    M[k] = std::exp(std::log(Mmin) + rng.uniform() * (std::log(Mmax) - std::log(Mmin)));
    c[k] = 10.0 * std::pow(M[k] / 1e12, -0.1) * std::pow(10.0, 0.1 * rng.normal());
    cx[k] = (2 * rng.uniform() - 1) * half;
    cy[k] = (2 * rng.uniform() - 1) * half;
  }
  const double Mtot = std::accumulate(M.begin(), M.end(), 0.0);
  // particles per halo proportional to mass, largest remainders
  std::vector<size_t> cnt(n_halos);
  std::vector<std::pair<double, int>> rem(n_halos);
  size_t used = 0;
  for (int k = 0; k < n_halos; ++k) {
    const double want = (double)n * M[k] / Mtot;
    cnt[k] = (size_t)std::floor(want);
    used += cnt[k];
    rem[k] = {want - std::floor(want), k};
  }
  std::sort(rem.begin(), rem.end(), [](const auto& a, const auto& b) {
    return a.first != b.first ? a.first > b.first : a.second < b.second;
  });
  for (size_t k = 0; used < n; ++k, ++used) cnt[rem[k % n_halos].second]++;
  const float mp = (float)(Mtot / (double)n);
  size_t off = 0;
  std::vector<Chunk> chunks;
  for (int k = 0; k < n_halos; ++k) {
    const double r200 = slsim_nfw_r200(M[k], z, Om, OL, h);
    add_halo_chunks(chunks, splitmix(seed ^ (0x5bd1e995ull * (k + 1))), off, cnt[k], r200 / c[k],
                    c[k], cx[k], cy[k]);
    off += cnt[k];
  }
  run_chunks(chunks, x, y);
  for (size_t i = 0; i < n; ++i) {
    m[i] = mp;
    s[i] = softening;
  }
  return SLSIM_OK;
}

}  // extern "C"
