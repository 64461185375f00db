// formats.cpp -- particle files in, pixel maps out.
//
// Host code of the library: the readers and writers on either side of the
// particle ray shooter.
//   * MakeParticleLenses (MultiPlane/particle_lens.cpp:22-528, include/
//     particle_halo.h:964-1080): csv3/csv4/csv5 tables, glmb particle files
//     (also the cache of smoothing sizes), Gadget-2 snapshots (include/
//     gadget.hh), centre of mass, bounding box, recentring, cuts, densest
//     particle and the unit conversion of CreateHalos.
//   * LensHaloParticles' ASCII position file and its sizes cache (include/
//     particle_halo.h:226-340, 453-632, 734-762).
//   * PixelMap::printFITS (ImageProcessing/pixelmap.cpp:1318-1388).
//
// Numbers are converted the way the reference converts them: std::stof is
// strtof plus two checks (nothing converted -> invalid_argument, ERANGE ->
// out_of_range); `istream >> T` is strtof/strtod of a token that has to be
// consumed whole.  Tokens that are plain decimal numbers take a strtof/strtod
// fast path; any other line is re-read with std::istringstream exactly as the
// reference reads it.  Text files are split into lines once and the lines are
// parsed on all host threads; results do not depend on the thread count.
// Float arithmetic keeps the reference's types (float products of float
// operands, float = float op double rounded once), so centres of mass, unit
// conversions and cuts give the same bits.  Smoothing sizes come from the
// device k-NN (slsim_smoothing_sizes_*).

#include <algorithm>
#include <atomic>
#include <cctype>
#include <cerrno>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <thread>
#include <vector>

#include <sys/stat.h>

#include "../../include/slsim_amd.h"

namespace slsim {
void set_error(const std::string& msg);
}

struct slsim_mpl_s {
  slsim_ctx ctx = nullptr;
  std::string filename;
  int nsmooth = 0;
  std::vector<slsim_particle_f32> data;
  std::vector<size_t> nparticles;
  std::vector<float> masses;
  double z_original = -1;
  double cm[3] = {0, 0, 0};
  double ll[3] = {0, 0, 0};
  double ur[3] = {0, 0, 0};
};

namespace {

using slsim::set_error;

constexpr double kPI = 3.14159265358979323846;
constexpr double kDegreesToRadians = 0.01745329251994;   // include/standard.h:55

int fail(int code, const std::string& msg) {
  set_error(msg);
  return code;
}

unsigned host_threads() {
  unsigned t = std::thread::hardware_concurrency();
  return std::max(1u, std::min(t, 16u));
}

// f(begin, end) over contiguous pieces of [0, n) on the host threads
template <class F>
void parallel_ranges(size_t n, size_t min_per_thread, F&& f) {
  unsigned nt = host_threads();
  nt = (unsigned)std::max<size_t>(1, std::min<size_t>(nt, n / std::max<size_t>(1, min_per_thread)));
  if (nt <= 1) {
    f(size_t(0), n);
    return;
  }
  std::vector<std::thread> th;
  const size_t chunk = (n + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    const size_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([&f, a, b] { f(a, b); });
  }
  for (auto& x : th) x.join();
}

bool file_exists(const std::string& p) {
  struct stat st;
  return stat(p.c_str(), &st) == 0;
}

int read_file(const std::string& path, std::string& buf) {
  FILE* f = std::fopen(path.c_str(), "rb");
  if (!f) return fail(SLSIM_ERR_IO, "Unable to open file " + path);
  if (fseeko(f, 0, SEEK_END) != 0) {
    std::fclose(f);
    return fail(SLSIM_ERR_IO, "Unable to read file " + path);
  }
  const off_t size = ftello(f);
  fseeko(f, 0, SEEK_SET);
  buf.resize((size_t)std::max<off_t>(0, size));
  const size_t got = buf.empty() ? 0 : std::fread(&buf[0], 1, buf.size(), f);
  std::fclose(f);
  if (got != buf.size()) return fail(SLSIM_ERR_IO, "Unable to read file " + path);
  return SLSIM_OK;
}

// Lines as std::getline returns them: split at '\n' (a '\r' stays in the
// line); a final '\n' ends the last line and does not start another.
struct Lines {
  const char* base = nullptr;
  std::vector<size_t> start;
  size_t last_end = 0;
  size_t count() const { return start.size(); }
  const char* ptr(size_t k) const { return base + start[k]; }
  size_t len(size_t k) const {
    return (k + 1 < start.size() ? start[k + 1] - 1 : last_end) - start[k];
  }
  std::string str(size_t k) const { return std::string(ptr(k), len(k)); }
};

Lines split_lines(const std::string& buf) {
  Lines L;
  L.base = buf.data();
  const size_t n = buf.size();
  if (n == 0) return L;
  L.last_end = buf[n - 1] == '\n' ? n - 1 : n;
  const unsigned nt = n < (1u << 22) ? 1u : host_threads();
  std::vector<std::vector<size_t>> part(nt);
  const size_t chunk = (n + nt - 1) / nt;
  auto scan = [&](unsigned t) {
    const size_t a = t * chunk, b = std::min(n, a + chunk);
    auto& v = part[t];
    const char* p = buf.data() + a;
    const char* e = buf.data() + b;
    while (p < e) {
      const char* q = static_cast<const char*>(std::memchr(p, '\n', (size_t)(e - p)));
      if (!q) break;
      const size_t next = (size_t)(q - buf.data()) + 1;
      if (next < n) v.push_back(next);
      p = q + 1;
    }
  };
  if (nt == 1) {
    scan(0);
  } else {
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; ++t) th.emplace_back(scan, t);
    for (auto& x : th) x.join();
  }
  size_t total = 1;
  for (auto& v : part) total += v.size();
  L.start.reserve(total);
  L.start.push_back(0);
  for (auto& v : part) L.start.insert(L.start.end(), v.begin(), v.end());
  return L;
}

// ---------------------------------------------------------- number parsing
// std::stof / std::stoi (libstdc++ __stoa): 0 ok, 1 invalid_argument, 2 out_of_range
int stof_like(const char* s, size_t len, float* v) {
  char small[96];
  std::string big;
  const char* p;
  if (len < sizeof small) {
    std::memcpy(small, s, len);
    small[len] = 0;
    p = small;
  } else {
    big.assign(s, len);
    p = big.c_str();
  }
  char* end;
  const int saved = errno;
  errno = 0;
  const float r = std::strtof(p, &end);
  const int e = errno;
  errno = saved;
  if (end == p) return 1;
  if (e == ERANGE) return 2;
  *v = r;
  return 0;
}

bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

// `istream >> T` for a token of digits, signs, '.', 'e' and 'E' that strto*
// consumes whole without ERANGE; false sends the caller to istringstream.
template <typename T>
bool token_number(const char* s, size_t len, T* v) {
  if (len == 0 || len >= 95) return false;
  char buf[96];
  for (size_t i = 0; i < len; ++i) {
    const char c = s[i];
    if (!((c >= '0' && c <= '9') || c == '+' || c == '-' || c == '.' || c == 'e' || c == 'E'))
      return false;
    buf[i] = c;
  }
  buf[len] = 0;
  char* end;
  const int saved = errno;
  errno = 0;
  T r;
  if (sizeof(T) == sizeof(float))
    r = (T)std::strtof(buf, &end);
  else
    r = (T)std::strtod(buf, &end);
  const int e = errno;
  errno = saved;
  if (end != buf + len || e != 0) return false;
  *v = r;
  return true;
}

// First `want` whitespace-separated tokens of a line as numbers of type T;
// false when the fast path does not apply.
template <typename T>
bool fast_tokens(const char* s, size_t len, int want, T* out) {
  size_t i = 0;
  for (int k = 0; k < want; ++k) {
    while (i < len && is_space(s[i])) ++i;
    const size_t a = i;
    while (i < len && !is_space(s[i])) ++i;
    if (!token_number<T>(s + a, i - a, &out[k])) return false;
  }
  return true;
}

// First error of a parallel parse: the lowest line wins, as in a serial read.
struct FirstError {
  std::atomic<size_t> line{SIZE_MAX};
  void note(size_t k) {
    size_t cur = line.load();
    while (k < cur && !line.compare_exchange_weak(cur, k)) {
    }
  }
};

// ------------------------------------------------------------------- csv
// MakeParticleLenses::readCSV (particle_lens.cpp:211-342) for 3, 4 and 5
// columns: leading '#' lines are skipped, the first data line must have at
// least `columns` comma-separated fields, every later line is data.
int read_csv(const std::string& filename, int columns, std::vector<slsim_particle_f32>& data,
             std::vector<size_t>& nparticles, std::vector<float>& masses) {
  std::string buf;
  if (!file_exists(filename))
    return fail(SLSIM_ERR_ARG, "file " + filename + " does not have enough columns (cannot open it).");
  int rc = read_file(filename, buf);
  if (rc) return rc;
  const Lines L = split_lines(buf);
  size_t f = 0;
  while (f < L.count() && L.len(f) > 0 && L.ptr(f)[0] == '#') ++f;
  size_t ncol = 1;
  if (f < L.count()) ncol += (size_t)std::count(L.ptr(f), L.ptr(f) + L.len(f), ',');
  if (ncol < (size_t)columns)
    return fail(SLSIM_ERR_ARG, "file " + filename + " does not have enough columns.");
  const size_t ntot = L.count() - f;
  data.assign(ntot, slsim_particle_f32{});
  FirstError err;
  parallel_ranges(ntot, 8192, [&](size_t a, size_t b) {
    float v[5];
    for (size_t k = a; k < b; ++k) {
      const char* s = L.ptr(f + k);
      const size_t len = L.len(f + k);
      size_t pos = 0;
      int bad = 0;
      for (int c = 0; c < columns && !bad; ++c) {
        const char* comma = static_cast<const char*>(std::memchr(s + pos, ',', len - pos));
        const size_t e = comma ? (size_t)(comma - s) : len;
        if (!comma && c + 1 < columns) bad = 3;   // fewer fields than the first line promised
        else bad = stof_like(s + pos, e - pos, &v[c]);
        pos = e + 1;
      }
      if (bad) {
        err.note(k);
        continue;
      }
      slsim_particle_f32& p = data[k];
      p.x[0] = v[0];
      p.x[1] = v[1];
      p.x[2] = v[2];
      p.Mass = columns >= 4 ? v[3] : 1.0f;
      if (columns >= 5) p.Size = v[4];
      p.type = 0;
    }
  });
  if (err.line.load() != SIZE_MAX) {
    const size_t k = err.line.load();
    // re-derive the reason for the first bad line serially
    const char* s = L.ptr(f + k);
    const size_t len = L.len(f + k);
    size_t pos = 0;
    int bad = 0;
    float v;
    for (int c = 0; c < columns && !bad; ++c) {
      const char* comma = static_cast<const char*>(std::memchr(s + pos, ',', len - pos));
      const size_t e = comma ? (size_t)(comma - s) : len;
      bad = (!comma && c + 1 < columns) ? 3 : stof_like(s + pos, e - pos, &v);
      pos = e + 1;
    }
    const char* what = bad == 1 ? "stof: no conversion" : bad == 2 ? "stof: out of range"
                                                                    : "fewer columns than the first data line";
    return fail(bad == 3 ? SLSIM_ERR_IO : SLSIM_ERR_ARG,
                "file " + filename + ", data line " + std::to_string(k + 1) + ": " + what);
  }
  nparticles.assign(6, 0);
  nparticles[0] = ntot;
  masses.assign(6, 0.0f);
  return SLSIM_OK;
}

// ------------------------------------------------------------------ glmb
// writeSizesB / readSizesB (particle_halo.h:1037-1073)
int glmb_write(const std::string& filename, const slsim_particle_f32* p, size_t n, int nsmooth,
               const size_t numbytype[6], double redshift) {
  FILE* f = std::fopen(filename.c_str(), "wb");
  if (!f) return fail(SLSIM_ERR_IO, "Unable to write to file " + filename);
  const int32_t ns = nsmooth;
  const uint64_t nt = n;
  uint64_t nb[6];
  for (int i = 0; i < 6; ++i) nb[i] = numbytype[i];
  bool ok = std::fwrite(&ns, 4, 1, f) == 1 && std::fwrite(&nt, 8, 1, f) == 1 &&
            std::fwrite(nb, 8, 6, f) == 6 && std::fwrite(&redshift, 8, 1, f) == 1 &&
            (n == 0 || std::fwrite(p, sizeof(slsim_particle_f32), n, f) == n);
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(SLSIM_ERR_IO, "Unable to write to file " + filename);
  return SLSIM_OK;
}

constexpr long kGlmbHeader = 4 + 8 + 48 + 8;

// 1 = the first read failed (the reference's `false`), 0 = ok, else an error code
int glmb_header(const std::string& filename, int* nsmooth, size_t* n, size_t numbytype[6],
                double* redshift) {
  FILE* f = std::fopen(filename.c_str(), "rb");
  if (!f) return -1;
  int32_t ns;
  uint64_t nt, nb[6];
  double z;
  if (std::fread(&ns, 4, 1, f) != 1) {
    std::fclose(f);
    return -1;
  }
  const bool ok = std::fread(&nt, 8, 1, f) == 1 && std::fread(nb, 8, 6, f) == 6 &&
                  std::fread(&z, 8, 1, f) == 1;
  std::fclose(f);
  if (!ok) return fail(SLSIM_ERR_IO, "glmb file " + filename + " ends inside its header");
  *nsmooth = ns;
  *n = nt;
  for (int i = 0; i < 6; ++i) numbytype[i] = nb[i];
  *redshift = z;
  return SLSIM_OK;
}

int glmb_records(const std::string& filename, slsim_particle_f32* p, size_t n) {
  FILE* f = std::fopen(filename.c_str(), "rb");
  if (!f) return fail(SLSIM_ERR_IO, "Unable to open file " + filename);
  bool ok = fseeko(f, kGlmbHeader, SEEK_SET) == 0 &&
            (n == 0 || std::fread(p, sizeof(slsim_particle_f32), n, f) == n);
  std::fclose(f);
  if (!ok) return fail(SLSIM_ERR_IO, "glmb file " + filename + " holds fewer particles than its header says");
  return SLSIM_OK;
}

int glmb_read(slsim_mpl_s& m, const std::string& filename, bool* found) {
  size_t n = 0, nb[6];
  int rc = glmb_header(filename, &m.nsmooth, &n, nb, &m.z_original);
  if (rc == -1) {
    *found = false;
    return SLSIM_OK;
  }
  if (rc) return rc;
  *found = true;
  m.nparticles.assign(nb, nb + 6);
  m.data.assign(n, slsim_particle_f32{});
  return glmb_records(filename, m.data.data(), n);
}

// -------------------------------------------------------------- Gadget-2
// GadgetFile (include/gadget.hh) as readGadget2 uses it (particle_lens.cpp:
// 344-389): SnapFormat 2 (blocks behind 8-byte label records), either byte
// order (a first record marker that reads 134217728 means swapped), one file;
// POS and MASS blocks; types in block order.
struct GadgetHead {
  int32_t npart[6];
  double masstab[6];
  double time, redshift, omega, lambda, hubble;
};

struct Snap {
  const std::string& buf;
  bool swap = false;
  uint32_t u32(size_t at) const {
    uint32_t v;
    std::memcpy(&v, buf.data() + at, 4);
    return swap ? __builtin_bswap32(v) : v;
  }
  double f64(size_t at) const {
    uint64_t v;
    std::memcpy(&v, buf.data() + at, 8);
    if (swap) v = __builtin_bswap64(v);
    double d;
    std::memcpy(&d, &v, 8);
    return d;
  }
  float f32(size_t at) const {
    uint32_t v = u32(at);
    float d;
    std::memcpy(&d, &v, 4);
    return d;
  }
  // find_block (gadget.hh:374-418): offset of the block's data (after its
  // leading record marker) and its size in bytes
  int find(const char* label, const std::string& fn, size_t* at, size_t* bytes) const {
    size_t pos = 0;
    for (;;) {
      if (pos + 4 > buf.size())
        return fail(SLSIM_ERR_IO, std::string("Block <") + label + "> not found in " + fn);
      if (u32(pos) != 8)
        return fail(SLSIM_ERR_IO, "incorrect format (blksize=" + std::to_string((int32_t)u32(pos)) +
                                      ") in " + fn + ": not a SnapFormat 2 Gadget file");
      if (pos + 16 > buf.size()) return fail(SLSIM_ERR_IO, "Gadget file " + fn + " is truncated");
      const int32_t bs = (int32_t)u32(pos + 8);
      const bool hit = std::memcmp(buf.data() + pos + 4, label, 4) == 0;
      pos += 16;
      if (hit) {
        if (bs < 8 || pos + (size_t)bs > buf.size())
          return fail(SLSIM_ERR_IO, std::string("Block <") + label + "> of " + fn + " is truncated");
        *at = pos + 4;
        *bytes = (size_t)bs - 8;
        return SLSIM_OK;
      }
      if (bs < 0) return fail(SLSIM_ERR_IO, "incorrect block size in " + fn);
      pos += (size_t)bs;
    }
  }
};

// GadgetFile(filename, data), openFile, readBlock("POS"), readBlock("MASS"):
// header and the raw particles (file units, types in block order).  With
// p == nullptr only the header is read.
int gadget2_raw(const std::string& fn, GadgetHead& h, std::vector<slsim_particle_f32>* out) {
  if (!file_exists(fn)) {
    if (file_exists(fn + ".0"))
      return fail(SLSIM_ERR_IO, "Gadget-2 snapshot " + fn + " is split into files " + fn +
                                    ".0, ...: only single-file snapshots are read");
    return fail(SLSIM_ERR_IO, "Cant open file <" + fn + ">");
  }
  std::string buf;
  int rc = read_file(fn, buf);
  if (rc) return rc;
  Snap s{buf};
  if (buf.size() < 4) return fail(SLSIM_ERR_IO, "Gadget file " + fn + " is empty");
  if (s.u32(0) == 134217728u) s.swap = true;
  size_t at, bytes;
  if ((rc = s.find("HEAD", fn, &at, &bytes))) return rc;
  if (at + 160 > buf.size()) return fail(SLSIM_ERR_IO, "Gadget header of " + fn + " is truncated");
  for (int i = 0; i < 6; ++i) h.npart[i] = (int32_t)s.u32(at + 4 * i);
  for (int i = 0; i < 6; ++i) h.masstab[i] = s.f64(at + 24 + 8 * i);
  h.time = s.f64(at + 72);
  h.redshift = s.f64(at + 80);
  h.omega = s.f64(at + 136);
  h.lambda = s.f64(at + 144);
  h.hubble = s.f64(at + 152);
  size_t ntot = 0, nmass = 0;
  for (int i = 0; i < 6; ++i) {
    if (h.npart[i] < 0 || !(h.masstab[i] >= 0))
      return fail(SLSIM_ERR_IO, "Gadget header of " + fn + " has a negative count or mass");
    ntot += (size_t)h.npart[i];
    if (h.masstab[i] == 0.0) nmass += (size_t)h.npart[i];
  }
  if (!out) return SLSIM_OK;
  out->assign(ntot, slsim_particle_f32{});
  auto& d = *out;
  {
    size_t k = 0;
    for (int t = 0; t < 6; ++t)
      for (int32_t i = 0; i < h.npart[t]; ++i) d[k++].type = t;
  }
  if ((rc = s.find("POS ", fn, &at, &bytes))) return rc;
  if (bytes != 12 * ntot)
    return fail(SLSIM_ERR_IO, "POS block of " + fn + " holds " + std::to_string(bytes) +
                                  " bytes for " + std::to_string(ntot) + " particles");
  parallel_ranges(ntot, 65536, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i)
      for (int k = 0; k < 3; ++k) d[i].x[k] = s.f32(at + 12 * i + 4 * k);
  });
  if ((rc = s.find("MASS", fn, &at, &bytes))) return rc;
  if (bytes != 4 * nmass)
    return fail(SLSIM_ERR_IO, "MASS block of " + fn + " holds " + std::to_string(bytes) +
                                  " bytes for " + std::to_string(nmass) + " individual masses");
  const float umass = 1.0f;
  size_t ic = 0;
  for (size_t i = 0; i < ntot; ++i) {
    const int t = d[i].type;
    if (h.masstab[t] > 0.0) d[i].Mass = (float)(h.masstab[t] * umass);
    else d[i].Mass = s.f32(at + 4 * ic++) * umass;
  }
  return SLSIM_OK;
}

// MakeParticleLenses::readGadget2 (particle_lens.cpp:344-389) before smoothing
int read_gadget2(slsim_mpl_s& m) {
  GadgetHead h;
  int rc = gadget2_raw(m.filename, h, &m.data);
  if (rc) return rc;
  auto& d = m.data;
  // readGadget2: comoving kpc/h -> physical Mpc/h, 1e10 Msun/h -> Msun/h
  m.z_original = h.redshift;
  const double a = 1.0e-3 / (1 + m.z_original);
  for (auto& p : d) {
    p.x[0] = (float)(p.x[0] * a);
    p.x[1] = (float)(p.x[1] * a);
    p.x[2] = (float)(p.x[2] * a);
    p.Mass = (float)(p.Mass * 1.0e10);
  }
  std::sort(d.begin(), d.end(), [](const slsim_particle_f32& a1, const slsim_particle_f32& a2) {
    return a1.type < a2.type;
  });
  m.nparticles.clear();
  m.masses.clear();
  for (int i = 0; i < 6; ++i) {
    m.nparticles.push_back((size_t)h.npart[i]);
    m.masses.push_back((float)h.masstab[i]);
  }
  return SLSIM_OK;
}

// ---------------------------------------------------------- smoothing
template <typename T>
int smooth_range_T(slsim_ctx ctx, const T* xyz, size_t n, int nsmooth, float* size) {
  std::vector<T> x(n), y(n), z(n);
  for (size_t i = 0; i < n; ++i) {
    x[i] = xyz[3 * i];
    y[i] = xyz[3 * i + 1];
    z[i] = xyz[3 * i + 2];
  }
  if (sizeof(T) == sizeof(float))
    return slsim_smoothing_sizes_f32(ctx, (const float*)x.data(), (const float*)y.data(),
                                     (const float*)z.data(), n, nsmooth, size);
  return slsim_smoothing_sizes_f64(ctx, (const double*)x.data(), (const double*)y.data(),
                                   (const double*)z.data(), n, nsmooth, size);
}

// LensHaloParticles<ParticleType<float>>::calculate_smoothing over data[a, a+n)
int smooth_particles(slsim_ctx ctx, slsim_particle_f32* p, size_t n, int nsmooth) {
  std::vector<float> xyz(3 * n), size(n);
  for (size_t i = 0; i < n; ++i)
    for (int k = 0; k < 3; ++k) xyz[3 * i + k] = p[i].x[k];
  const int rc = smooth_range_T<float>(ctx, xyz.data(), n, nsmooth, size.data());
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i) p[i].Size = size[i];
  return SLSIM_OK;
}

// centre of mass and bounding box (particle_lens.cpp:76-107, 120-131)
void mass_centre(slsim_mpl_s& m, bool box_from_first) {
  for (int i = 0; i < 3; ++i) m.cm[i] = 0;
  if (box_from_first && !m.data.empty())
    for (int i = 0; i < 3; ++i) m.ll[i] = m.ur[i] = m.data[0].x[i];
  double mass = 0;
  for (const auto& p : m.data) {
    for (int i = 0; i < 3; ++i) {
      const float w = p.x[i] * p.Mass;   // float * float, as there
      m.cm[i] += w;
      if (m.ll[i] > p.x[i]) m.ll[i] = p.x[i];
      if (m.ur[i] < p.x[i]) m.ur[i] = p.x[i];
    }
    mass += p.Mass;
  }
  for (int i = 0; i < 3; ++i) m.cm[i] /= mass;
}

void recenter(slsim_mpl_s& m, const double x[3]) {
  const double c[3] = {x[0], x[1], x[2]};
  for (auto& p : m.data)
    for (int i = 0; i < 3; ++i) p.x[i] = (float)(p.x[i] - c[i]);
  for (int i = 0; i < 3; ++i) {
    m.ll[i] -= c[i];
    m.ur[i] -= c[i];
    m.cm[i] -= c[i];
  }
}

void sort_by_type_if_mixed(slsim_mpl_s& m) {
  int ntypes = 0;
  for (size_t n : m.nparticles)
    if (n > 0) ++ntypes;
  if (ntypes > 1)
    std::sort(m.data.begin(), m.data.end(),
              [](const slsim_particle_f32& a1, const slsim_particle_f32& a2) { return a1.type < a2.type; });
}

template <class Far>
void cut(slsim_mpl_s& m, Far far) {
  size_t end = m.data.size(), it = 0, ntot = m.data.size();
  while (it != end) {
    if (far(m.data[it])) {
      --m.nparticles[m.data[it].type];
      --end;
      --ntot;
      std::swap(m.data[it], m.data[end]);
    } else {
      ++it;
    }
  }
  m.data.resize(ntot);
  sort_by_type_if_mixed(m);
}

// ------------------------------------------------------------- ASCII halo
// readPositionFileASCII's header (particle_halo.h:470-501), read as there
int ascii_header(std::istream& in, const std::string& fn, bool multimass, size_t* npoints) {
  std::string str, label;
  int count = 0;
  size_t np = 0;
  while (std::getline(in, str)) {
    std::stringstream ss(str);
    ss >> label;
    if (label == "#") {
      ss >> label;
      if (label == "nparticles") {
        ss >> np;
        ++count;
      }
      if (!multimass) {
        if (label == "mass") {
          float tmp_mass;
          ss >> tmp_mass;
          ++count;
        }
      }
    } else {
      break;
    }
    if (multimass && count == 1) break;
    if (!multimass && count == 2) break;
  }
  if (count == 0)
    return fail(SLSIM_ERR_IO, "File " + fn + " must have the header line: # nparticles ****");
  *npoints = np;
  return SLSIM_OK;
}

// the rest of the stream after the header, split into lines
int rest_lines(std::istream& in, std::string& buf) {
  std::ostringstream o;
  o << in.rdbuf();
  buf = o.str();
  return SLSIM_OK;
}

template <typename T>
int ascii_positions(const std::string& fn, bool multimass, size_t n, T* xyz, float* mass) {
  std::ifstream in(fn);
  if (!in.is_open()) return fail(SLSIM_ERR_IO, "Unable to open file " + fn);
  size_t np;
  int rc = ascii_header(in, fn, multimass, &np);
  if (rc) return rc;
  if (np != n)
    return fail(SLSIM_ERR_ARG, "file " + fn + " holds " + std::to_string(np) + " particles, not " +
                                   std::to_string(n));
  std::string buf;
  if (in.peek() != EOF) rest_lines(in, buf);
  const Lines L = split_lines(buf);
  // rows: every line not starting with '#', until n rows
  std::vector<size_t> rows;
  rows.reserve(n);
  for (size_t k = 0; k < L.count() && rows.size() < n; ++k)
    if (!(L.len(k) > 0 && L.ptr(k)[0] == '#')) rows.push_back(k);
  if (rows.size() != n)
    return fail(SLSIM_ERR_IO, "Number of data rows in " + fn +
                                  " does not match expected number of particles.");
  const int ncol = multimass ? 4 : 3;
  parallel_ranges(n, 8192, [&](size_t a, size_t b) {
    for (size_t r = a; r < b; ++r) {
      const size_t k = rows[r];
      T v[3];
      float mv = 0;
      bool ok = fast_tokens<T>(L.ptr(k), L.len(k), 3, v);
      if (ok && multimass) {
        // the fourth token, as float
        size_t i = 0, len = L.len(k);
        const char* s = L.ptr(k);
        for (int t = 0; t < 3; ++t) {
          while (i < len && is_space(s[i])) ++i;
          while (i < len && !is_space(s[i])) ++i;
        }
        while (i < len && is_space(s[i])) ++i;
        const size_t a4 = i;
        while (i < len && !is_space(s[i])) ++i;
        ok = token_number<float>(s + a4, i - a4, &mv);
      }
      if (!ok) {
        // the reference's own extraction; a failed field reads as 0
        std::istringstream ss(L.str(k));
        T p[3] = {0, 0, 0};
        float pm = 0;
        ss >> p[0];
        ss >> p[1];
        ss >> p[2];
        if (multimass) ss >> pm;
        for (int c = 0; c < 3; ++c) v[c] = p[c];
        mv = pm;
      }
      for (int c = 0; c < 3; ++c) xyz[3 * r + c] = v[c];
      mass[r] = multimass ? mv : 0.0f;
    }
  });
  (void)ncol;
  return SLSIM_OK;
}

// readSizesFile (particle_halo.h:557-632): 0 = read, 1 = absent or another
// Nsmooth (the reference's `false`), else an error code
int sizes_read(const std::string& fn, size_t n, int nsmooth, double min_size, float* size) {
  std::ifstream in(fn);
  if (!in.is_open()) return 1;
  std::string str, label;
  int count = 0;
  while (std::getline(in, str)) {
    std::stringstream ss(str);
    ss >> label;
    if (label == "#") {
      ss >> label;
      if (label == "nparticles") {
        size_t ntmp = 0;
        ss >> ntmp;
        if (ntmp != n)
          return fail(SLSIM_ERR_IO, "Number of particles in " + fn + " does not match expected number");
        ++count;
      }
      if (label == "nsmooth") {
        int nstmp = 0;
        ss >> nstmp;
        if (nstmp != nsmooth) return 1;
        ++count;
      }
    } else {
      break;
    }
    if (count == 2) break;
  }
  if (count != 2) return fail(SLSIM_ERR_IO, "File " + fn + " must have the header lines: # nparticles ****");
  std::string buf;
  if (in.peek() != EOF) rest_lines(in, buf);
  const Lines L = split_lines(buf);
  std::vector<size_t> rows;
  rows.reserve(n);
  for (size_t k = 0; k < L.count(); ++k)
    if (!(L.len(k) > 0 && L.ptr(k)[0] == '#')) {
      if (rows.size() == n)
        return fail(SLSIM_ERR_IO, "Number of data rows in " + fn +
                                      " does not match expected number of particles.");
      rows.push_back(k);
    }
  if (rows.size() != n)
    return fail(SLSIM_ERR_IO, "Number of data rows in " + fn +
                                  " does not match expected number of particles.");
  parallel_ranges(n, 8192, [&](size_t a, size_t b) {
    for (size_t r = a; r < b; ++r) {
      const size_t k = rows[r];
      float v;
      if (!fast_tokens<float>(L.ptr(k), L.len(k), 1, &v)) {
        std::istringstream ss(L.str(k));
        v = 0;
        ss >> v;
      }
      if (min_size > v) v = (float)min_size;
      size[r] = v;
    }
  });
  return 0;
}

// writeSizes (particle_halo.h:734-762): `ostream << float` is "%.6g"
int sizes_write(const std::string& fn, int nsmooth, const float* size, size_t n) {
  FILE* f = std::fopen(fn.c_str(), "w");
  if (!f) return fail(SLSIM_ERR_IO, "Unable to write to file " + fn);
  std::string out = "# nparticles " + std::to_string(n) + "\n# nsmooth " + std::to_string(nsmooth) + "\n";
  char line[64];
  for (size_t i = 0; i < n; ++i) {
    const int k = std::snprintf(line, sizeof line, "%.6g\n", (double)size[i]);
    out.append(line, (size_t)k);
  }
  bool ok = std::fwrite(out.data(), 1, out.size(), f) == out.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(SLSIM_ERR_IO, "Unable to write to file " + fn);
  return SLSIM_OK;
}

template <typename T>
int ascii_halo_load(slsim_ctx ctx, const char* filename, double redshift, int nsmooth,
                    const double theta_rotate[2], int recenter_, int multimass, double min_size,
                    double rescale_mass, size_t n, T* xyz, float* mass, float* size,
                    slsim_ascii_halo_info* info) {
  if (!filename || !xyz || !mass || !size || !info || n == 0)
    return fail(SLSIM_ERR_ARG, "ascii halo: null argument or no particles");
  const std::string fn(filename);
  int rc = ascii_positions<T>(fn, multimass != 0, n, xyz, mass);
  if (rc) return rc;
  for (size_t i = 0; i < n; ++i) size[i] = 0.0f;
  const std::string sizefile = fn + "." + std::to_string(nsmooth) + "sizes";
  rc = sizes_read(sizefile, n, nsmooth, min_size, size);
  if (rc > 1) return rc;
  if (rc == 1) {
    if ((rc = smooth_range_T<T>(ctx, xyz, n, nsmooth, size))) return rc;
    if ((rc = sizes_write(sizefile, nsmooth, size, n))) return rc;
    for (size_t i = 0; i < n; ++i)
      if (size[i] < min_size) size[i] = (float)min_size;
  }
  // particle_halo.h:275-291: the first particle keeps its mass
  double min_s = size[0];
  size_t smallest = 0;
  for (size_t i = 1; i < n; ++i) {
    if (size[i] < min_s) {
      min_s = size[i];
      smallest = i;
    }
    mass[i] = (float)(mass[i] * rescale_mass);
  }
  double dp[3] = {(double)xyz[3 * smallest], (double)xyz[3 * smallest + 1], (double)xyz[3 * smallest + 2]};
  const double sf = 1 / (1 + redshift);
  double mc[3] = {0, 0, 0}, mtot = 0;
  for (size_t i = 0; i < n; ++i) {
    for (int k = 0; k < 3; ++k) xyz[3 * i + k] = (T)(xyz[3 * i + k] * sf);
    const float mi = mass[multimass ? i : 0];
    for (int k = 0; k < 3; ++k) {
      const T w = xyz[3 * i + k] * (T)mi;
      mc[k] += w;
    }
    mtot += mi;
  }
  for (int k = 0; k < 3; ++k) mc[k] /= mtot;
  double rsize = 1.0e3;
  if (recenter_) {
    double r2max = 0;
    for (size_t i = 0; i < n; ++i) {
      T* p = xyz + 3 * i;
      for (int k = 0; k < 3; ++k) p[k] = (T)(p[k] - mc[k]);
      const T r2t = p[0] * p[0] + p[1] * p[1] + p[2] * p[2];
      const double r2 = r2t;
      if (r2 > r2max) r2max = r2;
    }
    rsize = std::sqrt(r2max);
    for (int k = 0; k < 3; ++k) {
      dp[k] -= mc[k];
      mc[k] = 0;
    }
  }
  // rotate_particles (particle_halo.h:634-680)
  const double tx = theta_rotate ? theta_rotate[0] : 0.0, ty = theta_rotate ? theta_rotate[1] : 0.0;
  if (!(tx == 0.0 && ty == 0.0)) {
    const double cx = std::cos(tx), sx = std::sin(tx), cy = std::cos(ty), sy = std::sin(ty);
    const double c[3][3] = {{cy, 0, -sy}, {-sy * sx, cx, -cy * sx}, {cx, sx, cy * cx}};
    auto rot = [&](double* v) {
      double t[3] = {0, 0, 0};
      for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r) t[r] += c[r][j] * v[j];
      for (int j = 0; j < 3; ++j) v[j] = t[j];
    };
    for (size_t i = 0; i < n; ++i) {
      T* p = xyz + 3 * i;
      double t[3] = {0, 0, 0};
      for (int j = 0; j < 3; ++j)
        for (int r = 0; r < 3; ++r) t[r] += c[r][j] * p[j];
      for (int j = 0; j < 3; ++j) p[j] = (T)t[j];
    }
    rot(mc);
    rot(dp);
  }
  info->mass = mtot;
  for (int k = 0; k < 3; ++k) {
    info->mcenter[k] = mc[k];
    info->densest_point[k] = dp[k];
  }
  info->rsize = rsize;
  return SLSIM_OK;
}

// ------------------------------------------------------------------ FITS
// FITS standard 4.0: 2880-byte records of 36 cards of 80 characters; fixed-
// format values end in column 30; data big-endian IEEE, zero-padded.
std::string pad80(std::string c) {
  if (c.size() > 80) c.resize(80);
  c.append(80 - c.size(), ' ');
  return c;
}

bool standard_key(const std::string& k) {
  if (k.empty() || k.size() > 8) return false;
  for (char ch : k)
    if (!((ch >= 'A' && ch <= 'Z') || (ch >= 'a' && ch <= 'z') || (ch >= '0' && ch <= '9') || ch == '-' ||
          ch == '_'))
      return false;
  return true;
}

std::string card(const std::string& key, const std::string& value, const std::string& comment,
                 bool is_string) {
  std::string c;
  if (standard_key(key)) {
    std::string k = key;
    for (auto& ch : k) ch = (char)std::toupper((unsigned char)ch);
    c = k;
    c.append(8 - k.size(), ' ');
    c += "= ";
    if (is_string) {
      c += value;
    } else {
      if (value.size() < 20) c.append(20 - value.size(), ' ');
      c += value;
    }
  } else {
    c = "HIERARCH " + key + " = " + value;
  }
  if (!comment.empty()) c += " / " + comment;
  return pad80(c);
}

std::string fits_double(double v) {
  char b[40];
  std::snprintf(b, sizeof b, "%.15G", v);
  std::string s(b);
  if (s.find('.') == std::string::npos && s.find("INF") == std::string::npos &&
      s.find("NAN") == std::string::npos) {
    const size_t e = s.find('E');
    if (e == std::string::npos) s += ".";
    else s.insert(e, ".");
  }
  return s;
}

std::string fits_string(const std::string& v) {
  std::string q;
  for (char ch : v) {
    q += ch;
    if (ch == '\'') q += '\'';
  }
  if (q.size() < 8) q.append(8 - q.size(), ' ');
  return "'" + q + "'";
}

const char* units_name(int u) {
  switch (u) {   // to_string(PixelMapUnits) (ImageProcessing/pixelize.cpp:41-65)
    case 0: return "not defined";
    case 1: return "surface brightness (ergs / s / cm**2) ";
    case 2: return "counts per sec";
    case 3: return "ADU";
    case 4: return "mass";
    case 5: return "mass density";
    default: return nullptr;
  }
}

}  // namespace

// ======================================================================= C ABI
extern "C" {

int slsim_mpl_create(slsim_ctx ctx, const char* filename, int format, int nsmooth, int recenter_,
                     int ignore_type_in_smoothing, slsim_mpl* out) {
  if (!filename || !out) return fail(SLSIM_ERR_ARG, "slsim_mpl_create: null argument");
  *out = nullptr;
  auto m = new slsim_mpl_s;
  m->ctx = ctx;
  m->filename = filename;
  m->nsmooth = nsmooth;
  int rc = SLSIM_OK;
  if (format == SLSIM_FMT_GLMB) {
    bool found = false;
    rc = glmb_read(*m, m->filename, &found);
    if (!rc && !found) rc = fail(SLSIM_ERR_ARG, " Cannot read file " + m->filename + " (Can't find file.)");
  } else {
    const std::string sizefile = m->filename + "_S" + std::to_string(nsmooth) + ".glmb";
    const bool cached = file_exists(sizefile);
    bool found = false;
    if (cached) {
      rc = glmb_read(*m, sizefile, &found);
      if (!rc && !found) rc = fail(SLSIM_ERR_IO, "Cannot read file " + sizefile);
    }
    if (!rc && !cached) {
      switch (format) {
        case SLSIM_FMT_GADGET2:
          rc = read_gadget2(*m);
          if (!rc) {
            if (!ignore_type_in_smoothing) {
              size_t skip = 0;
              for (int i = 0; i < 6 && !rc; ++i) {
                if (m->nparticles[i] > 0)
                  rc = smooth_particles(ctx, m->data.data() + skip, m->nparticles[i], nsmooth);
                skip += m->nparticles[i];
              }
            } else {
              rc = smooth_particles(ctx, m->data.data(), m->data.size(), nsmooth);
            }
          }
          break;
        case SLSIM_FMT_CSV3:
        case SLSIM_FMT_CSV4:
        case SLSIM_FMT_CSV5: {
          const int cols = format == SLSIM_FMT_CSV3 ? 3 : format == SLSIM_FMT_CSV4 ? 4 : 5;
          rc = read_csv(m->filename, cols, m->data, m->nparticles, m->masses);
          if (!rc && cols < 5) {
            size_t skip = 0;
            for (int i = 0; i < 6 && !rc; ++i) {
              if (m->nparticles[i] > 0)
                rc = smooth_particles(ctx, m->data.data() + skip, m->nparticles[i], nsmooth);
              skip += m->nparticles[i];
            }
          }
          break;
        }
        case SLSIM_FMT_CSV6:
          rc = fail(SLSIM_ERR_ARG,
                    "Data file formate is not supported in MakeParticleLens (csv6: the reference's "
                    "case falls through to its unsupported-format branch).");
          break;
        default:
          rc = fail(SLSIM_ERR_ARG, "Data file formate is not supported in MakeParticleLens.");
          break;
      }
      if (!rc) {
        if (m->nparticles.size() != 6) m->nparticles.resize(6, 0);
        rc = glmb_write(sizefile, m->data.data(), m->data.size(), nsmooth, m->nparticles.data(),
                        m->z_original);
      }
    }
  }
  if (!rc && m->data.empty()) rc = fail(SLSIM_ERR_IO, "no particles in " + m->filename);
  if (rc) {
    delete m;
    return rc;
  }
  if (m->masses.size() != 6) m->masses.assign(6, 0.0f);
  mass_centre(*m, true);
  if (recenter_) {
    for (auto& p : m->data)
      for (int i = 0; i < 3; ++i) p.x[i] = (float)(p.x[i] - m->cm[i]);
    for (int i = 0; i < 3; ++i) {
      m->ll[i] -= m->cm[i];
      m->ur[i] -= m->cm[i];
      m->cm[i] = 0;
    }
  }
  *out = m;
  return SLSIM_OK;
}

int slsim_mpl_create_glmb(slsim_ctx ctx, const char* filename, int recenter_, slsim_mpl* out) {
  if (!filename || !out) return fail(SLSIM_ERR_ARG, "slsim_mpl_create_glmb: null argument");
  *out = nullptr;
  auto m = new slsim_mpl_s;
  m->ctx = ctx;
  m->filename = filename;
  bool found = false;
  int rc = glmb_read(*m, m->filename, &found);
  if (!rc && !found) rc = fail(SLSIM_ERR_IO, "Cannot read file " + m->filename);
  if (!rc && m->data.empty()) rc = fail(SLSIM_ERR_IO, "no particles in " + m->filename);
  if (rc) {
    delete m;
    return rc;
  }
  m->masses.assign(6, 0.0f);
  mass_centre(*m, false);   // the box starts at the origin here (Point_3d zeroes itself)
  if (recenter_) {
    const double c[3] = {m->cm[0], m->cm[1], m->cm[2]};
    recenter(*m, c);
  }
  *out = m;
  return SLSIM_OK;
}

int slsim_mpl_destroy(slsim_mpl m) {
  delete m;
  return SLSIM_OK;
}

int slsim_mpl_data(slsim_mpl m, slsim_particle_f32** data, size_t* n) {
  if (!m || !data || !n) return fail(SLSIM_ERR_ARG, "slsim_mpl_data: null argument");
  *data = m->data.data();
  *n = m->data.size();
  return SLSIM_OK;
}

int slsim_mpl_get_info(slsim_mpl m, slsim_mpl_info* info) {
  if (!m || !info) return fail(SLSIM_ERR_ARG, "slsim_mpl_get_info: null argument");
  std::memset(info, 0, sizeof *info);
  for (int i = 0; i < 6; ++i) {
    info->nparticles[i] = i < (int)m->nparticles.size() ? m->nparticles[i] : 0;
    info->masses[i] = i < (int)m->masses.size() ? m->masses[i] : 0.0f;
  }
  info->z_original = m->z_original;
  for (int i = 0; i < 3; ++i) {
    info->cm[i] = m->cm[i];
    info->bbox_ll[i] = m->ll[i];
    info->bbox_ur[i] = m->ur[i];
  }
  info->nsmooth = m->nsmooth;
  return SLSIM_OK;
}

int slsim_mpl_recenter(slsim_mpl m, const double x[3]) {
  if (!m || !x) return fail(SLSIM_ERR_ARG, "slsim_mpl_recenter: null argument");
  recenter(*m, x);
  return SLSIM_OK;
}

int slsim_mpl_radial_cut(slsim_mpl m, const double center[3], double radius) {
  if (!m || !center) return fail(SLSIM_ERR_ARG, "slsim_mpl_radial_cut: null argument");
  const double radius2 = radius * radius;
  const double c0 = center[0], c1 = center[1], c2 = center[2];
  cut(*m, [&](const slsim_particle_f32& p) {
    const double r2 = (p.x[0] - c0) * (p.x[0] - c0) + (p.x[1] - c1) * (p.x[1] - c1) +
                      (p.x[2] - c2) * (p.x[2] - c2);
    return r2 > radius2;
  });
  return SLSIM_OK;
}

int slsim_mpl_cylindrical_cut(slsim_mpl m, const double center[2], double radius) {
  if (!m || !center) return fail(SLSIM_ERR_ARG, "slsim_mpl_cylindrical_cut: null argument");
  const double radius2 = radius * radius;
  const double c0 = center[0], c1 = center[1];
  cut(*m, [&](const slsim_particle_f32& p) {
    const double r2 = (p.x[0] - c0) * (p.x[0] - c0) + (p.x[1] - c1) * (p.x[1] - c1);
    return r2 > radius2;
  });
  return SLSIM_OK;
}

int slsim_mpl_densest_particle(slsim_mpl m, double x[3]) {
  if (!m || !x) return fail(SLSIM_ERR_ARG, "slsim_mpl_densest_particle: null argument");
  double dmax = -1;
  x[0] = x[1] = x[2] = 0;
  for (const auto& p : m->data) {
    const float df = p.Mass / p.Size / p.Size / p.Size;
    const double d = df;
    if (d > dmax) {
      dmax = d;
      for (int i = 0; i < 3; ++i) x[i] = p.x[i];
    }
  }
  return SLSIM_OK;
}

int slsim_mpl_physical_units(slsim_mpl m, float hubble) {
  if (!m) return fail(SLSIM_ERR_ARG, "slsim_mpl_physical_units: null argument");
  const double length_unit = 1.0 / hubble;
  const double mass_unit = 1.0 / hubble;
  for (auto& p : m->data) {
    for (int i = 0; i < 3; ++i) p.x[i] = (float)(p.x[i] * length_unit);
    p.Size = (float)(p.Size * length_unit);
    p.Mass = (float)(p.Mass * mass_unit);
  }
  return SLSIM_OK;
}

int slsim_glmb_write(const char* filename, const slsim_particle_f32* p, size_t n, int nsmooth,
                     const size_t numbytype[6], double redshift) {
  if (!filename || (!p && n) || !numbytype) return fail(SLSIM_ERR_ARG, "slsim_glmb_write: null argument");
  return glmb_write(filename, p, n, nsmooth, numbytype, redshift);
}

int slsim_glmb_read_header(const char* filename, int* nsmooth, size_t* n, size_t numbytype[6],
                           double* redshift) {
  if (!filename || !nsmooth || !n || !numbytype || !redshift)
    return fail(SLSIM_ERR_ARG, "slsim_glmb_read_header: null argument");
  const int rc = glmb_header(filename, nsmooth, n, numbytype, redshift);
  if (rc == -1) return fail(SLSIM_ERR_IO, std::string("Cannot read file ") + filename);
  return rc;
}

int slsim_glmb_read(const char* filename, slsim_particle_f32* p, size_t n) {
  if (!filename || (!p && n)) return fail(SLSIM_ERR_ARG, "slsim_glmb_read: null argument");
  return glmb_records(filename, p, n);
}

int slsim_gadget2_header(const char* filename, int32_t npart[6], double masstab[6], double* time,
                         double* redshift, double* omega, double* lambda, double* hubble) {
  if (!filename || !npart || !masstab || !time || !redshift || !omega || !lambda || !hubble)
    return fail(SLSIM_ERR_ARG, "slsim_gadget2_header: null argument");
  GadgetHead h;
  const int rc = gadget2_raw(filename, h, nullptr);
  if (rc) return rc;
  for (int i = 0; i < 6; ++i) {
    npart[i] = h.npart[i];
    masstab[i] = h.masstab[i];
  }
  *time = h.time;
  *redshift = h.redshift;
  *omega = h.omega;
  *lambda = h.lambda;
  *hubble = h.hubble;
  return SLSIM_OK;
}

int slsim_gadget2_read(const char* filename, slsim_particle_f32* p, size_t n) {
  if (!filename || (!p && n)) return fail(SLSIM_ERR_ARG, "slsim_gadget2_read: null argument");
  GadgetHead h;
  std::vector<slsim_particle_f32> d;
  const int rc = gadget2_raw(filename, h, &d);
  if (rc) return rc;
  if (d.size() != n)
    return fail(SLSIM_ERR_ARG, "slsim_gadget2_read: the file holds " + std::to_string(d.size()) +
                                   " particles, not " + std::to_string(n));
  std::copy(d.begin(), d.end(), p);
  return SLSIM_OK;
}

int slsim_ascii_halo_count(const char* filename, int multimass, size_t* n) {
  if (!filename || !n) return fail(SLSIM_ERR_ARG, "slsim_ascii_halo_count: null argument");
  std::ifstream in(filename);
  if (!in.is_open()) return fail(SLSIM_ERR_IO, std::string("Unable to open file ") + filename);
  return ascii_header(in, filename, multimass != 0, n);
}

int slsim_ascii_halo_load_f64(slsim_ctx ctx, const char* filename, double redshift, int nsmooth,
                              const double theta_rotate[2], int recenter_, int multimass,
                              double min_size, double rescale_mass, size_t n, double* xyz,
                              float* mass, float* size, slsim_ascii_halo_info* info) {
  return ascii_halo_load<double>(ctx, filename, redshift, nsmooth, theta_rotate, recenter_, multimass,
                                 min_size, rescale_mass, n, xyz, mass, size, info);
}

int slsim_ascii_halo_load_f32(slsim_ctx ctx, const char* filename, double redshift, int nsmooth,
                              const double theta_rotate[2], int recenter_, int multimass,
                              double min_size, double rescale_mass, size_t n, float* xyz,
                              float* mass, float* size, slsim_ascii_halo_info* info) {
  return ascii_halo_load<float>(ctx, filename, redshift, nsmooth, theta_rotate, recenter_, multimass,
                                min_size, rescale_mass, n, xyz, mass, size, info);
}

int slsim_fits_write_pixelmap(const char* filename, const void* map, int is_double, size_t nx,
                              size_t ny, double resolution, const double center[2], double ra,
                              double dec, int units, int flipx, int ctype) {
  if (!filename || !map || !center || nx == 0 || ny == 0)
    return fail(SLSIM_ERR_ARG, "slsim_fits_write_pixelmap: null argument or empty map");
  std::string fn(filename);
  if (fn.empty()) return fail(SLSIM_ERR_ARG, "Please enter a valid filename for the FITS file output");
  if (fn[0] == '!') fn.erase(0, 1);
  const char* uname = units_name(units);
  if (!uname) return fail(SLSIM_ERR_ARG, "No such unit");
  std::string h;
  h += card("SIMPLE", "T", "file does conform to FITS standard", false);
  h += card("BITPIX", is_double ? "-64" : "-32", "number of bits per data pixel", false);
  h += card("NAXIS", "2", "number of data axes", false);
  h += card("NAXIS1", std::to_string(nx), "length of data axis 1", false);
  h += card("NAXIS2", std::to_string(ny), "length of data axis 2", false);
  h += card("EXTEND", "T", "FITS dataset may contain extensions", false);
  h += card("WCSAXES", "2", "number of World Coordinate System axes", false);
  h += card("CRPIX1", fits_double(0.5 * ((double)(long)nx + 1)), "x-coordinate of reference pixel", false);
  h += card("CRPIX2", fits_double(0.5 * ((double)(long)ny + 1)), "y-coordinate of reference pixel", false);
  h += card("CROTA2", fits_double(0.0), "", false);
  h += card("CD1_1", fits_double(-180 * resolution / kPI), "partial of first axis coordinate w.r.t. x (deg)", false);
  h += card("CD1_2", fits_double(0.0), "partial of first axis coordinate w.r.t. y", false);
  h += card("CD2_1", fits_double(0.0), "partial of second axis coordinate w.r.t. x", false);
  h += card("CD2_2", fits_double(180 * resolution / kPI), "partial of second axis coordinate w.r.t. y (deg)", false);
  h += card("Nx", std::to_string(nx), "", false);
  h += card("Ny", std::to_string(ny), "", false);
  const double p1 = center[0] - (nx * resolution) / 2., p2 = center[0] + (nx * resolution) / 2.;
  h += card("range x", fits_double(p2 - p1), "radians", false);
  h += card("RA_global", fits_double(ra), "radians, center", false);
  h += card("DEC_global", fits_double(dec), "radians, center", false);
  h += card("center_x", fits_double(center[0]), "radians, center", false);
  h += card("center_y", fits_double(center[1]), "radians, center", false);
  h += card("CRVAL1", fits_double(center[0] / kDegreesToRadians), "RA, degrees", false);
  h += card("CRVAL2", fits_double(center[1] / kDegreesToRadians), "DEC, degrees", false);
  h += card("UNITS", fits_string(uname), "", true);
  if (ctype) {
    h += card("CTYPE1", fits_string("RA---TAN"), "the coordinate type for the first axis", true);
    h += card("CTYPE2", fits_string("DEC--TAN"), "the coordinate type for the second axis", true);
    h += card("RADESYS", fits_string("ICRS"), "", true);
    h += card("CUNIT1", fits_string("deg     "), "the coordinate unit for the first axis", true);
    h += card("CUNIT2", fits_string("deg     "), "the coordinate unit for the second axis", true);
  }
  h += pad80("END");
  if (h.size() % 2880) h.append(2880 - h.size() % 2880, ' ');
  const size_t esz = is_double ? 8 : 4, npix = nx * ny;
  std::vector<unsigned char> d(npix * esz);
  const unsigned char* src = static_cast<const unsigned char*>(map);
  parallel_ranges(ny, 64, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i)
      for (size_t j = 0; j < nx; ++j) {
        const size_t from = i * nx + j;
        const size_t to = flipx ? i * nx + (nx - 1 - j) : from;   // map_inv[Nx-j + i*Nx] = map[s++]
        if (esz == 8) {
          uint64_t v;
          std::memcpy(&v, src + 8 * from, 8);
          v = __builtin_bswap64(v);
          std::memcpy(&d[8 * to], &v, 8);
        } else {
          uint32_t v;
          std::memcpy(&v, src + 4 * from, 4);
          v = __builtin_bswap32(v);
          std::memcpy(&d[4 * to], &v, 4);
        }
      }
  });
  if (d.size() % 2880) d.resize(d.size() + (2880 - d.size() % 2880), 0);
  FILE* f = std::fopen(fn.c_str(), "wb");
  if (!f) return fail(SLSIM_ERR_IO, "Unable to write to file " + fn);
  bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size() &&
            std::fwrite(d.data(), 1, d.size(), f) == d.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) return fail(SLSIM_ERR_IO, "Unable to write to file " + fn);
  return SLSIM_OK;
}

}  // extern "C"
