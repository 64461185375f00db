// batch.hip -- many GridMap requests per launch: batches (with a schedule
// autotuner and HIP graph replay) and a serving loop with continuous
// batching.
//
// Both reuse the single-request kernels: every request keeps its own ray
// state; a plane step is one launch over a work list of (job, bundle)
// items, a job being one request at one plane.  Because a ray's result
// depends only on its own walk, the result of a request is the same in any
// batch, bit for bit.

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cosmo.h"
#include "slsim_internal.h"

using namespace slsim;

namespace {

constexpr int kNumCU = 256;

int err_arg(const char* m) {
  set_error(m);
  return SLSIM_ERR_ARG;
}

double now_s() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

struct DevBlock {
  slsim_ctx ctx = nullptr;
  void* p = nullptr;
  size_t bytes = 0;
  int get(slsim_ctx c, size_t b) {
    if (p && bytes >= b) return SLSIM_OK;
    release();
    ctx = c;
    int rc = c->alloc(&p, std::max<size_t>(b, 16));
    if (rc == SLSIM_OK) bytes = b;
    return rc;
  }
  void release() {
    if (p) ctx->release(p);
    p = nullptr;
    bytes = 0;
  }
  ~DevBlock() { release(); }
};

void carve_state(double* d, int64_t nslots, RayState* st) { *st = state_view(d, nslots); }

size_t request_bytes(int64_t nslots, int64_t N) {
  return kStateBytesPerSlot * (size_t)nslots + 48 * (size_t)N * (size_t)N + 256;
}

}  // namespace

// ===================================================================== batch
struct slsim_batch_s {
  slsim_ctx ctx = nullptr;
  int precision = SLSIM_FP64;
  struct Req {
    slsim_lens lens;
    int64_t N;
    double c0, c1, range, zs;
    int32_t jmax;
    double aal, bbl;
    DevBlock mem;         // state + outputs
    double* y = nullptr;
    double* A = nullptr;
    int64_t nplane_max = 0;   // particles of its largest plane (schedule order 2)
  };
  std::vector<std::unique_ptr<Req>> reqs;
  slsim_schedule sched{3, 0, 0, 0, 4, 0};
  bool prepared = false;
  int max_steps = 0;
  int64_t total_slots = 0;
  std::vector<GridReq> greq;
  DevBlock d_greq, d_jobs, d_items, d_ctr, d_err;
  uint32_t nfused = 0;        // fused schedule: bundles (d_items holds FusedItem)
  char fused_halos = 0;
  std::vector<size_t> job_off, item_off;
  std::vector<uint32_t> nitems;
  std::vector<char> step_halos;   // per step: the largest StepJob::hkind of its jobs
  hipGraphExec_t exec = nullptr;
  hipGraph_t graph = nullptr;

  void drop_graph() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
  }
  int prepare();
  int prepare_fused(const std::vector<std::vector<StepJob>>& steps, const std::vector<int>& ord);
  int enqueue();
  int run(slsim_shoot_stats* stats);
  ~slsim_batch_s() { drop_graph(); }
};

int slsim_batch_s::prepare() {
  drop_graph();
  const int tw = sched.tile_log2;
  const int th = 6 - tw;
  greq.clear();
  total_slots = 0;
  max_steps = 0;
  std::vector<std::vector<StepJob>> steps(reqs.size());
  for (size_t r = 0; r < reqs.size(); ++r) {
    Req& q = *reqs[r];
    int rc = q.lens->upload_trees();
    if (rc) return rc;
    const int64_t tiles_x = (q.N + (1 << tw) - 1) >> tw;
    const int64_t tiles_y = (q.N + (1 << th) - 1) >> th;
    const int64_t nslots = tiles_x * tiles_y * kWave;
    const size_t sbytes = kStateBytesPerSlot * (size_t)nslots;
    rc = q.mem.get(ctx, request_bytes(nslots, q.N));
    if (rc) return rc;
    char* base = static_cast<char*>(q.mem.p);
    GridReq g;
    std::memset(&g, 0, sizeof(g));
    g.N = q.N;
    g.tiles_x = tiles_x;
    g.slot_off = total_slots;
    g.nslots = nslots;
    g.c0 = q.c0;
    g.c1 = q.c1;
    g.range = q.range;
    g.tw = tw;
    q.jmax = 0;
    q.aal = q.bbl = 0;
    if (q.lens->plane_redshifts.size() > 1) uniform_source(q.lens, q.zs, &q.jmax, &q.aal, &q.bbl);
    g.jmax = q.jmax;
    g.aal = q.aal;
    g.bbl = q.bbl;
    carve_state(reinterpret_cast<double*>(base), nslots, &g.st);
    q.y = reinterpret_cast<double*>(base + ((sbytes + 255) & ~(size_t)255));
    q.A = q.y + 2 * q.N * q.N;
    g.y = q.y;
    g.A = q.A;
    greq.push_back(g);
    total_slots += nslots;
    make_steps(q.lens, g.st, nslots, nullptr, steps[r]);
    max_steps = std::max(max_steps, (int)q.jmax);
  }
  std::vector<int> ord(reqs.size());
  for (size_t r = 0; r < ord.size(); ++r) ord[r] = (int)r;
  if (sched.order == 2)
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) {
      return reqs[a]->nplane_max * greq[a].nslots > reqs[b]->nplane_max * greq[b].nslots;
    });
  if (sched.fuse) return prepare_fused(steps, ord);
  // per step: jobs of the requests with a plane left, and their work items
  std::vector<StepJob> all_jobs;
  std::vector<WorkItem> all_items;
  job_off.assign(max_steps + 1, 0);
  item_off.assign(max_steps + 1, 0);
  nitems.assign(max_steps, 0);
  step_halos.assign(max_steps, 0);
  for (int j = 0; j < max_steps; ++j) {
    job_off[j] = all_jobs.size();
    item_off[j] = all_items.size();
    std::vector<std::pair<uint32_t, int64_t>> live;   // (job index in step, bundles)
    for (int r : ord) {
      if (j >= reqs[r]->jmax) continue;
      live.push_back({(uint32_t)(all_jobs.size() - job_off[j]), greq[r].nslots / kWave});
      all_jobs.push_back(steps[r][j]);
      if (steps[r][j].nhtrees > 0)
        step_halos[j] = (char)std::max<int>(step_halos[j], steps[r][j].hkind);
    }
    if (sched.order == 1) {
      int64_t maxb = 0;
      for (auto& l : live) maxb = std::max(maxb, l.second);
      for (int64_t b = 0; b < maxb; ++b)
        for (auto& l : live)
          if (b < l.second) all_items.push_back({l.first, (uint32_t)b});
    } else {
      for (auto& l : live)
        for (int64_t b = 0; b < l.second; ++b) all_items.push_back({l.first, (uint32_t)b});
    }
    nitems[j] = (uint32_t)(all_items.size() - item_off[j]);
  }
  job_off[max_steps] = all_jobs.size();
  item_off[max_steps] = all_items.size();
  int rc;
  if ((rc = d_greq.get(ctx, sizeof(GridReq) * std::max<size_t>(greq.size(), 1))) ||
      (rc = d_jobs.get(ctx, sizeof(StepJob) * std::max<size_t>(all_jobs.size(), 1))) ||
      (rc = d_items.get(ctx, sizeof(WorkItem) * std::max<size_t>(all_items.size(), 1))) ||
      (rc = d_ctr.get(ctx, 64)) || (rc = d_err.get(ctx, 64)))
    return rc;
  hipStream_t st = ctx->stream;
  SLSIM_HIP_CHECK(hipMemcpyAsync(d_greq.p, greq.data(), sizeof(GridReq) * greq.size(),
                                 hipMemcpyHostToDevice, st));
  if (!all_jobs.empty())
    SLSIM_HIP_CHECK(hipMemcpyAsync(d_jobs.p, all_jobs.data(), sizeof(StepJob) * all_jobs.size(),
                                   hipMemcpyHostToDevice, st));
  if (!all_items.empty())
    SLSIM_HIP_CHECK(hipMemcpyAsync(d_items.p, all_items.data(),
                                   sizeof(WorkItem) * all_items.size(), hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  prepared = true;
  return SLSIM_OK;
}

// The fused schedule's lists: each request's plane steps contiguous, and one
// item per bundle that runs them in order (requests in `ord`; order 1
// interleaves the requests bundle by bundle).
int slsim_batch_s::prepare_fused(const std::vector<std::vector<StepJob>>& steps,
                                 const std::vector<int>& ord) {
  std::vector<StepJob> all_jobs;
  std::vector<uint32_t> first(reqs.size(), 0);
  fused_halos = 0;
  for (int r : ord) {
    first[r] = (uint32_t)all_jobs.size();
    for (int j = 0; j < reqs[r]->jmax; ++j) {
      all_jobs.push_back(steps[r][j]);
      if (steps[r][j].nhtrees > 0)
        fused_halos = (char)std::max<int>(fused_halos, steps[r][j].hkind);
    }
  }
  std::vector<FusedItem> items;
  auto add = [&](int r, int64_t b) {
    items.push_back({first[r], (uint32_t)b, (uint32_t)reqs[r]->jmax, 0u});
  };
  if (sched.order == 1) {
    int64_t maxb = 0;
    for (int r : ord) maxb = std::max(maxb, greq[r].nslots / kWave);
    for (int64_t b = 0; b < maxb; ++b)
      for (int r : ord)
        if (reqs[r]->jmax > 0 && b < greq[r].nslots / kWave) add(r, b);
  } else {
    for (int r : ord)
      if (reqs[r]->jmax > 0)
        for (int64_t b = 0; b < greq[r].nslots / kWave; ++b) add(r, b);
  }
  nfused = (uint32_t)items.size();
  int rc;
  if ((rc = d_greq.get(ctx, sizeof(GridReq) * std::max<size_t>(greq.size(), 1))) ||
      (rc = d_jobs.get(ctx, sizeof(StepJob) * std::max<size_t>(all_jobs.size(), 1))) ||
      (rc = d_items.get(ctx, sizeof(FusedItem) * std::max<size_t>(items.size(), 1))) ||
      (rc = d_ctr.get(ctx, 64)) || (rc = d_err.get(ctx, 64)))
    return rc;
  hipStream_t st = ctx->stream;
  SLSIM_HIP_CHECK(hipMemcpyAsync(d_greq.p, greq.data(), sizeof(GridReq) * greq.size(),
                                 hipMemcpyHostToDevice, st));
  if (!all_jobs.empty())
    SLSIM_HIP_CHECK(hipMemcpyAsync(d_jobs.p, all_jobs.data(), sizeof(StepJob) * all_jobs.size(),
                                   hipMemcpyHostToDevice, st));
  if (!items.empty())
    SLSIM_HIP_CHECK(hipMemcpyAsync(d_items.p, items.data(), sizeof(FusedItem) * items.size(),
                                   hipMemcpyHostToDevice, st));
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  prepared = true;
  return SLSIM_OK;
}

// The launch sequence of one batch run (init, one launch per plane step or
// the fused launch, finalize); captured into a graph when the schedule asks
// for it.
int slsim_batch_s::enqueue() {
  hipStream_t st = ctx->stream;
  uint32_t* err = static_cast<uint32_t*>(d_err.p);
  launch_zero_u32(st, err, 1);
  launch_init_batch(st, static_cast<GridReq*>(d_greq.p), (int)greq.size(), total_slots);
  const StepJob* jobs = static_cast<StepJob*>(d_jobs.p);
  const WorkItem* items = static_cast<WorkItem*>(d_items.p);
  if (sched.fuse)
    launch_plane_steps_fused(
        precision, sched.queue ? (unsigned)(kNumCU * std::max(1, sched.blocks_per_cu)) : 0u,
        nfused, st, jobs, static_cast<FusedItem*>(d_items.p), static_cast<uint32_t*>(d_ctr.p),
        err, fused_halos);
  for (int j = 0; j < (sched.fuse ? 0 : max_steps); ++j) {
    if (nitems[j] == 0) continue;
    if (sched.queue)
      launch_plane_step_queue(precision, (unsigned)(kNumCU * std::max(1, sched.blocks_per_cu)),
                              nitems[j], st, jobs + job_off[j], items + item_off[j],
                              static_cast<uint32_t*>(d_ctr.p), err, step_halos[j]);
    else
      launch_plane_step(precision, false, nitems[j], st, jobs + job_off[j], items + item_off[j],
                        err, step_halos[j]);
  }
  launch_finalize_batch(st, static_cast<GridReq*>(d_greq.p), (int)greq.size(), total_slots);
  SLSIM_HIP_CHECK(hipGetLastError());
  return SLSIM_OK;
}

int slsim_batch_s::run(slsim_shoot_stats* stats) {
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  int rc;
  if (!prepared && (rc = prepare())) return rc;
  hipStream_t st = ctx->stream;
  hipEvent_t e0, e1;
  SLSIM_HIP_CHECK(hipEventCreate(&e0));
  SLSIM_HIP_CHECK(hipEventCreate(&e1));
  const double t0 = now_s();
  if (sched.graph) {
    if (!exec) {
      SLSIM_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
      rc = enqueue();
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(st, &g);
      if (rc) return rc;
      if (e != hipSuccess) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        return SLSIM_ERR_HIP;
      }
      graph = g;
      SLSIM_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    }
    SLSIM_HIP_CHECK(hipEventRecord(e0, st));
    SLSIM_HIP_CHECK(hipGraphLaunch(exec, st));
    SLSIM_HIP_CHECK(hipEventRecord(e1, st));
  } else {
    SLSIM_HIP_CHECK(hipEventRecord(e0, st));
    if ((rc = enqueue())) return rc;
    SLSIM_HIP_CHECK(hipEventRecord(e1, st));
  }
  SLSIM_HIP_CHECK(hipStreamSynchronize(st));
  uint32_t herr = 0;
  SLSIM_HIP_CHECK(hipMemcpy(&herr, d_err.p, 4, hipMemcpyDeviceToHost));
  if (herr) {
    set_error("traversal stack overflow");
    return SLSIM_ERR_TOO_DEEP;
  }
  float ms = 0;
  SLSIM_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  // the precision form's guard, request by request (as a lone GridMap's)
  int32_t guarded = 0;
  if (precision == SLSIM_FP32C) {
    for (size_t r = 0; r < reqs.size(); ++r) {
      const GridReq& g = greq[r];
      if (g.jmax <= 0) continue;
      GuardMap m{};
      m.N = g.N;
      m.tiles_x = g.tiles_x;
      m.tw = g.tw;
      m.c0 = g.c0;
      m.c1 = g.c1;
      m.range = g.range;
      int32_t ng = 0;
      if ((rc = guard_refine_ctx(reqs[r]->lens, g.st, g.nslots, m, g.N * g.N, nullptr,
                                 reqs[r]->zs, g.y, g.A, &ng)))
        return rc;
      guarded += ng;
    }
  }
  if (stats) {
    std::memset(stats, 0, sizeof(*stats));
    stats->total_ms = 1e3 * (now_s() - t0);
    stats->walk_ms = ms;
    stats->n_walk_launches = max_steps;
    stats->n_guarded = guarded;
  }
  return SLSIM_OK;
}

extern "C" {

int slsim_batch_create(slsim_ctx ctx, int n, const slsim_lens* lenses, const int64_t* N,
                       const double* center, const double* range, const double* zs,
                       int precision, slsim_batch* out) {
  if (!ctx || !out || n < 1 || !lenses || !N || !center || !range) return err_arg("bad batch");
  if (precision < SLSIM_FP64 || precision > SLSIM_DIRECT64) return err_arg("bad precision");
  auto b = std::make_unique<slsim_batch_s>();
  b->ctx = ctx;
  b->precision = precision;
  for (int i = 0; i < n; ++i) {
    if (!lenses[i] || lenses[i]->ctx != ctx) return err_arg("lens of another context");
    if (N[i] < 1) return err_arg("GridMap size is < 1!");
    if (!(range[i] > 0)) return err_arg("GridMap range is <= 0!");
    auto q = std::make_unique<slsim_batch_s::Req>();
    q->lens = lenses[i];
    q->N = N[i];
    q->c0 = center[2 * i];
    q->c1 = center[2 * i + 1];
    q->range = range[i];
    q->zs = zs ? zs[i] : 0.0;
    for (auto& pl : lenses[i]->combined) {
      int64_t s = 0;
      for (slsim_plane p : pl) s += p->n;
      q->nplane_max = std::max(q->nplane_max, s);
    }
    b->reqs.push_back(std::move(q));
  }
  *out = b.release();
  return SLSIM_OK;
}

int slsim_batch_set_schedule(slsim_batch b, const slsim_schedule* s) {
  if (!b || !s) return err_arg("null argument");
  if (s->tile_log2 < 1 || s->tile_log2 > 5) return err_arg("tile_log2 must be in 1..5");
  if (s->fuse != 0 && s->fuse != 1) return err_arg("fuse must be 0 or 1");
  b->sched = *s;
  b->prepared = false;
  b->drop_graph();
  return SLSIM_OK;
}

int slsim_batch_get_schedule(slsim_batch b, slsim_schedule* s) {
  if (!b || !s) return err_arg("null argument");
  *s = b->sched;
  return SLSIM_OK;
}

int slsim_batch_run(slsim_batch b, slsim_shoot_stats* stats) {
  if (!b) return err_arg("null batch");
  return b->run(stats);
}

int slsim_batch_output(slsim_batch b, int i, double** y_dev, double** A_dev) {
  if (!b || i < 0 || i >= (int)b->reqs.size()) return err_arg("bad request index");
  if (!b->prepared) return err_arg("batch not run yet");
  if (y_dev) *y_dev = b->reqs[i]->y;
  if (A_dev) *A_dev = b->reqs[i]->A;
  return SLSIM_OK;
}

int slsim_batch_autotune(slsim_batch b, int repeats, slsim_schedule* best, double* base_ms,
                         double* best_ms) {
  if (!b) return err_arg("null batch");
  repeats = std::max(1, repeats);
  auto time_of = [&](const slsim_schedule& s, double* ms) -> int {
    int rc = slsim_batch_set_schedule(b, &s);
    if (rc) return rc;
    slsim_shoot_stats st;
    if ((rc = b->run(&st))) return rc;   // warm-up (and graph capture)
    double best_t = 1e300;
    for (int k = 0; k < repeats; ++k) {
      if ((rc = b->run(&st))) return rc;
      best_t = std::min(best_t, st.walk_ms);
    }
    *ms = best_t;
    return SLSIM_OK;
  };
  slsim_schedule base{3, 0, 0, 0, 4, 0};
  double tb;
  int rc = time_of(base, &tb);
  if (rc) return rc;
  slsim_schedule cur = base;
  cur.graph = 1;
  double tc;
  if ((rc = time_of(cur, &tc))) return rc;
  // greedy passes over the knobs: fused steps, tile, queue, order, queue grid
  const int fuses[] = {0, 1};
  const int tiles[] = {3, 2, 4};
  const int queues[] = {0, 1};
  const int orders[] = {0, 1, 2};
  const int bpcs[] = {2, 4, 8};
  for (int knob = 0; knob < 5; ++knob) {
    const int* vals = knob == 0 ? fuses : knob == 1 ? tiles : knob == 2 ? queues
                    : knob == 3 ? orders : bpcs;
    const int nv = (knob == 0 || knob == 2) ? 2 : 3;
    if (knob == 4 && !cur.queue) continue;
    for (int v = 0; v < nv; ++v) {
      slsim_schedule s = cur;
      int* f = knob == 0 ? &s.fuse : knob == 1 ? &s.tile_log2 : knob == 2 ? &s.queue
             : knob == 3 ? &s.order : &s.blocks_per_cu;
      if (*f == vals[v]) continue;
      *f = vals[v];
      double t;
      if ((rc = time_of(s, &t))) return rc;
      if (t < tc) {
        tc = t;
        cur = s;
      }
    }
  }
  if ((rc = slsim_batch_set_schedule(b, &cur))) return rc;
  if (best) *best = cur;
  if (base_ms) *base_ms = tb;
  if (best_ms) *best_ms = tc;
  return SLSIM_OK;
}

int slsim_batch_destroy(slsim_batch b) {
  if (!b) return SLSIM_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  delete b;
  return SLSIM_OK;
}

}  // extern "C"

// ==================================================================== server
struct slsim_server_s {
  slsim_ctx ctx = nullptr;
  size_t budget = 0, in_use = 0;
  hipStream_t stream = nullptr;
  std::thread th;
  std::mutex mu;
  std::condition_variable cv, cv_done;
  bool stop = false;
  uint64_t next_id = 1;
  enum State { PENDING, ACTIVE, DONE, FAILED };
  struct Req {
    uint64_t id;
    slsim_lens lens;
    int64_t N;
    double c0, c1, range, zs;
    int precision;
    // ray-list request (rayshooterInternal): nrays > 0, positions and
    // optional per-ray source redshifts copied at submit
    int64_t nrays = 0;
    std::vector<double> xh, zsh;
    double *xd = nullptr, *zsd = nullptr, *zpl = nullptr, *dl = nullptr;
    int64_t tiles_x, nslots;
    size_t bytes;
    void* mem = nullptr;
    RayState st;
    double* y = nullptr;
    double* A = nullptr;
    int32_t jmax = 0, step = 0;
    int64_t nbundles = 0, next_bundle = 0;   // progress within plane `step`
    double aal = 0, bbl = 0;
    std::vector<StepJob> steps;
    State state = PENDING;
    int status = SLSIM_OK;
    slsim_request_info info{};
  };
  std::deque<Req*> pending;
  std::vector<Req*> active;
  std::unordered_map<uint64_t, std::unique_ptr<Req>> all;
  slsim_server_stats stats{};
  // per-iteration work lists
  DevBlock d_jobs, d_items, d_err;
  StepJob* h_jobs = nullptr;
  WorkItem* h_items = nullptr;
  size_t h_jobs_cap = 0, h_items_cap = 0;
  double busy_since = -1;

  void loop();
  int admit(Req* r);
  int iterate();
  int stage(size_t njobs, size_t nitems);
};

int slsim_server_s::stage(size_t njobs, size_t nitems) {
  if (njobs > h_jobs_cap) {
    if (h_jobs) (void)hipHostFree(h_jobs);
    h_jobs_cap = std::max(njobs, 2 * h_jobs_cap);
    SLSIM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_jobs), sizeof(StepJob) * h_jobs_cap));
  }
  if (nitems > h_items_cap) {
    if (h_items) (void)hipHostFree(h_items);
    h_items_cap = std::max(nitems, 2 * h_items_cap);
    SLSIM_HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_items), sizeof(WorkItem) * h_items_cap));
  }
  int rc;
  if ((rc = d_jobs.get(ctx, sizeof(StepJob) * h_jobs_cap))) return rc;
  if ((rc = d_items.get(ctx, sizeof(WorkItem) * h_items_cap))) return rc;
  return SLSIM_OK;
}

// Allocates the request's state and outputs and initialises its rays.
int slsim_server_s::admit(Req* r) {
  int rc = ctx->scratch_get(&r->mem, r->bytes);
  if (rc) return rc;
  const size_t sbytes = kStateBytesPerSlot * (size_t)r->nslots;
  char* base = static_cast<char*>(r->mem);
  carve_state(reinterpret_cast<double*>(base), r->nslots, &r->st);
  if (r->nrays > 0) {
    // ray list: x, zs, y, A and the lens tables after the state
    slsim_lens_s* L = r->lens;
    if ((rc = L->upload_trees())) return rc;
    const int64_t n = r->nrays;
    const int nplanes = (int)L->plane_redshifts.size() - 1;
    double* p = reinterpret_cast<double*>(base + ((sbytes + 255) & ~(size_t)255));
    r->xd = p;   p += 2 * n;
    r->y = p;    p += 2 * n;
    r->A = p;    p += 4 * n;
    r->zpl = p;  p += L->plane_redshifts.size();
    r->dl = p;   p += L->Dl.size();
    r->zsd = r->zsh.empty() ? nullptr : p;
    SLSIM_HIP_CHECK(hipMemcpyAsync(r->xd, r->xh.data(), 16 * n, hipMemcpyHostToDevice, stream));
    if (r->zsd)
      SLSIM_HIP_CHECK(hipMemcpyAsync(r->zsd, r->zsh.data(), 8 * n, hipMemcpyHostToDevice, stream));
    SLSIM_HIP_CHECK(hipMemcpyAsync(r->zpl, L->plane_redshifts.data(),
                                   8 * L->plane_redshifts.size(), hipMemcpyHostToDevice, stream));
    SLSIM_HIP_CHECK(hipMemcpyAsync(r->dl, L->Dl.data(), 8 * L->Dl.size(), hipMemcpyHostToDevice,
                                   stream));
    r->jmax = 0;
    if (nplanes > 0) {
      if (r->zsd) r->jmax = nplanes;   // per-ray jmax in the state
      else uniform_source(L, 0.0, &r->jmax, &r->aal, &r->bbl);
    }
    make_steps(L, r->st, r->nslots, nullptr, r->steps);
    r->nbundles = r->nslots / kWave;
    r->next_bundle = 0;
    FlatLCDM cosmo{L->Om, L->OL, L->h};
    launch_init_array(stream, r->st, r->nslots, n, nullptr, r->xd, nplanes > 0 ? r->zsd : nullptr,
                      r->zsd ? 0 : r->jmax, r->aal, r->bbl, r->zpl, r->dl, nplanes, cosmo);
    SLSIM_HIP_CHECK(hipGetLastError());
    // xh and zsh live until the request is collected, after the stream has read them
    return SLSIM_OK;
  }
  r->y = reinterpret_cast<double*>(base + ((sbytes + 255) & ~(size_t)255));
  r->A = r->y + 2 * r->N * r->N;
  if ((rc = r->lens->upload_trees())) return rc;
  if (r->lens->plane_redshifts.size() > 1) uniform_source(r->lens, r->zs, &r->jmax, &r->aal, &r->bbl);
  make_steps(r->lens, r->st, r->nslots, nullptr, r->steps);
  r->nbundles = r->nslots / kWave;
  r->next_bundle = 0;
  launch_init_grid(stream, r->st, r->nslots, r->N, r->tiles_x, 3, r->c0, r->c1, r->range, r->jmax,
                   r->aal, r->bbl);
  SLSIM_HIP_CHECK(hipGetLastError());
  return SLSIM_OK;
}

// One scheduler iteration: up to kIterBundles bundles of plane steps, so that
// an iteration stays short however large the requests in flight are.  The
// oldest request gets up to a quarter of the iteration (no request starves);
// the rest goes to the requests with the least remaining work first.  A
// request takes bundles of its current plane only; its rays are independent,
// so how its bundles are split over iterations does not change a bit of its
// result.  One launch per precision.
constexpr int64_t kIterBundles = 32768;

int slsim_server_s::iterate() {
  std::vector<Req*> order;
  for (Req* r : active)
    if (r->step < r->jmax) order.push_back(r);
  if (order.empty()) return SLSIM_OK;
  auto remaining = [](const Req* r) {
    return (int64_t)(r->jmax - r->step) * r->nbundles - r->next_bundle;
  };
  Req* oldest = order.front();
  for (Req* r : order)
    if (r->info.t_submit < oldest->info.t_submit) oldest = r;
  std::stable_sort(order.begin(), order.end(),
                   [&](const Req* a, const Req* b) { return remaining(a) < remaining(b); });
  std::vector<std::pair<Req*, int64_t>> take;   // request, bundles this iteration
  int64_t left = kIterBundles;
  {
    const int64_t t = std::min(oldest->nbundles - oldest->next_bundle, kIterBundles / 4);
    take.push_back({oldest, t});
    left -= t;
  }
  for (Req* r : order) {
    if (left <= 0) break;
    if (r == oldest) {
      const int64_t more = std::min(r->nbundles - r->next_bundle - take[0].second, left);
      take[0].second += more;
      left -= more;
      continue;
    }
    const int64_t t = std::min(r->nbundles - r->next_bundle, left);
    take.push_back({r, t});
    left -= t;
  }
  size_t nit = 0;
  for (auto& t : take) nit += (size_t)t.second;
  int rc = stage(take.size(), nit);
  if (rc) return rc;
  // jobs and items of all precisions in one staging copy, grouped by precision
  size_t nj = 0, ni = 0;
  size_t job0[SLSIM_DIRECT64 + 2], item0[SLSIM_DIRECT64 + 2];
  for (int prec = SLSIM_FP64; prec <= SLSIM_DIRECT64; ++prec) {
    job0[prec] = nj;
    item0[prec] = ni;
    for (auto& t : take) {
      Req* r = t.first;
      if (r->precision != prec || t.second == 0) continue;
      h_jobs[nj] = r->steps[r->step];
      const uint32_t local = (uint32_t)(nj - job0[prec]);
      for (int64_t b = 0; b < t.second; ++b)
        h_items[ni++] = {local, (uint32_t)(r->next_bundle + b)};
      ++nj;
    }
  }
  job0[SLSIM_DIRECT64 + 1] = nj;
  item0[SLSIM_DIRECT64 + 1] = ni;
  SLSIM_HIP_CHECK(hipMemcpyAsync(d_jobs.p, h_jobs, sizeof(StepJob) * nj, hipMemcpyHostToDevice,
                                 stream));
  SLSIM_HIP_CHECK(hipMemcpyAsync(d_items.p, h_items, sizeof(WorkItem) * ni, hipMemcpyHostToDevice,
                                 stream));
  for (int prec = SLSIM_FP64; prec <= SLSIM_DIRECT64; ++prec) {
    const size_t n_items = item0[prec + 1] - item0[prec];
    if (n_items == 0) continue;
    int halos = 0;
    for (size_t q = job0[prec]; q < job0[prec + 1]; ++q)
      if (h_jobs[q].nhtrees > 0) halos = std::max(halos, (int)h_jobs[q].hkind);
    launch_plane_step(prec, false, (uint32_t)n_items, stream,
                      static_cast<StepJob*>(d_jobs.p) + job0[prec],
                      static_cast<WorkItem*>(d_items.p) + item0[prec],
                      static_cast<uint32_t*>(d_err.p), halos);
  }
  SLSIM_HIP_CHECK(hipGetLastError());
  // the staging buffers are reused next iteration
  SLSIM_HIP_CHECK(hipStreamSynchronize(stream));
  for (auto& t : take) {
    Req* r = t.first;
    r->next_bundle += t.second;
    if (r->next_bundle >= r->nbundles) {
      r->next_bundle = 0;
      ++r->step;
      ++r->info.iterations;
    }
  }
  return SLSIM_OK;
}

void slsim_server_s::loop() {
  (void)hipSetDevice(ctx->device);
  std::unique_lock<std::mutex> lk(mu);
  while (!stop) {
    // admission: FIFO while the next request's state and outputs fit
    while (!pending.empty() && in_use + pending.front()->bytes <= budget) {
      Req* r = pending.front();
      pending.pop_front();
      in_use += r->bytes;
      r->info.t_admit = now_s();
      const int rc = admit(r);
      if (rc) {
        r->state = FAILED;
        r->status = rc;
        ++stats.rejected;
        if (r->mem) ctx->scratch_put(r->mem);
        r->mem = nullptr;
        in_use -= r->bytes;
        cv_done.notify_all();
        continue;
      }
      ++stats.admitted;
      r->state = ACTIVE;
      active.push_back(r);
    }
    if (active.empty()) {
      if (busy_since >= 0) {
        stats.busy_s += now_s() - busy_since;
        busy_since = -1;
      }
      cv.wait(lk, [&] { return stop || (!pending.empty() && in_use + pending.front()->bytes <= budget); });
      continue;
    }
    if (busy_since < 0) busy_since = now_s();
    stats.max_active = std::max<int64_t>(stats.max_active, (int64_t)active.size());
    lk.unlock();
    int rc = iterate();
    // finalize the requests whose last plane is done
    std::vector<Req*> finished;
    for (Req* r : active) {
      if (rc || r->step >= r->jmax) {
        if (rc == SLSIM_OK && r->nrays > 0)
          launch_finalize_array(stream, r->st, r->nslots, r->nrays, nullptr, r->y, r->A);
        else if (rc == SLSIM_OK)
          launch_finalize_grid(stream, r->st, r->nslots, r->N, r->tiles_x, 3, r->y, r->A);
        finished.push_back(r);
      }
    }
    if (!finished.empty() && hipStreamSynchronize(stream) != hipSuccess) rc = SLSIM_ERR_HIP;
    uint32_t herr = 0;
    if (rc == SLSIM_OK && hipMemcpy(&herr, d_err.p, 4, hipMemcpyDeviceToHost) == hipSuccess && herr)
      rc = SLSIM_ERR_TOO_DEEP;
    // the precision form's guard of each finished request, as a lone shoot's
    for (Req* r : finished) {
      if (rc || r->precision != SLSIM_FP32C || r->jmax <= 0) continue;
      GuardMap m{};
      if (r->nrays > 0) {
        m.x = r->xd;
        m.zs = r->zsd;
        m.n = r->nrays;
      } else {
        m.N = r->N;
        m.tiles_x = r->tiles_x;
        m.tw = 3;
        m.c0 = r->c0;
        m.c1 = r->c1;
        m.range = r->range;
      }
      int32_t ng = 0;
      rc = guard_refine_ctx(r->lens, r->st, r->nslots, m, r->nrays > 0 ? r->nrays : r->N * r->N,
                            r->nrays > 0 ? r->zsd : nullptr, r->zs, r->y, r->A, &ng);
    }
    const double t = now_s();
    lk.lock();
    ++stats.iterations;
    for (Req* r : finished) {
      r->info.t_done = t;
      r->state = rc ? FAILED : DONE;
      r->status = rc;
      if (!rc) ++stats.completed;
      active.erase(std::find(active.begin(), active.end(), r));
    }
    if (!finished.empty()) cv_done.notify_all();
  }
}

extern "C" {

double slsim_now(void) { return now_s(); }

int slsim_server_create(slsim_ctx ctx, size_t hbm_budget_bytes, slsim_server* out) {
  if (!ctx || !out) return err_arg("null argument");
  SLSIM_HIP_CHECK(hipSetDevice(ctx->device));
  auto s = std::make_unique<slsim_server_s>();
  s->ctx = ctx;
  s->budget = hbm_budget_bytes ? hbm_budget_bytes : ctx->budget / 2;
  SLSIM_HIP_CHECK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  int rc = s->d_err.get(ctx, 64);
  if (rc) return rc;
  SLSIM_HIP_CHECK(hipMemset(s->d_err.p, 0, 64));
  s->stats.budget_bytes = (int64_t)s->budget;
  slsim_server_s* raw = s.get();
  s->th = std::thread([raw] { raw->loop(); });
  *out = s.release();
  return SLSIM_OK;
}

int slsim_server_submit(slsim_server s, slsim_lens lens, int64_t N, double c0, double c1,
                        double range, double zs, int precision, uint64_t* id) {
  if (!s || !lens || !id) return err_arg("null argument");
  if (lens->ctx != s->ctx) return err_arg("lens of another context");
  if (N < 1) return err_arg("GridMap size is < 1!");
  if (!(range > 0)) return err_arg("GridMap range is <= 0!");
  if (precision < SLSIM_FP64 || precision > SLSIM_DIRECT64) return err_arg("bad precision");
  auto r = std::make_unique<slsim_server_s::Req>();
  r->lens = lens;
  r->N = N;
  r->c0 = c0;
  r->c1 = c1;
  r->range = range;
  r->zs = zs;
  r->precision = precision;
  r->tiles_x = (N + 7) / 8;
  r->nslots = r->tiles_x * r->tiles_x * kWave;
  r->bytes = request_bytes(r->nslots, N);
  r->info.t_submit = now_s();
  r->info.rays = N * N;
  std::lock_guard<std::mutex> g(s->mu);
  ++s->stats.submitted;
  if (r->bytes > s->budget) {
    ++s->stats.rejected;
    set_error("request needs more HBM than the server budget");
    return SLSIM_ERR_BUDGET;
  }
  r->id = s->next_id++;
  *id = r->id;
  s->pending.push_back(r.get());
  s->all[r->id] = std::move(r);
  s->cv.notify_all();
  return SLSIM_OK;
}

int slsim_server_submit_rays(slsim_server s, slsim_lens lens, size_t n, const double* x,
                             const double* zs, int precision, uint64_t* id) {
  if (!s || !lens || !id || (n && !x)) return err_arg("null argument");
  if (lens->ctx != s->ctx) return err_arg("lens of another context");
  if (n == 0) return err_arg("no rays");
  if (n >= ((size_t)1 << 31)) return err_arg("too many rays in one request");
  if (precision < SLSIM_FP64 || precision > SLSIM_DIRECT64) return err_arg("bad precision");
  auto r = std::make_unique<slsim_server_s::Req>();
  r->lens = lens;
  r->N = 0;
  r->c0 = r->c1 = r->range = 0;
  r->zs = 0;
  r->precision = precision;
  r->nrays = (int64_t)n;
  r->xh.assign(x, x + 2 * n);
  if (zs) r->zsh.assign(zs, zs + n);
  r->tiles_x = 0;
  r->nslots = (((int64_t)n + kWave - 1) / kWave) * kWave;
  const size_t tables = lens->plane_redshifts.size() + lens->Dl.size();
  r->bytes = kStateBytesPerSlot * (size_t)r->nslots + 256 + 8 * (8 * n + (zs ? n : 0) + tables);
  r->info.t_submit = now_s();
  r->info.rays = (int64_t)n;
  std::lock_guard<std::mutex> g(s->mu);
  ++s->stats.submitted;
  if (r->bytes > s->budget) {
    ++s->stats.rejected;
    set_error("request needs more HBM than the server budget");
    return SLSIM_ERR_BUDGET;
  }
  r->id = s->next_id++;
  *id = r->id;
  s->pending.push_back(r.get());
  s->all[r->id] = std::move(r);
  s->cv.notify_all();
  return SLSIM_OK;
}

int slsim_server_wait(slsim_server s, uint64_t id, double* y_host, double* A_host,
                      slsim_request_info* info) {
  if (!s) return err_arg("null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return err_arg("unknown request id");
  slsim_server_s::Req* r = it->second.get();
  s->cv_done.wait(lk, [&] {
    return r->state == slsim_server_s::DONE || r->state == slsim_server_s::FAILED;
  });
  lk.unlock();
  int rc = r->status;
  if (rc == SLSIM_OK) {
    const size_t n = r->nrays > 0 ? (size_t)r->nrays : (size_t)r->N * r->N;
    if (y_host && hipMemcpy(y_host, r->y, 16 * n, hipMemcpyDeviceToHost) != hipSuccess) rc = SLSIM_ERR_HIP;
    if (A_host && hipMemcpy(A_host, r->A, 32 * n, hipMemcpyDeviceToHost) != hipSuccess) rc = SLSIM_ERR_HIP;
  }
  r->info.t_collect = now_s();
  r->info.status = rc;
  if (info) *info = r->info;
  lk.lock();
  if (r->mem) {
    // kept in the context's scratch cache: no hipFree (and its device-wide
    // synchronisation) on the request path
    s->ctx->scratch_put(r->mem);
    r->mem = nullptr;
    s->in_use -= r->bytes;
  }
  s->all.erase(id);
  s->cv.notify_all();
  if (rc) set_error("request failed");
  return rc;
}

int slsim_server_poll(slsim_server s, uint64_t id, int* done) {
  if (!s || !done) return err_arg("null argument");
  std::lock_guard<std::mutex> g(s->mu);
  auto it = s->all.find(id);
  if (it == s->all.end()) return err_arg("unknown request id");
  const int st = it->second->state;
  *done = (st == slsim_server_s::DONE || st == slsim_server_s::FAILED) ? 1 : 0;
  return SLSIM_OK;
}

int slsim_server_get_stats(slsim_server s, slsim_server_stats* st) {
  if (!s || !st) return err_arg("null argument");
  std::lock_guard<std::mutex> g(s->mu);
  *st = s->stats;
  st->in_use_bytes = (int64_t)s->in_use;
  if (s->busy_since >= 0) st->busy_s += now_s() - s->busy_since;
  return SLSIM_OK;
}

int slsim_server_destroy(slsim_server s) {
  if (!s) return SLSIM_OK;
  {
    std::lock_guard<std::mutex> g(s->mu);
    s->stop = true;
    s->cv.notify_all();
  }
  if (s->th.joinable()) s->th.join();
  (void)hipStreamSynchronize(s->stream);
  for (auto& kv : s->all)
    if (kv.second->mem) s->ctx->scratch_put(kv.second->mem);
  if (s->h_jobs) (void)hipHostFree(s->h_jobs);
  if (s->h_items) (void)hipHostFree(s->h_items);
  (void)hipStreamDestroy(s->stream);
  delete s;
  return SLSIM_OK;
}

}  // extern "C"
