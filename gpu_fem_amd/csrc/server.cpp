// Serving loop for linear-static requests (continuous admission under an HBM
// budget, overlapping solves on the device).
//
// The reference has no serving component; this is the request loop of the
// served workload: requests of differing meshes arrive over time, each is
// admitted when a solver slot is free and its device footprint fits the
// budget, and every admitted request runs the whole device path (symbolic
// phase on the worker's host thread, upload, assembly, factorisation,
// GMRES / GMRES-IR) on its own HIP stream, so the device work of concurrent
// requests overlaps.  Results are collected with gf_server_poll.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdlib>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "common.h"
#include "krylov.h"
#include "problem.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
  long long id = 0;
  int nnodes = 0, nelems = 0;
  std::vector<int> conn, bcs;
  std::vector<double> xpts, data, loads;
  gf_options opt{};
  Clock::time_point arrival;
};

struct Result {
  long long id = 0;
  std::vector<double> x;
  gf_solve_info info{};
  double latency_s = 0.0, queue_s = 0.0;
  int status = GF_OK;
};

}  // namespace

struct gf_server_s {
  int device = 0;
  size_t budget = 0;
  // device bytes counted against the budget: the reservation of every
  // request being solved plus the device blocks each worker context keeps
  // cached between requests
  size_t in_use = 0;
  std::mutex mu;
  std::condition_variable cv;      // queue / budget changes
  std::deque<Request> queue;
  std::deque<Result> done;
  std::vector<std::thread> workers;
  bool stopping = false;
  int active = 0;    // requests taken from the queue and not yet finished
  int busy = 0;      // workers holding a request's reservation
  int pressure = 0;  // workers waiting for budget: idle workers drop caches
  size_t peak_bytes = 0;
  long long completed = 0, failed = 0, escalated = 0;
  std::vector<double> latencies;
  Clock::time_point t0 = Clock::now();
  int class_bound[2] = {120, 360};  // sweep depth bounds of the size classes
  // batched GMRES engines, one per preconditioner precision (FP64, FP32)
  // and size class: the Krylov phase of concurrent requests of one class
  // runs as fused launches; each engine has its own stream (null: each
  // worker solves on its own stream)
  static constexpr int kClasses = 3;
  gf::BatchEngine* engine[3][kClasses] = {};
};

namespace {

// A lower bound of a request's device bytes, known before its symbolic
// phase: the diagonal blocks of K and of the preconditioner and the Krylov
// basis.  A request above the budget even by this bound can never be run.
size_t lower_bound_bytes(const Request& r) {
  const double n = r.nnodes;
  return (size_t)(2.0 * 288.0 * n + 48.0 * n * (r.opt.restart + 1));
}

constexpr int kMaxEscalations = 2;
constexpr double kEscalateAbove = 1e3;
constexpr int kMaxEscalatedFill = 64;

// One worker: takes the oldest queued request, runs its symbolic phase on
// the host (which fixes its device footprint exactly), waits until that
// footprint fits the budget, then runs the device path on its own stream.
void worker_main(gf_server_s* s) {
  gf_context ctx = nullptr;
  if (gf::context_create(s->device, true, &ctx) != GF_OK) return;
  size_t held = 0;  // this worker's share of in_use
  // give this worker's cached blocks back (hipFree) and its share of the
  // budget with them; called with the lock held, drops it meanwhile
  auto drop_cache = [&](std::unique_lock<std::mutex>& lk) {
    lk.unlock();
    gf::release_cache(ctx);
    lk.lock();
    s->in_use -= held;
    held = 0;
    s->cv.notify_all();
  };
  for (;;) {
    Request req;
    {
      std::unique_lock<std::mutex> lk(s->mu);
      for (;;) {
        s->cv.wait(lk, [&] {
          return s->stopping || !s->queue.empty() ||
                 (s->pressure > 0 && held > 0);
        });
        if (!s->queue.empty() || s->stopping) break;
        drop_cache(lk);  // another worker waits for budget
      }
      if (s->stopping && s->queue.empty()) break;
      req = std::move(s->queue.front());
      s->queue.pop_front();
      s->active++;
    }
    Result res;
    res.id = req.id;
    res.queue_s = std::chrono::duration<double>(Clock::now() - req.arrival).count();
    // A request whose solve ends far from the tolerance (relative residual
    // above kEscalateAbove * rtol: the preconditioner failed, not FP64's
    // attainable accuracy) or whose factorisation or GMRES breaks down is
    // run again with twice the ILU level of fill (up to kMaxEscalations
    // times, and not beyond ILU(kMaxEscalatedFill)); each attempt gives its
    // reservation back and is admitted again with its own footprint.
    int st = GF_OK;
    for (int attempt = 0;; attempt++) {
      gf::ProblemPlan plan;
      std::string why;
      st = gf::problem_plan(req.nnodes, req.nelems, req.conn.data(),
                            (int)req.bcs.size(), req.bcs.data(), req.opt, plan,
                            &why);
      if (st != GF_OK) break;
      const size_t bytes = plan.device_bytes(true);
      if (bytes > s->budget) {
        st = GF_ERR_OVER_BUDGET;
        break;
      }
      {
        // admission: this worker's footprint becomes its cache plus the
        // request (cached blocks may be reused, so this is an upper bound)
        std::unique_lock<std::mutex> lk(s->mu);
        for (;;) {
          const size_t others = s->in_use - held;
          if (others + held + bytes <= s->budget) break;
          if (held > 0) {
            drop_cache(lk);
            continue;
          }
          if (s->busy == 0 && others == 0) break;
          s->pressure++;
          s->cv.notify_all();
          s->cv.wait(lk);
          s->pressure--;
        }
        s->in_use += bytes;
        held += bytes;
        s->busy++;
        s->peak_bytes = std::max(s->peak_bytes, s->in_use);
      }
      gf_problem p = nullptr;
      st = gf::problem_create_planned(ctx, plan, req.conn.data(),
                                      req.xpts.data(), req.data.data(),
                                      req.bcs.data(), &p);
      if (st == GF_OK) st = gf_problem_assemble(p);
      if (st == GF_OK) st = gf_problem_factor(p);
      if (st == GF_OK) {
        res.x.resize(6 * (size_t)req.nnodes);
        // requests advance in lockstep inside an engine, so each engine
        // takes requests of similar sweep depth (dependency levels)
        int nl = 0, nu = 0;
        gf_ilu_info(gf_problem_ilu(p), nullptr, &nl, &nu);
        const int depth = std::max(nl, nu);
        const int cls = depth < s->class_bound[0] ? 0
                        : depth < s->class_bound[1] ? 1 : 2;
        gf::BatchEngine* eng = s->engine[req.opt.precision][cls];
        if (eng) {
          gf::BatchJob job;
          job.p = p;
          job.loads = req.loads.data();
          job.x = res.x.data();
          st = gf::batch_engine_solve(eng, &job);
          res.info = job.info;
        } else {
          st = gf_problem_solve(p, req.loads.data(), res.x.data(), 0,
                                &res.info);
        }
      }
      if (p) gf_problem_destroy(p);
      {
        // the problem's blocks went back to the context's cache
        std::lock_guard<std::mutex> lk(s->mu);
        s->in_use -= held;
        held = ctx->cached_bytes;
        s->in_use += held;
        s->busy--;
      }
      s->cv.notify_all();
      const bool numerical =
          (st == GF_ERR_NO_CONVERGENCE &&
           !(res.info.rel_resid <= kEscalateAbove * req.opt.rtol)) ||
          st == GF_ERR_ZERO_PIVOT || st == GF_ERR_BREAKDOWN;
      if (!numerical || attempt >= kMaxEscalations ||
          req.opt.lev_fill >= kMaxEscalatedFill)
        break;
      req.opt.lev_fill =
          std::min(2 * std::max(req.opt.lev_fill, 1), kMaxEscalatedFill);
      std::lock_guard<std::mutex> lk(s->mu);
      s->escalated++;
    }
    res.status = st;
    res.latency_s =
        std::chrono::duration<double>(Clock::now() - req.arrival).count();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->active--;
      if (st == GF_OK)
        s->completed++;
      else
        s->failed++;
      s->latencies.push_back(res.latency_s);
      s->done.push_back(std::move(res));
    }
    s->cv.notify_all();
  }
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->in_use -= held;
  }
  gf_context_destroy(ctx);
}

double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - i;
  return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" {

int gf_server_create(int device, double budget_gb, int slots,
                     gf_server* out) {
  if (!out || slots < 1 || slots > 64 || budget_gb <= 0.0) return GF_ERR_INVALID;
  auto* s = new gf_server_s;
  s->device = device;
  s->budget = (size_t)(budget_gb * 1e9);
  // GPUFEM_SERVE_BATCH=1: the Krylov phase of concurrent requests runs in
  // the batched engines (fused launches, graph replays); by default each
  // worker solves on its own stream, which measured faster on the served
  // mix (the batched solves advance in lockstep, so a small request's step
  // lasts as long as the largest request's step in its engine)
  const char* be = getenv("GPUFEM_SERVE_BATCH");
  if (be && atoi(be) == 1) {
    int st = GF_OK;
    for (int prec : {0, 2})
      for (int c = 0; c < gf_server_s::kClasses && st == GF_OK; c++)
        st = gf::batch_engine_create(device, prec, slots, &s->engine[prec][c]);
    if (st != GF_OK) {
      for (auto& row : s->engine)
        for (auto*& e : row) {
          gf::batch_engine_destroy(e);
          e = nullptr;
        }
      delete s;
      return st;
    }
  }
  for (int i = 0; i < slots; i++) s->workers.emplace_back(worker_main, s);
  *out = s;
  return GF_OK;
}

int gf_server_submit(gf_server s, long long id, int nnodes, int nelems,
                     const int* conn, const double* xpts, const double* data,
                     int nbcs, const int* bcs, const double* loads,
                     const gf_options* opt) {
  if (!s || nnodes <= 0 || nelems <= 0 || !conn || !xpts || !data || !loads ||
      nbcs < 0 || (nbcs && !bcs))
    return GF_ERR_INVALID;
  Request r;
  r.id = id;
  r.nnodes = nnodes;
  r.nelems = nelems;
  r.conn.assign(conn, conn + 4 * (size_t)nelems);
  r.xpts.assign(xpts, xpts + 3 * (size_t)nnodes);
  r.data.assign(data, data + 4 * (size_t)nelems);
  r.bcs.assign(bcs, bcs + nbcs);
  r.loads.assign(loads, loads + 6 * (size_t)nnodes);
  if (opt)
    r.opt = *opt;
  else
    gf_options_default(&r.opt);
  if (!gf::options_valid(r.opt)) return GF_ERR_INVALID;
  r.arrival = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (lower_bound_bytes(r) > s->budget) return GF_ERR_OVER_BUDGET;
    s->queue.push_back(std::move(r));
  }
  s->cv.notify_all();
  return GF_OK;
}

// Pops one finished request; returns 1 if one was returned, 0 if none is
// ready.  x must hold 6*nnodes doubles of that request (NULL to discard).
int gf_server_poll(gf_server s, long long* id, double* x, size_t x_len,
                   gf_solve_info* info, double* latency_s) {
  if (!s) return GF_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->done.empty()) return 0;
  Result r = std::move(s->done.front());
  s->done.pop_front();
  if (id) *id = r.id;
  if (x && x_len >= r.x.size())
    std::copy(r.x.begin(), r.x.end(), x);
  if (info) {
    *info = r.info;
    info->status = r.status;
  }
  if (latency_s) *latency_s = r.latency_s;
  return 1;
}

// stats: [0] completed [1] failed [2] queued [3] active [4] p50 latency s
// [5] p99 latency s [6] peak bytes admitted [7] elapsed s
int gf_server_stats(gf_server s, double* stats) {
  if (!s || !stats) return GF_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  stats[0] = (double)s->completed;
  stats[1] = (double)s->failed;
  stats[2] = (double)s->queue.size();
  stats[3] = (double)s->active;
  stats[4] = percentile(s->latencies, 0.50);
  stats[5] = percentile(s->latencies, 0.99);
  stats[6] = (double)s->peak_bytes;
  stats[7] = std::chrono::duration<double>(Clock::now() - s->t0).count();
  stats[8] = (double)s->escalated;
  return GF_OK;
}

// waits until every submitted request has finished
int gf_server_drain(gf_server s) {
  if (!s) return GF_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv.wait(lk, [&] { return s->queue.empty() && s->active == 0; });
  return GF_OK;
}

int gf_server_destroy(gf_server s) {
  if (!s) return GF_OK;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stopping = true;
  }
  s->cv.notify_all();
  for (auto& t : s->workers) t.join();
  for (auto& row : s->engine)
    for (auto* e : row) gf::batch_engine_destroy(e);
  delete s;
  return GF_OK;
}

}  // extern "C"
