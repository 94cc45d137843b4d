// tfs_server.cpp -- request serving on one GPU: continuous batching of solves.
//
// A case is one linear system (structure + values) registered by the client;
// its GPU copy (matrix, layouts, factor, work vectors) is made when its first
// solve is admitted and stays resident across its visits (time steps) until
// the client marks a solve as the case's last.  Requests queue in arrival
// order; each of W worker threads takes the first request whose case is not
// already being solved and whose GPU copy exists or fits in the HBM budget,
// and runs it on that case's own HIP stream -- so solves of different cases
// overlap on the device and a request arriving while others run starts as
// soon as a worker and the memory for it are free.  Each solve's arithmetic
// depends only on its own case (per-matrix state, fixed-order reductions), so
// a result is the same alone and in any mix.  The preconditioner sweeps use
// the static persistent schedule with each sweep capped at 1/W of the
// workgroups that fit on the GPU at once: the at most W sweeps in flight
// (one per worker) then all fit together, so none can wait on a workgroup
// that another sweep keeps out, and every other kernel terminates by itself.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "tfs_internal.hpp"
#include "tfs_solver.h"

namespace {

using Clock = std::chrono::steady_clock;

struct Case {
  int n = 0;
  long long nnz = 0;
  std::vector<int> row, col, dia;
  std::vector<double> val;
  int base = 0;
  tfs_matrix *A = nullptr;
  long long bytes = 0;     // device bytes held while resident (estimate until created)
  long long last_use = 0;  // admission tick of the latest visit (LRU eviction)
  int queued = 0;          // requests of this case still in the queue
  bool busy = false;
  bool released = false;
};

struct Request {
  int case_id = 0, method = 0, prec = 0, cap = 0;
  double tol = 0;
  bool release = false;
  std::vector<double> b, x;
  int niter = 0, status = 0;
  double ini_res = 0, fin_res = 0;
  double t_submit = 0, t_start = 0, t_end = 0;
  bool done = false;
  bool cancelled = false;
};

// Device bytes of a resident case before its first solve (then replaced by
// the exact figure), an upper bound: per entry, col + source index in the S
// layout, the caller-order and S-order FP64 values, the FP64 sliced ELL (col,
// source index, value; up to 1.25 nnz slots with padding), and col + value in
// the L or U triangle (off-diagonal entries); per row, row pointers, maps,
// diagonal copies, the factor in two orders, two staging vectors and up to 12
// work vectors (Cgs); plus the fixed per-matrix buffers (state, the partial
// sums of the reductions, flags).
long long estimate_bytes(int n, long long nnz) {
  return 24 * nnz + 20 * nnz + 12 * (nnz - n) + 200 * (long long)n + (320 << 10);
}

}  // namespace

struct tfs_server {
  long long budget = 0, used = 0, peak = 0;
  int resident = 0;
  long long completed = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<long long> queue;
  std::map<long long, std::unique_ptr<Request>> reqs;
  std::map<int, std::unique_ptr<Case>> cases;
  long long next_req = 1;
  long long tick = 0;
  int next_case = 1;
  bool stop = false;
  std::vector<std::thread> workers;
  int nworkers = 1;  // set before the worker threads start
  Clock::time_point t0 = Clock::now();

  double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }

  // Under the lock: makes room for `need` bytes by evicting idle resident
  // cases that have no queued request, least recently used first.  The
  // handles to free go to `evict` (freed by the caller outside the lock);
  // false, with nothing evicted, if even that would not make room.
  bool make_room(long long need, std::vector<tfs_matrix *> *evict) {
    std::vector<Case *> idle;
    long long freeable = 0;
    for (auto &kv : cases) {
      Case *c = kv.second.get();
      if (c->A && !c->busy && c->queued == 0) {
        idle.push_back(c);
        freeable += c->bytes;
      }
    }
    if (used - freeable + need > budget) return false;
    std::sort(idle.begin(), idle.end(),
              [](const Case *a, const Case *b) { return a->last_use < b->last_use; });
    for (Case *c : idle) {
      if (used + need <= budget) break;
      evict->push_back(c->A);
      c->A = nullptr;
      used -= c->bytes;
      resident--;
    }
    return true;
  }

  // Under the lock: the first queued request that can run now.
  bool pick(long long *id, std::vector<tfs_matrix *> *evict) {
    for (auto it = queue.begin(); it != queue.end(); ++it) {
      Request &r = *reqs[*it];
      Case &c = *cases[r.case_id];
      if (c.busy) continue;
      if (!c.A && used + c.bytes > budget && !make_room(c.bytes, evict)) continue;
      *id = *it;
      queue.erase(it);
      c.queued--;
      return true;
    }
    return false;
  }

  void worker() {
    for (;;) {
      long long id = 0;
      Request *r = nullptr;
      Case *c = nullptr;
      bool create = false;
      std::vector<tfs_matrix *> evict;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || pick(&id, &evict); });
        if (id == 0) return;
        r = reqs[id].get();
        c = cases[r->case_id].get();
        c->busy = true;
        c->last_use = ++tick;
        if (!c->A) {
          create = true;
          used += c->bytes;  // reserve the estimate
          peak = std::max(peak, used);
          resident++;
        }
        r->t_start = now();
      }
      for (tfs_matrix *e : evict) tfs_matrix_destroy(e);
      int rc = 0;
      if (create) {
        rc = tfs_matrix_create(c->n, c->nnz, c->row.data(), c->col.data(), c->dia.data(), c->base,
                               &c->A);
        if (rc == 0) rc = tfs_matrix_set_sweep_schedule(c->A, TFS_SWEEP_STATIC);
        if (rc == 0)
          tfs::set_sweep_grid(c->A, tfs::sweep_capacity(c->A) / nworkers);
        if (rc == 0) rc = tfs_matrix_set_values(c->A, c->val.data());
      }
      if (rc == 0) {
        r->x.assign((size_t)c->n, 0.0);
        std::vector<double> rr = r->b;
        int it = r->cap;
        double ini = 0.0, fin = 0.0;
        rc = tfs_solve(c->A, r->method, r->prec, r->x.data(), rr.data(), &it, r->tol, &ini, &fin);
        r->niter = it;
        r->ini_res = ini;
        r->fin_res = fin;
        // the exact device bytes, now that the solve has allocated its values
        // and work vectors (the reservation was an estimate)
        tfs_matrix_info info;
        if (tfs_matrix_get_info(c->A, &info) == 0) {
          std::lock_guard<std::mutex> lk(mu);
          used += info.device_bytes - c->bytes;
          peak = std::max(peak, used);
          c->bytes = info.device_bytes;
        }
      }
      r->status = rc;
      const bool free_case = r->release || (create && rc < 0 && rc != TFS_ERR_BREAKDOWN);
      if (free_case && c->A) {
        tfs_matrix_destroy(c->A);
        c->A = nullptr;
      }
      {
        std::lock_guard<std::mutex> lk(mu);
        if (free_case) {
          used -= c->bytes;
          resident--;
          c->released = r->release;
        }
        c->busy = false;
        r->t_end = now();
        r->done = true;
        completed++;
      }
      cv_done.notify_all();
      cv_work.notify_all();
    }
  }
};

extern "C" int tfs_server_create(long long hbm_budget_bytes, int workers, tfs_server **out) {
  if (!out || hbm_budget_bytes <= 0 || workers < 1 || workers > 64) return TFS_ERR_INVALID;
  const int rc = tfs_init(0);
  if (rc < 0) return rc;
  auto *S = new tfs_server;
  S->budget = hbm_budget_bytes;
  S->nworkers = workers;
  for (int w = 0; w < workers; ++w) S->workers.emplace_back([S] { S->worker(); });
  *out = S;
  return TFS_OK;
}

extern "C" int tfs_server_destroy(tfs_server *S) {
  if (!S) return TFS_OK;
  {
    std::unique_lock<std::mutex> lk(S->mu);
    S->cv_done.wait(lk, [&] { return S->queue.empty(); });
    S->stop = true;
  }
  S->cv_work.notify_all();
  for (auto &t : S->workers) t.join();
  for (auto &kv : S->cases)
    if (kv.second->A) tfs_matrix_destroy(kv.second->A);
  delete S;
  return TFS_OK;
}

extern "C" int tfs_server_add_case(tfs_server *S, int n, long long nnz, const int *row,
                                   const int *col, const int *dia, const double *val, int base,
                                   int *case_id) {
  if (!S || !case_id || n < 1 || nnz < n || !row || !col || !dia || !val) return TFS_ERR_INVALID;
  auto c = std::make_unique<Case>();
  c->n = n;
  c->nnz = nnz;
  c->row.assign(row, row + n + 1);
  c->col.assign(col, col + nnz);
  c->dia.assign(dia, dia + n);
  c->val.assign(val, val + nnz);
  c->base = base;
  c->bytes = estimate_bytes(n, nnz);
  std::lock_guard<std::mutex> lk(S->mu);
  if (c->bytes > S->budget) return TFS_ERR_NOMEM;  // could never be admitted
  *case_id = S->next_case++;
  S->cases[*case_id] = std::move(c);
  return TFS_OK;
}

extern "C" int tfs_server_submit(tfs_server *S, int case_id, int method, int prec,
                                 const double *b, double tol, int cap, int release,
                                 long long *req) {
  if (!S || !b || !req || method < TFS_CG || method > TFS_CGS) return TFS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  auto it = S->cases.find(case_id);
  if (it == S->cases.end() || it->second->released) return TFS_ERR_INVALID;
  auto r = std::make_unique<Request>();
  r->case_id = case_id;
  r->method = method;
  r->prec = prec;
  r->tol = tol;
  r->cap = cap;
  r->release = release != 0;
  r->b.assign(b, b + it->second->n);
  r->t_submit = S->now();
  *req = S->next_req++;
  S->reqs[*req] = std::move(r);
  S->queue.push_back(*req);
  it->second->queued++;
  S->cv_work.notify_one();
  return TFS_OK;
}

extern "C" int tfs_server_wait(tfs_server *S, long long req, double *x, int *niter,
                               double *ini_res, double *fin_res, int *status, double *t_submit,
                               double *t_start, double *t_end) {
  if (!S) return TFS_ERR_INVALID;
  std::unique_lock<std::mutex> lk(S->mu);
  auto it = S->reqs.find(req);
  if (it == S->reqs.end()) return TFS_ERR_INVALID;
  Request &r = *it->second;
  S->cv_done.wait(lk, [&] { return r.done; });
  if (x && !r.x.empty()) std::memcpy(x, r.x.data(), sizeof(double) * r.x.size());
  if (niter) *niter = r.niter;
  if (ini_res) *ini_res = r.ini_res;
  if (fin_res) *fin_res = r.fin_res;
  if (status) *status = r.status;
  if (t_submit) *t_submit = r.t_submit;
  if (t_start) *t_start = r.t_start;
  if (t_end) *t_end = r.t_end;
  S->reqs.erase(it);
  return TFS_OK;
}

extern "C" int tfs_server_cancel(tfs_server *S, long long req) {
  if (!S) return TFS_ERR_INVALID;
  {
    std::lock_guard<std::mutex> lk(S->mu);
    auto q = std::find(S->queue.begin(), S->queue.end(), req);
    if (q == S->queue.end()) return TFS_ERR_INVALID;  // running, done or unknown
    S->queue.erase(q);
    Request &r = *S->reqs[req];
    S->cases[r.case_id]->queued--;
    r.status = TFS_CANCELLED;
    r.cancelled = true;
    r.t_start = r.t_end = S->now();
    r.done = true;
  }
  S->cv_done.notify_all();
  S->cv_work.notify_all();  // an eviction may have been waiting on this request
  return TFS_OK;
}

extern "C" int tfs_server_stats(tfs_server *S, long long *hbm_in_use, long long *hbm_peak,
                                int *resident_cases, long long *completed) {
  if (!S) return TFS_ERR_INVALID;
  std::lock_guard<std::mutex> lk(S->mu);
  if (hbm_in_use) *hbm_in_use = S->used;
  if (hbm_peak) *hbm_peak = S->peak;
  if (resident_cases) *resident_cases = S->resident;
  if (completed) *completed = S->completed;
  return TFS_OK;
}
