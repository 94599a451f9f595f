// serve.cpp — request loop over one resident GLR (no reference counterpart:
// the reference's downstream calls are one synchronous Vec at a time,
// glr_common_impl.h:1007-1104).
//
// Callers submit solve / sample / logdet requests (a shift c and m
// right-hand sides each) from any thread.  Threads serve them as a
// pipeline over pinned staging slots: the stager forms a batch from
// everything queued whenever a slot is free (all requests of the oldest
// request's operation that fit in the batch's columns, each column with its
// own shift) and copies its columns into the slot; device threads (three
// by default), each with its own streams and Z-solve work buffers
// (SolveWork), run batches in place in their slots, so a batch that forms
// while others run starts on the device at once and they overlap there
// (their coarse multigrid levels leave most of the device idle); the
// finisher copies the results out to the callers and frees the slot.  The
// host copies of one batch overlap the device work of the others.
// Requests that arrive after a batch is formed wait for the next one
// (continuous batching at batch granularity).  The batch width is capped by
// a device workspace budget (memory-based admission); submit blocks while
// the queued columns exceed a host-side budget (back-pressure).  Every
// column of a batch is computed independently of the others (column-wise
// kernels, per-column shifts and coefficients, a fixed panel width), so a
// request's result does not depend on what it was batched with.
#include "engine.hpp"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>

using namespace lgh;

namespace {

using clk = std::chrono::steady_clock;

struct Req {
  int64_t id = 0;
  int op = 0, m = 0;
  double c = 0.0;
  const double *in = nullptr;
  int ldi = 0;
  double *out = nullptr;
  int ldo = 0;
  clk::time_point t_submit, t_done;
  int done = 0, rc = 0;
  std::string err;                // the failure's message (device thread's)
  double logdet = 0.0;
};

}  // namespace

// one pinned staging block and the batch it holds
struct Slot {
  double *pin = nullptr;          // n x max_cols, inputs then results
  std::vector<Req *> batch;
  int cols = 0, rc = 0;
  std::string err;                // message of a failed batch
};

struct lgh_serve {
  lgh_glr_t *g = nullptr;
  lgh_serve_opts_t o;
  int max_cols = 64;              // columns per device batch (multiple of 64)
  int ndev = 2;                   // device threads (batches at once)
  std::mutex mu;
  std::condition_variable cv_work, cv_done, cv_space, cv_ready, cv_drain;
  std::deque<Req *> queue;
  std::map<int64_t, Req *> reqs;
  int64_t next_id = 1;
  int queued_cols = 0;
  bool stop = false;
  std::thread stager, finisher;
  std::vector<std::thread> devices;
  std::vector<Slot> slot;         // 2 x ndev
  // slot ids: free, staged (to run), run (to drain); -1 ends a thread
  std::deque<int> free_slots, ready, drain;
  // the running batches (start, predicted seconds) per device thread and
  // moving averages of the device seconds per 64-column panel (per
  // operation) and of the copy-in seconds per column, which time the next
  // batch's formation
  std::vector<int> dev_busy;
  std::vector<clk::time_point> dev_t0;
  std::vector<double> dev_pred;
  double ema_panel[2] = { 0.0, 0.0 }, ema_col = 0.0;
  lgh_serve_stats_t st;
  int tier = 0;
};

// copy m columns of n doubles (column j: src[j] -> dst[j]), split over a
// few host threads when the block is large
static void
copy_cols_par (int n, int m, const double *const *src, double *const *dst)
{
  const size_t bytes = sizeof (double) * (size_t) n * m;
  const int nt = bytes > (64u << 20) ? 8 : 1;
  auto part = [&] (int t) {
    for (int j = t; j < m; j += nt) memcpy (dst[j], src[j], sizeof (double) * n);
  };
  if (nt == 1) { part (0); return; }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; t++) th.emplace_back (part, t);
  for (auto &x : th) x.join ();
}

static double
since (clk::time_point t0)
{
  return std::chrono::duration<double> (clk::now () - t0).count ();
}

static void
ema (double &e, double x)
{
  e = (e == 0.0) ? x : 0.7 * e + 0.3 * x;
}

// Forms batches into free slots and copies their inputs in.  A batch is
// formed when a device thread is idle, when a full batch is queued, or,
// while every device thread runs a batch, once the copy-in of what is
// queued would end about when the first of those batches does: as late as
// keeps the device fed, so that requests arriving meanwhile still join it.
static void
stager_loop (lgh_serve *s)
{
  const int n = s->g->n;
  for (;;) {
    int id;
    Slot *sl;
    {
      std::unique_lock<std::mutex> lk (s->mu);
      for (;;) {
        if (s->stop && s->queue.empty ()) {   // stop requested, nothing queued
          for (int d = 0; d < s->ndev; d++) s->ready.push_back (-1);
          s->cv_ready.notify_all ();
          return;
        }
        if (!s->queue.empty () && !s->free_slots.empty () && s->ready.empty ()) {
          int waiting = 0, busy = 0;
          for (const Req *r : s->queue) waiting += r->m;
          for (int d = 0; d < s->ndev; d++) busy += s->dev_busy[d];
          if (busy < s->ndev || waiting >= s->max_cols || s->stop) break;
          // every device thread busy: the earliest predicted end
          bool known = true;
          clk::time_point first = clk::time_point::max ();
          for (int d = 0; d < s->ndev; d++) {
            if (s->dev_pred[d] <= 0.0) { known = false; break; }
            const clk::time_point e = s->dev_t0[d]
              + std::chrono::duration_cast<clk::duration> (
                  std::chrono::duration<double> (s->dev_pred[d]));
            if (e < first) first = e;
          }
          if (known) {
            const double lead = s->ema_col * std::min (waiting, s->max_cols);
            const clk::time_point at = first
              - std::chrono::duration_cast<clk::duration> (std::chrono::duration<double> (lead));
            if (clk::now () >= at) break;
            s->cv_work.wait_until (lk, at);
            continue;
          }
        }
        s->cv_work.wait (lk);
      }
      id = s->free_slots.front ();
      s->free_slots.pop_front ();
      sl = &s->slot[id];
      sl->batch.clear ();
      const int op = s->queue.front ()->op;
      int cols = 0;
      for (auto it = s->queue.begin (); it != s->queue.end ();) {
        Req *r = *it;
        if (r->op == op && cols + r->m <= s->max_cols) {
          sl->batch.push_back (r);
          cols += r->m;
          it = s->queue.erase (it);
          if (cols == s->max_cols) break;
        }
        else ++it;
      }
      sl->cols = cols;
    }
    std::vector<const double *> src (sl->cols);
    std::vector<double *> dst (sl->cols);
    int j = 0;
    for (Req *r : sl->batch)
      for (int q = 0; q < r->m; q++, j++) {
        src[j] = r->in + (size_t) q * r->ldi;
        dst[j] = sl->pin + (size_t) j * n;
      }
    const clk::time_point t0 = clk::now ();
    copy_cols_par (n, sl->cols, src.data (), dst.data ());
    const double t = since (t0);
    std::lock_guard<std::mutex> lk (s->mu);
    s->st.stage_s += t;
    ema (s->ema_col, t / sl->cols);
    s->ready.push_back (id);
    s->cv_ready.notify_one ();
  }
}

// device thread d: runs staged batches on the device, in place in their
// slots, with its own streams and Z-solve work buffers
static void
device_loop (lgh_serve *s, int d)
{
  (void) hipSetDevice (0);
  const int n = s->g->n;
  {
    SolveWork work;
    SolveWorkScope scope (&work);
    for (;;) {
      int id;
      {
        std::unique_lock<std::mutex> lk (s->mu);
        s->cv_ready.wait (lk, [&] { return !s->ready.empty (); });
        id = s->ready.front ();
        s->ready.pop_front ();
      }
      if (id < 0) break;
      Slot &sl = s->slot[id];
      std::vector<double> c (sl.cols);
      int j = 0;
      for (Req *r : sl.batch)
        for (int q = 0; q < r->m; q++, j++) c[j] = r->c;
      const int kind = sl.batch[0]->op == LGH_SERVE_SOLVE ? 0 : 1;
      const int panels = (sl.cols + 63) / 64;
      const clk::time_point t0 = clk::now ();
      {
        std::lock_guard<std::mutex> lk (s->mu);
        s->dev_busy[d] = 1;
        int running = 0;
        for (int k = 0; k < s->ndev; k++) running += s->dev_busy[k];
        if (running > s->st.max_running) s->st.max_running = running;
        s->dev_t0[d] = t0;
        s->dev_pred[d] = s->ema_panel[kind] * panels;
        s->cv_work.notify_one ();
      }
      sl.rc = glr_run_block (s->g, kind == 0 ? 0 : 3, sl.cols, c.data (),
                             sl.pin, n, sl.pin, n, s->tier);
      // the messages are this (device) thread's: keep them with the batch
      if (sl.rc) {
        const char *m = lgh_petsc_last_error ();
        sl.err = (m && *m) ? m : lgh::last_error ();
      }
      else sl.err.clear ();
      const double t = since (t0);
      std::lock_guard<std::mutex> lk (s->mu);
      s->st.busy_s += t;
      if (!sl.rc) ema (s->ema_panel[kind], t / panels);
      s->dev_busy[d] = 0;
      s->drain.push_back (id);
      s->cv_drain.notify_one ();
      s->cv_work.notify_one ();
    }
    (void) dev_sync ();   // the work buffers are freed with no kernel left on them
  }
  dev_thread_exit (false);  // the device thread's streams; its blocks freed
  std::lock_guard<std::mutex> lk (s->mu);
  s->drain.push_back (-1);
  s->cv_drain.notify_one ();
}

// copies the results out, completes the requests and frees the slots
static void
finisher_loop (lgh_serve *s)
{
  const int n = s->g->n;
  int ended = 0;                  // device threads that have finished
  for (;;) {
    int id;
    {
      std::unique_lock<std::mutex> lk (s->mu);
      s->cv_drain.wait (lk, [&] { return !s->drain.empty (); });
      id = s->drain.front ();
      s->drain.pop_front ();
    }
    if (id < 0) {
      if (++ended == s->ndev) return;
      continue;
    }
    Slot &sl = s->slot[id];
    double t = 0.0;
    if (!sl.rc) {
      std::vector<const double *> src (sl.cols);
      std::vector<double *> dst (sl.cols);
      int j = 0;
      for (Req *r : sl.batch)
        for (int q = 0; q < r->m; q++, j++) {
          src[j] = sl.pin + (size_t) j * n;
          dst[j] = r->out + (size_t) q * r->ldo;
        }
      const clk::time_point t0 = clk::now ();
      copy_cols_par (n, sl.cols, src.data (), dst.data ());
      t = since (t0);
    }
    const clk::time_point td = clk::now ();
    std::lock_guard<std::mutex> lk (s->mu);
    for (Req *r : sl.batch) {
      r->rc = sl.rc;
      if (sl.rc) r->err = sl.err;
      r->done = 1;
      r->t_done = td;
      s->queued_cols -= r->m;
    }
    s->st.stage_s += t;
    s->st.batches++;
    s->st.columns += sl.cols;
    s->st.requests += (int64_t) sl.batch.size ();
    if (sl.cols > s->st.max_batch_cols) s->st.max_batch_cols = sl.cols;
    sl.batch.clear ();
    s->free_slots.push_back (id);
    s->cv_done.notify_all ();
    s->cv_space.notify_all ();
    s->cv_work.notify_one ();
  }
}

extern "C" lgh_serve_opts_t
lgh_serve_opts_default (void)
{
  lgh_serve_opts_t o;
  o.max_batch_cols = 256;
  o.device_budget_gb = 32.0;
  o.queue_cols = 1024;
  // three batches on the device at once: at 125 requests/s of the 10 000-
  // request mix on the 1024^2 build, p50/p99 78/208 ms against 99/289 ms
  // with two, and 1 470-1 530 against 1 250-1 370 RHS/s saturated
  o.device_streams = 3;
  return o;
}

// device bytes one batch column needs: the block driver's five n-row
// panels (host stage, in, out, Woodbury in/out), the mixed Chebyshev state
// (R x2 FP64, D x2 FP32, V-cycle in/out) and the V-cycle levels
static double
bytes_per_col (const lgh_glr_t *g)
{
  return (double) g->n * (5 * 8 + 2 * 8 + 2 * 4 + 2 * 4 + 3 * 4);
}

extern "C" int
lgh_serve_create (lgh_glr_t *g, const lgh_serve_opts_t *opts, lgh_serve_t **out)
{
  PetscCheck (g != NULL && out != NULL, 0, PETSC_ERR_ARG_NULL,
              "lgh_serve_create: NULL argument");
  const lgh_serve_opts_t o = opts ? *opts : lgh_serve_opts_default ();
  PetscCheck (o.max_batch_cols >= 64 && o.device_budget_gb > 0 && o.queue_cols >= 64
              && o.device_streams >= 1 && o.device_streams <= 8,
              0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_serve_create: need max_batch_cols >= 64, a positive device "
              "budget, queue_cols >= 64 and 1 .. 8 device streams");
  int cols = (o.max_batch_cols / 64) * 64;
  const int fit = (int) (o.device_budget_gb * 1e9 / bytes_per_col (g)) / 64 * 64;
  if (fit < cols) cols = fit;
  PetscCheck (cols >= 64, 0, PETSC_ERR_MEM,
              "lgh_serve_create: the device budget (%.1f GB) holds no 64-column "
              "batch at n = %d", o.device_budget_gb, g->n);
  lgh_serve *s = new lgh_serve ();
  s->g = g;
  s->o = o;
  s->max_cols = cols;
  memset (&s->st, 0, sizeof (s->st));
  // the blocked tier when the GLR's policy enables it at all: then every
  // batch, whatever its width, takes the same path
  s->tier = g->x.solve_blocked_min > 0 ? 1 : 0;
  s->ndev = o.device_streams;
  s->slot.resize ((size_t) 2 * s->ndev);
  s->dev_busy.assign ((size_t) s->ndev, 0);
  s->dev_t0.assign ((size_t) s->ndev, clk::now ());
  s->dev_pred.assign ((size_t) s->ndev, 0.0);
  const size_t bytes = sizeof (double) * (size_t) g->n * cols;
  for (size_t k = 0; k < s->slot.size (); k++) {
    s->slot[k].pin = (double *) host_pinned_alloc (bytes);
    if (!s->slot[k].pin) {
      for (size_t q = 0; q < k; q++) host_pinned_free (s->slot[q].pin);
      const int ns = (int) s->slot.size ();
      delete s;
      return lgh_petsc_error (PETSC_ERR_MEM, "lgh_serve_create: pinned staging "
                              "(%.1f GB) failed", ns * (double) bytes / 1e9);
    }
    s->free_slots.push_back ((int) k);
  }
  s->finisher = std::thread (finisher_loop, s);
  for (int d = 0; d < s->ndev; d++) s->devices.emplace_back (device_loop, s, d);
  s->stager = std::thread (stager_loop, s);
  *out = s;
  return 0;
}

extern "C" int
lgh_serve_submit (lgh_serve_t *s, int op, double c, int m, const double *in,
                  int ldi, double *out, int ldo, int64_t *ticket)
{
  PetscCheck (s != NULL && ticket != NULL, 0, PETSC_ERR_ARG_NULL,
              "lgh_serve_submit: NULL argument");
  PetscCheck (op == LGH_SERVE_SOLVE || op == LGH_SERVE_SAMPLE
              || op == LGH_SERVE_LOGDET, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_serve_submit: unknown operation %d", op);
  PetscCheck (c > s->g->floor, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_serve_submit: shift c must exceed the spectrum floor %g",
              s->g->floor);
  Req *r = new Req ();
  r->op = op;
  r->c = c;
  r->t_submit = clk::now ();
  if (op == LGH_SERVE_LOGDET) {
    // host-only: the sum over the resident eigenvalues
    r->logdet = lgh_glr_logdet (s->g, c);
    r->done = 1;
    r->t_done = clk::now ();
    std::lock_guard<std::mutex> lk (s->mu);
    r->id = s->next_id++;
    s->reqs[r->id] = r;
    s->st.requests++;
    *ticket = r->id;
    return 0;
  }
  if (!(m >= 1 && m <= s->max_cols && in && out && ldi >= s->g->n
        && ldo >= s->g->n)) {
    delete r;
    return lgh_petsc_error (PETSC_ERR_ARG_OUTOFRANGE,
                            "lgh_serve_submit: need 1 <= m <= %d and full "
                            "columns (ld >= n)", s->max_cols);
  }
  r->m = m;
  r->in = in; r->ldi = ldi;
  r->out = out; r->ldo = ldo;
  std::unique_lock<std::mutex> lk (s->mu);
  s->cv_space.wait (lk, [&] { return s->queued_cols + m <= s->o.queue_cols
                                     || s->queued_cols == 0; });
  r->id = s->next_id++;
  s->reqs[r->id] = r;
  s->queue.push_back (r);
  s->queued_cols += m;
  *ticket = r->id;
  s->cv_work.notify_one ();
  return 0;
}

extern "C" int
lgh_serve_wait (lgh_serve_t *s, int64_t ticket, double *logdet,
                double *latency_s)
{
  PetscCheck (s != NULL, 0, PETSC_ERR_ARG_NULL, "lgh_serve_wait: NULL server");
  std::unique_lock<std::mutex> lk (s->mu);
  auto it = s->reqs.find (ticket);
  PetscCheck (it != s->reqs.end (), 0, PETSC_ERR_ARG_WRONG,
              "lgh_serve_wait: unknown ticket %lld", (long long) ticket);
  Req *r = it->second;
  s->cv_done.wait (lk, [&] { return r->done != 0; });
  s->reqs.erase (it);
  lk.unlock ();
  if (logdet) *logdet = r->logdet;
  if (latency_s)
    *latency_s = std::chrono::duration<double> (r->t_done - r->t_submit).count ();
  const int rc = r->rc;
  const std::string err = r->err;
  delete r;
  if (rc) return lgh_petsc_error (PETSC_ERR_LIB, "lgh_serve_wait: batch failed: %s",
                                  err.c_str ());
  return 0;
}

extern "C" int
lgh_serve_get_stats (lgh_serve_t *s, lgh_serve_stats_t *st)
{
  PetscCheck (s != NULL && st != NULL, 0, PETSC_ERR_ARG_NULL,
              "lgh_serve_get_stats: NULL argument");
  std::lock_guard<std::mutex> lk (s->mu);
  *st = s->st;
  st->batch_cols = s->max_cols;
  return 0;
}

extern "C" void
lgh_serve_destroy (lgh_serve_t *s)
{
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk (s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all ();
  // the stager ends once the queue is empty; its end markers (one per
  // device thread) pass through the device threads to the finisher behind
  // every batch
  if (s->stager.joinable ()) s->stager.join ();
  for (auto &t : s->devices)
    if (t.joinable ()) t.join ();
  if (s->finisher.joinable ()) s->finisher.join ();
  for (auto &kv : s->reqs) delete kv.second;
  for (auto &sl : s->slot) host_pinned_free (sl.pin);
  delete s;
}
