// device.hip — context, streams, caching allocator, CSR upload, rocSOLVER
// glue.  One process drives one GPU (device 0).  Every host thread that
// calls the engine gets its own non-blocking stream at default priority,
// rocBLAS handle and info word (created on its first call); a thread's
// kernels all go to its stream, so the builds of a batch run concurrently
// from a pool of threads.
#include "lgh_dev.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <atomic>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

namespace lgh {

static thread_local std::string g_err;
void set_error (const std::string &m) { g_err = m; }
const char *last_error () { return g_err.c_str (); }

int
hip_fail (hipError_t e, const char *what, const char *file, int line)
{
  char buf[512];
  snprintf (buf, sizeof buf, "%s failed: %s (%s:%d)", what,
            hipGetErrorString (e), file, line);
  g_err = buf;
  return 2;
}

static bool g_init = false;
static std::mutex g_mu;

struct ThreadCtx {
  hipStream_t s = nullptr;
  rocblas_handle rb = nullptr;
  int *info = nullptr;
  hipStream_t main = nullptr;     // s outside any StreamScope
  hipStream_t aux = nullptr;
};
static thread_local ThreadCtx t_ctx;

static int
ctx_init (ThreadCtx &c)
{
  if (c.s) return 0;
  LGH_HIP (hipSetDevice (0));
  LGH_HIP (hipStreamCreateWithFlags (&c.s, hipStreamNonBlocking));
  c.main = c.s;
  if (rocblas_create_handle (&c.rb) != rocblas_status_success) {
    set_error ("rocblas_create_handle failed");
    return 4;
  }
  rocblas_set_stream (c.rb, c.s);
  // no atomics in rocBLAS reductions (rocSOLVER's potrf / trtri / syevd run
  // through them): results do not depend on the timing of other work
  rocblas_set_atomics_mode (c.rb, rocblas_atomics_not_allowed);
  LGH_HIP (hipMalloc (&c.info, sizeof (int) * 4));
  return 0;
}

int
dev_init ()
{
  {
    std::lock_guard<std::mutex> lk (g_mu);
    if (!g_init) {
      int ndev = 0;
      LGH_HIP (hipGetDeviceCount (&ndev));
      LGH_CHECK (ndev > 0, 3, "no HIP device visible: the engine needs a GPU");
      LGH_HIP (hipSetDevice (0));
      // the no-allocation entry points' scratch (kernels.hip), with the
      // device current
      LGH_TRY (scratch_init ());
      g_init = true;
    }
  }
  return ctx_init (t_ctx);
}

bool
dev_ready ()
{
  return g_init && t_ctx.s != nullptr;
}

hipStream_t
stream ()
{
  if (!t_ctx.s) (void) dev_init ();
  return t_ctx.s;
}
void *rocblas_h () { (void) stream (); return (void *) t_ctx.rb; }

// a thread that runs one of many concurrent instances (batch workers)
// gains nothing from overlapping a block's narrow last panel on a second
// stream: the other instances fill the device
static thread_local bool t_no_aux = false;
void set_aux_overlap (bool on) { t_no_aux = !on; }
static thread_local bool t_no_tail = false;
void set_tail_split (bool on) { t_no_tail = !on; }
bool tail_split () { return !t_no_tail; }

hipStream_t
aux_stream ()
{
  if (t_no_aux) return nullptr;
  (void) stream ();
  if (!t_ctx.aux
      && hipStreamCreateWithFlags (&t_ctx.aux, hipStreamNonBlocking) != hipSuccess)
    t_ctx.aux = nullptr;
  return t_ctx.aux;
}

StreamScope::StreamScope (hipStream_t s) : prev (stream ()) { t_ctx.s = s; }
StreamScope::~StreamScope () { t_ctx.s = prev; }

// Events that cross-stream waits were enqueued on are not destroyed while
// such a wait may still be pending: they are kept here and destroyed once
// the thread's main stream has been synchronized (every wait enqueued
// before then has completed; the auxiliary stream's own waits precede
// work that the main stream waited for).
static thread_local std::vector<hipEvent_t> t_retired;

void
event_retire (hipEvent_t e)
{
  if (e) t_retired.push_back (e);
}

static void
retire_flush ()
{
  if (t_ctx.s != t_ctx.main) return;
  for (hipEvent_t e : t_retired) (void) hipEventDestroy (e);
  t_retired.clear ();
}

int
dev_sync ()
{
  LGH_HIP (hipStreamSynchronize (t_ctx.s));
  retire_flush ();
  return 0;
}

static void free_list_release (hipStream_t s, bool keep);
static void stage_release ();

void
dev_thread_exit (bool keep_blocks)
{
  ThreadCtx &c = t_ctx;
  if (!c.main) return;
  (void) hipStreamSynchronize (c.main);
  free_list_release (c.main, keep_blocks);
  if (c.aux) {
    (void) hipStreamSynchronize (c.aux);
    free_list_release (c.aux, keep_blocks);
    (void) hipStreamDestroy (c.aux);
  }
  for (hipEvent_t e : t_retired) (void) hipEventDestroy (e);
  t_retired.clear ();
  stage_release ();
  randn_host_release ();
  (void) hipFree (c.info);
  rocblas_destroy_handle (c.rb);
  (void) hipStreamDestroy (c.main);
  c = ThreadCtx ();
}

// ---- live kernel timing -----------------------------------------------------
struct ProfRec { int kind, np, prec, W; hipEvent_t a, b; double bytes, flops; };
static int g_prof = 0;
static thread_local int g_btag = 0, g_ftag = 0;
static std::mutex g_pmu;                 // records and event pool
static std::vector<ProfRec> g_recs;
static std::vector<hipEvent_t> g_evpool;

static hipEvent_t
ev_get ()
{
  std::lock_guard<std::mutex> lk (g_pmu);
  if (!g_evpool.empty ()) { hipEvent_t e = g_evpool.back (); g_evpool.pop_back (); return e; }
  hipEvent_t e = nullptr;
  (void) hipEventCreate (&e);
  return e;
}

void prof_enable (int on) { g_prof = on; }
int prof_enabled () { return g_prof; }
void prof_set_b_tag (int on) { g_btag = on; }
int prof_b_tag () { return g_btag; }
void prof_set_fine_tag (int on) { g_ftag = on; }
int prof_fine_tag () { return g_ftag; }

ProfScope::ProfScope (int k, double by, double fl, int np_, int prec_, int W_)
  : kind (k), np (np_), prec (prec_), W (W_), bytes (by), flops (fl), active (g_prof),
    ev0 (nullptr)
{
  if (!active) return;
  hipEvent_t a = ev_get ();
  (void) hipEventRecord (a, t_ctx.s);
  ev0 = (void *) a;
}

ProfScope::~ProfScope ()
{
  if (!active) return;
  hipEvent_t b = ev_get ();
  (void) hipEventRecord (b, t_ctx.s);
  std::lock_guard<std::mutex> lk (g_pmu);
  g_recs.push_back ({ kind, np, prec, W, (hipEvent_t) ev0, b, bytes, flops });
}

int
prof_collect (int kind, int np, int prec, long *count, double *ms,
              double *bytes, double *flops)
{
  std::lock_guard<std::mutex> lk (g_pmu);
  long c = 0;
  double t = 0, by = 0, fl = 0;
  std::vector<ProfRec> keep;
  for (auto &r : g_recs) {
    if (r.kind != kind) { keep.push_back (r); continue; }
    if ((np < 0 || (r.np == np && r.W == 64)) && (prec < 0 || r.prec == prec)) {
      float e = 0.f;
      LGH_HIP (hipEventSynchronize (r.b));
      LGH_HIP (hipEventElapsedTime (&e, r.a, r.b));
      c++; t += e; by += r.bytes; fl += r.flops;
    }
    g_evpool.push_back (r.a);
    g_evpool.push_back (r.b);
  }
  g_recs.swap (keep);
  *count = c; *ms = t; *bytes = by; *flops = fl;
  return 0;
}

// ---- caching allocator ----------------------------------------------------
// Blocks are rounded up (256 B granularity below 1 MiB, 2 MiB above) and
// recycled by exact rounded size.  A freed block goes to the free list of
// the freeing thread's stream: work queued on that stream may still use
// it, and only later work on the same stream is ordered after that use.
// A thread that leaves the engine (dev_thread_exit) synchronizes its
// stream and hands its blocks to the shared list (key nullptr), which any
// stream may take from.
static std::map<hipStream_t, std::multimap<size_t, void *>> g_free;
static std::map<void *, size_t> g_live;
static size_t g_in_use = 0;

static size_t
round_up (size_t b)
{
  if (b < ((size_t) 1 << 20)) return (b + 255) & ~(size_t) 255;
  const size_t g = (size_t) 2 << 20;
  return (b + g - 1) / g * g;
}

// Cached (freed, not yet reused) bytes are bounded: a request takes a
// cached block of at most 1/8 more than it asks for, and a request that
// the cache cannot serve first releases every cached block when they
// exceed kCacheFrac of the device's memory.  Without the bound, work of
// ever-changing shapes (serving batches of 1 .. 256 columns) piled up
// exact-size blocks until the runtime's own allocations failed.
static size_t g_cached = 0;
static constexpr double kCacheFrac = 0.35;

static bool
take (hipStream_t key, size_t b, void **p, size_t *got)
{
  auto l = g_free.find (key);
  if (l == g_free.end ()) return false;
  auto it = l->second.lower_bound (b);
  if (it == l->second.end () || it->first > b + b / 8) return false;
  *p = it->second;
  *got = it->first;
  g_cached -= it->first;
  l->second.erase (it);
  return true;
}

// release every cached block (g_mu held); the device is idle first, so no
// stream still uses one
static void
flush_cache_locked ()
{
  (void) hipDeviceSynchronize ();
  for (auto &l : g_free)
    for (auto &kv : l.second) (void) hipFree (kv.second);
  g_free.clear ();
  g_cached = 0;
}

// a leaving thread's free list (its stream synchronized): to the shared
// list, or freed
static void
free_list_release (hipStream_t s, bool keep)
{
  std::lock_guard<std::mutex> lk (g_mu);
  auto l = g_free.find (s);
  if (l == g_free.end ()) return;
  if (keep) {
    auto &shared = g_free[nullptr];
    for (auto &kv : l->second) shared.insert (kv);
  }
  else {
    for (auto &kv : l->second) {
      (void) hipFree (kv.second);
      g_cached -= kv.first;
    }
  }
  g_free.erase (s);
}

void *
dmalloc (size_t bytes)
{
  if (dev_init ()) return nullptr;
  const hipStream_t s = t_ctx.s;
  const size_t cap = (size_t) (kCacheFrac * (double) dmem_total ());
  std::lock_guard<std::mutex> lk (g_mu);
  size_t b = round_up (bytes ? bytes : 1), got = 0;
  void *p = nullptr;
  if (!take (s, b, &p, &got) && !take (nullptr, b, &p, &got)) {
    if (cap && g_cached + b > cap) flush_cache_locked ();
    hipError_t e = hipMalloc (&p, b);
    if (e != hipSuccess) {
      // release every cached block and retry once
      flush_cache_locked ();
      e = hipMalloc (&p, b);
      if (e != hipSuccess) {
        char buf[256];
        snprintf (buf, sizeof buf, "device allocation of %zu bytes failed: %s",
                  b, hipGetErrorString (e));
        g_err = buf;
        return nullptr;
      }
    }
    got = b;
  }
  g_live[p] = got;
  g_in_use += got;
  return p;
}

void
dfree (void *p)
{
  if (!p) return;
  const hipStream_t s = t_ctx.s;
  std::lock_guard<std::mutex> lk (g_mu);
  auto it = g_live.find (p);
  if (it == g_live.end ()) return;
  g_in_use -= it->second;
  g_cached += it->second;
  g_free[s].insert ({ it->second, p });
  g_live.erase (it);
}

void
dtrim ()
{
  std::lock_guard<std::mutex> lk (g_mu);
  flush_cache_locked ();
}

size_t dbytes_in_use () { return g_in_use; }

// ---- host <-> device transfers -------------------------------------------
// A pageable block goes through the runtime's own staging at ~31 GB/s in
// (17 ms for 537 MB) and ~17 GB/s out, 9 GB/s into freshly allocated
// memory, whose pages fault one at a time; from pinned memory the DMA runs
// at ~57 GB/s and an 8-thread host copy at ~120 GB/s.  So large pageable
// blocks are staged through two pinned chunks: the host copy of one chunk
// (which also takes a fresh destination's page faults on 8 threads)
// overlaps the DMA of the other.
static constexpr size_t kStageChunk = size_t (64) << 20;

static std::mutex g_pin_mu;
static std::vector<std::pair<const char *, size_t>> g_pinned;

void *
host_pinned_alloc (size_t bytes)
{
  void *p = nullptr;
  if (hipHostMalloc (&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
    (void) hipGetLastError ();
    set_error ("hipHostMalloc of " + std::to_string (bytes) + " bytes failed");
    return nullptr;
  }
  std::lock_guard<std::mutex> lk (g_pin_mu);
  g_pinned.emplace_back ((const char *) p, bytes);
  return p;
}

void
host_pinned_free (void *p)
{
  if (!p) return;
  {
    std::lock_guard<std::mutex> lk (g_pin_mu);
    for (size_t i = 0; i < g_pinned.size (); i++)
      if (g_pinned[i].first == (const char *) p) {
        g_pinned.erase (g_pinned.begin () + (long) i);
        break;
      }
  }
  (void) hipHostFree (p);
}

static bool
own_pinned (const void *p, size_t bytes)
{
  const char *c = (const char *) p;
  std::lock_guard<std::mutex> lk (g_pin_mu);
  for (const auto &r : g_pinned)
    if (c >= r.first && c + bytes <= r.first + r.second) return true;
  return false;
}

// the calling thread's two pinned chunks and their events
struct Stage {
  char *pin[2] = { nullptr, nullptr };
  hipEvent_t ev[2] = { nullptr, nullptr };
  bool failed = false;            // could not be set up: copy directly
};
static thread_local Stage t_stage;

// nullptr when the chunks cannot be had (the caller then copies directly)
static Stage *
stage_get ()
{
  Stage *st = &t_stage;
  if (st->ev[1]) return st;
  if (st->failed) return nullptr;
  for (int b = 0; b < 2; b++) {
    if ((!st->pin[b]
         && hipHostMalloc ((void **) &st->pin[b], kStageChunk, hipHostMallocDefault)
            != hipSuccess)
        || (!st->ev[b]
            && hipEventCreateWithFlags (&st->ev[b], hipEventDisableTiming) != hipSuccess)) {
      (void) hipGetLastError ();
      st->failed = true;
      return nullptr;
    }
  }
  return st;
}

static void
stage_release ()
{
  for (int b = 0; b < 2; b++) {
    if (t_stage.pin[b]) (void) hipHostFree (t_stage.pin[b]);
    if (t_stage.ev[b]) (void) hipEventDestroy (t_stage.ev[b]);
  }
  t_stage = Stage ();
}

// host memcpy split over 8 threads in contiguous parts
static void
par_memcpy (char *dst, const char *src, size_t bytes)
{
  constexpr int NT = 8;
  const size_t part = (bytes + NT - 1) / NT;
  std::thread th[NT - 1];
  for (int t = 1; t < NT; t++) {
    const size_t o = std::min (bytes, part * t), e = std::min (bytes, o + part);
    th[t - 1] = std::thread ([=] { if (e > o) memcpy (dst + o, src + o, e - o); });
  }
  memcpy (dst, src, std::min (bytes, part));
  for (auto &x : th) x.join ();
}

int
h2d (void *dst, const void *src, size_t bytes)
{
  if (!bytes) return 0;
  hipStream_t s = stream ();
  Stage *st = (bytes < kStageChunk || own_pinned (src, bytes)) ? nullptr : stage_get ();
  if (!st)
    LGH_HIP (hipMemcpyAsync (dst, src, bytes, hipMemcpyHostToDevice, s));
  else {
    bool rec[2] = { false, false };
    for (size_t off = 0, i = 0; off < bytes; off += kStageChunk, i++) {
      const size_t len = std::min (kStageChunk, bytes - off);
      const int b = (int) (i & 1);
      if (rec[b]) LGH_HIP (hipEventSynchronize (st->ev[b]));   // chunk i - 2 sent
      par_memcpy (st->pin[b], (const char *) src + off, len);
      LGH_HIP (hipMemcpyAsync ((char *) dst + off, st->pin[b], len,
                               hipMemcpyHostToDevice, s));
      LGH_HIP (hipEventRecord (st->ev[b], s));
      rec[b] = true;
    }
  }
  LGH_HIP (hipStreamSynchronize (s));
  retire_flush ();
  return 0;
}

int
d2h (void *dst, const void *src, size_t bytes)
{
  if (!bytes) return 0;
  hipStream_t s = stream ();
  Stage *st = (bytes < kStageChunk || own_pinned (dst, bytes)) ? nullptr : stage_get ();
  if (!st)
    LGH_HIP (hipMemcpyAsync (dst, src, bytes, hipMemcpyDeviceToHost, s));
  else {
    const size_t nch = (bytes + kStageChunk - 1) / kStageChunk;
    auto len = [&] (size_t i) { return std::min (kStageChunk, bytes - i * kStageChunk); };
    for (size_t i = 0; i <= nch; i++) {
      // the DMA of chunk i runs while chunk i - 1 is copied out; chunk i's
      // buffer last held chunk i - 2, copied out in the previous iteration
      if (i < nch) {
        LGH_HIP (hipMemcpyAsync (st->pin[i & 1], (const char *) src + i * kStageChunk,
                                 len (i), hipMemcpyDeviceToHost, s));
        LGH_HIP (hipEventRecord (st->ev[i & 1], s));
      }
      if (i >= 1) {
        const size_t j = i - 1;
        LGH_HIP (hipEventSynchronize (st->ev[j & 1]));
        par_memcpy ((char *) dst + j * kStageChunk, st->pin[j & 1], len (j));
      }
    }
  }
  LGH_HIP (hipStreamSynchronize (s));
  retire_flush ();
  return 0;
}

int
dcopy (void *dst, const void *src, size_t bytes)
{
  if (!bytes) return 0;
  LGH_HIP (hipMemcpyAsync (dst, src, bytes, hipMemcpyDeviceToDevice, stream ()));
  return 0;
}

int
dcopy2d (void *dst, size_t dpitch, const void *src, size_t spitch,
         size_t width, size_t height)
{
  if (!width || !height) return 0;
  if (dpitch == width && spitch == width) return dcopy (dst, src, width * height);
  LGH_HIP (hipMemcpy2DAsync (dst, dpitch, src, spitch, width, height,
                             hipMemcpyDeviceToDevice, stream ()));
  return 0;
}

int
dev_block_alloc (size_t bytes, void **p)
{
  *p = nullptr;
  LGH_TRY (dev_init ());
  LGH_CHECK (bytes > 0, 8, "device block: size must be positive");
  LGH_HIP (hipMalloc (p, bytes));
  return 0;
}

void
dev_block_free (void *p)
{
  if (p) (void) hipFree (p);
}

bool
is_device_ptr (const void *p)
{
  hipPointerAttribute_t a;
  if (!p || hipPointerGetAttributes (&a, p) != hipSuccess) {
    (void) hipGetLastError ();
    return false;
  }
  return a.type == hipMemoryTypeDevice || a.type == hipMemoryTypeManaged;
}

bool
device_range_ok (const void *p, size_t bytes)
{
  if (!is_device_ptr (p)) return false;
  hipDeviceptr_t base = nullptr;
  size_t size = 0;
  if (hipMemGetAddressRange (&base, &size, (hipDeviceptr_t) p) != hipSuccess) {
    (void) hipGetLastError ();
    return false;
  }
  return (const char *) p + bytes <= (const char *) base + size;
}

// ---- CSR / panel matrices -------------------------------------------------
void
GroupCopy::release ()
{
  dfree (gp); dfree (gci); dfree (gv);
  gp = gci = nullptr;
  gv = nullptr;
  rg = 0;
  nbatch = 0;
}

void
TileCopy::release ()
{
  dfree (tp); dfree (tu); dfree (cb); dfree (es); dfree (trow); dfree (ev);
  *this = TileCopy ();
}

DevCSR::~DevCSR ()
{
  dfree (rp); dfree (ci); dfree (v); dfree (dinv);
  grp.release ();
  tile.release ();
  dfree (vs);
}

// Each group's union is the merge of its (sorted) rows; a row's entries
// keep their order inside it, so a product accumulated over the union in
// order adds the same terms in the same order as over the row (plus
// exact zeros).  The union is padded to whole batches of grp_batch(rg)
// entries with zero values at the last column (a valid index).  Two
// passes over the groups (sizes, then fill), each split over host threads.
int
csr_group_build (const DevCSR &A, const int *rp, const int *ci,
                 const double *v, int rg, GroupCopy &out)
{
  LGH_CHECK (rg == 4 || rg == 8, 8, "csr_group_build: rows per group 4 or 8");
  const int n = A.n, ng = (n + rg - 1) / rg, nb = grp_batch (rg);
  for (int i = 0; i < n; i++)
    for (int k = rp[i] + 1; k < rp[i + 1]; k++)
      LGH_CHECK (ci[k] > ci[k - 1], 8, "csr_group_build: unsorted row");
  std::vector<int> gp ((size_t) ng + 1, 0);
  // entries of group g's union; with cols/vals, also writes them (padded)
  auto merge = [&] (int g, int *cols, double *vals) -> int {
    int p[8], e[8];
    for (int r = 0; r < rg; r++) {
      const int i = g * rg + r;
      p[r] = (i < n) ? rp[i] : 0;
      e[r] = (i < n) ? rp[i + 1] : 0;
    }
    int cnt = 0, last = 0;
    for (;;) {
      int c = INT_MAX;
      for (int r = 0; r < rg; r++) if (p[r] < e[r]) c = std::min (c, ci[p[r]]);
      if (c == INT_MAX) break;
      for (int r = 0; r < rg; r++) {
        const bool hit = p[r] < e[r] && ci[p[r]] == c;
        if (vals) vals[(size_t) cnt * rg + r] = hit ? v[p[r]] : 0.0;
        if (hit) p[r]++;
      }
      if (cols) cols[cnt] = c;
      last = c;
      cnt++;
    }
    const int padded = (cnt + nb - 1) / nb * nb;
    for (int k = cnt; k < padded; k++) {
      if (cols) cols[k] = last;
      if (vals) for (int r = 0; r < rg; r++) vals[(size_t) k * rg + r] = 0.0;
    }
    return padded;
  };
  const int nt = (int) std::max (1u, std::min (16u, std::thread::hardware_concurrency ()));
  auto par = [&] (auto &&body) {
    std::vector<std::thread> th;
    for (int t = 0; t < nt; t++)
      th.emplace_back ([&, t] {
        for (int g = (int) ((int64_t) ng * t / nt); g < (int) ((int64_t) ng * (t + 1) / nt); g++)
          body (g);
      });
    for (auto &x : th) x.join ();
  };
  std::vector<int64_t> off ((size_t) ng + 1, 0);
  par ([&] (int g) { off[g + 1] = merge (g, nullptr, nullptr); });
  for (int g = 0; g < ng; g++) off[g + 1] += off[g];
  const int64_t nu = off[ng];
  LGH_CHECK (nu / nb < INT_MAX, 8, "csr_group_build: too many batches");
  for (int g = 0; g <= ng; g++) gp[g] = (int) (off[g] / nb);     // in batches
  std::vector<int> cols ((size_t) std::max<int64_t> (nu, 1));
  std::vector<double> vals ((size_t) std::max<int64_t> (nu, 1) * rg);
  par ([&] (int g) { merge (g, cols.data () + off[g], vals.data () + (size_t) off[g] * rg); });
  const size_t vb = prec_bytes (A.prec);
  GroupCopy c;
  c.rg = rg;
  c.nbatch = nu / nb;
  c.gp = (int *) dmalloc (sizeof (int) * ((size_t) ng + 1));
  c.gci = (int *) dmalloc (sizeof (int) * cols.size ());
  c.gv = dmalloc (vb * vals.size ());
  if (!c.gp || !c.gci || !c.gv) { c.release (); return 5; }
  int rc = h2d (c.gp, gp.data (), sizeof (int) * gp.size ());
  if (!rc) rc = h2d (c.gci, cols.data (), sizeof (int) * cols.size ());
  if (!rc) {
    if (A.prec == F64) rc = h2d (c.gv, vals.data (), sizeof (double) * vals.size ());
    else {
      std::vector<float> f (vals.begin (), vals.end ());
      rc = h2d (c.gv, f.data (), sizeof (float) * f.size ());
    }
  }
  if (rc) { c.release (); return rc; }
  out.release ();
  out = c;
  return 0;
}

// the rows of a tile of shape s (-1: none), local index e = w * rg + r.
// 2D tiles sit on an ntx x nty grid of tiles, tile t at column t % ntx of
// band t / ntx (the SpMM launches may deal them to the XCDs in strips)
static inline int
tile_row (const TileShape &s, int n, int ntx, int nty, int t, int e)
{
  const int T = s.rg * s.nw;
  if (!s.tx) {
    const int64_t r = (int64_t) t * T + e;
    return r < n ? (int) r : -1;
  }
  (void) nty;
  const int tcol = t % ntx, tband = t / ntx;
  // wavefront w: a (rg/wy) x wy block of the tile, wpl blocks across
  const int ty = T / s.tx, wx = s.rg / s.wy, wpl = s.tx / wx;
  const int w = e / s.rg, r = e % s.rg;
  const int x = tcol * s.tx + (w % wpl) * wx + r % wx;
  const int64_t y = (int64_t) tband * ty + (int64_t) (w / wpl) * s.wy + r / wx;
  const int64_t row = y * s.stride + x;
  return (x < s.stride && row < n) ? (int) row : -1;
}

// sorted union of the column indices of rows[0 .. nr) (skipping -1)
static void
rows_union (const int *rp, const int *ci, const int *rows, int nr,
            std::vector<int> &u)
{
  u.clear ();
  for (int e = 0; e < nr; e++)
    if (rows[e] >= 0) u.insert (u.end (), ci + rp[rows[e]], ci + rp[rows[e] + 1]);
  std::sort (u.begin (), u.end ());
  u.erase (std::unique (u.begin (), u.end ()), u.end ());
}

// A tile's union is sorted, so a wavefront's merged rows (each in column
// order) meet its entries chunk by chunk in column order: row r of the
// wavefront accumulates its own entries in the order of spmm64_k, plus
// exact zeros.  Two passes over the tiles (sizes, then fill), each split
// over host threads.
int
csr_tile_build (const DevCSR &A, const int *rp, const int *ci,
                const double *v, const TileShape &s, TileCopy &out)
{
  LGH_CHECK ((s.rg == 4 || s.rg == 8 || s.rg == 16) && s.nw >= 1
             && s.rg * s.nw <= 128 && s.ch >= 8 && s.ch <= 256 && s.wy >= 1
             && s.rg % s.wy == 0, 8, "csr_tile_build: tile shape");
  const int T = s.rg * s.nw;
  LGH_CHECK (!s.tx || (s.stride >= s.tx && T % s.tx == 0 && s.tx % (s.rg / s.wy) == 0
                       && (T / s.tx) % s.wy == 0),
             8, "csr_tile_build: 2D tile shape");
  const int n = A.n, nb = tile_batch (s.rg), nw = s.nw, ch = s.ch, rg = s.rg;
  for (int i = 0; i < n; i++)
    for (int k = rp[i] + 1; k < rp[i + 1]; k++)
      LGH_CHECK (ci[k] > ci[k - 1], 8, "csr_tile_build: unsorted row");
  int ntx = 0, nty = 0;
  int64_t nt64;
  if (s.tx) {
    const int64_t lines = ((int64_t) n + s.stride - 1) / s.stride;
    ntx = (s.stride + s.tx - 1) / s.tx;
    const int64_t ny = (lines + T / s.tx - 1) / (T / s.tx);
    LGH_CHECK (ny < INT_MAX / 2, 8, "csr_tile_build: too many tiles");
    nty = (int) ny;
    nt64 = (int64_t) ntx * nty;
  }
  else nt64 = ((int64_t) n + T - 1) / T;
  LGH_CHECK (nt64 < INT_MAX / 2, 8, "csr_tile_build: too many tiles");
  const int nt = (int) nt64;
  // per tile: chunks, and batches per (chunk, wavefront) in a flat list
  std::vector<int> nch ((size_t) nt, 0);
  std::vector<std::vector<int>> bcnt ((size_t) nt);
  // one tile; fill: write its chunks' rows and its batches from the offsets
  auto work = [&] (int t, bool fill, int *tu, int *es, double *ev,
                   int64_t bat0) {
    int rows[128];
    for (int e = 0; e < T; e++) rows[e] = tile_row (s, n, ntx, nty, t, e);
    std::vector<int> u;
    rows_union (rp, ci, rows, T, u);
    const int nc = (int) ((u.size () + ch - 1) / ch);
    if (fill)
      for (int j = 0; j < nc * ch; j++)
        tu[j] = j < (int) u.size () ? u[j] : u.back ();
    else {
      nch[t] = nc;
      bcnt[t].assign ((size_t) nc * nw, 0);
    }
    // per wavefront: its entries (slot, values) chunk by chunk
    std::vector<int64_t> bpos ((size_t) nc * nw);
    if (fill) {                       // a wavefront's batches: one run, chunk by chunk
      int64_t b = bat0;
      for (int w = 0; w < nw; w++)
        for (int c = 0; c < nc; c++) { bpos[(size_t) c * nw + w] = b; b += bcnt[t][(size_t) c * nw + w]; }
    }
    for (int w = 0; w < nw; w++) {
      int p[16], e[16];
      for (int r = 0; r < rg; r++) {
        const int i = rows[w * rg + r];
        p[r] = i >= 0 ? rp[i] : 0;
        e[r] = i >= 0 ? rp[i + 1] : 0;
      }
      int j = 0, cur = -1, inchunk = 0, last_slot = 0;
      int64_t ent = 0;                  // entries written in the current chunk
      auto close_chunk = [&] () {
        if (cur < 0) return;
        const int padded = (inchunk + nb - 1) / nb * nb;
        if (!fill) bcnt[t][(size_t) cur * nw + w] = padded / nb;
        else
          for (int k = inchunk; k < padded; k++) {
            const int64_t q = bpos[(size_t) cur * nw + w] * nb + k;
            es[q] = last_slot;
            for (int r = 0; r < rg; r++) ev[q * rg + r] = 0.0;
          }
      };
      for (;;) {
        int c = INT_MAX;
        for (int r = 0; r < rg; r++) if (p[r] < e[r]) c = std::min (c, ci[p[r]]);
        if (c == INT_MAX) break;
        while (u[j] < c) j++;
        const int chunk = j / ch, slot = j % ch;
        if (chunk != cur) { close_chunk (); cur = chunk; inchunk = 0; ent = 0; }
        if (fill) {
          const int64_t q = bpos[(size_t) cur * nw + w] * nb + ent;
          es[q] = slot;
          for (int r = 0; r < rg; r++) {
            const bool hit = p[r] < e[r] && ci[p[r]] == c;
            ev[q * rg + r] = hit ? v[p[r]] : 0.0;
          }
        }
        for (int r = 0; r < rg; r++) if (p[r] < e[r] && ci[p[r]] == c) p[r]++;
        last_slot = slot;
        inchunk++;
        ent++;
      }
      close_chunk ();
    }
  };
  const int nth = (int) std::max (1u, std::min (16u, std::thread::hardware_concurrency ()));
  auto par = [&] (auto &&body) {
    std::vector<std::thread> th;
    for (int k = 0; k < nth; k++)
      th.emplace_back ([&, k] {
        for (int t = (int) ((int64_t) nt * k / nth); t < (int) ((int64_t) nt * (k + 1) / nth); t++)
          body (t);
      });
    for (auto &x : th) x.join ();
  };
  par ([&] (int t) { work (t, false, nullptr, nullptr, nullptr, 0); });
  std::vector<int> tp ((size_t) nt + 1, 0);
  std::vector<int64_t> boff ((size_t) nt + 1, 0);
  for (int t = 0; t < nt; t++) {
    LGH_CHECK ((int64_t) tp[t] + nch[t] < INT_MAX / 2, 8, "csr_tile_build: too many chunks");
    tp[t + 1] = tp[t] + nch[t];
    int64_t b = 0;
    for (int x : bcnt[t]) b += x;
    boff[t + 1] = boff[t] + b;
  }
  const int64_t nchunk = tp[nt], nbatch = boff[nt];
  LGH_CHECK (nbatch < INT_MAX, 8, "csr_tile_build: too many batches");
  // (first, end) batch of (chunk, wavefront); a wavefront's batches of
  // one tile are one run (wavefront-major, then chunk)
  std::vector<int> cb ((size_t) nchunk * nw * 2, 0);
  for (int t = 0; t < nt; t++) {
    int64_t b = boff[t];
    for (int w = 0; w < nw; w++)
      for (int c = 0; c < nch[t]; c++) {
        const size_t k = ((size_t) (tp[t] + c) * nw + w) * 2;
        cb[k] = (int) b;
        b += bcnt[t][(size_t) c * nw + w];
        cb[k + 1] = (int) b;
      }
  }
  if (cb.empty ()) cb.assign (2, 0);
  const int64_t tailb = rg == 16 ? ch / 4 : 0;     // MFMA staging reads ch/4 batches
  std::vector<int> tu ((size_t) std::max<int64_t> (nchunk * ch, 1));
  std::vector<int> es ((size_t) std::max<int64_t> ((nbatch + tailb) * nb, 1), 0);
  std::vector<double> ev ((size_t) std::max<int64_t> ((nbatch + tailb) * nb * rg, 1), 0.0);
  par ([&] (int t) {
    work (t, true, tu.data () + (size_t) tp[t] * ch, es.data (), ev.data (), boff[t]);
  });
  std::vector<int> trow;
  if (s.tx) {
    trow.resize ((size_t) nt * T);
    for (int t = 0; t < nt; t++)
      for (int e = 0; e < T; e++) trow[(size_t) t * T + e] = tile_row (s, n, ntx, nty, t, e);
  }
  const size_t vb = prec_bytes (A.prec);
  TileCopy c;
  c.s = s;
  c.ntiles = nt;
  c.ntx = ntx;
  c.nchunk = nchunk;
  c.nbatch = nbatch;
  c.union_per_row = n ? (double) nchunk * ch / n : 0.0;
  c.tp = (int *) dmalloc (sizeof (int) * tp.size ());
  c.tu = (int *) dmalloc (sizeof (int) * tu.size ());
  c.cb = (int *) dmalloc (sizeof (int) * cb.size ());
  c.es = (int *) dmalloc (sizeof (int) * es.size ());
  c.ev = dmalloc (vb * ev.size ());
  if (s.tx) c.trow = (int *) dmalloc (sizeof (int) * trow.size ());
  if (!c.tp || !c.tu || !c.cb || !c.es || !c.ev || (s.tx && !c.trow)) { c.release (); return 5; }
  int rc = h2d (c.tp, tp.data (), sizeof (int) * tp.size ());
  if (!rc) rc = h2d (c.tu, tu.data (), sizeof (int) * tu.size ());
  if (!rc) rc = h2d (c.cb, cb.data (), sizeof (int) * cb.size ());
  if (!rc) rc = h2d (c.es, es.data (), sizeof (int) * es.size ());
  if (!rc && s.tx) rc = h2d (c.trow, trow.data (), sizeof (int) * trow.size ());
  if (!rc) {
    if (A.prec == F64) rc = h2d (c.ev, ev.data (), sizeof (double) * ev.size ());
    else {
      std::vector<float> f (ev.begin (), ev.end ());
      rc = h2d (c.ev, f.data (), sizeof (float) * f.size ());
    }
  }
  if (rc) { c.release (); return rc; }
  out.release ();
  out = c;
  return 0;
}

int
csr_tile_stride (int n, const int *rp, const int *ci, int tx, int rows)
{
  if (n < 4 * rows || rows % tx) return 0;
  const int ty = rows / tx;
  const int i0 = n / 2;
  std::vector<int> cand;
  for (int k = rp[i0]; k < rp[i0 + 1]; k++) {
    const int64_t d = (int64_t) ci[k] - i0;
    if (d > tx && d * ty < n / 2) cand.push_back ((int) d);
  }
  std::sort (cand.begin (), cand.end ());
  cand.erase (std::unique (cand.begin (), cand.end ()), cand.end ());
  if (cand.size () > 512) cand.resize (512);
  // union per row of a few sample tiles (rows r0 + a + s * b)
  std::vector<int> u, rr ((size_t) rows);
  auto cost = [&] (int s) -> double {
    double tot = 0.0;
    int cnt = 0;
    for (int q = 1; q <= 4; q++) {
      const int64_t r0 = (int64_t) n * q / 5;
      bool ok = true;
      for (int b = 0; b < ty && ok; b++)
        for (int a = 0; a < tx; a++) {
          const int64_t r = s ? r0 + a + (int64_t) s * b : r0 + (int64_t) b * tx + a;
          if (r >= n) { ok = false; break; }
          rr[(size_t) b * tx + a] = (int) r;
        }
      if (!ok) continue;
      rows_union (rp, ci, rr.data (), rows, u);
      tot += (double) u.size () / rows;
      cnt++;
    }
    return cnt ? tot / cnt : 1e300;
  };
  const double c1 = cost (0);
  int best = 0;
  double bc = 0.9 * c1;
  for (int s : cand) {
    const double c = cost (s);
    if (c < bc) { bc = c; best = s; }
  }
  return best;
}

// 1 / diagonal of a square host CSR (duplicates summed)
static std::vector<double>
host_dinv (int n, const int *rp, const int *ci, const double *v)
{
  std::vector<double> d ((size_t) n, 0.0);
  for (int i = 0; i < n; i++)
    for (int k = rp[i]; k < rp[i + 1]; k++)
      if (ci[k] == i) d[i] += v[k];
  for (int i = 0; i < n; i++) d[i] = 1.0 / d[i];
  return d;
}

// The products are formed on the host in the device copy's precision from
// the values and dinv exactly as uploaded (IEEE products, round to nearest,
// as the device rounds them), so a kernel that reads vs[k] sees the bits it
// would compute as v[k] * (b1 * dinv[ci[k]]).
int
csr_scale_cols (DevCSR &A, const int *rp, const int *ci, const double *v,
                double b1)
{
  LGH_CHECK (A.dinv && A.n == A.m && (A.prec == F64 || A.prec == F32), 8,
             "csr_scale_cols: needs a square F64/F32 matrix with dinv");
  const std::vector<double> d = host_dinv (A.n, rp, ci, v);
  const int64_t nnz = A.nnz;
  void *p = dmalloc (prec_bytes (A.prec) * (size_t) (nnz ? nnz : 1));
  if (!p) return 5;
  int rc;
  if (A.prec == F64) {
    std::vector<double> s ((size_t) nnz);
    for (int64_t k = 0; k < nnz; k++) s[k] = v[k] * (b1 * d[ci[k]]);
    rc = h2d (p, s.data (), sizeof (double) * (size_t) nnz);
  }
  else {
    const float bf = (float) b1;
    std::vector<float> s ((size_t) nnz);
    for (int64_t k = 0; k < nnz; k++)
      s[k] = (float) v[k] * (bf * (float) d[ci[k]]);
    rc = h2d (p, s.data (), sizeof (float) * (size_t) nnz);
  }
  if (rc) { dfree (p); return rc; }
  dfree (A.vs);
  A.vs = p;
  A.vs_b1 = b1;
  return 0;
}

int
csr_upload (int n, int m, const int *rp, const int *ci, const double *v,
            int prec, int want_dinv, DevCSR **out)
{
  LGH_TRY (dev_init ());
  DevCSR *A = new DevCSR ();
  A->n = n; A->m = m; A->nnz = rp[n]; A->prec = prec;
  const size_t vb = prec_bytes (prec);
  A->rp = (int *) dmalloc (sizeof (int) * (size_t) (n + 1));
  A->ci = (int *) dmalloc (sizeof (int) * (size_t) (A->nnz ? A->nnz : 1));
  A->v = dmalloc (vb * (size_t) (A->nnz ? A->nnz : 1));
  if (!A->rp || !A->ci || !A->v) { delete A; return 5; }
  int rc = h2d (A->rp, rp, sizeof (int) * (size_t) (n + 1));
  if (!rc) rc = h2d (A->ci, ci, sizeof (int) * (size_t) A->nnz);
  std::vector<float> vf;
  if (!rc) {
    if (prec == F64) rc = h2d (A->v, v, sizeof (double) * (size_t) A->nnz);
    else {
      vf.resize ((size_t) A->nnz);
      for (int64_t k = 0; k < A->nnz; k++) vf[k] = (float) v[k];
      rc = h2d (A->v, vf.data (), sizeof (float) * (size_t) A->nnz);
    }
  }
  if (!rc && want_dinv && n == m) {
    const std::vector<double> d = host_dinv (n, rp, ci, v);
    A->dinv = dmalloc (vb * (size_t) n);
    if (!A->dinv) rc = 5;
    else if (prec == F64) rc = h2d (A->dinv, d.data (), sizeof (double) * n);
    else {
      std::vector<float> df (d.begin (), d.end ());
      rc = h2d (A->dinv, df.data (), sizeof (float) * n);
    }
  }
  if (rc) { delete A; return rc; }
  *out = A;
  return 0;
}

int
pmat_alloc (PMat &m, int nrows, int ncols, int W, int prec)
{
  m.nrows = nrows; m.ncols = ncols; m.W = W; m.prec = prec;
  m.p = dmalloc (m.bytes () ? m.bytes () : 1);
  if (!m.p) return 5;
  return dzero (m.p, m.bytes ());
}

void
pmat_free (PMat &m)
{
  dfree (m.p);
  m.p = nullptr;
}

// ---- rocSOLVER ------------------------------------------------------------
// Diagnostic trace (environment variable LGH_TRACE_SOLVER = a file name):
// each factorization appends "name n hash(input) hash(output) info", the
// hashes (FNV-1a) over the n x n block it reads and writes, so that runs
// of the same builds can be compared call by call.  Off unless set.
static std::mutex g_trace_mu;

static const char *
trace_path ()
{
  const char *p = getenv ("LGH_TRACE_SOLVER");
  return (p && *p) ? p : nullptr;
}

static uint64_t
trace_hash (const double *A, int n, int lda, const double *w)
{
  std::vector<double> h ((size_t) std::max (lda, 1) * std::max (n, 1));
  if (n > 0 && d2h (h.data (), A, sizeof (double) * (size_t) lda * n)) return 0;
  uint64_t x = 1469598103934665603ULL;
  auto mix = [&] (double v) {
    uint64_t b;
    memcpy (&b, &v, 8);
    for (int k = 0; k < 8; k++) { x ^= (b >> (8 * k)) & 0xff; x *= 1099511628211ULL; }
  };
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) mix (h[i + (size_t) j * lda]);
  if (w) {
    std::vector<double> hw ((size_t) n);
    if (n > 0 && d2h (hw.data (), w, sizeof (double) * n)) return 0;
    for (double v : hw) mix (v);
  }
  return x;
}

static void
trace_line (const char *path, const char *name, int n, uint64_t hin,
            uint64_t hout, int info)
{
  std::lock_guard<std::mutex> lk (g_trace_mu);
  if (FILE *f = fopen (path, "a")) {
    fprintf (f, "%s %d %016llx %016llx %d\n", name, n, (unsigned long long) hin,
             (unsigned long long) hout, info);
    fclose (f);
  }
}

// The Cholesky factor and its inverse use the engine's own blocked kernels
// (k_potrf, k_trtri): rocSOLVER's dpotrf returned different factors for
// bit-identical Grams (and now and then a spurious failure) while other
// streams kept the device busy, which made batched builds differ from
// one-at-a-time ones.
int
sol_potrf (int upper, int n, double *A, int lda, int *info_host)
{
  const char *tp = trace_path ();
  const uint64_t hin = tp ? trace_hash (A, n, lda, nullptr) : 0;
  LGH_TRY (dzero (t_ctx.info, sizeof (int)));
  LGH_TRY (k_potrf (upper, n, A, lda, t_ctx.info));
  LGH_TRY (d2h (info_host, t_ctx.info, sizeof (int)));
  if (tp) trace_line (tp, "potrf", n, hin, trace_hash (A, n, lda, nullptr), *info_host);
  return 0;
}

int
sol_trtri (int upper, int n, double *A, int lda, int *info_host)
{
  const char *tp = trace_path ();
  const uint64_t hin = tp ? trace_hash (A, n, lda, nullptr) : 0;
  LGH_TRY (dzero (t_ctx.info, sizeof (int)));
  LGH_TRY (k_trtri (upper, n, A, lda, t_ctx.info));
  LGH_TRY (d2h (info_host, t_ctx.info, sizeof (int)));
  if (tp) trace_line (tp, "trtri", n, hin, trace_hash (A, n, lda, nullptr), *info_host);
  return 0;
}

int
sol_syevd (int n, double *A, int lda, double *w_dev, int *info_host)
{
  double *E = (double *) dmalloc (sizeof (double) * (size_t) (n > 0 ? n : 1));
  if (!E) return 5;
  const char *tp = trace_path ();
  const uint64_t hin = tp ? trace_hash (A, n, lda, nullptr) : 0;
  rocblas_status s = rocsolver_dsyevd (t_ctx.rb, rocblas_evect_original,
                                       rocblas_fill_upper, n, A, lda, w_dev,
                                       E, t_ctx.info);
  int rc = 0;
  if (s != rocblas_status_success) { set_error ("rocsolver_dsyevd failed"); rc = 6; }
  if (!rc) rc = d2h (info_host, t_ctx.info, sizeof (int));
  if (!rc && tp) trace_line (tp, "syevd", n, hin, trace_hash (A, n, lda, w_dev), *info_host);
  dfree (E);
  return rc;
}

int
sol_syev_vals (int n, double *A, int lda, double *w_dev, int *info_host)
{
  double *E = (double *) dmalloc (sizeof (double) * (size_t) (n > 0 ? n : 1));
  if (!E) return 5;
  rocblas_status s = rocsolver_dsyev (t_ctx.rb, rocblas_evect_none,
                                      rocblas_fill_upper, n, A, lda, w_dev, E,
                                      t_ctx.info);
  int rc = 0;
  if (s != rocblas_status_success) { set_error ("rocsolver_dsyev failed"); rc = 6; }
  if (!rc) rc = d2h (info_host, t_ctx.info, sizeof (int));
  dfree (E);
  return rc;
}

int
blas_dgemm (int ta, int tb, int m, int n, int k, double alpha, const double *A,
            int lda, const double *B, int ldb, double beta, double *C, int ldc)
{
  if (m == 0 || n == 0) return 0;
  (void) stream ();
  rocblas_status s = rocblas_dgemm (
      t_ctx.rb, ta ? rocblas_operation_transpose : rocblas_operation_none,
      tb ? rocblas_operation_transpose : rocblas_operation_none, m, n, k,
      &alpha, A, lda, B, ldb, &beta, C, ldc);
  LGH_CHECK (s == rocblas_status_success, 6, "rocblas_dgemm failed");
  return 0;
}

size_t
dmem_total ()
{
  static std::atomic<size_t> total (0);
  size_t t = total.load ();
  if (!t) {
    size_t fr = 0;
    (void) stream ();
    if (hipMemGetInfo (&fr, &t) != hipSuccess) t = 0;
    total.store (t);
  }
  return t;
}

}  // namespace lgh
