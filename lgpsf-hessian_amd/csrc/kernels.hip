// kernels.hip — gfx950 kernels of the GLR engine and their launchers.
//
// Row kernels (SpMM, fused Chebyshev step, residual) for W < 64 map one
// wavefront to 64/W rows of one panel: lane (r = lane / W, c = lane % W)
// owns column c of row r.  At W = 64 (every large block) a lane moves 16 B
// of a row, so a CSR neighbour costs one coalesced load of that neighbour's
// row by 8, 16 or 32 lanes (bf16, fp32, fp64 storage).  The tall-skinny products
// use MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32) on 64 x 64
// output tiles with split-K partials reduced in a fixed order
// (deterministic results).
#include "lgh_dev.hpp"
#include "randn.hpp"

#include <algorithm>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <type_traits>
#include <vector>

namespace lgh {

typedef double d4 __attribute__ ((ext_vector_type (4)));
typedef float f4 __attribute__ ((ext_vector_type (4)));

static constexpr int TPB = 256;

template <typename T> struct vt;
template <> struct vt<double> { static constexpr int prec = F64; };
template <> struct vt<float> { static constexpr int prec = F32; };

static inline unsigned
grid1 (size_t work, int per_block)
{
  size_t g = (work + per_block - 1) / per_block;
  return (unsigned) (g ? g : 1);
}

#define LGH_LAUNCH_CHECK() LGH_HIP (hipGetLastError ())

static double
csr_bytes (const DevCSR &A)
{
  return (double) A.nnz * (prec_bytes (A.prec) + 4) + 4.0 * (A.n + 1);
}


// W dispatch helper: f(std::integral_constant<int, W>)
template <int N> using IC = std::integral_constant<int, N>;
template <typename F>
static int
with_W (int W, F &&f)
{
  switch (W) {
  case 1: return f (IC<1> ());
  case 2: return f (IC<2> ());
  case 4: return f (IC<4> ());
  case 8: return f (IC<8> ());
  case 16: return f (IC<16> ());
  case 32: return f (IC<32> ());
  case 64: return f (IC<64> ());
  default: set_error ("unsupported panel width"); return 7;
  }
}

// ---------------------------------------------------------------------------
// fill / zero

__global__ void __launch_bounds__ (TPB)
zero_k (uint32_t *p, size_t nwords)
{
  size_t i = (size_t) blockIdx.x * TPB + threadIdx.x;
  size_t stride = (size_t) gridDim.x * TPB;
  for (; i < nwords; i += stride) p[i] = 0u;
}

__global__ void __launch_bounds__ (TPB)
zero_bytes_k (unsigned char *p, size_t n)
{
  size_t i = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (i < n) p[i] = 0;
}

int
dzero (void *dst, size_t bytes)
{
  if (!bytes) return 0;
  size_t nw = bytes / 4;
  if (nw) {
    unsigned g = grid1 (nw, TPB);
    if (g > 65536u * 4) g = 65536u * 4;
    hipLaunchKernelGGL (zero_k, dim3 (g), dim3 (TPB), 0, stream (),
                        (uint32_t *) dst, nw);
    LGH_LAUNCH_CHECK ();
  }
  size_t rem = bytes - nw * 4;
  if (rem) {
    hipLaunchKernelGGL (zero_bytes_k, dim3 (1), dim3 (TPB), 0, stream (),
                        (unsigned char *) dst + nw * 4, rem);
    LGH_LAUNCH_CHECK ();
  }
  return 0;
}

// ---------------------------------------------------------------------------
// hashed Gaussian fill
//
// Every draw is computed on the device with the correctly rounded log.  A
// draw whose value may depend on how the platform log rounds (randn.hpp:
// about 2.2 % of FP64 draws, ~2^-29 of FP32 ones) is also appended to a
// list (element offset, q, u1); the host evaluates the platform log of
// those q and randn_fix_k rewrites the draws, so the panel equals the
// reference's sketch bit for bit.  The list is gathered per block in LDS
// and appended with one global atomic per block; its order is arbitrary
// but every entry rewrites its own element, so the result is not.

static constexpr int RANDN_DPT = 16;        // draws per thread
static constexpr int RANDN_LDS_CAP = 512;   // list entries a block gathers

// elements e0 .. e1 - 1 of a panel layout of width W with nrows rows a
// panel (W = 1: column-major with leading dimension nrows); rows past
// nvalid are not written, columns past ncols are written as 0
template <typename T, int W>
__global__ void __launch_bounds__ (TPB)
randn_k (T *out, int nrows, int nvalid, int ncols, uint64_t seed,
         int64_t row0, int64_t col0, size_t e0, size_t e1, uint64_t *fe,
         double *fq, double *fu, unsigned *cnt, unsigned cap)
{
  __shared__ uint64_t se[RANDN_LDS_CAP];
  __shared__ double sq[RANDN_LDS_CAP], su[RANDN_LDS_CAP];
  __shared__ unsigned sn, sbase;
  if (threadIdx.x == 0) sn = 0u;
  __syncthreads ();
  const size_t base = e0 + (size_t) blockIdx.x * TPB * RANDN_DPT;
  for (int k = 0; k < RANDN_DPT; k++) {
    const size_t e = base + (size_t) k * TPB + threadIdx.x;
    if (e >= e1) break;
    const int c = (int) (e % W);
    const size_t rest = e / W;
    const int i = (int) (rest % nrows);
    const int j = (int) (rest / nrows) * W + c;
    if (i >= nvalid) continue;
    if (j >= ncols) { out[e] = (T) 0; continue; }
    bool f;
    double u1, q;
    out[e] = randn_at_flag<T> (seed, row0 + i, col0 + j, &f, &u1, &q);
    if (f) {
      const unsigned sl = atomicAdd (&sn, 1u);
      if (sl < (unsigned) RANDN_LDS_CAP) { se[sl] = e; sq[sl] = q; su[sl] = u1; }
      else {
        const unsigned g = atomicAdd (cnt, 1u);
        if (g < cap) { fe[g] = e; fq[g] = q; fu[g] = u1; }
      }
    }
  }
  __syncthreads ();
  const unsigned m = sn < (unsigned) RANDN_LDS_CAP ? sn : (unsigned) RANDN_LDS_CAP;
  if (m == 0u) return;
  if (threadIdx.x == 0) sbase = atomicAdd (cnt, m);
  __syncthreads ();
  for (unsigned sl = threadIdx.x; sl < m; sl += TPB) {
    const unsigned g = sbase + sl;
    if (g < cap) { fe[g] = se[sl]; fq[g] = sq[sl]; fu[g] = su[sl]; }
  }
}

// out[fe[k]] = the draw of (fu[k], fq[k]) with log(fq[k]) = fl[k]
template <typename T>
__global__ void __launch_bounds__ (TPB)
randn_fix_k (T *out, const uint64_t *fe, const double *fq, const double *fu,
             const double *fl, unsigned n)
{
  const unsigned k = blockIdx.x * TPB + threadIdx.x;
  if (k < n) out[fe[k]] = (T) randn_from (fu[k], fq[k], fl[k]);
}

// the calling thread's pinned buffer for the listed q (grown on demand)
struct RandnHost { double *p = nullptr; size_t n = 0; };
static thread_local RandnHost t_rh;

void
randn_host_release ()
{
  host_pinned_free (t_rh.p);
  t_rh = RandnHost ();
}

// x[k] = log(x[k]) with the platform log, over up to 16 threads (a build's
// sketch lists ~10 M draws; each log is ~20 ns)
static void
platform_log_inplace (double *x, size_t n)
{
  const int nt = n < 65536 ? 1
      : (int) std::max (1u, std::min (16u, std::thread::hardware_concurrency ()));
  auto part = [&] (int t) {
    const size_t o = n * t / nt, e = n * (t + 1) / nt;
    for (size_t k = o; k < e; k++) x[k] = ::log (x[k]);
  };
  std::vector<std::thread> th;
  for (int t = 1; t < nt; t++) th.emplace_back (part, t);
  part (0);
  for (auto &t : th) t.join ();
}

int
k_randn (PMat &out, uint64_t seed, int64_t row0, int64_t col0)
{
  const size_t tot = out.elems ();
  if (!tot) return 0;
  LGH_CHECK (out.prec == F64 || out.prec == F32, 8, "k_randn: FP64 or FP32 output");
  // list capacity: 1/32 of the draws (FP64: 2.2 % expected), 1/4096 for
  // FP32; a fill that lists more runs again with the exact capacity
  size_t cap = (out.prec == F64 ? tot / 32 : tot / 4096) + 4096;
  const unsigned nblk = (unsigned) ((tot + (size_t) TPB * RANDN_DPT - 1)
                                    / ((size_t) TPB * RANDN_DPT));
  for (int pass = 0; pass < 2; pass++) {
    cap = std::min (cap, (size_t) UINT32_MAX);
    char *buf = (char *) dmalloc (cap * 32 + 256);
    if (!buf) return 55;
    uint64_t *fe = (uint64_t *) buf;
    double *fq = (double *) (buf + cap * 8), *fu = (double *) (buf + cap * 16),
           *fl = (double *) (buf + cap * 24);
    unsigned *cnt = (unsigned *) (buf + cap * 32);
    auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      if (out.prec == F64)
        hipLaunchKernelGGL ((randn_k<double, W>), dim3 (nblk), dim3 (TPB), 0,
                            stream (), (double *) out.p, out.nrows, out.nrows,
                            out.ncols, seed, row0, col0, (size_t) 0, tot, fe,
                            fq, fu, cnt, (unsigned) cap);
      else
        hipLaunchKernelGGL ((randn_k<float, W>), dim3 (nblk), dim3 (TPB), 0,
                            stream (), (float *) out.p, out.nrows, out.nrows,
                            out.ncols, seed, row0, col0, (size_t) 0, tot, fe,
                            fq, fu, cnt, (unsigned) cap);
      LGH_LAUNCH_CHECK ();
      return 0;
    };
    unsigned n = 0;
    int rc = dzero (cnt, 4);
    if (!rc) rc = with_W (out.W, l);
    if (!rc) rc = d2h (&n, cnt, 4);
    if (!rc && n > cap && pass == 0) { dfree (buf); cap = n; continue; }
    if (!rc && n > cap) { set_error ("k_randn: flag list overflow"); rc = 8; }
    if (!rc && n) {
      if (n > t_rh.n) {
        randn_host_release ();
        const size_t want = std::max<size_t> (n + n / 4, 65536);
        t_rh.p = (double *) host_pinned_alloc (want * 8);
        if (!t_rh.p) rc = 55;
        else t_rh.n = want;
      }
      if (!rc) rc = d2h (t_rh.p, fq, (size_t) n * 8);
      if (!rc) {
        platform_log_inplace (t_rh.p, n);
        rc = h2d (fl, t_rh.p, (size_t) n * 8);
      }
      if (!rc) {
        if (out.prec == F64)
          hipLaunchKernelGGL ((randn_fix_k<double>), dim3 (grid1 (n, TPB)),
                              dim3 (TPB), 0, stream (), (double *) out.p, fe,
                              fq, fu, fl, n);
        else
          hipLaunchKernelGGL ((randn_fix_k<float>), dim3 (grid1 (n, TPB)),
                              dim3 (TPB), 0, stream (), (float *) out.p, fe,
                              fq, fu, fl, n);
        if (hipGetLastError () != hipSuccess) { set_error ("randn_fix_k launch"); rc = 76; }
      }
    }
    // (a freed block is reused only by later work on this stream)
    dfree (buf);
    return rc;
  }
  return 0;
}
// ---- copy bandwidth (measurement) ---------------------------------------------
// y = x over n 16-byte words, grid-stride; the achievable-HBM reference the
// roofline fractions are also quoted against
__global__ void __launch_bounds__ (TPB)
copy16_k (const uint4 *__restrict__ x, uint4 *__restrict__ y, size_t n)
{
  const size_t stride = (size_t) gridDim.x * TPB;
  for (size_t i = (size_t) blockIdx.x * TPB + threadIdx.x; i < n; i += stride)
    y[i] = x[i];
}

int
copy_gbps (size_t bytes, int reps, double *gbps)
{
  bytes = bytes / 16 * 16;
  LGH_CHECK (bytes > 0 && reps > 0, 63, "copy_gbps: bad arguments");
  void *a = dmalloc (bytes), *b = dmalloc (bytes);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = (a && b) ? 0 : 5;
  if (!rc) rc = dzero (a, bytes);
  if (!rc && (hipEventCreate (&e0) != hipSuccess || hipEventCreate (&e1) != hipSuccess)) {
    set_error ("copy_gbps: event creation failed");
    rc = 2;
  }
  float best = 0.f;
  for (int r = 0; r <= reps && !rc; r++) {     // the first launch is a warm-up
    float t = 0.f;
    (void) hipEventRecord (e0, stream ());
    hipLaunchKernelGGL (copy16_k, dim3 (256 * 64), dim3 (TPB), 0, stream (),
                        (const uint4 *) a, (uint4 *) b, bytes / 16);
    if (hipGetLastError () != hipSuccess) { set_error ("copy16_k launch"); rc = 76; break; }
    (void) hipEventRecord (e1, stream ());
    if (hipEventSynchronize (e1) != hipSuccess
        || hipEventElapsedTime (&t, e0, e1) != hipSuccess) { set_error ("copy timing"); rc = 2; }
    if (!rc && r > 0 && (best == 0.f || t < best)) best = t;
  }
  if (e0) (void) hipEventDestroy (e0);
  if (e1) (void) hipEventDestroy (e1);
  dfree (a);
  dfree (b);
  if (!rc) *gbps = 2.0 * (double) bytes / (best * 1e-3) / 1e9;
  return rc;
}

// ---- the fill without allocation --------------------------------------------
// lgh_hip_randn_fill_noalloc: the same draws into caller memory, for callers
// that cannot allocate device memory.  The flag list of the draws finished
// with the platform log lives in `scratch`, a file-scope device array of
// NA_SCRATCH_BYTES (one per GPU, held until the process ends): NA_CAP entries
// of (offset, q, u1, log q), then the counter.  A call fills at most
// NA_MAX_DRAWS draws (about 2.2 % of FP64 draws are listed, ~2^-29 of FP32
// ones); a fill that lists more than NA_CAP runs again in slices of NA_CAP
// elements, each of which cannot.  The listed q pass through a static host
// array.  Every call on a GPU shares its array (and the host array):
// overlapping calls race.
static constexpr size_t NA_CAP = size_t (1) << 20;
static constexpr size_t NA_SCRATCH_BYTES = NA_CAP * 32 + 256;
static constexpr size_t NA_MAX_DRAWS = size_t (1) << 25;
static __device__ __align__ (256) unsigned char scratch[NA_SCRATCH_BYTES];
static constexpr int NA_MAXDEV = 64;
static void *g_scratch_dev[NA_MAXDEV];
static double g_na_host[NA_CAP];

int
scratch_init ()
{
  int dev = 0;
  LGH_HIP (hipGetDevice (&dev));
  LGH_CHECK (dev >= 0 && dev < NA_MAXDEV, 8, "scratch_init: device index out of range");
  if (g_scratch_dev[dev]) return 0;
  void *p = nullptr;
  LGH_HIP (hipGetSymbolAddress (&p, HIP_SYMBOL (scratch)));
  g_scratch_dev[dev] = p;
  return 0;
}

int
randn_fill_noalloc (uint64_t seed, int64_t row0, int64_t col0, int nr, int nc,
                    int prec, void *out, int ld)
{
  LGH_CHECK (dev_ready (), 73, "lgh_hip_randn_fill_noalloc: call lgh_hip_init on "
             "this thread first (this entry point creates nothing)");
  LGH_CHECK (nr >= 0 && nc >= 0 && ld >= std::max (nr, 1) && (prec == 0 || prec == 1)
             && (out || (size_t) nr * nc == 0), 63,
             "lgh_hip_randn_fill_noalloc: bad arguments");
  const size_t tot = (size_t) ld * nc;
  LGH_CHECK ((size_t) nr * nc <= NA_MAX_DRAWS && tot <= 2 * NA_MAX_DRAWS, 60,
             "lgh_hip_randn_fill_noalloc: more than 33554432 draws (or a leading "
             "dimension above twice that per call): the no-allocation limit");
  if ((size_t) nr * nc == 0) return 0;
  int dev = 0;
  LGH_HIP (hipGetDevice (&dev));
  LGH_CHECK (dev >= 0 && dev < NA_MAXDEV && g_scratch_dev[dev], 73,
             "lgh_hip_randn_fill_noalloc: no scratch on this device");
  char *sp = (char *) g_scratch_dev[dev];
  uint64_t *fe = (uint64_t *) sp;
  double *fq = (double *) (sp + NA_CAP * 8), *fu = (double *) (sp + NA_CAP * 16),
         *fl = (double *) (sp + NA_CAP * 24);
  unsigned *cnt = (unsigned *) (sp + NA_CAP * 32);
  const hipStream_t s = stream ();
  // elements [e0, e1) of the column-major block; false: the list overflowed
  auto pass = [&] (size_t e0, size_t e1, bool *fits) -> int {
    *fits = true;
    LGH_TRY (dzero (cnt, 4));
    const unsigned nblk = (unsigned) ((e1 - e0 + (size_t) TPB * RANDN_DPT - 1)
                                      / ((size_t) TPB * RANDN_DPT));
    if (prec == 0)
      hipLaunchKernelGGL ((randn_k<double, 1>), dim3 (nblk), dim3 (TPB), 0, s,
                          (double *) out, ld, nr, nc, seed, row0, col0, e0, e1,
                          fe, fq, fu, cnt, (unsigned) NA_CAP);
    else
      hipLaunchKernelGGL ((randn_k<float, 1>), dim3 (nblk), dim3 (TPB), 0, s,
                          (float *) out, ld, nr, nc, seed, row0, col0, e0, e1,
                          fe, fq, fu, cnt, (unsigned) NA_CAP);
    LGH_LAUNCH_CHECK ();
    unsigned n = 0;
    LGH_HIP (hipMemcpyAsync (&n, cnt, 4, hipMemcpyDeviceToHost, s));
    LGH_HIP (hipStreamSynchronize (s));
    if (n > NA_CAP) { *fits = false; return 0; }
    if (n == 0) return 0;
    LGH_HIP (hipMemcpyAsync (g_na_host, fq, (size_t) n * 8, hipMemcpyDeviceToHost, s));
    LGH_HIP (hipStreamSynchronize (s));
    for (unsigned k = 0; k < n; k++) g_na_host[k] = ::log (g_na_host[k]);
    LGH_HIP (hipMemcpyAsync (fl, g_na_host, (size_t) n * 8, hipMemcpyHostToDevice, s));
    if (prec == 0)
      hipLaunchKernelGGL ((randn_fix_k<double>), dim3 (grid1 (n, TPB)), dim3 (TPB), 0, s,
                          (double *) out, fe, fq, fu, fl, n);
    else
      hipLaunchKernelGGL ((randn_fix_k<float>), dim3 (grid1 (n, TPB)), dim3 (TPB), 0, s,
                          (float *) out, fe, fq, fu, fl, n);
    LGH_LAUNCH_CHECK ();
    // the host array and the list are reused by the next pass
    LGH_HIP (hipStreamSynchronize (s));
    return 0;
  };
  bool fits = true;
  LGH_TRY (pass (0, tot, &fits));
  for (size_t e0 = 0; !fits && e0 < tot; e0 += NA_CAP) {
    bool f2 = true;
    LGH_TRY (pass (e0, std::min (tot, e0 + NA_CAP), &f2));
    LGH_CHECK (f2, 8, "lgh_hip_randn_fill_noalloc: flag list overflow");
  }
  return 0;
}


// ---------------------------------------------------------------------------
// W = 64 row kernels.  A row of a 64-column panel is 128 B (bf16), 256 B
// (fp32) or 512 B (fp64); each lane moves 16 B of it, so a row takes 8, 16
// or 32 lanes and a wavefront works on 8, 4 or 2 rows at once.
// The CSR entries of a row are taken 8 at a time: their column indices
// and values first, then the 8 neighbour-row loads all issued before any
// is used.  Entries past the row's end get weight 0, the row's own index
// (so every address stays in bounds) and no load: the load is masked off,
// which matters for 5- and 7-point rows.  The
// panel index is blockIdx.y: all blocks of one panel run before the next
// panel's, which keeps the neighbour rows of a row in L2.

template <typename T> struct VT;
template <> struct VT<float> { static constexpr int N = 4; };
template <> struct VT<double> { static constexpr int N = 2; };
template <> struct VT<bf16_t> { static constexpr int N = 8; };

__device__ inline void vld (const float *p, float (&x)[4])
{ const float4 v = *(const float4 *) p; x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w; }
__device__ inline void vld (const double *p, double (&x)[2])
{ const double2 v = *(const double2 *) p; x[0] = v.x; x[1] = v.y; }
__device__ inline void vst (float *p, const float (&x)[4])
{ *(float4 *) p = make_float4 (x[0], x[1], x[2], x[3]); }
__device__ inline void vst (double *p, const double (&x)[2])
{ *(double2 *) p = make_double2 (x[0], x[1]); }
// two fp32 values widened to fp64 (mixed kernel: fp32 D, Z beside fp64 R, X)
__device__ inline void vld (const float *p, double (&x)[2])
{ const float2 v = *(const float2 *) p; x[0] = v.x; x[1] = v.y; }
__device__ inline void vst (float *p, const double (&x)[2])
{ *(float2 *) p = make_float2 ((float) x[0], (float) x[1]); }

// bf16 storage (V-cycle vectors of the BF16 preconditioner tier): values
// widen to fp32 exactly; narrowing rounds to nearest even (the gfx950
// conversion instruction)
__device__ inline float bf2f (uint32_t h) { return __uint_as_float (h << 16); }
__device__ inline uint32_t f2bf (float f)
{
  return __builtin_bit_cast (uint16_t, (__bf16) f);
}
template <typename T, int N>
__device__ inline void vld (const bf16_t *p, T (&x)[N])
{
  static_assert (N == 2 || N == 4 || N == 8, "bf16 vector width");
  uint32_t w[N / 2];
  if constexpr (N == 8) {
    const uint4 v = *(const uint4 *) p;
    w[0] = v.x; w[1] = v.y; w[2] = v.z; w[3] = v.w;
  }
  else if constexpr (N == 4) {
    const uint2 v = *(const uint2 *) p;
    w[0] = v.x; w[1] = v.y;
  }
  else w[0] = *(const uint32_t *) p;
#pragma unroll
  for (int i = 0; i < N / 2; i++) {
    x[2 * i] = (T) bf2f (w[i] & 0xffffu);
    x[2 * i + 1] = (T) bf2f (w[i] >> 16);
  }
}
template <typename T, int N>
__device__ inline void vst (bf16_t *p, const T (&x)[N])
{
  static_assert (N == 2 || N == 4 || N == 8, "bf16 vector width");
  uint32_t w[N / 2];
#pragma unroll
  for (int i = 0; i < N / 2; i++)
    w[i] = f2bf ((float) x[2 * i]) | (f2bf ((float) x[2 * i + 1]) << 16);
  if constexpr (N == 8) *(uint4 *) p = make_uint4 (w[0], w[1], w[2], w[3]);
  else if constexpr (N == 4) *(uint2 *) p = make_uint2 (w[0], w[1]);
  else *(uint32_t *) p = w[0];
}

// 16 B of a panel row as loaded (the neighbour loads of a group stay in
// this raw form, 4 VGPRs each, and are widened to T where they are used)
template <typename S> struct Raw16;
template <> struct Raw16<float> { using type = float4; };
template <> struct Raw16<double> { using type = double2; };
template <> struct Raw16<bf16_t> { using type = uint4; };
template <typename S>
__device__ inline typename Raw16<S>::type rld (const S *p)
{ return *(const typename Raw16<S>::type *) p; }
__device__ inline void unpack (const float4 &r, float (&x)[4])
{ x[0] = r.x; x[1] = r.y; x[2] = r.z; x[3] = r.w; }
__device__ inline void unpack (const double2 &r, double (&x)[2])
{ x[0] = r.x; x[1] = r.y; }
__device__ inline void unpack (const uint4 &r, float (&x)[8])
{
  const uint32_t w[4] = { r.x, r.y, r.z, r.w };
#pragma unroll
  for (int i = 0; i < 4; i++) {
    x[2 * i] = __uint_as_float (w[i] << 16);
    x[2 * i + 1] = __uint_as_float (w[i] & 0xffff0000u);
  }
}

// four values of a panel row widened to fp64 (the mixed step's row map:
// 16 lanes per row); raw forms keep gathered neighbours at 2-4 VGPRs
__device__ inline void vld (const double *p, double (&x)[4])
{
  const double2 a = *(const double2 *) p, b = *(const double2 *) (p + 2);
  x[0] = a.x; x[1] = a.y; x[2] = b.x; x[3] = b.y;
}
__device__ inline void vst (double *p, const double (&x)[4])
{
  *(double2 *) p = make_double2 (x[0], x[1]);
  *(double2 *) (p + 2) = make_double2 (x[2], x[3]);
}
__device__ inline void vld (const float *p, double (&x)[4])
{ const float4 v = *(const float4 *) p; x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w; }
__device__ inline void vst (float *p, const double (&x)[4])
{ *(float4 *) p = make_float4 ((float) x[0], (float) x[1], (float) x[2], (float) x[3]); }
// streaming forms of the 4-value loads and stores (nontemporal: data that
// no other row of the launch reads, and that the next reader finds evicted
// anyway at fine-level sizes, does not displace the gathered neighbours)
typedef double ntd2 __attribute__ ((ext_vector_type (2)));
typedef float ntf4 __attribute__ ((ext_vector_type (4)));
typedef uint32_t ntu2 __attribute__ ((ext_vector_type (2)));
__device__ inline void ntld (const double *p, double (&x)[4])
{
  const ntd2 a = __builtin_nontemporal_load ((const ntd2 *) p);
  const ntd2 b = __builtin_nontemporal_load ((const ntd2 *) (p + 2));
  x[0] = a.x; x[1] = a.y; x[2] = b.x; x[3] = b.y;
}
__device__ inline void ntld (const float *p, double (&x)[4])
{
  const ntf4 v = __builtin_nontemporal_load ((const ntf4 *) p);
  x[0] = v.x; x[1] = v.y; x[2] = v.z; x[3] = v.w;
}
__device__ inline void ntst (double *p, const double (&x)[4])
{
  __builtin_nontemporal_store ((ntd2) { x[0], x[1] }, (ntd2 *) p);
  __builtin_nontemporal_store ((ntd2) { x[2], x[3] }, (ntd2 *) (p + 2));
}
__device__ inline void ntst (float *p, const double (&x)[4])
{
  __builtin_nontemporal_store (
      (ntf4) { (float) x[0], (float) x[1], (float) x[2], (float) x[3] },
      (ntf4 *) p);
}
__device__ inline void ntst (bf16_t *p, const double (&x)[4])
{
  const uint32_t w0 = f2bf ((float) x[0]) | (f2bf ((float) x[1]) << 16);
  const uint32_t w1 = f2bf ((float) x[2]) | (f2bf ((float) x[3]) << 16);
  __builtin_nontemporal_store ((ntu2) { w0, w1 }, (ntu2 *) p);
}

template <typename S> struct Raw4;
template <> struct Raw4<float> { using type = float4; };
template <> struct Raw4<bf16_t> { using type = uint2; };
__device__ inline float4 rld4 (const float *p) { return *(const float4 *) p; }
__device__ inline uint2 rld4 (const bf16_t *p) { return *(const uint2 *) p; }
__device__ inline void unpack4 (const float4 &r, double (&x)[4])
{ x[0] = r.x; x[1] = r.y; x[2] = r.z; x[3] = r.w; }
__device__ inline void unpack4 (const uint2 &r, double (&x)[4])
{
  x[0] = __uint_as_float (r.x << 16); x[1] = __uint_as_float (r.x & 0xffff0000u);
  x[2] = __uint_as_float (r.y << 16); x[3] = __uint_as_float (r.y & 0xffff0000u);
}

static constexpr int NG = 8;            // CSR entries per load group

template <typename T, int N>
__device__ inline void vzero (T (&x)[N])
{
#pragma unroll
  for (int i = 0; i < N; i++) x[i] = (T) 0;
}

// row map of a panel of width W (64, or 16 for a narrow last panel): W / N
// lanes per row, 16 B per lane
#define LGH_ROWMAP(T_)                                                        \
  constexpr int N = VT<T_>::N;                                                \
  constexpr int LPR = W / N;                                                  \
  const int row = blockIdx.x * (TPB / LPR) + (int) threadIdx.x / LPR;         \
  const size_t off = (size_t) blockIdx.y * n * W                              \
                     + (size_t) ((int) threadIdx.x % LPR) * N;                \
  if (row >= n) return;                                                       \
  const int k0 = rp[row], k1 = rp[row + 1];

// grid of a row kernel (panel width W) whose row map follows storage
// precision prec
static inline dim3
grid64 (int n, int np, int prec, int W = 64)
{
  const int rows = TPB / (W / (16 / (int) prec_bytes (prec)));
  return dim3 ((unsigned) ((n + rows - 1) / rows), (unsigned) np);
}
// the row kernels' widths: a full panel, or the narrow last panels of a
// block whose column count leaves at most 48 (16-wide) or 8 (8-wide) in
// it.  Every row map fits 8 columns: 16-byte lanes hold 8 BF16, 4 FP32 or
// 2 FP64 values.
static inline bool row_width (int W) { return W == 64 || W == 16 || W == 8; }
template <typename F>
static int
with_row_W (int W, F &&f)
{
  if (W == 64) return f (IC<64> ());
  if (W == 16) return f (IC<16> ());
  return f (IC<8> ());
}

// Y = alpha A X + beta Y (T arithmetic, S storage of X and Y)
template <typename TV, typename T, typename S, int W = 64>
__global__ void __launch_bounds__ (TPB)
spmm64_k (int n, int nx, const int *__restrict__ rp,
          const int *__restrict__ ci, const TV *__restrict__ v, T alpha,
          const S *__restrict__ X, T beta, S *__restrict__ Y)
{
  constexpr int N = VT<S>::N;
  constexpr int LPR = W / N;
  const int row = blockIdx.x * (TPB / LPR) + (int) threadIdx.x / LPR;
  const int lc = ((int) threadIdx.x % LPR) * N;
  if (row >= n) return;
  const S *Xp = X + (size_t) blockIdx.y * nx * W + lc;
  const int k0 = rp[row], k1 = rp[row + 1];
  T acc[N];
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = 0;
  for (int kb = k0; kb < k1; kb += NG) {
    int c[NG];
    T w[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      const bool in = kb + u < k1;
      c[u] = in ? ci[kb + u] : ci[k0];
      w[u] = in ? (T) v[kb + u] : (T) 0;
    }
    typename Raw16<S>::type x[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      if (kb + u < k1) x[u] = rld (Xp + (size_t) c[u] * W);
      else x[u] = {};
    }
#pragma unroll
    for (int u = 0; u < NG; u++) {
      T t[N];
      unpack (x[u], t);
#pragma unroll
      for (int i = 0; i < N; i++) acc[i] += w[u] * t[i];
    }
  }
  S *y = Y + ((size_t) blockIdx.y * n + row) * W + lc;
  if (beta == (T) 0) {
#pragma unroll
    for (int i = 0; i < N; i++) acc[i] *= alpha;
  }
  else {
    T yo[N];
    vld (y, yo);
#pragma unroll
    for (int i = 0; i < N; i++) acc[i] = alpha * acc[i] + beta * yo[i];
  }
  vst (y, acc);
}

// Y = alpha A X + beta Y over the row-grouped copy of A (DevCSR::grp),
// 64-column panels: one wavefront per group of RG rows and PPW panels.
// A lane moves 16 B of a panel row (VPL = 2 FP64 or 4 FP32 columns), so
// one gather instruction brings one neighbour row of VPL panels (lanes
// 0-31 panel q, 32-63 panel q + 1 for FP64; four quarters for FP32).  The
// wave walks the union of the group's column indices in batches of
// NB = 32 / RG entries, so a neighbour row of X that several rows of the
// group share is loaded once and used for each of them (a 127-entry PSF
// row shares ~90 % of its neighbours with the next row), and each entry's
// index and values serve PPW panels.  The group index is wave-uniform: a
// batch's NB indices and 32 values arrive by a few wide scalar loads.
// Row r accumulates its own entries in column order, as spmm64_k does,
// plus exact zeros (padding, and the other rows' entries): the result is
// the same bits.  Panels p0 + PPW * blockIdx.y + k; lanes whose panel is
// at or past pend read the last valid panel (the same addresses as their
// neighbours') and store nothing.
template <typename TV, typename T, typename S, int RG, int PPW>
__global__ void __launch_bounds__ (TPB)
spmm_grp_k (int n, int ngroups, int nx, int p0, int pend,
            const int *__restrict__ gp, const int *__restrict__ gci,
            const TV *__restrict__ gv, T alpha, const S *__restrict__ X,
            T beta, S *__restrict__ Y)
{
  constexpr int NB = 32 / RG;
  constexpr int VPL = 16 / (int) sizeof (S);      // columns per lane
  constexpr int LPR = 64 / VPL;                    // lanes per panel row
  constexpr int NL = PPW / VPL;                    // gathers per entry
  static_assert (PPW % VPL == 0, "panels per wave: a multiple of 16 B / element");
  typedef S V __attribute__ ((ext_vector_type (VPL)));
  const int lane = (int) threadIdx.x & 63;
  const int grp = __builtin_amdgcn_readfirstlane (
      (int) (blockIdx.x * (TPB / 64) + (threadIdx.x >> 6)));
  if (grp >= ngroups) return;
  const int pb = p0 + PPW * (int) blockIdx.y;
  const int c0 = (lane % LPR) * VPL;
  // panel of gather k of this lane, and its element offset from panel pb
  int pq[NL];
  size_t off[NL];
#pragma unroll
  for (int k = 0; k < NL; k++) {
    const int want = pb + k * VPL + lane / LPR;
    pq[k] = want < pend ? want : pend - 1;
    off[k] = (size_t) (pq[k] - pb) * nx * 64 + c0;
  }
  const S *Xb = X + (size_t) pb * nx * 64;
  T acc[RG][NL][VPL];
#pragma unroll
  for (int r = 0; r < RG; r++)
#pragma unroll
    for (int k = 0; k < NL; k++)
#pragma unroll
      for (int v = 0; v < VPL; v++) acc[r][k][v] = 0;
  // two-stage pipeline: the gathers of batch b + 1 are issued before the
  // products of batch b (scheduling barriers keep that order), so a wave
  // has up to two batches of neighbour rows in flight
  auto gather = [&] (int b, V (&x)[NB][NL]) {
    const int *cb = gci + (size_t) b * NB;
#pragma unroll
    for (int u = 0; u < NB; u++) {
      const S *row = Xb + (size_t) cb[u] * 64;
#pragma unroll
      for (int k = 0; k < NL; k++) x[u][k] = *(const V *) (row + off[k]);
    }
  };
  auto product = [&] (int b, const V (&x)[NB][NL]) {
    const TV *wb = gv + (size_t) b * NB * RG;
#pragma unroll
    for (int u = 0; u < NB; u++)
#pragma unroll
      for (int r = 0; r < RG; r++) {
        const T w = (T) wb[u * RG + r];
#pragma unroll
        for (int k = 0; k < NL; k++)
#pragma unroll
          for (int v = 0; v < VPL; v++) acc[r][k][v] += w * (T) x[u][k][v];
      }
  };
  const int b0 = gp[grp], b1 = gp[grp + 1];
  V xa[NB][NL], xb[NB][NL];
  int b = b0;
  if (b < b1) gather (b, xa);
  while (b < b1) {
    if (b + 1 < b1) gather (b + 1, xb);
    __builtin_amdgcn_sched_barrier (0);
    product (b, xa);
    __builtin_amdgcn_sched_barrier (0);
    if (++b >= b1) break;
    if (b + 1 < b1) gather (b + 1, xa);
    __builtin_amdgcn_sched_barrier (0);
    product (b, xb);
    __builtin_amdgcn_sched_barrier (0);
    ++b;
  }
#pragma unroll
  for (int k = 0; k < NL; k++) {
    if (pb + k * VPL + lane / LPR >= pend) continue;
#pragma unroll
    for (int r = 0; r < RG; r++) {
      const int row = grp * RG + r;
      if (row >= n) break;
      V *y = (V *) (Y + ((size_t) pq[k] * n + row) * 64 + c0);
      V o;
      if (beta == (T) 0) {
#pragma unroll
        for (int v = 0; v < VPL; v++) o[v] = (S) (alpha * acc[r][k][v]);
      }
      else {
        const V yo = *y;
#pragma unroll
        for (int v = 0; v < VPL; v++) o[v] = (S) (alpha * acc[r][k][v] + beta * (T) yo[v]);
      }
      *y = o;
    }
  }
}

// panels [p0, p0 + np) of Y = alpha A X + beta Y through the grouped copy:
// whole groups of ppw panels, then the remaining panels in one launch of
// the smallest group width that holds them
template <typename TV, typename T, typename S>
static int
spmm_grp_launch (const DevCSR &A, int ppw, int p0, int np, T alpha,
                 const S *X, T beta, S *Y)
{
  constexpr int VPL = 16 / (int) sizeof (S);
  const GroupCopy &g = A.grp;
  const int ng = (A.n + g.rg - 1) / g.rg;
  const unsigned gx = (unsigned) ((ng + TPB / 64 - 1) / (TPB / 64));
  const int pend = p0 + np;
  auto go = [&] (auto rgc, auto pc, int first, int groups) -> int {
    constexpr int RG = decltype (rgc)::value, PPW = decltype (pc)::value;
    if (groups <= 0) return 0;
    hipLaunchKernelGGL ((spmm_grp_k<TV, T, S, RG, PPW>),
                        dim3 (gx, (unsigned) groups), dim3 (TPB), 0,
                        stream (), A.n, ng, A.m, first, pend, g.gp, g.gci,
                        (const TV *) g.gv, alpha, X, beta, Y);
    LGH_LAUNCH_CHECK ();
    return 0;
  };
  auto by_w = [&] (auto rgc, int w, int first, int groups) -> int {
    if (w == VPL) return go (rgc, IC<VPL> (), first, groups);
    if (w == 2 * VPL) return go (rgc, IC<2 * VPL> (), first, groups);
    set_error ("spmm_grp_launch: panels per wave: 1 or 2 gathers of 16 B");
    return 8;
  };
  auto by_rg = [&] (auto rgc) -> int {
    const int whole = np / ppw;
    LGH_TRY (by_w (rgc, ppw, p0, whole));
    const int rem = np - whole * ppw;
    if (rem == 0) return 0;
    return by_w (rgc, rem <= VPL ? VPL : 2 * VPL, p0 + whole * ppw, 1);
  };
  if (g.rg == 4) return by_rg (IC<4> ());
  if (g.rg == 8) return by_rg (IC<8> ());
  set_error ("spmm_grp_launch: rows per group 4 or 8");
  return 8;
}

// the tile of this workgroup (-1: none) in launch order xcd: 0 the tiles
// in order; 1 dealt to the XCDs (dispatch order b mod 8) in contiguous
// runs; 2 (2D tiles, ntx a band) XCD x walks the x-th vertical strip of
// the tile grid band by band, so the tiles one XCD runs at a time are
// neighbours and share the rows they stage through its L2
__device__ inline int
tile_of_block (int ntiles, int xcd, int ntx)
{
  const int b = (int) blockIdx.x;
  if (xcd == 1) {
    const int per = (ntiles + 7) / 8, t = (b % 8) * per + b / 8;
    return t < ntiles ? t : -1;
  }
  if (xcd == 2) {
    const int nty = ntiles / ntx, sw = (ntx + 7) / 8, l = b / 8;
    const int tcol = (b % 8) * sw + l % sw, tband = l / sw;
    return (tcol < ntx && tband < nty) ? tband * ntx + tcol : -1;
  }
  return b < ntiles ? b : -1;
}

// Y = alpha A X + beta Y over the row-tile copy of A (DevCSR::tile), one
// workgroup of NW wavefronts per tile of NW * RG rows and one 16-byte piece
// of a panel row per lane (2 FP64 / 4 FP32 panels a launch row, so a staged
// row is 1 KB).  The tile's union of neighbour rows passes through LDS in
// chunks of CH rows, double-buffered; one barrier a chunk.  Chunk c + 1 is
// written to the free buffer after the products of chunk c, from registers
// loaded two chunks earlier (its loads have two chunks of products to
// land); the same loads bring the wavefront's values of that chunk into L2
// (one lane a 128-byte line), so that its scalar loads later hit there.  A
// neighbour row that many rows of the tile share crosses L1 once per tile,
// not once per row group.  Each wavefront walks its own rows' union, a
// batch's LDS slots and 16 values by wide scalar loads.  Row r of a
// wavefront accumulates its own entries in column order (chunks in order,
// slots in order), plus exact zeros: the bits of spmm64_k.  xcd = 1: tiles
// dealt to the XCDs (dispatch order b mod 8) in contiguous runs (2D strips
// per XCD are laid out by the tile builder itself).
template <typename TV, typename T, typename S, int RG, int NW, int CH>
__global__ void __launch_bounds__ (NW * 64)
spmm_tile_k (int n, int nx, int ntiles, int xcd, int ntx, int p0, int pend,
             const int *__restrict__ tp, const int *__restrict__ tu,
             const int *__restrict__ cb, const int *__restrict__ es,
             const TV *__restrict__ ev, const int *__restrict__ trow, T alpha,
             const S *__restrict__ X, T beta, S *__restrict__ Y,
             int *__restrict__ touch_sum)
{
  constexpr int NB = 16 / RG;                      // entries a batch
  constexpr int NV = NB * RG;                      // values a batch
  constexpr int VPL = 16 / (int) sizeof (S);       // columns per lane = panels a launch row
  constexpr int LPR = 64 / VPL;                    // lanes per panel row
  constexpr int ROWB = 1024;                       // bytes of a staged row
  constexpr int SPW = CH / NW;                     // staged rows per wavefront and chunk
  static_assert (CH % NW == 0, "chunk rows: a multiple of the wavefronts");
  typedef S V __attribute__ ((ext_vector_type (VPL)));
  __shared__ __attribute__ ((aligned (16))) unsigned char lds[2][CH * ROWB];
  const int t = tile_of_block (ntiles, xcd, ntx);
  if (t < 0) return;                               // whole workgroup
  const int wave = __builtin_amdgcn_readfirstlane ((int) (threadIdx.x >> 6));
  const int lane = (int) threadIdx.x & 63;
  const int pb = p0 + VPL * (int) blockIdx.y;
  const int want = pb + lane / LPR;
  const int pq = want < pend ? want : pend - 1;
  const int c0 = (lane % LPR) * VPL;
  const S *Xl = X + (size_t) pq * nx * 64 + c0;
  T acc[RG][VPL];
#pragma unroll
  for (int r = 0; r < RG; r++)
#pragma unroll
    for (int v = 0; v < VPL; v++) acc[r][v] = 0;
  const int cs = tp[t], ce = tp[t + 1];
  int sink = 0;
  if (cs < ce) {
    V s0[SPW], s1[SPW];
    auto fetch = [&] (int c, V (&st)[SPW]) {
      const int *u = tu + (size_t) c * CH + wave;
#pragma unroll
      for (int j = 0; j < SPW; j++) st[j] = *(const V *) (Xl + (size_t) u[j * NW] * 64);
      const int q0 = cb[((size_t) c * NW + wave) * 2], q1 = cb[((size_t) c * NW + wave) * 2 + 1];
      const size_t b0 = (size_t) q0 * NV * sizeof (TV), b1 = (size_t) q1 * NV * sizeof (TV);
      const size_t at = (b0 & ~(size_t) 127) + (size_t) lane * 128;
      if (at < b1) sink += *(const int *) ((const char *) ev + at);
    };
    auto put = [&] (const V (&st)[SPW], int buf) {
#pragma unroll
      for (int j = 0; j < SPW; j++)
        *(V *) (&lds[buf][(j * NW + wave) * ROWB + lane * 16]) = st[j];
    };
    auto products = [&] (int c, int buf) {
      const int q0 = cb[((size_t) c * NW + wave) * 2], q1 = cb[((size_t) c * NW + wave) * 2 + 1];
      const unsigned char *base = &lds[buf][lane * 16];
      for (int q = q0; q < q1; q++) {
        const int *sl = es + (size_t) q * NB;
        const TV *wv = ev + (size_t) q * NV;
        V x[NB];
#pragma unroll
        for (int e = 0; e < NB; e++) x[e] = *(const V *) (base + sl[e] * ROWB);
#pragma unroll
        for (int e = 0; e < NB; e++)
#pragma unroll
          for (int r = 0; r < RG; r++) {
            const T w = (T) wv[e * RG + r];
#pragma unroll
            for (int v = 0; v < VPL; v++) acc[r][v] += w * (T) x[e][v];
          }
      }
    };
    // one chunk: barrier (chunk c staged, the other buffer free), its
    // products, stage chunk c + 1 from st, refill st with chunk c + 3
    // (unconditionally: past the tile's end its last chunk again, so the
    // loads in flight are the same on every path and the wait before a
    // staging write covers only the set it writes)
    auto step = [&] (int c, int buf, V (&st)[SPW]) {
      __syncthreads ();
      products (c, buf);
      __builtin_amdgcn_sched_barrier (0);
      if (c + 1 < ce) put (st, buf ^ 1);
      fetch (min (c + 3, ce - 1), st);
      __builtin_amdgcn_sched_barrier (0);
    };
    fetch (cs, s0);
    put (s0, 0);
    fetch (min (cs + 1, ce - 1), s1);
    fetch (min (cs + 2, ce - 1), s0);
    for (int c = cs; c < ce; c += 2) {
      step (c, 0, s1);
      if (c + 1 >= ce) break;
      step (c + 1, 1, s0);
    }
  }
  // touch_sum (null from the launcher) takes the L2 touches' sum, so the
  // touch loads are not dropped as unused
  if (touch_sum) atomicAdd (touch_sum, sink);
  if (want >= pend) return;
#pragma unroll
  for (int r = 0; r < RG; r++) {
    const int e = wave * RG + r;
    const int row = trow ? trow[(size_t) t * (NW * RG) + e]
                         : ((int64_t) t * (NW * RG) + e < n ? t * (NW * RG) + e : -1);
    if (row < 0) continue;
    V *y = (V *) (Y + ((size_t) pq * n + row) * 64 + c0);
    V o;
    if (beta == (T) 0) {
#pragma unroll
      for (int v = 0; v < VPL; v++) o[v] = (S) (alpha * acc[r][v]);
    }
    else {
      const V yo = *y;
#pragma unroll
      for (int v = 0; v < VPL; v++) o[v] = (S) (alpha * acc[r][v] + beta * (T) yo[v]);
    }
    *y = o;
  }
}

// Y = alpha A X + beta Y (FP64 vectors) over a row-tile copy with 16 rows
// a wavefront (rg = 16: 4 x 4 blocks of a 2D tile, or 16 consecutive rows)
// on the matrix cores.  A wavefront's union of its rows' entries is walked
// in k-steps of 4 entries: the k-step's 16 x 4 block of values (A, lane
// i + 16 k holds row i, entry k) and the 4 x 16 column blocks of the 4
// staged neighbour rows (B, lane j + 16 k holds entry k, column j) feed
// v_mfma_f64_16x16x4f64, which on gfx950 rounds exactly as the FMA chain
// over k = 0..3 started from C (probed: every output of 2 000 random
// products); k-steps in column order therefore give each row the sums of
// spmm64_k in its order, plus exact zeros: the same bits.  The 128 columns
// of a launch row (2 panels) are 8 MFMA tiles: 16-column segments split
// into their even and odd columns, so one 16-byte LDS read gives a lane
// its column pair (2j, 2j + 1) of a segment of 32.  Everything the k-steps
// read is staged in LDS per chunk by loads of a fixed count (the chunk's
// X rows, and each wavefront's at most CH/4 A blocks and slots, which the
// copy pads past its end), two chunks ahead, double-buffered, one barrier
// a chunk.
template <typename TV, int NW, int CH, int PP>
__global__ void __launch_bounds__ (NW * 64)
spmm_mt_k (int n, int nx, int ntiles, int xcd, int ntx, int p0, int pend,
           const int *__restrict__ tp, const int *__restrict__ tu,
           const int *__restrict__ cb, const int *__restrict__ es,
           const TV *__restrict__ ev, const int *__restrict__ trow,
           double alpha, const double *__restrict__ X, double beta,
           double *__restrict__ Y)
{
  constexpr int KM = CH / 4;                       // k-steps of a wavefront in a chunk, at most
  constexpr int ROWB = PP * 512;                   // staged row: PP panels x 64 FP64
  constexpr int SPW = CH / NW;                     // staged X rows per wavefront and chunk
  constexpr int AL = KM * 64 * (int) sizeof (TV) / 16 / 64;   // 16-byte A loads per lane
  static_assert (CH % NW == 0 && (KM * 64 * (int) sizeof (TV)) % 1024 == 0, "chunk shape");
  static_assert (PP == 1 || PP == 2, "one or two panels a launch row");
  typedef double V __attribute__ ((ext_vector_type (2)));
  // a lane's piece of a staged row: 16 bytes of panel pb + lane / 32 (two
  // panels), or 8 bytes of panel pb (one)
  typedef typename std::conditional<PP == 2, V, double>::type P8;
  typedef uint4 A16;
  __shared__ __attribute__ ((aligned (16))) unsigned char ldx[2][CH * ROWB];
  __shared__ __attribute__ ((aligned (16))) TV lda[2][NW][KM * 64];
  __shared__ int lds_s[2][NW][KM * 4];
  const int t = tile_of_block (ntiles, xcd, ntx);
  if (t < 0) return;                               // whole workgroup
  const int wave = __builtin_amdgcn_readfirstlane ((int) (threadIdx.x >> 6));
  const int lane = (int) threadIdx.x & 63;
  const int pb = p0 + PP * (int) blockIdx.y;
  const int sq = PP == 2 ? min (pb + lane / 32, pend - 1) : pb;
  const double *Xl = X + (size_t) sq * nx * 64 + (PP == 2 ? (lane % 32) * 2 : lane);
  typedef double d4v __attribute__ ((ext_vector_type (4)));
  d4v acc[PP][2][2];
#pragma unroll
  for (int p = 0; p < PP; p++)
#pragma unroll
    for (int g = 0; g < 2; g++)
#pragma unroll
      for (int h = 0; h < 2; h++) acc[p][g][h] = (d4v) { 0.0, 0.0, 0.0, 0.0 };
  const int cs = tp[t], ce = tp[t + 1];
  if (cs < ce) {
    struct Stage { P8 x[SPW]; A16 a[AL]; int s; };
    Stage s0, s1;
    auto fetch = [&] (int c, Stage &st) {
      const int *u = tu + (size_t) c * CH + wave;
#pragma unroll
      for (int j = 0; j < SPW; j++) st.x[j] = *(const P8 *) (Xl + (size_t) u[j * NW] * 64);
      const int q0 = cb[((size_t) c * NW + wave) * 2];
      const A16 *av = (const A16 *) (ev + (size_t) q0 * 64);
#pragma unroll
      for (int m = 0; m < AL; m++) st.a[m] = av[m * 64 + lane];
      st.s = es[(size_t) q0 * 4 + (lane < KM * 4 ? lane : 0)];
    };
    auto put = [&] (const Stage &st, int buf) {
#pragma unroll
      for (int j = 0; j < SPW; j++)
        *(P8 *) (&ldx[buf][(j * NW + wave) * ROWB + lane * (int) sizeof (P8)]) = st.x[j];
#pragma unroll
      for (int m = 0; m < AL; m++) ((A16 *) lda[buf][wave])[m * 64 + lane] = st.a[m];
      if (lane < KM * 4) lds_s[buf][wave][lane] = st.s;
    };
    auto products = [&] (int c, int buf) {
      const int nk = cb[((size_t) c * NW + wave) * 2 + 1] - cb[((size_t) c * NW + wave) * 2];
      const int j = lane % 16, k = lane / 16;
      for (int ks = 0; ks < nk; ks++) {
        const double a = (double) lda[buf][wave][ks * 64 + lane];
        const int sl = lds_s[buf][wave][ks * 4 + k];
        const unsigned char *row = &ldx[buf][sl * ROWB + j * 16];
        V b[PP][2];
#pragma unroll
        for (int p = 0; p < PP; p++)
#pragma unroll
          for (int g = 0; g < 2; g++) b[p][g] = *(const V *) (row + p * 512 + g * 256);
#pragma unroll
        for (int p = 0; p < PP; p++)
#pragma unroll
          for (int g = 0; g < 2; g++) {
            acc[p][g][0] = __builtin_amdgcn_mfma_f64_16x16x4f64 (a, b[p][g].x, acc[p][g][0], 0, 0, 0);
            acc[p][g][1] = __builtin_amdgcn_mfma_f64_16x16x4f64 (a, b[p][g].y, acc[p][g][1], 0, 0, 0);
          }
      }
    };
    auto step = [&] (int c, int buf, Stage &st) {
      __syncthreads ();
      products (c, buf);
      __builtin_amdgcn_sched_barrier (0);
      if (c + 1 < ce) put (st, buf ^ 1);
      fetch (min (c + 3, ce - 1), st);
      __builtin_amdgcn_sched_barrier (0);
    };
    fetch (cs, s0);
    put (s0, 0);
    fetch (min (cs + 1, ce - 1), s1);
    fetch (min (cs + 2, ce - 1), s0);
    for (int c = cs; c < ce; c += 2) {
      step (c, 0, s1);
      if (c + 1 >= ce) break;
      step (c + 1, 1, s0);
    }
  }
  // D(i, j) of a tile: i = lane / 16 + 4 r, j = lane % 16; tile (p, g, h)
  // holds column 32 g + 2 j + h of panel pb + p
  const int j = lane % 16;
#pragma unroll
  for (int r = 0; r < 4; r++) {
    const int i = lane / 16 + 4 * r;
    const int e = wave * 16 + i;
    const int row = trow ? trow[(size_t) t * (NW * 16) + e]
                         : ((int64_t) t * (NW * 16) + e < n ? t * (NW * 16) + e : -1);
    if (row < 0) continue;
#pragma unroll
    for (int p = 0; p < PP; p++) {
      if (pb + p >= pend) continue;
#pragma unroll
      for (int g = 0; g < 2; g++) {
        V *y = (V *) (Y + ((size_t) (pb + p) * n + row) * 64 + 32 * g + 2 * j);
        V o;
        if (beta == 0.0) {
          o.x = alpha * acc[p][g][0][r];
          o.y = alpha * acc[p][g][1][r];
        }
        else {
          const V yo = *y;
          o.x = alpha * acc[p][g][0][r] + beta * yo.x;
          o.y = alpha * acc[p][g][1][r] + beta * yo.y;
        }
        *y = o;
      }
    }
  }
}

// the tile shapes the autotuner times (rows per wavefront, wavefronts,
// chunk rows); 2D variants of each are added when the matrix has a line
// stride (csr_tile_stride)
struct TileCand { int rg, nw, ch; };
static const TileCand k_tile_cands[] = {
  { 4, 16, 32 }, { 4, 16, 64 }, { 8, 8, 32 }, { 8, 8, 64 },
};
// the matrix-core kernel's shapes (16 rows a wavefront)
static const TileCand k_mt_cands[] = {
  { 16, 4, 16 }, { 16, 8, 16 }, { 16, 4, 32 }, { 16, 8, 32 },
};

// panels [p0, p0 + np) of Y = alpha A X + beta Y through the tile copy
template <typename TV, typename T, typename S>
static int
spmm_tile_launch (const DevCSR &A, int p0, int np, T alpha, const S *X,
                  T beta, S *Y)
{
  constexpr int VPL = 16 / (int) sizeof (S);
  const TileCopy &c = A.tile;
  if (np <= 0) return 0;
  const unsigned gx = (unsigned) (c.s.xcd == 1 ? (c.ntiles + 7) / 8 * 8
                                  : c.s.xcd == 2 ? 8 * ((c.ntx + 7) / 8) * (c.ntiles / c.ntx)
                                  : c.ntiles);
  const unsigned gy = (unsigned) ((np + VPL - 1) / VPL);
  auto go = [&] (auto rgc, auto nwc, auto chc) -> int {
    constexpr int RG = decltype (rgc)::value, NW = decltype (nwc)::value,
                  CH = decltype (chc)::value;
    hipLaunchKernelGGL ((spmm_tile_k<TV, T, S, RG, NW, CH>), dim3 (gx, gy),
                        dim3 (NW * 64), 0, stream (), A.n, A.m, c.ntiles,
                        c.s.xcd, c.ntx, p0, p0 + np, c.tp, c.tu, c.cb, c.es,
                        (const TV *) c.ev, c.trow, alpha, X, beta, Y, (int *) nullptr);
    LGH_LAUNCH_CHECK ();
    return 0;
  };
  const TileShape &s = c.s;
  if (s.rg == 16) {
    if constexpr (std::is_same<T, double>::value && std::is_same<S, double>::value) {
      // whole pairs of panels, then an odd last panel on its own
      auto mt = [&] (auto nwc, auto chc) -> int {
        constexpr int NW = decltype (nwc)::value, CH = decltype (chc)::value;
        const int pairs = np / 2;
        if (pairs > 0) {
          hipLaunchKernelGGL ((spmm_mt_k<TV, NW, CH, 2>), dim3 (gx, (unsigned) pairs),
                              dim3 (NW * 64), 0, stream (), A.n, A.m, c.ntiles, c.s.xcd, c.ntx,
                              p0, p0 + 2 * pairs, c.tp, c.tu, c.cb, c.es, (const TV *) c.ev,
                              c.trow, (double) alpha, X, (double) beta, Y);
          LGH_LAUNCH_CHECK ();
        }
        if (np % 2) {
          hipLaunchKernelGGL ((spmm_mt_k<TV, NW, CH, 1>), dim3 (gx, 1u),
                              dim3 (NW * 64), 0, stream (), A.n, A.m, c.ntiles, c.s.xcd, c.ntx,
                              p0 + np - 1, p0 + np, c.tp, c.tu, c.cb, c.es, (const TV *) c.ev,
                              c.trow, (double) alpha, X, (double) beta, Y);
          LGH_LAUNCH_CHECK ();
        }
        return 0;
      };
      if (s.nw == 4 && s.ch == 16) return mt (IC<4> (), IC<16> ());
      if (s.nw == 8 && s.ch == 16) return mt (IC<8> (), IC<16> ());
      if (s.nw == 4 && s.ch == 32) return mt (IC<4> (), IC<32> ());
      if (s.nw == 8 && s.ch == 32) return mt (IC<8> (), IC<32> ());
    }
    set_error ("spmm_tile_launch: matrix-core tile shape not compiled (FP64 vectors only)");
    return 8;
  }
  if (s.rg == 4 && s.nw == 16 && s.ch == 32) return go (IC<4> (), IC<16> (), IC<32> ());
  if (s.rg == 4 && s.nw == 16 && s.ch == 64) return go (IC<4> (), IC<16> (), IC<64> ());
  if (s.rg == 8 && s.nw == 8 && s.ch == 32) return go (IC<8> (), IC<8> (), IC<32> ());
  if (s.rg == 8 && s.nw == 8 && s.ch == 64) return go (IC<8> (), IC<8> (), IC<64> ());
  set_error ("spmm_tile_launch: tile shape not compiled");
  return 8;
}

// ---- Cholesky factor and triangular inverse (FP64, deterministic) ---------
// The CholQR2 steps' potrf / trtri (glr_common_impl.h:387, 402;
// glr_scalapack_impl.h:128-143) in 64 x 64 blocks: each block is factored,
// solved or updated by one workgroup in LDS with every sum in a fixed order
// and no atomics, so a factor depends only on its input (a library potrf
// gave different factors for identical Grams while other streams kept the
// device busy).  The upper variants work on the transpose (the lower
// triangle of A^T is the upper triangle of A); the other triangle is not
// written.  info: LAPACK's convention (first non-positive pivot / zero
// diagonal, 1-based), set once.

struct TriView {
  double *A;
  int lda, up;
  __device__ double &at (int i, int j) const
  { return up ? A[j + (size_t) i * lda] : A[i + (size_t) j * lda]; }
};

static constexpr int CB = 64;        // block size
static constexpr int CBP = CB + 1;   // LDS row pitch

// *info = min(*info, v) over the positive values (0: no failure yet)
__device__ static void
info_min (int *info, int v)
{
  int old = *info;
  while (old == 0 || old > v) {
    const int prev = atomicCAS (info, old, v);
    if (prev == old) break;
    old = prev;
  }
}

// factor the diagonal block k0 .. k0 + b - 1 in place (right-looking)
__global__ void __launch_bounds__ (TPB)
chol_diag_k (TriView V, int n, int k0, int *info)
{
  __shared__ double s[CB][CBP];
  const int b = min (CB, n - k0), t = (int) threadIdx.x;
  for (int e = t; e < CB * CB; e += TPB) {
    const int i = e % CB, j = e / CB;
    if (i < b && j <= i) s[i][j] = V.at (k0 + i, k0 + j);
  }
  __syncthreads ();
  for (int j = 0; j < b; j++) {
    if (t == 0) {
      const double d = s[j][j];
      if (!(d > 0.0) && *info == 0) *info = k0 + j + 1;
      s[j][j] = sqrt (d);
    }
    __syncthreads ();
    const double djj = s[j][j];
    for (int i = j + 1 + t; i < b; i += TPB) s[i][j] = s[i][j] / djj;
    __syncthreads ();
    const int m = b - j - 1;           // trailing (i, c), j < c <= i < b
    for (int e = t; e < m * m; e += TPB) {
      const int i = j + 1 + e % m, c = j + 1 + e / m;
      if (c <= i) s[i][c] -= s[i][j] * s[c][j];
    }
    __syncthreads ();
  }
  for (int e = t; e < CB * CB; e += TPB) {
    const int i = e % CB, j = e / CB;
    if (i < b && j <= i) V.at (k0 + i, k0 + j) = s[i][j];
  }
}

// block rows below the diagonal block k0: X = A(r, k) L(k, k)^-T
__global__ void __launch_bounds__ (TPB)
chol_panel_k (TriView V, int n, int k0)
{
  __shared__ double L[CB][CBP], X[CB][CBP];
  const int b = min (CB, n - k0), r0 = k0 + CB * (1 + (int) blockIdx.x);
  const int br = min (CB, n - r0), t = (int) threadIdx.x;
  for (int e = t; e < CB * CB; e += TPB) {
    const int i = e % CB, j = e / CB;
    if (i < b && j <= i) L[i][j] = V.at (k0 + i, k0 + j);
    if (i < br && j < b) X[i][j] = V.at (r0 + i, k0 + j);
  }
  __syncthreads ();
  if (t < br)
    for (int c = 0; c < b; c++) {
      double x = X[t][c];
      for (int j = 0; j < c; j++) x -= X[t][j] * L[c][j];
      X[t][c] = x / L[c][c];
    }
  __syncthreads ();
  for (int e = t; e < CB * CB; e += TPB) {
    const int i = e % CB, j = e / CB;
    if (i < br && j < b) V.at (r0 + i, k0 + j) = X[i][j];
  }
}

// C(bi, bj) -= P(bi) P(bj)^T over the trailing lower blocks, block pair
// number blockIdx.x in (bi, bj), bj <= bi order; P = the panel of column
// block k0 (b columns).  A thread owns a 4 x 4 piece of the 64 x 64 block.
__global__ void __launch_bounds__ (TPB)
chol_update_k (TriView V, int n, int k0)
{
  __shared__ double P[CB][CBP], Q[CB][CBP];
  int bi = 0, rem = (int) blockIdx.x;
  while (rem > bi) { rem -= bi + 1; bi++; }
  const int bj = rem;
  const int b = min (CB, n - k0);
  const int r0 = k0 + CB * (1 + bi), c0 = k0 + CB * (1 + bj);
  const int br = min (CB, n - r0), bc = min (CB, n - c0), t = (int) threadIdx.x;
  for (int e = t; e < CB * CB; e += TPB) {
    const int i = e % CB, j = e / CB;
    P[i][j] = (i < br && j < b) ? V.at (r0 + i, k0 + j) : 0.0;
    Q[i][j] = (i < bc && j < b) ? V.at (c0 + i, k0 + j) : 0.0;
  }
  __syncthreads ();
  const int ti = (t / 16) * 4, tj = (t % 16) * 4;
  double acc[4][4] = {};
  for (int kk = 0; kk < b; kk++)
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) acc[u][v] += P[ti + u][kk] * Q[tj + v][kk];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) {
      const int i = ti + u, j = tj + v;
      if (i < br && j < bc && (bi != bj || j <= i)) V.at (r0 + i, c0 + j) -= acc[u][v];
    }
}

// W(blk, blk) = L(blk, blk)^-1 for every diagonal block (one workgroup
// each; a thread per column)
__global__ void __launch_bounds__ (TPB)
trinv_diag_k (TriView L, double *W, int ldw, int n, int *info)
{
  __shared__ double s[CB][CBP], w[CB][CBP];
  const int k0 = CB * (int) blockIdx.x, b = min (CB, n - k0), t = (int) threadIdx.x;
  for (int e = t; e < CB * CB; e += TPB) {
    const int i = e % CB, j = e / CB;
    s[i][j] = (i < b && j <= i) ? L.at (k0 + i, k0 + j) : 0.0;
  }
  __syncthreads ();
  if (t < b) {
    const int j = t;
    if (s[j][j] == 0.0) info_min (info, k0 + j + 1);
    w[j][j] = 1.0 / s[j][j];
    for (int i = j + 1; i < b; i++) {
      double x = 0.0;
      for (int k = j; k < i; k++) x += s[i][k] * w[k][j];
      w[i][j] = -x / s[i][i];
    }
  }
  __syncthreads ();
  for (int e = t; e < CB * CB; e += TPB) {
    const int i = e % CB, j = e / CB;
    if (i < b && j <= i) W[(k0 + i) + (size_t) (k0 + j) * ldw] = w[i][j];
  }
}

// block row bi of the inverse: W(bi, bk) = -W(bi, bi) sum_{bj = bk}^{bi-1}
// L(bi, bj) W(bj, bk), block column bk = blockIdx.x < bi
__global__ void __launch_bounds__ (TPB)
trinv_row_k (TriView L, double *W, int ldw, int n, int bi)
{
  __shared__ double P[CB][CBP], Q[CB][CBP];
  const int bk = (int) blockIdx.x, t = (int) threadIdx.x;
  const int r0 = CB * bi, c0 = CB * bk, br = min (CB, n - r0);
  const int ti = (t / 16) * 4, tj = (t % 16) * 4;
  double acc[4][4] = {};
  for (int bj = bk; bj < bi; bj++) {
    const int m0 = CB * bj;
    __syncthreads ();
    for (int e = t; e < CB * CB; e += TPB) {
      const int i = e % CB, j = e / CB;
      P[i][j] = (i < br) ? L.at (r0 + i, m0 + j) : 0.0;     // L(bi, bj)
      Q[i][j] = W[(m0 + i) + (size_t) (c0 + j) * ldw];        // W(bj, bk)
    }
    __syncthreads ();
    for (int kk = 0; kk < CB; kk++)
#pragma unroll
      for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) acc[u][v] += P[ti + u][kk] * Q[kk][tj + v];
  }
  __syncthreads ();
  for (int e = t; e < CB * CB; e += TPB) {
    const int i = e % CB, j = e / CB;
    P[i][j] = (i < br && j <= i) ? W[(r0 + i) + (size_t) (r0 + j) * ldw] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++) Q[ti + u][tj + v] = acc[u][v];
  __syncthreads ();
  double o[4][4] = {};
  for (int kk = 0; kk < CB; kk++)
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
      for (int v = 0; v < 4; v++) o[u][v] += P[ti + u][kk] * Q[kk][tj + v];
#pragma unroll
  for (int u = 0; u < 4; u++)
#pragma unroll
    for (int v = 0; v < 4; v++)
      if (ti + u < br) W[(r0 + ti + u) + (size_t) (c0 + tj + v) * ldw] = -o[u][v];
}

// A's triangle = W's lower triangle (transposed for the upper variant)
__global__ void __launch_bounds__ (TPB)
trinv_store_k (TriView L, const double *W, int ldw, int n)
{
  const size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= (size_t) n * n) return;
  const int i = (int) (e % n), j = (int) (e / n);
  if (j <= i) L.at (i, j) = W[i + (size_t) j * ldw];
}

int
k_potrf (int upper, int n, double *A, int lda, int *info_dev)
{
  if (n <= 0) return 0;
  const TriView V { A, lda, upper };
  const int nb = (n + CB - 1) / CB;
  for (int kb = 0; kb < nb; kb++) {
    const int k0 = CB * kb, m = nb - kb - 1;
    hipLaunchKernelGGL (chol_diag_k, dim3 (1), dim3 (TPB), 0, stream (), V, n, k0, info_dev);
    LGH_LAUNCH_CHECK ();
    if (m == 0) break;
    hipLaunchKernelGGL (chol_panel_k, dim3 ((unsigned) m), dim3 (TPB), 0, stream (), V, n, k0);
    LGH_LAUNCH_CHECK ();
    hipLaunchKernelGGL (chol_update_k, dim3 ((unsigned) (m * (m + 1) / 2)), dim3 (TPB), 0,
                        stream (), V, n, k0);
    LGH_LAUNCH_CHECK ();
  }
  return 0;
}

int
k_trtri (int upper, int n, double *A, int lda, int *info_dev)
{
  if (n <= 0) return 0;
  const TriView V { A, lda, upper };
  const int nb = (n + CB - 1) / CB, ldw = nb * CB;
  double *W = (double *) dmalloc (sizeof (double) * (size_t) ldw * ldw);
  if (!W) return 5;
  int rc = dzero (W, sizeof (double) * (size_t) ldw * ldw);
  if (!rc) {
    hipLaunchKernelGGL (trinv_diag_k, dim3 ((unsigned) nb), dim3 (TPB), 0, stream (),
                        V, W, ldw, n, info_dev);
    if (hipGetLastError () != hipSuccess) { set_error ("trinv_diag_k launch"); rc = 76; }
  }
  for (int bi = 1; bi < nb && !rc; bi++) {
    hipLaunchKernelGGL (trinv_row_k, dim3 ((unsigned) bi), dim3 (TPB), 0, stream (),
                        V, W, ldw, n, bi);
    if (hipGetLastError () != hipSuccess) { set_error ("trinv_row_k launch"); rc = 76; }
  }
  if (!rc) {
    hipLaunchKernelGGL (trinv_store_k, dim3 (grid1 ((size_t) n * n, TPB)), dim3 (TPB), 0,
                        stream (), V, W, ldw, n);
    if (hipGetLastError () != hipSuccess) { set_error ("trinv_store_k launch"); rc = 76; }
  }
  dfree (W);
  return rc;
}

// ---- autotuner of the grouped SpMM ----------------------------------------
// At upload of a wide-row matrix, each candidate (4 or 8 rows per group;
// 2 or 4 panels per wavefront with FP64 vectors, 4 or 8 with FP32) is timed on the
// matrix itself over 4 zero panels with HIP events (best of 3 launches
// after one warm-up); the rows per group that run the FP64 product fastest
// (both products for FP32 values) are kept with the best panels per wave
// of each vector precision.  Decisions are cached per (rows, nonzeros,
// value precision), so later uploads of a matrix of that shape build one
// copy and time nothing.

namespace {
struct GrpTune {
  int rg = 0, ppw[2] = { 0, 0 }; double ms[2] = { 0, 0 };
  int tile = 0; TileShape ts;          // tile: the row-tile copy won
};
std::mutex g_tune_mu;
std::map<std::tuple<int, int64_t, int>, GrpTune> g_tune;
int g_force_rg = 0, g_force_ppw = 0;
int g_force_tile = -1;
}

void
spmm_group_force (int rg, int ppw)
{
  std::lock_guard<std::mutex> lk (g_tune_mu);
  g_force_rg = rg;
  g_force_ppw = ppw;
  if (rg) g_force_tile = -1;
}

void
spmm_tile_force (int cand)
{
  std::lock_guard<std::mutex> lk (g_tune_mu);
  g_force_tile = cand;
  if (cand >= 0) g_force_rg = 0;
}

bool
spmm_group_forced ()
{
  std::lock_guard<std::mutex> lk (g_tune_mu);
  return g_force_rg != 0 || g_force_tile >= 0;
}

std::string
spmm_group_report ()
{
  std::lock_guard<std::mutex> lk (g_tune_mu);
  std::string r;
  char buf[320];
  for (const auto &kv : g_tune) {
    const GrpTune &t = kv.second;
    snprintf (buf, sizeof buf, "%d %lld %d %d %d %d %.4f %.4f %d %d %d %d %d %d %d\n",
              std::get<0> (kv.first), (long long) std::get<1> (kv.first),
              std::get<2> (kv.first), t.rg, t.ppw[0], t.ppw[1], t.ms[0], t.ms[1],
              t.tile, t.ts.rg, t.ts.nw, t.ts.ch, t.ts.tx, t.ts.stride, t.ts.xcd);
    r += buf;
  }
  return r;
}

// milliseconds per panel of a product (after one warm-up, best of 3)
template <typename L>
static int
time_panels (int np, L &&launch, double *ms)
{
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LGH_HIP (hipEventCreate (&e0));
  if (hipEventCreate (&e1) != hipSuccess) { (void) hipEventDestroy (e0); return 2; }
  float best = 0.f;
  int rc = launch ();
  for (int rep = 0; rep < 3 && !rc; rep++) {
    float t = 0.f;
    if (hipEventRecord (e0, stream ()) != hipSuccess) rc = 2;
    if (!rc) rc = launch ();
    if (!rc && (hipEventRecord (e1, stream ()) != hipSuccess
                || hipEventSynchronize (e1) != hipSuccess
                || hipEventElapsedTime (&t, e0, e1) != hipSuccess)) {
      set_error ("SpMM timing failed");
      rc = 2;
    }
    if (!rc && (rep == 0 || t < best)) best = t;
  }
  (void) hipEventDestroy (e0);
  (void) hipEventDestroy (e1);
  *ms = (double) best / np;
  return rc;
}

template <typename TV, typename T>
static int
grp_time (const DevCSR &A, int ppw, const PMat &X, PMat &Y, double *ms)
{
  const int np = X.npanels ();
  return time_panels (np, [&] {
    return spmm_grp_launch<TV, T, T> (A, ppw, 0, np, (T) 1, (const T *) X.p,
                                      (T) 0, (T *) Y.p);
  }, ms);
}

template <typename TV, typename T>
static int
tile_time (const DevCSR &A, const PMat &X, PMat &Y, double *ms)
{
  const int np = X.npanels ();
  return time_panels (np, [&] {
    return spmm_tile_launch<TV, T, T> (A, 0, np, (T) 1, (const T *) X.p,
                                       (T) 0, (T *) Y.p);
  }, ms);
}

template <typename TV>
static int
grp_tune_one (const DevCSR &A, PMat (&X)[2], PMat (&Y)[2], GrpTune &t)
{
  static const int cand[2][3] = { { 2, 4, 0 }, { 4, 8, 0 } };
  for (int vp = 0; vp < 2; vp++) {
    t.ms[vp] = 0.0;
    for (int k = 0; k < 3 && cand[vp][k]; k++) {
      double ms = 0.0;
      const int rc = vp == 0 ? grp_time<TV, double> (A, cand[vp][k], X[0], Y[0], &ms)
                             : grp_time<TV, float> (A, cand[vp][k], X[1], Y[1], &ms);
      if (rc) return rc;
      if (t.ms[vp] == 0.0 || ms < t.ms[vp]) { t.ms[vp] = ms; t.ppw[vp] = cand[vp][k]; }
    }
  }
  return 0;
}

template <typename TV>
static int
tile_tune_one (const DevCSR &A, PMat (&X)[2], PMat (&Y)[2], double (&ms)[2])
{
  int rc = tile_time<TV, double> (A, X[0], Y[0], &ms[0]);
  if (!rc) rc = tile_time<TV, float> (A, X[1], Y[1], &ms[1]);
  return rc;
}

// the tile shape of candidate k: bits 0-1 the shape (k_tile_cands, or with
// bit 4 the matrix-core k_mt_cands), bit 2 2D tiles 16 rows wide on the
// matrix's line stride (1D when it has none; matrix-core wavefronts then
// take 4 x 4 blocks), bit 3 XCD order (runs of 1D tiles, strips of 2D)
static TileShape
tile_shape_of (int k, int n, const int *rp, const int *ci)
{
  const TileCand &c = (k & 16) ? k_mt_cands[k & 3] : k_tile_cands[k & 3];
  TileShape s;
  s.rg = c.rg; s.nw = c.nw; s.ch = c.ch;
  s.xcd = (k & 8) ? 1 : 0;
  if (k & 4) {
    const int rows = c.rg * c.nw, tx = 16;
    s.stride = csr_tile_stride (n, rp, ci, tx, rows);
    if (s.stride) {
      s.tx = tx;
      if (c.rg == 16) s.wy = 4;
    }
    if (s.tx && (k & 8)) s.xcd = 2;     // 2D: strips of the tile grid per XCD
  }
  return s;
}

int
csr_group_rows (DevCSR &A, const int *rp, const int *ci, const double *v)
{
  for (int i = 0; i < A.n; i++)
    for (int k = rp[i] + 1; k < rp[i + 1]; k++)
      if (ci[k] <= ci[k - 1]) return 0;           // unsorted: no copy
  const auto key = std::make_tuple (A.n, A.nnz, A.prec);
  int frg, fppw, ftile;
  GrpTune known;
  {
    std::lock_guard<std::mutex> lk (g_tune_mu);
    frg = g_force_rg;
    fppw = g_force_ppw;
    ftile = g_force_tile;
    auto it = g_tune.find (key);
    if (it != g_tune.end ()) known = it->second;
  }
  if (ftile >= 0) {
    A.grp.release ();
    return csr_tile_build (A, rp, ci, v, tile_shape_of (ftile, A.n, rp, ci), A.tile);
  }
  if (frg) {
    A.tile.release ();
    LGH_TRY (csr_group_build (A, rp, ci, v, frg, A.grp));
    A.grp_ppw[0] = fppw == 2 ? 2 : 4;
    A.grp_ppw[1] = 2 * A.grp_ppw[0];
    return 0;
  }
  if (known.tile) return csr_tile_build (A, rp, ci, v, known.ts, A.tile);
  if (known.rg) {
    LGH_TRY (csr_group_build (A, rp, ci, v, known.rg, A.grp));
    A.grp_ppw[0] = known.ppw[0];
    A.grp_ppw[1] = known.ppw[1];
    return 0;
  }
  PMat X[2], Y[2];
  int rc = 0;
  for (int vp = 0; vp < 2 && !rc; vp++) {
    rc = pmat_alloc (X[vp], A.m, 256, 64, vp == 0 ? F64 : F32);
    if (!rc) rc = pmat_alloc (Y[vp], A.n, 256, 64, vp == 0 ? F64 : F32);
  }
  GrpTune best;
  double best_cost = 0.0;
  auto cost_of = [&] (const double (&ms)[2]) {
    return A.prec == F64 ? ms[0] : ms[0] + ms[1];
  };
  for (int rg : { 4, 8 }) {
    if (rc) break;
    DevCSR C;                                    // candidate: the copy only
    C.n = A.n; C.m = A.m; C.nnz = A.nnz; C.prec = A.prec;
    rc = csr_group_build (A, rp, ci, v, rg, C.grp);
    GrpTune t;
    t.rg = rg;
    if (!rc) rc = A.prec == F64 ? grp_tune_one<double> (C, X, Y, t)
                                : grp_tune_one<float> (C, X, Y, t);
    const double cost = cost_of (t.ms);
    if (!rc && (best.rg == 0 || cost < best_cost)) {
      best = t;
      best_cost = cost;
      A.grp.release ();
      A.grp = C.grp;
      C.grp = GroupCopy ();                      // now owned by A
    }
  }
  // the row-tile copies: every shape as 2D tiles when the matrix has a
  // line stride (1D tiles otherwise; on a grid they never won), the
  // matrix-core shapes (FP64 vectors only) for FP64 values; each copy is
  // timed in plain and in XCD order (one copy serves both)
  const bool two_d = csr_tile_stride (A.n, rp, ci, 16, 64) > 0;
  for (int k = 0; k < 32 && !rc; k++) {
    if (k & 8) continue;
    if ((k & 16) && A.prec != F64) continue;
    if (((k & 4) != 0) != two_d) continue;
    const TileShape s = tile_shape_of (k, A.n, rp, ci);
    if ((k & 4) && !s.tx) continue;
    DevCSR C;
    C.n = A.n; C.m = A.m; C.nnz = A.nnz; C.prec = A.prec;
    rc = csr_tile_build (A, rp, ci, v, s, C.tile);
    int won = -1;
    for (int xo = 0; xo < 2 && !rc; xo++) {
      C.tile.s.xcd = xo ? (s.tx ? 2 : 1) : 0;
      double ms[2] = { 0, 0 };
      if (k & 16) rc = tile_time<double, double> (C, X[0], Y[0], &ms[0]);
      else rc = A.prec == F64 ? tile_tune_one<double> (C, X, Y, ms)
                              : tile_tune_one<float> (C, X, Y, ms);
      if (!rc && cost_of (ms) < best_cost) {
        best.tile = 1;
        best.ts = C.tile.s;
        best.ms[0] = ms[0];
        best.ms[1] = ms[1];
        best_cost = cost_of (ms);
        won = C.tile.s.xcd;
      }
    }
    if (!rc && won >= 0) {
      C.tile.s.xcd = won;
      A.tile.release ();
      A.tile = C.tile;
      C.tile = TileCopy ();                      // now owned by A
    }
  }
  for (int vp = 0; vp < 2; vp++) { pmat_free (X[vp]); pmat_free (Y[vp]); }
  if (rc) { A.grp.release (); A.tile.release (); return rc; }
  if (best.tile) A.grp.release ();
  else {
    A.grp_ppw[0] = best.ppw[0];
    A.grp_ppw[1] = best.ppw[1];
  }
  std::lock_guard<std::mutex> lk (g_tune_mu);
  g_tune[key] = best;
  return 0;
}

// Y = alpha A X + beta Y for matrices with long rows (an aggregation
// hierarchy's restriction to a small coarse level: hundreds of entries a
// row over few rows, which one lane group per row walks serially).  One
// block per (row, panel): the row's entries are cut into LCH chunks of
// equal length, the block's lane groups sum one chunk each (in order) into
// shared memory, and the chunk sums are added in chunk order.  The
// chunking depends only on the row, so the result does not depend on the
// panel width or the storage's lane map.
static constexpr int LCH = 16;
template <typename TV, typename T, typename S, int W>
__global__ void __launch_bounds__ (TPB)
spmm_long_k (int n, int nx, const int *__restrict__ rp,
             const int *__restrict__ ci, const TV *__restrict__ v, T alpha,
             const S *__restrict__ X, T beta, S *__restrict__ Y)
{
  constexpr int N = VT<S>::N;
  constexpr int LPC = W / N;            // lanes per chunk (one panel row)
  constexpr int G = TPB / LPC;          // lane groups per block
  __shared__ T part[LCH][W];
  const int row = blockIdx.x;
  const int g = (int) threadIdx.x / LPC;
  const int lc = ((int) threadIdx.x % LPC) * N;
  const S *Xp = X + (size_t) blockIdx.y * nx * W + lc;
  const int k0 = rp[row], k1 = rp[row + 1];
  const int L = (k1 - k0 + LCH - 1) / LCH;
  for (int s = g; s < LCH; s += G) {
    const int e0 = min (k1, k0 + s * L), e1 = min (k1, e0 + L);
    T acc[N];
#pragma unroll
    for (int i = 0; i < N; i++) acc[i] = 0;
    for (int kb = e0; kb < e1; kb += NG) {
      int c[NG];
      T w[NG];
#pragma unroll
      for (int u = 0; u < NG; u++) {
        const bool in = kb + u < e1;
        c[u] = in ? ci[kb + u] : ci[e0];
        w[u] = in ? (T) v[kb + u] : (T) 0;
      }
      typename Raw16<S>::type x[NG];
#pragma unroll
      for (int u = 0; u < NG; u++) {
        if (kb + u < e1) x[u] = rld (Xp + (size_t) c[u] * W);
        else x[u] = {};
      }
#pragma unroll
      for (int u = 0; u < NG; u++) {
        T t[N];
        unpack (x[u], t);
#pragma unroll
        for (int i = 0; i < N; i++) acc[i] += w[u] * t[i];
      }
    }
#pragma unroll
    for (int i = 0; i < N; i++) part[s][lc + i] = acc[i];
  }
  __syncthreads ();
  if (g != 0) return;
  T acc[N];
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = part[0][lc + i];
  for (int s = 1; s < LCH; s++)
#pragma unroll
    for (int i = 0; i < N; i++) acc[i] += part[s][lc + i];
  S *y = Y + ((size_t) blockIdx.y * n + row) * W + lc;
  if (beta == (T) 0) {
#pragma unroll
    for (int i = 0; i < N; i++) acc[i] *= alpha;
  }
  else {
    T yo[N];
    vld (y, yo);
#pragma unroll
    for (int i = 0; i < N; i++) acc[i] = alpha * acc[i] + beta * yo[i];
  }
  vst (y, acc);
}

// R = B - A X
template <typename TV, typename T, typename S, int W = 64>
__global__ void __launch_bounds__ (TPB)
resid64_k (int n, const int *__restrict__ rp, const int *__restrict__ ci,
           const TV *__restrict__ v, const S *__restrict__ B,
           const S *__restrict__ X, S *__restrict__ R)
{
  LGH_ROWMAP (S)
  T acc[N];
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = 0;
  for (int kb = k0; kb < k1; kb += NG) {
    int c[NG];
    T w[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      const bool in = kb + u < k1;
      c[u] = in ? ci[kb + u] : row;
      w[u] = in ? (T) v[kb + u] : (T) 0;
    }
    typename Raw16<S>::type x[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      if (kb + u < k1) x[u] = rld (X + off + (size_t) c[u] * W);
      else x[u] = {};
    }
#pragma unroll
    for (int u = 0; u < NG; u++) {
      T t[N];
      unpack (x[u], t);
#pragma unroll
      for (int i = 0; i < N; i++) acc[i] += w[u] * t[i];
    }
  }
  T b[N];
  vld (B + off + (size_t) row * W, b);
#pragma unroll
  for (int i = 0; i < N; i++) b[i] -= acc[i];
  vst (R + off + (size_t) row * W, b);
}

// fused Chebyshev step (zsolve_impl.h:104-141): Dn = a D + b Z with
// Z = dinv R (JAC) or Zs; X (+)= Dn; Rn = R - A Dn, with Dn recomputed at
// the neighbours from the old state (a pointwise function of it)
// (SP: storage of the preconditioned residual Zs and of Rpc, the copy of
// the new residual that feeds the preconditioner when it runs in a lower
// precision; Rpc may be null)
template <typename TV, typename T, bool JAC, bool HASD, typename SP, int W = 64>
__global__ void __launch_bounds__ (TPB)
cheb64_k (int n, const int *__restrict__ rp, const int *__restrict__ ci,
          const TV *__restrict__ v, const TV *__restrict__ dinv,
          const SP *__restrict__ Zs, T a, T b, const T *__restrict__ D,
          const T *__restrict__ R, T *__restrict__ X, int xmode,
          T *__restrict__ Dn, T *__restrict__ Rn, SP *__restrict__ Rpc)
{
  LGH_ROWMAP (T)
  // dnew at one position: (a D + b z), z = dj R (JAC) or Zs
  auto dnew = [&] (size_t q, T dj, T (&o)[N]) {
    T z[N], d[N];
    if (JAC) {
      vld (R + q, z);
#pragma unroll
      for (int i = 0; i < N; i++) z[i] = dj * z[i];
    }
    else vld (Zs + q, z);
    if (HASD) vld (D + q, d);
#pragma unroll
    for (int i = 0; i < N; i++) {
      const T bz = b * z[i];
      if (HASD) {
        const T ad = a * d[i];
        o[i] = ad + bz;
      }
      else o[i] = bz;
    }
  };
  T acc[N];
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = 0;
  for (int kb = k0; kb < k1; kb += NG) {
    int c[NG];
    T w[NG], dd[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      const bool in = kb + u < k1;
      c[u] = in ? ci[kb + u] : row;
      w[u] = in ? (T) v[kb + u] : (T) 0;
    }
#pragma unroll
    for (int u = 0; u < NG; u++) dd[u] = JAC ? (T) dinv[c[u]] : (T) 0;
    T y[NG][N];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      if (kb + u < k1) dnew (off + (size_t) c[u] * W, dd[u], y[u]);
      else vzero (y[u]);
    }
#pragma unroll
    for (int u = 0; u < NG; u++)
#pragma unroll
      for (int i = 0; i < N; i++) acc[i] += w[u] * y[u][i];
  }
  const size_t qi = off + (size_t) row * W;
  T di[N], r[N], x[N];
  dnew (qi, JAC ? (T) dinv[row] : (T) 0, di);
  vld (R + qi, r);
  if (xmode) {
    vld (X + qi, x);
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = x[i] + di[i];
  }
  else {
#pragma unroll
    for (int i = 0; i < N; i++) x[i] = di[i];
  }
#pragma unroll
  for (int i = 0; i < N; i++) r[i] -= acc[i];
  vst (Dn + qi, di);
  vst (X + qi, x);
  vst (Rn + qi, r);
  if (Rpc) vst (Rpc + qi, r);
}

// mixed-precision fused Chebyshev step: FP64 arithmetic on the residual R
// and the accumulator X; preconditioned residual Z in the preconditioner's
// precision SP (FP32 or BF16), direction D in FP32.  The new direction is
// rounded to FP32 once and that stored value is what X and R both see, so
// R stays the residual of X (up to R's own storage rounding) and the
// iteration reaches FP64 accuracy with a low-precision preconditioner (the
// rounding only perturbs the step).  Also
// writes R in SP for the next preconditioner application.  (A BF16
// direction saves 4 of ~44 B per element but its 2^-9 rounding, repeated
// over the 25 steps, costs the solve a factor ~2 in accuracy.)
// RI, RO, XT: storage of the input residual, the new residual and X
// (double; float in a solve's early phase, and R after its residual
// replacement, see cheb_mixed_run); XI: storage of the X that is read
// (Xin: X itself, or the early phase's FP32 X at the first update after
// the replacement); the arithmetic is FP64 either way.
template <bool HASD, typename SP, int W = 64, typename RI = double,
          typename RO = double, typename XT = RO, typename XI = XT>
__global__ void __launch_bounds__ (TPB)
cheb64m_k (int n, const int *__restrict__ rp, const int *__restrict__ ci,
           const double *__restrict__ v, const SP *__restrict__ Zs,
           double a, double b, const float *__restrict__ D,
           const RI *__restrict__ R, const XI *Xin, XT *X, int xmode,
           float *__restrict__ Dn, RO *__restrict__ Rn,
           SP *__restrict__ Rpc, const double *__restrict__ xsc)
{
  // 16 lanes per row, 4 values per lane (fewer, wider loads than the fp64
  // map: the gathered Z and D are 8-16 B per lane)
  constexpr int N = 4, LPR = W / N;
  const int row = blockIdx.x * (TPB / LPR) + (int) threadIdx.x / LPR;
  const size_t off = (size_t) blockIdx.y * n * W
                     + (size_t) ((int) threadIdx.x % LPR) * N;
  if (row >= n) return;
  const int k0 = rp[row], k1 = rp[row + 1];
  constexpr int NGM = 6;            // entries per group (5- and 7-point Z)
  double acc[N];
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = 0;
  for (int kb = k0; kb < k1; kb += NGM) {
    int c[NGM];
    double w[NGM];
#pragma unroll
    for (int u = 0; u < NGM; u++) {
      const bool in = kb + u < k1;
      c[u] = in ? ci[kb + u] : row;
      w[u] = in ? v[kb + u] : 0.0;
    }
    typename Raw4<SP>::type zr[NGM];
    typename Raw4<float>::type dr[NGM];
#pragma unroll
    for (int u = 0; u < NGM; u++) {
      zr[u] = {};
      dr[u] = {};
      if (kb + u < k1) {
        const size_t q = off + (size_t) c[u] * W;
        zr[u] = rld4 (Zs + q);
        if (HASD) dr[u] = rld4 (D + q);
      }
    }
    // the new direction at each neighbour, rounded to FP32 as stored
#pragma unroll
    for (int u = 0; u < NGM; u++) {
      double z[N], d[N];
      unpack4 (zr[u], z);
      unpack4 (dr[u], d);
#pragma unroll
      for (int i = 0; i < N; i++) {
        const double bz = b * z[i];
        const double y = (double) (float) (HASD ? a * d[i] + bz : bz);
        acc[i] += w[u] * y;
      }
    }
  }
  const size_t qi = off + (size_t) row * W;
  double di[N], r[N], x[N], dold[N];
  {
    double z[N];
    vld (Zs + qi, z);
    if (HASD) vld (D + qi, dold);
#pragma unroll
    for (int i = 0; i < N; i++) {
      const double bz = b * z[i];
      di[i] = (double) (float) (HASD ? a * dold[i] + bz : bz);
    }
  }
  vld (R + qi, r);
  // X update (MIX_X_*): the deferred forms add the previous direction
  // first, so X ends bit-identical to updating it at every step
  if (xmode == MIX_X_ADD || xmode == MIX_X_ADD2) vld (Xin + qi, x);
#pragma unroll
  for (int i = 0; i < N; i++) {
    if (xmode == MIX_X_SET) x[i] = di[i];
    else if (xmode == MIX_X_ADD) x[i] = x[i] + di[i];
    else if (xmode == MIX_X_SET2) x[i] = (HASD ? dold[i] : 0.0) + di[i];
    else if (xmode == MIX_X_ADD2) x[i] = (x[i] + (HASD ? dold[i] : 0.0)) + di[i];
  }
#pragma unroll
  for (int i = 0; i < N; i++) r[i] -= acc[i];
  ntst (Dn + qi, di);
  if (xmode != MIX_X_SKIP) {
    // the solve's last step may hand X over row-scaled (k_diag_scale's
    // product, rounded after X's own sum)
    if (xsc) {
      const double sr = xsc[row];
#pragma unroll
      for (int i = 0; i < N; i++) x[i] = sr * x[i];
    }
    vst (X + qi, x);
  }
  vst (Rn + qi, r);
  ntst (Rpc + qi, r);
}

// the mixed solve's residual replacement (zsolve.cpp cheb_mixed_run) in
// one pass: R = B - A X with resid64_k's sums over X32 widened (X in FP64
// is the FP32 X widened; the solve's next X update reads X32 itself), R
// stored in RT, and Rpc = R in the V-cycle's storage precision -- in place
// of a widening copy, an FP64 residual that reads X back, and a narrowing
// copy of R
template <typename SP, int W = 64, typename RT = double>
__global__ void __launch_bounds__ (TPB)
mix_switch_k (int n, const int *__restrict__ rp, const int *__restrict__ ci,
              const double *__restrict__ v, const double *__restrict__ B,
              const float *__restrict__ X32, RT *__restrict__ R,
              SP *__restrict__ Rpc)
{
  LGH_ROWMAP (double)
  double acc[N];
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = 0;
  for (int kb = k0; kb < k1; kb += NG) {
    int c[NG];
    double w[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      const bool in = kb + u < k1;
      c[u] = in ? ci[kb + u] : row;
      w[u] = in ? v[kb + u] : 0.0;
    }
    float2 x[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      if (kb + u < k1) x[u] = *(const float2 *) (X32 + off + (size_t) c[u] * W);
      else x[u] = {};
    }
#pragma unroll
    for (int u = 0; u < NG; u++) {
      const double t[N] = { (double) x[u].x, (double) x[u].y };
#pragma unroll
      for (int i = 0; i < N; i++) acc[i] += w[u] * t[i];
    }
  }
  const size_t q = off + (size_t) row * W;
  double b[N];
  vld (B + q, b);
#pragma unroll
  for (int i = 0; i < N; i++) b[i] -= acc[i];
  vst (R + q, b);
  vst (Rpc + q, b);
}

int
k_mix_switch (const DevCSR &A, const PMat &B, const PMat &X32, PMat &R,
              PMat &Rpc)
{
  const int W = B.W;
  LGH_CHECK (A.prec == F64 && B.prec == F64 && X32.prec == F32
             && (R.prec == F64 || R.prec == F32) && (Rpc.prec == F32 || Rpc.prec == BF16)
             && row_width (W) && X32.W == W && R.W == W && Rpc.W == W
             && B.nrows == A.n && A.n == A.m && X32.nrows == A.n
             && R.nrows == A.n && Rpc.nrows == A.n
             && X32.npanels () == B.npanels ()
             && R.npanels () == B.npanels () && Rpc.npanels () == B.npanels (),
             8, "k_mix_switch: shape/precision mismatch");
  const double vec = (double) A.n * W;
  ProfScope ps (PK_OTHER, B.npanels () * (csr_bytes (A)
                + vec * (8 + 4 + prec_bytes (R.prec) + prec_bytes (Rpc.prec))),
                B.npanels () * 2.0 * (double) A.nnz * W, B.npanels (), Rpc.prec, W);
  const dim3 g = grid64 (A.n, B.npanels (), F64, W);
  auto go = [&] (auto sp, auto wc, auto rt) -> int {
    using SP = decltype (sp);
    using RT = decltype (rt);
    constexpr int WW = decltype (wc)::value;
    hipLaunchKernelGGL ((mix_switch_k<SP, WW, RT>), g, dim3 (TPB), 0, stream (),
                        A.n, A.rp, A.ci, (const double *) A.v, (const double *) B.p,
                        (const float *) X32.p, (RT *) R.p, (SP *) Rpc.p);
    LGH_LAUNCH_CHECK ();
    return 0;
  };
  return with_row_W (W, [&] (auto wc) -> int {
    if (Rpc.prec == BF16)
      return R.prec == F64 ? go (bf16_t (), wc, 0.0) : go (bf16_t (), wc, 0.f);
    return R.prec == F64 ? go (0.f, wc, 0.0) : go (0.f, wc, 0.f);
  });
}

int
k_cheb_step_mixed (const DevCSR &A, const PMat &Zpc, double a, double b,
                   const PMat *D32, const PMat &R, PMat &X, int x_mode,
                   PMat &Dn32, PMat &Rn, PMat &Rnpc, const PMat *Xin,
                   const double *xscale)
{
  const bool xin32 = Xin && Xin->p != X.p;
  LGH_CHECK (!xscale || (X.prec == F64 && x_mode != MIX_X_SKIP), 8,
             "k_cheb_step_mixed: a row-scaled X is written by an FP64 X update");
  LGH_CHECK (!xin32 || (Xin->prec == F32 && X.prec == F64 && R.prec == F32
                        && Xin->W == X.W && Xin->nrows == X.nrows
                        && Xin->npanels () == X.npanels ()), 8,
             "k_cheb_step_mixed: X input shape/precision");
  const int W = R.W;
  // storage (R, Rn, X): all FP64; FP64, FP32, FP32 (a solve's first
  // step); all FP32 (its early phase); FP32, FP32, FP64 (after the
  // residual replacement)
  const bool combo = (R.prec == F64 && Rn.prec == F64 && X.prec == F64)
                     || (R.prec == F64 && Rn.prec == F32 && X.prec == F32)
                     || (R.prec == F32 && Rn.prec == F32);
  LGH_CHECK (A.prec == F64 && combo && row_width (W)
             && X.W == W && Zpc.W == W && Dn32.W == W && Rn.W == W
             && Rnpc.W == W && (!D32 || D32->W == W)
             && (Zpc.prec == F32 || Zpc.prec == BF16) && Rnpc.prec == Zpc.prec
             && Dn32.prec == F32 && (!D32 || D32->prec == F32)
             && R.nrows == A.n, 8,
             "k_cheb_step_mixed: shape/precision mismatch");
  LGH_CHECK (x_mode >= MIX_X_SET && x_mode <= MIX_X_ADD2
             && (D32 || (x_mode != MIX_X_SET2 && x_mode != MIX_X_ADD2)), 8,
             "k_cheb_step_mixed: bad X mode");
  const double vec = (double) A.n * W;
  const double rb = (double) prec_bytes (R.prec), ob = (double) prec_bytes (Rn.prec);
  const double xs = (double) prec_bytes (X.prec);
  const double xr = xin32 ? 4.0 : xs;
  const double xb = (x_mode == MIX_X_SKIP) ? 0 : (x_mode == MIX_X_SET
                     || x_mode == MIX_X_SET2) ? xs : xr + xs;
  const double pb = (double) prec_bytes (Zpc.prec);
  // PK_CHEB_MIXED times one kernel instantiation: the steps after the
  // residual replacement (FP32 R, FP64 X read and written); the first step
  // of a solve (no D), the early phase (FP32 X), the first X update after
  // the replacement (FP32 X read) and solves of fewer than 8 steps (FP64 R)
  // are credited to PK_OTHER
  ProfScope ps ((D32 && R.prec == F32 && X.prec == F64 && !xin32) ? PK_CHEB_MIXED
                : PK_OTHER,
                R.npanels () * (csr_bytes (A) + (xscale ? 8.0 * A.n : 0.0)
                + vec * (pb + (D32 ? 4 : 0) + rb + xb + 4 + ob + pb)),
                R.npanels () * (2.0 * (double) A.nnz + 6.0 * A.n) * W,
                R.npanels (), Zpc.prec, W);
  const dim3 g = grid64 (A.n, R.npanels (), F32, W);   // 4 columns per lane
  auto go = [&] (auto sp, auto wc, auto ri, auto ro, auto xt, auto xi) -> int {
    using SP = decltype (sp);
    using RI = decltype (ri);
    using RO = decltype (ro);
    using XT = decltype (xt);
    using XI = decltype (xi);
    constexpr int WW = decltype (wc)::value;
    const XI *xin = (const XI *) (xin32 ? Xin->p : X.p);
    if (D32)
      hipLaunchKernelGGL ((cheb64m_k<true, SP, WW, RI, RO, XT, XI>), g, dim3 (TPB), 0,
                          stream (), A.n, A.rp, A.ci, (const double *) A.v,
                          (const SP *) Zpc.p, a, b, (const float *) D32->p,
                          (const RI *) R.p, xin, (XT *) X.p, x_mode,
                          (float *) Dn32.p, (RO *) Rn.p, (SP *) Rnpc.p, xscale);
    else
      hipLaunchKernelGGL ((cheb64m_k<false, SP, WW, RI, RO, XT, XI>), g, dim3 (TPB), 0,
                          stream (), A.n, A.rp, A.ci, (const double *) A.v,
                          (const SP *) Zpc.p, a, b, nullptr,
                          (const RI *) R.p, xin, (XT *) X.p, x_mode,
                          (float *) Dn32.p, (RO *) Rn.p, (SP *) Rnpc.p, xscale);
    LGH_LAUNCH_CHECK ();
    return 0;
  };
  return with_row_W (W, [&] (auto wc) {
    auto by_r = [&] (auto sp) {
      if (R.prec == F64 && X.prec == F64) return go (sp, wc, 0.0, 0.0, 0.0, 0.0);
      if (R.prec == F64) return go (sp, wc, 0.0, 0.f, 0.f, 0.f);
      if (X.prec == F64)
        return xin32 ? go (sp, wc, 0.f, 0.f, 0.0, 0.f) : go (sp, wc, 0.f, 0.f, 0.0, 0.0);
      return go (sp, wc, 0.f, 0.f, 0.f, 0.f);
    };
    return Zpc.prec == BF16 ? by_r (bf16_t ()) : by_r (0.f);
  });
}

// one value of the degree-2 smoother's update: D1 = b1 dr r0, R1 = r0 - acc
// (acc = sum_j A_ij D1(j)), D2 = a2 D1 + b2 dr R1, out = (x0 +) D1 + D2.
// Every operation is rounded as written (no fused multiply-add), so the
// per-level smoother (smx_k) and the fused coarse cycle (mgc_k) round alike.
template <typename T>
__device__ __forceinline__ T
smx_update (T r0, T acc, T x0, T dr, T b1, T a2, T b2, bool xadd)
{
#pragma clang fp contract(off)
  const T d1 = b1 * (dr * r0);
  const T r1 = r0 - acc;
  const T bz = b2 * (dr * r1);
  const T d2 = a2 * d1 + bz;
  return xadd ? (x0 + d1) + d2 : d1 + d2;
}

// degree-2 correction-form smoother (zsolve_impl.h:385-404 with nu = 2)
// in one pass.  X2 = X0 + D1 + D2 needs R0 only at
// radius 1: R1(i) = R0(i) - sum_j A_ij D1(j) with D1(j) = b1 dinv_j R0(j),
// then D2(i) = a2 D1(i) + b2 dinv_i R1(i) (smx_update); D1, R1 and R2
// are never stored.  SC: v holds the level's scaled
// values A_ij (b1 dinv_j), rounded as the product below rounds them, so
// the neighbours' dinv is not gathered and the result is the same bits.
template <typename TV, typename T, typename S, bool XADD, int W = 64,
          bool SC = false>
__global__ void __launch_bounds__ (TPB)
smx_k (int n, const int *__restrict__ rp, const int *__restrict__ ci,
       const TV *__restrict__ v, const TV *__restrict__ dinv, T b1, T a2,
       T b2, const S *__restrict__ R0, S *__restrict__ X)
{
  LGH_ROWMAP (S)
  T acc[N];
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = 0;
  for (int kb = k0; kb < k1; kb += NG) {
    int c[NG];
    T w[NG], dd[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      const bool in = kb + u < k1;
      c[u] = in ? ci[kb + u] : row;
      w[u] = in ? (T) v[kb + u] : (T) 0;
    }
#pragma unroll
    for (int u = 0; u < NG; u++)
      dd[u] = SC ? w[u] : w[u] * (b1 * (T) dinv[c[u]]);
    typename Raw16<S>::type x[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      if (kb + u < k1) x[u] = rld (R0 + off + (size_t) c[u] * W);
      else x[u] = {};
    }
    // sum_j A_ij D1(j) with D1(j) = b1 dinv_j R0(j): one FMA per value
#pragma unroll
    for (int u = 0; u < NG; u++) {
      T t[N];
      unpack (x[u], t);
#pragma unroll
      for (int i = 0; i < N; i++) acc[i] += dd[u] * t[i];
    }
  }
  const size_t qi = off + (size_t) row * W;
  const T dr = (T) dinv[row];
  T r0[N], x0[N], out[N];
  vld (R0 + qi, r0);
  if (XADD) vld (X + qi, x0);
#pragma unroll
  for (int i = 0; i < N; i++)
    out[i] = smx_update (r0[i], acc[i], XADD ? x0[i] : (T) 0, dr, b1, a2, b2, XADD);
  vst (X + qi, out);
}

int
k_smooth2x (const DevCSR &A, double b1, double a2, double b2, const PMat &R0,
            PMat &X, int xadd)
{
  LGH_CHECK (row_width (R0.W) && X.W == R0.W && R0.prec == X.prec && A.dinv
             && R0.nrows == A.n && X.nrows == A.n
             && R0.npanels () == X.npanels ()
             && (R0.prec != BF16 || A.prec == F32), 8, "k_smooth2x: shape mismatch");
  const double vec = (double) A.n * R0.W * prec_bytes (R0.prec);
  const int kind = prof_fine_tag () ? (xadd ? PK_SM_POST : PK_SM_PRE) : PK_CHEB_JAC;
  ProfScope ps (kind, R0.npanels () * (csr_bytes (A) + (double) A.n * prec_bytes (A.prec)
                                      + (2 + (xadd ? 1 : 0)) * vec),
                R0.npanels () * (2.0 * (double) A.nnz + 9.0 * A.n) * R0.W,
                R0.npanels (), R0.prec, R0.W);
  const dim3 g = grid64 (A.n, R0.npanels (), R0.prec, R0.W);
  auto go = [&] (auto tv, auto t, auto st, auto wc) -> int {
    using TV = decltype (tv);
    using T = decltype (t);
    using S = decltype (st);
    constexpr int W = decltype (wc)::value;
    // the scaled values apply when they were made for this b1 and the
    // arithmetic is in their precision
    const bool sc = std::is_same<TV, T>::value && A.vs && A.vs_b1 == b1;
    const TV *vv = (const TV *) (sc ? A.vs : A.v);
    auto launch = [&] (auto xa, auto scc) {
      constexpr bool XA = decltype (xa)::value, SCC = decltype (scc)::value;
      hipLaunchKernelGGL ((smx_k<TV, T, S, XA, W, SCC>), g, dim3 (TPB), 0,
                          stream (), A.n, A.rp, A.ci, vv, (const TV *) A.dinv,
                          (T) b1, (T) a2, (T) b2, (const S *) R0.p, (S *) X.p);
    };
    using yes = std::true_type;
    using no = std::false_type;
    if (xadd) { if (sc) launch (yes (), yes ()); else launch (yes (), no ()); }
    else { if (sc) launch (no (), yes ()); else launch (no (), no ()); }
    LGH_LAUNCH_CHECK ();
    return 0;
  };
  return with_row_W (R0.W, [&] (auto wc) {
    if (R0.prec == BF16) return go (0.f, 0.f, bf16_t (), wc);
    if (R0.prec == F64)
      return A.prec == F64 ? go (0.0, 0.0, 0.0, wc) : go (0.f, 0.0, 0.0, wc);
    return A.prec == F64 ? go (0.0, 0.f, 0.f, wc) : go (0.f, 0.f, 0.f, wc);
  });
}

// ---------------------------------------------------------------------------
// SpMM: Y = alpha A X + beta Y

template <typename TV, typename T, int W>
__global__ void __launch_bounds__ (TPB)
spmm_k (int n, int nx, const int *__restrict__ rp, const int *__restrict__ ci,
        const TV *__restrict__ v, T alpha, const T *__restrict__ X, T beta,
        T *__restrict__ Y)
{
  constexpr int RPW = 64 / W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane % W;
  const int row = (blockIdx.x * (TPB / 64) + wave) * RPW + lane / W;
  if (row >= n) return;
  const size_t p = blockIdx.y;
  const T *Xp = X + p * (size_t) nx * W + c;
  T acc = 0;
  const int k1 = rp[row + 1];
  for (int k = rp[row]; k < k1; k++)
    acc += (T) v[k] * Xp[(size_t) ci[k] * W];
  T *y = Y + (p * (size_t) n + row) * W + c;
  *y = (beta == (T) 0) ? alpha * acc : alpha * acc + beta * *y;
}

template <typename TV, typename T, typename S = T>
static int
spmm_t (const DevCSR &A, double alpha, const PMat &X, double beta, PMat &Y)
{
  if constexpr (std::is_same<S, T>::value) {
    if (A.tile.tp && X.W == 64 && (A.tile.s.rg != 16 || std::is_same<T, double>::value))
      return spmm_tile_launch<TV, T, S> (A, 0, X.npanels (), (T) alpha,
                                         (const S *) X.p, (T) beta, (S *) Y.p);
    if (A.grp.gp && X.W == 64)
      return spmm_grp_launch<TV, T, S> (A, A.grp_ppw[std::is_same<T, double>::value ? 0 : 1],
                                        0, X.npanels (), (T) alpha,
                                        (const S *) X.p, (T) beta, (S *) Y.p);
  }
  // rows of 64 entries and more on average in a rectangular matrix (a
  // restriction onto a small coarse level) take the chunked kernel; a
  // square operator with wide rows (a PSF B) keeps the row-order sums of
  // spmm64_k, which its row-grouped copy reproduces bit for bit
  if (row_width (X.W) && A.n > 0 && A.n != A.m && A.nnz >= 64 * (int64_t) A.n)
    return with_row_W (X.W, [&] (auto wc) -> int {
      constexpr int W = decltype (wc)::value;
      hipLaunchKernelGGL ((spmm_long_k<TV, T, S, W>),
                          dim3 ((unsigned) A.n, (unsigned) X.npanels ()),
                          dim3 (TPB), 0, stream (), A.n, A.m, A.rp, A.ci,
                          (const TV *) A.v, (T) alpha, (const S *) X.p,
                          (T) beta, (S *) Y.p);
      LGH_LAUNCH_CHECK ();
      return 0;
    });
  if (row_width (X.W))
    return with_row_W (X.W, [&] (auto wc) -> int {
      constexpr int W = decltype (wc)::value;
      hipLaunchKernelGGL ((spmm64_k<TV, T, S, W>),
                          grid64 (A.n, X.npanels (), X.prec, W), dim3 (TPB), 0,
                          stream (), A.n, A.m, A.rp, A.ci, (const TV *) A.v,
                          (T) alpha, (const S *) X.p, (T) beta, (S *) Y.p);
      LGH_LAUNCH_CHECK ();
      return 0;
    });
  LGH_CHECK ((std::is_same<S, T>::value), 8,
             "k_spmm: bf16 storage needs W = 64 or 16");
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      constexpr int rows_per_block = (TPB / 64) * (64 / W);
      dim3 g (grid1 (A.n, rows_per_block), X.npanels ());
      hipLaunchKernelGGL ((spmm_k<TV, T, W>), g, dim3 (TPB), 0, stream (),
                          A.n, A.m, A.rp, A.ci, (const TV *) A.v, (T) alpha,
                          (const T *) X.p, (T) beta, (T *) Y.p);
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  return with_W (X.W, l);
}

int
k_spmm (const DevCSR &A, double alpha, const PMat &X, double beta, PMat &Y)
{
  LGH_CHECK (X.W == Y.W && X.prec == Y.prec && X.npanels () == Y.npanels ()
             && X.nrows == A.m && Y.nrows == A.n, 8, "k_spmm: shape mismatch");
  const double sx = (double) prec_bytes (X.prec), np_ = X.npanels ();
  ProfScope ps (prof_b_tag () ? PK_SPMM_B : PK_SPMM,
                np_ * (csr_bytes (A) + (double) A.m * X.W * sx
                       + (double) A.n * X.W * sx * (beta != 0.0 ? 2 : 1)),
                np_ * 2.0 * (double) A.nnz * X.W, (int) np_, X.prec, X.W);
  if (X.prec == BF16) {
    LGH_CHECK (A.prec == F32, 8, "k_spmm: bf16 vectors need fp32 values");
    return spmm_t<float, float, bf16_t> (A, alpha, X, beta, Y);
  }
  if (X.prec == F64)
    return A.prec == F64 ? spmm_t<double, double> (A, alpha, X, beta, Y)
                         : spmm_t<float, double> (A, alpha, X, beta, Y);
  return A.prec == F64 ? spmm_t<double, float> (A, alpha, X, beta, Y)
                       : spmm_t<float, float> (A, alpha, X, beta, Y);
}

// ---------------------------------------------------------------------------
// residual R = B - A X

template <typename TV, typename T, int W>
__global__ void __launch_bounds__ (TPB)
resid_k (int n, const int *__restrict__ rp, const int *__restrict__ ci,
         const TV *__restrict__ v, const T *__restrict__ B,
         const T *__restrict__ X, T *__restrict__ R)
{
  constexpr int RPW = 64 / W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane % W;
  const int row = (blockIdx.x * (TPB / 64) + wave) * RPW + lane / W;
  if (row >= n) return;
  const size_t off = (size_t) blockIdx.y * n * W + c;
  T acc = 0;
  const int k1 = rp[row + 1];
  for (int k = rp[row]; k < k1; k++)
    acc += (T) v[k] * X[off + (size_t) ci[k] * W];
  R[off + (size_t) row * W] = B[off + (size_t) row * W] - acc;
}

template <typename TV, typename T, typename S = T>
static int
resid_t (const DevCSR &A, const PMat &B, const PMat &X, PMat &R)
{
  if (row_width (X.W))
    return with_row_W (X.W, [&] (auto wc) -> int {
      constexpr int W = decltype (wc)::value;
      hipLaunchKernelGGL ((resid64_k<TV, T, S, W>),
                          grid64 (A.n, X.npanels (), X.prec, W), dim3 (TPB), 0,
                          stream (), A.n, A.rp, A.ci, (const TV *) A.v,
                          (const S *) B.p, (const S *) X.p, (S *) R.p);
      LGH_LAUNCH_CHECK ();
      return 0;
    });
  LGH_CHECK ((std::is_same<S, T>::value), 8,
             "k_resid: bf16 storage needs W = 64 or 16");
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      constexpr int rows_per_block = (TPB / 64) * (64 / W);
      dim3 g (grid1 (A.n, rows_per_block), X.npanels ());
      hipLaunchKernelGGL ((resid_k<TV, T, W>), g, dim3 (TPB), 0, stream (),
                          A.n, A.rp, A.ci, (const TV *) A.v, (const T *) B.p,
                          (const T *) X.p, (T *) R.p);
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  return with_W (X.W, l);
}

int
k_resid (const DevCSR &A, const PMat &B, const PMat &X, PMat &R)
{
  LGH_CHECK (X.W == R.W && B.W == R.W && X.prec == R.prec && B.prec == R.prec
             && X.nrows == A.n, 8, "k_resid: shape mismatch");
  ProfScope ps (prof_fine_tag () ? PK_RESID_FINE : PK_RESID,
                X.npanels () * (csr_bytes (A) + 3.0 * A.n * X.W * prec_bytes (X.prec)),
                X.npanels () * 2.0 * (double) A.nnz * X.W, X.npanels (), X.prec,
                X.W);
  if (X.prec == BF16) {
    LGH_CHECK (A.prec == F32, 8, "k_resid: bf16 vectors need fp32 values");
    return resid_t<float, float, bf16_t> (A, B, X, R);
  }
  if (X.prec == F64)
    return A.prec == F64 ? resid_t<double, double> (A, B, X, R)
                         : resid_t<float, double> (A, B, X, R);
  return A.prec == F64 ? resid_t<double, float> (A, B, X, R)
                       : resid_t<float, float> (A, B, X, R);
}

// ---------------------------------------------------------------------------
// fused Chebyshev step (zsolve_impl.h:104-141 recurrence, one launch per
// iteration).  With Z = P R for a diagonal P (Jacobi) or a precomputed
// Zsrc, the new direction Dn = a D + b Z is a pointwise function of the
// old state, so each row recomputes Dn at its CSR neighbours instead of
// reading a materialized Dn: R_new = R - A Dn, X (+)= Dn, one pass.

template <typename TV, typename T, int W, bool JAC, bool HASD>
__global__ void __launch_bounds__ (TPB)
cheb_k (int n, const int *__restrict__ rp, const int *__restrict__ ci,
        const TV *__restrict__ v, const TV *__restrict__ dinv,
        const T *__restrict__ Zs, T a, T b, const T *__restrict__ D,
        const T *__restrict__ R, T *__restrict__ X, int xmode,
        T *__restrict__ Dn, T *__restrict__ Rn)
{
  constexpr int RPW = 64 / W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane % W;
  const int row = (blockIdx.x * (TPB / 64) + wave) * RPW + lane / W;
  if (row >= n) return;
  const size_t off = (size_t) blockIdx.y * n * W + c;
  auto dnew = [&] (int j) -> T {
    const size_t q = off + (size_t) j * W;
    T z = JAC ? (T) dinv[j] * R[q] : Zs[q];
    T bz = b * z;
    if (HASD) {
      T ad = a * D[q];
      return ad + bz;
    }
    return bz;
  };
  const T di = dnew (row);
  T acc = 0;
  const int k1 = rp[row + 1];
  for (int k = rp[row]; k < k1; k++) acc += (T) v[k] * dnew (ci[k]);
  const size_t qi = off + (size_t) row * W;
  Dn[qi] = di;
  X[qi] = xmode ? X[qi] + di : di;
  Rn[qi] = R[qi] - acc;
}

template <typename TV, typename T, typename SP = T>
static int
cheb_t (const DevCSR &A, int jac, const PMat *Zs, double a, double b,
        const PMat *D, const PMat &R, PMat &X, int xmode, PMat &Dn, PMat &Rn,
        PMat *Rpc)
{
  if (row_width (R.W))
    return with_row_W (R.W, [&] (auto wc) -> int {
      constexpr int W = decltype (wc)::value;
      const dim3 g = grid64 (A.n, R.npanels (), vt<T>::prec, W);
      const SP *zp = Zs ? (const SP *) Zs->p : nullptr;
      const T *dp = D ? (const T *) D->p : nullptr;
      SP *rpc = Rpc ? (SP *) Rpc->p : nullptr;
#define LGH_CH64(J, H)                                                        \
      hipLaunchKernelGGL ((cheb64_k<TV, T, J, H, SP, W>), g, dim3 (TPB), 0,   \
                          stream (), A.n, A.rp, A.ci, (const TV *) A.v,        \
                          (const TV *) A.dinv, zp, (T) a, (T) b, dp,           \
                          (const T *) R.p, (T *) X.p, xmode, (T *) Dn.p,       \
                          (T *) Rn.p, rpc)
      if (jac) { if (dp) LGH_CH64 (true, true); else LGH_CH64 (true, false); }
      else { if (dp) LGH_CH64 (false, true); else LGH_CH64 (false, false); }
#undef LGH_CH64
      LGH_LAUNCH_CHECK ();
      return 0;
    });
  LGH_CHECK ((std::is_same<SP, T>::value) && !Rpc, 8,
             "k_cheb_step: a lower-precision preconditioner needs W = 64 or 16");
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      constexpr int rows_per_block = (TPB / 64) * (64 / W);
      dim3 g (grid1 (A.n, rows_per_block), R.npanels ());
      const T *zp = Zs ? (const T *) Zs->p : nullptr;
      const T *dp = D ? (const T *) D->p : nullptr;
#define LGH_CHEB(J, H)                                                        \
      hipLaunchKernelGGL ((cheb_k<TV, T, W, J, H>), g, dim3 (TPB), 0,        \
                          stream (), A.n, A.rp, A.ci, (const TV *) A.v,       \
                          (const TV *) A.dinv, zp, (T) a, (T) b, dp,          \
                          (const T *) R.p, (T *) X.p, xmode, (T *) Dn.p,         \
                          (T *) Rn.p)
      if (jac) { if (dp) LGH_CHEB (true, true); else LGH_CHEB (true, false); }
      else { if (dp) LGH_CHEB (false, true); else LGH_CHEB (false, false); }
#undef LGH_CHEB
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  return with_W (R.W, l);
}

int
k_cheb_step (const DevCSR &A, int jacobi, const PMat *Zsrc, double a,
             double b, const PMat *D, const PMat &R, PMat &X, int x_mode,
             PMat &Dn, PMat &Rn, PMat *Rpc)
{
  LGH_CHECK (R.nrows == A.n && X.W == R.W && Dn.W == R.W && Rn.W == R.W
             && X.prec == R.prec && (!jacobi || A.dinv), 8,
             "k_cheb_step: shape mismatch");
  const int pc = Zsrc ? Zsrc->prec : R.prec;
  LGH_CHECK ((pc == R.prec || (pc == BF16 && R.prec == F32 && row_width (R.W)))
             && (!Rpc || Rpc->prec == pc), 8,
             "k_cheb_step: preconditioner precision");
  // streams: read R, (D), (Zsrc), (X); write Dn, X, Rn, (Rpc)
  const double vec = (double) A.n * R.W * prec_bytes (R.prec);
  const double vpc = (double) A.n * R.W * prec_bytes (pc);
  const double streams = 1 + (D ? 1 : 0) + (x_mode ? 1 : 0) + 3;
  ProfScope ps (jacobi ? PK_CHEB_JAC : PK_CHEB_Z,
                R.npanels () * (csr_bytes (A) + streams * vec
                                + ((jacobi ? 0 : 1) + (Rpc ? 1 : 0)) * vpc
                                + (jacobi ? (double) A.n * prec_bytes (A.prec) : 0.0)),
                R.npanels () * (2.0 * (double) A.nnz + 6.0 * A.n) * R.W,
                R.npanels (), R.prec, R.W);
  if (pc == BF16)
    return A.prec == F64
        ? cheb_t<double, float, bf16_t> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn, Rpc)
        : cheb_t<float, float, bf16_t> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn, Rpc);
  if (R.prec == F64)
    return A.prec == F64
        ? cheb_t<double, double> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn, Rpc)
        : cheb_t<float, double> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn, Rpc);
  return A.prec == F64
      ? cheb_t<double, float> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn, Rpc)
      : cheb_t<float, float> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn, Rpc);
}




// scalar element access in any storage precision (value as T)
template <typename T> __device__ inline T ld_as (const float *p, size_t i) { return (T) p[i]; }
template <typename T> __device__ inline T ld_as (const double *p, size_t i) { return (T) p[i]; }
template <typename T> __device__ inline T ld_as (const bf16_t *p, size_t i) { return (T) bf2f (p[i].u); }
template <typename T> __device__ inline void st_as (float *p, size_t i, T v) { p[i] = (float) v; }
template <typename T> __device__ inline void st_as (double *p, size_t i, T v) { p[i] = (double) v; }
template <typename T> __device__ inline void st_as (bf16_t *p, size_t i, T v) { p[i].u = (uint16_t) f2bf ((float) v); }

// ---------------------------------------------------------------------------
// dense coarse apply X0 = Ainv B0 (Ainv row-major n0 x n0; S storage of B, X)

template <typename T, typename S, int W>
__global__ void __launch_bounds__ (TPB)
dense_apply_k (int n0, const T *__restrict__ Ainv, const S *__restrict__ B,
               S *__restrict__ X)
{
  constexpr int RPW = 64 / W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane % W;
  const int row = (blockIdx.x * (TPB / 64) + wave) * RPW + lane / W;
  if (row >= n0) return;
  const size_t off = (size_t) blockIdx.y * n0 * W + c;
  T acc = 0;
  for (int k = 0; k < n0; k++)
    acc += Ainv[(size_t) row * n0 + k] * ld_as<T> (B, off + (size_t) k * W);
  st_as<T> (X, off + (size_t) row * W, acc);
}

int
k_dense_apply (const void *Ainv, int n0, const PMat &B, PMat &X)
{
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      constexpr int rows_per_block = (TPB / 64) * (64 / W);
      dim3 g (grid1 (n0, rows_per_block), B.npanels ());
      if (B.prec == F64)
        hipLaunchKernelGGL ((dense_apply_k<double, double, W>), g, dim3 (TPB), 0,
                            stream (), n0, (const double *) Ainv,
                            (const double *) B.p, (double *) X.p);
      else if (B.prec == F32)
        hipLaunchKernelGGL ((dense_apply_k<float, float, W>), g, dim3 (TPB), 0,
                            stream (), n0, (const float *) Ainv,
                            (const float *) B.p, (float *) X.p);
      else
        hipLaunchKernelGGL ((dense_apply_k<float, bf16_t, W>), g, dim3 (TPB), 0,
                            stream (), n0, (const float *) Ainv,
                            (const bf16_t *) B.p, (bf16_t *) X.p);
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  LGH_CHECK (B.nrows == n0 && X.nrows == n0 && B.W == X.W && B.prec == X.prec,
             8, "k_dense_apply: shape mismatch");
  return with_W (B.W, l);
}

// ---------------------------------------------------------------------------
// fused coarse cycle (zsolve_impl.h:406-485 on the small levels).  Below a
// thousand rows a level's kernels are a few microseconds of work each, and
// a cycle spends its time between dependent launches.  The columns of a
// cycle are independent, so one workgroup can run every small level for its
// own N columns (16 bytes of a panel row) of one panel, ordering the
// passes by workgroup barriers.  Every level's slice of B, X and R lives in
// LDS for the whole launch (16 bytes a row): the top level's B comes in
// from HBM at the start and its X goes out at the end, and in between a
// pass reads only the CSR arrays (L2-resident) from memory.  One launch per
// coarse visit replaces six per level.

constexpr int MGC_TPB = 1024;

struct MgcLev {
  int n, nc;                             // rows; rows of level l - 1
  int sc;                                // vs holds the scaled values
  const int *rp, *ci;                    // A
  const void *v, *vs, *dinv;             // values, the smoother's, 1/diag
  const int *prp, *pci;                  // P (n x nc)
  const void *pv;
  const int *trp, *tci;                  // P^T (nc x n)
  const void *tv;
  int ob, ox, orr;                       // LDS offsets of B, X, R (bytes)
  // LDS offsets of the arrays above when the launch stages them
  int orp, oci, ov, ovs, odinv, oprp, opci, opv, otrp, otci, otv;
  double b1, a2, b2;
};
struct MgcArgs {
  MgcLev lev[MGC_MAX_LEV];
  const void *cinv;
  int ocinv;
  const void *Btop;                      // the top level's B, X panels (HBM)
  void *Xtop;
  int top, n0, nops;
  const uint4 *img;                      // the packed arrays (MgcImage), copied
  int img_dst, img_n16;                  // to LDS byte img_dst, 16 B at a time
  unsigned char op[MGC_MAX_OPS];
};
// LDS bytes of the levels' slices: B and X per level, R above level 0
static inline int
mgc_lds_bytes (int nlev, const int *rows)
{
  int b = 0;
  for (int l = 0; l < nlev; l++) b += 16 * rows[l] * (l > 0 ? 3 : 2);
  return b;
}
constexpr int MGC_MAX_LDS = 64 * 1024;

// sum over one CSR row of w(k, c) X(c) in the row kernels' order: groups of
// NG entries, the slots past the row's end adding 0 * 0
template <typename TV, typename T, typename S, int STRIDE, typename WF>
__device__ __forceinline__ void
mgc_rowsum (int k0, int k1, int pad, const int *ci, const TV *v, const S *Xb,
            WF wf, T (&acc)[VT<S>::N])
{
  constexpr int N = VT<S>::N;
#pragma unroll
  for (int i = 0; i < N; i++) acc[i] = 0;
  for (int kb = k0; kb < k1; kb += NG) {
    int c[NG];
    T w[NG], dd[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      const bool in = kb + u < k1;
      c[u] = in ? ci[kb + u] : pad;
      w[u] = in ? (T) v[kb + u] : (T) 0;
    }
#pragma unroll
    for (int u = 0; u < NG; u++) dd[u] = wf (w[u], c[u]);
    typename Raw16<S>::type x[NG];
#pragma unroll
    for (int u = 0; u < NG; u++) {
      if (kb + u < k1) x[u] = rld (Xb + (size_t) c[u] * STRIDE);
      else x[u] = {};
    }
#pragma unroll
    for (int u = 0; u < NG; u++) {
      T t[N];
      unpack (x[u], t);
#pragma unroll
      for (int i = 0; i < N; i++) acc[i] += dd[u] * t[i];
    }
  }
}

// TV matrix values, T arithmetic, S vector storage, TD the dense inverse
// CL: the levels' CSR arrays, 1/diag and the dense inverse are staged in
// LDS too (when they fit beside the slices), so a pass reads no memory
template <typename TV, typename T, typename S, typename TD, int W, bool CL>
__global__ void __launch_bounds__ (MGC_TPB)
mgc_k (const MgcArgs a)
{
  constexpr int N = VT<S>::N;
  typedef typename Raw16<S>::type raw_t;
  extern __shared__ uint4 mgc_lds[];
  char *sm = (char *) mgc_lds;
  const int tid = threadIdx.x, nt = blockDim.x;
  const size_t col = (size_t) blockIdx.x * N;
  const size_t pan = blockIdx.y;
  auto plain = [] (T w, int) { return w; };
  // the level slices: row r of a level's B / X / R at byte 16 r of its area
  auto vb = [&] (int l) { return (S *) (sm + a.lev[l].ob); };
  auto vx = [&] (int l) { return (S *) (sm + a.lev[l].ox); };
  auto vr = [&] (int l) { return (S *) (sm + a.lev[l].orr); };
  // an array of the levels: its LDS copy (CL) or its memory
  auto at = [&] (const void *g, int o) -> const void * {
    return CL ? (const void *) (sm + o) : g;
  };
  if (CL) {
    // one contiguous copy, four 16-byte loads in flight per thread
    uint4 *dst = (uint4 *) (sm + a.img_dst);
    for (int w = tid; w < a.img_n16; w += 4 * nt) {
      uint4 t[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (w + u * nt < a.img_n16) t[u] = a.img[w + u * nt];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (w + u * nt < a.img_n16) dst[w + u * nt] = t[u];
    }
  }
  {
    const MgcLev &L = a.lev[a.top];
    const S *g = (const S *) a.Btop + pan * L.n * W + col;
    for (int row = tid; row < L.n; row += nt)
      *(raw_t *) (vb (a.top) + (size_t) row * N) = rld (g + (size_t) row * W);
  }
  __syncthreads ();
  for (int q = 0; q < a.nops; q++) {
    const int kind = a.op[q] >> 4, l = a.op[q] & 15;
    const MgcLev &L = a.lev[l];
    const int n = L.n;
    const int *rp = (const int *) at (L.rp, L.orp);
    const int *ci = (const int *) at (L.ci, L.oci);
    if (kind == MGC_DENSE) {
      const TD *Ai = (const TD *) at (a.cinv, a.ocinv);
      const S *B = vb (0);
      S *X = vx (0);
      for (int row = tid; row < n; row += nt) {
        TD acc[N], b[N];
#pragma unroll
        for (int i = 0; i < N; i++) acc[i] = 0;
        for (int k = 0; k < n; k++) {
          const TD ak = Ai[(size_t) row * n + k];
          vld (B + (size_t) k * N, b);
#pragma unroll
          for (int i = 0; i < N; i++) acc[i] += ak * b[i];
        }
        vst (X + (size_t) row * N, acc);
      }
    }
    else if (kind == MGC_RESTRICT) {
      // B(l - 1) = P^T R, rows of level l - 1
      const S *R = vr (l);
      S *Bc = vb (l - 1);
      const int *trp = (const int *) at (L.trp, L.otrp);
      const int *tci = (const int *) at (L.tci, L.otci);
      const TV *tv = (const TV *) at (L.tv, L.otv);
      for (int row = tid; row < L.nc; row += nt) {
        const int k0 = trp[row], k1 = trp[row + 1];
        T acc[N];
        mgc_rowsum<TV, T, S, N> (k0, k1, k0 < k1 ? tci[k0] : 0, tci, tv, R,
                                 plain, acc);
        vst (Bc + (size_t) row * N, acc);
      }
    }
    else if (kind == MGC_PROLONG) {
      // X += P X(l - 1)
      const S *Xc = vx (l - 1);
      S *X = vx (l);
      const int *prp = (const int *) at (L.prp, L.oprp);
      const int *pci = (const int *) at (L.pci, L.opci);
      const TV *pv = (const TV *) at (L.pv, L.opv);
      for (int row = tid; row < n; row += nt) {
        const int k0 = prp[row], k1 = prp[row + 1];
        T acc[N], xo[N];
        mgc_rowsum<TV, T, S, N> (k0, k1, k0 < k1 ? pci[k0] : 0, pci, pv, Xc,
                                 plain, acc);
        vld (X + (size_t) row * N, xo);
#pragma unroll
        for (int i = 0; i < N; i++) acc[i] = acc[i] + xo[i];
        vst (X + (size_t) row * N, acc);
      }
    }
    else if (kind == MGC_RESID) {
      const S *B = vb (l), *X = vx (l);
      S *R = vr (l);
      const TV *v = (const TV *) at (L.v, L.ov);
      for (int row = tid; row < n; row += nt) {
        T acc[N], b[N];
        mgc_rowsum<TV, T, S, N> (rp[row], rp[row + 1], row, ci, v, X, plain,
                                 acc);
        vld (B + (size_t) row * N, b);
#pragma unroll
        for (int i = 0; i < N; i++) b[i] -= acc[i];
        vst (R + (size_t) row * N, b);
      }
    }
    else {
      // degree-2 smoother (smx_k): MGC_SM0 from B, MGC_SMX adds to X from R
      const bool xadd = kind == MGC_SMX;
      const S *R0 = xadd ? vr (l) : vb (l);
      S *X = vx (l);
      const TV *dinv = (const TV *) at (L.dinv, L.odinv);
      const TV *vs = (const TV *) at (L.vs, L.ovs);
      const T b1 = (T) L.b1, a2 = (T) L.a2, b2 = (T) L.b2;
      const bool sc = L.sc;
      auto scaled = [&] (T w, int c) {
        return sc ? w : w * (b1 * (T) dinv[c]);
      };
      for (int row = tid; row < n; row += nt) {
        T acc[N], r0[N], x0[N], out[N];
        mgc_rowsum<TV, T, S, N> (rp[row], rp[row + 1], row, ci, vs, R0,
                                 scaled, acc);
        const T dr = (T) dinv[row];
        vld (R0 + (size_t) row * N, r0);
        if (xadd) vld (X + (size_t) row * N, x0);
        else vzero (x0);
#pragma unroll
        for (int i = 0; i < N; i++)
          out[i] = smx_update (r0[i], acc[i], x0[i], dr, b1, a2, b2, xadd);
        vst (X + (size_t) row * N, out);
      }
    }
    __syncthreads ();
  }
  {
    const MgcLev &L = a.lev[a.top];
    S *g = (S *) a.Xtop + pan * L.n * W + col;
    for (int row = tid; row < L.n; row += nt)
      *(raw_t *) (g + (size_t) row * W) = *(const raw_t *) (vx (a.top) + (size_t) row * N);
  }
}

bool
mgc_supported (int nlev, const MgcLevelIn *lev, int n0)
{
  if (nlev < 2 || nlev > MGC_MAX_LEV || !lev[0].B || !lev[0].X) return false;
  const PMat &v0 = *lev[0].B;
  if (!row_width (v0.W) || v0.nrows != n0) return false;
  const int mprec = lev[1].A ? lev[1].A->prec : -1;
  // (TV, T, S): (double, double, double), (float, float, float),
  // (float, float, bf16)
  const bool ok_prec = (v0.prec == F64 && mprec == F64)
                       || (v0.prec != F64 && mprec == F32);
  if (!ok_prec) return false;
  for (int l = 0; l < nlev; l++) {
    const MgcLevelIn &L = lev[l];
    if (!L.B || !L.X || (l > 0 && !L.R)) return false;
    for (const PMat *m : { L.B, L.X, L.R })
      if (m && (m->W != v0.W || m->prec != v0.prec || m->nrows != L.B->nrows
                || m->npanels () != v0.npanels ()))
        return false;
    if (l == 0) continue;
    const DevCSR *A = L.A, *P = L.P, *PT = L.PT;
    if (!A || !P || !PT || !A->dinv || A->prec != mprec || P->prec != mprec
        || PT->prec != mprec || A->n != A->m || L.B->nrows != A->n
        || P->n != A->n || PT->m != A->n || P->m != PT->n
        || P->m != lev[l - 1].B->nrows)
      return false;
    // the per-level product's order: spmm64_k (row order), not the long-row
    // kernel
    for (const DevCSR *Q : { P, PT })
      if (Q->n != Q->m && Q->nnz >= 64 * (int64_t) Q->n) return false;
  }
  int rows[MGC_MAX_LEV];
  for (int l = 0; l < nlev; l++) rows[l] = lev[l].B->nrows;
  return mgc_lds_bytes (nlev, rows) <= MGC_MAX_LDS;
}

void
MgcImage::release ()
{
  dfree (p);
  p = nullptr;
  bytes = 0;
  nlev = 0;
}

int
k_mg_coarse (int nlev, const MgcLevelIn *lev, const void *cinv, int n0,
             int nops, const unsigned char *ops, MgcImage *img,
             std::mutex *img_mu)
{
  LGH_CHECK (nops >= 1 && nops <= MGC_MAX_OPS && mgc_supported (nlev, lev, n0),
             8, "k_mg_coarse: unsupported levels");
  MgcArgs a {};
  a.cinv = cinv;
  a.n0 = n0;
  a.nops = nops;
  double bytes = 0.0, flops = 0.0;
  const PMat &v0 = *lev[0].B;
  const double vb = prec_bytes (v0.prec) * (double) v0.W;
  for (int q = 0; q < nops; q++) {
    const int kind = ops[q] >> 4, l = ops[q] & 15;
    LGH_CHECK (l < nlev && kind <= MGC_DENSE && (kind == MGC_DENSE) == (l == 0),
               8, "k_mg_coarse: bad operation");
    a.op[q] = ops[q];
    const MgcLevelIn &L = lev[l];
    if (kind == MGC_DENSE) {
      bytes += (double) n0 * n0 * prec_bytes (v0.prec == F64 ? F64 : F32) + 2.0 * n0 * vb;
      flops += 2.0 * n0 * n0 * v0.W;
    }
    else if (kind == MGC_RESTRICT || kind == MGC_PROLONG) {
      const DevCSR &Q = kind == MGC_RESTRICT ? *L.PT : *L.P;
      bytes += csr_bytes (Q) + (double) (Q.m + Q.n * (kind == MGC_PROLONG ? 2 : 1)) * vb;
      flops += 2.0 * (double) Q.nnz * v0.W;
    }
    else {
      bytes += csr_bytes (*L.A) + 3.0 * L.A->n * vb;
      flops += 2.0 * (double) L.A->nnz * v0.W;
    }
  }
  int lds = 0;
  for (int l = 0; l < nlev; l++) {
    const MgcLevelIn &L = lev[l];
    MgcLev &o = a.lev[l];
    o.n = L.B->nrows;
    o.ob = lds; lds += 16 * o.n;
    o.ox = lds; lds += 16 * o.n;
    o.orr = -1;
    if (l == 0) continue;
    o.orr = lds; lds += 16 * o.n;
    o.nc = lev[l - 1].B->nrows;
    o.sc = L.A->vs && L.A->vs_b1 == L.b1;
    o.rp = L.A->rp; o.ci = L.A->ci;
    o.v = L.A->v;
    o.vs = o.sc ? L.A->vs : L.A->v;
    o.dinv = L.A->dinv;
    o.prp = L.P->rp; o.pci = L.P->ci; o.pv = L.P->v;
    o.trp = L.PT->rp; o.tci = L.PT->ci; o.tv = L.PT->v;
    o.b1 = L.b1; o.a2 = L.a2; o.b2 = L.b2;
  }
  a.top = nlev - 1;
  a.Btop = lev[nlev - 1].B->p;
  a.Xtop = lev[nlev - 1].X->p;
  LGH_CHECK (lds <= MGC_MAX_LDS, 8, "k_mg_coarse: levels exceed the LDS bound");
  // the CSR arrays, 1/diag and the dense inverse after the slices, each at
  // a 16-byte boundary, when all of them fit: packed once (MgcImage) in
  // this layout, so a launch stages them with one contiguous copy
  bool staged = false;
  if (img && img_mu) {
    const int base = lds;
    int off = base;
    bool fit = true;
    std::vector<std::pair<const void *, std::pair<int, size_t>>> parts;
    auto put = [&] (const void *src, size_t bytes, int *o) {
      off = (off + 15) / 16 * 16;
      *o = off;
      if (off + (int64_t) bytes > MGC_MAX_LDS) { fit = false; return; }
      parts.push_back ({ src, { off - base, bytes } });
      off += (int) bytes;
    };
    const size_t tdb = v0.prec == F64 ? 8 : 4;
    put (cinv, (size_t) n0 * n0 * tdb, &a.ocinv);
    for (int l = 1; l < nlev && fit; l++) {
      const MgcLevelIn &L = lev[l];
      MgcLev &o = a.lev[l];
      const size_t vbs = prec_bytes (L.A->prec);
      put (L.A->rp, 4 * ((size_t) L.A->n + 1), &o.orp);
      put (L.A->ci, 4 * (size_t) L.A->nnz, &o.oci);
      put (L.A->v, vbs * (size_t) L.A->nnz, &o.ov);
      if (o.sc) put (L.A->vs, vbs * (size_t) L.A->nnz, &o.ovs);
      else o.ovs = o.ov;
      put (L.A->dinv, vbs * (size_t) L.A->n, &o.odinv);
      put (L.P->rp, 4 * ((size_t) L.P->n + 1), &o.oprp);
      put (L.P->ci, 4 * (size_t) L.P->nnz, &o.opci);
      put (L.P->v, vbs * (size_t) L.P->nnz, &o.opv);
      put (L.PT->rp, 4 * ((size_t) L.PT->n + 1), &o.otrp);
      put (L.PT->ci, 4 * (size_t) L.PT->nnz, &o.otci);
      put (L.PT->v, vbs * (size_t) L.PT->nnz, &o.otv);
    }
    off = (off + 15) / 16 * 16;
    if (fit) {
      const size_t ib = (size_t) (off - base);
      std::lock_guard<std::mutex> lk (*img_mu);
      if (img->nlev != nlev || img->bytes != ib) {
        img->release ();
        img->p = dmalloc (ib);
        if (!img->p) return 5;
        int rc = dzero (img->p, ib);
        for (auto &pt : parts)
          if (!rc) rc = dcopy ((char *) img->p + pt.second.first, pt.first, pt.second.second);
        // the image is read by other threads' streams: complete it here
        if (!rc) rc = dev_sync ();
        if (rc) { img->release (); return rc; }
        img->bytes = ib;
        img->nlev = nlev;
      }
      a.img = (const uint4 *) img->p;
      a.img_dst = base;
      a.img_n16 = (int) (ib / 16);
      lds = off;
      staged = true;
    }
  }
  const int np = v0.npanels ();
  // one thread per row of the largest level (whole wavefronts, at most
  // MGC_TPB): the small launch fits beside other streams' work
  int rows = 64;
  for (int l = 0; l < nlev; l++) rows = std::max (rows, a.lev[l].n);
  const int tpb = std::min (MGC_TPB, (rows + 63) / 64 * 64);
  ProfScope ps (PK_MG_COARSE, np * bytes, np * flops, np, v0.prec, v0.W);
  return with_row_W (v0.W, [&] (auto wc) -> int {
    constexpr int W = decltype (wc)::value;
    auto go = [&] (auto cl) {
      constexpr bool CL = decltype (cl)::value;
      if (v0.prec == F64)
        hipLaunchKernelGGL ((mgc_k<double, double, double, double, W, CL>),
                            dim3 (W / VT<double>::N, np), dim3 (tpb), lds,
                            stream (), a);
      else if (v0.prec == F32)
        hipLaunchKernelGGL ((mgc_k<float, float, float, float, W, CL>),
                            dim3 (W / VT<float>::N, np), dim3 (tpb), lds,
                            stream (), a);
      else
        hipLaunchKernelGGL ((mgc_k<float, float, bf16_t, float, W, CL>),
                            dim3 (W / VT<bf16_t>::N, np), dim3 (tpb), lds,
                            stream (), a);
    };
    if (staged) go (std::true_type ());
    else go (std::false_type ());
    LGH_LAUNCH_CHECK ();
    return 0;
  });
}

// ---------------------------------------------------------------------------
// elementwise

template <typename T, int W>
__global__ void __launch_bounds__ (TPB)
diag_scale_k (size_t tot, int nrows, const double *__restrict__ d,
              const T *X, T *Y)
{
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= tot) return;
  int i = (int) ((e / W) % nrows);
  Y[e] = (T) (d[i] * (double) X[e]);
}

int
k_diag_scale (const PMat &X, const double *d, PMat &Y)
{
  LGH_CHECK (X.W == Y.W && X.prec == Y.prec && X.elems () == Y.elems (), 8,
             "k_diag_scale: shape mismatch");
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      unsigned g = grid1 (X.elems (), TPB);
      if (X.prec == F64)
        hipLaunchKernelGGL ((diag_scale_k<double, W>), dim3 (g), dim3 (TPB),
                            0, stream (), X.elems (), X.nrows, d,
                            (const double *) X.p, (double *) Y.p);
      else
        hipLaunchKernelGGL ((diag_scale_k<float, W>), dim3 (g), dim3 (TPB), 0,
                            stream (), X.elems (), X.nrows, d,
                            (const float *) X.p, (float *) Y.p);
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  return with_W (X.W, l);
}

template <typename T>
__global__ void __launch_bounds__ (TPB)
axpby_k (size_t tot, T a, const T *X, T b, T *Y)
{
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= tot) return;
  Y[e] = (b == (T) 0) ? a * X[e] : a * X[e] + b * Y[e];
}

int
k_axpby (double a, const PMat &X, double b, PMat &Y)
{
  LGH_CHECK (X.prec == Y.prec && X.elems () == Y.elems (), 8,
             "k_axpby: shape mismatch");
  unsigned g = grid1 (X.elems (), TPB);
  if (X.prec == F64)
    hipLaunchKernelGGL (axpby_k<double>, dim3 (g), dim3 (TPB), 0, stream (),
                        X.elems (), a, (const double *) X.p, b,
                        (double *) Y.p);
  else
    hipLaunchKernelGGL (axpby_k<float>, dim3 (g), dim3 (TPB), 0, stream (),
                        X.elems (), (float) a, (const float *) X.p, (float) b,
                        (float *) Y.p);
  LGH_LAUNCH_CHECK ();
  return 0;
}

template <typename TI, typename TO>
__global__ void __launch_bounds__ (TPB)
copy_cols_k (int nrows, int nc, const TI *X, int xW, int xc0, TO *Y, int yW,
             int yc0)
{
  // consecutive threads take consecutive columns of one row (contiguous
  // within a panel row)
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= (size_t) nrows * nc) return;
  int i = (int) (e / nc);
  int j = (int) (e % nc);
  int xj = xc0 + j, yj = yc0 + j;
  size_t xi = ((size_t) (xj / xW) * nrows + i) * xW + xj % xW;
  size_t yi = ((size_t) (yj / yW) * nrows + i) * yW + yj % yW;
  st_as<double> (Y, yi, ld_as<double> (X, xi));
}

// whole panels of equal width: one contiguous range, 4 elements a thread
// (one 8-32 B vector load and store; panel starts are 4-element aligned)
template <typename TI, typename TO>
__global__ void __launch_bounds__ (TPB)
convert_k (size_t count, const TI *__restrict__ X, TO *__restrict__ Y)
{
  const size_t e = ((size_t) blockIdx.x * TPB + threadIdx.x) * 4;
  if (e >= count) return;
  if (e + 4 <= count) {
    double x[4];
    vld (X + e, x);
    vst (Y + e, x);
    return;
  }
  for (size_t i = e; i < count; i++) st_as<double> (Y, i, ld_as<double> (X, i));
}

int
k_copy_cols (const PMat &X, int xc0, PMat &Y, int yc0, int nc)
{
  LGH_CHECK (X.nrows == Y.nrows && xc0 + nc <= X.npanels () * X.W
             && yc0 + nc <= Y.npanels () * Y.W, 8, "k_copy_cols: range");
  if (nc <= 0) return 0;
  // same panel width, panel-aligned, and either whole panels or running to
  // the logical end of both (the rest of the last panel is zero padding in
  // both): a straight conversion of one contiguous range
  const bool flat = X.W == Y.W && xc0 % X.W == 0 && yc0 % Y.W == 0
                    && ((size_t) X.nrows * X.W) % 4 == 0      // aligned panels
                    && (nc % X.W == 0 || (xc0 + nc >= X.ncols && yc0 + nc >= Y.ncols));
  const size_t cnt = (size_t) ((nc + X.W - 1) / X.W) * X.nrows * X.W;
  const unsigned gf = grid1 ((cnt + 3) / 4, TPB);
  unsigned g = grid1 ((size_t) X.nrows * nc, TPB);
#define LGH_CC(TI, TO)                                                        \
  do {                                                                        \
    if (flat)                                                                 \
      hipLaunchKernelGGL ((convert_k<TI, TO>), dim3 (gf), dim3 (TPB), 0,      \
                          stream (), cnt,                                     \
                          (const TI *) X.panel (xc0 / X.W),                   \
                          (TO *) Y.panel (yc0 / Y.W));                        \
    else                                                                      \
      hipLaunchKernelGGL ((copy_cols_k<TI, TO>), dim3 (g), dim3 (TPB), 0,     \
                          stream (), X.nrows, nc, (const TI *) X.p, X.W, xc0, \
                          (TO *) Y.p, Y.W, yc0);                              \
  } while (0)
  auto same = [&] (int a, int b) { return X.prec == a && Y.prec == b; };
  if (same (F64, F64)) LGH_CC (double, double);
  else if (same (F64, F32)) LGH_CC (double, float);
  else if (same (F32, F64)) LGH_CC (float, double);
  else if (same (F32, F32)) LGH_CC (float, float);
  else if (same (F64, BF16)) LGH_CC (double, bf16_t);
  else if (same (F32, BF16)) LGH_CC (float, bf16_t);
  else if (same (BF16, F32)) LGH_CC (bf16_t, float);
  else if (same (BF16, F64)) LGH_CC (bf16_t, double);
  else LGH_CC (bf16_t, bf16_t);
#undef LGH_CC
  LGH_LAUNCH_CHECK ();
  return 0;
}

// ---------------------------------------------------------------------------
// column reductions (deterministic two-stage)

static constexpr int RED_ROWS = 2048;   // rows per partial

template <typename T, int W, int OP>
__global__ void __launch_bounds__ (TPB)
colred_k (int nrows, const T *X, const T *Y, const double *s, double *part,
          int npanels)
{
  // block: one (chunk, panel); threads stride rows; lane c = tid % W
  __shared__ double red[TPB];
  const int c = threadIdx.x % W;
  const int rsub = threadIdx.x / W;
  constexpr int RSTEP = TPB / W;
  const int p = blockIdx.y;
  const int r0 = blockIdx.x * RED_ROWS;
  const int r1 = min (nrows, r0 + RED_ROWS);
  const size_t off = (size_t) p * nrows * W + c;
  const double sj = (OP == 1) ? s[p * W + c] : 0.0;
  double acc = 0.0;
  for (int i = r0 + rsub; i < r1; i += RSTEP) {
    double x = (double) X[off + (size_t) i * W];
    double y = (double) Y[off + (size_t) i * W];
    if (OP == 0) acc += x * y;
    else { double d = x - sj * y; acc += d * d; }
  }
  red[threadIdx.x] = acc;
  __syncthreads ();
  if (threadIdx.x < W) {
    double t = 0.0;
    for (int q = 0; q < RSTEP; q++) t += red[q * W + threadIdx.x];
    part[((size_t) blockIdx.x * npanels + p) * W + threadIdx.x] = t;
  }
}

__global__ void __launch_bounds__ (TPB)
colred_fin_k (int nchunks, int ncols_pad, const double *part, double *out)
{
  int j = blockIdx.x * TPB + threadIdx.x;
  if (j >= ncols_pad) return;
  double t = 0.0;
  for (int q = 0; q < nchunks; q++) t += part[(size_t) q * ncols_pad + j];
  out[j] = t;
}

template <int OP>
static int
colred (const PMat &X, const PMat &Y, const double *s, double *out)
{
  LGH_CHECK (X.W == Y.W && X.prec == Y.prec && X.elems () == Y.elems (), 8,
             "column reduction: shape mismatch");
  const int nch = (X.nrows + RED_ROWS - 1) / RED_ROWS;
  const int ncp = X.npanels () * X.W;
  double *part = (double *) dmalloc (sizeof (double) * (size_t) nch * ncp);
  if (!part) return 5;
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      dim3 g (nch, X.npanels ());
      if (X.prec == F64)
        hipLaunchKernelGGL ((colred_k<double, W, OP>), g, dim3 (TPB), 0,
                            stream (), X.nrows, (const double *) X.p,
                            (const double *) Y.p, s, part, X.npanels ());
      else
        hipLaunchKernelGGL ((colred_k<float, W, OP>), g, dim3 (TPB), 0,
                            stream (), X.nrows, (const float *) X.p,
                            (const float *) Y.p, s, part, X.npanels ());
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  int rc = with_W (X.W, l);
  if (!rc) {
    hipLaunchKernelGGL (colred_fin_k, dim3 (grid1 (ncp, TPB)), dim3 (TPB), 0,
                        stream (), nch, ncp, part, out);
    hipError_t e = hipGetLastError ();
    if (e != hipSuccess) rc = hip_fail (e, "colred_fin_k", __FILE__, __LINE__);
  }
  dfree (part);
  return rc;
}

int k_coldot (const PMat &X, const PMat &Y, double *out)
{ return colred<0> (X, Y, nullptr, out); }

int k_colresid2 (const PMat &X, const PMat &Y, const double *s, double *out)
{ return colred<1> (X, Y, s, out); }

// per-column CG updates
template <typename T, int W>
__global__ void __launch_bounds__ (TPB)
cg_xr_k (size_t tot, int nrows, const double *alpha, const T *P, const T *Q,
         T *X, T *R)
{
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= tot) return;
  const int j = (int) (e / ((size_t) nrows * W)) * W + (int) (e % W);
  const T a = (T) alpha[j];
  X[e] += a * P[e];
  R[e] -= a * Q[e];
}

template <typename T, int W>
__global__ void __launch_bounds__ (TPB)
cg_p_k (size_t tot, int nrows, const double *beta, const T *Z, T *P)
{
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= tot) return;
  const int j = (int) (e / ((size_t) nrows * W)) * W + (int) (e % W);
  P[e] = Z[e] + (T) beta[j] * P[e];
}

int
k_cg_xr (const double *alpha, const PMat &P, const PMat &Q, PMat &X, PMat &R)
{
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      unsigned g = grid1 (X.elems (), TPB);
      if (X.prec == F64)
        hipLaunchKernelGGL ((cg_xr_k<double, W>), dim3 (g), dim3 (TPB), 0,
                            stream (), X.elems (), X.nrows, alpha,
                            (const double *) P.p, (const double *) Q.p,
                            (double *) X.p, (double *) R.p);
      else
        hipLaunchKernelGGL ((cg_xr_k<float, W>), dim3 (g), dim3 (TPB), 0,
                            stream (), X.elems (), X.nrows, alpha,
                            (const float *) P.p, (const float *) Q.p,
                            (float *) X.p, (float *) R.p);
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  return with_W (X.W, l);
}

int
k_cg_p (const double *beta, const PMat &Z, PMat &P)
{
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      unsigned g = grid1 (P.elems (), TPB);
      if (P.prec == F64)
        hipLaunchKernelGGL ((cg_p_k<double, W>), dim3 (g), dim3 (TPB), 0,
                            stream (), P.elems (), P.nrows, beta,
                            (const double *) Z.p, (double *) P.p);
      else
        hipLaunchKernelGGL ((cg_p_k<float, W>), dim3 (g), dim3 (TPB), 0,
                            stream (), P.elems (), P.nrows, beta,
                            (const float *) Z.p, (float *) P.p);
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  return with_W (P.W, l);
}

// Y(:, j) = s_j X(:, j)
template <typename T, int W>
__global__ void __launch_bounds__ (TPB)
colscale_k (size_t tot, int nrows, const double *s, const T *X, T *Y)
{
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= tot) return;
  const int j = (int) (e / ((size_t) nrows * W)) * W + (int) (e % W);
  Y[e] = (T) (s[j] * (double) X[e]);
}

int
k_colscale (const double *s, const PMat &X, PMat &Y)
{
  LGH_CHECK (X.W == Y.W && X.prec == Y.prec && X.elems () == Y.elems (), 8,
             "k_colscale: shape mismatch");
  auto l = [&] (auto wt) -> int {
    constexpr int W = decltype (wt)::value;
    unsigned g = grid1 (X.elems (), TPB);
    if (X.prec == F64)
      hipLaunchKernelGGL ((colscale_k<double, W>), dim3 (g), dim3 (TPB), 0,
                          stream (), X.elems (), X.nrows, s,
                          (const double *) X.p, (double *) Y.p);
    else
      hipLaunchKernelGGL ((colscale_k<float, W>), dim3 (g), dim3 (TPB), 0,
                          stream (), X.elems (), X.nrows, s,
                          (const float *) X.p, (float *) Y.p);
    LGH_LAUNCH_CHECK ();
    return 0;
  };
  return with_W (X.W, l);
}

// ---------------------------------------------------------------------------
// MFMA tall-skinny products (W = 64 panels)

template <typename T> struct mfma;
// Result layouts (probed on gfx950): f64 16x16x4 puts D(lane/16 + 4 r,
// lane%16) in element r; f32 16x16x4 puts D(4 (lane/16) + r, lane%16).
template <> struct mfma<double> {
  typedef d4 acc_t;
  static __device__ inline acc_t op (double a, double b, acc_t c)
  { return __builtin_amdgcn_mfma_f64_16x16x4f64 (a, b, c, 0, 0, 0); }
  static __device__ inline int row (int lk, int r) { return lk + 4 * r; }
};
template <> struct mfma<float> {
  typedef f4 acc_t;
  static __device__ inline acc_t op (float a, float b, acc_t c)
  { return __builtin_amdgcn_mfma_f32_16x16x4f32 (a, b, c, 0, 0, 0); }
  static __device__ inline int row (int lk, int r) { return 4 * lk + r; }
};

// Blocks are numbered so that the blocks sharing an input tile land on one
// XCD back to back: the dispatcher hands block b to XCD b % 8, so block b
// takes group (b / 8) / nsib * 8 + b % 8 and sibling (b / 8) % nsib.  The
// first sibling reads the tile from HBM; the others find it in that XCD's L2.
__device__ inline void
xcd_split (int nsib, int &group, int &sib)
{
  const int b = blockIdx.x, s = b >> 3;
  group = (s / nsib) * 8 + (b & 7);
  sib = s % nsib;
}
static inline unsigned
xcd_blocks (int ngroups, int nsib)
{
  return (unsigned) (((ngroups + 7) / 8) * 8) * (unsigned) nsib;
}

// Gram partials: block (chunk, group) computes the four 64 x 64 tiles
// X_px^T Y_py of a gx x gy tile group (2 x 2; 4 x 1 or 1 x 4 when Y or X
// is a single panel), px in gx gp .. gx gp + gx - 1, py in gy gq .. gy gq
// + gy - 1, over rows [r0, r1) (the groups of one chunk are siblings).
// Wave w owns tile (gx gp + w / gy, gy gq + w % gy) as 4 x 4 MFMA 16 x 16
// tiles: per k-step of 4 rows it loads 4 A and 4 B fragments and issues 16
// MFMAs, and each panel row it reads is shared with the other waves of the
// block on the same panel (L1), so a panel chunk is read once per tile
// group instead of once per tile.
// Tiles past the panels, and in the symmetric case below the diagonal,
// are skipped; tile_of maps (px, py) to the pair slot of the reduction.
// MFMA 16x16x4 operand layout: A(i, k) at lane i + 16 k, B(k, j) at lane
// j + 16 k; result rows per mfma<T>::row.
template <typename T>
__global__ void __launch_bounds__ (TPB)
gram_k (int n, int nchunks, int ngroups, const T *__restrict__ X, int xp0,
        int npx, const T *__restrict__ Y, int yp0, int npy, int rows_per_chunk,
        const int2 *__restrict__ groups, const int *__restrict__ tile_of,
        int npairs, double *__restrict__ part, int gy)
{
  typedef typename mfma<T>::acc_t acc_t;
  int chunk, grp;
  xcd_split (ngroups, chunk, grp);
  if (chunk >= nchunks) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int2 gr = groups[grp];
  // a group is gx x gy tiles (gx gy = 4): 2 x 2, or 4 x 1 / 1 x 4 when one
  // side is a single panel (no wave left idle)
  const int gx = 4 / gy;
  const int px = gx * gr.x + wave / gy, py = gy * gr.y + wave % gy;
  if (px >= npx || py >= npy) return;
  const int slot = tile_of[px * npy + py];
  if (slot < 0) return;
  const T *Xp = X + (size_t) (xp0 + px) * n * 64;
  const T *Yp = Y + (size_t) (yp0 + py) * n * 64;
  const int r0 = chunk * rows_per_chunk;
  const int r1 = min (n, r0 + rows_per_chunk);
  const int li = lane & 15, lk = lane >> 4;
  acc_t acc[4][4];
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++)
      for (int r = 0; r < 4; r++) acc[a][b][r] = 0;
  // 4 k-steps (16 rows) per trip: all loads issued before the MFMAs (a
  // variant that issued the next trip's loads before this trip's MFMAs,
  // two trips of 2 k-steps in registers, ran at 34.5 against 42.8 TF/s).
  // FP64: a lane loads 16 bytes (columns 2 li, 2 li + 1 of each 32-column
  // half), so MFMA block q = 2 h + p holds the columns 32 h + 2 i + p
  // (i = 0..15); FP32: 4 bytes (block q holds columns 16 q + i).  The
  // epilogue maps the blocks back (col_of).
  constexpr bool W16 = sizeof (T) == 8;
  auto col_of = [&] (int q, int i) { return W16 ? 32 * (q >> 1) + 2 * i + (q & 1) : 16 * q + i; };
  constexpr int KU = 4;
  for (int k0 = r0; k0 < r1; k0 += 4 * KU) {
    T av[KU][4], bv[KU][4];
#pragma unroll
    for (int u = 0; u < KU; u++) {
      const int kr = k0 + 4 * u + lk;
      if (kr < r1) {
        const T *xr = Xp + (size_t) kr * 64;
        const T *yr = Yp + (size_t) kr * 64;
        if constexpr (W16) {
#pragma unroll
          for (int h = 0; h < 2; h++) {
            const double2 xa = *(const double2 *) (xr + 32 * h + 2 * li);
            const double2 ya = *(const double2 *) (yr + 32 * h + 2 * li);
            av[u][2 * h] = xa.x; av[u][2 * h + 1] = xa.y;
            bv[u][2 * h] = ya.x; bv[u][2 * h + 1] = ya.y;
          }
        }
        else {
#pragma unroll
          for (int q = 0; q < 4; q++) { av[u][q] = xr[16 * q + li]; bv[u][q] = yr[16 * q + li]; }
        }
      }
      else {
#pragma unroll
        for (int q = 0; q < 4; q++) { av[u][q] = (T) 0; bv[u][q] = (T) 0; }
      }
    }
#pragma unroll
    for (int u = 0; u < KU; u++)
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          acc[a][b] = mfma<T>::op (av[u][a], bv[u][b], acc[a][b]);
  }
  double *tile = part + ((size_t) chunk * npairs + slot) * 4096;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++)
      for (int r = 0; r < 4; r++) {
        const int row = col_of (a, mfma<T>::row (lk, r));
        const int col = col_of (b, li);
        tile[col * 64 + row] = (double) acc[a][b][r];
      }
}

// fixed-order reduction of the chunk partials: block (slice, pair) sums 16
// consecutive tile entries; thread (g, e) adds chunks g, g + 16, g + 32, ...
// in order, then thread g = 0 adds the 16 group sums in order (the result
// depends on the chunk count only, not on the launch)
__global__ void __launch_bounds__ (TPB)
gram_reduce_k (int nchunks, int npairs, const int2 *__restrict__ pairs,
               const double *__restrict__ part, double *G, int ldg, int sym)
{
  __shared__ double red[16][16];
  const int pair = blockIdx.y;
  const int el = threadIdx.x & 15, g = threadIdx.x >> 4;
  const int e = blockIdx.x * 16 + el;                 // 0..4095
  double t = 0.0;
  int q = g;
  for (; q + 48 < nchunks; q += 64) {
    const double a0 = part[((size_t) q * npairs + pair) * 4096 + e];
    const double a1 = part[((size_t) (q + 16) * npairs + pair) * 4096 + e];
    const double a2 = part[((size_t) (q + 32) * npairs + pair) * 4096 + e];
    const double a3 = part[((size_t) (q + 48) * npairs + pair) * 4096 + e];
    t += a0; t += a1; t += a2; t += a3;
  }
  for (; q < nchunks; q += 16)
    t += part[((size_t) q * npairs + pair) * 4096 + e];
  red[g][el] = t;
  __syncthreads ();
  if (g != 0) return;
  double v = red[0][el];
  for (int h = 1; h < 16; h++) v += red[h][el];
  const int2 pr = pairs[pair];
  const int row = e % 64, col = e / 64;
  G[(size_t) (pr.x * 64 + row) + (size_t) (pr.y * 64 + col) * ldg] = v;
  if (sym && pr.x != pr.y)
    G[(size_t) (pr.y * 64 + col) + (size_t) (pr.x * 64 + row) * ldg] = v;
}

int
k_gram (const PMat &X, int xp0, int npx, const PMat &Y, int yp0, int npy,
        double *G, int ldg, int sym)
{
  LGH_CHECK (X.W == 64 && Y.W == 64 && X.prec == Y.prec
             && X.nrows == Y.nrows && xp0 + npx <= X.npanels ()
             && yp0 + npy <= Y.npanels (), 8, "k_gram: shape mismatch");
  // tiles (pairs) in the reduction's order, and their 2 x 2 groups
  std::vector<int2> pv;
  std::vector<int> tile_of ((size_t) std::max (npx * npy, 1), -1);
  for (int i = 0; i < npx; i++)
    for (int j = (sym ? i : 0); j < npy; j++) {
      tile_of[(size_t) i * npy + j] = (int) pv.size ();
      pv.push_back (make_int2 (i, j));
    }
  const int npairs = (int) pv.size ();
  if (npairs == 0) return 0;
  // group shape: 2 x 2 tiles, or a column / row of 4 when one side has a
  // single panel (a symmetric Gram has npx == npy)
  const int gy = (npx == 1 && !sym) ? 4 : (npy == 1 && !sym) ? 1 : 2;
  const int gx = 4 / gy;
  std::vector<int2> gv;
  for (int a = 0; a < (npx + gx - 1) / gx; a++)
    for (int b = (sym ? a : 0); b < (npy + gy - 1) / gy; b++) gv.push_back (make_int2 (a, b));
  const int ngroups = (int) gv.size ();
  const int n = X.nrows;
  ProfScope ps (PK_GRAM, (double) n * 64 * prec_bytes (X.prec) * (npx + npy),
                2.0 * n * 4096.0 * npairs);
  // aim for ~4096 blocks; at least 256 rows per chunk, and at least 8
  // chunks: chunk c of every group runs on XCD c % 8 (xcd_split), so fewer
  // chunks would leave XCDs idle when there are many groups (wide blocks)
  int nch = 4096 / ngroups;
  if (nch < 8) nch = 8;
  int rpc = (n + nch - 1) / nch;
  if (rpc < 256) rpc = 256;
  rpc = (rpc + 3) / 4 * 4;
  nch = (n + rpc - 1) / rpc;
  if (nch < 1) nch = 1;
  int2 *pd = (int2 *) dmalloc (sizeof (int2) * (npairs + ngroups));
  int *td = (int *) dmalloc (sizeof (int) * tile_of.size ());
  double *part = (double *) dmalloc (sizeof (double) * 4096 * (size_t) nch * npairs);
  if (!pd || !td || !part) { dfree (pd); dfree (td); dfree (part); return 5; }
  int rc = h2d (pd, pv.data (), sizeof (int2) * npairs);
  if (!rc) rc = h2d (pd + npairs, gv.data (), sizeof (int2) * ngroups);
  if (!rc) rc = h2d (td, tile_of.data (), sizeof (int) * tile_of.size ());
  if (!rc) {
    const unsigned g = xcd_blocks (nch, ngroups);
    if (X.prec == F64)
      hipLaunchKernelGGL (gram_k<double>, dim3 (g), dim3 (TPB), 0, stream (),
                          n, nch, ngroups, (const double *) X.p, xp0, npx,
                          (const double *) Y.p, yp0, npy, rpc, pd + npairs, td,
                          npairs, part, gy);
    else
      hipLaunchKernelGGL (gram_k<float>, dim3 (g), dim3 (TPB), 0, stream (),
                          n, nch, ngroups, (const float *) X.p, xp0, npx,
                          (const float *) Y.p, yp0, npy, rpc, pd + npairs, td,
                          npairs, part, gy);
    hipError_t e = hipGetLastError ();
    if (e != hipSuccess) rc = hip_fail (e, "gram_k", __FILE__, __LINE__);
  }
  if (!rc) {
    hipLaunchKernelGGL (gram_reduce_k, dim3 (4096 / 16, npairs), dim3 (TPB),
                        0, stream (), nch, npairs, pd, part, G, ldg, sym);
    hipError_t e = hipGetLastError ();
    if (e != hipSuccess) rc = hip_fail (e, "gram_reduce_k", __FILE__, __LINE__);
  }
  dfree (pd);
  dfree (td);
  dfree (part);
  return rc;
}

// Y panel block = alpha X C + beta Y.  Block (row tile, out panel q) stages
// the 64 x 64 tile of each input panel and the matching 64 x 64 block of C
// in LDS; wave w owns output rows 16 w .. 16 w + 15 (4 MFMA tiles across).
// The out panels of one row tile are siblings (one XCD, back to back), and
// the next input panel's tile and C block are loaded into registers while
// the MFMAs of the current one run.  upper: C is block upper triangular
// (C[p, q] = 0 for p > q), so out panel q reads input panels 0..q only.
static constexpr int GPAD = 65;

template <typename T>
__global__ void __launch_bounds__ (TPB)
gemm_nn_k (int n, int ntiles, int npy, const T *__restrict__ X, int xp0,
           int npx, const double *__restrict__ C, int ldc, T alpha, T beta,
           T *__restrict__ Y, int yp0, int upper)
{
  typedef typename mfma<T>::acc_t acc_t;
  __shared__ T Xs[64 * GPAD];
  __shared__ T Cs[64 * GPAD];
  int tile, q;
  xcd_split (npy, tile, q);
  if (tile >= ntiles) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = tile * 64;
  const int pend = upper ? min (npx, q + 1) : npx;
  // element e = threadIdx.x + 256 i of a 64 x 64 tile: X row rr = e / 64,
  // col cc = e % 64; C row kk = e % 64, col jj = e / 64
  const int e64 = (int) threadIdx.x & 63, e6 = (int) threadIdx.x >> 6;
  constexpr int EPT = 4096 / TPB;
  T xr[EPT], cr[EPT];
  auto gload = [&] (int p) {
    const T *Xp = X + (size_t) (xp0 + p) * n * 64;
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int rr = e6 + (TPB / 64) * i;
      xr[i] = (r0 + rr < n) ? Xp[(size_t) (r0 + rr) * 64 + e64] : (T) 0;
      cr[i] = (T) C[(size_t) (p * 64 + e64) + (size_t) (q * 64 + rr) * ldc];
    }
  };
  acc_t acc[4];
  for (int b = 0; b < 4; b++)
    for (int r = 0; r < 4; r++) acc[b][r] = 0;
  if (pend > 0) gload (0);
  for (int p = 0; p < pend; p++) {
    __syncthreads ();
#pragma unroll
    for (int i = 0; i < EPT; i++) {
      const int rr = e6 + (TPB / 64) * i;
      Xs[rr * GPAD + e64] = xr[i];
      Cs[e64 * GPAD + rr] = cr[i];
    }
    __syncthreads ();
    if (p + 1 < pend) gload (p + 1);
    for (int k0 = 0; k0 < 64; k0 += 4) {
      const T a = Xs[(16 * wave + li) * GPAD + k0 + lk];
      for (int b = 0; b < 4; b++) {
        const T bv = Cs[(k0 + lk) * GPAD + 16 * b + li];
        acc[b] = mfma<T>::op (a, bv, acc[b]);
      }
    }
  }
  T *Yq = Y + (size_t) (yp0 + q) * n * 64;
  for (int b = 0; b < 4; b++)
    for (int r = 0; r < 4; r++) {
      const int row = r0 + 16 * wave + mfma<T>::row (lk, r);
      const int col = 16 * b + li;
      if (row < n) {
        T *y = Yq + (size_t) row * 64 + col;
        T v = alpha * (T) acc[b][r];
        *y = (beta == (T) 0) ? v : v + beta * *y;
      }
    }
}

int
k_gemm_nn (const PMat &X, int xp0, int npx, const double *C, int ldc,
           double alpha, double beta, PMat &Y, int yp0, int npy, int upper)
{
  LGH_CHECK (X.W == 64 && Y.W == 64 && X.prec == Y.prec
             && X.nrows == Y.nrows && xp0 + npx <= X.npanels ()
             && yp0 + npy <= Y.npanels (), 8, "k_gemm_nn: shape mismatch");
  // output panels may overlap input panels only in the one-panel in-place
  // form: the block reads its whole input tile before it writes
  const bool overlap = X.p == Y.p && yp0 < xp0 + npx && xp0 < yp0 + npy;
  LGH_CHECK (!overlap || (npx == 1 && npy == 1 && xp0 == yp0), 8,
             "k_gemm_nn: output panels overlap the input panels");
  if (npy == 0) return 0;
  double pairs = 0;
  for (int q = 0; q < npy; q++) pairs += upper ? std::min (npx, q + 1) : npx;
  ProfScope ps (PK_GEMM, (double) X.nrows * 64 * prec_bytes (X.prec)
                * (npx + npy * (beta != 0.0 ? 2 : 1)),
                2.0 * X.nrows * 4096.0 * pairs);
  const int ntiles = (X.nrows + 63) / 64;
  const dim3 g (xcd_blocks (ntiles, npy));
  if (X.prec == F64)
    hipLaunchKernelGGL (gemm_nn_k<double>, g, dim3 (TPB), 0, stream (),
                        X.nrows, ntiles, npy, (const double *) X.p, xp0, npx,
                        C, ldc, alpha, beta, (double *) Y.p, yp0, upper);
  else
    hipLaunchKernelGGL (gemm_nn_k<float>, g, dim3 (TPB), 0, stream (),
                        X.nrows, ntiles, npy, (const float *) X.p, xp0, npx,
                        C, ldc, (float) alpha, (float) beta, (float *) Y.p,
                        yp0, upper);
  LGH_LAUNCH_CHECK ();
  return 0;
}

// ---------------------------------------------------------------------------
// small dense helpers (fp64, column-major)

__global__ void
symmetrize_k (double *T, int n, int ld)
{
  int e = blockIdx.x * TPB + threadIdx.x;
  if (e >= n * n) return;
  int i = e % n, j = e / n;
  if (i < j) {
    double avg = 0.5 * (T[i + (size_t) j * ld] + T[j + (size_t) i * ld]);
    T[i + (size_t) j * ld] = avg;
    T[j + (size_t) i * ld] = avg;
  }
}

int
k_symmetrize (double *T, int n, int ld)
{
  hipLaunchKernelGGL (symmetrize_k, dim3 (grid1 ((size_t) n * n, TPB)),
                      dim3 (TPB), 0, stream (), T, n, ld);
  LGH_LAUNCH_CHECK ();
  return 0;
}

__global__ void
add_diag_k (double *A, int n, int ld, const double *s)
{
  int i = blockIdx.x * TPB + threadIdx.x;
  if (i < n) A[i + (size_t) i * ld] += *s;
}

int
k_add_diag (double *A, int n, int ld, const double *s)
{
  hipLaunchKernelGGL (add_diag_k, dim3 (grid1 (n, TPB)), dim3 (TPB), 0,
                      stream (), A, n, ld, s);
  LGH_LAUNCH_CHECK ();
  return 0;
}

__global__ void
lower_t_pad_k (const double *L, int n, int ldl, double *C, int ldc, int cc)
{
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= (size_t) ldc * cc) return;
  int i = (int) (e % ldc), j = (int) (e / ldc);
  // C(i, j) = Linv(j, i) for j >= i (transpose of the lower triangle)
  C[e] = (i < n && j < n && j >= i) ? L[j + (size_t) i * ldl] : 0.0;
}

int
k_lower_t_pad (const double *Linv, int n, int ldl, double *C, int ldc,
               int ccols)
{
  hipLaunchKernelGGL (lower_t_pad_k, dim3 (grid1 ((size_t) ldc * ccols, TPB)),
                      dim3 (TPB), 0, stream (), Linv, n, ldl, C, ldc, ccols);
  LGH_LAUNCH_CHECK ();
  return 0;
}

__global__ void
upper_pad_k (const double *R, int n, int ldr, double *C, int ldc, int cc)
{
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= (size_t) ldc * cc) return;
  int i = (int) (e % ldc), j = (int) (e / ldc);
  C[e] = (i < n && j < n && i <= j) ? R[i + (size_t) j * ldr] : 0.0;
}

int
k_upper_pad (const double *Rinv, int n, int ldr, double *C, int ldc,
             int ccols)
{
  hipLaunchKernelGGL (upper_pad_k, dim3 (grid1 ((size_t) ldc * ccols, TPB)),
                      dim3 (TPB), 0, stream (), Rinv, n, ldr, C, ldc, ccols);
  LGH_LAUNCH_CHECK ();
  return 0;
}

__global__ void
trace_k (const double *A, int n, int ld, double *out)
{
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double t = 0.0;
    for (int i = 0; i < n; i++) t += A[i + (size_t) i * ld];
    *out = t;
  }
}

int
k_trace (const double *A, int n, int ld, double *out_dev)
{
  hipLaunchKernelGGL (trace_k, dim3 (1), dim3 (64), 0, stream (), A, n, ld,
                      out_dev);
  LGH_LAUNCH_CHECK ();
  return 0;
}

}  // namespace lgh
