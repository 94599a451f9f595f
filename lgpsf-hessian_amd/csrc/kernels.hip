// kernels.hip — gfx950 kernels of the GLR engine and their launchers.
//
// Row kernels (SpMM, fused Chebyshev step, residual) map one wavefront to
// 64/W rows of one panel: lane (r = lane / W, c = lane % W) owns column c of
// row r.  At W = 64 a wave owns a whole row, so each CSR neighbour costs one
// coalesced 64-lane load of that neighbour's row.  The tall-skinny products
// use MFMA (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32) on 64 x 64
// output tiles with split-K partials reduced in a fixed order
// (deterministic results).
#include "lgh_dev.hpp"
#include "randn.hpp"

#include <type_traits>

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

template <typename T, int W>
__global__ void __launch_bounds__ (TPB)
randn_k (T *out, int nrows, int ncols, int npanels, uint64_t seed,
         int64_t row0, int64_t col0)
{
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  size_t tot = (size_t) npanels * nrows * W;
  if (e >= tot) return;
  int c = (int) (e % W);
  size_t rest = e / W;
  int i = (int) (rest % nrows);
  int p = (int) (rest / nrows);
  int j = p * W + c;
  out[e] = (j < ncols) ? (T) randn_at (seed, row0 + i, col0 + j) : (T) 0;
}

int
k_randn (PMat &out, uint64_t seed, int64_t row0, int64_t col0)
{
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      unsigned g = grid1 (out.elems (), TPB);
      if (out.prec == F64)
        hipLaunchKernelGGL ((randn_k<double, W>), dim3 (g), dim3 (TPB), 0,
                            stream (), (double *) out.p, out.nrows, out.ncols,
                            out.npanels (), seed, row0, col0);
      else
        hipLaunchKernelGGL ((randn_k<float, W>), dim3 (g), dim3 (TPB), 0,
                            stream (), (float *) out.p, out.nrows, out.ncols,
                            out.npanels (), seed, row0, col0);
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  return with_W (out.W, l);
}


// ---------------------------------------------------------------------------
// W = 64 row kernels: one wavefront per row.  The row's CSR entries are
// fetched 64 at a time with one coalesced load per array (lane k holds
// entry k) and broadcast with v_readlane, so the neighbour gathers of a
// row are independent loads that can all be in flight together.

__device__ inline int bcast (int v, int e) { return __builtin_amdgcn_readlane (v, e); }
__device__ inline float bcast (float v, int e)
{ return __int_as_float (__builtin_amdgcn_readlane (__float_as_int (v), e)); }
__device__ inline double bcast (double v, int e)
{
  const long long b = __double_as_longlong (v);
  const int lo = __builtin_amdgcn_readlane ((int) (b & 0xffffffffLL), e);
  const int hi = __builtin_amdgcn_readlane ((int) (b >> 32), e);
  return __longlong_as_double (((long long) hi << 32) | (unsigned int) lo);
}

__device__ inline int uniform_row (int per_block_row)
{
  return __builtin_amdgcn_readfirstlane (per_block_row);
}

// XCD-aware row tiling.  Work-groups are dispatched round-robin over the 8
// XCDs (each with its own L2); remapping block b to tile
// (b % 8) * (grid / 8) + b / 8 gives every XCD one contiguous range of
// rows, so the CSR neighbours of a row (the rows just before and after it)
// are found in that XCD's L2.  Each block owns ROWS64 consecutive rows;
// wave w takes rows w, w + 4, ... of the tile.  The grid is padded to a
// multiple of 8 blocks.
static constexpr int ROWS64 = 64;

__device__ inline int xcd_tile ()
{
  const int nb = gridDim.x;              // multiple of 8
  const int b = blockIdx.x;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

static inline unsigned
grid_rows64 (int n)
{
  unsigned g = (unsigned) ((n + ROWS64 - 1) / ROWS64);
  return (g + 7) / 8 * 8;
}

#define LGH_ROW_BEGIN(n)                                                       \
  {                                                                            \
    const int rbase_ = xcd_tile () * ROWS64 + (threadIdx.x >> 6);              \
    for (int t_ = 0; t_ < ROWS64 / (TPB / 64); t_++) {                         \
      const int row = uniform_row (rbase_ + t_ * (TPB / 64));                  \
      if (row >= (n)) break;
#define LGH_ROW_END                                                            \
    }                                                                          \
  }

// Y[p] = alpha A X[p] + beta Y[p] for NP panels per block (B read once)
template <typename TV, typename T, int NP>
__global__ void __launch_bounds__ (TPB)
spmm64_k (int n, int nx, const int *__restrict__ rp,
          const int *__restrict__ ci, const TV *__restrict__ v, T alpha,
          const T *__restrict__ X, T beta, T *__restrict__ Y, int npanels)
{
  const int lane = threadIdx.x & 63;
  const int p0 = blockIdx.y * NP;
  const int nph = min (NP, npanels - p0);
  const size_t xs = (size_t) nx * 64;
  const T *Xb = X + (size_t) p0 * xs + lane;
  LGH_ROW_BEGIN (n)
  T acc[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) acc[p] = 0;
  const int k0 = rp[row], k1 = rp[row + 1];
  for (int kb = k0; kb < k1; kb += 64) {
    const int kk = kb + lane;
    int c = 0;
    TV w = 0;
    if (kk < k1) { c = ci[kk]; w = v[kk]; }
    const int cnt = min (64, k1 - kb);
    int e = 0;
    for (; e + 4 <= cnt; e += 4) {
      const int c0 = bcast (c, e), c1 = bcast (c, e + 1);
      const int c2 = bcast (c, e + 2), c3 = bcast (c, e + 3);
      const T w0 = (T) bcast (w, e), w1 = (T) bcast (w, e + 1);
      const T w2 = (T) bcast (w, e + 2), w3 = (T) bcast (w, e + 3);
#pragma unroll
      for (int p = 0; p < NP; p++)
        if (p < nph) {
          const T *xp = Xb + p * xs;
          const T x0 = xp[(size_t) c0 * 64], x1 = xp[(size_t) c1 * 64];
          const T x2 = xp[(size_t) c2 * 64], x3 = xp[(size_t) c3 * 64];
          acc[p] += w0 * x0;
          acc[p] += w1 * x1;
          acc[p] += w2 * x2;
          acc[p] += w3 * x3;
        }
    }
    for (; e < cnt; e++) {
      const int c0 = bcast (c, e);
      const T w0 = (T) bcast (w, e);
#pragma unroll
      for (int p = 0; p < NP; p++)
        if (p < nph) acc[p] += w0 * Xb[p * xs + (size_t) c0 * 64];
    }
  }
#pragma unroll
  for (int p = 0; p < NP; p++)
    if (p < nph) {
      T *y = Y + ((size_t) (p0 + p) * n + row) * 64 + lane;
      *y = (beta == (T) 0) ? alpha * acc[p] : alpha * acc[p] + beta * *y;
    }
  LGH_ROW_END
}

// ---------------------------------------------------------------------------
// W = 64 row kernels over NP panels at once: a wavefront owns one row and
// NP panels of it, so the row's CSR entries (and dinv at the neighbours)
// are fetched once for all NP panels and the NP neighbour gathers of each
// entry are independent loads in flight together.

template <typename F>
static int
np_dispatch (int np, F &&f)
{
  if (np >= 5) return f (IC<8> ());
  if (np >= 3) return f (IC<4> ());
  if (np == 2) return f (IC<2> ());
  return f (IC<1> ());
}

template <typename TV, bool NEEDD, typename F>
__device__ __forceinline__ void
for_nbrs (int row, const int *__restrict__ rp, const int *__restrict__ ci,
          const TV *__restrict__ v, const TV *__restrict__ dinv, int lane,
          F &&f)
{
  const int k0 = rp[row], k1 = rp[row + 1];
  for (int kb = k0; kb < k1; kb += 64) {
    const int kk = kb + lane;
    int c = 0;
    TV w = 0, dd = 0;
    if (kk < k1) {
      c = ci[kk];
      w = v[kk];
      if (NEEDD) dd = dinv[c];
    }
    const int cnt = min (64, k1 - kb);
    for (int e = 0; e < cnt; e++)
      f (bcast (c, e), bcast (w, e), NEEDD ? bcast (dd, e) : (TV) 0);
  }
}

#define LGH_PANELS(NP)                                                        \
  const int lane = threadIdx.x & 63;                                          \
  const int p0 = blockIdx.y * (NP);                                           \
  const int nph = min ((NP), npanels - p0);                                   \
  const size_t ps = (size_t) n * 64;                                          \
  const size_t base = (size_t) p0 * ps + lane;

// R = B - A X
template <typename TV, typename T, int NP>
__global__ void __launch_bounds__ (TPB)
resid64_k (int n, int npanels, const int *__restrict__ rp,
           const int *__restrict__ ci, const TV *__restrict__ v,
           const T *__restrict__ B, const T *__restrict__ X,
           T *__restrict__ R)
{
  LGH_PANELS (NP)
  LGH_ROW_BEGIN (n)
  T acc[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) acc[p] = 0;
  for_nbrs<TV, false> (row, rp, ci, v, (const TV *) nullptr, lane,
                       [&] (int c, TV w, TV) {
#pragma unroll
    for (int p = 0; p < NP; p++)
      if (p < nph) acc[p] += (T) w * X[base + p * ps + (size_t) c * 64];
  });
#pragma unroll
  for (int p = 0; p < NP; p++)
    if (p < nph) {
      const size_t q = base + p * ps + (size_t) row * 64;
      R[q] = B[q] - acc[p];
    }
  LGH_ROW_END
}

// fused Chebyshev step (zsolve_impl.h:104-141): Dn = a D + b Z with
// Z = dinv R (JAC) or Zs; X (+)= Dn; Rn = R - A Dn (Dn recomputed at the
// neighbours)
template <typename TV, typename T, int NP, bool JAC, bool HASD>
__global__ void __launch_bounds__ (TPB)
cheb64_k (int n, int npanels, const int *__restrict__ rp,
          const int *__restrict__ ci, const TV *__restrict__ v,
          const TV *__restrict__ dinv, const T *__restrict__ Zs, T a, T b,
          const T *__restrict__ D, const T *__restrict__ R,
          T *__restrict__ X, int xmode, T *__restrict__ Dn,
          T *__restrict__ Rn)
{
  LGH_PANELS (NP)
  auto dnew = [&] (size_t q, T dj) -> T {
    T z = JAC ? dj * R[q] : Zs[q];
    T bz = b * z;
    if (HASD) {
      T ad = a * D[q];
      return ad + bz;
    }
    return bz;
  };
  LGH_ROW_BEGIN (n)
  T acc[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) acc[p] = 0;
  for_nbrs<TV, JAC> (row, rp, ci, v, dinv, lane, [&] (int c, TV w, TV dd) {
#pragma unroll
    for (int p = 0; p < NP; p++)
      if (p < nph) acc[p] += (T) w * dnew (base + p * ps + (size_t) c * 64, (T) dd);
  });
  const T dr = JAC ? (T) dinv[row] : (T) 0;
#pragma unroll
  for (int p = 0; p < NP; p++)
    if (p < nph) {
      const size_t qi = base + p * ps + (size_t) row * 64;
      const T di = dnew (qi, dr);
      Dn[qi] = di;
      X[qi] = xmode ? X[qi] + di : di;
      Rn[qi] = R[qi] - acc[p];
    }
  LGH_ROW_END
}

// mixed-precision fused Chebyshev step: FP64 residual R and accumulator X,
// FP32 direction D and preconditioned residual Z.  The new direction is
// rounded to FP32 once and that rounded value is what X and R both see,
// so R stays the FP64 residual of X and the iteration reaches FP64
// accuracy with an FP32 preconditioner.  Also writes R in FP32 for the
// next preconditioner application.
template <int NP, bool HASD>
__global__ void __launch_bounds__ (TPB)
cheb64m_k (int n, int npanels, const int *__restrict__ rp,
           const int *__restrict__ ci, const double *__restrict__ v,
           const float *__restrict__ Zs, double a, double b,
           const float *__restrict__ D, const double *__restrict__ R,
           double *__restrict__ X, int xmode, float *__restrict__ Dn,
           double *__restrict__ Rn, float *__restrict__ Rn32)
{
  LGH_PANELS (NP)
  auto dnew = [&] (size_t q) -> double {
    double bz = b * (double) Zs[q];
    double d = HASD ? a * (double) D[q] + bz : bz;
    return (double) (float) d;
  };
  LGH_ROW_BEGIN (n)
  double acc[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) acc[p] = 0;
  for_nbrs<double, false> (row, rp, ci, v, (const double *) nullptr, lane,
                           [&] (int c, double w, double) {
#pragma unroll
    for (int p = 0; p < NP; p++)
      if (p < nph) acc[p] += w * dnew (base + p * ps + (size_t) c * 64);
  });
#pragma unroll
  for (int p = 0; p < NP; p++)
    if (p < nph) {
      const size_t qi = base + p * ps + (size_t) row * 64;
      const double di = dnew (qi);
      Dn[qi] = (float) di;
      X[qi] = xmode ? X[qi] + di : di;
      const double rn = R[qi] - acc[p];
      Rn[qi] = rn;
      Rn32[qi] = (float) rn;
    }
  LGH_ROW_END
}

int
k_cheb_step_mixed (const DevCSR &A, const PMat &Z32, double a, double b,
                   const PMat *D32, const PMat &R, PMat &X, int x_mode,
                   PMat &Dn32, PMat &Rn, PMat &Rn32)
{
  LGH_CHECK (A.prec == F64 && R.prec == F64 && X.prec == F64 && R.W == 64
             && Z32.prec == F32 && Dn32.prec == F32 && Rn32.prec == F32
             && Rn.prec == F64 && R.nrows == A.n, 8,
             "k_cheb_step_mixed: shape/precision mismatch");
  const double vec = (double) A.n * 64;
  ProfScope ps (PK_CHEB_Z, R.npanels () * (csr_bytes (A)
                + vec * (4 + (D32 ? 4 : 0) + 8 + (x_mode ? 8 : 0) + 4 + 8 + 8 + 4)),
                R.npanels () * (2.0 * (double) A.nnz + 6.0 * A.n) * 64);
  const int np = R.npanels ();
  auto l = [&] (auto npt) -> int {
    constexpr int NP = decltype (npt)::value;
    dim3 g (grid_rows64 (A.n), (np + NP - 1) / NP);
    if (D32)
      hipLaunchKernelGGL ((cheb64m_k<NP, true>), g, dim3 (TPB), 0, stream (),
                          A.n, np, A.rp, A.ci, (const double *) A.v,
                          (const float *) Z32.p, a, b, (const float *) D32->p,
                          (const double *) R.p, (double *) X.p, x_mode,
                          (float *) Dn32.p, (double *) Rn.p, (float *) Rn32.p);
    else
      hipLaunchKernelGGL ((cheb64m_k<NP, false>), g, dim3 (TPB), 0, stream (),
                          A.n, np, A.rp, A.ci, (const double *) A.v,
                          (const float *) Z32.p, a, b, nullptr,
                          (const double *) R.p, (double *) X.p, x_mode,
                          (float *) Dn32.p, (double *) Rn.p, (float *) Rn32.p);
    LGH_LAUNCH_CHECK ();
    return 0;
  };
  return np_dispatch (np, l);
}

// degree-2 correction-form smoother (zsolve_impl.h:385-404 with nu = 2) in
// two fused passes.  With X = X0 on entry, step 1 gives D1 = b1 dinv R,
// R1 = R - A D1; step 2 gives D2 = a2 D1 + b2 dinv R1, R2 = R1 - A D2 and
// the smoothed X = X0 + D1 + D2.  X is touched once (in step 2), D2 is
// never stored, and R2 only when it is used (pre-smoothing: restriction).
template <typename TV, typename T, int NP>
__global__ void __launch_bounds__ (TPB)
sm1_k (int n, int npanels, const int *__restrict__ rp,
       const int *__restrict__ ci, const TV *__restrict__ v,
       const TV *__restrict__ dinv, T b, const T *__restrict__ Rin,
       T *__restrict__ D1, T *__restrict__ R1)
{
  LGH_PANELS (NP)
  LGH_ROW_BEGIN (n)
  T acc[NP];
#pragma unroll
  for (int p = 0; p < NP; p++) acc[p] = 0;
  for_nbrs<TV, true> (row, rp, ci, v, dinv, lane, [&] (int c, TV w, TV dd) {
#pragma unroll
    for (int p = 0; p < NP; p++)
      if (p < nph)
        acc[p] += (T) w * (b * ((T) dd * Rin[base + p * ps + (size_t) c * 64]));
  });
  const T dr = (T) dinv[row];
#pragma unroll
  for (int p = 0; p < NP; p++)
    if (p < nph) {
      const size_t qi = base + p * ps + (size_t) row * 64;
      const T rin = Rin[qi];
      D1[qi] = b * (dr * rin);
      R1[qi] = rin - acc[p];
    }
  LGH_ROW_END
}

template <typename TV, typename T, int NP, bool XADD, bool WR>
__global__ void __launch_bounds__ (TPB)
sm2_k (int n, int npanels, const int *__restrict__ rp,
       const int *__restrict__ ci, const TV *__restrict__ v,
       const TV *__restrict__ dinv, T a, T b, const T *__restrict__ D1,
       const T *__restrict__ R1, T *__restrict__ X, T *__restrict__ R2)
{
  LGH_PANELS (NP)
  auto d2 = [&] (size_t q, T dj) -> T {
    T bz = b * (dj * R1[q]);
    T ad = a * D1[q];
    return ad + bz;
  };
  LGH_ROW_BEGIN (n)
  const T dr = (T) dinv[row];
#pragma unroll
  for (int p = 0; p < NP; p++)
    if (p < nph) {
      const size_t qi = base + p * ps + (size_t) row * 64;
      const T di = d2 (qi, dr);
      const T d1 = D1[qi];
      X[qi] = XADD ? X[qi] + d1 + di : d1 + di;
    }
  if (WR) {
    T acc[NP];
#pragma unroll
    for (int p = 0; p < NP; p++) acc[p] = 0;
    for_nbrs<TV, true> (row, rp, ci, v, dinv, lane, [&] (int c, TV w, TV dd) {
#pragma unroll
      for (int p = 0; p < NP; p++)
        if (p < nph) acc[p] += (T) w * d2 (base + p * ps + (size_t) c * 64, (T) dd);
    });
#pragma unroll
    for (int p = 0; p < NP; p++)
      if (p < nph) {
        const size_t qi = base + p * ps + (size_t) row * 64;
        R2[qi] = R1[qi] - acc[p];
      }
  }
  LGH_ROW_END
}

template <typename TV, typename T>
static int
smooth2_t (const DevCSR &A, double b1, double a2, double b2, const PMat &Rin,
           PMat &D1, PMat &R1, PMat &X, int xadd, PMat *R2)
{
  const int np = Rin.npanels ();
  auto l = [&] (auto npt) -> int {
    constexpr int NP = decltype (npt)::value;
    dim3 g (grid_rows64 (A.n), (np + NP - 1) / NP);
    hipLaunchKernelGGL ((sm1_k<TV, T, NP>), g, dim3 (TPB), 0, stream (), A.n,
                        np, A.rp, A.ci, (const TV *) A.v, (const TV *) A.dinv,
                        (T) b1, (const T *) Rin.p, (T *) D1.p, (T *) R1.p);
    LGH_LAUNCH_CHECK ();
    T *r2 = R2 ? (T *) R2->p : nullptr;
#define LGH_SM2(XA, W_)                                                       \
    hipLaunchKernelGGL ((sm2_k<TV, T, NP, XA, W_>), g, dim3 (TPB), 0,         \
                        stream (), A.n, np, A.rp, A.ci, (const TV *) A.v,      \
                        (const TV *) A.dinv, (T) a2, (T) b2, (const T *) D1.p, \
                        (const T *) R1.p, (T *) X.p, r2)
    if (xadd) { if (R2) LGH_SM2 (true, true); else LGH_SM2 (true, false); }
    else { if (R2) LGH_SM2 (false, true); else LGH_SM2 (false, false); }
#undef LGH_SM2
    LGH_LAUNCH_CHECK ();
    return 0;
  };
  return np_dispatch (np, l);
}

int
k_smooth2 (const DevCSR &A, double b1, double a2, double b2, const PMat &Rin,
           PMat &D1, PMat &R1, PMat &X, int xadd, PMat *R2)
{
  LGH_CHECK (Rin.W == 64 && X.W == 64 && D1.W == 64 && R1.W == 64
             && Rin.prec == X.prec && A.dinv && Rin.nrows == A.n, 8,
             "k_smooth2: shape mismatch");
  const double vec = (double) A.n * 64 * prec_bytes (Rin.prec);
  const double streams = 3 + (2 + (xadd ? 1 : 0) + 1 + (R2 ? 1 : 0));
  ProfScope ps (PK_CHEB_JAC, Rin.npanels () * (2 * csr_bytes (A) + streams * vec),
                Rin.npanels () * (4.0 * (double) A.nnz + 10.0 * A.n) * 64);
  if (Rin.prec == F64)
    return A.prec == F64 ? smooth2_t<double, double> (A, b1, a2, b2, Rin, D1, R1, X, xadd, R2)
                         : smooth2_t<float, double> (A, b1, a2, b2, Rin, D1, R1, X, xadd, R2);
  return A.prec == F64 ? smooth2_t<double, float> (A, b1, a2, b2, Rin, D1, R1, X, xadd, R2)
                       : smooth2_t<float, float> (A, b1, a2, b2, Rin, D1, R1, X, xadd, R2);
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

template <typename TV, typename T>
static int
spmm_t (const DevCSR &A, double alpha, const PMat &X, double beta, PMat &Y)
{
  if (X.W == 64) {
    const int np = X.npanels ();
    const unsigned gx = grid_rows64 (A.n);
    auto launch = [&] (int NPsel) -> int {
#define LGH_SP64(NP)                                                          \
      hipLaunchKernelGGL ((spmm64_k<TV, T, NP>), dim3 (gx, (np + NP - 1) / NP), \
                          dim3 (TPB), 0, stream (), A.n, A.m, A.rp, A.ci,      \
                          (const TV *) A.v, (T) alpha, (const T *) X.p,        \
                          (T) beta, (T *) Y.p, np)
      switch (NPsel) {
      case 8: LGH_SP64 (8); break;
      case 4: LGH_SP64 (4); break;
      case 2: LGH_SP64 (2); break;
      default: LGH_SP64 (1); break;
      }
#undef LGH_SP64
      LGH_LAUNCH_CHECK ();
      return 0;
    };
    const int vp = vt<T>::prec;
    int &tuned = const_cast<DevCSR &> (A).spmm_np[vp];
    const int widest = np >= 5 ? 8 : np >= 3 ? 4 : np == 2 ? 2 : 1;
    if (tuned == 0 && np >= 2 && beta == 0.0 && A.nnz >= 16 * (int64_t) A.n) {
      // autotune once per (matrix, vector precision): panels per block trade
      // CSR re-reads (fewer panels) against gather working set (more panels)
      hipEvent_t e0, e1;
      LGH_HIP (hipEventCreate (&e0));
      LGH_HIP (hipEventCreate (&e1));
      float best = 1e30f;
      int best_np = widest;
      for (int cand = 1; cand <= widest; cand *= 2) {
        LGH_HIP (hipEventRecord (e0, stream ()));
        LGH_TRY (launch (cand));
        LGH_HIP (hipEventRecord (e1, stream ()));
        LGH_HIP (hipEventSynchronize (e1));
        float ms = 0.f;
        LGH_HIP (hipEventElapsedTime (&ms, e0, e1));
        if (ms < best) { best = ms; best_np = cand; }
      }
      (void) hipEventDestroy (e0);
      (void) hipEventDestroy (e1);
      tuned = best_np;
    }
    return launch (tuned ? tuned : widest);
  }
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
                np_ * 2.0 * (double) A.nnz * X.W);
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

template <typename TV, typename T>
static int
resid_t (const DevCSR &A, const PMat &B, const PMat &X, PMat &R)
{
  if (X.W == 64) {
    const int np = X.npanels ();
    auto l64 = [&] (auto npt) -> int {
      constexpr int NP = decltype (npt)::value;
      dim3 g (grid_rows64 (A.n), (np + NP - 1) / NP);
      hipLaunchKernelGGL ((resid64_k<TV, T, NP>), g, dim3 (TPB), 0, stream (),
                          A.n, np, A.rp, A.ci, (const TV *) A.v,
                          (const T *) B.p, (const T *) X.p, (T *) R.p);
      LGH_LAUNCH_CHECK ();
      return 0;
    };
    return np_dispatch (np, l64);
  }
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
  ProfScope ps (PK_RESID, X.npanels () * (csr_bytes (A)
                + 3.0 * A.n * X.W * prec_bytes (X.prec)),
                X.npanels () * 2.0 * (double) A.nnz * X.W);
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

template <typename TV, typename T>
static int
cheb_t (const DevCSR &A, int jac, const PMat *Zs, double a, double b,
        const PMat *D, const PMat &R, PMat &X, int xmode, PMat &Dn, PMat &Rn)
{
  if (R.W == 64) {
    const int np = R.npanels ();
    const T *zp = Zs ? (const T *) Zs->p : nullptr;
    const T *dp = D ? (const T *) D->p : nullptr;
    auto l64 = [&] (auto npt) -> int {
      constexpr int NP = decltype (npt)::value;
      dim3 g (grid_rows64 (A.n), (np + NP - 1) / NP);
#define LGH_CH64(J, H)                                                        \
      hipLaunchKernelGGL ((cheb64_k<TV, T, NP, J, H>), g, dim3 (TPB), 0,     \
                          stream (), A.n, np, A.rp, A.ci, (const TV *) A.v,    \
                          (const TV *) A.dinv, zp, (T) a, (T) b, dp,           \
                          (const T *) R.p, (T *) X.p, xmode, (T *) Dn.p,       \
                          (T *) Rn.p)
      if (jac) { if (dp) LGH_CH64 (true, true); else LGH_CH64 (true, false); }
      else { if (dp) LGH_CH64 (false, true); else LGH_CH64 (false, false); }
#undef LGH_CH64
      LGH_LAUNCH_CHECK ();
      return 0;
    };
    return np_dispatch (np, l64);
  }
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
             PMat &Dn, PMat &Rn)
{
  LGH_CHECK (R.nrows == A.n && X.W == R.W && Dn.W == R.W && Rn.W == R.W
             && X.prec == R.prec && (!jacobi || A.dinv), 8,
             "k_cheb_step: shape mismatch");
  // streams: read R, (D), (Zsrc), (X); write Dn, X, Rn
  const double vec = (double) A.n * R.W * prec_bytes (R.prec);
  const double streams = 1 + (D ? 1 : 0) + (jacobi ? 0 : 1) + (x_mode ? 1 : 0) + 3;
  ProfScope ps (jacobi ? PK_CHEB_JAC : PK_CHEB_Z,
                R.npanels () * (csr_bytes (A) + streams * vec
                                + (jacobi ? (double) A.n * prec_bytes (A.prec) : 0.0)),
                R.npanels () * (2.0 * (double) A.nnz + 6.0 * A.n) * R.W);
  if (R.prec == F64)
    return A.prec == F64
        ? cheb_t<double, double> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn)
        : cheb_t<float, double> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn);
  return A.prec == F64
      ? cheb_t<double, float> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn)
      : cheb_t<float, float> (A, jacobi, Zsrc, a, b, D, R, X, x_mode, Dn, Rn);
}




// ---------------------------------------------------------------------------
// dense coarse apply X0 = Ainv B0 (Ainv row-major n0 x n0)

template <typename T, int W>
__global__ void __launch_bounds__ (TPB)
dense_apply_k (int n0, const T *__restrict__ Ainv, const T *__restrict__ B,
               T *__restrict__ X)
{
  constexpr int RPW = 64 / W;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int c = lane % W;
  const int row = (blockIdx.x * (TPB / 64) + wave) * RPW + lane / W;
  if (row >= n0) return;
  const size_t off = (size_t) blockIdx.y * n0 * W + c;
  T acc = 0;
  for (int k = 0; k < n0; k++)
    acc += Ainv[(size_t) row * n0 + k] * B[off + (size_t) k * W];
  X[off + (size_t) row * W] = acc;
}

int
k_dense_apply (const void *Ainv, int n0, const PMat &B, PMat &X)
{
  auto l = [&] (auto wt) -> int {
      constexpr int W = decltype (wt)::value;
      constexpr int rows_per_block = (TPB / 64) * (64 / W);
      dim3 g (grid1 (n0, rows_per_block), B.npanels ());
      if (B.prec == F64)
        hipLaunchKernelGGL ((dense_apply_k<double, W>), g, dim3 (TPB), 0,
                            stream (), n0, (const double *) Ainv,
                            (const double *) B.p, (double *) X.p);
      else
        hipLaunchKernelGGL ((dense_apply_k<float, W>), g, dim3 (TPB), 0,
                            stream (), n0, (const float *) Ainv,
                            (const float *) B.p, (float *) X.p);
      LGH_LAUNCH_CHECK ();
      return 0;
  };
  LGH_CHECK (B.nrows == n0 && X.nrows == n0 && B.W == X.W, 8,
             "k_dense_apply: shape mismatch");
  return with_W (B.W, l);
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
  size_t e = (size_t) blockIdx.x * TPB + threadIdx.x;
  if (e >= (size_t) nrows * nc) return;
  int j = (int) (e / nrows);
  int i = (int) (e % nrows);
  int xj = xc0 + j, yj = yc0 + j;
  size_t xi = ((size_t) (xj / xW) * nrows + i) * xW + xj % xW;
  size_t yi = ((size_t) (yj / yW) * nrows + i) * yW + yj % yW;
  Y[yi] = (TO) X[xi];
}

int
k_copy_cols (const PMat &X, int xc0, PMat &Y, int yc0, int nc)
{
  LGH_CHECK (X.nrows == Y.nrows && xc0 + nc <= X.npanels () * X.W
             && yc0 + nc <= Y.npanels () * Y.W, 8, "k_copy_cols: range");
  if (nc <= 0) return 0;
  unsigned g = grid1 ((size_t) X.nrows * nc, TPB);
#define LGH_CC(TI, TO)                                                        \
  hipLaunchKernelGGL ((copy_cols_k<TI, TO>), dim3 (g), dim3 (TPB), 0,         \
                      stream (), X.nrows, nc, (const TI *) X.p, X.W, xc0,     \
                      (TO *) Y.p, Y.W, yc0)
  if (X.prec == F64 && Y.prec == F64) LGH_CC (double, double);
  else if (X.prec == F64) LGH_CC (double, float);
  else if (Y.prec == F64) LGH_CC (float, double);
  else LGH_CC (float, float);
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

// Gram partials: block (chunk, pair) computes the 64 x 64 tile
// X_pi^T Y_pj over rows [r0, r1).  Wave w owns rows 32*(w>>1).., cols
// 32*(w&1).. of the tile as 2 x 2 MFMA 16 x 16 tiles.  MFMA 16x16x4
// operand layout: A(i, k) at lane i + 16 k, B(k, j) at lane j + 16 k;
// result rows per mfma<T>::row.
template <typename T>
__global__ void __launch_bounds__ (TPB)
gram_k (int n, const T *__restrict__ X, int xp0, const T *__restrict__ Y,
        int yp0, int npy, int sym, int rows_per_chunk,
        const int2 *__restrict__ pairs, double *__restrict__ part)
{
  typedef typename mfma<T>::acc_t acc_t;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int2 pr = pairs[blockIdx.y];
  const T *Xp = X + (size_t) (xp0 + pr.x) * n * 64;
  const T *Yp = Y + (size_t) (yp0 + pr.y) * n * 64;
  const int r0 = blockIdx.x * rows_per_chunk;
  const int r1 = min (n, r0 + rows_per_chunk);
  const int mb = 32 * (wave >> 1), nb = 32 * (wave & 1);
  const int li = lane & 15, lk = lane >> 4;
  acc_t acc[2][2];
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int r = 0; r < 4; r++) acc[a][b][r] = 0;
  for (int k0 = r0; k0 < r1; k0 += 4) {
    const int kr = k0 + lk;
    T av[2], bv[2];
    if (kr < r1) {
      const T *xr = Xp + (size_t) kr * 64;
      const T *yr = Yp + (size_t) kr * 64;
      av[0] = xr[mb + li]; av[1] = xr[mb + 16 + li];
      bv[0] = yr[nb + li]; bv[1] = yr[nb + 16 + li];
    }
    else { av[0] = av[1] = bv[0] = bv[1] = (T) 0; }
    for (int a = 0; a < 2; a++)
      for (int b = 0; b < 2; b++)
        acc[a][b] = mfma<T>::op (av[a], bv[b], acc[a][b]);
  }
  double *tile = part + ((size_t) blockIdx.x * gridDim.y + blockIdx.y) * 4096;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int r = 0; r < 4; r++) {
        const int row = mb + 16 * a + mfma<T>::row (lk, r);
        const int col = nb + 16 * b + li;
        tile[col * 64 + row] = (double) acc[a][b][r];
      }
}

__global__ void __launch_bounds__ (TPB)
gram_reduce_k (int nchunks, int npairs, const int2 *__restrict__ pairs,
               const double *__restrict__ part, double *G, int ldg, int sym)
{
  const int pair = blockIdx.y;
  const int e = blockIdx.x * TPB + threadIdx.x;   // 0..4095
  if (e >= 4096) return;
  double t = 0.0;
  for (int q = 0; q < nchunks; q++)
    t += part[((size_t) q * npairs + pair) * 4096 + e];
  const int2 pr = pairs[pair];
  const int row = e % 64, col = e / 64;
  G[(size_t) (pr.x * 64 + row) + (size_t) (pr.y * 64 + col) * ldg] = t;
  if (sym && pr.x != pr.y)
    G[(size_t) (pr.y * 64 + col) + (size_t) (pr.x * 64 + row) * ldg] = t;
}

int
k_gram (const PMat &X, int xp0, int npx, const PMat &Y, int yp0, int npy,
        double *G, int ldg, int sym)
{
  LGH_CHECK (X.W == 64 && Y.W == 64 && X.prec == Y.prec
             && X.nrows == Y.nrows && xp0 + npx <= X.npanels ()
             && yp0 + npy <= Y.npanels (), 8, "k_gram: shape mismatch");
  std::vector<int2> pv;
  for (int i = 0; i < npx; i++)
    for (int j = (sym ? i : 0); j < npy; j++) pv.push_back (make_int2 (i, j));
  const int npairs = (int) pv.size ();
  if (npairs == 0) return 0;
  const int n = X.nrows;
  ProfScope ps (PK_GRAM, (double) n * 64 * prec_bytes (X.prec) * (npx + npy),
                2.0 * n * 4096.0 * npairs);
  // aim for ~4096 blocks; at least 256 rows per chunk
  int nch = 4096 / npairs;
  if (nch < 1) nch = 1;
  int rpc = (n + nch - 1) / nch;
  if (rpc < 256) rpc = 256;
  rpc = (rpc + 3) / 4 * 4;
  nch = (n + rpc - 1) / rpc;
  if (nch < 1) nch = 1;
  int2 *pd = (int2 *) dmalloc (sizeof (int2) * npairs);
  double *part = (double *) dmalloc (sizeof (double) * 4096 * (size_t) nch * npairs);
  if (!pd || !part) { dfree (pd); dfree (part); return 5; }
  int rc = h2d (pd, pv.data (), sizeof (int2) * npairs);
  if (!rc) {
    dim3 g (nch, npairs);
    if (X.prec == F64)
      hipLaunchKernelGGL (gram_k<double>, g, dim3 (TPB), 0, stream (), n,
                          (const double *) X.p, xp0, (const double *) Y.p,
                          yp0, npy, sym, rpc, pd, part);
    else
      hipLaunchKernelGGL (gram_k<float>, g, dim3 (TPB), 0, stream (), n,
                          (const float *) X.p, xp0, (const float *) Y.p, yp0,
                          npy, sym, rpc, pd, part);
    hipError_t e = hipGetLastError ();
    if (e != hipSuccess) rc = hip_fail (e, "gram_k", __FILE__, __LINE__);
  }
  if (!rc) {
    hipLaunchKernelGGL (gram_reduce_k, dim3 (4096 / TPB, npairs), dim3 (TPB),
                        0, stream (), nch, npairs, pd, part, G, ldg, sym);
    hipError_t e = hipGetLastError ();
    if (e != hipSuccess) rc = hip_fail (e, "gram_reduce_k", __FILE__, __LINE__);
  }
  dfree (pd);
  dfree (part);
  return rc;
}

// Y panel block = alpha X C + beta Y.  Block (rowtile, out panel q) stages
// the 64 x 64 tile of each input panel and the matching 64 x 64 block of C
// in LDS; wave w owns output rows 16 w .. 16 w + 15 (4 MFMA tiles across).
static constexpr int GPAD = 65;

template <typename T>
__global__ void __launch_bounds__ (TPB)
gemm_nn_k (int n, const T *__restrict__ X, int xp0, int npx,
           const double *__restrict__ C, int ldc, T alpha, T beta,
           T *__restrict__ Y, int yp0)
{
  typedef typename mfma<T>::acc_t acc_t;
  __shared__ T Xs[64 * GPAD];
  __shared__ T Cs[64 * GPAD];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int r0 = blockIdx.x * 64;
  const int q = blockIdx.y;
  acc_t acc[4];
  for (int b = 0; b < 4; b++)
    for (int r = 0; r < 4; r++) acc[b][r] = 0;
  for (int p = 0; p < npx; p++) {
    const T *Xp = X + (size_t) (xp0 + p) * n * 64;
    __syncthreads ();
    for (int e = threadIdx.x; e < 4096; e += TPB) {
      const int rr = e >> 6, cc = e & 63;
      Xs[rr * GPAD + cc] = (r0 + rr < n) ? Xp[(size_t) (r0 + rr) * 64 + cc] : (T) 0;
      // C block (k = p*64 + kk, j = q*64 + jj), stored Cs[kk][jj]
      const int kk = e & 63, jj = e >> 6;
      Cs[kk * GPAD + jj] = (T) C[(size_t) (p * 64 + kk) + (size_t) (q * 64 + jj) * ldc];
    }
    __syncthreads ();
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
           double alpha, double beta, PMat &Y, int yp0, int npy)
{
  LGH_CHECK (X.W == 64 && Y.W == 64 && X.prec == Y.prec
             && X.nrows == Y.nrows && xp0 + npx <= X.npanels ()
             && yp0 + npy <= Y.npanels (), 8, "k_gemm_nn: shape mismatch");
  if (npy == 0) return 0;
  ProfScope ps (PK_GEMM, (double) X.nrows * 64 * prec_bytes (X.prec)
                * (npx + npy * (beta != 0.0 ? 2 : 1)),
                2.0 * X.nrows * 4096.0 * npx * npy);
  dim3 g (grid1 (X.nrows, 64), npy);
  if (X.prec == F64)
    hipLaunchKernelGGL (gemm_nn_k<double>, g, dim3 (TPB), 0, stream (),
                        X.nrows, (const double *) X.p, xp0, npx, C, ldc,
                        alpha, beta, (double *) Y.p, yp0);
  else
    hipLaunchKernelGGL (gemm_nn_k<float>, g, dim3 (TPB), 0, stream (),
                        X.nrows, (const float *) X.p, xp0, npx, C, ldc,
                        (float) alpha, (float) beta, (float *) Y.p, yp0);
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
