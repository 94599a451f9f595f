// lgh_dev.hpp — internal device layer of the MI355X GLR engine.
//
// Data layout in HBM.  Every tall block (sketch, Q, Y, U, right-hand
// sides) is a PANEL MATRIX: an nrows x ncols matrix split into column
// panels of width W (W in {1,2,4,...,64}); panel p is an nrows x W
// row-major slab, so element (i, j) lives at
//     ((j / W) * nrows + i) * W + (j % W).
// With W = 64 a panel row is 128 B (bf16) / 256 B (fp32) / 512 B (fp64)
// contiguous: 8, 16 or 32 lanes load it as 16 B per lane, so every CSR
// neighbour gather is one coalesced row load, and elementwise work streams
// contiguously.  Columns past ncols in the last panel are padding and are
// kept at zero by every kernel.
//
// Sparse matrices are CSR with int32 row pointers / column indices and
// fp64 or fp32 values.  Small dense matrices (Gram, T, eigenvectors,
// coefficient blocks) are fp64 column-major.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <mutex>
#include <string>
#include <vector>

namespace lgh {

// BF16 is a storage precision only (arithmetic in fp32): the vectors of
// the BF16 V-cycle tier; its matrices are stored in fp32
enum Prec { F64 = 0, F32 = 1, BF16 = 2 };
inline size_t prec_bytes (int p) { return p == F64 ? 8 : p == F32 ? 4 : 2; }
struct bf16_t { uint16_t u; };

// ---- errors -------------------------------------------------------------
void set_error (const std::string &msg);
const char *last_error ();
int hip_fail (hipError_t e, const char *what, const char *file, int line);
#define LGH_HIP(expr)                                                        \
  do {                                                                       \
    hipError_t _e = (expr);                                                  \
    if (_e != hipSuccess) return ::lgh::hip_fail (_e, #expr, __FILE__, __LINE__); \
  } while (0)
#define LGH_TRY(expr)                                                        \
  do { int _rc = (expr); if (_rc) return _rc; } while (0)
#define LGH_CHECK(cond, code, msg)                                           \
  do { if (!(cond)) { ::lgh::set_error (msg); return (code); } } while (0)

// ---- context ------------------------------------------------------------
int  dev_init ();                 // idempotent; selects device 0 and
                                  // makes the calling thread's stream
// a worker thread leaves the engine: waits for its streams and hands its
// cached blocks to the shared list (keep_blocks: a batch worker, whose
// blocks the next batch reuses) or frees them (a serving thread, whose
// blocks of ever-changing batch widths nothing after it reuses: left in
// the cache they would push the next large request into a full flush)
void dev_thread_exit (bool keep_blocks = true);
// the library is initialized and the calling thread has its stream (so an
// engine call on it creates nothing)
bool dev_ready ();
// once per device, with it current (dev_init): the address of the
// no-allocation entry points' static scratch array
int  scratch_init ();
hipStream_t stream ();
// the thread's auxiliary stream (default priority, created on first use):
// work that may overlap the main stream's, ordered by events
hipStream_t aux_stream ();
// this thread's aux_stream () returns null (no overlap) when off
void set_aux_overlap (bool on);
// when off, this thread's Z-solves run a block's partial last panel as a
// whole 64-wide panel with the others instead of as narrow panels of its
// own (the same bits per column either way; fewer launches, more bytes)
void set_tail_split (bool on);
bool tail_split ();
// while alive, the thread's engine calls enqueue on s instead of its
// stream (allocator free lists follow: blocks freed here are reused on s)
struct StreamScope {
  hipStream_t prev;
  explicit StreamScope (hipStream_t s);
  ~StreamScope ();
};
int  dev_sync ();
void *rocblas_h ();               // rocblas_handle bound to stream()

// caching device allocator (no hipMalloc/hipFree churn inside builds)
void *dmalloc (size_t bytes);     // nullptr on failure (error set)
void  dfree (void *p);
void  dtrim ();                   // release cached blocks
size_t dbytes_in_use ();

// synchronous host <-> device copies on the thread's stream; blocks of
// 64 MiB and more in pageable memory move through two pinned 64 MiB chunks
// per thread, each chunk's host copy split over 8 threads and overlapping
// the DMA of the other
int h2d (void *dst, const void *src, size_t bytes);
int d2h (void *dst, const void *src, size_t bytes);
// pinned host blocks that h2d / d2h move directly (no staging)
void *host_pinned_alloc (size_t bytes);   // nullptr on failure (error set)
void  host_pinned_free (void *p);
int dzero (void *dst, size_t bytes);   // fill kernel (graph-safe)
// achieved HBM rate of a device copy of `bytes` (read + write counted),
// best of reps launches after a warm-up
int copy_gbps (size_t bytes, int reps, double *gbps);
int dcopy (void *dst, const void *src, size_t bytes);
// device-to-device copy of height rows of width bytes between pitched
// blocks (columns of a column-major matrix with leading dimensions)
int dcopy2d (void *dst, size_t dpitch, const void *src, size_t spitch,
             size_t width, size_t height);
// true when p is device (or managed) memory of the HIP runtime
bool is_device_ptr (const void *p);
// true when [p, p + bytes) lies inside one device allocation
bool device_range_ok (const void *p, size_t bytes);
// plain device blocks outside the caching allocator (a caller's own
// device memory, for the device-resident entry points)
int  dev_block_alloc (size_t bytes, void **p);
void dev_block_free (void *p);

// ---- live kernel timing (HIP events on the engine stream) ---------------
// Launchers bracket their kernel with events when profiling is enabled and
// credit the launch's algorithmic HBM bytes / flops to a kernel class.
enum ProfKind {
  PK_SPMM_B = 0,      // B X of the F sweep
  PK_SPMM = 1,        // other SpMMs (Z apply, P, P^T)
  PK_CHEB_JAC = 2,    // Jacobi Chebyshev steps and smoothers below the
                      // finest AMG level
  PK_CHEB_Z = 3,      // fused Chebyshev step, explicit preconditioned Z
  PK_RESID = 4,       // residuals below the finest AMG level
  PK_GRAM = 5,
  PK_GEMM = 6,
  PK_OTHER = 7,
  PK_CHEB_MIXED = 8,  // mixed-precision Chebyshev step after the
                      // residual replacement (FP32 R, FP64 X)
  PK_SM_PRE = 9,      // fused smoother, finest level, zero initial guess
  PK_SM_POST = 10,    // fused smoother, finest level, X += correction
  PK_RESID_FINE = 11, // residual on the finest AMG level
  PK_MG_COARSE = 12,  // the fused coarse levels of a cycle (one launch)
  PK_N = 13
};
void prof_enable (int on);
int  prof_enabled ();
void prof_set_b_tag (int on);      // next SpMMs are the F sweep's B product
int  prof_b_tag ();
void prof_set_fine_tag (int on);   // next launches run on the finest level
int  prof_fine_tag ();
struct ProfScope {
  int kind, np, prec, W; double bytes, flops; int active; void *ev0;
  // np, prec, W: panels, vector storage precision and panel width of the
  // launch (recorded so that one kernel instantiation at one shape can be
  // selected)
  ProfScope (int kind, double bytes, double flops, int np = 0, int prec = -1,
             int W = 64);
  ~ProfScope ();
};
// totals since the last collect of that kind (synchronizes the stream);
// np >= 0 / prec >= 0 count only launches over np full-width (64) panels /
// at that storage precision (the others are dropped too)
int prof_collect (int kind, int np, int prec, long *count, double *ms,
                  double *bytes, double *flops);

// ---- objects ------------------------------------------------------------
// Row-grouped copy of a wide-row matrix (the B of the F sweep): rows in
// groups of rg (4 or 8); group g holds the sorted union of its rows' column
// indices, padded to whole batches of grp_batch(rg) entries (zero values at
// the last column), batches gp[g] .. gp[g+1]; batch b has the column
// indices gci[b * nb .. + nb) and, per entry u, rg values
// gv[(b * nb + u) * rg + r] (0 where row r has no entry).  A batch holds 32
// values, which the SpMM reads with a few wide scalar loads.
constexpr int grp_batch (int rg) { return 32 / rg; }
struct GroupCopy {
  int rg = 0;
  int *gp = nullptr, *gci = nullptr;
  void *gv = nullptr;             // values, precision of the owning matrix
  int64_t nbatch = 0;
  void release ();
};

// Row-tile copy of a wide-row matrix for the LDS-staged SpMM
// (spmm_tile_k).  A tile is nw * rg rows (consecutive rows, or a tx-wide
// block of consecutive rows on successive lines `stride` rows apart, row
// list trow); wavefront w of the tile's workgroup owns rows w*rg .. +rg of
// it.  The tile's sorted union of column indices is cut into chunks of ch
// entries (the last one padded with its last index); chunk c stages the X
// rows tu[c*ch .. + ch) in LDS.  Wavefront w's own union (the merge of its
// rows) restricted to chunk c is batches cb[c*nw + w] .. cb[c*nw + w + 1];
// batch b has tile_batch(rg) entries, entry u at LDS slot es[b*nb + u] with
// rg values ev[(b*nb + u)*rg + r] (0 where row r has no entry; padding
// entries: value 0).  Tiles tp[t] .. tp[t+1] in chunks.  With rg = 16 (the
// MFMA kernel) a batch is one 4-entry k-step and its 64 values are in the
// MFMA's A-operand lane order (entry k, row i at k*16 + i); es and ev then
// carry ch/4 batches of zero padding past the last batch.
constexpr int tile_batch (int rg) { return rg == 16 ? 4 : 16 / rg; }
struct TileShape {
  int rg = 0, nw = 0, ch = 0;     // rows per wavefront, wavefronts, chunk
  int tx = 0, stride = 0;         // 0: consecutive rows; else a tx x (nw*rg/tx) block
  int xcd = 0;                    // launch order: 1 tiles dealt to the XCDs in
                                  // contiguous runs, 2 (2D) in strips of the tile grid
  int wy = 1;                     // 2D: lines a wavefront's rows span ((rg/wy) x wy)
};
struct TileCopy {
  TileShape s;
  int ntiles = 0, ntx = 0;        // ntx: tiles a band (2D), 0 for 1D
  int *tp = nullptr, *tu = nullptr, *cb = nullptr, *es = nullptr;
  int *trow = nullptr;            // null: consecutive rows
  void *ev = nullptr;             // values, precision of the owning matrix
  int64_t nchunk = 0, nbatch = 0;
  double union_per_row = 0.0;     // staged X rows per matrix row
  void release ();
};

struct DevCSR {
  int n = 0, m = 0;               // rows, cols
  int64_t nnz = 0;
  int *rp = nullptr, *ci = nullptr;
  void *v = nullptr;              // values, precision `prec`
  int prec = F64;
  void *dinv = nullptr;           // 1/diag (same precision), square only
  // row-grouped copy (csr_group_rows) and the panels per wavefront its
  // SpMM uses with FP64 / FP32 vectors (chosen by measurement)
  GroupCopy grp;
  int grp_ppw[2] = { 4, 8 };
  // row-tile copy (chosen by the same autotuner when it is faster; then
  // the grouped copy is not kept)
  TileCopy tile;
  // scaled values (csr_scale_cols): vs[k] = v[k] * (vs_b1 * dinv[ci[k]]),
  // each product rounded in precision `prec`
  void *vs = nullptr;
  double vs_b1 = 0.0;
  ~DevCSR ();
};
// upload host CSR (values fp64 on the host) at precision prec
int csr_upload (int n, int m, const int *rp, const int *ci, const double *v,
                int prec, int want_dinv, DevCSR **out);
// build the row-grouped copy with rg rows per group (same host CSR, rows
// with strictly increasing column indices) into out
int csr_group_build (const DevCSR &A, const int *rp, const int *ci,
                     const double *v, int rg, GroupCopy &out);
// build the row-tile copy of shape s (same host CSR, sorted rows) into out
int csr_tile_build (const DevCSR &A, const int *rp, const int *ci,
                    const double *v, const TileShape &s, TileCopy &out);
// the line stride of a grid-ordered matrix for 2D tiles of width tx: the
// offset, among the column offsets of a middle row, whose tx x (rows/tx)
// tiles have the smallest union per row (0: none smaller than 1D tiles)
int csr_tile_stride (int n, const int *rp, const int *ci, int tx, int rows);
// autotuner: build the row-grouped copies (rg 4 and 8), time their SpMM on
// the device for each panels-per-wave choice and vector precision, keep
// the fastest; decisions are cached per (rows, nonzeros, value precision).
// Rows not strictly sorted: no copy (returns 0).
int csr_group_rows (DevCSR &A, const int *rp, const int *ci, const double *v);
// force the grouped SpMM's configuration for later csr_group_rows calls:
// rg rows per group, ppw (2 or 4) panels per wave with FP64 vectors and
// twice that with FP32 (rg 0: autotune); the tests run every candidate
// through it
void spmm_group_force (int rg, int ppw);
// force the row-tile copy for later csr_group_rows calls: candidate cand
// (bits 0-1: shape, bit 2: 2D tiles over the matrix's line stride, bit 3:
// tiles in XCD runs); cand < 0: no tile forced
void spmm_tile_force (int cand);
// a configuration is forced (wide-row matrices of any size then get the copy)
bool spmm_group_forced ();
// the autotuner's decisions so far, one "n nnz vprec rg ppw64 ppw32 ms64"
// line per matrix shape
std::string spmm_group_report ();
// add the scaled values for the factor b1 to an uploaded square matrix with
// its dinv (same host CSR)
int csr_scale_cols (DevCSR &A, const int *rp, const int *ci, const double *v,
                    double b1);

struct PMat {
  void *p = nullptr;
  int nrows = 0, ncols = 0, W = 64, prec = F64;
  int npanels () const { return (ncols + W - 1) / W; }
  size_t elems () const { return (size_t) npanels () * nrows * W; }
  size_t bytes () const { return elems () * prec_bytes (prec); }
  void *panel (int q) const
  { return (char *) p + (size_t) q * nrows * W * prec_bytes (prec); }
};
int  pmat_alloc (PMat &m, int nrows, int ncols, int W, int prec);  // zeroed
void pmat_free (PMat &m);

// ---- kernels (all on stream(); asynchronous) ----------------------------
// hashed Gaussian block: element (i, j) = the reference's draw
// lgh_randn_at(seed, row0 + i, col0 + j) bit for bit (device arithmetic,
// the few draws that depend on the platform log's rounding finished with
// it on the host; synchronizes the stream when there are any)
int k_randn (PMat &out, uint64_t seed, int64_t row0, int64_t col0);
void randn_host_release ();        // the thread's pinned list buffer
// the same fill into caller memory (column-major, leading dimension ld,
// FP64 or FP32) with no allocation: the flag list lives in the static
// scratch array (see kernels.hip for the limits)
int randn_fill_noalloc (uint64_t seed, int64_t row0, int64_t col0, int nr,
                        int nc, int prec, void *out, int ld);
// Y = alpha A X + beta Y over all panels of X/Y (same W, prec of X/Y)
int k_spmm (const DevCSR &A, double alpha, const PMat &X, double beta,
            PMat &Y);
// Y = diag(d) X  (d: n doubles, fp64 device vector); Y may alias X
int k_diag_scale (const PMat &X, const double *d, PMat &Y);
// Y = a X + b Y elementwise (same shape)
int k_axpby (double a, const PMat &X, double b, PMat &Y);
// Y = X converted to Y's precision / panel width (same nrows/ncols), with
// column offsets: Y(:, yc0 + j) = X(:, xc0 + j), j < nc
int k_copy_cols (const PMat &X, int xc0, PMat &Y, int yc0, int nc);
// R = B - A X
int k_resid (const DevCSR &A, const PMat &B, const PMat &X, PMat &R);
// fused Chebyshev step (see kernels.hip):
//   Zj = dinv_j R_j (jacobi) or Zsrc_j;  Dn = a D + b Z;
//   X = X + Dn (x_mode 1) | X = Dn (x_mode 0);  Rn = R - A Dn
// (D may be null when a == 0)
// Zsrc may be BF16 beside FP32 R (W = 64); Rpc (optional, Zsrc's
// precision) receives a copy of Rn for the next preconditioner application
int k_cheb_step (const DevCSR &A, int jacobi, const PMat *Zsrc, double a,
                 double b, const PMat *D, const PMat &R, PMat &X, int x_mode,
                 PMat &Dn, PMat &Rn, PMat *Rpc = nullptr);
// mixed step: FP64 arithmetic (A values FP64), R and X stored in FP64 or
// FP32, FP32 D; the preconditioned residual Zpc and the copy Rnpc of the
// new residual (the next preconditioner input) are FP32 or BF16.  The new
// direction is rounded to FP32 once and X and R move by exactly the stored
// value.
// x_mode: how X takes the new direction Dn (D = the previous direction)
enum {
  MIX_X_SET = 0,      // X = Dn
  MIX_X_ADD = 1,      // X = X + Dn
  MIX_X_SKIP = 2,     // X untouched (Dn is added at the next step)
  MIX_X_SET2 = 3,     // X = D + Dn            (after a skipped first step)
  MIX_X_ADD2 = 4      // X = (X + D) + Dn      (after a skipped step)
};
// Xin: where an adding mode reads X (null: X itself; the early phase's
// FP32 X at the first update after the residual replacement, X FP64);
// xscale: X is written multiplied by xscale[row] (FP64 X; the last step)
int k_cheb_step_mixed (const DevCSR &A, const PMat &Zpc, double a, double b,
                       const PMat *D32, const PMat &R, PMat &X, int x_mode,
                       PMat &Dn32, PMat &Rn, PMat &Rnpc,
                       const PMat *Xin = nullptr,
                       const double *xscale = nullptr);
// the mixed solve's residual replacement in one pass: R = B - A X for X
// the FP32 X32 widened (formed in FP64, the sums of k_resid; stored in R's
// precision, FP64 or FP32), Rpc = R in Rpc's storage precision (FP32 or
// BF16, rounded from the FP64 value as k_copy_cols rounds)
int k_mix_switch (const DevCSR &A, const PMat &B, const PMat &X32, PMat &R,
                  PMat &Rpc);
// degree-2 Chebyshev-Jacobi smoother in one pass (W = 64):
// D1 = b1 dinv R0, R1 = R0 - A D1; D2 = a2 D1 + b2 dinv R1;
// X = (xadd ? X : 0) + D1 + D2 (D1, R1 recomputed at the neighbours)
int k_smooth2x (const DevCSR &A, double b1, double a2, double b2,
                const PMat &R0, PMat &X, int xadd);
// X0 = Ainv B0 (dense n0 x n0 row-major inverse in X's precision)
int k_dense_apply (const void *Ainv, int n0, const PMat &B, PMat &X);

// the coarse end of a multigrid cycle in one launch: levels 0 .. nlev - 1
// (level 0 the dense solve), the operations in ops[] (kind << 4 | level,
// kinds MGC_*), each a pass over one level's rows by one workgroup per 16
// bytes of a panel row.  Every row's arithmetic is that of k_smooth2x,
// k_resid, k_spmm (spmm64_k) and k_dense_apply in the same order, so the
// result equals the per-level launches bit for bit.
enum MgcOp {
  MGC_SM0 = 0,        // X = smoother(B), zero initial guess
  MGC_SMX = 1,        // X += smoother(R)
  MGC_RESID = 2,      // R = B - A X
  MGC_RESTRICT = 3,   // B(l - 1) = P^T R
  MGC_PROLONG = 4,    // X += P X(l - 1)
  MGC_DENSE = 5       // X(0) = Ainv B(0)
};
constexpr int MGC_MAX_LEV = 6, MGC_MAX_OPS = 256;
struct MgcLevelIn {
  const DevCSR *A = nullptr, *P = nullptr, *PT = nullptr;   // P, PT: l >= 1
  double b1 = 0, a2 = 0, b2 = 0;          // degree-2 smoother coefficients
  const PMat *B = nullptr, *X = nullptr, *R = nullptr;       // R: l >= 1
};
// levels and shapes the fused launch accepts (square CSR with dinv on
// levels >= 1, vectors of one precision and row width, P / P^T rows the
// per-level SpMM sums in row order)
bool mgc_supported (int nlev, const MgcLevelIn *lev, int n0);
// the fused levels' CSR arrays, 1/diag and dense inverse packed in the
// launch's LDS layout (one contiguous copy per launch when they fit beside
// the vector slices); built on first use for a depth and kept
struct MgcImage {
  int nlev = 0;                           // fused levels (0: not built)
  void *p = nullptr;
  size_t bytes = 0;
  void release ();
};
// img, img_mu: the hierarchy's image and the lock that builds it (several
// threads may run cycles on one hierarchy)
int k_mg_coarse (int nlev, const MgcLevelIn *lev, const void *cinv, int n0,
                 int nops, const unsigned char *ops, MgcImage *img,
                 std::mutex *img_mu);

// column-wise reductions: out[j] = sum_i X(i,j) Y(i,j) (fp64, ncols)
int k_coldot (const PMat &X, const PMat &Y, double *out_dev);
// out[j] = sum_i (X(i,j) - s_j Y(i,j))^2
int k_colresid2 (const PMat &X, const PMat &Y, const double *s_dev,
                 double *out_dev);
// per-column CG updates: X += alpha_j P, R -= alpha_j Q ; P = Z + beta_j P
int k_cg_xr (const double *alpha, const PMat &P, const PMat &Q, PMat &X,
             PMat &R);
int k_cg_p (const double *beta, const PMat &Z, PMat &P);
// Y(:, j) = s_j X(:, j)  (s: fp64 device, one per padded column)
int k_colscale (const double *s, const PMat &X, PMat &Y);

// ---- tall-skinny dense (W = 64 panel matrices) -------------------------
// G (fp64, col-major, ldg) = X[:, xc] ^T Y[:, yc] over whole panels
//   xp0/npx: X panels, yp0/npy: Y panels; G is (npx*64) x (npy*64)
int k_gram (const PMat &X, int xp0, int npx, const PMat &Y, int yp0, int npy,
            double *G, int ldg, int sym);
// Y[:, panels yp0..yp0+npy) = alpha X[:, panels xp0..xp0+npx) C + beta Y
//   C: fp64 col-major (npx*64) x (npy*64), ldc
//   upper: C is block upper triangular (zero blocks p > q are skipped)
int k_gemm_nn (const PMat &X, int xp0, int npx, const double *C, int ldc,
               double alpha, double beta, PMat &Y, int yp0, int npy,
               int upper = 0);

// ---- small dense helpers on device (fp64) ------------------------------
int k_symmetrize (double *T, int n, int ld);
int k_add_diag (double *A, int n, int ld, const double *shift_dev);
// C = transpose(lower(Linv)) embedded in a zeroed (ldc x nc) block
int k_lower_t_pad (const double *Linv, int n, int ldl, double *C, int ldc,
                   int ccols);
int k_upper_pad (const double *Rinv, int n, int ldr, double *C, int ldc,
                 int ccols);
int k_trace (const double *A, int n, int ld, double *out_dev);

// own FP64 Cholesky factor (A = L L^T, or U^T U with upper) and triangular
// inverse of the factor, in place on the triangle, 64 x 64 blocks in fixed
// order (results depend only on the input); *info_dev (device int, zeroed by
// the caller) receives LAPACK's info
int k_potrf (int upper, int n, double *A, int lda, int *info_dev);
int k_trtri (int upper, int n, double *A, int lda, int *info_dev);

// dense factorizations (device pointers, fp64): potrf / trtri run the own
// kernels above, the eigensolvers rocSOLVER
int sol_potrf (int upper, int n, double *A, int lda, int *info_host);
int sol_trtri (int upper, int n, double *A, int lda, int *info_host);
int sol_syevd (int n, double *A, int lda, double *w_dev, int *info_host);
// eigenvalues only (QR iteration on the tridiagonal form: a second method
// beside syevd's divide and conquer); A is destroyed
int sol_syev_vals (int n, double *A, int lda, double *w_dev, int *info_host);
// rocBLAS C = alpha op(A) op(B) + beta C on stream(); ta/tb: 1 = transpose
int blas_dgemm (int ta, int tb, int m, int n, int k, double alpha,
                const double *A, int lda, const double *B, int ldb,
                double beta, double *C, int ldc);
// hand over an event that a cross-stream wait was enqueued on: destroyed
// after the thread's next main-stream synchronization, not before
void event_retire (hipEvent_t e);
// device memory of the current GPU in bytes (0 if unknown)
size_t dmem_total ();

}  // namespace lgh
