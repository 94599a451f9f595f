/* lgh_hip.h — the thin C ABI of the gfx950 kernels and host set-up
 * routines under the GLR engine (plain pointers and sizes; no PETSc or
 * torch types).  Each kernel entry replaces one reference primitive:
 *
 *   lgh_hip_randn_fill   lgh_randn_at fills        glr_common_impl.h:441-443
 *   lgh_hip_spmm         per-column MatMult(B)     glr_common_impl.h:325-332
 *   lgh_hip_cheb_jacobi  lgh_zs_cheb_core + lgh_zs_rowscale_pc
 *                                                  zsolve_impl.h:104-152
 *   lgh_hip_gram         DGEMM("T","N") Gram / T   glr_common_impl.h:355, 481
 *   lgh_hip_gemm_nn      DGEMM("N","N") / DTRSM    glr_common_impl.h:402, 567
 *   lgh_amg_*            the PCGAMG hierarchy the reference harvests
 *                                                  zsolve_impl.h:200-356
 *
 * Host arrays in and out (column-major n x m blocks with a leading
 * dimension); the device data layout (panel matrices, see lgh_dev.hpp) is
 * internal.  All functions return 0 on success; lgh_hip_last_error gives
 * the message of the last failure.  Precision codes: 0 = FP64, 1 = FP32. */
#ifndef LGH_HIP_H
#define LGH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int         lgh_hip_init (void);
const char *lgh_hip_last_error (void);
int         lgh_hip_sync (void);
/* bytes currently held by the engine's device allocator */
int64_t     lgh_hip_bytes_in_use (void);
void        lgh_hip_trim (void);

/* live kernel timing: when enabled, every launch of the engine is bracketed
 * by HIP events on the engine stream and credited with its algorithmic
 * HBM bytes and flops.  kind: 0 B SpMM of the F sweep, 1 other SpMM,
 * 2 Jacobi Chebyshev steps and smoothers below the finest AMG level,
 * 3 fused Chebyshev step (explicit Z), 4 residual below the finest level,
 * 5 Gram (MFMA), 6 GEMM (MFMA), 7 other, 8 mixed-precision Chebyshev step
 * after the residual replacement (FP32 R, FP64 X),
 * 9 / 10 fused smoother on the finest level (pre: zero guess / post),
 * 11 residual on the finest level.  collect synchronizes and resets that
 * kind's totals; np >= 0 counts only launches over np 64-column panels,
 * prec >= 0 only launches whose vectors are stored in that precision
 * (0 FP64, 1 FP32, 2 BF16).  (No reference counterpart: measurement.) */
void lgh_hip_prof_enable (int on);
int  lgh_hip_prof_collect (int kind, int np, int prec, int64_t *count,
                           double *ms, double *bytes, double *flops);

/* achieved HBM bandwidth of a device-to-device copy of `bytes` (a 16-byte
 * vector copy kernel; read plus written bytes per second, best of reps
 * after one warm-up), the measured reference beside the nominal HBM peak.
 * (No reference counterpart: measurement.) */
int lgh_hip_copy_gbps (int64_t bytes, int reps, double *gbps);

/* plain device memory of the engine's HIP runtime, for callers that hold
 * no GPU array of their own (the device-resident lgh_glr_*_block_dev
 * calls take such blocks): alloc returns NULL on failure (message in
 * lgh_hip_last_error); put / get are synchronous host <-> device copies,
 * refusing a device side that is not device memory or is shorter than the
 * copy (one allocation must hold [ptr, ptr + bytes)).  (No reference
 * counterpart: the reference's blocks are host PETSc objects.) */
void *lgh_hip_dev_alloc (int64_t bytes);
void  lgh_hip_dev_free (void *p);
int   lgh_hip_dev_put (void *dst, const void *src, int64_t bytes);
int   lgh_hip_dev_get (void *dst, const void *src, int64_t bytes);

/* the device hashed-Gaussian arithmetic evaluated on the host (no GPU):
 * the draw with the correctly rounded log(q) */
double lgh_hip_randn_at_host (uint64_t seed, int64_t i, int64_t j);
/* the same over an nr x nc block (column-major, ld), in precision prec
 * (0 FP64, 1 FP32, returned as FP64), with flags(r, c) = 1 where the device
 * fill finishes the draw with the platform log because its value may depend
 * on that log's rounding (no GPU; no reference counterpart: it lets the
 * flag window be checked against the platform log on the host) */
int lgh_hip_randn_host_block (uint64_t seed, int64_t row0, int64_t col0,
                              int nr, int nc, int prec, double *vals, int ld,
                              unsigned char *flags, int ldf);

/* out(r, c) = lgh_randn_at(seed, row0 + r, col0 + c), the reference's draw
 * bit for bit, generated on the device with panel width W in precision
 * prec, returned as FP64 */
int lgh_hip_randn_fill (uint64_t seed, int64_t row0, int64_t col0, int nr,
                        int nc, int W, int prec, double *out, int ld);

/* No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * allocation).  The draws of lgh_hip_randn_fill, bit for bit, written into
 * caller device memory `out` (column-major, leading dimension ld >= nr;
 * prec 0 FP64, 1 FP32); rows nr .. ld-1 are left as they are.  Nothing is
 * allocated: the list of draws finished with the platform log lives in a
 * static 32 MiB device array per GPU (held until the process ends).  At
 * most 33554432 draws a call, larger requests are refused.  Call
 * lgh_hip_init on the calling thread first.  Calls that overlap on one GPU
 * race (they share the array).  Returns after the draws are written. */
int lgh_hip_randn_fill_noalloc (uint64_t seed, int64_t row0, int64_t col0,
                                int nr, int nc, int prec, void *out, int ld);

/* Y = A X (A: n x m CSR) on ncols columns */
int lgh_hip_spmm (int n, int m, const int *rp, const int *ci, const double *v,
                  int prec_vals, int prec_vec, int W, int ncols,
                  const double *X, int ldx, double *Y, int ldy);
/* the same through the row-grouped copy of A that the engine keeps for B
 * (rows in groups of 4 or 8, each group's union of column indices walked
 * once; W = 64; rows with strictly increasing column indices, else the
 * plain kernel runs): the same bits as lgh_hip_spmm.  The copy's rows per
 * group and the panels per wavefront are chosen by the autotuner, which
 * times the candidates on A on the device (no reference counterpart: the
 * reference multiplies B column by column, glr_common_impl.h:325-332, and
 * its only tunable is the prior's tile, prior.h:80) */
int lgh_hip_spmm_grouped (int n, int m, const int *rp, const int *ci,
                          const double *v, int prec_vals, int prec_vec,
                          int W, int ncols, const double *X, int ldx,
                          double *Y, int ldy);
/* force the grouped product's configuration for matrices uploaded later
 * (rg 4 or 8 rows per group, ppw 2 or 4 panels per wavefront with FP64
 * vectors, twice that with FP32; while forced, every wide-row B of the
 * engine gets the grouped copy, whatever its size; rg = 0 returns to the
 * autotuner and the size threshold of 2^24 nonzeros) */
void lgh_hip_spmm_group_config (int rg, int ppw);
/* force the row-tile copy instead (the LDS-staged product: tiles of 64
 * rows whose union of neighbour rows is staged in LDS chunk by chunk, the
 * same bits): cand bits 0-1 pick the shape (rows per wavefront x
 * wavefronts x chunk rows: 4x16x32, 4x16x64, 8x8x32, 8x8x64), bit 2 2D
 * tiles (16 rows wide, on the matrix's line stride found from its
 * pattern; 1D when it has none), bit 3 tiles dealt to the XCDs in
 * contiguous runs; cand < 0 ends the forcing (the autotuner, which times
 * every tile candidate beside the grouped ones, decides again) */
void lgh_hip_spmm_tile_config (int cand);
/* caps of the multigrid levels that run as one fused launch per coarse
 * visit of a cycle (the small levels above the dense solve, each with at
 * most max_rows rows and max_nnz nonzeros; negative: the defaults 1024
 * and 65536; max_rows = 0: one launch per operation and level).  Both
 * paths give the same bits; this selects the launch shape.  Change it only
 * while no other thread runs solves (a fused launch keeps a packed copy of
 * its levels' arrays, rebuilt when the fused depth changes). */
void lgh_hip_mg_coarse_config (int max_rows, int max_nnz);
/* the autotuner's decisions so far as text, one line per matrix shape:
 * "rows nonzeros value_prec rg ppw_fp64 ppw_fp32 ms_per_panel_fp64
 * ms_per_panel_fp32 tile tile_rg tile_nw tile_ch tile_tx tile_stride
 * tile_xcd" (tile = 1: the row-tile copy won, its shape follows; the
 * times are the winner's); returns the full length (buf may be null) */
int lgh_hip_spmm_group_report (char *buf, int len);

/* X = Chebyshev(nit) solve of A X = B with the Jacobi preconditioner on
 * the interval [lmin, lmax], zero initial guess, one fused kernel per
 * iteration */
int lgh_hip_cheb_jacobi (int n, const int *rp, const int *ci, const double *v,
                         int prec, int W, double lmin, double lmax, int nit,
                         int ncols, const double *B, int ldb, double *X,
                         int ldx);

/* G (a x b) = X^T Y, X: n x a, Y: n x b */
int lgh_hip_gram (int prec, int n, int a, const double *X, int ldx, int b,
                  const double *Y, int ldy, double *G, int ldg);
/* Y (n x b) = X C, X: n x a, C: a x b (a C that is zero below its 64 x 64
 * block diagonal takes the triangular product, skipping those blocks) */
int lgh_hip_gemm_nn (int prec, int n, int a, const double *X, int ldx, int b,
                     const double *C, int ldc, double *Y, int ldy);

/* ---- host AMG hierarchy (no GPU needed) ------------------------------ */
typedef struct lgh_amg lgh_amg_t;
int  lgh_amg_build (int n, const int *rp, const int *ci, const double *v,
                    int smax_krylov, double smoother_top, int nu_force,
                    lgh_amg_t **out);
void lgh_amg_free (lgh_amg_t *h);
int  lgh_amg_levels (const lgh_amg_t *h);
/* level l (0 = coarsest): operator CSR, interpolation from l-1 (l > 0),
 * smoother interval and degree; pointers are borrowed from h */
int  lgh_amg_level (const lgh_amg_t *h, int l, int *n, int *nc,
                    const int **rp, const int **ci, const double **v,
                    const int **prp, const int **pci, const double **pv,
                    double *slo, double *shi, int *nu);
/* cycle index of the device cycle: 1 V-cycle (default), 2 W-cycle (each
 * level between the finest and the coarsest corrects twice, the second
 * time from the first result); lgh_prior_create_mat picks it for its own
 * hierarchy (lgh_prior_info_t.cycle) */
int  lgh_amg_set_cycle (lgh_amg_t *h, int gamma);
/* X = cycle(B) on the device, ncols columns; prec 0 FP64, 1 FP32,
 * 2 BF16 vectors with FP32 matrices and arithmetic (W = 64, 16 or 8) */
int  lgh_amg_vcycle (const lgh_amg_t *h, int prec, int W, int ncols,
                     const double *B, int ldb, double *X, int ldx);

#ifdef __cplusplus
}
#endif
#endif
