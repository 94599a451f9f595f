// zsolve.hpp — device Z-solve tiers: fused Chebyshev-Jacobi, the block
// V-cycle on the host-built hierarchy, Chebyshev around the V-cycle, and
// block PCG (per-column coefficients) for the vector-solve path.
#pragma once
#include "lgh_dev.hpp"
#include "amg.hpp"

namespace lgh {

struct DevMG {
  int prec = F64;
  int nl = 0;
  int gamma = 1;                       // cycle index (AmgHierarchy::gamma)
  std::vector<DevCSR *> A, P, PT;      // per PCMG level (P, PT null at 0)
  std::vector<double> slo, shi;
  std::vector<int> nu, n;
  void *cinv = nullptr;                // dense row-major inverse of A_0
  int n0 = 0;
  MgcImage mgc_img;                    // the fused coarse launch's LDS image
  std::mutex mgc_mu;
  // per-level workspace for one (W, npanels) shape; slot 0 holds full
  // 64-column panels, slot 1 16-wide ones, slot 2 other widths (the
  // narrow last panels of blocks), so solves that alternate between them
  // reallocate nothing
  struct Work {
    int W = 0, np = -1;
    std::vector<PMat> R1, R2, D1, D2, Bl, Xl;
    void free ();
  } work[3];
  ~DevMG ();
  void free_work ();
};

int mg_upload (const AmgHierarchy &h, int prec, DevMG **out);
// X = cycle(B) at the finest level (zero initial guess; V- or W-cycle by
// g.gamma)
int mg_vcycle (DevMG &g, const PMat &B, PMat &X);
// caps of the levels the fused coarse launch (k_mg_coarse) takes: rows and
// nonzeros per level (negative: the defaults 1024 and 65536; rows 0: the
// per-level launches everywhere)
void mg_coarse_config (int max_rows, int max_nnz);

struct ChebWork {
  int W = 0, np = -1, prec = -1, nrows = 0;
  PMat R[2], D[2], Z;
  PMat Rpc, Zpc;                       // lower-precision V-cycle in / out
  PMat Bt, Xt;                         // the narrow panels' B and X (48 columns)
  ChebWork *tail = nullptr, *tail8 = nullptr;   // (see ChebMixedWork)
  ~ChebWork ();
  int ensure (const PMat &like);
};

// Both Chebyshev solves below run a block whose last 64-column panel holds
// at most 48 columns as its whole panels plus up to three 16-wide panels,
// or one 8-wide panel when it holds at most 8 (tail, tail8: their own work
// and copies of those columns, solved on the thread's auxiliary stream
// beside the whole panels), so the padding columns cost little; every
// column sees the same arithmetic at any width.
//
// X = Chebyshev(nit) solve of A X = B, zero initial guess
// (zsolve_impl.h:104-141 with the blocked-asolve contract of :629-633);
// preconditioner: the V-cycle when mg (in B's precision, or BF16 around
// FP32 vectors at W = 64), else Jacobi (A.dinv)
int cheb_solve (const DevCSR &A, DevMG *mg, double lmin, double lmax,
                int nit, const PMat &B, PMat &X, ChebWork &w);

// FP64-accurate Chebyshev around an FP32 or BF16 V-cycle (W = 64 panels):
// FP64 arithmetic; D in FP32; the V-cycle's input and output (Rpc, Zpc) in
// its own precision (k_cheb_step_mixed).  The first steps (an even count,
// about half) keep R and X in FP32; then X is widened, the residual
// B - A X is formed again in FP64 and the remaining steps (which damp the
// early rounding by their convergence factor) run with FP64 X and R
// stored in FP32 (its rounding is relative to the residual left by then).
// Solves of fewer than 8 steps keep FP64 R and X throughout.
struct ChebMixedWork {
  int W = 0, np = -1, nrows = 0, pprec = -1;
  PMat R[2], D[2], Rpc, Zpc;           // R: FP64 residuals (solves of fewer
                                       // than 8 steps only; made on demand)
  PMat R32[2], X32;                    // FP32 R (every step of a solve of 8
                                       // or more) and X (its early phase)
  PMat Bt, Xt;                         // the narrow panels' B and X (48 columns)
  ChebMixedWork *tail = nullptr, *tail8 = nullptr;
  ~ChebMixedWork ();
  int ensure (const PMat &like, int pc);
  int ensure64 ();                     // R[0], R[1] at the current shape
};
// xscale (n FP64 values or null): X is returned row-scaled, diag(xscale) X,
// with the bits of k_diag_scale applied after the solve
int cheb_solve_mixed (const DevCSR &A64, DevMG &mg, double lmin,
                      double lmax, int nit, const PMat &B, PMat &X,
                      ChebMixedWork &w, const double *xscale = nullptr);

// Work buffers of block Z-solves that a thread brings itself, so that
// several threads can solve with one prior at once (the serving loop's
// device threads): while a SolveWorkScope is alive on a thread, the
// prior's block solves and the V-cycles they run take their buffers from
// it instead of from the prior and the hierarchy.
struct SolveWork {
  ChebWork cw[2];
  ChebMixedWork cwm;
  std::vector<std::pair<const DevMG *, DevMG::Work *>> mg;   // 3 slots each
  DevMG::Work *mg_slots (const DevMG *g);
  ~SolveWork ();
};
SolveWork *solve_work ();              // the thread's, or null
struct SolveWorkScope {
  SolveWork *prev;
  explicit SolveWorkScope (SolveWork *w);
  ~SolveWorkScope ();
};

struct PcgInfo {
  int its = 0;
  int converged = 0;
  std::vector<double> alpha, beta;     // column 0 history (Lanczos)
};
// per-column PCG to ||z|| <= rtol ||z0|| (preconditioned norm, as KSPCG),
// preconditioner: V-cycle (mg) or Jacobi (dinv64: fp64 device vector)
int pcg_solve (const DevCSR &A, DevMG *mg, const double *dinv64,
               const PMat &B, PMat &X, double rtol, int maxit, PcgInfo *info);

}  // namespace lgh
