// amg.hpp — host-built smoothed-aggregation hierarchy for Z (the AMG that
// the reference obtains from PETSc GAMG: zsolve_impl.h:200-356 harvests it
// read-only).  Built once on the host, uploaded once, applied on the GPU.
#pragma once
#include <stdint.h>
#include <vector>

namespace lgh {

struct HostCSR {
  int n = 0, m = 0;
  std::vector<int> rp, ci;
  std::vector<double> v;
  int64_t nnz () const { return rp.empty () ? 0 : rp[n]; }
};

struct AmgOpts {
  double strength = 0.0;      // |a_ij| >= theta sqrt(|a_ii a_jj|)
  int coarse_max = 64;        // stop coarsening at <= this many unknowns
  int max_levels = 25;
  int smax_krylov = 1;        // 1: CG-Lanczos (40 its), 0: power (20 its)
  double smoother_top = 1.1;  // smoothing interval [0.1, top] * smax
  int nu_force = 0;           // 0: degree 2
};

struct AmgLevel {
  HostCSR A;                  // level operator
  HostCSR P;                  // interpolation from level l-1 (l > 0)
  std::vector<double> dinv;
  double smax = 0, slo = 0, shi = 0;
  int nu = 2;
};

struct AmgHierarchy {
  std::vector<AmgLevel> lev;  // lev[0] coarsest (PCMG numbering)
  std::vector<double> coarse_inv;   // dense row-major inverse of lev[0].A
  // cycle index: 1 V-cycle; 2 W-cycle (every level between the finest and
  // the coarsest corrects its coarse level twice, the second time from
  // the first correction's result)
  int gamma = 1;
};
// bytes moved per right-hand-side column by one application of the cycle
// (vectors in BF16, the level CSRs amortized over 64-column panels, and a
// launch-overhead term per level visit): the cost model that picks the
// cycle index
double amg_cycle_cost (const AmgHierarchy &h, int gamma);

// y = A x
void host_spmv (const HostCSR &A, const double *x, double *y);
HostCSR host_transpose (const HostCSR &A);
HostCSR host_spgemm (const HostCSR &A, const HostCSR &B);
// extreme eigenvalues of a symmetric tridiagonal (Sturm bisection)
void tridiag_extremes (int n, const double *d, const double *e, double *lo,
                       double *hi);
// host PCG with Jacobi and Lanczos extreme estimates (KSPCG +
// KSPComputeExtremeSingularValues semantics); returns iterations
int host_cg_lanczos (const HostCSR &A, const double *dinv, const double *b,
                     double rtol, int maxit, double *emin, double *emax);
int amg_setup (const HostCSR &A, const AmgOpts &o, AmgHierarchy &h);

}  // namespace lgh
