// ingest.cpp — B from the fit stage's row export (no reference counterpart
// in the engine's API: it replaces the lgh_fit_get_rows -> lgh_fit_get_mat
// round trip of fit.h:160-165 and glr.h:274-282 / fit_impl.hpp:893-968,
// where every entry goes through MatSetValues).  The rows are validated,
// sorted by column (duplicates summed), optionally checked for symmetry,
// and become an assembled AIJ Mat in one pass.
#include "petsc_impl.hpp"
#include "lgpsf_hessian/lgpsf_hessian.h"

#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

extern "C" int
lgh_mat_from_fit_rows (int n, const int *rowptr, const long *colgids,
                       const double *vals, double sym_tol, Mat *B)
{
  PetscCheck (rowptr != NULL && B != NULL && (rowptr[n] == 0 || (colgids && vals)),
              0, PETSC_ERR_ARG_NULL, "lgh_mat_from_fit_rows: NULL argument");
  PetscCheck (n > 0 && rowptr[0] == 0, 0, PETSC_ERR_ARG_OUTOFRANGE,
              "lgh_mat_from_fit_rows: need n > 0 and rowptr[0] == 0");
  std::vector<int> rp (n + 1, 0), ci;
  std::vector<double> v;
  ci.reserve (rowptr[n]);
  v.reserve (rowptr[n]);
  std::vector<int> idx;
  double vmax = 0.;
  for (int i = 0; i < n; i++) {
    const int p0 = rowptr[i], p1 = rowptr[i + 1];
    PetscCheck (p1 >= p0, 0, PETSC_ERR_ARG_OUTOFRANGE,
                "lgh_mat_from_fit_rows: rowptr decreases at row %d", i);
    for (int p = p0; p < p1; p++)
      PetscCheck (colgids[p] >= 0 && colgids[p] < n, 0, PETSC_ERR_ARG_OUTOFRANGE,
                  "lgh_mat_from_fit_rows: column %ld of row %d outside [0, %d)",
                  colgids[p], i, n);
    idx.resize (p1 - p0);
    std::iota (idx.begin (), idx.end (), p0);
    std::stable_sort (idx.begin (), idx.end (),
                      [&] (int a, int b) { return colgids[a] < colgids[b]; });
    for (size_t q = 0; q < idx.size (); q++) {
      const int c = (int) colgids[idx[q]];
      if (!ci.empty () && (int) ci.size () > rp[i] && ci.back () == c)
        v.back () += vals[idx[q]];
      else {
        ci.push_back (c);
        v.push_back (vals[idx[q]]);
      }
    }
    rp[i + 1] = (int) ci.size ();
  }
  for (double x : v) vmax = std::max (vmax, std::fabs (x));
  if (sym_tol >= 0.) {
    // B^T by a counting transpose; same structure and values within tol
    std::vector<int> tp (n + 1, 0), tc (ci.size ());
    std::vector<double> tv (v.size ());
    for (int c : ci) tp[c + 1]++;
    for (int i = 0; i < n; i++) tp[i + 1] += tp[i];
    std::vector<int> pos (tp.begin (), tp.end () - 1);
    for (int i = 0; i < n; i++)
      for (int p = rp[i]; p < rp[i + 1]; p++) {
        tc[pos[ci[p]]] = i;
        tv[pos[ci[p]]++] = v[p];
      }
    bool ok = tp == rp && tc == ci;
    for (size_t p = 0; ok && p < v.size (); p++)
      ok = std::fabs (v[p] - tv[p]) <= sym_tol * vmax;
    PetscCheck (ok, 0, PETSC_ERR_ARG_WRONG,
                "lgh_mat_from_fit_rows: B is not symmetric to %g relative",
                sym_tol);
  }
  return MatCreateSeqAIJWithArrays (PETSC_COMM_SELF, n, n, rp.data (),
                                    ci.data (), v.data (), B);
}
