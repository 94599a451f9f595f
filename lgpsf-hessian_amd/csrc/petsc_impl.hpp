// petsc_impl.hpp — object layouts of the PETSc-subset Vec / Mat.
#pragma once
#include <petsc.h>

#include <map>
#include <vector>

namespace lgh { struct DevCSR; }

struct lgh_obj_head { int refct = 1; };

enum { LGH_MAT_AIJ = 0, LGH_MAT_DENSE = 1 };

struct _p_Vec : lgh_obj_head {
  int n = 0;
  double *a = nullptr;
  std::vector<double> own;
};

struct _p_Mat : lgh_obj_head {
  int type = LGH_MAT_AIJ;
  int m = 0, n = 0;
  int assembled = 0;
  long state = 0;                       // bumped on every write access
  // AIJ
  std::vector<std::map<int, double>> rows;   // builder before assembly
  std::vector<int> rp, ci;
  std::vector<double> v;
  // DENSE (column-major, lda)
  std::vector<double> dense;
  double *a = nullptr;
  int lda = 0;
  // device mirror per precision (F64, F32)
  lgh::DevCSR *dev[2] = { nullptr, nullptr };
  long dev_state[2] = { -1, -1 };
  ~_p_Mat ();
};

namespace lgh {
// CSR view of an assembled Mat (dense Mats are expanded into the vectors)
int mat_host_csr (Mat A, const int **rp, const int **ci, const double **v,
                  std::vector<int> &rps, std::vector<int> &cis,
                  std::vector<double> &vs);
// cached device CSR of A at precision prec
int mat_device_csr (Mat A, int prec, int want_dinv, DevCSR **out);
}
