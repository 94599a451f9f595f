"""Writes the golden fixtures under tests/golden/ (run once; outputs are
committed).  Independent arithmetic only — nothing here calls the oracle:

* tg_eigs_{spd,indef}.npy  — LAPACK dsyevd (numpy.linalg.eigvalsh) of the
  dense F = M^{1/2} Z^{-1} B Z^{-1} M^{1/2} of tests/test_glr.cpp:91-123,
  sorted |lambda|-descending with the same selection sort (:115-122).
* tpm_xref.npy — direct sparse LU solve (scipy) of the 24 x 24 Laplacian +
  0.5 I system of tests/test_prior_mat.cpp:47-73, 119-127 (stands in for
  the reference's CG + GAMG solve at rtol 1e-13).
* randn_block.npz — the hashed Gaussian of glr_common_impl.h:35-73
  restated with Python arbitrary-precision integers masked to 64 bits and
  Python floats (IEEE double, libm log/sqrt): rows 0..63 x cols 0..7 for
  three seeds, plus column 0 of rows 0..4095 for the prior probe seed.
"""
import math
import os
import sys

import numpy as np
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "lgpsf-hessian_amd"))
import workloads as W  # noqa: E402

M64 = (1 << 64) - 1


def py_randn_at(seed, i, j):
    z = (seed + 0x9E3779B97F4A7C15 * (i + 1) + 0xC2B2AE3D27D4EB4F * (j + 1)) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    z = z ^ (z >> 31)
    s = z if z else 0x853C49E6748FEA9B
    while True:
        x = s
        x ^= x >> 12
        x ^= (x << 25) & M64
        x ^= x >> 27
        s = x
        u1 = float(((x * 2685821657736338717) & M64) >> 11) / 9007199254740992.0
        x = s
        x ^= x >> 12
        x ^= (x << 25) & M64
        x ^= x >> 27
        s = x
        u2 = float(((x * 2685821657736338717) & M64) >> 11) / 9007199254740992.0
        u1 = 2.0 * u1 - 1.0
        u2 = 2.0 * u2 - 1.0
        q = u1 * u1 + u2 * u2
        if 0.0 < q < 1.0:
            return u1 * math.sqrt(-2.0 * math.log(q) / q)


def sel_sort_absdesc(w):
    w = list(w)
    n = len(w)
    for i in range(n):
        for j in range(i + 1, n):
            if abs(w[j]) > abs(w[i]):
                w[i], w[j] = w[j], w[i]
    return np.array(w)


def main():
    for spd, name in ((True, "spd"), (False, "indef")):
        F = W.tg_dense_F(spd)
        w = np.linalg.eigvalsh(F)
        np.save(os.path.join(HERE, "tg_eigs_%s.npy" % name), sel_sort_absdesc(w))
    Z = W.tpm_Z()
    b = W.tg_probe(Z.shape[0])
    np.save(os.path.join(HERE, "tpm_xref.npy"), spla.spsolve(Z.tocsc(), b))
    seeds = [0, 271828, 0xB0BD5]
    blk = np.array([[[py_randn_at(s, i, j) for j in range(8)] for i in range(64)]
                    for s in seeds])
    col = np.array([py_randn_at(0xB0BD5, i, 0) for i in range(4096)])
    np.savez(os.path.join(HERE, "randn_block.npz"), seeds=np.array(seeds, dtype=np.uint64),
             block=blk, probe_col=col)


if __name__ == "__main__":
    main()
