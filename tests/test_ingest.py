"""B ingest from the fit stage's row export (lgh_mat_from_fit_rows): host
code, runs without a GPU.  Pinned against scipy on the same triplets."""
import numpy as np
import pytest
import scipy.sparse as sp

import workloads as W


def _rows_from(A, rng, dup=False):
    """the fit's CSR with each row's entries shuffled (and optionally one
    entry per row split into two duplicates)"""
    A = sp.csr_matrix(A)
    rp, cg, v = [0], [], []
    for i in range(A.shape[0]):
        cols = A.indices[A.indptr[i]:A.indptr[i + 1]].astype(np.int64)
        vals = A.data[A.indptr[i]:A.indptr[i + 1]].copy()
        if dup and len(cols):
            cols = np.append(cols, cols[0])
            vals = np.append(vals, 0.25 * vals[0])
            vals[0] *= 0.75
        perm = rng.permutation(len(cols))
        cg.extend(cols[perm]); v.extend(vals[perm])
        rp.append(len(cg))
    return np.array(rp), np.array(cg, dtype=np.int64), np.array(v)


def test_rows_to_mat_sorted_and_summed(lgh):
    B, _, _ = W.psf2d_problem(16)
    rng = np.random.default_rng(0)
    rp, cg, v = _rows_from(B, rng, dup=True)
    M = lgh.Mat.from_fit_rows(rp, cg, v, sym_tol=1e-14)
    C = M.csr()
    ref = sp.csr_matrix(B)
    ref.sort_indices()
    assert np.array_equal(C.indptr, ref.indptr)
    assert np.array_equal(C.indices, ref.indices)
    assert np.max(np.abs(C.data - ref.data)) <= 1e-15 * np.abs(ref.data).max()


def test_rows_refusals(lgh):
    B, _, _ = W.psf2d_problem(8)
    rng = np.random.default_rng(1)
    rp, cg, v = _rows_from(B, rng)
    bad = cg.copy()
    bad[3] = B.shape[0]                                  # column out of range
    with pytest.raises(lgh.LGHError):
        lgh.Mat.from_fit_rows(rp, bad, v)
    asym = v.copy()
    asym[0] *= 1.5                                        # breaks symmetry
    with pytest.raises(lgh.LGHError):
        lgh.Mat.from_fit_rows(rp, cg, asym, sym_tol=1e-12)
    lgh.Mat.from_fit_rows(rp, cg, asym)                  # no check requested
    with pytest.raises(lgh.LGHError):
        lgh.Mat.from_fit_rows(rp[::-1].copy(), cg, v)    # rowptr not monotone
