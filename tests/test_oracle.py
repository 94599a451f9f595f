"""The FP64 CPU oracle pinned against golden fixtures (numpy/LAPACK dense
oracle of tests/test_glr.cpp, scipy direct solve for
tests/test_prior_mat.cpp, an independent Python-integer restatement of the
hashed Gaussian) and against the identity gates of tests/test_glr.cpp at
the reference's own tolerances.  CPU only."""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN
import workloads as W


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


# ---- hashed Gaussian ------------------------------------------------------

def test_randn_matches_independent_restatement(oracle):
    g = np.load(os.path.join(GOLDEN, "randn_block.npz"))
    for s, blk in zip(g["seeds"], g["block"]):
        got = oracle.randn_fill(int(s), 0, 0, 64, 8)
        assert np.array_equal(got, blk)            # bit-exact
    col = oracle.randn_fill(0xB0BD5, 0, 0, 4096, 1)[:, 0]
    assert np.array_equal(col, g["probe_col"])


def test_randn_partition_and_extension_invariance(oracle):
    full = oracle.randn_fill(271828, 0, 0, 300, 20)
    # row slabs (partition independence) and column continuation (extend)
    a = oracle.randn_fill(271828, 0, 0, 117, 20)
    b = oracle.randn_fill(271828, 117, 0, 183, 20)
    assert np.array_equal(np.vstack([a, b]), full)
    c = oracle.randn_fill(271828, 0, 13, 300, 7)
    assert np.array_equal(c, full[:, 13:])


def test_randn_moments(oracle):
    x = oracle.randn_fill(0, 0, 0, 20000, 10).ravel()
    assert abs(x.mean()) < 0.01
    assert abs(x.var() - 1.0) < 0.01
    # fourth moment of N(0,1) is 3; KS-style check on the CDF at a few points
    assert abs(np.mean(x ** 4) - 3.0) < 0.06
    for t, p in ((-1.0, 0.158655), (0.0, 0.5), (1.5, 0.933193)):
        assert abs(np.mean(x <= t) - p) < 0.006


# ---- dense kernels --------------------------------------------------------

def test_dense_kernels(oracle):
    rng = np.random.default_rng(1)
    X = rng.standard_normal((200, 37))
    G = X.T @ X
    L, info = oracle.potrf_l(G)
    assert info == 0
    assert np.allclose(L, np.linalg.cholesky(G), rtol=1e-12, atol=1e-12)
    Y = oracle.trsm_rltn(L, X)
    assert np.allclose(Y.T @ Y, np.eye(37), atol=1e-12)
    A = rng.standard_normal((60, 60))
    A = A + A.T
    w, V, info = oracle.syev(A)
    assert info == 0
    assert np.allclose(w, np.linalg.eigvalsh(A), rtol=0, atol=1e-12 * np.abs(w).max())
    assert np.allclose(A @ V, V * w, atol=1e-12 * np.abs(w).max())
    assert np.allclose(V.T @ V, np.eye(60), atol=1e-13)
    # a rank-deficient Gram must fail potrf (the jitter path's trigger)
    Xd = np.hstack([X[:, :5], X[:, :1]])
    _, info = oracle.potrf_l(Xd.T @ Xd)
    assert info != 0


def test_cheb_count_closed_form(oracle):
    for lmin, lmax, tol in ((0.5, 2.0, 1e-10), (1e-3, 4.0, 1e-8), (0.9, 1.1, 1e-12)):
        k = lmax / lmin
        rho = (math.sqrt(k) - 1) / (math.sqrt(k) + 1)
        assert oracle.cheb_count(lmin, lmax, tol) == math.ceil(math.log(2 / tol) / math.log(1 / rho))


# ---- tests/test_glr.cpp gates ---------------------------------------------

@pytest.mark.parametrize("spd", [True, False])
@pytest.mark.parametrize("backend", [1, 2])
def test_glr_gates(oracle, spd, backend):
    """backend 2: the at-scale backend's build (BCGS2 orthonormalization,
    no residual sweep) against the same dense closed form."""
    c = 0.7
    B = oracle.CSR.from_dense(W.tg_B_exact(spd=spd))
    dense = np.load(os.path.join(GOLDEN, "tg_eigs_%s.npy" % ("spd" if spd else "indef")))
    zs = oracle.ZSolver.diag(W.tg_z())
    g = oracle.GLR(B, W.tg_m(), zs, ell=40, q_power=1, trunc_rel=1e-8, check=1,
                   backend=backend)
    rep = g.report
    assert rep["kept"] == 10
    lam = g.lam
    assert np.max(np.abs(lam - np.abs(dense[:10])) / np.abs(dense[:10])) < 1e-8
    assert rep["next_abs"] < 1e-6 * rep["lam_abs_max"]
    assert rep["resid_max"] < 1e-7 * rep["lam_abs_max"]
    if backend == 2:
        assert rep["resid_max"] == 0.0
        # the two orthonormalizations span the same space: the same Ritz values
        g1 = oracle.GLR(B, W.tg_m(), zs, ell=40, q_power=1, trunc_rel=1e-8)
        assert np.max(np.abs(lam - g1.lam) / np.abs(g1.lam)) < 1e-12
    if not spd:
        assert rep["n_negative_raw"] >= 5
    ref = np.sum(np.log(np.abs(dense[:10]) + c) - np.log(c))
    assert abs(g.logdet(c) - ref) < 1e-9 * abs(ref)
    v = W.tg_probe()
    if spd:
        y = g.apply(c, v)
        assert rel(g.solve(c, y), v) < 1e-8
        x = g.factor(c, 1, v)
        assert rel(g.factor(c, 0, x), y) < 1e-8
        y2 = g.factor(c, 1, v)
        x2 = g.factor(c, 3, y2)
        assert rel(x2, v) < 1e-8
        assert rel(g.sample(c, y2), x2) < 1e-14
        assert rel(g.pow(c, 1.0, g.pow(c, -1.0, v)), v) < 1e-10
        fl = 1.0 / np.sqrt(lam + c)
        assert rel(g.filter(fl, 1 / math.sqrt(c), v), g.pow(c, -0.5, v)) < 1e-12
    for incr in (False, True):
        g2 = oracle.GLR(B, W.tg_m(), zs, ell=30, q_power=1, trunc_rel=1e-8)
        g2.extend(10, incremental=incr, panel=16)
        assert g2.kept == g.kept
        assert np.max(np.abs(g2.lam - lam) / np.abs(lam)) < 1e-10


def test_glr_refusals(oracle):
    B = oracle.CSR.from_dense(W.tg_B_exact())
    zs = oracle.ZSolver.diag(W.tg_z())
    with pytest.raises(RuntimeError):
        oracle.GLR(B, W.tg_m(), zs, ell=121)
    g = oracle.GLR(B, W.tg_m(), zs, ell=40, trunc_rel=1e-8)
    with pytest.raises(ValueError):
        g.solve(0.0, W.tg_probe())
    with pytest.raises(RuntimeError):
        g.extend(0)
    with pytest.raises(RuntimeError):
        g.extend(200)


# ---- tests/test_prior_mat.cpp gates ---------------------------------------

def test_prior_mat_gates(oracle):
    Z = W.tpm_Z()
    n = Z.shape[0]
    A = oracle.CSR.from_scipy(Z)
    b = W.tg_probe(n)
    xref = np.load(os.path.join(GOLDEN, "tpm_xref.npy"))
    # vector solve (CG to 1e-12) and the interval estimate for mode 4
    zs0 = oracle.ZSolver.mat(A, mode=0, ksp_rtol=1e-12)
    assert rel(zs0.solve_vec(b), xref) < 1e-8
    _, its, emin, emax = oracle.cg(A, oracle.randn_fill(0xB0BD5, 0, 0, n, 1)[:, 0],
                                   rtol=1e-12, maxit=200)
    lmin, lmax = 0.9 * emin, 1.1 * emax
    dinv = 1.0 / Z.diagonal()
    ev = np.linalg.eigvalsh((sp.diags(np.sqrt(dinv)) @ Z @ sp.diags(np.sqrt(dinv))).toarray())
    assert lmin < ev[0] and lmax > ev[-1]        # margined interval brackets
    nit = oracle.cheb_count(lmin, lmax, 1e-10)
    zs4 = oracle.ZSolver.mat(A, mode=4, lmin=lmin, lmax=lmax, nit=nit)
    X = np.stack([b * (1 + 0.1 * j) for j in range(11)], 1)
    Y = zs4.solve_block(X)
    worst = max(rel(Y[:, j] / (1 + 0.1 * j), xref) for j in range(11))
    assert worst < 1e-7
    s1 = zs0.apply(b)[:, 0]
    assert rel(zs0.solve_vec(s1), b) < 1e-8


def test_thread_count_does_not_change_results(oracle, lgh):
    """The column-parallel sweeps (block Chebyshev-MG and Chebyshev-Jacobi
    Z-solves, sparse products) give the single-thread bits on 4 threads."""
    Bs, Zs, M = W.psf2d_problem(20)
    n = Zs.shape[0]
    h = lgh.AMG(Zs)
    levels = [{"A": oracle.CSR.from_scipy(l["A"]),
               "P": oracle.CSR.from_scipy(l["P"]) if l["P"] is not None else None,
               "slo": l["slo"], "shi": l["shi"], "nu": l["nu"]} for l in h.levels()]
    mg = oracle.MG(levels)
    X = oracle.randn_fill(9, 0, 0, n, 7)
    Zc = oracle.CSR.from_scipy(Zs)
    solvers = [oracle.ZSolver.mat(Zc, mode=3, mg=mg, lmin=0.5, lmax=1.1, nit=9),
               oracle.ZSolver.mat(Zc, mode=4, lmin=0.02, lmax=2.0, nit=30)]
    try:
        one = [zs.solve_block(X) for zs in solvers] + [oracle.csr_mm(oracle.CSR.from_scipy(Bs), X)]
        oracle.set_threads(4)
        four = [zs.solve_block(X) for zs in solvers] + [oracle.csr_mm(oracle.CSR.from_scipy(Bs), X)]
    finally:
        oracle.set_threads(1)
    for a, b in zip(one, four):
        assert np.array_equal(a, b)
