"""GLR engine parity through the reference C API on an MI355X:
tests/test_glr.cpp and tests/test_prior_mat.cpp restated (same problems,
same gates), plus GPU-vs-oracle comparisons on grid problems."""
import math
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import GOLDEN
import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def dev(lgh):
    lgh.init()
    return lgh


@pytest.mark.parametrize("spd", [True, False])
def test_glr_gates(dev, oracle, spd):
    """tests/test_glr.cpp:158-316 through the API (dense B, diagonal Z via
    the callbacks prior)."""
    c = 0.7
    Bd = W.tg_B_exact(spd=spd)
    dense = np.load(os.path.join(GOLDEN, "tg_eigs_%s.npy" % ("spd" if spd else "indef")))
    prior = dev.Prior.diagonal_callbacks(W.tg_z(), W.tg_m())
    B = dev.Mat.dense(Bd)
    opts = dev.glr_opts(ell=40, q_power=1, trunc_abs=0.0, trunc_rel=1e-8, nb=8,
                        panel=16, check=1)
    g = dev.GLR(B, prior, opts)
    rep = g.report
    assert rep["kept"] == 10
    lam = g.lam
    assert np.max(np.abs(lam - np.abs(dense[:10])) / np.abs(dense[:10])) < 1e-8
    assert rep["next_abs"] < 1e-6 * rep["lam_abs_max"]
    assert rep["resid_max"] < 1e-7 * rep["lam_abs_max"]
    if not spd:
        assert rep["n_negative_raw"] >= 5
    ref = np.sum(np.log(np.abs(dense[:10]) + c) - np.log(c))
    assert abs(g.logdet(c) - ref) < 1e-9 * abs(ref)
    # the oracle on the same inputs
    og = oracle.GLR(oracle.CSR.from_dense(Bd), W.tg_m(), oracle.ZSolver.diag(W.tg_z()),
                    ell=40, q_power=1, trunc_rel=1e-8)
    assert np.max(np.abs(lam - og.lam) / np.abs(og.lam)) < 1e-12
    v = W.tg_probe()
    if spd:
        y = g.apply(c, v)
        assert rel(y, og.apply(c, v)) < 1e-13
        x = g.solve(c, y)
        assert rel(x, v) < 1e-8
        assert rel(x, og.solve(c, y)) < 1e-12
        x = g.factor(c, 1, v)
        assert rel(g.factor(c, 0, x), y) < 1e-8
        y2 = g.factor(c, 1, v)
        x2 = g.factor(c, 3, y2)
        assert rel(x2, v) < 1e-8
        assert rel(g.sample(c, y2), x2) < 1e-14
        assert rel(g.pow(c, 1.0, g.pow(c, -1.0, v)), v) < 1e-10
        fl = 1.0 / np.sqrt(lam + c)
        assert rel(g.filter(fl, 1 / math.sqrt(c), v), g.pow(c, -0.5, v)) < 1e-12
        for which in range(4):
            assert rel(g.factor(c, which, v), og.factor(c, which, v)) < 1e-12
    # extend(10) from ell = 30 == one-shot at ell = 40
    g2 = dev.GLR(B, prior, dev.glr_opts(ell=30, q_power=1, trunc_rel=1e-8,
                                        nb=8, panel=16, check=1))
    g2.extend(10)
    assert g2.kept == g.kept
    assert np.max(np.abs(g2.lam - lam) / np.abs(lam)) < 1e-10


def test_glr_blocked_callbacks(dev, oracle):
    """The callback prior's blocked MATDENSE Z-solves (prior.h:53-54,
    glr_common_impl.h:206-234): a build through solveZ_blocked /
    solveZt_blocked takes no Vec solve, matches the oracle and equals the
    build through the Vec callbacks bit for bit (the same elementwise
    arithmetic)."""
    Bd = W.tg_B_exact(spd=True)
    B = dev.Mat.dense(Bd)
    opts = dev.glr_opts(ell=40, q_power=1, trunc_abs=0.0, trunc_rel=1e-8)
    calls = {}
    pb = dev.Prior.diagonal_callbacks(W.tg_z(), W.tg_m(), blocked=True, calls=calls)
    gb = dev.GLR(B, pb, opts)
    assert calls.get("solveZ_blocked", 0) > 0 and calls.get("solveZt_blocked", 0) > 0
    assert calls.get("solveZ", 0) == 0
    gv = dev.GLR(B, dev.Prior.diagonal_callbacks(W.tg_z(), W.tg_m()), opts)
    assert np.array_equal(gb.lam, gv.lam)
    assert np.array_equal(gb.U, gv.U)
    og = oracle.GLR(oracle.CSR.from_dense(Bd), W.tg_m(), oracle.ZSolver.diag(W.tg_z()),
                    ell=40, q_power=1, trunc_rel=1e-8)
    assert np.max(np.abs(gb.lam - og.lam) / np.abs(og.lam)) < 1e-12
    v = W.tg_probe()
    assert rel(gb.solve(0.7, v), og.solve(0.7, v)) < 1e-12


def test_glr_from_fit_rows_bitwise(dev):
    """B ingested from the fit stage's row export (lgh_mat_from_fit_rows:
    each row's entries shuffled, global column ids) builds the same GLR
    bit for bit as B assembled as AIJ (fit.h:160-165, glr.h:274-282)."""
    Bs, Zs, M = W.psf2d_problem(24)
    rng = np.random.default_rng(3)
    Bc = sp.csr_matrix(Bs)
    rp, cg, v = [0], [], []
    for i in range(Bc.shape[0]):
        cols = Bc.indices[Bc.indptr[i]:Bc.indptr[i + 1]].astype(np.int64)
        vals = Bc.data[Bc.indptr[i]:Bc.indptr[i + 1]]
        perm = rng.permutation(len(cols))
        cg.extend(cols[perm]); v.extend(vals[perm]); rp.append(len(cg))
    Bf = dev.Mat.from_fit_rows(np.array(rp), np.array(cg, dtype=np.int64),
                               np.array(v), sym_tol=1e-14)
    prior = dev.Prior.from_mat(Zs, M, blocked_mode=3)
    for mixed in (False, True):
        x = dev.xopts(mixed=mixed)
        g1 = dev.GLR(Bf, prior, dev.glr_opts(ell=30, seed=0), x=x)
        g2 = dev.GLR(dev.Mat.aij(Bs), prior, dev.glr_opts(ell=30, seed=0), x=x)
        assert np.array_equal(g1.lam, g2.lam) and np.array_equal(g1.U, g2.U)


@pytest.mark.parametrize("spd", [True, False])
def test_glr_implicit_u_backend(dev, oracle, spd):
    """LGH_GLR_SCALAPACK at communicator size 1 (glr_scalapack_impl.h:505-688
    build, 917-1076 implicit-U products): U = Q V is never formed, BCGS2
    orthonormalizes, and the check compares two eigensolvers.  The
    test_glr.cpp gates, the oracle's BCGS2 build, and the same operator
    actions as the replicated object."""
    c = 0.7
    Bd = W.tg_B_exact(spd=spd)
    dense = np.load(os.path.join(GOLDEN, "tg_eigs_%s.npy" % ("spd" if spd else "indef")))
    prior = dev.Prior.diagonal_callbacks(W.tg_z(), W.tg_m())
    B = dev.Mat.dense(Bd)
    o = dict(ell=40, q_power=1, trunc_abs=0.0, trunc_rel=1e-8, check=1)
    g = dev.GLR(B, prior, dev.glr_opts(backend=2, **o))
    gr = dev.GLR(B, prior, dev.glr_opts(backend=1, **o))
    og = oracle.GLR(oracle.CSR.from_dense(Bd), W.tg_m(), oracle.ZSolver.diag(W.tg_z()),
                    ell=40, q_power=1, trunc_rel=1e-8, backend=2)
    rep = g.report
    assert rep["kept"] == og.kept == 10
    lam = g.lam
    assert np.max(np.abs(lam - np.abs(dense[:10])) / np.abs(dense[:10])) < 1e-8
    assert np.max(np.abs(lam - og.lam) / np.abs(og.lam)) < 1e-12
    assert np.max(np.abs(lam - gr.lam) / np.abs(gr.lam)) < 1e-12
    assert 0.0 <= rep["resid_max"] < 1e-12       # eigensolver cross-check
    # (past the exact rank the raw values are rounding noise: absolute scale)
    for key in ("lam_raw_min", "lam_raw_max", "next_abs"):
        assert abs(rep[key] - og.report[key]) <= 1e-12 * rep["lam_abs_max"], key
    # U formed on request equals the oracle's up to column signs
    U, OU = g.U, og.U
    U = U * np.sign(np.sum(U * OU, axis=0))
    assert rel(U, OU) < 1e-10
    v = W.tg_probe()
    assert abs(g.logdet(c) - og.logdet(c)) < 1e-12 * abs(og.logdet(c))
    assert rel(g.pow(c, -0.5, v), og.pow(c, -0.5, v)) < 1e-11
    if spd:
        y = g.apply(c, v)
        for which in range(4):
            assert rel(g.factor(c, which, v), gr.factor(c, which, v)) < 1e-11
        assert rel(g.solve(c, y), og.solve(c, y)) < 1e-11
        assert rel(g.sample(c, v), og.sample(c, v)) < 1e-11
        cs = np.array([0.1, 0.7, 3.0])
        R = np.stack([v, np.cos(np.arange(v.size)), v[::-1].copy()], 1)
        Xb = g.solve_block(cs, R)
        for j in range(3):
            assert rel(Xb[:, j], g.solve(cs[j], R[:, j])) < 1e-12
    # extend stays implicit: == the oracle's incremental extension
    g2 = dev.GLR(B, prior, dev.glr_opts(backend=2, **dict(o, ell=30)))
    g2.extend(10)
    og2 = oracle.GLR(oracle.CSR.from_dense(Bd), W.tg_m(), oracle.ZSolver.diag(W.tg_z()),
                     ell=30, q_power=1, trunc_rel=1e-8, backend=2)
    og2.extend(10, incremental=True, panel=64)
    assert g2.kept == og2.kept == 10
    assert np.max(np.abs(g2.lam - og2.lam) / np.abs(og2.lam)) < 1e-12
    assert np.max(np.abs(g2.lam - lam) / np.abs(lam)) < 1e-10
    assert g2.report["resid_max"] == 0.0          # the extension runs no check


@pytest.mark.parametrize("treat", [0, 1, 2])
def test_glr_treatments_match_oracle(dev, oracle, treat):
    """FLIP / RELU / NONE (glr_common_impl.h:549-559) on the indefinite
    closed-form problem: kept count, treated spectrum, logdet and a solve
    equal the oracle's; RELU keeps zero-valued modes in kept."""
    Bd = W.tg_B_exact(spd=False)
    prior = dev.Prior.diagonal_callbacks(W.tg_z(), W.tg_m())
    g = dev.GLR(dev.Mat.dense(Bd), prior,
                dev.glr_opts(ell=40, trunc_rel=1e-8, treat=treat))
    og = oracle.GLR(oracle.CSR.from_dense(Bd), W.tg_m(),
                    oracle.ZSolver.diag(W.tg_z()), ell=40, q_power=1,
                    trunc_rel=1e-8, treat=treat)
    assert g.kept == og.kept == 10
    lam, olam = g.lam, og.lam
    assert np.max(np.abs(lam - olam)) < 1e-12 * np.max(np.abs(olam))
    if treat == 0:
        assert np.all(lam > 0)
    elif treat == 1:
        assert np.all(lam >= 0) and np.sum(lam == 0) >= 5
    else:
        assert np.sum(lam < 0) >= 5
    c = 0.7 + 2 * max(0.0, -lam.min())
    assert abs(g.logdet(c) - og.logdet(c)) < 1e-12 * abs(og.logdet(c))
    v = W.tg_probe()
    assert rel(g.solve(c, v), og.solve(c, v)) < 1e-12


@pytest.mark.parametrize("ell", [1, 120])
def test_glr_extreme_sketch_widths(dev, oracle, ell):
    """ell = 1 (one hashed column) and ell = N = 120 (a sketch as wide as
    the problem; the N = 120 rows also leave a ragged 64-row panel tail):
    kept modes and spectrum equal the oracle's; at ell = N the rank-10
    spectrum is the dense one."""
    Bd = W.tg_B_exact(spd=True)
    prior = dev.Prior.diagonal_callbacks(W.tg_z(), W.tg_m())
    g = dev.GLR(dev.Mat.dense(Bd), prior, dev.glr_opts(ell=ell, trunc_rel=1e-8))
    og = oracle.GLR(oracle.CSR.from_dense(Bd), W.tg_m(),
                    oracle.ZSolver.diag(W.tg_z()), ell=ell, q_power=1,
                    trunc_rel=1e-8)
    assert g.kept == og.kept == min(ell, 10)
    assert np.max(np.abs(g.lam - og.lam) / np.abs(og.lam)) < 1e-11
    v = W.tg_probe()
    assert rel(g.pow(0.7, 1.0, g.pow(0.7, -1.0, v)), v) < 1e-10
    if ell == 120:      # the model is exact: solve inverts the true apply
        dense = np.load(os.path.join(GOLDEN, "tg_eigs_spd.npy"))
        assert np.max(np.abs(g.lam - dense[:10]) / dense[:10]) < 1e-8
        assert rel(g.solve(0.7, g.apply(0.7, v)), v) < 1e-8


def test_glr_refusals(dev):
    prior = dev.Prior.diagonal_callbacks(W.tg_z(), W.tg_m())
    B = dev.Mat.dense(W.tg_B_exact())
    g = dev.GLR(B, prior, dev.glr_opts(ell=40, trunc_rel=1e-8))
    with pytest.raises(dev.LGHError):
        g.solve(0.0, W.tg_probe())
    with pytest.raises(dev.LGHError):
        g.extend(0)
    with pytest.raises(dev.LGHError):
        g.extend(1000)


@pytest.mark.parametrize("mode", [0, 2, 3, 4])
def test_prior_mat_gates(dev, oracle, mode):
    """tests/test_prior_mat.cpp:129-193 (tile 8) per blocked mode."""
    Z = W.tpm_Z()
    n = Z.shape[0]
    b = W.tg_probe(n)
    xref = np.load(os.path.join(GOLDEN, "tpm_xref.npy"))
    prior = dev.Prior.from_mat(Z, W.tg_m(n), blocked_mode=mode, tile=8)
    x = prior.solve(b, which=1)[:, 0]
    assert rel(x, xref) < 1e-8
    X = np.stack([b * (1 + 0.1 * j) for j in range(11)], 1)
    Y = prior.solve(X, which=0)
    worst = max(rel(Y[:, j] / (1 + 0.1 * j), xref) for j in range(11))
    assert worst < 1e-7
    s1 = Z @ b
    assert rel(prior.solve(s1, which=1)[:, 0], b) < 1e-8
    Rv = prior.apply(b)
    assert rel(Rv, Z @ ((Z @ b) / W.tg_m(n))) < 1e-13


def test_prior_mat_e2e(dev):
    """tests/test_prior_mat.cpp:221-277: low-rank dense B, mode-2 prior."""
    Z = W.tpm_Z()
    n = Z.shape[0]
    prior = dev.Prior.from_mat(Z, W.tg_m(n), blocked_mode=2, tile=8)
    B = dev.Mat.dense(W.tpm_B(n))
    g = dev.GLR(B, prior, dev.glr_opts(ell=16, trunc_rel=1e-6))
    assert g.report["kept"] == 6
    v = W.tg_probe(n)
    c = 0.7
    assert rel(g.solve(c, g.apply(c, v)), v) < 1e-6


@pytest.fixture(scope="module")
def grid(dev, oracle):
    """32 x 32 grid, the 2D isotropic PSF B and Z = M + 0.1 K, mode-3 prior; oracle built on the
    engine's own hierarchy (same Chebyshev-MG operator)."""
    side = 32
    Bs, Zs, M = W.psf2d_problem(side)
    prior = dev.Prior.from_mat(Zs, M, blocked_mode=3)
    info = prior.info()
    h = dev.AMG(Zs)
    levels = [{"A": oracle.CSR.from_scipy(l["A"]),
               "P": oracle.CSR.from_scipy(l["P"]) if l["P"] is not None else None,
               "slo": l["slo"], "shi": l["shi"], "nu": l["nu"]} for l in h.levels()]
    mg = oracle.MG(levels, gamma=info["cycle"])
    zs = oracle.ZSolver.mat(oracle.CSR.from_scipy(Zs), mode=3, mg=mg,
                            lmin=info["lmin"], lmax=info["lmax"], nit=info["nit"],
                            ksp_rtol=1e-12)
    return dict(B=Bs, Z=Zs, M=M, prior=prior, zs=zs, mg=mg, info=info, h=h)


def test_grid_blocked_solve_matches_oracle(dev, grid, oracle):
    n = grid["Z"].shape[0]
    X = oracle.randn_fill(7, 0, 0, n, 9)
    ref = grid["zs"].solve_block(X)
    assert rel(grid["prior"].solve(X, which=0), ref) < 1e-12
    # and it is Z^{-1} to the Chebyshev tolerance
    assert rel(grid["Z"] @ ref, X) < 1e-8
    # FP32 throughout: FP32-level agreement
    assert rel(grid["prior"].solve(X, which=0, prec=1), ref) < 1e-4
    # FP64 Chebyshev around the FP32 or BF16 V-cycle: the same FP64-accurate
    # Z^{-1} (each step rounds its direction once and moves X and R by it)
    for prec in (2, 3):
        mixed = grid["prior"].solve(X, which=0, prec=prec)
        assert rel(grid["Z"] @ mixed, X) < 1e-8, prec
        assert rel(mixed, ref) < 1e-8, prec
    # FP32 Chebyshev around the BF16 V-cycle: FP32-level agreement
    assert rel(grid["prior"].solve(X, which=0, prec=4), ref) < 1e-4


def test_grid_mixed_short_solve(dev, grid):
    """A loose Chebyshev tolerance gives solves of fewer than 8 steps, which
    the mixed tier runs with FP64 R and X throughout (no early FP32 phase):
    they reach the same residual as the FP64 tier's solve of that length."""
    n = grid["Z"].shape[0]
    p = dev.Prior.from_mat(grid["Z"], grid["M"], blocked_mode=3, cheb_rtol=1e-3)
    nit = p.info()["nit"]
    assert 1 <= nit < 8, nit
    X = dev.randn_fill(8, 0, 0, n, 9)
    ref = p.solve(X, which=0)
    r64 = rel(grid["Z"] @ ref, X)
    assert r64 < 1e-2
    for prec in (2, 3):
        Y = p.solve(X, which=0, prec=prec)
        assert rel(grid["Z"] @ Y, X) < 3 * r64, prec
        assert rel(Y, ref) < 0.1, prec


@pytest.mark.parametrize("policy", ["plain", "mixed", "mixed-fp32-vcycle",
                                    "mixed-fp32-sketch"])
def test_grid_glr_matches_oracle(dev, grid, oracle, policy):
    ell = 24
    og = oracle.GLR(oracle.CSR.from_scipy(grid["B"]), grid["M"], grid["zs"],
                    ell=ell, q_power=1, trunc_rel=1e-3, seed=0)
    mixed = policy != "plain"
    x = dev.xopts(mixed=mixed)
    if policy == "mixed-fp32-vcycle":
        x.prec_precond = dev.PREC_F32
    elif policy == "mixed-fp32-sketch":
        x.prec_sketch = dev.PREC_F32
        x.prec_sketch_precond = dev.PREC_F32
    else:
        assert not mixed or (x.prec_precond == dev.PREC_BF16
                             and x.prec_sketch == dev.PREC_F64
                             and x.prec_sketch_precond == dev.PREC_BF16)
    g = dev.GLR(grid["B"], grid["prior"], dev.glr_opts(ell=ell, seed=0), x=x)
    assert g.kept == og.kept
    if mixed:
        # normwise relative spectrum error (|dlam_i| / max|lam|); with an
        # FP32 sketch the unconverged trailing Ritz values of this small
        # sketch move with its rounding the way they move with any sketch
        # perturbation (the FP64-sketch default stays far inside these bounds)
        tol = 1e-6
        assert np.max(np.abs(g.lam - og.lam)) / og.lam[0] < tol
        k = og.kept // 2
        assert np.max(np.abs(g.lam[:k] - og.lam[:k]) / og.lam[:k]) < tol
    else:
        tol = 1e-9
        assert np.max(np.abs(g.lam - og.lam) / np.abs(og.lam)) < tol
    c = 1e-2
    rhs = oracle.randn_fill(11, 0, 0, grid["Z"].shape[0], 1)[:, 0]
    if mixed:
        # the mixed build's downstream results differ from the FP64 ones in
        # the trailing (unconverged) modes only: by less than 1e-4, and by
        # well under what another Gaussian sketch (seed 1) moves them
        g1 = dev.GLR(grid["B"], grid["prior"], dev.glr_opts(ell=ell, seed=1))
        for op in ("solve", "sample"):
            ref = getattr(og, op)(c, rhs)
            d_mixed = rel(getattr(g, op)(c, rhs), ref)
            d_seed = rel(getattr(g1, op)(c, rhs), ref)
            assert d_mixed < 1e-4 and d_mixed < 0.5 * d_seed, (op, d_mixed, d_seed)
        assert abs(g.logdet(c) - og.logdet(c)) < 1e-6 * abs(og.logdet(c))
        # and the mixed solve is the FP64 master formula of its own U, lam:
        # x = Z^-1 M^1/2 [U (f - f_c) U^T + f_c I] M^1/2 Z^-1 rhs (glr.h:14-19)
        U, lam, msq = g.U, g.lam, np.sqrt(grid["M"])
        w = msq * grid["zs"].solve_vec(rhs)
        t = U @ ((1.0 / (lam + c) - 1.0 / c) * (U.T @ w)) + w / c
        assert rel(g.solve(c, rhs), grid["zs"].solve_vec(msq * t)) < 1e-9
    else:
        assert rel(g.solve(c, rhs), og.solve(c, rhs)) < tol
        assert rel(g.sample(c, rhs), og.sample(c, rhs)) < tol
        assert abs(g.logdet(c) - og.logdet(c)) < tol * abs(og.logdet(c))
    # multi-shift block == per-column
    cs = np.array([1e-3, 1e-2, 0.5])
    R = oracle.randn_fill(12, 0, 0, grid["Z"].shape[0], 3)
    Xb = g.solve_block(cs, R)
    for j in range(3):
        assert rel(Xb[:, j], g.solve(cs[j], R[:, j])) < 1e-12
    if mixed:
        # the blocked tier solves a column with the same bits whatever the
        # width of the panel it lands in: 5 columns run as one 8-wide panel,
        # 20 as two 16-wide ones, 70 as a whole panel and an 8-wide tail
        R70 = oracle.randn_fill(13, 0, 0, grid["Z"].shape[0], 70)
        c70 = 10.0 ** np.linspace(-3, 0, 70)
        X70 = g.solve_block(c70, R70)
        assert np.array_equal(g.solve_block(c70[:5], R70[:, :5]), X70[:, :5])
        assert np.array_equal(g.solve_block(c70[:20], R70[:, :20]), X70[:, :20])
        assert np.array_equal(g.solve_block(c70[64:], R70[:, 64:]), X70[:, 64:])
        assert np.array_equal(g.sample_block(c70[:5], R70[:, :5]),
                              g.sample_block(c70, R70)[:, :5])
        # repeated narrow-panel solves reuse their staging and work buffers:
        # calls with new inputs, and 16-wide work between them, keep the
        # first call's bits
        x1 = g.solve_block(c70[:1], R70[:, :1])
        for _ in range(4):
            assert np.array_equal(g.solve_block(c70[:20], R70[:, :20]), X70[:, :20])
            assert np.array_equal(g.solve_block(c70[:1], R70[:, :1]), x1)
            assert np.array_equal(g.solve_block(c70[1:4], R70[:, 1:4]), X70[:, 1:4])
            assert np.array_equal(g.solve(c70[5], R70[:, 5]), X70[:, 5])


@pytest.mark.parametrize("mixed", [False, True])
def test_grid_block_ops_on_device(dev, grid, oracle, mixed):
    """lgh_glr_solve_block_dev / lgh_glr_sample_block_dev on blocks already
    in device memory: the bits of the host-block calls (a whole panel and
    an 8-wide tail, one column, columns with a leading dimension above n),
    the plain build's columns against the oracle, host memory refused."""
    import ctypes as C
    n = grid["Z"].shape[0]
    ell = 24
    g = dev.GLR(grid["B"], grid["prior"], dev.glr_opts(ell=ell, seed=0),
                x=dev.xopts(mixed=mixed))
    R = oracle.randn_fill(14, 0, 0, n, 70)
    cs = 10.0 ** np.linspace(-3, 0, 70)
    Rd = dev.DeviceBlock.from_host(R)
    Rs = dev.DeviceBlock.from_host(R, ld=n + 5)
    assert np.array_equal(Rs.to_host(), R) and np.array_equal(Rd.to_host(), R)
    for op in ("solve", "sample"):
        host = getattr(g, op + "_block")(cs, R)
        dev_op = getattr(g, op + "_block_device")
        assert np.array_equal(dev_op(cs, Rd).to_host(), host), op
        assert np.array_equal(dev_op(cs, Rs).to_host(), host), op
        assert np.array_equal(dev_op(cs[3:4], Rs.columns(3, 4)).to_host(),
                              getattr(g, op + "_block")(cs[3:4], R[:, 3:4])), op
    if not mixed:
        og = oracle.GLR(oracle.CSR.from_scipy(grid["B"]), grid["M"], grid["zs"],
                        ell=ell, q_power=1, trunc_rel=1e-3, seed=0)
        X = g.solve_block_device(cs[:3], Rd.columns(0, 3)).to_host()
        for j in range(3):
            assert rel(X[:, j], og.solve(cs[j], R[:, j])) < 1e-9
    # host pointers are refused before any device work
    Y = np.zeros((n, 2), order="F")
    cc = np.array([1e-2, 1e-2])
    rc = dev.lib().lgh_glr_solve_block_dev(
        g.h, 2, cc.ctypes.data_as(C.POINTER(C.c_double)), R.ctypes.data, n,
        Y.ctypes.data, n)
    assert rc != 0 and "device memory" in dev.lib().lgh_petsc_last_error().decode()
    with pytest.raises(ValueError):
        g.solve_block_device(cs[:2], R[:, :2])
    # a leading dimension below n and a shift at the floor are refused on
    # the host; no columns is a no-op
    pc = cc.ctypes.data_as(C.POINTER(C.c_double))
    assert dev.lib().lgh_glr_solve_block_dev(g.h, 2, pc, Rd.ptr, n - 1, Rd.ptr, n) != 0
    assert "leading dimensions" in dev.lib().lgh_petsc_last_error().decode()
    bad = np.array([0.0, 1e-2])
    with pytest.raises(dev.LGHError):
        g.solve_block_device(bad, Rd.columns(0, 2))
    assert dev.lib().lgh_glr_sample_block_dev(g.h, 0, pc, Rd.ptr, n, Rd.ptr, n) == 0
    # a block running past its allocation (71 columns of a 70-column one)
    # is refused before any device work
    c71 = np.full(71, 1e-2)
    Y71 = dev.DeviceBlock(n, 71)
    assert dev.lib().lgh_glr_solve_block_dev(
        g.h, 71, c71.ctypes.data_as(C.POINTER(C.c_double)), Rd.ptr, n, Y71.ptr, n) != 0
    assert "extends past" in dev.lib().lgh_petsc_last_error().decode()


@pytest.mark.parametrize("mixed", [False, True])
def test_grid_sweep_chunks_bitwise(dev, grid, mixed):
    """F sweeps in chunks of 1 or 2 panels (xopts.sweep_panels) give the
    same object bit for bit as the whole block at once: panels are solved
    independently (ell = 150: two whole panels and a 22-column tail)."""
    o = dev.glr_opts(ell=150, seed=0, trunc_rel=1e-6)
    ref = dev.GLR(grid["B"], grid["prior"], o, x=dev.xopts(mixed=mixed))
    for cp in (1, 2):
        g = dev.GLR(grid["B"], grid["prior"], o,
                    x=dev.xopts(mixed=mixed, sweep_panels=cp))
        assert g.kept == ref.kept
        assert np.array_equal(g.lam, ref.lam), cp
        assert np.array_equal(g.U, ref.U), cp
    # and on the implicit-U backend, with an extension in chunks too
    oi = dev.glr_opts(ell=150, seed=0, trunc_rel=1e-6, backend=2)
    a = dev.GLR(grid["B"], grid["prior"], oi, x=dev.xopts(mixed=mixed))
    b = dev.GLR(grid["B"], grid["prior"], oi, x=dev.xopts(mixed=mixed, sweep_panels=1))
    a.extend(70)
    b.extend(70)
    assert np.array_equal(a.lam, b.lam)


@pytest.mark.parametrize("mixed", [False, True])
def test_grid_fused_coarse_levels_bitwise(dev, grid, mixed):
    """A build (ell = 150: two whole panels and a 22-column tail) and a
    solve and sample, with the small AMG levels below the finest as one
    fused launch per coarse visit (the default), equal the same with one
    launch per operation and level bit for bit."""
    assert len(grid["h"].levels()) >= 3
    o = dev.glr_opts(ell=150, seed=0, trunc_rel=1e-6)
    x = dev.xopts(mixed=mixed)
    rhs = np.random.default_rng(8).standard_normal((grid["Z"].shape[0], 3))
    dev.prof_enable(True)
    try:
        dev.prof_collect("mg_coarse")
        f = dev.GLR(grid["B"], grid["prior"], o, x=x)
        fs, fx = f.solve_block(1e-2, rhs), f.sample(1e-2, rhs[:, 0])
        assert dev.prof_collect("mg_coarse")[0] > 0
        dev.mg_coarse_config(0)
        p = dev.GLR(grid["B"], grid["prior"], o, x=x)
        ps, px = p.solve_block(1e-2, rhs), p.sample(1e-2, rhs[:, 0])
        assert dev.prof_collect("mg_coarse")[0] == 0
    finally:
        dev.mg_coarse_config(-1, -1)
        dev.prof_enable(False)
    assert f.kept == p.kept
    assert np.array_equal(f.lam, p.lam)
    assert np.array_equal(f.U, p.U)
    assert np.array_equal(fs, ps) and np.array_equal(fx, px)


def test_grid_extend_incremental(dev, grid, oracle):
    g1 = dev.GLR(grid["B"], grid["prior"], dev.glr_opts(ell=30, seed=0, trunc_rel=1e-4))
    g2 = dev.GLR(grid["B"], grid["prior"], dev.glr_opts(ell=20, seed=0, trunc_rel=1e-4))
    g2.extend(10)
    og = oracle.GLR(oracle.CSR.from_scipy(grid["B"]), grid["M"], grid["zs"],
                    ell=20, q_power=1, trunc_rel=1e-4, seed=0)
    og.extend(10, incremental=True, panel=64)
    assert g2.kept == og.kept
    assert np.max(np.abs(g2.lam - og.lam) / np.abs(og.lam)) < 1e-9
    # incremental and one-shot span the same space in exact arithmetic:
    # leading (well-converged) modes agree
    k = min(5, g1.kept)
    assert np.max(np.abs(g2.lam[:k] - g1.lam[:k]) / np.abs(g1.lam[:k])) < 1e-6


@pytest.mark.parametrize("mixed", [False, True])
def test_grid3d_16_extend_twice(dev, oracle, mixed):
    """The 3D configuration's shape at 16^3 (n = 4096): Z = M + 0.1 K7
    (7-point, AMG tier), isotropic PSF B (sigma 1.5h, radius-2h ball,
    33 nnz/row), ell = 40 then lgh_glr_extend by 20 twice; engine vs the
    oracle's incremental extension on the same hierarchy."""
    Bs, Zs, M = W.psf3d_problem(16)
    assert abs(Bs.nnz / Bs.shape[0] - 33) < 8
    prior = dev.Prior.from_mat(Zs, M, blocked_mode=3)
    info = prior.info()
    h = dev.AMG(Zs)
    levels = [{"A": oracle.CSR.from_scipy(l["A"]),
               "P": oracle.CSR.from_scipy(l["P"]) if l["P"] is not None else None,
               "slo": l["slo"], "shi": l["shi"], "nu": l["nu"]} for l in h.levels()]
    zs = oracle.ZSolver.mat(oracle.CSR.from_scipy(Zs), mode=3, mg=oracle.MG(levels, gamma=info["cycle"]),
                            lmin=info["lmin"], lmax=info["lmax"], nit=info["nit"],
                            ksp_rtol=1e-12)
    og = oracle.GLR(oracle.CSR.from_scipy(Bs), M, zs, ell=40, q_power=1,
                    trunc_rel=1e-4, seed=0)
    g = dev.GLR(Bs, prior, dev.glr_opts(ell=40, seed=0, trunc_rel=1e-4),
                x=dev.xopts(mixed=mixed))
    # FP64: the oracle's arithmetic; mixed: FP64-accurate Z solves around the
    # BF16 V-cycle
    tol = 1e-7 if mixed else 1e-9
    for step in range(3):
        if step:
            g.extend(20)
            og.extend(20, incremental=True, panel=64)
        assert g.kept == og.kept
        assert np.max(np.abs(g.lam - og.lam) / np.abs(og.lam)) < tol, step
    c = 1e-2
    rhs = oracle.randn_fill(3, 0, 0, Zs.shape[0], 1)[:, 0]
    assert rel(g.solve(c, rhs), og.solve(c, rhs)) < (1e-6 if mixed else 1e-9)


def test_grid2d_128_full_size_matches_oracle(dev, oracle):
    """The CPU-runnable benchmark case at full size: 128 x 128 grid
    (n = 16 384), Z = M + 0.1 K, symmetrized Gaussian PSF B (radius 4h),
    ell = 64 + 10, seed 0, FP64; one solve and one sample at c = 1e-2 with
    N(0,1) right-hand sides.  Engine vs the FP64 oracle on the same AMG
    hierarchy; the mixed build vs the plain one."""
    Bs, Zs, M = W.psf2d_problem(128)
    n = Zs.shape[0]
    prior = dev.Prior.from_mat(Zs, M, blocked_mode=3)
    info = prior.info()
    h = dev.AMG(Zs)
    levels = [{"A": oracle.CSR.from_scipy(l["A"]),
               "P": oracle.CSR.from_scipy(l["P"]) if l["P"] is not None else None,
               "slo": l["slo"], "shi": l["shi"], "nu": l["nu"]} for l in h.levels()]
    mg = oracle.MG(levels, gamma=info["cycle"])
    zs = oracle.ZSolver.mat(oracle.CSR.from_scipy(Zs), mode=3, mg=mg,
                            lmin=info["lmin"], lmax=info["lmax"], nit=info["nit"],
                            ksp_rtol=1e-12)
    og = oracle.GLR(oracle.CSR.from_scipy(Bs), M, zs, ell=74, q_power=1, seed=0)
    g = dev.GLR(Bs, prior, dev.glr_opts(ell=74, seed=0))
    assert g.kept == og.kept and g.kept > 10
    assert np.max(np.abs(g.lam - og.lam) / np.abs(og.lam)) < 1e-9
    for key in ("lam_raw_min", "lam_raw_max", "next_abs"):
        assert abs(g.report[key] - og.report[key]) <= 1e-9 * abs(og.report[key])
    c = 1e-2
    rhs = oracle.randn_fill(1, 0, 0, n, 1)[:, 0]
    xi = oracle.randn_fill(2, 0, 0, n, 1)[:, 0]
    assert rel(g.solve(c, rhs), og.solve(c, rhs)) < 1e-9
    assert rel(g.sample(c, xi), og.sample(c, xi)) < 1e-9
    assert abs(g.logdet(c) - og.logdet(c)) < 1e-9 * abs(og.logdet(c))
    # the mixed policy (FP64 vectors around the BF16 V-cycle): its Z solves
    # are FP64-accurate to the Chebyshev tolerance, so every result stays
    # within 1e-7 of the FP64 build
    gm = dev.GLR(Bs, prior, dev.glr_opts(ell=74, seed=0), x=dev.xopts(mixed=True))
    assert gm.kept == g.kept
    assert np.max(np.abs(gm.lam - g.lam) / np.abs(g.lam)) < 1e-7
    assert rel(gm.solve(c, rhs), g.solve(c, rhs)) < 1e-7
    assert rel(gm.sample(c, xi), g.sample(c, xi)) < 1e-7


def test_c_consumer_of_public_api(tmp_path):
    """tests/c/test_glr_api.cpp: unchanged-API C++ consumer (the shape of
    tests/test_glr.cpp) linked against liblgh.so."""
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "tests", "c", "test_glr_api")
    if not os.path.exists(exe):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "c")], check=True)
    files = []
    for name in ("spd", "indef"):
        p = tmp_path / ("eigs_%s.txt" % name)
        np.savetxt(p, np.load(os.path.join(GOLDEN, "tg_eigs_%s.npy" % name)), fmt="%.17g")
        files.append(str(p))
    r = subprocess.run([exe] + files, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS test_glr_api" in r.stdout
