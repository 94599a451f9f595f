"""The measured configurations against the FP64 oracle on an MI355X.

* anisotropic PSF (the 1024 x 1024 build's operator family: Sigma(x) =
  R(pi x1) diag((3h)^2, (1.5h)^2) R^T truncated at Mahalanobis 3, ~127
  entries a row; Z = M + 0.05 K) at side 64: plain and mixed builds, with
  the plain and the row-grouped B product, against oracle.GLR on the same
  AMG hierarchy (spectrum, report, solve, sample, logdet); and at full size
  (n = 1 048 576) the property the build relies on: the blocked Z-solve
  tiers reach ||Z X - B|| / ||B|| < 1e-8;
* serving: requests with different shifts and right-hand-side counts,
  submitted concurrently to the request loop over a resident GLR, against
  oracle.solve / oracle.sample / oracle.logdet at the same c and right-hand
  sides, plain and mixed;
* batched uneven instances: instances of the 256-instance workload at
  their own sizes (n = 4096 ... 16641, ell = rank + 10), built as one batch,
  against one oracle build each."""
import threading

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def dev(lgh):
    lgh.init()
    return lgh


def oracle_zsolver(dev, oracle, Zs, prior):
    """oracle Z-solver with the device prior's operator: the engine's own
    AMG hierarchy, interval, Chebyshev count and cycle"""
    info = prior.info()
    h = dev.AMG(Zs)
    levels = [{"A": oracle.CSR.from_scipy(l["A"]),
               "P": oracle.CSR.from_scipy(l["P"]) if l["P"] is not None else None,
               "slo": l["slo"], "shi": l["shi"], "nu": l["nu"]} for l in h.levels()]
    mg = oracle.MG(levels, gamma=info["cycle"])
    return oracle.ZSolver.mat(oracle.CSR.from_scipy(Zs), mode=3, mg=mg,
                              lmin=info["lmin"], lmax=info["lmax"], nit=info["nit"],
                              ksp_rtol=1e-12)


@pytest.fixture(scope="module")
def aniso(dev, oracle):
    Bs, Zs, M = W.aniso_psf2d_problem(64)
    assert Bs.nnz > 120 * Bs.shape[0]
    prior = dev.Prior.from_mat(Zs, M, blocked_mode=3)
    zs = oracle_zsolver(dev, oracle, Zs, prior)
    # ell = 100: one whole 64-column panel and a 36-column tail (the narrow
    # panels of the Z-solves), q = 1, the default seed
    og = oracle.GLR(oracle.CSR.from_scipy(Bs), M, zs, ell=100, q_power=1, seed=271828)
    return dict(B=Bs, Z=Zs, M=M, prior=prior, zs=zs, og=og, n=Zs.shape[0])


@pytest.mark.parametrize("mixed", [False, True])
def test_aniso_psf_build_matches_oracle(dev, oracle, aniso, mixed):
    og = aniso["og"]
    x = dev.xopts(mixed=mixed)
    opts = dev.glr_opts(ell=100, q_power=1, seed=271828)
    g = dev.GLR(aniso["B"], aniso["prior"], opts, x=x)
    # the same build through the row-grouped B product (8-row groups, 4 and
    # 2 panels a wavefront): the same bits
    try:
        for rg, ppw in ((8, 4), (4, 2)):
            dev.spmm_group_config(rg, ppw)
            gg = dev.GLR(dev.Mat.aij(aniso["B"]), aniso["prior"], opts, x=x)
            assert np.array_equal(gg.lam, g.lam), (rg, ppw)
            assert np.array_equal(gg.U, g.U), (rg, ppw)
        # and through the row-tile copy (1D 4 x 16 rows; 2D 16 x 4 rows on
        # the grid's line stride, in XCD-run order; the matrix-core kernel
        # on 2D tiles of 4 x 4-row blocks)
        for cand in (0, 4 | 8 | 3, 16 | 4 | 1, 16 | 4 | 8 | 2):
            dev.spmm_tile_config(cand)
            gg = dev.GLR(dev.Mat.aij(aniso["B"]), aniso["prior"], opts, x=x)
            assert np.array_equal(gg.lam, g.lam), cand
            assert np.array_equal(gg.U, g.U), cand
    finally:
        dev.spmm_group_config(0, 0)
        dev.spmm_tile_config(-1)
    assert g.kept == og.kept and g.kept >= 10
    tol = 1e-7 if mixed else 1e-9
    assert np.max(np.abs(g.lam - og.lam) / np.abs(og.lam)) < tol
    for key in ("lam_abs_max", "lam_raw_max", "lam_raw_min", "next_abs"):
        assert abs(g.report[key] - og.report[key]) <= tol * abs(og.report[key]), key
    assert g.report["n_negative_raw"] == og.report["n_negative_raw"]
    n = aniso["n"]
    rhs = oracle.randn_fill(21, 0, 0, n, 1)[:, 0]
    xi = oracle.randn_fill(22, 0, 0, n, 1)[:, 0]
    for c in (1e-4, 1e-2, 3.0):
        assert rel(g.solve(c, rhs), og.solve(c, rhs)) < tol, c
        assert rel(g.sample(c, xi), og.sample(c, xi)) < tol, c
        assert abs(g.logdet(c) - og.logdet(c)) < tol * abs(og.logdet(c)), c


def test_aniso_psf_full_size_zsolve_residual(dev):
    """Z = M + 0.05 K on the 1024 x 1024 grid: the mode-3 tier (Chebyshev
    around the AMG V-cycle) in FP64, and the mixed tiers (FP64 Chebyshev
    around the BF16 or FP32 V-cycle) reach the Chebyshev tolerance."""
    Zs, M = W.prior_Z(1024, 2, 1.0, 0.05)
    n = Zs.shape[0]
    prior = dev.Prior.from_mat(Zs, M, blocked_mode=3)
    info = prior.info()
    assert info["levels"] >= 5 and info["nit"] <= 20
    X = np.random.default_rng(4).standard_normal((n, 6))
    for prec in (0, 2, 3):
        Y = prior.solve(X, which=0, prec=prec)
        r = np.linalg.norm(Zs @ Y - X, axis=0) / np.linalg.norm(X, axis=0)
        assert r.max() < 1e-8, (prec, r.max())


def _serve_requests(n, rng):
    """(op, c, X): 60 / 35 / 5 % solve / sample / logdet, c log-uniform in
    [1e-4, 10], 1..24 right-hand sides"""
    ops = rng.choice([0, 1, 2], size=14, p=[0.6, 0.35, 0.05])
    ops[3] = 2                                 # at least one logdet
    out = []
    for op in ops:
        c = 10.0 ** rng.uniform(-4, 1)
        m = int(rng.integers(1, 25))
        out.append((int(op), c, None if op == 2 else rng.standard_normal((n, m))))
    return out


@pytest.mark.parametrize("mixed", [False, True])
def test_served_results_match_oracle(dev, oracle, aniso, mixed):
    og = aniso["og"]
    g = dev.GLR(aniso["B"], aniso["prior"], dev.glr_opts(ell=100, seed=271828),
                x=dev.xopts(mixed=mixed))
    reqs = _serve_requests(aniso["n"], np.random.default_rng(31))
    srv = dev.Serve(g, max_batch_cols=64)
    res = [None] * len(reqs)

    def client(ids):
        ts = [(i, srv.submit(reqs[i][0], reqs[i][1], reqs[i][2])) for i in ids]
        for i, t in ts:
            res[i] = srv.wait(t)[0]

    th = [threading.Thread(target=client, args=(list(range(k, len(reqs), 3)),))
          for k in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    st = srv.stats()
    srv.close()
    assert st["requests"] == len(reqs)
    tol = 1e-7 if mixed else 1e-9
    for (op, c, X), r in zip(reqs, res):
        if op == 2:
            assert abs(r - og.logdet(c)) < tol * abs(og.logdet(c))
            continue
        fn = og.solve if op == 0 else og.sample
        for j in range(X.shape[1]):
            assert rel(r[:, j], fn(c, X[:, j])) < tol, (op, c, j)


def test_batched_instances_match_oracle(dev, oracle):
    """Instances 154, 107, 206 (sides 64, 65, 64; ranks 32, 64, 128) and 23
    (side 129, rank 64) of the 256-instance workload, built as one batch of
    the mixed policy: each against its oracle build (same hierarchy)."""
    ks = [154, 107, 206, 23]
    inst = [W.small_psf2d_instance(k) for k in ks]
    priors = dev.Prior.batch([(Z, M) for _, Z, M, _ in inst], workers=4, blocked_mode=3)
    x = dev.xopts(mixed=True)
    gs = dev.GLR.batch([(dev.Mat.aij(B), pr, dev.glr_opts(ell=ell, seed=0))
                        for (B, _, _, ell), pr in zip(inst, priors)], x=x, workers=4)
    for k, (B, Z, M, ell), pr, g in zip(ks, inst, priors, gs):
        zs = oracle_zsolver(dev, oracle, Z, pr)
        og = oracle.GLR(oracle.CSR.from_scipy(B), M, zs, ell=ell, q_power=1, seed=0)
        assert g.kept == og.kept, k
        assert np.max(np.abs(g.lam - og.lam) / np.abs(og.lam)) < 1e-7, k
        rhs = oracle.randn_fill(5, 0, 0, Z.shape[0], 1)[:, 0]
        assert rel(g.solve(1e-2, rhs), og.solve(1e-2, rhs)) < 1e-7, k
        assert rel(g.sample(1e-2, rhs), og.sample(1e-2, rhs)) < 1e-7, k
