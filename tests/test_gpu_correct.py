"""Stage-2.5 deflation correction on an MI355X: tests/test_deflate.cpp of
the reference restated through the API.  N = 120, B SPD of exact rank 10
(the GLR model of B is exact), diagonal Z and mass, true Hessian
H_d = B + DH with a planted signed error of exact rank 4, so a correction
with enough probes / applies recovers F_true = whiten(H_d) exactly; the
gates are the reference test's, against a dense oracle (numpy eigvalsh of
F_true, signed, |lambda|-descending)."""
import math

import numpy as np
import pytest

import workloads as W

pytestmark = pytest.mark.gpu

N, RANK_B, RANK_C = 120, 10, 4


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


def perr(k):
    i = np.arange(N)
    return np.cos(0.21 * (i + 2) * (k + 1)) + 0.3 * np.sin(0.4 * i - k)


def terr(k, err_scale):
    t = 0.08 / (k + 1.0) * (-1.0 if k % 2 else 1.0)
    return t * err_scale if k == 1 else t


def hd_matrix(err_scale):
    H = W.tg_B_exact(N, RANK_B, True)
    for k in range(RANK_C):
        p = perr(k)
        H = H + terr(k, err_scale) * np.outer(p, p)
    return H


def probes(k):
    i = np.arange(N)[:, None]
    l = np.arange(k)[None, :]
    return np.sin(0.03 * (i + 1) * (l + 2)) + 0.1 * np.cos(0.7 * i - 1.3 * l)


def oracle_true_eigs(err_scale):
    s = np.sqrt(W.tg_m(N)) / W.tg_z(N)
    F = s[:, None] * hd_matrix(err_scale) * s[None, :]
    w = np.linalg.eigvalsh(F)
    order = sorted(range(N), key=lambda i: -abs(w[i]))
    return w[order]


@pytest.fixture(scope="module")
def dev(lgh):
    lgh.init()
    return lgh


def build(dev, backend=0):
    prior = dev.Prior.diagonal_callbacks(W.tg_z(N), W.tg_m(N))
    B = dev.Mat.dense(W.tg_B_exact(N, RANK_B, True))
    return dev.GLR(B, prior, dev.glr_opts(ell=40, q_power=1, trunc_abs=0.0,
                                          trunc_rel=1e-8, nb=8, panel=16, check=0,
                                          backend=backend))


def identity_battery(g, c, v):
    """tests/test_deflate.cpp:200-241: the corrected apply, solve, factor,
    sample and pow are mutually consistent actions of the same (U', lam')"""
    y = g.apply(c, v)
    assert rel(g.solve(c, y), v) < 1e-8
    x = g.factor(c, 1, v)                       # GT
    assert rel(g.factor(c, 0, x), y) < 1e-8     # G(GT(v)) == apply(v)
    y2 = g.factor(c, 1, v)
    x2 = g.factor(c, 3, y2)                     # GINVT(GT(v))
    assert rel(x2, v) < 1e-8
    assert rel(g.sample(c, y2), x2) < 1e-14
    assert rel(g.pow(c, 1.0, g.pow(c, -1.0, v)), v) < 1e-10


def check_spectrum(g, oracle, nexpect, tol):
    lam = g.lam
    assert len(lam) == nexpect
    worst = np.max(np.abs(lam - oracle[:nexpect])) / abs(oracle[0])
    assert worst < tol, worst


@pytest.mark.parametrize("backend", [0, 2])
def test_value_pass(dev, backend):
    """Scenario A (tests/test_deflate.cpp:395-478); backend 2: an implicit-U
    object, whose U is formed when the correction grafts
    (correct_impl.h:338-345)"""
    c, k, n_qc = 0.7, 30, 6
    H = hd_matrix(1.0)
    g = build(dev, backend)
    V = probes(k)
    rep = g.correct_probes(V, H @ V, lambda x: H @ x, c0=1.0, applies=8, n_qc=n_qc)
    assert rep["basis"] == RANK_C
    assert rep["hd_applies"] == RANK_C
    assert rep["clamped"] == 0
    assert rep["qc_before"] > 1e-3
    assert rep["qc_after"] < 1e-7
    orc = oracle_true_eigs(1.0)
    check_spectrum(g, orc, RANK_B + RANK_C, 1e-7)
    lmin = min(0.0, orc[:RANK_B + RANK_C].min())
    assert abs(rep["floor"] - max(0.0, -lmin)) < 1e-8
    ref = np.sum(np.log(orc[:RANK_B + RANK_C] + c) - math.log(c))
    assert abs(g.logdet(c) - ref) < 1e-8 * abs(ref)
    identity_battery(g, c, W.tg_probe(N))
    with pytest.raises(dev.LGHError):           # extend after correct
        g.extend(5)
    with pytest.raises(dev.LGHError):           # correcting twice
        g.correct_probes(V, H @ V, lambda x: H @ x, c0=1.0, applies=8, n_qc=n_qc)


def test_fresh(dev):
    """Scenario A2 (tests/test_deflate.cpp:482-509)"""
    H = hd_matrix(1.0)
    g = build(dev)
    rep = g.correct(lambda x: H @ x, c0=1.0, applies=16)
    assert rep["basis"] == RANK_C
    assert rep["clamped"] == 0
    check_spectrum(g, oracle_true_eigs(1.0), RANK_B + RANK_C, 1e-6)
    identity_battery(g, 0.7, W.tg_probe(N))


def test_clamp(dev):
    """Scenario B (tests/test_deflate.cpp:513-550): a large negative error
    mode makes H_d(c0) indefinite; the clamp engages and the object stays
    consistent above its floor"""
    c0, k = 1.0, 30
    H = hd_matrix(400.0)
    g = build(dev)
    V = probes(k)
    rep = g.correct_probes(V, H @ V, lambda x: H @ x, c0=c0, applies=8)
    assert rep["clamped"] >= 1
    assert rep["floor"] < c0
    assert rep["d_min"] >= -1.0 + 0.05 - 1e-12
    identity_battery(g, 2.0 * c0, W.tg_probe(N))


def test_threshold(dev):
    """Scenario C (tests/test_deflate.cpp:554-600): nothing clears
    d_keep_min, the object is left uncorrected and stays extendable and
    correctable"""
    k = 30
    H = hd_matrix(1.0)
    g = build(dev)
    V = probes(k)
    rep = g.correct_probes(V, H @ V, lambda x: H @ x, c0=1.0, applies=8,
                           d_keep_min=10.0)
    assert rep["kept"] == 0
    assert rep["d_max"] != 0.0
    assert len(g.lam) == RANK_B
    g.extend(5)
    rep = g.correct_probes(V, H @ V, lambda x: H @ x, c0=1.0, applies=8)
    assert rep["kept"] > 0


def test_refusals(dev):
    H = hd_matrix(1.0)
    g = build(dev)
    V = probes(8)
    with pytest.raises(dev.LGHError):           # no probes left after n_qc
        g.correct_probes(V, H @ V, lambda x: H @ x, n_qc=8)
    with pytest.raises(dev.LGHError):
        g.correct_probes(V, H @ V, lambda x: H @ x, applies=0)
    with pytest.raises(dev.LGHError):
        g.correct(lambda x: H @ x, applies=1)
    with pytest.raises(dev.LGHError):
        g.correct(lambda x: H @ x, c0=0.0)
    assert not g.report.get("corrected", 0) and len(g.lam) == RANK_B


def test_grid_correction_reduces_error(dev):
    """The correction on a 32 x 32 grid (assembled Z, mode-3 prior, blocked
    Z solves): true Hessian = sparse B + a planted signed rank-3 error, on
    top of the sketch's truncation error.  Checked against dense algebra
    (n = 1024): the priced values d lie inside the spectrum of the whitened
    error E_w = S^- (F_true - U lam U^T) S^- (Rayleigh bounds), and the
    correction lowers |E_w|_F measured in the same S^- metric."""
    Bs, Zs, M = W.psf2d_problem(32)
    n = Zs.shape[0]
    c0 = 1e-2
    prior = dev.Prior.from_mat(Zs, M, blocked_mode=3)
    g = dev.GLR(Bs, prior, dev.glr_opts(ell=24, seed=0))
    U0, lam0 = g.U, g.lam
    x1 = W.grid_coords(32, 2)[0]
    P = np.stack([np.sin(np.pi * (j + 1) * x1[:, 0]) * np.cos(np.pi * j * x1[:, 1])
                  for j in range(3)], 1)
    scale = np.abs(Bs @ np.ones(n)).max()
    t = scale * np.array([0.3, -0.2, 0.1]) / (P ** 2).sum(0)

    def hd(x):
        return Bs @ x + P @ (t * (P.T @ x))

    rng = np.random.default_rng(1)
    V = rng.standard_normal((n, 40))
    HV = np.stack([hd(V[:, j]) for j in range(40)], 1)
    rep = g.correct_probes(V, HV, hd, c0=c0, applies=12, n_qc=8)
    assert rep["basis"] > 0 and rep["kept"] > 0 and rep["hd_applies"] == 12
    # dense truth
    ms = np.sqrt(M)
    Zi = np.linalg.inv(Zs.toarray())
    Hd = Bs.toarray() + (P * t) @ P.T
    Ft = ms[:, None] * (Zi @ Hd @ Zi) * ms[None, :]
    Sm = (U0 * ((lam0 + c0) ** -0.5 - c0 ** -0.5)) @ U0.T + c0 ** -0.5 * np.eye(n)
    E0 = Sm @ (Ft - (U0 * lam0) @ U0.T) @ Sm
    E0 = 0.5 * (E0 + E0.T)
    w = np.linalg.eigvalsh(E0)
    tol = 1e-8 * np.abs(w).max()
    assert w[0] - tol <= rep["d_min"] and rep["d_max"] <= w[-1] + tol
    U1, lam1 = g.U, g.lam
    E1 = Sm @ (Ft - (U1 * lam1) @ U1.T) @ Sm
    assert np.linalg.norm(E1) < 0.99 * np.linalg.norm(E0)
    # the battery's tolerances are the reference's (set at c = 0.7); at
    # c = c0 = 1e-2 the spread lam_max / c ~ 1e6 puts pow o pow^-1 at ~1e-10
    identity_battery(g, max(2.0 * rep["floor"], 1.0), rng.standard_normal(n))
