"""The C-ABI library loads and exports every function declared in
include/*.h (no GPU calls), and the host-side pieces that need no GPU."""
import glob
import math
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import ROOT
import workloads as W

DECL = re.compile(r"^(?:[A-Za-z_][\w\s\*]*?[\s\*])?([A-Za-z_]\w*)\s*\(", re.M)
SKIP = {"if", "for", "while", "switch", "return", "sizeof", "typedef"}


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        src = re.sub(r"#define[^\n]*(\\\n[^\n]*)*", "", src)
        src = re.sub(r"typedef[^;]*;", "", src)
        for m in DECL.finditer(src):
            nm = m.group(1)
            if nm not in SKIP and not nm.isupper():
                names.add(nm)
    return names


def test_every_declared_symbol_is_exported(lgh):
    L = lgh.lib()
    names = declared_functions()
    assert "lgh_glr_compute" in names and "lgh_hip_spmm" in names
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing


def test_opts_defaults_match_reference(lgh):
    o = lgh.glr_opts()          # glr_common_impl.h:295-312
    assert (o.ell, o.q_power, o.trunc_abs, o.trunc_rel, o.treat, o.seed,
            o.backend, o.nb, o.panel, o.grid2d, o.check) == \
        (200, 1, 0.0, 1e-3, 0, 271828, 0, 64, 64, 0, 1)
    p = lgh.prior_mat_opts()    # glr_common_impl.h:99-112
    assert (p.ksp_rtol, p.blocked_mode, p.tile, p.cheb_rtol, p.smax_krylov,
            p.smoother_top, p.nu_force, p.verbose) == (1e-12, 0, 64, 1e-10, 1, 1.1, 0, 0)


def test_device_randn_arithmetic_on_host(lgh, oracle):
    """The device draw arithmetic (correctly rounded log) evaluated on the
    host equals the oracle's glibc-log draw except where glibc's log is not
    correctly rounded; there the draw differs by one rounding of log(q)."""
    seeds = [0, 271828, 0xB0BD5]
    nbad = tot = 0
    worst = 0.0
    for s in seeds:
        ref = oracle.randn_fill(s, 0, 0, 2000, 20)
        for j in range(20):
            for i in range(0, 2000, 7):
                a = lgh.randn_at_host(s, i, j)
                tot += 1
                if a != ref[i, j]:
                    nbad += 1
                    worst = max(worst, abs(a - ref[i, j]) / abs(ref[i, j]))
    assert nbad / tot < 2e-3
    assert worst < 4e-16


@pytest.mark.parametrize("seed,n,c0,c1", [
    (0, 16384, 0, 74),              # the 128 x 128 grid's whole sketch (seed 0, l = 74)
    (271828, 1 << 20, 0, 400),      # the 1024 x 1024 grid's whole sketch (l = 400)
    (271828, 1 << 21, 400, 416),    # a slice of the 128^3 grid's extension columns
])
def test_randn_flag_window_covers_platform_log(lgh, oracle, seed, n, c0, c1):
    """Every draw of the device arithmetic that differs from the glibc-log
    draw (the oracle) is flagged for the platform-log finish, in FP64 and in
    FP32; and the flags stay few (2.2 % of FP64 draws)."""
    nflag = ndiff = tot = 0
    for a in range(c0, c1, 16):
        b = min(a + 16, c1)
        ref = oracle.randn_fill(seed, 0, a, n, b - a)
        v, f = lgh.randn_host_block(seed, 0, a, n, b - a, prec=lgh.PREC_F64)
        d = v != ref
        assert not (d & ~f).any(), "unflagged FP64 draw differs from the platform log"
        v32, f32 = lgh.randn_host_block(seed, 0, a, n, b - a, prec=lgh.PREC_F32)
        d32 = v32 != ref.astype(np.float32).astype(np.float64)
        assert not (d32 & ~f32).any(), "unflagged FP32 draw differs from the platform log"
        nflag += int(f.sum())
        ndiff += int(d.sum())
        tot += d.size
    assert ndiff > 0                      # the window is exercised
    assert nflag / tot < 0.03, nflag / tot


def test_amg_hierarchy_properties(lgh):
    Z, M = W.prior_Z(48, 2, 1.0, 0.1)
    h = lgh.AMG(Z)
    lv = h.levels()
    assert len(lv) >= 3
    assert lv[-1]["A"].shape == Z.shape
    assert abs(lv[-1]["A"] - Z).max() == 0.0
    for l in range(1, len(lv)):
        P, A, Ac = lv[l]["P"], lv[l]["A"], lv[l - 1]["A"]
        # Galerkin coarse operator
        G = (P.T @ A @ P).toarray()
        assert np.allclose(Ac.toarray(), G, rtol=1e-12, atol=1e-14 * abs(G).max())
        # prolongator has full column rank and every row is touched
        assert np.linalg.matrix_rank(P.toarray() if P.shape[1] < 2000 else P[:, :50].toarray()) == min(P.shape[1], 50 if P.shape[1] >= 2000 else P.shape[1])
        # smoother interval brackets lambda_max(D^-1 A) from below/above
        d = A.diagonal()
        lmax = np.max(np.linalg.eigvals(np.diag(1 / d) @ A.toarray()).real) if A.shape[0] <= 3000 else None
        if lmax is not None:
            assert lv[l]["shi"] >= lmax * 0.999
            assert math.isclose(lv[l]["slo"], 0.1 * lv[l]["shi"] / 1.1, rel_tol=1e-12)
        assert lv[l]["nu"] == 2
    assert lv[0]["A"].shape[0] <= 2000


def _oracle_levels(oracle, h):
    return [{"A": oracle.CSR.from_scipy(l["A"]),
             "P": oracle.CSR.from_scipy(l["P"]) if l["P"] is not None else None,
             "slo": l["slo"], "shi": l["shi"], "nu": l["nu"]} for l in h.levels()]


def test_oracle_w_cycle_properties(lgh, oracle):
    """The oracle's W-cycle (cycle index 2) pinned by what the mathematics
    fixes: it is a symmetric operator (the second coarse correction in
    correction form keeps the cycle self-adjoint); on a two-level hierarchy
    (coarse level solved exactly) it equals the V-cycle bit for bit; on
    three or more levels it reduces the A-norm of the error of one
    application below the V-cycle's."""
    rng = np.random.default_rng(11)
    Z, M = W.prior_Z(48, 2, 1.0, 0.1)
    h = lgh.AMG(Z)
    lv = _oracle_levels(oracle, h)
    assert len(lv) >= 3
    Wc, Vc = oracle.MG(lv, gamma=2), oracle.MG(lv, gamma=1)
    x, y = rng.standard_normal((2, Z.shape[0]))
    a, b = x @ Wc.vcycle(y)[:, 0], y @ Wc.vcycle(x)[:, 0]
    assert abs(a - b) <= 1e-12 * abs(a)
    e = rng.standard_normal(Z.shape[0])
    def anorm(v):
        return math.sqrt(v @ (Z @ v))
    ew = e - Wc.vcycle(Z @ e)[:, 0]
    ev = e - Vc.vcycle(Z @ e)[:, 0]
    assert anorm(ew) < anorm(ev) < anorm(e)
    two = lv[-2:]
    two[0] = dict(two[0], P=None, slo=0.0, shi=0.0, nu=0)
    B = rng.standard_normal((Z.shape[0], 3))
    assert np.array_equal(oracle.MG(two, gamma=2).vcycle(B), oracle.MG(two).vcycle(B))


def test_api_refusals_without_gpu(lgh):
    """Argument checks that precede any device work (glr_common_impl.h:651-664)."""
    L = lgh.lib()
    B = lgh.Mat.aij(sp.identity(10, format="csr"))
    mass = lgh.Vec(np.ones(10))
    import ctypes as C
    h = C.c_void_p()
    rc = L.lgh_prior_create_ksp(None, mass.h, None, C.byref(h))
    assert rc == 56                                           # PETSC_ERR_SUP
    cb = lgh.Callbacks()
    assert L.lgh_prior_create_callbacks(mass.h, C.byref(cb), C.byref(h)) == 85
    pr = lgh.Prior.diagonal_callbacks(np.ones(10), np.ones(10))
    o = lgh.glr_opts(ell=11)
    g = C.c_void_p()
    assert L.lgh_glr_compute(B.h, pr.h, C.byref(o), C.byref(g), None) == 63
    # a negative sweep chunk is refused before any device work
    o = lgh.glr_opts(ell=5, backend=2)
    x = lgh.xopts(sweep_panels=-1)
    assert L.lgh_glr_compute_x(B.h, pr.h, C.byref(o), C.byref(x), C.byref(g),
                               None) == 63
    # batches: nothing to do, and missing arrays (PETSC_ERR_ARG_NULL)
    assert L.lgh_glr_compute_batch(0, None, None, None, None, 0, None, None) == 0
    assert L.lgh_glr_compute_batch(2, None, None, None, None, 0, None, None) == 85
    assert L.lgh_prior_create_mat_batch(2, None, None, None, 0, None) == 85
