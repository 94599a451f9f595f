"""Kernel-level parity of the gfx950 kernels through the lgh_hip_* C ABI
against the FP64 oracle / numpy (run on an MI355X)."""
import numpy as np
import pytest
import scipy.sparse as sp

import workloads as W

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def dev(lgh):
    lgh.init()
    return lgh


def rand_csr(n, m, density, seed):
    A = sp.random(n, m, density=density, random_state=seed, format="csr")
    A.data = np.random.default_rng(seed).standard_normal(A.nnz)
    return A


@pytest.mark.parametrize("Wd", [64, 8, 1])
def test_randn_bitexact_vs_reference_draw(dev, oracle, Wd):
    """The device sketch equals the reference's glibc-log draw (the oracle)
    bit for bit, in FP64 and (rounded once) in FP32, for any panel width."""
    nr, nc = 1000, 70
    got = dev.randn_fill(271828, 5, 3, nr, nc, W=Wd)
    ref = oracle.randn_fill(271828, 5, 3, nr, nc)
    assert np.array_equal(got, ref)
    f32 = dev.randn_fill(271828, 5, 3, nr, nc, W=Wd, prec=dev.PREC_F32)
    assert np.array_equal(f32, ref.astype(np.float32).astype(np.float64))
    # the draws the device arithmetic rounds differently are among them
    host, flags = dev.randn_host_block(271828, 5, 3, nr, nc)
    assert (host != ref).any() and flags.any()


def test_randn_full_grid128_sketch_bitexact(dev, oracle):
    """The whole sketch of the 128 x 128 grid (seed 0, l = 74) and the
    prior's right-hand side equal the reference's draws bit for bit."""
    got = dev.randn_fill(0, 0, 0, 16384, 74)
    ref = oracle.randn_fill(0, 0, 0, 16384, 74)
    assert np.array_equal(got, ref)
    # extension columns continue j past l
    ext = dev.randn_fill(0, 0, 74, 16384, 20)
    assert np.array_equal(ext, oracle.randn_fill(0, 0, 74, 16384, 20))
    b = dev.randn_fill(0xB0BD5, 0, 0, 16384, 1)
    assert np.array_equal(b, oracle.randn_fill(0xB0BD5, 0, 0, 16384, 1))


def test_randn_large_fill_bitexact_sample(dev, oracle):
    """A 1024 x 1024 grid's first sketch panel (2^26 draws, ~1.5 M of them
    finished with the platform log): rows sampled across the panel."""
    n = 1 << 20
    got = dev.randn_fill(271828, 0, 0, n, 64)
    rows = np.arange(0, n, 97)
    ref = np.stack([oracle.randn_fill(271828, int(i), 0, 1, 64)[0] for i in rows[:2000]])
    assert np.array_equal(got[rows[:2000]], ref)
    blk = oracle.randn_fill(271828, n - 4096, 0, 4096, 64)
    assert np.array_equal(got[n - 4096:], blk)


@pytest.mark.parametrize("Wd", [64, 16, 1])
def test_spmm(dev, Wd):
    A = rand_csr(1003, 997, 0.01, 1)
    X = np.random.default_rng(2).standard_normal((997, 70))
    ref = A @ X
    assert rel(dev.spmm(A, X, W=Wd), ref) < 1e-13
    y32 = dev.spmm(A, X, prec_vals=dev.PREC_F32, prec_vec=dev.PREC_F32, W=Wd)
    assert rel(y32, ref) < 1e-5


@pytest.mark.parametrize("prec_vals", [0, 1])
def test_spmm_wide_rows_width_invariant(dev, prec_vals):
    """A PSF operator with wide rows (the 2D isotropic PSF, up to 49 entries a row):
    the 64-wide row kernel (grouped neighbour loads) is bit-identical to
    the generic 32-wide kernel (the same sums in the same order) and
    matches scipy."""
    B, _, _ = W.psf2d_problem(45)
    assert B.nnz >= 32 * B.shape[0]
    X = np.random.default_rng(5).standard_normal((B.shape[0], 150))
    y64 = dev.spmm(B, X, prec_vals=prec_vals, W=64)
    y32 = dev.spmm(B, X, prec_vals=prec_vals, W=32)
    assert np.array_equal(y64, y32)
    ref = B @ X if prec_vals == 0 else B.astype(np.float32).astype(np.float64) @ X
    assert rel(y64, ref) < 1e-13
    # the row-grouped copy (n = 2025 leaves a last group of 1 row): the same
    # sums in the same order plus exact zeros, FP64 and FP32 vectors, for
    # the autotuner's choice and for every candidate it times (4 or 8 rows
    # a group, 2 or 4 panels a wavefront with FP64 vectors, 4 or 8 with
    # FP32; 150 columns = 2 whole panels + 22, so the last launch has
    # panels without data)
    assert B.shape[0] % 8 == 1
    a32 = dev.spmm(B, X, prec_vals=1, prec_vec=dev.PREC_F32, W=64)
    try:
        for rg, ppw in [(0, 0), (4, 2), (4, 4), (8, 2), (8, 4)]:
            dev.spmm_group_config(rg, ppw)
            got = dev.spmm(B, X, prec_vals=prec_vals, W=64, grouped=True)
            assert np.array_equal(got, y64), (rg, ppw)
            b32 = dev.spmm(B, X, prec_vals=1, prec_vec=dev.PREC_F32, W=64, grouped=True)
            assert np.array_equal(a32, b32), (rg, ppw)
        # the row-tile copy (LDS-staged union of 64 rows): every shape, 1D
        # and 2D tiles (the 45-wide grid's line stride; partial tiles at
        # the grid's right edge and bottom), plain and XCD-run tile order;
        # candidates 16-31 run the FP64 products on the matrix cores (FP32
        # vectors fall back to the row kernel)
        for cand in range(32):
            dev.spmm_tile_config(cand)
            got = dev.spmm(B, X, prec_vals=prec_vals, W=64, grouped=True)
            assert np.array_equal(got, y64), cand
            b32 = dev.spmm(B, X, prec_vals=1, prec_vec=dev.PREC_F32, W=64, grouped=True)
            assert np.array_equal(a32, b32), cand
    finally:
        dev.spmm_group_config(0, 0)
        dev.spmm_tile_config(-1)
    rep = [r for r in dev.spmm_group_report() if r["n"] == B.shape[0]]
    assert rep and (rep[0]["rows_per_group"] in (4, 8) or rep[0]["tile"] == 1)


@pytest.mark.parametrize("prec", [0, 1])
def test_spmm_long_rows(dev, prec):
    """Rows of hundreds of entries (a restriction onto a small coarse level)
    take the chunked kernel: the same bits at panel widths 64 and 16, and
    scipy's product; an empty row and a one-entry row included."""
    A = rand_csr(90, 4000, 0.08, 7).tolil()
    A[3, :] = 0
    A[5, :] = 0
    A[5, 17] = 2.5
    A = A.tocsr()
    A.eliminate_zeros()
    assert A.nnz >= 64 * A.shape[0]
    X = np.random.default_rng(8).standard_normal((4000, 70))
    y64 = dev.spmm(A, X, prec_vals=prec, prec_vec=prec, W=64)
    y16 = dev.spmm(A, X, prec_vals=prec, prec_vec=prec, W=16)
    assert np.array_equal(y64, y16)
    ref = A @ X
    assert rel(y64, ref) < (1e-13 if prec == 0 else 1e-5)
    assert np.all(y64[3] == 0)


def test_cheb_jacobi_matches_oracle(dev, oracle):
    Z = W.tpm_Z()
    n = Z.shape[0]
    A = oracle.CSR.from_scipy(Z)
    b = W.tg_probe(n)
    B = np.stack([b * (1 + 0.1 * j) for j in range(11)], 1)
    dinv = 1.0 / Z.diagonal()
    lmin, lmax, nit = 0.1, 1.9, 37
    ref = oracle.cheb_jacobi(A, dinv, lmin, lmax, nit, B)
    for Wd in (64, 8):
        assert rel(dev.cheb_jacobi(Z, lmin, lmax, nit, B, W=Wd), ref) < 1e-13
    got32 = dev.cheb_jacobi(Z, lmin, lmax, nit, B, prec=dev.PREC_F32)
    assert rel(got32, ref) < 1e-5


@pytest.mark.parametrize("prec,tol", [(0, 1e-13), (1, 1e-5)])
def test_gram_and_gemm(dev, prec, tol):
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1003, 70))
    Y = rng.standard_normal((1003, 130))
    G = dev.gram(X, Y, prec=prec)
    assert rel(G, X.T @ Y) < tol
    Cm = rng.standard_normal((70, 130))
    assert rel(dev.gemm_nn(X, Cm, prec=prec), X @ Cm) < tol
    S = dev.gram(X, X, prec=prec)
    assert rel(S, X.T @ X) < tol


def test_gram_gemm_many_tiles(dev):
    """Row counts that leave partial row tiles and partial groups of eight
    tiles / chunks (the block order groups them per XCD), the symmetric
    Gram, and a block upper triangular C (the CholQR product, which skips
    the zero blocks) against numpy."""
    rng = np.random.default_rng(4)
    n = 40_037
    X = rng.standard_normal((n, 300))
    Y = rng.standard_normal((n, 70))
    assert rel(dev.gram(X, Y), X.T @ Y) < 1e-13
    assert rel(dev.gram(X, X), X.T @ X) < 1e-13
    Cu = np.triu(rng.standard_normal((300, 300)))
    assert rel(dev.gemm_nn(X, Cu), X @ Cu) < 1e-13
    Cb = Cu.copy()
    Cb[200, 10] = 1.0                       # one entry below the blocks
    assert rel(dev.gemm_nn(X, Cb), X @ Cb) < 1e-13
    Cr = rng.standard_normal((300, 130))
    assert rel(dev.gemm_nn(X, Cr), X @ Cr) < 1e-13


@pytest.mark.parametrize("gamma,dim", [(1, 2), (2, 2), (2, 3)])
def test_vcycle_matches_oracle(dev, oracle, gamma, dim):
    """V-cycle (gamma 1) and W-cycle (gamma 2) on a hierarchy of at least
    three levels, against the oracle cycle on the exported hierarchy; in
    the 3D hierarchy the restriction onto the 117-row coarsest level (164
    entries a row) takes the long-row kernel"""
    Z, M = W.prior_Z(40 if dim == 2 else 32, dim, 1.0, 0.1)
    h = dev.AMG(Z)
    h.set_cycle(gamma)
    levels = [{"A": oracle.CSR.from_scipy(l["A"]),
               "P": oracle.CSR.from_scipy(l["P"]) if l["P"] is not None else None,
               "slo": l["slo"], "shi": l["shi"], "nu": l["nu"]} for l in h.levels()]
    assert len(levels) >= 3
    mg = oracle.MG(levels, gamma=gamma)
    if gamma == 2:      # the two cycles differ on this hierarchy
        Bt = np.random.default_rng(4).standard_normal((Z.shape[0], 2))
        assert rel(mg.vcycle(Bt), oracle.MG(levels).vcycle(Bt)) > 1e-6
    B = np.random.default_rng(4).standard_normal((Z.shape[0], 5))
    ref = mg.vcycle(B)
    assert rel(h.vcycle(B), ref) < 1e-12
    # FP32: 2^-24 per operation, grown by the cycle (the 3D W-cycle sums
    # rows of 27-164 entries over four visits of its middle level: 2e-5)
    assert rel(h.vcycle(B, prec=dev.PREC_F32), ref) < (1e-5 if dim == 2 else 1e-4)
    assert rel(h.vcycle(B, W=4), ref) < 1e-12
    # BF16 storage (8-bit mantissa, 2^-9 relative rounding per stored value)
    assert rel(h.vcycle(B, prec=dev.PREC_BF16), ref) < 2e-2


@pytest.mark.parametrize("prec,dim", [(0, 2), (1, 2), (2, 2), (0, 3), (2, 3)])
def test_vcycle_narrow_panel_bitwise(dev, prec, dim):
    """16- and 8-wide panels (the narrow last panels of a block) run the
    same per-column arithmetic as a 64-wide one: identical bits, at every
    storage precision (in 3D with the long-row restriction onto the
    coarsest level)."""
    Z, M = W.prior_Z(64 if dim == 2 else 32, dim, 1.0, 0.1)
    h = dev.AMG(Z)
    B = np.random.default_rng(5).standard_normal((Z.shape[0], 13 if dim == 2 else 70))
    x64 = h.vcycle(B, prec=prec, W=64)
    x16 = h.vcycle(B, prec=prec, W=16)
    x8 = h.vcycle(B, prec=prec, W=8)
    assert np.array_equal(x64, x16)
    assert np.array_equal(x64, x8)


@pytest.mark.parametrize("prec", [0, 1, 2])
@pytest.mark.parametrize("gamma", [1, 2])
def test_vcycle_fused_coarse_bitwise(dev, prec, gamma):
    """The small levels below the 16 384-row finest level run as one fused
    launch per coarse visit with the bits of one launch per operation and
    level: the 319-row level and the 38-row dense solve (with a 4096-row
    cap too: the 2 752-row level's slices would not fit the launch's 64 KB
    of LDS, so it keeps its own launches); 64-wide (two panels), 16-wide
    (two panels) and 8-wide panels, V- and W-cycles, every storage
    precision."""
    Z, M = W.prior_Z(128, 2, 1.0, 0.1)
    h = dev.AMG(Z)
    sizes = [l["A"].shape[0] for l in h.levels()]
    assert len(sizes) == 4 and sizes[1] <= 1024 < sizes[2] <= 4096 < sizes[3]
    h.set_cycle(gamma)
    rng = np.random.default_rng(11 + prec)
    for Wd, m in ((64, 100), (16, 30), (8, 5)):
        B = rng.standard_normal((Z.shape[0], m))
        dev.prof_enable(True)
        try:
            dev.prof_collect("mg_coarse")
            dev.mg_coarse_config(0)
            xp = h.vcycle(B, prec=prec, W=Wd)
            per_level = dev.prof_collect("mg_coarse")[0]
            for cap in (-1, 4096):
                dev.mg_coarse_config(cap)
                xf = h.vcycle(B, prec=prec, W=Wd)
                fused = dev.prof_collect("mg_coarse")[0]
                # one launch per visit of the level above the fused ones
                # (the W-cycle visits the 2 752-row level twice)
                want = 2 if gamma == 2 else 1
                assert fused == want and per_level == 0, (Wd, cap, fused, per_level)
                assert np.array_equal(xf, xp), (Wd, cap, np.abs(xf - xp).max())
        finally:
            dev.mg_coarse_config(-1, -1)
            dev.prof_enable(False)


@pytest.mark.gpu
@pytest.mark.parametrize("n,m", [(1_000_003, 13), (1 << 20, 16), (1 << 20, 25)])
def test_large_host_blocks_round_trip(dev, n, m):
    """Host blocks of 64 MiB and more move through two pinned chunks per
    thread (104 MB: one whole chunk and a partial one; 128 MiB: exactly
    two; 200 MiB: three whole chunks and a partial fourth, so each chunk is
    reused, the host-to-device side waiting for its earlier copy and the
    device-to-host side alternating DMA and copy-out): Y = 2 I X through the
    SpMM entry point comes back exact."""
    A = sp.identity(n, format="csr") * 2.0
    X = np.random.default_rng(n).standard_normal((n, m))
    assert X.nbytes >= 64 << 20
    Y = dev.spmm(A, X)
    assert np.array_equal(Y, 2.0 * X)
