"""Synthetic inputs for the reference's closed-form test problems and the
benchmark configurations (pure numpy/scipy, host side).

* test_glr problem: tests/test_glr.cpp:43-65, 140-141 (N = 120, rank-10 B,
  diagonal Z and mass).
* test_prior_mat problem: tests/test_prior_mat.cpp:35-73 (24 x 24 5-point
  Laplacian + 0.5 I).
* grid problems: a uniform grid on the unit square/cube with lumped mass,
  the 5-/7-point Neumann stiffness (diag = neighbour count, as
  examples/heat/heat_source_inversion.cpp:321-335) and Gaussian point-spread
  B rows, for the benchmark configurations.

Every generator is deterministic; randomness (right-hand sides, amplitude
fields) is drawn from the hashed Gaussian of the oracle construction so the
CPU oracle and the device see identical inputs.
"""
import math

import numpy as np
import scipy.sparse as sp

# --------------------------------------------------------------------------
# tests/test_glr.cpp


def tg_factor_entry(k, i):
    return np.sin(0.1 * (i + 1) * (k + 3)) + 0.2 * np.cos(0.37 * i * k + k)


def tg_B(n=120, rank=10, spd=True):
    i = np.arange(n)
    B = np.zeros((n, n))
    for k in range(rank):
        sk = 2.0 / (k + 1.0)
        if not spd and (k % 2):
            sk = -sk
        a = tg_factor_entry(k, i)
        B += sk * np.outer(a, a)
    return B


def tg_B_exact(n=120, rank=10, spd=True):
    """B accumulated entry by entry in the order of tests/test_glr.cpp:50-59
    (so the dense matrix is bitwise the reference's)."""
    i = np.arange(n)
    fe = np.array([tg_factor_entry(k, i) for k in range(rank)])
    B = np.zeros((n, n))
    for k in range(rank):
        sk = 2.0 / (k + 1.0)
        if not spd and (k % 2):
            sk = -sk
        B = B + sk * fe[k][:, None] * fe[k][None, :]
    return B


def tg_z(n=120):
    return 1.0 + 0.01 * np.arange(n)


def tg_m(n=120):
    return 0.5 + 0.1 * (np.arange(n) % 7)


def tg_probe(n=120):
    return np.cos(0.05 * np.arange(n) + 1.0)


def tg_dense_F(spd=True, n=120, rank=10):
    m, z = tg_m(n), tg_z(n)
    s = np.sqrt(m) / z
    return s[:, None] * tg_B_exact(n, rank, spd) * s[None, :]


# --------------------------------------------------------------------------
# tests/test_prior_mat.cpp


def tpm_Z(m=24, shift=0.5):
    """2D 5-point Laplacian + shift*I (diag 4 + shift, off -1), Dirichlet-
    style (no boundary correction), exactly tests/test_prior_mat.cpp:56-67."""
    n = m * m
    rows, cols, vals = [], [], []
    for i in range(n):
        ix, iy = i % m, i // m
        rows.append(i); cols.append(i); vals.append(4.0 + shift)
        if ix > 0:
            rows.append(i); cols.append(i - 1); vals.append(-1.0)
        if ix < m - 1:
            rows.append(i); cols.append(i + 1); vals.append(-1.0)
        if iy > 0:
            rows.append(i); cols.append(i - m); vals.append(-1.0)
        if iy < m - 1:
            rows.append(i); cols.append(i + m); vals.append(-1.0)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(n, n))
    A.sort_indices()
    return A


def tpm_B(n, rank=6):
    i = np.arange(n)
    B = np.zeros((n, n))
    for k in range(rank):
        a = tg_factor_entry(k, i)
        B = B + 2.0 / (k + 1.0) * a[:, None] * a[None, :]
    return B


# --------------------------------------------------------------------------
# grid problems (benchmark configurations)


def grid_coords(side, dim):
    h = 1.0 / (side - 1)
    idx = np.indices((side,) * dim).reshape(dim, -1)[::-1]   # x fastest
    return idx.T.astype(np.float64) * h, h


def lumped_mass(side, dim):
    """h^dim per node, halved along each boundary direction (edges h^d/2,
    2D corners h^2/4: the tensor-product trapezoid rule)."""
    h = 1.0 / (side - 1)
    w1 = np.full(side, h)
    w1[0] = w1[-1] = 0.5 * h
    m = w1
    for _ in range(dim - 1):
        m = np.multiply.outer(w1, m).ravel()
    return m.ravel()


def neumann_laplacian(side, dim):
    """Unscaled FD graph Laplacian (diag = neighbour count, off = -1) with
    natural boundary: K x = sum over grid neighbours (x_i - x_j)."""
    n = side ** dim
    idx = np.arange(n).reshape((side,) * dim)
    rows, cols = [], []
    for ax in range(dim):
        a = np.take(idx, np.arange(side - 1), axis=dim - 1 - ax).ravel()
        b = np.take(idx, np.arange(1, side), axis=dim - 1 - ax).ravel()
        rows += [a, b]
        cols += [b, a]
    r = np.concatenate(rows)
    c = np.concatenate(cols)
    off = sp.csr_matrix((-np.ones(len(r)), (r, c)), shape=(n, n))
    deg = -np.asarray(off.sum(axis=1)).ravel()
    K = (off + sp.diags(deg)).tocsr()
    K.sort_indices()
    return K


def prior_Z(side, dim, mass_coef, k_coef):
    M = lumped_mass(side, dim)
    K = neumann_laplacian(side, dim)
    Z = (k_coef * K + sp.diags(mass_coef * M)).tocsr()
    Z.sort_indices()
    return Z, M


def psf_B(side, dim, offsets_fn, value_fn, dtype=np.float64):
    """B rows from a list of integer offsets (same for every node, clipped at
    the boundary) and a vectorized value function value_fn(i, j, d) of the
    row/column node indices and the offset vector, evaluated per offset.  Offsets are ordered so that column indices grow
    with the offset index, which makes the CSR rows sorted by construction
    (no COO sort)."""
    n = side ** dim
    X, h = grid_coords(side, dim)
    grid = np.indices((side,) * dim).reshape(dim, -1)[::-1].T  # (n, dim) ints
    offs = np.asarray(offsets_fn(h))
    strides = np.array([side ** a for a in range(dim)])
    offs = offs[np.argsort((offs * strides).sum(1), kind="stable")]
    vals = np.zeros((n, len(offs)), dtype=np.float64)
    for k, off in enumerate(offs):
        tgt = grid + off[None, :]
        ok = np.all((tgt >= 0) & (tgt < side), axis=1)
        i = np.nonzero(ok)[0]
        j = (tgt[ok] * strides).sum(axis=1)
        vals[i, k] = value_fn(i, j, np.broadcast_to(off * h, (len(i), dim)))
    mask = vals != 0.0
    rp = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(mask.sum(axis=1), out=rp[1:])
    ii, kk = np.nonzero(mask)
    ci = (ii + (offs[kk] * strides).sum(axis=1)).astype(np.int32)
    v = vals[ii, kk].astype(dtype)
    del vals, mask, ii, kk
    return sp.csr_matrix((v, ci, rp.astype(np.int32)), shape=(n, n))


def _ball_offsets(radius_cells, dim):
    R = int(math.floor(radius_cells + 1e-9))
    rng = np.arange(-R, R + 1)
    g = np.stack(np.meshgrid(*([rng] * dim), indexing="ij"), -1).reshape(-1, dim)
    return g[(g ** 2).sum(1) <= radius_cells ** 2 + 1e-9]


def psf2d_problem(side=128):
    """2D, Z = 1.0 M + 0.1 K, B: symmetrized Gaussian PSF
    a(x) exp(-|x-y|^2 / (2 (2h)^2)), radius 4h, symmetrized as
    (a(x_i) + a(x_j)) / 2, a(x) = 1 + 0.5 sin(2 pi x1) cos(2 pi x2)."""
    Z, M = prior_Z(side, 2, 1.0, 0.1)

    X, _ = grid_coords(side, 2)
    a = 1.0 + 0.5 * np.sin(2 * np.pi * X[:, 0]) * np.cos(2 * np.pi * X[:, 1])
    s = 2.0 / (side - 1)

    def val(i, j, d):
        return 0.5 * (a[i] + a[j]) * np.exp(-(d ** 2).sum(1) / (2 * s * s))

    B = psf_B(side, 2, lambda h: _ball_offsets(4.0, 2), val)
    return B, Z, M


def aniso_psf2d_problem(side=1024, dtype=np.float64):
    """2D anisotropic PSF: Sigma(x) = R(pi x1) diag((3h)^2, (1.5h)^2) R^T,
    truncated at Mahalanobis distance 3 (of the row's own covariance),
    symmetrized by averaging the two rows' values.  Z = M + 0.05 K."""
    Z, M = prior_Z(side, 2, 1.0, 0.05)
    h = 1.0 / (side - 1)

    X, _ = grid_coords(side, 2)
    th = np.pi * X[:, 0]
    c, s = np.cos(th), np.sin(th)
    a, b = 1.0 / (3 * h) ** 2, 1.0 / (1.5 * h) ** 2
    # Sigma^{-1} = R diag(a, b) R^T with R = [[c, -s], [s, c]]
    P_, Q_, R_ = a * c * c + b * s * s, (a - b) * c * s, a * s * s + b * c * c

    def val(i, j, d):
        out = np.zeros(len(i))
        for t in (i, j):
            p, q, r = P_[t], Q_[t], R_[t]
            md2 = p * d[:, 0] ** 2 + 2 * q * d[:, 0] * d[:, 1] + r * d[:, 1] ** 2
            out += np.where(md2 <= 9.0, np.exp(-0.5 * md2), 0.0)
        return 0.5 * out

    B = psf_B(side, 2, lambda hh: _ball_offsets(9.0, 2), val, dtype)
    return B, Z, M


def psf3d_problem(side=128):
    """3D isotropic PSF sigma = 1.5h truncated to a radius-2h ball, Z =
    M + 0.1 K7 (unscaled 7-point); B's values are FP32 numbers (the
    configuration's single-precision operator), held in a float64 CSR."""
    Z, M = prior_Z(side, 3, 1.0, 0.1)
    s = 1.5 / (side - 1)

    def val(i, j, d):
        return np.exp(-(d ** 2).sum(1) / (2 * s * s))

    B = psf_B(side, 3, lambda h: _ball_offsets(2.0, 3), val, dtype=np.float32)
    return B.astype(np.float64), Z, M


def small_psf2d_instance(k, seed=5):
    """Instance k of the batched configuration (256 small 2D problems):
    side s = round(exp(U(ln 64, ln 256))) (n = s^2 in [4096, 65536]), h =
    1/(s-1), PSF width sigma ~ U[1.5h, 3h] truncated at radius 3 sigma,
    amplitude field a(x) = 1 + 0.5 sin(2 pi (f1 x1 + p1)) cos(2 pi (f2 x2 +
    p2)) with f ~ U[0.5, 2], p ~ U[0, 1), B_ij = (a_i + a_j)/2 exp(-|x_i -
    x_j|^2 / (2 sigma^2)), Z = M + 0.1 K, rank r in {32, 64, 128} (uniform)
    and ell = r + 10.  Every draw comes from numpy's generator seeded with
    (seed, k), so instance k is the same whatever the batch.
    Returns (B, Z, M, ell)."""
    rng = np.random.default_rng([seed, k])
    side = int(round(math.exp(rng.uniform(math.log(64), math.log(256)))))
    h = 1.0 / (side - 1)
    sig = rng.uniform(1.5, 3.0)                 # in units of h
    f = rng.uniform(0.5, 2.0, size=2)
    ph = rng.uniform(0.0, 1.0, size=2)
    rank = int(rng.choice([32, 64, 128]))
    Z, M = prior_Z(side, 2, 1.0, 0.1)
    X, _ = grid_coords(side, 2)
    a = 1.0 + 0.5 * np.sin(2 * np.pi * (f[0] * X[:, 0] + ph[0])) \
        * np.cos(2 * np.pi * (f[1] * X[:, 1] + ph[1]))
    s = sig * h

    def val(i, j, d):
        return 0.5 * (a[i] + a[j]) * np.exp(-(d ** 2).sum(1) / (2 * s * s))

    B = psf_B(side, 2, lambda hh: _ball_offsets(3.0 * sig, 2), val)
    return B, Z, M, rank + 10
