"""ctypes bindings of the FP64 CPU oracle (oracle/lgh_oracle.c).

TEST INFRASTRUCTURE: only tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py import this package.  It is the parity
reference and the CPU baseline, never part of the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liblgh_oracle.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


class GlrOpts(C.Structure):
    _fields_ = [("ell", C.c_int), ("q_power", C.c_int),
                ("trunc_abs", C.c_double), ("trunc_rel", C.c_double),
                ("treat", C.c_int), ("seed", C.c_uint64), ("check", C.c_int),
                ("backend", C.c_int)]


class GlrReport(C.Structure):
    _fields_ = [("kept", C.c_int), ("lam_abs_max", C.c_double),
                ("lam_raw_min", C.c_double), ("lam_raw_max", C.c_double),
                ("next_abs", C.c_double), ("n_negative_raw", C.c_int),
                ("resid_max", C.c_double), ("t_operator", C.c_double),
                ("t_dense", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def _declare(L):
    L.orc_randn_at.restype = C.c_double
    L.orc_randn_at.argtypes = [C.c_uint64, C.c_int64, C.c_int64]
    L.orc_randn_fill.argtypes = [C.c_uint64, C.c_int64, C.c_int64, C.c_int64,
                                 C.c_int64, dp, C.c_int64]
    L.orc_set_threads.argtypes = [C.c_int]
    L.orc_get_threads.restype = C.c_int
    L.orc_csr_mm.argtypes = [C.c_int, ip, ip, dp, C.c_int, dp, C.c_int, dp,
                             C.c_int]
    L.orc_mg_create.restype = C.c_void_p
    L.orc_mg_create.argtypes = [C.c_int, C.c_int]
    L.orc_mg_set_level.restype = C.c_int
    L.orc_mg_set_level.argtypes = [C.c_void_p, C.c_int, C.c_int, ip, ip, dp,
                                   C.c_int, ip, ip, dp, C.c_double,
                                   C.c_double, C.c_int]
    L.orc_mg_free.argtypes = [C.c_void_p]
    L.orc_mg_set_cycle.argtypes = [C.c_void_p, C.c_int]
    L.orc_mg_vcycle.argtypes = [C.c_void_p, C.c_int, dp, C.c_int, dp, C.c_int]
    L.orc_zs_diag.restype = C.c_void_p
    L.orc_zs_diag.argtypes = [C.c_int, dp]
    L.orc_zs_mat.restype = C.c_void_p
    L.orc_zs_mat.argtypes = [C.c_int, ip, ip, dp, C.c_void_p, C.c_int,
                             C.c_double, C.c_double, C.c_int, C.c_double]
    L.orc_zs_free.argtypes = [C.c_void_p]
    L.orc_zs_apply.argtypes = [C.c_void_p, C.c_int, dp, C.c_int, dp, C.c_int]
    L.orc_zs_solve_vec.argtypes = [C.c_void_p, dp, dp]
    L.orc_zs_solve_block.argtypes = [C.c_void_p, C.c_int, dp, C.c_int, dp,
                                     C.c_int]
    L.orc_cheb_jacobi.argtypes = [C.c_int, ip, ip, dp, dp, C.c_double,
                                  C.c_double, C.c_int, C.c_int, dp, C.c_int,
                                  dp, C.c_int]
    L.orc_cheb_mg.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int,
                              C.c_int, dp, C.c_int, dp, C.c_int]
    L.orc_cheb_count.restype = C.c_int
    L.orc_cheb_count.argtypes = [C.c_double, C.c_double, C.c_double]
    L.orc_cg.restype = C.c_int
    L.orc_cg.argtypes = [C.c_int, ip, ip, dp, C.c_int, C.c_void_p, dp, dp,
                         C.c_double, C.c_int, dp, dp]
    L.orc_potrf_l.restype = C.c_int
    L.orc_potrf_l.argtypes = [C.c_int, dp, C.c_int]
    L.orc_potrf_u.restype = C.c_int
    L.orc_potrf_u.argtypes = [C.c_int, dp, C.c_int]
    L.orc_trsm_rltn.argtypes = [C.c_int, C.c_int, dp, C.c_int, dp, C.c_int]
    L.orc_trsm_runn.argtypes = [C.c_int, C.c_int, dp, C.c_int, dp, C.c_int]
    L.orc_syev.restype = C.c_int
    L.orc_syev.argtypes = [C.c_int, dp, C.c_int, dp]
    L.orc_glr_build.restype = C.c_void_p
    L.orc_glr_build.argtypes = [C.c_int, ip, ip, dp, dp, C.c_void_p,
                                C.POINTER(GlrOpts), C.POINTER(GlrReport)]
    L.orc_glr_extend.restype = C.c_int
    L.orc_glr_extend.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_int,
                                 C.POINTER(GlrReport)]
    L.orc_glr_free.argtypes = [C.c_void_p]
    L.orc_glr_kept.restype = C.c_int
    L.orc_glr_kept.argtypes = [C.c_void_p]
    L.orc_glr_lam.restype = dp
    L.orc_glr_lam.argtypes = [C.c_void_p]
    L.orc_glr_U.restype = dp
    L.orc_glr_U.argtypes = [C.c_void_p]
    L.orc_glr_apply_F.argtypes = [C.c_void_p, C.c_int, dp, C.c_int, dp,
                                  C.c_int]
    for nm in ("orc_glr_solve", "orc_glr_apply", "orc_glr_sample"):
        getattr(L, nm).restype = C.c_int
        getattr(L, nm).argtypes = [C.c_void_p, C.c_double, dp, dp]
    L.orc_glr_factor.restype = C.c_int
    L.orc_glr_factor.argtypes = [C.c_void_p, C.c_double, C.c_int, dp, dp]
    L.orc_glr_pow.restype = C.c_int
    L.orc_glr_pow.argtypes = [C.c_void_p, C.c_double, C.c_double, dp, dp]
    L.orc_glr_filter.restype = C.c_int
    L.orc_glr_filter.argtypes = [C.c_void_p, dp, C.c_double, dp, dp]
    L.orc_glr_logdet.restype = C.c_double
    L.orc_glr_logdet.argtypes = [C.c_void_p, C.c_double]


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _fcol(a):
    """column-major float64 copy"""
    return np.asfortranarray(a, dtype=np.float64)


class CSR:
    """Host CSR with int32 indices, float64 values (kept alive here)."""

    def __init__(self, n, rp, ci, v, m=None):
        self.n = int(n)
        self.m = int(m if m is not None else n)
        self.rp = np.ascontiguousarray(rp, dtype=np.int32)
        self.ci = np.ascontiguousarray(ci, dtype=np.int32)
        self.v = np.ascontiguousarray(v, dtype=np.float64)

    @classmethod
    def from_scipy(cls, A):
        A = A.tocsr()
        return cls(A.shape[0], A.indptr, A.indices, A.data, A.shape[1])

    @classmethod
    def from_dense(cls, D):
        import scipy.sparse as sp
        return cls.from_scipy(sp.csr_matrix(np.asarray(D)))

    def args(self):
        return (self.n, _i(self.rp), _i(self.ci), _d(self.v))


def set_threads(nt):
    """host threads for the column-parallel sweeps (results unchanged)"""
    lib().orc_set_threads(int(nt))


def randn_at(seed, i, j):
    return lib().orc_randn_at(seed, i, j)


def randn_fill(seed, row0, col0, nr, nc):
    out = np.zeros((nr, nc), dtype=np.float64, order="F")
    lib().orc_randn_fill(seed, row0, col0, nr, nc, _d(out), max(nr, 1))
    return out


def csr_mm(A, X):
    X = _fcol(X.reshape(X.shape[0], -1))
    Y = np.zeros((A.n, X.shape[1]), order="F")
    lib().orc_csr_mm(*A.args(), X.shape[1], _d(X), X.shape[0], _d(Y), A.n)
    return Y


class MG:
    """Multigrid hierarchy for the oracle cycle.

    levels: list finest-last of dicts {A: CSR, P: CSR or None, slo, shi, nu}
    (PCMG numbering: levels[0] is the coarsest, solved by dense LU);
    gamma: 1 V-cycle, 2 W-cycle."""

    def __init__(self, levels, tile=64, gamma=1):
        L = lib()
        self.h = L.orc_mg_create(len(levels), tile)
        L.orc_mg_set_cycle(self.h, gamma)
        self._keep = []
        for l, lv in enumerate(levels):
            A = lv["A"]
            P = lv.get("P")
            if P is not None:
                rc = L.orc_mg_set_level(self.h, l, A.n, _i(A.rp), _i(A.ci),
                                        _d(A.v), P.m, _i(P.rp), _i(P.ci),
                                        _d(P.v), lv["slo"], lv["shi"],
                                        lv["nu"])
            else:
                rc = L.orc_mg_set_level(self.h, l, A.n, _i(A.rp), _i(A.ci),
                                        _d(A.v), 0, None, None, None, 0., 0.,
                                        0)
            if rc != 0:
                raise RuntimeError("coarse LU failed")
        self.n = levels[-1]["A"].n

    def vcycle(self, B):
        B = _fcol(B.reshape(B.shape[0], -1))
        X = np.zeros_like(B, order="F")
        lib().orc_mg_vcycle(self.h, B.shape[1], _d(B), B.shape[0], _d(X),
                            B.shape[0])
        return X

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_mg_free(self.h)
            self.h = None


def cheb_jacobi(A, dinv, lmin, lmax, nit, B):
    B = _fcol(B.reshape(B.shape[0], -1))
    X = np.zeros_like(B, order="F")
    dinv = _f64(dinv)
    lib().orc_cheb_jacobi(*A.args(), _d(dinv), lmin, lmax, nit, B.shape[1],
                          _d(B), B.shape[0], _d(X), B.shape[0])
    return X


def cheb_mg(mg, lmin, lmax, nit, B):
    B = _fcol(B.reshape(B.shape[0], -1))
    X = np.zeros_like(B, order="F")
    lib().orc_cheb_mg(mg.h, lmin, lmax, nit, B.shape[1], _d(B), B.shape[0],
                      _d(X), B.shape[0])
    return X


def cheb_count(lmin, lmax, tol):
    return lib().orc_cheb_count(lmin, lmax, tol)


def cg(A, b, rtol=1e-12, maxit=10000, mg=None):
    b = _f64(b)
    x = np.zeros_like(b)
    emin = C.c_double()
    emax = C.c_double()
    its = lib().orc_cg(*A.args(), 1 if mg is not None else 0,
                       mg.h if mg is not None else None, _d(b), _d(x), rtol,
                       maxit, C.byref(emin), C.byref(emax))
    return x, its, emin.value, emax.value


class ZSolver:
    """Z-solver: diag (exact) or assembled Z with a block mode."""

    def __init__(self, ptr, keep):
        self.p = ptr
        self._keep = keep

    @classmethod
    def diag(cls, zdiag):
        zdiag = _f64(zdiag)
        return cls(lib().orc_zs_diag(len(zdiag), _d(zdiag)), [zdiag])

    @classmethod
    def mat(cls, A, mode=0, mg=None, lmin=0., lmax=0., nit=0, ksp_rtol=1e-12):
        p = lib().orc_zs_mat(*A.args(), mg.h if mg is not None else None,
                             mode, lmin, lmax, nit, ksp_rtol)
        return cls(p, [A, mg])

    def apply(self, X):
        X = _fcol(X.reshape(X.shape[0], -1))
        Y = np.zeros_like(X, order="F")
        lib().orc_zs_apply(self.p, X.shape[1], _d(X), X.shape[0], _d(Y),
                           X.shape[0])
        return Y

    def solve_vec(self, b):
        b = _f64(b)
        x = np.zeros_like(b)
        lib().orc_zs_solve_vec(self.p, _d(b), _d(x))
        return x

    def solve_block(self, X):
        X = _fcol(X.reshape(X.shape[0], -1))
        Y = np.zeros_like(X, order="F")
        lib().orc_zs_solve_block(self.p, X.shape[1], _d(X), X.shape[0],
                                 _d(Y), X.shape[0])
        return Y

    def __del__(self):
        if getattr(self, "p", None) and _lib is not None:
            _lib.orc_zs_free(self.p)
            self.p = None


def potrf_l(G):
    G = _fcol(G).copy(order="F")
    info = lib().orc_potrf_l(G.shape[0], _d(G), G.shape[0])
    return np.tril(G), info


def syev(A):
    A = _fcol(A).copy(order="F")
    w = np.zeros(A.shape[0])
    info = lib().orc_syev(A.shape[0], _d(A), A.shape[0], _d(w))
    return w, A, info


def trsm_rltn(L, X):
    X = _fcol(X).copy(order="F")
    L = _fcol(L)
    lib().orc_trsm_rltn(X.shape[0], X.shape[1], _d(L), L.shape[0], _d(X),
                        X.shape[0])
    return X


class GLR:
    """Oracle GLR object over (B CSR, mass, ZSolver)."""

    def __init__(self, B, mass, zs, ell=200, q_power=1, trunc_abs=0.0,
                 trunc_rel=1e-3, treat=0, seed=271828, check=1, backend=1):
        self.B = B
        self.mass = _f64(mass)
        self.zs = zs
        o = GlrOpts(ell, q_power, trunc_abs, trunc_rel, treat, seed, check,
                    backend)
        rep = GlrReport()
        self.g = lib().orc_glr_build(*B.args(), _d(self.mass), zs.p,
                                     C.byref(o), C.byref(rep))
        if not self.g:
            raise RuntimeError("oracle GLR build failed")
        self.report = rep.as_dict()
        self.n = B.n

    def extend(self, k_new, incremental=True, panel=64):
        rep = GlrReport()
        rc = lib().orc_glr_extend(self.g, k_new, 1 if incremental else 0,
                                  panel, C.byref(rep))
        if rc != 0:
            raise RuntimeError("oracle extend failed %d" % rc)
        self.report = rep.as_dict()
        return self.report

    @property
    def kept(self):
        return lib().orc_glr_kept(self.g)

    @property
    def lam(self):
        k = self.kept
        return np.ctypeslib.as_array(lib().orc_glr_lam(self.g),
                                     shape=(max(k, 1),))[:k].copy()

    @property
    def U(self):
        k = self.kept
        a = np.ctypeslib.as_array(lib().orc_glr_U(self.g),
                                  shape=(max(k, 1) * self.n,)).copy()
        return a.reshape((max(k, 1), self.n)).T[:, :k]

    def apply_F(self, X):
        X = _fcol(X.reshape(X.shape[0], -1))
        Y = np.zeros_like(X, order="F")
        lib().orc_glr_apply_F(self.g, X.shape[1], _d(X), X.shape[0], _d(Y),
                              X.shape[0])
        return Y

    def _vec(self, fn, *args):
        *scal, v = args
        v = _f64(v)
        out = np.zeros_like(v)
        rc = fn(self.g, *scal, _d(v), _d(out))
        if rc != 0:
            raise ValueError("oracle op refused (rc=%d)" % rc)
        return out

    def solve(self, c, rhs):
        return self._vec(lib().orc_glr_solve, c, rhs)

    def apply(self, c, v):
        return self._vec(lib().orc_glr_apply, c, v)

    def sample(self, c, xi):
        return self._vec(lib().orc_glr_sample, c, xi)

    def factor(self, c, which, v):
        return self._vec(lib().orc_glr_factor, c, which, v)

    def pow(self, c, p, v):
        return self._vec(lib().orc_glr_pow, c, p, v)

    def filter(self, f_lam, f_c, v):
        f_lam = _f64(f_lam)
        v = _f64(v)
        out = np.zeros_like(v)
        lib().orc_glr_filter(self.g, _d(f_lam), f_c, _d(v), _d(out))
        return out

    def logdet(self, c):
        return lib().orc_glr_logdet(self.g, c)

    def __del__(self):
        if getattr(self, "g", None) and _lib is not None:
            _lib.orc_glr_free(self.g)
            self.g = None
