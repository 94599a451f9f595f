"""lgpsf-hessian_amd — MI355X engine for lgpsf-hessian's GLR stage.

Python bindings (ctypes) of liblgh.so: the reference's C API
(lgh_prior_* / lgh_glr_*, include/lgpsf_hessian/*.h) over the PETSc-subset
Mat/Vec types, and the thin lgh_hip_* kernel ABI (include/lgh_hip.h).
The compute path is the HIP library; if it is missing or no GPU is
visible the calls fail with an error instead of falling back.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblgh.so")
_lib = None

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p

PREC_F64 = 0
PREC_F32 = 1
PREC_BF16 = 2          # V-cycle storage only (fp32 arithmetic)


class LGHError(RuntimeError):
    pass


def build(jobs=8):
    """Compile liblgh.so for gfx950 (always recompiles what changed)."""
    subprocess.run(["make", "-s", "-C", _HERE, "-j%d" % jobs], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LGHError("liblgh.so is not built (run build())")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


# ---- C structs ----------------------------------------------------------------

class GlrOpts(C.Structure):
    _fields_ = [("ell", C.c_int), ("q_power", C.c_int),
                ("trunc_abs", C.c_double), ("trunc_rel", C.c_double),
                ("treat", C.c_int), ("seed", C.c_ulong), ("backend", C.c_int),
                ("nb", C.c_int), ("panel", C.c_int), ("grid2d", C.c_int),
                ("check", C.c_int)]


class GlrReport(C.Structure):
    _fields_ = [("kept", C.c_int), ("lam_abs_max", C.c_double),
                ("lam_raw_min", C.c_double), ("lam_raw_max", C.c_double),
                ("next_abs", C.c_double), ("n_negative_raw", C.c_int),
                ("resid_max", C.c_double), ("t_operator", C.c_double),
                ("t_dense", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GlrXOpts(C.Structure):
    _fields_ = [("prec_sketch", C.c_int), ("prec_precond", C.c_int),
                ("prec_check", C.c_int), ("prec_sketch_precond", C.c_int),
                ("solve_blocked_min", C.c_int), ("sweep_panels", C.c_int)]


class ServeOpts(C.Structure):
    _fields_ = [("max_batch_cols", C.c_int), ("device_budget_gb", C.c_double),
                ("queue_cols", C.c_int), ("device_streams", C.c_int)]


class ServeStats(C.Structure):
    _fields_ = [("batches", C.c_int64), ("columns", C.c_int64),
                ("requests", C.c_int64), ("max_batch_cols", C.c_int),
                ("batch_cols", C.c_int), ("busy_s", C.c_double),
                ("stage_s", C.c_double), ("max_running", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class PriorMatOpts(C.Structure):
    _fields_ = [("ksp_rtol", C.c_double), ("blocked_mode", C.c_int),
                ("tile", C.c_int), ("cheb_rtol", C.c_double),
                ("smax_krylov", C.c_int), ("smoother_top", C.c_double),
                ("nu_force", C.c_int), ("verbose", C.c_int)]


class PriorInfo(C.Structure):
    _fields_ = [("is_mat", C.c_int), ("blocked_mode", C.c_int),
                ("nit", C.c_int), ("lmin", C.c_double), ("lmax", C.c_double),
                ("levels", C.c_int), ("n_coarse", C.c_int),
                ("bound_its", C.c_int), ("cycle", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


VECFN = C.CFUNCTYPE(None, vp, vp, vp)
MATFN = C.CFUNCTYPE(None, vp, vp, vp)


class Callbacks(C.Structure):
    _fields_ = [("applyZ", VECFN), ("applyZt", VECFN), ("solveZ", VECFN),
                ("solveZt", VECFN), ("solveZ_blocked", MATFN),
                ("solveZt_blocked", MATFN), ("ctx", vp)]


HESSFN = C.CFUNCTYPE(None, dp, dp, vp)     # lgh_hessian_fn


class CorrectOpts(C.Structure):
    _fields_ = [("c0", C.c_double), ("rank", C.c_int), ("applies", C.c_int),
                ("n_qc", C.c_int), ("q_power", C.c_int),
                ("clamp_eps", C.c_double), ("d_keep_min", C.c_double),
                ("drop_tol", C.c_double), ("seed", C.c_ulong),
                ("verbose", C.c_int)]


class CorrectReport(C.Structure):
    _fields_ = [("kept", C.c_int), ("basis", C.c_int), ("clamped", C.c_int),
                ("hd_applies", C.c_int), ("d_min", C.c_double),
                ("d_max", C.c_double), ("floor", C.c_double),
                ("qc_before", C.c_double), ("qc_after", C.c_double),
                ("t_operator", C.c_double), ("t_dense", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


# every exported function: name -> (restype, argtypes)
_SIGS = {
    # PETSc subset
    "lgh_petsc_last_error": (C.c_char_p, []),
    "VecCreateSeq": (C.c_int, [C.c_int, C.c_int, C.POINTER(vp)]),
    "VecDestroy": (C.c_int, [C.POINTER(vp)]),
    "VecGetArray": (C.c_int, [vp, C.POINTER(dp)]),
    "VecRestoreArray": (C.c_int, [vp, C.POINTER(dp)]),
    "VecGetLocalSize": (C.c_int, [vp, ip]),
    "MatCreateSeqAIJWithArrays": (C.c_int, [C.c_int, C.c_int, C.c_int, ip, ip,
                                            dp, C.POINTER(vp)]),
    "MatCreateSeqDense": (C.c_int, [C.c_int, C.c_int, C.c_int, dp,
                                    C.POINTER(vp)]),
    "MatDestroy": (C.c_int, [C.POINTER(vp)]),
    "MatDenseGetArray": (C.c_int, [vp, C.POINTER(dp)]),
    "MatDenseRestoreArray": (C.c_int, [vp, C.POINTER(dp)]),
    "MatGetSize": (C.c_int, [vp, ip, ip]),
    # reference API
    "lgh_prior_mat_opts_default": (PriorMatOpts, []),
    "lgh_prior_create_mat": (C.c_int, [vp, vp, C.POINTER(PriorMatOpts),
                                       C.POINTER(vp)]),
    "lgh_prior_create_callbacks": (C.c_int, [vp, C.POINTER(Callbacks),
                                             C.POINTER(vp)]),
    "lgh_prior_create_ksp": (C.c_int, [vp, vp, C.POINTER(PriorMatOpts),
                                       C.POINTER(vp)]),
    "lgh_prior_destroy": (None, [vp]),
    "lgh_prior_apply": (C.c_int, [vp, vp, vp]),
    "lgh_glr_opts_default": (GlrOpts, []),
    "lgh_glr_compute": (C.c_int, [vp, vp, C.POINTER(GlrOpts), C.POINTER(vp),
                                  C.POINTER(GlrReport)]),
    "lgh_glr_extend": (C.c_int, [vp, C.c_int, C.POINTER(GlrReport)]),
    "lgh_glr_destroy": (None, [vp]),
    "lgh_glr_solve": (C.c_int, [vp, C.c_double, vp, vp]),
    "lgh_glr_apply": (C.c_int, [vp, C.c_double, vp, vp]),
    "lgh_glr_sample": (C.c_int, [vp, C.c_double, vp, vp]),
    "lgh_glr_factor": (C.c_int, [vp, C.c_double, C.c_int, vp, vp]),
    "lgh_glr_logdet": (C.c_double, [vp, C.c_double]),
    "lgh_glr_pow": (C.c_int, [vp, C.c_double, C.c_double, vp, vp]),
    "lgh_glr_filter": (C.c_int, [vp, dp, C.c_double, vp, vp]),
    "lgh_glr_eigs": (C.c_int, [vp, ip, C.POINTER(dp)]),
    "lgh_glr_correct_opts_default": (CorrectOpts, []),
    "lgh_glr_correct_probes": (C.c_int, [vp, C.c_int, dp, dp, HESSFN, vp,
                                         C.POINTER(CorrectOpts),
                                         C.POINTER(CorrectReport)]),
    "lgh_glr_correct": (C.c_int, [vp, HESSFN, vp, C.POINTER(CorrectOpts),
                                  C.POINTER(CorrectReport)]),
    # extensions
    "lgh_glr_xopts_default": (GlrXOpts, []),
    "lgh_glr_xopts_mixed": (GlrXOpts, []),
    "lgh_glr_compute_x": (C.c_int, [vp, vp, C.POINTER(GlrOpts),
                                    C.POINTER(GlrXOpts), C.POINTER(vp),
                                    C.POINTER(GlrReport)]),
    "lgh_glr_compute_batch": (C.c_int, [C.c_int, C.POINTER(vp), C.POINTER(vp),
                                        C.POINTER(GlrOpts), C.POINTER(GlrXOpts),
                                        C.c_int, C.POINTER(vp),
                                        C.POINTER(GlrReport)]),
    "lgh_prior_create_mat_batch": (C.c_int, [C.c_int, C.POINTER(vp), C.POINTER(vp),
                                             C.POINTER(PriorMatOpts), C.c_int,
                                             C.POINTER(vp)]),
    "lgh_glr_solve_block": (C.c_int, [vp, C.c_int, dp, dp, C.c_int, dp,
                                      C.c_int]),
    "lgh_glr_sample_block": (C.c_int, [vp, C.c_int, dp, dp, C.c_int, dp,
                                       C.c_int]),
    "lgh_glr_solve_block_dev": (C.c_int, [vp, C.c_int, dp, vp, C.c_int, vp,
                                          C.c_int]),
    "lgh_glr_sample_block_dev": (C.c_int, [vp, C.c_int, dp, vp, C.c_int, vp,
                                           C.c_int]),
    "lgh_prior_get_info": (C.c_int, [vp, C.POINTER(PriorInfo)]),
    "lgh_prior_solve_host": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp,
                                       C.c_int, dp, C.c_int]),
    "lgh_glr_get_U": (C.c_int, [vp, dp, C.c_int]),
    "lgh_mat_from_fit_rows": (C.c_int, [C.c_int, ip, C.POINTER(C.c_long), dp,
                                        C.c_double, C.POINTER(vp)]),
    "MatSeqAIJGetCSR": (C.c_int, [vp, C.POINTER(ip), C.POINTER(ip), C.POINTER(dp)]),
    "lgh_serve_opts_default": (ServeOpts, []),
    "lgh_serve_create": (C.c_int, [vp, C.POINTER(ServeOpts), C.POINTER(vp)]),
    "lgh_serve_submit": (C.c_int, [vp, C.c_int, C.c_double, C.c_int, dp, C.c_int,
                                   dp, C.c_int, C.POINTER(C.c_int64)]),
    "lgh_serve_wait": (C.c_int, [vp, C.c_int64, dp, dp]),
    "lgh_serve_get_stats": (C.c_int, [vp, C.POINTER(ServeStats)]),
    "lgh_serve_destroy": (None, [vp]),
    # kernel ABI
    "lgh_hip_init": (C.c_int, []),
    "lgh_hip_last_error": (C.c_char_p, []),
    "lgh_hip_sync": (C.c_int, []),
    "lgh_hip_bytes_in_use": (C.c_int64, []),
    "lgh_hip_trim": (None, []),
    "lgh_hip_prof_enable": (None, [C.c_int]),
    "lgh_hip_prof_collect": (C.c_int, [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int64), dp, dp, dp]),
    "lgh_hip_copy_gbps": (C.c_int, [C.c_int64, C.c_int, dp]),
    "lgh_hip_dev_alloc": (vp, [C.c_int64]),
    "lgh_hip_dev_free": (None, [vp]),
    "lgh_hip_dev_put": (C.c_int, [vp, vp, C.c_int64]),
    "lgh_hip_dev_get": (C.c_int, [vp, vp, C.c_int64]),
    "lgh_hip_randn_at_host": (C.c_double, [C.c_uint64, C.c_int64, C.c_int64]),
    "lgh_hip_randn_host_block": (C.c_int, [C.c_uint64, C.c_int64, C.c_int64, C.c_int,
                                           C.c_int, C.c_int, dp, C.c_int,
                                           C.POINTER(C.c_ubyte), C.c_int]),
    "lgh_hip_randn_fill": (C.c_int, [C.c_uint64, C.c_int64, C.c_int64, C.c_int,
                                     C.c_int, C.c_int, C.c_int, dp, C.c_int]),
    "lgh_hip_spmm": (C.c_int, [C.c_int, C.c_int, ip, ip, dp, C.c_int, C.c_int,
                               C.c_int, C.c_int, dp, C.c_int, dp, C.c_int]),
    "lgh_hip_spmm_grouped": (C.c_int, [C.c_int, C.c_int, ip, ip, dp, C.c_int,
                                       C.c_int, C.c_int, C.c_int, dp, C.c_int,
                                       dp, C.c_int]),
    "lgh_hip_randn_fill_noalloc": (C.c_int, [C.c_uint64, C.c_int64, C.c_int64, C.c_int,
                                             C.c_int, C.c_int, C.c_void_p, C.c_int]),
    "lgh_hip_spmm_group_config": (None, [C.c_int, C.c_int]),
    "lgh_hip_spmm_tile_config": (None, [C.c_int]),
    "lgh_hip_mg_coarse_config": (None, [C.c_int, C.c_int]),
    "lgh_hip_spmm_group_report": (C.c_int, [C.c_char_p, C.c_int]),
    "lgh_hip_cheb_jacobi": (C.c_int, [C.c_int, ip, ip, dp, C.c_int, C.c_int,
                                      C.c_double, C.c_double, C.c_int, C.c_int,
                                      dp, C.c_int, dp, C.c_int]),
    "lgh_hip_gram": (C.c_int, [C.c_int, C.c_int, C.c_int, dp, C.c_int, C.c_int,
                               dp, C.c_int, dp, C.c_int]),
    "lgh_hip_gemm_nn": (C.c_int, [C.c_int, C.c_int, C.c_int, dp, C.c_int,
                                  C.c_int, dp, C.c_int, dp, C.c_int]),
    "lgh_amg_build": (C.c_int, [C.c_int, ip, ip, dp, C.c_int, C.c_double,
                                C.c_int, C.POINTER(vp)]),
    "lgh_amg_free": (None, [vp]),
    "lgh_amg_levels": (C.c_int, [vp]),
    "lgh_amg_level": (C.c_int, [vp, C.c_int, ip, ip, C.POINTER(ip),
                                C.POINTER(ip), C.POINTER(dp), C.POINTER(ip),
                                C.POINTER(ip), C.POINTER(dp), dp, dp, ip]),
    "lgh_amg_vcycle": (C.c_int, [vp, C.c_int, C.c_int, C.c_int, dp, C.c_int,
                                 dp, C.c_int]),
    "lgh_amg_set_cycle": (C.c_int, [vp, C.c_int]),
}


def _declare(L):
    for name, (res, args) in _SIGS.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args


def _d(a):
    return a.ctypes.data_as(dp)


def _i(a):
    return a.ctypes.data_as(ip)


def _hip_check(rc, what):
    if rc != 0:
        msg = lib().lgh_hip_last_error().decode()
        raise LGHError("%s failed (%d): %s" % (what, rc, msg))


def _petsc_check(rc, what):
    if rc != 0:
        msg = lib().lgh_petsc_last_error().decode()
        raise LGHError("%s failed (PETSc code %d): %s" % (what, rc, msg))


def _fcol(a):
    a = np.asarray(a, dtype=np.float64)
    if a.ndim == 1:
        a = a[:, None]
    return np.asfortranarray(a)


# ---- PETSc-subset objects -------------------------------------------------------

class Vec:
    def __init__(self, n_or_array):
        L = lib()
        h = vp()
        if np.isscalar(n_or_array):
            n = int(n_or_array)
            _petsc_check(L.VecCreateSeq(0, n, C.byref(h)), "VecCreateSeq")
            self.h, self.n = h, n
        else:
            a = np.asarray(n_or_array, dtype=np.float64).ravel()
            _petsc_check(L.VecCreateSeq(0, len(a), C.byref(h)), "VecCreateSeq")
            self.h, self.n = h, len(a)
            self.set(a)

    def array(self):
        p = dp()
        lib().VecGetArray(self.h, C.byref(p))
        return np.ctypeslib.as_array(p, shape=(self.n,))

    def set(self, a):
        self.array()[:] = a

    def get(self):
        return self.array().copy()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.VecDestroy(C.byref(self.h))
            self.h = None


class Mat:
    def __init__(self, handle, keep=()):
        self.h = handle
        self._keep = keep

    @classmethod
    def aij(cls, A):
        """AIJ Mat from a scipy sparse matrix (values copied)."""
        import scipy.sparse as sp
        A = sp.csr_matrix(A)
        A.sort_indices()
        rp = np.ascontiguousarray(A.indptr, dtype=np.int32)
        ci = np.ascontiguousarray(A.indices, dtype=np.int32)
        v = np.ascontiguousarray(A.data, dtype=np.float64)
        h = vp()
        _petsc_check(lib().MatCreateSeqAIJWithArrays(0, A.shape[0], A.shape[1],
                                                     _i(rp), _i(ci), _d(v),
                                                     C.byref(h)),
                     "MatCreateSeqAIJWithArrays")
        return cls(h)

    @classmethod
    def from_fit_rows(cls, rowptr, colgids, vals, sym_tol=-1.0):
        """AIJ Mat from the fit stage's row export (lgh_mat_from_fit_rows):
        row pointers, 64-bit global column ids, values"""
        rp = np.ascontiguousarray(rowptr, dtype=np.int32)
        cg = np.ascontiguousarray(colgids, dtype=np.int64)
        v = np.ascontiguousarray(vals, dtype=np.float64)
        h = vp()
        _petsc_check(lib().lgh_mat_from_fit_rows(
            len(rp) - 1, _i(rp), cg.ctypes.data_as(C.POINTER(C.c_long)), _d(v),
            sym_tol, C.byref(h)), "lgh_mat_from_fit_rows")
        return cls(h)

    def csr(self):
        """scipy CSR copy of an assembled AIJ Mat"""
        import scipy.sparse as sp
        i, j, a = ip(), ip(), dp()
        _petsc_check(lib().MatSeqAIJGetCSR(self.h, C.byref(i), C.byref(j), C.byref(a)),
                     "MatSeqAIJGetCSR")
        m, n = C.c_int(), C.c_int()
        _petsc_check(lib().MatGetSize(self.h, C.byref(m), C.byref(n)), "MatGetSize")
        rp = np.ctypeslib.as_array(i, shape=(m.value + 1,)).copy()
        nnz = int(rp[-1])
        ci = np.ctypeslib.as_array(j, shape=(max(nnz, 1),))[:nnz].copy()
        v = np.ctypeslib.as_array(a, shape=(max(nnz, 1),))[:nnz].copy()
        return sp.csr_matrix((v, ci, rp), shape=(m.value, n.value))

    @classmethod
    def dense(cls, D):
        D = np.asfortranarray(D, dtype=np.float64)
        h = vp()
        _petsc_check(lib().MatCreateSeqDense(0, D.shape[0], D.shape[1], _d(D),
                                             C.byref(h)), "MatCreateSeqDense")
        return cls(h, keep=(D,))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.MatDestroy(C.byref(self.h))
            self.h = None


# ---- reference API ----------------------------------------------------------------

def prior_mat_opts(**kw):
    o = lib().lgh_prior_mat_opts_default()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def glr_opts(**kw):
    o = lib().lgh_glr_opts_default()
    for k, v in kw.items():
        setattr(o, k, v)
    return o


def xopts(mixed=False, **kw):
    x = lib().lgh_glr_xopts_mixed() if mixed else lib().lgh_glr_xopts_default()
    for k, v in kw.items():
        setattr(x, k, v)
    return x


class Prior:
    def __init__(self, h, keep=()):
        self.h = h
        self._keep = keep

    @classmethod
    def from_mat(cls, Z, mass, **opts):
        Zm = Z if isinstance(Z, Mat) else Mat.aij(Z)
        mv = mass if isinstance(mass, Vec) else Vec(mass)
        o = prior_mat_opts(**opts)
        h = vp()
        _petsc_check(lib().lgh_prior_create_mat(Zm.h, mv.h, C.byref(o),
                                                C.byref(h)),
                     "lgh_prior_create_mat")
        return cls(h, keep=(Zm, mv))

    @classmethod
    def batch(cls, pairs, workers=0, **opts):
        """priors of many (Z, mass) pairs at once (lgh_prior_create_mat_batch)"""
        mats = [(Z if isinstance(Z, Mat) else Mat.aij(Z),
                 m if isinstance(m, Vec) else Vec(m)) for Z, m in pairs]
        k = len(mats)
        Zs = (vp * k)(*[z.h for z, _ in mats])
        Ms = (vp * k)(*[m.h for _, m in mats])
        os_ = (PriorMatOpts * k)(*[prior_mat_opts(**opts) for _ in range(k)])
        hs = (vp * k)()
        _petsc_check(lib().lgh_prior_create_mat_batch(k, Zs, Ms, os_, workers, hs),
                     "lgh_prior_create_mat_batch")
        return [cls(vp(hs[i]), keep=mats[i]) for i in range(k)]

    @classmethod
    def diagonal_callbacks(cls, zdiag, mass, blocked=False, calls=None):
        """Prior with Z = diag(zdiag) supplied through the Vec callbacks
        (the exact "fake solver" prior of tests/test_glr.cpp:67-85); with
        blocked=True also through the MATDENSE block callbacks
        solveZ_blocked / solveZt_blocked (prior.h:53-54).  calls (a dict)
        counts the invocations of each callback."""
        L = lib()
        zd = np.asarray(zdiag, dtype=np.float64).copy()
        n = len(zd)
        cnt = calls if calls is not None else {}

        def arr(h):
            p = dp()
            L.VecGetArray(h, C.byref(p))
            return np.ctypeslib.as_array(p, shape=(n,))

        def marr(h):
            rows, cols = C.c_int(), C.c_int()
            L.MatGetSize(h, C.byref(rows), C.byref(cols))
            p = dp()
            L.MatDenseGetArray(h, C.byref(p))
            return np.ctypeslib.as_array(p, shape=(cols.value, rows.value))

        def applyZ(x, y, ctx):
            cnt["applyZ"] = cnt.get("applyZ", 0) + 1
            arr(y)[:] = arr(x) * zd

        def solveZ(x, y, ctx):
            cnt["solveZ"] = cnt.get("solveZ", 0) + 1
            arr(y)[:] = arr(x) / zd

        def solveZ_blocked(X, Y, ctx):
            cnt["solveZ_blocked"] = cnt.get("solveZ_blocked", 0) + 1
            marr(Y)[:] = marr(X) / zd[None, :]

        def solveZt_blocked(X, Y, ctx):
            cnt["solveZt_blocked"] = cnt.get("solveZt_blocked", 0) + 1
            marr(Y)[:] = marr(X) / zd[None, :]

        cbs = (VECFN(applyZ), VECFN(solveZ))
        mfs = (MATFN(solveZ_blocked), MATFN(solveZt_blocked)) if blocked \
            else (MATFN(), MATFN())
        cbs = cbs + mfs
        cb = Callbacks(cbs[0], VECFN(), cbs[1], VECFN(), mfs[0], mfs[1], None)
        mv = mass if isinstance(mass, Vec) else Vec(mass)
        h = vp()
        _petsc_check(L.lgh_prior_create_callbacks(mv.h, C.byref(cb), C.byref(h)),
                     "lgh_prior_create_callbacks")
        return cls(h, keep=(mv, cbs, cb, zd))

    def info(self):
        i = PriorInfo()
        lib().lgh_prior_get_info(self.h, C.byref(i))
        return i.as_dict()

    def solve(self, X, which=0, prec=PREC_F64):
        X = _fcol(X)
        Y = np.zeros(X.shape, order="F")   # calloc: pages fault in on the copy-out threads
        _petsc_check(lib().lgh_prior_solve_host(self.h, which, prec, X.shape[1],
                                                _d(X), X.shape[0], _d(Y),
                                                X.shape[0]),
                     "lgh_prior_solve_host")
        return Y

    def apply(self, x):
        xv, yv = Vec(x), Vec(len(x))
        _petsc_check(lib().lgh_prior_apply(self.h, xv.h, yv.h), "lgh_prior_apply")
        return yv.get()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.lgh_prior_destroy(self.h)
            self.h = None


class DeviceBlock:
    """An n x m FP64 column-major block in device memory (leading dimension
    ld >= n), allocated with the engine's own HIP runtime
    (lgh_hip_dev_alloc): the input and output of GLR.solve_block_device /
    sample_block_device.  columns(j0, j1) is a view of columns j0..j1-1 that
    keeps its parent alive."""

    def __init__(self, n, m, ld=None, _parent=None, _ptr=None):
        self.n, self.m = int(n), int(m)
        self.ld = self.n if ld is None else int(ld)
        if self.ld < self.n:
            raise ValueError("DeviceBlock: ld must be >= n")
        self._parent = _parent
        if _ptr is not None:
            self.ptr = _ptr
        else:
            p = lib().lgh_hip_dev_alloc(8 * self.ld * max(self.m, 1))
            if not p:
                raise LGHError("lgh_hip_dev_alloc failed: %s"
                               % lib().lgh_hip_last_error().decode())
            self.ptr = p

    @classmethod
    def from_host(cls, A, ld=None):
        A = _fcol(A)
        b = cls(A.shape[0], A.shape[1], ld)
        b.put(A)
        return b

    def put(self, A):
        A = _fcol(A)
        if A.shape != (self.n, self.m):
            raise ValueError("DeviceBlock.put: shape mismatch")
        for j0, j1 in self._runs():
            src = np.asfortranarray(A[:, j0:j1]) if self.ld != self.n else A
            _hip_check(lib().lgh_hip_dev_put(self._col(j0), src.ctypes.data,
                                             8 * self.n * (j1 - j0)), "lgh_hip_dev_put")

    def to_host(self):
        A = np.zeros((self.n, self.m), order="F")
        for j0, j1 in self._runs():
            part = np.zeros((self.n, j1 - j0), order="F") if self.ld != self.n else A
            _hip_check(lib().lgh_hip_dev_get(part.ctypes.data, self._col(j0),
                                             8 * self.n * (j1 - j0)), "lgh_hip_dev_get")
            if part is not A:
                A[:, j0:j1] = part
        return A

    def columns(self, j0, j1):
        return DeviceBlock(self.n, j1 - j0, self.ld, _parent=self, _ptr=self._col(j0))

    def _col(self, j):
        return self.ptr + 8 * self.ld * j

    def _runs(self):
        # contiguous blocks move whole; strided ones column by column
        return [(0, self.m)] if self.ld == self.n else [(j, j + 1) for j in range(self.m)]

    def __del__(self):
        if getattr(self, "_parent", 1) is None and getattr(self, "ptr", None) \
                and _lib is not None:
            _lib.lgh_hip_dev_free(self.ptr)
            self.ptr = None


class GLR:
    def __init__(self, B, prior, opts=None, x=None, **kw):
        self.Bm = B if isinstance(B, Mat) else Mat.aij(B)
        self.prior = prior
        o = opts if opts is not None else glr_opts(**kw)
        rep = GlrReport()
        h = vp()
        if x is None:
            rc = lib().lgh_glr_compute(self.Bm.h, prior.h, C.byref(o),
                                       C.byref(h), C.byref(rep))
        else:
            rc = lib().lgh_glr_compute_x(self.Bm.h, prior.h, C.byref(o),
                                         C.byref(x), C.byref(h), C.byref(rep))
        _petsc_check(rc, "lgh_glr_compute")
        self.h = h
        self.report = rep.as_dict()
        n = C.c_int()
        lib().MatGetSize(self.Bm.h, C.byref(n), None)
        self.n = n.value

    @classmethod
    def batch(cls, items, x=None, workers=0):
        """GLRs of many (B, prior, opts) triples at once
        (lgh_glr_compute_batch); x: shared extended options or None"""
        k = len(items)
        Bms = [B if isinstance(B, Mat) else Mat.aij(B) for B, _, _ in items]
        Bs = (vp * k)(*[b.h for b in Bms])
        Ps = (vp * k)(*[p.h for _, p, _ in items])
        os_ = (GlrOpts * k)(*[o for _, _, o in items])
        hs = (vp * k)()
        reps = (GlrReport * k)()
        _petsc_check(lib().lgh_glr_compute_batch(k, Bs, Ps, os_,
                                                 C.byref(x) if x is not None else None,
                                                 workers, hs, reps),
                     "lgh_glr_compute_batch")
        out = []
        for i in range(k):
            g = cls.__new__(cls)
            g.Bm, g.prior, g.h = Bms[i], items[i][1], vp(hs[i])
            g.report = reps[i].as_dict()
            n = C.c_int()
            lib().MatGetSize(g.Bm.h, C.byref(n), None)
            g.n = n.value
            out.append(g)
        return out

    @staticmethod
    def _hess(fn, n):
        """wrap fn(x) -> H_d x (numpy, length n) as an lgh_hessian_fn"""
        def cb(xin, yout, ctx):
            x = np.ctypeslib.as_array(xin, shape=(n,))
            np.ctypeslib.as_array(yout, shape=(n,))[:] = fn(x)
        return HESSFN(cb)

    @staticmethod
    def _copts(kw):
        o = lib().lgh_glr_correct_opts_default()
        for k, v in kw.items():
            setattr(o, k, v)
        return o

    def correct_probes(self, V, HV, hess, **kw):
        """stage-2.5 value pass (lgh_glr_correct_probes); V, HV: n x k"""
        V, HV = _fcol(V), _fcol(HV)
        o, rep, cb = self._copts(kw), CorrectReport(), self._hess(hess, self.n)
        _petsc_check(lib().lgh_glr_correct_probes(self.h, V.shape[1], _d(V), _d(HV),
                                                  cb, None, C.byref(o), C.byref(rep)),
                     "lgh_glr_correct_probes")
        return rep.as_dict()

    def correct(self, hess, **kw):
        """stage-2.5 fresh randomized correction (lgh_glr_correct)"""
        o, rep, cb = self._copts(kw), CorrectReport(), self._hess(hess, self.n)
        _petsc_check(lib().lgh_glr_correct(self.h, cb, None, C.byref(o), C.byref(rep)),
                     "lgh_glr_correct")
        return rep.as_dict()

    def extend(self, k_new):
        rep = GlrReport()
        _petsc_check(lib().lgh_glr_extend(self.h, k_new, C.byref(rep)),
                     "lgh_glr_extend")
        self.report = rep.as_dict()
        return self.report

    @property
    def lam(self):
        k = C.c_int()
        p = dp()
        lib().lgh_glr_eigs(self.h, C.byref(k), C.byref(p))
        if k.value == 0:
            return np.zeros(0)
        return np.ctypeslib.as_array(p, shape=(k.value,)).copy()

    @property
    def kept(self):
        return len(self.lam)

    @property
    def U(self):
        k = self.kept
        U = np.zeros((self.n, k), order="F")
        if k:
            _petsc_check(lib().lgh_glr_get_U(self.h, _d(U), self.n), "lgh_glr_get_U")
        return U

    def _vop(self, fn, name, *scal, v):
        xv, yv = Vec(v), Vec(self.n)
        _petsc_check(fn(self.h, *scal, xv.h, yv.h), name)
        return yv.get()

    def solve(self, c, rhs):
        return self._vop(lib().lgh_glr_solve, "lgh_glr_solve", c, v=rhs)

    def apply(self, c, v):
        return self._vop(lib().lgh_glr_apply, "lgh_glr_apply", c, v=v)

    def sample(self, c, xi):
        return self._vop(lib().lgh_glr_sample, "lgh_glr_sample", c, v=xi)

    def factor(self, c, which, v):
        return self._vop(lib().lgh_glr_factor, "lgh_glr_factor", c, which, v=v)

    def pow(self, c, p, v):
        return self._vop(lib().lgh_glr_pow, "lgh_glr_pow", c, p, v=v)

    def filter(self, f_lam, f_c, v):
        f = np.ascontiguousarray(f_lam, dtype=np.float64)
        return self._vop(lib().lgh_glr_filter, "lgh_glr_filter", _d(f), f_c, v=v)

    def logdet(self, c):
        return lib().lgh_glr_logdet(self.h, c)

    def _block(self, fn, name, c, X):
        X = _fcol(X)
        m = X.shape[1]
        cc = np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64), (m,)))
        Y = np.zeros(X.shape, order="F")   # calloc: pages fault in on the copy-out threads
        _petsc_check(fn(self.h, m, _d(cc), _d(X), X.shape[0], _d(Y), X.shape[0]), name)
        return Y

    def solve_block(self, c, X):
        return self._block(lib().lgh_glr_solve_block, "lgh_glr_solve_block", c, X)

    def sample_block(self, c, X):
        return self._block(lib().lgh_glr_sample_block, "lgh_glr_sample_block", c, X)

    def _block_device(self, fn, name, c, X):
        if not isinstance(X, DeviceBlock) or X.n != self.n:
            raise ValueError("%s: need a DeviceBlock with n = %d rows" % (name, self.n))
        m = X.m
        cc = np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64), (m,)))
        Y = DeviceBlock(self.n, m)
        _petsc_check(fn(self.h, m, _d(cc), X.ptr, X.ld, Y.ptr, Y.ld), name)
        return Y

    def solve_block_device(self, c, X):
        return self._block_device(lib().lgh_glr_solve_block_dev,
                                  "lgh_glr_solve_block_dev", c, X)

    def sample_block_device(self, c, X):
        return self._block_device(lib().lgh_glr_sample_block_dev,
                                  "lgh_glr_sample_block_dev", c, X)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.lgh_glr_destroy(self.h)
            self.h = None


SERVE_SOLVE, SERVE_SAMPLE, SERVE_LOGDET = 0, 1, 2


class Serve:
    """Request loop over a resident GLR (lgh_serve_*).  submit() returns a
    ticket; the input block is kept alive here until wait() returns the
    result (solve / sample: an n x m array; logdet: a float) and the
    submit-to-done latency in seconds.  Submit and wait may be called from
    several Python threads (the ctypes calls release the GIL)."""

    def __init__(self, glr, **kw):
        o = lib().lgh_serve_opts_default()
        for k, v in kw.items():
            setattr(o, k, v)
        h = vp()
        _petsc_check(lib().lgh_serve_create(glr.h, C.byref(o), C.byref(h)),
                     "lgh_serve_create")
        self.h, self.glr, self.n = h, glr, glr.n
        self._live = {}
        import threading
        self._lock = threading.Lock()

    def submit(self, op, c, X=None, out=None):
        """out: optional preallocated n x m Fortran-ordered float64 result"""
        t = C.c_int64()
        if op == SERVE_LOGDET:
            _petsc_check(lib().lgh_serve_submit(self.h, op, c, 0, None, 0, None, 0,
                                                C.byref(t)), "lgh_serve_submit")
            out = None
        else:
            X = _fcol(X)
            if out is None:
                out = np.empty_like(X, order="F")
            if out.shape != X.shape or not out.flags.f_contiguous or out.dtype != np.float64:
                raise ValueError("Serve.submit: out must be an n x m float64 Fortran array")
            _petsc_check(lib().lgh_serve_submit(self.h, op, c, X.shape[1], _d(X),
                                                X.shape[0], _d(out), X.shape[0],
                                                C.byref(t)), "lgh_serve_submit")
        with self._lock:
            self._live[t.value] = (X, out)
        return t.value

    def wait(self, ticket):
        ld, lat = C.c_double(), C.c_double()
        _petsc_check(lib().lgh_serve_wait(self.h, ticket, C.byref(ld), C.byref(lat)),
                     "lgh_serve_wait")
        with self._lock:
            X, out = self._live.pop(ticket)
        return (ld.value if out is None else out), lat.value

    def stats(self):
        s = ServeStats()
        _petsc_check(lib().lgh_serve_get_stats(self.h, C.byref(s)), "lgh_serve_get_stats")
        return s.as_dict()

    def close(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.lgh_serve_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


# ---- kernel ABI -------------------------------------------------------------------

def init():
    _hip_check(lib().lgh_hip_init(), "lgh_hip_init")


PROF_KINDS = {"spmm_B": 0, "spmm": 1, "cheb_jacobi": 2, "cheb_z": 3,
              "resid": 4, "gram": 5, "gemm": 6, "other": 7, "cheb_mixed": 8,
              "smooth_pre_fine": 9, "smooth_post_fine": 10, "resid_fine": 11,
              "mg_coarse": 12}


def prof_enable(on=True):
    lib().lgh_hip_prof_enable(1 if on else 0)


def prof_collect(kind, np_=-1, prec=-1):
    """(launches, total ms, algorithmic bytes, flops) since the last collect
    of that kind; np_ >= 0 / prec >= 0 count only launches over np_ panels /
    with vectors stored in that precision."""
    c = C.c_int64()
    ms, by, fl = C.c_double(), C.c_double(), C.c_double()
    _hip_check(lib().lgh_hip_prof_collect(PROF_KINDS.get(kind, kind), np_, prec, C.byref(c),
                                          C.byref(ms), C.byref(by), C.byref(fl)),
               "lgh_hip_prof_collect")
    return c.value, ms.value, by.value, fl.value


def randn_at_host(seed, i, j):
    return lib().lgh_hip_randn_at_host(seed, i, j)


def copy_gbps(nbytes=1 << 30, reps=3):
    """achieved device-to-device copy bandwidth (GB/s, read + write)"""
    g = C.c_double()
    _hip_check(lib().lgh_hip_copy_gbps(nbytes, reps, C.byref(g)), "lgh_hip_copy_gbps")
    return g.value


def randn_host_block(seed, row0, col0, nr, nc, prec=PREC_F64):
    """The device draw arithmetic on the host (correctly rounded log) over a
    block, and the flags of the draws the device fill finishes with the
    platform log: (values, flags)."""
    vals = np.zeros((nr, nc), order="F")
    flags = np.zeros((nr, nc), dtype=np.uint8, order="F")
    _hip_check(lib().lgh_hip_randn_host_block(
        seed, row0, col0, nr, nc, prec, _d(vals), max(nr, 1),
        flags.ctypes.data_as(C.POINTER(C.c_ubyte)), max(nr, 1)),
        "lgh_hip_randn_host_block")
    return vals, flags.astype(bool)


def randn_fill(seed, row0, col0, nr, nc, W=64, prec=PREC_F64):
    out = np.zeros((nr, nc), order="F")
    _hip_check(lib().lgh_hip_randn_fill(seed, row0, col0, nr, nc, W, prec,
                                        _d(out), nr), "lgh_hip_randn_fill")
    return out


def _csr(A):
    import scipy.sparse as sp
    A = sp.csr_matrix(A)
    A.sort_indices()
    return (np.ascontiguousarray(A.indptr, dtype=np.int32),
            np.ascontiguousarray(A.indices, dtype=np.int32),
            np.ascontiguousarray(A.data, dtype=np.float64))


def spmm(A, X, prec_vals=PREC_F64, prec_vec=PREC_F64, W=64, grouped=False):
    rp, ci, v = _csr(A)
    X = _fcol(X)
    Y = np.zeros((A.shape[0], X.shape[1]), order="F")
    fn = lib().lgh_hip_spmm_grouped if grouped else lib().lgh_hip_spmm
    _hip_check(fn(A.shape[0], A.shape[1], _i(rp), _i(ci), _d(v), prec_vals,
                  prec_vec, W, X.shape[1], _d(X), X.shape[0], _d(Y), A.shape[0]),
               "lgh_hip_spmm")
    return Y


def spmm_group_config(rg=0, ppw=0):
    """Force the grouped SpMM's rows per group / panels per wave for later
    uploads (rg = 0: the autotuner decides)."""
    lib().lgh_hip_spmm_group_config(rg, ppw)


def spmm_tile_config(cand=-1):
    """Force the row-tile (LDS-staged) SpMM copy for later uploads: cand
    bits 0-1 the shape, bit 2 2D tiles, bit 3 XCD-run order; -1 ends it."""
    lib().lgh_hip_spmm_tile_config(cand)


def mg_coarse_config(max_rows=-1, max_nnz=-1):
    """Caps of the multigrid levels that run as one fused launch per coarse
    visit (negative: defaults 1024 rows, 65536 nonzeros; max_rows = 0: one
    launch per operation and level).  Both give the same bits."""
    lib().lgh_hip_mg_coarse_config(max_rows, max_nnz)


def spmm_group_report():
    """The autotuner's decisions: list of dicts, one per matrix shape."""
    n = lib().lgh_hip_spmm_group_report(None, 0)
    buf = C.create_string_buffer(n + 1)
    lib().lgh_hip_spmm_group_report(buf, n + 1)
    out = []
    for line in buf.value.decode().splitlines():
        f = line.split()
        out.append({"n": int(f[0]), "nnz": int(f[1]), "value_prec": int(f[2]),
                    "rows_per_group": int(f[3]), "panels_per_wave_fp64": int(f[4]),
                    "panels_per_wave_fp32": int(f[5]), "ms_per_panel_fp64": float(f[6]),
                    "ms_per_panel_fp32": float(f[7]), "tile": int(f[8]),
                    "tile_rows_per_wave": int(f[9]), "tile_waves": int(f[10]),
                    "tile_chunk": int(f[11]), "tile_tx": int(f[12]),
                    "tile_stride": int(f[13]), "tile_xcd": int(f[14])})
    return out


def cheb_jacobi(A, lmin, lmax, nit, B, prec=PREC_F64, W=64):
    rp, ci, v = _csr(A)
    B = _fcol(B)
    X = np.zeros(B.shape, order="F")
    _hip_check(lib().lgh_hip_cheb_jacobi(A.shape[0], _i(rp), _i(ci), _d(v), prec,
                                         W, lmin, lmax, nit, B.shape[1], _d(B),
                                         B.shape[0], _d(X), B.shape[0]),
               "lgh_hip_cheb_jacobi")
    return X


def gram(X, Y, prec=PREC_F64):
    X, Y = _fcol(X), _fcol(Y)
    G = np.zeros((X.shape[1], Y.shape[1]), order="F")
    _hip_check(lib().lgh_hip_gram(prec, X.shape[0], X.shape[1], _d(X), X.shape[0],
                                  Y.shape[1], _d(Y), Y.shape[0], _d(G),
                                  X.shape[1]), "lgh_hip_gram")
    return G


def gemm_nn(X, Cm, prec=PREC_F64):
    X, Cm = _fcol(X), _fcol(Cm)
    Y = np.zeros((X.shape[0], Cm.shape[1]), order="F")
    _hip_check(lib().lgh_hip_gemm_nn(prec, X.shape[0], X.shape[1], _d(X),
                                     X.shape[0], Cm.shape[1], _d(Cm),
                                     Cm.shape[0], _d(Y), X.shape[0]),
               "lgh_hip_gemm_nn")
    return Y


class AMG:
    """Host-built hierarchy (lgh_amg_build); levels() exports it."""

    def __init__(self, Z, smax_krylov=1, smoother_top=1.1, nu_force=0):
        rp, ci, v = _csr(Z)
        h = vp()
        _hip_check(lib().lgh_amg_build(Z.shape[0], _i(rp), _i(ci), _d(v),
                                       smax_krylov, smoother_top, nu_force,
                                       C.byref(h)), "lgh_amg_build")
        self.h = h

    def levels(self):
        import scipy.sparse as sp
        L = lib()
        out = []
        for l in range(L.lgh_amg_levels(self.h)):
            n, nc, nu = C.c_int(), C.c_int(), C.c_int()
            rp, ci, prp, pci = ip(), ip(), ip(), ip()
            v, pv = dp(), dp()
            slo, shi = C.c_double(), C.c_double()
            L.lgh_amg_level(self.h, l, C.byref(n), C.byref(nc), C.byref(rp),
                            C.byref(ci), C.byref(v), C.byref(prp), C.byref(pci),
                            C.byref(pv), C.byref(slo), C.byref(shi), C.byref(nu))
            nn = n.value
            r = np.ctypeslib.as_array(rp, shape=(nn + 1,)).copy()
            nnz = int(r[-1])
            A = sp.csr_matrix((np.ctypeslib.as_array(v, shape=(max(nnz, 1),))[:nnz].copy(),
                               np.ctypeslib.as_array(ci, shape=(max(nnz, 1),))[:nnz].copy(), r),
                              shape=(nn, nn))
            P = None
            if l > 0:
                pr = np.ctypeslib.as_array(prp, shape=(nn + 1,)).copy()
                pn = int(pr[-1])
                P = sp.csr_matrix((np.ctypeslib.as_array(pv, shape=(pn,)).copy(),
                                   np.ctypeslib.as_array(pci, shape=(pn,)).copy(), pr),
                                  shape=(nn, nc.value))
            out.append({"A": A, "P": P, "slo": slo.value, "shi": shi.value,
                        "nu": nu.value})
        return out

    def set_cycle(self, gamma):
        """1: V-cycle, 2: W-cycle (lgh_amg_set_cycle)"""
        _hip_check(lib().lgh_amg_set_cycle(self.h, gamma), "lgh_amg_set_cycle")

    def vcycle(self, B, prec=PREC_F64, W=64):
        B = _fcol(B)
        X = np.zeros(B.shape, order="F")
        _hip_check(lib().lgh_amg_vcycle(self.h, prec, W, B.shape[1], _d(B),
                                        B.shape[0], _d(X), B.shape[0]),
                   "lgh_amg_vcycle")
        return X

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.lgh_amg_free(self.h)
            self.h = None
