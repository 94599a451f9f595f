"""GLR build + downstream benchmark on one MI355X.

Workload (the plain-vs-mixed GPU build configuration): 2D 1024 x 1024 grid,
n = 1,048,576; B an anisotropic Gaussian PSF (angle pi*x1, axes 3h / 1.5h,
Mahalanobis <= 3, ~127 nnz/row); Z = M + 0.05 K; ell = 400 (q = 1) in
64-column panels; Z^{-1} by blocked Chebyshev around the AMG V-cycle.
A step is one mixed-precision lgh_glr_compute on inputs already resident in
HBM.  Reported beside it: the plain FP64 build, the error of the mixed
build against it, solve RHS/s with p50/p99 latency, the roofline of the
dominant kernel (HIP events, algorithmic bytes), and the FP64 CPU oracle
timed on a bounded sample of the same build.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "lgpsf-hessian_amd")
sys.path.insert(0, ROOT)
sys.path.insert(0, PKG_DIR)

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_PEAK_TFS = 157.3
# profiler kind -> the kernel instantiation it times (rocprofv3 name), for
# the V-cycle storage type of the FP64 sweeps
def kernel_of(kind, vc):
    st = {"bf16": "lgh::bf16_t", "f32": "float"}[vc]
    return {"cheb_mixed": "lgh::cheb64m_k<true, %s, 64, float, float, double, double>" % st,
            "smooth_pre_fine": "lgh::smx_k<float, float, %s, false, 64, true>" % st,
            "smooth_post_fine": "lgh::smx_k<float, float, %s, true, 64, true>" % st,
            "resid_fine": "lgh::resid64_k<float, float, %s, 64>" % st}[kind]


def load_pkg():
    spec = importlib.util.spec_from_file_location(
        "lgpsf_hessian_amd", os.path.join(PKG_DIR, "__init__.py"),
        submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["lgpsf_hessian_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def pmc_traffic(kernel, panels):
    """HBM bytes per launch of `kernel` at `panels` panels from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration recorded there), or None when it holds no such line."""
    path = os.path.join(ROOT, "profiles", "glr_build_grid2d_1024_pmc_traffic.md")
    try:
        with open(path) as f:
            for line in f:
                t = line.split()
                if (len(t) >= 4 and t[0] == "traffic_bytes_per_launch"
                        and " ".join(t[1:-2]) == kernel and int(t[-2]) == panels):
                    return float(t[-1])
    except OSError:
        pass
    return None


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(B, Z, M, lgh, ell, q, kept, budget_s, gamma=1):
    """FP64 oracle (single thread) on a bounded sample: whole F sweeps of a
    few sketch columns with the same Chebyshev-MG operator, scaled to the
    (q+2) ell + kept columns of one build; the dense stages are not in the
    sample (they are < 1% of the oracle's build time here)."""
    import oracle as O
    h = lgh.AMG(Z)
    levels = [{"A": O.CSR.from_scipy(l["A"]),
               "P": O.CSR.from_scipy(l["P"]) if l["P"] is not None else None,
               "slo": l["slo"], "shi": l["shi"], "nu": l["nu"]} for l in h.levels()]
    mg = O.MG(levels, gamma=gamma)
    return mg, levels


def serve_bench(lgh, g, n, nreq, pool, seed=5, rate=None):
    """Request loop on the resident mixed GLR: 60% solve / 35% sample / 5%
    logdet, c log-uniform in [1e-4, 10], RHS count geometric (mean 8)
    truncated to [1, 64], entries from a pool of hashed-Gaussian columns.
    First a burst of min(nreq/3, 1000) requests measures saturation, then nreq
    requests arrive as a Poisson process at 0.7 x the saturated rate (or
    at `rate` requests/s when given)."""
    import threading
    rng = np.random.default_rng(seed)

    def make(k):
        out = []
        for _ in range(k):
            u = rng.uniform()
            op = lgh.SERVE_SOLVE if u < 0.6 else lgh.SERVE_SAMPLE if u < 0.95 else lgh.SERVE_LOGDET
            m = int(min(64, rng.geometric(1.0 / 8.0)))
            off = int(rng.integers(0, 64 - m + 1))
            out.append((op, 10.0 ** rng.uniform(-4, 1), pool[:, off:off + m]))
        return out

    srv = lgh.Serve(g)
    free = {}                       # reusable result blocks by width
    flock = threading.Lock()

    def take(m):
        with flock:
            lst = free.get(m)
            if lst:
                return lst.pop()
        return np.empty((n, m), order="F")

    def give(a):
        if a is not None and not isinstance(a, float):
            with flock:
                free.setdefault(a.shape[1], []).append(a)

    def run(reqs, times):
        """submit reqs at times (s from start) on this thread, collect on
        another; returns (elapsed s, latencies, RHS columns)"""
        import queue
        lat, cols = [None] * len(reqs), 0
        qq = queue.Queue()

        def waiter():
            for _ in range(len(reqs)):
                i, t = qq.get()
                res, lat[i] = srv.wait(t)
                give(res)

        th = threading.Thread(target=waiter)
        th.start()
        t0 = time.perf_counter()
        for i, ((op, c, X), ta) in enumerate(zip(reqs, times)):
            dt = t0 + ta - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            if op == lgh.SERVE_LOGDET:
                qq.put((i, srv.submit(op, c)))
            else:
                cols += X.shape[1]
                qq.put((i, srv.submit(op, c, X, out=take(X.shape[1]))))
        th.join()
        return time.perf_counter() - t0, np.array(lat), cols

    warm = make(8)
    run(warm, [0.0] * len(warm))
    burst = make(max(min(nreq // 3, 1000), 16))
    t_b, lat_b, cols_b = run(burst, [0.0] * len(burst))
    sat_rps = len(burst) / t_b
    rate = 0.7 * sat_rps if rate is None else rate
    reqs = make(nreq)
    arr = np.cumsum(rng.exponential(1.0 / rate, size=nreq))
    t_p, lat_p, cols_p = run(reqs, list(arr))
    st = srv.stats()
    srv.close()
    return {"requests": nreq, "saturation_req_per_s": sat_rps,
            "saturation_rhs_per_s": cols_b / t_b,
            "poisson_rate_req_per_s": rate,
            "rhs_per_s": cols_p / t_p, "req_per_s": nreq / t_p,
            "latency_p50_s": float(np.percentile(lat_p, 50)),
            "latency_p99_s": float(np.percentile(lat_p, 99)),
            "batch_cols": st["batch_cols"], "batches": st["batches"],
            "mean_batch_cols": st["columns"] / max(st["batches"], 1),
            "worker_busy_s": st["busy_s"], "worker_stage_s": st["stage_s"]}


def dominant_roofline(lgh, args, prof, npan, vpc):
    """roofline object of the dominant single full-width kernel kind"""
    single = ("cheb_mixed", "smooth_pre_fine", "smooth_post_fine", "resid_fine")
    dom = max(single, key=lambda k: prof[k][1])
    c, ms, by, fl = prof[dom]
    avg_ms = ms / c
    ach = by / c / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
            "kernel": kernel_of(dom, args.vcycle), "panels": npan,
            "avg_launch_ms": avg_ms, "launches": c,
            "algorithmic_bytes_per_launch": by / c}


def grid3d_bench(args, W, lgh):
    """3D configuration: 128^3 grid (n = 2 097 152), Z = M + 0.1 K7, B an
    isotropic PSF (sigma 1.5h, radius-2h ball, 33 nnz/row) with FP32
    values; mixed build at ell = 400, then lgh_glr_extend by 200 twice
    (ell 800), each timed; the extended spectrum is checked against a
    one-shot ell = 800 build (extend == one-shot for the hashed sketch)."""
    side = 128 if args.side == 1024 else args.side
    t0 = time.time()
    Bs, Zs, M = W.psf3d_problem(side)
    n = Zs.shape[0]
    log("3D problem n=%d nnz(B)=%d (%.1f/row) gen %.1fs"
        % (n, Bs.nnz, Bs.nnz / n, time.time() - t0))
    lgh.init()
    t0 = time.time()
    prior = lgh.Prior.from_mat(Zs, M, blocked_mode=args.mode)
    info = prior.info()
    log("prior create %.1fs %s" % (time.time() - t0, info))
    Bm = lgh.Mat.aij(Bs)
    pcode = {"bf16": lgh.PREC_BF16, "f32": lgh.PREC_F32}
    mixed = lgh.xopts(mixed=True, prec_precond=pcode[args.vcycle],
                      prec_sketch_precond=pcode[args.sketch_vcycle])
    ell0, ext = 400, 200
    opts = lgh.glr_opts(ell=ell0, q_power=1, seed=0)
    for w in range(args.warmup):
        t0 = time.time()
        g = lgh.GLR(Bm, prior, opts, x=mixed)
        log("warmup build %d: %.2fs kept=%d" % (w, time.time() - t0, g.kept))
        del g
    lgh.prof_enable(True)
    for k in lgh.PROF_KINDS:
        lgh.prof_collect(k)
    lgh.lib().lgh_hip_sync()
    t0 = time.perf_counter()
    g = None
    for s in range(args.steps):
        g = None
        g = lgh.GLR(Bm, prior, opts, x=mixed)
    lgh.lib().lgh_hip_sync()
    t_build = (time.perf_counter() - t0) / args.steps
    lgh.prof_enable(False)
    npan = ell0 // 64 if 0 < ell0 % 64 <= 48 else (ell0 + 63) // 64
    vpc = pcode[args.vcycle]
    single = ("cheb_mixed", "smooth_pre_fine", "smooth_post_fine", "resid_fine")
    prof = {k: lgh.prof_collect(k, npan, vpc) if k in single else lgh.prof_collect(k)
            for k in lgh.PROF_KINDS}
    log("3D mixed build %.3fs kept=%d report=%s" % (t_build, g.kept, g.report))
    for k, (c, ms, by, fl) in prof.items():
        if c:
            log("  %-16s launches %6d  %9.2f ms  avg %.3f ms  %7.1f GB/s  %6.2f TF/s"
                % (k, c, ms, ms / c, by / ms / 1e6, fl / ms / 1e9))
    roof = dominant_roofline(lgh, args, prof, npan, vpc)
    t_ext = []
    for e in range(2):
        lgh.lib().lgh_hip_sync()
        t0 = time.perf_counter()
        rep = g.extend(ext)
        lgh.lib().lgh_hip_sync()
        t_ext.append(time.perf_counter() - t0)
        log("extend +%d -> ell %d: %.3fs kept=%d" % (ext, ell0 + ext * (e + 1),
                                                     t_ext[-1], rep["kept"]))
    peak_gb = lgh.lib().lgh_hip_bytes_in_use() / 1e9
    # the achievable HBM rate on this box (a 1 GiB device copy), beside the
    # nominal 8 TB/s
    copy_peak = lgh.copy_gbps(1 << 30, 5)
    for r in (roof,):
        if r:
            r["copy_peak_gbs"] = copy_peak
            r["frac_of_copy_peak"] = r["achieved"] / copy_peak
    log("device copy: %.0f GB/s" % copy_peak)
    lam_ext = np.array(g.lam)
    kept_ext = g.kept
    del g
    lgh.lib().lgh_hip_trim()
    t0 = time.perf_counter()
    g8 = lgh.GLR(Bm, prior, lgh.glr_opts(ell=ell0 + 2 * ext, q_power=1, seed=0),
                 x=mixed)
    lgh.lib().lgh_hip_sync()
    t_one = time.perf_counter() - t0
    k = min(kept_ext, g8.kept)
    agree = float(np.max(np.abs(lam_ext[:k] - g8.lam[:k]) / np.abs(g8.lam[:k])))
    log("one-shot ell=%d build %.3fs kept=%d; extend vs one-shot spectrum rel %.2e"
        % (ell0 + 2 * ext, t_one, g8.kept, agree))
    line = {
        "metric": "GLR build s + solve RHS/s, p99 latency; mixed vs FP64 plain at rel err <=1e-6",
        "value": t_build, "unit": "s/build", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_build * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": "3D %d^3 grid n=%d, isotropic PSF B (sigma "
                               "1.5h, radius 2h, %.1f nnz/row, FP32 values), "
                               "Z=M+0.1K7, ell=%d q=1, then extend +%d twice, "
                               "Z^-1 = Chebyshev(%d) around AMG V-cycle"
                               % (side, n, Bs.nnz / n, ell0, ext, info["nit"]),
                   "n": n, "ell": ell0, "kept": kept_ext,
                   "vcycle_fp64_sweeps": args.vcycle,
                   "spmm_autotune": [r for r in lgh.spmm_group_report() if r["n"] == n]},
        "roofline": roof,
        "extend_s": t_ext, "extend_cols": ext, "ell_final": ell0 + 2 * ext,
        "kept_after_extends": kept_ext, "one_shot_build_s": t_one,
        "kept_one_shot": g8.kept, "extend_vs_one_shot_spectrum_rel": agree,
        "device_bytes_in_use_gb": peak_gb,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def _small_instance_gen(k):
    import workloads as W
    B, Z, M, ell = W.small_psf2d_instance(k)
    return B, Z, M, ell


def batch_bench(args, W, lgh):
    """Batched configuration: 256 small 2D instances (workloads.small_psf2d_instance:
    n = 4k..64k, ell = rank + 10 with rank in {32, 64, 128}), mixed
    precision.  Inputs are generated on the host (process pool, outside the
    timed region); timed: the priors (host AMG set-up + interval estimate)
    and the GLR builds of all instances, one call at a time and as one
    batch (`value`: the batched builds)."""
    import multiprocessing as mp
    t0 = time.time()
    with mp.get_context("fork").Pool(min(16, os.cpu_count() or 1)) as pool:
        inst = pool.map(_small_instance_gen, range(args.instances), chunksize=1)
    rows = sum(z.shape[0] for _, z, _, _ in inst)
    nnz = sum(b.nnz for b, _, _, _ in inst)
    log("batch: %d instances, %d rows, %d nnz(B) (%.1f/row), gen %.1fs"
        % (len(inst), rows, nnz, nnz / rows, time.time() - t0))
    lgh.init()
    mats = [(lgh.Mat.aij(b), z, m, ell) for b, z, m, ell in inst]
    mixed = lgh.xopts(mixed=True)
    # warm-up: one instance of each rank (kernels loaded, allocator primed)
    for j in range(min(3, len(mats))):
        Bm, z, m, ell = mats[j]
        pr = lgh.Prior.from_mat(z, m, blocked_mode=3)
        lgh.GLR(Bm, pr, lgh.glr_opts(ell=ell, q_power=1, seed=0), x=mixed)
    t_prior = t_seq = None
    kept = None
    if not args.skip_sequential:
        lgh.lib().lgh_hip_sync()
        t0 = time.perf_counter()
        priors = [lgh.Prior.from_mat(z, m, blocked_mode=3) for _, z, m, _ in mats]
        lgh.lib().lgh_hip_sync()
        t_prior = time.perf_counter() - t0
        infos = [pr.info() for pr in priors]
        t0 = time.perf_counter()
        kept, lams = [], []
        for (Bm, z, m, ell), pr in zip(mats, priors):
            g = lgh.GLR(Bm, pr, lgh.glr_opts(ell=ell, q_power=1, seed=0), x=mixed)
            kept.append(g.kept)
            lams.append(np.array(g.lam))
            del g
        lgh.lib().lgh_hip_sync()
        t_seq = time.perf_counter() - t0
        log("batch one at a time: priors %.2fs, builds %.2fs (%.1f instances/s)"
            % (t_prior, t_seq, len(mats) / t_seq))
        del priors
    # the same instances as batches (lgh_prior_create_mat_batch,
    # lgh_glr_compute_batch): host worker threads, one stream each
    lgh.lib().lgh_hip_sync()
    t0 = time.perf_counter()
    bpri = lgh.Prior.batch([(z, m) for _, z, m, _ in mats],
                           workers=args.batch_workers, blocked_mode=3)
    t_bprior = time.perf_counter() - t0
    t0 = time.perf_counter()
    gs = lgh.GLR.batch([(Bm, pr, lgh.glr_opts(ell=ell, q_power=1, seed=0))
                        for (Bm, _, _, ell), pr in zip(mats, bpri)],
                       x=mixed, workers=args.batch_workers)
    lgh.lib().lgh_hip_sync()
    t_batch = time.perf_counter() - t0
    kept_b = [g.kept for g in gs]
    # the batched builds equal the single calls bit for bit (the workers only
    # change which stream runs an instance)
    same_bits = None if kept is None else all(
        np.array_equal(np.array(g.lam), l) for g, l in zip(gs, lams))
    if kept is not None and not same_bits:
        bad = [i for i, (g, l) in enumerate(zip(gs, lams))
               if not np.array_equal(np.array(g.lam), l)]
        worst = max(float(np.max(np.abs(np.array(gs[i].lam) - lams[i])
                                 / np.abs(lams[i]))) if len(lams[i]) else 0.0
                    for i in bad)
        pdiff = [i for i, (a, b) in enumerate(zip(infos, bpri)) if a != b.info()]
        log("batch vs one at a time: %d instances differ (first %s, max rel %.2e); "
            "priors differ: %d %s" % (len(bad), bad[:8], worst, len(pdiff), pdiff[:8]))
    del gs
    log("batch (%d workers): priors %.2fs, builds %.2fs (%.1f instances/s)%s"
        % (args.batch_workers or 8, t_bprior, t_batch, len(mats) / t_batch,
           "" if kept is None else ", kept " + ("same" if kept_b == kept else "DIFFERENT")
           + ", spectra " + ("bit-identical" if same_bits else "DIFFERENT")))
    line = {
        "metric": "GLR build s + solve RHS/s, p99 latency; mixed vs FP64 plain at rel err <=1e-6",
        "value": t_batch, "unit": "s/batch", "n_gpus": 1, "steps": 1,
        "warmup": 1, "ms_per_step": t_batch * 1e3, "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "%d independent 2D instances n=%d..%d (sum %d rows), "
                               "PSF sigma U[1.5h,3h] radius 3 sigma (%.1f nnz/row), "
                               "Z=M+0.1K, ell=rank+10, rank in {32,64,128}, mixed"
                               % (len(mats), min(z.shape[0] for _, z, _, _ in inst),
                                  max(z.shape[0] for _, z, _, _ in inst), rows,
                                  nnz / rows),
                   "instances": len(mats)},
        "instances_per_s": len(mats) / t_batch, "priors_s": t_bprior,
        "sequential_builds_s": t_seq, "sequential_priors_s": t_prior,
        "batch_workers": args.batch_workers or 8,
        "kept_total": int(sum(kept_b)),
        "kept_same_as_sequential": None if kept is None else kept_b == kept,
        "spectra_bitwise_same_as_sequential": same_bits,
        "cpu_baseline": None,
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--side", type=int, default=1024)
    ap.add_argument("--ell", type=int, default=400)
    ap.add_argument("--mode", type=int, default=3)
    ap.add_argument("--cpu-cols", type=int, default=64,
                    help="F-sweep columns the CPU baseline measures (one panel)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--vcycle", choices=("bf16", "f32"), default="bf16",
                    help="V-cycle storage precision inside the FP64 sweeps")
    ap.add_argument("--sketch-vcycle", choices=("bf16", "f32"), default="bf16",
                    help="V-cycle storage precision inside the FP32 sweeps (the "
                         "residual-gate sweep; the sketch too with --sketch f32)")
    ap.add_argument("--sketch", choices=("f32", "f64"), default=None,
                    help="vector precision of the sketch sweep (default: the "
                         "mixed policy's)")
    ap.add_argument("--extend", type=int, default=200,
                    help="columns of the timed lgh_glr_extend (0: skip)")
    ap.add_argument("--serve-rate", type=float, default=None,
                    help="Poisson arrival rate of the serving run in requests/s "
                         "(default 0.7 x the measured saturation)")
    ap.add_argument("--serve-requests", type=int, default=10000,
                    help="requests of the Poisson serving run (0: skip)")
    ap.add_argument("--workload", choices=("grid2d", "grid3d", "batch"),
                    default="grid2d",
                    help="grid2d: the 2D 1024^2 build (the headline line); "
                         "grid3d: the 3D 128^3 build at ell=400, extended by 200 "
                         "twice; batch: 256 small 2D instances built as batches")
    ap.add_argument("--instances", type=int, default=256,
                    help="batch: number of instances")
    ap.add_argument("--batch-workers", type=int, default=0,
                    help="batch: concurrent build workers (0: the library's default)")
    ap.add_argument("--skip-sequential", action="store_true",
                    help="batch: time the batches only")
    ap.add_argument("--backend", choices=("replicated", "implicit"),
                    default="replicated",
                    help="grid2d: U explicit (LGH_GLR_REPLICATED) or implicit "
                         "U = Q V (LGH_GLR_SCALAPACK at one process)")
    ap.add_argument("--sweep-panels", type=int, default=0,
                    help="64-column panels per F sweep chunk (0: all at once "
                         "unless device memory is short)")
    ap.add_argument("--solve-blocked-min", type=int, default=None,
                    help="downstream calls on at least this many columns use "
                         "the blocked Z-solve tier (default: the mixed "
                         "policy's)")
    ap.add_argument("--mg-coarse-rows", type=int, default=-1,
                    help="multigrid levels of at most this many rows run as one "
                         "fused launch per coarse visit (0: one launch per "
                         "operation and level; -1: the library's default)")
    ap.add_argument("--copy-calib", action="store_true",
                    help="only run a 1 GiB device-to-device copy three times (the "
                         "known-size kernel that scales rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calib:
        import torch
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("copied 3 x %d bytes" % (x.numel() * 4))
        return
    if args.gpus != 1:
        print("bench.py runs on one GPU only (the engine has no multi-GPU "
              "path); use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import workloads as W
    lgh = load_pkg()
    if args.mg_coarse_rows >= 0:
        lgh.mg_coarse_config(args.mg_coarse_rows, -1)
    if args.workload == "grid3d":
        return grid3d_bench(args, W, lgh)
    if args.workload == "batch":
        return batch_bench(args, W, lgh)
    t0 = time.time()
    Bs, Zs, M = W.aniso_psf2d_problem(args.side)
    n = Zs.shape[0]
    log("problem n=%d nnz(B)=%d (%.1f/row) gen %.1fs" % (n, Bs.nnz, Bs.nnz / n,
                                                           time.time() - t0))
    lgh.init()
    t0 = time.time()
    prior = lgh.Prior.from_mat(Zs, M, blocked_mode=args.mode)
    info = prior.info()
    log("prior create %.1fs %s" % (time.time() - t0, info))
    Bm = lgh.Mat.aij(Bs)
    backend = {"replicated": 1, "implicit": 2}[args.backend]
    opts = lgh.glr_opts(ell=args.ell, q_power=1, seed=0, backend=backend)
    pcode = {"bf16": lgh.PREC_BF16, "f32": lgh.PREC_F32}
    mixed = lgh.xopts(mixed=True, prec_precond=pcode[args.vcycle],
                      prec_sketch_precond=pcode[args.sketch_vcycle],
                      sweep_panels=args.sweep_panels)
    if args.sketch:
        mixed.prec_sketch = {"f32": lgh.PREC_F32, "f64": lgh.PREC_F64}[args.sketch]
    if args.solve_blocked_min is not None:
        mixed.solve_blocked_min = args.solve_blocked_min
    plain = lgh.xopts(mixed=False, sweep_panels=args.sweep_panels)

    for w in range(args.warmup):
        t0 = time.time()
        g = lgh.GLR(Bm, prior, opts, x=mixed)
        log("warmup build %d: %.2fs kept=%d" % (w, time.time() - t0, g.kept))
        del g

    lgh.prof_enable(True)
    for k in lgh.PROF_KINDS:
        lgh.prof_collect(k)
    lgh.lib().lgh_hip_sync()
    t0 = time.perf_counter()
    reps = []
    g = None
    for s in range(args.steps):
        g = lgh.GLR(Bm, prior, opts, x=mixed)
        reps.append(g.report)
    lgh.lib().lgh_hip_sync()
    t_mixed = (time.perf_counter() - t0) / args.steps
    lgh.prof_enable(False)
    # one kernel instantiation at one launch shape per kind below: the
    # full-width launches on the finest level (all whole 64-column panels;
    # a last panel of at most 48 columns runs as 16-wide panels)
    npan = args.ell // 64 if 0 < args.ell % 64 <= 48 else (args.ell + 63) // 64
    npan_b = (args.ell + 63) // 64
    if 0 < args.sweep_panels < npan:      # chunked sweeps: the full chunks
        npan = args.sweep_panels
        npan_b = args.sweep_panels
    vpc = pcode[args.vcycle]
    single = ("cheb_mixed", "smooth_pre_fine", "smooth_post_fine", "resid_fine")
    prof = {k: lgh.prof_collect(k, npan, vpc) if k in single else lgh.prof_collect(k)
            for k in lgh.PROF_KINDS if k != "spmm_B"}
    # the B SpMM of the FP64 sweeps (all ceil(ell / 64) panels of a block in
    # one launch; the residual check's FP32 one-panel product is not this
    # kernel instantiation)
    prof["spmm_B"] = lgh.prof_collect("spmm_B", npan_b, lgh.PREC_F64)
    kept_build = g.kept
    log("mixed build %.3fs kept=%d report=%s" % (t_mixed, kept_build, reps[-1]))
    for k, (c, ms, by, fl) in prof.items():
        if c:
            log("  %-16s launches %6d  %9.2f ms  avg %.3f ms  %7.1f GB/s  %6.2f TF/s"
                % (k, c, ms, ms / c, by / ms / 1e6, fl / ms / 1e9))

    # dominant single kernel (by total time) -> roofline
    dom = max(single, key=lambda k: prof[k][1])
    c, ms, by, fl = prof[dom]
    roof = None
    if c:
        avg_ms = ms / c
        ach = by / c / (avg_ms * 1e-3) / 1e9
        roof = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": ach / HBM_PEAK_GBS, "traffic": None,
                "kernel": kernel_of(dom, args.vcycle), "panels": npan,
                "avg_launch_ms": avg_ms, "launches": c,
                "algorithmic_bytes_per_launch": by / c}
        tr = pmc_traffic(kernel_of(dom, args.vcycle), npan)
        if tr is not None:
            roof["traffic"] = tr
            roof["traffic_source"] = "profiles/glr_build_grid2d_1024_pmc_traffic.md"
    else:
        log("no full-width launches at %d panels: roofline not reported" % npan)
    sp = prof["spmm_B"]
    spmm_roof = None
    if sp[0]:
        a = sp[2] / sp[1] / 1e6
        spmm_roof = {"kernel": "spmm_B", "achieved": a, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                     "avg_launch_ms": sp[1] / sp[0]}
    mf = {}
    for k in ("gram", "gemm"):
        c_, ms_, by_, fl_ = prof[k]
        if c_:
            mf[k] = {"achieved_tflops": fl_ / ms_ / 1e9, "peak_fp64_tflops": FP64_PEAK_TFS,
                     "frac": fl_ / ms_ / 1e9 / FP64_PEAK_TFS}
    peak_gb = lgh.lib().lgh_hip_bytes_in_use() / 1e9
    # the achievable HBM rate on this box (a 1 GiB device copy), beside the
    # nominal 8 TB/s
    copy_peak = lgh.copy_gbps(1 << 30, 5)
    for r in (roof, spmm_roof):
        if r:
            r["copy_peak_gbs"] = copy_peak
            r["frac_of_copy_peak"] = r["achieved"] / copy_peak
    log("device copy: %.0f GB/s" % copy_peak)

    extra = {}
    if not args.no_plain:
        t0 = time.perf_counter()
        gp = lgh.GLR(Bm, prior, opts, x=plain)
        lgh.lib().lgh_hip_sync()
        t_plain = time.perf_counter() - t0
        k = min(g.kept, gp.kept)
        spec_err = float(np.max(np.abs(g.lam[:k] - gp.lam[:k]) / np.abs(gp.lam[:k])))
        extra.update(plain_build_s=t_plain, kept_plain=gp.kept,
                     mixed_vs_fp64_spectrum_rel=spec_err,
                     resid_max_plain=gp.report["resid_max"],
                     resid_max_mixed=reps[-1]["resid_max"])
        log("plain build %.3fs kept=%d; mixed-vs-fp64 spectrum rel err %.2e"
            % (t_plain, gp.kept, spec_err))

    # downstream: batched multi-shift solves (host arrays); right-hand sides
    # from the device hashed Gaussian
    rng = np.random.default_rng(0)
    m = 64
    R = lgh.randn_fill(99, 0, 0, n, m)
    cs = 10.0 ** rng.uniform(-4, 1, size=m)
    X = g.solve_block(cs, R)
    t0 = time.perf_counter()
    X = g.solve_block(cs, R)
    t_blk = time.perf_counter() - t0
    # single-RHS latency: one untimed call (it sets up the narrow panel's
    # work buffers, as the untimed block solve above does for 64 columns),
    # then 20 timed calls
    g.solve(cs[0], R[:, 0])
    lat = []
    for j in range(20):
        t1 = time.perf_counter()
        g.solve(cs[j], R[:, j])
        lat.append(time.perf_counter() - t1)
    extra.update(solve_rhs_per_s=m / t_blk, solve_block_m=m,
                 solve_latency_p50_s=float(np.percentile(lat, 50)),
                 solve_latency_p99_s=float(np.percentile(lat, 99)))
    # the same solves on right-hand sides already in HBM (lgh_glr_solve_block_dev:
    # no host transfer), one untimed call each first
    Rd = lgh.DeviceBlock.from_host(R)
    Xd = g.solve_block_device(cs, Rd)
    t0 = time.perf_counter()
    Xd = g.solve_block_device(cs, Rd)
    t_dev = time.perf_counter() - t0
    same = bool(np.array_equal(Xd.to_host(), X))
    g.solve_block_device(cs[:1], Rd.columns(0, 1))
    lat_d = []
    for j in range(20):
        t1 = time.perf_counter()
        g.solve_block_device(cs[j], Rd.columns(j, j + 1))
        lat_d.append(time.perf_counter() - t1)
    extra.update(solve_rhs_per_s_resident=m / t_dev, solve_resident_same_bits=same,
                 solve_latency_resident_p50_s=float(np.percentile(lat_d, 50)))
    del Rd, Xd
    if not args.no_plain:
        t0 = time.perf_counter()
        Xp = gp.solve_block(cs, R)
        extra["plain_solve_rhs_per_s"] = m / (time.perf_counter() - t0)
        extra["mixed_vs_fp64_solve_rel"] = float(
            np.max(np.linalg.norm(X - Xp, axis=0) / np.linalg.norm(Xp, axis=0)))
        S = g.sample_block(cs, R)
        Sp = gp.sample_block(cs, R)
        extra["mixed_vs_fp64_sample_rel"] = float(
            np.max(np.linalg.norm(S - Sp, axis=0) / np.linalg.norm(Sp, axis=0)))
        ld = np.array([g.logdet(c) for c in cs])
        ldp = np.array([gp.logdet(c) for c in cs])
        extra["mixed_vs_fp64_logdet_rel"] = float(np.max(np.abs(ld - ldp) / np.abs(ldp)))
        del S, Sp, Xp
    log("solve block m=%d: %.1f RHS/s (%.1f resident in HBM, same bits %s); "
        "single-RHS p50 %.4fs (%.4fs resident)" %
        (m, m / t_blk, extra["solve_rhs_per_s_resident"],
         extra["solve_resident_same_bits"], extra["solve_latency_p50_s"],
         extra["solve_latency_resident_p50_s"]))
    del X
    # incremental extension of the resident mixed GLR by 200 columns
    # (hashed columns ell..ell+199, power pass, BCGS2, bordered projection);
    # the served object below keeps its 56 modes
    if args.extend > 0:
        lgh.lib().lgh_hip_sync()
        t0 = time.perf_counter()
        rep = g.extend(args.extend)
        lgh.lib().lgh_hip_sync()
        extra.update(extend_cols=args.extend, extend_s=time.perf_counter() - t0,
                     kept_after_extend=rep["kept"])
        log("extend +%d: %.3fs kept=%d" % (args.extend, extra["extend_s"], rep["kept"]))

    # continuous-batching request loop on the (extended) resident GLR
    if args.serve_requests > 0:
        extra["serve"] = serve_bench(lgh, g, n, args.serve_requests, R,
                                     rate=args.serve_rate)
        log("serve:", extra["serve"])

    cpu = None
    if not args.no_cpu:
        import oracle as O
        mg, levels = cpu_baseline(Bs, Zs, M, lgh, args.ell, 1, kept_build, 0,
                                  gamma=info["cycle"])
        zs = O.ZSolver.mat(O.CSR.from_scipy(Zs), mode=3, mg=mg, lmin=info["lmin"],
                           lmax=info["lmax"], nit=info["nit"], ksp_rtol=1e-12)
        msq = np.sqrt(M)
        Bc = O.CSR.from_scipy(Bs)
        # one whole 64-column panel of an F sweep, its columns spread over
        # the host threads (OMP_NUM_THREADS, else every core): the oracle's
        # column-parallel sweeps give the single-thread bits
        threads = int(os.environ.get("OMP_NUM_THREADS") or 0) or (os.cpu_count() or 1)
        threads = max(1, min(threads, os.cpu_count() or 1))
        O.set_threads(threads)
        Xs = O.randn_fill(0, 0, 0, n, args.cpu_cols)
        t0 = time.perf_counter()
        Wk = Xs * msq[:, None]
        T = zs.solve_block(Wk)
        Wk = O.csr_mm(Bc, T)
        T = zs.solve_block(Wk)
        t_cols = time.perf_counter() - t0
        O.set_threads(1)
        cols_build = 3 * args.ell + kept_build
        cpu = {"value": t_cols / args.cpu_cols * cols_build, "unit": "s/build",
               "cores": threads, "kind": "port",
               "sample": "one full %d-column F-sweep panel (two Chebyshev-MG Z-solves "
                         "and the B product) of the FP64 oracle on %d threads, "
                         "measured %.1fs; the build's %d F columns (%.1f panels) "
                         "extrapolated from it; dense stages excluded"
                         % (args.cpu_cols, threads, t_cols, cols_build,
                            cols_build / args.cpu_cols),
               "measured_s": t_cols, "host_nproc": os.cpu_count()}
        log("cpu baseline:", cpu)

    line = {
        "metric": "GLR build s + solve RHS/s, p99 latency; mixed vs FP64 plain at rel err <=1e-6",
        "value": t_mixed,
        "unit": "s/build",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_mixed * 1e3,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "2D %dx%d grid n=%d, anisotropic PSF B "
                               "(%.1f nnz/row), Z=M+0.05K, ell=%d q=1, w=64, "
                               "Z^-1 = Chebyshev(%d) around AMG V-cycle"
                               % (args.side, args.side, n, Bs.nnz / n, args.ell,
                                  info["nit"]),
                   "n": n, "ell": args.ell, "kept": kept_build,
                   "backend": args.backend, "sweep_panels": args.sweep_panels,
                   "solve_blocked_min": mixed.solve_blocked_min,
                   "vcycle_fp64_sweeps": args.vcycle,
                   "vcycle_fp32_sweeps": args.sketch_vcycle,
                   "spmm_autotune": [r for r in lgh.spmm_group_report() if r["n"] == n]},
        "roofline": roof,
        "spmm_roofline": spmm_roof,
        "mfma": mf,
        "device_bytes_in_use_gb": peak_gb,
        "cpu_baseline": cpu,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
