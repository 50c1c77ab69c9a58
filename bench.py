#!/usr/bin/env python3
"""Benchmark: sigma_min grid points per second at relative error <= 1e-4.

One "step" = psTriRecy over the full 200 x 200 recursion grid of the workload (40000 shifts,
zigzag order) with the batched mixed solver, matrix resident in HBM.  Prints ONE JSON line.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload toeplitz|laser|...]

--workload skewlap3d runs psPrecRecy (the preconditioned two-stage solver) on the 3-D skew
Laplacian (n = 15625) over the leading --max-points points of its 200 x 200 traversal.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(name: str):
    import scipy.linalg as sla
    from rls_amd import workloads as WL
    if name == "toeplitz":
        n = 2000
        C = WL.toeplitz_real(n).astype(np.complex128)
        normC = float(sla.svdvals(C)[0])
    elif name == "laser":
        n = 5000
        C = WL.laser(n)
        normC = 1.0
    elif name == "sprand":
        n = 4000
        C = WL.sprand(n=n, density=1e-3, norm2=3.2)
        normC = 3.2
    elif name == "cfd":
        n = 23560
        C = WL.banded_cfd(n=n)
        normC = 645.74
    else:
        raise SystemExit(f"unknown workload {name}")
    g = WL.GRIDS[name]
    return C, normC, g, n


TRUTH_SAMPLES_DEFAULT = 200


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="laser", choices=["laser", "toeplitz", "sprand", "cfd", "skewlap3d"])
    ap.add_argument("--window", type=int, default=0, help="max shifts per speculative window (0 = auto)")
    ap.add_argument("--dense-path", default="auto", choices=["auto", "hessenberg", "schur"],
                    help="dense preprocessing: Hessenberg form on the device with a Hessenberg LU per "
                         "refine shift, or the Schur form (device Hessenberg + host QR iterations); "
                         "auto = schur for n <= 8000 (fastest grid; its host QR is timed as "
                         "preprocessing), hessenberg above")
    ap.add_argument("--no-compare-paths", action="store_true",
                    help="skip the end-to-end (preprocessing + grid) run of the other dense path")
    ap.add_argument("--no-autotune", action="store_true",
                    help="skip the measured choice of the window cap (psTriRecy) or of the sparse "
                         "solve's chunk widths (psPrecRecy); both leave the results unchanged")
    ap.add_argument("--sparse-path", default="auto", choices=["auto", "dense", "factor"],
                    help="sparse C: densified (dense path) or kept sparse with a host LU per refine "
                         "shift; auto = dense for n <= 40000")
    ap.add_argument("--max-points", type=int, default=0,
                    help="solve only the first N points of the traversal (0 = the whole grid)")
    ap.add_argument("--plain-points", type=int, default=4000,
                    help="leading traversal points timed with the plain FP64 solver")
    ap.add_argument("--mid-points", type=int, default=2000,
                    help="points of a mid-grid segment timed with both solvers (0 = skip)")
    ap.add_argument("--invlanc-points", type=int, default=40,
                    help="leading traversal points also solved by psInvLanc on the device (0 = skip)")
    ap.add_argument("--truth-samples", type=int, default=TRUTH_SAMPLES_DEFAULT,
                    help="grid points checked against FP64 ground truth: the 16 smallest computed sigma, "
                         "then half of the rest among the smallest 5 %% and half uniformly at random "
                         "(skewlap3d: 24 random points unless given)")
    ap.add_argument("--cpu-seconds", type=float, default=24.0,
                    help="CPU-restatement time, split over --cpu-segments segments of the traversal")
    ap.add_argument("--cpu-segments", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--acceptance-checks", type=int, default=0, metavar="K",
                    help="for up to K sampled points whose sigma differs from sigma_min by more than "
                         "1e-4, continue the oracle from the GPU state a few points before and report "
                         "its sigma there")
    ap.add_argument("--adjoint", action="store_true",
                    help="keep the certified singular vectors and time d sigma/dz (and the "
                         "Toeplitz-diagonal gradient for --workload toeplitz) after the steps")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device-to-device copy of MIB MiB (x5) and exit: the "
                         "reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        import torch
        n = args.copy_calibration * (1 << 20) // 8
        a = torch.ones(n, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"copied {args.copy_calibration} MiB x5")
        return 0
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (the grid would "
                                   "shard by sub-rectangles across GPUs, not implemented)"}))
        return 2

    if args.workload == "skewlap3d":
        return bench_prec(args)
    import rls_amd as R
    log(f"workload {args.workload}: building matrix")
    C, normC, g, n = build_workload(args.workload)
    x_st, x_ed, y_st, y_ed, nx, ny, k = g
    rec = R.recursion_grid(x_st, x_ed, y_st, y_ed, nx, ny, k)
    if args.max_points > 0:
        rec.Z = rec.Z[: args.max_points]
    import scipy.sparse as sp
    # sparse configurations: kept sparse with a host LU per refine shift (the reference's
    # structure) when the dense operator would not fit, else taken dense
    sparse = sp.issparse(C) and (args.sparse_path == "factor" or (args.sparse_path == "auto" and n > 40000))
    Cd = C.toarray() if sp.issparse(C) and not sparse else C
    dense_path = args.dense_path if args.dense_path != "auto" else ("schur" if n <= 8000 else "hessenberg")
    hess = not sparse and not args.adjoint and dense_path == "hessenberg"
    t0 = time.perf_counter()
    pre_parts = {}
    T = H = Q = None
    if sparse:
        solver = R.SparseTriSolver(C)
        path = "sparse nnz=%d (host LU per refine)" % C.nnz
    elif args.adjoint:
        T, Q = R.schur_factor(Cd)     # host Schur with vectors (gradients in C's basis)
        solver = R.TriSolver(T)
        path = "dense (Schur with vectors)"
    elif hess:
        # pre_process on the device: Hessenberg form, neighbour-pivoted LU per refine shift
        solver = R.TriSolver.from_dense(Cd, timings=pre_parts)
        H = solver.operator()
        path = "dense (device Hessenberg, Hessenberg LU per refine)"
    else:
        # GPU Hessenberg reduction + host QR iterations (no Schur vectors needed)
        T = R.schur_preprocess(Cd, 0, timings=pre_parts)
        solver = R.TriSolver(T)
        path = "dense (Schur: device Hessenberg + host QR)"
    t_pre = time.perf_counter() - t0
    log(f"{path} n={n} preprocess {t_pre:.1f}s "
        f"{' '.join(f'{k} {v:.2f}s' for k, v in pre_parts.items())}, normC={normC:.6g}")
    from oracle import rls_oracle as O
    if sparse:
        oproblem = lambda: O.sparse_tri_problem(C)     # noqa: E731
    elif hess:
        oproblem = lambda: O.dense_hess_problem(H)     # noqa: E731
    else:
        oproblem = lambda: O.dense_tri_problem(T)      # noqa: E731
    if args.adjoint and not sparse:
        solver.set_schur_vectors(Q)
        solver.keep_vectors(True)
    ctl = R.PsTriRecyControl(tol=1e-4, normC=normC, p=60, recy_tau=1.6, mode="batched", progress=1,
                             window=args.window)
    nZ = len(rec.Z)
    tuned = None
    if not args.no_autotune and args.window == 0:
        # the window cap by measurement on this traversal (every cap gives the same
        # trajectory); the candidates' solves also warm the device up
        from rls_amd import autotune as AT
        ctl_q = dataclasses.replace(ctl, progress=0)
        tuned = AT.tune_window(solver, rec.Z, ctl_q, rec, use_cache=False)
        ctl = dataclasses.replace(ctl, window=tuned["window"])
        log(f"autotune window: {tuned['window']} ({', '.join(f'{k}: {v:.2f}s' for k, v in tuned['seconds'].items())}; "
            f"max rel diff {tuned['max_rel_diff']:.1e} at point {tuned['max_rel_diff_point']}, "
            f"decisions differ {tuned['decisions_differ']})")
        if tuned["max_rel_diff"] > 1e-8:
            # the candidates' sigma at the point where they differ most, against ground truth
            from oracle import truth
            i = tuned["max_rel_diff_point"]
            if sp.issparse(C):
                t_i = float(truth.sparse_sigma_min(C, rec.Z[i:i + 1])[0])
            else:
                import scipy.linalg as sla
                t_i = float(sla.svdvals(np.asarray(Cd) - rec.Z[i] * np.eye(n))[-1]) if n <= 6000 else None
            tuned["sigma_min_at_point"] = t_i
            if t_i:
                log("  at that point: " + ", ".join(f"cap {k}: rel err {abs(v - t_i) / t_i:.1e}"
                                                   for k, v in tuned["sigmas_at_point"].items()))

    for w in range(args.warmup):
        t0 = time.perf_counter()
        solver.solve(rec.Z, ctl, rec)
        log(f"warmup {w}: {time.perf_counter() - t0:.2f}s")
    times = []
    sig = st = None
    solver.enable_kernel_timing(True)
    for s in range(args.steps):
        t0 = time.perf_counter()
        sig, st = solver.solve(rec.Z, ctl, rec)
        dt = time.perf_counter() - t0
        times.append(dt)
        log(f"step {s}: {dt:.3f}s refined={int((st.lobpcg_it > 0).sum())} windows={st.n_windows}")
    total = sum(times)
    value = nZ * len(times) / total
    kt = solver.kernel_timing()
    solver.enable_kernel_timing(False)
    # end to end (preprocessing + one grid) for both dense paths: the Schur route pays host QR
    # iterations once, the Hessenberg route a device factor per refine shift
    e2e = None
    if not sparse and not args.adjoint and not args.no_compare_paths and n <= 8000:
        e2e = {("hessenberg" if hess else "schur"): {"preprocess_s": t_pre, "grid_s": total / len(times),
                                                      "pts_per_s": nZ / (t_pre + total / len(times))}}
        t0 = time.perf_counter()
        other_parts = {}
        if hess:
            T2 = R.schur_preprocess(Cd, 0, timings=other_parts)
            s2 = R.TriSolver(T2)
        else:
            s2 = R.TriSolver.from_dense(Cd, timings=other_parts)
        t_pre2 = time.perf_counter() - t0
        s2.solve(rec.Z, dataclasses.replace(ctl, progress=0), rec)    # warm-up of the second context
        t0 = time.perf_counter()
        _, st2 = s2.solve(rec.Z, dataclasses.replace(ctl, progress=0), rec)
        g2 = time.perf_counter() - t0
        e2e["schur" if hess else "hessenberg"] = {"preprocess_s": t_pre2, "preprocess_parts": other_parts,
                                                   "grid_s": g2, "pts_per_s": nZ / (t_pre2 + g2),
                                                   "refined_points": int((st2.lobpcg_it > 0).sum())}
        del s2
        log("end to end: " + ", ".join(f"{k} {v['preprocess_s']:.2f}s + {v['grid_s']:.2f}s = "
                                       f"{v['pts_per_s']:.0f} pts/s" for k, v in e2e.items()))
    log("kernel families: " + ", ".join(f"{k} {v['ms']:.0f} ms / {v['launches']} launches"
                                         for k, v in kt.items()))
    # roofline of the dominant kernel family (largest event-timed share of the step)
    dom = max(kt, key=lambda k: kt[k]["ms"])
    d = kt[dom]
    avg_s = d["ms"] / 1e3 / max(d["launches"], 1)
    if dom in ("trsm", "factor", "iter"):
        # "iter": a LOBPCG iteration replayed from its captured graph, priced by the bytes of
        # its two shifted solves (the triangle of the factor read once each)
        ach = d["bytes"] / max(d["launches"], 1) / avg_s / 1e9
        kname = {"trsm": "k_trsm_persistent", "factor": "k_hlu_factor",
                 "iter": "LOBPCG iteration graph (2 x k_trsm_persistent + small kernels)"}[dom]
        roof = {"kernel": kname, "bound": "hbm",
                "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0, "traffic": None}
    else:
        ach = d["flops"] / max(d["launches"], 1) / avg_s / 1e12
        roof = {"kernel": "k_zgemm_mfma" if dom == "gemm" else "k_heev_rr/k_heev_wave",
                "bound": "mfma", "achieved": ach, "peak": 78.6, "unit": "TFLOP/s",
                "frac": ach / 78.6, "traffic": None}
    # HBM bytes per launch of the dominant family, from the FETCH_SIZE / WRITE_SIZE counter
    # runs committed for this workload (profiles/traffic_<workload>.json), when available
    tpath = os.path.join(ROOT, "profiles", f"traffic_{args.workload}{n}.json")
    if dom == "heev" and os.path.exists(tpath):
        with open(tpath) as fh:
            tr_ = json.load(fh)
        roof["traffic"] = tr_["eigensolver_family_bytes_per_launch"]
        roof["traffic_GBps"] = roof["traffic"] / avg_s / 1e9
    roof["launches"] = d["launches"]
    roof["avg_launch_us"] = 1e6 * avg_s
    roof["share_of_step"] = d["ms"] / 1e3 / (total / len(times)) / len(times)

    adj = None
    if args.adjoint:
        t0 = time.perf_counter()
        dsdz = solver.sigma_grad()
        t_z = time.perf_counter() - t0
        t0 = time.perf_counter()
        dsdt = solver.grad_toeplitz() if args.workload == "toeplitz" else None
        t_t = time.perf_counter() - t0
        from oracle import truth
        errs = []
        for i in np.random.default_rng(2).choice(nZ, size=min(2, nZ), replace=False):
            _, u, v = truth.svd_triplet(np.asarray(C.toarray() if hasattr(C, "toarray") else C), rec.Z[i])
            uv = np.vdot(u, v)
            errs.append(float(np.abs(dsdz[i] - np.array([-uv.real, uv.imag])).max()))
        adj = {"points": nZ, "dsdz_s": t_z, "dsdt_s": t_t if dsdt is not None else None,
               "dsdz_max_abs_err_vs_svd": max(errs), "svd_samples": len(errs)}
        log(f"adjoint: dsdz {t_z:.2f}s, dsdt {t_t:.2f}s, |err| vs SVD vectors {max(errs):.2e}")

    # plain FP64 one-shift-at-a-time solver against the batched one on the same points:
    # (1) the leading points of the traversal (the recycle space still growing), and
    # (2) a mid-grid segment at steady state: both solvers resume from the state the batched
    #     solver reaches at the segment start (the trajectories of the two modes are identical)
    npl = min(args.plain_points, nZ)
    plain = None
    if npl > 0:
        ctl_plain = R.PsTriRecyControl(tol=1e-4, normC=normC, p=60, recy_tau=1.6, mode="plain", progress=1)
        t0 = time.perf_counter()
        sig_p, st_p = solver.solve(rec.Z[:npl], ctl_plain, rec)
        t_plain = time.perf_counter() - t0
        plain_pts = npl / t_plain
        t0 = time.perf_counter()
        sig_b, _ = solver.solve(rec.Z[:npl], ctl, rec)
        t_b = time.perf_counter() - t0
        log(f"plain FP64 {npl} pts {t_plain:.2f}s ({plain_pts:.0f} pts/s); batched same pts {t_b:.2f}s")
        plain_agree = float(np.nanmax(np.abs(sig_p - sig[:npl]) / sig[:npl]))
        plain = {"value": plain_pts, "points": npl, "batched_same_points": npl / t_b,
                 "speedup": (npl / t_b) / plain_pts, "max_rel_diff_vs_batched": plain_agree}
    mid = None
    nmid = min(args.mid_points, nZ // 2)
    if nmid > 0:
        P = nZ // 2
        seg = rec.Z[P:P + nmid]
        ctl_res_b = R.PsTriRecyControl(tol=1e-4, normC=normC, p=60, recy_tau=1.6, mode="batched",
                                       progress=1, window=args.window, resume=True)
        ctl_res_p = R.PsTriRecyControl(tol=1e-4, normC=normC, p=60, recy_tau=1.6, mode="plain",
                                       progress=1, resume=True)
        solver.solve(rec.Z[:P], ctl, rec)
        t0 = time.perf_counter()
        sig_mp, _ = solver.solve(seg, ctl_res_p, rec)
        t_mp = time.perf_counter() - t0
        solver.solve(rec.Z[:P], ctl, rec)
        t0 = time.perf_counter()
        sig_mb, st_mb = solver.solve(seg, ctl_res_b, rec)
        t_mb = time.perf_counter() - t0
        mid = {"first_point": int(P), "points": int(nmid), "plain_pts_per_s": nmid / t_mp,
               "batched_pts_per_s": nmid / t_mb, "speedup": t_mp / t_mb,
               "refined_points": int((st_mb.lobpcg_it > 0).sum()),
               "max_rel_diff": float(np.max(np.abs(sig_mp - sig_mb) / sig_mb)),
               "max_rel_diff_vs_full_grid": float(np.max(np.abs(sig_mb - sig[P:P + nmid]) / sig[P:P + nmid]))}
        log(f"mid-grid points {P}..{P + nmid}: plain {nmid / t_mp:.0f} pts/s, batched {nmid / t_mb:.0f} pts/s "
            f"({t_mp / t_mb:.1f}x)")

    # the reference's baseline, psInvLanc (inverse Lanczos per point, psInvLanc.jl:87-261), on
    # the device over the leading points, with the reference's default control
    inv = None
    ninv = min(args.invlanc_points, nZ)
    if ninv > 0:
        from rls_amd import invlanc as IL

        class _Points:
            Z = rec.Z[:ninv]
        ictl = IL.PsInvLancControl(tol=1e-4)
        t0 = time.perf_counter()
        if sparse:
            sig_i, st_i = IL.psInvLanc(C, _Points, ictl)
        else:
            sig_i, st_i = IL.psInvLanc(None, _Points, ictl, T=T, H=H)
        t_i = time.perf_counter() - t0
        inv = {"points": int(ninv), "pts_per_s": ninv / t_i, "mv_mean": float(st_i.mv.mean()),
               "converged": bool(st_i.converged.all()),
               "max_rel_diff_vs_batched": float(np.max(np.abs(sig_i - sig[:ninv]) / sig[:ninv])),
               "control": "tol=1e-4 krylovdim=20 maxiter=800, a new random start per point"}
        log(f"psInvLanc {ninv} points: {ninv / t_i:.1f} pts/s, {st_i.mv.mean():.1f} operator "
            f"applications per point")

    # accuracy against FP64 ground truth on sampled grid points: the full singular spectrum
    # (dense SVD) for n <= 2000, else block inverse iteration on the Schur factor.  Besides the
    # error against sigma_min, report the distance to the nearest singular value: the
    # psTriRecy acceptance test certifies a singular triplet, which in a tight cluster of
    # singular values need not be the smallest one.
    from oracle import truth
    idx = truth_sample(sig, args.truth_samples)
    t0 = time.perf_counter()
    rel_any = None
    if sp.issparse(C):
        tr = np.empty(len(idx))
        for j, i in enumerate(idx):   # a sparse factorization per point: one progress line each
            t1 = time.perf_counter()
            tr[j] = truth.sparse_sigma_min(C, rec.Z[i:i + 1])[0]
            log(f"truth point {int(i)}: sigma_min {tr[j]:.6e} ({time.perf_counter() - t1:.1f}s)")
    elif n <= 2000:
        import scipy.linalg as sla
        Ca = np.asarray(C)
        spectra = [sla.svdvals(Ca - z * np.eye(n)) for z in rec.Z[idx]]
        tr = np.array([sv[-1] for sv in spectra])
        rel_any = np.array([np.min(np.abs(sv - s_)) / s_ for sv, s_ in zip(spectra, sig[idx])])
    else:
        # inverse iteration on a Schur factor of C (the truth's own preprocessing when the
        # solver ran on the Hessenberg form)
        Tt = T if T is not None else R.schur_preprocess(Cd, 0)
        tr = truth.tri_sigma_min(Tt, rec.Z[idx], block=16)
    rel = np.abs(sig[idx] - tr) / tr
    worst = [{"point": int(idx[j]), "sigma": float(sig[idx[j]]), "sigma_min": float(tr[j]), "rel_err": float(rel[j]),
              "refined": bool(st.lobpcg_it[idx[j]] > 0)} for j in np.argsort(-rel)[:5]]
    log(f"truth on {len(idx)} points {time.perf_counter() - t0:.1f}s: max rel {rel.max():.2e}"
        + (f", to nearest singular value {rel_any.max():.2e}" if rel_any is not None else ""))

    acc_checks = None
    if args.acceptance_checks > 0:
        acc_checks = acceptance_checks(args, solver, ctl, sig, rec, oproblem, n, idx, rel, tr,
                                       (x_st, x_ed, y_st, y_ed, nx, ny, k), normC)
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_segments(args, solver, ctl, sig, st, rec, oproblem, n, sparse, normC,
                                    (x_st, x_ed, y_st, y_ed, nx, ny, k))

    out = {
        "metric": "sigma_min grid pts/sec at rel err<=1e-4: batched mixed vs plain FP64 on GPU",
        "value": value,
        "unit": "grid pts/s",
        "n_gpus": 1,
        "steps": len(times),
        "warmup": args.warmup,
        "ms_per_step": 1000.0 * total / len(times),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128",
        "data": "synthetic",
        "config": {"workload": f"{args.workload} {path} n={n}, "
                               f"grid {nx}x{ny} (k={k}), psTriRecy tol=1e-4 p=60 r=6 recy_tau=1.6, "
                               f"normC={normC:.6g}"
                               + (f", first {nZ} traversal points" if args.max_points > 0 else ""),
                   "grid_points": nZ},
        "max_rel_err": float(rel.max()),
        "max_rel_err_nearest_singular_value": None if rel_any is None else float(rel_any.max()),
        "rel_err_samples": int(len(idx)), "rel_err_worst": worst,
        "rel_err_above_1e-4": int((rel > 1e-4).sum()),
        "plain_fp64": None if plain is None else {**plain, "mid_grid": mid},
        "psInvLanc": inv,
        "refined_points": int((st.lobpcg_it > 0).sum()),
        "windows": int(st.n_windows),
        "window_shifts": int(st.n_window_shifts),
        "window_shifts_per_point": st.n_window_shifts / nZ,
        "host": host_info(),
        "preprocess_s": t_pre, "preprocess_parts": pre_parts or None, "autotune": tuned, "end_to_end": e2e,
        "refine_factor_s": getattr(solver, "factor_time_s", None),
        "refine_factors": getattr(solver, "factor_count", None),
        "cpu_baseline": cpu,
        "roofline": roof,
        "kernel_families": kt,
    }
    if adj is not None:
        out["adjoint"] = adj
    if acc_checks is not None:
        out["acceptance_checks"] = acc_checks
    print(json.dumps(out), flush=True)
    return 0


def acceptance_checks(args, solver, ctl, sig, rec, oproblem, n, idx, rel, tr, grid, normC, back=8):
    """For up to args.acceptance_checks sampled points i whose sigma is more than 1e-4 from
    sigma_min: the GPU state at i - back (the traversal resumed in pieces, one pass in total)
    continued in the oracle over points i - back .. i; reports the oracle's sigma at i beside
    the GPU's and sigma_min.  Equal values show the reference's acceptance rule certifying the
    same non-minimal singular value."""
    import rls_amd as R
    from oracle import rls_oracle as O
    miss = sorted(int(idx[j]) for j in np.argsort(-rel) if rel[j] > 1e-4)[: args.acceptance_checks]
    out = []
    pos = 0
    for i in miss:
        P = max(0, i - back)
        if P > pos:
            c = ctl if pos == 0 else dataclasses.replace(ctl, resume=True)
            solver.solve(rec.Z[pos:P], c, rec)
            pos = P
        st0 = solver.state() if pos > 0 else None
        orec = O.recursion_grid(*grid)
        orec.Z = orec.Z[P:i + 1]
        osig, ost, _ = O.ps_tri_recy(oproblem(), orec,
                                     O.PsTriRecyControl(tol=1e-4, normC=normC, p=60, recy_tau=1.6), state=st0,
                                     X_init=None if st0 else R.initial_block(n, ctl.r, ctl.seed))
        j = list(idx).index(i)
        out.append({"point": i, "gpu_sigma": float(sig[i]), "oracle_sigma": float(osig[-1]),
                    "sigma_min": float(tr[j]), "rel_gpu_vs_oracle": float(abs(sig[i] - osig[-1]) / osig[-1]),
                    "rel_vs_sigma_min": float(rel[j]), "oracle_refined": bool(ost.refined[-1])})
        log(f"acceptance check at {i}: gpu {sig[i]:.8f} oracle {osig[-1]:.8f} sigma_min {tr[j]:.8f}")
    return out


def truth_sample(sig, count, seed=1):
    """Grid points checked against ground truth: the 16 smallest computed sigma (points next
    to an eigenvalue, where the recycled sigma's relative accuracy is lowest), then half of
    the rest among the smallest 5 % and half uniformly at random."""
    nZ = len(sig)
    count = min(count, nZ)
    rng = np.random.default_rng(seed)
    order = np.argsort(sig)
    lowest = order[: min(16, count)]
    n_small = max(0, count // 2 - len(lowest))
    pool = np.setdiff1d(order[: max(n_small + len(lowest), nZ // 20)], lowest)
    small = rng.choice(pool, size=min(n_small, len(pool)), replace=False)
    rest = np.setdiff1d(np.arange(nZ), np.concatenate([lowest, small]))
    uni = rng.choice(rest, size=max(0, count - len(lowest) - len(small)), replace=False)
    return np.sort(np.concatenate([lowest, small, uni]))


def host_info():
    """CPU model, logical CPUs this process may use, and the BLAS thread setting."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except AttributeError:
        ncpu = os.cpu_count()
    blas = None
    try:
        from threadpoolctl import threadpool_info
        blas = [{"api": d.get("internal_api"), "threads": d.get("num_threads")} for d in threadpool_info()]
    except Exception:
        pass
    return {"cpu_model": model, "cpus_available": ncpu, "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "blas": blas}


def cpu_baseline_segments(args, solver, ctl, sig, st, rec, oproblem, n, sparse, normC, grid):
    """The numpy/OpenBLAS FP64 restatement (the oracle) timed on segments spread across the
    traversal.  A segment starting at point P > 0 continues from the state the GPU solver
    reaches there (rls_tri_state_get: recycle space, block, acceptance state), so the CPU
    runs the same warm trajectory; the first segment starts cold, as the grid does.  Each
    segment runs for cpu_seconds / segments; the sigma it computes is compared with the
    GPU's on the same points."""
    import rls_amd as R
    from oracle import rls_oracle as O
    nZ = len(rec.Z)
    nseg = max(1, args.cpu_segments)
    starts = sorted(set(int(f * nZ) for f in np.arange(nseg) / nseg))
    per = args.cpu_seconds / len(starts)
    octl = O.PsTriRecyControl(tol=1e-4, normC=normC, p=60, recy_tau=1.6)
    problem = oproblem()

    class Stop(Exception):
        pass

    segs, tot_pts, tot_s, max_rel, mism = [], 0, 0.0, 0.0, 0
    for P in starts:
        state = None
        if P > 0:
            if sparse:
                continue
            solver.solve(rec.Z[:P], ctl, rec)        # the GPU state at point P
            state = solver.state()
        orec = O.recursion_grid(*grid)
        orec.Z = orec.Z[P:]
        done = [0]
        t_start = time.perf_counter()

        def cb(i, *a):
            done[0] = i + 1
            if time.perf_counter() - t_start > per:
                raise Stop()
        out = {}

        def keep(i, z, what, *a):
            out[i] = (a[-1], what.startswith("refine"))
            cb(i)
        x0 = None if state is not None else R.initial_block(n, ctl.r, ctl.seed)
        try:
            O.ps_tri_recy(problem, orec, octl, on_point=keep, state=state, X_init=x0)
        except Stop:
            pass
        dt = time.perf_counter() - t_start
        m = done[0]
        tot_pts += m
        tot_s += dt
        if m:
            csig = np.array([out[i][0] for i in range(m)])
            cref = np.array([out[i][1] for i in range(m)])
            rel = float(np.max(np.abs(csig - sig[P:P + m]) / sig[P:P + m]))
            mm = int((cref != (st.lobpcg_it[P:P + m] > 0)).sum())
            max_rel = max(max_rel, rel)
            mism += mm
            segs.append({"first_point": P, "points": m, "seconds": dt, "pts_per_s": m / dt,
                         "refined": int(cref.sum()), "max_rel_diff_vs_gpu": rel, "decisions_differ": mm})
        log(f"cpu segment at {P}: {m} points in {dt:.1f}s ({m / max(dt, 1e-9):.1f} pts/s)")
    hi = host_info()
    threads = None
    for b in hi.get("blas") or []:
        threads = max(threads or 0, b.get("threads") or 0)
    return {"value": tot_pts / tot_s, "unit": "grid pts/s", "cores": threads or hi["cpus_available"],
            "kind": "port",
            "sample": f"{tot_pts} points in {len(segs)} segments starting at "
                      f"{[sg['first_point'] for sg in segs]} of the same zigzag traversal (mid-grid "
                      f"segments continue from the GPU solver's state there), numpy/OpenBLAS FP64 "
                      f"restatement, {tot_s:.1f}s",
            "segments": segs, "max_rel_diff_vs_gpu": max_rel, "decisions_differ": mism,
            "cpu_model": hi["cpu_model"], "cpus_available": hi["cpus_available"]}


def cpu_baseline_prec(args, A, rec, ctl, perm, sig, st):
    """The oracle's psPrecRecy (numpy/scipy FP64 restatement) on the leading points of the same
    traversal, from the same start blocks, with the threshold IC built by the library's host
    C++ routine (the restated one takes minutes per build at this size), for cpu_seconds;
    its sigma is compared with the GPU's on the same points."""
    import scipy.sparse as sp
    import rls_amd as R
    from oracle import rls_ic as OI
    from oracle import rls_prec as OP
    from rls_amd import prec as RP

    def ic_builder(C, z, p, c):
        f = RP.ic_build(RP.normal_lower(C, z), p, c.hsl_tol, c.hsl_cntl_1, c.hsl_cntl_2, c.hsl_icntl_1,
                        c.hsl_icntl_2)
        L = sp.csc_matrix(f.to_scipy())
        invp = np.empty(len(p), np.int64)
        invp[p] = np.arange(len(p))
        return OI.SparseLLTP(L, np.asarray(p), invp, f.alpha, f.n_shifts)

    octl = OP.PsPrecRecyControl(**{k: getattr(ctl, k) for k in OP.PsPrecRecyControl.__dataclass_fields__
                                   if hasattr(ctl, k)})
    seed_base = ctl.seed * 1000003
    rb = lambda nn, rr, tag: R.initial_block(nn, rr, seed_base + tag)  # noqa: E731
    out = {}
    t_start = time.perf_counter()

    class Stop(Exception):
        pass

    def cb(i, z, what, s_):
        out[i] = (s_, "main(" in what)
        log(f"cpu baseline point {i}: {what} ({time.perf_counter() - t_start:.1f}s)")
        if time.perf_counter() - t_start > args.cpu_seconds:
            raise Stop()
    try:
        OP.ps_prec_recy(A, rec, octl, perm=perm, rand_block=rb, on_point=cb, ic_builder=ic_builder)
    except Stop:
        pass
    dt = time.perf_counter() - t_start
    m = len(out)
    csig = np.array([out[i][0] for i in range(m)])
    rel = float(np.max(np.abs(csig - sig[:m]) / sig[:m])) if m else None
    hi = host_info()
    threads = max([b.get("threads") or 0 for b in (hi.get("blas") or [])] or [hi["cpus_available"]])
    log(f"cpu baseline (oracle psPrecRecy) {m} points in {dt:.1f}s; max rel diff vs GPU {rel}")
    return {"value": m / dt, "unit": "grid pts/s", "cores": threads, "kind": "port",
            "sample": f"first {m} points of the same zigzag traversal from the same start blocks, numpy/scipy "
                      f"FP64 restatement of psPrecRecy with the threshold IC from the library's host C++ "
                      f"build, {dt:.1f}s",
            "max_rel_diff_vs_gpu": rel, "cpu_model": hi["cpu_model"], "cpus_available": hi["cpus_available"]}


def bench_prec(args):
    """psPrecRecy on the C2 configuration (example/skewlap3d/skewlap3d_psPrecRecy.jl:16-45)
    over the leading points of the 200 x 200 traversal."""
    import rls_amd as R
    from rls_amd import prec as RP
    from rls_amd import workloads as WL
    N = 26                                    # (N-1)^3 = 15625 unknowns
    A = WL.skewlap3d(N)
    n = A.shape[0]
    normC = WL.norm2_estimate(A)
    g = WL.GRIDS["skewlap3d"]
    rec = R.recursion_grid(*g)
    npts = args.max_points if args.max_points > 0 else 2000
    rec.Z = rec.Z[:npts]
    ctl = RP.PsPrecRecyControl(
        tol=1e-4, main_normC=normC, main_tau=1.6, gap_tau=1e-10, tol_pred=1e-3, pred_tau=1.6,
        idx_cutoff_pred=0.01, pred_orth_type="shiftchol", pred_conv_thresh_type=1,
        lambda_lower_min=0.01, lambda_upper_max=30.0, r_main=6, p_main=60, p_min_main=1,
        main_implict_update=False, main_implict_recycle=True, r_pred=40, p_pred=20, p_min_pred=1,
        pred_implict_update=True, pred_implict_recycle=True, hsl_tol=1e-3, hsl_cntl_1=1e-6,
        hsl_icntl_1=200, hsl_icntl_2=500, hsl_cntl_2=1e-4, gamma_tol=1e-3, lambda_n_tol=1e-3)
    # the C++ driver prints a progress line every ~5 % of the traversal
    os.environ.setdefault("RLS_PROGRESS", "1")
    t0 = time.perf_counter()
    perm = RP.column_ordering(A)
    solver = RP.PrecSolver(A)
    t_pre = time.perf_counter() - t0
    log(f"skewlap3d n={n} nnz={A.nnz} normC={normC:.6g}; ordering + device setup {t_pre:.1f}s")
    for w in range(args.warmup):
        RP.psPrecRecy(A, rec, ctl, perm=perm, solver=solver)
    tuned = None
    if not args.no_autotune and args.warmup > 0:
        # chunk widths of the IC solves by measurement on the factor the warm-up left
        from rls_amd import autotune as AT
        from rls_amd import _capi
        tuned = AT.tune_sparse_chunks(solver.ctx, n, _capi.RLS_OP_KINV, use_cache=False)
        log(f"autotune IC-solve chunk widths: k<=8 -> {tuned['small']}, k>8 -> {tuned['large']} "
            f"(k=6 ms {tuned['ms_k6']}, k=40 ms {tuned['ms_k40']}, identical outputs {tuned['identical_outputs']})")
    solver.enable_kernel_timing(True)
    times = []
    sig = st = None
    for s_ in range(args.steps):
        t0 = time.perf_counter()
        sig, st = RP.psPrecRecy(A, rec, ctl, perm=perm, solver=solver)
        times.append(time.perf_counter() - t0)
        log(f"step {s_}: {times[-1]:.2f}s main refines={st.n_pts} pred refines={st.n_pred_pts} "
            f"preconditioner builds={st.n_precond_rebuild_pts}")
    kt = solver.kernel_timing()
    solver.enable_kernel_timing(False)
    total = sum(times)
    value = npts * len(times) / total
    d = kt["trsm"]
    avg_s = d["ms"] / 1e3 / max(d["launches"], 1)
    ach = d["bytes"] / max(d["launches"], 1) / avg_s / 1e9 if d["launches"] else None
    roof = {"kernel": "k_sptrsv (IC factor L / L^H)", "bound": "hbm", "achieved": ach, "peak": 8000.0,
            "unit": "GB/s", "frac": None if ach is None else ach / 8000.0, "traffic": None,
            "launches": d["launches"], "avg_launch_us": 1e6 * avg_s,
            "share_of_step": d["ms"] / 1e3 / total}
    # HBM bytes per IC solve from the FETCH_SIZE / WRITE_SIZE counter runs committed for this
    # workload, when available
    tpath = os.path.join(ROOT, "profiles", f"traffic_skewlap3d{n}.json")
    if os.path.exists(tpath) and d["launches"]:
        with open(tpath) as fh:
            tr_ = json.load(fh)
        roof["traffic"] = tr_["sptrsv_bytes_per_launch"]
        roof["traffic_GBps"] = roof["traffic"] / avg_s / 1e9
    from oracle import truth
    # a sparse factorization per sample (~1.6 s each at this size): 24 samples unless asked
    n_truth = args.truth_samples if args.truth_samples != TRUTH_SAMPLES_DEFAULT else 24
    idx = np.random.default_rng(1).choice(npts, size=min(n_truth, npts), replace=False)
    tr = np.empty(len(idx))
    for j, i in enumerate(idx):       # a sparse factorization per point: one progress line each
        t0 = time.perf_counter()
        tr[j] = truth.sparse_sigma_min(A, rec.Z[i:i + 1])[0]
        log(f"truth point {int(i)}: sigma_min {tr[j]:.6e} ({time.perf_counter() - t0:.1f}s)")
    rel = np.abs(sig[idx] - tr) / tr
    log(f"truth on {len(idx)} points: max rel {rel.max():.2e}; trsm {d['ms']:.0f} ms / {d['launches']} launches")
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_prec(args, A, rec, ctl, perm, sig, st)
    out = {
        "metric": "sigma_min grid pts/sec at rel err<=1e-4: batched mixed vs plain FP64 on GPU",
        "value": value, "unit": "grid pts/s", "n_gpus": 1, "steps": len(times), "warmup": args.warmup,
        "ms_per_step": 1000.0 * total / len(times), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c128", "data": "synthetic",
        "config": {"workload": f"skewlap3d n={n} (3-D skew Laplacian, N=26), psPrecRecy tol=1e-4, "
                               f"r_main=6 p_main=60 r_pred=40 p_pred=20, IC hsl_tol=1e-3 lsize=200 "
                               f"rsize=500, first {npts} of the 200x200 traversal",
                   "grid_points": npts},
        "max_rel_err": float(rel.max()), "rel_err_samples": int(len(idx)),
        "main_refines": int(st.n_pts), "pred_refines": int(st.n_pred_pts), "autotune": tuned,
        "preconditioner_builds": int(st.n_precond_rebuild_pts),
        "stage_seconds": {"precond": st.precond_time_ns / 1e9, "main_recycle": st.main_recycle_time_ns / 1e9,
                          "main_refine": st.main_refine_time_ns / 1e9, "main_update": st.main_update_time_ns / 1e9,
                          "pred_recycle": st.pred_recycle_time_ns / 1e9, "pred_refine": st.pred_refine_time_ns / 1e9,
                          "pred_update": st.pred_update_time_ns / 1e9},
        "preprocess_s": t_pre, "cpu_baseline": cpu, "roofline": roof, "kernel_families": kt,
        "host": host_info(),
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
