#!/usr/bin/env python3
"""Benchmark of the EFIE hot path on one MI355X: fill + LU + 361-RHS solve at N = 12160.

Workload: PEC plate 3 x 3 wavelengths at 300 MHz, 64 x 64 squares split into right triangles
(T = 8192, N = 12160 RWG unknowns), Dunavant degree 5, 361 monostatic plane waves
(theta = 0..90 deg in 0.25 deg steps, phi = 0, E_theta = 1) solved as one block.

A step is one pass of the path: impedance fill (complex128, quadrature mode), plane-wave RHS
block, LU, triangular solves, and -- in the headline form -- complex64 LU (without pivoting; the
pivoted LU takes over if a pivot is zero or the refinement stalls) with FP64-residual iterative
refinement to max_c ||b_c - Z x_c|| / ||b_c|| <= 1e-10.  The plain complex128 form is
timed the same way.  value = F / t_step with F = 8N^3/3 + 8N^2 nrhs (LAPACK flop convention).

Prints one JSON line.  Multi-GPU runs are not part of this build: --gpus > 1 exits.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FREQ = 300e6
NRHS = 361
DEGREE = 5
PEAK_FP32_MFMA = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_FP64_MFMA = 78.6   # TFLOP/s, MI355X dense FP64 matrix
PEAK_HBM = 8000.0       # GB/s
# HBM bytes per timed trailing-update launch of cgemm_sub<float> (K = 128) at N = 12160, nb = 128
# (average over the 188 launches of one factorisation), from rocprofv3 FETCH_SIZE / WRITE_SIZE
# runs scaled by a copy kernel of known size (profiles/plate_N12160_c64np_hbm_traffic.md for the
# default form, 699.7e6; plate_N12160_c64ir_hbm_traffic.md for the pivoted one, 700.1e6);
# algorithmic bytes per launch: 419.3e6.
GEMM_F32_TRAFFIC = {"c64_np_ir": 699.7e6, "c64_ir": 700.1e6}
# HBM bytes per timed launch of cgemm_sub_f16 at N = 12160, nb = 128, averaged over the 235 launches of
# one factorisation in paired steps (rocprofv3 FETCH_SIZE / WRITE_SIZE, scaled by a copy kernel of
# known size: profiles/plate_N12160_c64np_hbm_traffic.md); algorithmic bytes per launch 176.7e6.
GEMM_F16_TRAFFIC = 237.6e6
# HBM bytes of one quadrature-mode fill at N = 12160 (tripair_f64 + assemble + near_f64; same profile).
FILL_TRAFFIC = 31.1e9
PEAK_FP16_MFMA = 2500.0  # TFLOP/s, MI355X dense FP16 matrix


def lu_flops(n, nrhs):
    return 8.0 * n ** 3 / 3.0 + 8.0 * n ** 2 * nrhs


def run_step(ctx, form, th, tol, maxit, mode="quad"):
    ctx.fill(FREQ, mode, "c128")
    ctx.rhs_planewave(FREQ, th, 0.0, 1.0, 0.0)
    ctx.factor(form)
    return ctx.solve(tol=tol, maxit=maxit)


def timed(ctx, form, th, steps, tol, maxit, mode="quad"):
    ctx.set_profiling(True)
    ctx.synchronize()
    t0 = time.perf_counter()
    infos = [run_step(ctx, form, th, tol, maxit, mode) for _ in range(steps)]
    ctx.synchronize()
    t1 = time.perf_counter()
    prof = ctx.profile()
    ctx.set_profiling(False)
    return (t1 - t0) / steps, infos, prof


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(threads, n_target):
    """The reference's CPU path restated: the oracle's OpenMP fill (compute_matrix's loop structure,
    src/emw_solver.cpp:45-142), its plane-wave RHS block, and LAPACK zgetrf/zgetrs through scipy
    (solve_general, src/lapack_wrappers.cpp:53-133).  Timed on plates of the headline's mesh density
    (64 cells per 3 wavelengths) with N ~ 1000, 2000 and 4000 and 361 incidences; t_fill = a N^2,
    t_factor = b N^3 and t_solve = c N^2 fitted by least squares and extrapolated to the headline's
    N, whose algorithmic flops divided by that time give the value."""
    import oracle
    import scipy.linalg as sla
    from frico_amd import meshes
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    lam = meshes.wavelength(FREQ)
    th = np.radians(np.arange(NRHS) * 0.25)
    rows = []
    for n in (18, 26, 37):
        sm = meshes.plate(3.0 * lam * n / 64.0, n)
        t0 = time.perf_counter()
        om = oracle.OracleMesh(sm)
        Z = om.fill(FREQ, DEGREE)
        B = np.stack([om.rhs_planewave(FREQ, t, 0.0, 1.0, 0.0, DEGREE) for t in th], axis=1)
        t1 = time.perf_counter()
        lu = sla.lu_factor(Z, overwrite_a=True, check_finite=False)
        t2 = time.perf_counter()
        sla.lu_solve(lu, B, check_finite=False)
        t3 = time.perf_counter()
        rows.append((om.N, t1 - t0, t2 - t1, t3 - t2))
    Ns = np.array([r[0] for r in rows], float)
    fit = lambda t, p: float(np.dot(Ns ** p, t) / np.dot(Ns ** p, Ns ** p))  # least squares t = c N^p
    a, b, c = fit([r[1] for r in rows], 2), fit([r[2] for r in rows], 3), fit([r[3] for r in rows], 2)
    t_target = a * n_target ** 2 + b * n_target ** 3 + c * n_target ** 2
    meas = ", ".join(f"N={r[0]}: fill {r[1]:.2f} s, zgetrf {r[2]:.2f} s, zgetrs {r[3]:.2f} s" for r in rows)
    return {"value": lu_flops(n_target, NRHS) / t_target / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "extrapolated": True, "seconds_at_target": t_target, "cpu": _cpu_model(), "nproc": os.cpu_count(),
            "lapack": "scipy (bundled OpenBLAS) zgetrf/zgetrs",
            "sample": f"plates of the headline's mesh density, {NRHS} RHS: oracle fill (C++ OpenMP) + scipy LAPACK; "
                      f"{meas}; fitted a N^2 + b N^3 + c N^2 extrapolated to N={n_target}: {t_target:.1f} s"}


def batch_bench(args):
    """Uneven batch: 32 strip dipoles / loops with N ~ U[500, 4000] at 16 frequencies from 100 to
    500 MHz, complex64 fill, complex64 LU + FP64 refinement to 1e-10, graphs and autotuning on.
    One timed pass after an untimed one; value = systems per second."""
    # concurrent contexts need hardware queues of their own (read when HIP initialises; the
    # environment's own setting, HIP's default of 4, would serialise more than four workers)
    os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(4, args.workers)))
    import frico_amd
    from frico_amd import meshes
    ms = meshes.batch_workload()
    fr = meshes.batch_frequencies()
    with frico_amd.Batch(0, workers=args.workers) as b:
        for sm in ms:
            b.add_mesh(sm)
        b.antenna(fr, fill="c64", form=args.batch_form, tol=1e-10)  # warm-up: allocation, tuning table
        t0 = time.perf_counter()
        z, info = b.antenna(fr, fill="c64", form=args.batch_form, tol=1e-10)
        t = time.perf_counter() - t0
        nb, ngraph = b.block_sizes()
        used = b.workers_used()
    N = []
    with frico_amd.Context(0) as ctx:
        for sm in ms:
            ctx.load_mesh(sm, 5)
            N.append(ctx.N)
    F = sum(8.0 * n ** 3 / 3.0 + 8.0 * n ** 2 for n in N) * len(fr)
    nsys = len(ms) * len(fr)
    ok = all(i.converged and i.max_rel_residual <= 1e-10 for row in info for i in row)
    tf = F / t / 1e12
    out = {"metric": "uneven batch systems/s (fill + LU + refinement + Zin)", "value": nsys / t,
           "unit": "systems/s", "n_gpus": 1, "higher_is_better": True, "seconds": t,
           "tflops": tf, "dtype": "c64 fill + c64 LU + c128 residual", "data": "synthetic",
           # the whole batch's LU flops (8N^3/3 + 8N^2 a system) over its wall time, against the
           # dense FP32 MFMA peak (the complex64 LU's arithmetic class)
           "roofline": {"bound": "mfma", "achieved": tf, "peak": PEAK_FP32_MFMA, "unit": "TFLOP/s",
                        "frac": tf / PEAK_FP32_MFMA, "traffic": None},
           "config": {"workload": f"{len(ms)} strip dipoles/loops, N {min(N)}..{max(N)}, {len(fr)} frequencies "
                                  f"100..500 MHz", "workers": used, "hw_queues": os.environ["GPU_MAX_HW_QUEUES"],
                      "graphs": True, "form": args.batch_form,
                      "block_widths": sorted(set(nb.tolist())), "graph_meshes": ngraph},
           "all_converged": ok}
    print(json.dumps(out))


def approx_bench(args):
    """Approximate-mode plate (the same N = 12160 mesh and 361 incidences; centroid fill of
    compute_matrix_approx), solved as run_context solves it, with solve_symmetric: the complex64
    L D L^T + FP64 refinement, beside the pivoted complex64 LU + refinement and the complex128 LU of
    the same system.  value = (8N^3/3 + 8N^2 nrhs) / t_step, the LU's flop count, so that the
    forms compare by time."""
    import frico_amd
    from frico_amd import meshes
    sm = meshes.workload_plate(FREQ, args.n)
    th = np.radians(np.arange(NRHS) * 0.25)
    ctx = frico_amd.Context(0)
    ctx.set_block(args.nb)
    ctx.load_mesh(sm, DEGREE)
    N, T = ctx.N, ctx.T
    F = lu_flops(N, NRHS)
    res = {}
    sols = {}
    for form in ("c64_sym_ir", "c64_ir", "c128"):
        for _ in range(args.warmup):
            run_step(ctx, form, th, args.tol, 10, "approx")
        t, infos, prof = timed(ctx, form, th, args.steps, args.tol, 10, "approx")
        sols[form] = ctx.solution()
        res[form] = {"ms_per_step": t * 1e3, "tflops": F / t / 1e12, "factored_form": ctx.factored_form,
                     "iterations": infos[-1].iterations, "max_rel_residual": infos[-1].max_rel_residual,
                     "phases_ms": {k: prof[k] / args.steps
                                   for k in ("fill_ms", "rhs_ms", "factor_ms", "solve_ms", "refine_ms")},
                     "gemm_tflops": prof["gemm_flops"] / max(prof["gemm_ms"], 1e-9) / 1e9}
        print(f"# approx {form} {t * 1e3:.2f} ms/step {infos[-1]} form {ctx.factored_form}", file=sys.stderr,
              flush=True)
    ref = sols["c128"]
    for form in ("c64_sym_ir", "c64_ir"):
        res[form]["solution_rel_err_vs_c128"] = float(np.max(np.abs(sols[form] - ref)) / np.max(np.abs(ref)))
    sym = res["c64_sym_ir"]
    out = {"metric": "N=12k approximate-mode fill+LDL^T+361-RHS time, TFLOP/s (LU flop count) at 1e-10 residual",
           "value": sym["tflops"], "unit": "TFLOP/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": sym["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
           "dtype": "c64 LDL^T + c128 fill/residual", "data": "synthetic (generated plate mesh, analytic plane waves)",
           "config": {"workload": f"PEC plate 3x3 lambda @300 MHz, {args.n}x{args.n} squares, N={N}, T={T}, "
                                  f"{NRHS} plane waves, approximate-mode fill", "N": N, "nrhs": NRHS,
                      "tol": args.tol, "block": args.nb},
           "forms": res,
           "speedup_sym_vs_lu": res["c64_ir"]["ms_per_step"] / sym["ms_per_step"],
           "speedup_sym_vs_c128": res["c128"]["ms_per_step"] / sym["ms_per_step"]}
    ctx.close()
    print(json.dumps(out))


def nearfield_bench(args):
    """Near-field sampling plane on the N = 12160 plate: eval_fields(mesh) (one current vector, E
    and H at every vertex of a -P plane) after one normal-incidence solve.  The plane is
    "xz:3l:3l:h" (y = 0, across the plate), h = lambda/100: 301 x 301 = 90601 points, 7.4e8
    point-triangle evaluations.  value = evaluations per second over frico_fields calls (host
    arrays in and out, as the reference's caller gets them)."""
    import frico_amd
    from frico_amd import meshes
    sm = meshes.workload_plate(FREQ, args.n)
    lam = meshes.wavelength(FREQ)
    ctx = frico_amd.Context(0)
    ctx.load_mesh(sm, DEGREE)
    N, T = ctx.N, ctx.T
    ctx.fill(FREQ, "quad", "c128")
    ctx.rhs_planewave(FREQ, [0.0], [0.0], 1.0, 0.0)
    ctx.factor("c64_ir")
    ctx.solve(tol=args.tol)
    spec = f"xz:{3 * lam}:{3 * lam}:{lam / 100}"
    sp = frico_amd.parse_sampling_plane(spec)
    V, _ = frico_amd.sampling_plane_mesh(sp)
    P = V.shape[0]
    for _ in range(args.warmup):
        ctx.fields(FREQ, V)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        ctx.fields(FREQ, V)
    ctx.synchronize()
    t = (time.perf_counter() - t0) / args.steps
    out = {"metric": "near-field plane point-triangle evaluations per second (eval_fields on a -P plane)",
           "value": P * T / t, "unit": "evals/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (generated plate mesh, its normal-incidence currents)",
           "config": {"workload": f"PEC plate 3x3 lambda, N={N}, T={T}; plane {spec}", "points": P, "triangles": T}}
    if not args.no_cpu:
        import oracle
        om = oracle.OracleMesh(sm)
        I = ctx.solution()[:, 0]
        AJ, AdivJ, _ = om.scatter(I)
        sample = V[:: max(1, P // 4000)]
        t0 = time.perf_counter()
        om.fields(FREQ, AJ, AdivJ, sample)
        tc = time.perf_counter() - t0
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = {"value": sample.shape[0] * T / tc, "unit": "evals/s", "cores": threads, "kind": "port",
                               "sample": f"{sample.shape[0]} of the plane's points, oracle eval_fields (C++ OpenMP), "
                                         f"{tc:.2f} s"}
    ctx.close()
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--nb", type=int, default=128)
    ap.add_argument("--n", type=int, default=64, help="plate cells per side (64 -> N=12160)")
    ap.add_argument("--tol", type=float, default=1e-10)
    ap.add_argument("--form", default="c64_np_ir", choices=["c64_ir", "c64_np_ir"],
                    help="complex64 factorisation of the headline form: LU without pivoting (the "
                         "default; falls back to the pivoted LU on a zero pivot or a stalled "
                         "refinement), or the pivoted LU")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--workload", choices=["plate", "batch", "approx", "nearfield"], default="plate",
                    help="plate: the N=12160 headline; batch: the uneven 32-mesh x 16-frequency sweep; "
                         "approx: the same plate in approximate mode with the symmetric solve; "
                         "nearfield: E and H on a 90601-point sampling plane over that plate")
    ap.add_argument("--workers", type=int, default=4,
                    help="batch workload: concurrent contexts (lockstep units: 725 systems/s at 4, 694 at 8)")
    ap.add_argument("--batch-form", default="c64_np_ir", choices=["c64_ir", "c64_np_ir"],
                    help="batch workload: factorisation form")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              "(multi-GPU sharding is design-only)", file=sys.stderr)
        sys.exit(2)

    if args.workload == "batch":
        batch_bench(args)
        return
    if args.workload == "approx":
        approx_bench(args)
        return
    if args.workload == "nearfield":
        nearfield_bench(args)
        return

    import frico_amd
    from frico_amd import meshes

    sm = meshes.workload_plate(FREQ, args.n)
    th = np.radians(np.arange(NRHS) * 0.25)
    ctx = frico_amd.Context(0)
    ctx.set_block(args.nb)
    ctx.load_mesh(sm, DEGREE)
    N, T = ctx.N, ctx.T
    F = lu_flops(N, NRHS)
    print(f"# N={N} T={T} nrhs={NRHS} F={F:.4e}", file=sys.stderr, flush=True)

    # ---- headline form: complex64 LU + FP64 refinement -------------------------------------
    for _ in range(args.warmup):
        run_step(ctx, args.form, th, args.tol, 10)
    print(f"# warmup done ({args.form})", file=sys.stderr, flush=True)
    t_ir, infos_ir, prof_ir = timed(ctx, args.form, th, args.steps, args.tol, 10)
    form_used = ctx.factored_form
    X_ir = ctx.solution()
    # a fresh exact FP64 residual B - Z X of the timed answer (untimed): the refinement's own figure
    # after its first sweep is the carried residual updated by fp32 products of the corrections
    true_res = float(np.max(ctx.residual()))
    print(f"# {args.form} {t_ir * 1e3:.2f} ms/step {infos_ir[-1]} (factored form {form_used}), "
          f"fresh residual {true_res:.3e}", file=sys.stderr, flush=True)

    # ---- error against the FP64 oracle's answer (tests/golden fixture), untimed ---------------
    def oracle_error(form):
        fx = os.path.join(ROOT, "tests", "golden", "plate64_c3_monostatic.npz")
        if (args.n, args.nb) != (64, 128) or not os.path.exists(fx):
            return None
        g = np.load(fx)
        sig, inf = ctx.monostatic_sweep([FREQ], g["theta"], 0.0, form=form, tol=args.tol, R=float(g["R"]))
        X = ctx.solution()
        return {"sigma_max_rel": float(np.max(np.abs(sig[0] - g["sigma"])) / np.max(np.abs(g["sigma"]))),
                "solution_sample_max_rel": float(np.max(np.abs(X[g["x_rows"], :] - g["x_vals"])) /
                                                 np.max(np.abs(g["x_vals"]))),
                "max_rel_residual": inf[0].max_rel_residual,
                "fixture": "tests/golden/plate64_c3_monostatic.npz (oracle FP64 fill + LAPACK zgetrf/zgetrs)"}
    err_ir = oracle_error(args.form)
    print(f"# {args.form} vs oracle: {err_ir}", file=sys.stderr, flush=True)

    # ---- plain complex128 form ---------------------------------------------------------------
    fp64 = None
    if not args.no_fp64:
        for _ in range(args.warmup):
            run_step(ctx, "c128", th, args.tol, 10)
        t_64, _, prof_64 = timed(ctx, "c128", th, args.steps, args.tol, 10)
        X_64 = ctx.solution()
        # residual of the complex128 solution against a fresh Z (untimed)
        ctx.fill(FREQ, "quad", "c128")
        res64 = float(np.max(ctx.residual()))
        err = float(np.max(np.abs(X_ir - X_64)) / np.max(np.abs(X_64)))
        err_64 = oracle_error("c128")
        fp64 = {"ms_per_step": t_64 * 1e3, "tflops": F / t_64 / 1e12, "max_rel_residual": res64,
                "error_vs_oracle": err_64,
                "phases_ms": {k: prof_64[k] / args.steps for k in ("fill_ms", "rhs_ms", "factor_ms", "solve_ms")},
                "gemm_tflops": prof_64["gemm_flops"] / max(prof_64["gemm_ms"], 1e-9) / 1e9,
                "c64_ir_solution_rel_err_vs_c128": err}
        print(f"# c128 {t_64 * 1e3:.2f} ms/step res {res64:.2e}; c64_ir vs c128 err {err:.2e}",
              file=sys.stderr, flush=True)

    gemm_tf = prof_ir["gemm_flops"] / max(prof_ir["gemm_ms"], 1e-9) / 1e9
    gemm_gbs = prof_ir["gemm_bytes"] / max(prof_ir["gemm_ms"], 1e-9) / 1e6
    gemm_avg_ms = prof_ir["gemm_ms"] / max(prof_ir["gemm_launches"], 1)
    f16 = args.form == "c64_np_ir" and os.environ.get("FRICO_GEMM_F16", "1") != "0"
    if f16:
        # The trailing products on fp16 MFMA: a K = 128 update reads and writes C once per 3072 fp16
        # flops an entry (192 flop/B, under the MI355X ridge of 312: HBM-bound); the paired steps'
        # K = 256 updates do twice the flops per C byte (384 flop/B, just over the ridge).  The line
        # gives the HBM view over all launches and, beside it, the same launches' fp16 MFMA rate
        # (three split-term products of four real MFMAs per complex multiply-add: 3 x 8 MNK).
        roofline = {"bound": "hbm", "kernel": "cgemm_sub_f16 (LU trailing update, fp16 MFMA on split operands)",
                    "achieved": gemm_gbs, "peak": PEAK_HBM, "unit": "GB/s", "frac": gemm_gbs / PEAK_HBM,
                    "traffic": GEMM_F16_TRAFFIC if (args.n, args.nb) == (64, 128) else None,
                    "algorithmic_bytes_per_launch": prof_ir["gemm_bytes"] / max(prof_ir["gemm_launches"], 1),
                    "launches": prof_ir["gemm_launches"], "avg_launch_ms": gemm_avg_ms,
                    "complex_tflops": gemm_tf, "fp16_mfma_tflops": 3.0 * gemm_tf,
                    "fp16_mfma_peak_tflops": PEAK_FP16_MFMA, "fp16_mfma_frac": 3.0 * gemm_tf / PEAK_FP16_MFMA}
    else:
        roofline = {"bound": "mfma", "kernel": "cgemm_sub<float> (LU trailing update)",
                    "achieved": gemm_tf, "peak": PEAK_FP32_MFMA, "unit": "TFLOP/s",
                    "frac": gemm_tf / PEAK_FP32_MFMA,
                    "traffic": GEMM_F32_TRAFFIC.get(args.form) if (args.n, args.nb) == (64, 128) else None,
                    "launches": prof_ir["gemm_launches"], "avg_launch_ms": gemm_avg_ms}
    fill_ms = prof_ir["fill_ms"] / args.steps
    out = {
        "metric": "N=12k fill+LU+361-RHS time, TFLOP/s: FP32+IR vs FP64 GPU at 1e-10 residual",
        "value": F / t_ir / 1e12,
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_ir * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("c64 LU (trailing products on fp16 MFMA, split operands) + c128 fill/residual" if f16
                  else "c64 LU + c128 fill/residual"),
        "data": "synthetic (generated plate mesh, analytic plane-wave excitations)",
        "config": {"workload": f"PEC plate 3x3 lambda @300 MHz, {args.n}x{args.n} squares, N={N}, T={T}, "
                               f"{NRHS} plane waves theta 0..90 deg step 0.25, Dunavant degree {DEGREE}",
                   "N": N, "nrhs": NRHS,
                   "form": ("complex64 LU without pivoting + FP64 iterative refinement"
                            if args.form == "c64_np_ir" else "complex64 LU (partial pivoting) + FP64 iterative refinement"),
                   "factored_form": {0: "c128", 1: "c64 LU, partial pivoting", 2: "c64 LDL^T",
                                     3: "c64 LU, no pivoting"}.get(form_used, str(form_used)),
                   "tol": args.tol, "block": args.nb},
        "roofline": roofline,
        "refinement": {"iterations": infos_ir[-1].iterations, "max_rel_residual": infos_ir[-1].max_rel_residual,
                       "fresh_fp64_max_rel_residual": true_res},
        "error_vs_oracle": err_ir,
        "phases_ms": {k: prof_ir[k] / args.steps for k in ("fill_ms", "rhs_ms", "factor_ms", "solve_ms", "refine_ms")},
        "fill": {"ms": fill_ms, "z_write_GBps": 16.0 * N * N / (fill_ms * 1e-3) / 1e9,
                 "tripair_ms": prof_ir["tripair_ms"] / args.steps,
                 "geval_per_s": 49.0 * T * T / max(prof_ir["tripair_ms"] / args.steps * 1e-3, 1e-12),
                 # measured HBM bytes of the fill (the triangle-pair moments' round trip through HBM)
                 "hbm_traffic_GB": FILL_TRAFFIC / 1e9 if (args.n, args.nb) == (64, 128) else None,
                 "hbm_frac": (FILL_TRAFFIC / (fill_ms * 1e-3) / 1e9 / PEAK_HBM) if (args.n, args.nb) == (64, 128)
                 else None},
        "fp64_form": fp64,
        "speedup_vs_fp64": (fp64["ms_per_step"] / (t_ir * 1e3)) if fp64 else None,
    }
    if not args.no_cpu:
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        out["cpu_baseline"] = cpu_baseline(threads, N)
    ctx.close()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
