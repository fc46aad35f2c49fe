"""Benchmark: time to relative residual 1e-8 of two-level BDDC + PCG on the
128^3 Q1 Poisson cube (2,146,689 dofs), 512 cubic subdomains of 16^3
elements, seeded uniform[-1,1] right-hand side, on one MI355X.

A step = bddcml_setup_preconditioner + bddcml_solve (the time-to-solution of
the reference harness, examples/poisson_on_cube.f90:520-560).  Subdomain data
is uploaded through bddcml_upload_subdomain_data before each step, outside
the timed region.  value = dofs / (mean seconds per step).

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "time-to-rel-residual 1e-8 (mixed vs FP64 GPU vs host CPU), dofs/s, HBM GB/s"
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline():
    """Oracle (CPU FP64 restatement) on a bounded sample of the same
    workload: 5x5x5 subdomains of 16^3 elements (80^3 mesh, 531,441 dofs),
    the subdomain size of the 128^3 case, solved to 1e-8 (10-20 s of CPU
    work on the host cores)."""
    from oracle import oracle as orc
    from bddcml_amd.cube import CubeProblem
    orc.build()
    p = CubeProblem(16, 5, rhs="random", seed=2)
    subs = list(p.subdomains())
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = orc.Oracle(p.nsub, p.nnod, p.ndof, 3, 3, p.numbase)
    for d in subs:
        o.upload(d)
    t0 = time.perf_counter()
    o.setup(1, 1, 0)
    it, reason, cond = o.solve(0, 1e-8, 500, 50)
    dt = time.perf_counter() - t0
    o.close()
    return dict(value=p.ndof / dt, unit="dofs/s", cores=threads, kind="port",
                sample=f"oracle setup+PCG to 1e-8 on 80^3 Q1 Poisson, 125 subdomains of 16^3 "
                       f"({p.ndof} dofs, {it} iterations, {dt:.2f} s, OpenMP over subdomains); "
                       "the oracle forms each Schur complement explicitly (skyline Cholesky of "
                       "the interior in natural order, one solve per interface column) and "
                       "factors the augmented matrix by dense LU: not the reference's default "
                       "implicit-Schur path with nested-dissection MUMPS factors, which cannot "
                       "be built here")


def profiled_traffic(args, kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary of the default workload (null for other
    workloads or when the summary is absent)."""
    if (args.workload, args.nel, args.nsub, args.ordering, args.nd_leaf, args.levels,
            args.explicit_inverse) != ("cube", 16, 8, 0, 512, 2, 2):
        return None
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "hbm_traffic_128cube_pmc.md")
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(f"traffic_bytes_per_launch {kernel} ="):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def gradient_workload(args):
    """-div(k grad u) = f on 64^3 Q1 elements, 512 subdomains of 8^3, k per
    element log-normal (mu 0, sigma 1, seed 5), J = 0.5 |u - u_t|^2 (u_t
    uniform[-1,1], seed 6).  Timed: forward setup + solve, adjoint solve on
    the same factorization, element gradient kernel.  Checked: dJ/dk against
    central finite differences (h = 1e-6 k_e) on 50 seeded elements."""
    from bddcml_amd.adjoint import AdjointSolver
    from bddcml_amd.cube import CubeProblem, lognormal, uniform01
    nel, ns = 8, 8
    p0 = CubeProblem(nel, ns)
    k = lognormal(5, np.arange(p0.nelem))
    ut = 2.0 * uniform01(6, np.arange(p0.ndof, dtype=np.uint64)) - 1.0
    make = lambda kk: CubeProblem(nel, ns, rhs="random", seed=5, coef=lambda e: kk[e])
    sol = AdjointSolver(make, p0.kref, tol=1e-12)
    times = []
    for step in range(args.warmup + args.steps):
        u = sol.forward(k)
        tm = sol.b.timings()
        tf = tm["setup_total"] + tm["solve_total"]  # setup_preconditioner + solve
        t1 = time.perf_counter()
        lam = sol.adjoint(u - ut)
        t2 = time.perf_counter()
        g = sol.gradient(u, lam)
        t3 = time.perf_counter()
        log(f"step {step}: forward setup+solve {tf:.3f} s, adjoint (new data + solve) "
            f"{t2 - t1:.3f} s, gradient kernel {t3 - t2:.4f} s, iterations {sol.iterations[-2:]}")
        if step >= args.warmup:
            times.append((tf, t2 - t1, t3 - t2))
    J = lambda uu: 0.5 * np.sum((uu - ut) ** 2)
    rng_e = (uniform01(7, np.arange(50, dtype=np.uint64)) * p0.nelem).astype(int)
    errs = []
    for e in rng_e:
        h = 1e-6 * k[e]
        kp = k.copy(); kp[e] += h
        km = k.copy(); km[e] -= h
        fd = (J(sol.forward(kp)) - J(sol.forward(km))) / (2 * h)
        errs.append(abs(fd - g[e]) / np.abs(g).max())
    sol.close()
    tf, ta, tg = (float(np.mean([t[i] for t in times])) for i in range(3))
    ms = 1000.0 * (tf + ta + tg)
    line = {
        "metric": METRIC, "value": p0.ndof / (tf + ta + tg), "unit": "dofs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "adjoint gradient: -div(k grad u)=f, 64^3 Q1, 512 subdomains of "
                               "8^3, log-normal k (seed 5), J=0.5|u-u_t|^2 (seed 6)",
                   "dofs": p0.ndof, "elements": p0.nelem,
                   "forward_ms": round(1000 * tf, 2), "adjoint_ms": round(1000 * ta, 2),
                   "gradient_kernel_ms": round(1000 * tg, 3),
                   "fd_check_elements": 50, "fd_max_rel_error": float(max(errs))},
        "note": "value: dofs / (forward setup+solve + adjoint solve on the same factorization "
                "+ element gradient kernel)",
    }
    print(json.dumps(line), flush=True)


def adaptive_workload(args):
    """Adaptive constraints: 3D elasticity (E=1, nu=0.3) on an n^3 Q1 mesh
    with 27 seeded stiff layers (E jump 1e6) split into seeded random boxes;
    face eigenproblems by batched LOBPCG (block 10, tolerance 1e-6, threshold
    10), stiffness weights, 2 levels, PCG to 1e-8.  Timed per step:
    setup_preconditioner (arithmetic setup, LOBPCG, second interface phase)
    + solve; the same problem without adaptive constraints is solved once
    for the condition number it starts from."""
    import bddcml_amd
    from bddcml_amd.cube import BoxProblem
    p = BoxProblem(args.box_n, args.box_nsub, seed=3, kind="elasticity", layers=27, jump=1e6)
    subs = list(p.subdomains())
    sizes = [d["ndofs"] for d in subs]
    log(f"adaptive workload: {p.nsub} subdomains ({min(sizes)}..{max(sizes)} dofs), {p.ndof} dofs")

    def one(adaptive):
        b = bddcml_amd.Bddcml()
        b.init(2, [p.nsub, 1], numbase=p.numbase,
               verbose_level=int(os.environ.get("BENCH_VERBOSE", "0")))
        for k, v in (("adaptive_threshold", 10.0), ("adaptive_tol", 1e-6), ("adaptive_maxit", 15),
                     ("adaptive_nvec", 10), ("explicit_inverse", args.explicit_inverse)):
            b.set_option(k, v)
        for d in subs:
            b.upload_subdomain_data(d)
        t0 = time.perf_counter()
        b.setup_preconditioner(1, 0, 1, 1, 1, adaptive, 0, 1)
        t1 = time.perf_counter()
        it, reason, cond = b.solve(0, args.tol, 3000, 300)
        t2 = time.perf_counter()
        pairs = b.adaptive_pairs() if adaptive else []
        ndofc = b.counts()["ndofc"]
        b.finalize()
        if reason != 0:
            raise SystemExit(f"PCG did not converge: reason {reason}")
        return t1 - t0, t2 - t1, it, cond, pairs, ndofc

    s0, v0, it0, c0, _, nc0 = one(0)
    log(f"arithmetic constraints: setup {s0:.3f} s solve {v0:.3f} s, {it0} iterations, "
        f"cond {c0:.4g}, {nc0} coarse dofs")
    times = []
    for step in range(args.warmup + args.steps):
        ts, tv, it, cond, pairs, ndofc = one(1)
        nad = sum(q["nadaptive"] for q in pairs)
        lit = max((q["iterations"] for q in pairs), default=0)
        log(f"step {step}: setup {ts:.3f} s solve {tv:.3f} s, {it} iterations, cond {cond:.4g}, "
            f"{len(pairs)} face pairs, {nad} adaptive constraints, LOBPCG iterations <= {lit}, "
            f"{ndofc} coarse dofs")
        if step >= args.warmup:
            times.append((ts, tv))
    ts, tv = (float(np.mean([t[i] for t in times])) for i in range(2))
    line = {
        "metric": METRIC, "value": p.ndof / (ts + tv), "unit": "dofs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": 1000.0 * (ts + tv),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"adaptive constraints: Q1 elasticity {p.nel}^3, 27 seeded layers "
                               f"E=1e6, {p.nsub} random boxes, LOBPCG block 10 tol 1e-6 threshold "
                               "10, weights type 1, 2 levels, PCG to "
                               f"{args.tol:g}",
                   "dofs": p.ndof, "subdomains": p.nsub,
                   "subdomain_dofs": [min(sizes), max(sizes)],
                   "setup_ms": round(1000 * ts, 1), "solve_ms": round(1000 * tv, 1),
                   "iterations": it, "condition_number": cond, "face_pairs": len(pairs),
                   "adaptive_constraints": nad, "coarse_dofs": ndofc,
                   "without_adaptivity": {"iterations": it0, "condition_number": c0,
                                          "setup_ms": round(1000 * s0, 1),
                                          "solve_ms": round(1000 * v0, 1)}},
    }
    print(json.dumps(line), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="cube", choices=("cube", "boxes", "gradient", "adaptive"),
                    help="cube: Q1 Poisson on a cube split into equal cubes (the reported "
                         "workload); boxes: Q1 elasticity with stiff layers on uneven "
                         "subdomains from seeded random box splits")
    ap.add_argument("--box-n", type=int, default=96, help="boxes: elements per cube edge")
    ap.add_argument("--box-nsub", type=int, default=400, help="boxes: number of subdomains")
    ap.add_argument("--weights", type=int, default=0, choices=(0, 1, 3),
                    help="interface weights: 0 cardinality, 1 diagonal stiffness, 3 unit")
    ap.add_argument("--nel", type=int, default=16, help="elements per subdomain edge")
    ap.add_argument("--nsub", type=int, default=8, help="subdomains per cube edge")
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--levels", type=int, default=2, choices=(2, 3, 4, 5),
                    help="3 or more: a second level of --nsub2 subdomains replaces the coarse "
                         "solve, each further level a quarter of the one before (at least 2)")
    ap.add_argument("--nsub2", type=int, default=8, help="subdomains of the second level")
    ap.add_argument("--explicit-inverse", type=int, default=2, choices=(0, 1, 2),
                    help="explicit inverse of S' for the local preconditioner: 0 never, 1 at "
                         "setup, 2 after 40 applies (the library's default)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--set", action="append", default=[], metavar="KEY=VALUE",
                    help="further library option (bddcml_amd_set_option), repeatable")
    ap.add_argument("--pcg-graph", type=int, default=1, choices=(0, 1),
                    help="PCG iterations from the third on replayed from one captured graph")
    ap.add_argument("--ordering", type=int, default=0,
                    help="interior ordering: 0 tile-aligned nested dissection, 1 RCM")
    ap.add_argument("--precision", default="all", choices=("all", "both", "64", "32", "refined"),
                    help="64: FP64 throughout; 32: FP32 factor of the augmented interface "
                         "matrices (mixed precision); refined: FP32 interior factorization, "
                         "Schur complements and S' factor with FP64 iterative refinement; "
                         "both: 64 and 32; all: 64, 32 and refined (FP64 is the reported "
                         "value, the others are added to the line)")
    ap.add_argument("--nd-leaf", type=int, default=512,
                    help="nested dissection: largest part (dofs) left whole")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); every subdomain of the "
              "level is a batched slot on that GPU, multi-GPU sharding is not built.",
              file=sys.stderr)
        sys.exit(2)

    import bddcml_amd
    from bddcml_amd.cube import CubeProblem

    # plain HBM copy of known size: the reference for the HBM byte counters
    # of a profile of this run, and the bandwidth a copy reaches on this GPU
    copy_bytes = 1 << 30
    copy_ms = bddcml_amd.selftest_copy(copy_bytes)
    copy_gbs = 2 * copy_bytes / (copy_ms * 1e-3) / 1e9
    log(f"k_copy_calibration: {copy_bytes} bytes read + written in {copy_ms:.3f} ms "
        f"({copy_gbs:.0f} GB/s)")

    if args.workload == "gradient":
        return gradient_workload(args)
    if args.workload == "adaptive":
        return adaptive_workload(args)
    nsub_next = [args.nsub2]
    for _ in range(args.levels - 3):
        nsub_next.append(max(2, nsub_next[-1] // 4))
    levels_txt = ("2 levels" if args.levels == 2 else
                  f"{args.levels} levels ({', '.join(map(str, nsub_next))} subdomains on "
                  f"levels 2..{args.levels - 1})")
    if args.explicit_inverse != 2:
        levels_txt += (", explicit inverse of S' at setup" if args.explicit_inverse == 1 else
                       ", no explicit inverse of S'")
    t0 = time.perf_counter()
    if args.workload == "cube":
        p = CubeProblem(args.nel, args.nsub, rhs="random", seed=2)
        workload = (f"Q1 Poisson {p.nel}^3 unit cube, {p.nsub} subdomains of {args.nel}^3 "
                    f"elements, {levels_txt}, PCG to {args.tol:g}")
    else:
        from bddcml_amd.cube import BoxProblem
        p = BoxProblem(args.box_n, args.box_nsub, seed=3, kind="elasticity", layers=27,
                       jump=1e4)
        workload = (f"Q1 elasticity (E=1, nu=0.3) {p.nel}^3 unit cube, 27 seeded layers E=1e4, "
                    f"{p.nsub} subdomains from seeded random box splits, {levels_txt}, weights "
                    f"type {args.weights}, PCG to {args.tol:g}")
    subs = list(p.subdomains())
    sizes = [d["ndofs"] for d in subs]
    log(f"generated {p.nsub} subdomains ({min(sizes)}..{max(sizes)} dofs), {p.ndof} dofs in "
        f"{time.perf_counter() - t0:.1f} s")

    def run(mode):
        """warmup + timed steps of one precision mode (64, 32 or "refined");
        each step is a fresh init + upload (untimed) and setup_preconditioner
        + solve (timed)"""
        precision = 64 if mode == 64 else 32
        tag = "refined" if mode == "refined" else f"fp{precision}"
        step_times, results, tims = [], [], []
        for step in range(args.warmup + args.steps):
            b = bddcml_amd.Bddcml()
            nsublev = [p.nsub, 1] if args.levels == 2 else [p.nsub] + nsub_next + [1]
            b.init(args.levels, nsublev, numbase=p.numbase,
                   verbose_level=int(os.environ.get("BENCH_VERBOSE", "0")))
            b.set_option("interior_ordering", args.ordering)
            b.set_option("nd_leaf", args.nd_leaf)
            b.set_option("precision", precision)
            b.set_option("refinement", 1 if mode == "refined" else 0)
            b.set_option("explicit_inverse", args.explicit_inverse)
            b.set_option("pcg_graph", args.pcg_graph)
            for kv in args.set:
                k, v = kv.split("=")
                if b.set_option(k, float(v)) != 0:
                    raise SystemExit(f"unknown option or value: {kv}")
            t0 = time.perf_counter()
            for d in subs:
                b.upload_subdomain_data(d)
            t_up = time.perf_counter() - t0
            t0 = time.perf_counter()
            b.setup_preconditioner(1, 0, 1, 1, 1, 0, 0, args.weights)
            t1 = time.perf_counter()
            it, reason, cond = b.solve(0, args.tol, 2000, 200)
            t2 = time.perf_counter()
            tm = b.timings()
            hist = b.residual_history()
            if step == args.warmup + args.steps - 1 and mode != "refined":
                # outside the timed region: the true residual of the interface
                # problem, |g - S u| / |g|, with the converged solution
                ui = b.interface_solution()
                g = b.reduced_rhs()
                tm["true_relres"] = float(np.linalg.norm(g - b.apply_schur(ui)) / np.linalg.norm(g))
                log(f"[{tag}] true interface residual |g - S u|/|g| = "
                    f"{tm['true_relres']:.3e}")
            elif mode == "refined":
                # the refinement stops on the FP64 interface residual of the
                # exact Schur complements (the library's own S is FP32 here)
                tm["true_relres"] = tm["refine_relres"]
            b.finalize()
            if reason != 0:
                raise SystemExit(f"PCG did not converge: reason {reason}")
            log(f"[{tag}] step {step}: upload {t_up:.2f} s setup {t1 - t0:.3f} s "
                f"solve {t2 - t1:.3f} s iters {it} cond {cond:.4f} final relres {hist[-1]:.3e}")
            log("  " + json.dumps({k: round(v, 4) for k, v in tm.items()}))
            if step >= args.warmup:
                step_times.append(t2 - t0)
                results.append((it, cond, hist[-1]))
                tims.append(tm)
        ms = 1000.0 * float(np.mean(step_times))
        return ms, results[-1], tims[-1]

    modes = {"all": [64, 32, "refined"], "both": [64, 32], "64": [64], "32": [32],
             "refined": ["refined"]}[args.precision]
    runs = {m: run(m) for m in modes}
    head = modes[0]
    ms, (iters, cond, relres), tm = runs[head]
    # the interior factorization's tiles: FP64, or FP32 in the refined form
    gtype, gpeak = ("float", PEAK_FP32_TFLOPS) if head == "refined" else ("double", PEAK_FP64_TFLOPS)
    value = p.ndof / (ms / 1000.0)
    schur_tf = tm["schur_flops"] / (tm["ev_schur_ms"] * 1e-3) / 1e12
    gemv_gbs = tm["gemv_bytes"] / (tm["ev_gemv_ms"] * 1e-3) / 1e9 if tm["ev_gemv_ms"] > 0 else 0.0
    interior_tf = tm["interior_flops"] / (tm["ev_interior_ms"] * 1e-3) / 1e12
    kapply_gbs = tm["kapply_bytes"] / (tm["ev_kapply_ms"] * 1e-3) / 1e9 if tm["ev_kapply_ms"] > 0 else 0.0
    log(f"local preconditioner apply {tm['ev_kapply_ms']:.3f} ms ({kapply_gbs:.0f} GB/s of "
        f"{tm['kapply_bytes'] / 1e9:.2f} GB)")
    log(f"interior factorisation {tm['ev_interior_ms']:.1f} ms ({interior_tf:.1f} TFLOP/s), "
        f"Schur block GEMM {tm['ev_schur_ms']:.1f} ms ({schur_tf:.1f} TFLOP/s), "
        f"interface {tm['ev_interface_ms']:.1f} ms, coarse {tm['ev_coarse_ms']:.1f} ms, "
        f"S GEMV {tm['ev_gemv_ms']:.3f} ms ({gemv_gbs:.0f} GB/s)")
    # dominant kernel: the GEMM launches of the interior factorization
    # (k_tile_gemm<double>), flops from their task lists over their summed
    # event time, i.e. average flops per launch / average launch duration
    gemm_tf = tm["gemm_flops"] / (tm["ev_gemm_ms"] * 1e-3) / 1e12 if tm["ev_gemm_ms"] > 0 else 0.0
    nl = max(int(tm["gemm_launches"]), 1)
    log(f"interior GEMM launches (k_tile_gemm<{gtype}>): {nl} launches, "
        f"{tm['gemm_flops'] / nl:.4e} flops and {tm['ev_gemm_ms'] / nl:.3f} ms per launch on "
        f"average ({gemm_tf:.1f} TFLOP/s)")
    schur_kernel = "k_schur_gemm" if head != "refined" else "k_schur_gemm_f32"
    roofline = dict(bound="mfma", achieved=round(gemm_tf, 3), peak=gpeak,
                    unit="TFLOP/s", frac=round(gemm_tf / gpeak, 4),
                    traffic=profiled_traffic(args, f"k_tile_gemm<{gtype}> interior"),
                    kernel=f"k_tile_gemm<{gtype}> (interior factorization launches)",
                    launches=nl,
                    flops_per_launch=tm["gemm_flops"] / nl,
                    ms_per_launch=round(tm["ev_gemm_ms"] / nl, 4),
                    others={schur_kernel: {"achieved": round(schur_tf, 3),
                                           "frac": round(schur_tf / gpeak, 4),
                                           "flops_per_launch": tm["schur_flops"],
                                           "ms_per_launch": round(tm["ev_schur_ms"], 3),
                                           "traffic": profiled_traffic(args, schur_kernel)},
                            f"k_symv<{gtype}>": {"achieved": round(gemv_gbs, 1), "unit": "GB/s",
                                                 "frac": round(gemv_gbs / PEAK_HBM_GBS, 4),
                                                 "bytes_per_launch": tm["gemv_bytes"]}})
    if head == "refined" and tm["ev_resid_ms"] > 0:
        rg = tm["resid_bytes"] / (tm["ev_resid_ms"] * 1e-3) / 1e9
        roofline["others"]["k_resid_rows"] = {"achieved": round(rg, 1), "unit": "GB/s",
                                              "frac": round(rg / PEAK_HBM_GBS, 4),
                                              "bytes_per_launch": tm["resid_bytes"],
                                              "ms_per_launch": round(tm["ev_resid_ms"], 4)}
    line = {
        "metric": METRIC, "value": value, "unit": "dofs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": {64: "f64", 32: "f64+f32", "refined": "f32+f64"}[head],
        "data": "synthetic",
        "config": {"workload": workload,
                   "dofs": p.ndof, "subdomains": p.nsub, "iterations": iters,
                   "subdomain_dofs": [min(sizes), max(sizes)],
                   "condition_number": round(cond, 4), "final_relres": relres,
                   "precision": {64: "fp64",
                                 32: "mixed: FP32 interface factor, FP64 Schur operator/Krylov/"
                                     "coarse",
                                 "refined": "refined: FP32 interior factorization, Schur "
                                            "complements and interface factor, FP64 iterative "
                                            "refinement"}[head],
                   "s_gemv_GBps": round(gemv_gbs, 1), "hbm_copy_GBps": round(copy_gbs, 1),
                   "pc_apply_GBps": round(kapply_gbs, 1),
                   "true_interface_relres": tm.get("true_relres")},
        "roofline": roofline,
    }
    if head == 64 and 32 in runs:
        ms32, (it32, c32, rr32), tm32 = runs[32]
        line["mixed"] = {"value": p.ndof / (ms32 / 1000.0), "unit": "dofs/s", "ms_per_step": ms32,
                         "dtype": "f64+f32", "iterations": it32, "final_relres": rr32,
                         "true_interface_relres": tm32.get("true_relres"),
                         "interface_phase_ms": round(tm32["ev_interface_ms"], 2),
                         "pc_apply_ms": round(tm32["ev_kapply_ms"], 3),
                         "note": "FP32 factor of S' = S + C^T W C, its sweeps and Z; "
                                 "Schur operator, Krylov and coarse problem FP64"}
        log(f"mixed precision: {ms32:.1f} ms/step ({p.ndof / (ms32 / 1000.0):.3e} dofs/s), "
            f"{it32} iterations, final relres {rr32:.3e}")
    if head == 64 and "refined" in runs:
        msr, (itr, cr, rrr), tmr = runs["refined"]
        gemm32 = (tmr["gemm_flops"] / (tmr["ev_gemm_ms"] * 1e-3) / 1e12
                  if tmr["ev_gemm_ms"] > 0 else 0.0)
        schur32 = tmr["schur_flops"] / (tmr["ev_schur_ms"] * 1e-3) / 1e12
        resid_gbs = (tmr["resid_bytes"] / (tmr["ev_resid_ms"] * 1e-3) / 1e9
                     if tmr["ev_resid_ms"] > 0 else 0.0)
        factor64 = tm["ev_interior_ms"] + tm["ev_schur_ms"]
        factor32 = tmr["ev_interior_ms"] + tmr["ev_schur_ms"]
        line["refined"] = {
            "value": p.ndof / (msr / 1000.0), "unit": "dofs/s", "ms_per_step": msr,
            "dtype": "f32+f64", "iterations": itr, "corrections": int(tmr["refinements"]),
            "final_interface_relres": tmr["refine_relres"],
            "interior_factorization_ms": round(tmr["ev_interior_ms"], 2),
            "schur_gemm_ms": round(tmr["ev_schur_ms"], 2),
            "factorization_ms_fp64_path": round(factor64, 2),
            "factorization_speedup_vs_fp64": round(factor64 / factor32, 3) if factor32 > 0 else None,
            "interface_phase_ms": round(tmr["ev_interface_ms"], 2),
            "krylov_s": round(tmr["krylov"], 4), "refinement_s": round(tmr["refine"], 4),
            "postprocess_s": round(tmr["postprocess"], 4),
            "interior_gemm_TFLOPs_f32": round(gemm32, 2),
            "interior_gemm_frac_f32": round(gemm32 / PEAK_FP32_TFLOPS, 4),
            "schur_gemm_TFLOPs_f32": round(schur32, 2),
            "fp64_residual_ms": round(tmr["ev_resid_ms"], 4),
            "fp64_residual_GBps": round(resid_gbs, 1),
            "note": "FP32 interior factorization, Schur complements, S' factor and sweeps; "
                    "FP64 residuals of the subdomain triplets, Krylov, coarse problem"}
        log(f"refined FP32: {msr:.1f} ms/step ({p.ndof / (msr / 1000.0):.3e} dofs/s), {itr} "
            f"iterations, {int(tmr['refinements'])} corrections, interface residual "
            f"{tmr['refine_relres']:.3e}; interior {tmr['ev_interior_ms']:.1f} ms (GEMM "
            f"{gemm32:.1f} TFLOP/s), Schur GEMM {tmr['ev_schur_ms']:.1f} ms ({schur32:.1f} "
            f"TFLOP/s), FP64 residual {tmr['ev_resid_ms']:.3f} ms ({resid_gbs:.0f} GB/s)")
    if not args.no_cpu_baseline and args.workload == "cube":
        t0 = time.perf_counter()
        line["cpu_baseline"] = cpu_baseline()
        log(f"cpu baseline took {time.perf_counter() - t0:.1f} s")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
