#!/usr/bin/env python3
"""Benchmark of the numeric multifrontal LDLᵀ on one MI355X: factor + solve to a backward error of 1e-10, mixed
precision against the plain FP64 GPU factor.

The mixed path is the learned precision selector's choice of a cut of the supernode tree's levels for this matrix
(made once, untimed, after the FP64 leg whose pivot growth it reads): FP32 fronts on MFMA below the cut and FP64
above, refined in FP64 by conjugate gradients; FP64 fronts throughout when it predicts no cut pays (also when no
trained weights are present), FP32 throughout when the best cut is the top. The FP32-fronts path is also timed on its
own ("fp32_fronts_cg"). The roofline is that of the dominant kernel of the FP32 leg when the choice is FP32 fronts,
else of the FP64 leg.

Workload: the 500 × 500 quad4 shell plate (1,494,006 DOF), interior nodes jittered ±20 % of the element size, element
thickness log-uniform in [1, 4] mm, E = 70 GPa, ν = 0.3, clamped edges, 16 load cases of independent N(0, 1) forces on
the translational DOFs (rotations 0); nested-dissection order and symbolic factor on the host (untimed), the matrix
and the right-hand sides resident in device memory before the timed region.

A step is one factorisation and one refined solve of the 16 right-hand sides. One JSON line on stdout; progress on
stderr.

    python bench.py [--steps K] [--warmup W] [--scale 500] [--nrhs 16]
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "factor+solve s at 1e-10 backward err: mixed FP32+IR vs FP64 GPU; % MFMA peak"
PEAK_FP64 = 78.6   # TFLOP/s, MI355X dense FP64 matrix
PEAK_FP32 = 157.3  # TFLOP/s, MI355X dense FP32 matrix
PEAK_HBM = 8000.0  # GB/s, MI355X HBM3E
# The HIP kernel behind each profile kind, as a rocprofv3 trace names it ({T}: double or float).
KIND_KERNELS = {"small_fronts": "smallFrontKernel<{T}>", "assembly": "assembleKernel<{T}>",
                "diagonal_blocks": "diagonalBlockKernel<{T}>",
                "panel_solve": "updateTilesKernel<{T},false,false>",
                "panel_update_tiles": "updateTilesKernel<{T},false,false>",
                "panel_update_fused_tiles": "updateTilesKernel<{T},false,true>",
                "update_matrix_tiles": "extendTilesKernel<{T},64>",
                "update_matrix_big_tiles": "extendTilesKernel<{T},128>",
                "forward_triangles": "forwardKernel<{T}>", "backward_triangles": "backwardKernel<{T}>",
                "residual": "multiplyRowsKernel", "forward_rows": "forwardRowsKernel<{T}>",
                "backward_columns": "backwardColumnsKernel<{T}>", "triangle_products": "triangleProductsKernel<{T}>",
                "triangle_inverses": "triangleInverseKernel<{T}>"}
TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "hbm_traffic_shell_plate_c2.txt")
BYTE_KINDS = ("assembly", "forward_triangles", "backward_triangles", "residual", "forward_rows", "backward_columns",
              "triangle_products")


def log(message):
    print(f"[bench {time.strftime('%H:%M:%S')}] {message}", file=sys.stderr, flush=True)


def cpu_baseline(workload, b, threads, keep_solution=False):
    """The reference's own FP64 factor (AVX2 kernel family, the family it selects itself) and solve on the host
    cores, on the same symbolic structure and matrix: oracle/_ref, timed as one factorisation + one solve of the
    16 right-hand sides. Returns the cpu_baseline object, or None when the compiled reference is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    if _oracle.reference() is None:
        return None
    import numpy as np
    h, sizes, flops, arr = _oracle.ref_symbolic(workload.graph_offsets, workload.adjacency, workload.weights,
                                                order=workload.symbolic.node_order.copy())
    try:
        assert np.array_equal(arr["rows"], workload.symbolic.rows_all), "the reference's structure differs"
        x = np.array(b, order="F", copy=True)
        rc, msg, inertia, counts, times, _, _ = _oracle.ref_factor_solve(h, workload.matrix, b=x, kernel=1,
                                                                         threads=threads)
    finally:
        _oracle.reference().ref_sym_free(h)
    if rc != 0:
        raise RuntimeError("reference factorisation failed: " + msg)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    total = float(times[0] + times[1])
    extra = {"x": x} if keep_solution else {}
    return {**extra, "value": round(total, 4), "unit": "s", "cores": threads, "kind": "reference",
            "factor_s": round(float(times[0]), 4), "solve_s": round(float(times[1]), 4),
            "gflops_factor": round(flops / times[0] / 1e9, 1), "cpu": model,
            "hardware_concurrency": int(_oracle.reference().ref_hardware_threads()),
            "sample": "the whole workload: one FP64 factorisation (SupernodalFactor::factorise, avx2 family, "
                      f"{threads} threads) and one solve of the {b.shape[1]} right-hand sides (FactorSolve::solve)"
                      + ("; value is the median of 3 runs" if keep_solution else "")}


def reference_order_flops(workload, threads):
    """The factorisation flops of the workload under the reference's own nested-dissection order
    (GeometricDissection, through oracle/_ref), beside this library's order: the orders differ, so the structures do
    too. None when the compiled reference is absent."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle
    if _oracle.reference() is None:
        return None
    h, _, flops, _ = _oracle.ref_symbolic(workload.graph_offsets, workload.adjacency, workload.weights,
                                          coords=workload.coordinates, method=2, threads=threads)
    _oracle.reference().ref_sym_free(h)
    return float(flops)


def cylinder_main(args):
    """The buckling shifts of the stiffened cylinder (360 × 200 quad4 skin, 24 stringers, 429,840 DOF): the first 20
    buckling factors by bisection on GPU Sturm counts, then 8 shifts across them, each an indefinite FP64
    factorisation (Bunch–Kaufman) and a solve on the GPU; the CPU baseline is the reference's indefinite factor of
    the same 8 matrices on the host cores, whose Sturm counts must equal the GPU's."""
    import numpy as np
    import torch

    from o_o_solver_amd import buckling
    from o_o_solver_amd.pencil import combine, scale_symmetric

    if not torch.cuda.is_available():
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)
    record = buckling.run(360, 200, 20, 8, 1e-7, log=log)
    b = record.pop("buckling")
    w = b.workload
    shifts = record["shifts"]
    baseline = None
    if args.cpu_baseline == "auto":
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        import _oracle
        if _oracle.reference() is not None:
            threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
            h, _, _, _ = _oracle.ref_symbolic(w.graph_offsets, w.adjacency, w.weights,
                                              order=np.array(w.symbolic.node_order, np.int32))
            seconds = []
            try:
                for r in shifts:
                    sigma = r["shift"]
                    matrix = scale_symmetric(combine(1.0, b.stiffness, -sigma, b.load), b.pencil.scale)
                    rc, msg, inertia, counts, times, _, _ = _oracle.ref_factor_solve(h, matrix, pivoting=1,
                                                                                     threads=threads)
                    if rc != 0:
                        raise RuntimeError("reference factorisation failed: " + msg)
                    r["reference_count"] = int(inertia[1])
                    r["counts_equal"] = r["reference_count"] == r["count"]
                    r["reference_delayed"], r["reference_two_by_two"] = int(counts[0]), int(counts[1])
                    seconds.append(float(times[0]))
                    log(f"reference at {sigma:.6g}: count {inertia[1]} ({'equal' if r['counts_equal'] else 'DIFFERS'}"
                        f"), {times[0]:.2f} s")
            finally:
                _oracle.reference().ref_sym_free(h)
            baseline = {"value": round(float(np.mean(seconds)), 4), "unit": "s", "cores": threads,
                        "kind": "reference",
                        "sample": "the reference's indefinite FP64 factorisation (Bunch-Kaufman, avx2 family) of "
                                  "each of the 8 shifted matrices, mean seconds per shift"}
    gpu = [r["factor_s"] for r in shifts]
    out = {"metric": "indefinite FP64 factor s per buckling shift (Bunch-Kaufman), Sturm counts vs reference",
           "value": round(float(np.mean(gpu)), 5), "unit": "s", "n_gpus": 1, "steps": len(shifts), "warmup": 0,
           "ms_per_step": round(float(np.mean(gpu)) * 1e3, 3), "higher_is_better": False, "scaling": "weak",
           "vs_baseline": None, "dtype": "f64", "data": "synthetic",
           "config": {"workload": record["workload"], "roots": 20, "shifts": 8},
           "roots": record["roots"], "shifts": shifts, "min_relative_gap": record["min_relative_gap"],
           "bisection_factorisations": record["bisection_factorisations"],
           "counts_equal": all(r.get("counts_equal", True) for r in shifts), "cpu_baseline": baseline}
    print(json.dumps(out), flush=True)


def statics_main(args, kind):
    """A statics workload other than the headline: kind "hex", the 64 × 64 × 64 hexahedral block (811,200 DOF;
    per-element E lognormal, median 200 GPa, σ = 1.5 in ln; base fixed; 4 load cases of N(0, 1) forces on the top
    face), or kind "batch", 512 independent clamped plates (sides U{20..80} elements, t ~ U[1, 5] mm; one N(0, 1) load
    case each) factorised as one block-diagonal matrix. A step is one factorisation and one solve to a backward error
    of 1e-10, with FP64 fronts and with FP32 fronts refined by conjugate gradients. The CPU baseline is the
    reference's FP64 factor and solve of the same matrix; the GPU solutions' error against its solution is
    reported."""
    import numpy as np
    import torch

    from o_o_solver_amd import FactorOptions, FactorSolve, NotConvergedError, Precision, SupernodalFactor, models

    if not torch.cuda.is_available():
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)
    t = time.perf_counter()
    if kind == "hex":
        workload = models.hex_block(64, seed=20261204, sigma=1.5)
        sym = workload.analyse()
        b = workload.loads("top", nrhs=4, seed=20261204)
        description = (f"hex8 block 64^3, {workload.n} DOF, E lognormal (median 200 GPa, sigma 1.5), base fixed; 4 "
                       "N(0,1) top-face load cases")
        metric = "factor+solve s at 1e-10 backward err (64^3 hex block): learned precision vs FP64 GPU"
    else:
        workload = models.plate_batch_c4(512)
        sym = workload.analyse()
        b = workload.loads("random", nrhs=1, seed=20261205)
        description = (f"512 clamped quad4 plates, sides U{{20..80}} elements, t U[1,5] mm, {workload.n} DOF in one "
                       "block-diagonal factor; one N(0,1) load case each")
        metric = "factor+solve s at 1e-10 backward err (512 batched plates): learned precision vs FP64 GPU"
    log(f"{workload.name}: {workload.n} DOF, {sym.supernode_count} supernodes, {sym.flop_count:.3e} flops "
        f"({time.perf_counter() - t:.1f} s)")
    plan = sym.plan()
    dm = plan.matrix(workload.matrix)
    bd = torch.tensor(b, device="cuda:0")
    tuning = {}
    for precision in (Precision.FP64, Precision.FP32):
        big, tuned, streams = plan.autotune(dm, precision)
        tuning[precision.name.lower()] = {"big_tile_launches": big, "launches_tuned": tuned, "streams": streams,
                                          "launch_mode": plan.launch_mode(precision)}
    log(f"autotune: {tuning}")
    results = {}

    def leg(name, options, steps):
        times = []
        for k in range(args.warmup + args.steps):
            torch.cuda.synchronize()
            t = time.perf_counter()
            f = SupernodalFactor.factorise(sym, workload.matrix, options, device_matrix=dm)
            try:
                x, rep = FactorSolve(f).solve_refined(bd, target=args.target, max_steps=steps)
            except NotConvergedError as e:
                x, rep = None, None
                log(f"{name}: not converged ({e})")
            torch.cuda.synchronize()
            elapsed = time.perf_counter() - t
            if k >= args.warmup:
                times.append(elapsed)
            factor_s = f.factor_seconds
            del f
        results[name] = {"seconds": round(float(np.mean(times)), 4), "factor_s": round(factor_s, 4),
                         "factor_tflops": round(sym.flop_count / factor_s / 1e12, 2),
                         "solves": rep.steps if rep else None,
                         "backward_error": rep.backward_error if rep else None,
                         "x": x.cpu().numpy() if x is not None else None}
        log(f"{name}: {results[name]['seconds'] * 1e3:.1f} ms/step, factor {factor_s * 1e3:.1f} ms "
            f"({results[name]['factor_tflops']} TFLOP/s), solves {results[name]['solves']}, eta "
            f"{results[name]['backward_error']}")

    from o_o_solver_amd.selector import PrecisionSelector
    leg("fp64", FactorOptions(precision=Precision.FP64), 4)
    # The learned selector's cut for this matrix (after the FP64 leg, whose pivot growth it reads).
    selector = PrecisionSelector.trained()
    cut, p = selector.choose(plan, dm) if selector is not None else (0, np.zeros(0))
    chosen = "fp64" if cut == 0 else ("fp32" if cut >= p.size else "mixed")
    log(f"selector: p(cut pays) = {[round(float(v), 4) for v in p]}, chose {chosen} ({cut} of {plan.levels} levels)")
    leg("fp32_fronts_cg", FactorOptions(precision=Precision.FP32), args.max_steps)
    if chosen == "mixed":
        leg("mixed_cut", FactorOptions(precision=Precision.MIXED, fp32_levels=cut), args.max_steps)
    learned = results[{"fp64": "fp64", "fp32": "fp32_fronts_cg", "mixed": "mixed_cut"}[chosen]]
    baseline = None
    if args.cpu_baseline == "auto":
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        log(f"cpu baseline: the reference's avx2 factor + solve on {threads} threads")
        baseline = cpu_baseline(workload, b, threads, keep_solution=True)
    if baseline is not None:
        xr = baseline.pop("x")
        for name in results:
            x = results[name]["x"]
            if x is not None:
                results[name]["error_vs_reference"] = float(np.max(np.abs(x - xr)) / np.max(np.abs(xr)))
    for name in results:
        results[name].pop("x")
    value = learned["seconds"]
    out = {"metric": metric, "value": value, "unit": "s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "ms_per_step": round(value * 1e3, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
           "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": description, "flops": sym.flop_count, "supernodes": sym.supernode_count},
           "learned": {"precision_chosen": chosen, "fp32_levels": int(cut), "levels": plan.levels,
                       "selector_probability": float(p.max()) if p.size else None},
           "speedup_learned_vs_fp64": round(results["fp64"]["seconds"] / value, 3),
           "autotune": tuning, **results, "cpu_baseline": baseline}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--scale", type=int, default=500, help="elements per side of the plate")
    ap.add_argument("--nrhs", type=int, default=16)
    ap.add_argument("--target", type=float, default=1e-10)
    ap.add_argument("--max-steps", type=int, default=30, help="refinement steps of the mixed path at most")
    ap.add_argument("--cpu-baseline", choices=["auto", "off"], default="auto")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: the host's CPU share, at most 16")
    ap.add_argument("--workload", choices=["plate", "cylinder", "hex", "batch"], default="plate",
                    help="plate: the headline (factor + solve of the 1.5M-DOF shell plate); cylinder: the buckling "
                         "shifts of the stiffened cylinder (indefinite factors and their Sturm counts); hex: the "
                         "64^3 hexahedral block (3D fronts), FP64 and FP32 fronts + refinement; batch: 512 plates "
                         "factorised as one")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (the factor of one matrix on one MI355X); "
              "run it with --gpus 1", file=sys.stderr)
        sys.exit(2)
    if args.workload == "cylinder":
        cylinder_main(args)
        return
    if args.workload in ("hex", "batch"):
        statics_main(args, args.workload)
        return

    import numpy as np
    import torch

    import o_o_solver_amd as oos
    from o_o_solver_amd import profiling
    from o_o_solver_amd import (FactorOptions, FactorSolve, NotConvergedError, Precision, PrecisionInsufficientError,
                                SupernodalFactor, models)

    if not torch.cuda.is_available():
        print("bench.py: no GPU", file=sys.stderr)
        sys.exit(1)
    device = torch.device("cuda:0")

    t = time.perf_counter()
    workload = models.shell_plate_c2(args.scale)
    log(f"model {workload.name}: {workload.n} DOF ({time.perf_counter() - t:.1f} s)")
    t = time.perf_counter()
    sym = workload.analyse()
    log(f"order + symbolic: {sym.supernode_count} supernodes, {sym.flop_count:.3e} flops "
        f"({time.perf_counter() - t:.1f} s)")
    b = workload.loads("random", nrhs=args.nrhs, seed=20261203)
    t = time.perf_counter()
    plan = sym.plan()
    dm = plan.matrix(workload.matrix)
    bd = torch.tensor(b, device=device)
    torch.cuda.synchronize()
    stats = plan.stats()
    log(f"plan + upload ({time.perf_counter() - t:.1f} s): {stats}")

    # The autotuner picks the update-matrix tile size per level by timing both on this matrix (setup, untimed).
    t = time.perf_counter()
    tuning = {}
    for precision in (Precision.FP64, Precision.FP32):
        big, tuned, streams = plan.autotune(dm, precision)
        tuning[precision.name.lower()] = {"big_tile_launches": big, "launches_tuned": tuned, "streams": streams,
                                          "launch_mode": plan.launch_mode(precision)}
    log(f"autotune ({time.perf_counter() - t:.1f} s): {tuning}")

    from o_o_solver_amd.selector import PrecisionSelector
    selector = PrecisionSelector.trained()
    choice = {"probability": None, "precision": "fp64", "fp32_levels": 0, "levels": stats["levels"],
              "selector_ms": None}

    def choose():
        # The learned selector's cut for this matrix, made once (it reads the pivot growth of the plan's last
        # factorisation, the FP64 leg's): the levels below it in FP32, the rest in FP64.
        if selector is None:
            return
        t = time.perf_counter()
        cut, p = selector.choose(plan, dm)
        choice["selector_ms"] = round((time.perf_counter() - t) * 1e3, 1)
        choice["probability"] = float(p.max()) if p.size else None
        choice["fp32_levels"] = cut
        choice["precision"] = "fp64" if cut == 0 else ("fp32" if cut >= p.size else "mixed")
        choice["probabilities"] = [round(float(v), 4) for v in p]

    # A step is one call of the combined factorisation and refined solve (the first solve's forward sweep of the
    # lower levels runs beside the factorisation of the top ones; the same results as the two calls). The last
    # solution of each path is kept for its error against the reference's.
    last_x = {}

    def fp32_step():
        try:
            f, x, rep = SupernodalFactor.factorise_and_solve(sym, workload.matrix, bd,
                                                             FactorOptions(precision=Precision.FP32),
                                                             device_matrix=dm, target=args.target,
                                                             max_steps=args.max_steps)
            fallback = False
        except (NotConvergedError, PrecisionInsufficientError) as e:
            # The FP32 factor is not accurate enough (its refinement stalls, or its pivots fail the tests): the FP64
            # factor decides.
            e.factor = None
            f, x, rep = SupernodalFactor.factorise_and_solve(sym, workload.matrix, bd, device_matrix=dm,
                                                             target=args.target, max_steps=4)
            fallback = True
        last_x["fp32"] = x
        return f.factor_seconds, rep, fallback

    def cut_step():
        # FP32 fronts below the selector's cut, FP64 above, refined by conjugate gradients.
        try:
            f, x, rep = SupernodalFactor.factorise_and_solve(
                sym, workload.matrix, bd, FactorOptions(precision=Precision.MIXED, fp32_levels=choice["fp32_levels"]),
                device_matrix=dm, target=args.target, max_steps=args.max_steps)
            fallback = False
        except (NotConvergedError, PrecisionInsufficientError) as e:
            e.factor = None
            f, x, rep = SupernodalFactor.factorise_and_solve(sym, workload.matrix, bd, device_matrix=dm,
                                                             target=args.target, max_steps=4)
            fallback = True
        last_x["cut"] = x
        return f.factor_seconds, rep, fallback

    def mixed_step():
        # The learned selector's choice: FP64 fronts, FP32 fronts, or FP32 below a cut of the levels.
        step = {"fp64": fp64_step, "fp32": fp32_step, "mixed": cut_step}[choice["precision"]]
        result = step()
        last_x["mixed"] = last_x[{"fp64": "fp64", "fp32": "fp32", "mixed": "cut"}[choice["precision"]]]
        return result

    def fp64_step():
        f, x, rep = SupernodalFactor.factorise_and_solve(sym, workload.matrix, bd, device_matrix=dm,
                                                         target=args.target, max_steps=4)
        last_x["fp64"] = x
        return f.factor_seconds, rep, False

    def run(step, label, profile):
        for k in range(args.warmup):
            fs, rep, fb = step()
            log(f"{label} warmup {k}: factor {fs * 1e3:.1f} ms, {rep.steps} solves, eta {rep.backward_error:.2e}"
                f"{' (fp64 fallback)' if fb else ''}")
        # The timed steps run as in production (launch sequences replayed as graphs); one more step afterwards, with
        # HIP events around every launch, gives the per-kernel times and the roofline.
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        results = [step() for _ in range(args.steps)]
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        prof = None
        if profile:
            plan.profile(True)
            step()
            torch.cuda.synchronize()
            prof = plan.profile_read()
            plan.profile(False)
        fs = [r[0] for r in results]
        reps = [r[1] for r in results]
        log(f"{label}: {elapsed / args.steps * 1e3:.1f} ms/step, factor {np.mean(fs) * 1e3:.1f} ms, "
            f"eta {max(r.backward_error for r in reps):.2e}, solves {[r.steps for r in reps]}")
        return elapsed / args.steps, float(np.mean(fs)), reps, any(r[2] for r in results), prof

    fp64_s, fp64_factor, fp64_reps, _, fp64_prof = run(fp64_step, "fp64", True)
    choose()
    log(f"selector ({choice['selector_ms']} ms): p(cut pays) = {choice.get('probabilities')}, chose "
        f"{choice['precision']} ({choice['fp32_levels']} of {choice['levels']} levels in FP32)")
    fp32_s, fp32_factor, fp32_reps, fp32_fallback, fp32_prof = run(fp32_step, "fp32 fronts + cg", True)
    mixed_s, mixed_factor, mixed_reps, fallback, _ = run(mixed_step, "mixed", False)

    # The roofline of the HIP kernel that takes the most device time in the path the mixed step ran (the profile
    # kinds summed per kernel: the three tile kinds are one kernel), with its HBM traffic per launch from the
    # committed rocprofv3 counter summary of this workload when it lists that kernel.
    chose_fp32 = choice["precision"] == "fp32"
    prof = fp32_prof if chose_fp32 else fp64_prof
    kernel_type = "float" if chose_fp32 and not fp32_fallback else "double"
    per_kernel = {}
    for kind, (sec, count, work) in prof.items():
        if count == 0:
            continue
        name = KIND_KERNELS[kind].format(T=kernel_type)
        acc = per_kernel.setdefault(name, {"seconds": 0.0, "launches": 0, "work": 0.0, "kinds": []})
        acc["seconds"] += sec
        acc["launches"] += count
        acc["work"] += work
        acc["kinds"].append(kind)
    dominant = max(per_kernel, key=lambda k: per_kernel[k]["seconds"])
    top = per_kernel[dominant]
    seconds, launches, work = top["seconds"], top["launches"], top["work"]  # of the one profiled step
    traffic = profiling.read_traffic(TRAFFIC_SUMMARY, dominant)
    if top["kinds"][0] in BYTE_KINDS:
        achieved = work / seconds / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM, "unit": "GB/s",
                    "frac": round(achieved / PEAK_HBM, 4)}
    else:
        peak = PEAK_FP32 if kernel_type == "float" else PEAK_FP64
        achieved = work / seconds / 1e12
        roofline = {"bound": "mfma", "achieved": round(achieved, 2), "peak": peak, "unit": "TFLOP/s",
                    "frac": round(achieved / peak, 4)}
    roofline.update({"traffic": round(traffic) if traffic else None, "hip_kernel": dominant, "kinds": top["kinds"],
                     "launches_per_step": launches, "avg_launch_ms": round(seconds / launches * 1e3, 4),
                     "work_per_launch": work / launches,
                     "share_of_device_time": round(seconds / sum(v[0] for v in prof.values()), 3),
                     "traffic_source": os.path.relpath(TRAFFIC_SUMMARY, ROOT) if traffic else None})
    def breakdown(profile, peak_flops):
        # Per kind: its time, launches and rate (GB/s for the HBM-bound kinds, TFLOP/s for the others) and the rate's
        # fraction of its roofline (HBM 8 TB/s, or the MFMA peak of the precision).
        out = {}
        for k, v in profile.items():
            rate = v[2] / v[0] / (1e9 if k in BYTE_KINDS else 1e12) if v[0] > 0 else None
            peak = PEAK_HBM if k in BYTE_KINDS else peak_flops
            out[k] = {"ms_per_step": round(v[0] * 1e3, 3), "launches_per_step": v[1],
                      "rate": round(rate, 2) if rate is not None else None,
                      "roofline_frac": round(rate / peak, 4) if rate is not None else None}
        return out
    kernels = breakdown(fp32_prof, PEAK_FP32)

    # The solutions on the host, for their error against the reference's FP64 solution.
    solutions = {k: v.cpu().numpy() for k, v in last_x.items()}
    baseline = None
    reference_flops = None
    errors = {}
    if args.cpu_baseline == "auto":
        threads = args.cpu_threads or min(16, len(os.sched_getaffinity(0)))
        log(f"cpu baseline: the reference's avx2 factor + solve on {threads} threads, median of 3")
        runs = [cpu_baseline(workload, b, threads, keep_solution=True) for _ in range(3)]
        if runs[0]:
            runs.sort(key=lambda r: r["value"])
            baseline = runs[1]
            baseline["runs_s"] = sorted(r["value"] for r in runs)
            reference_x = baseline.pop("x")
            for r in runs:
                r.pop("x", None)
            scale = float(np.max(np.abs(reference_x)))
            errors = {k: float(np.max(np.abs(v - reference_x)) / scale) for k, v in solutions.items()}
            log(f"cpu baseline: {baseline['value']} s ({baseline['gflops_factor']} GFLOP/s factor); errors vs the "
                f"reference's solution {errors}")
            single = cpu_baseline(workload, b, 1)
            baseline["one_thread"] = {k: single[k] for k in ("value", "factor_s", "solve_s", "gflops_factor")}
            log(f"cpu baseline, 1 thread: {single['value']} s ({single['gflops_factor']} GFLOP/s factor)")
            reference_flops = reference_order_flops(workload, threads)
            log(f"flops under the reference's own order (GeometricDissection): {reference_flops:.4e} "
                f"(this library's order: {sym.flop_count:.4e})")

    flops = sym.flop_count
    out = {
        "metric": METRIC,
        "value": round(mixed_s, 5),
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed_s * 1e3, 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"shell plate {args.scale}x{args.scale} quad4, {workload.n} DOF, jitter 0.2, "
                               f"t log-uniform [1,4] mm, clamped; {args.nrhs} N(0,1) load cases",
                   "nrhs": args.nrhs, "target_backward_error": args.target, "flops": flops,
                   "flops_reference_order": reference_flops,
                   "supernodes": sym.supernode_count, "levels": stats["levels"]},
        "fp64": {"seconds": round(fp64_s, 5), "factor_s": round(fp64_factor, 5),
                 "solve_s": round(float(np.mean([r.seconds for r in fp64_reps])), 5),
                 "solves": max(r.steps for r in fp64_reps),
                 "backward_error": max(r.backward_error for r in fp64_reps),
                 "factor_tflops": round(flops / fp64_factor / 1e12, 2),
                 "mfma_frac": round(flops / fp64_factor / 1e12 / PEAK_FP64, 4)},
        "mixed": {"seconds": round(mixed_s, 5), "factor_s": round(mixed_factor, 5),
                  "solve_s": round(float(np.mean([r.seconds for r in mixed_reps])), 5),
                  "solves": max(r.steps for r in mixed_reps),
                  "backward_error": max(r.backward_error for r in mixed_reps),
                  "precision_chosen": choice["precision"], "fp32_levels": choice["fp32_levels"],
                  "levels": choice["levels"], "selector_probability": choice["probability"],
                  "selector_ms": choice["selector_ms"], "fp64_fallback": fallback},
        "fp32_fronts_cg": {"seconds": round(fp32_s, 5), "factor_s": round(fp32_factor, 5),
                           "solve_s": round(float(np.mean([r.seconds for r in fp32_reps])), 5),
                           "solves": max(r.steps for r in fp32_reps),
                           "backward_error": max(r.backward_error for r in fp32_reps),
                           "fp64_fallback": fp32_fallback,
                           "factor_tflops": round(flops / fp32_factor / 1e12, 2),
                           "mfma_frac": round(flops / fp32_factor / 1e12 / PEAK_FP32, 4)},
        "speedup_mixed_vs_fp64": round(fp64_s / mixed_s, 3),
        # ‖x − x_ref‖∞ / ‖x_ref‖∞ against the reference's FP64 CPU factor and solve of the same matrix and loads.
        "error_vs_reference": {k: float(f"{v:.3e}") for k, v in errors.items()} or None,
        "roofline": roofline,
        "autotune": tuning,
        "kernels_fp32_fronts_cg": kernels,
        "kernels_fp64": breakdown(fp64_prof, PEAK_FP64),
        "cpu_baseline": baseline,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
