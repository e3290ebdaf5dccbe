"""Trains the learned precision selector on the GPU from real outcomes and writes its weights.

For a family of synthetic models (shell plates over a range of sizes, thicknesses and thickness contrasts; hexahedral
blocks over a range of sizes and stiffness contrasts) it factorises and solves each to a backward error of 1e-10 on
the device (16 load cases, one combined call each, the median of two timed calls after a warm one) with FP64 fronts
throughout and with every cut of the plan's levels: cut l factorises the levels below l in FP32 (l = L: FP32 fronts
throughout) and refines by conjugate gradients. A sample is one (model, cut) pair: its features are
oosg_selector_cut_features after the FP64 factorisation (whose pivot growth the features read), its label the
measured gain clip(0.5 + t_FP64 / t_cut − 1.02, 0, 1): above 1/2 when the cut reaches the target in less than 0.98 of
the FP64 time, and the larger the gain, the larger the label (0 when it does not converge), so that the cut of the
highest probability is the one predicted to gain the most. Every fourth model is held out; the selector is trained
(full-batch Adam, the library's kernels) on the others, and the time of its choices on the held-out models is set
against FP64 and FP32 fronts throughout:

    python -m o_o_solver_amd.train_selector [--out DIR]   # DIR/selector_weights.npz, DIR/selector_data.npz
    python -m o_o_solver_amd.train_selector --measured DIR/selector_data.npz   # retrain on measured times
    python -m o_o_solver_amd.train_selector --measured FILE --measure-from I   # keep models < I, measure the rest
"""

import argparse
import gc
import itertools
import os
import sys
import time

import numpy as np


def problems():
    for n, t, ratio in itertools.product((60, 120, 200, 300, 400), (2e-3, 2e-2), (1.0, 4.0)):
        yield (f"plate {n} t={t} x{ratio}", "plate",
               dict(nx=n, jitter=0.2, seed=n + int(1e4 * t), thickness=(t, t * ratio),
                    thickness_mode="constant" if ratio == 1.0 else "log-uniform"))
    for n, t, ratio in itertools.product((20, 32, 48), (1e-3, 2e-2), (1.0, 30.0)):
        yield (f"plate {n} t={t} x{ratio}", "plate",
               dict(nx=n, jitter=0.2, seed=n + int(1e4 * t), thickness=(t, t * ratio),
                    thickness_mode="constant" if ratio == 1.0 else "log-uniform"))
    for n, t, ratio in itertools.product((300, 400), (5e-2,), (1.0, 4.0)):
        yield (f"plate {n} t={t} x{ratio}", "plate",
               dict(nx=n, jitter=0.2, seed=n + int(1e4 * t), thickness=(t, t * ratio),
                    thickness_mode="constant" if ratio == 1.0 else "log-uniform"))
    for n, sigma in itertools.product((8, 14, 20, 28, 36), (0.0, 2.0)):
        yield f"hex {n} sigma={sigma}", "hex", dict(nx=n, seed=n, sigma=sigma)
    for n, sigma in itertools.product((44, 52, 60), (0.0, 1.0, 2.0, 3.0)):
        yield f"hex {n} sigma={sigma}", "hex", dict(nx=n, seed=n, sigma=sigma)
    for n, t, ratio in itertools.product((450, 500), (1e-3, 2e-3, 2e-2), (1.0, 4.0)):
        yield (f"plate {n} t={t} x{ratio}", "plate",
               dict(nx=n, jitter=0.2, seed=n + int(1e4 * t), thickness=(t, t * ratio),
                    thickness_mode="constant" if ratio == 1.0 else "log-uniform"))


def choice_times(probability, times, fp64):
    """The time of the selector's choice: the cut of the highest probability if it is at least 1/2, else FP64."""
    best = int(np.argmax(probability))
    return times[best] if probability[best] >= 0.5 else fp64


def gain_labels(times, fp64):
    """The label of each cut from its measured time and the FP64 time (infinite: not converged)."""
    with np.errstate(divide="ignore"):
        ratio = np.where(np.isfinite(times), fp64 / times, 0.0)
    return np.clip(0.5 + ratio - 1.02, 0.0, 1.0).astype(np.float32)


def measure(args, start=0):
    """Times every model from index `start` on with FP64 fronts and with every cut; the samples' features, times and
    models."""
    import torch

    from o_o_solver_amd import (FactorOptions, NotConvergedError, Precision, PrecisionInsufficientError,
                                SupernodalFactor, models)
    from o_o_solver_amd.selector import PrecisionSelector

    features, times, problem, cuts = [], [], [], []
    names, fp64_times, fp32_times, held_out = [], [], [], []
    for index, (name, kind, kw) in enumerate(problems()):
        if index < start:
            continue
        t0 = time.perf_counter()
        w = models.plate(**kw) if kind == "plate" else models.hex_block(kw["nx"], seed=kw["seed"], sigma=kw["sigma"])
        sym = w.analyse()
        plan = sym.plan()
        dm = plan.matrix(w.matrix)
        for precision in (Precision.FP64, Precision.FP32):
            plan.autotune(dm, precision)
        b = torch.tensor(w.loads("random", nrhs=16, seed=1), device="cuda:0")
        levels = plan.levels

        def timed(options):
            # A warm call first (graphs captured, buffers in place), then the median of two timed ones.
            runs, ok = [], True
            for rep in range(3):
                torch.cuda.synchronize()
                t = time.perf_counter()
                try:
                    f, _, _ = SupernodalFactor.factorise_and_solve(sym, w.matrix, b, options, device_matrix=dm,
                                                                   target=1e-10, max_steps=args.budget)
                    del f
                except (NotConvergedError, PrecisionInsufficientError) as e:
                    e.factor = None
                    ok = False
                torch.cuda.synchronize()
                if rep > 0:
                    runs.append(time.perf_counter() - t)
            return ok, float(np.median(runs))

        _, t64 = timed(FactorOptions(precision=Precision.FP64))
        f = PrecisionSelector.cut_features(plan, dm)  # the pivot growth of the FP64 factorisation
        cut_times = []
        for cut in range(1, levels + 1):
            options = (FactorOptions(precision=Precision.FP32) if cut == levels else
                       FactorOptions(precision=Precision.MIXED, fp32_levels=cut))
            ok, t = timed(options)
            cut_times.append(t if ok else np.inf)
            features.append(f[cut - 1])
            times.append(t if ok else np.inf)
            problem.append(index)
            cuts.append(cut)
        names.append(name)
        fp64_times.append(t64)
        fp32_times.append(cut_times[-1])
        held_out.append(index % 4 == 3)
        best = int(np.argmin(cut_times))
        pay = int(np.sum(gain_labels(np.array(cut_times), t64) > 0.5))
        print(f"[train] {name}: n={w.n} levels={levels} fp64 {t64 * 1e3:.1f} ms, fp32 {cut_times[-1] * 1e3:.1f} ms, "
              f"best cut {best + 1} {cut_times[best] * 1e3:.1f} ms, {pay} cuts pay "
              f"({time.perf_counter() - t0:.1f} s)", file=sys.stderr, flush=True)
        # The plans hold device memory; drop this model's before the next one plans its own.
        del dm, plan, sym, b, w, timed
        gc.collect()
        torch.cuda.empty_cache()
    return dict(features=np.array(features, np.float32), times=np.array(times), problem=np.array(problem, np.int32),
                cut=np.array(cuts, np.int32), names=np.array(names), fp64_times=np.array(fp64_times),
                fp32_times=np.array(fp32_times), held_out=np.array(held_out, bool))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    ap.add_argument("--budget", type=int, default=30, help="solves of the refinement at most")
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--measured", help="a selector_data.npz whose measured times to train on")
    ap.add_argument("--measure-from", type=int, default=None,
                    help="with --measured: measure the models from this index on and add them to its data")
    args = ap.parse_args()
    from o_o_solver_amd.selector import PrecisionSelector

    keys = ("features", "times", "problem", "cut", "names", "fp64_times", "fp32_times", "held_out")
    if args.measured:
        with np.load(args.measured, allow_pickle=False) as d:
            data = {k: d[k] for k in keys}
        if args.measure_from is not None:
            kept = data["problem"] < args.measure_from
            models_kept = args.measure_from
            data = {k: (v[kept] if k in ("features", "times", "problem", "cut") else v[:models_kept])
                    for k, v in data.items()}
            more = measure(args, args.measure_from)
            data = {k: np.concatenate([data[k], more[k]]) for k in keys}
    else:
        data = measure(args)
    x, times, problem = data["features"], data["times"], data["problem"]
    y = gain_labels(times, data["fp64_times"][problem])
    held = data["held_out"]
    train = ~held[problem]
    selector = PrecisionSelector(seed=20261205)
    loss = selector.train(x[train], y[train], steps=args.steps, learning_rate=5e-3)
    p = selector.predict(x)
    accuracy = float(np.mean((p[train] >= 0.5) == (y[train] > 0.5)))
    held_accuracy = float(np.mean((p[~train] >= 0.5) == (y[~train] > 0.5))) if (~train).any() else float("nan")
    totals = {"learned": 0.0, "fp64": 0.0, "fp32": 0.0, "best": 0.0}
    for q in np.flatnonzero(held):
        rows = problem == q
        totals["learned"] += choice_times(p[rows], times[rows], data["fp64_times"][q])
        totals["fp64"] += data["fp64_times"][q]
        totals["fp32"] += data["fp32_times"][q]
        totals["best"] += min(data["fp64_times"][q], times[rows].min())
    print(f"[train] {len(data['names'])} models, {len(y)} cuts ({int(np.sum(y > 0.5))} pay); loss {loss:.4f}, "
          f"accuracy {accuracy:.3f} (training), {held_accuracy:.3f} (held out); held-out time (ms): "
          + ", ".join(f"{k} {v * 1e3:.1f}" for k, v in totals.items()), file=sys.stderr, flush=True)
    selector.save(os.path.join(args.out, "selector_weights.npz"))
    np.savez(os.path.join(args.out, "selector_data.npz"), labels=y, probability=p, **data)


if __name__ == "__main__":
    main()
