"""Trains the learned precision selector on the GPU from real outcomes and writes its weights.

For a family of synthetic models (shell plates over a range of sizes, thicknesses and thickness contrasts; hexahedral
blocks over a range of stiffness contrasts) it solves each to a backward error of 1e-10 both ways on the device: FP32
fronts refined with conjugate gradients (within a budget of solves), and FP64 fronts with iterative refinement; the
label is 1 when the FP32 way converges and takes less time. The features are those of oosg_selector_features. The
selector is then trained (full-batch Adam) and its parameters, with the data, written:

    python -m o_o_solver_amd.train_selector [--out DIR]   # DIR/selector_weights.npz, DIR/selector_data.npz
"""

import argparse
import itertools
import os
import sys
import time

import numpy as np


def problems():
    for n, t, ratio in itertools.product((120, 220, 320, 400), (2e-3, 2e-2), (1.0, 4.0)):
        yield (f"plate {n} t={t} x{ratio}", "plate",
               dict(nx=n, jitter=0.2, seed=n + int(1e4 * t), thickness=(t, t * ratio),
                    thickness_mode="constant" if ratio == 1.0 else "log-uniform"))
    for n, t, ratio, jitter in itertools.product((16, 28, 44, 64), (1e-3, 4e-3, 2e-2, 8e-2), (1.0, 4.0, 30.0),
                                                 (0.0, 0.2)):
        yield (f"plate {n} t={t} x{ratio} j={jitter}", "plate",
               dict(nx=n, jitter=jitter, seed=n + int(1e4 * t), thickness=(t, t * ratio),
                    thickness_mode="constant" if ratio == 1.0 else "log-uniform"))
    for n, sigma in itertools.product((6, 10, 14, 20, 28, 36), (0.0, 1.0, 2.0, 3.0)):
        yield f"hex {n} sigma={sigma}", "hex", dict(nx=n, seed=n, sigma=sigma)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.dirname(os.path.abspath(__file__)))
    ap.add_argument("--budget", type=int, default=30, help="solves of the refinement at most")
    ap.add_argument("--steps", type=int, default=2000)
    args = ap.parse_args()
    import torch

    from o_o_solver_amd import FactorOptions, FactorSolve, NotConvergedError, Precision, SupernodalFactor, models
    from o_o_solver_amd.selector import PrecisionSelector

    features, labels, names, solves = [], [], [], []
    for name, kind, kw in problems():
        t0 = time.perf_counter()
        w = models.plate(**kw) if kind == "plate" else models.hex_block(kw["nx"], seed=kw["seed"], sigma=kw["sigma"])
        sym = w.analyse()
        plan = sym.plan()
        dm = plan.matrix(w.matrix)
        f = PrecisionSelector.features(plan, dm)
        b = torch.tensor(w.loads("random", nrhs=4, seed=1), device="cuda:0")
        def timed(precision, budget):
            # A warm run first (graphs captured, buffers in place), then the timed one.
            result = None
            for _ in range(2):
                torch.cuda.synchronize()
                t = time.perf_counter()
                factor = SupernodalFactor.factorise(sym, w.matrix, FactorOptions(precision=precision),
                                                    device_matrix=dm)
                try:
                    _, rep = FactorSolve(factor).solve_refined(b, target=1e-10, max_steps=budget)
                    result = (True, rep.steps)
                except NotConvergedError:
                    result = (False, budget)
                torch.cuda.synchronize()
                result = result + (time.perf_counter() - t,)
                del factor
            return result
        ok32, steps, t32 = timed(Precision.FP32, args.budget)
        ok64, _, t64 = timed(Precision.FP64, 4)
        ok = 1.0 if ok32 and (not ok64 or t32 < t64) else 0.0
        features.append(f)
        labels.append(ok)
        names.append(name)
        solves.append(steps)
        print(f"[train] {name}: n={w.n} label={int(ok)} fp32 {t32 * 1e3:.1f} ms ({steps} solves"
              f"{'' if ok32 else ', not converged'}) fp64 {t64 * 1e3:.1f} ms ({time.perf_counter() - t0:.1f} s)",
              file=sys.stderr, flush=True)
        del dm, plan
    x = np.array(features, np.float32)
    y = np.array(labels, np.float32)
    selector = PrecisionSelector(seed=20261205)
    loss = selector.train(x, y, steps=args.steps, learning_rate=5e-3)
    accuracy = float(np.mean((selector.predict(x) >= 0.5) == (y == 1)))
    print(f"[train] {len(y)} problems, {int(y.sum())} faster with FP32 fronts; loss {loss:.4f}, "
          f"training accuracy {accuracy:.3f}", file=sys.stderr, flush=True)
    selector.save(os.path.join(args.out, "selector_weights.npz"))
    np.savez(os.path.join(args.out, "selector_data.npz"), features=x, labels=y, names=np.array(names),
             solves=np.array(solves, np.int32))


if __name__ == "__main__":
    main()
