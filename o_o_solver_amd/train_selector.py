"""Trains the learned precision selector on the GPU from real outcomes and writes its weights.

For a family of synthetic models (shell plates over a range of sizes, thicknesses and thickness contrasts; hexahedral
blocks over a range of stiffness contrasts) it factorises each with FP32 fronts and refines with conjugate gradients
to a backward error of 1e-10 within a budget of solves; the label is 1 when that succeeds. The features are those of
oosg_selector_features. The selector is then trained (full-batch Adam) and its parameters, with the data, written:

    python -m o_o_solver_amd.train_selector [--out DIR]   # DIR/selector_weights.npz, DIR/selector_data.npz
"""

import argparse
import itertools
import os
import sys
import time

import numpy as np


def problems():
    for n, t, ratio, jitter in itertools.product((16, 28, 44, 64), (1e-3, 4e-3, 2e-2, 8e-2), (1.0, 4.0, 30.0),
                                                 (0.0, 0.2)):
        yield (f"plate {n} t={t} x{ratio} j={jitter}", "plate",
               dict(nx=n, jitter=jitter, seed=n + int(1e4 * t), thickness=(t, t * ratio),
                    thickness_mode="constant" if ratio == 1.0 else "log-uniform"))
    for n, sigma in itertools.product((6, 10, 14), (0.0, 1.0, 2.0, 3.0)):
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
        f32 = SupernodalFactor.factorise(sym, w.matrix, FactorOptions(precision=Precision.FP32), device_matrix=dm)
        try:
            _, rep = FactorSolve(f32).solve_refined(b, target=1e-10, max_steps=args.budget)
            ok, steps = 1.0, rep.steps
        except NotConvergedError:
            ok, steps = 0.0, args.budget
        features.append(f)
        labels.append(ok)
        names.append(name)
        solves.append(steps)
        print(f"[train] {name}: n={w.n} label={int(ok)} solves={steps} ({time.perf_counter() - t0:.1f} s)",
              file=sys.stderr, flush=True)
        del f32, dm, plan
    x = np.array(features, np.float32)
    y = np.array(labels, np.float32)
    selector = PrecisionSelector(seed=20261205)
    loss = selector.train(x, y, steps=args.steps, learning_rate=5e-3)
    accuracy = float(np.mean((selector.predict(x) >= 0.5) == (y == 1)))
    print(f"[train] {len(y)} problems, {int(y.sum())} converged with FP32 fronts; loss {loss:.4f}, "
          f"training accuracy {accuracy:.3f}", file=sys.stderr, flush=True)
    selector.save(os.path.join(args.out, "selector_weights.npz"))
    np.savez(os.path.join(args.out, "selector_data.npz"), features=x, labels=y, names=np.array(names),
             solves=np.array(solves, np.int32))


if __name__ == "__main__":
    main()
