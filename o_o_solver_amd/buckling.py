"""Buckling-shift factorisations of the stiffened cylinder: the inner step of a buckling analysis (SOL 105) on the GPU.

The buckling factors λ of (K + λ·K_G)φ = 0 (K_G the geometric stiffness of an axial compression, negative
semi-definite) are the eigenvalues of the pencil (K, G) with G = −K_G. The number of them below a shift σ is the
number of negative pivots of K − σ·G (Sylvester): the Sturm count of ShiftedPencil.factorise(σ), an indefinite
(Bunch–Kaufman) FP64 factorisation of S·(K − σG)·S on the GPU. The reference's BucklingPencil
(src/linalg/eigen/BucklingPencil.cpp:74-97) shifts the same pencil.

  roots(count)   the first `count` buckling factors, bracketed by bisection on the Sturm counts;
  shifts(...)    σ_i = λ1 + (i + ½)(λ_count − λ1)/k, i < k: shifts spread across the first roots.

    python -m o_o_solver_amd.buckling [--circ 360] [--axial 200] [--roots 20] [--shifts 8]

prints one JSON line: the roots, the shifts, the GPU Sturm count at each, each factorisation's time, and each shift's
relative gap to the nearest root.
"""

import argparse
import json
import sys
import time

import numpy as np

from . import models
from .pencil import ShiftedPencil


class BucklingShifts:
    """The buckling pencil of a workload with a geometric stiffness, factorised at shifts on the GPU."""

    def __init__(self, workload, log=None):
        self.workload = workload
        self.log = log
        self.symbolic = workload.symbolic if workload.symbolic is not None else workload.analyse()
        self.stiffness = workload.matrix
        geometric = workload.geometric()
        # G = −K_G, positive semi-definite under compression.
        self.load = type(geometric)(geometric.n, geometric.offsets, geometric.rows, -geometric.values)
        self.pencil = ShiftedPencil(self.symbolic, self.stiffness, self.load, transform="buckling")
        self.counts = {}
        self.seconds = []

    def count_below(self, shift):
        if shift not in self.counts:
            f = self.pencil.factorise(shift)
            self.counts[shift] = f.count_below
            self.seconds.append(f.factor.factor_seconds)
            if self.log is not None and len(self.counts) % 10 == 0:
                self.log(f"{len(self.counts)} factorisations; ν({shift:.6g}) = {f.count_below}")
        return self.counts[shift]

    def roots(self, count, guess=1.0, tolerance=1e-6):
        """Brackets [lo_q, hi_q] of the first `count` buckling factors, each of relative width <= tolerance: the
        Sturm counts at lo_q are < q and at hi_q >= q. Returns (lo, hi) arrays."""
        hi = guess
        while self.count_below(hi) < count:
            hi = 2.0 * hi
        intervals = [(0.0, hi)]  # ν(0) = 0: K is positive definite
        lows, highs = np.full(count, np.nan), np.full(count, np.nan)
        while intervals:
            a, b = intervals.pop()
            ca, cb = self.count_below(a) if a > 0 else 0, self.count_below(b)
            if cb == ca or ca >= count:
                continue
            if b - a <= tolerance * b:
                for q in range(ca, min(cb, count)):
                    lows[q], highs[q] = a, b
                continue
            mid = 0.5 * (a + b)
            intervals.append((mid, b))
            intervals.append((a, mid))
        return lows, highs

    @staticmethod
    def shifts(first, last, k):
        return [first + (i + 0.5) * (last - first) / k for i in range(k)]


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--circ", type=int, default=360)
    ap.add_argument("--axial", type=int, default=200)
    ap.add_argument("--roots", type=int, default=20)
    ap.add_argument("--shifts", type=int, default=8)
    ap.add_argument("--tolerance", type=float, default=1e-5)
    args = ap.parse_args(argv)
    record = run(args.circ, args.axial, args.roots, args.shifts, args.tolerance)
    record.pop("buckling")
    print(json.dumps(record))


def run(circ, axial, root_count, shift_count, tolerance, log=None):
    """The workload: roots by bisection, then a factorisation and a solve at each shift. Returns the record."""

    if log is None:
        def log(message):
            print(f"[buckling {time.strftime('%H:%M:%S')}] {message}", file=sys.stderr, flush=True)

    t = time.perf_counter()
    w = models.cylinder(circ, axial)
    w.analyse()
    log(f"{w.name}: {w.n} DOF, {w.symbolic.supernode_count} supernodes ({time.perf_counter() - t:.1f} s)")
    b = BucklingShifts(w, log)
    t = time.perf_counter()
    lows, highs = b.roots(root_count, guess=100.0, tolerance=tolerance)
    roots = 0.5 * (lows + highs)
    log(f"roots by bisection ({len(b.counts)} factorisations, {time.perf_counter() - t:.1f} s): {roots.tolist()}")
    shifts = BucklingShifts.shifts(roots[0], roots[-1], shift_count)
    records = []
    rng = np.random.default_rng(20261206)
    for sigma in shifts:
        f = b.pencil.factorise(sigma)
        rhs = rng.standard_normal((w.n, 2))
        x = b.pencil.solve(f, rhs)
        shifted = b.stiffness.multiply(x) - sigma * b.load.multiply(x)
        # The normwise backward error max|r| / (‖K − σG‖∞·max|x| + max|b|), ‖·‖∞ bounded by ‖K‖∞ + σ‖G‖∞.
        norm = b.stiffness.norm_inf() + sigma * b.load.norm_inf()
        eta = float(np.max(np.abs(rhs - shifted)) / (norm * np.max(np.abs(x)) + np.max(np.abs(rhs))))
        record = {"shift": sigma, "count": int(f.count_below), "factor_s": round(f.factor.factor_seconds, 5),
                  "delayed_pivots": f.factor.delayed_pivot_count, "two_by_two": f.factor.two_by_two_pivot_count,
                  "relative_gap_to_root": float(np.min(np.abs(roots - sigma)) / sigma),
                  "backward_error": eta}
        log(json.dumps(record))
        records.append(record)
    return {"workload": f"stiffened cylinder R=1 m L=2 m, {circ}x{axial} quad4, {circ // 15} stringers, "
                        f"N_x=-1 kN/m; {w.n} DOF",
            "dof": w.n, "roots": roots.tolist(), "shifts": records,
            "min_relative_gap": min(r["relative_gap_to_root"] for r in records),
            "bisection_factorisations": len(b.counts), "buckling": b}


if __name__ == "__main__":
    main()
