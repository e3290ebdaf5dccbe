"""Shifted pencils K − σM on the GPU factor, and their Sturm counts.

The host side of the reference's eigenvalue pencils, as far as they feed the numeric factorisation:
  combine / scale_symmetric   SymmetricCsc::combine and ::scaleSymmetric (src/linalg/sparse/SymmetricCsc.cpp:209-252)
  pencil_scale_of             pencilScaleOf (src/linalg/eigen/PencilScaling.cpp:7-60): powers of two near 1/sqrt(K_ii)
  SturmCheck                  include/oos/linalg/eigen/SturmCheck.hpp:53-72, src/linalg/eigen/SturmCheck.cpp:24-54
  ShiftedPencil               the factorise(σ) / solve(block) / multiplyMass part of ShiftedPencil
                              (include/oos/linalg/eigen/ShiftedPencil.hpp:34-88): S·(K − σM)·S factorised in indefinite
                              mode on the device; its negative pivots are ν(σ), the eigenvalues of Kφ = λMφ below σ.
"""

from dataclasses import dataclass

import numpy as np

from . import FactorOptions, FactorSolve, PivotingMode, SupernodalFactor, SymmetricCsc


class SolverError(RuntimeError):
    """base::SolverError: the Sturm counts contradict each other."""


def combine(alpha, a, beta, b):
    """alpha·A + beta·B of two matrices of one structure: two products and a sum, each rounded (no fused
    multiply-add), the reference's bits."""
    if a.n != b.n or not (np.array_equal(a.offsets, b.offsets) and np.array_equal(a.rows, b.rows)):
        raise ValueError(f"SymmetricCsc::combine: the matrices must have the same structure, found dimensions {a.n} "
                         f"and {b.n} with {a.offsets[-1]} and {b.offsets[-1]} terms or other rows; assemble both on "
                         "one sparse pattern and convert them with one DOF selection")
    left = np.multiply(alpha, a.values)
    right = np.multiply(beta, b.values)
    return SymmetricCsc(a.n, a.offsets, a.rows, left + right)


def scale_symmetric(matrix, scale):
    """S·A·S: each term (i, j) becomes scale[i]·a_ij·scale[j], in that order."""
    scale = np.asarray(scale, dtype=np.float64)
    if scale.shape != (matrix.n,):
        raise ValueError(f"SymmetricCsc::scaleSymmetric: expected {matrix.n} scales (one per equation), found "
                         f"{scale.size}")
    columns = np.repeat(np.arange(matrix.n), np.diff(matrix.offsets))
    values = (scale[matrix.rows] * matrix.values) * scale[columns]
    return SymmetricCsc(matrix.n, matrix.offsets, matrix.rows, values)


def pencil_scale_of(stiffness):
    """Per equation 2^(−⌊e/2⌋) for K_jj = m·2^e (m in [1/2, 1)): the scaled diagonal lies in [1/4, 2); 1 where the
    diagonal term is missing, not positive or not finite."""
    starts = stiffness.offsets[:-1]
    has = (stiffness.offsets[1:] > starts)
    diagonal = np.zeros(stiffness.n)
    first = np.where(has, starts, 0)
    is_diag = has & (stiffness.rows[first] == np.arange(stiffness.n))
    diagonal[is_diag] = stiffness.values[first[is_diag]]
    scale = np.ones(stiffness.n)
    ok = (diagonal > 0.0) & np.isfinite(diagonal)
    _, exponent = np.frexp(diagonal[ok])
    # C's integer division truncates toward zero.
    scale[ok] = np.ldexp(1.0, -np.fix(exponent / 2).astype(np.int64))
    return scale


@dataclass(frozen=True)
class SturmPoint:
    shift: float
    count_below: int


@dataclass(frozen=True)
class SturmVerification:
    verified_up_to: float
    verified_count: int
    has_mismatch: bool
    mismatch: tuple = None  # (lower, upper, expected, found)


class SturmCheck:
    """Sturm points ν(σ) kept sorted; a point that contradicts another (a count that decreases as the shift grows,
    or two counts at one shift) raises SolverError."""

    def __init__(self):
        self.points = []

    def add_point(self, point):
        shifts = [p.shift for p in self.points]
        at = int(np.searchsorted(shifts, point.shift, side="left"))
        same = at < len(self.points) and self.points[at].shift == point.shift
        below_next = at == len(self.points) or (self.points[at].count_below == point.count_below if same
                                                  else self.points[at].count_below >= point.count_below)
        above_previous = at == 0 or self.points[at - 1].count_below <= point.count_below
        if not below_next or not above_previous:
            other = self.points[at] if not below_next else self.points[at - 1]
            raise SolverError(f"eigenvalue analysis: the Sturm counts contradict each other: {point.count_below} "
                              f"below {point.shift} against {other.count_below} below {other.shift}; the "
                              "factorisations of K - sigma*M are not reliable")
        if not same:
            self.points.insert(at, point)

    def count_below(self, shift):
        for p in self.points:
            if p.shift == shift:
                return p.count_below
        raise ValueError(f"SturmCheck::countBelow: no Sturm point at the shift {shift}; add it first")

    def verify(self, sorted_eigenvalues):
        """From the lowest point upward, the intervals [a, b) whose number of eigenvalues found equals the Sturm
        difference are verified; the first that does not is the mismatch."""
        values = np.asarray(sorted_eigenvalues, dtype=np.float64)
        lowest = self.points[0]
        verified_up_to, verified = lowest.shift, 0
        for lo, hi in zip(self.points[:-1], self.points[1:]):
            expected = hi.count_below - lo.count_below
            found = int(np.count_nonzero((values >= lo.shift) & (values < hi.shift)))
            if found != expected:
                return SturmVerification(verified_up_to, verified, True, (lo.shift, hi.shift, expected, found))
            verified_up_to = hi.shift
            verified += found
        return SturmVerification(verified_up_to, verified, False, None)


@dataclass
class ShiftFactorisation:
    shift: float
    factor: SupernodalFactor
    scale: np.ndarray

    @property
    def count_below(self):
        """ν(σ): the negative pivots of S·(K − σM)·S, the eigenvalues below σ (Sylvester)."""
        return self.factor.inertia.negative

    @property
    def inertia(self):
        return self.factor.inertia


class ShiftedPencil:
    """K − σM on one symbolic structure (K and M assembled on one pattern, factor order). Each factorise(σ) forms
    S·(K − σM)·S on the host (O(nnz)), copies its values into the plan's device matrix and factorises it in indefinite
    mode on the GPU; solve(block) returns (K − σM)⁻¹·block = S·(S·(K − σM)·S)⁻¹·S·block."""

    def __init__(self, symbolic, stiffness, mass, device=0):
        if not (np.array_equal(stiffness.offsets, mass.offsets) and np.array_equal(stiffness.rows, mass.rows)):
            raise ValueError("ShiftedPencil: K and M must have one structure")
        self.symbolic = symbolic
        self.stiffness = stiffness
        self.mass = mass
        self.scale = pencil_scale_of(stiffness)
        self.plan = symbolic.plan(device)
        self.device_matrix = None
        self.sturm = SturmCheck()

    @property
    def dimension(self):
        return self.stiffness.n

    def factorise(self, shift):
        shifted = scale_symmetric(combine(1.0, self.stiffness, -float(shift), self.mass), self.scale)
        if self.device_matrix is None:
            self.device_matrix = self.plan.matrix(shifted)
        else:
            self.device_matrix.set_values(shifted.values)
        factor = SupernodalFactor.factorise(self.symbolic, shifted, FactorOptions(pivoting=PivotingMode.INDEFINITE),
                                            device_matrix=self.device_matrix)
        result = ShiftFactorisation(float(shift), factor, self.scale)
        self.sturm.add_point(SturmPoint(float(shift), result.count_below))
        return result

    def solve(self, factorisation, block):
        """(K − σM)⁻¹·block for an n × k host block."""
        b = np.asarray(block, dtype=np.float64)
        s = self.scale.reshape(-1, 1) if b.ndim == 2 else self.scale
        return s * FactorSolve(factorisation.factor).solve(s * b)

    def multiply_mass(self, block):
        """M·block (host arrays; the symmetric product from the lower CSC)."""
        return self.mass.multiply(block)
