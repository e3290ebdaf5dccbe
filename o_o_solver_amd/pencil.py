"""Shifted pencils K − σM on the GPU factor, and their Sturm counts.

The host side of the reference's eigenvalue pencils, as far as they feed the numeric factorisation:
  combine / scale_symmetric   SymmetricCsc::combine and ::scaleSymmetric (src/linalg/sparse/SymmetricCsc.cpp:209-252)
  pencil_scale_of             pencilScaleOf (src/linalg/eigen/PencilScaling.cpp:7-60): powers of two near 1/sqrt(K_ii)
  SturmCheck                  include/oos/linalg/eigen/SturmCheck.hpp:53-72, src/linalg/eigen/SturmCheck.cpp:24-54
  ShiftedPencil               ShiftedPencil (include/oos/linalg/eigen/ShiftedPencil.hpp:34-88) held on the device:
                              S·(K − σM)·S formed and factorised in indefinite mode there, its negative pivots ν(σ) the
                              eigenvalues of Kφ = λMφ below σ; the solves, M- and inner products on device blocks.
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
    """The pencil (K, M) of K φ = λ M φ on one symbolic structure (K and M assembled on one pattern, factor order), as
    the reference's eigenvalue solvers see it (ShiftedPencil, include/oos/linalg/eigen/ShiftedPencil.hpp:55-88), held
    on the device: K and M are uploaded once; factorise(σ) forms S·(K − σM)·S on the device
    (oosg_matrix_set_combination: SymmetricCsc::combine and scaleSymmetric with the reference's rounding) and
    factorises it in indefinite mode; the factor stays resident for the solves of its blocks; the products with M and
    with the inner-product matrix run on the device. Blocks are torch tensors on the device (or host arrays, copied).

    transform: "vibration" (the inner product of M) or "buckling" (M = B = −K_G, the inner product of K)."""

    def __init__(self, symbolic, stiffness, mass, device=0, transform="vibration"):
        if not (np.array_equal(stiffness.offsets, mass.offsets) and np.array_equal(stiffness.rows, mass.rows)):
            raise ValueError("ShiftedPencil: K and M must have one structure")
        if transform not in ("vibration", "buckling"):
            raise ValueError("ShiftedPencil: transform is 'vibration' or 'buckling'")
        self.symbolic = symbolic
        self.stiffness = stiffness
        self.mass = mass
        self._transform = transform
        self.scale = pencil_scale_of(stiffness)
        self.plan = symbolic.plan(device)
        self.stiffness_device = self.plan.matrix(stiffness)
        self.mass_device = self.plan.matrix(mass)
        self.device_matrix = self.plan.matrix(stiffness)  # S·(K − σM)·S of the last shift
        self.last = None
        self.sturm = SturmCheck()
        self._scale_device = None

    @property
    def dimension(self):
        return self.stiffness.n

    def transform(self):
        return self._transform

    def factorise(self, shift):
        """ShiftedPencil::factorise(σ): the factor of S·(K − σM)·S (formed on the device), kept for the solves; a
        singular factor is returned, not raised (is_singular); its Sturm count is added to the check."""
        self.device_matrix.set_combination(1.0, self.stiffness_device, -float(shift), self.mass_device, self.scale)
        factor = SupernodalFactor.factorise(self.symbolic, None, FactorOptions(pivoting=PivotingMode.INDEFINITE),
                                            device_matrix=self.device_matrix)
        result = ShiftFactorisation(float(shift), factor, self.scale)
        self.last = result
        self.sturm.add_point(SturmPoint(float(shift), result.count_below))
        return result

    def _scale_on(self, device):
        import torch
        if self._scale_device is None or self._scale_device.device != device:
            self._scale_device = torch.tensor(self.scale, device=device)
        return self._scale_device

    def solve(self, factorisation, block):
        """(K − σM)⁻¹·block = S·(S·(K − σM)·S)⁻¹·S·block with `factorisation` (the last one: ShiftedPencil::solve),
        block n × k: a device tensor (solved on the device) or a host array."""
        if hasattr(block, "data_ptr"):
            s = self._scale_on(block.device)
            s = s.reshape(-1, 1) if block.dim() == 2 else s
            return s * FactorSolve(factorisation.factor).solve((s * block).contiguous())
        b = np.asarray(block, dtype=np.float64)
        s = self.scale.reshape(-1, 1) if b.ndim == 2 else self.scale
        return s * FactorSolve(factorisation.factor).solve(s * b)

    def multiply_mass(self, block):
        """M·block (ShiftedPencil::multiplyMass): on the device for a device tensor, on the host otherwise."""
        if hasattr(block, "data_ptr"):
            return self.mass_device.multiply(block)
        return self.mass.multiply(block)

    def multiply_inner_product(self, block):
        """W·block with W the inner product's matrix: M (vibration) or K (buckling) (multiplyInnerProduct)."""
        if self._transform == "buckling":
            if hasattr(block, "data_ptr"):
                return self.stiffness_device.multiply(block)
            return self.stiffness.multiply(block)
        return self.multiply_mass(block)

    def diagonal_rayleigh_bound(self):
        """min K_ii / M_ii over M_ii > 0, K_ii > 0 (vibration), min K_ii / |M_ii| over M_ii ≠ 0 (buckling); +∞ when
        there is none (diagonalRayleighBound)."""
        k = _diagonal(self.stiffness)
        m = _diagonal(self.mass)
        ok = (k > 0) & ((m != 0) if self._transform == "buckling" else (m > 0))
        return float(np.min(k[ok] / np.abs(m[ok]))) if np.any(ok) else float("inf")


def _diagonal(matrix):
    starts = matrix.offsets[:-1]
    has = matrix.offsets[1:] > starts
    first = np.where(has, starts, 0)
    is_diag = has & (matrix.rows[first] == np.arange(matrix.n))
    d = np.zeros(matrix.n)
    d[is_diag] = matrix.values[first[is_diag]]
    return d
