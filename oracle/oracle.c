/*
 * ORACLE — test infrastructure only. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * liboracle.so, and only as the checker; the product path never links or calls it.
 *
 * A plain-C restatement of the reference's numeric LDLᵀ path (jcdulas/O_o_Solver, src/linalg), FP64, compiled with
 * -ffp-contract=off so that every product is rounded before it is added, as the reference family of the dense kernels
 * is (cmake/OosKernels.cmake:15-40). Each function names the reference lines it follows.
 *
 * What pins each function (tests/test_oracle.py):
 *   orc_ldlt_nopivot, orc_front_nopivot   bit-equal to the reference family (oracle/_ref) on random SPD fronts;
 *                                         the tridiagonal [-1, 2, -1] closed form d_k = (k + 1)/k; reconstruction.
 *   orc_bunch_kaufman                     bit-equal to the reference family, with and without delayed candidates;
 *                                         inertia of diag(±1) and of [[0, 1], [1, 0]]; reconstruction P A Pᵀ = L D Lᵀ.
 *   orc_multifrontal_pd                   panels bit-equal to the reference's factor (reference family) on plates
 *                                         and grids; inertia and ratio-test known answers of
 *                                         checks/linalg/FactorEdgeChecks.cpp:151-189; fronts equal to the dense
 *                                         Schur complements on integer matrices.
 *   orc_solve                             residual of manufactured solutions; within 1e-12 of the reference's solve.
 *   orc_spmv                              bit-equal to the reference's SymmetricCsc::multiply; dense product.
 *   orc_multifrontal_partial              K* and the interior panels bit-equal to the reference's PartialFactor;
 *                                         K* equal to the dense K_bb − K_bi·K_ii⁻¹·K_ib.
 *   orc_partial_condense, _recover        within 1e-12 of the reference's condenseLoads and recover; the dense
 *                                         f_b − K_bi·K_ii⁻¹·f_i and K_ii⁻¹·(f_i − K_ib·u_b).
 *   orc_csc_from_blocks                   bit-equal to the reference's SymmetricCsc::fromBlocks (random graphs, DOF
 *                                         counts, eliminations and orders); the dense matrix of the blocks.
 */

#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---- dense kernels ------------------------------------------------------------------------------------------- */

/* Right-looking LDLᵀ without pivoting of the n × n lower-stored `a`, columns [0, c) eliminated: the outer-product
 * form of src/linalg/kernels/reference/DenseBlockKernels.cpp:368-393. Element (i, j) receives the terms k = 0, 1, ...
 * in increasing order, each L(i, k)·W(j, k) with W(j, k) = L(j, k)·d_k. Returns the columns eliminated (the index of
 * the first zero or non-finite pivot). */
int64_t orc_ldlt_nopivot(double *a, int64_t lda, int64_t n, int64_t c)
{
    for (int64_t k = 0; k < c; ++k) {
        const double pivot = a[k + k * lda];
        if (pivot == 0.0 || !isfinite(pivot))
            return k;
        double *l = a + k * lda;
        for (int64_t i = k + 1; i < n; ++i)
            l[i] = l[i] / pivot;
        for (int64_t j = k + 1; j < n; ++j) {
            const double w = l[j] * pivot;
            double *t = a + j * lda;
            for (int64_t i = j; i < n; ++i)
                t[i] -= l[i] * w;
        }
    }
    return c;
}

/* The same factorisation on a front stored in two parts (the reference's SplitLdltOperands,
 * include/oos/linalg/kernels/DenseKernelEntryPoints.hpp:197-210; factorFrontWithoutPivoting,
 * DenseBlockKernels.cpp:395-446): element (i, j) of the front is panel[i + j·ldp] for j < c, else
 * contribution[(i − c) + (j − c)·ldc]. The trailing block is left untouched when a pivot fails (the reference
 * applies the trailing terms only once every column is eliminated). */
int64_t orc_front_nopivot(double *panel, int64_t ldp, double *contribution, int64_t ldc, int64_t c, int64_t r)
{
    const int64_t n = c + r;
    for (int64_t k = 0; k < c; ++k) {
        const double pivot = panel[k + k * ldp];
        if (pivot == 0.0 || !isfinite(pivot))
            return k;
        double *l = panel + k * ldp;
        for (int64_t i = k + 1; i < n; ++i)
            l[i] = l[i] / pivot;
        for (int64_t j = k + 1; j < c; ++j) {
            const double w = l[j] * pivot;
            double *t = panel + j * ldp;
            for (int64_t i = j; i < n; ++i)
                t[i] -= l[i] * w;
        }
    }
    if (contribution == NULL)
        return c;
    /* The trailing block: its element (i, j) receives the terms k = 0 .. c − 1 in order, the products of the final
     * L(i, k) and W(j, k) = L(j, k)·d_k, as in the unblocked algorithm. */
    for (int64_t k = 0; k < c; ++k) {
        const double pivot = panel[k + k * ldp];
        const double *l = panel + k * ldp + c;
        for (int64_t j = 0; j < r; ++j) {
            const double w = l[j] * pivot;
            double *t = contribution + j * ldc;
            for (int64_t i = j; i < r; ++i)
                t[i] -= l[i] * w;
        }
    }
    return c;
}

/* The 2 × 2 inverse in the form of LAPACK dsytf2 (src/linalg/dense/LdltArithmetic.cpp:50-56). */
static void invert_two_by_two(double a, double b, double c, double *d11, double *d22, double *s)
{
    *d11 = c / b;
    *d22 = a / b;
    const double t = 1.0 / (*d11 * *d22 - 1.0);
    *s = t / b;
}

/* The symmetric interchange of rows and columns s < t of the lower-stored `a`
 * (DenseBlockKernels.cpp:69-79). */
static void swap_symmetric(double *a, int64_t lda, int64_t n, int64_t s, int64_t t)
{
#define SWAP(x, y)                                                                                                     \
    do {                                                                                                               \
        const double tmp_ = (x);                                                                                       \
        (x) = (y);                                                                                                     \
        (y) = tmp_;                                                                                                    \
    } while (0)
    for (int64_t j = 0; j < s; ++j)
        SWAP(a[s + j * lda], a[t + j * lda]);
    SWAP(a[s + s * lda], a[t + t * lda]);
    for (int64_t i = s + 1; i < t; ++i)
        SWAP(a[i + s * lda], a[t + i * lda]);
    for (int64_t i = t + 1; i < n; ++i)
        SWAP(a[i + s * lda], a[i + t * lda]);
#undef SWAP
}

/* Bunch–Kaufman LDLᵀ restricted to the first `candidates` columns (DenseBlockKernels.cpp:448-580): α = (1 + √17)/8
 * (include/oos/linalg/kernels/DenseKernelEntryPoints.hpp:234), a candidate whose partner row is not a candidate is
 * pivoted alone when |a_kk| >= 0.01·colmax (hpp:241) and delayed otherwise (swapped with the last untried candidate);
 * passes are retried over the delayed candidates while a pass eliminates something. permutation[k] = the original
 * index at position k; subdiagonal[k] = e_k of a 2 × 2 block starting at k. Returns the columns eliminated. */
int64_t orc_bunch_kaufman(double *a, int64_t lda, int64_t n, int64_t candidates, int64_t *permutation,
                          double *subdiagonal)
{
    const double alpha = 0.6403882032022076;
    const double threshold = 0.01;
#define AT(i, j) a[(i) + (j) * lda]
    for (int64_t i = 0; i < candidates; ++i) {
        permutation[i] = i;
        subdiagonal[i] = 0.0;
    }
    int64_t k = 0, end = candidates, passStart = 0;
    for (;;) {
        while (k < end) {
            const double absakk = fabs(AT(k, k));
            double colmax = 0.0;
            int64_t imax = k;
            for (int64_t i = k + 1; i < n; ++i)
                if (fabs(AT(i, k)) > colmax) {
                    colmax = fabs(AT(i, k));
                    imax = i;
                }
            if (absakk == 0.0 && colmax == 0.0) {
                k += 1;
                continue;
            }
            int64_t blockSize = 1, swapWith = k;
            if (absakk < alpha * colmax && imax >= end) {
                if (absakk < threshold * colmax) {
                    --end;
                    if (k != end) {
                        swap_symmetric(a, lda, n, k, end);
                        const int64_t tmp = permutation[k];
                        permutation[k] = permutation[end];
                        permutation[end] = tmp;
                    }
                    continue;
                }
            } else if (absakk < alpha * colmax) {
                double rowmax = 0.0;
                for (int64_t j = k; j < imax; ++j)
                    rowmax = fmax(rowmax, fabs(AT(imax, j)));
                for (int64_t i = imax + 1; i < n; ++i)
                    rowmax = fmax(rowmax, fabs(AT(i, imax)));
                if (absakk >= alpha * colmax * (colmax / rowmax))
                    swapWith = k;
                else if (fabs(AT(imax, imax)) >= alpha * rowmax)
                    swapWith = imax;
                else {
                    swapWith = imax;
                    blockSize = 2;
                }
            }
            const int64_t moved = k + blockSize - 1;
            if (swapWith != moved) {
                swap_symmetric(a, lda, n, moved, swapWith);
                const int64_t tmp = permutation[moved];
                permutation[moved] = permutation[swapWith];
                permutation[swapWith] = tmp;
            }
            if (blockSize == 1) {
                const double pivot = AT(k, k);
                for (int64_t i = k + 1; i < n; ++i)
                    AT(i, k) = AT(i, k) / pivot;
                for (int64_t j = k + 1; j < n; ++j) {
                    const double w = AT(j, k) * pivot;
                    for (int64_t i = j; i < n; ++i)
                        AT(i, j) -= AT(i, k) * w;
                }
            } else {
                const double d0 = AT(k, k), e = AT(k + 1, k), d1 = AT(k + 1, k + 1);
                double d11, d22, s;
                invert_two_by_two(d0, e, d1, &d11, &d22, &s);
                for (int64_t i = k + 2; i < n; ++i) {
                    const double r0 = AT(i, k), r1 = AT(i, k + 1);
                    AT(i, k) = s * (d11 * r0 - r1);
                    AT(i, k + 1) = s * (d22 * r1 - r0);
                }
                subdiagonal[k] = e;
                AT(k + 1, k) = 0.0;
                for (int64_t j = k + 2; j < n; ++j) {
                    const double w0 = AT(j, k) * d0 + AT(j, k + 1) * e;
                    const double w1 = AT(j, k) * e + AT(j, k + 1) * d1;
                    for (int64_t i = j; i < n; ++i) {
                        AT(i, j) -= AT(i, k) * w0;
                        AT(i, j) -= AT(i, k + 1) * w1;
                    }
                }
            }
            k += blockSize;
        }
        if (end == candidates || k == passStart)
            break;
        passStart = k;
        end = candidates;
    }
#undef AT
    return k;
}

/* The inertia of a block diagonal D (1 × 1 by sign; 2 × 2 by the sign of the determinant, then the trace):
 * src/linalg/supernodal/MultifrontalDriver.cpp:764-795, LdltArithmetic.cpp:70-108. */
void orc_inertia(const double *diag, const double *subdiag, int64_t n, int64_t *inertia)
{
    int64_t k = 0;
    while (k < n) {
        const double pivot = diag[k];
        if (subdiag != NULL && k + 1 < n && subdiag[k] != 0.0) {
            const double second = diag[k + 1];
            const double det = pivot * second - subdiag[k] * subdiag[k];
            if (det < 0.0) {
                ++inertia[0];
                ++inertia[1];
            } else if (det > 0.0) {
                inertia[pivot > 0.0 ? 0 : 1] += 2;
            } else {
                ++inertia[2];
                const double trace = pivot + second;
                ++inertia[trace > 0.0 ? 0 : trace < 0.0 ? 1 : 2];
            }
            k += 2;
            continue;
        }
        ++inertia[pivot > 0.0 ? 0 : pivot < 0.0 ? 1 : 2];
        k += 1;
    }
}

/* ---- sparse ---------------------------------------------------------------------------------------------------- */

/* y = A·x with A symmetric stored as its lower CSC, the reference's serial order (src/linalg/sparse/
 * SymmetricCsc.cpp:479-495): y[i] += a_ij·x_j, then y[j] += a_ij·x_i for i != j, column after column. */
void orc_spmv(int32_t n, const int64_t *offsets, const int32_t *rows, const double *values, const double *x, double *y)
{
    for (int32_t i = 0; i < n; ++i)
        y[i] = 0.0;
    for (int32_t j = 0; j < n; ++j)
        for (int64_t k = offsets[j]; k < offsets[j + 1]; ++k) {
            const int32_t i = rows[k];
            y[i] += values[k] * x[j];
            if (i != j)
                y[j] += values[k] * x[i];
        }
}

/* ---- multifrontal factorisation (positive definite) ------------------------------------------------------------ */

/* Position of `row` in the front of supernode s (its columns first, then its rows below), -1 when absent. */
static int64_t front_position(const orc_symbolic *sym, int32_t s, int32_t row)
{
    const int32_t first = sym->first_column[s], c = sym->column_count[s];
    if (row >= first && row < first + c)
        return row - first;
    const int32_t *r = sym->rows + sym->row_start[s];
    int64_t lo = 0, hi = sym->row_start[s + 1] - sym->row_start[s];
    while (lo < hi) {
        const int64_t mid = (lo + hi) / 2;
        if (r[mid] < row)
            lo = mid + 1;
        else
            hi = mid;
    }
    if (lo < sym->row_start[s + 1] - sym->row_start[s] && r[lo] == row)
        return c + lo;
    return -1;
}

/* The multifrontal LDLᵀ of a positive definite matrix (PivotingMode::PositiveDefinite), supernode after supernode
 * in their order (the reference's orders all give the same bits: include/oos/linalg/supernodal/
 * SupernodalFactor.hpp:41-53): the front of s is cleared, receives the CSC entries of its columns, then the update
 * matrices of its children in increasing child order (src/linalg/supernodal/FrontAssembly.cpp:166-300); its columns
 * are eliminated in place, panel n × c (leading dimension n) and contribution r × r (MultifrontalDriver.cpp:257-284,
 * 315-344). The panel's entries above the diagonal are zero. Then the zero-pivot test, the inertia and the ratio test
 * of MultifrontalDriver::finish (MultifrontalDriver.cpp:704-820).
 *
 * panels: sized Σ c·(c + r), the symbolic shape (panel_offset given by the caller's cumulative sum).
 * out->status: 0 done; 1 a matrix entry outside the structure (std::invalid_argument); 3 a zero or non-finite pivot
 * (SingularityError), out->zero_pivot_count zero pivots, the first at out->first_zero_pivot.
 * diag (n values): d of each equation. */
static void multifrontal(const orc_symbolic *sym, const int64_t *offsets, const int32_t *rows, const double *values,
                         double max_ratio, double *panels, double *diag, orc_result *out, int32_t trailing,
                         double *schur)
{
    const int32_t S = sym->supernode_count;
    memset(out, 0, sizeof(*out));
    out->first_zero_pivot = -1;
    out->first_issue = -1;
    int64_t *panelOffset = (int64_t *)calloc((size_t)S + 1, sizeof(int64_t));
    double **updates = (double **)calloc((size_t)S, sizeof(double *));
    int8_t *status = (int8_t *)calloc((size_t)S, 1); /* 0 pending, 1 done, 2 zero pivot, 3 skipped */
    int32_t *zeroAt = (int32_t *)calloc((size_t)S, sizeof(int32_t));
    /* children lists in increasing order */
    int32_t *childCount = (int32_t *)calloc((size_t)S + 1, sizeof(int32_t));
    for (int32_t s = 0; s < S; ++s) {
        const int64_t c = sym->column_count[s], r = sym->row_start[s + 1] - sym->row_start[s];
        panelOffset[s + 1] = panelOffset[s] + c * (c + r);
        if (sym->parent[s] >= 0)
            ++childCount[sym->parent[s] + 1];
    }
    for (int32_t s = 0; s < S; ++s)
        childCount[s + 1] += childCount[s];
    int32_t *children = (int32_t *)malloc(sizeof(int32_t) * (size_t)(childCount[S] > 0 ? childCount[S] : 1));
    int32_t *fill = (int32_t *)calloc((size_t)S, sizeof(int32_t));
    for (int32_t s = 0; s < S; ++s)
        if (sym->parent[s] >= 0) {
            const int32_t p = sym->parent[s];
            children[childCount[p] + fill[p]++] = s;
        }

    for (int32_t s = 0; s < S && out->status != 1; ++s) {
        const int32_t c = sym->column_count[s], first = sym->first_column[s];
        const int64_t r = sym->row_start[s + 1] - sym->row_start[s];
        const int64_t n = c + r;
        int ready = 1;
        for (int32_t k = childCount[s]; k < childCount[s + 1]; ++k)
            ready = ready && status[children[k]] == 1;
        if (!ready) {
            status[s] = 3;
            for (int32_t k = childCount[s]; k < childCount[s + 1]; ++k) {
                free(updates[children[k]]);
                updates[children[k]] = NULL;
            }
            continue;
        }
        double *panel = panels + panelOffset[s];
        double *update = r > 0 ? (double *)calloc((size_t)(r * r), sizeof(double)) : NULL;
        memset(panel, 0, sizeof(double) * (size_t)(n * c));
        /* the matrix entries of its columns */
        for (int32_t j = 0; j < c && out->status != 1; ++j) {
            const int32_t column = first + j;
            for (int64_t k = offsets[column]; k < offsets[column + 1]; ++k) {
                const int32_t row = rows[k];
                const int64_t at = front_position(sym, s, row);
                if (row < column || at < 0) {
                    out->status = 1;
                    break;
                }
                panel[at + (int64_t)j * n] += values[k];
            }
        }
        if (out->status == 1) {
            free(update);
            break;
        }
        /* the children's update matrices, in increasing child order */
        for (int32_t q = childCount[s]; q < childCount[s + 1]; ++q) {
            const int32_t child = children[q];
            const int64_t rc = sym->row_start[child + 1] - sym->row_start[child];
            const int32_t *crows = sym->rows + sym->row_start[child];
            int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * (size_t)(rc > 0 ? rc : 1));
            for (int64_t i = 0; i < rc; ++i)
                pos[i] = front_position(sym, s, crows[i]);
            const double *u = updates[child];
            for (int64_t jj = 0; jj < rc; ++jj) {
                const int64_t pj = pos[jj];
                for (int64_t ii = jj; ii < rc; ++ii) {
                    const int64_t pi = pos[ii];
                    const double v = u[ii + jj * rc];
                    if (pj < c)
                        panel[pi + pj * n] += v;
                    else
                        update[(pi - c) + (pj - c) * r] += v;
                }
            }
            free(pos);
            free(updates[child]);
            updates[child] = NULL;
        }
        if (s == trailing) {
            /* The trailing front is K* itself (MultifrontalDriver::assembleTrailing, MultifrontalDriver.cpp:567-599):
             * handed over with both triangles, never eliminated. */
            for (int32_t j = 0; j < c; ++j)
                for (int32_t i = j; i < c; ++i) {
                    schur[i + (int64_t)j * c] = panel[i + (int64_t)j * n];
                    schur[j + (int64_t)i * c] = panel[i + (int64_t)j * n];
                }
            status[s] = 1;
            continue;
        }
        const int64_t eliminated = orc_front_nopivot(panel, n, update, r, c, r);
        if (eliminated < c) {
            status[s] = 2;
            zeroAt[s] = first + (int32_t)eliminated;
            free(update);
            continue;
        }
        /* zero above the diagonal (storePanel's convention) */
        for (int32_t j = 1; j < c; ++j)
            for (int32_t i = 0; i < j; ++i)
                panel[i + (int64_t)j * n] = 0.0;
        updates[s] = update;
        status[s] = 1;
    }

    if (out->status == 0) {
        for (int32_t s = 0; s < S; ++s)
            if (status[s] == 2) {
                if (out->zero_pivot_count == 0)
                    out->first_zero_pivot = zeroAt[s];
                ++out->zero_pivot_count;
            }
        if (out->zero_pivot_count > 0)
            out->status = 3;
    }
    if (out->status == 0) {
        /* the inertia and the ratio test, panel by panel in elimination order */
        double largest = 0.0;
        int32_t largestAt = -1;
        for (int32_t s = 0; s < S; ++s) {
            if (s == trailing)
                continue; /* not eliminated: the boundary's equations */
            const int32_t c = sym->column_count[s], first = sym->first_column[s];
            const int64_t n = c + sym->row_start[s + 1] - sym->row_start[s];
            for (int32_t k = 0; k < c; ++k) {
                const double pivot = panels[panelOffset[s] + k + (int64_t)k * n];
                const int32_t equation = first + k;
                if (diag != NULL)
                    diag[equation] = pivot;
                ++out->inertia[pivot > 0.0 ? 0 : pivot < 0.0 ? 1 : 2];
                double a = 0.0;
                if (offsets[equation] < offsets[equation + 1] && rows[offsets[equation]] == equation)
                    a = values[offsets[equation]];
                const double ratio = a / pivot;
                if (largestAt < 0 || ratio > largest) {
                    largest = ratio;
                    largestAt = equation;
                }
                const int negative = ratio < 0.0 || pivot < 0.0;
                if (ratio > max_ratio || negative) {
                    if (out->issue_count == 0) {
                        out->first_issue = equation;
                        out->first_issue_negative = negative;
                    }
                    ++out->issue_count;
                }
            }
        }
        out->largest_ratio = largest;
        out->largest_ratio_equation = largestAt;
    }
    for (int32_t s = 0; s < S; ++s)
        free(updates[s]);
    free(updates);
    free(panelOffset);
    free(status);
    free(zeroAt);
    free(childCount);
    free(children);
    free(fill);
}

void orc_multifrontal_pd(const orc_symbolic *sym, const int64_t *offsets, const int32_t *rows, const double *values,
                         double max_ratio, double *panels, double *diag, orc_result *out)
{
    multifrontal(sym, offsets, rows, values, max_ratio, panels, diag, out, -1, NULL);
}

/* PartialFactor::factorise (src/linalg/supernodal/PartialFactor.cpp:65-83): the same factorisation with the last
 * supernode (the trailing set: trailing_dofs columns, no rows below) assembled but not eliminated; its front, matrix
 * entries then the children's update matrices in increasing child order, is K* (trailing_dofs square, both
 * triangles, into schur). The inertia and the ratio test cover the interior equations only. */
void orc_multifrontal_partial(const orc_symbolic *sym, const int64_t *offsets, const int32_t *rows,
                              const double *values, double max_ratio, double *panels, double *diag, double *schur,
                              orc_result *out)
{
    const int32_t S = sym->supernode_count;
    multifrontal(sym, offsets, rows, values, max_ratio, panels, diag, out, S - 1, schur);
}

/* ---- solve ----------------------------------------------------------------------------------------------------- */

/* X ← A⁻¹X with the factor of orc_multifrontal_pd (positive definite, symbolic shape): the forward sweep
 * (X_F ← L_FF⁻¹X_F, X_R −= L_RF·X_F, supernodes in increasing order), the diagonal (X ← D⁻¹X) and the backward sweep
 * (X_F −= L_RFᵀ·X_R, X_F ← L_FF⁻ᵀX_F, supernodes in decreasing order): src/linalg/supernodal/FactorSolve.cpp:176-190,
 * SolveSweeps.cpp:272-294. x is n × nrhs column-major (leading dimension ldx). The chains differ from the reference's
 * (its subtree accumulators regroup the terms), so the solve agrees with it to rounding, not to the bit. */
static void sweep_forward(const orc_symbolic *sym, const double *panels, const int64_t *panelOffset, int32_t skip,
                          double *b)
{
    for (int32_t s = 0; s < sym->supernode_count; ++s) {
        if (s == skip)
            continue;
        const int32_t c = sym->column_count[s], first = sym->first_column[s];
        const int64_t r = sym->row_start[s + 1] - sym->row_start[s], n = c + r;
        const double *p = panels + panelOffset[s];
        const int32_t *rws = sym->rows + sym->row_start[s];
        for (int32_t q = 0; q < c; ++q) {
            const double xq = b[first + q];
            for (int32_t i = q + 1; i < c; ++i)
                b[first + i] -= p[i + (int64_t)q * n] * xq;
            for (int64_t i = 0; i < r; ++i)
                b[rws[i]] -= p[c + i + (int64_t)q * n] * xq;
        }
    }
}

static void sweep_diagonal(const orc_symbolic *sym, const double *panels, const int64_t *panelOffset, int32_t skip,
                           double *b)
{
    for (int32_t s = 0; s < sym->supernode_count; ++s) {
        if (s == skip)
            continue;
        const int32_t c = sym->column_count[s], first = sym->first_column[s];
        const int64_t n = c + sym->row_start[s + 1] - sym->row_start[s];
        const double *p = panels + panelOffset[s];
        for (int32_t q = 0; q < c; ++q)
            b[first + q] = b[first + q] / p[q + (int64_t)q * n];
    }
}

static void sweep_backward(const orc_symbolic *sym, const double *panels, const int64_t *panelOffset, int32_t skip,
                           double *b)
{
    for (int32_t s = sym->supernode_count - 1; s >= 0; --s) {
        if (s == skip)
            continue;
        const int32_t c = sym->column_count[s], first = sym->first_column[s];
        const int64_t r = sym->row_start[s + 1] - sym->row_start[s], n = c + r;
        const double *p = panels + panelOffset[s];
        const int32_t *rws = sym->rows + sym->row_start[s];
        for (int32_t q = c - 1; q >= 0; --q) {
            double sum = b[first + q];
            for (int64_t i = r - 1; i >= 0; --i)
                sum -= p[c + i + (int64_t)q * n] * b[rws[i]];
            for (int32_t i = c - 1; i > q; --i)
                sum -= p[i + (int64_t)q * n] * b[first + i];
            b[first + q] = sum;
        }
    }
}

static int64_t *panel_offsets(const orc_symbolic *sym)
{
    const int32_t S = sym->supernode_count;
    int64_t *panelOffset = (int64_t *)calloc((size_t)S + 1, sizeof(int64_t));
    for (int32_t s = 0; s < S; ++s) {
        const int64_t c = sym->column_count[s], r = sym->row_start[s + 1] - sym->row_start[s];
        panelOffset[s + 1] = panelOffset[s] + c * (c + r);
    }
    return panelOffset;
}

void orc_solve(const orc_symbolic *sym, const double *panels, double *x, int64_t ldx, int64_t nrhs)
{
    int64_t *panelOffset = panel_offsets(sym);
    for (int64_t col = 0; col < nrhs; ++col) {
        double *b = x + col * ldx;
        sweep_forward(sym, panels, panelOffset, -1, b);
        sweep_diagonal(sym, panels, panelOffset, -1, b);
        sweep_backward(sym, panels, panelOffset, -1, b);
    }
    free(panelOffset);
}

/* PartialFactor::condenseLoads and recover (src/linalg/supernodal/PartialFactor.cpp:117-158) with the factor of
 * orc_multifrontal_partial: the forward sweep over the interior supernodes leaves f* = f_b − K_bi·K_ii⁻¹·f_i in the
 * trailing rows (condense: stop there). Recover: the trailing rows then replaced by u_b (trailing_values, trailing ×
 * nrhs, leading dimension ldt), the interior's diagonal and backward sweeps give u_i = K_ii⁻¹·(f_i − K_ib·u_b). */
void orc_partial_condense(const orc_symbolic *sym, const double *panels, double *x, int64_t ldx, int64_t nrhs)
{
    int64_t *panelOffset = panel_offsets(sym);
    for (int64_t col = 0; col < nrhs; ++col)
        sweep_forward(sym, panels, panelOffset, sym->supernode_count - 1, x + col * ldx);
    free(panelOffset);
}

void orc_partial_recover(const orc_symbolic *sym, const double *panels, double *x, int64_t ldx, int64_t nrhs,
                         const double *trailing_values, int64_t ldt)
{
    const int32_t T = sym->supernode_count - 1;
    const int32_t first = sym->first_column[T], c = sym->column_count[T];
    int64_t *panelOffset = panel_offsets(sym);
    for (int64_t col = 0; col < nrhs; ++col) {
        double *b = x + col * ldx;
        sweep_forward(sym, panels, panelOffset, T, b);
        for (int32_t i = 0; i < c; ++i)
            b[first + i] = trailing_values[i + col * ldt];
        sweep_diagonal(sym, panels, panelOffset, T, b);
        sweep_backward(sym, panels, panelOffset, T, b);
    }
    free(panelOffset);
}

/* ---- CSC conversion ------------------------------------------------------------------------------------------ */

/* SymmetricCsc::fromBlocks (src/linalg/sparse/SymmetricCsc.cpp:41-200): the lower-triangle terms of the node blocks
 * (block b of column node j couples node block_rows[b] >= j; its values column-major at block_value_offsets[b],
 * dofs(row node) × dofs(j); a diagonal block read on and below its diagonal), the kept DOFs only
 * (equation_of_dof[d] >= 0), each term at (max(p, q), min(p, q)) of its equations; every column's rows ascending.
 * rows == NULL: offsets (kept + 1) only. Returns the terms. */
int64_t orc_csc_from_blocks(int32_t nodes, const int64_t *first_dof, const int64_t *column_start,
                            const int32_t *block_rows, const int64_t *block_value_offsets, const double *values,
                            const int32_t *equation_of_dof, int32_t kept, int64_t *offsets, int32_t *rows,
                            double *csc_values)
{
    for (int32_t q = 0; q <= kept; ++q)
        offsets[q] = 0;
    for (int pass = 0; pass < (rows != NULL ? 2 : 1); ++pass) {
        int64_t *cursor = NULL;
        if (pass == 1) {
            cursor = (int64_t *)malloc(sizeof(int64_t) * (size_t)(kept > 0 ? kept : 1));
            for (int32_t q = 0; q < kept; ++q)
                cursor[q] = offsets[q];
        }
        for (int32_t j = 0; j < nodes; ++j) {
            const int64_t fj = first_dof[j];
            const int cj = (int)(first_dof[j + 1] - fj);
            for (int64_t b = column_start[j]; b < column_start[j + 1]; ++b) {
                const int32_t node = block_rows[b];
                const int64_t fr = first_dof[node];
                const int cr = (int)(first_dof[node + 1] - fr);
                const double *v = values + block_value_offsets[b];
                for (int c = 0; c < cj; ++c) {
                    const int32_t q = equation_of_dof[fj + c];
                    if (q < 0)
                        continue;
                    for (int r = node == j ? c : 0; r < cr; ++r) {
                        const int32_t p = equation_of_dof[fr + r];
                        if (p < 0)
                            continue;
                        const int32_t column = p < q ? p : q, row = p < q ? q : p;
                        if (pass == 0)
                            ++offsets[column + 1];
                        else {
                            const int64_t at = cursor[column]++;
                            rows[at] = row;
                            csc_values[at] = v[(int64_t)c * cr + r];
                        }
                    }
                }
            }
        }
        if (pass == 0)
            for (int32_t q = 0; q < kept; ++q)
                offsets[q + 1] += offsets[q];
        free(cursor);
    }
    if (rows != NULL)
        for (int32_t q = 0; q < kept; ++q)
            /* insertion sort of the column by row (each row once) */
            for (int64_t a = offsets[q] + 1; a < offsets[q + 1]; ++a) {
                const int32_t row = rows[a];
                const double value = csc_values[a];
                int64_t b = a - 1;
                while (b >= offsets[q] && rows[b] > row) {
                    rows[b + 1] = rows[b];
                    csc_values[b + 1] = csc_values[b];
                    --b;
                }
                rows[b + 1] = row;
                csc_values[b + 1] = value;
            }
    return offsets[kept];
}
