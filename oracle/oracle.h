/* ORACLE — test infrastructure only (see oracle.c). Never linked by the product. */
#ifndef OOS_ORACLE_H
#define OOS_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* The symbolic factor as the numeric phase reads it (factor order): supernode s has column_count[s] columns from
 * first_column[s] and the rows rows[row_start[s] .. row_start[s + 1]) below them, ascending; parent[s] = -1 at a
 * root. */
typedef struct orc_symbolic {
    int32_t n;
    int32_t supernode_count;
    const int32_t *first_column;
    const int32_t *column_count;
    const int32_t *parent;
    const int64_t *row_start;
    const int32_t *rows;
} orc_symbolic;

typedef struct orc_result {
    int32_t status;
    int32_t zero_pivot_count;
    int32_t first_zero_pivot;
    int32_t first_issue;
    int32_t first_issue_negative;
    int32_t largest_ratio_equation;
    int64_t issue_count;
    int64_t inertia[3];
    double largest_ratio;
} orc_result;

int64_t orc_ldlt_nopivot(double *a, int64_t lda, int64_t n, int64_t c);
int64_t orc_front_nopivot(double *panel, int64_t ldp, double *contribution, int64_t ldc, int64_t c, int64_t r);
int64_t orc_bunch_kaufman(double *a, int64_t lda, int64_t n, int64_t candidates, int64_t *permutation,
                          double *subdiagonal);
void orc_inertia(const double *diag, const double *subdiag, int64_t n, int64_t *inertia);
void orc_spmv(int32_t n, const int64_t *offsets, const int32_t *rows, const double *values, const double *x,
              double *y);
void orc_multifrontal_pd(const orc_symbolic *sym, const int64_t *offsets, const int32_t *rows, const double *values,
                         double max_ratio, double *panels, double *diag, orc_result *out);
void orc_solve(const orc_symbolic *sym, const double *panels, double *x, int64_t ldx, int64_t nrhs);
void orc_multifrontal_partial(const orc_symbolic *sym, const int64_t *offsets, const int32_t *rows,
                              const double *values, double max_ratio, double *panels, double *diag, double *schur,
                              orc_result *out);
void orc_partial_condense(const orc_symbolic *sym, const double *panels, double *x, int64_t ldx, int64_t nrhs);
int64_t orc_csc_from_blocks(int32_t nodes, const int64_t *first_dof, const int64_t *column_start,
                            const int32_t *block_rows, const int64_t *block_value_offsets, const double *values,
                            const int32_t *equation_of_dof, int32_t kept, int64_t *offsets, int32_t *rows,
                            double *csc_values);
void orc_partial_recover(const orc_symbolic *sym, const double *panels, double *x, int64_t ldx, int64_t nrhs,
                         const double *trailing_values, int64_t ldt);

#ifdef __cplusplus
}
#endif

#endif
