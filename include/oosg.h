/*
 * oosg — the numeric phase of the supernodal multifrontal LDLᵀ (factor and solve) on one AMD Instinct MI355X
 * (gfx950), behind a plain C ABI: plain pointers and sizes, status codes, no exceptions across the boundary.
 *
 * It replaces, for a caller of jcdulas/O_o_Solver, the numeric phase behind:
 *   SupernodalFactor::factorise      include/oos/linalg/supernodal/SupernodalFactor.hpp:176-183
 *   SupernodalFactor accessors       include/oos/linalg/supernodal/SupernodalFactor.hpp:192-233
 *   FactorSolve::solve               include/oos/linalg/supernodal/FactorSolve.hpp:77-81
 *   SymmetricCsc (the input)         include/oos/linalg/sparse/SymmetricCsc.hpp:14-108
 * The ordering and the symbolic analysis stay on the host (SymbolicFactor::analyse,
 * include/oos/linalg/symbolic/SymbolicFactor.hpp:47-50): their result is this library's structural input.
 *
 * Status codes map onto the reference's exceptions (include/oos/base/errors/): OOSG_INVALID_ARGUMENT is
 * std::invalid_argument, OOSG_OUT_OF_MEMORY is MemoryBudgetError (raised before any numeric work), OOSG_SINGULAR is
 * SingularityError, OOSG_INTERNAL is InternalError. oosg_last_error() gives the message of the last failure on the
 * calling thread.
 *
 * Ownership: the caller owns every array it passes; the library copies what it keeps. Handles own device memory.
 * Threading: a plan and the matrices and factors made with it belong to one device and one stream; calls on one plan
 * are serialised by the caller (distinct plans may be used from distinct threads). A factor is immutable once made.
 * Streams: the library works on streams of its own. A call that reads device memory the caller gives first waits for
 * the work queued on the caller's stream (oosg_plan_set_stream; the null stream by default), and every call returns
 * only when its own device work is done, so its outputs may be used on any stream at once.
 */
#ifndef OOSG_H
#define OOSG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum oosg_status {
    OOSG_OK = 0,
    OOSG_INVALID_ARGUMENT = 1,
    OOSG_OUT_OF_MEMORY = 2,
    OOSG_SINGULAR = 3,
    OOSG_CANCELLED = 4,
    OOSG_INTERNAL = 5,
    OOSG_HIP_ERROR = 6,
    OOSG_NOT_CONVERGED = 7,
    /* An FP32 factor met a zero or non-finite pivot, or failed the ratio test, where single precision may be the cause:
     * the matrix is not shown singular; factorise it in FP64 (which decides, with OOSG_SINGULAR). */
    OOSG_PRECISION_INSUFFICIENT = 8
} oosg_status;

/* The symbolic factor, factor order (SymbolicFactor, include/oos/linalg/symbolic/SymbolicFactor.hpp:37-125):
 * supernode s owns the columns [first_column[s], first_column[s] + column_count[s]), its rows below them are
 * rows[row_start[s] .. row_start[s + 1]) (ascending), parent[s] > s or -1 at a root; supernodes in postorder. */
typedef struct oosg_symbolic_desc {
    int32_t n;
    int32_t supernode_count;
    const int32_t *first_column;
    const int32_t *column_count;
    const int32_t *parent;
    const int64_t *row_start;
    const int32_t *rows;
} oosg_symbolic_desc;

/* A symmetric matrix as its lower triangle by columns, each column's rows ascending from its diagonal, structural
 * zeros kept (SymmetricCsc, include/oos/linalg/sparse/SymmetricCsc.hpp:14-20). Host pointers. */
typedef struct oosg_csc_desc {
    int32_t n;
    const int64_t *offsets;
    const int32_t *rows;
    const double *values;
} oosg_csc_desc;

typedef enum oosg_pivoting {
    OOSG_POSITIVE_DEFINITE = 0, /* PivotingMode::PositiveDefinite (SupernodalFactor.hpp:34-42) */
    /* PivotingMode::Indefinite, FP64 only: Bunch-Kaufman pivoting among each front's candidates (its columns and
     * those its children delayed), 1 x 1 and 2 x 2 pivots, candidates that cannot be eliminated delayed to the parent
     * (MultifrontalDriver.cpp:346-436, DenseBlockKernels.cpp:448-580). The inertia counts 2 x 2 blocks by the sign of
     * their determinant. No ratio test, as in the reference. A root front that would delay is OOSG_INTERNAL. */
    OOSG_INDEFINITE = 1
} oosg_pivoting;

typedef enum oosg_precision {
    OOSG_FP64 = 0, /* every front in double precision */
    OOSG_FP32 = 1, /* fronts in single precision on the diagonally scaled matrix; solves refine in double */
    OOSG_MIXED = 2 /* the fronts of the lowest fp32_levels levels (the subtrees below a cut) in single precision, the
                      others in double, all on the diagonally scaled matrix; solves refine in double */
} oosg_precision;

typedef struct oosg_factor_options {
    int32_t pivoting;  /* oosg_pivoting */
    int32_t precision; /* oosg_precision */
    double max_ratio;  /* PARAM MaxDiagonalRatio (PivotPolicy, PivotDiagnostics.hpp:34-44); 0: 1e7 */
    int32_t bailout;   /* PARAM SingularityBailOut: < 0 keeps a factor that fails the ratio test */
    /* The diagonal the ratio test divides by d_j: n values (host memory, factor order; a negative value keeps the
     * matrix diagonal of that equation) or null for the matrix diagonal everywhere (FactorOptions::referenceDiagonal,
     * SupernodalFactor.hpp:75; ratioTestDiagonal, PivotDiagnostics.hpp:108-120). */
    const double *reference_diagonal;
    /* OOSG_MIXED: the levels (heights in the supernode tree, from the leaves) factorised in FP32; clamped to the
     * tree's levels. */
    int32_t fp32_levels;
} oosg_factor_options;

/* The inertia, pivot counts and ratio test of a factor (SupernodalFactor::inertia, pivotDiagnostics,
 * SupernodalFactor.hpp:192-205; MultifrontalDriver::finish, src/linalg/supernodal/MultifrontalDriver.cpp:704-820). */
typedef struct oosg_factor_info {
    int64_t positive, negative, zero;
    int64_t delayed_pivots;
    int64_t two_by_two_pivots;
    int64_t issue_count;          /* pivots failing the ratio test */
    int32_t first_issue_equation; /* -1 when none */
    int32_t first_issue_negative; /* 1: a negative pivot */
    double largest_ratio;
    int32_t largest_ratio_equation;
    int32_t precision;
    double factor_seconds;        /* device time of the numeric factorisation */
    double flops;                 /* SymbolicFactor::flopCount() of the structure */
    int32_t fp32_levels;          /* OOSG_MIXED: the levels in FP32 */
} oosg_factor_info;

/* What a refined solve did. */
typedef struct oosg_solve_report {
    int32_t steps;             /* solves with the factor */
    int32_t converged;         /* 1: the backward error reached the target */
    double backward_error;     /* max over columns of ‖b − A·x‖∞ / (‖A‖∞·‖x‖∞ + ‖b‖∞) */
    double seconds;            /* device time of the solve */
} oosg_solve_report;

typedef struct oosg_plan oosg_plan;
typedef struct oosg_matrix oosg_matrix;
typedef struct oosg_factor oosg_factor;

/* The message of the last failure on this thread. */
const char *oosg_last_error(void);
/* The equations the last OOSG_SINGULAR (or OOSG_PRECISION_INSUFFICIENT) on this thread names, in elimination order:
 * the zero pivots, or the pivots failing the ratio test; at most OOSG_RECORDED_ISSUE_COUNT are kept, *count receives
 * how many (the DOFs of SingularityError::dofs(), include/oos/base/errors/SingularityError.hpp:43-50; the cap of
 * PivotDiagnostics::RECORDED_ISSUE_COUNT, include/oos/linalg/supernodal/PivotDiagnostics.hpp:70-75). Up to
 * `capacity` of them are written to `equations` (may be null). */
#define OOSG_RECORDED_ISSUE_COUNT 1000
oosg_status oosg_last_error_equations(int32_t *equations, int64_t capacity, int64_t *count);
/* What the last OOSG_OUT_OF_MEMORY on this thread asked for (MemoryShortfall of MemoryBudgetError,
 * include/oos/base/errors/MemoryBudgetError.hpp:24-36): the bytes requested, those left in the budget, the budget. */
typedef struct oosg_memory_shortfall {
    uint64_t requested_bytes;
    uint64_t available_bytes;
    uint64_t limit_bytes;
} oosg_memory_shortfall;
oosg_status oosg_last_error_shortfall(oosg_memory_shortfall *shortfall);

/* The plan of the numeric phase for one symbolic factor on `device`: the supernode tree by levels, the relative
 * indices of the extend-add, the schedule of the kernels and the device memory (panels, update matrices), all
 * reserved before any numeric work (the memory-budget contract of SupernodalFactor::factorise,
 * SupernodalFactor.hpp:130-135). hbm_budget_bytes = 0: the device's free memory. A plan serves every factorisation of
 * matrices of its structure. */
oosg_status oosg_plan_create(const oosg_symbolic_desc *symbolic, int32_t device, uint64_t hbm_budget_bytes,
                             oosg_plan **out);
/* The device memory oosg_plan_create reserves for `symbolic` (SupernodalFactor::estimateMemoryBytes,
 * SupernodalFactor.hpp:185-189): the structure, the update matrices, one factor's panels, inverses and pivots, and the
 * solves' and refinements' vectors for 16 right-hand sides. In indefinite mode delayed pivots grow fronts and panels
 * beyond it (reserved when they happen). No device is touched. */
oosg_status oosg_estimate_memory_bytes(const oosg_symbolic_desc *symbolic, int32_t pivoting, uint64_t *bytes);
/* The bytes of the plan's budget in use now (the plan's own memory and that of its live factors) and the budget. */
oosg_status oosg_plan_memory(const oosg_plan *plan, uint64_t *reserved_bytes, uint64_t *budget_bytes);
/* The caller's stream (a hipStream_t; null: the null stream): calls that read the caller's device memory wait for the
 * work queued on it before their first launch. */
oosg_status oosg_plan_set_stream(oosg_plan *plan, void *stream);
/* Cancellation (FactorOptions::cancellation, SupernodalFactor.hpp:78-80; base::CancellationToken): `cancelled(context)`
 * is polled before a factorisation, between its levels where it is launched one launch at a time, before it is
 * finished, before each solve of a refinement and between the 16-column tiles of a solve; a non-zero answer stops the
 * call with OOSG_CANCELLED once the work already queued is done (nothing is left half written: no factor is
 * returned). Null: no polling. */
typedef int32_t (*oosg_cancel_fn)(void *context);
oosg_status oosg_plan_set_cancellation(oosg_plan *plan, oosg_cancel_fn cancelled, void *context);
void oosg_plan_destroy(oosg_plan *plan);
/* {supernodes, levels, panel values, update values, small fronts, large fronts, kernel launches per factorisation,
 * MFMA update tiles}: 8 values. */
oosg_status oosg_plan_stats(const oosg_plan *plan, int64_t *stats);

/* A matrix of the plan's structure copied into device memory (SymmetricCsc::fromCompressedColumns,
 * SymmetricCsc.hpp:66-72). OOSG_INVALID_ARGUMENT when an entry lies outside the structure of the symbolic factor
 * (the test of src/linalg/supernodal/FrontAssembly.cpp:22-28). */
oosg_status oosg_matrix_create(oosg_plan *plan, const oosg_csc_desc *matrix, oosg_matrix **out);
/* New values for the same structure (the shifted matrices K − σB of an eigenvalue solver). */
oosg_status oosg_matrix_set_values(oosg_matrix *matrix, const double *values);
/* New values for `matrix` from two matrices of its plan and structure, on the device: SymmetricCsc::combine then
 * scaleSymmetric (src/linalg/sparse/SymmetricCsc.cpp:209-252), the shifted pencil S·(K − σ·M)·S of an eigenvalue
 * solver (BucklingPencil.cpp:74-97): each value is scale[row]·(alpha·a + beta·b)·scale[column], the two products,
 * the sum and the two scalings each rounded in that order (no fused multiply-add): the reference's bits. scale: n
 * values (host memory) or null for none. matrix may be a or b. The derived arrays (the full rows of the residual
 * products, the diagonal, ‖A‖∞, the FP32 scales) are recomputed on the device; nothing passes through the host but
 * the diagonal. OOSG_INVALID_ARGUMENT when the structures differ (the message of SymmetricCsc::combine). */
oosg_status oosg_matrix_set_combination(oosg_matrix *matrix, double alpha, const oosg_matrix *a, double beta,
                                        const oosg_matrix *b, const double *scale);
void oosg_matrix_destroy(oosg_matrix *matrix);

/* ---- the CSC of a block matrix on the device (SymmetricCsc::fromBlocks) ------------------------------------------ */

/* A matrix assembled at node granularity (SparsePattern + BlockSparseMatrix, include/oos/linalg/sparse/
 * SparsePattern.hpp:13-22, BlockSparseMatrix.hpp:14-36), in device memory. */
typedef struct oosg_block_matrix_desc {
    int32_t node_count;
    const int64_t *first_dof;           /* node_count + 1: node n owns the DOFs [first_dof[n], first_dof[n + 1]), 1-6 */
    const int64_t *column_start;        /* node_count + 1: the blocks of node column j */
    const int32_t *block_rows;          /* per block, its row node: the diagonal block first, then rows below ascending */
    const int64_t *block_value_offsets; /* blocks + 1: block b column-major, dofs(its row) × dofs(column), in values */
    const double *values;
} oosg_block_matrix_desc;

/* SymmetricCsc::fromBlocks (SymmetricCsc.hpp:40-42, SymmetricCsc.cpp:41-200) on the device: the lower CSC of the kept
 * DOFs in equation order (equation_of_dof: per node-major DOF its equation, -1 when eliminated:
 * DofSelection::equationOfDof; dof_count = first_dof[node_count]), each column's rows ascending, the values copied
 * from the blocks (a diagonal block read on and below its diagonal): the reference's bits. Every pointer is device
 * memory; the work is queued on `stream` (a hipStream_t; null: the null stream) and finished on return. Call once with
 * rows == NULL for offsets (kept + 1) and *nnz, then with rows and values of *nnz terms. OOSG_INVALID_ARGUMENT for a
 * block matrix or a selection that is not well formed (DOF counts, block order, value offsets, equations that are not
 * a numbering of the kept DOFs). */
oosg_status oosg_csc_from_blocks_device(int32_t device, const oosg_block_matrix_desc *blocks,
                                        const int32_t *equation_of_dof, int64_t dof_count, int32_t kept,
                                        int64_t *offsets, int32_t *rows, double *values, int64_t *nnz, void *stream);
/* A matrix of the plan from a CSC in device memory (oosg_csc_from_blocks_device's): its structure is read once to the
 * host for the plan's positions, as oosg_matrix_create does. */
oosg_status oosg_matrix_create_device(oosg_plan *plan, int32_t n, const int64_t *offsets, const int32_t *rows,
                                      const double *values, oosg_matrix **out);
/* New values of a matrix from device memory, in the order of its CSC (a matrix of the same structure converted on the
 * device): its diagonal, FP32 scales and rows' copy follow on the device, no host copy of the values. */
oosg_status oosg_matrix_set_values_device(oosg_matrix *matrix, const double *values);

/* The numeric factorisation (SupernodalFactor::factorise). OOSG_SINGULAR for a zero or non-finite pivot in positive
 * definite mode, or when the ratio test fails with bailout >= 0; the message names the equations. */
oosg_status oosg_factorise(oosg_plan *plan, const oosg_matrix *matrix, const oosg_factor_options *options,
                           oosg_factor **out);
void oosg_factor_destroy(oosg_factor *factor);

/* ---- out-of-core factorisation (OutOfCoreFactor, include/oos/linalg/supernodal/OutOfCoreFactor.hpp:14-95) -------
 *
 * A plan whose factorisations keep their panels in pinned host memory (the spool, ScratchSpool's in-memory mode,
 * ScratchSpool.hpp:44-60): each level's panels and inverses are made in one of two device windows and copied to the
 * spool once the level is done, so the device holds two levels' panels, not the factor's. The solves (every solve
 * entry point) copy the levels back into the windows, the next level's copy beside this level's sweep. The factor and
 * the solves give the in-core plan's bits. Positive definite mode (FP64 or FP32 fronts). */
oosg_status oosg_plan_create_out_of_core(const oosg_symbolic_desc *symbolic, int32_t device,
                                         uint64_t hbm_budget_bytes, oosg_plan **out);
/* What the spool holds and what the factorisation wrote and the solves read back (OutOfCoreTraffic,
 * OutOfCoreFactor.hpp:14-20). */
typedef struct oosg_out_of_core_traffic {
    uint64_t spooled_bytes;
    uint64_t written_bytes;
    uint64_t read_bytes;
    uint64_t read_requests;
} oosg_out_of_core_traffic;
oosg_status oosg_factor_out_of_core_traffic(const oosg_factor *factor, oosg_out_of_core_traffic *traffic);

/* ---- partial factorisation (PartialFactor, include/oos/linalg/supernodal/PartialFactor.hpp:47-75) ---------------
 *
 * The plan of a symbolic factor analysed with a trailing set (the boundary of a unit, ordered last: its last supernode,
 * of trailing_dof_count columns and no rows below): oosg_factorise on it eliminates the interior only, in positive
 * definite mode and FP64 (PartialFactor::factorise, PartialFactor.cpp:65-83: OOSG_INVALID_ARGUMENT otherwise, with the
 * reference's message), and assembles the trailing front, the Schur complement
 *     K* = K_bb − K_bi·K_ii⁻¹·K_ib
 * (the matrix entries of its columns, then its children's update matrices in child order). The factor's inertia,
 * ratio test and issues cover the interior equations. Its solves are the two calls below (the whole-matrix solves
 * return OOSG_INVALID_ARGUMENT). trailing_dof_count = 0 gives a whole plan (oosg_plan_create). */
oosg_status oosg_plan_create_partial(const oosg_symbolic_desc *symbolic, int32_t trailing_dof_count, int32_t device,
                                     uint64_t hbm_budget_bytes, oosg_plan **out);
/* K* (PartialFactor::schurComplement, PartialFactor.hpp:85-86): t × t, both triangles, into host `out` (leading
 * dimension ld >= t), in the order of the trailing equations. */
oosg_status oosg_partial_schur_complement(const oosg_factor *factor, double *out, int64_t ld);
/* f* = f_b − K_bi·K_ii⁻¹·f_i for each column of the host `loads` (n rows, leading dimension ld, nrhs columns), into the
 * host `condensed` (t × nrhs, leading dimension ldc) (PartialFactor::condenseLoads, PartialFactor.cpp:117-133). */
oosg_status oosg_partial_condense_loads(const oosg_factor *factor, const double *loads, int64_t ld, int64_t nrhs,
                                        double *condensed, int64_t ldc);
/* Recovery (PartialFactor::recover, PartialFactor.cpp:135-158): on entry the interior rows of the host
 * `loads_to_solution` (n × nrhs, leading dimension ld) hold f_i, and `trailing_values` (t × nrhs, leading dimension
 * ldt) holds u_b; on return every row holds u: u_i = K_ii⁻¹·(f_i − K_ib·u_b), u_b copied. */
oosg_status oosg_partial_recover(const oosg_factor *factor, double *loads_to_solution, int64_t ld, int64_t nrhs,
                                 const double *trailing_values, int64_t ldt);
/* The checksum of a matrix's structure and values (host), by which a recomputed interior factor is refused for another
 * matrix (PartialFactor::isFactorisationOf and recomputeInteriorFactor, PartialFactor.cpp:40-60): 64-bit FNV-1a over
 * the dimension, the n + 1 column offsets, the row indices (32 bits each) and the values' bits. */
oosg_status oosg_csc_checksum(const oosg_csc_desc *matrix, uint64_t *checksum);
oosg_status oosg_factor_info_get(const oosg_factor *factor, oosg_factor_info *info);
/* One pivot that fails the ratio test (PivotIssue, include/oos/linalg/supernodal/PivotDiagnostics.hpp:54-62). */
typedef struct oosg_pivot_issue {
    int32_t equation;
    int32_t kind; /* 0 RatioAboveMaximum, 1 NegativePivot (PivotIssueKind) */
    double matrix_diagonal; /* the diagonal of the test: the matrix's, or the reference given for the equation */
    double factor_diagonal;
    double ratio;
} oosg_pivot_issue;
/* The issues of a factor kept despite them (bailout < 0), in elimination order: the first OOSG_RECORDED_ISSUE_COUNT
 * (PivotDiagnostics::issues(), PivotDiagnostics.hpp:78-81; info.issue_count counts all). Up to `capacity` are
 * written; *recorded receives how many are kept. */
oosg_status oosg_factor_issues(const oosg_factor *factor, oosg_pivot_issue *issues, int64_t capacity,
                               int64_t *recorded);
/* The panel of supernode s in the reference's layout (FactorPanel, SupernodalFactor.hpp:110-125): values
 * (c + r) × e column-major, L below the diagonal, D on it, zero above; rows[k] the equation of row k. Either output
 * may be null; *eliminated receives e. For tests and diagnostics. */
oosg_status oosg_factor_panel(const oosg_factor *factor, int32_t supernode, double *values, int32_t *rows,
                              int32_t *eliminated);
/* The shape of that panel: its rows (c + r, or c + d + r with d columns delayed into the front by its children in
 * indefinite mode) and its eliminated columns e (FactorPanel::rows.size(), eliminatedCount()). */
oosg_status oosg_factor_panel_shape(const oosg_factor *factor, int32_t supernode, int32_t *row_count,
                                    int32_t *eliminated);
/* The subdiagonal of D over that panel's e pivots: e_k of a 2 × 2 block starting at k, 0 elsewhere
 * (FactorPanel::subdiagonal, SupernodalFactor.hpp:110-125). All zero in positive definite mode. */
oosg_status oosg_factor_subdiagonal(const oosg_factor *factor, int32_t supernode, double *subdiagonal);
/* D of each equation (n values): for a 2 × 2 block, its diagonal entries. */
oosg_status oosg_factor_diagonal(const oosg_factor *factor, double *d);

/* X ← A⁻¹·B in place (FactorSolve::solve, FactorSolve.hpp:77-81): b is n × nrhs column-major, leading dimension
 * ldb >= n, host memory, any nrhs (tiles of 16 columns through the plan's own vectors: nothing is allocated). With an
 * FP32 factor the solve applies the factor once (no refinement). */
oosg_status oosg_solve(const oosg_factor *factor, double *b, int64_t ldb, int64_t nrhs);
/* The same on device memory: b_device is n × nrhs column-major (leading dimension ldb) in the plan's device memory. */
oosg_status oosg_solve_device(const oosg_factor *factor, double *b_device, int64_t ldb, int64_t nrhs);
/* x = A⁻¹·b by iterative refinement in double precision until the backward error is at most `target` (or
 * max_steps solves): x_device ← 0, then d = factor⁻¹·r, x += d, r = b − A·x with `matrix` (the matrix factorised).
 * b_device and x_device are n × nrhs column-major device arrays (leading dimension ld), any nrhs: tiles of 16 columns
 * (FactorSolve.hpp:62-67 tiles its right-hand sides likewise), refined one after the other in the plan's vectors; the
 * report gives the most solves of a tile and the largest backward error. OOSG_NOT_CONVERGED when the target is not
 * met (the report says how far it got). */
oosg_status oosg_solve_refined_device(const oosg_factor *factor, const oosg_matrix *matrix, const double *b_device,
                                      double *x_device, int64_t ld, int64_t nrhs, double target, int32_t max_steps,
                                      oosg_solve_report *report);
/* No counterpart in the reference: oosg_solve_refined_device for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). The same refinement, the same
 * kernels in the same order and the same results, but its working vectors (the solve's and the refinement's, and a
 * pivoted factor's sweep vectors) are laid out in a file-scope device array of OOSG_NOALLOC_SCRATCH_BYTES that the
 * library holds on each device from the first plan made there until the process ends; the call allocates and frees
 * nothing and captures no graph. OOSG_INVALID_ARGUMENT when the factor's vectors for 16 right-hand-side columns do
 * not fit the array: 16 × (7·n + the rows below all supernodes) doubles and the sweeps' partial sums, so n up to
 * about 200,000 equations for a shell plate. Every call on a device shares that device's array: calls that overlap on one device (from
 * several threads or plans) race; the array keeps the last call's values, and each call clears what it uses. */
#define OOSG_NOALLOC_SCRATCH_BYTES (256ull << 20)
oosg_status oosg_solve_refined_device_noalloc(const oosg_factor *factor, const oosg_matrix *matrix,
                                              const double *b_device, double *x_device, int64_t ld, int64_t nrhs,
                                              double target, int32_t max_steps, oosg_solve_report *report);
/* oosg_factorise followed by oosg_solve_refined_device, as the reference's caller runs them back to back
 * (FsetSolver.cpp:136 and :171); no counterpart of its own in the reference. Where the factorisation is launched one
 * launch at a time (OOSG_LAUNCH_EAGER_STREAMS, or graphs off) and the tree ends in levels of one supernode (the top
 * separators), the first solve's forward sweep of the levels below them runs on a stream of its own beside their
 * factorisation, whose panel steps leave most of the GPU idle. *factor receives the factor whenever the
 * factorisation succeeds (also with OOSG_NOT_CONVERGED); the caller destroys it. Same results as the two calls. */
oosg_status oosg_factorise_solve_device(oosg_plan *plan, const oosg_matrix *matrix, const oosg_factor_options *options,
                                        const double *b_device, double *x_device, int64_t ld, int64_t nrhs,
                                        double target, int32_t max_steps, oosg_factor **factor,
                                        oosg_solve_report *report);
/* y = A·x for n × nrhs column-major device arrays (SymmetricCsc::multiply, SymmetricCsc.hpp:86-96). */
oosg_status oosg_matrix_multiply_device(const oosg_matrix *matrix, const double *x_device, double *y_device,
                                        int64_t ld, int64_t nrhs);

/* Diagnostic: one 16 × 16 × 4 MFMA product in FP64 and in FP32 (A row-major 16 × 4, B row-major 4 × 16, D row-major
 * 16 × 16, host arrays), run on the current device: the check of the lane layout the update tiles rely on. */
oosg_status oosg_debug_mfma_probe(const double *a, const double *b, double *d, const float *af, const float *bf,
                                  float *df);

/* Per-kernel device times, by HIP events recorded on the plan's stream around every launch while enabled
 * (enabling resets the sums; launches are eager while enabled). OOSG_PROFILE_KINDS kinds, in order: 0 small fronts
 * (fused assembly and elimination), 1 assembly of large fronts, 2 diagonal blocks, 3 panel solves (rows below a
 * diagonal block), 4 panel updates (MFMA tiles), 5 update matrices (MFMA tiles), 6 forward sweep gathers and
 * triangles, 7 backward sweep triangles, 8 residual products, 9 forward sweep row tiles (L_RF), 10 backward sweep
 * column blocks (L_RFᵀ), 11 products with the wide triangles' whole inverses (both sweeps), 12 the whole inverses
 * (factorisation), 13 update matrices on 128 x 128 tiles, 14 panel updates whose last tile factorises the next
 * diagonal block (fused). work[k] is the algorithmic work of the launches timed: flops for kinds 0, 2-5 and 12-14;
 * bytes for kinds 1 (the panels' lower parts written and the children's update-matrix parts read, the matrix's own
 * entries not counted) and 6-11 (the values of L or of an inverse read once, and the right-hand-side rows read or
 * written once; the matrix and the vectors of a residual). Arrays of OOSG_PROFILE_KINDS values; any may be null. */
#define OOSG_PROFILE_KINDS 15
oosg_status oosg_plan_profile(oosg_plan *plan, int32_t enable);
oosg_status oosg_plan_profile_read(const oosg_plan *plan, double *seconds, int64_t *launches, double *work);

/* Replays of the launch sequences as HIP graphs (default on): a factorisation's launches, and a solve's sweeps, are
 * captured once per set of buffers and replayed by one graph launch (the role of the reference's reusable solve plan,
 * SolvePlan.hpp:36-144, on the device side). Off: every kernel is launched eagerly. The profiler launches eagerly. */
oosg_status oosg_plan_use_graphs(oosg_plan *plan, int32_t enable);

/* The tile size of the update-matrix products (no counterpart in the reference): 0 64 x 64 tiles everywhere, 1
 * 128 x 128 tiles everywhere, 2 (default) by level, as oosg_plan_autotune measured, or before any measurement big
 * tiles where a level has at least 2048 of them. Both sizes give the same bits (the same k-steps of the same MFMA). */
oosg_status oosg_plan_tile_policy(oosg_plan *plan, int32_t policy);
/* The streams of the factorisation in `precision` (no counterpart in the reference): 1, the plan's stream only
 * (the default), or 3, the bulk of each step's trailing update and the whole inverses beside the critical path, ordered
 * by events. The results are the same bits either way. */
oosg_status oosg_plan_streams(oosg_plan *plan, int32_t precision, int32_t count);
/* How the factorisation in `precision` is launched (no counterpart in the reference): OOSG_LAUNCH_GRAPH, one HIP graph
 * on the plan's stream (the default); OOSG_LAUNCH_GRAPH_STREAMS, one graph captured from the three streams
 * (oosg_plan_streams(…, 3)); OOSG_LAUNCH_EAGER_STREAMS, the three streams launched one launch at a time, for runtimes
 * that do not overlap a graph's branches as they overlap the streams. With graphs off (oosg_plan_use_graphs) every
 * mode launches one launch at a time. The same bits in every mode. */
#define OOSG_LAUNCH_GRAPH 0
#define OOSG_LAUNCH_GRAPH_STREAMS 1
#define OOSG_LAUNCH_EAGER_STREAMS 2
oosg_status oosg_plan_launch_mode(oosg_plan *plan, int32_t precision, int32_t mode);
oosg_status oosg_plan_get_launch_mode(const oosg_plan *plan, int32_t precision, int32_t *mode);
/* The autotuner (no counterpart in the reference): factorises `matrix` (of the plan's structure) twice with each tile
 * size in `precision`, times every update-matrix launch with HIP events, and keeps per level the faster size (policy
 * 2). Then it times whole factorisations in each launch mode (oosg_plan_launch_mode; the three-stream modes put the
 * trailing updates' bulk and the whole inverses on two more streams beside the critical path), and keeps the fastest.
 * *big_launches receives the launches that chose 128 x 128 tiles, *tuned_launches those measured, *streams the streams
 * kept (1 or 3); any may be null. The factors are discarded; the captured graphs are rebuilt on the next
 * factorisation. */
oosg_status oosg_plan_autotune(oosg_plan *plan, const oosg_matrix *matrix, int32_t precision, int32_t *big_launches,
                               int32_t *tuned_launches, int32_t *streams);

/* Diagnostic: the factorisation's launches in order, with their profile kind, level and algorithmic flops, and the
 * device time of each in the last autotuner run, one launch at a time on one stream, the update matrices with the tile
 * size kept (ms; -1 before one). Call with null arrays for the count. */
oosg_status oosg_plan_launch_times(const oosg_plan *plan, int32_t *count, int32_t *kinds, int32_t *levels, float *ms,
                                   double *flops);

/* Blocks until the plan's stream is idle. */
oosg_status oosg_plan_synchronize(oosg_plan *plan);

/* ---- the learned precision selector (no counterpart in the reference) -------------------------------------- */

/* A small MLP (8 features → 32 → 32 → 1, tanh hidden layers, logistic output) that gives, for a cut of a plan's
 * levels (oosg_selector_cut_features), p(factorising the subtrees below the cut in FP32 and the rest in FP64, refined
 * by conjugate gradients, reaches the target backward error faster than FP64 fronts throughout); FP32 on the device,
 * forward, backward and Adam steps as HIP kernels with deterministic reductions. Arrays are host memory. */
#define OOSG_SELECTOR_FEATURES 8
#define OOSG_SELECTOR_PARAMETERS 1377
typedef struct oosg_selector oosg_selector;
/* Parameters drawn Glorot-uniform from `seed` (biases zero). */
oosg_status oosg_selector_create(int32_t device, uint64_t seed, oosg_selector **out);
void oosg_selector_destroy(oosg_selector *selector);
/* The flat parameter array: W1 (32 × 8 row-major), b1 (32), W2 (32 × 32), b2 (32), W3 (32), b3 (1). */
oosg_status oosg_selector_parameters(const oosg_selector *selector, float *params);
oosg_status oosg_selector_set_parameters(oosg_selector *selector, const float *params);
/* The 8 features of each cut of a plan's levels, for the per-subtree choice of OOSG_MIXED: cut l (1 <= l <= L, the
 * plan's levels) factorises the subtrees below level l (the levels 0 .. l - 1) in FP32 and the rest in FP64; row l - 1
 * of `features` holds: l / L; the share of the flops below the cut; log10 of the largest front order (c + r) below
 * the cut − 2.5 and at or above it − 2.5; log10 n − 5; the mean and the minimum over the equations below the cut of
 * log10 of the scaled matrix's diagonal dominance |a_jj| / Σ_i≠j |a_ij| (clamped to [−3, 3]); log10(1 + g) / 4 for the
 * largest pivot growth g = |a_jj / d_j| of those equations in the plan's last factorisation (0 before the first).
 * *count receives L; with features null nothing else is done, otherwise capacity must be at least L. No counterpart
 * in the reference (the learned precision choice is this library's). */
oosg_status oosg_selector_cut_features(const oosg_plan *plan, const oosg_matrix *matrix, float *features,
                                       int32_t capacity, int32_t *count);
/* p for count × 8 features. */
oosg_status oosg_selector_predict(oosg_selector *selector, const float *features, int64_t count, float *probability);
/* The mean binary cross-entropy over count samples (labels in [0, 1]: y·log p + (1 − y)·log(1 − p), negated) and its
 * gradient by the parameters. */
oosg_status oosg_selector_gradient(oosg_selector *selector, const float *features, const float *labels, int64_t count,
                                   float *loss, float *gradient);
/* `steps` full-batch Adam steps (learning rate lr); *loss receives the loss before the last step. */
oosg_status oosg_selector_train(oosg_selector *selector, const float *features, const float *labels, int64_t count,
                                int32_t steps, float learning_rate, float *loss);

/* ---- host analysis (the reference keeps these on the host too) --------------------------------------------- */

/* Geometric nested dissection of a node graph (adjacency without self loops, both directions) on 3 coordinates per
 * node: out_order[p] = the node at position p (the role of ConstrainedOrdering::order,
 * include/oos/linalg/ordering/ConstrainedOrdering.hpp:41-60; this project's own dissection). Returns 0 on success. */
int oosa_order_dissection(int32_t n, const int64_t *offsets, const int32_t *adjacency, const int32_t *weights,
                          const double *coords, int32_t leaf_nodes, int32_t *out_order);
/* The symbolic analysis (SymbolicFactor::analyse, include/oos/linalg/symbolic/SymbolicFactor.hpp:47-50): the same
 * structure as the reference for the same graph, weights and order. Returns a handle, null on failure. */
void *oosa_analyse(int32_t n, const int64_t *offsets, const int32_t *adjacency, const int32_t *weights,
                   const int32_t *order, int32_t trailing, int amalgamate);
void oosa_free(void *symbolic);
/* {nodes, DOFs, supernodes, row total, panel values, factor entries, exact entries, trailing DOFs}, and the flop
 * count. */
void oosa_sizes(void *symbolic, int64_t *sizes, double *flops);
void oosa_export(void *symbolic, int32_t *node_order, int32_t *first_column, int32_t *column_count, int32_t *parent,
                 int64_t *row_start, int32_t *rows);
void oosa_memory(void *symbolic, int64_t *values);
const char *oosa_last_error(void);

#ifdef __cplusplus
}
#endif

#endif
