// Shared declarations of the device kernels (kernels.hip) and the host driver (oosg.cpp).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

namespace oosg
{

// Columns of a panel step of the blocked elimination of a large front.
constexpr int PANEL_BLOCK = 64;
// Fronts of at most this order are assembled and eliminated in one workgroup, the whole front in LDS.
constexpr int SMALL_FRONT_MAX = 96;
// Elements of a front a workgroup of the assembly kernel holds in LDS.
constexpr int ASSEMBLY_BLOCK_VALUES = 4096;
// Supernodes with at least this many columns run the column blocks of their sweeps as steps over many workgroups.
constexpr int WIDE_COLUMNS = 192;

// The symbolic structure in device memory (read-only during a factorisation or a solve).
struct DevicePlan
{
    const int32_t* first;       // [S] first column
    const int32_t* ncol;        // [S] c
    const int32_t* nrow;        // [S] r
    const int64_t* panelOff;    // [S] panel offsets (elements)
    const int64_t* updateOff;   // [S] update matrix offsets (elements)
    const int64_t* rowStart;    // [S + 1]
    const int32_t* rows;        // rows below each supernode's columns (equations)
    const int32_t* rel;         // [row total] position of each row of s in its parent's front
    const int32_t* childStart;  // [S + 1]
    const int32_t* children;    // children in increasing order
    const int64_t* linvOff;     // [S] offsets of the inverses of the diagonal blocks (PANEL_BLOCK columns each)
    // The pull table of the tiny supernodes' forward sweep: position q of the front of s is hit by the update rows
    // (indices into the rows of all supernodes) pulls[pullBase[s] + q], in child order, -1 past the last, at most
    // pullWidth <= TINY_PULLS of them anywhere; pullBase[s] is -1 for a supernode without children or not tiny.
    const int32_t* pullBase;    // [S]
    const int4* pulls;
    int32_t pullWidth;
};

// A matrix in device memory: its lower CSC, the front position of each entry (within the front of the supernode
// owning its column) and, for the FP32 factor, the power-of-two scale of each equation (null: no scaling).
struct DeviceMatrix
{
    const int64_t* offsets;
    const int32_t* rows;
    const double* values;
    const int32_t* position; // row of each entry in its front
    const int32_t* column;   // column of each entry in its front (local to its supernode)
    const double* scale;
};

// A block of a large front, columns [j0, j1) by rows [i0, i1), and its extend-add items [itemStart, itemEnd).
struct AssemblyBlock
{
    int32_t supernode, j0, j1, i0, i1, itemStart, itemEnd, pad;
};

// The part of a child's update matrix that lands in a block: its columns [q0, q1) and, of each column jj, the rows
// [max(jj, ia), ib) (positions in the child's row list; the ranges found on the host, once per plan).
struct AssemblyItem
{
    int32_t child, q0, q1, ia, ib, pad;
};

// One 64 × 64 MFMA tile (element offsets): mode 0 C −= A·diag(d)·Bᵀ with A, B, d in the panels and C in the panels
// (target 0) or the update matrices (target 1); mode 1 C = A·Bᵀ·diag(d)⁻¹ with B an inverse of a diagonal block.
// An update-matrix tile of extendTilesKernel (64 × 64 or 128 × 128) also adds its children's contributions: its
// 64 × 64 quarters q take the assembly items [itemStarts[items + q], itemStarts[items + q + 1]); rowBase and colBase are
// the front positions of its element (0, 0).
struct UpdateTile
{
    int64_t cOff, aOff, bOff, dOff, ld, ldb, ldc;
    int32_t m, n, k, target, lower, mode;
    int32_t items, supernode, rowBase, colBase;
};

// A pivoted front of the indefinite factorisation (device pointers): F, N × N column-major, positions [own columns c,
// delayed columns d, rows]; W, N × C, the updated columns (L·D) by position; perm, diag, sub: per candidate position,
// its index among the candidates before the interchanges, D, and the subdiagonal of a 2 × 2 block (0 otherwise).
struct PivotFront
{
    double* F;
    double* W;
    int32_t* perm;
    double* diag;
    double* sub;
    int32_t N, C, c, d, first, pad;
};

// Where a front's Bunch–Kaufman elimination stands: the next position k, the end of the candidates still to try in
// this pass, the start of the pass, done (1: e = k), and the 2 × 2 blocks taken.
struct PivotState
{
    int32_t k, end, passStart, done, twoByTwo, pad0, pad1, pad2;
};

// Columns [j0, j1) of a child's update matrix (M × M at src, leading dimension lds) added into its parent's front
// (dst, leading dimension ldd) at the positions map[0 .. M).
struct ExtendItem
{
    const double* src;
    double* dst;
    const int32_t* map;
    int32_t lds, ldd, M, j0, j1, pad;
};

struct CopyRange
{
    const double* src;
    double* dst;
    int64_t count;
};

// The values of the inverses of the diagonal blocks of a supernode of c columns (blocks of PANEL_BLOCK, each kb²).
__host__ __device__ inline int64_t blockInverseValues(int64_t c)
{
    const int64_t full = c / PANEL_BLOCK, rest = c % PANEL_BLOCK;
    return full * PANEL_BLOCK * PANEL_BLOCK + rest * rest;
}

// Supernodes with at least WIDE_COLUMNS columns also keep the whole inverse X = L_FF⁻¹ of their unit lower triangle
// (c × c, column-major, leading dimension c) in the inverses' array, after their diagonal blocks' inverses: at
// linvOff[s] + blockInverseValues(c).
__host__ __device__ inline bool keepsTriangleInverse(int c)
{
    return c >= WIDE_COLUMNS;
}

// The sweeps' L_RF products of a big supernode run in chunks of SWEEP_CHUNK blocks of 64 (columns of a forward row
// tile, rows of a backward column block), each into a slot of partial sums that one more launch subtracts, in chunk
// order, where the products are longer than one chunk: the top levels' few big supernodes then spread over many
// workgroups. Items int4 {supernode, tile or block, chunk (-1: the whole range, straight into the result), slot};
// subtractions int4 {first target row (of X, or of the update vectors U with bit 30 of .w), rows, first slot,
// chunks}.
constexpr int SWEEP_CHUNK = 8;

// Small supernodes (at most SWEEP_INLINE_COLUMNS columns and SWEEP_INLINE_ROWS rows below them) do their L_RF
// products in the sweeps inside the workgroup of their triangle, with no row tiles or column blocks of their own: the
// low levels' thousands of small supernodes then take one launch and one workgroup each per sweep.
constexpr int SWEEP_INLINE_COLUMNS = 64, SWEEP_INLINE_ROWS = 128;
// Supernodes swept by one wave each (tinyForwardKernel, tinyBackwardKernel): at most TINY_COLUMNS columns and
// TINY_ROWS rows below them.
constexpr int TINY_COLUMNS = 64, TINY_ROWS = 128;
// At most this many children's update rows may hit one position of a tiny supernode's front (else it is swept by the
// kernels of the other supernodes).
constexpr int TINY_PULLS = 4;
// The other supernodes of at most MEDIUM_COLUMNS columns are swept by a workgroup each (mediumForwardKernel,
// mediumBackwardKernel), after their gathers; they have no row tiles or column blocks.
constexpr int MEDIUM_COLUMNS = 128;
constexpr int MEDIUM_ROW_INDICES = 2048; // row indices a medium backward sweep stages in LDS
__host__ __device__ inline bool isMediumSupernode(int c)
{
    return c > 0 && c <= MEDIUM_COLUMNS;
}
__host__ __device__ inline bool isTinySupernode(int c, int r)
{
    return c > 0 && c <= TINY_COLUMNS && r <= TINY_ROWS;
}
__host__ __device__ inline bool sweepRowsInline(int c, int r)
{
    return c <= SWEEP_INLINE_COLUMNS && r <= SWEEP_INLINE_ROWS;
}

// Sets the LDS limits of the kernels that need more than the default; once per process and device, before any launch.
hipError_t configureKernels();

template <typename T>
hipError_t launchSmallFronts(const DevicePlan& p, const DeviceMatrix& a, T* panels, T* updates, T* linv,
                             const int32_t* list, int count, int maxOrder, double* diagOut, int32_t* zeroAt,
                             hipStream_t stream);
template <typename T> size_t smallFrontBytes(int maxOrder);
template <typename T>
hipError_t launchAssembly(const DevicePlan& p, const DeviceMatrix& a, T* panels, T* updates,
                          const AssemblyBlock* blocks, const AssemblyItem* items, int count, hipStream_t stream);
template <typename T>
hipError_t launchDiagonalBlocks(const DevicePlan& p, T* panels, T* linv, const int32_t* fronts, int frontCount, int k0,
                                double* diagOut, int32_t* zeroAt, hipStream_t stream);
// A step of the whole inverses X = L_FF⁻¹ of the wide supernodes' triangles (items (s, I << 16 | J)): step K >= 0
// updates T_IJ = −Σ_{K'<=K} L_IK'·X_K'J for I > K >= J; step −1 sets X_IJ = L_II⁻¹·T_IJ (I > J) and X_II = L_II⁻¹.
template <typename T>
hipError_t launchTriangleInverse(const DevicePlan& p, const T* panels, T* linv, const int2* items, int count, int step,
                                 hipStream_t stream);
// The triangles of the wide supernodes in a sweep: items int4 {s, I, chunk, slot} write into their slot of partial
// sums the chunk's part (SWEEP_CHUNK blocks J) of Σ_{J<=I} X_IJ·X(block J) (forward) or Σ_{J>=I} X_JIᵀ·X(block J)
// (transpose); the subtraction launch's items with bit 29 (assign) then set X's rows of block I to the sum.
template <typename TP>
hipError_t launchTriangleProducts(const DevicePlan& p, const TP* linv, const double* X, int k, const int4* items,
                                  int count, int transpose, double* partials, hipStream_t stream);

// The update matrices: C = −A·diag(d)·Bᵀ + Σ children's contributions, on TS × TS tiles (64 or 128; the same bits
// for both sizes).
template <typename T>
hipError_t launchExtendTiles(const DevicePlan& p, T* panels, T* updates, const UpdateTile* tiles, int count,
                             const int32_t* itemStarts, const AssemblyItem* items, int tileSize, hipStream_t stream);
// Panel solves (mode 1) and panel updates (mode 0); a mode-2 tile also factorises the diagonal block it finishes
// (writing L, D, L11⁻¹, diagOut and zeroAt as the diagonal-block kernel does).
template <typename T>
hipError_t launchUpdateTiles(const DevicePlan& p, T* panels, T* updates, T* linv, const UpdateTile* tiles, int count,
                             double* diagOut, int32_t* zeroAt, bool fused, hipStream_t stream);
// One level of the forward sweep: the supernodes of `list`, then the row tiles (supernode, tile) of its big ones.
template <typename TP>
hipError_t launchForward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, double* U, int k,
                         const int32_t* list, int count, const int4* rowTiles, int tileCount, double* partials,
                         hipStream_t stream, bool gathered = false);
// The gathers of a level's other supernodes, by chunks of GATHER_CHUNK front positions: X(cols) += and U(rows) = the
// children's update rows, before launchForward(gathered). A chunk: its first position q0 in the front of its supernode
// (order c + r, c columns from `first`, update rows from `rowStart`), the front's first position in the gather lists.
constexpr int GATHER_CHUNK = 16;
struct GatherChunk
{
    int32_t q0, order, columns, first, rowStart, positions, pad0, pad1;
};
hipError_t launchGather(const GatherChunk* chunks, int chunkCount, const int32_t* gatherStart,
                        const int32_t* gatherRows, double* X, double* U, int k, hipStream_t stream);
// One level of the backward sweep: the column blocks (supernode, block) of its big supernodes, then all of `list`.
// The sweeps of tiny supernodes (isTinySupernode and a pull table entry when they have children): one wave each.
template <typename TP>
hipError_t launchTinyForward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, double* U, int k,
                             const int32_t* list, int count, hipStream_t stream);
template <typename TP>
hipError_t launchMediumForward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, double* U, int k,
                               const int32_t* list, int count, hipStream_t stream);
template <typename TP>
hipError_t launchMediumBackward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, int k,
                                const int32_t* list, int count, hipStream_t stream);
template <typename TP>
hipError_t launchTinyBackward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, int k,
                              const int32_t* list, int count, hipStream_t stream);
template <typename TP>
hipError_t launchBackward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, int k, const int32_t* list,
                          int count, const int4* columnBlocks, int blockCount, double* partials, hipStream_t stream);
// X or U rows −= Σ partial sums (the chunks of the subtractions' items), summed in chunk order.
hipError_t launchSubtractPartials(double* X, double* U, int k, const int4* items, int count, const double* partials,
                                  hipStream_t stream);
hipError_t launchDiagonalSolve(double* X, const double* d, int64_t n, int k, hipStream_t stream);
// A partial factor's condensation: row-major X's trailing columns [first, first + c) += their children's update rows.
hipError_t launchTrailingGather(double* X, const double* U, int k, int first, int c, const int32_t* pullStart,
                                const int32_t* pulls, hipStream_t stream);

hipError_t launchFillInt32(int32_t* data, int64_t count, int32_t value, hipStream_t stream);
hipError_t launchToRowMajor(const double* src, int64_t ld, double* dst, int64_t n, int k, const double* scale,
                            hipStream_t stream);
// With norm: the column maxima |S·src| into it too (bits, zeroed by the caller).
hipError_t launchFromRowMajor(const double* src, double* dst, int64_t ld, int64_t n, int k, const double* scale,
                              int accumulate, hipStream_t stream, unsigned long long* norm = nullptr);
// Right-hand sides src (column-major n × k, leading dimension ld) into B (row-major) and X1 = S·src (row-major; null:
// none; scale null: a copy), with the column maxima |src| into norm (bits, zeroed by the caller; null: none): one pass.
hipError_t launchRightHandSides(const double* src, int64_t ld, double* B, double* X1, const double* scale, int64_t n,
                                int k, unsigned long long* norm, hipStream_t stream);
hipError_t launchResidual(const int64_t* offsets, const int32_t* cols, const double* values, const double* x,
                          const double* b, double* out, int64_t ld, int64_t n, int k, hipStream_t stream);

hipError_t launchFillDouble(double* data, int64_t ld, int64_t n, int k, double value, hipStream_t stream);

// Gram products: their partial sums array holds (GRAM_GROUPS + 16) × GRAM_PAIRS × 256 values (the workgroups' sums,
// then the sums of 16 runs of them).
constexpr int GRAM_PAIRS = 4;
constexpr int GRAM_GROUPS = 1024; // workgroups of a Gram product (a contiguous range of rows each)
// Gram products of up to GRAM_PAIRS pairs of row-major n × k vectors (k <= 16): out[q·256 + 16a + b] =
// Σ_i U_q(i, a)·V_q(i, b), zero beyond k; a fixed order of summation (deterministic).
struct GramPairs
{
    const double* u[GRAM_PAIRS];
    const double* v[GRAM_PAIRS];
    int pairs;
};
hipError_t launchBlockGram(const GramPairs& pairs, int64_t n, int k, double* partial, double* out,
                           hipStream_t stream);
// Y(i, :) = c0·Y(i, :) + U(i, :)·C for row-major n × k vectors (k <= 16) and a 16 × 16 coefficient block C (row a,
// column b): given by value, or read from device memory (the result of a Gram product) times `scale`. Y must not be U.
struct BlockTerm
{
    const double* u;
    const double* deviceC; // non-null: C = scale·deviceC (16 × 16 on the device)
    double scale;
    double c0;
    double c[256]; // C by value where deviceC is null
};
hipError_t launchBlockCombine(double* y, const BlockTerm& term, int64_t n, int k, hipStream_t stream);
// One block conjugate-gradient step over row-major n × k vectors, from W (in z), A·W (in p) and T, U = T·α (16 × 16
// each at coef and coef + 256 on the device): Q = (A·W)·T, P = W·T (over A·W), X += W·U, R −= (A·W)·U, Z = R (over
// W); the column maxima of |S⁻¹·R| and |S·X| into norms[0, 16) and norms[16, 32) (bits; the caller zeroes them;
// scale null: S = I).
hipError_t launchBlockStep(double* z, double* p, double* q, double* x, double* r, const double* coef,
                           const double* scale, int64_t n, int k, unsigned long long* norms, hipStream_t stream);

// Row-major n × k vectors (k <= 16): out = b − A·x (b null: A·x) with the CSR copy of A, by its row groups (rows
// [groups[g], groups[g + 1]), at most 8, with the same columns), or with a row scaling S (powers of two) the scaled
// matrix's out = S·(b − A·S·x); column max |v| (bits into out, which the caller zeroes), of S·v or S⁻¹·v (divide)
// with a scaling; dst = src·diag(scale) by rows; x += z.
hipError_t launchMultiplyRows(const int64_t* offsets, const int32_t* cols, const double* values,
                              const int32_t* groups, int64_t groupCount, const double* x, const double* b, double* out,
                              int k, hipStream_t stream, const double* scale = nullptr);
hipError_t launchColumnMaxRows(const double* v, int64_t n, int k, unsigned long long* out, hipStream_t stream,
                               const double* scale = nullptr, int divide = 0);
hipError_t launchScaleRows(const double* src, double* dst, const double* scale, int64_t n, int k, hipStream_t stream);
hipError_t launchAddRows(double* x, const double* z, int64_t n, int k, hipStream_t stream);
hipError_t launchPlainUpdateTiles(double* fronts, const double* w, const UpdateTile* tiles, int count,
                                  hipStream_t stream);
hipError_t launchBunchKaufman(const PivotFront* fronts, PivotState* states, const int32_t* list, int count, int nb,
                              hipStream_t stream);
hipError_t launchPivotMatrix(const DeviceMatrix& a, const PivotFront* fronts, int count, hipStream_t stream);
hipError_t launchPivotExtendAdd(const ExtendItem* items, int count, hipStream_t stream);
hipError_t launchCopyRanges(const CopyRange* ranges, int count, hipStream_t stream);
// A mixed factor's boundary: update matrices written in FP32 by the levels below the cut, read in FP64 above it. Each
// range (at most 65536 values) is staged through tmp: phase 0 copies src to tmp for every range, phase 1 widens tmp
// into dst (dst may overlap the sources of other ranges, never a range's own tmp).
struct WidenRange
{
    const float* src;
    float* tmp;
    double* dst;
    int64_t count;
};
hipError_t launchWidenRanges(const WidenRange* ranges, int count, hipStream_t stream);
hipError_t launchUnitBlockInverses(const DevicePlan& p, const double* panels, double* linv, const int2* items,
                                   int count, hipStream_t stream);
hipError_t launchBlockDiagonalSolve(double* X, const double* d, const double* sub, int64_t n, int k,
                                    hipStream_t stream);
hipError_t launchPermuteRows(const double* src, double* dst, const int32_t* perm, int64_t n, int k, bool scatter,
                             hipStream_t stream);
hipError_t launchMfmaProbe(const double* a, const double* b, double* d, const float* af, const float* bf, float* df,
                           hipStream_t stream);

// The pivots' statistics of a positive definite factor (the ratio test of MultifrontalDriver::finish): the pivot of
// equation j is d_j, or d_j / s_j² with the scales s of an FP32 factor; its ratio is the diagonal a_jj (or the
// reference diagonal where that is >= 0) over it. Counts by sign; the largest ratio (as an ordered key, NaN ratios
// left out) and the first equation that has it; the pivots that fail the test (ratio > maxRatio, or negative ratio or
// pivot), the first of them and whether it fails by a negative value.
struct PivotStatistics
{
    unsigned long long positive, negative, zero, issues;
    unsigned long long largestKey;
    int32_t largestAt, firstIssue, firstIssueNegative, pad;
};
hipError_t launchPivotStatistics(const double* diag, const double* scale, const double* diagonal,
                                 const double* reference, int64_t n, double maxRatio, PivotStatistics* out,
                                 hipStream_t stream);
// Per supernode s < count the largest |a_jj / d_j| over its columns (its pivot growth), 0 for a supernode whose
// columns are not all among the first `eliminated` equations.
hipError_t launchSupernodeGrowth(const double* diag, const double* scale, const double* diagonal, const int32_t* first,
                                 const int32_t* ncol, int32_t count, int64_t eliminated, float* out,
                                 hipStream_t stream);

// A block matrix on a node pattern (SparsePattern + BlockSparseMatrix): node j owns the DOFs [firstDof[j],
// firstDof[j + 1]); its block column is [columnStart[j], columnStart[j + 1]), the diagonal block first, then the blocks
// of the rows blockRows[b] > j ascending, each column-major (dofs(row) × dofs(j)) at values + blockValueOffsets[b].
struct BlockMatrixView
{
    int32_t nodes;
    const int64_t* firstDof;
    const int64_t* columnStart;
    const int32_t* blockRows;
    const int64_t* blockValueOffsets;
    const double* values;
};

// SymmetricCsc::fromBlocks on the device. Validation (flag bits: DOF counts, block columns, block rows, value offsets,
// equations; hits: per equation, zeroed by the caller), term counts per column (cursor null) or their scatter to the
// slots cursor[column]++, the exclusive scan of the counts (pieceSums: scanPieces(n) values), the columns put in row
// order.
hipError_t launchBlocksValidate(const BlockMatrixView& m, const int32_t* equationOfDof, int64_t dofCount,
                                int32_t kept, int64_t blockCount, int64_t valueCount, int32_t* hits, int32_t* flag,
                                hipStream_t stream);
hipError_t launchBlockTerms(const BlockMatrixView& m, const int32_t* equationOfDof, int32_t* counts,
                            unsigned long long* cursor, int32_t* rows, double* values, hipStream_t stream);
int64_t scanPieces(int64_t n);
hipError_t launchExclusiveScan(const int32_t* counts, int64_t n, int64_t* offsets, int64_t* pieceSums,
                               hipStream_t stream);
hipError_t launchSortColumns(const int64_t* offsets, int32_t n, const int32_t* rowsIn, const double* valuesIn,
                             int32_t* rowsOut, double* valuesOut, hipStream_t stream);

// The values of a matrix from two of its structure (SymmetricCsc::combine, then ::scaleSymmetric when scale is not
// null): out[k] = scale[row]·(alpha·a[k] + beta·b[k])·scale[column], every product and the sum rounded in that
// order (b null: scale[row]·a[k]·scale[column]); the diagonal of each column (0 where not stored) and the
// power-of-two scale of the FP32 factor. out may be
// a or b.
hipError_t launchCombineValues(int32_t n, const int64_t* offsets, const int32_t* rows, double alpha, const double* a,
                               double beta, const double* b, const double* scale, double* out, double* diagonal,
                               double* fp32Scale, hipStream_t stream);
// The full rows of the residual product from the lower values (csrValues[e] = values[map[e]]) and ‖A‖∞: each row's
// |values| summed in its column order, the largest sum as an ordered key (atomicMax; *normKey cleared first).
hipError_t launchRowValues(int32_t n, const int64_t* csrOffsets, const int64_t* map, const double* values,
                           double* csrValues, unsigned long long* normKey, hipStream_t stream);

// The scratch of the refined solve that allocates nothing (oosg_solve_refined_device_noalloc): a file-scope device
// array of SCRATCH_BYTES in scratch.hip, one copy per device, held until the process ends. scratchAddress gives its
// address on `device` (which must be current), taken once per device with hipGetSymbolAddress and kept.
constexpr size_t SCRATCH_BYTES = size_t(256) << 20;
hipError_t scratchAddress(int device, void** out);

} // namespace oosg
