// Test-infrastructure driver over the compiled reference (jcdulas/O_o_Solver, `base` + `linalg`), built by
// oracle/build_ref.py into oracle/_ref/liboos_ref.so. Only the tests and the CPU-baseline leg of bench.py load it, as
// the checker: it is never part of the product path.
//
// It exposes, as plain C functions, the reference calls the parity tests compare against:
//   ref_analyse          ConstrainedOrdering::order + SymbolicFactor::analyse
//                        (include/oos/linalg/ordering/ConstrainedOrdering.hpp:41-60,
//                         include/oos/linalg/symbolic/SymbolicFactor.hpp:47-50)
//   ref_factor_solve     SupernodalFactor::factorise + FactorSolve::solve
//                        (include/oos/linalg/supernodal/SupernodalFactor.hpp:181-183,
//                         include/oos/linalg/supernodal/FactorSolve.hpp:77-81)
//   ref_spmv             SymmetricCsc::multiply (include/oos/linalg/sparse/SymmetricCsc.hpp:86-90)
//   ref_analyse_trailing SymbolicFactor::analyse of a given order with a trailing set (SymbolicFactor.hpp:41-50)
//   ref_partial          PartialFactor::factorise, schurComplement, condenseLoads, recover
//                        (include/oos/linalg/supernodal/PartialFactor.hpp:75-101)
//   ref_pattern_*, ref_selection, ref_from_blocks
//                        SparsePattern of a node graph, DofSelection, SymmetricCsc::fromBlocks
//                        (include/oos/linalg/sparse/SparsePattern.hpp:24-27, DofSelection.hpp:24-27,
//                         SymmetricCsc.hpp:40-42)
//   ref_combine          SymmetricCsc::combine (src/linalg/sparse/SymmetricCsc.cpp:209-236)
//   ref_scale_symmetric  SymmetricCsc::scaleSymmetric (src/linalg/sparse/SymmetricCsc.cpp:239-252)
//   ref_pencil_scale_of  pencilScaleOf (src/linalg/eigen/PencilScaling.cpp:63-78)
//   ref_dense_*          the reference kernel family's LDLᵀ entry points
//                        (include/oos/linalg/kernels/DenseKernelEntryPoints.hpp:253-264)
// Errors come back as status codes: 1 std::invalid_argument, 2 MemoryBudgetError, 3 SingularityError, 4 other.

#include "oos/base/errors/MemoryBudgetError.hpp"
#include "oos/base/errors/SingularityError.hpp"
#include "oos/base/math/Vec3.hpp"
#include "oos/base/memory/MemoryBudget.hpp"
#include "oos/base/threading/ThreadPool.hpp"
#include "oos/linalg/dense/DenseKernels.hpp"
#include "oos/linalg/eigen/PencilScaling.hpp"
#include "oos/linalg/kernels/DenseKernelEntryPoints.hpp"
#include "oos/linalg/ordering/ConstrainedOrdering.hpp"
#include "oos/linalg/ordering/NodeGraph.hpp"
#include "oos/linalg/ordering/NodePermutation.hpp"
#include "oos/linalg/sparse/BlockSparseMatrix.hpp"
#include "oos/linalg/sparse/DofSelection.hpp"
#include "oos/linalg/sparse/SparsePattern.hpp"
#include "oos/linalg/sparse/SymmetricCsc.hpp"
#include "oos/linalg/supernodal/FactorSolve.hpp"
#include "oos/linalg/supernodal/PartialFactor.hpp"
#include "oos/linalg/supernodal/SupernodalFactor.hpp"
#include "oos/linalg/symbolic/SymbolicFactor.hpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

using namespace oos;

namespace
{

constexpr std::uint64_t BUDGET = base::MemoryBudget::UNLIMITED;

void copyMessage(const char* text, char* err, int errlen)
{
    if (err == nullptr || errlen <= 0)
        return;
    std::strncpy(err, text, static_cast<std::size_t>(errlen - 1));
    err[errlen - 1] = '\0';
}

template <typename F> int guarded(F&& body, char* err, int errlen)
{
    try
    {
        body();
        return 0;
    }
    catch (const std::invalid_argument& e)
    {
        copyMessage(e.what(), err, errlen);
        return 1;
    }
    catch (const base::MemoryBudgetError& e)
    {
        copyMessage(e.what(), err, errlen);
        return 2;
    }
    catch (const base::SingularityError& e)
    {
        copyMessage(e.what(), err, errlen);
        return 3;
    }
    catch (const std::exception& e)
    {
        copyMessage(e.what(), err, errlen);
        return 4;
    }
}

// What ref_analyse keeps: the budget must outlive the symbolic factor's buffers.
struct Analysis
{
    std::unique_ptr<base::MemoryBudget> budget;
    linalg::SymbolicFactor symbolic;
    std::unique_ptr<linalg::NodeGraph> graph;
    std::vector<std::int32_t> weights;
};

} // namespace

extern "C" {

// method: 0 automatic (ConstrainedOrdering::chooseMethod), 1 minimum degree, 2 nested dissection, 3 natural,
// 4 the order given in `given_order` (node at each position). amalgamate: 1 default thresholds, 0 none.
void* ref_analyse(std::int32_t node_count, const std::int64_t* adj_offsets, const std::int32_t* adj,
                  const std::int32_t* weights, const double* coords, const std::int32_t* given_order, int method,
                  int amalgamate, int threads, int* status, char* err, int errlen)
{
    auto analysis = std::make_unique<Analysis>();
    *status = guarded(
        [&]
        {
            analysis->budget = std::make_unique<base::MemoryBudget>(BUDGET);
            base::MemoryBudget& budget = *analysis->budget;
            const std::int64_t adjacency = adj_offsets[node_count];
            analysis->graph = std::make_unique<linalg::NodeGraph>(linalg::NodeGraph::fromAdjacency(
                node_count, std::span<const std::int64_t>(adj_offsets, static_cast<std::size_t>(node_count) + 1),
                std::span<const std::int32_t>(adj, static_cast<std::size_t>(adjacency)), budget));
            analysis->weights.assign(weights, weights + node_count);
            std::vector<base::Vec3> points;
            if (coords != nullptr)
                for (std::int32_t v = 0; v < node_count; ++v)
                    points.emplace_back(coords[3 * v], coords[3 * v + 1], coords[3 * v + 2]);
            linalg::NodePermutation order;
            if (method == 4)
                order = linalg::NodePermutation(
                    std::span<const std::int32_t>(given_order, static_cast<std::size_t>(node_count)), budget);
            else
            {
                const linalg::OrderingMethod chosen =
                    method == 1   ? linalg::OrderingMethod::MinimumDegree
                    : method == 2 ? linalg::OrderingMethod::NestedDissection
                    : method == 3 ? linalg::OrderingMethod::Natural
                                  : linalg::OrderingMethod::Automatic;
                std::unique_ptr<base::ThreadPool> pool;
                if (threads > 1)
                    pool = std::make_unique<base::ThreadPool>(static_cast<std::size_t>(threads));
                order = linalg::ConstrainedOrdering().order(*analysis->graph, analysis->weights, points, {}, chosen,
                                                            budget, pool.get());
            }
            const linalg::AmalgamationThresholds thresholds =
                amalgamate != 0 ? linalg::AmalgamationThresholds{} : linalg::AmalgamationThresholds::none();
            analysis->symbolic =
                linalg::SymbolicFactor::analyse(*analysis->graph, analysis->weights, order, 0, thresholds, budget);
        },
        err, errlen);
    if (*status != 0)
        return nullptr;
    return analysis.release();
}

// The symbolic factor of the order given in `given_order` (node at each position) whose last `trailing` positions
// are the trailing set (kept last, forming the last supernode).
void* ref_analyse_trailing(std::int32_t node_count, const std::int64_t* adj_offsets, const std::int32_t* adj,
                           const std::int32_t* weights, const std::int32_t* given_order, std::int32_t trailing,
                           int amalgamate, int* status, char* err, int errlen)
{
    auto analysis = std::make_unique<Analysis>();
    *status = guarded(
        [&]
        {
            analysis->budget = std::make_unique<base::MemoryBudget>(BUDGET);
            base::MemoryBudget& budget = *analysis->budget;
            const std::int64_t adjacency = adj_offsets[node_count];
            analysis->graph = std::make_unique<linalg::NodeGraph>(linalg::NodeGraph::fromAdjacency(
                node_count, std::span<const std::int64_t>(adj_offsets, static_cast<std::size_t>(node_count) + 1),
                std::span<const std::int32_t>(adj, static_cast<std::size_t>(adjacency)), budget));
            analysis->weights.assign(weights, weights + node_count);
            const linalg::NodePermutation order(
                std::span<const std::int32_t>(given_order, static_cast<std::size_t>(node_count)), budget);
            const linalg::AmalgamationThresholds thresholds =
                amalgamate != 0 ? linalg::AmalgamationThresholds{} : linalg::AmalgamationThresholds::none();
            analysis->symbolic = linalg::SymbolicFactor::analyse(*analysis->graph, analysis->weights, order,
                                                                 trailing, thresholds, budget);
        },
        err, errlen);
    if (*status != 0)
        return nullptr;
    return analysis.release();
}

void ref_sym_free(void* handle)
{
    delete static_cast<Analysis*>(handle);
}

// Sizes: [node count, dof count, supernode count, row total, panel values, factor entries, exact entries].
void ref_sym_sizes(void* handle, std::int64_t* sizes, double* flops)
{
    const linalg::SymbolicFactor& s = static_cast<Analysis*>(handle)->symbolic;
    sizes[0] = s.nodeCount();
    sizes[1] = s.dofCount();
    sizes[2] = s.supernodeCount();
    std::int64_t rows = 0;
    for (std::int32_t k = 0; k < s.supernodeCount(); ++k)
        rows += s.rowCount(k);
    sizes[3] = rows;
    sizes[4] = s.panelValueCount();
    sizes[5] = s.factorEntryCount();
    sizes[6] = s.exactEntryCount();
    *flops = s.flopCount();
}

// node_order[nodeCount]; first_column, column_count, parent [supernodes]; row_start [supernodes + 1]; rows [row total].
void ref_sym_export(void* handle, std::int32_t* node_order, std::int32_t* first_column, std::int32_t* column_count,
                    std::int32_t* parent, std::int64_t* row_start, std::int32_t* rows)
{
    const linalg::SymbolicFactor& s = static_cast<Analysis*>(handle)->symbolic;
    for (std::int32_t p = 0; p < s.nodeCount(); ++p)
        node_order[p] = s.nodeOrder().nodeAt(p);
    std::int64_t at = 0;
    for (std::int32_t k = 0; k < s.supernodeCount(); ++k)
    {
        first_column[k] = s.firstColumn(k);
        column_count[k] = s.columnCount(k);
        parent[k] = s.supernodes().parent(k);
        row_start[k] = at;
        for (const std::int32_t r : s.rows(k))
            rows[at++] = r;
    }
    row_start[s.supernodeCount()] = at;
}

// The factorisation and solve of the reference on the symbolic factor of `handle`.
// kernel: 0 reference family, 1 avx2, 2 avx512. pivoting: 0 positive definite, 1 indefinite.
// b: n × nrhs column-major, solved in place (may be null with nrhs = 0).
// inertia[3], counts[4] = {delayed, 2x2, issue count, equation of the first issue}, times[2] = factor s, solve s.
// panels (optional, sized panelValueCount) receives the panels in the symbolic shape (positive definite mode).
// diag (optional, n values) receives d per equation in pivot order of the panels: diag[rows_of_panel[k]].
int ref_factor_solve(void* handle, std::int32_t n, const std::int64_t* offsets, const std::int32_t* rows,
                     const double* values, int kernel, int threads, int pivoting, double max_ratio, int bailout,
                     double* b, std::int64_t nrhs, std::int64_t* inertia, std::int64_t* counts, double* times,
                     double* panels, double* diag, char* err, int errlen)
{
    Analysis& analysis = *static_cast<Analysis*>(handle);
    return guarded(
        [&]
        {
            base::MemoryBudget budget(BUDGET);
            const std::int64_t nnz = offsets[n];
            const linalg::SymmetricCsc matrix = linalg::SymmetricCsc::fromCompressedColumns(
                n, std::span<const std::int64_t>(offsets, static_cast<std::size_t>(n) + 1),
                std::span<const std::int32_t>(rows, static_cast<std::size_t>(nnz)),
                std::span<const double>(values, static_cast<std::size_t>(nnz)), budget);
            std::unique_ptr<base::ThreadPool> pool;
            if (threads > 1)
                pool = std::make_unique<base::ThreadPool>(static_cast<std::size_t>(threads));
            linalg::DenseKernelOptions kernelOptions;
            kernelOptions.pool = pool.get();
            const linalg::KernelKind kind = kernel == 1   ? linalg::KernelKind::Avx2
                                            : kernel == 2 ? linalg::KernelKind::Avx512
                                                          : linalg::KernelKind::Reference;
            const linalg::DenseKernels kernels(kind, kernelOptions);
            linalg::FactorOptions options;
            options.pivoting =
                pivoting == 1 ? linalg::PivotingMode::Indefinite : linalg::PivotingMode::PositiveDefinite;
            options.pivotPolicy.maxRatio = max_ratio;
            options.pivotPolicy.bailout = bailout;
            const auto t0 = std::chrono::steady_clock::now();
            const linalg::SupernodalFactor factor =
                linalg::SupernodalFactor::factorise(analysis.symbolic, matrix, kernels, options, budget);
            const auto t1 = std::chrono::steady_clock::now();
            inertia[0] = static_cast<std::int64_t>(factor.inertia().positive);
            inertia[1] = static_cast<std::int64_t>(factor.inertia().negative);
            inertia[2] = static_cast<std::int64_t>(factor.inertia().zero);
            counts[0] = factor.delayedPivotCount();
            counts[1] = factor.twoByTwoPivotCount();
            counts[2] = factor.pivotDiagnostics().issueCount();
            counts[3] = factor.pivotDiagnostics().issues().empty() ? -1
                                                                   : factor.pivotDiagnostics().issues().front().equation;
            if (panels != nullptr || diag != nullptr)
            {
                std::int64_t at = 0;
                for (std::int32_t s = 0; s < factor.supernodeCount(); ++s)
                {
                    const linalg::FactorPanel panel = factor.panel(s);
                    const std::size_t m = panel.rows.size();
                    const std::size_t e = panel.eliminatedCount();
                    if (panels != nullptr)
                        for (std::size_t j = 0; j < e; ++j)
                            for (std::size_t i = 0; i < m; ++i)
                                panels[at++] = panel.values(i, j);
                    if (diag != nullptr)
                        for (std::size_t k = 0; k < e; ++k)
                            diag[panel.rows[k]] = panel.values(k, k);
                }
            }
            double solveSeconds = 0.0;
            if (b != nullptr && nrhs > 0)
            {
                const linalg::DenseMatrixView view{b, static_cast<std::size_t>(n), static_cast<std::size_t>(nrhs),
                                                   static_cast<std::size_t>(n)};
                const auto s0 = std::chrono::steady_clock::now();
                linalg::FactorSolve(factor, kernels).solve(view);
                const auto s1 = std::chrono::steady_clock::now();
                solveSeconds = std::chrono::duration<double>(s1 - s0).count();
            }
            times[0] = std::chrono::duration<double>(t1 - t0).count();
            times[1] = solveSeconds;
        },
        err, errlen);
}

// The reference's PartialFactor of the matrix on the symbolic factor of `handle` (a trailing set), reference family
// unless kernel is 1 (avx2) or 2 (avx512). schur (t × t, t = trailing DOFs) receives K*; panels (optional) the
// interior panels in the symbolic shape, diag (optional, n values) the interior pivots, inertia[3] the interior's.
// With loads (n × nrhs column-major): condensed (t × nrhs) receives condenseLoads(loads); with trailing_values
// (t × nrhs) too, recovered (n × nrhs) receives recover(loads, trailing_values).
int ref_partial(void* handle, std::int32_t n, const std::int64_t* offsets, const std::int32_t* rows,
                const double* values, int kernel, double max_ratio, double* schur, double* panels, double* diag,
                std::int64_t* inertia, const double* loads, std::int64_t nrhs, double* condensed,
                const double* trailing_values, double* recovered, char* err, int errlen)
{
    Analysis& analysis = *static_cast<Analysis*>(handle);
    return guarded(
        [&]
        {
            base::MemoryBudget budget(BUDGET);
            const std::int64_t nnz = offsets[n];
            const linalg::SymmetricCsc matrix = linalg::SymmetricCsc::fromCompressedColumns(
                n, std::span<const std::int64_t>(offsets, static_cast<std::size_t>(n) + 1),
                std::span<const std::int32_t>(rows, static_cast<std::size_t>(nnz)),
                std::span<const double>(values, static_cast<std::size_t>(nnz)), budget);
            const linalg::KernelKind kind = kernel == 1   ? linalg::KernelKind::Avx2
                                            : kernel == 2 ? linalg::KernelKind::Avx512
                                                          : linalg::KernelKind::Reference;
            const linalg::DenseKernels kernels(kind, linalg::DenseKernelOptions{});
            linalg::FactorOptions options;
            options.pivoting = linalg::PivotingMode::PositiveDefinite;
            options.pivotPolicy.maxRatio = max_ratio;
            const linalg::PartialFactor partial =
                linalg::PartialFactor::factorise(analysis.symbolic, matrix, kernels, options, budget);
            const auto t = static_cast<std::size_t>(partial.trailingDofCount());
            const linalg::DenseMatrix& kstar = partial.schurComplement();
            for (std::size_t j = 0; j < t; ++j)
                for (std::size_t i = 0; i < t; ++i)
                    schur[i + j * t] = kstar(i, j);
            const linalg::SupernodalFactor& interior = partial.interiorFactor();
            inertia[0] = static_cast<std::int64_t>(interior.inertia().positive);
            inertia[1] = static_cast<std::int64_t>(interior.inertia().negative);
            inertia[2] = static_cast<std::int64_t>(interior.inertia().zero);
            if (panels != nullptr || diag != nullptr)
            {
                std::int64_t at = 0;
                for (std::int32_t s = 0; s < interior.supernodeCount(); ++s)
                {
                    const linalg::FactorPanel panel = interior.panel(s);
                    const std::size_t m = panel.rows.size();
                    const std::size_t e = panel.eliminatedCount();
                    if (panels != nullptr)
                        for (std::size_t j = 0; j < e; ++j)
                            for (std::size_t i = 0; i < m; ++i)
                                panels[at++] = panel.values(i, j);
                    if (diag != nullptr)
                        for (std::size_t k = 0; k < e; ++k)
                            diag[panel.rows[k]] = panel.values(k, k);
                }
            }
            if (loads == nullptr || nrhs <= 0)
                return;
            const auto rowsN = static_cast<std::size_t>(n), columns = static_cast<std::size_t>(nrhs);
            const linalg::ConstDenseMatrixView given{loads, rowsN, columns, rowsN};
            if (condensed != nullptr)
            {
                const linalg::DenseMatrix out = partial.condenseLoads(given, kernels, budget);
                for (std::size_t j = 0; j < columns; ++j)
                    for (std::size_t i = 0; i < t; ++i)
                        condensed[i + j * t] = out(i, j);
            }
            if (trailing_values != nullptr && recovered != nullptr)
            {
                std::memcpy(recovered, loads, sizeof(double) * rowsN * columns);
                partial.recover(linalg::DenseMatrixView{recovered, rowsN, columns, rowsN},
                                linalg::ConstDenseMatrixView{trailing_values, t, columns, std::max<std::size_t>(t, 1)},
                                kernels);
            }
        },
        err, errlen);
}

namespace
{

// The node graph and pattern of ref_pattern_* and ref_from_blocks.
struct PatternOf
{
    base::MemoryBudget budget{BUDGET};
    std::unique_ptr<linalg::NodeGraph> graph;
    std::unique_ptr<linalg::SparsePattern> pattern;

    PatternOf(std::int32_t nodes, const std::int64_t* offsets, const std::int32_t* adjacency,
              const std::uint8_t* dofCounts)
    {
        graph = std::make_unique<linalg::NodeGraph>(linalg::NodeGraph::fromAdjacency(
            nodes, std::span<const std::int64_t>(offsets, static_cast<std::size_t>(nodes) + 1),
            std::span<const std::int32_t>(adjacency, static_cast<std::size_t>(offsets[nodes])), budget));
        pattern = std::make_unique<linalg::SparsePattern>(
            *graph, std::span<const std::uint8_t>(dofCounts, static_cast<std::size_t>(nodes)), budget);
    }
};

} // namespace

// The reference's SparsePattern of a node graph: sizes[0] = blocks, sizes[1] = values; then its arrays (first_dof and
// column_start nodes + 1, block_rows blocks, block_value_offsets blocks + 1).
int ref_pattern_export(std::int32_t nodes, const std::int64_t* offsets, const std::int32_t* adjacency,
                       const std::uint8_t* dof_counts, std::int64_t* sizes, std::int64_t* first_dof,
                       std::int64_t* column_start, std::int32_t* block_rows, std::int64_t* block_value_offsets,
                       char* err, int errlen)
{
    return guarded(
        [&]
        {
            const PatternOf p(nodes, offsets, adjacency, dof_counts);
            const linalg::SparsePattern& pattern = *p.pattern;
            sizes[0] = pattern.blockCount();
            sizes[1] = pattern.valueCount();
            if (first_dof == nullptr)
                return;
            for (std::int32_t v = 0; v <= nodes; ++v)
            {
                first_dof[v] = v < nodes ? pattern.firstDof(v) : pattern.dofCount();
                column_start[v] = v < nodes ? pattern.columnBegin(v) : pattern.blockCount();
            }
            for (std::int64_t b = 0; b < pattern.blockCount(); ++b)
            {
                block_rows[b] = pattern.blockRow(b);
                block_value_offsets[b] = pattern.blockValueOffset(b);
            }
            block_value_offsets[pattern.blockCount()] = pattern.valueCount();
        },
        err, errlen);
}

// The reference's DofSelection: equation_of_dof (one per node-major DOF), *kept the equations.
int ref_selection(std::int32_t nodes, const std::uint8_t* dof_counts, const std::uint8_t* eliminated,
                  std::int64_t total_dofs, const std::int32_t* node_order, std::int32_t order_count,
                  std::int32_t* equation_of_dof, std::int32_t* kept, char* err, int errlen)
{
    return guarded(
        [&]
        {
            base::MemoryBudget budget(BUDGET);
            const linalg::DofSelection selection(
                std::span<const std::uint8_t>(dof_counts, static_cast<std::size_t>(nodes)),
                std::span<const std::uint8_t>(eliminated, eliminated != nullptr ? static_cast<std::size_t>(total_dofs)
                                                                                 : 0),
                std::span<const std::int32_t>(node_order, static_cast<std::size_t>(order_count)), budget);
            for (std::int64_t d = 0; d < selection.totalDofCount(); ++d)
                equation_of_dof[d] = selection.equationOfDof(d);
            *kept = selection.keptDofCount();
        },
        err, errlen);
}

// SymmetricCsc::fromBlocks of the block matrix with `values` on the pattern of the graph, under the selection of
// ref_selection's arguments: *nnz, then (when rows is not null) offsets (kept + 1), rows, csc_values.
int ref_from_blocks(std::int32_t nodes, const std::int64_t* offsets, const std::int32_t* adjacency,
                    const std::uint8_t* dof_counts, const double* values, const std::uint8_t* eliminated,
                    std::int64_t total_dofs, const std::int32_t* node_order, std::int32_t order_count, int threads,
                    std::int64_t* nnz, std::int64_t* csc_offsets, std::int32_t* rows, double* csc_values, char* err,
                    int errlen)
{
    return guarded(
        [&]
        {
            const PatternOf p(nodes, offsets, adjacency, dof_counts);
            base::MemoryBudget budget(BUDGET);
            linalg::BlockSparseMatrix matrix(*p.pattern, budget, "blocks");
            std::copy_n(values, static_cast<std::size_t>(p.pattern->valueCount()), matrix.values().data());
            const linalg::DofSelection selection(
                std::span<const std::uint8_t>(dof_counts, static_cast<std::size_t>(nodes)),
                std::span<const std::uint8_t>(eliminated, eliminated != nullptr ? static_cast<std::size_t>(total_dofs)
                                                                                 : 0),
                std::span<const std::int32_t>(node_order, static_cast<std::size_t>(order_count)), budget);
            std::unique_ptr<base::ThreadPool> pool;
            if (threads > 1)
                pool = std::make_unique<base::ThreadPool>(static_cast<std::size_t>(threads));
            const linalg::SymmetricCsc csc = linalg::SymmetricCsc::fromBlocks(matrix, selection, budget, pool.get());
            *nnz = csc.nonZeroCount();
            if (rows == nullptr)
                return;
            std::copy(csc.columnOffsets().begin(), csc.columnOffsets().end(), csc_offsets);
            std::copy(csc.rowIndices().begin(), csc.rowIndices().end(), rows);
            std::copy(csc.values().begin(), csc.values().end(), csc_values);
        },
        err, errlen);
}

// y = A·x with A symmetric, stored as its lower triangle (the reference's fixed term order).
int ref_spmv(std::int32_t n, const std::int64_t* offsets, const std::int32_t* rows, const double* values,
             const double* x, double* y, char* err, int errlen)
{
    return guarded(
        [&]
        {
            base::MemoryBudget budget(BUDGET);
            const std::int64_t nnz = offsets[n];
            const linalg::SymmetricCsc matrix = linalg::SymmetricCsc::fromCompressedColumns(
                n, std::span<const std::int64_t>(offsets, static_cast<std::size_t>(n) + 1),
                std::span<const std::int32_t>(rows, static_cast<std::size_t>(nnz)),
                std::span<const double>(values, static_cast<std::size_t>(nnz)), budget);
            matrix.multiply(std::span<const double>(x, static_cast<std::size_t>(n)),
                            std::span<double>(y, static_cast<std::size_t>(n)));
        },
        err, errlen);
}

namespace
{

linalg::SymmetricCsc cscOf(std::int32_t n, const std::int64_t* offsets, const std::int32_t* rows, const double* values,
                           base::MemoryBudget& budget)
{
    const std::int64_t nnz = offsets[n];
    return linalg::SymmetricCsc::fromCompressedColumns(
        n, std::span<const std::int64_t>(offsets, static_cast<std::size_t>(n) + 1),
        std::span<const std::int32_t>(rows, static_cast<std::size_t>(nnz)),
        std::span<const double>(values, static_cast<std::size_t>(nnz)), budget);
}

} // namespace

// out = alpha·A + beta·B (one structure, the values of A and B given).
int ref_combine(std::int32_t n, const std::int64_t* offsets, const std::int32_t* rows, double alpha,
                const double* a_values, double beta, const double* b_values, double* out, char* err, int errlen)
{
    return guarded(
        [&]
        {
            base::MemoryBudget budget(BUDGET);
            const linalg::SymmetricCsc a = cscOf(n, offsets, rows, a_values, budget);
            const linalg::SymmetricCsc b = cscOf(n, offsets, rows, b_values, budget);
            const linalg::SymmetricCsc c = linalg::SymmetricCsc::combine(alpha, a, beta, b, budget);
            const auto v = c.values();
            std::copy(v.begin(), v.end(), out);
        },
        err, errlen);
}

// values ← S·A·S in place.
int ref_scale_symmetric(std::int32_t n, const std::int64_t* offsets, const std::int32_t* rows, double* values,
                        const double* scale, char* err, int errlen)
{
    return guarded(
        [&]
        {
            base::MemoryBudget budget(BUDGET);
            linalg::SymmetricCsc a = cscOf(n, offsets, rows, values, budget);
            a.scaleSymmetric(std::span<const double>(scale, static_cast<std::size_t>(n)));
            const auto v = a.values();
            std::copy(v.begin(), v.end(), values);
        },
        err, errlen);
}

// scale = pencilScaleOf(K).
int ref_pencil_scale_of(std::int32_t n, const std::int64_t* offsets, const std::int32_t* rows, const double* values,
                        double* scale, char* err, int errlen)
{
    return guarded(
        [&]
        {
            base::MemoryBudget budget(BUDGET);
            const linalg::SymmetricCsc k = cscOf(n, offsets, rows, values, budget);
            const base::BudgetedBuffer<double> s = linalg::pencilScaleOf(k, budget);
            std::copy(s.data(), s.data() + n, scale);
        },
        err, errlen);
}

// The reference family's in-place LDLᵀ of a front in two parts (SplitLdltOperands): returns the columns eliminated.
std::int64_t ref_dense_front_nopivot(double* panel, std::int64_t ldp, double* contribution, std::int64_t ldc,
                                     std::int64_t c, std::int64_t r)
{
    linalg::SplitLdltOperands operands;
    operands.panel = panel;
    operands.ldp = static_cast<std::size_t>(ldp);
    operands.contribution = contribution;
    operands.ldc = static_cast<std::size_t>(ldc);
    operands.pivotColumnCount = static_cast<std::size_t>(c);
    operands.rowCount = static_cast<std::size_t>(r);
    return static_cast<std::int64_t>(linalg::kernels::reference::factorFrontWithoutPivoting(operands));
}

// The reference family's Bunch-Kaufman LDLᵀ with candidates (PivotedLdltOperands): returns the columns eliminated.
std::int64_t ref_dense_bunch_kaufman(double* a, std::int64_t lda, std::int64_t n, std::int64_t candidates,
                                     std::int64_t* permutation, double* subdiagonal)
{
    std::vector<std::size_t> perm(static_cast<std::size_t>(candidates));
    linalg::PivotedLdltOperands operands;
    operands.a = a;
    operands.lda = static_cast<std::size_t>(lda);
    operands.size = static_cast<std::size_t>(n);
    operands.permutation = perm.data();
    operands.subdiagonal = subdiagonal;
    operands.candidateCount = static_cast<std::size_t>(candidates);
    const std::size_t eliminated = linalg::kernels::reference::factorWithPivoting(operands);
    for (std::int64_t k = 0; k < candidates; ++k)
        permutation[k] = static_cast<std::int64_t>(perm[static_cast<std::size_t>(k)]);
    return static_cast<std::int64_t>(eliminated);
}

int ref_hardware_threads()
{
    return static_cast<int>(std::thread::hardware_concurrency());
}

} // extern "C"
