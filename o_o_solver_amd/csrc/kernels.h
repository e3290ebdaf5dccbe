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
constexpr int ASSEMBLY_BLOCK_VALUES = 8192;
// Supernodes with at least this many rows below their columns spread the products of their sweeps over workgroups.
constexpr int BIG_ROWS = 256;
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
};

// A matrix in device memory: its lower CSC, the front position of each entry (within the front of the supernode
// owning its column) and, for the FP32 factor, the power-of-two scale of each equation (null: no scaling).
struct DeviceMatrix
{
    const int64_t* offsets;
    const int32_t* rows;
    const double* values;
    const int32_t* position;
    const double* scale;
};

struct AssemblyBlock
{
    int32_t supernode, j0, j1, i0, i1, pad;
};

// One 64 × 64 MFMA tile (element offsets): mode 0 C −= A·diag(d)·Bᵀ with A, B, d in the panels and C in the panels
// (target 0) or the update matrices (target 1); mode 1 C = A·Bᵀ·diag(d)⁻¹ with B an inverse of a diagonal block.
struct UpdateTile
{
    int64_t cOff, aOff, bOff, dOff, ld, ldb, ldc;
    int32_t m, n, k, target, lower, mode;
};

template <typename T>
hipError_t launchSmallFronts(const DevicePlan& p, const DeviceMatrix& a, T* panels, T* updates, T* linv,
                             const int32_t* list, int count, int maxOrder, double* diagOut, int32_t* zeroAt,
                             hipStream_t stream);
template <typename T> size_t smallFrontBytes(int maxOrder);
template <typename T>
hipError_t launchAssembly(const DevicePlan& p, const DeviceMatrix& a, T* panels, T* updates,
                          const AssemblyBlock* blocks, int count, hipStream_t stream);
template <typename T>
hipError_t launchDiagonalBlocks(const DevicePlan& p, T* panels, T* linv, const int32_t* fronts, int frontCount, int k0,
                                double* diagOut, int32_t* zeroAt, hipStream_t stream);
template <typename T>
hipError_t launchUpdateTiles(T* panels, T* updates, const T* linv, const UpdateTile* tiles, int count,
                             hipStream_t stream);
// One level of the forward sweep: the supernodes of `list`, then the row tiles (supernode, tile) of its big ones.
template <typename TP>
hipError_t launchForward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, double* U, int k,
                         const int32_t* list, int count, const int2* rowTiles, int tileCount, hipStream_t stream);
// One level of the backward sweep: the column blocks (supernode, block) of its big supernodes, then all of `list`.
template <typename TP>
hipError_t launchBackward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, int k, const int32_t* list,
                          int count, const int2* columnBlocks, int blockCount, hipStream_t stream);
// One step of the sweeps of the wide supernodes of a level (items (supernode, tile or block), −1: finalise a block).
template <typename TP>
hipError_t launchForwardWide(const DevicePlan& p, const TP* panels, const TP* linv, double* X, double* U, int k,
                             const int2* items, int count, int step, hipStream_t stream);
template <typename TP>
hipError_t launchBackwardWide(const DevicePlan& p, const TP* panels, const TP* linv, double* X, int k,
                              const int2* items, int count, int step, hipStream_t stream);
hipError_t launchDiagonalSolve(double* X, const double* d, int64_t n, int k, hipStream_t stream);

hipError_t launchFillInt32(int32_t* data, int64_t count, int32_t value, hipStream_t stream);
hipError_t launchToRowMajor(const double* src, int64_t ld, double* dst, int64_t n, int k, const double* scale,
                            hipStream_t stream);
hipError_t launchFromRowMajor(const double* src, double* dst, int64_t ld, int64_t n, int k, const double* scale,
                              int accumulate, hipStream_t stream);
hipError_t launchResidual(const int64_t* offsets, const int32_t* cols, const double* values, const double* x,
                          const double* b, double* out, int64_t ld, int64_t n, int k, hipStream_t stream);
hipError_t launchColumnMax(const double* v, int64_t ld, int64_t n, int k, unsigned long long* out,
                           hipStream_t stream);
hipError_t launchFillDouble(double* data, int64_t ld, int64_t n, int k, double value, hipStream_t stream);
hipError_t launchAxpyColumns(double* x, const double* d, int64_t ld, int64_t n, int k, hipStream_t stream);
hipError_t launchMfmaProbe(const double* a, const double* b, double* d, const float* af, const float* bf, float* df,
                           hipStream_t stream);

} // namespace oosg
