// Device kernels of the numeric multifrontal LDLᵀ for gfx950 (CDNA4): 64-wide wavefronts, MFMA, LDS.
//
// Data layout in HBM (element offsets fixed by the plan, the same for FP64 and FP32 storage):
//   panels   supernode s at panel_off[s]: its front's first c columns over all n = c + r rows, column-major,
//            leading dimension n (the reference's FactorPanel, include/oos/linalg/supernodal/SupernodalFactor.hpp:
//            110-125): L below the diagonal, D on it, zero above.
//   updates  supernode s at update_off[s]: its update matrix r × r, lower triangle, column-major, leading dimension r
//            (the trailing block of the front, MultifrontalDriver.cpp:257-284), alive until its parent is assembled.
//   X        right-hand sides n × k row-major (the k values of an equation side by side).
//   U        per supernode the r × k update vector of the forward sweep, row-major, at row_start[s]·k.
//
// Arithmetic: each front element is assembled in the reference's order — zero, the matrix entries of its column,
// then the children's update matrices in increasing child order (src/linalg/supernodal/FrontAssembly.cpp:166-300) —
// and eliminated by the LDLᵀ of src/linalg/kernels/reference/DenseBlockKernels.cpp:368-446 (L(i, k) = a(i, k)/d_k,
// W(j, k) = L(j, k)·d_k, a(i, j) −= L(i, k)·W(j, k)). Blocked fronts sum the terms of a block with MFMA, whose
// rounding differs from the reference's chains: results agree to rounding, which is the parity contract of the GPU
// factor (inertia and structure are exact).

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <utility>

#include "kernels.h"

namespace oosg
{

typedef double double4_t __attribute__((ext_vector_type(4)));
typedef float float4_t __attribute__((ext_vector_type(4)));

// ---- helpers -------------------------------------------------------------------------------------------------------

template <typename T> __device__ inline bool isBadPivot(T v)
{
    return v == T(0) || !isfinite(v);
}

// The first position of `value` or more in the increasing rel[0, count).
__device__ inline int lowerBound(const int32_t* rel, int count, int value)
{
    int lo = 0, hi = count;
    while (lo < hi)
    {
        const int mid = (lo + hi) >> 1;
        if (rel[mid] < value)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// f(i, j) for i in [0, rows), j in [0, cols), spread over the workgroup without integer division: rows across the
// threads when they are many, else rows across the lanes of a wave and columns across the waves.
template <typename F> __device__ inline void forEach2D(int rows, int cols, F f)
{
    if (rows >= int(blockDim.x))
    {
        for (int j = 0; j < cols; ++j)
            for (int i = threadIdx.x; i < rows; i += blockDim.x)
                f(i, j);
    }
    else
    {
        const int lane = threadIdx.x & 63, groups = blockDim.x >> 6;
        for (int j = threadIdx.x >> 6; j < cols; j += groups)
            for (int i = lane; i < rows; i += 64)
                f(i, j);
    }
}

// The rank-one step of column k on the lower triangle of rows and columns [k + 1, end): F(i, j) −= L(i, k)·W(j, k),
// W(j, k) = L(j, k)·d_k (L(·, k) in lk), the reference's chain (DenseBlockKernels.cpp:376-390); a column per wave,
// its rows across the lanes.
template <typename T> __device__ inline void rankOneStep(T* F, int ld, int k, int end, const T* lk, T pivot)
{
    const int lane = threadIdx.x & 63, groups = blockDim.x >> 6;
    for (int j = k + 1 + (threadIdx.x >> 6); j < end; j += groups)
    {
        const T w = lk[j] * pivot;
        for (int i = j + lane; i < end; i += 64)
            F[i + j * ld] -= lk[i] * w;
    }
}

// The same restricted to the panel of a small front: columns (k, cols), rows [j, rows).
template <typename T>
__device__ inline void panelRankOneStep(T* F, int ld, int k, int cols, int rows, const T* lk, T pivot)
{
    const int lane = threadIdx.x & 63, groups = blockDim.x >> 6;
    for (int j = k + 1 + (threadIdx.x >> 6); j < cols; j += groups)
    {
        const T w = lk[j] * pivot;
        for (int i = j + lane; i < rows; i += 64)
            F[i + j * ld] -= lk[i] * w;
    }
}

// The value as an opaque register: a load whose result feeds a select stays unconditional (the compiler would
// otherwise move it into a branch and wait for it there), so the loads of a loop issue back to back.
template <typename T> __device__ inline T pinned(T v)
{
    asm volatile("" : "+v"(v));
    return v;
}

__global__ void fillInt32Kernel(int32_t* data, int64_t count, int32_t value)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
        data[i] = value;
}

// ---- unit lower inverses of the diagonal blocks ------------------------------------------------------------------
//
// The inverse X = L_JJ⁻¹ of each diagonal block of PANEL_BLOCK columns is accumulated during the elimination itself
// (Gauss-Jordan on [L | I]): once column k of L is known, X(i, j) −= L(i, k)·X(k, j) for the rows i > k of its block
// and the columns j <= k, every pair in parallel (X(k, k) = 1). X(i, j), i > j, is kept at F[j + i·ld], in the block's
// upper triangle, which the elimination leaves free. The inverses turn the panel solves into MFMA products and the
// in-block triangles of the sweeps into small products.

// The step of column k (local to a block starting at base = q0 + q0·ld, kb columns): `lk` holds L(i, k) at index i.
template <typename T>
__device__ inline void inverseStep(T* F, int ld, int q0, int kb, int k, const T* lk)
{
    const int lane = threadIdx.x & 63, groups = blockDim.x >> 6;
    for (int j = q0 + (threadIdx.x >> 6); j <= k; j += groups)
    {
        const T xkj = j == k ? T(1) : F[j + k * ld];
        for (int i = k + 1 + lane; i < q0 + kb; i += 64)
            F[j + i * ld] -= lk[i] * xkj;
    }
}

// Writes the inverse accumulated in the upper triangle of the block at (q0, q0): column-major kb × kb, with its unit
// diagonal and zeros above.
template <typename T>
__device__ inline void writeInverse(const T* F, int ld, int q0, int kb, T* out)
{
    forEach2D(kb, kb, [&](int i, int j) { out[i + j * kb] = i == j ? T(1) : (i > j ? F[(q0 + j) + (q0 + i) * ld] : T(0)); });
}

// ---- MFMA blocks ---------------------------------------------------------------------------------------------

template <typename T> struct Mfma;

// Operands of a 16 × 16 × 4 block: lane l holds A(l mod 16, l / 16) and B(l / 16, l mod 16); its four results are
// D(row(l, v), l mod 16), v = 0..3, with the row mapping of each instruction (checked on the device by
// oosg_debug_mfma_probe, which writes through these same functions).
template <> struct Mfma<double>
{
    typedef double4_t Acc;
    __device__ static inline Acc step(double a, double b, Acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

template <> struct Mfma<float>
{
    typedef float4_t Acc;
    __device__ static inline Acc step(float a, float b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// ---- small fronts: assembly and elimination fused, the whole front in LDS ------------------------------------------

// One workgroup per front of at most SMALL_FRONT_MAX rows: the front is assembled in LDS (zero, matrix entries,
// children in order), its c columns are eliminated right-looking within the panel (two barriers per column), the
// update matrix is one MFMA product from the panel, and the panel, the update matrix and the inverses of the diagonal
// blocks are written once.
template <typename T>
__global__ __launch_bounds__(256) void smallFrontKernel(DevicePlan p, DeviceMatrix a, T* panels, T* updates, T* linv,
                                                        const int32_t* list, double* diagOut, int32_t* zeroAt)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T* F = reinterpret_cast<T*>(smem);
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    T* lcol = F + n * n;
    const int tid = threadIdx.x, nt = blockDim.x;

    for (int idx = tid; idx < n * n; idx += nt)
        F[idx] = T(0);
    __syncthreads();
    // The matrix entries of the supernode's columns, one range of the CSC (distinct positions: no two threads meet).
    {
        const int64_t kb = a.offsets[first], ke = a.offsets[first + c];
        for (int64_t k = kb + tid; k < ke; k += nt)
        {
            const int j = a.column[k];
            double v = a.values[k];
            if (a.scale)
                v = v * a.scale[a.rows[k]] * a.scale[first + j];
            F[a.position[k] + j * n] += T(v);
        }
    }
    // The children's update matrices, one after the other.
    for (int q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
    {
        __syncthreads();
        const int ch = p.children[q];
        const int rc = p.nrow[ch];
        const int32_t* rel = p.rel + p.rowStart[ch];
        const T* u = updates + p.updateOff[ch];
        // The child's lower part by consecutive rows, eight entries per thread per pass, every load issued before any
        // is used (a child's entries land on distinct positions).
        constexpr int UNROLL = 8;
        const int total = rc * rc;
        for (int base = tid; base < total; base += 256 * UNROLL)
        {
            int ri[UNROLL], ci[UNROLL];
            T v[UNROLL];
            bool in[UNROLL];
#pragma unroll
            for (int e = 0; e < UNROLL; ++e)
            {
                const int idx = base + 256 * e;
                const int jj = min(idx / rc, rc - 1), i = idx % rc;
                in[e] = idx < total && i >= jj;
                const int ic = in[e] ? i : jj;
                ri[e] = rel[ic];
                ci[e] = rel[jj];
                v[e] = u[ic + jj * rc];
            }
#pragma unroll
            for (int e = 0; e < UNROLL; ++e)
                if (in[e])
                    F[pinned(ri[e]) + pinned(ci[e]) * n] += pinned(v[e]);
        }
    }
    __syncthreads();

    // Right-looking elimination of the c columns within the panel (the reference's chain, column by column).
    for (int k = 0; k < c; ++k)
    {
        const T pivot = F[k + k * n];
        if (tid == 0 && isBadPivot(pivot))
            atomicMin(&zeroAt[s], k);
        for (int i = k + 1 + tid; i < n; i += nt)
            lcol[i] = F[i + k * n] / pivot;
        __syncthreads();
        panelRankOneStep(F, n, k, c, n, lcol, pivot);
        for (int i = k + 1 + tid; i < n; i += nt)
            F[i + k * n] = lcol[i];
        const int q0 = (k / PANEL_BLOCK) * PANEL_BLOCK;
        inverseStep(F, n, q0, min(PANEL_BLOCK, c - q0), k, lcol);
        __syncthreads();
    }

    T* panel = panels + p.panelOff[s];
    forEach2D(n, c, [&](int i, int j) { panel[i + j * n] = i < j ? T(0) : F[i + j * n]; });
    // The update matrix U = F_RR − L_R·D·L_Rᵀ as one product on 16 × 16 MFMA tiles (its lower tiles over the four
    // waves), K = c in steps of four, from the front in LDS.
    T* update = updates + p.updateOff[s];
    {
        typedef typename Mfma<T>::Acc Acc;
        const int wave = tid >> 6, lane = tid & 63;
        const int tiles = (r + 15) / 16;
        for (int q = wave; q < tiles * (tiles + 1) / 2; q += 4)
        {
            int J = 0, I = q;
            while (I >= tiles - J)
            {
                I -= tiles - J;
                ++J;
            }
            I += J;
            const int col = 16 * J + (lane & 15), colIn = min(col, r - 1);
            Acc acc;
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int row = min(16 * I + Mfma<T>::row(lane, v), r - 1);
                acc[v] = F[(c + row) + (c + colIn) * n];
            }
            const int rowA = min(16 * I + (lane & 15), r - 1);
            for (int k0 = 0; k0 < c; k0 += 4)
            {
                const int k = k0 + (lane >> 4), kIn = min(k, c - 1);
                const T a = F[(c + rowA) + kIn * n], b = F[(c + colIn) + kIn * n] * F[kIn + kIn * n];
                acc = Mfma<T>::step(k < c ? -a : T(0), k < c ? b : T(0), acc);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int row = 16 * I + Mfma<T>::row(lane, v);
                if (row < r && col < r && row >= col)
                    update[row + col * r] = acc[v];
            }
        }
    }
    for (int k = tid; k < c; k += nt)
        diagOut[first + k] = double(F[k + k * n]);
    for (int q0 = 0; q0 < c; q0 += PANEL_BLOCK)
        writeInverse(F, n, q0, min(PANEL_BLOCK, c - q0), linv + p.linvOff[s] + q0 * PANEL_BLOCK);
}

// ---- large fronts --------------------------------------------------------------------------------------------------

// Assembly of a block of a front: columns [j0, j1), rows [i0, i1), accumulated in LDS in the reference's order (zero,
// the matrix entries, the children in increasing order), then its lower part written to the panel (columns < c; the
// panel's values above the diagonal are never read, nor written) or to the update matrix. The matrix entries of the block's columns are one range
// of the CSC, spread over the threads; a child's part, its columns by consecutive rows, too, so that every thread has
// several independent loads in flight whatever the block's shape.
template <typename T>
__global__ __launch_bounds__(256) void assembleKernel(DevicePlan p, DeviceMatrix a, T* panels, T* updates,
                                                      const AssemblyBlock* blocks, const AssemblyItem* items)
{
    extern __shared__ __align__(16) unsigned char smem[];
    T* B = reinterpret_cast<T*>(smem);
    const AssemblyBlock blk = blocks[blockIdx.x];
    const int s = blk.supernode;
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    const int rowsIn = blk.i1 - blk.i0, colsIn = blk.j1 - blk.j0;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < rowsIn * colsIn; idx += nt)
        B[idx] = T(0);
    __syncthreads();
    if (blk.j0 < c)
    {
        const int64_t kb = a.offsets[first + blk.j0], ke = a.offsets[first + min(blk.j1, c)];
        for (int64_t k = kb + tid; k < ke; k += nt)
        {
            const int pos = a.position[k];
            if (pos < blk.i0 || pos >= blk.i1)
                continue;
            const int j = a.column[k];
            double v = a.values[k];
            if (a.scale)
                v = v * a.scale[a.rows[k]] * a.scale[first + j];
            B[(pos - blk.i0) + (j - blk.j0) * rowsIn] += T(v);
        }
    }
    for (int q = blk.itemStart; q < blk.itemEnd; ++q)
    {
        __syncthreads();
        const AssemblyItem it = items[q];
        const int rc = p.nrow[it.child];
        const int32_t* rel = p.rel + p.rowStart[it.child];
        const T* u = updates + p.updateOff[it.child];
        // The child's columns [q0, q1) × rows [ia, ib) (its lower part), spread over all the threads by consecutive
        // rows; eight entries per thread per pass, every load issued before any is used.
        const int h = it.ib - it.ia, total = h * (it.q1 - it.q0);
        constexpr int UNROLL = 8;
        for (int base = tid; base < total; base += 256 * UNROLL)
        {
            int ci[UNROLL], ri[UNROLL];
            T v[UNROLL];
            bool in[UNROLL];
#pragma unroll
            for (int e = 0; e < UNROLL; ++e)
            {
                const int idx = base + 256 * e;
                const int jj = it.q0 + idx / h, i = it.ia + idx % h;
                in[e] = idx < total && i >= jj;
                const int jc = min(jj, it.q1 - 1), ic = in[e] ? i : it.ia;
                ci[e] = rel[jc];
                ri[e] = rel[ic];
                v[e] = u[ic + int64_t(jc) * rc];
            }
#pragma unroll
            for (int e = 0; e < UNROLL; ++e)
                if (in[e])
                    B[(pinned(ri[e]) - blk.i0) + (pinned(ci[e]) - blk.j0) * rowsIn] += pinned(v[e]);
        }
    }
    __syncthreads();
    // The offsets read once: the stores below could alias the plan's arrays as far as the compiler knows.
    T* const panel = panels + p.panelOff[s];
    T* const update = updates + p.updateOff[s];
    forEach2D(rowsIn, colsIn, [&](int il, int jl) {
        const int i = blk.i0 + il, j = blk.j0 + jl;
        if (i < j)
            return;
        if (j < c)
            panel[i + int64_t(j) * n] = B[il + jl * rowsIn];
        else
            update[(i - c) + int64_t(j - c) * r] = B[il + jl * rowsIn];
    });
}

// The value of lane `lane` (uniform across the wave) on every lane: v_readlane into a scalar register, a few cycles,
// where a general shuffle (ds_bpermute) waits on LDS latency.
__device__ inline float broadcastLane(float v, int lane)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, v), lane));
}
__device__ inline double broadcastLane(double v, int lane)
{
    const long long bits = __builtin_bit_cast(long long, v);
    const unsigned lo = unsigned(__builtin_amdgcn_readlane(int(bits), lane));
    const unsigned hi = unsigned(__builtin_amdgcn_readlane(int(bits >> 32), lane));
    return __builtin_bit_cast(double, (static_cast<unsigned long long>(hi) << 32) | lo);
}

// MFMA tiles, 64 × 64 per workgroup, four waves of 32 × 32 in 16 × 16 × 4 blocks:
//   mode 0  C −= A·diag(d)·Bᵀ  the trailing update of a panel step, and the update matrix of a front;
//   mode 1  C  = A·Bᵀ·diag(d)⁻¹ with B = L11⁻¹ of the diagonal block: the panel solve (rows below a factorised
//           diagonal block, L(i, blk) = F(i, blk)·L11⁻ᵀ·D⁻¹), in place (C = A: the tile reads its rows before writing).
// A(i, p) = panels[aOff + i + p·ld]; mode 0: B(j, p) = panels[bOff + j + p·ld]; mode 1: B(j, p) = linv[bOff + j + p·ldb];
// d_p = panels[dOff + p·(ld + 1)].

// The diagonal block [k0, k0 + kb) of a front's panel factorised in place (kb <= PANEL_BLOCK), with the inverse of
// its unit lower factor; one workgroup of four waves per front, the block in LDS (rows and columns from kb on hold
// the identity, which the elimination leaves unchanged):
//   elimination by sub-panels of 16 columns: wave 0 eliminates the sub-panel's columns one after the other, a row per
//     lane, the pivot d_k and W(j, k) = L(j, k)·d_k broadcast from their lanes by v_readlane (the reference's chain,
//     DenseBlockKernels.cpp:376-390, inside the sub-panel); then the waves subtract L·W of the sub-panel from the
//     trailing lower 16 × 16 tiles with MFMA;
//   inverse X = L⁻¹: the four diagonal tiles by forward substitution (wave w: X_ww, lane j its column j), then the
//     tiles below, by distance d = I − J = 1, 2, 3: X_IJ = −X_II·(Σ_{K=J}^{I−1} L_IK·X_KJ), two MFMA products, the
//     first one's accumulator used directly as the second one's operand (its k-steps follow the accumulator rows).
// LDS: S[j·LD + i] holds L(i, j) for i > j, D(j) for i = j, and X(j, i) for i < j (the inverse, transposed, in the
// upper triangle, which the elimination leaves free).
// Factors the block held in S (LD = PANEL_BLOCK + 1: S[j·LD + i] = F(k0 + i, k0 + j) for i >= j, the identity
// outside the kb × kb block) and writes it back: L and D into the panel, L11⁻¹ into the inverses, D into diagOut; a
// zero or non-finite pivot is recorded in zeroAt. One workgroup of four waves.
template <typename T>
__device__ void factorBlockInLds(T* S, const DevicePlan& p, int s, int k0, int kb, T* panels, T* linv,
                                 double* diagOut, int32_t* zeroAt);

template <typename T>
__global__ __launch_bounds__(256, 2) void diagonalBlockKernel(DevicePlan p, T* panels, T* linv, const int32_t* list,
                                                           int k0, double* diagOut, int32_t* zeroAt)
{
    constexpr int B = PANEL_BLOCK, LD = PANEL_BLOCK + 1;
    __shared__ T S[B * LD];
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], n = c + p.nrow[s];
    const int kb = min(B, c - k0);
    T* panel = panels + p.panelOff[s];
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int64_t base = k0 + int64_t(k0) * n;

    // Every load is issued (from the block's first element where the position is outside the lower triangle) and
    // the value selected afterwards: no branch waits on an earlier load.
    // Thread (wave, lane) moves rows i = lane of the columns j = wave + 4q.
    {
        const T* src = panel + base + int64_t(wave) * n;
        T v[B / 4];
#pragma unroll
        for (int q = 0; q < B / 4; ++q)
        {
            // Columns from kb on lie outside the panel (past the end of the panel array for the last supernode):
            // their loads read the block's first element instead.
            const int j = wave + 4 * q;
            const T* column = j < kb ? src : panel + base;
            v[q] = column[(lane < kb && j < kb && lane >= j) ? lane : 0];
            src += 4 * int64_t(n);
        }
#pragma unroll
        for (int q = 0; q < B / 4; ++q)
        {
            const int j = wave + 4 * q;
            const bool inside = lane < kb && j < kb && lane >= j;
            S[j * LD + lane] = inside ? pinned(v[q]) : T(lane == j ? 1 : 0);
        }
    }
    __syncthreads();
    factorBlockInLds(S, p, s, k0, kb, panels, linv, diagOut, zeroAt);
}

template <typename T>
__device__ void factorBlockInLds(T* S, const DevicePlan& p, int s, int k0, int kb, T* panels, T* linv,
                                 double* diagOut, int32_t* zeroAt)
{
    constexpr int B = PANEL_BLOCK, LD = PANEL_BLOCK + 1, SP = 16;
    static_assert(B == 64, "four sub-panels of 16 columns");
    typedef typename Mfma<T>::Acc Acc;
    const int c = p.ncol[s], n = c + p.nrow[s];
    T* panel = panels + p.panelOff[s];
    const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
    const int64_t base = k0 + int64_t(k0) * n;

    for (int c0 = 0; c0 < kb; c0 += SP)
    {
        if (wave == 0)
        {
            T a[SP]; // row `lane` of the sub-panel
#pragma unroll
            for (int t = 0; t < SP; ++t)
                a[t] = S[(c0 + t) * LD + lane];
#pragma unroll
            for (int k = 0; k < SP; ++k)
            {
                const int K = c0 + k;
                if (K >= kb)
                    break; // the identity beyond the block: nothing to eliminate
                const T pivot = broadcastLane(a[k], K);
                if (lane == 0 && isBadPivot(pivot))
                    atomicMin(&zeroAt[s], k0 + K);
                const T l = a[k] / pivot; // L(lane, K) on the lanes below K
                const T w = l * pivot;
#pragma unroll
                for (int t = k + 1; t < SP; ++t)
                    a[t] -= l * broadcastLane(w, c0 + t);
                if (lane > K)
                    a[k] = l;
            }
#pragma unroll
            for (int t = 0; t < SP; ++t)
                if (lane >= c0 + t)
                    S[(c0 + t) * LD + lane] = a[t];
        }
        __syncthreads();
        // Trailing tiles (I, J), c0/16 < J <= I < 4, in the order J, then I; those whose rows lie beyond the block
        // hold the identity, which the update (with L zero there) leaves as it is.
        const int P = c0 / SP;
        const int tiles = (3 - P) * (4 - P) / 2;
        for (int q = wave; q < tiles; q += 4)
        {
            int J = P + 1, I = q;
            while (I >= 4 - J)
            {
                I -= 4 - J;
                ++J;
            }
            I += J;
            if (SP * I >= kb)
                continue;
            const int col = SP * J + (lane & 15);
            Acc acc;
#pragma unroll
            for (int v = 0; v < 4; ++v)
                acc[v] = S[col * LD + SP * I + Mfma<T>::row(lane, v)];
#pragma unroll
            for (int kk = 0; kk < SP; kk += 4)
            {
                const int pc = c0 + kk + (lane >> 4);
                const T av = -S[pc * LD + SP * I + (lane & 15)];
                const T bv = S[pc * LD + col] * S[pc * LD + pc];
                acc = Mfma<T>::step(av, bv, acc);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
                S[col * LD + SP * I + Mfma<T>::row(lane, v)] = acc[v];
        }
        __syncthreads();
    }

    // Diagonal tiles of the inverse: lane j (lanes from 16 on repeat lane j mod 16 and store nothing); a tile beyond
    // the block is the identity's, whose inverse's zeros the upper triangle already holds.
    if (SP * wave < kb)
    {
        const int J0 = SP * wave, j = lane & 15;
        T x[SP];
#pragma unroll
        for (int i = 0; i < SP; ++i)
            x[i] = T(i == j ? 1 : 0);
#pragma unroll
        for (int k = 0; k < SP; ++k)
        {
#pragma unroll
            for (int i = k + 1; i < SP; ++i)
                x[i] -= S[(J0 + k) * LD + J0 + i] * x[k];
            asm volatile("" ::: "memory"); // one column's reads in flight at a time (registers)
        }
        if (lane < SP)
        {
#pragma unroll
            for (int i = 1; i < SP; ++i)
                if (i > j)
                    S[(J0 + i) * LD + J0 + j] = x[i];
        }
    }
    __syncthreads();
    for (int d = 1; d < 4; ++d)
    {
        // Tiles whose rows lie beyond the block: L is zero there, and so is X.
        if (wave < 4 - d && SP * (wave + d) < kb)
        {
            const int J = wave, I = wave + d;
            const int m = lane & 15;
            Acc t = {0, 0, 0, 0};
            for (int K = J; K < I; ++K)
            {
#pragma unroll
                for (int kk = 0; kk < SP; kk += 4)
                {
                    const int rk = SP * K + kk + (lane >> 4), cj = SP * J + m;
                    const T av = S[rk * LD + SP * I + m];                                 // L(16I + m, rk)
                    const T bv = rk > cj ? S[rk * LD + cj] : T(rk == cj ? 1 : 0);           // X(rk, cj)
                    t = Mfma<T>::step(av, bv, t);
                }
            }
            Acc u = {0, 0, 0, 0};
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int kr = Mfma<T>::row(lane, v);
                const T xv = m > kr ? S[(SP * I + m) * LD + SP * I + kr] : T(m == kr ? 1 : 0); // X_II(m, kr)
                u = Mfma<T>::step(xv, t[v], u);
            }
#pragma unroll
            for (int v = 0; v < 4; ++v)
                S[(SP * I + Mfma<T>::row(lane, v)) * LD + SP * J + m] = -u[v];
        }
        __syncthreads();
    }

    T* out = linv + p.linvOff[s] + int64_t(k0) * B;
    T* dst = panel + base + int64_t(wave) * n;
    for (int j = wave; j < kb; j += 4)
    {
        if (lane < kb)
        {
            if (lane >= j)
                dst[lane] = S[j * LD + lane];
            out[lane + j * kb] = lane == j ? T(1) : (lane > j ? S[lane * LD + j] : T(0));
        }
        dst += 4 * int64_t(n);
    }
    if (tid < kb)
        diagOut[p.first[s] + k0 + tid] = double(S[tid * LD + tid]);
}

constexpr int TILE = 64;
constexpr int TILE_K = 32; // terms of a chunk staged in LDS

// The tile a workgroup takes: consecutive tiles of the list (the tiles of one front and panel step, which share their
// A and B blocks) go to the workgroups of one XCD (the hardware deals workgroups to the 8 XCDs in turn), so that the
// shared blocks are read into one L2.
__device__ inline int tileOfWorkgroup(int count)
{
    const int b = blockIdx.x, full = count - count % 8;
    return b < full ? (b & 7) * (full >> 3) + (b >> 3) : b;
}

// Thread t stages row t mod 64 of A and of B for the terms t / 64 + 4u (u < TILE_K / 4) of a chunk; each chunk's loads
// are issued while the previous chunk is multiplied (registers as the second buffer), from positions clamped into the
// operands; mode 0 accumulates onto C itself (the accumulators start from the C tile, A enters negated).
// PLAIN: C −= A·Bᵀ with B in the third array (leading dimension ldb) and no diagonal: the trailing updates of the
// pivoted fronts, whose B holds the rows of L·D already (D has 2 × 2 blocks there).
// Mode 2 (a diagonal tile of the next panel step's block): after its update the tile goes to LDS and its block is
// factorised there (factorBlockInLds, the diagonal-block step fused into the update that finishes the block);
// rowBase holds the block's first column and colBase its width.
template <typename T, bool PLAIN = false, bool FUSE = false>
__global__ __launch_bounds__(256, FUSE ? 2 : 1) void updateTilesKernel(DevicePlan p, T* panels, T* updates, T* linv,
                                                                       const UpdateTile* tiles, int count,
                                                                       double* diagOut, int32_t* zeroAt)
{
    constexpr int PER = TILE * TILE_K / 256;
    constexpr int SHARED = 2 * TILE_K * (TILE + 4) > TILE * (TILE + 1) ? 2 * TILE_K * (TILE + 4) : TILE * (TILE + 1);
    __shared__ T lds[SHARED];
    T (*As)[TILE + 4] = reinterpret_cast<T (*)[TILE + 4]>(lds);
    T (*Bs)[TILE + 4] = reinterpret_cast<T (*)[TILE + 4]>(lds + TILE_K * (TILE + 4));
    const UpdateTile t = tiles[tileOfWorkgroup(count)];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const bool solve = !PLAIN && t.mode == 1;
    const T* A = panels + t.aOff;
    const T* Bm = (solve || PLAIN ? linv : panels) + t.bOff;
    const T* dv = panels + t.dOff;
    const int64_t ld = t.ld, ldb = solve || PLAIN ? t.ldb : t.ld;
    T* C = (PLAIN || t.target == 0 ? panels : updates) + t.cOff;

    typename Mfma<T>::Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int i = min(wm + a * 16 + Mfma<T>::row(lane, v), t.m - 1);
                const int j = min(wn + b * 16 + (lane & 15), t.n - 1);
                acc[a][b][v] = solve ? T(0) : C[i + int64_t(j) * t.ldc];
            }

    const int row = tid & (TILE - 1), kq = tid >> 6;
    const T* Ar = A + min(row, t.m - 1);
    const T* Br = Bm + min(row, t.n - 1);
    // A solve divides its columns by their pivots: loaded now, used after the products.
    T pivots[2] = {T(1), T(1)};
    if (solve)
    {
#pragma unroll
        for (int b = 0; b < 2; ++b)
            pivots[b] = dv[min(wn + b * 16 + (lane & 15), t.n - 1) * (ld + 1)];
    }
    T ra[PER], rb[PER], rd[PER];
    const auto load = [&](int p0) {
#pragma unroll
        for (int u = 0; u < PER; ++u)
        {
            const int64_t pp = min(p0 + kq + 4 * u, t.k - 1);
            ra[u] = Ar[pp * ld];
            rb[u] = Br[pp * ldb];
            rd[u] = solve || PLAIN ? T(1) : dv[pp * (ld + 1)];
        }
    };
    load(0);
    for (int p0 = 0; p0 < t.k; p0 += TILE_K)
    {
        __syncthreads(); // the previous chunk's products have read the LDS
#pragma unroll
        for (int u = 0; u < PER; ++u)
        {
            const int kk = kq + 4 * u;
            const bool term = p0 + kk < t.k;
            const T av = pinned(ra[u]), bv = pinned(rb[u]), dd = pinned(rd[u]);
            As[kk][row] = term && row < t.m ? (solve ? av : -av) : T(0);
            Bs[kk][row] = term && row < t.n ? (solve || PLAIN ? bv : bv * dd) : T(0);
        }
        __syncthreads();
        if (p0 + TILE_K < t.k)
            load(p0 + TILE_K);
#pragma unroll
        for (int kk = 0; kk < TILE_K; kk += 4)
        {
            const int ka = kk + (lane >> 4);
            T av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
                av[a] = As[ka][wm + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < 2; ++b)
                bv[b] = Bs[ka][wn + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = Mfma<T>::step(av[a], bv[b], acc[a][b]);
        }
    }
    if (FUSE && t.mode == 2)
    {
        // The block [k0, k0 + kb)², now final, into LDS (the identity outside it), and factorised there.
        constexpr int LD = TILE + 1;
        const int kb = t.colBase;
        __syncthreads(); // the products have read the LDS
        for (int idx = tid; idx < TILE * LD; idx += 256)
            lds[idx] = T(idx / LD == idx % LD ? 1 : 0);
        __syncthreads();
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int i = wm + a * 16 + Mfma<T>::row(lane, v);
                    const int j = wn + b * 16 + (lane & 15);
                    if (i < kb && j < kb && i >= j)
                        lds[j * LD + i] = acc[a][b][v];
                    else if (i >= kb && i < t.m && j < t.n)
                        C[i + int64_t(j) * t.ldc] = acc[a][b][v]; // rows below a partial block: updated as usual
                }
        __syncthreads();
        factorBlockInLds(lds, p, t.supernode, t.rowBase, kb, panels, linv, diagOut, zeroAt);
        return;
    }
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int i = wm + a * 16 + Mfma<T>::row(lane, v);
                const int j = wn + b * 16 + (lane & 15);
                if (i >= t.m || j >= t.n || (t.lower && i < j))
                    continue;
                C[i + int64_t(j) * t.ldc] = solve ? acc[a][b][v] / pivots[b] : acc[a][b][v];
            }
}

// The update matrices U = −L_R·D·L_Rᵀ + Σ children's contributions (the extend-add of their update matrices into the
// rows and columns below the supernode's own), on TS × TS tiles, TS = 64 or 128: four waves of TS/2 × TS/2 (NB × NB
// MFMA blocks each), K staged through LDS by KC with the next chunk loaded into registers meanwhile. The products
// start from zero; then, a 64 × 64 quarter at a time, the tile goes through LDS, its children's parts are added in the
// children's order (the assembly items of that quarter, a barrier between children: each child's positions are
// distinct), and the quarter is written once. With 128 × 128 tiles each operand value loaded feeds twice the flops.
// The k-steps (four terms per MFMA, in increasing order) and the order of the additions do not depend on TS, so both
// sizes give the same bits.
template <typename T, int TS>
__global__ __launch_bounds__(256) void extendTilesKernel(DevicePlan p, T* panels, T* updates, const UpdateTile* tiles,
                                                         int count, const int32_t* itemStarts,
                                                         const AssemblyItem* items)
{
    constexpr int NB = TS / 32;            // MFMA blocks per wave and dimension
    constexpr int KC = TS == 64 ? 32 : 16; // terms of a chunk
    constexpr int PER = TS * KC / 256;     // operand values a thread stages per chunk
    constexpr int LD = TS + 4;
    constexpr int QLD = 65; // a 64 × 64 quarter in LDS
    constexpr int SHARED = 2 * KC * LD > 64 * QLD ? 2 * KC * LD : 64 * QLD;
    __shared__ T lds[SHARED];
    T (*As)[LD] = reinterpret_cast<T (*)[LD]>(lds);
    T (*Bs)[LD] = reinterpret_cast<T (*)[LD]>(lds + KC * LD);
    const UpdateTile t = tiles[tileOfWorkgroup(count)];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * (TS / 2), wn = (wave & 1) * (TS / 2);
    const T* A = panels + t.aOff;
    const T* Bm = panels + t.bOff;
    const T* dv = panels + t.dOff;
    const int64_t ld = t.ld;

    typename Mfma<T>::Acc acc[NB][NB];
#pragma unroll
    for (int a = 0; a < NB; ++a)
#pragma unroll
        for (int b = 0; b < NB; ++b)
            acc[a][b] = typename Mfma<T>::Acc{0, 0, 0, 0};

    const int row = tid & (TS - 1), kq = tid / TS, kstep = 256 / TS;
    const T* Ar = A + min(row, t.m - 1);
    const T* Br = Bm + min(row, t.n - 1);
    T ra[PER], rb[PER], rd[PER];
    const auto load = [&](int p0) {
#pragma unroll
        for (int u = 0; u < PER; ++u)
        {
            const int64_t pp = min(p0 + kq + kstep * u, t.k - 1);
            ra[u] = Ar[pp * ld];
            rb[u] = Br[pp * ld];
            rd[u] = dv[pp * (ld + 1)];
        }
    };
    load(0);
    for (int p0 = 0; p0 < t.k; p0 += KC)
    {
        __syncthreads(); // the previous chunk's products have read the LDS
#pragma unroll
        for (int u = 0; u < PER; ++u)
        {
            const int kk = kq + kstep * u;
            const bool term = p0 + kk < t.k;
            const T av = pinned(ra[u]), bv = pinned(rb[u]), dd = pinned(rd[u]);
            As[kk][row] = term && row < t.m ? -av : T(0);
            Bs[kk][row] = term && row < t.n ? bv * dd : T(0);
        }
        __syncthreads();
        if (p0 + KC < t.k)
            load(p0 + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4)
        {
            const int ka = kk + (lane >> 4);
            T av[NB], bv[NB];
#pragma unroll
            for (int a = 0; a < NB; ++a)
                av[a] = As[ka][wm + a * 16 + (lane & 15)];
#pragma unroll
            for (int b = 0; b < NB; ++b)
                bv[b] = Bs[ka][wn + b * 16 + (lane & 15)];
#pragma unroll
            for (int a = 0; a < NB; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b)
                    acc[a][b] = Mfma<T>::step(av[a], bv[b], acc[a][b]);
        }
    }

    // The quarters: (qi, qj) over 64-row and 64-column halves (one quarter for TS = 64).
    T(*Q)[QLD] = reinterpret_cast<T (*)[QLD]>(lds);
    T* C = updates + t.cOff;
#pragma unroll
    for (int q = 0; q < (TS / 64) * (TS / 64); ++q)
    {
        const int qi = (q / (TS / 64)) * 64, qj = (q % (TS / 64)) * 64;
        const int qm = min(64, t.m - qi), qn = min(64, t.n - qj);
        if (qm <= 0 || qn <= 0 || (t.lower && qi + qm - 1 < qj))
            continue; // uniform: every thread skips the same quarters
        __syncthreads(); // the LDS is free (products or the previous quarter done)
        // The waves' accumulators that fall in this quarter, into Q (transposed: Q[column][row]).
#pragma unroll
        for (int a = 0; a < NB; ++a)
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int i = wm + a * 16 + Mfma<T>::row(lane, v) - qi;
                    const int j = wn + b * 16 + (lane & 15) - qj;
                    if (i >= 0 && i < 64 && j >= 0 && j < 64)
                        Q[j][i] = acc[a][b][v];
                }
        __syncthreads();
        const int rowBase = t.rowBase + qi, colBase = t.colBase + qj;
        const int itemEnd = itemStarts[t.items + q + 1];
        for (int x = itemStarts[t.items + q]; x < itemEnd; ++x)
        {
            // The child's block, columns [q0, q1) × rows [ia, ib) (its lower part), spread over all the threads by
            // consecutive rows; eight entries per thread per pass, every load issued before any is used.
            const AssemblyItem it = items[x];
            const int rc = p.nrow[it.child];
            const int32_t* rel = p.rel + p.rowStart[it.child];
            const T* u = updates + p.updateOff[it.child];
            const int h = it.ib - it.ia, total = h * (it.q1 - it.q0);
            constexpr int UNROLL = 8;
            for (int base = tid; base < total; base += 256 * UNROLL)
            {
                int ci[UNROLL], ri[UNROLL];
                T v[UNROLL];
                bool in[UNROLL];
#pragma unroll
                for (int e = 0; e < UNROLL; ++e)
                {
                    const int idx = base + 256 * e;
                    const int jj = it.q0 + idx / h, i = it.ia + idx % h;
                    in[e] = idx < total && i >= jj;
                    const int jc = min(jj, it.q1 - 1), ic = in[e] ? i : it.ia;
                    ci[e] = rel[jc];
                    ri[e] = rel[ic];
                    v[e] = u[ic + int64_t(jc) * rc];
                }
#pragma unroll
                for (int e = 0; e < UNROLL; ++e)
                    if (in[e])
                        Q[pinned(ci[e]) - colBase][pinned(ri[e]) - rowBase] += pinned(v[e]);
            }
            __syncthreads(); // the next child may add to the same positions
        }
        // The quarter, written once (the lower part only on the diagonal).
        for (int idx = tid; idx < 64 * 64; idx += 256)
        {
            const int i = idx & 63, j = idx >> 6;
            const int ti = qi + i, tj = qj + j;
            if (i < qm && j < qn && !(t.lower && ti < tj))
                C[ti + int64_t(tj) * t.ldc] = Q[j][i];
        }
    }
}

// ---- whole inverses of the wide supernodes' triangles -----------------------------------------------------------
//
// X = L_FF⁻¹ (c × c, leading dimension c) by blocks of 64, right-looking: with T_IJ = −Σ_{K<I} L_IK·X_KJ (I > J), X_IJ =
// L_II⁻¹·T_IJ. Step K (one launch) takes the items (s, I, J), I > K >= J: T_IJ = −L_IK·X_KJ (J = K, X_KK = L_KK⁻¹) or
// T_IJ −= L_IK·X_KJ (J < K), where X_KJ = L_KK⁻¹·T_KJ is formed in the workgroup from T_KJ (final since step K − 1; no
// item of step K writes row block K). The last launch (items (s, I, J), I >= J) sets X_IJ = L_II⁻¹·T_IJ in place, or
// X_II = L_II⁻¹ from the diagonal block's inverse. Blocks other than the last are full (64), so only row block I can
// be narrower (ib). Four waves of 32 × 32 over a 64 × 64 tile, 16 × 16 × 4 MFMA blocks; products of 64 terms staged
// through LDS in chunks of 32.

// The operands of a product of 64 terms for a 64 × 64 tile, held in registers: a thread stages 16 values of A(i, t)
// (row i = tid mod 64, terms kq + 4u of each half) and 16 of B(t, j) (term bk of each half, columns bj + 8u). All of
// an item's operands are loaded before its first product, so its global loads wait once, not once per chunk.
template <typename T> struct TileTerms
{
    T a[16], b[16];
};

// Loads A(i, t) = a(i, t) for i < rowsA and B(t, j) = b(t, j) for t < terms; a and b are given positions clamped into
// range, and the values beyond are zeroed when staged.
template <typename T, typename FA, typename FB>
__device__ inline void loadTileTerms(TileTerms<T>& s, int rowsA, int terms, FA a, FB b)
{
    const int tid = threadIdx.x;
    const int row = tid & (TILE - 1), kq = tid >> 6;
    const int bk = tid & 31, bj = tid >> 5;
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
#pragma unroll
        for (int u = 0; u < 8; ++u)
            s.a[8 * h + u] = a(min(row, rowsA - 1), min(32 * h + kq + 4 * u, terms - 1));
#pragma unroll
        for (int u = 0; u < 8; ++u)
            s.b[8 * h + u] = b(min(32 * h + bk, terms - 1), bj + 8 * u);
    }
}

// acc += A·B over the staged terms, through LDS in two chunks of 32. Four waves of 32 × 32 over the tile, 16 × 16 × 4
// MFMA blocks.
template <typename T>
__device__ inline void multiplyTileTerms(typename Mfma<T>::Acc (&acc)[2][2], T (*As)[TILE + 4], T (*Bs)[TILE + 4],
                                         const TileTerms<T>& s, int rowsA, int terms)
{
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    const int row = tid & (TILE - 1), kq = tid >> 6;
    const int bk = tid & 31, bj = tid >> 5;
#pragma unroll
    for (int h = 0; h < 2; ++h)
    {
        __syncthreads(); // the previous chunk's products have read the LDS
#pragma unroll
        for (int u = 0; u < 8; ++u)
            As[kq + 4 * u][row] = row < rowsA && 32 * h + kq + 4 * u < terms ? pinned(s.a[8 * h + u]) : T(0);
#pragma unroll
        for (int u = 0; u < 8; ++u)
            Bs[bk][bj + 8 * u] = 32 * h + bk < terms ? pinned(s.b[8 * h + u]) : T(0);
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < 32; kk += 4)
        {
            const int ka = kk + (lane >> 4);
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
                    acc[x][y] = Mfma<T>::step(As[ka][wm + x * 16 + (lane & 15)], Bs[ka][wn + y * 16 + (lane & 15)],
                                              acc[x][y]);
        }
    }
}

template <typename T> __device__ inline void zeroTile(typename Mfma<T>::Acc (&acc)[2][2])
{
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
            acc[x][y] = typename Mfma<T>::Acc{0, 0, 0, 0};
}

// The element (i, j) of the tile that acc[x][y][v] holds on this lane.
__device__ inline int tileRow(int x, int v, int lane, int wave, bool fp64)
{
    return (wave >> 1) * 32 + x * 16 + (fp64 ? (lane >> 4) + 4 * v : 4 * (lane >> 4) + v);
}

template <typename T>
__global__ __launch_bounds__(256) void triangleInverseKernel(DevicePlan p, const T* panels, T* linv, const int2* items,
                                                             int step)
{
    __shared__ T As[32][TILE + 4];
    __shared__ T Bs[32][TILE + 4];
    const int2 item = items[blockIdx.x];
    const int s = item.x, I = item.y >> 16, J = item.y & 0xFFFF;
    const int c = p.ncol[s], n = c + p.nrow[s];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    const int wn = (wave & 1) * 32;
    T* X = linv + p.linvOff[s] + blockInverseValues(c);
    const int i0 = TILE * I, j0 = TILE * J, ib = min(TILE, c - i0);
    const bool fp64 = sizeof(T) == 8;
    typename Mfma<T>::Acc acc[2][2];
    if (step < 0)
    {
        // A row block's final values: X_II from the diagonal block's inverse, or X_IJ = L_II⁻¹·T_IJ (T_IJ complete:
        // before step I, or in the last launch).
        const T* inv = linv + p.linvOff[s] + int64_t(i0) * TILE; // ib × ib, leading dimension ib
        if (I == J)
        {
            for (int j = wave; j < ib; j += 4)
                if (lane < ib)
                    X[(i0 + lane) + int64_t(i0 + j) * c] = inv[lane + j * ib];
            return;
        }
        TileTerms<T> terms;
        loadTileTerms<T>(terms, ib, ib, [&](int i, int t) { return inv[i + t * ib]; },
                         [&](int t, int j) { return X[(i0 + t) + int64_t(j0 + j) * c]; });
        zeroTile<T>(acc);
        multiplyTileTerms<T>(acc, As, Bs, terms, ib, ib);
        __syncthreads(); // every wave has read T_IJ before any writes X_IJ over it
    }
    else
    {
        // The terms of steps K0 .. K, in that order (K0 = K − 1 for a paired step, flag bit 16: step K − 1 did only
        // the items of row block K): T_IJ −= L_I,Kt·X_Kt,J for J < Kt (row block Kt is final before step Kt), or
        // T_IKt = −L_I,Kt·L_KtKt⁻¹ for J = Kt. The accumulators start from T_IJ (J < K0) or zero, L enters negated.
        const int K = step & 0xFFFF, K0 = ((step >> 16) & 1) ? K - 1 : K;
        const T* L = panels + p.panelOff[s];
        if (J < K0)
        {
#pragma unroll
            for (int x = 0; x < 2; ++x)
#pragma unroll
                for (int y = 0; y < 2; ++y)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        acc[x][y][v] = X[(i0 + min(tileRow(x, v, lane, wave, fp64), ib - 1)) +
                                         int64_t(j0 + wn + y * 16 + (lane & 15)) * c];
        }
        else
            zeroTile<T>(acc);
        // The operands of both steps are loaded before the first product (step Kt has a term when J <= Kt; J <= K).
        TileTerms<T> terms[2];
#pragma unroll
        for (int e = 0; e < 2; ++e)
        {
            const int Kt = K0 + e;
            if (Kt > K || J > Kt)
                continue; // the same for the whole workgroup
            const int kt0 = TILE * Kt;
            const T* invKt = linv + p.linvOff[s] + int64_t(kt0) * TILE; // 64 × 64 (block Kt is full)
            if (J < Kt)
                loadTileTerms<T>(terms[e], ib, TILE, [&](int i, int t) { return -L[(i0 + i) + int64_t(kt0 + t) * n]; },
                                 [&](int t, int j) { return X[(kt0 + t) + int64_t(j0 + j) * c]; });
            else
                loadTileTerms<T>(terms[e], ib, TILE, [&](int i, int t) { return -L[(i0 + i) + int64_t(kt0 + t) * n]; },
                                 [&](int t, int j) { return invKt[t + j * TILE]; });
        }
#pragma unroll
        for (int e = 0; e < 2; ++e)
            if (K0 + e <= K && J <= K0 + e)
                multiplyTileTerms<T>(acc, As, Bs, terms[e], ib, TILE);
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int i = tileRow(x, v, lane, wave, fp64), j = wn + y * 16 + (lane & 15);
                if (i < ib)
                    X[(i0 + i) + int64_t(j0 + j) * c] = acc[x][y][v];
            }
}

__global__ void mfmaProbeKernel(const double* a, const double* b, double* d, const float* af, const float* bf,
                                float* df)
{
    const int lane = threadIdx.x;
    double4_t acc = {0, 0, 0, 0};
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[(lane & 15) * 4 + (lane >> 4)], b[(lane >> 4) * 16 + (lane & 15)],
                                               acc, 0, 0, 0);
    float4_t accf = {0, 0, 0, 0};
    accf = __builtin_amdgcn_mfma_f32_16x16x4f32(af[(lane & 15) * 4 + (lane >> 4)], bf[(lane >> 4) * 16 + (lane & 15)],
                                                accf, 0, 0, 0);
    for (int v = 0; v < 4; ++v)
    {
        d[Mfma<double>::row(lane, v) * 16 + (lane & 15)] = acc[v];
        df[Mfma<float>::row(lane, v) * 16 + (lane & 15)] = accf[v];
    }
}

// ---- solves --------------------------------------------------------------------------------------------------------
//
// Multifrontal sweeps, level by level, in double precision (panels of either precision). Per supernode s, with
// y = [X(cols); update vector]:
//   forward   gather the children's update vectors (child order); by blocks J of 64 columns: y_J ← L_JJ⁻¹·y_J, then
//             the rows below the block −= L(rows, J)·y_J (MFMA block products); X(cols) keeps L⁻¹y, D⁻¹ is applied
//             to all of X after the last level;
//   backward  X(cols) −= L_RFᵀ·X(rows), then by blocks J from the last: y_J −= L(rows after J, J)ᵀ·X(those rows),
//             X_J ← L_JJ⁻ᵀ·y_J.
// The products with L_RF (the rows below a supernode's columns) are spread over workgroups for every supernode: row
// tiles of 64 in the forward sweep (after the level's triangles), column blocks of 64 in the backward sweep (before
// them); each workgroup owns its outputs, so there is no reduction across workgroups. A workgroup per supernode (or,
// for the wide ones, per step and row tile) does the triangle L_FF.

constexpr int SB = 64;       // rows or columns of a staged panel block
constexpr int SBP = SB + 1;  // its padded leading dimension in LDS
constexpr int KMAX = 16;     // right-hand sides per sweep

// A block blk[q·SBP + i] = L(row0 + i, col0 + q) in the panels' precision, zero outside [0, rows) × [0, cols)
// (1 <= rows, cols <= SB), staged through registers: load() issues the global loads (thread t: row t mod 64 of the
// columns t / 64 + 4u, positions clamped into the block), store() writes LDS; a workgroup of 256 threads. Between
// the two, the loads are in flight while the workgroup computes.
template <typename TP> struct StagedBlock
{
    TP v[SB / 4];
    int rows, cols;
    __device__ inline void load(const TP* L, int64_t ld, int row0, int rows_, int col0, int cols_)
    {
        rows = rows_;
        cols = cols_;
        const int i = threadIdx.x & (SB - 1), q0 = threadIdx.x >> 6;
        const TP* src = L + (row0 + min(i, rows - 1)) + int64_t(col0) * ld;
#pragma unroll
        for (int u = 0; u < SB / 4; ++u)
            v[u] = src[int64_t(min(q0 + 4 * u, cols - 1)) * ld];
    }
    __device__ inline void store(TP* blk) const
    {
        const int i = threadIdx.x & (SB - 1), q0 = threadIdx.x >> 6;
#pragma unroll
        for (int u = 0; u < SB / 4; ++u)
        {
            const int q = q0 + 4 * u;
            blk[q * SBP + i] = (i < rows && q < cols) ? pinned(v[u]) : TP(0);
        }
    }
};

template <typename TP>
__device__ inline void stageBlock(TP* blk, const TP* L, int64_t ld, int row0, int rows, int col0, int cols)
{
    StagedBlock<TP> b;
    b.load(L, ld, row0, rows, col0, cols);
    b.store(blk);
}

// Wave w: acc(i, h) += Σ_q blk(i, q)·v(q, h) for its 16 rows i = 16w + ..., v[q·KMAX + h] (in FP64).
template <typename TP> __device__ inline double4_t rowsTimesVectors(const TP* blk, const double* v, double4_t acc)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll 4
    for (int kk = 0; kk < SB; kk += 4)
    {
        const double a = double(blk[(kk + (lane >> 4)) * SBP + 16 * w + (lane & 15)]);
        const double b = v[(kk + (lane >> 4)) * KMAX + (lane & 15)];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Wave w: acc(q, h) += Σ_i blk(i, q)·v(i, h) for its 16 columns q = 16w + ..., v[i·KMAX + h] (in FP64).
template <typename TP> __device__ inline double4_t columnsTimesVectors(const TP* blk, const double* v, double4_t acc)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll 4
    for (int kk = 0; kk < SB; kk += 4)
    {
        const double a = double(blk[(16 * w + (lane & 15)) * SBP + kk + (lane >> 4)]);
        const double b = v[(kk + (lane >> 4)) * KMAX + (lane & 15)];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// The output row (within a 64-row tile) of acc[v] on this lane, and its column h.
__device__ inline int accRow(int v)
{
    return 16 * (threadIdx.x >> 6) + ((threadIdx.x & 63) >> 4) + 4 * v;
}
__device__ inline int accColumn()
{
    return threadIdx.x & 15;
}

// Wide supernodes solve their triangles by the whole inverse (triangleProductsKernel).
__device__ inline bool isWideSupernode(int c)
{
    return keepsTriangleInverse(c);
}

// ys ← inv·ys (transpose: invᵀ·ys) for the unit lower inverse of a diagonal block (column-major qb × qb), staged in
// blk and applied by MFMA; ys is SB × KMAX, zero beyond its rows and columns.
template <typename TP>
__device__ inline void applyBlockInverse(const TP* inv, int qb, TP* blk, double* ys, bool transpose)
{
    stageBlock(blk, inv, qb, 0, qb, 0, qb);
    __syncthreads();
    const double4_t zero = {0, 0, 0, 0};
    const double4_t acc = transpose ? columnsTimesVectors(blk, ys, zero) : rowsTimesVectors(blk, ys, zero);
    __syncthreads();
    const int h = accColumn();
#pragma unroll
    for (int v = 0; v < 4; ++v)
        ys[accRow(v) * KMAX + h] = acc[v];
    __syncthreads();
}

// SB rows of k right-hand-side values (X row-major, n × k) for ys (SB × KMAX, zero padded), staged through registers:
// the rows row0 .. row0 + qb (load), or rows gathered by the row list of a supernode (gather: rows[i0 − c + i], or
// first + i0 + i below c); positions clamped into the rows (1 <= qb <= SB).
struct StagedRows
{
    static constexpr int PER = SB * KMAX / 256;
    double v[PER];
    int count, k;
    __device__ inline void load(const double* X, int64_t row0, int qb, int k_)
    {
        count = qb;
        k = k_;
#pragma unroll
        for (int u = 0; u < PER; ++u)
        {
            const int idx = threadIdx.x + 256 * u;
            v[u] = X[(row0 + min(idx >> 4, qb - 1)) * k + min(idx & (KMAX - 1), k - 1)];
        }
    }
    __device__ inline void gather(const double* X, const int32_t* rows, int first, int c, int i0, int ib, int k_)
    {
        count = ib;
        k = k_;
        int64_t row[PER];
#pragma unroll
        for (int u = 0; u < PER; ++u)
        {
            const int idx = threadIdx.x + 256 * u;
            const int at = i0 + min(idx >> 4, ib - 1);
            row[u] = at < c ? int64_t(first + at) : int64_t(rows[max(at - c, 0)]);
        }
#pragma unroll
        for (int u = 0; u < PER; ++u)
        {
            const int idx = threadIdx.x + 256 * u;
            v[u] = X[row[u] * k + min(idx & (KMAX - 1), k - 1)];
        }
    }
    __device__ inline void store(double* ys) const
    {
#pragma unroll
        for (int u = 0; u < PER; ++u)
        {
            const int idx = threadIdx.x + 256 * u;
            const int i = idx >> 4, h = idx & (KMAX - 1);
            ys[idx] = (i < count && h < k) ? pinned(v[u]) : 0.0;
        }
    }
};

__device__ inline void loadBlockRows(double* ys, const double* X, int64_t row0, int qb, int k)
{
    StagedRows r;
    r.load(X, row0, qb, k);
    r.store(ys);
}

// The rows row0 .. row0 + qb of X ← ys.
__device__ inline void storeBlockRows(const double* ys, double* X, int64_t row0, int qb, int k)
{
    for (int idx = threadIdx.x; idx < qb * KMAX; idx += blockDim.x)
    {
        const int i = idx >> 4, h = idx & (KMAX - 1);
        if (h < k)
            X[(row0 + i) * k + h] = ys[idx];
    }
}

template <typename TP>
__global__ __launch_bounds__(256) void forwardKernel(DevicePlan p, const TP* panels, const TP* linv, double* X,
                                                     double* U, int k, const int32_t* list, int gathered)
{
    __shared__ TP blk[SB * SBP];
    __shared__ double ys[SB * KMAX];
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    if (gathered && isWideSupernode(c))
        return; // its gather is done (gatherKernel), its triangle is a product with the whole inverse
    const TP* L = panels + p.panelOff[s];
    double* u = U + p.rowStart[s] * k;
    const int tid = threadIdx.x, nt = blockDim.x;
    if (!gathered)
        for (int idx = tid; idx < r * k; idx += nt)
            u[idx] = 0.0;
    for (int q = p.childStart[s]; q < p.childStart[s + 1] && !gathered; ++q)
    {
        __syncthreads();
        const int ch = p.children[q];
        const int rc = p.nrow[ch];
        const int32_t* rel = p.rel + p.rowStart[ch];
        const double* uc = U + p.rowStart[ch] * k;
        // Eight elements (i, h) of the child's rc × k values per thread per pass (distinct targets: a child's rows
        // map to distinct rows of s): the child's values and positions are loaded first (from elements clamped into
        // its values), then the eight targets, then the eight sums are stored.
        constexpr int UNROLL = 8;
        const int total = rc * k;
        for (int base = tid; base < total; base += UNROLL * nt)
        {
            int pos[UNROLL], col[UNROLL];
            double v[UNROLL];
#pragma unroll
            for (int t = 0; t < UNROLL; ++t)
            {
                const int idx = min(base + t * nt, total - 1);
                const int i = idx / k;
                col[t] = idx - i * k;
                pos[t] = rel[i];
                v[t] = uc[idx];
            }
            double* y[UNROLL];
            double old[UNROLL];
#pragma unroll
            for (int t = 0; t < UNROLL; ++t)
            {
                const int at = pinned(pos[t]);
                y[t] = at < c ? X + int64_t(first + at) * k + col[t] : u + int64_t(at - c) * k + col[t];
                old[t] = *y[t];
            }
#pragma unroll
            for (int t = 0; t < UNROLL; ++t)
                if (base + t * nt < total)
                    *y[t] = pinned(old[t]) + pinned(v[t]);
        }
    }
    __syncthreads();
    if (isWideSupernode(c))
        return; // its triangle is a product with the whole inverse (triangleProductsKernel)
    // The triangle L_FF, and the rows below (L_RF) of a small supernode; a larger one's are the row tiles of
    // forwardRowsKernel.
    const int rowsEnd = sweepRowsInline(c, r) ? n : c;
    for (int q0 = 0; q0 < c; q0 += SB)
    {
        const int qb = min(SB, c - q0);
        // The block's rows of X, its inverse and the first rows below it, loaded together.
        StagedRows own;
        own.load(X, first + q0, qb, k);
        StagedBlock<TP> inverse, below;
        inverse.load(linv + p.linvOff[s] + int64_t(q0) * SB, qb, 0, qb, 0, qb);
        if (q0 + qb < rowsEnd)
            below.load(L, n, q0 + qb, min(SB, rowsEnd - q0 - qb), q0, qb);
        own.store(ys);
        inverse.store(blk);
        __syncthreads();
        // y_J ← L_JJ⁻¹·y_J.
        {
            const double4_t out = rowsTimesVectors(blk, ys, double4_t{0, 0, 0, 0});
            __syncthreads();
            const int h = accColumn();
#pragma unroll
            for (int v = 0; v < 4; ++v)
                ys[accRow(v) * KMAX + h] = out[v];
            __syncthreads();
        }
        storeBlockRows(ys, X, first + q0, qb, k);
        // The rows below the block, 64 at a time, the next 64 loaded into registers while these are multiplied.
        for (int i0 = q0 + qb; i0 < rowsEnd; i0 += SB)
        {
            below.store(blk);
            __syncthreads();
            if (i0 + SB < rowsEnd)
                below.load(L, n, i0 + SB, min(SB, rowsEnd - i0 - SB), q0, qb);
            const double4_t acc = rowsTimesVectors(blk, ys, double4_t{0, 0, 0, 0});
            const int h = accColumn();
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int i = i0 + accRow(v);
                if (i < rowsEnd && h < k)
                {
                    double* y = i < c ? X + int64_t(first + i) * k : u + int64_t(i - c) * k;
                    y[h] -= acc[v];
                }
            }
            __syncthreads();
        }
        __syncthreads();
    }
}

// ---- the sweeps of tiny supernodes: a wave each ---------------------------------------------------------------------
//
// Supernodes with at most TINY_COLUMNS columns and TINY_ROWS rows below them (the many small supernodes of the
// lowest levels) are swept by one wave each, TINY_WAVES waves a workgroup, with FP64 MFMA products (16 × 16 × 4) on
// values read straight from HBM into registers and no LDS, so that occupancy is bounded by registers alone. Lane
// l = 16·lq + lr holds a 16-row tile of a block of vectors as the MFMA result layout: its register v holds row lq + 4v,
// column (right-hand side) lr. That is also the B-operand layout of k-step v of a product whose inner dimension runs
// over those 16 rows, so one product's result is the next one's operand in place.
//
// The children's update vectors reach a front by the pull table: position q of the front of s (a column for q < c,
// else row q − c below) is hit by the update rows pulls[pullBase[s] + q] (at most TINY_PULLS, in child order, -1 past
// the last); they are added to the front in that order, as the assembly of the factorisation does.
constexpr int TINY_WAVES = 4;

// The update rows that land on position q of the front of s (base = pullBase[s] >= 0).
__device__ inline int4 pullsOf(const DevicePlan& p, int base, int q)
{
    return p.pulls[base + q];
}

// y + the values of column h of the update rows e (in order; -1: none; row 0 stands in for none as an address). The
// loads are issued together, the sums formed after them.
__device__ inline double addPulled(const DevicePlan& p, int4 e, double y, const double* U, int k, int h)
{
    const double u0 = U[int64_t(max(e.x, 0)) * k + h];
    const double u1 = p.pullWidth > 1 ? U[int64_t(max(e.y, 0)) * k + h] : 0.0;
    const double u2 = p.pullWidth > 2 ? U[int64_t(max(e.z, 0)) * k + h] : 0.0;
    const double u3 = p.pullWidth > 3 ? U[int64_t(max(e.w, 0)) * k + h] : 0.0;
    y = e.x >= 0 ? y + u0 : y;
    y = e.y >= 0 ? y + u1 : y;
    y = e.z >= 0 ? y + u2 : y;
    y = e.w >= 0 ? y + u3 : y;
    return y;
}

// Forward: y_F = X(cols) + pulls, y_F ← L_FF⁻¹·y_F (the inverse of the unit lower triangle, c × c column-major),
// X(cols) = y_F, U(s) = pulls(rows) − L_RF·y_F.
template <typename TP>
__global__ __launch_bounds__(64 * TINY_WAVES) void tinyForwardKernel(DevicePlan p, const TP* panels, const TP* linv,
                                                                     double* X, double* U, int k,
                                                                     const int32_t* list, int count)
{
    const int at = blockIdx.x * TINY_WAVES + (threadIdx.x >> 6);
    if (at >= count)
        return; // the whole wave: no barrier follows
    const int s = list[at];
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s], base = p.pullBase[s];
    const int ct = (c + 15) >> 4, rt = (r + 15) >> 4;
    const bool column = lr < k;
    double4_t y[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int i = 16 * t + lq + 4 * v;
            const bool valid = t < ct && i < c && column;
            const double x = X[valid ? int64_t(first + i) * k + lr : 0];
            y[t][v] = valid ? x : 0.0;
        }
    if (base >= 0)
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (t < ct)
            {
                int4 e[4];
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    e[v] = pullsOf(p, base, min(16 * t + lq + 4 * v, c - 1));
#pragma unroll
                for (int v = 0; v < 4; ++v)
                {
                    const int i = 16 * t + lq + 4 * v;
                    if (i < c && column)
                        y[t][v] = addPulled(p, e[v], y[t][v], U, k, lr);
                }
            }
    // y_F ← L_FF⁻¹·y_F, tiles from the last: tile t reads tiles 0..t only.
    const TP* inv = linv + p.linvOff[s];
#pragma unroll
    for (int t = 3; t >= 0; --t)
        if (t < ct)
        {
            const int i = 16 * t + lr;
            double a[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
            {
                const int j = 4 * q + lq; // column 16·kt + 4m + lq, kt = q / 4, m = q % 4
                const bool valid = (q >> 2) <= t && i < c && j <= i;
                const double value = double(inv[valid ? i + j * c : 0]);
                a[q] = valid ? value : 0.0;
            }
            double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if ((q >> 2) <= t)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], y[q >> 2][q & 3], acc, 0, 0, 0);
            y[t] = acc;
        }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int i = 16 * t + lq + 4 * v;
            if (t < ct && i < c && column)
                X[int64_t(first + i) * k + lr] = y[t][v];
        }
    // U(s) = pulls(rows) − L_RF·y_F, a 16-row tile at a time; the pull entries of the next tile are loaded with this
    // tile's values.
    const TP* L = panels + p.panelOff[s];
    double* u = U + p.rowStart[s] * k;
    int4 e[4];
    const int4 none = {-1, -1, -1, -1};
#pragma unroll
    for (int v = 0; v < 4; ++v)
        e[v] = base >= 0 && r > 0 ? pullsOf(p, base, c + min(lq + 4 * v, r - 1)) : none;
    for (int t = 0; t < rt; ++t)
    {
        const int rho = 16 * t + lr; // the A operand's row
        TP a[16];
#pragma unroll
        for (int q = 0; q < 16; ++q)
        {
            const int j = 4 * q + lq;
            const bool valid = (q >> 2) < ct && rho < r && j < c;
            const TP value = L[valid ? (c + rho) + int64_t(j) * n : 0];
            a[q] = valid ? value : TP(0);
        }
        double4_t acc = {0.0, 0.0, 0.0, 0.0};
        int4 next[4];
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int i = 16 * t + lq + 4 * v;
            if (base >= 0 && i < r && column)
                acc[v] = addPulled(p, e[v], 0.0, U, k, lr);
            next[v] = base >= 0 && t + 1 < rt ? pullsOf(p, base, c + min(i + 16, r - 1)) : none;
        }
#pragma unroll
        for (int q = 0; q < 16; ++q)
            if ((q >> 2) < ct)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(-double(a[q]), y[q >> 2][q & 3], acc, 0, 0, 0);
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int i = 16 * t + lq + 4 * v;
            if (i < r && column)
                u[int64_t(i) * k + lr] = acc[v];
            e[v] = next[v];
        }
    }
}

// Backward: x_F = X(cols) − L_RFᵀ·X(rows), then X(cols) = L_FF⁻ᵀ·x_F.
template <typename TP>
__global__ __launch_bounds__(64 * TINY_WAVES) void tinyBackwardKernel(DevicePlan p, const TP* panels, const TP* linv,
                                                                      double* X, int k, const int32_t* list, int count)
{
    const int at = blockIdx.x * TINY_WAVES + (threadIdx.x >> 6);
    if (at >= count)
        return; // the whole wave: no barrier follows
    const int s = list[at];
    const int lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    const int ct = (c + 15) >> 4;
    const bool column = lr < k;
    double4_t x[4];
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int i = 16 * t + lq + 4 * v;
            const bool valid = t < ct && i < c && column;
            const double value = X[valid ? int64_t(first + i) * k + lr : 0];
            x[t][v] = valid ? value : 0.0;
        }
    // x_F −= L_RFᵀ·X(rows): k-steps of 4 rows below, 8 at a time. The row indices (r <= 128) are loaded once, two
    // per lane (rows lane and 64 + lane), and handed to the lanes that gather by shuffles; each batch's X rows are
    // gathered while the previous batch is multiplied.
    const TP* L = panels + p.panelOff[s];
    const int32_t* rows = p.rows + p.rowStart[s];
    const int32_t index0 = r > 0 ? rows[min(lane, r - 1)] : 0;
    const int32_t index1 = r > 0 ? rows[min(64 + lane, r - 1)] : 0;
    // Batch r0's X rows: row r0 + 4m + lq (r0 + 4m < 64 for every lane of a batch or for none).
    const auto gatherBatch = [&](int r0, double (&b)[8])
    {
#pragma unroll
        for (int m = 0; m < 8; ++m)
        {
            const int rho = r0 + 4 * m + lq;
            const bool valid = rho < r && column;
            const int32_t row = __shfl(r0 + 4 * m < 64 ? index0 : index1, rho & 63);
            const double value = X[valid ? int64_t(row) * k + lr : 0];
            b[m] = valid ? value : 0.0;
        }
    };
    double next[8];
    gatherBatch(0, next);
    for (int r0 = 0; r0 < r; r0 += 32)
    {
        double b[8];
#pragma unroll
        for (int m = 0; m < 8; ++m)
            b[m] = next[m];
        if (r0 + 32 < r)
            gatherBatch(r0 + 32, next);
#pragma unroll
        for (int t = 0; t < 4; ++t)
            if (t < ct)
            {
                const int j = 16 * t + lr; // the A operand's row: column j of L
                double a[8];
#pragma unroll
                for (int m = 0; m < 8; ++m)
                {
                    const int rho = r0 + 4 * m + lq;
                    const bool valid = rho < r && j < c;
                    const double value = double(L[valid ? (c + rho) + int64_t(j) * n : 0]);
                    a[m] = valid ? -value : 0.0;
                }
#pragma unroll
                for (int m = 0; m < 8; ++m)
                    x[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[m], x[t], 0, 0, 0);
            }
    }
    // X(cols) = L_FF⁻ᵀ·x_F, tiles from the first: tile t reads tiles t..3 only.
    const TP* inv = linv + p.linvOff[s];
#pragma unroll
    for (int t = 0; t < 4; ++t)
        if (t < ct)
        {
            const int i = 16 * t + lr; // row i of the transposed inverse: column i of the inverse
            double a[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
            {
                const int j = 4 * q + lq;
                const bool valid = (q >> 2) >= t && (q >> 2) < ct && j < c && j >= i;
                const double value = double(inv[valid ? j + i * c : 0]);
                a[q] = valid ? value : 0.0;
            }
            double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if ((q >> 2) >= t && (q >> 2) < ct)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], x[q >> 2][q & 3], acc, 0, 0, 0);
            x[t] = acc;
        }
#pragma unroll
    for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int i = 16 * t + lq + 4 * v;
            if (t < ct && i < c && column)
                X[int64_t(first + i) * k + lr] = x[t][v];
        }
}

// ---- the sweeps of medium supernodes: a workgroup each ---------------------------------------------------------------
//
// Supernodes of at most MEDIUM_COLUMNS columns that are not tiny are swept by one workgroup of four waves each, with
// FP64 MFMA products: the right-hand sides of the front's columns in LDS (MEDIUM_COLUMNS × KMAX doubles, zero beyond
// its rows and columns), the triangle by the inverses of its diagonal blocks of PANEL_BLOCK columns (two at most), a
// wave per 16-row tile of a block; the rows below in 16-row tiles spread over the waves (forward), the columns' tiles
// two per wave over every row below (backward). Their gathers are gatherKernel's, before them.

// Wave-level: a 16-row tile (rows 16·tile within the block) of Inv·Y (TRANSPOSE: Invᵀ·Y) for the unit lower inverse of
// the diagonal block of qb columns from row q0 (qb × qb column-major at inv), Y the block's rows of ys; the result in
// the MFMA result layout.
template <typename TP, bool TRANSPOSE>
__device__ inline double4_t blockInverseTile(const TP* inv, int qb, int q0, const double* ys, int tile, int lr, int lq)
{
    const int i = 16 * tile + lr;
    double a[16];
#pragma unroll
    for (int q = 0; q < 16; ++q)
    {
        const int j = 4 * q + lq;
        const bool valid = i < qb && j < qb && (TRANSPOSE ? j >= i : j <= i);
        const double value = double(inv[valid ? (TRANSPOSE ? j + i * qb : i + j * qb) : 0]);
        a[q] = valid ? value : 0.0;
    }
    double4_t acc = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 16; ++q)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], ys[(q0 + 4 * q + lq) * KMAX + lr], acc, 0, 0, 0);
    return acc;
}

// Stores a tile in the result layout to the rows row0 + lq + 4v of ys.
__device__ inline void storeTile(double* ys, int row0, double4_t acc, int lr, int lq)
{
#pragma unroll
    for (int v = 0; v < 4; ++v)
        ys[(row0 + lq + 4 * v) * KMAX + lr] = acc[v];
}

__device__ inline double4_t loadTile(const double* ys, int row0, int lr, int lq)
{
    double4_t t;
#pragma unroll
    for (int v = 0; v < 4; ++v)
        t[v] = ys[(row0 + lq + 4 * v) * KMAX + lr];
    return t;
}

// Forward: y_F = X(cols) (gathered), y_F ← L_FF⁻¹·y_F block by block, X(cols) = y_F, U(s) −= L_RF·y_F.
template <typename TP>
__global__ __launch_bounds__(256) void mediumForwardKernel(DevicePlan p, const TP* panels, const TP* linv, double* X,
                                                           double* U, int k, const int32_t* list)
{
    __shared__ double ys[MEDIUM_COLUMNS * KMAX];
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    const int ct = (c + 15) >> 4;
    for (int idx = tid; idx < MEDIUM_COLUMNS * KMAX; idx += 256)
    {
        const int i = idx >> 4, h = idx & (KMAX - 1);
        const bool valid = i < c && h < k;
        const double x = X[valid ? int64_t(first + i) * k + h : 0];
        ys[idx] = valid ? x : 0.0;
    }
    __syncthreads();
    const TP* L = panels + p.panelOff[s];
    const TP* inv = linv + p.linvOff[s];
    const int qb0 = min(c, PANEL_BLOCK);
    double4_t acc = {0.0, 0.0, 0.0, 0.0};
    if (16 * w < qb0)
        acc = blockInverseTile<TP, false>(inv, qb0, 0, ys, w, lr, lq);
    __syncthreads();
    if (16 * w < qb0)
        storeTile(ys, 16 * w, acc, lr, lq);
    __syncthreads();
    if (c > PANEL_BLOCK)
    {
        const int qb1 = c - PANEL_BLOCK;
        // y_1 −= L_10·z_0.
        if (16 * w < qb1)
        {
            const int i = 16 * w + lr;
            acc = loadTile(ys, PANEL_BLOCK + 16 * w, lr, lq);
            double a[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
            {
                const int j = 4 * q + lq;
                const bool valid = i < qb1;
                const double value = double(L[valid ? (PANEL_BLOCK + i) + int64_t(j) * n : 0]);
                a[q] = valid ? -value : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q)
                acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], ys[(4 * q + lq) * KMAX + lr], acc, 0, 0, 0);
        }
        __syncthreads();
        if (16 * w < qb1)
            storeTile(ys, PANEL_BLOCK + 16 * w, acc, lr, lq);
        __syncthreads();
        if (16 * w < qb1)
            acc = blockInverseTile<TP, false>(inv + PANEL_BLOCK * PANEL_BLOCK, qb1, PANEL_BLOCK, ys, w, lr, lq);
        __syncthreads();
        if (16 * w < qb1)
            storeTile(ys, PANEL_BLOCK + 16 * w, acc, lr, lq);
        __syncthreads();
    }
    for (int idx = tid; idx < c * KMAX; idx += 256)
    {
        const int i = idx >> 4, h = idx & (KMAX - 1);
        if (h < k)
            X[int64_t(first + i) * k + h] = ys[idx];
    }
    // U(s) −= L_RF·y_F, 16-row tiles over the waves, the columns' k-steps 16 at a time.
    double* u = U + p.rowStart[s] * k;
    const int rt = (r + 15) >> 4, steps = 4 * ct;
    for (int t = w; t < rt; t += 4)
    {
        const int rho = 16 * t + lr; // the A operand's row
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int i = 16 * t + lq + 4 * v;
            const bool valid = i < r && lr < k;
            const double value = u[valid ? int64_t(i) * k + lr : 0];
            acc[v] = valid ? value : 0.0;
        }
        for (int q0 = 0; q0 < steps; q0 += 16)
        {
            double a[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
            {
                const int j = 4 * (q0 + q) + lq;
                const bool valid = rho < r && j < c;
                const double value = double(L[valid ? (c + rho) + int64_t(j) * n : 0]);
                a[q] = valid ? -value : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q)
                if (q0 + q < steps)
                    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], ys[(4 * (q0 + q) + lq) * KMAX + lr], acc, 0, 0,
                                                               0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int i = 16 * t + lq + 4 * v;
            if (i < r && lr < k)
                u[int64_t(i) * k + lr] = acc[v];
        }
    }
}

// Backward: x_F = X(cols) − L_RFᵀ·X(rows) (wave w: the columns' tiles w and w + 4 over every row below), then
// X(cols) = L_FF⁻ᵀ·x_F block by block from the last.
template <typename TP>
__global__ __launch_bounds__(256) void mediumBackwardKernel(DevicePlan p, const TP* panels, const TP* linv, double* X,
                                                            int k, const int32_t* list)
{
    __shared__ double xs[MEDIUM_COLUMNS * KMAX];
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63, lr = lane & 15, lq = lane >> 4;
    const int ct = (c + 15) >> 4;
    const TP* L = panels + p.panelOff[s];
    const int32_t* rows = p.rows + p.rowStart[s];
    // The first MEDIUM_ROW_INDICES row indices through LDS (one coalesced load), the rest read where used.
    __shared__ int32_t rowIndex[MEDIUM_ROW_INDICES];
    for (int i = tid; i < min(r, MEDIUM_ROW_INDICES); i += 256)
        rowIndex[i] = rows[i];
    __syncthreads();
    // x_F −= L_RFᵀ·X(rows): wave w takes the batches of 32 rows w, w + 4, ... over every column tile, its partial sums
    // in registers; each batch's X rows (rows r0 + 4m + lq) are gathered while the previous batch is multiplied.
    constexpr int TILES = MEDIUM_COLUMNS / 16;
    double4_t part[TILES];
#pragma unroll
    for (int t = 0; t < TILES; ++t)
        part[t] = double4_t{0.0, 0.0, 0.0, 0.0};
    const auto gatherBatch = [&](int r0, double (&b)[8])
    {
#pragma unroll
        for (int m = 0; m < 8; ++m)
        {
            const int rho = r0 + 4 * m + lq;
            const bool valid = rho < r && lr < k;
            const int32_t row = rho < MEDIUM_ROW_INDICES ? rowIndex[rho < r ? rho : 0] : rows[rho < r ? rho : 0];
            const double value = X[valid ? int64_t(row) * k + lr : 0];
            b[m] = valid ? value : 0.0;
        }
    };
    double next[8];
    gatherBatch(32 * w, next);
    for (int r0 = 32 * w; r0 < r; r0 += 128)
    {
        double b[8];
#pragma unroll
        for (int m = 0; m < 8; ++m)
            b[m] = next[m];
        if (r0 + 128 < r)
            gatherBatch(r0 + 128, next);
#pragma unroll
        for (int t = 0; t < TILES; ++t)
            if (t < ct)
            {
                const int j = 16 * t + lr; // the A operand's row: column j of L
                double a[8];
#pragma unroll
                for (int m = 0; m < 8; ++m)
                {
                    const int rho = r0 + 4 * m + lq;
                    const bool valid = rho < r && j < c;
                    const double value = double(L[valid ? (c + rho) + int64_t(j) * n : 0]);
                    a[m] = valid ? -value : 0.0;
                }
#pragma unroll
                for (int m = 0; m < 8; ++m)
                    part[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[m], b[m], part[t], 0, 0, 0);
            }
    }
    // x_F = X(cols) + the waves' sums, added in wave order through LDS (every tile written by wave 0: zeros beyond
    // the columns).
    for (int q = 0; q < 4; ++q)
    {
        if (w == q)
#pragma unroll
            for (int t = 0; t < TILES; ++t)
            {
                double4_t base;
                if (q == 0)
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                    {
                        const int i = 16 * t + lq + 4 * v;
                        const bool valid = i < c && lr < k;
                        const double value = X[valid ? int64_t(first + i) * k + lr : 0];
                        base[v] = valid ? value : 0.0;
                    }
                else if (t < ct)
                    base = loadTile(xs, 16 * t, lr, lq);
                if (q == 0 || t < ct)
                {
#pragma unroll
                    for (int v = 0; v < 4; ++v)
                        base[v] += part[t][v];
                    storeTile(xs, 16 * t, base, lr, lq);
                }
            }
        __syncthreads();
    }
    const TP* inv = linv + p.linvOff[s];
    const int qb0 = min(c, PANEL_BLOCK);
    double4_t t0 = {0.0, 0.0, 0.0, 0.0};
    if (c > PANEL_BLOCK)
    {
        const int qb1 = c - PANEL_BLOCK;
        if (16 * w < qb1)
            t0 = blockInverseTile<TP, true>(inv + PANEL_BLOCK * PANEL_BLOCK, qb1, PANEL_BLOCK, xs, w, lr, lq);
        __syncthreads();
        if (16 * w < qb1)
            storeTile(xs, PANEL_BLOCK + 16 * w, t0, lr, lq);
        __syncthreads();
        // x_0 −= L_10ᵀ·z_1.
        {
            const int i = 16 * w + lr; // row i of L_10ᵀ: column i of L
            t0 = loadTile(xs, 16 * w, lr, lq);
            double a[16];
#pragma unroll
            for (int q = 0; q < 16; ++q)
            {
                const int j = 4 * q + lq; // row PANEL_BLOCK + j of L
                const bool valid = j < qb1;
                const double value = double(L[valid ? (PANEL_BLOCK + j) + int64_t(i) * n : 0]);
                a[q] = valid ? -value : 0.0;
            }
#pragma unroll
            for (int q = 0; q < 16; ++q)
                t0 = __builtin_amdgcn_mfma_f64_16x16x4f64(a[q], xs[(PANEL_BLOCK + 4 * q + lq) * KMAX + lr], t0, 0, 0,
                                                          0);
        }
        __syncthreads();
        storeTile(xs, 16 * w, t0, lr, lq);
        __syncthreads();
    }
    if (16 * w < qb0)
        t0 = blockInverseTile<TP, true>(inv, qb0, 0, xs, w, lr, lq);
    __syncthreads();
    if (16 * w < qb0)
        storeTile(xs, 16 * w, t0, lr, lq);
    __syncthreads();
    for (int idx = tid; idx < c * KMAX; idx += 256)
    {
        const int i = idx >> 4, h = idx & (KMAX - 1);
        if (h < k)
            X[int64_t(first + i) * k + h] = xs[idx];
    }
}

// The update vectors of the big supernodes of a level: U(s)(rows of the tile) −= L_RF(tile, :)·X(cols(s)).
template <typename TP>
__global__ __launch_bounds__(256) void forwardRowsKernel(DevicePlan p, const TP* panels, const double* X, double* U,
                                                         int k, const int4* tiles, double* partials)
{
    __shared__ TP blk[SB * SBP];
    __shared__ double ys[SB * KMAX];
    const int4 tile = tiles[blockIdx.x];
    const int s = tile.x;
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    const int i0 = c + tile.y * SB;
    const int rows = min(SB, n - i0);
    // The columns of this chunk (all of them without chunks).
    const int qBegin = tile.z < 0 ? 0 : tile.z * SWEEP_CHUNK * SB;
    const int qEnd = tile.z < 0 ? c : min(c, qBegin + SWEEP_CHUNK * SB);
    const TP* L = panels + p.panelOff[s];
    double4_t acc = {0, 0, 0, 0};
    // Block q0 + SB is loaded while block q0 is multiplied.
    StagedBlock<TP> nextL;
    StagedRows nextY;
    nextL.load(L, n, i0, rows, qBegin, min(SB, qEnd - qBegin));
    nextY.load(X, first + qBegin, min(SB, qEnd - qBegin), k);
    for (int q0 = qBegin; q0 < qEnd; q0 += SB)
    {
        __syncthreads(); // the previous block's products have read the LDS
        nextL.store(blk);
        nextY.store(ys);
        __syncthreads();
        if (q0 + SB < qEnd)
        {
            const int qb = min(SB, qEnd - q0 - SB);
            nextL.load(L, n, i0, rows, q0 + SB, qb);
            nextY.load(X, first + q0 + SB, qb, k);
        }
        acc = rowsTimesVectors(blk, ys, acc);
    }
    const int h = accColumn();
    if (tile.w >= 0)
    {
        double* out = partials + int64_t(tile.w) * SB * KMAX;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            out[accRow(v) * KMAX + h] = acc[v];
        return;
    }
    double* u = U + p.rowStart[s] * k;
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const int i = accRow(v);
        if (i < rows && h < k)
            u[int64_t(i0 - c + i) * k + h] -= acc[v];
    }
}

// X or U rows −= the sum of an item's chunks of partial sums, in chunk order (or = with bit 29 of .w); one workgroup
// per item.
__global__ __launch_bounds__(256) void subtractPartialsKernel(double* X, double* U, int k, const int4* items,
                                                              const double* partials)
{
    const int4 item = items[blockIdx.x];
    const bool toU = (item.w >> 30) & 1, assign = (item.w >> 29) & 1;
    const int chunks = item.w & ((1 << 29) - 1);
    double* dst = (toU ? U : X) + int64_t(item.x) * k;
    for (int idx = threadIdx.x; idx < item.y * KMAX; idx += blockDim.x)
    {
        const int i = idx / KMAX, h = idx % KMAX;
        if (h >= k)
            continue;
        double sum = 0.0;
        for (int ch = 0; ch < chunks; ++ch)
            sum += partials[int64_t(item.z + ch) * SB * KMAX + idx];
        double& at = dst[int64_t(i) * k + h];
        at = assign ? sum : at - sum;
    }
}

// Lanes of a row of a row-major n × k array (k <= 16) in the vector kernels: the smallest power of two >= k, so a
// workgroup of 256 threads covers 256 >> vectorShift(k) rows a pass (a thread per row for one right-hand side).
__host__ __device__ constexpr int vectorShift(int k)
{
    return k <= 1 ? 0 : k <= 2 ? 1 : k <= 4 ? 2 : k <= 8 ? 3 : 4;
}

// Thread threadIdx.x of a vector kernel: column h (idle when h >= k) of the rows slot, slot + rows, ... of its
// workgroup's passes.
struct VectorLanes
{
    int h, slot, rows;
    __device__ explicit VectorLanes(int k)
    {
        const int shift = vectorShift(k);
        h = threadIdx.x & ((1 << shift) - 1);
        slot = threadIdx.x >> shift;
        rows = 256 >> shift;
    }
    __device__ int64_t firstRow() const { return blockIdx.x * int64_t(rows) + slot; }
    __device__ int64_t stride() const { return int64_t(gridDim.x) * rows; }
};

// X ← D⁻¹·X, X row-major n × k.
__global__ __launch_bounds__(256) void diagonalSolveKernel(double* X, const double* d, int64_t n, int k)
{
    const VectorLanes v(k);
    if (v.h >= k)
        return;
    for (int64_t i = v.firstRow(); i < n; i += v.stride())
        X[i * k + v.h] /= d[i];
}

__device__ inline void gatherRows(double* xs, const double* X, const int32_t* rows, int first, int c, int i0, int ib,
                                  int k)
{
    StagedRows r;
    r.gather(X, rows, first, c, i0, ib, k);
    r.store(xs);
}

// X(cols of block J) −= L_RF(:, J)ᵀ·X(rows), for the column blocks of the big and wide supernodes of a level.
template <typename TP>
__global__ __launch_bounds__(256) void backwardColumnsKernel(DevicePlan p, const TP* panels, double* X, int k,
                                                             const int4* blocks, double* partials)
{
    __shared__ TP blk[SB * SBP];
    __shared__ double xs[SB * KMAX];
    const int4 block = blocks[blockIdx.x];
    const int s = block.x;
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    const int q0 = block.y * SB, qb = min(SB, c - q0);
    // The rows of this chunk (all of L_RF without chunks).
    const int iBegin = block.z < 0 ? c : c + block.z * SWEEP_CHUNK * SB;
    const int iEnd = block.z < 0 ? n : min(n, iBegin + SWEEP_CHUNK * SB);
    const TP* L = panels + p.panelOff[s];
    const int32_t* rows = p.rows + p.rowStart[s];
    double4_t acc = {0, 0, 0, 0};
    // Row block i0 + SB is loaded while block i0 is multiplied.
    StagedBlock<TP> nextL;
    StagedRows nextX;
    nextL.load(L, n, iBegin, min(SB, iEnd - iBegin), q0, qb);
    nextX.gather(X, rows, first, c, iBegin, min(SB, iEnd - iBegin), k);
    for (int i0 = iBegin; i0 < iEnd; i0 += SB)
    {
        __syncthreads(); // the previous block's products have read the LDS
        nextL.store(blk);
        nextX.store(xs);
        __syncthreads();
        if (i0 + SB < iEnd)
        {
            const int ib = min(SB, iEnd - i0 - SB);
            nextL.load(L, n, i0 + SB, ib, q0, qb);
            nextX.gather(X, rows, first, c, i0 + SB, ib, k);
        }
        acc = columnsTimesVectors(blk, xs, acc);
    }
    const int h = accColumn();
    if (block.w >= 0)
    {
        double* out = partials + int64_t(block.w) * SB * KMAX;
#pragma unroll
        for (int v = 0; v < 4; ++v)
            out[accRow(v) * KMAX + h] = acc[v];
        return;
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const int q = accRow(v);
        if (q < qb && h < k)
            X[int64_t(first + q0 + q) * k + h] -= acc[v];
    }
}

template <typename TP>
__global__ __launch_bounds__(256) void backwardKernel(DevicePlan p, const TP* panels, const TP* linv, double* X, int k,
                                                      const int32_t* list)
{
    __shared__ TP blk[SB * SBP];
    __shared__ double xs[SB * KMAX];
    __shared__ double ys[SB * KMAX];
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    if (isWideSupernode(c) || c == 0)
        return; // backwardColumnsKernel and triangleProductsKernel did it; or no columns (all delayed)
    const TP* L = panels + p.panelOff[s];
    const int32_t* rows = p.rows + p.rowStart[s];
    const int h = accColumn();
    const int lastBlock = ((c - 1) / SB) * SB;
    for (int q0 = lastBlock; q0 >= 0; q0 -= SB)
    {
        const int qb = min(SB, c - q0);
        // The block's inverse and its own rows of X, loaded with the first products' operands (used after them).
        StagedBlock<TP> inverse;
        inverse.load(linv + p.linvOff[s] + int64_t(q0) * SB, qb, 0, qb, 0, qb);
        StagedRows own;
        own.load(X, first + q0, qb, k);
        // Dot products of the block's columns with the final rows after it: the supernode's own columns after the
        // block, then the rows of L_RF (unless a big supernode, whose column blocks did them).
        double4_t acc = {0, 0, 0, 0};
        // A larger supernode's rows of L_RF are the column blocks of backwardColumnsKernel.
        const int rowsEnd = sweepRowsInline(c, r) ? n : c;
        // Tile i0 + SB's rows of X and block of L are loaded into registers while tile i0 is multiplied (the first
        // with the inverse and the block's own rows).
        StagedRows nextRows;
        StagedBlock<TP> nextBlock;
        const auto loadTile = [&](int i0)
        {
            const int ib = min(SB, rowsEnd - i0);
            nextRows.gather(X, rows, first, c, i0, ib, k);
            nextBlock.load(L, n, i0, ib, q0, qb);
        };
        if (q0 + qb < rowsEnd)
            loadTile(q0 + qb);
        for (int i0 = q0 + qb; i0 < rowsEnd; i0 += SB)
        {
            nextRows.store(xs);
            nextBlock.store(blk);
            __syncthreads();
            if (i0 + SB < rowsEnd)
                loadTile(i0 + SB);
            acc = columnsTimesVectors(blk, xs, acc);
            __syncthreads();
        }
        own.store(ys); // zero beyond the block's rows and the k columns
        __syncthreads();
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int q = accRow(v);
            if (q < qb && h < k)
                ys[q * KMAX + h] -= acc[v];
        }
        inverse.store(blk);
        __syncthreads();
        // X_J ← L_JJ⁻ᵀ·y_J.
        {
            const double4_t out = columnsTimesVectors(blk, ys, double4_t{0, 0, 0, 0});
            __syncthreads();
#pragma unroll
            for (int v = 0; v < 4; ++v)
                ys[accRow(v) * KMAX + h] = out[v];
            __syncthreads();
        }
        storeBlockRows(ys, X, first + q0, qb, k);
        __syncthreads();
    }
}

// The triangle of a wide supernode in a sweep, by the whole inverse X = L_FF⁻¹: item (s, I) writes rows of block I
// of W: Σ_{J<=I} X_IJ·x_J (forward) or Σ_{J>=I} X_JIᵀ·x_J (transpose), x = X(cols of s); every item reads x only.
template <typename TP>
__global__ __launch_bounds__(256) void triangleProductsKernel(DevicePlan p, const TP* linv, const double* X, int k,
                                                              const int4* items, int transpose, double* partials)
{
    __shared__ TP blk[SB * SBP];
    __shared__ double ys[SB * KMAX];
    const int4 item = items[blockIdx.x];
    const int s = item.x, I = item.y;
    const int c = p.ncol[s], nb = (c + SB - 1) / SB;
    const int first = p.first[s];
    const TP* Xinv = linv + p.linvOff[s] + blockInverseValues(c);
    const int ib = min(SB, c - SB * I);
    double4_t acc = {0, 0, 0, 0};
    // The blocks J of this chunk (SWEEP_CHUNK of them) of the item's range, J <= I (forward) or J >= I (transpose).
    const int rangeFirst = transpose ? I : 0, rangeEnd = transpose ? nb : I + 1;
    const int jFirst = rangeFirst + item.z * SWEEP_CHUNK, jEnd = min(rangeEnd, jFirst + SWEEP_CHUNK);
    // Block J + 1 is loaded while block J is multiplied.
    StagedBlock<TP> nextB;
    StagedRows nextY;
    const auto loadJ = [&](int J)
    {
        const int jb = min(SB, c - SB * J);
        nextY.load(X, first + SB * J, jb, k);
        if (transpose)
            nextB.load(Xinv, c, SB * J, jb, SB * I, ib);
        else
            nextB.load(Xinv, c, SB * I, ib, SB * J, jb);
    };
    loadJ(jFirst);
    for (int J = jFirst; J < jEnd; ++J)
    {
        __syncthreads(); // the previous block's products have read the LDS
        nextB.store(blk);
        nextY.store(ys);
        __syncthreads();
        if (J + 1 < jEnd)
            loadJ(J + 1);
        acc = transpose ? columnsTimesVectors(blk, ys, acc) : rowsTimesVectors(blk, ys, acc);
    }
    // Into the item's slot of partial sums: the chunks' sums become the block's rows of X once every item has read X.
    const int h = accColumn();
    double* out = partials + int64_t(item.w) * SB * KMAX;
#pragma unroll
    for (int v = 0; v < 4; ++v)
        out[accRow(v) * KMAX + h] = acc[v];
}

// ---- vectors and the sparse product --------------------------------------------------------------------------------

// Column-major (ld) ↔ row-major k-wide layout, with an optional per-equation scale (exact powers of two).
__global__ void toRowMajorKernel(const double* src, int64_t ld, double* dst, int64_t n, int k, const double* scale)
{
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n * k;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / k;
        const int h = int(idx - i * k);
        const double v = src[i + h * ld];
        dst[idx] = scale ? v * scale[i] : v;
    }
}

// The column maxima of a workgroup's rows (m[h] for its thread's rows) into norm[h] (bits, atomicMax; NaN wins),
// through the wave's shuffles and LDS: every thread of the workgroup calls it.
__device__ inline void columnMaxima(const double (&m)[16], int k, unsigned long long* norm)
{
    __shared__ double slots[4][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int h = 0; h < 16; ++h)
    {
        double v = m[h];
#pragma unroll
        for (int offset = 32; offset > 0; offset >>= 1)
        {
            const double o = __shfl_xor(v, offset);
            v = o > v || o != o ? o : v;
        }
        if (lane == 0)
            slots[wave][h] = v;
    }
    __syncthreads();
    if (threadIdx.x < k)
    {
        double v = slots[0][threadIdx.x];
        for (int w = 1; w < 4; ++w)
        {
            const double o = slots[w][threadIdx.x];
            v = o > v || o != o ? o : v;
        }
        atomicMax(&norm[threadIdx.x], static_cast<unsigned long long>(__double_as_longlong(v)));
    }
}

// Right-hand sides into the solve's vectors, 256 rows per workgroup pass through an LDS tile (the column-major loads
// coalesced per column, the row-major stores coalesced over the tile's contiguous k·256 values): B = src row-major,
// X1 = S·src row-major (X1 null: none; S null: a copy), and the column maxima |src| into norm (bits, zeroed by the
// caller; null: none).
__global__ __launch_bounds__(256) void rightHandSidesKernel(const double* src, int64_t ld, double* B, double* X1,
                                                           const double* scale, int64_t n, int k,
                                                           unsigned long long* norm)
{
    __shared__ double tile[256 * 17];
    __shared__ double rowScale[256];
    double m[16];
#pragma unroll
    for (int h = 0; h < 16; ++h)
        m[h] = 0.0;
    const int t = threadIdx.x;
    for (int64_t r0 = blockIdx.x * int64_t(256); r0 < n; r0 += int64_t(gridDim.x) * 256)
    {
        const int rows = int(min(int64_t(256), n - r0));
        const int64_t i = r0 + t;
#pragma unroll
        for (int h = 0; h < 16; ++h)
            if (h < k && t < rows)
            {
                const double v = src[i + h * ld];
                tile[t * 17 + h] = v;
                const double a = fabs(v);
                m[h] = a > m[h] || a != a ? a : m[h];
            }
        if (t < rows)
            rowScale[t] = scale ? scale[i] : 1.0;
        __syncthreads();
        const int values = rows * k;
        for (int e = t; e < values; e += 256)
        {
            const int row = e / k, h = e - row * k;
            const double v = tile[row * 17 + h];
            B[r0 * k + e] = v;
            if (X1)
                X1[r0 * k + e] = scale ? v * rowScale[row] : v;
        }
        __syncthreads(); // the tile is read before the next pass writes it
    }
    if (norm)
        columnMaxima(m, k, norm);
}

// dst (column-major, leading dimension ld) = S·src (src row-major n × k; S null: the identity), or += with
// accumulate; with norm, the column maxima |S·src| into it (bits, zeroed by the caller). 256 rows per workgroup pass
// through an LDS tile (the row-major loads coalesced over the tile's contiguous values, the stores per column).
__global__ __launch_bounds__(256) void columnsFromRowsKernel(const double* src, double* dst, int64_t ld, int64_t n,
                                                            int k, const double* scale, int accumulate,
                                                            unsigned long long* norm)
{
    __shared__ double tile[256 * 17];
    double m[16];
#pragma unroll
    for (int h = 0; h < 16; ++h)
        m[h] = 0.0;
    const int t = threadIdx.x;
    for (int64_t r0 = blockIdx.x * int64_t(256); r0 < n; r0 += int64_t(gridDim.x) * 256)
    {
        const int rows = int(min(int64_t(256), n - r0));
        const int values = rows * k;
        for (int e = t; e < values; e += 256)
        {
            const int row = e / k, h = e - row * k;
            tile[row * 17 + h] = src[r0 * k + e];
        }
        __syncthreads();
        const int64_t i = r0 + t;
        const double s = scale && t < rows ? scale[i] : 1.0;
#pragma unroll
        for (int h = 0; h < 16; ++h)
            if (h < k && t < rows)
            {
                const double v = scale ? tile[t * 17 + h] * s : tile[t * 17 + h];
                if (accumulate)
                    dst[i + h * ld] += v;
                else
                    dst[i + h * ld] = v;
                const double a = fabs(v);
                m[h] = a > m[h] || a != a ? a : m[h];
            }
        __syncthreads(); // the tile is read before the next pass writes it
    }
    if (norm)
        columnMaxima(m, k, norm);
}

__global__ void fromRowMajorKernel(const double* src, double* dst, int64_t ld, int64_t n, int k, const double* scale,
                                   int accumulate)
{
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n * k;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / k;
        const int h = int(idx - i * k);
        double v = src[idx];
        if (scale)
            v *= scale[i];
        if (accumulate)
            dst[i + h * ld] += v;
        else
            dst[i + h * ld] = v;
    }
}

// r = b − A·x (or y = A·x when b is null) with the full symmetric CSR, column-major vectors of leading dimension ld;
// one thread per (row, column), the columns of a row side by side.
__global__ void residualKernel(const int64_t* offsets, const int32_t* cols, const double* values, const double* x,
                               const double* b, double* out, int64_t ld, int64_t n, int k)
{
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n * k;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx / k;
        const int h = int(idx - i * k);
        const double* xc = x + h * ld;
        double sum = 0.0;
        for (int64_t e = offsets[i]; e < offsets[i + 1]; ++e)
            sum += values[e] * xc[cols[e]];
        out[i + h * ld] = b ? b[i + h * ld] - sum : sum;
    }
}

__global__ void fillDoubleKernel(double* data, int64_t ld, int64_t n, int k, double value)
{
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n * k;
         idx += int64_t(gridDim.x) * blockDim.x)
        data[(idx % n) + (idx / n) * ld] = value;
}

// ---- Krylov vectors: n × k row-major (k <= 16), a row per 1, 2, 4, 8 or 16 lanes (VectorLanes) --------------------
//
// No integer division, the k values of a row are one contiguous run of loads, and no lane idles for k a power of two.

constexpr int VECTOR_ROWS = 16; // row groups of a workgroup's pass in the sparse product

// out = b − A·x (b null: out = A·x), A the full symmetric matrix by rows (the CSR copy of the lower CSC), by row
// groups: runs of up to ROW_GROUP consecutive rows with the same columns (the degrees of freedom of one node). A
// group per 16 lanes (four groups in flight per wave), lane h its column h: each x row is read once for the whole
// group, and each row's products are summed in its columns' order. With a row scaling S (powers of two, so every
// scaling is exact) the product is the scaled matrix's: out = S·(b − A·S·x), each value a_ij taken as a_ij·s_j.
constexpr int ROW_GROUP = 8;

// The sparse product's workgroup index, XCD by XCD: the dispatcher hands workgroup b to XCD b mod 8, so XCD x takes
// the contiguous range [x·G/8, (x + 1)·G/8) of row groups (G, the grid, a multiple of 8), and rows that share columns
// read their x rows through the same L2.
__device__ inline int64_t xcdOrderedBlock()
{
    const int perXcd = gridDim.x >> 3;
    return int64_t(blockIdx.x & 7) * perXcd + (blockIdx.x >> 3);
}

// multiplyRowsKernel for one right-hand side: the 16 lanes of a group share its entries (lane j the entries j, j + 16,
// ...), and the lanes' sums of each row are added by a fixed butterfly over the 16 lanes.
__device__ inline void multiplyRowsSingle(const int64_t* offsets, const int32_t* cols, const double* values,
                                          const int32_t* groups, int64_t groupCount, const double* x,
                                          const double* b, double* out, const double* scale)
{
    const int lane = threadIdx.x & 15;
    for (int64_t g = xcdOrderedBlock() * VECTOR_ROWS + (threadIdx.x >> 4); g < groupCount;
         g += int64_t(gridDim.x) * VECTOR_ROWS)
    {
        const int32_t r0 = groups[g], rows = groups[g + 1] - r0;
        const int64_t e0 = offsets[r0], degree = offsets[r0 + 1] - e0;
        int64_t start[ROW_GROUP];
        double sum[ROW_GROUP];
#pragma unroll
        for (int t = 0; t < ROW_GROUP; ++t)
        {
            start[t] = offsets[r0 + min(t, rows - 1)];
            sum[t] = 0.0;
        }
        for (int64_t e = lane; e < degree; e += 16)
        {
            const int32_t c = cols[e0 + e];
            const double xv = scale ? x[c] * scale[c] : x[c];
#pragma unroll
            for (int t = 0; t < ROW_GROUP; ++t)
                if (t < rows)
                    sum[t] += values[start[t] + e] * xv;
        }
        // Every lane of the group runs the butterfly (the group's lanes share g, so all of them are here).
        double mine = 0.0;
#pragma unroll
        for (int t = 0; t < ROW_GROUP; ++t)
        {
#pragma unroll
            for (int offset = 8; offset > 0; offset >>= 1)
                sum[t] += __shfl_xor(sum[t], offset, 16);
            mine = lane == t ? sum[t] : mine;
        }
        if (lane < rows)
        {
            const double y = b ? b[r0 + lane] - mine : mine;
            out[r0 + lane] = scale ? y * scale[r0 + lane] : y;
        }
    }
}

__global__ __launch_bounds__(256) void multiplyRowsKernel(const int64_t* offsets, const int32_t* cols,
                                                         const double* values, const int32_t* groups,
                                                         int64_t groupCount, const double* x, const double* b,
                                                         double* out, int k, const double* scale)
{
    if (k == 1)
    {
        multiplyRowsSingle(offsets, cols, values, groups, groupCount, x, b, out, scale);
        return;
    }
    // Lanes h >= k load and broadcast like the others (their x column clamped) and store nothing.
    const int lane = threadIdx.x & 15, h = min(lane, k - 1);
    for (int64_t g = xcdOrderedBlock() * VECTOR_ROWS + (threadIdx.x >> 4); g < groupCount;
         g += int64_t(gridDim.x) * VECTOR_ROWS)
    {
        const int32_t r0 = groups[g], rows = groups[g + 1] - r0;
        const int64_t e0 = offsets[r0], degree = offsets[r0 + 1] - e0;
        int64_t start[ROW_GROUP];
        double sum[ROW_GROUP];
#pragma unroll
        for (int t = 0; t < ROW_GROUP; ++t)
        {
            start[t] = offsets[r0 + min(t, rows - 1)];
            sum[t] = 0.0;
        }
        // Runs of 16 entries: lane j loads entry j's column and the group's values in it (coalesced, positions
        // clamped into the run); each entry's column and values are then broadcast from its lane, in the entries'
        // order, so every row's products are summed in its columns' order.
        for (int64_t e = 0; e < degree; e += 16)
        {
            const int run = int(min(int64_t(16), degree - e));
            const int j = min(lane, run - 1);
            const int cj = cols[e0 + e + j];
            const double sj = scale ? scale[cj] : 1.0;
            double vj[ROW_GROUP];
#pragma unroll
            for (int t = 0; t < ROW_GROUP; ++t)
                vj[t] = t < rows ? values[start[t] + e + j] * sj : 0.0;
            if (run == 16)
            {
                double xv[16];
#pragma unroll
                for (int u = 0; u < 16; ++u)
                    xv[u] = x[int64_t(__shfl(cj, u, 16)) * k + h];
#pragma unroll
                for (int u = 0; u < 16; ++u)
#pragma unroll
                    for (int t = 0; t < ROW_GROUP; ++t)
                        sum[t] += __shfl(vj[t], u, 16) * xv[u];
            }
            else
                for (int u = 0; u < run; ++u)
                {
                    const double xu = x[int64_t(__shfl(cj, u, 16)) * k + h];
#pragma unroll
                    for (int t = 0; t < ROW_GROUP; ++t)
                        sum[t] += __shfl(vj[t], u, 16) * xu;
                }
        }
        if (lane >= k)
            continue;
#pragma unroll
        for (int t = 0; t < ROW_GROUP; ++t)
            if (t < rows)
            {
                const int64_t at = int64_t(r0 + t) * k + h;
                const double y = b ? b[at] - sum[t] : sum[t];
                out[at] = scale ? y * scale[r0 + t] : y;
            }
    }
}

// Per column the largest |v| (bits of a non-negative double, ordered as the values; NaN wins): the workgroup's 16
// row slots combined in LDS, then one atomic per workgroup and column. With a row scaling (powers of two) the
// maximum is that of S·v (divide: S⁻¹·v).
__global__ __launch_bounds__(256) void columnMaxRowsKernel(const double* v, int64_t n, int k, unsigned long long* out,
                                                          const double* scale, int divide)
{
    __shared__ double slots[256];
    const VectorLanes lanes(k);
    double m = 0.0;
    if (lanes.h < k)
        for (int64_t i = lanes.firstRow(); i < n; i += lanes.stride())
        {
            const double value = v[i * k + lanes.h];
            const double a = fabs(scale ? (divide ? value / scale[i] : value * scale[i]) : value);
            m = a > m || a != a ? a : m;
        }
    slots[threadIdx.x] = m;
    // The row slots' maxima by a fixed tree, column h ending in slots[h].
    for (int half = lanes.rows >> 1; half > 0; half >>= 1)
    {
        __syncthreads();
        if (lanes.slot < half)
        {
            const double a = slots[threadIdx.x + (half << vectorShift(k))];
            const double b = slots[threadIdx.x];
            slots[threadIdx.x] = a > b || a != a ? a : b;
        }
    }
    __syncthreads();
    if (threadIdx.x < k)
        atomicMax(&out[threadIdx.x], static_cast<unsigned long long>(__double_as_longlong(slots[threadIdx.x])));
}

// ---- matrix values on the device ----------------------------------------------------------------------------------

__global__ __launch_bounds__(256) void combineValuesKernel(int32_t n, const int64_t* offsets, const int32_t* rows,
                                                           double alpha, const double* a, double beta,
                                                           const double* b, const double* scale, double* out,
                                                           double* diagonal, double* fp32Scale)
{
#pragma clang fp contract(off)
    const int32_t j = int32_t(blockIdx.x) * 256 + int32_t(threadIdx.x);
    if (j >= n)
        return;
    const int64_t k0 = offsets[j], k1 = offsets[j + 1];
    for (int64_t k = k0; k < k1; ++k)
    {
        // Two products and a sum, each rounded (b null: the values of a as they are); then S·A·S as
        // scale[i]·a_ij·scale[j], left to right.
        double v = a[k];
        if (b != nullptr)
        {
            const double left = alpha * a[k];
            const double right = beta * b[k];
            v = left + right;
        }
        if (scale != nullptr)
        {
            const double rowScaled = scale[rows[k]] * v;
            v = rowScaled * scale[j];
        }
        out[k] = v;
    }
    const double d = (k0 < k1 && rows[k0] == j) ? out[k0] : 0.0;
    diagonal[j] = d;
    double s = 1.0;
    if (d > 0.0 && isfinite(d))
    {
        int e = 0;
        (void)frexp(d, &e);
        s = ldexp(1.0, -(e / 2));
    }
    fp32Scale[j] = s;
}

__global__ __launch_bounds__(256) void rowValuesKernel(int32_t n, const int64_t* csrOffsets, const int64_t* map,
                                                       const double* values, double* csrValues,
                                                       unsigned long long* normKey)
{
    const int32_t i = int32_t(blockIdx.x) * 256 + int32_t(threadIdx.x);
    double sum = 0.0;
    if (i < n)
        for (int64_t e = csrOffsets[i]; e < csrOffsets[i + 1]; ++e)
        {
            const double v = values[map[e]];
            csrValues[e] = v;
            sum += fabs(v);
        }
    // Non-negative doubles order as their bit patterns; a row whose sum is not a number is passed over, as the
    // host's std::max passes it over.
    const unsigned long long key = sum == sum ? __double_as_longlong(sum) : 0ull;
    unsigned long long top = key;
    for (int offset = 32; offset > 0; offset >>= 1)
    {
        const unsigned long long other = __shfl_xor(top, offset);
        top = other > top ? other : top;
    }
    if ((threadIdx.x & 63) == 0)
        atomicMax(normKey, top);
}

// ---- block conjugate gradients: Gram products and block combinations -------------------------------------------
//
// Gram products U_qᵀ·V_q of n × k row-major vectors by 16 × 16 × 4 FP64 MFMA, four rows at a time: lane l holds
// U(row + l/16, l mod 16) as A(a = l mod 16, i = l/16) and V(row + l/16, l mod 16) as B(i, b = l mod 16) (each lane
// loads one value of each operand, the 4 rows × k values coalesced). A workgroup takes a fixed contiguous range of
// rows, its waves rows 4w, 4w + 16, ...; the waves' sums are added in wave order and the workgroups' in index order.
__global__ __launch_bounds__(256) void blockGramKernel(GramPairs g, int64_t n, int k, double* partial)
{
    __shared__ double sums[4][GRAM_PAIRS][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t per = ((n + gridDim.x - 1) / gridDim.x + 3) / 4 * 4;
    const int64_t r0 = min(n, int64_t(blockIdx.x) * per), r1 = min(n, r0 + per);
    const int col = lane & 15;
    double4_t acc[GRAM_PAIRS];
#pragma unroll
    for (int q = 0; q < GRAM_PAIRS; ++q)
        acc[q] = double4_t{0, 0, 0, 0};
    for (int64_t row = r0 + 4 * wave; row < r1; row += 16)
    {
        const int64_t i = row + (lane >> 4);
        const bool ok = i < r1 && col < k;
        const int64_t at = ok ? i * k + col : 0;
        // A pair whose U is the first pair's takes the first pair's loaded value (one read of a shared vector).
        const double a0 = ok ? g.u[0][at] : 0.0;
#pragma unroll
        for (int q = 0; q < GRAM_PAIRS; ++q)
            if (q < g.pairs)
            {
                const double a = q == 0 || g.u[q] == g.u[0] ? a0 : (ok ? g.u[q][at] : 0.0);
                const double b = ok ? g.v[q][at] : 0.0;
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[q], 0, 0, 0);
            }
    }
#pragma unroll
    for (int q = 0; q < GRAM_PAIRS; ++q)
#pragma unroll
        for (int v = 0; v < 4; ++v)
            sums[wave][q][((lane >> 4) + 4 * v) * 16 + col] = acc[q][v];
    __syncthreads();
    for (int q = 0; q < g.pairs; ++q)
    {
        const int e = threadIdx.x;
        partial[(int64_t(blockIdx.x) * GRAM_PAIRS + q) * 256 + e] =
            ((sums[0][q][e] + sums[1][q][e]) + sums[2][q][e]) + sums[3][q][e];
    }
}

// out[q·256 + e] = Σ_g partial[(g·GRAM_PAIRS + q)·256 + e] in two fixed-order stages: workgroup (q, run) adds the
// GRAM_RUNS consecutive workgroups' partial sums of its run (all its loads in flight in four batches) into
// runs[(q·GRAM_RUN_COUNT + run)·256 + e], then sumGramRunsKernel adds the runs in order.
constexpr int GRAM_RUN_COUNT = 16, GRAM_RUNS = GRAM_GROUPS / GRAM_RUN_COUNT;
static_assert(GRAM_GROUPS % (4 * GRAM_RUN_COUNT) == 0, "runs of whole batches");
__global__ __launch_bounds__(256) void sumGramPartialsKernel(const double* partial, double* runs)
{
    const int q = blockIdx.x / GRAM_RUN_COUNT, run = blockIdx.x % GRAM_RUN_COUNT, e = threadIdx.x;
    const int g0 = run * GRAM_RUNS;
    double sum = 0.0;
    for (int g = 0; g < GRAM_RUNS; g += GRAM_RUNS / 4)
    {
        double v[GRAM_RUNS / 4];
#pragma unroll
        for (int t = 0; t < GRAM_RUNS / 4; ++t)
            v[t] = partial[(int64_t(g0 + g + t) * GRAM_PAIRS + q) * 256 + e];
#pragma unroll
        for (int t = 0; t < GRAM_RUNS / 4; ++t)
            sum += v[t];
    }
    runs[(q * GRAM_RUN_COUNT + run) * 256 + e] = sum;
}

__global__ __launch_bounds__(256) void sumGramRunsKernel(const double* runs, double* out)
{
    const int q = blockIdx.x, e = threadIdx.x;
    double v[GRAM_RUN_COUNT];
#pragma unroll
    for (int t = 0; t < GRAM_RUN_COUNT; ++t)
        v[t] = runs[(q * GRAM_RUN_COUNT + t) * 256 + e];
    double sum = 0.0;
#pragma unroll
    for (int t = 0; t < GRAM_RUN_COUNT; ++t)
        sum += v[t];
    out[q * 256 + e] = sum;
}

// Y(rows, :) = c0·Y(rows, :) + U(rows, :)·C, a wave per 16 rows: four 16 × 16 × 4 FP64 MFMA steps over the k
// columns of U (lane l: A(l mod 16, l / 16) = U(row0 + l mod 16, 4s + l / 16), B(l / 16, l mod 16) = C(4s + l / 16,
// l mod 16)), then each lane adds c0·Y to its four results and stores them.
__global__ __launch_bounds__(256) void blockCombineKernel(double* y, BlockTerm m, int64_t n, int k)
{
    const int lane = threadIdx.x & 63;
    const int64_t row0 = ((int64_t(blockIdx.x) * 256 + threadIdx.x) >> 6) * 16;
    if (row0 >= n)
        return; // the whole wave
    const int r = lane & 15, q = lane >> 4;
    const int64_t i = row0 + r;
    double4_t acc = {0, 0, 0, 0};
#pragma unroll
    for (int step = 0; step < 4; ++step)
    {
        const int a = 4 * step + q;
        const double av = (i < n && a < k) ? m.u[i * k + a] : 0.0;
        double bv = 0.0;
        if (a < k && r < k)
            bv = m.deviceC != nullptr ? m.scale * m.deviceC[a * 16 + r] : m.c[a * 16 + r];
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc, 0, 0, 0);
    }
#pragma unroll
    for (int v = 0; v < 4; ++v)
    {
        const int64_t row = row0 + q + 4 * v;
        if (row < n && r < k)
        {
            double* at = y + row * k + r;
            *at = (m.c0 == 0.0 ? 0.0 : m.c0 * *at) + acc[v];
        }
    }
}

// The step of block conjugate gradients in one pass, a wave per 16 rows (grid-stride): with W (in z) and A·W (in p)
// and the coefficient blocks T and U = T·α (16 × 16 each, row a, column j, at coef and coef + 256 on the device),
//   Q = (A·W)·T,  P = W·T,  X += W·U,  R −= (A·W)·U,  Z = R,
// each product by four 16 × 16 × 4 FP64 MFMA steps over the k columns (the operands' lanes as in
// blockCombineKernel). P overwrites A·W and Z overwrites W row block by row block, after the wave has read them.
// The column maxima of |S⁻¹·R| and |S·X| (the backward error's norms of the unscaled vectors) go to norms[0, 16)
// and norms[16, 32) (bits, atomicMax; zeroed by the caller) once per workgroup.
__global__ __launch_bounds__(256) void blockStepKernel(double* z, double* p, double* q, double* x, double* r,
                                                       const double* coef, const double* scale, int64_t n, int k,
                                                       unsigned long long* norms)
{
    __shared__ double slots[2][4][16];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    // This lane's B operands: T(4s + lq, lr) and U(4s + lq, lr).
    double bt[4], bu[4];
#pragma unroll
    for (int step = 0; step < 4; ++step)
    {
        const int a = 4 * step + lq;
        const bool valid = a < k && lr < k;
        bt[step] = valid ? coef[a * 16 + lr] : 0.0;
        bu[step] = valid ? coef[256 + a * 16 + lr] : 0.0;
    }
    double maxR = 0.0, maxX = 0.0;
    const int64_t blocks = (n + 15) / 16;
    for (int64_t blk = int64_t(blockIdx.x) * 4 + wave; blk < blocks; blk += int64_t(gridDim.x) * 4)
    {
        const int64_t row0 = blk * 16, i = row0 + lr;
        double aw[4], w[4];
#pragma unroll
        for (int step = 0; step < 4; ++step)
        {
            const int a = 4 * step + lq;
            const bool valid = i < n && a < k;
            aw[step] = valid ? p[i * k + a] : 0.0;
            w[step] = valid ? z[i * k + a] : 0.0;
        }
        double4_t dq = {0, 0, 0, 0}, dp = {0, 0, 0, 0}, dx = {0, 0, 0, 0}, dr = {0, 0, 0, 0};
#pragma unroll
        for (int step = 0; step < 4; ++step)
        {
            dq = __builtin_amdgcn_mfma_f64_16x16x4f64(aw[step], bt[step], dq, 0, 0, 0);
            dp = __builtin_amdgcn_mfma_f64_16x16x4f64(w[step], bt[step], dp, 0, 0, 0);
            dx = __builtin_amdgcn_mfma_f64_16x16x4f64(w[step], bu[step], dx, 0, 0, 0);
            dr = __builtin_amdgcn_mfma_f64_16x16x4f64(aw[step], bu[step], dr, 0, 0, 0);
        }
#pragma unroll
        for (int v = 0; v < 4; ++v)
        {
            const int64_t row = row0 + lq + 4 * v;
            if (row < n && lr < k)
            {
                const int64_t at = row * k + lr;
                const double xn = x[at] + dx[v];
                const double rn = r[at] - dr[v];
                q[at] = dq[v];
                p[at] = dp[v];
                x[at] = xn;
                r[at] = rn;
                z[at] = rn;
                const double s = scale ? scale[row] : 1.0;
                const double ar = fabs(rn / s), ax = fabs(xn * s);
                maxR = ar > maxR || ar != ar ? ar : maxR;
                maxX = ax > maxX || ax != ax ? ax : maxX;
            }
        }
    }
    // Column lr's maxima over the wave's four row quarters, then over the four waves.
#pragma unroll
    for (int offset = 16; offset < 64; offset <<= 1)
    {
        const double or_ = __shfl_xor(maxR, offset), ox = __shfl_xor(maxX, offset);
        maxR = or_ > maxR || or_ != or_ ? or_ : maxR;
        maxX = ox > maxX || ox != ox ? ox : maxX;
    }
    if (lq == 0)
    {
        slots[0][wave][lr] = maxR;
        slots[1][wave][lr] = maxX;
    }
    __syncthreads();
    if (threadIdx.x < 32)
    {
        const int which = threadIdx.x >> 4, h = threadIdx.x & 15;
        double m = slots[which][0][h];
        for (int w2 = 1; w2 < 4; ++w2)
        {
            const double a = slots[which][w2][h];
            m = a > m || a != a ? a : m;
        }
        if (h < k)
            atomicMax(&norms[16 * which + h], static_cast<unsigned long long>(__double_as_longlong(m)));
    }
}

// x += z (row-major n × k).
__global__ __launch_bounds__(256) void addRowsKernel(double* x, const double* z, int64_t n, int k)
{
    const VectorLanes v(k);
    const int h = v.h;
    if (h >= k)
        return;
    for (int64_t i = v.firstRow(); i < n; i += v.stride())
        x[i * k + h] += z[i * k + h];
}

// dst(i, h) = src(i, h)·scale[i] (scale null: a copy).
__global__ __launch_bounds__(256) void scaleRowsKernel(const double* src, double* dst, const double* scale, int64_t n,
                                                      int k)
{
    const VectorLanes v(k);
    const int h = v.h;
    if (h >= k)
        return;
    for (int64_t i = v.firstRow(); i < n; i += v.stride())
        dst[i * k + h] = scale ? src[i * k + h] * scale[i] : src[i * k + h];
}

// ---- launchers -----------------------------------------------------------------------------------------------------

namespace
{
int blocksFor(int64_t work, int threads)
{
    const int64_t b = (work + threads - 1) / threads;
    return int(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}

// Workgroups of a vector kernel over n rows of k values (over n row groups of the sparse product: k = 16).
int vectorGroups(int64_t n, int k)
{
    const int64_t rows = 256 >> vectorShift(k);
    const int64_t g = (n + rows - 1) / rows;
    return int(g < 1 ? 1 : (g > 4096 ? 4096 : g));
}
} // namespace

template <typename T> size_t smallFrontBytes(int maxOrder)
{
    return size_t(maxOrder) * maxOrder * sizeof(T) + size_t(maxOrder) * sizeof(T);
}

template <typename T>
hipError_t launchSmallFronts(const DevicePlan& p, const DeviceMatrix& a, T* panels, T* updates, T* linv,
                             const int32_t* list, int count, int maxOrder, double* diagOut, int32_t* zeroAt,
                             hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    const size_t bytes = smallFrontBytes<T>(maxOrder);
    hipLaunchKernelGGL(smallFrontKernel<T>, dim3(count), dim3(256), bytes, stream, p, a, panels, updates, linv, list,
                       diagOut, zeroAt);
    return hipGetLastError();
}

template <typename T>
hipError_t launchAssembly(const DevicePlan& p, const DeviceMatrix& a, T* panels, T* updates,
                          const AssemblyBlock* blocks, const AssemblyItem* items, int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    const size_t bytes = size_t(ASSEMBLY_BLOCK_VALUES) * sizeof(T);
    hipLaunchKernelGGL(assembleKernel<T>, dim3(count), dim3(256), bytes, stream, p, a, panels, updates, blocks,
                       items);
    return hipGetLastError();
}

template <typename T>
hipError_t launchDiagonalBlocks(const DevicePlan& p, T* panels, T* linv, const int32_t* fronts, int frontCount, int k0,
                                double* diagOut, int32_t* zeroAt, hipStream_t stream)
{
    if (frontCount == 0)
        return hipSuccess;
    // One wave per block: its barriers cost next to nothing.
    hipLaunchKernelGGL(diagonalBlockKernel<T>, dim3(frontCount), dim3(256), 0, stream, p, panels, linv, fronts, k0,
                       diagOut, zeroAt);
    return hipGetLastError();
}

template <typename T>
hipError_t launchUpdateTiles(const DevicePlan& p, T* panels, T* updates, T* linv, const UpdateTile* tiles, int count,
                             double* diagOut, int32_t* zeroAt, bool fused, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    if (fused)
        hipLaunchKernelGGL((updateTilesKernel<T, false, true>), dim3(count), dim3(256), 0, stream, p, panels, updates,
                           linv, tiles, count, diagOut, zeroAt);
    else
        hipLaunchKernelGGL((updateTilesKernel<T, false, false>), dim3(count), dim3(256), 0, stream, p, panels,
                           updates, linv, tiles, count, diagOut, zeroAt);
    return hipGetLastError();
}

template <typename T>
hipError_t launchExtendTiles(const DevicePlan& p, T* panels, T* updates, const UpdateTile* tiles, int count,
                             const int32_t* itemStarts, const AssemblyItem* items, int tileSize, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    if (tileSize == 128)
        hipLaunchKernelGGL((extendTilesKernel<T, 128>), dim3(count), dim3(256), 0, stream, p, panels, updates, tiles,
                           count, itemStarts, items);
    else if (tileSize == 64)
        hipLaunchKernelGGL((extendTilesKernel<T, 64>), dim3(count), dim3(256), 0, stream, p, panels, updates, tiles,
                           count, itemStarts, items);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launchPlainUpdateTiles(double* fronts, const double* w, const UpdateTile* tiles, int count,
                                  hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL((updateTilesKernel<double, true>), dim3(count), dim3(256), 0, stream, DevicePlan{}, fronts,
                       nullptr, const_cast<double*>(w), tiles, count, nullptr, nullptr);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchTinyForward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, double* U, int k,
                             const int32_t* list, int count, hipStream_t stream)
{
    if (count > 0)
        hipLaunchKernelGGL(tinyForwardKernel<TP>, dim3((count + TINY_WAVES - 1) / TINY_WAVES), dim3(64 * TINY_WAVES), 0,
                           stream, p, panels, linv, X, U, k, list, count);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchTinyBackward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, int k,
                              const int32_t* list, int count, hipStream_t stream)
{
    if (count > 0)
        hipLaunchKernelGGL(tinyBackwardKernel<TP>, dim3((count + TINY_WAVES - 1) / TINY_WAVES), dim3(64 * TINY_WAVES),
                           0, stream, p, panels, linv, X, k, list, count);
    return hipGetLastError();
}

// The forward sweep's gathers of a level's other supernodes, every front position at once: a workgroup per chunk of
// GATHER_CHUNK positions of one front (the chunk's record holds what the threads need of its supernode), 16 threads
// (the right-hand sides) per position. A column position adds its children's update rows to X (child order); a row
// position sets its row of U(s) to their sum (0 without children).
__global__ __launch_bounds__(256) void gatherKernel(const GatherChunk* chunks, const int32_t* gatherStart,
                                                    const int32_t* gatherRows, double* X, double* U, int k)
{
    const GatherChunk chunk = chunks[blockIdx.x];
    const int q = chunk.q0 + int(threadIdx.x >> 4), h = int(threadIdx.x & (KMAX - 1));
    if (q >= chunk.order || h >= k)
        return;
    const int64_t gp = int64_t(chunk.positions) + q;
    const int32_t e0 = gatherStart[gp], e1 = gatherStart[gp + 1];
    if (q < chunk.columns)
    {
        if (e1 == e0)
            return;
        double* x = X + int64_t(chunk.first + q) * k + h;
        double y = *x;
        for (int32_t e = e0; e < e1; ++e)
            y += U[int64_t(gatherRows[e]) * k + h];
        *x = y;
    }
    else
    {
        double y = 0.0;
        for (int32_t e = e0; e < e1; ++e)
            y += U[int64_t(gatherRows[e]) * k + h];
        U[(int64_t(chunk.rowStart) + q - chunk.columns) * k + h] = y;
    }
}

hipError_t launchGather(const GatherChunk* chunks, int chunkCount, const int32_t* gatherStart,
                        const int32_t* gatherRows, double* X, double* U, int k, hipStream_t stream)
{
    if (chunkCount > 0)
        hipLaunchKernelGGL(gatherKernel, dim3(unsigned(chunkCount)), dim3(256), 0, stream, chunks, gatherStart,
                           gatherRows, X, U, k);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchMediumForward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, double* U, int k,
                               const int32_t* list, int count, hipStream_t stream)
{
    if (count > 0)
        hipLaunchKernelGGL(mediumForwardKernel<TP>, dim3(count), dim3(256), 0, stream, p, panels, linv, X, U, k, list);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchMediumBackward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, int k,
                                const int32_t* list, int count, hipStream_t stream)
{
    if (count > 0)
        hipLaunchKernelGGL(mediumBackwardKernel<TP>, dim3(count), dim3(256), 0, stream, p, panels, linv, X, k, list);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchForward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, double* U, int k,
                         const int32_t* list, int count, const int4* rowTiles, int tileCount, double* partials,
                         hipStream_t stream, bool gathered)
{
    if (count > 0)
        hipLaunchKernelGGL(forwardKernel<TP>, dim3(count), dim3(256), 0, stream, p, panels, linv, X, U, k, list,
                           gathered ? 1 : 0);
    if (tileCount > 0)
        hipLaunchKernelGGL(forwardRowsKernel<TP>, dim3(tileCount), dim3(256), 0, stream, p, panels, X, U, k,
                           rowTiles, partials);
    return hipGetLastError();
}

hipError_t launchSubtractPartials(double* X, double* U, int k, const int4* items, int count, const double* partials,
                                  hipStream_t stream)
{
    if (count > 0)
        hipLaunchKernelGGL(subtractPartialsKernel, dim3(count), dim3(256), 0, stream, X, U, k, items, partials);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchBackward(const DevicePlan& p, const TP* panels, const TP* linv, double* X, int k, const int32_t* list,
                          int count, const int4* columnBlocks, int blockCount, double* partials, hipStream_t stream)
{
    if (blockCount > 0)
        hipLaunchKernelGGL(backwardColumnsKernel<TP>, dim3(blockCount), dim3(256), 0, stream, p, panels, X, k,
                           columnBlocks, partials);
    if (count > 0)
        hipLaunchKernelGGL(backwardKernel<TP>, dim3(count), dim3(256), 0, stream, p, panels, linv, X, k, list);
    return hipGetLastError();
}

template <typename T>
hipError_t launchTriangleInverse(const DevicePlan& p, const T* panels, T* linv, const int2* items, int count, int step,
                                 hipStream_t stream)
{
    if (count > 0)
        hipLaunchKernelGGL(triangleInverseKernel<T>, dim3(count), dim3(256), 0, stream, p, panels, linv, items, step);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchTriangleProducts(const DevicePlan& p, const TP* linv, const double* X, int k, const int4* items,
                                  int count, int transpose, double* partials, hipStream_t stream)
{
    if (count > 0)
        hipLaunchKernelGGL(triangleProductsKernel<TP>, dim3(count), dim3(256), 0, stream, p, linv, X, k, items,
                           transpose, partials);
    return hipGetLastError();
}

// The condensation of loads into the trailing columns of a partial factor (row-major X, n × k): X(first + j) += the
// update rows that land on trailing column j, U(pulls[pullStart[j] .. pullStart[j + 1])), in child order.
__global__ __launch_bounds__(256) void trailingGatherKernel(double* X, const double* U, int k, int first, int c,
                                                            const int32_t* pullStart, const int32_t* pulls)
{
    const int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (idx >= int64_t(c) * k)
        return;
    const int j = int(idx / k), h = int(idx - int64_t(j) * k);
    double y = X[int64_t(first + j) * k + h];
    for (int e = pullStart[j]; e < pullStart[j + 1]; ++e)
        y += U[int64_t(pulls[e]) * k + h];
    X[int64_t(first + j) * k + h] = y;
}

hipError_t launchTrailingGather(double* X, const double* U, int k, int first, int c, const int32_t* pullStart,
                                const int32_t* pulls, hipStream_t stream)
{
    const int64_t total = int64_t(c) * k;
    if (total > 0)
        hipLaunchKernelGGL(trailingGatherKernel, dim3(unsigned((total + 255) / 256)), dim3(256), 0, stream, X, U, k,
                           first, c, pullStart, pulls);
    return hipGetLastError();
}

hipError_t launchDiagonalSolve(double* X, const double* d, int64_t n, int k, hipStream_t stream)
{
    hipLaunchKernelGGL(diagonalSolveKernel, dim3(vectorGroups(n, k)), dim3(256), 0, stream, X, d, n, k);
    return hipGetLastError();
}

hipError_t launchFillInt32(int32_t* data, int64_t count, int32_t value, hipStream_t stream)
{
    hipLaunchKernelGGL(fillInt32Kernel, dim3(blocksFor(count, 256)), dim3(256), 0, stream, data, count, value);
    return hipGetLastError();
}

hipError_t launchToRowMajor(const double* src, int64_t ld, double* dst, int64_t n, int k, const double* scale,
                            hipStream_t stream)
{
    hipLaunchKernelGGL(toRowMajorKernel, dim3(blocksFor(n * k, 256)), dim3(256), 0, stream, src, ld, dst, n, k, scale);
    return hipGetLastError();
}

hipError_t launchFromRowMajor(const double* src, double* dst, int64_t ld, int64_t n, int k, const double* scale,
                              int accumulate, hipStream_t stream, unsigned long long* norm)
{
    if (norm != nullptr)
    {
        const int groups = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256)));
        hipLaunchKernelGGL(columnsFromRowsKernel, dim3(groups), dim3(256), 0, stream, src, dst, ld, n, k, scale,
                           accumulate, norm);
    }
    else
        hipLaunchKernelGGL(fromRowMajorKernel, dim3(blocksFor(n * k, 256)), dim3(256), 0, stream, src, dst, ld, n, k,
                           scale, accumulate);
    return hipGetLastError();
}

hipError_t launchRightHandSides(const double* src, int64_t ld, double* B, double* X1, const double* scale, int64_t n,
                                int k, unsigned long long* norm, hipStream_t stream)
{
    const int groups = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(1024, (n + 255) / 256)));
    hipLaunchKernelGGL(rightHandSidesKernel, dim3(groups), dim3(256), 0, stream, src, ld, B, X1, scale, n, k, norm);
    return hipGetLastError();
}

hipError_t launchResidual(const int64_t* offsets, const int32_t* cols, const double* values, const double* x,
                          const double* b, double* out, int64_t ld, int64_t n, int k, hipStream_t stream)
{
    hipLaunchKernelGGL(residualKernel, dim3(blocksFor(n * k, 256)), dim3(256), 0, stream, offsets, cols, values, x, b,
                       out, ld, n, k);
    return hipGetLastError();
}


hipError_t launchFillDouble(double* data, int64_t ld, int64_t n, int k, double value, hipStream_t stream)
{
    hipLaunchKernelGGL(fillDoubleKernel, dim3(blocksFor(n * k, 256)), dim3(256), 0, stream, data, ld, n, k, value);
    return hipGetLastError();
}


// The dynamic LDS limits of the kernels that take more than the default, set once per process before any launch
// (outside any graph capture).
hipError_t configureKernels()
{
    const void* kernels[4] = {reinterpret_cast<const void*>(&smallFrontKernel<double>),
                              reinterpret_cast<const void*>(&smallFrontKernel<float>),
                              reinterpret_cast<const void*>(&assembleKernel<double>),
                              reinterpret_cast<const void*>(&assembleKernel<float>)};
    const size_t bytes[4] = {smallFrontBytes<double>(SMALL_FRONT_MAX), smallFrontBytes<float>(SMALL_FRONT_MAX),
                             size_t(ASSEMBLY_BLOCK_VALUES) * sizeof(double), size_t(ASSEMBLY_BLOCK_VALUES) * sizeof(float)};
    for (int q = 0; q < 4; ++q)
    {
        const hipError_t e = hipFuncSetAttribute(kernels[q], hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes[q]));
        if (e != hipSuccess)
            return e;
    }
    return hipSuccess;
}

hipError_t launchMultiplyRows(const int64_t* offsets, const int32_t* cols, const double* values,
                              const int32_t* groups, int64_t groupCount, const double* x, const double* b, double* out,
                              int k, hipStream_t stream, const double* scale)
{
    // A workgroup per VECTOR_ROWS row groups (one pass), the grid a multiple of the 8 XCDs.
    const int64_t blocks = std::min<int64_t>((groupCount + VECTOR_ROWS - 1) / VECTOR_ROWS, int64_t(1) << 24);
    const int grid = static_cast<int>(std::max<int64_t>(8, (blocks + 7) / 8 * 8));
    hipLaunchKernelGGL(multiplyRowsKernel, dim3(grid), dim3(256), 0, stream, offsets, cols, values, groups,
                       groupCount, x, b, out, k, scale);
    return hipGetLastError();
}

// ---- the CSC of a block matrix (SymmetricCsc::fromBlocks) ----------------------------------------------------------
//
// A thread per column node j walks its blocks (the diagonal block first, then the blocks of the rows below it) and
// their lower-triangle terms of kept DOFs: a term of equations p, q lands in column min(p, q) at row max(p, q). The
// counts per column are atomic integer sums (exact); the terms are scattered to atomic slots of their column, then
// each column is put in row order by ranks (every row appears once in a column), so the result does not depend on the
// order of the atomics.

// Bits of the validation flag.
constexpr int BLOCKS_BAD_DOFS = 1, BLOCKS_BAD_COLUMNS = 2, BLOCKS_BAD_ROWS = 4, BLOCKS_BAD_VALUES = 8,
              BLOCKS_BAD_EQUATIONS = 16;

__global__ __launch_bounds__(256) void blocksValidateKernel(BlockMatrixView m, const int32_t* equationOfDof,
                                                            int64_t dofCount, int32_t kept, int64_t blockCount,
                                                            int64_t valueCount, int32_t* hits, int32_t* flag)
{
    const int32_t j = int32_t(blockIdx.x) * 256 + int32_t(threadIdx.x);
    if (j >= m.nodes)
        return;
    int bad = 0;
    const int64_t fj = m.firstDof[j];
    const int64_t dj = m.firstDof[j + 1] - fj;
    if (dj < 1 || dj > 6 || fj < 0 || fj + dj > dofCount)
        bad |= BLOCKS_BAD_DOFS;
    else
        for (int c = 0; c < int(dj); ++c)
        {
            const int32_t q = equationOfDof[fj + c];
            if (q < -1 || q >= kept)
                bad |= BLOCKS_BAD_EQUATIONS;
            else if (q >= 0)
                atomicAdd(&hits[q], 1);
        }
    const int64_t b0 = m.columnStart[j], b1 = m.columnStart[j + 1];
    if (b0 < 0 || b1 < b0 || b1 > blockCount)
        bad |= BLOCKS_BAD_COLUMNS;
    else if (!(bad & BLOCKS_BAD_DOFS))
        for (int64_t b = b0; b < b1; ++b)
        {
            const int32_t row = m.blockRows[b];
            // The diagonal block first, then the rows below ascending (each node pair once).
            if (b == b0 ? row != j : (row <= m.blockRows[b - 1] || row >= m.nodes))
            {
                bad |= BLOCKS_BAD_ROWS;
                break;
            }
            const int64_t dr = m.firstDof[row + 1] - m.firstDof[row];
            const int64_t v0 = m.blockValueOffsets[b], v1 = m.blockValueOffsets[b + 1];
            if (dr < 1 || dr > 6 || v0 < 0 || v1 > valueCount || v1 - v0 != dr * dj)
            {
                bad |= BLOCKS_BAD_VALUES;
                break;
            }
        }
    if (bad)
        atomicOr(flag, bad);
}

// Every kept equation is the equation of exactly one DOF.
__global__ __launch_bounds__(256) void equationHitsKernel(const int32_t* hits, int32_t kept, int32_t* flag)
{
    const int32_t q = int32_t(blockIdx.x) * 256 + int32_t(threadIdx.x);
    if (q < kept && hits[q] != 1)
        atomicOr(flag, BLOCKS_BAD_EQUATIONS);
}

// Counts the terms of each column (scatter = false), or writes them to the slots cursor[column]++ (scatter = true).
template <bool SCATTER>
__global__ __launch_bounds__(256) void blockTermsKernel(BlockMatrixView m, const int32_t* equationOfDof,
                                                        int32_t* counts, unsigned long long* cursor, int32_t* rows,
                                                        double* values)
{
    const int32_t j = int32_t(blockIdx.x) * 256 + int32_t(threadIdx.x);
    if (j >= m.nodes)
        return;
    const int64_t fj = m.firstDof[j];
    const int dj = int(m.firstDof[j + 1] - fj);
    for (int64_t b = m.columnStart[j]; b < m.columnStart[j + 1]; ++b)
    {
        const int32_t node = m.blockRows[b];
        const int64_t fr = m.firstDof[node];
        const int dr = int(m.firstDof[node + 1] - fr);
        const double* v = m.values + m.blockValueOffsets[b];
        for (int c = 0; c < dj; ++c)
        {
            const int32_t q = equationOfDof[fj + c];
            if (q < 0)
                continue;
            for (int r = node == j ? c : 0; r < dr; ++r)
            {
                const int32_t p = equationOfDof[fr + r];
                if (p < 0)
                    continue;
                const int32_t column = min(p, q);
                if (SCATTER)
                {
                    const unsigned long long at = atomicAdd(&cursor[column], 1ull);
                    rows[at] = max(p, q);
                    values[at] = v[int64_t(c) * dr + r];
                }
                else
                    atomicAdd(&counts[column], 1);
            }
        }
    }
}

// Exclusive scan of n int32 counts into int64 offsets (n + 1), three launches: sums of SCAN_ITEMS-element pieces,
// their exclusive scan (one workgroup, a carry across passes), the pieces' scans plus their bases.
constexpr int SCAN_ITEMS = 1024;

__device__ inline int64_t blockInclusiveScan(int64_t v, int64_t* lds)
{
    // Hillis–Steele over the 256 threads' values.
    const int t = threadIdx.x;
    lds[t] = v;
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1)
    {
        const int64_t add = t >= d ? lds[t - d] : 0;
        __syncthreads();
        lds[t] += add;
        __syncthreads();
    }
    const int64_t out = lds[t];
    __syncthreads();
    return out;
}

__global__ __launch_bounds__(256) void scanPieceSumsKernel(const int32_t* counts, int64_t n, int64_t* sums)
{
    __shared__ int64_t lds[256];
    const int64_t base = int64_t(blockIdx.x) * SCAN_ITEMS;
    int64_t v = 0;
    for (int u = 0; u < SCAN_ITEMS / 256; ++u)
    {
        const int64_t i = base + threadIdx.x * (SCAN_ITEMS / 256) + u;
        v += i < n ? counts[i] : 0;
    }
    v = blockInclusiveScan(v, lds);
    if (threadIdx.x == 255)
        sums[blockIdx.x] = v;
}

__global__ __launch_bounds__(256) void scanSumsKernel(int64_t* sums, int64_t pieces)
{
    __shared__ int64_t lds[256];
    __shared__ int64_t total;
    int64_t carry = 0;
    for (int64_t p0 = 0; p0 < pieces; p0 += 256)
    {
        const int64_t i = p0 + threadIdx.x;
        const int64_t v = i < pieces ? sums[i] : 0;
        const int64_t inclusive = blockInclusiveScan(v, lds);
        if (i < pieces)
            sums[i] = carry + inclusive - v; // exclusive
        if (threadIdx.x == 255)
            total = inclusive;
        __syncthreads();
        carry += total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(256) void scanPiecesKernel(const int32_t* counts, int64_t n, const int64_t* bases,
                                                        int64_t* offsets)
{
    __shared__ int64_t lds[256];
    constexpr int PER = SCAN_ITEMS / 256;
    const int64_t base = int64_t(blockIdx.x) * SCAN_ITEMS + threadIdx.x * PER;
    int64_t local[PER];
    int64_t v = 0;
    for (int u = 0; u < PER; ++u)
    {
        local[u] = base + u < n ? counts[base + u] : 0;
        v += local[u];
    }
    int64_t before = bases[blockIdx.x] + blockInclusiveScan(v, lds) - v;
    for (int u = 0; u < PER; ++u)
    {
        if (base + u < n)
            offsets[base + u] = before;
        before += local[u];
        if (base + u == n - 1)
            offsets[n] = before;
    }
}

// Each column in row order: a wave per column ranks its terms (the rows below it in the column), then writes each at
// its rank. Columns of at most RANK_LDS terms are ranked from LDS, longer ones from memory.
constexpr int RANK_LDS = 256;

__global__ __launch_bounds__(256) void sortColumnsKernel(const int64_t* offsets, int32_t n, const int32_t* rowsIn,
                                                         const double* valuesIn, int32_t* rowsOut, double* valuesOut)
{
    __shared__ int32_t lds[4][RANK_LDS];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int32_t column = int32_t(blockIdx.x) * 4 + w;
    const bool active = column < n;
    const int64_t k0 = active ? offsets[column] : 0;
    const int m = active ? int(offsets[column + 1] - k0) : 0;
    const int32_t* rows = rowsIn + k0;
    const bool inLds = m <= RANK_LDS;
    if (inLds)
        for (int e = lane; e < m; e += 64)
            lds[w][e] = rows[e];
    __syncthreads();
    for (int e = lane; e < m; e += 64)
    {
        const int32_t row = rows[e];
        int rank = 0;
        if (inLds)
            for (int f = 0; f < m; ++f)
                rank += lds[w][f] < row ? 1 : 0;
        else
            for (int f = 0; f < m; ++f)
                rank += rows[f] < row ? 1 : 0;
        rowsOut[k0 + rank] = row;
        valuesOut[k0 + rank] = valuesIn[k0 + e];
    }
}

hipError_t launchBlocksValidate(const BlockMatrixView& m, const int32_t* equationOfDof, int64_t dofCount,
                                int32_t kept, int64_t blockCount, int64_t valueCount, int32_t* hits, int32_t* flag,
                                hipStream_t stream)
{
    if (m.nodes > 0)
        hipLaunchKernelGGL(blocksValidateKernel, dim3(unsigned((m.nodes + 255) / 256)), dim3(256), 0, stream, m,
                           equationOfDof, dofCount, kept, blockCount, valueCount, hits, flag);
    if (kept > 0)
        hipLaunchKernelGGL(equationHitsKernel, dim3(unsigned((kept + 255) / 256)), dim3(256), 0, stream, hits, kept,
                           flag);
    return hipGetLastError();
}

hipError_t launchBlockTerms(const BlockMatrixView& m, const int32_t* equationOfDof, int32_t* counts,
                            unsigned long long* cursor, int32_t* rows, double* values, hipStream_t stream)
{
    if (m.nodes == 0)
        return hipSuccess;
    const dim3 grid(unsigned((m.nodes + 255) / 256));
    if (cursor == nullptr)
        hipLaunchKernelGGL(blockTermsKernel<false>, grid, dim3(256), 0, stream, m, equationOfDof, counts, nullptr,
                           nullptr, nullptr);
    else
        hipLaunchKernelGGL(blockTermsKernel<true>, grid, dim3(256), 0, stream, m, equationOfDof, nullptr, cursor,
                           rows, values);
    return hipGetLastError();
}

int64_t scanPieces(int64_t n)
{
    return (n + SCAN_ITEMS - 1) / SCAN_ITEMS;
}

hipError_t launchExclusiveScan(const int32_t* counts, int64_t n, int64_t* offsets, int64_t* pieceSums,
                               hipStream_t stream)
{
    if (n == 0)
    {
        hipLaunchKernelGGL(fillInt32Kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<int32_t*>(offsets), 2, 0);
        return hipGetLastError();
    }
    const int64_t pieces = scanPieces(n);
    hipLaunchKernelGGL(scanPieceSumsKernel, dim3(unsigned(pieces)), dim3(256), 0, stream, counts, n, pieceSums);
    hipLaunchKernelGGL(scanSumsKernel, dim3(1), dim3(256), 0, stream, pieceSums, pieces);
    hipLaunchKernelGGL(scanPiecesKernel, dim3(unsigned(pieces)), dim3(256), 0, stream, counts, n, pieceSums, offsets);
    return hipGetLastError();
}

hipError_t launchSortColumns(const int64_t* offsets, int32_t n, const int32_t* rowsIn, const double* valuesIn,
                             int32_t* rowsOut, double* valuesOut, hipStream_t stream)
{
    if (n > 0)
        hipLaunchKernelGGL(sortColumnsKernel, dim3(unsigned((n + 3) / 4)), dim3(256), 0, stream, offsets, n, rowsIn,
                           valuesIn, rowsOut, valuesOut);
    return hipGetLastError();
}

__global__ __launch_bounds__(256) void widenRangesKernel(const WidenRange* ranges, int phase)
{
    const WidenRange r = ranges[blockIdx.x];
    for (int64_t i = threadIdx.x; i < r.count; i += 256)
    {
        if (phase == 0)
            r.tmp[i] = r.src[i];
        else
            r.dst[i] = double(r.tmp[i]);
    }
}

hipError_t launchWidenRanges(const WidenRange* ranges, int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(widenRangesKernel, dim3(count), dim3(256), 0, stream, ranges, 0);
    hipLaunchKernelGGL(widenRangesKernel, dim3(count), dim3(256), 0, stream, ranges, 1);
    return hipGetLastError();
}

hipError_t launchCombineValues(int32_t n, const int64_t* offsets, const int32_t* rows, double alpha, const double* a,
                               double beta, const double* b, const double* scale, double* out, double* diagonal,
                               double* fp32Scale, hipStream_t stream)
{
    hipLaunchKernelGGL(combineValuesKernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, n,
                       offsets, rows, alpha, a, beta, b, scale, out, diagonal, fp32Scale);
    return hipGetLastError();
}

hipError_t launchRowValues(int32_t n, const int64_t* csrOffsets, const int64_t* map, const double* values,
                           double* csrValues, unsigned long long* normKey, hipStream_t stream)
{
    hipLaunchKernelGGL(fillInt32Kernel, dim3(1), dim3(64), 0, stream, reinterpret_cast<int32_t*>(normKey), 2, 0);
    hipLaunchKernelGGL(rowValuesKernel, dim3(static_cast<unsigned>((n + 255) / 256)), dim3(256), 0, stream, n,
                       csrOffsets, map, values, csrValues, normKey);
    return hipGetLastError();
}

hipError_t launchBlockGram(const GramPairs& pairs, int64_t n, int k, double* partial, double* out,
                           hipStream_t stream)
{
    hipLaunchKernelGGL(blockGramKernel, dim3(GRAM_GROUPS), dim3(256), 0, stream, pairs, n, k, partial);
    double* runs = partial + int64_t(GRAM_GROUPS) * GRAM_PAIRS * 256;
    hipLaunchKernelGGL(sumGramPartialsKernel, dim3(pairs.pairs * GRAM_RUN_COUNT), dim3(256), 0, stream, partial, runs);
    hipLaunchKernelGGL(sumGramRunsKernel, dim3(pairs.pairs), dim3(256), 0, stream, runs, out);
    return hipGetLastError();
}

hipError_t launchBlockCombine(double* y, const BlockTerm& term, int64_t n, int k, hipStream_t stream)
{
    const int64_t waves = (n + 15) / 16;
    hipLaunchKernelGGL(blockCombineKernel, dim3(static_cast<unsigned>(std::max<int64_t>(1, (waves + 3) / 4))),
                       dim3(256), 0, stream, y, term, n, k);
    return hipGetLastError();
}

hipError_t launchBlockStep(double* z, double* p, double* q, double* x, double* r, const double* coef,
                           const double* scale, int64_t n, int k, unsigned long long* norms, hipStream_t stream)
{
    const int64_t groups = ((n + 15) / 16 + 3) / 4;
    const int grid = static_cast<int>(std::max<int64_t>(1, std::min<int64_t>(groups, 2048)));
    hipLaunchKernelGGL(blockStepKernel, dim3(grid), dim3(256), 0, stream, z, p, q, x, r, coef, scale, n, k, norms);
    return hipGetLastError();
}

hipError_t launchColumnMaxRows(const double* v, int64_t n, int k, unsigned long long* out, hipStream_t stream,
                               const double* scale, int divide)
{
    hipLaunchKernelGGL(columnMaxRowsKernel, dim3(std::min(vectorGroups(n, k), 1024)), dim3(256), 0, stream, v, n, k,
                       out, scale, divide);
    return hipGetLastError();
}

hipError_t launchAddRows(double* x, const double* z, int64_t n, int k, hipStream_t stream)
{
    hipLaunchKernelGGL(addRowsKernel, dim3(vectorGroups(n, k)), dim3(256), 0, stream, x, z, n, k);
    return hipGetLastError();
}

hipError_t launchScaleRows(const double* src, double* dst, const double* scale, int64_t n, int k, hipStream_t stream)
{
    hipLaunchKernelGGL(scaleRowsKernel, dim3(vectorGroups(n, k)), dim3(256), 0, stream, src, dst, scale, n, k);
    return hipGetLastError();
}

// ---- pivot statistics -------------------------------------------------------------------------------------------

// A double as an unsigned key of the same order (negative values below positive ones).
__device__ inline unsigned long long orderedKey(double x)
{
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    return (b >> 63) ? ~b : b | 0x8000000000000000ull;
}

struct PivotTerms
{
    double pivot, ratio;
};

__device__ inline PivotTerms pivotTerms(const double* diag, const double* scale, const double* diagonal,
                                        const double* reference, int64_t j)
{
    const double pivot = scale != nullptr ? diag[j] / (scale[j] * scale[j]) : diag[j];
    const double a = reference != nullptr && reference[j] >= 0.0 ? reference[j] : diagonal[j];
    return {pivot, a / pivot};
}

// The pivot growth of each supernode: the largest ratio |a_jj / d_j| (pivotTerms) over its columns, a wave per
// supernode (0 for the supernodes whose columns are not all eliminated; infinity where a ratio is not a number).
__global__ __launch_bounds__(256) void supernodeGrowthKernel(const double* diag, const double* scale,
                                                             const double* diagonal, const int32_t* first,
                                                             const int32_t* ncol, int32_t count, int64_t eliminated,
                                                             float* out)
{
    const int32_t s = int32_t(blockIdx.x) * 4 + int32_t(threadIdx.x >> 6);
    if (s >= count)
        return; // the whole wave
    const int lane = threadIdx.x & 63;
    const int32_t f = first[s], c = ncol[s];
    float m = 0.0f;
    if (int64_t(f) + c <= eliminated)
        for (int i = lane; i < c; i += 64)
        {
            const double ratio = fabs(pivotTerms(diag, scale, diagonal, nullptr, f + i).ratio);
            const float g = isnan(ratio) ? INFINITY : float(fmin(ratio, 3.0e38));
            m = fmaxf(m, g);
        }
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1)
        m = fmaxf(m, __shfl_xor(m, offset));
    if (lane == 0)
        out[s] = m;
}

__global__ void pivotStatisticsInitKernel(PivotStatistics* out)
{
    if (threadIdx.x == 0)
        *out = PivotStatistics{0, 0, 0, 0, 0, INT32_MAX, INT32_MAX, 0, 0};
}

// Counts, the largest key and the first failing equation: per thread, then per workgroup through LDS, then one
// atomic per workgroup and quantity.
__global__ __launch_bounds__(256) void pivotStatisticsKernel(const double* diag, const double* scale,
                                                             const double* diagonal, const double* reference, int64_t n,
                                                             double maxRatio, PivotStatistics* out)
{
    __shared__ unsigned long long sums[4][256];
    __shared__ unsigned long long keys[256];
    __shared__ int32_t firsts[256];
    unsigned long long positive = 0, negative = 0, zero = 0, issues = 0, key = 0;
    int32_t first = INT32_MAX;
    for (int64_t j = blockIdx.x * int64_t(256) + threadIdx.x; j < n; j += int64_t(gridDim.x) * 256)
    {
        const PivotTerms t = pivotTerms(diag, scale, diagonal, reference, j);
        positive += t.pivot > 0.0;
        negative += t.pivot < 0.0;
        zero += !(t.pivot > 0.0) && !(t.pivot < 0.0);
        if (!isnan(t.ratio))
            key = max(key, orderedKey(t.ratio));
        if (t.ratio > maxRatio || t.ratio < 0.0 || t.pivot < 0.0)
        {
            ++issues;
            first = min(first, static_cast<int32_t>(j));
        }
    }
    const int tid = threadIdx.x;
    sums[0][tid] = positive;
    sums[1][tid] = negative;
    sums[2][tid] = zero;
    sums[3][tid] = issues;
    keys[tid] = key;
    firsts[tid] = first;
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1)
    {
        if (tid < half)
        {
            for (int q = 0; q < 4; ++q)
                sums[q][tid] += sums[q][tid + half];
            keys[tid] = max(keys[tid], keys[tid + half]);
            firsts[tid] = min(firsts[tid], firsts[tid + half]);
        }
        __syncthreads();
    }
    if (tid == 0)
    {
        atomicAdd(&out->positive, sums[0][0]);
        atomicAdd(&out->negative, sums[1][0]);
        atomicAdd(&out->zero, sums[2][0]);
        atomicAdd(&out->issues, sums[3][0]);
        atomicMax(&out->largestKey, keys[0]);
        atomicMin(&out->firstIssue, firsts[0]);
    }
}

// The first equation whose ratio has the largest key, and whether the first failing pivot fails by a negative value.
__global__ __launch_bounds__(256) void pivotLargestKernel(const double* diag, const double* scale,
                                                          const double* diagonal, const double* reference, int64_t n,
                                                          PivotStatistics* out)
{
    const unsigned long long largest = out->largestKey;
    const int32_t firstIssue = out->firstIssue;
    int32_t at = INT32_MAX;
    for (int64_t j = blockIdx.x * int64_t(256) + threadIdx.x; j < n; j += int64_t(gridDim.x) * 256)
    {
        const PivotTerms t = pivotTerms(diag, scale, diagonal, reference, j);
        if (!isnan(t.ratio) && orderedKey(t.ratio) == largest)
            at = min(at, static_cast<int32_t>(j));
        if (j == firstIssue)
            out->firstIssueNegative = t.ratio < 0.0 || t.pivot < 0.0 ? 1 : 0;
    }
    __shared__ int32_t ats[256];
    ats[threadIdx.x] = at;
    __syncthreads();
    for (int half = 128; half > 0; half >>= 1)
    {
        if (threadIdx.x < half)
            ats[threadIdx.x] = min(ats[threadIdx.x], ats[threadIdx.x + half]);
        __syncthreads();
    }
    if (threadIdx.x == 0)
        atomicMin(&out->largestAt, ats[0]);
}

hipError_t launchPivotStatistics(const double* diag, const double* scale, const double* diagonal,
                                 const double* reference, int64_t n, double maxRatio, PivotStatistics* out,
                                 hipStream_t stream)
{
    hipLaunchKernelGGL(pivotStatisticsInitKernel, dim3(1), dim3(64), 0, stream, out);
    if (n > 0)
    {
        const int groups = static_cast<int>(std::min<int64_t>(1024, (n + 255) / 256));
        hipLaunchKernelGGL(pivotStatisticsKernel, dim3(groups), dim3(256), 0, stream, diag, scale, diagonal,
                           reference, n, maxRatio, out);
        hipLaunchKernelGGL(pivotLargestKernel, dim3(groups), dim3(256), 0, stream, diag, scale, diagonal, reference,
                           n, out);
    }
    return hipGetLastError();
}

hipError_t launchSupernodeGrowth(const double* diag, const double* scale, const double* diagonal, const int32_t* first,
                                 const int32_t* ncol, int32_t count, int64_t eliminated, float* out,
                                 hipStream_t stream)
{
    if (count <= 0)
        return hipSuccess;
    hipLaunchKernelGGL(supernodeGrowthKernel, dim3((count + 3) / 4), dim3(256), 0, stream, diag, scale, diagonal, first,
                       ncol, count, eliminated, out);
    return hipGetLastError();
}

hipError_t launchMfmaProbe(const double* a, const double* b, double* d, const float* af, const float* bf, float* df,
                           hipStream_t stream)
{
    hipLaunchKernelGGL(mfmaProbeKernel, dim3(1), dim3(64), 0, stream, a, b, d, af, bf, df);
    return hipGetLastError();
}

#define OOSG_INSTANTIATE(T)                                                                                            \
    template hipError_t launchSmallFronts<T>(const DevicePlan&, const DeviceMatrix&, T*, T*, T*, const int32_t*, int,  \
                                             int, double*, int32_t*, hipStream_t);                                     \
    template hipError_t launchAssembly<T>(const DevicePlan&, const DeviceMatrix&, T*, T*, const AssemblyBlock*,        \
                                          const AssemblyItem*, int, hipStream_t);                                      \
    template hipError_t launchDiagonalBlocks<T>(const DevicePlan&, T*, T*, const int32_t*, int, int, double*,          \
                                                int32_t*, hipStream_t);                                                \
    template hipError_t launchUpdateTiles<T>(const DevicePlan&, T*, T*, T*, const UpdateTile*, int, double*, int32_t*, \
                                             bool, hipStream_t);                                                       \
    template hipError_t launchExtendTiles<T>(const DevicePlan&, T*, T*, const UpdateTile*, int, const int32_t*,        \
                                             const AssemblyItem*, int, hipStream_t);                                   \
    template hipError_t launchForward<T>(const DevicePlan&, const T*, const T*, double*, double*, int, const int32_t*, \
                                         int, const int4*, int, double*, hipStream_t, bool);                           \
    template hipError_t launchBackward<T>(const DevicePlan&, const T*, const T*, double*, int, const int32_t*, int,     \
                                          const int4*, int, double*, hipStream_t);                                     \
    template hipError_t launchMediumForward<T>(const DevicePlan&, const T*, const T*, double*, double*, int,           \
                                               const int32_t*, int, hipStream_t);                                      \
    template hipError_t launchMediumBackward<T>(const DevicePlan&, const T*, const T*, double*, int, const int32_t*,   \
                                                int, hipStream_t);                                                     \
    template hipError_t launchTinyForward<T>(const DevicePlan&, const T*, const T*, double*, double*, int,             \
                                             const int32_t*, int, hipStream_t);                                        \
    template hipError_t launchTinyBackward<T>(const DevicePlan&, const T*, const T*, double*, int, const int32_t*, int, \
                                              hipStream_t);                                                            \
    template hipError_t launchTriangleInverse<T>(const DevicePlan&, const T*, T*, const int2*, int, int, hipStream_t); \
    template hipError_t launchTriangleProducts<T>(const DevicePlan&, const T*, const double*, int, const int4*, int,   \
                                                  int, double*, hipStream_t);                                          \
    template size_t smallFrontBytes<T>(int);

OOSG_INSTANTIATE(double)
OOSG_INSTANTIATE(float)

} // namespace oosg
