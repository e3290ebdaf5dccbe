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

#include <cstdint>

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

__global__ void fillInt32Kernel(int32_t* data, int64_t count, int32_t value)
{
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < count; i += int64_t(gridDim.x) * blockDim.x)
        data[i] = value;
}

// ---- small fronts: assembly and elimination fused, the whole front in LDS ------------------------------------------

// One workgroup per front of at most SMALL_FRONT_MAX rows: the front is assembled in LDS (zero, matrix entries,
// children in order), its c columns are eliminated right-looking (two barriers per column), and the panel and the
// update matrix are written once.
template <typename T>
__global__ __launch_bounds__(256) void smallFrontKernel(DevicePlan p, DeviceMatrix a, T* panels, T* updates,
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
    // The matrix entries of the supernode's columns (distinct positions: no two threads meet).
    for (int j = 0; j < c; ++j)
    {
        const int64_t b = a.offsets[first + j], e = a.offsets[first + j + 1];
        const double sj = a.scale ? a.scale[first + j] : 1.0;
        for (int64_t k = b + tid; k < e; k += nt)
        {
            double v = a.values[k];
            if (a.scale)
                v = v * a.scale[a.rows[k]] * sj;
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
        for (int idx = tid; idx < rc * rc; idx += nt)
        {
            const int i = idx % rc, jj = idx / rc;
            if (i >= jj)
                F[rel[i] + rel[jj] * n] += u[idx];
        }
    }
    __syncthreads();

    // Right-looking elimination of the c columns.
    for (int k = 0; k < c; ++k)
    {
        const T pivot = F[k + k * n];
        if (tid == 0 && isBadPivot(pivot))
            atomicMin(&zeroAt[s], k);
        for (int i = k + 1 + tid; i < n; i += nt)
            lcol[i] = F[i + k * n] / pivot;
        __syncthreads();
        const int m = n - k - 1;
        for (int idx = tid; idx < m * m; idx += nt)
        {
            const int i = k + 1 + idx % m, j = k + 1 + idx / m;
            if (i >= j)
                F[i + j * n] -= lcol[i] * (lcol[j] * pivot);
        }
        for (int i = k + 1 + tid; i < n; i += nt)
            F[i + k * n] = lcol[i];
        __syncthreads();
    }

    T* panel = panels + p.panelOff[s];
    for (int idx = tid; idx < n * c; idx += nt)
    {
        const int i = idx % n, j = idx / n;
        panel[idx] = i < j ? T(0) : F[idx];
    }
    T* update = updates + p.updateOff[s];
    for (int idx = tid; idx < r * r; idx += nt)
    {
        const int i = idx % r, j = idx / r;
        if (i >= j)
            update[idx] = F[(c + i) + (c + j) * n];
    }
    for (int k = tid; k < c; k += nt)
        diagOut[first + k] = double(F[k + k * n]);
}

// ---- large fronts --------------------------------------------------------------------------------------------------

// Assembly of a block of a front: columns [j0, j1), rows [i0, i1), accumulated in LDS in the reference's order, then
// written to the panel (columns < c, the rows above the diagonal as zeros) or to the update matrix (lower part).
template <typename T>
__global__ __launch_bounds__(256) void assembleKernel(DevicePlan p, DeviceMatrix a, T* panels, T* updates,
                                                      const AssemblyBlock* blocks)
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
    for (int j = blk.j0; j < min(blk.j1, c); ++j)
    {
        const int64_t b = a.offsets[first + j], e = a.offsets[first + j + 1];
        const double sj = a.scale ? a.scale[first + j] : 1.0;
        for (int64_t k = b + tid; k < e; k += nt)
        {
            const int pos = a.position[k];
            if (pos < blk.i0 || pos >= blk.i1)
                continue;
            double v = a.values[k];
            if (a.scale)
                v = v * a.scale[a.rows[k]] * sj;
            B[(pos - blk.i0) + (j - blk.j0) * rowsIn] += T(v);
        }
    }
    for (int q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
    {
        __syncthreads();
        const int ch = p.children[q];
        const int rc = p.nrow[ch];
        const int32_t* rel = p.rel + p.rowStart[ch];
        const T* u = updates + p.updateOff[ch];
        const int q0 = lowerBound(rel, rc, blk.j0), q1 = lowerBound(rel, rc, blk.j1);
        const int ia = lowerBound(rel, rc, blk.i0), ib = lowerBound(rel, rc, blk.i1);
        for (int jj = q0; jj < q1; ++jj)
        {
            const int lo = max(jj, ia);
            const int jcol = rel[jj] - blk.j0;
            for (int i = lo + tid; i < ib; i += nt)
                B[(rel[i] - blk.i0) + jcol * rowsIn] += u[i + int64_t(jj) * rc];
        }
    }
    __syncthreads();
    for (int idx = tid; idx < rowsIn * colsIn; idx += nt)
    {
        const int il = idx % rowsIn, jl = idx / rowsIn;
        const int i = blk.i0 + il, j = blk.j0 + jl;
        if (j < c)
            panels[p.panelOff[s] + i + int64_t(j) * n] = B[idx];
        else if (i >= j)
            updates[p.updateOff[s] + (i - c) + int64_t(j - c) * r] = B[idx];
    }
}

// The diagonal block [k0, k0 + kb) of a front's panel factorised in place (kb <= PANEL_BLOCK), one workgroup per front.
template <typename T>
__global__ __launch_bounds__(256) void diagonalBlockKernel(DevicePlan p, T* panels, const int32_t* list, int k0,
                                                           double* diagOut, int32_t* zeroAt)
{
    __shared__ T D[PANEL_BLOCK * (PANEL_BLOCK + 1)];
    __shared__ T lcol[PANEL_BLOCK];
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], n = c + p.nrow[s];
    const int kb = min(PANEL_BLOCK, c - k0);
    const int ld = PANEL_BLOCK + 1;
    T* panel = panels + p.panelOff[s];
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < kb * kb; idx += nt)
    {
        const int i = idx % kb, j = idx / kb;
        D[i + j * ld] = panel[(k0 + i) + int64_t(k0 + j) * n];
    }
    __syncthreads();
    for (int k = 0; k < kb; ++k)
    {
        const T pivot = D[k + k * ld];
        if (tid == 0 && isBadPivot(pivot))
            atomicMin(&zeroAt[s], k0 + k);
        for (int i = k + 1 + tid; i < kb; i += nt)
            lcol[i] = D[i + k * ld] / pivot;
        __syncthreads();
        const int m = kb - k - 1;
        for (int idx = tid; idx < m * m; idx += nt)
        {
            const int i = k + 1 + idx % m, j = k + 1 + idx / m;
            if (i >= j)
                D[i + j * ld] -= lcol[i] * (lcol[j] * pivot);
        }
        for (int i = k + 1 + tid; i < kb; i += nt)
            D[i + k * ld] = lcol[i];
        __syncthreads();
    }
    for (int idx = tid; idx < kb * kb; idx += nt)
    {
        const int i = idx % kb, j = idx / kb;
        if (i >= j)
            panel[(k0 + i) + int64_t(k0 + j) * n] = D[i + j * ld];
    }
    for (int k = tid; k < kb; k += nt)
        diagOut[p.first[s] + k0 + k] = double(D[k + k * ld]);
}

// The rows below a factorised diagonal block: L(i, blk) = F(i, blk)·L11⁻ᵀ·D⁻¹, one row per thread held in registers
// (PANEL_BLOCK values), L11 and D broadcast from LDS. A chunk is (front, first row); rows [row0, row0 + 256) ∩ [k0 + kb, n).
template <typename T>
__global__ __launch_bounds__(256) void panelSolveKernel(DevicePlan p, T* panels, const int2* chunks, int k0)
{
    __shared__ T L[PANEL_BLOCK * PANEL_BLOCK];
    __shared__ T d[PANEL_BLOCK];
    const int2 chunk = chunks[blockIdx.x];
    const int s = chunk.x;
    const int c = p.ncol[s], n = c + p.nrow[s];
    const int kb = min(PANEL_BLOCK, c - k0);
    T* panel = panels + p.panelOff[s];
    const int tid = threadIdx.x;
    for (int idx = tid; idx < PANEL_BLOCK * PANEL_BLOCK; idx += blockDim.x)
    {
        const int i = idx % PANEL_BLOCK, j = idx / PANEL_BLOCK;
        // L stored row-major in LDS: L[i * B + j] = L11(i, j), identity outside the block.
        T v = T(i == j ? 1 : 0);
        if (i < kb && j < kb && i > j)
            v = panel[(k0 + i) + int64_t(k0 + j) * n];
        L[i * PANEL_BLOCK + j] = v;
    }
    for (int k = tid; k < PANEL_BLOCK; k += blockDim.x)
        d[k] = k < kb ? panel[(k0 + k) + int64_t(k0 + k) * n] : T(1);
    __syncthreads();
    const int i = chunk.y + tid;
    if (i >= n)
        return;
    T x[PANEL_BLOCK];
#pragma unroll
    for (int k = 0; k < PANEL_BLOCK; ++k)
        x[k] = k < kb ? panel[i + int64_t(k0 + k) * n] : T(0);
    // z·L11ᵀ = x, z = L(i, blk)·D.
#pragma unroll
    for (int k = 0; k < PANEL_BLOCK; ++k)
    {
        T z = x[k];
#pragma unroll
        for (int q = 0; q < k; ++q)
            z -= x[q] * L[k * PANEL_BLOCK + q];
        x[k] = z;
    }
#pragma unroll
    for (int k = 0; k < PANEL_BLOCK; ++k)
        if (k < kb)
            panel[i + int64_t(k0 + k) * n] = x[k] / d[k];
}

// MFMA tiles of C −= A·diag(d)·Bᵀ (64 × 64 per workgroup, four waves of 32 × 32, 16 × 16 × 4 MFMA blocks): the
// trailing update of a panel step and the update matrix of a front. A(i, p) = panels[aOff + i + p·ld],
// B(j, p) = panels[bOff + j + p·ld], d_p = panels[dOff + p·(ld + 1)].
template <typename T> struct Mfma;

template <> struct Mfma<double>
{
    typedef double4_t Acc;
    __device__ static inline Acc step(double a, double b, Acc c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <> struct Mfma<float>
{
    typedef float4_t Acc;
    __device__ static inline Acc step(float a, float b, Acc c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};

constexpr int TILE = 64;
constexpr int TILE_K = 16;

template <typename T>
__global__ __launch_bounds__(256) void updateTilesKernel(T* panels, T* updates, const UpdateTile* tiles)
{
    __shared__ T As[TILE_K][TILE + 4];
    __shared__ T Bs[TILE_K][TILE + 4];
    const UpdateTile t = tiles[blockIdx.x];
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const int wm = (wave >> 1) * 32, wn = (wave & 1) * 32;
    typename Mfma<T>::Acc acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
            acc[a][b] = typename Mfma<T>::Acc{0, 0, 0, 0};
    const T* A = panels + t.aOff;
    const T* Bm = panels + t.bOff;
    const T* dv = panels + t.dOff;
    const int64_t ld = t.ld;
    for (int p0 = 0; p0 < t.k; p0 += TILE_K)
    {
        // Stage A and W = B·D: 64 rows × 16 terms each, 4 elements per thread, coalesced along the rows.
#pragma unroll
        for (int q = 0; q < (TILE * TILE_K) / 256; ++q)
        {
            const int idx = tid + q * 256;
            const int row = idx & (TILE - 1), kk = idx >> 6;
            const int pp = p0 + kk;
            T av = T(0), bv = T(0);
            if (pp < t.k)
            {
                if (row < t.m)
                    av = A[row + pp * ld];
                if (row < t.n)
                    bv = Bm[row + pp * ld] * dv[pp * (ld + 1)];
            }
            As[kk][row] = av;
            Bs[kk][row] = bv;
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < TILE_K; kk += 4)
        {
            const int ka = kk + (lane >> 4);
#pragma unroll
            for (int a = 0; a < 2; ++a)
            {
                const T av = As[ka][wm + a * 16 + (lane & 15)];
#pragma unroll
                for (int b = 0; b < 2; ++b)
                {
                    const T bv = Bs[ka][wn + b * 16 + (lane & 15)];
                    acc[a][b] = Mfma<T>::step(av, bv, acc[a][b]);
                }
            }
        }
        __syncthreads();
    }
    T* C = (t.target == 0 ? panels : updates) + t.cOff;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int v = 0; v < 4; ++v)
            {
                const int i = wm + a * 16 + 4 * (lane >> 4) + v;
                const int j = wn + b * 16 + (lane & 15);
                if (i < t.m && j < t.n && (!t.lower || i >= j))
                    C[i + int64_t(j) * t.ldc] -= acc[a][b][v];
            }
}

// ---- solves --------------------------------------------------------------------------------------------------------

constexpr int SOLVE_BLOCK = 64;

// Forward sweep of a level, one workgroup per supernode, K right-hand sides at most (k actual): y = [X(cols); 0] plus
// the children's update vectors in child order, y ← L⁻¹·y on the front (unit lower), X(cols) = D⁻¹·y(cols),
// U(s) = y(rows) for the parent.
template <typename TP, int K>
__global__ __launch_bounds__(256) void forwardKernel(DevicePlan p, const TP* panels, double* X, double* U, int k,
                                                     const int32_t* list)
{
    __shared__ double ys[SOLVE_BLOCK][K];
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    const TP* L = panels + p.panelOff[s];
    double* u = U + p.rowStart[s] * k;
    const int tid = threadIdx.x, nt = blockDim.x;
    for (int idx = tid; idx < r * k; idx += nt)
        u[idx] = 0.0;
    for (int q = p.childStart[s]; q < p.childStart[s + 1]; ++q)
    {
        __syncthreads();
        const int ch = p.children[q];
        const int rc = p.nrow[ch];
        const int32_t* rel = p.rel + p.rowStart[ch];
        const double* uc = U + p.rowStart[ch] * k;
        for (int idx = tid; idx < rc * k; idx += nt)
        {
            const int i = idx / k, h = idx - i * k;
            const int pos = rel[i];
            if (pos < c)
                X[int64_t(first + pos) * k + h] += uc[idx];
            else
                u[int64_t(pos - c) * k + h] += uc[idx];
        }
    }
    __syncthreads();
    for (int q0 = 0; q0 < c; q0 += SOLVE_BLOCK)
    {
        const int qb = min(SOLVE_BLOCK, c - q0);
        for (int idx = tid; idx < qb * k; idx += nt)
        {
            const int i = idx / k, h = idx - i * k;
            ys[i][h] = X[int64_t(first + q0 + i) * k + h];
        }
        __syncthreads();
        // In-block triangle: the columns one after the other.
        for (int q = 0; q < qb; ++q)
        {
            for (int idx = tid; idx < (qb - q - 1) * k; idx += nt)
            {
                const int i = q + 1 + idx / k, h = idx % k;
                ys[i][h] -= double(L[(q0 + i) + int64_t(q0 + q) * n]) * ys[q][h];
            }
            __syncthreads();
        }
        for (int idx = tid; idx < qb * k; idx += nt)
        {
            const int i = idx / k, h = idx - i * k;
            X[int64_t(first + q0 + i) * k + h] = ys[i][h];
        }
        // The rows below the block: y(i) −= Σ_q L(i, q)·y(q), one row per thread, K sums in registers.
        for (int i = q0 + qb + tid; i < n; i += nt)
        {
            double acc[K];
#pragma unroll
            for (int h = 0; h < K; ++h)
                acc[h] = 0.0;
            for (int q = 0; q < qb; ++q)
            {
                const double l = double(L[i + int64_t(q0 + q) * n]);
#pragma unroll
                for (int h = 0; h < K; ++h)
                    acc[h] += l * ys[q][h];
            }
            double* y = i < c ? X + int64_t(first + i) * k : u + int64_t(i - c) * k;
#pragma unroll
            for (int h = 0; h < K; ++h)
                if (h < k)
                    y[h] -= acc[h];
        }
        __syncthreads();
    }
    for (int idx = tid; idx < c * k; idx += nt)
    {
        const int i = idx / k;
        X[int64_t(first) * k + idx] /= double(L[i + int64_t(i) * n]);
    }
}

// Backward sweep of a level (top down), one workgroup per supernode: X(cols) −= L_RFᵀ·X(rows), then
// X(cols) ← L_FF⁻ᵀ·X(cols), by blocks of columns from the last.
template <typename TP, int K>
__global__ __launch_bounds__(256) void backwardKernel(DevicePlan p, const TP* panels, double* X, int k,
                                                      const int32_t* list)
{
    __shared__ double ys[SOLVE_BLOCK][K];
    const int s = list[blockIdx.x];
    const int c = p.ncol[s], r = p.nrow[s], n = c + r, first = p.first[s];
    const TP* L = panels + p.panelOff[s];
    const int32_t* rows = p.rows + p.rowStart[s];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, waves = blockDim.x >> 6;
    for (int q0 = ((c - 1) / SOLVE_BLOCK) * SOLVE_BLOCK; q0 >= 0; q0 -= SOLVE_BLOCK)
    {
        const int qb = min(SOLVE_BLOCK, c - q0);
        // Dot products of the block's columns with the final rows below the block, one column per wave at a time.
        for (int q = wave; q < qb; q += waves)
        {
            double acc[K];
#pragma unroll
            for (int h = 0; h < K; ++h)
                acc[h] = 0.0;
            const TP* col = L + int64_t(q0 + q) * n;
            for (int i = q0 + qb + lane; i < n; i += 64)
            {
                const double l = double(col[i]);
                const double* x = i < c ? X + int64_t(first + i) * k : X + int64_t(rows[i - c]) * k;
#pragma unroll
                for (int h = 0; h < K; ++h)
                    if (h < k)
                        acc[h] += l * x[h];
            }
#pragma unroll
            for (int h = 0; h < K; ++h)
            {
                double v = acc[h];
#pragma unroll
                for (int off = 32; off > 0; off >>= 1)
                    v += __shfl_xor(v, off);
                acc[h] = v;
            }
            if (lane == 0)
#pragma unroll
                for (int h = 0; h < K; ++h)
                    if (h < k)
                        ys[q][h] = X[int64_t(first + q0 + q) * k + h] - acc[h];
        }
        __syncthreads();
        // In-block triangle, from the last column up.
        for (int q = qb - 1; q >= 0; --q)
        {
            for (int idx = tid; idx < q * k; idx += blockDim.x)
            {
                const int i = idx / k, h = idx % k;
                ys[i][h] -= double(L[(q0 + q) + int64_t(q0 + i) * n]) * ys[q][h];
            }
            __syncthreads();
        }
        for (int idx = tid; idx < qb * k; idx += blockDim.x)
        {
            const int i = idx / k, h = idx - i * k;
            X[int64_t(first + q0 + i) * k + h] = ys[i][h];
        }
        __syncthreads();
    }
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

// Per column the largest |v| (as the bits of a non-negative double, whose order is the order of the values).
__global__ void columnMaxKernel(const double* v, int64_t ld, int64_t n, int k, unsigned long long* out)
{
    const int h = blockIdx.y;
    double m = 0.0;
    for (int64_t i = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
    {
        const double a = fabs(v[i + h * ld]);
        m = a > m || a != a ? a : m;
    }
    for (int off = 32; off > 0; off >>= 1)
    {
        const double o = __shfl_xor(m, off);
        m = o > m || o != o ? o : m;
    }
    if ((threadIdx.x & 63) == 0)
        atomicMax(&out[h], static_cast<unsigned long long>(__double_as_longlong(m)));
}

__global__ void fillDoubleKernel(double* data, int64_t ld, int64_t n, int k, double value)
{
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n * k;
         idx += int64_t(gridDim.x) * blockDim.x)
        data[(idx % n) + (idx / n) * ld] = value;
}

__global__ void axpyColumnsKernel(double* x, const double* d, int64_t ld, int64_t n, int k)
{
    for (int64_t idx = blockIdx.x * int64_t(blockDim.x) + threadIdx.x; idx < n * k;
         idx += int64_t(gridDim.x) * blockDim.x)
    {
        const int64_t i = idx % n;
        const int h = int(idx / n);
        x[i + h * ld] += d[i + h * ld];
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
        d[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = acc[v];
        df[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = accf[v];
    }
}

// ---- launchers -----------------------------------------------------------------------------------------------------

namespace
{
int blocksFor(int64_t work, int threads)
{
    const int64_t b = (work + threads - 1) / threads;
    return int(b < 1 ? 1 : (b > 65536 ? 65536 : b));
}
} // namespace

template <typename T> size_t smallFrontBytes(int maxOrder)
{
    return size_t(maxOrder) * maxOrder * sizeof(T) + size_t(maxOrder) * sizeof(T);
}

template <typename T>
hipError_t launchSmallFronts(const DevicePlan& p, const DeviceMatrix& a, T* panels, T* updates, const int32_t* list,
                             int count, int maxOrder, double* diagOut, int32_t* zeroAt, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    const size_t bytes = smallFrontBytes<T>(maxOrder);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&smallFrontKernel<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(smallFrontKernel<T>, dim3(count), dim3(256), bytes, stream, p, a, panels, updates, list, diagOut,
                       zeroAt);
    return hipGetLastError();
}

template <typename T>
hipError_t launchAssembly(const DevicePlan& p, const DeviceMatrix& a, T* panels, T* updates,
                          const AssemblyBlock* blocks, int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    const size_t bytes = size_t(ASSEMBLY_BLOCK_VALUES) * sizeof(T);
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&assembleKernel<T>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, int(bytes));
    if (e != hipSuccess)
        return e;
    hipLaunchKernelGGL(assembleKernel<T>, dim3(count), dim3(256), bytes, stream, p, a, panels, updates, blocks);
    return hipGetLastError();
}

template <typename T>
hipError_t launchDiagonalBlocks(const DevicePlan& p, T* panels, const int32_t* fronts, int frontCount, int k0,
                                double* diagOut, int32_t* zeroAt, hipStream_t stream)
{
    if (frontCount == 0)
        return hipSuccess;
    hipLaunchKernelGGL(diagonalBlockKernel<T>, dim3(frontCount), dim3(256), 0, stream, p, panels, fronts, k0, diagOut,
                       zeroAt);
    return hipGetLastError();
}

template <typename T>
hipError_t launchPanelSolve(const DevicePlan& p, T* panels, const int2* chunks, int chunkCount, int k0,
                            hipStream_t stream)
{
    if (chunkCount == 0)
        return hipSuccess;
    hipLaunchKernelGGL(panelSolveKernel<T>, dim3(chunkCount), dim3(256), 0, stream, p, panels, chunks, k0);
    return hipGetLastError();
}

template <typename T>
hipError_t launchUpdateTiles(T* panels, T* updates, const UpdateTile* tiles, int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(updateTilesKernel<T>, dim3(count), dim3(256), 0, stream, panels, updates, tiles);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchForward(const DevicePlan& p, const TP* panels, double* X, double* U, int k, const int32_t* list,
                         int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    if (k == 1)
        hipLaunchKernelGGL((forwardKernel<TP, 1>), dim3(count), dim3(256), 0, stream, p, panels, X, U, k, list);
    else if (k <= 4)
        hipLaunchKernelGGL((forwardKernel<TP, 4>), dim3(count), dim3(256), 0, stream, p, panels, X, U, k, list);
    else
        hipLaunchKernelGGL((forwardKernel<TP, 16>), dim3(count), dim3(256), 0, stream, p, panels, X, U, k, list);
    return hipGetLastError();
}

template <typename TP>
hipError_t launchBackward(const DevicePlan& p, const TP* panels, double* X, int k, const int32_t* list, int count,
                          hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    if (k == 1)
        hipLaunchKernelGGL((backwardKernel<TP, 1>), dim3(count), dim3(256), 0, stream, p, panels, X, k, list);
    else if (k <= 4)
        hipLaunchKernelGGL((backwardKernel<TP, 4>), dim3(count), dim3(256), 0, stream, p, panels, X, k, list);
    else
        hipLaunchKernelGGL((backwardKernel<TP, 16>), dim3(count), dim3(256), 0, stream, p, panels, X, k, list);
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
                              int accumulate, hipStream_t stream)
{
    hipLaunchKernelGGL(fromRowMajorKernel, dim3(blocksFor(n * k, 256)), dim3(256), 0, stream, src, dst, ld, n, k,
                       scale, accumulate);
    return hipGetLastError();
}

hipError_t launchResidual(const int64_t* offsets, const int32_t* cols, const double* values, const double* x,
                          const double* b, double* out, int64_t ld, int64_t n, int k, hipStream_t stream)
{
    hipLaunchKernelGGL(residualKernel, dim3(blocksFor(n * k, 256)), dim3(256), 0, stream, offsets, cols, values, x, b,
                       out, ld, n, k);
    return hipGetLastError();
}

hipError_t launchColumnMax(const double* v, int64_t ld, int64_t n, int k, unsigned long long* out, hipStream_t stream)
{
    const int blocks = blocksFor(n, 256) > 1024 ? 1024 : blocksFor(n, 256);
    hipLaunchKernelGGL(columnMaxKernel, dim3(blocks, k), dim3(256), 0, stream, v, ld, n, k, out);
    return hipGetLastError();
}

hipError_t launchFillDouble(double* data, int64_t ld, int64_t n, int k, double value, hipStream_t stream)
{
    hipLaunchKernelGGL(fillDoubleKernel, dim3(blocksFor(n * k, 256)), dim3(256), 0, stream, data, ld, n, k, value);
    return hipGetLastError();
}

hipError_t launchAxpyColumns(double* x, const double* d, int64_t ld, int64_t n, int k, hipStream_t stream)
{
    hipLaunchKernelGGL(axpyColumnsKernel, dim3(blocksFor(n * k, 256)), dim3(256), 0, stream, x, d, ld, n, k);
    return hipGetLastError();
}

hipError_t launchMfmaProbe(const double* a, const double* b, double* d, const float* af, const float* bf, float* df,
                           hipStream_t stream)
{
    hipLaunchKernelGGL(mfmaProbeKernel, dim3(1), dim3(64), 0, stream, a, b, d, af, bf, df);
    return hipGetLastError();
}

#define OOSG_INSTANTIATE(T)                                                                                            \
    template hipError_t launchSmallFronts<T>(const DevicePlan&, const DeviceMatrix&, T*, T*, const int32_t*, int, int, \
                                             double*, int32_t*, hipStream_t);                                          \
    template hipError_t launchAssembly<T>(const DevicePlan&, const DeviceMatrix&, T*, T*, const AssemblyBlock*, int,   \
                                          hipStream_t);                                                                \
    template hipError_t launchDiagonalBlocks<T>(const DevicePlan&, T*, const int32_t*, int, int, double*, int32_t*,    \
                                                hipStream_t);                                                          \
    template hipError_t launchPanelSolve<T>(const DevicePlan&, T*, const int2*, int, int, hipStream_t);               \
    template hipError_t launchUpdateTiles<T>(T*, T*, const UpdateTile*, int, hipStream_t);                             \
    template hipError_t launchForward<T>(const DevicePlan&, const T*, double*, double*, int, const int32_t*, int,      \
                                         hipStream_t);                                                                 \
    template hipError_t launchBackward<T>(const DevicePlan&, const T*, double*, int, const int32_t*, int, hipStream_t); \
    template size_t smallFrontBytes<T>(int);

OOSG_INSTANTIATE(double)
OOSG_INSTANTIATE(float)

} // namespace oosg
