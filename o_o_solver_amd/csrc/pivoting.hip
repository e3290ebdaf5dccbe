// The indefinite factorisation with Bunch–Kaufman pivoting: the kernels of the pivoted fronts, and the solve-side
// kernels of a pivoted factor (unit lower block inverses, the 2 × 2 block diagonal, the pivot permutation).
//
// A pivoted front is a full N × N column-major array (leading dimension N, lower triangle used) whose positions are
// [the supernode's own c columns, the d columns its children delayed, its r rows]; the first C = c + d are the
// candidates. Its elimination runs in blocks of columns on one workgroup per front (bunchKaufmanKernel), each column
// computed left-looking from the block's columns (W = the updated columns, L·D), then the columns of the candidates
// to the right get C −= L·Wᵀ by MFMA tiles (updateTilesKernel<double, true>) and, after the last block, the rows
// below the candidates likewise with K = e (the eliminated count). The pivot tests are those of
// src/linalg/kernels/reference/DenseBlockKernels.cpp:448-580: α = (1 + √17)/8, and a candidate whose partner row is
// not a candidate is pivoted alone when |a_kk| >= 0.01·colmax and delayed otherwise (moved behind the candidates still
// to try); passes over the delayed candidates repeat while a pass eliminates something.

#include <hip/hip_runtime.h>

#include <cstdint>

#include "kernels.h"

namespace oosg
{

namespace
{

constexpr double BK_ALPHA = 0.6403882032022076; // (1 + √17) / 8
constexpr double BK_DELAY = 0.01;                // DenseKernelEntryPoints.hpp:241
constexpr int BK_THREADS = 256;

struct MaxAt
{
    double v;
    int i;
};

// The larger value, the smaller index on a tie: a scan in increasing index order with a strict test keeps the first.
__device__ inline MaxAt larger(MaxAt a, MaxAt b)
{
    return (b.v > a.v || (b.v == a.v && b.i < a.i)) ? b : a;
}

__device__ inline MaxAt blockMax(MaxAt m, double* rv, int* ri)
{
    const int tid = threadIdx.x;
    rv[tid] = m.v;
    ri[tid] = m.i;
    __syncthreads();
    for (int half = BK_THREADS / 2; half > 0; half >>= 1)
    {
        if (tid < half)
        {
            const MaxAt o = larger(MaxAt{rv[tid], ri[tid]}, MaxAt{rv[tid + half], ri[tid + half]});
            rv[tid] = o.v;
            ri[tid] = o.i;
        }
        __syncthreads();
    }
    const MaxAt out{rv[0], ri[0]};
    __syncthreads();
    return out;
}

__device__ inline void swapValues(double& a, double& b)
{
    const double t = a;
    a = b;
    b = t;
}

// W(i, dst) = A(i, col) − Σ_{t ∈ [k0, k)} L(i, t)·W(col, t) for the rows i in [k, N): the column col of the front
// updated by the block's eliminated columns (those before k0 are in the stored values already). A(i, col) is read
// from the lower triangle: column col below its diagonal, row col before it.
__device__ inline void updatedColumn(const double* F, double* W, int N, int k0, int k, int col, int dst, double* wr)
{
    const int tid = threadIdx.x, terms = k - k0;
    for (int t = tid; t < terms; t += BK_THREADS)
        wr[t] = W[col + int64_t(k0 + t) * N];
    __syncthreads();
    for (int i = k + tid; i < N; i += BK_THREADS)
    {
        double a = i >= col ? F[i + int64_t(col) * N] : F[col + int64_t(i) * N];
        for (int t = 0; t < terms; ++t)
            a -= F[i + int64_t(k0 + t) * N] * wr[t];
        W[i + int64_t(dst) * N] = a;
    }
    __syncthreads();
}

// The symmetric interchange of positions p < q of the front (DenseBlockKernels.cpp:69-79), over the stored values:
// rows p and q of the columns before p (the eliminated columns' L, and the columns still to come), the parts between,
// below and on the diagonal; with rows p and q of the block's W columns and the permutation.
__device__ inline void interchange(double* F, double* W, int32_t* perm, int N, int k0, int k, int p, int q)
{
    const int tid = threadIdx.x;
    for (int t = tid; t < p; t += BK_THREADS)
        swapValues(F[p + int64_t(t) * N], F[q + int64_t(t) * N]);
    for (int i = p + 1 + tid; i < q; i += BK_THREADS)
        swapValues(F[i + int64_t(p) * N], F[q + int64_t(i) * N]);
    for (int i = q + 1 + tid; i < N; i += BK_THREADS)
        swapValues(F[i + int64_t(p) * N], F[i + int64_t(q) * N]);
    for (int t = k0 + tid; t < k; t += BK_THREADS)
        swapValues(W[p + int64_t(t) * N], W[q + int64_t(t) * N]);
    if (tid == 0)
    {
        swapValues(F[p + int64_t(p) * N], F[q + int64_t(q) * N]);
        const int32_t e = perm[p];
        perm[p] = perm[q];
        perm[q] = e;
    }
    __syncthreads();
}

// One block step of the Bunch–Kaufman elimination of each front of the list (one workgroup per front): from the
// front's state, columns are eliminated until `nb` of them are (one more when the last pivot is 2 × 2) or the
// candidates are done. The state records where the next step starts.
__global__ __launch_bounds__(BK_THREADS) void bunchKaufmanKernel(const PivotFront* fronts, PivotState* states,
                                                                 const int32_t* list, int nb)
{
    __shared__ double wr[2 * PANEL_BLOCK + 8];
    __shared__ double rv[BK_THREADS];
    __shared__ int ri[BK_THREADS];
    const int q = list[blockIdx.x];
    const PivotFront fr = fronts[q];
    const PivotState st = states[q];
    if (st.done)
        return;
    double* F = fr.F;
    double* W = fr.W;
    int32_t* perm = fr.perm;
    const int N = fr.N, C = fr.C, tid = threadIdx.x;
    const int k0 = st.k;
    int k = st.k, end = st.end, passStart = st.passStart, twoByTwo = st.twoByTwo, done = 0;
    for (;;)
    {
        if (k == end)
        {
            if (end == C || k == passStart)
            {
                done = 1;
                break;
            }
            passStart = k; // another pass over the delayed candidates
            end = C;
            continue;
        }
        if (k - k0 >= nb)
            break;
        updatedColumn(F, W, N, k0, k, k, k, wr);
        const double absakk = fabs(W[k + int64_t(k) * N]);
        MaxAt m{0.0, k};
        for (int i = k + 1 + tid; i < N; i += BK_THREADS)
        {
            const double v = fabs(W[i + int64_t(k) * N]);
            if (v > m.v)
                m = MaxAt{v, i};
        }
        m = blockMax(m, rv, ri);
        const double colmax = m.v;
        const int imax = m.i;
        if (absakk == 0.0 && colmax == 0.0)
        {
            // A zero column: a zero pivot, nothing to eliminate with it.
            for (int i = k + 1 + tid; i < N; i += BK_THREADS)
                F[i + int64_t(k) * N] = 0.0;
            if (tid == 0)
            {
                F[k + int64_t(k) * N] = 0.0;
                fr.diag[k] = 0.0;
                fr.sub[k] = 0.0;
            }
            __syncthreads();
            k += 1;
            continue;
        }
        int blockSize = 1, swapWith = k;
        if (absakk < BK_ALPHA * colmax && imax >= end)
        {
            if (absakk < BK_DELAY * colmax)
            {
                // The partner is not a candidate and the diagonal is too small: delay the column.
                --end;
                if (k != end)
                    interchange(F, W, perm, N, k0, k, k, end);
                continue;
            }
        }
        else if (absakk < BK_ALPHA * colmax)
        {
            // Column imax, updated, in W's column k + 1 (imax < end <= C, so k + 1 < C).
            updatedColumn(F, W, N, k0, k, imax, k + 1, wr);
            double rowmax = 0.0;
            for (int i = k + tid; i < N; i += BK_THREADS)
                if (i != imax)
                    rowmax = fmax(rowmax, fabs(W[i + int64_t(k + 1) * N]));
            rowmax = blockMax(MaxAt{rowmax, 0}, rv, ri).v;
            if (absakk >= BK_ALPHA * colmax * (colmax / rowmax))
                swapWith = k;
            else if (fabs(W[imax + int64_t(k + 1) * N]) >= BK_ALPHA * rowmax)
                swapWith = imax;
            else
            {
                swapWith = imax;
                blockSize = 2;
            }
        }
        const int moved = k + blockSize - 1;
        if (swapWith != moved)
        {
            interchange(F, W, perm, N, k0, k, moved, swapWith);
            if (blockSize == 1)
            {
                // The pivot column is column imax, its rows k and imax exchanged.
                for (int i = k + tid; i < N; i += BK_THREADS)
                {
                    const int from = i == k ? swapWith : (i == swapWith ? k : i);
                    W[i + int64_t(k) * N] = W[from + int64_t(k + 1) * N];
                }
            }
            else if (tid == 0)
            {
                swapValues(W[(k + 1) + int64_t(k) * N], W[swapWith + int64_t(k) * N]);
                swapValues(W[(k + 1) + int64_t(k + 1) * N], W[swapWith + int64_t(k + 1) * N]);
            }
            __syncthreads();
        }
        if (blockSize == 1)
        {
            const double d = W[k + int64_t(k) * N];
            for (int i = k + 1 + tid; i < N; i += BK_THREADS)
                F[i + int64_t(k) * N] = W[i + int64_t(k) * N] / d;
            if (tid == 0)
            {
                F[k + int64_t(k) * N] = d;
                fr.diag[k] = d;
                fr.sub[k] = 0.0;
            }
        }
        else
        {
            // LdltArithmetic's invertTwoByTwo: d11 = c/b, d22 = a/b, t = 1/(d11·d22 − 1), s = t/b.
            const double d0 = W[k + int64_t(k) * N], e = W[(k + 1) + int64_t(k) * N];
            const double d1 = W[(k + 1) + int64_t(k + 1) * N];
            const double d11 = d1 / e, d22 = d0 / e;
            const double t = 1.0 / (d11 * d22 - 1.0);
            const double s = t / e;
            for (int i = k + 2 + tid; i < N; i += BK_THREADS)
            {
                const double r0 = W[i + int64_t(k) * N], r1 = W[i + int64_t(k + 1) * N];
                F[i + int64_t(k) * N] = s * (d11 * r0 - r1);
                F[i + int64_t(k + 1) * N] = s * (d22 * r1 - r0);
            }
            if (tid == 0)
            {
                F[k + int64_t(k) * N] = d0;
                F[(k + 1) + int64_t(k) * N] = 0.0;
                F[(k + 1) + int64_t(k + 1) * N] = d1;
                fr.diag[k] = d0;
                fr.diag[k + 1] = d1;
                fr.sub[k] = e;
                fr.sub[k + 1] = 0.0;
            }
            ++twoByTwo;
        }
        __syncthreads();
        k += blockSize;
    }
    if (tid == 0)
        states[q] = PivotState{k, end, passStart, done, twoByTwo, 0, 0, 0};
}

// The matrix entries of each front's own columns, at their positions (rows past the own columns move down by the
// delayed count). The front is zero before.
__global__ __launch_bounds__(256) void pivotMatrixKernel(DeviceMatrix a, const PivotFront* fronts)
{
    const PivotFront fr = fronts[blockIdx.x];
    const int64_t kb = a.offsets[fr.first], ke = a.offsets[fr.first + fr.c];
    for (int64_t q = kb + threadIdx.x; q < ke; q += blockDim.x)
    {
        const int pos = a.position[q], j = a.column[q];
        const int at = pos < fr.c ? pos : pos + fr.d;
        fr.F[at + int64_t(j) * fr.N] += a.values[q];
    }
}

// The extend-add of one child's update matrix (columns [j0, j1) of it) into its parent's front: entry (i, j), i >= j,
// goes to the parent's positions (map[i], map[j]) ordered into the lower triangle. A child's positions are distinct,
// so no two threads meet; children of one parent go in separate launches, in order.
__global__ __launch_bounds__(256) void pivotExtendAddKernel(const ExtendItem* items)
{
    const ExtendItem it = items[blockIdx.x];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int j = it.j0 + wave; j < it.j1; j += 4)
    {
        const int pj = it.map[j];
        for (int i = j + lane; i < it.M; i += 64)
        {
            const int pi = it.map[i];
            const int r = max(pi, pj), c = min(pi, pj);
            it.dst[r + int64_t(c) * it.ldd] += it.src[i + int64_t(j) * it.lds];
        }
    }
}

// Contiguous copies (the eliminated columns of the fronts into the panels).
__global__ __launch_bounds__(256) void copyRangesKernel(const CopyRange* ranges)
{
    const CopyRange r = ranges[blockIdx.x];
    for (int64_t q = threadIdx.x; q < r.count; q += blockDim.x)
        r.dst[q] = r.src[q];
}

// The inverse of a 64-column diagonal block (s, J) of a unit lower triangle L_FF (the strictly lower values of the
// panel; the diagonal taken as 1): kb × kb column-major at linvOff[s] + 64·64·J, leading dimension kb. Column j of
// the inverse by forward substitution on one lane each.
__global__ __launch_bounds__(64) void unitBlockInverseKernel(DevicePlan p, const double* panels, double* linv,
                                                             const int2* items)
{
    __shared__ double Ls[PANEL_BLOCK][PANEL_BLOCK + 1];
    __shared__ double Xs[PANEL_BLOCK][PANEL_BLOCK + 1];
    const int2 item = items[blockIdx.x];
    const int s = item.x, q0 = item.y * PANEL_BLOCK;
    const int c = p.ncol[s], n = c + p.nrow[s];
    const int kb = min(PANEL_BLOCK, c - q0);
    const double* L = panels + p.panelOff[s] + q0 + int64_t(q0) * n;
    const int lane = threadIdx.x;
    for (int j = 0; j < kb; ++j)
        if (lane < kb)
            Ls[lane][j] = lane > j ? L[lane + int64_t(j) * n] : 0.0;
    __syncthreads();
    if (lane < kb)
    {
        const int j = lane;
        for (int i = 0; i < kb; ++i)
            Xs[i][j] = i == j ? 1.0 : 0.0;
        for (int i = j + 1; i < kb; ++i)
        {
            double x = 0.0;
            for (int t = j; t < i; ++t)
                x -= Ls[i][t] * Xs[t][j];
            Xs[i][j] = x;
        }
    }
    __syncthreads();
    double* out = linv + p.linvOff[s] + int64_t(q0) * PANEL_BLOCK;
    for (int j = 0; j < kb; ++j)
        if (lane < kb)
            out[lane + j * kb] = Xs[lane][j];
}

// X ← D⁻¹·X for a block diagonal D (X row-major n × k): 1 × 1 pivots divide; a 2 × 2 block (sub[i] != 0, rows i and
// i + 1) is solved by LdltArithmetic's inverse, x0 = s·(d11·b0 − b1), x1 = s·(d22·b1 − b0).
__global__ __launch_bounds__(256) void blockDiagonalSolveKernel(double* X, const double* d, const double* sub,
                                                                int64_t n, int k)
{
    const int h = threadIdx.x & 15;
    const int64_t i = blockIdx.x * int64_t(blockDim.x >> 4) + (threadIdx.x >> 4);
    if (h >= k || i >= n)
        return;
    if (i > 0 && sub[i - 1] != 0.0)
        return; // the second row of a block, solved with the first
    if (sub[i] != 0.0 && i + 1 < n)
    {
        const double a = d[i], e = sub[i], c = d[i + 1];
        const double d11 = c / e, d22 = a / e;
        const double t = 1.0 / (d11 * d22 - 1.0);
        const double s = t / e;
        const double b0 = X[i * k + h], b1 = X[(i + 1) * k + h];
        X[i * k + h] = s * (d11 * b0 - b1);
        X[(i + 1) * k + h] = s * (d22 * b1 - b0);
        return;
    }
    X[i * k + h] = X[i * k + h] / d[i];
}

// dst[i] = src[perm[i]] (gather) or dst[perm[i]] = src[i] (scatter), rows of k values.
__global__ __launch_bounds__(256) void permuteRowsKernel(const double* src, double* dst, const int32_t* perm,
                                                         int64_t n, int k, int scatter)
{
    const int h = threadIdx.x & 15;
    const int64_t i = blockIdx.x * int64_t(blockDim.x >> 4) + (threadIdx.x >> 4);
    if (h >= k || i >= n)
        return;
    const int64_t at = perm[i];
    if (scatter)
        dst[at * k + h] = src[i * k + h];
    else
        dst[i * k + h] = src[at * k + h];
}

int rowGroups(int64_t n)
{
    return static_cast<int>((n + 15) / 16);
}

} // namespace

hipError_t launchBunchKaufman(const PivotFront* fronts, PivotState* states, const int32_t* list, int count, int nb,
                              hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    if (nb < 1 || nb > 2 * PANEL_BLOCK)
        return hipErrorInvalidValue;
    hipLaunchKernelGGL(bunchKaufmanKernel, dim3(count), dim3(BK_THREADS), 0, stream, fronts, states, list, nb);
    return hipGetLastError();
}

hipError_t launchPivotMatrix(const DeviceMatrix& a, const PivotFront* fronts, int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(pivotMatrixKernel, dim3(count), dim3(256), 0, stream, a, fronts);
    return hipGetLastError();
}

hipError_t launchPivotExtendAdd(const ExtendItem* items, int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(pivotExtendAddKernel, dim3(count), dim3(256), 0, stream, items);
    return hipGetLastError();
}

hipError_t launchCopyRanges(const CopyRange* ranges, int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(copyRangesKernel, dim3(count), dim3(256), 0, stream, ranges);
    return hipGetLastError();
}

hipError_t launchUnitBlockInverses(const DevicePlan& p, const double* panels, double* linv, const int2* items,
                                   int count, hipStream_t stream)
{
    if (count == 0)
        return hipSuccess;
    hipLaunchKernelGGL(unitBlockInverseKernel, dim3(count), dim3(64), 0, stream, p, panels, linv, items);
    return hipGetLastError();
}

hipError_t launchBlockDiagonalSolve(double* X, const double* d, const double* sub, int64_t n, int k,
                                    hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(blockDiagonalSolveKernel, dim3(rowGroups(n)), dim3(256), 0, stream, X, d, sub, n, k);
    return hipGetLastError();
}

hipError_t launchPermuteRows(const double* src, double* dst, const int32_t* perm, int64_t n, int k, bool scatter,
                             hipStream_t stream)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(permuteRowsKernel, dim3(rowGroups(n)), dim3(256), 0, stream, src, dst, perm, n, k,
                       scatter ? 1 : 0);
    return hipGetLastError();
}

} // namespace oosg
