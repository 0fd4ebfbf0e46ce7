// lu.hip — the step-0 solve: dense FP64 LU on the GPU, and solves with it.
//
// Stands in for KSPSetOperators + KSPSolve (src/SolverStepLinearT.cpp:503-504).  The reference
// refactors every step; here the n >= 3 system matrix (2 a1 + a3) M_0 + a1 M_1 (after the column
// exchange) is the same every step, so it is factored once and inverted once, and each step is
// one product with the inverse.
//
// lu_factor first runs a blocked LU without row interchanges (panel width 64: diagonal block in
// LDS, the rows below and the block row to the right by triangular solves, an LDS-tiled FP64 GEMM
// for the trailing matrix) and checks every multiplier for |l| <= 1: then partial pivoting would
// have kept every row in place and the factors are the partially pivoted ones.  Otherwise it
// restores the matrix and runs blocked right-looking LU with partial pivoting: a recursive panel
// factorisation with one-workgroup register leaves, row interchanges applied to the other
// columns, a triangular solve for the U12 block row and the same GEMM.  Solves use the same
// pieces column-blocked over the right-hand sides.  The large trailing-matrix and block-solve
// updates are plain GEMMs and go to rocBLAS (dgemm, atomics not allowed, so results are the same
// run to run); the small ones inside the panels stay on the LDS-tiled kernel below.
#include <algorithm>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>

#include "step_device.h"

namespace tdbem {

namespace {

constexpr int NB = 64;          // panel width

// ---------------------------------------------------------------------------------------------
// Panel factorisation, recursive (64 -> 32 -> 16 -> 8 columns).  A leaf of kLeafW columns over
// rows [c0, n) is factored by one workgroup directly in the row-major matrix: per column a block
// argmax picks the pivot (ties keep the smaller row), the two rows are swapped inside the leaf,
// and one pass over the rows below scales the column, updates the rest of the leaf and tracks the
// largest |a| of the next column, so each column costs one sweep of m x kLeafW values (L2
// resident).  Each leaf applies its row swaps to the other columns of its 64-column panel and
// k_laswp to the columns outside the panel; the stored factors
// satisfy P A = L U with L unit lower (below the diagonal) and U upper.
// ---------------------------------------------------------------------------------------------
constexpr int kLeafW = 8;
constexpr int kLeafThreads = 1024;
constexpr int kLeafUnroll = 4;    // rows in flight per thread

__device__ __forceinline__ void argmax_merge(double& v, int& i, double ov, int oi)
{
    if (ov > v || (ov == v && oi < i)) {
        v = ov;
        i = oi;
    }
}

// The leaf's row interchanges piv[c0 .. c0 + w) applied, in order, to the other columns of its
// panel [j0, j0 + pnb) (one thread per column; piv is complete after the block barrier).
__device__ __forceinline__ void leaf_panel_swaps(double* __restrict__ A, int n, int c0, int w,
                                                 int j0, int pnb, const int* __restrict__ piv)
{
    __syncthreads();
    const int t = threadIdx.x;
    if (t >= pnb - w) return;
    const int col = j0 + t + (j0 + t >= c0 ? w : 0);
    for (int c = c0; c < c0 + w; c++) {
        const int p = piv[c];
        if (p != c) {
            const double a = A[(int64_t)c * n + col];
            A[(int64_t)c * n + col] = A[(int64_t)p * n + col];
            A[(int64_t)p * n + col] = a;
        }
    }
}

__global__ void __launch_bounds__(kLeafThreads)
    k_leaf(double* __restrict__ A, int n, int c0, int w, int j0, int pnb, int* __restrict__ piv,
           int* __restrict__ status)
{
    __shared__ double sv[kLeafThreads / 64];
    __shared__ int si[kLeafThreads / 64];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    double best = -1.0;
    int bi = n;
    for (int r = c0 + tid; r < n; r += kLeafThreads) argmax_merge(best, bi, fabs(A[(int64_t)r * n + c0]), r);
    for (int c = c0; c < c0 + w; c++) {
        const int cc = c - c0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            argmax_merge(best, bi, __shfl_xor(best, off, 64), __shfl_xor(bi, off, 64));
        if (lane == 0) {
            sv[wid] = best;
            si[wid] = bi;
        }
        __syncthreads();
        double pm = sv[0];
        int p = si[0];
#pragma unroll
        for (int q = 1; q < kLeafThreads / 64; q++) argmax_merge(pm, p, sv[q], si[q]);
        // no candidate (a column of NaN after a zero pivot): keep the row, the matrix is singular
        if (p >= n) p = c;
        if (tid == 0) {
            piv[c] = p;
            if (!(pm > 0.0)) *status = 1;
        }
        if (p != c && tid < w) {
            double* a = A + (int64_t)c * n + c0 + tid;
            double* b = A + (int64_t)p * n + c0 + tid;
            const double t = *a;
            *a = *b;
            *b = t;
        }
        __syncthreads();   // swapped rows visible; sv/si free for the next column
        const double* prow = A + (int64_t)c * n + c0;
        const double d = prow[cc];
        double u[kLeafW];
#pragma unroll
        for (int j = 0; j < kLeafW; j++) u[j] = (j > cc && j < w) ? prow[j] : 0.0;
        best = -1.0;
        bi = n;
        for (int r0 = c + 1 + tid; r0 < n; r0 += kLeafThreads * kLeafUnroll) {
            double x[kLeafUnroll][kLeafW];
#pragma unroll
            for (int q = 0; q < kLeafUnroll; q++) {
                const int r = r0 + q * kLeafThreads;
#pragma unroll
                for (int j = 0; j < kLeafW; j++)
                    x[q][j] = (r < n && j >= cc && j < w) ? A[(int64_t)r * n + c0 + j] : 0.0;
            }
#pragma unroll
            for (int q = 0; q < kLeafUnroll; q++) {
                const int r = r0 + q * kLeafThreads;
                if (r >= n) continue;
                double xc = 0.0;
#pragma unroll
                for (int j = 0; j < kLeafW; j++)
                    if (j == cc) xc = x[q][j];
                const double l = xc / d;
                double* row = A + (int64_t)r * n + c0;
#pragma unroll
                for (int j = 0; j < kLeafW; j++) {
                    if (j == cc) {
                        row[j] = l;
                    } else if (j > cc && j < w) {
                        const double v = x[q][j] - l * u[j];
                        row[j] = v;
                        if (j == cc + 1) argmax_merge(best, bi, fabs(v), r);
                    }
                }
            }
        }
        __syncthreads();   // the column's updates visible before the next pivot row is read
    }
    leaf_panel_swaps(A, n, c0, w, j0, pnb, piv);
}

// The same leaf with the m x kLeafW panel held in registers (thread t owns local rows
// t + 1024 q, q < RPT): per column one block argmax, one exchange of the two swapped rows through
// LDS (the new pivot row doubles as the broadcast row), then the update in registers.  The panel
// is read and written once.
template <int RPT>
__global__ void __launch_bounds__(kLeafThreads)
    k_leaf_reg(double* __restrict__ A, int n, int c0, int w, int j0, int pnb,
               int* __restrict__ piv, int* __restrict__ status)
{
    __shared__ double sv[kLeafThreads / 64];
    __shared__ int si[kLeafThreads / 64];
    __shared__ double rowc[kLeafW], rowp[kLeafW];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const int m = n - c0;
    double x[RPT][kLeafW];
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int lr = tid + q * kLeafThreads;
#pragma unroll
        for (int j = 0; j < kLeafW; j++)
            x[q][j] = (lr < m && j < w) ? A[(int64_t)(c0 + lr) * n + c0 + j] : 0.0;
    }
    for (int cc = 0; cc < w; cc++) {
        const int c = c0 + cc;
        double best = -1.0;
        int bi = n;
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int lr = tid + q * kLeafThreads;
            double v = 0.0;
#pragma unroll
            for (int j = 0; j < kLeafW; j++)
                if (j == cc) v = x[q][j];
            if (lr >= cc && lr < m) argmax_merge(best, bi, fabs(v), c0 + lr);
        }
#pragma unroll
        for (int off = 32; off > 0; off >>= 1)
            argmax_merge(best, bi, __shfl_xor(best, off, 64), __shfl_xor(bi, off, 64));
        if (lane == 0) {
            sv[wid] = best;
            si[wid] = bi;
        }
        __syncthreads();
        double pm = sv[0];
        int p = si[0];
#pragma unroll
        for (int q = 1; q < kLeafThreads / 64; q++) argmax_merge(pm, p, sv[q], si[q]);
        // no candidate (a column of NaN after a zero pivot): keep the row, the matrix is singular
        if (p >= n) p = c;
        if (tid == 0) {
            piv[c] = p;
            if (!(pm > 0.0)) *status = 1;
        }
        // rows cc (thread cc, q = 0) and lp = p - c0 trade places; rowp is the new pivot row
        const int lp = p - c0, tp_ = lp % kLeafThreads, qp = lp / kLeafThreads;
        if (tid == cc)
#pragma unroll
            for (int j = 0; j < kLeafW; j++) rowc[j] = x[0][j];
        if (tid == tp_)
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (q == qp)
#pragma unroll
                    for (int j = 0; j < kLeafW; j++) rowp[j] = x[q][j];
        __syncthreads();
        double u[kLeafW];
#pragma unroll
        for (int j = 0; j < kLeafW; j++) u[j] = rowp[j];
        if (tid == tp_)
#pragma unroll
            for (int q = 0; q < RPT; q++)
                if (q == qp)
#pragma unroll
                    for (int j = 0; j < kLeafW; j++) x[q][j] = rowc[j];
        if (tid == cc)
#pragma unroll
            for (int j = 0; j < kLeafW; j++) x[0][j] = u[j];
        double d = 0.0;
#pragma unroll
        for (int j = 0; j < kLeafW; j++)
            if (j == cc) d = u[j];
#pragma unroll
        for (int q = 0; q < RPT; q++) {
            const int lr = tid + q * kLeafThreads;
            if (lr <= cc || lr >= m) continue;
            double xc = 0.0;
#pragma unroll
            for (int j = 0; j < kLeafW; j++)
                if (j == cc) xc = x[q][j];
            const double l = xc / d;
#pragma unroll
            for (int j = 0; j < kLeafW; j++) {
                if (j == cc)
                    x[q][j] = l;
                else if (j > cc && j < w)
                    x[q][j] -= l * u[j];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < RPT; q++) {
        const int lr = tid + q * kLeafThreads;
        if (lr >= m) continue;
#pragma unroll
        for (int j = 0; j < kLeafW; j++)
            if (j < w) A[(int64_t)(c0 + lr) * n + c0 + j] = x[q][j];
    }
    leaf_panel_swaps(A, n, c0, w, j0, pnb, piv);
}

// Row interchanges of panel [j0, j0 + nb) applied, in order, to the columns outside the panel.
__global__ void k_laswp(double* __restrict__ A, int n, int j0, int nb, const int* __restrict__ piv)
{
    int col = blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= n - nb) return;
    if (col >= j0) col += nb;
    for (int c = j0; c < j0 + nb; c++) {
        const int p = piv[c];
        if (p != c) {
            const double a = A[(int64_t)c * n + col];
            A[(int64_t)c * n + col] = A[(int64_t)p * n + col];
            A[(int64_t)p * n + col] = a;
        }
    }
}

// X = L11^{-1} X for the nb <= 64 rows starting at j0 (L unit lower, from the panel), columns
// [c0, c0 + ncols) of a row-major matrix B with leading dimension ldb.  Thread per column; the
// column is held in registers and L11 in LDS, both with their rows in reverse order, so forward
// substitution runs as the back substitution of k_trsm_upper.  L and B may be one matrix but the
// regions read through L and written through B do not overlap.
__global__ void __launch_bounds__(128)
    k_trsm_lower_unit(const double* __restrict__ L, int ldl, int j0, int nb,
                      double* __restrict__ B, int ldb, int brow0, int c0, int ncols)
{
    __shared__ double Ls[NB][NB + 1];
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
        const int r = t / NB, s = t % NB;
        // Ls[r][s] = L11[nb-1-r][nb-1-s] for s > r (strictly lower part of L11), 1 on the
        // diagonal, identity padding for rows past nb
        Ls[r][s] = (s > r && s < nb) ? L[(int64_t)(j0 + nb - 1 - r) * ldl + j0 + nb - 1 - s]
                                     : (r == s ? 1.0 : 0.0);
    }
    __syncthreads();
    const int col = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= c0 + ncols) return;
    double* xb = B + (int64_t)brow0 * ldb + col;
    double x[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) x[r] = r < nb ? xb[(int64_t)(nb - 1 - r) * ldb] : 0.0;
#pragma unroll
    for (int r = NB - 1; r >= 0; r--) {
        double v = x[r];
#pragma unroll
        for (int s = r + 1; s < NB; s++) v -= Ls[r][s] * x[s];
        x[r] = v / Ls[r][r];
    }
#pragma unroll
    for (int r = 0; r < NB; r++)
        if (r < nb) xb[(int64_t)(nb - 1 - r) * ldb] = x[r];
}

// X = U11^{-1} X (U upper with its diagonal) for the nb <= 64 rows starting at j0.
__global__ void __launch_bounds__(128)
    k_trsm_upper(const double* __restrict__ U, int ldu, int j0, int nb, double* __restrict__ B,
                 int ldb, int brow0, int c0, int ncols)
{
    __shared__ double Us[NB][NB + 1];
    for (int t = threadIdx.x; t < NB * NB; t += blockDim.x) {
        const int r = t / NB, s = t % NB;
        // identity outside the nb x nb block keeps the unrolled solve exact for nb < 64
        Us[r][s] = (r < nb && s < nb) ? U[(int64_t)(j0 + r) * ldu + j0 + s] : (r == s ? 1.0 : 0.0);
    }
    __syncthreads();
    const int col = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    if (col >= c0 + ncols) return;
    double* xb = B + (int64_t)brow0 * ldb + col;
    double x[NB];
#pragma unroll
    for (int r = 0; r < NB; r++) x[r] = r < nb ? xb[(int64_t)r * ldb] : 0.0;
#pragma unroll
    for (int r = NB - 1; r >= 0; r--) {
        double v = x[r];
#pragma unroll
        for (int s = r + 1; s < NB; s++) v -= Us[r][s] * x[s];
        x[r] = v / Us[r][r];
    }
#pragma unroll
    for (int r = 0; r < NB; r++)
        if (r < nb) xb[(int64_t)r * ldb] = x[r];
}

// C[m x n] -= A[m x k] * B[k x n], all row-major with leading dimensions.  64x64 output tile per
// 256-thread block, 4x4 per thread, K staged through LDS 32 at a time.
constexpr int TM = 64, TN = 64, TK = 32;
__global__ void __launch_bounds__(256)
    k_gemm_sub(int m, int n, int k, const double* __restrict__ A, int lda,
               const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc)
{
    __shared__ double As[TK][TM + 1];
    __shared__ double Bs[TK][TN + 1];
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    const int row0 = blockIdx.y * TM, col0 = blockIdx.x * TN;
    double acc[4][4] = {};
    for (int k0 = 0; k0 < k; k0 += TK) {
        for (int t = threadIdx.x; t < TM * TK; t += 256) {
            const int r = t / TK, s = t % TK;
            const int gr = row0 + r, gs = k0 + s;
            As[s][r] = (gr < m && gs < k) ? A[(int64_t)gr * lda + gs] : 0.0;
        }
        for (int t = threadIdx.x; t < TK * TN; t += 256) {
            const int s = t / TN, cc = t % TN;
            const int gs = k0 + s, gc = col0 + cc;
            Bs[s][cc] = (gs < k && gc < n) ? B[(int64_t)gs * ldb + gc] : 0.0;
        }
        __syncthreads();
#pragma unroll 4
        for (int s = 0; s < TK; s++) {
            double a[4], b[4];
#pragma unroll
            for (int u = 0; u < 4; u++) a[u] = As[s][ty + 16 * u];
#pragma unroll
            for (int v = 0; v < 4; v++) b[v] = Bs[s][tx + 16 * v];
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int v = 0; v < 4; v++) acc[u][v] += a[u] * b[v];
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 4; u++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int gr = row0 + ty + 16 * u, gc = col0 + tx + 16 * v;
            if (gr < m && gc < n) C[(int64_t)gr * ldc + gc] -= acc[u][v];
        }
}

// Updates at least this large (m n k) go to rocBLAS.
constexpr double kBlasMinWork = 64.0 * 1024 * 1024;

// C -= A B on rocBLAS.  One handle per device for the whole process, created on first use and
// kept until the process ends (it holds the library's kernels and workspace on the device, about
// 0.2 GB); threads share it under a lock held while the stream is set and the product enqueued.
void blas_gemm_sub(int m, int n, int k, const double* A, int lda, const double* B, int ldb,
                   double* C, int ldc, hipStream_t s)
{
    static std::mutex mu;
    static std::map<int, rocblas_handle> handles;
    int dev = 0;
    TD_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(mu);
    rocblas_handle& h = handles[dev];
    if (!h) {
        if (rocblas_create_handle(&h) != rocblas_status_success) {
            h = nullptr;
            throw std::runtime_error("HIP error: rocblas_create_handle failed");
        }
        rocblas_set_atomics_mode(h, rocblas_atomics_not_allowed);
    }
    if (rocblas_set_stream(h, s) != rocblas_status_success)
        throw std::runtime_error("HIP error: rocblas_set_stream failed");
    // row-major C -= A B is column-major C^T -= B^T A^T: the same buffers, operands swapped
    const double alpha = -1.0, beta = 1.0;
    const rocblas_status st = rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_none, n,
                                            m, k, &alpha, B, ldb, A, lda, &beta, C, ldc);
    if (st != rocblas_status_success)
        throw std::runtime_error("HIP error: rocblas_dgemm failed (status " +
                                 std::to_string((int)st) + ")");
}

void gemm_sub(int m, int n, int k, const double* A, int lda, const double* B, int ldb, double* C,
              int ldc, hipStream_t s)
{
    if (m <= 0 || n <= 0 || k <= 0) return;
    if ((double)m * n * k >= kBlasMinWork) {
        blas_gemm_sub(m, n, k, A, lda, B, ldb, C, ldc, s);
        return;
    }
    dim3 grid((n + TN - 1) / TN, (m + TM - 1) / TM);
    hipLaunchKernelGGL(k_gemm_sub, grid, dim3(256), 0, s, m, n, k, A, lda, B, ldb, C, ldc);
    TD_HIP(hipGetLastError());
}

// B_out[r][:] = B_in[perm[r]][:]
__global__ void k_permute_rows(const double* __restrict__ Bin, const int* __restrict__ perm,
                               int n, int ncols, double* __restrict__ Bout)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * ncols) return;
    const int r = (int)(t / ncols), c = (int)(t % ncols);
    Bout[t] = Bin[(int64_t)perm[r] * ncols + c];
}

__global__ void k_perm_identity(const int* __restrict__ perm, int n, double* __restrict__ B)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)n * n) return;
    const int r = (int)(t / n), c = (int)(t % n);
    B[t] = (perm[r] == c) ? 1.0 : 0.0;
}

// y = A x + beta y for a dense n x n row-major A: one 256-thread block per row.
__global__ void __launch_bounds__(256)
    k_gemv(const double* __restrict__ A, int n, const double* __restrict__ x,
           double* __restrict__ y, double beta)
{
    __shared__ double red[256];
    const int r = blockIdx.x;
    const double* row = A + (int64_t)r * n;
    double s = 0.0;
    for (int c = threadIdx.x; c < n; c += 256) s += row[c] * x[c];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) y[r] = red[0] + (beta != 0.0 ? beta * y[r] : 0.0);
}

std::vector<int> host_permutation(const int* dpiv, int n, hipStream_t s)
{
    std::vector<int> piv(n), perm(n);
    TD_HIP(hipMemcpyAsync(piv.data(), dpiv, n * sizeof(int), hipMemcpyDeviceToHost, s));
    TD_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < n; i++) perm[i] = i;
    for (int c = 0; c < n; c++) std::swap(perm[c], perm[piv[c]]);
    return perm;
}

// B := U^{-1} L^{-1} B for B of n x ncols.  identity: B is the n x n identity, so L^{-1} B is
// unit lower triangular and the forward pass of row block b touches only its first
// j0 + nb columns (a third of the work of the full forward pass).
void solve_inplace(const double* LU, int n, double* B, int ncols, hipStream_t s,
                   bool identity = false)
{
    const int nblk = (n + NB - 1) / NB;
    const int tpb = 128;
    for (int b = 0; b < nblk; b++) {
        const int j0 = b * NB, nb = std::min(NB, n - j0);
        const int cols = identity ? j0 + nb : ncols;
        hipLaunchKernelGGL(k_trsm_lower_unit, dim3((cols + tpb - 1) / tpb), dim3(tpb), 0, s, LU,
                           n, j0, nb, B, ncols, j0, 0, cols);
        TD_HIP(hipGetLastError());
        const int rest = n - j0 - nb;
        gemm_sub(rest, cols, nb, LU + (int64_t)(j0 + nb) * n + j0, n, B + (int64_t)j0 * ncols,
                 ncols, B + (int64_t)(j0 + nb) * ncols, ncols, s);
    }
    for (int b = nblk - 1; b >= 0; b--) {
        const int j0 = b * NB, nb = std::min(NB, n - j0);
        hipLaunchKernelGGL(k_trsm_upper, dim3((ncols + tpb - 1) / tpb), dim3(tpb), 0, s, LU, n,
                           j0, nb, B, ncols, j0, 0, ncols);
        TD_HIP(hipGetLastError());
        gemm_sub(j0, ncols, nb, LU + j0, n, B + (int64_t)j0 * ncols, ncols, B, ncols, s);
    }
}

}  // namespace

namespace {

// Factors columns [c0, c0 + w) over rows [c0, n) inside panel [j0, j0 + pnb); the leaves apply
// their row swaps to every column of the panel.
void panel_rec(double* A, int n, int c0, int w, int j0, int pnb, int* piv, int* status,
               hipStream_t s)
{
    if (w <= kLeafW) {
        const int rpt = (n - c0 + kLeafThreads - 1) / kLeafThreads;
        const dim3 g1(1), b1(kLeafThreads);
        switch (rpt) {
        case 1: hipLaunchKernelGGL(k_leaf_reg<1>, g1, b1, 0, s, A, n, c0, w, j0, pnb, piv, status); break;
        case 2: hipLaunchKernelGGL(k_leaf_reg<2>, g1, b1, 0, s, A, n, c0, w, j0, pnb, piv, status); break;
        case 3: hipLaunchKernelGGL(k_leaf_reg<3>, g1, b1, 0, s, A, n, c0, w, j0, pnb, piv, status); break;
        case 4: hipLaunchKernelGGL(k_leaf_reg<4>, g1, b1, 0, s, A, n, c0, w, j0, pnb, piv, status); break;
        case 5: hipLaunchKernelGGL(k_leaf_reg<5>, g1, b1, 0, s, A, n, c0, w, j0, pnb, piv, status); break;
        default: hipLaunchKernelGGL(k_leaf, g1, b1, 0, s, A, n, c0, w, j0, pnb, piv, status); break;
        }
        TD_HIP(hipGetLastError());
        return;
    }
    const int h = (w / 2 + kLeafW - 1) / kLeafW * kLeafW;
    panel_rec(A, n, c0, h, j0, pnb, piv, status, s);
    hipLaunchKernelGGL(k_trsm_lower_unit, dim3(1), dim3(64), 0, s, A, n, c0, h, A, n, c0, c0 + h,
                       w - h);
    TD_HIP(hipGetLastError());
    gemm_sub(n - c0 - h, w - h, h, A + (int64_t)(c0 + h) * n + c0, n,
             A + (int64_t)c0 * n + c0 + h, n, A + (int64_t)(c0 + h) * n + c0 + h, n, s);
    panel_rec(A, n, c0 + h, w - h, j0, pnb, piv, status, s);
}

// ---------------------------------------------------------------------------------------------
// Blocked LU without row interchanges.  Partial pivoting leaves a row in place exactly when no
// entry below the diagonal exceeds the diagonal in magnitude (ties keep the diagonal), i.e. when
// every multiplier satisfies |l| <= 1.  The step-0 matrices of the march are of that kind (their
// multipliers stay well below 1), so lu_factor first factors without interchanges, checking
// |l| <= 1 for every multiplier, and falls back to the pivoted panels only when a check fails.
// Per panel of NB columns: the NB x NB diagonal block (one wavefront, rows in registers), the
// rows below it against U11 (thread per row), the block row right of it against L11, and the
// trailing GEMM.
// ---------------------------------------------------------------------------------------------

// Diagonal block A[j0 .. j0 + nb) x [j0 .. j0 + nb), unit-lower L and upper U in place.  One
// wavefront, lane r holding row r in registers (identity past nb); at step c lane c publishes
// the pivot row through LDS and every lane below it eliminates column c.  The operations per
// element are those of the textbook right-looking elimination, in the same order.
__global__ void __launch_bounds__(64)
    k_diag_nopiv(double* __restrict__ A, int n, int j0, int nb, unsigned* __restrict__ bad)
{
    __shared__ double P[NB];
    const int r = threadIdx.x;
    double* row = A + (int64_t)(j0 + r) * n + j0;
    double a[NB];
#pragma unroll
    for (int j = 0; j < NB; j++) a[j] = (r < nb && j < nb) ? row[j] : (r == j ? 1.0 : 0.0);
    bool ok = true;
#pragma unroll
    for (int c = 0; c < NB; c++) {
        if (c < nb) {   // uniform
            if (r == c) {
#pragma unroll
                for (int j = c; j < NB; j++) P[j] = a[j];
            }
            __syncthreads();
            if (r > c) {
                const double l = a[c] / P[c];
                ok = ok && fabs(l) <= 1.0;   // false for NaN as well
                a[c] = l;
#pragma unroll
                for (int j = c + 1; j < NB; j++) a[j] -= l * P[j];
            }
            __syncthreads();   // P is rewritten by the next pivot lane
        }
    }
    if (r < nb) {
#pragma unroll
        for (int j = 0; j < NB; j++)
            if (j < nb) row[j] = a[j];
    }
    if (!ok) atomicOr(bad, 1u);
}

// L21 = A21 U11^{-1} for the rows [j0 + nb, n): each thread solves one row x U11 = a,
// x_c = (a_c - sum_{k<c} x_k U_kc) / U_cc, with U11 in LDS (identity past nb).
__global__ void __launch_bounds__(64)
    k_trsm_rows_nopiv(double* __restrict__ A, int n, int j0, int nb, unsigned* __restrict__ bad)
{
    __shared__ double Us[NB][NB + 1];
    for (int q = threadIdx.x; q < NB * NB; q += blockDim.x) {
        const int r = q / NB, c = q % NB;
        Us[r][c] = (r < nb && c < nb) ? A[(int64_t)(j0 + r) * n + j0 + c] : (r == c ? 1.0 : 0.0);
    }
    __syncthreads();
    const int row = j0 + nb + blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n) return;
    double* a = A + (int64_t)row * n + j0;
    double x[NB];
#pragma unroll
    for (int c = 0; c < NB; c++) x[c] = c < nb ? a[c] : 0.0;
    bool ok = true;
#pragma unroll
    for (int c = 0; c < NB; c++) {
        double v = x[c];
#pragma unroll
        for (int k = 0; k < c; k++) v -= x[k] * Us[k][c];
        x[c] = v / Us[c][c];
        ok = ok && (c >= nb || fabs(x[c]) <= 1.0);
    }
#pragma unroll
    for (int c = 0; c < NB; c++)
        if (c < nb) a[c] = x[c];
    if (!ok) atomicOr(bad, 1u);
}

// Returns false (A partly overwritten) when some multiplier has |l| > 1 or is not finite.
bool lu_nopiv(double* A, int n, hipStream_t s)
{
    DBuf<unsigned> bad(s);
    bad.alloc(1);
    TD_HIP(hipMemsetAsync(bad.p, 0, sizeof(unsigned), s));
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nb = std::min(NB, n - j0);
        hipLaunchKernelGGL(k_diag_nopiv, dim3(1), dim3(64), 0, s, A, n, j0, nb, bad.p);
        TD_HIP(hipGetLastError());
        const int rest = n - j0 - nb;
        if (rest <= 0) break;
        hipLaunchKernelGGL(k_trsm_rows_nopiv, dim3((rest + 63) / 64), dim3(64), 0, s, A, n, j0,
                           nb, bad.p);
        TD_HIP(hipGetLastError());
        const int tpb = 128;
        hipLaunchKernelGGL(k_trsm_lower_unit, dim3((rest + tpb - 1) / tpb), dim3(tpb), 0, s, A,
                           n, j0, nb, A, n, j0, j0 + nb, rest);
        TD_HIP(hipGetLastError());
        gemm_sub(rest, rest, nb, A + (int64_t)(j0 + nb) * n + j0, n,
                 A + (int64_t)j0 * n + j0 + nb, n, A + (int64_t)(j0 + nb) * n + j0 + nb, n, s);
    }
    unsigned h_bad = 0;
    TD_HIP(hipMemcpyAsync(&h_bad, bad.p, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    TD_HIP(hipStreamSynchronize(s));
    return h_bad == 0;
}

__global__ void k_iota(int* __restrict__ piv, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) piv[i] = i;
}

}  // namespace

void lu_factor(double* A, int n, int* piv, hipStream_t s)
{
    {
        // without interchanges first (see lu_nopiv); the copy restores A when a check fails
        DBuf<double> keep(s);
        keep.alloc((size_t)n * n);
        TD_HIP(hipMemcpyAsync(keep.p, A, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
        if (lu_nopiv(A, n, s)) {
            hipLaunchKernelGGL(k_iota, dim3((n + 255) / 256), dim3(256), 0, s, piv, n);
            TD_HIP(hipGetLastError());
            return;
        }
        TD_HIP(hipMemcpyAsync(A, keep.p, (size_t)n * n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    lu_factor_pivoted(A, n, piv, s);
}

void lu_factor_pivoted(double* A, int n, int* piv, hipStream_t s)
{
    DBuf<int> status(s);
    status.alloc(1);
    TD_HIP(hipMemsetAsync(status.p, 0, sizeof(int), s));
    for (int j0 = 0; j0 < n; j0 += NB) {
        const int nb = std::min(NB, n - j0);
        panel_rec(A, n, j0, nb, j0, nb, piv, status.p, s);
        if (n - nb > 0) {
            hipLaunchKernelGGL(k_laswp, dim3((n - nb + 255) / 256), dim3(256), 0, s, A, n, j0, nb,
                               piv);
            TD_HIP(hipGetLastError());
        }
        const int rest = n - j0 - nb;
        if (rest > 0) {
            const int tpb = 128;
            hipLaunchKernelGGL(k_trsm_lower_unit, dim3((rest + tpb - 1) / tpb), dim3(tpb), 0, s,
                               A, n, j0, nb, A, n, j0, j0 + nb, rest);
            TD_HIP(hipGetLastError());
            gemm_sub(rest, rest, nb, A + (int64_t)(j0 + nb) * n + j0, n,
                     A + (int64_t)j0 * n + j0 + nb, n, A + (int64_t)(j0 + nb) * n + j0 + nb, n,
                     s);
        }
    }
    int h_status = 0;
    TD_HIP(hipMemcpyAsync(&h_status, status.p, sizeof(int), hipMemcpyDeviceToHost, s));
    TD_HIP(hipStreamSynchronize(s));
    if (h_status) throw std::runtime_error("step-0 system matrix is singular");
}

void lu_solve(const double* LU, int n, const int* piv, double* B, int ncols, hipStream_t s)
{
    std::vector<int> perm = host_permutation(piv, n, s);
    DBuf<int> dperm(s);
    DBuf<double> tmp(s);
    dperm.upload(perm.data(), n, s);
    tmp.alloc((size_t)n * ncols);
    const int64_t tot = (int64_t)n * ncols;
    hipLaunchKernelGGL(k_permute_rows, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, B,
                       dperm.p, n, ncols, tmp.p);
    TD_HIP(hipGetLastError());
    solve_inplace(LU, n, tmp.p, ncols, s);
    TD_HIP(hipMemcpyAsync(B, tmp.p, tot * sizeof(double), hipMemcpyDeviceToDevice, s));
    TD_HIP(hipStreamSynchronize(s));
}

void lu_inverse(const double* LU, int n, const int* piv, double* Ainv, hipStream_t s)
{
    std::vector<int> perm = host_permutation(piv, n, s);
    DBuf<int> dperm(s);
    dperm.upload(perm.data(), n, s);
    const int64_t tot = (int64_t)n * n;
    hipLaunchKernelGGL(k_perm_identity, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s,
                       dperm.p, n, Ainv);
    TD_HIP(hipGetLastError());
    bool identity = true;
    for (int i = 0; i < n && identity; i++) identity = perm[i] == i;
    solve_inplace(LU, n, Ainv, n, s, identity);
    TD_HIP(hipStreamSynchronize(s));
}

void dense_solve(int n, const double* hA, int nrhs, double* hB, bool inverse, bool pivoted)
{
    hipStream_t s = nullptr;
    DBuf<double> A, B;
    DBuf<int> piv;
    A.upload(hA, (size_t)n * n, s);
    B.upload(hB, (size_t)n * nrhs, s);
    piv.alloc(n);
    if (pivoted)
        lu_factor_pivoted(A.p, n, piv.p, s);
    else
        lu_factor(A.p, n, piv.p, s);
    if (!inverse) {
        lu_solve(A.p, n, piv.p, B.p, nrhs, s);
    } else {
        DBuf<double> Ai, X;
        Ai.alloc((size_t)n * n);
        X.alloc((size_t)n * nrhs);
        lu_inverse(A.p, n, piv.p, Ai.p, s);
        // X = Ai B, column by column of B through the GEMV (diagnostic, not fast)
        DBuf<double> col, out;
        col.alloc(n);
        out.alloc(n);
        std::vector<double> hc(n), ho(n);
        for (int r = 0; r < nrhs; r++) {
            for (int i = 0; i < n; i++) hc[i] = hB[(size_t)i * nrhs + r];
            TD_HIP(hipMemcpy(col.p, hc.data(), n * sizeof(double), hipMemcpyHostToDevice));
            gemv(Ai.p, n, col.p, out.p, 0.0, s);
            TD_HIP(hipMemcpy(ho.data(), out.p, n * sizeof(double), hipMemcpyDeviceToHost));
            for (int i = 0; i < n; i++) hB[(size_t)i * nrhs + r] = ho[i];
        }
        return;
    }
    TD_HIP(hipMemcpy(hB, B.p, (size_t)n * nrhs * sizeof(double), hipMemcpyDeviceToHost));
}

void gemv(const double* A, int n, const double* x, double* y, double beta, hipStream_t s)
{
    hipLaunchKernelGGL(k_gemv, dim3(n), dim3(256), 0, s, A, n, x, y, beta);
    TD_HIP(hipGetLastError());
}

namespace {

// Row-major n x n -> fragment order (zeros in the padding).  One thread per fragment element.
__global__ void k_to_frag(const double* __restrict__ A, int n, double* __restrict__ Af,
                          int64_t total)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int e = (int)(t & 7), lane = (int)((t >> 3) & 63);
    const int64_t fc = t >> 9;
    const int C = frag_cols(n);
    const int c = (int)(fc % C), T = (int)(fc / C);
    const int row = 16 * T + (lane & 15), col = 32 * c + 8 * (lane >> 4) + e;
    Af[t] = (row < n && col < n) ? A[(int64_t)row * n + col] : 0.0;
}

// Step solve (see apply_wave): one wavefront per (16-row tile, column split).
template <int NT>
__global__ void __launch_bounds__(64 * kApplyWaves)
    k_apply_frag(const double* __restrict__ Af, int n, const double* __restrict__ Xf, int nrhs,
                 int nsplit, double* __restrict__ parts)
{
    const int g = blockIdx.x * kApplyWaves + (threadIdx.x >> 6);
    if (g < frag_rows(n) / 16 * nsplit) apply_wave<NT>(Af, n, Xf, nrhs, nsplit, g, parts);
}

// One v_mfma_f64_16x16x4 with the lane layouts k_apply_frag assumes, for the layout self-test.
__global__ void k_mfma_f64_selftest(const double* __restrict__ A, const double* __restrict__ B,
                                    double* __restrict__ D)
{
    const int lane = threadIdx.x;
    f64x4 d = {0.0, 0.0, 0.0, 0.0};
    d = __builtin_amdgcn_mfma_f64_16x16x4f64(A[(lane & 15) * 4 + (lane >> 4)],
                                             B[(lane >> 4) * 16 + (lane & 15)], d, 0, 0, 0);
    for (int v = 0; v < 4; v++) D[((lane >> 4) + 4 * v) * 16 + (lane & 15)] = d[v];
}

}  // namespace

int64_t frag_elems(int n, int rows) { return (int64_t)frag_rows(rows) * frag_cols(n) * 32; }

void to_frag(const double* A, int n, double* Af, hipStream_t s)
{
    const int64_t total = frag_elems(n, n);
    hipLaunchKernelGGL(k_to_frag, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, A, n,
                       Af, total);
    TD_HIP(hipGetLastError());
}

int apply_splits(int n, int target_waves)
{
    // about target_waves wavefronts, at least 4 fragments (128 columns) per split
    const int nTr = frag_rows(n) / 16, C = frag_cols(n);
    return std::max(1, std::min((target_waves + nTr - 1) / nTr, C / 4));
}

void apply_frag(const double* Af, int n, const double* Xf, int nrhs, double* parts,
                hipStream_t s)
{
    const int nsplit = apply_splits(n);
    const int waves = frag_rows(n) / 16 * nsplit;
    const dim3 grid((waves + kApplyWaves - 1) / kApplyWaves), block(64 * kApplyWaves);
    switch ((nrhs + 15) / 16) {
    case 1: hipLaunchKernelGGL(k_apply_frag<1>, grid, block, 0, s, Af, n, Xf, nrhs, nsplit, parts); break;
    case 2: hipLaunchKernelGGL(k_apply_frag<2>, grid, block, 0, s, Af, n, Xf, nrhs, nsplit, parts); break;
    case 3: hipLaunchKernelGGL(k_apply_frag<3>, grid, block, 0, s, Af, n, Xf, nrhs, nsplit, parts); break;
    case 4: hipLaunchKernelGGL(k_apply_frag<4>, grid, block, 0, s, Af, n, Xf, nrhs, nsplit, parts); break;
    default: throw std::invalid_argument("step solve: at most 64 right-hand sides");
    }
    TD_HIP(hipGetLastError());
}

void mfma_f64_selftest(const double* hA, const double* hB, double* hD)
{
    DBuf<double> A, B, D;
    A.upload(hA, 16 * 4, nullptr);
    B.upload(hB, 4 * 16, nullptr);
    D.alloc(16 * 16);
    hipLaunchKernelGGL(k_mfma_f64_selftest, dim3(1), dim3(64), 0, nullptr, A.p, B.p, D.p);
    TD_HIP(hipGetLastError());
    TD_HIP(hipMemcpy(hD, D.p, 16 * 16 * sizeof(double), hipMemcpyDeviceToHost));
}

}  // namespace tdbem
