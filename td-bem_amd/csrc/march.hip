// march.hip — hot path 2: the history convolution and the step-0 solve of
// SolverStepLinearT::Solve_Averaging (src/SolverStepLinearT.cpp:314-544), FP64.
//
// With u_0 = t_0 = 0 (src/SolverT.cpp:117-125, src/SolverStepLinearT.cpp:50) the three averaged
// lag sums of the reference ("Eqn M+1", "Eqn M", "Eqn M-1", :408-482) collapse to one pass over
// the lag store:
//     hist_s = sum_{k=0}^{min(s,Kc)} ( M^G_k z^t_k - M^H_k z^u_k ),
//     z_0 = (a2 - a1) x_{s-1},   z_1 = a3 x_{s-1} + a2 x_{s-2},
//     z_k = y_{s-k} = a1 x_{s-k+1} + a3 x_{s-k} + a2 x_{s-k-1}   (k >= 2),
// where x = (u, t) after the boundary-condition post-processing.  The y vectors live in a ring of
// Kc + 2 slots: slot s-1 holds z_1 until x_s is known and then becomes y_{s-1}; slot s holds z_0.
// Slots of negative time are never written and stay zero, which is what x_m = 0 (m <= 0) needs.
//
// Each step: conv (stream the shell store once), r = f(s) gF + gU + hist, x = A^{-1} r with the
// inverse of the constant n >= 3 matrix, then the post-processing of :512-536.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "step_device.h"

namespace tdbem {

struct March {
    // every buffer stream-ordered on the context's stream: a re-initialisation or the context's
    // end frees them without a device-wide synchronisation (which would wait for the kernels of
    // every other context's set-up and march)
    explicit March(hipStream_t st)
        : Af(st), Xf(st), parts(st), gF(st), gU(st), wF(st), wU(st), ring(st), hist(st), U(st),
          T(st), is_ubc(st), is_fbc(st), ubc_val(st), fbc_val(st), fac(st), d_amp(st),
          d_rise(st), zs(st), partial(st), carry(st)
    {
    }
    int N = 0, nrhs = 1, R = 0, s = 0;
    double a1 = 0, a2 = 0, a3 = 0;
    std::vector<double> amp, t_rise;
    DBuf<double> Af;        // inverse of the n >= 3 system matrix, FP64-MFMA fragment order
    DBuf<double> Xf;        // right-hand sides of the step solve, same order (16 per tile row)
    DBuf<double> parts;     // [nsplit][nrhs][Np] column-split partial solutions
    int nsplit = 1, Np = 0;
    DBuf<double> gF, gU;    // G~ y split into the FBC part (scaled by f) and the UBC part
    DBuf<double> wF, wU;    // A_2^{-1} (G~_2 y_F), A_2^{-1} (G~_2 y_U) for the first solve
    DBuf<double> ring;      // [R][2][N][nrhs], part 0 = u, part 1 = t (node-major)
    DBuf<double> hist;      // [nrhs][N]
    DBuf<double> U, T;      // [numstep][nrhs][N]
    DBuf<int> is_ubc, is_fbc;
    DBuf<double> ubc_val, fbc_val;   // dense per dof
    DBuf<double> fac;       // [nrhs] load factor of the current step
    DBuf<double> d_amp, d_rise;
    int precision = 0;             // 1: split-BF16 MFMA convolution
    DBuf<float> zs;                // the ring as split bf16 terms in MFMA operand order (mixed)
    DBuf<double> partial;          // per-wave partial sums of the mixed convolution
    DBuf<double> carry;            // time-blocked: the next step's history part, summed
    int64_t zslot = 0;
};

namespace {

constexpr int kConvThreads = 256;

inline unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

__global__ void k_fac(const double* __restrict__ amp, const double* __restrict__ rise, double t,
                      int nrhs, double* __restrict__ fac)
{
    const int r = threadIdx.x;
    if (r < nrhs) fac[r] = load_factor(amp[r], rise[r], t);
}

__global__ void k_combine(const double* __restrict__ G0, const double* __restrict__ H0,
                          const double* __restrict__ G1, const double* __restrict__ H1, double c0,
                          double c1, const int* __restrict__ is_ubc, int N,
                          double* __restrict__ A, double* __restrict__ Gc, double* __restrict__ Hc)
{
    // A = H_comb with UBC columns replaced by -G_comb (ExchangeColumns, src/SolverT.cpp:265-342)
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * N) return;
    const int col = (int)(t % N);
    const double g = c0 * G0[t] + (G1 ? c1 * G1[t] : 0.0);
    const double h = c0 * H0[t] + (H1 ? c1 * H1[t] : 0.0);
    A[t] = is_ubc[col] ? -g : h;
    Gc[t] = g;
    Hc[t] = h;
}

__global__ void k_pack2(const double* __restrict__ a, const double* __restrict__ b, int N,
                        double* __restrict__ out)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= N) return;
    out[2 * d] = a[d];
    out[2 * d + 1] = b[d];
}

__global__ void k_unpack2(const double* __restrict__ in, int N, double* __restrict__ a,
                          double* __restrict__ b)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= N) return;
    a[d] = in[2 * d];
    b[d] = in[2 * d + 1];
}

// Xf = r (nrhs right-hand sides [nrhs][N]) in the fragment order of the step solve
__global__ void k_frag_pack(const double* __restrict__ r, int N, int nrhs, double* __restrict__ Xf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * nrhs) return;
    const int k = (int)(t / N), d = (int)(t % N);
    Xf[frag_index(N, k, d)] = r[t];
}

// x += scale x the sum of the column-split partials of the step solve ([nsplit][nrhs][Np])
__global__ void k_parts_add(const double* __restrict__ parts, int nsplit, int Np, int N,
                            int nrhs, double scale, double* __restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * nrhs) return;
    const int k = (int)(t / N), d = (int)(t % N);
    double v = 0.0;
    for (int p = 0; p < nsplit; p++) v += parts[((int64_t)p * nrhs + k) * Np + d];
    x[t] += scale * v;
}

__global__ void k_sub(const double* __restrict__ a, const double* __restrict__ b, int64_t n,
                      double* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = a[t] - b[t];
}

__global__ void k_negate(double* __restrict__ v, int N)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < N) v[d] = -v[d];
}

// hist[r][3i+p] = sum over row-i entries of G z^t - H z^u, one block per row node.
template <int NRHS>
__global__ void __launch_bounds__(kConvThreads)
    k_conv(const double* __restrict__ vals, const uint32_t* __restrict__ idx,
           const int64_t* __restrict__ row_chunk, int N, const double* __restrict__ ring, int R,
           int s, double* __restrict__ hist)
{
    __shared__ double red[kConvThreads / 64][3 * NRHS];
    const int i = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    double acc[NRHS][3];
#pragma unroll
    for (int r = 0; r < NRHS; r++) acc[r][0] = acc[r][1] = acc[r][2] = 0.0;
    const int64_t c1 = row_chunk[i + 1];
    // U chunks per wavefront and iteration: their loads are all issued before the products
    constexpr int U = NRHS == 1 ? 2 : 1;
    constexpr int W = kConvThreads / 64;
    for (int64_t c = row_chunk[i] + wave; c < c1; c += U * W) {
        // the store is read once per step: non-temporal loads keep the history ring in cache
        double g[U][kVals], zv[U][2][3 * NRHS];
#pragma unroll
        for (int q = 0; q < U; q++) {
            const int64_t cq = c + q * W;
            if (cq >= c1) {
#pragma unroll
                for (int v = 0; v < kVals; v++) g[q][v] = 0.0;
#pragma unroll
                for (int v = 0; v < 3 * NRHS; v++) zv[q][0][v] = zv[q][1][v] = 0.0;
                continue;
            }
            const uint32_t id = __builtin_nontemporal_load(&idx[cq * kChunk + lane]);
            const int j = (int)(id >> kKBits), k = (int)(id & kMaxLag);
            int slot = (s - k) % R;
            if (slot < 0) slot += R;
            const double* v = vals + cq * (kVals * kChunk) + lane;
#pragma unroll
            for (int w = 0; w < kVals; w++) g[q][w] = __builtin_nontemporal_load(&v[64 * w]);
            // node-major ring: the 3 NRHS values of node j's u (then t) are contiguous
            const double* zu = ring + ((int64_t)slot * 2 * N + 3 * j) * NRHS;
            const double* zt = zu + (int64_t)N * NRHS;
#pragma unroll
            for (int w = 0; w < 3 * NRHS; w++) {
                zv[q][0][w] = zu[w];
                zv[q][1][w] = zt[w];
            }
        }
#pragma unroll
        for (int q = 0; q < U; q++) {
            const double g00 = g[q][0], g01 = g[q][1], g02 = g[q][2], g11 = g[q][3],
                         g12 = g[q][4], g22 = g[q][5];
            const double h00 = g[q][6], h01 = g[q][7], h02 = g[q][8], h10 = g[q][9],
                         h11 = g[q][10], h12 = g[q][11], h20 = g[q][12], h21 = g[q][13],
                         h22 = g[q][14];
#pragma unroll
            for (int r = 0; r < NRHS; r++) {
                const double u0 = zv[q][0][r], u1 = zv[q][0][NRHS + r], u2 = zv[q][0][2 * NRHS + r];
                const double t0 = zv[q][1][r], t1 = zv[q][1][NRHS + r], t2 = zv[q][1][2 * NRHS + r];
                acc[r][0] += g00 * t0 + g01 * t1 + g02 * t2 - (h00 * u0 + h01 * u1 + h02 * u2);
                acc[r][1] += g01 * t0 + g11 * t1 + g12 * t2 - (h10 * u0 + h11 * u1 + h12 * u2);
                acc[r][2] += g02 * t0 + g12 * t1 + g22 * t2 - (h20 * u0 + h21 * u1 + h22 * u2);
            }
        }
    }
#pragma unroll
    for (int r = 0; r < NRHS; r++)
#pragma unroll
        for (int p = 0; p < 3; p++) {
            double x = acc[r][p];
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
            if (lane == 0) red[wave][3 * r + p] = x;
        }
    __syncthreads();
    if (threadIdx.x < 3 * NRHS) {
        double x = 0.0;
#pragma unroll
        for (int w = 0; w < kConvThreads / 64; w++) x += red[w][threadIdx.x];
        const int r = threadIdx.x / 3, p = threadIdx.x % 3;
        hist[(int64_t)r * N + 3 * i + p] = x;
    }
}

// x_1 = f_r(dt) wF + wU  (the n == 2 solve, src/SolverStepLinearT.cpp:355-378), written as
// split 0 of the step-solve partials x[r][Np]
__global__ void k_first(const double* __restrict__ wF, const double* __restrict__ wU,
                        const double* __restrict__ fac, int N, int Np, int nrhs,
                        double* __restrict__ x)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)N * nrhs) return;
    const int r = (int)(t / N), d = (int)(t % N);
    x[(int64_t)r * Np + d] = fac[r] * wF[d] + wU[d];
}

// r = f_r gF + gU + hist, written in the fragment order of the step solve
__global__ void k_rhs(const double* __restrict__ gF, const double* __restrict__ gU,
                      const double* __restrict__ hist, const double* __restrict__ fac, int N,
                      int nrhs, double* __restrict__ Xf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)N * nrhs) rhs_elem(gF, gU, hist, fac, N, Xf, t);
}

// The step's end (see finalize_elem).
__global__ void k_finalize(const double* __restrict__ parts, int nparts, int Np,
                           const int* __restrict__ is_ubc,
                           const int* __restrict__ is_fbc, const double* __restrict__ ubc_val,
                           const double* __restrict__ fbc_val, const double* __restrict__ fac,
                           int N, int nrhs, int s, int R, double a1, double a2, double a3,
                           double* __restrict__ U, double* __restrict__ T, double* __restrict__ ring)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < (int64_t)N * nrhs)
        finalize_elem(parts, nparts, Np, is_ubc, is_fbc, ubc_val, fbc_val, fac, N, nrhs, s, R, a1,
                      a2, a3, U, T, ring, t);
}

// x = A2^{-1} b for the two right-hand sides ([2][N]) of the first solve, by iterative
// refinement with the inverse of the n >= 3 matrix A = c0 (A2 + (a1 / c0) M_1) (column-exchanged
// alike), c0 = 2 a1 + a3: x <- x + c0 A^{-1} (b - A2 x) contracts with the spectral radius of
// (A / c0)^{-1} (a1 / c0) M_1 (about 0.2 for the sphere and 0.5 for the bar at a1 = a2 = 1/4;
// 0 when a1 = 0).  Stops at a residual of 1e-15 of |A2 x| + |b|; returns false when it stalls
// above 1e-14 or runs out of iterations.
bool refine_first(const double* A2, int N, const double* Af, double c0, const double* b,
                  double* x, hipStream_t s)
{
    const int Np = frag_rows(N), nsplit = apply_splits(N);
    const int64_t n2 = 2LL * N;
    DBuf<double> Xf(s), parts(s), r(s), ax(s);
    Xf.alloc((size_t)frag_elems(N, 2));
    parts.alloc((size_t)nsplit * 2 * Np);
    r.alloc(n2);
    ax.alloc(n2);
    TD_HIP(hipMemsetAsync(Xf.p, 0, Xf.bytes(), s));   // padding stays zero
    TD_HIP(hipMemsetAsync(x, 0, n2 * sizeof(double), s));
    TD_HIP(hipMemcpyAsync(r.p, b, n2 * sizeof(double), hipMemcpyDeviceToDevice, s));
    std::vector<double> hb(n2), hr(n2), hax(n2);
    TD_HIP(hipMemcpyAsync(hb.data(), b, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
    TD_HIP(hipStreamSynchronize(s));
    auto norm = [&](const std::vector<double>& v, int k) {
        double q = 0.0;
        for (int d = 0; d < N; d++) q += v[(size_t)k * N + d] * v[(size_t)k * N + d];
        return std::sqrt(q);
    };
    const double bn[2] = {norm(hb, 0), norm(hb, 1)};
    double prev = 1e300;
    // the residual is read back (and the host waits) after every kCheck iterations only
    constexpr int kCheck = 4;
    for (int it = 0; it < 200; it++) {
        hipLaunchKernelGGL(k_frag_pack, dim3(blocks(n2)), dim3(256), 0, s, r.p, N, 2, Xf.p);
        apply_frag(Af, N, Xf.p, 2, parts.p, s);
        hipLaunchKernelGGL(k_parts_add, dim3(blocks(n2)), dim3(256), 0, s, parts.p, nsplit, Np, N,
                           2, c0, x);
        gemv(A2, N, x, ax.p, 0.0, s);
        gemv(A2, N, x + N, ax.p + N, 0.0, s);
        hipLaunchKernelGGL(k_sub, dim3(blocks(n2)), dim3(256), 0, s, b, ax.p, n2, r.p);
        TD_HIP(hipGetLastError());
        if (it % kCheck != kCheck - 1) continue;
        TD_HIP(hipMemcpyAsync(hr.data(), r.p, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
        TD_HIP(hipMemcpyAsync(hax.data(), ax.p, n2 * sizeof(double), hipMemcpyDeviceToHost, s));
        TD_HIP(hipStreamSynchronize(s));
        double rel = 0.0;
        for (int k = 0; k < 2; k++) {
            const double scale = norm(hax, k) + bn[k];
            const double rn = norm(hr, k);
            if (!std::isfinite(rn)) return false;
            if (scale > 0) rel = std::max(rel, rn / scale);
        }
        if (rel <= 1e-15) return true;
        // stalled: kCheck more iterations gained less than a factor 0.7
        if (it >= 2 * kCheck && rel > 0.7 * prev) return rel <= 1e-14;
        prev = rel;
    }
    return false;
}

}  // namespace

void conv_fp64(Context& c, const double* ring, int R, int s, int nrhs, double* hist,
               hipStream_t st)
{
    const int N = 3 * c.nnd;
    const dim3 grid(c.nnd), block(kConvThreads);
    const double *v = c.store.vals.p;
    const uint32_t* idx = c.store.idx.p;
    const int64_t* rc = c.store.row_chunk.p;
    switch (nrhs) {
    case 1: hipLaunchKernelGGL(k_conv<1>, grid, block, 0, st, v, idx, rc, N, ring, R, s, hist); break;
    case 2: hipLaunchKernelGGL(k_conv<2>, grid, block, 0, st, v, idx, rc, N, ring, R, s, hist); break;
    case 4: hipLaunchKernelGGL(k_conv<4>, grid, block, 0, st, v, idx, rc, N, ring, R, s, hist); break;
    case 8: hipLaunchKernelGGL(k_conv<8>, grid, block, 0, st, v, idx, rc, N, ring, R, s, hist); break;
    default: throw std::runtime_error("nrhs must be 1, 2, 4 or 8");
    }
    TD_HIP(hipGetLastError());
}

void march_free(Context& c)
{
    delete c.march;
    c.march = nullptr;
}

void march_init(Context& c, double a1, double a2, int nrhs, const double* amp,
                const double* t_rise, int n_ubc, const int* ubc_dof, const double* ubc_val,
                int n_fbc, const int* fbc_dof, const double* fbc_val)
{
    // src/SolverStepLinearT.cpp:317-322
    if (!(a1 >= 0 && a1 <= 1 && a2 >= 0 && a2 < 1 && a1 + a2 <= 1))
        throw std::runtime_error("Alpha is out of range");
    if (c.precision == 0 && !(nrhs == 1 || nrhs == 2 || nrhs == 4 || nrhs == 8))
        throw std::invalid_argument("FP64 march: nrhs must be 1, 2, 4 or 8");
    if (c.precision == 1 && !(nrhs == 16 || nrhs == 32 || nrhs == 48 || nrhs == 64))
        throw std::invalid_argument("mixed march: nrhs must be 16, 32, 48 or 64");
    const int N = 3 * c.nnd;
    // the boundary data is checked before anything is freed or built: a rejected call leaves
    // the context as it was
    for (int q = 0; q < n_fbc; q++)
        if (fbc_dof[q] < 0 || fbc_dof[q] >= N) throw std::invalid_argument("FBC dof out of range");
    for (int q = 0; q < n_ubc; q++)
        if (ubc_dof[q] < 0 || ubc_dof[q] >= N) throw std::invalid_argument("UBC dof out of range");
    if (c.caller_last > 0) {
        // caller-supplied lags: the step-0 matrices need M_0 and M_1 complete, i.e. the blocks
        // of steps 1 and 2 (of step 1 alone when the store keeps one lag)
        if (c.caller_last < std::min(2, c.Kc))
            throw std::logic_error("UpdateGHLinear of steps 1 and 2 must come before the march");
        if (c.caller_stale || !c.store_ready) caller_store(c);
    } else if (!c.store_ready) {
        build_lag_store(c);
    }
    const auto t0 = std::chrono::steady_clock::now();
    march_free(c);
    // the new march state becomes the context's only when it is complete (a failure below
    // leaves no march, never a partly built one)
    std::unique_ptr<March> owner(new March(c.stream));
    March& M = *owner;
    M.N = N;
    M.nrhs = nrhs;
    M.a1 = a1;
    M.a2 = a2;
    M.a3 = 1 - a1 - a2;
    M.R = c.Kc + 2;
    M.amp.assign(nrhs, 1.0);
    M.t_rise.assign(nrhs, 0.0);
    for (int r = 0; r < nrhs; r++) {
        if (amp) M.amp[r] = amp[r];
        if (t_rise) M.t_rise[r] = t_rise[r];
    }
    // dense boundary data (INSERT order of SetLoad, src/SolverT.cpp:366-369)
    std::vector<int> isu(N, 0), isf(N, 0);
    std::vector<double> uv(N, 0.0), fv(N, 0.0), yF(N, 0.0), yU(N, 0.0);
    for (int q = 0; q < n_fbc; q++) {
        const int d = fbc_dof[q];
        isf[d] = 1;
        fv[d] = fbc_val[q];
    }
    for (int q = 0; q < n_ubc; q++) {
        const int d = ubc_dof[q];
        isu[d] = 1;
        uv[d] = ubc_val[q];
    }
    for (int d = 0; d < N; d++) {
        yF[d] = (isf[d] && !isu[d]) ? fv[d] : 0.0;
        yU[d] = isu[d] ? uv[d] : 0.0;
    }
    M.is_ubc.upload(isu.data(), N, c.stream);
    M.is_fbc.upload(isf.data(), N, c.stream);
    M.ubc_val.upload(uv.data(), N, c.stream);
    M.fbc_val.upload(fv.data(), N, c.stream);
    DBuf<double> dyF(c.stream), dyU(c.stream);
    dyF.upload(yF.data(), N, c.stream);
    dyU.upload(yU.data(), N, c.stream);

    const int64_t NN = (int64_t)N * N;
    DBuf<double> G0(c.stream), H0(c.stream), G1(c.stream), H1(c.stream), A(c.stream),
        Gc(c.stream), Hc(c.stream);
    G0.alloc(NN);
    H0.alloc(NN);
    G1.alloc(NN);
    H1.alloc(NN);
    A.alloc(NN);
    Gc.alloc(NN);
    Hc.alloc(NN);
    lag_dense(c, 0, G0.p, H0.p);
    lag_dense(c, 1, G1.p, H1.p);
    DBuf<int> piv(c.stream);
    piv.alloc(N);
    M.gF.alloc(N);
    M.gU.alloc(N);
    M.wF.alloc(N);
    M.wU.alloc(N);

    // n >= 3: (2 a1 + a3) M_0 + a1 M_1 (src/SolverStepLinearT.cpp:384-393)
    hipLaunchKernelGGL(k_combine, dim3(blocks(NN)), dim3(256), 0, c.stream, G0.p, H0.p, G1.p,
                       H1.p, 2 * a1 + M.a3, a1, M.is_ubc.p, N, A.p, Gc.p, Hc.p);
    TD_HIP(hipGetLastError());
    gemv(Gc.p, N, dyF.p, M.gF.p, 0.0, c.stream);
    gemv(Hc.p, N, dyU.p, M.gU.p, 0.0, c.stream);
    hipLaunchKernelGGL(k_negate, dim3(blocks(N)), dim3(256), 0, c.stream, M.gU.p, N);
    lu_factor(A.p, N, piv.p, c.stream);
    {
        // the inverse goes to fragment order; Gc holds it in row-major order meanwhile
        lu_inverse(A.p, N, piv.p, Gc.p, c.stream);
        M.Af.alloc((size_t)frag_elems(N, N));
        to_frag(Gc.p, N, M.Af.p, c.stream);
    }
    // n == 2 (src/SolverStepLinearT.cpp:355-378): A_2 = M_0 exchanged (into G1; G1, H1 and Hc
    // are free now), gF2 = G_0 y_F, gU2 = -H_0 y_U; solved by refinement with the inverse above,
    // or by its own factorisation when that does not converge
    hipLaunchKernelGGL(k_combine, dim3(blocks(NN)), dim3(256), 0, c.stream, G0.p, H0.p,
                       (const double*)nullptr, (const double*)nullptr, 1.0, 0.0, M.is_ubc.p, N,
                       G1.p, H1.p, Hc.p);
    TD_HIP(hipGetLastError());
    {
        DBuf<double> b2(c.stream), x2(c.stream);
        b2.alloc((size_t)2 * N);
        x2.alloc((size_t)2 * N);
        gemv(H1.p, N, dyF.p, b2.p, 0.0, c.stream);
        gemv(Hc.p, N, dyU.p, b2.p + N, 0.0, c.stream);
        hipLaunchKernelGGL(k_negate, dim3(blocks(N)), dim3(256), 0, c.stream, b2.p + N, N);
        if (refine_first(G1.p, N, M.Af.p, 2 * a1 + M.a3, b2.p, x2.p, c.stream)) {
            TD_HIP(hipMemcpyAsync(M.wF.p, x2.p, N * sizeof(double), hipMemcpyDeviceToDevice,
                                  c.stream));
            TD_HIP(hipMemcpyAsync(M.wU.p, x2.p + N, N * sizeof(double), hipMemcpyDeviceToDevice,
                                  c.stream));
        } else {
            DBuf<double> B(c.stream);
            B.alloc((size_t)2 * N);
            hipLaunchKernelGGL(k_pack2, dim3(blocks(N)), dim3(256), 0, c.stream, b2.p, b2.p + N,
                               N, B.p);
            lu_factor(G1.p, N, piv.p, c.stream);
            lu_solve(G1.p, N, piv.p, B.p, 2, c.stream);
            hipLaunchKernelGGL(k_unpack2, dim3(blocks(N)), dim3(256), 0, c.stream, B.p, N,
                               M.wF.p, M.wU.p);
        }
        TD_HIP(hipStreamSynchronize(c.stream));
    }

    M.precision = c.precision;
    if (M.precision == 1) {
        tp_build(c);
        tp_plan(c, nrhs);
        M.zslot = tp_zslot_elems(c, nrhs);
        M.zs.alloc((size_t)M.R * M.zslot);
        TD_HIP(hipMemsetAsync(M.zs.p, 0, M.zs.bytes(), c.stream));
        M.partial.alloc((size_t)tp_partial_elems(c, nrhs));
        M.carry.alloc((size_t)tp_carry_elems(c, nrhs));
        TD_HIP(hipMemsetAsync(M.carry.p, 0, M.carry.bytes(), c.stream));
    }
    M.ring.alloc((size_t)M.R * nrhs * 2 * N);
    TD_HIP(hipMemsetAsync(M.ring.p, 0, M.ring.bytes(), c.stream));
    if (M.precision != 1) M.hist.alloc((size_t)nrhs * N);   // the mixed path sums into Xf
    M.Np = frag_rows(N);
    M.nsplit = apply_splits(N);
    M.Xf.alloc((size_t)frag_elems(N, nrhs));
    TD_HIP(hipMemsetAsync(M.Xf.p, 0, M.Xf.bytes(), c.stream));   // padding stays zero
    M.parts.alloc((size_t)M.nsplit * nrhs * M.Np);
    M.fac.alloc(nrhs);
    M.d_amp.upload(M.amp.data(), nrhs, c.stream);
    M.d_rise.upload(M.t_rise.data(), nrhs, c.stream);
    const int nhist = std::max(c.numstep, 1);
    M.U.alloc((size_t)nhist * nrhs * N);
    M.T.alloc((size_t)nhist * nrhs * N);
    TD_HIP(hipMemsetAsync(M.U.p, 0, M.U.bytes(), c.stream));
    TD_HIP(hipMemsetAsync(M.T.p, 0, M.T.bytes(), c.stream));
    TD_HIP(hipStreamSynchronize(c.stream));
    M.s = 0;
    c.march = owner.release();
    c.tm.init_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

int march_step_count(const Context& c) { return c.march ? c.march->s : 0; }

int64_t march_state_bytes(Context& c, int nrhs)
{
    // the allocations of march_init that outlive it, term by term
    const int64_t N = 3LL * c.nnd, R = c.Kc + 2, Np = frag_rows((int)N);
    const int64_t nsplit = apply_splits((int)N);
    int64_t b = 0;
    b += frag_elems((int)N, (int)N) * 8;                 // Af
    b += frag_elems((int)N, nrhs) * 8;                   // Xf
    b += nsplit * nrhs * Np * 8;                         // parts
    b += 4 * N * 8;                                      // gF, gU, wF, wU
    b += 2 * N * 4 + 2 * N * 8;                          // is_ubc, is_fbc, ubc_val, fbc_val
    b += R * nrhs * 2 * N * 8;                           // ring
    b += 3LL * nrhs * 8;                                 // fac, amp, rise
    b += 2LL * std::max(c.numstep, 1) * nrhs * N * 8;    // U, T
    if (c.precision == 1) {
        b += R * tp_zslot_elems(c, nrhs) * 4;            // zs
        std::vector<int4> ch;
        std::vector<int> gch;
        tp_chunks(c, tp_chunk_target(c, 4), 0, ch, gch);
        const int NT = nrhs / 16;
        b += (int64_t)ch.size() * 2 * kGroup * 3 * NT * 256 * 8;   // partial
        b += (int64_t)ch.size() * 16 + (int64_t)gch.size() * 4;   // chunk, grp_chunk
        b += tp_carry_elems(c, nrhs) * 8;                         // carry
    } else {
        b += (int64_t)nrhs * N * 8;                      // hist
    }
    return b;
}

int64_t march_init_transient_bytes(const Context& c)
{
    // the seven dense N x N matrices of the step-0 set-up, the pivots and the first solve's
    // refinement buffers
    const int64_t N = 3LL * c.nnd;
    return 7 * N * N * 8 + N * 4 + 2 * frag_elems((int)N, 2) * 8 + 16 * N * 8;
}

void march_describe(Context& c, BatchScen& d)
{
    if (!c.march || c.march->precision != 1)
        throw std::logic_error("batch: every context needs a mixed-precision march_init");
    March& M = *c.march;
    d.ring = M.ring.p;
    d.gF = M.gF.p;
    d.gU = M.gU.p;
    d.Xf = M.Xf.p;
    d.Af = M.Af.p;
    d.parts = nullptr;
    d.is_ubc = M.is_ubc.p;
    d.is_fbc = M.is_fbc.p;
    d.ubc_val = M.ubc_val.p;
    d.fbc_val = M.fbc_val.p;
    d.fac = M.fac.p;
    d.amp = M.d_amp.p;
    d.rise = M.d_rise.p;
    d.U = M.U.p;
    d.T = M.T.p;
    d.N = M.N;
    d.Np = M.Np;
    d.nrhs = M.nrhs;
    d.R = M.R;
    d.numstep = c.numstep;
    d.nsplit = 0;
    d.dt = c.dt;
    d.a1 = M.a1;
    d.a2 = M.a2;
    d.a3 = M.a3;
    d.zs = M.zs.p;
    d.carry = M.carry.p;
}

void march_set_step(Context& c, int s)
{
    if (c.march) c.march->s = s;
}

void march_run(Context& c, int nsteps, bool wait)
{
    if (!c.march) throw std::runtime_error("march not initialised");
    March& M = *c.march;
    const int N = M.N;
    const int64_t tot = (int64_t)N * M.nrhs;
    const auto t0 = std::chrono::steady_clock::now();
    // Without a wait the steps are only queued; per-kernel timing needs the wait.
    KernelTimer conv_timer(c.stream, c.profile && wait),
        solve_timer(c.stream, c.profile == 1 && wait),
        reduce_timer(c.stream, c.profile == 1 && wait && M.precision == 1);
    const size_t steps = (size_t)std::max(0, std::min(nsteps, c.numstep - 1 - M.s));
    conv_timer.reserve(steps);
    solve_timer.reserve(steps);
    reduce_timer.reserve(steps);
    for (int it = 0; it < nsteps; it++) {
        if (M.s + 1 >= c.numstep) break;  // histories hold x_0 .. x_{numstep-1}
        if (c.caller_last > 0) {
            // caller-supplied lags: step s sums lags up to min(s, Kc); lag k is complete once
            // the blocks of step k + 1 are in (or k = Kc, which has no phi_2 part)
            const int complete = c.caller_last >= c.Kc ? c.Kc : c.caller_last - 1;
            if (std::min(M.s + 1, c.Kc) > complete)
                throw std::logic_error("UpdateGHLinear of step " + std::to_string(M.s + 2) +
                                       " is needed before this step");
            if (c.caller_stale) {
                TD_HIP(hipStreamSynchronize(c.stream));
                caller_store(c);
            }
        }
        const int s = ++M.s;
        const bool mixed = M.precision == 1 && s >= 2;
        if (!mixed)
            hipLaunchKernelGGL(k_fac, dim3(1), dim3(64), 0, c.stream, M.d_amp.p, M.d_rise.p,
                               s * c.dt, M.nrhs, M.fac.p);
        if (s == 1) {
            hipLaunchKernelGGL(k_first, dim3(blocks(tot)), dim3(256), 0, c.stream, M.wF.p, M.wU.p,
                               M.fac.p, N, M.Np, M.nrhs, M.parts.p);
        } else {
            if (mixed) {
                // per step: split-bf16 history copy + load factors, convolution, reduction into the
                // right-hand side, step solve, finalize
                tp_prepare_step(c, M.ring.p, N, M.nrhs, s, M.R, M.zs.p, M.d_amp.p, M.d_rise.p,
                                s * c.dt, M.fac.p);
                conv_timer.begin();
                tp_conv(c, M.zs.p, M.nrhs, M.R, s, M.partial.p);
                conv_timer.end();
                reduce_timer.begin();
                tp_reduce_rhs(c, M.partial.p, M.nrhs, N, s, M.gF.p, M.gU.p, M.fac.p,
                              M.carry.p, M.Xf.p);
                reduce_timer.end();
            } else {
                conv_timer.begin();
                conv_fp64(c, M.ring.p, M.R, s, M.nrhs, M.hist.p, c.stream);
                conv_timer.end();
                hipLaunchKernelGGL(k_rhs, dim3(blocks(tot)), dim3(256), 0, c.stream, M.gF.p,
                                   M.gU.p, M.hist.p, M.fac.p, N, M.nrhs, M.Xf.p);
            }
            solve_timer.begin();
            apply_frag(M.Af.p, N, M.Xf.p, M.nrhs, M.parts.p, c.stream);
            solve_timer.end();
        }
        hipLaunchKernelGGL(k_finalize, dim3(blocks(tot)), dim3(256), 0, c.stream, M.parts.p,
                           s == 1 ? 1 : M.nsplit, M.Np,
                           M.is_ubc.p, M.is_fbc.p, M.ubc_val.p, M.fbc_val.p, M.fac.p, N, M.nrhs,
                           s, M.R, M.a1, M.a2, M.a3, M.U.p, M.T.p, M.ring.p);
        TD_HIP(hipGetLastError());
    }
    if (!wait) return;
    TD_HIP(hipStreamSynchronize(c.stream));
    c.tm.march_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (c.profile) {
        c.tm.conv_launches += conv_timer.count();
        c.tm.conv_ms += conv_timer.collect();
        c.tm.solve_ms += solve_timer.collect();
        c.tm.reduce_ms += reduce_timer.collect();
    }
}

void march_history(Context& c, int rhs, double* U, double* T)
{
    if (!c.march) throw std::runtime_error("march not initialised");
    March& M = *c.march;
    if (rhs < 0 || rhs >= M.nrhs) throw std::runtime_error("rhs index out of range");
    const int N = M.N;
    for (int s = 0; s < c.numstep; s++) {
        const int64_t off = ((int64_t)s * M.nrhs + rhs) * N;
        TD_HIP(hipMemcpyAsync(U + (int64_t)s * N, M.U.p + off, N * sizeof(double),
                              hipMemcpyDeviceToHost, c.stream));
        TD_HIP(hipMemcpyAsync(T + (int64_t)s * N, M.T.p + off, N * sizeof(double),
                              hipMemcpyDeviceToHost, c.stream));
    }
    TD_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace tdbem

namespace tdbem {

void march_state(Context& c, int rhs, int s, double* u, double* t)
{
    if (!c.march) throw std::runtime_error("march not initialised");
    March& M = *c.march;
    if (rhs < 0 || rhs >= M.nrhs) throw std::runtime_error("rhs index out of range");
    const int64_t off = ((int64_t)s * M.nrhs + rhs) * M.N;
    TD_HIP(hipMemcpyAsync(u, M.U.p + off, M.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    TD_HIP(hipMemcpyAsync(t, M.T.p + off, M.N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    TD_HIP(hipStreamSynchronize(c.stream));
}

}  // namespace tdbem
