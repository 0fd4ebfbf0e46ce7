// caller.hip — lag matrices from caller-supplied per-step blocks, the role of
// SolverStepLinearT::UpdateGHLinear (src/SolverStepLinearT.cpp:58-97) when the G/H blocks of each
// step come from elsewhere (an FE-BE coupling code, a modified element) instead of the device's
// own assembly.  UpdateGHLinear, called at fCurrStep = n with the step's phi_1 blocks (G_0, H_0)
// and phi_2 blocks (G_1, H_1), INSERTs the phi_2 blocks into lag 0 at n = 1 and ADDs them to lag
// n - 1 after, and INSERTs the phi_1 blocks into lag n, for n <= MaxStep.  Here the blocks land
// in a dense database of the Kc + 1 lags on the device in exactly that way; before a march step
// needs a lag, the shell store is built from the database: each node pair keeps the lags whose
// 3 x 3 block of G or H holds a nonzero, so the march, its kernels and its results are those of a
// store the device assembled itself.
#include <stdexcept>
#include <string>
#include <vector>

#include "step_device.h"

namespace tdbem {

namespace {

inline unsigned grid_of(int64_t n) { return (unsigned)((n + 255) / 256); }

// blk = [G0; G1; H0; H1], each [nrow][N] (rows r0 .. r0 + nrow - 1 of the lags); db [Kc+1][2][N][N]
__global__ void k_db_update(double* __restrict__ db, int64_t N, int64_t r0, int64_t nrow, int step,
                            const double* __restrict__ blk)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = nrow * N;
    if (t >= n) return;
    const int64_t NN = N * N, off = r0 * N + t;
    const double g0 = blk[t], g1 = blk[n + t], h0 = blk[2 * n + t], h1 = blk[3 * n + t];
    if (step == 1) {
        db[off] = g1;
        db[NN + off] = h1;
    } else {
        db[(int64_t)(step - 1) * 2 * NN + off] += g1;
        db[((int64_t)(step - 1) * 2 + 1) * NN + off] += h1;
    }
    db[(int64_t)step * 2 * NN + off] = g0;
    db[((int64_t)step * 2 + 1) * NN + off] = h0;
}

// Lag range of node pair (i, j): the lags 0..K whose 3 x 3 block of G or H holds a nonzero.
__global__ void k_db_hull(const double* __restrict__ db, int64_t N, int nnd, int K,
                          int2* __restrict__ hull)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nnd * nnd) return;
    const int i = (int)(t / nnd), j = (int)(t % nnd);
    const int64_t NN = N * N;
    int lo = 1 << 30, hi = -1;
    for (int k = 0; k <= K; k++) {
        bool any = false;
        for (int p = 0; p < 3; p++)
            for (int q = 0; q < 3; q++) {
                const int64_t o = (int64_t)(3 * i + p) * N + 3 * j + q;
                any = any || db[(int64_t)k * 2 * NN + o] != 0.0 || db[((int64_t)k * 2 + 1) * NN + o] != 0.0;
            }
        if (any) {
            lo = min(lo, k);
            hi = k;
        }
    }
    hull[t] = hi < 0 ? make_int2(1, 0) : make_int2(lo, hi);
}

__device__ __forceinline__ int chunk_row(const int64_t* row_chunk, int nnd, int64_t c)
{
    int lo = 0, hi = nnd;   // row_chunk[lo] <= c < row_chunk[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (row_chunk[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Store entries (i, j, k) from the database: G's upper triangle and H row-major (the store's 15
// values).  A G block that is not symmetric to 1e-12 of its largest value sets *asym.
__global__ void __launch_bounds__(kChunk)
    k_db_to_store(const double* __restrict__ db, int64_t N, const uint32_t* __restrict__ idx,
                  const int64_t* __restrict__ row_chunk, const int64_t* __restrict__ row_valid,
                  int nnd, double* __restrict__ vals, unsigned* __restrict__ asym)
{
    const int64_t c = blockIdx.x;
    const int i = chunk_row(row_chunk, nnd, c);
    const int lane = threadIdx.x;
    double v[kVals];
#pragma unroll
    for (int t = 0; t < kVals; t++) v[t] = 0.0;
    if ((c - row_chunk[i]) * kChunk + lane < row_valid[i]) {
        const uint32_t id = idx[c * kChunk + lane];
        const int j = (int)(id >> kKBits), k = (int)(id & kMaxLag);
        const int64_t NN = N * N;
        const double* G = db + (int64_t)k * 2 * NN + (int64_t)(3 * i) * N + 3 * j;
        const double* H = G + NN;
        double g[3][3], gmax = 0.0;
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                g[p][q] = G[(int64_t)p * N + q];
                gmax = fmax(gmax, fabs(g[p][q]));
                v[6 + 3 * p + q] = H[(int64_t)p * N + q];
            }
        v[0] = g[0][0];
        v[1] = g[0][1];
        v[2] = g[0][2];
        v[3] = g[1][1];
        v[4] = g[1][2];
        v[5] = g[2][2];
        const double d = fmax(fabs(g[0][1] - g[1][0]), fmax(fabs(g[0][2] - g[2][0]), fabs(g[1][2] - g[2][1])));
        if (d > 1e-12 * gmax) atomicOr(asym, 1u);
    }
    double* o = vals + c * (kVals * kChunk);
#pragma unroll
    for (int t = 0; t < kVals; t++) o[t * kChunk + lane] = v[t];
}

}  // namespace

void caller_drop(Context& c)
{
    c.caller_db.reset();
    if (c.caller_last > 0) c.store_ready = false;
    c.caller_last = 0;
    c.caller_stale = false;
}

void caller_update(Context& c, int step, int lower, int upper, const double* G0,
                   const double* G1, const double* H0, const double* H1)
{
    if (c.precision != 0)
        throw std::invalid_argument("caller-supplied lag matrices feed the FP64 march");
    if (c.kernel != 4)
        throw std::invalid_argument("the lag store is built for the Step kernel");
    if (step < 1) throw std::invalid_argument("step must be >= 1");
    if (lower < 0 || upper >= c.nnd || upper < lower)
        throw std::invalid_argument("row range out of bounds");
    if (!G0 || !G1 || !H0 || !H1) throw std::invalid_argument("null block");
    // UpdateGHLinear keeps nothing past MaxStep (src/SolverStepLinearT.cpp:65)
    if (step > c.Kc) return;
    const int64_t N = 3LL * c.nnd, NN = N * N;
    if (c.caller_last == 0) {
        // entering this mode: the database replaces a store the device built, and a march on it
        march_free(c);
        c.store_ready = false;
        c.tp.ready = false;
        c.caller_db.alloc((size_t)(c.Kc + 1) * 2 * NN);
        TD_HIP(hipMemsetAsync(c.caller_db.p, 0, c.caller_db.bytes(), c.stream));
    }
    const int64_t nrow = 3LL * (upper - lower + 1), n = nrow * N;
    DBuf<double> stage(c.stream);
    stage.alloc((size_t)(4 * n));
    const double* src[4] = {G0, G1, H0, H1};
    for (int b = 0; b < 4; b++)
        TD_HIP(hipMemcpyAsync(stage.p + b * n, src[b], n * sizeof(double), hipMemcpyHostToDevice,
                              c.stream));
    hipLaunchKernelGGL(k_db_update, dim3(grid_of(n)), dim3(256), 0, c.stream, c.caller_db.p, N,
                       3LL * lower, nrow, step, stage.p);
    TD_HIP(hipGetLastError());
    TD_HIP(hipStreamSynchronize(c.stream));
    c.caller_last = std::max(c.caller_last, step);
    c.caller_stale = true;
}

void caller_store(Context& c)
{
    if (c.caller_last == 0) throw std::logic_error("no lag matrices supplied");
    c.store_ready = false;
    const int64_t N = 3LL * c.nnd, nn = (int64_t)c.nnd * c.nnd;
    std::vector<int2> hull(nn);
    {
        DBuf<int2> d_hull(c.stream);
        d_hull.alloc(nn);
        hipLaunchKernelGGL(k_db_hull, dim3(grid_of(nn)), dim3(256), 0, c.stream, c.caller_db.p, N,
                           c.nnd, std::min(c.caller_last, c.Kc), d_hull.p);
        TD_HIP(hipGetLastError());
        TD_HIP(hipMemcpyAsync(hull.data(), d_hull.p, nn * sizeof(int2), hipMemcpyDeviceToHost,
                              c.stream));
        TD_HIP(hipStreamSynchronize(c.stream));
    }
    alloc_store_for_hull(c, hull.data());
    LagStore& S = c.store;
    DBuf<unsigned> asym(c.stream);
    asym.alloc(1);
    TD_HIP(hipMemsetAsync(asym.p, 0, sizeof(unsigned), c.stream));
    if (S.n_chunks > 0)
        hipLaunchKernelGGL(k_db_to_store, dim3((unsigned)S.n_chunks), dim3(kChunk), 0, c.stream,
                           c.caller_db.p, N, S.idx.p, S.row_chunk.p, S.row_valid.p, c.nnd,
                           S.vals.p, asym.p);
    TD_HIP(hipGetLastError());
    unsigned flag = 0;
    TD_HIP(hipMemcpyAsync(&flag, asym.p, sizeof(unsigned), hipMemcpyDeviceToHost, c.stream));
    TD_HIP(hipStreamSynchronize(c.stream));
    if (flag)
        throw std::invalid_argument(
            "UpdateGHLinear: a 3 x 3 node block of G is not symmetric (the lag store keeps the "
            "upper triangle of G's blocks)");
    c.store_ready = true;
    c.store_precision = 0;
    c.store_lags = c.Kc;
    c.caller_stale = false;
}

}  // namespace tdbem
