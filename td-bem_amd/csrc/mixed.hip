// mixed.hip — the mixed-precision history convolution: split-BF16 (three terms) MFMA products
// accumulated in FP32 inside a 32-wide k-tile and in FP64 across tiles, blocks and lags, for R
// right-hand sides at once (R = 16, 32, 48 or 64).
//
// Layout.  Nodes are reordered along a Morton curve and grouped into tiles of 16 (48 dofs).  For
// each pair of tiles (I, J) the lags k in the union of their node-pair shells are stored as dense
// 48 x 96 blocks A[(3 ii + p)][(6 jj + 3 part + q)] with part 0 = G[p][q] (times z^t) and
// part 1 = -H[p][q] (times z^u), each value split as a1 + a2 + a3 (bf16, 24 significant bits) and
// laid out in MFMA fragment order, so a wavefront loads each 16x32 operand as one 1 KiB line.
// The history ring (FP64) is mirrored as split-BF16 fragments per (slot, column tile).
//
// Per block the products a1 z1 + a1 z2 + a2 z1 + a1 z3 + a2 z2 + a3 z1 (all terms down to 2^-24
// relative) run on v_mfma_f32_16x16x32_bf16; per 32-wide k-tile the leading product and the sum of
// the corrections leave FP32 for FP64 accumulators, so rounding does not grow with the number of
// blocks and lags summed.
#include <algorithm>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "tdbem_internal.h"

namespace tdbem {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

namespace {

constexpr int kTile = 16;                 // nodes per tile
constexpr int kRows = 3 * kTile;          // 48
constexpr int kCols = 6 * kTile;          // 96
constexpr int kBlockElems = 3 * 3 * 3 * 512;  // terms x m-tiles x k-tiles x (64 lanes x 8)
constexpr int kSplitWaves = 4;

__device__ __forceinline__ unsigned short bf16_rne(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_float(unsigned short b)
{
    return __uint_as_float(((unsigned int)b) << 16);
}

// v = b1 + b2 + b3 + O(2^-24 |v|), each a bf16; residuals formed exactly in FP64.
__device__ __forceinline__ void split3(double v, unsigned short* b)
{
    b[0] = bf16_rne((float)v);
    const double r1 = v - (double)bf16_to_float(b[0]);
    b[1] = bf16_rne((float)r1);
    const double r2 = r1 - (double)bf16_to_float(b[1]);
    b[2] = bf16_rne((float)r2);
}

// Position of A[r][c] of term t inside a block (16x16x32 A-operand: lane = row + 16 * (k / 8)).
__device__ __forceinline__ int a_offset(int t, int r, int c)
{
    const int m = r >> 4, i = r & 15, kt = c >> 5, kk = c & 31;
    return ((t * 3 + m) * 3 + kt) * 512 + (i + 16 * (kk >> 3)) * 8 + (kk & 7);
}

__device__ __forceinline__ int row_of_chunk(const int64_t* row_chunk, int nnd, int64_t c)
{
    int lo = 0, hi = nnd;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (row_chunk[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// FP64 node-pair store -> split-BF16 tile blocks.  One 64-thread block per store chunk.
__global__ void __launch_bounds__(kChunk)
    k_tp_fill(const double* __restrict__ vals, const uint32_t* __restrict__ idx,
              const int64_t* __restrict__ row_chunk, const int64_t* __restrict__ row_valid,
              int nnd, const int* __restrict__ pos, int nT, const int64_t* __restrict__ pair_off,
              const int* __restrict__ pair_lo, unsigned short* __restrict__ tp)
{
    const int64_t c = blockIdx.x;
    const int i = row_of_chunk(row_chunk, nnd, c);
    const int lane = threadIdx.x;
    if ((c - row_chunk[i]) * kChunk + lane >= row_valid[i]) return;
    const uint32_t id = idx[c * kChunk + lane];
    const int j = (int)(id >> kKBits), k = (int)(id & kMaxLag);
    const int pi = pos[i], pj = pos[j];
    const int I = pi / kTile, ii = pi % kTile, J = pj / kTile, jj = pj % kTile;
    const int64_t base = pair_off[(int64_t)I * nT + J];
    const int64_t blk = base + (k - pair_lo[(int64_t)I * nT + J]);
    unsigned short* B = tp + blk * kBlockElems;
    double v[kVals];
    const double* s = vals + c * (kVals * kChunk) + lane;
#pragma unroll
    for (int t = 0; t < kVals; t++) v[t] = s[t * kChunk];
    const int sym[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            unsigned short g[3], h[3];
            split3(v[sym[p][q]], g);
            split3(-v[6 + 3 * p + q], h);
            const int r = 3 * ii + p;
#pragma unroll
            for (int t = 0; t < 3; t++) {
                B[a_offset(t, r, 6 * jj + q)] = g[t];
                B[a_offset(t, r, 6 * jj + 3 + q)] = h[t];
            }
        }
}

// Split-BF16 copy of ring slot `slot` in B-operand order (lane = rhs + 16 * (k / 8)).
__global__ void k_zsplit(const double* __restrict__ ring, int N, int nrhs, int nT,
                         const int* __restrict__ perm, int slot, int NT,
                         unsigned short* __restrict__ zs)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nT * kCols * nrhs) return;
    const int r = (int)(t % nrhs);
    const int c = (int)((t / nrhs) % kCols);
    const int J = (int)(t / ((int64_t)nrhs * kCols));
    const int node = perm[J * kTile + c / 6];
    const int part = (c % 6) / 3, q = c % 3;
    // block column part 0 multiplies z^t (ring part 1), part 1 multiplies z^u (ring part 0)
    const int rpart = 1 - part;
    const double v = node < 0 ? 0.0 : ring[((int64_t)(slot * nrhs + r) * 2 + rpart) * N + 3 * node + q];
    unsigned short b[3];
    split3(v, b);
    const int kt = c >> 5, kk = c & 31, nt = r >> 4, n = r & 15;
    const int lane = n + 16 * (kk >> 3), e = kk & 7;
    unsigned short* base = zs + (((int64_t)slot * nT + J) * 3) * 3 * NT * 512;
#pragma unroll
    for (int tm = 0; tm < 3; tm++) base[((tm * 3 + kt) * NT + nt) * 512 + lane * 8 + e] = b[tm];
}

template <int NT>
__global__ void __launch_bounds__(64 * kSplitWaves)
    k_conv_mixed(const bf16x8* __restrict__ tp, const uint32_t* __restrict__ bidx,
                 const int64_t* __restrict__ row_blk, const bf16x8* __restrict__ zs, int nT,
                 int Rring, int s, int nsplit, double* __restrict__ partial)
{
    const int I = blockIdx.x, sp = blockIdx.y;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t b0 = row_blk[I], b1 = row_blk[I + 1];
    const int chunks = nsplit * kSplitWaves;
    const int64_t per = (b1 - b0 + chunks - 1) / chunks;
    const int64_t lo = b0 + (int64_t)(sp * kSplitWaves + w) * per;
    const int64_t hi = min(b1, lo + per);
    double acc[3][NT][4];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[m][n][v] = 0.0;
    for (int64_t b = lo; b < hi; b++) {
        const uint32_t id = bidx[b];
        const int J = (int)(id >> kKBits), k = (int)(id & kMaxLag);
        int slot = (s - k) % Rring;
        if (slot < 0) slot += Rring;
        const bf16x8* A = tp + b * (kBlockElems / 8);
        const bf16x8* Z = zs + (((int64_t)slot * nT + J) * 3) * 3 * NT * 64;
#pragma unroll
        for (int kt = 0; kt < 3; kt++) {
            bf16x8 a[3][3], z[3][NT];
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int m = 0; m < 3; m++) a[t][m] = A[((t * 3 + m) * 3 + kt) * 64 + lane];
#pragma unroll
            for (int t = 0; t < 3; t++)
#pragma unroll
                for (int n = 0; n < NT; n++) z[t][n] = Z[((t * 3 + kt) * NT + n) * 64 + lane];
            // The leading product a1 z1 and the five corrections (2^-8 smaller and below) are
            // summed in separate FP32 registers and both moved to FP64 after every 32-wide
            // k-tile: each FP32 rounding then acts on one 32-term dot product only.
#pragma unroll
            for (int m = 0; m < 3; m++)
#pragma unroll
                for (int n = 0; n < NT; n++) {
                    const f32x4 zero = {0.f, 0.f, 0.f, 0.f};
                    const f32x4 dm =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][m], z[0][n], zero, 0, 0, 0);
                    f32x4 dc =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2][m], z[0][n], zero, 0, 0, 0);
                    dc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][m], z[1][n], dc, 0, 0, 0);
                    dc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][m], z[2][n], dc, 0, 0, 0);
                    dc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1][m], z[0][n], dc, 0, 0, 0);
                    dc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0][m], z[1][n], dc, 0, 0, 0);
#pragma unroll
                    for (int v = 0; v < 4; v++)
                        acc[m][n][v] += (double)dm[v] + (double)dc[v];
                }
        }
    }
    double* out = partial + (((int64_t)I * chunks + sp * kSplitWaves + w) * 3 * NT) * 256;
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int n = 0; n < NT; n++)
#pragma unroll
            for (int v = 0; v < 4; v++) out[((m * NT + n) * 64 + lane) * 4 + v] = acc[m][n][v];
}

// hist[r][dof] = sum over the partials of a row tile (fixed order).  16x16 output fragment:
// lane = col + 16 * (row / 4), element row % 4.
__global__ void k_tp_reduce(const double* __restrict__ partial, int nT, int chunks, int NT,
                            const int* __restrict__ perm, int N, int nrhs,
                            double* __restrict__ hist)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nT * kRows * nrhs) return;
    const int r = (int)(t % nrhs);
    const int row = (int)((t / nrhs) % kRows);
    const int I = (int)(t / ((int64_t)nrhs * kRows));
    const int node = perm[I * kTile + row / 3];
    if (node < 0) return;
    const int m = row >> 4, i = row & 15, nt = r >> 4, n = r & 15;
    const int lane = n + 16 * (i >> 2), v = i & 3;
    double sum = 0.0;
    for (int c = 0; c < chunks; c++)
        sum += partial[((((int64_t)I * chunks + c) * 3 * NT + m * NT + nt) * 64 + lane) * 4 + v];
    hist[(int64_t)r * N + 3 * node + row % 3] = sum;
}

// One 16x16x32 bf16 product with the lane layouts assumed above, for the layout self-test.
__global__ void k_mfma_selftest(const float* __restrict__ A, const float* __restrict__ B,
                                float* __restrict__ D)
{
    const int lane = threadIdx.x;
    bf16x8 a, b;
    for (int e = 0; e < 8; e++) {
        const int kk = 8 * (lane >> 4) + e;
        a[e] = (short)bf16_rne(A[(lane & 15) * 32 + kk]);
        b[e] = (short)bf16_rne(B[kk * 16 + (lane & 15)]);
    }
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, d, 0, 0, 0);
    for (int v = 0; v < 4; v++) D[(4 * (lane >> 4) + v) * 16 + (lane & 15)] = d[v];
}

uint32_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    auto spread = [](uint32_t v) {
        v &= 0x3FF;
        v = (v | (v << 16)) & 0x030000FF;
        v = (v | (v << 8)) & 0x0300F00F;
        v = (v | (v << 4)) & 0x030C30C3;
        v = (v | (v << 2)) & 0x09249249;
        return v;
    };
    return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

}  // namespace

void tp_build(Context& c)
{
    if (!c.store_ready) build_lag_store(c);
    TPStore& T = c.tp;
    if (T.ready) return;
    const int nnd = c.nnd;
    // Morton order of the nodes
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int i = 0; i < nnd; i++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], c.h_coords[3 * i + d]);
            hi[d] = std::max(hi[d], c.h_coords[3 * i + d]);
        }
    std::vector<uint32_t> key(nnd);
    for (int i = 0; i < nnd; i++) {
        uint32_t q[3];
        for (int d = 0; d < 3; d++) {
            const double span = hi[d] - lo[d];
            q[d] = span > 0 ? (uint32_t)std::min(1023.0, (c.h_coords[3 * i + d] - lo[d]) / span * 1023.0)
                            : 0u;
        }
        key[i] = morton3(q[0], q[1], q[2]);
    }
    std::vector<int> order(nnd);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    T.nT = (nnd + kTile - 1) / kTile;
    T.perm.assign((size_t)T.nT * kTile, -1);
    T.pos.assign(nnd, 0);
    for (int p = 0; p < nnd; p++) {
        T.perm[p] = order[p];
        T.pos[order[p]] = p;
    }
    // tile-pair lag hulls from the node-pair hulls
    const int nT = T.nT;
    std::vector<int> plo((size_t)nT * nT, 1 << 30), phi((size_t)nT * nT, -1);
    const auto& H = c.store.h_hull;
    for (int i = 0; i < nnd; i++) {
        const int I = T.pos[i] / kTile;
        for (int j = 0; j < nnd; j++) {
            const int2 h = H[(size_t)i * nnd + j];
            if (h.y < h.x) continue;
            const size_t q = (size_t)I * nT + T.pos[j] / kTile;
            plo[q] = std::min(plo[q], h.x);
            phi[q] = std::max(phi[q], h.y);
        }
    }
    std::vector<int64_t> off((size_t)nT * nT, -1);
    std::vector<uint32_t> bidx;
    T.h_row_blk.assign(nT + 1, 0);
    int64_t nb = 0;
    for (int I = 0; I < nT; I++) {
        T.h_row_blk[I] = nb;
        for (int J = 0; J < nT; J++) {
            const size_t q = (size_t)I * nT + J;
            if (phi[q] < plo[q]) continue;
            off[q] = nb;
            for (int k = plo[q]; k <= phi[q]; k++) {
                bidx.push_back(((uint32_t)J << kKBits) | (uint32_t)k);
                nb++;
            }
        }
    }
    T.h_row_blk[nT] = nb;
    T.nblocks = nb;
    T.d_perm.upload(T.perm.data(), T.perm.size(), c.stream);
    T.d_pos.upload(T.pos.data(), T.pos.size(), c.stream);
    T.bidx.upload(bidx.data(), bidx.size(), c.stream);
    T.row_blk.upload(T.h_row_blk.data(), T.h_row_blk.size(), c.stream);
    DBuf<int64_t> d_off;
    DBuf<int> d_plo;
    d_off.upload(off.data(), off.size(), c.stream);
    d_plo.upload(plo.data(), plo.size(), c.stream);
    T.vals.alloc((size_t)std::max<int64_t>(nb, 1) * kBlockElems);
    TD_HIP(hipMemsetAsync(T.vals.p, 0, T.vals.bytes(), c.stream));
    if (c.store.n_chunks > 0)
        hipLaunchKernelGGL(k_tp_fill, dim3((unsigned)c.store.n_chunks), dim3(kChunk), 0, c.stream,
                           c.store.vals.p, c.store.idx.p, c.store.row_chunk.p, c.store.row_valid.p,
                           nnd, T.d_pos.p, nT, d_off.p, d_plo.p, T.vals.p);
    TD_HIP(hipGetLastError());
    TD_HIP(hipStreamSynchronize(c.stream));
    T.ready = true;
}

void tp_zsplit(Context& c, const double* ring, int N, int nrhs, int slot, unsigned short* zs)
{
    const TPStore& T = c.tp;
    const int64_t n = (int64_t)T.nT * kCols * nrhs;
    hipLaunchKernelGGL(k_zsplit, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, ring,
                       N, nrhs, T.nT, T.d_perm.p, slot, nrhs / 16, zs);
    TD_HIP(hipGetLastError());
}

int64_t tp_zslot_elems(const Context& c, int nrhs) { return (int64_t)c.tp.nT * 9 * (nrhs / 16) * 512; }

int tp_nsplit(const Context& c)
{
    // about 8192 wavefronts in flight over the row tiles
    return std::max(1, std::min(64, 2048 / std::max(1, c.tp.nT)));
}

void tp_conv(Context& c, const unsigned short* zs, int nrhs, int Rring, int s, double* partial)
{
    const TPStore& T = c.tp;
    const int nsplit = tp_nsplit(c);
    const int NT = nrhs / 16;
    dim3 grid(T.nT, nsplit);
    const bf16x8* tpv = reinterpret_cast<const bf16x8*>(T.vals.p);
    const bf16x8* z = reinterpret_cast<const bf16x8*>(zs);
    switch (NT) {
    case 1: hipLaunchKernelGGL(k_conv_mixed<1>, grid, dim3(64 * kSplitWaves), 0, c.stream, tpv, T.bidx.p, T.row_blk.p, z, T.nT, Rring, s, nsplit, partial); break;
    case 2: hipLaunchKernelGGL(k_conv_mixed<2>, grid, dim3(64 * kSplitWaves), 0, c.stream, tpv, T.bidx.p, T.row_blk.p, z, T.nT, Rring, s, nsplit, partial); break;
    case 3: hipLaunchKernelGGL(k_conv_mixed<3>, grid, dim3(64 * kSplitWaves), 0, c.stream, tpv, T.bidx.p, T.row_blk.p, z, T.nT, Rring, s, nsplit, partial); break;
    case 4: hipLaunchKernelGGL(k_conv_mixed<4>, grid, dim3(64 * kSplitWaves), 0, c.stream, tpv, T.bidx.p, T.row_blk.p, z, T.nT, Rring, s, nsplit, partial); break;
    default: throw std::runtime_error("mixed march needs nrhs in {16, 32, 48, 64}");
    }
    TD_HIP(hipGetLastError());
}

void tp_reduce(Context& c, const double* partial, int nrhs, double* hist, int N)
{
    const TPStore& T = c.tp;
    const int chunks = tp_nsplit(c) * kSplitWaves;
    const int NT = nrhs / 16;
    const int64_t n = (int64_t)T.nT * kRows * nrhs;
    hipLaunchKernelGGL(k_tp_reduce, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream,
                       partial, T.nT, chunks, NT, T.d_perm.p, N, nrhs, hist);
    TD_HIP(hipGetLastError());
}

int64_t tp_partial_elems(const Context& c, int nrhs)
{
    return (int64_t)c.tp.nT * tp_nsplit(c) * kSplitWaves * 3 * (nrhs / 16) * 256;
}

void mfma_selftest(const float* hA, const float* hB, float* hD)
{
    DBuf<float> A, B, D;
    A.upload(hA, 16 * 32, nullptr);
    B.upload(hB, 32 * 16, nullptr);
    D.alloc(16 * 16);
    hipLaunchKernelGGL(k_mfma_selftest, dim3(1), dim3(64), 0, nullptr, A.p, B.p, D.p);
    TD_HIP(hipGetLastError());
    TD_HIP(hipMemcpy(hD, D.p, 16 * 16 * sizeof(float), hipMemcpyDeviceToHost));
}

}  // namespace tdbem
