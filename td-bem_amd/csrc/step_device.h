// step_device.h — device code of one marching step, shared by the single-context kernels
// (march.hip, mixed.hip, lu.hip) and the batched kernels over many scenarios (batch.hip).
#pragma once

#include "tdbem_internal.h"

namespace tdbem {

typedef __attribute__((__vector_size__(8 * sizeof(short)))) short bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;
typedef double f64x4 __attribute__((ext_vector_type(4)));

constexpr int kTile = 16;                 // nodes per tile of the mixed-precision store
constexpr int kRows = 3 * kTile;          // 48
constexpr int kCols = 6 * kTile;          // 96
constexpr int kSubFloats = 512;           // FP32 per stored sub-tile: [2][64 lanes][4]
constexpr int kApplyWaves = 4;            // wavefronts per step-solve workgroup

// f_r(t): src/SolverT.cpp:379-394 returns 1; a per-RHS amplitude and a linear rise over t_rise
// (0 = Heaviside) generalise it.
__device__ __forceinline__ double load_factor(double amp, double rise, double t)
{
    return rise > 0 ? amp * (t < rise ? t / rise : 1.0) : amp;
}

__device__ __forceinline__ unsigned short bf16_rne(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

// Workgroup b of a grid of nb is dispatched to XCD b % 8 (round robin over the eight XCDs of an
// MI355X, each with its own L2).  Its logical chunk: XCD x takes the x-th contiguous eighth of
// the chunks, in order, so the workgroups resident together on one XCD walk neighbouring row-tile
// groups and share their history blocks in that XCD's L2.  A bijection of [0, nb).
__device__ __forceinline__ int xcd_chunk(int b, int nb)
{
    constexpr int kXcd = 8;
    const int q = nb / kXcd, r = nb % kXcd, x = b % kXcd, i = b / kXcd;
    return x * q + min(x, r) + i;
}

// x = hi + mid + lo exactly for an FP32 x: hi keeps the leading 8 significant bits (truncated),
// mid the next 8 of the remainder x - hi, lo the rest (at most 8 bits, so it is a bf16 exactly).
// Eight values, as two f32x4 in 16x16x32 A-operand order, become three bf16x8 operands.
__device__ __forceinline__ void split_f32x8(const f32x4& a0, const f32x4& a1, bf16x8& hi,
                                            bf16x8& mid, bf16x8& lo)
{
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const float x = e < 4 ? a0[e] : a1[e - 4];
        const unsigned xb = __float_as_uint(x);
        const float r1 = x - __uint_as_float(xb & 0xFFFF0000u);
        const unsigned rb = __float_as_uint(r1);
        const float r2 = r1 - __uint_as_float(rb & 0xFFFF0000u);
        hi[e] = (short)(xb >> 16);
        mid[e] = (short)(rb >> 16);
        lo[e] = (short)(__float_as_uint(r2) >> 16);
    }
}

// Element t (< nT * 96 * nrhs) of the copy of ring slot `slot`, rounded to FP32 and split
// exactly into three bf16 terms as split_f32x8 splits a value (once per step here, not in every
// workgroup that stages it), in B-operand order (lane = rhs + 16 * (k / 8), element k % 8),
// laid out [slot][J][k-tile][rhs tile][term][64 lanes][8].  The ring holds [slot][part][N][rhs]
// with part 0 = u, part 1 = t; block column part 0 multiplies z^t.
__device__ __forceinline__ void zcopy_elem(const double* __restrict__ ring, int N, int nrhs,
                                           int nT, const int* __restrict__ perm, int slot,
                                           int NT, unsigned short* __restrict__ zs, int64_t t)
{
    const int r = (int)(t % nrhs);
    const int c = (int)((t / nrhs) % kCols);
    const int J = (int)(t / ((int64_t)nrhs * kCols));
    const int node = perm[J * kTile + c / 6];
    const int part = (c % 6) / 3, q = c % 3;
    const int rpart = 1 - part;
    const double v =
        node < 0 ? 0.0 : ring[(((int64_t)slot * 2 + rpart) * N + 3 * node + q) * nrhs + r];
    const int kt = c >> 5, kk = c & 31, nt = r >> 4, n = r & 15;
    const int lane = n + 16 * (kk >> 3), e = kk & 7;
    const float x = (float)v;
    const unsigned xb = __float_as_uint(x);
    const float r1 = x - __uint_as_float(xb & 0xFFFF0000u);
    const unsigned rb = __float_as_uint(r1);
    const float r2 = r1 - __uint_as_float(rb & 0xFFFF0000u);
    const int64_t base = ((((int64_t)slot * nT + J) * 3 + kt) * NT + nt) * 3;
    zs[((base + 0) * 64 + lane) * 8 + e] = (unsigned short)(xb >> 16);
    zs[((base + 1) * 64 + lane) * 8 + e] = (unsigned short)(rb >> 16);
    zs[((base + 2) * 64 + lane) * 8 + e] = (unsigned short)(__float_as_uint(r2) >> 16);
}

// Lags below kNearLags are the near lags of the two-step (time-blocked) convolution: their
// history at step s + 1 is not known at step s (z_0, z_1 and y_{s-1} need x_s), so they are
// convolved at every step for that step alone; a lag k >= kNearLags is convolved every other
// step, at the steps s with s = k (mod 2), for both s (history slot s - k) and s + 1 (slot
// s + 1 - k, complete at step s since s + 1 - k <= s - 2).  Each stored sub-tile is thus read
// once per two steps.
constexpr int kNearLags = 3;

// One workgroup's share of a convolution chunk: the steps [e0, e1) of one of its group's merged
// lists (the near lags and the far lags of one parity), at step s, for the 16 load cases of
// right-hand-side tile nt0 of NT.  Wavefront w takes row tile w of the group and both target
// steps, s and s + 1 (s alone for a near lag), and writes its 48 x 16 FP64 partial sums of
// target t to the columns of tile nt0 of out[t] ([3][NT][64][4]) unless the row tile does not
// exist.  The NT workgroups of a chunk run next to each other on one XCD and read the same
// stored sub-tiles at about the same time, so those come from HBM once and from L2 after.
// All wavefronts of the workgroup call it together (a barrier per step): the history k-tiles of
// a step (for each target) are loaded once per workgroup into one of two LDS buffers, one step
// ahead of their use, and shared by the group's row tiles; each wavefront streams its own stored
// FP32 sub-tiles and splits each into three bf16 operands in registers (split_f32x8) just before
// its products, which serve both targets.
__device__ __forceinline__ void conv_chunk(const f32x4* __restrict__ tp,
                                           const TPEntry* __restrict__ ent,
                                           const bf16x8* __restrict__ zs, int nT, int NT, int nt0,
                                           int Rring, int s, int64_t e0, int64_t e1, int w,
                                           bool write, double* __restrict__ out0,
                                           double* __restrict__ out1)
{
    // one (slot, column tile) history block: in HBM, [k-tile][rhs tile][term] lines of 64 lanes
    // x 8 bf16 (zcopy_elem); in LDS (the workgroup's rhs tile), [term][k-tile] lines, one block
    // per target
    constexpr int kUnits = 3;                              // per target: k-tiles
    constexpr int kPer = (2 * kUnits + kGroup - 1) / kGroup;  // units per wavefront
    __shared__ bf16x8 zl[2][2][9 * 64];
    // the store, the lists and the history are global memory: loads through global (not flat)
    // instructions even where the pointers come from a descriptor table, so they count against
    // the vector-memory counter only and an LDS wait does not wait for them
    typedef const __attribute__((address_space(1))) f32x4 gf32x4;
    typedef const __attribute__((address_space(1))) bf16x8 gbf16x8;
    typedef const __attribute__((address_space(1))) TPEntry gentry;
    gf32x4* tpg = (gf32x4*)tp;
    gbf16x8* zsg = (gbf16x8*)zs;
    gentry* entg = (gentry*)ent;
    const int lane = threadIdx.x & 63;
    // w is the same in every lane of a wavefront: the entry fields are scalar loads
    w = __builtin_amdgcn_readfirstlane(w);
    double acc[2][3][4];
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
        for (int m = 0; m < 3; m++)
#pragma unroll
            for (int v = 0; v < 4; v++) acc[t][m][v] = 0.0;
    // wavefront w stages units w, w + 4 of an entry's blocks; unit U = target * 3 + k-tile; the
    // target-(s + 1) block of a near lag is not staged
    bf16x8 zn[kPer][3];
    const int64_t zblock = 3LL * NT * 3 * 64;   // bf16x8 per (slot, column tile)
    auto zload = [&](uint32_t id, unsigned kmask) {
        const int J = (int)(id >> kKBits), k = (int)(id & kMaxLag);
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const int U = q * kGroup + w;
            const int t = U / kUnits, kt = U % kUnits;
            if (U < 2 * kUnits && ((kmask >> kt) & 1u) && !(t == 1 && k < kNearLags)) {
                int slot = (s + t - k) % Rring;
                if (slot < 0) slot += Rring;
                gbf16x8* Zg = zsg + ((int64_t)slot * nT + J) * zblock +
                              (int64_t)(kt * NT + nt0) * 3 * 64 + lane;
                zn[q][0] = Zg[0];
                zn[q][1] = Zg[64];
                zn[q][2] = Zg[128];
            }
        }
    };
    auto zstore = [&](int buf, uint32_t id, unsigned kmask) {
        const int k = (int)(id & kMaxLag);
#pragma unroll
        for (int q = 0; q < kPer; q++) {
            const int U = q * kGroup + w;
            const int t = U / kUnits, kt = U % kUnits;
            if (U < 2 * kUnits && ((kmask >> kt) & 1u) && !(t == 1 && k < kNearLags)) {
#pragma unroll
                for (int term = 0; term < 3; term++)
                    zl[buf][t][(term * 3 + kt) * 64 + lane] = zn[q][term];
            }
        }
    };
    // this wavefront's stored sub-tiles of an entry go to a[u] by sub-tile position u; the next
    // entry's sub-tiles of k-tile kt are requested as soon as the products of k-tile kt have
    // consumed a[kt], a[3+kt], a[6+kt], so their latency overlaps the rest of the entry and the
    // barrier.  With one rhs tile (NT = 1) nothing else reads them: non-temporal loads, so they
    // do not evict history blocks; otherwise the chunk's other workgroups read them from L2.
    f32x4 a[9][2];
    const bool shared = NT > 1;
    auto aload = [&](unsigned mk, uint32_t first, int kt) {
        gf32x4* A = tpg + (int64_t)first * (kSubFloats / 4) + lane;
#pragma unroll
        for (int m = 0; m < 3; m++) {
            const int u = 3 * m + kt;
            if ((mk >> u) & 1u) {
                const int q = __builtin_popcount(mk & ((1u << u) - 1u));
#pragma unroll
                for (int h = 0; h < 2; h++)
                    a[u][h] = shared ? A[(q * 2 + h) * 64]
                                     : __builtin_nontemporal_load(&A[(q * 2 + h) * 64]);
            }
        }
    };
    if (e0 < e1) {
        const gentry& E0 = entg[e0];
        zload(E0.id, E0.kmask);
        zstore(0, E0.id, E0.kmask);
        const unsigned m0 = E0.msk[w];
        const uint32_t f0 = E0.sub[w];
#pragma unroll
        for (int kt = 0; kt < 3; kt++) aload(m0, f0, kt);
    }
    __syncthreads();
    for (int64_t e = e0; e < e1; e++) {
        const int cur = (int)((e - e0) & 1);
        const gentry& E = entg[e];
        const unsigned msk = E.msk[w];
        const bool far = (int)(E.id & kMaxLag) >= kNearLags;
        const bool more = e + 1 < e1;
        unsigned knext = 0, mnext = 0;
        uint32_t fnext = 0, idnext = 0;
        if (more) {
            const gentry& En = entg[e + 1];
            idnext = En.id;
            knext = En.kmask;
            mnext = En.msk[w];
            fnext = En.sub[w];
            zload(idnext, knext);
        }
        // The leading product and the corrections accumulate apart in FP32, over the entry's
        // three k-tiles (96 columns), and then move to FP64.
        f32x4 dm[2][3], dc[2][3];
#pragma unroll
        for (int t = 0; t < 2; t++)
#pragma unroll
            for (int m = 0; m < 3; m++) dm[t][m] = dc[t][m] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kt = 0; kt < 3; kt++) {
            const unsigned rows = (msk >> kt) & 0x49u;   // bits kt, 3 + kt, 6 + kt
            if (rows) {
                bf16x8 z[2][3];
#pragma unroll
                for (int t = 0; t < 2; t++)
#pragma unroll
                    for (int q = 0; q < 3; q++) z[t][q] = zl[cur][t][(q * 3 + kt) * 64 + lane];
#pragma unroll
                for (int m = 0; m < 3; m++) {
                    const int u = 3 * m + kt;
                    if (!((msk >> u) & 1u)) continue;
                    bf16x8 a1, a2, a3;
                    split_f32x8(a[u][0], a[u][1], a1, a2, a3);
                    // the leading product a1 z1 and the seven corrections (smallest first) sum
                    // in separate FP32 registers.  Dropped: a3 z3, below 2^-28 of |a| |z|.
#pragma unroll
                    for (int t = 0; t < 2; t++) {
                        if (t == 1 && !far) continue;
                        f32x4 d = dm[t][m], c = dc[t][m];
                        d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, z[t][0], d, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, z[t][1], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, z[t][2], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, z[t][2], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a3, z[t][0], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, z[t][1], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a2, z[t][0], c, 0, 0, 0);
                        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a1, z[t][1], c, 0, 0, 0);
                        dm[t][m] = d;
                        dc[t][m] = c;
                    }
                }
            }
            if (more) aload(mnext, fnext, kt);
        }
#pragma unroll
        for (int t = 0; t < 2; t++) {
            if (t == 1 && !far) continue;
#pragma unroll
            for (int m = 0; m < 3; m++) {
                if (!((msk >> (3 * m)) & 7u)) continue;
#pragma unroll
                for (int v = 0; v < 4; v++)
                    acc[t][m][v] += (double)dm[t][m][v] + (double)dc[t][m][v];
            }
        }
        if (more) zstore(cur ^ 1, idnext, knext);
        __syncthreads();
    }
    if (!write) return;
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int64_t o = ((int64_t)(m * NT + nt0) * 64 + lane) * 4 + v;
            out0[o] = acc[0][m][v];
            out1[o] = acc[1][m][v];
        }
}

// Element t (< nT * 48 * nrhs) of the history sums of step s (sum0) and of the part of step
// s + 1 convolved at step s (sum1): fixed-order sums over the chunks of the row tile's group that
// ran at step s (those of the lists of lag parity p = s mod 2).  Partial of chunk c, target tg,
// row tile w of its group: [c][tg][w][3][NT][64][4], 16x16 output fragment lane = col + 16 *
// (row / 4), element row % 4; t runs through a row tile's partials in that order, so the reads
// of consecutive threads are contiguous.  grp_chunk: [2][nG + 1], the first chunk of each group
// in the lists of parity 0 and 1.  Returns false for the padding rows of the last tile (no dof).
__device__ __forceinline__ bool reduce_sum(const double* __restrict__ partial,
                                           const int* __restrict__ grp_chunk, int nG, int p,
                                           int NT, const int* __restrict__ perm, int nrhs,
                                           int64_t t, int& r, int& dof, double& sum0,
                                           double& sum1)
{
    const int per = kRows * nrhs;   // = 3 NT 256
    const int I = (int)(t / per), o = (int)(t % per);
    const int v = o & 3, lane = (o >> 2) & 63, mn = o >> 8;
    const int m = mn / NT, nt = mn % NT;
    const int row = 16 * m + 4 * (lane >> 4) + v;
    r = 16 * nt + (lane & 15);
    const int node = perm[I * kTile + row / 3];
    if (node < 0) return false;
    dof = 3 * node + row % 3;
    const int g = I / kGroup, w = I % kGroup;
    double a0 = 0.0, a1 = 0.0;
    const int* gc = grp_chunk + p * (nG + 1);
    for (int c = gc[g]; c < gc[g + 1]; c++) {
        a0 += partial[(((int64_t)c * 2) * kGroup + w) * per + o];
        a1 += partial[(((int64_t)c * 2 + 1) * kGroup + w) * per + o];
    }
    sum0 = a0;
    sum1 = a1;
    return true;
}

// Element t (< N * nrhs): r = f_r gF + gU + hist, written in the fragment order of the step solve.
__device__ __forceinline__ double rhs_value(double fac, double gF, double gU, double hist)
{
    return fac * gF + gU + hist;
}

__device__ __forceinline__ void rhs_elem(const double* __restrict__ gF,
                                         const double* __restrict__ gU,
                                         const double* __restrict__ hist,
                                         const double* __restrict__ fac, int N,
                                         double* __restrict__ Xf, int64_t t)
{
    const int r = (int)(t / N), d = (int)(t % N);
    Xf[frag_index(N, r, d)] = rhs_value(fac[r], gF[d], gU[d], hist[t]);
}

// reduce_sum and rhs_elem in one pass: the summed history goes straight into the step solve's
// right-hand side (every dof of every rhs is one element t of the reduction).  carry[t] holds
// the part of step s's history convolved at step s - 1 and receives the part of step s + 1's
// convolved now (the same thread reads, then writes, element t).
__device__ __forceinline__ void reduce_rhs_elem(const double* __restrict__ partial,
                                                const int* __restrict__ grp_chunk, int nG, int p,
                                                int NT,
                                                const int* __restrict__ perm, int N, int nrhs,
                                                const double* __restrict__ gF,
                                                const double* __restrict__ gU,
                                                const double* __restrict__ fac,
                                                double* __restrict__ carry,
                                                double* __restrict__ Xf, int64_t t)
{
    int r, d;
    double sum0, sum1;
    if (!reduce_sum(partial, grp_chunk, nG, p, NT, perm, nrhs, t, r, d, sum0, sum1)) return;
    sum0 += carry[t];
    carry[t] = sum1;
    Xf[frag_index(N, r, d)] = rhs_value(fac[r], gF[d], gU[d], sum0);
}

// Chunk of a convolution plan: (group | lag parity << 24, first entry, end entry, scenario).
constexpr int kChunkParityShift = 24;

// Wavefront g (< frag_rows(n) / 16 * nsplit) of parts[p][r][i] = sum over the columns j of
// split p of A[i][j] X[r][j], A and X in fragment order: per 32-column fragment it loads 4 KiB of
// A (64 B contiguous per lane) and NT fragments of X and runs 8 v_mfma_f64_16x16x4 per rhs tile.
template <int NT>
__device__ __forceinline__ void apply_wave(const double* __restrict__ Af, int n,
                                           const double* __restrict__ Xf, int nrhs, int nsplit,
                                           int g, double* __restrict__ parts)
{
    const int lane = threadIdx.x & 63;
    const int nTr = frag_rows(n) / 16, C = frag_cols(n);
    const int cper = (C + nsplit - 1) / nsplit;
    const int T = g % nTr, p = g / nTr;
    const int c0 = p * cper, c1 = min(C, c0 + cper);
    f64x4 acc[NT];
#pragma unroll
    for (int q = 0; q < NT; q++) acc[q] = f64x4{0.0, 0.0, 0.0, 0.0};
    const double* a = Af + ((int64_t)T * C) * 512 + lane * 8;
    const double* x = Xf + lane * 8;
    // two fragments in flight: the loads of fragment c + 1 are issued before the products of c
    double av[2][8], xv[2][NT][8];
    auto load = [&](int b, int c) {
#pragma unroll
        for (int e = 0; e < 8; e++) av[b][e] = a[(int64_t)c * 512 + e];
#pragma unroll
        for (int q = 0; q < NT; q++)
#pragma unroll
            for (int e = 0; e < 8; e++) xv[b][q][e] = x[((int64_t)q * C + c) * 512 + e];
    };
    auto mul = [&](int b) {
#pragma unroll
        for (int q = 0; q < NT; q++)
#pragma unroll
            for (int e = 0; e < 8; e++)
                acc[q] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[b][e], xv[b][q][e], acc[q], 0, 0, 0);
    };
    if (c0 < c1) load(0, c0);
    int c = c0;
    for (; c + 1 < c1; c += 2) {
        load(1, c + 1);
        mul(0);
        if (c + 2 < c1) load(0, c + 2);
        mul(1);
    }
    if (c < c1) mul(0);
    // D fragment of the f64 16x16x4 product: lane = rhs + 16 * (row % 4), element row / 4
    const int Np = frag_rows(n);
#pragma unroll
    for (int q = 0; q < NT; q++) {
        const int r = 16 * q + (lane & 15);
        if (r >= nrhs) continue;
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int i = 16 * T + (lane >> 4) + 4 * v;
            parts[((int64_t)p * nrhs + r) * Np + i] = acc[q][v];
        }
    }
}

// Element t (< N * nrhs) of the step's end: x = sum of the column-split partials (fixed order),
// the boundary post-processing (src/SolverStepLinearT.cpp:512-536) and the ring update.
__device__ __forceinline__ void finalize_elem(const double* __restrict__ parts, int nparts,
                                              int Np, const int* __restrict__ is_ubc,
                                              const int* __restrict__ is_fbc,
                                              const double* __restrict__ ubc_val,
                                              const double* __restrict__ fbc_val,
                                              const double* __restrict__ fac, int N, int nrhs,
                                              int s, int R, double a1, double a2, double a3,
                                              double* __restrict__ U, double* __restrict__ T,
                                              double* __restrict__ ring, int64_t t)
{
    const int r = (int)(t / N), d = (int)(t % N);
    double xv = 0.0;
    for (int p = 0; p < nparts; p++) xv += parts[((int64_t)p * nrhs + r) * Np + d];
    const double u = is_ubc[d] ? ubc_val[d] : xv;
    double tr = is_fbc[d] ? fbc_val[d] * fac[r] : 0.0;
    // fDofHigh is inclusive (src/SolverT.cpp:179) and the test is u_dof < fDofHigh (:527)
    if (is_ubc[d] && d < N - 1) tr = xv;
    const int64_t hs = ((int64_t)s * nrhs + r) * N + d;
    U[hs] = u;
    T[hs] = tr;
    const double up = s >= 2 ? U[hs - (int64_t)nrhs * N] : 0.0;
    const double tp = s >= 2 ? T[hs - (int64_t)nrhs * N] : 0.0;
    auto at = [&](int m, int part) -> double* {
        int sl = m % R;
        if (sl < 0) sl += R;
        return ring + (((int64_t)sl * 2 + part) * N + d) * nrhs + r;
    };
    *at(s - 1, 0) += a1 * u;            // y_{s-1} complete
    *at(s - 1, 1) += a1 * tr;
    *at(s, 0) = a3 * u + a2 * up;       // z_1 of step s+1
    *at(s, 1) = a3 * tr + a2 * tp;
    *at(s + 1, 0) = (a2 - a1) * u;      // z_0 of step s+1
    *at(s + 1, 1) = (a2 - a1) * tr;
}

}  // namespace tdbem
