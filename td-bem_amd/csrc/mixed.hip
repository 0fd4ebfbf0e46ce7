// mixed.hip — the mixed-precision history convolution: split-BF16 (three terms) MFMA products
// accumulated in FP32 inside a 32-wide k-tile and in FP64 across tiles, blocks and lags, for R
// right-hand sides at once (R = 16, 32, 48 or 64).
//
// Layout.  Nodes are ordered by recursive coordinate bisection and grouped into tiles of 16
// (48 dofs).  For each pair of tiles (I, J) and each lag k in the union of their node-pair shells
// the 48 x 96 block A[(3 ii + p)][(6 jj + 3 part + q)] (part 0 = G[p][q], times z^t; part 1 =
// -H[p][q], times z^u) is cut into 3 x 3 sub-tiles of 16 x 32; only sub-tiles that some node pair
// reaches at lag k are stored, in FP32, laid out in MFMA A-operand order so that a wavefront
// loads each 16 x 32 operand as two contiguous 1 KiB lines (2 KiB per sub-tile).  The history
// ring (FP64) is mirrored in FP32, in MFMA B-operand order, per (slot, column tile); a workgroup
// splits each history block it stages into LDS into three bf16 terms.
//
// Work.  A workgroup takes 4 row tiles (one per wavefront, neighbours in the bisection order)
// and its wavefronts walk the same merged (J, k) list of the four, each skipping the steps where
// its row tile has no block.  Walking in step, they fetch the history fragment of
// (J, slot of s - k) close together in time, so it comes from HBM once and from L1/L2 after.
//
// Per sub-tile each FP32 value is split exactly in registers as a1 + a2 + a3 (bf16, 8 + 8 + 8
// significant bits; z likewise) and the products of all term pairs but a3 z3 (below 2^-28 of
// |a| |z|) run on v_mfma_f32_16x16x32_bf16; the leading
// product and the sum of the corrections leave FP32 for FP64 accumulators after every sub-tile,
// so rounding does not grow with the number of blocks and lags summed.  Storing FP32 instead of
// the three bf16 terms moves 2 KiB instead of 3 KiB per sub-tile through HBM, the bound.
#include <algorithm>
#include <array>
#include <cstring>
#include <numeric>
#include <stdexcept>
#include <vector>

#include "step_device.h"

namespace tdbem {

namespace {

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

// Tile store from a segment of the FP64 node-pair store: the rows of the positions P0 .. of the
// bisection order (one batch of the lag kernel's slot pass), in the store's layout ([chunk][15]
// [64] entries, each row's (j, k) entries in order), with each pair's lag range hull[li][j] and
// first entry off[li][j] (li = position - P0).  One wavefront per stored sub-tile (row tile I,
// column tile J, lag k, sub-tile u = 3 m + kt) of the row tiles the segment holds.  Lane L holds
// block row 16 m + L % 16 and the 8 block columns 32 kt + 8 (L / 16) .. + 7 (16x16x32 A-operand
// order); it looks each of its (at most two) node pairs up once, reads the values -- the
// segment's entries of consecutive lags are consecutive, so the sub-tiles of one tile pair's
// lags, which neighbouring workgroups on one XCD take, share their lines -- and writes columns
// 0..3 of its eight to the sub-tile's first 1 KiB line and 4..7 to the second.  Block row
// 3 ii + p, column 6 jj + 3 part + q holds G[p][q] (part 0) or -H[p][q] (part 1) of node pair
// (perm[16 I + ii], perm[16 J + jj]): the FP64 store's value rounded once to FP32.
__global__ void __launch_bounds__(256)
    k_tp_from_segment(const double* __restrict__ seg, const int2* __restrict__ hull,
                      const int64_t* __restrict__ off, const int* __restrict__ perm, int P0,
                      int nnd, const uint2* __restrict__ sub_id, int64_t s0, int64_t s1,
                      f32x4* __restrict__ tp)
{
    // consecutive sub-tiles (the lags of one tile pair) on one XCD
    const int64_t sidx = s0 + (int64_t)xcd_chunk(blockIdx.x, gridDim.x) * 4 + (threadIdx.x >> 6);
    if (sidx >= s1) return;
    const int lane = threadIdx.x & 63;
    const uint2 id = sub_id[sidx];
    const int I = (int)(id.x >> 16), J = (int)(id.x & 0xFFFFu);
    const int k = (int)(id.y >> 4), u = (int)(id.y & 15u);
    const int m = u / 3, kt = u % 3;
    const int r = 16 * m + (lane & 15);
    const int pos = I * kTile + r / 3;
    const int i = perm[pos], p = r % 3;
    const int64_t li = pos - P0;
    const int sym[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
    // the lane's columns c0 .. c0 + 7 (c0 a multiple of 8) cover node columns j0 = c0 / 6 and
    // j0 + 1; both pairs are looked up at once and every value load is issued before any is used
    // (the segment holds at least one chunk whenever this kernel runs, so a missing entry reads
    // entry 0 and is replaced by zero)
    const int c0 = 32 * kt + 8 * (lane >> 4);
    const int j0 = c0 / 6;
    int64_t en[2];
#pragma unroll
    for (int b = 0; b < 2; b++) {
        const int j = perm[J * kTile + j0 + b];
        const int64_t cell = li * nnd + (j >= 0 ? j : 0);
        const int2 h = hull[cell];
        const int64_t o = off[cell];
        en[b] = (i >= 0 && j >= 0 && h.x <= k && k <= h.y) ? o + (k - h.x) : -1;
    }
    double v[8];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int c = c0 + e;
        const int b = c / 6 - j0;
        const int part = (c % 6) / 3, q = c % 3;
        const int comp = part == 0 ? sym[p][q] : 6 + 3 * p + q;
        const int64_t x = en[b] >= 0 ? en[b] : 0;
        v[e] = seg[((x / kChunk) * kVals + comp) * kChunk + x % kChunk];
    }
    f32x4 out[2];
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int c = c0 + e;
        const int b = c / 6 - j0;
        out[e >> 2][e & 3] = en[b] >= 0 ? (float)((c % 6) >= 3 ? -v[e] : v[e]) : 0.0f;
    }
#pragma unroll
    for (int h = 0; h < 2; h++) tp[(sidx * 2 + h) * 64 + lane] = out[h];
}

// Rows lower..upper of the dense G and H of lag k from the tile store (the FP32 values): one
// workgroup per merged-list entry, one wavefront per row tile of the entry's group (a near lag
// sits in both lists of its group and is written twice, with the same values).
__global__ void __launch_bounds__(64 * kGroup)
    k_tile_to_dense(const f32x4* __restrict__ tp, const TPEntry* __restrict__ ent,
                    const int64_t* __restrict__ grp_off, int nL, const int* __restrict__ perm,
                    int nT, int k, int lower, int upper, int64_t N, double* __restrict__ G,
                    double* __restrict__ H)
{
    const int64_t e = blockIdx.x;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const TPEntry E = ent[e];
    if ((int)(E.id & kMaxLag) != k) return;
    int lo = 0, hi = nL;   // list of entry e: the last L with grp_off[L] <= e; group L / 2
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (grp_off[mid] <= e)
            lo = mid;
        else
            hi = mid;
    }
    const int I = (lo >> 1) * kGroup + w, J = (int)(E.id >> kKBits);
    if (I >= nT) return;
    const unsigned msk = E.msk[w];
    for (int u = 0; u < 9; u++) {
        if (!((msk >> u) & 1u)) continue;
        const int64_t sidx = E.sub[w] + __builtin_popcount(msk & ((1u << u) - 1u));
        const int m = u / 3, kt = u % 3;
        const int r = 16 * m + (lane & 15);
        const int i = perm[I * kTile + r / 3], p = r % 3;
        if (i < lower || i > upper) continue;
#pragma unroll
        for (int e8 = 0; e8 < 8; e8++) {
            const int c = 32 * kt + 8 * (lane >> 4) + e8;
            const int jj = c / 6, part = (c % 6) / 3, q = c % 3;
            const int j = perm[J * kTile + jj];
            if (j < 0) continue;
            const float v = tp[(sidx * 2 + (e8 >> 2)) * 64 + lane][e8 & 3];
            const int64_t o = (int64_t)(3 * (i - lower) + p) * N + 3 * j + q;
            if (part == 0)
                G[o] = v;
            else
                H[o] = -(double)v;
        }
    }
}

// FP32 copies (see zcopy_elem) of the ring slots of steps s - 2, s - 1, s (the ones the
// previous step's finalize changed), and the load factors f_r(t) of the step in block 0.
__global__ void k_zcopy3(const double* __restrict__ ring, int N, int nrhs, int nT,
                         const int* __restrict__ perm, int s, int R, int NT,
                         unsigned short* __restrict__ zs, const double* __restrict__ amp,
                         const double* __restrict__ rise, double t_load,
                         double* __restrict__ fac)
{
    if (blockIdx.x == 0 && (int)threadIdx.x < nrhs)
        fac[threadIdx.x] = load_factor(amp[threadIdx.x], rise[threadIdx.x], t_load);
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t per = (int64_t)nT * kCols * nrhs;
    if (t >= 3 * per) return;
    const int m = (int)(t / per);
    int slot = (s - 2 + m) % R;
    if (slot < 0) slot += R;
    zcopy_elem(ring, N, nrhs, nT, perm, slot, NT, zs, t - m * per);
}

// One workgroup per (chunk, right-hand-side tile): chunks base .. of the plan, the NT
// workgroups of a chunk next to each other on one XCD, one wavefront per row tile of the chunk's
// group (conv_chunk).  A launch runs the chunks of the lag parity of step s.
__global__ void __launch_bounds__(64 * kGroup, 2)
    k_conv_mixed(const f32x4* __restrict__ tp, const TPEntry* __restrict__ ent,
                 const int4* __restrict__ chunk, int base, const bf16x8* __restrict__ zs, int nT,
                 int NT, int Rring, int s, double* __restrict__ partial)
{
    const int L = xcd_chunk(blockIdx.x, gridDim.x);
    const int c = base + L / NT, nt0 = L % NT;
    const int4 cw = chunk[c];
    const int w = threadIdx.x >> 6;
    const int g = cw.x & ((1 << kChunkParityShift) - 1);
    // partial sums of chunk c: [c][target][row tile][3][NT][64][4]
    double* out = partial + ((int64_t)c * 2 * kGroup + w) * 3 * NT * 256;
    conv_chunk(tp, ent, zs, nT, NT, nt0, Rring, s, cw.y, cw.z, w, g * kGroup + w < nT, out,
               out + (int64_t)kGroup * 3 * NT * 256);
}

// Step solve right-hand side from the partial sums of the history at step s (see
// reduce_rhs_elem).
__global__ void k_tp_reduce_rhs(const double* __restrict__ partial, int nT,
                                const int* __restrict__ grp_chunk, int nG, int s, int NT,
                                const int* __restrict__ perm, int N, int nrhs,
                                const double* __restrict__ gF, const double* __restrict__ gU,
                                const double* __restrict__ fac, double* __restrict__ carry,
                                double* __restrict__ Xf)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nT * kRows * nrhs) return;
    reduce_rhs_elem(partial, grp_chunk, nG, s & 1, NT, perm, N, nrhs, gF, gU, fac, carry, Xf, t);
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

// Sorts idx[lo, hi) along its widest coordinate.
void sort_widest(std::vector<int>& idx, int lo, int hi, const std::vector<double>& X)
{
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int a = lo; a < hi; a++)
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], X[3 * idx[a] + d]);
            mx[d] = std::max(mx[d], X[3 * idx[a] + d]);
        }
    int ax = 0;
    for (int d = 1; d < 3; d++)
        if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
    std::stable_sort(idx.begin() + lo, idx.begin() + hi,
                     [&](int a, int b) { return X[3 * a + ax] < X[3 * b + ax]; });
}

// Recursive halving of idx[lo, hi) across the widest coordinate, down to single nodes.
void halve(std::vector<int>& idx, int lo, int hi, const std::vector<double>& X)
{
    if (hi - lo <= 1) return;
    sort_widest(idx, lo, hi, X);
    const int mid = lo + (hi - lo) / 2;
    halve(idx, lo, mid, X);
    halve(idx, mid, hi, X);
}

// Recursive coordinate bisection of idx[lo, hi): split across the widest coordinate so that the
// left part holds a whole number of tiles; leaves are tiles of kTile nodes.  Inside a full tile
// the nodes are split across its widest coordinate into parts of 5, 6 and 5 -- the nodes whose
// rows fall in 16-row sub-tiles 0, 0-1 .. 1-2 and 2 (node 5 has rows 15-17, node 10 rows 30-32;
// columns likewise, 6 per node in 32-column sub-tiles) -- each ordered by recursive halving, so
// the node pairs of one 16 x 32 sub-tile are close together and their lag ranges overlap more:
// 4.7% fewer stored sub-tiles than the tile's plain sort order on the 1536-quad sphere.
void bisect(std::vector<int>& idx, int lo, int hi, const std::vector<double>& X)
{
    const int n = hi - lo;
    if (n <= kTile) {
        if (n < kTile) {
            halve(idx, lo, hi, X);
            return;
        }
        static_assert(kTile == 16, "in-tile parts are sized for 16-node tiles");
        sort_widest(idx, lo, hi, X);
        halve(idx, lo, lo + 5, X);
        halve(idx, lo + 5, lo + 11, X);
        halve(idx, lo + 11, hi, X);
        return;
    }
    sort_widest(idx, lo, hi, X);
    const int ntiles = (n + kTile - 1) / kTile;
    const int left = ntiles / 2 * kTile;
    bisect(idx, lo, lo + left, X);
    bisect(idx, lo + left, hi, X);
}

}  // namespace

void tp_plan_structure(Context& c)
{
    TPStore& T = c.tp;
    if (T.planned) return;
    const int nnd = c.nnd;
    std::vector<int> order(nnd);
    std::iota(order.begin(), order.end(), 0);
    bisect(order, 0, nnd, c.h_coords);
    T.nT = (nnd + kTile - 1) / kTile;
    T.nG = (T.nT + kGroup - 1) / kGroup;
    T.perm.assign((size_t)T.nT * kTile, -1);
    T.pos.assign(nnd, 0);
    for (int p = 0; p < nnd; p++) {
        T.perm[p] = order[p];
        T.pos[order[p]] = p;
    }
    const int nT = T.nT;
    // lag range of every sub-tile (m, kt) of every tile pair, from the node-pair hulls: node ii
    // has rows 3 ii .. 3 ii + 2 (m-tiles (3 ii) / 16 and (3 ii + 2) / 16), node jj columns
    // 6 jj .. 6 jj + 5 (k-tiles (6 jj) / 32 and (6 jj + 5) / 32)
    const size_t npair = (size_t)nT * nT;
    std::vector<int> slo(npair * 9, 1 << 30), shi(npair * 9, -1);
    const HostBuf<int2>& H = c.pin_hull;   // node-pair lag ranges of the store just built
    // row tiles in parallel: each writes only the sub-tile ranges of its own tile pairs (I, J)
    parallel_for(nT, [&](int64_t Ilo, int64_t Ihi) {
        for (int64_t pi = Ilo * kTile; pi < Ihi * kTile; pi++) {
            const int i = T.perm[pi];
            if (i < 0) continue;
            const int I = (int)(pi / kTile), ii = (int)(pi % kTile);
            const int m0 = (3 * ii) >> 4, m1 = (3 * ii + 2) >> 4;
            for (int pj = 0; pj < nT * kTile; pj++) {
                const int j = T.perm[pj];
                if (j < 0) continue;
                const int2 h = H[(size_t)i * nnd + j];
                if (h.y < h.x) continue;
                const int J = pj / kTile, jj = pj % kTile;
                const int k0 = (6 * jj) >> 5, k1 = (6 * jj + 5) >> 5;
                int* lo = &slo[((size_t)I * nT + J) * 9];
                int* hi = &shi[((size_t)I * nT + J) * 9];
                for (int m = m0; m <= m1; m++)
                    for (int kt = k0; kt <= k1; kt++) {
                        lo[3 * m + kt] = std::min(lo[3 * m + kt], h.x);
                        hi[3 * m + kt] = std::max(hi[3 * m + kt], h.y);
                    }
            }
        }
    });
    // blocks: every lag of each tile-pair hull; sub-tiles stored where the lag is in range.
    // Counted per tile pair in parallel, offsets by a prefix sum, then filled in parallel.
    std::vector<int> plo(npair, 1 << 30), phi(npair, -1);
    std::vector<int64_t> off(npair, -1), qsub(npair + 1, 0), qblk(npair + 1, 0);
    auto block_mask = [&](size_t q, int k) {
        uint16_t mk = 0;
        for (int u = 0; u < 9; u++)
            if (slo[q * 9 + u] <= k && k <= shi[q * 9 + u]) mk |= (uint16_t)(1u << u);
        return mk;
    };
    parallel_for((int64_t)npair, [&](int64_t lo, int64_t hi) {
        for (int64_t q = lo; q < hi; q++) {
            for (int u = 0; u < 9; u++) {
                plo[q] = std::min(plo[q], slo[q * 9 + u]);
                phi[q] = std::max(phi[q], shi[q * 9 + u]);
            }
            int64_t ns = 0;
            for (int k = plo[q]; k <= phi[q]; k++) ns += __builtin_popcount(block_mask(q, k));
            qsub[q + 1] = ns;
            qblk[q + 1] = phi[q] >= plo[q] ? phi[q] - plo[q] + 1 : 0;
        }
    }, 256);
    for (size_t q = 0; q < npair; q++) {
        qsub[q + 1] += qsub[q];
        qblk[q + 1] += qblk[q];
    }
    const int64_t nb = qblk[npair], nsub = qsub[npair];
    if (nsub >= (int64_t)UINT32_MAX || nb >= (int64_t)INT32_MAX)
        throw std::runtime_error("tile store too large");
    std::vector<uint32_t> sub_off(nb);
    std::vector<uint16_t> msk(nb);
    std::vector<uint2> sub_id(nsub);   // (I << 16 | J, k << 4 | u) of each stored sub-tile
    std::vector<int64_t> qnz(npair, 0);
    parallel_for((int64_t)npair, [&](int64_t lo, int64_t hi) {
        for (int64_t q = lo; q < hi; q++) {
            if (phi[q] < plo[q]) continue;
            off[q] = qblk[q];
            const uint32_t IJ = ((uint32_t)(q / nT) << 16) | (uint32_t)(q % nT);
            int64_t b = qblk[q], sub = qsub[q];
            for (int k = plo[q]; k <= phi[q]; k++, b++) {
                const uint16_t mk = block_mask(q, k);
                sub_off[b] = (uint32_t)sub;
                msk[b] = mk;
                for (int u = 0; u < 9; u++)
                    if ((mk >> u) & 1u) sub_id[sub++] = make_uint2(IJ, ((uint32_t)k << 4) | (uint32_t)u);
                qnz[q] += mk != 0;
            }
        }
    }, 256);
    int64_t nz = 0;
    for (int64_t x : qnz) nz += x;
    T.nblocks = nz;
    T.nsub = nsub;
    // merged (J, k) lists of the groups of kGroup row tiles, built per group in parallel: list
    // 2 g + p holds the near lags (k < kNearLags) and the far lags of parity p (the time-blocked
    // convolution runs the lists of one parity per step)
    std::vector<std::vector<TPEntry>> gent(2 * (size_t)T.nG);
    parallel_for(T.nG, [&](int64_t glo, int64_t ghi) {
        for (int64_t g = glo; g < ghi; g++) {
            for (int J = 0; J < nT; J++) {
                int klo = 1 << 30, khi = -1;
                for (int w = 0; w < kGroup; w++) {
                    const int64_t I = g * kGroup + w;
                    if (I >= nT) continue;
                    const size_t q = (size_t)I * nT + J;
                    klo = std::min(klo, plo[q]);
                    khi = std::max(khi, phi[q]);
                }
                for (int k = klo; k <= khi; k++) {
                    TPEntry e;
                    e.id = ((uint32_t)J << kKBits) | (uint32_t)k;
                    e.kmask = 0;
                    bool any = false;
                    for (int w = 0; w < kGroup; w++) {
                        e.sub[w] = 0;
                        e.msk[w] = 0;
                        const int64_t I = g * kGroup + w;
                        if (I >= nT) continue;
                        const size_t q = (size_t)I * nT + J;
                        if (off[q] < 0 || k < plo[q] || k > phi[q]) continue;
                        const int64_t b = off[q] + (k - plo[q]);
                        if (!msk[b]) continue;
                        e.sub[w] = sub_off[b];
                        e.msk[w] = msk[b];
                        for (int m = 0; m < 3; m++) e.kmask |= (msk[b] >> (3 * m)) & 7u;
                        any = true;
                    }
                    if (!any) continue;
                    if (k < kNearLags) {
                        gent[2 * g].push_back(e);
                        gent[2 * g + 1].push_back(e);
                    } else {
                        gent[2 * g + (k & 1)].push_back(e);
                    }
                }
            }
        }
    });
    std::vector<TPEntry> ent;
    std::vector<int64_t> goff(2 * (size_t)T.nG + 1, 0);
    for (size_t L = 0; L < gent.size(); L++) {
        goff[L] = (int64_t)ent.size();
        ent.insert(ent.end(), gent[L].begin(), gent[L].end());
    }
    goff[gent.size()] = (int64_t)ent.size();
    T.nentries = (int64_t)ent.size();
    T.h_grp_off = goff;
    // cost of an entry: its stored sub-tiles, the staged history k-tiles (a far lag's for two
    // targets), a fixed part
    T.h_work.resize(ent.size());
    for (size_t e = 0; e < ent.size(); e++) {
        int nsubs = 0;
        for (int w = 0; w < kGroup; w++) nsubs = std::max(nsubs, __builtin_popcount(ent[e].msk[w]));
        const int tg = (int)(ent[e].id & kMaxLag) < kNearLags ? 1 : 2;
        // the slowest wavefront of the step, the history staging, the step's latency
        T.h_work[e] = (float)(nsubs + 0.5 * tg * __builtin_popcount(ent[e].kmask) + 4.0);
    }
    // per lag class (near, far even, far odd): in-shell node-pair lags and stored sub-tiles
    // (what a launch of either parity streams), entries per parity list
    {
        const HostBuf<int2>& Hh = c.pin_hull;
        std::vector<std::array<int64_t, 3>> pl(nT);
        parallel_for(nT, [&](int64_t Ilo, int64_t Ihi) {
            for (int64_t I = Ilo; I < Ihi; I++) {
                std::array<int64_t, 3> cnt{0, 0, 0};
                for (int ii = 0; ii < kTile; ii++) {
                    const int i = T.perm[I * kTile + ii];
                    if (i < 0) continue;
                    for (int j = 0; j < nnd; j++) {
                        const int2 h = Hh[(size_t)i * nnd + j];
                        if (h.y < h.x) continue;
                        cnt[0] += std::max(0, std::min(h.y, kNearLags - 1) - h.x + 1);
                        const int lo = std::max(h.x, kNearLags);
                        if (h.y < lo) continue;
                        // lags lo..h.y: the even ones, then the odd ones
                        const int ev = h.y / 2 - (lo + 1) / 2 + 1;
                        cnt[1] += ev;
                        cnt[2] += (h.y - lo + 1) - ev;
                    }
                }
                pl[I] = cnt;
            }
        });
        for (int q = 0; q < 3; q++) T.pair_lags_class[q] = T.sub_class[q] = 0;
        for (int I = 0; I < nT; I++)
            for (int q = 0; q < 3; q++) T.pair_lags_class[q] += pl[I][q];
        for (const uint2& id : sub_id) {
            const int k = (int)(id.y >> 4);
            T.sub_class[k < kNearLags ? 0 : 1 + (k & 1)]++;
        }
        for (int p = 0; p < 2; p++) {
            T.ent_parity[p] = 0;
            for (int g = 0; g < T.nG; g++)
                T.ent_parity[p] += goff[2 * g + p + 1] - goff[2 * g + p];
        }
    }
    T.plan_nrhs = 0;
    T.d_perm.upload(T.perm.data(), T.perm.size(), c.stream);
    T.d_pos.upload(T.pos.data(), T.pos.size(), c.stream);
    T.ent.upload(ent.data(), ent.size(), c.stream);
    T.d_grp_off.upload(goff.data(), goff.size(), c.stream);
    // stored sub-tiles are ordered by tile pair (I, J): those of row tile I are a range
    T.h_tile_sub.assign(nT + 1, 0);
    for (int I = 0; I <= nT; I++) T.h_tile_sub[I] = qsub[(size_t)I * nT];
    T.h_sub_id = std::move(sub_id);
    TD_HIP(hipStreamSynchronize(c.stream));
    T.planned = true;
}

void tp_alloc(Context& c)
{
    TPStore& T = c.tp;
    if (!T.planned) throw std::logic_error("tile store not planned");
    if ((int64_t)T.h_sub_id.size() != T.nsub)
        throw std::logic_error("tile store plan without its sub-tile list");
    T.vals.alloc((size_t)std::max<int64_t>(T.nsub, 1) * kSubFloats);
    // the sub-tile list lives for one build only: stream-ordered, so that its free at the end
    // of the build does not wait for the whole device
    T.d_sub_id.order_on(c.stream);
    T.d_sub_id.upload(T.h_sub_id.data(), T.h_sub_id.size(), c.stream);
    // a large list is not kept on the host: the next build plans again
    if (T.nsub * (int64_t)sizeof(uint2) > kPinnedKeepBytes) std::vector<uint2>().swap(T.h_sub_id);
}

void tp_gather_segment(Context& c, int I0, int I1, const double* seg, const int2* hull,
                       const int64_t* off)
{
    TPStore& T = c.tp;
    if (!T.planned || !T.vals.p) throw std::logic_error("tile store not planned");
    // the kernel reads the sub-tile list and writes the sub-tiles: both must be the plan's size
    if (!T.d_sub_id.p || (int64_t)T.d_sub_id.n != T.nsub ||
        (int64_t)T.vals.n != std::max<int64_t>(T.nsub, 1) * kSubFloats)
        throw std::logic_error("tile store buffers do not match its plan");
    if (I0 < 0 || I1 > T.nT || I0 > I1) throw std::logic_error("row tiles out of range");
    const int64_t s0 = T.h_tile_sub[I0], s1 = T.h_tile_sub[I1];
    if (s1 <= s0) return;
    hipLaunchKernelGGL(k_tp_from_segment, dim3((unsigned)((s1 - s0 + 3) / 4)), dim3(256), 0,
                       c.stream, seg, hull, off, T.d_perm.p, I0 * kTile, c.nnd, T.d_sub_id.p, s0,
                       s1, reinterpret_cast<f32x4*>(T.vals.p));
    TD_HIP(hipGetLastError());
}

void tp_finish(Context& c)
{
    TPStore& T = c.tp;
    TD_HIP(hipStreamSynchronize(c.stream));
    T.d_sub_id.reset();
    // the plan stays valid for the next build only while its sub-tile list is kept
    if ((int64_t)T.h_sub_id.size() != T.nsub) T.planned = false;
    T.ready = true;
}

void tp_build(Context& c)
{
    if (!c.store_ready) build_lag_store(c);
    if (!c.tp.ready) throw std::logic_error("mixed-precision lag store without its tile store");
}

void tile_rows(Context& c, int k, int lower, int upper, double* dG, double* dH)
{
    const TPStore& T = c.tp;
    const int64_t N = 3LL * c.nnd;
    const int64_t n = 3LL * (upper - lower + 1) * N;
    TD_HIP(hipMemsetAsync(dG, 0, n * sizeof(double), c.stream));
    TD_HIP(hipMemsetAsync(dH, 0, n * sizeof(double), c.stream));
    if (!T.ready || T.nentries == 0) return;
    hipLaunchKernelGGL(k_tile_to_dense, dim3((unsigned)T.nentries), dim3(64 * kGroup), 0,
                       c.stream, reinterpret_cast<const f32x4*>(T.vals.p), T.ent.p,
                       T.d_grp_off.p, 2 * T.nG, T.d_perm.p, T.nT, k, lower, upper, N, dG, dH);
    TD_HIP(hipGetLastError());
}

void tp_prepare_step(Context& c, const double* ring, int N, int nrhs, int s, int R, float* zs,
                     const double* amp, const double* rise, double t_load, double* fac)
{
    const TPStore& T = c.tp;
    const int64_t n = 3LL * T.nT * kCols * nrhs;
    hipLaunchKernelGGL(k_zcopy3, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c.stream, ring,
                       N, nrhs, T.nT, T.d_perm.p, s, R, nrhs / 16,
                       reinterpret_cast<unsigned short*>(zs), amp, rise, t_load, fac);
    TD_HIP(hipGetLastError());
}

// FP32-sized words per ring slot of the history copy: per (column tile, k-tile, rhs tile) three
// bf16 terms of 64 lanes x 8
int64_t tp_zslot_elems(const Context& c, int nrhs) { return (int64_t)c.tp.nT * 3 * (nrhs / 16) * 768; }

int tp_resident_groups(int nrhs, int device)
{
    if (nrhs < 16 || nrhs > 64 || nrhs % 16)
        throw std::invalid_argument("mixed march needs nrhs in {16, 32, 48, 64}");
    int per_cu = 1, cus = 1;
    TD_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(&k_conv_mixed), 64 * kGroup, 0));
    TD_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
    return std::max(1, per_cu) * std::max(1, cus);
}

double tp_work(const Context& c)
{
    double total = 0;
    for (float x : c.tp.h_work) total += x;
    return total;
}

void tp_chunks(const Context& c, int target, int sc, std::vector<int4>& ch, std::vector<int>& gch)
{
    const TPStore& T = c.tp;
    target = std::max(target, 2 * T.nG);
    const double total = tp_work(c);
    const size_t gbase = gch.size();
    gch.resize(gbase + 2 * ((size_t)T.nG + 1), 0);
    for (int p = 0; p < 2; p++) {
        int* gc = gch.data() + gbase + (size_t)p * (T.nG + 1);
        for (int g = 0; g < T.nG; g++) {
            const int first = (int)ch.size();
            gc[g] = first;
            const int64_t e0 = T.h_grp_off[2 * g + p], e1 = T.h_grp_off[2 * g + p + 1];
            double wg = 0;
            for (int64_t e = e0; e < e1; e++) wg += T.h_work[e];
            const int n = std::max(1, (int)std::lround(total > 0 ? target * wg / total : 1.0));
            const double step = wg / n;
            const int gx = g | (p << kChunkParityShift);
            double acc = 0;
            int64_t start = e0;
            for (int64_t e = e0; e < e1; e++) {
                acc += T.h_work[e];
                // close the chunk once it holds its share (the last one takes the rest)
                const int made = (int)ch.size() - first;
                if (acc >= step * (made + 1) && made < n - 1) {
                    ch.push_back(make_int4(gx, (int)start, (int)(e + 1), sc));
                    start = e + 1;
                }
            }
            ch.push_back(make_int4(gx, (int)start, (int)e1, sc));
        }
        gc[T.nG] = (int)ch.size();
    }
}

// Chunks of a context's merged lists: one per kChunkWork / rounds units of entry work (about
// 2 000 for the 1536-quad sphere at rounds = 4, four rounds of the workgroups an MI355X holds at
// once), at least one per group.  A function of the context alone -- not of the device or of the
// other scenarios of a batch -- so a scenario's partial sums, and its results bit for bit, are
// the same marched alone and in any batch at the same rounds.
constexpr double kChunkWork = 2400.0;

int tp_chunk_target(const Context& c, int rounds)
{
    return std::max(c.tp.nG, (int)std::lround(rounds * tp_work(c) / kChunkWork));
}

void tp_plan(Context& c, int nrhs, int rounds)
{
    TPStore& T = c.tp;
    if (T.plan_nrhs == nrhs && T.plan_rounds == rounds) return;
    std::vector<int4> ch;
    std::vector<int> gch;
    tp_chunks(c, tp_chunk_target(c, rounds), 0, ch, gch);
    T.nchunks = (int)ch.size();
    T.nch[1] = 0;
    for (const int4& q : ch) T.nch[1] += (q.x >> kChunkParityShift) & 1;
    T.nch[0] = T.nchunks - T.nch[1];
    T.chunk.upload(ch.data(), ch.size(), c.stream);
    T.grp_chunk.upload(gch.data(), gch.size(), c.stream);
    TD_HIP(hipStreamSynchronize(c.stream));
    T.plan_nrhs = nrhs;
    T.plan_rounds = rounds;
}

void tp_conv(Context& c, const float* zs, int nrhs, int Rring, int s, double* partial)
{
    const TPStore& T = c.tp;
    if (T.plan_nrhs != nrhs) throw std::logic_error("mixed convolution not planned for nrhs");
    if (nrhs < 16 || nrhs > 64 || nrhs % 16)
        throw std::runtime_error("mixed march needs nrhs in {16, 32, 48, 64}");
    const int NT = nrhs / 16;
    // the chunks of lag parity s mod 2 (the plan holds parity 0's first), NT workgroups each
    const int base = (s & 1) ? T.nch[0] : 0;
    const int n = T.nch[s & 1];
    if (n == 0) return;
    hipLaunchKernelGGL(k_conv_mixed, dim3((unsigned)(n * NT)), dim3(64 * kGroup), 0, c.stream,
                       reinterpret_cast<const f32x4*>(T.vals.p), T.ent.p, T.chunk.p, base,
                       reinterpret_cast<const bf16x8*>(zs), T.nT, NT, Rring, s, partial);
    TD_HIP(hipGetLastError());
}

void tp_reduce_rhs(Context& c, const double* partial, int nrhs, int N, int s, const double* gF,
                   const double* gU, const double* fac, double* carry, double* Xf)
{
    const TPStore& T = c.tp;
    const int NT = nrhs / 16;
    const int64_t n = (int64_t)T.nT * kRows * nrhs;
    hipLaunchKernelGGL(k_tp_reduce_rhs, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                       c.stream, partial, T.nT, T.grp_chunk.p, T.nG, s, NT, T.d_perm.p, N, nrhs,
                       gF, gU, fac, carry, Xf);
    TD_HIP(hipGetLastError());
}

int64_t tp_partial_elems(const Context& c, int nrhs)
{
    return (int64_t)c.tp.nchunks * 2 * kGroup * 3 * (nrhs / 16) * 256;
}

int64_t tp_carry_elems(const Context& c, int nrhs) { return (int64_t)c.tp.nT * kRows * nrhs; }

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
