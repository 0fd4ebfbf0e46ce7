// assemble.hip — hot path 1: G/H influence-matrix assembly on CDNA4.
//
// Replaces ElastoDynamicsT::G_H_Driver / TDElementStepLinearT::FormGH_* (src/ElastoDynamicsT.cpp:
// 108-280, src/TDElementStepLinearT.cpp:28-626).  Two products:
//   * assemble_step_dense: the per-step dense row blocks fG_1, fG_2, fH_1, fH_2 exactly as
//     G_H_Driver leaves them (parity and the FE-BE matrix export);
//   * build_lag_store: every lag matrix M_k = phi_1(step k) + phi_2(step k+1) of
//     SolverStepLinearT::UpdateGHLinear (src/SolverStepLinearT.cpp:58-97), for k = 0..Kc, kept
//     only inside each node pair's wavefront shell (lags where the Step kernel can be nonzero).
//
// Work decomposition (both products): one thread per (collocation node i, element e, lag/part)
// integrates the element over its quadrature points, skipping points whose Step kernel is zero
// at that lag (before the P wave arrives or after both waves have passed), and writes per-node
// 3x3 blocks to a scratch slot; a second kernel gathers, for each (i, column node j), the slots
// of the elements around j.  The gather is deterministic (fixed order over the node->element
// adjacency), so results do not depend on scheduling.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <stdexcept>

#include "step_device.h"

namespace tdbem {

namespace {

constexpr int kBlock = 256;
// The lag kernel's workgroups are single wavefronts of at most 256 registers (two per SIMD for
// 4-node elements): beside a running march they start wherever one SIMD has room, instead of
// waiting for a whole CU.
constexpr int kLagBlock = 64;

inline unsigned grid_for(int64_t n, int block = kBlock)
{
    return (unsigned)((n + block - 1) / block);
}

template <int ENND>
__device__ __forceinline__ int singular_node(const double* X, const double* coll)
{
    // src/ElastoDynamicsT.cpp:166-175: the last element node within 1e-9 (squared) wins.
    int s = -1;
#pragma unroll
    for (int a = 0; a < ENND; a++) {
        const double dx = X[3 * a] - coll[0], dy = X[3 * a + 1] - coll[1],
                     dz = X[3 * a + 2] - coll[2];
        if (dx * dx + dy * dy + dz * dz < 1e-9) s = a;
    }
    return s;
}

template <int ENND>
__device__ __forceinline__ void load_element(const double* coords, const int* ien, int e,
                                             double* X)
{
#pragma unroll
    for (int a = 0; a < ENND; a++) {
        const int g = ien[e * ENND + a];
        X[3 * a] = coords[3 * g];
        X[3 * a + 1] = coords[3 * g + 1];
        X[3 * a + 2] = coords[3 * g + 2];
    }
}

// Visits every quadrature point of pair (coll, element): f(x, n, wJ, N[ENND]).
// Regular: nq x nq Gauss tensor rule from the precomputed table (src/TDElementStepLinearT.cpp:
// 52-61).  Singular: two Duffy triangles of nq x nq (:436-530); 8-node elements are split into
// bilinear sub-elements whose shape values land on their parent nodes (:149-175, :597-626).
template <int ENND, typename F>
__device__ __forceinline__ void for_each_point(const double* X, int sing, const double* geo_e,
                                               const double* gauss, int nq, F&& f)
{
    constexpr int GS = 7 + ENND;
    if (sing < 0) {
        const int nq2 = nq * nq;
        for (int q = 0; q < nq2; q++) {
            const double* p = geo_e + q * GS;
            f(p, p + 3, p[6], p + 7);
        }
        return;
    }
    const double* gx = gauss;
    const double* gw = gauss + nq;
    if constexpr (ENND == 4) {
        for (int tri = 0; tri < 2; tri++)
            for (int q1 = 0; q1 < nq; q1++)
                for (int q2 = 0; q2 < nq; q2++) {
                    double e1, e2, Jt, N[4], x[3], n[3], J;
                    duffy(sing, tri, gx[q1], gx[q2], e1, e2, Jt);
                    shape_eval<4>(X, e1, e2, N, x, n, J);
                    f(x, n, gw[q1] * gw[q2] * Jt * J, N);
                }
    } else {
        int ids[4], nsub, cid;
        sub_element(sing, 0, ids, nsub, cid);
        for (int se = 0; se < nsub; se++) {
            sub_element(sing, se, ids, nsub, cid);
            double Xs[12];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                Xs[3 * k] = X[3 * ids[k]];
                Xs[3 * k + 1] = X[3 * ids[k] + 1];
                Xs[3 * k + 2] = X[3 * ids[k] + 2];
            }
            for (int tri = 0; tri < 2; tri++)
                for (int q1 = 0; q1 < nq; q1++)
                    for (int q2 = 0; q2 < nq; q2++) {
                        double e1, e2, Jt, N4[4], x[3], n[3], J, N[8];
                        duffy(cid, tri, gx[q1], gx[q2], e1, e2, Jt);
                        shape_eval<4>(Xs, e1, e2, N4, x, n, J);
#pragma unroll
                        for (int a = 0; a < 8; a++) N[a] = 0.0;
#pragma unroll
                        for (int k = 0; k < 4; k++) N[ids[k]] = N4[k];
                        f(x, n, gw[q1] * gw[q2] * Jt * J, N);
                    }
        }
    }
}

// Time-integral combination S for one point: sum over three time levels of cI*I(t) + cJ*J(t).
// Returns the number of levels evaluated (and ir = 1/r), 0 when the combination is zero: every level with a nonzero coefficient is before
// the P-wave arrival (t <= r/c_p), or (when the coefficients sum to zero, as for the Step
// kernel) every such level is in the static regime t - r/c_s >= dt, where HL(t) no longer
// depends on t (src/GreenFunctionT.cpp:251-314 with b = dt for both waves).
struct TimeCombo {
    double t[3];
    double cI[3], cJ[3];
    bool static_ok;
};

__device__ __forceinline__ int eval_s(const TimeCombo& tc, double r, double dt, double inv_dt,
                                      const Material& m, double* S, double& ir)
{
    const double inv_cp = m.inv_cp, inv_cs = m.inv_cs;
    const double ta = r * inv_cp;         // P-wave arrival
    const double ts = r * inv_cs + dt;    // static from here on
    bool any = false, all_static = true;
#pragma unroll
    for (int l = 0; l < 3; l++) {
        if (tc.cI[l] == 0.0 && tc.cJ[l] == 0.0) continue;
        if (tc.t[l] > ta) any = true;
        if (tc.t[l] < ts) all_static = false;
    }
    if (!any || (tc.static_ok && all_static)) return 0;
    ir = 1.0 / r;
    const double inv_r2 = ir * ir;
#pragma unroll
    for (int l = 0; l < 6; l++) S[l] = 0.0;
    int levels = 0;
#pragma unroll
    for (int l = 0; l < 3; l++) {
        if ((tc.cI[l] == 0.0 && tc.cJ[l] == 0.0) || !(tc.t[l] > ta)) continue;
        accumulate_hl<double>(tc.t[l], dt, inv_dt, r, inv_r2, inv_cp, inv_cs, m.inv_cp2,
                              m.inv_cs2, tc.cI[l], tc.cJ[l], S);
        levels++;
    }
    return levels;
}

// The lag combination M_k = phi_1(k) + phi_2(k + 1) of the lag store (src/SolverStepLinearT.cpp:
// 65-97, each phi the Step kernel HL(t) - HL(t - dt)): per point
//     S = I(t_k) - I(t_{k-1}) + cJ (J(t_{k+1}) - J(t_k)),   t_l = l dt,
// cJ = 0 at the last lag k = Kc.  It is the TimeCombo (-1, 0 | 1, -cJ | 0, cJ) with the I
// coefficients left out: a level contributes only after the P-wave arrival t > r/c_p > 0, and at
// k = 0 neither t_{k-1} = -dt nor t_k = 0 passes that test.
struct LagCombo {
    double t0, t1, t2;   // (k - 1) dt, k dt, (k + 1) dt
    double cJ;           // 1, or 0 at k = Kc
};

// r = sqrt(rr) and ir = 1/r from the hardware reciprocal square root (v_rsq_f64, about 23 good
// bits) by two Goldschmidt iterations on (g, h) -> (sqrt(rr), 1/(2 sqrt(rr))) and one Newton
// correction of the root (Markstein, "IA-64 and Elementary Functions", 2000): r to within an
// ulp of the rounded root, ir to a few ulps, in 12 instructions instead of a rounded square root
// and a divide.  rr is a squared distance between distinct points (normal, positive).
__device__ __forceinline__ void sqrt_and_recip(double rr, double& r, double& ir)
{
    const double y = __builtin_amdgcn_rsq(rr);
    double g = rr * y, h = 0.5 * y;
    double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    h = fma(h, e, h);
    r = fma(fma(-g, g, rr), h, g);
    ir = h + h;
}

// One wave's hat integrals at one time level in units of dt (hat_integrals x dt; the caller
// scales the level sum by 1/dt once): the I hat (phi_1) when WANT_I, the J hat (phi_2) when
// WANT_J.  t_arr = r/c of the wave, as the caller compared it; TREF_POS: the caller knows
// t > t_arr, so I5 needs no test for t - r/c < 0.
template <bool WANT_I, bool WANT_J, bool TREF_POS>
__device__ __forceinline__ void lag_hat(double t, double dt, double t_arr, double inv_c2,
                                        double ir2_12, double& I1, double& I3, double& I5,
                                        double& J1, double& J3, double& J5)
{
    const double t_ref = t - t_arr;
    const double b = fmin(fmax(t_ref, 0.0), dt);
    const double b2 = b * b;
    const double hb2 = 0.5 * b2;
    const double t2 = t * t;
    if constexpr (WANT_I) {
        const double pI = b * (12.0 * dt * t2 + b * (-(6.0 * t2 + 12.0 * t * dt) +
                                                     b * ((8.0 * t + 4.0 * dt) - 3.0 * b)));
        const double q = dt * b - hb2;
        I1 = pI * ir2_12 - q * inv_c2;
        I3 = q;
        I5 = (TREF_POS || t_ref >= 0.0) ? dt - b : 0.0;
    }
    if constexpr (WANT_J) {
        const double pJ = b2 * (-6.0 * t2 + b * (8.0 * t - 3.0 * b));
        J1 = -pJ * ir2_12 - hb2 * inv_c2;
        J3 = hb2;
        J5 = t_ref < dt ? b : 0.0;
    }
}

// S of the lag combination at one point (as eval_s, whose zero tests it restates for this
// combination): levels t_{k+1} (J only), t_k (I and J), t_{k-1} (I only), each formed for the P
// and the S wave.  Returns the number of levels evaluated.
__device__ __forceinline__ int eval_s_lag(const LagCombo& lc, double r, double ir, double dt,
                                          double inv_dt, const Material& m, double* S)
{
    const double ta = r * m.inv_cp;        // P-wave arrival
    const double tsa = r * m.inv_cs;       // S-wave arrival
    const bool any = lc.cJ != 0.0 ? lc.t2 > ta : lc.t1 > ta;
    if (!any || lc.t0 >= tsa + dt) return 0;
    const double ir2_12 = ir * ir * (1.0 / 12.0);
    const double icp2 = m.inv_cp2, ics2 = m.inv_cs2;
    double x1, x3, x5, y1, y3, y5;   // P wave
    double u1, u3, u5, v1, v3, v5;   // S wave
#pragma unroll
    for (int l = 0; l < 6; l++) S[l] = 0.0;
    int levels = 0;
    if (lc.cJ != 0.0 && lc.t2 > ta) {
        lag_hat<false, true, true>(lc.t2, dt, ta, icp2, ir2_12, x1, x3, x5, y1, y3, y5);
        lag_hat<false, true, false>(lc.t2, dt, tsa, ics2, ir2_12, u1, u3, u5, v1, v3, v5);
        S[0] = y1;
        S[1] = v1;
        S[2] = y3;
        S[3] = v3;
        S[4] = y5;
        S[5] = v5;
        levels++;
    }
    if (lc.t1 > ta) {
        const double cJ = lc.cJ;
        lag_hat<true, true, true>(lc.t1, dt, ta, icp2, ir2_12, x1, x3, x5, y1, y3, y5);
        lag_hat<true, true, false>(lc.t1, dt, tsa, ics2, ir2_12, u1, u3, u5, v1, v3, v5);
        S[0] += x1 - cJ * y1;
        S[1] += u1 - cJ * v1;
        S[2] += x3 - cJ * y3;
        S[3] += u3 - cJ * v3;
        S[4] += x5 - cJ * y5;
        S[5] += u5 - cJ * v5;
        levels++;
    }
    if (lc.t0 > ta) {
        lag_hat<true, false, true>(lc.t0, dt, ta, icp2, ir2_12, x1, x3, x5, y1, y3, y5);
        lag_hat<true, false, false>(lc.t0, dt, tsa, ics2, ir2_12, u1, u3, u5, v1, v3, v5);
        S[0] -= x1;
        S[1] -= u1;
        S[2] -= x3;
        S[3] -= u3;
        S[4] -= x5;
        S[5] -= u5;
        levels++;
    }
#pragma unroll
    for (int l = 0; l < 6; l++) S[l] *= inv_dt;
    return levels;
}

// Distance, reciprocal and S of one point for either combination.
__device__ __forceinline__ int point_s(const TimeCombo& tc, double rr, double dt, double inv_dt,
                                       const Material& m, double* S, double& r, double& ir)
{
    r = sqrt(rr);
    return eval_s(tc, r, dt, inv_dt, m, S, ir);
}

__device__ __forceinline__ int point_s(const LagCombo& lc, double rr, double dt, double inv_dt,
                                       const Material& m, double* S, double& r, double& ir)
{
    sqrt_and_recip(rr, r, ir);
    return eval_s_lag(lc, r, ir, dt, inv_dt, m, S);
}

// ---------------------------------------------------------------------------------------------
// Geometry table: shape functions, positions, unit normals and w*J at the regular Gauss points.
// ---------------------------------------------------------------------------------------------
template <int ENND>
__global__ void k_geom(const double* __restrict__ coords, const int* __restrict__ ien,
                       const double* __restrict__ gauss, int nq, int nel, double* __restrict__ geo)
{
    const int nq2 = nq * nq;
    const int tid = blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nel * nq2) return;
    const int e = tid / nq2, q = tid % nq2;
    const int q1 = q / nq, q2 = q % nq;
    double X[3 * ENND], N[ENND], x[3], n[3], J;
    load_element<ENND>(coords, ien, e, X);
    shape_eval<ENND>(X, gauss[q1], gauss[q2], N, x, n, J);
    double* o = geo + ((size_t)e * nq2 + q) * (7 + ENND);
    o[0] = x[0];
    o[1] = x[1];
    o[2] = x[2];
    o[3] = n[0];
    o[4] = n[1];
    o[5] = n[2];
    o[6] = gauss[nq + q1] * gauss[nq + q2] * J;
#pragma unroll
    for (int a = 0; a < ENND; a++) o[7 + a] = N[a];
}

// ---------------------------------------------------------------------------------------------
// C-term: C(i) = - sum_e sum_q wJ * (Kelvin traction * dt/2)^T, added by AssembleC to the
// diagonal block (src/ElastoDynamicsT.cpp:204-205, :261-280).  One block per collocation node,
// fixed-shape tree reduction.
// ---------------------------------------------------------------------------------------------
template <int ENND>
__global__ void __launch_bounds__(kBlock) k_cterm(const double* __restrict__ coords,
                                                  const int* __restrict__ ien,
                                                  const double* __restrict__ geo,
                                                  const double* __restrict__ gauss, int nq,
                                                  int nel, double dt, Material m,
                                                  double* __restrict__ cterm)
{
    __shared__ double red[9][kBlock];
    const int i = blockIdx.x;
    const double coll[3] = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    double acc[9];
#pragma unroll
    for (int v = 0; v < 9; v++) acc[v] = 0.0;
    const int nq2 = nq * nq;
    for (int e = threadIdx.x; e < nel; e += blockDim.x) {
        double X[3 * ENND];
        load_element<ENND>(coords, ien, e, X);
        const int sing = singular_node<ENND>(X, coll);
        for_each_point<ENND>(X, sing, geo + (size_t)e * nq2 * (7 + ENND), gauss, nq,
                             [&](const double* x, const double* n, double wJ, const double*) {
                                 double d[3] = {x[0] - coll[0], x[1] - coll[1], x[2] - coll[2]};
                                 const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
                                 const double ir = 1.0 / r;
                                 d[0] *= ir;
                                 d[1] *= ir;
                                 d[2] *= ir;
                                 double KT[9];
                                 kelvin_block(d, n, r, dt, m, KT);
#pragma unroll
                                 for (int v = 0; v < 9; v++) acc[v] -= wJ * KT[v];
                             });
    }
#pragma unroll
    for (int v = 0; v < 9; v++) red[v][threadIdx.x] = acc[v];
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int v = 0; v < 9; v++) red[v][threadIdx.x] += red[v][threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x < 9) cterm[9 * i + threadIdx.x] = red[threadIdx.x][0];
}

// ---------------------------------------------------------------------------------------------
// Shell bounds per (node, element) pair: lags k with M_k possibly nonzero.
// P_n (the Step kernel at step n) is nonzero only for r/c_p < n dt < r/c_s + 2 dt at some point;
// M_k uses P_k and P_{k+1}, so k runs over [n_lo - 1, n_hi] clipped to [0, Kc].
// ---------------------------------------------------------------------------------------------
template <int ENND>
__global__ void k_pair_range(const double* __restrict__ coords, const int* __restrict__ ien,
                             const double* __restrict__ geo, const double* __restrict__ gauss,
                             int nq, int nnd, int nel, double dt, Material m, int Kc,
                             int2* __restrict__ pair_k)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)nnd * nel) return;
    const int i = (int)(tid / nel), e = (int)(tid % nel);
    const double coll[3] = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    double X[3 * ENND];
    load_element<ENND>(coords, ien, e, X);
    const int sing = singular_node<ENND>(X, coll);
    double rmin = 1e300, rmax = 0.0;
    for_each_point<ENND>(X, sing, geo + (size_t)e * nq * nq * (7 + ENND), gauss, nq,
                         [&](const double* x, const double*, double, const double*) {
                             const double d0 = x[0] - coll[0], d1 = x[1] - coll[1],
                                          d2 = x[2] - coll[2];
                             const double r = sqrt(d0 * d0 + d1 * d1 + d2 * d2);
                             rmin = fmin(rmin, r);
                             rmax = fmax(rmax, r);
                         });
    const double tol = 1e-9;
    const int n_lo = max(1, (int)floor(rmin / (m.cp * dt) - tol) + 1);
    const int n_hi = min(Kc, (int)ceil(rmax / (m.cs * dt) + tol) + 1);
    int2 out;
    if (n_lo > Kc || n_hi < n_lo) {
        out.x = 1;
        out.y = 0;
    } else {
        out.x = n_lo - 1;
        out.y = n_hi;
    }
    pair_k[tid] = out;
}

// Hull of the pair shells of the elements around column node j, for each (i, j).
__global__ void k_node_hull(const int2* __restrict__ pair_k, const int* __restrict__ adj_off,
                            const int* __restrict__ adj, int nnd, int nel,
                            int2* __restrict__ hull)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)nnd * nnd) return;
    const int i = (int)(tid / nnd), j = (int)(tid % nnd);
    int lo = 1 << 30, hi = -1;
    for (int a = adj_off[j]; a < adj_off[j + 1]; a++) {
        const int e = adj[a] >> 3;
        const int2 pk = pair_k[(int64_t)i * nel + e];
        if (pk.y < pk.x) continue;
        lo = min(lo, pk.x);
        hi = max(hi, pk.y);
    }
    hull[tid] = hi < lo ? make_int2(1, 0) : make_int2(lo, hi);
}

// Entry index (j << kKBits | k) of every in-shell (row, j, k) from the pair hulls and first
// entries of nrows rows ([nrows][nnd] each).
__global__ void k_fill_idx(const int2* __restrict__ hull, const int64_t* __restrict__ off,
                           int nnd, int64_t nrows, uint32_t* __restrict__ idx)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= nrows * nnd) return;
    const int j = (int)(tid % nnd);
    const int2 h = hull[tid];
    const int64_t o = off[tid];
    for (int k = h.x; k <= h.y; k++) idx[o + (k - h.x)] = ((uint32_t)j << kKBits) | (uint32_t)k;
}

// Slot list of one row batch of the lag kernel: slot pair_slot[p] + j, j below the in-shell lag
// count of pair p = (row r, element e), belongs to p.  rows: the batch's row nodes (-1: a
// padding row, no slots) when they are not i0, i0 + 1, ...  One thread per pair; each wavefront
// then writes its 64 pairs' runs one pair at a time with all lanes, so the stores are
// contiguous.  The blocks have kBlock threads, whole wavefronts.
__global__ void __launch_bounds__(kBlock)
    k_slot_fill(const int64_t* __restrict__ pair_slot, const int2* __restrict__ pair_k,
                const int* __restrict__ rows, int i0, int nel, int64_t npairs,
                int* __restrict__ slot_pair)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int lane = (int)(threadIdx.x & 63);
    int cnt = 0;
    int64_t off = 0;
    if (p < npairs) {
        const int i = rows ? rows[p / nel] : i0 + (int)(p / nel);
        if (i >= 0) {
            const int2 k = pair_k[(int64_t)i * nel + (int)(p % nel)];
            cnt = k.y >= k.x ? k.y - k.x + 1 : 0;
            off = pair_slot[p];
        }
    }
    const int64_t first = p - lane;   // the wavefront's first pair
    for (int l = 0; l < 64; l++) {
        const int c = __builtin_amdgcn_readlane(cnt, l);
        if (c == 0) continue;
        const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)off, l);
        const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(off >> 32), l);
        const int64_t o = (int64_t)(((uint64_t)hi << 32) | lo);
        for (int j = lane; j < c; j += 64) slot_pair[o + j] = (int)(first + l);
    }
}

// Slots of the pairs with a node at the collocation point: complemented, so that the regular
// lag kernel skips them and the singular one takes them.
__global__ void k_slot_mark(const int64_t* __restrict__ sing_slots, int64_t nsing,
                            int* __restrict__ slot_pair)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsing) return;
    const int64_t s = sing_slots[t];
    slot_pair[s] = ~slot_pair[s];
}

// ---------------------------------------------------------------------------------------------
// Element-pair integration into scratch slots.  Mode LAG: slot = (pair, k) gives
// M_k = [k>=1](I(t_k) - I(t_{k-1})) + [k+1<=Kc](J(t_{k+1}) - J(t_k)).  Mode STEP: slot =
// (pair, part) gives phi_1 (part 0) or phi_2 (part 1) of step n; for the Heaviside kernel
// (src/GreenFunctionT.cpp:126-129) there is no (t - dt) level.
// ---------------------------------------------------------------------------------------------
// The rules of for_each_point with the regular points' shape functions evaluated in place of
// the mesh's geometry table (the element entry integrates pairs of no mesh).
template <int ENND, typename F>
__device__ __forceinline__ void for_each_point_direct(const double* X, int sing,
                                                      const double* gauss, int nq, F&& f)
{
    if (sing >= 0) {
        for_each_point<ENND>(X, sing, nullptr, gauss, nq, f);
        return;
    }
    for (int q1 = 0; q1 < nq; q1++)
        for (int q2 = 0; q2 < nq; q2++) {
            double N[ENND], x[3], n[3], J;
            shape_eval<ENND>(X, gauss[q1], gauss[q2], N, x, n, J);
            f(x, n, gauss[nq + q1] * gauss[nq + q2] * J, N);
        }
}

// Per-point addition after a point's kernel blocks are accumulated: none.
struct NoPointExtra {
    template <typename Acc>
    __device__ __forceinline__ void operator()(const double*, const double*, double, Acc&) const
    {
    }
};

// FP64 integration of a pair over the points that points(f) visits (for_each_point: regular
// rule for sing < 0, Duffy triangles / sub-elements for a singular node), accumulated into
// out[ENND][15].  extra(x, n, wJ, acc) runs at every point after its blocks are added.  Returns
// (points with a nonzero time combination, time levels evaluated at them).
template <int ENND, typename Points, typename Combo, typename Extra = NoPointExtra>
__device__ __forceinline__ int2 integrate_points(Points&& points, const double* coll,
                                                 const Combo& tc, double dt,
                                                 const Material& m, double* out, int64_t ostride,
                                                 Extra&& extra = Extra{})
{
    int active = 0, levels = 0;
    const double inv_dt = 1.0 / dt;
    double acc[ENND][kVals];
#pragma unroll
    for (int a = 0; a < ENND; a++)
#pragma unroll
        for (int v = 0; v < kVals; v++) acc[a][v] = 0.0;
    points([&](const double* x, const double* n, double wJ, const double* N) {
        double d[3] = {x[0] - coll[0], x[1] - coll[1], x[2] - coll[2]};
        double S[6], r, ir;
        const int nl = point_s(tc, d[0] * d[0] + d[1] * d[1] + d[2] * d[2], dt, inv_dt, m, S,
                               r, ir);
        if (nl) {
            active++;
            levels += nl;
            d[0] *= ir;
            d[1] *= ir;
            d[2] *= ir;
            double blk[kVals];
            blocks_from_s<double>(S, d, n, r, ir, ir * ir, m, blk);
#pragma unroll
            for (int a = 0; a < ENND; a++) {
                const double wN = wJ * N[a];
#pragma unroll
                for (int v = 0; v < kVals; v++) acc[a][v] += wN * blk[v];
            }
        }
        extra(x, n, wJ, acc);
    });
#pragma unroll
    for (int a = 0; a < ENND; a++)
#pragma unroll
        for (int v = 0; v < kVals; v++) out[(a * kVals + v) * ostride] = acc[a][v];
    return make_int2(active, levels);
}

// Value v of node a goes to out[(a * 15 + v) * ostride].  REGULAR: the caller guarantees that
// no node of e lies at the collocation point, so the singular rules are not compiled in (the
// regular path then fits two waves per SIMD).
template <int ENND, bool REGULAR = false, typename Combo = TimeCombo>
__device__ __forceinline__ int2 integrate_pair(const double* coords, const int* ien,
                                              const double* geo, const double* gauss, int nq,
                                              int i, int e, const Combo& tc, double dt,
                                              const Material& m, double* out, int64_t ostride = 1)
{
    const double coll[3] = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    double X[3 * ENND];
    load_element<ENND>(coords, ien, e, X);
    const int sing = REGULAR ? -1 : singular_node<ENND>(X, coll);
    const double* geo_e = geo + (size_t)e * nq * nq * (7 + ENND);
    return integrate_points<ENND>(
        [&](auto&& f) { for_each_point<ENND>(X, sing, geo_e, gauss, nq, f); }, coll, tc, dt, m,
        out, ostride);
}

// The reference's 20-point time rule (src/GreenFunctionT.cpp:1042-1164, 9 digits), for the
// device and (below) the host.
__constant__ double kTime20x[20] = {
    -0.076526521, 0.076526521, -0.227785851, 0.227785851, -0.373706089, 0.373706089,
    -0.510867002, 0.510867002, -0.636053681, 0.636053681, -0.746331906, 0.746331906,
    -0.839116972, 0.839116972, -0.912234428, 0.912234428, -0.963971927, 0.963971927,
    -0.993128599, 0.993128599};
__constant__ double kTime20w[20] = {
    0.152753387, 0.152753387, 0.149172986, 0.149172986, 0.142096109, 0.142096109, 0.131688638,
    0.131688638, 0.118194532, 0.118194532, 0.10193012, 0.10193012, 0.083276742, 0.083276742,
    0.062672048, 0.062672048, 0.04060143, 0.04060143, 0.017614007, 0.017614007};

const double kTime20xHost[20] = {
    -0.076526521, 0.076526521, -0.227785851, 0.227785851, -0.373706089, 0.373706089,
    -0.510867002, 0.510867002, -0.636053681, 0.636053681, -0.746331906, 0.746331906,
    -0.839116972, 0.839116972, -0.912234428, 0.912234428, -0.963971927, 0.963971927,
    -0.993128599, 0.993128599};
const double kTime20wHost[20] = {
    0.152753387, 0.152753387, 0.149172986, 0.149172986, 0.142096109, 0.142096109, 0.131688638,
    0.131688638, 0.118194532, 0.118194532, 0.10193012, 0.10193012, 0.083276742, 0.083276742,
    0.062672048, 0.062672048, 0.04060143, 0.04060143, 0.017614007, 0.017614007};

// The reference's 40-point rule (src/GreenFunctionT.cpp:1082-1139; the same values as the
// 40-point shape-function rule, src/ShapeFunctionT.cpp:134-159), for the ramped-Heaviside kernel.
__constant__ double kTime40x[40] = {
    -0.0387724175060508, 0.0387724175060508, -0.1160840706752552, 0.1160840706752552,
    -0.1926975807013711, 0.1926975807013711, -0.2681521850072537, 0.2681521850072537,
    -0.3419940908257585, 0.3419940908257585, -0.4137792043716050, 0.4137792043716050,
    -0.4830758016861787, 0.4830758016861787, -0.5494671250951282, 0.5494671250951282,
    -0.6125538896679802, 0.6125538896679802, -0.6719566846141796, 0.6719566846141796,
    -0.7273182551899271, 0.7273182551899271, -0.7783056514265194, 0.7783056514265194,
    -0.8246122308333117, 0.8246122308333117, -0.8659595032122595, 0.8659595032122595,
    -0.9020988069688743, 0.9020988069688743, -0.9328128082786765, 0.9328128082786765,
    -0.9579168192137917, 0.9579168192137917, -0.9772599499837743, 0.9772599499837743,
    -0.9907262386994570, 0.9907262386994570, -0.9982377097105593, 0.9982377097105593};
__constant__ double kTime40w[40] = {
    0.0775059479784248, 0.0775059479784248, 0.0770398181642480, 0.0770398181642480,
    0.0761103619006262, 0.0761103619006262, 0.0747231690579683, 0.0747231690579683,
    0.0728865823958041, 0.0728865823958041, 0.0706116473912868, 0.0706116473912868,
    0.0679120458152339, 0.0679120458152339, 0.0648040134566010, 0.0648040134566010,
    0.0613062424929289, 0.0613062424929289, 0.0574397690993916, 0.0574397690993916,
    0.0532278469839368, 0.0532278469839368, 0.0486958076350722, 0.0486958076350722,
    0.0438709081856733, 0.0438709081856733, 0.0387821679744720, 0.0387821679744720,
    0.0334601952825478, 0.0334601952825478, 0.0279370069800234, 0.0279370069800234,
    0.0222458491941670, 0.0222458491941670, 0.0164210583819079, 0.0164210583819079,
    0.0104982845311528, 0.0104982845311528, 0.0045212770985332, 0.0045212770985332};

// Pulse kernels of the TDElementRampHeaviLinearT element loop: kernel 5, the cubic B-spline
// (Compute(5): Int_Stokes_CubicBSpline_Linear, src/GreenFunctionT.cpp:625-658, 20-point time
// rule), and kernel 6, the ramped Heaviside (Int_Stokes_RampedHeaviside_Linear, :589-622,
// 40-point time rule).  phi_{part+1} blocks of pair (i, e) at time t at every spatial point; the
// Kelvin parts go to the C-term (assemble_step_dense).
template <int ENND, typename Points, typename Extra = NoPointExtra>
__device__ __forceinline__ void pulse_points(Points&& points, const double* coll, double t,
                                             double dt, const Material& m, int part, int kernel,
                                             double* out, Extra&& extra = Extra{})
{
    double acc[ENND][kVals];
#pragma unroll
    for (int a = 0; a < ENND; a++)
#pragma unroll
        for (int v = 0; v < kVals; v++) acc[a][v] = 0.0;
    const bool ramp = kernel == 6;
    const double bw = 0.7 * dt, slope = 2 / dt;
    const int ntq = ramp ? 40 : 20;
    points([&](const double* x, const double* n, double wJ, const double* N) {
        double d[3] = {x[0] - coll[0], x[1] - coll[1], x[2] - coll[2]};
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double ir = 1.0 / r;
        d[0] *= ir;
        d[1] *= ir;
        d[2] *= ir;
        double blk[kVals];
#pragma unroll
        for (int v = 0; v < kVals; v++) blk[v] = 0.0;
        for (int q = 0; q < ntq; q++) {
            const double ab = ramp ? kTime40x[q] : kTime20x[q];
            const double w = ramp ? kTime40w[q] : kTime20w[q];
            const double phi = part == 0 ? 0.5 * (1 - ab) : 0.5 * (1 + ab);
            const double tt = 0.5 * (1 + ab) * dt;
            double b[kVals];
            if (ramp)
                ramp_blocks(t - tt, slope, r, d, n, m, b);
            else
                bspline_blocks(t - tt, bw, r, d, n, m, b);
            const double wt = w * (dt * 0.5) * phi;
#pragma unroll
            for (int v = 0; v < kVals; v++) blk[v] += wt * b[v];
        }
#pragma unroll
        for (int a = 0; a < ENND; a++) {
            const double wN = wJ * N[a];
#pragma unroll
            for (int v = 0; v < kVals; v++) acc[a][v] += wN * blk[v];
        }
        extra(x, n, wJ, acc);
    });
#pragma unroll
    for (int a = 0; a < ENND; a++)
#pragma unroll
        for (int v = 0; v < kVals; v++) out[a * kVals + v] = acc[a][v];
}

template <int ENND>
__device__ __forceinline__ void integrate_pair_pulse(const double* coords, const int* ien,
                                                     const double* geo, const double* gauss,
                                                     int nq, int i, int e, double t, double dt,
                                                     const Material& m, int part, int kernel,
                                                     double* out)
{
    const double coll[3] = {coords[3 * i], coords[3 * i + 1], coords[3 * i + 2]};
    double X[3 * ENND];
    load_element<ENND>(coords, ien, e, X);
    const int sing = singular_node<ENND>(X, coll);
    const double* geo_e = geo + (size_t)e * nq * nq * (7 + ENND);
    pulse_points<ENND>([&](auto&& f) { for_each_point<ENND>(X, sing, geo_e, gauss, nq, f); },
                       coll, t, dt, m, part, kernel, out);
}

// Lag slots of one row batch, value v of element node a of slot s at scratch[(a * 15 + v) *
// nslots + s] (consecutive slots of a pair are consecutive lags, so the gather's lanes read
// runs of consecutive addresses).  slot_pair[s] is the pair of slot s, complemented (~p) for the
// pairs whose element has a node within 1e-3 of the collocation point: the SINGULAR launch runs
// over those slots (list sing_slots, nsing entries) with the full integrator, the regular launch
// over all the others with the regular-only one.
template <int ENND, bool SINGULAR>
__global__ void __launch_bounds__(kLagBlock, ENND == 4 ? 2 : 1)
    k_pair_lags(const double* __restrict__ coords, const int* __restrict__ ien,
                const double* __restrict__ geo, const double* __restrict__ gauss, int nq, int nel,
                double dt, Material m, int Kc, int i0, int64_t nslots,
                const int* __restrict__ slot_pair, const int64_t* __restrict__ pair_slot,
                const int2* __restrict__ pair_k, const int64_t* __restrict__ sing_slots,
                int64_t nsing, double* __restrict__ scratch,
                unsigned long long* __restrict__ counter, const int* __restrict__ rows)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t s;
    int p;
    if (SINGULAR) {
        if (t >= nsing) return;
        s = sing_slots[t];
        p = ~slot_pair[s];
    } else {
        if (t >= nslots) return;
        s = t;
        p = slot_pair[s];
        if (p < 0) return;
    }
    // rows: the batch's row nodes when they are not the contiguous range i0..
    const int i = rows ? rows[p / nel] : i0 + p / nel, e = p % nel;
    const int k = pair_k[(int64_t)i * nel + e].x + (int)(s - pair_slot[p]);
    LagCombo lc;
    lc.t0 = (k - 1) * dt;
    lc.t1 = k * dt;
    lc.t2 = (k + 1) * dt;
    lc.cJ = (k + 1 <= Kc) ? 1.0 : 0.0;
    const int2 n = integrate_pair<ENND, !SINGULAR>(coords, ien, geo, gauss, nq, i, e, lc, dt, m,
                                                   scratch + s, nslots);
    if (counter) {
        atomicAdd(counter, (unsigned long long)n.x);
        atomicAdd(counter + 1, (unsigned long long)n.y);
    }
}

template <int ENND>
__global__ void __launch_bounds__(kBlock)
    k_pair_step(const double* __restrict__ coords, const int* __restrict__ ien,
                const double* __restrict__ geo, const double* __restrict__ gauss, int nq, int nel,
                double dt, Material m, int step, int kernel, int i0, int nrow_nodes,
                double* __restrict__ scratch)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= (int64_t)nrow_nodes * nel * 2) return;
    const int part = (int)(s & 1);
    const int64_t p = s >> 1;
    const int i = i0 + (int)(p / nel), e = (int)(p % nel);
    if (kernel == 5 || kernel == 6) {
        integrate_pair_pulse<ENND>(coords, ien, geo, gauss, nq, i, e, step * dt, dt, m, part,
                                   kernel, scratch + s * (ENND * kVals));
        return;
    }
    const int heaviside = kernel == 2;
    TimeCombo tc;
    tc.t[0] = step * dt - dt;   // fT - fT_1, as Stokes_Step_Linear forms it
    tc.t[1] = step * dt;
    tc.t[2] = 0.0;
    const double prev = heaviside ? 0.0 : -1.0;
    tc.cI[0] = part == 0 ? prev : 0.0;
    tc.cJ[0] = part == 1 ? prev : 0.0;
    tc.cI[1] = part == 0 ? 1.0 : 0.0;
    tc.cJ[1] = part == 1 ? 1.0 : 0.0;
    tc.cI[2] = 0.0;
    tc.cJ[2] = 0.0;
    tc.static_ok = !heaviside;
    integrate_pair<ENND>(coords, ien, geo, gauss, nq, i, e, tc, dt, m,
                         scratch + s * (ENND * kVals));
}

// Element matrices of TDElementT::FormGH_Regular / FormGH_Singular (src/TDElementT.h:37-48;
// src/TDElementStepLinearT.cpp:28-187, src/TDElementHeavisideLinearT.cpp,
// src/TDElementRampHeaviLinearT.cpp) for npairs (collocation point, element) pairs of no mesh,
// thread per (pair, part): the phi_{part+1} matrices GE and HE [3][3 ENND] and the Kelvin block
// C [3][3] (weight cw1 / cw2), by the point rules and kernels of the assembly.  X [npairs][ENND][3],
// coll [npairs][3], sing [npairs] (-1: regular rule, a: Duffy rules about node a).  out holds six
// arrays: GE_1, GE_2, HE_1, HE_2 of npairs * 9 ENND doubles, then C_1, C_2 of npairs * 9.
// A pair whose sing is out of range gets NaN.
template <int ENND>
__global__ void __launch_bounds__(kBlock)
    k_element(const double* __restrict__ X_all, const double* __restrict__ coll_all,
              const int* __restrict__ sing_all, int64_t npairs, const double* __restrict__ gauss,
              int nq, int kernel, double t, double dt, Material m, double cw1, double cw2,
              double* __restrict__ out)
{
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= npairs * 2) return;
    const int part = (int)(s & 1);
    const int64_t p = s >> 1;
    constexpr int W = 9 * ENND;
    double* GE = out + (part * npairs + p) * W;
    double* HE = out + ((2 + part) * npairs + p) * W;
    double* C = out + 4 * npairs * W + (part * npairs + p) * 9;
    const int sing = sing_all[p];
    if (sing < -1 || sing >= ENND) {
        const double nan = __builtin_nan("");
        for (int v = 0; v < W; v++) GE[v] = HE[v] = nan;
        for (int v = 0; v < 9; v++) C[v] = nan;
        return;
    }
    double X[3 * ENND], coll[3];
#pragma unroll
    for (int v = 0; v < 3 * ENND; v++) X[v] = X_all[p * 3 * ENND + v];
#pragma unroll
    for (int d = 0; d < 3; d++) coll[d] = coll_all[3 * p + d];
    auto points = [&](auto&& f) { for_each_point_direct<ENND>(X, sing, gauss, nq, f); };
    const double cw = part == 0 ? cw1 : cw2;
    // Kelvin part at one point: KT = Kelvin traction * dt/2, transposed as the C block takes it
    auto kelvin_at = [&](const double* x, const double* n, double* KT) {
        double d[3] = {x[0] - coll[0], x[1] - coll[1], x[2] - coll[2]};
        const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        const double ir = 1.0 / r;
        d[0] *= ir;
        d[1] *= ir;
        d[2] *= ir;
        kelvin_block(d, n, r, dt, m, KT);
    };
    // Singular elements of the Heaviside and pulse kernels subtract their Kelvin part from the
    // singular node's H columns at every point, after the point's Stokes part, instead of adding
    // it to C (src/TDElementHeavisideLinearT.cpp:489-497, src/TDElementRampHeaviLinearT.cpp:
    // 499-506): near the node the two nearly cancel, and the per-point difference keeps the
    // digits a separate sum would lose.
    const bool into_h = sing >= 0 && kernel != 4;
    auto extra = [&](const double* x, const double* n, double wJ, auto& acc) {
        if (!into_h || cw == 0.0) return;
        double KT[9];
        kelvin_at(x, n, KT);
        const double f = cw * wJ;
#pragma unroll
        for (int a = 0; a < ENND; a++)
            if (a == sing)
#pragma unroll
                for (int v = 0; v < 9; v++) acc[a][6 + v] -= f * KT[v];
    };
    double acc[ENND * kVals];
    if (kernel == 5 || kernel == 6) {
        pulse_points<ENND>(points, coll, t, dt, m, part, kernel, acc, extra);
    } else {
        // Stokes_Step_Linear (HL(t) - HL(t - dt)) or the Heaviside kernel's HL(t), as k_pair_step
        const bool heaviside = kernel == 2;
        TimeCombo tc;
        tc.t[0] = t - dt;
        tc.t[1] = t;
        tc.t[2] = 0.0;
        const double prev = heaviside ? 0.0 : -1.0;
        tc.cI[0] = part == 0 ? prev : 0.0;
        tc.cJ[0] = part == 1 ? prev : 0.0;
        tc.cI[1] = part == 0 ? 1.0 : 0.0;
        tc.cJ[1] = part == 1 ? 1.0 : 0.0;
        tc.cI[2] = 0.0;
        tc.cJ[2] = 0.0;
        tc.static_ok = !heaviside;
        integrate_points<ENND>(points, coll, tc, dt, m, acc, 1, extra);
    }
    // C = cw * (-sum wJ K^T) for the others
    double c[9];
#pragma unroll
    for (int v = 0; v < 9; v++) c[v] = 0.0;
    if (cw != 0.0 && !into_h) {
        points([&](const double* x, const double* n, double wJ, const double*) {
            double KT[9];
            kelvin_at(x, n, KT);
#pragma unroll
            for (int v = 0; v < 9; v++) c[v] -= wJ * KT[v];
        });
#pragma unroll
        for (int v = 0; v < 9; v++) c[v] *= cw;
    }
#pragma unroll
    for (int a = 0; a < ENND; a++)
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const int lo = r < q ? r : q, hi = r < q ? q : r;
                const int o = lo == 0 ? hi : (lo == 1 ? 2 + hi : 5);
                GE[r * 3 * ENND + 3 * a + q] = acc[a * kVals + o];
                HE[r * 3 * ENND + 3 * a + q] = acc[a * kVals + 6 + 3 * r + q];
            }
#pragma unroll
    for (int v = 0; v < 9; v++) C[v] = c[v];
}

// Dense row blocks of one step: thread per (row node il, column node j).
__device__ __forceinline__ void put_block(double* M, int64_t ncol, int il, int j, const double* g,
                                          bool sym)
{
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            double v;
            if (sym) {
                const int a = p < q ? p : q, b = p < q ? q : p;
                const int o = a == 0 ? b : (a == 1 ? 2 + b : 5);
                v = g[o];
            } else {
                v = g[3 * p + q];
            }
            M[(int64_t)(3 * il + p) * ncol + 3 * j + q] = v;
        }
}

template <int ENND>
__global__ void k_dense_gather(const double* __restrict__ scratch, const int* __restrict__ adj_off,
                               const int* __restrict__ adj, const double* __restrict__ cterm,
                               int nnd, int nel, int i0, int nrow_nodes, double c1s, double c2s,
                               double ext,
                               double* __restrict__ G1, double* __restrict__ G2,
                               double* __restrict__ H1, double* __restrict__ H2)
{
    const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (tid >= (int64_t)nrow_nodes * nnd) return;
    const int il = (int)(tid / nnd), j = (int)(tid % nnd);
    double v1[kVals], v2[kVals];
#pragma unroll
    for (int v = 0; v < kVals; v++) v1[v] = v2[v] = 0.0;
    for (int a = adj_off[j]; a < adj_off[j + 1]; a++) {
        const int e = adj[a] >> 3, loc = adj[a] & 7;
        const double* s = scratch + ((int64_t)il * nel + e) * 2 * (ENND * kVals);
#pragma unroll
        for (int v = 0; v < kVals; v++) {
            v1[v] += s[loc * kVals + v];
            v2[v] += s[ENND * kVals + loc * kVals + v];
        }
    }
    const int i = i0 + il;
    if (j == i) {
#pragma unroll
        for (int p = 0; p < 3; p++)
#pragma unroll
            for (int q = 0; q < 3; q++) {
                const double c = cterm[9 * i + 3 * p + q], x = p == q ? ext : 0.0;
                v1[6 + 3 * p + q] += (c1s != 0.0 ? c1s * c : 0.0) + x;
                v2[6 + 3 * p + q] += (c2s != 0.0 ? c2s * c : 0.0) + x;
            }
    }
    const int64_t ncol = 3LL * nnd;
    put_block(G1, ncol, il, j, v1, true);
    put_block(G2, ncol, il, j, v2, true);
    put_block(H1, ncol, il, j, v1 + 6, false);
    put_block(H2, ncol, il, j, v2 + 6, false);
}

// Finds the row node owning a chunk (row_chunk is sorted, length nnd+1).
__device__ __forceinline__ int row_of_chunk(const int64_t* row_chunk, int nnd, int64_t c)
{
    int lo = 0, hi = nnd;  // invariant: row_chunk[lo] <= c < row_chunk[hi]
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (row_chunk[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Gathers lag-store entries of row nodes [i0, i1): one 64-thread block per chunk.
template <int ENND>
__global__ void __launch_bounds__(kChunk)
    k_store_gather(const double* __restrict__ scratch, const int64_t* __restrict__ pair_slot,
                   const int2* __restrict__ pair_k, const int* __restrict__ adj_off,
                   const int* __restrict__ adj, const double* __restrict__ cterm,
                   const int64_t* __restrict__ row_chunk, const int64_t* __restrict__ row_valid,
                   const uint32_t* __restrict__ idx, int nrows, int nel, int i0, double ext,
                   int64_t chunk0, int64_t nslots, double* __restrict__ vals,
                   const int* __restrict__ rows)
{
    // row r of the store layout: node r (slots from node i0 on), or with `rows` node rows[r]
    // (a batch of the lag kernel's slot pass in its own row order, slots from position 0)
    const int64_t c = chunk0 + blockIdx.x;
    const int r = row_of_chunk(row_chunk, nrows, c);
    const int i = rows ? rows[r] : r;
    const int lane = threadIdx.x;
    const int64_t in_row = (c - row_chunk[r]) * kChunk + lane;
    double v[kVals];
#pragma unroll
    for (int t = 0; t < kVals; t++) v[t] = 0.0;
    if (in_row < row_valid[r]) {
        const uint32_t id = idx[c * kChunk + lane];
        const int j = (int)(id >> kKBits), k = (int)(id & kMaxLag);
        for (int a = adj_off[j]; a < adj_off[j + 1]; a++) {
            const int e = adj[a] >> 3, loc = adj[a] & 7;
            const int2 pk = pair_k[(int64_t)i * nel + e];
            if (k < pk.x || k > pk.y) continue;
            const int64_t p = (int64_t)(rows ? r : i - i0) * nel + e;
            const double* s = scratch + (pair_slot[p] + (k - pk.x)) + loc * kVals * nslots;
#pragma unroll
            for (int t = 0; t < kVals; t++) v[t] += s[t * nslots];
        }
        if (j == i && k <= 1) {
            // C-term and the exterior-problem term go to fH_1 and fH_2 of step 1, i.e. into
            // M_1 (phi_1 part) and M_0 (phi_2 part): src/ElastoDynamicsT.cpp:204-225.
#pragma unroll
            for (int p = 0; p < 3; p++)
#pragma unroll
                for (int q = 0; q < 3; q++)
                    v[6 + 3 * p + q] += cterm[9 * i + 3 * p + q] + (p == q ? ext : 0.0);
        }
    }
    double* o = vals + c * (kVals * kChunk);
#pragma unroll
    for (int t = 0; t < kVals; t++) o[t * kChunk + lane] = v[t];
}

// Dense rows of G and H of one lag k from the store: the chunks from chunk0 on belong to row
// nodes lower.., whose rows land at 3 (i - lower) of the [3 (upper - lower + 1)][3 nnd] block.
__global__ void __launch_bounds__(kChunk)
    k_store_to_dense(const double* __restrict__ vals, const uint32_t* __restrict__ idx,
                     const int64_t* __restrict__ row_chunk, const int64_t* __restrict__ row_valid,
                     int nnd, int k, int lower, int64_t chunk0, double* __restrict__ G,
                     double* __restrict__ H)
{
    const int64_t c = chunk0 + blockIdx.x;
    const int i = row_of_chunk(row_chunk, nnd, c);
    const int lane = threadIdx.x;
    const int64_t in_row = (c - row_chunk[i]) * kChunk + lane;
    if (in_row >= row_valid[i]) return;
    const uint32_t id = idx[c * kChunk + lane];
    if ((int)(id & kMaxLag) != k) return;
    const int j = (int)(id >> kKBits);
    double v[kVals];
    const double* s = vals + c * (kVals * kChunk) + lane;
#pragma unroll
    for (int t = 0; t < kVals; t++) v[t] = s[t * kChunk];
    const int64_t ncol = 3LL * nnd;
    put_block(G, ncol, i - lower, j, v, true);
    put_block(H, ncol, i - lower, j, v + 6, false);
}

template <int ENND>
void launch_geom(Context& c)
{
    const int n = c.nel * c.nq * c.nq;
    hipLaunchKernelGGL(k_geom<ENND>, dim3(grid_for(n)), dim3(kBlock), 0, c.stream, c.coords.p,
                       c.ien.p, c.gauss.p, c.nq, c.nel, c.geo.p);
    TD_HIP(hipGetLastError());
}

}  // namespace

// ---------------------------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------------------------
void prepare_geometry(Context& c)
{
    const int GS = 7 + c.ennd;
    c.geo.alloc((size_t)c.nel * c.nq * c.nq * GS);
    if (c.ennd == 4)
        launch_geom<4>(c);
    else
        launch_geom<8>(c);
    // node -> (element, local node) adjacency, in element order
    std::vector<int> cnt(c.nnd + 1, 0);
    for (int e = 0; e < c.nel; e++)
        for (int a = 0; a < c.ennd; a++) cnt[c.h_ien[(size_t)e * c.ennd + a] + 1]++;
    for (int j = 0; j < c.nnd; j++) cnt[j + 1] += cnt[j];
    c.h_adj_off = cnt;
    c.h_adj.assign(cnt[c.nnd], 0);
    std::vector<int> fill(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < c.nel; e++)
        for (int a = 0; a < c.ennd; a++) {
            const int j = c.h_ien[(size_t)e * c.ennd + a];
            c.h_adj[fill[j]++] = (e << 3) | a;
        }
    c.adj_off.upload(c.h_adj_off.data(), c.h_adj_off.size(), c.stream);
    c.adj.upload(c.h_adj.data(), c.h_adj.size(), c.stream);
}

void compute_cterm(Context& c)
{
    if (c.cterm_ready) return;
    c.cterm.alloc((size_t)c.nnd * 9);
    if (c.ennd == 4)
        hipLaunchKernelGGL(k_cterm<4>, dim3(c.nnd), dim3(kBlock), 0, c.stream, c.coords.p,
                           c.ien.p, c.geo.p, c.gauss.p, c.nq, c.nel, c.dt, c.mat, c.cterm.p);
    else
        hipLaunchKernelGGL(k_cterm<8>, dim3(c.nnd), dim3(kBlock), 0, c.stream, c.coords.p,
                           c.ien.p, c.geo.p, c.gauss.p, c.nq, c.nel, c.dt, c.mat, c.cterm.p);
    TD_HIP(hipGetLastError());
    c.cterm_ready = true;
}

void compute_pair_ranges(Context& c)
{
    if (c.ranges_ready) return;
    const int64_t np = (int64_t)c.nnd * c.nel;
    c.pair_k.alloc(np);
    if (c.ennd == 4)
        hipLaunchKernelGGL(k_pair_range<4>, dim3(grid_for(np)), dim3(kBlock), 0, c.stream,
                           c.coords.p, c.ien.p, c.geo.p, c.gauss.p, c.nq, c.nnd, c.nel, c.dt,
                           c.mat, c.Kc, c.pair_k.p);
    else
        hipLaunchKernelGGL(k_pair_range<8>, dim3(grid_for(np)), dim3(kBlock), 0, c.stream,
                           c.coords.p, c.ien.p, c.geo.p, c.gauss.p, c.nq, c.nnd, c.nel, c.dt,
                           c.mat, c.Kc, c.pair_k.p);
    TD_HIP(hipGetLastError());
    c.ranges_ready = true;
}

namespace {
// Scratch budget for the per-pair slots of one row batch.
constexpr int64_t kScratchBytes = 1LL << 30;
}  // namespace

// Weights of the Kelvin block (the C-term) in the phi_1 and phi_2 matrices at time t: the Step
// kernel at its first step only (first: fT == fT_1, src/TDElementStepLinearT.cpp:103), Heaviside
// every step; B-spline and ramp every step with the phi_1 / phi_2 weights of their load pulse,
// Int_CubicB_Linear (src/GreenFunctionT.cpp:733-759) and Int_RampedHeaviside_Linear (:705-731),
// over dt/2.
void kelvin_weights(int kernel, double t, double dt, bool first, double& c1s, double& c2s)
{
    c1s = c2s = 0.0;
    if (kernel == 2 || (kernel == 4 && first)) {
        c1s = c2s = 1.0;
    } else if (kernel == 5) {
        const double bw = 0.7 * dt, J = dt * 0.5;
        double g1 = 0.0, g2 = 0.0;
        for (int q = 0; q < 20; q++) {
            const double ab = kTime20xHost[q], w = kTime20wHost[q];
            const double phi1 = 0.5 * (1 - ab), phi2 = 0.5 * (1 + ab);
            double g, dg;
            cubic_b1(t - phi2 * dt, bw, g, dg);
            g1 += phi1 * g * w * J;
            g2 += phi2 * g * w * J;
        }
        c1s = g1 / (0.5 * dt);
        c2s = g2 / (0.5 * dt);
    } else if (kernel == 6) {
        const double k = 2 / dt, J = dt * 0.5;
        const double *gx, *gw;
        gauss_table(40, &gx, &gw);
        double g1 = 0.0, g2 = 0.0;
        for (int q = 0; q < 40; q++) {
            const double phi1 = 0.5 * (1 - gx[q]), phi2 = 0.5 * (1 + gx[q]);
            const double g = ramp_load(t - phi2 * dt, k);
            g1 += phi1 * g * gw[q] * J;
            g2 += phi2 * g * gw[q] * J;
        }
        c1s = g1 / (0.5 * dt);
        c2s = g2 / (0.5 * dt);
    }
}

void assemble_step_dense(Context& c, int step, int lower, int upper, double* dG1, double* dG2,
                         double* dH1, double* dH2, double* work, int64_t work_doubles)
{
    if (work && !c.cterm_ready) throw std::logic_error("C-term not computed");
    compute_cterm(c);
    // C-term (Kelvin * dt/2 summed over the elements) into H1 and H2
    double c1s, c2s;
    kelvin_weights(c.kernel, step * c.dt, c.dt, step == 1, c1s, c2s);
    const double ext = (c.if_ext && step == 1) ? 0.5 * c.dt : 0.0;
    // per-pair slots of one row batch: in a buffer of this call, or in the caller's workspace
    // (work, work_doubles: nothing is allocated and the call does not wait for its kernels)
    const int64_t per_node = (int64_t)c.nel * 2 * c.ennd * kVals;
    if (work && per_node > work_doubles)
        throw std::invalid_argument("one collocation row's element slots exceed the workspace");
    const int rows_per_batch =
        (int)std::max<int64_t>(1, (work ? work_doubles * (int64_t)sizeof(double) : kScratchBytes) /
                                      (per_node * (int64_t)sizeof(double)));
    DBuf<double> own(c.stream);
    if (!work)
        own.alloc((size_t)std::min(rows_per_batch, upper - lower + 1) * per_node);
    double* const slots = work ? work : own.p;
    const int64_t ncol = 3LL * c.nnd;
    for (int b0 = lower; b0 <= upper; b0 += rows_per_batch) {
        const int nb = std::min(rows_per_batch, upper - b0 + 1);
        const int64_t nslots = (int64_t)nb * c.nel * 2;
        const int64_t roff = (int64_t)(b0 - lower) * 3 * ncol;
        if (c.ennd == 4) {
            hipLaunchKernelGGL(k_pair_step<4>, dim3(grid_for(nslots)), dim3(kBlock), 0, c.stream,
                               c.coords.p, c.ien.p, c.geo.p, c.gauss.p, c.nq, c.nel, c.dt, c.mat,
                               step, c.kernel, b0, nb, slots);
            hipLaunchKernelGGL(k_dense_gather<4>, dim3(grid_for((int64_t)nb * c.nnd)),
                               dim3(kBlock), 0, c.stream, slots, c.adj_off.p, c.adj.p,
                               c.cterm.p, c.nnd, c.nel, b0, nb, c1s, c2s, ext, dG1 + roff,
                               dG2 + roff, dH1 + roff, dH2 + roff);
        } else {
            hipLaunchKernelGGL(k_pair_step<8>, dim3(grid_for(nslots)), dim3(kBlock), 0, c.stream,
                               c.coords.p, c.ien.p, c.geo.p, c.gauss.p, c.nq, c.nel, c.dt, c.mat,
                               step, c.kernel, b0, nb, slots);
            hipLaunchKernelGGL(k_dense_gather<8>, dim3(grid_for((int64_t)nb * c.nnd)),
                               dim3(kBlock), 0, c.stream, slots, c.adj_off.p, c.adj.p,
                               c.cterm.p, c.nnd, c.nel, b0, nb, c1s, c2s, ext, dG1 + roff,
                               dG2 + roff, dH1 + roff, dH2 + roff);
        }
        TD_HIP(hipGetLastError());
    }
    // a buffer of this call must outlive the kernels that use it
    if (!work) TD_HIP(hipStreamSynchronize(c.stream));
}

void element_form(Element& el, int64_t npairs, const double* X, const double* coll,
                  const int* sing, double t, double t_1, double* out)
{
    if (npairs <= 0) return;
    // the Step kernel's Kelvin block at its first step only: fT == fT_1, compared exactly as the
    // reference compares them
    double cw1, cw2;
    kelvin_weights(el.kernel, t, t_1, t == t_1, cw1, cw2);
    const unsigned grid = grid_for(2 * npairs);
    if (el.ennd == 4)
        hipLaunchKernelGGL(k_element<4>, dim3(grid), dim3(kBlock), 0, el.stream, X, coll, sing,
                           npairs, el.gauss.p, el.nq, el.kernel, t, t_1, el.mat, cw1, cw2, out);
    else
        hipLaunchKernelGGL(k_element<8>, dim3(grid), dim3(kBlock), 0, el.stream, X, coll, sing,
                           npairs, el.gauss.p, el.nq, el.kernel, t, t_1, el.mat, cw1, cw2, out);
    TD_HIP(hipGetLastError());
}

namespace {

// Pair ranges clipped to lags <= kmax (the FP64 store of the mixed-precision build keeps the
// lags of the step-0 matrices only).
__global__ void k_clip_pairs(const int2* __restrict__ in, int64_t n, int kmax,
                             int2* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const int2 v = in[t];
    out[t] = (v.y < v.x || v.x > kmax) ? make_int2(1, 0) : make_int2(v.x, min(v.y, kmax));
}

struct ShellPlan {
    int2* hull = nullptr;            // [nnd*nnd] (the context's pinned copy)
    int64_t* off = nullptr;          // [nnd*nnd] entry offsets (pinned)
    std::vector<int64_t> row_chunk;  // [nnd+1]
    std::vector<int64_t> row_valid;  // [nnd]
    int64_t nonzero = 0;
};

void plan_from_hull(Context& c, ShellPlan& P, bool offsets);

// Node-pair hulls (lag range of each (i, j), from the pair ranges dpk of the elements around j)
// into the pinned c.pin_hull, entries per row node and their total; with offsets = true also the
// store layout (row chunks, first entry of every pair in c.pin_off).
void plan_shells(Context& c, ShellPlan& P, const int2* dpk, bool offsets = true)
{
    const int64_t nn = (int64_t)c.nnd * c.nnd;
    DBuf<int2> hull(c.stream);
    hull.alloc(nn);
    hipLaunchKernelGGL(k_node_hull, dim3(grid_for(nn)), dim3(kBlock), 0, c.stream, dpk,
                       c.adj_off.p, c.adj.p, c.nnd, c.nel, hull.p);
    TD_HIP(hipGetLastError());
    c.pin_hull.reserve(nn);
    P.hull = c.pin_hull.p;
    TD_HIP(hipMemcpyAsync(P.hull, hull.p, nn * sizeof(int2), hipMemcpyDeviceToHost,
                          c.stream));
    TD_HIP(hipStreamSynchronize(c.stream));
    plan_from_hull(c, P, offsets);
}

// Entries per row node, row chunks and (offsets = true) each pair's first entry, from the
// node-pair hulls in P.hull.
void plan_from_hull(Context& c, ShellPlan& P, bool offsets)
{
    const int64_t nn = (int64_t)c.nnd * c.nnd;
    P.row_chunk.assign(c.nnd + 1, 0);
    P.row_valid.assign(c.nnd, 0);
    P.nonzero = 0;
    // entries per row node, then each row's first chunk, then the entry offset of every pair
    const int nnd = c.nnd;
    parallel_for(nnd, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            int64_t n = 0;
            for (int j = 0; j < nnd; j++) {
                const int2 h = P.hull[i * nnd + j];
                if (h.y >= h.x) n += h.y - h.x + 1;
            }
            P.row_valid[i] = n;
        }
    });
    int64_t chunk = 0;
    for (int i = 0; i < nnd; i++) {
        P.row_chunk[i] = chunk;
        P.nonzero += P.row_valid[i];
        chunk += (P.row_valid[i] + kChunk - 1) / kChunk;
    }
    P.row_chunk[nnd] = chunk;
    if (!offsets) return;
    c.pin_off.reserve(nn);
    P.off = c.pin_off.p;
    parallel_for(nnd, [&](int64_t lo, int64_t hi) {
        for (int64_t i = lo; i < hi; i++) {
            int64_t o = P.row_chunk[i] * kChunk;
            for (int j = 0; j < nnd; j++) {
                const int2 h = P.hull[i * nnd + j];
                P.off[i * nnd + j] = o;
                if (h.y >= h.x) o += h.y - h.x + 1;
            }
        }
    });
}

// One pass of the lag kernel over the rows `order` (row node per position, -1 = none; empty =
// every node in order), in batches of whole multiples of `align` positions whose slots fit the
// scratch budget.  hpk / dpk: the pair ranges (host, device) the slots cover.  After the lag
// kernel of a batch of positions [b0, b1), consume(b0, b1, scratch, nslots, pair_slot) launches
// the gather from its slots; the slot buffers are reused by the next batch in stream order.
template <int ENND, typename Consume>
void slot_pass(Context& c, const std::vector<int>& order, int align, const int2* hpk,
               const int2* dpk, int kmax, Consume&& consume)
{
    // the host ranges clipped to lags <= kmax as k_clip_pairs clips the device ones
    auto range = [kmax](int2 k) {
        return (k.y < k.x || k.x > kmax) ? make_int2(1, 0) : make_int2(k.x, std::min(k.y, kmax));
    };
    const bool ident = order.empty();
    const int npos = ident ? c.nnd : (int)order.size();
    auto node_at = [&](int pos) { return ident ? pos : order[pos]; };
    const int64_t slot_bytes = (int64_t)ENND * kVals * sizeof(double);
    const bool prof = c.profile == 1;
    KernelTimer lag_timer(c.stream, prof), gather_timer(c.stream, prof);
    if (prof && !c.counters.p) {
        c.counters.alloc(2);
        TD_HIP(hipMemsetAsync(c.counters.p, 0, 2 * sizeof(unsigned long long), c.stream));
    }
    // slots (lags of in-shell element pairs) per position, then batches of whole `align` groups
    // of positions whose slots fit the scratch budget
    std::vector<int64_t> rs(npos, 0);
    parallel_for(npos, [&](int64_t lo, int64_t hi) {
        for (int64_t q = lo; q < hi; q++) {
            const int i = node_at((int)q);
            if (i < 0) continue;
            int64_t n = 0;
            for (int e = 0; e < c.nel; e++) {
                const int2 k = range(hpk[(int64_t)i * c.nel + e]);
                if (k.y >= k.x) n += k.y - k.x + 1;
            }
            rs[q] = n;
        }
    });
    struct Batch { int b0, b1; int64_t slots; };
    std::vector<Batch> batches;
    int64_t max_slots = 1;
    for (int b0 = 0; b0 < npos;) {
        int b1 = b0;
        int64_t slots = 0;
        while (b1 < npos) {
            const int g1 = std::min(npos, b1 + align);
            int64_t add = 0;
            for (int q = b1; q < g1; q++) add += rs[q];
            if (b1 > b0 && (slots + add) * slot_bytes > kScratchBytes) break;
            slots += add;
            b1 = g1;
        }
        batches.push_back({b0, b1, slots});
        max_slots = std::max(max_slots, slots);
        b0 = b1;
    }
    DBuf<double> scratch(c.stream);
    DBuf<int> slot_pair(c.stream), d_rows(c.stream);
    DBuf<int64_t> pair_slot(c.stream);
    scratch.reserve((size_t)max_slots * ENND * kVals);
    int max_rows = 1;
    for (const Batch& bt : batches) max_rows = std::max(max_rows, bt.b1 - bt.b0);
    slot_pair.reserve((size_t)max_slots);
    pair_slot.reserve((size_t)max_rows * c.nel);
    if (!ident) d_rows.reserve((size_t)max_rows);
    // The host builds batch b + 1's pair offsets while the device works on batch b: two host
    // sets, each reused only after the event that follows its upload; the device buffers are
    // rewritten in stream order after the previous batch's kernels.  The slot list itself is
    // expanded on the device (k_slot_fill) from the pairs' first slots.  (Pageable host sets: a
    // pinned allocation would be freed with hipHostFree, which waits for the whole device.)
    std::vector<int64_t> h_pair_slot[2];
    std::vector<int64_t> h_sing[2];
    std::vector<int> h_rows[2];
    DBuf<int64_t> sing_slots(c.stream);
    hipEvent_t uploaded[2];
    TD_HIP(hipEventCreateWithFlags(&uploaded[0], hipEventDisableTiming));
    TD_HIP(hipEventCreateWithFlags(&uploaded[1], hipEventDisableTiming));
    for (size_t b = 0; b < batches.size(); b++) {
        const int set = (int)(b & 1);
        const int b0 = batches[b].b0, b1 = batches[b].b1;
        const int64_t slots = batches[b].slots;
        const int nb = b1 - b0;
        if (b >= 2) TD_HIP(hipEventSynchronize(uploaded[set]));
        auto& hp = h_pair_slot[set];
        auto& hr = h_rows[set];
        hp.resize((size_t)nb * c.nel);
        hr.resize(nb);
        std::vector<int64_t> row_start(nb + 1, 0);
        for (int r = 0; r < nb; r++) {
            row_start[r + 1] = row_start[r] + rs[b0 + r];
            hr[r] = node_at(b0 + r);
        }
        // slots of the pairs with a node at the collocation point, per row
        std::vector<std::vector<int64_t>> near_slots(nb);
        parallel_for(nb, [&](int64_t lo, int64_t hi) {
            for (int64_t r = lo; r < hi; r++) {
                const int i = hr[r];
                int64_t sl = row_start[r];
                if (i < 0) {
                    for (int e = 0; e < c.nel; e++) hp[r * c.nel + e] = sl;
                    continue;
                }
                const double* xc = &c.h_coords[3 * (int64_t)i];
                for (int e = 0; e < c.nel; e++) {
                    hp[r * c.nel + e] = sl;
                    const int2 k = range(hpk[(int64_t)i * c.nel + e]);
                    if (k.y < k.x) continue;
                    // the device's singular test is a squared distance below 1e-9; a pair goes
                    // to the singular launch well before that (squared distance below 1e-6)
                    bool near = false;
                    for (int a = 0; a < ENND; a++) {
                        const double* xa = &c.h_coords[3 * (int64_t)c.h_ien[(int64_t)e * ENND + a]];
                        const double dx = xa[0] - xc[0], dy = xa[1] - xc[1], dz = xa[2] - xc[2];
                        near = near || dx * dx + dy * dy + dz * dz < 1e-6;
                    }
                    const int64_t n = k.y - k.x + 1;
                    if (near)
                        for (int64_t t = 0; t < n; t++) near_slots[r].push_back(sl + t);
                    sl += n;
                }
            }
        });
        auto& hsing = h_sing[set];
        hsing.clear();
        for (const auto& v : near_slots) hsing.insert(hsing.end(), v.begin(), v.end());
        const int64_t npairs = (int64_t)nb * c.nel;
        pair_slot.upload(hp.data(), (size_t)npairs, c.stream);
        if (!ident) d_rows.upload(hr.data(), hr.size(), c.stream);
        if (!hsing.empty()) sing_slots.upload(hsing.data(), hsing.size(), c.stream);
        TD_HIP(hipEventRecord(uploaded[set], c.stream));
        if (slots > 0) {
            hipLaunchKernelGGL(k_slot_fill, dim3(grid_for(npairs)), dim3(kBlock), 0, c.stream,
                               pair_slot.p, dpk, ident ? nullptr : d_rows.p, b0, c.nel, npairs,
                               slot_pair.p);
            TD_HIP(hipGetLastError());
            if (!hsing.empty()) {
                hipLaunchKernelGGL(k_slot_mark, dim3(grid_for((int64_t)hsing.size())),
                                   dim3(kBlock), 0, c.stream, sing_slots.p,
                                   (int64_t)hsing.size(), slot_pair.p);
                TD_HIP(hipGetLastError());
            }
        }
        const int* rows = ident ? nullptr : d_rows.p;
        if (slots > 0) {
            lag_timer.begin();
            hipLaunchKernelGGL((k_pair_lags<ENND, false>), dim3(grid_for(slots, kLagBlock)),
                               dim3(kLagBlock), 0,
                               c.stream, c.coords.p, c.ien.p, c.geo.p, c.gauss.p, c.nq, c.nel,
                               c.dt, c.mat, c.Kc, b0, slots, slot_pair.p, pair_slot.p, dpk,
                               sing_slots.p, (int64_t)0, scratch.p,
                               prof ? c.counters.p : nullptr, rows);
            TD_HIP(hipGetLastError());
            const int64_t nsing = (int64_t)hsing.size();
            if (nsing > 0) {
                hipLaunchKernelGGL((k_pair_lags<ENND, true>), dim3(grid_for(nsing, kLagBlock)),
                                   dim3(kLagBlock), 0, c.stream, c.coords.p, c.ien.p, c.geo.p,
                                   c.gauss.p, c.nq, c.nel, c.dt, c.mat, c.Kc, b0, slots,
                                   slot_pair.p, pair_slot.p, dpk, sing_slots.p, nsing,
                                   scratch.p, prof ? c.counters.p : nullptr, rows);
                TD_HIP(hipGetLastError());
            }
            lag_timer.end();
        }
        gather_timer.begin();
        consume(b0, b1, (const double*)scratch.p, slots, (const int64_t*)pair_slot.p);
        gather_timer.end();
    }
    TD_HIP(hipStreamSynchronize(c.stream));
    TD_HIP(hipEventDestroy(uploaded[0]));
    TD_HIP(hipEventDestroy(uploaded[1]));
    if (prof) {
        c.tm.lag_launches += lag_timer.count();
        c.tm.lag_kernel_ms += lag_timer.collect();
        c.tm.gather_ms += gather_timer.collect();
    }
}

// The FP64 store of the rows' in-shell lags (pair ranges hpk / dpk, layout P): slots of
// contiguous row ranges, gathered per (i, j, k) entry by k_store_gather.
template <int ENND>
void fill_store(Context& c, const ShellPlan& P, const int2* hpk, const int2* dpk, int kmax)
{
    LagStore& S = c.store;
    const double ext = c.if_ext ? 0.5 * c.dt : 0.0;
    slot_pass<ENND>(c, std::vector<int>(), 1, hpk, dpk, kmax,
                    [&](int i0, int i1, const double* scratch, int64_t nslots,
                        const int64_t* pair_slot) {
                        const int64_t ch0 = P.row_chunk[i0], ch1 = P.row_chunk[i1];
                        if (ch1 <= ch0) return;
                        hipLaunchKernelGGL(k_store_gather<ENND>, dim3((unsigned)(ch1 - ch0)),
                                           dim3(kChunk), 0, c.stream, scratch, pair_slot, dpk,
                                           c.adj_off.p, c.adj.p, c.cterm.p, S.row_chunk.p,
                                           S.row_valid.p, S.idx.p, c.nnd, c.nel, i0, ext, ch0,
                                           nslots, S.vals.p, (const int*)nullptr);
                        TD_HIP(hipGetLastError());
                    });
}

// Allocates the FP64 store of layout P (pair hulls in P.hull, offsets in P.off) and its entry
// index; the values are left to the caller.
void alloc_store(Context& c, const ShellPlan& P)
{
    LagStore& S = c.store;
    S.n_chunks = P.row_chunk[c.nnd];
    S.n_entries = S.n_chunks * kChunk;
    S.nonzero_entries = P.nonzero;
    S.h_row_chunk = P.row_chunk;
    S.h_row_valid = P.row_valid;
    S.vals.alloc(std::max<int64_t>(S.n_chunks, 1) * kVals * kChunk);
    S.idx.alloc(std::max<int64_t>(S.n_entries, 1));
    TD_HIP(hipMemsetAsync(S.idx.p, 0, S.idx.bytes(), c.stream));
    S.row_chunk.upload(P.row_chunk.data(), P.row_chunk.size(), c.stream);
    S.row_valid.upload(P.row_valid.data(), P.row_valid.size(), c.stream);
    const int64_t nn = (int64_t)c.nnd * c.nnd;
    {
        DBuf<int2> d_hull(c.stream);
        DBuf<int64_t> d_off(c.stream);
        d_off.upload(P.off, nn, c.stream);
        d_hull.upload(P.hull, nn, c.stream);
        hipLaunchKernelGGL(k_fill_idx, dim3(grid_for(nn)), dim3(kBlock), 0, c.stream, d_hull.p,
                           d_off.p, c.nnd, (int64_t)c.nnd, S.idx.p);
        TD_HIP(hipGetLastError());
    }
}

// Allocates the FP64 store of layout P and fills it from the lag kernel's slots.
template <int ENND>
void build_fp64_store(Context& c, const ShellPlan& P, const int2* hpk, const int2* dpk,
                      int kmax = kMaxLag)
{
    alloc_store(c, P);
    fill_store<ENND>(c, P, hpk, dpk, kmax);
}

// The pair ranges of c.pair_k on the host (pinned, kept between builds unless large).
const int2* host_pair_ranges(Context& c)
{
    const int64_t np = (int64_t)c.nnd * c.nel;
    c.pin_pk.reserve(np);
    TD_HIP(hipMemcpyAsync(c.pin_pk.p, c.pair_k.p, np * sizeof(int2), hipMemcpyDeviceToHost,
                          c.stream));
    TD_HIP(hipStreamSynchronize(c.stream));
    return c.pin_pk.p;
}

// Mixed-precision store: the tile store of every lag, filled straight from the lag kernel's
// slots in row tiles of the bisection order (no FP64 copy of the whole store is ever held),
// and an FP64 store of lags 0 and 1 only, from which march_init forms the step-0 matrices.
// Rows of one batch of the tile pass in the FP64 store's layout: the entries of row li (node
// rows[li], -1 = none) from chunk rc[li] on, rv[li] of them, each pair's lag range and first
// entry in hull / off [li][nnd].  Two host sets, reused after the event of their upload, as the
// slot pass's lists.
struct SegmentRows {
    std::vector<int64_t> rc, rv, off;
    std::vector<int2> hull;
    std::vector<int> rows;
};

template <int ENND>
void build_mixed(Context& c)
{
    const int64_t np = (int64_t)c.nnd * c.nel;
    const int nnd = c.nnd;
    // node-pair hulls of every lag (host: the plan and each batch's segment layout)
    ShellPlan Pf;
    plan_shells(c, Pf, c.pair_k.p, false);
    c.pair_lags = Pf.nonzero;
    if (!c.tp.planned) tp_plan_structure(c);
    const int2* hpk = host_pair_ranges(c);
    // every lag into the tile store, row tiles in bisection order: per batch of the lag kernel,
    // its rows' FP64 store entries (k_store_gather into a segment), then the sub-tiles of its row
    // tiles from the segment
    tp_alloc(c);
    const double ext = c.if_ext ? 0.5 * c.dt : 0.0;
    {
        DBuf<double> seg(c.stream);
        DBuf<uint32_t> seg_idx(c.stream);
        DBuf<int2> seg_hull(c.stream);
        DBuf<int64_t> seg_off(c.stream), seg_rc(c.stream), seg_rv(c.stream);
        DBuf<int> seg_rows(c.stream);
        SegmentRows h[2];
        hipEvent_t uploaded[2];
        TD_HIP(hipEventCreateWithFlags(&uploaded[0], hipEventDisableTiming));
        TD_HIP(hipEventCreateWithFlags(&uploaded[1], hipEventDisableTiming));
        int nbatch = 0;
        try {
            slot_pass<ENND>(
                c, c.tp.perm, kTile, hpk, c.pair_k.p, kMaxLag,
                [&](int b0, int b1, const double* scratch, int64_t nslots,
                    const int64_t* pair_slot) {
                    const int set = nbatch & 1;
                    if (nbatch >= 2) TD_HIP(hipEventSynchronize(uploaded[set]));
                    nbatch++;
                    SegmentRows& H = h[set];
                    const int nb = b1 - b0;
                    H.rc.assign(nb + 1, 0);
                    H.rv.assign(nb, 0);
                    H.rows.resize(nb);
                    for (int li = 0; li < nb; li++) {
                        const int i = c.tp.perm[b0 + li];
                        H.rows[li] = i;
                        H.rv[li] = i >= 0 ? Pf.row_valid[i] : 0;
                        H.rc[li + 1] = H.rc[li] + (H.rv[li] + kChunk - 1) / kChunk;
                    }
                    H.off.resize((size_t)nb * nnd);
                    H.hull.resize((size_t)nb * nnd);
                    parallel_for(nb, [&](int64_t lo, int64_t hi) {
                        for (int64_t li = lo; li < hi; li++) {
                            const int i = H.rows[li];
                            int64_t o = H.rc[li] * kChunk;
                            for (int j = 0; j < nnd; j++) {
                                const int2 hj = i >= 0 ? Pf.hull[(int64_t)i * nnd + j]
                                                       : make_int2(1, 0);
                                H.hull[li * nnd + j] = hj;
                                H.off[li * nnd + j] = o;
                                if (hj.y >= hj.x) o += hj.y - hj.x + 1;
                            }
                        }
                    });
                    const int64_t chunks = H.rc[nb];
                    seg_rc.upload(H.rc.data(), H.rc.size(), c.stream);
                    seg_rv.upload(H.rv.data(), H.rv.size(), c.stream);
                    seg_rows.upload(H.rows.data(), H.rows.size(), c.stream);
                    seg_hull.upload(H.hull.data(), H.hull.size(), c.stream);
                    seg_off.upload(H.off.data(), H.off.size(), c.stream);
                    TD_HIP(hipEventRecord(uploaded[set], c.stream));
                    if (chunks == 0) return;
                    seg.reserve((size_t)chunks * kChunk * kVals);
                    seg_idx.reserve((size_t)chunks * kChunk);
                    const int64_t ncell = (int64_t)nb * nnd;
                    hipLaunchKernelGGL(k_fill_idx, dim3(grid_for(ncell)), dim3(kBlock), 0,
                                       c.stream, seg_hull.p, seg_off.p, nnd, (int64_t)nb,
                                       seg_idx.p);
                    hipLaunchKernelGGL(k_store_gather<ENND>, dim3((unsigned)chunks),
                                       dim3(kChunk), 0, c.stream, scratch, pair_slot,
                                       c.pair_k.p, c.adj_off.p, c.adj.p, c.cterm.p, seg_rc.p,
                                       seg_rv.p, seg_idx.p, nb, c.nel, 0, ext, (int64_t)0,
                                       nslots, seg.p, (const int*)seg_rows.p);
                    TD_HIP(hipGetLastError());
                    tp_gather_segment(c, b0 / kTile, b1 / kTile, seg.p, seg_hull.p, seg_off.p);
                });
        } catch (...) {
            (void)hipStreamSynchronize(c.stream);
            (void)hipEventDestroy(uploaded[0]);
            (void)hipEventDestroy(uploaded[1]);
            throw;
        }
        TD_HIP(hipStreamSynchronize(c.stream));
        TD_HIP(hipEventDestroy(uploaded[0]));
        TD_HIP(hipEventDestroy(uploaded[1]));
    }
    tp_finish(c);
    // lags 0 and 1 in FP64 (their planning replaces the pinned node-pair hulls)
    {
        DBuf<int2> pk01(c.stream);
        pk01.alloc(np);
        hipLaunchKernelGGL(k_clip_pairs, dim3(grid_for(np)), dim3(kBlock), 0, c.stream,
                           c.pair_k.p, np, 1, pk01.p);
        TD_HIP(hipGetLastError());
        ShellPlan P01;
        plan_shells(c, P01, pk01.p);
        // the slot pass clips the host ranges to lags 0..1 itself, as pk01 on the device
        build_fp64_store<ENND>(c, P01, hpk, pk01.p, 1);
    }
}

}  // namespace

void alloc_store_for_hull(Context& c, const int2* hull)
{
    const int64_t nn = (int64_t)c.nnd * c.nnd;
    c.pin_hull.reserve(nn);
    std::copy(hull, hull + nn, c.pin_hull.p);
    ShellPlan P;
    P.hull = c.pin_hull.p;
    plan_from_hull(c, P, true);
    c.pair_lags = P.nonzero;
    alloc_store(c, P);
    TD_HIP(hipStreamSynchronize(c.stream));
}

namespace {

// Largest slot count and row count of the batches slot_pass forms over `order` (see there).
template <int ENND>
void slot_batches(Context& c, const std::vector<int>& order, int align, const int2* hpk,
                  int64_t& max_slots, int& max_rows, const std::vector<int64_t>* row_entries = nullptr,
                  int64_t* max_seg_chunks = nullptr)
{
    const bool ident = order.empty();
    const int npos = ident ? c.nnd : (int)order.size();
    std::vector<int64_t> rs(npos, 0);
    parallel_for(npos, [&](int64_t lo, int64_t hi) {
        for (int64_t q = lo; q < hi; q++) {
            const int i = ident ? (int)q : order[q];
            if (i < 0) continue;
            int64_t n = 0;
            for (int e = 0; e < c.nel; e++) {
                const int2 k = hpk[(int64_t)i * c.nel + e];
                if (k.y >= k.x) n += k.y - k.x + 1;
            }
            rs[q] = n;
        }
    });
    const int64_t slot_bytes = (int64_t)ENND * kVals * sizeof(double);
    max_slots = 1;
    max_rows = 1;
    for (int b0 = 0; b0 < npos;) {
        int b1 = b0;
        int64_t slots = 0;
        while (b1 < npos) {
            const int g1 = std::min(npos, b1 + align);
            int64_t add = 0;
            for (int q = b1; q < g1; q++) add += rs[q];
            if (b1 > b0 && (slots + add) * slot_bytes > kScratchBytes) break;
            slots += add;
            b1 = g1;
        }
        max_slots = std::max(max_slots, slots);
        max_rows = std::max(max_rows, b1 - b0);
        if (row_entries && max_seg_chunks) {
            // the batch's rows in the FP64 store's layout (the tile pass's segment)
            int64_t ch = 0;
            for (int q = b0; q < b1; q++) {
                const int i = ident ? q : order[q];
                if (i >= 0) ch += ((*row_entries)[i] + kChunk - 1) / kChunk;
            }
            *max_seg_chunks = std::max(*max_seg_chunks, ch);
        }
        b0 = b1;
    }
}

// Device bytes of one slot pass's buffers (scratch, slot lists, row map).
int64_t slot_pass_bytes(const Context& c, int64_t max_slots, int max_rows)
{
    return max_slots * ((int64_t)c.ennd * kVals * 8 + 4 + 8) + (int64_t)max_rows * (c.nel * 8 + 4);
}

}  // namespace

void predict_store(Context& c, int64_t& store, int64_t& transient)
{
    compute_pair_ranges(c);
    const int64_t nn = (int64_t)c.nnd * c.nnd, np = (int64_t)c.nnd * c.nel;
    const int64_t rows_bytes = (int64_t)(c.nnd + 1) * 8 + (int64_t)c.nnd * 8;
    const int64_t entry_bytes = kChunk * (int64_t)(kVals * sizeof(double) + sizeof(uint32_t));
    const int2* hpk = host_pair_ranges(c);
    int64_t max_slots = 0;
    int max_rows = 0;
    if (c.precision == 1) {
        ShellPlan Pf;
        plan_shells(c, Pf, c.pair_k.p, false);
        c.pair_lags = Pf.nonzero;
        if (!c.tp.planned) tp_plan_structure(c);
        const TPStore& T = c.tp;
        // the FP64 store of lags 0 and 1
        DBuf<int2> pk01(c.stream);
        pk01.alloc(np);
        hipLaunchKernelGGL(k_clip_pairs, dim3(grid_for(np)), dim3(kBlock), 0, c.stream,
                           c.pair_k.p, np, 1, pk01.p);
        TD_HIP(hipGetLastError());
        ShellPlan P01;
        plan_shells(c, P01, pk01.p, false);
        const int64_t mini = P01.row_chunk[c.nnd] * entry_bytes + rows_bytes;
        store = T.nsub * (int64_t)kSubFloats * 4 + T.nentries * (int64_t)sizeof(TPEntry) +
                (int64_t)T.nT * kTile * 4 + (int64_t)c.nnd * 4 + (int64_t)(T.nG + 1) * 8 + mini;
        int64_t seg_chunks = 0;
        if (c.ennd == 4)
            slot_batches<4>(c, T.perm, kTile, hpk, max_slots, max_rows, &Pf.row_valid, &seg_chunks);
        else
            slot_batches<8>(c, T.perm, kTile, hpk, max_slots, max_rows, &Pf.row_valid, &seg_chunks);
        // the tile pass (sub-tile list, slot buffers and one batch's FP64 segment: values, entry
        // index, row tables and the pairs' hulls and first entries) or the lags-0..1 pass
        const int64_t segment = seg_chunks * kChunk * (kVals * 8 + 4) +
                                (int64_t)max_rows * c.nnd * 16 + (int64_t)max_rows * 28 + 8;
        transient = std::max(nn * 16 + np * 8,
                             T.nsub * 8 + slot_pass_bytes(c, max_slots, max_rows) + segment);
    } else {
        ShellPlan P;
        plan_shells(c, P, c.pair_k.p, false);
        store = P.row_chunk[c.nnd] * entry_bytes + rows_bytes;
        if (c.ennd == 4)
            slot_batches<4>(c, std::vector<int>(), 1, hpk, max_slots, max_rows);
        else
            slot_batches<8>(c, std::vector<int>(), 1, hpk, max_slots, max_rows);
        transient = std::max(nn * 16, slot_pass_bytes(c, max_slots, max_rows));
    }
}

int64_t predict_store_bytes(Context& c)
{
    int64_t store = 0, transient = 0;
    predict_store(c, store, transient);
    return store;
}

void build_lag_store(Context& c)
{
    if (c.Kc > kMaxLag) throw std::runtime_error("lag count exceeds the store's 10-bit lag index");
    const auto t0 = std::chrono::steady_clock::now();
    c.store_ready = false;
    c.tp.ready = false;
    c.store_precision = c.precision;
    if (c.profile == 1 && c.counters.p)
        TD_HIP(hipMemsetAsync(c.counters.p, 0, 2 * sizeof(unsigned long long), c.stream));
    compute_cterm(c);
    compute_pair_ranges(c);
    if (c.precision == 1) {
        if (c.ennd == 4)
            build_mixed<4>(c);
        else
            build_mixed<8>(c);
        c.store_lags = std::min(1, c.Kc);
    } else {
        // a tile store of an earlier mixed-precision build is released
        c.tp.vals.reset();
        c.tp.ent.reset();
        c.tp.planned = false;
        ShellPlan P;
        plan_shells(c, P, c.pair_k.p);
        c.pair_lags = P.nonzero;
        const int2* hpk = host_pair_ranges(c);
        if (c.ennd == 4)
            build_fp64_store<4>(c, P, hpk, c.pair_k.p);
        else
            build_fp64_store<8>(c, P, hpk, c.pair_k.p);
        c.store_lags = c.Kc;
    }
    TD_HIP(hipStreamSynchronize(c.stream));
    if (c.profile == 1 && c.counters.p) {
        unsigned long long n[2] = {0, 0};
        TD_HIP(hipMemcpy(n, c.counters.p, sizeof(n), hipMemcpyDeviceToHost));
        c.tm.active_points += (int64_t)n[0];
        c.tm.active_levels += (int64_t)n[1];
    }
    // the pinned planning copies stay for the next build unless they are large
    if ((int64_t)c.nnd * c.nnd * 24 > kPinnedKeepBytes) {
        c.pin_off.reserve(0);
        c.pin_pk.reserve(0);
        c.pin_hull.reserve(0);
    }
    c.store_ready = true;
    c.tm.store_ms +=
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void lag_dense(Context& c, int k, double* dG, double* dH) { lag_rows(c, k, 0, c.nnd - 1, dG, dH); }

void lag_rows(Context& c, int k, int lower, int upper, double* dG, double* dH)
{
    if (!c.store_ready) throw std::runtime_error("lag store not built");
    const int64_t n = 3LL * (upper - lower + 1) * 3LL * c.nnd;
    TD_HIP(hipMemsetAsync(dG, 0, n * sizeof(double), c.stream));
    TD_HIP(hipMemsetAsync(dH, 0, n * sizeof(double), c.stream));
    const int64_t ch0 = c.store.h_row_chunk[lower], ch1 = c.store.h_row_chunk[upper + 1];
    if (k < 0 || k > c.Kc || ch1 <= ch0) return;
    hipLaunchKernelGGL(k_store_to_dense, dim3((unsigned)(ch1 - ch0)), dim3(kChunk), 0, c.stream,
                       c.store.vals.p, c.store.idx.p, c.store.row_chunk.p, c.store.row_valid.p,
                       c.nnd, k, lower, ch0, dG, dH);
    TD_HIP(hipGetLastError());
}

}  // namespace tdbem
