// tdbem_device.h — device-side math for the TD-BEM hot paths on CDNA4 (gfx950).
//
// The Stokes fundamental solution of a Heaviside point load, integrated in closed form against
// the two linear time hats (src/GreenFunctionT.cpp:241-416), is written here as
//     U_ik = (ua * rh_i rh_k + ub * d_ik) / (4 pi rho r)
//     T_ik = (al * rh_i rh_k + be * n_i rh_k + ga * n_k rh_i + de * d_ik) / (4 pi)
// with six scalars built from six time integrals S_1..S_6 (the reference's I1..I6, or J1..J6).
// Because every output is linear in S, the Step kernel HL(t) - HL(t - dt)
// (src/GreenFunctionT.cpp:424-446) and the lag combination M_k = phi_1(k) + phi_2(k+1) of
// src/SolverStepLinearT.cpp:65-97 are formed on S before the 3x3 tensors are built.
#pragma once

#include <hip/hip_runtime.h>

namespace tdbem {

constexpr double kPi = 3.14159265358979323846;

struct Material {
    double E, nu, rho, mu, cp, cs;
};

// Six closed-form integrals of one wave against the phi_1 ("I") and phi_2 ("J") hats:
// src/GreenFunctionT.cpp:251-281 (c = c_p) and :285-314 (c = c_s).  out: I1,I3,I5 / J1,J3,J5.
template <typename R>
__device__ __forceinline__ void hat_integrals(R t, R dt, R inv_dt, R r, R inv_c, R inv_c2,
                                              R inv_r2, R& I1, R& I3, R& I5, R& J1, R& J3,
                                              R& J5)
{
    const R t_ref = t - r * inv_c;
    R b;
    if (dt <= t_ref) {
        b = dt;
        I5 = R(0);
        J5 = R(0);
    } else if (R(0) <= t_ref) {
        b = t_ref;
        I5 = (dt - t_ref) * inv_dt;
        J5 = t_ref * inv_dt;
    } else {
        b = R(0);
        I5 = R(0);
        J5 = R(0);
    }
    const R b2 = b * b;
    const R t2 = t * t;
    // (-3b^4 + (8t+4dt) b^3 - (6t^2+12 t dt) b^2 + 12 dt t^2 b) / (12 r^2)
    const R pI = b * (R(12) * dt * t2 + b * (-(R(6) * t2 + R(12) * t * dt) +
                                             b * ((R(8) * t + R(4) * dt) - R(3) * b)));
    const R pJ = b2 * (-R(6) * t2 + b * (R(8) * t - R(3) * b));
    const R q = dt * b - R(0.5) * b2;
    I1 = (pI * inv_r2 * R(1.0 / 12.0) - q * inv_c2) * inv_dt;
    J1 = (-pJ * inv_r2 * R(1.0 / 12.0) - R(0.5) * b2 * inv_c2) * inv_dt;
    I3 = q * inv_dt;
    J3 = R(0.5) * b2 * inv_dt;
}

// S[0..5] += cI * (I1, I2, I3, I4, I5, I6)(t) + cJ * (J1, ..., J6)(t)
template <typename R>
__device__ __forceinline__ void accumulate_hl(R t, R dt, R inv_dt, R r, R inv_r2, R inv_cp,
                                              R inv_cs, R inv_cp2, R inv_cs2, R cI, R cJ,
                                              R* S)
{
    R I1, I3, I5, J1, J3, J5, I2, I4, I6, J2, J4, J6;
    hat_integrals<R>(t, dt, inv_dt, r, inv_cp, inv_cp2, inv_r2, I1, I3, I5, J1, J3, J5);
    hat_integrals<R>(t, dt, inv_dt, r, inv_cs, inv_cs2, inv_r2, I2, I4, I6, J2, J4, J6);
    S[0] += cI * I1 + cJ * J1;
    S[1] += cI * I2 + cJ * J2;
    S[2] += cI * I3 + cJ * J3;
    S[3] += cI * I4 + cJ * J4;
    S[4] += cI * I5 + cJ * J5;
    S[5] += cI * I6 + cJ * J6;
}

// FP32 form of hat_integrals for the mixed-precision assembly.  With a = t - r/c (formed in FP64
// by the caller) and d2 = 2 r / c, (t - tau)^2 / r^2 - 1/c^2 = (a - tau)(a - tau + d2) / r^2, so
//   I1 = 1/(dt r^2) int_0^b (dt - tau)(a - tau)(a - tau + d2) dtau,
//   J1 = 1/(dt r^2) int_0^b tau (a - tau)(a - tau + d2) dtau,
// cubic integrands whose factors are all >= 0 on [0, b] (b <= a): two-point Gauss-Legendre is
// exact for them and adds only positive terms, so nothing cancels near the wave's arrival (the
// expanded form above loses the digits of t^2/r^2 - 1/c^2 there).  I3, I5, J3, J5 as above.
__device__ __forceinline__ void hat_integrals_f32(float a, float dt, float inv_dt, float d2,
                                                  float inv_r2, float& I1, float& I3, float& I5,
                                                  float& J1, float& J3, float& J5)
{
    float b;
    if (dt <= a) {
        b = dt;
        I5 = 0.f;
        J5 = 0.f;
    } else if (0.f <= a) {
        b = a;
        I5 = (dt - a) * inv_dt;
        J5 = a * inv_dt;
    } else {
        b = 0.f;
        I5 = 0.f;
        J5 = 0.f;
    }
    const float h = 0.5f * b, g = 0.57735026918962576f * h;  // 1/sqrt(3)
    const float tm = h - g, tp = h + g;
    const float em = (a - tm) * (a - tm + d2), ep = (a - tp) * (a - tp + d2);
    const float sc = h * inv_r2 * inv_dt;
    I1 = ((dt - tm) * em + (dt - tp) * ep) * sc;
    J1 = (tm * em + tp * ep) * sc;
    I3 = (dt * b - 0.5f * b * b) * inv_dt;
    J3 = 0.5f * b * b * inv_dt;
}

// S[0..5] += cI * (I1..I6)(t) + cJ * (J1..J6)(t) in FP32; the retarded times t - r/c in FP64.
__device__ __forceinline__ void accumulate_hl_f32(double t, double r, float dt, float inv_dt,
                                                  float inv_r2, double inv_cp, double inv_cs,
                                                  float cI, float cJ, float* S)
{
    float I1, I3, I5, J1, J3, J5, I2, I4, I6, J2, J4, J6;
    hat_integrals_f32((float)(t - r * inv_cp), dt, inv_dt, (float)(2.0 * r * inv_cp), inv_r2, I1,
                      I3, I5, J1, J3, J5);
    hat_integrals_f32((float)(t - r * inv_cs), dt, inv_dt, (float)(2.0 * r * inv_cs), inv_r2, I2,
                      I4, I6, J2, J4, J6);
    S[0] += cI * I1 + cJ * J1;
    S[1] += cI * I2 + cJ * J2;
    S[2] += cI * I3 + cJ * J3;
    S[3] += cI * I4 + cJ * J4;
    S[4] += cI * I5 + cJ * J5;
    S[5] += cI * I6 + cJ * J6;
}

// From S to the 15 block values of one point, already transposed into the global block layout
// (row = collocation component p, column = node component q):
//   G[p][q] = U[q][p] (symmetric: g00 g01 g02 g11 g12 g22)
//   H[p][q] = T[q][p] (h00 h01 h02 h10 h11 h12 h20 h21 h22)
// multATB(DU, NMatrix) in src/TDElementStepLinearT.cpp:91-101 is what puts U[q][p] at [p][q].
template <typename R>
__device__ __forceinline__ void blocks_from_s(const R* S, const R* rh, const R* n, R r,
                                              R inv_r, R inv_r2, const Material& m, R* out)
{
    const R cs = R(m.cs), cp = R(m.cp);
    const R inv_cs2 = R(1) / (cs * cs);
    const R k2 = (cs / cp) * (cs / cp);
    const R k3 = k2 * (cs / cp);
    const R A1 = R(0.5) * (S[0] - S[1]);
    const R A2 = S[2] / (cp * cp) - S[3] * inv_cs2;
    const R A3 = S[3];
    const R B2 = S[3] - k2 * S[2];
    const R B3 = S[5] - k3 * S[4];
    const R B4 = S[2] + r / cp * S[4];
    const R B5 = S[3] + r / cs * S[5];
    const R rdn = rh[0] * n[0] + rh[1] * n[1] + rh[2] * n[2];
    const R uscale = R(1) / (R(4 * kPi) * R(m.rho) * r);
    const R ua = (R(3) * A1 + A2) * uscale;
    const R ub = (-A1 + A3 * inv_cs2) * uscale;
    const R w = (R(6) * cs * cs * A1 - R(2) * B2) * inv_r2;
    const R t4 = R(1) / R(4 * kPi);
    const R al = (rdn * (-R(30) * cs * cs * A1 + R(12) * B2) * inv_r2 +
                  R(2) * rdn * B3 * inv_r / cs) * t4;
    const R be = (w - (R(1) - R(2) * k2) * B4 * inv_r2) * t4;
    const R ga = (w - B5 * inv_r2) * t4;
    const R de = (rdn * w - rdn * B5 * inv_r2) * t4;
    // G (symmetric)
    out[0] = ua * rh[0] * rh[0] + ub;
    out[1] = ua * rh[0] * rh[1];
    out[2] = ua * rh[0] * rh[2];
    out[3] = ua * rh[1] * rh[1] + ub;
    out[4] = ua * rh[1] * rh[2];
    out[5] = ua * rh[2] * rh[2] + ub;
    // H[p][q] = al rh_p rh_q + be rh_p n_q + ga n_p rh_q + de d_pq
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++)
            out[6 + 3 * p + q] = al * rh[p] * rh[q] + be * rh[p] * n[q] + ga * n[p] * rh[q] +
                                 (p == q ? de : R(0));
}

// Kelvin traction * dt/2 (src/GreenFunctionT.cpp:205-231), returned transposed as the C-term
// block adds it: C[p][q] -= wJ * K[q][p] (src/TDElementStepLinearT.cpp:103-109).
__device__ __forceinline__ void kelvin_block(const double* rh, const double* n, double r,
                                             double dt, const Material& m, double* KT)
{
    const double rdn = rh[0] * n[0] + rh[1] * n[1] + rh[2] * n[2];
    const double c = 0.5 * dt / (-(8 * kPi) * (1 - m.nu) * r * r);
    const double a = 1 - 2 * m.nu;
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            // K[q][p] with i=q, k=p: ((a d + 3 r_q r_p) rdn - a (r_p n_q - r_q n_p))
            const double v = ((p == q ? a : 0.0) + 3 * rh[q] * rh[p]) * rdn -
                             a * (rh[p] * n[q] - rh[q] * n[p]);
            KT[3 * p + q] = v * c;
        }
}

// Bilinear / serendipity shape functions (src/ShapeFunctionT.cpp:236-366).
template <int ENND>
__device__ __forceinline__ void shape_eval(const double* X, double e1, double e2, double* N,
                                           double* x, double* nrm, double& J)
{
    double dN1[ENND], dN2[ENND];
    if constexpr (ENND == 4) {
        N[0] = 0.25 * (1 - e1) * (1 - e2);
        N[1] = 0.25 * (1 + e1) * (1 - e2);
        N[2] = 0.25 * (1 + e1) * (1 + e2);
        N[3] = 0.25 * (1 - e1) * (1 + e2);
        dN1[0] = -0.25 * (1 - e2);
        dN1[1] = 0.25 * (1 - e2);
        dN1[2] = 0.25 * (1 + e2);
        dN1[3] = -0.25 * (1 + e2);
        dN2[0] = -0.25 * (1 - e1);
        dN2[1] = -0.25 * (1 + e1);
        dN2[2] = 0.25 * (1 + e1);
        dN2[3] = 0.25 * (1 - e1);
    } else {
        N[0] = 0.25 * (1 - e1) * (1 - e2) * (-e1 - e2 - 1);
        N[1] = 0.25 * (1 + e1) * (1 - e2) * (e1 - e2 - 1);
        N[2] = 0.25 * (1 + e1) * (1 + e2) * (e1 + e2 - 1);
        N[3] = 0.25 * (1 - e1) * (1 + e2) * (-e1 + e2 - 1);
        N[4] = 0.5 * (1 - e1 * e1) * (1 - e2);
        N[5] = 0.5 * (1 + e1) * (1 - e2 * e2);
        N[6] = 0.5 * (1 - e1 * e1) * (1 + e2);
        N[7] = 0.5 * (1 - e1) * (1 - e2 * e2);
        dN1[0] = -0.25 * (1 - e1) * (1 - e2) - 0.25 * (1 - e2) * (-1 - e1 - e2);
        dN1[1] = 0.25 * (1 + e1) * (1 - e2) + 0.25 * (1 - e2) * (-1 + e1 - e2);
        dN1[2] = 0.25 * (1 + e1) * (1 + e2) + 0.25 * (1 + e2) * (-1 + e1 + e2);
        dN1[3] = -0.25 * (1 - e1) * (1 + e2) - 0.25 * (1 + e2) * (-1 - e1 + e2);
        dN1[4] = -e1 * (1 - e2);
        dN1[5] = 0.5 * (1 - e2 * e2);
        dN1[6] = -e1 * (1 + e2);
        dN1[7] = -0.5 * (1 - e2 * e2);
        dN2[0] = -0.25 * (1 - e1) * (1 - e2) - 0.25 * (1 - e1) * (-1 - e1 - e2);
        dN2[1] = -0.25 * (1 + e1) * (1 - e2) - 0.25 * (1 + e1) * (-1 + e1 - e2);
        dN2[2] = 0.25 * (1 + e1) * (1 + e2) + 0.25 * (1 + e1) * (-1 + e1 + e2);
        dN2[3] = 0.25 * (1 - e1) * (1 + e2) + 0.25 * (1 - e1) * (-1 - e1 + e2);
        dN2[4] = -0.5 * (1 - e1 * e1);
        dN2[5] = -(1 + e1) * e2;
        dN2[6] = 0.5 * (1 - e1 * e1);
        dN2[7] = -(1 - e1) * e2;
    }
    double d1[3] = {0, 0, 0}, d2[3] = {0, 0, 0};
    x[0] = x[1] = x[2] = 0;
#pragma unroll
    for (int a = 0; a < ENND; a++)
#pragma unroll
        for (int d = 0; d < 3; d++) {
            x[d] += N[a] * X[3 * a + d];
            d1[d] += dN1[a] * X[3 * a + d];
            d2[d] += dN2[a] * X[3 * a + d];
        }
    nrm[0] = d1[1] * d2[2] - d1[2] * d2[1];
    nrm[1] = d1[2] * d2[0] - d1[0] * d2[2];
    nrm[2] = d1[0] * d2[1] - d1[1] * d2[0];
    J = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
    const double iJ = 1.0 / J;
    nrm[0] *= iJ;
    nrm[1] *= iJ;
    nrm[2] *= iJ;
}

// Duffy map of the two triangles around corner `c` of [-1,1]^2 (src/TDElementStepLinearT.cpp:
// 408-427, 445-449, 526-530).  Returns the parent coordinates and the triangle Jacobian.
__device__ __forceinline__ void duffy(int c, int tri, double et1, double et2, double& e1,
                                      double& e2, double& Jt)
{
    const double c0 = (c == 1 || c == 2) ? 1.0 : -1.0;
    const double c1 = (c >= 2) ? 1.0 : -1.0;
    Jt = 0.5 * (1 - et1);
    if (tri == 0) {
        e1 = c0 * (1 - 0.5 * (1 - et1) * (1 + c0 * c1 * et2));
        e2 = c1 * et1;
    } else {
        e1 = c0 * et1;
        e2 = c1 * (1 - 0.5 * (1 - et1) * (1 - c0 * c1 * et2));
    }
}

// Sub-elements used for singular integration of 8-node quads (src/TDElementStepLinearT.cpp:
// 192-386): node ids of up to two bilinear sub-elements, and the singular corner in each.
__device__ __forceinline__ void sub_element(int singular, int se, int* ids, int& nsub, int& cid)
{
    const int tab[8][2][4] = {{{0, 1, 2, 3}, {0, 0, 0, 0}}, {{0, 1, 2, 3}, {0, 0, 0, 0}},
                              {{0, 1, 2, 3}, {0, 0, 0, 0}}, {{0, 1, 2, 3}, {0, 0, 0, 0}},
                              {{4, 6, 3, 0}, {4, 1, 2, 6}}, {{5, 2, 3, 7}, {5, 7, 0, 1}},
                              {{6, 3, 0, 4}, {6, 4, 1, 2}}, {{7, 0, 1, 5}, {7, 5, 2, 3}}};
    nsub = singular < 4 ? 1 : 2;
    cid = singular < 4 ? singular : 0;
#pragma unroll
    for (int k = 0; k < 4; k++) ids[k] = tab[singular][se][k];
}

}  // namespace tdbem
