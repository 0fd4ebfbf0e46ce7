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
    // derived once on the host (make_material), so that the kernels multiply instead of divide
    double inv_cp, inv_cs, inv_cp2, inv_cs2;   // 1/c_p, 1/c_s, 1/c_p^2, 1/c_s^2
    double k2, k3;                             // (c_s/c_p)^2, (c_s/c_p)^3
    double inv_4pi_rho;                        // 1 / (4 pi rho)
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

// From S to the 15 block values of one point, already transposed into the global block layout
// (row = collocation component p, column = node component q):
//   G[p][q] = U[q][p] (symmetric: g00 g01 g02 g11 g12 g22)
//   H[p][q] = T[q][p] (h00 h01 h02 h10 h11 h12 h20 h21 h22)
// multATB(DU, NMatrix) in src/TDElementStepLinearT.cpp:91-101 is what puts U[q][p] at [p][q].
template <typename R>
__device__ __forceinline__ void blocks_from_s(const R* S, const R* rh, const R* n, R r,
                                              R inv_r, R inv_r2, const Material& m, R* out)
{
    const R cs = R(m.cs);
    const R inv_cs2 = R(m.inv_cs2);
    const R k2 = R(m.k2);
    const R k3 = R(m.k3);
    const R A1 = R(0.5) * (S[0] - S[1]);
    const R A2 = S[2] * R(m.inv_cp2) - S[3] * inv_cs2;
    const R A3 = S[3];
    const R B2 = S[3] - k2 * S[2];
    const R B3 = S[5] - k3 * S[4];
    const R B4 = S[2] + r * R(m.inv_cp) * S[4];
    const R B5 = S[3] + r * R(m.inv_cs) * S[5];
    const R rdn = rh[0] * n[0] + rh[1] * n[1] + rh[2] * n[2];
    const R uscale = R(m.inv_4pi_rho) * inv_r;
    const R ua = (R(3) * A1 + A2) * uscale;
    const R ub = (-A1 + A3 * inv_cs2) * uscale;
    const R w = (R(6) * cs * cs * A1 - R(2) * B2) * inv_r2;
    const R t4 = R(1) / R(4 * kPi);
    const R al = (rdn * (-R(30) * cs * cs * A1 + R(12) * B2) * inv_r2 +
                  R(2) * rdn * B3 * inv_r * R(m.inv_cs)) * t4;
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

// ---------------------------------------------------------------------------------------------
// Cubic B-spline temporal kernel (Compute(5), src/GreenFunctionT.cpp:522-585, 625-658, 733-998):
// the Stokes solution for a cubic B-spline load pulse of width bw = 0.7 dt, integrated against
// the linear hats by the 20-point time rule.
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline double heav(double t) { return t >= 0 ? 1.0 : 0.0; }

// B1(t / T) and its time derivative (src/GreenFunctionT.cpp:780-815)
__host__ __device__ inline void cubic_b1(double t, double T, double& B1, double& DB1)
{
    t = t / T;
    const double tp2 = t * t, tp3 = tp2 * t;
    const double h0 = heav(t), h1 = heav(t - 0.25), h2 = heav(t - 0.5), h3 = heav(t - 0.75),
                 h4 = heav(t - 1);
    B1 = (h0 - h1) * (32.0 / 3.0) * tp3;
    B1 += (h1 - h2) * (-32 * tp3 + 32 * tp2 + -8 * t + 2.0 / 3.0);
    B1 += (h2 - h3) * (32 * tp3 + -64 * tp2 + 40 * t + -22.0 / 3.0);
    B1 += (h3 - h4) * (-32.0 / 3.0 * tp3 + 32 * tp2 + -32 * t + 32.0 / 3.0);
    double d = (h0 - h1) * 32 * tp2;
    d += (h1 - h2) * (-96 * tp2 + 64 * t - 8);
    d += (h2 - h3) * (96 * tp2 - 128 * t + 40);
    d += (h3 - h4) * (-32 * tp2 + 64 * t - 32);
    DB1 = d / T;
}

// One spline segment [T1, T2] of the time convolution of the spline with the Stokes time
// function (the four blocks of src/GreenFunctionT.cpp:857-995); t1, t2 = r/c_p, r/c_s.
__device__ __forceinline__ double bspline_segment(double t, double r, double t1, double t2,
                                                  double T1, double T2, double pc, double pq,
                                                  double pl, double p0)
{
    const double tp2 = t * t, tp3 = tp2 * t, tp4 = tp3 * t, tp5 = tp4 * t;
    const double a2 = t1 * t1, a3 = a2 * t1, a4 = a3 * t1, a5 = a4 * t1;
    const double b2 = t2 * t2, b3 = b2 * t2, b4 = b3 * t2, b5 = b4 * t2;
    const double J1c = 4 * a5 - 15 * a4 * t + 20 * a3 * tp2 - 10 * a2 * tp3;
    const double J2c = 4 * b5 - 15 * b4 * t + 20 * b3 * tp2 - 10 * b2 * tp3;
    const double J1q = -3 * a4 + 8 * a3 * t - 6 * a2 * tp2;
    const double J2q = -3 * b4 + 8 * b3 * t - 6 * b2 * tp2;
    const double J1l = 2 * a3 - 3 * a2 * t;
    const double J2l = 2 * b3 - 3 * b2 * t;
    const double J10 = -a2, J20 = -b2;
    const double c2 = T1 * T1, c3 = c2 * T1, c4 = c3 * T1, c5 = c4 * T1;
    const double d2 = T2 * T2, d3 = d2 * T2, d4 = d3 * T2, d5 = d4 * T2;
    const double K1c = tp5 - 5 * t * c4 + 4 * c5;
    const double K2c = tp5 - 5 * t * d4 + 4 * d5;
    const double K1q = tp4 - 4 * t * c3 + 3 * c4;
    const double K2q = tp4 - 4 * t * d3 + 3 * d4;
    const double K1l = (t - T1) * (t - T1) * (t + 2 * T1);
    const double K2l = (t - T2) * (t - T2) * (t + 2 * T2);
    const double K10 = (t - T1) * (t - T1), K20 = (t - T2) * (t - T2);
    const double h11 = heav(t - t1 - T1), h21 = heav(t - t2 - T1);
    const double h12 = heav(t - t1 - T2), h22 = heav(t - t2 - T2);
    const double rr = r * r;
    double Ic = h11 * (J1c + K1c) - h21 * (J2c + K1c);
    Ic -= (h12 * (J1c + K2c) - h22 * (J2c + K2c));
    Ic = Ic / (20.0 * rr);
    double Iq = h11 * (J1q + K1q) - h21 * (J2q + K1q);
    Iq -= (h12 * (J1q + K2q) - h22 * (J2q + K2q));
    Iq = Iq / (12.0 * rr);
    double Il = h11 * (J1l + K1l) - h21 * (J2l + K1l);
    Il -= (h12 * (J1l + K2l) - h22 * (J2l + K2l));
    Il = Il / (6.0 * rr);
    double I0 = h11 * (J10 + K10) - h21 * (J20 + K10);
    I0 -= (h12 * (J10 + K20) - h22 * (J20 + K20));
    I0 = I0 / (2.0 * rr);
    return pc * Ic + pq * Iq + pl * Il + p0 * I0;
}

// src/GreenFunctionT.cpp:817-998: the first segment carries only the cubic term.
__device__ __forceinline__ double cubic_b1_int(double r, double cp, double cs, double t, double T)
{
    const double t1 = r / cp, t2 = r / cs;
    const double Tp2 = T * T, Tp3 = Tp2 * T;
    double B = bspline_segment(t, r, t1, t2, 0, 0.25 * T, 32.0 / (3.0 * Tp3), 0, 0, 0);
    B += bspline_segment(t, r, t1, t2, 0.25 * T, 0.5 * T, -32.0 / Tp3, 32.0 / Tp2, -8.0 / T,
                         2.0 / 3.0);
    B += bspline_segment(t, r, t1, t2, 0.5 * T, 0.75 * T, 32.0 / Tp3, -64.0 / Tp2, 40 / T,
                         -22.0 / 3.0);
    B += bspline_segment(t, r, t1, t2, 0.75 * T, T, -32.0 / (3.0 * Tp3), 32 / Tp2, -32.0 / T,
                         32.0 / 3.0);
    return B;
}

// The Stokes displacement and traction of a pulsed point load, written as the 15 block values of
// one point in the global block layout (G[p][q] = U[q][p], H[p][q] = T[q][p]), from the load's
// value GC and rate DGC at the two retarded times t - r/c_p, t - r/c_s and the integral I1 of the
// load against the near-field term: the common tail of Stokes_CubicBSpline
// (src/GreenFunctionT.cpp:522-585) and Stokes_RampedHeaviside (:449-519).
__device__ __forceinline__ void pulse_blocks(double GC1, double GC2, double DGC1, double DGC2,
                                             double I1, double r, const double* rh,
                                             const double* n, const Material& m, double* out)
{
    const double cp = m.cp, cs = m.cs;
    const double DI2 = GC1 / (cp * cp) - GC2 / (cs * cs);
    const double DI3 = GC2;
    const double k = cs / cp;
    const double TI2 = GC2 - k * k * GC1;
    const double TI3 = DGC2 - k * k * k * DGC1;
    const double TI4 = GC1 + (r / cp) * DGC1;
    const double TI5 = GC2 + (r / cs) * DGC2;
    const double rdn = rh[0] * n[0] + rh[1] * n[1] + rh[2] * n[2];
    const double divider = 4 * kPi * m.rho * r;
    const double RR = r * r;
    // U is symmetric: G = U^T = U
    const int gi[6] = {0, 0, 0, 1, 1, 2}, gk[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int v = 0; v < 6; v++) {
        const int i = gi[v], kk = gk[v];
        const double dik = i == kk ? 1.0 : 0.0, rirk = rh[i] * rh[kk];
        out[v] = ((3 * rirk - dik) * I1 + rirk * DI2 + dik / (cs * cs) * DI3) / divider;
    }
#pragma unroll
    for (int p = 0; p < 3; p++)
#pragma unroll
        for (int q = 0; q < 3; q++) {
            const int i = q, kk = p;  // H[p][q] = T[q][p]
            const double dik = i == kk ? 1.0 : 0.0, rirk = rh[i] * rh[kk];
            const double nirk = n[i] * rh[kk], nkri = n[kk] * rh[i];
            const double s1 = -6 * cs * cs * (5 * rirk * rdn - (nirk + nkri + dik * rdn)) / RR;
            const double s2 = 2 * (6 * rirk * rdn - (nirk + nkri + dik * rdn)) / RR;
            const double s3 = 2 * rirk * rdn / (cs * r);
            const double s4 = -nirk * (1 - 2 * k * k) / RR;
            const double s5 = -(dik * rdn + nkri) / RR;
            out[6 + 3 * p + q] =
                (s1 * I1 + s2 * TI2 + s3 * TI3 + s4 * TI4 + s5 * TI5) / (4.0 * kPi);
        }
}

// Stokes_CubicBSpline (src/GreenFunctionT.cpp:522-585) at time t: spline pulse of width bw.
__device__ __forceinline__ void bspline_blocks(double t, double bw, double r, const double* rh,
                                               const double* n, const Material& m, double* out)
{
    double GC1, GC2, DGC1, DGC2;
    cubic_b1(t - r / m.cp, bw, GC1, DGC1);
    cubic_b1(t - r / m.cs, bw, GC2, DGC2);
    pulse_blocks(GC1, GC2, DGC1, DGC2, cubic_b1_int(r, m.cp, m.cs, t, bw), r, rh, n, m, out);
}

// ---------------------------------------------------------------------------------------------
// Ramped-Heaviside temporal kernel (src/GreenFunctionT.cpp:449-519, 589-622, 661-680, 705-731):
// the load g(t) = min(max(k t, 0), 1) with slope k = 2/dt (SetTime, :111-117).  No Compute case
// of the reference reaches these functions; here they are kernel 6, in the element loop of
// TDElementRampHeaviLinearT with the 40-point time rule.
// ---------------------------------------------------------------------------------------------
__host__ __device__ inline double ramp_load(double t, double k)
{
    // src/GreenFunctionT.cpp:722
    return k * t * (heav(t) - heav(t - 1 / k)) + heav(t - 1 / k);
}

// One wave's part of I1 in Stokes_RampedHeaviside (src/GreenFunctionT.cpp:466-469): the ramp
// k (t - tau) for t - r/c > tau > max(0, ...) and the constant 1 after it, against the
// near-field time function (tau^2 / r^2 - 1 / c^2).
__device__ __forceinline__ double ramp_i1(double t, double k, double r, double c)
{
    const double tr = t / r, ic = 1 / c;
    const double ramp = k * ((tr * tr - ic * ic) * t / 2 - (tr * tr * tr - ic * ic * ic) * r / 3);
    const double tk = t - 1 / k, tkr = tk / r;
    const double tail = (1 - k * t) * (tkr * tkr - ic * ic) / 2 + k * r * (tkr * tkr * tkr - ic * ic * ic) / 3;
    return ramp * heav(t - r / c) + tail * heav(tk - r / c);
}

// Stokes_RampedHeaviside (src/GreenFunctionT.cpp:449-519) at time t.
__device__ __forceinline__ void ramp_blocks(double t, double k, double r, const double* rh,
                                            const double* n, const Material& m, double* out)
{
    const double cp = m.cp, cs = m.cs;
    const double HC1 = heav(t - r / cp), HC2 = heav(t - r / cs);
    const double HKC1 = heav(t - 1 / k - r / cp), HKC2 = heav(t - 1 / k - r / cs);
    const double GC1 = k * (t - r / cp) * (HC1 - HKC1) + HKC1;
    const double GC2 = k * (t - r / cs) * (HC2 - HKC2) + HKC2;
    const double DGC1 = k * (HC1 - HKC1), DGC2 = k * (HC2 - HKC2);
    const double I1 = ramp_i1(t, k, r, cp) - ramp_i1(t, k, r, cs);
    pulse_blocks(GC1, GC2, DGC1, DGC2, I1, r, rh, n, m, out);
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
