/*
 * oracle.c — CPU restatement of TD-BEM's G/H lag assembly and time marching (FP64, plain C).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).  Nothing in td-bem_amd/ links or loads this file.
 *
 * What pins each part (tests/test_oracle_*.py):
 *   - or_shape_eval, or_kelvin_*, or_stokes_*, or_element, or_assemble_step: element blocks and
 *     assembled G/H rows compared against the reference's own classes compiled from source into
 *     oracle/_ref (when /root/reference is present) and against the golden fixtures in
 *     tests/golden/ written from them; plus invariants: partition of unity and planar area,
 *     exact zero before the P-wave arrives, static (Kelvin) limit of Stokes_Heaviside_Linear,
 *     step kernel = difference of two Heaviside kernels, brute-force time quadrature.
 *   - or_build_lags: the UpdateGHLinear bookkeeping, pinned by rebuilding lags from
 *     or_assemble_step outputs in the test.
 *   - or_march_literal: PETSc is not installed, so the reference solver cannot run anywhere
 *     here ("parity unpinned" for the reference's own KSP iterates).  Pinned by: the fused form
 *     (or_march_fused) derived independently; alpha1=alpha2=0 reducing to plain collocation
 *     (the recurrence of Solve_Direct, src/SolverStepLinearT.cpp:185-308); linearity in the load;
 *     a brute-force solve of the full block-lower-triangular space-time system.
 *   - or_lu_*: residual checks against numpy.linalg.solve.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* ------------------------------------------------------------------------------------------ */
/* Material: src/GreenFunctionT.cpp:76-93                                                      */
/* ------------------------------------------------------------------------------------------ */
void or_material_set(or_material* m, double E, double nu, double rho)
{
    double lam;
    m->E = E;
    m->nu = nu;
    m->rho = rho;
    m->mu = E / (2 * (1 + nu));
    lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    m->cp = sqrt((lam + 2 * m->mu) / rho);
    m->cs = sqrt(m->mu / rho);
}

/* ------------------------------------------------------------------------------------------ */
/* Gauss tables.  The shape-function tables follow src/ShapeFunctionT.cpp:83-163, including    */
/* the sign of the 19th 20-point abscissa (:121, +0.993128599 where -0.993128599 is the         */
/* Legendre node); the time tables follow src/GreenFunctionT.cpp:1046-1159 (no such slip).      */
/* ------------------------------------------------------------------------------------------ */
static const double g6_x[6] = {0.6612093864662645, -0.6612093864662645, -0.2386191860831969,
                               0.2386191860831969, -0.9324695142031521, 0.9324695142031521};
static const double g6_w[6] = {0.3607615730481386, 0.3607615730481386, 0.4679139345726910,
                               0.4679139345726910, 0.1713244923791704, 0.1713244923791704};
static const double g10_x[10] = {-0.148874339, 0.148874339, -0.433395394, 0.433395394,
                                 -0.679409568, 0.679409568, -0.865063367, 0.865063367,
                                 -0.973906529, 0.973906529};
static const double g10_w[10] = {0.295524225, 0.295524225, 0.269266719, 0.269266719, 0.219086363,
                                 0.219086363, 0.149451349, 0.149451349, 0.066671344, 0.066671344};
static const double g20_x_time[20] = {
    -0.076526521, 0.076526521, -0.227785851, 0.227785851, -0.373706089, 0.373706089,
    -0.510867002, 0.510867002, -0.636053681, 0.636053681, -0.746331906, 0.746331906,
    -0.839116972, 0.839116972, -0.912234428, 0.912234428, -0.963971927, 0.963971927,
    -0.993128599, 0.993128599};
static const double g20_x_shape[20] = {
    -0.076526521, 0.076526521, -0.227785851, 0.227785851, -0.373706089, 0.373706089,
    -0.510867002, 0.510867002, -0.636053681, 0.636053681, -0.746331906, 0.746331906,
    -0.839116972, 0.839116972, -0.912234428, 0.912234428, -0.963971927, 0.963971927,
    0.993128599, 0.993128599};
static const double g20_w[20] = {0.152753387, 0.152753387, 0.149172986, 0.149172986, 0.142096109,
                                 0.142096109, 0.131688638, 0.131688638, 0.118194532, 0.118194532,
                                 0.10193012, 0.10193012, 0.083276742, 0.083276742, 0.062672048,
                                 0.062672048, 0.04060143, 0.04060143, 0.017614007, 0.017614007};
static const double g40_x[40] = {
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
static const double g40_w[40] = {
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

int or_gauss_shape(int nq, const double** x, const double** w)
{
    switch (nq) {
    case 6: *x = g6_x; *w = g6_w; return 0;
    case 10: *x = g10_x; *w = g10_w; return 0;
    case 20: *x = g20_x_shape; *w = g20_w; return 0;
    case 40: *x = g40_x; *w = g40_w; return 0;
    default: return -1; /* "ShapeFunctionT::SetNQ, Unsupported number of Gaussian points" */
    }
}

int or_gauss_time(int nq, const double** x, const double** w)
{
    switch (nq) {
    case 6: *x = g6_x; *w = g6_w; return 0;
    case 10: *x = g10_x; *w = g10_w; return 0;
    case 20: *x = g20_x_time; *w = g20_w; return 0;
    case 40: *x = g40_x; *w = g40_w; return 0;
    default: return -1;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Shape functions: src/ShapeFunctionT.cpp:236-366                                             */
/* ------------------------------------------------------------------------------------------ */
static void cross3(const double* a, const double* b, double* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

static double norm3(const double* a)
{
    /* MathOperationT::VecNormal (src/MathOperationT.cpp:374-389) sums pow(.,2) then sqrt. */
    return sqrt(pow(a[0], 2) + pow(a[1], 2) + pow(a[2], 2));
}

int or_shape_eval(int etype, const double* X, double e1, double e2, double* N, double* x,
                  double* n, double* J)
{
    double dN[16];
    double d1[3] = {0, 0, 0}, d2[3] = {0, 0, 0};
    int ennd, a, d;
    if (etype == 1) {
        ennd = 4;
        N[0] = 0.25 * (1 - e1) * (1 - e2);
        N[1] = 0.25 * (1 + e1) * (1 - e2);
        N[2] = 0.25 * (1 + e1) * (1 + e2);
        N[3] = 0.25 * (1 - e1) * (1 + e2);
        dN[0] = -0.25 * (1 - e2);
        dN[1] = 0.25 * (1 - e2);
        dN[2] = 0.25 * (1 + e2);
        dN[3] = -0.25 * (1 + e2);
        dN[4] = -0.25 * (1 - e1);
        dN[5] = -0.25 * (1 + e1);
        dN[6] = 0.25 * (1 + e1);
        dN[7] = 0.25 * (1 - e1);
    } else if (etype == 2) {
        ennd = 8;
        N[0] = 0.25 * (1 - e1) * (1 - e2) * (-e1 - e2 - 1);
        N[1] = 0.25 * (1 + e1) * (1 - e2) * (e1 - e2 - 1);
        N[2] = 0.25 * (1 + e1) * (1 + e2) * (e1 + e2 - 1);
        N[3] = 0.25 * (1 - e1) * (1 + e2) * (-e1 + e2 - 1);
        N[4] = 0.5 * (1 - pow(e1, 2)) * (1 - e2);
        N[5] = 0.5 * (1 + e1) * (1 - pow(e2, 2));
        N[6] = 0.5 * (1 - pow(e1, 2)) * (1 + e2);
        N[7] = 0.5 * (1 - e1) * (1 - pow(e2, 2));
        dN[0] = -0.25 * (1 - e1) * (1 - e2) - 0.25 * (1 - e2) * (-1 - e1 - e2);
        dN[1] = 0.25 * (1 + e1) * (1 - e2) + 0.25 * (1 - e2) * (-1 + e1 - e2);
        dN[2] = 0.25 * (1 + e1) * (1 + e2) + 0.25 * (1 + e2) * (-1 + e1 + e2);
        dN[3] = -0.25 * (1 - e1) * (1 + e2) - 0.25 * (1 + e2) * (-1 - e1 + e2);
        dN[4] = -e1 * (1 - e2);
        dN[5] = 0.5 * (1 - pow(e2, 2));
        dN[6] = -e1 * (1 + e2);
        dN[7] = -0.5 * (1 - pow(e2, 2));
        dN[8] = -0.25 * (1 - e1) * (1 - e2) - 0.25 * (1 - e1) * (-1 - e1 - e2);
        dN[9] = -0.25 * (1 + e1) * (1 - e2) - 0.25 * (1 + e1) * (-1 + e1 - e2);
        dN[10] = 0.25 * (1 + e1) * (1 + e2) + 0.25 * (1 + e1) * (-1 + e1 + e2);
        dN[11] = 0.25 * (1 - e1) * (1 + e2) + 0.25 * (1 - e1) * (-1 - e1 + e2);
        dN[12] = -0.5 * (1 - pow(e1, 2));
        dN[13] = -(1 + e1) * e2;
        dN[14] = 0.5 * (1 - pow(e1, 2));
        dN[15] = -(1 - e1) * e2;
    } else {
        return -1; /* "unsupported Element Type" */
    }
    x[0] = x[1] = x[2] = 0;
    for (a = 0; a < ennd; a++)
        for (d = 0; d < 3; d++) x[d] += N[a] * X[a * 3 + d];
    for (a = 0; a < ennd; a++)
        for (d = 0; d < 3; d++) {
            d1[d] += dN[a] * X[a * 3 + d];
            d2[d] += dN[a + ennd] * X[a * 3 + d];
        }
    cross3(d1, d2, n);
    *J = norm3(n);
    for (d = 0; d < 3; d++) n[d] /= *J;
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Fundamental solutions: src/GreenFunctionT.cpp                                               */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    double r, rv[3], n[3];
} or_spatial;

/* GreenFunctionT::SetSpatial (src/GreenFunctionT.cpp:95-109). */
static void set_spatial(or_spatial* s, const double* dir, const double* nrm)
{
    double R;
    int i;
    s->r = norm3(dir);
    for (i = 0; i < 3; i++) s->rv[i] = dir[i] / s->r;
    R = norm3(nrm);
    for (i = 0; i < 3; i++) s->n[i] = nrm[i] / R;
}

static double dot3(const double* a, const double* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

static void kelvin_scaled(const or_material* m, const or_spatial* s, double scale, double* out)
{
    /* src/GreenFunctionT.cpp:147-173 (scale 1) and :205-231 (scale dt/2). */
    double rdn = dot3(s->rv, s->n);
    double divider = -(8 * M_PI) * (1 - m->nu) * s->r * s->r;
    int i, k;
    for (i = 0; i < 3; i++)
        for (k = 0; k < 3; k++) {
            double dik = (i == k) ? 1.0 : 0.0;
            double rirk = s->rv[i] * s->rv[k];
            double rink = s->rv[i] * s->n[k];
            double rkni = s->rv[k] * s->n[i];
            double v = ((1 - 2 * m->nu) * dik + 3 * rirk) * rdn - (1 - 2 * m->nu) * (rkni - rink);
            v /= divider;
            out[i * 3 + k] = v * scale;
        }
}

void or_kelvin_traction(const or_material* m, const double* dir, const double* nrm, double out[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    kelvin_scaled(m, &s, 1.0, out);
}

void or_kelvin_linear(const or_material* m, const double* dir, const double* nrm, double dt,
                      double out[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    kelvin_scaled(m, &s, 0.5 * dt, out);
}

/* Closed-form time integrals of one wave against the two linear hats
 * (src/GreenFunctionT.cpp:251-281 for c_p; :285-314 for c_s). */
static void hat_integrals(double t, double dt, double r, double c, double* I1, double* J1,
                          double* I3, double* J3, double* I5, double* J5)
{
    double b;
    double t_ref = t - r / c;
    if (dt <= t_ref) {
        b = dt;
        *I5 = 0;
        *J5 = 0;
    } else if (0 <= t_ref && dt > t_ref) {
        b = t_ref;
        *I5 = (dt - t_ref) / dt;
        *J5 = t_ref / dt;
    } else {
        b = 0;
        *I5 = 0;
        *J5 = 0;
    }
    *I1 = (-3 * pow(b, 4) + (8 * t + 4 * dt) * pow(b, 3) - (6 * pow(t, 2) + 12 * t * dt) * pow(b, 2) +
           12 * dt * pow(t, 2) * b) / (12 * pow(r, 2)) -
          (dt * b - pow(b, 2) / 2.0) / pow(c, 2);
    *I1 = *I1 / dt;
    *J1 = -(-3 * pow(b, 4) + 8 * t * pow(b, 3) - 6 * pow(t, 2) * pow(b, 2)) / (12 * pow(r, 2)) -
          (pow(b, 2) / 2.0) / pow(c, 2);
    *J1 = *J1 / dt;
    *I3 = (dt * b - pow(b, 2) / 2.0) / dt;
    *J3 = pow(b, 2) / (2.0 * dt);
}

static void heaviside_linear(const or_material* m, const or_spatial* s, double t, double dt,
                             double* DU1, double* DU2, double* DT1, double* DT2)
{
    /* src/GreenFunctionT.cpp:241-416 */
    double I1, I2, I3, I4, I5, I6, J1, J2, J3, J4, J5, J6;
    double It1, It2, It3, It4, It5, Jt1, Jt2, Jt3, Jt4, Jt5;
    double cp = m->cp, cs = m->cs, r = s->r, rdn, divider;
    int i, k;
    hat_integrals(t, dt, r, cp, &I1, &J1, &I3, &J3, &I5, &J5);
    hat_integrals(t, dt, r, cs, &I2, &J2, &I4, &J4, &I6, &J6);

    It1 = 0.5 * (I1 - I2);
    It2 = 1 / pow(cp, 2) * I3 - 1 / pow(cs, 2) * I4;
    It3 = I4;
    Jt1 = 0.5 * (J1 - J2);
    Jt2 = 1.0 / pow(cp, 2) * J3 - 1.0 / pow(cs, 2) * J4;
    Jt3 = J4;
    divider = 4 * M_PI * m->rho * r;
    for (i = 0; i < 3; i++)
        for (k = 0; k < 3; k++) {
            double dik = (i == k) ? 1.0 : 0.0;
            double rirk = s->rv[i] * s->rv[k];
            double s1 = 3 * rirk - dik, s2 = rirk, s3 = dik / pow(cs, 2);
            DU1[i * 3 + k] = (s1 * It1 + s2 * It2 + s3 * It3) / divider;
            DU2[i * 3 + k] = (s1 * Jt1 + s2 * Jt2 + s3 * Jt3) / divider;
        }

    It1 = 0.5 * (I1 - I2);
    It2 = I4 - pow(cs / cp, 2) * I3;
    It3 = I6 - pow(cs / cp, 3) * I5;
    It4 = I3 + r / cp * I5;
    It5 = I4 + r / cs * I6;
    Jt1 = 0.5 * (J1 - J2);
    Jt2 = J4 - pow(cs / cp, 2) * J3;
    Jt3 = J6 - pow(cs / cp, 3) * J5;
    Jt4 = J3 + r / cp * J5;
    Jt5 = J4 + r / cs * J6;
    rdn = dot3(s->rv, s->n);
    for (i = 0; i < 3; i++)
        for (k = 0; k < 3; k++) {
            double dik = (i == k) ? 1.0 : 0.0;
            double RR = pow(r, 2);
            double rirk = s->rv[i] * s->rv[k];
            double nirk = s->n[i] * s->rv[k];
            double nkri = s->n[k] * s->rv[i];
            double s1 = -6 * pow(cs, 2) * (5 * rirk * rdn - (nirk + nkri + dik * rdn)) / RR;
            double s2 = 2 * (6 * rirk * rdn - (nirk + nkri + dik * rdn)) / RR;
            double s3 = 2 * rirk * rdn / (cs * r);
            double s4 = -nirk * (1 - 2 * pow(cs / cp, 2)) / RR;
            double s5 = -(dik * rdn + nkri) / RR;
            DT1[i * 3 + k] = (s1 * It1 + s2 * It2 + s3 * It3 + s4 * It4 + s5 * It5) / (4.0 * M_PI);
            DT2[i * 3 + k] = (s1 * Jt1 + s2 * Jt2 + s3 * Jt3 + s4 * Jt4 + s5 * Jt5) / (4.0 * M_PI);
        }
}

void or_stokes_heaviside_linear(const or_material* m, const double* dir, const double* nrm,
                                double t, double dt, double DU1[9], double DU2[9], double DT1[9],
                                double DT2[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    heaviside_linear(m, &s, t, dt, DU1, DU2, DT1, DT2);
}

static void step_linear(const or_material* m, const or_spatial* s, double t, double dt,
                        double* DU1, double* DU2, double* DT1, double* DT2)
{
    /* src/GreenFunctionT.cpp:424-446: HL(t) - HL(t - dt); t - dt formed as fT - fT_1. */
    double a1[9], a2[9], b1[9], b2[9];
    int i;
    heaviside_linear(m, s, t, dt, DU1, DU2, DT1, DT2);
    heaviside_linear(m, s, t - dt, dt, a1, a2, b1, b2);
    for (i = 0; i < 9; i++) {
        DU1[i] = 1 * DU1[i] + -1 * a1[i];
        DU2[i] = 1 * DU2[i] + -1 * a2[i];
        DT1[i] = 1 * DT1[i] + -1 * b1[i];
        DT2[i] = 1 * DT2[i] + -1 * b2[i];
    }
}

void or_stokes_step_linear(const or_material* m, const double* dir, const double* nrm, double t,
                           double dt, double DU1[9], double DU2[9], double DT1[9], double DT2[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    step_linear(m, &s, t, dt, DU1, DU2, DT1, DT2);
}

/* --- cubic B-spline kernels: src/GreenFunctionT.cpp:522-585, 625-658, 682-701, 733-998 ----- */
static double heav(double t) { return t >= 0 ? 1.0 : 0.0; } /* :771-777 */

static void cubic_b1(double* B1, double* DB1, double t, double T)
{
    /* src/GreenFunctionT.cpp:780-815 */
    double tp2, tp3;
    t = t / T;
    tp2 = pow(t, 2);
    tp3 = pow(t, 3);
    *B1 = (heav(t) - heav(t - 0.25)) * (32.0 / 3.0) * tp3;
    *B1 += (heav(t - 0.25) - heav(t - 0.5)) * (-32 * tp3 + 32 * tp2 + -8 * t + 2.0 / 3.0);
    *B1 += (heav(t - 0.5) - heav(t - 0.75)) * (32 * tp3 + -64 * tp2 + 40 * t + -22.0 / 3.0);
    *B1 += (heav(t - 0.75) - heav(t - 1)) * (-32.0 / 3.0 * tp3 + 32 * tp2 + -32 * t + 32.0 / 3.0);
    *DB1 = (heav(t) - heav(t - 0.25)) * 32 * tp2;
    *DB1 += (heav(t - 0.25) - heav(t - 0.5)) * (-96 * tp2 + 64 * t - 8);
    *DB1 += (heav(t - 0.5) - heav(t - 0.75)) * (96 * tp2 - 128 * t + 40);
    *DB1 += (heav(t - 0.75) - heav(t - 1)) * (-32 * tp2 + 64 * t - 32);
    *DB1 = *DB1 / T;
}

/* One segment [T1,T2] of the spline integral (the four blocks of :857-995 share this form). */
static double bspline_segment(double t, double r, double t1, double t2, double T1, double T2,
                              double pc, double pq, double pl, double p0)
{
    double tp5 = pow(t, 5), tp4 = pow(t, 4), tp3 = pow(t, 3), tp2 = pow(t, 2);
    double t1p5 = pow(t1, 5), t1p4 = pow(t1, 4), t1p3 = pow(t1, 3), t1p2 = pow(t1, 2);
    double t2p5 = pow(t2, 5), t2p4 = pow(t2, 4), t2p3 = pow(t2, 3), t2p2 = pow(t2, 2);
    double J1c = 4 * t1p5 - 15 * t1p4 * t + 20 * t1p3 * tp2 - 10 * t1p2 * tp3;
    double J2c = 4 * t2p5 - 15 * t2p4 * t + 20 * t2p3 * tp2 - 10 * t2p2 * tp3;
    double J1q = -3 * t1p4 + 8 * t1p3 * t - 6 * t1p2 * tp2;
    double J2q = -3 * t2p4 + 8 * t2p3 * t - 6 * t2p2 * tp2;
    double J1l = 2 * t1p3 - 3 * t1p2 * t;
    double J2l = 2 * t2p3 - 3 * t2p2 * t;
    double J10 = -t1p2, J20 = -t2p2;
    double K1c = tp5 - 5 * t * pow(T1, 4) + 4 * pow(T1, 5);
    double K2c = tp5 - 5 * t * pow(T2, 4) + 4 * pow(T2, 5);
    double K1q = tp4 - 4 * t * pow(T1, 3) + 3 * pow(T1, 4);
    double K2q = tp4 - 4 * t * pow(T2, 3) + 3 * pow(T2, 4);
    double K1l = pow(t - T1, 2) * (t + 2 * T1);
    double K2l = pow(t - T2, 2) * (t + 2 * T2);
    double K10 = pow(t - T1, 2), K20 = pow(t - T2, 2);
    double h11 = heav(t - t1 - T1), h21 = heav(t - t2 - T1);
    double h12 = heav(t - t1 - T2), h22 = heav(t - t2 - T2);
    double Ic, Iq, Il, I0;
    Ic = h11 * (J1c + K1c) - h21 * (J2c + K1c);
    Ic -= (h12 * (J1c + K2c) - h22 * (J2c + K2c));
    Ic = Ic / (20.0 * pow(r, 2));
    Iq = h11 * (J1q + K1q) - h21 * (J2q + K1q);
    Iq -= (h12 * (J1q + K2q) - h22 * (J2q + K2q));
    Iq = Iq / (12.0 * pow(r, 2));
    Il = h11 * (J1l + K1l) - h21 * (J2l + K1l);
    Il -= (h12 * (J1l + K2l) - h22 * (J2l + K2l));
    Il = Il / (6.0 * pow(r, 2));
    I0 = h11 * (J10 + K10) - h21 * (J20 + K10);
    I0 -= (h12 * (J10 + K20) - h22 * (J20 + K20));
    I0 = I0 / (2.0 * pow(r, 2));
    return pc * Ic + pq * Iq + pl * Il + p0 * I0;
}

static double cubic_b1_int(double r, double C1, double C2, double t, double T)
{
    /* src/GreenFunctionT.cpp:817-998.  The first segment carries only the cubic term. */
    double t1 = r / C1, t2 = r / C2;
    double B = 0;
    double Tp3 = pow(T, 3), Tp2 = pow(T, 2);
    B += bspline_segment(t, r, t1, t2, 0, 0.25 * T, 32.0 / (3.0 * Tp3), 0, 0, 0);
    B += bspline_segment(t, r, t1, t2, 0.25 * T, 0.5 * T, -32.0 / Tp3, 32.0 / Tp2, -8.0 / T,
                         2.0 / 3.0);
    B += bspline_segment(t, r, t1, t2, 0.5 * T, 0.75 * T, 32.0 / Tp3, -64.0 / Tp2, 40 / T,
                         -22.0 / 3.0);
    B += bspline_segment(t, r, t1, t2, 0.75 * T, T, -32.0 / (3.0 * Tp3), 32 / Tp2, -32.0 / T,
                         32.0 / 3.0);
    return B;
}

static void stokes_cubic_bspline(const or_material* m, const or_spatial* s, double t, double bw,
                                 double* U, double* Tr)
{
    /* src/GreenFunctionT.cpp:522-585 */
    double GC1, GC2, DGC1, DGC2, I1, DI2, DI3, TI2, TI3, TI4, TI5, rdn, divider;
    double cp = m->cp, cs = m->cs, r = s->r;
    int i, k;
    cubic_b1(&GC1, &DGC1, t - r / cp, bw);
    cubic_b1(&GC2, &DGC2, t - r / cs, bw);
    I1 = cubic_b1_int(r, cp, cs, t, bw);
    DI2 = GC1 / pow(cp, 2) - GC2 / pow(cs, 2);
    DI3 = GC2;
    TI2 = GC2 - pow(cs / cp, 2) * GC1;
    TI3 = DGC2 - pow(cs / cp, 3) * DGC1;
    TI4 = GC1 + (r / cp) * DGC1;
    TI5 = GC2 + (r / cs) * DGC2;
    rdn = dot3(s->rv, s->n);
    divider = 4 * M_PI * m->rho * r;
    for (i = 0; i < 3; i++)
        for (k = 0; k < 3; k++) {
            double dik = (i == k) ? 1.0 : 0.0;
            double RR = pow(r, 2);
            double rirk = s->rv[i] * s->rv[k];
            double nirk = s->n[i] * s->rv[k];
            double nkri = s->n[k] * s->rv[i];
            double s1 = -6 * pow(cs, 2) * (5 * rirk * rdn - (nirk + nkri + dik * rdn)) / RR;
            double s2 = 2 * (6 * rirk * rdn - (nirk + nkri + dik * rdn)) / RR;
            double s3 = 2 * rirk * rdn / (cs * r);
            double s4 = -nirk * (1 - 2 * pow(cs / cp, 2)) / RR;
            double s5 = -(dik * rdn + nkri) / RR;
            U[i * 3 + k] = ((3 * rirk - dik) * I1 + rirk * DI2 + dik / pow(cs, 2) * DI3) / divider;
            Tr[i * 3 + k] = (s1 * I1 + s2 * TI2 + s3 * TI3 + s4 * TI4 + s5 * TI5) / (4.0 * M_PI);
        }
}

static void bspline_linear(const or_material* m, const or_spatial* s, double t, double dt,
                           double* DU1, double* DU2, double* DT1, double* DT2)
{
    /* src/GreenFunctionT.cpp:625-658 with the 20-point time rule and width 0.7 dt (:116). */
    const double *gx, *gw;
    double U[9], Tr[9], bw = 0.7 * dt;
    int q, i;
    or_gauss_time(20, &gx, &gw);
    for (i = 0; i < 9; i++) DU1[i] = DU2[i] = DT1[i] = DT2[i] = 0;
    for (q = 0; q < 20; q++) {
        double ab = gx[q], w = gw[q];
        double phi1 = 0.5 * (1 - ab), phi2 = 0.5 * (1 + ab), J = dt * 0.5;
        double tt = phi2 * dt;
        stokes_cubic_bspline(m, s, t - tt, bw, U, Tr);
        for (i = 0; i < 9; i++) {
            DU1[i] = 1 * DU1[i] + w * J * phi1 * U[i];
            DU2[i] = 1 * DU2[i] + w * J * phi2 * U[i];
            DT1[i] = 1 * DT1[i] + w * J * phi1 * Tr[i];
            DT2[i] = 1 * DT2[i] + w * J * phi2 * Tr[i];
        }
    }
}

void or_stokes_bspline_linear(const or_material* m, const double* dir, const double* nrm,
                              double t, double dt, double DU1[9], double DU2[9], double DT1[9],
                              double DT2[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    bspline_linear(m, &s, t, dt, DU1, DU2, DT1, DT2);
}

static void kelvin_bspline(const or_material* m, const or_spatial* s, double t, double dt,
                           double* K1, double* K2)
{
    /* Int_CubicB_Linear (src/GreenFunctionT.cpp:733-759) times KelvinTraction (:682-701). */
    const double *gx, *gw;
    double g1 = 0, g2 = 0, KT[9], bw = 0.7 * dt;
    int q, i;
    or_gauss_time(20, &gx, &gw);
    for (q = 0; q < 20; q++) {
        double ab = gx[q], w = gw[q];
        double phi1 = 0.5 * (1 - ab), phi2 = 0.5 * (1 + ab), J = dt * 0.5;
        double tt = t - phi2 * dt, g, dg;
        cubic_b1(&g, &dg, tt, bw);
        g1 += phi1 * g * w * J;
        g2 += phi2 * g * w * J;
    }
    kelvin_scaled(m, s, 1.0, KT);
    for (i = 0; i < 9; i++) {
        K1[i] = KT[i] * g1;
        K2[i] = KT[i] * g2;
    }
}

void or_kelvin_bspline_linear(const or_material* m, const double* dir, const double* nrm, double t,
                              double dt, double K1[9], double K2[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    kelvin_bspline(m, &s, t, dt, K1, K2);
}

/* --- ramped-Heaviside kernels: src/GreenFunctionT.cpp:449-519, 589-622, 661-680, 705-731 ---- */
/* Load g(t) = min(max(k t, 0), 1) with slope k = 2/dt (SetTime, :111-117). */
static void stokes_ramped_heaviside(const or_material* m, const or_spatial* s, double t,
                                    double k, double* U, double* Tr)
{
    /* src/GreenFunctionT.cpp:449-519 */
    double cp = m->cp, cs = m->cs, r = s->r;
    double HC1 = heav(t - r / cp), HC2 = heav(t - r / cs);
    double HKC1 = heav(t - 1 / k - r / cp), HKC2 = heav(t - 1 / k - r / cs);
    double GC1 = k * (t - r / cp) * (HC1 - HKC1) + HKC1;
    double GC2 = k * (t - r / cs) * (HC2 - HKC2) + HKC2;
    double DGC1 = k * (HC1 - HKC1), DGC2 = k * (HC2 - HKC2);
    double I1, DI2, DI3, TI2, TI3, TI4, TI5, rdn, divider;
    int i, kk;
    I1 = k * ((pow(t / r, 2) - 1 / pow(cp, 2)) * t / 2 - (pow(t / r, 3) - 1 / pow(cp, 3)) * r / 3) * HC1;
    I1 = I1 - k * ((pow(t / r, 2) - 1 / pow(cs, 2)) * t / 2 - (pow(t / r, 3) - 1 / pow(cs, 3)) * r / 3) * HC2;
    I1 = I1 + ((1 - k * t) * (pow(t - 1 / k, 2) / pow(r, 2) - 1 / pow(cp, 2)) / 2 +
               k * r * (pow(t - 1 / k, 3) / pow(r, 3) - 1 / pow(cp, 3)) / 3) * HKC1;
    I1 = I1 - ((1 - k * t) * (pow(t - 1 / k, 2) / pow(r, 2) - 1 / pow(cs, 2)) / 2 +
               k * r * (pow(t - 1 / k, 3) / pow(r, 3) - 1 / pow(cs, 3)) / 3) * HKC2;
    DI2 = GC1 / pow(cp, 2) - GC2 / pow(cs, 2);
    DI3 = GC2;
    TI2 = GC2 - pow(cs / cp, 2) * GC1;
    TI3 = DGC2 - pow(cs / cp, 3) * DGC1;
    TI4 = GC1 + (r / cp) * DGC1;
    TI5 = GC2 + (r / cs) * DGC2;
    rdn = dot3(s->rv, s->n);
    divider = 4 * M_PI * m->rho * r;
    for (i = 0; i < 3; i++)
        for (kk = 0; kk < 3; kk++) {
            double dik = (i == kk) ? 1.0 : 0.0;
            double RR = pow(r, 2);
            double rirk = s->rv[i] * s->rv[kk];
            double nirk = s->n[i] * s->rv[kk];
            double nkri = s->n[kk] * s->rv[i];
            double s1 = -6 * pow(cs, 2) * (5 * rirk * rdn - (nirk + nkri + dik * rdn)) / RR;
            double s2 = 2 * (6 * rirk * rdn - (nirk + nkri + dik * rdn)) / RR;
            double s3 = 2 * rirk * rdn / (cs * r);
            double s4 = -nirk * (1 - 2 * pow(cs / cp, 2)) / RR;
            double s5 = -(dik * rdn + nkri) / RR;
            U[i * 3 + kk] = ((3 * rirk - dik) * I1 + rirk * DI2 + dik / pow(cs, 2) * DI3) / divider;
            Tr[i * 3 + kk] = (s1 * I1 + s2 * TI2 + s3 * TI3 + s4 * TI4 + s5 * TI5) / (4.0 * M_PI);
        }
}

static void ramp_linear(const or_material* m, const or_spatial* s, double t, double dt,
                        double* DU1, double* DU2, double* DT1, double* DT2)
{
    /* src/GreenFunctionT.cpp:589-622: the 40-point time rule over the hats. */
    const double *gx, *gw;
    double U[9], Tr[9];
    int q, i;
    or_gauss_time(40, &gx, &gw);
    for (i = 0; i < 9; i++) DU1[i] = DU2[i] = DT1[i] = DT2[i] = 0;
    for (q = 0; q < 40; q++) {
        double ab = gx[q], w = gw[q];
        double phi1 = 0.5 * (1 - ab), phi2 = 0.5 * (1 + ab), J = dt * 0.5;
        stokes_ramped_heaviside(m, s, t - phi2 * dt, 2 / dt, U, Tr);
        for (i = 0; i < 9; i++) {
            DU1[i] += w * J * phi1 * U[i];
            DU2[i] += w * J * phi2 * U[i];
            DT1[i] += w * J * phi1 * Tr[i];
            DT2[i] += w * J * phi2 * Tr[i];
        }
    }
}

void or_ramp_weights(double t, double dt, double g[2])
{
    /* Int_RampedHeaviside_Linear (src/GreenFunctionT.cpp:705-731). */
    const double *gx, *gw;
    double k = 2 / dt;
    int q;
    or_gauss_time(40, &gx, &gw);
    g[0] = g[1] = 0;
    for (q = 0; q < 40; q++) {
        double ab = gx[q], w = gw[q];
        double phi1 = 0.5 * (1 - ab), phi2 = 0.5 * (1 + ab), J = dt * 0.5;
        double tt = t - phi2 * dt;
        double gg = k * tt * (heav(tt) - heav(tt - 1 / k)) + heav(tt - 1 / k);
        g[0] += phi1 * gg * w * J;
        g[1] += phi2 * gg * w * J;
    }
}

static void kelvin_ramp(const or_material* m, const or_spatial* s, double t, double dt,
                        double* K1, double* K2)
{
    /* Int_Kelvin_RampedHeaviside_Linear (src/GreenFunctionT.cpp:661-680). */
    double g[2], KT[9];
    int i;
    or_ramp_weights(t, dt, g);
    kelvin_scaled(m, s, 1.0, KT);
    for (i = 0; i < 9; i++) {
        K1[i] = KT[i] * g[0];
        K2[i] = KT[i] * g[1];
    }
}

void or_stokes_ramp_point(const or_material* m, const double* dir, const double* nrm, double t,
                          double dt, double U[9], double T[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    stokes_ramped_heaviside(m, &s, t, 2 / dt, U, T);
}

void or_stokes_ramp_linear(const or_material* m, const double* dir, const double* nrm, double t,
                           double dt, double DU1[9], double DU2[9], double DT1[9], double DT2[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    ramp_linear(m, &s, t, dt, DU1, DU2, DT1, DT2);
}

void or_kelvin_ramp_linear(const or_material* m, const double* dir, const double* nrm, double t,
                           double dt, double K1[9], double K2[9])
{
    or_spatial s;
    set_spatial(&s, dir, nrm);
    kelvin_ramp(m, &s, t, dt, K1, K2);
}

/* ------------------------------------------------------------------------------------------ */
/* Element integration                                                                         */
/* ------------------------------------------------------------------------------------------ */
typedef struct {
    int kernel, step_index;
    double t, dt;
    const or_material* m;
} or_time;

/* Accumulate one quadrature point into the element blocks: the multATB + VecPlus pairs of
 * src/TDElementStepLinearT.cpp:91-101 give GE[i][3a+k] += wJ N_a DU[k][i]. */
static void add_point(int ennd, const double* N, double wJ, const double* DU1, const double* DU2,
                      const double* DT1, const double* DT2, double* GE1, double* GE2, double* HE1,
                      double* HE2)
{
    int i, a, k;
    for (i = 0; i < 3; i++)
        for (a = 0; a < ennd; a++)
            for (k = 0; k < 3; k++) {
                int o = i * 3 * ennd + 3 * a + k;
                GE1[o] = 1 * GE1[o] + wJ * (DU1[k * 3 + i] * N[a]);
                GE2[o] = 1 * GE2[o] + wJ * (DU2[k * 3 + i] * N[a]);
                HE1[o] = 1 * HE1[o] + wJ * (DT1[k * 3 + i] * N[a]);
                HE2[o] = 1 * HE2[o] + wJ * (DT2[k * 3 + i] * N[a]);
            }
}

/* Kernel evaluation at one point: Compute(kernel) + the getters. */
static void eval_kernel(const or_time* tm, const or_spatial* s, double* DU1, double* DU2,
                        double* DT1, double* DT2, double* K1, double* K2)
{
    if (tm->kernel == OR_KERNEL_STEP) {
        kelvin_scaled(tm->m, s, 0.5 * tm->dt, K1);
        memcpy(K2, K1, 9 * sizeof(double));
        step_linear(tm->m, s, tm->t, tm->dt, DU1, DU2, DT1, DT2);
    } else if (tm->kernel == OR_KERNEL_HEAVISIDE) {
        kelvin_scaled(tm->m, s, 0.5 * tm->dt, K1);
        memcpy(K2, K1, 9 * sizeof(double));
        heaviside_linear(tm->m, s, tm->t, tm->dt, DU1, DU2, DT1, DT2);
    } else if (tm->kernel == OR_KERNEL_BSPLINE) {
        bspline_linear(tm->m, s, tm->t, tm->dt, DU1, DU2, DT1, DT2);
        kelvin_bspline(tm->m, s, tm->t, tm->dt, K1, K2);
    } else {
        ramp_linear(tm->m, s, tm->t, tm->dt, DU1, DU2, DT1, DT2);
        kelvin_ramp(tm->m, s, tm->t, tm->dt, K1, K2);
    }
}

/* Does this point add its Kelvin term to C (regular elements, and singular elements of the Step
 * kernel) — Step: only at step 1 (src/TDElementStepLinearT.cpp:103,501,577); Heaviside and
 * B-spline: every step (src/TDElementHeavisideLinearT.cpp:102-104,
 * src/TDElementRampHeaviLinearT.cpp:100-103). */
static int kelvin_active(const or_time* tm)
{
    return tm->kernel == OR_KERNEL_STEP ? (tm->step_index == 1) : 1;
}

static void element_regular(const or_time* tm, int etype, int nq, const double* coll,
                            const double* X, double* GE1, double* GE2, double* HE1, double* HE2,
                            double* C1, double* C2)
{
    /* src/TDElementStepLinearT.cpp:28-125 (and the variants' FormGH_Regular). */
    const double *gx, *gw;
    int ennd = etype == 1 ? 4 : 8, q1, q2, i, j;
    or_gauss_shape(nq, &gx, &gw);
    for (q1 = 0; q1 < nq; q1++)
        for (q2 = 0; q2 < nq; q2++) {
            double N[8], x[3], n[3], J, w, dir[3];
            double DU1[9], DU2[9], DT1[9], DT2[9], K1[9], K2[9];
            or_spatial s;
            or_shape_eval(etype, X, gx[q1], gx[q2], N, x, n, &J);
            w = gw[q1] * gw[q2];
            for (i = 0; i < 3; i++) dir[i] = 1 * x[i] + -1 * coll[i];
            set_spatial(&s, dir, n);
            eval_kernel(tm, &s, DU1, DU2, DT1, DT2, K1, K2);
            add_point(ennd, N, w * J, DU1, DU2, DT1, DT2, GE1, GE2, HE1, HE2);
            if (kelvin_active(tm))
                for (i = 0; i < 3; i++)
                    for (j = 0; j < 3; j++) {
                        C1[i * 3 + j] = 1 * C1[i * 3 + j] + -w * J * K1[j * 3 + i];
                        C2[i * 3 + j] = 1 * C2[i * 3 + j] + -w * J * K2[j * 3 + i];
                    }
        }
}

/* FormGH_Singular_1 for a bilinear (sub-)element whose local node `singular` is the collocation
 * point: two Duffy triangles (src/TDElementStepLinearT.cpp:389-592).  Step kernel: Kelvin goes
 * to C at step 1; Heaviside/B-spline: subtracted from the singular node's columns every step
 * (src/TDElementHeavisideLinearT.cpp:489-497, src/TDElementRampHeaviLinearT.cpp:499-506). */
static int element_singular_bilinear(const or_time* tm, int nq, const double* coll,
                                     const double* X, int singular, double* GE1, double* GE2,
                                     double* HE1, double* HE2, double* C1, double* C2)
{
    const double *gx, *gw;
    double ec0, ec1;
    int tri, q1, q2, i, j;
    const int ennd = 4;
    or_gauss_shape(nq, &gx, &gw);
    switch (singular) {
    case 0: ec0 = -1; ec1 = -1; break;
    case 1: ec0 = 1; ec1 = -1; break;
    case 2: ec0 = 1; ec1 = 1; break;
    case 3: ec0 = -1; ec1 = 1; break;
    default: return -1; /* "collocation id is out of range" */
    }
    for (i = 0; i < 36; i++) GE1[i] = GE2[i] = HE1[i] = HE2[i] = 0;
    for (i = 0; i < 9; i++) C1[i] = C2[i] = 0;
    for (tri = 0; tri < 2; tri++)
        for (q1 = 0; q1 < nq; q1++)
            for (q2 = 0; q2 < nq; q2++) {
                double et1 = gx[q1], et2 = gx[q2], w = gw[q1] * gw[q2];
                double Jt = 0.5 * (1 - et1), e1, e2;
                double N[4], x[3], n[3], J, dir[3];
                double DU1[9], DU2[9], DT1[9], DT2[9], K1[9], K2[9];
                or_spatial s;
                if (tri == 0) {
                    e1 = ec0 * (1 - 0.5 * (1 - et1) * (1 + ec0 * ec1 * et2));
                    e2 = ec1 * et1;
                } else {
                    e1 = ec0 * et1;
                    e2 = ec1 * (1 - 0.5 * (1 - et1) * (1 - ec0 * ec1 * et2));
                }
                or_shape_eval(1, X, e1, e2, N, x, n, &J);
                for (i = 0; i < 3; i++) dir[i] = x[i] - coll[i];
                set_spatial(&s, dir, n);
                eval_kernel(tm, &s, DU1, DU2, DT1, DT2, K1, K2);
                add_point(ennd, N, w * Jt * J, DU1, DU2, DT1, DT2, GE1, GE2, HE1, HE2);
                if (tm->kernel == OR_KERNEL_STEP) {
                    if (tm->step_index == 1)
                        for (i = 0; i < 3; i++)
                            for (j = 0; j < 3; j++) {
                                C1[i * 3 + j] = 1 * C1[i * 3 + j] + -w * J * Jt * K1[j * 3 + i];
                                C2[i * 3 + j] = 1 * C2[i * 3 + j] + -w * J * Jt * K2[j * 3 + i];
                            }
                } else {
                    for (i = 0; i < 3; i++)
                        for (j = 0; j < 3; j++) {
                            HE1[i * 3 * ennd + 3 * singular + j] -= w * Jt * J * K1[j * 3 + i];
                            HE2[i * 3 * ennd + 3 * singular + j] -= w * Jt * J * K2[j * 3 + i];
                        }
                }
            }
    return 0;
}

/* SubElement (src/TDElementStepLinearT.cpp:192-386): 8-node quads integrated singularly as one
 * or two bilinear sub-elements; ids map sub-element nodes to parent nodes. */
static const int sub_ids[8][2][4] = {
    {{0, 1, 2, 3}, {0, 0, 0, 0}}, {{0, 1, 2, 3}, {0, 0, 0, 0}}, {{0, 1, 2, 3}, {0, 0, 0, 0}},
    {{0, 1, 2, 3}, {0, 0, 0, 0}}, {{4, 6, 3, 0}, {4, 1, 2, 6}}, {{5, 2, 3, 7}, {5, 7, 0, 1}},
    {{6, 3, 0, 4}, {6, 4, 1, 2}}, {{7, 0, 1, 5}, {7, 5, 2, 3}}};

int or_element(int kernel, int etype, int nq, const or_material* m, const double* coll,
               const double* X, double t, double dt, int step_index, int singular, double* GE1,
               double* GE2, double* HE1, double* HE2, double* C1, double* C2)
{
    or_time tm;
    const double *gx, *gw;
    int ennd, i;
    if (etype != 1 && etype != 2) return -1;
    if (kernel != OR_KERNEL_STEP && kernel != OR_KERNEL_HEAVISIDE && kernel != OR_KERNEL_BSPLINE &&
        kernel != OR_KERNEL_RAMP)
        return -1;
    if (or_gauss_shape(nq, &gx, &gw)) return -1;
    ennd = etype == 1 ? 4 : 8;
    tm.kernel = kernel;
    tm.step_index = step_index;
    tm.t = t;
    tm.dt = dt;
    tm.m = m;
    for (i = 0; i < 3 * 3 * ennd; i++) GE1[i] = GE2[i] = HE1[i] = HE2[i] = 0;
    for (i = 0; i < 9; i++) C1[i] = C2[i] = 0;
    if (singular < 0) {
        element_regular(&tm, etype, nq, coll, X, GE1, GE2, HE1, HE2, C1, C2);
        return 0;
    }
    if (singular >= ennd) return -1;
    if (etype == 1)
        return element_singular_bilinear(&tm, nq, coll, X, singular, GE1, GE2, HE1, HE2, C1, C2);
    {
        /* EType 2: AssembleSubElement (src/TDElementStepLinearT.cpp:597-626). */
        int nsub = singular < 4 ? 1 : 2, se, ni, ri, ci;
        int cid = singular < 4 ? singular : 0;
        for (se = 0; se < nsub; se++) {
            double Xs[12], g1[36], g2[36], h1[36], h2[36], c1[9], c2[9];
            const int* ids = sub_ids[singular][se];
            for (ni = 0; ni < 4; ni++)
                for (i = 0; i < 3; i++) Xs[ni * 3 + i] = X[ids[ni] * 3 + i];
            if (element_singular_bilinear(&tm, nq, coll, Xs, cid, g1, g2, h1, h2, c1, c2))
                return -1;
            for (ni = 0; ni < 4; ni++)
                for (ri = 0; ri < 3; ri++)
                    for (ci = 0; ci < 3; ci++) {
                        int dst = ri * 24 + 3 * ids[ni] + ci, src = ri * 12 + 3 * ni + ci;
                        GE1[dst] += g1[src];
                        GE2[dst] += g2[src];
                        HE1[dst] += h1[src];
                        HE2[dst] += h2[src];
                    }
            if (kernel == OR_KERNEL_STEP)
                for (i = 0; i < 9; i++) {
                    C1[i] += c1[i];
                    C2[i] += c2[i];
                }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* Assembly driver: src/ElastoDynamicsT.cpp:108-280                                             */
/* ------------------------------------------------------------------------------------------ */
int or_assemble_step(int kernel, int etype, int nq, const or_material* m, int nnd,
                     const double* coords, int nel, const int* ien, int if_ext, int lower,
                     int upper, double t, double dt, int step_index, double* G1, double* G2,
                     double* H1, double* H2, int nthreads)
{
    int ennd = etype == 1 ? 4 : 8;
    long ncol = 3L * nnd;
    long nrow = 3L * (upper - lower + 1);
    int err = 0, ni;
    const double *gx, *gw;
    if ((etype != 1 && etype != 2) || or_gauss_shape(nq, &gx, &gw)) return -1;
    if (lower < 0 || upper >= nnd || upper < lower) return -1;
    memset(G1, 0, sizeof(double) * nrow * ncol);
    memset(G2, 0, sizeof(double) * nrow * ncol);
    memset(H1, 0, sizeof(double) * nrow * ncol);
    memset(H2, 0, sizeof(double) * nrow * ncol);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1) reduction(| : err)
#endif
    for (ni = lower; ni <= upper; ni++) {
        double coll[3], X[24];
        double GE1[72], GE2[72], HE1[72], HE2[72], C1[9], C2[9];
        long loc = ni - lower;
        int e, a, i, j, d;
        (void)nthreads;
        for (d = 0; d < 3; d++) coll[d] = coords[ni * 3 + d];
        for (e = 0; e < nel; e++) {
            int singular = -1;
            for (a = 0; a < ennd; a++)
                for (d = 0; d < 3; d++) X[a * 3 + d] = coords[ien[e * ennd + a] * 3 + d];
            for (a = 0; a < ennd; a++) {
                double r2 = pow(X[a * 3] - coll[0], 2) + pow(X[a * 3 + 1] - coll[1], 2) +
                            pow(X[a * 3 + 2] - coll[2], 2);
                if (r2 < 1e-9) singular = a;
            }
            if (or_element(kernel, etype, nq, m, coll, X, t, dt, step_index, singular, GE1, GE2,
                           HE1, HE2, C1, C2)) {
                err |= 1;
                continue;
            }
            /* AssembleGH (:235-259) */
            for (a = 0; a < ennd; a++) {
                long nid = ien[e * ennd + a];
                for (i = 0; i < 3; i++)
                    for (j = 0; j < 3; j++) {
                        long g = (i + loc * 3) * ncol + nid * 3 + j;
                        int o = i * ennd * 3 + a * 3 + j;
                        G1[g] += GE1[o];
                        G2[g] += GE2[o];
                        H1[g] += HE1[o];
                        H2[g] += HE2[o];
                    }
            }
            /* AssembleC (:261-280) */
            for (i = 0; i < 3; i++)
                for (j = 0; j < 3; j++) {
                    long g = (loc * 3 + i) * ncol + (long)ni * 3 + j;
                    H1[g] += C1[i * 3 + j];
                    H2[g] += C2[i * 3 + j];
                }
        }
        if (if_ext && step_index == 1) /* :212-225 */
            for (d = 0; d < 3; d++) {
                long g = (loc * 3 + d) * ncol + (long)ni * 3 + d;
                H1[g] += 0.5 * dt;
                H2[g] += 0.5 * dt;
            }
    }
    return err ? -1 : 0;
}

int or_max_step(const or_material* m, int nnd, const double* coords, double dt)
{
    /* src/SolverT.cpp:211-250: Cs from E, nu, rho directly (same value as WaveSpeed). */
    double Mu = m->E / (2.0 * (1.0 + m->nu));
    double Cs = sqrt(Mu / m->rho);
    double max_dis = 0;
    int i, j, d;
    for (i = 0; i < nnd - 1; i++)
        for (j = i + 1; j < nnd; j++) {
            double tmp = 0;
            for (d = 0; d < 3; d++) tmp += pow(coords[i * 3 + d] - coords[j * 3 + d], 2);
            tmp = sqrt(tmp);
            if (tmp > max_dis) max_dis = tmp;
        }
    return (int)ceil(max_dis / (Cs * dt)) + 1;
}

int or_build_lags(int kernel, int etype, int nq, const or_material* m, int nnd,
                  const double* coords, int nel, const int* ien, int if_ext, double dt,
                  int numstep, int K, double** MG, double** MH, int nthreads)
{
    long N = 3L * nnd, NN = N * N, i;
    int n, k, last = numstep < K ? numstep : K;
    double *g0 = malloc(sizeof(double) * NN), *g1 = malloc(sizeof(double) * NN);
    double *h0 = malloc(sizeof(double) * NN), *h1 = malloc(sizeof(double) * NN);
    int err = 0;
    if (!g0 || !g1 || !h0 || !h1) err = -1;
    for (k = 0; k <= K && !err; k++) {
        memset(MG[k], 0, sizeof(double) * NN);
        memset(MH[k], 0, sizeof(double) * NN);
    }
    for (n = 1; n <= last && !err; n++) {
        /* main.cpp:89-105: t = n dt; G_0 := fG_1 (phi_1), G_1 := fG_2 (phi_2). */
        if (or_assemble_step(kernel, etype, nq, m, nnd, coords, nel, ien, if_ext, 0, nnd - 1,
                             n * dt, dt, n, g0, g1, h0, h1, nthreads)) {
            err = -1;
            break;
        }
        /* SolverStepLinearT::UpdateGHLinear (src/SolverStepLinearT.cpp:65-97). */
        for (i = 0; i < NN; i++) {
            if (n == 1) {
                MG[0][i] = g1[i];
                MH[0][i] = h1[i];
            } else {
                MG[n - 1][i] += g1[i];
                MH[n - 1][i] += h1[i];
            }
            MG[n][i] = g0[i];
            MH[n][i] = h0[i];
        }
    }
    free(g0);
    free(g1);
    free(h0);
    free(h1);
    return err;
}

/* ------------------------------------------------------------------------------------------ */
/* Dense LU (stands in for KSPSolve; the reference's np=1 GMRES+ILU on a dense matrix is a     */
/* direct solve to rounding).                                                                  */
/* ------------------------------------------------------------------------------------------ */
int or_lu_factor(int n, double* A, int* piv)
{
    int k, i, j;
    for (k = 0; k < n; k++) {
        int p = k;
        double amax = fabs(A[(long)k * n + k]);
        for (i = k + 1; i < n; i++)
            if (fabs(A[(long)i * n + k]) > amax) {
                amax = fabs(A[(long)i * n + k]);
                p = i;
            }
        piv[k] = p;
        if (amax == 0) return -2;
        if (p != k)
            for (j = 0; j < n; j++) {
                double tmp = A[(long)k * n + j];
                A[(long)k * n + j] = A[(long)p * n + j];
                A[(long)p * n + j] = tmp;
            }
        /* rows are independent: each is updated by one thread in the same order as serially */
#pragma omp parallel for private(j) schedule(static) if (n - k > 256)
        for (i = k + 1; i < n; i++) {
            double l = A[(long)i * n + k] / A[(long)k * n + k];
            double* ri = A + (long)i * n;
            const double* rk = A + (long)k * n;
            A[(long)i * n + k] = l;
            for (j = k + 1; j < n; j++) ri[j] -= l * rk[j];
        }
    }
    return 0;
}

void or_lu_solve(int n, const double* LU, const int* piv, double* b)
{
    int i, j;
    for (i = 0; i < n; i++)
        if (piv[i] != i) {
            double tmp = b[i];
            b[i] = b[piv[i]];
            b[piv[i]] = tmp;
        }
    for (i = 0; i < n; i++) {
        double s = b[i];
        for (j = 0; j < i; j++) s -= LU[(long)i * n + j] * b[j];
        b[i] = s;
    }
    for (i = n - 1; i >= 0; i--) {
        double s = b[i];
        for (j = i + 1; j < n; j++) s -= LU[(long)i * n + j] * b[j];
        b[i] = s / LU[(long)i * n + i];
    }
}

/* ------------------------------------------------------------------------------------------ */
/* Time marching                                                                               */
/* ------------------------------------------------------------------------------------------ */
double or_load_factor(const or_march_cfg* c, double t)
{
    if (c->t_rise > 0) return c->amp * (t < c->t_rise ? t / c->t_rise : 1.0);
    return c->amp;
}

/* y += alpha * A x  (dense N x N) */
static void gemv_add(int N, double alpha, const double* A, const double* x, double* y)
{
    int i, j;
    /* one row per thread, each summed in the serial order */
#pragma omp parallel for private(j) schedule(static) if (N > 256)
    for (i = 0; i < N; i++) {
        double s = 0;
        const double* row = A + (long)i * N;
        for (j = 0; j < N; j++) s += row[j] * x[j];
        y[i] += alpha * s;
    }
}

static int alpha_ok(double a1, double a2)
{
    /* src/SolverStepLinearT.cpp:317-322 */
    return a1 >= 0 && a1 <= 1 && a2 >= 0 && a2 < 1 && a1 + a2 <= 1;
}

/* SetLoad (src/SolverT.cpp:344-377): FBC * f at FBC dofs, then UBC values (INSERT order). */
static void set_load(const or_march_cfg* c, double f, double* y)
{
    int i;
    memset(y, 0, sizeof(double) * c->N);
    for (i = 0; i < c->n_fbc; i++) y[c->fbc_dof[i]] = c->fbc_val[i] * f;
    for (i = 0; i < c->n_ubc; i++) y[c->ubc_dof[i]] = c->ubc_val[i];
}

/* ExchangeColumns (src/SolverT.cpp:265-342) on dense G, H. */
static void exchange_columns(const or_march_cfg* c, double* G, double* H)
{
    int N = c->N, u, i;
    for (u = 0; u < c->n_ubc; u++) {
        int col = c->ubc_dof[u];
        for (i = 0; i < N; i++) {
            double g = G[(long)i * N + col], h = H[(long)i * N + col];
            G[(long)i * N + col] = -h;
            H[(long)i * N + col] = -g;
        }
    }
}

/* Post-processing of src/SolverStepLinearT.cpp:512-536.  fDofHigh is 3*nnd-1 on one rank and
 * the test u_dof < fDofHigh skips the last dof (:527); that is kept. */
static void post_process(const or_march_cfg* c, double f, const double* x, double* u, double* t)
{
    int i, N = c->N;
    memcpy(u, x, sizeof(double) * N);
    for (i = 0; i < c->n_ubc; i++) u[c->ubc_dof[i]] = c->ubc_val[i];
    memset(t, 0, sizeof(double) * N);
    for (i = 0; i < c->n_fbc; i++) t[c->fbc_dof[i]] = c->fbc_val[i] * f;
    for (i = 0; i < c->n_ubc; i++) {
        int d = c->ubc_dof[i];
        if (d >= 0 && d < N - 1) t[d] = x[d];
    }
}

int or_march_literal(const or_march_cfg* c, double* const* MG, double* const* MH, double* U,
                     double* T)
{
    int N = c->N, K = c->K, n, s, fi, rc = 0;
    long NN = (long)N * N, i;
    double a1 = c->a1, a2 = c->a2, a3;
    double *Gc, *Hc, *Mt, *y, *rhs;
    int* piv;
    if (!alpha_ok(a1, a2)) return -1;
    a3 = 1 - a1 - a2;
    Gc = malloc(sizeof(double) * NN);
    Hc = malloc(sizeof(double) * NN);
    Mt = malloc(sizeof(double) * NN);
    y = malloc(sizeof(double) * N);
    rhs = malloc(sizeof(double) * N);
    piv = malloc(sizeof(int) * N);
    memset(U, 0, sizeof(double) * N);
    memset(T, 0, sizeof(double) * N);
    /* x_m for m >= numstep are never formed; histories have numstep rows. */
    for (n = 2; n <= c->numstep && rc == 0; n++) {
        double f;
        s = n - 1; /* solve_step */
        f = or_load_factor(c, s * c->dt);
        set_load(c, f, y);
        if (n == 2) {
            /* :355-378 */
            memcpy(Gc, MG[0], sizeof(double) * NN);
            memcpy(Hc, MH[0], sizeof(double) * NN);
            exchange_columns(c, Gc, Hc);
            memset(rhs, 0, sizeof(double) * N);
            gemv_add(N, 1.0, Gc, y, rhs);
            /* bgn[1] * t_0 with t_0 = 0 (:377) adds nothing. */
        } else {
            /* LHS (:384-393) */
            for (i = 0; i < NN; i++) {
                Gc[i] = 0;
                Hc[i] = 0;
                Gc[i] += (2 * a1 + a3) * MG[0][i];
                Gc[i] += a1 * MG[1][i];
                Hc[i] += (2 * a1 + a3) * MH[0][i];
                Hc[i] += a1 * MH[1][i];
            }
            exchange_columns(c, Gc, Hc);
            memset(rhs, 0, sizeof(double) * N);
            gemv_add(N, 1.0, Gc, y, rhs); /* :400 */
            /* Eqn M+1, lag 1 (:408-418) */
            for (i = 0; i < NN; i++) Mt[i] = -(a1 * MH[2][i] - a1 * MH[0][i]);
            gemv_add(N, 1.0, Mt, U + (long)(s - 1) * N, rhs);
            for (i = 0; i < NN; i++) Mt[i] = a1 * MG[2][i] - a1 * MG[0][i];
            gemv_add(N, 1.0, Mt, T + (long)(s - 1) * N, rhs);
            /* Eqn M+1, lags >= 2 (:420-430) */
            for (fi = 3; fi <= (s + 1 < K ? s + 1 : K); fi++) {
                gemv_add(N, -a1, MH[fi], U + (long)(s + 1 - fi) * N, rhs);
                gemv_add(N, a1, MG[fi], T + (long)(s + 1 - fi) * N, rhs);
            }
            /* Eqn M (:433-443) */
            for (fi = 1; fi <= (s - 1 < K ? s - 1 : K); fi++) {
                gemv_add(N, -a3, MH[fi], U + (long)(s - fi) * N, rhs);
                gemv_add(N, a3, MG[fi], T + (long)(s - fi) * N, rhs);
            }
            /* bgn terms (:446-455, :473-482) multiply u_0 = t_0 = 0. */
            /* Eqn M-1 (:460-470) */
            for (fi = 0; fi <= (s - 2 < K ? s - 2 : K); fi++) {
                gemv_add(N, -a2, MH[fi], U + (long)(s - 1 - fi) * N, rhs);
                gemv_add(N, a2, MG[fi], T + (long)(s - 1 - fi) * N, rhs);
            }
        }
        /* KSPSolve (:503-504) */
        memcpy(Mt, Hc, sizeof(double) * NN);
        if (or_lu_factor(N, Mt, piv)) {
            rc = -2;
            break;
        }
        or_lu_solve(N, Mt, piv, rhs);
        post_process(c, f, rhs, U + (long)s * N, T + (long)s * N);
    }
    free(Gc);
    free(Hc);
    free(Mt);
    free(y);
    free(rhs);
    free(piv);
    return rc;
}

int or_march_fused(const or_march_cfg* c, double* const* MG, double* const* MH, double* U,
                   double* T)
{
    /* hist_s = sum_{k=0}^{min(s,K)} (MG_k z^t_k - MH_k z^u_k) with
     *   z_0 = (a2 - a1) x_{s-1},  z_1 = a3 x_{s-1} + a2 x_{s-2},
     *   z_k = a1 x_{s+1-k} + a3 x_{s-k} + a2 x_{s-1-k}  (k >= 2),  x_m = 0 for m <= 0.
     * The LHS for n >= 3 is the same matrix every step, so it is factored once. */
    int N = c->N, K = c->K, n, s, k, j, rc = 0;
    long NN = (long)N * N, i;
    double a1 = c->a1, a2 = c->a2, a3;
    double *Gc = NULL, *Hc = NULL, *y, *rhs, *zu, *zt;
    double *LU2, *LU3;
    int *piv2, *piv3;
    if (!alpha_ok(a1, a2)) return -1;
    a3 = 1 - a1 - a2;
    Gc = malloc(sizeof(double) * NN);
    Hc = malloc(sizeof(double) * NN);
    LU2 = malloc(sizeof(double) * NN);
    LU3 = malloc(sizeof(double) * NN);
    piv2 = malloc(sizeof(int) * N);
    piv3 = malloc(sizeof(int) * N);
    y = malloc(sizeof(double) * N);
    rhs = malloc(sizeof(double) * N);
    zu = malloc(sizeof(double) * N);
    zt = malloc(sizeof(double) * N);
    memset(U, 0, sizeof(double) * N);
    memset(T, 0, sizeof(double) * N);
    for (n = 2; n <= c->numstep && rc == 0; n++) {
        double f;
        s = n - 1;
        f = or_load_factor(c, s * c->dt);
        set_load(c, f, y);
        memset(rhs, 0, sizeof(double) * N);
        if (n == 2) {
            memcpy(Gc, MG[0], sizeof(double) * NN);
            memcpy(LU2, MH[0], sizeof(double) * NN);
            exchange_columns(c, Gc, LU2);
            if (or_lu_factor(N, LU2, piv2)) {
                rc = -2;
                break;
            }
            gemv_add(N, 1.0, Gc, y, rhs);
            or_lu_solve(N, LU2, piv2, rhs);
        } else {
            if (n == 3) {
                for (i = 0; i < NN; i++) {
                    Gc[i] = (2 * a1 + a3) * MG[0][i] + a1 * MG[1][i];
                    LU3[i] = (2 * a1 + a3) * MH[0][i] + a1 * MH[1][i];
                }
                exchange_columns(c, Gc, LU3);
                if (or_lu_factor(N, LU3, piv3)) {
                    rc = -2;
                    break;
                }
            }
            gemv_add(N, 1.0, Gc, y, rhs);
            for (k = 0; k <= (s < K ? s : K); k++) {
                for (j = 0; j < N; j++) {
                    double xu1 = s - 1 >= 1 ? U[(long)(s - 1) * N + j] : 0;
                    double xt1 = s - 1 >= 1 ? T[(long)(s - 1) * N + j] : 0;
                    if (k == 0) {
                        zu[j] = (a2 - a1) * xu1;
                        zt[j] = (a2 - a1) * xt1;
                    } else if (k == 1) {
                        double xu2 = s - 2 >= 1 ? U[(long)(s - 2) * N + j] : 0;
                        double xt2 = s - 2 >= 1 ? T[(long)(s - 2) * N + j] : 0;
                        zu[j] = a3 * xu1 + a2 * xu2;
                        zt[j] = a3 * xt1 + a2 * xt2;
                    } else {
                        int m = s - k;
                        double su = 0, st = 0;
                        if (m + 1 >= 1) { su += a1 * U[(long)(m + 1) * N + j]; st += a1 * T[(long)(m + 1) * N + j]; }
                        if (m >= 1) { su += a3 * U[(long)m * N + j]; st += a3 * T[(long)m * N + j]; }
                        if (m - 1 >= 1) { su += a2 * U[(long)(m - 1) * N + j]; st += a2 * T[(long)(m - 1) * N + j]; }
                        zu[j] = su;
                        zt[j] = st;
                    }
                }
                gemv_add(N, 1.0, MG[k], zt, rhs);
                gemv_add(N, -1.0, MH[k], zu, rhs);
            }
            or_lu_solve(N, LU3, piv3, rhs);
        }
        post_process(c, f, rhs, U + (long)s * N, T + (long)s * N);
    }
    free(Gc);
    free(Hc);
    free(LU2);
    free(LU3);
    free(piv2);
    free(piv3);
    free(y);
    free(rhs);
    free(zu);
    free(zt);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* Two-region solver SolverMRT (src/SolverMRT.cpp), restated literally: the global matrices   */
/* assembled from the regions' lag matrices through the interior / interface index maps       */
/* (InterfaceIndexing :223-568, Assemble_GH_R1/R2 :685-850), the load of SetLoad (:935-1026)  */
/* and the marching of Solve_Direct (:1035-1099) or Solve_Averaging (:1101-1249).             */
/* ------------------------------------------------------------------------------------------ */
double or_mrt_load_factor(const or_mrt_cfg* c, double t)
{
    /* src/SolverMRT.cpp:937-939 (p0 = 5.46366e9, decay 2000); kind 1: amp * min(1, t/t_rise) */
    if (c->load_kind == 0) return 5.46366e9 * t * exp(-2000 * t);
    if (c->t_rise > 0) return c->amp * (t < c->t_rise ? t / c->t_rise : 1.0);
    return c->amp;
}

/* Global index maps.  Unknowns (columns of H): interior dofs of region 1, of region 2, interface
 * displacements, interface tractions (both in region 1's interface order); columns of G: the
 * interior dofs of region 1 then region 2.  Rows: region 1's dofs then region 2's.
 * hcol[r][d]: H column of region r's dof d; gcol[r][d]: G column (interior dofs) or -1;
 * tcol[r][d]: traction column of an interface dof or -1.  Returns -1 on an interface mismatch. */
static int mrt_maps(const or_mrt_cfg* c, int* hcol[2], int* gcol[2], int* tcol[2])
{
    const int nz = c->n_iface, NZ = 3 * nz;
    const int nint1 = 3 * (c->nnd[0] - nz), nint2 = 3 * (c->nnd[1] - nz);
    int r, i, d;
    int* slot2 = malloc(sizeof(int) * (nz > 0 ? nz : 1));
    /* region 2's interface nodes matched to region 1's by coordinates (:440-477) */
    for (i = 0; i < nz; i++) {
        const double* x2 = c->coords[1] + 3 * c->iface[1][i];
        int match = -1, j;
        for (j = 0; j < nz; j++) {
            const double* x1 = c->coords[0] + 3 * c->iface[0][j];
            double dx = x1[0] - x2[0], dy = x1[1] - x2[1], dz = x1[2] - x2[2];
            if (sqrt(dx * dx + dy * dy + dz * dz) < 1e-10) {
                match = j;
                break;
            }
        }
        if (match < 0) {
            free(slot2);
            return -1; /* "Interface mismatch, interface nodes must match one-to-one" */
        }
        slot2[i] = match;
    }
    for (r = 0; r < 2; r++) {
        const int nnd = c->nnd[r];
        int* lab = calloc(nnd, sizeof(int));
        int k = 0;
        for (i = 0; i < nz; i++) lab[c->iface[r][i]] = 1 + i;
        for (i = 0; i < nnd; i++)
            for (d = 0; d < 3; d++) {
                const int dof = 3 * i + d;
                if (!lab[i]) {
                    /* interior nodes in ascending order (:235-246, 359-370) */
                    hcol[r][dof] = (r == 0 ? 0 : nint1) + k;
                    gcol[r][dof] = (r == 0 ? 0 : nint1) + k;
                    tcol[r][dof] = -1;
                    k++;
                } else {
                    const int zi = lab[i] - 1, slot = r == 0 ? zi : slot2[zi];
                    hcol[r][dof] = nint1 + nint2 + 3 * slot + d;
                    gcol[r][dof] = -1;
                    tcol[r][dof] = nint1 + nint2 + NZ + 3 * slot + d;
                }
            }
        free(lab);
    }
    free(slot2);
    return 0;
}

int or_mrt_march(const or_mrt_cfg* c, double* const* MG1, double* const* MH1, double* const* MG2,
                 double* const* MH2, double* X)
{
    const int nz = c->n_iface, K = c->K;
    const int N[2] = {3 * c->nnd[0], 3 * c->nnd[1]};
    const int NH = N[0] + N[1], NG = NH - 6 * nz;
    double* const* MG[2] = {MG1, MG2};
    double* const* MH[2] = {MH1, MH2};
    int *hcol[2], *gcol[2], *tcol[2];
    double **Gg, **Hg, *Gc, *Hc, *load, *T, *rhs, *Gt, *Ht;
    int *piv, r, i, d, k, n, rc = 0;
    for (r = 0; r < 2; r++) {
        hcol[r] = malloc(sizeof(int) * N[r]);
        gcol[r] = malloc(sizeof(int) * N[r]);
        tcol[r] = malloc(sizeof(int) * N[r]);
    }
    if (mrt_maps(c, hcol, gcol, tcol)) {
        for (r = 0; r < 2; r++) { free(hcol[r]); free(gcol[r]); free(tcol[r]); }
        return -1;
    }
    /* global lag matrices (Assemble_GH_R1: region 1's interface G enters H negated; R2: as is) */
    Gg = malloc(sizeof(double*) * (K + 1));
    Hg = malloc(sizeof(double*) * (K + 1));
    for (k = 0; k <= K; k++) {
        Gg[k] = calloc((long)NH * NG, sizeof(double));
        Hg[k] = calloc((long)NH * NH, sizeof(double));
        for (r = 0; r < 2; r++) {
            const int row0 = r == 0 ? 0 : N[0];
            for (i = 0; i < N[r]; i++)
                for (d = 0; d < N[r]; d++) {
                    const double g = MG[r][k][(long)i * N[r] + d], h = MH[r][k][(long)i * N[r] + d];
                    const long row = row0 + i;
                    Hg[k][row * NH + hcol[r][d]] = h;
                    if (gcol[r][d] >= 0) Gg[k][row * NG + gcol[r][d]] = g;
                    else Hg[k][row * NH + tcol[r][d]] = r == 0 ? -g : g;
                }
        }
    }
    Gc = malloc(sizeof(double) * (long)NH * NG);
    Hc = malloc(sizeof(double) * (long)NH * NH);
    Gt = malloc(sizeof(double) * (long)NH * NG);
    Ht = malloc(sizeof(double) * (long)NH * NH);
    T = calloc((long)(c->numstep + 1) * NG, sizeof(double));
    load = malloc(sizeof(double) * NG);
    rhs = malloc(sizeof(double) * NH);
    piv = malloc(sizeof(int) * NH);
    memset(X, 0, sizeof(double) * (long)(c->numstep + 1) * NH);
    for (n = 1; n <= c->numstep && rc == 0; n++) {
        /* SetLoad at fCurrStep = n: interior FBC values times f(n dt) (:935-1026) */
        const double f = or_mrt_load_factor(c, n * c->dt);
        int s = c->scheme == 0 ? n : n - 1, m, fi;
        memset(load, 0, sizeof(double) * NG);
        for (r = 0; r < 2; r++)
            for (i = 0; i < c->n_fbc[r]; i++) {
                const int dof = c->fbc_dof[r][i];
                if (gcol[r][dof] >= 0) load[gcol[r][dof]] = c->fbc_val[r][i] * f;
            }
        if (c->scheme == 1 && n == 1) continue;   /* Solve_Averaging: nothing at step 1 */
        memset(rhs, 0, sizeof(double) * NH);
        if (c->scheme == 0 || n == 2) {
            memcpy(Gc, Gg[0], sizeof(double) * (long)NH * NG);
            memcpy(Hc, Hg[0], sizeof(double) * (long)NH * NH);
        } else {
            for (i = 0; i < NH * NG; i++) Gc[i] = 4 * Gg[0][i] + (K >= 1 ? Gg[1][i] : 0);
            for (i = 0; i < NH * NH; i++) Hc[i] = 4 * Hg[0][i] + (K >= 1 ? Hg[1][i] : 0);
        }
        for (i = 0; i < NH; i++) {
            double acc = 0;
            for (m = 0; m < NG; m++) acc += Gc[(long)i * NG + m] * load[m];
            rhs[i] += acc;
        }
#define MRT_TERM(w, lag, step)                                                                   \
    do {                                                                                        \
        const double* Gk = Gg[lag];                                                             \
        const double* Hk = Hg[lag];                                                             \
        const double* tv = T + (long)(step) * NG;                                               \
        const double* xv = X + (long)(step) * NH;                                               \
        for (i = 0; i < NH; i++) {                                                              \
            double acc = 0;                                                                     \
            for (m = 0; m < NG; m++) acc += Gk[(long)i * NG + m] * tv[m];                       \
            for (m = 0; m < NH; m++) acc -= Hk[(long)i * NH + m] * xv[m];                       \
            rhs[i] += (w) * acc;                                                                \
        }                                                                                       \
    } while (0)
        if (c->scheme == 0) {
            /* :1059-1076 */
            for (fi = 1; fi <= (n < K ? n : K); fi++) MRT_TERM(1.0, fi, n - fi);
        } else if (n >= 3) {
            /* :1150-1230, t_0 = x_0 = 0 makes the bgn terms vanish */
            for (i = 0; i < NH * NG; i++) Gt[i] = (K >= 2 ? Gg[2][i] : 0) + 2 * (K >= 1 ? Gg[1][i] : 0);
            for (i = 0; i < NH * NH; i++) Ht[i] = (K >= 2 ? Hg[2][i] : 0) + 2 * (K >= 1 ? Hg[1][i] : 0);
            {
                const double* tv = T + (long)(s - 1) * NG;
                const double* xv = X + (long)(s - 1) * NH;
                for (i = 0; i < NH; i++) {
                    double acc = 0;
                    for (m = 0; m < NG; m++) acc += Gt[(long)i * NG + m] * tv[m];
                    for (m = 0; m < NH; m++) acc -= Ht[(long)i * NH + m] * xv[m];
                    rhs[i] += acc;
                }
            }
            for (fi = 3; fi <= (s < K ? s : K); fi++) MRT_TERM(1.0, fi, s + 1 - fi);
            for (fi = 2; fi <= (s - 1 < K ? s - 1 : K); fi++) MRT_TERM(2.0, fi, s - fi);
            for (fi = 1; fi <= (s - 2 < K ? s - 2 : K); fi++) MRT_TERM(1.0, fi, s - 1 - fi);
        }
#undef MRT_TERM
        memcpy(Ht, Hc, sizeof(double) * (long)NH * NH);
        if (or_lu_factor(NH, Ht, piv)) rc = -2;
        else {
            or_lu_solve(NH, Ht, piv, rhs);
            memcpy(X + (long)s * NH, rhs, sizeof(double) * NH);
            memcpy(T + (long)s * NG, load, sizeof(double) * NG);
        }
    }
    for (k = 0; k <= K; k++) { free(Gg[k]); free(Hg[k]); }
    free(Gg); free(Hg); free(Gc); free(Hc); free(Gt); free(Ht); free(T); free(load); free(rhs);
    free(piv);
    for (r = 0; r < 2; r++) { free(hcol[r]); free(gcol[r]); free(tcol[r]); }
    return rc;
}

int or_mrt_extract(const or_mrt_cfg* c, int region, const double* x, double* u)
{
    /* ExtractField (:1251-1301): region displacements from the global unknowns */
    int* hcol[2], *gcol[2], *tcol[2], r, d, rc;
    for (r = 0; r < 2; r++) {
        hcol[r] = malloc(sizeof(int) * 3 * c->nnd[r]);
        gcol[r] = malloc(sizeof(int) * 3 * c->nnd[r]);
        tcol[r] = malloc(sizeof(int) * 3 * c->nnd[r]);
    }
    rc = mrt_maps(c, hcol, gcol, tcol);
    if (!rc)
        for (d = 0; d < 3 * c->nnd[region]; d++) u[d] = x[hcol[region][d]];
    for (r = 0; r < 2; r++) { free(hcol[r]); free(gcol[r]); free(tcol[r]); }
    return rc;
}
