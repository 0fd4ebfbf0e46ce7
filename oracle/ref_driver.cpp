// ref_driver.cpp — a C-callable harness around the reference's own element and Green's-function
// classes, compiled from their unmodified sources (see oracle/Makefile, target `ref`).
//
// TEST INFRASTRUCTURE ONLY: it exists so the oracle and the golden fixtures can be checked
// against the reference's compiled code.  The reference's ElastoDynamicsT.cpp cannot be compiled
// here (it includes PETSc headers that the image does not have), so ref_assemble_step restates
// its 40-line pair loop and scatter (src/ElastoDynamicsT.cpp:108-280) around the reference's
// TDElement*::FormGH_* calls.
#include "GreenFunctionT.h"
#include "ShapeFunctionT.h"
#include "TDElementHeavisideLinearT.h"
#include "TDElementRampHeaviLinearT.h"
#include "TDElementStepLinearT.h"

#include <cmath>
#include <cstring>
#include <vector>

using namespace TD_BEM;

namespace {

TDElementT* make_element(int kernel, int etype)
{
    switch (kernel) {
    case 2: return new TDElementHeavisideLinearT(etype);
    case 4: return new TDElementStepLinearT(etype);
    case 5: return new TDElementRampHeaviLinearT(etype);
    default: return nullptr;
    }
}

}  // namespace

extern "C" {

// One pair through TDElement*::FormGH_Regular / FormGH_Singular.  t and dt are passed exactly as
// main.cpp:91-92 forms them (t = n*dt), so the reference's fT==fT_1 test sees step 1 as it would.
int ref_element(int kernel, int etype, int nq, double E, double nu, double rho,
                const double* coll, const double* X, double t, double dt, int singular,
                double* GE1, double* GE2, double* HE1, double* HE2, double* C1, double* C2)
{
    try {
        TDElementT* el = make_element(kernel, etype);
        if (!el) return -1;
        int ennd = etype == 1 ? 4 : 8;
        std::vector<double> xc(X, X + 3 * ennd), cc(coll, coll + 3);
        el->SetNQ(nq);
        el->SetMaterial(E, nu, rho);
        el->SetCoords(xc.data());
        el->SetCollocation(cc.data());
        el->SetTime(t, dt);
        if (singular < 0)
            el->FormGH_Regular();
        else
            el->FormGH_Singular(singular);
        const double *g1, *g2, *h1, *h2, *c1, *c2;
        el->GetGHLinear(&g1, &g2, &h1, &h2, &c1, &c2);
        std::memcpy(GE1, g1, sizeof(double) * 9 * ennd);
        std::memcpy(GE2, g2, sizeof(double) * 9 * ennd);
        std::memcpy(HE1, h1, sizeof(double) * 9 * ennd);
        std::memcpy(HE2, h2, sizeof(double) * 9 * ennd);
        std::memcpy(C1, c1, sizeof(double) * 9);
        std::memcpy(C2, c2, sizeof(double) * 9);
        delete el;
        return 0;
    } catch (const char*) {
        return -1;
    }
}

// GreenFunctionT at one point: Compute(kernel), then the getters.  out = DU1|DU2|DT1|DT2|K1|K2.
int ref_green(int kernel, double E, double nu, double rho, const double* dir, const double* nrm,
              double t, double dt, double* out)
{
    GreenFunctionT g;
    g.SetMaterial(E, nu, rho);
    g.SetSpatial(dir, nrm);
    g.SetTime(t, dt);
    g.Compute(kernel);
    double *du1, *du2, *dt1, *dt2;
    g.GetStokesLinear(&du1, &du2, &dt1, &dt2);
    std::memcpy(out, du1, 72);
    std::memcpy(out + 9, du2, 72);
    std::memcpy(out + 18, dt1, 72);
    std::memcpy(out + 27, dt2, 72);
    if (kernel == 5) {
        double *k1, *k2;
        g.GetKelvinRampedLinear(&k1, &k2);
        std::memcpy(out + 36, k1, 72);
        std::memcpy(out + 45, k2, 72);
    } else {
        std::memcpy(out + 36, g.GetKelvinLinear(), 72);
        std::memcpy(out + 45, g.GetKelvinLinear(), 72);
    }
    return 0;
}

// The ramped-Heaviside functions of GreenFunctionT, which no Compute case reaches: the 40-point
// time integrals Int_Stokes_RampedHeaviside_Linear (src/GreenFunctionT.cpp:589-622) and
// Int_Kelvin_RampedHeaviside_Linear (:661-680), out = DU1|DU2|DT1|DT2|K1|K2; the closed form
// Stokes_RampedHeaviside (:449-519) at time t, point = U|T; and Int_RampedHeaviside_Linear
// (:705-731), g = g1|g2.
int ref_green_ramp(double E, double nu, double rho, const double* dir, const double* nrm,
                   double t, double dt, double* out, double* point, double* g)
{
    GreenFunctionT gf;
    gf.SetMaterial(E, nu, rho);
    gf.SetSpatial(dir, nrm);
    gf.SetTime(t, dt);
    gf.Int_Stokes_RampedHeaviside_Linear(out, out + 9, out + 18, out + 27);
    gf.Int_Kelvin_RampedHeaviside_Linear();
    double *k1, *k2;
    gf.GetKelvinRampedLinear(&k1, &k2);
    std::memcpy(out + 36, k1, 72);
    std::memcpy(out + 45, k2, 72);
    gf.Stokes_RampedHeaviside(point, point + 9, t);
    gf.Int_RampedHeaviside_Linear(g[0], g[1]);
    return 0;
}

// ShapeFunctionT::Evaluate(eta1, eta2): out = N[ennd] | x[3] | n[3] | J.
int ref_shape_eval(int etype, const double* X, double e1, double e2, double* out)
{
    try {
        ShapeFunctionT s(etype);
        int ennd = s.GetENND();
        std::vector<double> xc(X, X + 3 * ennd);
        s.SetCoords(xc.data());
        s.Evaluate(e1, e2);
        std::memcpy(out, s.GetShapeValues(), sizeof(double) * ennd);
        std::memcpy(out + ennd, s.GetX(), 24);
        std::memcpy(out + ennd + 3, s.GetOutNormal(), 24);
        out[ennd + 6] = s.GetJacobian();
        return 0;
    } catch (const char*) {
        return -1;
    }
}

// The pair loop of ElastoDynamicsT::G_H_Driver around the reference element class.
int ref_assemble_step(int kernel, int etype, int nq, double E, double nu, double rho, int nnd,
                      const double* coords, int nel, const int* ien, int if_ext, int lower,
                      int upper, double t, double dt, double* G1, double* G2, double* H1,
                      double* H2)
{
    try {
        TDElementT* el = make_element(kernel, etype);
        if (!el) return -1;
        const int ennd = etype == 1 ? 4 : 8, sd = 3;
        const long ncol = 3L * nnd, nrow = 3L * (upper - lower + 1);
        std::memset(G1, 0, sizeof(double) * nrow * ncol);
        std::memset(G2, 0, sizeof(double) * nrow * ncol);
        std::memset(H1, 0, sizeof(double) * nrow * ncol);
        std::memset(H2, 0, sizeof(double) * nrow * ncol);
        std::vector<double> ele_nodes(sd * ennd), collocation(sd);
        el->SetNQ(nq);
        el->SetMaterial(E, nu, rho);
        el->SetCoords(ele_nodes.data());
        el->SetCollocation(collocation.data());
        el->SetTime(t, dt);
        for (int ni = lower; ni <= upper; ni++) {
            for (int d = 0; d < sd; d++) collocation[d] = coords[ni * sd + d];
            for (int e = 0; e < nel; e++) {
                for (int a = 0; a < ennd; a++)
                    for (int d = 0; d < sd; d++)
                        ele_nodes[a * sd + d] = coords[ien[e * ennd + a] * sd + d];
                int singular = -1;
                for (int a = 0; a < ennd; a++) {
                    double r2 = std::pow(ele_nodes[a * sd] - collocation[0], 2) +
                                std::pow(ele_nodes[a * sd + 1] - collocation[1], 2) +
                                std::pow(ele_nodes[a * sd + 2] - collocation[2], 2);
                    if (r2 < 1e-9) singular = a;
                }
                if (singular == -1)
                    el->FormGH_Regular();
                else
                    el->FormGH_Singular(singular);
                const double *g1, *g2, *h1, *h2, *c1, *c2;
                el->GetGHLinear(&g1, &g2, &h1, &h2, &c1, &c2);
                long loc = ni - lower;
                for (int a = 0; a < ennd; a++) {
                    long nid = ien[e * ennd + a];
                    for (int i = 0; i < sd; i++)
                        for (int j = 0; j < sd; j++) {
                            long g = (i + loc * sd) * ncol + nid * sd + j;
                            int o = i * (ennd * sd) + a * sd + j;
                            G1[g] += g1[o];
                            G2[g] += g2[o];
                            H1[g] += h1[o];
                            H2[g] += h2[o];
                        }
                }
                for (int i = 0; i < sd; i++)
                    for (int j = 0; j < sd; j++) {
                        long g = (loc * sd + i) * ncol + (long)ni * sd + j;
                        H1[g] += c1[i * sd + j];
                        H2[g] += c2[i * sd + j];
                    }
            }
            if (if_ext && t == dt) {
                long loc = ni - lower;
                for (int d = 0; d < sd; d++) {
                    long g = (loc * sd + d) * ncol + (long)ni * sd + d;
                    H1[g] += 0.5 * dt;
                    H2[g] += 0.5 * dt;
                }
            }
        }
        delete el;
        return 0;
    } catch (const char*) {
        return -1;
    }
}

}  // extern "C"
