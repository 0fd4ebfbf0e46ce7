/*
 * oracle.h — CPU restatement of the two TD-BEM hot paths, in plain C (FP64).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/, the smoke() entry of
 * __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The product path
 * (td-bem_amd/) never links, loads or calls anything under oracle/.
 *
 * Every function cites the reference file:line it restates (paths relative to the
 * XiaoyongBai/TD-BEM tree).  Pinning of each function (goldens from the reference's own compiled
 * classes in oracle/_ref, closed forms, invariants) is listed in the header of oracle.c.
 */
#ifndef TDBEM_ORACLE_H
#define TDBEM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Temporal kernels selected by GreenFunctionT::Compute(type) (src/GreenFunctionT.cpp:119-145). */
enum {
    OR_KERNEL_HEAVISIDE = 2, /* Kelvin_Linear + Stokes_Heaviside_Linear, TDElementHeavisideLinearT */
    OR_KERNEL_STEP = 4,      /* Kelvin_Linear + Stokes_Step_Linear, TDElementStepLinearT (main path) */
    OR_KERNEL_BSPLINE = 5,   /* cubic B-spline, TDElementRampHeaviLinearT */
    OR_KERNEL_RAMP = 6       /* ramped Heaviside (no Compute case reaches it) in the element loop
                              * of TDElementRampHeaviLinearT */
};

typedef struct {
    double E, nu, rho;   /* inputs */
    double mu, cp, cs;   /* derived by or_material_set */
} or_material;

/* GreenFunctionT::SetMaterial + WaveSpeed (src/GreenFunctionT.cpp:76-93). */
void or_material_set(or_material* m, double E, double nu, double rho);

/* Gauss tables of ShapeFunctionT::SetNQ (src/ShapeFunctionT.cpp:71-169) and of
 * GreenFunctionT::Gaussian_Abs/Weight (src/GreenFunctionT.cpp:1042-1164).  Returns 0 or -1. */
int or_gauss_shape(int nq, const double** abscissae, const double** weights);
int or_gauss_time(int nq, const double** abscissae, const double** weights);

/* ShapeFunctionT::Evaluate(eta1, eta2) (src/ShapeFunctionT.cpp:236-366).
 * etype 1: 4-node bilinear, 2: 8-node serendipity.  X is [ennd][3] row-major.
 * Outputs N[ennd], x[3], unit normal n[3], Jacobian J.  Returns 0 or -1. */
int or_shape_eval(int etype, const double* X, double e1, double e2,
                  double* N, double* x, double* n, double* J);

/* GreenFunctionT::KelvinTraction (src/GreenFunctionT.cpp:147-173): 3x3 row-major [i][k]. */
void or_kelvin_traction(const or_material* m, const double* dir, const double* nrm, double out[9]);
/* GreenFunctionT::Kelvin_Linear (src/GreenFunctionT.cpp:205-231): traction * dt/2. */
void or_kelvin_linear(const or_material* m, const double* dir, const double* nrm, double dt,
                      double out[9]);
/* GreenFunctionT::Stokes_Heaviside_Linear (src/GreenFunctionT.cpp:241-416) at time t. */
void or_stokes_heaviside_linear(const or_material* m, const double* dir, const double* nrm,
                                double t, double dt, double DU1[9], double DU2[9],
                                double DT1[9], double DT2[9]);
/* GreenFunctionT::Stokes_Step_Linear (src/GreenFunctionT.cpp:424-446): HL(t) - HL(t-dt). */
void or_stokes_step_linear(const or_material* m, const double* dir, const double* nrm,
                           double t, double dt, double DU1[9], double DU2[9],
                           double DT1[9], double DT2[9]);
/* GreenFunctionT::Int_Stokes_CubicBSpline_Linear (src/GreenFunctionT.cpp:625-658) and
 * Int_Kelvin_CubicBSpline_Linear (:682-701). */
void or_stokes_bspline_linear(const or_material* m, const double* dir, const double* nrm,
                              double t, double dt, double DU1[9], double DU2[9],
                              double DT1[9], double DT2[9]);
void or_kelvin_bspline_linear(const or_material* m, const double* dir, const double* nrm,
                              double t, double dt, double K1[9], double K2[9]);
/* Ramped-Heaviside load g(t) = min(max(2 t / dt, 0), 1): Stokes_RampedHeaviside at time t
 * (src/GreenFunctionT.cpp:449-519), Int_Stokes_RampedHeaviside_Linear (:589-622),
 * Int_Kelvin_RampedHeaviside_Linear (:661-680) and Int_RampedHeaviside_Linear (:705-731). */
void or_stokes_ramp_point(const or_material* m, const double* dir, const double* nrm, double t,
                          double dt, double U[9], double T[9]);
void or_stokes_ramp_linear(const or_material* m, const double* dir, const double* nrm, double t,
                           double dt, double DU1[9], double DU2[9], double DT1[9], double DT2[9]);
void or_kelvin_ramp_linear(const or_material* m, const double* dir, const double* nrm, double t,
                           double dt, double K1[9], double K2[9]);
void or_ramp_weights(double t, double dt, double g[2]);

/* One (collocation point, element) pair: TDElement{Step,Heaviside,RampHeavi}LinearT::
 * FormGH_Regular / FormGH_Singular (src/TDElementStepLinearT.cpp:28-626 and the two variants).
 * GE1, GE2, HE1, HE2 are [3][3*ennd] row-major, C1, C2 are [3][3].  step_index is the 1-based step (the
 * reference tests fT==fT_1, i.e. step 1).  singular = -1 or the local node of the collocation
 * point.  Returns 0 or -1. */
int or_element(int kernel, int etype, int nq, const or_material* m, const double* coll,
               const double* X, double t, double dt, int step_index, int singular,
               double* GE1, double* GE2, double* HE1, double* HE2, double* C1, double* C2);

/* ElastoDynamicsT::G_H_Driver + AssembleGH + AssembleC (src/ElastoDynamicsT.cpp:108-280).
 * Rows of collocation nodes [lower, upper] (inclusive), outputs dense
 * [3*(upper-lower+1)][3*nnd] row-major: G1 = fG_1 (phi_1 part, "G_0" in main.cpp),
 * G2 = fG_2 (phi_2 part, "G_1" in main.cpp), H1, H2.  nthreads<=0 uses the OpenMP default. */
int or_assemble_step(int kernel, int etype, int nq, const or_material* m, int nnd,
                     const double* coords, int nel, const int* ien, int if_ext, int lower,
                     int upper, double t, double dt, int step_index, double* G1, double* G2,
                     double* H1, double* H2, int nthreads);

/* SolverT::ComputeMaxStep (src/SolverT.cpp:211-250). */
int or_max_step(const or_material* m, int nnd, const double* coords, double dt);

/* Lag store of SolverStepLinearT::UpdateGHLinear (src/SolverStepLinearT.cpp:58-175), driven as
 * main.cpp:89-121 drives it: assembly for steps 1..min(numstep, K), then
 *   M[n-1] (+)= phi_2 part of step n (INSERT when n==1), M[n] = phi_1 part of step n.
 * MG/MH point at K+1 dense N x N buffers (N = 3*nnd), zeroed by this call. */
int or_build_lags(int kernel, int etype, int nq, const or_material* m, int nnd,
                  const double* coords, int nel, const int* ien, int if_ext, double dt,
                  int numstep, int K, double** MG, double** MH, int nthreads);

typedef struct {
    int N, K, numstep;
    double dt, a1, a2;
    int n_ubc;
    const int* ubc_dof;
    const double* ubc_val;
    int n_fbc;
    const int* fbc_dof;
    const double* fbc_val;
    /* load history f(t) = amp * min(1, t/t_rise) (t_rise <= 0: Heaviside).  The reference's
     * SolverT::LoadTimeVariation (src/SolverT.cpp:379-394) is amp=1, t_rise=0. */
    double amp, t_rise;
} or_march_cfg;

/* f(t) of the load history (generalisation of src/SolverT.cpp:379-394). */
double or_load_factor(const or_march_cfg* c, double t);

/* SolverStepLinearT::Solve_Averaging restated literally (src/SolverStepLinearT.cpp:314-544,
 * SolverT::SetLoad :344-377, ExchangeColumns :265-342), with a partial-pivoting LU in place of
 * PETSc's KSPSolve.  U, T: [numstep][N] histories x_0 .. x_{numstep-1} as RecordResult sees them
 * (x_0 = 0).  Returns 0, -1 (bad alpha: "Alpha is out of range"), -2 (singular matrix). */
int or_march_literal(const or_march_cfg* c, double* const* MG, double* const* MH, double* U,
                     double* T);

/* The same recurrence written as hist_s = sum_k M_k z_k (one pass over the lag store). */
int or_march_fused(const or_march_cfg* c, double* const* MG, double* const* MH, double* U,
                   double* T);

/* Two-region solver (SolverMRT, src/SolverMRT.cpp).  Region r: nnd[r] nodes with coordinates
 * coords[r] (for matching the interface nodes), n_iface interface nodes iface[r] (0-based, same
 * count in both regions; region 2's are matched to region 1's by position), traction loads
 * fbc_dof/val[r] (interior dofs only count; SolverMRT ignores displacement conditions).
 * scheme 0: Solve_Direct (what SolverMRT::Solve runs, :1028-1033); 1: Solve_Averaging.
 * load_kind 0: f(t) = 5.46366e9 t exp(-2000 t) of SetLoad (:937-939); 1: amp * min(1, t/t_rise).
 * The load of step n is f(n dt) in both schemes, as SetLoad forms it from fCurrStep. */
typedef struct {
    int nnd[2];
    const double* coords[2];
    int n_iface;
    const int* iface[2];
    int n_fbc[2];
    const int* fbc_dof[2];
    const double* fbc_val[2];
    int K, numstep, scheme, load_kind;
    double dt, amp, t_rise;
} or_mrt_cfg;
double or_mrt_load_factor(const or_mrt_cfg* c, double t);
/* X: [numstep+1][3 (nnd[0] + nnd[1])] global unknowns fDis_DB_Global (interior displacements of
 * region 1, of region 2, interface displacements, interface tractions), x_0 = 0.  MG1/MH1 (MG2/MH2)
 * are region 1's (2's) K+1 lag matrices of or_build_lags.  Returns 0, -1 (interface mismatch),
 * -2 (singular matrix). */
int or_mrt_march(const or_mrt_cfg* c, double* const* MG1, double* const* MH1, double* const* MG2,
                 double* const* MH2, double* X);
/* ExtractField (:1251-1301): region `region`'s displacements u[3 nnd] from global unknowns x. */
int or_mrt_extract(const or_mrt_cfg* c, int region, const double* x, double* u);

/* Dense LU with partial pivoting (the KSPSolve stand-in), exposed for tests. */
int or_lu_factor(int n, double* A, int* piv);
void or_lu_solve(int n, const double* LU, const int* piv, double* b);

#ifdef __cplusplus
}
#endif

#endif
