// tdbem_host.h — C++ host classes with the reference's interfaces, over the C ABI of
// include/tdbem.h.  A maintainer can swap them in where main.cpp (src/main.cpp:22-149) uses
// ModelManagerT, TDElementStepLinearT, ElastoDynamicsT and SolverStepLinearT: the same method
// names, argument meanings, ownership (borrowed pointers from getters) and error convention
// (throw const char*).  The compute runs on the GPU; nothing here evaluates kernels.
#pragma once

#include <string>
#include <vector>

#include "../../include/tdbem.h"

namespace TD_BEM {

// Input deck reader (src/InputT.cpp:84-994): same keywords, delimiters " \t=;,", 1-based ids in
// the file, atoi/atof parsing, and the reference's error messages.
class InputT {
public:
    void ReadInput(const char* name);
    int GetFormulation() const { return fFormulation; }
    int GetNumStep() const { return fNumStep; }
    double GetDT() const { return fDT; }
    int GetTimeInterpolation() const { return fInterpolation; }
    int GetNQ() const { return fNQ; }
    int GetSD() const { return fSD; }
    int GetIfExt() const { return fIfExt; }
    int GetNND() const { return fNND; }
    int GetElementType() const { return fEType; }
    int GetENND() const { return fENND; }
    int GetNEL() const { return fNEL; }
    const int* GetIEN() const { return fIEN.data(); }
    const double* GetCoords() const { return fCoords.data(); }
    void GetMaterialConstants(double& E, double& nu, double& rho) const
    {
        E = fE;
        nu = fNu;
        rho = fDensity;
    }
    void BoundaryCondition(int& UBC_num, const int** UBC_DOFs, const double** UBC_Values,
                           int& FBC_num, const int** FBC_DOFs, const double** FBC_Values) const;
    int GetNumInterface() const { return (int)fInterface.size(); }
    const int* GetInterface() const { return fInterface.data(); }

private:
    int fFormulation = 1, fInterpolation = 1, fNumStep = 1, fNQ = 6;
    double fDT = 0, fE = 1, fNu = 0, fDensity = 1;
    int fSD = 3, fIfExt = 0, fNND = 0, fNEL = 0, fEType = 1, fENND = 4;
    int fNND_Aux = 0, fNEL_Aux = 0, fEType_Aux = 1, fENND_Aux = 4;
    std::vector<double> fCoords, fCoords_Aux;
    std::vector<int> fIEN, fIEN_Aux;
    int fUBC_num = 0, fFBC_num = 0, fNumInterface = 0;
    std::vector<int> fUBC_DOFs, fFBC_DOFs, fInterface;
    std::vector<double> fUBC_Values, fFBC_Values;
};

// src/ModelManagerT.cpp:356-545 (single process: the local boundary data is the global one).
class ModelManagerT {
public:
    void ReadInput(const char* name);
    int GetFormulation() const { return fInput.GetFormulation(); }
    int GetNumStep() const { return fInput.GetNumStep(); }
    double GetDT() const { return fInput.GetDT(); }
    int GetTimeInterpolation() const { return fInput.GetTimeInterpolation(); }
    int GetNQ() const { return fInput.GetNQ(); }
    int GetSD() const { return fInput.GetSD(); }
    int GetIfExt() const { return fInput.GetIfExt(); }
    int GetNND() const { return fInput.GetNND(); }
    int GetElementType() const { return fInput.GetElementType(); }
    int GetENND() const { return fInput.GetENND(); }
    int GetNEL() const { return fInput.GetNEL(); }
    const int* GetIEN() const { return fInput.GetIEN(); }
    const double* GetCoords() const { return fInput.GetCoords(); }
    void GetMaterialConstants(double& E, double& nu, double& rho) const
    {
        fInput.GetMaterialConstants(E, nu, rho);
    }
    int GetNumInterface() const { return fInput.GetNumInterface(); }
    const int* GetInterface() const { return fInput.GetInterface(); }
    void Distributed_BC(int low, int high);
    void BoundaryCondition_Local(int& UBC_num, const int** UBC_DOFs, const double** UBC_Values,
                                 int& FBC_num, const int** FBC_DOFs,
                                 const double** FBC_Values) const;
    void BoundaryCondition_Global(int& UBC_num, const int** UBC_DOFs, const double** UBC_Values,
                                  int& FBC_num, const int** FBC_DOFs,
                                  const double** FBC_Values) const;

private:
    InputT fInput;
    std::vector<int> fUBC_local, fFBC_local;
    std::vector<double> fUBCv_local, fFBCv_local;
};

// One (collocation point, element) pair, src/TDElementT.h:16-115.  Evaluated on the GPU by a
// reusable element evaluator (tdbem_element_form), made at the first Form and made again when
// the rule or the material changes.
class TDElementT {
public:
    explicit TDElementT(int type, int kernel);
    virtual ~TDElementT();
    TDElementT(const TDElementT&) = delete;
    TDElementT& operator=(const TDElementT&) = delete;
    void SetCollocation(double* collocation) { fCollocation = collocation; }
    void SetCoords(double* coords) { fCoords = coords; }
    virtual void SetTime(double t, double t_1)
    {
        fT = t;
        fT_1 = t_1;
    }
    void SetNQ(int nq) { fNQ = nq; }
    void SetMaterial(double E, double nu, double density)
    {
        fE = E;
        fNu = nu;
        fDensity = density;
    }
    virtual void FormGH_Regular() { Form(-1); }
    virtual void FormGH_Singular(int singular) { Form(singular); }
    void GetGHLinear(const double** GE_1, const double** GE_2, const double** HE_1,
                     const double** HE_2, const double** C_1, const double** C_2) const;

protected:
    void Form(int singular);
    tdbem_element* fEval = nullptr;
    int fEvalNQ = 0;
    double fEvalMat[3] = {0, 0, 0};
    int fType, fKernel, fNQ = 6;
    double fE = 1, fNu = 0, fDensity = 1, fT = 1, fT_1 = 1;
    double* fCollocation = nullptr;
    double* fCoords = nullptr;
    std::vector<double> fGE_1, fGE_2, fHE_1, fHE_2, fC_1, fC_2;
};

class TDElementStepLinearT : public TDElementT {
public:
    explicit TDElementStepLinearT(int type) : TDElementT(type, TDBEM_KERNEL_STEP) {}
};

class TDElementHeavisideLinearT : public TDElementT {
public:
    explicit TDElementHeavisideLinearT(int type) : TDElementT(type, TDBEM_KERNEL_HEAVISIDE) {}
};

// src/TDElementRampHeaviLinearT.h: the cubic B-spline pulse (GreenFunctionT::Compute(5))
class TDElementRampHeaviLinearT : public TDElementT {
public:
    explicit TDElementRampHeaviLinearT(int type) : TDElementT(type, TDBEM_KERNEL_BSPLINE) {}
};

// src/ElastoDynamicsT.h:19-55.  G_H_Driver fills the four dense row blocks on the GPU and copies
// them to host buffers owned by this object (valid until the next call).
class ElastoDynamicsT {
public:
    explicit ElastoDynamicsT(ModelManagerT* model);
    ~ElastoDynamicsT();
    void SetLowerUpper(int lower, int upper);
    void G_H_Driver(double t, double t_1);
    void GetGHLinear(double** G_1, double** G_2, double** H_1, double** H_2);
    tdbem_context* Context() const { return fCtx; }

private:
    ModelManagerT* fModel;
    tdbem_context* fCtx = nullptr;
    int fLower = 0, fUpper = -1;
    std::vector<double> fG_1, fG_2, fH_1, fH_2;
};

// src/SolverT.h:21-127 + src/SolverStepLinearT.h.  The lag store and the march live on the GPU:
// UpdateGHLinear accepts the reference's call (and ignores the host copies, which the GPU builds
// itself in one pass), Solve advances one step, RecordResult appends displacement.txt.
class SolverStepLinearT {
public:
    explicit SolverStepLinearT(ModelManagerT* model);
    ~SolverStepLinearT();
    void SetTimeControl(double dt, int num_steps);
    void Initialize();
    void SetCurrStep(int curr_step) { fCurrStep = curr_step; }
    int GetMaxStep() const { return fMaxStep; }
    void GetNodeLowHigh(int* low, int* high) const;
    void GetDofLowHigh(int* low, int* high) const;
    void UpdateGHLinear(const double* G_0, const double* G_1, const double* H_0,
                        const double* H_1);
    void SetLoad() {}
    void Solve(double a1, double a2);
    void RecordResult(const char* path = "displacement.txt");
    double LoadTimeVariation(double) const { return 1.0; }  // src/SolverT.cpp:379-394
    // Full histories (x_0 .. x_{numstep-1}) after the march.
    void GetHistory(std::vector<double>& U, std::vector<double>& T);
    // src/SolverT.cpp:670-732 with one process: "<prefix><step>.txt", %15.5e rows, no header.
    void WriteGH(const char* name_prefix, int step, int num_row, int num_column,
                 const double* data, int comm_size = 1, int comm_rank = 0);
    // The same matrices at full precision (no reference counterpart): "<prefix><step>.bin",
    // two int64 (rows, columns) then the row-major doubles.
    void WriteGH_Binary(const char* name_prefix, int step, int num_row, int num_column,
                        const double* data);

private:
    ModelManagerT* fModel;
    tdbem_context* fCtx = nullptr;
    int fMaxStep = 0, fNumStep = 0, fCurrStep = 0, fN = 0;
    double fDt = 0;
    bool fInit = false;
};

// src/SolverMRT.h:24-134: two regions (model_1, model_2) coupled on their interface nodes, each
// on its own GPU context; lag stores, the global step matrix and the march live on the GPU
// (tdbem_mrt_*).  Solve runs Solve_Direct as the reference's does; Solve_Averaging is the other
// scheme (one scheme per run).  The loads are each model's FBC list times the pulse of SetLoad.
// RecordResult appends the time and the region-2 dof the reference records (46*3-1, its "coarse
// mesh" line) of step fCurrStep-1 to displacement.txt and writes the OutPutVTK files
// BEM_1st_<n>.vtk and BEM_2nd_<n>.vtk.
class SolverMRT {
public:
    SolverMRT(ModelManagerT* model_1, ModelManagerT* model_2);
    ~SolverMRT();
    void SetTimeControl(double dt, int num_steps);
    void SetCurrStep(int curr_step) { fCurrStep = curr_step; }
    void SetMaxStep(int max_step);
    int GetMaxStep() const { return fMaxStep; }
    void GetNodeLowHigh(int* low_1, int* high_1, int* low_2, int* high_2) const;
    void GetDofLowHigh(int* low_1, int* high_1, int* low_2, int* high_2) const;
    void UpdateGHLinear(const double*, const double*, const double*, const double*,
                        const double*, const double*, const double*, const double*)
    {
        // the lag stores are assembled on the GPU in one pass by the first solve
    }
    void SetLoad() {}
    void Solve() { Solve_Direct(); }
    void Solve_Direct();
    void Solve_Averaging();
    void ExtractField() {}
    void RecordResult(const char* path = "displacement.txt", int record_dof = 46 * 3 - 1,
                      bool vtk = true);
    void OutPutVTK(std::string filename, int region);
    // displacements u [3 nnd_r] of region r (0 or 1) at step s (0 .. numstep)
    void GetField(int region, int step, std::vector<double>& u);

private:
    void Advance(int scheme, int target);
    ModelManagerT* fModel[2];
    tdbem_context* fCtx[2] = {nullptr, nullptr};
    tdbem_mrt* fMrt = nullptr;
    int fMaxStep = 10, fNumStep = 0, fCurrStep = 0, fScheme = -1;
    double fDt = 0;
};

}  // namespace TD_BEM
