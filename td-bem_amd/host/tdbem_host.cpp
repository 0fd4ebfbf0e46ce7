// tdbem_host.cpp — the reference's host classes over the C ABI (see tdbem_host.h).
#include "tdbem_host.h"

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>

namespace TD_BEM {

namespace {

// strtok with the reference's delimiters (src/InputT.cpp:21)
std::vector<std::string> tokens(const std::string& line)
{
    std::vector<std::string> out;
    std::string cur;
    for (char ch : line) {
        if (ch == ' ' || ch == '\t' || ch == '=' || ch == ';' || ch == ',' || ch == '\r') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else {
            cur.push_back(ch);
        }
    }
    if (!cur.empty()) out.push_back(cur);
    return out;
}

[[noreturn]] void fail(const char* msg) { throw msg; }

void check(tdbem_context* ctx, int rc)
{
    if (rc != TDBEM_OK) {
        static thread_local std::string msg;
        msg = tdbem_last_error(ctx);
        throw msg.c_str();
    }
}

}  // namespace

// ------------------------------------------------------------------------------------------
// InputT
// ------------------------------------------------------------------------------------------
void InputT::ReadInput(const char* name)
{
    std::ifstream fin(name);
    if (!fin.good()) fail("file not found");
    std::vector<std::string> lines;
    std::string l;
    while (std::getline(fin, l)) lines.push_back(l);
    size_t i = 0;
    auto next = [&]() -> std::vector<std::string> {
        if (i >= lines.size()) fail("unexpected end of file");
        return tokens(lines[i++]);
    };
    auto read_list_i = [&](int n, std::vector<int>& out) {
        out.clear();
        while ((int)out.size() < n) {
            auto t = next();
            if (t.empty()) fail("empty line inside a list");
            for (auto& s : t)
                if ((int)out.size() < n) out.push_back(std::atoi(s.c_str()) - 1);
        }
    };
    auto read_list_d = [&](int n, std::vector<double>& out) {
        out.clear();
        while ((int)out.size() < n) {
            auto t = next();
            if (t.empty()) fail("empty line inside a list");
            for (auto& s : t)
                if ((int)out.size() < n) out.push_back(std::atof(s.c_str()));
        }
    };
    while (i < lines.size()) {
        auto t = tokens(lines[i++]);
        if (t.empty()) continue;
        const std::string& k = t[0];
        auto val = [&](const char* msg) -> const char* {
            if (t.size() < 2) fail(msg);
            return t[1].c_str();
        };
        if (k == "formulation") fFormulation = std::atoi(val("if_ext is not set properly"));
        else if (k == "numstep") fNumStep = std::atoi(val("number of step to run is not set properly"));
        else if (k == "dt") fDT = std::atof(val("time step size is not set properly"));
        else if (k == "interpolation") fInterpolation = (int)std::atof(val("interpolation is not set properly"));
        else if (k == "nq") fNQ = (int)std::atof(val("number of Gaussian points (nq) is not set properly"));
        else if (k == "E") fE = std::atof(val("Young's modulus is not set properly"));
        else if (k == "nu") fNu = std::atof(val("possion's ratio is not set properly"));
        else if (k == "density") fDensity = std::atof(val("density is not set properly"));
        else if (k == "sd") fSD = std::atoi(val("spatial dimension is not set properly"));
        else if (k == "if_ext") fIfExt = std::atoi(val("if_ext is not set properly"));
        else if (k == "nnd") fNND = std::atoi(val("number of nodes is not set properly"));
        else if (k == "nel") fNEL = std::atoi(val("number of nodes is not set properly"));
        else if (k == "EType") {
            fEType = std::atoi(val("number of nodes is not set properly"));
            if (fEType == 1) fENND = 4;
            else if (fEType == 2) fENND = 8;
            else fail("InputT::ReadInput, unsupported element type");
        } else if (k == "nodes") {
            fCoords.assign(3 * (size_t)fNND, 0.0);
            for (int n = 0; n < fNND; n++) {
                auto r = next();
                for (int d = 0; d < 3 && d < (int)r.size(); d++)
                    fCoords[3 * n + d] = std::atof(r[d].c_str());
            }
        } else if (k == "IEN") {
            fIEN.assign((size_t)fNEL * fENND, 0);
            for (int e = 0; e < fNEL; e++) {
                auto r = next();
                // the reference hands atoi a NULL token on a short row (src/InputT.cpp:531-535)
                if ((int)r.size() < fENND) fail("IEN row has too few node ids");
                for (int a = 0; a < fENND; a++) fIEN[(size_t)e * fENND + a] = std::atoi(r[a].c_str()) - 1;
            }
        } else if (fFormulation == 3 && k == "nnd_aux") fNND_Aux = std::atoi(val("number of nodes is not set properly"));
        else if (fFormulation == 3 && k == "nel_aux") fNEL_Aux = std::atoi(val("number of nodes is not set properly"));
        else if (fFormulation == 3 && k == "etype_aux") {
            fEType_Aux = std::atoi(val("number of nodes is not set properly"));
            // src/InputT.cpp:595-597 tests fEType for the 8-node case
            if (fEType_Aux == 1) fENND_Aux = 4;
            else if (fEType == 2) fENND_Aux = 8;
            else fail("InputT::ReadInput, unsupported element type");
        } else if (fFormulation == 3 && k == "nodes_aux") {
            fCoords_Aux.assign(3 * (size_t)fNND_Aux, 0.0);
            for (int n = 0; n < fNND_Aux; n++) {
                auto r = next();
                for (int d = 0; d < 3 && d < (int)r.size(); d++)
                    fCoords_Aux[3 * n + d] = std::atof(r[d].c_str());
            }
        } else if (fFormulation == 3 && k == "IEN_aux") {
            fIEN_Aux.assign((size_t)fNEL_Aux * fENND_Aux, 0);
            for (int e = 0; e < fNEL_Aux; e++) {
                auto r = next();
                if ((int)r.size() < fENND_Aux) fail("IEN_aux row has too few node ids");
                for (int a = 0; a < fENND_Aux; a++)
                    fIEN_Aux[(size_t)e * fENND_Aux + a] = std::atoi(r[a].c_str()) - 1;
            }
        } else if (k == "UBC_num") fUBC_num = std::atoi(val("number of displacement BC is not set properly"));
        else if (k == "UBC_dof") read_list_i(fUBC_num, fUBC_DOFs);
        else if (k == "UBC_values") read_list_d(fUBC_num, fUBC_Values);
        else if (k == "FBC_num") fFBC_num = std::atoi(val("number of displacement BC is not set properly"));
        else if (k == "FBC_dof") read_list_i(fFBC_num, fFBC_DOFs);
        else if (k == "FBC_values") read_list_d(fFBC_num, fFBC_Values);
        else if (k == "num_interface") fNumInterface = std::atoi(val("number of interface nodes is not set properly"));
        else if (k == "interface") read_list_i(fNumInterface, fInterface);
        else fail("InputT::ReadInput, unaccepted key word");
    }
    if (fUBC_DOFs.size() != (size_t)fUBC_num) fUBC_DOFs.resize(fUBC_num, 0);
    if (fUBC_Values.size() != (size_t)fUBC_num) fUBC_Values.resize(fUBC_num, 0.0);
    if (fFBC_DOFs.size() != (size_t)fFBC_num) fFBC_DOFs.resize(fFBC_num, 0);
    if (fFBC_Values.size() != (size_t)fFBC_num) fFBC_Values.resize(fFBC_num, 0.0);
}

void InputT::BoundaryCondition(int& UBC_num, const int** UBC_DOFs, const double** UBC_Values,
                               int& FBC_num, const int** FBC_DOFs,
                               const double** FBC_Values) const
{
    UBC_num = fUBC_num;
    FBC_num = fFBC_num;
    *UBC_DOFs = fUBC_DOFs.data();
    *UBC_Values = fUBC_Values.data();
    *FBC_DOFs = fFBC_DOFs.data();
    *FBC_Values = fFBC_Values.data();
}

// ------------------------------------------------------------------------------------------
// ModelManagerT
// ------------------------------------------------------------------------------------------
void ModelManagerT::ReadInput(const char* name)
{
    fInput.ReadInput(name);
    Distributed_BC(0, 3 * fInput.GetNND() - 1);
}

void ModelManagerT::Distributed_BC(int low, int high)
{
    // src/ModelManagerT.cpp:441-533: keep the dofs in [low, high]
    int nu, nf;
    const int *ud, *fd;
    const double *uv, *fv;
    fInput.BoundaryCondition(nu, &ud, &uv, nf, &fd, &fv);
    fUBC_local.clear();
    fUBCv_local.clear();
    fFBC_local.clear();
    fFBCv_local.clear();
    for (int q = 0; q < nu; q++)
        if (ud[q] >= low && ud[q] <= high) {
            fUBC_local.push_back(ud[q]);
            fUBCv_local.push_back(uv[q]);
        }
    for (int q = 0; q < nf; q++)
        if (fd[q] >= low && fd[q] <= high) {
            fFBC_local.push_back(fd[q]);
            fFBCv_local.push_back(fv[q]);
        }
}

void ModelManagerT::BoundaryCondition_Local(int& UBC_num, const int** UBC_DOFs,
                                            const double** UBC_Values, int& FBC_num,
                                            const int** FBC_DOFs,
                                            const double** FBC_Values) const
{
    UBC_num = (int)fUBC_local.size();
    FBC_num = (int)fFBC_local.size();
    *UBC_DOFs = fUBC_local.data();
    *UBC_Values = fUBCv_local.data();
    *FBC_DOFs = fFBC_local.data();
    *FBC_Values = fFBCv_local.data();
}

void ModelManagerT::BoundaryCondition_Global(int& UBC_num, const int** UBC_DOFs,
                                             const double** UBC_Values, int& FBC_num,
                                             const int** FBC_DOFs,
                                             const double** FBC_Values) const
{
    fInput.BoundaryCondition(UBC_num, UBC_DOFs, UBC_Values, FBC_num, FBC_DOFs, FBC_Values);
}

// ------------------------------------------------------------------------------------------
// TDElementT
// ------------------------------------------------------------------------------------------
TDElementT::TDElementT(int type, int kernel) : fType(type), fKernel(kernel)
{
    if (type != 1 && type != 2)
        fail("TDElementT::TDElementT, unsupported element type. There is no sub element for this element");
    const int ennd = type == 1 ? 4 : 8;
    fGE_1.assign(9 * ennd, 0.0);
    fGE_2 = fGE_1;
    fHE_1 = fGE_1;
    fHE_2 = fGE_1;
    fC_1.assign(9, 0.0);
    fC_2 = fC_1;
}

TDElementT::~TDElementT()
{
    if (fEval) tdbem_element_destroy(fEval);
}

void TDElementT::Form(int singular)
{
    if (!fCoords || !fCollocation) fail("TDElementT: coordinates or collocation point not set");
    if (!fEval || fEvalNQ != fNQ || fEvalMat[0] != fE || fEvalMat[1] != fNu ||
        fEvalMat[2] != fDensity) {
        if (fEval) tdbem_element_destroy(fEval);
        fEval = nullptr;
        if (tdbem_element_create(fType, fKernel, fNQ, fE, fNu, fDensity, &fEval) != TDBEM_OK) {
            static thread_local std::string msg;
            msg = tdbem_element_last_error(nullptr);
            throw msg.c_str();
        }
        fEvalNQ = fNQ;
        fEvalMat[0] = fE;
        fEvalMat[1] = fNu;
        fEvalMat[2] = fDensity;
    }
    if (tdbem_element_form(fEval, 1, fCoords, fCollocation, &singular, fT, fT_1, fGE_1.data(),
                           fGE_2.data(), fHE_1.data(), fHE_2.data(), fC_1.data(),
                           fC_2.data()) != TDBEM_OK) {
        static thread_local std::string msg;
        msg = tdbem_element_last_error(fEval);
        throw msg.c_str();
    }
}

void TDElementT::GetGHLinear(const double** GE_1, const double** GE_2, const double** HE_1,
                             const double** HE_2, const double** C_1, const double** C_2) const
{
    *GE_1 = fGE_1.data();
    *GE_2 = fGE_2.data();
    *HE_1 = fHE_1.data();
    *HE_2 = fHE_2.data();
    *C_1 = fC_1.data();
    *C_2 = fC_2.data();
}

// ------------------------------------------------------------------------------------------
// ElastoDynamicsT
// ------------------------------------------------------------------------------------------
ElastoDynamicsT::ElastoDynamicsT(ModelManagerT* model) : fModel(model)
{
    double E, nu, rho;
    model->GetMaterialConstants(E, nu, rho);
    if (tdbem_create(model->GetNND(), model->GetCoords(), model->GetNEL(), model->GetIEN(),
                     model->GetElementType(), model->GetNQ(), E, nu, rho, model->GetIfExt(),
                     model->GetDT(), model->GetNumStep(), TDBEM_KERNEL_STEP,
                     &fCtx) != TDBEM_OK) {
        static thread_local std::string msg;
        msg = tdbem_last_error(nullptr);
        throw msg.c_str();
    }
    SetLowerUpper(0, model->GetNND() - 1);
}

ElastoDynamicsT::~ElastoDynamicsT() { tdbem_destroy(fCtx); }

void ElastoDynamicsT::SetLowerUpper(int lower, int upper)
{
    fLower = lower;
    fUpper = upper;
    const size_t n = (size_t)3 * (upper - lower + 1) * 3 * fModel->GetNND();
    fG_1.assign(n, 0.0);
    fG_2.assign(n, 0.0);
    fH_1.assign(n, 0.0);
    fH_2.assign(n, 0.0);
}

void ElastoDynamicsT::G_H_Driver(double t, double t_1)
{
    if (std::fabs(t_1 - fModel->GetDT()) > 1e-15 * std::fabs(t_1))
        fail("ElastoDynamicsT::G_H_Driver, t_1 differs from the model time step");
    const int step = (int)std::llround(t / t_1);
    check(fCtx, tdbem_assemble_step(fCtx, step, fLower, fUpper, fG_1.data(), fG_2.data(),
                                    fH_1.data(), fH_2.data()));
}

void ElastoDynamicsT::GetGHLinear(double** G_1, double** G_2, double** H_1, double** H_2)
{
    *G_1 = fG_1.data();
    *G_2 = fG_2.data();
    *H_1 = fH_1.data();
    *H_2 = fH_2.data();
}

// ------------------------------------------------------------------------------------------
// SolverStepLinearT
// ------------------------------------------------------------------------------------------
SolverStepLinearT::SolverStepLinearT(ModelManagerT* model) : fModel(model)
{
    fN = 3 * model->GetNND();
    SetTimeControl(model->GetDT(), model->GetNumStep());
}

SolverStepLinearT::~SolverStepLinearT() { tdbem_destroy(fCtx); }

void SolverStepLinearT::SetTimeControl(double dt, int num_steps)
{
    fDt = dt;
    fNumStep = num_steps;
    if (fCtx) tdbem_destroy(fCtx);
    fCtx = nullptr;
    double E, nu, rho;
    fModel->GetMaterialConstants(E, nu, rho);
    if (tdbem_create(fModel->GetNND(), fModel->GetCoords(), fModel->GetNEL(), fModel->GetIEN(),
                     fModel->GetElementType(), fModel->GetNQ(), E, nu, rho, fModel->GetIfExt(),
                     dt, num_steps, TDBEM_KERNEL_STEP, &fCtx) != TDBEM_OK) {
        static thread_local std::string msg;
        msg = tdbem_last_error(nullptr);
        throw msg.c_str();
    }
    fMaxStep = tdbem_max_step(fCtx);
    fInit = false;
}

void SolverStepLinearT::Initialize() {}

void SolverStepLinearT::GetNodeLowHigh(int* low, int* high) const
{
    *low = 0;
    *high = fModel->GetNND() - 1;
}

void SolverStepLinearT::GetDofLowHigh(int* low, int* high) const
{
    // src/SolverT.cpp:178-179 (fDofHigh is inclusive)
    *low = 0;
    *high = 3 * (fModel->GetNND() - 1) + 2;
}

void SolverStepLinearT::UpdateGHLinear(const double* G_0, const double* G_1, const double* H_0,
                                       const double* H_1)
{
    // With the step's blocks (the rows of this solver's nodes): the reference's accumulation
    // into the lag matrices, on the device (tdbem_update_gh_linear).  Without them (null): the
    // lag matrices are assembled on the GPU in one pass by the first Solve.
    if (!G_0 || !G_1 || !H_0 || !H_1) return;
    int low, high;
    GetNodeLowHigh(&low, &high);
    check(fCtx, tdbem_update_gh_linear(fCtx, fCurrStep, low, high, G_0, G_1, H_0, H_1));
}

void SolverStepLinearT::Solve(double a1, double a2)
{
    if (!(a1 >= 0 && a1 <= 1 && a2 >= 0 && a2 < 1 && a1 + a2 <= 1)) fail("Alpha is out of range");
    // step 1 solves nothing (src/SolverStepLinearT.cpp:349-354): the march starts at step 2, when
    // lags 0 and 1 of supplied blocks are complete
    if (!fInit && fCurrStep >= 2) {
        int nu, nf;
        const int *ud, *fd;
        const double *uv, *fv;
        fModel->BoundaryCondition_Local(nu, &ud, &uv, nf, &fd, &fv);
        check(fCtx, tdbem_march_init(fCtx, a1, a2, nu, ud, uv, nf, fd, fv, 1, nullptr, nullptr));
        fInit = true;
    }
    // step n solves x_{n-1}
    while (fInit && tdbem_march_step_index(fCtx) < fCurrStep - 1)
        check(fCtx, tdbem_march_run(fCtx, 1));
}

void SolverStepLinearT::GetHistory(std::vector<double>& U, std::vector<double>& T)
{
    U.assign((size_t)fNumStep * fN, 0.0);
    T.assign((size_t)fNumStep * fN, 0.0);
    if (fInit) check(fCtx, tdbem_get_history(fCtx, 0, U.data(), T.data()));
}

void SolverStepLinearT::RecordResult(const char* path)
{
    // src/SolverStepLinearT.cpp:549-603: time, load, u[2], t[2] of x_{n-1}
    std::ofstream f;
    if (fCurrStep == 1) {
        f.open(path, std::ios::out | std::ios::trunc);
        f << std::setw(10) << "%time" << std::setw(15) << "load" << std::setw(15)
          << "displacement" << std::setw(15) << "traction" << std::endl;
    } else {
        f.open(path, std::ios::out | std::ios::app);
    }
    double u2 = 0, t2 = 0;
    if (fInit && fCurrStep >= 2) {
        std::vector<double> u(fN), t(fN);
        check(fCtx, tdbem_get_state(fCtx, 0, fCurrStep - 1, u.data(), t.data()));
        u2 = u[2];
        t2 = t[2];
    }
    const double time = fDt * (fCurrStep - 1);
    f << std::setprecision(6);
    f << std::setw(10) << time << std::setw(15) << LoadTimeVariation(time) << std::setw(15) << u2
      << std::setw(15) << t2 << std::endl;
}

void SolverStepLinearT::WriteGH(const char* name_prefix, int step, int num_row, int num_column,
                                const double* data, int, int)
{
    std::ostringstream name;
    name << name_prefix << step << ".txt";
    std::FILE* f = std::fopen(name.str().c_str(), "w");
    if (!f) fail("SolverT::WriteGH, cannot open the output file");
    for (int i = 0; i < num_row; i++) {
        for (int j = 0; j < num_column; j++) std::fprintf(f, "%15.5e", data[(size_t)i * fN + j]);
        std::fputc('\n', f);
    }
    std::fclose(f);
}

void SolverStepLinearT::WriteGH_Binary(const char* name_prefix, int step, int num_row,
                                       int num_column, const double* data)
{
    // full precision: "<prefix><step>.bin" = int64 rows, int64 columns, then the row-major
    // doubles (little-endian, as the host stores them)
    std::ostringstream name;
    name << name_prefix << step << ".bin";
    std::FILE* f = std::fopen(name.str().c_str(), "wb");
    if (!f) fail("SolverT::WriteGH, cannot open the output file");
    const long long dims[2] = {num_row, num_column};
    bool ok = std::fwrite(dims, sizeof(dims), 1, f) == 1;
    for (int i = 0; i < num_row && ok; i++)
        ok = std::fwrite(data + (size_t)i * fN, sizeof(double), num_column, f) == (size_t)num_column;
    if (std::fclose(f) != 0 || !ok) fail("SolverT::WriteGH, cannot write the output file");
}

// ------------------------------------------------------------------------------------------
// SolverMRT
// ------------------------------------------------------------------------------------------
namespace {

void check_mrt(int rc)
{
    if (rc != TDBEM_OK) {
        static thread_local std::string msg;
        msg = tdbem_last_error(nullptr);
        throw msg.c_str();
    }
}

}  // namespace

SolverMRT::SolverMRT(ModelManagerT* model_1, ModelManagerT* model_2) : fModel{model_1, model_2}
{
    // src/SolverMRT.cpp:22-60: dt and numstep come from model 1, fMaxStep starts at 10
    if (model_1->GetNumInterface() != model_2->GetNumInterface())
        fail("Interface mismatch, interface nodes must match one-to-one");
    SetTimeControl(model_1->GetDT(), model_1->GetNumStep());
}

SolverMRT::~SolverMRT()
{
    tdbem_mrt_destroy(fMrt);
    tdbem_destroy(fCtx[0]);
    tdbem_destroy(fCtx[1]);
}

void SolverMRT::SetTimeControl(double dt, int num_steps)
{
    fDt = dt;
    fNumStep = num_steps;
    tdbem_mrt_destroy(fMrt);
    fMrt = nullptr;
    for (int r = 0; r < 2; r++) {
        tdbem_destroy(fCtx[r]);
        fCtx[r] = nullptr;
        double E, nu, rho;
        fModel[r]->GetMaterialConstants(E, nu, rho);
        check_mrt(tdbem_create(fModel[r]->GetNND(), fModel[r]->GetCoords(), fModel[r]->GetNEL(),
                               fModel[r]->GetIEN(), fModel[r]->GetElementType(),
                               fModel[r]->GetNQ(), E, nu, rho, fModel[r]->GetIfExt(), dt,
                               num_steps, TDBEM_KERNEL_STEP, &fCtx[r]));
    }
    check_mrt(tdbem_mrt_create(fCtx[0], fCtx[1], fMaxStep, fModel[0]->GetNumInterface(),
                               fModel[0]->GetInterface(), fModel[1]->GetInterface(), &fMrt));
    fScheme = -1;
}

void SolverMRT::SetMaxStep(int max_step)
{
    if (fScheme >= 0) fail("SetMaxStep after the first solve");
    fMaxStep = max_step;
    SetTimeControl(fDt, fNumStep);
}

void SolverMRT::GetNodeLowHigh(int* low_1, int* high_1, int* low_2, int* high_2) const
{
    *low_1 = 0;
    *high_1 = fModel[0]->GetNND() - 1;
    *low_2 = 0;
    *high_2 = fModel[1]->GetNND() - 1;
}

void SolverMRT::GetDofLowHigh(int* low_1, int* high_1, int* low_2, int* high_2) const
{
    *low_1 = 0;
    *high_1 = 3 * fModel[0]->GetNND() - 1;
    *low_2 = 0;
    *high_2 = 3 * fModel[1]->GetNND() - 1;
}

void SolverMRT::Advance(int scheme, int target)
{
    if (fScheme < 0) {
        int nu[2], nf[2];
        const int *ud[2], *fd[2];
        const double *uv[2], *fv[2];
        for (int r = 0; r < 2; r++)
            fModel[r]->BoundaryCondition_Global(nu[r], &ud[r], &uv[r], nf[r], &fd[r], &fv[r]);
        check_mrt(tdbem_mrt_init(fMrt, scheme, TDBEM_MRT_LOAD_PULSE, 1.0, 0.0, nf[0], fd[0], fv[0],
                                 nf[1], fd[1], fv[1]));
        fScheme = scheme;
    } else if (fScheme != scheme) {
        fail("Solve_Direct and Solve_Averaging cannot be mixed in one run");
    }
    while (tdbem_mrt_step_index(fMrt) < target) check_mrt(tdbem_mrt_run(fMrt, 1));
}

// src/SolverMRT.cpp:1035-1099: step fCurrStep solves x_{fCurrStep}
void SolverMRT::Solve_Direct() { Advance(TDBEM_MRT_DIRECT, fCurrStep); }

// src/SolverMRT.cpp:1101-1249: step fCurrStep solves x_{fCurrStep-1}
void SolverMRT::Solve_Averaging() { Advance(TDBEM_MRT_AVERAGING, fCurrStep - 1); }

void SolverMRT::GetField(int region, int step, std::vector<double>& u)
{
    if (region < 0 || region > 1) fail("region must be 0 or 1");
    const int n = tdbem_mrt_num_unknowns(fMrt);
    std::vector<double> all(n, 0.0);
    if (fScheme >= 0) check_mrt(tdbem_mrt_get_state(fMrt, step, nullptr, all.data(), nullptr));
    const int n0 = 3 * fModel[0]->GetNND();
    if (region == 0)
        u.assign(all.begin(), all.begin() + n0);
    else
        u.assign(all.begin() + n0, all.end());
}

void SolverMRT::RecordResult(const char* path, int record_dof, bool vtk)
{
    // src/SolverMRT.cpp:1325-1383: region 2's dof of step fCurrStep-1 at time dt (fCurrStep-1)
    std::ofstream f;
    if (fCurrStep == 1)
        f.open(path, std::ios::out | std::ios::trunc);
    else
        f.open(path, std::ios::out | std::ios::app);
    std::vector<double> u2;
    GetField(1, fCurrStep - 1, u2);
    if (record_dof < 0 || record_dof >= (int)u2.size()) fail("record dof out of range");
    f << std::setprecision(6);
    f << std::setw(10) << fDt * (fCurrStep - 1) << std::setw(15) << u2[record_dof] << std::endl;
    if (vtk) {
        OutPutVTK("BEM_1st_", 0);
        OutPutVTK("BEM_2nd_", 1);
    }
}

void SolverMRT::OutPutVTK(std::string filename, int region)
{
    // src/SolverMRT.cpp:1386-1466: legacy VTK, quads (cell type 9), displacement of fCurrStep-1
    ModelManagerT* m = fModel[region];
    std::vector<double> u;
    GetField(region, fCurrStep - 1, u);
    filename += std::to_string(fCurrStep) + ".vtk";
    std::ofstream f(filename, std::ios::out | std::ios::trunc);
    const int nnd = m->GetNND(), nel = m->GetNEL();
    const int* ien = m->GetIEN();
    const double* X = m->GetCoords();
    f << "# vtk DataFile Version 3.0 " << std::endl;
    f << "Boundary element result" << std::endl;
    f << "ASCII" << std::endl;
    f << "DATASET UNSTRUCTURED_GRID " << std::endl;
    f << "POINTS " << nnd << " float" << std::endl;
    f << std::setprecision(6);
    for (int n = 0; n < nnd; n++)
        f << std::setw(15) << X[3 * n] << std::setw(15) << X[3 * n + 1] << std::setw(15)
          << X[3 * n + 2] << std::endl;
    f << "CELLS  " << nel << " " << nel * 5 << std::endl;
    const int ennd = m->GetENND();
    for (int e = 0; e < nel; e++)
        f << std::setw(5) << 4 << std::setw(5) << ien[ennd * e] << std::setw(5)
          << ien[ennd * e + 1] << std::setw(5) << ien[ennd * e + 2] << std::setw(5)
          << ien[ennd * e + 3] << std::endl;
    f << "CELL_TYPES " << nel << std::endl;
    for (int e = 0; e < nel; e++) f << 9 << std::endl;
    f << "POINT_DATA " << nnd << std::endl;
    f << "VECTORS " << "Displacement " << "double" << std::endl;
    for (int n = 0; n < nnd; n++)
        f << std::setw(15) << u[3 * n] << std::setw(15) << u[3 * n + 1] << std::setw(15)
          << u[3 * n + 2] << std::endl;
}

}  // namespace TD_BEM
