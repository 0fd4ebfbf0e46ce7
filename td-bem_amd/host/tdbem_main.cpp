// tdbem_main.cpp — command-line driver with the step loop of src/main.cpp:22-149 on the GPU.
//
//   tdbem_main input.txt [--write-gh] [--write-gh-binary] [--history out.bin] [--caller-lags]
//   tdbem_main region1.txt --mrt region2.txt [--averaging] [--max-step K] [--no-vtk]
//
// Reads the deck and alpha.txt (a1 a2) from the working directory like the reference, marches
// numstep steps, and writes displacement.txt in the reference's format.  --write-gh also writes
// the FE-BE matrices G1_step<k>.txt, G2_step<k>.txt, H1_step<k>.txt, H2_step<k>.txt for every
// assembled step (the calls commented out at main.cpp:113-116, with their G1/G2 naming: file G1
// holds fG_2 = phi_2 part, file G2 holds fG_1); --write-gh-binary writes the same matrices at full
// precision as G1_step<k>.bin ... (int64 rows, int64 columns, row-major doubles).  --history writes the full-precision u and t
// histories as raw doubles [2][numstep][3 nnd].  --mrt runs the two-region solver SolverMRT
// (src/SolverMRT.cpp; the reference ships no driver for it) on the two decks with the step loop
// UpdateGHLinear / SetLoad / Solve / RecordResult: Solve_Direct as SolverMRT::Solve does, or
// Solve_Averaging with --averaging; fMaxStep 10 unless --max-step; displacement.txt gets the
// reference's recorded region-2 dof and BEM_1st_<n>.vtk / BEM_2nd_<n>.vtk the region fields
// (not with --no-vtk).  --caller-lags runs main.cpp's own data flow: G_H_Driver assembles each
// step's blocks (ElastoDynamicsT) and UpdateGHLinear hands them to the solver, which accumulates
// the lag matrices from them (tdbem_update_gh_linear) instead of assembling its store itself.
// Unlike the reference, errors exit non-zero.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <vector>

#include "tdbem_host.h"

using namespace TD_BEM;

int main(int argc, char** argv)
{
    if (argc <= 1) {
        std::cerr << "Input file not specified, try ./TD_BEM input.txt" << std::endl;
        return 2;
    }
    bool write_gh = false, write_bin = false, averaging = false, vtk = true, caller_lags = false;
    const char* history = nullptr;
    const char* mrt = nullptr;
    int max_step = 10;
    for (int a = 2; a < argc; a++) {
        if (!std::strcmp(argv[a], "--write-gh"))
            write_gh = true;
        else if (!std::strcmp(argv[a], "--write-gh-binary"))
            write_bin = true;
        else if (!std::strcmp(argv[a], "--history") && a + 1 < argc)
            history = argv[++a];
        else if (!std::strcmp(argv[a], "--mrt") && a + 1 < argc)
            mrt = argv[++a];
        else if (!std::strcmp(argv[a], "--averaging"))
            averaging = true;
        else if (!std::strcmp(argv[a], "--caller-lags"))
            caller_lags = true;
        else if (!std::strcmp(argv[a], "--no-vtk"))
            vtk = false;
        else if (!std::strcmp(argv[a], "--max-step") && a + 1 < argc)
            max_step = std::atoi(argv[++a]);
        else {
            std::cerr << "unknown option " << argv[a] << std::endl;
            return 2;
        }
    }
    if (mrt) {
        try {
            ModelManagerT model_1, model_2;
            model_1.ReadInput(argv[1]);
            model_2.ReadInput(mrt);
            SolverMRT solver(&model_1, &model_2);
            if (max_step != 10) solver.SetMaxStep(max_step);
            const int num_step = model_1.GetNumStep();
            for (int curr_step = 1; curr_step <= num_step; curr_step++) {
                solver.SetCurrStep(curr_step);
                solver.UpdateGHLinear(nullptr, nullptr, nullptr, nullptr, nullptr, nullptr,
                                      nullptr, nullptr);
                solver.SetLoad();
                if (averaging)
                    solver.Solve_Averaging();
                else
                    solver.Solve();
                solver.RecordResult("displacement.txt", 46 * 3 - 1, vtk);
            }
        } catch (const char* msg) {
            std::cerr << msg << std::endl;
            return 1;
        }
        std::cout << "The computation terminated properly" << std::endl;
        return 0;
    }
    try {
        ModelManagerT model;
        model.ReadInput(argv[1]);
        SolverStepLinearT solver(&model);
        int dof_low, dof_high;
        solver.GetDofLowHigh(&dof_low, &dof_high);
        model.Distributed_BC(dof_low, dof_high);
        const int num_step = model.GetNumStep();
        const double dt = model.GetDT();
        solver.Initialize();

        double a1, a2;
        std::ifstream alpha("alpha.txt");
        if (!(alpha >> a1 >> a2)) throw "alpha.txt not found or unreadable (expected: a1 a2)";
        std::cout << "***********************************\n"
                  << "alpha_1=" << a1 << "  alpha_2=" << a2 << "\n"
                  << "***********************************" << std::endl;

        ElastoDynamicsT* ED =
            (write_gh || write_bin || caller_lags) ? new ElastoDynamicsT(&model) : nullptr;
        const int max_step = solver.GetMaxStep();
        for (int curr_step = 1; curr_step <= num_step; curr_step++) {
            double *G_0 = nullptr, *G_1 = nullptr, *H_0 = nullptr, *H_1 = nullptr;
            if (ED && curr_step <= max_step) {
                ED->G_H_Driver(curr_step * dt, dt);
                ED->GetGHLinear(&G_0, &G_1, &H_0, &H_1);
                const int rows = dof_high - dof_low + 1, cols = 3 * model.GetNND();
                if (write_gh) {
                    solver.WriteGH("G1_step", curr_step - 1, rows, cols, G_1);
                    solver.WriteGH("G2_step", curr_step - 1, rows, cols, G_0);
                    solver.WriteGH("H1_step", curr_step - 1, rows, cols, H_1);
                    solver.WriteGH("H2_step", curr_step - 1, rows, cols, H_0);
                }
                if (write_bin) {
                    solver.WriteGH_Binary("G1_step", curr_step - 1, rows, cols, G_1);
                    solver.WriteGH_Binary("G2_step", curr_step - 1, rows, cols, G_0);
                    solver.WriteGH_Binary("H1_step", curr_step - 1, rows, cols, H_1);
                    solver.WriteGH_Binary("H2_step", curr_step - 1, rows, cols, H_0);
                }
            }
            solver.SetCurrStep(curr_step);
            if (caller_lags)
                solver.UpdateGHLinear(G_0, G_1, H_0, H_1);   // main.cpp:121
            else
                solver.UpdateGHLinear(nullptr, nullptr, nullptr, nullptr);
            solver.SetLoad();
            solver.Solve(a1, a2);
            solver.RecordResult();
        }
        delete ED;
        if (history) {
            std::vector<double> U, T;
            solver.GetHistory(U, T);
            std::FILE* f = std::fopen(history, "wb");
            if (!f) throw "cannot open the history file";
            std::fwrite(U.data(), sizeof(double), U.size(), f);
            std::fwrite(T.data(), sizeof(double), T.size(), f);
            std::fclose(f);
        }
    } catch (const char* msg) {
        std::cerr << msg << std::endl;
        return 1;
    }
    std::cout << "The computation terminated properly" << std::endl;
    return 0;
}
