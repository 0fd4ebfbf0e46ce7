// tdbem_element.cpp — element matrices of (collocation point, element) pairs through the host
// classes TDElementStepLinearT / TDElementHeavisideLinearT / TDElementRampHeaviLinearT, the call
// sequence of src/ElastoDynamicsT.cpp:136-192 (SetNQ, SetMaterial, SetCoords, SetCollocation,
// SetTime, FormGH_Regular or FormGH_Singular, GetGHLinear) for a file of pairs.
//
//   tdbem_element pairs.bin out.bin
//
// pairs.bin: records of int32 kernel, etype, nq, singular; float64 E, nu, density, t, t_1,
// collocation[3], coords[8][3] (the first ennd nodes used).  out.bin: per record float64 GE_1,
// GE_2, HE_1, HE_2 as [3][24] (columns past 3 ennd zero), then C_1, C_2 as [3][3].
#include <cstdint>
#include <cstdio>
#include <iostream>
#include <map>
#include <memory>
#include <utility>
#include <vector>

#include "tdbem_host.h"

using namespace TD_BEM;

namespace {

struct Record {
    int32_t kernel, etype, nq, singular;
    double E, nu, density, t, t_1, coll[3], X[24];
};

std::unique_ptr<TDElementT> make_element(int kernel, int etype)
{
    switch (kernel) {
    case TDBEM_KERNEL_STEP: return std::make_unique<TDElementStepLinearT>(etype);
    case TDBEM_KERNEL_HEAVISIDE: return std::make_unique<TDElementHeavisideLinearT>(etype);
    case TDBEM_KERNEL_BSPLINE: return std::make_unique<TDElementRampHeaviLinearT>(etype);
    default: throw "tdbem_element: unsupported kernel";
    }
}

}  // namespace

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::cerr << "usage: tdbem_element pairs.bin out.bin" << std::endl;
        return 2;
    }
    std::FILE* in = std::fopen(argv[1], "rb");
    if (!in) {
        std::cerr << "cannot open " << argv[1] << std::endl;
        return 2;
    }
    std::vector<Record> recs;
    for (;;) {
        Record r;
        if (std::fread(&r.kernel, sizeof(int32_t), 4, in) != 4) break;
        if (std::fread(&r.E, sizeof(double), 5 + 3 + 24, in) != 32) {
            std::cerr << "truncated record" << std::endl;
            return 2;
        }
        recs.push_back(r);
    }
    std::fclose(in);
    std::vector<double> out;
    out.reserve(recs.size() * 306);
    try {
        // one element object per class and type, reused across pairs as G_H_Driver reuses its
        // fElement
        std::map<std::pair<int, int>, std::unique_ptr<TDElementT>> elems;
        for (Record& r : recs) {
            auto& el = elems[{r.kernel, r.etype}];
            if (!el) el = make_element(r.kernel, r.etype);
            const int ennd = r.etype == 1 ? 4 : 8;
            el->SetNQ(r.nq);
            el->SetMaterial(r.E, r.nu, r.density);
            el->SetCoords(r.X);
            el->SetCollocation(r.coll);
            el->SetTime(r.t, r.t_1);
            if (r.singular < 0)
                el->FormGH_Regular();
            else
                el->FormGH_Singular(r.singular);
            const double *g1, *g2, *h1, *h2, *c1, *c2;
            el->GetGHLinear(&g1, &g2, &h1, &h2, &c1, &c2);
            for (const double* m : {g1, g2, h1, h2})
                for (int i = 0; i < 3; i++)
                    for (int c = 0; c < 24; c++) out.push_back(c < 3 * ennd ? m[i * 3 * ennd + c] : 0.0);
            out.insert(out.end(), c1, c1 + 9);
            out.insert(out.end(), c2, c2 + 9);
        }
    } catch (const char* msg) {
        std::cerr << msg << std::endl;
        return 1;
    }
    std::FILE* f = std::fopen(argv[2], "wb");
    if (!f) {
        std::cerr << "cannot open " << argv[2] << std::endl;
        return 2;
    }
    std::fwrite(out.data(), sizeof(double), out.size(), f);
    std::fclose(f);
    return 0;
}
