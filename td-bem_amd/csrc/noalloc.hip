// noalloc.hip — the allocation-free variant of the per-step dense assembly
// (tdbem_assemble_step_device_noalloc): the element slots of each row batch live in a file-scope
// device array of fixed size instead of a buffer allocated per call, for callers that cannot
// allocate (code under hipGraph capture, tight serving loops).  Each GPU the library runs on
// holds its own copy of the array until the process ends; calls on one GPU share it, so calls
// that overlap on one GPU (contexts on different streams) race, and the array keeps the last
// call's contents (nothing in it needs to start at zero: every slot is written before it is read).
#include <mutex>
#include <stdexcept>
#include <vector>

#include "tdbem_internal.h"

namespace tdbem {

namespace {

constexpr int64_t kNoAllocBytes = 64LL << 20;   // 64 MiB per GPU

}  // namespace

// The workspace.  In host code this name is only a stand-in for the symbol: its device address
// (one per GPU) comes from hipGetSymbolAddress, and only that address is passed to kernels.
static __device__ __align__(256) double scratch[kNoAllocBytes / sizeof(double)];

namespace {

std::mutex g_mu;
std::vector<double*> g_addr;   // device address of `scratch`, per device

// The array's address on `device`, looked up once per device with that device current.
double* scratch_on(int device)
{
    std::lock_guard<std::mutex> lock(g_mu);
    if ((int)g_addr.size() <= device) g_addr.resize(device + 1, nullptr);
    if (!g_addr[device]) {
        int current = 0;
        TD_HIP(hipGetDevice(&current));
        if (current != device) TD_HIP(hipSetDevice(device));
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (current != device) (void)hipSetDevice(current);
        TD_HIP(e);
        if (!p) throw std::runtime_error("HIP error: no device address for the workspace");
        g_addr[device] = static_cast<double*>(p);
    }
    return g_addr[device];
}

}  // namespace

int64_t noalloc_scratch_bytes() { return kNoAllocBytes; }

void noalloc_prepare(int device) { (void)scratch_on(device); }

void assemble_step_dense_noalloc(Context& c, int step, int lower, int upper, double* dG1,
                                 double* dG2, double* dH1, double* dH2)
{
    const int64_t per_row = (int64_t)c.nel * 2 * c.ennd * kVals * (int64_t)sizeof(double);
    if (per_row > kNoAllocBytes)
        throw std::invalid_argument(
            "tdbem_assemble_step_device_noalloc: one collocation row needs " +
            std::to_string(per_row) + " bytes of element slots, more than the " +
            std::to_string(kNoAllocBytes) + "-byte workspace (at most " +
            std::to_string(kNoAllocBytes / (2 * c.ennd * kVals * (int64_t)sizeof(double))) +
            " elements)");
    assemble_step_dense(c, step, lower, upper, dG1, dG2, dH1, dH2, scratch_on(c.device),
                        kNoAllocBytes / (int64_t)sizeof(double));
}

}  // namespace tdbem
