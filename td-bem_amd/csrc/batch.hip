// batch.hip — many independent time-marching scenarios (uneven DOF and step counts) stepped
// together.  Each scenario is a context with a mixed-precision march; one step of every scenario
// is six launches in total, whatever the number of scenarios:
//   advance + load factors, FP32 history copy, MFMA convolution, reduction into the
//   right-hand sides, step solve, finalize,
// each launch covering all scenarios through a table of per-scenario descriptors and a device
// step counter per scenario (scenarios past their last step drop out).  A run captures
// `steps_per_graph` steps of this sequence into one hipGraph and replays it, so one graph launch
// advances every scenario by that many steps.  The convolution work of all scenarios is cut into
// one work-balanced chunk table (one workgroup per chunk); the number of workgroup rounds and the
// width of the step solve's column splits are picked by batch_tune, which times the candidates on
// the batch itself.
#include <algorithm>
#include <chrono>
#include <map>
#include <memory>
#include <stdexcept>
#include <vector>

#include "step_device.h"

namespace tdbem {

namespace {

constexpr int kThreads = 256;

// scenario of block b: the last sc with off[sc] <= b (off[0] = 0, off[S] = total blocks)
__device__ __forceinline__ int scen_of(const int* __restrict__ off, int S, int b)
{
    int lo = 0, hi = S;
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (off[mid] <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Advances every scenario with steps left (s + 1 < numstep) and forms its load factors.
__global__ void k_b_step(const BatchScen* __restrict__ D, int* __restrict__ sdev,
                         int* __restrict__ act)
{
    const int sc = blockIdx.x;
    const BatchScen& d = D[sc];
    __shared__ int s_new, a;
    if (threadIdx.x == 0) {
        int s = sdev[sc];
        a = s + 1 < d.numstep;
        if (a) {
            s += 1;
            sdev[sc] = s;
        }
        act[sc] = a;
        s_new = s;
    }
    __syncthreads();
    const int r = threadIdx.x;
    if (a && r < d.nrhs) d.fac[r] = load_factor(d.amp[r], d.rise[r], s_new * d.dt);
}

// The three ring slots s - 2 .. s changed by the previous step, in FP32 operand order.
__global__ void k_b_zcopy(const BatchScen* __restrict__ D, const int* __restrict__ off, int S,
                           const int* __restrict__ sdev, const int* __restrict__ act)
{
    const int sc = scen_of(off, S, blockIdx.x);
    if (!act[sc]) return;
    const BatchScen& d = D[sc];
    const int64_t t = (int64_t)(blockIdx.x - off[sc]) * kThreads + threadIdx.x;
    const int64_t per = (int64_t)d.nT * kCols * d.nrhs;
    if (t >= 3 * per) return;
    const int m = (int)(t / per);
    int slot = (sdev[sc] - 2 + m) % d.R;
    if (slot < 0) slot += d.R;
    zcopy_elem(d.ring, d.N, d.nrhs, d.nT, d.perm, slot, d.nrhs / 16,
               reinterpret_cast<unsigned short*>(d.zs), t - m * per);
}

// One workgroup per (chunk slot, right-hand-side tile), one wavefront per row tile of the
// chunk's group (conv_chunk).  A member with n0 chunks of lag parity 0 and n1 of parity 1 has
// max(n0, n1) slots (slot_off: first slot of each member); at step s slot j of the member runs
// its chunk j of parity s mod 2, if there is one, so the grid holds no workgroup of the parity
// that does not run.
__global__ void __launch_bounds__(64 * kGroup, 2)
    k_b_conv(const BatchScen* __restrict__ D, const int* __restrict__ slot_off, int S,
             const int4* __restrict__ chunk, int NT, const int* __restrict__ sdev,
             const int* __restrict__ act, double* __restrict__ partial)
{
    const int L = xcd_chunk(blockIdx.x, gridDim.x);
    const int slot = L / NT, nt0 = L % NT;
    const int sc = scen_of(slot_off, S, slot);
    if (!act[sc]) return;
    const BatchScen& d = D[sc];
    const int s = sdev[sc], p = s & 1;
    const int j = slot - slot_off[sc];
    if (j >= (p ? d.nch1 : d.nch0)) return;
    const int c = d.ch0 + (p ? d.nch0 : 0) + j;
    const int4 cw = chunk[c];
    const int w = threadIdx.x >> 6;
    const int g = cw.x & ((1 << kChunkParityShift) - 1);
    double* out = partial + ((int64_t)c * 2 * kGroup + w) * 3 * NT * 256;
    conv_chunk(reinterpret_cast<const f32x4*>(d.tp), d.ent, reinterpret_cast<const bf16x8*>(d.zs),
               d.nT, NT, nt0, d.R, s, cw.y, cw.z, w, g * kGroup + w < d.nT, out,
               out + (int64_t)kGroup * 3 * NT * 256);
}

// the reduction of the partial sums, written as the step solve's right-hand side
__global__ void k_b_reduce_rhs(const BatchScen* __restrict__ D, const int* __restrict__ off,
                               int S, const int* __restrict__ sdev, const int* __restrict__ act,
                               const double* __restrict__ partial, const int* __restrict__ gch)
{
    const int sc = scen_of(off, S, blockIdx.x);
    if (!act[sc]) return;
    const BatchScen& d = D[sc];
    const int64_t t = (int64_t)(blockIdx.x - off[sc]) * kThreads + threadIdx.x;
    if (t >= (int64_t)d.nT * kRows * d.nrhs) return;
    reduce_rhs_elem(partial, gch + d.gch0, d.nG, sdev[sc] & 1, d.nrhs / 16, d.perm, d.N, d.nrhs,
                    d.gF, d.gU, d.fac, d.carry, d.Xf, t);
}

template <int NT>
__global__ void __launch_bounds__(64 * kApplyWaves)
    k_b_apply(const BatchScen* __restrict__ D, const int* __restrict__ off, int S,
              const int* __restrict__ act)
{
    const int sc = scen_of(off, S, blockIdx.x);
    if (!act[sc]) return;
    const BatchScen& d = D[sc];
    const int g = (blockIdx.x - off[sc]) * kApplyWaves + (threadIdx.x >> 6);
    if (g < frag_rows(d.N) / 16 * d.nsplit)
        apply_wave<NT>(d.Af, d.N, d.Xf, d.nrhs, d.nsplit, g, d.parts);
}

__global__ void k_b_finalize(const BatchScen* __restrict__ D, const int* __restrict__ off, int S,
                             const int* __restrict__ sdev, const int* __restrict__ act)
{
    const int sc = scen_of(off, S, blockIdx.x);
    if (!act[sc]) return;
    const BatchScen& d = D[sc];
    const int64_t t = (int64_t)(blockIdx.x - off[sc]) * kThreads + threadIdx.x;
    if (t < (int64_t)d.N * d.nrhs)
        finalize_elem(d.parts, d.nsplit, d.Np, d.is_ubc, d.is_fbc, d.ubc_val, d.fbc_val, d.fac,
                      d.N, d.nrhs, sdev[sc], d.R, d.a1, d.a2, d.a3, d.U, d.T, d.ring, t);
}

int64_t ceil_div(int64_t a, int64_t b) { return (a + b - 1) / b; }

}  // namespace

struct Batch {
    std::vector<Context*> ctx;   // members, in the order they joined
    int S = 0, nrhs = 0, device = -1;
    hipStream_t stream = nullptr;
    bool dirty = true;           // membership changed since the last plan
    std::vector<BatchScen> h_desc;
    DBuf<BatchScen> desc;
    DBuf<int> sdev, act;
    std::map<Context*, std::unique_ptr<DBuf<double>>> parts;
    DBuf<int> off_z, off_r, off_rhs, off_a;
    int nb_z = 0, nb_r = 0, nb_rhs = 0, nb_a = 0;
    DBuf<int4> chunk;
    DBuf<int> slot_off;          // [S + 1] first convolution slot of each member
    int nslots = 0;
    DBuf<int> gch;
    DBuf<double> partial;
    int nchunks = 0;
    int rounds = 4, apply_waves = 4096;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int graph_steps = 0;

    ~Batch()
    {
        drop_graph();
        if (stream) {
            (void)hipStreamSynchronize(stream);
            (void)hipStreamDestroy(stream);
        }
        for (Context* c : ctx) c->batch_refs--;
    }
    void drop_graph()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_steps = 0;
    }
};

namespace {

// chunk table, partial sums, step-solve splits and block tables for the current tunables
void batch_plan(Batch& B)
{
    B.drop_graph();
    if (B.S == 0) return;
    const int NT = B.nrhs / 16;
    std::vector<int4> ch;
    std::vector<int> gch;
    std::vector<int> oz(B.S + 1, 0), orr(B.S + 1, 0), orhs(B.S + 1, 0), oa(B.S + 1, 0),
        os(B.S + 1, 0);
    for (int sc = 0; sc < B.S; sc++) {
        Context& c = *B.ctx[sc];
        BatchScen& d = B.h_desc[sc];
        d.gch0 = (int)gch.size();
        // each scenario is cut as it would be alone (its results do not depend on the mix);
        // its chunks of parity 0 come first, then those of parity 1
        d.ch0 = (int)ch.size();
        tp_chunks(c, tp_chunk_target(c, B.rounds), sc, ch, gch);
        d.nch0 = 0;
        for (size_t q = d.ch0; q < ch.size(); q++) d.nch0 += ((ch[q].x >> kChunkParityShift) & 1) == 0;
        d.nch1 = (int)ch.size() - d.ch0 - d.nch0;
        os[sc + 1] = os[sc] + std::max(d.nch0, d.nch1);
        d.nsplit = apply_splits(d.N, B.apply_waves);
        const size_t need = (size_t)d.nsplit * d.nrhs * d.Np;
        auto& pb = B.parts[&c];
        if (!pb) pb.reset(new DBuf<double>());
        pb->alloc(need);
        d.parts = pb->p;
        oz[sc + 1] = oz[sc] + (int)ceil_div(3LL * d.nT * kCols * d.nrhs, kThreads);
        orr[sc + 1] = orr[sc] + (int)ceil_div((int64_t)d.nT * kRows * d.nrhs, kThreads);
        orhs[sc + 1] = orhs[sc] + (int)ceil_div((int64_t)d.N * d.nrhs, kThreads);
        oa[sc + 1] = oa[sc] + (int)ceil_div((int64_t)frag_rows(d.N) / 16 * d.nsplit, kApplyWaves);
    }
    B.nchunks = (int)ch.size();
    B.chunk.upload(ch.data(), ch.size(), B.stream);
    B.slot_off.upload(os.data(), os.size(), B.stream);
    B.nslots = os[B.S];
    B.gch.upload(gch.data(), gch.size(), B.stream);
    B.partial.alloc((size_t)B.nchunks * 2 * kGroup * 3 * NT * 256);
    B.off_z.upload(oz.data(), oz.size(), B.stream);
    B.off_r.upload(orr.data(), orr.size(), B.stream);
    B.off_rhs.upload(orhs.data(), orhs.size(), B.stream);
    B.off_a.upload(oa.data(), oa.size(), B.stream);
    B.nb_z = oz[B.S];
    B.nb_r = orr[B.S];
    B.nb_rhs = orhs[B.S];
    B.nb_a = oa[B.S];
    B.desc.upload(B.h_desc.data(), B.h_desc.size(), B.stream);
    TD_HIP(hipStreamSynchronize(B.stream));
}

void launch_conv(Batch& B)
{
    const int NT = B.nrhs / 16;
    if (NT < 1 || NT > 4 || B.nrhs % 16)
        throw std::invalid_argument("batch: nrhs must be 16, 32, 48 or 64");
    if (B.nslots == 0) return;
    hipLaunchKernelGGL(k_b_conv, dim3((unsigned)(B.nslots * NT)), dim3(64 * kGroup), 0,
                       B.stream, B.desc.p, B.slot_off.p, B.S, B.chunk.p, NT, B.sdev.p, B.act.p,
                       B.partial.p);
    TD_HIP(hipGetLastError());
}

void launch_apply(Batch& B)
{
    const dim3 g((unsigned)B.nb_a), b(64 * kApplyWaves);
    switch (B.nrhs / 16) {
    case 1: hipLaunchKernelGGL(k_b_apply<1>, g, b, 0, B.stream, B.desc.p, B.off_a.p, B.S, B.act.p); break;
    case 2: hipLaunchKernelGGL(k_b_apply<2>, g, b, 0, B.stream, B.desc.p, B.off_a.p, B.S, B.act.p); break;
    case 3: hipLaunchKernelGGL(k_b_apply<3>, g, b, 0, B.stream, B.desc.p, B.off_a.p, B.S, B.act.p); break;
    case 4: hipLaunchKernelGGL(k_b_apply<4>, g, b, 0, B.stream, B.desc.p, B.off_a.p, B.S, B.act.p); break;
    default: throw std::invalid_argument("batch: nrhs must be 16, 32, 48 or 64");
    }
    TD_HIP(hipGetLastError());
}

// one step of every scenario with steps left
void launch_step(Batch& B)
{
    hipLaunchKernelGGL(k_b_step, dim3(B.S), dim3(64), 0, B.stream, B.desc.p, B.sdev.p, B.act.p);
    hipLaunchKernelGGL(k_b_zcopy, dim3(B.nb_z), dim3(kThreads), 0, B.stream, B.desc.p,
                       B.off_z.p, B.S, B.sdev.p, B.act.p);
    launch_conv(B);
    hipLaunchKernelGGL(k_b_reduce_rhs, dim3(B.nb_r), dim3(kThreads), 0, B.stream, B.desc.p,
                       B.off_r.p, B.S, B.sdev.p, B.act.p, B.partial.p, B.gch.p);
    launch_apply(B);
    hipLaunchKernelGGL(k_b_finalize, dim3(B.nb_rhs), dim3(kThreads), 0, B.stream, B.desc.p,
                       B.off_rhs.p, B.S, B.sdev.p, B.act.p);
    TD_HIP(hipGetLastError());
}

// device step counters and activity flags from the contexts' host step indices
void batch_sync_state(Batch& B)
{
    std::vector<int> s(B.S), a(B.S);
    for (int sc = 0; sc < B.S; sc++) {
        s[sc] = march_step_count(*B.ctx[sc]);
        a[sc] = s[sc] + 1 < B.ctx[sc]->numstep;
    }
    TD_HIP(hipMemcpyAsync(B.sdev.p, s.data(), B.S * sizeof(int), hipMemcpyHostToDevice, B.stream));
    TD_HIP(hipMemcpyAsync(B.act.p, a.data(), B.S * sizeof(int), hipMemcpyHostToDevice, B.stream));
    TD_HIP(hipStreamSynchronize(B.stream));
}

template <typename F>
double time_launches(Batch& B, F&& launch, int reps)
{
    launch();  // warm-up
    hipEvent_t e0, e1;
    TD_HIP(hipEventCreate(&e0));
    TD_HIP(hipEventCreate(&e1));
    TD_HIP(hipEventRecord(e0, B.stream));
    for (int r = 0; r < reps; r++) launch();
    TD_HIP(hipEventRecord(e1, B.stream));
    TD_HIP(hipEventSynchronize(e1));
    float ms = 0;
    TD_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return ms / reps;
}

}  // namespace

namespace {

// Descriptors, device step tables and the plan for the current members, after a change of
// membership (a graph captured for the previous members is dropped by batch_plan).
void batch_refresh(Batch& B)
{
    if (!B.dirty) return;
    B.S = (int)B.ctx.size();
    B.h_desc.assign(B.S, BatchScen{});
    for (int sc = 0; sc < B.S; sc++) {
        Context& c = *B.ctx[sc];
        BatchScen& d = B.h_desc[sc];
        march_describe(c, d);
        d.tp = c.tp.vals.p;
        d.ent = c.tp.ent.p;
        d.perm = c.tp.d_perm.p;
        d.nT = c.tp.nT;
        d.nG = c.tp.nG;
    }
    for (auto it = B.parts.begin(); it != B.parts.end();) {
        if (std::find(B.ctx.begin(), B.ctx.end(), it->first) == B.ctx.end())
            it = B.parts.erase(it);
        else
            ++it;
    }
    B.sdev.reserve(std::max(B.S, 1));
    B.act.reserve(std::max(B.S, 1));
    batch_plan(B);
    B.dirty = false;
}

}  // namespace

void batch_add(Batch& B, Context* c)
{
    if (!c) throw std::invalid_argument("batch: null context");
    if (std::find(B.ctx.begin(), B.ctx.end(), c) != B.ctx.end())
        throw std::invalid_argument("batch: the context is already in the batch");
    if (c->device != B.device) throw std::invalid_argument("batch: contexts on different GPUs");
    BatchScen d{};
    march_describe(*c, d);
    if (B.nrhs != 0 && d.nrhs != B.nrhs)
        throw std::invalid_argument("batch: contexts differ in nrhs");
    TD_HIP(hipStreamSynchronize(c->stream));
    // step 1 (the n == 2 solve) of a scenario that has not started, on its own stream
    if (march_step_count(*c) == 0) march_run(*c, 1);
    B.nrhs = d.nrhs;
    B.ctx.push_back(c);
    c->batch_refs++;
    B.dirty = true;
}

void batch_remove(Batch& B, Context* c)
{
    auto it = std::find(B.ctx.begin(), B.ctx.end(), c);
    if (it == B.ctx.end()) throw std::invalid_argument("batch: the context is not in the batch");
    TD_HIP(hipStreamSynchronize(B.stream));
    B.ctx.erase(it);
    c->batch_refs--;
    B.dirty = true;
}

int batch_size(const Batch& B) { return (int)B.ctx.size(); }

Batch* batch_create(const std::vector<Context*>& ctxs)
{
    std::unique_ptr<Batch> B(new Batch());
    if (ctxs.empty()) {
        TD_HIP(hipGetDevice(&B->device));
    } else {
        if (!ctxs[0]) throw std::invalid_argument("batch: null context");
        B->device = ctxs[0]->device;
    }
    TD_HIP(hipSetDevice(B->device));
    TD_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
    for (Context* c : ctxs) batch_add(*B, c);
    batch_refresh(*B);
    return B.release();
}

void batch_destroy(Batch* B) { delete B; }

void batch_tune(Batch& B, double* report)
{
    batch_refresh(B);
    if (B.S == 0) {
        if (report) report[0] = report[1] = report[2] = report[3] = 0.0;
        return;
    }
    batch_sync_state(B);
    const int rounds_c[] = {1, 2, 4, 8};
    const int waves_c[] = {1024, 2048, 4096, 8192};
    double best_conv = 1e300, best_apply = 1e300;
    int best_r = B.rounds, best_w = B.apply_waves;
    for (int r : rounds_c) {
        B.rounds = r;
        batch_plan(B);
        const double ms = time_launches(B, [&] { launch_conv(B); }, 2);
        if (ms < best_conv) {
            best_conv = ms;
            best_r = r;
        }
    }
    B.rounds = best_r;
    for (int w : waves_c) {
        B.apply_waves = w;
        batch_plan(B);
        const double ms = time_launches(B, [&] { launch_apply(B); }, 2);
        if (ms < best_apply) {
            best_apply = ms;
            best_w = w;
        }
    }
    B.apply_waves = best_w;
    batch_plan(B);
    if (report) {
        report[0] = best_r;
        report[1] = best_w;
        report[2] = best_conv;
        report[3] = best_apply;
    }
}

void batch_set_params(Batch& B, int rounds, int apply_waves)
{
    if (rounds < 1 || rounds > 64 || apply_waves < 64 || apply_waves > (1 << 20))
        throw std::invalid_argument("batch: tunables out of range");
    B.rounds = rounds;
    B.apply_waves = apply_waves;
    if (!B.dirty) batch_plan(B);   // otherwise the next refresh plans with them
}

int batch_run(Batch& B, int steps_per_graph, int max_replays)
{
    if (steps_per_graph < 1 || steps_per_graph > 4096)
        throw std::invalid_argument("batch: steps_per_graph out of range");
    batch_refresh(B);
    if (B.S == 0) return 0;
    batch_sync_state(B);
    int remaining = 0;
    for (Context* c : B.ctx)
        remaining = std::max(remaining, c->numstep - 1 - march_step_count(*c));
    if (remaining <= 0) return 0;
    if (!B.exec || B.graph_steps != steps_per_graph) {
        B.drop_graph();
        TD_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        try {
            for (int k = 0; k < steps_per_graph; k++) launch_step(B);
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(B.stream, &g);
            if (g) (void)hipGraphDestroy(g);
            throw;
        }
        TD_HIP(hipStreamEndCapture(B.stream, &B.graph));
        TD_HIP(hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
        B.graph_steps = steps_per_graph;
    }
    int replays = (int)ceil_div(remaining, steps_per_graph);
    if (max_replays > 0) replays = std::min(replays, max_replays);
    for (int r = 0; r < replays; r++) TD_HIP(hipGraphLaunch(B.exec, B.stream));
    TD_HIP(hipStreamSynchronize(B.stream));
    std::vector<int> s(B.S);
    TD_HIP(hipMemcpy(s.data(), B.sdev.p, B.S * sizeof(int), hipMemcpyDeviceToHost));
    for (int sc = 0; sc < B.S; sc++) march_set_step(*B.ctx[sc], s[sc]);
    return replays;
}

}  // namespace tdbem
