// capi.hip — the C ABI of include/tdbem.h over the HIP implementation.
#include <cmath>
#include <memory>
#include <cstring>
#include <mutex>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/tdbem.h"
#include "tdbem_internal.h"

struct tdbem_context {
    tdbem::Context c;
};

struct tdbem_batch {
    tdbem::Batch* b = nullptr;
    ~tdbem_batch() { tdbem::batch_destroy(b); }
};

struct tdbem_element {
    tdbem::Element e;
};

struct tdbem_mrt {
    tdbem::Mrt* m = nullptr;
    ~tdbem_mrt() { tdbem::mrt_destroy(m); }
};

namespace tdbem {

void check_hip(hipError_t e, const char* what)
{
    if (e != hipSuccess)
        throw std::runtime_error(std::string("HIP error: ") + hipGetErrorString(e) + " in " + what);
}

Material make_material(double E, double nu, double rho)
{
    Material m;
    m.E = E;
    m.nu = nu;
    m.rho = rho;
    m.mu = E / (2 * (1 + nu));
    const double lam = E * nu / ((1 + nu) * (1 - 2 * nu));
    m.cp = std::sqrt((lam + 2 * m.mu) / rho);
    m.cs = std::sqrt(m.mu / rho);
    m.inv_cp = 1.0 / m.cp;
    m.inv_cs = 1.0 / m.cs;
    m.inv_cp2 = 1.0 / (m.cp * m.cp);
    m.inv_cs2 = 1.0 / (m.cs * m.cs);
    m.k2 = (m.cs / m.cp) * (m.cs / m.cp);
    m.k3 = m.k2 * (m.cs / m.cp);
    m.inv_4pi_rho = 1.0 / (4 * kPi * rho);
    return m;
}

Element::~Element()
{
    gauss.reset();
    in.reset();
    out.reset();
    sing.reset();
    if (stream) (void)hipStreamDestroy(stream);
}

Context::~Context()
{
    // the stream-ordered buffers go before the stream they are ordered on
    march_free(*this);
    tp.d_sub_id.reset();
    if (stream) (void)hipStreamDestroy(stream);
}

namespace {

thread_local std::string g_create_error;

// Stream-ordered temporaries (DBuf with a stream) come from the device's default memory pool;
// it keeps up to 16 GiB of freed memory between uses instead of returning it at every
// synchronisation, so the next set-up reuses it without mapping memory again.
void keep_pool_memory(int device)
{
    static std::mutex mu;
    static std::set<int> done;
    std::lock_guard<std::mutex> lock(mu);
    if (!done.insert(device).second) return;
    hipMemPool_t pool = nullptr;
    TD_HIP(hipDeviceGetDefaultMemPool(&pool, device));
    uint64_t keep = 16ull << 30;
    TD_HIP(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep));
}

// Gauss tables of the shape-function quadrature (src/ShapeFunctionT.cpp:83-163).  The 10- and
// 20-point rows carry the reference's 9 digits, and the 19th 20-point abscissa keeps the
// reference's sign (+0.993128599), so results match the reference for every nq it accepts.
const double kG6x[6] = {0.6612093864662645, -0.6612093864662645, -0.2386191860831969,
                        0.2386191860831969, -0.9324695142031521, 0.9324695142031521};
const double kG6w[6] = {0.3607615730481386, 0.3607615730481386, 0.4679139345726910,
                        0.4679139345726910, 0.1713244923791704, 0.1713244923791704};
const double kG10x[10] = {-0.148874339, 0.148874339, -0.433395394, 0.433395394, -0.679409568,
                          0.679409568, -0.865063367, 0.865063367, -0.973906529, 0.973906529};
const double kG10w[10] = {0.295524225, 0.295524225, 0.269266719, 0.269266719, 0.219086363,
                          0.219086363, 0.149451349, 0.149451349, 0.066671344, 0.066671344};
const double kG20x[20] = {-0.076526521, 0.076526521, -0.227785851, 0.227785851, -0.373706089,
                          0.373706089, -0.510867002, 0.510867002, -0.636053681, 0.636053681,
                          -0.746331906, 0.746331906, -0.839116972, 0.839116972, -0.912234428,
                          0.912234428, -0.963971927, 0.963971927, 0.993128599, 0.993128599};
const double kG20w[20] = {0.152753387, 0.152753387, 0.149172986, 0.149172986, 0.142096109,
                          0.142096109, 0.131688638, 0.131688638, 0.118194532, 0.118194532,
                          0.10193012, 0.10193012, 0.083276742, 0.083276742, 0.062672048,
                          0.062672048, 0.04060143, 0.04060143, 0.017614007, 0.017614007};
const double kG40x[40] = {
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
const double kG40w[40] = {
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

}  // namespace

bool gauss_table(int nq, const double** x, const double** w)
{
    switch (nq) {
    case 6: *x = kG6x; *w = kG6w; return true;
    case 10: *x = kG10x; *w = kG10w; return true;
    case 20: *x = kG20x; *w = kG20w; return true;
    case 40: *x = kG40x; *w = kG40w; return true;
    default: return false;
    }
}

namespace {

int max_step(const std::vector<double>& X, int nnd, double E, double nu, double rho, double dt)
{
    // SolverT::ComputeMaxStep (src/SolverT.cpp:211-250)
    const double Mu = E / (2.0 * (1.0 + nu));
    const double Cs = std::sqrt(Mu / rho);
    double max_dis = 0;
    for (int i = 0; i < nnd - 1; i++)
        for (int j = i + 1; j < nnd; j++) {
            double t = 0;
            for (int d = 0; d < 3; d++) {
                const double v = X[3 * i + d] - X[3 * j + d];
                t += v * v;
            }
            t = std::sqrt(t);
            if (t > max_dis) max_dis = t;
        }
    return (int)std::ceil(max_dis / (Cs * dt)) + 1;
}

// Refuses a call that would free or rebuild state a batch reads (TDBEM_ERR_STATE).
void not_batched(const Context& c, const char* what)
{
    if (c.batch_refs > 0)
        throw std::logic_error(std::string(what) + ": the context is in a batch");
}

template <typename F>
int guarded(tdbem_context* ctx, F&& f)
{
    if (!ctx) return TDBEM_ERR_ARG;
    try {
        f(ctx->c);
        ctx->c.err.clear();
        return TDBEM_OK;
    } catch (const std::invalid_argument& e) {
        ctx->c.err = e.what();
        return TDBEM_ERR_ARG;
    } catch (const std::logic_error& e) {
        ctx->c.err = e.what();
        return TDBEM_ERR_STATE;
    } catch (const std::runtime_error& e) {
        ctx->c.err = e.what();
        const std::string m = e.what();
        if (m.find("singular") != std::string::npos) return TDBEM_ERR_NUMERIC;
        if (m.find("HIP") != std::string::npos) return TDBEM_ERR_HIP;
        return TDBEM_ERR_ARG;
    } catch (const std::exception& e) {
        ctx->c.err = e.what();
        return TDBEM_ERR_HIP;
    }
}

// 16 B per thread per iteration, grid-stride: reads and writes n16 x 16 bytes once each.
__global__ void __launch_bounds__(256)
    k_stream_copy(const float4* __restrict__ src, float4* __restrict__ dst, int64_t n16)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
        dst[i] = src[i];
}

// Read-only stream: each workgroup reads one contiguous span of kReadSpan 16 B words with 4
// non-temporal loads in flight per thread, folded into one word that is written only if it
// matches a value the data never holds (so the loads stay).  Contiguous spans per workgroup
// stream faster than a grid-stride walk on this part (about 7.1 against 6.5-7.0 TB/s).
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
constexpr int64_t kReadSpan = 16384;
__global__ void __launch_bounds__(256)
    k_stream_read(const u32x4* __restrict__ src, int64_t n16, unsigned* __restrict__ sink)
{
    const int64_t b0 = (int64_t)blockIdx.x * kReadSpan;
    const int64_t b1 = b0 + kReadSpan < n16 ? b0 + kReadSpan : n16;
    u32x4 acc = {0u, 0u, 0u, 0u};
    for (int64_t i = b0 + threadIdx.x; i < b1; i += 4 * 256) {
        u32x4 v[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            const int64_t j = i + u * 256;
            v[u] = j < b1 ? __builtin_nontemporal_load(&src[j]) : u32x4{0u, 0u, 0u, 0u};
        }
#pragma unroll
        for (int u = 0; u < 4; u++) acc ^= v[u];
    }
    const unsigned x = acc.x ^ acc.y ^ acc.z ^ acc.w;
    if (x == 0x7fc01234u) sink[0] = x;
}

}  // namespace

double stream_read(size_t bytes, int launches)
{
    hipStream_t s;
    TD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    double ms = 0;
    {
        DBuf<u32x4> a;
        DBuf<unsigned> sink;
        const int64_t n16 = (int64_t)(bytes / 16);
        a.alloc((size_t)n16);
        sink.alloc(1);
        TD_HIP(hipMemsetAsync(a.p, 0, a.bytes(), s));
        KernelTimer t(s, true);
        for (int l = 0; l < launches; l++) {
            t.begin();
            hipLaunchKernelGGL(k_stream_read, dim3((unsigned)((n16 + kReadSpan - 1) / kReadSpan)),
                               dim3(256), 0, s, a.p, n16, sink.p);
            t.end();
        }
        TD_HIP(hipGetLastError());
        ms = t.collect() / launches;
        TD_HIP(hipStreamSynchronize(s));
    }
    (void)hipStreamDestroy(s);
    return ms;
}

double stream_copy(size_t bytes, int launches)
{
    hipStream_t s;
    TD_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    double ms = 0;
    {
        DBuf<float4> a, b;
        const int64_t n16 = (int64_t)(bytes / 16);
        a.alloc((size_t)n16);
        b.alloc((size_t)n16);
        TD_HIP(hipMemsetAsync(a.p, 0, a.bytes(), s));
        KernelTimer t(s, true);
        for (int l = 0; l < launches; l++) {
            t.begin();
            hipLaunchKernelGGL(k_stream_copy, dim3(8192), dim3(256), 0, s, a.p, b.p, n16);
            t.end();
        }
        TD_HIP(hipGetLastError());
        ms = t.collect() / launches;
        TD_HIP(hipStreamSynchronize(s));
    }
    (void)hipStreamDestroy(s);
    return ms;
}

}  // namespace tdbem

using namespace tdbem;

namespace {
template <typename F>
int batch_guarded(F&& f)
{
    try {
        f();
        return TDBEM_OK;
    } catch (const std::invalid_argument& e) {
        g_create_error = e.what();
        return TDBEM_ERR_ARG;
    } catch (const std::logic_error& e) {
        g_create_error = e.what();
        return TDBEM_ERR_STATE;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return TDBEM_ERR_HIP;
    }
}

template <typename F>
int element_guarded(tdbem_element* el, F&& f)
{
    if (!el) return TDBEM_ERR_ARG;
    const int rc = batch_guarded([&] { f(el->e); });
    el->e.err = rc == TDBEM_OK ? std::string() : g_create_error;
    return rc;
}
}  // namespace

extern "C" {

int tdbem_create(int nnd, const double* coords, int nel, const int* ien, int etype, int nq,
                 double E, double nu, double rho, int if_ext, double dt, int numstep, int kernel,
                 tdbem_context** out)
{
    if (!out) return TDBEM_ERR_ARG;
    *out = nullptr;
    tdbem_context* ctx = nullptr;
    try {
        const double *gx, *gw;
        if (nnd <= 0 || nel <= 0 || !coords || !ien)
            throw std::invalid_argument("empty mesh");
        if (etype != 1 && etype != 2)
            throw std::invalid_argument("InputT::ReadInput, unsupported element type");
        if (!gauss_table(nq, &gx, &gw))
            throw std::invalid_argument(
                "ShapeFunctionT::SetNQ, Unsupported number of Gaussian points");
        if (kernel != TDBEM_KERNEL_STEP && kernel != TDBEM_KERNEL_HEAVISIDE &&
            kernel != TDBEM_KERNEL_BSPLINE && kernel != TDBEM_KERNEL_RAMP)
            throw std::invalid_argument("unsupported temporal kernel");
        if (!(dt > 0) || numstep < 0 || !(E > 0) || !(rho > 0) || !(nu > -1 && nu < 0.5))
            throw std::invalid_argument("bad material or time step");
        const int ennd = etype == 1 ? 4 : 8;
        for (int64_t q = 0; q < (int64_t)nel * ennd; q++)
            if (ien[q] < 0 || ien[q] >= nnd) throw std::invalid_argument("IEN node id out of range");
        ctx = new tdbem_context();
        Context& c = ctx->c;
        TD_HIP(hipGetDevice(&c.device));
        TD_HIP(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
        keep_pool_memory(c.device);
        c.nnd = nnd;
        c.nel = nel;
        c.etype = etype;
        c.ennd = ennd;
        c.nq = nq;
        c.kernel = kernel;
        c.if_ext = if_ext ? 1 : 0;
        c.numstep = numstep;
        c.dt = dt;
        c.mat = make_material(E, nu, rho);
        c.h_coords.assign(coords, coords + 3 * (size_t)nnd);
        c.h_ien.assign(ien, ien + (size_t)nel * ennd);
        c.K = max_step(c.h_coords, nnd, E, nu, rho, dt);
        c.Kc = std::min(c.K, numstep);
        c.coords.upload(c.h_coords.data(), c.h_coords.size(), c.stream);
        c.ien.upload(c.h_ien.data(), c.h_ien.size(), c.stream);
        std::vector<double> g(2 * nq);
        std::memcpy(g.data(), gx, nq * sizeof(double));
        std::memcpy(g.data() + nq, gw, nq * sizeof(double));
        c.gauss.upload(g.data(), g.size(), c.stream);
        prepare_geometry(c);
        // the C-term now, so no later assembly call needs to allocate it; and the device's
        // workspace address for the allocation-free entry point
        compute_cterm(c);
        noalloc_prepare(c.device);
        TD_HIP(hipStreamSynchronize(c.stream));
        *out = ctx;
        return TDBEM_OK;
    } catch (const std::invalid_argument& e) {
        g_create_error = e.what();
        delete ctx;
        return TDBEM_ERR_ARG;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        delete ctx;
        return TDBEM_ERR_HIP;
    }
}

int tdbem_destroy(tdbem_context* ctx)
{
    if (!ctx) return TDBEM_OK;
    if (ctx->c.batch_refs > 0) {
        ctx->c.err = "tdbem_destroy: the context is in a batch";
        return TDBEM_ERR_STATE;
    }
    (void)hipStreamSynchronize(ctx->c.stream);
    delete ctx;
    return TDBEM_OK;
}

const char* tdbem_last_error(const tdbem_context* ctx)
{
    return ctx ? ctx->c.err.c_str() : g_create_error.c_str();
}

int tdbem_max_step(const tdbem_context* ctx) { return ctx ? ctx->c.K : TDBEM_ERR_ARG; }

int tdbem_assemble_step_device(tdbem_context* ctx, int step, int lower, int upper, double* G1,
                               double* G2, double* H1, double* H2)
{
    return guarded(ctx, [&](Context& c) {
        if (step < 1) throw std::invalid_argument("step must be >= 1");
        if (lower < 0 || upper >= c.nnd || upper < lower)
            throw std::invalid_argument("row range out of bounds");
        if (!G1 || !G2 || !H1 || !H2) throw std::invalid_argument("null output");
        assemble_step_dense(c, step, lower, upper, G1, G2, H1, H2);
    });
}

int tdbem_assemble_step_device_noalloc(tdbem_context* ctx, int step, int lower, int upper,
                                       double* G1, double* G2, double* H1, double* H2)
{
    return guarded(ctx, [&](Context& c) {
        if (step < 1) throw std::invalid_argument("step must be >= 1");
        if (lower < 0 || upper >= c.nnd || upper < lower)
            throw std::invalid_argument("row range out of bounds");
        if (!G1 || !G2 || !H1 || !H2) throw std::invalid_argument("null output");
        assemble_step_dense_noalloc(c, step, lower, upper, G1, G2, H1, H2);
    });
}

int tdbem_assemble_step(tdbem_context* ctx, int step, int lower, int upper, double* G1,
                        double* G2, double* H1, double* H2)
{
    return guarded(ctx, [&](Context& c) {
        if (step < 1) throw std::invalid_argument("step must be >= 1");
        if (lower < 0 || upper >= c.nnd || upper < lower)
            throw std::invalid_argument("row range out of bounds");
        if (!G1 || !G2 || !H1 || !H2) throw std::invalid_argument("null output");
        const size_t n = (size_t)3 * (upper - lower + 1) * 3 * c.nnd;
        DBuf<double> b[4];
        for (auto& x : b) x.alloc(n);
        assemble_step_dense(c, step, lower, upper, b[0].p, b[1].p, b[2].p, b[3].p);
        double* outs[4] = {G1, G2, H1, H2};
        for (int q = 0; q < 4; q++)
            TD_HIP(hipMemcpyAsync(outs[q], b[q].p, n * sizeof(double), hipMemcpyDeviceToHost,
                                  c.stream));
        TD_HIP(hipStreamSynchronize(c.stream));
    });
}

int tdbem_build_lag_store(tdbem_context* ctx)
{
    return guarded(ctx, [&](Context& c) {
        not_batched(c, "tdbem_build_lag_store");
        if (c.kernel != TDBEM_KERNEL_STEP)
            throw std::invalid_argument("the lag store is built for the Step kernel");
        caller_drop(c);   // the device assembles the lags itself
        build_lag_store(c);
    });
}

int64_t tdbem_lag_store_bytes(tdbem_context* ctx)
{
    int64_t v = -1;
    int rc = guarded(ctx, [&](Context& c) {
        if (c.store_ready)
            v = (int64_t)(c.store.vals.bytes() + c.store.idx.bytes() + c.tp.vals.bytes() +
                          c.tp.ent.bytes());
        else
            v = predict_store_bytes(c);
    });
    return rc == TDBEM_OK ? v : rc;
}

int tdbem_get_lag(tdbem_context* ctx, int k, double* G, double* H)
{
    return guarded(ctx, [&](Context& c) {
        if (c.caller_last > 0 && (c.caller_stale || !c.store_ready)) caller_store(c);
        if (!c.store_ready) throw std::logic_error("lag store not built");
        if (k < 0 || k > c.Kc) throw std::invalid_argument("lag index out of range");
        const size_t N = 3 * (size_t)c.nnd;
        DBuf<double> g, h;
        g.alloc(N * N);
        h.alloc(N * N);
        if (c.store_precision == 1)
            tile_rows(c, k, 0, c.nnd - 1, g.p, h.p);
        else
            lag_dense(c, k, g.p, h.p);
        TD_HIP(hipMemcpyAsync(G, g.p, N * N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        TD_HIP(hipMemcpyAsync(H, h.p, N * N * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        TD_HIP(hipStreamSynchronize(c.stream));
    });
}

int tdbem_get_lag_rows(tdbem_context* ctx, int k, int lower, int upper, double* G, double* H)
{
    return guarded(ctx, [&](Context& c) {
        if (c.caller_last > 0 && (c.caller_stale || !c.store_ready)) caller_store(c);
        if (!c.store_ready) throw std::logic_error("lag store not built");
        if (k < 0 || k > c.Kc) throw std::invalid_argument("lag index out of range");
        if (lower < 0 || upper >= c.nnd || upper < lower)
            throw std::invalid_argument("row range out of bounds");
        if (!G || !H) throw std::invalid_argument("null output");
        const size_t n = (size_t)3 * (upper - lower + 1) * 3 * c.nnd;
        DBuf<double> g(c.stream), h(c.stream);
        g.alloc(n);
        h.alloc(n);
        if (c.store_precision == 1)
            tile_rows(c, k, lower, upper, g.p, h.p);
        else
            lag_rows(c, k, lower, upper, g.p, h.p);
        TD_HIP(hipMemcpyAsync(G, g.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        TD_HIP(hipMemcpyAsync(H, h.p, n * sizeof(double), hipMemcpyDeviceToHost, c.stream));
        TD_HIP(hipStreamSynchronize(c.stream));
    });
}

int tdbem_update_gh_linear(tdbem_context* ctx, int step, int lower, int upper, const double* G0,
                           const double* G1, const double* H0, const double* H1)
{
    return guarded(ctx, [&](Context& c) {
        not_batched(c, "tdbem_update_gh_linear");
        caller_update(c, step, lower, upper, G0, G1, H0, H1);
    });
}

int tdbem_predict_bytes(tdbem_context* ctx, int nrhs, int64_t* resident, int64_t* peak)
{
    return guarded(ctx, [&](Context& c) {
        if (!resident || !peak) throw std::invalid_argument("null output");
        if (c.kernel != TDBEM_KERNEL_STEP)
            throw std::invalid_argument("the lag store is built for the Step kernel");
        if (c.precision == TDBEM_PRECISION_FP64 && nrhs != 1 && nrhs != 2 && nrhs != 4 &&
            nrhs != 8)
            throw std::invalid_argument("FP64 march: nrhs must be 1, 2, 4 or 8");
        if (c.precision == TDBEM_PRECISION_MIXED_BF16X3 && nrhs != 16 && nrhs != 32 &&
            nrhs != 48 && nrhs != 64)
            throw std::invalid_argument("mixed march: nrhs must be 16, 32, 48 or 64");
        const int64_t nq2 = (int64_t)c.nq * c.nq;
        // what the context holds from its creation and the assembly tables of the first build
        const int64_t base = (int64_t)c.nnd * 3 * 8 + (int64_t)c.nel * c.ennd * 4 + 2LL * c.nq * 8 +
                             c.nel * nq2 * (7 + c.ennd) * 8 + (int64_t)(c.nnd + 1) * 4 +
                             (int64_t)c.nel * c.ennd * 4 + (int64_t)c.nnd * 9 * 8 +
                             (int64_t)c.nnd * c.nel * 8;
        int64_t store = 0, transient = 0;
        predict_store(c, store, transient);
        const int64_t state = march_state_bytes(c, nrhs);
        *resident = base + store + state;
        *peak = std::max(base + store + transient,
                         base + store + state + march_init_transient_bytes(c));
    });
}

int tdbem_device_memory(int64_t* free_bytes, int64_t* total_bytes)
{
    if (!free_bytes || !total_bytes) return TDBEM_ERR_ARG;
    try {
        int dev = 0;
        TD_HIP(hipGetDevice(&dev));
        TD_HIP(hipDeviceSynchronize());
        hipMemPool_t pool = nullptr;
        TD_HIP(hipDeviceGetDefaultMemPool(&pool, dev));
        TD_HIP(hipMemPoolTrimTo(pool, 0));
        size_t f = 0, t = 0;
        TD_HIP(hipMemGetInfo(&f, &t));
        *free_bytes = (int64_t)f;
        *total_bytes = (int64_t)t;
        return TDBEM_OK;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return TDBEM_ERR_HIP;
    }
}

int tdbem_march_init(tdbem_context* ctx, double a1, double a2, int n_ubc, const int* ubc_dof,
                     const double* ubc_val, int n_fbc, const int* fbc_dof, const double* fbc_val,
                     int nrhs, const double* amp, const double* t_rise)
{
    return guarded(ctx, [&](Context& c) {
        not_batched(c, "tdbem_march_init");
        if (c.kernel != TDBEM_KERNEL_STEP)
            throw std::invalid_argument("the march runs the Step kernel (SolverStepLinearT)");
        if (!(a1 >= 0 && a1 <= 1 && a2 >= 0 && a2 < 1 && a1 + a2 <= 1))
            throw std::invalid_argument("Alpha is out of range");
        if (c.precision == TDBEM_PRECISION_FP64 && nrhs != 1 && nrhs != 2 && nrhs != 4 &&
            nrhs != 8)
            throw std::invalid_argument("FP64 march: nrhs must be 1, 2, 4 or 8");
        if (c.precision == TDBEM_PRECISION_MIXED_BF16X3 && nrhs != 16 && nrhs != 32 &&
            nrhs != 48 && nrhs != 64)
            throw std::invalid_argument("mixed march: nrhs must be 16, 32, 48 or 64");
        if ((n_ubc > 0 && (!ubc_dof || !ubc_val)) || (n_fbc > 0 && (!fbc_dof || !fbc_val)))
            throw std::invalid_argument("null boundary-condition array");
        march_init(c, a1, a2, nrhs, amp, t_rise, n_ubc, ubc_dof, ubc_val, n_fbc, fbc_dof,
                   fbc_val);
    });
}

int tdbem_march_run(tdbem_context* ctx, int nsteps)
{
    return guarded(ctx, [&](Context& c) {
        not_batched(c, "tdbem_march_run");
        if (!c.march) throw std::logic_error("march not initialised");
        march_run(c, nsteps);
    });
}

int tdbem_march_enqueue(tdbem_context* ctx, int nsteps)
{
    return guarded(ctx, [&](Context& c) {
        not_batched(c, "tdbem_march_enqueue");
        if (!c.march) throw std::logic_error("march not initialised");
        march_run(c, nsteps, false);
    });
}

int tdbem_query(tdbem_context* ctx)
{
    if (!ctx) return TDBEM_ERR_ARG;
    const hipError_t e = hipStreamQuery(ctx->c.stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    ctx->c.err = std::string("HIP error: ") + hipGetErrorString(e);
    return TDBEM_ERR_HIP;
}

int tdbem_march_step_index(const tdbem_context* ctx)
{
    return ctx ? march_step_count(ctx->c) : TDBEM_ERR_ARG;
}

int tdbem_get_history(tdbem_context* ctx, int rhs, double* U, double* T)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.march) throw std::logic_error("march not initialised");
        if (!U || !T) throw std::invalid_argument("null output");
        march_history(c, rhs, U, T);
    });
}

int tdbem_get_state(tdbem_context* ctx, int rhs, int s, double* u, double* t)
{
    return guarded(ctx, [&](Context& c) {
        if (!c.march) throw std::logic_error("march not initialised");
        if (!u || !t) throw std::invalid_argument("null output");
        if (s < 0 || s >= c.numstep) throw std::invalid_argument("step index out of range");
        march_state(c, rhs, s, u, t);
    });
}

int tdbem_synchronize(tdbem_context* ctx)
{
    return guarded(ctx, [&](Context& c) { TD_HIP(hipStreamSynchronize(c.stream)); });
}

int tdbem_set_precision(tdbem_context* ctx, int precision)
{
    return guarded(ctx, [&](Context& c) {
        not_batched(c, "tdbem_set_precision");
        if (precision != c.precision) caller_drop(c);
        if (precision != TDBEM_PRECISION_FP64 && precision != TDBEM_PRECISION_MIXED_BF16X3)
            throw std::invalid_argument("unknown precision");
        c.precision = precision;
        // the lag store's regular pairs are assembled in the precision of the context
        if (c.store_ready && c.store_precision != precision) {
            c.store_ready = false;
            c.tp.ready = false;
        }
    });
}

int tdbem_mfma_selftest(const float* A, const float* B, float* D)
{
    if (!A || !B || !D) return TDBEM_ERR_ARG;
    try {
        mfma_selftest(A, B, D);
        return TDBEM_OK;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return TDBEM_ERR_HIP;
    }
}

int tdbem_mfma_f64_selftest(const double* A, const double* B, double* D)
{
    if (!A || !B || !D) return TDBEM_ERR_ARG;
    try {
        mfma_f64_selftest(A, B, D);
        return TDBEM_OK;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return TDBEM_ERR_HIP;
    }
}

int tdbem_dense_solve(int n, const double* A, int nrhs, double* B, int mode)
{
    if (n < 1 || nrhs < 1 || !A || !B || (mode & ~3)) return TDBEM_ERR_ARG;
    try {
        dense_solve(n, A, nrhs, B, (mode & TDBEM_SOLVE_INVERSE) != 0,
                    (mode & TDBEM_SOLVE_PIVOTED) != 0);
        return TDBEM_OK;
    } catch (const std::runtime_error& e) {
        g_create_error = e.what();
        return std::string(e.what()).find("singular") != std::string::npos ? TDBEM_ERR_NUMERIC
                                                                           : TDBEM_ERR_HIP;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return TDBEM_ERR_HIP;
    }
}

int tdbem_stream_read(size_t bytes, int launches, double* avg_ms)
{
    if (bytes == 0 || bytes % 16 || launches < 1 || !avg_ms) return TDBEM_ERR_ARG;
    try {
        *avg_ms = stream_read(bytes, launches);
        return TDBEM_OK;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return TDBEM_ERR_HIP;
    }
}

int tdbem_stream_copy(size_t bytes, int launches, double* avg_ms)
{
    if (bytes == 0 || bytes % 16 || launches < 1 || !avg_ms) return TDBEM_ERR_ARG;
    try {
        *avg_ms = stream_copy(bytes, launches);
        return TDBEM_OK;
    } catch (const std::exception& e) {
        g_create_error = e.what();
        return TDBEM_ERR_HIP;
    }
}

int tdbem_set_profiling(tdbem_context* ctx, int level)
{
    return guarded(ctx, [&](Context& c) {
        c.profile = level == 0 ? 0 : (level == TDBEM_PROFILE_CONV ? 2 : 1);
    });
}

int tdbem_reset_timings(tdbem_context* ctx)
{
    return guarded(ctx, [&](Context& c) { c.tm = Context::Timings(); });
}

int tdbem_get_timings(tdbem_context* ctx, double* out)
{
    return guarded(ctx, [&](Context& c) {
        if (!out) throw std::invalid_argument("null output");
        const auto& t = c.tm;
        const double v[TDBEM_NUM_TIMINGS] = {
            t.store_ms, t.init_ms, t.march_ms, t.lag_kernel_ms, t.gather_ms, t.conv_ms,
            t.solve_ms, (double)t.lag_launches, (double)t.conv_launches, (double)t.active_points,
            (double)c.pair_lags, (double)c.store.n_entries, (double)c.tp.nsub,
            t.reduce_ms, (double)c.tp.nblocks, (double)c.tp.nentries, (double)t.active_levels,
            (double)c.tp.pair_lags_class[0], (double)c.tp.pair_lags_class[1],
            (double)c.tp.pair_lags_class[2], (double)c.tp.sub_class[0],
            (double)c.tp.sub_class[1], (double)c.tp.sub_class[2], (double)c.tp.ent_parity[0],
            (double)c.tp.ent_parity[1]};
        std::memcpy(out, v, sizeof(v));
    });
}


int tdbem_batch_create(tdbem_context* const* ctxs, int count, tdbem_batch** out)
{
    if ((!ctxs && count > 0) || count < 0 || !out) return TDBEM_ERR_ARG;
    *out = nullptr;
    return batch_guarded([&] {
        std::vector<Context*> v;
        for (int i = 0; i < count; i++) {
            if (!ctxs[i]) throw std::invalid_argument("batch: null context");
            v.push_back(&ctxs[i]->c);
        }
        std::unique_ptr<tdbem_batch> b(new tdbem_batch());
        b->b = batch_create(v);
        *out = b.release();
    });
}

int tdbem_batch_add(tdbem_batch* b, tdbem_context* ctx)
{
    if (!b || !ctx) return TDBEM_ERR_ARG;
    return batch_guarded([&] { batch_add(*b->b, &ctx->c); });
}

int tdbem_batch_remove(tdbem_batch* b, tdbem_context* ctx)
{
    if (!b || !ctx) return TDBEM_ERR_ARG;
    return batch_guarded([&] { batch_remove(*b->b, &ctx->c); });
}

int tdbem_batch_size(const tdbem_batch* b) { return b ? batch_size(*b->b) : TDBEM_ERR_ARG; }

int tdbem_batch_tune(tdbem_batch* b, double* report)
{
    if (!b) return TDBEM_ERR_ARG;
    return batch_guarded([&] { batch_tune(*b->b, report); });
}

int tdbem_batch_set_params(tdbem_batch* b, int rounds, int solve_waves)
{
    if (!b) return TDBEM_ERR_ARG;
    return batch_guarded([&] { batch_set_params(*b->b, rounds, solve_waves); });
}

int tdbem_batch_run(tdbem_batch* b, int steps_per_graph, int max_replays, int* launches)
{
    if (!b) return TDBEM_ERR_ARG;
    return batch_guarded([&] {
        const int n = batch_run(*b->b, steps_per_graph, max_replays);
        if (launches) *launches = n;
    });
}

int tdbem_batch_destroy(tdbem_batch* b)
{
    delete b;
    return TDBEM_OK;
}

int tdbem_set_max_step(tdbem_context* ctx, int max_step)
{
    return guarded(ctx, [&](Context& c) {
        not_batched(c, "tdbem_set_max_step");
        caller_drop(c);
        set_max_step(c, max_step);
    });
}

int tdbem_mrt_create(tdbem_context* region1, tdbem_context* region2, int max_step, int n_iface,
                     const int* iface1, const int* iface2, tdbem_mrt** out)
{
    if (!region1 || !region2 || !out) return TDBEM_ERR_ARG;
    *out = nullptr;
    return batch_guarded([&] {
        not_batched(region1->c, "tdbem_mrt_create");
        not_batched(region2->c, "tdbem_mrt_create");
        std::unique_ptr<tdbem_mrt> m(new tdbem_mrt());
        m->m = mrt_create(&region1->c, &region2->c, max_step, n_iface, iface1, iface2);
        *out = m.release();
    });
}

int tdbem_mrt_init(tdbem_mrt* m, int scheme, int load_kind, double amp, double t_rise,
                   int n_fbc1, const int* fbc_dof1, const double* fbc_val1, int n_fbc2,
                   const int* fbc_dof2, const double* fbc_val2)
{
    if (!m) return TDBEM_ERR_ARG;
    return batch_guarded([&] {
        const int n[2] = {n_fbc1, n_fbc2};
        const int* dof[2] = {fbc_dof1, fbc_dof2};
        const double* val[2] = {fbc_val1, fbc_val2};
        if (n_fbc1 < 0 || n_fbc2 < 0) throw std::invalid_argument("negative FBC count");
        mrt_init(*m->m, scheme, load_kind, amp, t_rise, n, dof, val);
    });
}

int tdbem_mrt_run(tdbem_mrt* m, int nsteps)
{
    if (!m) return TDBEM_ERR_ARG;
    return batch_guarded([&] { mrt_run(*m->m, nsteps); });
}

int tdbem_mrt_step_index(const tdbem_mrt* m) { return m ? mrt_step_count(*m->m) : TDBEM_ERR_ARG; }

int tdbem_mrt_num_unknowns(const tdbem_mrt* m) { return m ? mrt_unknowns(*m->m) : TDBEM_ERR_ARG; }

int tdbem_mrt_get_state(tdbem_mrt* m, int s, double* x, double* u, double* t)
{
    if (!m) return TDBEM_ERR_ARG;
    return batch_guarded([&] { mrt_state(*m->m, s, x, u, t); });
}

int tdbem_mrt_get_timings(const tdbem_mrt* m, double* init_ms, double* run_ms)
{
    if (!m || !init_ms || !run_ms) return TDBEM_ERR_ARG;
    mrt_timings(*m->m, init_ms, run_ms);
    return TDBEM_OK;
}

int tdbem_mrt_destroy(tdbem_mrt* m)
{
    delete m;
    return TDBEM_OK;
}

int tdbem_element_create(int etype, int kernel, int nq, double E, double nu, double density,
                         tdbem_element** out)
{
    if (!out) return TDBEM_ERR_ARG;
    *out = nullptr;
    tdbem_element* el = nullptr;
    const int rc = batch_guarded([&] {
        const double *gx, *gw;
        if (etype != 1 && etype != 2)
            throw std::invalid_argument(
                "TDElementT::TDElementT, unsupported element type. There is no sub element for "
                "this element");
        if (!gauss_table(nq, &gx, &gw))
            throw std::invalid_argument(
                "ShapeFunctionT::SetNQ, Unsupported number of Gaussian points");
        if (kernel != TDBEM_KERNEL_STEP && kernel != TDBEM_KERNEL_HEAVISIDE &&
            kernel != TDBEM_KERNEL_BSPLINE && kernel != TDBEM_KERNEL_RAMP)
            throw std::invalid_argument("unsupported temporal kernel");
        if (!(E > 0) || !(density > 0) || !(nu > -1 && nu < 0.5))
            throw std::invalid_argument("bad material");
        el = new tdbem_element();
        Element& e = el->e;
        TD_HIP(hipGetDevice(&e.device));
        TD_HIP(hipStreamCreateWithFlags(&e.stream, hipStreamNonBlocking));
        keep_pool_memory(e.device);
        e.etype = etype;
        e.ennd = etype == 1 ? 4 : 8;
        e.nq = nq;
        e.kernel = kernel;
        e.mat = make_material(E, nu, density);
        std::vector<double> g(2 * nq);
        std::memcpy(g.data(), gx, nq * sizeof(double));
        std::memcpy(g.data() + nq, gw, nq * sizeof(double));
        e.gauss.upload(g.data(), g.size(), e.stream);
        TD_HIP(hipStreamSynchronize(e.stream));
    });
    if (rc != TDBEM_OK) {
        delete el;
        return rc;
    }
    *out = el;
    return TDBEM_OK;
}

int tdbem_element_destroy(tdbem_element* el)
{
    delete el;
    return TDBEM_OK;
}

const char* tdbem_element_last_error(const tdbem_element* el)
{
    return el ? el->e.err.c_str() : g_create_error.c_str();
}

int tdbem_element_form(tdbem_element* el, int64_t npairs, const double* coords,
                       const double* collocation, const int* singular, double t, double t_1,
                       double* GE_1, double* GE_2, double* HE_1, double* HE_2, double* C_1,
                       double* C_2)
{
    return element_guarded(el, [&](Element& e) {
        if (npairs < 0) throw std::invalid_argument("npairs < 0");
        if (npairs == 0) return;
        if (!coords || !collocation || !singular || !GE_1 || !GE_2 || !HE_1 || !HE_2 || !C_1 ||
            !C_2)
            throw std::invalid_argument("null array");
        if (!(t_1 > 0) || !std::isfinite(t)) throw std::invalid_argument("bad time");
        for (int64_t p = 0; p < npairs; p++)
            if (singular[p] < -1 || singular[p] >= e.ennd)
                throw std::invalid_argument(
                    "TDElementT::FormGH_Singular, collocation id is out of range");
        const int64_t nx = npairs * 3 * e.ennd, w = npairs * 9 * e.ennd;
        e.in.reserve((size_t)(nx + 3 * npairs));
        e.sing.reserve((size_t)npairs);
        e.out.reserve((size_t)(4 * w + 18 * npairs));
        TD_HIP(hipMemcpyAsync(e.in.p, coords, nx * sizeof(double), hipMemcpyHostToDevice,
                              e.stream));
        TD_HIP(hipMemcpyAsync(e.in.p + nx, collocation, 3 * npairs * sizeof(double),
                              hipMemcpyHostToDevice, e.stream));
        TD_HIP(hipMemcpyAsync(e.sing.p, singular, npairs * sizeof(int), hipMemcpyHostToDevice,
                              e.stream));
        element_form(e, npairs, e.in.p, e.in.p + nx, e.sing.p, t, t_1, e.out.p);
        double* dst[6] = {GE_1, GE_2, HE_1, HE_2, C_1, C_2};
        for (int b = 0; b < 6; b++) {
            const int64_t off = b < 4 ? b * w : 4 * w + (b - 4) * 9 * npairs;
            const int64_t len = b < 4 ? w : 9 * npairs;
            TD_HIP(hipMemcpyAsync(dst[b], e.out.p + off, len * sizeof(double),
                                  hipMemcpyDeviceToHost, e.stream));
        }
        TD_HIP(hipStreamSynchronize(e.stream));
    });
}

int tdbem_element_form_device(tdbem_element* el, int64_t npairs, const double* coords,
                              const double* collocation, const int* singular, double t,
                              double t_1, double* out)
{
    return element_guarded(el, [&](Element& e) {
        if (npairs < 0) throw std::invalid_argument("npairs < 0");
        if (npairs == 0) return;
        if (!coords || !collocation || !singular || !out)
            throw std::invalid_argument("null array");
        if (!(t_1 > 0) || !std::isfinite(t)) throw std::invalid_argument("bad time");
        element_form(e, npairs, coords, collocation, singular, t, t_1, out);
    });
}

int tdbem_element_synchronize(tdbem_element* el)
{
    return element_guarded(el, [&](Element& e) { TD_HIP(hipStreamSynchronize(e.stream)); });
}

}  // extern "C"
