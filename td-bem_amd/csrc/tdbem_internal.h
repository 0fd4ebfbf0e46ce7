// tdbem_internal.h — host-side context and launch declarations shared by the HIP translation
// units of libtdbem.  Not part of the public C ABI (include/tdbem.h).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <thread>
#include <vector>

#include "tdbem_device.h"

namespace tdbem {

// Throws a std::runtime_error carrying `msg` and the HIP error string when `e` is not success.
void check_hip(hipError_t e, const char* what);
#define TD_HIP(call) ::tdbem::check_hip((call), #call)

// Device buffer owned by a context (freed by ~DBuf).  A buffer constructed with a stream is
// stream-ordered: hipMallocAsync / hipFreeAsync on that stream (the device's memory pool, which
// keep_pool_memory makes keep freed memory), for temporaries used on that stream only; hipFree
// synchronises the whole device, so the temporaries of set-ups running on several streams at
// once would otherwise wait for each other's kernels.
template <typename T>
struct DBuf {
    T* p = nullptr;
    size_t n = 0;
    size_t cap = 0;
    hipStream_t ord = nullptr;
    bool ordered = false;
    DBuf() = default;
    explicit DBuf(hipStream_t s) : ord(s), ordered(true) {}
    DBuf(const DBuf&) = delete;
    DBuf& operator=(const DBuf&) = delete;
    ~DBuf() { reset(); }
    // stream-ordered on s from the next allocation on (what the buffer holds is freed first)
    void order_on(hipStream_t s)
    {
        reset();
        ord = s;
        ordered = true;
    }
    void reset()
    {
        if (p) {
            if (ordered)
                (void)hipFreeAsync(p, ord);
            else
                (void)hipFree(p);
        }
        p = nullptr;
        n = cap = 0;
    }
    void alloc(size_t count)
    {
        if (count == n && p) return;
        reset();
        if (count == 0) return;
        if (ordered)
            TD_HIP(hipMallocAsync((void**)&p, count * sizeof(T), ord));
        else
            TD_HIP(hipMalloc(&p, count * sizeof(T)));
        n = cap = count;
    }
    // n = count, keeping the allocation when it is large enough (no hipFree / hipMalloc, which
    // synchronise the device, in loops that resize a buffer)
    void reserve(size_t count)
    {
        if (p && count <= cap) {
            n = count;
            return;
        }
        alloc(count);
    }
    void upload(const T* h, size_t count, hipStream_t s)
    {
        reserve(count);
        if (count) TD_HIP(hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s));
    }
    size_t bytes() const { return n * sizeof(T); }
};

// Pinned host copies of the shell planning (24 B per node pair) larger than this are released
// after use instead of being kept for the next build.
constexpr int64_t kPinnedKeepBytes = 256LL << 20;

// Pinned host buffer (hipHostMalloc) that keeps its allocation across resizes: large transfers
// between the shell planning on the host and the device run at full bus speed.
template <typename T>
struct HostBuf {
    T* p = nullptr;
    size_t n = 0, cap = 0;
    HostBuf() = default;
    HostBuf(const HostBuf&) = delete;
    HostBuf& operator=(const HostBuf&) = delete;
    ~HostBuf()
    {
        if (p) (void)hipHostFree(p);
    }
    // reserve(0) releases the allocation
    void reserve(size_t count)
    {
        if (p && count <= cap && count > 0) {
            n = count;
            return;
        }
        if (p) (void)hipHostFree(p);
        p = nullptr;
        n = cap = 0;
        if (count == 0) return;
        TD_HIP(hipHostMalloc(reinterpret_cast<void**>(&p), count * sizeof(T), hipHostMallocDefault));
        n = cap = count;
    }
    T& operator[](size_t i) { return p[i]; }
    const T& operator[](size_t i) const { return p[i]; }
};

// f(lo, hi) over [0, n) split into contiguous ranges on up to 16 host threads, at least `grain`
// items each (host-side planning loops over rows of node pairs).
template <typename F>
void parallel_for(int64_t n, F&& f, int64_t grain = 1)
{
    const int64_t nt = std::min<int64_t>(
        {16, (int64_t)std::max(1u, std::thread::hardware_concurrency()),
         std::max<int64_t>(n / std::max<int64_t>(grain, 1), 1)});
    if (nt <= 1) {
        f((int64_t)0, n);
        return;
    }
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int64_t t = 0; t < nt; t++) th.emplace_back([&f, t, n, nt] { f(n * t / nt, n * (t + 1) / nt); });
    for (auto& x : th) x.join();
}

// Brackets kernel launches with HIP events on one stream; collect() waits and sums the times.
struct KernelTimer {
    hipStream_t s = nullptr;
    bool on = false;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;
    size_t used = 0;   // pairs recorded since the last collect; the rest are kept for reuse
    KernelTimer(hipStream_t stream, bool enabled) : s(stream), on(enabled) {}
    ~KernelTimer()
    {
        collect();
        for (auto& p : ev) {
            (void)hipEventDestroy(p.first);
            (void)hipEventDestroy(p.second);
        }
    }
    // creates the events of n launches up front, outside the launch sequence
    void reserve(size_t n)
    {
        while (on && ev.size() < n) {
            hipEvent_t a, b;
            TD_HIP(hipEventCreate(&a));
            TD_HIP(hipEventCreate(&b));
            ev.emplace_back(a, b);
        }
    }
    void begin()
    {
        if (!on) return;
        if (used == ev.size()) {
            hipEvent_t a, b;
            TD_HIP(hipEventCreate(&a));
            TD_HIP(hipEventCreate(&b));
            ev.emplace_back(a, b);
        }
        TD_HIP(hipEventRecord(ev[used++].first, s));
    }
    void end()
    {
        if (on && used > 0) TD_HIP(hipEventRecord(ev[used - 1].second, s));
    }
    int64_t count() const { return (int64_t)used; }
    double collect()
    {
        double ms = 0;
        if (used > 0) (void)hipEventSynchronize(ev[used - 1].second);
        for (size_t i = 0; i < used; i++) {
            float t = 0;
            if (hipEventElapsedTime(&t, ev[i].first, ev[i].second) == hipSuccess) ms += t;
        }
        used = 0;
        return ms;
    }
};

// Entries of the shell lag store: 15 values per (row node i, column node j, lag k):
// G[p][q] symmetric (g00 g01 g02 g11 g12 g22) then H[p][q] row-major.  Entries of one row node
// are ordered by (j, k) and padded to a multiple of 64 (one wavefront); values are stored in
// 64-entry chunks as [chunk][15][64] so each component load of a wavefront is one 512-byte line.
constexpr int kVals = 15;
constexpr int kChunk = 64;
constexpr int kKBits = 10;  // entry index = (j << kKBits) | k
constexpr int kMaxLag = (1 << kKBits) - 1;

struct LagStore {
    int64_t n_entries = 0;        // padded
    int64_t n_chunks = 0;
    DBuf<double> vals;            // [n_chunks][15][64]
    DBuf<uint32_t> idx;           // [n_entries]
    DBuf<int64_t> row_chunk;      // [nnd + 1] chunk offsets per row node
    DBuf<int64_t> row_valid;      // [nnd] entries before padding, per row node
    std::vector<int64_t> h_row_chunk, h_row_valid;
    int64_t nonzero_entries = 0;  // before padding
};

// Row tiles per group of the mixed-precision convolution: one wavefront each in a workgroup.
constexpr int kGroup = 4;

// One step of a tile group's merged list: column tile J and lag k (id = J << 10 | k); for each
// of the group's row tiles the first stored sub-tile of its block at (J, k) and the block's
// sub-tile mask (0: no block); the k-tiles any of them uses.
struct TPEntry {
    uint32_t sub[kGroup];
    uint16_t msk[kGroup];
    uint32_t id;
    uint32_t kmask;
};

// Tile-pair store of the mixed-precision convolution (mixed.hip): nodes ordered by recursive
// coordinate bisection into tiles of 16; per (row tile, column tile, lag) a 48 x 96 block of
// which only the non-zero 16 x 32 sub-tiles are stored, FP32, in MFMA operand order.
struct TPStore {
    int nT = 0, nG = 0;              // row tiles, groups of 4 row tiles
    std::vector<int> perm, pos;      // tile position -> node (-1 = padding); node -> position
    DBuf<int> d_perm, d_pos;
    int64_t nblocks = 0;             // blocks with a non-zero sub-tile
    int64_t nsub = 0;                // stored sub-tiles
    int64_t nentries = 0;            // merged-list entries over all groups
    DBuf<float> vals;                // [nsub][2][64][4] FP32, 16x16x32 A-operand order
    // merged lists, two per group: list 2 g + p holds the near lags (k < kNearLags) and the far
    // lags of parity p of group g, lists in order
    DBuf<TPEntry> ent;
    std::vector<int64_t> h_grp_off;  // [2 nG + 1] first entry of each list
    std::vector<float> h_work;       // per entry: relative cost (bytes moved)
    // per lag class (near, far even, far odd): in-shell node-pair lags, stored sub-tiles; entries
    // of the parity-0 and parity-1 lists
    int64_t pair_lags_class[3] = {0, 0, 0}, sub_class[3] = {0, 0, 0}, ent_parity[2] = {0, 0};
    // launch plan of k_conv_mixed for one right-hand-side count: one workgroup per chunk, chunks
    // of equal work inside each list, the parity-0 lists' chunks first
    int plan_nrhs = 0, plan_rounds = 0;
    int nchunks = 0, nch[2] = {0, 0};
    DBuf<int4> chunk;                // (group | parity << 24, first entry, end entry, 0)
    DBuf<int> grp_chunk;             // [2][nG + 1] first chunk of each group per parity
    DBuf<int64_t> d_grp_off;         // [2 nG + 1] first entry of each list
    // planning (tp_plan_structure) before the sub-tiles are filled: the stored sub-tiles'
    // (I << 16 | J, k << 4 | u), ordered by tile pair, and the first sub-tile of each row tile
    std::vector<uint2> h_sub_id;
    DBuf<uint2> d_sub_id;
    std::vector<int64_t> h_tile_sub; // [nT + 1]
    bool planned = false;            // structure planned from the node-pair hulls
    bool ready = false;              // sub-tiles filled
};

struct Context {
    // problem
    int nnd = 0, nel = 0, etype = 1, ennd = 4, nq = 6, kernel = 4, if_ext = 0;
    int numstep = 0, K = 0, Kc = 0;
    double dt = 0;
    Material mat{};
    hipStream_t stream = nullptr;
    int device = 0;
    std::string err;

    std::vector<double> h_coords;
    std::vector<int> h_ien;
    DBuf<double> coords;   // [nnd][3]
    DBuf<int> ien;         // [nel][ennd]
    DBuf<double> gauss;    // [2][nq] abscissae then weights (shape table)
    DBuf<double> geo;      // [nel][nq*nq][7+ennd]: x(3) n(3) wJ N(ennd)
    DBuf<int> adj_off;     // [nnd+1]
    DBuf<int> adj;         // (e << 3) | a
    std::vector<int> h_adj_off, h_adj;
    DBuf<double> cterm;    // [nnd][9] sum over elements of the Kelvin C-term
    bool cterm_ready = false;
    DBuf<int2> pair_k;     // [nnd*nel] (k_lo, k_hi), k_hi < k_lo when the pair is empty
    // host copies of the shell planning (pinned, kept between builds): lag range and first store
    // entry of every node pair [nnd*nnd] (the tile-pair copy reads pin_hull), pair_k [nnd*nel]
    HostBuf<int2> pin_hull;
    HostBuf<int64_t> pin_off;
    HostBuf<int2> pin_pk;
    bool ranges_ready = false;

    LagStore store;
    bool store_ready = false;
    // in-shell (row node, column node, lag) entries of the full store; the FP64 store holds all
    // of them (store_lags = Kc), the mixed-precision build only those of lags 0 and 1, the rest
    // living in the tile store (store_lags = 1)
    int64_t pair_lags = 0;
    int store_lags = 0;
    // lag matrices supplied by the caller step by step (UpdateGHLinear): a dense database of
    // Kc + 1 lags of G and H on the device, the last step supplied, and whether the store was
    // built from it before the latest supply
    DBuf<double> caller_db;   // [Kc + 1][2][N][N]
    int caller_last = 0;      // 0: not in this mode
    bool caller_stale = false;
    TPStore tp;
    int precision = 0;  // 0: FP64 march, 1: split-BF16 MFMA convolution (mixed)
    int store_precision = 0;  // precision the lag store was assembled with

    // march state (march.hip)
    struct March* march = nullptr;
    // batches holding this context (batch.hip): while nonzero, the calls that would free or
    // re-initialise the march state or the store the batch reads are refused
    int batch_refs = 0;

    // Optional timing of the hot kernels with HIP events on the context stream.
    int profile = 0;   // 0 off, 1 every timed kernel, 2 the history convolution only
    struct Timings {
        double store_ms = 0, init_ms = 0, march_ms = 0;  // host wall clock, synchronised
        double lag_kernel_ms = 0, gather_ms = 0, conv_ms = 0, solve_ms = 0, reduce_ms = 0;
        int64_t lag_launches = 0, conv_launches = 0;
        int64_t active_points = 0;  // (quadrature point, lag) evaluations of the lag kernel
        int64_t active_levels = 0;  // time levels evaluated at them (1 to 3 each)
    } tm;
    DBuf<unsigned long long> counters;  // device counters of the lag kernel

    ~Context();
};

// Material constants with the wave speeds (capi.hip).
Material make_material(double E, double nu, double rho);

// A reusable element evaluator: TDElementT (src/TDElementT.h:16-115) for (collocation point,
// element) pairs of no mesh, on one device and stream; its buffers grow with the largest call.
struct Element {
    int device = 0, etype = 1, ennd = 4, nq = 6, kernel = 4;
    Material mat{};
    hipStream_t stream = nullptr;
    DBuf<double> gauss;             // [nq] abscissae then [nq] weights
    DBuf<double> in, out;           // staging of host-pointer calls
    DBuf<int> sing;
    std::string err;
    ~Element();
};

// assemble.hip
// Element matrices of npairs pairs (device pointers, stream-ordered): X [npairs][ennd][3], coll
// [npairs][3], sing [npairs]; out: GE_1, GE_2, HE_1, HE_2 (npairs * 9 ennd each), C_1, C_2
// (npairs * 9 each), the arrays of TDElementT::GetGHLinear after SetTime(t, t_1) and
// FormGH_Regular (sing -1) or FormGH_Singular(sing).
void element_form(Element& el, int64_t npairs, const double* X, const double* coll,
                  const int* sing, double t, double t_1, double* out);
void prepare_geometry(Context& c);
void compute_cterm(Context& c);
void compute_pair_ranges(Context& c);
// Dense rows of one step (G_H_Driver).  work == nullptr: the element slots live in a buffer of
// the call, which waits for its kernels; else in the caller's device workspace of work_doubles
// (no allocation, no wait; the C-term must already be computed).
void assemble_step_dense(Context& c, int step, int lower, int upper, double* dG1, double* dG2,
                         double* dH1, double* dH2, double* work = nullptr,
                         int64_t work_doubles = 0);
// The allocation-free variant: the workspace is the library's per-GPU static array (noalloc.hip)
void assemble_step_dense_noalloc(Context& c, int step, int lower, int upper, double* dG1,
                                 double* dG2, double* dH1, double* dH2);
// bytes of that array; noalloc_prepare looks its address up on a device (tdbem_create does, once
// per device, so the entry point itself only reads the kept address)
int64_t noalloc_scratch_bytes();
void noalloc_prepare(int device);
void build_lag_store(Context& c);
// Allocates the FP64 store whose node pair (i, j) holds lags hull[i * nnd + j] (.x..y; empty
// when y < x) and fills its entry index; the values are left to the caller (caller.hip).
void alloc_store_for_hull(Context& c, const int2* hull);

// caller.hip: the lag matrices of caller-supplied per-step blocks (UpdateGHLinear)
// caller_update: the blocks of one step (rows of collocation nodes lower..upper, host) into the
// dense lag database on the device, as UpdateGHLinear accumulates them; caller_store: the FP64
// store from that database (lags with any nonzero 3 x 3 block); caller_drop: leaves the mode.
void caller_update(Context& c, int step, int lower, int upper, const double* G0,
                   const double* G1, const double* H0, const double* H1);
void caller_store(Context& c);
void caller_drop(Context& c);
void lag_dense(Context& c, int k, double* dG, double* dH);   // [N][N] each, device
// rows of collocation nodes lower..upper of lag k: [3 (upper - lower + 1)][N] each, device
void lag_rows(Context& c, int k, int lower, int upper, double* dG, double* dH);
int64_t predict_store_bytes(Context& c);
// device bytes of the store the next build_lag_store makes (store) and the most its build holds
// besides (transient: node-pair tables, slot buffers, sub-tile list), planned without building
void predict_store(Context& c, int64_t& store, int64_t& transient);

// march.hip
void march_init(Context& c, double a1, double a2, int nrhs, const double* amp,
                const double* t_rise, int n_ubc, const int* ubc_dof, const double* ubc_val,
                int n_fbc, const int* fbc_dof, const double* fbc_val);
void march_run(Context& c, int nsteps, bool wait = true);
int march_step_count(const Context& c);
void march_history(Context& c, int rhs, double* U, double* T);
void march_state(Context& c, int rhs, int s, double* u, double* t);
void march_free(Context& c);
// device bytes of march_init's state for nrhs right-hand sides (held until the march is freed)
// and of its temporaries (the dense step-0 matrices); the store must be planned (predict_store)
int64_t march_state_bytes(Context& c, int nrhs);
int64_t march_init_transient_bytes(const Context& c);
// k_conv: hist[r][3i+p] = sum over the FP64 lag store of G z^t - H z^u, z from ring slot
// (s - k) mod R of the node-major layout [R][2][3 nnd][nrhs]; nrhs in {1, 2, 4, 8}
void conv_fp64(Context& c, const double* ring, int R, int s, int nrhs, double* hist,
               hipStream_t st);

// mrt.hip: the two-region solver (SolverMRT) over two FP64 region contexts
struct Mrt;
// SetMaxStep: the lag store holds M_0 .. M_K (K capped at numstep); a change drops the store
void set_max_step(Context& c, int K);
Mrt* mrt_create(Context* c1, Context* c2, int max_step, int nz, const int* if1, const int* if2);
void mrt_destroy(Mrt* M);
void mrt_init(Mrt& M, int scheme, int load_kind, double amp, double t_rise, const int* n_fbc,
              const int* const* fbc_dof, const double* const* fbc_val);
void mrt_run(Mrt& M, int nsteps);
int mrt_unknowns(const Mrt& M);
int mrt_step_count(const Mrt& M);
void mrt_state(Mrt& M, int s, double* x, double* u, double* t);
void mrt_timings(const Mrt& M, double* init_ms, double* run_ms);

// One scenario of a batched march, as the batched kernels see it (batch.hip): the march state
// and the mixed-precision store of one context.
struct BatchScen {
    // march (march.hip)
    double* ring;
    const double* gF;
    const double* gU;
    double* Xf;
    const double* Af;
    double* parts;          // column-split partial solutions, owned by the batch
    const int* is_ubc;
    const int* is_fbc;
    const double* ubc_val;
    const double* fbc_val;
    double* fac;
    const double* amp;
    const double* rise;
    double* U;
    double* T;
    int N, Np, nrhs, R, numstep, nsplit;
    double dt, a1, a2, a3;
    // mixed-precision store (mixed.hip)
    const void* tp;
    const TPEntry* ent;
    float* zs;              // the ring as split bf16 terms (zcopy_elem)
    const int* perm;
    double* carry;          // time-blocked convolution: the next step's part of the history sum
    int nT, nG, gch0;       // gch0: this scenario's first entry in the batch's group-chunk table
    // its chunks in the batch's chunk table: the first, and how many of lag parity 0 (those
    // first) and 1
    int ch0, nch0, nch1;
};
// fills the march part of d (requires a mixed-precision march_init); march_set_step stores the
// step a batch reached
void march_describe(Context& c, BatchScen& d);
void march_set_step(Context& c, int s);

// batch.hip: scenarios stepped together (see the file's header)
struct Batch;
Batch* batch_create(const std::vector<Context*>& ctxs);
void batch_destroy(Batch* b);
// membership changes take effect at the next run or tune (a new plan and graph); a member's
// batch_refs is held until it leaves or the batch is destroyed
void batch_add(Batch& b, Context* c);
void batch_remove(Batch& b, Context* c);
int batch_size(const Batch& b);
// times the tunables on the batch; report = {rounds, apply waves, conv ms, apply ms}
void batch_tune(Batch& b, double* report);
void batch_set_params(Batch& b, int rounds, int apply_waves);
// steps every scenario to its last step (or max_replays graph launches); returns the launches
int batch_run(Batch& b, int steps_per_graph, int max_replays);

// mixed.hip
// Tile store of the mixed-precision march.  tp_plan_structure plans it from the node-pair hulls
// (c.pin_hull) on the host: bisection order, stored sub-tiles, merged lists and their work;
// tp_alloc allocates the sub-tiles; tp_gather_segment fills those of row tiles [I0, I1) from a
// segment of the FP64 store holding their rows (bisection positions 16 I0 ..: values seg, pair
// hulls and first entries [row][nnd]); tp_finish marks it complete.  tp_build checks that a
// built mixed-precision store has its tiles.
void tp_plan_structure(Context& c);
void tp_alloc(Context& c);
void tp_gather_segment(Context& c, int I0, int I1, const double* seg, const int2* hull,
                       const int64_t* off);
void tp_finish(Context& c);
void tp_build(Context& c);
// rows lower..upper of lag k from the tile store (FP32 values), dense [3 (upper-lower+1)][N]
void tile_rows(Context& c, int k, int lower, int upper, double* dG, double* dH);
// start of step s of the mixed march (one launch): FP32 copies of ring slots s - 2 .. s in the
// B-operand order of the mixed convolution, and the load factors fac[r] = f_r(t_load)
void tp_prepare_step(Context& c, const double* ring, int N, int nrhs, int s, int R, float* zs,
                     const double* amp, const double* rise, double t_load, double* fac);
int64_t tp_zslot_elems(const Context& c, int nrhs);
// plans the chunks of k_conv_mixed for nrhs right-hand sides (after tp_build)
void tp_plan(Context& c, int nrhs, int rounds = 4);
// chunk count of context c's convolution at `rounds` (a function of its work alone)
int tp_chunk_target(const Context& c, int rounds);
// appends `target` work-balanced chunks of context c's merged lists (chunk.w = sc) and the
// absolute first chunk of each of its groups (nG + 1 entries) to gch
void tp_chunks(const Context& c, int target, int sc, std::vector<int4>& chunk,
               std::vector<int>& gch);
// resident workgroups of k_conv_mixed for nrhs right-hand sides over the whole GPU
int tp_resident_groups(int nrhs, int device);
double tp_work(const Context& c);
int64_t tp_partial_elems(const Context& c, int nrhs);
// doubles of the time-blocked convolution's carry (the part of the next step's history sum
// convolved at this step), zero before the first step
int64_t tp_carry_elems(const Context& c, int nrhs);
// k_conv_mixed: per-wave FP64 partial sums of the history at step s (time-blocked: the far lags
// of parity s mod 2 for steps s and s + 1, the near lags for step s)
void tp_conv(Context& c, const float* zs, int nrhs, int Rring, int s, double* partial);
// fixed-order sum of the partials of step s, written as the step solve's right-hand side
// Xf = f_r gF + gU + (history sum + carry) in fragment order; the carry takes the next step's
void tp_reduce_rhs(Context& c, const double* partial, int nrhs, int N, int s, const double* gF,
                   const double* gU, const double* fac, double* carry, double* Xf);
void mfma_selftest(const float* A, const float* B, float* D);
double stream_copy(size_t bytes, int launches);
double stream_read(size_t bytes, int launches);

// lu.hip: in-place LU with partial pivoting of a dense row-major n x n device matrix, and
// solves with it.  piv is a device int array of length n.
// Gauss-Legendre rule of the reference for nq in {6, 10, 20, 40} (capi.hip); false otherwise.
bool gauss_table(int nq, const double** x, const double** w);

// P A = L U in place (row-major, piv as LAPACK's getrf): without interchanges when partial
// pivoting would make none, else with them; lu_factor_pivoted always takes the pivoted path.
void lu_factor(double* A, int n, int* piv, hipStream_t s);
void lu_factor_pivoted(double* A, int n, int* piv, hipStream_t s);
// Host-to-host X = A^{-1} B with lu_factor + lu_solve, or through lu_inverse (diagnostic);
// pivoted: factor with lu_factor_pivoted.
void dense_solve(int n, const double* A, int nrhs, double* B, bool inverse, bool pivoted);
void lu_solve(const double* LU, int n, const int* piv, double* B, int ncols, hipStream_t s);
void lu_inverse(const double* LU, int n, const int* piv, double* Ainv, hipStream_t s);

// small dense helpers (lu.hip)
void gemv(const double* A, int n, const double* x, double* y, double beta, hipStream_t s);

// Step solve with the inverse in v_mfma_f64_16x16x4 fragment order (lu.hip).  A matrix of
// n x n is held as 16-row x 32-column fragments: element (i, j) of fragment (T, c) sits at
// ((T * frag_cols(n) + c) * 64 + lane) * 8 + e with i = lane % 16, j = 8 * (lane / 16) + e,
// rows padded to frag_rows(n) and columns to 32 * frag_cols(n) with zeros.  Right-hand sides
// use the same order with 16 of them per fragment row ("rhs tile").
__host__ __device__ inline int frag_rows(int n) { return (n + 15) / 16 * 16; }
__host__ __device__ inline int frag_cols(int n) { return (n + 31) / 32; }
__host__ __device__ inline int64_t frag_index(int n, int row, int col)
{
    const int T = row >> 4, i = row & 15, c = col >> 5, kk = col & 31;
    return (((int64_t)T * frag_cols(n) + c) * 64 + i + 16 * (kk >> 3)) * 8 + (kk & 7);
}
int64_t frag_elems(int n, int rows);  // doubles of a fragment array of `rows` rows
void to_frag(const double* A, int n, double* Af, hipStream_t s);
// column splits of the step solve; its partial sums are parts[split][nrhs][frag_rows(n)]
int apply_splits(int n, int target_waves = 4096);
// parts[p][r][i] = sum over the columns j of split p of A[i][j] X[r][j]
void apply_frag(const double* Af, int n, const double* Xf, int nrhs, double* parts,
                hipStream_t s);
void mfma_f64_selftest(const double* A, const double* B, double* D);

}  // namespace tdbem
