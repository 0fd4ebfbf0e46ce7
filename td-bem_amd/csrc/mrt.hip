// mrt.hip — the two-region solver SolverMRT (src/SolverMRT.cpp) on the GPU, FP64.
//
// Unknowns follow the reference's global vectors (InterfaceIndexing, :223-568): the interior
// dofs of region 1, those of region 2, the interface displacements, then the interface
// tractions of region 1, the interface in region 1's node order (region 2's interface nodes are
// matched to region 1's by coordinates, :440-477).  Rows are region 1's dofs then region 2's.
// Region 1's interface G enters the system matrix negated (Assemble_GH_R1, :685-767), region 2's
// as it is (Assemble_GH_R2, :769-850): tau is region 1's interface traction, region 2 carries
// -tau.
//
// Each global lag matrix is a column permutation of the two regions' own lag matrices, so the
// history never forms a global matrix: region r's rows of it are
//     hist_r = sum_k ( G^r_k z^t_{r,k} - H^r_k z^u_{r,k} )
// over region r's own FP64 shell lag store (k_conv), with the region's displacements and
// tractions (its interior loads, +-tau on the interface) kept in a ring exactly as the
// single-region march keeps them (march.hip).  The z's are those of Solve_Direct (:1035-1099,
// a1 = a2 = 0, a3 = 1: z_k = x_{s-k}) or of Solve_Averaging (:1101-1249, a1 = a2 = 1, a3 = 2:
// the (1, 2, 1) weights of its three loops, with 4 M_0 + M_1 as the step matrix; its first
// solve uses M_0 alone).  Each step: two convolutions, r = f gF + hist, x = A^{-1} r with the
// inverse of the global step matrix in FP64-MFMA fragment order, then one kernel that writes the
// global unknowns, the region fields of ExtractField (:1251-1301) and the two rings.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <memory>
#include <stdexcept>
#include <vector>

#include "step_device.h"

namespace tdbem {

struct Mrt {
    Context* c[2] = {nullptr, nullptr};
    hipStream_t stream = nullptr;
    int K = 10, nz = 0, N[2] = {0, 0}, NH = 0, numstep = 0;
    double dt = 0;
    std::vector<int> h_hcol, h_tcol;   // per region dof (region 2's after region 1's)
    DBuf<int> hcol, tcol;              // column of the dof's displacement / traction (-1: none)
    // march state
    bool ready = false;
    int scheme = 0, load_kind = 0, s = 0, R = 0, Np = 0, nsplit = 1;
    double amp = 1, t_rise = 0, a1 = 0, a2 = 0, a3 = 1;
    DBuf<double> lv;       // [NH] interior traction loads per region dof (0 on the interface)
    DBuf<double> gF;       // [NH] G_c lv of the step matrix
    DBuf<double> w;        // [NH] M_0^{-1} (G_0 lv): the first solve of Solve_Averaging
    DBuf<double> Af, Xf, parts, hist;
    DBuf<double> ring[2];  // [R][2][N_r]
    DBuf<double> X, U, T;  // [numstep + 1][NH]: global unknowns; region fields u, t
    double init_ms = 0, run_ms = 0;

    ~Mrt()
    {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

inline unsigned blocks(int64_t n, int b = 256) { return (unsigned)((n + b - 1) / b); }

// out = c0 a + c1 b (b may be null)
__global__ void k_lincomb(const double* __restrict__ a, const double* __restrict__ b, double c0,
                          double c1, int64_t n, double* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = c0 * a[t] + (b ? c1 * b[t] : 0.0);
}

// Region block of the global step matrix: row row0 + i, column hcol[d] gets H_c[i][d]; an
// interface dof also puts sign * G_c[i][d] in its traction column.
__global__ void k_mrt_place(const double* __restrict__ Gc, const double* __restrict__ Hc, int Nr,
                            int row0, const int* __restrict__ hcol, const int* __restrict__ tcol,
                            double sign, int NH, double* __restrict__ A)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nr * Nr) return;
    const int i = (int)(t / Nr), d = (int)(t % Nr);
    double* row = A + (int64_t)(row0 + i) * NH;
    row[hcol[d]] = Hc[t];
    if (tcol[d] >= 0) row[tcol[d]] = sign * Gc[t];
}

// r = f gF + hist in the fragment order of the step solve (one right-hand side)
__global__ void k_mrt_rhs(const double* __restrict__ gF, const double* __restrict__ hist, double f,
                          int NH, double* __restrict__ Xf)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < NH) Xf[frag_index(NH, 0, d)] = f * gF[d] + hist[d];
}

__global__ void k_mrt_first(const double* __restrict__ w, double f, int NH,
                            double* __restrict__ parts)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d < NH) parts[d] = f * w[d];
}

// Thread e: global unknown e of step s (sum of the column-split partials) and region dof e
// (region 1's dofs, then region 2's): its displacement and traction of step s (ExtractField)
// and its ring update, as finalize_elem does for one region.
__global__ void k_mrt_finalize(const double* __restrict__ parts, int nparts, int Np,
                               const int* __restrict__ hcol, const int* __restrict__ tcol,
                               const double* __restrict__ lv, double f, int N0, int N1, int s,
                               int R, double a1, double a2, double a3, double* __restrict__ X,
                               double* __restrict__ U, double* __restrict__ T,
                               double* __restrict__ ring0, double* __restrict__ ring1)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int NH = N0 + N1;
    if (e >= NH) return;
    auto xsum = [&](int j) {
        double v = 0.0;
        for (int p = 0; p < nparts; p++) v += parts[(int64_t)p * Np + j];
        return v;
    };
    const int64_t hs = (int64_t)s * NH;
    X[hs + e] = xsum(e);
    const int r = e < N0 ? 0 : 1, d = r ? e - N0 : e, Nr = r ? N1 : N0;
    const double u = xsum(hcol[e]);
    const double tr = tcol[e] >= 0 ? (r ? -xsum(tcol[e]) : xsum(tcol[e])) : lv[e] * f;
    U[hs + e] = u;
    T[hs + e] = tr;
    const double up = U[hs - NH + e], tp = T[hs - NH + e];   // step s - 1 >= 0
    double* ring = r ? ring1 : ring0;
    auto at = [&](int m, int part) -> double* {
        int sl = m % R;
        if (sl < 0) sl += R;
        return ring + ((int64_t)sl * 2 + part) * Nr + d;
    };
    *at(s - 1, 0) += a1 * u;            // y_{s-1} complete
    *at(s - 1, 1) += a1 * tr;
    *at(s, 0) = a3 * u + a2 * up;       // z_1 of step s + 1
    *at(s, 1) = a3 * tr + a2 * tp;
    *at(s + 1, 0) = (a2 - a1) * u;      // z_0 of step s + 1
    *at(s + 1, 1) = (a2 - a1) * tr;
}

double ms_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
        .count();
}

}  // namespace

void set_max_step(Context& c, int K)
{
    if (K < 0) throw std::invalid_argument("max step must be >= 0");
    const int Kc = std::min(K, c.numstep);
    c.K = K;
    if (Kc == c.Kc) return;
    march_free(c);
    c.Kc = Kc;
    c.store_ready = false;
    c.ranges_ready = false;
    c.tp.ready = false;
    c.tp.planned = false;
}

Mrt* mrt_create(Context* c1, Context* c2, int max_step, int nz, const int* if1, const int* if2)
{
    if (!c1 || !c2 || c1 == c2) throw std::invalid_argument("two distinct region contexts needed");
    if (nz < 0 || (nz > 0 && (!if1 || !if2))) throw std::invalid_argument("bad interface arrays");
    if (max_step < 0) throw std::invalid_argument("max step must be >= 0");
    if (c1->device != c2->device)
        throw std::invalid_argument("the region contexts must be on the same GPU");
    for (Context* c : {c1, c2}) {
        if (c->kernel != 4)
            throw std::invalid_argument("the two-region solver runs the Step kernel");
        if (c->precision != 0)
            throw std::invalid_argument("the two-region solver needs FP64 region contexts");
    }
    if (c1->numstep != c2->numstep || c1->dt != c2->dt)
        throw std::invalid_argument("the regions must share dt and numstep");
    std::unique_ptr<Mrt> M(new Mrt());
    M->c[0] = c1;
    M->c[1] = c2;
    M->K = max_step;
    M->nz = nz;
    M->numstep = c1->numstep;
    M->dt = c1->dt;
    M->N[0] = 3 * c1->nnd;
    M->N[1] = 3 * c2->nnd;
    M->NH = M->N[0] + M->N[1];
    const int* ifc[2] = {if1, if2};
    for (int r = 0; r < 2; r++) {
        if (nz > M->c[r]->nnd) throw std::invalid_argument("more interface nodes than nodes");
        std::vector<char> seen(M->c[r]->nnd, 0);
        for (int i = 0; i < nz; i++) {
            const int n = ifc[r][i];
            if (n < 0 || n >= M->c[r]->nnd) throw std::invalid_argument("interface node out of range");
            if (seen[n]) throw std::invalid_argument("interface node listed twice");
            seen[n] = 1;
        }
    }
    // region 2's interface nodes matched to region 1's by position (:440-477)
    std::vector<int> slot2(nz, -1), taken(nz, 0);
    for (int i = 0; i < nz; i++) {
        const double* x2 = &c2->h_coords[3 * (size_t)if2[i]];
        for (int j = 0; j < nz; j++) {
            const double* x1 = &c1->h_coords[3 * (size_t)if1[j]];
            const double dx = x1[0] - x2[0], dy = x1[1] - x2[1], dz = x1[2] - x2[2];
            if (std::sqrt(dx * dx + dy * dy + dz * dz) < 1e-10) {
                slot2[i] = j;
                break;
            }
        }
        if (slot2[i] < 0 || taken[slot2[i]])
            throw std::invalid_argument("Interface mismatch, interface nodes must match one-to-one");
        taken[slot2[i]] = 1;
    }
    // column maps: interior dofs in ascending node order (:235-246, :359-370)
    const int nint0 = M->N[0] - 3 * nz, nint1 = M->N[1] - 3 * nz;
    M->h_hcol.assign(M->NH, -1);
    M->h_tcol.assign(M->NH, -1);
    for (int r = 0; r < 2; r++) {
        std::vector<int> lab(M->c[r]->nnd, -1);
        for (int i = 0; i < nz; i++) lab[ifc[r][i]] = i;
        int k = r == 0 ? 0 : nint0;
        const int base = r == 0 ? 0 : M->N[0];
        for (int n = 0; n < M->c[r]->nnd; n++)
            for (int d = 0; d < 3; d++) {
                const int e = base + 3 * n + d;
                if (lab[n] < 0) {
                    M->h_hcol[e] = k++;
                } else {
                    const int sl = r == 0 ? lab[n] : slot2[lab[n]];
                    M->h_hcol[e] = nint0 + nint1 + 3 * sl + d;
                    M->h_tcol[e] = nint0 + nint1 + 3 * nz + 3 * sl + d;
                }
            }
    }
    TD_HIP(hipSetDevice(c1->device));
    TD_HIP(hipStreamCreate(&M->stream));
    M->hcol.upload(M->h_hcol.data(), M->NH, M->stream);
    M->tcol.upload(M->h_tcol.data(), M->NH, M->stream);
    TD_HIP(hipStreamSynchronize(M->stream));
    // SetMaxStep (:1318-1321): the lag stores hold M_0 .. M_K, M_K without its step-(K+1) part
    set_max_step(*c1, max_step);
    set_max_step(*c2, max_step);
    return M.release();
}

void mrt_destroy(Mrt* M) { delete M; }

int mrt_unknowns(const Mrt& M) { return M.NH; }
int mrt_step_count(const Mrt& M) { return M.s; }

double mrt_load_factor(const Mrt& M, double t)
{
    // SetLoad (:935-939): p0 t exp(-2000 t), p0 = 5.46366e9; kind 1: the load history
    // amp * min(1, t / t_rise) of the single-region solver
    if (M.load_kind == 0) return 5.46366e9 * t * std::exp(-2000.0 * t);
    if (M.t_rise > 0) return M.amp * (t < M.t_rise ? t / M.t_rise : 1.0);
    return M.amp;
}

void mrt_init(Mrt& M, int scheme, int load_kind, double amp, double t_rise, const int* n_fbc,
              const int* const* fbc_dof, const double* const* fbc_val)
{
    if (scheme != 0 && scheme != 1) throw std::invalid_argument("scheme must be 0 or 1");
    if (load_kind != 0 && load_kind != 1) throw std::invalid_argument("load kind must be 0 or 1");
    const auto t0 = std::chrono::steady_clock::now();
    for (int r = 0; r < 2; r++) {
        Context& c = *M.c[r];
        if (c.store_ready && c.store_precision != 0) c.store_ready = false;
        if (!c.store_ready) build_lag_store(c);
    }
    M.ready = false;
    M.scheme = scheme;
    M.load_kind = load_kind;
    M.amp = amp;
    M.t_rise = t_rise;
    M.a1 = M.a2 = scheme == 1 ? 1.0 : 0.0;
    M.a3 = scheme == 1 ? 2.0 : 1.0;
    M.R = std::max(M.c[0]->Kc, M.c[1]->Kc) + 2;
    const int NH = M.NH;
    hipStream_t st = M.stream;
    // interior loads, last value of a dof wins (INSERT_VALUES of SetLoad)
    std::vector<double> lv(NH, 0.0);
    for (int r = 0; r < 2; r++) {
        const int base = r == 0 ? 0 : M.N[0];
        if (n_fbc[r] > 0 && (!fbc_dof[r] || !fbc_val[r]))
            throw std::invalid_argument("null boundary-condition array");
        for (int q = 0; q < n_fbc[r]; q++) {
            const int d = fbc_dof[r][q];
            if (d < 0 || d >= M.N[r]) throw std::invalid_argument("FBC dof out of range");
            if (M.h_tcol[base + d] < 0) lv[base + d] = fbc_val[r][q];
        }
    }
    M.lv.upload(lv.data(), NH, st);

    // region lags 0 and 1, dense
    DBuf<double> G[2][2], H[2][2], Gc, Hc;
    for (int r = 0; r < 2; r++) {
        const int64_t NN = (int64_t)M.N[r] * M.N[r];
        for (int k = 0; k < 2; k++) {
            G[r][k].alloc(NN);
            H[r][k].alloc(NN);
            lag_dense(*M.c[r], k, G[r][k].p, H[r][k].p);
        }
        TD_HIP(hipStreamSynchronize(M.c[r]->stream));
    }
    const int64_t NN0 = (int64_t)std::max(M.N[0], M.N[1]) * std::max(M.N[0], M.N[1]);
    Gc.alloc(NN0);
    Hc.alloc(NN0);
    DBuf<double> A, Ainv, tmp;
    DBuf<int> piv;
    A.alloc((size_t)NH * NH);
    piv.alloc(NH);
    // A = sum over regions of the placed c0 M_0 + c1 M_1; gF = G_c lv
    auto build = [&](double c0, double c1, double* gF) {
        TD_HIP(hipMemsetAsync(A.p, 0, A.bytes(), st));
        for (int r = 0; r < 2; r++) {
            const int Nr = M.N[r], base = r == 0 ? 0 : M.N[0];
            const int64_t NN = (int64_t)Nr * Nr;
            hipLaunchKernelGGL(k_lincomb, dim3(blocks(NN)), dim3(256), 0, st, G[r][0].p,
                               c1 != 0.0 ? G[r][1].p : nullptr, c0, c1, NN, Gc.p);
            hipLaunchKernelGGL(k_lincomb, dim3(blocks(NN)), dim3(256), 0, st, H[r][0].p,
                               c1 != 0.0 ? H[r][1].p : nullptr, c0, c1, NN, Hc.p);
            hipLaunchKernelGGL(k_mrt_place, dim3(blocks(NN)), dim3(256), 0, st, Gc.p, Hc.p, Nr,
                               base, M.hcol.p + base, M.tcol.p + base, r == 0 ? -1.0 : 1.0, NH,
                               A.p);
            TD_HIP(hipGetLastError());
            gemv(Gc.p, Nr, M.lv.p + base, gF + base, 0.0, st);
        }
    };
    M.gF.alloc(NH);
    if (scheme == 1) {
        // first solve of Solve_Averaging: M_0 x_1 = G_0 lv f
        M.w.alloc(NH);
        build(1.0, 0.0, M.w.p);
        lu_factor(A.p, NH, piv.p, st);
        lu_solve(A.p, NH, piv.p, M.w.p, 1, st);
        build(4.0, 1.0, M.gF.p);
    } else {
        build(1.0, 0.0, M.gF.p);
    }
    lu_factor(A.p, NH, piv.p, st);
    Ainv.alloc((size_t)NH * NH);
    lu_inverse(A.p, NH, piv.p, Ainv.p, st);
    M.Af.alloc((size_t)frag_elems(NH, NH));
    to_frag(Ainv.p, NH, M.Af.p, st);

    for (int r = 0; r < 2; r++) {
        M.ring[r].alloc((size_t)M.R * 2 * M.N[r]);
        TD_HIP(hipMemsetAsync(M.ring[r].p, 0, M.ring[r].bytes(), st));
    }
    M.hist.alloc(NH);
    M.Np = frag_rows(NH);
    M.nsplit = apply_splits(NH);
    M.Xf.alloc((size_t)frag_elems(NH, 1));
    TD_HIP(hipMemsetAsync(M.Xf.p, 0, M.Xf.bytes(), st));   // padding stays zero
    M.parts.alloc((size_t)M.nsplit * M.Np);
    const size_t hist_elems = (size_t)(M.numstep + 1) * NH;
    for (DBuf<double>* b : {&M.X, &M.U, &M.T}) {
        b->alloc(hist_elems);
        TD_HIP(hipMemsetAsync(b->p, 0, b->bytes(), st));
    }
    TD_HIP(hipStreamSynchronize(st));
    M.s = 0;
    M.ready = true;
    M.init_ms += ms_since(t0);
}

void mrt_run(Mrt& M, int nsteps)
{
    if (!M.ready) throw std::logic_error("two-region solver not initialised");
    const auto t0 = std::chrono::steady_clock::now();
    const int NH = M.NH;
    hipStream_t st = M.stream;
    // Solve_Direct solves x_1 .. x_numstep; Solve_Averaging solves x_{n-1} at n = 2 .. numstep
    const int last = M.scheme == 0 ? M.numstep : M.numstep - 1;
    for (int it = 0; it < nsteps && M.s < last; it++) {
        const int s = ++M.s;
        const double f = mrt_load_factor(M, (M.scheme == 0 ? s : s + 1) * M.dt);
        int nparts = M.nsplit;
        if (M.scheme == 1 && s == 1) {
            hipLaunchKernelGGL(k_mrt_first, dim3(blocks(NH)), dim3(256), 0, st, M.w.p, f, NH,
                               M.parts.p);
            nparts = 1;
        } else {
            conv_fp64(*M.c[0], M.ring[0].p, M.R, s, 1, M.hist.p, st);
            conv_fp64(*M.c[1], M.ring[1].p, M.R, s, 1, M.hist.p + M.N[0], st);
            hipLaunchKernelGGL(k_mrt_rhs, dim3(blocks(NH)), dim3(256), 0, st, M.gF.p, M.hist.p, f,
                               NH, M.Xf.p);
            apply_frag(M.Af.p, NH, M.Xf.p, 1, M.parts.p, st);
        }
        hipLaunchKernelGGL(k_mrt_finalize, dim3(blocks(NH)), dim3(256), 0, st, M.parts.p, nparts,
                           M.Np, M.hcol.p, M.tcol.p, M.lv.p, f, M.N[0], M.N[1], s, M.R, M.a1,
                           M.a2, M.a3, M.X.p, M.U.p, M.T.p, M.ring[0].p, M.ring[1].p);
        TD_HIP(hipGetLastError());
    }
    TD_HIP(hipStreamSynchronize(st));
    M.run_ms += ms_since(t0);
}

void mrt_state(Mrt& M, int s, double* x, double* u, double* t)
{
    if (!M.ready) throw std::logic_error("two-region solver not initialised");
    if (s < 0 || s > M.numstep) throw std::invalid_argument("step out of range");
    const int64_t off = (int64_t)s * M.NH;
    const size_t bytes = sizeof(double) * M.NH;
    if (x) TD_HIP(hipMemcpyAsync(x, M.X.p + off, bytes, hipMemcpyDeviceToHost, M.stream));
    if (u) TD_HIP(hipMemcpyAsync(u, M.U.p + off, bytes, hipMemcpyDeviceToHost, M.stream));
    if (t) TD_HIP(hipMemcpyAsync(t, M.T.p + off, bytes, hipMemcpyDeviceToHost, M.stream));
    TD_HIP(hipStreamSynchronize(M.stream));
}

void mrt_timings(const Mrt& M, double* init_ms, double* run_ms)
{
    *init_ms = M.init_ms;
    *run_ms = M.run_ms;
}

}  // namespace tdbem
