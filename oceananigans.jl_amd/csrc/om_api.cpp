// om_api.cpp — C ABI of the MI355X NonhydrostaticModel time step (see include/oceananigans_mi355x.h).
//
// Owns the device state of one model: prognostic fields, tendencies Gⁿ/G⁻, pressures, closure
// fields, grid metrics and the Poisson solver plan; sequences the kernels of one RK3 substage
// on the handle's stream.  Reference sequencing:
//   time_step!               src/TimeSteppers/runge_kutta_3.jl:103-168
//   rk3_substep!             src/Models/NonhydrostaticModels/nonhydrostatic_rk3_substep.jl:31-63
//   update_state!            src/Models/NonhydrostaticModels/update_nonhydrostatic_model_state.jl:22-62
//   pressure correction      src/Models/NonhydrostaticModels/pressure_correction.jl:6-106
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <limits>
#include <new>
#include <string>
#include <type_traits>
#include <vector>

#include "om_internal.h"
#include "om_batch.h"
#include "om_kernels.h"
#include "om_learned.h"

using namespace om;

namespace om {
Recorder *&thread_recorder() {
    static thread_local Recorder *r = nullptr;
    return r;
}
}  // namespace om

namespace {

std::string g_create_error;

struct HipError {
    hipError_t e;
    const char *what;
};

#define HIPCHECK(expr)                                                      \
    do {                                                                    \
        hipError_t _e = (expr);                                             \
        if (_e != hipSuccess) throw HipError{_e, #expr};                    \
    } while (0)

struct ArgError {
    std::string msg;
};

// A conjugate-gradient solve stopped at maxiter above its tolerance.  The reference stops there
// silently and the step goes on (conjugate_gradient_solver.jl:279-284); the library does the
// same and reports it when the call returns, as OM_ENOTCONV, without poisoning the handle.
struct NotConverged {
    std::string msg;
};

// ---------------------------------------------------------------------------------------------
// Grid metrics in FT (src/Grids/grid_generation.jl:34-95 stretched, :117-156 uniform)
// ---------------------------------------------------------------------------------------------
template <class T> struct HostGrid {
    int N[3], H[3], topo[3];
    T dx, dy;
    std::vector<T> dzc, dzf;   // index k + Hz (k from -Hz .. Nz + Hz), 0-based device k
    T Lx, Ly, Lz;
    bool stretched = false;
    std::vector<double> zc;   // interior cell-centre heights, device k = 0 .. Nz-1
};

template <class T> static T spacing(double L, int N) { return (T)((long double)L / (long double)N); }

template <class T> static void make_grid(HostGrid<T> &g, const om_config_t &c) {
    for (int d = 0; d < 3; ++d) { g.N[d] = c.N[d]; g.topo[d] = c.topology[d]; g.H[d] = c.topology[d] == OM_FLAT ? 0 : c.H[d]; }
    g.dx = g.topo[0] == OM_FLAT ? T(1) : spacing<T>(c.extent[0], c.N[0]);
    g.dy = g.topo[1] == OM_FLAT ? T(1) : spacing<T>(c.extent[1], c.N[1]);
    g.Lx = g.topo[0] == OM_FLAT ? T(1) : (T)c.extent[0];
    g.Ly = g.topo[1] == OM_FLAT ? T(1) : (T)c.extent[1];
    const int Nz = g.N[2], Hz = g.H[2];
    const int n = Nz + 2 * Hz + 1;
    g.dzc.assign(n, T(1));
    g.dzf.assign(n, T(1));
    if (g.topo[2] == OM_FLAT) { g.Lz = T(1); return; }
    if (!c.z_faces) {
        const T dz = spacing<T>(c.extent[2], Nz);
        for (auto &x : g.dzc) x = dz;
        for (auto &x : g.dzf) x = dz;
        g.Lz = (T)c.extent[2];
        // node positions of a regular grid (range(z₀, step = Lz/Nz), rectilinear_grid.jl)
        const long double d = (long double)c.extent[2] / Nz;
        for (int kd = 0; kd < Nz; ++kd) g.zc.push_back((double)((long double)c.origin[2] + (kd + 0.5L) * d));
        return;
    }
    g.stretched = true;
    // faces F[m] for reference face index m - Hz + 1, m = 0 .. Nz + 2Hz
    std::vector<T> Fi(Nz + 1);
    for (int q = 0; q <= Nz; ++q) Fi[q] = (T)c.z_faces[q];
    std::vector<T> lo(Hz), hi(Hz);
    for (int q = 0; q < Hz; ++q) {
        if (g.topo[2] == OM_BOUNDED) { lo[q] = Fi[1] - Fi[0]; hi[q] = Fi[Nz] - Fi[Nz - 1]; }
        else { lo[q] = Fi[Nz - Hz + q + 1] - Fi[Nz - Hz + q]; hi[q] = Fi[q + 1] - Fi[q]; }
    }
    std::vector<T> hir(hi.rbegin(), hi.rend());
    std::vector<T> F;
    for (int q = 0; q < Hz; ++q) { T s = 0; for (int r = q; r < Hz; ++r) s += lo[r]; F.push_back(Fi[0] - s); }
    for (int q = 0; q <= Nz; ++q) F.push_back(Fi[q]);
    std::vector<T> top;
    for (int q = 0; q < Hz; ++q) { T s = 0; for (int r = q; r < Hz; ++r) s += hir[r]; top.push_back(Fi[Nz] + s); }
    for (int q = Hz - 1; q >= 0; --q) F.push_back(top[q]);
    const int TC = Nz + 2 * Hz;
    std::vector<T> C(TC);
    for (int q = 0; q < TC; ++q) C[q] = (F[q + 1] + F[q]) / T(2);
    // device index kd = reference k - 1; array slot kd + Hz
    for (int kd = -Hz; kd < Nz + Hz; ++kd) g.dzc[kd + Hz] = F[kd + Hz + 1] - F[kd + Hz];
    for (int kd = -Hz + 1; kd < Nz + Hz; ++kd) g.dzf[kd + Hz] = C[kd + Hz] - C[kd + Hz - 1];
    for (int kd = 0; kd < Nz; ++kd) g.zc.push_back((double)C[kd + Hz]);
    g.dzf[0] = g.dzf[1];
    g.dzc[n - 1] = g.dzc[n - 2];
    g.dzf[n - 1] = g.dzf[n - 2];
    g.Lz = Fi[Nz] - Fi[0];
}

// poisson_eigenvalues.jl:8-32 (evaluated in double with L/N rounded in FT, then rounded to FT)
template <class T> static std::vector<T> eigen(int N, T L, int topo, int extension) {
    std::vector<T> lam(N * extension, T(0));
    const double h = (double)(T)(L / T(N));
    for (int m = 0; m < N * extension; ++m) {
        double v = 0;
        if (topo == OM_PERIODIC) v = std::pow(2 * std::sin(m * M_PI / N) / h, 2);
        else if (topo == OM_BOUNDED) {
            int mm = m < N ? m : 2 * N - m;   // even extension: λ(2N - m) = λ(m)
            v = std::pow(2 * std::sin(mm * M_PI / (2.0 * N)) / h, 2);
        }
        lam[m] = (T)v;
    }
    return lam;
}

template <class T> static std::vector<typename Vec2<T>::type> twiddles(int n) {
    std::vector<typename Vec2<T>::type> tw(n > 0 ? n : 1);
    for (int m = 0; m < n; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)n;
        tw[m].x = (T)std::cos(a);
        tw[m].y = (T)std::sin(a);
    }
    return tw;
}

struct ModelBase {
    virtual ~ModelBase() {}
};

template <class T> struct ModelT : ModelBase {
    om_config_t cfg;
    HostGrid<T> hg;
    DevGrid<T> g;
    DevPhysics<T> ph;
    StateView<T> s;
    PoissonPlan<T> P;
    hipStream_t st = nullptr;
    int device = 0;                    // the device current at create; the handle's stream lives there
    int nt = 0, nf = 0;
    bool has_pHY = false;
    int pnx = 0, pny = 0, pnz = 0;     // parent array size (pnz: one more for w on a Bounded z)
    size_t lead = 0;                   // elements before the parent array in a field's block
    size_t fsize = 0;                  // elements per field block
    std::vector<void *> allocs;
    T *fields_block = nullptr;
    T *rhs_dense = nullptr, *phi_dense = nullptr;
    // clock
    double time = 0, last_dt = INFINITY, last_stage_dt = INFINITY;
    int64_t iteration = 0;
    int stage = 1;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0;
    // optional per-phase timing: a pool of events recorded around each phase on the stream
    bool phase_timing = false;
    std::vector<hipEvent_t> evpool;
    size_t evused = 0;
    std::vector<int> evphase;
    double phase_ms[OM_NPHASES] = {0};
    int phase_n[OM_NPHASES] = {0};
    // conjugate-gradient Poisson solver work arrays (pressure_solver == OM_SOLVER_CONJUGATE_GRADIENT)
    bool cg = false;
    double *cg_r = nullptr, *cg_b = nullptr;       // residual and right-hand side (FP64, padded layout)
    double *cg_x = nullptr, *cg_p = nullptr, *cg_q = nullptr;   // iterate, search direction, A·p (FP64)
    T *cg_z = nullptr;                             // preconditioned residual (model precision)
    T *cg_t = nullptr;                             // Thomas-sweep ratios of the columnwise preconditioner
    double *diag_partial = nullptr, *diag_out = nullptr, *diag_host = nullptr;   // diagnostics reductions
    double *cg_partial = nullptr, *cg_out = nullptr, *cg_host = nullptr;
    int cg_iterations = 0;
    double cg_res_last = 0, cg_true_res_last = 0;
    int cg_unconverged = 0;            // solves of the current call that stopped at maxiter above tolerance
    std::string cg_unconverged_msg;
    // learned correction to the FFT preconditioner (OM_PRECOND_LEARNED) and its training data
    NetState net;
    float *ln_in3 = nullptr, *ln_maskd = nullptr;  // current network input (3 channels, dense), active mask
    double ln_nactive = 0;
    double *cg_rprev = nullptr;                    // r of the previous iteration (flexible CG)
    double *cg_xsave = nullptr;                    // the warm start of a learned solve, for a restart
    int ln_learned_its = 0, ln_fell_back = 0;
    int log_iters = 0, log_cap = 0, log_count = 0, log_next = 0;   // pair logging (om_learned_log)
    int64_t log_total = 0;                                         // pairs logged since om_learned_log
    std::vector<float *> pair_in3, pair_tgt;       // logged pairs (device, dense)
    std::vector<double> pair_energy;               // tᵀ(−V∇²)t of each pair's target
    float *ln_gout = nullptr;                      // dense output gradient of the energy loss
    std::vector<double *> snap_x;                  // iterates of the logged iterations of the current solve
    std::vector<float *> snap_in3;
    std::vector<double> snap_isz;
    // launch-shape choices (om_set_tuning)
    int tune_tendency_kernel = 0;      // 0: tiled z-marching kernel where supported, 1: thread per cell
    int tune_kchunk = 16;              // z levels per block of the tiled tendency kernel
    int tune_kchunk_closure = 16;      // z levels per block of the tiled closure kernel
    int tune_learned_iterations = 0;   // leading CG iterations that apply the learned correction (0: all)
    int tune_step_graphs = 1;          // om_time_step replays captured two-step graphs
    int tune_refine = 0;               // FP64 iterative refinement of FFT / Fourier–tridiagonal solves
    RefineOp refop{};                  // the solver's operator for the refinement residual
    T *ref_r = nullptr, *ref_d = nullptr;   // dense residual and correction
    double *ref_b64 = nullptr;         // FP64 right-hand side of om_poisson_solve

    struct PhaseScope {
        ModelT *M;
        int phase;
        size_t slot;
        PhaseScope(ModelT *m, int p) : M(m), phase(p), slot(0) {
            if (!M->phase_timing) return;
            if (M->evused + 2 > M->evpool.size()) {
                for (int q = 0; q < 256; ++q) { hipEvent_t e; HIPCHECK(hipEventCreate(&e)); M->evpool.push_back(e); }
                M->evphase.resize(M->evpool.size());
            }
            slot = M->evused;
            M->evused += 2;
            M->evphase[slot] = p;
            HIPCHECK(hipEventRecord(M->evpool[slot], M->st));
        }
        ~PhaseScope() noexcept(false) {
            if (!M->phase_timing) return;
            HIPCHECK(hipEventRecord(M->evpool[slot + 1], M->st));
        }
    };
    void collect_phase_times() {
        for (int p = 0; p < OM_NPHASES; ++p) { phase_ms[p] = 0; phase_n[p] = 0; }
        for (size_t q = 0; q + 1 < evused; q += 2) {
            float ms = 0;
            HIPCHECK(hipEventElapsedTime(&ms, evpool[q], evpool[q + 1]));
            phase_ms[evphase[q]] += ms;
            phase_n[evphase[q]] += 1;
        }
        evused = 0;
    }

    ~ModelT() override {
        if (st) (void)hipStreamSynchronize(st);
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (void *p : allocs) (void)hipFree(p);
        if (cg_host) (void)hipHostFree(cg_host);
        if (diag_host) (void)hipHostFree(diag_host);
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
        for (auto e : evpool) (void)hipEventDestroy(e);
        if (st) (void)hipStreamDestroy(st);
    }

    template <class U> U *dalloc(size_t n) {
        void *p = nullptr;
        HIPCHECK(hipMalloc(&p, n * sizeof(U) + 256));
        HIPCHECK(hipMemset(p, 0, n * sizeof(U) + 256));
        allocs.push_back(p);
        return reinterpret_cast<U *>(p);
    }
    void dfree(void *p) {
        for (size_t q = 0; q < allocs.size(); ++q)
            if (allocs[q] == p) { HIPCHECK(hipFree(p)); allocs.erase(allocs.begin() + q); return; }
    }
    template <class U> const U *upload(const std::vector<U> &h) {
        U *d = dalloc<U>(h.size());
        HIPCHECK(hipMemcpy(d, h.data(), h.size() * sizeof(U), hipMemcpyHostToDevice));
        return d;
    }

    // Field layout: rows of sy elements (pnx rounded up to 128 bytes), planes of sz = sy·pny, and a
    // lead of `lead` elements before the parent array so that every interior row starts on a
    // 128-byte line (an interior row of a kernel then touches Nx·size/128 lines, not one more).
    // Element strides are shared by the T fields and the FP64 CG vectors.
    template <class U> U *field_origin(U *base) const { return base + lead + (size_t)g.Hx + (size_t)g.Hy * g.sy + (size_t)g.Hz * g.sz; }
    T *field_base(const T *origin) const { return const_cast<T *>(origin) - ((size_t)g.Hx + (size_t)g.Hy * g.sy + (size_t)g.Hz * g.sz); }
    size_t parent_index(int x, int y, int z) const { return (size_t)x + (size_t)g.sy * y + (size_t)g.sz * z; }

    void init(const om_config_t &c) {
        cfg = c;
        nt = c.n_tracers;
        nf = 3 + nt;
        make_grid(hg, c);
        HIPCHECK(hipGetDevice(&device));
        HIPCHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        HIPCHECK(hipEventCreate(&ev0));
        HIPCHECK(hipEventCreate(&ev1));
        // device grid
        g.Nx = hg.N[0]; g.Ny = hg.N[1]; g.Nz = hg.N[2];
        g.Hx = hg.H[0]; g.Hy = hg.H[1]; g.Hz = hg.H[2];
        g.tx = hg.topo[0]; g.ty = hg.topo[1]; g.tz = hg.topo[2];
        pnx = g.Nx + 2 * g.Hx; pny = g.Ny + 2 * g.Hy; pnz = g.Nz + 2 * g.Hz + 1;
        {
            const int line = (int)(128 / sizeof(T));   // elements per 128-byte line
            g.sy = (pnx + line - 1) / line * line;
            g.sz = g.sy * pny;
            lead = (size_t)((line - g.Hx % line) % line);
        }
        g.dx = hg.dx; g.dy = hg.dy; g.idx = T(1) / hg.dx; g.idy = T(1) / hg.dy;
        g.Az = hg.dx * hg.dy; g.iAz = T(1) / g.Az;
        std::vector<T> idzc(hg.dzc.size()), idzf(hg.dzf.size());
        for (size_t q = 0; q < idzc.size(); ++q) { idzc[q] = T(1) / hg.dzc[q]; idzf[q] = T(1) / hg.dzf[q]; }
        g.dzc = upload(hg.dzc) + g.Hz;
        g.dzf = upload(hg.dzf) + g.Hz;
        g.idzc = upload(idzc) + g.Hz;
        g.idzf = upload(idzf) + g.Hz;
        {
            std::vector<double> idzf64(hg.dzf.size());
            for (size_t q = 0; q < idzf64.size(); ++q) idzf64[q] = 1.0 / (double)hg.dzf[q];
            g.idzf64 = upload(idzf64) + g.Hz;
        }
        g.stretched = hg.stretched;
        g.amdr = nullptr;
        if (g.tz != OM_FLAT) {
            std::vector<T> amd((size_t)(g.Nz + 2) * 5);
            const T Dfx = T(2) * hg.dx, Dfy = T(2) * hg.dy;
            for (int k = -1; k <= g.Nz; ++k) {
                const T Dfz = T(2) * hg.dzc[(size_t)(k + g.Hz)];
                T *e = &amd[(size_t)(k + 1) * 5];
                e[0] = Dfx / Dfz; e[1] = Dfz / Dfx; e[2] = Dfy / Dfz; e[3] = Dfz / Dfy;
                e[4] = T(3) / ((T(1) / (Dfx * Dfx) + T(1) / (Dfy * Dfy)) + T(1) / (Dfz * Dfz));
            }
            g.amdr = upload(amd);
        }
        // GridFittedBottom: kbot = the number of cells of a column whose centre lies at or below
        // the bottom height, compared in FT (grid_fitted_bottom.jl:118-147)
        g.kbot = nullptr;
        g.kbot_top = 0;
        g.kfix = nullptr;
        g.kfix_max = 0;
        g.fix_list = nullptr;
        g.fix_count = 0;
        if (c.bottom_height) {
            std::vector<int> kb((size_t)g.Nx * g.Ny, 0);
            for (int j = 0; j < g.Ny; ++j)
                for (int i = 0; i < g.Nx; ++i) {
                    const T zb = (T)c.bottom_height[(size_t)i + (size_t)g.Nx * j];
                    int n = 0;
                    for (int kd = 0; kd < g.Nz; ++kd)
                        if ((T)hg.zc[kd] <= zb) n = kd + 1;
                    kb[(size_t)i + (size_t)g.Nx * j] = n;
                }
            g.kbot = upload(kb);
            g.kbot_top = *std::max_element(kb.begin(), kb.end());
            // kfix: the window maximum of kbot over ±kImmersedReach columns (periodic wrap), plus the
            // reach, on columns whose window holds an immersed cell; max filter along x, then along y
            const int R = kImmersedReach, Nx = g.Nx, Ny = g.Ny;
            std::vector<int> wx(kb.size()), kf(kb.size());
            auto wrap = [](int a, int n) { return ((a % n) + n) % n; };
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i) {
                    int m = 0;
                    for (int d = -R; d <= R; ++d) m = std::max(m, kb[(size_t)wrap(i + d, Nx) + (size_t)Nx * j]);
                    wx[(size_t)i + (size_t)Nx * j] = m;
                }
            int kfm = 0;
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i) {
                    int m = 0;
                    for (int d = -R; d <= R; ++d) m = std::max(m, wx[(size_t)i + (size_t)Nx * wrap(j + d, Ny)]);
                    const int f = m > 0 ? std::min(g.Nz, m + R) : 0;
                    kf[(size_t)i + (size_t)Nx * j] = f;
                    kfm = std::max(kfm, f);
                }
            g.kfix = upload(kf);
            g.kfix_max = kfm;
            // the fix-up launches' blocks: every (block, level) holding a cell below kfix, level-major
            const int nbx = (Nx + kGenericBlockX - 1) / kGenericBlockX, nby = (Ny + kGenericBlockY - 1) / kGenericBlockY;
            if (nbx > 0xffff || nby > 0x7fff) throw ArgError{"grid too wide for the immersed fix-up list"};
            std::vector<int> bmax((size_t)nbx * nby, 0);
            for (int j = 0; j < Ny; ++j)
                for (int i = 0; i < Nx; ++i) {
                    int &m = bmax[(size_t)(i / kGenericBlockX) + (size_t)nbx * (j / kGenericBlockY)];
                    m = std::max(m, kf[(size_t)i + (size_t)Nx * j]);
                }
            std::vector<int2> fl;
            for (int k = 0; k < kfm; ++k)
                for (int by = 0; by < nby; ++by)
                    for (int bx = 0; bx < nbx; ++bx)
                        if (bmax[(size_t)bx + (size_t)nbx * by] > k) fl.push_back(make_int2(bx | (by << 16), k));
            g.fix_count = (int)fl.size();
            if (!fl.empty()) g.fix_list = upload(fl);
        }
        // physics
        ph.advection = c.advection; ph.closure = c.closure; ph.buoyancy = c.buoyancy; ph.coriolis = c.coriolis;
        ph.n_tracers = nt; ph.b_tracer = c.buoyancy_tracer; ph.T_tracer = c.T_tracer; ph.S_tracer = c.S_tracer;
        ph.f = (T)c.f; ph.g = (T)c.gravitational_acceleration; ph.alpha = (T)c.thermal_expansion; ph.beta = (T)c.haline_contraction;
        ph.c0 = (T)c.closure_params[0]; ph.c1 = (T)c.closure_params[1]; ph.c2 = (T)c.closure_params[2];
        for (int f = 0; f < kMaxFields; ++f) {
            ph.bc_bottom[f] = c.bc_bottom[f]; ph.bc_top[f] = c.bc_top[f];
            ph.bc_bottom_value[f] = (T)c.bc_bottom_value[f]; ph.bc_top_value[f] = (T)c.bc_top_value[f];
        }
        if (c.hydrostatic_pressure < 0)
            has_pHY = c.topology[2] == OM_BOUNDED && c.buoyancy != OM_BUOYANCY_NONE;
        else has_pHY = c.hydrostatic_pressure == 1;
        ph.hydrostatic = has_pHY ? 1 : 0;
        // fields: one block; order U[nf], Gn[nf], Gm[nf], pNHS, pHY, nu, kap[nt]
        fsize = lead + (size_t)g.sz * pnz;
        fsize = (fsize + 63) & ~(size_t)63;   // 256-byte aligned field blocks
        const bool diff_pHY = sizeof(T) == 4 && has_pHY;   // FP32: pHY′ differences as fields of their own
        const int nblocks = 3 * nf + 3 + nt + (diff_pHY ? 2 : 0);
        fields_block = dalloc<T>(fsize * nblocks);
        memset(&s, 0, sizeof(s));
        int q = 0;
        for (int f = 0; f < nf; ++f) s.U[f] = field_origin(fields_block + fsize * (q++));
        for (int f = 0; f < nf; ++f) s.Gn[f] = field_origin(fields_block + fsize * (q++));
        for (int f = 0; f < nf; ++f) s.Gm[f] = field_origin(fields_block + fsize * (q++));
        s.pNHS = field_origin(fields_block + fsize * (q++));
        s.pHY = field_origin(fields_block + fsize * (q++));
        s.nu = field_origin(fields_block + fsize * (q++));
        for (int n = 0; n < nt; ++n) s.kap[n] = field_origin(fields_block + fsize * (q++));
        if (diff_pHY) {
            s.dxpHY = field_origin(fields_block + fsize * (q++));
            s.dypHY = field_origin(fields_block + fsize * (q++));
        }
        // Poisson plan
        const size_t ncell = (size_t)g.Nx * g.Ny * g.Nz;
        rhs_dense = dalloc<T>(ncell);
        phi_dense = dalloc<T>(ncell);
        const int Mx = g.Nx / 2 + 1;
        {   // spectral rows padded to whole 128-byte lines, so the y, z and tridiagonal passes' groups
            // of consecutive kx read aligned lines (rows narrower than a line stay unpadded)
            const int line = (int)(128 / sizeof(typename Vec2<T>::type));
            P.Ms = Mx >= line ? (Mx + line - 1) / line * line : Mx;
        }
        P.S = dalloc<typename Vec2<T>::type>((size_t)P.Ms * g.Ny * g.Nz);
        const int xline = (g.Nx % 2 == 0) ? g.Nx / 2 : g.Nx;
        P.tw_xh = upload(twiddles<T>(xline));
        {
            std::vector<typename Vec2<T>::type> post(Mx);
            for (int k = 0; k < Mx; ++k) {
                const double a = -2.0 * M_PI * (double)k / (double)g.Nx;
                post[k].x = (T)std::cos(a);
                post[k].y = (T)std::sin(a);
            }
            P.post = upload(post);
        }
        P.tw_y = upload(twiddles<T>(g.Ny));
        // the Fourier–tridiagonal plan serves the FT solver and, on stretched z, the CG preconditioner
        P.tridiagonal = (c.pressure_solver == OM_SOLVER_FOURIER_TRIDIAGONAL ||
                         (c.pressure_solver == OM_SOLVER_CONJUGATE_GRADIENT && hg.stretched)) ? 1 : 0;
        P.zmode = g.tz != OM_BOUNDED ? 0 : (g.Nz % 2 == 0 ? 1 : 2);
        P.Lz_line = P.zmode == 2 ? 2 * g.Nz : g.Nz;
        P.Lz_norm = g.tz == OM_BOUNDED ? 2 * g.Nz : g.Nz;
        P.tw_z = upload(twiddles<T>(P.Lz_line));
        {
            std::vector<typename Vec2<T>::type> c(g.Nz);
            for (int k = 0; k < g.Nz; ++k) {
                const double a = -M_PI * (double)k / (2.0 * g.Nz);
                c[k].x = (T)std::cos(a);
                c[k].y = (T)std::sin(a);
            }
            P.tw_dct = upload(c);
        }
        P.lx = upload(eigen<T>(g.Nx, hg.Lx, g.tx, 1));
        P.ly = upload(eigen<T>(g.Ny, hg.Ly, g.ty, 1));
        P.lz = upload(eigen<T>(g.Nz, hg.Lz, g.tz, g.tz == OM_BOUNDED ? 2 : 1));
        P.eps10 = T(10) * std::numeric_limits<T>::epsilon();
        {
            // the inverse squared spacings behind the solver's eigenvalues (eigen(): h = T(L / N))
            auto ih2 = [](T L, int N, int topo) { if (topo == OM_FLAT) return 0.0; const double h = (double)(T)(L / T(N)); return 1.0 / (h * h); };
            refop.ihx2 = ih2(hg.Lx, g.Nx, g.tx);
            refop.ihy2 = ih2(hg.Ly, g.Ny, g.ty);
            refop.ihz2 = P.tridiagonal ? -1.0 : ih2(hg.Lz, g.Nz, g.tz);
            if (P.tridiagonal) {
                // FP64 spacings of the column from the configured face heights (uniform without them)
                const int Nz = g.Nz;
                std::vector<double> zf(Nz + 1), zcen(Nz), idzc(Nz), idzf(Nz + 1, 0.0);
                for (int k = 0; k <= Nz; ++k) zf[k] = c.z_faces ? c.z_faces[k] : c.origin[2] + c.extent[2] * k / Nz;
                for (int k = 0; k < Nz; ++k) {
                    idzc[k] = 1.0 / (zf[k + 1] - zf[k]);
                    zcen[k] = 0.5 * (zf[k] + zf[k + 1]);
                }
                for (int k = 1; k < Nz; ++k) idzf[k] = 1.0 / (zcen[k] - zcen[k - 1]);
                refop.idzc = upload(idzc);
                refop.idzf = upload(idzf);
            }
        }
        diag_partial = dalloc<double>(3 * 1024);
        diag_out = dalloc<double>(4);
        HIPCHECK(hipHostMalloc((void **)&diag_host, 4 * sizeof(double), hipHostMallocDefault));
        if (c.pressure_solver == OM_SOLVER_CONJUGATE_GRADIENT) {
            cg = true;
            cg_r = field_origin(dalloc<double>(fsize)); cg_b = field_origin(dalloc<double>(fsize));
            cg_x = field_origin(dalloc<double>(fsize)); cg_p = field_origin(dalloc<double>(fsize));
            cg_q = field_origin(dalloc<double>(fsize)); cg_z = field_origin(dalloc<T>(fsize));
            cg_partial = dalloc<double>((size_t)cg_partials_needed(g.Nx, g.Ny, g.Nz));
            cg_out = dalloc<double>(kCgSlots);
            if (cg_precond() == OM_PRECOND_COLUMNWISE_TRIDIAGONAL) cg_t = dalloc<T>((size_t)g.Nx * g.Ny * g.Nz);
            if (cg_precond() == OM_PRECOND_LEARNED) {
                cg_rprev = field_origin(dalloc<double>(fsize));
                cg_xsave = field_origin(dalloc<double>(fsize));
                net_alloc();
            }
            HIPCHECK(hipHostMalloc((void **)&cg_host, 4 * sizeof(double), hipHostMallocDefault));
        }
        HIPCHECK(hipDeviceSynchronize());
    }

    Ctx<T> ctx() const {
        Ctx<T> C;
        C.g = g; C.s = s; C.ph = ph;
        const int nt = ph.n_tracers;
        C.s.btr = (ph.b_tracer >= 0 && ph.b_tracer < nt) ? s.U[3 + ph.b_tracer] : nullptr;
        C.s.Ttr = (ph.T_tracer >= 0 && ph.T_tracer < nt) ? s.U[3 + ph.T_tracer] : nullptr;
        C.s.Str = (ph.S_tracer >= 0 && ph.S_tracer < nt) ? s.U[3 + ph.S_tracer] : nullptr;
        return C;
    }

    // ---- halo fills
    // p's halo fill deferred by a pressure correction that read no halo of p: it joins the next
    // halo launch (p is not written in between), or flush_p_halo fills it at the end of a call
    bool p_halo_pending = false;
    void add_pending_p(HaloJob<T> &J) {
        if (!p_halo_pending) return;
        for (int n = 0; n < J.nf; ++n) if (J.f[n] == s.pNHS) { p_halo_pending = false; return; }
        const int n = J.nf++;
        J.f[n] = s.pNHS;
        J.kind_bottom[n] = OM_BC_DEFAULT; J.kind_top[n] = OM_BC_DEFAULT;
        p_halo_pending = false;
    }
    void flush_p_halo() { if (p_halo_pending) halos_aux({}); }
    // the velocities' halo fill before the Poisson solve, deferred inside a time step when the
    // solve's source reads no halo: it joins the next prognostic fill (update_state's), with the
    // impenetrable w faces set there; nothing reads the velocity halos in between
    bool u_halo_pending = false;
    void halos_prognostic(uint32_t mask, bool fill_normal) {
        if (u_halo_pending) { mask |= 0x7u; fill_normal = true; u_halo_pending = false; }
        HaloJob<T> J;
        memset(&J, 0, sizeof(J));
        for (int f = 0; f < nf; ++f) {
            if (!(mask & (1u << f))) continue;
            const int n = J.nf++;
            J.f[n] = s.U[f];
            if (f == 2) { J.kind_bottom[n] = fill_normal ? -1 : -2; J.kind_top[n] = J.kind_bottom[n]; }
            else {
                J.kind_bottom[n] = cfg.bc_bottom[f]; J.kind_top[n] = cfg.bc_top[f];
                J.val_bottom[n] = (T)cfg.bc_bottom_value[f]; J.val_top[n] = (T)cfg.bc_top_value[f];
            }
        }
        add_pending_p(J);
        PhaseScope ps(this, OM_PHASE_HALOS);
        launch_halos(g, J, st);
    }
    void halos_aux(std::initializer_list<T *> fs) {
        HaloJob<T> J;
        memset(&J, 0, sizeof(J));
        for (T *p : fs) {
            const int n = J.nf++;
            J.f[n] = p;
            J.kind_bottom[n] = OM_BC_DEFAULT; J.kind_top[n] = OM_BC_DEFAULT;
        }
        add_pending_p(J);
        if (J.nf == 0) return;
        PhaseScope ps(this, OM_PHASE_HALOS);
        launch_halos(g, J, st);
    }
    uint32_t all_mask() const { return (1u << nf) - 1u; }

    // ---- phases
    void compute_tendencies() {
        PhaseScope ps(this, OM_PHASE_TENDENCIES);
        const Ctx<T> C = ctx();
        if (tune_tendency_kernel == 0 && tiled_tendencies_supported(C)) {
            launch_tend_tiled(C, tune_kchunk, st);
            if (g.kbot) launch_tendencies(C, st, true);   // cells within reach of the immersed bottom
        } else {
            launch_tendencies(C, st);
        }
    }
    void update_state() {
        if (g.kbot) {   // mask immersed tracers (update_nonhydrostatic_model_state.jl:30-33)
            PhaseScope ps(this, OM_PHASE_HALOS);
            for (int n = 0; n < nt; ++n) launch_mask(g, s.U[3 + n], 0, st);
        }
        halos_prognostic(all_mask(), false);
        // νₑ, κₑ and pHY′ all take default (copy) halo fills, and neither the closure nor the
        // pHY′ kernel reads the others: one fill after both serves the three
        HaloJob<T> J;
        memset(&J, 0, sizeof(J));
        auto add = [&](T *f) { J.f[J.nf] = f; J.kind_bottom[J.nf] = OM_BC_DEFAULT; J.kind_top[J.nf] = OM_BC_DEFAULT; J.nf++; };
        if (cfg.closure == OM_CLOSURE_AMD || cfg.closure == OM_CLOSURE_SMAGORINSKY_LILLY) {
            {
                PhaseScope ps(this, OM_PHASE_CLOSURE);
                const Ctx<T> C = ctx();
                if (tune_tendency_kernel == 0 && tiled_closure_supported(C)) {
                    launch_closure_tiled(C, tune_kchunk_closure, st);
                    if (g.kbot) launch_closure(C, st, true);   // cells within reach of the immersed bottom
                } else {
                    launch_closure(C, st);
                }
            }
            add(s.nu);
            if (cfg.closure == OM_CLOSURE_AMD) for (int n = 0; n < nt; ++n) add(s.kap[n]);
        }
        if (has_pHY) {
            { PhaseScope ps(this, OM_PHASE_HYDROSTATIC); launch_hydrostatic(ctx(), st); }
            add(s.pHY);
        }
        if (J.nf > 0) { PhaseScope ps(this, OM_PHASE_HALOS); launch_halos(g, J, st); }
        compute_tendencies();
    }
    // defer: called inside a time step, whose update_state fills the velocity halos next
    void pressure_correction(T dtau, bool defer = false) {
        if (g.kbot) {   // mask immersed velocities (pressure_correction.jl:7)
            PhaseScope ps(this, OM_PHASE_HALOS);
            for (int f = 0; f < 3; ++f) launch_mask(g, s.U[f], f + 1, st);
        }
        if (defer && !cg && !tune_refine && g.tz == OM_BOUNDED && poisson_source_reads_no_halo(P, ctx()))
            u_halo_pending = true;
        else
            halos_prognostic(0x7u, true);
        const T eps = std::numeric_limits<T>::epsilon();
        const T dtp = dtau > eps ? dtau : eps;
        if (cg) {
            // solve_for_pressure.jl:149-161: warm start from p·Δt★, b = active·δ·V, solve V∇²ϕ = b;
            // then U −= ∇ϕ and p = ϕ / max(eps, Δt) (pressure_correction.jl:82-106)
            {
                PhaseScope ps(this, OM_PHASE_POISSON);
                const T dts = std::isfinite(dtau) ? dtp : T(0);
                cg_xio(g, s.pNHS, cg_x, dts, 0, st);
                cg_source(ctx(), (const T *)nullptr, cg_b, st);
                cg_solve();
                cg_xio(g, s.pNHS, cg_x, T(1), 1, st);
            }
            halos_aux({s.pNHS});
            {
                PhaseScope ps(this, OM_PHASE_CORRECTION);
                launch_correct(ctx(), T(1), 1, st);
                cg_scale(g, s.pNHS, dtp, 1, st);
            }
            return;
        }
        {
            PhaseScope ps(this, OM_PHASE_POISSON);
            launch_poisson(P, ctx(), (const T *)nullptr, s.pNHS, 0, dtp, st);
            if (tune_refine) refine(s.pNHS, g.sy, g.sz, nullptr, (double)dtp);
        }
        // periodic x and y without an immersed bottom or a refinement pass: the correction reads
        // p's neighbours across the edges by wrapped index, and p's halo fill joins the next
        // halo launch (update_state's), one launch fewer per substage
        const bool wrap = !tune_refine && !g.kbot && g.tx == OM_PERIODIC && g.ty == OM_PERIODIC &&
                          (g.tz == OM_BOUNDED || g.tz == OM_PERIODIC);
        if (!wrap) halos_aux({s.pNHS});
        { PhaseScope ps(this, OM_PHASE_CORRECTION); launch_correct(ctx(), dtp, 0, st, wrap ? 1 : 0); }
        p_halo_pending = wrap;
    }

    // ---- conjugate-gradient solve of V∇²x = b (conjugate_gradient_solver.jl:177-284) with the
    // zero-mean gauge; x holds the initial guess, cg_b the right-hand side.  ρ, p·Ap and the gauge
    // sums stay on the device (the kernels that use them form α, β and the means); the host reads
    // the norm once per iteration for the convergence test.
    double cg_read_slot(int slot) {
        HIPCHECK(hipMemcpyAsync(cg_host, cg_out + slot, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return cg_host[0];
    }
    double cg_read(int n) {
        HIPCHECK(hipMemcpyAsync(cg_host, cg_out, n * sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        return cg_host[0];
    }
    int cg_precond() const {
        const int pk = cfg.cg_preconditioner;
        // the reference's default: the FFT solver of the underlying grid on an immersed grid
        // (conjugate_gradient_poisson_solver.jl:141-147), the diagonally dominant one otherwise
        if (pk == OM_PRECOND_DEFAULT) return g.kbot ? OM_PRECOND_FFT : OM_PRECOND_DIAGONALLY_DOMINANT;
        return pk;
    }
    // ---- learned preconditioner: network buffers (allocated on first use; zero weights give a
    // zero correction, i.e. the FFT preconditioner)
    void net_alloc() {
        if (net.w) return;
        if (!net_dims_ok(g.Nx, g.Ny, g.Nz))
            throw ArgError{"the learned preconditioner's four levels need Nx a multiple of 128 and Ny, Nz multiples of 8"};
        if (g.tz != OM_BOUNDED) throw ArgError{"the learned preconditioner needs a Bounded z"};
        net.d = NetDims{g.Nx, g.Ny, g.Nz};
        net_place_buffers(net, dalloc<char>(net_arena_bytes(net.d)));
        net.w = dalloc<float>(NET_LAYERS * NET_WPL);
        net.m = dalloc<float>(NET_LAYERS * NET_WPL);
        net.v = dalloc<float>(NET_LAYERS * NET_WPL);
        net.dw = dalloc<float>(NET_LAYERS * NET_WPL);
        net.wf = dalloc<uint16_t>(NET_LAYERS * 14 * 64 * 8);
        net.wb = dalloc<uint16_t>(NET_LAYERS * 14 * 64 * 8);
        const size_t nc = (size_t)g.Nx * g.Ny * g.Nz;
        net.out = dalloc<float>(nc);
        net.wpartial = dalloc<float>((size_t)std::max(net.nchunk, 512) * 27 * 256);
        net.lpartial = dalloc<double>(NET_RED_BLOCKS);
        net.lsum = dalloc<double>(1);
        ln_in3 = dalloc<float>(3 * nc);
        ln_maskd = dalloc<float>(nc);
        ln_mask(g, ln_maskd, st);
        net_pack(net, st);
        std::vector<float> hm(nc);
        HIPCHECK(hipMemcpyAsync(hm.data(), ln_maskd, nc * sizeof(float), hipMemcpyDeviceToHost, st));
        HIPCHECK(hipStreamSynchronize(st));
        ln_nactive = 0;
        for (float v : hm) ln_nactive += v;
    }
    // the network input's scales 1/s_r, 1/s_z (s_r = rms(r V⁻¹), s_z = rms(z_FFT) over the active cells)
    bool ln_prepare(double &isr, double &isz) {
        ln_scales(g, cg_r, cg_z, cg_partial, cg_out, st);
        cg_read(3);
        const double na = std::max(cg_host[2], 1.0);
        const double sr = std::sqrt(cg_host[0] / na), sz = std::sqrt(cg_host[1] / na);
        if (!(sr > 0) || !(sz > 0)) return false;
        isr = 1.0 / sr;
        isz = 1.0 / sz;
        return true;
    }
    void ln_log_snapshot(int k, double isz) {
        const size_t nc = (size_t)g.Nx * g.Ny * g.Nz;
        while ((int)snap_x.size() <= k) {
            snap_x.push_back(field_origin(dalloc<double>(fsize)));
            snap_in3.push_back(dalloc<float>(3 * nc));
            snap_isz.push_back(0);
        }
        cg_copy_d(g, cg_x, snap_x[k], st);
        HIPCHECK(hipMemcpyAsync(snap_in3[k], ln_in3, 3 * nc * sizeof(float), hipMemcpyDeviceToDevice, st));
        snap_isz[k] = isz;
    }
    // after a converged solve: pairs (input of iteration k, (x_final − x_k)/s_z − ẑ_k with its mean removed)
    void ln_log_pairs(int nlogged) {
        const size_t nc = (size_t)g.Nx * g.Ny * g.Nz;
        for (int k = 0; k < nlogged; ++k) {
            if ((int)pair_in3.size() < log_cap) {
                pair_in3.push_back(dalloc<float>(3 * nc));
                pair_tgt.push_back(dalloc<float>(nc));
                pair_energy.push_back(0);
            }
            const int slot = log_next;
            log_next = (log_next + 1) % log_cap;
            log_count = std::min(log_count + 1, log_cap);
            ++log_total;
            HIPCHECK(hipMemcpyAsync(pair_in3[slot], snap_in3[k], 3 * nc * sizeof(float), hipMemcpyDeviceToDevice, st));
            ln_target(g, cg_x, snap_x[k], snap_in3[k], snap_isz[k], pair_tgt[slot], cg_partial, cg_out, st);
            cg_read(2);
            const float mean = (float)(cg_host[0] / std::max(cg_host[1], 1.0));
            ln_center(g, pair_tgt[slot], mean, st);
            ln_energy(g, (const float *)nullptr, pair_tgt[slot], cg_q, 0.0, (float *)nullptr, cg_partial, cg_out, st);
            pair_energy[slot] = cg_read(1);
        }
    }

    void note_unconverged(int64_t it, double res, double tol) {
        ++cg_unconverged;
        char buf[256];
        if (std::isfinite(res))
            snprintf(buf, sizeof buf, "conjugate-gradient solve stopped at maxiter = %lld with |r| = %.3e above the tolerance %.3e "
                     "(%d solve(s) of this call; the step completed as the reference's does)", (long long)it, res, tol, cg_unconverged);
        else
            snprintf(buf, sizeof buf, "conjugate-gradient solve stopped after %lld iterations with a non-finite residual "
                     "(%d solve(s) of this call)", (long long)it, cg_unconverged);
        cg_unconverged_msg = buf;
    }
    // throws NotConverged (after the call's work is done) when a solve of the call stopped at maxiter
    void report_unconverged() {
        if (cg_unconverged) { cg_unconverged = 0; throw NotConverged{cg_unconverged_msg}; }
    }
    void cg_solve() {
        const int pk = cg_precond();
        if (pk == OM_PRECOND_LEARNED) cg_copy_d(g, cg_x, cg_xsave, st);   // the warm start, kept for a restart
        cg_solve_with(pk);
        // the recomputed FP64 residual judges the learned preconditioner: a solve it leaves above
        // a hundred times the tolerance is redone by FFT-preconditioned CG from the warm start
        // (FFT-preconditioned CG itself ends with the recomputed residual up to ~15× the recursive
        // one: the usual drift of the recursively updated residual).  Continuing from the learned
        // iterate instead was seen to diverge on the 256×256×128 seamount.
        if (pk == OM_PRECOND_LEARNED && !(cg_true_res_last <= 100.0 * std::max(cg_tol_last, 1e-300))) {
            const int it0 = cg_iterations;
            const double tol0 = cg_tol_last;
            ln_fell_back = 1;
            cg_copy_d(g, cg_xsave, cg_x, st);
            cg_solve_with(OM_PRECOND_FFT, tol0);   // the original absolute tolerance, not one relative to the small residual left
            cg_iterations += it0;
        }
        if (pk == OM_PRECOND_LEARNED && !(cg_res_last <= cg_tol_last)) note_unconverged(cg_iterations, cg_res_last, cg_tol_last);
    }
    double cg_tol_last = 0;
    void cg_solve_with(const int pk, const double tol_override = -1.0) {
        double *x = cg_x;
        const double seps = std::sqrt((double)std::numeric_limits<T>::epsilon());
        const double reltol = cfg.cg_reltol < 0 ? (double)(T)seps : cfg.cg_reltol;
        const double abstol = cfg.cg_abstol < 0 ? (double)(T)seps : cfg.cg_abstol;
        const int64_t maxiter = cfg.cg_maxiter > 0 ? cfg.cg_maxiter : (int64_t)g.Nx * g.Ny * g.Nz;
        const double ncell = (double)g.Nx * (double)g.Ny * (double)g.Nz;
        cg_halo_r(g, x, st);
        cg_residual(g, x, cg_b, cg_r, cg_partial, cg_out, st);
        const double res0 = std::sqrt(cg_read(1));
        const double tol = tol_override >= 0 ? tol_override : std::max(reltol * res0, abstol);
        cg_tol_last = tol;
        // with the FFT or learned preconditioner the gauge sweeps also write its source r V⁻¹
        T *const fsrc = (pk == OM_PRECOND_FFT || pk == OM_PRECOND_LEARNED) ? rhs_dense : nullptr;
        // the loop's first VolumeInverseNorm (step 0: no update, so p, q and the scalars are not read)
        cg_gauge_norm(g, x, cg_p, cg_r, cg_q, cg_out, 0, 0, ncell, fsrc, cg_partial, cg_out + kCgSlotGauge, st);
        double res = std::sqrt(cg_read_slot(kCgSlotGauge));
        int64_t it = 0;
        // flexible CG for the learned preconditioner (a nonlinear map): β = zₖ·(rₖ − rₖ₋₁) / (zₖ₋₁·rₖ₋₁);
        // three iterations in a row that do not bring ‖V⁻¹r‖ below 0.999 of its best value so far hand
        // over to the FFT preconditioner for the rest of the solve
        const bool flex = pk == OM_PRECOND_LEARNED;
        bool fallback = false;
        int stall = 0, nlogged = 0;
        double res_best = res;
        const bool logging = log_iters > 0 && (pk == OM_PRECOND_FFT || pk == OM_PRECOND_LEARNED) && net.w;
        if (flex) { ln_learned_its = 0; ln_fell_back = 0; }
        while (it < maxiter && !(res <= tol)) {
            if (pk == OM_PRECOND_DIAGONALLY_DOMINANT) { cg_halo_r(g, cg_r, st); cg_dd(g, cg_r, cg_z, st); }
            else if (pk == OM_PRECOND_FFT || pk == OM_PRECOND_LEARNED) {
                launch_poisson(P, ctx(), rhs_dense, cg_z, 0, T(1), st);   // source written by the last gauge sweep
                const bool log_this = logging && it < log_iters;
                const bool learn_this = pk == OM_PRECOND_LEARNED && !fallback &&
                                        (tune_learned_iterations == 0 || it < tune_learned_iterations);
                double isr = 0, isz = 0;
                if ((log_this || learn_this) && ln_prepare(isr, isz)) {
                    if (log_this) {   // the dense input is kept with the pair
                        ln_input(g, cg_r, cg_z, isr, isz, ln_in3, st);
                        ln_log_snapshot((int)it, isz);
                        nlogged = (int)it + 1;
                    }
                    if (learn_this) {
                        if (log_this) net_input_dense(net, ln_in3, st);
                        else {
                            ln_x0(g, cg_r, cg_z, isr, isz, net.b.x0, st);
                            net_input_halo(net, st);
                        }
                        net_forward(net, net.out, st);
                        ln_correct(g, cg_z, net.out, 1.0 / isz, st);
                        ++ln_learned_its;
                    }
                }
            } else if (pk == OM_PRECOND_COLUMNWISE_TRIDIAGONAL) cg_ct(g, cg_r, cg_z, cg_t, st);
            else cg_copy_r(g, cg_r, cg_z, st);
            // ρ (and z·r_prev), p·Ap with Σp and Σq, and the gauge sweep's sums of x and r stay on the
            // device (kCgSlot*): α, β and the gauge means are formed by the kernels that use them, and
            // the host reads one norm per iteration
            const int itp = (int)(it & 1);
            if (flex && it > 0) cg_dot2(g, cg_z, cg_r, cg_rprev, cg_partial, cg_out + cg_slot_rho(itp), st);
            else cg_dot_zr(g, cg_z, cg_r, cg_partial, cg_out + cg_slot_rho(itp), st);
            cg_update_p(g, cg_p, cg_z, cg_out, it == 0 ? 0 : 2 + itp, (flex && it > 0) ? 1 : 0, st);   // with p's halos
            if (flex) cg_copy_d(g, cg_r, cg_rprev, st);
            cg_apply(g, cg_p, cg_q, cg_partial, cg_out + kCgSlotPQ, st);
            // x += αp, r −= αq, the zero-mean gauge and the norm in one sweep
            cg_gauge_norm(g, x, cg_p, cg_r, cg_q, cg_out, itp, 1, ncell, fsrc, cg_partial, cg_out + kCgSlotGauge, st);
            res = std::sqrt(cg_read_slot(kCgSlotGauge));
            it += 1;
            if (!std::isfinite(res)) break;   // reported as not converged; iterating on cannot recover
            // a learned correction that sends the flexible iteration off (‖V⁻¹r‖ a thousand times its
            // best) ends the solve: cg_solve restarts it from the warm start with the FFT preconditioner
            if (flex && res > 1e3 * res_best) break;
            if (flex && !fallback) {
                stall = res < 0.999 * res_best ? 0 : stall + 1;
                if (stall >= 3) { fallback = true; ln_fell_back = 1; }
            }
            res_best = std::min(res_best, res);
        }
        cg_iterations = (int)it;
        cg_res_last = res;
        if (!(res <= tol) && tol_override < 0 && !(pk == OM_PRECOND_LEARNED)) note_unconverged(it, res, tol);
        if (logging && nlogged > 0 && res <= tol) ln_log_pairs(nlogged);
        cg_halo_r(g, x, st);
        cg_true_residual(g, x, cg_b, cg_q, cg_partial, cg_out, st);
        const double emean = cg_read(1) / ncell;
        cg_gauged_norm(g, cg_q, emean, cg_partial, cg_out, st);
        cg_true_res_last = std::sqrt(cg_read(1));
    }
    void rk3_substep(double dt, T gamma, T zeta, bool first, bool defer = false) {
        const T kdt = (T)dt;
        const T dtau = first ? (T)(dt * (double)gamma) : (T)(dt * (double)(gamma + zeta));
        { PhaseScope ps(this, OM_PHASE_RK_UPDATE); launch_rk3(ctx(), kdt, gamma, zeta, first ? 1 : 0, nf, st); }
        pressure_correction(dtau, defer);
    }
    void set_refinement(int on) {
        if (on && (cg || g.kbot)) throw ArgError{"iterative refinement applies to the FFT and Fourier-tridiagonal solvers on a grid without an immersed bottom"};
        if (on && !ref_r) {
            const size_t n = (size_t)g.Nx * g.Ny * g.Nz;
            ref_r = dalloc<T>(n); ref_d = dalloc<T>(n); ref_b64 = dalloc<double>(n);
        }
        tune_refine = on;
    }
    // one refinement step of the solution p (strides psy, psz) of L p = b (b64, or ∇·U / divisor)
    void refine(T *p, int64_t psy, int64_t psz, const double *b64, double divisor) {
        launch_poisson_residual(ctx(), refop, p, psy, psz, b64, divisor, ref_r, st);
        launch_poisson(P, ctx(), ref_r, ref_d, 1, T(1), st);
        launch_add_correction(g, p, psy, psz, ref_d, st);
    }
    void cache_previous_tendencies() {
        // G⁻ ← Gⁿ by exchanging the buffers: the next tendency evaluation overwrites every
        // interior point of Gⁿ, so the copy is not needed.
        for (int f = 0; f < nf; ++f) { T *t = s.Gm[f]; s.Gm[f] = s.Gn[f]; s.Gn[f] = t; }
        swap_parity ^= 1;
    }
    // RK3 coefficients (runge_kutta_3.jl:36-48)
    static constexpr double kG1 = 8.0 / 15, kG2 = 5.0 / 12, kG3 = 3.0 / 4, kZ2 = -17.0 / 60, kZ3 = -5.0 / 12;
    // the device work of one time step: every launch depends only on Δt and the buffer assignment
    void step_launches(double dt) {
        const T g1 = T(kG1), g2 = T(kG2), g3 = T(kG3), z2 = T(kZ2), z3 = T(kZ3);
        rk3_substep(dt, g1, T(0), true, true);
        cache_previous_tendencies();
        update_state();
        rk3_substep(dt, g2, z2, false, true);
        cache_previous_tendencies();
        update_state();
        rk3_substep(dt, g3, z3, false, true);
        cache_previous_tendencies();
        update_state();
    }
    // the clock of one time step (runge_kutta_3.jl:103-168: tick! per substage, the last one
    // corrected so that the step ends at t + Δt)
    void step_clock(double dt) {
        const T g1 = T(kG1), g2 = T(kG2), z2 = T(kZ2);
        const double tn1 = time + dt;
        time += dt * (double)g1;
        time += dt * (double)(g2 + z2);
        const double corrected = tn1 - time;
        time += corrected; iteration += 1; stage = 1; last_dt = dt; last_stage_dt = corrected;
    }
    void time_step(double dt) {
        if (iteration == 0) {
            last_dt = dt;
            last_stage_dt = dt * (double)T(kG1);
            update_state();
        }
        step_launches(dt);
        step_clock(dt);
    }

    // ---- captured steps: two time steps recorded once into a hipGraph and replayed.  Two steps
    // return G⁻/Gⁿ to the buffers they started in (three exchanges per step), so a replay is valid
    // whenever the exchange parity matches the one at capture.  CG solves read scalars on the
    // host and are never captured.
    hipGraphExec_t graph_exec = nullptr;
    double graph_dt = 0;
    int graph_parity = 0;
    int swap_parity = 0;   // parity of the G⁻/Gⁿ exchanges so far
    bool capturable() const { return !cg; }
    int launch_version = 0;   // changes with every change of a launch choice (tuning, autotuner)
    void drop_graph() {
        ++launch_version;
        if (graph_exec) { HIPCHECK(hipGraphExecDestroy(graph_exec)); graph_exec = nullptr; }
    }
    void capture(double dt) {
        drop_graph();
        const bool pt = phase_timing;
        phase_timing = false;
        hipGraph_t gr = nullptr;
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        step_launches(dt);
        step_launches(dt);
        HIPCHECK(hipStreamEndCapture(st, &gr));
        phase_timing = pt;
        HIPCHECK(hipGraphInstantiate(&graph_exec, gr, nullptr, nullptr, 0));
        HIPCHECK(hipGraphDestroy(gr));
        graph_dt = dt;
        graph_parity = swap_parity;
    }
    // enqueue (without waiting) up to two steps of Δt; returns the number of steps issued
    int issue_steps(double dt, int left, bool use_graph) {
        if (!use_graph || !capturable() || iteration == 0 || left < 2) { time_step(dt); return 1; }
        if (!graph_exec || graph_dt != dt) capture(dt);
        if (graph_parity != swap_parity) { time_step(dt); return 1; }
        HIPCHECK(hipGraphLaunch(graph_exec, st));
        step_clock(dt);
        step_clock(dt);
        return 2;
    }

    // ---- member tables: the launches of two steps of Δt (the sequence a replay repeats) are
    // recorded into R instead of run.  The model's stream is captured meanwhile, so a launch that
    // has no member-table form (it bypasses the recorder) lands in the captured graph, not on the
    // device, and is reported.
    void record_two_steps(double dt, Recorder &R) {
        if (!capturable()) throw ArgError{"the conjugate-gradient solver's steps read scalars on the host"};
        if (iteration == 0) throw ArgError{"a model's first step is not recorded (it begins with update_state)"};
        HIPCHECK(hipStreamSynchronize(st));
        const bool pt = phase_timing;
        phase_timing = false;
        hipGraph_t gr = nullptr;
        HIPCHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        thread_recorder() = &R;
        bool threw = false;
        try {
            step_launches(dt);
            step_launches(dt);
        } catch (...) {
            threw = true;
        }
        thread_recorder() = nullptr;
        phase_timing = pt;
        const hipError_t ec = hipStreamEndCapture(st, &gr);
        size_t nodes = 0;
        if (ec == hipSuccess && gr) (void)hipGraphGetNodes(gr, nullptr, &nodes);
        if (gr) (void)hipGraphDestroy(gr);
        if (threw) throw ArgError{"recording the step failed"};
        HIPCHECK(ec);
        if (nodes != 0) throw ArgError{"a launch of this configuration's step has no member-table form"};
    }

    // ---- autotuner: the z-chunk length of the tiled kernels, timed on this model's own state
    int autotune() {
        const Ctx<T> C = ctx();
        if (!tiled_tendencies_supported(C)) return tune_kchunk;
        const bool clo = (cfg.closure == OM_CLOSURE_AMD || cfg.closure == OM_CLOSURE_SMAGORINSKY_LILLY) && tiled_closure_supported(C);
        std::vector<int> cand;
        for (int kc : {4, 8, 12, 16, 24, 32, 64}) if (kc <= g.Nz) cand.push_back(kc);
        if (cand.empty()) cand.push_back(g.Nz);
        // each kernel is timed on its own, so each gets its own chunk length
        auto fastest = [&](auto launch) {
            float best = INFINITY;
            int pick = cand[0];
            for (int kc : cand) {
                launch(kc);
                HIPCHECK(hipEventRecord(ev0, st));
                for (int r = 0; r < 5; ++r) launch(kc);
                HIPCHECK(hipEventRecord(ev1, st));
                HIPCHECK(hipEventSynchronize(ev1));
                float ms = 0;
                HIPCHECK(hipEventElapsedTime(&ms, ev0, ev1));
                if (ms < best) { best = ms; pick = kc; }
            }
            return pick;
        };
        // the re-evaluations leave νₑ, κₑ and Gⁿ as they were (same state, same values); on an
        // immersed grid the cells below kfix are redone by the generic kernels, the closure's before
        // the tendency kernel is timed (its last tiled launch then reads the final νₑ, κₑ)
        if (clo) tune_kchunk_closure = fastest([&](int kc) { launch_closure_tiled(C, kc, st); });
        if (clo && g.kbot) launch_closure(C, st, true);
        tune_kchunk = fastest([&](int kc) { launch_tend_tiled(C, kc, st); });
        if (g.kbot) launch_tendencies(C, st, true);
        drop_graph();
        return tune_kchunk;
    }

    T *field_ptr(int id, bool &is_w) {
        is_w = false;
        if (id >= 0 && id < nf) { is_w = id == 2; return s.U[id]; }
        if (id >= OM_FIELD_GN && id < OM_FIELD_GN + nf) { is_w = id - OM_FIELD_GN == 2; return s.Gn[id - OM_FIELD_GN]; }
        if (id >= OM_FIELD_GM && id < OM_FIELD_GM + nf) { is_w = id - OM_FIELD_GM == 2; return s.Gm[id - OM_FIELD_GM]; }
        if (id == OM_FIELD_PNHS) return s.pNHS;
        if (id == OM_FIELD_PHY) return s.pHY;
        if (id == OM_FIELD_NU) return s.nu;
        if (id >= OM_FIELD_KAPPA0 && id < OM_FIELD_KAPPA0 + nt) return s.kap[id - OM_FIELD_KAPPA0];
        return nullptr;
    }
    // interior extent along z of a field (w on a Bounded z has Nz + 1 faces)
    int nz_interior(bool is_w) const { return (is_w && g.tz == OM_BOUNDED) ? g.Nz + 1 : g.Nz; }
    int nz_parent(bool is_w) const { return (is_w && g.tz == OM_BOUNDED) ? g.Nz + 2 * g.Hz + 1 : g.Nz + 2 * g.Hz; }

    void set_field(int id, const double *h, bool with_halos) {
        bool is_w;
        T *o = field_ptr(id, is_w);
        if (!o) throw ArgError{"unknown field id"};
        HIPCHECK(hipStreamSynchronize(st));
        const size_t n = (size_t)g.sz * nz_parent(is_w);   // the parent array's span
        std::vector<T> par(n);
        T *base = field_base(o);
        HIPCHECK(hipMemcpy(par.data(), base, n * sizeof(T), hipMemcpyDeviceToHost));
        size_t t = 0;
        if (with_halos) {
            for (int z = 0; z < nz_parent(is_w); ++z) for (int y = 0; y < pny; ++y) for (int x = 0; x < pnx; ++x)
                par[parent_index(x, y, z)] = (T)h[t++];
        } else {
            const int nzi = nz_interior(is_w);
            for (int k = 0; k < nzi; ++k) for (int j = 0; j < g.Ny; ++j) for (int i = 0; i < g.Nx; ++i)
                par[parent_index(i + g.Hx, j + g.Hy, k + g.Hz)] = (T)h[t++];
        }
        HIPCHECK(hipMemcpy(base, par.data(), n * sizeof(T), hipMemcpyHostToDevice));
    }
    void get_field(int id, double *h, bool with_halos) {
        bool is_w;
        T *o = field_ptr(id, is_w);
        if (!o) throw ArgError{"unknown field id"};
        HIPCHECK(hipStreamSynchronize(st));
        const size_t n = (size_t)g.sz * nz_parent(is_w);
        std::vector<T> par(n);
        HIPCHECK(hipMemcpy(par.data(), field_base(o), n * sizeof(T), hipMemcpyDeviceToHost));
        size_t t = 0;
        if (with_halos) {
            for (int z = 0; z < nz_parent(is_w); ++z) for (int y = 0; y < pny; ++y) for (int x = 0; x < pnx; ++x)
                h[t++] = (double)par[parent_index(x, y, z)];
        } else {
            const int nzi = nz_interior(is_w);
            for (int k = 0; k < nzi; ++k) for (int j = 0; j < g.Ny; ++j) for (int i = 0; i < g.Nx; ++i)
                h[t++] = (double)par[parent_index(i + g.Hx, j + g.Hy, k + g.Hz)];
        }
    }
};

}  // namespace

struct om_model {
    int precision = 8;
    ModelBase *impl = nullptr;
    std::string err;
    bool broken = false;
};

template <class F> static int guarded(om_model *m, F &&fn) {
    if (!m) return OM_EINVAL;
    if (m->broken) { m->err = "handle unusable after an earlier HIP failure: " + m->err; return OM_EHIP; }
    try {
        fn();
        return OM_OK;
    } catch (const HipError &e) {
        m->err = std::string(e.what) + ": " + hipGetErrorString(e.e);
        m->broken = true;
        return OM_EHIP;
    } catch (const ArgError &e) {
        m->err = e.msg;
        return OM_EINVAL;
    } catch (const NotConverged &e) {
        m->err = e.msg;
        return OM_ENOTCONV;
    } catch (const std::bad_alloc &) {
        m->err = "host allocation failed";
        return OM_ENOMEM;
    } catch (...) {
        m->err = "unexpected C++ exception";
        return OM_EINVAL;
    }
}

#define WITH_MODEL(m, body)                                                          \
    guarded(m, [&]() {                                                              \
        if ((m)->precision == 8) { auto *M = static_cast<ModelT<double> *>((m)->impl); body; } \
        else { auto *M = static_cast<ModelT<float> *>((m)->impl); body; }           \
    })

static std::string validate(const om_config_t *c) {
    if (!c) return "null config";
    if (c->precision != 4 && c->precision != 8) return "precision must be 4 (FP32) or 8 (FP64)";
    for (int d = 0; d < 3; ++d) {
        if (c->N[d] < 1) return "grid size must be positive";
        if (c->topology[d] < 0 || c->topology[d] > 2) return "unknown topology";
        if (c->topology[d] == OM_FLAT && c->N[d] != 1) return "a Flat dimension must have size 1";
    }
    if (c->topology[0] != OM_PERIODIC || c->topology[1] != OM_PERIODIC)
        return "x and y must be Periodic on the MI355X path";
    if (c->N[0] < 2) return "Nx must be at least 2";
    for (int d = 0; d < 3; ++d)
        if (c->topology[d] == OM_PERIODIC && c->N[d] < c->H[d]) return "a Periodic dimension needs at least as many cells as halo points";
    if (c->n_tracers < 0 || c->n_tracers > OM_MAX_TRACERS) return "n_tracers out of range";
    int need = c->advection == OM_ADV_WENO5 ? 3 : 1;
    if (c->closure != OM_CLOSURE_NONE && need < 2) need = 2;
    for (int d = 0; d < 3; ++d) {
        if (c->topology[d] == OM_FLAT) continue;
        if (c->H[d] < need) return "halo too small for the advection scheme / closure (WENO5 needs 3)";
        if (c->topology[d] != OM_FLAT && c->advection == OM_ADV_WENO5 && c->N[d] < 3)
            return "WENO5 needs at least 3 cells per non-Flat dimension";
    }
    if (c->closure != OM_CLOSURE_NONE && c->closure != OM_CLOSURE_SCALAR_DIFFUSIVITY && c->topology[2] == OM_FLAT)
        return "AMD / Smagorinsky closures need a three-dimensional grid";
    if (c->z_faces && c->topology[2] != OM_BOUNDED) return "stretched z requires a Bounded z";
    if (c->z_faces && c->pressure_solver == OM_SOLVER_FFT)
        return "a stretched z grid needs the Fourier-tridiagonal or the conjugate-gradient pressure solver";
    if (c->pressure_solver < 0 || c->pressure_solver > OM_SOLVER_CONJUGATE_GRADIENT) return "unknown pressure solver";
    if (c->pressure_solver == OM_SOLVER_CONJUGATE_GRADIENT) {
        for (int d = 0; d < 3; ++d)
            if (c->topology[d] == OM_FLAT) return "the conjugate-gradient solver needs a three-dimensional grid";
        if (c->cg_preconditioner < OM_PRECOND_DEFAULT || c->cg_preconditioner > OM_PRECOND_LEARNED) return "unknown CG preconditioner";
        if (c->cg_preconditioner == OM_PRECOND_LEARNED && (c->N[0] % 16 != 0 || c->topology[2] != OM_BOUNDED))
            return "the learned preconditioner needs Nx to be a multiple of 16 and a Bounded z";
    }
    if (c->pressure_solver == OM_SOLVER_FOURIER_TRIDIAGONAL && c->topology[2] != OM_BOUNDED)
        return "the Fourier-tridiagonal solver needs a Bounded z";
    {
        // FFT lines are staged in LDS (om_kernels.h: fft_lines_per_block)
        const size_t e = 2 * (size_t)c->precision;
        const int xl = c->N[0] % 2 == 0 ? c->N[0] / 2 : c->N[0];
        const int Nz = c->N[2];
        bool ok = om::fft_lines_per_block(xl, xl, e) > 0 && om::fft_lines_per_block(c->N[1], c->N[1], e) > 0;
        const bool z_tridiagonal = c->pressure_solver == OM_SOLVER_FOURIER_TRIDIAGONAL || c->z_faces;
        if (c->topology[2] != OM_FLAT && !z_tridiagonal) {
            const int L = (c->topology[2] == OM_BOUNDED && Nz % 2 == 1) ? 2 * Nz : Nz;
            ok = ok && om::fft_lines_per_block(L, L + (c->topology[2] == OM_BOUNDED && Nz % 2 == 0 ? Nz : 0), e) > 0;
        }
        if (!ok) return "grid too large for the FFT pressure solver: a line of Nx/2, Ny or Nz points and its tables must fit in 64 KiB of LDS";
    }
    if (c->bottom_height) {
        if (c->topology[2] != OM_BOUNDED) return "an immersed bottom needs a Bounded z";
        if (c->advection != OM_ADV_WENO5 && c->advection != OM_ADV_NONE)
            return "an immersed bottom supports WENO(order=5) advection or none";
    }
    if (c->buoyancy == OM_BUOYANCY_TRACER && (c->buoyancy_tracer < 0 || c->buoyancy_tracer >= c->n_tracers))
        return "buoyancy tracer index out of range";
    if (c->buoyancy == OM_BUOYANCY_LINEAR_EOS &&
        (c->T_tracer < 0 || c->T_tracer >= c->n_tracers || c->S_tracer < 0 || c->S_tracer >= c->n_tracers))
        return "T/S tracer index out of range";
    return "";
}

extern "C" {

int om_abi_version(void) { return OM_ABI_VERSION; }

int om_create(const om_config_t *cfg, om_model_t **out) {
    if (!out) return OM_EINVAL;
    *out = nullptr;
    std::string v = validate(cfg);
    if (!v.empty()) { g_create_error = v; return OM_EINVAL; }
    auto *m = new (std::nothrow) om_model();
    if (!m) return OM_ENOMEM;
    m->precision = cfg->precision;
    try {
        if (m->precision == 8) { auto *M = new ModelT<double>(); m->impl = M; M->init(*cfg); }
        else { auto *M = new ModelT<float>(); m->impl = M; M->init(*cfg); }
    } catch (const HipError &e) {
        g_create_error = std::string(e.what) + ": " + hipGetErrorString(e.e);
        delete m->impl;
        delete m;
        return OM_EHIP;
    } catch (const ArgError &e) {
        g_create_error = e.msg;
        delete m->impl;
        delete m;
        return OM_EINVAL;
    }
    *out = m;
    return OM_OK;
}

int om_destroy(om_model_t *m) {
    if (!m) return OM_EINVAL;
    delete m->impl;
    delete m;
    return OM_OK;
}

const char *om_last_error(const om_model_t *m) {
    if (!m) return g_create_error.c_str();
    return m->err.c_str();
}

int om_field_info(om_model_t *m, int32_t field, om_field_info_t *info) {
    if (!info) return OM_EINVAL;
    return WITH_MODEL(m, {
        bool is_w;
        auto *o = M->field_ptr(field, is_w);
        if (!o) throw ArgError{"unknown field id"};
        info->data = M->field_base(o);
        info->element_size = m->precision;
        info->parent_size[0] = M->pnx;
        info->parent_size[1] = M->pny;
        info->parent_size[2] = M->nz_parent(is_w);
        info->offset[0] = 1 - M->g.Hx;
        info->offset[1] = 1 - M->g.Hy;
        info->offset[2] = 1 - M->g.Hz;
        info->stride[0] = 1;
        info->stride[1] = M->g.sy;
        info->stride[2] = M->g.sz;
    });
}

int om_set_field(om_model_t *m, int32_t field, const double *host, int32_t with_halos) {
    if (!host) return OM_EINVAL;
    return WITH_MODEL(m, M->set_field(field, host, with_halos != 0));
}

int om_get_field(om_model_t *m, int32_t field, double *host, int32_t with_halos) {
    if (!host) return OM_EINVAL;
    return WITH_MODEL(m, M->get_field(field, host, with_halos != 0));
}

int om_fill_halos(om_model_t *m, uint32_t mask, int32_t fill_normal_flow) {
    return WITH_MODEL(m, {
        M->halos_prognostic(mask & M->all_mask(), fill_normal_flow != 0);
        if (mask & (1u << 16)) M->halos_aux({M->s.pNHS});
        HIPCHECK(hipStreamSynchronize(M->st));
    });
}

int om_update_state(om_model_t *m) {
    return WITH_MODEL(m, { M->update_state(); HIPCHECK(hipStreamSynchronize(M->st)); HIPCHECK(hipGetLastError()); });
}

int om_compute_tendencies(om_model_t *m) {
    return WITH_MODEL(m, { M->compute_tendencies(); HIPCHECK(hipStreamSynchronize(M->st)); HIPCHECK(hipGetLastError()); });
}

int om_rk3_substep(om_model_t *m, double dt, double gamma, double zeta, int32_t first_stage) {
    return WITH_MODEL(m, {
        using TT = std::remove_pointer_t<decltype(M->rhs_dense)>;
        M->cg_unconverged = 0;
        M->rk3_substep(dt, (TT)gamma, (TT)zeta, first_stage != 0);
        M->flush_p_halo();
        HIPCHECK(hipStreamSynchronize(M->st));
        HIPCHECK(hipGetLastError());
        M->report_unconverged();
    });
}

int om_cache_previous_tendencies(om_model_t *m) { return WITH_MODEL(m, M->cache_previous_tendencies()); }

int om_pressure_correction(om_model_t *m, double dt) {
    return WITH_MODEL(m, {
        using TT = std::remove_pointer_t<decltype(M->rhs_dense)>;
        M->cg_unconverged = 0;
        M->pressure_correction((TT)dt);
        M->flush_p_halo();
        HIPCHECK(hipStreamSynchronize(M->st));
        HIPCHECK(hipGetLastError());
        M->report_unconverged();
    });
}

int om_time_step(om_model_t *m, double dt, int32_t nsteps) {
    if (nsteps < 0) return OM_EINVAL;
    return WITH_MODEL(m, {
        M->evused = 0;
        M->cg_unconverged = 0;
        HIPCHECK(hipEventRecord(M->ev0, M->st));
        // pairs of steps replay the model's captured two-step graph (the same launches, so the
        // same results) unless phases are being timed or the graph form is switched off
        const bool graphs = M->tune_step_graphs && !M->phase_timing;
        for (int left = nsteps; left > 0;) left -= M->issue_steps(dt, left, graphs);
        HIPCHECK(hipEventRecord(M->ev1, M->st));
        HIPCHECK(hipEventSynchronize(M->ev1));
        HIPCHECK(hipGetLastError());
        float ms = 0;
        HIPCHECK(hipEventElapsedTime(&ms, M->ev0, M->ev1));
        M->last_ms = ms;
        if (M->phase_timing) M->collect_phase_times();
        M->report_unconverged();
    });
}

int om_get_clock(const om_model_t *mc, double *time, int64_t *iteration, int32_t *stage, double *last_dt,
                 double *last_stage_dt) {
    om_model_t *m = const_cast<om_model_t *>(mc);
    return WITH_MODEL(m, {
        if (time) *time = M->time;
        if (iteration) *iteration = M->iteration;
        if (stage) *stage = M->stage;
        if (last_dt) *last_dt = M->last_dt;
        if (last_stage_dt) *last_stage_dt = M->last_stage_dt;
    });
}

int om_set_clock(om_model_t *m, double time, int64_t iteration, int32_t stage, double last_dt, double last_stage_dt) {
    return WITH_MODEL(m, {
        M->time = time; M->iteration = iteration; M->stage = stage; M->last_dt = last_dt; M->last_stage_dt = last_stage_dt;
    });
}

int om_poisson_solve(om_model_t *m, const double *rhs, double *phi) {
    if (!rhs || !phi) return OM_EINVAL;
    return WITH_MODEL(m, {
        using TT = std::remove_pointer_t<decltype(M->rhs_dense)>;
        const size_t n = (size_t)M->g.Nx * M->g.Ny * M->g.Nz;
        std::vector<TT> h(n);
        for (size_t q = 0; q < n; ++q) h[q] = (TT)rhs[q];
        HIPCHECK(hipMemcpy(M->rhs_dense, h.data(), n * sizeof(TT), hipMemcpyHostToDevice));
        M->cg_unconverged = 0;
        if (M->cg) {
            // ∇²ϕ = rhs as V∇²ϕ = V·rhs from a zero initial guess
            cg_source(M->ctx(), M->rhs_dense, M->cg_b, M->st);
            cg_xio(M->g, M->cg_z, M->cg_x, (TT)0, 0, M->st);
            M->cg_solve();
            cg_gather(M->g, M->cg_x, M->phi_dense, M->st);
        } else {
            launch_poisson(M->P, M->ctx(), M->rhs_dense, M->phi_dense, 1, (TT)1, M->st);
            if (M->tune_refine) {
                HIPCHECK(hipMemcpyAsync(M->ref_b64, rhs, n * sizeof(double), hipMemcpyHostToDevice, M->st));
                M->refine(M->phi_dense, M->g.Nx, (int64_t)M->g.Nx * M->g.Ny, M->ref_b64, 1.0);
            }
        }
        HIPCHECK(hipStreamSynchronize(M->st));
        HIPCHECK(hipGetLastError());
        HIPCHECK(hipMemcpy(h.data(), M->phi_dense, n * sizeof(TT), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; ++q) phi[q] = (double)h[q];
        M->report_unconverged();
    });
}

int om_divergence(om_model_t *m, double *out) {
    if (!out) return OM_EINVAL;
    return WITH_MODEL(m, {
        using TT = std::remove_pointer_t<decltype(M->rhs_dense)>;
        const size_t n = (size_t)M->g.Nx * M->g.Ny * M->g.Nz;
        launch_divergence(M->ctx(), M->rhs_dense, M->st);
        HIPCHECK(hipStreamSynchronize(M->st));
        HIPCHECK(hipGetLastError());
        std::vector<TT> h(n);
        HIPCHECK(hipMemcpy(h.data(), M->rhs_dense, n * sizeof(TT), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n; ++q) out[q] = (double)h[q];
    });
}

int om_set_phase_timing(om_model_t *m, int32_t enable) {
    return WITH_MODEL(m, { M->phase_timing = enable != 0; M->evused = 0; });
}

int om_last_solve_info(const om_model_t *mc, int32_t *iterations, double *residual_norm, double *true_residual_norm) {
    om_model_t *m = const_cast<om_model_t *>(mc);
    if (!m) return OM_EINVAL;
    return WITH_MODEL(m, {
        if (!M->cg) throw std::invalid_argument("the model's pressure solver is not the conjugate-gradient solver");
        if (iterations) *iterations = M->cg_iterations;
        if (residual_norm) *residual_norm = M->cg_res_last;
        if (true_residual_norm) *true_residual_norm = M->cg_true_res_last;
    });
}

int om_cell_advection_timescale(om_model_t *m, double *tau) {
    if (!m || !tau) return OM_EINVAL;
    return WITH_MODEL(m, {
        using TT = std::remove_pointer_t<decltype(M->rhs_dense)>;
        launch_adv_timescale(M->ctx(), M->diag_partial, M->diag_out, M->st);
        HIPCHECK(hipMemcpyAsync(M->diag_host, M->diag_out, 3 * sizeof(double), hipMemcpyDeviceToHost, M->st));
        HIPCHECK(hipStreamSynchronize(M->st));
        if (M->diag_host[2] > 0) *tau = std::numeric_limits<double>::quiet_NaN();
        else *tau = (double)(TT(1) / (TT)std::max(M->diag_host[0], 0.0));
    });
}

int om_field_statistics(om_model_t *m, int32_t field, int32_t with_halos, double *maximum, double *minimum,
                        int64_t *nan_count) {
    if (!m) return OM_EINVAL;
    return WITH_MODEL(m, {
        bool is_w;
        auto *o = M->field_ptr(field, is_w);
        if (!o) throw ArgError{"unknown field id"};
        launch_field_stats(M->g, o, with_halos ? 1 : 0, (is_w && M->g.tz == OM_BOUNDED) ? 1 : 0, M->diag_partial, M->diag_out, M->st);
        HIPCHECK(hipMemcpyAsync(M->diag_host, M->diag_out, 3 * sizeof(double), hipMemcpyDeviceToHost, M->st));
        HIPCHECK(hipStreamSynchronize(M->st));
        if (maximum) *maximum = M->diag_host[0];
        if (minimum) *minimum = M->diag_host[1];
        if (nan_count) *nan_count = (int64_t)M->diag_host[2];
    });
}

int om_field_statistics_noalloc(om_model_t *m, int32_t field, int32_t with_halos, double *maximum, double *minimum,
                                int64_t *nan_count) {
    if (!m) return OM_EINVAL;
    return WITH_MODEL(m, {
        bool is_w;
        auto *o = M->field_ptr(field, is_w);
        if (!o) throw ArgError{"unknown field id"};
        // the reduction's partials and results take a fixed 3 · 1024 + 3 doubles whatever
        // the field's size; fields beyond kNoallocMaxElements are refused
        const int64_t n = (int64_t)(with_halos ? M->pnx : M->g.Nx) * (with_halos ? M->pny : M->g.Ny) *
                          (with_halos ? M->nz_parent(is_w) : M->nz_interior(is_w));
        if (n > kNoallocMaxElements)
            throw ArgError{"om_field_statistics_noalloc: the field has more than 2^40 elements"};
        int err = 0;
        double *scr = diag_scratch(M->device, &err);
        if (!scr) {
            if (err > 0) throw (HipError{(hipError_t)err, "hipGetSymbolAddress(scratch)"});
            throw ArgError{"om_field_statistics_noalloc: device index beyond the scratch table"};
        }
        if (diag_partial_doubles() + 3 > diag_scratch_doubles()) throw ArgError{"scratch smaller than the reduction"};
        double *out = scr + diag_partial_doubles();
        launch_field_stats(M->g, o, with_halos ? 1 : 0, (is_w && M->g.tz == OM_BOUNDED) ? 1 : 0, scr, out, M->st);
        HIPCHECK(hipGetLastError());
        double h[3];
        HIPCHECK(hipMemcpyAsync(h, out, 3 * sizeof(double), hipMemcpyDeviceToHost, M->st));
        HIPCHECK(hipStreamSynchronize(M->st));
        if (maximum) *maximum = h[0];
        if (minimum) *minimum = h[1];
        if (nan_count) *nan_count = (int64_t)h[2];
    });
}

int om_learned_set_weights(om_model_t *m, const float *weights) {
    if (!m || !weights) return OM_EINVAL;
    return WITH_MODEL(m, {
        M->net_alloc();
        const size_t nb = (size_t)NET_LAYERS * NET_WPL * sizeof(float);
        HIPCHECK(hipMemcpyAsync(M->net.w, weights, nb, hipMemcpyHostToDevice, M->st));
        HIPCHECK(hipMemsetAsync(M->net.m, 0, nb, M->st));
        HIPCHECK(hipMemsetAsync(M->net.v, 0, nb, M->st));
        M->net.adam_step = 0;
        net_pack(M->net, M->st);
        HIPCHECK(hipStreamSynchronize(M->st));
    });
}

int om_learned_get_weights(om_model_t *m, float *weights) {
    if (!m || !weights) return OM_EINVAL;
    return WITH_MODEL(m, {
        M->net_alloc();
        HIPCHECK(hipMemcpyAsync(weights, M->net.w, (size_t)NET_LAYERS * NET_WPL * sizeof(float), hipMemcpyDeviceToHost, M->st));
        HIPCHECK(hipStreamSynchronize(M->st));
    });
}

int om_learned_log(om_model_t *m, int32_t iterations_per_solve, int32_t capacity) {
    if (!m || iterations_per_solve < 0 || capacity < 0) return OM_EINVAL;
    return WITH_MODEL(m, {
        if (!M->cg) throw ArgError{"pairs are logged from conjugate-gradient solves"};
        M->net_alloc();
        // a new call starts a new set of pairs (buffers are reused)
        M->log_iters = capacity > 0 ? iterations_per_solve : 0;
        M->log_cap = capacity;
        M->log_count = 0;
        M->log_total = 0;
        M->log_next = 0;
        while ((int)M->pair_in3.size() > capacity) {
            M->dfree(M->pair_in3.back());
            M->dfree(M->pair_tgt.back());
            M->pair_in3.pop_back();
            M->pair_tgt.pop_back();
            M->pair_energy.pop_back();
        }
    });
}

int om_learned_pairs(om_model_t *m, int32_t *count) {
    if (!m || !count) return OM_EINVAL;
    return WITH_MODEL(m, { *count = M->log_count; });
}

int om_learned_pairs_logged(om_model_t *m, int64_t *total) {
    if (!m || !total) return OM_EINVAL;
    return WITH_MODEL(m, { *total = M->log_total; });
}

int om_learned_train(om_model_t *m, int32_t steps, double lr, double *loss_first, double *loss_last) {
    if (!m || steps < 0 || !(lr > 0)) return OM_EINVAL;
    return WITH_MODEL(m, {
        if (M->log_count == 0) throw ArgError{"no logged pairs to train on"};
        NetState &S = M->net;
        if (!M->ln_gout) M->ln_gout = M->template dalloc<float>((size_t)M->g.Nx * M->g.Ny * M->g.Nz);
        // loss: the A-energy of the remaining error relative to that of the target,
        // (δ − t)ᵀ(−V∇²)(δ − t) / tᵀ(−V∇²)t, the quantity a CG preconditioner should make small
        for (int s = 0; s < steps; ++s) {
            const int pr = s % M->log_count;
            const double et = std::max(M->pair_energy[pr], 1e-300);
            net_input_dense(S, M->pair_in3[pr], M->st);
            net_forward(S, S.out, M->st);
            ln_energy(M->g, S.out, M->pair_tgt[pr], M->cg_q, 1.0 / et, M->ln_gout, M->cg_partial, M->cg_out, M->st);
            if (s == 0 || s == steps - 1) {
                const double e = M->cg_read(1) / et;
                if (s == 0 && loss_first) *loss_first = e;
                if (s == steps - 1 && loss_last) *loss_last = e;
            }
            net_output_grad(S, M->ln_gout, M->st);
            net_backward(S, nullptr, M->st);
            net_adam(S, (float)lr, M->st);
        }
        HIPCHECK(hipStreamSynchronize(M->st));
    });
}

int om_learned_evaluate(om_model_t *m, double *mean_loss) {
    if (!m || !mean_loss) return OM_EINVAL;
    return WITH_MODEL(m, {
        if (M->log_count == 0) throw ArgError{"no logged pairs to evaluate on"};
        NetState &S = M->net;
        double sum = 0;
        for (int pr = 0; pr < M->log_count; ++pr) {
            net_input_dense(S, M->pair_in3[pr], M->st);
            net_forward(S, S.out, M->st);
            ln_energy(M->g, S.out, M->pair_tgt[pr], M->cg_q, 0.0, (float *)nullptr, M->cg_partial, M->cg_out, M->st);
            sum += M->cg_read(1) / std::max(M->pair_energy[pr], 1e-300);
        }
        *mean_loss = sum / M->log_count;
    });
}

int om_learned_forward(om_model_t *m, const float *input3, float *output) {
    if (!m || !input3 || !output) return OM_EINVAL;
    return WITH_MODEL(m, {
        M->net_alloc();
        const size_t nc = (size_t)M->g.Nx * M->g.Ny * M->g.Nz;
        HIPCHECK(hipMemcpyAsync(M->ln_in3, input3, 3 * nc * sizeof(float), hipMemcpyHostToDevice, M->st));
        net_input_dense(M->net, M->ln_in3, M->st);
        net_forward(M->net, M->net.out, M->st);
        HIPCHECK(hipMemcpyAsync(output, M->net.out, nc * sizeof(float), hipMemcpyDeviceToHost, M->st));
        HIPCHECK(hipStreamSynchronize(M->st));
    });
}

int om_learned_backward(om_model_t *m, const float *input3, const float *target, const float *mask, double *loss, float *dweights,
                        float *dinput3) {
    if (!m || !input3 || !target || !mask) return OM_EINVAL;
    return WITH_MODEL(m, {
        M->net_alloc();
        const size_t nc = (size_t)M->g.Nx * M->g.Ny * M->g.Nz;
        float *dev = M->template dalloc<float>(5 * nc);   // target, mask, input gradient
        double n = 0;
        for (size_t q = 0; q < nc; ++q) n += mask[q];
        const float inv_n = (float)(1.0 / std::max(n, 1.0));
        HIPCHECK(hipMemcpyAsync(M->ln_in3, input3, 3 * nc * sizeof(float), hipMemcpyHostToDevice, M->st));
        HIPCHECK(hipMemcpyAsync(dev, target, nc * sizeof(float), hipMemcpyHostToDevice, M->st));
        HIPCHECK(hipMemcpyAsync(dev + nc, mask, nc * sizeof(float), hipMemcpyHostToDevice, M->st));
        net_input_dense(M->net, M->ln_in3, M->st);
        net_forward(M->net, M->net.out, M->st);
        net_loss(M->net, M->net.out, dev, dev + nc, inv_n, M->st);
        net_backward(M->net, dev + 2 * nc, M->st);
        double h = 0;
        HIPCHECK(hipMemcpyAsync(&h, M->net.lsum, sizeof(double), hipMemcpyDeviceToHost, M->st));
        if (dweights)
            HIPCHECK(hipMemcpyAsync(dweights, M->net.dw, (size_t)NET_LAYERS * NET_WPL * sizeof(float), hipMemcpyDeviceToHost, M->st));
        if (dinput3) HIPCHECK(hipMemcpyAsync(dinput3, dev + 2 * nc, 3 * nc * sizeof(float), hipMemcpyDeviceToHost, M->st));
        HIPCHECK(hipStreamSynchronize(M->st));
        if (loss) *loss = h * (double)inv_n;
        M->dfree(dev);
    });
}

int om_learned_energy_loss(om_model_t *m, const float *output, const float *target, double *energy, float *doutput) {
    if (!m || !output || !target || !energy) return OM_EINVAL;
    return WITH_MODEL(m, {
        if (!M->cg) throw ArgError{"the energy uses the conjugate-gradient solver's operator"};
        const size_t nc = (size_t)M->g.Nx * M->g.Ny * M->g.Nz;
        float *dev = M->template dalloc<float>(3 * nc);   // y, t, dE/dy (a test hook: allocated per call)
        HIPCHECK(hipMemcpyAsync(dev, output, nc * sizeof(float), hipMemcpyHostToDevice, M->st));
        HIPCHECK(hipMemcpyAsync(dev + nc, target, nc * sizeof(float), hipMemcpyHostToDevice, M->st));
        ln_energy(M->g, dev, dev + nc, M->cg_q, 1.0, dev + 2 * nc, M->cg_partial, M->cg_out, M->st);
        *energy = M->cg_read(1);
        if (doutput) HIPCHECK(hipMemcpyAsync(doutput, dev + 2 * nc, nc * sizeof(float), hipMemcpyDeviceToHost, M->st));
        HIPCHECK(hipStreamSynchronize(M->st));
        M->dfree(dev);
    });
}

int om_learned_last_solve(om_model_t *m, int32_t *learned_iterations, int32_t *fell_back) {
    if (!m) return OM_EINVAL;
    return WITH_MODEL(m, {
        if (learned_iterations) *learned_iterations = M->ln_learned_its;
        if (fell_back) *fell_back = M->ln_fell_back;
    });
}

static bool known_tuning_key(int32_t key) {
    return key == OM_TUNE_TENDENCY_KERNEL || key == OM_TUNE_KCHUNK || key == OM_TUNE_LEARNED_ITERATIONS ||
           key == OM_TUNE_POISSON_REFINEMENT || key == OM_TUNE_STEP_GRAPHS || key == OM_TUNE_CONV_TILED ||
           key == OM_TUNE_WGRAD_MFMA;
}

int om_autotune(om_model_t *m, int32_t *kchunk) {
    if (!m) return OM_EINVAL;
    return WITH_MODEL(m, {
        const int kc = M->autotune();
        if (kchunk) *kchunk = kc;
        HIPCHECK(hipGetLastError());
    });
}

// The member-table form of an ensemble step: position p of every member's recorded two-step
// launch sequence becomes one launch over a device table of the members' arguments, and the whole
// sequence one hipGraph.  The members may be split into groups, each with its own tables, graph and
// stream, so that the groups' graphs run concurrently and one group's bandwidth-bound phases
// overlap another's compute-bound ones.
struct BatchGroup {
    std::vector<size_t> members;      // indices into the ensemble's member list
    std::vector<void *> tables;       // device argument tables, one per position
    hipGraphExec_t exec = nullptr;
    hipStream_t st = nullptr;
};

struct BatchPlan {
    double dt = 0;
    int groups_wanted = 1;
    std::vector<int> parity;          // each member's G⁻/Gⁿ exchange parity when recorded
    std::vector<int> version;         // each member's launch_version when recorded
    std::vector<BatchGroup> groups;
    int positions = 0;
    bool ready() const { return !groups.empty() && groups[0].exec; }
    void clear() {
        for (BatchGroup &g : groups) {
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
            for (void *t : g.tables) (void)hipFree(t);
            if (g.st) (void)hipStreamDestroy(g.st);
        }
        groups.clear();
        parity.clear();
        version.clear();
        positions = 0;
    }
    ~BatchPlan() { clear(); }
};

struct om_ensemble {
    std::vector<om_model_t *> members;
    std::string err;
    BatchPlan plan;
};

namespace {
int member_parity(om_model_t *m) {
    int p = 0;
    (void)WITH_MODEL(m, { p = M->swap_parity; });
    return p;
}
int member_version(om_model_t *m) {
    int v = 0;
    (void)WITH_MODEL(m, { v = M->launch_version; });
    return v;
}

// records every member's two-step sequence, checks that the sequences match launch for launch,
// splits the members into groups of about equal block counts, and per group uploads the member
// tables and captures the batched launches into one graph on the group's stream
void build_plan(om_ensemble *e, double dt) {
    BatchPlan &P = e->plan;
    P.clear();
    const size_t n = e->members.size();
    std::vector<Recorder> rec(n);
    int device = -1;
    for (size_t q = 0; q < n; ++q) {
        om_model_t *m = e->members[q];
        int dev = -1;
        const int rc = WITH_MODEL(m, { M->record_two_steps(dt, rec[q]); dev = M->device; });
        if (rc != OM_OK) throw ArgError{"member " + std::to_string(q) + ": " + m->err};
        if (q == 0) device = dev;
        else if (dev != device) throw ArgError{"member " + std::to_string(q) + " lives on another device"};
    }
    const size_t npos = rec[0].recs.size();
    for (size_t q = 1; q < n; ++q) {
        if (rec[q].recs.size() != npos) throw ArgError{"member " + std::to_string(q) + "'s step has another launch sequence than member 0's"};
        for (size_t p = 0; p < npos; ++p) {
            const LaunchRec &a = rec[0].recs[p], &b = rec[q].recs[p];
            if (a.kind != b.kind || a.block.x != b.block.x || a.block.y != b.block.y || a.block.z != b.block.z)
                throw ArgError{"member " + std::to_string(q) + "'s step has another launch sequence than member 0's"};
        }
    }
    // groups: members in order, each group closed once it holds its share of the blocks
    const int G = std::max(1, std::min<int>(P.groups_wanted, (int)n));
    std::vector<uint64_t> work(n, 0);
    uint64_t total = 0;
    for (size_t q = 0; q < n; ++q) {
        for (const LaunchRec &r : rec[q].recs) work[q] += (uint64_t)r.grid.x * r.grid.y * r.grid.z;
        total += work[q];
    }
    P.groups.resize(G);
    {
        uint64_t acc = 0;
        int gi = 0;
        for (size_t q = 0; q < n; ++q) {
            if (gi < G - 1 && acc >= (uint64_t)(gi + 1) * total / G && !P.groups[gi].members.empty()) ++gi;
            P.groups[gi].members.push_back(q);
            acc += work[q];
        }
        while (!P.groups.empty() && P.groups.back().members.empty()) P.groups.pop_back();
    }
    for (BatchGroup &grp : P.groups) {
        HIPCHECK(hipStreamCreateWithFlags(&grp.st, hipStreamNonBlocking));
        const int gn = (int)grp.members.size();
        std::vector<uint32_t> blocks(npos);
        std::vector<size_t> lds(npos);
        for (size_t p = 0; p < npos; ++p) {
            const size_t es = rec[0].recs[p].entry_size;
            std::vector<unsigned char> host(es * gn);
            uint64_t begin = 0;
            size_t l = 0;
            for (int q = 0; q < gn; ++q) {
                const LaunchRec &r = rec[grp.members[q]].recs[p];
                const uint64_t nb = (uint64_t)r.grid.x * r.grid.y * r.grid.z;
                r.fill(host.data() + q * es, r, (uint32_t)begin);
                begin += (nb + 7) & ~uint64_t(7);   // member ranges start on a multiple of 8 (XCD residue kept)
                l = std::max(l, r.lds);
            }
            if (begin > 0x7fffffffull) throw ArgError{"the ensemble's launch has more than 2^31 blocks"};
            void *t = nullptr;
            HIPCHECK(hipMalloc(&t, host.size()));
            grp.tables.push_back(t);
            HIPCHECK(hipMemcpy(t, host.data(), host.size(), hipMemcpyHostToDevice));
            blocks[p] = (uint32_t)begin;
            lds[p] = l;
        }
        hipGraph_t gr = nullptr;
        HIPCHECK(hipStreamBeginCapture(grp.st, hipStreamCaptureModeThreadLocal));
        for (size_t p = 0; p < npos; ++p) {
            const LaunchRec &r = rec[grp.members[0]].recs[p];
            r.launch_many(grp.tables[p], gn, blocks[p], r.block, lds[p], grp.st);
        }
        HIPCHECK(hipStreamEndCapture(grp.st, &gr));
        const hipError_t ei = hipGraphInstantiate(&grp.exec, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        HIPCHECK(ei);
    }
    P.dt = dt;
    P.positions = (int)npos;
    for (size_t q = 0; q < n; ++q) {
        P.parity.push_back(member_parity(e->members[q]));
        P.version.push_back(member_version(e->members[q]));
    }
}
}  // namespace

int om_ensemble_create(om_model_t *const *members, int32_t n, om_ensemble_t **out) {
    if (!out || !members || n < 1) return OM_EINVAL;
    *out = nullptr;
    auto *e = new om_ensemble;
    for (int q = 0; q < n; ++q) {
        om_model_t *m = members[q];
        if (!m || m->broken) { delete e; g_create_error = "ensemble member " + std::to_string(q) + " is null or unusable"; return OM_EINVAL; }
        bool ok = true;
        const int rc = WITH_MODEL(m, { ok = M->capturable(); });
        if (rc != OM_OK || !ok) {
            delete e;
            g_create_error = "ensemble member " + std::to_string(q) + " uses the conjugate-gradient solver, whose steps cannot be captured";
            return OM_EINVAL;
        }
        for (om_model_t *x : e->members)
            if (x == m) { delete e; g_create_error = "ensemble member " + std::to_string(q) + " is listed twice"; return OM_EINVAL; }
        e->members.push_back(m);
    }
    *out = e;
    return OM_OK;
}

int om_ensemble_destroy(om_ensemble_t *e) {
    delete e;
    return OM_OK;
}

// OM_ENSEMBLE_BATCHED: members at iteration 0 take their first step on their own; then every
// pair of steps of all members is one replay of the member-table graph; a remainder (members with
// an odd count) is stepped member by member
static int ensemble_batched(om_ensemble_t *e, double dt, int32_t nsteps) {
    const size_t n = e->members.size();
    std::vector<int> left(n, nsteps);
    auto each = [&](auto &&fn) -> int {
        for (size_t q = 0; q < n; ++q) {
            om_model_t *m = e->members[q];
            const int rc = fn(q, m);
            if (rc != OM_OK) { e->err = "member " + std::to_string(q) + ": " + m->err; return rc; }
        }
        return OM_OK;
    };
    int rc = each([&](size_t q, om_model_t *m) {
        return WITH_MODEL(m, {
            if (left[q] > 0 && M->iteration == 0) { M->time_step(dt); left[q] -= 1; }
            HIPCHECK(hipStreamSynchronize(M->st));
        });
    });
    if (rc != OM_OK) return rc;
    int pairs = nsteps;
    for (size_t q = 0; q < n; ++q) pairs = std::min(pairs, left[q] / 2);
    if (pairs > 0) {
        BatchPlan &P = e->plan;
        bool stale = !P.ready() || P.dt != dt || P.parity.size() != n;
        for (size_t q = 0; q < n && !stale; ++q)
            stale = member_parity(e->members[q]) != P.parity[q] || member_version(e->members[q]) != P.version[q];
        if (stale) {
            try {
                build_plan(e, dt);
            } catch (const ArgError &a) {
                e->plan.clear();
                e->err = a.msg;
                return OM_EINVAL;
            } catch (const HipError &h) {
                e->plan.clear();
                e->err = std::string(h.what) + ": " + hipGetErrorString(h.e);
                return OM_EHIP;
            }
        }
        for (int r = 0; r < pairs; ++r) {
            for (BatchGroup &grp : P.groups) {
                const hipError_t he = hipGraphLaunch(grp.exec, grp.st);
                if (he != hipSuccess) { e->err = std::string("hipGraphLaunch: ") + hipGetErrorString(he); return OM_EHIP; }
            }
        }
        for (BatchGroup &grp : P.groups) {
            const hipError_t hs = hipStreamSynchronize(grp.st);
            if (hs != hipSuccess) { e->err = std::string("ensemble stream: ") + hipGetErrorString(hs); return OM_EHIP; }
        }
        rc = each([&](size_t q, om_model_t *m) {
            return WITH_MODEL(m, { for (int r = 0; r < 2 * pairs; ++r) M->step_clock(dt); });
        });
        if (rc != OM_OK) return rc;
        for (size_t q = 0; q < n; ++q) left[q] -= 2 * pairs;
    }
    return each([&](size_t q, om_model_t *m) {
        return WITH_MODEL(m, {
            for (int r = 0; r < left[q]; ++r) M->time_step(dt);
            HIPCHECK(hipStreamSynchronize(M->st));
            HIPCHECK(hipGetLastError());
        });
    });
}

int om_ensemble_time_step(om_ensemble_t *e, double dt, int32_t nsteps, int32_t use_graphs) {
    if (!e || nsteps < 0) return OM_EINVAL;
    if (use_graphs < OM_ENSEMBLE_DIRECT || use_graphs > OM_ENSEMBLE_BATCHED) return OM_EINVAL;
    if (use_graphs == OM_ENSEMBLE_BATCHED) return ensemble_batched(e, dt, nsteps);
    // members advance in rounds of up to two steps each, so every member's stream holds work
    // at once; then every stream is drained
    std::vector<int> left(e->members.size(), nsteps);
    bool more = nsteps > 0;
    while (more) {
        more = false;
        for (size_t q = 0; q < e->members.size(); ++q) {
            if (left[q] <= 0) continue;
            om_model_t *m = e->members[q];
            int done = 0;
            const int rc = WITH_MODEL(m, { done = M->issue_steps(dt, left[q], use_graphs != 0); });
            if (rc != OM_OK) { e->err = "member " + std::to_string(q) + ": " + m->err; return rc; }
            left[q] -= done;
            more = more || left[q] > 0;
        }
    }
    for (size_t q = 0; q < e->members.size(); ++q) {
        om_model_t *m = e->members[q];
        const int rc = WITH_MODEL(m, {
            HIPCHECK(hipStreamSynchronize(M->st));
            HIPCHECK(hipGetLastError());
        });
        if (rc != OM_OK) { e->err = "member " + std::to_string(q) + ": " + m->err; return rc; }
    }
    return OM_OK;
}

const char *om_ensemble_last_error(const om_ensemble_t *e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int om_ensemble_set_batch_groups(om_ensemble_t *e, int32_t groups) {
    if (!e || groups < 1 || groups > 64) return OM_EINVAL;
    if (groups != e->plan.groups_wanted) {
        e->plan.clear();
        e->plan.groups_wanted = groups;
    }
    return OM_OK;
}

int om_set_tuning(om_model_t *m, int32_t key, int32_t value) {
    if (!m) return OM_EINVAL;
    if (key == OM_TUNE_TENDENCY_KERNEL && (value < 0 || value > 1)) return OM_EINVAL;
    if (key == OM_TUNE_KCHUNK && (value < 1 || value > 1 << 20)) return OM_EINVAL;
    if (key == OM_TUNE_LEARNED_ITERATIONS && value < 0) return OM_EINVAL;
    if (key == OM_TUNE_POISSON_REFINEMENT && (value < 0 || value > 1)) return OM_EINVAL;
    if (key == OM_TUNE_STEP_GRAPHS && (value < 0 || value > 1)) return OM_EINVAL;
    if (key == OM_TUNE_CONV_TILED && (value < 0 || value > 1)) return OM_EINVAL;
    if (key == OM_TUNE_WGRAD_MFMA && (value < 0 || value > 1)) return OM_EINVAL;
    if (!known_tuning_key(key)) return OM_EINVAL;
    return WITH_MODEL(m, {
        if (key == OM_TUNE_TENDENCY_KERNEL) M->tune_tendency_kernel = value;
        else if (key == OM_TUNE_KCHUNK) M->tune_kchunk = M->tune_kchunk_closure = value;
        else if (key == OM_TUNE_LEARNED_ITERATIONS) M->tune_learned_iterations = value;
        else if (key == OM_TUNE_STEP_GRAPHS) M->tune_step_graphs = value;
        else if (key == OM_TUNE_CONV_TILED) M->net.tiled_conv = value;
        else if (key == OM_TUNE_WGRAD_MFMA) M->net.mfma_wgrad = value;
        else M->set_refinement(value);
        M->drop_graph();   // a captured step holds the launch choices it was recorded with
    });
}

int om_get_tuning(const om_model_t *mc, int32_t key, int32_t *value) {
    om_model_t *m = const_cast<om_model_t *>(mc);
    if (!m || !value) return OM_EINVAL;
    if (!known_tuning_key(key)) return OM_EINVAL;
    return WITH_MODEL(m, {
        *value = key == OM_TUNE_TENDENCY_KERNEL      ? M->tune_tendency_kernel
               : key == OM_TUNE_KCHUNK               ? M->tune_kchunk
               : key == OM_TUNE_LEARNED_ITERATIONS   ? M->tune_learned_iterations
               : key == OM_TUNE_STEP_GRAPHS          ? M->tune_step_graphs
               : key == OM_TUNE_CONV_TILED           ? M->net.tiled_conv
               : key == OM_TUNE_WGRAD_MFMA           ? M->net.mfma_wgrad
                                                     : M->tune_refine;
    });
}

int om_last_step_timing(const om_model_t *mc, double *ms_total, double *ms_per_phase, int32_t *launches_per_phase,
                        int32_t nphases) {
    om_model_t *m = const_cast<om_model_t *>(mc);
    if (nphases < 0 || nphases > OM_NPHASES) return OM_EINVAL;
    return WITH_MODEL(m, {
        if (ms_total) *ms_total = M->last_ms;
        for (int p = 0; p < nphases; ++p) {
            if (ms_per_phase) ms_per_phase[p] = M->phase_ms[p];
            if (launches_per_phase) launches_per_phase[p] = M->phase_n[p];
        }
    });
}

}  // extern "C"
