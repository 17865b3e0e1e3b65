// fcs_lib.hip -- host side and C ABI (include/fcs.h) of the gfx950 chemistry
// integrator, with its size-independent kernels.  Linked with one object per
// padded system size (fcs_kernels_np.hip) and fcs_mlp.hip into libfcs.so.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "fcs.h"
#include "fcs_kernels.h"
#include "fcs_mech_tables.h"
#include "fcs_mlp_internal.h"

// ============================================================== kernels
// The per-size kernel templates live in fcs_kernels.h (one translation unit
// per padded size); the size-independent kernels are here.

// Clears the active-list counters and the deltaTMin accumulator (a kernel
// rather than memset nodes, so the sequence can be captured in a graph).
__global__ void k_reset(unsigned int* counts, unsigned long long* dtMinBits, unsigned long long* nAccepted)
{
    counts[0] = counts[1] = counts[2] = counts[3] = 0u;   // [3]: gated proposals of the learned seed
    *dtMinBits = 0x7ff0000000000000ull;  // +inf
    nAccepted[0] = nAccepted[1] = 0ull;   // accepted, verified
}

// One thread per cell: cold cells are finished at once
// (FastChemistryModel_transientSolve.H:70-76); reacting cells get their
// initial state y = [max(0, Y), T] (src/ODE/OptRodas34/OptRodas34.C:72-81)
// and join the active list.
__global__ __launch_bounds__(256) void k_setup(DevOpts o, int ns, CellArrays a, int* list, unsigned int* count)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nc) return;
    const size_t nc = (size_t)a.nc;
    const double T = a.T0[c];
    if (!(T > o.Treact)) {
        for (int i = 0; i < ns; ++i) a.RR[i * nc + c] = 0.0;
        a.status[c] = CELL_SKIPPED;
        a.dtcOut[c] = a.dtcIn[c];
        a.Tend[c] = T;
        a.nAtt[c] = 0;
        return;
    }
    for (int i = 0; i < ns; ++i) {
        const double Y = a.Y0[i * nc + c];
        a.Yw[i * nc + c] = Y > 0.0 ? Y : 0.0;
    }
    a.Yw[ns * nc + c] = T;
    a.x[c] = 0.0;
    a.dxTry[c] = a.dtcIn[c];
    a.dx[c] = 0.0;
    a.dxTry0[c] = 0.0;
    a.flags[c] = o.solver == SOLVER_SEULEX ? SX_FIRST : 0;
    a.dxFirst[c] = 0.0;
    a.nStep[c] = 0;
    a.nAtt[c] = 0;
    if (o.solver == SOLVER_SEULEX) {
        a.theta[c] = 0.0;    // set at the first step (OptSeulex.C:291-294)
        a.kTarg[c] = 0;
    }
    list[atomicAdd(count, 1u)] = c;
}

// Cells still in an active list when a run reaches its launch cap (a bound
// that the per-cell attempt cap keeps out of reach) are finished as failed,
// as finish_cell finishes a cell over maxSteps: status CELL_MAXSTEPS, the
// input deltaTChem, RR = 0 and the temperature of the last accepted state.
__global__ __launch_bounds__(256) void k_fail_list(DevOpts o, int ns, CellArrays a, const int* list,
                                                   const unsigned int* count)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    const int c = list[i];
    const size_t nc = (size_t)a.nc;
    for (int s = 0; s < ns; ++s) a.RR[s * nc + c] = 0.0;
    const double dtc = a.dtcIn[c];
    a.status[c] = CELL_MAXSTEPS;
    a.Tend[c] = a.Yw[(size_t)ns * nc + c];
    a.dtcOut[c] = dtc < o.dtChemMax ? dtc : o.dtChemMax;
    if (dtc > 0.0) atomicMin(a.dtMinBits, (unsigned long long)__double_as_longlong(dtc));
}

// Network inputs of the cells of a list, one thread per entry: row i =
// [T, log p, Y_0 .. Y_ns-1, log deltaT] in raw units (the network holds its
// own normalisation).  Y is the clipped initial state k_setup prepared.
__global__ void k_features(int ns, CellArrays a, const int* list, const unsigned int* count, float* X)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    const int c = list[i];
    const size_t nc = (size_t)a.nc;
    float* x = X + (size_t)i * (ns + 3);
    x[0] = (float)a.Yw[(size_t)ns * nc + c];
    x[1] = (float)log(a.p0[c]);
    for (int s = 0; s < ns; ++s) x[2 + s] = (float)a.Yw[(size_t)s * nc + c];
    x[ns + 2] = (float)log(a.dtCell ? a.dtCell[c] : a.dt);
}

// The learned seed's gate, one thread per entry of the reacting-cell list:
// a proposal whose first sub-step h is at least half the flow step goes to
// the gated list (its position in list0) for verification; every other cell
// joins list 1 at once, in seed mode starting from min(h, deltaT) when h is a
// positive number.  Keeping the gate out of the verification kernel leaves
// the verification attempt outside any branch.
__global__ void k_gate(int ns, CellArrays a, const int* list, const unsigned int* count, const float* P, int seed,
                       int* listG, unsigned int* countG, int* listOut, unsigned int* countOut,
                       unsigned long long* nAccepted)
{
    const unsigned int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= *count) return;
    const int c = list[i];
    const double dt = a.dtCell ? a.dtCell[c] : a.dt;
    const double h = exp((double)P[(size_t)i * (ns + 2)]);
    if (h >= 0.5 * dt) {
        listG[atomicAdd(countG, 1u)] = (int)i;
        atomicAdd(nAccepted + 1, 1ull);
        return;
    }
    if (seed && h == h && h > 0.0) a.dxTry[c] = h < dt ? h : dt;
    listOut[atomicAdd(countOut, 1u)] = c;
}

// FastChemistryModel::Qdot (src/ChemistryModel/FastChemistryModel.C:
// 646-675): Qdot = -sum_i Hf_i RR_i, species in order; one thread per cell.
__global__ __launch_bounds__(256) void k_qdot(int nc, int ns, const double* Hf, const double* RR, double* Qdot)
{
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    double q = 0.0;
    for (int i = 0; i < ns; ++i) q -= Hf[i] * RR[(size_t)i * nc + c];
    Qdot[c] = q;
}

// ============================================================== host side
#define FCS_DECLARE_KSET(N) extern const KernelSet fcs_kset_##N;
FCS_FOR_EACH_NP(FCS_DECLARE_KSET)
FCS_FOR_EACH_NP_LARGE(FCS_DECLARE_KSET)
#undef FCS_DECLARE_KSET

namespace {

#define FCS_KSET_PTR(N) &fcs_kset_##N,
const KernelSet* const kKernels[] = {FCS_FOR_EACH_NP(FCS_KSET_PTR) FCS_FOR_EACH_NP_LARGE(FCS_KSET_PTR)};
#undef FCS_KSET_PTR

// Kernels of the smallest padded size holding n rows.  With the reaction
// table sizes given (nr >= 0), a workgroup-form size that keeps the matrix in
// LDS is passed over when its LDS would exceed a CU's; the next size, with
// the same species-table layout, then takes the mechanism.
const KernelSet* pick_kernels(int n, int nr = -1, int nTB = 0, int nPlog = 0)
{
    for (const KernelSet* k : kKernels) {
        if (k->NP < n) continue;
        if (nr >= 0 && jlds_of(k->NP) && lds_doubles(k->NP, nr, nTB, nPlog) * sizeof(double) > kMaxLdsBytes)
            continue;
        return k;
    }
    return nullptr;
}

#define HIP_OK(x)                                  \
    do {                                           \
        if ((x) != hipSuccess) return FCS_E_HIP;   \
    } while (0)


template <class T>
int grow(T*& p, size_t n)
{
    if (p) (void)hipFree(p);
    p = nullptr;
    return hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)) == hipSuccess ? FCS_OK : FCS_E_NOMEM;
}

template <class F>
int with_device(size_t bytes, F&& f)
{
    void* p = nullptr;
    if (hipMalloc(&p, std::max<size_t>(bytes, 8)) != hipSuccess) return FCS_E_NOMEM;
    int rc = f((char*)p);
    (void)hipFree(p);
    return rc;
}

}  // namespace

struct fcs_mech {
    DevMech dm;
    DevMech* dmDev = nullptr;   // device copy of dm, read by the kernels
    void* dbuf = nullptr;
    int ns = 0, nr = 0;
    std::vector<int> perm;   // device order -> file order
};

// Device buffers of one batch of cells (one allocation) plus the pinned
// host words and events of the termination poll.
struct DevCells {
    static constexpr int kPoll = 4;   // step launches in flight ahead of the poll
    size_t cap = 0;
    int ns = 0;
    void* base = nullptr;
    double *T0 = nullptr, *p0 = nullptr, *rho0 = nullptr, *Y0 = nullptr, *dtcIn = nullptr, *dtcOut = nullptr,
           *dtField = nullptr, *dxFirst = nullptr,
           *RR = nullptr, *Tend = nullptr, *Yw = nullptr, *x = nullptr, *dxTry = nullptr, *dx = nullptr,
           *dxTry0 = nullptr, *Y0s = nullptr, *Scl = nullptr, *theta = nullptr, *Jy = nullptr;
    int *status = nullptr, *nAtt = nullptr, *flags = nullptr, *nStep = nullptr, *list0 = nullptr, *list1 = nullptr,
        *kTarg = nullptr, *listG = nullptr;   // listG: gated proposals (positions in list0) of the learned seed
    size_t jyPerCell = 0;                    // doubles of Jy per cell (Seulex), 0 otherwise
    double* Jg = nullptr;                    // workgroup form: per-workgroup matrix slices
    size_t jgDoubles = 0;
    unsigned int* counts = nullptr;          // [3] active-list lengths
    unsigned long long* dtMin = nullptr;
    unsigned long long* nAccepted = nullptr; // learned proposals accepted in the last run
    float *Xf = nullptr, *Pf = nullptr;      // network inputs / outputs of the reacting cells
    size_t fcap = 0;
    long long lastPredicted = 0;
    // launch shape (autotuned or set by the caller): grid of the step kernel
    // (0: occupancy x CUs) and step launches per captured hipGraph (0: eager)
    int gridOverride = 0, graphLaunches = 0;
    hipGraphExec_t graphExec = nullptr;
    struct GraphKey { double dt; int nc, grid, lts, solver; } gkey{0.0, 0, 0, 0, 0};
    unsigned int* hostCount = nullptr;       // pinned [kPoll]
    hipEvent_t ev[kPoll] = {};
    // optional per-launch timing of the step kernel (event pair per launch)
    bool profile = false;
    std::vector<hipEvent_t> pev;
    long long lastLaunches = 0;
    double lastStepMs = 0.0;
    // profiling: the active-list length each step launch of the last run
    // read (eager launches), for the per-wavefront load report
    std::vector<long long> listLen;

    // jyDoubles: per-cell Jacobian storage of the Seulex solver (0 for the
    // Rosenbrock solvers).
    int alloc(size_t nc, int nsp, size_t jyDoubles)
    {
        release();
        const size_t d = align256(nc * 8), i = align256(nc * 4);
        const size_t nsd = align256(nc * nsp * 8), nwd = align256(nc * (nsp + 1) * 8);
        const size_t jyb = align256(nc * jyDoubles * 8);
        const size_t total = 13 * d + 2 * nsd + 3 * nwd + 8 * i + jyb + 3 * 256;
        if (hipMalloc(&base, total) != hipSuccess) {
            base = nullptr;
            return FCS_E_NOMEM;
        }
        char* q = (char*)base;
        auto take = [&](size_t bytes) { char* r = q; q += bytes; return r; };
        T0 = (double*)take(d); p0 = (double*)take(d); rho0 = (double*)take(d); dtcIn = (double*)take(d);
        dtcOut = (double*)take(d); Tend = (double*)take(d); x = (double*)take(d); dxTry = (double*)take(d);
        dx = (double*)take(d); dxTry0 = (double*)take(d); theta = (double*)take(d); dtField = (double*)take(d);
        dxFirst = (double*)take(d);
        Y0 = (double*)take(nsd); RR = (double*)take(nsd); Yw = (double*)take(nwd);
        Y0s = (double*)take(nwd); Scl = (double*)take(nwd);
        status = (int*)take(i); nAtt = (int*)take(i); flags = (int*)take(i); nStep = (int*)take(i);
        list0 = (int*)take(i); list1 = (int*)take(i); kTarg = (int*)take(i); listG = (int*)take(i);
        Jy = jyDoubles ? (double*)take(jyb) : nullptr;
        jyPerCell = jyDoubles;
        counts = (unsigned int*)take(256);
        dtMin = (unsigned long long*)take(256);
        nAccepted = (unsigned long long*)take(256);
        if (hipHostMalloc((void**)&hostCount, kPoll * sizeof(unsigned int), hipHostMallocDefault) != hipSuccess) {
            hostCount = nullptr;
            release();
            return FCS_E_NOMEM;
        }
        for (auto& e : ev)
            if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) {
                e = nullptr;
                release();
                return FCS_E_HIP;
            }
        cap = nc;
        ns = nsp;
        return FCS_OK;
    }
    int ensure_features(size_t rows, int nsp)
    {
        if (rows <= fcap) return FCS_OK;
        if (Xf) (void)hipFree(Xf), Xf = nullptr;
        if (Pf) (void)hipFree(Pf), Pf = nullptr;
        fcap = 0;
        if (hipMalloc(&Xf, rows * (nsp + 3) * 4) != hipSuccess || hipMalloc(&Pf, rows * (nsp + 2) * 4) != hipSuccess)
            return FCS_E_NOMEM;
        fcap = rows;
        return FCS_OK;
    }
    // per-workgroup matrix slices of the workgroup form (kept; regrown on demand)
    int ensure_jg(size_t doubles)
    {
        if (doubles <= jgDoubles) return FCS_OK;
        drop_graph();   // a captured graph holds the old address
        if (Jg) (void)hipFree(Jg), Jg = nullptr;
        jgDoubles = 0;
        if (hipMalloc(&Jg, doubles * 8) != hipSuccess) {
            Jg = nullptr;
            return FCS_E_NOMEM;
        }
        jgDoubles = doubles;
        return FCS_OK;
    }
    void drop_graph()
    {
        if (graphExec) (void)hipGraphExecDestroy(graphExec);
        graphExec = nullptr;
    }
    void release()
    {
        drop_graph();
        if (Xf) (void)hipFree(Xf), Xf = nullptr;
        if (Pf) (void)hipFree(Pf), Pf = nullptr;
        fcap = 0;
        if (Jg) (void)hipFree(Jg), Jg = nullptr;
        jgDoubles = 0;
        if (base) (void)hipFree(base);
        if (hostCount) (void)hipHostFree(hostCount);
        for (auto& e : ev)
            if (e) (void)hipEventDestroy(e), e = nullptr;
        for (auto& e : pev) (void)hipEventDestroy(e);
        pev.clear();
        base = nullptr;
        hostCount = nullptr;
        cap = 0;
    }
    CellArrays view(int nc, double dt, bool lts) const
    {
        CellArrays a;
        a.nc = nc;
        a.dt = dt;
        a.dtCell = lts ? dtField : nullptr;
        a.T0 = T0; a.p0 = p0; a.rho0 = rho0; a.Y0 = Y0; a.dtcIn = dtcIn;
        a.dtcOut = dtcOut; a.RR = RR; a.Tend = Tend; a.status = status; a.nAtt = nAtt; a.dtMinBits = dtMin;
        a.Yw = Yw; a.x = x; a.dxTry = dxTry; a.dx = dx; a.dxTry0 = dxTry0; a.flags = flags; a.nStep = nStep;
        a.Y0s = Y0s; a.Scl = Scl; a.theta = theta; a.Jy = Jy; a.kTarg = kTarg; a.dxFirst = dxFirst;
        a.Jg = Jg;
        return a;
    }
};

struct fcs_solver {
    const fcs_mech* mech = nullptr;
    fcs_solver_opts opts{};
    const KernelSet* ks = nullptr;
    size_t ldsBytes = 0;
    int grid = 0;
    hipStream_t stream = nullptr;
    DevCells cells;   // scratch of fcs_solve
};

struct fcs_batch {
    fcs_solver* s = nullptr;
    int64_t nc = 0;
    DevCells cells;
    fcs_mlp* mlp = nullptr;   // learned seed (fcs_batch_set_learned)
    int learnedMode = 0;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    double lastDt = 0.0;
};

extern "C" {

const char* fcs_version(void) { return "fcs 0.1 gfx950"; }

int fcs_device_count(int* count)
{
    if (!count) return FCS_E_ARG;
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    *count = c;
    return FCS_OK;
}

int fcs_mech_create(const fcs_mech_desc* d, fcs_mech** out)
{
    if (!d || !out || d->n_species < 1 || d->n_reactions < 0) return FCS_E_ARG;
    const KernelSet* kset = pick_kernels(d->n_species + 1);
    if (!kset) return FCS_E_UNSUPPORTED;
    fcs_mech* m = new (std::nothrow) fcs_mech();
    if (!m) return FCS_E_NOMEM;
    bool copyFailed = false;
    auto place = [&](const std::vector<unsigned char>& host) -> void* {
        if (hipMalloc(&m->dbuf, host.size()) != hipSuccess) {
            m->dbuf = nullptr;
            return nullptr;
        }
        if (hipMemcpy(m->dbuf, host.data(), host.size(), hipMemcpyHostToDevice) != hipSuccess) {
            copyFailed = true;
            return nullptr;
        }
        return m->dbuf;
    };
    const int rc = build_mech_tables(d, kset->NP, kset->G, place, m->dm, m->perm);
    if (rc != FCS_OK) {
        if (m->dbuf) (void)hipFree(m->dbuf);
        delete m;
        return copyFailed ? FCS_E_HIP : rc;
    }
    DevMech& dm = m->dm;
#ifdef FCS_PHASE_TIMING
    if (hipMalloc((void**)&dm.phase, PH_SLOTS * PH_SPREAD * sizeof(unsigned long long)) != hipSuccess ||
        hipMemset(dm.phase, 0, PH_SLOTS * PH_SPREAD * sizeof(unsigned long long)) != hipSuccess) {
        (void)hipFree(m->dbuf);
        delete m;
        return FCS_E_NOMEM;
    }
#endif
    if (hipMalloc((void**)&m->dmDev, sizeof(DevMech)) != hipSuccess ||
        hipMemcpy(m->dmDev, &dm, sizeof(DevMech), hipMemcpyHostToDevice) != hipSuccess) {
        fcs_mech_destroy(m);
        return FCS_E_NOMEM;
    }
    m->ns = d->n_species;
    m->nr = d->n_reactions;
    *out = m;
    return FCS_OK;
}

int fcs_mech_destroy(fcs_mech* m)
{
    if (!m) return FCS_E_ARG;
    if (m->dbuf) (void)hipFree(m->dbuf);
    if (m->dmDev) (void)hipFree(m->dmDev);
#ifdef FCS_PHASE_TIMING
    if (m->dm.phase) (void)hipFree(m->dm.phase);
#endif
    delete m;
    return FCS_OK;
}

#ifdef FCS_PHASE_TIMING
// Profiling build only: the per-phase cycle sums of every attempt made with
// this mechanism since the last reset (PH_* slots), optionally cleared.
int fcs_phase_read(fcs_mech* m, unsigned long long* out, int reset)
{
    if (!m || !out) return FCS_E_ARG;
    HIP_OK(hipDeviceSynchronize());
    std::vector<unsigned long long> h((size_t)PH_SLOTS * PH_SPREAD);
    HIP_OK(hipMemcpy(h.data(), m->dm.phase, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (int k = 0; k < PH_SLOTS; ++k) {
        out[k] = 0;
        for (int j = 0; j < PH_SPREAD; ++j) out[k] += h[(size_t)k * PH_SPREAD + j];
    }
    if (reset) HIP_OK(hipMemset(m->dm.phase, 0, h.size() * sizeof(unsigned long long)));
    return FCS_OK;
}
#endif

static int solver_grid(fcs_solver* s)
{
    int dev = 0, cus = 0, per = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, s->ks->step[s->opts.solver], s->ks->G, s->ldsBytes));
    s->grid = std::max(1, per) * std::max(1, cus);
    return FCS_OK;
}

int fcs_solver_create(const fcs_mech* mech, const fcs_solver_opts* opts, fcs_solver** out)
{
    if (!mech || !opts || !out) return FCS_E_ARG;
    if (opts->solver < 0 || opts->solver > FCS_SOLVER_ROSENBROCK23 || opts->jacobian < 0 || opts->jacobian > 1 ||
        opts->maxSteps < 1 || opts->maxSteps > 10000000 || !(opts->absTol > 0) || !(opts->relTol > 0))
        return FCS_E_ARG;
    const KernelSet* ks = pick_kernels(mech->ns + 1, mech->dm.nrq, mech->dm.nTB, mech->dm.nPlog);
    if (!ks || !ks->step[opts->solver]) return FCS_E_UNSUPPORTED;   // no Seulex in the workgroup form
    fcs_solver* s = new (std::nothrow) fcs_solver();
    if (!s) return FCS_E_NOMEM;
    s->mech = mech;
    s->opts = *opts;
    s->ks = ks;
    s->ldsBytes = lds_doubles(ks->NP, mech->dm.nrq, mech->dm.nTB, mech->dm.nPlog) * sizeof(double);
    if (s->ldsBytes > kMaxLdsBytes) {   // reaction arrays beyond one CU's LDS
        delete s;
        return FCS_E_UNSUPPORTED;
    }
    if (hipStreamCreate(&s->stream) != hipSuccess) {
        delete s;
        return FCS_E_HIP;
    }
    if (s->ldsBytes > 65536)
        for (const void* f : {ks->step[0], ks->step[1], ks->step[2], ks->step[3], ks->step[4], ks->deriv, ks->jac,
                              ks->lu, ks->attempt[0], ks->attempt[1], ks->attempt[3], ks->attempt[4], ks->calculate,
                              ks->tc, ks->verify})
            if (f) (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->ldsBytes);
    if (solver_grid(s) != FCS_OK) {
        (void)hipStreamDestroy(s->stream);
        delete s;
        return FCS_E_HIP;
    }
    *out = s;
    return FCS_OK;
}

int fcs_solver_destroy(fcs_solver* s)
{
    if (!s) return FCS_E_ARG;
    s->cells.release();
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return FCS_OK;
}

// Per-cell Jacobian storage of the Seulex solver (doubles), 0 otherwise.
static size_t jy_doubles(const fcs_solver* s)
{
    // per-cell Jy of the Seulex solver: NP columns of 64 (wavefront form) or NP rows
    const size_t NP = (size_t)s->ks->NP;
    return s->opts.solver == FCS_SOLVER_SEULEX ? NP * (NP > 64 ? NP : (size_t)WAVE) : 0;
}

static DevOpts dev_opts(const fcs_solver_opts& o)
{
    DevOpts d;
    d.solver = o.solver;
    d.jac = o.jacobian;
    d.maxSteps = o.maxSteps;
    d.absTol = o.absTol;
    d.relTol = o.relTol;
    d.Treact = o.Treact;
    d.dtChemMax = o.deltaTChemMax;
    return d;
}

// FastChemistryModel::solve(deltaT) over the nc cells held in d (inputs
// already on the device): setup, then one attempt per active cell per step
// launch until the active list is empty.  The list length after each launch
// is copied to pinned memory behind an event; the host waits only on the
// launch kPoll-1 places back, so the GPU always has step launches queued.
// Every cell finishes within 101 * maxSteps attempts (maxSteps accepted
// steps plus the rejection cap), which bounds the launch count.
static int run_cells(fcs_solver* s, DevCells& d, int nc, double dt, bool lts = false, fcs_mlp* mlp = nullptr,
                     int learnedMode = 0)
{
    hipStream_t st = s->stream;
    CellArrays a = d.view(nc, dt, lts);
    const DevMech* M = s->mech->dmDev;   // kernels stage it in LDS
    DevOpts o = dev_opts(s->opts);
    int ns = s->mech->ns;
    {
        void* args[] = {&d.counts, &d.dtMin, &d.nAccepted};
        HIP_OK(hipLaunchKernel((const void*)k_reset, dim3(1), dim3(1), args, 0, st));
    }
    if (nc > 0) {
        unsigned int* c0 = d.counts;
        void* args[] = {&o, &ns, &a, &d.list0, &c0};
        HIP_OK(hipLaunchKernel((const void*)k_setup, dim3((nc + 255) / 256), dim3(256), args, 0, st));
    }
    const void* f = s->ks->step[s->opts.solver];
    const int grid = std::max(1, std::min(d.gridOverride > 0 ? d.gridOverride : s->grid, nc));
    const int blk = s->ks->G;
    if (s->ks->NP > 64 && !jlds_of(s->ks->NP)) {
        if (d.ensure_jg((size_t)grid * ((size_t)s->ks->NP * s->ks->NP + s->ks->NP)) != FCS_OK) return FCS_E_NOMEM;
        a.Jg = d.Jg;
    }
    // learned seed: network proposals for the reacting cells, verified by one
    // Rodas34 step each; the rejected ones form list 1 and the step loop
    // starts there
    long long it0 = 0;
    d.lastPredicted = 0;
    if (mlp && learnedMode && nc > 0) {
        HIP_OK(hipMemcpyAsync(d.hostCount, d.counts, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        const unsigned int n0 = d.hostCount[0];
        d.lastPredicted = n0;
        if (n0 > 0) {
            if (d.ensure_features(n0, ns) != FCS_OK) return FCS_E_NOMEM;
            {
                unsigned int* c0 = d.counts;
                void* args[] = {&ns, &a, &d.list0, &c0, &d.Xf};
                HIP_OK(hipLaunchKernel((const void*)k_features, dim3((n0 + 255) / 256), dim3(256), args, 0, st));
            }
            const int rc = fcs_mlp_forward_device(mlp, (int)n0, ns + 3, ns + 2, d.Xf, d.Pf, st);
            if (rc != FCS_OK) return rc;
            int seed = learnedMode == 2 ? 1 : 0;
            unsigned int* c0 = d.counts;
            unsigned int* c1 = d.counts + 1;
            unsigned int* cG = d.counts + 3;
            {   // proposals with h >= deltaT/2 to the gated list, the others straight to list 1
                void* args[] = {&ns, &a, &d.list0, &c0, &d.Pf, &seed, &d.listG, &cG, &d.list1, &c1, &d.nAccepted};
                HIP_OK(hipLaunchKernel((const void*)k_gate, dim3((n0 + 255) / 256), dim3(256), args, 0, st));
            }
            void* args[] = {&M, &o, &a, &d.list0, &d.listG, &cG, &d.Pf, &seed, &d.list1, &c1, &d.nAccepted};
            HIP_OK(hipLaunchKernel(s->ks->verify, dim3(std::max(1, std::min(grid, (int)n0))), dim3(blk), args,
                                   s->ldsBytes, st));
            it0 = 1;
        }
    }
    const long long maxLaunches = 101LL * s->opts.maxSteps + 2 * DevCells::kPoll + 2;
    // step launch number `it` reads list it%2 and counter it%3, appends to the
    // other list and counter (it+1)%3, and clears counter (it+2)%3
    auto launch_args = [&](long long it, int*& lin, unsigned int*& cin, int*& lout, unsigned int*& cout,
                           unsigned int*& czero) {
        lin = (it & 1) ? d.list1 : d.list0;
        lout = (it & 1) ? d.list0 : d.list1;
        cin = d.counts + it % 3;
        cout = d.counts + (it + 1) % 3;
        czero = d.counts + (it + 2) % 3;
    };
    // launch cap reached: let the queued launches drain, then finish the
    // cells of the last output list as failed (FCS_E_CELLS)
    auto abandon = [&](int* list, unsigned int* count) -> int {
        HIP_OK(hipStreamSynchronize(st));
        unsigned int left = 0;
        HIP_OK(hipMemcpy(&left, count, sizeof left, hipMemcpyDeviceToHost));
        if (left > 0) {
            void* args[] = {&o, &ns, &a, &list, &count};
            HIP_OK(hipLaunchKernel((const void*)k_fail_list, dim3((left + 255) / 256), dim3(256), args, 0, st));
            HIP_OK(hipStreamSynchronize(st));
        }
        return FCS_E_CELLS;
    };
    d.lastStepMs = 0.0;
    if (d.graphLaunches > 0) {
        // hipGraph replays of G step launches (G a multiple of 6, so every
        // replay starts at the same list / counter phase); a few eager
        // launches first bring the phase to 0.  Launches after the list has
        // emptied return at once.  The host polls the last launch's counter
        // of the replay one replay back.
        const int G = d.graphLaunches;
        long long k = 0;
        for (; (k + it0) % 6 != 0; ++k) {
            int *lin, *lout;
            unsigned int *cin, *cout, *czero;
            launch_args(k + it0, lin, cin, lout, cout, czero);
            void* args[] = {&M, &o, &a, &lin, &cin, &lout, &cout, &czero};
            HIP_OK(hipLaunchKernel(f, dim3(grid), dim3(blk), args, s->ldsBytes, st));
        }
        const DevCells::GraphKey key{dt, nc, grid, lts ? 1 : 0, s->opts.solver};
        if (!d.graphExec || std::memcmp(&key, &d.gkey, sizeof key) != 0) {
            d.drop_graph();
            hipGraph_t g = nullptr;
            HIP_OK(hipStreamSynchronize(st));
            HIP_OK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            for (int j = 0; j < G; ++j) {
                int *lin, *lout;
                unsigned int *cin, *cout, *czero;
                launch_args(j, lin, cin, lout, cout, czero);
                void* args[] = {&M, &o, &a, &lin, &cin, &lout, &cout, &czero};
                if (hipLaunchKernel(f, dim3(grid), dim3(blk), args, s->ldsBytes, st) != hipSuccess) {
                    (void)hipStreamEndCapture(st, &g);
                    if (g) (void)hipGraphDestroy(g);
                    return FCS_E_HIP;
                }
            }
            HIP_OK(hipStreamEndCapture(st, &g));
            const hipError_t e = hipGraphInstantiate(&d.graphExec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                d.graphExec = nullptr;
                return FCS_E_HIP;
            }
            d.gkey = key;
        }
        unsigned int* lastOut = d.counts + (G - 1 + 1) % 3;   // output counter of launch G-1
        const long long maxReplays = maxLaunches / G + 2;
        while (d.pev.size() < 2) {
            hipEvent_t e;
            HIP_OK(hipEventCreate(&e));
            d.pev.push_back(e);
        }
        for (long long r = 0;; ++r) {
            if (d.profile) HIP_OK(hipEventRecord(d.pev[0], st));
            HIP_OK(hipGraphLaunch(d.graphExec, st));
            if (d.profile) {
                HIP_OK(hipEventRecord(d.pev[1], st));
                HIP_OK(hipEventSynchronize(d.pev[1]));
                float ms = 0.f;
                HIP_OK(hipEventElapsedTime(&ms, d.pev[0], d.pev[1]));
                d.lastStepMs += ms;
            }
            d.lastLaunches = k + (r + 1) * G;
            const int slot = (int)(r % 2);
            HIP_OK(hipMemcpyAsync(d.hostCount + slot, lastOut, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
            HIP_OK(hipEventRecord(d.ev[slot], st));
            if (r >= 1) {
                const int bs = (int)((r - 1) % 2);
                HIP_OK(hipEventSynchronize(d.ev[bs]));
                if (d.hostCount[bs] == 0u) break;
            }
            if (r > maxReplays) return abandon((G - 1) & 1 ? d.list0 : d.list1, lastOut);
        }
        HIP_OK(hipStreamSynchronize(st));
        return FCS_OK;
    }
    d.listLen.clear();
    if (d.profile) {   // the first launch's list length (later ones come with the polls)
        unsigned int* c0 = d.counts + it0 % 3;
        HIP_OK(hipMemcpyAsync(d.hostCount, c0, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        d.listLen.push_back(d.hostCount[0]);
    }
    for (long long k = 0;; ++k) {
        const long long it = k + it0;
        int *lin, *lout;
        unsigned int *cin, *cout, *czero;
        launch_args(it, lin, cin, lout, cout, czero);
        void* args[] = {&M, &o, &a, &lin, &cin, &lout, &cout, &czero};
        if (d.profile) {
            while (d.pev.size() < 2 * (size_t)(k + 1)) {
                hipEvent_t e;
                HIP_OK(hipEventCreate(&e));
                d.pev.push_back(e);
            }
            HIP_OK(hipEventRecord(d.pev[2 * k], st));
        }
        HIP_OK(hipLaunchKernel(f, dim3(grid), dim3(blk), args, s->ldsBytes, st));
        if (d.profile) HIP_OK(hipEventRecord(d.pev[2 * k + 1], st));
        d.lastLaunches = k + 1;
        const int slot = (int)(k % DevCells::kPoll);
        HIP_OK(hipMemcpyAsync(d.hostCount + slot, cout, sizeof(unsigned int), hipMemcpyDeviceToHost, st));
        HIP_OK(hipEventRecord(d.ev[slot], st));
        const long long back = k - (DevCells::kPoll - 1);
        if (back >= 0) {
            const int bs = (int)(back % DevCells::kPoll);
            HIP_OK(hipEventSynchronize(d.ev[bs]));
            if (d.hostCount[bs] == 0u) break;
            if (d.profile) d.listLen.push_back(d.hostCount[bs]);   // list read by launch back + 1
        }
        if (k > maxLaunches) return abandon(lout, cout);
    }
    d.lastStepMs = 0.0;
    if (d.profile) {
        HIP_OK(hipStreamSynchronize(st));
        for (long long i = 0; i < d.lastLaunches; ++i) {
            float ms = 0.f;
            HIP_OK(hipEventElapsedTime(&ms, d.pev[2 * i], d.pev[2 * i + 1]));
            d.lastStepMs += ms;
        }
    }
    return FCS_OK;
}

// solve(deltaT) / solve(scalarField deltaT) over host arrays: dtField == null
// selects the serial contract (FastChemistryModel_transientSolve.H:3-104,
// deltaTMin capped by 2*deltaT), else the local-time-step one
// (FastChemistryModel_localEulerSolve.H:3-101, per-cell deltaT, plain
// minimum).
static int solve_host(fcs_solver* s, int64_t nCells, double deltaT, const double* dtField, const double* T0,
                      const double* p0, const double* rho0, const double* const* Y0, double* deltaTChem,
                      double* const* RR, double* deltaTMin, int32_t* cellStatus, double* Tend, int32_t* nAttempts)
{
    if (!s || nCells < 0 || nCells > 0x7fffffff || (!dtField && !(deltaT > 0)) ||
        (nCells > 0 && (!T0 || !p0 || !rho0 || !Y0 || !deltaTChem || !RR)))
        return FCS_E_ARG;
    const int ns = s->mech->ns;
    const size_t nc = (size_t)nCells;
    for (size_t c = 0; dtField && c < nc; ++c)
        if (!(dtField[c] > 0)) return FCS_E_ARG;
    if (nc == 0) {
        if (deltaTMin) *deltaTMin = dtField ? OF_GREAT : std::min(OF_GREAT, 2 * deltaT);
        return FCS_OK;
    }
    DevCells& d = s->cells;
    const size_t jy = jy_doubles(s);
    if ((nc > d.cap || jy > d.jyPerCell) && d.alloc(nc, ns, jy) != FCS_OK) return FCS_E_NOMEM;
    hipStream_t st = s->stream;
    HIP_OK(hipMemcpyAsync(d.T0, T0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d.p0, p0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d.rho0, rho0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d.dtcIn, deltaTChem, nc * 8, hipMemcpyHostToDevice, st));
    if (dtField) HIP_OK(hipMemcpyAsync(d.dtField, dtField, nc * 8, hipMemcpyHostToDevice, st));
    for (int i = 0; i < ns; ++i)
        HIP_OK(hipMemcpyAsync(d.Y0 + (size_t)i * nc, Y0[i], nc * 8, hipMemcpyHostToDevice, st));
    int rc = run_cells(s, d, (int)nc, deltaT, dtField != nullptr);
    if (rc != FCS_OK && rc != FCS_E_CELLS) return rc;   // failed cells: results are still written
    for (int i = 0; i < ns; ++i)
        HIP_OK(hipMemcpyAsync(RR[i], d.RR + (size_t)i * nc, nc * 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipMemcpyAsync(deltaTChem, d.dtcOut, nc * 8, hipMemcpyDeviceToHost, st));
    std::vector<int> stat(nc);
    HIP_OK(hipMemcpyAsync(stat.data(), d.status, nc * 4, hipMemcpyDeviceToHost, st));
    if (Tend) HIP_OK(hipMemcpyAsync(Tend, d.Tend, nc * 8, hipMemcpyDeviceToHost, st));
    if (nAttempts) HIP_OK(hipMemcpyAsync(nAttempts, d.nAtt, nc * 4, hipMemcpyDeviceToHost, st));
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, d.dtMin, 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    double dmin;
    std::memcpy(&dmin, &bits, 8);
    dmin = std::min(dmin, OF_GREAT);
    if (deltaTMin) *deltaTMin = dtField ? dmin : std::min(dmin, 2 * deltaT);
    bool bad = false;
    for (size_t c = 0; c < nc; ++c) {
        if (cellStatus) cellStatus[c] = stat[c];
        if (stat[c] & (CELL_MAXSTEPS | CELL_NONFINITE)) bad = true;
    }
    return bad ? FCS_E_CELLS : FCS_OK;
}

int fcs_solve_ex(fcs_solver* s, int64_t nCells, double deltaT, const double* T0, const double* p0,
                 const double* rho0, const double* const* Y0, double* deltaTChem, double* const* RR,
                 double* deltaTMin, int32_t* cellStatus, double* Tend, int32_t* nAttempts)
{
    return solve_host(s, nCells, deltaT, nullptr, T0, p0, rho0, Y0, deltaTChem, RR, deltaTMin, cellStatus, Tend,
                      nAttempts);
}

int fcs_solve(fcs_solver* s, int64_t nCells, double deltaT, const double* T0, const double* p0, const double* rho0,
              const double* const* Y0, double* deltaTChem, double* const* RR, double* deltaTMin, int32_t* cellStatus)
{
    return solve_host(s, nCells, deltaT, nullptr, T0, p0, rho0, Y0, deltaTChem, RR, deltaTMin, cellStatus, nullptr,
                      nullptr);
}

int fcs_solve_lts(fcs_solver* s, int64_t nCells, const double* deltaT, const double* T0, const double* p0,
                  const double* rho0, const double* const* Y0, double* deltaTChem, double* const* RR,
                  double* deltaTMin, int32_t* cellStatus, double* Tend, int32_t* nAttempts)
{
    if (!deltaT && nCells > 0) return FCS_E_ARG;
    return solve_host(s, nCells, 0.0, deltaT, T0, p0, rho0, Y0, deltaTChem, RR, deltaTMin, cellStatus, Tend,
                      nAttempts);
}

// ------------------------------------------------------------ batches
int fcs_batch_create(fcs_solver* s, int64_t nCells, fcs_batch** out)
{
    if (!s || !out || nCells < 1 || nCells > 0x7fffffff) return FCS_E_ARG;
    fcs_batch* b = new (std::nothrow) fcs_batch();
    if (!b) return FCS_E_NOMEM;
    b->s = s;
    b->nc = nCells;
    if (b->cells.alloc((size_t)nCells, s->mech->ns, jy_doubles(s)) != FCS_OK || hipEventCreate(&b->e0) != hipSuccess ||
        hipEventCreate(&b->e1) != hipSuccess) {
        fcs_batch_destroy(b);
        return FCS_E_NOMEM;
    }
    *out = b;
    return FCS_OK;
}

int fcs_batch_destroy(fcs_batch* b)
{
    if (!b) return FCS_E_ARG;
    b->cells.release();
    if (b->e0) (void)hipEventDestroy(b->e0);
    if (b->e1) (void)hipEventDestroy(b->e1);
    delete b;
    return FCS_OK;
}

int fcs_batch_upload(fcs_batch* b, const double* T0, const double* p0, const double* rho0, const double* Y0,
                     const double* deltaTChem)
{
    if (!b || !T0 || !p0 || !rho0 || !Y0 || !deltaTChem) return FCS_E_ARG;
    const size_t nc = (size_t)b->nc, ns = (size_t)b->s->mech->ns;
    hipStream_t st = b->s->stream;
    DevCells& d = b->cells;
    HIP_OK(hipMemcpyAsync(d.T0, T0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d.p0, p0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d.rho0, rho0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d.Y0, Y0, nc * ns * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(d.dtcIn, deltaTChem, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    return FCS_OK;
}

int fcs_batch_run(fcs_batch* b, double deltaT, float* elapsed_ms)
{
    if (!b || !(deltaT > 0)) return FCS_E_ARG;
    fcs_solver* s = b->s;
    b->lastDt = deltaT;
    HIP_OK(hipEventRecord(b->e0, s->stream));
    int rc = run_cells(s, b->cells, (int)b->nc, deltaT, false, b->mlp, b->learnedMode);
    if (rc != FCS_OK && rc != FCS_E_CELLS) return rc;
    HIP_OK(hipEventRecord(b->e1, s->stream));
    HIP_OK(hipEventSynchronize(b->e1));
    if (elapsed_ms) HIP_OK(hipEventElapsedTime(elapsed_ms, b->e0, b->e1));
    return rc;
}

int fcs_batch_set_launch(fcs_batch* b, int32_t grid, int32_t graphLaunches)
{
    if (!b || grid < 0 || graphLaunches < 0 || graphLaunches % 6 != 0 || graphLaunches > 6000) return FCS_E_ARG;
    b->cells.gridOverride = grid;
    b->cells.graphLaunches = graphLaunches;
    b->cells.drop_graph();
    return FCS_OK;
}

int fcs_batch_autotune(fcs_batch* b, double deltaT, int32_t* grid, int32_t* graphLaunches)
{
    if (!b || !(deltaT > 0)) return FCS_E_ARG;
    fcs_solver* s = b->s;
    int dev = 0, cus = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int per = std::max(1, s->grid / std::max(cus, 1));
    float best = 0.f;
    int bestGrid = 0, bestGraph = 0;
    const bool wasProfile = b->cells.profile;
    b->cells.profile = false;
    for (int g : {0, 24})
        for (int k = 1; k <= per; ++k) {
            const int gr = k * cus;
            b->cells.gridOverride = gr;
            b->cells.graphLaunches = g;
            b->cells.drop_graph();
            float ms = 0.f;
            int rc = run_cells(s, b->cells, (int)b->nc, deltaT, false, b->mlp, b->learnedMode);   // warm-up / capture
            if (rc == FCS_OK) {
                HIP_OK(hipEventRecord(b->e0, s->stream));
                rc = run_cells(s, b->cells, (int)b->nc, deltaT, false, b->mlp, b->learnedMode);
            }
            if (rc != FCS_OK) {
                b->cells.profile = wasProfile;
                return rc;
            }
            HIP_OK(hipEventRecord(b->e1, s->stream));
            HIP_OK(hipEventSynchronize(b->e1));
            HIP_OK(hipEventElapsedTime(&ms, b->e0, b->e1));
            if (bestGrid == 0 || ms < best) {
                best = ms;
                bestGrid = gr;
                bestGraph = g;
            }
        }
    b->cells.profile = wasProfile;
    b->cells.gridOverride = bestGrid;
    b->cells.graphLaunches = bestGraph;
    b->cells.drop_graph();
    if (grid) *grid = bestGrid;
    if (graphLaunches) *graphLaunches = bestGraph;
    return FCS_OK;
}

int fcs_batch_set_learned(fcs_batch* b, fcs_mlp* mlp, int32_t mode)
{
    if (!b || mode < FCS_LEARNED_OFF || mode > FCS_LEARNED_SEED || (mode != FCS_LEARNED_OFF && !mlp)) return FCS_E_ARG;
    if (mode != FCS_LEARNED_OFF) {
        if (b->s->opts.solver == FCS_SOLVER_SEULEX || !b->s->ks->verify) return FCS_E_UNSUPPORTED;
        const int ns = b->s->mech->ns;
        // width check: a one-row forward pass on a scratch input
        float* x = nullptr;
        if (hipMalloc(&x, (size_t)(2 * ns + 5) * 4) != hipSuccess) return FCS_E_NOMEM;
        HIP_OK(hipMemset(x, 0, (size_t)(2 * ns + 5) * 4));
        const int rc = fcs_mlp_forward_device(mlp, 1, ns + 3, ns + 2, x, x + ns + 3, b->s->stream);
        (void)hipStreamSynchronize(b->s->stream);
        (void)hipFree(x);
        if (rc != FCS_OK) return rc;
    }
    b->mlp = mode ? mlp : nullptr;
    b->learnedMode = mode;
    return FCS_OK;
}

int fcs_batch_learned_stats(fcs_batch* b, int64_t* predicted, int64_t* verified, int64_t* accepted)
{
    if (!b) return FCS_E_ARG;
    unsigned long long acc[2] = {0, 0};
    HIP_OK(hipMemcpy(acc, b->cells.nAccepted, 16, hipMemcpyDeviceToHost));
    if (predicted) *predicted = b->cells.lastPredicted;
    if (verified) *verified = b->learnedMode ? (int64_t)acc[1] : 0;
    if (accepted) *accepted = b->learnedMode ? (int64_t)acc[0] : 0;
    return FCS_OK;
}

int fcs_batch_first_steps(fcs_batch* b, double* dxFirst)
{
    if (!b || !dxFirst) return FCS_E_ARG;
    HIP_OK(hipMemcpy(dxFirst, b->cells.dxFirst, (size_t)b->nc * 8, hipMemcpyDeviceToHost));
    return FCS_OK;
}

int fcs_batch_profile(fcs_batch* b, int enable)
{
    if (!b) return FCS_E_ARG;
    b->cells.profile = enable != 0;
    return FCS_OK;
}

int fcs_batch_stats(fcs_batch* b, int64_t* stepLaunches, double* stepKernelMs)
{
    if (!b) return FCS_E_ARG;
    if (stepLaunches) *stepLaunches = b->cells.lastLaunches;
    if (stepKernelMs) *stepKernelMs = b->cells.lastStepMs;
    return FCS_OK;
}

int fcs_batch_load_report(fcs_batch* b, int64_t* listLen, int32_t cap, int32_t* nLaunches, int32_t* grid)
{
    if (!b || cap < 0 || (cap > 0 && !listLen)) return FCS_E_ARG;
    const DevCells& d = b->cells;
    const int n = (int)d.listLen.size();
    for (int i = 0; i < n && i < cap; ++i) listLen[i] = d.listLen[i];
    if (nLaunches) *nLaunches = n;
    if (grid) *grid = std::max(1, std::min(d.gridOverride > 0 ? d.gridOverride : b->s->grid, (int)b->nc));
    return FCS_OK;
}

int fcs_batch_download(fcs_batch* b, double* RR, double* deltaTChem, double* deltaTMin, int32_t* cellStatus,
                       double* Tend, int32_t* nAttempts)
{
    if (!b) return FCS_E_ARG;
    const size_t nc = (size_t)b->nc, ns = (size_t)b->s->mech->ns;
    hipStream_t st = b->s->stream;
    DevCells& d = b->cells;
    if (RR) HIP_OK(hipMemcpyAsync(RR, d.RR, nc * ns * 8, hipMemcpyDeviceToHost, st));
    if (deltaTChem) HIP_OK(hipMemcpyAsync(deltaTChem, d.dtcOut, nc * 8, hipMemcpyDeviceToHost, st));
    if (cellStatus) HIP_OK(hipMemcpyAsync(cellStatus, d.status, nc * 4, hipMemcpyDeviceToHost, st));
    if (Tend) HIP_OK(hipMemcpyAsync(Tend, d.Tend, nc * 8, hipMemcpyDeviceToHost, st));
    if (nAttempts) HIP_OK(hipMemcpyAsync(nAttempts, d.nAtt, nc * 4, hipMemcpyDeviceToHost, st));
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, d.dtMin, 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (deltaTMin) {
        double dmin;
        std::memcpy(&dmin, &bits, 8);
        dmin = std::min(dmin, OF_GREAT);
        *deltaTMin = std::min(dmin, 2 * b->lastDt);
    }
    return FCS_OK;
}

// ------------------------------------------------------------ no-allocation solve
// fcs_solve_noalloc keeps the integrator's per-cell state, the active lists
// and the counters in this array, one copy per GPU, held until the process
// ends.  Host code must not pass the array's name to a kernel (it is a
// host-side stand-in there): the device address comes from
// hipGetSymbolAddress, once per device, and is kept in gScratchPtr.
constexpr size_t kScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void* gScratchPtr[kMaxDevices] = {};

int scratch_of_current_device(char** out)
{
    int dev = -1;
    HIP_OK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return FCS_E_HIP;
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchPtr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return FCS_E_HIP;
        gScratchPtr[dev] = p;
    }
    *out = (char*)gScratchPtr[dev];
    return FCS_OK;
}

// Carves the per-cell state of nc cells (and the Seulex extras when
// jyDoubles > 0) out of the scratch at base; returns the bytes used.
size_t carve_scratch(char* base, size_t nc, int ns, size_t jyDoubles, CellArrays& a, int*& list0, int*& list1,
                     unsigned int*& counts, unsigned long long*& nAccepted)
{
    size_t off = 0;
    auto take = [&](size_t bytes) { char* r = base + off; off += align256(bytes); return r; };
    counts = (unsigned int*)take(4 * sizeof(unsigned int));
    a.dtMinBits = (unsigned long long*)take(8);
    nAccepted = (unsigned long long*)take(16);
    a.Yw = (double*)take(nc * (ns + 1) * 8);
    a.x = (double*)take(nc * 8);
    a.dxTry = (double*)take(nc * 8);
    a.dx = (double*)take(nc * 8);
    a.dxTry0 = (double*)take(nc * 8);
    a.dxFirst = (double*)take(nc * 8);
    a.flags = (int*)take(nc * 4);
    a.nStep = (int*)take(nc * 4);
    a.nAtt = (int*)take(nc * 4);
    list0 = (int*)take(nc * 4);
    list1 = (int*)take(nc * 4);
    a.Y0s = a.Scl = a.theta = a.Jy = nullptr;
    a.kTarg = nullptr;
    if (jyDoubles) {
        a.Y0s = (double*)take(nc * (ns + 1) * 8);
        a.Scl = (double*)take(nc * (ns + 1) * 8);
        a.theta = (double*)take(nc * 8);
        a.kTarg = (int*)take(nc * 4);
        a.Jy = (double*)take(nc * jyDoubles * 8);
    }
    return off;
}
}  // namespace

int fcs_solve_noalloc(fcs_solver* s, int64_t nCells, double deltaT, const double* T0, const double* p0,
                      const double* rho0, const double* Y0, const double* deltaTChemIn, double* deltaTChemOut,
                      double* RR, int32_t* cellStatus, double* Tend, double* deltaTMin)
{
    if (!s || nCells < 0 || nCells > 0x7fffffff || !(deltaT > 0) ||
        (nCells > 0 && (!T0 || !p0 || !rho0 || !Y0 || !deltaTChemIn || !deltaTChemOut || !RR || !cellStatus ||
                        !Tend)))
        return FCS_E_ARG;
    // the workgroup form with its matrix in HBM needs per-workgroup slices
    // beyond this scratch
    if (s->ks->NP > 64 && !jlds_of(s->ks->NP)) return FCS_E_UNSUPPORTED;
    const int ns = s->mech->ns;
    const size_t nc = (size_t)nCells;
    char* base = nullptr;
    const int rc0 = scratch_of_current_device(&base);
    if (rc0 != FCS_OK) return rc0;
    CellArrays a{};
    int *list0 = nullptr, *list1 = nullptr;
    unsigned int* counts = nullptr;
    unsigned long long* nAccepted = nullptr;
    if (carve_scratch(base, nc, ns, jy_doubles(s), a, list0, list1, counts, nAccepted) > kScratchBytes)
        return FCS_E_SCRATCH;
    a.nc = (int)nc;
    a.dt = deltaT;
    a.dtCell = nullptr;
    a.T0 = T0; a.p0 = p0; a.rho0 = rho0; a.Y0 = Y0; a.dtcIn = deltaTChemIn;
    a.dtcOut = deltaTChemOut; a.RR = RR; a.Tend = Tend; a.status = cellStatus;
    a.Jg = nullptr;
    hipStream_t st = s->stream;
    const DevMech* M = s->mech->dmDev;   // kernels stage it in LDS
    DevOpts o = dev_opts(s->opts);
    int nsp = ns;
    {
        void* args[] = {&counts, &a.dtMinBits, &nAccepted};
        HIP_OK(hipLaunchKernel((const void*)k_reset, dim3(1), dim3(1), args, 0, st));
    }
    if (nc > 0) {
        unsigned int* c0 = counts;
        void* args[] = {&o, &nsp, &a, &list0, &c0};
        HIP_OK(hipLaunchKernel((const void*)k_setup, dim3((unsigned)((nc + 255) / 256)), dim3(256), args, 0, st));
        const void* f = s->ks->step[s->opts.solver];
        const int grid = std::max(1, std::min(s->grid, (int)nc));
        // same list / counter rotation as run_cells; the host reads the
        // active-list length after every kPollEvery launches with a
        // synchronous copy into a stack word (no pinned buffer, no events)
        constexpr int kPollEvery = 8;
        const long long maxLaunches = 101LL * s->opts.maxSteps + 2 * kPollEvery + 2;
        for (long long k = 0;; ++k) {
            int* lin = (k & 1) ? list1 : list0;
            int* lout = (k & 1) ? list0 : list1;
            unsigned int* cin = counts + k % 3;
            unsigned int* cout = counts + (k + 1) % 3;
            unsigned int* czero = counts + (k + 2) % 3;
            void* args[] = {&M, &o, &a, &lin, &cin, &lout, &cout, &czero};
            HIP_OK(hipLaunchKernel(f, dim3(grid), dim3(s->ks->G), args, s->ldsBytes, st));
            if ((k + 1) % kPollEvery == 0) {
                unsigned int left = 0;
                HIP_OK(hipMemcpyAsync(&left, cout, sizeof left, hipMemcpyDeviceToHost, st));
                HIP_OK(hipStreamSynchronize(st));
                if (left == 0u) break;
            }
            if (k > maxLaunches) {
                // out of reach while every cell obeys its attempt cap: finish
                // the cells still in flight as failed, as run_cells does
                HIP_OK(hipStreamSynchronize(st));
                unsigned int left = 0;
                HIP_OK(hipMemcpy(&left, cout, sizeof left, hipMemcpyDeviceToHost));
                void* fargs[] = {&o, &nsp, &a, &lout, &cout};
                if (left > 0)
                    HIP_OK(hipLaunchKernel((const void*)k_fail_list, dim3((left + 255) / 256), dim3(256), fargs, 0,
                                           st));
                break;
            }
        }
    }
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, a.dtMinBits, 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (deltaTMin) {
        double dmin;
        std::memcpy(&dmin, &bits, 8);
        dmin = std::min(dmin, OF_GREAT);
        *deltaTMin = std::min(dmin, 2 * deltaT);
    }
    return FCS_OK;
}

// ------------------------------------------------------------ kernel-level entries

// HBM matrix doubles per workgroup of the workgroup form with the matrix in
// HBM (0 otherwise).
static size_t jg_per_block(const fcs_solver* s)
{
    const size_t NP = (size_t)s->ks->NP;
    return NP > 64 && !jlds_of((int)NP) ? NP * NP + NP : 0;
}

int fcs_eval_derivatives(fcs_solver* s, int64_t nS, const double* p, double* YT, double* dYTdt)
{
    if (!s || nS < 1 || nS > 0x7fffffff || !p || !YT || !dYTdt) return FCS_E_ARG;
    const int n = s->mech->ns + 1;
    const size_t nv = (size_t)n * nS;
    return with_device((nS + 2 * nv) * 8, [&](char* base) -> int {
        double* dp = (double*)base;
        double* dy = dp + nS;
        double* dd = dy + nv;
        HIP_OK(hipMemcpy(dp, p, nS * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dy, YT, nv * 8, hipMemcpyHostToDevice));
        const DevMech* M = s->mech->dmDev;   // kernels stage it in LDS
        int ns32 = (int)nS;
        void* args[] = {&M, &ns32, &dp, &dy, &dd};
        HIP_OK(hipLaunchKernel(s->ks->deriv, dim3((unsigned)nS), dim3(s->ks->G), args, s->ldsBytes, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(YT, dy, nv * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(dYTdt, dd, nv * 8, hipMemcpyDeviceToHost));
        return FCS_OK;
    });
}

int fcs_eval_jacobian(fcs_solver* s, int64_t nS, const double* p, double* YT, double* dYTdt, double* Jy)
{
    if (!s || nS < 1 || nS > 0x7fffffff || !p || !YT || !dYTdt || !Jy) return FCS_E_ARG;
    const int n = s->mech->ns + 1;
    const size_t nv = (size_t)n * nS, nm = (size_t)n * n * nS, ng = jg_per_block(s) * nS;
    return with_device((nS + 2 * nv + nm + ng) * 8, [&](char* base) -> int {
        double* dp = (double*)base;
        double* dy = dp + nS;
        double* df = dy + nv;
        double* dJ = df + nv;
        double* dG = ng ? dJ + nm : nullptr;
        HIP_OK(hipMemcpy(dp, p, nS * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dy, YT, nv * 8, hipMemcpyHostToDevice));
        const DevMech* M = s->mech->dmDev;   // kernels stage it in LDS
        int jt = s->opts.jacobian, ns32 = (int)nS;
        void* args[] = {&M, &jt, &ns32, &dp, &dy, &df, &dJ, &dG};
        HIP_OK(hipLaunchKernel(s->ks->jac, dim3((unsigned)nS), dim3(s->ks->G), args, s->ldsBytes, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(YT, dy, nv * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(dYTdt, df, nv * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(Jy, dJ, nm * 8, hipMemcpyDeviceToHost));
        return FCS_OK;
    });
}

int fcs_eval_attempt(fcs_solver* s, int64_t nS, const double* p, const double* dx, double* YT, double* YTnew,
                     double* err)
{
    if (!s || nS < 1 || nS > 0x7fffffff || !p || !dx || !YT || !YTnew || !err) return FCS_E_ARG;
    if (s->opts.solver == FCS_SOLVER_SEULEX) return FCS_E_UNSUPPORTED;
    const int n = s->mech->ns + 1;
    const size_t nv = (size_t)n * nS, ng = jg_per_block(s) * nS;
    return with_device((3 * nS + 2 * nv + ng) * 8, [&](char* base) -> int {
        double* dp = (double*)base;
        double* ddx = dp + nS;
        double* de = ddx + nS;
        double* dy = de + nS;
        double* dn = dy + nv;
        double* dG = ng ? dn + nv : nullptr;
        HIP_OK(hipMemcpy(dp, p, nS * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(ddx, dx, nS * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dy, YT, nv * 8, hipMemcpyHostToDevice));
        const DevMech* M = s->mech->dmDev;   // kernels stage it in LDS
        DevOpts o = dev_opts(s->opts);
        int ns32 = (int)nS;
        void* args[] = {&M, &o, &ns32, &dp, &ddx, &dy, &dn, &de, &dG};
        HIP_OK(hipLaunchKernel(s->ks->attempt[s->opts.solver], dim3((unsigned)nS), dim3(s->ks->G), args,
                               s->ldsBytes, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(YT, dy, nv * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(YTnew, dn, nv * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(err, de, nS * 8, hipMemcpyDeviceToHost));
        return FCS_OK;
    });
}

// Per-species host fields -> device [ns][nc] (rho, T, p first), one
// wavefront per cell; out receives nOut rows of nc doubles.
static int field_kernel(fcs_solver* s, const void* kern, int64_t nCells, const double* rho, const double* T,
                        const double* p, const double* const* Y, int nOut, double* const* out)
{
    if (!s || nCells < 0 || nCells > 0x7fffffff || (nCells > 0 && (!rho || !T || !p || !Y || !out))) return FCS_E_ARG;
    const int ns = s->mech->ns;
    const size_t nc = (size_t)nCells;
    if (nc == 0) return FCS_OK;
    for (int i = 0; i < ns; ++i)
        if (!Y[i]) return FCS_E_ARG;
    for (int i = 0; i < nOut; ++i)
        if (!out[i]) return FCS_E_ARG;
    return with_device((3 + ns + nOut) * nc * 8, [&](char* base) -> int {
        double* dRho = (double*)base;
        double* dT = dRho + nc;
        double* dp = dT + nc;
        double* dY = dp + nc;
        double* dOut = dY + (size_t)ns * nc;
        hipStream_t st = s->stream;
        HIP_OK(hipMemcpyAsync(dRho, rho, nc * 8, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(dT, T, nc * 8, hipMemcpyHostToDevice, st));
        HIP_OK(hipMemcpyAsync(dp, p, nc * 8, hipMemcpyHostToDevice, st));
        for (int i = 0; i < ns; ++i)
            HIP_OK(hipMemcpyAsync(dY + (size_t)i * nc, Y[i], nc * 8, hipMemcpyHostToDevice, st));
        const DevMech* M = s->mech->dmDev;   // kernels stage it in LDS
        int nc32 = (int)nc;
        void* args[] = {&M, &nc32, &dRho, &dT, &dp, &dY, &dOut};
        HIP_OK(hipLaunchKernel(kern, dim3((unsigned)nc), dim3(s->ks->G), args, s->ldsBytes, st));
        for (int i = 0; i < nOut; ++i)
            HIP_OK(hipMemcpyAsync(out[i], dOut + (size_t)i * nc, nc * 8, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        return FCS_OK;
    });
}

int fcs_calculate(fcs_solver* s, int64_t nCells, const double* rho, const double* T, const double* p,
                  const double* const* Y, double* const* RR)
{
    if (!s) return FCS_E_ARG;
    return field_kernel(s, s->ks->calculate, nCells, rho, T, p, Y, s->mech->ns, RR);
}

int fcs_tc(fcs_solver* s, int64_t nCells, const double* rho, const double* T, const double* p,
           const double* const* Y, double* tc)
{
    if (!s) return FCS_E_ARG;
    double* out[1] = {tc};
    return field_kernel(s, s->ks->tc, nCells, rho, T, p, Y, 1, out);
}

int fcs_qdot(fcs_solver* s, int64_t nCells, const double* const* RR, double* Qdot)
{
    if (!s || nCells < 0 || nCells > 0x7fffffff || (nCells > 0 && (!RR || !Qdot))) return FCS_E_ARG;
    const int ns = s->mech->ns;
    const size_t nc = (size_t)nCells;
    if (nc == 0) return FCS_OK;
    for (int i = 0; i < ns; ++i)
        if (!RR[i]) return FCS_E_ARG;
    return with_device((ns + 1) * nc * 8, [&](char* base) -> int {
        double* dRR = (double*)base;
        double* dQ = dRR + (size_t)ns * nc;
        hipStream_t st = s->stream;
        for (int i = 0; i < ns; ++i)
            HIP_OK(hipMemcpyAsync(dRR + (size_t)i * nc, RR[i], nc * 8, hipMemcpyHostToDevice, st));
        const double* hf = s->mech->dm.Hf.p;
        int nc32 = (int)nc, ns32 = ns;
        void* args[] = {&nc32, &ns32, &hf, &dRR, &dQ};
        HIP_OK(hipLaunchKernel((const void*)k_qdot, dim3((unsigned)((nc + 255) / 256)), dim3(256), args, 0, st));
        HIP_OK(hipMemcpyAsync(Qdot, dQ, nc * 8, hipMemcpyDeviceToHost, st));
        HIP_OK(hipStreamSynchronize(st));
        return FCS_OK;
    });
}

int fcs_lu_solve(fcs_solver* s, int64_t nSys, const double* A, double* b)
{
    if (!s || nSys < 1 || nSys > 0x7fffffff || !A || !b) return FCS_E_ARG;
    const int n = s->mech->ns + 1;
    const size_t nm = (size_t)n * n * nSys, nv = (size_t)n * nSys, ng = jg_per_block(s) * nSys;
    return with_device((nm + nv + ng) * 8, [&](char* base) -> int {
        double* dA = (double*)base;
        double* db = dA + nm;
        double* dG = ng ? db + nv : nullptr;
        HIP_OK(hipMemcpy(dA, A, nm * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(db, b, nv * 8, hipMemcpyHostToDevice));
        const DevMech* M = s->mech->dmDev;   // kernels stage it in LDS
        int n32 = (int)nSys;
        void* args[] = {&M, &n32, &dA, &db, &dG};
        HIP_OK(hipLaunchKernel(s->ks->lu, dim3((unsigned)nSys), dim3(s->ks->G), args, s->ldsBytes, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(b, db, nv * 8, hipMemcpyDeviceToHost));
        return FCS_OK;
    });
}

}  // extern "C"
