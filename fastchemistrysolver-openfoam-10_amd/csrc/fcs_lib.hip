// fcs_lib.hip -- kernels and C ABI (include/fcs.h) of the gfx950 chemistry
// integrator.  Built with hipcc --offload-arch=gfx950 into libfcs.so.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "fcs.h"
#include "fcs_integrate.h"

using namespace fcs;

// Padded system sizes compiled into the library (T sits at NP-1).
#define FCS_FOR_EACH_NP(X) X(12) X(56)

// ============================================================== kernels
// Persistent cell queue: every wavefront takes the next cell index from an
// atomic counter until the batch is drained, so cells with very different
// sub-step counts balance across the chip without host involvement.  This
// replaces the reference's MPI load balancer
// (src/ChemistryModel/FastChemistryModel_transientSolve.H:106-686).
template <int NP, int SOLVER>
__global__ __launch_bounds__(64) void k_integrate(DevMech M, DevOpts o, int nc, double dt, const double* T0,
                                                  const double* p0, const double* rho0, const double* Y0,
                                                  const double* dtcIn, double* dtcOut, double* RR,
                                                  int* status, double* Tend, int* nAttOut,
                                                  unsigned long long* dtMinBits, unsigned int* counter)
{
    extern __shared__ double lds_raw[];
    const Lds L = make_lds(lds_raw, NP, M.nr, M.nTB);
    const int lane = lane_id();
    const bool sp = lane < M.ns;
    for (;;) {
        unsigned int c = 0;
        if (lane == 0) c = atomicAdd(counter, 1u);
        c = __shfl(c, 0, WAVE);
        if (c >= (unsigned int)nc) break;
        const double T = T0[c];
        double Yin = 0.0;
        if (sp) Yin = Y0[(size_t)lane * nc + c];
        if (!(T > o.Treact)) {
            // src/ChemistryModel/FastChemistryModel_transientSolve.H:70-76
            if (sp) RR[(size_t)lane * nc + c] = 0.0;
            if (lane == 0) {
                status[c] = CELL_SKIPPED;
                dtcOut[c] = dtcIn[c];
                if (Tend) Tend[c] = T;
                if (nAttOut) nAttOut[c] = 0;
            }
            continue;
        }
        const double p = p0[c];
        // y00 = [max(0, Y), T] (src/ODE/OptRodas34/OptRodas34.C:72-81)
        double y = sp ? (Yin > 0.0 ? Yin : 0.0) : (lane == NP - 1 ? T : 0.0);
        double dxTry = dtcIn[c];
        int nAtt = 0;
        int st;
        if (SOLVER == SOLVER_SEULEX) st = seulex_integrate<NP>(M, L, o, p, dt, y, dxTry, nAtt);
        else st = rosenbrock_integrate<NP>(M, L, o, SOLVER == SOLVER_ROS34 ? kRos34 : kRodas34, p, dt, y, dxTry,
                                          nAtt);
        if (wave_any(!(y == y) || isinf(y))) st |= CELL_NONFINITE;
        // writeback (OptRodas34.C:114-117, transientSolve.H:59-67)
        if (sp) {
            const double Yn = y > 0.0 ? y : 0.0;
            const double r0 = rho0[c];
            RR[(size_t)lane * nc + c] = (Yn * r0 - Yin * r0) / dt;
        }
        if (lane == NP - 1 && Tend) Tend[c] = y;
        if (lane == 0) {
            status[c] = st;
            if (nAttOut) nAttOut[c] = nAtt;
            dtcOut[c] = dxTry < o.dtChemMax ? dxTry : o.dtChemMax;
            if (dxTry > 0.0) atomicMin(dtMinBits, (unsigned long long)__double_as_longlong(dxTry));
        }
    }
}

// Resets the queue counter and the deltaTMin accumulator on the stream
// (a kernel rather than memset/memcpy nodes, so the sequence can be captured).
__global__ void k_queue_init(unsigned int* counter, unsigned long long* dtMinBits)
{
    *counter = 0u;
    *dtMinBits = 0x7ff0000000000000ull;  // +inf
}

// One wavefront per state: derivatives.  External layout [n][nS] with T at
// row ns.
template <int NP>
__global__ __launch_bounds__(64) void k_derivatives(DevMech M, int nS, const double* p, double* YT, double* out)
{
    extern __shared__ double lds_raw[];
    const Lds L = make_lds(lds_raw, NP, M.nr, M.nTB);
    const int s = blockIdx.x, lane = lane_id(), ns = M.ns;
    const int ext = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    double y = ext >= 0 ? YT[(size_t)ext * nS + s] : 0.0;
    const double d = derivatives<NP>(M, L, p[s], log(p[s]), y);
    if (ext >= 0) {
        YT[(size_t)ext * nS + s] = y;
        out[(size_t)ext * nS + s] = d;
    }
}

template <int NP>
__global__ __launch_bounds__(64) void k_jacobian(DevMech M, int jacType, int nS, const double* p, double* YT,
                                                 double* f, double* J)
{
    extern __shared__ double lds_raw[];
    const Lds L = make_lds(lds_raw, NP, M.nr, M.nTB);
    const int s = blockIdx.x, lane = lane_id(), ns = M.ns, n = ns + 1;
    const int ext = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    double y = ext >= 0 ? YT[(size_t)ext * nS + s] : 0.0;
    double a[NP];
    const double d = jacobian<NP>(M, L, p[s], log(p[s]), jacType, y, a);
    if (ext >= 0) {
        YT[(size_t)ext * nS + s] = y;
        f[(size_t)ext * nS + s] = d;
        double* row = J + ((size_t)s * n + ext) * n;
#pragma unroll
        for (int j = 0; j < NP; ++j) {
            const int ej = j < ns ? j : (j == NP - 1 ? ns : -1);
            if (ej >= 0) row[ej] = a[j];
        }
    }
}

template <int NP>
__global__ __launch_bounds__(64) void k_lu_solve(DevMech M, int nSys, const double* A, double* b)
{
    extern __shared__ double lds_raw[];
    const Lds L = make_lds(lds_raw, NP, M.nr, M.nTB);
    const int s = blockIdx.x, lane = lane_id(), ns = M.ns, n = ns + 1;
    const int ext = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    double a[NP];
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const int ej = j < ns ? j : (j == NP - 1 ? ns : -1);
        a[j] = (ext >= 0 && ej >= 0) ? A[((size_t)s * n + ext) * n + ej] : ((lane == j && lane < NP) ? 1.0 : 0.0);
    }
    const LuState st = lu_factor<NP>(L, a);
    const double x = lu_solve<NP>(a, st, ext >= 0 ? b[(size_t)s * n + ext] : 0.0);
    if (ext >= 0) b[(size_t)s * n + ext] = x;
}

// ============================================================== host side
namespace {

struct KernelSet {
    int NP;
    const void* integrate[3];   // by FCS_SOLVER_*
    const void* deriv;
    const void* jac;
    const void* lu;
};

#define FCS_KSET(N)                                                                                        \
    {N,                                                                                                    \
     {(const void*)k_integrate<N, SOLVER_RODAS34>, (const void*)k_integrate<N, SOLVER_ROS34>,              \
      (const void*)k_integrate<N, SOLVER_SEULEX>},                                                         \
     (const void*)k_derivatives<N>, (const void*)k_jacobian<N>, (const void*)k_lu_solve<N>},
const KernelSet kKernels[] = {FCS_FOR_EACH_NP(FCS_KSET)};
#undef FCS_KSET

const KernelSet* pick_kernels(int n)
{
    for (const auto& k : kKernels)
        if (k.NP >= n) return &k;
    return nullptr;
}

#define HIP_OK(x)                                  \
    do {                                           \
        if ((x) != hipSuccess) return FCS_E_HIP;   \
    } while (0)

size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

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
    void* dbuf = nullptr;
    int ns = 0, nr = 0;
    std::vector<int> perm;   // device order -> file order
};

struct fcs_batch;

struct fcs_solver {
    const fcs_mech* mech = nullptr;
    fcs_solver_opts opts{};
    const KernelSet* ks = nullptr;
    size_t ldsBytes = 0;
    int grid = 0;
    hipStream_t stream = nullptr;
    // scratch buffers of fcs_solve
    size_t cap = 0;
    double *T0 = nullptr, *p0 = nullptr, *rho0 = nullptr, *Y0 = nullptr, *dtcIn = nullptr, *dtcOut = nullptr,
           *RR = nullptr, *Tend = nullptr;
    int *status = nullptr, *nAtt = nullptr;
    unsigned long long* dtMin = nullptr;
    unsigned int* counter = nullptr;
};

struct fcs_batch {
    fcs_solver* s = nullptr;
    int64_t nc = 0;
    double *T0 = nullptr, *p0 = nullptr, *rho0 = nullptr, *Y0 = nullptr, *dtcIn = nullptr, *dtcOut = nullptr,
           *RR = nullptr, *Tend = nullptr;
    int *status = nullptr, *nAtt = nullptr;
    unsigned long long* dtMin = nullptr;
    unsigned int* counter = nullptr;
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
    if (!d || !out || d->n_species < 1 || d->n_species > 63 || d->n_reactions < 0) return FCS_E_ARG;
    const int ns = d->n_species, nr = d->n_reactions;
    // ---- order reactions by rate form (file order within a form)
    std::vector<int> perm;
    for (int t = 0; t <= FCS_RXN_PLOG; ++t)
        for (int r = 0; r < nr; ++r)
            if (d->rxn_type[r] == t) perm.push_back(r);
    if ((int)perm.size() != nr) return FCS_E_ARG;
    for (int r = 0; r < nr; ++r) {
        if (d->rxn_n_lhs[r] < 1 || d->rxn_n_lhs[r] > 3 || d->rxn_n_rhs[r] < 1 || d->rxn_n_rhs[r] > 3)
            return FCS_E_ARG;
        for (int q = 0; q < d->rxn_n_lhs[r]; ++q)
            if (d->rxn_lhs[3 * r + q] < 0 || d->rxn_lhs[3 * r + q] >= ns) return FCS_E_ARG;
        for (int q = 0; q < d->rxn_n_rhs[r]; ++q)
            if (d->rxn_rhs[3 * r + q] < 0 || d->rxn_rhs[3 * r + q] >= ns) return FCS_E_ARG;
    }
    const int pmax = d->plog_max_points > 0 ? d->plog_max_points : 1;
    // ---- host images of every table
    std::vector<double> W(64, 0.0), invW(64, 0.0), Tc(64, 1000.0), clo(7 * 64, 0.0), chi(7 * 64, 0.0);
    double TlowMin = d->T_low[0], ThighMax = d->T_high[0], TcMin = d->T_common[0], TcMax = d->T_common[0];
    for (int s = 0; s < ns; ++s) {
        W[s] = d->W[s];
        invW[s] = 1.0 / d->W[s];
        Tc[s] = d->T_common[s];
        for (int k = 0; k < 7; ++k) {
            clo[k * 64 + s] = d->coeffs_low[7 * s + k];
            chi[k * 64 + s] = d->coeffs_high[7 * s + k];
        }
        TlowMin = std::min(TlowMin, d->T_low[s]);
        ThighMax = std::max(ThighMax, d->T_high[s]);
        TcMin = std::min(TcMin, d->T_common[s]);
        TcMax = std::max(TcMax, d->T_common[s]);
    }
    std::vector<int> rtype(nr), rrev(nr), rnl(nr), rnr(nr), rsp(6 * std::max(nr, 1), 0), rtb(nr, -1), raux(nr, -1);
    std::vector<double> rA(nr), rbeta(nr), rTa(nr), rA0(nr), rbeta0(nr), rTa0(nr);
    std::vector<double> troe, eff, plogLnp, plogLnA, plogBeta, plogTa;
    std::vector<int> plogN;
    int nTB = 0;
    for (int r = 0; r < nr; ++r) {
        const int f = perm[r];
        const int t = d->rxn_type[f];
        rtype[r] = t;
        rrev[r] = d->rxn_reversible[f] ? 1 : 0;
        rnl[r] = d->rxn_n_lhs[f];
        rnr[r] = d->rxn_n_rhs[f];
        for (int q = 0; q < 3; ++q) {
            rsp[q * nr + r] = q < rnl[r] ? d->rxn_lhs[3 * f + q] : 0;
            rsp[(3 + q) * nr + r] = q < rnr[r] ? d->rxn_rhs[3 * f + q] : 0;
        }
        rA[r] = d->A[f];
        rbeta[r] = d->beta[f];
        rTa[r] = d->Ta[f];
        rA0[r] = d->A0 ? d->A0[f] : 0.0;
        rbeta0[r] = d->beta0 ? d->beta0[f] : 0.0;
        rTa0[r] = d->Ta0 ? d->Ta0[f] : 0.0;
        if (t == FCS_RXN_THIRD_BODY || t == FCS_RXN_LINDEMANN || t == FCS_RXN_TROE) {
            if (!d->efficiencies) return FCS_E_ARG;
            rtb[r] = nTB++;
            for (int s = 0; s < 64; ++s) eff.push_back(s < ns ? d->efficiencies[(size_t)f * ns + s] : 0.0);
        }
        if (t == FCS_RXN_TROE) {
            if (!d->troe) return FCS_E_ARG;
            raux[r] = (int)troe.size() / 4;
            for (int q = 0; q < 4; ++q) troe.push_back(d->troe[4 * (size_t)f + q]);
        }
        if (t == FCS_RXN_PLOG) {
            const int np = d->plog_n[f];
            if (np < 1 || np > pmax) return FCS_E_ARG;
            raux[r] = (int)plogN.size();
            plogN.push_back(np);
            for (int q = 0; q < pmax; ++q) {
                const int qq = q < np ? q : np - 1;
                const size_t k = (size_t)f * pmax + qq;
                plogLnp.push_back(std::log(d->plog_p[k]));
                plogLnA.push_back(std::log(d->plog_A[k]));
                plogBeta.push_back(d->plog_beta[k]);
                plogTa.push_back(d->plog_Ta[k]);
            }
        }
    }
    std::vector<double> effT((size_t)64 * std::max(nTB, 1), 0.0);
    for (int t = 0; t < nTB; ++t)
        for (int s = 0; s < 64; ++s) effT[(size_t)s * nTB + t] = eff[(size_t)t * 64 + s];
    // ---- per-species lists (net stoichiometry), lane-interleaved
    std::vector<std::vector<std::pair<int, double>>> lst(ns), lstTB(ns);
    for (int r = 0; r < nr; ++r) {
        std::vector<double> nu(ns, 0.0);
        for (int q = 0; q < rnl[r]; ++q) nu[rsp[q * nr + r]] -= 1.0;
        for (int q = 0; q < rnr[r]; ++q) nu[rsp[(3 + q) * nr + r]] += 1.0;
        for (int s = 0; s < ns; ++s) {
            if (nu[s] == 0.0) continue;
            lst[s].push_back({r, nu[s]});
            if (rtb[r] >= 0) lstTB[s].push_back({rtb[r], nu[s]});
        }
    }
    int maxDeg = 1, maxDegTB = 1;
    for (int s = 0; s < ns; ++s) {
        maxDeg = std::max(maxDeg, (int)lst[s].size());
        maxDegTB = std::max(maxDegTB, (int)lstTB[s].size());
    }
    std::vector<int> spR((size_t)maxDeg * 64, 0), spTB((size_t)maxDegTB * 64, 0);
    std::vector<double> spNu((size_t)maxDeg * 64, 0.0), spTBnu((size_t)maxDegTB * 64, 0.0);
    for (int s = 0; s < ns; ++s) {
        for (size_t e = 0; e < lst[s].size(); ++e) {
            spR[e * 64 + s] = lst[s][e].first;
            spNu[e * 64 + s] = lst[s][e].second;
        }
        for (size_t e = 0; e < lstTB[s].size(); ++e) {
            spTB[e * 64 + s] = lstTB[s][e].first;
            spTBnu[e * 64 + s] = lstTB[s][e].second;
        }
    }
    // ---- one device allocation
    struct Piece { const void* src; size_t bytes; size_t off; };
    std::vector<Piece> pieces;
    size_t total = 0;
    auto add = [&](const void* src, size_t bytes) {
        pieces.push_back({src, bytes, total});
        total += align256(bytes > 0 ? bytes : 8);
        return pieces.size() - 1;
    };
    auto vd = [&](const std::vector<double>& v) { return add(v.data(), v.size() * sizeof(double)); };
    auto vi = [&](const std::vector<int>& v) { return add(v.data(), v.size() * sizeof(int)); };
    const size_t iW = vd(W), iInvW = vd(invW), iTc = vd(Tc), iClo = vd(clo), iChi = vd(chi);
    const size_t iType = vi(rtype), iRev = vi(rrev), iNl = vi(rnl), iNr = vi(rnr), iSp = vi(rsp), iTb = vi(rtb),
                 iAux = vi(raux);
    const size_t iA = vd(rA), iB = vd(rbeta), iTa = vd(rTa), iA0 = vd(rA0), iB0 = vd(rbeta0), iTa0 = vd(rTa0);
    const size_t iTroe = vd(troe), iEff = vd(eff), iEffT = vd(effT);
    const size_t iPn = vi(plogN), iPp = vd(plogLnp), iPA = vd(plogLnA), iPb = vd(plogBeta), iPT = vd(plogTa);
    const size_t iSpR = vi(spR), iSpNu = vd(spNu), iSpTB = vi(spTB), iSpTBnu = vd(spTBnu);
    std::vector<unsigned char> host(total, 0);
    for (const auto& p : pieces)
        if (p.bytes) std::memcpy(host.data() + p.off, p.src, p.bytes);
    fcs_mech* m = new (std::nothrow) fcs_mech();
    if (!m) return FCS_E_NOMEM;
    if (hipMalloc(&m->dbuf, total) != hipSuccess) {
        delete m;
        return FCS_E_NOMEM;
    }
    if (hipMemcpy(m->dbuf, host.data(), total, hipMemcpyHostToDevice) != hipSuccess) {
        (void)hipFree(m->dbuf);
        delete m;
        return FCS_E_HIP;
    }
    auto P = [&](size_t i) { return (void*)((char*)m->dbuf + pieces[i].off); };
    DevMech& dm = m->dm;
    dm.ns = ns;
    dm.nr = nr;
    dm.NPm = 64;
    dm.nTB = nTB;
    dm.pmax = pmax;
    dm.maxDeg = maxDeg;
    dm.maxDegTB = maxDegTB;
    dm.TlowMin = TlowMin;
    dm.ThighMax = ThighMax;
    dm.TcMin = TcMin;
    dm.TcMax = TcMax;
    dm.W = (const double*)P(iW);
    dm.invW = (const double*)P(iInvW);
    dm.Tc = (const double*)P(iTc);
    dm.clo = (const double*)P(iClo);
    dm.chi = (const double*)P(iChi);
    dm.rtype = (const int*)P(iType);
    dm.rrev = (const int*)P(iRev);
    dm.rnl = (const int*)P(iNl);
    dm.rnr = (const int*)P(iNr);
    dm.rsp = (const int*)P(iSp);
    dm.rtb = (const int*)P(iTb);
    dm.raux = (const int*)P(iAux);
    dm.rA = (const double*)P(iA);
    dm.rbeta = (const double*)P(iB);
    dm.rTa = (const double*)P(iTa);
    dm.rA0 = (const double*)P(iA0);
    dm.rbeta0 = (const double*)P(iB0);
    dm.rTa0 = (const double*)P(iTa0);
    dm.troe = (const double*)P(iTroe);
    dm.eff = (const double*)P(iEff);
    dm.effT = (const double*)P(iEffT);
    dm.plog_n = (const int*)P(iPn);
    dm.plog_lnp = (const double*)P(iPp);
    dm.plog_lnA = (const double*)P(iPA);
    dm.plog_beta = (const double*)P(iPb);
    dm.plog_Ta = (const double*)P(iPT);
    dm.sp_r = (const int*)P(iSpR);
    dm.sp_nu = (const double*)P(iSpNu);
    dm.sptb_t = (const int*)P(iSpTB);
    dm.sptb_nu = (const double*)P(iSpTBnu);
    m->ns = ns;
    m->nr = nr;
    m->perm = perm;
    *out = m;
    return FCS_OK;
}

int fcs_mech_destroy(fcs_mech* m)
{
    if (!m) return FCS_E_ARG;
    if (m->dbuf) (void)hipFree(m->dbuf);
    delete m;
    return FCS_OK;
}

static int solver_grid(fcs_solver* s)
{
    int dev = 0, cus = 0, per = 0;
    HIP_OK(hipGetDevice(&dev));
    HIP_OK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    HIP_OK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, s->ks->integrate[s->opts.solver], 64,
                                                        s->ldsBytes));
    s->grid = std::max(1, per) * std::max(1, cus);
    return FCS_OK;
}

int fcs_solver_create(const fcs_mech* mech, const fcs_solver_opts* opts, fcs_solver** out)
{
    if (!mech || !opts || !out) return FCS_E_ARG;
    if (opts->solver < 0 || opts->solver > 2 || opts->jacobian < 0 || opts->jacobian > 1 || opts->maxSteps < 1 ||
        !(opts->absTol > 0) || !(opts->relTol > 0))
        return FCS_E_ARG;
    const KernelSet* ks = pick_kernels(mech->ns + 1);
    if (!ks) return FCS_E_UNSUPPORTED;
    fcs_solver* s = new (std::nothrow) fcs_solver();
    if (!s) return FCS_E_NOMEM;
    s->mech = mech;
    s->opts = *opts;
    s->ks = ks;
    s->ldsBytes = lds_doubles(ks->NP, mech->nr, mech->dm.nTB) * sizeof(double);
    if (hipStreamCreate(&s->stream) != hipSuccess) {
        delete s;
        return FCS_E_HIP;
    }
    if (s->ldsBytes > 65536)
        for (const void* f : {ks->integrate[0], ks->integrate[1], ks->integrate[2], ks->deriv, ks->jac, ks->lu})
            (void)hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->ldsBytes);
    if (solver_grid(s) != FCS_OK) {
        (void)hipStreamDestroy(s->stream);
        delete s;
        return FCS_E_HIP;
    }
    *out = s;
    return FCS_OK;
}

static void free_scratch(fcs_solver* s)
{
    for (void* p : {(void*)s->T0, (void*)s->p0, (void*)s->rho0, (void*)s->Y0, (void*)s->dtcIn, (void*)s->dtcOut,
                    (void*)s->RR, (void*)s->Tend, (void*)s->status, (void*)s->nAtt, (void*)s->dtMin,
                    (void*)s->counter})
        if (p) (void)hipFree(p);
    s->T0 = s->p0 = s->rho0 = s->Y0 = s->dtcIn = s->dtcOut = s->RR = s->Tend = nullptr;
    s->status = s->nAtt = nullptr;
    s->dtMin = nullptr;
    s->counter = nullptr;
    s->cap = 0;
}

int fcs_solver_destroy(fcs_solver* s)
{
    if (!s) return FCS_E_ARG;
    free_scratch(s);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return FCS_OK;
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

// Launches the queue kernel on stream st over device arrays.
static int launch_integrate(fcs_solver* s, int nc, double dt, const double* T0, const double* p0,
                            const double* rho0, const double* Y0, const double* dtcIn, double* dtcOut, double* RR,
                            int* status, double* Tend, int* nAtt, unsigned long long* dtMin, unsigned int* counter)
{
    hipLaunchKernelGGL(k_queue_init, dim3(1), dim3(1), 0, s->stream, counter, dtMin);
    HIP_OK(hipGetLastError());
    DevMech M = s->mech->dm;
    DevOpts o = dev_opts(s->opts);
    void* args[] = {&M, &o, &nc, &dt, &T0, &p0, &rho0, &Y0, &dtcIn, &dtcOut, &RR, &status, &Tend, &nAtt, &dtMin,
                    &counter};
    const int grid = std::min(s->grid, std::max(nc, 1));
    HIP_OK(hipLaunchKernel(s->ks->integrate[s->opts.solver], dim3(grid), dim3(64), args, s->ldsBytes, s->stream));
    return FCS_OK;
}


int fcs_solve_ex(fcs_solver* s, int64_t nCells, double deltaT, const double* T0, const double* p0,
                 const double* rho0, const double* const* Y0, double* deltaTChem, double* const* RR,
                 double* deltaTMin, int32_t* cellStatus, double* Tend, int32_t* nAttempts)
{
    if (!s || nCells < 0 || nCells > 0x7fffffff || !(deltaT > 0) || (nCells > 0 && (!T0 || !p0 || !rho0 || !Y0 ||
                                                                                   !deltaTChem || !RR)))
        return FCS_E_ARG;
    const int ns = s->mech->ns;
    const size_t nc = (size_t)nCells;
    if (nc == 0) {
        if (deltaTMin) *deltaTMin = std::min(OF_GREAT, 2 * deltaT);
        return FCS_OK;
    }
    if (nc > s->cap) {
        free_scratch(s);
        int rc = FCS_OK;
        rc |= grow(s->T0, nc) | grow(s->p0, nc) | grow(s->rho0, nc) | grow(s->Y0, nc * ns) | grow(s->dtcIn, nc) |
              grow(s->dtcOut, nc) | grow(s->RR, nc * ns) | grow(s->Tend, nc) | grow(s->status, nc) |
              grow(s->nAtt, nc) | grow(s->dtMin, 1) | grow(s->counter, 1);
        if (rc != FCS_OK) {
            free_scratch(s);
            return FCS_E_NOMEM;
        }
        s->cap = nc;
    }
    HIP_OK(hipMemcpyAsync(s->T0, T0, nc * 8, hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipMemcpyAsync(s->p0, p0, nc * 8, hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipMemcpyAsync(s->rho0, rho0, nc * 8, hipMemcpyHostToDevice, s->stream));
    HIP_OK(hipMemcpyAsync(s->dtcIn, deltaTChem, nc * 8, hipMemcpyHostToDevice, s->stream));
    for (int i = 0; i < ns; ++i)
        HIP_OK(hipMemcpyAsync(s->Y0 + (size_t)i * nc, Y0[i], nc * 8, hipMemcpyHostToDevice, s->stream));
    int rc = launch_integrate(s, (int)nc, deltaT, s->T0, s->p0, s->rho0, s->Y0, s->dtcIn, s->dtcOut, s->RR,
                              s->status, s->Tend, s->nAtt, s->dtMin, s->counter);
    if (rc != FCS_OK) return rc;
    for (int i = 0; i < ns; ++i)
        HIP_OK(hipMemcpyAsync(RR[i], s->RR + (size_t)i * nc, nc * 8, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipMemcpyAsync(deltaTChem, s->dtcOut, nc * 8, hipMemcpyDeviceToHost, s->stream));
    std::vector<int> st(nc);
    HIP_OK(hipMemcpyAsync(st.data(), s->status, nc * 4, hipMemcpyDeviceToHost, s->stream));
    if (Tend) HIP_OK(hipMemcpyAsync(Tend, s->Tend, nc * 8, hipMemcpyDeviceToHost, s->stream));
    if (nAttempts) HIP_OK(hipMemcpyAsync(nAttempts, s->nAtt, nc * 4, hipMemcpyDeviceToHost, s->stream));
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, s->dtMin, 8, hipMemcpyDeviceToHost, s->stream));
    HIP_OK(hipStreamSynchronize(s->stream));
    double dmin;
    std::memcpy(&dmin, &bits, 8);
    dmin = std::min(dmin, OF_GREAT);
    if (deltaTMin) *deltaTMin = std::min(dmin, 2 * deltaT);
    bool bad = false;
    for (size_t c = 0; c < nc; ++c) {
        if (cellStatus) cellStatus[c] = st[c];
        if (st[c] & (CELL_MAXSTEPS | CELL_NONFINITE)) bad = true;
    }
    return bad ? FCS_E_CELLS : FCS_OK;
}

int fcs_solve(fcs_solver* s, int64_t nCells, double deltaT, const double* T0, const double* p0, const double* rho0,
              const double* const* Y0, double* deltaTChem, double* const* RR, double* deltaTMin, int32_t* cellStatus)
{
    return fcs_solve_ex(s, nCells, deltaT, T0, p0, rho0, Y0, deltaTChem, RR, deltaTMin, cellStatus, nullptr,
                        nullptr);
}

// ------------------------------------------------------------ batches
int fcs_batch_create(fcs_solver* s, int64_t nCells, fcs_batch** out)
{
    if (!s || !out || nCells < 1 || nCells > 0x7fffffff) return FCS_E_ARG;
    fcs_batch* b = new (std::nothrow) fcs_batch();
    if (!b) return FCS_E_NOMEM;
    b->s = s;
    b->nc = nCells;
    const size_t nc = (size_t)nCells, ns = (size_t)s->mech->ns;
    int rc = grow(b->T0, nc) | grow(b->p0, nc) | grow(b->rho0, nc) | grow(b->Y0, nc * ns) | grow(b->dtcIn, nc) |
             grow(b->dtcOut, nc) | grow(b->RR, nc * ns) | grow(b->Tend, nc) | grow(b->status, nc) |
             grow(b->nAtt, nc) | grow(b->dtMin, 1) | grow(b->counter, 1);
    if (rc != FCS_OK || hipEventCreate(&b->e0) != hipSuccess || hipEventCreate(&b->e1) != hipSuccess) {
        fcs_batch_destroy(b);
        return FCS_E_NOMEM;
    }
    *out = b;
    return FCS_OK;
}

int fcs_batch_destroy(fcs_batch* b)
{
    if (!b) return FCS_E_ARG;
    for (void* p : {(void*)b->T0, (void*)b->p0, (void*)b->rho0, (void*)b->Y0, (void*)b->dtcIn, (void*)b->dtcOut,
                    (void*)b->RR, (void*)b->Tend, (void*)b->status, (void*)b->nAtt, (void*)b->dtMin,
                    (void*)b->counter})
        if (p) (void)hipFree(p);
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
    HIP_OK(hipMemcpyAsync(b->T0, T0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(b->p0, p0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(b->rho0, rho0, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(b->Y0, Y0, nc * ns * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipMemcpyAsync(b->dtcIn, deltaTChem, nc * 8, hipMemcpyHostToDevice, st));
    HIP_OK(hipStreamSynchronize(st));
    return FCS_OK;
}

int fcs_batch_run(fcs_batch* b, double deltaT, float* elapsed_ms)
{
    if (!b || !(deltaT > 0)) return FCS_E_ARG;
    fcs_solver* s = b->s;
    b->lastDt = deltaT;
    HIP_OK(hipEventRecord(b->e0, s->stream));
    int rc = launch_integrate(s, (int)b->nc, deltaT, b->T0, b->p0, b->rho0, b->Y0, b->dtcIn, b->dtcOut, b->RR,
                              b->status, b->Tend, b->nAtt, b->dtMin, b->counter);
    if (rc != FCS_OK) return rc;
    HIP_OK(hipEventRecord(b->e1, s->stream));
    if (elapsed_ms) {
        HIP_OK(hipEventSynchronize(b->e1));
        HIP_OK(hipEventElapsedTime(elapsed_ms, b->e0, b->e1));
    }
    return FCS_OK;
}

int fcs_batch_download(fcs_batch* b, double* RR, double* deltaTChem, double* deltaTMin, int32_t* cellStatus,
                       double* Tend, int32_t* nAttempts)
{
    if (!b) return FCS_E_ARG;
    const size_t nc = (size_t)b->nc, ns = (size_t)b->s->mech->ns;
    hipStream_t st = b->s->stream;
    if (RR) HIP_OK(hipMemcpyAsync(RR, b->RR, nc * ns * 8, hipMemcpyDeviceToHost, st));
    if (deltaTChem) HIP_OK(hipMemcpyAsync(deltaTChem, b->dtcOut, nc * 8, hipMemcpyDeviceToHost, st));
    if (cellStatus) HIP_OK(hipMemcpyAsync(cellStatus, b->status, nc * 4, hipMemcpyDeviceToHost, st));
    if (Tend) HIP_OK(hipMemcpyAsync(Tend, b->Tend, nc * 8, hipMemcpyDeviceToHost, st));
    if (nAttempts) HIP_OK(hipMemcpyAsync(nAttempts, b->nAtt, nc * 4, hipMemcpyDeviceToHost, st));
    unsigned long long bits = 0;
    HIP_OK(hipMemcpyAsync(&bits, b->dtMin, 8, hipMemcpyDeviceToHost, st));
    HIP_OK(hipStreamSynchronize(st));
    if (deltaTMin) {
        double dmin;
        std::memcpy(&dmin, &bits, 8);
        dmin = std::min(dmin, OF_GREAT);
        *deltaTMin = std::min(dmin, 2 * b->lastDt);
    }
    return FCS_OK;
}

// ------------------------------------------------------------ kernel-level entries

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
        DevMech M = s->mech->dm;
        int ns32 = (int)nS;
        void* args[] = {&M, &ns32, &dp, &dy, &dd};
        HIP_OK(hipLaunchKernel(s->ks->deriv, dim3((unsigned)nS), dim3(64), args, s->ldsBytes, s->stream));
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
    const size_t nv = (size_t)n * nS, nm = (size_t)n * n * nS;
    return with_device((nS + 2 * nv + nm) * 8, [&](char* base) -> int {
        double* dp = (double*)base;
        double* dy = dp + nS;
        double* df = dy + nv;
        double* dJ = df + nv;
        HIP_OK(hipMemcpy(dp, p, nS * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(dy, YT, nv * 8, hipMemcpyHostToDevice));
        DevMech M = s->mech->dm;
        int jt = s->opts.jacobian, ns32 = (int)nS;
        void* args[] = {&M, &jt, &ns32, &dp, &dy, &df, &dJ};
        HIP_OK(hipLaunchKernel(s->ks->jac, dim3((unsigned)nS), dim3(64), args, s->ldsBytes, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(YT, dy, nv * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(dYTdt, df, nv * 8, hipMemcpyDeviceToHost));
        HIP_OK(hipMemcpy(Jy, dJ, nm * 8, hipMemcpyDeviceToHost));
        return FCS_OK;
    });
}

int fcs_lu_solve(fcs_solver* s, int64_t nSys, const double* A, double* b)
{
    if (!s || nSys < 1 || nSys > 0x7fffffff || !A || !b) return FCS_E_ARG;
    const int n = s->mech->ns + 1;
    const size_t nm = (size_t)n * n * nSys, nv = (size_t)n * nSys;
    return with_device((nm + nv) * 8, [&](char* base) -> int {
        double* dA = (double*)base;
        double* db = dA + nm;
        HIP_OK(hipMemcpy(dA, A, nm * 8, hipMemcpyHostToDevice));
        HIP_OK(hipMemcpy(db, b, nv * 8, hipMemcpyHostToDevice));
        DevMech M = s->mech->dm;
        int n32 = (int)nSys;
        void* args[] = {&M, &n32, &dA, &db};
        HIP_OK(hipLaunchKernel(s->ks->lu, dim3((unsigned)nSys), dim3(64), args, s->ldsBytes, s->stream));
        HIP_OK(hipStreamSynchronize(s->stream));
        HIP_OK(hipMemcpy(b, db, nv * 8, hipMemcpyDeviceToHost));
        return FCS_OK;
    });
}

}  // extern "C"
