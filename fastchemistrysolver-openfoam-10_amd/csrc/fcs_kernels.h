// fcs_kernels.h -- the per-cell kernel templates of the integrator (flow-step
// kernels, proposal verification, kernel-level entries, field evaluations)
// and the table of one padded size's kernels.  Every padded size is
// instantiated in a translation unit of its own (fcs_kernels_np.hip built
// with -DFCS_NP=<size>), so the sizes compile in parallel; fcs_lib.hip holds
// the host side and reaches the kernels only through the tables.
#pragma once

#include <hip/hip_runtime.h>

#include "fcs.h"
#include "fcs_integrate.h"

using namespace fcs;

// Padded system sizes compiled into the library (T sits at NP-1).
#define FCS_FOR_EACH_NP(X) X(12) X(24) X(40) X(56) X(64)
// workgroup form (matrix in HBM): padded sizes, multiples of 16
#define FCS_FOR_EACH_NP_LARGE(X) X(80) X(96) X(112) X(128) X(176) X(256)
// LDS of one compute unit: the most one cell's workgroup (or wavefront) can take
constexpr size_t kMaxLdsBytes = 160 * 1024;

// ============================================================== kernels
// Device arrays of one batch of cells: the flow-step inputs and outputs of
// FastChemistryModel::solve (src/ChemistryModel/
// FastChemistryModel_transientSolve.H:3-104) and the integrator state that
// each cell carries between step launches.  All per-cell arrays are SoA over
// cells; Y0/RR/Yw are [rows][nc] (Yw has the temperature as row ns).
struct CellArrays {
    int nc;
    double dt;                       // flow time step deltaT
    const double* dtCell;            // per-cell deltaT of the local-time-step solve, or null
    const double *T0, *p0, *rho0, *Y0, *dtcIn;
    double *dtcOut, *RR, *Tend;
    int *status, *nAtt;
    unsigned long long* dtMinBits;   // min over cells of the new deltaTChem (bits)
    // integrator state
    double *Yw, *x, *dxTry, *dx, *dxTry0;
    int *flags, *nStep;
    // Seulex only: step-start state and scale vector [ns+1][nc], theta_ and
    // kTarg_ [nc], and the per-cell Jacobian [nc][NP*64] kept across steps
    double *Y0s, *Scl, *theta, *Jy;
    int* kTarg;
    double* dxFirst;                 // first accepted sub-step of the last run (0: none yet)
    double* Jg;                      // workgroup form: matrix slices [grid][NP*NP + NP]
};

// Copies the mechanism descriptor from device memory into the last
// kMechLdsDoubles of the cell's LDS and returns the copy.  The kernels take
// the descriptor by pointer and read its fields from LDS where they use them:
// held in SGPRs from the kernel arguments instead (45 pointers and sizes),
// they outgrow the scalar register file, and the compiler spills them to
// VGPR lanes and reloads them inside every loop.
template <int NP>
__device__ __forceinline__ const DevMech& stage_mech(const DevMech* Mg, double* lds_raw)
{
    const int nrq = Mg->nrq, nTB = Mg->nTB, nPlog = Mg->nPlog;
    DevMech* Ms = reinterpret_cast<DevMech*>(lds_raw + lds_doubles(NP, nrq, nTB, nPlog) - kMechLdsDoubles);
    const unsigned int* src = reinterpret_cast<const unsigned int*>(Mg);
    unsigned int* dst = reinterpret_cast<unsigned int*>(Ms);
    for (int w = tid<NP>(); w < (int)(sizeof(DevMech) / 4); w += Cell<NP>::G) dst[w] = src[w];
    __syncthreads();
    return *Ms;
}

// HBM matrix slice of this workgroup (workgroup form with the matrix in
// HBM), else null.
template <int NP>
__device__ __forceinline__ double* jslice(double* Jg)
{
    if constexpr (Cell<NP>::kLarge && !Cell<NP>::kJLds) return Jg + (size_t)blockIdx.x * ((size_t)NP * NP + NP);
    else return nullptr;
}

// Writes the flow-step results of a finished cell (OptRodas34.C:114-117,
// FastChemistryModel_transientSolve.H:59-67): RR = (max(0, y) rho0 - Y0 rho0)
// / deltaT (or (max(0, y) - Y0) rho0 / deltaT_c for the local-time-step
// solve), the new deltaTChem (the input one when the integration failed)
// capped by deltaTChemMax, and the deltaTMin reduction.
template <int NP>
__device__ void finish_cell(const DevMech& M, const DevOpts& o, const Lds& L, const CellArrays& a, int c, int st,
                            double y, double dxTry, int nAtt)
{
    const int lane = tid<NP>();
    const size_t nc = (size_t)a.nc;
    if (gany<NP>(!(y == y) || isinf(y), L)) st |= CELL_NONFINITE;
    if (lane < M.ns) {
        const double Yn = y > 0.0 ? y : 0.0;
        const double r0 = a.rho0[c];
        const double Y0 = a.Y0[lane * nc + c];
        // FastChemistryModel_transientSolve.H:64-67 / FastChemistryModel_localEulerSolve.H:58-61
        a.RR[lane * nc + c] = a.dtCell ? (Yn - Y0) * r0 / a.dtCell[c] : (Yn * r0 - Y0 * r0) / a.dt;
    }
    if (lane == NP - 1) a.Tend[c] = y;
    if (lane == 0) {
        const double dtc = (st & (CELL_MAXSTEPS | CELL_NONFINITE)) ? a.dtcIn[c] : dxTry;
        a.status[c] = st;
        a.nAtt[c] = nAtt;
        a.dtcOut[c] = dtc < o.dtChemMax ? dtc : o.dtChemMax;
        if (dtc > 0.0) atomicMin(a.dtMinBits, (unsigned long long)__double_as_longlong(dtc));
    }
}

// One Rosenbrock attempt for every cell of the active list (one wavefront
// per cell, cells strided over the grid).  Cells still in flight are
// appended to the next list; finished cells write their results.  The
// counter *countZero (the list after next) is cleared for the next launch.
// Every loop here is bounded by the list length read at entry.
template <int NP, int SOLVER>
__global__ __launch_bounds__(Cell<NP>::G, Cell<NP>::kMinWaves) void k_step(const DevMech* Mg, DevOpts o, CellArrays a, const int* listIn,
                                                      const unsigned int* countIn, int* listOut,
                                                      unsigned int* countOut, unsigned int* countZero)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog, jslice<NP>(a.Jg));
    const int lane = tid<NP>();
    const int ns = M.ns;
    const size_t nc = (size_t)a.nc;
    const int row = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    const unsigned int n = *countIn;
    if (blockIdx.x == 0 && lane == 0) *countZero = 0u;
    for (unsigned int i = blockIdx.x; i < n; i += gridDim.x) {
        const int c = __builtin_amdgcn_readfirstlane(listIn[i]);   // wave-uniform cell index
        double y = row >= 0 ? a.Yw[row * nc + c] : 0.0;
        double x = a.x[c], dxTry = a.dxTry[c], dx = a.dx[c], dxTry0 = a.dxTry0[c];
        int flags = a.flags[c], nStep = a.nStep[c], nAtt = a.nAtt[c];
        const double xEnd = a.dtCell ? a.dtCell[c] : a.dt;
        const int nStep0 = nStep;
        double dxAcc = 0.0;
        const double p = a.p0[c];
        stage_plog<NP>(M, L, log(p));
        const int st = rosenbrock_step<NP>(M, L, o, tableau_of<SOLVER>(), p, xEnd, y,
                                           x, dxTry, dx, dxTry0, flags, nStep, nAtt, dxAcc);
        if (lane == 0 && nStep0 == 0 && dxAcc > 0.0 && a.dxFirst[c] == 0.0) a.dxFirst[c] = dxAcc;
        if (st == STEP_CONTINUE) {
            if (row >= 0) a.Yw[row * nc + c] = y;
            if (lane == 0) {
                a.x[c] = x;
                a.dxTry[c] = dxTry;
                a.dx[c] = dx;
                a.dxTry0[c] = dxTry0;
                a.flags[c] = flags;
                a.nStep[c] = nStep;
                a.nAtt[c] = nAtt;
                listOut[atomicAdd(countOut, 1u)] = c;
            }
        } else {
            finish_cell<NP>(M, o, L, a, c, st, y, dxTry, nAtt);
        }
    }
}

// One pass of the Seulex extrapolation loop for every cell of the active
// list (same list protocol as k_step).  The per-cell Jacobian lives in
// a.Jy between launches.
template <int NP>
__global__ __launch_bounds__(Cell<NP>::G, Cell<NP>::kMinWaves) void k_step_seulex(const DevMech* Mg, DevOpts o, CellArrays a,
                                                             const int* listIn, const unsigned int* countIn,
                                                             int* listOut, unsigned int* countOut,
                                                             unsigned int* countZero)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog, jslice<NP>(a.Jg));
    const int lane = tid<NP>();
    const int ns = M.ns;
    const size_t nc = (size_t)a.nc;
    const int row = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    const unsigned int n = *countIn;
    if (blockIdx.x == 0 && lane == 0) *countZero = 0u;
    for (unsigned int i = blockIdx.x; i < n; i += gridDim.x) {
        const int c = __builtin_amdgcn_readfirstlane(listIn[i]);   // wave-uniform cell index
        double y = row >= 0 ? a.Yw[row * nc + c] : 0.0;
        double y0s = row >= 0 ? a.Y0s[row * nc + c] : 0.0;
        double scale = row >= 0 ? a.Scl[row * nc + c] : o.absTol;
        SeulexCell sc;
        sc.x = a.x[c];
        sc.dxTry = a.dxTry[c];
        sc.dxNew = a.dx[c];
        sc.dxTry0 = a.dxTry0[c];
        sc.theta = a.theta[c];
        sc.kTarg = a.kTarg[c];
        sc.flags = a.flags[c];
        sc.nStep = a.nStep[c];
        sc.nAtt = a.nAtt[c];
        double* JyG = a.Jy + (size_t)c * NP * jy_stride<NP>();
        stage_plog<NP>(M, L, log(a.p0[c]));
        const int st = seulex_pass<NP>(M, L, o, a.p0[c], a.dtCell ? a.dtCell[c] : a.dt, JyG, y, y0s, scale, sc);
        if (st == STEP_CONTINUE) {
            if (row >= 0) {
                a.Yw[row * nc + c] = y;
                a.Y0s[row * nc + c] = y0s;
                a.Scl[row * nc + c] = scale;
            }
            if (lane == 0) {
                a.x[c] = sc.x;
                a.dxTry[c] = sc.dxTry;
                a.dx[c] = sc.dxNew;
                a.dxTry0[c] = sc.dxTry0;
                a.theta[c] = sc.theta;
                a.kTarg[c] = sc.kTarg;
                a.flags[c] = sc.flags;
                a.nStep[c] = sc.nStep;
                a.nAtt[c] = sc.nAtt;
                listOut[atomicAdd(countOut, 1u)] = c;
            }
        } else {
            finish_cell<NP>(M, o, L, a, c, st, y, sc.dxTry, sc.nAtt);
        }
    }
}

// Verification of the network's proposals, one wavefront per gated cell
// (listG: positions in the reacting-cell list `list`, whose proposals have
// h >= deltaT/2; k_gate sends every other cell on).  Row i of P holds
// [ln h, z_0 .. z_ns-1, z_T]: h is the proposed first sub-step of an
// integration that starts by trying the whole flow step, and the increments
// are tolerance-scaled and compressed, d_k = sinh(z_k) (absTol + relTol
// |y0_k|).  The proposed end state y0 + d is accepted when one FP64 Rodas34
// step over the whole flow step from y0 passes its own error test
// (norm <= 1) and agrees with the proposal to the same tolerance (max-norm of
// |yT - yP| / (absTol + relTol max(|yT|, |yP|)) <= 1); the cell then finishes
// with the Rodas34 result and the step proposal of an accepted step.  A
// rejected cell joins the output list unchanged, so the integrator then runs
// on it exactly as without the network -- or, with seed != 0, starting from
// min(h, deltaT) instead of deltaTChem.  The attempt sits in straight-line
// code of the cell loop: under a conditional branch the compiled attempt
// has produced wrong error norms on this device (the same code without the
// branch, and in every other kernel, is exact).
template <int NP>
__global__ __launch_bounds__(Cell<NP>::G) void k_verify(const DevMech* Mg, DevOpts o, CellArrays a, const int* list,
                                                        const int* listG, const unsigned int* countG, const float* P,
                                                        int seed, int* listOut, unsigned int* countOut,
                                                        unsigned long long* nAccepted)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog, jslice<NP>(a.Jg));
    const int lane = tid<NP>();
    const int ns = M.ns;
    const size_t nc = (size_t)a.nc;
    const int row = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    const unsigned int n = *countG;
    for (unsigned int g = blockIdx.x; g < n; g += gridDim.x) {
        const int i = __builtin_amdgcn_readfirstlane(listG[g]);   // wave-uniform
        const int c = __builtin_amdgcn_readfirstlane(list[i]);
        const float* pr = P + (size_t)i * (ns + 2);
        const double dt = a.dtCell ? a.dtCell[c] : a.dt;
        double y0 = row >= 0 ? a.Yw[row * nc + c] : 0.0;
        const double h = exp((double)pr[0]);
        const double yP = row >= 0 ? y0 + sinh((double)pr[1 + row]) * (o.absTol + o.relTol * fabs(y0)) : 0.0;
        double yT = 0.0;
        const double p = a.p0[c];
        stage_plog<NP>(M, L, log(p));
        const double err = rosenbrock_attempt<NP>(M, L, o, kRodas34, p, log(p), dt, y0, yT);
        double r = 0.0;
        if (row >= 0) {
            const double m = fabs(yT) > fabs(yP) ? fabs(yT) : fabs(yP);
            r = fabs(yT - yP) / (o.absTol + o.relTol * m);
        }
        const bool bad = gany<NP>(!(r == r), L);
        const double agree = gmax<NP>(r, L);
        const bool accept = (err <= 1.0) && !bad && (agree <= 1.0);
        if (accept) {
            finish_cell<NP>(M, o, L, a, c, FCS_CELL_LEARNED, yT, accepted_dxTry(err, dt), 1);
            if (lane == 0) atomicAdd(nAccepted, 1ull);
        } else if (lane == 0) {
            if (seed) a.dxTry[c] = h < dt ? h : dt;   // h >= deltaT/2 > 0 here
            listOut[atomicAdd(countOut, 1u)] = c;
        }
    }
}

// One wavefront per state: derivatives.  External layout [n][nS] with T at
// row ns.
template <int NP>
__global__ __launch_bounds__(Cell<NP>::G) void k_derivatives(const DevMech* Mg, int nS, const double* p, double* YT,
                                                             double* out)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog);
    const int s = blockIdx.x, lane = tid<NP>(), ns = M.ns;
    const int ext = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    double y = ext >= 0 ? YT[(size_t)ext * nS + s] : 0.0;
    stage_plog<NP>(M, L, log(p[s]));
    const double d = derivatives<NP>(M, L, p[s], log(p[s]), y);
    if (ext >= 0) {
        YT[(size_t)ext * nS + s] = y;
        out[(size_t)ext * nS + s] = d;
    }
}

template <int NP>
__global__ __launch_bounds__(Cell<NP>::G) void k_jacobian(const DevMech* Mg, int jacType, int nS, const double* p,
                                                          double* YT, double* f, double* J, double* Jg)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog, jslice<NP>(Jg));
    const int s = blockIdx.x, lane = tid<NP>(), ns = M.ns, n = ns + 1;
    const int ext = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    double y = ext >= 0 ? YT[(size_t)ext * nS + s] : 0.0;
    stage_plog<NP>(M, L, log(p[s]));
    const double d = jacobian<NP>(M, L, p[s], log(p[s]), jacType, y);
    if (ext >= 0) {
        YT[(size_t)ext * nS + s] = y;
        f[(size_t)ext * nS + s] = d;
        double* row = J + ((size_t)s * n + ext) * n;
        for (int j = 0; j < NP; ++j) {
            const int ej = j < ns ? j : (j == NP - 1 ? ns : -1);
            if (ej >= 0) row[ej] = L.J[j * Cell<NP>::JSt + lane];
        }
    }
}

// One Rosenbrock attempt per state (OptRodas34::Rodas34Solve,
// src/ODE/OptRodas34/OptRodas34.C:291-879, or
// OptRosenbrock34::Rosenbrock34Solve, src/ODE/OptRosenbrock34/
// OptRosenbrock34.C:281-677).  External layout [n][nS], T at row ns; the
// start state is clipped in place as the Jacobian evaluation does.
template <int NP, int SOLVER>
__global__ __launch_bounds__(Cell<NP>::G) void k_attempt(const DevMech* Mg, DevOpts o, int nS, const double* p,
                                                         const double* dx, double* YT, double* YTnew, double* err,
                                                         double* Jg)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog, jslice<NP>(Jg));
    const int s = blockIdx.x, lane = tid<NP>(), ns = M.ns;
    const int ext = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    double y = ext >= 0 ? YT[(size_t)ext * nS + s] : 0.0;
    double yT = 0.0;
    stage_plog<NP>(M, L, log(p[s]));
    const double e = rosenbrock_attempt<NP>(M, L, o, tableau_of<SOLVER>(), p[s], log(p[s]),
                                            dx[s], y, yT);
    if (ext >= 0) {
        YT[(size_t)ext * nS + s] = y;
        YTnew[(size_t)ext * nS + s] = yT;
    }
    if (lane == 0) err[s] = e;
}

template <int NP>
__global__ __launch_bounds__(Cell<NP>::G) void k_lu_solve(const DevMech* Mg, int nSys, const double* A, double* b,
                                                          double* Jg)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog, jslice<NP>(Jg));
    const int s = blockIdx.x, lane = tid<NP>(), ns = M.ns, n = ns + 1;
    const int ext = lane < ns ? lane : (lane == NP - 1 ? ns : -1);
    if (!Cell<NP>::kLarge || lane < NP)
        for (int j = 0; j < NP; ++j) {
            const int ej = j < ns ? j : (j == NP - 1 ? ns : -1);
            L.J[j * Cell<NP>::JSt + lane] =
                (ext >= 0 && ej >= 0) ? A[((size_t)s * n + ext) * n + ej] : ((lane == j && lane < NP) ? 1.0 : 0.0);
        }
    wave_sync();
    const auto st = lu_factor_any<NP>(M, L);
    const double x = lu_solve_any<NP>(L, st, ext >= 0 ? b[(size_t)s * n + ext] : 0.0);
    if (ext >= 0) b[(size_t)s * n + ext] = x;
}

// Reaction-rate state of one cell from (rho, T, p, Y) as the reference's
// calculate() and tc() form it: c_i = rho Y_i / W_i without clipping,
// thermo and rate constants at T clamped to the fit range.  Leaves C, expG
// and Mtb in LDS; returns this lane's concentration (0 beyond ns).
template <int NP>
__device__ inline double cell_rate_state(const DevMech& M, const Lds& L, int nc, int c, const double* rho,
                                         const double* T, const double* Y, double& Tcl, double& lnT, double& iT)
{
    const int lane = tid<NP>();
    const bool sp = lane < M.ns;
    const SpeciesData S = load_species<NP>(M, lane);
    const double Cl = sp ? rho[c] * Y[(size_t)lane * nc + c] / S.W : 0.0;
    Tcl = clampT(M, T[c]);
    lnT = log(Tcl);
    iT = 1.0 / Tcl;
    Thermo th{0, 0, 0, 0, 0};
    if (sp) th = species_thermo(M, S, Tcl, lnT, iT);
    L.C[lane] = Cl;
    const double eg = sp ? exp(th.negG) : 0.0;
    L.expG[lane] = eg;
    if constexpr (!Cell<NP>::kLarge) {
        L.dBdT[lane] = sp ? 1.0 / eg : 0.0;   // exp(+G/RT) for the division-free 1/Kc
        if (lane == 0) set_identity_slot(L, 1.0);
    }
    wave_sync();
    third_body_sums<NP>(M, L);
    wave_sync();
    return Cl;
}

// FastChemistryModel::calculate (src/ChemistryModel/FastChemistryModel.C:
// 700-740): RR_i = W_i dNdtByV_i(p, T, c), one wavefront per cell, fields
// [ns][nc].
template <int NP>
__global__ __launch_bounds__(Cell<NP>::G) void k_calculate(const DevMech* Mg, int nc, const double* rho, const double* T,
                                                           const double* p, const double* Y, double* RR)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog);
    const int c = blockIdx.x, lane = tid<NP>();
    double Tcl, lnT, iT;
    stage_plog<NP>(M, L, log(p[c]));
    (void)cell_rate_state<NP>(M, L, nc, c, rho, T, Y, Tcl, lnT, iT);
    reaction_round<false, NP>(M, L, Tcl, lnT, iT, log(p[c]));
    wave_sync();
    if (lane < M.ns) RR[(size_t)lane * nc + c] = production_rate<NP>(M, L) * M.W[lane];
}

// FastChemistryModel::tc (src/ChemistryModel/FastChemistryModel.C:588-643)
// with OptReaction::Tc (src/Reaction/Tc.C:4-834): tc = sumW /
// sumWRateByCTot * sum(c), or vGreat when no reaction proceeds.
template <int NP>
__global__ __launch_bounds__(Cell<NP>::G) void k_tc(const DevMech* Mg, int nc, const double* rho, const double* T,
                                                    const double* p, const double* Y, double* tc)
{
    extern __shared__ double lds_raw[];
    const DevMech& M = stage_mech<NP>(Mg, lds_raw);
    const Lds L = make_lds(lds_raw, NP, M.nrq, M.nTB, M.nPlog);
    const int c = blockIdx.x;
    double Tcl, lnT, iT;
    stage_plog<NP>(M, L, log(p[c]));
    const double Cl = cell_rate_state<NP>(M, L, nc, c, rho, T, Y, Tcl, lnT, iT);
    const double sumc = gsum<NP>(Cl, L);
    const double2 sw = time_scale_sums<NP>(M, L, Tcl, lnT, iT, log(p[c]));
    if (tid<NP>() == 0) tc[c] = sw.y == 0.0 ? OF_VGREAT : sw.x / sw.y * sumc;
}

// The kernels of one padded size, as the host launches them.
struct KernelSet {
    int NP;
    int G;                      // threads per cell
    const void* step[5];        // indexed by FCS_SOLVER_*
    const void* verify;
    const void* deriv;
    const void* jac;
    const void* lu;
    const void* attempt[5];     // indexed by FCS_SOLVER_* (none for Seulex)
    const void* calculate;
    const void* tc;
};

// Table entry of padded size N (one translation unit per size defines it).
#define FCS_KSET_INIT(N)                                                                                   \
    {N, group_of(N),                                                                                       \
     {(const void*)k_step<N, SOLVER_RODAS34>, (const void*)k_step<N, SOLVER_ROS34>,                        \
      (const void*)k_step_seulex<N>, (const void*)k_step<N, SOLVER_RODAS23>,                               \
      (const void*)k_step<N, SOLVER_ROS23>},                                                               \
     (const void*)k_verify<N>,                                                                             \
     (const void*)k_derivatives<N>, (const void*)k_jacobian<N>, (const void*)k_lu_solve<N>,                \
     {(const void*)k_attempt<N, SOLVER_RODAS34>, (const void*)k_attempt<N, SOLVER_ROS34>, nullptr,         \
      (const void*)k_attempt<N, SOLVER_RODAS23>, (const void*)k_attempt<N, SOLVER_ROS23>},                 \
     (const void*)k_calculate<N>, (const void*)k_tc<N>}
