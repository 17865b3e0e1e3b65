// fcs_device.h -- gfx950 device code of the per-cell stiff chemistry
// integrator (one cell per 64-lane wavefront).
//
// Layout of one cell inside a wavefront (NP = padded system size, a multiple
// of 4, NP <= 64):
//   lane s < ns        species s: state y[s] = Y_s, matrix row s in a[0..NP-1]
//   ns <= lane < NP-1  padding: zero state, identity-like rows
//   lane NP-1          temperature: y = T, matrix row of dT/dt
// Matrix columns follow the same order, so every register index is a
// compile-time constant; the temperature column is a[NP-1].  Placing T at
// NP-1 instead of ns leaves the pivot decisions of the reference's block-4
// LU unchanged: padding rows hold zeros in every real column, so they never
// win a pivot search (src/LUsolver/LUsolver.C:106-246,503-650).
//
// Reaction-level work (rate constants, equilibrium constants) is spread over
// lanes (lane l handles reactions l, l+64, ...); species-level gathers use a
// per-species reaction list, so no floating-point atomics are used and every
// result is bit-reproducible for a fixed launch shape and for any other.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fcs {

constexpr double RU = 8314.470066505449722171761095523834228515625;  // OptReaction.H:404
constexpr double PSTD = 1.0e5;                                         // OptReaction.H:409
constexpr double KC_FLOOR = 1.4901171103413047e-8;                     // ReactionFunctionI.H:17
constexpr double TROE_SMALL = 2.2e-16;                                 // ReactionRate.C:475
constexpr double PIVOT_EPS = 2.2204460492503131e-16;                   // LUsolver.C:155
constexpr double OF_SMALL = 1.0e-15;    // OpenFOAM-10 doubleScalarSmall
constexpr double OF_GREAT = 1.0e15;     // OpenFOAM-10 doubleScalarGreat
constexpr double OF_VGREAT = 1.0e300;   // OpenFOAM-10 doubleScalarVGreat
constexpr int WAVE = 64;

enum { RXN_ARR = 0, RXN_TB = 1, RXN_LIND = 2, RXN_TROE = 3, RXN_PLOG = 4 };
enum { SOLVER_RODAS34 = 0, SOLVER_ROS34 = 1, SOLVER_SEULEX = 2 };
enum { CELL_SKIPPED = 1, CELL_MAXSTEPS = 2, CELL_NONFINITE = 4 };

// Device-resident mechanism tables (built on the host by fcs_mech_create).
// Reactions are sorted by rate form so that lanes of one reaction round
// mostly take the same branch.
struct DevMech {
    int ns, nr, NPm;          // species, reactions, padded size used for tables
    int nTB;                  // third-body + falloff rows
    int pmax;                 // PLOG points per row
    int maxDeg, maxDegTB;     // per-species list lengths (padded)
    double TlowMin, ThighMax, TcMin, TcMax;
    const double* W;          // [64]  (0 beyond ns)
    const double* invW;       // [64]  (0 beyond ns)
    const double* Tc;         // [64]
    const double* clo;        // [7*64] coefficient-major: c[k*64 + s]
    const double* chi;        // [7*64]
    const int* rtype;         // [nr]
    const int* rrev;          // [nr]
    const int* rnl;           // [nr]
    const int* rnr;           // [nr]
    const int* rsp;           // [6*nr] slot-major: rsp[q*nr + r], q<3 lhs, q>=3 rhs
    const int* rtb;           // [nr]  eff row or -1
    const int* raux;          // [nr]  troe / plog row
    const double* rA;         // [nr]  forward (k_inf)
    const double* rbeta;
    const double* rTa;
    const double* rA0;        // [nr]  k_0 (falloff)
    const double* rbeta0;
    const double* rTa0;
    const double* troe;       // [4*nTroe] (alpha, Tsss, Ts, Tss) per row
    const double* eff;        // [nTB*64] row-major, zero beyond ns
    const double* effT;       // [64*nTB] species-major for M sums
    const int* plog_n;        // [nPlog]
    const double* plog_lnp;   // [nPlog*pmax]
    const double* plog_lnA;
    const double* plog_beta;
    const double* plog_Ta;
    const int* sp_r;          // [maxDeg*64] lane-interleaved (reaction)
    const double* sp_nu;      // [maxDeg*64] net stoichiometric coefficient (0 = pad)
    const int* sptb_t;        // [maxDegTB*64] eff row of third-body/falloff entries
    const double* sptb_nu;    // [maxDegTB*64]
    const unsigned long long* rpack;    // [nr] species of slots 0-5 in bytes 0-5, nl, nr in bytes 6, 7
    const unsigned long long* sp_pack;  // [maxDeg*64] rpack of each per-species list entry
};

__device__ __forceinline__ int pk_species(unsigned long long w, int q) { return (int)((w >> (8 * q)) & 255u); }
__device__ __forceinline__ int pk_nl(unsigned long long w) { return (int)((w >> 48) & 255u); }
__device__ __forceinline__ int pk_nr(unsigned long long w) { return (int)(w >> 56); }

struct DevOpts {
    int solver, jac, maxSteps;
    double absTol, relTol, Treact, dtChemMax;
};

// ---------------------------------------------------------------- lane utils
__device__ __forceinline__ int lane_id() { return __lane_id(); }

__device__ __forceinline__ double readlane_d(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

__device__ __forceinline__ int readlane_i(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

// Butterfly sum: every lane ends with the same bits (each step adds a
// commutative pair).
__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, WAVE);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, WAVE));
    return v;
}
__device__ __forceinline__ int first_lane(bool pred)
{
    const unsigned long long b = __ballot(pred);
    return b ? (__ffsll((long long)b) - 1) : -1;
}
// Orders LDS traffic between the lanes of the (single-wave) workgroup.
__device__ __forceinline__ void wave_sync() { __syncthreads(); }

// ---------------------------------------------------------------- LDS layout
// The per-cell matrix lives in LDS, lane-interleaved column-major: element
// (row i, column j) at J[j*64 + i], row i owned by lane i.  Column accesses
// of one row index across the lanes are conflict-free; the pivot row of the
// LU is read by all lanes at one address (broadcast).
struct Lds {
    double* C;     // [64] concentrations
    double* expG;  // [64] exp(-G/RT)
    double* dBdT;  // [64] d(-G/RT)/dT
    double* H;     // [64] species enthalpy (T-row column sums)
    double* Q;     // [64] dY/dt * cp (T-row column sums)
    double* rq;    // [nrP] q (rates) or kf (Jacobian)
    double* rk;    // [nrP] kr (Jacobian)
    double* rdT;   // [nrP] dq/dT (Jacobian)
    double* rw;    // [nTBP] third-body partial (Jacobian)
    double* J;     // [NP*64] Jacobian, then iteration matrix and its LU factors
    int* ip;       // [64] lane holding each logical row after the LU
};

__host__ __device__ inline int round_up(int x, int m) { return (x + m - 1) / m * m; }

__host__ __device__ inline size_t lds_doubles(int NP, int nr, int nTB)
{
    return 5 * 64 + 3 * (size_t)round_up(nr, 2) + (size_t)round_up(nTB > 0 ? nTB : 1, 2) + (size_t)NP * 64 + 32;
}

__device__ inline Lds make_lds(double* base, int NP, int nr, int nTB)
{
    Lds L;
    const int nrP = round_up(nr, 2), nTBP = round_up(nTB > 0 ? nTB : 1, 2);
    L.C = base;
    L.expG = base + 64;
    L.dBdT = base + 128;
    L.H = base + 192;
    L.Q = base + 256;
    L.rq = base + 320;
    L.rk = L.rq + nrP;
    L.rdT = L.rk + nrP;
    L.rw = L.rdT + nrP;
    L.J = L.rw + nTBP;
    L.ip = (int*)(L.J + (size_t)NP * 64);
    return L;
}

// ---------------------------------------------------------------- thermo
// Mass-based NASA-7 thermo (src/Reaction/ThermoFunctionI.H:33-123); the
// coefficient choice follows setPtrCoeffs (ThermoFunctionI.H:1-31).
struct Thermo { double cp, ha, negG, dBdT, dcpdT; };

__device__ inline Thermo species_thermo(const DevMech& M, int s, double T, double lnT, double iT)
{
    const bool low = (T < M.TcMin) ? true : (T > M.TcMax ? false : (T < M.Tc[s]));
    const double* c = low ? M.clo : M.chi;
    const double a0 = c[0 * 64 + s], a1 = c[1 * 64 + s], a2 = c[2 * 64 + s], a3 = c[3 * 64 + s],
                 a4 = c[4 * 64 + s], a5 = c[5 * 64 + s], a6 = c[6 * 64 + s];
    const double r = RU * M.invW[s];
    Thermo t;
    t.cp = ((((a4 * T + a3) * T + a2) * T + a1) * T + a0) * r;
    t.ha = (((((a4 * 0.2 * T + a3 * 0.25) * T + a2 * (1.0 / 3.0)) * T + a1 * 0.5) * T + a0) * T + a5) * r;
    t.negG = (((a4 * 0.05 * T + a3 * (1.0 / 12.0)) * T + a2 * (1.0 / 6.0)) * T + a1 * 0.5) * T
             - a5 * iT + a6 + (lnT - 1.0) * a0;
    t.dBdT = ((a4 * 0.2 * T + a3 * 0.25) * T + a2 * (1.0 / 3.0)) * T + a1 * 0.5 + (a5 * iT + a0) * iT;
    t.dcpdT = (((4.0 * a4 * T + 3.0 * a3) * T + 2.0 * a2) * T + a1) * r;
    return t;
}

__device__ __forceinline__ double clampT(const DevMech& M, double T)
{
    T = T < M.TlowMin ? M.TlowMin : T;   // src/Reaction/ReactionRate.C:14-15
    return T > M.ThighMax ? M.ThighMax : T;
}

// ---------------------------------------------------------------- rates
struct RxnK { double kf, dkfdT, dkfdM; };

// Rate constant of reaction r (kf includes M for third-body and falloff
// forms).  JAC selects the Jacobian-path variants, which also return
// dkf/dT and dkf/dM (src/Reaction/ReactionRate.C:84-672,
// src/Reaction/Jacobian.C:108-910).
template <bool JAC>
__device__ inline RxnK rate_constant(const DevMech& M, const Lds& L, int r, int t, double T, double lnT,
                                     double iT, double lnp)
{
    RxnK k{0.0, 0.0, 0.0};
    if (t == RXN_PLOG) {
        const int row = M.raux[r];
        const int np = M.plog_n[row];
        const double* lp = M.plog_lnp + (size_t)row * M.pmax;
        const double* la = M.plog_lnA + (size_t)row * M.pmax;
        const double* b = M.plog_beta + (size_t)row * M.pmax;
        const double* ta = M.plog_Ta + (size_t)row * M.pmax;
        double lk, dlk;
        if (lnp <= lp[0]) {
            lk = la[0] + b[0] * lnT - ta[0] * iT;
            dlk = (b[0] + ta[0] * iT) * iT;
        } else if (lnp >= lp[np - 1]) {
            lk = la[np - 1] + b[np - 1] * lnT - ta[np - 1] * iT;
            dlk = (b[np - 1] + ta[np - 1] * iT) * iT;
        } else {
            int j = 0;
            for (int q = 0; q < np - 1; ++q)
                if (lp[q] <= lnp && lnp <= lp[q + 1]) { j = q; break; }
            const double f = (lnp - lp[j]) / (lp[j + 1] - lp[j]);
            lk = la[j] + b[j] * lnT - ta[j] * iT
               + (la[j + 1] - la[j] + (b[j + 1] - b[j]) * lnT - (ta[j + 1] - ta[j]) * iT) * f;
            dlk = iT * (b[j] + ta[j] * iT + (b[j + 1] - b[j] + (ta[j + 1] - ta[j]) * iT) * f);
        }
        k.kf = exp(lk);
        if (JAC) k.dkfdT = k.kf * dlk;
        return k;
    }
    const double kk = M.rA[r] * exp(M.rbeta[r] * lnT - M.rTa[r] * iT);
    const double dk = JAC ? kk * (M.rbeta[r] + M.rTa[r] * iT) * iT : 0.0;
    if (t == RXN_ARR) {
        k.kf = kk;
        k.dkfdT = dk;
        return k;
    }
    // third-body concentration M = sum_j eff_rj C_j
    const int tb = M.rtb[r];
    double Mc = 0.0;
    for (int j0 = 0; j0 < M.ns; j0 += 8) {   // batches of 8 independent loads
        double e[8], c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u < M.ns) {
                e[u] = M.effT[(size_t)(j0 + u) * M.nTB + tb];
                c[u] = L.C[j0 + u];
            }
#pragma unroll
        for (int u = 0; u < 8; ++u)
            if (j0 + u < M.ns) Mc += e[u] * c[u];
    }
    if (t == RXN_TB) {
        k.kf = kk * Mc;
        k.dkfdT = dk * Mc;
        k.dkfdM = kk;
        return k;
    }
    const double kinf = kk, dkinf = dk;
    const double k0 = M.rA0[r] * exp(M.rbeta0[r] * lnT - M.rTa0[r] * iT);
    const double dk0 = JAC ? k0 * (M.rbeta0[r] + M.rTa0[r] * iT) * iT : 0.0;
    if (t == RXN_LIND) {
        const double Pr = k0 * Mc / kinf;
        const double tmp = 1.0 / (1.0 + Pr);
        if (JAC) {
            const double dPrdT = (Mc * dk0 - Pr * dkinf) / kinf;
            k.dkfdT = tmp * Pr * dkinf + tmp * tmp * dPrdT * kinf;
            k.dkfdM = k0 * tmp * tmp;
        }
        k.kf = Mc * tmp * k0;
        return k;
    }
    // Troe
    const double* tr = M.troe + 4 * (size_t)M.raux[r];
    const double alpha = tr[0], Tsss = tr[1], Ts = tr[2], Tss = tr[3];
    const double e1 = exp(-T / Tsss), e2 = exp(-Tss * iT), e3 = exp(-T / Ts);
    const double Fcent = (1.0 - alpha) * e1 + alpha * e3 + e2;
    const double logFc = log10(Fcent > TROE_SMALL ? Fcent : TROE_SMALL);
    const double c = -0.4 - 0.67 * logFc;
    const double nn = 0.75 - 1.27 * logFc;
    if (!JAC) {
        double Pr = k0 * Mc / kinf;
        Pr = Pr > TROE_SMALL ? Pr : TROE_SMALL;
        const double logPr = log10(Pr);
        const double x1 = nn - 0.14 * (logPr + c);
        const double x2 = (logPr + c) / x1;
        const double F = exp10(logFc / (1.0 + x2 * x2));
        k.kf = Mc * (k0 * F / (1.0 + Pr));
        return k;
    }
    const double Pr0 = Mc * k0 / kinf;
    const bool prOk = Pr0 >= TROE_SMALL;
    const double Pr = Pr0 + 1e-100;
    const double logPr = log10(Pr > TROE_SMALL ? Pr : TROE_SMALL);
    const double x1 = nn - 0.14 * (logPr + c);
    const double x2 = (logPr + c) / x1;
    const double x3 = 1.0 + x2 * x2;
    const double x4 = logFc / x3;
    const double F = exp10(x4);
    const double ln10 = 2.302585092994045684;
    const double dFcdT = (alpha - 1.0) / Tsss * e1 - alpha / Ts * e3 + e2 * Tss * iT * iT;
    const double dlogFcdT = (Fcent >= TROE_SMALL) ? dFcdT / Fcent / ln10 : 0.0;
    const double dcdT = -0.67 * dlogFcdT, dndT = -1.27 * dlogFcdT;
    const double dx1dT = dndT - 0.14 * dcdT;
    const double dx2dT = (dcdT - x2 * dx1dT) / x1;
    const double dx3dT = 2.0 * x2 * dx2dT;
    const double dx4dT = (dlogFcdT - x4 * dx3dT) / x3;
    double dFdT = ln10 * F * dx4dT;
    const double dlogPrdPr = prOk ? 1.0 / (Pr * ln10) : 0.0;
    const double dx1dPr = -0.14 * dlogPrdPr;
    const double dx2dPr = (dlogPrdPr - x2 * dx1dPr) / x1;
    const double dx3dPr = 2.0 * x2 * dx2dPr;
    const double dx4dPr = -x4 * dx3dPr / x3;
    const double dFdPr = ln10 * F * dx4dPr;
    const double dPrdT = (Mc * dk0 - Pr * dkinf) / kinf;
    dFdT += dFdPr * dPrdT;
    const double i1 = 1.0 / (1.0 + Pr);
    k.dkfdT = F * i1 * (Pr * dkinf) + kinf * F * i1 * i1 * dPrdT + k0 * i1 * dFdT * Mc;
    k.dkfdM = k0 * i1 * (F * i1 + Pr * dFdPr);
    k.kf = (F * k0 * i1) * Mc;
    return k;
}

// Kc in concentration units (src/Reaction/ReactionFunctionI.H:1-52,
// ChemicalEquilibriumFunctionI.H:14-25) and, with JAC, d ln Kc / dT
// (src/Reaction/JacobianFunctionI.H:1150-1158).
template <bool JAC>
__device__ inline double equilibrium(const Lds& L, const int* s, int nl, int nrh, double PbyRT, double iT,
                                     double* dlnKc)
{
    double num = 1.0, den = 1.0, sdb = 0.0;
    for (int a = 0; a < nrh; ++a) { num *= L.expG[s[3 + a]]; if (JAC) sdb += L.dBdT[s[3 + a]]; }
    for (int a = 0; a < nl; ++a) { den *= L.expG[s[a]]; if (JAC) sdb -= L.dBdT[s[a]]; }
    const int dn = nrh - nl;
    double pw = 1.0;
    if (dn > 0) for (int q = 0; q < dn; ++q) pw *= PbyRT;
    if (dn < 0) for (int q = 0; q < -dn; ++q) pw /= PbyRT;
    if (JAC) *dlnKc = sdb - dn * iT;
    return (num / den) * pw;
}

// Reaction round: lane l handles reactions l, l+64, ...
//  !JAC: rq[r] = q_r
//   JAC: rq[r] = kf, rk[r] = kr, rdT[r] = dq/dT, rw[tb] = dkf/dM CF - dkr/dM CR
template <bool JAC>
__device__ inline void reaction_round(const DevMech& M, const Lds& L, double T, double lnT, double iT,
                                      double lnp)
{
    const double PbyRT = PSTD / (RU * T);
    for (int r = lane_id(); r < M.nr; r += WAVE) {
        const int t = M.rtype[r];
        const unsigned long long pk = M.rpack[r];
        const int nl = pk_nl(pk), nrh = pk_nr(pk);
        int s[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) s[q] = pk_species(pk, q);
        const RxnK k = rate_constant<JAC>(M, L, r, t, T, lnT, iT, lnp);
        double kr = 0.0, dkrdT = 0.0, invKc = 0.0;
        if (M.rrev[r]) {
            double dlnKc = 0.0;
            double Kc = equilibrium<JAC>(L, s, nl, nrh, PbyRT, iT, &dlnKc);
            const bool floored = !(Kc > KC_FLOOR);
            Kc = floored ? KC_FLOOR : Kc;
            invKc = 1.0 / Kc;
            kr = k.kf * invKc;
            if (JAC) dkrdT = k.dkfdT * invKc - (floored ? 0.0 : kr * dlnKc);
        }
        double CF = 1.0, CR = 1.0;
        for (int a = 0; a < nl; ++a) CF *= L.C[s[a]];
        for (int a = 0; a < nrh; ++a) CR *= L.C[s[3 + a]];
        if (!JAC) {
            L.rq[r] = k.kf * CF - kr * CR;
        } else {
            L.rq[r] = k.kf;
            L.rk[r] = kr;
            L.rdT[r] = k.dkfdT * CF - dkrdT * CR;
            if (t == RXN_TB || t == RXN_LIND || t == RXN_TROE)
                L.rw[M.rtb[r]] = k.dkfdM * CF - (k.dkfdM * invKc) * CR;
        }
    }
}

// ---------------------------------------------------------------- RHS
// FastChemistryModel::derivatives (src/ChemistryModel/FastChemistryModel.C:
// 194-273, 340-417): y (this lane's element) is clipped in place for species
// lanes; returns this lane's element of dy/dt.
template <int NP>
__device__ inline double derivatives(const DevMech& M, const Lds& L, double p, double lnp, double& y)
{
    const int lane = lane_id();
    const bool sp = lane < M.ns;
    if (sp) y = y > 0.0 ? y : 0.0;
    const double T = readlane_d(y, NP - 1);
    const double RuTByP = RU * T / p;                 // unclamped T for the density
    const double invW = sp ? M.invW[lane] : 0.0;
    const double invrho = wave_sum(sp ? y * invW * RuTByP : 0.0);
    const double rho = 1.0 / invrho;
    const double Tc = clampT(M, T), lnT = log(Tc), iT = 1.0 / Tc;
    Thermo th{0, 0, 0, 0, 0};
    if (sp) th = species_thermo(M, lane, Tc, lnT, iT);
    L.C[lane] = sp ? rho * invW * y : 0.0;
    L.expG[lane] = sp ? exp(th.negG) : 0.0;
    wave_sync();
    reaction_round<false>(M, L, Tc, lnT, iT, lnp);
    wave_sync();
    double om = 0.0;
    if (sp) {
        for (int e0 = 0; e0 < M.maxDeg; e0 += 8) {   // batches of 8 list entries
            double nu[8], q[8];
            int r[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const bool in = e0 + u < M.maxDeg;
                nu[u] = in ? M.sp_nu[(e0 + u) * WAVE + lane] : 0.0;
                r[u] = in ? M.sp_r[(e0 + u) * WAVE + lane] : 0;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) q[u] = nu[u] != 0.0 ? L.rq[r[u]] : 0.0;
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (nu[u] != 0.0) om += nu[u] * q[u];
        }
    }
    const double Yd = sp ? om * (M.W[lane] * invrho) : 0.0;
    const double cpM = wave_sum(sp ? y * th.cp : 0.0);
    const double hY = wave_sum(sp ? th.ha * Yd : 0.0);
    wave_sync();
    return sp ? Yd : (lane == NP - 1 ? -hY / cpM : 0.0);
}

// ---------------------------------------------------------------- Jacobian
// FastChemistryModel::jacobian (src/ChemistryModel/FastChemistryModel.C:
// 489-572) -> OptReaction::ddNdtByVdcTp (src/Reaction/Jacobian.C:6-1001) and
// the dC/dY conversion.  The exact conversion uses the rank-1 identity
//   sum_k Jc_ik (rho/W_k)(delta_kj - (MW/W_j) Y_k)
//     = Jc_ij rho/W_j - (MW/W_j) sum_k Jc_ik C_k
// (O(n^2) instead of the reference's dense O(n^3) product,
// Jacobian.C:2814-2857); fast keeps the diagonal dC/dY only
// (Jacobian.C:4570-4841).  y is clipped in place; Jy is left in L.J (row i
// at lane i, columns 0..NP-1, temperature column NP-1); returns f(y) for
// this lane.
template <int NP>
__device__ inline double jacobian(const DevMech& M, const Lds& L, double p, double lnp, int jacType, double& y)
{
    const int lane = lane_id();
    const int ns = M.ns;
    const bool sp = lane < ns;
    if (sp) y = y > 0.0 ? y : 0.0;
    const double T = clampT(M, readlane_d(y, NP - 1));
    const double lnT = log(T), iT = 1.0 / T;
    const double invW = sp ? M.invW[lane] : 0.0;
    Thermo th{0, 0, 0, 0, 0};
    if (sp) th = species_thermo(M, lane, T, lnT, iT);
    const double RuTByP = RU * T / p;
    const double invrho = wave_sum(sp ? RuTByP * (y * invW) : 0.0);
    const double MW = 1.0 / wave_sum(sp ? y * invW : 0.0);
    const double rho = 1.0 / invrho;
    const double cpM = wave_sum(sp ? y * th.cp : 0.0);
    const double dcpMdT = wave_sum(sp ? y * th.dcpdT : 0.0);
    const double Cl = sp ? rho * y * invW : 0.0;
    L.C[lane] = Cl;
    L.expG[lane] = sp ? exp(th.negG) : 0.0;
    L.dBdT[lane] = th.dBdT;
    double* Jrow = L.J + lane;   // element j of this lane's row at Jrow[j*64]
    for (int j = 0; j < NP; ++j) Jrow[j * WAVE] = 0.0;
    wave_sync();
    reaction_round<true>(M, L, T, lnT, iT, lnp);
    wave_sync();
    // species rows: sparse partials dq/dC into this lane's row.  The list is
    // walked in blocks of 4 entries: all loads of a block are issued before
    // its (sequential, order-preserving) updates.
    double om = 0.0, jT = 0.0;
    if (sp) {
        for (int e0 = 0; e0 < M.maxDeg; e0 += 4) {
            double nu[4], kf[4], kr[4], rd[4], c[4][6];
            unsigned long long pk[4];
            int r[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool in = e0 + u < M.maxDeg;
                nu[u] = in ? M.sp_nu[(e0 + u) * WAVE + lane] : 0.0;
                r[u] = in ? M.sp_r[(e0 + u) * WAVE + lane] : 0;
                pk[u] = in ? M.sp_pack[(e0 + u) * WAVE + lane] : 0ull;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                kf[u] = L.rq[r[u]];
                kr[u] = L.rk[r[u]];
                rd[u] = L.rdT[r[u]];
#pragma unroll
                for (int q = 0; q < 6; ++q) c[u][q] = L.C[pk_species(pk[u], q)];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (nu[u] == 0.0) continue;
                // (slot loops fully unrolled with guards: register-resident c[][])
                const int nl = pk_nl(pk[u]), nrh = pk_nr(pk[u]);
                double CF = 1.0, CR = 1.0;
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (q < nl) CF *= c[u][q];
                    if (q < nrh) CR *= c[u][3 + q];
                }
                om += nu[u] * (kf[u] * CF - kr[u] * CR);
                jT += nu[u] * rd[u];
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    if (q >= nl) continue;
                    double d = kf[u];
#pragma unroll
                    for (int b = 0; b < 3; ++b) if (b < nl && b != q) d *= c[u][b];
                    Jrow[pk_species(pk[u], q) * WAVE] += nu[u] * d;
                }
                if (kr[u] != 0.0) {
#pragma unroll
                    for (int q = 0; q < 3; ++q) {
                        if (q >= nrh) continue;
                        double d = kr[u];
#pragma unroll
                        for (int b = 0; b < 3; ++b) if (b < nrh && b != q) d *= c[u][3 + b];
                        Jrow[pk_species(pk[u], 3 + q) * WAVE] -= nu[u] * d;
                    }
                }
            }
        }
        // third-body / falloff partials: nu * w_r * eff_rj on every column
        // (species-major efficiencies: the lanes' loads of one column share a
        // row).  Blocks of 4 entries x 4 columns; each element still takes
        // the entries' contributions in list order.
        for (int e0 = 0; e0 < M.maxDegTB; e0 += 4) {
            double coef[4];
            int t[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const bool in = e0 + u < M.maxDegTB;
                const double nu = in ? M.sptb_nu[(e0 + u) * WAVE + lane] : 0.0;
                t[u] = in ? M.sptb_t[(e0 + u) * WAVE + lane] : 0;
                coef[u] = nu != 0.0 ? nu * L.rw[t[u]] : 0.0;
            }
            const bool any = coef[0] != 0.0 || coef[1] != 0.0 || coef[2] != 0.0 || coef[3] != 0.0;
            if (!any) continue;
            for (int j0 = 0; j0 < ns; j0 += 4) {
                double v[4], ef[4][4];
#pragma unroll
                for (int w = 0; w < 4; ++w)
                    if (j0 + w < ns) {
                        v[w] = Jrow[(j0 + w) * WAVE];
#pragma unroll
                        for (int u = 0; u < 4; ++u) ef[u][w] = M.effT[(size_t)(j0 + w) * M.nTB + t[u]];
                    }
#pragma unroll
                for (int w = 0; w < 4; ++w)
                    if (j0 + w < ns) {
                        double acc = v[w];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (coef[u] != 0.0) acc += coef[u] * ef[u][w];
                        Jrow[(j0 + w) * WAVE] = acc;
                    }
            }
        }
    }
    // dC/dY conversion of the species rows, T column
    const double WiByRho = sp ? M.W[lane] * invrho : 0.0;
    const double Yd = sp ? om * WiByRho : 0.0;
    if (sp) {
        double sC = 0.0;
        for (int j0 = 0; j0 < ns; j0 += 8) {
            double v[8], c[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j0 + u < ns) {
                    v[u] = Jrow[(j0 + u) * WAVE];
                    c[u] = L.C[j0 + u];
                }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j0 + u < ns) sC += v[u] * c[u];
        }
        const double sCx = jacType == 0 ? sC : 0.0;   // fast: diagonal dC/dY only
        for (int j0 = 0; j0 < ns; j0 += 8) {
            double v[8], iw[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j0 + u < ns) {
                    v[u] = Jrow[(j0 + u) * WAVE];
                    iw[u] = M.invW[j0 + u];
                }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j0 + u < ns) {
                    const double iWj = iw[u];
                    Jrow[(j0 + u) * WAVE] = jacType == 0
                        ? WiByRho * (v[u] * (rho * iWj) - (MW * iWj) * sCx) + (MW * iWj) * Yd
                        : WiByRho * (v[u] * (rho * iWj)) + (MW * iWj) * Yd;
                }
        }
        Jrow[(NP - 1) * WAVE] = WiByRho * (jT - sC * iT) + iT * Yd;
    }
    // temperature row (src/Reaction/Jacobian.C:3743-3791): lane j < NP forms
    // column j's sum over the species rows and writes row NP-1, column j
    const double dTdt = -wave_sum(sp ? Yd * th.ha : 0.0) / cpM;
    L.H[lane] = sp ? th.ha : 0.0;
    L.Q[lane] = sp ? Yd * th.cp : 0.0;
    wave_sync();
    // (the sums read species rows only, so row NP-1 can be written at once)
    if (lane < NP) {
        const bool tcol = lane == NP - 1;
        double colsum = 0.0;
        const int i0 = lane % ns;   // rotated start: the lanes hit different banks
        for (int q0 = 0; q0 < ns; q0 += 8) {
            double v[8], h[8], qq[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (q0 + u < ns) {
                    int i = i0 + q0 + u;
                    i = i >= ns ? i - ns : i;
                    v[u] = L.J[lane * WAVE + i];
                    h[u] = L.H[i];
                    qq[u] = tcol ? L.Q[i] : 0.0;
                }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (q0 + u < ns) colsum += v[u] * h[u] + qq[u];
        }
        const double rowT = !tcol ? (sp ? (-colsum - th.cp * dTdt) / cpM : 0.0)
                                  : (-colsum - dTdt * dcpMdT) / cpM;
        L.J[lane * WAVE + (NP - 1)] = rowT;
    }
    wave_sync();
    return sp ? Yd : (lane == NP - 1 ? dTdt : 0.0);
}

// ---------------------------------------------------------------- LU
// Block-4 LU with pivoting restricted to each 4-row diagonal block
// (src/LUsolver/LUsolver.C:106-246 full blocks, :503-551 / :553-650 tails)
// on the LDS matrix.  Rows never move: pos is the logical row a lane holds,
// swapped as the reference swaps rows, and ip[r] is the lane holding
// logical row r when the factorisation ends.  Right-looking elimination:
// the pivot row is read by all lanes at one address.
struct LuState { double invD; int pos; int ipl; };   // ipl: lane holding logical row `lane`

// |value| of column-k entry in the lane that holds logical row r.
__device__ __forceinline__ double cand(double ak, int pos, int r)
{
    return fabs(readlane_d(ak, first_lane(pos == r)));
}

template <int NP>
__device__ inline LuState lu_factor(const Lds& L)
{
    static_assert(NP % 4 == 0, "padded system size must be a multiple of 4");
    const int lane = lane_id();
    const bool row = lane < NP;
    double* Jrow = L.J + lane;
    int pos = lane;          // lanes >= NP hold no row
    double invD = 0.0;
    for (int k = 0; k < NP; ++k) {
        const int k0 = k & ~3, d = k & 3;
        double ak = Jrow[k * WAVE];
        // --- pivot choice (uniform) ---
        int pick = k;
        if (d == 0) {
            double temp = 0.0;
            int im = 0;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const double v = cand(ak, pos, k0 + q);
                if (temp < v) { im = q; temp = v; }
            }
            pick = k0 + im;
        } else if (d == 1) {
            double t = cand(ak, pos, k0 + 1);
            int im = 1;
            const double v2 = cand(ak, pos, k0 + 2);
            if (t < v2) {
                im = 2;
                t = readlane_d(ak, first_lane(pos == k0 + 2));   // signed, LUsolver.C:177
            }
            const double v3 = cand(ak, pos, k0 + 3);
            if (t < v3) im = 3;
            pick = k0 + im;
        } else if (d == 2) {
            const double v2 = cand(ak, pos, k0 + 2);
            const double v3 = cand(ak, pos, k0 + 3);
            if (v2 < v3) pick = k0 + 3;
        }
        if (pick != k) {
            if (pos == k) pos = pick;
            else if (pos == pick) pos = k;
        }
        const int pl = first_lane(pos == k);
        double piv = readlane_d(ak, pl);
        if (piv == 0.0) {                        // LUsolver.C:155
            piv = PIVOT_EPS;
            if (lane == pl) Jrow[k * WAVE] = PIVOT_EPS;
        }
        const double inv = 1.0 / piv;
        if (lane == pl) invD = inv;
        const double* prow = L.J + pl;           // pivot row, broadcast reads
        // Every lane runs the update (uniform control flow); only rows below
        // the pivot (pos > k) take it, and only in columns j > k.  The column
        // range starts at a multiple of 4 (NP is one), so every block of 4
        // is whole: its 8 loads are issued before its 4 stores.
        const bool elim = row && pos > k;
        const double l = ak * inv;
        if (elim) Jrow[k * WAVE] = l;
        for (int j0 = (k + 1) & ~3; j0 < NP; j0 += 4) {
            double v[4], pv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                v[u] = Jrow[(j0 + u) * WAVE];
                pv[u] = prow[(j0 + u) * WAVE];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u)
                Jrow[(j0 + u) * WAVE] = (elim && j0 + u > k) ? v[u] - l * pv[u] : v[u];
        }
        wave_sync();
    }
    if (row) L.ip[pos] = lane;
    wave_sync();
    const int ipl = row ? L.ip[lane] : lane;
    wave_sync();
    return LuState{invD, pos, ipl};
}

// Solves (L U) x = P b; b is this lane's element in the original row order
// (lane i = row i); returns x_i for lane i (src/LUsolver/LUsolver.C:
// 3922-4111: swaps, unit-L forward, U backward scaled by invD).
template <int NP>
__device__ inline double lu_solve(const Lds& L, const LuState& st, double b)
{
    const int lane = lane_id();
    const double* Jrow = L.J + lane;
#pragma unroll
    for (int r = 0; r < NP - 1; ++r) {
        const int src = readlane_i(st.ipl, r);
        const double yr = readlane_d(b, src);
        const double lr = Jrow[r * WAVE];            // unconditional: prefetchable
        const double nb = b - lr * yr;
        b = (st.pos > r && lane < NP) ? nb : b;
    }
    double x = 0.0;
#pragma unroll
    for (int r = NP - 1; r >= 0; --r) {
        const int src = readlane_i(st.ipl, r);
        if (lane == src) b *= st.invD;
        const double xr = readlane_d(b, src);
        const double ur = Jrow[r * WAVE];
        const double nb = b - ur * xr;
        b = st.pos < r ? nb : b;
        if (lane == r) x = xr;
    }
    return x;
}

}  // namespace fcs
