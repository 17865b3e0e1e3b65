// fcs_integrate.h -- gfx950 device code of the ODE steppers: Rodas3(4),
// ROS3(4), Rodas3(2), ROS3(2) and linearly-implicit Euler extrapolation (Seulex), each with the
// reference's step-size controller, run by one wavefront per cell.
#pragma once
#include "fcs_device.h"
#include "fcs_lu_large.h"

namespace fcs {

// Controller constants: src/ODE/OptRodas34/OptRodas34.H:105-112
constexpr double SAFE = 0.9, ALPHA_INC = 0.2, ALPHA_DEC = 0.25, MIN_SCALE = 0.2, MAX_SCALE = 10.0;

// Rodas3(4) coefficients (Sandu et al. 1997): src/ODE/OptRodas34/OptRodas34.H:116-149
namespace rodas {
constexpr double d1 = 0.25, d2 = -0.1043, d3 = 0.1035, d4 = -0.3620000000000023e-01;
constexpr double a21 = 0.1544e1, a31 = 0.9466785280815826, a32 = 0.2557011698983284;
constexpr double a41 = 0.3314825187068521e1, a42 = 0.2896124015972201e1, a43 = 0.9986419139977817;
constexpr double a51 = 0.1221224509226641e1, a52 = 0.6019134481288629e1, a53 = 0.1253708332932087e2,
                 a54 = -0.6878860361058950;
constexpr double c21 = -0.56688e1, c31 = -0.2430093356833875e1, c32 = -0.2063599157091915;
constexpr double c41 = -0.1073529058151375, c42 = -0.9594562251023355e1, c43 = -0.2047028614809616e2;
constexpr double c51 = 0.7496443313967647e1, c52 = -0.1024680431464352e2, c53 = -0.3399990352819905e2,
                 c54 = 0.1170890893206160e2;
constexpr double c61 = 0.8083246795921522e1, c62 = -0.7981132988064893e1, c63 = -0.3152159432874371e2,
                 c64 = 0.1631930543123136e2, c65 = -0.6058818238834054e1;
constexpr double invGamma = 4.0;
}  // namespace rodas

// ROS3(4) (Shampine 1982): src/ODE/OptRosenbrock34/OptRosenbrock34.H:127-158
namespace ros {
constexpr double a21 = 2.0, a31 = 48.0 / 25.0, a32 = 6.0 / 25.0;
constexpr double c21 = -8.0, c31 = 372.0 / 25.0, c32 = 12.0 / 5.0;
constexpr double c41 = -112.0 / 125.0, c42 = -54.0 / 125.0, c43 = -2.0 / 5.0;
constexpr double b1 = 19.0 / 9.0, b2 = 1.0 / 2.0, b3 = 25.0 / 108.0, b4 = 125.0 / 108.0;
constexpr double e1 = 34.0 / 108.0, e2 = 7.0 / 36.0, e4 = 125.0 / 108.0;
constexpr double d1 = 1.0 / 2.0, d2 = -3.0 / 2.0, d3 = 605.0 / 250.0, d4 = 29.0 / 250.0;
constexpr double invGamma = 2.0;
}  // namespace ros

// Rodas3(2) (Sandu et al. 1997): src/ODE/OptRodas23/OptRodas23.H
namespace rodas23 {
constexpr double d1 = 1.0 / 2.0, d2 = 3.0 / 2.0;
constexpr double a31 = 2.0;
constexpr double c21 = 4.0, c31 = 1.0, c32 = -1.0, c41 = 1.0, c42 = -1.0, c43 = -8.0 / 3.0;
constexpr double invGamma = 1.0 / (1.0 / 2.0);
}  // namespace rodas23

// ROS3 with embedded order 2 (Sandu et al. 1997):
// src/ODE/OptRosenbrock23/OptRosenbrock23.H
namespace ros23 {
constexpr double a21 = 1.0;
constexpr double c21 = -1.0156171083877702091975600115545;
constexpr double c31 = 4.0759956452537699824805835358067, c32 = 9.2076794298330791242156818474003;
constexpr double b1 = 1.0, b2 = 6.1697947043828245592553615689730, b3 = -0.4277225654321857332623837380651;
constexpr double e1 = 0.5, e2 = -2.9079558716805469821718236208017, e3 = 0.2235406989781156962736090927619;
constexpr double gamma = 0.43586652150845899941601945119356;
constexpr double d1 = 0.43586652150845899941601945119356, d2 = 0.24291996454816804366592249683314,
                 d3 = 2.1851380027664058511513169485832;
constexpr double invGamma = 1.0 / gamma;
}  // namespace ros23

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// Max-norm error (src/ODE/OptRodas34/OptRodas34.C:822-877); NaN in any
// lane is returned as NaN.
template <int NP>
__device__ inline double error_norm(const DevOpts& o, const Lds& L, double y0, double yT, double err)
{
    const int lane = tid<NP>();
    double r = 0.0;
    if (lane < NP) {
        const double a0 = fabs(y0), a1 = fabs(yT);
        r = fabs(err) / (o.absTol + o.relTol * (a1 > a0 ? a1 : a0));
    }
    const bool bad = gany<NP>(!(r == r), L);
    const double mx = gmax<NP>(r, L);
    return bad ? __builtin_nan("") : mx;
}

// Iteration matrix diag*I - J in place, in batches of FB elements whose
// loads are all in flight together (one round trip to LDS or to the HBM
// slice per batch rather than per element).
template <int NP>
__device__ inline void form_matrix(const Lds& L, double diag)
{
    constexpr int JSt = Cell<NP>::JSt;
    constexpr int FB = NP % 16 == 0 ? 16 : (NP % 8 == 0 ? 8 : 4);
    const int lane = tid<NP>();
    double* Jrow = L.J + lane;
    if (!Cell<NP>::kLarge || lane < NP) {
        for (int j0 = 0; j0 < NP; j0 += FB) {
            double v[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) v[u] = Jrow[(j0 + u) * JSt];
#pragma unroll
            for (int u = 0; u < FB; ++u) Jrow[(j0 + u) * JSt] = (lane == j0 + u ? diag : 0.0) - v[u];
        }
    }
    wave_sync();
}

// LU factor / solve of either form.
template <int NP>
__device__ inline auto lu_factor_any(const DevMech& M, const Lds& L)
{
    if constexpr (Cell<NP>::kLarge) return lu_factor_large<NP>(M, L);
    else return lu_factor<NP>(M, L);
}
template <int NP, class St>
__device__ inline double lu_solve_any(const Lds& L, const St& st, double b)
{
    if constexpr (Cell<NP>::kLarge) return lu_solve_large<NP>(L, st, b);
    else return lu_solve<NP>(L, st, b);
}

// Rosenbrock tableau in the transformed (Hairer-Wanner) form the reference
// codes stage by stage.  Stage s (0-based) solves
//   (I/(gamma dx) - J) k_s = f(y0 + sum_j A[s][j] k_j) + D[s] dx f(y0)
//                            + sum_j C[s][j]/dx k_j           (s > 0)
//   (I/(gamma dx) - J) k_0 = (1 + D[0] dx) f(y0)
// with f(y0) returned by the Jacobian evaluation (quirk of the reference).
// A stage with newF = 0 reuses the previous stage's f.
struct Tableau {
    int stages;
    double invGamma;
    int newF[6];
    double A[6][6];
    double C[6][6];
    double D[6];
    double B[6];    // solution weights (final y = y0 + sum B k)
    double E[6];    // error weights   (err = sum E k)
};

// Rodas3(4): src/ODE/OptRodas34/OptRodas34.H:116-149 and the stage sums of
// OptRodas34.C:345-820.  Stage 5 evaluates f at y0 + dy with
// dy = sum a5j k_j, stage 6 at y0 + dy + k5; the new state is
// y0 + dy + k5 + k6 and the error estimate is k6.
__constant__ const Tableau kRodas34 = {
    6, rodas::invGamma, {0, 1, 1, 1, 1, 1},
    {{0, 0, 0, 0, 0, 0},
     {rodas::a21, 0, 0, 0, 0, 0},
     {rodas::a31, rodas::a32, 0, 0, 0, 0},
     {rodas::a41, rodas::a42, rodas::a43, 0, 0, 0},
     {rodas::a51, rodas::a52, rodas::a53, rodas::a54, 0, 0},
     {rodas::a51, rodas::a52, rodas::a53, rodas::a54, 1.0, 0}},
    {{0, 0, 0, 0, 0, 0},
     {rodas::c21, 0, 0, 0, 0, 0},
     {rodas::c31, rodas::c32, 0, 0, 0, 0},
     {rodas::c41, rodas::c42, rodas::c43, 0, 0, 0},
     {rodas::c51, rodas::c52, rodas::c53, rodas::c54, 0, 0},
     {rodas::c61, rodas::c62, rodas::c63, rodas::c64, rodas::c65, 0}},
    {rodas::d1, rodas::d2, rodas::d3, rodas::d4, 0, 0},
    {rodas::a51, rodas::a52, rodas::a53, rodas::a54, 1.0, 1.0},
    {0, 0, 0, 0, 0, 1.0}};

// ROS3(4): src/ODE/OptRosenbrock34/OptRosenbrock34.H:127-158 and
// OptRosenbrock34.C:340-640 (stage 4 reuses the stage-3 derivative).
__constant__ const Tableau kRos34 = {
    4, ros::invGamma, {0, 1, 1, 0, 0, 0},
    {{0, 0, 0, 0, 0, 0},
     {ros::a21, 0, 0, 0, 0, 0},
     {ros::a31, ros::a32, 0, 0, 0, 0},
     {ros::a31, ros::a32, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0}},
    {{0, 0, 0, 0, 0, 0},
     {ros::c21, 0, 0, 0, 0, 0},
     {ros::c31, ros::c32, 0, 0, 0, 0},
     {ros::c41, ros::c42, ros::c43, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0}},
    {ros::d1, ros::d2, ros::d3, ros::d4, 0, 0},
    {ros::b1, ros::b2, ros::b3, ros::b4, 0, 0},
    {ros::e1, ros::e2, 0.0, ros::e4, 0, 0}};

// Rodas3(2): OptRodas23.C:245-311.  Stage 2 takes f(y0) (newF = 0 with the
// carried f initialised to f(y0)); stage 3 evaluates f at y0 + a31 k1 and
// stage 4 at y0 + a31 k1 + k3; the new state is y0 + a31 k1 + k3 + k4 and the
// error estimate is k4.
__constant__ const Tableau kRodas23 = {
    4, rodas23::invGamma, {0, 0, 1, 1, 0, 0},
    {{0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {rodas23::a31, 0, 0, 0, 0, 0},
     {rodas23::a31, 0, 1.0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0}},
    {{0, 0, 0, 0, 0, 0},
     {rodas23::c21, 0, 0, 0, 0, 0},
     {rodas23::c31, rodas23::c32, 0, 0, 0, 0},
     {rodas23::c41, rodas23::c42, rodas23::c43, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0}},
    {rodas23::d1, rodas23::d2, 0, 0, 0, 0},
    {rodas23::a31, 0, 1.0, 1.0, 0, 0},
    {0, 0, 0, 1.0, 0, 0}};

// ROS3(2): OptRosenbrock23.C:253-290 (stage 3 reuses the stage-2 derivative).
__constant__ const Tableau kRos23 = {
    3, ros23::invGamma, {0, 1, 0, 0, 0, 0},
    {{0, 0, 0, 0, 0, 0},
     {ros23::a21, 0, 0, 0, 0, 0},
     {ros23::a21, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0}},
    {{0, 0, 0, 0, 0, 0},
     {ros23::c21, 0, 0, 0, 0, 0},
     {ros23::c31, ros23::c32, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0},
     {0, 0, 0, 0, 0, 0}},
    {ros23::d1, ros23::d2, ros23::d3, 0, 0, 0},
    {ros23::b1, ros23::b2, ros23::b3, 0, 0, 0},
    {ros23::e1, ros23::e2, ros23::e3, 0, 0, 0}};

// Tableau of a Rosenbrock solver id (SOLVER_*; Seulex has none).
template <int SOLVER>
__device__ __forceinline__ const Tableau& tableau_of()
{
    static_assert(SOLVER != SOLVER_SEULEX, "Seulex is not a Rosenbrock method");
    return SOLVER == SOLVER_ROS34 ? kRos34 : SOLVER == SOLVER_RODAS23 ? kRodas23 : SOLVER == SOLVER_ROS23 ? kRos23 : kRodas34;
}

// One Rosenbrock attempt from y0 with step dx.  y0 is clipped in place by
// the Jacobian (src/ChemistryModel/FastChemistryModel.C:498-501); yT
// receives the candidate state.  Returns the error norm.
template <int NP>
__device__ double rosenbrock_attempt(const DevMech& M, const Lds& L, const DevOpts& o, const Tableau& tb,
                                     double p, double lnp, double dx, double& y0, double& yT)
{
    const double invdx = 1.0 / dx;
    const bool first = tid<NP>() == 0;
    const double dfdx = jacobian<NP>(M, L, p, lnp, o.jac, y0);
    PhaseClock clk;
    form_matrix<NP>(L, 1.0 * invdx * tb.invGamma);
    clk.mark(M, first, PH_FORM);
    const auto st = lu_factor_any<NP>(M, L);
    clk.mark(M, first, PH_LU);
    double k[6] = {0, 0, 0, 0, 0, 0};
    double dydx = dfdx;   // a stage without its own f before any f evaluation takes f(y0)
    for (int s = 0; s < tb.stages; ++s) {
        double rhs;
        if (s == 0) {
            rhs = (dx * tb.D[0] + 1) * dfdx;
        } else {
            if (tb.newF[s]) {
                double yS = y0;
#pragma unroll
                for (int j = 0; j < 5; ++j) yS += tb.A[s][j] * k[j];
                clk.mark(M, first, PH_STAGE);
                dydx = derivatives<NP>(M, L, p, lnp, yS);
                clk = PhaseClock();
            }
            rhs = dydx + (dx * tb.D[s]) * dfdx;
#pragma unroll
            for (int j = 0; j < 5; ++j) rhs += (tb.C[s][j] * invdx) * k[j];
        }
        clk.mark(M, first, PH_STAGE);
        const double ks = lu_solve_any<NP>(L, st, rhs);
        clk.mark(M, first, PH_SOLVE);
#pragma unroll
        for (int j = 0; j < 6; ++j)
            if (j == s) k[j] = ks;
    }
    double err = 0.0;
    yT = y0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        yT += tb.B[j] * k[j];
        err += tb.E[j] * k[j];
    }
    const double norm = error_norm<NP>(o, L, y0, yT, err);
    clk.mark(M, first, PH_ERR);
    return norm;
}

// Step-size proposal after an accepted step of size dx with error norm Err
// (src/ODE/OptRodas34/OptRodas34.C:276-285).
__device__ __forceinline__ double accepted_dxTry(double Err, double dx)
{
    const double ratio = pow(MAX_SCALE / SAFE, -1.0 / ALPHA_INC);
    if (Err > ratio) {
        double f = SAFE * pow(Err, -ALPHA_INC);
        f = f > MIN_SCALE ? f : MIN_SCALE;
        f = f < MAX_SCALE ? f : MAX_SCALE;
        return f * dx;
    }
    return SAFE * MAX_SCALE * dx;
}

// Step-controller state of one cell between launches (flags bits).
enum { STEP_LAST = 1, STEP_RETRY = 2 };
// Return value of a step that leaves the cell in flight.
constexpr int STEP_CONTINUE = -1;

// One attempt of ODESolve + adaptiveSolve of the Rosenbrock solvers
// (src/ODE/OptRodas34/OptRodas34.C:123-286), written as a resumable state
// machine: the controller variables of the reference's loops (x, dxTry, the
// trial step dx of a step being retried, dxTry0, last, the step and attempt
// counters) are carried by the caller between calls.  Each call makes
// exactly one attempt; it returns STEP_CONTINUE while the cell is in
// flight, else its final status (0 when xEnd is reached).  Rejections are
// capped at 100 * maxSteps attempts in total, a guard the reference does not
// have (its rejection loop is unbounded).  dxAcc receives the step size
// when this call accepts a step (0 otherwise).
template <int NP>
__device__ int rosenbrock_step(const DevMech& M, const Lds& L, const DevOpts& o, const Tableau& tb, double p,
                               double xEnd, double& y0, double& x, double& dxTry, double& dx, double& dxTry0,
                               int& flags, int& nStep, int& nAtt, double& dxAcc)
{
    dxAcc = 0.0;
    const double xStart = 0.0;
    if (!(flags & STEP_RETRY)) {
        dxTry0 = dxTry;
        if ((x + dxTry - xEnd) * (x + dxTry - xStart) > 0) {
            flags |= STEP_LAST;
            dxTry = xEnd - x;
        }
        dx = dxTry;
    }
    double yT = 0.0;
    const double Err = rosenbrock_attempt<NP>(M, L, o, tb, p, log(p), dx, y0, yT);
    ++nAtt;
    if (!(Err == Err)) return CELL_NONFINITE;
    if (Err > 1) {
        if (nAtt > 100 * o.maxSteps) return CELL_MAXSTEPS;
        const double sc = SAFE * pow(Err, -ALPHA_DEC);
        dx *= sc > MIN_SCALE ? sc : MIN_SCALE;
        flags |= STEP_RETRY;
        return STEP_CONTINUE;
    }
    flags &= ~STEP_RETRY;
    x += dx;
    y0 = yT;
    dxAcc = dx;
    dxTry = accepted_dxTry(Err, dx);
    if ((x - xEnd) * (xEnd - xStart) >= 0) {
        if (nStep > 0 && (flags & STEP_LAST)) dxTry = dxTry0;
        return 0;
    }
    if (++nStep >= o.maxSteps) return CELL_MAXSTEPS;
    return STEP_CONTINUE;
}

// ---------------------------------------------------------------- Seulex
// src/ODE/OptSeulex/OptSeulex.C:33-731.  Per-lane state: the extrapolation
// table (kMax rows), the step-start state y0_ and the scale vector; the
// Jacobian Jy stays resident in the LDS buffer between Jacobian updates.
constexpr int SX_KMAX = 12, SX_IMAX = 13;

struct SeulexConst {
    int nSeq[SX_IMAX];
    double cpu[SX_IMAX];
    double coeff[SX_IMAX][SX_IMAX];
};

// Step-number sequence, work model and extrapolation coefficients
// (src/ODE/OptSeulex/OptSeulex.C:63-89).
constexpr SeulexConst make_seulex_const()
{
    SeulexConst s{};
    const double cpuFunc = 1, cpuJac = 5, cpuLU = 1, cpuSolve = 1;
    s.nSeq[0] = 2;
    s.nSeq[1] = 3;
    for (int i = 2; i < SX_IMAX; ++i) s.nSeq[i] = 2 * s.nSeq[i - 2];
    s.cpu[0] = cpuJac + cpuLU + s.nSeq[0] * (cpuFunc + cpuSolve);
    for (int k = 0; k < SX_KMAX; ++k) s.cpu[k + 1] = s.cpu[k] + (s.nSeq[k + 1] - 1) * (cpuFunc + cpuSolve) + cpuLU;
    for (int k = 0; k < SX_IMAX; ++k)
        for (int l = 0; l < SX_IMAX; ++l)
            s.coeff[k][l] = (l < k) ? 1.0 / ((double)s.nSeq[k] / s.nSeq[l] - 1.0) : 0.0;
    return s;
}
__constant__ const SeulexConst kSeulex = make_seulex_const();

// The Jacobian Jy of a Seulex cell persists across its steps (it is
// re-evaluated only when theta > jacRedo, src/ODE/OptSeulex/OptSeulex.C:
// 314-318, 483-490), so it lives in HBM between step launches, per cell,
// column-major: element (i, j) at JyG[j*JYS + i] (JYS = 64 for the
// wavefront form, NP for the workgroup form).  Each seul() forms its
// iteration matrix from it.
template <int NP>
__device__ __forceinline__ constexpr int jy_stride() { return Cell<NP>::kLarge ? NP : WAVE; }

template <int NP>
__device__ __forceinline__ void jy_store(const Lds& L, double* JyG)
{
    constexpr int FB = NP % 16 == 0 ? 16 : (NP % 8 == 0 ? 8 : 4);   // loads in flight together
    const int lane = tid<NP>();
    if (!Cell<NP>::kLarge || lane < NP)
        for (int j0 = 0; j0 < NP; j0 += FB) {
            double v[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) v[u] = L.J[(j0 + u) * Cell<NP>::JSt + lane];
#pragma unroll
            for (int u = 0; u < FB; ++u) JyG[(j0 + u) * jy_stride<NP>() + lane] = v[u];
        }
}

// Iteration matrix diag*I - Jy in the LDS matrix, Jy read from HBM.
template <int NP>
__device__ __forceinline__ void form_matrix_from(const Lds& L, const double* JyG, double diag)
{
    constexpr int FB = NP % 16 == 0 ? 16 : (NP % 8 == 0 ? 8 : 4);   // loads in flight together
    const int lane = tid<NP>();
    double* Jrow = L.J + lane;
    if (!Cell<NP>::kLarge || lane < NP)
        for (int j0 = 0; j0 < NP; j0 += FB) {
            double v[FB];
#pragma unroll
            for (int u = 0; u < FB; ++u) v[u] = JyG[(j0 + u) * jy_stride<NP>() + lane];
#pragma unroll
            for (int u = 0; u < FB; ++u) Jrow[(j0 + u) * Cell<NP>::JSt] = (lane == j0 + u ? diag : 0.0) - v[u];
        }
    wave_sync();
}

// seul(): src/ODE/OptSeulex/OptSeulex.C:562-706, written with a single RHS
// evaluation site.  y0s (this lane's y0_) is clipped in place by the first
// RHS evaluation; returns false when the theta divergence test fails.
template <int NP>
__device__ bool seul(const DevMech& M, const Lds& L, const double* JyG, double p, double lnp, double& y0s,
                     double dxTot, int k, int nSteps, double scale, double& theta, double& ySeq)
{
    const int lane = tid<NP>();
    const double invdx = nSteps / dxTot;
    form_matrix_from<NP>(L, JyG, invdx);
    const auto st = lu_factor_any<NP>(M, L);
    double yT = 0.0, dy = 0.0;
    int nn = 0;
    bool thetaPhase = false;
    double dy1 = 0.0;
    for (;;) {
        double ev = (nn == 0) ? y0s : yT;
        const double f = derivatives<NP>(M, L, p, lnp, ev);
        if (nn == 0) y0s = ev; else yT = ev;
        dy = lu_solve_any<NP>(L, st, thetaPhase ? f - dy * invdx : f);
        if (nn == 0) yT = y0s;
        if (thetaPhase) {
            const bool huge = gany<NP>(lane < NP && fabs(dy) > sqrt(OF_VGREAT), L);
            const double s2 = gsum<NP>(lane < NP ? (dy / scale) * (dy / scale) : 0.0, L);
            double dy2 = huge ? OF_VGREAT : s2;
            dy2 = sqrt(dy2);
            const double den = 1.0 > dy1 + OF_SMALL ? 1.0 : dy1 + OF_SMALL;
            theta = dy2 / den;
            if (theta > 1.0) return false;
            thetaPhase = false;
            continue;
        }
        ++nn;
        if (nn >= nSteps) break;
        yT += dy;
        if (nn == 1 && k <= 1) {
            dy1 = sqrt(gsum<NP>(lane < NP ? (dy / scale) * (dy / scale) : 0.0, L));
            thetaPhase = true;
        }
    }
    ySeq = yT + dy;
    return true;
}

// Seulex controller state of one cell between launches.
enum { SX_FIRST = 1, SX_LAST = 2, SX_PREVREJECT = 4, SX_JACUPD = 8, SX_INSTEP = 16 };

struct SeulexCell {
    double x, dxTry, dxNew, dxTry0, theta;
    int kTarg, flags, nStep, nAtt;
};

// One pass of the extrapolation loop of OptSeulex (ODESolve + SeulexSolve,
// src/ODE/OptSeulex/OptSeulex.C:171-559): the body of SeulexSolve's
// `while (firstk || step.reject)` loop, i.e. columns k = 0..kTarg+1 from the
// step-start state, preceded by the step set-up when a new step begins and
// followed by the step-size / order update when the pass is accepted.  The
// controller variables of the reference (theta_, kTarg_, the stepState
// flags, dxNew, the step counters) are carried by the caller between calls,
// with the per-lane state y, the step-start state y0s and the scale vector
// (k11).  temp_ and dxOpt_ are only read at columns set earlier in the same
// pass (temp_[0] = great at every pass), so they stay local.  Returns
// STEP_CONTINUE while the cell is in flight, else its final status.
template <int NP>
__device__ int seulex_pass(const DevMech& M, const Lds& L, const DevOpts& o, double p, double xEnd,
                           double* JyG, double& y, double& y0s, double& scale, SeulexCell& c)
{
    const SeulexConst& S = kSeulex;
    const int lane = tid<NP>();
    const double lnp = log(p);
    const double kF1 = 0.7, kF2 = 0.9, sF1 = 0.8, sF2 = 0.93, sF3 = 0.1, sF4 = 4, sF5 = 0.5;
    const double jacRedo = o.relTol < 1e-4 ? o.relTol : 1e-4;
    const double xStart = 0.0;
    if (!(c.flags & SX_INSTEP)) {
        // ODESolve loop head (OptSeulex.C:198-207) and SeulexSolve set-up (:283-318)
        c.dxTry0 = c.dxTry;
        if ((c.x + c.dxTry - xEnd) * (c.x + c.dxTry - xStart) > 0) {
            c.flags |= SX_LAST;
            c.dxTry = xEnd - c.x;
        }
        y0s = y;
        if (c.flags & (SX_FIRST | SX_PREVREJECT)) c.theta = 2 * jacRedo;
        if (c.flags & SX_FIRST) {
            const double logTol = -log10(o.relTol + o.absTol) * 0.6 + 0.5;
            int kt = (int)logTol;
            kt = kt < SX_KMAX - 1 ? kt : SX_KMAX - 1;
            c.kTarg = kt > 1 ? kt : 1;
        }
        scale = o.absTol + o.relTol * fabs(y);
        c.flags &= ~SX_JACUPD;
        if (c.theta > jacRedo) {
            (void)jacobian<NP>(M, L, p, lnp, o.jac, y);
            jy_store<NP>(L, JyG);
            c.flags |= SX_JACUPD;
        }
        c.dxNew = fabs(c.dxTry);
        c.flags |= SX_INSTEP;
    }
    double dxOpt[SX_IMAX], temp[SX_IMAX];
    double table[SX_KMAX];
#pragma unroll
    for (int q = 0; q < SX_KMAX; ++q) table[q] = 0.0;
    temp[0] = OF_GREAT;
    const bool first = c.flags & SX_FIRST, last = c.flags & SX_LAST;
    bool prevReject = c.flags & SX_PREVREJECT;
    const double dx = c.dxNew;
    bool reject = false;
    double errOld = 0;
    int k = 0;
    for (k = 0; k <= c.kTarg + 1; ++k) {
        double ySeq = 0.0;
        const bool ok = seul<NP>(M, L, JyG, p, lnp, y0s, dx, k, S.nSeq[k], scale, c.theta, ySeq);
        ++c.nAtt;
        if (!ok) { reject = true; c.dxNew = fabs(dx) * sF5; break; }
        if (k == 0) {
            y = ySeq;
            continue;
        }
        // table[k-1] = ySeq, then extrapolate (OptSeulex.C:711-731)
#pragma unroll
        for (int q = 0; q < SX_KMAX; ++q)
            if (q == k - 1) table[q] = ySeq;
#pragma unroll
        for (int j = SX_KMAX - 1; j > 0; --j)
            if (j <= k - 1) table[j - 1] = table[j] + S.coeff[k][j] * (table[j] - table[j - 1]);
        y = table[0] + S.coeff[k][0] * (table[0] - y);
        scale = o.absTol + o.relTol * fabs(y0s);
        const double dd = (y - table[0]) / scale;
        const double err = sqrt(gsum<NP>(lane < NP ? dd * dd : 0.0, L) / (M.ns + 1));
        if (!(err == err)) return CELL_NONFINITE;
        if (err > 1 / OF_SMALL || (k > 1 && err >= errOld)) {
            reject = true;
            c.dxNew = fabs(dx) * sF5;
            break;
        }
        errOld = 4 * err < 1 ? 4 * err : 1;
        const double expo = 1.0 / (k + 1);
        const double facmin = pow(sF3, expo);
        double fac;
        if (err == 0) fac = 1.0 / facmin;
        else {
            fac = sF2 / pow(err / sF1, expo);
            const double hi = 1 / facmin < fac ? 1 / facmin : fac;
            fac = facmin / sF4 > hi ? facmin / sF4 : hi;
        }
        dxOpt[k] = fabs(dx * fac);
        temp[k] = S.cpu[k] / dxOpt[k];
        if ((first || last) && err <= 1.0) break;
        if (k == c.kTarg - 1 && !prevReject && !first && !last) {
            if (err <= 1.0) break;
            else if (err > S.nSeq[c.kTarg] * S.nSeq[c.kTarg + 1] * 4.0) {
                reject = true;
                c.kTarg = k;
                if (c.kTarg > 1 && temp[k - 1] < kF1 * temp[k]) c.kTarg--;
                c.dxNew = dxOpt[c.kTarg];
                break;
            }
        }
        if (k == c.kTarg) {
            if (err <= 1.0) break;
            else if (err > S.nSeq[k + 1] * 2.0) {
                reject = true;
                if (c.kTarg > 1 && temp[k - 1] < kF1 * temp[k]) c.kTarg--;
                c.dxNew = dxOpt[c.kTarg];
                break;
            }
        }
        if (k == c.kTarg + 1) {
            if (err > 1.0) {
                reject = true;
                if (c.kTarg > 1 && temp[c.kTarg - 1] < kF1 * temp[c.kTarg]) c.kTarg--;
                c.dxNew = dxOpt[c.kTarg];
            }
            break;
        }
    }
    if (reject) {
        c.flags |= SX_PREVREJECT;
        if (!(c.flags & SX_JACUPD)) {
            c.theta = 2.0 * jacRedo;
            (void)jacobian<NP>(M, L, p, lnp, o.jac, y);
            jy_store<NP>(L, JyG);
            c.flags |= SX_JACUPD;
        }
        if (c.nAtt > 100 * o.maxSteps) return CELL_MAXSTEPS;
        return STEP_CONTINUE;
    }
    if (c.nAtt > 100 * o.maxSteps) return CELL_MAXSTEPS;
    // accepted pass: step-size and order update (OptSeulex.C:495-557)
    c.flags &= ~SX_INSTEP;
    c.x += dx;
    int kopt;
    if (k == 1) kopt = 2;
    else if (k <= c.kTarg) {
        kopt = k;
        if (temp[k - 1] < kF1 * temp[k]) kopt = k - 1;
        else if (temp[k] < kF2 * temp[k - 1]) kopt = (k + 1 < SX_KMAX - 1) ? k + 1 : SX_KMAX - 1;
    } else {
        kopt = k - 1;
        if (k > 2 && temp[k - 2] < kF1 * temp[k - 1]) kopt = k - 2;
        if (temp[k] < kF2 * temp[kopt]) kopt = (k < SX_KMAX - 1) ? k : SX_KMAX - 1;
    }
    double dxNew;
    if (prevReject) {
        c.kTarg = kopt < k ? kopt : k;
        dxNew = fabs(dx) < dxOpt[c.kTarg] ? fabs(dx) : dxOpt[c.kTarg];
        c.flags &= ~SX_PREVREJECT;
    } else {
        if (kopt <= k) dxNew = dxOpt[kopt];
        else {
            if (k < c.kTarg && temp[k] < kF2 * temp[k - 1]) dxNew = dxOpt[k] * S.cpu[kopt + 1] / S.cpu[k];
            else dxNew = dxOpt[k] * S.cpu[kopt] / S.cpu[k];
        }
        c.kTarg = kopt;
    }
    c.dxTry = dxNew;
    // ODESolve loop tail (OptSeulex.C:236-251)
    if ((c.x - xEnd) * (xEnd - xStart) >= 0) {
        if (c.nStep > 0 && last) c.dxTry = c.dxTry0;
        return 0;
    }
    c.flags &= ~SX_FIRST;
    if (++c.nStep >= o.maxSteps) return CELL_MAXSTEPS;
    return STEP_CONTINUE;
}

}  // namespace fcs
