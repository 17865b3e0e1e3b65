/*
 * fcs.h -- C ABI of the MI355X (gfx950) per-cell stiff chemistry integrator.
 *
 * This header is the drop-in boundary for the chemistry step of the
 * OpenFOAM-10 plug-in `FastChemistryModel`.  Every entry point takes plain
 * pointers and sizes; no C++ or framework types cross the boundary.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *   - solver selection `chemistryType { solver OptRodas34|OptRosenbrock34|
 *     OptRodas23|OptRosenbrock23|OptSeulex; method FastChemistryModel; }`
 *       src/odeChemistrySolvers.C:66-72, src/ChemistryModel/FastChemistryModel.H:230
 *   - Coeffs keys absTol/relTol/maxSteps        src/ODE/OptRodas34/OptRodas34.C:40-43,
 *                                                src/ODE/OptSeulex/OptSeulex.C:40-46
 *   - keys jacobian exact|fast, Treact          src/ChemistryModel/FastChemistryModel.C:54-59,66
 *   - scalar solve(const scalar deltaT)          src/ChemistryModel/FastChemistryModel.H:292,
 *                                                FastChemistryModel_transientSolve.H:3-104
 *   - mechanism tables built by readInfo        src/Reaction/OptReaction.C:348-1339
 *
 * Threading: a solver handle is used by one host thread at a time.  All
 * entry points return 0 on success and a negative FCS_E* code on failure;
 * the library never aborts the process (the reference's FatalError on
 * maxSteps becomes a per-cell status bit, see FCS_CELL_*).
 */
#ifndef FCS_H
#define FCS_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes ------------------------------------------------------ */
#define FCS_OK              0
#define FCS_E_ARG          -1   /* invalid argument / size                     */
#define FCS_E_HIP          -2   /* HIP runtime error (no device, launch fail)   */
#define FCS_E_UNSUPPORTED  -3   /* mechanism outside the compiled size range    */
#define FCS_E_NOMEM        -4   /* device or host allocation failed             */
#define FCS_E_CELLS        -5   /* one or more cells failed (see cellStatus)    */
#define FCS_E_SCRATCH      -6   /* input beyond fcs_solve_noalloc's scratch     */

/* ---- per-cell status bits (cellStatus output) ------------------------- */
#define FCS_CELL_OK            0
#define FCS_CELL_SKIPPED       1   /* T0 <= Treact: RR = 0, not integrated      */
#define FCS_CELL_MAXSTEPS      2   /* maxSteps exceeded (reference: FatalError) */
#define FCS_CELL_NONFINITE     4   /* NaN/Inf in the state or error norm        */
#define FCS_CELL_LEARNED       8   /* finished by a verified learned proposal   */

/* ---- reaction rate forms (fcs_mech_desc.rxn_type) --------------------- */
#define FCS_RXN_ARRHENIUS   0   /* (ir)reversibleArrhenius                      */
#define FCS_RXN_THIRD_BODY  1   /* (ir)reversibleThirdBodyArrhenius             */
#define FCS_RXN_LINDEMANN   2   /* (ir)reversibleArrheniusLindemannFallOff      */
#define FCS_RXN_TROE        3   /* (ir)reversibleArrheniusTroeFallOff           */
#define FCS_RXN_PLOG        4   /* (ir)reversiblePlog (log-p interpolation)     */
#define FCS_RXN_SRI         5   /* (ir)reversibleArrheniusSRIFallOff            */

/* ---- ODE solvers (chemistryType.solver) -------------------------------- */
#define FCS_SOLVER_RODAS34        0   /* "OptRodas34"       */
#define FCS_SOLVER_ROSENBROCK34   1   /* "OptRosenbrock34"  */
#define FCS_SOLVER_SEULEX         2   /* "OptSeulex"        */
#define FCS_SOLVER_RODAS23        3   /* "OptRodas23"       */
#define FCS_SOLVER_ROSENBROCK23   4   /* "OptRosenbrock23"  */

#define FCS_JACOBIAN_EXACT  0   /* jacobian exact; (default)                     */
#define FCS_JACOBIAN_FAST   1   /* jacobian fast;  diagonal dC/dY only           */

#define FCS_MAX_STOICH      3   /* reactant slots per reaction                    */
#define FCS_MAX_SLOTS      16   /* product slots per reaction (slot_stride)       */

/*
 * Flat mechanism description (species thermo + reactions, file order).
 * Units follow the reference: W in kg/kmol, concentrations kmol/m^3,
 * A in (m^3/kmol)^(order-1)/s, Ta in K, pressures in Pa.
 * Stoichiometry is integer and expanded: "2OH" is two slots holding OH
 * (src/Reaction/OptReaction.C:99-126).
 * Falloff reactions carry k_inf in (A,beta,Ta) and k_0 in (A0,beta0,Ta0).
 * efficiencies is a dense [n_reactions x n_species] row-major table read for
 * third-body and falloff reactions only (may be NULL when there are none).
 * troe is [n_reactions x 4] = (alpha, Tsss, Ts, Tss) read for TROE rows.
 * PLOG rows use plog_n[r] points stored at [r*plog_max_points + k].
 * rxn_reversible: 0 irreversible, 1 reverse rate kf/Kc, 2 nonEquilibrium:
 * the reverse rate is its own Arrhenius (A_rev, beta_rev, Ta_rev), times the
 * same third-body concentration for third-body rows
 * (nonEquilibriumReversible[ThirdBody]Arrhenius, OptReaction.C:617-700).
 * sri is [n_reactions x 5] = (a, b, c, d, e) read for SRI rows; sri, A_rev,
 * beta_rev and Ta_rev may be NULL when no row needs them.
 * rxn_lhs / rxn_rhs hold slot_stride slots per reaction (slot_stride 0 means
 * FCS_MAX_STOICH; at most FCS_MAX_SLOTS).  Reactant sides have 1..3 slots,
 * product sides 1..slot_stride.  A reaction with more than 3 product slots
 * takes the reference's general path (RFG / JFG, ReactionFunctionI.H:648-720,
 * JacobianFunctionI.H:2965-3060): a Kc reverse rate runs over every product
 * slot (RFG floors Kc at 1.49011611938476e-8, JFG at 1.4901171103413047e-8),
 * and a nonEquilibrium row has no reverse rate there.
 */
typedef struct fcs_mech_desc {
    int32_t n_species;
    int32_t n_reactions;
    const double* W;            /* [ns]                                   */
    const double* T_low;        /* [ns]                                   */
    const double* T_high;       /* [ns]                                   */
    const double* T_common;     /* [ns]                                   */
    const double* coeffs_high;  /* [ns*7] NASA-7, T >= T_common           */
    const double* coeffs_low;   /* [ns*7] NASA-7, T <  T_common           */
    const int32_t* rxn_type;    /* [nr] FCS_RXN_*                         */
    const int32_t* rxn_reversible; /* [nr] 1: reverse rate from Kc         */
    const int32_t* rxn_n_lhs;   /* [nr] 1..3                              */
    const int32_t* rxn_n_rhs;   /* [nr] 1..slot_stride                    */
    const int32_t* rxn_lhs;     /* [nr*slot_stride] species index per slot */
    const int32_t* rxn_rhs;     /* [nr*slot_stride]                       */
    const double* A;            /* [nr] forward Arrhenius / k_inf         */
    const double* beta;
    const double* Ta;
    const double* A0;           /* [nr] low-pressure k_0 (falloff rows)   */
    const double* beta0;
    const double* Ta0;
    const double* efficiencies; /* [nr*ns]                                */
    const double* troe;         /* [nr*4]                                 */
    int32_t plog_max_points;
    const int32_t* plog_n;      /* [nr]                                   */
    const double* plog_p;       /* [nr*plog_max_points]                   */
    const double* plog_A;
    const double* plog_beta;
    const double* plog_Ta;
    const double* sri;          /* [nr*5]                                 */
    const double* A_rev;        /* [nr] reverse Arrhenius (rxn_reversible == 2) */
    const double* beta_rev;
    const double* Ta_rev;
    int32_t slot_stride;        /* slots per reaction in rxn_lhs / rxn_rhs (0: 3) */
} fcs_mech_desc;

/* Solver options: the chemistryProperties keys of the reference. */
typedef struct fcs_solver_opts {
    int32_t solver;          /* FCS_SOLVER_*                                 */
    int32_t jacobian;        /* FCS_JACOBIAN_*                               */
    double  absTol;          /* <Solver>Coeffs.absTol                        */
    double  relTol;          /* <Solver>Coeffs.relTol                        */
    int32_t maxSteps;        /* <Solver>Coeffs.maxSteps (reference default 10000) */
    double  Treact;          /* cells with T0 <= Treact are skipped (default 0)   */
    double  deltaTChemMax;   /* maxChemicalTimeStep (OpenFOAM default: great)     */
} fcs_solver_opts;

typedef struct fcs_mech   fcs_mech;     /* device-resident mechanism tables   */
typedef struct fcs_solver fcs_solver;   /* mechanism + options + device state */

/* Library / device information. */
int fcs_device_count(int* count);
const char* fcs_version(void);

/* Mechanism: replaces OptReaction::readInfo (src/Reaction/OptReaction.C:348). */
int fcs_mech_create(const fcs_mech_desc* desc, fcs_mech** out);
int fcs_mech_destroy(fcs_mech* mech);

/* Solver: replaces the chemistrySolver construction selected by
 * makeChemistrySolver (src/odeChemistrySolvers.C:66-72). */
int fcs_solver_create(const fcs_mech* mech, const fcs_solver_opts* opts,
                      fcs_solver** out);
int fcs_solver_destroy(fcs_solver* solver);

/*
 * fcs_solve: FastChemistryModel::solve(const scalar deltaT), serial branch
 * (src/ChemistryModel/FastChemistryModel_transientSolve.H:3-104) over host
 * arrays.  Fields are per species, one pointer each (OpenFOAM stores one
 * volScalarField per species):
 *   T0, p0, rho0  [nCells]        old-time temperature, pressure, density
 *   Y0[i]         [nCells]        old-time mass fractions, i < n_species
 *   deltaTChem    [nCells] in/out per-cell chemical sub-step
 *   RR[i]         [nCells] out    (Y - Y0) * rho0 / deltaT, 0 for skipped cells
 *   deltaTMin     out             min(min over integrated cells of deltaTChem,
 *                                     2*deltaT)
 *   cellStatus    [nCells] out    FCS_CELL_* bits (may be NULL)
 * Returns FCS_E_CELLS when any cell reports MAXSTEPS or NONFINITE.
 */
int fcs_solve(fcs_solver* solver, int64_t nCells, double deltaT,
              const double* T0, const double* p0, const double* rho0,
              const double* const* Y0, double* deltaTChem,
              double* const* RR, double* deltaTMin, int32_t* cellStatus);

/* Same contract, also returning the integrated end state T (the reference
 * keeps it local) and the number of Rodas/Rosenbrock attempts or Seulex
 * columns per cell (diagnostics; any output may be NULL). */
int fcs_solve_ex(fcs_solver* solver, int64_t nCells, double deltaT,
                 const double* T0, const double* p0, const double* rho0,
                 const double* const* Y0, double* deltaTChem,
                 double* const* RR, double* deltaTMin, int32_t* cellStatus,
                 double* Tend, int32_t* nAttempts);

/*
 * fcs_solve_lts: FastChemistryModel::solve(const scalarField& deltaT), the
 * local-time-step variant (src/ChemistryModel/FastChemistryModel.H:296,
 * FastChemistryModel_localEulerSolve.H:3-101): cell c integrates over
 * deltaT[c] > 0, RR = (Y - Y0) * rho0 / deltaT[c], and deltaTMin is the
 * minimum over integrated cells of the new deltaTChem (great when none is
 * integrated), without the 2*deltaT cap of fcs_solve.  Tend and nAttempts
 * as in fcs_solve_ex (may be NULL).
 */
int fcs_solve_lts(fcs_solver* solver, int64_t nCells, const double* deltaT,
                  const double* T0, const double* p0, const double* rho0,
                  const double* const* Y0, double* deltaTChem,
                  double* const* RR, double* deltaTMin, int32_t* cellStatus,
                  double* Tend, int32_t* nAttempts);

/*
 * Field evaluations at the current state (per-species host pointers as in
 * fcs_solve; rho, T, p [nCells]; c_i = rho Y_i / W_i, not clipped).
 *
 * fcs_calculate: FastChemistryModel::calculate (src/ChemistryModel/
 *   FastChemistryModel.C:700-740): RR[i] = W_i * dNdtByV_i(p, T, c).
 * fcs_tc: FastChemistryModel::tc (FastChemistryModel.C:588-643) with
 *   OptReaction::Tc (src/Reaction/Tc.C:4-834): chemical time scale
 *   sumW / sumWRateByCTot * sum(c), 1e300 (vGreat) where nothing reacts.
 * fcs_qdot: FastChemistryModel::Qdot (FastChemistryModel.C:646-675):
 *   Qdot = -sum_i Hf_i RR[i], Hf_i the mass-specific enthalpy of formation
 *   (low-range NASA-7 fit at 298.15 K, as OpenFOAM-10's janafThermo::Hf).
 */
int fcs_calculate(fcs_solver* solver, int64_t nCells, const double* rho,
                  const double* T, const double* p, const double* const* Y,
                  double* const* RR);
int fcs_tc(fcs_solver* solver, int64_t nCells, const double* rho,
           const double* T, const double* p, const double* const* Y,
           double* tc);
int fcs_qdot(fcs_solver* solver, int64_t nCells, const double* const* RR,
             double* Qdot);

/*
 * Device-resident batch: inputs uploaded once, integrated in place in HBM.
 * Layout in HBM is structure-of-arrays, species-major: Y[i*nCells + c].
 */
typedef struct fcs_batch fcs_batch;
int fcs_batch_create(fcs_solver* solver, int64_t nCells, fcs_batch** out);
int fcs_batch_destroy(fcs_batch* batch);
int fcs_batch_upload(fcs_batch* batch, const double* T0, const double* p0,
                     const double* rho0, const double* Y0 /* [ns*nCells] */,
                     const double* deltaTChem);
/* Runs the integration of every cell of the batch over deltaT on the
 * batch's stream, restarting from the uploaded state each call.  When
 * elapsed_ms is non-NULL the call waits and returns the time between HIP
 * events recorded on that stream around the kernel launches.  Per-cell
 * failures are reported through the downloaded cellStatus; FCS_E_CELLS is
 * returned only if the run reaches its launch cap (101 * maxSteps step
 * launches, out of reach while every cell obeys its own attempt cap), after
 * the queued launches have drained and the cells still in flight have been
 * finished as FCS_CELL_MAXSTEPS. */
int fcs_batch_run(fcs_batch* batch, double deltaT, float* elapsed_ms);
/* Profiling of fcs_batch_run (no counterpart in the reference): with
 * enable != 0 each step launch is bracketed by a pair of HIP events on the
 * solver's stream; fcs_batch_stats reports the launch count of the last run
 * and the summed device duration of its step kernels (0 when disabled). */
int fcs_batch_profile(fcs_batch* batch, int enable);
int fcs_batch_stats(fcs_batch* batch, int64_t* stepLaunches, double* stepKernelMs);
/* Per-wavefront load of the last profiled run (no counterpart in the
 * reference; it reports what its MPI load balancer,
 * FastChemistryModel_transientSolve.H:106-686, evens out): the active-list
 * length each step launch read, in launch order (at most cap of them; the
 * count in *nLaunches), and the grid of the step launches.  A launch over n
 * cells gives wavefront w (0 <= w < grid) ceil((n - w) / grid) attempts, so
 * the lengths fix every wavefront's load.  Recorded with fcs_batch_profile
 * on (eager launches); empty otherwise. */
int fcs_batch_load_report(fcs_batch* batch, int64_t* listLen, int32_t cap, int32_t* nLaunches, int32_t* grid);
int fcs_batch_download(fcs_batch* batch, double* RR /* [ns*nCells] */,
                       double* deltaTChem, double* deltaTMin,
                       int32_t* cellStatus, double* Tend, int32_t* nAttempts);

/*
 * No-allocation solve(deltaT) (no counterpart in the reference; for callers
 * that cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The variant of
 * fcs_batch_upload + fcs_batch_run + fcs_batch_download over arrays the
 * caller already holds in device memory, in the batch layout
 * (Y0 and RR [ns*nCells], species-major); plain integrator only, same
 * kernels and results as fcs_batch_run.  It calls no hipMalloc, hipFree or
 * other allocating routine: the integrator state, active lists and counters
 * live in a library-owned 32 MiB device array, one per GPU, held until the
 * process ends; every call on a GPU shares it, so calls that overlap on one
 * GPU race.  About (ns+1)*8 + 64 bytes per cell (Rosenbrock solvers), so
 * some 70,000 cells at 53 species; larger inputs return FCS_E_SCRATCH.
 * FCS_E_UNSUPPORTED for mechanisms whose workgroup form keeps its matrix in
 * HBM (more than 111 species).  deltaTMin (host, may be NULL) as
 * fcs_batch_download returns it. */
int fcs_solve_noalloc(fcs_solver* solver, int64_t nCells, double deltaT,
                      const double* T0, const double* p0, const double* rho0,
                      const double* Y0 /* [ns*nCells] */, const double* deltaTChemIn,
                      double* deltaTChemOut, double* RR /* [ns*nCells] */,
                      int32_t* cellStatus, double* Tend, double* deltaTMin);

/*
 * Launch shape of a batch's step kernel (no counterpart in the reference:
 * the device work lists replace the MPI load balancer).  grid: workgroups
 * of the step kernel (0: occupancy x CUs, the default).  graphLaunches: 0
 * launches every step kernel from the host; a positive multiple of 6
 * (at most 6000) captures that many step launches in a hipGraph, replayed
 * until every cell has finished (launches after the lists have emptied
 * return at once; results are identical to eager launches).
 * fcs_batch_autotune times every grid of 1..occupancy workgroups per CU
 * with eager launches and with 24-launch graphs on the batch's uploaded
 * cells (two runs each), keeps the fastest and returns it.
 */
int fcs_batch_set_launch(fcs_batch* batch, int32_t grid, int32_t graphLaunches);
int fcs_batch_autotune(fcs_batch* batch, double deltaT, int32_t* grid, int32_t* graphLaunches);

/*
 * Learned seed on a batch (no counterpart in the reference).  The network
 * (fcs_mlp) maps [T, ln p, Y_0..Y_ns-1, ln deltaT] of each reacting cell to
 * [ln h, z_0..z_ns-1, z_T]: h is the first accepted sub-step of an
 * integration that starts by trying the whole flow step, and the end-state
 * increments are d_k = sinh(z_k) * (absTol + relTol*|y0_k|).  With mode
 * FCS_LEARNED_VERIFY, cells with h >= deltaT/2 get their proposed end state
 * verified: it is accepted only if one FP64 Rodas34 step over the whole
 * flow step from the initial state passes relTol / absTol and agrees with
 * the proposal to the same tolerance, and the cell then finishes with that
 * Rodas34 result.  Every other cell runs the plain integrator from its
 * original state and deltaTChem, making the same accept/reject decisions as
 * without the network.  FCS_LEARNED_SEED starts those cells from min(h,
 * deltaT) instead of deltaTChem.  FCS_E_ARG for other network widths,
 * FCS_E_UNSUPPORTED for OptSeulex.  The batch keeps the network pointer:
 * destroy the batch first.
 */
#define FCS_LEARNED_OFF     0
#define FCS_LEARNED_VERIFY  1
#define FCS_LEARNED_SEED    2
typedef struct fcs_mlp fcs_mlp;
int fcs_batch_set_learned(fcs_batch* batch, fcs_mlp* mlp, int32_t mode);
/* cells given a proposal in the last run, proposals verified (h >= deltaT/2)
 * and proposals accepted (cellStatus FCS_CELL_LEARNED) */
int fcs_batch_learned_stats(fcs_batch* batch, int64_t* predicted, int64_t* verified, int64_t* accepted);
/* first accepted Rosenbrock sub-step of each cell in the last run (0 when the
 * cell took none: skipped, finished by a learned proposal, or OptSeulex);
 * with deltaTChem set to deltaT on input, the training target h of the
 * network's first output */
int fcs_batch_first_steps(fcs_batch* batch, double* dxFirst /* [nCells] */);

/*
 * Kernel-level entry points over host arrays (batch of independent states,
 * species-major SoA, n = n_species + 1).  They run the same device code the
 * integrator runs and exist for parity checks and profiling.
 */
/* FastChemistryModel::derivatives (src/ChemistryModel/FastChemistryModel.C:194-273):
 * YT [n*nStates] (Y clipped >= 0 on output) -> dYTdt [n*nStates] */
int fcs_eval_derivatives(fcs_solver* solver, int64_t nStates, const double* p,
                         double* YT, double* dYTdt);
/* FastChemistryModel::jacobian (src/ChemistryModel/FastChemistryModel.C:489-572):
 * YT [n*nStates] -> f(y) [n*nStates] and Jy row-major [nStates][n][n] */
int fcs_eval_jacobian(fcs_solver* solver, int64_t nStates, const double* p,
                      double* YT, double* dYTdt, double* Jy);
/* One Rosenbrock attempt per state with step dx[s]: OptRodas34::Rodas34Solve
 * (src/ODE/OptRodas34/OptRodas34.C:291-879),
 * OptRosenbrock34::Rosenbrock34Solve (src/ODE/OptRosenbrock34/
 * OptRosenbrock34.C:281-677), OptRodas23::Rodas23Solve
 * (src/ODE/OptRodas23/OptRodas23.C:198-314) or
 * OptRosenbrock23::Rosenbrock23Solve (src/ODE/OptRosenbrock23/
 * OptRosenbrock23.C:204-294), by the solver's option.  YT [n*nStates] start
 * states (clipped >= 0 on output, as the Jacobian evaluation clips them),
 * YTnew [n*nStates] candidate states, err [nStates] error norms.
 * FCS_E_UNSUPPORTED for OptSeulex. */
int fcs_eval_attempt(fcs_solver* solver, int64_t nStates, const double* p,
                     const double* dx, double* YT, double* YTnew, double* err);
/* LUsolver::Block4LUDecompose + xSolve (src/LUsolver/LUsolver.C:3058-3171):
 * A [nSys][n][n] row-major, b [nSys][n] -> x in b.  n = n_species + 1. */
int fcs_lu_solve(fcs_solver* solver, int64_t nSys, const double* A, double* b);

/*
 * Learned seed of the integrator: a fully connected network with SiLU
 * hidden layers, evaluated and trained on the GPU with BF16 MFMA tiles and
 * FP32 accumulation (FP32 master weights, Adam).  No counterpart in the
 * reference: it proposes each cell's first chemical sub-step and its end
 * state, which the FP64 Rodas34 integrator verifies (fcs_solve_learned).
 * widths[0..nLayers] are the layer widths (input first).  Parameters are
 * FP32 in a flat padded layout (per layer W [pad16(in)][pad16(out)] then
 * b [pad16(out)], padding zero); inputs and outputs are normalised with
 * the statistics of fcs_mlp_set_norm (default mean 0, std 1).
 */
int fcs_mlp_create(int32_t nLayers, const int32_t* widths, uint64_t seed, fcs_mlp** out);
int fcs_mlp_destroy(fcs_mlp* mlp);
int64_t fcs_mlp_param_count(const fcs_mlp* mlp);
int fcs_mlp_get_params(fcs_mlp* mlp, float* params);
int fcs_mlp_set_params(fcs_mlp* mlp, const float* params);
int fcs_mlp_set_norm(fcs_mlp* mlp, const float* inMean, const float* inStd,
                     const float* outMean, const float* outStd);
/* X [n][widths[0]] raw inputs -> Y [n][widths[nLayers]] raw outputs */
int fcs_mlp_forward(fcs_mlp* mlp, int64_t n, const float* X, float* Y);
/* Mean-squared error of the normalised outputs against targets T (raw
 * units) and its gradient with respect to every parameter (grad may be NULL). */
int fcs_mlp_loss_grad(fcs_mlp* mlp, int64_t n, const float* X, const float* T,
                      float* loss, float* grad);
/* Adam on shuffled mini-batches; lossOut[e] = mean loss of epoch e (may be NULL). */
int fcs_mlp_train(fcs_mlp* mlp, int64_t n, const float* X, const float* T,
                  int32_t epochs, int32_t batch, float lr, uint64_t seed,
                  float* lossOut);

#ifdef __cplusplus
}
#endif

#endif /* FCS_H */
