/*
 * fcs_oracle.h -- CPU restatement of the FastChemistryModel hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the gfx950
 * kernels: only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product path (the HIP library behind include/fcs.h)
 * never calls it.
 */
#ifndef FCS_ORACLE_H
#define FCS_ORACLE_H

#include <stdint.h>
#include "../include/fcs.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_mech orc_mech;

orc_mech* orc_mech_create(const fcs_mech_desc* d);
void orc_mech_destroy(orc_mech* m);
int orc_n_species(const orc_mech* m);

/* Per-species mass-specific thermo at T (no clamping here):
 * cp, ha [J/kg/K, J/kg], negGbyRT = -G/(RT), dBdT = d(-G/RT)/dT, dcpdT. */
void orc_thermo(const orc_mech* m, double T, double* cp, double* ha,
                double* negGbyRT, double* dBdT, double* dcpdT);

/* Molar production rates omega [kmol/m^3/s] from C [kmol/m^3] (T clamped
 * to [min Tlow, max Thigh] inside, as the reference does). */
void orc_calculate(const orc_mech* m, double rho, double T, double p, const double* Y, double* RR);
double orc_tc(const orc_mech* m, double rho, double T, double p, const double* Y);
double orc_hf(const orc_mech* m, int i);
double orc_qdot(const orc_mech* m, const double* RR);
void orc_rates(const orc_mech* m, double p, double T, const double* C,
               double* omega);

/* FastChemistryModel::derivatives: YT[n] (Y clipped in place) -> dYTdt[n]. */
void orc_derivatives(const orc_mech* m, double p, double* YT, double* dYTdt);

/* FastChemistryModel::jacobian: YT[n] (clipped in place) -> f(YT) in
 * dYTdt[n] and Jy[n*n] row-major. jac = FCS_JACOBIAN_EXACT|FAST.
 * Jc (optional, may be NULL) receives the concentration Jacobian. */
void orc_jacobian(const orc_mech* m, double p, double* YT, double* dYTdt,
                  double* Jy, int jac, double* Jc);

/* LUsolver: block-4 factorisation with pivoting restricted to each 4-row
 * diagonal block; piv[n], invD[n]. */
void orc_lu_factor(int n, double* a, int32_t* piv, double* invD);
void orc_lu_solve(int n, const double* a, const int32_t* piv,
                  const double* invD, double* b);

/* Test hook: 1 = Rosenbrock stages with df/dt = 0 (textbook form). */
void orc_set_textbook_rosenbrock(int on);

/* Step log of the last single-cell integration (for controller checks). */
typedef struct orc_step_rec {
    double x, dx, err;
    int32_t accepted;
    int32_t order;     /* Seulex column index k at exit, else 0 */
} orc_step_rec;

/* Integrate one cell: y[n] in/out (Y, T), p, over deltaT from dxTry (in/out).
 * Returns 0, or FCS_CELL_MAXSTEPS / FCS_CELL_NONFINITE. nAttempts out. */
int orc_integrate_cell(const orc_mech* m, const fcs_solver_opts* o, double p,
                       double* y, double deltaT, double* dxTry,
                       int32_t* nAttempts, orc_step_rec* log, int32_t logCap,
                       int32_t* logLen);

/* One Rosenbrock attempt (Rodas34Solve / Rosenbrock34Solve) from y[n]
 * (clipped in place) with step dx: candidate state to yNew[n], returns the
 * error norm (NaN for Seulex). */
double orc_rosenbrock_attempt(const orc_mech* m, const fcs_solver_opts* o, double p,
                              double dx, double* y, double* yNew);

/* FastChemistryModel::solve(deltaT) over a batch, species-major arrays
 * Y0/RR [ns*nCells]; nThreads > 1 uses OpenMP with a dynamic schedule. */
int orc_solve(const orc_mech* m, const fcs_solver_opts* o, int64_t nCells,
              double deltaT, const double* T0, const double* p0,
              const double* rho0, const double* Y0, double* deltaTChem,
              double* RR, double* deltaTMin, int32_t* cellStatus,
              double* Tend, int32_t* nAttempts, int nThreads);
/* local-time-step twin: src/ChemistryModel/FastChemistryModel_localEulerSolve.H:3-101 */
int orc_solve_lts(const orc_mech* m, const fcs_solver_opts* o, int64_t nCells,
              const double* deltaT, const double* T0, const double* p0,
              const double* rho0, const double* Y0, double* deltaTChem,
              double* RR, double* deltaTMin, int32_t* cellStatus,
              double* Tend, int32_t* nAttempts, int nThreads);

#ifdef __cplusplus
}
#endif
#endif
