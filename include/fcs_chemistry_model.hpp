// fcs_chemistry_model.hpp -- host-side C++ mirror of the FastChemistryModel
// plug-in interface over the C ABI of fcs.h (header-only, C++17).
//
// It is what an OpenFOAM-10 chemistry model class forwards to: it reads the
// plug-in's chemistryProperties keys, owns the mechanism and solver handles,
// and runs solve(deltaT), the local-time-step solve(deltaT field),
// calculate(), tc() and Qdot() on per-cell field pointers (one pointer per
// species, as OpenFOAM stores one field per species).  Errors the reference
// turns into FatalError are thrown as fcs::ChemistryError.
//
// Reference interfaces (paths relative to the reference tree):
//   chemistryType { solver ...; method FastChemistryModel; }  src/odeChemistrySolvers.C:66-72
//   jacobian exact|fast (default exact), Treact (default 0)    src/ChemistryModel/FastChemistryModel.C:54-66
//   <solver>Coeffs { absTol; relTol; maxSteps 10000; }         src/ODE/OptRodas34/OptRodas34.C:40-43,
//                                                              src/ODE/OptSeulex/OptSeulex.C:40-43
//   solve(scalar deltaT) / solve(scalarField deltaT)           src/ChemistryModel/FastChemistryModel.H:292,296
//   calculate(), tc(), Qdot()                                  src/ChemistryModel/FastChemistryModel.C:588-740
// initialChemicalTimeStep, maxChemicalTimeStep and `chemistry on|off` are
// read by OpenFOAM-10's basicChemistryModel (not in the reference tree);
// their defaults here are OpenFOAM's: great (1e15) and on.
#pragma once

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "fcs.h"

namespace fcs {

struct ChemistryError : std::runtime_error {
    int code;
    ChemistryError(const std::string& what, int rc) : std::runtime_error(what), code(rc) {}
};

// The options the plug-in reads from chemistryProperties.
struct ChemistryOptions {
    std::string solverName = "OptRodas34";
    fcs_solver_opts opts{FCS_SOLVER_RODAS34, FCS_JACOBIAN_EXACT, 0.0, 0.0, 10000, 0.0, 1.0e15};
    double initialChemicalTimeStep = 1.0e15;
    bool chemistry = true;
};

inline const char* const* solver_names()
{
    static const char* const names[] = {"OptRodas34", "OptRosenbrock34", "OptSeulex", "OptRodas23",
                                        "OptRosenbrock23", nullptr};
    return names;
}

// Reads the options through a dictionary view with
//   bool found(const std::string& key) and std::string lookup(const std::string& key),
// keys of sub-dictionaries written "sub.key" (chemistryType.solver,
// OptRodas34Coeffs.absTol).  Missing mandatory keys and unknown names throw,
// as the reference's lookups and NamedEnum reads do.
template <class Dict>
ChemistryOptions read_chemistry_options(const Dict& d)
{
    ChemistryOptions o;
    auto num = [&](const std::string& k) -> double {
        const std::string v = d.lookup(k);
        std::size_t used = 0;
        double x = 0.0;
        try {
            x = std::stod(v, &used);
        } catch (const std::exception&) {
            used = 0;
        }
        if (used != v.size()) throw ChemistryError("keyword " + k + " is not a number: " + v, FCS_E_ARG);
        return x;
    };
    o.solverName = d.lookup("chemistryType.solver");
    int kind = -1;
    for (int i = 0; solver_names()[i]; ++i)
        if (o.solverName == solver_names()[i]) kind = i;
    if (kind < 0) {
        std::string valid;
        for (int i = 0; solver_names()[i]; ++i) valid += std::string(" ") + solver_names()[i];
        throw ChemistryError("Unknown chemistry solver " + o.solverName + "; valid solvers:" + valid, FCS_E_ARG);
    }
    if (d.found("chemistryType.method") && d.lookup("chemistryType.method") != "FastChemistryModel")
        throw ChemistryError("chemistryType.method must be FastChemistryModel, not " +
                             d.lookup("chemistryType.method"), FCS_E_ARG);
    o.opts.solver = kind;
    if (d.found("jacobian")) {
        const std::string j = d.lookup("jacobian");
        if (j == "exact") o.opts.jacobian = FCS_JACOBIAN_EXACT;
        else if (j == "fast") o.opts.jacobian = FCS_JACOBIAN_FAST;
        else throw ChemistryError("jacobian must be exact or fast, not " + j, FCS_E_ARG);
    }
    o.opts.Treact = d.found("Treact") ? num("Treact") : 0.0;
    const std::string c = o.solverName + "Coeffs.";
    if (!d.found(c + "absTol") || !d.found(c + "relTol"))
        throw ChemistryError("keywords absTol and relTol are mandatory in " + o.solverName + "Coeffs", FCS_E_ARG);
    o.opts.absTol = num(c + "absTol");
    o.opts.relTol = num(c + "relTol");
    o.opts.maxSteps = d.found(c + "maxSteps") ? (int32_t)num(c + "maxSteps") : 10000;
    o.initialChemicalTimeStep = d.found("initialChemicalTimeStep") ? num("initialChemicalTimeStep") : 1.0e15;
    o.opts.deltaTChemMax = d.found("maxChemicalTimeStep") ? num("maxChemicalTimeStep") : 1.0e15;
    if (d.found("chemistry")) {
        const std::string v = d.lookup("chemistry");
        o.chemistry = !(v == "off" || v == "false" || v == "no" || v == "0");
    }
    return o;
}

// Per-cell fields of one solve: old-time T, p, rho and mass fractions
// (one pointer per species), the model's deltaTChem (in/out) and RR (out).
struct ChemistryFields {
    int64_t nCells = 0;
    const double* T0 = nullptr;
    const double* p0 = nullptr;
    const double* rho0 = nullptr;
    const double* const* Y0 = nullptr;
    double* deltaTChem = nullptr;
    double* const* RR = nullptr;
};

class ChemistryModel {
public:
    ChemistryModel(const fcs_mech_desc& mech, const ChemistryOptions& o) : opts_(o), ns_(mech.n_species)
    {
        check(fcs_mech_create(&mech, &mech_), "fcs_mech_create");
        const int rc = fcs_solver_create(mech_, &opts_.opts, &solver_);
        if (rc != FCS_OK) {
            fcs_mech_destroy(mech_);
            mech_ = nullptr;
            check(rc, "fcs_solver_create (" + opts_.solverName + ")");
        }
    }
    ~ChemistryModel()
    {
        if (solver_) fcs_solver_destroy(solver_);
        if (mech_) fcs_mech_destroy(mech_);
    }
    ChemistryModel(const ChemistryModel&) = delete;
    ChemistryModel& operator=(const ChemistryModel&) = delete;

    const ChemistryOptions& options() const { return opts_; }
    int nSpecie() const { return ns_; }

    // FastChemistryModel::solve(const scalar deltaT): returns
    // min(deltaTMin, 2 deltaT), or great with chemistry off.
    double solve(double deltaT, const ChemistryFields& f, int32_t* cellStatus = nullptr)
    {
        if (!opts_.chemistry) return 1.0e15;
        double dmin = 0.0;
        std::vector<int32_t> st;
        if (!cellStatus) st.resize((size_t)f.nCells), cellStatus = st.data();
        const int rc = fcs_solve(solver_, f.nCells, deltaT, f.T0, f.p0, f.rho0, f.Y0, f.deltaTChem, f.RR, &dmin,
                                 cellStatus);
        cells_failed(rc, cellStatus, f.nCells);
        return dmin;
    }

    // FastChemistryModel::solve(const scalarField& deltaT): local time steps.
    double solve(const double* deltaT, const ChemistryFields& f, int32_t* cellStatus = nullptr)
    {
        if (!opts_.chemistry) return 1.0e15;
        double dmin = 0.0;
        std::vector<int32_t> st;
        if (!cellStatus) st.resize((size_t)f.nCells), cellStatus = st.data();
        const int rc = fcs_solve_lts(solver_, f.nCells, deltaT, f.T0, f.p0, f.rho0, f.Y0, f.deltaTChem, f.RR,
                                     &dmin, cellStatus, nullptr, nullptr);
        cells_failed(rc, cellStatus, f.nCells);
        return dmin;
    }

    void calculate(int64_t nCells, const double* rho, const double* T, const double* p, const double* const* Y,
                   double* const* RR)
    {
        check(fcs_calculate(solver_, nCells, rho, T, p, Y, RR), "fcs_calculate");
    }
    void tc(int64_t nCells, const double* rho, const double* T, const double* p, const double* const* Y,
            double* tcOut)
    {
        check(fcs_tc(solver_, nCells, rho, T, p, Y, tcOut), "fcs_tc");
    }
    void Qdot(int64_t nCells, const double* const* RR, double* QdotOut)
    {
        check(fcs_qdot(solver_, nCells, RR, QdotOut), "fcs_qdot");
    }

private:
    static void check(int rc, const std::string& what)
    {
        if (rc != FCS_OK) throw ChemistryError(what + " failed with code " + std::to_string(rc), rc);
    }
    // maxSteps overflow and non-finite states: the reference calls FatalError
    // (src/ODE/OptRodas34/OptRodas34.C:193-195); the first failing cell is named.
    void cells_failed(int rc, const int32_t* st, int64_t n) const
    {
        if (rc == FCS_OK) return;
        if (rc == FCS_E_CELLS)
            for (int64_t c = 0; c < n; ++c) {
                if (st[c] & FCS_CELL_MAXSTEPS)
                    throw ChemistryError("Integration steps greater than maximum " +
                                         std::to_string(opts_.opts.maxSteps) + " in cell " + std::to_string(c),
                                         rc);
                if (st[c] & FCS_CELL_NONFINITE)
                    throw ChemistryError("non-finite state in cell " + std::to_string(c), rc);
            }
        check(rc, "fcs_solve");
    }

    ChemistryOptions opts_;
    int ns_ = 0;
    fcs_mech* mech_ = nullptr;
    fcs_solver* solver_ = nullptr;
};

}  // namespace fcs
