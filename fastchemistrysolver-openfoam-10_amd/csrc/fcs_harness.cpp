// fcs_harness.cpp -- host harness standing in for OpenFOAM around the
// ChemistryModel adapter (include/fcs_chemistry_model.hpp).
//
// It mimics what an OpenFOAM-10 solver does with the plug-in: one field per
// species, old-time T / p / rho fields, the model's deltaTChem field
// initialised to initialChemicalTimeStep, and a time loop that calls
// solve(deltaT), then advances the species with the returned reaction rates
// (Y += deltaT RR / rho0, the chemistry part of the species equation), and
// finally evaluates calculate(), tc() and Qdot() at the end state.
//
// Usage:
//   fcs_harness CASE_DIR --options-only
//       prints the options read from CASE_DIR/options.txt (no device use)
//   fcs_harness CASE_DIR DELTA_T N_STEPS OUT_FILE [--lts]
//       CASE_DIR/options.txt  "key value" lines (sub-dictionary keys as sub.key)
//       CASE_DIR/mech.bin     mechanism arrays (see read_mech)
//       CASE_DIR/fields.bin   int64 nCells, then T, p, rho [nCells], Y [ns][nCells]
//       --lts                 solve(scalarField deltaT) with deltaT * (1 + c mod 3) per cell
//   OUT_FILE: per step deltaTMin [N_STEPS]; then Y [ns][nCells], deltaTChem,
//   RR of the last step [ns][nCells], calculate() RR [ns][nCells], tc, Qdot.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "fcs_chemistry_model.hpp"

namespace {

struct FlatDict {
    std::map<std::string, std::string> kv;
    bool found(const std::string& k) const { return kv.count(k) != 0; }
    std::string lookup(const std::string& k) const
    {
        auto it = kv.find(k);
        if (it == kv.end()) throw fcs::ChemistryError("keyword " + k + " is undefined", FCS_E_ARG);
        return it->second;
    }
};

FlatDict read_dict(const std::string& path)
{
    FlatDict d;
    std::ifstream f(path);
    if (!f) throw std::runtime_error("cannot open " + path);
    std::string line;
    while (std::getline(f, line)) {
        std::istringstream s(line);
        std::string k, v;
        if (s >> k >> v) d.kv[k] = v;
    }
    return d;
}

struct Reader {
    std::ifstream f;
    explicit Reader(const std::string& p) : f(p, std::ios::binary)
    {
        if (!f) throw std::runtime_error("cannot open " + p);
    }
    template <class T>
    std::vector<T> get(size_t n)
    {
        std::vector<T> v(n);
        if (n && !f.read(reinterpret_cast<char*>(v.data()), (std::streamsize)(n * sizeof(T))))
            throw std::runtime_error("short read");
        return v;
    }
};

// Mechanism arrays in fcs_mech_desc order: int32 ns, nr, pmax, slot stride
// ss; doubles W, T_low, T_high, T_common [ns], coeffs_high, coeffs_low
// [7 ns]; int32 rxn_type, rxn_reversible, rxn_n_lhs, rxn_n_rhs [nr],
// rxn_lhs, rxn_rhs [ss nr]; doubles A, beta, Ta, A0, beta0, Ta0 [nr], efficiencies [nr ns],
// troe [4 nr]; int32 plog_n [nr]; doubles plog_p, plog_A, plog_beta,
// plog_Ta [nr pmax], sri [5 nr], A_rev, beta_rev, Ta_rev [nr].
struct MechData {
    std::vector<double> d[22];
    std::vector<int32_t> i[7];
    fcs_mech_desc desc{};
};

void read_mech(const std::string& path, MechData& m)
{
    Reader r(path);
    const auto h = r.get<int32_t>(4);
    const size_t ns = h[0], nr = h[1], pm = h[2], ss = h[3];
    int k = 0;
    for (int q = 0; q < 4; ++q) m.d[k++] = r.get<double>(ns);
    m.d[k++] = r.get<double>(7 * ns);
    m.d[k++] = r.get<double>(7 * ns);
    for (int q = 0; q < 4; ++q) m.i[q] = r.get<int32_t>(nr);
    m.i[4] = r.get<int32_t>(ss * nr);
    m.i[5] = r.get<int32_t>(ss * nr);
    for (int q = 0; q < 6; ++q) m.d[k++] = r.get<double>(nr);
    m.d[k++] = r.get<double>(nr * ns);
    m.d[k++] = r.get<double>(4 * nr);
    m.i[6] = r.get<int32_t>(nr);
    for (int q = 0; q < 4; ++q) m.d[k++] = r.get<double>(nr * pm);
    m.d[k++] = r.get<double>(5 * nr);
    for (int q = 0; q < 3; ++q) m.d[k++] = r.get<double>(nr);
    fcs_mech_desc& s = m.desc;
    s.n_species = (int32_t)ns;
    s.n_reactions = (int32_t)nr;
    s.slot_stride = (int32_t)ss;
    s.W = m.d[0].data(); s.T_low = m.d[1].data(); s.T_high = m.d[2].data(); s.T_common = m.d[3].data();
    s.coeffs_high = m.d[4].data(); s.coeffs_low = m.d[5].data();
    s.rxn_type = m.i[0].data(); s.rxn_reversible = m.i[1].data(); s.rxn_n_lhs = m.i[2].data();
    s.rxn_n_rhs = m.i[3].data(); s.rxn_lhs = m.i[4].data(); s.rxn_rhs = m.i[5].data();
    s.A = m.d[6].data(); s.beta = m.d[7].data(); s.Ta = m.d[8].data();
    s.A0 = m.d[9].data(); s.beta0 = m.d[10].data(); s.Ta0 = m.d[11].data();
    s.efficiencies = m.d[12].data(); s.troe = m.d[13].data();
    s.plog_max_points = (int32_t)pm; s.plog_n = m.i[6].data();
    s.plog_p = m.d[14].data(); s.plog_A = m.d[15].data(); s.plog_beta = m.d[16].data(); s.plog_Ta = m.d[17].data();
    s.sri = m.d[18].data(); s.A_rev = m.d[19].data(); s.beta_rev = m.d[20].data(); s.Ta_rev = m.d[21].data();
}

std::vector<double*> rows(std::vector<double>& a, size_t ns, size_t nc)
{
    std::vector<double*> p(ns);
    for (size_t i = 0; i < ns; ++i) p[i] = a.data() + i * nc;
    return p;
}

}  // namespace

int main(int argc, char** argv)
{
    try {
        if (argc < 3) {
            std::cerr << "usage: fcs_harness CASE_DIR (--options-only | DELTA_T N_STEPS OUT_FILE [--lts])\n";
            return 2;
        }
        const std::string dir = argv[1];
        const fcs::ChemistryOptions o = fcs::read_chemistry_options(read_dict(dir + "/options.txt"));
        if (std::strcmp(argv[2], "--options-only") == 0) {
            std::printf("solver=%s\nsolver_id=%d\njacobian=%d\nabsTol=%.17g\nrelTol=%.17g\nmaxSteps=%d\n"
                        "Treact=%.17g\ndeltaTChemMax=%.17g\ninitialChemicalTimeStep=%.17g\nchemistry=%d\n",
                        o.solverName.c_str(), o.opts.solver, o.opts.jacobian, o.opts.absTol, o.opts.relTol,
                        o.opts.maxSteps, o.opts.Treact, o.opts.deltaTChemMax, o.initialChemicalTimeStep,
                        o.chemistry ? 1 : 0);
            return 0;
        }
        if (argc < 5) throw std::runtime_error("missing DELTA_T N_STEPS OUT_FILE");
        const double deltaT = std::stod(argv[2]);
        const int nSteps = std::stoi(argv[3]);
        const bool lts = argc > 5 && std::strcmp(argv[5], "--lts") == 0;
        MechData md;
        read_mech(dir + "/mech.bin", md);
        const size_t ns = (size_t)md.desc.n_species;
        Reader fr(dir + "/fields.bin");
        const size_t nc = (size_t)fr.get<int64_t>(1)[0];
        std::vector<double> T = fr.get<double>(nc), p = fr.get<double>(nc), rho = fr.get<double>(nc);
        std::vector<double> Y = fr.get<double>(ns * nc);
        fcs::ChemistryModel model(md.desc, o);
        std::vector<double> dtc(nc, o.initialChemicalTimeStep), RR(ns * nc, 0.0), dtField(nc);
        for (size_t c = 0; c < nc; ++c) dtField[c] = deltaT * (1.0 + (double)(c % 3));
        std::vector<double> Y0 = Y;
        std::vector<double*> Yp = rows(Y0, ns, nc), RRp = rows(RR, ns, nc);
        std::vector<const double*> Y0c(Yp.begin(), Yp.end());
        std::vector<double> dmins;
        for (int s = 0; s < nSteps; ++s) {
            Y0 = Y;   // old-time mass fractions of this step
            fcs::ChemistryFields f;
            f.nCells = (int64_t)nc;
            f.T0 = T.data();
            f.p0 = p.data();
            f.rho0 = rho.data();
            f.Y0 = Y0c.data();
            f.deltaTChem = dtc.data();
            f.RR = RRp.data();
            dmins.push_back(lts ? model.solve(dtField.data(), f) : model.solve(deltaT, f));
            for (size_t i = 0; i < ns; ++i)
                for (size_t c = 0; c < nc; ++c)
                    Y[i * nc + c] = Y0[i * nc + c] + (lts ? dtField[c] : deltaT) * RR[i * nc + c] / rho[c];
        }
        std::vector<double> RRc(ns * nc), tc(nc), Qd(nc);
        std::vector<double*> RRcp = rows(RRc, ns, nc);
        std::vector<double*> Yq = rows(Y, ns, nc);
        std::vector<const double*> Yc(Yq.begin(), Yq.end());
        std::vector<const double*> RRconst(RRp.begin(), RRp.end());
        model.calculate((int64_t)nc, rho.data(), T.data(), p.data(), Yc.data(), RRcp.data());
        model.tc((int64_t)nc, rho.data(), T.data(), p.data(), Yc.data(), tc.data());
        model.Qdot((int64_t)nc, RRconst.data(), Qd.data());
        std::ofstream out(argv[4], std::ios::binary);
        auto put = [&](const std::vector<double>& v) {
            out.write(reinterpret_cast<const char*>(v.data()), (std::streamsize)(v.size() * sizeof(double)));
        };
        put(dmins);
        put(Y);
        put(dtc);
        put(RR);
        put(RRc);
        put(tc);
        put(Qd);
        if (!out) throw std::runtime_error("cannot write output");
        std::printf("harness: %d steps over %zu cells, solver %s\n", nSteps, nc, o.solverName.c_str());
        return 0;
    } catch (const fcs::ChemistryError& e) {
        std::fprintf(stderr, "FOAM FATAL ERROR (chemistry): %s\n", e.what());
        return 3;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "fcs_harness: %s\n", e.what());
        return 1;
    }
}
