/*---------------------------------------------------------------------------*\
  FastChemistryModelGPU -- see FastChemistryModelGPU.H.

  Not compiled in this repository (OpenFOAM-10 is not available to its
  build); the adapter core it calls, fcs::ChemistryModel, is compiled and
  tested through csrc/fcs_harness.cpp.

  Reference counterparts: constructor src/ChemistryModel/FastChemistryModel.C:
  45-161, solve FastChemistryModel_transientSolve.H:3-104 and
  FastChemistryModel_localEulerSolve.H:3-101, tc / Qdot / calculate
  FastChemistryModel.C:588-740, reaction reading src/Reaction/OptReaction.C:
  348-1175.
\*---------------------------------------------------------------------------*/
#include "FastChemistryModelGPU.H"
#include "fvm.H"

#include <map>
#include <sstream>

// * * * * * * * * * * * * * * mechanism arrays  * * * * * * * * * * * * * //

template<class ThermoType>
struct Foam::FastChemistryModelGPU<ThermoType>::MechArrays
{
    std::vector<double> W, Tlow, Thigh, Tcommon, hi, lo, A, beta, Ta, A0, beta0, Ta0, eff, troe, sri,
        Arev, betarev, Tarev, plogP, plogA, plogBeta, plogTa;
    std::vector<int32_t> type, rev, nl, nr, lhs, rhs, plogN;
    int32_t pmax = 1;
    int32_t stride = FCS_MAX_STOICH;   // slots per reaction in lhs / rhs
    fcs_mech_desc desc{};
};


namespace
{

// dictionary view with dotted sub-dictionary keys for read_chemistry_options
struct FoamDictView
{
    const Foam::dictionary& d;

    const Foam::dictionary* owner(const std::string& key, std::string& leaf) const
    {
        const auto dot = key.find('.');
        if (dot == std::string::npos) { leaf = key; return &d; }
        const Foam::word sub(key.substr(0, dot));
        leaf = key.substr(dot + 1);
        return d.found(sub) ? &d.subDict(sub) : nullptr;
    }
    bool found(const std::string& key) const
    {
        std::string leaf;
        const Foam::dictionary* o = owner(key, leaf);
        return o && o->found(Foam::word(leaf));
    }
    std::string lookup(const std::string& key) const
    {
        std::string leaf;
        const Foam::dictionary* o = owner(key, leaf);
        if (!o || !o->found(Foam::word(leaf)))
        {
            FatalIOErrorInFunction(d) << "keyword " << key << " is undefined" << Foam::exit(Foam::FatalIOError);
        }
        Foam::ITstream& is = o->lookup(Foam::word(leaf));
        Foam::token t(is);
        if (t.isWord()) return t.wordToken();
        if (t.isNumber()) { std::ostringstream s; s.precision(17); s << t.number(); return s.str(); }
        if (t.isString()) return t.stringToken();
        return std::string();
    }
};

// one side of "A + 2B = C": species indices, integer coefficients expanded
// into repeated slots (OptReaction.C:99-126)
std::vector<int32_t> side(const std::string& text, const Foam::hashedWordList& species)
{
    std::vector<int32_t> out;
    std::istringstream s(text);
    std::string term;
    while (std::getline(s, term, '+'))
    {
        const auto b = term.find_first_not_of(' '), e = term.find_last_not_of(' ');
        if (b == std::string::npos) continue;
        term = term.substr(b, e - b + 1);
        std::size_t k = 0;
        while (k < term.size() && std::isdigit((unsigned char)term[k])) ++k;
        const int coef = k ? std::stoi(term.substr(0, k)) : 1;
        const Foam::word name(term.substr(k));
        for (int q = 0; q < coef; ++q) out.push_back(species[name]);
    }
    return out;
}

} // namespace


template<class ThermoType>
void Foam::FastChemistryModelGPU<ThermoType>::readMechanism
(
    const multiComponentMixture<ThermoType>& mixture
)
{
    MechArrays& m = *mechArrays_;
    const PtrList<ThermoType>& th = mixture.specieThermos();
    forAll(th, i)
    {
        // janafThermo keeps mass-based coefficients; the library takes the
        // dictionary's molar NASA-7 form
        const scalar R = th[i].R();
        m.W.push_back(th[i].W());
        m.Tlow.push_back(th[i].Tlow());
        m.Thigh.push_back(th[i].Thigh());
        m.Tcommon.push_back(th[i].Tcommon());
        for (int k = 0; k < 7; ++k) m.hi.push_back(th[i].highCpCoeffs()[k]/R);
        for (int k = 0; k < 7; ++k) m.lo.push_back(th[i].lowCpCoeffs()[k]/R);
    }
    const hashedWordList& species = this->thermo().composition().species();
    const dictionary& reactions = this->subDict("reactions");
    const label ns = species.size();
    std::vector<std::vector<double>> plog;
    std::vector<std::vector<int32_t>> lhsSides, rhsSides;
    forAllConstIter(dictionary, reactions, iter)
    {
        const dictionary& r = reactions.subDict(iter().keyword());
        const word type(r.lookup("type"));
        const bool noneq = type.find("nonEquilibriumReversible") == 0;
        const bool irrev = type.find("irreversible") == 0;
        const word form =
            noneq ? word(type.substr(24)) : irrev ? word(type.substr(12)) : word(type.substr(10));
        int32_t t = -1;
        if (form == "Arrhenius") t = FCS_RXN_ARRHENIUS;
        else if (form == "ThirdBodyArrhenius") t = FCS_RXN_THIRD_BODY;
        else if (form == "ArrheniusLindemannFallOff") t = FCS_RXN_LINDEMANN;
        else if (form == "ArrheniusTroeFallOff") t = FCS_RXN_TROE;
        else if (form == "ArrheniusSRIFallOff") t = FCS_RXN_SRI;
        else if (form == "ArrheniusPLOG" || form == "Plog") t = FCS_RXN_PLOG;
        else
        {
            FatalErrorInFunction << "Unsupported reaction type " << type << exit(FatalError);
        }
        const string eq(r.lookup("reaction"));
        const auto at = eq.find('=');
        const std::vector<int32_t> L = side(eq.substr(0, at), species), Rs = side(eq.substr(at + 1), species);
        m.type.push_back(t);
        m.rev.push_back(noneq ? 2 : irrev ? 0 : 1);
        m.nl.push_back(L.size());
        m.nr.push_back(Rs.size());
        lhsSides.push_back(L);
        rhsSides.push_back(Rs);
        auto arr = [](const dictionary& a, std::vector<double>& A, std::vector<double>& b, std::vector<double>& T)
        {
            A.push_back(a.lookup<scalar>("A"));
            b.push_back(a.lookup<scalar>("beta"));
            T.push_back(a.lookup<scalar>("Ta"));
        };
        const bool falloff = t == FCS_RXN_LINDEMANN || t == FCS_RXN_TROE || t == FCS_RXN_SRI;
        if (noneq) arr(r.subDict("forward"), m.A, m.beta, m.Ta);
        else if (falloff) arr(r.subDict("kInf"), m.A, m.beta, m.Ta);
        else if (t != FCS_RXN_PLOG) arr(r, m.A, m.beta, m.Ta);
        else { m.A.push_back(0); m.beta.push_back(0); m.Ta.push_back(0); }
        if (falloff) arr(r.subDict("k0"), m.A0, m.beta0, m.Ta0);
        else { m.A0.push_back(0); m.beta0.push_back(0); m.Ta0.push_back(0); }
        if (noneq) arr(r.subDict("reverse"), m.Arev, m.betarev, m.Tarev);
        else { m.Arev.push_back(0); m.betarev.push_back(0); m.Tarev.push_back(0); }
        std::vector<double> e(ns, 1.0);
        if (t == FCS_RXN_THIRD_BODY || falloff)
        {
            const List<Tuple2<word, scalar>> c
            (
                noneq ? r.subDict("forward").lookup("coeffs")
              : falloff ? r.subDict("thirdBodyEfficiencies").lookup("coeffs")
              : r.lookup("coeffs")
            );
            forAll(c, k) e[species[c[k].first()]] = c[k].second();
        }
        m.eff.insert(m.eff.end(), e.begin(), e.end());
        const dictionary* F = r.found("F") ? &r.subDict("F") : nullptr;
        for (const char* k : {"alpha", "Tsss", "Ts", "Tss"})
            m.troe.push_back(t == FCS_RXN_TROE ? F->lookup<scalar>(k) : 0.0);
        for (const char* k : {"a", "b", "c", "d", "e"})
            m.sri.push_back(t == FCS_RXN_SRI ? F->lookup<scalar>(k) : 0.0);
        std::vector<double> row;
        if (t == FCS_RXN_PLOG)
        {
            const List<FixedList<scalar, 4>> d
            (
                r.found("ArrheniusData") ? r.lookup("ArrheniusData") : r.lookup("coeffs")
            );
            forAll(d, k) for (int q = 0; q < 4; ++q) row.push_back(d[k][q]);
            m.pmax = std::max<int32_t>(m.pmax, d.size());
        }
        plog.push_back(row);
    }
    const size_t nr = m.type.size();
    // slots per reaction: 3, or the longest product side (irreversible and
    // nonEquilibrium reactions may carry up to FCS_MAX_SLOTS products)
    m.stride = FCS_MAX_STOICH;
    for (const auto& Rs : rhsSides) m.stride = std::max<int32_t>(m.stride, Rs.size());
    m.lhs.assign(nr*m.stride, 0);
    m.rhs.assign(nr*m.stride, 0);
    for (size_t k = 0; k < nr; ++k)
    {
        for (size_t q = 0; q < lhsSides[k].size() && q < size_t(m.stride); ++q) m.lhs[k*m.stride + q] = lhsSides[k][q];
        for (size_t q = 0; q < rhsSides[k].size(); ++q) m.rhs[k*m.stride + q] = rhsSides[k][q];
    }
    m.plogP.assign(nr*m.pmax, 0); m.plogA.assign(nr*m.pmax, 0);
    m.plogBeta.assign(nr*m.pmax, 0); m.plogTa.assign(nr*m.pmax, 0);
    m.plogN.assign(nr, 0);
    for (size_t k = 0; k < nr; ++k)
    {
        const size_t np = plog[k].size()/4;
        m.plogN[k] = np;
        for (size_t q = 0; q < np; ++q)
        {
            m.plogP[k*m.pmax + q] = plog[k][4*q];
            m.plogA[k*m.pmax + q] = plog[k][4*q + 1];
            m.plogBeta[k*m.pmax + q] = plog[k][4*q + 2];
            m.plogTa[k*m.pmax + q] = plog[k][4*q + 3];
        }
    }
    fcs_mech_desc& d = m.desc;
    d.n_species = ns; d.n_reactions = nr;
    d.W = m.W.data(); d.T_low = m.Tlow.data(); d.T_high = m.Thigh.data(); d.T_common = m.Tcommon.data();
    d.coeffs_high = m.hi.data(); d.coeffs_low = m.lo.data();
    d.rxn_type = m.type.data(); d.rxn_reversible = m.rev.data(); d.rxn_n_lhs = m.nl.data();
    d.rxn_n_rhs = m.nr.data(); d.rxn_lhs = m.lhs.data(); d.rxn_rhs = m.rhs.data();
    d.slot_stride = m.stride;
    d.A = m.A.data(); d.beta = m.beta.data(); d.Ta = m.Ta.data();
    d.A0 = m.A0.data(); d.beta0 = m.beta0.data(); d.Ta0 = m.Ta0.data();
    d.efficiencies = m.eff.data(); d.troe = m.troe.data();
    d.plog_max_points = m.pmax; d.plog_n = m.plogN.data(); d.plog_p = m.plogP.data();
    d.plog_A = m.plogA.data(); d.plog_beta = m.plogBeta.data(); d.plog_Ta = m.plogTa.data();
    d.sri = m.sri.data(); d.A_rev = m.Arev.data(); d.beta_rev = m.betarev.data(); d.Ta_rev = m.Tarev.data();
}


// * * * * * * * * * * * * * * * * Constructors  * * * * * * * * * * * * * * //

template<class ThermoType>
Foam::FastChemistryModelGPU<ThermoType>::FastChemistryModelGPU
(
    const fluidReactionThermo& thermo
)
:
    basicChemistryModel(thermo),
    Yvf_(this->thermo().composition().Y()),
    nSpecie_(Yvf_.size()),
    RR_(nSpecie_),
    mechArrays_(new MechArrays)
{
    forAll(RR_, fieldi)
    {
        RR_.set
        (
            fieldi,
            new volScalarField::Internal
            (
                IOobject
                (
                    "RR." + Yvf_[fieldi].name(),
                    this->mesh().time().timeName(),
                    this->mesh(),
                    IOobject::NO_READ,
                    IOobject::NO_WRITE
                ),
                thermo.T().mesh(),
                dimensionedScalar(dimMass/dimVolume/dimTime, 0)
            )
        );
    }
    readMechanism
    (
        refCast<const multiComponentMixture<ThermoType>>(this->thermo())
    );
    try
    {
        model_.reset
        (
            new fcs::ChemistryModel
            (
                mechArrays_->desc,
                fcs::read_chemistry_options(FoamDictView{*this})
            )
        );
    }
    catch (const fcs::ChemistryError& e)
    {
        FatalErrorInFunction << e.what() << exit(FatalError);
    }
}


template<class ThermoType>
Foam::FastChemistryModelGPU<ThermoType>::~FastChemistryModelGPU()
{}


// * * * * * * * * * * * * * * * Member Functions  * * * * * * * * * * * * * //

template<class ThermoType>
Foam::label Foam::FastChemistryModelGPU<ThermoType>::nReaction() const
{
    return mechArrays_->desc.n_reactions;
}


template<class ThermoType>
Foam::fcs::ChemistryFields Foam::FastChemistryModelGPU<ThermoType>::fields
(
    const scalarField& T0,
    const scalarField& p0,
    const scalarField& rho0,
    std::vector<const double*>& Y0,
    std::vector<double*>& RR
)
{
    fcs::ChemistryFields f;
    f.nCells = T0.size();
    f.T0 = T0.cdata();
    f.p0 = p0.cdata();
    f.rho0 = rho0.cdata();
    Y0.resize(nSpecie_);
    RR.resize(nSpecie_);
    for (label i = 0; i < nSpecie_; i++)
    {
        Y0[i] = Yvf_[i].oldTime().primitiveField().cdata();
        RR[i] = RR_[i].data();
    }
    f.Y0 = Y0.data();
    f.deltaTChem = this->deltaTChem_.data();
    f.RR = RR.data();
    return f;
}


template<class ThermoType>
template<class DeltaTType>
Foam::scalar Foam::FastChemistryModelGPU<ThermoType>::solveImpl(const DeltaTType& deltaT)
{
    basicChemistryModel::correct();
    if (!this->chemistry_)
    {
        return great;
    }
    tmp<volScalarField> trho0(this->thermo().rho0());
    std::vector<const double*> Y0;
    std::vector<double*> RR;
    const fcs::ChemistryFields f = fields
    (
        this->thermo().T().oldTime().primitiveField(),
        this->thermo().p().oldTime().primitiveField(),
        trho0().primitiveField(),
        Y0,
        RR
    );
    try
    {
        return model_->solve(deltaT, f);
    }
    catch (const fcs::ChemistryError& e)
    {
        FatalErrorInFunction << e.what() << exit(FatalError);
    }
    return great;
}


template<class ThermoType>
Foam::scalar Foam::FastChemistryModelGPU<ThermoType>::solve(const scalar deltaT)
{
    return solveImpl(double(deltaT));
}


template<class ThermoType>
Foam::scalar Foam::FastChemistryModelGPU<ThermoType>::solve(const scalarField& deltaT)
{
    return solveImpl(deltaT.cdata());
}


template<class ThermoType>
void Foam::FastChemistryModelGPU<ThermoType>::calculate()
{
    if (!this->chemistry_) return;
    tmp<volScalarField> trho(this->thermo().rho());
    std::vector<const double*> Y(nSpecie_);
    std::vector<double*> RR(nSpecie_);
    for (label i = 0; i < nSpecie_; i++)
    {
        Y[i] = Yvf_[i].primitiveField().cdata();
        RR[i] = RR_[i].data();
    }
    model_->calculate
    (
        trho().size(), trho().primitiveField().cdata(), this->thermo().T().primitiveField().cdata(),
        this->thermo().p().primitiveField().cdata(), Y.data(), RR.data()
    );
}


template<class ThermoType>
Foam::tmp<Foam::volScalarField> Foam::FastChemistryModelGPU<ThermoType>::tc() const
{
    tmp<volScalarField> ttc
    (
        volScalarField::New
        (
            "tc", this->mesh(), dimensionedScalar(dimTime, small),
            extrapolatedCalculatedFvPatchScalarField::typeName
        )
    );
    if (this->chemistry_)
    {
        tmp<volScalarField> trho(this->thermo().rho());
        std::vector<const double*> Y(nSpecie_);
        for (label i = 0; i < nSpecie_; i++) Y[i] = Yvf_[i].primitiveField().cdata();
        const_cast<fcs::ChemistryModel&>(*model_).tc
        (
            trho().size(), trho().primitiveField().cdata(), this->thermo().T().primitiveField().cdata(),
            this->thermo().p().primitiveField().cdata(), Y.data(), ttc.ref().primitiveFieldRef().data()
        );
    }
    ttc.ref().correctBoundaryConditions();
    return ttc;
}


template<class ThermoType>
Foam::tmp<Foam::volScalarField> Foam::FastChemistryModelGPU<ThermoType>::Qdot() const
{
    tmp<volScalarField> tQdot
    (
        volScalarField::New("Qdot", this->mesh_, dimensionedScalar(dimEnergy/dimVolume/dimTime, 0))
    );
    if (this->chemistry_)
    {
        std::vector<const double*> RR(nSpecie_);
        for (label i = 0; i < nSpecie_; i++) RR[i] = RR_[i].cdata();
        const_cast<fcs::ChemistryModel&>(*model_).Qdot
        (
            tQdot().size(), RR.data(), tQdot.ref().primitiveFieldRef().data()
        );
    }
    return tQdot;
}


template<class ThermoType>
Foam::tmp<Foam::DimensionedField<Foam::scalar, Foam::volMesh>>
Foam::FastChemistryModelGPU<ThermoType>::calculateRR(const label, const label) const
{
    // not supported by the reference either (FastChemistryModel.C:678-697)
    FatalErrorInFunction << "This function is not supported and should not be used" << abort(FatalError);
    return tmp<volScalarField::Internal>(nullptr);
}
