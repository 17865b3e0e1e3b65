/*---------------------------------------------------------------------------*\
  FastChemistryModelGPU -- OpenFOAM-10 chemistry model that forwards the
  FastChemistryModel plug-in's interface to the gfx950 library (fcs.h)
  through fcs::ChemistryModel (include/fcs_chemistry_model.hpp).

  Source of an optional wmake target: OpenFOAM is not part of this
  repository's build, which exercises the same adapter through the host
  harness csrc/fcs_harness.cpp instead.

  Mirrors src/ChemistryModel/FastChemistryModel.H of the reference:
  TypeName "FastChemistryModel", solve(scalar), solve(scalarField),
  calculate(), tc(), Qdot(), RR(i), nSpecie(), nReaction().
\*---------------------------------------------------------------------------*/
#ifndef FastChemistryModelGPU_H
#define FastChemistryModelGPU_H

#include "basicChemistryModel.H"
#include "multiComponentMixture.H"
#include "fcs_chemistry_model.hpp"

#include <memory>
#include <vector>

namespace Foam
{

template<class ThermoType>
class FastChemistryModelGPU
:
    public basicChemistryModel
{
    // Private data

        //- Mass fractions of the mixture
        PtrList<volScalarField>& Yvf_;

        //- Number of species
        label nSpecie_;

        //- Reaction rates [kg/m^3/s], one field per species
        PtrList<volScalarField::Internal> RR_;

        //- Mechanism arrays handed to the library (fcs_mech_desc storage)
        struct MechArrays;
        std::unique_ptr<MechArrays> mechArrays_;

        //- The device-side model
        std::unique_ptr<fcs::ChemistryModel> model_;

        //- Reads the mechanism from the reactions / thermo dictionaries
        //  into flat arrays (the forms OptReaction::readInfo accepts)
        void readMechanism(const multiComponentMixture<ThermoType>& mixture);

        //- Per-cell field pointers of one solve
        fcs::ChemistryFields fields
        (
            const scalarField& T0,
            const scalarField& p0,
            const scalarField& rho0,
            std::vector<const double*>& Y0,
            std::vector<double*>& RR
        );

        template<class DeltaTType>
        scalar solveImpl(const DeltaTType& deltaT);

public:

    TypeName("FastChemistryModel");

    FastChemistryModelGPU(const fluidReactionThermo& thermo);
    FastChemistryModelGPU(const FastChemistryModelGPU&) = delete;
    virtual ~FastChemistryModelGPU();

    virtual label nSpecie() const { return nSpecie_; }
    virtual label nReaction() const;

    virtual void calculate();
    virtual scalar solve(const scalar deltaT);
    virtual scalar solve(const scalarField& deltaT);
    virtual tmp<volScalarField> tc() const;
    virtual tmp<volScalarField> Qdot() const;

    virtual const volScalarField::Internal& RR(const label i) const { return RR_[i]; }
    virtual volScalarField::Internal& RR(const label i) { return RR_[i]; }
    virtual tmp<volScalarField::Internal> calculateRR(const label reactionI, const label speciei) const;

    void operator=(const FastChemistryModelGPU&) = delete;
};

} // End namespace Foam

#ifdef NoRepository
    #include "FastChemistryModelGPU.C"
#endif

#endif
