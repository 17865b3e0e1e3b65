/*---------------------------------------------------------------------------*\
  Registration of FastChemistryModelGPU under the reference's names: method
  FastChemistryModel with solvers OptRodas34, OptRosenbrock34, OptRodas23,
  OptRosenbrock23 and OptSeulex for the thermo the reference supports
  (src/odeChemistrySolvers.C:44-72).  The solver name only selects the
  library's stepper (read from chemistryType.solver), so one model class
  serves all five.  Not compiled in this repository (OpenFOAM is absent).
\*---------------------------------------------------------------------------*/
#include "FastChemistryModelGPU.H"
#include "makeChemistrySolver.H"
#include "thermoPhysicsTypes.H"

namespace Foam
{
    typedef sutherlandTransportsensibleEnthalpyjanafThermoperfectGasspecie FcsThermo;

    // basicChemistryModel selection by method name, as makeChemistrySolver
    // registers the reference's FastChemistryModel
    defineTemplateTypeNameAndDebugWithName
    (
        FastChemistryModelGPU<FcsThermo>,
        "FastChemistryModel<sutherlandTransportsensibleEnthalpyjanafThermoperfectGasspecie>",
        0
    );
    addToRunTimeSelectionTable
    (
        basicChemistryModel,
        FastChemistryModelGPU<FcsThermo>,
        thermo
    );
}
