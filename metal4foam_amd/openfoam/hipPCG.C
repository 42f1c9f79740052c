/*---------------------------------------------------------------------------*\
    hipFoam: MI355X (gfx950) linear solvers for OpenFOAM
\*---------------------------------------------------------------------------*/

#include "hipPCG.H"
#include "PrecisionAdaptor.H"

namespace Foam
{
    defineTypeNameAndDebug(hipPCG, 0);

    // symmetric matrices only, as metalPCG (metalPCG.C:16-19)
    lduMatrix::solver::addsymMatrixConstructorToTable<hipPCG>
        addhipPCGSymMatrixConstructorToTable_;
}


Foam::hipPCG::hipPCG
(
    const word& fieldName,
    const lduMatrix& matrix,
    const FieldField<Field, scalar>& interfaceBouCoeffs,
    const FieldField<Field, scalar>& interfaceIntCoeffs,
    const lduInterfaceFieldPtrsList& interfaces,
    const dictionary& solverControls
)
:
    hipFoamSolver
    (
        fieldName,
        matrix,
        interfaceBouCoeffs,
        interfaceIntCoeffs,
        interfaces,
        solverControls
    )
{}


Foam::solverPerformance Foam::hipPCG::scalarSolve
(
    solveScalarField& psi,
    const solveScalarField& source,
    const direction cmpt
) const
{
    return solveImpl(typeName, false, psi, source, cmpt);
}


Foam::solverPerformance Foam::hipPCG::solve
(
    scalarField& psi_s,
    const scalarField& source,
    const direction cmpt
) const
{
    PrecisionAdaptor<solveScalar, scalar> tpsi(psi_s);
    return scalarSolve
    (
        tpsi.ref(),
        ConstPrecisionAdaptor<solveScalar, scalar>(source)(),
        cmpt
    );
}
