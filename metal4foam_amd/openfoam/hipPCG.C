/*---------------------------------------------------------------------------*\
    hipFoam: MI355X (gfx950) linear solvers for OpenFOAM

    The OpenFOAM solver-class structure of this file (run-time selection,
    constructor, solve/scalarSolve, solveImpl) follows metal4Foam
    (scotCW/metal4Foam, src/metal4Foam/solvers/metalPCG.C), Copyright (C) 2026
    metal4Foam authors, and OpenFOAM's lduMatrix::solver API.  Both are
    licensed GPL-3.0-or-later; this file is distributed under the same terms.
    SPDX-License-Identifier: GPL-3.0-or-later
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
