/*---------------------------------------------------------------------------*\
    hipFoam: MI355X (gfx950) linear solvers for OpenFOAM

    The OpenFOAM solver-class structure of this file (run-time selection,
    constructor, solve/scalarSolve, solveImpl) follows metal4Foam
    (scotCW/metal4Foam, src/metal4Foam/solvers/metalPBiCGStab.C), Copyright (C) 2026
    metal4Foam authors, and OpenFOAM's lduMatrix::solver API.  Both are
    licensed GPL-3.0-or-later; this file is distributed under the same terms.
    SPDX-License-Identifier: GPL-3.0-or-later
\*---------------------------------------------------------------------------*/

#include "hipPBiCGStab.H"
#include "PrecisionAdaptor.H"

namespace Foam
{
    defineTypeNameAndDebug(hipPBiCGStab, 0);

    // symmetric and asymmetric matrices, as metalPBiCGStab
    // (metalPBiCGStab.C:16-22)
    lduMatrix::solver::addsymMatrixConstructorToTable<hipPBiCGStab>
        addhipPBiCGStabSymMatrixConstructorToTable_;

    lduMatrix::solver::addasymMatrixConstructorToTable<hipPBiCGStab>
        addhipPBiCGStabAsymMatrixConstructorToTable_;
}


Foam::hipPBiCGStab::hipPBiCGStab
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


Foam::solverPerformance Foam::hipPBiCGStab::scalarSolve
(
    solveScalarField& psi,
    const solveScalarField& source,
    const direction cmpt
) const
{
    return solveImpl(typeName, true, psi, source, cmpt);
}


Foam::solverPerformance Foam::hipPBiCGStab::solve
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
