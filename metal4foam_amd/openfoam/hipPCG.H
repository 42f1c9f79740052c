/*---------------------------------------------------------------------------*\
    hipFoam: MI355X (gfx950) linear solvers for OpenFOAM

    The OpenFOAM solver-class structure of this file (run-time selection,
    constructor, solve/scalarSolve, solveImpl) follows metal4Foam
    (scotCW/metal4Foam, src/metal4Foam/solvers/metalPCG.H), Copyright (C) 2026
    metal4Foam authors, and OpenFOAM's lduMatrix::solver API.  Both are
    licensed GPL-3.0-or-later; this file is distributed under the same terms.
    SPDX-License-Identifier: GPL-3.0-or-later

Class
    Foam::hipPCG

Group
    grpLduMatrixSolvers

Description
    Preconditioned conjugate gradient for symmetric lduMatrices on an AMD
    Instinct MI355X: FP32 Chebyshev-Jacobi PCG iterations with FP64 iterative
    refinement, every step on the GPU.  Registered in the symmetric-matrix
    table only, like metalPCG (metalPCG.C:16-19).

    Selected in fvSolution like the built-in solvers:
    \verbatim
    p
    {
        solver          hipPCG;
        tolerance       1e-06;
        relTol          0.05;
        hip
        {
            maxRestarts     8;
            profile         false;
        }
    }
    \endverbatim
    with the library loaded from controlDict: libs (hipFoam);

SourceFiles
    hipPCG.C

\*---------------------------------------------------------------------------*/

#ifndef Foam_hipPCG_H
#define Foam_hipPCG_H

#include "hipFoamSolver.H"

namespace Foam
{

class hipPCG
:
    public hipFoamSolver
{
public:

    //- Runtime type information
    TypeName("hipPCG");

        hipPCG
        (
            const word& fieldName,
            const lduMatrix& matrix,
            const FieldField<Field, scalar>& interfaceBouCoeffs,
            const FieldField<Field, scalar>& interfaceIntCoeffs,
            const lduInterfaceFieldPtrsList& interfaces,
            const dictionary& solverControls
        );

        virtual ~hipPCG() = default;

        //- Solve in solveScalar precision
        virtual solverPerformance scalarSolve
        (
            solveScalarField& psi,
            const solveScalarField& source,
            const direction cmpt = 0
        ) const;

        //- Solve the matrix with this solver
        virtual solverPerformance solve
        (
            scalarField& psi,
            const scalarField& source,
            const direction cmpt = 0
        ) const;
};

} // End namespace Foam

#endif
