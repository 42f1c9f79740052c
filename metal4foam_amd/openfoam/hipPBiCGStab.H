/*---------------------------------------------------------------------------*\
    hipFoam: MI355X (gfx950) linear solvers for OpenFOAM

    The OpenFOAM solver-class structure of this file (run-time selection,
    constructor, solve/scalarSolve, solveImpl) follows metal4Foam
    (scotCW/metal4Foam, src/metal4Foam/solvers/metalPBiCGStab.H), Copyright (C) 2026
    metal4Foam authors, and OpenFOAM's lduMatrix::solver API.  Both are
    licensed GPL-3.0-or-later; this file is distributed under the same terms.
    SPDX-License-Identifier: GPL-3.0-or-later

Class
    Foam::hipPBiCGStab

Group
    grpLduMatrixSolvers

Description
    Preconditioned bi-conjugate gradient stabilised for symmetric and
    asymmetric lduMatrices on an AMD Instinct MI355X: FP32 Jacobi BiCGStab
    iterations with FP64 iterative refinement, every step on the GPU.
    Registered in the symmetric- and asymmetric-matrix tables, like
    metalPBiCGStab (metalPBiCGStab.C:16-22).

    Selected in fvSolution like the built-in solvers:
    \verbatim
    U
    {
        solver          hipPBiCGStab;
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
    hipPBiCGStab.C

\*---------------------------------------------------------------------------*/

#ifndef Foam_hipPBiCGStab_H
#define Foam_hipPBiCGStab_H

#include "hipFoamSolver.H"

namespace Foam
{

class hipPBiCGStab
:
    public hipFoamSolver
{
public:

    //- Runtime type information
    TypeName("hipPBiCGStab");

        hipPBiCGStab
        (
            const word& fieldName,
            const lduMatrix& matrix,
            const FieldField<Field, scalar>& interfaceBouCoeffs,
            const FieldField<Field, scalar>& interfaceIntCoeffs,
            const lduInterfaceFieldPtrsList& interfaces,
            const dictionary& solverControls
        );

        virtual ~hipPBiCGStab() = default;

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
