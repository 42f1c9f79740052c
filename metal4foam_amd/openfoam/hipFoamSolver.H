/*---------------------------------------------------------------------------*\
    hipFoam: MI355X (gfx950) linear solvers for OpenFOAM

    The OpenFOAM solver-class structure of this file (run-time selection,
    constructor, solve/scalarSolve, solveImpl) follows metal4Foam
    (scotCW/metal4Foam, src/metal4Foam/solvers/metalFoamSolver.H), Copyright (C) 2026
    metal4Foam authors, and OpenFOAM's lduMatrix::solver API.  Both are
    licensed GPL-3.0-or-later; this file is distributed under the same terms.
    SPDX-License-Identifier: GPL-3.0-or-later

Class
    Foam::hipFoamSolver

Description
    Common base of hipPCG and hipPBiCGStab: an lduMatrix::solver that hands
    the whole refined solve -- initial FP64 residual and normFactor, sumA,
    constant-mode deflation, Gershgorin bound, FP32 Krylov iterations, FP64
    iterative refinement, constant-mode correction, hybrid polish -- to the
    HIP library through its C-ABI (include/hipfoam.h).  It replaces the
    reference's metalFoamSolver (src/metal4Foam/solvers/metalFoamSolver.H,
    metalFoamSolver.C:77-437), whose FP64 steps ran on the CPU.

    Optional controls, in a "hip" sub-dictionary of the solver entry:
        deflate            auto | on | off | pinned   (auto; pinned: auto
                           detection, inner right-hand side projected along
                           A 1, for sources with a non-zero sum on a matrix
                           with pinned reference cells)
        chebyshevOrder     8                 (hipPCG; 1 = Jacobi)
        chebyshevRatio     30
        maxRestarts        8
        iterationsPerSync  1
        profile            false             per-solve timing line
        verbose            false             per-restart residual trace
        device             0                 HIP device ordinal

    Requirements checked at compile time: double-precision scalars and
    32-bit labels (the C-ABI takes double and int32 arrays as they are).

SourceFiles
    hipFoamSolver.C

\*---------------------------------------------------------------------------*/

#ifndef Foam_hipFoamSolver_H
#define Foam_hipFoamSolver_H

#include "lduMatrix.H"

struct hipfoam_context;

namespace Foam
{

class hipFoamSolver
:
    public lduMatrix::solver
{
protected:

        //- The "hip" sub-dictionary
        dictionary hipDict_;

        //- Library context of this (field, mesh): created on first use and
        //  kept, because the solver object is re-created on every solve
        hipfoam_context* context() const;

        //- True when a solverInfo function object registered the
        //  initial (or final) residual field of this field name
        bool residualFieldRequested(const bool initial) const;

        //- The refined solve shared by hipPCG (useBiCGStab = false) and
        //  hipPBiCGStab (useBiCGStab = true)
        solverPerformance solveImpl
        (
            const word& solverName,
            const bool useBiCGStab,
            solveScalarField& psi,
            const solveScalarField& source,
            const direction cmpt
        ) const;

public:

        hipFoamSolver
        (
            const word& fieldName,
            const lduMatrix& matrix,
            const FieldField<Field, scalar>& interfaceBouCoeffs,
            const FieldField<Field, scalar>& interfaceIntCoeffs,
            const lduInterfaceFieldPtrsList& interfaces,
            const dictionary& solverControls
        );

        hipFoamSolver(const hipFoamSolver&) = delete;
        void operator=(const hipFoamSolver&) = delete;

        virtual ~hipFoamSolver() = default;
};

} // End namespace Foam

#endif
