/*---------------------------------------------------------------------------*\
    hipFoam: MI355X (gfx950) linear solvers for OpenFOAM

    The OpenFOAM solver-class structure of this file (run-time selection,
    constructor, solve/scalarSolve, solveImpl) follows metal4Foam
    (scotCW/metal4Foam, src/metal4Foam/solvers/metalFoamSolver.C), Copyright (C) 2026
    metal4Foam authors, and OpenFOAM's lduMatrix::solver API.  Both are
    licensed GPL-3.0-or-later; this file is distributed under the same terms.
    SPDX-License-Identifier: GPL-3.0-or-later
\*---------------------------------------------------------------------------*/

#include "hipFoamSolver.H"
#include "hipfoam.h"
#include "PrecisionAdaptor.H"
#include "scalarIOField.H"

#include <map>
#include <memory>
#include <sstream>
#include <string>

static_assert
(
    sizeof(Foam::solveScalar) == 8,
    "hipFoam needs a double-precision OpenFOAM build (WM_PRECISION_OPTION=DP)"
);
static_assert
(
    sizeof(Foam::label) == 4,
    "hipFoam needs 32-bit labels (WM_LABEL_SIZE=32)"
);

namespace
{

// Owns one library context for the life of the process
struct ContextHolder
{
    hipfoam_context* ctx = nullptr;
    ~ContextHolder() { if (ctx) hipfoam_destroy(ctx); }
};

} // anonymous namespace


Foam::hipFoamSolver::hipFoamSolver
(
    const word& fieldName,
    const lduMatrix& matrix,
    const FieldField<Field, scalar>& interfaceBouCoeffs,
    const FieldField<Field, scalar>& interfaceIntCoeffs,
    const lduInterfaceFieldPtrsList& interfaces,
    const dictionary& solverControls
)
:
    lduMatrix::solver
    (
        fieldName,
        matrix,
        interfaceBouCoeffs,
        interfaceIntCoeffs,
        interfaces,
        solverControls
    ),
    hipDict_(solverControls.subOrEmptyDict("hip"))
{}


bool Foam::hipFoamSolver::residualFieldRequested(const bool initial) const
{
    const lduMesh& mesh = matrix_.mesh();
    return
        mesh.hasDb()
     && mesh.thisDb().foundObject<scalarIOField>
        (
            IOobject::scopedName
            (
                initial ? "initialResidual" : "residual",
                fieldName_
            )
        );
}


hipfoam_context* Foam::hipFoamSolver::context() const
{
    // keyed like the reference's backend cache: field name and mesh identity
    static std::map<std::string, std::unique_ptr<ContextHolder>> contexts;

    std::ostringstream key;
    key << fieldName_ << '@' << static_cast<const void*>(&matrix_.mesh());

    std::unique_ptr<ContextHolder>& slot = contexts[key.str()];
    if (!slot)
    {
        slot.reset(new ContextHolder);
        const label device = hipDict_.getOrDefault<label>("device", 0);
        hipfoam_create(device, &slot->ctx);
        Info<< "    " << hipfoam_description(slot->ctx) << nl;
        // "content" (default): any rewrite of the addressing is detected;
        // "pointers": the reference's key (sizes and addressing pointers)
        const word check = hipDict_.getOrDefault<word>("topologyCheck", "content");
        if (check != "content" && check != "pointers")
        {
            FatalErrorInFunction
                << "hip.topologyCheck must be content or pointers, not "
                << check << exit(FatalError);
        }
        hipfoam_set_topology_check
        (
            slot->ctx,
            check == "pointers" ? HIPFOAM_TOPO_POINTERS : HIPFOAM_TOPO_CONTENT
        );
        // how the inner loop runs on small meshes (every choice gives the
        // same results): "persistent" auto|on|off, the loop as one launch on
        // one XCD; "localLoop" auto|off, in one workgroup.  Off suits a
        // caller whose own kernels on other streams occupy the GPU.
        const word persistent = hipDict_.getOrDefault<word>("persistent", "auto");
        if (persistent != "auto" && persistent != "on" && persistent != "off")
        {
            FatalErrorInFunction
                << "hip.persistent must be auto, on or off, not "
                << persistent << exit(FatalError);
        }
        const word localLoop = hipDict_.getOrDefault<word>("localLoop", "auto");
        if (localLoop != "auto" && localLoop != "off")
        {
            FatalErrorInFunction
                << "hip.localLoop must be auto or off, not "
                << localLoop << exit(FatalError);
        }
        hipfoam_schedule sched;
        hipfoam_get_schedule(slot->ctx, &sched);
        sched.persistent =
            persistent == "auto" ? -1 : (persistent == "on" ? 1 : 0);
        sched.localLoop = localLoop == "auto" ? -1 : 0;
        hipfoam_set_schedule(slot->ctx, &sched);
    }
    return slot->ctx;
}


Foam::solverPerformance Foam::hipFoamSolver::solveImpl
(
    const word& solverName,
    const bool useBiCGStab,
    solveScalarField& psi,
    const solveScalarField& source,
    const direction cmpt
) const
{
    // Coupled interfaces would need their contributions inside every
    // device-resident iteration; like the reference, run serially without
    // coupled patches (metalFoamSolver.C:86-99)
    forAll(interfaces_, patchi)
    {
        if (interfaces_.set(patchi))
        {
            FatalErrorInFunction
                << solverName << " does not support coupled boundaries "
                << "(processor/cyclic/AMI); run serially on a mesh "
                << "without coupled patches."
                << exit(FatalError);
        }
    }

    solverPerformance solverPerf(solverName, fieldName_);

    hipfoam_context* ctx = context();
    if (!hipfoam_good(ctx))
    {
        FatalErrorInFunction
            << solverName << ": HIP backend unavailable: "
            << hipfoam_description(ctx)
            << exit(FatalError);
    }

    const lduAddressing& addr = matrix_.mesh().lduAddr();
    const label nCells = addr.size();
    const label nFaces = addr.upperAddr().size();

    int32_t rc = hipfoam_set_topology
    (
        ctx,
        nCells,
        nFaces,
        addr.lowerAddr().cdata(),
        addr.upperAddr().cdata(),
        addr.ownerStartAddr().cdata(),
        addr.losortAddr().cdata(),
        addr.losortStartAddr().cdata()
    );

    hipfoam_controls ctl;
    hipfoam_default_controls(&ctl);
    ctl.tolerance = tolerance_;
    ctl.relTol = relTol_;
    ctl.maxIter = maxIter_;
    ctl.minIter = minIter_;
    const word deflate = hipDict_.getOrDefault<word>("deflate", "auto");
    ctl.deflate =
        deflate == "off" ? 0 : deflate == "on" ? 1 : deflate == "pinned" ? 3 : 2;
    ctl.chebyshevOrder = hipDict_.getOrDefault<label>("chebyshevOrder", 8);
    ctl.chebyshevRatio = hipDict_.getOrDefault<scalar>("chebyshevRatio", 30.0);
    ctl.maxRestarts = hipDict_.getOrDefault<label>("maxRestarts", 8);
    ctl.iterationsPerSync = hipDict_.getOrDefault<label>("iterationsPerSync", 1);
    ctl.verbose = hipDict_.getOrDefault<bool>("verbose", false) ? 1 : 0;
    const bool profile = hipDict_.getOrDefault<bool>("profile", false);

    // lower is passed only for PBiCGStab on an asymmetric matrix
    // (metalFoamSolver.C:211-220)
    const scalar* lowerPtr =
        (useBiCGStab && matrix_.hasLower()) ? matrix_.lower().cdata() : nullptr;

    // lduMatrix::setResidualField stores nothing unless a solverInfo
    // function object registered the field: form FP64 residual fields on the
    // host only then.  The solve itself forms b - A psi on the GPU.
    if (residualFieldRequested(true))
    {
        solveScalarField wA(nCells);
        matrix_.Amul(wA, psi, interfaceBouCoeffs_, interfaces_, cmpt);
        const solveScalarField rA0(source - wA);
        matrix().setResidualField
        (
            ConstPrecisionAdaptor<scalar, solveScalar>(rA0)(),
            fieldName_,
            true
        );
    }
    const bool wantResidual = residualFieldRequested(false);
    solveScalarField rA(wantResidual ? nCells : 0);
    hipfoam_performance perf;

    if (rc == HIPFOAM_OK)
    {
        rc = hipfoam_solve
        (
            ctx,
            useBiCGStab ? HIPFOAM_PBICGSTAB : HIPFOAM_PCG,
            matrix_.diag().cdata(),
            matrix_.upper().cdata(),
            lowerPtr,
            psi.data(),
            source.cdata(),
            &ctl,
            &perf,
            wantResidual ? rA.data() : nullptr
        );
    }
    if (rc != HIPFOAM_OK)
    {
        FatalErrorInFunction
            << solverName << ": " << hipfoam_last_error(ctx)
            << exit(FatalError);
    }

    if ((log_ >= 2) || (lduMatrix::debug >= 2))
    {
        Info<< "   Normalisation factor = " << perf.normFactor << endl;
    }

    solverPerf.initialResidual() = perf.initialResidual;
    solverPerf.finalResidual() = perf.finalResidual;
    solverPerf.nIterations() = perf.nIterations;
    solverPerf.checkConvergence(tolerance_, relTol_, log_);

    // residual field of the final FP64 state (metalFoamSolver.C:429-434)
    if (wantResidual)
    {
        matrix().setResidualField
        (
            ConstPrecisionAdaptor<scalar, solveScalar>(rA)(),
            fieldName_,
            false
        );
    }

    if (profile)
    {
        Info<< "    " << solverName << " profile: cells=" << nCells
            << " iters=" << perf.nIterations
            << " restarts=" << perf.restarts
            << " deflated=" << perf.deflated
            << " polished=" << perf.polished
            << " tConvert=" << 1000*perf.tConvert << "ms"
            << " tInner=" << 1000*perf.tInner << "ms"
            << " tGpu=" << 1000*perf.tGpu << "ms"
            << " tTotal=" << 1000*perf.tTotal << "ms"
            << nl;
    }

    return solverPerf;
}
