/*---------------------------------------------------------------------------*\
    Test driver for the hipFoam OpenFOAM plugin (test infrastructure).

    Plays the part of fvMatrix::solve: builds an lduMesh/lduMatrix from a
    case file, selects the solver by name through lduMatrix::solver::New and
    calls solve(psi, source) the way OpenFOAM does -- a fresh solver object
    per solve, the library context cached by the plugin across solves.

    Usage: foam_driver <case.bin> <solverControls> <out.bin> [nSolves]

    case.bin (little-endian):
        int32  nCells, nFaces, hasLower, nInterfaces, coupledPatch (-1: none),
               wantInitialResidual, wantResidual, reserved
        int32  lowerAddr[nFaces], upperAddr[nFaces], ownerStart[nCells+1],
               losort[nFaces], losortStart[nCells+1]
        double diag[nCells], upper[nFaces], lower[nFaces] (if hasLower),
               source[nCells], psi[nCells]

    out.bin: double psi[nCells], then initialResidual[nCells] and
    residual[nCells] for the fields requested.

    stdout, one line per solve:
        RESULT <solver> initial=<r0> final=<r> iters=<k> converged=<0|1>
    A FatalError prints "FATAL <message>" and exits with status 3.
\*---------------------------------------------------------------------------*/

#include "lduMatrix.H"

#include <cstdio>
#include <fstream>

using namespace Foam;

namespace
{

template<class T>
void readArray(std::ifstream& is, Field<T>& f, const label n)
{
    f = Field<T>(n);
    is.read(reinterpret_cast<char*>(f.data()), sizeof(T)*size_t(n));
    if (!is) throw error("case file truncated");
}

void writeArray(std::ofstream& os, const scalar* v, const label n)
{
    os.write(reinterpret_cast<const char*>(v), sizeof(scalar)*size_t(n));
}

} // anonymous namespace


int main(int argc, char** argv)
{
    if (argc < 4)
    {
        std::fprintf
        (
            stderr,
            "usage: %s case.bin solverControls out.bin [nSolves]\n",
            argv[0]
        );
        return 2;
    }
    const int nSolves = argc > 4 ? std::atoi(argv[4]) : 1;

    try
    {
        std::ifstream is(argv[1], std::ios::binary);
        if (!is) throw error(std::string("cannot open ") + argv[1]);
        int32_t hdr[8];
        is.read(reinterpret_cast<char*>(hdr), sizeof(hdr));
        if (!is) throw error("case header truncated");
        const label n = hdr[0], nFaces = hdr[1];

        lduMesh mesh;
        lduAddressing& addr = mesh.addr();
        addr.nCells = n;
        readArray(is, addr.lower, nFaces);
        readArray(is, addr.upper, nFaces);
        readArray(is, addr.ownerStart, n + 1);
        readArray(is, addr.losort, nFaces);
        readArray(is, addr.losortStart, n + 1);
        if (hdr[5]) mesh.db().registerObject("initialResidual:x");
        if (hdr[6]) mesh.db().registerObject("residual:x");

        lduMatrix matrix(mesh);
        readArray(is, matrix.diag(), n);
        readArray(is, matrix.upper(), nFaces);
        if (hdr[2]) readArray(is, matrix.lower(), nFaces);
        scalarField source, psi;
        readArray(is, source, n);
        readArray(is, psi, n);

        // one opaque non-null pointer marks the coupled patch
        lduInterfaceFieldPtrsList interfaces(hdr[3]);
        static const int coupledTag = 0;
        if (hdr[4] >= 0) interfaces.set(hdr[4], &coupledTag);
        FieldField<Field, scalar> bouCoeffs, intCoeffs;

        std::ifstream ds(argv[2]);
        if (!ds) throw error(std::string("cannot open ") + argv[2]);
        const dictionary controls = dictionary::parse(ds);

        for (int k = 0; k < nSolves; ++k)
        {
            std::unique_ptr<lduMatrix::solver> solver =
                lduMatrix::solver::New
                (
                    "x", matrix, bouCoeffs, intCoeffs, interfaces, controls
                );
            const solverPerformance perf = solver->solve(psi, source);
            std::printf
            (
                "RESULT %s initial=%.17g final=%.17g iters=%d converged=%d\n",
                perf.solverName().c_str(),
                perf.initialResidual(),
                perf.finalResidual(),
                int(perf.nIterations()),
                int(perf.converged())
            );
            std::fflush(stdout);
        }

        std::ofstream os(argv[3], std::ios::binary);
        writeArray(os, psi.cdata(), n);
        for (const char* name : {"initialResidual:x", "residual:x"})
        {
            if (const auto* r = mesh.thisDb().stored(name))
            {
                writeArray(os, r->data(), label(r->size()));
            }
        }
        if (!os) throw error(std::string("cannot write ") + argv[3]);
    }
    catch (const error& e)
    {
        std::string msg(e.what());
        for (char& ch : msg) if (ch == '\n') ch = ' ';
        std::printf("FATAL %s\n", msg.c_str());
        std::fflush(stdout);
        return 3;
    }
    return 0;
}
