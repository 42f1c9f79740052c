/*---------------------------------------------------------------------------*\
    Stand-in for OpenFOAM's scalarIOField.H (test infrastructure): only the
    type name is needed, as the key of objectRegistry::foundObject.
\*---------------------------------------------------------------------------*/

#ifndef Foam_shim_scalarIOField_H
#define Foam_shim_scalarIOField_H

#include "lduMatrix.H"

namespace Foam
{
    typedef Field<scalar> scalarIOField;
}

#endif
