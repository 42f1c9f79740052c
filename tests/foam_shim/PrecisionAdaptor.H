/*---------------------------------------------------------------------------*\
    Stand-in for OpenFOAM's PrecisionAdaptor.H (test infrastructure).  In a
    double-precision build with solveScalar == scalar both adaptors refer to
    the field they are given, which is the only case the hipFoam plugin
    supports.
\*---------------------------------------------------------------------------*/

#ifndef Foam_shim_PrecisionAdaptor_H
#define Foam_shim_PrecisionAdaptor_H

#include "lduMatrix.H"

#include <type_traits>

namespace Foam
{

template<class Type, class InputType>
class ConstPrecisionAdaptor
{
    static_assert(std::is_same<Type, InputType>::value, "same precision");
    const Field<Type>& field_;

public:
    explicit ConstPrecisionAdaptor(const Field<InputType>& input)
    :
        field_(input)
    {}

    const Field<Type>& operator()() const { return field_; }
};

template<class Type, class InputType>
class PrecisionAdaptor
{
    static_assert(std::is_same<Type, InputType>::value, "same precision");
    Field<Type>& field_;

public:
    explicit PrecisionAdaptor(Field<InputType>& input) : field_(input) {}

    Field<Type>& ref() { return field_; }
};

} // End namespace Foam

#endif
