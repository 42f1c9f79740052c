/*---------------------------------------------------------------------------*\
    Minimal stand-in for the parts of OpenFOAM's API that the hipFoam solver
    plugin (metal4foam_amd/openfoam) uses, so that the plugin's own sources
    compile and run in tests without an OpenFOAM installation.

    Test infrastructure only.  Class and member names follow the public
    OpenFOAM (v2506) headers of the same names; the behaviour restated here
    is the part the plugin relies on:
      - lduMatrix::solver run-time selection: New() picks the symmetric-
        matrix table for a matrix without lower coefficients and the
        asymmetric table otherwise, and raises FatalError for an unknown
        solver name (lduMatrixSolver.C);
      - solver controls read by lduMatrix::solver::readControls: maxIter
        (default 1000), minIter (0), tolerance (1e-6), relTol (0), log (1);
      - SolverPerformance::checkConvergence: converged when
        final < tolerance, or relTol > 1e-20 and final < relTol*initial;
      - lduMatrix::Amul without interfaces: diag, then upper[f] psi[u]
        into row l and lower[f] psi[l] into row u;
      - lduMatrix::setResidualField stores the field only when the mesh's
        registry holds "initialResidual:<field>" / "residual:<field>";
      - FatalError: the message is thrown as Foam::error (OpenFOAM's
        FOAM_ABORT-free, throwExceptions mode).
\*---------------------------------------------------------------------------*/

#ifndef Foam_shim_lduMatrix_H
#define Foam_shim_lduMatrix_H

#include <cctype>
#include <cstdint>
#include <cstdlib>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

namespace Foam
{

typedef int32_t label;
typedef double scalar;
typedef double solveScalar;
typedef uint8_t direction;

class word : public std::string
{
public:
    word() = default;
    word(const char* s) : std::string(s) {}
    word(const std::string& s) : std::string(s) {}
};

// ---- output ---------------------------------------------------------------

inline std::ostream& Info = std::cout;
constexpr char nl = '\n';
using std::endl;

// ---- errors ---------------------------------------------------------------

class error : public std::runtime_error
{
public:
    explicit error(const std::string& msg) : std::runtime_error(msg) {}
};

class errorStream
{
    std::ostringstream msg_;

public:
    errorStream& operator()(const char* function, const char*, int)
    {
        msg_.str("");
        msg_ << "--> FOAM FATAL ERROR: (in " << function << ")\n    ";
        return *this;
    }

    template<class T>
    errorStream& operator<<(const T& v)
    {
        msg_ << v;
        return *this;
    }

    [[noreturn]] void exitNow() { throw error(msg_.str()); }
};

inline errorStream FatalError;

struct errorExit
{
    errorStream* err;
};

inline errorExit exit(errorStream& err, const int = 1)
{
    return errorExit{&err};
}

inline errorStream& operator<<(errorStream& err, const errorExit& e)
{
    e.err->exitNow();
}

#define FatalErrorInFunction \
    ::Foam::FatalError(__PRETTY_FUNCTION__, __FILE__, __LINE__)

// ---- run-time type information --------------------------------------------

#define TypeName(TypeNameString)                                              \
    static const char* typeName_() { return TypeNameString; }                \
    static const ::Foam::word typeName;                                       \
    static int debug;                                                         \
    virtual const ::Foam::word& type() const { return typeName; }

#define defineTypeNameAndDebug(Type, DebugSwitch)                             \
    const ::Foam::word Type::typeName(Type::typeName_());                     \
    int Type::debug(DebugSwitch)

// ---- containers -----------------------------------------------------------

template<class T>
class Field
{
    std::vector<T> v_;

public:
    Field() = default;
    explicit Field(const label n) : v_(n) {}
    Field(const label n, const T& value) : v_(n, value) {}
    Field(std::vector<T> v) : v_(std::move(v)) {}

    label size() const { return static_cast<label>(v_.size()); }
    bool empty() const { return v_.empty(); }
    T* data() { return v_.data(); }
    const T* cdata() const { return v_.data(); }
    T& operator[](const label i) { return v_[i]; }
    const T& operator[](const label i) const { return v_[i]; }
    T* begin() { return v_.data(); }
    T* end() { return v_.data() + v_.size(); }
    const T* begin() const { return v_.data(); }
    const T* end() const { return v_.data() + v_.size(); }
};

template<class T>
Field<T> operator-(const Field<T>& a, const Field<T>& b)
{
    Field<T> r(a.size());
    for (label i = 0; i < a.size(); ++i) r[i] = a[i] - b[i];
    return r;
}

typedef Field<scalar> scalarField;
typedef Field<solveScalar> solveScalarField;
typedef Field<label> labelList;
typedef Field<label> labelUList;

template<template<class> class FieldType, class Type>
class FieldField : public std::vector<FieldType<Type>>
{};

#define forAll(list, i) \
    for (::Foam::label i = 0; i < (list).size(); ++i)

//- Pointers to the coupled interface fields: a null entry is an uncoupled
//  patch, a non-null one a processor/cyclic/AMI coupling
class lduInterfaceFieldPtrsList
{
    std::vector<const void*> ptrs_;

public:
    lduInterfaceFieldPtrsList() = default;
    explicit lduInterfaceFieldPtrsList(const label n) : ptrs_(n, nullptr) {}
    label size() const { return static_cast<label>(ptrs_.size()); }
    bool set(const label i) const { return ptrs_[i] != nullptr; }
    void set(const label i, const void* p) { ptrs_[i] = p; }
};

// ---- dictionary -----------------------------------------------------------

class dictionary
{
    std::map<std::string, std::string> entries_;
    std::map<std::string, dictionary> subDicts_;

    static bool parseBool(const std::string& s)
    {
        if (s == "true" || s == "on" || s == "yes" || s == "1") return true;
        if (s == "false" || s == "off" || s == "no" || s == "none"
         || s == "0")
        {
            return false;
        }
        throw error("bad bool '" + s + "'");
    }

public:
    void add(const word& key, const std::string& value)
    {
        entries_[key] = value;
    }

    dictionary& subDictOrAdd(const word& key) { return subDicts_[key]; }

    bool found(const word& key) const
    {
        return entries_.count(key) || subDicts_.count(key);
    }

    dictionary subOrEmptyDict(const word& key) const
    {
        const auto it = subDicts_.find(key);
        return it == subDicts_.end() ? dictionary() : it->second;
    }

    template<class T>
    T get(const word& key) const
    {
        const auto it = entries_.find(key);
        if (it == entries_.end())
        {
            throw error("keyword '" + key + "' is undefined");
        }
        if constexpr (std::is_same<T, bool>::value)
        {
            return parseBool(it->second);
        }
        else if constexpr (std::is_same<T, word>::value)
        {
            return word(it->second);
        }
        else
        {
            std::istringstream is(it->second);
            T v{};
            is >> v;
            if (is.fail()) throw error("bad value for '" + key + "'");
            return v;
        }
    }

    template<class T>
    T getOrDefault(const word& key, const T& deflt) const
    {
        return entries_.count(key) ? get<T>(key) : deflt;
    }

    //- Parse "key value;" entries and "key { ... }" sub-dictionaries
    static dictionary parse(std::istream& is)
    {
        std::vector<std::string> tokens;
        std::string tok;
        char ch;
        auto flush = [&]()
        {
            if (!tok.empty()) tokens.push_back(tok);
            tok.clear();
        };
        while (is.get(ch))
        {
            if (ch == '{' || ch == '}' || ch == ';')
            {
                flush();
                tokens.push_back(std::string(1, ch));
            }
            else if (std::isspace(static_cast<unsigned char>(ch)))
            {
                flush();
            }
            else
            {
                tok += ch;
            }
        }
        flush();
        size_t pos = 0;
        return parseTokens(tokens, pos);
    }

private:
    static dictionary parseTokens
    (
        const std::vector<std::string>& t,
        size_t& pos
    )
    {
        dictionary d;
        while (pos < t.size() && t[pos] != "}")
        {
            const std::string key = t[pos++];
            if (pos < t.size() && t[pos] == "{")
            {
                ++pos;
                d.subDicts_[key] = parseTokens(t, pos);
                ++pos;  // '}'
            }
            else
            {
                std::string value;
                while (pos < t.size() && t[pos] != ";")
                {
                    value += (value.empty() ? "" : " ") + t[pos++];
                }
                ++pos;  // ';'
                d.entries_[key] = value;
            }
        }
        return d;
    }
};

// ---- mesh, registry, addressing -------------------------------------------

class IOobject
{
public:
    static word scopedName(const std::string& scope, const word& name)
    {
        return word(scope + ':' + name);
    }
};

class objectRegistry
{
    std::set<std::string> names_;
    mutable std::map<std::string, std::vector<scalar>> stored_;

public:
    void registerObject(const word& name) { names_.insert(name); }

    template<class T>
    bool foundObject(const word& name) const
    {
        return names_.count(name) > 0;
    }

    void store(const word& name, const scalar* v, const label n) const
    {
        stored_[name].assign(v, v + n);
    }

    const std::vector<scalar>* stored(const word& name) const
    {
        const auto it = stored_.find(name);
        return it == stored_.end() ? nullptr : &it->second;
    }
};

class lduAddressing
{
public:
    label nCells = 0;
    labelList lower, upper, ownerStart, losort, losortStart;

    label size() const { return nCells; }
    const labelUList& lowerAddr() const { return lower; }
    const labelUList& upperAddr() const { return upper; }
    const labelUList& ownerStartAddr() const { return ownerStart; }
    const labelUList& losortAddr() const { return losort; }
    const labelUList& losortStartAddr() const { return losortStart; }
};

class lduMesh
{
    lduAddressing addr_;
    objectRegistry db_;

public:
    lduAddressing& addr() { return addr_; }
    objectRegistry& db() { return db_; }
    const lduAddressing& lduAddr() const { return addr_; }
    bool hasDb() const { return true; }
    const objectRegistry& thisDb() const { return db_; }
};

// ---- solver performance ---------------------------------------------------

class solverPerformance
{
    word solverName_, fieldName_;
    scalar initialResidual_ = 0, finalResidual_ = 0;
    label nIterations_ = 0;
    bool converged_ = false;

public:
    solverPerformance(const word& solverName, const word& fieldName)
    :
        solverName_(solverName),
        fieldName_(fieldName)
    {}

    const word& solverName() const { return solverName_; }
    const word& fieldName() const { return fieldName_; }
    scalar& initialResidual() { return initialResidual_; }
    scalar initialResidual() const { return initialResidual_; }
    scalar& finalResidual() { return finalResidual_; }
    scalar finalResidual() const { return finalResidual_; }
    label& nIterations() { return nIterations_; }
    label nIterations() const { return nIterations_; }
    bool converged() const { return converged_; }

    bool checkConvergence
    (
        const scalar tolerance,
        const scalar relTol,
        const int logLevel = 0
    )
    {
        converged_ =
            finalResidual_ < tolerance
         || (relTol > 1e-20 && finalResidual_ < relTol*initialResidual_);
        return converged_;
    }
};

// ---- lduMatrix ------------------------------------------------------------

class lduMatrix
{
    const lduMesh& mesh_;
    scalarField diag_, upper_, lower_;
    bool hasLower_ = false;

public:

    static inline int debug = 0;

    explicit lduMatrix(const lduMesh& mesh) : mesh_(mesh) {}

    const lduMesh& mesh() const { return mesh_; }
    const lduAddressing& lduAddr() const { return mesh_.lduAddr(); }

    scalarField& diag() { return diag_; }
    scalarField& upper() { return upper_; }
    scalarField& lower()
    {
        hasLower_ = true;
        return lower_;
    }
    const scalarField& diag() const { return diag_; }
    const scalarField& upper() const { return upper_; }
    const scalarField& lower() const { return hasLower_ ? lower_ : upper_; }

    bool hasLower() const { return hasLower_; }
    bool symmetric() const { return !hasLower_; }
    bool asymmetric() const { return hasLower_; }

    void Amul
    (
        solveScalarField& Apsi,
        const solveScalarField& psi,
        const FieldField<Field, scalar>&,
        const lduInterfaceFieldPtrsList&,
        const direction
    ) const
    {
        const lduAddressing& a = lduAddr();
        const scalarField& L = lower();
        for (label i = 0; i < a.size(); ++i) Apsi[i] = diag_[i]*psi[i];
        for (label f = 0; f < a.lower.size(); ++f)
        {
            Apsi[a.upper[f]] += L[f]*psi[a.lower[f]];
            Apsi[a.lower[f]] += upper_[f]*psi[a.upper[f]];
        }
    }

    void setResidualField
    (
        const scalarField& residual,
        const word& fieldName,
        const bool initial
    ) const
    {
        const word name = IOobject::scopedName
        (
            initial ? "initialResidual" : "residual",
            fieldName
        );
        if (mesh_.hasDb() && mesh_.thisDb().foundObject<void>(name))
        {
            mesh_.thisDb().store(name, residual.cdata(), residual.size());
        }
    }

    // ---- solver base and run-time selection -------------------------------

    class solver
    {
    protected:

        word fieldName_;
        const lduMatrix& matrix_;
        const FieldField<Field, scalar>& interfaceBouCoeffs_;
        const FieldField<Field, scalar>& interfaceIntCoeffs_;
        lduInterfaceFieldPtrsList interfaces_;
        dictionary controlDict_;
        int log_ = 1;
        label minIter_ = 0;
        label maxIter_ = 1000;
        scalar tolerance_ = 1e-6;
        scalar relTol_ = 0;

    public:

        typedef std::function
        <
            std::unique_ptr<solver>
            (
                const word&,
                const lduMatrix&,
                const FieldField<Field, scalar>&,
                const FieldField<Field, scalar>&,
                const lduInterfaceFieldPtrsList&,
                const dictionary&
            )
        > constructor;

        typedef std::map<std::string, constructor> constructorTable;

        static constructorTable& symMatrixConstructorTable()
        {
            static constructorTable t;
            return t;
        }

        static constructorTable& asymMatrixConstructorTable()
        {
            static constructorTable t;
            return t;
        }

        template<class SolverType>
        static constructor makeConstructor()
        {
            return []
            (
                const word& fieldName,
                const lduMatrix& matrix,
                const FieldField<Field, scalar>& bou,
                const FieldField<Field, scalar>& in,
                const lduInterfaceFieldPtrsList& interfaces,
                const dictionary& controls
            )
            {
                return std::unique_ptr<solver>
                (
                    new SolverType
                    (
                        fieldName, matrix, bou, in, interfaces, controls
                    )
                );
            };
        }

        template<class SolverType>
        struct addsymMatrixConstructorToTable
        {
            addsymMatrixConstructorToTable()
            {
                symMatrixConstructorTable()[SolverType::typeName] =
                    makeConstructor<SolverType>();
            }
        };

        template<class SolverType>
        struct addasymMatrixConstructorToTable
        {
            addasymMatrixConstructorToTable()
            {
                asymMatrixConstructorTable()[SolverType::typeName] =
                    makeConstructor<SolverType>();
            }
        };

        static std::unique_ptr<solver> New
        (
            const word& fieldName,
            const lduMatrix& matrix,
            const FieldField<Field, scalar>& interfaceBouCoeffs,
            const FieldField<Field, scalar>& interfaceIntCoeffs,
            const lduInterfaceFieldPtrsList& interfaces,
            const dictionary& solverControls
        )
        {
            const word name = solverControls.get<word>("solver");
            constructorTable& table =
                matrix.symmetric()
              ? symMatrixConstructorTable()
              : asymMatrixConstructorTable();
            const auto it = table.find(name);
            if (it == table.end())
            {
                std::string valid;
                for (const auto& kv : table) valid += " " + kv.first;
                FatalErrorInFunction
                    << "Unknown "
                    << (matrix.symmetric() ? "symmetric" : "asymmetric")
                    << " matrix solver " << name
                    << "\n    Valid solvers:" << valid
                    << exit(FatalError);
            }
            return it->second
            (
                fieldName,
                matrix,
                interfaceBouCoeffs,
                interfaceIntCoeffs,
                interfaces,
                solverControls
            );
        }

        solver
        (
            const word& fieldName,
            const lduMatrix& matrix,
            const FieldField<Field, scalar>& interfaceBouCoeffs,
            const FieldField<Field, scalar>& interfaceIntCoeffs,
            const lduInterfaceFieldPtrsList& interfaces,
            const dictionary& solverControls
        )
        :
            fieldName_(fieldName),
            matrix_(matrix),
            interfaceBouCoeffs_(interfaceBouCoeffs),
            interfaceIntCoeffs_(interfaceIntCoeffs),
            interfaces_(interfaces),
            controlDict_(solverControls)
        {
            log_ = controlDict_.getOrDefault<int>("log", 1);
            minIter_ = controlDict_.getOrDefault<label>("minIter", 0);
            maxIter_ = controlDict_.getOrDefault<label>("maxIter", 1000);
            tolerance_ = controlDict_.getOrDefault<scalar>("tolerance", 1e-6);
            relTol_ = controlDict_.getOrDefault<scalar>("relTol", 0);
        }

        virtual ~solver() = default;

        virtual const word& type() const = 0;
        const word& fieldName() const { return fieldName_; }
        const lduMatrix& matrix() const { return matrix_; }

        virtual solverPerformance solve
        (
            scalarField& psi,
            const scalarField& source,
            const direction cmpt = 0
        ) const = 0;

        virtual solverPerformance scalarSolve
        (
            solveScalarField& psi,
            const solveScalarField& source,
            const direction cmpt = 0
        ) const = 0;
    };
};

} // End namespace Foam

#endif
