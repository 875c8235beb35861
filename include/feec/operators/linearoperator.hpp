// feec/operators/linearoperator.hpp -- LinearOperator of FEEC++
// (operators/linearoperator.hpp:27-95): dimensions, apply(dest, src, scaling)
// and the operator* that allocates the product.
#pragma once

#include <string>

#include "../base/include.hpp"
#include "floatvector.hpp"

class LinearOperator {
public:
    LinearOperator(int dimout, int dimin) : dimout_(dimout), dimin_(dimin) {}
    virtual ~LinearOperator() = default;
    int getdimout() const { return dimout_; }
    int getdimin() const { return dimin_; }
    bool is_square() const { return dimout_ == dimin_; }
    virtual void check() const { assert(dimout_ >= 0 && dimin_ >= 0); }
    virtual std::string text() const = 0;
    // dest = scaling * (this) * src
    virtual void apply(FloatVector& dest, const FloatVector& src, Float scaling) const = 0;
    void apply(FloatVector& dest, const FloatVector& src) const { apply(dest, src, 1.); }

private:
    int dimout_;
    int dimin_;
};

// operators/linearoperator.hpp:87-95
inline FloatVector operator*(const LinearOperator& op, const FloatVector& vec)
{
    FloatVector ret(op.getdimout());
    op.apply(ret, vec, 1.);
    return ret;
}
