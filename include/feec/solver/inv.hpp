// feec/solver/inv.hpp -- InverseOperator<MatrixCSR> of FEEC++
// (solver/inv.hpp:18-76, 151-201): a LinearOperator whose apply solves
// op x = src by CG with Eisenstat-trick Rainbow SSOR (omega 1), the variant
// the reference's OpenMP build uses, starting from the previous solution.
//
// Many callers (block Hodge solvers, mixed formulations) apply the same
// inverse many times.  Here the matrix, its colouring and the solve's
// schedules stay on the GPU between applications; each apply copies src in
// and the result out.  The operator is read at every apply as the reference
// does: when its arrays differ from the ones the device copy was built from
// (a host copy of them, compared byte for byte), the device copy is rebuilt.
// Only MatrixCSR operators have this inverse; the reference's generic one (a
// conjugate residual method on any LinearOperator) is not part of this path.
#pragma once

#include <cstdint>
#include <memory>
#include <string>

#include "../base/include.hpp"
#include "../operators/floatvector.hpp"
#include "../operators/linearoperator.hpp"
#include "../sparse/matcsr.hpp"

template <typename Operator>
class InverseOperator;

template <>
class InverseOperator<MatrixCSR> : public LinearOperator {
public:
    InverseOperator() = delete;
    explicit InverseOperator(const MatrixCSR& op, Float precision, int print_modulo = -1);
    InverseOperator(const InverseOperator&);
    InverseOperator& operator=(const InverseOperator&) = delete;
    InverseOperator(InverseOperator&&) noexcept;
    ~InverseOperator() noexcept override;

    void check() const override;
    std::string text() const override;

    using LinearOperator::apply;
    // dest = scaling * op^-1 src (solver/inv.hpp:151-201)
    void apply(FloatVector& dest, const FloatVector& src, Float scaling) const override;

    // iteration count of the last apply
    int last_iterations() const { return last_iterations_; }

    bool use_previous_sol = true;

private:
    struct Device;  // the device-resident system

    const MatrixCSR& op;
    Float precision;
    int print_modulo;
    mutable FloatVector previous_sol;
    mutable std::unique_ptr<Device> device_;
    mutable int last_iterations_ = 0;
};
