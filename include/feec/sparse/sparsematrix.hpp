// feec/sparse/sparsematrix.hpp -- SparseMatrix of FEEC++ (sparse/sparsematrix.hpp:25-150):
// a host-owned list of (row, column, value) entries, the format assembly
// produces before MatrixCSR(const SparseMatrix&) turns it into CSR.
#pragma once

#include <functional>
#include <initializer_list>
#include <string>
#include <vector>

#include "../base/include.hpp"
#include "../operators/floatvector.hpp"
#include "../operators/linearoperator.hpp"

class SparseMatrix : public LinearOperator {
public:
    struct MatrixEntry {
        int row;
        int column;
        Float value;
    };

    enum class MatrixEntrySorting : bool { rowwise, columnwise };

    explicit SparseMatrix(int dimout, int dimin, int numentries = 0,
                          const std::function<MatrixEntry(int)>& generator =
                              [](int) -> MatrixEntry { return MatrixEntry{0, 0, notanumber}; });
    explicit SparseMatrix(int dimout, int dimin, const std::vector<MatrixEntry>& entries);
    explicit SparseMatrix(int dimout, int dimin, const std::initializer_list<MatrixEntry>& entries);

    SparseMatrix() = delete;
    SparseMatrix(const SparseMatrix&) = default;
    SparseMatrix& operator=(const SparseMatrix&) = default;
    SparseMatrix(SparseMatrix&&) noexcept = default;
    SparseMatrix& operator=(SparseMatrix&&) noexcept = default;
    ~SparseMatrix() noexcept override = default;

    void check() const override;
    std::string text() const override;

    using LinearOperator::apply;
    // dest = 0; dest[row] += scaling * value * src[column], entries in order
    // (sparse/sparsematrix.cpp:198-219).  Host loop: an assembly-side type.
    void apply(FloatVector& dest, const FloatVector& add, Float scaling) const override;

    void scale(Float s);
    bool is_finite() const;

    const std::vector<MatrixEntry>& getentries() const { return entries; }
    std::vector<MatrixEntry>& getentries() { return entries; }
    int getnumberofentries() const { return (int)entries.size(); }

    // (sparse/sparsematrix.cpp:307-326): non-decreasing (row, column) or (column, row)
    bool is_sorted(MatrixEntrySorting manner = MatrixEntrySorting::rowwise) const;
    // std::sort by (row, column) or (column, row) (sparse/sparsematrix.cpp:366-403)
    const SparseMatrix& sortentries(MatrixEntrySorting manner = MatrixEntrySorting::rowwise) const;
    // drop exact zeros, sort, add up entries with equal (row, column) in
    // sorted order (sparse/sparsematrix.cpp:405-466)
    const SparseMatrix& sortandcompressentries(
        MatrixEntrySorting manner = MatrixEntrySorting::rowwise) const;

    void reserve(int n) const { entries.reserve(n); }
    const MatrixEntry& getentry(int i) const { return entries[i]; }
    MatrixEntry& getentry(int i) { return entries[i]; }
    void setentry(int i, int r, int c, Float v) { setentry(i, MatrixEntry{r, c, v}); }
    void setentry(int i, MatrixEntry entry);
    void appendentry(int r, int c, Float v) { appendentry(MatrixEntry{r, c, v}); }
    void appendentry(MatrixEntry entry);
    void clearentries() { entries.clear(); }

private:
    mutable std::vector<MatrixEntry> entries;
};
