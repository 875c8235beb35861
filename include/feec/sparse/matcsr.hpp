// feec/sparse/matcsr.hpp -- MatrixCSR of FEEC++ (sparse/matcsr.hpp:25-164).
// A host-owned CSR matrix (A = row offsets, C = columns, V = values) whose
// product apply() runs on the GPU through the C-ABI of libsimplefem_amd.
#pragma once

#include <string>
#include <utility>
#include <vector>

#include "../base/include.hpp"
#include "../operators/floatvector.hpp"
#include "../operators/linearoperator.hpp"
#include "sparsematrix.hpp"

class MatrixCSR : public LinearOperator {
public:
    // from a list of entries (sparse/matcsr.cpp:43-89): an unsorted list is
    // first sorted and compressed in place (SparseMatrix::sortandcompressentries);
    // a sorted one is taken as it is, duplicates and zeros included
    explicit MatrixCSR(const SparseMatrix& matrix);
    explicit MatrixCSR(int rows, int columns, const std::vector<int>& A, const std::vector<int>& C,
                       const std::vector<Float>& V);
    explicit MatrixCSR(int rows, int columns, std::vector<int>&& A, std::vector<int>&& C,
                       std::vector<Float>&& V);
    explicit MatrixCSR(int rows, int columns);  // zero matrix

    MatrixCSR() = delete;
    MatrixCSR(const MatrixCSR&);
    MatrixCSR& operator=(const MatrixCSR&);
    MatrixCSR(MatrixCSR&&) noexcept;
    MatrixCSR& operator=(MatrixCSR&&) noexcept;
    ~MatrixCSR() noexcept override = default;

    void check() const override;
    std::string text() const override;

    using LinearOperator::apply;
    // dest = scaling * M * src (sparse/matcsr.cpp:216-244), on the GPU
    void apply(FloatVector& dest, const FloatVector& add, Float scaling) const override;

    void scale(Float s);
    bool is_finite() const;
    FloatVector getDiagonal() const;
    MatrixCSR getTranspose() const;

    const int* getA() const { return A.data(); }
    const int* getC() const { return C.data(); }
    const Float* getV() const { return V.data(); }
    int getnumberofentries() const { return (int)V.size(); }
    int getnumberofzeroentries() const;
    int getmaxrowwidth() const;
    Float eigenvalueupperbound() const;
    // sort each row by column, merge duplicates, drop zeros (sparse/matcsr.cpp:585-703)
    void compressentries() const;
    std::size_t memorysize() const;

private:
    std::vector<int> A;
    std::vector<int> C;
    std::vector<Float> V;
};

MatrixCSR MatrixCSRAddition(const MatrixCSR& mat1, const MatrixCSR& mat2, Float s1, Float s2);
// mat1 * mat2: all products v1*v2, then sorted and compressed (sparse/matcsr.cpp:767-849)
MatrixCSR MatrixCSRMultiplication(const MatrixCSR& mat1, const MatrixCSR& mat2);
// mat1 * mat2 over nonzero pairs only, summed per output column in loop
// order, then compressed (sparse/matcsr.cpp:857-962)
MatrixCSR MatrixCSRMultiplication_reduced(const MatrixCSR& mat1, const MatrixCSR& mat2);
// factor^T * center * factor without compression: every product
// (v_j * v_i) * v_k as its own entry (sparse/matcsr.cpp:1067-1178)
MatrixCSR Conjugation(const MatrixCSR& center, const MatrixCSR& factor);

inline MatrixCSR operator&(const MatrixCSR& m1, const MatrixCSR& m2) { return MatrixCSRMultiplication(m1, m2); }

inline MatrixCSR operator+(const MatrixCSR& m1, const MatrixCSR& m2) { return MatrixCSRAddition(m1, m2, 1., 1.); }
inline MatrixCSR operator-(const MatrixCSR& m1, const MatrixCSR& m2) { return MatrixCSRAddition(m1, m2, 1., -1.); }
inline MatrixCSR operator*(Float s, const MatrixCSR& mat)
{
    MatrixCSR ret(mat);
    ret.scale(s);
    return ret;
}
inline MatrixCSR operator*(const MatrixCSR& mat, Float s) { return s * mat; }
