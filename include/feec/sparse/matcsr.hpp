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

class MatrixCSR : public LinearOperator {
public:
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

inline MatrixCSR operator+(const MatrixCSR& m1, const MatrixCSR& m2) { return MatrixCSRAddition(m1, m2, 1., 1.); }
inline MatrixCSR operator-(const MatrixCSR& m1, const MatrixCSR& m2) { return MatrixCSRAddition(m1, m2, 1., -1.); }
inline MatrixCSR operator*(Float s, const MatrixCSR& mat)
{
    MatrixCSR ret(mat);
    ret.scale(s);
    return ret;
}
inline MatrixCSR operator*(const MatrixCSR& mat, Float s) { return s * mat; }
