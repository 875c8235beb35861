// sparse_ops_test.cpp -- runs one host-side sparse operation of the C++ drop-in
// on matrices read from stdin and writes the resulting CSR matrix to stdout
// (little-endian binary), for tests/test_dropin_cpu.py.
//
//   in : int32 op; op 0: int32 rows, cols, n; int32 row[n], col[n]; float64 val[n]
//        (MatrixCSR(SparseMatrix)); ops 1..6: two CSR matrices, each int32 rows,
//        cols, nnz; int32 A[rows+1], C[nnz]; float64 V[nnz]; then float64 s1, s2
//        (1 product, 2 reduced product, 3 Conjugation(m1, m2), 4 transpose of
//        m1, 5 MatrixCSRAddition(m1, m2, s1, s2), 6 m1.compressentries())
//   out: int32 rows, cols, nnz; int32 A[rows+1], C[nnz]; float64 V[nnz]
#include <cstdint>
#include <cstdio>
#include <vector>

#include "feec/sparse/matcsr.hpp"
#include "feec/sparse/sparsematrix.hpp"

namespace {

template <typename T>
std::vector<T> get(size_t n)
{
    std::vector<T> v(n);
    if (n && std::fread(v.data(), sizeof(T), n, stdin) != n) throw 1;
    return v;
}

int get1() { return get<int32_t>(1)[0]; }

MatrixCSR get_csr()
{
    const int rows = get1(), cols = get1(), nnz = get1();
    std::vector<int> A = get<int>(rows + 1);
    std::vector<int> C = get<int>(nnz);
    std::vector<double> V = get<double>(nnz);
    return MatrixCSR(rows, cols, A, C, V);
}

template <typename T>
void put(const T* p, size_t n)
{
    if (n) std::fwrite(p, sizeof(T), n, stdout);
}

void put_csr(const MatrixCSR& m)
{
    const int rows = m.getdimout(), cols = m.getdimin(), nnz = m.getA()[rows];
    put(&rows, 1);
    put(&cols, 1);
    put(&nnz, 1);
    put(m.getA(), rows + 1);
    put(m.getC(), nnz);
    put(m.getV(), nnz);
}

}  // namespace

int main()
{
    try {
        const int op = get1();
        if (op == 0) {
            const int rows = get1(), cols = get1(), n = get1();
            std::vector<int> r = get<int>(n), c = get<int>(n);
            std::vector<double> v = get<double>(n);
            SparseMatrix sm(rows, cols);
            for (int i = 0; i < n; i++) sm.appendentry(r[i], c[i], v[i]);
            put_csr(MatrixCSR(sm));
            return 0;
        }
        MatrixCSR m1 = get_csr();
        MatrixCSR m2 = get_csr();
        const double s1 = get<double>(1)[0], s2 = get<double>(1)[0];
        switch (op) {
        case 1: put_csr(MatrixCSRMultiplication(m1, m2)); break;
        case 2: put_csr(MatrixCSRMultiplication_reduced(m1, m2)); break;
        case 3: put_csr(Conjugation(m1, m2)); break;
        case 4: put_csr(m1.getTranspose()); break;
        case 5: put_csr(MatrixCSRAddition(m1, m2, s1, s2)); break;
        case 6: m1.compressentries(); put_csr(m1); break;
        default: return 2;
        }
    } catch (int) {
        std::fprintf(stderr, "short input\n");
        return 3;
    }
    return 0;
}
