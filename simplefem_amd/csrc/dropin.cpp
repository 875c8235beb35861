// dropin.cpp -- C++ drop-in for FEEC++'s sparse solve path (libsimplefem_amd_cxx.so).
//
// FloatVector, MatrixCSR, Rainbow and the six ConjugateGradientSolverCSR*
// functions with the reference's declarations (include/feec/...).  Host data
// stays in host containers as in the reference; products and solves go to the
// GPU through the C-ABI (include/simplefem_amd.h).  The C-ABI status of a call
// is turned into the reference's failure behaviour: a message and abort().
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <memory>
#include <string>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <random>

#include "../../include/feec/operators/floatvector.hpp"
#include "../../include/feec/operators/linearoperator.hpp"
#include "../../include/feec/solver/inv.hpp"
#include "../../include/feec/solver/sparsesolver.hpp"
#include "../../include/feec/solver/systemsparsesolver.hpp"
#include "../../include/feec/sparse/matcsr.hpp"
#include "../../include/feec/sparse/rainbow.hpp"
#include "../../include/feec/sparse/sparsematrix.hpp"

#include <set>
#include "../../include/simplefem_amd.h"

namespace {

void require(int status, const char* what)
{
    if (status == SFA_OK) return;
    std::fprintf(stderr, "%s failed (%d): %s\n", what, status, sfa_last_error());
    std::abort();
}

void require_true(bool cond, const char* what)
{
    if (cond) return;
    std::fprintf(stderr, "check failed: %s\n", what);
    std::abort();
}

// Row-wise sort, duplicate merge and zero removal with the operations of
// sparse/matcsr.cpp:585-687 (so duplicate sums round the same way).
void sort_and_compress(std::vector<int>& A, std::vector<int>& C, std::vector<Float>& V)
{
    const int rows = (int)A.size() - 1;
    std::vector<int> kept(rows);
    for (int r = 0; r < rows; r++) {
        for (int i = A[r]; i < A[r + 1]; i++)
            for (int j = i + 1; j < A[r + 1]; j++) {
                if (C[i] == C[j]) {
                    V[i] += V[j];
                    V[j] = 0.;
                    continue;
                }
                if (C[i] > C[j]) {
                    std::swap(C[i], C[j]);
                    std::swap(V[i], V[j]);
                }
            }
        for (int i = A[r]; i < A[r + 1]; i++)  // zeros to the end, order kept otherwise
            for (int j = A[r] + 1; j < A[r + 1]; j++)
                if (V[j - 1] == 0. && V[j] != 0.) {
                    std::swap(C[j - 1], C[j]);
                    std::swap(V[j - 1], V[j]);
                }
        int k = A[r];
        while (k < A[r + 1] && V[k] != 0.) k++;
        kept[r] = k - A[r];
    }
    int at = 0;
    for (int r = 0; r < rows; r++)
        for (int i = A[r]; i < A[r] + kept[r]; i++) {
            C[at] = C[i];
            V[at] = V[i];
            at++;
        }
    A[0] = 0;
    for (int r = 0; r < rows; r++) A[r + 1] = A[r] + kept[r];
    C.resize(A[rows]);
    V.resize(A[rows]);
}

}  // namespace

// ============================================================== FloatVector

FloatVector::FloatVector(int dim, Float initialvalue) : dimension(dim), pointer(nullptr)
{
    require_true(dim >= 0, "FloatVector dimension >= 0");
    pointer = new Float[dim > 0 ? dim : 1];
    std::fill_n(pointer, dim, initialvalue);
}

FloatVector::FloatVector(const FloatVector& src, Float scaling) : FloatVector(src.dimension)
{
    for (int i = 0; i < dimension; i++) pointer[i] = scaling * src.pointer[i];
}

FloatVector::FloatVector(const std::vector<Float>& v, Float scaling) : FloatVector((int)v.size())
{
    for (int i = 0; i < dimension; i++) pointer[i] = scaling * v[i];
}

FloatVector::FloatVector(const std::vector<int>& v, Float scaling) : FloatVector((int)v.size())
{
    for (int i = 0; i < dimension; i++) pointer[i] = scaling * v[i];
}

FloatVector::FloatVector(const std::initializer_list<Float>& l) : FloatVector((int)l.size())
{
    int i = 0;
    for (Float f : l) pointer[i++] = f;
}

FloatVector::FloatVector(int dim, const std::function<Float(int)>& generator, Float scaling)
    : FloatVector(dim)
{
    for (int i = 0; i < dim; i++) pointer[i] = scaling * generator(i);
}

FloatVector::FloatVector(const FloatVector& src) : FloatVector(src.dimension)
{
    std::copy_n(src.pointer, dimension, pointer);
}

FloatVector::FloatVector(FloatVector&& src) noexcept : dimension(src.dimension), pointer(src.pointer)
{
    src.dimension = 0;
    src.pointer = nullptr;
}

FloatVector& FloatVector::operator=(const FloatVector& vec)
{
    require_true(dimension == vec.dimension, "FloatVector assignment of equal dimension");
    if (this != &vec) std::copy_n(vec.pointer, dimension, pointer);
    return *this;
}

FloatVector& FloatVector::operator=(FloatVector&& vec) noexcept
{
    if (this != &vec) {
        delete[] pointer;
        dimension = vec.dimension;
        pointer = vec.pointer;
        vec.dimension = 0;
        vec.pointer = nullptr;
    }
    return *this;
}

FloatVector::~FloatVector() noexcept { delete[] pointer; }

void FloatVector::check() const { require_true(dimension >= 0, "FloatVector::check"); }

std::string FloatVector::text() const
{
    return "float vector of dimension: " + std::to_string(dimension);
}

std::string FloatVector::data_as_text(bool indexed, bool rowwise) const
{
    std::string s;
    for (int i = 0; i < dimension; i++) {
        if (indexed) s += std::to_string(i) + ": ";
        s += std::to_string(pointer[i]);
        s += rowwise ? " " : "\n";
    }
    return s;
}

Float FloatVector::setentry(int p, Float value) { return pointer[p] = value; }
Float FloatVector::getentry(int p) const { return pointer[p]; }
void FloatVector::setentries(Float v) { std::fill_n(pointer, dimension, v); }

void FloatVector::random()
{
    static thread_local std::default_random_engine engine(0);
    std::normal_distribution<Float> dist(0., 1.);
    for (int i = 0; i < dimension; i++) pointer[i] = dist(engine);
}

void FloatVector::random_within_range(Float min, Float max)
{
    static thread_local std::default_random_engine engine(0);
    std::uniform_real_distribution<Float> dist(min, max);
    for (int i = 0; i < dimension; i++) pointer[i] = dist(engine);
}

void FloatVector::to_absolute()
{
    for (int i = 0; i < dimension; i++) pointer[i] = absolute(pointer[i]);
}

void FloatVector::zero() { std::fill_n(pointer, dimension, 0.); }

FloatVector& FloatVector::normalize()
{
    const Float alpha = norm();
    for (int i = 0; i < dimension; i++) pointer[i] = pointer[i] / alpha;
    return *this;
}

FloatVector& FloatVector::scale(Float f)
{
    for (int i = 0; i < dimension; i++) pointer[i] *= f;
    return *this;
}

FloatVector& FloatVector::scaleinverse(Float d)
{
    for (int i = 0; i < dimension; i++) pointer[i] /= d;
    return *this;
}

FloatVector& FloatVector::shift(Float delta)
{
    for (int i = 0; i < dimension; i++) pointer[i] += delta;
    return *this;
}

FloatVector& FloatVector::shiftnegative(Float delta) { return shift(-delta); }

FloatVector& FloatVector::reciprocal()
{
    for (int i = 0; i < dimension; i++) pointer[i] = 1. / pointer[i];
    return *this;
}

FloatVector FloatVector::getslice(int base, int len) const
{
    require_true(base >= 0 && len >= 0 && base + len <= dimension, "FloatVector::getslice bounds");
    FloatVector ret(len);
    std::copy_n(pointer + base, len, ret.pointer);
    return ret;
}

void FloatVector::setslice(int base, int len, Float v)
{
    require_true(base >= 0 && len >= 0 && base + len <= dimension, "FloatVector::setslice bounds");
    std::fill_n(pointer + base, len, v);
}

void FloatVector::setslice(int base, const FloatVector& src)
{
    require_true(base >= 0 && base + src.dimension <= dimension, "FloatVector::setslice bounds");
    std::copy_n(src.pointer, src.dimension, pointer + base);
}

void FloatVector::addslice(int base, const FloatVector& summand, Float scaling)
{
    require_true(base >= 0 && base + summand.dimension <= dimension, "FloatVector::addslice bounds");
    for (int i = 0; i < summand.dimension; i++) pointer[base + i] += scaling * summand.pointer[i];
}

void FloatVector::copydatafrom(const FloatVector& src) { copydatafrom(1., src); }

void FloatVector::copydatafrom(Float scaling, const FloatVector& src)
{
    require_true(dimension == src.dimension, "copydatafrom of equal dimension");
    for (int i = 0; i < dimension; i++) pointer[i] = scaling * src.pointer[i];
}

void FloatVector::generatedatafrom(const std::function<Float(int)>& g)
{
    for (int i = 0; i < dimension; i++) pointer[i] = g(i);
}

void FloatVector::adddatafrom(const FloatVector& s) { adddatafrom(1., 1., s); }
void FloatVector::adddatafrom(Float ss, const FloatVector& s) { adddatafrom(1., ss, s); }

// operators/floatvector.cpp:554-560: dest = sd * dest + ss * summand
void FloatVector::adddatafrom(Float sd, Float ss, const FloatVector& s)
{
    require_true(dimension == s.dimension, "adddatafrom of equal dimension");
    for (int i = 0; i < dimension; i++) pointer[i] = sd * pointer[i] + ss * s.pointer[i];
}

// long double accumulation (operators/floatvector.cpp:563-571)
Float FloatVector::scalarproductwith(const FloatVector& right) const
{
    require_true(dimension == right.dimension, "scalar product of equal dimension");
    long double ret = 0.;
    for (int i = 0; i < dimension; i++) ret += pointer[i] * right.pointer[i];
    return (Float)ret;
}

Float FloatVector::min() const
{
    Float ret = pointer[0];
    for (int i = 1; i < dimension; i++) ret = (ret < pointer[i]) ? ret : pointer[i];
    return ret;
}

Float FloatVector::average() const { return sum() / dimension; }

Float FloatVector::sum() const
{
    long double ret = 0.;
    for (int i = 0; i < dimension; i++) ret += pointer[i];
    return (Float)ret;
}

Float FloatVector::norm() const { return std::sqrt(norm_sq()); }

Float FloatVector::norm_sq() const
{
    long double ret = 0.;
    for (int i = 0; i < dimension; i++) ret += pointer[i] * pointer[i];
    return (Float)ret;
}

Float FloatVector::norm(const LinearOperator& op) const { return std::sqrt(norm_sq(op)); }

Float FloatVector::norm_sq(const LinearOperator& op) const
{
    long double ret = (*this) * (op * (*this));
    return (Float)ret;
}

Float FloatVector::maxnorm() const
{
    Float ret = 0.;
    for (int i = 0; i < dimension; i++) ret = maximum(ret, absolute(pointer[i]));
    return ret;
}

Float FloatVector::sumnorm() const
{
    long double ret = 0.;
    for (int i = 0; i < dimension; i++) ret = ret + absolute(pointer[i]);
    return (Float)ret;
}

bool FloatVector::is_finite() const
{
    for (int i = 0; i < dimension; i++)
        if (!std::isfinite(pointer[i])) return false;
    return true;
}

bool FloatVector::is_zero() const
{
    for (int i = 0; i < dimension; i++)
        if (pointer[i] != 0.) return false;
    return true;
}

bool FloatVector::is_positive() const
{
    for (int i = 0; i < dimension; i++)
        if (!(pointer[i] > 0.)) return false;
    return true;
}

bool FloatVector::is_negative() const
{
    for (int i = 0; i < dimension; i++)
        if (!(pointer[i] < 0.)) return false;
    return true;
}

bool FloatVector::is_nonnegative() const
{
    for (int i = 0; i < dimension; i++)
        if (!(pointer[i] >= 0.)) return false;
    return true;
}

bool FloatVector::is_nonpositive() const
{
    for (int i = 0; i < dimension; i++)
        if (!(pointer[i] <= 0.)) return false;
    return true;
}

bool FloatVector::is_numerically_small(Float threshold) const { return norm() < threshold; }

bool FloatVector::is_equal_to(const FloatVector& l, const FloatVector& r)
{
    require_true(l.dimension == r.dimension, "comparison of equal dimension");
    for (int i = 0; i < l.dimension; i++)
        if (l.pointer[i] != r.pointer[i]) return false;
    return true;
}

// ============================================================== MatrixCSR

MatrixCSR::MatrixCSR(int rows, int columns, const std::vector<int>& A_, const std::vector<int>& C_,
                     const std::vector<Float>& V_)
    : LinearOperator(rows, columns), A(A_), C(C_), V(V_)
{
    check();
}

MatrixCSR::MatrixCSR(int rows, int columns, std::vector<int>&& A_, std::vector<int>&& C_,
                     std::vector<Float>&& V_)
    : LinearOperator(rows, columns), A(std::move(A_)), C(std::move(C_)), V(std::move(V_))
{
    check();
}

MatrixCSR::MatrixCSR(int rows, int columns)
    : LinearOperator(rows, columns), A(rows + 1, 0), C(), V()
{
    check();
}

MatrixCSR::MatrixCSR(const MatrixCSR& m) : LinearOperator(m.getdimout(), m.getdimin()), A(m.A), C(m.C), V(m.V) {}

MatrixCSR& MatrixCSR::operator=(const MatrixCSR& m)
{
    require_true(getdimin() == m.getdimin() && getdimout() == m.getdimout(), "MatrixCSR assignment");
    if (this != &m) {
        A = m.A;
        C = m.C;
        V = m.V;
    }
    return *this;
}

MatrixCSR::MatrixCSR(MatrixCSR&& m) noexcept
    : LinearOperator(m.getdimout(), m.getdimin()), A(std::move(m.A)), C(std::move(m.C)), V(std::move(m.V))
{
}

MatrixCSR& MatrixCSR::operator=(MatrixCSR&& m) noexcept
{
    if (this != &m) {
        A = std::move(m.A);
        C = std::move(m.C);
        V = std::move(m.V);
    }
    return *this;
}

// sparse/matcsr.cpp:178-194
void MatrixCSR::check() const
{
    require_true((int)A.size() == getdimout() + 1, "MatrixCSR: A has rows+1 entries");
    require_true(C.size() == V.size(), "MatrixCSR: C and V have equal length");
    for (int p = 1; p <= getdimout(); p++) require_true(A[p - 1] <= A[p], "MatrixCSR: A ascending");
    require_true(A[getdimout()] == (int)V.size(), "MatrixCSR: A ends at nnz");
    for (size_t i = 0; i < C.size(); i++)
        require_true(0 <= C[i] && C[i] < getdimin() && std::isfinite(V[i]), "MatrixCSR: entries");
}

std::string MatrixCSR::text() const
{
    return "CSRMatrix " + std::to_string(getdimout()) + "x" + std::to_string(getdimin()) + " with " +
           std::to_string(V.size()) + " entries";
}

void MatrixCSR::apply(FloatVector& dest, const FloatVector& add, Float scaling) const
{
    require_true(getdimin() == add.getdimension() && getdimout() == dest.getdimension(),
                 "MatrixCSR::apply dimensions");
    require_true(&dest != &add, "MatrixCSR::apply: dest and source differ");
    if (getdimout() == 0) return;
    if (getdimin() == 0) {
        dest.zero();
        return;
    }
    require(sfa_MatrixCSR_apply(getdimout(), getdimin(), A.data(), C.data(), V.data(), dest.raw(),
                                add.raw(), scaling),
            "MatrixCSR::apply");
}

void MatrixCSR::scale(Float s)
{
    for (auto& v : V) v *= s;
}

bool MatrixCSR::is_finite() const
{
    for (Float v : V)
        if (!std::isfinite(v)) return false;
    return true;
}

FloatVector MatrixCSR::getDiagonal() const
{
    require_true(getdimin() == getdimout(), "getDiagonal of a square matrix");
    FloatVector ret(getdimin(), 0.);
    if (getdimin() > 0)
        require(sfa_MatrixCSR_getDiagonal(getdimin(), A.data(), C.data(), V.data(), ret.raw()),
                "MatrixCSR::getDiagonal");
    return ret;
}

// sparse/matcsr.cpp:291-350: counting transpose, then sort and compress.
MatrixCSR MatrixCSR::getTranspose() const
{
    const int rows = getdimout(), cols = getdimin();
    const int nnz = (int)V.size();
    std::vector<int> B(cols + 1, 0), D(nnz, 0), Z(cols, 0);
    std::vector<Float> W(nnz, 0.);
    for (int i = 0; i < nnz; i++) Z[C[i]]++;
    for (int c = 1; c <= cols; c++) B[c] = B[c - 1] + Z[c - 1];
    for (int r = 0; r < rows; r++)
        for (int i = A[r]; i < A[r + 1]; i++) {
            const int c = C[i];
            const int at = B[c] + (--Z[c]);
            D[at] = r;
            W[at] = V[i];
        }
    sort_and_compress(B, D, W);
    return MatrixCSR(cols, rows, std::move(B), std::move(D), std::move(W));
}

int MatrixCSR::getnumberofzeroentries() const
{
    int ret = 0;
    for (Float v : V)
        if (v == 0.) ret++;
    return ret;
}

int MatrixCSR::getmaxrowwidth() const
{
    int ret = 0;
    for (int r = 0; r < getdimout(); r++) ret = std::max(ret, A[r + 1] - A[r]);
    return ret;
}

Float MatrixCSR::eigenvalueupperbound() const
{
    Float ret = 0.;
    for (int r = 0; r < getdimout(); r++) {
        Float row = 0.;
        for (int c = A[r]; c < A[r + 1]; c++) row += absolute(V[c]);
        ret = maximum(ret, row);
    }
    return ret;
}

void MatrixCSR::compressentries() const
{
    auto& a = const_cast<std::vector<int>&>(A);
    auto& c = const_cast<std::vector<int>&>(C);
    auto& v = const_cast<std::vector<Float>&>(V);
    sort_and_compress(a, c, v);
}

std::size_t MatrixCSR::memorysize() const
{
    return sizeof(*this) + A.size() * sizeof(int) + C.size() * sizeof(int) + V.size() * sizeof(Float);
}

// sparse/matcsr.cpp:707-765: concatenate the rows, then sort and compress.
MatrixCSR MatrixCSRAddition(const MatrixCSR& m1, const MatrixCSR& m2, Float s1, Float s2)
{
    require_true(m1.getdimout() == m2.getdimout() && m1.getdimin() == m2.getdimin(),
                 "MatrixCSRAddition dimensions");
    const int rows = m1.getdimout();
    const int* A1 = m1.getA();
    const int* A2 = m2.getA();
    std::vector<int> A(rows + 1, 0);
    for (int r = 1; r <= rows; r++) A[r] = A[r - 1] + (A1[r] - A1[r - 1]) + (A2[r] - A2[r - 1]);
    std::vector<int> C(A[rows]);
    std::vector<Float> V(A[rows]);
    for (int r = 0; r < rows; r++) {
        int at = A[r];
        for (int k = A1[r]; k < A1[r + 1]; k++, at++) C[at] = m1.getC()[k], V[at] = s1 * m1.getV()[k];
        for (int k = A2[r]; k < A2[r + 1]; k++, at++) C[at] = m2.getC()[k], V[at] = s2 * m2.getV()[k];
    }
    sort_and_compress(A, C, V);
    return MatrixCSR(rows, m1.getdimin(), std::move(A), std::move(C), std::move(V));
}

// Products of at least this many pairs of entries run on the GPU
// (sfa_MatrixCSRMultiplication, sfa_MatrixCSRMultiplication_reduced): the
// same products and the same compression, each row's run by one GPU thread,
// so the result is the same array for array.  The compression's work grows
// with the square of a row's product count and one GPU thread takes a step
// many times slower than a host core, so the GPU pays only when that work is
// spread over many rows: the sum over rows of (products)^2 must be at least
// kSpread times the largest row's.
constexpr long long kDeviceProducts = 1LL << 19;
constexpr double kSpread = 64.;

static bool product_on_device(const MatrixCSR& m1, const MatrixCSR& m2)
{
    const int *A1 = m1.getA(), *C1 = m1.getC(), *A2 = m2.getA();
    long long products = 0;
    double sum_sq = 0., max_sq = 0.;
    for (int r = 0; r < m1.getdimout(); r++) {
        long long row = 0;
        for (int i = A1[r]; i < A1[r + 1]; i++) row += A2[C1[i] + 1] - A2[C1[i]];
        products += row;
        const double sq = (double)row * (double)row;
        sum_sq += sq;
        max_sq = std::max(max_sq, sq);
    }
    return products >= kDeviceProducts && sum_sq >= kSpread * max_sq;
}

static MatrixCSR device_product(const MatrixCSR& m1, const MatrixCSR& m2, bool reduced,
                                const char* what)
{
    auto fn = reduced ? sfa_MatrixCSRMultiplication_reduced : sfa_MatrixCSRMultiplication;
    sfa_csr h = nullptr;
    require(fn(m1.getdimout(), m1.getdimin(), m1.getA(), m1.getC(), m1.getV(), m2.getdimout(),
               m2.getdimin(), m2.getA(), m2.getC(), m2.getV(), &h),
            what);
    int r = 0, c = 0, nnz = 0;
    require(sfa_csr_sizes(h, &r, &c, &nnz), what);
    std::vector<int> A(r + 1), C(nnz);
    std::vector<Float> V(nnz);
    require(sfa_csr_copy(h, A.data(), C.data(), V.data()), what);
    sfa_csr_destroy(h);
    return MatrixCSR(r, c, std::move(A), std::move(C), std::move(V));
}

// mat1 * mat2 (sparse/matcsr.cpp:767-849): row r collects, for each entry
// (r, c1) of mat1 and each entry (c1, c2) of mat2 in order, the product
// v1 * v2 at column c2; sort_and_compress then merges and drops zeros.
MatrixCSR MatrixCSRMultiplication(const MatrixCSR& m1, const MatrixCSR& m2)
{
    require_true(m1.getdimin() == m2.getdimout(), "MatrixCSRMultiplication dimensions");
    const int rows = m1.getdimout();
    const int *A1 = m1.getA(), *C1 = m1.getC(), *A2 = m2.getA(), *C2 = m2.getC();
    const Float *V1 = m1.getV(), *V2 = m2.getV();
    if (product_on_device(m1, m2)) return device_product(m1, m2, false, "MatrixCSRMultiplication");
    std::vector<int> A(rows + 1, 0);
    for (int r = 0; r < rows; r++) {
        int count = 0;
        for (int i = A1[r]; i < A1[r + 1]; i++) count += A2[C1[i] + 1] - A2[C1[i]];
        A[r + 1] = A[r] + count;
    }
    std::vector<int> C(A[rows]);
    std::vector<Float> V(A[rows]);
    for (int r = 0; r < rows; r++) {
        int at = A[r];
        for (int i = A1[r]; i < A1[r + 1]; i++)
            for (int k = A2[C1[i]]; k < A2[C1[i] + 1]; k++, at++) {
                C[at] = C2[k];
                V[at] = V1[i] * V2[k];
            }
    }
    sort_and_compress(A, C, V);
    return MatrixCSR(rows, m2.getdimin(), std::move(A), std::move(C), std::move(V));
}

// sparse/matcsr.cpp:857-962: the output pattern of row r is the set of
// columns c2 reached through a pair of nonzero values; each output value
// starts at 0 and adds v1 * v2 over those pairs in (c1, c2) loop order.
MatrixCSR MatrixCSRMultiplication_reduced(const MatrixCSR& m1, const MatrixCSR& m2)
{
    require_true(m1.getdimin() == m2.getdimout(), "MatrixCSRMultiplication_reduced dimensions");
    if (product_on_device(m1, m2))
        return device_product(m1, m2, true, "MatrixCSRMultiplication_reduced");
    const int rows = m1.getdimout();
    const int *A1 = m1.getA(), *C1 = m1.getC(), *A2 = m2.getA(), *C2 = m2.getC();
    const Float *V1 = m1.getV(), *V2 = m2.getV();
    std::vector<int> A(rows + 1, 0);
    std::vector<int> C;
    std::vector<Float> V;
    std::vector<int> at(m2.getdimin(), -1);  // output position of a column in the current row
    for (int r = 0; r < rows; r++) {
        std::set<int> cols;
        for (int i = A1[r]; i < A1[r + 1]; i++)
            for (int k = A2[C1[i]]; k < A2[C1[i] + 1]; k++)
                if (V1[i] != 0. && V2[k] != 0.) cols.insert(C2[k]);
        const int base = (int)C.size();
        for (int c : cols) {
            at[c] = (int)C.size();
            C.push_back(c);
            V.push_back(0.);
        }
        for (int i = A1[r]; i < A1[r + 1]; i++)
            for (int k = A2[C1[i]]; k < A2[C1[i] + 1]; k++)
                if (V1[i] != 0. && V2[k] != 0.) V[at[C2[k]]] += V1[i] * V2[k];
        A[r + 1] = (int)C.size();
        for (int p = base; p < (int)C.size(); p++) at[C[p]] = -1;
    }
    sort_and_compress(A, C, V);
    return MatrixCSR(rows, m2.getdimin(), std::move(A), std::move(C), std::move(V));
}

// sparse/matcsr.cpp:1067-1178: for each row r of center and factor, each
// entry (r, ci) of center, (r, cj) of factor and (ci, ck) of factor gives the
// entry (cj, ck) with value (vj * vi) * vk; entries of an output row are
// placed from the end of its range backwards, in that loop order, and are
// neither sorted nor merged.
MatrixCSR Conjugation(const MatrixCSR& center, const MatrixCSR& factor)
{
    require_true(center.is_square() && center.getdimin() == factor.getdimout(),
                 "Conjugation dimensions");
    const int N = center.getdimout();
    const int L = factor.getdimin();
    const int *cA = center.getA(), *cC = center.getC(), *fA = factor.getA(), *fC = factor.getC();
    const Float *cV = center.getV(), *fV = factor.getV();
    std::vector<int> left(L, 0);  // entries still to place in each output row
    for (int r = 0; r < N; r++) {
        int per = 0;
        for (int i = cA[r]; i < cA[r + 1]; i++) per += fA[cC[i] + 1] - fA[cC[i]];
        for (int j = fA[r]; j < fA[r + 1]; j++) left[fC[j]] += per;
    }
    std::vector<int> A(L + 1, 0);
    for (int c = 0; c < L; c++) A[c + 1] = A[c] + left[c];
    std::vector<int> C(A[L], -1);
    std::vector<Float> V(A[L], notanumber);
    for (int r = 0; r < N; r++)
        for (int i = cA[r]; i < cA[r + 1]; i++)
            for (int j = fA[r]; j < fA[r + 1]; j++) {
                const int cj = fC[j];
                const Float vjvi = fV[j] * cV[i];
                const int ci = cC[i];
                for (int k = fA[ci]; k < fA[ci + 1]; k++) {
                    const int p = A[cj] + (--left[cj]);
                    C[p] = fC[k];
                    V[p] = vjvi * fV[k];
                }
            }
    return MatrixCSR(L, L, std::move(A), std::move(C), std::move(V));
}

// ============================================================== SparseMatrix

SparseMatrix::SparseMatrix(int dimout, int dimin, int numentries,
                           const std::function<MatrixEntry(int)>& generator)
    : LinearOperator(dimout, dimin), entries()
{
    require_true(numentries >= 0, "SparseMatrix: numentries >= 0");
    entries.reserve(numentries);
    for (int i = 0; i < numentries; i++) appendentry(generator(i));
}

SparseMatrix::SparseMatrix(int dimout, int dimin, const std::vector<MatrixEntry>& list)
    : LinearOperator(dimout, dimin), entries()
{
    for (const MatrixEntry& e : list) appendentry(e);
}

SparseMatrix::SparseMatrix(int dimout, int dimin, const std::initializer_list<MatrixEntry>& list)
    : LinearOperator(dimout, dimin), entries()
{
    for (const MatrixEntry& e : list) appendentry(e);
}

void SparseMatrix::check() const
{
    for (const MatrixEntry& e : entries)
        require_true(0 <= e.row && e.row < getdimout() && 0 <= e.column && e.column < getdimin(),
                     "SparseMatrix: entry in range");
}

std::string SparseMatrix::text() const
{
    std::string ret = "SparseMatrix " + std::to_string(getdimout()) + "x" + std::to_string(getdimin());
    for (const MatrixEntry& e : entries)
        ret += "\n" + std::to_string(e.row) + " " + std::to_string(e.column) + " : " +
               std::to_string(e.value);
    return ret;
}

void SparseMatrix::apply(FloatVector& dest, const FloatVector& add, Float scaling) const
{
    require_true(getdimin() == add.getdimension() && getdimout() == dest.getdimension(),
                 "SparseMatrix::apply dimensions");
    dest.zero();
    for (const MatrixEntry& e : entries) dest[e.row] += scaling * e.value * add[e.column];
}

void SparseMatrix::scale(Float s)
{
    for (MatrixEntry& e : entries) e.value *= s;
}

bool SparseMatrix::is_finite() const
{
    for (const MatrixEntry& e : entries)
        if (!std::isfinite(e.value)) return false;
    return true;
}

bool SparseMatrix::is_sorted(MatrixEntrySorting manner) const
{
    const bool rowwise = manner == MatrixEntrySorting::rowwise;
    for (size_t i = 1; i < entries.size(); i++) {
        const MatrixEntry& a = entries[i - 1];
        const MatrixEntry& b = entries[i];
        const int a1 = rowwise ? a.row : a.column, a2 = rowwise ? a.column : a.row;
        const int b1 = rowwise ? b.row : b.column, b2 = rowwise ? b.column : b.row;
        if (a1 > b1 || (a1 == b1 && a2 > b2)) return false;
    }
    return true;
}

// The standard library's (unstable) sort with the strict (major, minor)
// order: entries with equal keys end up in the order this sort leaves them,
// which the compression then sums in -- the same sort over the same input as
// the reference's, so the same sums.
const SparseMatrix& SparseMatrix::sortentries(MatrixEntrySorting manner) const
{
    if (manner == MatrixEntrySorting::rowwise)
        std::sort(entries.begin(), entries.end(), [](const MatrixEntry& a, const MatrixEntry& b) {
            return a.row < b.row || (a.row == b.row && a.column < b.column);
        });
    else
        std::sort(entries.begin(), entries.end(), [](const MatrixEntry& a, const MatrixEntry& b) {
            return a.column < b.column || (a.column == b.column && a.row < b.row);
        });
    return *this;
}

const SparseMatrix& SparseMatrix::sortandcompressentries(MatrixEntrySorting manner) const
{
    entries.erase(std::remove_if(entries.begin(), entries.end(),
                                 [](const MatrixEntry& e) { return e.value == 0.; }),
                  entries.end());
    sortentries(manner);
    if (entries.empty()) {
        // the reference's compaction always keeps dest+1 entries: an empty
        // list becomes one value-initialised entry (0, 0, 0.)
        // (sparse/sparsematrix.cpp:438-458)
        if (getdimout() > 0 && getdimin() > 0) entries.resize(1, MatrixEntry{0, 0, 0.});
        return *this;
    }
    size_t last = 0;  // the entry that equal keys are added into
    for (size_t i = 1; i < entries.size(); i++) {
        if (entries[i].row == entries[last].row && entries[i].column == entries[last].column)
            entries[last].value += entries[i].value;
        else
            entries[++last] = entries[i];
    }
    entries.resize(last + 1);
    return *this;
}

void SparseMatrix::setentry(int i, MatrixEntry entry)
{
    require_true(0 <= i && i < (int)entries.size(), "SparseMatrix::setentry index");
    entries[i] = entry;
}

void SparseMatrix::appendentry(MatrixEntry entry)
{
    require_true(0 <= entry.row && entry.row < getdimout() && 0 <= entry.column &&
                     entry.column < getdimin(),
                 "SparseMatrix: entry in range");
    entries.push_back(entry);
}

// sparse/matcsr.cpp:43-89
MatrixCSR::MatrixCSR(const SparseMatrix& matrix)
    : LinearOperator(matrix.getdimout(), matrix.getdimin()), A(), C(), V()
{
    matrix.check();
    if (!matrix.is_sorted()) matrix.sortandcompressentries();
    const auto& list = matrix.getentries();
    const int rows = matrix.getdimout();
    A.assign(rows + 1, 0);
    C.resize(list.size());
    V.resize(list.size());
    for (size_t i = 0; i < list.size(); i++) {
        A[list[i].row + 1]++;
        C[i] = list[i].column;
        V[i] = list[i].value;
    }
    for (int r = 1; r <= rows; r++) A[r] += A[r - 1];
}

// ============================================================== Rainbow

Rainbow::Rainbow(const MatrixCSR& mat, bool do_shuffle)
    : num_rows(mat.getdimout()), num_colors(0), F(), B(), R()
{
    require_true(mat.is_square(), "Rainbow of a square matrix");
    require_true(!do_shuffle, "Rainbow: the shuffled painting order is not supported");
    const int n = num_rows;
    F.assign(n, 0);
    R.assign(n, 0);
    std::vector<int> Bfull(n + 1, 0);
    if (n > 0) require(sfa_Rainbow(n, mat.getA(), mat.getC(), F.data(), Bfull.data(), R.data(), &num_colors),
                       "Rainbow");
    else num_colors = 1;
    B.assign(Bfull.begin(), Bfull.begin() + num_colors + 1);
}

// sparse/rainbow.cpp:133-152
void Rainbow::check() const
{
    for (int c = 0; c < num_colors; c++)
        for (int i = B[c]; i < B[c + 1]; i++) require_true(F[R[i]] == c, "Rainbow: F agrees with B/R");
}

std::string Rainbow::text() const { return "num_colors: " + std::to_string(num_colors); }

// ============================================================== solvers

int ConjugateGradientSolverCSR(const int N, Float* x, const Float* b, const int* csrrows,
                               const int* csrcolumns, const Float* csrvalues, Float* residual,
                               Float precision, int print_modulo)
{
    sfa_solve_info info{};
    require(sfa_ConjugateGradientSolverCSR(N, x, b, csrrows, csrcolumns, csrvalues, residual,
                                           precision, print_modulo, &info),
            "ConjugateGradientSolverCSR");
    return info.iterations;
}

int ConjugateGradientSolverCSR_DiagonalPreconditioner(const int N, Float* x, const Float* b,
                                                      const int* csrrows, const int* csrcolumns,
                                                      const Float* csrvalues, Float* residual,
                                                      Float precision, int print_modulo,
                                                      const Float* precon)
{
    sfa_solve_info info{};
    require(sfa_ConjugateGradientSolverCSR_DiagonalPreconditioner(
                N, x, b, csrrows, csrcolumns, csrvalues, residual, precision, print_modulo, precon,
                &info),
            "ConjugateGradientSolverCSR_DiagonalPreconditioner");
    return info.iterations;
}

int ConjugateGradientSolverCSR_SSOR(const int N, Float* x, const Float* b, const int* csrrows,
                                    const int* csrcolumns, const Float* csrvalues, Float* residual,
                                    Float precision, int print_modulo, const Float* diagonal,
                                    Float omega)
{
    sfa_solve_info info{};
    require(sfa_ConjugateGradientSolverCSR_SSOR(N, x, b, csrrows, csrcolumns, csrvalues, residual,
                                                precision, print_modulo, diagonal, omega, &info),
            "ConjugateGradientSolverCSR_SSOR");
    return info.iterations;
}

int HodgeConjugateResidualSolverCSR_SSOR(
    const int N, const int L, Float* x, const Float* b, const int* Arows, const int* Acolumns,
    const Float* Avalues, const int* Brows, const int* Bcolumns, const Float* Bvalues,
    const int* Btrows, const int* Btcolumns, const Float* Btvalues, const int* Crows,
    const int* Ccolumns, const Float* Cvalues, Float* residual, Float precision, int print_modulo,
    Float inneriteration_precision, int inneriteration_print_modulo)
{
    sfa_solve_info info{};
    require(sfa_HodgeConjugateResidualSolverCSR_SSOR(
                N, L, x, b, Arows, Acolumns, Avalues, Brows, Bcolumns, Bvalues, Btrows, Btcolumns,
                Btvalues, Crows, Ccolumns, Cvalues, residual, precision, print_modulo,
                inneriteration_precision, inneriteration_print_modulo, &info),
            "HodgeConjugateResidualSolverCSR_SSOR");
    return info.iterations;
}

int HodgeConjugateResidualSolverCSR(
    const int N, const int L, Float* x, const Float* b, const int* Arows, const int* Acolumns,
    const Float* Avalues, const int* Brows, const int* Bcolumns, const Float* Bvalues,
    const int* Btrows, const int* Btcolumns, const Float* Btvalues, const int* Crows,
    const int* Ccolumns, const Float* Cvalues, Float* residual, Float precision, int print_modulo,
    Float inneriteration_precision, int inneriteration_print_modulo)
{
    sfa_solve_info info{};
    require(sfa_HodgeConjugateResidualSolverCSR(
                N, L, x, b, Arows, Acolumns, Avalues, Brows, Bcolumns, Bvalues, Btrows, Btcolumns,
                Btvalues, Crows, Ccolumns, Cvalues, residual, precision, print_modulo,
                inneriteration_precision, inneriteration_print_modulo, &info),
            "HodgeConjugateResidualSolverCSR");
    return info.iterations;
}

int HodgeConjugateResidualSolverCSR_diagonal(
    const int N, const int L, Float* x, const Float* b, const int* Arows, const int* Acolumns,
    const Float* Avalues, const int* Brows, const int* Bcolumns, const Float* Bvalues,
    const int* Btrows, const int* Btcolumns, const Float* Btvalues, const int* Crows,
    const int* Ccolumns, const Float* Cvalues, Float* residual, Float precision, int print_modulo,
    Float inneriteration_precision, int inneriteration_print_modulo)
{
    sfa_solve_info info{};
    require(sfa_HodgeConjugateResidualSolverCSR_diagonal(
                N, L, x, b, Arows, Acolumns, Avalues, Brows, Bcolumns, Bvalues, Btrows, Btcolumns,
                Btvalues, Crows, Ccolumns, Cvalues, residual, precision, print_modulo,
                inneriteration_precision, inneriteration_print_modulo, &info),
            "HodgeConjugateResidualSolverCSR_diagonal");
    return info.iterations;
}

int HodgeConjugateResidualSolverCSR_textbook(
    const int N, const int L, Float* x, const Float* b, const int* Arows, const int* Acolumns,
    const Float* Avalues, const int* Brows, const int* Bcolumns, const Float* Bvalues,
    const int* Btrows, const int* Btcolumns, const Float* Btvalues, const int* Crows,
    const int* Ccolumns, const Float* Cvalues, Float* residual, Float precision, int print_modulo,
    Float inneriteration_precision, int inneriteration_print_modulo)
{
    sfa_solve_info info{};
    require(sfa_HodgeConjugateResidualSolverCSR_textbook(
                N, L, x, b, Arows, Acolumns, Avalues, Brows, Bcolumns, Bvalues, Btrows, Btcolumns,
                Btvalues, Crows, Ccolumns, Cvalues, residual, precision, print_modulo,
                inneriteration_precision, inneriteration_print_modulo, &info),
            "HodgeConjugateResidualSolverCSR_textbook");
    return info.iterations;
}

int ConjugateGradientSolverCSR_SSOR_Eisenstat(const int N, Float* x, const Float* b,
                                              const int* csrrows, const int* csrcolumns,
                                              const Float* csrvalues, Float* residual,
                                              Float precision, int print_modulo,
                                              const Float* diagonal, Float omega)
{
    sfa_solve_info info{};
    require(sfa_ConjugateGradientSolverCSR_SSOR_Eisenstat(N, x, b, csrrows, csrcolumns, csrvalues,
                                                          residual, precision, print_modulo,
                                                          diagonal, omega, &info),
            "ConjugateGradientSolverCSR_SSOR_Eisenstat");
    return info.iterations;
}

int ConjugateGradientSolverCSR_Rainbow(const int N, Float* x, const Float* b, const int* csrrows,
                                       const int* csrcolumns, const Float* csrvalues,
                                       Float* residual, Float precision, int print_modulo,
                                       const Float* diagonal, Float omega, int num_colors,
                                       const int* F, const int* B, const int* R)
{
    sfa_solve_info info{};
    require(sfa_ConjugateGradientSolverCSR_Rainbow(N, x, b, csrrows, csrcolumns, csrvalues,
                                                   residual, precision, print_modulo, diagonal,
                                                   omega, num_colors, F, B, R, &info),
            "ConjugateGradientSolverCSR_Rainbow");
    return info.iterations;
}

int ConjugateGradientSolverCSR_Eisenstat_Rainbow(const int N, Float* x, const Float* b,
                                                 const int* csrrows, const int* csrcolumns,
                                                 const Float* csrvalues, Float* residual,
                                                 Float precision, int print_modulo,
                                                 const Float* diagonal, Float omega,
                                                 int num_colors, const int* F, const int* B,
                                                 const int* R)
{
    sfa_solve_info info{};
    require(sfa_ConjugateGradientSolverCSR_Eisenstat_Rainbow(
                N, x, b, csrrows, csrcolumns, csrvalues, residual, precision, print_modulo,
                diagonal, omega, num_colors, F, B, R, &info),
            "ConjugateGradientSolverCSR_Eisenstat_Rainbow");
    return info.iterations;
}

// ------------------------------------------------------------ InverseOperator

struct InverseOperator<MatrixCSR>::Device {
    sfa_system sys = nullptr;
    // the arrays the device copy was built from, to see whether the operator
    // has changed since (compared with memcmp: exact, at memory bandwidth)
    int rows = -1, cols = -1;
    std::vector<int> A, C;
    std::vector<Float> V;
    ~Device()
    {
        if (sys) (void)sfa_system_destroy(sys);
    }
    bool same_as(const MatrixCSR& m) const
    {
        const int r = m.getdimout(), c = m.getdimin(), nnz = m.getnumberofentries();
        return r == rows && c == cols && (size_t)nnz == C.size() &&
               std::memcmp(A.data(), m.getA(), sizeof(int) * (size_t)(r + 1)) == 0 &&
               std::memcmp(C.data(), m.getC(), sizeof(int) * (size_t)nnz) == 0 &&
               std::memcmp(V.data(), m.getV(), sizeof(Float) * (size_t)nnz) == 0;
    }
    void keep(const MatrixCSR& m)
    {
        rows = m.getdimout();
        cols = m.getdimin();
        const int nnz = m.getnumberofentries();
        A.assign(m.getA(), m.getA() + rows + 1);
        C.assign(m.getC(), m.getC() + nnz);
        V.assign(m.getV(), m.getV() + nnz);
    }
};

InverseOperator<MatrixCSR>::InverseOperator(const MatrixCSR& op_, Float precision_, int print_modulo_)
    : LinearOperator(op_.getdimout(), op_.getdimin()),
      op(op_),
      precision(precision_),
      print_modulo(print_modulo_),
      previous_sol(op_.getdimin(), 0.)
{
    require_true(op.getdimin() == op.getdimout(), "InverseOperator of a non-square operator");
    if (print_modulo >= 0) LOG << "Inverse created" << nl;
}

InverseOperator<MatrixCSR>::InverseOperator(const InverseOperator& o)
    : LinearOperator(o.getdimout(), o.getdimin()),
      use_previous_sol(o.use_previous_sol),
      op(o.op),
      precision(o.precision),
      print_modulo(o.print_modulo),
      previous_sol(o.previous_sol)
{
}

InverseOperator<MatrixCSR>::InverseOperator(InverseOperator&& o) noexcept
    : LinearOperator(o.getdimout(), o.getdimin()),
      use_previous_sol(o.use_previous_sol),
      op(o.op),
      precision(o.precision),
      print_modulo(o.print_modulo),
      previous_sol(std::move(o.previous_sol)),
      device_(std::move(o.device_)),
      last_iterations_(o.last_iterations_)
{
}

InverseOperator<MatrixCSR>::~InverseOperator() noexcept = default;

void InverseOperator<MatrixCSR>::check() const
{
    op.check();
    previous_sol.check();
    require_true(op.getdimin() == op.getdimout(), "InverseOperator of a non-square operator");
    require_true(op.getdimin() == previous_sol.getdimension(), "InverseOperator dimension");
}

std::string InverseOperator<MatrixCSR>::text() const
{
    std::string inner = op.text(), out;
    size_t at = 0;
    while (at < inner.size()) {  // tab_each_line
        const size_t nl_at = inner.find('\n', at);
        const size_t end = nl_at == std::string::npos ? inner.size() : nl_at + 1;
        out += "\t" + inner.substr(at, end - at);
        at = end;
    }
    return "Inverse Operator " + std::to_string(getdimout()) + "x" + std::to_string(getdimin()) +
           "\n" + out;
}

void InverseOperator<MatrixCSR>::apply(FloatVector& dest, const FloatVector& src, Float scaling) const
{
    check();
    src.check();
    dest.check();
    require_true(getdimin() == src.getdimension(), "InverseOperator::apply: src dimension");
    require_true(getdimout() == dest.getdimension(), "InverseOperator::apply: dest dimension");
    if (use_previous_sol) dest.copydatafrom(previous_sol);
    else dest.zero();
    // the reference reads the operator at every apply: rebuild the device copy
    // when its arrays have changed
    if (!device_ || !device_->same_as(op)) {
        device_.reset();
        std::unique_ptr<Device> dev(new Device());
        const FloatVector diagonal = op.getDiagonal();
        const Rainbow rainbow(op);
        require(sfa_system_create(SFA_METHOD_EISENSTAT_RAINBOW, op.getdimin(), op.getA(), op.getC(),
                                  op.getV(), diagonal.raw(), 1.0, rainbow.num_colors,
                                  rainbow.F.data(), rainbow.B.data(), rainbow.R.data(), &dev->sys),
                "InverseOperator: sfa_system_create");
        dev->keep(op);
        device_ = std::move(dev);
    }
    FloatVector res(dest);
    sfa_solve_info info{};
    require(sfa_system_solve(device_->sys, dest.raw(), src.raw(), res.raw(), precision, print_modulo,
                             0, &info),
            "InverseOperator: sfa_system_solve");
    last_iterations_ = info.iterations;
    dest.scale(scaling);
    if (use_previous_sol) previous_sol.copydatafrom(dest);
}
