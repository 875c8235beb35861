// feec/operators/floatvector.hpp -- FloatVector of FEEC++
// (operators/floatvector.hpp:21-338): an owning host vector of Float, NaN
// initialised, with the arithmetic the solve path and its callers use.
// Reductions accumulate in long double as the reference does
// (operators/floatvector.cpp:563-626).
#pragma once

#include <functional>
#include <initializer_list>
#include <string>
#include <vector>

#include "../base/include.hpp"

class LinearOperator;

class FloatVector final {
public:
    explicit FloatVector(int dim, Float initialvalue = notanumber);
    explicit FloatVector(const FloatVector&, Float scaling);
    explicit FloatVector(const std::vector<Float>&, Float scaling = 1.);
    explicit FloatVector(const std::vector<int>&, Float scaling = 1.);
    explicit FloatVector(const std::initializer_list<Float>& l);
    explicit FloatVector(int dimension, const std::function<Float(int)>& generator,
                         Float scaling = 1.);

    FloatVector() = delete;
    FloatVector(const FloatVector&);
    FloatVector(FloatVector&&) noexcept;
    FloatVector& operator=(const FloatVector& vec);
    FloatVector& operator=(FloatVector&& vec) noexcept;
    ~FloatVector() noexcept;

    void check() const;
    std::string text() const;
    std::string data_as_text(bool indexed = true, bool rowwise = false) const;

    FloatVector clone() const { return FloatVector(*this); }

    int getdimension() const { return dimension; }
    Float setentry(int p, Float value);
    Float getentry(int p) const;
    Float& at(int p) & { return pointer[p]; }
    const Float& at(int p) const& { return pointer[p]; }
    Float& operator[](int p) & { return pointer[p]; }
    const Float& operator[](int p) const& { return pointer[p]; }

    void setentries(Float uniform_value);
    void random();  // standard Gaussian entries
    void random_within_range(Float min, Float max);
    void to_absolute();
    void zero();
    void clear() { zero(); }

    FloatVector& normalize();
    FloatVector& scale(Float factor);
    FloatVector& scaleinverse(Float divisor);
    FloatVector& shift(Float delta);
    FloatVector& shiftnegative(Float delta);
    FloatVector& reciprocal();

    FloatVector getslice(int base, int len) const;
    void setslice(int base, int len, Float uniform_value);
    void setslice(int base, const FloatVector& source);
    void addslice(int base, const FloatVector& summand, Float scaling);

    void copydatafrom(const FloatVector& source);
    void copydatafrom(Float scaling, const FloatVector& source);
    void generatedatafrom(const std::function<Float(int)>& generator);
    void adddatafrom(const FloatVector& summand);
    void adddatafrom(Float scalingsummand, const FloatVector& summand);
    void adddatafrom(Float scalingdest, Float scalingsummand, const FloatVector& summand);
    Float scalarproductwith(const FloatVector& right) const;

    Float min() const;
    Float average() const;
    Float sum() const;
    Float norm() const;
    Float norm_sq() const;
    Float norm(const LinearOperator& op) const;
    Float norm_sq(const LinearOperator& op) const;
    Float maxnorm() const;
    Float sumnorm() const;
    Float l2norm() const { return norm(); }

    bool is_finite() const;
    bool is_zero() const;
    bool is_positive() const;
    bool is_negative() const;
    bool is_nonnegative() const;
    bool is_nonpositive() const;
    bool is_numerically_small(Float threshold = desired_closeness) const;

    std::vector<Float> getdata() const { return std::vector<Float>(pointer, pointer + dimension); }
    Float* raw() { return pointer; }
    const Float* raw() const { return pointer; }
    std::size_t memorysize() const { return sizeof(*this) + sizeof(Float) * dimension; }

    static bool is_equal_to(const FloatVector& left, const FloatVector& right);
    friend inline bool operator==(const FloatVector& l, const FloatVector& r) { return is_equal_to(l, r); }
    friend inline bool operator!=(const FloatVector& l, const FloatVector& r) { return !is_equal_to(l, r); }

private:
    int dimension;
    Float* pointer;
};

// operators/floatvector.hpp:357-431
inline FloatVector operator+(const FloatVector& vec) { return vec; }
inline FloatVector operator-(const FloatVector& vec) { return FloatVector(vec, -1.); }
inline FloatVector operator*=(FloatVector& vec, Float s) { vec.scale(s); return vec; }
inline FloatVector operator/=(FloatVector& vec, Float s) { vec.scaleinverse(s); return vec; }
inline FloatVector operator+=(FloatVector& vec, const FloatVector& src) { vec.adddatafrom(src); return vec; }
inline FloatVector operator-=(FloatVector& vec, const FloatVector& src) { vec.adddatafrom(-1., src); return vec; }
inline FloatVector operator+(const FloatVector& l, const FloatVector& r)
{
    FloatVector ret(l);
    ret.adddatafrom(r);
    return ret;
}
inline FloatVector operator-(const FloatVector& l, const FloatVector& r)
{
    FloatVector ret(l);
    ret.adddatafrom(-1., r);
    return ret;
}
inline FloatVector operator*(Float s, const FloatVector& vec) { return FloatVector(vec, s); }
inline FloatVector operator/(const FloatVector& vec, Float s) { return FloatVector(vec, 1. / s); }
inline Float operator*(const FloatVector& l, const FloatVector& r) { return l.scalarproductwith(r); }

inline std::ostream& operator<<(std::ostream& out, const FloatVector& vec)
{
    return out << vec.text();
}

inline FloatVector unitvector(int d, int i)
{
    FloatVector ret(d, 0.);
    ret[i] = 1.;
    return ret;
}
