// feec/base/include.hpp -- the part of FEEC++'s base layer the sparse solve
// path uses (base/base.hpp:87-128, base/base.hpp:229-283): the Float type,
// tolerance constants, absolute/maximum, and a LOG stream.
#pragma once

#include <cassert>
#include <cmath>
#include <cstdio>
#include <iostream>
#include <limits>

typedef double Float;

static const constexpr Float notanumber = std::numeric_limits<Float>::quiet_NaN();
static const constexpr Float machine_epsilon = std::numeric_limits<Float>::epsilon();
// sqrt(DBL_EPSILON) = 2^-26
static const constexpr Float desired_precision = 1.4901161193847656e-08;
static const constexpr Float desired_closeness = 1.4901161193847656e-08;

namespace Constants {
static const constexpr Float pi = 3.14159265358979323846264338327950288;
}

template <typename T>
inline constexpr T absolute(const T& x)
{
    if (x >= 0) return x;
    if (x <= 0) return -x;
    return x;
}

template <typename T>
inline constexpr T maximum(const T& a, const T& b)
{
    if (a >= b) return a;
    if (a <= b) return b;
    return a;
}

#define LOG std::cout
static const char nl = '\n';
