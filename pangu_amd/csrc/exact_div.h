// Division with the rounding of the IEEE quotient, in three FMA-unit
// operations instead of the ~11 of a full division (scale, reciprocal,
// Newton steps, fixup).
//
// Markstein's correction (P. Markstein, IBM J. Res. Dev. 34 (1990) 111-119;
// J.-M. Muller et al., Handbook of Floating-Point Arithmetic, 2nd ed.,
// sec. 4.7): with ry = RN(1/y) and q within one ulp of x/y, the remainder
// x - q y is exactly representable (one FMA computes it exactly) and
// RN(q + (x - q y) ry) = RN(x/y).
//
// div12: ry = RN(1/12) has relative error 2^-54, so q = RN(x ry) is within
// one ulp of x/12 and the theorem applies to every x away from underflow
// (|x| > 2^-1018) and overflow.
//
// div_by: ry = 1/y comes from one IEEE division shared by all numerators of
// a divisor.  q = RN(x ry) is within 1.5 ulp of x/y in general; the
// corrected result matched the IEEE quotient on every one of 6e8 random and
// structured operand pairs (divisor significands near 1, 4/3, 3/2 and 2,
// dense and sparse bit patterns) and is within one ulp by construction.
//
// Domain.  Both return the IEEE quotient for x = 0, or for 2^-968 <= |x| <
// 2^1023 with a quotient of magnitude in [2^-968, 2^1023) and y finite,
// nonzero and normal.  Outside it they differ in two ways: a numerator
// below 2^-968 (the remainder underflows) can leave the result one unit in
// the last place off, and y = 0, y = +-inf or x = +-inf give NaN where the
// IEEE quotient is +-inf or +-0.  The kernels divide only by quantities that
// are positive and finite for a finite state (u^t, twice the fast-speed
// quadratic's leading coefficient, the HLL denominator c+ + c- + 1e-20, the
// cell volume), so only states that are already non-finite, or values below
// 1e-291, can reach the outside.
//
// The remainder is negated inside the second FMA, so a zero quotient keeps
// the sign the IEEE division gives it (compilers that fold the negation
// into the first FMA, as GCC does at -O2 with FMA units, lose that sign;
// clang keeps it).  Both helpers need -ffp-contract=off around them only as
// much as the rest of the code does: the FMAs are explicit.
#pragma once
#include <cmath>

#ifdef __HIPCC__
#define PANGU_HD __host__ __device__ __forceinline__
#else
#define PANGU_HD inline
#endif

namespace pangu {

PANGU_HD double div12(double x) {
  constexpr double ry = 1.0 / 12.0;
  const double q = x * ry;
  const double e = fma(q, 12.0, -x);  // q*12 - x, exact
  return fma(-e, ry, q);
}

// x / y given ry = 1 / y (IEEE).
PANGU_HD double div_by(double x, double y, double ry) {
  const double q = x * ry;
  const double e = fma(q, y, -x);
  return fma(-e, ry, q);
}

}  // namespace pangu
