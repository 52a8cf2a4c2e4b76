// Operation-counting scalar for the oracle (TEST INFRASTRUCTURE ONLY).
//
// Built with -DORACLE_COUNTING, `real` is this class: every arithmetic
// operation, comparison and math-library call on it increments a counter, so
// one run of the oracle measures the algorithmic operation count of each
// task per zone (the reference's algorithm, which the oracle restates).
// Values are unchanged: each operation is the double operation.  The counters
// are plain globals; the counting build is single-threaded (no OpenMP).
#pragma once
#include <cmath>
#include <cstdint>
#include <type_traits>

namespace oracle {

enum OpKind {
  OP_ADD = 0,   // + and - (binary)
  OP_MUL,       // *
  OP_DIV,       // /
  OP_SQRT,
  OP_CMP,       // < <= > >= == !=
  OP_MINMAX,    // fmin, fmax
  OP_FABS,
  OP_EXP,
  OP_LOG,
  OP_POW,
  OP_SIN,       // sin and cos
  OP_N
};
extern uint64_t g_ops[OP_N];
extern uint64_t g_task_ops[14][OP_N];  // per ORT_* task (oracle_api.h)

struct Real {
  double v;
  Real() = default;
  constexpr Real(double x) : v(x) {}
  operator double() const { return v; }
  Real &operator+=(Real o) { ++g_ops[OP_ADD]; v += o.v; return *this; }
  Real &operator-=(Real o) { ++g_ops[OP_ADD]; v -= o.v; return *this; }
  Real &operator*=(Real o) { ++g_ops[OP_MUL]; v *= o.v; return *this; }
  Real &operator/=(Real o) { ++g_ops[OP_DIV]; v /= o.v; return *this; }
  Real operator-() const { return Real(-v); }
  Real operator+() const { return *this; }
};

template <class A, class B>
using mixed_t = std::enable_if_t<(std::is_same<A, Real>::value || std::is_same<B, Real>::value) &&
                                     (std::is_same<A, Real>::value || std::is_arithmetic<A>::value) &&
                                     (std::is_same<B, Real>::value || std::is_arithmetic<B>::value),
                                 int>;
inline double val(Real x) { return x.v; }
template <class T, std::enable_if_t<std::is_arithmetic<T>::value, int> = 0>
inline double val(T x) { return (double)x; }

#define ORACLE_BINOP(op, kind)                                     \
  template <class A, class B, mixed_t<A, B> = 0>                   \
  inline Real operator op(A a, B b) {                              \
    ++g_ops[kind];                                                 \
    return Real(val(a) op val(b));                                 \
  }
ORACLE_BINOP(+, OP_ADD)
ORACLE_BINOP(-, OP_ADD)
ORACLE_BINOP(*, OP_MUL)
ORACLE_BINOP(/, OP_DIV)
#undef ORACLE_BINOP
#define ORACLE_CMPOP(op)                                           \
  template <class A, class B, mixed_t<A, B> = 0>                   \
  inline bool operator op(A a, B b) {                              \
    ++g_ops[OP_CMP];                                               \
    return val(a) op val(b);                                       \
  }
ORACLE_CMPOP(<)
ORACLE_CMPOP(<=)
ORACLE_CMPOP(>)
ORACLE_CMPOP(>=)
ORACLE_CMPOP(==)
ORACLE_CMPOP(!=)
#undef ORACLE_CMPOP

}  // namespace oracle

// The oracle calls the math library qualified (std::sqrt ...), so the
// counting overloads live in std.
namespace std {
inline oracle::Real sqrt(oracle::Real x) { ++oracle::g_ops[oracle::OP_SQRT]; return std::sqrt(x.v); }
inline oracle::Real fabs(oracle::Real x) { ++oracle::g_ops[oracle::OP_FABS]; return std::fabs(x.v); }
inline oracle::Real exp(oracle::Real x) { ++oracle::g_ops[oracle::OP_EXP]; return std::exp(x.v); }
inline oracle::Real log(oracle::Real x) { ++oracle::g_ops[oracle::OP_LOG]; return std::log(x.v); }
inline oracle::Real sin(oracle::Real x) { ++oracle::g_ops[oracle::OP_SIN]; return std::sin(x.v); }
inline oracle::Real cos(oracle::Real x) { ++oracle::g_ops[oracle::OP_SIN]; return std::cos(x.v); }
inline bool isfinite(oracle::Real x) { return std::isfinite(x.v); }
template <class A, class B, oracle::mixed_t<A, B> = 0>
inline oracle::Real pow(A a, B b) { ++oracle::g_ops[oracle::OP_POW]; return std::pow(oracle::val(a), oracle::val(b)); }
template <class A, class B, oracle::mixed_t<A, B> = 0>
inline oracle::Real fmax(A a, B b) { ++oracle::g_ops[oracle::OP_MINMAX]; return std::fmax(oracle::val(a), oracle::val(b)); }
template <class A, class B, oracle::mixed_t<A, B> = 0>
inline oracle::Real fmin(A a, B b) { ++oracle::g_ops[oracle::OP_MINMAX]; return std::fmin(oracle::val(a), oracle::val(b)); }
}  // namespace std
