// opcount.cpp -- TEST INFRASTRUCTURE ONLY.
//
// The oracle (camr_oracle.c) compiled as C++ with `double` replaced by an
// operation-counting real, to count the algorithmic FP64 operations of the
// reference's path: add/sub = 1, mul = 1, div = 1, sqrt = 1 (no FMA: the
// reference CPU build does not contract), comparisons, min/max, fabs,
// copysign and negation free.  The count is attributed to the stage
// running at the time (ORA_STAGE markers in camr_oracle.c).  The type is a
// trivially copyable wrapper of one double, so every array, struct and
// argument keeps the layout and calling convention of the plain build and
// the ctypes bindings load this library unchanged.  Single-threaded: the
// counters are plain globals.
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <type_traits>

enum { ORA_NSTAGE = 9, ORA_NOP = 4 };
static unsigned long long g_ops[ORA_NSTAGE][ORA_NOP];  // add, mul, div, sqrt
static int ora_stage = 0;

struct cnt_real {
    double v;
    cnt_real() = default;
    cnt_real(double x) : v(x) {}
    explicit operator double() const { return v; }
    explicit operator int() const { return (int)v; }
    explicit operator long() const { return (long)v; }
    explicit operator long long() const { return (long long)v; }
    explicit operator bool() const { return v != 0.0; }
};
static_assert(std::is_trivially_copyable<cnt_real>::value && sizeof(cnt_real) == 8,
              "cnt_real must have the layout of double");

static inline void op(int k) { ++g_ops[ora_stage][k]; }
static inline cnt_real operator+(cnt_real a, cnt_real b) { op(0); return a.v + b.v; }
static inline cnt_real operator-(cnt_real a, cnt_real b) { op(0); return a.v - b.v; }
static inline cnt_real operator*(cnt_real a, cnt_real b) { op(1); return a.v * b.v; }
static inline cnt_real operator/(cnt_real a, cnt_real b) { op(2); return a.v / b.v; }
static inline cnt_real operator-(cnt_real a) { return -a.v; }
static inline cnt_real operator+(cnt_real a) { return a; }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline cnt_real operator+(cnt_real a, T b) { return a + cnt_real((double)b); }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline cnt_real operator+(T a, cnt_real b) { return cnt_real((double)a) + b; }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline cnt_real operator-(cnt_real a, T b) { return a - cnt_real((double)b); }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline cnt_real operator-(T a, cnt_real b) { return cnt_real((double)a) - b; }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline cnt_real operator*(cnt_real a, T b) { return a * cnt_real((double)b); }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline cnt_real operator*(T a, cnt_real b) { return cnt_real((double)a) * b; }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline cnt_real operator/(cnt_real a, T b) { return a / cnt_real((double)b); }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline cnt_real operator/(T a, cnt_real b) { return cnt_real((double)a) / b; }
template <class T> static inline cnt_real &operator+=(cnt_real &a, T b) { return a = a + b; }
template <class T> static inline cnt_real &operator-=(cnt_real &a, T b) { return a = a - b; }
template <class T> static inline cnt_real &operator*=(cnt_real &a, T b) { return a = a * b; }
template <class T> static inline cnt_real &operator/=(cnt_real &a, T b) { return a = a / b; }
static inline double rv(cnt_real a) { return a.v; }
template <class T, class = typename std::enable_if<std::is_arithmetic<T>::value>::type>
static inline double rv(T a) { return (double)a; }
#define CMP(OP)                                                                   \
    template <class A, class B,                                                   \
              class = typename std::enable_if<std::is_same<A, cnt_real>::value || \
                                              std::is_same<B, cnt_real>::value>::type> \
    static inline bool operator OP(A a, B b) { return rv(a) OP rv(b); }
CMP(<) CMP(>) CMP(<=) CMP(>=) CMP(==) CMP(!=)
#undef CMP
static inline cnt_real sqrt(cnt_real a) { op(3); return ::sqrt(a.v); }
static inline cnt_real fabs(cnt_real a) { return ::fabs(a.v); }
static inline cnt_real copysign(cnt_real a, cnt_real b) { return ::copysign(a.v, b.v); }
static inline cnt_real fmin(cnt_real a, cnt_real b) { return ::fmin(a.v, b.v); }
static inline cnt_real fmax(cnt_real a, cnt_real b) { return ::fmax(a.v, b.v); }

#define ORA_OPCOUNT 1
#define double cnt_real
#include "camr_oracle.c"
#undef double

extern "C" {
/* ops[s * 4 + k]: stage s (0 fill/clean/primitives, 1 traces, 2 first Riemann solves, 3 transdo + Riemann, 4 nodal divu,
 * 5 transdd + final Riemann + flux adjustment, 6 p div u + update + Saxpy,
 * 7 clean_state after the update + CFL estimate, 8 flattening coefficients),
 * op k (add/sub, mul, div, sqrt). */
void ora_opcount_get(unsigned long long *ops) { memcpy(ops, g_ops, sizeof(g_ops)); }
void ora_opcount_reset(void) { memset(g_ops, 0, sizeof(g_ops)); }
}
