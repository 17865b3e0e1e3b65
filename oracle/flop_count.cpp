/*
 * flop_count.cpp -- op-counting build of the oracle (test infrastructure).
 *
 * fcs_oracle.c is compiled here as C++ with every `double` replaced by a
 * scalar that counts the arithmetic applied to it, so one call of
 * orc_rosenbrock_attempt (one Rosenbrock attempt: Jacobian, LU, stage
 * solves and right-hand sides) yields the attempt's operation counts:
 * add/sub, mul, div, sqrt, and exp / log / log10 / pow (each a
 * transcendental).  Comparisons, fabs, min/max and copies are not counted.
 * With ORC_CONVERSION_RANK1 the exact dC/dY conversion takes its rank-1
 * form (the algorithmic work) instead of the reference's dense product.
 * tests/golden/make_flops.py runs it and freezes tests/golden/flops.json.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#include <stdint.h>
#include <stddef.h>
#include <alloca.h>

namespace counted {

struct Counts {
    unsigned long long add, mul, div, sqrt_, exp_, log_, pow_;
};
static Counts g;

struct cd {
    double v;
    cd() = default;
    constexpr cd(double x) : v(x) {}
    explicit constexpr operator double() const { return v; }
    explicit constexpr operator int() const { return (int)v; }
    cd& operator+=(cd o) { ++g.add; v += o.v; return *this; }
    cd& operator-=(cd o) { ++g.add; v -= o.v; return *this; }
    cd& operator*=(cd o) { ++g.mul; v *= o.v; return *this; }
    cd& operator/=(cd o) { ++g.div; v /= o.v; return *this; }
};
static inline cd operator+(cd a, cd b) { ++g.add; return cd(a.v + b.v); }
static inline cd operator-(cd a, cd b) { ++g.add; return cd(a.v - b.v); }
static inline cd operator*(cd a, cd b) { ++g.mul; return cd(a.v * b.v); }
static inline cd operator/(cd a, cd b) { ++g.div; return cd(a.v / b.v); }
static inline cd operator-(cd a) { return cd(-a.v); }
static inline cd operator+(cd a) { return a; }
#define FC_MIXED(OP)                                                              \
    static inline cd operator OP(cd a, double b) { return a OP cd(b); }           \
    static inline cd operator OP(double a, cd b) { return cd(a) OP b; }           \
    static inline cd operator OP(cd a, int b) { return a OP cd((double)b); }      \
    static inline cd operator OP(int a, cd b) { return cd((double)a) OP b; }
FC_MIXED(+) FC_MIXED(-) FC_MIXED(*) FC_MIXED(/)
#define FC_CMP(OP)                                                                \
    static inline bool operator OP(cd a, cd b) { return a.v OP b.v; }             \
    static inline bool operator OP(cd a, double b) { return a.v OP b; }           \
    static inline bool operator OP(double a, cd b) { return a OP b.v; }           \
    static inline bool operator OP(cd a, int b) { return a.v OP b; }              \
    static inline bool operator OP(int a, cd b) { return a OP b.v; }
FC_CMP(<) FC_CMP(>) FC_CMP(<=) FC_CMP(>=) FC_CMP(==) FC_CMP(!=)
static inline cd fabs(cd a) { return cd(::fabs(a.v)); }
static inline cd sqrt(cd a) { ++g.sqrt_; return cd(::sqrt(a.v)); }
static inline cd exp(cd a) { ++g.exp_; return cd(::exp(a.v)); }
static inline cd log(cd a) { ++g.log_; return cd(::log(a.v)); }
static inline cd log10(cd a) { ++g.log_; return cd(::log10(a.v)); }
static inline cd pow(cd a, cd b) { ++g.pow_; return cd(::pow(a.v, b.v)); }
static inline cd pow(cd a, double b) { return pow(a, cd(b)); }
static inline cd pow(double a, cd b) { return pow(cd(a), b); }
static inline double pow(double a, double b) { ++g.pow_; return ::pow(a, b); }
static inline bool isinf(cd a) { return ::isinf(a.v); }
static inline bool isnan(cd a) { return ::isnan(a.v); }

#define double cd
#define ORC_CONVERSION_RANK1
#include "fcs_oracle.c"
#undef double

}  // namespace counted

// One Rosenbrock attempt of the counted oracle from state y (n values) with
// step dx at pressure p; desc and opts are the C-ABI fcs_mech_desc and
// fcs_solver_opts (same layout here: the counted scalar is one double).
// out receives add, mul, div, sqrt, exp, log, pow counts.  Returns the
// attempt's error norm.
extern "C" double fcc_attempt(const void* desc, const void* opts, double p, double dx, const double* y,
                              unsigned long long* out)
{
    using namespace counted;
    orc_mech* m = orc_mech_create(static_cast<const fcs_mech_desc*>(desc));
    if (!m) return -1.0;
    const int n = orc_n_species(m) + 1;
    cd* yy = static_cast<cd*>(malloc(sizeof(cd) * n));
    cd* yn = static_cast<cd*>(malloc(sizeof(cd) * n));
    for (int i = 0; i < n; ++i) yy[i] = cd(y[i]);
    g = Counts{};
    const cd e = orc_rosenbrock_attempt(m, static_cast<const fcs_solver_opts*>(opts), cd(p), cd(dx), yy, yn);
    const Counts c = g;
    out[0] = c.add; out[1] = c.mul; out[2] = c.div; out[3] = c.sqrt_;
    out[4] = c.exp_; out[5] = c.log_; out[6] = c.pow_;
    free(yy);
    free(yn);
    orc_mech_destroy(m);
    return e.v;
}
