// flopcount.cpp -- exact count of the floating-point operations the
// reference's godunov_fine performs per cell-update, taken from the oracle.
//
// TEST INFRASTRUCTURE ONLY, like the oracle it instruments: only tests/ and
// the documentation of the measurement use it; the product never links it.
//
// The oracle's translation unit (rh_oracle.cpp, which restates the reference
// operation for operation, fp-contract off) is compiled here once more with
// its floating-point type replaced by Flop: a double that tallies every
// addition/subtraction, multiplication, division, square root and comparison
// it takes part in, and otherwise computes exactly what a double computes
// (same operations, same order), so the counted run's results are
// bit-identical to the plain oracle's (checked by the test that uses it).
// Sign operations (fabs, unary minus) and copies are not counted.

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace flopcount {

struct Tally {
    unsigned long long add = 0, mul = 0, div = 0, sqrt = 0, cmp = 0, pow = 0;
};
static Tally g_tally;

struct Flop {
    double v;
    Flop() = default;
    constexpr Flop(double x) : v(x) {}
    // Explicit, so that a conditional expression mixing Flop and double
    // literals converts the literal (the oracle's gravity lookups do that).
    constexpr explicit operator double() const { return v; }
    Flop& operator+=(Flop o) { ++g_tally.add; v = v + o.v; return *this; }
    Flop& operator-=(Flop o) { ++g_tally.add; v = v - o.v; return *this; }
    Flop& operator*=(Flop o) { ++g_tally.mul; v = v * o.v; return *this; }
    Flop& operator/=(Flop o) { ++g_tally.div; v = v / o.v; return *this; }
    Flop operator-() const { return Flop(-v); }
    Flop operator+() const { return *this; }
};

#define FLOP_BINARY(OP, FIELD)                                                               \
    inline Flop operator OP(Flop a, Flop b) { ++g_tally.FIELD; return Flop(a.v OP b.v); }    \
    inline Flop operator OP(Flop a, double b) { ++g_tally.FIELD; return Flop(a.v OP b); }    \
    inline Flop operator OP(double a, Flop b) { ++g_tally.FIELD; return Flop(a OP b.v); }    \
    inline Flop operator OP(Flop a, int b) { ++g_tally.FIELD; return Flop(a.v OP b); }       \
    inline Flop operator OP(int a, Flop b) { ++g_tally.FIELD; return Flop(a OP b.v); }
FLOP_BINARY(+, add)
FLOP_BINARY(-, add)
FLOP_BINARY(*, mul)
FLOP_BINARY(/, div)
#undef FLOP_BINARY

#define FLOP_COMPARE(OP)                                                                   \
    inline bool operator OP(Flop a, Flop b) { ++g_tally.cmp; return a.v OP b.v; }          \
    inline bool operator OP(Flop a, double b) { ++g_tally.cmp; return a.v OP b; }          \
    inline bool operator OP(double a, Flop b) { ++g_tally.cmp; return a OP b.v; }          \
    inline bool operator OP(Flop a, int b) { ++g_tally.cmp; return a.v OP b; }             \
    inline bool operator OP(int a, Flop b) { ++g_tally.cmp; return a OP b.v; }
FLOP_COMPARE(<)
FLOP_COMPARE(>)
FLOP_COMPARE(<=)
FLOP_COMPARE(>=)
FLOP_COMPARE(==)
FLOP_COMPARE(!=)
#undef FLOP_COMPARE

}  // namespace flopcount

// The oracle calls std::sqrt, std::fabs and std::pow on its floating-point
// values.  While it is compiled here, its `std` names this namespace, which
// adds overloads for the counting type and reaches everything else of std
// through the using-directive (the standard headers the oracle includes are
// included above, before the renaming).
namespace flopstd {
using namespace std;
inline flopcount::Flop sqrt(flopcount::Flop a) {
    ++flopcount::g_tally.sqrt;
    return flopcount::Flop(std::sqrt(a.v));
}
inline flopcount::Flop fabs(flopcount::Flop a) { return flopcount::Flop(std::fabs(a.v)); }
inline flopcount::Flop pow(flopcount::Flop a, flopcount::Flop b) {
    ++flopcount::g_tally.pow;
    return flopcount::Flop(std::pow(a.v, b.v));
}
inline flopcount::Flop pow(flopcount::Flop a, int b) {
    ++flopcount::g_tally.pow;
    return flopcount::Flop(std::pow(a.v, b));
}
}  // namespace flopstd

using flopcount::Flop;

#define double Flop
#define std flopstd
#include "rh_oracle.cpp"
#undef std
#undef double

extern "C" {

// set_unew + godunov_fine of level ilevel on a copy of the given mesh (the
// same arrays or_mesh describes, state in doubles), through the counting
// oracle.  unew_out [nvar*ncell] receives the resulting unew; counts[6] the
// tallies (add/sub, mul, div, sqrt, compare, pow) of the godunov_fine call
// alone.  Returns the oracle's status.
int or_flop_count(int ncoarse, int ngridmax, int nx, int ny, int nz, int nlevelmax, const int* father,
                  const int* son, const int* nbor, const int* next, const int* headl, const double* uold,
                  int ilevel, double gamma, int slope_type, int riemann, int interpol_type, double dt, double dx,
                  double* unew_out, unsigned long long counts[6]) {
    const int ncell = ncoarse + 8 * ngridmax;
    std::vector<Flop> u((size_t)5 * ncell), un((size_t)5 * ncell);
    for (size_t i = 0; i < u.size(); ++i) u[i] = Flop(uold[i]);
    or_mesh m{};
    m.ncoarse = ncoarse;
    m.ngridmax = ngridmax;
    m.ncell = ncell;
    m.nvar = 5;
    m.nx = nx;
    m.ny = ny;
    m.nz = nz;
    m.nlevelmax = nlevelmax;
    m.ncpu = 1;
    m.nboundary = 0;
    m.myid = 1;
    m.father = father;
    m.son = son;
    m.nbor = nbor;
    m.next = next;
    m.headl = headl;
    m.uold = u.data();
    m.unew = un.data();
    m.f = nullptr;
    m.bound_type = nullptr;
    m.nbound_type = 0;
    or_params p{};
    p.gamma = gamma;
    p.slope_type = slope_type;
    p.riemann = riemann;
    p.interpol_type = interpol_type;
    p.interpol_var = 0;
    p.boxlen = 1.0;
    or_set_unew(&m, ilevel);
    flopcount::g_tally = flopcount::Tally{};
    const int st = or_godunov_fine(&m, &p, ilevel, dt, dx);
    const flopcount::Tally t = flopcount::g_tally;
    counts[0] = t.add;
    counts[1] = t.mul;
    counts[2] = t.div;
    counts[3] = t.sqrt;
    counts[4] = t.cmp;
    counts[5] = t.pow;
    for (size_t i = 0; i < un.size(); ++i) unew_out[i] = un[i].v;
    return st;
}

}  // extern "C"
