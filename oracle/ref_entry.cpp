// ref_entry.cpp -- C entry points over the reference's own physics sources.
//
// TEST INFRASTRUCTURE ONLY.  Compiled together with the reference's
// RiemannSolver.cpp, SlopeLimiter.cpp, EquationOfState.cpp, Parameters.cpp and
// Constants.cpp, exactly where they lie under the reference tree, into
// oracle/_ref/libref_kernels.so (see oracle/Makefile).  Those five files
// include no logging or MPI header, so they build on their own; the rest of
// the reference (HydroSolver.cpp, AmrGrid.cpp, ...) includes spdlog, which
// this image does not have, and is therefore not built.  The tests use these
// entry points to pin the oracle's Riemann solvers, slope limiter and sound
// speed bit-for-bit against the reference implementation.
#include "ramses/solvers/hydro/RiemannSolver.hpp"
#include "ramses/solvers/hydro/SlopeLimiter.hpp"
#include "ramses/solvers/physics/EquationOfState.hpp"

using ramses::real_t;

extern "C" {

// kind: 0=llf 1=hll 2=hllc 3=exact 4=acoustic
void ref_riemann(int kind, const double* ql, const double* qr, double* flux, double gamma) {
    real_t a[20] = {0}, b[20] = {0}, f[20] = {0};
    for (int i = 0; i < 5; ++i) { a[i] = ql[i]; b[i] = qr[i]; }
    switch (kind) {
        case 1: ramses::RiemannSolver::solve_hll(a, b, f, gamma); break;
        case 2: ramses::RiemannSolver::solve_hllc(a, b, f, gamma); break;
        case 3: ramses::RiemannSolver::solve_godunov_nr(a, b, f, gamma); break;
        case 4: ramses::RiemannSolver::solve_acoustic(a, b, f, gamma); break;
        default: ramses::RiemannSolver::solve_llf(a, b, f, gamma); break;
    }
    for (int i = 0; i < 5; ++i) flux[i] = f[i];
}

double ref_compute_slope(double ql, double qc, double qr, int slope_type, double theta, double nu) {
    return ramses::SlopeLimiter::compute_slope(ql, qc, qr, slope_type, theta, nu);
}

double ref_get_cs2(double d, double p, double gamma) {
    return ramses::RiemannSolver::get_cs2(d, p, gamma);
}

}  // extern "C"
