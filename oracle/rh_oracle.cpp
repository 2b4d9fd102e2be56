// rh_oracle.cpp -- CPU restatement of RAMSES-CPP's per-level Godunov hydro path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
// implementation under ramses-cpp_amd/.  Only tests/, __graft_entry__.smoke()
// and the cpu_baseline leg of bench.py may load the library built from it,
// and only to check or time the reference algorithm.  The product never
// links or calls it.
//
// It restates, operation by operation and in the same association order, the
// reference's single-threaded FP64 path for NDIM=3, nvar=5 (nener=0, no
// passive scalars), with the reference's own (bug-compatible) semantics:
//   * godunov_fine         src/solvers/hydro/HydroSolver.cpp:18-573
//   * set_unew / set_uold  src/solvers/hydro/HydroSolver.cpp:575-611
//   * ctoprim              src/solvers/hydro/HydroSolver.cpp:613-651
//   * compute_slopes       src/solvers/hydro/HydroSolver.cpp:653-678
//   * trace                src/solvers/hydro/HydroSolver.cpp:680-732
//   * interpol_hydro       src/solvers/hydro/HydroSolver.cpp:734-876
//   * compute_courant_step src/solvers/hydro/HydroSolver.cpp:878-977
//   * Riemann solvers      src/solvers/hydro/RiemannSolver.cpp:15-192,309-427
//   * compute_slope        src/solvers/hydro/SlopeLimiter.cpp:7-60
//   * get_cs2              src/solvers/physics/EquationOfState.cpp:26-51
//   * neighbour lookup     src/core/AmrGrid.cpp:389-444,513-550,
//                          tables src/core/Constants.cpp:8-43
//   * uniform mesh build   src/core/Simulation.cpp:226-234 driving
//                          src/core/TreeUpdater.cpp:27-150 (all cells flagged),
//                          free list src/core/AmrGrid.cpp:64-69,259-268,289-302
//
// Parity status of each function is stated in the repository's design notes.
// Build: plain g++ -O2 -ffp-contract=off (x86-64 SSE2, no FMA), see Makefile.

#include "rh_oracle_core.h"

// Planted mistakes, for the tests that show the composite checks can fail
// (tests/test_oracle_composite.py builds the oracle once per value into a
// temporary directory).  0, the default and the only value the Makefile
// builds, is the faithful restatement.
//   1  the 1/8 reflux into the neighbour is dropped
//   2  the reflux has the wrong sign
//   3  refluxes land before the cell's clamp (the clamps run after the
//      whole level instead of at each cell's turn)
//   4  the reflux factor is 1/4 instead of 1/8
//   5  the two face states of every face are swapped (transposed operands)
#ifndef RH_ORACLE_PLANT
#define RH_ORACLE_PLANT 0
#endif

namespace {

// HydroSolver.cpp:613-651 (nener=0, barotropic_eos off)
void ctoprim(const double u[NV], double q[NV], double gamma) {
    double smallr = 1e-10;
    double smallc = 1e-10;
    double smalle = smallc * smallc / gamma / (gamma - 1.0);
    double d = smax(u[0], smallr);
    q[0] = d;
    double oneoverrho = 1.0 / d;
    double eken = 0.0;
    for (int i = 1; i <= 3; ++i) {
        q[i] = u[i] * oneoverrho;
        eken += 0.5 * q[i] * q[i];
    }
    double erad = 0.0;
    double eint = smax(u[4] * oneoverrho - eken - erad, smalle);
    q[4] = (gamma - 1.0) * d * eint;
}

// EquationOfState.cpp:26-51 (q ignored: nener=0)
double get_cs2(double density, double pressure, double gamma) {
    double cs2 = gamma * pressure;
    return smax(cs2 / smax(density, 1e-10), 1e-20);
}

// SlopeLimiter.cpp:7-60
double compute_slope(double ql, double qc, double qr, int slope_type, double theta, double nu) {
    if (slope_type == 0) return 0.0;
    double dlft = qc - ql;
    double drgt = qr - qc;
    if (slope_type == 4 || slope_type == 5) {
        double dlft_s, drgt_s;
        if (slope_type == 4) {
            dlft_s = 2.0 / (1.0 + nu) * dlft;
            drgt_s = 2.0 / (1.0 - nu) * drgt;
        } else {
            if (nu >= 0.0) {
                dlft_s = 2.0 / (nu + 1e-10) * dlft;
                drgt_s = 2.0 / (1.0 - nu) * drgt;
            } else {
                dlft_s = 2.0 / (1.0 + nu) * dlft;
                drgt_s = 2.0 / (-nu + 1e-10) * drgt;
            }
        }
        double dcen = 0.5 * (dlft + drgt);
        double dsgn = (dcen >= 0.0) ? 1.0 : -1.0;
        double dlim = smin(std::fabs(dlft_s), std::fabs(drgt_s));
        if (dlft * drgt <= 0.0) dlim = 0.0;
        return dsgn * dlim;
    }
    if (slope_type == 6) return 0.5 * (dlft + drgt);
    if (dlft * drgt <= 0.0) return 0.0;
    double dcen = 0.5 * (dlft + drgt);
    double dsgn = (dcen >= 0.0) ? 1.0 : -1.0;
    double dlim = 0.0;
    if (slope_type == 1) {
        dlim = smin(std::fabs(dlft), std::fabs(drgt));
    } else if (slope_type == 2 || slope_type == 3) {
        dlim = smin3(2.0 * std::fabs(dlft), 2.0 * std::fabs(drgt), std::fabs(dcen));
    } else if (slope_type == 7) {
        dlim = 2.0 * std::fabs(dlft) * std::fabs(drgt) / (std::fabs(dlft) + std::fabs(drgt));
    } else if (slope_type == 8) {
        dlim = smin3(theta * std::fabs(dlft), theta * std::fabs(drgt), std::fabs(dcen));
    } else {
        dlim = smin(std::fabs(dlft), std::fabs(drgt));
    }
    return dsgn * dlim;
}

// HydroSolver.cpp:680-732
void trace(const double q[NV], const double dq[NV], double dt, double dx, double qm[NV], double qp[NV],
           double gamma) {
    double r = smax(q[0], 1e-10), u = q[1], p = smax(q[4], r * 1e-20);
    double sr0 = -u * dq[0] - dq[1] * r;
    double su0 = -u * dq[1] - dq[4] / r;
    double cs2 = get_cs2(r, p, gamma);
    double sp0 = -u * dq[4] - dq[1] * r * cs2;
    for (int iv = 0; iv < NV; ++iv) {
        double dqi = dq[iv];
        double dq_dt = 0;
        if (iv == 0) dq_dt = sr0;
        else if (iv == 1) dq_dt = su0;
        else if (iv == 4) dq_dt = sp0;
        qp[iv] = q[iv] - 0.5 * dx * dqi + 0.5 * dt * dq_dt;
        qm[iv] = q[iv] + 0.5 * dx * dqi + 0.5 * dt * dq_dt;
        if (iv == 0) {
            if (qp[iv] < 1e-10) qp[iv] = r;
            if (qm[iv] < 1e-10) qm[iv] = r;
        }
    }
}

// RiemannSolver.cpp:385-402
void prim_to_cons(const double q[NV], double u[NV], double gamma) {
    u[0] = q[0];
    double v2 = 0;
    for (int i = 1; i <= 3; ++i) {
        u[i] = q[0] * q[i];
        v2 += q[i] * q[i];
    }
    double e_kin = 0.5 * q[0] * v2;
    double e_int = q[4] / (gamma - 1.0);
    double e_nonthermal = 0.0;
    u[4] = e_kin + e_int + e_nonthermal;
}

// RiemannSolver.cpp:404-427
void compute_flux(const double q[NV], double f[NV], double gamma) {
    double rho = q[0];
    double u = q[1];
    double p = q[4];
    double v2 = 0;
    for (int i = 1; i <= 3; ++i) v2 += q[i] * q[i];
    double e_kin = 0.5 * rho * v2;
    double e_int = p / (gamma - 1.0);
    double e_nonthermal = 0.0;
    double p_tot = p;
    double E = e_kin + e_int + e_nonthermal;
    f[0] = rho * u;
    f[1] = rho * u * u + p_tot;
    for (int i = 2; i <= 3; ++i) f[i] = rho * u * q[i];
    f[4] = (E + p_tot) * u;
}

// RiemannSolver.cpp:15-30
void solve_llf(const double ql[NV], const double qr[NV], double flux[NV], double gamma) {
    double fl[NV], fr[NV], ul[NV], ur[NV];
    compute_flux(ql, fl, gamma);
    compute_flux(qr, fr, gamma);
    prim_to_cons(ql, ul, gamma);
    prim_to_cons(qr, ur, gamma);
    double al = std::sqrt(get_cs2(ql[0], ql[4], gamma));
    double ar = std::sqrt(get_cs2(qr[0], qr[4], gamma));
    double a_max = smax(std::fabs(ql[1]) + al, std::fabs(qr[1]) + ar);
    for (int i = 0; i < NV; ++i) flux[i] = 0.5 * (fl[i] + fr[i]) - 0.5 * a_max * (ur[i] - ul[i]);
}

// RiemannSolver.cpp:32-66
void solve_hll(const double ql[NV], const double qr[NV], double flux[NV], double gamma) {
    double fl[NV], fr[NV], ul[NV], ur[NV];
    compute_flux(ql, fl, gamma);
    compute_flux(qr, fr, gamma);
    prim_to_cons(ql, ul, gamma);
    prim_to_cons(qr, ur, gamma);
    double al = std::sqrt(get_cs2(ql[0], ql[4], gamma));
    double ar = std::sqrt(get_cs2(qr[0], qr[4], gamma));
    double sl = smin(0.0, smin(ql[1] - al, qr[1] - ar));
    double sr = smax(0.0, smax(ql[1] + al, qr[1] + ar));
    if (sr - sl < 1e-20) {
        for (int i = 0; i < NV; ++i) flux[i] = 0.5 * (fl[i] + fr[i]);
    } else {
        double inv_srs = 1.0 / (sr - sl);
        for (int i = 0; i < NV; ++i) flux[i] = (sr * fl[i] - sl * fr[i] + sr * sl * (ur[i] - ul[i])) * inv_srs;
    }
}

// RiemannSolver.cpp:68-192
void solve_hllc(const double ql[NV], const double qr[NV], double flux[NV], double gamma) {
    const double smallc = 1e-10;
    const double smallr = smallc;
    const double smallp = smallc * smallc / gamma;
    double rl = smax(ql[0], smallr);
    double ul = ql[1];
    double pl = smax(ql[4], rl * smallp);
    double rr = smax(qr[0], smallr);
    double ur = qr[1];
    double pr = smax(qr[4], rr * smallp);
    double cl = std::sqrt(get_cs2(rl, ql[4], gamma));
    double cr = std::sqrt(get_cs2(rr, qr[4], gamma));
    double sl = smin(ul, ur) - smax(cl, cr);
    double sr = smax(ul, ur) + smax(cl, cr);
    if (sl > 0.0) {
        compute_flux(ql, flux, gamma);
    } else if (sr < 0.0) {
        compute_flux(qr, flux, gamma);
    } else {
        double rcl = rl * (ul - sl);
        double rcr = rr * (sr - ur);
        double div = rcr + rcl;
        if (div < 1e-5) {
            double fl[NV], fr[NV];
            compute_flux(ql, fl, gamma);
            compute_flux(qr, fr, gamma);
            for (int i = 0; i < NV; ++i) flux[i] = 0.5 * (fl[i] + fr[i]);
            return;
        }
        double ustar = (rcr * ur + rcl * ul + (pl - pr)) / div;
        ustar = smax(-1e6, smin(1e6, ustar));
        double pstar = (rcr * pl + rcl * pr + rcl * rcr * (ul - ur)) / div;
        pstar = smax(pstar, 0.0);
        if (ustar > 0.0) {
            double rstarl = rl * (sl - ul) / (smin(sl - ustar, -1e-10));
            double qs[NV] = {0, 0, 0, 0, 0};
            qs[0] = rstarl;
            qs[1] = ustar;
            for (int i = 2; i <= 3; ++i) qs[i] = ql[i];
            double sum_prad_star = 0.0;
            qs[4] = smax(pstar - sum_prad_star, rstarl * smallp);
            double fl[NV], uv[NV], usv[NV];
            compute_flux(ql, fl, gamma);
            prim_to_cons(ql, uv, gamma);
            prim_to_cons(qs, usv, gamma);
            usv[4] = ((sl - ul) * uv[4] - pl * ul + pstar * ustar) / (smin(sl - ustar, -1e-10));
            for (int i = 0; i < NV; ++i) flux[i] = fl[i] + sl * (usv[i] - uv[i]);
        } else {
            double rstarr = rr * (sr - ur) / (smax(sr - ustar, 1e-10));
            double qs[NV] = {0, 0, 0, 0, 0};
            qs[0] = rstarr;
            qs[1] = ustar;
            for (int i = 2; i <= 3; ++i) qs[i] = qr[i];
            double sum_prad_star = 0.0;
            qs[4] = smax(pstar - sum_prad_star, rstarr * smallp);
            double fr[NV], uv[NV], usv[NV];
            compute_flux(qr, fr, gamma);
            prim_to_cons(qr, uv, gamma);
            prim_to_cons(qs, usv, gamma);
            usv[4] = ((sr - ur) * uv[4] - pr * ur + pstar * ustar) / (smax(sr - ustar, 1e-10));
            for (int i = 0; i < NV; ++i) flux[i] = fr[i] + sr * (usv[i] - uv[i]);
        }
    }
}

// RiemannSolver.cpp:309-383
void solve_acoustic(const double ql[NV], const double qr[NV], double flux[NV], double gamma) {
    const double smallc = 1e-10;
    const double smallr = smallc;
    const double smallp = smallc * smallc / gamma;
    double rl = smax(ql[0], smallr);
    double ul = ql[1];
    double pl = smax(ql[4], rl * smallp);
    double rr = smax(qr[0], smallr);
    double ur = qr[1];
    double pr = smax(qr[4], rr * smallp);
    double cl = std::sqrt(get_cs2(rl, ql[4], gamma));
    double cr = std::sqrt(get_cs2(rr, qr[4], gamma));
    double wl = cl * rl;
    double wr = cr * rr;
    double pstar = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) / (wl + wr + 1e-20);
    double ustar = ((wr * ur + wl * ul) + (pl - pr)) / (wl + wr + 1e-20);
    double sgnm = (ustar >= 0.0) ? 1.0 : -1.0;
    double ro, uo, po, co;
    if (sgnm == 1.0) {
        ro = rl; uo = ul; po = pl; co = cl;
    } else {
        ro = rr; uo = ur; po = pr; co = cr;
    }
    double rstar = ro + (pstar - po) / smax(co * co, 1e-20);
    rstar = smax(rstar, smallr);
    double cstar = std::sqrt(smax(gamma * pstar / rstar, 1e-20));
    cstar = smax(cstar, smallc);
    double spout = co - sgnm * uo;
    double spin = cstar - sgnm * ustar;
    double ushock = 0.5 * (spin + spout);
    ushock = smax(ushock, -sgnm * ustar);
    if (pstar >= po) {
        spout = ushock;
        spin = spout;
    }
    double rg, ug, pg;
    if (spout < 0.0) {
        rg = ro; ug = uo; pg = po;
    } else if (spin >= 0.0) {
        rg = rstar; ug = ustar; pg = pstar;
    } else {
        double frac = spout / (spout - spin + 1e-20);
        rg = frac * rstar + (1.0 - frac) * ro;
        ug = frac * ustar + (1.0 - frac) * uo;
        pg = frac * pstar + (1.0 - frac) * po;
    }
    double qg[NV] = {0, 0, 0, 0, 0};
    qg[0] = rg;
    qg[1] = ug;
    for (int i = 2; i <= 3; ++i) qg[i] = (sgnm == 1.0) ? ql[i] : qr[i];
    double sum_prad_star = 0.0;
    qg[4] = smax(pg - sum_prad_star, rg * smallp);
    compute_flux(qg, flux, gamma);
}

// RiemannSolver.cpp:194-307 (debug print at :283-288 omitted: no numeric effect)
void solve_exact(const double ql[NV], const double qr[NV], double flux[NV], double gamma) {
    const double smallc = 1e-10;
    const double smallr = smallc;
    const double smallp = smallc * smallc / gamma;
    const int niter = 10;
    const double gamma6 = (gamma + 1.0) / (2.0 * gamma);
    double rl = smax(ql[0], smallr);
    double ul = ql[1];
    double pl = smax(ql[4], rl * smallp);
    double rr = smax(qr[0], smallr);
    double ur = qr[1];
    double pr = smax(qr[4], rr * smallp);
    double cl_sq = gamma * pl * rl;
    double cr_sq = gamma * pr * rr;
    double wl = std::sqrt(smax(cl_sq, 0.0));
    double wr = std::sqrt(smax(cr_sq, 0.0));
    double pstar = (wr * pl + wl * pr + wl * wr * (ul - ur)) / (wl + wr + 1e-20);
    pstar = smax(pstar, 0.0);
    double pold = pstar;
    for (int iter = 0; iter < niter; ++iter) {
        double wwl_sq = cl_sq * (1.0 + gamma6 * (pold - pl) / (pl + 1e-20));
        double wwr_sq = cr_sq * (1.0 + gamma6 * (pold - pr) / (pr + 1e-20));
        double wwl = std::sqrt(smax(wwl_sq, 0.0));
        double wwr = std::sqrt(smax(wwr_sq, 0.0));
        double ql_nr = (2.0 * std::pow(wwl, 3)) / (wwl * wwl + cl_sq + 1e-20);
        double qr_nr = (2.0 * std::pow(wwr, 3)) / (wwr * wwr + cr_sq + 1e-20);
        double usl = ul - (pold - pl) / (wwl + 1e-20);
        double usr = ur + (pold - pr) / (wwr + 1e-20);
        double delp = (qr_nr * ql_nr / (qr_nr + ql_nr + 1e-20)) * (usl - usr);
        delp = smax(delp, -pold);
        pold = pold + delp;
        if (std::fabs(delp / (pold + 1e-10)) < 1e-6) break;
    }
    pstar = pold;
    double wwl_f = std::sqrt(smax(cl_sq * (1.0 + gamma6 * (pstar - pl) / (pl + 1e-20)), 0.0));
    double wwr_f = std::sqrt(smax(cr_sq * (1.0 + gamma6 * (pstar - pr) / (pr + 1e-20)), 0.0));
    double ustar = 0.5 * (ul + (pl - pstar) / (wwl_f + 1e-20) + ur - (pr - pstar) / (wwr_f + 1e-20));
    double ro, uo, po;
    double sgnm = (ustar > 0) ? 1.0 : -1.0;
    if (sgnm == 1.0) {
        ro = rl; uo = ul; po = pl;
    } else {
        ro = rr; uo = ur; po = pr;
    }
    double rstar_val;
    if (pstar >= po) {
        rstar_val = ro * ((gamma + 1.0) * pstar + (gamma - 1.0) * po) / ((gamma - 1.0) * pstar + (gamma + 1.0) * po);
    } else {
        rstar_val = ro * std::pow(pstar / (po + 1e-20), 1.0 / gamma);
    }
    rstar_val = smax(rstar_val, smallr);
    double cl_h = std::sqrt(get_cs2(rl, pl, gamma));
    double cr_h = std::sqrt(get_cs2(rr, pr, gamma));
    double sl_h = smin(ul, ur) - smax(cl_h, cr_h);
    double sr_h = smax(ul, ur) + smax(cl_h, cr_h);
    if (sl_h > 0.0) {
        ro = rl; uo = ul; po = pl;
    } else if (sr_h < 0.0) {
        ro = rr; uo = ur; po = pr;
    } else {
        ro = rstar_val; uo = ustar; po = pstar;
    }
    double e_int = po / (gamma - 1.0);
    double e_kin = 0.5 * ro * uo * uo;
    for (int i = 2; i <= 3; ++i) {
        double vt = (ustar > 0.0) ? ql[i] : qr[i];
        e_kin += 0.5 * ro * vt * vt;
    }
    double E = e_int + e_kin;
    flux[0] = ro * uo;
    flux[1] = ro * uo * uo + po;
    for (int i = 2; i <= 3; ++i) flux[i] = flux[0] * ((ustar > 0.0) ? ql[i] : qr[i]);
    flux[4] = (E + po) * uo;
}

void riemann(int kind, const double ql[NV], const double qr[NV], double flux[NV], double gamma) {
    switch (kind) {
        case 2: solve_hllc(ql, qr, flux, gamma); break;
        case 1: solve_hll(ql, qr, flux, gamma); break;
        case 3: solve_exact(ql, qr, flux, gamma); break;
        case 4: solve_acoustic(ql, qr, flux, gamma); break;
        default: solve_llf(ql, qr, flux, gamma); break;
    }
}

// Ghost-state interpolation limiter of the 3D coarse branch
// (HydroSolver.cpp:391-404): interpol_type 1 minmod/2, 2 MonCen/4, 3 central/4.
double ghost_slope(double dlft, double drgt, int interpol_type) {
    double s = 0.0;
    if (interpol_type > 0 && dlft * drgt > 0.0) {
        double sgn = (dlft >= 0.0) ? 1.0 : -1.0;
        if (interpol_type == 1) s = sgn * 0.5 * smin(std::fabs(dlft), std::fabs(drgt));
        else if (interpol_type == 2) s = sgn * smin3(std::fabs(dlft), std::fabs(drgt), 0.25 * std::fabs(dlft + drgt));
        else if (interpol_type == 3) s = 0.25 * (dlft + drgt);
    }
    return s;
}

struct Solver {
    Ctx c;
    double gamma;
    int ilevel;
    double dt, dx;
    std::vector<int> cell_levels;
    std::vector<int> octs;
    std::vector<int> lpos;           // grid -> position in octs (or -1)
    std::vector<double> qr_, ql_;    // per (pos*8+ic-1, idim, iv): right / left face states
    int error = 0;

    void load_u(int icell, double u[NV]) const {
        for (int iv = 1; iv <= NV; ++iv) u[iv - 1] = c.uold(icell, iv);
    }

    // HydroSolver.cpp:653-678
    void compute_slopes(int idc, const int icn[6], int idim, double dq[NV], int slope_type) {
        if (slope_type == 0) {
            for (int iv = 0; iv < NV; ++iv) dq[iv] = 0.0;
            return;
        }
        double ql[NV], qc[NV], qr[NV], uc[NV];
        load_u(idc, uc);
        ctoprim(uc, qc, gamma);
        auto nbq = [&](int id_n, double qn[NV]) {
            if (id_n > 0) {
                double un[NV];
                load_u(id_n, un);
                ctoprim(un, qn, gamma);
            } else {
                for (int iv = 0; iv < NV; ++iv) qn[iv] = qc[iv];
                if (c.bound_is_reflect(-id_n)) qn[1 + idim] *= -1.0;
            }
        };
        nbq(icn[idim * 2], ql);
        nbq(icn[idim * 2 + 1], qr);
        double nu = qc[1 + idim] * dt / dx;
        for (int iv = 0; iv < NV; ++iv) {
            if ((slope_type == 5 || slope_type == 6) && iv != 0)
                dq[iv] = 0.0;
            else
                dq[iv] = compute_slope(ql[iv], qc[iv], qr[iv], slope_type, 1.5, nu) / dx;
        }
    }

    double& QR(int lc, int idim, int iv) { return qr_[((size_t)lc * 3 + idim) * NV + iv]; }
    double& QL(int lc, int idim, int iv) { return ql_[((size_t)lc * 3 + idim) * NV + iv]; }

    // Trace pass, HydroSolver.cpp:93-115
    void trace_pass(size_t noct_limit) {
        size_t n = octs.size() < noct_limit ? octs.size() : noct_limit;
        for (size_t pos = 0; pos < n; ++pos) {
            int ig = octs[pos];
            int ign[7];
            c.nbor_grids(ig, ign);
            for (int ic = 1; ic <= 8; ++ic) {
                int idc0 = c.m->ncoarse + (ic - 1) * c.m->ngridmax + ig - 1;
                int icn[6];
                c.nbor_cells(ign, ic, icn, ig);
                double uc[NV], qc[NV];
                load_u(idc0 + 1, uc);
                ctoprim(uc, qc, gamma);
                int lc = (int)pos * 8 + ic - 1;
                for (int idim = 0; idim < 3; ++idim) {
                    double dq[NV], qrot[NV];
                    for (int iv = 0; iv < NV; ++iv) qrot[iv] = qc[iv];
                    compute_slopes(idc0 + 1, icn, idim, dq, c.p->slope_type);
                    if (idim > 0) {
                        swapd(qrot[1], qrot[1 + idim]);
                        swapd(dq[1], dq[1 + idim]);
                    }
                    double qm[NV], qp[NV];
                    trace(qrot, dq, dt, dx, qm, qp, gamma);
                    if (idim > 0) {
                        swapd(qm[1], qm[1 + idim]);
                        swapd(qp[1], qp[1 + idim]);
                    }
                    for (int iv = 0; iv < NV; ++iv) {
                        QR(lc, idim, iv) = qm[iv];
                        QL(lc, idim, iv) = qp[iv];
                    }
                }
            }
        }
    }

    // Flux pass for one cell, HydroSolver.cpp:118-555 (3D branches).
    void compute_fluxes(int idc0, const int icn[6], int lc) {
        const or_mesh* m = c.m;
        if (m->son[idc0] != 0) return;
        double dtdx = dt / dx;
        double flux_sum[NV] = {0, 0, 0, 0, 0};
        for (int idim = 0; idim < 3; ++idim) {
            for (int side = 0; side < 2; ++side) {
                int id_n = icn[idim * 2 + side];
                double qlf[NV], qrf[NV], flux[NV];
                bool is_refined = (id_n > 0 && (m->son[id_n - 1] > 0 || cell_levels[id_n] > ilevel));
                if (is_refined) {
                    for (int iv = 0; iv < NV; ++iv) flux[iv] = 0.0;
                } else {
                    if (id_n <= 0) {
                        // Boundary branch, HydroSolver.cpp:130-145
                        int ibound = -id_n;
                        int btype = 1;
                        if (ibound > 0 && ibound <= m->nbound_type) btype = m->bound_type[ibound - 1];
                        double un[NV], qn[NV], qmn[NV], qpn[NV], dq0[NV] = {0, 0, 0, 0, 0};
                        load_u(idc0 + 1, un);
                        if (btype == 1) un[1 + idim] *= -1.0;
                        ctoprim(un, qn, gamma);
                        if (idim > 0) swapd(qn[1], qn[1 + idim]);
                        trace(qn, dq0, dt, dx, qmn, qpn, gamma);
                        if (idim > 0) {
                            swapd(qmn[1], qmn[1 + idim]);
                            swapd(qpn[1], qpn[1 + idim]);
                        }
                        if (side == 0) {
                            for (int iv = 0; iv < NV; ++iv) { qlf[iv] = qmn[iv]; qrf[iv] = QL(lc, idim, iv); }
                        } else {
                            for (int iv = 0; iv < NV; ++iv) { qlf[iv] = QR(lc, idim, iv); qrf[iv] = qpn[iv]; }
                        }
                    } else if (!m->son[id_n - 1] && cell_levels[id_n] == ilevel) {
                        // Same-level branch (HydroSolver.cpp:148-161).  It reads
                        // traces of a cell outside this level's list, which the
                        // reference never computed in this call: unreachable on
                        // meshes built by the reference; reported as an error.
                        error = 2;
                        for (int iv = 0; iv < NV; ++iv) flux[iv] = 0.0;
                        goto accumulate;
                    } else {
                        // Coarse-ghost branch, 3D form (HydroSolver.cpp:162-218, 357-453)
                        double uc[NV];
                        load_u(id_n, uc);
                        int icn_p[6] = {0, 0, 0, 0, 0, 0};
                        if (id_n <= m->ncoarse) {
                            c.nbor_cells_coarse(id_n, icn_p);
                        } else {
                            int ig_p = ((id_n - m->ncoarse - 1) % m->ngridmax) + 1;
                            int ic_p = ((id_n - m->ncoarse - 1) / m->ngridmax) + 1;
                            int ign_p[7];
                            c.nbor_grids(ig_p, ign_p);
                            c.nbor_cells(ign_p, ic_p, icn_p, ig_p);
                        }
                        int id_l = icn_p[idim * 2];
                        int id_r = icn_p[idim * 2 + 1];
                        double ul[NV] = {0, 0, 0, 0, 0}, ur[NV] = {0, 0, 0, 0, 0};
                        auto nb_u = [&](int id, double uo[NV]) {
                            if (id > 0) {
                                load_u(id, uo);
                            } else {
                                for (int iv = 0; iv < NV; ++iv) uo[iv] = uc[iv];
                                if (c.bound_is_reflect(-id)) uo[1 + idim] *= -1.0;
                            }
                        };
                        nb_u(id_l, ul);
                        nb_u(id_r, ur);
                        // ctoprim on the unrotated states, then rotate the
                        // primitives (both interpol_var paths give this).
                        double qc[NV], ql[NV], qr[NV];
                        ctoprim(uc, qc, gamma);
                        ctoprim(ul, ql, gamma);
                        ctoprim(ur, qr, gamma);
                        if (idim > 0) {
                            swapd(qc[1], qc[1 + idim]);
                            swapd(ql[1], ql[1 + idim]);
                            swapd(qr[1], qr[1 + idim]);
                        }
                        double dx_coarse = 2.0 * dx;
                        double dqc[NV];
                        for (int iv = 0; iv < NV; ++iv) {
                            double dlft = qc[iv] - ql[iv];
                            double drgt = qr[iv] - qc[iv];
                            dqc[iv] = ghost_slope(dlft, drgt, c.p->interpol_type) / dx_coarse;
                        }
                        double qmn[NV], qpn[NV];
                        trace(qc, dqc, dt, dx_coarse, qmn, qpn, gamma);
                        if (idim > 0) {
                            swapd(qmn[1], qmn[1 + idim]);
                            swapd(qpn[1], qpn[1 + idim]);
                        }
                        if (side == 0) {
                            for (int iv = 0; iv < NV; ++iv) { qlf[iv] = qmn[iv]; qrf[iv] = QL(lc, idim, iv); }
                        } else {
                            for (int iv = 0; iv < NV; ++iv) { qlf[iv] = QR(lc, idim, iv); qrf[iv] = qpn[iv]; }
                        }
                    }
                    if (idim > 0) {
                        swapd(qlf[1], qlf[1 + idim]);
                        swapd(qrf[1], qrf[1 + idim]);
                    }
                    if (RH_ORACLE_PLANT == 5)
                        for (int iv = 0; iv < NV; ++iv) swapd(qlf[iv], qrf[iv]);
                    riemann(c.p->riemann, qlf, qrf, flux, gamma);
                    if (idim > 0) swapd(flux[1], flux[1 + idim]);
                    // Refluxing, HydroSolver.cpp:476-482
                    if (id_n > 0 && cell_levels[id_n] < ilevel && RH_ORACLE_PLANT != 1) {
                        double factor = RH_ORACLE_PLANT == 4 ? 1.0 / (1 << 2) : 1.0 / (1 << 3);
                        double sign = (side == 0) ? 1.0 : -1.0;
                        if (RH_ORACLE_PLANT == 2) sign = -sign;
                        for (int iv = 1; iv <= NV; ++iv) c.unew(id_n, iv) -= dtdx * sign * flux[iv - 1] * factor;
                    }
                }
            accumulate:
                double sign = (side == 0) ? 1.0 : -1.0;
                for (int iv = 0; iv < NV; ++iv) flux_sum[iv] += sign * flux[iv];
            }
        }
        for (int iv = 1; iv <= NV; ++iv) c.unew(idc0 + 1, iv) = c.unew(idc0 + 1, iv) + dtdx * flux_sum[iv - 1];
        if (RH_ORACLE_PLANT != 3) clamp_cell(idc0);
    }

    // Clamps, HydroSolver.cpp:534-547
    void clamp_cell(int idc0) {
        double d_curr = sclamp(c.unew(idc0 + 1, 1), 1e-10, 1e6);
        c.unew(idc0 + 1, 1) = d_curr;
        for (int i = 1; i <= 3; ++i)
            c.unew(idc0 + 1, 1 + i) = sclamp(c.unew(idc0 + 1, 1 + i), -d_curr * 10.0, d_curr * 10.0);
        double v2_curr = 0.0;
        for (int i = 1; i <= 3; ++i) {
            double v = c.unew(idc0 + 1, 1 + i) / d_curr;
            v2_curr += v * v;
        }
        double ei_curr = c.unew(idc0 + 1, 5) - 0.5 * d_curr * v2_curr;
        if (ei_curr < d_curr * 1e-10 / (gamma - 1.0)) {
            double e_non = 0;
            c.unew(idc0 + 1, 5) = d_curr * 1e-10 / (gamma - 1.0) + 0.5 * d_curr * v2_curr + e_non;
        }
        c.unew(idc0 + 1, 5) = smin(c.unew(idc0 + 1, 5), d_curr * 1e8);
    }

    int run(size_t noct_limit) {
        const or_mesh* m = c.m;
        // cell_levels, HydroSolver.cpp:22-38
        cell_levels.assign((size_t)m->ncell + 1, 0);
        for (int ig = 1; ig <= m->ngridmax; ++ig) {
            if (m->father[ig - 1] > 0) {
                int level = 1;
                int cur = ig;
                while (cur > 0) {
                    int fc = m->father[cur - 1];
                    if (fc <= m->ncoarse) break;
                    cur = ((fc - m->ncoarse - 1) % m->ngridmax) + 1;
                    level++;
                }
                for (int ic = 1; ic <= 8; ++ic) cell_levels[(size_t)m->ncoarse + (ic - 1) * m->ngridmax + ig] = level;
            }
        }
        // Oct list, HydroSolver.cpp:43-47
        octs.clear();
        if (ilevel > 0) {
            int ig = c.headl(m->myid, ilevel);
            while (ig > 0) {
                octs.push_back(ig);
                ig = m->next[ig - 1];
            }
        }
        if (octs.empty()) return 0;
        size_t n = octs.size() < noct_limit ? octs.size() : noct_limit;
        qr_.assign(n * 8 * 3 * NV, 0.0);
        ql_.assign(n * 8 * 3 * NV, 0.0);
        trace_pass(noct_limit);
        for (size_t pos = 0; pos < n; ++pos) {
            int ig = octs[pos];
            int ign[7];
            c.nbor_grids(ig, ign);
            for (int ic = 1; ic <= 8; ++ic) {
                int idc0 = m->ncoarse + (ic - 1) * m->ngridmax + ig - 1;
                int icn[6];
                c.nbor_cells(ign, ic, icn, ig);
                compute_fluxes(idc0, icn, (int)pos * 8 + ic - 1);
            }
        }
        if (RH_ORACLE_PLANT == 3)
            for (size_t pos = 0; pos < n; ++pos)
                for (int ic = 1; ic <= 8; ++ic) {
                    int idc0 = m->ncoarse + (ic - 1) * m->ngridmax + octs[pos] - 1;
                    if (m->son[idc0] == 0) clamp_cell(idc0);
                }
        return error;
    }
};

}  // namespace

extern "C" {

// Status: 0 ok, 1 ilevel<=0 (coarse-level branch not restated), 2 same-level
// branch reached (unreachable on reference-built meshes).
int or_godunov_fine(const or_mesh* m, const or_params* p, int ilevel, double dt, double dx) {
    if (ilevel <= 0) return 1;
    Solver s{Ctx{m, p}, p->gamma, ilevel, dt, dx};
    return s.run((size_t)-1);
}

// Same computation restricted to the first max_octs octs of the level list
// (used to time a bounded sample of a large level on the host).
int or_godunov_fine_sample(const or_mesh* m, const or_params* p, int ilevel, double dt, double dx,
                           long long max_octs) {
    if (ilevel <= 0) return 1;
    Solver s{Ctx{m, p}, p->gamma, ilevel, dt, dx};
    return s.run((size_t)max_octs);
}

// HydroSolver.cpp:575-590 (ilevel >= 1)
void or_set_unew(const or_mesh* m, int ilevel) {
    Ctx c{m, nullptr};
    if (ilevel == 0)
        for (int i = 1; i <= m->ncoarse; ++i)
            for (int iv = 1; iv <= NV; ++iv) c.unew(i, iv) = c.uold(i, iv);
    int ig = c.headl(m->myid, ilevel);
    while (ig > 0) {
        for (int ic = 1; ic <= 8; ++ic) {
            int idc = m->ncoarse + (ic - 1) * m->ngridmax + ig;
            for (int iv = 1; iv <= NV; ++iv) c.unew(idc, iv) = c.uold(idc, iv);
        }
        ig = m->next[ig - 1];
    }
}

// HydroSolver.cpp:592-611
void or_set_uold(const or_mesh* m, int ilevel) {
    Ctx c{m, nullptr};
    if (ilevel == 0)
        for (int i = 1; i <= m->ncoarse; ++i)
            if (m->son[i - 1] == 0)
                for (int iv = 1; iv <= NV; ++iv) m->uold[(size_t)(iv - 1) * m->ncell + i - 1] = c.unew(i, iv);
    int ig = c.headl(m->myid, ilevel);
    while (ig > 0) {
        for (int ic = 1; ic <= 8; ++ic) {
            int idc = m->ncoarse + (ic - 1) * m->ngridmax + ig;
            if (m->son[idc - 1] == 0)
                for (int iv = 1; iv <= NV; ++iv) m->uold[(size_t)(iv - 1) * m->ncell + idc - 1] = c.unew(idc, iv);
        }
        ig = m->next[ig - 1];
    }
}

// HydroSolver.cpp:980-998: gravity kick of the level's leaf cells into
// unew (f_vec zero when m->f is null, as AmrGrid::allocate leaves it).
void or_add_gravity_source_terms(const or_mesh* m, int ilevel, double dt) {
    Ctx c{m, nullptr};
    auto f = [&](int idc, int idim) { return m->f ? m->f[(size_t)(idim - 1) * m->ncell + (idc - 1)] : 0.0; };
    int ig = c.headl(m->myid, ilevel);
    while (ig > 0) {
        for (int ic = 1; ic <= 8; ++ic) {
            int idc = m->ncoarse + (ic - 1) * m->ngridmax + ig;
            if (m->son[idc - 1] > 0) continue;
            double d = sclamp(c.uold(idc, 1), 1e-10, 1e6);
            for (int idim = 1; idim <= 3; ++idim) {
                double fv = f(idc, idim);
                c.unew(idc, 1 + idim) += d * fv * dt;
                c.unew(idc, 1 + idim) = sclamp(c.unew(idc, 1 + idim), -d * 10.0, d * 10.0);
                c.unew(idc, 5) += c.uold(idc, 1 + idim) * fv * dt + 0.5 * d * fv * fv * dt * dt;
            }
        }
        ig = m->next[ig - 1];
    }
}

// HydroSolver.cpp:1000-1043 (nener=0, barotropic EOS off): gravity kick of
// the level's leaf cells in uold, keeping their internal energy.
void or_synchro_hydro_fine(const or_mesh* m, int ilevel, double dt, double gamma) {
    Ctx c{m, nullptr};
    auto f = [&](int idc, int idim) { return m->f ? m->f[(size_t)(idim - 1) * m->ncell + (idc - 1)] : 0.0; };
    auto U = [&](int idc, int iv) -> double& { return m->uold[(size_t)(iv - 1) * m->ncell + idc - 1]; };
    int ig = c.headl(m->myid, ilevel);
    while (ig > 0) {
        for (int ic = 1; ic <= 8; ++ic) {
            int idc = m->ncoarse + (ic - 1) * m->ngridmax + ig;
            if (m->son[idc - 1] > 0) continue;
            double d = sclamp(U(idc, 1), 1e-10, 1e6);
            double v2_old = 0;
            for (int idim = 1; idim <= 3; ++idim) {
                double v = U(idc, 1 + idim) / d;
                v2_old += v * v;
            }
            double e_nonthermal = 0.0;
            double e_int = smax(U(idc, 5) - 0.5 * d * v2_old - e_nonthermal, d * 1e-10 / (gamma - 1.0));
            for (int idim = 1; idim <= 3; ++idim) {
                U(idc, 1 + idim) += d * f(idc, idim) * dt;
                U(idc, 1 + idim) = sclamp(U(idc, 1 + idim), -d * 10.0, d * 10.0);
            }
            double v2_new = 0;
            for (int idim = 1; idim <= 3; ++idim) {
                double v = U(idc, 1 + idim) / d;
                v2_new += v * v;
            }
            U(idc, 5) = e_int + 0.5 * d * v2_new + e_nonthermal;
        }
        ig = m->next[ig - 1];
    }
}

// HydroSolver.cpp:878-977 (barotropic_eos off, nener=0)
double or_courant(const or_mesh* m, int ilevel, double dx, double gamma, double courant_factor, double boxlen) {
    Ctx c{m, nullptr};
    double dt_max = boxlen / 1e-10;
    auto cell = [&](int id) {
        double d = smax(c.uold(id, 1), 1e-10), v2 = 0.0;
        for (int i = 1; i <= 3; ++i) {
            double v = c.uold(id, 1 + i) / d;
            v2 += v * v;
        }
        double sum_gamma_p = 0.0;
        double e_nonthermal = 0.0;
        double p = smax((c.uold(id, 5) - 0.5 * d * v2 - e_nonthermal) * (gamma - 1.0), d * 1e-10);
        p = smin(p, d * 1e2);
        double cs2 = get_cs2(d, p, gamma);
        cs2 += sum_gamma_p / d;
        double cs = std::sqrt(smax(cs2, 1e-10));
        double u_wave = 3 * cs;
        for (int idim = 1; idim <= 3; ++idim) u_wave += std::fabs(c.uold(id, 1 + idim) / d);
        double a_grav = 0.0;
        for (int idim = 1; idim <= 3; ++idim)
            a_grav += std::fabs(m->f ? m->f[(size_t)(idim - 1) * m->ncell + (id - 1)] : 0.0);
        double alpha = a_grav * dx / (u_wave * u_wave);
        alpha = smax(alpha, 0.0001);
        double dt_cand = (dx / u_wave) * (std::sqrt(1.0 + 2.0 * courant_factor * alpha) - 1.0) / alpha;
        dt_max = smin(dt_max, dt_cand);
    };
    if (ilevel == 0) {
        for (int id = 1; id <= m->ncoarse; ++id)
            if (!(m->son[id - 1] > 0)) cell(id);
    } else {
        int ig = c.headl(m->myid, ilevel);
        while (ig > 0) {
            for (int ic = 1; ic <= 8; ++ic) {
                int id = m->ncoarse + (ic - 1) * m->ngridmax + ig;
                if (m->son[id - 1] > 0) continue;
                cell(id);
            }
            ig = m->next[ig - 1];
        }
    }
    return dt_max;
}

// HydroSolver.cpp:734-876 (nener=0, barotropic off); u1/u2 rows hold nvar values.
void or_interpol_hydro(const or_params* p, int nvar, const double* u1 /*[7][64]*/, double* u2 /*[8][64]*/) {
    double gam = p->gamma;
    int interpol_var = p->interpol_var;
    int interpol_type = p->interpol_type;
    double q1[7][64] = {{0}};
    double q2[8][64] = {{0}};
    auto U1 = [&](int i, int iv) { return u1[i * 64 + iv]; };
    if (interpol_var == 0) {
        for (int i = 0; i < 7; ++i)
            for (int iv = 0; iv < nvar; ++iv) q1[i][iv] = U1(i, iv);
    } else if (interpol_var == 1) {
        for (int i = 0; i < 7; ++i) {
            double d = smax(U1(i, 0), 1e-10);
            q1[i][0] = U1(i, 0);
            double e_kin = 0.0;
            for (int idim = 1; idim <= 3; ++idim) {
                q1[i][idim] = U1(i, idim);
                e_kin += 0.5 * U1(i, idim) * U1(i, idim) / d;
            }
            double e_nonthermal = 0.0;
            q1[i][4] = U1(i, 4) - e_kin - e_nonthermal;
            for (int iv = 5; iv < nvar; ++iv) q1[i][iv] = U1(i, iv);
        }
    } else {
        for (int i = 0; i < 7; ++i) {
            double uu[NV], qq[NV];
            for (int iv = 0; iv < NV; ++iv) uu[iv] = U1(i, iv);
            ctoprim(uu, qq, gam);
            for (int iv = 0; iv < NV; ++iv) q1[i][iv] = qq[iv];
            // passive scalars (nvar > 5) follow ctoprim's q = u / rho
            double d = smax(U1(i, 0), 1e-10);
            double oor = 1.0 / d;
            for (int iv = 5; iv < nvar; ++iv) q1[i][iv] = U1(i, iv) * oor;
        }
    }
    for (int iv = 0; iv < nvar; ++iv) {
        double slopes[3] = {0, 0, 0};
        for (int idim = 0; idim < 3; ++idim) {
            double dlft = q1[0][iv] - q1[2 * idim + 1][iv];
            double drgt = q1[2 * idim + 2][iv] - q1[0][iv];
            if (dlft * drgt > 0.0) {
                double sgn = (dlft >= 0.0) ? 1.0 : -1.0;
                int itype = interpol_type;
                if (interpol_type == 4) itype = (interpol_var == 2 && iv >= 1 && iv <= 3) ? 3 : 2;
                if (itype == 1) slopes[idim] = sgn * 0.5 * smin(std::fabs(dlft), std::fabs(drgt));
                else if (itype == 2) slopes[idim] = sgn * smin3(std::fabs(dlft), std::fabs(drgt), 0.25 * std::fabs(dlft + drgt));
                else if (itype == 3) slopes[idim] = 0.25 * (dlft + drgt);
                else slopes[idim] = 0.0;
            }
        }
        for (int i = 0; i < 8; ++i) {
            int ix = i & 1, iy = (i & 2) >> 1, iz = (i & 4) >> 2;
            q2[i][iv] = q1[0][iv] + (ix - 0.5) * slopes[0] + (iy - 0.5) * slopes[1] + (iz - 0.5) * slopes[2];
        }
    }
    auto U2 = [&](int i, int iv) -> double& { return u2[i * 64 + iv]; };
    if (interpol_var == 0) {
        for (int i = 0; i < 8; ++i)
            for (int iv = 0; iv < nvar; ++iv) U2(i, iv) = q2[i][iv];
    } else if (interpol_var == 1) {
        for (int i = 0; i < 8; ++i) {
            double d = smax(q2[i][0], 1e-10);
            U2(i, 0) = d;
            double e_kin = 0.0;
            for (int idim = 1; idim <= 3; ++idim) {
                U2(i, idim) = q2[i][idim];
                e_kin += 0.5 * q2[i][idim] * q2[i][idim] / d;
            }
            double e_nonthermal = 0.0;
            U2(i, 4) = q2[i][4] + e_kin + e_nonthermal;
            for (int iv = 5; iv < nvar; ++iv) U2(i, iv) = q2[i][iv];
        }
    } else {
        double oneover_n2d = 1.0 / 8;
        for (int d_idx = 1; d_idx <= 3; ++d_idx) {
            double mom_sum = 0.0;
            for (int i = 0; i < 8; ++i) mom_sum += q2[i][0] * q2[i][d_idx] * oneover_n2d;
            double mom_err = mom_sum - U1(0, d_idx);
            for (int i = 0; i < 8; ++i) {
                double d = smax(q2[i][0], 1e-10);
                q2[i][d_idx] = (d * q2[i][d_idx] - mom_err) / d;
            }
        }
        for (int i = 0; i < 8; ++i) {
            double d = smax(q2[i][0], 1e-10);
            double pp = smax(q2[i][4], 1e-20);
            U2(i, 0) = d;
            double v2 = 0;
            for (int d_idx = 1; d_idx <= 3; ++d_idx) {
                U2(i, d_idx) = d * q2[i][d_idx];
                v2 += q2[i][d_idx] * q2[i][d_idx];
            }
            double e_thermal = pp / (gam - 1.0), e_kinetic = 0.5 * d * v2, e_nonthermal = 0.0;
            U2(i, 4) = e_thermal + e_kinetic + e_nonthermal;
            for (int iv = 5; iv < nvar; ++iv) U2(i, iv) = d * q2[i][iv];
        }
    }
}

// ---- single-function entry points for unit checks -------------------------
void or_ctoprim(const double* u, double* q, double gamma) { ctoprim(u, q, gamma); }
void or_trace(const double* q, const double* dq, double dt, double dx, double* qm, double* qp, double gamma) {
    trace(q, dq, dt, dx, qm, qp, gamma);
}
double or_compute_slope(double ql, double qc, double qr, int slope_type, double theta, double nu) {
    return compute_slope(ql, qc, qr, slope_type, theta, nu);
}
double or_ghost_slope(double dlft, double drgt, int interpol_type) { return ghost_slope(dlft, drgt, interpol_type); }
double or_get_cs2(double d, double p, double gamma) { return get_cs2(d, p, gamma); }
void or_riemann(int kind, const double* ql, const double* qr, double* flux, double gamma) {
    riemann(kind, ql, qr, flux, gamma);
}
void or_prim_to_cons(const double* q, double* u, double gamma) { prim_to_cons(q, u, gamma); }
void or_compute_flux(const double* q, double* f, double gamma) { compute_flux(q, f, gamma); }

// ---- uniform mesh, built the way the reference builds it -------------------
// Simulation.cpp:226-234: for il_ref in 1..levelmin, for il in 1..levelmax:
// flag_fine (every cell below levelmin flagged, TreeUpdater.cpp:410-456) then
// make_grid_fine (TreeUpdater.cpp:27-150), with the free list handing out
// grids 1,2,3,... (AmrGrid.cpp:64-69, 259-268) and each new oct appended to
// level list ilevel+1 (AmrGrid.cpp:289-302).  Only the first pass creates
// octs; later passes find every flagged cell already refined.  Arrays must be
// zero-initialised by the caller except next/prev (free list, set here).
// Returns the number of grids used, or -1 if ngridmax is too small.
int or_build_uniform(int nx, int ny, int nz, int levelmin, int nlevelmax, int ngridmax, int* father, int* son,
                     int* nbor, int* next, int* prev, int* headl, int* taill, int* numbl) {
    int ncoarse = nx * ny * nz;
    int ncell = ncoarse + 8 * ngridmax;
    for (int i = 1; i <= ngridmax; ++i) {
        next[i - 1] = (i < ngridmax) ? i + 1 : 0;
        prev[i - 1] = (i > 1) ? i - 1 : 0;
    }
    int headf = 1, tailf = ngridmax;
    (void)tailf;
    int used = 0;
    auto get_free = [&]() {
        int ig = headf;
        if (ig == 0) return 0;
        headf = next[ig - 1];
        if (headf != 0) prev[headf - 1] = 0;
        next[ig - 1] = 0;
        prev[ig - 1] = 0;
        return ig;
    };
    auto add_to_list = [&](int ig, int il) {
        if (headl[il] == 0) {
            headl[il] = ig;
            taill[il] = ig;
            prev[ig - 1] = 0;
            next[ig - 1] = 0;
        } else {
            int tail = taill[il];
            next[tail - 1] = ig;
            prev[ig - 1] = tail;
            next[ig - 1] = 0;
            taill[il] = ig;
        }
        numbl[il]++;
    };
    or_mesh mv{};
    mv.ncoarse = ncoarse;
    mv.ngridmax = ngridmax;
    mv.ncell = ncell;
    mv.nvar = NV;
    mv.nx = nx; mv.ny = ny; mv.nz = nz;
    mv.nlevelmax = nlevelmax;
    mv.ncpu = 1; mv.nboundary = 0; mv.myid = 1;
    mv.father = father; mv.son = son; mv.nbor = nbor; mv.next = next; mv.headl = headl;
    Ctx c{&mv, nullptr};
    auto nbor_of_coarse = [&](int ic, int idim, int side) {
        int idx = ic - 1;
        int iz = idx / (nx * ny);
        int rem = idx % (nx * ny);
        int iy = rem / nx;
        int ix = rem % nx;
        int x[3] = {ix, iy, iz};
        int n[3] = {nx, ny, nz};
        if (side == 0) x[idim] = (x[idim] - 1 + n[idim]) % n[idim];
        else x[idim] = (x[idim] + 1) % n[idim];
        return x[2] * nx * ny + x[1] * nx + x[0] + 1;
    };
    for (int il = 1; il <= nlevelmax; ++il) {
        if (il >= nlevelmax) break;          // make_grid_fine returns at il >= nlevelmax
        if (!(il < levelmin)) break;         // flag_fine flags nothing new at il >= levelmin
        std::vector<int> cells;
        if (il == 1) {
            for (int ic = 1; ic <= ncoarse; ++ic)
                if (son[ic - 1] == 0) cells.push_back(ic);
        } else {
            int ig = headl[il];
            while (ig > 0) {
                for (int ic = 1; ic <= 8; ++ic) {
                    int idc = ncoarse + (ic - 1) * ngridmax + ig - 1;
                    if (son[idc] == 0) cells.push_back(idc + 1);
                }
                ig = next[ig - 1];
            }
        }
        for (int cc : cells) {
            int new_ig = get_free();
            if (new_ig == 0) return -1;
            used++;
            son[cc - 1] = new_ig;
            father[new_ig - 1] = cc;
            if (il == 1) {
                for (int idim = 0; idim < 3; ++idim)
                    for (int side = 0; side < 2; ++side)
                        nbor[(idim * 2 + side) * ngridmax + new_ig - 1] = nbor_of_coarse(cc, idim, side);
            } else {
                int idc = cc - 1;
                int ig = ((idc - ncoarse) % ngridmax) + 1;
                int ic = ((idc - ncoarse) / ngridmax) + 1;
                int ign[7];
                c.nbor_grids(ig, ign);
                int icn[6];
                c.nbor_cells(ign, ic, icn, ig);
                for (int i = 0; i < 6; ++i) nbor[i * ngridmax + new_ig - 1] = icn[i];
            }
            add_to_list(new_ig, il + 1);
        }
    }
    return used;
}

}  // extern "C"

// ---- refinement of flagged cells ------------------------------------------
// TreeUpdater::make_grid_fine (TreeUpdater.cpp:27-150) for an explicit set of
// cells of level list `ilevel` (1-based cell ids, in the order the reference
// would visit them: list order, then cell order) instead of flag1/flag2.
// New grids come from the free list starting at *headf (AmrGrid.cpp:259-268),
// children get the prolongation of interpol_hydro (with u1 built from the
// parent cell and its six neighbours, TreeUpdater.cpp:103-138) in both uold
// and unew, and each new oct is appended to level list ilevel+1.  Assumes the
// free list has room (no resize).  xg is updated as TreeUpdater.cpp:78-101.
// Returns the number of octs created, -1 when the free list ran out, -2 when
// level list ilevel+1 is beyond nlevelmax.
extern "C" int or_refine(or_mesh* mm, const or_params* p, int ilevel, const int* cells, int ncells, int* prev,
                         int* taill, int* numbl, double* xg, int* headf, int* son_w, int* father_w, int* nbor_w,
                         int* next_w, int* headl_w) {
    or_mesh& m = *mm;
    Ctx c{&m, p};
    if (ilevel < 1 || ilevel + 1 > m.nlevelmax) return -2;  // list ilevel+1 must exist
    int created = 0;
    auto get_free = [&]() {
        int ig = *headf;
        if (ig == 0) return 0;
        *headf = next_w[ig - 1];
        if (*headf != 0) prev[*headf - 1] = 0;
        next_w[ig - 1] = 0;
        prev[ig - 1] = 0;
        return ig;
    };
    auto add_to_list = [&](int ig, int il) {
        if (headl_w[il] == 0) {
            headl_w[il] = ig;
            taill[il] = ig;
            prev[ig - 1] = 0;
            next_w[ig - 1] = 0;
        } else {
            int tail = taill[il];
            next_w[tail - 1] = ig;
            prev[ig - 1] = tail;
            next_w[ig - 1] = 0;
            taill[il] = ig;
        }
        numbl[il]++;
    };
    for (int k = 0; k < ncells; ++k) {
        int ic_coarse = cells[k];
        int idc = ic_coarse - 1;
        if (son_w[idc] != 0) continue;
        int new_ig = get_free();
        if (new_ig == 0) return -1;
        created++;
        son_w[idc] = new_ig;
        father_w[new_ig - 1] = idc + 1;
        int icn_nb[6] = {0, 0, 0, 0, 0, 0};
        if (ilevel == 1) {
            c.nbor_cells_coarse(ic_coarse, icn_nb);
            for (int f = 0; f < 6; ++f) nbor_w[f * m.ngridmax + new_ig - 1] = icn_nb[f];
            int idx = ic_coarse - 1;
            int x[3] = {idx % m.nx, (idx % (m.nx * m.ny)) / m.nx, idx / (m.nx * m.ny)};
            for (int d = 0; d < 3; ++d) xg[d * m.ngridmax + new_ig - 1] = (double)x[d] + 0.5;
        } else {
            int ig = ((idc - m.ncoarse) % m.ngridmax) + 1;
            int ic = ((idc - m.ncoarse) / m.ngridmax) + 1;
            int ign[7];
            c.nbor_grids(ig, ign);
            c.nbor_cells(ign, ic, icn_nb, ig);
            for (int f = 0; f < 6; ++f) nbor_w[f * m.ngridmax + new_ig - 1] = icn_nb[f];
            double dx_level = 1.0 / (double)(1 << (ilevel - 1));
            int ixyz[3] = {(ic - 1) & 1, ((ic - 1) & 2) >> 1, ((ic - 1) & 4) >> 2};
            for (int d = 0; d < 3; ++d)
                xg[d * m.ngridmax + new_ig - 1] = xg[d * m.ngridmax + ig - 1] + (double)(ixyz[d] - 0.5) * dx_level;
        }
        double u1[7][64] = {{0}}, u2[8][64] = {{0}};
        for (int iv = 1; iv <= NV; ++iv) u1[0][iv - 1] = c.uold(ic_coarse, iv);
        for (int i = 1; i <= 6; ++i) {
            int id_n = icn_nb[i - 1];
            if (id_n > 0) {
                for (int iv = 1; iv <= NV; ++iv) u1[i][iv - 1] = c.uold(id_n, iv);
            } else {
                for (int iv = 1; iv <= NV; ++iv) u1[i][iv - 1] = u1[0][iv - 1];
            }
        }
        or_interpol_hydro(p, NV, &u1[0][0], &u2[0][0]);
        for (int isc = 1; isc <= 8; ++isc) {
            int id_child = m.ncoarse + (isc - 1) * m.ngridmax + new_ig - 1;
            for (int iv = 1; iv <= NV; ++iv) {
                m.uold[(size_t)(iv - 1) * m.ncell + id_child] = u2[isc - 1][iv - 1];
                m.unew[(size_t)(iv - 1) * m.ncell + id_child] = u2[isc - 1][iv - 1];
            }
        }
        add_to_list(new_ig, ilevel + 1);
    }
    return created;
}
