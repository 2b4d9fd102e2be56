// rh_physics.h -- per-cell physics of the Godunov hydro path, as device
// functions.  Every routine is templated on the arithmetic type T: T=double
// is the bit-faithful FP64 form, T=float the compute half of the mixed form.
//
// Bit-faithfulness rules for T=double (must match the reference compiled by
// g++ for x86-64 without FMA):
//   * the whole library is compiled with -ffp-contract=off, no fast-math;
//   * every expression keeps the reference's left-to-right association;
//   * std::max(a,b) is (a<b)?b:a, std::min(a,b) is (b<a)?b:a,
//     std::clamp(v,lo,hi) is (v<lo)?lo:((hi<v)?hi:v), std::min({a,b,c}) keeps
//     the first smallest element -- written out as ternaries so NaN and -0
//     behave as in libstdc++;
//   * '/' and sqrt are the correctly rounded IEEE operations.
// State vectors are 5 components: rho, 3 velocity (or momentum), p (or E).
// Rotated frames put the normal velocity at index 1 by swapping 1<->1+idim.
#pragma once
#include <hip/hip_runtime.h>

#define RH_DEV __device__ __forceinline__

namespace rh {

template <typename T> RH_DEV T smax(T a, T b) { return (a < b) ? b : a; }
template <typename T> RH_DEV T smin(T a, T b) { return (b < a) ? b : a; }
template <typename T> RH_DEV T sclamp(T v, T lo, T hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
template <typename T> RH_DEV T smin3(T a, T b, T c) {
    T m = a;
    m = (b < m) ? b : m;
    m = (c < m) ? c : m;
    return m;
}
template <typename T> RH_DEV T sabs(T a) { return __builtin_fabs(a); }
// Division and square root: correctly rounded for double (the FP64 form);
// hardware reciprocal / square root (about 1 ulp) for float, which only the
// mixed form's FP32 half uses.
RH_DEV double ssqrt(double a) { return __builtin_sqrt(a); }
RH_DEV float ssqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
RH_DEV double dv(double a, double b) { return a / b; }
RH_DEV float dv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

// Cell width of the slope and trace divisions.  When dx is a power of two
// (EXACT_INV) its reciprocal is exact, and a*(1/dx) rounds the same real
// number as a/dx once, so the FP64 form multiplies instead of dividing and
// stays bit-identical.  Otherwise, and in the FP32 half of the mixed form,
// over() is dv().
template <typename T, bool EXACT_INV> struct Len {
    T dx, inv;
};
template <typename T, bool E> RH_DEV T over(T a, const Len<T, E>& l) {
    if constexpr (E && sizeof(T) == 8) return a * l.inv;
    else return dv(a, l.dx);
}
// The coarse cell width 2*dx of the ghost-state interpolation (its
// reciprocal 0.5/dx stays exact).
template <typename T, bool E> RH_DEV Len<T, E> coarse(const Len<T, E>& l) { return {T(2.0) * l.dx, T(0.5) * l.inv}; }

template <typename T> RH_DEV void swap_(T& a, T& b) {
    T t = a;
    a = b;
    b = t;
}

// Rotate a 5-vector into the frame of dimension d (swap 1 <-> 1+d).
// Written with value selects so a runtime d never turns q into an indexed
// (scratch-memory) array.
template <typename T> RH_DEV void rotate(T q[5], int d) {
    const T a = q[1], b = q[2], c = q[3];
    q[1] = (d == 1) ? b : ((d == 2) ? c : a);
    q[2] = (d == 1) ? a : b;
    q[3] = (d == 2) ? a : c;
}

// Per-call constants derived from the namelist.
template <typename T> struct Gas {
    T gamma;     // ratio of specific heats
    T gm1;       // gamma - 1
    T smalle;    // ((1e-10*1e-10)/gamma)/(gamma-1), ctoprim floor
    T smallp;    // (1e-10*1e-10)/gamma, Riemann pressure floor factor
    T inv_gm1;   // 1/(gamma-1), used only by the mixed form's FP32 Riemann solve
};

// Conservative -> primitive (HydroSolver.cpp:613-651, nener=0).
template <typename T> RH_DEV void ctoprim(const T u[5], T q[5], const Gas<T>& g) {
    T d = smax(u[0], T(1e-10));
    q[0] = d;
    T oneoverrho = dv(T(1.0), d);
    T eken = T(0.0);
    q[1] = u[1] * oneoverrho;
    eken = eken + T(0.5) * q[1] * q[1];
    q[2] = u[2] * oneoverrho;
    eken = eken + T(0.5) * q[2] * q[2];
    q[3] = u[3] * oneoverrho;
    eken = eken + T(0.5) * q[3] * q[3];
    T erad = T(0.0);
    T eint = smax(u[4] * oneoverrho - eken - erad, g.smalle);
    q[4] = g.gm1 * d * eint;
}

// Sound speed squared (EquationOfState.cpp:26-51, no barotropic EOS, nener=0).
template <typename T> RH_DEV T cs2(T density, T pressure, const Gas<T>& g) {
    T c = g.gamma * pressure;
    return smax(dv(c, smax(density, T(1e-10))), T(1e-20));
}

// TVD slope limiter (SlopeLimiter.cpp:7-60), theta = 1.5.
template <typename T> RH_DEV T limiter(T ql, T qc, T qr, int slope_type, T nu) {
    if (slope_type == 0) return T(0.0);
    T dlft = qc - ql;
    T drgt = qr - qc;
    if (slope_type == 4 || slope_type == 5) {
        T dlft_s, drgt_s;
        if (slope_type == 4) {
            dlft_s = dv(T(2.0), T(1.0) + nu) * dlft;
            drgt_s = dv(T(2.0), T(1.0) - nu) * drgt;
        } else if (nu >= T(0.0)) {
            dlft_s = dv(T(2.0), nu + T(1e-10)) * dlft;
            drgt_s = dv(T(2.0), T(1.0) - nu) * drgt;
        } else {
            dlft_s = dv(T(2.0), T(1.0) + nu) * dlft;
            drgt_s = dv(T(2.0), -nu + T(1e-10)) * drgt;
        }
        T dcen = T(0.5) * (dlft + drgt);
        T dsgn = (dcen >= T(0.0)) ? T(1.0) : T(-1.0);
        T dlim = smin(sabs(dlft_s), sabs(drgt_s));
        if (dlft * drgt <= T(0.0)) dlim = T(0.0);
        return dsgn * dlim;
    }
    if (slope_type == 6) return T(0.5) * (dlft + drgt);
    if (dlft * drgt <= T(0.0)) return T(0.0);
    T dcen = T(0.5) * (dlft + drgt);
    T dsgn = (dcen >= T(0.0)) ? T(1.0) : T(-1.0);
    T dlim;
    if (slope_type == 1) {
        dlim = smin(sabs(dlft), sabs(drgt));
    } else if (slope_type == 2 || slope_type == 3) {
        dlim = smin3(T(2.0) * sabs(dlft), T(2.0) * sabs(drgt), sabs(dcen));
    } else if (slope_type == 7) {
        dlim = dv(T(2.0) * sabs(dlft) * sabs(drgt), sabs(dlft) + sabs(drgt));
    } else if (slope_type == 8) {
        dlim = smin3(T(1.5) * sabs(dlft), T(1.5) * sabs(drgt), sabs(dcen));
    } else {
        dlim = smin(sabs(dlft), sabs(drgt));
    }
    return dsgn * dlim;
}

// Interpolation limiter of the coarse-ghost state (HydroSolver.cpp:391-404).
template <typename T> RH_DEV T ghost_limiter(T dlft, T drgt, int interpol_type) {
    T s = T(0.0);
    if (interpol_type > 0 && dlft * drgt > T(0.0)) {
        T sgn = (dlft >= T(0.0)) ? T(1.0) : T(-1.0);
        if (interpol_type == 1) s = sgn * T(0.5) * smin(sabs(dlft), sabs(drgt));
        else if (interpol_type == 2) s = sgn * smin3(sabs(dlft), sabs(drgt), T(0.25) * sabs(dlft + drgt));
        else if (interpol_type == 3) s = T(0.25) * (dlft + drgt);
    }
    return s;
}

// MUSCL-Hancock predictor in a rotated frame (HydroSolver.cpp:680-732), with
// the cell's density floor r, normal velocity u and sound speed squared c2
// given (shared by every trace of the same cell).
template <typename T>
RH_DEV void trace_core(const T q[5], const T dq[5], T r, T u, T c2, T hdx, T hdt, T qm[5], T qp[5]) {
    T sr0 = -u * dq[0] - dq[1] * r;
    T su0 = -u * dq[1] - dv(dq[4], r);
    T sp0 = -u * dq[4] - dq[1] * r * c2;
    T src[5] = {sr0, su0, T(0), T(0), sp0};
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) {
        qp[iv] = q[iv] - hdx * dq[iv] + hdt * src[iv];
        qm[iv] = q[iv] + hdx * dq[iv] + hdt * src[iv];
    }
    if (qp[0] < T(1e-10)) qp[0] = r;
    if (qm[0] < T(1e-10)) qm[0] = r;
}

template <typename T>
RH_DEV void trace(const T q[5], const T dq[5], T dt, T dx, T qm[5], T qp[5], const Gas<T>& g) {
    T r = smax(q[0], T(1e-10));
    T p = smax(q[4], r * T(1e-20));
    trace_core(q, dq, r, q[1], cs2(r, p, g), T(0.5) * dx, T(0.5) * dt, qm, qp);
}

// Slopes of the trace pass (compute_slopes, HydroSolver.cpp:653-678) from the
// rotated primitives of the left neighbour a, the cell b, the right
// neighbour c.
template <typename T, bool E>
RH_DEV void own_slopes(const T a[5], const T b[5], const T c[5], int slope_type, T dt, const Len<T, E>& len,
                       T dq[5]) {
    if (slope_type == 0) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) dq[iv] = T(0.0);
    } else {
        T nu = over(b[1] * dt, len);
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            if ((slope_type == 5 || slope_type == 6) && iv != 0) dq[iv] = T(0.0);
            else dq[iv] = over(limiter(a[iv], b[iv], c[iv], slope_type, nu), len);
        }
    }
}

// Interpolation slopes of a coarse-ghost state (HydroSolver.cpp:391-404) over
// the coarse width in lc, from rotated primitives.
template <typename T, bool E>
RH_DEV void ghost_slopes(const T a[5], const T b[5], const T c[5], int interpol_type, const Len<T, E>& lc, T dq[5]) {
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) dq[iv] = over(ghost_limiter(b[iv] - a[iv], c[iv] - b[iv], interpol_type), lc);
}

// Ghost face states only (a coarser neighbour's), rotated primitives.
template <typename T, bool E>
RH_DEV void ghost_trace_rot(const T a[5], const T b[5], const T c[5], int interpol_type, T dt, const Len<T, E>& len,
                            T qm[5], T qp[5], const Gas<T>& g) {
    const Len<T, E> lc = coarse(len);
    T dq[5];
    ghost_slopes(a, b, c, interpol_type, lc, dq);
    trace(b, dq, dt, lc.dx, qm, qp, g);
}

template <typename T> RH_DEV void rotated(const T q[5], T out[5], int d) {
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) out[iv] = q[iv];
    rotate(out, d);
}

// Own face states of a cell along dimension d (HydroSolver.cpp:93-115 with
// compute_slopes :653-678).  ql/qc/qr are the unrotated primitives of the
// left neighbour, the cell and the right neighbour; outputs are rotated.
template <typename T, bool E>
RH_DEV void own_trace(const T ql[5], const T qc[5], const T qr[5], int d, int slope_type, T dt, const Len<T, E>& len,
                      T qm[5], T qp[5], const Gas<T>& g) {
    T a[5], b[5], c[5];
    rotated(ql, a, d);
    rotated(qc, b, d);
    rotated(qr, c, d);
    T dq[5];
    own_slopes(a, b, c, slope_type, dt, len, dq);
    trace(b, dq, dt, len.dx, qm, qp, g);
}

// Ghost face states of a neighbour cell n along dimension d, interpolated
// with interpol_type slopes over 2*dx (HydroSolver.cpp:357-440).  Inputs are
// unrotated primitives of n's left neighbour, n, n's right neighbour.
template <typename T, bool E>
RH_DEV void ghost_trace(const T ql[5], const T qc[5], const T qr[5], int d, int interpol_type, T dt,
                        const Len<T, E>& len, T qm[5], T qp[5], const Gas<T>& g) {
    T a[5], b[5], c[5];
    rotated(ql, a, d);
    rotated(qc, b, d);
    rotated(qr, c, d);
    ghost_trace_rot(a, b, c, interpol_type, dt, len, qm, qp, g);
}

// Ghost of a boundary face: the cell's own state, normal momentum flipped for
// reflective boundaries, traced with zero slopes (HydroSolver.cpp:130-145).
template <typename T>
RH_DEV void boundary_trace(const T u_own[5], bool reflect, int d, T dt, T dx, T qm[5], T qp[5], const Gas<T>& g) {
    T u[5];
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) u[iv] = u_own[iv];
    if (reflect) {
#pragma unroll
        for (int i = 1; i <= 3; ++i)
            if (i == 1 + d) u[i] = u[i] * T(-1.0);
    }
    T q[5];
    ctoprim(u, q, g);
    rotate(q, d);
    T dq[5] = {T(0), T(0), T(0), T(0), T(0)};
    trace(q, dq, dt, dx, qm, qp, g);
}

// RiemannSolver.cpp:385-402
template <typename T> RH_DEV void prim_to_cons(const T q[5], T u[5], const Gas<T>& g) {
    u[0] = q[0];
    T v2 = T(0);
    u[1] = q[0] * q[1];
    v2 = v2 + q[1] * q[1];
    u[2] = q[0] * q[2];
    v2 = v2 + q[2] * q[2];
    u[3] = q[0] * q[3];
    v2 = v2 + q[3] * q[3];
    T e_kin = T(0.5) * q[0] * v2;
    T e_int = dv(q[4], g.gm1);
    T e_nonthermal = T(0.0);
    u[4] = e_kin + e_int + e_nonthermal;
}

// RiemannSolver.cpp:404-427
template <typename T> RH_DEV void phys_flux(const T q[5], T f[5], const Gas<T>& g) {
    T rho = q[0];
    T u = q[1];
    T p = q[4];
    T v2 = T(0);
    v2 = v2 + q[1] * q[1];
    v2 = v2 + q[2] * q[2];
    v2 = v2 + q[3] * q[3];
    T e_kin = T(0.5) * rho * v2;
    T e_int = dv(p, g.gm1);
    T e_nonthermal = T(0.0);
    T p_tot = p;
    T E = e_kin + e_int + e_nonthermal;
    f[0] = rho * u;
    f[1] = rho * u * u + p_tot;
    f[2] = rho * u * q[2];
    f[3] = rho * u * q[3];
    f[4] = (E + p_tot) * u;
}

// RiemannSolver.cpp:15-30
template <typename T> RH_DEV void riemann_llf(const T ql[5], const T qr[5], T flux[5], const Gas<T>& g) {
    T fl[5], fr[5], ul[5], ur[5];
    phys_flux(ql, fl, g);
    phys_flux(qr, fr, g);
    prim_to_cons(ql, ul, g);
    prim_to_cons(qr, ur, g);
    T al = ssqrt(cs2(ql[0], ql[4], g));
    T ar = ssqrt(cs2(qr[0], qr[4], g));
    T a_max = smax(sabs(ql[1]) + al, sabs(qr[1]) + ar);
#pragma unroll
    for (int i = 0; i < 5; ++i) flux[i] = T(0.5) * (fl[i] + fr[i]) - T(0.5) * a_max * (ur[i] - ul[i]);
}

// RiemannSolver.cpp:32-66
template <typename T> RH_DEV void riemann_hll(const T ql[5], const T qr[5], T flux[5], const Gas<T>& g) {
    T fl[5], fr[5], ul[5], ur[5];
    phys_flux(ql, fl, g);
    phys_flux(qr, fr, g);
    prim_to_cons(ql, ul, g);
    prim_to_cons(qr, ur, g);
    T al = ssqrt(cs2(ql[0], ql[4], g));
    T ar = ssqrt(cs2(qr[0], qr[4], g));
    T sl = smin(T(0.0), smin(ql[1] - al, qr[1] - ar));
    T sr = smax(T(0.0), smax(ql[1] + al, qr[1] + ar));
    if (sr - sl < T(1e-20)) {
#pragma unroll
        for (int i = 0; i < 5; ++i) flux[i] = T(0.5) * (fl[i] + fr[i]);
    } else {
        T inv_srs = dv(T(1.0), sr - sl);
#pragma unroll
        for (int i = 0; i < 5; ++i) flux[i] = (sr * fl[i] - sl * fr[i] + sr * sl * (ur[i] - ul[i])) * inv_srs;
    }
}

// RiemannSolver.cpp:68-192 (nener=0)
template <typename T> RH_DEV void riemann_hllc(const T ql[5], const T qr[5], T flux[5], const Gas<T>& g) {
    const T smallr = T(1e-10);
    T rl = smax(ql[0], smallr);
    T ul = ql[1];
    T pl = smax(ql[4], rl * g.smallp);
    T rr = smax(qr[0], smallr);
    T ur = qr[1];
    T pr = smax(qr[4], rr * g.smallp);
    T cl = ssqrt(cs2(rl, ql[4], g));
    T cr = ssqrt(cs2(rr, qr[4], g));
    T cmax = smax(cl, cr);
    T sl = smin(ul, ur) - cmax;
    T sr = smax(ul, ur) + cmax;
    if (sl > T(0.0)) {
        phys_flux(ql, flux, g);
        return;
    }
    if (sr < T(0.0)) {
        phys_flux(qr, flux, g);
        return;
    }
    T rcl = rl * (ul - sl);
    T rcr = rr * (sr - ur);
    T div = rcr + rcl;
    if (div < T(1e-5)) {
        T fl[5], fr[5];
        phys_flux(ql, fl, g);
        phys_flux(qr, fr, g);
#pragma unroll
        for (int i = 0; i < 5; ++i) flux[i] = T(0.5) * (fl[i] + fr[i]);
        return;
    }
    T ustar = dv(rcr * ur + rcl * ul + (pl - pr), div);
    ustar = smax(T(-1e6), smin(T(1e6), ustar));
    T pstar = dv(rcr * pl + rcl * pr + rcl * rcr * (ul - ur), div);
    pstar = smax(pstar, T(0.0));
    // One side of the contact: the left star state for ustar > 0, otherwise
    // the right one; both branches share this code with mirrored inputs.
    const bool left = ustar > T(0.0);
    const T* qs_src = left ? ql : qr;
    T rk = left ? rl : rr;
    T uk = left ? ul : ur;
    T pk = left ? pl : pr;
    T sk = left ? sl : sr;
    T den = left ? smin(sl - ustar, T(-1e-10)) : smax(sr - ustar, T(1e-10));
    T rstar = dv(rk * (sk - uk), den);
    T qs[5];
    qs[0] = rstar;
    qs[1] = ustar;
    qs[2] = qs_src[2];
    qs[3] = qs_src[3];
    T sum_prad_star = T(0.0);
    qs[4] = smax(pstar - sum_prad_star, rstar * g.smallp);
    T fk[5], uk_v[5], us_v[5];
    phys_flux(qs_src, fk, g);
    prim_to_cons(qs_src, uk_v, g);
    prim_to_cons(qs, us_v, g);
    us_v[4] = dv((sk - uk) * uk_v[4] - pk * uk + pstar * ustar, den);
#pragma unroll
    for (int i = 0; i < 5; ++i) flux[i] = fk[i] + sk * (us_v[i] - uk_v[i]);
}

// RiemannSolver.cpp:309-383 (nener=0)
template <typename T> RH_DEV void riemann_acoustic(const T ql[5], const T qr[5], T flux[5], const Gas<T>& g) {
    const T smallr = T(1e-10);
    const T smallc = T(1e-10);
    T rl = smax(ql[0], smallr);
    T ul = ql[1];
    T pl = smax(ql[4], rl * g.smallp);
    T rr = smax(qr[0], smallr);
    T ur = qr[1];
    T pr = smax(qr[4], rr * g.smallp);
    T cl = ssqrt(cs2(rl, ql[4], g));
    T cr = ssqrt(cs2(rr, qr[4], g));
    T wl = cl * rl;
    T wr = cr * rr;
    T pstar = dv((wr * pl + wl * pr) + wl * wr * (ul - ur), wl + wr + T(1e-20));
    T ustar = dv((wr * ur + wl * ul) + (pl - pr), wl + wr + T(1e-20));
    T sgnm = (ustar >= T(0.0)) ? T(1.0) : T(-1.0);
    bool lft = sgnm == T(1.0);
    T ro = lft ? rl : rr, uo = lft ? ul : ur, po = lft ? pl : pr, co = lft ? cl : cr;
    T rstar = ro + dv(pstar - po, smax(co * co, T(1e-20)));
    rstar = smax(rstar, smallr);
    T cstar = ssqrt(smax(dv(g.gamma * pstar, rstar), T(1e-20)));
    cstar = smax(cstar, smallc);
    T spout = co - sgnm * uo;
    T spin = cstar - sgnm * ustar;
    T ushock = T(0.5) * (spin + spout);
    ushock = smax(ushock, -sgnm * ustar);
    if (pstar >= po) {
        spout = ushock;
        spin = spout;
    }
    T rg, ug, pg;
    if (spout < T(0.0)) {
        rg = ro; ug = uo; pg = po;
    } else if (spin >= T(0.0)) {
        rg = rstar; ug = ustar; pg = pstar;
    } else {
        T frac = dv(spout, spout - spin + T(1e-20));
        rg = frac * rstar + (T(1.0) - frac) * ro;
        ug = frac * ustar + (T(1.0) - frac) * uo;
        pg = frac * pstar + (T(1.0) - frac) * po;
    }
    T qg[5];
    qg[0] = rg;
    qg[1] = ug;
    qg[2] = lft ? ql[2] : qr[2];
    qg[3] = lft ? ql[3] : qr[3];
    T sum_prad_star = T(0.0);
    qg[4] = smax(pg - sum_prad_star, rg * g.smallp);
    phys_flux(qg, flux, g);
}

template <int RIEMANN, typename T>
RH_DEV void riemann(const T ql[5], const T qr[5], T flux[5], const Gas<T>& g) {
    if constexpr (RIEMANN == 2) riemann_hllc(ql, qr, flux, g);
    else if constexpr (RIEMANN == 1) riemann_hll(ql, qr, flux, g);
    else if constexpr (RIEMANN == 4) riemann_acoustic(ql, qr, flux, g);
    else riemann_llf(ql, qr, flux, g);
}

// ---------------------------------------------------------------------------
// Mixed form: the whole Riemann solve -- sound speeds, wave speeds, star
// states, physical fluxes and the flux assembly -- is FP32 on FP32 face
// states (hardware reciprocal and square root); the flux is widened to FP64
// for the flux difference, the reflux and the update, which stay FP64.  The
// branch structure follows the FP64 solvers above; the primitive <->
// conservative conversions multiply by 1/(gamma-1).

RH_DEV void prim_to_cons_f(const float q[5], float u[5], const Gas<float>& g) {
    u[0] = q[0];
    u[1] = q[0] * q[1];
    u[2] = q[0] * q[2];
    u[3] = q[0] * q[3];
    u[4] = 0.5f * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + q[4] * g.inv_gm1;
}

RH_DEV void phys_flux_f(const float q[5], float f[5], const Gas<float>& g) {
    const float mu = q[0] * q[1];
    const float E = 0.5f * q[0] * (q[1] * q[1] + q[2] * q[2] + q[3] * q[3]) + q[4] * g.inv_gm1;
    f[0] = mu;
    f[1] = mu * q[1] + q[4];
    f[2] = mu * q[2];
    f[3] = mu * q[3];
    f[4] = (E + q[4]) * q[1];
}

RH_DEV void riemann_hllc_f(const float ql[5], const float qr[5], float flux[5], const Gas<float>& gf) {
    const float rl = smax(ql[0], 1e-10f), ul = ql[1], pl = smax(ql[4], rl * gf.smallp);
    const float rr = smax(qr[0], 1e-10f), ur = qr[1], pr = smax(qr[4], rr * gf.smallp);
    const float cl = ssqrt(cs2(rl, ql[4], gf));
    const float cr = ssqrt(cs2(rr, qr[4], gf));
    const float cmax = smax(cl, cr);
    const float sl = smin(ul, ur) - cmax;
    const float sr = smax(ul, ur) + cmax;
    if (sl > 0.0f) {
        phys_flux_f(ql, flux, gf);
        return;
    }
    if (sr < 0.0f) {
        phys_flux_f(qr, flux, gf);
        return;
    }
    const float rcl = rl * (ul - sl);
    const float rcr = rr * (sr - ur);
    const float div = rcr + rcl;
    if (div < 1e-5f) {
        float fl[5], fr[5];
        phys_flux_f(ql, fl, gf);
        phys_flux_f(qr, fr, gf);
#pragma unroll
        for (int i = 0; i < 5; ++i) flux[i] = 0.5f * (fl[i] + fr[i]);
        return;
    }
    const float idiv = __builtin_amdgcn_rcpf(div);
    float ustar = (rcr * ur + rcl * ul + (pl - pr)) * idiv;
    ustar = smax(-1e6f, smin(1e6f, ustar));
    float pstar = (rcr * pl + rcl * pr + rcl * rcr * (ul - ur)) * idiv;
    pstar = smax(pstar, 0.0f);
    // One side of the contact, as in riemann_hllc.
    const bool left = ustar > 0.0f;
    const float* qk = left ? ql : qr;
    const float rk = left ? rl : rr, uk = left ? ul : ur, pk = left ? pl : pr, sk = left ? sl : sr;
    const float den = left ? smin(sl - ustar, -1e-10f) : smax(sr - ustar, 1e-10f);
    const float iden = __builtin_amdgcn_rcpf(den);
    const float rstar = rk * (sk - uk) * iden;
    float qs[5], fk[5], ukv[5], usv[5];
    qs[0] = rstar;
    qs[1] = ustar;
    qs[2] = qk[2];
    qs[3] = qk[3];
    qs[4] = smax(pstar, rstar * gf.smallp);
    phys_flux_f(qk, fk, gf);
    prim_to_cons_f(qk, ukv, gf);
    prim_to_cons_f(qs, usv, gf);
    usv[4] = ((sk - uk) * ukv[4] - pk * uk + pstar * ustar) * iden;
#pragma unroll
    for (int i = 0; i < 5; ++i) flux[i] = fk[i] + sk * (usv[i] - ukv[i]);
}

RH_DEV void riemann_hll_f(const float ql[5], const float qr[5], float flux[5], const Gas<float>& gf) {
    const float al = ssqrt(cs2(ql[0], ql[4], gf));
    const float ar = ssqrt(cs2(qr[0], qr[4], gf));
    const float sl = smin(0.0f, smin(ql[1] - al, qr[1] - ar));
    const float sr = smax(0.0f, smax(ql[1] + al, qr[1] + ar));
    float fl[5], fr[5];
    phys_flux_f(ql, fl, gf);
    phys_flux_f(qr, fr, gf);
    if (sr - sl < 1e-20f) {
#pragma unroll
        for (int i = 0; i < 5; ++i) flux[i] = 0.5f * (fl[i] + fr[i]);
        return;
    }
    float ul[5], ur[5];
    prim_to_cons_f(ql, ul, gf);
    prim_to_cons_f(qr, ur, gf);
    const float inv = __builtin_amdgcn_rcpf(sr - sl);
#pragma unroll
    for (int i = 0; i < 5; ++i) flux[i] = (sr * fl[i] - sl * fr[i] + sr * sl * (ur[i] - ul[i])) * inv;
}

RH_DEV void riemann_llf_f(const float ql[5], const float qr[5], float flux[5], const Gas<float>& gf) {
    const float al = ssqrt(cs2(ql[0], ql[4], gf));
    const float ar = ssqrt(cs2(qr[0], qr[4], gf));
    const float a_max = smax(sabs(ql[1]) + al, sabs(qr[1]) + ar);
    float fl[5], fr[5], ul[5], ur[5];
    phys_flux_f(ql, fl, gf);
    phys_flux_f(qr, fr, gf);
    prim_to_cons_f(ql, ul, gf);
    prim_to_cons_f(qr, ur, gf);
#pragma unroll
    for (int i = 0; i < 5; ++i) flux[i] = 0.5f * (fl[i] + fr[i]) - 0.5f * a_max * (ur[i] - ul[i]);
}

RH_DEV void riemann_acoustic_f(const float ql[5], const float qr[5], float flux[5], const Gas<float>& gf) {
    const float rl = smax(ql[0], 1e-10f), ul = ql[1], pl = smax(ql[4], rl * gf.smallp);
    const float rr = smax(qr[0], 1e-10f), ur = qr[1], pr = smax(qr[4], rr * gf.smallp);
    const float cl = ssqrt(cs2(rl, ql[4], gf)), cr = ssqrt(cs2(rr, qr[4], gf));
    const float wl = cl * rl, wr = cr * rr;
    const float iw = __builtin_amdgcn_rcpf(wl + wr + 1e-20f);
    const float pstar = ((wr * pl + wl * pr) + wl * wr * (ul - ur)) * iw;
    const float ustar = ((wr * ur + wl * ul) + (pl - pr)) * iw;
    const float sgnm = (ustar >= 0.0f) ? 1.0f : -1.0f;
    const bool lft = sgnm == 1.0f;
    const float ro = lft ? rl : rr, uo = lft ? ul : ur, po = lft ? pl : pr, co = lft ? cl : cr;
    float rstar = smax(ro + dv(pstar - po, smax(co * co, 1e-20f)), 1e-10f);
    float cstar = smax(ssqrt(smax(dv(gf.gamma * pstar, rstar), 1e-20f)), 1e-10f);
    float spout = co - sgnm * uo;
    float spin = cstar - sgnm * ustar;
    float ushock = smax(0.5f * (spin + spout), -sgnm * ustar);
    if (pstar >= po) {
        spout = ushock;
        spin = spout;
    }
    float rg, ug, pg;
    if (spout < 0.0f) {
        rg = ro;
        ug = uo;
        pg = po;
    } else if (spin >= 0.0f) {
        rg = rstar;
        ug = ustar;
        pg = pstar;
    } else {
        const float frac = dv(spout, spout - spin + 1e-20f);
        rg = frac * rstar + (1.0f - frac) * ro;
        ug = frac * ustar + (1.0f - frac) * uo;
        pg = frac * pstar + (1.0f - frac) * po;
    }
    float qg[5];
    qg[0] = rg;
    qg[1] = ug;
    qg[2] = lft ? ql[2] : qr[2];
    qg[3] = lft ? ql[3] : qr[3];
    qg[4] = smax(pg, rg * gf.smallp);
    phys_flux_f(qg, flux, gf);
}

// The mixed form's face flux (FP32).
template <int RIEMANN>
RH_DEV void riemann_f(const float ql[5], const float qr[5], float flux[5], const Gas<float>& gf) {
    if constexpr (RIEMANN == 2) riemann_hllc_f(ql, qr, flux, gf);
    else if constexpr (RIEMANN == 1) riemann_hll_f(ql, qr, flux, gf);
    else if constexpr (RIEMANN == 4) riemann_acoustic_f(ql, qr, flux, gf);
    else riemann_llf_f(ql, qr, flux, gf);
}

}  // namespace rh
