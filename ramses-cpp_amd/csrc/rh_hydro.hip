// rh_hydro.hip -- gfx950 kernels of the per-level Godunov hydro path.
//
// Kernels (one level per launch):
//   k_set_unew      unew <- uold for every cell of the level's octs
//   k_set_uold      uold <- unew for the level's leaf cells
//   k_courant_*     minimum Courant time step over the level's leaf cells
//                   (wavefront shuffles, then one block-level pass)
//   k_godunov       fluxes, ordered reflux accumulation and clamps
//   k_coarse_reflux ordered reflux into coarser leaf cells (AMR faces)
//   k_interpol      prolongation of coarse cells into new octs
//
// The Godunov update is gather-only and deterministic: every leaf cell owns
// its unew entry and rebuilds, in registers, every contribution the
// reference's sequential loop adds to it -- its own six face fluxes and the
// six fluxes its same-level neighbours compute for the faces they share --
// then adds them in the reference's processing order (oct position in the
// level list x 8 + cell, then face).  No atomics, no inter-thread exchange.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "rh_kernels.h"
#include "rh_physics.h"

namespace rh {

namespace {

constexpr int OCTS_PER_BLOCK = 32;  // 32 octs x 8 cells = 256 threads
constexpr int BLOCK = OCTS_PER_BLOCK * 8;

RH_DEV void load_u(const double* __restrict__ u, int ncell, int icell, double out[5]) {
    const size_t i = (size_t)icell - 1;
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) out[iv] = u[(size_t)iv * ncell + i];
}

RH_DEV int cell_id(const MeshDev& m, int igrid, int ic0) { return m.ncoarse + ic0 * m.ngridmax + igrid; }

// Index of cell ic0 among the four cells on an oct face normal to d.
RH_DEV int face_slot(int ic0, int d) { return ((ic0 >> (d + 1)) << d) | (ic0 & ((1 << d) - 1)); }

template <typename T> RH_DEV void copy5(const T a[5], T b[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) b[i] = a[i];
}

__global__ __launch_bounds__(BLOCK) void k_set_unew(MeshDev m, LevelDev lv) {
    const int t = threadIdx.x;
    const int p = blockIdx.x * OCTS_PER_BLOCK + (t % OCTS_PER_BLOCK);
    const int ic0 = t / OCTS_PER_BLOCK;
    if (p >= lv.noct) return;
    const size_t c = (size_t)cell_id(m, lv.grid[p], ic0) - 1;
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + c] = m.uold[(size_t)iv * m.ncell + c];
}

__global__ __launch_bounds__(BLOCK) void k_set_uold(MeshDev m, LevelDev lv) {
    const int t = threadIdx.x;
    const int p = blockIdx.x * OCTS_PER_BLOCK + (t % OCTS_PER_BLOCK);
    const int ic0 = t / OCTS_PER_BLOCK;
    if (p >= lv.noct) return;
    if (!((lv.leaf[p] >> ic0) & 1)) return;
    const size_t c = (size_t)cell_id(m, lv.grid[p], ic0) - 1;
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) m.uold[(size_t)iv * m.ncell + c] = m.unew[(size_t)iv * m.ncell + c];
}

// NaN-skipping minimum: a NaN candidate never replaces the running value,
// as with std::min(dt_max, dt_cand) in the reference loop.
RH_DEV double min_skip_nan(double a, double b) { return (b < a) ? b : a; }

RH_DEV double wave_min(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = min_skip_nan(v, o);
    }
    return v;
}

// HydroSolver.cpp:878-977 per leaf cell, then a block minimum.
__global__ __launch_bounds__(BLOCK) void k_courant_partial(MeshDev m, LevelDev lv, double dx, double gamma,
                                                           double courant_factor, double* part) {
    __shared__ double s_min[BLOCK / 64];
    const int t = threadIdx.x;
    const int p = blockIdx.x * OCTS_PER_BLOCK + (t % OCTS_PER_BLOCK);
    const int ic0 = t / OCTS_PER_BLOCK;
    double cand = __builtin_inf();
    if (p < lv.noct && ((lv.leaf[p] >> ic0) & 1)) {
        const int id = cell_id(m, lv.grid[p], ic0);
        double u[5];
        load_u(m.uold, m.ncell, id, u);
        double d = smax(u[0], 1e-10), v2 = 0.0;
        for (int i = 1; i <= 3; ++i) {
            double v = u[i] / d;
            v2 = v2 + v * v;
        }
        double sum_gamma_p = 0.0;
        double e_nonthermal = 0.0;
        double pr = smax((u[4] - 0.5 * d * v2 - e_nonthermal) * (gamma - 1.0), d * 1e-10);
        pr = smin(pr, d * 1e2);
        double c2 = smax(gamma * pr / smax(d, 1e-10), 1e-20);
        c2 = c2 + sum_gamma_p / d;
        double cs = ssqrt(smax(c2, 1e-10));
        double u_wave = 3 * cs;
        for (int i = 1; i <= 3; ++i) u_wave = u_wave + sabs(u[i] / d);
        double a_grav = 0.0;
        for (int i = 1; i <= 3; ++i) a_grav = a_grav + sabs(m.f ? m.f[(size_t)(i - 1) * m.ncell + id - 1] : 0.0);
        double alpha = a_grav * dx / (u_wave * u_wave);
        alpha = smax(alpha, 0.0001);
        double dt_cand = (dx / u_wave) * (ssqrt(1.0 + 2.0 * courant_factor * alpha) - 1.0) / alpha;
        cand = min_skip_nan(cand, dt_cand);
    }
    cand = wave_min(cand);
    if ((t & 63) == 0) s_min[t >> 6] = cand;
    __syncthreads();
    if (t == 0) {
        double v = s_min[0];
        for (int w = 1; w < BLOCK / 64; ++w) v = min_skip_nan(v, s_min[w]);
        part[blockIdx.x] = v;
    }
}

__global__ __launch_bounds__(256) void k_courant_final(const double* part, int npart, double dt_init, double* dt) {
    __shared__ double s_min[4];
    const int t = threadIdx.x;
    double v = __builtin_inf();
    for (int i = t; i < npart; i += 256) v = min_skip_nan(v, part[i]);
    v = wave_min(v);
    if ((t & 63) == 0) s_min[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double r = dt_init;
        for (int w = 0; w < 4; ++w) r = min_skip_nan(r, s_min[w]);
        *dt = r;
    }
}

// ---------------------------------------------------------------------------
// Godunov update (FP64, bit-faithful).

struct Nbr {
    int cell;  // 1-based cell index (<= 0: boundary)
    int pos;   // oct position in the level list, or -1 for a record
    int ic0;   // cell within that oct
    int rec;   // record index when pos < 0
};

template <int RIEMANN>
__global__ __launch_bounds__(BLOCK) void k_godunov_fp64(MeshDev m, LevelDev lv, GasParams gp, const double* d_dt,
                                                         double dx) {
    const int t = threadIdx.x;
    const int p = blockIdx.x * OCTS_PER_BLOCK + (t % OCTS_PER_BLOCK);
    const int ic0 = t / OCTS_PER_BLOCK;
    if (p >= lv.noct) return;
    if (!((lv.leaf[p] >> ic0) & 1)) return;

    const Gas<double> g{gp.gamma, gp.gm1, gp.smalle, gp.smallp};
    const int slope_type = gp.slope_type, interpol_type = gp.interpol_type;
    const double dt = *d_dt;
    const double dtdx = dt / dx;
    const int noct = lv.noct;
    const int cC = cell_id(m, lv.grid[p], ic0);

    double uC[5], qC[5];
    load_u(m.uold, m.ncell, cC, uC);
    ctoprim(uC, qC, g);

    double fs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double rin[6][5];
    long long kin[6];

    auto same = [&](int pos, int ic) {
        Nbr n;
        n.cell = cell_id(m, lv.grid[pos], ic);
        n.pos = pos;
        n.ic0 = ic;
        n.rec = -1;
        return n;
    };
    auto from_rec = [&](int code) {
        Nbr n;
        n.rec = -1 - code;
        n.cell = lv.rec[n.rec].x;
        n.pos = -1;
        n.ic0 = 0;
        return n;
    };
    // Primitive state of a cell as read by the slope / ghost stencils: a
    // boundary index falls back to `fallback` (the reference copies the
    // centre cell's state there; bound_type is never filled).
    auto prim = [&](int icell, const double fallback[5], double q[5]) {
        if (icell > 0) {
            double u[5];
            load_u(m.uold, m.ncell, icell, u);
            ctoprim(u, q, g);
        } else {
            copy5(fallback, q);
        }
    };

#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
        const int b = 1 << d;
        const int par = (ic0 >> d) & 1;
        const int A = lv.nbo[(size_t)(2 * d) * noct + p];
        const int B = lv.nbo[(size_t)(2 * d + 1) * noct + p];
        const Nbr L1 = par ? same(p, ic0 ^ b) : (A >= 0 ? same(A, ic0 ^ b) : from_rec(A));
        const Nbr R1 = par ? (B >= 0 ? same(B, ic0 ^ b) : from_rec(B)) : same(p, ic0 ^ b);

        double qL1[5], qR1[5];
        prim(L1.cell, qC, qL1);
        prim(R1.cell, qC, qR1);

        // Own face states (trace pass) and this cell's ghost states as its
        // neighbours see them (coarse-ghost branch with n = this cell).
        double OCm[5], OCp[5], GCm[5], GCp[5];
        own_trace(qL1, qC, qR1, d, slope_type, dt, dx, OCm, OCp, g);
        ghost_trace(qL1, qC, qR1, d, interpol_type, dt, dx, GCm, GCp, g);

        // ---- minus face (side 0): ql = ghost qm of L1, qr = own qp
        {
            double F[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            long long K = -1;
            if (L1.pos >= 0) {
                if ((lv.leaf[L1.pos] >> L1.ic0) & 1) {
                    double qL2[5];
                    const int L2cell = A >= 0 ? cell_id(m, lv.grid[A], ic0) : lv.rec[-1 - A].x;
                    prim(L2cell, qL1, qL2);
                    double Gm[5], Gp[5];
                    ghost_trace(qL2, qL1, qC, d, interpol_type, dt, dx, Gm, Gp, g);
                    riemann<RIEMANN>(Gm, OCp, F, g);
                    rotate(F, d);
                    // L1's own plus-face flux into this cell (reflux).
                    double Om[5], Op[5], Fin[5];
                    own_trace(qL2, qL1, qC, d, slope_type, dt, dx, Om, Op, g);
                    riemann<RIEMANN>(Om, GCp, Fin, g);
                    rotate(Fin, d);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) rin[2 * d][iv] = dtdx * -1.0 * Fin[iv] * 0.125;
                    K = ((long long)L1.pos * 8 + L1.ic0) * 8 + (2 * d + 1);
                }
            } else {
                const int4 rc = lv.rec[L1.rec];
                if (rc.w & REC_REFINED_DEV) {
                } else if (rc.w & REC_BOUNDARY_DEV) {
                    double Bm[5], Bp[5];
                    boundary_trace(uC, true, d, dt, dx, Bm, Bp, g);
                    riemann<RIEMANN>(Bm, OCp, F, g);
                    rotate(F, d);
                } else {
                    double qn[5], qnl[5], qnr[5];
                    prim(rc.x, qC, qn);
                    prim(rc.y, qn, qnl);
                    prim(rc.z, qn, qnr);
                    double Gm[5], Gp[5];
                    ghost_trace(qnl, qn, qnr, d, interpol_type, dt, dx, Gm, Gp, g);
                    riemann<RIEMANN>(Gm, OCp, F, g);
                    rotate(F, d);
                    const int slot = lv.rec_slot[L1.rec] + face_slot(ic0, d);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) lv.slot_r[(size_t)slot * 5 + iv] = dtdx * 1.0 * F[iv] * 0.125;
                }
            }
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) fs[iv] = fs[iv] + 1.0 * F[iv];
            kin[2 * d] = K;
        }
        // ---- plus face (side 1): ql = own qm, qr = ghost qp of R1
        {
            double F[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            long long K = -1;
            if (R1.pos >= 0) {
                if ((lv.leaf[R1.pos] >> R1.ic0) & 1) {
                    double qR2[5];
                    const int R2cell = B >= 0 ? cell_id(m, lv.grid[B], ic0) : lv.rec[-1 - B].x;
                    prim(R2cell, qR1, qR2);
                    double Gm[5], Gp[5];
                    ghost_trace(qC, qR1, qR2, d, interpol_type, dt, dx, Gm, Gp, g);
                    riemann<RIEMANN>(OCm, Gp, F, g);
                    rotate(F, d);
                    // R1's own minus-face flux into this cell (reflux).
                    double Om[5], Op[5], Fin[5];
                    own_trace(qC, qR1, qR2, d, slope_type, dt, dx, Om, Op, g);
                    riemann<RIEMANN>(GCm, Op, Fin, g);
                    rotate(Fin, d);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) rin[2 * d + 1][iv] = dtdx * 1.0 * Fin[iv] * 0.125;
                    K = ((long long)R1.pos * 8 + R1.ic0) * 8 + (2 * d);
                }
            } else {
                const int4 rc = lv.rec[R1.rec];
                if (rc.w & REC_REFINED_DEV) {
                } else if (rc.w & REC_BOUNDARY_DEV) {
                    double Bm[5], Bp[5];
                    boundary_trace(uC, true, d, dt, dx, Bm, Bp, g);
                    riemann<RIEMANN>(OCm, Bp, F, g);
                    rotate(F, d);
                } else {
                    double qn[5], qnl[5], qnr[5];
                    prim(rc.x, qC, qn);
                    prim(rc.y, qn, qnl);
                    prim(rc.z, qn, qnr);
                    double Gm[5], Gp[5];
                    ghost_trace(qnl, qn, qnr, d, interpol_type, dt, dx, Gm, Gp, g);
                    riemann<RIEMANN>(OCm, Gp, F, g);
                    rotate(F, d);
                    const int slot = lv.rec_slot[R1.rec] + face_slot(ic0, d);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) lv.slot_r[(size_t)slot * 5 + iv] = dtdx * -1.0 * F[iv] * 0.125;
                }
            }
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) fs[iv] = fs[iv] + -1.0 * F[iv];
            kin[2 * d + 1] = K;
        }
    }

    // ---- ordered accumulation (HydroSolver.cpp:476-547 in processing order)
    // Sort the six incoming contributions by key (invalid ones last), then
    // subtract those ordered before this cell, apply the cell's own update
    // and clamps, and subtract the rest.
    const long long KOWN = ((long long)p * 8 + ic0) * 8 + 6;
    const long long KMAX = 0x7fffffffffffffffLL;
    long long ks[6];
    int ix[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        ks[f] = kin[f] < 0 ? KMAX : kin[f];
        ix[f] = f;
    }
    // Sorting network for 6 elements (12 compare-exchanges).
#define RH_CX(a, b)                                   \
    {                                                 \
        const bool sw = ks[b] < ks[a];                \
        const long long k0 = sw ? ks[b] : ks[a];      \
        const long long k1 = sw ? ks[a] : ks[b];      \
        const int i0 = sw ? ix[b] : ix[a];            \
        const int i1 = sw ? ix[a] : ix[b];            \
        ks[a] = k0;                                   \
        ks[b] = k1;                                   \
        ix[a] = i0;                                   \
        ix[b] = i1;                                   \
    }
    RH_CX(1, 2) RH_CX(4, 5) RH_CX(0, 2) RH_CX(3, 5) RH_CX(0, 1) RH_CX(3, 4)
    RH_CX(1, 4) RH_CX(0, 3) RH_CX(2, 5) RH_CX(1, 3) RH_CX(2, 4) RH_CX(2, 3)
#undef RH_CX
    double rs[6][5];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            double v = rin[0][iv];
#pragma unroll
            for (int f = 1; f < 6; ++f) v = (ix[k] == f) ? rin[f][iv] : v;
            rs[k][iv] = v;
        }
    }
    double u[5];
    load_u(m.unew, m.ncell, cC, u);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const bool take = ks[k] < KOWN;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            double nv = u[iv] - rs[k][iv];
            u[iv] = take ? nv : u[iv];
        }
    }
    // Own update and clamps, HydroSolver.cpp:489-547.
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) u[iv] = u[iv] + dtdx * fs[iv];
    {
        double d_curr = sclamp(u[0], 1e-10, 1e6);
        u[0] = d_curr;
        for (int i = 1; i <= 3; ++i) u[i] = sclamp(u[i], -d_curr * 10.0, d_curr * 10.0);
        double v2_curr = 0.0;
        for (int i = 1; i <= 3; ++i) {
            double v = u[i] / d_curr;
            v2_curr = v2_curr + v * v;
        }
        double ei_curr = u[4] - 0.5 * d_curr * v2_curr;
        if (ei_curr < d_curr * 1e-10 / gp.gm1) {
            double e_non = 0;
            u[4] = d_curr * 1e-10 / gp.gm1 + 0.5 * d_curr * v2_curr + e_non;
        }
        u[4] = smin(u[4], d_curr * 1e8);
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const bool take = ks[k] > KOWN && ks[k] != KMAX;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            double nv = u[iv] - rs[k][iv];
            u[iv] = take ? nv : u[iv];
        }
    }
    const size_t ci = (size_t)cC - 1;
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + ci] = u[iv];
}

// Reflux into coarser leaf cells, one thread per target cell, contributions
// in processing order.
__global__ __launch_bounds__(256) void k_coarse_reflux(MeshDev m, LevelDev lv) {
    const int t = blockIdx.x * 256 + threadIdx.x;
    if (t >= lv.ntgt) return;
    const size_t c = (size_t)lv.tgt_cell[t] - 1;
    double u[5];
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) u[iv] = m.unew[(size_t)iv * m.ncell + c];
    for (int k = lv.tgt_off[t]; k < lv.tgt_off[t + 1]; ++k) {
        const size_t s = (size_t)lv.tgt_slot[k] * 5;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) u[iv] = u[iv] - lv.slot_r[s + iv];
    }
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + c] = u[iv];
}

// HydroSolver.cpp:734-876 for nvar = 5, one thread per new oct.
__global__ __launch_bounds__(64) void k_interpol(int n_octs, const double* __restrict__ u1,
                                                 double* __restrict__ u2, GasParams gp, int interpol_var) {
    const int o = blockIdx.x * 64 + threadIdx.x;
    if (o >= n_octs) return;
    const double* U1 = u1 + (size_t)o * 7 * 64;
    double* U2 = u2 + (size_t)o * 8 * 64;
    const Gas<double> g{gp.gamma, gp.gm1, gp.smalle, gp.smallp};
    const int it = gp.interpol_type;
    double q1[7][5];
    for (int i = 0; i < 7; ++i) {
        const double* ui = U1 + i * 64;
        if (interpol_var == 0) {
            for (int iv = 0; iv < 5; ++iv) q1[i][iv] = ui[iv];
        } else if (interpol_var == 1) {
            double d = smax(ui[0], 1e-10);
            q1[i][0] = ui[0];
            double e_kin = 0.0;
            for (int k = 1; k <= 3; ++k) {
                q1[i][k] = ui[k];
                e_kin = e_kin + 0.5 * ui[k] * ui[k] / d;
            }
            double e_nonthermal = 0.0;
            q1[i][4] = ui[4] - e_kin - e_nonthermal;
        } else {
            double uu[5];
            for (int iv = 0; iv < 5; ++iv) uu[iv] = ui[iv];
            ctoprim(uu, q1[i], g);
        }
    }
    double q2[8][5];
    for (int iv = 0; iv < 5; ++iv) {
        double sl[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < 3; ++d) {
            double dlft = q1[0][iv] - q1[2 * d + 1][iv];
            double drgt = q1[2 * d + 2][iv] - q1[0][iv];
            if (dlft * drgt > 0.0) {
                double sgn = (dlft >= 0.0) ? 1.0 : -1.0;
                int itype = it;
                if (it == 4) itype = (interpol_var == 2 && iv >= 1 && iv <= 3) ? 3 : 2;
                if (itype == 1) sl[d] = sgn * 0.5 * smin(sabs(dlft), sabs(drgt));
                else if (itype == 2) sl[d] = sgn * smin3(sabs(dlft), sabs(drgt), 0.25 * sabs(dlft + drgt));
                else if (itype == 3) sl[d] = 0.25 * (dlft + drgt);
                else sl[d] = 0.0;
            }
        }
        for (int i = 0; i < 8; ++i) {
            const int ix = i & 1, iy = (i & 2) >> 1, iz = (i & 4) >> 2;
            q2[i][iv] = q1[0][iv] + (ix - 0.5) * sl[0] + (iy - 0.5) * sl[1] + (iz - 0.5) * sl[2];
        }
    }
    if (interpol_var == 0) {
        for (int i = 0; i < 8; ++i)
            for (int iv = 0; iv < 5; ++iv) U2[i * 64 + iv] = q2[i][iv];
    } else if (interpol_var == 1) {
        for (int i = 0; i < 8; ++i) {
            double d = smax(q2[i][0], 1e-10);
            U2[i * 64 + 0] = d;
            double e_kin = 0.0;
            for (int k = 1; k <= 3; ++k) {
                U2[i * 64 + k] = q2[i][k];
                e_kin = e_kin + 0.5 * q2[i][k] * q2[i][k] / d;
            }
            double e_nonthermal = 0.0;
            U2[i * 64 + 4] = q2[i][4] + e_kin + e_nonthermal;
        }
    } else {
        const double oneover_n2d = 1.0 / 8;
        for (int k = 1; k <= 3; ++k) {
            double mom_sum = 0.0;
            for (int i = 0; i < 8; ++i) mom_sum = mom_sum + q2[i][0] * q2[i][k] * oneover_n2d;
            double mom_err = mom_sum - U1[k];
            for (int i = 0; i < 8; ++i) {
                double d = smax(q2[i][0], 1e-10);
                q2[i][k] = (d * q2[i][k] - mom_err) / d;
            }
        }
        for (int i = 0; i < 8; ++i) {
            double d = smax(q2[i][0], 1e-10);
            double pp = smax(q2[i][4], 1e-20);
            U2[i * 64 + 0] = d;
            double v2 = 0;
            for (int k = 1; k <= 3; ++k) {
                U2[i * 64 + k] = d * q2[i][k];
                v2 = v2 + q2[i][k] * q2[i][k];
            }
            double e_thermal = pp / g.gm1, e_kinetic = 0.5 * d * v2, e_nonthermal = 0.0;
            U2[i * 64 + 4] = e_thermal + e_kinetic + e_nonthermal;
        }
    }
}

inline int blocks_for(int noct) { return (noct + OCTS_PER_BLOCK - 1) / OCTS_PER_BLOCK; }

}  // namespace

int courant_blocks(const LevelDev& lv) { return blocks_for(lv.noct); }

hipError_t launch_set_unew(const MeshDev& m, const LevelDev& lv, hipStream_t stream) {
    if (lv.noct == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_unew, dim3(blocks_for(lv.noct)), dim3(BLOCK), 0, stream, m, lv);
    return hipGetLastError();
}

hipError_t launch_set_uold(const MeshDev& m, const LevelDev& lv, hipStream_t stream) {
    if (lv.noct == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_uold, dim3(blocks_for(lv.noct)), dim3(BLOCK), 0, stream, m, lv);
    return hipGetLastError();
}

hipError_t launch_courant(const MeshDev& m, const LevelDev& lv, double dx, double gamma, double courant_factor,
                          double dt_init, double* d_part, int npart, double* d_dt, hipStream_t stream) {
    int nb = blocks_for(lv.noct);
    if (nb > npart) return hipErrorInvalidValue;
    if (nb > 0)
        hipLaunchKernelGGL(k_courant_partial, dim3(nb), dim3(BLOCK), 0, stream, m, lv, dx, gamma, courant_factor,
                           d_part);
    hipLaunchKernelGGL(k_courant_final, dim3(1), dim3(256), 0, stream, (const double*)d_part, nb, dt_init, d_dt);
    return hipGetLastError();
}

hipError_t launch_godunov(const MeshDev& m, const LevelDev& lv, const GasParams& gp, int riemann, int precision,
                          const double* d_dt, double dx, hipStream_t stream) {
    if (lv.noct == 0) return hipSuccess;
    if (precision != 0) return hipErrorNotSupported;
    dim3 grid(blocks_for(lv.noct)), block(BLOCK);
    switch (riemann) {
        case 0: hipLaunchKernelGGL(k_godunov_fp64<0>, grid, block, 0, stream, m, lv, gp, d_dt, dx); break;
        case 1: hipLaunchKernelGGL(k_godunov_fp64<1>, grid, block, 0, stream, m, lv, gp, d_dt, dx); break;
        case 2: hipLaunchKernelGGL(k_godunov_fp64<2>, grid, block, 0, stream, m, lv, gp, d_dt, dx); break;
        case 4: hipLaunchKernelGGL(k_godunov_fp64<4>, grid, block, 0, stream, m, lv, gp, d_dt, dx); break;
        default: return hipErrorNotSupported;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || lv.ntgt == 0) return e;
    hipLaunchKernelGGL(k_coarse_reflux, dim3((lv.ntgt + 255) / 256), dim3(256), 0, stream, m, lv);
    return hipGetLastError();
}

hipError_t launch_interpol(int n_octs, const double* d_u1, double* d_u2, const GasParams& gp, int interpol_var,
                           hipStream_t stream) {
    if (n_octs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_interpol, dim3((n_octs + 63) / 64), dim3(64), 0, stream, n_octs, d_u1, d_u2, gp,
                       interpol_var);
    return hipGetLastError();
}

}  // namespace rh
