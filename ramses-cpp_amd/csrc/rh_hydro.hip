// rh_hydro.hip -- gfx950 kernels of the per-level Godunov hydro path.
//
// Kernels (one level per launch unless stated):
//   k_set_unew / k_set_uold   unew <- uold for every cell of the level's octs /
//                             uold <- unew for its leaf cells
//   k_courant_partial/_final  minimum Courant time step over the level's leaf
//                             cells (wavefront shuffles, block minima, then one
//                             1024-thread pass; the first can also do set_unew)
//   k_step_join               set_uold, the next Courant candidates and
//                             set_unew in one pass between two steps
//   k_courant_boxes           the Courant minimum per box of an ensemble (64-bit atomicMin
//                             on an order-preserving key of the double;
//                             k_fill_keys / k_keys_to_dt around it)
//   k_godunov                 per-cell form: fluxes, ordered reflux
//                             accumulation and clamps
//   k_godunov_oct             oct-in-lanes form of the same update (siblings by
//                             cross-lane shuffle, ordered accumulation in LDS)
//   k_coarse_reflux(_wave)    ordered reflux into coarser leaf cells
//   k_interpol / k_new_octs   prolongation of coarse cells into new octs
//   k_restrict, k_set_uold_restrict
//                             adaptive-run restriction (the subcycling
//                             time-step bookkeeping runs in k_courant_final)
//   k_smooth_bits, k_grad_bits
//                             flag_fine's smoothing passes and gradient
//                             criteria over a level list
//   k_gravity_source / k_synchro, k_pack_hydro
//                             gravity kicks; snapshot packing
//
// The Godunov update is deterministic and uses no floating-point atomics:
// every leaf cell owns its unew entry and gathers every contribution the
// reference's sequential loop adds to it -- its own six face fluxes and the
// six fluxes its same-level neighbours compute for the faces they share --
// then adds them in the reference's processing order (oct position in the
// level list x 8 + cell, then face).  Contributions to coarser cells are
// collected per target and subtracted in the same order by k_coarse_reflux.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <type_traits>
#include <utility>

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

// Streaming kernels (copies, the Courant reduction, the fused step boundary)
// run one thread per oct, looping over its 8 cells: a wave then reads 64
// consecutive octs' values of one cell index and variable, 512 contiguous
// bytes per load on a uniform level (4.6 TB/s against 3.6 for one thread
// per cell).
constexpr int SBLOCK = 256;
inline int stream_blocks(int noct) { return (noct + SBLOCK - 1) / SBLOCK; }

// Copies of an oct's cells (mask bit ic0: cell ic0 is copied) from src to
// dst, which never overlap: all 40 loads are issued before the first store,
// so a short level costs one memory round trip per thread, not eight.
RH_DEV void copy_oct(const MeshDev& m, int g, int mask, const double* __restrict__ src, double* __restrict__ dst) {
    double v[8][5];
#pragma unroll
    for (int ic0 = 0; ic0 < 8; ++ic0) {
        const size_t c = (size_t)cell_id(m, g, ic0) - 1;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) v[ic0][iv] = ((mask >> ic0) & 1) ? src[(size_t)iv * m.ncell + c] : 0.0;
    }
#pragma unroll
    for (int ic0 = 0; ic0 < 8; ++ic0) {
        if (!((mask >> ic0) & 1)) continue;
        const size_t c = (size_t)cell_id(m, g, ic0) - 1;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) dst[(size_t)iv * m.ncell + c] = v[ic0][iv];
    }
}

__global__ __launch_bounds__(SBLOCK) void k_set_unew(MeshDev m, LevelDev lv) {
    const int p = blockIdx.x * SBLOCK + threadIdx.x;
    if (p >= lv.noct) return;
    copy_oct(m, lv.grid[p], 0xff, m.uold, m.unew);
}

__global__ __launch_bounds__(SBLOCK) void k_set_uold(MeshDev m, LevelDev lv) {
    const int p = blockIdx.x * SBLOCK + threadIdx.x;
    if (p >= lv.noct) return;
    copy_oct(m, lv.grid[p], lv.leaf[p], m.unew, m.uold);
}

RH_DEV double fval(const MeshDev& m, size_t c, int idim) {
    return m.f ? m.f[(size_t)(idim - 1) * m.ncell + c] : 0.0;
}

// HydroSolver.cpp:980-998 per leaf cell of the level: gravity kick into unew.
__global__ __launch_bounds__(BLOCK) void k_gravity_source(MeshDev m, LevelDev lv, double dt) {
    const int t = threadIdx.x;
    const int p = blockIdx.x * OCTS_PER_BLOCK + (t % OCTS_PER_BLOCK);
    const int ic0 = t / OCTS_PER_BLOCK;
    if (p >= lv.noct) return;
    if (!((lv.leaf[p] >> ic0) & 1)) return;
    const size_t c = (size_t)cell_id(m, lv.grid[p], ic0) - 1;
    const double d = sclamp(m.uold[c], 1e-10, 1e6);
    double e = m.unew[(size_t)4 * m.ncell + c];
    for (int idim = 1; idim <= 3; ++idim) {
        const double f = fval(m, c, idim);
        const size_t k = (size_t)idim * m.ncell + c;
        double mom = m.unew[k] + d * f * dt;
        mom = sclamp(mom, -d * 10.0, d * 10.0);
        m.unew[k] = mom;
        e = e + (m.uold[k] * f * dt + 0.5 * d * f * f * dt * dt);
    }
    m.unew[(size_t)4 * m.ncell + c] = e;
}

// HydroSolver.cpp:1000-1043 per leaf cell of the level: gravity kick of uold
// keeping the internal energy (nener = 0, no barotropic EOS).
__global__ __launch_bounds__(BLOCK) void k_synchro(MeshDev m, LevelDev lv, double dt, double gamma) {
    const int t = threadIdx.x;
    const int p = blockIdx.x * OCTS_PER_BLOCK + (t % OCTS_PER_BLOCK);
    const int ic0 = t / OCTS_PER_BLOCK;
    if (p >= lv.noct) return;
    if (!((lv.leaf[p] >> ic0) & 1)) return;
    const size_t c = (size_t)cell_id(m, lv.grid[p], ic0) - 1;
    double u[5];
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) u[iv] = m.uold[(size_t)iv * m.ncell + c];
    const double d = sclamp(u[0], 1e-10, 1e6);
    double v2_old = 0;
    for (int idim = 1; idim <= 3; ++idim) {
        const double v = u[idim] / d;
        v2_old = v2_old + v * v;
    }
    const double e_nonthermal = 0.0;
    const double e_int = smax(u[4] - 0.5 * d * v2_old - e_nonthermal, d * 1e-10 / (gamma - 1.0));
    for (int idim = 1; idim <= 3; ++idim) {
        u[idim] = u[idim] + d * fval(m, c, idim) * dt;
        u[idim] = sclamp(u[idim], -d * 10.0, d * 10.0);
    }
    double v2_new = 0;
    for (int idim = 1; idim <= 3; ++idim) {
        const double v = u[idim] / d;
        v2_new = v2_new + v * v;
    }
    u[4] = e_int + 0.5 * d * v2_new + e_nonthermal;
#pragma unroll
    for (int iv = 1; iv < 5; ++iv) m.uold[(size_t)iv * m.ncell + c] = u[iv];
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

// Courant candidate of one leaf cell, HydroSolver.cpp:878-977.
RH_DEV double courant_cand_u(const MeshDev& m, const double u[5], int id, double dx, double gamma,
                             double courant_factor) {
    double d = smax(u[0], 1e-10), v2 = 0.0;
    for (int i = 1; i <= 3; ++i) {
        double v = u[i] / d;
        v2 = v2 + v * v;
    }
    // The reference subtracts a non-thermal energy of 0 and adds a
    // non-thermal sound speed of 0/d: both leave every value unchanged (c2
    // is at least 1e-20 or NaN, and d is NaN only where the velocities and
    // so the result are), so they are left out.
    double pr = smax((u[4] - 0.5 * d * v2) * (gamma - 1.0), d * 1e-10);
    pr = smin(pr, d * 1e2);
    double c2 = smax(gamma * pr / smax(d, 1e-10), 1e-20);
    double cs = ssqrt(smax(c2, 1e-10));
    double u_wave = 3 * cs;
    for (int i = 1; i <= 3; ++i) u_wave = u_wave + sabs(u[i] / d);
    if (!m.f) {
        // No gravity: a_grav = 0, so alpha = max(0 * dx / u_wave^2, 1e-4) =
        // 1e-4 wherever u_wave is a number (it is at least 3e-5); where it
        // is NaN the result is NaN either way.
        const double s1 = ssqrt(1.0 + 2.0 * courant_factor * 1e-4) - 1.0;
        return (dx / u_wave) * s1 / 1e-4;
    }
    double a_grav = 0.0;
    for (int i = 1; i <= 3; ++i) a_grav = a_grav + sabs(m.f[(size_t)(i - 1) * m.ncell + id - 1]);
    double alpha = a_grav * dx / (u_wave * u_wave);
    alpha = smax(alpha, 0.0001);
    return (dx / u_wave) * (ssqrt(1.0 + 2.0 * courant_factor * alpha) - 1.0) / alpha;
}

RH_DEV double courant_cand(const MeshDev& m, int id, double dx, double gamma, double courant_factor) {
    double u[5];
    load_u(m.uold, m.ncell, id, u);
    return courant_cand_u(m, u, id, dx, gamma, courant_factor);
}

// Block minimum of per-thread candidates into part[blockIdx.x].
RH_DEV void block_min(double cand, double* part) {
    __shared__ double s_min[SBLOCK / 64];
    const int t = threadIdx.x;
    cand = wave_min(cand);
    if ((t & 63) == 0) s_min[t >> 6] = cand;
    __syncthreads();
    if (t == 0) {
        double v = s_min[0];
        for (int w = 1; w < SBLOCK / 64; ++w) v = min_skip_nan(v, s_min[w]);
        part[blockIdx.x] = v;
    }
}

// Block minimum over an NT-thread block (NT a multiple of 64).
template <int NT> RH_DEV void block_min_n(double cand, double* part) {
    __shared__ double s_min[NT / 64];
    const int t = threadIdx.x;
    cand = wave_min(cand);
    if ((t & 63) == 0) s_min[t >> 6] = cand;
    __syncthreads();
    if (t == 0) {
        double v = s_min[0];
        for (int w = 1; w < NT / 64; ++w) v = min_skip_nan(v, s_min[w]);
        part[blockIdx.x] = v;
    }
}

// Per leaf cell, then a block minimum (one thread per oct).  SET_UNEW: the
// same pass also copies every cell of the oct from uold into unew
// (set_unew, which reads the same values), for a driver that runs the two
// back to back.
template <bool SET_UNEW>
__global__ __launch_bounds__(SBLOCK) void k_courant_partial(MeshDev m, LevelDev lv, double dx, double gamma,
                                                            double courant_factor, double* part) {
    const int p = blockIdx.x * SBLOCK + threadIdx.x;
    double cand = __builtin_inf();
    if (p < lv.noct) {
        const int g = lv.grid[p];
        const int lm = lv.leaf[p];
        if constexpr (SET_UNEW) {
            // All eight cells loaded before the copies are stored (uold and
            // unew never overlap), then the candidates of the leaf cells.
            double v[8][5];
            const double* __restrict__ src = m.uold;
            double* __restrict__ dst = m.unew;
#pragma unroll
            for (int ic0 = 0; ic0 < 8; ++ic0) {
                const size_t c = (size_t)cell_id(m, g, ic0) - 1;
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) v[ic0][iv] = src[(size_t)iv * m.ncell + c];
            }
#pragma unroll
            for (int ic0 = 0; ic0 < 8; ++ic0) {
                const size_t c = (size_t)cell_id(m, g, ic0) - 1;
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) dst[(size_t)iv * m.ncell + c] = v[ic0][iv];
            }
#pragma unroll
            for (int ic0 = 0; ic0 < 8; ++ic0)
                if ((lm >> ic0) & 1)
                    cand = min_skip_nan(cand, courant_cand_u(m, v[ic0], cell_id(m, g, ic0), dx, gamma, courant_factor));
        } else {
            for (int ic0 = 0; ic0 < 8; ++ic0)
                if ((lm >> ic0) & 1) cand = min_skip_nan(cand, courant_cand(m, cell_id(m, g, ic0), dx, gamma, courant_factor));
        }
    }
    block_min(cand, part);
}

// The end of one step of a level and the start of the next in one pass
// over the level's cells: set_uold (uold <- unew on leaf cells), the Courant
// candidates of the new state, then set_unew (unew <- uold; for a leaf cell
// it already holds that value, so only non-leaf cells are written).  The
// state and the block minima are exactly those of the three kernels run in
// that order, for 80 B per leaf cell instead of 200.
__global__ __launch_bounds__(SBLOCK) void k_step_join(MeshDev m, LevelDev lv, double dx, double gamma,
                                                      double courant_factor, double* part) {
    const int p = blockIdx.x * SBLOCK + threadIdx.x;
    double cand = __builtin_inf();
    if (p < lv.noct) {
        const int g = lv.grid[p];
        const int lm = lv.leaf[p];
        for (int ic0 = 0; ic0 < 8; ++ic0) {
            const int id = cell_id(m, g, ic0);
            const size_t c = (size_t)id - 1;
            double u[5];
            if ((lm >> ic0) & 1) {
                load_u(m.unew, m.ncell, id, u);
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) m.uold[(size_t)iv * m.ncell + c] = u[iv];
                cand = min_skip_nan(cand, courant_cand_u(m, u, id, dx, gamma, courant_factor));
            } else {
                load_u(m.uold, m.ncell, id, u);
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + c] = u[iv];
            }
        }
    }
    block_min(cand, part);
}

// Ordered 64-bit key of a double: unsigned comparison of keys orders them as
// the doubles (-0 before +0), and NaN maps to the largest key, so a minimum
// over keys never picks a NaN candidate -- the reference's std::min loop
// skips them too.  The minimum of a set is order-independent, so atomics
// give a deterministic result.
RH_DEV unsigned long long dt_key(double v) {
    if (v != v) return ~0ull;
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
RH_DEV double key_dt(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)b);
}

// Ensemble Courant step: per-box minimum into keys[box].  Octs of a box are
// contiguous in the concatenated list, so a block whose first and last oct
// share a box reduces in registers and LDS and issues one atomic; a block
// straddling boxes lets every leaf cell issue its own.
// JOIN: also the end of the previous step and the start of this one, as in
// k_step_join (set_uold on leaf cells, the candidates from the new state,
// set_unew on non-leaf cells).
template <bool JOIN>
__global__ __launch_bounds__(SBLOCK) void k_courant_boxes(MeshDev m, LevelDev lv, double gamma, double courant_factor,
                                                          unsigned long long* keys) {
    __shared__ double s_min[SBLOCK / 64];
    const int t = threadIdx.x;
    const int p0 = blockIdx.x * SBLOCK;
    const int p = p0 + t;
    const int plast = min(p0 + SBLOCK, lv.noct) - 1;
    const int b0 = lv.obox[p0], b1 = lv.obox[plast];
    double cand = __builtin_inf();
    int bx = b0;
    if (p < lv.noct) {
        bx = lv.obox[p];
        const double dx = lv.box_len[bx];
        const int g = lv.grid[p];
        const int lm = lv.leaf[p];
        for (int ic0 = 0; ic0 < 8; ++ic0) {
            const int id = cell_id(m, g, ic0);
            const bool leafc = (lm >> ic0) & 1;
            if constexpr (JOIN) {
                const size_t c = (size_t)id - 1;
                double u[5];
                if (leafc) {
                    load_u(m.unew, m.ncell, id, u);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) m.uold[(size_t)iv * m.ncell + c] = u[iv];
                    cand = min_skip_nan(cand, courant_cand_u(m, u, id, dx, gamma, courant_factor));
                } else {
                    load_u(m.uold, m.ncell, id, u);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + c] = u[iv];
                }
            } else if (leafc) {
                cand = min_skip_nan(cand, courant_cand(m, id, dx, gamma, courant_factor));
            }
        }
    }
    if (b0 == b1) {
        cand = wave_min(cand);
        if ((t & 63) == 0) s_min[t >> 6] = cand;
        __syncthreads();
        if (t == 0) {
            double v = s_min[0];
            for (int w = 1; w < SBLOCK / 64; ++w) v = min_skip_nan(v, s_min[w]);
            atomicMin(&keys[b0], dt_key(v));
        }
    } else if (cand != __builtin_inf()) {
        atomicMin(&keys[bx], dt_key(cand));
    }
}

__global__ __launch_bounds__(256) void k_fill_keys(unsigned long long* keys, int n, double value) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) keys[i] = dt_key(value);
}

__global__ __launch_bounds__(256) void k_keys_to_dt(const unsigned long long* keys, int n, double* dt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < n) dt[i] = key_dt(keys[i]);
}

// One workgroup of 1024 threads; each keeps four independent minima so that
// four loads are in flight per round instead of a chain of dependent ones
// (the partial minima number one per 32 octs, 65,536 at 2^21 octs).
RH_DEV void amr_times(double* T, int op, int il, int lmin, double nsub_prev, double nsub);

// at.T non-null: after storing the minimum, the adaptive run's time-step
// bookkeeping of the level step (op 0 of amr_times, then at.then_op when
// set) in the same thread.
__global__ __launch_bounds__(1024) void k_courant_final(const double* part, int npart, double dt_init, double* dt,
                                                        AmrTimesArgs at) {
    __shared__ double s_min[16];
    const int t = threadIdx.x;
    double v0 = __builtin_inf(), v1 = v0, v2 = v0, v3 = v0;
    int i = t;
    for (; i + 3 * 1024 < npart; i += 4 * 1024) {
        v0 = min_skip_nan(v0, part[i]);
        v1 = min_skip_nan(v1, part[i + 1024]);
        v2 = min_skip_nan(v2, part[i + 2 * 1024]);
        v3 = min_skip_nan(v3, part[i + 3 * 1024]);
    }
    for (; i < npart; i += 1024) v0 = min_skip_nan(v0, part[i]);
    double v = min_skip_nan(min_skip_nan(v0, v1), min_skip_nan(v2, v3));
    v = wave_min(v);
    if ((t & 63) == 0) s_min[t >> 6] = v;
    __syncthreads();
    if (t == 0) {
        double r = dt_init;
        for (int w = 0; w < 16; ++w) r = min_skip_nan(r, s_min[w]);
        *dt = r;
        if (at.T) {
            amr_times(at.T, 0, at.il, at.lmin, at.nsub_prev, 0.0);
            if (at.then_op) amr_times(at.T, at.then_op, at.il, at.lmin, 0.0, at.then_nsub);
        }
    }
}

// ---------------------------------------------------------------------------
// Godunov update (FP64, bit-faithful).

// Time step and cell width of oct p: the kernel arguments for a single mesh,
// its box's entries for an ensemble.
RH_DEV void step_of(const LevelDev& lv, int p, const double* d_dt, double& dt, double& dx, double& inv_dx) {
    if (lv.obox) {
        const int bx = lv.obox[p];
        dt = d_dt[bx];
        dx = lv.box_len[bx];
        inv_dx = lv.box_len[lv.nbox + bx];
    } else {
        dt = *d_dt;
    }
}

struct Nbr {
    int cell;  // 1-based cell index (<= 0: boundary)
    int pos;   // oct position in the level list, or -1 for a record
    int ic0;   // cell within that oct
    int rec;   // record index when pos < 0
};

template <typename TC> RH_DEV void narrow(const double a[5], TC b[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) b[i] = (TC)a[i];
}

// Primitive state of a cell for the slopes, traces and Riemann problems:
// FP64 ctoprim in the FP64 form; in the mixed form the conservative state is
// narrowed to FP32 and converted there.
template <typename TC>
RH_DEV void face_prim(const double u[5], TC q[5], const Gas<double>& gd, const Gas<TC>& gc) {
    if constexpr (sizeof(TC) == 8) {
        ctoprim(u, q, gd);
    } else {
        TC uc[5];
        narrow(u, uc);
        ctoprim(uc, q, gc);
    }
}

// One face's Riemann problem: the FP64 form calls the bit-faithful solver on
// double states; the mixed form solves it in FP32 (its flux is widened to
// FP64 where it enters the update or a reflux).  ql/qr are rotated; F comes
// back rotated.
template <int RIEMANN, bool MIXED, typename TC>
RH_DEV void solve_face(const TC ql[5], const TC qr[5], TC F[5], const Gas<TC>& gc, const Gas<double>& gd) {
    if constexpr (MIXED) riemann_f<RIEMANN>(ql, qr, F, gc);
    else riemann<RIEMANN>(ql, qr, F, gd);
}

// Octs per workgroup of the Godunov kernel (8 threads per oct).
constexpr int G_OCTS = 16;
constexpr int G_BLOCK = G_OCTS * 8;

// A leaf cell's own update and clamps, HydroSolver.cpp:489-547.
RH_DEV void own_update(double u[5], const double fs[5], double dtdx, const GasParams& gp) {
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) u[iv] = u[iv] + dtdx * fs[iv];
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

// Sort network for 6 keys (12 compare-exchanges) carrying slot indices.
RH_DEV void sort6(long long ks[6], int ix[6]) {
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
}

constexpr long long KMAX = 0x7fffffffffffffffLL;

// Ordered accumulation of one leaf cell (HydroSolver.cpp:476-547 in
// processing order).  s_r/s_k hold, for this thread, the six incoming
// contributions and their keys (negative key: none): sorted by key, those
// ordered before the cell are subtracted, then the cell's own update and
// clamps are applied, then the rest.
template <int NT>
RH_DEV void ordered_apply(const MeshDev& m, int cC, long long own_key, const double fs[5], double dtdx,
                          const GasParams& gp, double (*s_r)[NT], long long (*s_k)[NT], int t,
                          const double* unew_pre = nullptr, double* u_out = nullptr) {
    const long long KOWN = own_key * 8 + 6;
    long long ks[6];
    int ix[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const long long kf = s_k[f][t];
        ks[f] = kf < 0 ? KMAX : kf;
        ix[f] = f;
    }
    sort6(ks, ix);
    double rs[6][5];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) rs[k][iv] = s_r[ix[k] * 5 + iv][t];
    }
    double u[5];
    if (unew_pre) copy5(unew_pre, u);
    else load_u(m.unew, m.ncell, cC, u);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const bool take = ks[k] < KOWN;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            const double nv = u[iv] - rs[k][iv];
            u[iv] = take ? nv : u[iv];
        }
    }
    own_update(u, fs, dtdx, gp);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const bool take = ks[k] > KOWN && ks[k] != KMAX;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            const double nv = u[iv] - rs[k][iv];
            u[iv] = take ? nv : u[iv];
        }
    }
    const size_t ci = (size_t)cC - 1;
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + ci] = u[iv];
    if (u_out) copy5(u, u_out);
}

// One thread per cell of the level's octs; non-leaf cells exit.  For each
// dimension the thread builds the 5-cell line L2 L1 C R1 R2 through its oct
// and the two neighbour octs, computes its own two face fluxes and the two
// fluxes its same-level neighbours L1/R1 compute for the shared faces, then
// adds everything into unew in the reference's processing order.
template <int RIEMANN, bool MIXED, bool EXACT>
__global__ __launch_bounds__(G_BLOCK) void k_godunov(MeshDev m, LevelDev lv, GasParams gp, const double* d_dt,
                                                      double dx, double inv_dx) {
    using TC = typename std::conditional<MIXED, float, double>::type;
    // Per-thread reflux payload: the six incoming contributions and their
    // processing keys, kept in LDS so they can be read back in sorted order.
    __shared__ double s_r[30][G_BLOCK];
    __shared__ long long s_k[6][G_BLOCK];
    const int t = threadIdx.x;
    const int p = blockIdx.x * G_OCTS + (t % G_OCTS);
    const int ic0 = t / G_OCTS;
    if (p >= lv.noct) return;
    if (!((lv.leaf[p] >> ic0) & 1)) return;

    const Gas<double> gd{gp.gamma, gp.gm1, gp.smalle, gp.smallp, 1.0 / gp.gm1};
    const Gas<TC> gc{(TC)gp.gamma, (TC)gp.gm1, (TC)gp.smalle, (TC)gp.smallp, (TC)(1.0 / gp.gm1)};
    const int slope_type = gp.slope_type, interpol_type = gp.interpol_type;
    double dt;
    step_of(lv, p, d_dt, dt, dx, inv_dx);
    const double dtdx = dt / dx;
    const TC dtc = (TC)dt, dxc = (TC)dx;
    const Len<TC, EXACT> len{dxc, (TC)inv_dx};
    const int noct = lv.noct;
    const int cC = cell_id(m, lv.grid[p], ic0);

    double uC[5];
    TC qC[5];
    load_u(m.uold, m.ncell, cC, uC);
    {
        face_prim(uC, qC, gd, gc);
    }

    double fs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};

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
    // Primitive state of a cell as read by the slope / ghost stencils
    // (face_prim): a boundary index falls back to `fallback` (the reference
    // copies the centre cell's state there; bound_type is never filled).
    auto prim = [&](int icell, const TC fallback[5], TC q[5]) {
        if (icell > 0) {
            double u[5];
            load_u(m.uold, m.ncell, icell, u);
            face_prim(u, q, gd, gc);
        } else {
            copy5(fallback, q);
        }
    };
    // Boundary face: own state with the normal momentum flipped, zero slopes
    // (HydroSolver.cpp:130-145).
    auto boundary = [&](int d, TC qm[5], TC qp[5]) {
        double u[5];
        load_u(m.uold, m.ncell, cC, u);
#pragma unroll
        for (int i = 1; i <= 3; ++i)
            if (i == 1 + d) u[i] = u[i] * -1.0;
        TC q[5];
        face_prim(u, q, gd, gc);
        rotate(q, d);
        TC dq[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
        trace(q, dq, dtc, dxc, qm, qp, gc);
    };

#pragma unroll 1
    for (int d = 0; d < 3; ++d) {
        const int b = 1 << d;
        const int par = (ic0 >> d) & 1;
        const int A = lv.nbo[(size_t)(2 * d) * noct + p];
        const int B = lv.nbo[(size_t)(2 * d + 1) * noct + p];
        const Nbr L1 = par ? same(p, ic0 ^ b) : (A >= 0 ? same(A, ic0 ^ b) : from_rec(A));
        const Nbr R1 = par ? (B >= 0 ? same(B, ic0 ^ b) : from_rec(B)) : same(p, ic0 ^ b);

        TC qL1[5], qR1[5];
        prim(L1.cell, qC, qL1);
        prim(R1.cell, qC, qR1);

        // Own face states (trace pass) and this cell's ghost states as its
        // neighbours see them (coarse-ghost branch with n = this cell).
        TC OCm[5], OCp[5], GCm[5], GCp[5];
        own_trace(qL1, qC, qR1, d, slope_type, dtc, len, OCm, OCp, gc);
        ghost_trace(qL1, qC, qR1, d, interpol_type, dtc, len, GCm, GCp, gc);

        // ---- minus face (side 0): ql = ghost qm of L1, qr = own qp
        {
            TC F[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
            long long K = -1;
            if (L1.pos >= 0) {
                if ((lv.leaf[L1.pos] >> L1.ic0) & 1) {
                    TC qL2[5];
                    const int L2cell = A >= 0 ? cell_id(m, lv.grid[A], ic0) : lv.rec[-1 - A].x;
                    prim(L2cell, qL1, qL2);
                    TC Gm[5], Gp[5];
                    ghost_trace(qL2, qL1, qC, d, interpol_type, dtc, len, Gm, Gp, gc);
                    solve_face<RIEMANN, MIXED>(Gm, OCp, F, gc, gd);
                    rotate(F, d);
                    // L1's own plus-face flux into this cell (reflux).
                    TC Om[5], Op[5];
                    TC Fin[5];
                    own_trace(qL2, qL1, qC, d, slope_type, dtc, len, Om, Op, gc);
                    solve_face<RIEMANN, MIXED>(Om, GCp, Fin, gc, gd);
                    rotate(Fin, d);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) s_r[(2 * d) * 5 + iv][t] = dtdx * -1.0 * Fin[iv] * 0.125;
                    K = ((long long)L1.pos * 8 + L1.ic0) * 8 + (2 * d + 1);
                }
            } else {
                const int4 rc = lv.rec[L1.rec];
                if (rc.w & REC_REFINED_DEV) {
                } else if (rc.w & REC_BOUNDARY_DEV) {
                    TC Bm[5], Bp[5];
                    boundary(d, Bm, Bp);
                    solve_face<RIEMANN, MIXED>(Bm, OCp, F, gc, gd);
                    rotate(F, d);
                } else {
                    TC qn[5], qnl[5], qnr[5];
                    prim(rc.x, qC, qn);
                    prim(rc.y, qn, qnl);
                    prim(rc.z, qn, qnr);
                    TC Gm[5], Gp[5];
                    ghost_trace(qnl, qn, qnr, d, interpol_type, dtc, len, Gm, Gp, gc);
                    solve_face<RIEMANN, MIXED>(Gm, OCp, F, gc, gd);
                    rotate(F, d);
                    const int slot = lv.rec_slot[L1.rec] + face_slot(ic0, d);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) lv.slot_r[(size_t)slot * 5 + iv] = dtdx * 1.0 * F[iv] * 0.125;
                }
            }
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) fs[iv] = fs[iv] + 1.0 * F[iv];
            s_k[2 * d][t] = K;
        }
        // ---- plus face (side 1): ql = own qm, qr = ghost qp of R1
        {
            TC F[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
            long long K = -1;
            if (R1.pos >= 0) {
                if ((lv.leaf[R1.pos] >> R1.ic0) & 1) {
                    TC qR2[5];
                    const int R2cell = B >= 0 ? cell_id(m, lv.grid[B], ic0) : lv.rec[-1 - B].x;
                    prim(R2cell, qR1, qR2);
                    TC Gm[5], Gp[5];
                    ghost_trace(qC, qR1, qR2, d, interpol_type, dtc, len, Gm, Gp, gc);
                    solve_face<RIEMANN, MIXED>(OCm, Gp, F, gc, gd);
                    rotate(F, d);
                    // R1's own minus-face flux into this cell (reflux).
                    TC Om[5], Op[5];
                    TC Fin[5];
                    own_trace(qC, qR1, qR2, d, slope_type, dtc, len, Om, Op, gc);
                    solve_face<RIEMANN, MIXED>(GCm, Op, Fin, gc, gd);
                    rotate(Fin, d);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) s_r[(2 * d + 1) * 5 + iv][t] = dtdx * 1.0 * Fin[iv] * 0.125;
                    K = ((long long)R1.pos * 8 + R1.ic0) * 8 + (2 * d);
                }
            } else {
                const int4 rc = lv.rec[R1.rec];
                if (rc.w & REC_REFINED_DEV) {
                } else if (rc.w & REC_BOUNDARY_DEV) {
                    TC Bm[5], Bp[5];
                    boundary(d, Bm, Bp);
                    solve_face<RIEMANN, MIXED>(OCm, Bp, F, gc, gd);
                    rotate(F, d);
                } else {
                    TC qn[5], qnl[5], qnr[5];
                    prim(rc.x, qC, qn);
                    prim(rc.y, qn, qnl);
                    prim(rc.z, qn, qnr);
                    TC Gm[5], Gp[5];
                    ghost_trace(qnl, qn, qnr, d, interpol_type, dtc, len, Gm, Gp, gc);
                    solve_face<RIEMANN, MIXED>(OCm, Gp, F, gc, gd);
                    rotate(F, d);
                    const int slot = lv.rec_slot[R1.rec] + face_slot(ic0, d);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) lv.slot_r[(size_t)slot * 5 + iv] = dtdx * -1.0 * F[iv] * 0.125;
                }
            }
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) fs[iv] = fs[iv] + -1.0 * F[iv];
            s_k[2 * d + 1][t] = K;
        }
    }

    ordered_apply<G_BLOCK>(m, cC, (long long)p * 8 + ic0, fs, dtdx, gp, s_r, s_k, t);
}

// The ordered accumulation of ordered_apply with the six incoming face
// fluxes in LDS (slot f = face f of the receiving cell: the neighbour's own
// flux on the shared face, in the compute type of the form) and their keys in
// registers (negative: none).  Slot f's contribution is the reference's
// ((dtdx * sign) * F) * 0.125 with the neighbour's face sign: -1 for a
// neighbour on the minus side (its plus face), +1 on the plus side.
// With CONTRIB the slots already hold the FP64 contributions themselves.
template <int NT, bool CONTRIB = false, typename TF>
RH_DEV void ordered_apply_flux(const MeshDev& m, int cC, long long own_key, const double fs[5], double dtdx,
                               const GasParams& gp, const TF (*s_f)[NT], const long long kin[6], int t,
                               const double* unew_pre, double* u_out) {
    const long long KOWN = own_key * 8 + 6;
    long long ks[6];
    int ix[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        ks[f] = kin[f] < 0 ? KMAX : kin[f];
        ix[f] = f;
    }
    sort6(ks, ix);
    double rs[6][5];
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double sdt = dtdx * ((ix[k] & 1) ? 1.0 : -1.0);
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            const double v = (double)s_f[ix[k] * 5 + iv][t];
            rs[k][iv] = CONTRIB ? v : sdt * v * 0.125;
        }
    }
    double u[5];
    copy5(unew_pre, u);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const bool take = ks[k] < KOWN;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            const double nv = u[iv] - rs[k][iv];
            u[iv] = take ? nv : u[iv];
        }
    }
    own_update(u, fs, dtdx, gp);
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const bool take = ks[k] > KOWN && ks[k] != KMAX;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            const double nv = u[iv] - rs[k][iv];
            u[iv] = take ? nv : u[iv];
        }
    }
    const size_t ci = (size_t)cC - 1;
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + ci] = u[iv];
    if (u_out) copy5(u, u_out);
}

// Keys of the six incoming contributions of cell ic0 of oct p, from the
// tables (face slot f; -1: none): the sibling's when it is a leaf, X's when
// its oct is a same-level neighbour and X a leaf.
RH_DEV void incoming_keys(const LevelDev& lv, int p, int ic0, int lm, long long kin[6]) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const bool par = (ic0 >> d) & 1;
        const int icS = ic0 ^ (1 << d);
        const long long kS = ((lm >> icS) & 1) ? ((long long)p * 8 + icS) * 8 + (par ? 2 * d + 1 : 2 * d) : -1;
        const int E = lv.nbo[(size_t)(2 * d + (par ? 1 : 0)) * lv.noct + p];
        long long kX = -1;
        if (E >= 0 && ((lv.leaf[E] >> icS) & 1)) kX = ((long long)E * 8 + icS) * 8 + (par ? 2 * d : 2 * d + 1);
        kin[2 * d] = par ? kS : kX;
        kin[2 * d + 1] = par ? kX : kS;
    }
}

// ---------------------------------------------------------------------------
// Oct-in-lanes variant: the eight cells of an oct sit in eight adjacent lanes.
// Per dimension each cell has one neighbour inside its oct (the sibling, lane
// ^ (1<<d)) and one across the oct face (X, in the neighbour oct E).  The
// sibling's primitives, traces and its own flux on the shared face come by
// cross-lane shuffle; only X and the cell beyond it are loaded.  Per cell:
// 4 traces and 3 Riemann solves per dimension (the per-cell kernel above
// needs 6 and 4).  The code is the same for both parities of a cell along d
// (operands chosen with selects), so lanes never diverge on geometry.
template <typename T> RH_DEV void shfl5(const T in[5], T out[5], int mask) {
#pragma unroll
    for (int i = 0; i < 5; ++i) out[i] = __shfl_xor(in[i], mask, 64);
}
template <typename T> RH_DEV void sel5(bool c, const T a[5], const T b[5], T out[5]) {
#pragma unroll
    for (int i = 0; i < 5; ++i) out[i] = c ? a[i] : b[i];
}

// Workgroup index -> oct range, XCD-aware.  The dispatcher deals
// workgroups round-robin to the 8 XCDs, each with its own L2, so workgroup
// x runs on XCD x % 8.  Renumbering gives each XCD one contiguous share of
// the list (XCD k takes ranges k*q.. in order, q = n/8, the first n%8 XCDs
// one more), so octs that are neighbours in the list's Morton order, whose
// cells each other's line gathers read, share an L2 instead of every
// neighbour pair fetching the same lines from HBM on two XCDs.  A bijection
// of [0, n) for any n.
RH_DEV int xcd_block(int x, int n) {
    const int xcd = x & 7, idx = x >> 3, q = n >> 3, r = n & 7;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Octs per workgroup: 16 (two waves, kernel variant 1) or 8 (one wave,
// variant 2); rh_autotune times both.
template <int OCTS> constexpr int oct_block() { return OCTS * 8; }

// set_unew (uold -> unew on every cell of the level's octs) with the memory
// access pattern of k_godunov_oct's own-cell loads and unew stores: the
// eight cells of an oct in eight adjacent lanes, 8 octs per 64-lane
// workgroup, workgroups in xcd_block order, so a wave's access to one
// variable touches eight 64-byte pieces (one per cell index; eight
// consecutive grids each on a uniform level).  A copy of known size whose
// byte counters calibrate those of the Godunov kernel.
__global__ __launch_bounds__(64) void k_set_unew_oct_lanes(MeshDev m, LevelDev lv) {
    const int t = threadIdx.x;
    const int p = xcd_block(blockIdx.x, gridDim.x) * 8 + (t >> 3);
    if (p >= lv.noct) return;
    const size_t c = (size_t)cell_id(m, lv.grid[p], t & 7) - 1;
    double v[5];
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) v[iv] = m.uold[(size_t)iv * m.ncell + c];
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + c] = v[iv];
}

// Occupancy: the FP64 form runs two waves per SIMD (at most 256 VGPRs), the
// mixed form three (at most 168).  LDS per oct: the incoming contributions
// and their keys, 2.25 KB in the FP64 form, and the FP32 incoming fluxes,
// 0.9 KB, in the mixed form.
// The registers above what the arithmetic needs hold the next dimension's
// neighbour states, gathered ahead of their use.
//
// Fused step (a level whose cells are all leaves and whose neighbours are
// all at the same level, stepped with uold and unew exchanged between
// steps): `pre` is uold, which is what set_unew would have copied into unew,
// and, except on the last step, each leaf cell also computes the Courant
// candidate of its new state -- the state set_uold would copy into uold --
// with a block minimum into cpart[blockIdx.x] (a single mesh) or, for an
// ensemble, into its box's key in ckeys as k_courant_boxes does.  Otherwise
// pre is unew and cpart and ckeys are null.
template <int RIEMANN, bool MIXED, bool EXACT, int OCTS>
__global__ __launch_bounds__(OCTS * 8) __attribute__((amdgpu_waves_per_eu(MIXED ? 3 : 2, 8)))
void k_godunov_oct(MeshDev m, LevelDev lv, GasParams gp, const double* d_dt,
                                                          double dx, double inv_dx, const double* pre,
                                                          double* cpart, unsigned long long* ckeys,
                                                          double courant_factor) {
    using TC = typename std::conditional<MIXED, float, double>::type;
    constexpr int NT = oct_block<OCTS>();
    // The six incoming face contributions of each cell (slot f*5 + iv).
    // The FP64 form stores each contribution ((dtdx*sign)*F)*0.125 and its
    // key as it is found; the mixed form keeps the FP32 flux F (the
    // contribution is formed in FP64 when it is applied) and recomputes the
    // keys from the tables at the end, which halves its LDS so that three
    // waves share a SIMD.
    __shared__ TC s_f[30][NT];
    __shared__ long long s_k[MIXED ? 1 : 6][NT];
    const int t = threadIdx.x;
    const int ic0 = t & 7;
    const int praw = xcd_block(blockIdx.x, gridDim.x) * OCTS + (t >> 3);
    const int noct = lv.noct;
    // Lanes past the end of the list work on the last oct (valid memory) and
    // write nothing, so every lane of an oct group takes part in the shuffles.
    const bool valid = praw < noct;
    const int p = valid ? praw : noct - 1;
    const int lm = lv.leaf[p];
    const bool leaf = valid && ((lm >> ic0) & 1);

    const Gas<double> gd{gp.gamma, gp.gm1, gp.smalle, gp.smallp, 1.0 / gp.gm1};
    const Gas<TC> gc{(TC)gp.gamma, (TC)gp.gm1, (TC)gp.smalle, (TC)gp.smallp, (TC)(1.0 / gp.gm1)};
    const int slope_type = gp.slope_type, interpol_type = gp.interpol_type;
    double dt;
    step_of(lv, p, d_dt, dt, dx, inv_dx);
    const double dtdx = dt / dx;
    const TC dtc = (TC)dt, dxc = (TC)dx;
    const Len<TC, EXACT> len{dxc, (TC)inv_dx};
    const int cC = cell_id(m, lv.grid[p], ic0);

    double uC[5];
    TC qC[5];
    load_u(m.uold, m.ncell, cC, uC);
    {
        face_prim(uC, qC, gd, gc);
    }
    double fs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};

    auto prim = [&](int icell, const TC fallback[5], TC q[5]) {
        if (icell > 0) {
            double u[5];
            load_u(m.uold, m.ncell, icell, u);
            face_prim(u, q, gd, gc);
        } else {
            copy5(fallback, q);
        }
    };
    auto to_prim = [&](const double u[5], TC q[5]) {
        face_prim(u, q, gd, gc);
    };

    // Across the oct face along dimension d: the neighbour oct E (or a
    // record), the cell X beyond the face and, when X's oct is a same-level
    // leaf, the cell X2 beyond X.  The next dimension's X and X2 are
    // gathered while the current one is computed: two waves per SIMD cannot
    // hide the table -> grid -> uold chain of dependent loads on their own.
    struct Across {
        int E;
        int4 rc;
        int Xcell, X2cell;
    };
    auto across = [&](int d) {
        Across a;
        const bool pd = (ic0 >> d) & 1;
        const int icS = ic0 ^ (1 << d);
        a.E = lv.nbo[(size_t)(2 * d + (pd ? 1 : 0)) * noct + p];
        a.rc = make_int4(0, 0, 0, 0);
        a.X2cell = 0;
        if (a.E >= 0) {
            const int g = lv.grid[a.E];
            a.Xcell = cell_id(m, g, icS);
            if ((lv.leaf[a.E] >> icS) & 1) a.X2cell = cell_id(m, g, ic0);
        } else {
            a.rc = lv.rec[-1 - a.E];
            a.Xcell = a.rc.x;
        }
        return a;
    };
    auto gather = [&](const Across& a, double uX[5], double uX2[5]) {
        if (a.Xcell > 0) load_u(m.uold, m.ncell, a.Xcell, uX);
        if (a.X2cell > 0) load_u(m.uold, m.ncell, a.X2cell, uX2);
    };
    Across nx = across(0);
    double uXn[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, uX2n[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    gather(nx, uXn, uX2n);

    // Unrolling the dimension loop makes the rotations by a constant
    // dimension free; it pays for every form except FP64 HLLC, whose code
    // would triple and lose more in instruction fetch than it gains.
    constexpr int DIM_UNROLL = (MIXED || RIEMANN != 2) ? 3 : 1;
#pragma unroll DIM_UNROLL
    for (int d = 0; d < 3; ++d) {
        const int b = 1 << d;
        const bool par = (ic0 >> d) & 1;  // 0: sibling on the plus side, X on the minus side
        const int icS = ic0 ^ b;
        const Across cur = nx;
        double uX[5], uX2[5];
        copy5(uXn, uX);
        copy5(uX2n, uX2);
        if (d < 2) {
            nx = across(d + 1);
            gather(nx, uXn, uX2n);
        } else {
            // Last dimension: the cell's own unew entry before this update
            // (no other thread writes it in this kernel) takes the freed
            // registers.
            load_u(pre, m.ncell, cC, uXn);
        }
        const int E = cur.E;
        const int4 rc = cur.rc;
        TC qS[5], qX[5];
        shfl5(qC, qS, b);
        if (cur.Xcell > 0) to_prim(uX, qX);
        else copy5(qC, qX);
        TC qL1[5], qR1[5];
        sel5(par, qS, qX, qL1);
        sel5(par, qX, qS, qR1);
        TC OCm[5], OCp[5], GCm[5], GCp[5];
        own_trace(qL1, qC, qR1, d, slope_type, dtc, len, OCm, OCp, gc);
        ghost_trace(qL1, qC, qR1, d, interpol_type, dtc, len, GCm, GCp, gc);
        TC GSm[5], GSp[5];
        shfl5(GCm, GSm, b);
        shfl5(GCp, GSp, b);

        // Internal face: this cell's own flux on the sibling side.
        const bool sib_leaf = (lm >> icS) & 1;
        TC Fint[5];
        {
            TC ql[5], qr[5];
            sel5(par, GSm, OCm, ql);
            sel5(par, OCp, GSp, qr);
            solve_face<RIEMANN, MIXED>(ql, qr, Fint, gc, gd);
            rotate(Fint, d);
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) Fint[iv] = sib_leaf ? Fint[iv] : TC(0);
        }
        TC Fsib[5];
        shfl5(Fint, Fsib, b);  // the sibling's own flux on this cell's side

        // External face (across the oct boundary).
        TC Fext[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
        TC Fin[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
        long long Kin = -1;
        if (E >= 0) {
            if (cur.X2cell > 0) {
                TC qX2[5];
                to_prim(uX2, qX2);
                TC a5[5], c5[5];
                sel5(par, qC, qX2, a5);
                sel5(par, qX2, qC, c5);
                TC GXm[5], GXp[5], OXm[5], OXp[5];
                ghost_trace(a5, qX, c5, d, interpol_type, dtc, len, GXm, GXp, gc);
                own_trace(a5, qX, c5, d, slope_type, dtc, len, OXm, OXp, gc);
                TC ql[5], qr[5];
                sel5(par, OCm, GXm, ql);
                sel5(par, GXp, OCp, qr);
                solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
                rotate(Fext, d);
                sel5(par, GCm, OXm, ql);
                sel5(par, OXp, GCp, qr);
                solve_face<RIEMANN, MIXED>(ql, qr, Fin, gc, gd);
                rotate(Fin, d);
                Kin = ((long long)E * 8 + icS) * 8 + (par ? 2 * d : 2 * d + 1);
            }
        } else if (rc.w & REC_REFINED_DEV) {
        } else if (rc.w & REC_BOUNDARY_DEV) {
            // The cell's own state, read again rather than kept in registers
            // through the loop for this rare branch.
            double u[5];
            load_u(m.uold, m.ncell, cC, u);
#pragma unroll
            for (int i = 1; i <= 3; ++i)
                if (i == 1 + d) u[i] = u[i] * -1.0;
            TC qb[5];
            face_prim(u, qb, gd, gc);
            rotate(qb, d);
            TC dq0[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
            TC Bm[5], Bp[5];
            trace(qb, dq0, dtc, dxc, Bm, Bp, gc);
            TC ql[5], qr[5];
            sel5(par, OCm, Bm, ql);
            sel5(par, Bp, OCp, qr);
            solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
            rotate(Fext, d);
        } else {
            TC qnl[5], qnr[5];
            prim(rc.y, qX, qnl);
            prim(rc.z, qX, qnr);
            TC Gm[5], Gp[5];
            ghost_trace(qnl, qX, qnr, d, interpol_type, dtc, len, Gm, Gp, gc);
            TC ql[5], qr[5];
            sel5(par, OCm, Gm, ql);
            sel5(par, Gp, OCp, qr);
            solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
            rotate(Fext, d);
            if (leaf) {
                const int slot = lv.rec_slot[-1 - E] + face_slot(ic0, d);
                const double sg = par ? -1.0 : 1.0;
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) lv.slot_r[(size_t)slot * 5 + iv] = dtdx * sg * Fext[iv] * 0.125;
            }
        }

        // Own fluxes in face order (minus face, then plus face).
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            const double Fm = (double)(par ? Fint[iv] : Fext[iv]);
            const double Fp = (double)(par ? Fext[iv] : Fint[iv]);
            fs[iv] = fs[iv] + 1.0 * Fm;
            fs[iv] = fs[iv] + -1.0 * Fp;
        }
        // Incoming fluxes: the sibling's own flux on the shared face (its plus
        // face when it is on the minus side) and X's.
        const int slotS = 2 * d + (par ? 0 : 1);
        const int slotX = 2 * d + (par ? 1 : 0);
        if constexpr (MIXED) {
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) {
                s_f[5 * slotS + iv][t] = Fsib[iv];
                s_f[5 * slotX + iv][t] = Fin[iv];
            }
        } else {
            const double sgS = par ? -1.0 : 1.0;
            const double sgX = par ? 1.0 : -1.0;
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) {
                s_f[5 * slotS + iv][t] = dtdx * sgS * Fsib[iv] * 0.125;
                s_f[5 * slotX + iv][t] = dtdx * sgX * Fin[iv] * 0.125;
            }
            s_k[slotS][t] = sib_leaf ? ((long long)p * 8 + icS) * 8 + (par ? 2 * d + 1 : 2 * d) : -1;
            s_k[slotX][t] = Kin;
        }
    }
    long long kin[6];
    if constexpr (MIXED) {
        incoming_keys(lv, p, ic0, lm, kin);
    } else {
#pragma unroll
        for (int f = 0; f < 6; ++f) kin[f] = s_k[f][t];
    }
    if (!cpart && !ckeys) {
        if (!leaf) return;
        ordered_apply_flux<NT, !MIXED>(m, cC, (long long)p * 8 + ic0, fs, dtdx, gp, s_f, kin, t, uXn, nullptr);
        return;
    }
    double cand = __builtin_inf();
    if (leaf) {
        double un[5];
        ordered_apply_flux<NT, !MIXED>(m, cC, (long long)p * 8 + ic0, fs, dtdx, gp, s_f, kin, t, uXn, un);
        cand = courant_cand_u(m, un, cC, dx, gp.gamma, courant_factor);
    }
    if (cpart) {
        block_min_n<NT>(cand, cpart);
        return;
    }
    // Ensemble: one atomic per workgroup when its octs are all in one box,
    // else one per cell (the minimum of a set is order-independent).
    const int pfirst = praw - (t >> 3);
    const int b0 = lv.obox[pfirst], b1 = lv.obox[min(pfirst + OCTS, noct) - 1];
    if (b0 == b1) {
        __shared__ double s_bmin[NT / 64];
        cand = wave_min(cand);
        if ((t & 63) == 0) s_bmin[t >> 6] = cand;
        __syncthreads();
        if (t == 0) {
            double v = s_bmin[0];
            for (int w = 1; w < NT / 64; ++w) v = min_skip_nan(v, s_bmin[w]);
            atomicMin(&ckeys[b0], dt_key(v));
        }
    } else if (cand != __builtin_inf()) {
        atomicMin(&ckeys[lv.obox[p]], dt_key(cand));
    }
}

// ---------------------------------------------------------------------------
// Dimension-parallel variant (variant 6), for levels with few octs, whose
// update is bound by the length of one wave's instruction stream rather than
// by throughput: three waves per 8 octs, wave d computing dimension d of the
// oct-in-lanes update for the same 64 cells -- its own fluxes on the two
// faces along d and the two incoming contributions of d, exactly as one
// iteration of k_godunov_oct's dimension loop -- then, after a barrier, the
// first wave adds the three dimensions' own fluxes in dimension order and
// applies the contributions in processing order.  The same operations on
// the same operands as k_godunov_oct, so the results are identical; each
// wave executes about a third of the instructions.  `pre`, cpart, ckeys and
// courant_factor as in k_godunov_oct (cpart: one minimum per workgroup).
template <int RIEMANN, bool MIXED, bool EXACT>
__global__ __launch_bounds__(192) __attribute__((amdgpu_waves_per_eu(MIXED ? 3 : 2, 8)))
void k_godunov_dims(MeshDev m, LevelDev lv, GasParams gp, const double* d_dt, double dx, double inv_dx,
                    const double* pre, double* cpart, unsigned long long* ckeys, double courant_factor) {
    using TC = typename std::conditional<MIXED, float, double>::type;
    // Contributions and keys as in k_godunov_oct (slots 2d, 2d+1 written by
    // wave d), and the own fluxes of dimensions 1 and 2 (minus face, then
    // plus face) for the first wave.
    __shared__ TC s_f[30][64];
    __shared__ long long s_k[MIXED ? 1 : 6][64];
    __shared__ TC s_own[2][10][64];
    const int t = threadIdx.x & 63;
    const int d = threadIdx.x >> 6;  // the same for every lane of a wave
    const int ic0 = t & 7;
    const int praw = xcd_block(blockIdx.x, gridDim.x) * 8 + (t >> 3);
    const int noct = lv.noct;
    // Lanes past the end of the list work on the last oct (valid memory) and
    // write nothing, so every lane of an oct group takes part in the shuffles.
    const bool valid = praw < noct;
    const int p = valid ? praw : noct - 1;
    const int lm = lv.leaf[p];
    const bool leaf = valid && ((lm >> ic0) & 1);

    const Gas<double> gd{gp.gamma, gp.gm1, gp.smalle, gp.smallp, 1.0 / gp.gm1};
    const Gas<TC> gc{(TC)gp.gamma, (TC)gp.gm1, (TC)gp.smalle, (TC)gp.smallp, (TC)(1.0 / gp.gm1)};
    const int slope_type = gp.slope_type, interpol_type = gp.interpol_type;
    double dt;
    step_of(lv, p, d_dt, dt, dx, inv_dx);
    const double dtdx = dt / dx;
    const TC dtc = (TC)dt, dxc = (TC)dx;
    const Len<TC, EXACT> len{dxc, (TC)inv_dx};
    const int cC = cell_id(m, lv.grid[p], ic0);

    double uC[5];
    TC qC[5];
    load_u(m.uold, m.ncell, cC, uC);
    face_prim(uC, qC, gd, gc);

    auto prim = [&](int icell, const TC fallback[5], TC q[5]) {
        if (icell > 0) {
            double u[5];
            load_u(m.uold, m.ncell, icell, u);
            face_prim(u, q, gd, gc);
        } else {
            copy5(fallback, q);
        }
    };

    // Across the oct face along d: the neighbour oct E (or a record), the
    // cell X beyond the face and, when X is a leaf, the cell X2 beyond it.
    const int b = 1 << d;
    const bool par = (ic0 >> d) & 1;  // 0: sibling on the plus side, X on the minus side
    const int icS = ic0 ^ b;
    const int E = lv.nbo[(size_t)(2 * d + (par ? 1 : 0)) * noct + p];
    int4 rc = make_int4(0, 0, 0, 0);
    int Xcell, X2cell = 0;
    if (E >= 0) {
        const int g = lv.grid[E];
        Xcell = cell_id(m, g, icS);
        if ((lv.leaf[E] >> icS) & 1) X2cell = cell_id(m, g, ic0);
    } else {
        rc = lv.rec[-1 - E];
        Xcell = rc.x;
    }
    double uX[5], uX2[5];
    if (Xcell > 0) load_u(m.uold, m.ncell, Xcell, uX);
    if (X2cell > 0) load_u(m.uold, m.ncell, X2cell, uX2);

    TC qS[5], qX[5];
    shfl5(qC, qS, b);
    if (Xcell > 0) face_prim(uX, qX, gd, gc);
    else copy5(qC, qX);
    TC qL1[5], qR1[5];
    sel5(par, qS, qX, qL1);
    sel5(par, qX, qS, qR1);
    TC OCm[5], OCp[5], GCm[5], GCp[5];
    own_trace(qL1, qC, qR1, d, slope_type, dtc, len, OCm, OCp, gc);
    ghost_trace(qL1, qC, qR1, d, interpol_type, dtc, len, GCm, GCp, gc);
    TC GSm[5], GSp[5];
    shfl5(GCm, GSm, b);
    shfl5(GCp, GSp, b);

    // Internal face: this cell's own flux on the sibling side.
    const bool sib_leaf = (lm >> icS) & 1;
    TC Fint[5];
    {
        TC ql[5], qr[5];
        sel5(par, GSm, OCm, ql);
        sel5(par, OCp, GSp, qr);
        solve_face<RIEMANN, MIXED>(ql, qr, Fint, gc, gd);
        rotate(Fint, d);
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) Fint[iv] = sib_leaf ? Fint[iv] : TC(0);
    }
    TC Fsib[5];
    shfl5(Fint, Fsib, b);  // the sibling's own flux on this cell's side

    // External face (across the oct boundary).
    TC Fext[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
    TC Fin[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
    long long Kin = -1;
    if (E >= 0) {
        if (X2cell > 0) {
            TC qX2[5];
            face_prim(uX2, qX2, gd, gc);
            TC a5[5], c5[5];
            sel5(par, qC, qX2, a5);
            sel5(par, qX2, qC, c5);
            TC GXm[5], GXp[5], OXm[5], OXp[5];
            ghost_trace(a5, qX, c5, d, interpol_type, dtc, len, GXm, GXp, gc);
            own_trace(a5, qX, c5, d, slope_type, dtc, len, OXm, OXp, gc);
            TC ql[5], qr[5];
            sel5(par, OCm, GXm, ql);
            sel5(par, GXp, OCp, qr);
            solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
            rotate(Fext, d);
            sel5(par, GCm, OXm, ql);
            sel5(par, OXp, GCp, qr);
            solve_face<RIEMANN, MIXED>(ql, qr, Fin, gc, gd);
            rotate(Fin, d);
            Kin = ((long long)E * 8 + icS) * 8 + (par ? 2 * d : 2 * d + 1);
        }
    } else if (rc.w & REC_REFINED_DEV) {
    } else if (rc.w & REC_BOUNDARY_DEV) {
        double u[5];
        copy5(uC, u);
#pragma unroll
        for (int i = 1; i <= 3; ++i)
            if (i == 1 + d) u[i] = u[i] * -1.0;
        TC qb[5];
        face_prim(u, qb, gd, gc);
        rotate(qb, d);
        TC dq0[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
        TC Bm[5], Bp[5];
        trace(qb, dq0, dtc, dxc, Bm, Bp, gc);
        TC ql[5], qr[5];
        sel5(par, OCm, Bm, ql);
        sel5(par, Bp, OCp, qr);
        solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
        rotate(Fext, d);
    } else {
        TC qnl[5], qnr[5];
        prim(rc.y, qX, qnl);
        prim(rc.z, qX, qnr);
        TC Gm[5], Gp[5];
        ghost_trace(qnl, qX, qnr, d, interpol_type, dtc, len, Gm, Gp, gc);
        TC ql[5], qr[5];
        sel5(par, OCm, Gm, ql);
        sel5(par, Gp, OCp, qr);
        solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
        rotate(Fext, d);
        if (leaf) {
            const int slot = lv.rec_slot[-1 - E] + face_slot(ic0, d);
            const double sg = par ? -1.0 : 1.0;
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) lv.slot_r[(size_t)slot * 5 + iv] = dtdx * sg * Fext[iv] * 0.125;
        }
    }

    // Own fluxes in face order (minus face, then plus face).
    TC Fm[5], Fp[5];
    sel5(par, Fint, Fext, Fm);
    sel5(par, Fext, Fint, Fp);
    if (d > 0) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            s_own[d - 1][iv][t] = Fm[iv];
            s_own[d - 1][5 + iv][t] = Fp[iv];
        }
    }
    // Incoming fluxes: the sibling's own flux on the shared face and X's.
    const int slotS = 2 * d + (par ? 0 : 1);
    const int slotX = 2 * d + (par ? 1 : 0);
    if constexpr (MIXED) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            s_f[5 * slotS + iv][t] = Fsib[iv];
            s_f[5 * slotX + iv][t] = Fin[iv];
        }
    } else {
        const double sgS = par ? -1.0 : 1.0;
        const double sgX = par ? 1.0 : -1.0;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            s_f[5 * slotS + iv][t] = dtdx * sgS * Fsib[iv] * 0.125;
            s_f[5 * slotX + iv][t] = dtdx * sgX * Fin[iv] * 0.125;
        }
        s_k[slotS][t] = sib_leaf ? ((long long)p * 8 + icS) * 8 + (par ? 2 * d + 1 : 2 * d) : -1;
        s_k[slotX][t] = Kin;
    }
    __syncthreads();
    if (d != 0) return;

    // The first wave: the own fluxes of the three dimensions in order, then
    // the ordered application of the six contributions.
    double fs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int iv = 0; iv < 5; ++iv) {
        fs[iv] = fs[iv] + 1.0 * (double)Fm[iv];
        fs[iv] = fs[iv] + -1.0 * (double)Fp[iv];
    }
#pragma unroll
    for (int dd = 0; dd < 2; ++dd)
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            fs[iv] = fs[iv] + 1.0 * (double)s_own[dd][iv][t];
            fs[iv] = fs[iv] + -1.0 * (double)s_own[dd][5 + iv][t];
        }
    double upre[5];
    load_u(pre, m.ncell, cC, upre);
    long long kin[6];
    if constexpr (MIXED) {
        incoming_keys(lv, p, ic0, lm, kin);
    } else {
#pragma unroll
        for (int f = 0; f < 6; ++f) kin[f] = s_k[f][t];
    }
    if (!cpart && !ckeys) {
        if (!leaf) return;
        ordered_apply_flux<64, !MIXED>(m, cC, (long long)p * 8 + ic0, fs, dtdx, gp, s_f, kin, t, upre, nullptr);
        return;
    }
    double cand = __builtin_inf();
    if (leaf) {
        double un[5];
        ordered_apply_flux<64, !MIXED>(m, cC, (long long)p * 8 + ic0, fs, dtdx, gp, s_f, kin, t, upre, un);
        cand = courant_cand_u(m, un, cC, dx, gp.gamma, courant_factor);
    }
    // One wave holds the workgroup's cells: its minimum is the block's.
    if (cpart) {
        cand = wave_min(cand);
        if (t == 0) cpart[blockIdx.x] = cand;
        return;
    }
    const int pfirst = praw - (t >> 3);
    const int b0 = lv.obox[pfirst], b1 = lv.obox[min(pfirst + 8, noct) - 1];
    if (b0 == b1) {
        cand = wave_min(cand);
        if (t == 0) atomicMin(&ckeys[b0], dt_key(cand));
    } else if (cand != __builtin_inf()) {
        atomicMin(&ckeys[lv.obox[p]], dt_key(cand));
    }
}

// ---------------------------------------------------------------------------
// Brick variant: OCTS consecutive octs of the list per workgroup, the cells
// of an oct in eight adjacent lanes as above.  On a uniform level, whose
// list is the Morton order of the octs, 64 / 32 / 16 consecutive octs are a
// 4x4x4 / 4x4x2 / 2x2x4 brick of octs, so most external face neighbours X
// of a cell are cells of the same workgroup.  Per dimension:
//  A. every lane computes its cell's own and ghost face states (2 traces),
//     sends the ghost state its sibling needs by shuffle and puts the one
//     its external neighbour needs in LDS (s_g); a leaf cell whose X is a
//     same-level leaf outside the workgroup registers a halo item;
//  B. the halo items, compacted into the first lanes of the workgroup,
//     compute what the neighbouring workgroup's lane would have: X's own and
//     ghost face states and X's flux on the shared face (the contribution to
//     the cell), and leave X's ghost state in the cell's s_g entry;
//  C. every lane solves its two own faces (2 Riemann problems) with the
//     neighbours' ghost states, shuffles the internal flux to its sibling
//     and writes its external flux, as a contribution, into X's LDS slot
//     when X is in the workgroup.
// Per cell and dimension that is 2 traces and 2 solves, plus the halo items'
// 2 traces and 1 solve for the brick's boundary cells (a quarter of the cells
// of a 4x4x4 brick), against 4 traces and 3 solves in k_godunov_oct.  The
// same operations on the same operands give the same values, so the results
// are identical to the other variants for any level order; only the share of
// halo items depends on it.  Contributions are kept per face slot in LDS and
// applied in processing order as in ordered_apply, with the keys recomputed
// from the tables at the end instead of stored.

// Position of the k-th set bit (0-based) of a 64-bit mask that has more
// than k set bits.
RH_DEV int nth_set_bit(unsigned long long mask, int k) {
    int pos = 0;
#pragma unroll
    for (int w = 32; w >= 1; w >>= 1) {
        const unsigned long long low = mask & ((1ull << w) - 1ull);
        const int c = __popcll(low);
        const bool up = k >= c;
        k = up ? k - c : k;
        mask = up ? (mask >> w) : low;
        pos += up ? w : 0;
    }
    return pos;
}

template <int OCTS> constexpr int brick_block() { return OCTS * 8; }

// Occupancy: two waves per SIMD in the FP64 form (at most 256 VGPRs), three
// in the mixed form (at most 168).  LDS per lane: six incoming face fluxes
// and one ghost state in the compute type (280 B in the FP64 form, 140 B in
// the mixed form), plus 16 B of halo cell ids: 148 KB for a 64-oct workgroup
// in the FP64 form, so one such workgroup fills a CU; 32- and 16-oct
// workgroups let two or four share one.
// `pre`, cpart, ckeys and courant_factor as in k_godunov_oct.
template <int RIEMANN, bool MIXED, bool EXACT, int OCTS>
__global__ __launch_bounds__(OCTS * 8) __attribute__((amdgpu_waves_per_eu(MIXED ? 3 : 2, 8)))
void k_godunov_brick(MeshDev m, LevelDev lv, GasParams gp, const double* d_dt, double dx, double inv_dx,
                     const double* pre, double* cpart, unsigned long long* ckeys, double courant_factor) {
    using TC = typename std::conditional<MIXED, float, double>::type;
    constexpr int NT = brick_block<OCTS>();
    constexpr int NW = NT / 64;
    __shared__ TC s_f[30][NT];  // the six incoming face fluxes of each cell
    __shared__ TC s_g[5][NT];
    // Halo registration of a dimension (ballot masks per wave, by dimension
    // parity) and, per registering lane, the cells its halo item reads: the
    // cell, X and X2.
    __shared__ unsigned long long s_bal[2][NW];
    __shared__ int4 s_hc[NT];
    const int t = threadIdx.x;
    const int ic0 = t & 7;
    const int p0 = xcd_block(blockIdx.x, gridDim.x) * OCTS;
    const int praw = p0 + (t >> 3);
    const int noct = lv.noct;
    const int pend = min(p0 + OCTS, noct);  // octs [p0, pend) are this workgroup's
    // Lanes past the end of the list work on the last oct (valid memory) and
    // write nothing, so every lane of an oct group takes part in the shuffles.
    const bool valid = praw < noct;
    const int p = valid ? praw : noct - 1;
    const int lm = lv.leaf[p];
    const bool leaf = valid && ((lm >> ic0) & 1);

    const Gas<double> gd{gp.gamma, gp.gm1, gp.smalle, gp.smallp, 1.0 / gp.gm1};
    const Gas<TC> gc{(TC)gp.gamma, (TC)gp.gm1, (TC)gp.smalle, (TC)gp.smallp, (TC)(1.0 / gp.gm1)};
    const int slope_type = gp.slope_type, interpol_type = gp.interpol_type;
    double dt;
    step_of(lv, p, d_dt, dt, dx, inv_dx);
    const double dtdx = dt / dx;
    const TC dtc = (TC)dt, dxc = (TC)dx;
    const Len<TC, EXACT> len{dxc, (TC)inv_dx};
    const int cC = cell_id(m, lv.grid[p], ic0);

    TC qC[5];
    {
        double u[5];
        load_u(m.uold, m.ncell, cC, u);
        face_prim(u, qC, gd, gc);
    }
    double fs[5] = {0.0, 0.0, 0.0, 0.0, 0.0};

    auto to_prim = [&](const double u[5], TC q[5]) {
        face_prim(u, q, gd, gc);
    };
    auto prim = [&](int icell, const TC fallback[5], TC q[5]) {
        if (icell > 0) {
            double u[5];
            load_u(m.uold, m.ncell, icell, u);
            to_prim(u, q);
        } else {
            copy5(fallback, q);
        }
    };

    // Across the oct face along dimension d, as in k_godunov_oct; the next
    // dimension's X and X2 are gathered while the current one is computed.
    struct Across {
        int E;
        int4 rc;
        int Xcell, X2cell;
    };
    auto across = [&](int d) {
        Across a;
        const bool pd = (ic0 >> d) & 1;
        const int icS = ic0 ^ (1 << d);
        a.E = lv.nbo[(size_t)(2 * d + (pd ? 1 : 0)) * noct + p];
        a.rc = make_int4(0, 0, 0, 0);
        a.X2cell = 0;
        if (a.E >= 0) {
            const int g = lv.grid[a.E];
            a.Xcell = cell_id(m, g, icS);
            if ((lv.leaf[a.E] >> icS) & 1) a.X2cell = cell_id(m, g, ic0);
        } else {
            a.rc = lv.rec[-1 - a.E];
            a.Xcell = a.rc.x;
        }
        return a;
    };
    Across nx = across(0);
    double uXn[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (nx.Xcell > 0) load_u(m.uold, m.ncell, nx.Xcell, uXn);

    // A leaf cell whose X along d is a same-level leaf outside the workgroup
    // registers a halo item for d.
    auto register_halo = [&](const Across& a, int d) {
        const bool halo = leaf && a.E >= 0 && a.X2cell > 0 && !(a.E >= p0 && a.E < pend);
        const unsigned long long bal = __ballot(halo);
        if ((t & 63) == 0) s_bal[d & 1][t >> 6] = bal;
        s_hc[t] = make_int4(cC, a.Xcell, a.X2cell, 0);
    };
    auto halo_count = [&](int d) {
        int n = 0;
#pragma unroll
        for (int w = 0; w < NW; ++w) n += __popcll(s_bal[d & 1][w]);
        return n;
    };
    // Lane of the h-th registered cell of dimension d (in lane order).
    auto halo_lane = [&](int d, int h) {
        int w = 0, k = h;
#pragma unroll
        for (int i = 0; i < NW - 1; ++i) {
            const int c = __popcll(s_bal[d & 1][w]);
            const bool on = k >= c;
            k = on ? k - c : k;
            w = on ? w + 1 : w;
        }
        return w * 64 + nth_set_bit(s_bal[d & 1][w], k);
    };
    // Lane t takes halo item t of dimension d: its cell's lane, and the
    // states of the cell, X and X2 gathered one dimension ahead of their
    // use (a lane without an item loads its own cell, a valid address).
    int h_tc = t;
    double hC[5], hX[5], hX2[5];
    auto halo_gather = [&](int d) {
        const bool has = t < halo_count(d);
        h_tc = has ? halo_lane(d, t) : t;
        const int4 ids = has ? s_hc[h_tc] : make_int4(cC, cC, cC, 0);
        load_u(m.uold, m.ncell, ids.x, hC);
        load_u(m.uold, m.ncell, ids.y, hX);
        load_u(m.uold, m.ncell, ids.z, hX2);
    };
    register_halo(nx, 0);
    __syncthreads();
    halo_gather(0);

    constexpr int DIM_UNROLL = (MIXED || RIEMANN != 2) ? 3 : 1;
#pragma unroll DIM_UNROLL
    for (int d = 0; d < 3; ++d) {
        const int b = 1 << d;
        const bool par = (ic0 >> d) & 1;  // 0: sibling on the plus side, X on the minus side
        const int icS = ic0 ^ b;
        const Across cur = nx;
        double uX[5];
        copy5(uXn, uX);
        if (d < 2) {
            nx = across(d + 1);
            if (nx.Xcell > 0) load_u(m.uold, m.ncell, nx.Xcell, uXn);
        } else {
            // Last dimension: the cell's own unew entry before this update
            // (no other lane writes it in this kernel).
            load_u(pre, m.ncell, cC, uXn);
        }
        const int E = cur.E;
        const bool xleaf = E >= 0 && cur.X2cell > 0;     // X is a same-level leaf
        const bool xin = E >= p0 && E < pend;            // X's oct is in this workgroup
        const int lane_x = (E - p0) * 8 + icS;           // X's lane when xin

        // A. Own and ghost face states of the cell.
        TC OCm[5], OCp[5], Gsib[5];
        {
            TC qS[5], qX[5];
            shfl5(qC, qS, b);
            if (cur.Xcell > 0) to_prim(uX, qX);
            else copy5(qC, qX);
            TC qL1[5], qR1[5];
            sel5(par, qS, qX, qL1);
            sel5(par, qX, qS, qR1);
            own_trace(qL1, qC, qR1, d, slope_type, dtc, len, OCm, OCp, gc);
            TC GCm[5], GCp[5];
            ghost_trace(qL1, qC, qR1, d, interpol_type, dtc, len, GCm, GCp, gc);
            // Toward the sibling: its left neighbour's qm / right neighbour's
            // qp; toward X the other one.
            TC Gown[5];
            sel5(par, GCp, GCm, Gown);
            shfl5(Gown, Gsib, b);
            TC Gext[5];
            sel5(par, GCm, GCp, Gext);
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) s_g[iv][t] = Gext[iv];
        }
        if (d < 2) register_halo(nx, d + 1);
        __syncthreads();

        // B. Halo items: lane h takes the h-th registered cell (in lane
        // order), whose states it gathered during the previous dimension.
        if (t < halo_count(d)) {
            const int tc = h_tc;
            const int pc = p0 + (tc >> 3), icc = tc & 7;
            // The cell's time step and width (its box's, in an ensemble).
            double dth, dxh = dx, ivh = inv_dx;
            step_of(lv, pc, d_dt, dth, dxh, ivh);
            const TC dtch = (TC)dth;
            const Len<TC, EXACT> lenh{(TC)dxh, (TC)ivh};
            const bool parc = (icc >> d) & 1;
            TC qc[5], qx[5], qx2[5];
            to_prim(hC, qc);
            to_prim(hX, qx);
            to_prim(hX2, qx2);
            TC a5[5], c5[5];
            sel5(parc, qc, qx2, a5);
            sel5(parc, qx2, qc, c5);
            TC GXm[5], GXp[5], OXm[5], OXp[5];
            ghost_trace(a5, qx, c5, d, interpol_type, dtch, lenh, GXm, GXp, gc);
            own_trace(a5, qx, c5, d, slope_type, dtch, lenh, OXm, OXp, gc);
            TC Gc[5];
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) Gc[iv] = s_g[iv][tc];
            // X's own flux on the shared face: its side 0 when X is on the
            // cell's plus side, else its side 1.
            TC Fin[5];
            {
                TC ql[5], qr[5];
                sel5(parc, Gc, OXm, ql);
                sel5(parc, OXp, Gc, qr);
                solve_face<RIEMANN, MIXED>(ql, qr, Fin, gc, gd);
                rotate(Fin, d);
            }
            const int slotX = 2 * d + (parc ? 1 : 0);
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) s_f[5 * slotX + iv][tc] = Fin[iv];
            // The cell's s_g entry now holds X's ghost state toward the cell
            // (no other lane reads it: X is outside the workgroup).
            TC GXc[5];
            sel5(parc, GXp, GXm, GXc);
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) s_g[iv][tc] = GXc[iv];
        }
        if (d < 2) halo_gather(d + 1);
        __syncthreads();

        // C. The cell's two own faces.
        const bool sib_leaf = (lm >> icS) & 1;
        TC Fint[5];
        {
            TC ql[5], qr[5];
            sel5(par, Gsib, OCm, ql);
            sel5(par, OCp, Gsib, qr);
            solve_face<RIEMANN, MIXED>(ql, qr, Fint, gc, gd);
            rotate(Fint, d);
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) Fint[iv] = sib_leaf ? Fint[iv] : TC(0);
        }
        TC Fsib[5];
        shfl5(Fint, Fsib, b);  // the sibling's own flux on this cell's side

        TC Fext[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
        const int4 rc = cur.rc;
        if (E >= 0) {
            if (xleaf) {
                const int src = xin ? lane_x : t;
                TC GX[5];
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) GX[iv] = s_g[iv][src];
                TC ql[5], qr[5];
                sel5(par, OCm, GX, ql);
                sel5(par, GX, OCp, qr);
                solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
                rotate(Fext, d);
                if (xin && leaf) {
                    // This cell's flux is X's incoming flux on this face.
                    const int slotXX = 2 * d + (par ? 0 : 1);
#pragma unroll
                    for (int iv = 0; iv < 5; ++iv) s_f[5 * slotXX + iv][lane_x] = Fext[iv];
                }
            }
        } else if (rc.w & REC_REFINED_DEV) {
        } else if (rc.w & REC_BOUNDARY_DEV) {
            double u[5];
            load_u(m.uold, m.ncell, cC, u);
#pragma unroll
            for (int i = 1; i <= 3; ++i)
                if (i == 1 + d) u[i] = u[i] * -1.0;
            TC qb[5];
            face_prim(u, qb, gd, gc);
            rotate(qb, d);
            TC dq0[5] = {TC(0), TC(0), TC(0), TC(0), TC(0)};
            TC Bm[5], Bp[5];
            trace(qb, dq0, dtc, dxc, Bm, Bp, gc);
            TC ql[5], qr[5];
            sel5(par, OCm, Bm, ql);
            sel5(par, Bp, OCp, qr);
            solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
            rotate(Fext, d);
        } else {
            TC qX[5], qnl[5], qnr[5];
            prim(rc.x, qC, qX);
            prim(rc.y, qX, qnl);
            prim(rc.z, qX, qnr);
            TC Gm[5], Gp[5];
            ghost_trace(qnl, qX, qnr, d, interpol_type, dtc, len, Gm, Gp, gc);
            TC ql[5], qr[5];
            sel5(par, OCm, Gm, ql);
            sel5(par, Gp, OCp, qr);
            solve_face<RIEMANN, MIXED>(ql, qr, Fext, gc, gd);
            rotate(Fext, d);
            if (leaf) {
                const int slot = lv.rec_slot[-1 - E] + face_slot(ic0, d);
                const double sg = par ? -1.0 : 1.0;
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) lv.slot_r[(size_t)slot * 5 + iv] = dtdx * sg * Fext[iv] * 0.125;
            }
        }

        // Own fluxes in face order (minus face, then plus face).
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            const double Fm = (double)(par ? Fint[iv] : Fext[iv]);
            const double Fp = (double)(par ? Fext[iv] : Fint[iv]);
            fs[iv] = fs[iv] + 1.0 * Fm;
            fs[iv] = fs[iv] + -1.0 * Fp;
        }
        // The sibling's own flux on the shared face.
        const int slotS = 2 * d + (par ? 0 : 1);
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) s_f[5 * slotS + iv][t] = Fsib[iv];
        // s_g and s_bal are rewritten by the next dimension; the last
        // dimension's barrier also publishes every contribution slot.
        __syncthreads();
    }

    long long kin[6];
    incoming_keys(lv, p, ic0, lm, kin);
    if (!cpart && !ckeys) {
        if (!leaf) return;
        ordered_apply_flux<NT>(m, cC, (long long)p * 8 + ic0, fs, dtdx, gp, s_f, kin, t, uXn, nullptr);
        return;
    }
    double cand = __builtin_inf();
    if (leaf) {
        double un[5];
        ordered_apply_flux<NT>(m, cC, (long long)p * 8 + ic0, fs, dtdx, gp, s_f, kin, t, uXn, un);
        cand = courant_cand_u(m, un, cC, dx, gp.gamma, courant_factor);
    }
    if (cpart) {
        block_min_n<NT>(cand, cpart);
        return;
    }
    const int b0 = lv.obox[p0], b1 = lv.obox[pend - 1];
    if (b0 == b1) {
        __shared__ double s_bmin[NW];
        cand = wave_min(cand);
        if ((t & 63) == 0) s_bmin[t >> 6] = cand;
        __syncthreads();
        if (t == 0) {
            double v = s_bmin[0];
            for (int w = 1; w < NW; ++w) v = min_skip_nan(v, s_bmin[w]);
            atomicMin(&ckeys[b0], dt_key(v));
        }
    } else if (cand != __builtin_inf()) {
        atomicMin(&ckeys[lv.obox[p]], dt_key(cand));
    }
}

// Reflux into coarser leaf cells, contributions in processing order.  Two
// forms with identical results (each target's contributions subtracted one
// after another in list order), chosen by the number of target cells:
//  - k_coarse_reflux_wave, one wave per target: the wave loads up to 64
//    contributions at once (slot index, then the five values) into LDS and
//    lane 0 subtracts them in order.  Latency-optimal for the few hundred to
//    few thousand targets of a small level, whose lists run to dozens.
//  - k_coarse_reflux, eight lanes per target (below): for levels with many
//    targets, whose lists are short.
__global__ __launch_bounds__(64) void k_coarse_reflux_wave(MeshDev m, LevelDev lv) {
    __shared__ double s_v[64][5];
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    if (t >= lv.ntgt) return;
    const size_t c = (size_t)lv.tgt_cell[t] - 1;
    double u[5];
    if (lane == 0) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) u[iv] = m.unew[(size_t)iv * m.ncell + c];
    }
    const int beg = lv.tgt_off[t], end = lv.tgt_off[t + 1];
    for (int k = beg; k < end; k += 64) {
        const int n = min(64, end - k);
        if (lane < n) {
            const size_t slot = (size_t)lv.tgt_slot[k + lane];
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) s_v[lane][iv] = lv.slot_r[slot * 5 + iv];
        }
        __syncthreads();
        if (lane == 0) {
            for (int j = 0; j < n; ++j)
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) u[iv] = u[iv] - s_v[j][iv];
        }
        __syncthreads();
    }
    if (lane == 0) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + c] = u[iv];
    }
}

// Reflux, eight lanes per target cell (32 targets per 256-thread block).  The eight lanes
// of a group load the target's next 32 contributions (slot index, then the
// five values; four per lane, all loads independent) at once, and the
// group's first lane subtracts them in order, taking each from its lane by
// shuffle.  A coarse cell next to a finer region can receive dozens of
// contributions: it costs one round trip per 32 of them instead of a chain of
// dependent loads, while the many targets with a few contributions each still
// fill the lanes.
constexpr int RF_LANES = 8;
constexpr int RF_PER_LANE = 4;
constexpr int RF_CHUNK = RF_LANES * RF_PER_LANE;
constexpr int RF_BLOCK = 256;
// Levels with fewer targets use the wave-per-target form.
constexpr int RF_GROUP_MIN_TARGETS = 4096;

__global__ __launch_bounds__(RF_BLOCK) void k_coarse_reflux(MeshDev m, LevelDev lv) {
    const long long t = ((long long)blockIdx.x * RF_BLOCK + threadIdx.x) / RF_LANES;
    const int j = threadIdx.x & (RF_LANES - 1);
    // The group's first lane within the wave; a group never straddles waves.
    const int lead = (threadIdx.x & 63) & ~(RF_LANES - 1);
    // Whole groups leave together, so every shuffle below reads active lanes.
    if (t >= lv.ntgt) return;
    const size_t c = (size_t)lv.tgt_cell[t] - 1;
    double u[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (j == 0) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) u[iv] = m.unew[(size_t)iv * m.ncell + c];
    }
    const int beg = lv.tgt_off[t], end = lv.tgt_off[t + 1];
    for (int k = beg; k < end; k += RF_CHUNK) {
        const int n = min(RF_CHUNK, end - k);
        // Contribution k + r*RF_LANES + j sits in lane j, register row r.
        double v[RF_PER_LANE][5];
#pragma unroll
        for (int r = 0; r < RF_PER_LANE; ++r) {
            const int e = r * RF_LANES + j;
            const size_t slot = e < n ? (size_t)lv.tgt_slot[k + e] : 0;
#pragma unroll
            for (int iv = 0; iv < 5; ++iv) v[r][iv] = e < n ? lv.slot_r[slot * 5 + iv] : 0.0;
        }
#pragma unroll
        for (int r = 0; r < RF_PER_LANE; ++r) {
            const int nr = min(RF_LANES, n - r * RF_LANES);  // group-uniform
            for (int i = 0; i < nr; ++i) {
#pragma unroll
                for (int iv = 0; iv < 5; ++iv) {
                    const double w = __shfl(v[r][iv], lead + i, 64);
                    u[iv] = u[iv] - w;
                }
            }
        }
    }
    if (j == 0) {
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) m.unew[(size_t)iv * m.ncell + c] = u[iv];
    }
}

// HydroSolver.cpp:734-876 for nvar = 5: prolongation of one coarse cell and
// its six neighbours (U1(i, iv), i = 0..6) into the eight children
// (U2(i, iv, value), i = 0..7).
template <class In, class Out>
RH_DEV void interpol_oct(In U1, Out U2, const GasParams& gp, int interpol_var) {
    const Gas<double> g{gp.gamma, gp.gm1, gp.smalle, gp.smallp, 1.0 / gp.gm1};
    const int it = gp.interpol_type;
    double q1[7][5];
    for (int i = 0; i < 7; ++i) {
        if (interpol_var == 0) {
            for (int iv = 0; iv < 5; ++iv) q1[i][iv] = U1(i, iv);
        } else if (interpol_var == 1) {
            double d = smax(U1(i, 0), 1e-10);
            q1[i][0] = U1(i, 0);
            double e_kin = 0.0;
            for (int k = 1; k <= 3; ++k) {
                q1[i][k] = U1(i, k);
                e_kin = e_kin + 0.5 * U1(i, k) * U1(i, k) / d;
            }
            double e_nonthermal = 0.0;
            q1[i][4] = U1(i, 4) - e_kin - e_nonthermal;
        } else {
            double uu[5];
            for (int iv = 0; iv < 5; ++iv) uu[iv] = U1(i, iv);
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
            for (int iv = 0; iv < 5; ++iv) U2(i, iv, q2[i][iv]);
    } else if (interpol_var == 1) {
        for (int i = 0; i < 8; ++i) {
            double d = smax(q2[i][0], 1e-10);
            U2(i, 0, d);
            double e_kin = 0.0;
            for (int k = 1; k <= 3; ++k) {
                U2(i, k, q2[i][k]);
                e_kin = e_kin + 0.5 * q2[i][k] * q2[i][k] / d;
            }
            double e_nonthermal = 0.0;
            U2(i, 4, q2[i][4] + e_kin + e_nonthermal);
        }
    } else {
        const double oneover_n2d = 1.0 / 8;
        for (int k = 1; k <= 3; ++k) {
            double mom_sum = 0.0;
            for (int i = 0; i < 8; ++i) mom_sum = mom_sum + q2[i][0] * q2[i][k] * oneover_n2d;
            double mom_err = mom_sum - U1(0, k);
            for (int i = 0; i < 8; ++i) {
                double d = smax(q2[i][0], 1e-10);
                q2[i][k] = (d * q2[i][k] - mom_err) / d;
            }
        }
        for (int i = 0; i < 8; ++i) {
            double d = smax(q2[i][0], 1e-10);
            double pp = smax(q2[i][4], 1e-20);
            U2(i, 0, d);
            double v2 = 0;
            for (int k = 1; k <= 3; ++k) {
                U2(i, k, d * q2[i][k]);
                v2 = v2 + q2[i][k] * q2[i][k];
            }
            double e_thermal = pp / g.gm1, e_kinetic = 0.5 * d * v2, e_nonthermal = 0.0;
            U2(i, 4, e_thermal + e_kinetic + e_nonthermal);
        }
    }
}

// interpol_hydro over independent [7][64] -> [8][64] blocks, one thread each.
__global__ __launch_bounds__(64) void k_interpol(int n_octs, const double* __restrict__ u1,
                                                 double* __restrict__ u2, GasParams gp, int interpol_var) {
    const int o = blockIdx.x * 64 + threadIdx.x;
    if (o >= n_octs) return;
    const double* U1 = u1 + (size_t)o * 7 * 64;
    double* U2 = u2 + (size_t)o * 8 * 64;
    interpol_oct([&](int i, int iv) { return U1[i * 64 + iv]; },
                 [&](int i, int iv, double v) { U2[i * 64 + iv] = v; }, gp, interpol_var);
}

// ---------------------------------------------------------------------------
// Adaptive-run kernels (the tree itself is edited on the host).

// New octs of make_grid_fine (TreeUpdater.cpp:115-147): for oct o, cells[8o]
// is the refined cell, cells[8o+1..8o+6] its six neighbours (<= 0: use the
// cell itself; no reflective boundaries exist), cells[8o+7] the new grid.
// The eight children get the prolongated state in uold and unew.
__global__ __launch_bounds__(64) void k_new_octs(MeshDev m, int n, const int* __restrict__ cells, GasParams gp,
                                                 int interpol_var, int* __restrict__ father, int* __restrict__ nbor) {
    const int o = blockIdx.x * 64 + threadIdx.x;
    if (o >= n) return;
    const int* c = cells + (size_t)o * 8;
    const int grid = c[7];
    // The new oct's links, as the host tree holds them: father = the refined
    // cell, nbor[f] = its six neighbour cells.
    father[grid - 1] = c[0];
#pragma unroll
    for (int f = 0; f < 6; ++f) nbor[(size_t)f * m.ngridmax + grid - 1] = c[1 + f];
    interpol_oct(
        [&](int i, int iv) {
            const int id = c[i] > 0 ? c[i] : c[0];
            return m.uold[(size_t)iv * m.ncell + (size_t)id - 1];
        },
        [&](int i, int iv, double v) {
            const size_t id = (size_t)cell_id(m, grid, i) - 1;
            m.uold[(size_t)iv * m.ncell + id] = v;
            m.unew[(size_t)iv * m.ncell + id] = v;
        },
        gp, interpol_var);
}

// Time-step bookkeeping of the adaptive run's subcycling (Simulation.cpp:
// 522-635) on the device, so that no fine step waits on the host for its
// Courant step; run by k_courant_final's first thread.  T as laid out by
// AMR_T_* in rh_kernels.h.
// op 0: after the level's Courant step (its minimum in T[AMR_T_DTC]):
//   dtold[il] = dtnew[il]; dtnew[il] = min(dtc, dtnew[il-1]/nsub_prev) for
//   il > lmin, else dtc (std::min's operand order); used[il] = dtnew[il].
// op 1: a level without finer octs: dtold/dtnew[il+1] = dtnew[il]/nsub, then
//   t += used[il].  op 2: the finest level: t += used[il].
RH_DEV void amr_times(double* T, int op, int il, int lmin, double nsub_prev, double nsub) {
    double* dtnew = T + AMR_T_DTNEW;
    double* dtold = T + AMR_T_DTOLD;
    double* used = T + AMR_T_USED;
    if (op == 0) {
        dtold[il] = dtnew[il];
        double v = T[AMR_T_DTC];
        if (il > lmin) {
            const double w = dtnew[il - 1] / nsub_prev;
            v = (w < v) ? w : v;
        }
        dtnew[il] = v;
        used[il] = v;
    } else {
        if (op == 1) {
            const double w = dtnew[il] / nsub;
            dtold[il + 1] = w;
            dtnew[il + 1] = w;
        }
        T[AMR_T_TIME] = T[AMR_T_TIME] + used[il];
    }
}

// restrict_fine (TreeUpdater.cpp:195-215): the father cell of every oct of
// the finer list gets the mean of its eight children's uold.
// Mean of the oct's eight children, summed in cell order, into its father cell.
RH_DEV void restrict_oct(const MeshDev& m, int g, const int* __restrict__ father) {
    const int fc = father[g - 1];
    if (fc <= 0) return;
    for (int iv = 0; iv < 5; ++iv) {
        double sum = 0;
        for (int ic = 0; ic < 8; ++ic) sum += m.uold[(size_t)iv * m.ncell + (size_t)cell_id(m, g, ic) - 1];
        m.uold[(size_t)iv * m.ncell + (size_t)fc - 1] = sum / 8;
    }
}

__global__ __launch_bounds__(256) void k_restrict(MeshDev m, LevelDev fine, const int* __restrict__ father) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= fine.noct) return;
    restrict_oct(m, fine.grid[p], father);
}

// set_uold of one level and the restriction of the next finer level into it
// (the two calls that end a level step of an adaptive run) in one launch:
// blocks [0, nb_set) copy the level's leaf cells as k_set_uold does, the
// remaining blocks average the finer octs into their father cells as
// k_restrict does.  The first part writes leaf cells of the level from unew,
// the second refined cells of the level from the finer level's uold, so the
// two touch disjoint cells and need no ordering between them.
__global__ __launch_bounds__(SBLOCK) void k_set_uold_restrict(MeshDev m, LevelDev lv, LevelDev fine,
                                                              const int* __restrict__ father, int nb_set) {
    if ((int)blockIdx.x < nb_set) {
        const int p = blockIdx.x * SBLOCK + threadIdx.x;
        if (p >= lv.noct) return;
        copy_oct(m, lv.grid[p], lv.leaf[p], m.unew, m.uold);
        return;
    }
    const int p = ((int)blockIdx.x - nb_set) * SBLOCK + threadIdx.x;
    if (p >= fine.noct) return;
    restrict_oct(m, fine.grid[p], father);
}

struct FlagState {
    double d, v[3], p;
};

// TreeUpdater.cpp:342-384 (hydro, nener = 0); zeros outside 1..ncell.
RH_DEV FlagState flag_state(const MeshDev& m, int id, double gamma) {
    FlagState s{0.0, {0.0, 0.0, 0.0}, 0.0};
    if (id <= 0 || id > m.ncell) return s;
    double u[5];
    load_u(m.uold, m.ncell, id, u);
    s.d = smax(u[0], 1e-10);
    double v2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        s.v[k] = u[1 + k] / s.d;
        v2 += s.v[k] * s.v[k];
    }
    double ek = 0.5 * s.d * v2;
    double em = 0.0, e_nonthermal = 0.0;
    s.p = smax((u[4] - ek - em - e_nonthermal) * (gamma - 1.0), s.d * 1e-10);
    return s;
}

RH_DEV double err_grad(double l, double c, double r, double floor) {
    return 2.0 * smax(sabs(r - c) / (r + c + floor), sabs(c - l) / (c + l + floor));
}

RH_DEV double err_grad_u(double vl, double vc, double vr, double pl, double pc, double pr, double dl, double dc,
                         double dr, double gamma) {
    const double floor = 1e-10;
    const double cg = ssqrt(smax(gamma * pl / dl, floor * floor));
    const double cm = ssqrt(smax(gamma * pc / dc, floor * floor));
    const double cd = ssqrt(smax(gamma * pr / dr, floor * floor));
    return 2.0 * smax(sabs(vr - vc) / (cd + cm + sabs(vr) + sabs(vc) + floor),
                      sabs(vc - vl) / (cm + cg + sabs(vc) + sabs(vl) + floor));
}

// Refinement flags of a level list on the device, one byte per oct in list
// order (bit ic-1: cell ic is flagged), so that flag_fine's passes over the
// list (TreeUpdater.cpp:328-582) run where the state and the level tables
// are and only the final flags go back to the host tree.

// Gradient criteria of flag_fine (TreeUpdater.cpp:482-526) for every cell of
// a level list, OR-ed into the oct's flag byte: a cell is flagged when a
// density, pressure or velocity error exceeds its threshold along some
// direction (the reference tests only unflagged cells; testing every cell and
// OR-ing gives the same flags).  The neighbour along a face is the same-level
// cell when its oct exists (get_nbor_cells_exact), else the grid's nbor
// entry; a neighbour outside 1..ncell is replaced by the cell itself.  The
// eight cells of an oct are eight adjacent lanes; the first writes the byte.
__global__ __launch_bounds__(256) void k_grad_bits(MeshDev m, LevelDev lv, const int* __restrict__ nbor, double ed,
                                                   double ep, double ev, double gamma, unsigned char* __restrict__ bits) {
    const int k = blockIdx.x * 256 + threadIdx.x;
    const bool valid = k < lv.noct * 8;
    const int p = valid ? k >> 3 : lv.noct - 1, ic0 = k & 7;
    bool ok = false;
    if (valid) {
        const int g = lv.grid[p];
        const int c = cell_id(m, g, ic0);
        const FlagState sc = flag_state(m, c, gamma);
        for (int d = 0; d < 3 && !ok; ++d) {
            int idn[2];
            for (int side = 0; side < 2; ++side) {
                const int par = (ic0 >> d) & 1;
                if (par != side) {
                    idn[side] = cell_id(m, g, ic0 ^ (1 << d));  // sibling inside the oct
                } else {
                    const int E = lv.nbo[(size_t)(2 * d + side) * lv.noct + p];
                    idn[side] = E >= 0 ? cell_id(m, lv.grid[E], ic0 ^ (1 << d))
                                       : nbor[(size_t)(2 * d + side) * m.ngridmax + g - 1];
                }
            }
            const FlagState sl = (idn[0] > 0 && idn[0] <= m.ncell) ? flag_state(m, idn[0], gamma) : sc;
            const FlagState sr = (idn[1] > 0 && idn[1] <= m.ncell) ? flag_state(m, idn[1], gamma) : sc;
            if (ed > 0.0 && err_grad(sl.d, sc.d, sr.d, 1e-10) > ed) ok = true;
            else if (ep > 0.0 && err_grad(sl.p, sc.p, sr.p, 1e-10) > ep) ok = true;
            else if (ev > 0.0 &&
                     err_grad_u(sl.v[d], sc.v[d], sr.v[d], sl.p, sc.p, sr.p, sl.d, sc.d, sr.d, gamma) > ev)
                ok = true;
        }
    }
    const unsigned long long b = __ballot(ok);
    const int lane = threadIdx.x & 63;
    if (valid && ic0 == 0) bits[p] = (unsigned char)(bits[p] | ((b >> (lane & 56)) & 0xFFull));
}

// One smooth_fine pass (TreeUpdater.cpp:217-268) over a level list: every
// unflagged cell with a flagged face neighbour at the same level (its
// sibling, or the cell of the neighbour oct on that side when that oct
// exists) becomes flagged, all decided from the flags before the pass
// (`in`), written to `out`.  One thread per oct.
__global__ __launch_bounds__(256) void k_smooth_bits(LevelDev lv, const unsigned char* __restrict__ in,
                                                     unsigned char* __restrict__ out) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= lv.noct) return;
    const int b = in[p];
    int nb[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
        const int E = lv.nbo[(size_t)f * lv.noct + p];
        nb[f] = E >= 0 ? in[E] : 0;
    }
    int r = b;
#pragma unroll
    for (int ic0 = 0; ic0 < 8; ++ic0) {
        bool any = false;
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            const int icS = ic0 ^ (1 << d);
            const int par = (ic0 >> d) & 1;
            // Sibling on the side opposite to the cell's parity, the
            // neighbour oct's cell icS on the cell's own side.
            any = any || ((b >> icS) & 1) || ((nb[2 * d + par] >> icS) & 1);
        }
        if (!((b >> ic0) & 1) && any) r |= 1 << ic0;
    }
    out[p] = (unsigned char)r;
}

// One level of a snapshot's hydro file (RamsesWriter.cpp:163-233), as the
// 32-bit words of its Fortran records: the il and ncache records, then (when
// n > 0) 40 records of n primitive values -- cell position ic outer,
// variable inner -- each framed by its byte count, then nbnd pairs of (il, 0)
// records for the boundary domains.  Thread i packs grid i of the level list
// (grids[i]); threads 0..39 also write the frames, thread 0 the small
// records.  A double sits at an odd word index, so it is stored as two words.
__global__ __launch_bounds__(256) void k_pack_hydro(MeshDev m, const int* __restrict__ grids, int n, int il,
                                                    int nbnd, double gamma, uint32_t* __restrict__ out) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const size_t rw = 2 + 2 * (size_t)n;  // words per value record
    const size_t body = n > 0 ? 40 * rw : 0;
    if (i == 0) {
        const uint32_t head[6] = {4u, (uint32_t)il, 4u, 4u, (uint32_t)n, 4u};
        for (int w = 0; w < 6; ++w) out[w] = head[w];
        for (int b = 0; b < nbnd; ++b) {
            const uint32_t tail[6] = {4u, (uint32_t)il, 4u, 4u, 0u, 4u};
            for (int w = 0; w < 6; ++w) out[6 + body + 6 * (size_t)b + w] = tail[w];
        }
    }
    if (n > 0 && i < 40) {
        out[6 + i * rw] = 8u * (uint32_t)n;
        out[6 + i * rw + rw - 1] = 8u * (uint32_t)n;
    }
    if (i >= n) return;
    const int g = grids[i];
    const double smallr = 1e-10;
    for (int ic0 = 0; ic0 < 8; ++ic0) {
        const size_t c = (size_t)cell_id(m, g, ic0) - 1;
        double u[5];
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) u[iv] = m.uold[(size_t)iv * m.ncell + c];
        const double d = u[0] < smallr ? smallr : u[0];
        double q[5];
        q[0] = d;
        double v2 = 0.0;
#pragma unroll
        for (int j = 1; j <= 3; ++j) {
            q[j] = u[j] / d;
            v2 += q[j] * q[j];
        }
        const double p = (u[4] - 0.5 * d * v2 - 0.0) * (gamma - 1.0);
        q[4] = p < 0.0 ? 0.0 : p;
#pragma unroll
        for (int iv = 0; iv < 5; ++iv) {
            const unsigned long long bits = (unsigned long long)__double_as_longlong(q[iv]);
            const size_t w = 6 + (size_t)(ic0 * 5 + iv) * rw + 1 + 2 * (size_t)i;
            out[w] = (uint32_t)bits;
            out[w + 1] = (uint32_t)(bits >> 32);
        }
    }
}

inline int blocks_for(int noct) { return (noct + OCTS_PER_BLOCK - 1) / OCTS_PER_BLOCK; }

}  // namespace

int courant_blocks(const LevelDev& lv) { return stream_blocks(lv.noct); }
static_assert(NOALLOC_COURANT_MAX_OCTS == (long long)NOALLOC_COURANT_PARTS * SBLOCK, "octs per Courant block");

hipError_t launch_set_unew(const MeshDev& m, const LevelDev& lv, hipStream_t stream) {
    if (lv.noct == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_unew, dim3(stream_blocks(lv.noct)), dim3(SBLOCK), 0, stream, m, lv);
    return hipGetLastError();
}

hipError_t launch_set_unew_oct_lanes(const MeshDev& m, const LevelDev& lv, hipStream_t stream) {
    if (lv.noct == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_unew_oct_lanes, dim3((lv.noct + 7) / 8), dim3(64), 0, stream, m, lv);
    return hipGetLastError();
}

hipError_t launch_set_uold(const MeshDev& m, const LevelDev& lv, hipStream_t stream) {
    if (lv.noct == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_uold, dim3(stream_blocks(lv.noct)), dim3(SBLOCK), 0, stream, m, lv);
    return hipGetLastError();
}

hipError_t launch_gravity_source(const MeshDev& m, const LevelDev& lv, double dt, hipStream_t stream) {
    if (lv.noct == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gravity_source, dim3(blocks_for(lv.noct)), dim3(BLOCK), 0, stream, m, lv, dt);
    return hipGetLastError();
}

hipError_t launch_synchro(const MeshDev& m, const LevelDev& lv, double dt, double gamma, hipStream_t stream) {
    if (lv.noct == 0) return hipSuccess;
    hipLaunchKernelGGL(k_synchro, dim3(blocks_for(lv.noct)), dim3(BLOCK), 0, stream, m, lv, dt, gamma);
    return hipGetLastError();
}

hipError_t launch_step_join(const MeshDev& m, const LevelDev& lv, double dx, double gamma, double courant_factor,
                            double dt_init, double* d_part, int npart, double* d_dt, hipStream_t stream) {
    int nb = stream_blocks(lv.noct);
    if (nb > npart) return hipErrorInvalidValue;
    if (nb > 0)
        hipLaunchKernelGGL(k_step_join, dim3(nb), dim3(SBLOCK), 0, stream, m, lv, dx, gamma, courant_factor, d_part);
    hipLaunchKernelGGL(k_courant_final, dim3(1), dim3(1024), 0, stream, (const double*)d_part, nb, dt_init, d_dt,
                       AmrTimesArgs{});
    return hipGetLastError();
}

int godunov_blocks(const LevelDev& lv, int variant) {
    const int octs = variant == 1   ? 16
                     : variant == 2 ? 8
                     : variant == 3 ? 16
                     : variant == 4 ? 32
                     : variant == 5 ? 64
                     : variant == 6 ? 8
                                    : G_OCTS;
    return (lv.noct + octs - 1) / octs;
}

hipError_t launch_courant_final(const double* d_part, int nparts, double dt_init, double* d_dt, hipStream_t stream) {
    hipLaunchKernelGGL(k_courant_final, dim3(1), dim3(1024), 0, stream, d_part, nparts, dt_init, d_dt,
                       AmrTimesArgs{});
    return hipGetLastError();
}

hipError_t launch_courant(const MeshDev& m, const LevelDev& lv, double dx, double gamma, double courant_factor,
                          double dt_init, double* d_part, int npart, double* d_dt, hipStream_t stream,
                          bool set_unew, const AmrTimesArgs& at) {
    if (at.T && (at.il < 1 || at.il + 1 >= AMR_T_LEVELS || at.then_op < 0 || at.then_op > 2))
        return hipErrorInvalidValue;
    int nb = stream_blocks(lv.noct);
    if (nb > npart) return hipErrorInvalidValue;
    if (nb > 0 && set_unew)
        hipLaunchKernelGGL(k_courant_partial<true>, dim3(nb), dim3(SBLOCK), 0, stream, m, lv, dx, gamma,
                           courant_factor, d_part);
    else if (nb > 0)
        hipLaunchKernelGGL(k_courant_partial<false>, dim3(nb), dim3(SBLOCK), 0, stream, m, lv, dx, gamma,
                           courant_factor, d_part);
    hipLaunchKernelGGL(k_courant_final, dim3(1), dim3(1024), 0, stream, (const double*)d_part, nb, dt_init, d_dt,
                       at);
    return hipGetLastError();
}

hipError_t launch_courant_boxes(const MeshDev& m, const LevelDev& lv, double gamma, double courant_factor,
                                unsigned long long* keys, hipStream_t stream, bool join) {
    if (lv.noct == 0) return hipSuccess;
    if (!lv.obox || !lv.box_len) return hipErrorInvalidValue;
    if (join)
        hipLaunchKernelGGL(k_courant_boxes<true>, dim3(stream_blocks(lv.noct)), dim3(SBLOCK), 0, stream, m, lv, gamma,
                           courant_factor, keys);
    else
        hipLaunchKernelGGL(k_courant_boxes<false>, dim3(stream_blocks(lv.noct)), dim3(SBLOCK), 0, stream, m, lv, gamma,
                           courant_factor, keys);
    return hipGetLastError();
}

hipError_t launch_fill_keys(unsigned long long* keys, int n, double value, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fill_keys, dim3((n + 255) / 256), dim3(256), 0, stream, keys, n, value);
    return hipGetLastError();
}

hipError_t launch_keys_to_dt(const unsigned long long* keys, int n, double* dt, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_keys_to_dt, dim3((n + 255) / 256), dim3(256), 0, stream, keys, n, dt);
    return hipGetLastError();
}

namespace {

template <int R, bool MIXED, bool EXACT>
void launch_form(int variant, const MeshDev& m, const LevelDev& lv, const GasParams& gp, const double* d_dt,
                 double dx, double inv_dx, const FusedCourant& fc, hipStream_t stream) {
    const double* pre = fc.pre_from_uold ? m.uold : m.unew;
    if (variant == 6) {
        hipLaunchKernelGGL((k_godunov_dims<R, MIXED, EXACT>), dim3((lv.noct + 7) / 8), dim3(192), 0, stream, m, lv,
                           gp, d_dt, dx, inv_dx, pre, fc.part, fc.keys, fc.courant_factor);
    } else if (variant >= 3) {
        const int octs = variant == 3 ? 16 : variant == 4 ? 32 : 64;
        const dim3 grid((lv.noct + octs - 1) / octs);
        if (variant == 3)
            hipLaunchKernelGGL((k_godunov_brick<R, MIXED, EXACT, 16>), grid, dim3(brick_block<16>()), 0, stream, m,
                               lv, gp, d_dt, dx, inv_dx, pre, fc.part, fc.keys, fc.courant_factor);
        else if (variant == 4)
            hipLaunchKernelGGL((k_godunov_brick<R, MIXED, EXACT, 32>), grid, dim3(brick_block<32>()), 0, stream, m,
                               lv, gp, d_dt, dx, inv_dx, pre, fc.part, fc.keys, fc.courant_factor);
        else
            hipLaunchKernelGGL((k_godunov_brick<R, MIXED, EXACT, 64>), grid, dim3(brick_block<64>()), 0, stream, m,
                               lv, gp, d_dt, dx, inv_dx, pre, fc.part, fc.keys, fc.courant_factor);
    } else if (variant == 1)
        hipLaunchKernelGGL((k_godunov_oct<R, MIXED, EXACT, 16>), dim3((lv.noct + 15) / 16), dim3(oct_block<16>()), 0,
                           stream, m, lv, gp, d_dt, dx, inv_dx, pre, fc.part, fc.keys, fc.courant_factor);
    else if (variant == 2)
        hipLaunchKernelGGL((k_godunov_oct<R, MIXED, EXACT, 8>), dim3((lv.noct + 7) / 8), dim3(oct_block<8>()), 0,
                           stream, m, lv, gp, d_dt, dx, inv_dx, pre, fc.part, fc.keys, fc.courant_factor);
    else
        hipLaunchKernelGGL((k_godunov<R, MIXED, EXACT>), dim3((lv.noct + G_OCTS - 1) / G_OCTS), dim3(G_BLOCK), 0,
                           stream, m, lv, gp, d_dt, dx, inv_dx);
}

// The FP32 half of the mixed form divides through dv() whatever dx is, so
// only the FP64 form has an exact-reciprocal instantiation.
template <int R>
void launch_riemann(int variant, bool mixed, bool exact, const MeshDev& m, const LevelDev& lv, const GasParams& gp,
                    const double* d_dt, double dx, double inv_dx, const FusedCourant& fc, hipStream_t stream) {
    if (mixed) launch_form<R, true, false>(variant, m, lv, gp, d_dt, dx, inv_dx, fc, stream);
    else if (exact) launch_form<R, false, true>(variant, m, lv, gp, d_dt, dx, inv_dx, fc, stream);
    else launch_form<R, false, false>(variant, m, lv, gp, d_dt, dx, inv_dx, fc, stream);
}

}  // namespace

hipError_t launch_godunov(const MeshDev& m, const LevelDev& lv, const GasParams& gp, int riemann, int precision,
                          int variant, const double* d_dt, double dx, double inv_dx, hipStream_t stream,
                          const FusedCourant& fc) {
    if (lv.noct == 0) return hipSuccess;
    if (precision != 0 && precision != 1) return hipErrorNotSupported;
    if (variant < 0 || variant >= KERNEL_VARIANTS) return hipErrorInvalidValue;
    // The fused Courant epilogue needs the oct-in-lanes or brick kernel and a
    // level with no coarser neighbours (no reflux targets, no records); the
    // pre-update state from uold needs one of those kernels.
    if ((fc.part || fc.keys) && (variant == 0 || lv.ntgt != 0 || lv.nrec != 0 || lv.nleaf != 8LL * lv.noct))
        return hipErrorInvalidValue;
    if (fc.pre_from_uold && variant == 0) return hipErrorInvalidValue;
    if ((fc.part && (lv.obox || fc.keys)) || (fc.keys && !lv.obox)) return hipErrorInvalidValue;
    const bool mixed = precision == 1;
    const bool exact = inv_dx != 0.0;
    switch (riemann) {
        case 0: launch_riemann<0>(variant, mixed, exact, m, lv, gp, d_dt, dx, inv_dx, fc, stream); break;
        case 1: launch_riemann<1>(variant, mixed, exact, m, lv, gp, d_dt, dx, inv_dx, fc, stream); break;
        case 2: launch_riemann<2>(variant, mixed, exact, m, lv, gp, d_dt, dx, inv_dx, fc, stream); break;
        case 4: launch_riemann<4>(variant, mixed, exact, m, lv, gp, d_dt, dx, inv_dx, fc, stream); break;
        default: return hipErrorNotSupported;
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || lv.ntgt == 0) return e;
    if (lv.ntgt < RF_GROUP_MIN_TARGETS) {
        hipLaunchKernelGGL(k_coarse_reflux_wave, dim3(lv.ntgt), dim3(64), 0, stream, m, lv);
    } else {
        const size_t rf_threads = (size_t)lv.ntgt * RF_LANES;
        hipLaunchKernelGGL(k_coarse_reflux, dim3((unsigned)((rf_threads + RF_BLOCK - 1) / RF_BLOCK)),
                           dim3(RF_BLOCK), 0, stream, m, lv);
    }
    return hipGetLastError();
}

hipError_t launch_new_octs(const MeshDev& m, int n, const int* d_cells, const GasParams& gp, int interpol_var,
                           int* d_father, int* d_nbor, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_new_octs, dim3((n + 63) / 64), dim3(64), 0, stream, m, n, d_cells, gp, interpol_var,
                       d_father, d_nbor);
    return hipGetLastError();
}

hipError_t launch_restrict(const MeshDev& m, const LevelDev& fine, const int* d_father, hipStream_t stream) {
    if (fine.noct == 0) return hipSuccess;
    hipLaunchKernelGGL(k_restrict, dim3((fine.noct + 255) / 256), dim3(256), 0, stream, m, fine, d_father);
    return hipGetLastError();
}

hipError_t launch_set_uold_restrict(const MeshDev& m, const LevelDev& lv, const LevelDev& fine, const int* d_father,
                                    hipStream_t stream) {
    const int nb_set = stream_blocks(lv.noct), nb_fine = stream_blocks(fine.noct);
    if (nb_set + nb_fine == 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_uold_restrict, dim3(nb_set + nb_fine), dim3(SBLOCK), 0, stream, m, lv, fine, d_father,
                       nb_set);
    return hipGetLastError();
}

hipError_t launch_flag_bits(const MeshDev& m, const LevelDev& lv, const int* d_nbor, double ed, double ep,
                            double ev, double gamma, int nexpand, unsigned char* d_a, unsigned char* d_b,
                            unsigned char** d_result, hipStream_t stream) {
    *d_result = d_a;
    if (lv.noct == 0) return hipSuccess;
    const dim3 go((lv.noct + 255) / 256), gc((lv.noct * 8 + 255) / 256), blk(256);
    unsigned char* cur = d_a;
    unsigned char* oth = d_b;
    hipLaunchKernelGGL(k_smooth_bits, go, blk, 0, stream, lv, (const unsigned char*)cur, oth);
    std::swap(cur, oth);
    if (ed > 0.0 || ep > 0.0 || ev > 0.0)
        hipLaunchKernelGGL(k_grad_bits, gc, blk, 0, stream, m, lv, d_nbor, ed, ep, ev, gamma, cur);
    for (int i = 0; i < nexpand; ++i) {
        hipLaunchKernelGGL(k_smooth_bits, go, blk, 0, stream, lv, (const unsigned char*)cur, oth);
        std::swap(cur, oth);
    }
    *d_result = cur;
    return hipGetLastError();
}

hipError_t launch_interpol(int n_octs, const double* d_u1, double* d_u2, const GasParams& gp, int interpol_var,
                           hipStream_t stream) {
    if (n_octs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_interpol, dim3((n_octs + 63) / 64), dim3(64), 0, stream, n_octs, d_u1, d_u2, gp,
                       interpol_var);
    return hipGetLastError();
}

size_t pack_hydro_words(int n, int nbnd) { return 6 + (n > 0 ? 40 * (2 + 2 * (size_t)n) : 0) + 6 * (size_t)nbnd; }

hipError_t launch_pack_hydro(const MeshDev& m, const int* d_grids, int n, int il, int nbnd, double gamma,
                             uint32_t* d_out, hipStream_t stream) {
    const int threads = n > 40 ? n : 40;
    hipLaunchKernelGGL(k_pack_hydro, dim3((threads + 255) / 256), dim3(256), 0, stream, m, d_grids, n, il, nbnd,
                       gamma, d_out);
    return hipGetLastError();
}

}  // namespace rh

// ---------------------------------------------------------------------------
// Scratch of the allocation-free Courant step (rh_courant_noalloc): block
// minima, then the time step.  Every call on one GPU shares this array, so
// calls that overlap on one GPU race.  Nothing in it has to start at zero:
// k_courant_partial writes every block minimum that k_courant_final reads,
// and k_courant_final writes the time step.
static __device__ __align__(256) double scratch[rh::NOALLOC_SCRATCH_DOUBLES];

namespace rh {

namespace {
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};
}  // namespace

hipError_t noalloc_scratch(int device, double** ptr) {
    if (!ptr || device < 0 || device >= SCRATCH_MAX_DEVICES) return hipErrorInvalidValue;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        // The array's name is a host-side stand-in; its address on this
        // device is what kernels and copies must be given.
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        g_scratch_ptr[device] = (double*)p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

}  // namespace rh
