/*
 * rhydro_driver.h -- host-side driver helpers of the standalone build.
 *
 * These have no counterpart among the reference's hydro virtuals: they stand
 * in for the parts of the reference's Simulation / TreeUpdater / Initializer
 * that a run needs around the hot path (uniform mesh construction, region
 * and seeded initial conditions) and expose the per-kernel timers used by
 * the benchmark.  A drop-in user keeps the reference's own mesh and
 * initialisation and needs none of this.
 */
#ifndef RHYDRO_DRIVER_H
#define RHYDRO_DRIVER_H

#include "rhydro.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Number of grids a uniform mesh of nx*ny*nz coarse cells refined down to
 * level list `levelmin` needs (the reference's numbering: level list L holds
 * ncoarse*8^(L-2) octs). */
long long rh_uniform_ngrid(int nx, int ny, int nz, int levelmin);

/* Builds the uniform mesh the reference builds for levelmin == levelmax
 * (Simulation.cpp:226-234 with TreeUpdater::make_grid_fine): grids numbered
 * 1,2,3,... breadth first, children in cell order, level list L+1 filled from
 * level list L in list order.  Arrays are caller-allocated, sized as in
 * AmrGrid::allocate (AmrGrid.cpp:13-70) for the given ngridmax, and fully
 * written here: father[ngridmax], son[ncell], nbor[6*ngridmax],
 * next/prev[ngridmax] (level lists, then the free list from the first unused
 * grid), headl/taill/numbl[nlevelmax+1] (ncpu = 1), xg[3*ngridmax].
 * Returns the number of grids used or a negative rh_status. */
long long rh_build_uniform(int nx, int ny, int nz, int levelmin, int nlevelmax, int ngridmax, int* father,
                           int* son, int* nbor, int* next, int* prev, int* headl, int* taill, int* numbl,
                           double* xg);

/* Region initial conditions (Initializer.cpp:31-157, types 'square'/'cube'
 * and 'sphere'), applied to every cell of every level list and the coarse
 * cells, then restricted upward as Simulation::initialize's final
 * legacy_init_flow does (Simulation.cpp:218-224,257; TreeUpdater.cpp:195-215).
 * Region arrays have nreg entries; type[i] is 0 = square/cube, 1 = sphere.
 * unew is set equal to uold. */
int rh_init_regions(const rh_mesh* mesh, const double* xg, double boxlen, double gamma, int nreg, const int* type,
                    const double* x_center, const double* y_center, const double* z_center,
                    const double* length_x, const double* length_y, const double* length_z, const double* d_region,
                    const double* u_region, const double* v_region, const double* w_region,
                    const double* p_region);

/* Shock-tube initial condition with an arbitrary plane: cells whose centre x
 * has (x - point).normal < 0 get the primitive state left[5] = (d, u, v, w,
 * p), the others right[5]; cell centres, the upward restriction and
 * unew = uold as in rh_init_regions.  For the ensemble benchmark's Sod boxes
 * (no counterpart in the reference's region types). */
int rh_init_shock_tube(const rh_mesh* mesh, const double* xg, double boxlen, double gamma, const double* normal,
                       const double* point, const double* left, const double* right);

/* Seeded random state on the leaf cells of level list `level`, in list order
 * and cell order: rho = exp(s*Z1), p = exp(s*Z2), v_i = a*c_s*Z with
 * c_s = sqrt(gamma*p/rho), E = p/(gamma-1) + rho*v^2/2, where the Z are
 * standard normals from std::mt19937_64(seed) via Box-Muller (pairs of
 * draws u1,u2 in (0,1]: sqrt(-2 ln u1) * cos(2 pi u2), then sin).  Every
 * other cell of the mesh gets rho=1, v=0, p=1.  unew = uold. */
int rh_init_lognormal(const rh_mesh* mesh, int level, unsigned long long seed, double gamma, double sigma,
                      double vel_amp);

/* Per-kernel device time accumulated on the context's stream (HIP events
 * around each launch) since the last reset: kernel 0 = godunov,
 * 1 = courant, 2 = set_unew, 3 = set_uold, 4 = the fused pass between two
 * steps of rh_run_uniform (set_uold, courant and set_unew in one kernel).
 * ms_total/count per kernel. */
int rh_timers_enable(rh_ctx* ctx, int enable);
int rh_timers_read(rh_ctx* ctx, int kernel, double* ms_total, long long* count);
int rh_timers_reset(rh_ctx* ctx);

/* Godunov kernel variant of a context: 0 = one thread per cell with the
 * octs of a workgroup strided across lanes; 1 = the eight cells of an oct in
 * adjacent lanes, sharing sibling states and fluxes by cross-lane shuffles,
 * 16 octs per workgroup; 2 = the same with 8 octs per workgroup;
 * 3, 4, 5 = bricks of 16, 32, 64 consecutive octs per workgroup that also
 * share face states and fluxes between neighbouring octs through LDS (a
 * 2x2x4, 4x4x2, 4x4x4 brick of octs on a uniform level); 6 = 8 octs per
 * workgroup of three waves, each computing one dimension of the update for
 * the same cells.  All produce identical results.  -1 (the default of a new
 * context) chooses per level: 6 for the FP64 form, and for the mixed form 6
 * on levels of fewer than 16,384 octs and 2 on larger ones (the fastest
 * measured at 256^3 cells and on small levels).  rh_autotune and
 * rh_set_kernel_variant fix one variant for every level of the context;
 * rh_get_kernel_variant reports -1 while the default is in use. */
#define RH_KERNEL_VARIANTS 7
int rh_set_kernel_variant(rh_ctx* ctx, int variant);
int rh_get_kernel_variant(rh_ctx* ctx, int* variant);

/* Times each Godunov kernel variant on level `ilevel` of the bound mesh
 * (eight warm-up launches, then per variant four untimed launches and `reps`
 * timed ones back to back, HIP events on the context's stream around each;
 * the launches read the pre-update state from uold and, on a uniform
 * periodic level, end with the next step's Courant candidates, as the
 * multi-step calls do), keeps the fastest for the context and reports it in
 * *best_variant and the average ms per launch in
 * ms_per_variant[RH_KERNEL_VARIANTS] (either pointer may be NULL).  unew is
 * restored after every trial. */
int rh_autotune(rh_ctx* ctx, int ilevel, double dt, double dx, int reps, int* best_variant, double* ms_per_variant);

/* Multi-step calls whose steps run with uold and unew exchanged (levels of
 * leaf cells with same-level neighbours only: rh_run_uniform, and
 * rh_run_ensemble on periodic boxes) replay their launch sequence from a HIP
 * graph captured on first use and kept while the level tables, state
 * buffers, parameters and step count stay the same (on by default; eager
 * launches while the timers are on).  rh_graph_stats reports how many graphs
 * the context captured and launched. */
int rh_set_graphs(rh_ctx* ctx, int enable);
int rh_graph_stats(rh_ctx* ctx, long long* captured, long long* launched);

/* rh_set_unew's copy (uold -> unew on every cell of the level's octs) done
 * with the Godunov kernel's memory access pattern -- the eight cells of an
 * oct in eight adjacent lanes, eight octs per 64-lane workgroup -- so that
 * the byte counters of a copy of known size calibrate the Godunov kernel's.
 * Same result as rh_set_unew. */
int rh_set_unew_oct_lanes(rh_ctx* ctx, int ilevel);

/* Leaf cells of a level in the bound mesh (cell-updates per godunov call). */
int rh_level_leaf_cells(rh_ctx* ctx, int ilevel, long long* nleaf);

#ifdef __cplusplus
}
#endif

#endif /* RHYDRO_DRIVER_H */
