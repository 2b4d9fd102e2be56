/*
 * rhydro_amr.h -- adaptive runs around the GPU hydro path.
 *
 * Counterpart of the reference's adaptive time integration: the tree updater
 * (TreeUpdater.cpp:27-636: flag_fine, smooth_fine, ensure_ref_rules,
 * make_grid_fine, remove_grid_fine, restrict_fine), the level and free lists
 * of AmrGrid (AmrGrid.cpp:13-70,259-302) and Simulation::initialize's
 * refinement loops, run's coarse step and the subcycled amr_step
 * (Simulation.cpp:160-255,355-395,522-635), for ncpu = 1.  A drop-in user
 * keeps the reference's own driver and only swaps the hydro solver
 * (rhydro.h); this driver is for runs that keep the whole state on the GPU.
 *
 * The state stays in HBM for the whole run.  The hydro steps, the gradient
 * refinement criteria, the prolongation into new octs and the restriction
 * run as HIP kernels; the host edits only the tree (integer arrays) and
 * downloads one byte per cell of a list to combine the refinement flags.
 */
#ifndef RHYDRO_AMR_H
#define RHYDRO_AMR_H

#include "rhydro.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rh_amr rh_amr;

/* Namelist of an adaptive run (the keys the reference reads, with its
 * defaults: nexpand 1, courant_factor 0.8, err_grad_* -1 = off). */
typedef struct {
    int nx, ny, nz;             /* amr_params nx, ny, nz */
    int ngridmax;               /* grids available to the run (no resizing) */
    int levelmin, levelmax;     /* amr_params levelmin, levelmax */
    double courant_factor;      /* hydro_params courant_factor */
    double err_grad_d, err_grad_p, err_grad_v;  /* refine_params */
    int nexpand;                /* amr_params nexpand */
    /* Initial condition: kind 0 = regions (Initializer.cpp:31-157; type 0 =
     * square/cube, 1 = sphere; nreg entries per array); kind 1 =
     * Kelvin-Helmholtz slab: rho = 2 for |y - 0.5| < 0.25 else 1,
     * u = +0.5 inside / -0.5 outside, v = kh_amp*sin(4 pi x), p = 2.5. */
    int ic_kind;
    int nreg;
    const int* reg_type;
    const double *x_center, *y_center, *z_center, *length_x, *length_y, *length_z;
    const double *d_region, *u_region, *v_region, *w_region, *p_region;
    double kh_amp;
} rh_amr_params;

/* Creates the run and builds its initial mesh and state as
 * Simulation::initialize does (region ICs, refinement to levelmax).
 * hydro->boxlen is the box size. */
int rh_amr_create(const rh_params* hydro, const rh_amr_params* amr, rh_amr** out);
int rh_amr_destroy(rh_amr* run);

/* nsteps coarse steps of Simulation::run (Simulation.cpp:355-395). */
int rh_amr_step(rh_amr* run, int nsteps);

/* Copies uold/unew to the host arrays and describes them (pointers stay
 * valid until the next rh_amr call).  headl/numbl have nlevelmax+1 entries. */
int rh_amr_mesh(rh_amr* run, rh_mesh* mesh, const int** numbl, const double** xg);

/* Time, fine-step count, and leaf cells updated by godunov_fine since
 * creation (the cell-update count of the benchmark metric).  The run keeps
 * its time and time steps in device memory: this call reads them back and so
 * waits for the run's queued work. */
int rh_amr_stats(rh_amr* run, double* t, int* nstep, long long* cell_updates);

/* The run's hydro context (kernel timers, variant), owned by the run. */
rh_ctx* rh_amr_context(rh_amr* run);

/* Host wall time (ms, accumulated since creation) of the phases of the
 * coarse steps: ms[0] hydro calls (Courant, set_unew, godunov_fine, set_uold,
 * each waiting for its kernels), ms[1] flag_fine, ms[2] make_grid_fine,
 * ms[3] remove_grid_fine, ms[4] restrict_fine, ms[5] the part of all
 * phases (and of creation) spent rebuilding level tables after tree edits,
 * and ms[6] the part spent waiting for the stream before the flags' copy
 * back is read (which includes every launch queued before it). */
int rh_amr_timers(rh_amr* run, double* ms);

#ifdef __cplusplus
}
#endif

#endif /* RHYDRO_AMR_H */
