/*
 * rhydro_ensemble.h -- the batched form of the hydro path: many independent
 * meshes ("boxes") advanced together on one GPU.
 *
 * No counterpart in the reference, which runs one simulation per process
 * (Simulation::run, Simulation.cpp:298-520).  An ensemble stands in for N such
 * processes running uniform single-level steps (levelmin == levelmax): each
 * box keeps its own mesh, level, cell width and Courant time step, and its
 * result is bit-identical to running it alone through rh_run_uniform (same
 * kernels, same operations).  All boxes advance with one launch per kernel per
 * step, so small boxes share the GPU instead of each under-filling it.
 */
#ifndef RHYDRO_ENSEMBLE_H
#define RHYDRO_ENSEMBLE_H

#include "rhydro.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Binds nbox meshes to ctx (replacing any single-mesh or ensemble binding).
 * Each mesh is described as for rh_bind (host arrays stay owned by the
 * caller; gravity fields are not supported here).  The topology of every box
 * is read now and at the first rh_run_ensemble; bind again after changing
 * any box's topology.  mode is RH_HOST_COHERENT (each run call copies every
 * box's uold/unew in and out) or RH_DEVICE_RESIDENT (copies only at bind
 * and rh_sync).  While an ensemble is bound, the single-mesh entry points
 * (rh_set_unew, rh_godunov_fine, rh_courant, rh_run_uniform, ...) return
 * RH_ERR_ARG; rh_sync copies every box. */
int rh_bind_ensemble(rh_ctx* ctx, int nbox, const rh_mesh* meshes, int mode);

/* Advances every box nsteps uniform steps at level levels[b] with cell width
 * dx[b]: per step and box, dt_b = courant(levels[b]) -> set_unew ->
 * godunov_fine(dt_b) -> set_uold, as rh_run_uniform does for one mesh.
 * dt_out (may be NULL) receives nsteps*nbox time steps, step-major
 * (dt_out[k*nbox + b]). */
int rh_run_ensemble(rh_ctx* ctx, const int* levels, const double* dx, double courant_factor, int nsteps,
                    double* dt_out);

/* Total leaf cells of the boxes at the given levels (cell-updates per step). */
int rh_ensemble_leaf_cells(rh_ctx* ctx, const int* levels, long long* nleaf);

#ifdef __cplusplus
}
#endif

#endif /* RHYDRO_ENSEMBLE_H */
