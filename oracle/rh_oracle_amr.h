// rh_oracle_amr.h -- C interface of the oracle's adaptive run
// (rh_oracle_amr.cpp).  TEST INFRASTRUCTURE ONLY.
#pragma once
#include "rh_oracle_core.h"
#include "rh_oracle_snapshot.h"

extern "C" {

// Run configuration (the namelist keys the restated code reads).
typedef struct {
    int nx, ny, nz;
    int ngridmax;
    int levelmin, levelmax;
    double boxlen, gamma, courant_factor;
    double err_grad_d, err_grad_p, err_grad_v;
    int nexpand;
    or_params hydro;
    // Initial condition: kind 0 = regions (square/cube = 0, sphere = 1), kind
    // 1 = Kelvin-Helmholtz slab (rho 2 for |y - 0.5| < 0.25 else 1, u = +0.5
    // inside / -0.5 outside, v = amp*sin(4 pi x), p = 2.5).
    int ic_kind;
    int nreg;
    const int* reg_type;
    const double *x_center, *y_center, *z_center, *length_x, *length_y, *length_z;
    const double *d_region, *u_region, *v_region, *w_region, *p_region;
    double kh_amp;
} or_amr_config;


// View of the run's arrays (owned by the run; valid until the next call).
typedef struct {
    int ncoarse, ngridmax, ncell, nlevelmax;
    const int *father, *son, *nbor, *next, *headl, *numbl, *flag1;
    const double *xg, *uold, *unew;
    double t;
    int nstep;
    const double* dtnew;  // [33]
    long long cell_updates;
    int status;
} or_amr_state;

void* or_amr_create(const or_amr_config* cfg);
int or_amr_step(void* h, int nsteps);
void or_amr_get(void* h, or_amr_state* s);
void or_amr_destroy(void* h);
int or_amr_dump(void* h, const or_snapshot_info* info, int iout, const char* root);

}  // extern "C"
