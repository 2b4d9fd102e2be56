// rh_oracle_snapshot.h -- what the oracle's snapshot writer reads (test
// infrastructure only): the SnapshotInfo fields with the namelist values the
// info file prints, and the AmrGrid arrays of a tree (ncpu = 1), 1-based ids.
#pragma once

extern "C" {

typedef struct {
    double t;
    int nstep, nstep_coarse, noutput, iout, nener;
    double einit, mass_tot_0, rho_tot;
    double omega_m, omega_l, omega_k, omega_b, h0, aexp_ini, boxlen_ini;
    double aexp, hexp, aexp_old, epot_tot_int, epot_tot_old, mass_sph;
    int levelmin, levelmax;
    double boxlen, units_length, units_density, units_time;
} or_snapshot_info;

// headl/taill/numbl [(nlevelmax+1)*(ncpu+nboundary)]; flag1 NULL = 0,
// cpu_map NULL = 1, phi/f NULL = 0.
typedef struct {
    int nx, ny, nz, ncoarse, ngridmax, ncell, nlevelmax, ncpu, nboundary;
    double gamma;
    const int *headl, *taill, *numbl;
    int headf, tailf, numbf;
    const int *next, *prev, *father, *nbor, *son, *flag1, *cpu_map;
    const double* xg;
    const double *phi, *f;
    long long npart;
} or_snapshot_tree;

// Simulation::dump_snapshot into <root>/output_NNNNN from host arrays (uold
// [5*ncell] column-major).  0 on success, -1 on a file error.
int or_dump_snapshot(const or_snapshot_tree* t, const double* uold, const or_snapshot_info* info,
                     const char* root);

}  // extern "C"
