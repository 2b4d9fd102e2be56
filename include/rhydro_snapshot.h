/*
 * rhydro_snapshot.h -- RAMSES snapshots of GPU-resident runs.
 *
 * Counterpart of Simulation::dump_snapshot (Simulation.cpp:678-732) and the
 * RamsesWriter it drives (RamsesWriter.cpp:18-475, RamsesWriter.hpp:14-32):
 * the directory output_NNNNN with amr_NNNNN.out00001, hydro_NNNNN.out00001,
 * grav_NNNNN.out00001 (Fortran unformatted records: int32 byte count, payload,
 * int32 byte count), header_NNNNN.txt, hydro_file_descriptor.txt,
 * part_file_descriptor.txt and info_NNNNN.txt, byte for byte as the reference
 * writes them, so that the reference's reader (tests/visu/visu_ramses.py)
 * loads them.  ncpu = 1 (rank 0 writes everything, file suffix 00001).
 *
 * The hydro records are packed on the GPU from the device-resident state
 * (primitive variables in file order, record markers included) and reach
 * the host in one copy per level; the tree records are written from the
 * host's integer arrays, where the tree lives.  A drop-in user keeps the
 * reference's own writer and needs none of this.
 */
#ifndef RHYDRO_SNAPSHOT_H
#define RHYDRO_SNAPSHOT_H

#include "rhydro.h"
#include "rhydro_amr.h"

#ifdef __cplusplus
extern "C" {
#endif

/* SnapshotInfo (RamsesWriter.hpp:14-32) plus the namelist values the info
 * file prints (params::levelmin, nlevelmax, boxlen, units_*). */
typedef struct {
    double t;
    int nstep, nstep_coarse, noutput, iout, nener;
    double einit, mass_tot_0, rho_tot;
    double omega_m, omega_l, omega_k, omega_b, h0, aexp_ini, boxlen_ini;
    double aexp, hexp, aexp_old, epot_tot_int, epot_tot_old, mass_sph;
    int levelmin, levelmax;
    double boxlen, units_length, units_density, units_time;
} rh_snapshot_info;

/* The AmrGrid members the writers read (AmrGrid.hpp:15-33), 1-based ids as in
 * rh_mesh.  headl/taill/numbl are [(nlevelmax+1)*(ncpu+nboundary)] indexed
 * ilevel*(ncpu+nboundary) + icpu-1; next/prev/father [ngridmax]; nbor
 * [6*ngridmax]; xg [3*ngridmax]; son/flag1/cpu_map [ncell]; phi [ncell] and
 * f [3*ncell] column-major.  flag1 NULL reads as 0, cpu_map NULL as 1, phi
 * and f NULL as 0 (what AmrGrid::allocate leaves without gravity). */
typedef struct {
    int nx, ny, nz, ncoarse, ngridmax, ncell, nlevelmax, ncpu, nboundary;
    double gamma;
    const int *headl, *taill, *numbl;
    int headf, tailf, numbf;
    const int *next, *prev, *father, *nbor, *son, *flag1, *cpu_map;
    const double* xg;
    const double *phi, *f;
    long long npart;
} rh_snapshot_tree;

/* Simulation::dump_snapshot (Simulation.cpp:678-732) for the mesh bound to
 * ctx: creates <root>/output_NNNNN (NNNNN = info->iout) and writes every file
 * listed above.  The hydro file comes from the context's device state (in
 * host-coherent mode the host arrays are uploaded first, as every entry
 * point does).  The tree must describe the bound mesh (same ncoarse,
 * ngridmax, level lists); ncpu must be 1 and the mesh nvar 5. */
int rh_dump_snapshot(rh_ctx* ctx, const rh_snapshot_tree* tree, const rh_snapshot_info* info, const char* root);

/* RamsesWriter::write_hydro (RamsesWriter.cpp:146-235) alone, to `path`. */
int rh_write_hydro(rh_ctx* ctx, const rh_snapshot_tree* tree, const rh_snapshot_info* info, const char* path);

/* RamsesWriter::write_amr (RamsesWriter.cpp:18-144) alone, to `path`; host
 * arrays only (no context). */
int rh_write_amr(const rh_snapshot_tree* tree, const rh_snapshot_info* info, const char* path);

/* dump_snapshot of an adaptive run (rhydro_amr.h): the run supplies t,
 * nstep, nstep_coarse, levelmin, levelmax, boxlen and its tree; info (may be
 * NULL: no outputs, aexp = 1, units 1) supplies the rest; iout names the
 * directory. */
int rh_amr_dump(rh_amr* run, const rh_snapshot_info* info, int iout, const char* root);

#ifdef __cplusplus
}
#endif

#endif /* RHYDRO_SNAPSHOT_H */
