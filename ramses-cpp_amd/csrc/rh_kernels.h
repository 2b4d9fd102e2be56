// rh_kernels.h -- device-side layout of one level's tables and the kernel
// launch helpers shared by the C-ABI implementation.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace rh {

// Record flags (same values as the host tables in rh_tables.h).
constexpr int REC_REFINED_DEV = 1;
constexpr int REC_BOUNDARY_DEV = 2;
constexpr int REC_REFLUX_DEV = 4;

// Device copy of one level's tables (rh_tables.h describes the contents).
struct LevelDev {
    int noct = 0;
    int nrec = 0;
    int nslots = 0;
    int ntgt = 0;
    long long nleaf = 0;
    int* grid = nullptr;        // [noct]
    uint8_t* leaf = nullptr;    // [noct]
    int* nbo = nullptr;         // [6*noct]
    int4* rec = nullptr;        // [nrec] {n, nl, nr, flags}
    int* rec_slot = nullptr;    // [nrec]
    int* tgt_cell = nullptr;    // [ntgt]
    int* tgt_off = nullptr;     // [ntgt+1]
    int* tgt_slot = nullptr;    // [nslots used]
    double* slot_r = nullptr;   // [5*nslots] coarse reflux contributions
    // Ensemble form: the box of each oct and, per box, its cell width dx
    // (box_len[b]) and exact reciprocal (box_len[nbox+b], 0 when dx is not a
    // power of two).  Null for a single mesh, whose dx is a kernel argument.
    int nbox = 0;
    int* obox = nullptr;              // [noct]
    const double* box_len = nullptr;  // [2*nbox]
};

// Geometry / state of the bound mesh on the device.
struct MeshDev {
    int ncoarse, ngridmax, ncell;
    int nx, ny, nz;
    double* uold;   // [5*ncell]
    double* unew;   // [5*ncell]
    const double* f;  // [3*ncell] or null
};

struct GasParams {
    double gamma, gm1, smalle, smallp;
    int slope_type, interpol_type;
};

// Launchers (defined in rh_hydro.hip).  All enqueue on `stream`.
hipError_t launch_set_unew(const MeshDev& m, const LevelDev& lv, hipStream_t stream);
hipError_t launch_set_uold(const MeshDev& m, const LevelDev& lv, hipStream_t stream);
// The same copy as launch_set_unew with the Godunov kernel's lane mapping
// (k_set_unew_oct_lanes), for calibrating byte counters.
hipError_t launch_set_unew_oct_lanes(const MeshDev& m, const LevelDev& lv, hipStream_t stream);
// Gravity source terms of the level's leaf cells: the kick into unew
// (add_gravity_source_terms) and the kick of uold at constant internal energy
// (synchro_hydro_fine).  m.f null means f = 0.
hipError_t launch_gravity_source(const MeshDev& m, const LevelDev& lv, double dt, hipStream_t stream);
hipError_t launch_synchro(const MeshDev& m, const LevelDev& lv, double dt, double gamma, hipStream_t stream);
// An adaptive run's level step whose time-step bookkeeping (op 0 of
// amr_times in rh_hydro.hip, on table T) the Courant reduction does once its
// minimum is stored; T null: none.
struct AmrTimesArgs {
    double* T = nullptr;
    int il = 0, lmin = 0;
    double nsub_prev = 1.0;
    // then_op 1 or 2: the level has no finer octs, so the bookkeeping's op
    // that ends its step (1 below the finest level, with then_nsub
    // sub-steps; 2 on it) runs right after op 0, in the same thread.
    int then_op = 0;
    double then_nsub = 0.0;
};
// Courant minimum into *d_dt (device).  d_part has >= courant_blocks(lv) doubles.
// set_unew: the same pass also does set_unew (uold -> unew on every cell of
// the level's octs).
hipError_t launch_courant(const MeshDev& m, const LevelDev& lv, double dx, double gamma, double courant_factor,
                          double dt_init, double* d_part, int npart, double* d_dt, hipStream_t stream,
                          bool set_unew = false, const AmrTimesArgs& at = AmrTimesArgs{});
int courant_blocks(const LevelDev& lv);
// Allocation-free Courant step: its block minima and the resulting time step
// live in a file-scope device array (one copy per GPU, held until the process
// ends) of NOALLOC_SCRATCH_DOUBLES doubles -- NOALLOC_COURANT_PARTS block
// minima, then the time step -- so a level of at most
// NOALLOC_COURANT_PARTS * 256 octs needs no device allocation.
constexpr int NOALLOC_COURANT_PARTS = 32768;
constexpr int NOALLOC_SCRATCH_DOUBLES = NOALLOC_COURANT_PARTS + 32;
constexpr long long NOALLOC_COURANT_MAX_OCTS = (long long)NOALLOC_COURANT_PARTS * 256;
// Device address of that array on `device`, looked up once per device (with
// the device current) and cached; later calls only read the cache.
hipError_t noalloc_scratch(int device, double** ptr);
// set_uold of one step, then the Courant minimum (into *d_dt) and set_unew of
// the next, fused into one pass over the level; same d_part requirement.
hipError_t launch_step_join(const MeshDev& m, const LevelDev& lv, double dx, double gamma, double courant_factor,
                            double dt_init, double* d_part, int npart, double* d_dt, hipStream_t stream);
// Godunov update of one level; dt is read from device memory (*d_dt).
// variant 0: one thread per cell, octs strided across lanes; variants 1 and
// 2: the eight cells of an oct in adjacent lanes, sharing by cross-lane
// shuffles, 16 (variant 1) or 8 (variant 2) octs per workgroup; variants 3,
// 4, 5: bricks of 16, 32, 64 consecutive octs per workgroup that share face
// states and fluxes between octs through LDS as well; variant 6: 8 octs per
// workgroup of three waves, one per dimension (for levels too small to fill
// the GPU).
// dt is d_dt[0] for a single mesh and d_dt[box] for an ensemble.  inv_dx is
// 1/dx when dx is a power of two (the FP64 form then multiplies by it, with
// identical results), else 0; an ensemble reads dx and 1/dx per box and
// passes inv_dx != 0 when every box has an exact reciprocal.
constexpr int KERNEL_VARIANTS = 7;
// pre_from_uold: the launch reads the cells' pre-update unew from uold, for
// a caller that knows unew equals uold on the level's leaf cells (what
// set_unew copies; variants 1 to 6).  Fused step of a single-level run or an
// ensemble: each leaf cell also takes the Courant candidate of its new state
// (what set_uold copies into uold) into a block minimum in
// part[godunov_blocks(lv, variant)] (single mesh, part non-null) or into
// its box's key in keys (ensemble, keys non-null; as launch_courant_boxes),
// only for variants 1 to 6 on levels with only leaf cells and same-level
// neighbours (no records, no reflux targets).
struct FusedCourant {
    bool pre_from_uold = false;
    double* part = nullptr;
    unsigned long long* keys = nullptr;
    double courant_factor = 0.0;
};
hipError_t launch_godunov(const MeshDev& m, const LevelDev& lv, const GasParams& gp, int riemann, int precision,
                          int variant, const double* d_dt, double dx, double inv_dx, hipStream_t stream,
                          const FusedCourant& fc = FusedCourant{});
int godunov_blocks(const LevelDev& lv, int variant);
// Minimum of part[0..nparts) and dt_init into *d_dt (the Courant step).
hipError_t launch_courant_final(const double* d_part, int nparts, double dt_init, double* d_dt, hipStream_t stream);
// Ensemble Courant step: keys[b] <- min(keys[b], key(dt candidate)) over the
// leaf cells of box b (keys ordered like the doubles they encode, NaN last).
// join: fused with set_uold before it and set_unew after it, as
// launch_step_join.
hipError_t launch_courant_boxes(const MeshDev& m, const LevelDev& lv, double gamma, double courant_factor,
                                unsigned long long* keys, hipStream_t stream, bool join = false);
// keys[0..n) <- key(value).
hipError_t launch_fill_keys(unsigned long long* keys, int n, double value, hipStream_t stream);
// dt[i] <- the double encoded by keys[i], i < n.
hipError_t launch_keys_to_dt(const unsigned long long* keys, int n, double* dt, hipStream_t stream);
// Adaptive run: prolongation into new octs (d_cells: 8 ints per new oct,
// the refined cell, its six neighbours, the new grid; the same values go to
// the new grid's entries of d_father [ngridmax] and d_nbor [6*ngridmax]) and
// restriction of a list into its fathers.
hipError_t launch_new_octs(const MeshDev& m, int n, const int* d_cells, const GasParams& gp, int interpol_var,
                           int* d_father, int* d_nbor, hipStream_t stream);
hipError_t launch_restrict(const MeshDev& m, const LevelDev& fine, const int* d_father, hipStream_t stream);
// set_uold of lv and the restriction of fine (the next finer level) into it, one launch.
hipError_t launch_set_uold_restrict(const MeshDev& m, const LevelDev& lv, const LevelDev& fine, const int* d_father,
                                    hipStream_t stream);
// Adaptive run's time-step table on the device (doubles): dtnew, dtold and
// the dt each level's last step used, AMR_T_LEVELS entries each, then the
// time t and the latest Courant minimum.
constexpr int AMR_T_LEVELS = 33;
constexpr int AMR_T_DTNEW = 0;
constexpr int AMR_T_DTOLD = AMR_T_LEVELS;
constexpr int AMR_T_USED = 2 * AMR_T_LEVELS;
constexpr int AMR_T_TIME = 3 * AMR_T_LEVELS;
constexpr int AMR_T_DTC = 3 * AMR_T_LEVELS + 1;
constexpr int AMR_T_SIZE = 3 * AMR_T_LEVELS + 2;

// flag_fine's passes over a level list on the device (flags: one byte per
// oct in list order, bit ic-1 for cell ic): one smooth_fine pass, the
// gradient criteria OR-ed in (when a threshold is positive), then nexpand
// smooth_fine passes.  d_a holds the initial flags; d_a and d_b alternate
// between the passes, and *d_result names the one holding the final flags.
hipError_t launch_flag_bits(const MeshDev& m, const LevelDev& lv, const int* d_nbor, double ed, double ep,
                            double ev, double gamma, int nexpand, unsigned char* d_a, unsigned char* d_b,
                            unsigned char** d_result, hipStream_t stream);
hipError_t launch_interpol(int n_octs, const double* d_u1, double* d_u2, const GasParams& gp, int interpol_var,
                           hipStream_t stream);
// Snapshot hydro records of one level list (n grids, d_grids in list order)
// as 32-bit words, pack_hydro_words(n, nbnd) of them, into d_out.
size_t pack_hydro_words(int n, int nbnd);
hipError_t launch_pack_hydro(const MeshDev& m, const int* d_grids, int n, int il, int nbnd, double gamma,
                             uint32_t* d_out, hipStream_t stream);

}  // namespace rh
