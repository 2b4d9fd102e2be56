/*
 * rhydro.h -- C-ABI of the MI355X-native per-level Godunov hydro path.
 *
 * These entry points replace, one for one, the virtual methods of the
 * reference's hydro solver (include/ramses/solvers/hydro/HydroSolver.hpp:19-29
 * of RAMSES-CPP).  A HydroSolver subclass in the reference's patch/ directory
 * forwards each virtual to the matching rh_* call.  All functions return an
 * rh_status (0 = success); rh_last_error() gives the message of the last
 * failure on a context.  Plain C types only: no HIP or torch types cross this
 * boundary.
 *
 * Threading: a context is not thread-safe.  One context per AmrGrid, one HIP
 * stream per context.  Calls on a context are ordered on its stream.
 */
#ifndef RHYDRO_H
#define RHYDRO_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    RH_OK = 0,
    RH_ERR_ARG = 1,          /* invalid argument or call order            */
    RH_ERR_HIP = 2,          /* HIP runtime error (message has details)   */
    RH_ERR_UNSUPPORTED = 3,  /* configuration outside the GPU path        */
    RH_ERR_TOPOLOGY = 4,     /* mesh the reference itself cannot produce  */
    RH_ERR_NO_DEVICE = 5,    /* no usable GPU                             */
    RH_ERR_HOST = 6          /* host-side failure: memory exhausted or an
                                internal exception (message has details) */
} rh_status;

/* Riemann solver ids.  The adapter resolves the namelist string exactly as
 * the reference's face loop does (HydroSolver.cpp:456-461): "hllc" -> HLLC,
 * "hll" -> HLL, "exact" -> EXACT, "acoustic" -> ACOUSTIC, anything else
 * (including a missing key) -> LLF. */
typedef enum {
    RH_RIEMANN_LLF = 0,
    RH_RIEMANN_HLL = 1,
    RH_RIEMANN_HLLC = 2,
    RH_RIEMANN_EXACT = 3,
    RH_RIEMANN_ACOUSTIC = 4
} rh_riemann;

typedef enum {
    RH_PREC_FP64 = 0,  /* bit-faithful FP64 form                               */
    RH_PREC_MIXED = 1  /* FP32 slopes/trace/Riemann, FP64 update and reflux    */
} rh_precision;

/* Residency of the cell state between calls. */
typedef enum {
    RH_HOST_COHERENT = 0,   /* host uold/unew valid after every call (AMR runs) */
    RH_DEVICE_RESIDENT = 1  /* state stays in HBM; host copies only via rh_sync */
} rh_mode;

enum { RH_SYNC_TO_HOST = 1, RH_SYNC_TO_DEVICE = 2 };
enum { RH_UOLD = 1, RH_UNEW = 2 };

/* Namelist parameters of the path (hydro_params, refine_params). */
typedef struct {
    double gamma;        /* hydro_params.gamma (default 1.4)                 */
    int slope_type;      /* hydro_params.slope_type (default 1), 0..8        */
    int riemann;         /* rh_riemann, resolved as described above          */
    int interpol_type;   /* refine_params.interpol_type (default 1)          */
    int interpol_var;    /* refine_params.interpol_var (default 0)           */
    int nvar;            /* must be 5 (NDIM=3, nener=0, no passive scalars)  */
    int nener;           /* must be 0                                        */
    int barotropic_eos;  /* must be 0                                        */
    double boxlen;       /* amr_params.boxlen (Courant start value)          */
    int precision;       /* rh_precision                                     */
    int device;          /* HIP device ordinal                               */
} rh_params;

/* View of the reference's AmrGrid arrays (AmrGrid.hpp:15-33,104-107).  All
 * index arrays are 1-based as in the reference; state is column-major
 * uold[(ivar-1)*ncell + (icell-1)].  The caller owns every array and bumps
 * mesh_epoch whenever topology (father/son/nbor/next/headl) changes. */
typedef struct {
    int ncoarse, ngridmax, ncell, nvar;
    int nx, ny, nz;
    int nlevelmax, ncpu, nboundary, myid;
    const int* father;   /* [ngridmax]                                    */
    const int* son;      /* [ncell]                                       */
    const int* nbor;     /* [6*ngridmax], (j-1)*ngridmax + igrid-1        */
    const int* next;     /* [ngridmax]                                    */
    const int* headl;    /* [(nlevelmax+1)*(ncpu+nboundary)]              */
    double* uold;        /* [nvar*ncell]                                  */
    double* unew;        /* [nvar*ncell]                                  */
    const double* f;     /* [3*ncell] gravity or NULL (treated as zero)   */
    long long mesh_epoch;
} rh_mesh;

typedef struct rh_ctx rh_ctx;

/* Replaces the constructor HydroSolver(AmrGrid&, Config&)
 * (HydroSolver.hpp:16) and the namelist reads of HydroSolver.cpp:55,62,456
 * and :216-217. */
int rh_create(const rh_params* params, rh_ctx** out);

/* Replaces the destructor (HydroSolver.cpp:16). */
int rh_destroy(rh_ctx* ctx);

/* Binds (or re-binds) the AmrGrid arrays.  Uploads topology and state; the
 * level tables are rebuilt lazily when mesh_epoch or any pointer changes.
 * No reference counterpart: the reference reads AmrGrid members directly. */
int rh_bind(rh_ctx* ctx, const rh_mesh* mesh, int mode);

/* Replaces HydroSolver::set_unew(ilevel) (HydroSolver.cpp:575-590). */
int rh_set_unew(rh_ctx* ctx, int ilevel);

/* Replaces HydroSolver::godunov_fine(ilevel, dt, dx) (HydroSolver.cpp:18-573).
 * Bug-compatible with the reference, including its refluxing order. */
int rh_godunov_fine(rh_ctx* ctx, int ilevel, double dt, double dx);

/* Replaces HydroSolver::set_uold(ilevel) (HydroSolver.cpp:592-611). */
int rh_set_uold(rh_ctx* ctx, int ilevel);

/* Replaces HydroSolver::compute_courant_step(ilevel, dx, gamma,
 * courant_factor) (HydroSolver.cpp:878-977); the result goes to *dt_out. */
int rh_courant(rh_ctx* ctx, int ilevel, double dx, double gamma, double courant_factor, double* dt_out);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (serving loops, harnesses that forbid it): the Courant step of
 * rh_courant, with its block minima and result in a static device array of
 * the library instead of buffers the context allocates.  It calls no device
 * allocation or free.  Limits: the level may hold at most 8,388,608 octs
 * (RH_ERR_ARG beyond); its tables must already be built for the bound mesh
 * by an earlier call on that level (rh_courant, rh_set_unew,
 * rh_godunov_fine), else RH_ERR_ARG.  Every context on one GPU shares the
 * array: calls that overlap on one GPU race. */
int rh_courant_noalloc(rh_ctx* ctx, int ilevel, double dx, double gamma, double courant_factor, double* dt_out);

/* Replaces HydroSolver::interpol_hydro(u1[7][64], u2[8][64])
 * (HydroSolver.cpp:734-876), the prolongation used when refining.  u1 rows
 * are the coarse cell and its 6 neighbours; u2 rows the 8 children.  Runs
 * on the GPU; n_octs independent prolongations per call (u1 is
 * [n_octs][7][64], u2 is [n_octs][8][64]). */
int rh_interpol_hydro(rh_ctx* ctx, int n_octs, const double* u1, double* u2);

/* Replaces HydroSolver::add_gravity_source_terms(ilevel, dt)
 * (HydroSolver.cpp:980-998): for every leaf cell of the level,
 * unew momentum += rho*f*dt (clamped to +-10 rho) and unew energy +=
 * uold momentum . f dt + rho f^2 dt^2 / 2, with the bound f (zero when the
 * mesh has no f, as AmrGrid::allocate leaves f_vec). */
int rh_add_gravity_source_terms(rh_ctx* ctx, int ilevel, double dt);

/* Replaces HydroSolver::synchro_hydro_fine(ilevel, dt)
 * (HydroSolver.cpp:1000-1043, nener = 0, no barotropic EOS): the same kick
 * applied to uold at constant internal energy (gamma from rh_params). */
int rh_synchro_hydro_fine(rh_ctx* ctx, int ilevel, double dt);

/* Explicit host<->device copies of uold and/or unew (what = RH_UOLD|RH_UNEW).
 * No reference counterpart (the reference state lives on the host). */
int rh_sync(rh_ctx* ctx, int direction, int what);

/* Text of the last error on ctx (or of rh_create when ctx is NULL). */
const char* rh_last_error(const rh_ctx* ctx);

/* One uniform coarse step of a single-level run, entirely on the device:
 * dt = courant(ilevel) -> set_unew -> godunov_fine -> set_uold, repeated
 * nsteps times, the time step staying in device memory.  Equivalent to
 * Simulation::amr_step(ilevel, 1) with levelmin == levelmax
 * (Simulation.cpp:522-635).  dt_out (may be NULL) receives the nsteps time
 * steps. */
int rh_run_uniform(rh_ctx* ctx, int ilevel, double dx, double courant_factor, int nsteps, double* dt_out);

#ifdef __cplusplus
}
#endif

#endif /* RHYDRO_H */
