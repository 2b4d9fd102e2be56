// rh_context.h -- the context behind the C-ABI handle (rh_ctx) and the host
// helpers shared by the single-mesh entry points (rh_api.cpp) and the
// ensemble form (rh_ensemble.cpp).  Internal to the library.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <utility>
#include <vector>

#include "rh_kernels.h"
#include "rh_tables.h"
#include "rhydro.h"

namespace rhi {

struct DevLevel {
    long long epoch = -2;
    // Sizes of the mesh the tables were built and range-checked against; a
    // binding of other sizes rebuilds them even under the same epoch.
    int ngridmax = -1;
    int ncell = -1;
    rh::LevelDev dev;
    std::vector<void*> allocs;  // separately allocated tables (ensemble form)
    // A single mesh's tables: one device block (the tables, then the reflux
    // slots) filled from one pinned host block by one copy, both kept and
    // reused across rebuilds while large enough, so a rebuild neither waits
    // for the stream nor frees device memory.  `staged` is recorded after the
    // copy from hblock; the next rebuild waits on it before refilling hblock.
    void* dblock = nullptr;
    size_t dcap = 0;
    void* hblock = nullptr;
    size_t hcap = 0;
    hipEvent_t staged = nullptr;
};

enum { TK_GODUNOV = 0, TK_COURANT = 1, TK_SET_UNEW = 2, TK_SET_UOLD = 3, TK_STEP_JOIN = 4, TK_N = 5 };

// One mesh of an ensemble: its host arrays and where its cells sit in the
// packed device state (coarse cells from coff+1, grids from goff+1).
struct EnsBox {
    rh_mesh mesh{};
    long long coff = 0, goff = 0;
};

// A multi-step launch sequence captured into a HIP graph, with the bytes of
// every value its launches were captured with (kernel arguments, device
// pointers, counts); it is replayed while a new call's values are the same.
struct StepGraph {
    hipGraphExec_t exec = nullptr;
    std::vector<unsigned char> key;
};

// Builds a StepGraph key field by field (no struct padding in it).
struct GraphKey {
    std::vector<unsigned char> b;
    template <typename T> GraphKey& add(const T& v) {
        const unsigned char* p = reinterpret_cast<const unsigned char*>(&v);
        b.insert(b.end(), p, p + sizeof(T));
        return *this;
    }
};

}  // namespace rhi

struct rh_ctx {
    rh_params p{};
    int mode = RH_HOST_COHERENT;
    hipStream_t stream = nullptr;
    std::string err;
    bool bound = false;
    // Single mesh: the caller's arrays.  Ensemble: a packed forest of every
    // box (ncoarse, ngridmax, ncell summed; no host arrays).
    rh_mesh mesh{};
    double* d_uold = nullptr;
    double* d_unew = nullptr;
    double* d_f = nullptr;
    size_t state_cap = 0;  // doubles per state array
    size_t f_cap = 0;
    rh::TopoIndex topo;
    bool topo_valid = false;
    std::vector<rhi::DevLevel> levels;
    double* d_part = nullptr;
    int npart = 0;
    double* d_dt = nullptr;
    int ndt = 0;
    double* d_u1 = nullptr;
    double* d_u2 = nullptr;
    int interpol_cap = 0;
    // Godunov kernel variant (rh_set_kernel_variant, rh_autotune); -1: the
    // default, chosen per level by variant_for.
    int variant = -1;
    // Device address of the allocation-free Courant step's scratch array on
    // this context's GPU (looked up by rh_create), or null with the lookup's
    // error in scratch_err.
    double* scratch = nullptr;
    hipError_t scratch_err = hipSuccess;
    bool timers = false;
    double t_ms[rhi::TK_N] = {0, 0, 0, 0, 0};
    double table_ms = 0.0;  // host time spent rebuilding level tables
    long long t_n[rhi::TK_N] = {0, 0, 0, 0, 0};
    // Event pairs recorded around launches, summed when the timers are read.
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, std::pair<hipEvent_t, hipEvent_t>>> ev_pending;
    // Ensemble form (rh_bind_ensemble / rh_run_ensemble).
    bool ens = false;
    std::vector<rhi::EnsBox> boxes;
    std::vector<int> ens_levels;  // level of each box the tables were built for
    rhi::DevLevel ens_level;      // concatenated tables of every box
    double* d_boxlen = nullptr;   // [2*nbox]: dx, then 1/dx (0 when not exact)
    unsigned long long* d_keys = nullptr;  // [nkeys]: ordered keys of the per-box dt
    int nkeys = 0;
    // Multi-step runs with exchanged state buffers (rh_run_uniform,
    // rh_run_ensemble) replay a captured graph of their launches unless
    // graphs are off (rh_set_graphs) or the timers are on.
    bool graphs = true;
    rhi::StepGraph g_uniform, g_ensemble;
    long long graph_captures = 0, graph_launches = 0;
};

namespace rhi {

extern std::string g_create_error;

int fail(rh_ctx* c, int code, const std::string& msg);
int hip_fail(rh_ctx* c, hipError_t e, const char* what);

#define RH_HIP(c, call)                                      \
    do {                                                     \
        hipError_t _e = (call);                              \
        if (_e != hipSuccess) return rhi::hip_fail(c, _e, #call); \
    } while (0)

void free_level(DevLevel& L);
rh::MeshDev mesh_dev(const rh_ctx* c);
rh::GasParams gas(const rh_ctx* c, double gamma);
int ensure_dt(rh_ctx* c, int n);
// Level tables of ilevel for the bound mesh, rebuilt when the epoch moved.
int get_level(rh_ctx* c, int ilevel, const rh::LevelDev** out);
void collect_timers(rh_ctx* c);
// A single mesh is bound (not an ensemble); else an error on c.
int check_bound(rh_ctx* c);
// Host-coherent mode: uploads the host state (and f) before a call.
int pre_call(rh_ctx* c);
// Courant minimum of a level into d_out (with set_unew: the level's
// set_unew in the same pass; with at.T: the adaptive run's time-step
// bookkeeping after it), and godunov_fine with dt read from d_dt, both
// device pointers (device-resident single mesh; no host wait).
int courant_dev(rh_ctx* c, int ilevel, double dx, double gamma, double courant_factor, double* d_out,
                bool set_unew = false, const rh::AmrTimesArgs& at = rh::AmrTimesArgs{});
// unew_is_uold: the caller knows that unew equals uold on every leaf cell of
// the level (set_unew ran and nothing wrote unew since), so the kernel reads
// the cells' pre-update state from uold, which it reads anyway.
int godunov_dev(rh_ctx* c, int ilevel, const double* d_dt, double dx, bool unew_is_uold = false);
// rh_set_uold(ilevel) followed by the restriction of level ilevel+1 into
// its father cells (d_father: father cell of every grid, device memory), in
// one launch; device-resident meshes only.
int set_uold_restrict_dev(rh_ctx* c, int ilevel, const int* d_father);
// Frees every device buffer of the current binding (single or ensemble).
void unbind(rh_ctx* c);

// The Godunov kernel variant of a launch on lv: the context's when one is
// set, else the measured default -- one wave per dimension (variant 6) for
// the FP64 form, and for the mixed form on levels of fewer than
// AUTO_DIMS_MAX_OCTS octs; 8 octs per workgroup (variant 2) otherwise.
constexpr int AUTO_DIMS_MAX_OCTS = 16384;
int variant_for(const rh_ctx* c, const rh::LevelDev& lv);

void drop_graph(StepGraph& g);
// The launch-relevant values of a level's tables, a mesh and the gas
// parameters, appended to a graph key.
void key_add(GraphKey& k, const rh::LevelDev& lv);
void key_add(GraphKey& k, const rh::MeshDev& m);
void key_add(GraphKey& k, const rh::GasParams& gp);
// Runs `issue` -- which enqueues a fixed launch sequence on c->stream and
// returns an RH_* status -- through the graph g: captured from it when g is
// empty or was captured with another key, then launched.  Eager (issue
// runs as is) when graphs are off or the timers are on.  issue must not
// allocate, free, synchronise or copy from pageable memory.
template <class F> int run_graph(rh_ctx* c, StepGraph& g, const GraphKey& key, F&& issue) {
    if (!c->graphs || c->timers) return issue();
    if (!g.exec || g.key != key.b) {
        // A launch of the old graph may still be queued.
        if (g.exec) RH_HIP(c, hipStreamSynchronize(c->stream));
        drop_graph(g);
        RH_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        const int st = issue();
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (st != RH_OK || e != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            return st != RH_OK ? st : hip_fail(c, e, "hipStreamEndCapture");
        }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return hip_fail(c, ei, "hipGraphInstantiate");
        g.exec = exec;
        g.key = key.b;
        ++c->graph_captures;
    }
    RH_HIP(c, hipGraphLaunch(g.exec, c->stream));
    ++c->graph_launches;
    return RH_OK;
}

// Allocates a device copy of src, owned by L (freed by free_level).
template <typename T> hipError_t upload(DevLevel& L, T*& dst, const std::vector<T>& src, hipStream_t s) {
    dst = nullptr;
    if (src.empty()) return hipSuccess;
    void* p = nullptr;
    hipError_t e = hipMalloc(&p, src.size() * sizeof(T));
    if (e != hipSuccess) return e;
    L.allocs.push_back(p);
    dst = (T*)p;
    return hipMemcpyAsync(p, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

// When dx is a normal power of two, its reciprocal (exact) for the kernels'
// multiply-instead-of-divide path; 0 otherwise.
double exact_inverse(double dx);

// HIP events around one launch on the context's stream, when timers are on.
hipEvent_t take_event(rh_ctx* c);
struct Timed {
    rh_ctx* c;
    int k;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    Timed(rh_ctx* cc, int kk) : c(cc), k(kk) {
        if (!c->timers) return;
        e0 = take_event(c);
        e1 = take_event(c);
        if (e0) (void)hipEventRecord(e0, c->stream);
    }
    void done() {
        if (!c->timers || !e0 || !e1) return;
        (void)hipEventRecord(e1, c->stream);
        c->ev_pending.push_back({k, {e0, e1}});
        if (c->ev_pending.size() > 4096) collect_timers(c);
    }
};

// Ensemble state transfer (rh_ensemble.cpp).
int ens_upload_state(rh_ctx* c, int what);
int ens_download_state(rh_ctx* c, int what);

}  // namespace rhi
