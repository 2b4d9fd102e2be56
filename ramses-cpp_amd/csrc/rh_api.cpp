// rh_api.cpp -- the C-ABI (rhydro.h): context, binding, residency and the
// per-call orchestration of the level kernels.  Host code; kernels live in
// rh_hydro.hip.
#include <hip/hip_runtime.h>

#include <chrono>
#include <utility>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "rh_context.h"
#include "rh_guard.h"
#include "rh_kernels.h"
#include "rh_tables.h"
#include "rhydro.h"
#include "rhydro_driver.h"
#include "rhydro_ensemble.h"

static_assert(rh::REC_REFINED == rh::REC_REFINED_DEV, "record flags");
static_assert(rh::REC_BOUNDARY == rh::REC_BOUNDARY_DEV, "record flags");
static_assert(rh::REC_REFLUX == rh::REC_REFLUX_DEV, "record flags");

namespace rhi {

std::string g_create_error;

int fail(rh_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

int hip_fail(rh_ctx* c, hipError_t e, const char* what) {
    return fail(c, RH_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}


// Callers have synchronised the stream (no queued kernel reads the tables).
void free_level(DevLevel& L) {
    for (void* p : L.allocs) (void)hipFree(p);
    L.allocs.clear();
    if (L.staged) (void)hipEventSynchronize(L.staged);
    if (L.dblock) (void)hipFree(L.dblock);
    if (L.hblock) (void)hipHostFree(L.hblock);
    if (L.staged) (void)hipEventDestroy(L.staged);
    L.dblock = L.hblock = nullptr;
    L.dcap = L.hcap = 0;
    L.staged = nullptr;
    L.dev = rh::LevelDev{};
    L.epoch = -2;
    L.ngridmax = L.ncell = -1;
}


rh::MeshDev mesh_dev(const rh_ctx* c) {
    rh::MeshDev m;
    m.ncoarse = c->mesh.ncoarse;
    m.ngridmax = c->mesh.ngridmax;
    m.ncell = c->mesh.ncell;
    m.nx = c->mesh.nx;
    m.ny = c->mesh.ny;
    m.nz = c->mesh.nz;
    m.uold = c->d_uold;
    m.unew = c->d_unew;
    m.f = c->d_f;
    return m;
}

rh::GasParams gas(const rh_ctx* c, double gamma) {
    rh::GasParams g;
    g.gamma = gamma;
    g.gm1 = gamma - 1.0;
    double smallc = 1e-10;
    g.smalle = smallc * smallc / gamma / (gamma - 1.0);
    g.smallp = smallc * smallc / gamma;
    g.slope_type = c->p.slope_type;
    g.interpol_type = c->p.interpol_type;
    return g;
}

// Level tables for ilevel, rebuilt when the mesh epoch moved.
int get_level(rh_ctx* c, int ilevel, const rh::LevelDev** out) {
    if (ilevel < 1 || ilevel > c->mesh.nlevelmax) return fail(c, RH_ERR_ARG, "ilevel out of range");
    struct Clock {
        rh_ctx* c;
        std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
        ~Clock() {
            c->table_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        }
    } clock{c};
    if ((int)c->levels.size() <= ilevel) c->levels.resize(ilevel + 1);
    DevLevel& L = c->levels[ilevel];
    if (L.epoch != c->mesh.mesh_epoch || L.ngridmax != c->mesh.ngridmax || L.ncell != c->mesh.ncell) {
        // The topology index is built only when some level's tables are.
        if (!c->topo_valid || c->topo.epoch != c->mesh.mesh_epoch) {
            std::string e;
            int st = rh::build_topo_index(c->mesh, c->topo, e);
            if (st != RH_OK) return fail(c, st, e);
            c->topo_valid = true;
        }
        rh::LevelHost h;
        std::string e;
        int st = rh::build_level(c->mesh, c->topo, ilevel, h, e);
        if (st == RH_OK) st = rh::validate_level(h, c->mesh.ngridmax, c->mesh.ncell, e);
        if (st != RH_OK) {
            L.epoch = -2;
            return fail(c, st, e);
        }
        const int nrec = (int)h.rec_n.size();
        std::vector<int4> rec(nrec);
        for (int r = 0; r < nrec; ++r) rec[r] = make_int4(h.rec_n[r], h.rec_nl[r], h.rec_nr[r], h.rec_flags[r]);
        // Layout of the block: each table at a 256-byte boundary, the reflux
        // slots (device only) last.
        struct Part {
            const void* src;
            size_t bytes;
            size_t off;
        };
        Part parts[8] = {{h.grid.data(), h.grid.size() * sizeof(int), 0},
                         {h.leaf.data(), h.leaf.size(), 0},
                         {h.nbo.data(), h.nbo.size() * sizeof(int), 0},
                         {rec.data(), rec.size() * sizeof(int4), 0},
                         {h.rec_slot.data(), h.rec_slot.size() * sizeof(int), 0},
                         {h.tgt_cell.data(), h.tgt_cell.size() * sizeof(int), 0},
                         {h.tgt_off.data(), h.tgt_off.size() * sizeof(int), 0},
                         {h.tgt_slot.data(), h.tgt_slot.size() * sizeof(int), 0}};
        size_t up = 0;
        for (Part& pt : parts) {
            pt.off = up;
            up += (pt.bytes + 255) & ~(size_t)255;
        }
        const size_t slot_bytes = (size_t)h.nslots * 5 * sizeof(double);
        const size_t total = up + slot_bytes;
        if (!L.staged) RH_HIP(c, hipEventCreateWithFlags(&L.staged, hipEventDisableTiming));
        // The previous copy out of the host block has finished before it is
        // refilled (normally long before).
        RH_HIP(c, hipEventSynchronize(L.staged));
        if (up > L.hcap) {
            if (L.hblock) (void)hipHostFree(L.hblock);
            L.hblock = nullptr;
            L.hcap = 0;
            const size_t cap = up + up / 4 + 4096;
            RH_HIP(c, hipHostMalloc(&L.hblock, cap, hipHostMallocDefault));
            L.hcap = cap;
        }
        if (total > L.dcap) {
            // Queued kernels may still read the old block.
            RH_HIP(c, hipStreamSynchronize(c->stream));
            if (L.dblock) (void)hipFree(L.dblock);
            L.dblock = nullptr;
            L.dcap = 0;
            const size_t cap = total + total / 4 + 4096;
            RH_HIP(c, hipMalloc(&L.dblock, cap));
            L.dcap = cap;
        }
        for (const Part& pt : parts)
            if (pt.bytes) std::memcpy((char*)L.hblock + pt.off, pt.src, pt.bytes);
        // Stream-ordered after every kernel queued on the old tables.
        RH_HIP(c, hipMemcpyAsync(L.dblock, L.hblock, up, hipMemcpyHostToDevice, c->stream));
        RH_HIP(c, hipEventRecord(L.staged, c->stream));
        char* base = (char*)L.dblock;
        rh::LevelDev& d = L.dev;
        d = rh::LevelDev{};
        d.noct = h.noct;
        d.nrec = nrec;
        d.nslots = h.nslots;
        d.ntgt = (int)h.tgt_cell.size();
        d.nleaf = h.nleaf;
        auto at = [&](int k) { return parts[k].bytes ? (void*)(base + parts[k].off) : nullptr; };
        d.grid = (int*)at(0);
        d.leaf = (uint8_t*)at(1);
        d.nbo = (int*)at(2);
        d.rec = (int4*)at(3);
        d.rec_slot = (int*)at(4);
        d.tgt_cell = (int*)at(5);
        d.tgt_off = (int*)at(6);
        d.tgt_slot = (int*)at(7);
        d.slot_r = slot_bytes ? (double*)(base + up) : nullptr;
        L.epoch = c->mesh.mesh_epoch;
        L.ngridmax = c->mesh.ngridmax;
        L.ncell = c->mesh.ncell;
    }
    *out = &L.dev;
    return RH_OK;
}

int ensure_part(rh_ctx* c, int n) {
    if (n <= c->npart) return RH_OK;
    if (c->d_part) (void)hipFree(c->d_part);
    c->d_part = nullptr;
    RH_HIP(c, hipMalloc(&c->d_part, (size_t)n * sizeof(double)));
    c->npart = n;
    return RH_OK;
}

int ensure_dt(rh_ctx* c, int n) {
    if (n <= c->ndt) return RH_OK;
    if (c->d_dt) (void)hipFree(c->d_dt);
    c->d_dt = nullptr;
    RH_HIP(c, hipMalloc(&c->d_dt, (size_t)n * sizeof(double)));
    c->ndt = n;
    return RH_OK;
}

size_t state_bytes(const rh_ctx* c) { return (size_t)c->mesh.nvar * c->mesh.ncell * sizeof(double); }

int upload_state(rh_ctx* c, int what) {
    if (c->ens) return ens_upload_state(c, what);
    if (what & RH_UOLD)
        RH_HIP(c, hipMemcpyAsync(c->d_uold, c->mesh.uold, state_bytes(c), hipMemcpyHostToDevice, c->stream));
    if (what & RH_UNEW)
        RH_HIP(c, hipMemcpyAsync(c->d_unew, c->mesh.unew, state_bytes(c), hipMemcpyHostToDevice, c->stream));
    return RH_OK;
}

int download_state(rh_ctx* c, int what) {
    if (c->ens) return ens_download_state(c, what);
    if (what & RH_UOLD)
        RH_HIP(c, hipMemcpyAsync(c->mesh.uold, c->d_uold, state_bytes(c), hipMemcpyDeviceToHost, c->stream));
    if (what & RH_UNEW)
        RH_HIP(c, hipMemcpyAsync(c->mesh.unew, c->d_unew, state_bytes(c), hipMemcpyDeviceToHost, c->stream));
    RH_HIP(c, hipStreamSynchronize(c->stream));
    return RH_OK;
}

int check_bound(rh_ctx* c) {
    if (!c) return RH_ERR_ARG;
    if (!c->bound) return fail(c, RH_ERR_ARG, "no mesh bound (call rh_bind first)");
    if (c->ens) return fail(c, RH_ERR_ARG, "an ensemble is bound: its boxes advance through rh_run_ensemble");
    return RH_OK;
}

// In host-coherent mode the host may have changed uold/unew since the last
// call: copy them in first.
int pre_call(rh_ctx* c) {
    if (c->mode != RH_HOST_COHERENT) return RH_OK;
    int st = upload_state(c, RH_UOLD | RH_UNEW);
    if (st) return st;
    // The host's gravity solver may have updated f since the last call.
    if (c->d_f && c->mesh.f)
        RH_HIP(c, hipMemcpyAsync(c->d_f, c->mesh.f, (size_t)3 * c->mesh.ncell * sizeof(double),
                                 hipMemcpyHostToDevice, c->stream));
    return RH_OK;
}

int post_call(rh_ctx* c, int what) {
    if (c->mode == RH_HOST_COHERENT) return download_state(c, what);
    return RH_OK;
}

hipEvent_t take_event(rh_ctx* c) {
    if (!c->ev_pool.empty()) {
        hipEvent_t e = c->ev_pool.back();
        c->ev_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

// Sums the elapsed time of every recorded pair (waits for the stream).
void collect_timers(rh_ctx* c) {
    if (c->ev_pending.empty()) return;
    (void)hipStreamSynchronize(c->stream);
    for (auto& pe : c->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, pe.second.first, pe.second.second) == hipSuccess) {
            c->t_ms[pe.first] += ms;
            c->t_n[pe.first] += 1;
        }
        c->ev_pool.push_back(pe.second.first);
        c->ev_pool.push_back(pe.second.second);
    }
    c->ev_pending.clear();
}

double exact_inverse(double dx) {
    int e = 0;
    const double m = std::frexp(dx, &e);
    if (!(dx > 0.0) || m != 0.5 || e < -1000 || e > 1000) return 0.0;
    return 1.0 / dx;
}

int variant_for(const rh_ctx* c, const rh::LevelDev& lv) {
    if (c->variant >= 0) return c->variant;
    return (c->p.precision == RH_PREC_FP64 || lv.noct < AUTO_DIMS_MAX_OCTS) ? 6 : 2;
}

void drop_graph(StepGraph& g) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    g.key.clear();
}

void key_add(GraphKey& k, const rh::LevelDev& lv) {
    k.add(lv.noct).add(lv.nrec).add(lv.nslots).add(lv.ntgt).add(lv.nleaf).add(lv.grid).add(lv.leaf).add(lv.nbo);
    k.add(lv.rec).add(lv.rec_slot).add(lv.tgt_cell).add(lv.tgt_off).add(lv.tgt_slot).add(lv.slot_r);
    k.add(lv.nbox).add(lv.obox).add(lv.box_len);
}

void key_add(GraphKey& k, const rh::MeshDev& m) {
    k.add(m.ncoarse).add(m.ngridmax).add(m.ncell).add(m.nx).add(m.ny).add(m.nz).add(m.uold).add(m.unew).add(m.f);
}

void key_add(GraphKey& k, const rh::GasParams& gp) {
    k.add(gp.gamma).add(gp.gm1).add(gp.smalle).add(gp.smallp).add(gp.slope_type).add(gp.interpol_type);
}

void unbind(rh_ctx* c) {
    (void)hipStreamSynchronize(c->stream);
    drop_graph(c->g_uniform);
    drop_graph(c->g_ensemble);
    for (auto& L : c->levels) free_level(L);
    free_level(c->ens_level);
    for (void* p : {(void*)c->d_uold, (void*)c->d_unew, (void*)c->d_f, (void*)c->d_boxlen, (void*)c->d_keys})
        if (p) (void)hipFree(p);
    c->d_uold = c->d_unew = c->d_f = c->d_boxlen = nullptr;
    c->d_keys = nullptr;
    c->state_cap = c->f_cap = 0;
    c->nkeys = 0;
    c->topo_valid = false;
    c->ens = false;
    c->boxes.clear();
    c->ens_levels.clear();
    c->bound = false;
    c->mesh = rh_mesh{};
}

int do_courant(rh_ctx* c, const rh::LevelDev* lv, double dx, double gamma, double cf, double* d_out,
               bool set_unew = false, const rh::AmrTimesArgs& at = rh::AmrTimesArgs{}) {
    int st = ensure_part(c, rh::courant_blocks(*lv) + 1);
    if (st) return st;
    double dt_init = c->p.boxlen / 1e-10;
    Timed tm(c, TK_COURANT);
    RH_HIP(c, rh::launch_courant(mesh_dev(c), *lv, dx, gamma, cf, dt_init, c->d_part, c->npart, d_out, c->stream,
                                 set_unew, at));
    tm.done();
    return RH_OK;
}

int do_godunov(rh_ctx* c, const rh::LevelDev* lv, const double* d_dt, double dx, bool pre_from_uold = false) {
    Timed tm(c, TK_GODUNOV);
    rh::FusedCourant fc;
    const int var = variant_for(c, *lv);
    fc.pre_from_uold = pre_from_uold && var != 0;
    hipError_t e = rh::launch_godunov(mesh_dev(c), *lv, gas(c, c->p.gamma), c->p.riemann, c->p.precision, var,
                                      d_dt, dx, exact_inverse(dx), c->stream, fc);
    if (e == hipErrorNotSupported) return fail(c, RH_ERR_UNSUPPORTED, "riemann/precision combination not supported");
    if (e != hipSuccess) return hip_fail(c, e, "godunov launch");
    tm.done();
    return RH_OK;
}

// godunov_fine and the Courant step with the time step in device memory
// (the adaptive driver's device-side bookkeeping): the same launches as
// rh_godunov_fine / rh_courant, without the host round trip.  Single mesh in
// device-resident mode only.
int courant_dev(rh_ctx* c, int ilevel, double dx, double gamma, double courant_factor, double* d_out,
                bool set_unew, const rh::AmrTimesArgs& at) {
    int st = check_bound(c);
    if (st) return st;
    if (c->mode != RH_DEVICE_RESIDENT) return fail(c, RH_ERR_ARG, "device time steps need a device-resident mesh");
    if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
    const rh::LevelDev* lv;
    if ((st = get_level(c, ilevel, &lv))) return st;
    return do_courant(c, lv, dx, gamma, courant_factor, d_out, set_unew, at);
}

int godunov_dev(rh_ctx* c, int ilevel, const double* d_dt, double dx, bool unew_is_uold) {
    int st = check_bound(c);
    if (st) return st;
    if (c->mode != RH_DEVICE_RESIDENT) return fail(c, RH_ERR_ARG, "device time steps need a device-resident mesh");
    if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
    const rh::LevelDev* lv;
    if ((st = get_level(c, ilevel, &lv))) return st;
    return do_godunov(c, lv, d_dt, dx, unew_is_uold);
}

int set_uold_restrict_dev(rh_ctx* c, int ilevel, const int* d_father) {
    int st = check_bound(c);
    if (st) return st;
    if (c->mode != RH_DEVICE_RESIDENT) return fail(c, RH_ERR_ARG, "the fused restriction needs a device-resident mesh");
    if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
    // The finer level first: get_level may grow the level vector, which the
    // second call then leaves alone, so both pointers stay valid.
    const rh::LevelDev *lv, *fine;
    if ((st = get_level(c, ilevel + 1, &fine))) return st;
    if ((st = get_level(c, ilevel, &lv))) return st;
    Timed tm(c, TK_SET_UOLD);
    RH_HIP(c, rh::launch_set_uold_restrict(mesh_dev(c), *lv, *fine, d_father, c->stream));
    tm.done();
    return RH_OK;
}

}  // namespace rhi

using namespace rhi;

extern "C" {

int rh_create(const rh_params* params, rh_ctx** out) {
    return guard(&g_create_error, nullptr, [&]() -> int {
        if (!params || !out) return fail(nullptr, RH_ERR_ARG, "null argument");
        *out = nullptr;
        if (params->nvar != 5 || params->nener != 0)
            return fail(nullptr, RH_ERR_UNSUPPORTED, "only nvar=5 (NDIM=3, nener=0, no passive scalars) is supported");
        if (params->barotropic_eos) return fail(nullptr, RH_ERR_UNSUPPORTED, "barotropic EOS is not supported");
        if (params->slope_type < 0 || params->slope_type > 8) return fail(nullptr, RH_ERR_ARG, "slope_type outside 0..8");
        if (params->riemann < 0 || params->riemann > 4) return fail(nullptr, RH_ERR_ARG, "unknown riemann id");
        if (params->riemann == RH_RIEMANN_EXACT)
            return fail(nullptr, RH_ERR_UNSUPPORTED, "the exact (Newton-Raphson) Riemann solver is not on the GPU path");
        if (params->precision != RH_PREC_FP64 && params->precision != RH_PREC_MIXED)
            return fail(nullptr, RH_ERR_ARG, "unknown precision");
        if (!(params->gamma > 1.0)) return fail(nullptr, RH_ERR_ARG, "gamma must exceed 1");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
            return fail(nullptr, RH_ERR_NO_DEVICE, "no HIP device available");
        if (params->device < 0 || params->device >= ndev) return fail(nullptr, RH_ERR_ARG, "device ordinal out of range");
        std::unique_ptr<rh_ctx> c(new rh_ctx());
        c->p = *params;
        hipError_t e = hipSetDevice(params->device);
        if (e != hipSuccess) return fail(nullptr, RH_ERR_HIP, std::string("hipSetDevice: ") + hipGetErrorString(e));
        e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
        if (e != hipSuccess) return fail(nullptr, RH_ERR_HIP, std::string("hipStreamCreate: ") + hipGetErrorString(e));
        // Once per device (cached after the first context): the address of
        // the allocation-free Courant step's scratch array on this GPU.
        c->scratch_err = rh::noalloc_scratch(params->device, &c->scratch);
        if (c->scratch_err != hipSuccess) c->scratch = nullptr;
        *out = c.release();
        return RH_OK;
    });
}

int rh_destroy(rh_ctx* c) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_OK;
        (void)hipSetDevice(c->p.device);
        if (c->stream) (void)hipStreamSynchronize(c->stream);
        unbind(c);
        for (void* p : {(void*)c->d_part, (void*)c->d_dt, (void*)c->d_u1, (void*)c->d_u2})
            if (p) (void)hipFree(p);
        collect_timers(c);
        for (hipEvent_t e : c->ev_pool) (void)hipEventDestroy(e);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return RH_OK;
    });
}

int rh_bind(rh_ctx* c, const rh_mesh* m, int mode) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c || !m) return RH_ERR_ARG;
        if (mode != RH_HOST_COHERENT && mode != RH_DEVICE_RESIDENT) return fail(c, RH_ERR_ARG, "unknown mode");
        if (m->nvar != 5) return fail(c, RH_ERR_UNSUPPORTED, "mesh nvar must be 5");
        if (m->ncoarse != m->nx * m->ny * m->nz || m->ncell != m->ncoarse + 8 * m->ngridmax)
            return fail(c, RH_ERR_ARG, "ncoarse/ncell inconsistent with nx,ny,nz,ngridmax");
        if (m->ncpu < 1 || m->myid < 1 || m->myid > m->ncpu + m->nboundary || m->nlevelmax < 1)
            return fail(c, RH_ERR_ARG, "bad ncpu/myid/nlevelmax");
        if (!m->father || !m->son || !m->nbor || !m->next || !m->headl || !m->uold || !m->unew)
            return fail(c, RH_ERR_ARG, "null mesh array");
        RH_HIP(c, hipSetDevice(c->p.device));
        if (c->ens) unbind(c);
        const size_t n = (size_t)m->nvar * m->ncell;
        bool topo_changed = !c->bound || c->mesh.mesh_epoch != m->mesh_epoch || c->mesh.father != m->father ||
                            c->mesh.son != m->son || c->mesh.nbor != m->nbor || c->mesh.next != m->next ||
                            c->mesh.headl != m->headl || c->mesh.ncell != m->ncell;
        if (n > c->state_cap) {
            if (c->d_uold) (void)hipFree(c->d_uold);
            if (c->d_unew) (void)hipFree(c->d_unew);
            c->d_uold = c->d_unew = nullptr;
            c->state_cap = 0;
            RH_HIP(c, hipMalloc(&c->d_uold, n * sizeof(double)));
            RH_HIP(c, hipMalloc(&c->d_unew, n * sizeof(double)));
            c->state_cap = n;
        }
        c->mesh = *m;
        c->mode = mode;
        c->bound = true;
        if (topo_changed) {
            c->topo_valid = false;
            for (auto& L : c->levels) free_level(L);
        }
        int st = upload_state(c, RH_UOLD | RH_UNEW);
        if (st) return st;
        if (m->f) {
            size_t nf = (size_t)3 * m->ncell;
            if (nf > c->f_cap) {
                if (c->d_f) (void)hipFree(c->d_f);
                c->d_f = nullptr;
                RH_HIP(c, hipMalloc(&c->d_f, nf * sizeof(double)));
                c->f_cap = nf;
            }
            RH_HIP(c, hipMemcpyAsync(c->d_f, m->f, nf * sizeof(double), hipMemcpyHostToDevice, c->stream));
        } else if (c->d_f) {
            (void)hipFree(c->d_f);
            c->d_f = nullptr;
            c->f_cap = 0;
        }
        RH_HIP(c, hipStreamSynchronize(c->stream));
        return RH_OK;
    });
}

int rh_set_unew(rh_ctx* c, int ilevel) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        Timed tm(c, TK_SET_UNEW);
        RH_HIP(c, rh::launch_set_unew(mesh_dev(c), *lv, c->stream));
        tm.done();
        return post_call(c, RH_UNEW);
    });
}

int rh_set_unew_oct_lanes(rh_ctx* c, int ilevel) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        Timed tm(c, TK_SET_UNEW);
        RH_HIP(c, rh::launch_set_unew_oct_lanes(mesh_dev(c), *lv, c->stream));
        tm.done();
        return post_call(c, RH_UNEW);
    });
}

int rh_set_uold(rh_ctx* c, int ilevel) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        Timed tm(c, TK_SET_UOLD);
        RH_HIP(c, rh::launch_set_uold(mesh_dev(c), *lv, c->stream));
        tm.done();
        return post_call(c, RH_UOLD);
    });
}

int rh_add_gravity_source_terms(rh_ctx* c, int ilevel, double dt) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        RH_HIP(c, rh::launch_gravity_source(mesh_dev(c), *lv, dt, c->stream));
        return post_call(c, RH_UNEW);
    });
}

int rh_synchro_hydro_fine(rh_ctx* c, int ilevel, double dt) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        RH_HIP(c, rh::launch_synchro(mesh_dev(c), *lv, dt, c->p.gamma, c->stream));
        return post_call(c, RH_UOLD);
    });
}

int rh_godunov_fine(rh_ctx* c, int ilevel, double dt, double dx) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        if ((st = ensure_dt(c, 1))) return st;
        RH_HIP(c, hipMemcpyAsync(c->d_dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
        if ((st = do_godunov(c, lv, c->d_dt, dx))) return st;
        RH_HIP(c, hipStreamSynchronize(c->stream));  // dt staging value is on the stack
        return post_call(c, RH_UNEW);
    });
}

int rh_courant(rh_ctx* c, int ilevel, double dx, double gamma, double courant_factor, double* dt_out) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (!dt_out) return fail(c, RH_ERR_ARG, "null dt_out");
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        if ((st = ensure_dt(c, 1))) return st;
        if ((st = do_courant(c, lv, dx, gamma, courant_factor, c->d_dt))) return st;
        RH_HIP(c, hipMemcpyAsync(dt_out, c->d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        RH_HIP(c, hipStreamSynchronize(c->stream));
        return RH_OK;
    });
}

int rh_courant_noalloc(rh_ctx* c, int ilevel, double dx, double gamma, double courant_factor, double* dt_out) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (!dt_out) return fail(c, RH_ERR_ARG, "null dt_out");
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if (!c->scratch) return hip_fail(c, c->scratch_err, "scratch address of the allocation-free Courant step");
        // Building a level's tables allocates, so they must already be current
        // for the bound mesh (built by an earlier call on this level).
        if (ilevel >= (int)c->levels.size() || c->levels[ilevel].epoch != c->mesh.mesh_epoch ||
            c->levels[ilevel].ngridmax != c->mesh.ngridmax || c->levels[ilevel].ncell != c->mesh.ncell)
            return fail(c, RH_ERR_ARG,
                        "level tables not built for the bound mesh: call rh_courant, rh_set_unew or rh_godunov_fine "
                        "on this level first");
        const rh::LevelDev* lv = &c->levels[ilevel].dev;
        if (lv->noct > rh::NOALLOC_COURANT_MAX_OCTS)
            return fail(c, RH_ERR_ARG,
                        "level has " + std::to_string(lv->noct) + " octs; the allocation-free Courant step takes at most " +
                            std::to_string(rh::NOALLOC_COURANT_MAX_OCTS));
        if ((st = pre_call(c))) return st;
        double* part = c->scratch;
        double* d_dt = c->scratch + rh::NOALLOC_COURANT_PARTS;
        RH_HIP(c, rh::launch_courant(mesh_dev(c), *lv, dx, gamma, courant_factor, c->p.boxlen / 1e-10, part,
                                     rh::NOALLOC_COURANT_PARTS, d_dt, c->stream));
        RH_HIP(c, hipMemcpyAsync(dt_out, d_dt, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        RH_HIP(c, hipStreamSynchronize(c->stream));
        return RH_OK;
    });
}

int rh_interpol_hydro(rh_ctx* c, int n_octs, const double* u1, double* u2) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        if (n_octs < 0 || (n_octs > 0 && (!u1 || !u2))) return fail(c, RH_ERR_ARG, "bad interpol arguments");
        if (n_octs == 0) return RH_OK;
        RH_HIP(c, hipSetDevice(c->p.device));
        if (n_octs > c->interpol_cap) {
            if (c->d_u1) (void)hipFree(c->d_u1);
            if (c->d_u2) (void)hipFree(c->d_u2);
            c->d_u1 = c->d_u2 = nullptr;
            c->interpol_cap = 0;
            RH_HIP(c, hipMalloc(&c->d_u1, (size_t)n_octs * 7 * 64 * sizeof(double)));
            RH_HIP(c, hipMalloc(&c->d_u2, (size_t)n_octs * 8 * 64 * sizeof(double)));
            c->interpol_cap = n_octs;
        }
        RH_HIP(c, hipMemcpyAsync(c->d_u1, u1, (size_t)n_octs * 7 * 64 * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        // Rows beyond nvar are returned unchanged from what the caller passed.
        RH_HIP(c, hipMemcpyAsync(c->d_u2, u2, (size_t)n_octs * 8 * 64 * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        RH_HIP(c, rh::launch_interpol(n_octs, c->d_u1, c->d_u2, gas(c, c->p.gamma), c->p.interpol_var, c->stream));
        RH_HIP(c, hipMemcpyAsync(u2, c->d_u2, (size_t)n_octs * 8 * 64 * sizeof(double), hipMemcpyDeviceToHost,
                                 c->stream));
        RH_HIP(c, hipStreamSynchronize(c->stream));
        return RH_OK;
    });
}

int rh_sync(rh_ctx* c, int direction, int what) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        if (!c->bound) return fail(c, RH_ERR_ARG, "no mesh bound (call rh_bind first)");
        int st;
        if (direction == RH_SYNC_TO_HOST) return download_state(c, what);
        if (direction == RH_SYNC_TO_DEVICE) {
            if ((st = upload_state(c, what))) return st;
            RH_HIP(c, hipStreamSynchronize(c->stream));
            return RH_OK;
        }
        return fail(c, RH_ERR_ARG, "unknown sync direction");
    });
}

const char* rh_last_error(const rh_ctx* c) { return c ? c->err.c_str() : g_create_error.c_str(); }

int rh_run_uniform(rh_ctx* c, int ilevel, double dx, double courant_factor, int nsteps, double* dt_out) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (nsteps < 0) return fail(c, RH_ERR_ARG, "nsteps < 0");
        if (ilevel <= 0) return fail(c, RH_ERR_UNSUPPORTED, "level 0 (coarse cells) is not on the GPU path");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        if ((st = ensure_dt(c, nsteps > 0 ? nsteps : 1))) return st;
        const rh::MeshDev m = mesh_dev(c);
        // A level of leaf cells whose neighbours are all at the same level (a
        // uniform periodic level) steps with uold and unew exchanged instead of
        // copied: the Godunov launch of step k reads the state from one buffer,
        // writes the new state into the other (set_unew's copy is the buffer it
        // reads, set_uold's copy is the exchange) and computes the next step's
        // Courant candidates from the values it writes.  At the end the level's
        // cells of both buffers hold the final state, as after set_uold.
        const int var = variant_for(c, *lv);
        const bool fused = var != 0 && lv->ntgt == 0 && lv->nrec == 0 && lv->nleaf == 8LL * lv->noct &&
                           lv->noct > 0;
        if (fused && nsteps > 0) {
            const int nb = rh::godunov_blocks(*lv, var);
            // Both sizes up front: the launches may be captured into a graph,
            // where no allocation may happen.
            if ((st = ensure_part(c, std::max(nb, rh::courant_blocks(*lv)) + 1))) return st;
            const rh::GasParams gp = gas(c, c->p.gamma);
            const double inv_dx = exact_inverse(dx), dt_init = c->p.boxlen / 1e-10;
            auto issue = [&]() -> int {
                int s2;
                rh::MeshDev mk = m;
                for (int k = 0; k < nsteps; ++k) {
                    if (k == 0 && (s2 = do_courant(c, lv, dx, c->p.gamma, courant_factor, c->d_dt))) return s2;
                    const bool next = k + 1 < nsteps;
                    rh::FusedCourant fc;
                    fc.pre_from_uold = true;
                    fc.part = next ? c->d_part : nullptr;
                    fc.courant_factor = courant_factor;
                    {
                        Timed tm(c, TK_GODUNOV);
                        hipError_t e = rh::launch_godunov(mk, *lv, gp, c->p.riemann, c->p.precision, var,
                                                          c->d_dt + k, dx, inv_dx, c->stream, fc);
                        if (e == hipErrorNotSupported)
                            return fail(c, RH_ERR_UNSUPPORTED, "riemann/precision combination not supported");
                        if (e != hipSuccess) return hip_fail(c, e, "godunov launch");
                        tm.done();
                    }
                    if (next) {
                        Timed tm(c, TK_COURANT);
                        RH_HIP(c, rh::launch_courant_final(c->d_part, nb, dt_init, c->d_dt + k + 1, c->stream));
                        tm.done();
                    }
                    std::swap(mk.uold, mk.unew);
                }
                // mk.uold holds the final state; copy the level's cells into
                // the other buffer (set_unew's copy, uold -> unew).
                Timed tm(c, TK_SET_UOLD);
                RH_HIP(c, rh::launch_set_unew(mk, *lv, c->stream));
                tm.done();
                return RH_OK;
            };
            GraphKey key;
            key_add(key, *lv);
            key_add(key, m);
            key_add(key, gp);
            key.add(c->p.riemann).add(c->p.precision).add(var).add(dx).add(inv_dx).add(courant_factor);
            key.add(dt_init).add(nsteps).add(nb).add(c->d_dt).add(c->d_part).add(c->npart);
            if ((st = run_graph(c, c->g_uniform, key, issue))) return st;
        }
        // Otherwise step k is courant -> set_unew -> godunov_fine -> set_uold.
        // Between two steps, set_uold of the first and courant + set_unew of the
        // second run as one fused pass (k_step_join), with the same state and
        // time step.
        for (int k = 0; k < (fused ? 0 : nsteps); ++k) {
            if (k == 0) {
                if ((st = do_courant(c, lv, dx, c->p.gamma, courant_factor, c->d_dt))) return st;
                Timed tm(c, TK_SET_UNEW);
                RH_HIP(c, rh::launch_set_unew(m, *lv, c->stream));
                tm.done();
            }
            // set_unew, or the join pass, has just made unew equal uold on the
            // level's cells.
            if ((st = do_godunov(c, lv, c->d_dt + k, dx, true))) return st;
            if (k + 1 < nsteps) {
                if ((st = ensure_part(c, rh::courant_blocks(*lv) + 1))) return st;
                Timed tm(c, TK_STEP_JOIN);
                RH_HIP(c, rh::launch_step_join(m, *lv, dx, c->p.gamma, courant_factor, c->p.boxlen / 1e-10, c->d_part,
                                               c->npart, c->d_dt + k + 1, c->stream));
                tm.done();
            } else {
                Timed tm(c, TK_SET_UOLD);
                RH_HIP(c, rh::launch_set_uold(m, *lv, c->stream));
                tm.done();
            }
        }
        if (dt_out && nsteps > 0)
            RH_HIP(c, hipMemcpyAsync(dt_out, c->d_dt, (size_t)nsteps * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
        RH_HIP(c, hipStreamSynchronize(c->stream));
        return post_call(c, RH_UOLD | RH_UNEW);
    });
}

int rh_set_graphs(rh_ctx* c, int enable) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        c->graphs = enable != 0;
        if (!c->graphs) {
            (void)hipSetDevice(c->p.device);
            (void)hipStreamSynchronize(c->stream);
            drop_graph(c->g_uniform);
            drop_graph(c->g_ensemble);
        }
        return RH_OK;
    });
}

int rh_graph_stats(rh_ctx* c, long long* captured, long long* launched) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        if (captured) *captured = c->graph_captures;
        if (launched) *launched = c->graph_launches;
        return RH_OK;
    });
}

int rh_set_kernel_variant(rh_ctx* c, int variant) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        if (variant < -1 || variant >= RH_KERNEL_VARIANTS) return fail(c, RH_ERR_ARG, "unknown kernel variant");
        c->variant = variant;
        return RH_OK;
    });
}

int rh_autotune(rh_ctx* c, int ilevel, double dt, double dx, int reps, int* best_variant, double* ms_per_variant) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        if (reps < 1) return fail(c, RH_ERR_ARG, "reps < 1");
        if ((st = pre_call(c))) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        if ((st = ensure_dt(c, 1))) return st;
        RH_HIP(c, hipMemcpyAsync(c->d_dt, &dt, sizeof(double), hipMemcpyHostToDevice, c->stream));
        // A saved copy of unew is restored after each trial launch, so the
        // bound state is left as it was.
        const size_t bytes = state_bytes(c);
        double* saved = nullptr;
        RH_HIP(c, hipMalloc(&saved, bytes));
        if (const hipError_t e = hipMemcpyAsync(saved, c->d_unew, bytes, hipMemcpyDeviceToDevice, c->stream);
            e != hipSuccess) {
            (void)hipFree(saved);
            return hip_fail(c, e, "save unew");
        }
        constexpr int NV = RH_KERNEL_VARIANTS;
        static_assert(NV == rh::KERNEL_VARIANTS, "variant count of the C-ABI and the kernels");
        // Each trial launches as the multi-step runs do: the pre-update state
        // read from uold (variant 0 reads it from unew, its only form) and, on
        // a uniform periodic level, the next step's Courant candidates at the
        // end (into the context's partial-minimum scratch).  The result is
        // discarded by restoring unew.
        const bool epilogue = lv->ntgt == 0 && lv->nrec == 0 && lv->nleaf == 8LL * lv->noct && lv->noct > 0;
        int nb_max = 0;
        for (int v = 1; v < NV; ++v) nb_max = std::max(nb_max, rh::godunov_blocks(*lv, v));
        if (epilogue && (st = ensure_part(c, nb_max + 1))) {
            (void)hipFree(saved);
            return st;
        }
        const rh::GasParams gp = gas(c, c->p.gamma);
        const double inv_dx = exact_inverse(dx);
        auto launch = [&](int v) -> int {
            rh::FusedCourant fc;
            fc.pre_from_uold = v != 0;
            if (v != 0 && epilogue) {
                fc.part = c->d_part;
                fc.courant_factor = 0.8;
            }
            const hipError_t e = rh::launch_godunov(mesh_dev(c), *lv, gp, c->p.riemann, c->p.precision, v, c->d_dt,
                                                    dx, inv_dx, c->stream, fc);
            if (e == hipErrorNotSupported) return fail(c, RH_ERR_UNSUPPORTED, "riemann/precision combination not supported");
            if (e != hipSuccess) return hip_fail(c, e, "godunov launch");
            return RH_OK;
        };
        auto restore = [&]() -> int {
            if (hipMemcpyAsync(c->d_unew, saved, bytes, hipMemcpyDeviceToDevice, c->stream) != hipSuccess)
                return fail(c, RH_ERR_HIP, "restore unew");
            return RH_OK;
        };
        std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
        int rc = RH_OK;
        for (auto& e : ev)
            if (rc == RH_OK && hipEventCreate(&e) != hipSuccess) rc = fail(c, RH_ERR_HIP, "autotune events");
        const int keep = c->variant;
        // Launches after the GPU has been idle run slower for a few
        // milliseconds while its clocks rise, and the first few launches
        // after another kernel can run slower too: warm the GPU up first,
        // then time each variant over `reps` launches that follow four
        // untimed launches of the same variant, back to back.
        for (int w = 0; w < 8 && rc == RH_OK; ++w)
            if ((rc = launch(variant_for(c, *lv))) == RH_OK) rc = restore();
        double total[NV] = {};
        for (int v = 0; v < NV && rc == RH_OK; ++v) {
            for (int r = -4; r < reps && rc == RH_OK; ++r) {
                if (r >= 0) (void)hipEventRecord(ev[2 * r], c->stream);
                rc = launch(v);
                if (r >= 0) (void)hipEventRecord(ev[2 * r + 1], c->stream);
                if (rc == RH_OK) rc = restore();
            }
            if (rc != RH_OK) break;
            if (hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = fail(c, RH_ERR_HIP, "autotune synchronisation");
                break;
            }
            for (int r = 0; r < reps; ++r) {
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]);
                total[v] += ms;
            }
        }
        int best = keep;
        double best_ms = 0.0;
        for (int v = 0; v < NV && rc == RH_OK; ++v) {
            const double avg = total[v] / reps;
            if (ms_per_variant) ms_per_variant[v] = avg;
            if (v == 0 || avg < best_ms) {
                best_ms = avg;
                best = v;
            }
        }
        (void)hipStreamSynchronize(c->stream);
        (void)hipFree(saved);
        for (hipEvent_t e : ev)
            if (e) (void)hipEventDestroy(e);
        c->variant = rc == RH_OK ? best : keep;
        if (best_variant) *best_variant = c->variant;
        return rc;
    });
}

int rh_get_kernel_variant(rh_ctx* c, int* variant) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c || !variant) return RH_ERR_ARG;
        *variant = c->variant;
        return RH_OK;
    });
}

int rh_timers_enable(rh_ctx* c, int enable) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        c->timers = enable != 0;
        return RH_OK;
    });
}

int rh_timers_read(rh_ctx* c, int kernel, double* ms_total, long long* count) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c || kernel < 0 || kernel >= TK_N) return RH_ERR_ARG;
        collect_timers(c);
        if (ms_total) *ms_total = c->t_ms[kernel];
        if (count) *count = c->t_n[kernel];
        return RH_OK;
    });
}

int rh_timers_reset(rh_ctx* c) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        collect_timers(c);
        for (int k = 0; k < TK_N; ++k) {
            c->t_ms[k] = 0;
            c->t_n[k] = 0;
        }
        return RH_OK;
    });
}

int rh_level_leaf_cells(rh_ctx* c, int ilevel, long long* nleaf) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        int st = check_bound(c);
        if (st) return st;
        const rh::LevelDev* lv;
        if ((st = get_level(c, ilevel, &lv))) return st;
        if (nleaf) *nleaf = lv->nleaf;
        return RH_OK;
    });
}

}  // extern "C"
