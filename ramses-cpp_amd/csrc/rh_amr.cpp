// rh_amr.cpp -- adaptive runs with the state resident on the GPU
// (rhydro_amr.h).
//
// The host owns the tree (AmrGrid's integer arrays, ncpu = 1) and edits it as
// the reference's tree updater does; every operation on the state runs on the
// device: the hydro step through the context's kernels, flag_fine's passes
// over a level list (smoothing and gradient criteria, k_smooth_bits and
// k_grad_bits: one byte per oct goes each way), the prolongation into new
// octs (k_new_octs) and the restriction (k_restrict).
// After a tree edit the context's mesh epoch moves, so its level tables are
// rebuilt on next use, and the father/nbor arrays the AMR kernels read are
// uploaded again.
//
// Equivalences relied on (ncpu = 1, nboundary = 0):
//   * cpu_map is 1 everywhere and authorize_fine sets flag2 = 1 for every
//     cell before make_grid_fine reads it, so neither array is kept;
//   * the neighbours make_grid_fine reads (TreeUpdater.cpp:82-118) are
//     get_nbor_cells / get_nbor_of_coarse of the refined cell, i.e.
//     rh::nbor_cells_of_cell (periodic coarse wrap, no boundaries);
//   * new octs of one make_grid_fine call read only cells of the refined
//     list and write only their own children, so they are prolongated
//     together in one launch after the tree edits.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "rh_context.h"
#include "rh_guard.h"
#include "rh_kernels.h"
#include "rh_pool.h"
#include "rh_tables.h"
#include "rhydro.h"
#include "rhydro_amr.h"
#include "rhydro_driver.h"
#include "rhydro_snapshot.h"
#include "rh_snapshot.h"

struct rh_amr {
    rh_ctx* ctx = nullptr;
    bool pool_user = false;  // holds a reference to the worker pool
    rh_params hp{};
    rh_amr_params ap{};
    // AmrGrid (AmrGrid.hpp:15-33).
    int nx = 1, ny = 1, nz = 1, ncoarse = 1, ngridmax = 1, ncell = 1, nlevelmax = 1;
    double boxlen = 1.0, gamma = 1.4;
    std::vector<int> father, son, nbor, next, prev, flag1, headl, taill, numbl;
    std::vector<double> xg, uold, unew;
    int headf = 0, tailf = 0, numbf = 0;
    // Simulation (Simulation.hpp).
    int levelmin = 1;
    std::vector<int> nsubcycle;
    // Host copies of the device time-step table (d_times), refreshed by
    // sync_times before the run's time is reported.
    std::vector<double> dtnew, dtold;
    double t = 0.0;
    int nstep = 0;
    long long cell_updates = 0;
    std::vector<int> reg_type;
    std::vector<double> reg[11];  // x,y,z centre, lengths x,y,z, d, u, v, w, p
    // Device buffers of the AMR kernels.
    int* d_father = nullptr;
    int* d_nbor = nullptr;
    int* d_cells = nullptr;
    int ncells_cap = 0;
    // Pinned staging of the new octs' index lists (ncells_cap * 8 ints); the
    // next make_grid_fine waits on cells_staged before refilling it.
    int* h_cells = nullptr;
    hipEvent_t cells_staged = nullptr;
    unsigned char* d_flags = nullptr;  // 8*ngridmax bytes: two flag buffers of flag_fine
    unsigned char* h_flags = nullptr;  // ngridmax bytes of pinned host memory: a list's flags
    double* d_times = nullptr;  // rh::AMR_T_SIZE doubles
    // Each level list's grids in list order, and the edit count of the list
    // they were read at (list_version is bumped by every edit of the list).
    std::vector<std::vector<int>> list_cache;
    std::vector<long long> list_cache_version, list_version;
    bool topo_dirty = true;
    long long epoch = 1;
    rh_mesh mesh{};
    std::string err;
    // Host wall time per phase of the coarse steps (ms): hydro calls, flag_fine,
    // make_grid_fine, remove_grid_fine, restrict_fine.
    double phase_ms[5] = {0, 0, 0, 0, 0};
    // The part of the phases spent waiting for the stream (the flags' copy
    // back, which also waits for every launch queued before it).
    double wait_ms = 0.0;
};

using namespace rhi;

namespace {

int afail(rh_amr* a, int code, const std::string& msg) {
    a->err = msg;
    if (a->ctx) a->ctx->err = msg;
    return code;
}

#define RH_AHIP(a, call)                                                         \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess) return afail(a, RH_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define RH_ACHK(call)            \
    do {                         \
        int _s = (call);         \
        if (_s != RH_OK) return _s; \
    } while (0)

int cell(const rh_amr* a, int igrid, int ic) { return a->ncoarse + (ic - 1) * a->ngridmax + igrid; }

enum { PH_HYDRO = 0, PH_FLAG = 1, PH_MAKE = 2, PH_REMOVE = 3, PH_RESTRICT = 4 };

struct Phase {
    rh_amr* a;
    int k;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    ~Phase() {
        a->phase_ms[k] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
};

void refresh_mesh(rh_amr* a) {
    rh_mesh& m = a->mesh;
    m.ncoarse = a->ncoarse;
    m.ngridmax = a->ngridmax;
    m.ncell = a->ncell;
    m.nvar = 5;
    m.nx = a->nx;
    m.ny = a->ny;
    m.nz = a->nz;
    m.nlevelmax = a->nlevelmax;
    m.ncpu = 1;
    m.nboundary = 0;
    m.myid = 1;
    m.father = a->father.data();
    m.son = a->son.data();
    m.nbor = a->nbor.data();
    m.next = a->next.data();
    m.headl = a->headl.data();
    m.uold = a->uold.data();
    m.unew = a->unew.data();
    m.f = nullptr;
    m.mesh_epoch = a->epoch;
}

// Level list k gained or lost octs.  The tables of level il are built from
// lists il-1 (the coarser neighbours of its records and their neighbours),
// il (its octs and same-level neighbours) and il+1 (its leaf masks), so the
// tables of levels k-1, k and k+1 are stale and the others stay valid: they
// are carried to the new mesh epoch.  The device copies of father/nbor stay
// current without an upload: kernels read them only for listed grids,
// k_new_octs writes the entries of every grid it creates, and a removed
// grid's stale entries are unused until a new oct takes the grid (and
// writes them again).
void tree_changed(rh_amr* a, int k) {
    const long long old = a->epoch;
    a->epoch++;
    a->ctx->mesh.mesh_epoch = a->epoch;
    a->mesh.mesh_epoch = a->epoch;
    std::vector<rhi::DevLevel>& levels = a->ctx->levels;
    for (int il = 1; il < (int)levels.size(); ++il)
        if (levels[il].epoch == old && (il < k - 1 || il > k + 1)) levels[il].epoch = a->epoch;
    if (k >= 0 && k < (int)a->list_version.size()) a->list_version[k]++;
}

// The whole host father/nbor once, before the first kernel that reads them
// (the octs made before it are in the host arrays too).
int upload_topology(rh_amr* a) {
    if (!a->topo_dirty) return RH_OK;
    RH_AHIP(a, hipMemcpyAsync(a->d_father, a->father.data(), (size_t)a->ngridmax * sizeof(int),
                              hipMemcpyHostToDevice, a->ctx->stream));
    RH_AHIP(a, hipMemcpyAsync(a->d_nbor, a->nbor.data(), (size_t)6 * a->ngridmax * sizeof(int),
                              hipMemcpyHostToDevice, a->ctx->stream));
    a->topo_dirty = false;
    return RH_OK;
}

// The device time-step table into the host copies (waits for the stream).
int sync_times(rh_amr* a) {
    std::vector<double> T(rh::AMR_T_SIZE);
    RH_AHIP(a, hipSetDevice(a->hp.device));
    RH_AHIP(a, hipMemcpyAsync(T.data(), a->d_times, T.size() * sizeof(double), hipMemcpyDeviceToHost,
                              a->ctx->stream));
    RH_AHIP(a, hipStreamSynchronize(a->ctx->stream));
    for (int l = 0; l < rh::AMR_T_LEVELS && l < (int)a->dtnew.size(); ++l) {
        a->dtnew[l] = T[rh::AMR_T_DTNEW + l];
        a->dtold[l] = T[rh::AMR_T_DTOLD + l];
    }
    a->t = T[rh::AMR_T_TIME];
    return RH_OK;
}

// ---- AmrGrid lists (AmrGrid.cpp:13-70, 259-302) ---------------------------

void allocate(rh_amr* a) {
    a->ncoarse = a->nx * a->ny * a->nz;
    a->ncell = a->ncoarse + 8 * a->ngridmax;
    const int nl = a->nlevelmax + 1;
    a->headl.assign(nl, 0);
    a->taill.assign(nl, 0);
    a->numbl.assign(nl, 0);
    a->list_cache.assign(nl, std::vector<int>());
    a->list_cache_version.assign(nl, -1);
    a->list_version.assign(nl, 0);
    a->next.assign(a->ngridmax, 0);
    a->prev.assign(a->ngridmax, 0);
    a->father.assign(a->ngridmax, 0);
    a->son.assign(a->ncell, 0);
    a->nbor.assign((size_t)6 * a->ngridmax, 0);
    a->flag1.assign(a->ncell, 0);
    a->xg.assign((size_t)3 * a->ngridmax, 0.0);
    a->uold.assign((size_t)5 * a->ncell, 0.0);
    a->unew.assign((size_t)5 * a->ncell, 0.0);
    a->headf = 1;
    a->tailf = a->ngridmax;
    for (int i = 1; i <= a->ngridmax; ++i) {
        a->next[i - 1] = (i < a->ngridmax) ? i + 1 : 0;
        a->prev[i - 1] = (i > 1) ? i - 1 : 0;
    }
    a->numbf = a->ngridmax;
}

int get_free_grid(rh_amr* a) {
    if (a->headf == 0) return 0;
    int ig = a->headf;
    a->headf = a->next[ig - 1];
    if (a->headf == 0) a->tailf = 0;
    else a->prev[a->headf - 1] = 0;
    a->next[ig - 1] = 0;
    a->prev[ig - 1] = 0;
    a->numbf--;
    return ig;
}

void free_grid(rh_amr* a, int ig) {
    for (int i = 0; i < 6; ++i) a->nbor[(size_t)i * a->ngridmax + ig - 1] = 0;
    int fc = a->father[ig - 1];
    if (fc > 0) a->son[fc - 1] = 0;
    a->father[ig - 1] = 0;
    for (int ic = 1; ic <= 8; ++ic) {
        int idc = cell(a, ig, ic) - 1;
        a->son[idc] = 0;
        a->flag1[idc] = 0;
    }
    if (a->tailf == 0) {
        a->headf = a->tailf = ig;
        a->prev[ig - 1] = a->next[ig - 1] = 0;
    } else {
        a->next[a->tailf - 1] = ig;
        a->prev[ig - 1] = a->tailf;
        a->next[ig - 1] = 0;
        a->tailf = ig;
    }
    a->numbf++;
}

void add_to_list(rh_amr* a, int ig, int il) {
    if (a->headl[il] == 0) {
        a->headl[il] = a->taill[il] = ig;
        a->prev[ig - 1] = a->next[ig - 1] = 0;
    } else {
        int tail = a->taill[il];
        a->next[tail - 1] = ig;
        a->prev[ig - 1] = tail;
        a->next[ig - 1] = 0;
        a->taill[il] = ig;
    }
    a->numbl[il]++;
}

int count(const rh_amr* a, int il) { return (il >= 0 && il <= a->nlevelmax) ? a->numbl[il] : 0; }

// ---- geometry (AmrGrid.hpp:69-80, AmrGrid.cpp:331-387, 474-511) -----------

int depth_of(const rh_amr* a, int ig) {
    int level = 1, cur = ig;
    while (cur > 0) {
        int fc = a->father[cur - 1];
        if (fc <= a->ncoarse) break;
        cur = ((fc - a->ncoarse - 1) % a->ngridmax) + 1;
        level++;
    }
    return level;
}

int cell_level(const rh_amr* a, int icell) {
    if (icell <= a->ncoarse) return 0;
    return depth_of(a, ((icell - a->ncoarse - 1) % a->ngridmax) + 1);
}

void cell_center(const rh_amr* a, int icell, double xc[3]) {
    xc[0] = xc[1] = xc[2] = 0.5 * a->boxlen;
    if (icell <= a->ncoarse) {
        int idx = icell - 1;
        int iz = idx / (a->nx * a->ny);
        idx %= (a->nx * a->ny);
        int iy = idx / a->nx, ix = idx % a->nx;
        double dxc = a->boxlen / std::max({a->nx, a->ny, a->nz});
        xc[0] = (ix + 0.5) * dxc;
        xc[1] = (iy + 0.5) * dxc;
        xc[2] = (iz + 0.5) * dxc;
        return;
    }
    int ig = ((icell - a->ncoarse - 1) % a->ngridmax) + 1;
    int ic0 = (icell - a->ncoarse - 1) / a->ngridmax;
    int ixyz[3] = {ic0 & 1, (ic0 & 2) >> 1, (ic0 & 4) >> 2};
    double dx = a->boxlen / (double)(a->nx * (1 << depth_of(a, ig)));
    double scale = a->boxlen / (double)a->nx;
    for (int d = 0; d < 3; ++d) xc[d] = a->xg[(size_t)d * a->ngridmax + ig - 1] * scale + (double)(ixyz[d] - 0.5) * dx;
}

int find_cell(const rh_amr* a, const double x[3], int level) {
    double dxc = a->boxlen / std::max({a->nx, a->ny, a->nz});
    int ix = std::clamp((int)(x[0] / dxc), 0, a->nx - 1);
    int iy = std::clamp((int)(x[1] / dxc), 0, a->ny - 1);
    int iz = std::clamp((int)(x[2] / dxc), 0, a->nz - 1);
    int cur = iz * a->nx * a->ny + iy * a->nx + ix + 1;
    int lev = 1;
    const double scale = a->boxlen / (double)a->nx;
    while (lev <= level && a->son[cur - 1] > 0) {
        int ig = a->son[cur - 1];
        lev++;
        int ixc = (x[0] > a->xg[(size_t)0 * a->ngridmax + ig - 1] * scale) ? 1 : 0;
        int iyc = (x[1] > a->xg[(size_t)1 * a->ngridmax + ig - 1] * scale) ? 1 : 0;
        int izc = (x[2] > a->xg[(size_t)2 * a->ngridmax + ig - 1] * scale) ? 1 : 0;
        cur = cell(a, ig, izc * 4 + iyc * 2 + ixc + 1);
    }
    return cur;
}

// The cell at offset o (each -1, 0, 1) from cell ic0 of oct g, at the same
// level, reached through the nbor/son links: 0 when a link does not lead to
// an oct at exactly the expected position (a coarser region, or a diagonal
// reached only around a coarser one); the caller then asks find_cell.  The
// expected positions are exact (xg values are dyadic).
int walk_neighbour(const rh_amr* a, int g, int ic0, const int o[3], double w) {
    const int n[3] = {a->nx, a->ny, a->nz};
    int bits = 0;
    for (int d = 0; d < 3; ++d) {
        const int q = ((ic0 >> d) & 1) + o[d];
        const int s = q < 0 ? -1 : (q > 1 ? 1 : 0);
        bits |= (q & 1) << d;
        if (s == 0) continue;
        const int c1 = a->nbor[(size_t)(2 * d + (s > 0 ? 1 : 0)) * a->ngridmax + g - 1];
        if (c1 <= 0 || c1 > a->ncell) return 0;
        const int g2 = a->son[c1 - 1];
        if (g2 <= 0) return 0;
        for (int e = 0; e < 3; ++e) {
            double want = a->xg[(size_t)e * a->ngridmax + g - 1] + (e == d ? s * w : 0.0);
            if (want < 0.0) want += n[e];
            else if (want >= n[e]) want -= n[e];
            if (a->xg[(size_t)e * a->ngridmax + g2 - 1] != want) return 0;
        }
        g = g2;
    }
    return cell(a, g, bits + 1);
}

// Every 27-neighbour of a cell exists at its level and is refined
// (AmrGrid.cpp:474-511 finds them by coordinates).
bool neighbours_refined_by_search(const rh_amr* a, int icell);

// The 4x4x4 cells around oct g at g's level, positions -1..2 per dimension
// relative to g's cell 0 (index ((z+1)*4 + y+1)*4 + x+1), each reached
// through the nbor/son links as walk_neighbour does, 0 where a walk does not
// land exactly.  w is the width of an oct of g's depth D in coarse-cell
// units: a depth-1 oct fills one coarse cell and each refinement halves it
// (make_grid_fine's offsets), so 2/2^D.
void cells_around(const rh_amr* a, int g, int out[64]) {
    const double w = 2.0 / (double)(1 << depth_of(a, g));
    for (int z = -1; z <= 2; ++z)
        for (int y = -1; y <= 2; ++y)
            for (int x = -1; x <= 2; ++x) {
                const int o[3] = {x, y, z};
                out[((z + 1) * 4 + (y + 1)) * 4 + (x + 1)] = walk_neighbour(a, g, 0, o, w);
            }
}

// neighbours_refined_by_search for cell ic0 (0-based) of oct g, from g's
// cells_around: a walk from cell ic0 with offset o takes the same links as
// one from cell 0 with offset o + bits(ic0), so the block holds the 27
// neighbours.  Any neighbour the links do not reach exactly (a coarser
// region) sends the cell to the search.
bool neighbours_refined_in_block(const rh_amr* a, int g, int ic0, const int block[64]) {
    const int bx = ic0 & 1, by = (ic0 >> 1) & 1, bz = (ic0 >> 2) & 1;
    bool any_fallback = false;
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                const int c = block[((bz + dz + 1) * 4 + (by + dy + 1)) * 4 + (bx + dx + 1)];
                if (c == 0) {
                    any_fallback = true;
                    continue;
                }
                if (a->son[c - 1] == 0) return false;
            }
    return any_fallback ? neighbours_refined_by_search(a, a->ncoarse + ic0 * a->ngridmax + g) : true;
}

bool neighbours_refined_by_search(const rh_amr* a, int icell) {
    int level = cell_level(a, icell);
    double xc[3];
    cell_center(a, icell, xc);
    double dxl = a->boxlen / (double)(a->nx * (1 << level));
    for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
            for (int dx = -1; dx <= 1; ++dx) {
                double x[3] = {xc[0] + dx * dxl, xc[1] + dy * dxl, xc[2] + dz * dxl};
                for (int d = 0; d < 3; ++d) {
                    if (x[d] < 0.0) x[d] += a->boxlen;
                    else if (x[d] >= a->boxlen) x[d] -= a->boxlen;
                }
                int c = find_cell(a, x, level);
                if (cell_level(a, c) != level) return false;
                if (a->son[c - 1] == 0) return false;
            }
    return true;
}

// ---- host parallelism ------------------------------------------------------

// The grids of a level list, in list order: walked again only after the
// list was edited (every edit goes through tree_changed).  The reference
// stays valid until the list's next edit.
const std::vector<int>& list_grids(rh_amr* a, int il) {
    if (a->list_cache_version[il] != a->list_version[il]) {
        std::vector<int>& g = a->list_cache[il];
        g.clear();
        g.reserve(a->numbl[il]);
        for (int ig = a->headl[il]; ig > 0; ig = a->next[ig - 1]) g.push_back(ig);
        a->list_cache_version[il] = a->list_version[il];
    }
    return a->list_cache[il];
}

// ---- tree updater (TreeUpdater.cpp) ---------------------------------------

bool child_active(const rh_amr* a, int idc0) {
    int child = a->son[idc0];
    if (child <= 0) return false;
    for (int ics = 1; ics <= 8; ++ics) {
        int s = cell(a, child, ics) - 1;
        if (a->son[s] > 0 || a->flag1[s] == 1) return true;
    }
    return false;
}

// The six same-level neighbour octs of oct ig (0 where none), face order
// 2*dim + side: the son of each father-level neighbour cell
// (get_nbor_cells_exact, AmrGrid.cpp:449-472, takes its cells from them).
void nbor_octs(const rh_amr* a, int ig, int out[6]) {
    for (int f = 0; f < 6; ++f) {
        int nb = a->nbor[(size_t)f * a->ngridmax + ig - 1];
        out[f] = (nb > 0 && nb <= a->ncell) ? a->son[nb - 1] : 0;
    }
}

void smooth_fine(rh_amr* a, int il, const std::vector<int>& grids) {
    if (il == 1) {
        std::vector<int> ok(a->ncoarse, 0);
        for (int ic = 1; ic <= a->ncoarse; ++ic) {
            if (a->flag1[ic - 1] != 0) continue;
            int icn[6];
            rh::nbor_cells_of_cell(a->mesh, ic, icn);
            int nf = 0;
            for (int k = 0; k < 6; ++k)
                if (icn[k] > 0 && icn[k] <= a->ncoarse && a->flag1[icn[k] - 1] == 1) nf++;
            if (nf > 0) ok[ic - 1] = 1;
        }
        for (int i = 0; i < a->ncoarse; ++i)
            if (ok[i]) a->flag1[i] = 1;
        return;
    }
    // Two passes over the octs: which unflagged cells have a flagged
    // same-level face neighbour (from the flags before this pass), then
    // flag them.  A cell's neighbour along d is its sibling inside the oct
    // or the cell of the neighbour oct on that side (get_nbor_cells_exact),
    // each oct's six neighbour octs looked up once.
    std::vector<unsigned char> ok(grids.size(), 0);  // bit ic-1 of each oct
    parallel_for((int)grids.size(), [&](int k) {
        const int ig = grids[k];
        int on[6];
        bool looked = false;
        unsigned char bits = 0;
        for (int ic = 1; ic <= 8; ++ic) {
            if (a->flag1[cell(a, ig, ic) - 1] != 0) continue;
            if (!looked) {
                nbor_octs(a, ig, on);
                looked = true;
            }
            const int ic0 = ic - 1;
            bool any = false;
            for (int d = 0; d < 3 && !any; ++d)
                for (int side = 0; side < 2 && !any; ++side) {
                    const int icS = (ic0 ^ (1 << d)) + 1;
                    const int par = (ic0 >> d) & 1;
                    const int og = par != side ? ig : on[2 * d + side];
                    any = og > 0 && a->flag1[cell(a, og, icS) - 1] == 1;
                }
            if (any) bits |= (unsigned char)(1u << ic0);
        }
        ok[k] = bits;
    });
    parallel_for((int)grids.size(), [&](int k) {
        for (int ic = 1; ic <= 8; ++ic)
            if ((ok[k] >> (ic - 1)) & 1) a->flag1[cell(a, grids[k], ic) - 1] = 1;
    });
}

void test_flag(rh_amr* a, int il, const std::vector<int>& grids) {
    if (il == 1) {
        for (int ic = 1; ic <= a->ncoarse; ++ic)
            if (child_active(a, ic - 1)) a->flag1[ic - 1] = 1;
        return;
    }
    // Each cell's test reads only its own children, which are not flagged
    // here, so the octs are independent.
    parallel_for((int)grids.size(), [&](int k) {
        for (int ic = 1; ic <= 8; ++ic) {
            int idc = cell(a, grids[k], ic) - 1;
            if (child_active(a, idc)) a->flag1[idc] = 1;
        }
    });
}

void ensure_ref_rules(rh_amr* a, const std::vector<int>& grids) {
    // Octs of a list whose fathers are cells of one oct G were made together,
    // so they sit in runs; a run takes G's 4x4x4 cells once and finds each
    // father's 27 neighbours in it.  Coarse fathers, and a non-cubic coarse
    // grid (where the coordinate search scales positions by boxlen/max(n) at
    // the coarse level and by boxlen/nx below it), take the search.
    const bool cubic = a->nx == a->ny && a->nx == a->nz;
    auto father_oct = [&](int ig) {
        const int fc = a->father[ig - 1];
        return fc > a->ncoarse ? ((fc - a->ncoarse - 1) % a->ngridmax) + 1 : 0;
    };
    auto run_start = [&](size_t k) {
        return k == 0 || father_oct(grids[k]) != father_oct(grids[k - 1]) || father_oct(grids[k]) == 0;
    };
    // The octs [k, end) of one run.
    auto do_run = [&](size_t k, size_t end) {
        int block[64];
        bool have = false;
        for (size_t j = k; j < end; ++j) {
            const int ig = grids[j];
            const int fc = a->father[ig - 1];
            bool ok;
            if (fc <= 0) {
                ok = false;
            } else if (fc <= a->ncoarse || !cubic) {
                ok = neighbours_refined_by_search(a, fc);
            } else {
                const int g = ((fc - a->ncoarse - 1) % a->ngridmax) + 1;
                if (!have) {
                    cells_around(a, g, block);
                    have = true;
                }
                ok = neighbours_refined_in_block(a, g, (fc - a->ncoarse - 1) / a->ngridmax, block);
            }
            if (!ok)
                for (int ic = 1; ic <= 8; ++ic) a->flag1[cell(a, ig, ic) - 1] = 0;
        }
    };
    // Each oct's flags depend on the tree only, so the runs go to the worker
    // threads in any grouping.
    const size_t n = grids.size();
    if (n < (1u << 16)) {
        // A run costs 64 neighbour walks and, next to a coarser region, up
        // to eight coordinate searches (microseconds), so the worker threads
        // take lists of a few dozen runs already.
        std::vector<int> start;
        start.reserve(n / 4 + 2);
        for (size_t k = 0; k < n; ++k)
            if (run_start(k)) start.push_back((int)k);
        start.push_back((int)n);
        parallel_for((int)start.size() - 1, [&](int r) { do_run(start[r], start[r + 1]); }, 32);
        return;
    }
    // Long lists: finding the runs is itself split over the threads.  The
    // list is cut into chunks; a chunk takes the runs that start in it, to
    // their ends (which may lie in the next chunk).
    const int nchunks = (int)std::min<size_t>(n / 128, 8192);
    parallel_for(
        nchunks,
        [&](int c) {
            const size_t lo = n * (size_t)c / (size_t)nchunks, hi = n * (size_t)(c + 1) / (size_t)nchunks;
            size_t k = lo;
            while (k < hi && !run_start(k)) ++k;
            while (k < hi) {
                size_t end = k + 1;
                while (end < n && !run_start(end)) ++end;
                do_run(k, end);
                k = end;
            }
        },
        2);
}

// Gradient criteria of the coarse cells (TreeUpdater.cpp:459-480), on the
// host from a copy of their state (the coarse level is not a level list).
struct CState {
    double d, v[3], p;
};

CState cstate(const rh_amr* a, const std::vector<double>& u, int id) {
    CState s{0.0, {0.0, 0.0, 0.0}, 0.0};
    if (id <= 0 || id > a->ncoarse) return s;
    auto U = [&](int iv) { return u[(size_t)(iv - 1) * a->ncoarse + id - 1]; };
    s.d = std::max(U(1), 1e-10);
    double v2 = 0.0;
    for (int k = 0; k < 3; ++k) {
        s.v[k] = U(2 + k) / s.d;
        v2 += s.v[k] * s.v[k];
    }
    double ek = 0.5 * s.d * v2;
    double em = 0.0, e_non = 0.0;
    s.p = std::max((U(5) - ek - em - e_non) * (a->gamma - 1.0), s.d * 1e-10);
    return s;
}

double egrad(double l, double c, double r, double fl) {
    return 2.0 * std::max(std::abs(r - c) / (r + c + fl), std::abs(c - l) / (c + l + fl));
}

int coarse_gradients(rh_amr* a) {
    std::vector<double> u((size_t)5 * a->ncoarse);
    for (int iv = 0; iv < 5; ++iv)
        RH_AHIP(a, hipMemcpyAsync(u.data() + (size_t)iv * a->ncoarse, a->ctx->d_uold + (size_t)iv * a->ncell,
                                  (size_t)a->ncoarse * sizeof(double), hipMemcpyDeviceToHost, a->ctx->stream));
    RH_AHIP(a, hipStreamSynchronize(a->ctx->stream));
    const double ed = a->ap.err_grad_d, ep = a->ap.err_grad_p, ev = a->ap.err_grad_v;
    for (int i = 1; i <= a->ncoarse; ++i) {
        if (a->flag1[i - 1] != 0) continue;
        CState sc = cstate(a, u, i);
        bool ok = false;
        for (int d = 0; d < 3 && !ok; ++d) {
            int nxd = d == 0 ? a->nx : (d == 1 ? a->ny : a->nz);
            int nskip = d == 0 ? 1 : (d == 1 ? a->nx : a->nx * a->ny);
            int il = i - nskip;
            if (il < 1) il += nxd * nskip;
            int ir = i + nskip;
            if (ir > a->ncoarse) ir -= nxd * nskip;
            CState sl = cstate(a, u, il), sr = cstate(a, u, ir);
            if (ed > 0.0 && egrad(sl.d, sc.d, sr.d, 1e-10) > ed) ok = true;
            else if (ep > 0.0 && egrad(sl.p, sc.p, sr.p, 1e-10) > ep) ok = true;
            else if (ev > 0.0) {
                const double fl = 1e-10, g = a->gamma;
                double cg = std::sqrt(std::max(g * sl.p / sl.d, fl * fl));
                double cm = std::sqrt(std::max(g * sc.p / sc.d, fl * fl));
                double cd = std::sqrt(std::max(g * sr.p / sr.d, fl * fl));
                double e = 2.0 * std::max(std::abs(sr.v[d] - sc.v[d]) /
                                              (cd + cm + std::abs(sr.v[d]) + std::abs(sc.v[d]) + fl),
                                          std::abs(sc.v[d] - sl.v[d]) /
                                              (cm + cg + std::abs(sc.v[d]) + std::abs(sl.v[d]) + fl));
                if (e > ev) ok = true;
            }
        }
        if (ok) a->flag1[i - 1] = 1;
    }
    return RH_OK;
}

// TreeUpdater.cpp:328-583.  The coarse level (il = 1) is flagged on the
// host.  For a level list the host sets the initial flags (every cell below
// levelmin, else the cells whose son has a refined or flagged cell) and the
// device runs the passes over the list -- smooth_fine, the gradient
// criteria, nexpand more smooth_fine passes (launch_flag_bits) -- with one
// byte of flags per oct going each way; ensure_ref_rules, which reads the
// tree, stays on the host.  test_flag (applied when il > levelmin and icount
// < nsubcycle) would flag the cells whose son has a refined or flagged cell:
// the initial flags already hold them and flags only grow in between, so it
// changes nothing and is not run.
int flag_fine(rh_amr* a, int il, int icount, int nsub) {
    if (il >= a->nlevelmax) return RH_OK;
    const int lmin = a->levelmin;
    if (il == 1) {
        for (int i = 0; i < a->ncoarse; ++i) a->flag1[i] = 0;
        for (int i = 1; i <= a->ncoarse; ++i)
            if (il < lmin || child_active(a, i - 1)) a->flag1[i - 1] = 1;
        static const std::vector<int> no_grids;
        if (il >= lmin) smooth_fine(a, il, no_grids);
        if (il >= lmin && (a->ap.err_grad_d > 0.0 || a->ap.err_grad_p > 0.0 || a->ap.err_grad_v > 0.0))
            RH_ACHK(coarse_gradients(a));
        if (il > lmin && icount < nsub) test_flag(a, il, no_grids);
        for (int i = 0; i < a->ap.nexpand; ++i) smooth_fine(a, il, no_grids);
        return RH_OK;
    }
    // The list does not change during the passes (they only set flags).
    const std::vector<int>& grids = list_grids(a, il);
    const int n = (int)grids.size();
    unsigned char* hf = a->h_flags;
    parallel_for(n, [&](int k) {
        unsigned char bits = 0;
        for (int ic = 1; ic <= 8; ++ic)
            if (il < lmin || child_active(a, cell(a, grids[k], ic) - 1)) bits |= (unsigned char)(1u << (ic - 1));
        hf[k] = bits;
    });
    // Below levelmin every cell is flagged and the passes change nothing.
    if (il >= lmin && n > 0) {
        const rh::LevelDev* lv;
        RH_ACHK(get_level(a->ctx, il, &lv));
        if (lv->noct != n) return afail(a, RH_ERR_HOST, "level list and level tables disagree");
        RH_ACHK(upload_topology(a));
        const hipStream_t stream = a->ctx->stream;
        unsigned char* res = nullptr;
        RH_AHIP(a, hipMemcpyAsync(a->d_flags, hf, (size_t)n, hipMemcpyHostToDevice, stream));
        RH_AHIP(a, rh::launch_flag_bits(mesh_dev(a->ctx), *lv, a->d_nbor, a->ap.err_grad_d, a->ap.err_grad_p,
                                        a->ap.err_grad_v, a->gamma, a->ap.nexpand, a->d_flags,
                                        a->d_flags + a->ngridmax, &res, stream));
        RH_AHIP(a, hipMemcpyAsync(hf, res, (size_t)n, hipMemcpyDeviceToHost, stream));
        const auto w0 = std::chrono::steady_clock::now();
        RH_AHIP(a, hipStreamSynchronize(stream));
        a->wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    }
    parallel_for(n, [&](int k) {
        for (int ic = 1; ic <= 8; ++ic) a->flag1[cell(a, grids[k], ic) - 1] = (hf[k] >> (ic - 1)) & 1;
    });
    if (il > lmin && icount < nsub) ensure_ref_rules(a, grids);
    return RH_OK;
}

// TreeUpdater.cpp:27-150: topology on the host, prolongation on the device.
int make_grid_fine(rh_amr* a, int il) {
    if (il >= a->nlevelmax) return RH_OK;
    std::vector<int> cells;
    if (il == 1) {
        for (int ic = 1; ic <= a->ncoarse; ++ic)
            if (a->flag1[ic - 1] == 1 && a->son[ic - 1] == 0) cells.push_back(ic);
    } else {
        // Flagged leaf cells in list order (oct by oct, cell by cell), found
        // by the worker threads on long lists (a short scan costs less than
        // waking them) and gathered in that order.
        const std::vector<int>& grids = list_grids(a, il);
        std::vector<unsigned char> hit(grids.size(), 0);
        parallel_for(
            (int)grids.size(),
            [&](int k) {
                unsigned char bits = 0;
                for (int ic = 1; ic <= 8; ++ic) {
                    const int idc = cell(a, grids[k], ic);
                    if (a->flag1[idc - 1] == 1 && a->son[idc - 1] == 0) bits |= (unsigned char)(1u << (ic - 1));
                }
                hit[k] = bits;
            },
            1 << 17);
        for (size_t k = 0; k < grids.size(); ++k)
            if (hit[k])
                for (int ic = 1; ic <= 8; ++ic)
                    if ((hit[k] >> (ic - 1)) & 1) cells.push_back(cell(a, grids[k], ic));
    }
    if (cells.empty()) return RH_OK;
    std::vector<int> idx;
    idx.reserve(cells.size() * 8);
    for (int ic_coarse : cells) {
        int new_ig = get_free_grid(a);
        if (new_ig == 0) {
            tree_changed(a, il + 1);  // the octs made so far are in the list
            return afail(a, RH_ERR_ARG, "ngridmax exhausted (the run does not resize)");
        }
        a->son[ic_coarse - 1] = new_ig;
        a->father[new_ig - 1] = ic_coarse;
        int icn[6];
        rh::nbor_cells_of_cell(a->mesh, ic_coarse, icn);
        for (int f = 0; f < 6; ++f) a->nbor[(size_t)f * a->ngridmax + new_ig - 1] = icn[f];
        if (il == 1) {
            int idx0 = ic_coarse - 1;
            int x[3] = {idx0 % a->nx, (idx0 % (a->nx * a->ny)) / a->nx, idx0 / (a->nx * a->ny)};
            for (int d = 0; d < 3; ++d) a->xg[(size_t)d * a->ngridmax + new_ig - 1] = (double)x[d] + 0.5;
        } else {
            int ig = ((ic_coarse - 1 - a->ncoarse) % a->ngridmax) + 1;
            int ic0 = (ic_coarse - 1 - a->ncoarse) / a->ngridmax;
            int ixyz[3] = {ic0 & 1, (ic0 & 2) >> 1, (ic0 & 4) >> 2};
            for (int d = 0; d < 3; ++d) {
                double dxl = 1.0 / (double)(1 << (il - 1));
                a->xg[(size_t)d * a->ngridmax + new_ig - 1] =
                    a->xg[(size_t)d * a->ngridmax + ig - 1] + (double)(ixyz[d] - 0.5) * dxl;
            }
        }
        idx.push_back(ic_coarse);
        for (int f = 0; f < 6; ++f) idx.push_back(icn[f]);
        idx.push_back(new_ig);
        add_to_list(a, new_ig, il + 1);
    }
    const int n = (int)cells.size();
    if (!a->cells_staged) RH_AHIP(a, hipEventCreateWithFlags(&a->cells_staged, hipEventDisableTiming));
    // The previous copy out of the staging buffer has finished (and, when the
    // buffers grow, every kernel that reads d_cells).
    RH_AHIP(a, hipEventSynchronize(a->cells_staged));
    if (n > a->ncells_cap) {
        RH_AHIP(a, hipStreamSynchronize(a->ctx->stream));
        if (a->d_cells) (void)hipFree(a->d_cells);
        if (a->h_cells) (void)hipHostFree(a->h_cells);
        a->d_cells = nullptr;
        a->h_cells = nullptr;
        a->ncells_cap = 0;
        const int cap = n + n / 4 + 64;
        RH_AHIP(a, hipMalloc(&a->d_cells, (size_t)cap * 8 * sizeof(int)));
        RH_AHIP(a, hipHostMalloc((void**)&a->h_cells, (size_t)cap * 8 * sizeof(int), hipHostMallocDefault));
        a->ncells_cap = cap;
    }
    std::memcpy(a->h_cells, idx.data(), idx.size() * sizeof(int));
    RH_AHIP(a, hipMemcpyAsync(a->d_cells, a->h_cells, idx.size() * sizeof(int), hipMemcpyHostToDevice,
                              a->ctx->stream));
    RH_AHIP(a, hipEventRecord(a->cells_staged, a->ctx->stream));
    RH_AHIP(a, rh::launch_new_octs(mesh_dev(a->ctx), n, a->d_cells, gas(a->ctx, a->gamma), a->hp.interpol_var,
                                   a->d_father, a->d_nbor, a->ctx->stream));
    tree_changed(a, il + 1);
    return RH_OK;
}

// TreeUpdater.cpp:152-193
void remove_grid_fine(rh_amr* a, int il) {
    if (il >= a->nlevelmax) return;
    // Whether an oct goes depends only on its father's flag and its own
    // cells' sons, which no removal from this list changes (a removal clears
    // the removed oct's cells and its father's son link): decided on the
    // worker threads for long lists, then the octs are unlinked in list
    // order.  The cached list stays as it is until tree_changed.
    const std::vector<int>& grids = list_grids(a, il);
    std::vector<unsigned char> drop(grids.size(), 0);
    parallel_for(
        (int)grids.size(),
        [&](int k) {
            const int ig = grids[k];
            const int fc = a->father[ig - 1];
            if (fc <= 0 || a->flag1[fc - 1] != 0) return;
            for (int ic = 1; ic <= 8; ++ic)
                if (a->son[cell(a, ig, ic) - 1] > 0) return;
            drop[k] = 1;
        },
        1 << 17);
    bool changed = false;
    for (size_t k = 0; k < grids.size(); ++k) {
        if (!drop[k]) continue;
        const int ig = grids[k];
        for (int ic = 1; ic <= 8; ++ic) a->flag1[cell(a, ig, ic) - 1] = 0;
        int p = a->prev[ig - 1], n = a->next[ig - 1];
        if (p > 0) a->next[p - 1] = n;
        else a->headl[il] = n;
        if (n > 0) a->prev[n - 1] = p;
        else a->taill[il] = p;
        a->numbl[il]--;
        free_grid(a, ig);
        changed = true;
    }
    if (changed) tree_changed(a, il);
}

// TreeUpdater.cpp:195-215 on the device.
int restrict_fine(rh_amr* a, int il) {
    if (il + 1 > a->nlevelmax || count(a, il + 1) == 0) return RH_OK;
    const rh::LevelDev* lv;
    RH_ACHK(get_level(a->ctx, il + 1, &lv));
    RH_ACHK(upload_topology(a));
    RH_AHIP(a, rh::launch_restrict(mesh_dev(a->ctx), *lv, a->d_father, a->ctx->stream));
    return RH_OK;
}

// Initializer.cpp:19-157 (apply_all) computed on the host for every coarse
// and listed cell, then uploaded.  Cells that no region covers keep their
// state, so the current device state is the starting point.
int apply_all(rh_amr* a) {
    const bool kh = a->ap.ic_kind == 1;
    auto set = [&](int idc, double d, double u, double v, double w, double p) {
        auto U = [&](int iv) -> double& { return a->uold[(size_t)(iv - 1) * a->ncell + idc - 1]; };
        U(1) = d;
        U(2) = d * u;
        U(3) = d * v;
        U(4) = d * w;
        double v2 = u * u;
        v2 += v * v;
        v2 += w * w;
        double e_kin = 0.5 * d * v2;
        double e_int = p / (a->gamma - 1.0);
        double e_nonthermal = 0.0;
        U(5) = e_kin + e_int + e_nonthermal;
    };
    auto apply = [&](int idc) {
        double xc[3];
        cell_center(a, idc, xc);
        const double x = xc[0], y = xc[1], z = xc[2];
        if (kh) {
            const double pi = 3.14159265358979323846;
            const bool in = std::abs(y - 0.5) < 0.25;
            set(idc, in ? 2.0 : 1.0, in ? 0.5 : -0.5, a->ap.kh_amp * std::sin(4.0 * pi * x), 0.0, 2.5);
            return;
        }
        for (size_t r = 0; r < a->reg_type.size(); ++r) {
            const double xr = a->reg[0][r], yr = a->reg[1][r], zr = a->reg[2][r];
            bool in;
            if (a->reg_type[r] == 0)
                in = std::abs(x - xr) < 0.5 * a->reg[3][r] && std::abs(y - yr) < 0.5 * a->reg[4][r] &&
                     std::abs(z - zr) < 0.5 * a->reg[5][r];
            else
                in = std::pow(x - xr, 2) + std::pow(y - yr, 2) + std::pow(z - zr, 2) < std::pow(a->reg[3][r], 2);
            if (in) set(idc, a->reg[6][r], a->reg[7][r], a->reg[8][r], a->reg[9][r], a->reg[10][r]);
        }
    };
    RH_AHIP(a, hipMemcpyAsync(a->uold.data(), a->ctx->d_uold, a->uold.size() * sizeof(double),
                              hipMemcpyDeviceToHost, a->ctx->stream));
    RH_AHIP(a, hipStreamSynchronize(a->ctx->stream));
    for (int i = 1; i <= a->ncoarse; ++i) apply(i);
    for (int il = 1; il <= a->nlevelmax; ++il)
        for (int ig = a->headl[il]; ig > 0; ig = a->next[ig - 1])
            for (int ic = 1; ic <= 8; ++ic) apply(cell(a, ig, ic));
    RH_AHIP(a, hipMemcpyAsync(a->ctx->d_uold, a->uold.data(), a->uold.size() * sizeof(double),
                              hipMemcpyHostToDevice, a->ctx->stream));
    RH_AHIP(a, hipStreamSynchronize(a->ctx->stream));
    return RH_OK;
}

int legacy_init_flow(rh_amr* a) {
    RH_ACHK(apply_all(a));
    for (int il = a->nlevelmax - 1; il >= 0; --il) RH_ACHK(restrict_fine(a, il));
    return RH_OK;
}

// Simulation.cpp:160-255 (refinement part of initialize).
int initialize(rh_amr* a) {
    RH_ACHK(apply_all(a));
    a->nsubcycle.assign(33, 1);
    for (int i = a->levelmin; i < 33; ++i) a->nsubcycle[i] = 2;
    a->dtnew.assign(33, 0.0);
    a->dtold.assign(33, 0.0);
    const int lmin = a->levelmin, lmax = a->nlevelmax;
    for (int il_ref = 1; il_ref <= lmin; ++il_ref)
        for (int il = 1; il <= lmax; ++il) {
            RH_ACHK(flag_fine(a, il, 1, 2));
            RH_ACHK(make_grid_fine(a, il));
        }
    for (int il_ref = lmin + 1; il_ref <= lmax; ++il_ref) {
        RH_ACHK(legacy_init_flow(a));
        for (int il = 1; il <= lmax; ++il) {
            RH_ACHK(flag_fine(a, il, 1, 2));
            RH_ACHK(make_grid_fine(a, il));
        }
        if (count(a, il_ref) == 0) break;
    }
    a->nstep = 0;
    return legacy_init_flow(a);
}

// Simulation.cpp:522-635
int amr_step(rh_amr* a, int il, int icount) {
    if (il > a->nlevelmax) return RH_OK;
    if (count(a, il) == 0 && il > 0) return RH_OK;
    if (a->levelmin < a->nlevelmax && (il == a->levelmin || icount > 1)) {
        {
            Phase ph{a, PH_MAKE};
            for (int i = il; i <= a->nlevelmax; ++i) RH_ACHK(make_grid_fine(a, i + 1));
        }
        Phase ph{a, PH_REMOVE};
        for (int i = a->nlevelmax; i >= il; --i) remove_grid_fine(a, i + 1);
    }
    const double dx = a->boxlen / (double)(a->nx * (1 << il));
    // dtold/dtnew/t live in d_times: the Courant minimum, the subcycling
    // limit and the time update are stream-ordered on the device.
    const int nsub = il < (int)a->nsubcycle.size() ? a->nsubcycle[il] : 1;
    const bool has_finer = il < a->nlevelmax && count(a, il + 1) > 0;
    {
        Phase ph{a, PH_HYDRO};
        // The Courant pass also does the level's set_unew (it copies the
        // uold values the Courant step reads), and its reduction the time
        // step bookkeeping (op 0 of amr_times).
        rh::AmrTimesArgs at;
        at.T = a->d_times;
        at.il = il;
        at.lmin = a->levelmin;
        at.nsub_prev = (double)a->nsubcycle[il - 1];
        // A level without finer octs takes no sub-steps: the bookkeeping that
        // ends its step (op 1, or op 2 on the finest level) follows op 0 in
        // the same launch.
        if (!has_finer) {
            at.then_op = il < a->nlevelmax ? 1 : 2;
            at.then_nsub = il < a->nlevelmax ? (double)nsub : 0.0;
        }
        RH_ACHK(courant_dev(a->ctx, il, dx, a->gamma, a->ap.courant_factor, a->d_times + rh::AMR_T_DTC, true, at));
    }
    if (has_finer) {
        for (int i = 1; i <= nsub; ++i) RH_ACHK(amr_step(a, il + 1, i));
    } else {
        a->nstep++;
    }
    {
        Phase ph{a, PH_HYDRO};
        long long nleaf = 0;
        RH_ACHK(rh_level_leaf_cells(a->ctx, il, &nleaf));
        a->cell_updates += nleaf;
        // Without finer octs no reflux reached this level's unew since its
        // set_unew (in the Courant pass), so unew still equals uold there.
        const bool no_finer = il >= a->nlevelmax || count(a, il + 1) == 0;
        RH_ACHK(godunov_dev(a->ctx, il, a->d_times + rh::AMR_T_USED + il, dx, no_finer));
    }
    if (il < a->nlevelmax && count(a, il + 1) > 0) {
        // set_uold of this level and restrict_fine of the next one in one
        // launch (they write disjoint cells of this level).
        Phase ph{a, PH_RESTRICT};
        RH_ACHK(upload_topology(a));
        RH_ACHK(set_uold_restrict_dev(a->ctx, il, a->d_father));
    } else {
        Phase ph{a, PH_HYDRO};
        RH_ACHK(rh_set_uold(a->ctx, il));
    }
    Phase ph{a, PH_FLAG};
    return flag_fine(a, il + 1, icount, il > 0 ? a->nsubcycle[il - 1] : 1);
}

int step_coarse(rh_amr* a) {
    if (a->levelmin < a->nlevelmax) {
        Phase ph{a, PH_MAKE};
        for (int il = 1; il <= a->levelmin - 1; ++il) RH_ACHK(make_grid_fine(a, il));
    }
    RH_ACHK(amr_step(a, a->levelmin, 1));
    {
        Phase ph{a, PH_FLAG};
        for (int il = a->levelmin; il >= 1; --il) RH_ACHK(flag_fine(a, il, 2, 2));
    }
    if (a->levelmin < a->nlevelmax) {
        Phase ph{a, PH_RESTRICT};
        for (int il = a->levelmin - 1; il >= 0; --il) RH_ACHK(restrict_fine(a, il));
    }
    return RH_OK;
}

void destroy(rh_amr* a) {
    if (a->ctx) {
        (void)hipSetDevice(a->hp.device);
        (void)hipStreamSynchronize(a->ctx->stream);
        for (void* p : {(void*)a->d_father, (void*)a->d_nbor, (void*)a->d_cells, (void*)a->d_flags, (void*)a->d_times})
            if (p) (void)hipFree(p);
        if (a->h_flags) (void)hipHostFree(a->h_flags);
        if (a->h_cells) (void)hipHostFree(a->h_cells);
        if (a->cells_staged) (void)hipEventDestroy(a->cells_staged);
        rh_destroy(a->ctx);
    }
    if (a->pool_user) pool_release();
    delete a;
}

}  // namespace

extern "C" {

int rh_amr_create(const rh_params* hydro, const rh_amr_params* ap, rh_amr** out) {
    return guard(&g_create_error, nullptr, [&]() -> int {
        if (!hydro || !ap || !out) return RH_ERR_ARG;
        *out = nullptr;
        if (ap->levelmin < 1 || ap->levelmax < ap->levelmin || ap->levelmax > 30 || ap->ngridmax < 1 || ap->nx < 1 ||
            ap->ny < 1 || ap->nz < 1 || ap->nexpand < 0 || (ap->ic_kind == 0 && ap->nreg > 0 && !ap->reg_type))
            return RH_ERR_ARG;
        const long long ncell = (long long)ap->nx * ap->ny * ap->nz + 8LL * ap->ngridmax;
        if (ncell > INT_MAX / 5) return RH_ERR_ARG;
        rh_amr* a = new rh_amr();
        // Destroys the run on every early return or exception until it is
        // handed out.
        struct Owner {
            rh_amr* a;
            ~Owner() {
                if (a) destroy(a);
            }
        } own{a};
        pool_acquire();
        a->pool_user = true;
        a->hp = *hydro;
        a->ap = *ap;
        a->nx = ap->nx;
        a->ny = ap->ny;
        a->nz = ap->nz;
        a->ngridmax = ap->ngridmax;
        a->nlevelmax = ap->levelmax;
        a->levelmin = ap->levelmin;
        a->boxlen = hydro->boxlen;
        a->gamma = hydro->gamma;
        if (ap->ic_kind == 0)
            for (int r = 0; r < ap->nreg; ++r) {
                a->reg_type.push_back(ap->reg_type[r]);
                const double* cols[11] = {ap->x_center, ap->y_center, ap->z_center, ap->length_x, ap->length_y,
                                          ap->length_z, ap->d_region, ap->u_region, ap->v_region, ap->w_region,
                                          ap->p_region};
                for (int k = 0; k < 11; ++k) a->reg[k].push_back(cols[k][r]);
            }
        a->ap.reg_type = nullptr;  // the run keeps its own copy
        int st = rh_create(hydro, &a->ctx);
        if (st != RH_OK) return st;
        allocate(a);
        refresh_mesh(a);
        st = rh_bind(a->ctx, &a->mesh, RH_DEVICE_RESIDENT);
        if (st == RH_OK) {
            hipError_t e = hipMalloc(&a->d_father, (size_t)a->ngridmax * sizeof(int));
            if (e == hipSuccess) e = hipMalloc(&a->d_nbor, (size_t)6 * a->ngridmax * sizeof(int));
            if (e == hipSuccess) e = hipMalloc(&a->d_flags, (size_t)8 * a->ngridmax);
            if (e == hipSuccess) e = hipHostMalloc((void**)&a->h_flags, (size_t)a->ngridmax, hipHostMallocDefault);
            if (e == hipSuccess) e = hipMalloc(&a->d_times, (size_t)rh::AMR_T_SIZE * sizeof(double));
            if (e == hipSuccess) {
                const std::vector<double> zero(rh::AMR_T_SIZE, 0.0);  // dtnew = dtold = 0, t = 0
                e = hipMemcpy(a->d_times, zero.data(), zero.size() * sizeof(double), hipMemcpyHostToDevice);
            }
            if (e != hipSuccess) st = afail(a, RH_ERR_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        if (st == RH_OK) st = initialize(a);
        if (st != RH_OK) {
            g_create_error = a->err.empty() ? rh_last_error(a->ctx) : a->err;
            return st;
        }
        own.a = nullptr;
        *out = a;
        return RH_OK;
    });
}

int rh_amr_destroy(rh_amr* a) {
    return guard(a ? &a->err : &g_create_error, (a && a->ctx) ? &a->ctx->err : nullptr, [&]() -> int {
        if (a) destroy(a);
        return RH_OK;
    });
}

int rh_amr_step(rh_amr* a, int nsteps) {
    return guard(a ? &a->err : &g_create_error, (a && a->ctx) ? &a->ctx->err : nullptr, [&]() -> int {
        if (!a || nsteps < 0) return RH_ERR_ARG;
        RH_AHIP(a, hipSetDevice(a->hp.device));
        for (int k = 0; k < nsteps; ++k) RH_ACHK(step_coarse(a));
        RH_AHIP(a, hipStreamSynchronize(a->ctx->stream));
        return RH_OK;
    });
}

int rh_amr_mesh(rh_amr* a, rh_mesh* mesh, const int** numbl, const double** xg) {
    return guard(a ? &a->err : &g_create_error, (a && a->ctx) ? &a->ctx->err : nullptr, [&]() -> int {
        if (!a) return RH_ERR_ARG;
        RH_ACHK(rh_sync(a->ctx, RH_SYNC_TO_HOST, RH_UOLD | RH_UNEW));
        refresh_mesh(a);
        if (mesh) *mesh = a->mesh;
        if (numbl) *numbl = a->numbl.data();
        if (xg) *xg = a->xg.data();
        return RH_OK;
    });
}

int rh_amr_stats(rh_amr* a, double* t, int* nstep, long long* cell_updates) {
    return guard(a ? &a->err : &g_create_error, (a && a->ctx) ? &a->ctx->err : nullptr, [&]() -> int {
        if (!a) return RH_ERR_ARG;
        RH_ACHK(sync_times(a));
        if (t) *t = a->t;
        if (nstep) *nstep = a->nstep;
        if (cell_updates) *cell_updates = a->cell_updates;
        return RH_OK;
    });
}

rh_ctx* rh_amr_context(rh_amr* a) { return a ? a->ctx : nullptr; }

int rh_amr_dump(rh_amr* a, const rh_snapshot_info* info, int iout, const char* root) {
    return guard(a ? &a->err : &g_create_error, (a && a->ctx) ? &a->ctx->err : nullptr, [&]() -> int {
        if (!a || !root) return RH_ERR_ARG;
        rh_snapshot_info in{};
        if (info) {
            in = *info;
        } else {  // SnapshotInfo's defaults (RamsesWriter.hpp:24-31), no outputs, unit scales 1
            in.aexp_ini = in.boxlen_ini = in.aexp = in.aexp_old = 1.0;
            in.units_length = in.units_density = in.units_time = 1.0;
        }
        // What Simulation::dump_snapshot takes from the run (Simulation.cpp:683-692).
        RH_ACHK(sync_times(a));
        in.t = a->t;
        in.nstep = in.nstep_coarse = a->nstep;
        in.iout = iout;
        in.levelmin = a->levelmin;
        in.levelmax = a->nlevelmax;
        in.boxlen = a->boxlen;
        rh_snapshot_tree t{};
        t.nx = a->nx;
        t.ny = a->ny;
        t.nz = a->nz;
        t.ncoarse = a->ncoarse;
        t.ngridmax = a->ngridmax;
        t.ncell = a->ncell;
        t.nlevelmax = a->nlevelmax;
        t.ncpu = 1;
        t.nboundary = 0;
        t.gamma = a->gamma;
        t.headl = a->headl.data();
        t.taill = a->taill.data();
        t.numbl = a->numbl.data();
        t.headf = a->headf;
        t.tailf = a->tailf;
        t.numbf = a->numbf;
        t.next = a->next.data();
        t.prev = a->prev.data();
        t.father = a->father.data();
        t.nbor = a->nbor.data();
        t.son = a->son.data();
        t.flag1 = a->flag1.data();
        t.cpu_map = nullptr;  // every cell belongs to rank 1
        t.xg = a->xg.data();
        return rhi::dump_snapshot(a->ctx, &t, &in, root);
    });
}

int rh_amr_timers(rh_amr* a, double* ms) {
    return guard(a ? &a->err : &g_create_error, (a && a->ctx) ? &a->ctx->err : nullptr, [&]() -> int {
        if (!a || !ms) return RH_ERR_ARG;
        for (int k = 0; k < 5; ++k) ms[k] = a->phase_ms[k];
        ms[5] = a->ctx->table_ms;
        ms[6] = a->wait_ms;
        return RH_OK;
    });
}

}  // extern "C"
