// rh_ensemble.cpp -- the batched (ensemble) form of the C-ABI: many
// independent meshes ("boxes"), each advanced at its own level with its own
// cell width and its own Courant time step, by one launch per kernel per
// step for all boxes together.
//
// The boxes are packed into one forest on the device.  With NC = sum of the
// boxes' ncoarse and NG = sum of their ngridmax, box b's coarse cell i becomes
// coff_b + i and its grid g becomes goff_b + g, so the packed cell index keeps
// the reference's formula ncoarse + (ic-1)*ngridmax + igrid with NC and NG
// (AmrGrid.hpp:104-107).  Each box's level tables are built by the
// single-mesh builder and concatenated: oct positions, records and reflux
// slots are offset, cell ids mapped.  Boxes never share a face, so every
// contribution keeps its box's processing order and each box's result is
// exactly what a single-mesh context computes for it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <utility>
#include <climits>
#include <cstring>
#include <string>
#include <vector>

#include "rh_context.h"
#include "rh_guard.h"
#include "rh_kernels.h"
#include "rh_tables.h"
#include "rhydro.h"
#include "rhydro_ensemble.h"

using namespace rhi;

namespace {

// Packed index of cell `icell` (1-based) of box b; indices <= 0 (boundary
// markers) pass through.
long long map_cell(const EnsBox& b, const rh_mesh& tot, long long icell) {
    if (icell <= 0) return icell;
    const rh_mesh& m = b.mesh;
    if (icell <= m.ncoarse) return b.coff + icell;
    const long long k = icell - m.ncoarse - 1;
    const long long ic0 = k / m.ngridmax, g = k % m.ngridmax + 1;
    return (long long)tot.ncoarse + ic0 * tot.ngridmax + b.goff + g;
}

int check_box(rh_ctx* c, const rh_mesh& m, int b) {
    const std::string at = "box " + std::to_string(b) + ": ";
    if (m.nvar != 5) return fail(c, RH_ERR_UNSUPPORTED, at + "mesh nvar must be 5");
    if (m.ncoarse != m.nx * m.ny * m.nz || m.ncell != m.ncoarse + 8 * m.ngridmax || m.ngridmax < 1)
        return fail(c, RH_ERR_ARG, at + "ncoarse/ncell inconsistent with nx,ny,nz,ngridmax");
    if (m.ncpu < 1 || m.myid < 1 || m.myid > m.ncpu + m.nboundary || m.nlevelmax < 1)
        return fail(c, RH_ERR_ARG, at + "bad ncpu/myid/nlevelmax");
    if (!m.father || !m.son || !m.nbor || !m.next || !m.headl || !m.uold || !m.unew)
        return fail(c, RH_ERR_ARG, at + "null mesh array");
    if (m.f) return fail(c, RH_ERR_UNSUPPORTED, at + "gravity fields are not supported in an ensemble");
    return RH_OK;
}

// Copies one state array of every box between its host arrays and the
// packed device array: per variable, the coarse cells, then the eight cell
// columns (one 2-D copy: 8 rows of ngridmax doubles).
int transfer(rh_ctx* c, double* dev, bool up, bool uold) {
    const rh_mesh& T = c->mesh;
    const hipMemcpyKind kind = up ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
    for (const EnsBox& b : c->boxes) {
        const rh_mesh& m = b.mesh;
        double* host = uold ? m.uold : m.unew;
        for (int iv = 0; iv < 5; ++iv) {
            double* h0 = host + (size_t)iv * m.ncell;
            double* d0 = dev + (size_t)iv * T.ncell;
            double* hc = h0;
            double* dc = d0 + b.coff;
            const size_t nc = (size_t)m.ncoarse * sizeof(double);
            if (up) RH_HIP(c, hipMemcpyAsync(dc, hc, nc, kind, c->stream));
            else RH_HIP(c, hipMemcpyAsync(hc, dc, nc, kind, c->stream));
            double* hg = h0 + m.ncoarse;
            double* dg = d0 + T.ncoarse + b.goff;
            const size_t hp = (size_t)m.ngridmax * sizeof(double), dp = (size_t)T.ngridmax * sizeof(double);
            if (up) RH_HIP(c, hipMemcpy2DAsync(dg, dp, hg, hp, hp, 8, kind, c->stream));
            else RH_HIP(c, hipMemcpy2DAsync(hg, hp, dg, dp, hp, 8, kind, c->stream));
        }
    }
    return RH_OK;
}

// Concatenated tables of every box at its level (levels[b]).
int build_tables(rh_ctx* c, const int* levels) {
    const int nbox = (int)c->boxes.size();
    const rh_mesh& T = c->mesh;
    rh::LevelHost all;
    std::vector<int> obox;
    std::vector<rh::LevelHost> hb(nbox);
    long long noct = 0, nrec = 0;
    for (int b = 0; b < nbox; ++b) {
        const EnsBox& B = c->boxes[b];
        rh::TopoIndex ti;
        std::string e;
        int st = rh::build_topo_index(B.mesh, ti, e);
        if (st == RH_OK) st = rh::build_level(B.mesh, ti, levels[b], hb[b], e);
        if (st != RH_OK) return fail(c, st, "box " + std::to_string(b) + ": " + e);
        noct += hb[b].noct;
        nrec += (long long)hb[b].rec_n.size();
    }
    if (noct > INT_MAX / 8 || nrec > INT_MAX / 8) return fail(c, RH_ERR_ARG, "ensemble too large");
    all.noct = (int)noct;
    all.grid.reserve(noct);
    all.leaf.reserve(noct);
    all.nbo.assign((size_t)6 * noct, 0);
    obox.reserve(noct);
    int P = 0, R = 0, S = 0;
    for (int b = 0; b < nbox; ++b) {
        const EnsBox& B = c->boxes[b];
        const rh::LevelHost& h = hb[b];
        for (int p = 0; p < h.noct; ++p) {
            all.grid.push_back((int)(B.goff + h.grid[p]));
            all.leaf.push_back(h.leaf[p]);
            obox.push_back(b);
            for (int f = 0; f < 6; ++f) {
                const int v = h.nbo[(size_t)f * h.noct + p];
                all.nbo[(size_t)f * noct + P + p] = v >= 0 ? P + v : -1 - (R + (-1 - v));
            }
        }
        for (size_t r = 0; r < h.rec_n.size(); ++r) {
            all.rec_n.push_back((int)map_cell(B, T, h.rec_n[r]));
            all.rec_nl.push_back((int)map_cell(B, T, h.rec_nl[r]));
            all.rec_nr.push_back((int)map_cell(B, T, h.rec_nr[r]));
            all.rec_flags.push_back(h.rec_flags[r]);
            all.rec_slot.push_back(h.rec_slot[r] < 0 ? -1 : S + h.rec_slot[r]);
        }
        const int base = (int)all.tgt_slot.size();
        for (size_t i = 0; i < h.tgt_cell.size(); ++i) {
            all.tgt_cell.push_back((int)map_cell(B, T, h.tgt_cell[i]));
            all.tgt_off.push_back(base + h.tgt_off[i]);
        }
        for (int sl : h.tgt_slot) all.tgt_slot.push_back(S + sl);
        all.nleaf += h.nleaf;
        P += h.noct;
        R += (int)h.rec_n.size();
        S += h.nslots;
    }
    all.tgt_off.push_back((int)all.tgt_slot.size());
    all.nslots = S;
    {
        std::string e;
        const int st = rh::validate_level(all, T.ngridmax, T.ncell, e);
        if (st != RH_OK) return fail(c, st, "ensemble: " + e);
        for (size_t p = 0; p < obox.size(); ++p)
            if (obox[p] < 0 || obox[p] >= nbox) return fail(c, RH_ERR_TOPOLOGY, "ensemble: oct box out of range");
    }

    DevLevel& L = c->ens_level;
    free_level(L);
    rh::LevelDev& d = L.dev;
    d.noct = all.noct;
    d.nrec = (int)all.rec_n.size();
    d.nslots = all.nslots;
    d.ntgt = (int)all.tgt_cell.size();
    d.nleaf = all.nleaf;
    std::vector<int4> rec(d.nrec);
    for (int r = 0; r < d.nrec; ++r)
        rec[r] = make_int4(all.rec_n[r], all.rec_nl[r], all.rec_nr[r], all.rec_flags[r]);
    RH_HIP(c, upload(L, d.grid, all.grid, c->stream));
    RH_HIP(c, upload(L, d.leaf, all.leaf, c->stream));
    RH_HIP(c, upload(L, d.nbo, all.nbo, c->stream));
    RH_HIP(c, upload(L, d.rec, rec, c->stream));
    RH_HIP(c, upload(L, d.rec_slot, all.rec_slot, c->stream));
    RH_HIP(c, upload(L, d.tgt_cell, all.tgt_cell, c->stream));
    RH_HIP(c, upload(L, d.tgt_off, all.tgt_off, c->stream));
    RH_HIP(c, upload(L, d.tgt_slot, all.tgt_slot, c->stream));
    RH_HIP(c, upload(L, d.obox, obox, c->stream));
    if (d.nslots > 0) {
        void* p = nullptr;
        RH_HIP(c, hipMalloc(&p, (size_t)d.nslots * 5 * sizeof(double)));
        L.allocs.push_back(p);
        d.slot_r = (double*)p;
    }
    d.nbox = nbox;
    d.box_len = c->d_boxlen;
    RH_HIP(c, hipStreamSynchronize(c->stream));  // host vectors go out of scope
    L.epoch = 0;
    c->ens_levels.assign(levels, levels + nbox);
    return RH_OK;
}

}  // namespace

namespace rhi {

int ens_upload_state(rh_ctx* c, int what) {
    int st;
    if ((what & RH_UOLD) && (st = transfer(c, c->d_uold, true, true))) return st;
    if ((what & RH_UNEW) && (st = transfer(c, c->d_unew, true, false))) return st;
    return RH_OK;
}

int ens_download_state(rh_ctx* c, int what) {
    int st;
    if ((what & RH_UOLD) && (st = transfer(c, c->d_uold, false, true))) return st;
    if ((what & RH_UNEW) && (st = transfer(c, c->d_unew, false, false))) return st;
    RH_HIP(c, hipStreamSynchronize(c->stream));
    return RH_OK;
}

}  // namespace rhi

extern "C" {

int rh_bind_ensemble(rh_ctx* c, int nbox, const rh_mesh* meshes, int mode) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        if (nbox < 1 || !meshes) return fail(c, RH_ERR_ARG, "an ensemble needs at least one mesh");
        if (mode != RH_HOST_COHERENT && mode != RH_DEVICE_RESIDENT) return fail(c, RH_ERR_ARG, "unknown mode");
        long long nc = 0, ng = 0;
        int nlevelmax = 1;
        for (int b = 0; b < nbox; ++b) {
            int st = check_box(c, meshes[b], b);
            if (st) return st;
            nc += meshes[b].ncoarse;
            ng += meshes[b].ngridmax;
            if (meshes[b].nlevelmax > nlevelmax) nlevelmax = meshes[b].nlevelmax;
        }
        const long long ncell = nc + 8 * ng;
        if (ncell > INT_MAX) return fail(c, RH_ERR_ARG, "ensemble has more than 2^31-1 cells");
        RH_HIP(c, hipSetDevice(c->p.device));
        unbind(c);
        c->boxes.resize(nbox);
        long long coff = 0, goff = 0;
        for (int b = 0; b < nbox; ++b) {
            c->boxes[b].mesh = meshes[b];
            c->boxes[b].coff = coff;
            c->boxes[b].goff = goff;
            coff += meshes[b].ncoarse;
            goff += meshes[b].ngridmax;
        }
        rh_mesh& T = c->mesh;
        T = rh_mesh{};
        T.ncoarse = (int)nc;
        T.ngridmax = (int)ng;
        T.ncell = (int)ncell;
        T.nvar = 5;
        T.nlevelmax = nlevelmax;
        const size_t n = (size_t)5 * ncell;
        RH_HIP(c, hipMalloc(&c->d_uold, n * sizeof(double)));
        RH_HIP(c, hipMalloc(&c->d_unew, n * sizeof(double)));
        c->state_cap = n;
        RH_HIP(c, hipMalloc(&c->d_boxlen, (size_t)2 * nbox * sizeof(double)));
        c->ens = true;
        c->bound = true;
        c->mode = mode;
        int st = ens_upload_state(c, RH_UOLD | RH_UNEW);
        if (st) return st;
        RH_HIP(c, hipStreamSynchronize(c->stream));
        return RH_OK;
    });
}

int rh_run_ensemble(rh_ctx* c, const int* levels, const double* dx, double courant_factor, int nsteps,
                    double* dt_out) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        if (!c->bound || !c->ens) return fail(c, RH_ERR_ARG, "no ensemble bound (call rh_bind_ensemble first)");
        if (!levels || !dx) return fail(c, RH_ERR_ARG, "null levels or dx");
        if (nsteps < 0) return fail(c, RH_ERR_ARG, "nsteps < 0");
        const int nbox = (int)c->boxes.size();
        if ((long long)nbox * (nsteps > 0 ? nsteps : 1) > INT_MAX / 2) return fail(c, RH_ERR_ARG, "too many steps");
        std::vector<double> len(2 * (size_t)nbox);
        bool all_exact = true;
        for (int b = 0; b < nbox; ++b) {
            if (levels[b] < 1 || levels[b] > c->boxes[b].mesh.nlevelmax)
                return fail(c, RH_ERR_ARG, "box " + std::to_string(b) + ": level out of range");
            if (!(dx[b] > 0.0)) return fail(c, RH_ERR_ARG, "box " + std::to_string(b) + ": dx must be positive");
            len[b] = dx[b];
            len[nbox + b] = exact_inverse(dx[b]);
            all_exact = all_exact && len[nbox + b] != 0.0;
        }
        RH_HIP(c, hipSetDevice(c->p.device));
        int st;
        if (c->mode == RH_HOST_COHERENT && (st = ens_upload_state(c, RH_UOLD | RH_UNEW))) return st;
        RH_HIP(c, hipMemcpyAsync(c->d_boxlen, len.data(), len.size() * sizeof(double), hipMemcpyHostToDevice,
                                 c->stream));
        if (c->ens_levels.size() != (size_t)nbox || !std::equal(c->ens_levels.begin(), c->ens_levels.end(), levels))
            if ((st = build_tables(c, levels))) return st;
        const rh::LevelDev& lv = c->ens_level.dev;
        const int nk = nbox * (nsteps > 0 ? nsteps : 1);
        if ((st = ensure_dt(c, nk))) return st;
        if (nk > c->nkeys) {
            if (c->d_keys) (void)hipFree(c->d_keys);
            c->d_keys = nullptr;
            c->nkeys = 0;
            RH_HIP(c, hipMalloc(&c->d_keys, (size_t)nk * sizeof(unsigned long long)));
            c->nkeys = nk;
        }
        const rh::MeshDev m = mesh_dev(c);
        const rh::GasParams gp = gas(c, c->p.gamma);
        const double dt_init = c->p.boxlen / 1e-10;
        // As in rh_run_uniform: boxes whose levels hold only leaf cells with
        // same-level neighbours step with uold and unew exchanged, each Godunov
        // launch taking the next step's per-box Courant keys from the values it
        // writes; a final copy leaves the last state in both buffers.  Those
        // launches are replayed from a captured graph (run_graph).
        const int var = variant_for(c, lv);
        const bool exchanged = var != 0 && lv.ntgt == 0 && lv.nrec == 0 && lv.nleaf == 8LL * lv.noct &&
                               lv.noct > 0;
        if (exchanged && nsteps > 0) {
            auto issue = [&]() -> int {
                RH_HIP(c, rh::launch_fill_keys(c->d_keys, nk, dt_init, c->stream));
                rh::MeshDev mk = m;
                for (int k = 0; k < nsteps; ++k) {
                    unsigned long long* keys = c->d_keys + (size_t)k * nbox;
                    double* dts = c->d_dt + (size_t)k * nbox;
                    const bool next = k + 1 < nsteps;
                    if (k == 0) {
                        Timed tm(c, TK_COURANT);
                        RH_HIP(c, rh::launch_courant_boxes(mk, lv, c->p.gamma, courant_factor, keys, c->stream));
                        RH_HIP(c, rh::launch_keys_to_dt(keys, nbox, dts, c->stream));
                        tm.done();
                    }
                    rh::FusedCourant fc;
                    fc.pre_from_uold = true;
                    fc.keys = next ? keys + nbox : nullptr;
                    fc.courant_factor = courant_factor;
                    {
                        Timed tm(c, TK_GODUNOV);
                        hipError_t e = rh::launch_godunov(mk, lv, gp, c->p.riemann, c->p.precision, var, dts,
                                                          0.0, all_exact ? 1.0 : 0.0, c->stream, fc);
                        if (e == hipErrorNotSupported)
                            return fail(c, RH_ERR_UNSUPPORTED, "riemann/precision combination not supported");
                        if (e != hipSuccess) return hip_fail(c, e, "godunov launch");
                        tm.done();
                    }
                    if (next) {
                        Timed tm(c, TK_COURANT);
                        RH_HIP(c, rh::launch_keys_to_dt(keys + nbox, nbox, dts + nbox, c->stream));
                        tm.done();
                    }
                    std::swap(mk.uold, mk.unew);
                }
                Timed tm(c, TK_SET_UOLD);
                RH_HIP(c, rh::launch_set_unew(mk, lv, c->stream));
                tm.done();
                return RH_OK;
            };
            GraphKey key;
            key_add(key, lv);
            key_add(key, m);
            key_add(key, gp);
            key.add(c->p.riemann).add(c->p.precision).add(var).add(all_exact).add(courant_factor);
            key.add(dt_init).add(nsteps).add(nbox).add(nk).add(c->d_keys).add(c->d_dt);
            if ((st = run_graph(c, c->g_ensemble, key, issue))) return st;
        } else {
            RH_HIP(c, rh::launch_fill_keys(c->d_keys, nk, dt_init, c->stream));
        }
        // Otherwise set_uold of one step and the Courant step and set_unew of the
        // next run as one fused pass.
        for (int k = 0; k < (exchanged ? 0 : nsteps); ++k) {
            unsigned long long* keys = c->d_keys + (size_t)k * nbox;
            double* dts = c->d_dt + (size_t)k * nbox;
            if (k == 0) {
                {
                    Timed tm(c, TK_COURANT);
                    RH_HIP(c, rh::launch_courant_boxes(m, lv, c->p.gamma, courant_factor, keys, c->stream));
                    RH_HIP(c, rh::launch_keys_to_dt(keys, nbox, dts, c->stream));
                    tm.done();
                }
                Timed tm(c, TK_SET_UNEW);
                RH_HIP(c, rh::launch_set_unew(m, lv, c->stream));
                tm.done();
            }
            {
                // set_unew, or the join pass, has just made unew equal uold.
                rh::FusedCourant fc;
                fc.pre_from_uold = var != 0;
                Timed tm(c, TK_GODUNOV);
                hipError_t e = rh::launch_godunov(m, lv, gp, c->p.riemann, c->p.precision, var, dts, 0.0,
                                                  all_exact ? 1.0 : 0.0, c->stream, fc);
                if (e == hipErrorNotSupported)
                    return fail(c, RH_ERR_UNSUPPORTED, "riemann/precision combination not supported");
                if (e != hipSuccess) return hip_fail(c, e, "godunov launch");
                tm.done();
            }
            if (k + 1 < nsteps) {
                unsigned long long* next_keys = keys + nbox;
                Timed tm(c, TK_STEP_JOIN);
                RH_HIP(c, rh::launch_courant_boxes(m, lv, c->p.gamma, courant_factor, next_keys, c->stream, true));
                RH_HIP(c, rh::launch_keys_to_dt(next_keys, nbox, dts + nbox, c->stream));
                tm.done();
            } else {
                Timed tm(c, TK_SET_UOLD);
                RH_HIP(c, rh::launch_set_uold(m, lv, c->stream));
                tm.done();
            }
        }
        if (dt_out && nsteps > 0)
            RH_HIP(c, hipMemcpyAsync(dt_out, c->d_dt, (size_t)nsteps * nbox * sizeof(double), hipMemcpyDeviceToHost,
                                     c->stream));
        RH_HIP(c, hipStreamSynchronize(c->stream));
        if (c->mode == RH_HOST_COHERENT) return ens_download_state(c, RH_UOLD | RH_UNEW);
        return RH_OK;
    });
}

int rh_ensemble_leaf_cells(rh_ctx* c, const int* levels, long long* nleaf) {
    return guard(c ? &c->err : &g_create_error, nullptr, [&]() -> int {
        if (!c) return RH_ERR_ARG;
        if (!c->bound || !c->ens) return fail(c, RH_ERR_ARG, "no ensemble bound (call rh_bind_ensemble first)");
        if (!levels) return fail(c, RH_ERR_ARG, "null levels");
        const int nbox = (int)c->boxes.size();
        for (int b = 0; b < nbox; ++b)
            if (levels[b] < 1 || levels[b] > c->boxes[b].mesh.nlevelmax)
                return fail(c, RH_ERR_ARG, "box " + std::to_string(b) + ": level out of range");
        if (c->ens_levels.size() != (size_t)nbox || !std::equal(c->ens_levels.begin(), c->ens_levels.end(), levels)) {
            int st = build_tables(c, levels);
            if (st) return st;
        }
        if (nleaf) *nleaf = c->ens_level.dev.nleaf;
        return RH_OK;
    });
}

}  // extern "C"
