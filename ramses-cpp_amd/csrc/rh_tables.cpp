// rh_tables.cpp -- host construction of the per-level index tables (see
// rh_tables.h).  Index preprocessing only; no state arithmetic happens here.
#include "rh_tables.h"

#include <algorithm>
#include <cstdio>
#include <numeric>

namespace rh {

namespace {

// Octant geometry of the reference (Constants.cpp:8-43): the neighbour of
// cell ic (0-based ic0 = ix + 2*iy + 4*iz) across dimension d lies in the
// same oct when bit d of ic0 differs from the side (0 = minus, 1 = plus),
// otherwise in the neighbouring oct in direction 2*d+side, and in both cases
// at cell ic0 ^ (1<<d).
inline bool crosses(int ic0, int d, int side) { return ((ic0 >> d) & 1) == side; }

struct MeshView {
    const rh_mesh& m;
    int headl(int il) const {
        if (il < 0 || il > m.nlevelmax) return 0;
        return m.headl[(size_t)il * (m.ncpu + m.nboundary) + (m.myid - 1)];
    }
    int nbor(int dir, int igrid) const {
        int v = m.nbor[(size_t)dir * m.ngridmax + igrid - 1];
        return (v > m.ncell) ? 0 : v;  // AmrGrid.cpp:394-396
    }
    int cell(int igrid, int ic0) const { return m.ncoarse + ic0 * m.ngridmax + igrid; }
    int grid_of(int icell) const { return ((icell - m.ncoarse - 1) % m.ngridmax) + 1; }
    int ic0_of(int icell) const { return (icell - m.ncoarse - 1) / m.ngridmax; }
};

// AmrGrid.cpp:513-550 with no boundary regions: periodic wrap.
void coarse_nbors(const rh_mesh& m, int icell, int icn[6]) {
    for (int i = 0; i < 6; ++i) icn[i] = 0;
    if (icell <= 0 || icell > m.ncoarse) return;
    int idx = icell - 1;
    int nx = m.nx, ny = m.ny, nz = m.nz;
    int x[3] = {idx % nx, (idx % (nx * ny)) / nx, idx / (nx * ny)};
    int n[3] = {nx, ny, nz};
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s) {
            int c[3] = {x[0], x[1], x[2]};
            c[d] = (c[d] + (s == 0 ? -1 : 1) + n[d]) % n[d];
            icn[2 * d + s] = c[2] * nx * ny + c[1] * nx + c[0] + 1;
        }
}

// AmrGrid.cpp:402-444 for cell (igrid, ic0).
void oct_cell_nbors(const rh_mesh& m, int igrid, int ic0, int icn[6]) {
    MeshView v{m};
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s) {
            int f = 2 * d + s;
            int ic_pos = ic0 ^ (1 << d);
            if (!crosses(ic0, d, s)) {
                icn[f] = v.cell(igrid, ic_pos);
                continue;
            }
            int ig = v.nbor(f, igrid);
            if (ig > 0) {
                int child = (ig <= m.ncell) ? m.son[ig - 1] : 0;
                icn[f] = child > 0 ? v.cell(child, ic_pos) : ig;
            } else if (ig < 0) {
                icn[f] = ig;
            } else {
                int ifather = m.father[igrid - 1];
                int fn[6];
                nbor_cells_of_cell(m, ifather, fn);
                icn[f] = fn[f];
            }
        }
}

}  // namespace

void nbor_cells_of_cell(const rh_mesh& m, int icell, int icn[6]) {
    if (icell <= m.ncoarse) {
        coarse_nbors(m, icell, icn);
        return;
    }
    MeshView v{m};
    oct_cell_nbors(m, v.grid_of(icell), v.ic0_of(icell), icn);
}

int build_topo_index(const rh_mesh& m, TopoIndex& ti, std::string& err) {
    MeshView v{m};
    // Every index the table builder and the kernels follow from the level
    // lists must stay inside its array: the coarse cells' sons, and for every
    // grid in a list its father, its list link and its cells' sons.  (Grids
    // in no list are never reached from one.)
    auto son_ok = [&](int icell) {
        const int s = m.son[icell - 1];
        if (s >= 0 && s <= m.ngridmax) return true;
        err = "son[" + std::to_string(icell) + "] = " + std::to_string(s) + " outside 0..ngridmax";
        return false;
    };
    for (int i = 1; i <= m.ncoarse; ++i)
        if (!son_ok(i)) return RH_ERR_TOPOLOGY;
    // The index arrays keep their size between builds for one ngridmax, and
    // only the entries the previous build wrote are cleared.
    if (ti.glev.size() != (size_t)m.ngridmax + 1) {
        ti.glev.assign((size_t)m.ngridmax + 1, 0);
        ti.gpos.assign((size_t)m.ngridmax + 1, -1);
        ti.depth.assign((size_t)m.ngridmax + 1, 0);
    } else {
        for (const int g : ti.touched) {
            ti.glev[g] = 0;
            ti.gpos[g] = -1;
            ti.depth[g] = 0;
        }
    }
    ti.touched.clear();
    ti.epoch = -1;
    std::vector<int>& listed = ti.listed;
    listed.clear();
    for (int il = 1; il <= m.nlevelmax; ++il) {
        int ig = v.headl(il);
        int pos = 0;
        while (ig > 0) {
            if (ig > m.ngridmax) {
                err = "level list " + std::to_string(il) + " holds grid index beyond ngridmax";
                return RH_ERR_TOPOLOGY;
            }
            if (ti.glev[ig] != 0) {
                err = "grid " + std::to_string(ig) + " appears in two level lists (or a list is cyclic)";
                return RH_ERR_TOPOLOGY;
            }
            if (m.father[ig - 1] < 0 || m.father[ig - 1] > m.ncell || m.next[ig - 1] < 0 ||
                m.next[ig - 1] > m.ngridmax) {
                err = "father/next of grid " + std::to_string(ig) + " out of range";
                return RH_ERR_TOPOLOGY;
            }
            for (int ic = 0; ic < 8; ++ic)
                if (!son_ok(m.ncoarse + ic * m.ngridmax + ig)) return RH_ERR_TOPOLOGY;
            ti.glev[ig] = il;
            ti.gpos[ig] = pos++;
            listed.push_back(ig);
            ti.touched.push_back(ig);
            ig = m.next[ig - 1];
        }
    }
    // Father-chain depth, as HydroSolver.cpp:22-38 computes cell_levels: the
    // number of grids on the chain up to the first whose father is a coarse
    // cell (or 0).  Each chain is walked only up to a grid whose depth is
    // already known, so the whole pass is linear in the grids.  Grids with
    // father 0 keep depth 0 (the reference never asks for theirs).
    std::vector<int> chain;
    for (const int ig : listed) {
        if (m.father[ig - 1] <= 0 || ti.depth[ig] != 0) continue;
        chain.clear();
        int cur = ig, base = 0;
        for (;;) {
            chain.push_back(cur);
            const int fc = m.father[cur - 1];
            if (fc <= m.ncoarse) break;
            cur = ((fc - m.ncoarse - 1) % m.ngridmax) + 1;
            if (ti.depth[cur] != 0) {
                base = ti.depth[cur];
                break;
            }
            if ((int)chain.size() > 64) {
                err = "father chain longer than 64 from grid " + std::to_string(ig);
                return RH_ERR_TOPOLOGY;
            }
        }
        if (base + (int)chain.size() > 65) {
            err = "father chain longer than 64 from grid " + std::to_string(ig);
            return RH_ERR_TOPOLOGY;
        }
        for (int k = (int)chain.size() - 1, d = base + 1; k >= 0; --k, ++d)
            if (m.father[chain[k] - 1] > 0) {
                ti.depth[chain[k]] = d;
                if (ti.glev[chain[k]] == 0) ti.touched.push_back(chain[k]);
            }
    }
    ti.epoch = m.mesh_epoch;
    return RH_OK;
}

int build_level(const rh_mesh& m, const TopoIndex& ti, int level, LevelHost& out, std::string& err) {
    MeshView v{m};
    out = LevelHost{};
    out.level = level;
    auto cell_level = [&](int icell) -> int {  // cell_levels[icell] of the reference
        if (icell <= m.ncoarse) return 0;
        return ti.depth[v.grid_of(icell)];
    };
    // Oct list in processing order.
    for (int ig = v.headl(level); ig > 0; ig = m.next[ig - 1]) out.grid.push_back(ig);
    const int noct = (int)out.grid.size();
    out.noct = noct;
    out.leaf.assign(noct, 0);
    out.nbo.assign((size_t)6 * noct, 0);
    for (int p = 0; p < noct; ++p) {
        int g = out.grid[p];
        if (ti.depth[g] != level - 1) {
            err = "oct " + std::to_string(g) + " of level list " + std::to_string(level) + " has depth " +
                  std::to_string(ti.depth[g]) + " (expected level-1)";
            return RH_ERR_TOPOLOGY;
        }
        uint8_t lm = 0;
        for (int ic0 = 0; ic0 < 8; ++ic0)
            if (m.son[v.cell(g, ic0) - 1] == 0) lm |= (uint8_t)(1u << ic0);
        out.leaf[p] = lm;
        out.nleaf += __builtin_popcount(lm);
    }
    // Neighbour octs / records.
    for (int p = 0; p < noct; ++p) {
        int g = out.grid[p];
        for (int f = 0; f < 6; ++f) {
            int d = f >> 1;
            int nb = v.nbor(f, g);
            int n = 0;
            if (nb > 0) {
                int s = m.son[nb - 1];
                if (s > 0) {
                    if (ti.glev[s] != level) {
                        err = "neighbour oct " + std::to_string(s) + " of oct " + std::to_string(g) +
                              " is not in level list " + std::to_string(level);
                        return RH_ERR_TOPOLOGY;
                    }
                    out.nbo[(size_t)f * noct + p] = ti.gpos[s];
                    continue;
                }
                n = nb;
            } else if (nb < 0) {
                n = nb;
            } else {
                int fn[6];
                nbor_cells_of_cell(m, m.father[g - 1], fn);
                n = fn[f];
            }
            int r = (int)out.rec_n.size();
            int flags = 0, nl = 0, nr = 0;
            if (n <= 0) {
                flags = REC_BOUNDARY;
            } else {
                bool refined = m.son[n - 1] > 0 || cell_level(n) > level;
                if (refined) {
                    flags = REC_REFINED;
                } else {
                    if (cell_level(n) == level) {
                        err = "face neighbour " + std::to_string(n) + " of oct " + std::to_string(g) +
                              " reaches the reference's same-level branch (HydroSolver.cpp:148-161)";
                        return RH_ERR_TOPOLOGY;
                    }
                    flags = REC_REFLUX;
                    int icn[6];
                    nbor_cells_of_cell(m, n, icn);
                    nl = icn[2 * d];
                    nr = icn[2 * d + 1];
                }
            }
            out.rec_n.push_back(n);
            out.rec_nl.push_back(nl);
            out.rec_nr.push_back(nr);
            out.rec_flags.push_back(flags);
            int slot = -1;
            if (flags & REC_REFLUX) {
                slot = out.nslots;
                out.nslots += 4;
            }
            out.rec_slot.push_back(slot);
            out.nbo[(size_t)f * noct + p] = -1 - r;
        }
    }
    // Mutual adjacency: the oct across face f of p must see p across the
    // opposite face (true for every mesh the reference's TreeUpdater builds).
    for (int p = 0; p < noct; ++p)
        for (int f = 0; f < 6; ++f) {
            int q = out.nbo[(size_t)f * noct + p];
            if (q < 0) continue;
            int back = out.nbo[(size_t)(f ^ 1) * noct + q];
            if (back != p) {
                err = "octs " + std::to_string(out.grid[p]) + " and " + std::to_string(out.grid[q]) +
                      " are not mutual neighbours across face " + std::to_string(f);
                return RH_ERR_TOPOLOGY;
            }
        }
    // Coarse reflux targets, ordered per target by processing key.
    struct Entry {
        int target;
        long long key;
        int slot;
    };
    std::vector<Entry> ent;
    for (int p = 0; p < noct; ++p)
        for (int f = 0; f < 6; ++f) {
            int q = out.nbo[(size_t)f * noct + p];
            if (q >= 0) continue;
            int r = -1 - q;
            if (!(out.rec_flags[r] & REC_REFLUX)) continue;
            int d = f >> 1, s = f & 1;
            for (int ic0 = 0; ic0 < 8; ++ic0) {
                if (!crosses(ic0, d, s)) continue;
                if (!((out.leaf[p] >> ic0) & 1)) continue;
                int j = ((ic0 >> (d + 1)) << d) | (ic0 & ((1 << d) - 1));
                ent.push_back({out.rec_n[r], ((long long)p * 8 + ic0) * 8 + f, out.rec_slot[r] + j});
            }
        }
    std::sort(ent.begin(), ent.end(), [](const Entry& a, const Entry& b) {
        return a.target != b.target ? a.target < b.target : a.key < b.key;
    });
    for (size_t i = 0; i < ent.size(); ++i) {
        if (i == 0 || ent[i].target != ent[i - 1].target) {
            out.tgt_cell.push_back(ent[i].target);
            out.tgt_off.push_back((int)i);
        }
        out.tgt_slot.push_back(ent[i].slot);
    }
    out.tgt_off.push_back((int)ent.size());
    return RH_OK;
}

int validate_level(const LevelHost& h, int ngridmax, long long ncell, std::string& err) {
    auto bad = [&](const std::string& what, long long i, long long v) {
        err = "level " + std::to_string(h.level) + " table check: " + what + "[" + std::to_string(i) +
              "] = " + std::to_string(v) + " is out of range";
        return RH_ERR_TOPOLOGY;
    };
    const long long noct = h.noct;
    const long long nrec = (long long)h.rec_n.size();
    if ((long long)h.grid.size() != noct || (long long)h.leaf.size() != noct || (long long)h.nbo.size() != 6 * noct)
        return bad("table size", 0, (long long)h.grid.size());
    if ((long long)h.rec_nl.size() != nrec || (long long)h.rec_nr.size() != nrec ||
        (long long)h.rec_flags.size() != nrec || (long long)h.rec_slot.size() != nrec)
        return bad("record size", 0, nrec);
    for (long long p = 0; p < noct; ++p)
        if (h.grid[p] < 1 || h.grid[p] > ngridmax) return bad("grid", p, h.grid[p]);
    for (long long i = 0; i < 6 * noct; ++i) {
        const long long q = h.nbo[i];
        if (q >= noct || (q < 0 && -1 - q >= nrec)) return bad("nbo", i, q);
    }
    long long nleaf = 0;
    for (long long p = 0; p < noct; ++p) nleaf += __builtin_popcount(h.leaf[p]);
    if (nleaf != h.nleaf) return bad("nleaf", 0, h.nleaf);
    for (long long r = 0; r < nrec; ++r) {
        const int fl = h.rec_flags[r];
        if (fl != REC_REFINED && fl != REC_BOUNDARY && fl != REC_REFLUX) return bad("rec_flags", r, fl);
        // A boundary record's cell is the boundary code itself (<= 0); the
        // kernels read the others' cells, and a reflux record's nl/nr, from
        // the state (an nl/nr <= 0 falls back to the neighbour's own state).
        if (fl == REC_BOUNDARY ? h.rec_n[r] > 0 : (h.rec_n[r] < 1 || h.rec_n[r] > ncell))
            return bad("rec_n", r, h.rec_n[r]);
        if (h.rec_nl[r] > ncell) return bad("rec_nl", r, h.rec_nl[r]);
        if (h.rec_nr[r] > ncell) return bad("rec_nr", r, h.rec_nr[r]);
        const int s = h.rec_slot[r];
        if (fl == REC_REFLUX ? (s < 0 || (long long)s + 4 > h.nslots) : s != -1) return bad("rec_slot", r, s);
    }
    const long long ntgt = (long long)h.tgt_cell.size();
    if ((long long)h.tgt_off.size() != ntgt + 1 || h.tgt_off[0] != 0 ||
        h.tgt_off[ntgt] != (long long)h.tgt_slot.size())
        return bad("tgt_off", ntgt, h.tgt_off.empty() ? -1 : h.tgt_off.back());
    for (long long t = 0; t < ntgt; ++t) {
        if (h.tgt_cell[t] < 1 || h.tgt_cell[t] > ncell) return bad("tgt_cell", t, h.tgt_cell[t]);
        if (h.tgt_off[t + 1] < h.tgt_off[t]) return bad("tgt_off", t + 1, h.tgt_off[t + 1]);
    }
    for (size_t i = 0; i < h.tgt_slot.size(); ++i)
        if (h.tgt_slot[i] < 0 || h.tgt_slot[i] >= h.nslots) return bad("tgt_slot", (long long)i, h.tgt_slot[i]);
    return RH_OK;
}

}  // namespace rh
