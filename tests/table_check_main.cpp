// Host-only check of the level-table builder and its range check
// (ramses-cpp_amd/csrc/rh_tables.cpp), compiled and run by
// tests/test_table_check.py with the host compiler; no GPU involved.
//
// Mesh: one coarse cell (nx=ny=nz=1), its oct (grid 1, level list 2) and one
// refined cell of that oct (grid 2, level list 3), linked as the reference's
// make_grid_fine links a new oct (TreeUpdater.cpp:27-150): father = the
// refined cell, nbor = that cell's six face neighbours.  Level 3 then has six
// faces onto coarser leaf cells (reflux records, slots and target lists) and
// level 2 one refined cell.  The valid tables must pass the range check; each
// planted corruption of an index the kernels follow must fail it.
#include <cstdio>
#include <functional>
#include <string>
#include <vector>

#include "rh_tables.h"

namespace rhi {
std::string g_create_error;  // defined by the library's C-ABI file otherwise
}

static int failures = 0;

static void expect(bool ok, const std::string& what) {
    std::printf("%s %s\n", ok ? "ok  " : "FAIL", what.c_str());
    if (!ok) ++failures;
}

int main() {
    const int ngridmax = 4, ncoarse = 1, ncell = ncoarse + 8 * ngridmax, nlevelmax = 3;
    std::vector<int> father(ngridmax, 0), son(ncell, 0), nbor(6 * ngridmax, 0), next(ngridmax, 0),
        headl(nlevelmax + 1, 0);
    std::vector<double> uold(5 * ncell, 1.0), unew(5 * ncell, 1.0);
    rh_mesh m{};
    m.ncoarse = ncoarse;
    m.ngridmax = ngridmax;
    m.ncell = ncell;
    m.nvar = 5;
    m.nx = m.ny = m.nz = 1;
    m.nlevelmax = nlevelmax;
    m.ncpu = 1;
    m.nboundary = 0;
    m.myid = 1;
    m.father = father.data();
    m.son = son.data();
    m.nbor = nbor.data();
    m.next = next.data();
    m.headl = headl.data();
    m.uold = uold.data();
    m.unew = unew.data();
    m.mesh_epoch = 1;
    // Grid 1: the coarse cell's oct; every face neighbour is coarse cell 1
    // (periodic wrap of a one-cell coarse grid).
    son[0] = 1;
    father[0] = 1;
    for (int f = 0; f < 6; ++f) nbor[f * ngridmax + 0] = 1;
    headl[2] = 1;
    // Grid 2: child of cell ic=1 of grid 1 (cell id ncoarse + 0*ngridmax + 1).
    const int refined = ncoarse + 0 * ngridmax + 1;
    son[refined - 1] = 2;
    father[1] = refined;
    int icn[6];
    rh::nbor_cells_of_cell(m, refined, icn);
    for (int f = 0; f < 6; ++f) nbor[f * ngridmax + 1] = icn[f];
    headl[3] = 2;

    rh::TopoIndex ti;
    std::string e;
    expect(rh::build_topo_index(m, ti, e) == RH_OK, "topology index: " + e);
    rh::LevelHost l2, l3;
    expect(rh::build_level(m, ti, 2, l2, e) == RH_OK, "level 2 tables: " + e);
    expect(rh::build_level(m, ti, 3, l3, e) == RH_OK, "level 3 tables: " + e);
    expect(rh::validate_level(l2, ngridmax, ncell, e) == RH_OK, "level 2 passes the range check");
    expect(rh::validate_level(l3, ngridmax, ncell, e) == RH_OK, "level 3 passes the range check");
    expect(l3.rec_n.size() == 6 && l3.nslots == 24 && !l3.tgt_cell.empty(), "level 3 has six reflux records");
    expect(l2.nleaf == 7, "level 2 has seven leaf cells");

    using Plant = std::function<void(rh::LevelHost&)>;
    const std::vector<std::pair<std::string, Plant>> plants = {
        {"grid beyond ngridmax", [&](rh::LevelHost& h) { h.grid[0] = ngridmax + 1; }},
        {"grid 0", [](rh::LevelHost& h) { h.grid[0] = 0; }},
        {"nbo beyond the oct list", [](rh::LevelHost& h) { h.nbo[0] = h.noct; }},
        {"nbo record code beyond the records", [](rh::LevelHost& h) { h.nbo[0] = -1 - (int)h.rec_n.size(); }},
        {"record cell beyond ncell", [&](rh::LevelHost& h) { h.rec_n[0] = ncell + 1; }},
        {"reflux record cell 0", [](rh::LevelHost& h) { h.rec_n[0] = 0; }},
        {"record nl beyond ncell", [&](rh::LevelHost& h) { h.rec_nl[1] = ncell + 1; }},
        {"record nr beyond ncell", [&](rh::LevelHost& h) { h.rec_nr[2] = ncell + 8; }},
        {"unknown record flag", [](rh::LevelHost& h) { h.rec_flags[0] = 3; }},
        {"reflux slot past nslots", [](rh::LevelHost& h) { h.rec_slot[5] = h.nslots - 2; }},
        {"negative reflux slot", [](rh::LevelHost& h) { h.rec_slot[0] = -1; }},
        {"target cell beyond ncell", [&](rh::LevelHost& h) { h.tgt_cell[0] = ncell + 1; }},
        {"target slot past nslots", [](rh::LevelHost& h) { h.tgt_slot.back() = h.nslots; }},
        {"target offsets out of order", [](rh::LevelHost& h) {
             if (h.tgt_off.size() > 2) std::swap(h.tgt_off[1], h.tgt_off[h.tgt_off.size() - 2]);
             h.tgt_off[1] = h.tgt_off.back() + 1;
         }},
        {"leaf count mismatch", [](rh::LevelHost& h) { h.nleaf += 1; }},
        {"short nbo table", [](rh::LevelHost& h) { h.nbo.pop_back(); }},
    };
    for (const auto& pl : plants) {
        rh::LevelHost h = l3;
        pl.second(h);
        std::string err;
        const int st = rh::validate_level(h, ngridmax, ncell, err);
        expect(st == RH_ERR_TOPOLOGY && !err.empty(), "planted " + pl.first + " is rejected (" + err + ")");
    }
    // Tables checked against one mesh size do not pass for a smaller mesh.
    expect(rh::validate_level(l3, 1, ncoarse + 8, e) == RH_ERR_TOPOLOGY, "tables of a larger mesh are rejected");
    std::printf("%d failures\n", failures);
    return failures == 0 ? 0 : 1;
}
