// rh_tables.h -- per-level index tables of the Godunov path, built on the host
// from the reference's 1-based AmrGrid topology arrays once per mesh epoch.
//
// For the octs of one level list (in list order, the reference's processing
// order, HydroSolver.cpp:43-47) the tables hold, per oct position:
//   grid[pos]        the oct's grid index igrid (1-based)
//   leaf[pos]        bit (ic-1) set when cell ic has no son
//   nbo[d*noct+pos]  d = 2*idim+side: position (same list) of the oct holding
//                    the face neighbours in that direction, or -1-r where r
//                    indexes a record for a neighbour that is not an oct of
//                    this level (coarser leaf cell, refined cell, boundary)
// and per record r: the neighbour cell n (AmrGrid.cpp:402-444 semantics), n's
// own neighbours nl/nr along the face's dimension (used to interpolate n's
// ghost state, HydroSolver.cpp:178-191) and flags.  Records whose neighbour is
// a coarser leaf own 4 consecutive "coarse reflux slots" (one per cell on the
// face); the coarse-target lists order those slots per target cell by the
// reference's processing key (pos*8+ic-1)*8 + face.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

#include "rhydro.h"

namespace rh {

enum : int {
    REC_REFINED = 1,   // neighbour has a son (or is finer): zero flux, no reflux
    REC_BOUNDARY = 2,  // neighbour index <= 0: reflective boundary branch
    REC_REFLUX = 4     // neighbour is a coarser leaf: reflux into it
};

struct TopoIndex {
    long long epoch = -1;
    std::vector<int> glev;   // list index of each grid (0 = in no list)
    std::vector<int> gpos;   // position of each grid within its list
    std::vector<int> depth;  // father-chain depth = the reference's cell_levels value
    std::vector<int> listed; // every grid of the level lists, list by list
    std::vector<int> touched;  // grids whose entries above the build wrote
};

struct LevelHost {
    int level = 0;
    int noct = 0;
    std::vector<int> grid;
    std::vector<uint8_t> leaf;
    std::vector<int> nbo;
    std::vector<int> rec_n, rec_nl, rec_nr, rec_flags, rec_slot;
    int nslots = 0;
    std::vector<int> tgt_cell, tgt_off, tgt_slot;
    long long nleaf = 0;  // leaf cells of the level (cell-updates per call)
};

// Builds glev/gpos/depth for every list.  Returns RH_OK or an error code
// with a message in err.
int build_topo_index(const rh_mesh& m, TopoIndex& ti, std::string& err);

// Builds the tables of one level.  Returns RH_OK, RH_ERR_TOPOLOGY (mesh the
// reference cannot produce: inconsistent neighbours, an oct of the list whose
// depth is not level-1, a neighbour reaching the reference's dead same-level
// branch) with a message in err.
int build_level(const rh_mesh& m, const TopoIndex& ti, int level, LevelHost& out, std::string& err);

// Checks every index the kernels follow from a level's tables before they
// are uploaded: grid ids within [1, ngridmax]; each nbo entry an oct position
// of the level or a record code; record cells at most ncell (a boundary code
// <= 0 only where the kernels read it as one); reflux slots inside nslots;
// target lists well formed.  A table that fails never reaches a launch:
// RH_ERR_TOPOLOGY with the first offending entry in err.
int validate_level(const LevelHost& h, int ngridmax, long long ncell, std::string& err);

// Neighbour cells of a cell (1-based), as AmrGrid::get_nbor_cells /
// get_nbor_cells_coarse return them.
void nbor_cells_of_cell(const rh_mesh& m, int icell, int icn[6]);

}  // namespace rh
