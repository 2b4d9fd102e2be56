// rh_snapshot.h -- snapshot writer shared by the C-ABI entry points of
// rh_snapshot.cpp and the adaptive driver (rh_amr.cpp).  Internal.
#pragma once
#include "rhydro_snapshot.h"

namespace rhi {

// Simulation::dump_snapshot for the mesh bound to c, described by t.
int dump_snapshot(rh_ctx* c, const rh_snapshot_tree* t, const rh_snapshot_info* in, const char* root);

}  // namespace rhi
