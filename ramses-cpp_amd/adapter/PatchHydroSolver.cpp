// PatchHydroSolver.cpp -- drop-in HydroSolver for RAMSES-CPP's patch seam.
//
// Copy this file into the reference's patch/ directory (globbed into the
// library by CMakeLists.txt:47-53), add -DRAMSES_HAS_PATCH_HYDRO so the
// default factory (SolverFactory.cpp:5-9) steps aside, and link librhydro.so.
// Every hydro virtual of HydroSolver.hpp:19-29 except get_diagnostics (an
// empty function in the reference) is forwarded to the matching C-ABI call of
// include/rhydro.h.
//
// The reference mutates the AmrGrid on the host between calls (refinement,
// restriction, flagging), so an adaptive run's context runs in
// RH_HOST_COHERENT mode: state is copied in before and out after each call,
// and the mesh is re-bound whenever the topology arrays change (detected by
// comparing them with the copy taken at the last binding, and by pointer).
// A uniform run (levelmin == levelmax) with no host-side physics between the
// hydro calls keeps the state on the GPU instead: nothing on the host reads
// or writes uold/unew between those calls there (amr_step neither refines,
// restricts nor flags), so the context is device-resident, the topology is
// bound once, and only uold comes back, once per step after set_uold, for
// the snapshots and diagnostics that read it.
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "ramses/core/Config.hpp"
#include "ramses/core/MpiManager.hpp"
#include "ramses/core/SolverFactory.hpp"
#include "ramses/solvers/hydro/HydroSolver.hpp"
#include "rhydro.h"

namespace ramses {

class PatchHydroSolver : public HydroSolver {
public:
    PatchHydroSolver(AmrGrid& grid, Config& config) : HydroSolver(grid, config) {}
    ~PatchHydroSolver() override {
        if (ctx_) rh_destroy(ctx_);
    }

    void godunov_fine(int ilevel, real_t dt, real_t dx) override {
        ensure_bound();
        check(rh_godunov_fine(ctx_, ilevel, dt, dx), "rh_godunov_fine");
    }

    void set_uold(int ilevel) override {
        ensure_bound();
        check(rh_set_uold(ctx_, ilevel), "rh_set_uold");
        if (resident_) check(rh_sync(ctx_, RH_SYNC_TO_HOST, RH_UOLD), "rh_sync");
    }

    void set_unew(int ilevel) override {
        ensure_bound();
        check(rh_set_unew(ctx_, ilevel), "rh_set_unew");
    }

    real_t compute_courant_step(int ilevel, real_t dx, real_t gamma, real_t courant_factor) override {
        ensure_bound();
        double dt = 0.0;
        check(rh_courant(ctx_, ilevel, dx, gamma, courant_factor, &dt), "rh_courant");
        return dt;
    }

    void add_gravity_source_terms(int ilevel, real_t dt) override {
        ensure_bound();
        check(rh_add_gravity_source_terms(ctx_, ilevel, dt), "rh_add_gravity_source_terms");
    }

    void synchro_hydro_fine(int ilevel, real_t dt) override {
        ensure_bound();
        check(rh_synchro_hydro_fine(ctx_, ilevel, dt), "rh_synchro_hydro_fine");
    }

    void interpol_hydro(const real_t u1[7][64], real_t u2[8][64]) override {
        ensure_context();
        check(rh_interpol_hydro(ctx_, 1, &u1[0][0], &u2[0][0]), "rh_interpol_hydro");
    }

private:
    rh_ctx* ctx_ = nullptr;
    long long epoch_ = 0;
    bool resident_ = false;  // uniform run: state kept on the GPU
    // Topology arrays and state pointers at the last binding.
    std::vector<int> topo_copy_[5];
    const void* ptrs_[8] = {};

    void check(int st, const char* what) {
        if (st != RH_OK) throw std::runtime_error(std::string(what) + ": " + rh_last_error(ctx_));
    }

    // The namelist is parsed after the factory builds the solver
    // (Simulation.cpp:27 then :43), so parameters are read on first use,
    // with the reference's keys and defaults.
    void ensure_context() {
        if (ctx_) return;
        rh_params p{};
        p.gamma = grid_.gamma;
        p.slope_type = config_.get_int("hydro_params", "slope_type", 1);
        const std::string r = config_.get("hydro_params", "riemann", "llf");  // HydroSolver.cpp:456
        p.riemann = r == "hllc" ? RH_RIEMANN_HLLC
                    : r == "hll" ? RH_RIEMANN_HLL
                    : r == "exact" ? RH_RIEMANN_EXACT
                    : r == "acoustic" ? RH_RIEMANN_ACOUSTIC
                                      : RH_RIEMANN_LLF;
        p.interpol_type = config_.get_int("refine_params", "interpol_type", 1);
        p.interpol_var = config_.get_int("refine_params", "interpol_var",
                                         config_.get_int("hydro_params", "interpol_var", 0));
        p.nvar = grid_.nvar;
        p.nener = nener_;
        p.barotropic_eos = config_.get_bool("cooling_params", "barotropic_eos", false) ? 1 : 0;
        p.boxlen = config_.get_double("amr_params", "boxlen", 1.0);
        p.precision = RH_PREC_FP64;
        p.device = 0;
        check(rh_create(&p, &ctx_), "rh_create");
        // Host code between the hydro calls of amr_step (Simulation.cpp:
        // 522-635): refinement, restriction and flagging only when levelmin <
        // levelmax; cooling, turbulence forcing, sinks and stars when enabled.
        const bool uniform = config_.get_int("amr_params", "levelmin", 1) == config_.get_int("amr_params", "levelmax", 1);
        const bool host_physics = config_.get_bool("cooling_params", "cooling", false) ||
                                  config_.get_bool("run_params", "turb", false) ||
                                  config_.get_bool("run_params", "sink", false) ||
                                  config_.get_bool("run_params", "star", false);
        resident_ = uniform && !host_physics;
    }

    // True when the topology or any array pointer differs from the last
    // binding (then the copy is refreshed).  memcmp of the saved arrays: an
    // exact test that streams them at memory bandwidth.
    bool topology_changed() {
        const std::vector<int>* arrs[5] = {&grid_.father, &grid_.son, &grid_.nbor, &grid_.next, &grid_.headl_vec};
        const void* ptrs[8] = {grid_.father.data(), grid_.son.data(),      grid_.nbor.data(),     grid_.next.data(),
                               grid_.headl_vec.data(), grid_.uold_vec.data(), grid_.unew_vec.data(), grid_.f_vec.data()};
        bool changed = epoch_ == 0;
        for (int i = 0; i < 8 && !changed; ++i) changed = ptrs[i] != ptrs_[i];
        for (int i = 0; i < 5 && !changed; ++i)
            changed = arrs[i]->size() != topo_copy_[i].size() ||
                      std::memcmp(arrs[i]->data(), topo_copy_[i].data(), arrs[i]->size() * sizeof(int)) != 0;
        if (!changed) return false;
        for (int i = 0; i < 5; ++i) topo_copy_[i] = *arrs[i];
        for (int i = 0; i < 8; ++i) ptrs_[i] = ptrs[i];
        return true;
    }

    void ensure_bound() {
        ensure_context();
        // A device-resident (uniform) run never changes its topology: it is
        // bound once.
        if (resident_) {
            if (epoch_ > 0) return;
        } else if (!topology_changed()) {
            return;
        }
        ++epoch_;
        rh_mesh m{};
        m.ncoarse = grid_.ncoarse;
        m.ngridmax = grid_.ngridmax;
        m.ncell = grid_.ncell;
        m.nvar = grid_.nvar;
        m.nx = grid_.nx;
        m.ny = grid_.ny;
        m.nz = grid_.nz;
        m.nlevelmax = grid_.nlevelmax;
        m.ncpu = grid_.ncpu;
        m.nboundary = grid_.nboundary;
        m.myid = MpiManager::instance().rank() + 1;
        m.father = grid_.father.data();
        m.son = grid_.son.data();
        m.nbor = grid_.nbor.data();
        m.next = grid_.next.data();
        m.headl = grid_.headl_vec.data();
        m.uold = grid_.uold_vec.data();
        m.unew = grid_.unew_vec.data();
        m.f = grid_.f_vec.size() == (size_t)3 * grid_.ncell ? grid_.f_vec.data() : nullptr;
        m.mesh_epoch = epoch_;
        check(rh_bind(ctx_, &m, resident_ ? RH_DEVICE_RESIDENT : RH_HOST_COHERENT), "rh_bind");
    }
};

std::unique_ptr<HydroSolver> create_hydro_solver(AmrGrid& grid, Config& config) {
    return std::make_unique<PatchHydroSolver>(grid, config);
}

}  // namespace ramses
