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
// restriction, flagging), so the context runs in RH_HOST_COHERENT mode: state
// is copied in before and out after each call, and the mesh is re-bound
// whenever the topology arrays change (detected by pointer and checksum).
#include <cstdint>
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
    uint64_t topo_sum_ = 0;
    const void* topo_ptrs_[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};

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
    }

    static uint64_t fnv(uint64_t h, const std::vector<int>& v) {
        for (int x : v) {
            h ^= (uint32_t)x;
            h *= 1099511628211ull;
        }
        return h;
    }

    void ensure_bound() {
        ensure_context();
        const void* ptrs[5] = {grid_.father.data(), grid_.son.data(), grid_.nbor.data(), grid_.next.data(),
                               grid_.headl_vec.data()};
        uint64_t sum = 1469598103934665603ull;
        sum = fnv(sum, grid_.father);
        sum = fnv(sum, grid_.son);
        sum = fnv(sum, grid_.nbor);
        sum = fnv(sum, grid_.next);
        sum = fnv(sum, grid_.headl_vec);
        bool same = sum == topo_sum_ && epoch_ > 0;
        for (int i = 0; i < 5; ++i) same = same && ptrs[i] == topo_ptrs_[i];
        if (same) return;
        for (int i = 0; i < 5; ++i) topo_ptrs_[i] = ptrs[i];
        topo_sum_ = sum;
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
        check(rh_bind(ctx_, &m, RH_HOST_COHERENT), "rh_bind");
    }
};

std::unique_ptr<HydroSolver> create_hydro_solver(AmrGrid& grid, Config& config) {
    return std::make_unique<PatchHydroSolver>(grid, config);
}

}  // namespace ramses
