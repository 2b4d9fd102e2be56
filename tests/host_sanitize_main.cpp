// Host code under AddressSanitizer and UndefinedBehaviorSanitizer, compiled
// and run by tests/test_host_sanitizers.py with the host compiler; no GPU.
//
// Exercises the library's host-only parts (uniform mesh construction and
// initial conditions, rh_driver.cpp; the per-level table builder and its
// range check, rh_tables.cpp) and the oracle (uniform Godunov steps, an
// adaptive Sedov run with its tree updates, a snapshot dump), so an
// out-of-bounds access or undefined operation in any of them ends the run
// with a sanitizer report.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "rh_oracle_amr.h"
#include "rh_tables.h"
#include "rhydro.h"
#include "rhydro_driver.h"

namespace rhi {
std::string g_create_error;  // defined by the library's C-ABI file otherwise
}

extern "C" {
int or_godunov_fine(const or_mesh* m, const or_params* p, int ilevel, double dt, double dx);
void or_set_unew(const or_mesh* m, int ilevel);
void or_set_uold(const or_mesh* m, int ilevel);
double or_courant(const or_mesh* m, int ilevel, double dx, double gamma, double courant_factor, double boxlen);
}

static int failures = 0;
#define CHECK(cond)                                                     \
    do {                                                                \
        if (!(cond)) {                                                  \
            std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #cond); \
            ++failures;                                                 \
        }                                                               \
    } while (0)

static void uniform_run(int level) {
    const int nlevelmax = level;
    const long long ng = rh_uniform_ngrid(1, 1, 1, level);
    const int ngridmax = (int)ng + 8;
    const int ncoarse = 1, ncell = ncoarse + 8 * ngridmax;
    std::vector<int> father(ngridmax), son(ncell), nbor(6 * ngridmax), next(ngridmax), prev(ngridmax);
    std::vector<int> headl(nlevelmax + 1), taill(nlevelmax + 1), numbl(nlevelmax + 1);
    std::vector<double> xg(3 * (size_t)ngridmax), uold(5 * (size_t)ncell), unew(5 * (size_t)ncell);
    CHECK(rh_build_uniform(1, 1, 1, level, nlevelmax, ngridmax, father.data(), son.data(), nbor.data(), next.data(),
                           prev.data(), headl.data(), taill.data(), numbl.data(), xg.data()) == ng);
    rh_mesh rm{};
    rm.ncoarse = ncoarse;
    rm.ngridmax = ngridmax;
    rm.ncell = ncell;
    rm.nvar = 5;
    rm.nx = rm.ny = rm.nz = 1;
    rm.nlevelmax = nlevelmax;
    rm.ncpu = 1;
    rm.nboundary = 0;
    rm.myid = 1;
    rm.father = father.data();
    rm.son = son.data();
    rm.nbor = nbor.data();
    rm.next = next.data();
    rm.headl = headl.data();
    rm.uold = uold.data();
    rm.unew = unew.data();
    rm.mesh_epoch = 1;
    CHECK(rh_init_lognormal(&rm, level, 7, 1.4, 0.5, 0.5) == RH_OK);

    // Tables of every level, range-checked as before an upload.
    rh::TopoIndex topo;
    std::string err;
    CHECK(rh::build_topo_index(rm, topo, err) == RH_OK);
    for (int il = 1; il <= level; ++il) {
        rh::LevelHost h;
        CHECK(rh::build_level(rm, topo, il, h, err) == RH_OK);
        CHECK(rh::validate_level(h, ngridmax, ncell, err) == RH_OK);
    }

    // Oracle steps on the same arrays.
    or_mesh om{};
    om.ncoarse = ncoarse;
    om.ngridmax = ngridmax;
    om.ncell = ncell;
    om.nvar = 5;
    om.nx = om.ny = om.nz = 1;
    om.nlevelmax = nlevelmax;
    om.ncpu = 1;
    om.nboundary = 0;
    om.myid = 1;
    om.father = father.data();
    om.son = son.data();
    om.nbor = nbor.data();
    om.next = next.data();
    om.headl = headl.data();
    om.uold = uold.data();
    om.unew = unew.data();
    const or_params p{1.4, 1, 2, 1, 0, 1.0};
    const double dx = 1.0 / (1 << level);
    for (int k = 0; k < 2; ++k) {
        const double dt = or_courant(&om, level, dx, 1.4, 0.8, 1.0);
        CHECK(dt > 0.0);
        or_set_unew(&om, level);
        CHECK(or_godunov_fine(&om, &p, level, dt, dx) == 0);
        or_set_uold(&om, level);
    }
    std::printf("ok   uniform level %d: tables and two oracle steps\n", level);
}

static void adaptive_run(const char* root) {
    const int reg_type[2] = {0, 1};
    const double xc[2] = {0.5, 0.5}, yc[2] = {0.5, 0.5}, zc[2] = {0.5, 0.5};
    const double h = 1.0 / 8;  // cell width at level 4
    const double lx[2] = {10.0, 1.5 * h}, ly[2] = {10.0, 1.5 * h}, lz[2] = {10.0, 1.5 * h};
    const double d[2] = {1.0, 1.0}, u[2] = {0, 0}, v[2] = {0, 0}, w[2] = {0, 0};
    const double pr[2] = {1e-5, 0.4 / (8 * h * h * h)};
    or_amr_config c{};
    c.nx = c.ny = c.nz = 1;
    c.ngridmax = 6000;
    c.levelmin = 4;
    c.levelmax = 6;
    c.boxlen = 1.0;
    c.gamma = 1.4;
    c.courant_factor = 0.8;
    c.err_grad_d = c.err_grad_p = 0.1;
    c.err_grad_v = -1.0;
    c.nexpand = 1;
    c.hydro = or_params{1.4, 1, 2, 1, 0, 1.0};
    c.ic_kind = 0;
    c.nreg = 2;
    c.reg_type = reg_type;
    c.x_center = xc;
    c.y_center = yc;
    c.z_center = zc;
    c.length_x = lx;
    c.length_y = ly;
    c.length_z = lz;
    c.d_region = d;
    c.u_region = u;
    c.v_region = v;
    c.w_region = w;
    c.p_region = pr;
    void* a = or_amr_create(&c);
    CHECK(a != nullptr);
    if (!a) return;
    CHECK(or_amr_step(a, 3) == 0);
    or_amr_state s{};
    or_amr_get(a, &s);
    CHECK(s.status == 0 && s.numbl[5] > 0 && s.numbl[6] > 0);
    or_snapshot_info info{};
    info.t = s.t;
    info.nstep = s.nstep;
    info.levelmin = 4;
    info.levelmax = 6;
    info.boxlen = 1.0;
    info.aexp = 1.0;
    info.units_length = info.units_density = info.units_time = 1.0;
    CHECK(or_amr_dump(a, &info, 1, root) == 0);
    std::printf("ok   adaptive Sedov run, levels 4-6, 3 coarse steps, lists %d/%d/%d, snapshot written\n",
                s.numbl[4], s.numbl[5], s.numbl[6]);
    or_amr_destroy(a);
}

int main(int argc, char** argv) {
    if (argc < 2) {
        std::printf("usage: %s <snapshot root>\n", argv[0]);
        return 2;
    }
    uniform_run(3);
    uniform_run(5);
    adaptive_run(argv[1]);
    std::printf("%d failures\n", failures);
    return failures ? 1 : 0;
}
