// rh_driver.cpp -- host driver helpers of the standalone build
// (rhydro_driver.h): uniform mesh construction and initial conditions.
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "rhydro.h"
#include "rh_guard.h"
#include "rhydro_driver.h"

namespace {

int cell_of(int ncoarse, int ngridmax, int igrid, int ic0) { return ncoarse + ic0 * ngridmax + igrid; }

// Depth of a grid along its father chain (the reference's cell level).
int grid_depth(const rh_mesh& m, int ig) {
    int level = 1;
    int cur = ig;
    while (cur > 0) {
        int fc = m.father[cur - 1];
        if (fc <= m.ncoarse) break;
        cur = ((fc - m.ncoarse - 1) % m.ngridmax) + 1;
        level++;
    }
    return level;
}

int headl_of(const rh_mesh& m, int il) {
    if (il < 0 || il > m.nlevelmax) return 0;
    return m.headl[(size_t)il * (m.ncpu + m.nboundary) + (m.myid - 1)];
}


// Calls fn(icell, x, y, z) for every coarse cell and every cell of every
// level list, at the cell centres AmrGrid::get_cell_center gives
// (AmrGrid.cpp:333-363); then restricts upward as restrict_fine does for
// il = nlevelmax-1 .. 0 (TreeUpdater.cpp:195-215) and sets unew = uold.
template <typename Fn> void fill_cells(const rh_mesh& m, const double* xg, double boxlen, Fn fn) {
    auto U = [&](int icell, int iv) -> double& { return m.uold[(size_t)(iv - 1) * m.ncell + icell - 1]; };
    int nmax = m.nx > m.ny ? m.nx : m.ny;
    nmax = nmax > m.nz ? nmax : m.nz;
    double dx_coarse = boxlen / nmax;
    for (int i = 1; i <= m.ncoarse; ++i) {
        int idx = i - 1;
        int iz = idx / (m.nx * m.ny);
        idx %= (m.nx * m.ny);
        int iy = idx / m.nx, ix = idx % m.nx;
        fn(i, (ix + 0.5) * dx_coarse, (iy + 0.5) * dx_coarse, (iz + 0.5) * dx_coarse);
    }
    for (int il = 1; il <= m.nlevelmax; ++il) {
        for (int ig = headl_of(m, il); ig > 0; ig = m.next[ig - 1]) {
            int level = grid_depth(m, ig);
            double dx = boxlen / (double)(m.nx * (1 << level));
            double scale = boxlen / (double)m.nx;
            for (int ic0 = 0; ic0 < 8; ++ic0) {
                int ixyz[3] = {ic0 & 1, (ic0 & 2) >> 1, (ic0 & 4) >> 2};
                double xc[3];
                for (int d = 0; d < 3; ++d)
                    xc[d] = xg[(size_t)d * m.ngridmax + ig - 1] * scale + (double)(ixyz[d] - 0.5) * dx;
                fn(cell_of(m.ncoarse, m.ngridmax, ig, ic0), xc[0], xc[1], xc[2]);
            }
        }
    }
    for (int il = m.nlevelmax - 1; il >= 0; --il) {
        for (int ig = headl_of(m, il + 1); ig > 0; ig = m.next[ig - 1]) {
            int fc = m.father[ig - 1];
            if (fc <= 0) continue;
            for (int iv = 1; iv <= m.nvar; ++iv) {
                double sum = 0;
                for (int ic0 = 0; ic0 < 8; ++ic0) sum += U(cell_of(m.ncoarse, m.ngridmax, ig, ic0), iv);
                U(fc, iv) = sum / 8;
            }
        }
    }
    for (size_t i = 0; i < (size_t)m.nvar * m.ncell; ++i) m.unew[i] = m.uold[i];
}

// Conservative state of (d, u, v, w, p) into cell idc (Initializer.cpp:104-121).
void set_prim(const rh_mesh& m, int idc, double gamma, double dr, double ur, double vr, double wr, double pr) {
    auto U = [&](int iv) -> double& { return m.uold[(size_t)(iv - 1) * m.ncell + idc - 1]; };
    U(1) = dr;
    U(2) = dr * ur;
    U(3) = dr * vr;
    U(4) = dr * wr;
    double v2 = ur * ur;
    v2 += vr * vr;
    v2 += wr * wr;
    double e_kin = 0.5 * dr * v2;
    double e_int = pr / (gamma - 1.0);
    double e_nonthermal = 0.0;
    U(5) = e_kin + e_int + e_nonthermal;
}

}  // namespace

extern "C" {

long long rh_uniform_ngrid(int nx, int ny, int nz, int levelmin) {
    long long nc = (long long)nx * ny * nz;
    long long tot = 0, per = nc;
    for (int l = 2; l <= levelmin; ++l) {
        tot += per;
        per *= 8;
    }
    return tot;
}

long long rh_build_uniform(int nx, int ny, int nz, int levelmin, int nlevelmax, int ngridmax, int* father, int* son,
                           int* nbor, int* next, int* prev, int* headl, int* taill, int* numbl, double* xg) {
    if (nx < 1 || ny < 1 || nz < 1 || levelmin < 1 || nlevelmax < levelmin || ngridmax < 1) return -RH_ERR_ARG;
    const long long need = rh_uniform_ngrid(nx, ny, nz, levelmin);
    if (need > ngridmax) return -RH_ERR_ARG;
    const int ncoarse = nx * ny * nz;
    const long long ncell = (long long)ncoarse + 8LL * ngridmax;
    for (long long i = 0; i < ncell; ++i) son[i] = 0;
    for (int i = 0; i < ngridmax; ++i) {
        father[i] = 0;
        next[i] = 0;
        prev[i] = 0;
        for (int f = 0; f < 6; ++f) nbor[(size_t)f * ngridmax + i] = 0;
        for (int d = 0; d < 3; ++d) xg[(size_t)d * ngridmax + i] = 0.0;
    }
    for (int l = 0; l <= nlevelmax; ++l) headl[l] = taill[l] = numbl[l] = 0;
    auto nb = [&](int f, int g) -> int& { return nbor[(size_t)f * ngridmax + g - 1]; };
    auto X = [&](int d, int g) -> double& { return xg[(size_t)d * ngridmax + g - 1]; };
    int ng = 0;
    auto append = [&](int g, int l) {
        if (headl[l] == 0) {
            headl[l] = taill[l] = g;
        } else {
            next[taill[l] - 1] = g;
            prev[g - 1] = taill[l];
            taill[l] = g;
        }
        numbl[l]++;
    };
    // Level list 2: one oct per coarse cell (TreeUpdater.cpp:33-39,78-89).
    if (levelmin >= 2 && nlevelmax > 1) {
        for (int ic = 1; ic <= ncoarse; ++ic) {
            int g = ++ng;
            son[ic - 1] = g;
            father[g - 1] = ic;
            int idx = ic - 1;
            int x[3] = {idx % nx, (idx % (nx * ny)) / nx, idx / (nx * ny)};
            int n[3] = {nx, ny, nz};
            for (int d = 0; d < 3; ++d)
                for (int s = 0; s < 2; ++s) {
                    int c[3] = {x[0], x[1], x[2]};
                    c[d] = (s == 0) ? (c[d] - 1 + n[d]) % n[d] : (c[d] + 1) % n[d];
                    nb(2 * d + s, g) = c[2] * nx * ny + c[1] * nx + c[0] + 1;
                }
            for (int d = 0; d < 3; ++d) X(d, g) = (double)x[d] + 0.5;
            append(g, 2);
        }
    }
    // Level lists 3..levelmin, each from the previous list in list order
    // (TreeUpdater.cpp:41-49,90-101,148).
    for (int l = 2; l < levelmin && l < nlevelmax; ++l) {
        const double dx_level = 1.0 / (double)(1 << (l - 1));
        for (int pg = headl[l]; pg > 0; pg = next[pg - 1]) {
            for (int ic0 = 0; ic0 < 8; ++ic0) {
                int g = ++ng;
                int pc = cell_of(ncoarse, ngridmax, pg, ic0);
                son[pc - 1] = g;
                father[g - 1] = pc;
                for (int d = 0; d < 3; ++d)
                    for (int s = 0; s < 2; ++s) {
                        int f = 2 * d + s;
                        int ic_pos = ic0 ^ (1 << d);
                        bool cross = ((ic0 >> d) & 1) == s;
                        int v;
                        if (!cross) {
                            v = cell_of(ncoarse, ngridmax, pg, ic_pos);
                        } else {
                            int ig = nb(f, pg);
                            int child = ig > 0 ? son[ig - 1] : 0;
                            v = child > 0 ? cell_of(ncoarse, ngridmax, child, ic_pos) : ig;
                        }
                        nb(f, g) = v;
                    }
                int ixyz[3] = {ic0 & 1, (ic0 >> 1) & 1, (ic0 >> 2) & 1};
                for (int d = 0; d < 3; ++d) X(d, g) = X(d, pg) + (double)(ixyz[d] - 0.5) * dx_level;
                append(g, l + 1);
            }
        }
    }
    // Remaining grids form the free list (AmrGrid.cpp:64-69 after the
    // allocations above).
    for (int g = ng + 1; g <= ngridmax; ++g) {
        next[g - 1] = (g < ngridmax) ? g + 1 : 0;
        prev[g - 1] = (g > ng + 1) ? g - 1 : 0;
    }
    return ng;
}

int rh_init_regions(const rh_mesh* mp, const double* xg, double boxlen, double gamma, int nreg, const int* type,
                    const double* x_center, const double* y_center, const double* z_center,
                    const double* length_x, const double* length_y, const double* length_z, const double* d_region,
                    const double* u_region, const double* v_region, const double* w_region,
                    const double* p_region) {
    return rhi::guard(&rhi::g_create_error, nullptr, [&]() -> int {
        if (!mp || !xg || nreg < 0) return RH_ERR_ARG;
        const rh_mesh& m = *mp;
        fill_cells(m, xg, boxlen, [&](int idc, double x, double y, double z) {
            for (int ir = 0; ir < nreg; ++ir) {
                bool in = false;
                if (type[ir] == 0) {
                    in = std::fabs(x - x_center[ir]) < 0.5 * length_x[ir] &&
                         std::fabs(y - y_center[ir]) < 0.5 * length_y[ir] &&
                         std::fabs(z - z_center[ir]) < 0.5 * length_z[ir];
                } else {
                    double r2 = std::pow(x - x_center[ir], 2) + std::pow(y - y_center[ir], 2) +
                                std::pow(z - z_center[ir], 2);
                    in = r2 < std::pow(length_x[ir], 2);
                }
                if (!in) continue;
                set_prim(m, idc, gamma, d_region[ir], u_region[ir], v_region[ir], w_region[ir], p_region[ir]);
            }
        });
        return RH_OK;
    });
}

int rh_init_shock_tube(const rh_mesh* mp, const double* xg, double boxlen, double gamma, const double* normal,
                       const double* point, const double* left, const double* right) {
    return rhi::guard(&rhi::g_create_error, nullptr, [&]() -> int {
        if (!mp || !xg || !normal || !point || !left || !right) return RH_ERR_ARG;
        const rh_mesh& m = *mp;
        fill_cells(m, xg, boxlen, [&](int idc, double x, double y, double z) {
            const double s = (x - point[0]) * normal[0] + (y - point[1]) * normal[1] + (z - point[2]) * normal[2];
            const double* q = s < 0.0 ? left : right;
            set_prim(m, idc, gamma, q[0], q[1], q[2], q[3], q[4]);
        });
        return RH_OK;
    });
}

int rh_init_lognormal(const rh_mesh* mp, int level, unsigned long long seed, double gamma, double sigma,
                      double vel_amp) {
    return rhi::guard(&rhi::g_create_error, nullptr, [&]() -> int {
        if (!mp) return RH_ERR_ARG;
        const rh_mesh& m = *mp;
        auto U = [&](int icell, int iv) -> double& { return m.uold[(size_t)(iv - 1) * m.ncell + icell - 1]; };
        for (int i = 1; i <= m.ncell; ++i) {
            U(i, 1) = 1.0;
            U(i, 2) = 0.0;
            U(i, 3) = 0.0;
            U(i, 4) = 0.0;
            U(i, 5) = 1.0 / (gamma - 1.0);
        }
        std::mt19937_64 rng(seed);
        bool have = false;
        double spare = 0.0;
        const double two_pi = 6.283185307179586;
        auto normal = [&]() {
            if (have) {
                have = false;
                return spare;
            }
            double u1 = ((double)(rng() >> 11) + 1.0) * 0x1.0p-53;
            double u2 = ((double)(rng() >> 11) + 1.0) * 0x1.0p-53;
            double r = std::sqrt(-2.0 * std::log(u1));
            spare = r * std::sin(two_pi * u2);
            have = true;
            return r * std::cos(two_pi * u2);
        };
        for (int ig = headl_of(m, level); ig > 0; ig = m.next[ig - 1]) {
            for (int ic0 = 0; ic0 < 8; ++ic0) {
                int c = cell_of(m.ncoarse, m.ngridmax, ig, ic0);
                if (m.son[c - 1] != 0) continue;
                double rho = std::exp(sigma * normal());
                double p = std::exp(sigma * normal());
                double cs = std::sqrt(gamma * p / rho);
                double v[3];
                for (int d = 0; d < 3; ++d) v[d] = vel_amp * cs * normal();
                U(c, 1) = rho;
                U(c, 2) = rho * v[0];
                U(c, 3) = rho * v[1];
                U(c, 4) = rho * v[2];
                U(c, 5) = p / (gamma - 1.0) + 0.5 * rho * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
            }
        }
        for (size_t i = 0; i < (size_t)m.nvar * m.ncell; ++i) m.unew[i] = m.uold[i];
        return RH_OK;
    });
}

}  // extern "C"
