// rh_oracle_core.h -- shared definitions of the CPU oracle (test
// infrastructure only; see rh_oracle.cpp): the mesh view, the namelist
// parameters, libstdc++'s min/max/clamp semantics, the octant tables and the
// neighbour lookup of AmrGrid, restated.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

// std::max / std::min / std::clamp exactly as libstdc++ defines them.
inline double smax(double a, double b) { return (a < b) ? b : a; }
inline double smin(double a, double b) { return (b < a) ? b : a; }
inline double sclamp(double v, double lo, double hi) { return (v < lo) ? lo : ((hi < v) ? hi : v); }
// std::min({a,b,c}): first smallest element.
inline double smin3(double a, double b, double c) {
    double m = a;
    if (b < m) m = b;
    if (c < m) m = c;
    return m;
}
inline void swapd(double& a, double& b) { double t = a; a = b; b = t; }

constexpr int NV = 5;  // rho, rho*u, rho*v, rho*w, E

// Constants.cpp:8-43
const int III[3][2][8] = {
    {{1, 0, 1, 0, 1, 0, 1, 0}, {0, 2, 0, 2, 0, 2, 0, 2}},
    {{3, 3, 0, 0, 3, 3, 0, 0}, {0, 0, 4, 4, 0, 0, 4, 4}},
    {{5, 5, 5, 5, 0, 0, 0, 0}, {0, 0, 0, 0, 6, 6, 6, 6}}};
const int JJJ[3][2][8] = {
    {{2, 1, 4, 3, 6, 5, 8, 7}, {2, 1, 4, 3, 6, 5, 8, 7}},
    {{3, 4, 1, 2, 7, 8, 5, 6}, {3, 4, 1, 2, 7, 8, 5, 6}},
    {{5, 6, 7, 8, 1, 2, 3, 4}, {5, 6, 7, 8, 1, 2, 3, 4}}};

}  // namespace

extern "C" {

// Mesh view over AmrGrid-shaped arrays (AmrGrid.hpp:15-33, 104-107).
typedef struct {
    int ncoarse, ngridmax, ncell, nvar;
    int nx, ny, nz;
    int nlevelmax, ncpu, nboundary, myid;
    const int* father;   // [ngridmax]
    const int* son;      // [ncell]
    const int* nbor;     // [6*ngridmax], (j-1)*ngridmax + igrid-1
    const int* next;     // [ngridmax]
    const int* headl;    // [(nlevelmax+1)*(ncpu+nboundary)]
    double* uold;        // [nvar*ncell], (ivar-1)*ncell + icell-1
    double* unew;        // [nvar*ncell]
    const double* f;     // [3*ncell] or null (gravity, zero in all configs)
    const int* bound_type;  // [nbound_type] (empty in every config)
    int nbound_type;
} or_mesh;

// Namelist-derived parameters of the path.
//   riemann: 0=llf 1=hll 2=hllc 3=exact 4=acoustic (HydroSolver.cpp:456-461)
typedef struct {
    double gamma;
    int slope_type;
    int riemann;
    int interpol_type;
    int interpol_var;
    double boxlen;
} or_params;

}  // extern "C"

namespace {

struct Ctx {
    const or_mesh* m;
    const or_params* p;
    double uold(int icell, int iv) const { return m->uold[(size_t)(iv - 1) * m->ncell + (icell - 1)]; }
    double& unew(int icell, int iv) const { return m->unew[(size_t)(iv - 1) * m->ncell + (icell - 1)]; }

    int headl(int icpu, int ilevel) const {
        if (m->ncpu == 0) return 0;
        if (ilevel < 0 || ilevel > m->nlevelmax || icpu < 1 || icpu > m->ncpu + m->nboundary) return 0;
        return m->headl[(size_t)ilevel * (m->ncpu + m->nboundary) + (icpu - 1)];
    }

    // AmrGrid.cpp:389-400
    void nbor_grids(int igrid, int ign[7]) const {
        ign[0] = igrid;
        for (int i = 1; i <= 6; ++i) {
            int val = m->nbor[(size_t)(i - 1) * m->ngridmax + igrid - 1];
            ign[i] = (val > m->ncell) ? 0 : val;
        }
    }

    // AmrGrid.cpp:513-550 (nboundary boundaries never match, periodic wrap)
    void nbor_cells_coarse(int icell, int icn[6]) const {
        for (int i = 0; i < 6; ++i) icn[i] = 0;
        if (icell <= 0 || icell > m->ncoarse) return;
        int idx = icell - 1;
        int nx = m->nx, ny = m->ny, nz = m->nz;
        int iz = idx / (nx * ny);
        int rem = idx % (nx * ny);
        int iy = rem / nx;
        int ix = rem % nx;
        int n[3] = {nx, ny, nz};
        for (int idim = 0; idim < 3; ++idim) {
            for (int side = 0; side < 2; ++side) {
                int c[3] = {ix, iy, iz};
                c[idim] += (side == 0) ? -1 : 1;
                if (c[idim] < 0 || c[idim] >= n[idim]) {
                    // No boundary regions exist (nboundary == 0 in every
                    // config; boundary_params is never parsed), so the
                    // reference falls back to a periodic wrap.
                    int p[3] = {ix, iy, iz};
                    p[idim] = (p[idim] + (side == 0 ? -1 : 1) + n[idim]) % n[idim];
                    icn[idim * 2 + side] = p[2] * nx * ny + p[1] * nx + p[0] + 1;
                } else {
                    icn[idim * 2 + side] = c[2] * nx * ny + c[1] * nx + c[0] + 1;
                }
            }
        }
    }

    // AmrGrid.cpp:402-444
    void nbor_cells(const int ign[7], int ic, int icn[6], int igrid) const {
        for (int i = 0; i < 6; ++i) icn[i] = 0;
        if (igrid <= 0) return;
        int ifather = m->father[igrid - 1];
        for (int idim = 0; idim < 3; ++idim) {
            for (int inbor = 0; inbor < 2; ++inbor) {
                int ig_idx = III[idim][inbor][ic - 1];
                int ic_pos = JJJ[idim][inbor][ic - 1];
                int ig = ign[ig_idx];
                if (ig > 0) {
                    if (ig_idx == 0) {
                        icn[idim * 2 + inbor] = m->ncoarse + (ic_pos - 1) * m->ngridmax + ig;
                    } else {
                        int ig_child = (ig <= m->ncell) ? m->son[ig - 1] : 0;
                        if (ig_child > 0)
                            icn[idim * 2 + inbor] = m->ncoarse + (ic_pos - 1) * m->ngridmax + ig_child;
                        else
                            icn[idim * 2 + inbor] = ig;
                    }
                } else if (ig < 0) {
                    icn[idim * 2 + inbor] = ig;
                } else {
                    int ifn[6] = {0, 0, 0, 0, 0, 0};
                    if (ifather <= m->ncoarse) {
                        nbor_cells_coarse(ifather, ifn);
                    } else {
                        int p_ig = ((ifather - m->ncoarse - 1) % m->ngridmax) + 1;
                        int p_ign[7];
                        nbor_grids(p_ig, p_ign);
                        int p_ic = ((ifather - m->ncoarse - 1) / m->ngridmax) + 1;
                        nbor_cells(p_ign, p_ic, ifn, p_ig);
                    }
                    icn[idim * 2 + inbor] = ifn[idim * 2 + inbor];
                }
            }
        }
    }

    bool bound_is_reflect(int ib) const {
        return ib > 0 && ib <= m->nbound_type && m->bound_type[ib - 1] == 1;
    }
};

}  // namespace

extern "C" {
// Hydro path of the oracle (rh_oracle.cpp), used by the AMR driver restatement.
int or_godunov_fine(const or_mesh* m, const or_params* p, int ilevel, double dt, double dx);
void or_set_unew(const or_mesh* m, int ilevel);
void or_set_uold(const or_mesh* m, int ilevel);
double or_courant(const or_mesh* m, int ilevel, double dx, double gamma, double courant_factor, double boxlen);
void or_interpol_hydro(const or_params* p, int nvar, const double* u1, double* u2);
}
