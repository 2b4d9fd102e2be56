// rh_oracle_amr.cpp -- CPU restatement of RAMSES-CPP's adaptive time
// integration around the hydro path: the tree updater, the level lists and
// free list of AmrGrid, the region initial conditions and the subcycled
// Simulation loop.
//
// TEST INFRASTRUCTURE ONLY, like rh_oracle.cpp: the parity checker of the
// library's AMR driver.  Only tests/ and bench.py's CPU baseline load it.
//
// Restated, in the reference's order of operations (ncpu = 1, nboundary = 0,
// NDIM = 3, nvar = 5, no particles / gravity / MHD):
//   * AmrGrid::allocate, get_free_grid, free_grid, add_to_level_list,
//     get_cell_center, find_cell_by_coords, get_nbor_cells_exact,
//     get_27_cell_neighbors                src/core/AmrGrid.cpp:13-70,259-302,
//                                          331-388,449-511; AmrGrid.hpp:69-80
//   * TreeUpdater::make_grid_fine, remove_grid_fine, restrict_fine,
//     smooth_fine, test_flag, flag_fine, authorize_fine, ensure_ref_rules
//                                          src/core/TreeUpdater.cpp:15-636
//   * Initializer::apply_all / region_condinit (region ICs)
//                                          src/Initializer.cpp:19-157
//   * Simulation::initialize (refinement loops), run (one coarse step) and
//     amr_step                             src/core/Simulation.cpp:160-255,
//                                          355-395,522-635
// with the hydro calls going to the oracle's own godunov_fine, set_unew,
// set_uold, compute_courant_step and interpol_hydro (rh_oracle.cpp).
// AmrGrid::resize_grids is not restated: a run that exhausts the free list
// stops with an error instead (size ngridmax for the run).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "rh_oracle_amr.h"
#include "rh_oracle_core.h"
#include "rh_oracle_snapshot.h"


namespace {

struct Region {
    int type;
    double xc, yc, zc, lx, ly, lz, d, u, v, w, p;
};

struct Amr {
    // AmrGrid (AmrGrid.hpp:15-33), ncpu = 1, nboundary = 0, myid = 1.
    int nx, ny, nz, ncoarse, ngridmax, ncell, nlevelmax;
    double boxlen, gamma;
    std::vector<int> father, son, nbor, next, prev, flag1, flag2, cpu_map;
    std::vector<int> headl, taill, numbl;  // [nlevelmax + 1]
    std::vector<double> xg, uold, unew;
    int headf = 0, tailf = 0, numbf = 0;
    // Simulation state (Simulation.hpp).
    int levelmin;
    double courant_factor, ed, ep, ev;
    int nexpand;
    std::vector<int> nsubcycle;
    std::vector<double> dtnew, dtold;
    double t = 0.0;
    int nstep = 0;
    or_params hp;
    int ic_kind = 0;
    std::vector<Region> regions;
    double kh_amp = 0.0;
    // Cell-updates (leaf cells of every godunov_fine call) since creation.
    long long cell_updates = 0;
    int status = 0;  // nonzero once the free list ran out

    double& U(int icell, int iv) { return uold[(size_t)(iv - 1) * ncell + (icell - 1)]; }
    double& N(int icell, int iv) { return unew[(size_t)(iv - 1) * ncell + (icell - 1)]; }
    int cell(int igrid, int ic) const { return ncoarse + (ic - 1) * ngridmax + igrid; }

    or_mesh view() {
        or_mesh m{};
        m.ncoarse = ncoarse;
        m.ngridmax = ngridmax;
        m.ncell = ncell;
        m.nvar = 5;
        m.nx = nx;
        m.ny = ny;
        m.nz = nz;
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
        m.f = nullptr;
        m.bound_type = nullptr;
        m.nbound_type = 0;
        return m;
    }

    // AmrGrid.cpp:13-70
    void allocate(int nx_, int ny_, int nz_, int ngridmax_, int nlevelmax_) {
        nx = std::max(nx_, 1);
        ny = std::max(ny_, 1);
        nz = std::max(nz_, 1);
        ngridmax = std::max(ngridmax_, 1);
        nlevelmax = std::max(nlevelmax_, 1);
        ncoarse = nx * ny * nz;
        ncell = ncoarse + 8 * ngridmax;
        headl.assign(nlevelmax + 1, 0);
        taill.assign(nlevelmax + 1, 0);
        numbl.assign(nlevelmax + 1, 0);
        next.assign(ngridmax, 0);
        prev.assign(ngridmax, 0);
        father.assign(ngridmax, 0);
        son.assign(ncell, 0);
        nbor.assign((size_t)6 * ngridmax, 0);
        flag1.assign(ncell, 0);
        flag2.assign(ncell, 0);
        cpu_map.assign(ncell, 1);
        xg.assign((size_t)3 * ngridmax, 0.0);
        uold.assign((size_t)5 * ncell, 0.0);
        unew.assign((size_t)5 * ncell, 0.0);
        headf = 1;
        tailf = ngridmax;
        for (int i = 1; i <= ngridmax; ++i) {
            next[i - 1] = (i < ngridmax) ? i + 1 : 0;
            prev[i - 1] = (i > 1) ? i - 1 : 0;
        }
        numbf = ngridmax;
    }

    // AmrGrid.cpp:259-268 (no resize: 0 when the free list is empty)
    int get_free_grid() {
        if (headf == 0) return 0;
        int igrid = headf;
        headf = next[igrid - 1];
        if (headf == 0) tailf = 0;
        else prev[headf - 1] = 0;
        next[igrid - 1] = 0;
        prev[igrid - 1] = 0;
        numbf--;
        return igrid;
    }

    // AmrGrid.cpp:270-287
    void free_grid(int igrid) {
        if (igrid <= 0) return;
        for (int i = 0; i < 6; ++i) nbor[(size_t)i * ngridmax + igrid - 1] = 0;
        int father_cell = father[igrid - 1];
        if (father_cell > 0) son[father_cell - 1] = 0;
        father[igrid - 1] = 0;
        for (int ic = 1; ic <= 8; ++ic) {
            int idc = ncoarse + (ic - 1) * ngridmax + igrid - 1;
            son[idc] = 0;
            flag1[idc] = 0;
            flag2[idc] = 0;
        }
        if (tailf == 0) {
            headf = igrid;
            tailf = igrid;
            prev[igrid - 1] = 0;
            next[igrid - 1] = 0;
        } else {
            next[tailf - 1] = igrid;
            prev[igrid - 1] = tailf;
            next[igrid - 1] = 0;
            tailf = igrid;
        }
        numbf++;
    }

    // AmrGrid.cpp:289-302
    void add_to_level_list(int igrid, int ilevel) {
        if (headl[ilevel] == 0) {
            headl[ilevel] = igrid;
            taill[ilevel] = igrid;
            prev[igrid - 1] = 0;
            next[igrid - 1] = 0;
        } else {
            int tail = taill[ilevel];
            next[tail - 1] = igrid;
            prev[igrid - 1] = tail;
            next[igrid - 1] = 0;
            taill[ilevel] = igrid;
        }
        numbl[ilevel]++;
    }

    int count(int ilevel) const { return (ilevel >= 0 && ilevel <= nlevelmax) ? numbl[ilevel] : 0; }

    // AmrGrid.hpp:69-80
    int cell_level(int icell) const {
        if (icell <= ncoarse) return 0;
        int ig = ((icell - ncoarse - 1) % ngridmax) + 1;
        int level = 1, cur = ig;
        while (cur > 0) {
            int fc = father[cur - 1];
            if (fc <= ncoarse) break;
            cur = ((fc - ncoarse - 1) % ngridmax) + 1;
            level++;
        }
        return level;
    }

    // AmrGrid.cpp:331-365
    void cell_center(int icell, double xc[3]) const {
        xc[0] = 0.5 * boxlen;
        xc[1] = 0.5 * boxlen;
        xc[2] = 0.5 * boxlen;
        if (icell <= ncoarse) {
            int idx = icell - 1;
            int iz = idx / (nx * ny);
            idx %= (nx * ny);
            int iy = idx / nx, ix = idx % nx;
            double dxc = boxlen / std::max({nx, ny, nz});
            xc[0] = (ix + 0.5) * dxc;
            xc[1] = (iy + 0.5) * dxc;
            xc[2] = (iz + 0.5) * dxc;
        } else {
            int ig = ((icell - ncoarse - 1) % ngridmax) + 1;
            int ic = ((icell - ncoarse - 1) / ngridmax) + 1;
            int idc = ic - 1;
            int ixyz[3] = {idc & 1, (idc & 2) >> 1, (idc & 4) >> 2};
            int level = 1, cur = ig;
            while (cur > 0) {
                int fc = father[cur - 1];
                if (fc <= ncoarse) break;
                cur = ((fc - ncoarse - 1) % ngridmax) + 1;
                level++;
            }
            double dx = boxlen / (double)(nx * (1 << level));
            double scale = boxlen / (double)nx;
            for (int d = 0; d < 3; ++d) xc[d] = xg[(size_t)d * ngridmax + ig - 1] * scale + (double)(ixyz[d] - 0.5) * dx;
        }
    }

    // AmrGrid.cpp:367-387
    int find_cell_by_coords(const double x[3], int level) const {
        double dxc = boxlen / std::max({nx, ny, nz});
        int ix = std::clamp((int)(x[0] / dxc), 0, nx - 1);
        int iy = std::clamp((int)(x[1] / dxc), 0, ny - 1);
        int iz = std::clamp((int)(x[2] / dxc), 0, nz - 1);
        int cur = iz * nx * ny + iy * nx + ix + 1;
        int lev = 1;
        while (lev <= level && son[cur - 1] > 0) {
            int ig = son[cur - 1];
            lev++;
            double scale = boxlen / (double)nx;
            int ixc = (x[0] > xg[(size_t)0 * ngridmax + ig - 1] * scale) ? 1 : 0;
            int iyc = (x[1] > xg[(size_t)1 * ngridmax + ig - 1] * scale) ? 1 : 0;
            int izc = (x[2] > xg[(size_t)2 * ngridmax + ig - 1] * scale) ? 1 : 0;
            int ic = izc * 4 + iyc * 2 + ixc + 1;
            cur = ncoarse + (ic - 1) * ngridmax + ig;
        }
        return cur;
    }

    // AmrGrid.cpp:449-472
    void nbor_cells_exact(const int ign[7], int ic, int icn[6]) const {
        for (int i = 0; i < 6; ++i) icn[i] = 0;
        for (int idim = 0; idim < 3; ++idim)
            for (int inbor = 0; inbor < 2; ++inbor) {
                int ig_idx = III[idim][inbor][ic - 1];
                int ic_pos = JJJ[idim][inbor][ic - 1];
                int ig = ign[ig_idx];
                int slot = idim * 2 + inbor;
                if (ig_idx == 0) {
                    icn[slot] = ncoarse + (ic_pos - 1) * ngridmax + ig;
                } else if (ig > 0) {
                    int ig_child = (ig <= ncell) ? son[ig - 1] : 0;
                    if (ig_child > 0) icn[slot] = ncoarse + (ic_pos - 1) * ngridmax + ig_child;
                }
            }
    }

    // AmrGrid.cpp:474-511
    void neighbors27(int icell, int nb[27]) const {
        for (int i = 0; i < 27; ++i) nb[i] = 0;
        if (icell <= 0 || icell > ncell) return;
        int level = cell_level(icell);
        double xc[3];
        cell_center(icell, xc);
        double dxl = boxlen / (double)(nx * (1 << level));
        for (int dz = -1; dz <= 1; ++dz)
            for (int dy = -1; dy <= 1; ++dy)
                for (int dx = -1; dx <= 1; ++dx) {
                    double x[3] = {xc[0], xc[1], xc[2]};
                    x[0] += dx * dxl;
                    x[1] += dy * dxl;
                    x[2] += dz * dxl;
                    for (int d = 0; d < 3; ++d) {
                        if (x[d] < 0.0) x[d] += boxlen;
                        else if (x[d] >= boxlen) x[d] -= boxlen;
                    }
                    int c = find_cell_by_coords(x, level);
                    if (cell_level(c) == level) nb[(dx + 1) * 9 + (dy + 1) * 3 + (dz + 1)] = c;
                }
    }

    // TreeUpdater.cpp:15-25
    int nbor_of_coarse(int ic, int idim, int side) const {
        int idx = ic - 1;
        int iz = idx / (nx * ny);
        int rem = idx % (nx * ny);
        int iy = rem / nx, ix = rem % nx;
        int x[3] = {ix, iy, iz};
        int n[3] = {nx, ny, nz};
        if (side == 0) x[idim] = (x[idim] - 1 + n[idim]) % n[idim];
        else x[idim] = (x[idim] + 1) % n[idim];
        return x[2] * nx * ny + x[1] * nx + x[0] + 1;
    }

    // TreeUpdater.cpp:585-591
    void authorize_fine(int ilevel) {
        if (ilevel >= nlevelmax) return;
        for (int ic = 0; ic < ncell; ++ic) flag2[ic] = (cpu_map[ic] == 1) ? 1 : 0;
    }

    // TreeUpdater.cpp:27-150
    void make_grid_fine(int ilevel) {
        if (ilevel >= nlevelmax) return;
        authorize_fine(ilevel);
        std::vector<int> cells;
        if (ilevel == 1) {
            for (int ic = 1; ic <= ncoarse; ++ic)
                if (flag1[ic - 1] == 1 && flag2[ic - 1] == 1 && son[ic - 1] == 0) cells.push_back(ic);
        } else {
            for (int ig = headl[ilevel]; ig > 0; ig = next[ig - 1])
                for (int ic = 1; ic <= 8; ++ic) {
                    int idc = ncoarse + (ic - 1) * ngridmax + ig - 1;
                    if (flag1[idc] == 1 && flag2[idc] == 1 && son[idc] == 0) cells.push_back(idc + 1);
                }
        }
        or_mesh mv = view();
        Ctx c{&mv, &hp};
        for (size_t k = 0; k < cells.size(); ++k) {
            int ic_coarse = cells[k];
            int idc = ic_coarse - 1;
            int new_ig = get_free_grid();
            if (new_ig == 0) {
                status = -1;
                return;
            }
            son[idc] = new_ig;
            father[new_ig - 1] = idc + 1;
            if (ilevel == 1) {
                for (int idim = 0; idim < 3; ++idim)
                    for (int side = 0; side < 2; ++side)
                        nbor[(size_t)(idim * 2 + side) * ngridmax + new_ig - 1] = nbor_of_coarse(ic_coarse, idim, side);
                int idx = ic_coarse - 1;
                int x[3];
                x[2] = idx / (nx * ny);
                idx %= (nx * ny);
                x[1] = idx / nx;
                x[0] = idx % nx;
                for (int d = 0; d < 3; ++d) xg[(size_t)d * ngridmax + new_ig - 1] = (double)x[d] + 0.5;
            } else {
                int ig = ((idc - ncoarse) % ngridmax) + 1;
                int ic = ((idc - ncoarse) / ngridmax) + 1;
                int ign[7];
                c.nbor_grids(ig, ign);
                int icn[6];
                c.nbor_cells(ign, ic, icn, ig);
                for (int i = 0; i < 6; ++i) nbor[(size_t)i * ngridmax + new_ig - 1] = icn[i];
                for (int d = 0; d < 3; ++d) {
                    double dxl = 1.0 / (double)(1 << (ilevel - 1));
                    int ixyz[3] = {(ic - 1) & 1, ((ic - 1) & 2) >> 1, ((ic - 1) & 4) >> 2};
                    xg[(size_t)d * ngridmax + new_ig - 1] =
                        xg[(size_t)d * ngridmax + ig - 1] + (double)(ixyz[d] - 0.5) * dxl;
                }
            }
            int icn_c[6] = {0, 0, 0, 0, 0, 0};
            if (ilevel == 1) {
                for (int idim = 0; idim < 3; ++idim) {
                    icn_c[idim * 2] = nbor_of_coarse(ic_coarse, idim, 0);
                    icn_c[idim * 2 + 1] = nbor_of_coarse(ic_coarse, idim, 1);
                }
            } else {
                int igc = ((ic_coarse - 1 - ncoarse) % ngridmax) + 1;
                int icp = ((ic_coarse - 1 - ncoarse) / ngridmax) + 1;
                int ign[7];
                c.nbor_grids(igc, ign);
                c.nbor_cells(ign, icp, icn_c, igc);
            }
            double u1[7][64] = {{0}}, u2[8][64] = {{0}};
            for (int iv = 1; iv <= 5; ++iv) u1[0][iv - 1] = U(ic_coarse, iv);
            for (int i = 1; i <= 6; ++i) {
                int idn = icn_c[i - 1];
                if (idn > 0) {
                    for (int iv = 1; iv <= 5; ++iv) u1[i][iv - 1] = U(idn, iv);
                } else {
                    // bound_type is never filled: no reflective flip.
                    for (int iv = 1; iv <= 5; ++iv) u1[i][iv - 1] = u1[0][iv - 1];
                }
            }
            or_interpol_hydro(&hp, 5, &u1[0][0], &u2[0][0]);
            for (int isc = 1; isc <= 8; ++isc) {
                int child = ncoarse + (isc - 1) * ngridmax + new_ig;
                for (int iv = 1; iv <= 5; ++iv) {
                    U(child, iv) = u2[isc - 1][iv - 1];
                    N(child, iv) = u2[isc - 1][iv - 1];
                }
                cpu_map[child - 1] = 1;
            }
            add_to_level_list(new_ig, ilevel + 1);
        }
    }

    // TreeUpdater.cpp:152-193
    void remove_grid_fine(int ilevel) {
        if (ilevel >= nlevelmax) return;
        int ig = headl[ilevel];
        while (ig > 0) {
            int nxt = next[ig - 1];
            int fc = father[ig - 1];
            bool remove = false;
            if (fc > 0 && flag1[fc - 1] == 0) {
                remove = true;
                for (int ic = 1; ic <= 8; ++ic)
                    if (son[ncoarse + (ic - 1) * ngridmax + ig - 1] > 0) {
                        remove = false;
                        break;
                    }
            }
            if (remove) {
                for (int ic = 1; ic <= 8; ++ic) flag1[ncoarse + (ic - 1) * ngridmax + ig - 1] = 0;
                int p = prev[ig - 1], n = next[ig - 1];
                if (p > 0) next[p - 1] = n;
                else headl[ilevel] = n;
                if (n > 0) prev[n - 1] = p;
                else taill[ilevel] = p;
                numbl[ilevel]--;
                free_grid(ig);
            }
            ig = nxt;
        }
    }

    // TreeUpdater.cpp:195-215
    void restrict_fine(int ilevel) {
        if (ilevel + 1 < 0 || ilevel + 1 > nlevelmax) return;
        for (int ig = headl[ilevel + 1]; ig > 0;) {
            int nxt = next[ig - 1];
            int fc = father[ig - 1];
            if (fc > 0) {
                for (int iv = 1; iv <= 5; ++iv) {
                    double sum = 0;
                    for (int ic = 1; ic <= 8; ++ic) sum += U(cell(ig, ic), iv);
                    U(fc, iv) = sum / 8;
                }
            }
            ig = nxt;
        }
    }

    // TreeUpdater.cpp:217-267
    void smooth_fine(int ilevel) {
        std::vector<int> ok(ncell, 0);
        if (ilevel == 1) {
            or_mesh mv = view();
            Ctx c{&mv, &hp};
            for (int ic = 1; ic <= ncoarse; ++ic) {
                if (flag1[ic - 1] != 0) continue;
                int icn[6];
                c.nbor_cells_coarse(ic, icn);
                int nf = 0;
                for (int k = 0; k < 6; ++k)
                    if (icn[k] > 0 && icn[k] <= ncoarse && flag1[icn[k] - 1] == 1) nf++;
                if (nf > 0) ok[ic - 1] = 1;
            }
        } else {
            or_mesh mv = view();
            Ctx c{&mv, &hp};
            for (int ig = headl[ilevel]; ig > 0; ig = next[ig - 1]) {
                int ign[7];
                c.nbor_grids(ig, ign);
                for (int ic = 1; ic <= 8; ++ic) {
                    int idc = ncoarse + (ic - 1) * ngridmax + ig - 1;
                    if (flag1[idc] != 0) continue;
                    int icn[6];
                    nbor_cells_exact(ign, ic, icn);
                    int nf = 0;
                    for (int k = 0; k < 6; ++k)
                        if (icn[k] > 0 && icn[k] <= ncell && flag1[icn[k] - 1] == 1) nf++;
                    if (nf > 0) ok[idc] = 1;
                }
            }
        }
        for (int i = 0; i < ncell; ++i)
            if (ok[i] == 1) flag1[i] = 1;
    }

    // "ok" test of TreeUpdater.cpp:276-287 / 293-306: the cell has a son
    // grid with a refined or flagged child.
    bool child_active(int idc0) const {
        int child = son[idc0];
        if (child <= 0) return false;
        for (int ics = 1; ics <= 8; ++ics) {
            int s = ncoarse + (ics - 1) * ngridmax + child - 1;
            if (son[s] > 0 || flag1[s] == 1) return true;
        }
        return false;
    }

    // TreeUpdater.cpp:270-311
    void test_flag(int ilevel) {
        if (ilevel == 1) {
            for (int ic = 1; ic <= ncoarse; ++ic)
                if (child_active(ic - 1)) flag1[ic - 1] = 1;
        } else {
            for (int ig = headl[ilevel]; ig > 0; ig = next[ig - 1])
                for (int ic = 1; ic <= 8; ++ic) {
                    int idc = ncoarse + (ic - 1) * ngridmax + ig - 1;
                    if (child_active(idc)) flag1[idc] = 1;
                }
        }
    }

    struct CellState {
        double d = 0, v[3] = {0, 0, 0}, p = 0;
    };

    // TreeUpdater.cpp:342-384 (hydro, nener = 0)
    CellState cell_state(int idx) {
        CellState s;
        if (idx <= 0 || idx > ncell) return s;
        s.d = std::max(U(idx, 1), (double)1e-10);
        double v2 = 0.0;
        for (int idim = 0; idim < 3; ++idim) {
            s.v[idim] = U(idx, 2 + idim) / s.d;
            v2 += s.v[idim] * s.v[idim];
        }
        double ek = 0.5 * s.d * v2;
        double em = 0.0;
        double e_tot = U(idx, 5);
        double e_nonthermal = 0.0;
        s.p = std::max((e_tot - ek - em - e_nonthermal) * (gamma - 1.0), s.d * (double)1e-10);
        return s;
    }

    // TreeUpdater.cpp:386-391
    static double err_grad(double l, double c, double r, double floor) {
        return 2.0 * std::max(std::abs(r - c) / (r + c + floor), std::abs(c - l) / (c + l + floor));
    }

    // TreeUpdater.cpp:393-404
    double err_grad_u(double vl, double vc, double vr, double pl, double pc, double pr, double dl, double dc,
                      double dr) const {
        double floor = 1e-10;
        double cg = std::sqrt(std::max(gamma * pl / dl, floor * floor));
        double cm = std::sqrt(std::max(gamma * pc / dc, floor * floor));
        double cd = std::sqrt(std::max(gamma * pr / dr, floor * floor));
        return 2.0 * std::max(std::abs(vr - vc) / (cd + cm + std::abs(vr) + std::abs(vc) + floor),
                              std::abs(vc - vl) / (cm + cg + std::abs(vc) + std::abs(vl) + floor));
    }

    // Gradient test of one cell's three directions (TreeUpdater.cpp:463-480,
    // 497-526): true when any error exceeds its threshold.
    bool gradient_flag(const CellState& c, const CellState& l, const CellState& r, int idim, double ed_,
                       double ep_, double ev_) const {
        if (ed_ > 0.0 && err_grad(l.d, c.d, r.d, 1e-10) > ed_) return true;
        if (ep_ > 0.0 && err_grad(l.p, c.p, r.p, 1e-10) > ep_) return true;
        if (ev_ > 0.0 && err_grad_u(l.v[idim], c.v[idim], r.v[idim], l.p, c.p, r.p, l.d, c.d, r.d) > ev_) return true;
        return false;
    }

    // TreeUpdater.cpp:328-583 (eb2 <= 0: no magnetic criterion)
    void flag_fine(int ilevel, double ed_, double ep_, double ev_, int nexp, int icount, int nsub) {
        if (ilevel >= nlevelmax) return;
        const int lmin = levelmin;
        if (ilevel == 1) {
            for (int i = 0; i < ncoarse; ++i) flag1[i] = 0;
            for (int i = 1; i <= ncoarse; ++i) {
                cpu_map[i - 1] = 1;
                if (ilevel < lmin) flag1[i - 1] = 1;
                else if (child_active(i - 1)) flag1[i - 1] = 1;
            }
        } else {
            for (int ig = headl[ilevel]; ig > 0; ig = next[ig - 1])
                for (int ic = 1; ic <= 8; ++ic) {
                    int idc = ncoarse + (ic - 1) * ngridmax + ig - 1;
                    flag1[idc] = 0;
                    if (ilevel < lmin) flag1[idc] = 1;
                    else if (child_active(idc)) flag1[idc] = 1;
                }
        }
        if (ilevel >= lmin) smooth_fine(ilevel);
        if (ilevel >= lmin && (ed_ > 0.0 || ep_ > 0.0 || ev_ > 0.0)) {
            if (ilevel == 1) {
                for (int i = 1; i <= ncoarse; ++i) {
                    if (flag1[i - 1] != 0) continue;
                    bool ok = false;
                    CellState sc = cell_state(i);
                    for (int idim = 0; idim < 3 && !ok; ++idim) {
                        int nxd = (idim == 0) ? nx : ((idim == 1) ? ny : nz);
                        int nskip = (idim == 0) ? 1 : ((idim == 1) ? nx : nx * ny);
                        int il = i - nskip;
                        if (il < 1) il += nxd * nskip;
                        int ir = i + nskip;
                        if (ir > ncoarse) ir -= nxd * nskip;
                        ok = gradient_flag(sc, cell_state(il), cell_state(ir), idim, ed_, ep_, ev_);
                    }
                    if (ok) flag1[i - 1] = 1;
                }
            } else {
                or_mesh mv = view();
                Ctx c{&mv, &hp};
                for (int ig = headl[ilevel]; ig > 0; ig = next[ig - 1]) {
                    int ign[7];
                    c.nbor_grids(ig, ign);
                    for (int ic = 1; ic <= 8; ++ic) {
                        int idc = ncoarse + (ic - 1) * ngridmax + ig - 1;
                        if (flag1[idc] != 0) continue;
                        bool ok = false;
                        CellState sc = cell_state(idc + 1);
                        int icn[6];
                        nbor_cells_exact(ign, ic, icn);
                        for (int idim = 0; idim < 3 && !ok; ++idim) {
                            int il = icn[idim * 2];
                            if (il == 0) il = nbor[(size_t)(idim * 2) * ngridmax + ig - 1];
                            int ir = icn[idim * 2 + 1];
                            if (ir == 0) ir = nbor[(size_t)(idim * 2 + 1) * ngridmax + ig - 1];
                            CellState sl = (il > 0 && il <= ncell) ? cell_state(il) : sc;
                            CellState sr = (ir > 0 && ir <= ncell) ? cell_state(ir) : sc;
                            ok = gradient_flag(sc, sl, sr, idim, ed_, ep_, ev_);
                        }
                        if (ok) flag1[idc] = 1;
                    }
                }
            }
        }
        if (ilevel > lmin && icount < nsub) test_flag(ilevel);
        for (int i = 0; i < nexp; ++i) smooth_fine(ilevel);
        if (ilevel > lmin && icount < nsub) ensure_ref_rules(ilevel);
    }

    // TreeUpdater.cpp:593-636
    void ensure_ref_rules(int ilevel) {
        for (int ig = headl[ilevel]; ig > 0; ig = next[ig - 1]) {
            int fc = father[ig - 1];
            bool ok = true;
            if (fc > 0) {
                int nb[27];
                neighbors27(fc, nb);
                for (int i = 0; i < 27 && ok; ++i) {
                    if (nb[i] == 0) ok = false;
                    else if (nb[i] > 0 && son[nb[i] - 1] == 0) ok = false;
                }
            } else {
                ok = false;
            }
            if (!ok)
                for (int ic = 1; ic <= 8; ++ic) flag1[ncoarse + (ic - 1) * ngridmax + ig - 1] = 0;
        }
    }

    // Initializer.cpp:19-157 for ilevel (region ICs) or the slab.
    void apply_ic(int idc) {
        double xc[3];
        cell_center(idc, xc);
        double x = xc[0], y = xc[1], z = xc[2];
        auto set = [&](double d, double u, double v, double w, double p) {
            U(idc, 1) = d;
            U(idc, 2) = d * u;
            U(idc, 3) = d * v;
            U(idc, 4) = d * w;
            double v2 = u * u;
            v2 += v * v;
            v2 += w * w;
            double e_kin = 0.5 * d * v2;
            double e_int = p / (gamma - 1.0);
            double e_nonthermal = 0.0;
            U(idc, 5) = e_kin + e_int + e_nonthermal;
        };
        if (ic_kind == 1) {
            const double pi = 3.14159265358979323846;
            bool in = std::abs(y - 0.5) < 0.25;
            set(in ? 2.0 : 1.0, in ? 0.5 : -0.5, kh_amp * std::sin(4.0 * pi * x), 0.0, 2.5);
            return;
        }
        for (const Region& r : regions) {
            bool in = false;
            if (r.type == 0) {
                in = std::abs(x - r.xc) < 0.5 * r.lx && std::abs(y - r.yc) < 0.5 * r.ly && std::abs(z - r.zc) < 0.5 * r.lz;
            } else {
                double r2 = std::pow(x - r.xc, 2) + std::pow(y - r.yc, 2) + std::pow(z - r.zc, 2);
                in = r2 < std::pow(r.lx, 2);
            }
            if (in) set(r.d, r.u, r.v, r.w, r.p);
        }
    }

    // Initializer.cpp:19-29 (apply_all over levels 0..nlevelmax)
    void apply_all() {
        for (int i = 1; i <= ncoarse; ++i) apply_ic(i);
        for (int il = 1; il <= nlevelmax; ++il)
            for (int ig = headl[il]; ig > 0; ig = next[ig - 1])
                for (int ic = 1; ic <= 8; ++ic) apply_ic(cell(ig, ic));
    }

    void legacy_init_flow() {
        apply_all();
        for (int il = nlevelmax - 1; il >= 0; --il) restrict_fine(il);
    }

    // Simulation.cpp:160-255 (the refinement part of initialize)
    void initialize() {
        apply_all();
        nsubcycle.assign(33, 1);
        for (int i = levelmin; i < 33; ++i) nsubcycle[i] = 2;
        dtnew.assign(33, 0.0);
        dtold.assign(33, 0.0);
        const int lmin = levelmin, lmax = nlevelmax;
        for (int il_ref = 1; il_ref <= lmin; ++il_ref)
            for (int il = 1; il <= lmax; ++il) {
                flag_fine(il, ed, ep, ev, nexpand, 1, 2);
                make_grid_fine(il);
            }
        for (int il_ref = lmin + 1; il_ref <= lmax; ++il_ref) {
            legacy_init_flow();
            for (int il = 1; il <= lmax; ++il) {
                flag_fine(il, ed, ep, ev, nexpand, 1, 2);
                make_grid_fine(il);
            }
            if (count(il_ref) == 0) break;
        }
        nstep = 0;
        legacy_init_flow();
    }

    long long leaf_cells(int ilevel) const {
        long long n = 0;
        for (int ig = headl[ilevel]; ig > 0; ig = next[ig - 1])
            for (int ic = 1; ic <= 8; ++ic)
                if (son[cell(ig, ic) - 1] == 0) n++;
        return n;
    }

    // Simulation.cpp:522-635 (no MHD, RT, turbulence, sinks, stars or cooling)
    void amr_step(int ilevel, int icount) {
        if (ilevel > nlevelmax) return;
        if (count(ilevel) == 0 && ilevel > 0) return;
        if (levelmin < nlevelmax && (ilevel == levelmin || icount > 1)) {
            for (int i = ilevel; i <= nlevelmax; ++i) make_grid_fine(i + 1);
            for (int i = nlevelmax; i >= ilevel; --i) remove_grid_fine(i + 1);
        }
        double dx = boxlen / (double)(nx * (1 << ilevel));
        dtold[ilevel] = dtnew[ilevel];
        or_mesh mv = view();
        dtnew[ilevel] = or_courant(&mv, ilevel, dx, gamma, courant_factor, hp.boxlen);
        if (ilevel > levelmin)
            dtnew[ilevel] = std::min(dtnew[ilevel], dtnew[ilevel - 1] / (double)nsubcycle[ilevel - 1]);
        double dt = dtnew[ilevel];
        or_set_unew(&mv, ilevel);
        int nsub = (ilevel < (int)nsubcycle.size()) ? nsubcycle[ilevel] : 1;
        if (ilevel < nlevelmax) {
            if (count(ilevel + 1) > 0) {
                for (int i = 1; i <= nsub; ++i) amr_step(ilevel + 1, i);
            } else {
                dtold[ilevel + 1] = dtnew[ilevel] / (double)nsub;
                dtnew[ilevel + 1] = dtnew[ilevel] / (double)nsub;
                t += dt;
                nstep++;
            }
        } else {
            t += dt;
            nstep++;
        }
        mv = view();
        cell_updates += leaf_cells(ilevel);
        if (or_godunov_fine(&mv, &hp, ilevel, dt, dx) != 0) status = -3;
        or_set_uold(&mv, ilevel);
        if (ilevel < nlevelmax) restrict_fine(ilevel);
        flag_fine(ilevel + 1, ed, ep, ev, nexpand, icount, ilevel > 0 ? nsubcycle[ilevel - 1] : 1);
    }

    // Simulation.cpp:355-395: one coarse step of the run loop.
    void step_coarse() {
        if (levelmin < nlevelmax)
            for (int il = 1; il <= levelmin - 1; ++il) make_grid_fine(il);
        amr_step(levelmin, 1);
        for (int il = levelmin; il >= 1; --il) flag_fine(il, ed, ep, ev, nexpand, 2, 2);
        if (levelmin < nlevelmax)
            for (int il = levelmin - 1; il >= 0; --il) restrict_fine(il);
    }
};

}  // namespace

extern "C" {

void* or_amr_create(const or_amr_config* cfg) {
    if (!cfg || cfg->levelmin < 1 || cfg->levelmax < cfg->levelmin) return nullptr;
    Amr* a = new Amr();
    a->allocate(cfg->nx, cfg->ny, cfg->nz, cfg->ngridmax, cfg->levelmax);
    a->boxlen = cfg->boxlen;
    a->gamma = cfg->gamma;
    a->levelmin = cfg->levelmin;
    a->courant_factor = cfg->courant_factor;
    a->ed = cfg->err_grad_d;
    a->ep = cfg->err_grad_p;
    a->ev = cfg->err_grad_v;
    a->nexpand = cfg->nexpand;
    a->hp = cfg->hydro;
    a->ic_kind = cfg->ic_kind;
    a->kh_amp = cfg->kh_amp;
    for (int i = 0; i < cfg->nreg; ++i)
        a->regions.push_back({cfg->reg_type[i], cfg->x_center[i], cfg->y_center[i], cfg->z_center[i],
                              cfg->length_x[i], cfg->length_y[i], cfg->length_z[i], cfg->d_region[i],
                              cfg->u_region[i], cfg->v_region[i], cfg->w_region[i], cfg->p_region[i]});
    a->initialize();
    if (a->status != 0) {
        delete a;
        return nullptr;
    }
    return a;
}

int or_amr_step(void* h, int nsteps) {
    Amr* a = (Amr*)h;
    for (int k = 0; k < nsteps && a->status == 0; ++k) a->step_coarse();
    return a->status;
}

void or_amr_get(void* h, or_amr_state* s) {
    Amr* a = (Amr*)h;
    s->ncoarse = a->ncoarse;
    s->ngridmax = a->ngridmax;
    s->ncell = a->ncell;
    s->nlevelmax = a->nlevelmax;
    s->father = a->father.data();
    s->son = a->son.data();
    s->nbor = a->nbor.data();
    s->next = a->next.data();
    s->headl = a->headl.data();
    s->numbl = a->numbl.data();
    s->flag1 = a->flag1.data();
    s->xg = a->xg.data();
    s->uold = a->uold.data();
    s->unew = a->unew.data();
    s->t = a->t;
    s->nstep = a->nstep;
    s->dtnew = a->dtnew.data();
    s->cell_updates = a->cell_updates;
    s->status = a->status;
}

void or_amr_destroy(void* h) { delete (Amr*)h; }

// dump_snapshot of the run (Simulation.cpp:678-732): t, nstep, levelmin,
// levelmax and boxlen from the run, the rest from *info; gamma from the run.
int or_amr_dump(void* h, const or_snapshot_info* info, int iout, const char* root) {
    Amr* a = (Amr*)h;
    or_snapshot_info in = *info;
    in.t = a->t;
    in.nstep = in.nstep_coarse = a->nstep;
    in.iout = iout;
    in.levelmin = a->levelmin;
    in.levelmax = a->nlevelmax;
    in.boxlen = a->boxlen;
    or_snapshot_tree t{};
    t.nx = a->nx;
    t.ny = a->ny;
    t.nz = a->nz;
    t.ncoarse = a->ncoarse;
    t.ngridmax = a->ngridmax;
    t.ncell = a->ncell;
    t.nlevelmax = a->nlevelmax;
    t.ncpu = 1;
    t.nboundary = 0;
    t.gamma = a->gamma;
    t.headl = a->headl.data();
    t.taill = a->taill.data();
    t.numbl = a->numbl.data();
    t.headf = a->headf;
    t.tailf = a->tailf;
    t.numbf = a->numbf;
    t.next = a->next.data();
    t.prev = a->prev.data();
    t.father = a->father.data();
    t.nbor = a->nbor.data();
    t.son = a->son.data();
    t.flag1 = a->flag1.data();
    t.cpu_map = a->cpu_map.data();
    t.xg = a->xg.data();
    return or_dump_snapshot(&t, a->uold.data(), &in, root);
}

}  // extern "C"
