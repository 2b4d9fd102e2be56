// rh_oracle_snapshot.cpp -- CPU restatement of the reference's snapshot
// writer (test infrastructure only: loaded by tests/ to check the library's
// files byte for byte, never by the product).
//
// Follows Simulation::dump_snapshot (Simulation.cpp:678-732) and
// RamsesWriter::write_amr (RamsesWriter.cpp:18-144), write_hydro (:146-235),
// write_header (:237-266), write_header_file (:268-273),
// write_hydro_descriptor (:275-313), write_particles_descriptor (:428-441)
// and write_grav (:443-473), for NDIM = 3, no MHD, nener = 0, nvar = 5 and
// rank 0 of 1.  The info file is written through an ostream with the same
// sticky std::scientific / setprecision(16) the reference uses.
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <string>
#include <vector>

#include "rh_oracle_snapshot.h"

namespace {

struct Writer {
    std::ofstream out;
    explicit Writer(const std::string& p) : out(p, std::ios::binary) {}
    // One Fortran unformatted record (RamsesWriter.cpp:20-25).
    void rec(const void* data, size_t size) {
        const int32_t s = (int32_t)size;
        out.write((const char*)&s, 4);
        out.write((const char*)data, (std::streamsize)size);
        out.write((const char*)&s, 4);
    }
    template <typename T> void rec(const std::vector<T>& v) { rec(v.data(), v.size() * sizeof(T)); }
    void i(int32_t v) { rec(&v, 4); }
    void d(double v) { rec(&v, 8); }
};

struct View {
    const or_snapshot_tree* t;
    int slot(int icpu, int il) const { return il * (t->ncpu + t->nboundary) + icpu - 1; }
    int numbl(int icpu, int il) const { return t->numbl[slot(icpu, il)]; }
    int headl(int icpu, int il) const { return t->headl[slot(icpu, il)]; }
    int taill(int icpu, int il) const { return t->taill[slot(icpu, il)]; }
    int flag1(long i) const { return t->flag1 ? t->flag1[i] : 0; }
    int cpu(long i) const { return t->cpu_map ? t->cpu_map[i] : 1; }
    double phi(long i) const { return t->phi ? t->phi[i] : 0.0; }
    // AmrGrid::f(icell, idim) with a 1-based cell id (AmrGrid.hpp:33).
    double f(long icell, int idim) const { return t->f ? t->f[(long)(idim - 1) * t->ncell + icell - 1] : 0.0; }
    std::vector<int> grids(int il) const {
        std::vector<int> g;
        for (int ig = headl(1, il); ig > 0; ig = t->next[ig - 1]) g.push_back(ig);
        return g;
    }
};

bool write_amr(const View& v, const or_snapshot_info& in, const std::string& path) {
    const or_snapshot_tree& t = *v.t;
    Writer w(path);
    w.i(t.ncpu);
    w.i(3);
    int32_t nxyz[3] = {t.nx, t.ny, t.nz};
    w.rec(nxyz, 12);
    w.i(t.nlevelmax);
    w.i(t.ngridmax);
    w.i(t.nboundary);
    int32_t ngrid_tot = 0;
    for (int l = 1; l <= t.nlevelmax; ++l)
        for (int c = 1; c <= t.ncpu; ++c) ngrid_tot += v.numbl(c, l);
    w.i(ngrid_tot);
    w.d(in.boxlen);
    int32_t nout[3] = {in.noutput, in.iout, in.iout};
    w.rec(nout, 12);
    w.rec(std::vector<double>(in.noutput, 0.0));
    w.rec(std::vector<double>(in.noutput, 1.0));
    w.d(in.t);
    w.rec(std::vector<double>(t.nlevelmax, 0.0));
    w.rec(std::vector<double>(t.nlevelmax, 0.0));
    int32_t st[2] = {in.nstep, in.nstep_coarse};
    w.rec(st, 8);
    double a[3] = {in.einit, in.mass_tot_0, in.rho_tot};
    w.rec(a, 24);
    double b[7] = {in.omega_m, in.omega_l, in.omega_k, in.omega_b, in.h0, in.aexp_ini, in.boxlen_ini};
    w.rec(b, 56);
    double c[5] = {in.aexp, in.hexp, in.aexp_old, in.epot_tot_int, in.epot_tot_old};
    w.rec(c, 40);
    w.d(in.mass_sph);
    const int n_cl = t.ncpu * t.nlevelmax;
    std::vector<int32_t> hl(n_cl), tl(n_cl), nl(n_cl);
    for (int l = 1; l <= t.nlevelmax; ++l)
        for (int cc = 1; cc <= t.ncpu; ++cc) {
            const int k = (l - 1) * t.ncpu + cc - 1;
            hl[k] = v.headl(cc, l);
            tl[k] = v.taill(cc, l);
            nl[k] = v.numbl(cc, l);
        }
    w.rec(hl);
    w.rec(tl);
    w.rec(nl);
    std::vector<int32_t> numbtot(10 * t.nlevelmax, 0);
    for (int l = 1; l <= t.nlevelmax; ++l) numbtot[(l - 1) * 10] = v.numbl(1, l);
    w.rec(numbtot);
    if (t.nboundary > 0)
        for (int k = 0; k < 3; ++k) w.rec(std::vector<int32_t>(t.nboundary * t.nlevelmax, 0));
    int32_t mem[5] = {t.headf, t.tailf, t.numbf, 0, 0};
    w.rec(mem, 20);
    char ord[128];
    std::memset(ord, 0, sizeof ord);
    std::memcpy(ord, "hilbert", 7);
    w.rec(ord, 128);
    std::vector<double> key(t.ncpu + 1, 0.0);
    key[t.ncpu] = 1.0;
    w.rec(key);
    std::vector<int32_t> cs(t.ncoarse), cf(t.ncoarse), cc(t.ncoarse);
    for (int i = 0; i < t.ncoarse; ++i) {
        cs[i] = t.son[i];
        cf[i] = v.flag1(i);
        cc[i] = v.cpu(i);
    }
    w.rec(cs);
    w.rec(cf);
    w.rec(cc);
    for (int il = 1; il <= t.nlevelmax; ++il) {
        for (int ib = 1; ib <= t.nboundary + t.ncpu; ++ib) {
            const int ncache = ib <= t.ncpu ? v.numbl(ib, il) : 0;
            if (ncache <= 0) continue;
            const std::vector<int> g = v.grids(il);
            const long G = t.ngridmax;
            std::vector<int32_t> iv(g.size());
            std::vector<double> dv(g.size());
            auto grid_int = [&](auto fn) {
                for (size_t k = 0; k < g.size(); ++k) iv[k] = fn(g[k]);
                w.rec(iv);
            };
            grid_int([](int ig) { return ig; });
            grid_int([&](int ig) { return t.next[ig - 1]; });
            grid_int([&](int ig) { return t.prev[ig - 1]; });
            for (int d = 0; d < 3; ++d) {
                for (size_t k = 0; k < g.size(); ++k) dv[k] = t.xg[d * G + g[k] - 1];
                w.rec(dv);
            }
            grid_int([&](int ig) { return t.father[ig - 1]; });
            for (int d = 0; d < 6; ++d) grid_int([&](int ig) { return t.nbor[d * G + ig - 1]; });
            for (int k = 0; k < 8; ++k) grid_int([&](int ig) { return t.son[t.ncoarse + k * G + ig - 1]; });
            for (int k = 0; k < 8; ++k) grid_int([&](int ig) { return v.cpu(t.ncoarse + k * G + ig - 1); });
            for (int k = 0; k < 8; ++k) grid_int([&](int ig) { return v.flag1(t.ncoarse + k * G + ig - 1); });
        }
    }
    return (bool)w.out;
}

bool write_hydro(const View& v, const double* uold, const std::string& path) {
    const or_snapshot_tree& t = *v.t;
    auto U = [&](long icell, int ivar) { return uold[(long)(ivar - 1) * t.ncell + icell - 1]; };
    Writer w(path);
    w.i(t.ncpu);
    w.i(5);
    w.i(3);
    w.i(t.nlevelmax);
    w.i(t.nboundary);
    w.d(t.gamma);
    const double smallr = 1e-10;
    for (int il = 1; il <= t.nlevelmax; ++il) {
        for (int icpu = 1; icpu <= t.ncpu + t.nboundary; ++icpu) {
            const int ncache = icpu <= t.ncpu ? v.numbl(icpu, il) : 0;
            w.i(il);
            w.i(ncache);
            if (ncache <= 0) continue;
            const std::vector<int> g = v.grids(il);
            std::vector<double> tmp(ncache);
            for (int ic = 0; ic < 8; ++ic) {
                for (int ivar = 1; ivar <= 5; ++ivar) {
                    for (int i = 0; i < ncache; ++i) {
                        const long idc = (long)t.ncoarse + (long)ic * t.ngridmax + g[i];
                        const double d = std::max(U(idc, 1), smallr);
                        if (ivar == 1) {
                            tmp[i] = d;
                        } else if (ivar <= 4) {
                            tmp[i] = U(idc, ivar) / d;
                        } else {
                            double v2 = 0;
                            for (int j = 1; j <= 3; ++j) {
                                const double vj = U(idc, 1 + j) / d;
                                v2 += vj * vj;
                            }
                            const double e_nonthermal = 0.0;
                            tmp[i] = std::max((U(idc, 5) - 0.5 * d * v2 - e_nonthermal) * (t.gamma - 1.0), 0.0);
                        }
                    }
                    w.rec(tmp);
                }
            }
        }
    }
    return (bool)w.out;
}

bool write_grav(const View& v, const std::string& path) {
    const or_snapshot_tree& t = *v.t;
    Writer w(path);
    w.i(t.ncpu);
    w.i(4);
    int32_t ngrid_tot = 0;
    for (int l = 1; l <= t.nlevelmax; ++l) ngrid_tot += v.numbl(1, l);
    w.i(ngrid_tot);
    w.i(t.nlevelmax);
    w.i(t.nboundary);
    std::vector<double> tmp(t.ncoarse);
    for (int i = 1; i <= t.ncoarse; ++i) tmp[i - 1] = v.phi(i - 1);
    w.rec(tmp);
    for (int d = 1; d <= 3; ++d) {
        for (int i = 1; i <= t.ncoarse; ++i) tmp[i - 1] = v.f(i, d);
        w.rec(tmp);
    }
    for (int il = 1; il <= t.nlevelmax; ++il) {
        for (int icpu = 1; icpu <= t.ncpu + t.nboundary; ++icpu) {
            const int ncache = icpu <= t.ncpu ? v.numbl(icpu, il) : 0;
            if (ncache <= 0) continue;
            w.i(il);
            w.i(ncache);
            const std::vector<int> g = v.grids(il);
            std::vector<double> p(ncache);
            for (int ic = 0; ic < 8; ++ic) {
                for (int i = 0; i < ncache; ++i) p[i] = v.phi((long)t.ncoarse + (long)ic * t.ngridmax + g[i] - 1);
                w.rec(p);
                for (int d = 1; d <= 3; ++d) {
                    for (int i = 0; i < ncache; ++i) p[i] = v.f((long)t.ncoarse + (long)ic * t.ngridmax + g[i], d);
                    w.rec(p);
                }
            }
        }
    }
    return (bool)w.out;
}

bool write_info(const View& v, const or_snapshot_info& in, const std::string& path) {
    const or_snapshot_tree& t = *v.t;
    std::ofstream file(path);
    file << "ncpu        = " << t.ncpu << std::endl;
    file << "ndim        = " << 3 << std::endl;
    file << "nx          = " << t.nx << std::endl;
    file << "ny          = " << t.ny << std::endl;
    file << "nz          = " << t.nz << std::endl;
    file << "ncoarse     = " << t.ncoarse << std::endl;
    file << "nboundary   = " << t.nboundary << std::endl;
    file << "levelmin    = " << in.levelmin << std::endl;
    file << "levelmax    = " << in.levelmax << std::endl;
    file << "ngridmax    = " << t.ngridmax << std::endl;
    file << "nstep       = " << in.nstep << std::endl;
    file << "nstep_coarse= " << in.nstep_coarse << std::endl;
    file << "noutput     = " << in.noutput << std::endl;
    file << "time        = " << std::scientific << std::setprecision(16) << in.t << std::endl;
    file << "boxlen      = " << std::scientific << std::setprecision(16) << in.boxlen / (double)t.nx << std::endl;
    file << "omega_m     = " << in.omega_m << std::endl;
    file << "omega_l     = " << in.omega_l << std::endl;
    file << "omega_k     = " << in.omega_k << std::endl;
    file << "omega_b     = " << in.omega_b << std::endl;
    file << "h0          = " << in.h0 << std::endl;
    file << "aexp        = " << in.aexp << std::endl;
    file << "unit_l      = " << in.units_length << std::endl;
    file << "unit_d      = " << in.units_density << std::endl;
    file << "unit_t      = " << in.units_time << std::endl;
    file << "nener       = " << in.nener << std::endl;
    file << "nvar        = " << 5 << std::endl;
    file << "ordering type= hilbert" << std::endl;
    return (bool)file;
}

bool write_lines(const std::string& path, const std::vector<std::string>& lines) {
    std::ofstream file(path);
    for (const auto& l : lines) file << l << std::endl;
    return (bool)file;
}

std::string numbered(const std::string& dir, const char* prefix, int iout, const char* ext, bool rank) {
    std::stringstream ss;
    ss << dir << "/" << prefix << "_" << std::setfill('0') << std::setw(5) << iout << ext;
    if (rank) ss << std::setfill('0') << std::setw(5) << 1;
    return ss.str();
}

}  // namespace

extern "C" int or_dump_snapshot(const or_snapshot_tree* t, const double* uold, const or_snapshot_info* info,
                                const char* root) {
    std::stringstream ssd;
    ssd << root << "/output_" << std::setfill('0') << std::setw(5) << info->iout;
    const std::string dir = ssd.str();
    if (mkdir(dir.c_str(), 0777) != 0 && errno != EEXIST) return -1;
    const View v{t};
    const int iout = info->iout;
    bool ok = write_amr(v, *info, numbered(dir, "amr", iout, ".out", true));
    ok = ok && write_hydro(v, uold, numbered(dir, "hydro", iout, ".out", true));
    ok = ok && write_grav(v, numbered(dir, "grav", iout, ".out", true));
    ok = ok && write_lines(numbered(dir, "header", iout, ".txt", false),
                           {"# Family Count", "undefined " + std::to_string(t->npart), "Particle fields"});
    ok = ok && write_lines(dir + "/hydro_file_descriptor.txt",
                           {"# version: 1", "1, density, double", "2, velocity_x, double", "3, velocity_y, double",
                            "4, velocity_z, double", "5, pressure, double"});
    ok = ok && write_lines(dir + "/part_file_descriptor.txt",
                           {"# version: 1", "1, position_x, double", "2, position_y, double",
                            "3, position_z, double", "4, velocity_x, double", "5, velocity_y, double",
                            "6, velocity_z, double", "7, mass, double", "8, identity, int", "9, level, int",
                            "10, family, int", "11, tag, int", "12, birth_time, double",
                            "13, metallicity, double"});
    ok = ok && write_info(v, *info, numbered(dir, "info", iout, ".txt", false));
    return ok ? 0 : -1;
}
