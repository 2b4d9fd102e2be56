// rh_snapshot.cpp -- RAMSES snapshots of the bound mesh (rhydro_snapshot.h).
// The tree, gravity and text files are written from the host arrays, where
// the tree lives; the hydro file is packed level by level on the GPU from the
// device state (k_pack_hydro: primitives in file order with their record
// frames) and written with one fwrite per level.  Host code.
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "rh_context.h"
#include "rh_guard.h"
#include "rh_kernels.h"
#include "rh_snapshot.h"
#include "rhydro_snapshot.h"

namespace rhi {

namespace {

// A binary file of Fortran unformatted records: int32 byte count, payload,
// int32 byte count (RamsesWriter.cpp:20-25).
struct Records {
    FILE* f = nullptr;
    bool ok = false;
    explicit Records(const std::string& path) : f(std::fopen(path.c_str(), "wb")), ok(f != nullptr) {}
    ~Records() {
        if (f) std::fclose(f);
    }
    void raw(const void* p, size_t n) {
        if (ok && n > 0 && std::fwrite(p, 1, n, f) != n) ok = false;
    }
    void rec(const void* p, size_t n) {
        const int32_t s = (int32_t)n;
        raw(&s, 4);
        raw(p, n);
        raw(&s, 4);
    }
    template <typename T> void vec(const std::vector<T>& v) { rec(v.data(), v.size() * sizeof(T)); }
    void i32(int v) {
        const int32_t x = v;
        rec(&x, 4);
    }
    void f64(double v) { rec(&v, 8); }
    bool finish() {
        if (!f) return false;
        const bool good = ok && std::fclose(f) == 0;
        f = nullptr;
        return good;
    }
};

int nd(const rh_snapshot_tree* t) { return t->ncpu + t->nboundary; }
int numbl(const rh_snapshot_tree* t, int icpu, int il) { return t->numbl[(size_t)il * nd(t) + icpu - 1]; }
int headl(const rh_snapshot_tree* t, int icpu, int il) { return t->headl[(size_t)il * nd(t) + icpu - 1]; }
int taill(const rh_snapshot_tree* t, int icpu, int il) { return t->taill[(size_t)il * nd(t) + icpu - 1]; }
int flag1(const rh_snapshot_tree* t, size_t i) { return t->flag1 ? t->flag1[i] : 0; }
int cpu_map(const rh_snapshot_tree* t, size_t i) { return t->cpu_map ? t->cpu_map[i] : 1; }

std::string out_name(const std::string& dir, const char* prefix, int iout, const char* ext, bool rank) {
    char b[64];
    std::snprintf(b, sizeof b, "/%s_%05d%s%s", prefix, iout, ext, rank ? "00001" : "");
    return dir + b;
}

}  // namespace

int check_tree(rh_ctx* c, const rh_snapshot_tree* t) {
    if (!t) return fail(c, RH_ERR_ARG, "snapshot: tree is NULL");
    if (t->ncpu != 1) return fail(c, RH_ERR_UNSUPPORTED, "snapshot: only ncpu = 1 is supported");
    if (t->nboundary < 0 || t->nlevelmax < 1 || t->ngridmax < 1 || t->nx < 1 || t->ny < 1 || t->nz < 1 ||
        t->ncoarse != t->nx * t->ny * t->nz || t->ncell != t->ncoarse + 8 * t->ngridmax)
        return fail(c, RH_ERR_ARG, "snapshot: inconsistent tree dimensions");
    if (!t->headl || !t->taill || !t->numbl || !t->next || !t->prev || !t->father || !t->nbor || !t->son || !t->xg)
        return fail(c, RH_ERR_ARG, "snapshot: a required tree array is NULL");
    return RH_OK;
}

// Grids of level list il of cpu 1 in list order; their number must be
// numbl(1, il), as the reference's writers assume.
int level_grids(rh_ctx* c, const rh_snapshot_tree* t, int il, std::vector<int>& g) {
    g.clear();
    const int n = numbl(t, 1, il);
    int ig = headl(t, 1, il);
    while (ig > 0) {
        if (ig > t->ngridmax || (int)g.size() >= n)
            return fail(c, RH_ERR_ARG, "snapshot: level list " + std::to_string(il) + " does not match numbl");
        g.push_back(ig);
        ig = t->next[ig - 1];
    }
    if ((int)g.size() != n)
        return fail(c, RH_ERR_ARG, "snapshot: level list " + std::to_string(il) + " does not match numbl");
    return RH_OK;
}

// RamsesWriter::write_amr (RamsesWriter.cpp:18-144).
int write_amr_file(rh_ctx* c, const rh_snapshot_tree* t, const rh_snapshot_info* in, const std::string& path) {
    int st = check_tree(c, t);
    if (st) return st;
    const int L = t->nlevelmax, ncpu = t->ncpu, nb = t->nboundary;
    Records r(path);
    if (!r.ok) return fail(c, RH_ERR_ARG, "snapshot: cannot open " + path);
    r.i32(ncpu);
    r.i32(3);
    const int32_t nxyz[3] = {t->nx, t->ny, t->nz};
    r.rec(nxyz, 12);
    r.i32(L);
    r.i32(t->ngridmax);
    r.i32(nb);
    int ngrid_tot = 0;
    for (int l = 1; l <= L; ++l)
        for (int k = 1; k <= ncpu; ++k) ngrid_tot += numbl(t, k, l);
    r.i32(ngrid_tot);
    r.f64(in->boxlen);
    const int32_t nout[3] = {in->noutput, in->iout, in->iout};
    r.rec(nout, 12);
    r.vec(std::vector<double>((size_t)std::max(in->noutput, 0), 0.0));  // tout
    r.vec(std::vector<double>((size_t)std::max(in->noutput, 0), 1.0));  // aout
    r.f64(in->t);
    r.vec(std::vector<double>((size_t)L, 0.0));  // dtold
    r.vec(std::vector<double>((size_t)L, 0.0));  // dtnew
    const int32_t steps[2] = {in->nstep, in->nstep_coarse};
    r.rec(steps, 8);
    const double c1[3] = {in->einit, in->mass_tot_0, in->rho_tot};
    r.rec(c1, sizeof c1);
    const double c2[7] = {in->omega_m, in->omega_l, in->omega_k, in->omega_b, in->h0, in->aexp_ini, in->boxlen_ini};
    r.rec(c2, sizeof c2);
    const double c3[5] = {in->aexp, in->hexp, in->aexp_old, in->epot_tot_int, in->epot_tot_old};
    r.rec(c3, sizeof c3);
    r.f64(in->mass_sph);
    std::vector<int32_t> lst((size_t)ncpu * L);
    for (int l = 1; l <= L; ++l)
        for (int k = 1; k <= ncpu; ++k) lst[(size_t)(l - 1) * ncpu + k - 1] = headl(t, k, l);
    r.vec(lst);
    for (int l = 1; l <= L; ++l)
        for (int k = 1; k <= ncpu; ++k) lst[(size_t)(l - 1) * ncpu + k - 1] = taill(t, k, l);
    r.vec(lst);
    for (int l = 1; l <= L; ++l)
        for (int k = 1; k <= ncpu; ++k) lst[(size_t)(l - 1) * ncpu + k - 1] = numbl(t, k, l);
    r.vec(lst);
    std::vector<int32_t> numbtot((size_t)10 * L, 0);
    for (int l = 1; l <= L; ++l) numbtot[(size_t)(l - 1) * 10] = numbl(t, 1, l);
    r.vec(numbtot);
    if (nb > 0) {
        const std::vector<int32_t> zb((size_t)nb * L, 0);
        r.vec(zb);  // headb
        r.vec(zb);  // tailb
        r.vec(zb);  // numbb
    }
    const int32_t mem[5] = {t->headf, t->tailf, t->numbf, 0, 0};
    r.rec(mem, 20);
    char ordering[128] = {0};
    std::strncpy(ordering, "hilbert", sizeof ordering);
    r.rec(ordering, 128);
    std::vector<double> bound_key((size_t)ncpu + 1, 0.0);
    bound_key[ncpu] = 1.0;
    r.vec(bound_key);
    std::vector<int32_t> tmp((size_t)t->ncoarse);
    for (int i = 0; i < t->ncoarse; ++i) tmp[i] = t->son[i];
    r.vec(tmp);
    for (int i = 0; i < t->ncoarse; ++i) tmp[i] = flag1(t, i);
    r.vec(tmp);
    for (int i = 0; i < t->ncoarse; ++i) tmp[i] = cpu_map(t, i);
    r.vec(tmp);

    std::vector<int> g;
    std::vector<double> dv;
    for (int il = 1; il <= L; ++il) {
        // Domains ncpu+1..ncpu+nboundary have no grids here: nothing written.
        if ((st = level_grids(c, t, il, g))) return st;
        const size_t n = g.size();
        if (n == 0) continue;
        const size_t G = (size_t)t->ngridmax;
        tmp.resize(n);
        dv.resize(n);
        auto grid_rec = [&](const int* a, size_t off) {
            for (size_t i = 0; i < n; ++i) tmp[i] = a[off + g[i] - 1];
            r.vec(tmp);
        };
        for (size_t i = 0; i < n; ++i) tmp[i] = g[i];
        r.vec(tmp);
        grid_rec(t->next, 0);
        grid_rec(t->prev, 0);
        for (int d = 0; d < 3; ++d) {
            for (size_t i = 0; i < n; ++i) dv[i] = t->xg[d * G + g[i] - 1];
            r.vec(dv);
        }
        grid_rec(t->father, 0);
        for (int d = 0; d < 6; ++d) grid_rec(t->nbor, d * G);
        for (int k = 0; k < 8; ++k) grid_rec(t->son, t->ncoarse + k * G);
        for (int k = 0; k < 8; ++k) {
            for (size_t i = 0; i < n; ++i) tmp[i] = cpu_map(t, t->ncoarse + k * G + g[i] - 1);
            r.vec(tmp);
        }
        for (int k = 0; k < 8; ++k) {
            for (size_t i = 0; i < n; ++i) tmp[i] = flag1(t, t->ncoarse + k * G + g[i] - 1);
            r.vec(tmp);
        }
    }
    if (!r.finish()) return fail(c, RH_ERR_ARG, "snapshot: write failed on " + path);
    return RH_OK;
}

// RamsesWriter::write_hydro (RamsesWriter.cpp:146-235) from the device state.
int write_hydro_file(rh_ctx* c, const rh_snapshot_tree* t, const rh_snapshot_info* in, const std::string& path) {
    (void)in;
    int st = check_bound(c);
    if (st) return st;
    if ((st = check_tree(c, t))) return st;
    if (c->mesh.nvar != 5) return fail(c, RH_ERR_UNSUPPORTED, "snapshot: nvar must be 5");
    if (t->ncoarse != c->mesh.ncoarse || t->ngridmax != c->mesh.ngridmax || t->ncell != c->mesh.ncell)
        return fail(c, RH_ERR_ARG, "snapshot: the tree does not describe the bound mesh");
    if ((st = pre_call(c))) return st;

    const int L = t->nlevelmax, nb = t->nboundary;
    std::vector<std::vector<int>> lists((size_t)L + 1);
    int nmax = 0;
    for (int il = 1; il <= L; ++il) {
        if ((st = level_grids(c, t, il, lists[il]))) return st;
        nmax = std::max(nmax, (int)lists[il].size());
    }
    const size_t wmax = rh::pack_hydro_words(nmax, nb);
    int* d_grids = nullptr;
    uint32_t* d_out = nullptr;
    uint32_t* h_out = nullptr;
    struct Free {
        int*& a;
        uint32_t*& b;
        uint32_t*& h;
        ~Free() {
            if (a) (void)hipFree(a);
            if (b) (void)hipFree(b);
            if (h) (void)hipHostFree(h);
        }
    } guard{d_grids, d_out, h_out};
    RH_HIP(c, hipMalloc((void**)&d_grids, (size_t)std::max(nmax, 1) * sizeof(int)));
    RH_HIP(c, hipMalloc((void**)&d_out, wmax * sizeof(uint32_t)));
    RH_HIP(c, hipHostMalloc((void**)&h_out, wmax * sizeof(uint32_t), hipHostMallocDefault));

    Records r(path);
    if (!r.ok) return fail(c, RH_ERR_ARG, "snapshot: cannot open " + path);
    r.i32(t->ncpu);
    r.i32(c->mesh.nvar);
    r.i32(3);
    r.i32(L);
    r.i32(nb);
    r.f64(t->gamma);
    const rh::MeshDev m = mesh_dev(c);
    for (int il = 1; il <= L; ++il) {
        const std::vector<int>& g = lists[il];
        const int n = (int)g.size();
        if (n > 0)
            RH_HIP(c, hipMemcpyAsync(d_grids, g.data(), (size_t)n * sizeof(int), hipMemcpyHostToDevice, c->stream));
        RH_HIP(c, rh::launch_pack_hydro(m, d_grids, n, il, nb, t->gamma, d_out, c->stream));
        const size_t w = rh::pack_hydro_words(n, nb);
        RH_HIP(c, hipMemcpyAsync(h_out, d_out, w * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        RH_HIP(c, hipStreamSynchronize(c->stream));
        r.raw(h_out, w * sizeof(uint32_t));
    }
    if (!r.finish()) return fail(c, RH_ERR_ARG, "snapshot: write failed on " + path);
    return RH_OK;
}

// RamsesWriter::write_grav (RamsesWriter.cpp:443-473): phi and f (zero when
// the tree has none).  Unlike the hydro file, a domain without grids gets no
// (il, ncache) records.
int write_grav_file(rh_ctx* c, const rh_snapshot_tree* t, const std::string& path) {
    int st = check_tree(c, t);
    if (st) return st;
    const int L = t->nlevelmax;
    const size_t N = (size_t)t->ncell, G = (size_t)t->ngridmax;
    auto phi = [&](size_t i) { return t->phi ? t->phi[i] : 0.0; };
    auto f = [&](size_t i, int d) { return t->f ? t->f[(size_t)(d - 1) * N + i] : 0.0; };
    Records r(path);
    if (!r.ok) return fail(c, RH_ERR_ARG, "snapshot: cannot open " + path);
    r.i32(t->ncpu);
    r.i32(4);
    int ngrid_tot = 0;
    for (int l = 1; l <= L; ++l) ngrid_tot += numbl(t, 1, l);
    r.i32(ngrid_tot);
    r.i32(L);
    r.i32(t->nboundary);
    std::vector<double> tmp((size_t)t->ncoarse);
    for (int i = 0; i < t->ncoarse; ++i) tmp[i] = phi(i);
    r.vec(tmp);
    for (int d = 1; d <= 3; ++d) {
        for (int i = 0; i < t->ncoarse; ++i) tmp[i] = f(i, d);
        r.vec(tmp);
    }
    std::vector<int> g;
    for (int il = 1; il <= L; ++il) {
        if ((st = level_grids(c, t, il, g))) return st;
        const size_t n = g.size();
        if (n == 0) continue;
        r.i32(il);
        r.i32((int)n);
        tmp.resize(n);
        for (int k = 0; k < 8; ++k) {
            const size_t base = (size_t)t->ncoarse + k * G;
            for (size_t i = 0; i < n; ++i) tmp[i] = phi(base + g[i] - 1);
            r.vec(tmp);
            for (int d = 1; d <= 3; ++d) {
                for (size_t i = 0; i < n; ++i) tmp[i] = f(base + g[i] - 1, d);
                r.vec(tmp);
            }
        }
    }
    if (!r.finish()) return fail(c, RH_ERR_ARG, "snapshot: write failed on " + path);
    return RH_OK;
}

namespace {

int write_text(rh_ctx* c, const std::string& path, const std::string& text) {
    FILE* f = std::fopen(path.c_str(), "w");
    if (!f) return fail(c, RH_ERR_ARG, "snapshot: cannot open " + path);
    const bool ok = std::fwrite(text.data(), 1, text.size(), f) == text.size();
    if (std::fclose(f) != 0 || !ok) return fail(c, RH_ERR_ARG, "snapshot: write failed on " + path);
    return RH_OK;
}

std::string line_i(const char* key, long long v) { return std::string(key) + std::to_string(v) + "\n"; }

// The info file prints its reals after a sticky std::scientific /
// setprecision(16): every real is %.16e.
std::string line_e(const char* key, double v) {
    char b[64];
    std::snprintf(b, sizeof b, "%.16e", v);
    return std::string(key) + b + "\n";
}

}  // namespace

// RamsesWriter::write_header (RamsesWriter.cpp:237-266).
std::string info_text(const rh_snapshot_tree* t, const rh_snapshot_info* in) {
    std::string s;
    s += line_i("ncpu        = ", t->ncpu);
    s += line_i("ndim        = ", 3);
    s += line_i("nx          = ", t->nx);
    s += line_i("ny          = ", t->ny);
    s += line_i("nz          = ", t->nz);
    s += line_i("ncoarse     = ", t->ncoarse);
    s += line_i("nboundary   = ", t->nboundary);
    s += line_i("levelmin    = ", in->levelmin);
    s += line_i("levelmax    = ", in->levelmax);
    s += line_i("ngridmax    = ", t->ngridmax);
    s += line_i("nstep       = ", in->nstep);
    s += line_i("nstep_coarse= ", in->nstep_coarse);
    s += line_i("noutput     = ", in->noutput);
    s += line_e("time        = ", in->t);
    s += line_e("boxlen      = ", in->boxlen / (double)t->nx);
    s += line_e("omega_m     = ", in->omega_m);
    s += line_e("omega_l     = ", in->omega_l);
    s += line_e("omega_k     = ", in->omega_k);
    s += line_e("omega_b     = ", in->omega_b);
    s += line_e("h0          = ", in->h0);
    s += line_e("aexp        = ", in->aexp);
    s += line_e("unit_l      = ", in->units_length);
    s += line_e("unit_d      = ", in->units_density);
    s += line_e("unit_t      = ", in->units_time);
    s += line_i("nener       = ", in->nener);
    s += line_i("nvar        = ", 5);
    s += "ordering type= hilbert\n";
    return s;
}

// Simulation::dump_snapshot (Simulation.cpp:678-732) with rank 0 of 1.
int dump_snapshot(rh_ctx* c, const rh_snapshot_tree* t, const rh_snapshot_info* in, const char* root) {
    if (!in || !root) return fail(c, RH_ERR_ARG, "snapshot: info and root must not be NULL");
    int st = check_tree(c, t);
    if (st) return st;
    if (in->nener != 0) return fail(c, RH_ERR_UNSUPPORTED, "snapshot: nener must be 0");
    char b[32];
    std::snprintf(b, sizeof b, "/output_%05d", in->iout);
    const std::string dir = std::string(root) + b;
    if (::mkdir(dir.c_str(), 0777) != 0 && errno != EEXIST)
        return fail(c, RH_ERR_ARG, "snapshot: cannot create " + dir + ": " + std::strerror(errno));
    if ((st = write_amr_file(c, t, in, out_name(dir, "amr", in->iout, ".out", true)))) return st;
    if ((st = write_hydro_file(c, t, in, out_name(dir, "hydro", in->iout, ".out", true)))) return st;
    if ((st = write_grav_file(c, t, out_name(dir, "grav", in->iout, ".out", true)))) return st;
    // RamsesWriter::write_header_file (RamsesWriter.cpp:268-273).
    if ((st = write_text(c, out_name(dir, "header", in->iout, ".txt", false),
                         "# Family Count\nundefined " + std::to_string(t->npart) + "\nParticle fields\n")))
        return st;
    // RamsesWriter::write_hydro_descriptor (RamsesWriter.cpp:275-313), NDIM 3,
    // no MHD, nener 0, no passive scalars.
    if ((st = write_text(c, dir + "/hydro_file_descriptor.txt",
                         "# version: 1\n1, density, double\n2, velocity_x, double\n3, velocity_y, double\n"
                         "4, velocity_z, double\n5, pressure, double\n")))
        return st;
    // RamsesWriter::write_particles_descriptor (RamsesWriter.cpp:428-441).
    if ((st = write_text(c, dir + "/part_file_descriptor.txt",
                         "# version: 1\n1, position_x, double\n2, position_y, double\n3, position_z, double\n"
                         "4, velocity_x, double\n5, velocity_y, double\n6, velocity_z, double\n7, mass, double\n"
                         "8, identity, int\n9, level, int\n10, family, int\n11, tag, int\n12, birth_time, double\n"
                         "13, metallicity, double\n")))
        return st;
    std::snprintf(b, sizeof b, "/info_%05d.txt", in->iout);
    return write_text(c, dir + b, info_text(t, in));
}

}  // namespace rhi

extern "C" {

int rh_dump_snapshot(rh_ctx* ctx, const rh_snapshot_tree* tree, const rh_snapshot_info* info, const char* root) {
    return rhi::guard(ctx ? &ctx->err : &rhi::g_create_error, nullptr, [&]() -> int {
        if (!ctx) return RH_ERR_ARG;
        return rhi::dump_snapshot(ctx, tree, info, root);
    });
}

int rh_write_hydro(rh_ctx* ctx, const rh_snapshot_tree* tree, const rh_snapshot_info* info, const char* path) {
    return rhi::guard(ctx ? &ctx->err : &rhi::g_create_error, nullptr, [&]() -> int {
        if (!ctx || !path) return RH_ERR_ARG;
        return rhi::write_hydro_file(ctx, tree, info, path);
    });
}

int rh_write_amr(const rh_snapshot_tree* tree, const rh_snapshot_info* info, const char* path) {
    return rhi::guard(&rhi::g_create_error, nullptr, [&]() -> int {
        if (!info || !path) return rhi::fail(nullptr, RH_ERR_ARG, "snapshot: info and path must not be NULL");
        return rhi::write_amr_file(nullptr, tree, info, path);
    });
}

}  // extern "C"
