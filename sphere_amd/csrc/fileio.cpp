// fileio.cpp -- the versioned binary format v2.15 in C++ (the field order of
// DEM::readbin / DEM::writebin, src/file_io.cpp:30-731).  The reader fills
// a HostState; the writer serialises one.
#include "host_state.h"

#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>

namespace {

struct In {
    const std::vector<char> &b;
    size_t o = 0;
    bool ok = true;
    explicit In(const std::vector<char> &bb) : b(bb) {}
    template <class T> T get()
    {
        T v{};
        if (o + sizeof(T) > b.size()) { ok = false; return v; }
        std::memcpy(&v, b.data() + o, sizeof(T));
        o += sizeof(T);
        return v;
    }
    template <class T> void arr(T *dst, size_t n)
    {
        if (o + n * sizeof(T) > b.size()) { ok = false; return; }
        if (n) std::memcpy(dst, b.data() + o, n * sizeof(T));
        o += n * sizeof(T);
    }
};

struct Out {
    std::vector<char> b;
    template <class T> void put(T v)
    {
        const char *p = reinterpret_cast<const char *>(&v);
        b.insert(b.end(), p, p + sizeof(T));
    }
    template <class T> void arr(const T *src, size_t n)
    {
        const char *p = reinterpret_cast<const char *>(src);
        b.insert(b.end(), p, p + n * sizeof(T));
    }
};

}  // namespace

int read_state_file(const char *path, int fluid, HostState &hs, std::string &err)
{
    std::ifstream ifs(path, std::ios_base::binary);
    if (!ifs) {
        err = std::string("Could not read input binary file '") + path + "'";
        return SPH_ERR_FILE;
    }
    std::vector<char> data((std::istreambuf_iterator<char>(ifs)), std::istreambuf_iterator<char>());
    In in(data);
    sph_state &h = hs.h;
    std::memset(&h, 0, sizeof(h));
    double version = in.get<double>();
    if (version != SPH_VERSION) {
        char m[256];
        snprintf(m, sizeof(m), "The input file '%s' is written by sphere version %g, which is "
                 "incompatible with this version (%g).", path, version, SPH_VERSION);
        err = m;
        return SPH_ERR_FILE;
    }
    h.nd = in.get<int32_t>();
    h.np = in.get<uint32_t>();
    if (h.nd != 3) { err = "Dimensionality mismatch between dataset and this SPHERE program"; return SPH_ERR_FILE; }
    h.time_dt = in.get<double>();
    h.time_current = in.get<double>();
    h.time_total = in.get<double>();
    h.time_file_dt = in.get<double>();
    h.time_step_count = in.get<uint32_t>();
    in.arr(h.origo, 3);
    in.arr(h.L, 3);
    in.arr(h.num, 3);
    h.periodic = in.get<int32_t>();
    h.adaptive = in.get<int32_t>();
    const size_t np = h.np;
    if (!in.ok || np * 228 > data.size()) { err = "input binary file is truncated"; return SPH_ERR_FILE; }
    hs.alloc_particles(np);
    in.arr(hs.x.data(), np * 4);
    in.arr(hs.xyzsum.data(), np * 3);
    in.arr(hs.vel.data(), np * 4);
    in.arr(hs.force.data(), np * 3);
    in.arr(hs.angpos.data(), np * 3);
    in.arr(hs.angvel.data(), np * 3);
    in.arr(hs.torque.data(), np * 3);
    in.arr(hs.es_dot.data(), np);
    in.arr(hs.es.data(), np);
    in.arr(hs.ev_dot.data(), np);
    in.arr(hs.ev.data(), np);
    in.arr(hs.p.data(), np);
    in.arr(h.g, 3);
    h.k_n = in.get<double>(); h.k_t = in.get<double>(); h.k_r = in.get<double>();
    h.E = in.get<double>(); h.gamma_n = in.get<double>(); h.gamma_t = in.get<double>();
    h.gamma_r = in.get<double>(); h.mu_s = in.get<double>(); h.mu_d = in.get<double>();
    h.mu_r = in.get<double>(); h.gamma_wn = in.get<double>(); h.gamma_wt = in.get<double>();
    h.mu_ws = in.get<double>(); h.mu_wd = in.get<double>(); h.rho = in.get<double>();
    h.contactmodel = in.get<uint32_t>();
    h.kappa = in.get<double>(); h.db = in.get<double>(); h.V_b = in.get<double>();
    h.nw = in.get<uint32_t>();
    if (h.nw > SPH_MAXWALLS) {
        err = "MAXWALLS is smaller than the number of walls specified in the input file";
        return SPH_ERR_FILE;
    }
    in.arr(h.wmode, h.nw);
    in.arr(h.w_nx, (size_t)h.nw * 4);
    in.arr(h.w_mvfd, (size_t)h.nw * 4);
    h.sigma0_A = in.get<double>(); h.sigma0_f = in.get<double>(); h.tau_x = in.get<double>();
    h.lambda_bar = in.get<double>();
    h.nb0 = in.get<uint32_t>();
    h.sigma_b = in.get<double>(); h.tau_b = in.get<double>();
    const size_t nb = h.nb0;
    if (!in.ok || nb * 72 > data.size()) { err = "input binary file is truncated"; return SPH_ERR_FILE; }
    hs.alloc_bonds(nb);
    in.arr(hs.bonds.data(), nb * 2);
    in.arr(hs.bdn.data(), nb);
    in.arr(hs.bdt.data(), nb * 3);
    in.arr(hs.bon.data(), nb);
    in.arr(hs.bot.data(), nb * 3);
    h.fluid = fluid ? 1 : 0;
    if (fluid) {
        h.cfd_solver = in.get<int32_t>();
        if (h.cfd_solver != 1) {
            err = "Value of cfd_solver not understood or not the Darcy solver (" +
                  std::to_string(h.cfd_solver) + ")";
            return SPH_ERR_FILE;
        }
        h.mu = in.get<double>();
        const size_t cells = (size_t)h.num[0] * h.num[1] * h.num[2];
        if (cells * 52 > data.size()) { err = "input binary file is truncated"; return SPH_ERR_FILE; }
        hs.alloc_cells(cells);
        for (size_t c = 0; c < cells; ++c) {
            hs.v_f[3 * c] = in.get<double>();
            hs.v_f[3 * c + 1] = in.get<double>();
            hs.v_f[3 * c + 2] = in.get<double>();
            hs.p_f[c] = in.get<double>();
            hs.phi[c] = in.get<double>();
            hs.dphi[c] = in.get<double>();
        }
        h.rho_f = in.get<double>(); h.p_mod_A = in.get<double>();
        h.p_mod_f = in.get<double>(); h.p_mod_phi = in.get<double>();
        h.bc_xn = in.get<int32_t>(); h.bc_xp = in.get<int32_t>();
        h.bc_yn = in.get<int32_t>(); h.bc_yp = in.get<int32_t>();
        h.bc_bot = in.get<int32_t>(); h.bc_top = in.get<int32_t>();
        h.free_slip_bot = in.get<int32_t>(); h.free_slip_top = in.get<int32_t>();
        h.bc_bot_flux = in.get<double>(); h.bc_top_flux = in.get<double>();
        in.arr(hs.p_constant.data(), cells);
        h.tolerance = in.get<double>();
        h.maxiter = in.get<uint32_t>();
        h.ndem = in.get<uint32_t>();
        h.c_phi = in.get<double>();
        hs.f_p.assign(np * 3, 0.0);
        in.arr(hs.f_p.data(), np * 3);
        h.beta_f = in.get<double>();
        h.k_c = in.get<double>();
    }
    in.arr(hs.color.data(), np);
    if (!in.ok) { err = "input binary file is truncated"; return SPH_ERR_FILE; }
    return SPH_OK;
}

int write_state_file(const char *path, const HostState &hs, std::string &err)
{
    const sph_state &h = hs.h;
    Out o;
    o.put<double>(SPH_VERSION);
    o.put<int32_t>(h.nd);
    o.put<uint32_t>(h.np);
    o.put(h.time_dt); o.put(h.time_current); o.put(h.time_total); o.put(h.time_file_dt);
    o.put<uint32_t>(h.time_step_count);
    o.arr(h.origo, 3);
    o.arr(h.L, 3);
    o.arr(h.num, 3);
    o.put<int32_t>(h.periodic);
    o.put<int32_t>(h.adaptive);
    const size_t np = h.np;
    o.arr(hs.x.data(), np * 4);
    o.arr(hs.xyzsum.data(), np * 3);
    o.arr(hs.vel.data(), np * 4);
    o.arr(hs.force.data(), np * 3);
    o.arr(hs.angpos.data(), np * 3);
    o.arr(hs.angvel.data(), np * 3);
    o.arr(hs.torque.data(), np * 3);
    o.arr(hs.es_dot.data(), np);
    o.arr(hs.es.data(), np);
    o.arr(hs.ev_dot.data(), np);
    o.arr(hs.ev.data(), np);
    o.arr(hs.p.data(), np);
    o.arr(h.g, 3);
    o.put(h.k_n); o.put(h.k_t); o.put(h.k_r); o.put(h.E); o.put(h.gamma_n); o.put(h.gamma_t);
    o.put(h.gamma_r); o.put(h.mu_s); o.put(h.mu_d); o.put(h.mu_r); o.put(h.gamma_wn);
    o.put(h.gamma_wt); o.put(h.mu_ws); o.put(h.mu_wd); o.put(h.rho);
    o.put<uint32_t>(h.contactmodel);
    o.put(h.kappa); o.put(h.db); o.put(h.V_b);
    o.put<uint32_t>(h.nw);
    o.arr(h.wmode, h.nw);
    o.arr(h.w_nx, (size_t)h.nw * 4);
    o.arr(h.w_mvfd, (size_t)h.nw * 4);
    o.put(h.sigma0_A); o.put(h.sigma0_f); o.put(h.tau_x);
    o.put(h.lambda_bar);
    o.put<uint32_t>(h.nb0);
    o.put(h.sigma_b); o.put(h.tau_b);
    const size_t nb = h.nb0;
    o.arr(hs.bonds.data(), nb * 2);
    o.arr(hs.bdn.data(), nb);
    o.arr(hs.bdt.data(), nb * 3);
    o.arr(hs.bon.data(), nb);
    o.arr(hs.bot.data(), nb * 3);
    if (h.fluid) {
        o.put<int32_t>(h.cfd_solver);
        o.put(h.mu);
        const size_t cells = (size_t)h.num[0] * h.num[1] * h.num[2];
        for (size_t c = 0; c < cells; ++c) {
            o.put(hs.v_f[3 * c]); o.put(hs.v_f[3 * c + 1]); o.put(hs.v_f[3 * c + 2]);
            o.put(hs.p_f[c]); o.put(hs.phi[c]); o.put(hs.dphi[c]);
        }
        o.put(h.rho_f); o.put(h.p_mod_A); o.put(h.p_mod_f); o.put(h.p_mod_phi);
        o.put<int32_t>(h.bc_xn); o.put<int32_t>(h.bc_xp); o.put<int32_t>(h.bc_yn);
        o.put<int32_t>(h.bc_yp); o.put<int32_t>(h.bc_bot); o.put<int32_t>(h.bc_top);
        o.put<int32_t>(h.free_slip_bot); o.put<int32_t>(h.free_slip_top);
        o.put(h.bc_bot_flux); o.put(h.bc_top_flux);
        o.arr(hs.p_constant.data(), cells);
        o.put(h.tolerance);
        o.put<uint32_t>(h.maxiter);
        o.put<uint32_t>(h.ndem);
        o.put(h.c_phi);
        o.arr(hs.f_p.data(), np * 3);
        o.put(h.beta_f);
        o.put(h.k_c);
    }
    o.arr(hs.color.data(), np);
    std::ofstream ofs(path, std::ios_base::binary);
    if (!ofs) {
        err = std::string("could create output binary file '") + path + "'";
        return SPH_ERR_FILE;
    }
    ofs.write(o.b.data(), (std::streamsize)o.b.size());
    if (!ofs) { err = std::string("write failed: ") + path; return SPH_ERR_FILE; }
    return SPH_OK;
}
