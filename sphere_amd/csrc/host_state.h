// host_state.h -- owning host image of one state (header scalars + arrays in
// the layout of include/sphere_hip.h).
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

#include "../../include/sphere_hip.h"

struct HostState {
    sph_state h{};  // scalars (array pointers unused)
    std::vector<double> x, xyzsum, vel, force, angpos, angvel, torque;
    std::vector<double> es_dot, es, ev_dot, ev, p;
    std::vector<int32_t> color;
    std::vector<uint32_t> bonds;
    std::vector<double> bdn, bdt, bon, bot;
    std::vector<double> v_f, p_f, phi, dphi, f_p;
    std::vector<int32_t> p_constant;

    void alloc_particles(size_t np)
    {
        x.assign(np * 4, 0.0); xyzsum.assign(np * 3, 0.0); vel.assign(np * 4, 0.0);
        force.assign(np * 3, 0.0); angpos.assign(np * 3, 0.0); angvel.assign(np * 3, 0.0);
        torque.assign(np * 3, 0.0); es_dot.assign(np, 0.0); es.assign(np, 0.0);
        ev_dot.assign(np, 0.0); ev.assign(np, 0.0); p.assign(np, 0.0); color.assign(np, 0);
    }
    void alloc_bonds(size_t nb)
    {
        bonds.assign(nb * 2, 0u); bdn.assign(nb, 0.0); bdt.assign(nb * 3, 0.0);
        bon.assign(nb, 0.0); bot.assign(nb * 3, 0.0);
    }
    void alloc_cells(size_t cells)
    {
        v_f.assign(cells * 3, 0.0); p_f.assign(cells, 0.0); phi.assign(cells, 0.0);
        dphi.assign(cells, 0.0); p_constant.assign(cells, 0);
    }
};

int read_state_file(const char *path, int fluid, HostState &hs, std::string &err);
int write_state_file(const char *path, const HostState &hs, std::string &err);
