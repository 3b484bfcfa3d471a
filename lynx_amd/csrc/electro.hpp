// Local-pseudopotential set-up of the SCF (electro.cpp).  Internal.
#pragma once

#include "common.hpp"

#include <vector>

namespace lynx {

struct ElecHost {
    std::vector<double> b, b_ref;    // pseudocharge and reference charge (Nd)
    std::vector<double> vc;          // correction potential V_ref - V_J (Nd)
    std::vector<double> rho_at;      // isolated-atom valence densities, scaled to Nelectron (Nd)
    std::vector<double> rho_core;    // model core density (Nd; zero without NLCC)
    double eself = 0, ec = 0, int_b = 0;
    int nelectron = 0, natom = 0, n_boxes = 0;
    bool has_nlcc = false;
};

void build_electrostatics(const Lattice &lat, const Stencil &st, const std::vector<Psp8> &psps,
                          const std::vector<int> &types, const std::vector<double> &pos_nc, ElecHost &out);

// Local and core-correction forces per atom, lattice-aligned components
// (3 per atom): dphi and dvc are the FD gradients of phi and Vc along the
// three lattice directions on the grid; vxc (Nd, or null) weights the
// gradient of each atom's core density.
void local_forces(const Lattice &lat, const Stencil &st, const std::vector<Psp8> &psps,
                  const std::vector<int> &types, const std::vector<double> &pos_nc, const ElecHost &E,
                  const double *const dphi[3], const double *const dvc[3], const double *vxc,
                  std::vector<double> &f_loc, std::vector<double> &f_xc);

}  // namespace lynx
