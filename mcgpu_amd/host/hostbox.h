// Host-resident box: the SimBox fields the energy path needs, as owned SoA
// vectors (SimBox.h:18-612 keeps the same rows behind Real** tables).
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "mcgpu_energy.h"

namespace mcgpu {

constexpr double kBoltz = 0.00198717;  // kcal/mol/K (SimBox.h:16)

struct HostBox {
  double L[3] = {0, 0, 0};
  double cutoff = 0, temperature = 0, kT = 0;
  double max_translate = 0, max_rotate = 0;
  uint32_t mc_seed = 1;
  std::string primary_line = "1";  // 1-based primary definition per type

  std::vector<double> x, y, z;
  std::vector<double> sigma, epsilon, charge;
  std::vector<std::string> site_name;
  std::vector<int32_t> mol_start, mol_len, mol_pidx_start, mol_pidx_count;
  std::vector<int32_t> mol_type, pidx;

  // where the box's move sequence stopped (a .state written by this
  // library after a run); absent for fresh boxes
  bool has_resume = false;
  mcgpu_resume_state resume{};

  int32_t n_mol() const { return (int32_t)mol_start.size(); }
  int32_t n_site() const { return (int32_t)x.size(); }

  mcgpu_box_desc desc() const {
    mcgpu_box_desc d{};
    d.n_molecules = n_mol();
    d.n_sites = n_site();
    d.n_primary = (int32_t)pidx.size();
    for (int k = 0; k < 3; k++) d.box[k] = L[k];
    d.cutoff = cutoff;
    d.kT = kT;
    d.max_translate = max_translate;
    d.max_rotate = max_rotate;
    d.x = x.data();
    d.y = y.data();
    d.z = z.data();
    d.sigma = sigma.data();
    d.epsilon = epsilon.data();
    d.charge = charge.data();
    d.mol_start = mol_start.data();
    d.mol_len = mol_len.data();
    d.mol_pidx_start = mol_pidx_start.data();
    d.mol_pidx_count = mol_pidx_count.data();
    d.primary_index = pidx.data();
    return d;
  }

  static HostBox from_desc(const mcgpu_box_desc& d);
};

// Validates a descriptor; returns an empty string when it is usable.
std::string validate_desc(const mcgpu_box_desc& d);

// Thread-local error message shared by every C entry point of the library.
void set_error(const std::string& msg);

}  // namespace mcgpu
