// Host-side view of a device-resident box (shared by the single-box and
// batched translation units).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "device_common.h"
#include "mcgpu_energy.h"

namespace mcgpu {

void set_error(const std::string& msg);
std::string validate_desc(const mcgpu_box_desc& d);

#define MCGPU_HIP_TRY(call)                                                  \
  do {                                                                       \
    hipError_t e_ = (call);                                                  \
    if (e_ != hipSuccess) {                                                  \
      ::mcgpu::set_error(std::string(#call) + " failed: " +                  \
                         hipGetErrorString(e_));                             \
      return MCGPU_ERR_HIP;                                                  \
    }                                                                        \
  } while (0)

// Result of a fused trial move / molecule energy launch.
struct MoveOut {
  double e_old;
  double e_new;
  double de;
  int32_t accepted;
  int32_t source;  // neighbour source used (MCGPU_NB_*)
  int32_t stale;   // Verlet mode: a committed primary left its skin
  int32_t reserved;
};

// Molecule signatures -> site parameter table, plus per-molecule base index.
struct ParamTable {
  std::vector<SiteParam64> p64;
  std::vector<SiteParam32> p32;
  std::vector<int32_t> mol_param;
  std::vector<MolMeta> meta;
  int32_t single_primary = 1;
  double spread = 0.0;  // max distance of a primary from the first primary
  double reach = 0.0;   // max distance of a site from the first primary
  int32_t max_len = 0;
  // every molecule alike (see BoxView::uni_len); zeros otherwise
  bool uniform = true;
  int32_t uni_len = 0, uni_param = 0, uni_p0 = 0;
};
ParamTable build_param_table(const mcgpu_box_desc& d);

// BoxView::reach2 for a largest site distance `reach` from the first primary.
inline double reach2_of(double reach) { return 2.0 * reach + 1e-6; }

// Largest distance of a site of one molecule from its first primary (site
// pf of the molecule); xyz = x block, y block, z block of len sites each.
double molecule_reach(const double* x, const double* y, const double* z, int len,
                      int pf);

// Cell grid geometry for a box: cells of width >= rc + 2 * spread.
void cell_geometry(const double L[3], double cutoff, double spread, int n[3],
                   double inv_w[3]);

int ensure_device();

// Launch helpers shared by the single-box and batched paths (stream-ordered,
// no synchronisation).
int launch_refresh_primary(const BoxView& bv, int n_pidx, hipStream_t s);
int launch_cells_build(const BoxView& bv, hipStream_t s);
int launch_volume_scale(const BoxView& bv, double f, hipStream_t s);

}  // namespace mcgpu
