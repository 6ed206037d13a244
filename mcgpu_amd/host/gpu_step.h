// GPUStep: the reference's SimulationStep strategy interface
// (SimulationStep.h:24-94) served by the MI355X energy path.  A metrosim
// build selects it the way Simulation.cpp:140-157 selects BruteForceStep;
// its methods keep the reference's names, arguments and meaning.
#pragma once

#include <cstdint>
#include <vector>

#include "../csrc/rng31.h"
#include "hostbox.h"
#include "mcgpu_energy.h"

namespace mcgpu {

class GPUStep {
 public:
  // Creates the device box from `box` (which stays the host mirror, like
  // SimBox under GPUCopy) with the given pair arithmetic and neighbour mode.
  // skin <= 0 takes the reference's Verlet shell, 1.2875 rc
  // (VerletStep.cpp:133): skin = 0.2875 rc.
  GPUStep(HostBox* box, int precision, int neighbor, double skin = 0.0);
  ~GPUStep();
  int status() const { return status_; }

  // SimulationStep.cpp:25-27 (molecule index clamped to N-1; see .cpp)
  int chooseMolecule(Rng31& rng) const;
  // SimulationStep.cpp:15-18 (intramolecular terms disabled, .h:15)
  double calcMoleculeEnergy(int currMol, int startMol);
  // BruteForceStep.cpp:15-46 semantics on the device
  double calcMolecularEnergyContribution(int currMol, int startMol);
  // SimulationStep.cpp:520-566: move on the host mirror, then the device
  void changeMolecule(int molIdx, Rng31& rng);
  // SimulationStep.cpp:731-749
  void rollback(int molIdx);
  // SimulationStep.cpp:40-71 (subLJ/subCharge added as the reference does)
  double calcSystemEnergy(double& subLJ, double& subCharge, int numMolecules);

  mcgpu_box* device_box() { return dev_; }

 private:
  HostBox* hb_;
  mcgpu_box* dev_ = nullptr;
  int status_ = 0;
  std::vector<double> rollback_;  // 3 x len, x block / y block / z block
  std::vector<double> scratch_;
};

// Rigid trial move of the reference (SimulationStep.cpp:520-566, rotations
// :387-422, keepMoleculeInBox :710-729) applied to molecule `mol` of `hb`,
// written into out (x block, y block, z block); hb is not modified.
void propose_move(const HostBox& hb, int mol, int vertex, const double d[3],
                  const double rot_deg[3], double* out);

// Box.cpp:84-109: wrap a molecule by its first primary site.
void keep_in_box(HostBox& hb, int mol);

}  // namespace mcgpu
