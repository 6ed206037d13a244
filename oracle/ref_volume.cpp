// Test infrastructure only: calls the reference's own NPT acceptance test,
// SimCalcs::acceptVolumeMove (SimulationStep.cpp:476-488), compiled from the
// reference sources by oracle/Makefile into oracle/_ref/.  Nothing in the
// product links or runs this.
//
// A SimBox is built from a .state file the way Simulation::run does it
// (createBox, SimBoxBuilder::build, GPUCopy::copyIn in serial mode,
// Simulation.cpp:40-56,125-127); then, per input line on stdin
//   seed n_molecules kT v_old Lx Ly Lz dE pressure
// the box edges the function reads (GPUCopy::sizePtr) are set to Lx Ly Lz,
// sb->numMolecules and sb->kT to the line's values, the generator is seeded
// (MathLibrary.cpp:27-29), and one line is printed: the function's decision
// (1 accepted, 0 rejected).
//
// usage: ref_volume <input.state>  < grid  > decisions
#include <cstdio>
#include <cstdlib>

#include "Metropolis/Box.h"
#include "Metropolis/BruteForceStep.h"
#include "Metropolis/GPUCopy.h"
#include "Metropolis/SerialSim/SerialCalcs.h"
#include "Metropolis/SimBox.h"
#include "Metropolis/SimBoxBuilder.h"
#include "Metropolis/SimulationArgs.h"
#include "Metropolis/SimulationStep.h"
#include "Metropolis/Utilities/FileUtilities.h"
#include "Metropolis/Utilities/MathLibrary.h"

int main(int argc, char** argv) {
  if (argc < 2) {
    std::fprintf(stderr, "usage: %s <input.state> < grid\n", argv[0]);
    return 2;
  }
  SimulationArgs args;
  args.filePath = argv[1];
  args.fileType = InputFile::State;
  args.simulationMode = SimulationMode::Serial;
  args.strategy = Strategy::BruteForce;
  args.deviceIndex = 0;
  args.stepCount = 0;
  args.verboseOutput = false;
  args.useNeighborList = false;
  args.neighborListInterval = 0;
  args.ewaldRadius = 0;
  args.volumeInterval = 0;
  args.statusInterval = 0;
  args.stateInterval = -1;

  SBScanner* sbScanner = new SBScanner();
  long stepStart = 0, simSteps = 0;
  Box* box = SerialCalcs::createBox(args, &stepStart, &simSteps, sbScanner);
  if (box == NULL) {
    std::fprintf(stderr, "ref_volume: could not build box from %s\n", argv[1]);
    return 1;
  }
  SimBoxBuilder builder(false, sbScanner);
  SimBox* sb = builder.build(box);
  GPUCopy::setParallel(false);
  SimulationStep* step = new BruteForceStep(sb);  // sets SimCalcs::sb
  GPUCopy::copyIn(sb);
  (void)step;

  long s;
  int n;
  double kT, v_old, L[3], de, p;
  while (std::scanf("%ld %d %lf %lf %lf %lf %lf %lf %lf", &s, &n, &kT, &v_old, &L[0],
                    &L[1], &L[2], &de, &p) == 9) {
    Real* size = GPUCopy::sizePtr();
    for (int d = 0; d < 3; d++) size[d] = L[d];
    sb->numMolecules = n;
    sb->kT = kT;
    seed((int)s);
    std::printf("%d\n", SimCalcs::acceptVolumeMove(de, v_old, p) ? 1 : 0);
  }
  return 0;
}
