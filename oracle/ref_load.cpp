// Test infrastructure only: loads a .config (or .state) input through the
// reference's own loader, compiled from the reference sources by
// oracle/Makefile into oracle/_ref/, and dumps the box the energy path sees.
// Nothing in the product links or runs this.
//
// Sequence (the start-up half of Simulation::run):
//   createBox -> loadBoxData       Simulation.cpp:40-56, FileUtilities.cpp:19-110
//     ConfigScanner / OplsScanner / ZmatrixScanner, buildMoleculeXYZ,
//     buildBoxData, generatefccBox
//   keepMoleculeInBox per molecule Simulation.cpp:70-72
//   calcEnergy_LRC                 Simulation.cpp:85
//   SimBoxBuilder::build           Simulation.cpp:125-127
//   calcSystemEnergy               Simulation.cpp:186-188
//
// usage: ref_load <input.config|.state> <dump-out>
// Dump layout (text, doubles as %.17g):
//   env <nmol> <nsite> <npidx> <Lx> <Ly> <Lz> <cutoff> <T> <kT> <maxT>
//       <maxR> <seed> <steps>
//   energy <E0_pairs> <lrc>
//   mol <start> <len> <type> <pidx_start> <pidx_count>      (nmol lines)
//   pidx <absolute site index>                              (npidx lines)
//   site <x> <y> <z> <sigma> <epsilon> <charge> <name>      (nsite lines)
#include <cstdio>
#include <cstdlib>
#include <string>

#include "Metropolis/Box.h"
#include "Metropolis/BruteForceStep.h"
#include "Metropolis/GPUCopy.h"
#include "Metropolis/SerialSim/SerialCalcs.h"
#include "Metropolis/SimBox.h"
#include "Metropolis/SimBoxBuilder.h"
#include "Metropolis/SimulationArgs.h"
#include "Metropolis/Utilities/FileUtilities.h"

static bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() &&
         s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

int main(int argc, char** argv) {
  if (argc < 3) {
    std::fprintf(stderr, "usage: %s <input.config|.state> <dump-out>\n",
                 argv[0]);
    return 2;
  }
  SimulationArgs args;
  args.filePath = argv[1];
  args.fileType = ends_with(args.filePath, ".state") ? InputFile::State
                                                     : InputFile::Configuration;
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
    std::fprintf(stderr, "ref_load: could not build box from %s\n", argv[1]);
    return 1;
  }
  for (int m = 0; m < box->environment->numOfMolecules; m++)
    box->keepMoleculeInBox(m);
  const Real lrc = SerialCalcs::calcEnergy_LRC(box);

  SimBoxBuilder builder(false, sbScanner);
  SimBox* sb = builder.build(box);
  GPUCopy::setParallel(false);
  SimulationStep* step = new BruteForceStep(sb);
  GPUCopy::copyIn(sb);
  Real subLJ = 0, subCharge = 0;
  const Real e0 = step->calcSystemEnergy(subLJ, subCharge, sb->numMolecules);

  FILE* f = std::fopen(argv[2], "w");
  if (!f) return 1;
  std::fprintf(f,
               "env %d %d %d %.17g %.17g %.17g %.17g %.17g %.17g %.17g %.17g "
               "%d %ld\n",
               sb->numMolecules, sb->numAtoms, sb->numPIdxes, sb->size[0],
               sb->size[1], sb->size[2], sb->cutoff, sb->temperature, sb->kT,
               sb->maxTranslate, sb->maxRotate, box->environment->randomseed,
               simSteps);
  std::fprintf(f, "energy %.17g %.17g\n", e0, lrc);
  for (int m = 0; m < sb->numMolecules; m++)
    std::fprintf(f, "mol %d %d %d %d %d\n", sb->moleculeData[MOL_START][m],
                 sb->moleculeData[MOL_LEN][m], sb->moleculeData[MOL_TYPE][m],
                 sb->moleculeData[MOL_PIDX_START][m],
                 sb->moleculeData[MOL_PIDX_COUNT][m]);
  for (int p = 0; p < sb->numPIdxes; p++)
    std::fprintf(f, "pidx %d\n", sb->primaryIndexes[p]);
  // Site names come from the Box (SimBox keeps none), in the same order.
  int a = 0;
  for (int m = 0; m < sb->numMolecules; m++) {
    for (int k = 0; k < box->molecules[m].numOfAtoms; k++, a++) {
      const std::string* nm = box->molecules[m].atoms[k].name;
      std::fprintf(f, "site %.17g %.17g %.17g %.17g %.17g %.17g %s\n",
                   sb->atomCoordinates[0][a], sb->atomCoordinates[1][a],
                   sb->atomCoordinates[2][a], sb->atomData[ATOM_SIGMA][a],
                   sb->atomData[ATOM_EPSILON][a], sb->atomData[ATOM_CHARGE][a],
                   (nm && !nm->empty()) ? nm->c_str() : "?");
    }
  }
  std::fclose(f);
  return 0;
}
