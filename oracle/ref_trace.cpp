// Test infrastructure only: drives the reference MCGPU CPU path (compiled
// from the reference sources by oracle/Makefile into oracle/_ref/) and
// writes a per-move trace.  Nothing in the product links or runs this.
//
// It performs, with the reference's own functions and classes, the start-up
// and step sequence of Simulation::run (Simulation.cpp:65-264) with the
// BruteForce strategy in serial mode:
//   createBox + seed            Simulation.cpp:40-56
//   keepMoleculeInBox per mol   Simulation.cpp:70-72
//   calcEnergy_LRC              Simulation.cpp:85
//   SimBoxBuilder::build        Simulation.cpp:125-127
//   calcSystemEnergy + LRC      Simulation.cpp:186-188
//   step loop                   Simulation.cpp:219-264
// and, per move, prints one line: step mol oldE newE dE accepted.
//
// usage: ref_trace <input .state|.config> <steps> <trace-out|-> [final-coords-out]
// With "-" no trace is written; the step loop's wall time is printed
// ("Loop Seconds") so the reference can serve as a timed CPU baseline.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "Metropolis/Box.h"
#include "Metropolis/BruteForceStep.h"
#include "Metropolis/Ensemble.h"
#include "Metropolis/GPUCopy.h"
#include "Metropolis/SerialSim/SerialCalcs.h"
#include "Metropolis/SimBox.h"
#include "Metropolis/SimBoxBuilder.h"
#include "Metropolis/SimulationArgs.h"
#include "Metropolis/SimulationStep.h"
#include "Metropolis/Utilities/FileUtilities.h"
#include "Metropolis/Utilities/MathLibrary.h"

static bool ends_with(const std::string& s, const std::string& suf) {
  return s.size() >= suf.size() &&
         s.compare(s.size() - suf.size(), suf.size(), suf) == 0;
}

int main(int argc, char** argv) {
  if (argc < 4) {
    std::fprintf(stderr,
                 "usage: %s <input.state|.config> <steps> <trace-out> "
                 "[final-coords-out]\n",
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
  args.stepCount = std::atoi(argv[2]);
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
    std::fprintf(stderr, "ref_trace: could not build box from %s\n", argv[1]);
    return 1;
  }
  seed(box->environment->randomseed);
  if (args.stepCount > 0) simSteps = args.stepCount;

  for (int m = 0; m < box->environment->numOfMolecules; m++)
    box->keepMoleculeInBox(m);
  Real lrc = SerialCalcs::calcEnergy_LRC(box);

  SimBoxBuilder builder(false, sbScanner);
  SimBox* sb = builder.build(box);
  GPUCopy::setParallel(false);
  SimulationStep* step = new BruteForceStep(sb);
  NVTEnsemble ensemble;
  GPUCopy::copyIn(sb);

  Real subLJ = 0, subCharge = 0;
  Real energy = step->calcSystemEnergy(subLJ, subCharge, sb->numMolecules);
  Real e0_pairs = energy;
  energy += lrc;

  const bool tracing = std::string(argv[3]) != "-";
  FILE* out = tracing ? std::fopen(argv[3], "w") : nullptr;
  if (tracing && !out) return 1;
  FILE* hdr = out ? out : stdout;
  std::fprintf(hdr, "# seed %d molecules %d sites %d\n",
               box->environment->randomseed, sb->numMolecules, sb->numAtoms);
  std::fprintf(hdr, "# E0_pairs %.17g LRC %.17g E0 %.17g\n", e0_pairs, lrc,
               energy);
  const auto t0 = std::chrono::steady_clock::now();
  long accepted = 0, rejected = 0;
  for (long move = stepStart; move < stepStart + simSteps; move++) {
    ensemble.update();
    int idx = step->chooseMolecule(sb);
    Real oldE = step->calcMoleculeEnergy(idx, 0);
    step->changeMolecule(idx, sb);
    Real newE = step->calcMoleculeEnergy(idx, 0);
    int acc = ensemble.accept(newE - oldE) ? 1 : 0;
    if (acc) {
      accepted++;
      energy += newE - oldE;
    } else {
      rejected++;
      step->rollback(idx, sb);
    }
    sb->stepNum++;
    if (out)
      std::fprintf(out, "%ld %d %.17g %.17g %.17g %d\n", move, idx, oldE, newE,
                   newE - oldE, acc);
  }
  const double loop_s =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  std::fprintf(hdr, "# final %.17g accepted %ld rejected %ld\n", energy,
               accepted, rejected);
  if (out) std::fclose(out);
  std::printf("Loop Seconds: %.6f\n", loop_s);

  if (argc > 4) {
    FILE* fc = std::fopen(argv[4], "w");
    if (!fc) return 1;
    for (int a = 0; a < sb->numAtoms; a++)
      std::fprintf(fc, "%.17g %.17g %.17g\n", sb->atomCoordinates[0][a],
                   sb->atomCoordinates[1][a], sb->atomCoordinates[2][a]);
    std::fclose(fc);
  }
  std::printf("Final Energy: %.3f\nAccepted Moves: %ld\nRejected Moves: %ld\n",
              energy, accepted, rejected);
  return 0;
}
