// metrosim: the reference's command-line simulator over the MI355X energy
// path.  Same invocation, options, console summary, state files, PDB
// snapshot and .results file as the reference's front end, so scripts and
// test harnesses written for it (TestUtil.cpp:76-144) run unchanged:
//   options                 CommandParsing.cpp:36-255
//   strategy names          SimulationArgs.cpp:4-16
//   input routes            FileUtilities.cpp:19-110 (.config, .state)
//   run + outputs           Simulation.cpp:65-339, :341-417
// The whole step loop runs on the GPU (the batched persistent kernel with
// the box's glibc stream on the device), so -s and -p select the same
// engine; the move sequence is the reference's for the seed.
//
//   metrosim <file.config|file.state> [-s|-p] [-n steps] [-i status]
//            [-I state] [--name title] [-S strategy] [-l interval] [-k]
//            [--mixed] [-h] [-V]
#include <getopt.h>

#include <algorithm>
#include <cerrno>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <string>
#include <vector>

#include "mcgpu_host.h"

namespace {

const char* kApp = "metrosim";
constexpr int kLongName = 400;
constexpr int kLongMixed = 401;
constexpr int kLongVolume = 402;
constexpr long kDefaultSteps = 100;        // .state runs (FileUtilities.cpp:17)
constexpr double kAtmToKcalMolA3 = 1.4584e-5;

struct Args {
  std::string input;
  bool serial = false, parallel = false, verbose = false, mixed = false;
  long steps = 0;
  int status_interval = 1000;  // CommandParsing.h:21
  int state_interval = 0;      // 0: final state only; < 0: none
  std::string name, strategy;
  bool neighbor = false;
  double max_volume_delta = 0.01;
};

bool to_int(const char* s, long& v) {
  char* end = nullptr;
  errno = 0;
  const long r = std::strtol(s, &end, 10);
  if (errno || end == s || *end) return false;
  v = r;
  return true;
}

void help() {
  std::printf(
      "\nUsage: %s <inputfile> [options]\n\n"
      "Input File Types\n=================\n"
      "\t.config\t: a fresh run from the configuration's z-matrix and OPLS\n"
      "\t\t  parameters (FCC start)\n"
      "\t.state\t: resume a saved run\n\n"
      "Operation Flags\n====================\n"
      "--parallel\t(-p)\tRun on the GPU (the only engine of this build).\n"
      "--serial\t(-s)\tAccepted for compatibility: also runs on the GPU.\n"
      "--verbose\t(-k)\tPrint status updates.\n"
      "--mixed\t\t\tFP32 LJ terms and 1/r estimates; FP64 sums and Metropolis test.\n\n"
      "Simulation Run Parameters\n==========================\n"
      "--name <title>\t\tBase name of the .results, .pdb and .state files.\n"
      "--steps <count>\t(-n)\tTrial moves to run (> 0).\n"
      "--neighbor <interval>\t(-l)\tUse cell lists (exact: the same molecule\n"
      "\t\t\tpairs as brute force).\n"
      "--status-interval <n>\t(-i)\tSteps between status lines (with -k).\n"
      "--state-interval <n>\t(-I)\tSteps between .state snapshots; 0 = final\n"
      "\t\t\tstate only, < 0 = none.  Files: <name>_<step>.state\n"
      "--strategy <name>\t(-S)\tbrute-force, proximity-matrix,\n"
      "\t\t\tneighbor-linked-cell or verlet-list (all exact).\n"
      "--max-volume-delta <f>\tNPT: relative volume-move size (default 0.01).\n\n"
      "NPT runs take \"ensemble=npt\", \"pressure\" (atm) and\n"
      "\"volume-interval\" from the .config file.\n\n",
      kApp);
}

// Returns 0 to run, 1 for a clean exit (help/version), -1 on a usage error.
int parse_args(int argc, char** argv, Args& a) {
  opterr = 0;
  const char* short_opts = ":i:I:n:sphVkl:S:";
  static struct option long_opts[] = {
      {"status-interval", required_argument, 0, 'i'},
      {"state-interval", required_argument, 0, 'I'},
      {"steps", required_argument, 0, 'n'},
      {"serial", no_argument, 0, 's'},
      {"parallel", no_argument, 0, 'p'},
      {"help", no_argument, 0, 'h'},
      {"version", no_argument, 0, 'V'},
      {"verbose", no_argument, 0, 'k'},
      {"neighbor", required_argument, 0, 'l'},
      {"name", required_argument, 0, kLongName},
      {"strategy", required_argument, 0, 'S'},
      {"mixed", no_argument, 0, kLongMixed},
      {"max-volume-delta", required_argument, 0, kLongVolume},
      {0, 0, 0, 0}};
  int c, idx;
  long v;
  while ((c = getopt_long(argc, argv, short_opts, long_opts, &idx)) != -1) {
    switch (c) {
      case 'V':
        std::printf("%s: MCGPU Monte Carlo Simulator (MI355X energy path)\n", kApp);
        return 1;
      case 'h':
        help();
        return 1;
      case kLongName:
        a.name = optarg;
        break;
      case kLongMixed:
        a.mixed = true;
        break;
      case kLongVolume:
        a.max_volume_delta = std::atof(optarg);
        if (!(a.max_volume_delta > 0 && a.max_volume_delta < 0.5)) {
          std::fprintf(stderr, "%s: --max-volume-delta must lie in (0, 0.5)\n", kApp);
          return -1;
        }
        break;
      case 'i':
        if (!to_int(optarg, v)) {
          std::fprintf(stderr, "%s:  --status_interval (-i): Invalid status interval\n", kApp);
          return -1;
        }
        if (v < 0) {
          std::fprintf(stderr, "%s:  --status_interval (-i): Status Interval must be non-negative\n", kApp);
          return -1;
        }
        a.status_interval = (int)v;
        break;
      case 'I':
        if (!to_int(optarg, v)) {
          std::fprintf(stderr, "%s:  --state_interval (-I): Invalid state interval\n", kApp);
          return -1;
        }
        a.state_interval = (int)v;
        break;
      case 'n':
        if (!to_int(optarg, v)) {
          std::fprintf(stderr, "%s:  --steps (-n): Invalid step count\n", kApp);
          return -1;
        }
        if (v <= 0) {
          std::fprintf(stderr, "%s:  --steps (-n): Step count must be greater than zero\n", kApp);
          return -1;
        }
        a.steps = v;
        break;
      case 's':
        a.serial = true;
        break;
      case 'p':
        a.parallel = true;
        break;
      case 'k':
        a.verbose = true;
        break;
      case 'l':
        a.neighbor = true;
        if (to_int(optarg, v) && v < 1) {
          std::fprintf(stderr, "%s:  --neighborlist_interval (-l): Neighborlist Interval must be greater than 0\n", kApp);
          return -1;
        }
        break;
      case 'S':
        a.strategy = optarg;
        break;
      case ':':
        if (optopt == 'l') {
          a.neighbor = true;
          break;
        }
        std::fprintf(stderr, "%s: A required argument is missing for %s\n", kApp,
                     argv[optind - 1]);
        return -1;
      case '?':
      default:
        if (optopt)
          std::fprintf(stderr, "%s: Unknown option -%c\n", kApp, optopt);
        else
          std::fprintf(stderr, "%s: Unknown option %s\n", kApp, argv[optind - 1]);
        return -1;
    }
  }
  const int nargs = argc - optind;
  if (nargs < 1) {
    std::fprintf(stderr, "%s: Input file not specified\n", kApp);
    return -1;
  }
  if (nargs > 1) {
    std::fprintf(stderr, "%s: Too many input files specified\n", kApp);
    return -1;
  }
  if (a.serial && a.parallel) {
    std::fprintf(stderr, "%s: Cannot specify both GPU and CPU modes\n", kApp);
    return -1;
  }
  a.input = argv[optind];
  return 0;
}

bool ends_with(const std::string& s, const char* suf) {
  const size_t n = std::strlen(suf);
  return s.size() > n && s.compare(s.size() - n, n, suf) == 0;
}

// Strategy names of SimulationArgs.cpp:4-16 -> neighbour source.  Every
// strategy of this build sums exactly the brute-force molecule set.
bool strategy_neighbor(const std::string& s, int& neighbor) {
  if (s.empty() || s == "brute" || s == "brute-force" || s == "prox" ||
      s == "proximity-matrix") {
    neighbor = MCGPU_NEIGHBOR_BRUTE;
    return true;
  }
  if (s == "nlc" || s == "neighbor-linked-cell" || s == "verlet" ||
      s == "verlet-list") {
    neighbor = MCGPU_NEIGHBOR_CELL;
    return true;
  }
  return false;
}

std::string now_string() {
  const std::time_t t = std::time(nullptr);
  std::tm tm{};
  localtime_r(&t, &tm);
  char buf[80];
  std::strftime(buf, sizeof(buf), "%Y-%m-%d %X", &tm);
  return buf;
}

[[noreturn]] void die(const char* what) {
  std::fprintf(stderr, "Error: %s: %s\n", what, mcgpu_last_error());
  std::exit(EXIT_FAILURE);
}

}  // namespace

int main(int argc, char** argv) {
  Args a;
  const int pr = parse_args(argc, argv, a);
  if (pr != 0) return EXIT_FAILURE;  // help and version exit so too

  // ---- input (before any device work, as the reference loads first)
  mcgpu_hostbox* hb = nullptr;
  mcgpu_config_info info{};
  long steps = 0, start_step = 0;
  std::string ensemble, state_dir, pdb_dir, cfg_strategy;
  double pressure_atm = 0.0;
  int volume_interval = 0;
  if (ends_with(a.input, ".config")) {
    if (mcgpu_config_read(a.input.c_str(), &hb, &info) != MCGPU_OK) {
      std::fprintf(stderr, "Error: loadBoxData(): %s\n", mcgpu_last_error());
      std::fprintf(stderr, "Error: Could not build from file: %s\n", a.input.c_str());
      return EXIT_FAILURE;
    }
    steps = (long)info.steps;
    if (a.name.empty()) a.name = info.sim_name;
    ensemble = info.ensemble;
    state_dir = info.state_output;
    pdb_dir = info.pdb_output;
    cfg_strategy = info.strategy;
    pressure_atm = info.pressure;
    volume_interval = info.volume_interval;
  } else if (ends_with(a.input, ".state")) {
    int64_t st = 0;
    if (mcgpu_state_read(a.input.c_str(), &hb, &st) != MCGPU_OK) {
      std::fprintf(stderr, "Error: loadBoxData(): %s\n", mcgpu_last_error());
      std::fprintf(stderr, "Error: Could not build from file: %s\n", a.input.c_str());
      return EXIT_FAILURE;
    }
    steps = kDefaultSteps;
    start_step = (long)st;
  } else {
    std::fprintf(stderr, "%s: Must specify a config or state file\n", kApp);
    return EXIT_FAILURE;
  }
  if (a.steps > 0) steps = a.steps;
  // a .state written by this program after a run carries the run's ensemble
  mcgpu_resume_state resume{};
  int32_t resumed = 0;
  mcgpu_hostbox_get_resume(hb, &resume, &resumed);
  if (resumed && resume.ensemble == MCGPU_NPT && ensemble.empty()) {
    ensemble = "npt";
    volume_interval = resume.volume_interval;
    pressure_atm = resume.pressure / kAtmToKcalMolA3;
  }
  int neighbor = MCGPU_NEIGHBOR_BRUTE;
  const std::string strat = a.strategy.empty() ? cfg_strategy : a.strategy;
  if (!strategy_neighbor(strat, neighbor)) {
    std::fprintf(stderr, "%s: Unknown energy calculation strategy specified\n", kApp);
    return EXIT_FAILURE;
  }
  if (a.neighbor) neighbor = MCGPU_NEIGHBOR_CELL;
  const bool npt = ensemble == "npt";
  if (npt && volume_interval < 1) {
    std::fprintf(stderr, "Error: an npt run needs volume-interval > 0\n");
    return EXIT_FAILURE;
  }

  double temperature = 0.0;
  uint32_t seed = 0;
  mcgpu_hostbox_info(hb, &temperature, &seed);
  mcgpu_box_desc desc{};
  mcgpu_hostbox_desc(hb, &desc);
  double lrc = 0.0;
  mcgpu_lrc_for_run(&desc, npt ? MCGPU_NPT : MCGPU_NVT, &lrc);
  std::printf("Using seed: %d\n", (int)seed);

  // ---- device
  int ndev = 0;
  if (mcgpu_device_count(&ndev) != MCGPU_OK || ndev < 1) {
    std::printf("ERROR: Cannot find suitable GPU!\n");
    return EXIT_FAILURE;
  }
  std::printf("%d GPU device(s) found; running on GPU.\n", ndev);
  if (!a.name.empty()) std::printf("Simulation Name: %s\n", a.name.c_str());
  std::printf("Running %ld time steps", steps);
  if (start_step > 0) std::printf(" starting from step %ld", start_step);
  std::printf("\n");

  mcgpu_batch* batch = nullptr;
  if (mcgpu_batch_create(1, 0, &batch) != MCGPU_OK) die("mcgpu_batch_create");
  mcgpu_run_params p{};
  p.seed = seed;
  p.precision = a.mixed ? MCGPU_MIXED : MCGPU_FP64;
  p.ensemble = npt ? MCGPU_NPT : MCGPU_NVT;
  p.volume_interval = npt ? volume_interval : 0;
  p.steps = steps;
  p.pressure = pressure_atm * kAtmToKcalMolA3;
  p.max_volume_delta = npt ? a.max_volume_delta : 0.0;
  p.energy_offset = lrc;
  p.neighbor = neighbor;
  p.record_trace = 0;
  int32_t slot = -1;
  const auto t0 = std::chrono::steady_clock::now();
  // a .state written after a run of this program carries the generator
  // state, running energy and NPT schedule: the run continues that move
  // sequence (the reference re-seeds instead, Simulation.cpp:51)
  if (resumed) {
    if (mcgpu_batch_add_resumed(batch, &desc, &p, &resume, &slot) != MCGPU_OK)
      die("mcgpu_batch_add_resumed");
    if (a.verbose) std::printf("Continuing the move sequence of step %ld\n", start_step);
  } else if (mcgpu_batch_add(batch, &desc, &p, &slot) != MCGPU_OK) {
    die("mcgpu_batch_add");
  }

  const std::string base = a.name.empty() ? "untitled" : a.name;
  auto save_state = [&](long step) {
    double *x, *y, *z;
    mcgpu_hostbox_coords(hb, &x, &y, &z);
    if (mcgpu_batch_download(batch, slot, x, y, z) != MCGPU_OK) die("mcgpu_batch_download");
    mcgpu_box_stats st{};
    mcgpu_batch_stats(batch, slot, &st);
    mcgpu_hostbox_set_edges(hb, st.box);
    mcgpu_resume_state rs{};
    if (mcgpu_batch_get_resume(batch, slot, &rs) != MCGPU_OK) die("mcgpu_batch_get_resume");
    mcgpu_hostbox_set_resume(hb, &rs);
    std::string path = state_dir.empty() ? "" : state_dir + "/";
    path += base + "_" + std::to_string(step) + ".state";
    if (a.verbose) std::printf("Saving state file %s\n", path.c_str());
    if (mcgpu_state_write(hb, path.c_str(), step) != MCGPU_OK) die("mcgpu_state_write");
  };

  // ---- step loop: the device runs the moves; the host stops it at the
  // status and state boundaries
  mcgpu_box_stats st{};
  mcgpu_batch_stats(batch, slot, &st);
  const long status_every = (a.verbose && a.status_interval > 0) ? a.status_interval : 0;
  const long state_every = a.state_interval > 0 ? a.state_interval : 0;
  constexpr long kChunk = 20000;  // moves per launch at most
  long done = 0;
  while (true) {
    if (status_every && done % status_every == 0 && done < steps)
      std::printf("Step %ld:\n--Current Energy: %g\n\n", start_step + done, st.energy);
    if (state_every && done > 0 && done % state_every == 0 && done < steps)
      save_state(start_step + done);
    if (done >= steps) break;
    long next = steps;
    if (status_every) next = std::min(next, (done / status_every + 1) * status_every);
    if (state_every) next = std::min(next, (done / state_every + 1) * state_every);
    while (done < next) {
      const long k = std::min(next - done, kChunk);
      if (mcgpu_batch_run(batch, (int32_t)k, 1) != MCGPU_OK) die("mcgpu_batch_run");
      if (mcgpu_batch_stats(batch, slot, &st) != MCGPU_OK) die("mcgpu_batch_stats");
      if (st.fault) {
        std::fprintf(stderr, "Error: the box stopped on the device: %s\n", mcgpu_last_error());
        return EXIT_FAILURE;
      }
      done = (long)st.steps_done;
    }
  }
  const double run_s =
      std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();

  // ---- outputs (Simulation.cpp:266-339)
  {
    double *x, *y, *z;
    mcgpu_hostbox_coords(hb, &x, &y, &z);
    if (mcgpu_batch_download(batch, slot, x, y, z) != MCGPU_OK) die("mcgpu_batch_download");
    mcgpu_hostbox_set_edges(hb, st.box);
    std::string pdb = pdb_dir.empty() ? "" : pdb_dir + "/";
    pdb += (a.name.empty() ? std::string("run") : a.name) + ".pdb";
    if (mcgpu_pdb_write(hb, pdb.c_str()) != MCGPU_OK)
      std::fprintf(stderr, "warning: %s\n", mcgpu_last_error());
  }
  if (status_every)
    std::printf("Step %ld:\r\n--Current Energy: %g\n", start_step + steps, st.energy);
  if (a.state_interval >= 0) save_state(start_step + steps);

  const long accepted = (long)st.accepted;
  const long rejected = steps - accepted;
  std::printf("\nFinished running %ld steps\n", steps);
  std::printf("LJ-Energy Subtotal: %.3f\n", 0.0);
  std::printf("Charge Energy Subtotal: %.3f\n", 0.0);
  std::printf("Energy Long-range Correction: %.3f\n", lrc);
  std::printf("Ewald Energy: %.3f\n", 0.0);
  std::printf("Final Energy: %.3f\n", st.energy);
  std::printf("Run Time: %.3f seconds\n", run_s);
  std::printf("Accepted Moves: %ld\n", accepted);
  std::printf("Rejected Moves: %ld\n", rejected);
  std::printf("Acceptance Ratio: %.2f%%\n", 100.0 * accepted / (double)(accepted + rejected));
  if (npt)
    std::printf("Volume Moves: %lld accepted of %lld\n", (long long)st.volume_accepted,
                (long long)st.volume_moves);

  const std::string results = (a.name.empty() ? std::string("run") : a.name) + ".results";
  std::ofstream rf(results);
  rf << "######### MCGPU Results File #############" << std::endl;
  rf << "[Information]" << std::endl;
  rf << "Timestamp = " << now_string() << std::endl;
  if (!a.name.empty()) rf << "Simulation-Name = " << a.name << std::endl;
  rf << "Simulation-Mode = GPU" << std::endl;
  rf << "Starting-Step = " << start_step << std::endl;
  rf << "Steps = " << steps << std::endl;
  rf << "Molecule-Count = " << desc.n_molecules << std::endl << std::endl;
  rf << "[Results]" << std::endl;
  rf << "LJ-Energy Subtotal: " << 0.0 << std::endl;
  rf << "Charge Energy Subtotal: " << 0.0 << std::endl;
  rf << "Final-Energy = " << st.energy << std::endl;
  rf << "Run-Time = " << run_s << " seconds" << std::endl;
  rf << "Accepted-Moves = " << accepted << std::endl;
  rf << "Rejected-Moves = " << rejected << std::endl;
  rf << "Acceptance-Rate = " << 100.0f * accepted / (float)(accepted + rejected) << "%"
     << std::endl;
  rf.close();

  mcgpu_batch_destroy(batch);
  mcgpu_hostbox_free(hb);
  std::printf("Finishing simulation...\n\n");
  return EXIT_SUCCESS;
}
