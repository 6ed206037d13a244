/*
 * mcgpu_energy.h -- C ABI of the MI355X-native MCGPU energy hot path.
 *
 * One shared library (mcgpu_amd/libmcgpu.so) exports every function below.
 * Plain pointers and sizes only; host buffers are owned by the caller, device
 * buffers by the library.  Every call returns 0 (MCGPU_OK) or a negative
 * error code; mcgpu_last_error() gives the message of the calling thread's
 * last failure.  A handle is not thread-safe: one host thread per handle.
 *
 * Reference interfaces each entry point replaces (paths relative to the
 * MCGPU source tree):
 *   SimBox SoA + GPUCopy residency     src/Metropolis/SimBox.h:18-612,
 *                                      src/Metropolis/GPUCopy.cpp:97-222
 *   calcMolecularEnergyContribution    src/Metropolis/SimulationStep.h:54,
 *                                      src/Metropolis/BruteForceStep.cpp:15-79
 *   calcSystemEnergy                   src/Metropolis/SimulationStep.h:93,
 *                                      src/Metropolis/SimulationStep.cpp:40-71
 *   changeMolecule / rollback          src/Metropolis/SimulationStep.h:74,83,
 *                                      src/Metropolis/SimulationStep.cpp:520-566,731-749
 *   NLC / Verlet rebuild               src/Metropolis/NLCStep.cpp:125-253,
 *                                      src/Metropolis/VerletStep.cpp:119-207
 *   NPT resize + volume acceptance     src/Metropolis/SimulationStep.cpp:476-518,
 *                                      src/Metropolis/Ensemble.cpp:9-28
 *   MC step loop (batched, on device)  src/Metropolis/Simulation.cpp:219-264
 */
#ifndef MCGPU_ENERGY_H
#define MCGPU_ENERGY_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define MCGPU_OK 0
#define MCGPU_ERR_INVALID (-1)  /* bad argument or descriptor */
#define MCGPU_ERR_HIP (-2)      /* a HIP runtime call failed */
#define MCGPU_ERR_NOMEM (-3)    /* device or host allocation failed / budget */
#define MCGPU_ERR_RANGE (-4)    /* index out of range */
#define MCGPU_ERR_STATE (-5)    /* call not valid in the current state */
#define MCGPU_ERR_NODEVICE (-6) /* no HIP device visible */

/* ---- pair arithmetic ---- */
#define MCGPU_FP64 0  /* every term and sum in FP64 */
#define MCGPU_MIXED 1 /* 1/r estimate and LJ terms in FP32 (near-contact
                         LJ in FP64); Coulomb from the FP32 1/r with one
                         FP64 Newton step; displacements, sums, coordinates
                         and Metropolis in FP64 */

/* ---- neighbour structure used by energy calls ---- */
#define MCGPU_NEIGHBOR_BRUTE 0  /* scan every molecule (BruteForceStep) */
#define MCGPU_NEIGHBOR_CELL 1   /* cell list, width >= rc + primary spread */
#define MCGPU_NEIGHBOR_VERLET 2 /* Verlet list with skin */

/* ---- ensembles ---- */
#define MCGPU_NVT 0
#define MCGPU_NPT 1

/*
 * Box descriptor: the SimBox fields the energy path reads
 * (SimBox.h: atomCoordinates, atomData, moleculeData rows MOL_START,
 * MOL_LEN, MOL_PIDX_START, MOL_PIDX_COUNT, primaryIndexes, size, cutoff,
 * kT, maxTranslate, maxRotate).  All arrays are host memory, read during
 * mcgpu_box_create only.  Sites with sigma < 0 or epsilon < 0 are dummy
 * sites: they move with their molecule but never interact
 * (BruteForceStep.cpp:64-65).
 */
typedef struct mcgpu_box_desc {
  int32_t n_molecules;
  int32_t n_sites;
  int32_t n_primary; /* length of primary_index */
  int32_t reserved;
  double box[3];        /* edge lengths, Angstrom */
  double cutoff;        /* molecule cutoff on primary sites, Angstrom */
  double kT;            /* kcal/mol (T * 0.00198717) */
  double max_translate; /* Angstrom */
  double max_rotate;    /* degrees */
  const double* x;      /* [n_sites] */
  const double* y;
  const double* z;
  const double* sigma;   /* [n_sites], Angstrom */
  const double* epsilon; /* [n_sites], kcal/mol */
  const double* charge;  /* [n_sites], e */
  const int32_t* mol_start;      /* [n_molecules] first site */
  const int32_t* mol_len;        /* [n_molecules] site count */
  const int32_t* mol_pidx_start; /* [n_molecules] into primary_index */
  const int32_t* mol_pidx_count; /* [n_molecules] */
  const int32_t* primary_index;  /* [n_primary] absolute site index */
} mcgpu_box_desc;

/* Largest molecule the kernels accept (sites per molecule). */
#define MCGPU_MAX_MOL_SITES 32

typedef struct mcgpu_box mcgpu_box;

const char* mcgpu_last_error(void);
int mcgpu_device_count(int* count);
/* Blocks until all work the library queued on the device is done. */
int mcgpu_device_synchronize(void);

/* Creates the device-resident box (replaces SimBoxBuilder::build +
 * GPUCopy::copyIn).  precision: MCGPU_FP64 or MCGPU_MIXED. */
int mcgpu_box_create(const mcgpu_box_desc* desc, int precision,
                     mcgpu_box** out);
int mcgpu_box_destroy(mcgpu_box* box);
int mcgpu_box_set_precision(mcgpu_box* box, int precision);
/* Selects the neighbour structure and (re)builds it.  kind: MCGPU_NEIGHBOR_*.
 * skin is used for MCGPU_NEIGHBOR_VERLET only (Angstrom). */
int mcgpu_neighbor_rebuild(mcgpu_box* box, int kind, double skin);
/* Neighbour-structure counters (VerletCalcs::updateVerlet,
 * VerletStep.cpp:186-207, rebuilds when a primary moves more than skin/2;
 * this box rebuilds before the next evaluation after such a commit):
 * builds = Verlet list builds since creation (the first included);
 * fallbacks = evaluations that scanned every molecule although cell or
 * Verlet lists were selected (results stay exact, only slower). */
int mcgpu_box_neighbor_stats(mcgpu_box* box, int64_t* builds, int64_t* fallbacks);

/* Whole-box coordinate transfer (GPUCopy::copyIn / copyOut).  Coordinates
 * handed in here, to mcgpu_set_molecule_coords or to mcgpu_trial_move need
 * not be wrapped into [0, L): distances take the reference's single wrap
 * (SimulationStep.cpp:377-385), and cell lists bin a molecule by the periodic
 * image of its first primary (up to one box length outside). */
int mcgpu_upload_coords(mcgpu_box* box, const double* x, const double* y,
                        const double* z);
int mcgpu_download_coords(mcgpu_box* box, double* x, double* y, double* z);
/* Overwrites one molecule's sites: xyz = x block, y block, z block, each
 * mol_len doubles (the shape of SimBox::rollBackCoordinates). */
int mcgpu_set_molecule_coords(mcgpu_box* box, int32_t mol, const double* xyz);
int mcgpu_get_molecule_coords(mcgpu_box* box, int32_t mol, double* xyz);

/* E_mol(start) = sum over j in [start_mol, N), j != mol, molecules in range,
 * of all real site pairs -- exactly calcMolecularEnergyContribution
 * (BruteForceStep.cpp:15-46).  Synchronous. */
int mcgpu_molecule_energy(mcgpu_box* box, int32_t mol, int32_t start_mol,
                          double* energy);

/* Sum_i E_i(i): the full-system intermolecular energy
 * (SimulationStep.cpp:40-71 without the LRC).  Synchronous. */
int mcgpu_system_energy(mcgpu_box* box, double* energy);
/* Number of real site pairs summed by the last mcgpu_system_energy call. */
int mcgpu_system_pair_count(mcgpu_box* box, int64_t* pairs);

/* mcgpu_system_energy without allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): the same
 * kernels and result as mcgpu_system_energy, with the per-molecule partials
 * and totals in a fixed device array of the library (2 MiB per GPU, held
 * until the process ends) instead of the box's buffers; the call makes no
 * hipMalloc/hipFree.  Boxes of more than 131,072 molecules are refused with
 * MCGPU_ERR_RANGE.  Calls that overlap on one GPU share that array and race:
 * serialise them.  Synchronous. */
int mcgpu_system_energy_noalloc(mcgpu_box* box, double* energy);

/* One trial move, fused: evaluates E_mol(0) at the current (old) and at the
 * proposed coordinates in one pass, applies the Metropolis test of
 * SimulationStep.cpp:794-803 with the caller's uniform draw u (used only when
 * dE >= 0), and commits the proposed coordinates on acceptance.
 * new_xyz: x block, y block, z block (mol_len each).  Synchronous. */
typedef struct mcgpu_move_result {
  double old_energy;
  double new_energy;
  double delta;   /* new_energy - old_energy */
  int32_t accepted;
  int32_t reserved;
} mcgpu_move_result;
int mcgpu_trial_move(mcgpu_box* box, int32_t mol, const double* new_xyz,
                     double u, mcgpu_move_result* result);

/* NPT: rescales the box by factor (molecules translated rigidly, primary
 * site position multiplied by factor; replaces SimCalcs::resizeBox) and
 * rebuilds the neighbour structure. */
int mcgpu_volume_scale(mcgpu_box* box, double factor);
int mcgpu_box_size(mcgpu_box* box, double* edges3);

/* Device timing of the last kernel launched by a synchronous box call, in
 * milliseconds (HIP events on the box's stream; enable first). */
int mcgpu_box_enable_timing(mcgpu_box* box, int on);
int mcgpu_last_kernel_ms(mcgpu_box* box, double* ms);

/* Test hook (no counterpart in the reference): evaluates the device sin,
 * cos and exp used for trial-move rotations and the Metropolis test. */
int mcgpu_debug_math(const double* x, int32_t n, double* sin_out,
                     double* cos_out, double* exp_out);

/* Test hook (no counterpart in the reference): evaluates the device FP64
 * 1/sqrt(x) of the pair energies, which stands for the reference's
 * r = sqrt(r2) and its divisions by r2 and r (SimulationStep.cpp:120,
 * 344-367). */
int mcgpu_debug_rsqrt(const double* x, int32_t n, double* out);

/* ------------------------------------------------------------------ */
/* Batched boxes: many independent boxes advanced by one persistent launch.
 * Each box carries its own device RNG (a glibc TYPE_3 stream seeded like
 * srand(seed), MathLibrary.cpp:27-33), its own Metropolis test and energy.
 * A sweep slot = one trial move for every active box. */

typedef struct mcgpu_run_params {
  uint32_t seed;          /* MC seed (srand) */
  int32_t precision;      /* MCGPU_FP64 / MCGPU_MIXED */
  int32_t ensemble;       /* MCGPU_NVT / MCGPU_NPT */
  int32_t volume_interval;/* NPT: a volume move every this many moves */
  int64_t steps;          /* trial moves to run for this box */
  double pressure;        /* NPT, kcal/mol/A^3 (1 atm = 1.4584e-5) */
  double max_volume_delta;/* NPT, in (0, 0.5): volume step dV = U(-1,1) *
                             this * V0, V0 the volume at admission (a fixed,
                             symmetric step width) */
  double energy_offset;   /* added to the start energy (e.g. the LRC) */
  int32_t neighbor;       /* MCGPU_NEIGHBOR_BRUTE or MCGPU_NEIGHBOR_CELL */
  int32_t record_trace;   /* nonzero: keep per-move (mol, dE, accepted) */
} mcgpu_run_params;

typedef struct mcgpu_box_stats {
  int64_t steps_done;
  int64_t accepted;
  int64_t volume_moves;
  int64_t volume_accepted;
  double energy;  /* running total energy (start energy + accepted dE) */
  double start_energy;
  double box[3];  /* current edges (change under NPT) */
  /* Work done so far: candidate molecules examined and sites loaded (in-range
   * molecules' sites plus the moving molecule's old and proposed sites).
   * Algorithmic bytes = 24 * (work_candidates + work_sites). */
  int64_t work_candidates;
  int64_t work_sites;
  int32_t done;
  int32_t fault;  /* nonzero: box stopped (see mcgpu_last_error on add) */
  /* Moves of a cell-list box evaluated by scanning every molecule (the
   * lists were incomplete after a cell overflowed, or the move's cell range
   * exceeded the scan's limit); results stay exact, only slower. */
  int64_t neighbor_fallbacks;
} mcgpu_box_stats;

typedef struct mcgpu_batch mcgpu_batch;

/* hbm_budget_bytes: admission limit for all resident boxes (0 = no limit). */
int mcgpu_batch_create(int32_t max_boxes, size_t hbm_budget_bytes,
                       mcgpu_batch** out);
int mcgpu_batch_destroy(mcgpu_batch* b);
/* Makes the batch's device (the one current at mcgpu_batch_create) current
 * on the calling thread, for a batch driven from another host thread. */
int mcgpu_batch_bind_thread(mcgpu_batch* b);
/* Device bytes a box would take (for admission decisions). */
int mcgpu_box_footprint(const mcgpu_box_desc* desc,
                        const mcgpu_run_params* params, size_t* bytes);
/* Adds a box; computes its start energy on the device. Returns
 * MCGPU_ERR_NOMEM if the budget or slot count is exhausted. */
int mcgpu_batch_add(mcgpu_batch* b, const mcgpu_box_desc* desc,
                    const mcgpu_run_params* params, int32_t* slot);
/* Frees a slot.  It waits for a tick in flight only when that tick runs the
 * slot's box (a box that was done or paused when the tick was launched is
 * freed at once). */
int mcgpu_batch_remove(mcgpu_batch* b, int32_t slot);

/* Everything a box needs to continue its move sequence exactly where it
 * stopped: its generator state, running energy and LRC, the NPT step width,
 * and counters since its first start.  The reference re-seeds on resume
 * (Simulation.cpp:51, with the seed its .state reader loses), so its resumed
 * runs start a new stream; a box added with this record continues the old
 * one, and its moves, decisions and energies equal those of the run that was
 * not interrupted. */
typedef struct mcgpu_resume_state {
  int32_t rng_words[31];   /* glibc TYPE_3 state words */
  int32_t rng_front;       /* front index (rear = front - 3 mod 31) */
  int32_t rng_rear;
  int32_t volume_pending;  /* NPT: the volume move due at steps_done is not
                              made yet (a run that ended on a volume stop) */
  int64_t steps_done;      /* moves since the first start */
  int64_t accepted;
  int64_t volume_moves;    /* NPT volume moves since the first start */
  int64_t volume_accepted;
  double energy;           /* running total energy */
  double lrc;              /* current LRC (NPT: rescaled with 1/V) */
  double volume_step;      /* NPT half-width of the volume step, A^3 */
  /* the run's ensemble settings, for a caller that continues the run from
     a .state file alone (mcgpu_batch_add_resumed takes its params) */
  int32_t ensemble;        /* MCGPU_NVT / MCGPU_NPT */
  int32_t volume_interval;
  double pressure;         /* kcal/mol/A^3 */
} mcgpu_resume_state;
/* Reads a box's resume record (after the work queued so far). */
int mcgpu_batch_get_resume(mcgpu_batch* b, int32_t slot, mcgpu_resume_state* out);
/* mcgpu_batch_add for a box that continues from a resume record: params
 * give the run (its steps are the moves of this run; params->seed and
 * energy_offset are not used); the generator, running energy, LRC and NPT
 * step width come from r, and the NPT volume schedule keeps its phase
 * (volume moves after every volume_interval moves counted from the first
 * start). */
int mcgpu_batch_add_resumed(mcgpu_batch* b, const mcgpu_box_desc* desc,
                            const mcgpu_run_params* params,
                            const mcgpu_resume_state* r, int32_t* slot);
/* Advances every active box by up to `sweeps` moves in one persistent
 * launch (captured and replayed as a HIP graph when the box set is
 * unchanged).  NPT boxes that stopped at a volume move in the previous tick
 * take it first, beside the other boxes, and then move until the same
 * deadline.  Asynchronous unless sync != 0. */
int mcgpu_batch_run(mcgpu_batch* b, int32_t sweeps, int32_t sync);
/* A box's statistics after the work queued so far (waits for a tick in
 * flight). */
int mcgpu_batch_stats(mcgpu_batch* b, int32_t slot, mcgpu_box_stats* out);
/* The same as of the last tick consumed (by mcgpu_batch_run or any call that
 * waits), without waiting for a tick in flight: a serving loop launches the
 * next tick first and then reads, retires and admits while it runs.  No
 * reference counterpart. */
int mcgpu_batch_peek(mcgpu_batch* b, int32_t slot, mcgpu_box_stats* out);
int mcgpu_batch_download(mcgpu_batch* b, int32_t slot, double* x, double* y,
                         double* z);
/* Copies the recorded trace (if record_trace was set): up to cap moves. */
int mcgpu_batch_trace(mcgpu_batch* b, int32_t slot, int32_t* mol,
                      double* delta, int8_t* accepted, int64_t cap,
                      int64_t* n);
/* One NPT volume move as decided on the device (record_trace boxes): moves
 * done before it, V, V' (product of the proposed edges; <= 0 for a proposal
 * rejected outright), the running energy before it, E' = pairs at V' plus
 * the rescaled LRC, the uniform u of the test, and the decision
 * exp(N ln(V'/V) - (E' - E + P (V' - V)) / kT) >= u, which is
 * SimCalcs::acceptVolumeMove (SimulationStep.cpp:476-488). */
typedef struct mcgpu_volume_record {
  int64_t step;
  double v_old;
  double v_new;
  double e_old;
  double e_new;
  double u;
  int32_t accepted;
  int32_t reserved;
} mcgpu_volume_record;
/* Copies up to cap volume-move records of an NPT box added with
 * record_trace (MCGPU_ERR_STATE otherwise); n = records copied. */
int mcgpu_batch_volume_trace(mcgpu_batch* b, int32_t slot,
                             mcgpu_volume_record* out, int64_t cap, int64_t* n);
/* Per-launch time limit (microseconds, 0 = none): a box stops its sweeps
 * when its work-group has run this long, so small boxes do not wait idle
 * for the slowest one.  Results do not depend on it (each box's move
 * sequence is fixed by its RNG stream), only how moves split over launches. */
int mcgpu_batch_set_tick_budget(mcgpu_batch* b, double microseconds);
/* Arena bytes in use, peak, and capacity (the admission budget). */
int mcgpu_batch_memory(mcgpu_batch* b, size_t* used, size_t* peak,
                       size_t* capacity);
/* Work-group size the batch uses for a box (256, 512 or 1024 lanes). */
int mcgpu_batch_group(mcgpu_batch* b, int32_t slot, int32_t* threads);
/* Work-group-size tuner (on by default; active only with a tick budget):
 * per box class (neighbour source, precision, largest molecule, molecule
 * count in quarter-octaves) the first box of the class runs one full tick
 * at each size and the size with the most moves per tick is kept for the
 * class.  Off: sizes come from a density estimate of the per-move work. */
int mcgpu_batch_set_autotune(mcgpu_batch* b, int32_t on);
/* The tuned classes: key, chosen lanes, and moves per tick measured at
 * 256 / 512 / 1024 lanes (3 per class); n = number of tuned classes. */
int mcgpu_batch_tuning(mcgpu_batch* b, int64_t* keys, int32_t* threads,
                       double* moves_per_tick, int32_t cap, int32_t* n);
/* Host seconds spent so far waiting for ticks to finish, queueing NPT
 * volume moves, and launching ticks. */
int mcgpu_batch_host_seconds(mcgpu_batch* b, double* wait, double* volume,
                             double* launch);
/* Average device duration (ms) of the batch kernel's main launch (the
 * boxes that do not start with a volume move) over the mcgpu_batch_run
 * calls since the previous query that had one, and their count. */
int mcgpu_batch_kernel_ms(mcgpu_batch* b, double* avg_ms, int64_t* launches);

#ifdef __cplusplus
}
#endif
#endif
