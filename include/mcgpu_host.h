/*
 * mcgpu_host.h -- host-side C ABI around the energy path: synthetic
 * workload generators, the .state format, the plain (one launch per move)
 * MC driver and the serving loop.  Exported by mcgpu_amd/libmcgpu.so.
 *
 * Reference interfaces:
 *   .state read/write        src/Metropolis/Utilities/StateScanner.cpp:18-195,
 *                            235-281, 391-597; FileUtilities.cpp:119-283
 *   MC driver loop           src/Metropolis/Simulation.cpp:65-264
 *   RNG                      src/Metropolis/Utilities/MathLibrary.cpp:27-33
 *   LRC                      src/Metropolis/SerialSim/SerialCalcs.cpp:37-108
 */
#ifndef MCGPU_HOST_H
#define MCGPU_HOST_H

#include "mcgpu_energy.h"

#ifdef __cplusplus
extern "C" {
#endif

/* A host-resident box (owned arrays), convertible to an mcgpu_box_desc. */
typedef struct mcgpu_hostbox mcgpu_hostbox;

#define MCGPU_WL_WATER 0    /* rigid 3-site TIP3P water */
#define MCGPU_WL_BICYCLIC 1 /* rigid 16-site planar bicyclic (9 heavy + 7 H) */
#define MCGPU_WL_MIXTURE 2  /* water + 6-site methanol */

typedef struct mcgpu_workload_params {
  int32_t kind;        /* MCGPU_WL_* */
  int32_t n_molecules; /* total molecules */
  int32_t n_second;    /* MIXTURE: how many are methanol */
  int32_t lattice_jitter_permille; /* lattice jitter, 1/1000 of spacing */
  double density;      /* g/cm^3 (ignored when edge > 0) */
  double edge;         /* cubic box edge in A, or 0 to derive from density */
  double temperature;  /* K */
  double cutoff;       /* A; clamped to edge/2 */
  double max_translate;/* A */
  double max_rotate;   /* degrees */
  uint64_t seed;       /* generator seed (geometry, parameters, placement) */
  uint32_t mc_seed;    /* MC seed recorded in the box (srand) */
  int32_t reserved;
} mcgpu_workload_params;

int mcgpu_workload_generate(const mcgpu_workload_params* p,
                            mcgpu_hostbox** out);
int mcgpu_hostbox_free(mcgpu_hostbox* h);
/* Descriptor pointing into h's arrays (valid until h is freed or changed). */
int mcgpu_hostbox_desc(mcgpu_hostbox* h, mcgpu_box_desc* out);
int mcgpu_hostbox_info(mcgpu_hostbox* h, double* temperature,
                       uint32_t* mc_seed);
/* Writable views of the coordinate arrays (for loading final states). */
int mcgpu_hostbox_coords(mcgpu_hostbox* h, double** x, double** y, double** z);

/* Sets the box edges (e.g. after NPT volume moves, before a state write). */
int mcgpu_hostbox_set_edges(mcgpu_hostbox* h, const double* edges3);

/* .state files, full precision (%.17g) so a reload is bit exact.  A box
 * that carries a resume record (mcgpu_hostbox_set_resume) gets it written on
 * a last line that the reference's reader skips (StateScanner.cpp:54-195);
 * mcgpu_state_read restores it. */
int mcgpu_state_write(mcgpu_hostbox* h, const char* path, int64_t step);
int mcgpu_state_read(const char* path, mcgpu_hostbox** out, int64_t* step);
/* Attaches (r != NULL) or clears the box's resume record, and reads it
 * (*present = 0 when the box has none). */
int mcgpu_hostbox_set_resume(mcgpu_hostbox* h, const mcgpu_resume_state* r);
int mcgpu_hostbox_get_resume(mcgpu_hostbox* h, mcgpu_resume_state* r,
                             int32_t* present);
/* PDB snapshot in the reference's column layout (Simulation.cpp:362-417). */
int mcgpu_pdb_write(mcgpu_hostbox* h, const char* path);

/* .config runs (metrosim's z-matrix input route).  Replaces loadBoxData's
 * Configuration branch (FileUtilities.cpp:19-110): ConfigScanner
 * (ConfigScanner.cpp:20-265), OplsScanner (OplsScanner.cpp:18-101),
 * ZmatrixScanner (ZmatrixScanner.cpp:21-215), buildMoleculeXYZ
 * (MathLibrary.cpp:751-902), buildBoxData + generatefccBox
 * (FileUtilities.cpp:285-642), then keepMoleculeInBox per molecule
 * (Simulation.cpp:70-72) and addPrimaryIndexes (SimBoxBuilder.cpp:282-299).
 * Relative .par / z-matrix paths are opened from the current directory, as
 * the reference does.  The box holds the requested number of molecules. */
typedef struct mcgpu_config_info {
  int64_t steps;          /* "steps" */
  int32_t n_molecules;    /* "molecules" */
  int32_t n_types;        /* molecule types in the z-matrix */
  int32_t random_seed;    /* "random-seed" (the MC seed, also in the box) */
  int32_t volume_interval;/* "volume-interval" (0 if absent) */
  int32_t ewald_radius;   /* "ewald-radius" (0 if absent) */
  int32_t reserved;
  double pressure;        /* "pressure" (0 if absent) */
  char sim_name[256];     /* "sim-name" */
  char strategy[64];      /* "strategy" */
  char ensemble[64];      /* "ensemble" */
  char state_output[1024];/* "state-output" */
  char pdb_output[1024];  /* "pdb-output" */
} mcgpu_config_info;

/* Returns MCGPU_ERR_INVALID (with the message in mcgpu_last_error) for an
 * unreadable or malformed file, an unknown key, an OPLS type missing from
 * the .par file, or a primary-atom line whose group count differs from the
 * z-matrix's molecule types.  info may be NULL. */
int mcgpu_config_read(const char* path, mcgpu_hostbox** out,
                      mcgpu_config_info* info);

/* The reference's LRC (SerialCalcs.cpp:37-108) where it is defined -- its
 * pair loop needs molecules 0 and 1 of equal length; other boxes get the
 * K-species form -- and the K-species form itself. */
int mcgpu_lrc_reference(const mcgpu_box_desc* d, double* e);
int mcgpu_lrc_species(const mcgpu_box_desc* d, double* e);
/* The LRC a run starts from (its energy_offset): for MCGPU_NVT the
 * reference's formula, so start and final energies match the reference's
 * (Simulation.cpp:85); for MCGPU_NPT the K-species form, which volume moves
 * rescale with 1/V and whose value enters their acceptance test (the
 * reference's two-solvent formula counts N/2 molecules of molecules 0's and
 * 1's kinds, wrong for any other mixture). */
int mcgpu_lrc_for_run(const mcgpu_box_desc* d, int32_t ensemble, double* e);

/* ---- plain driver: host RNG + one fused trial-move launch per move ---- */
typedef struct mcgpu_plain_result {
  double start_energy; /* pairs + LRC */
  double lrc;
  double final_energy;
  int64_t accepted;
  int64_t rejected;
  double loop_seconds;    /* wall time of the step loop */
  double startup_seconds; /* wall time of the start-up energy */
  int64_t neighbor_builds;    /* Verlet list builds (mcgpu_box_neighbor_stats) */
  int64_t neighbor_fallbacks; /* evaluations that scanned every molecule */
} mcgpu_plain_result;

/* Runs `steps` NVT moves of box h (mutated to the final state) with the
 * reference's draw order and Metropolis rule.  Optional trace arrays of
 * length steps (NULL to skip).  neighbor: MCGPU_NEIGHBOR_*. */
int mcgpu_plain_run(mcgpu_hostbox* h, int32_t precision, int32_t neighbor,
                    int64_t steps, int32_t* trace_mol, double* trace_de,
                    int8_t* trace_acc, mcgpu_plain_result* out);

/* The same loop through the reference-shaped strategy calls (two
 * molecule-energy calls, changeMolecule, rollback per move), i.e. what a
 * drop-in GPU SimulationStep does. */
int mcgpu_strategy_run(mcgpu_hostbox* h, int64_t steps, int32_t* trace_mol,
                       double* trace_de, int8_t* trace_acc,
                       mcgpu_plain_result* out);

/* ---- serving loop ---- */
typedef struct mcgpu_server_config {
  int32_t max_boxes;      /* resident slots */
  int32_t sweeps_per_tick;/* at most this many moves per box per launch */
  size_t hbm_budget_bytes;/* admission budget */
  double tick_budget_us;  /* per-launch time limit (0 = none) */
} mcgpu_server_config;

/* Job states (mcgpu_job_result.state). */
#define MCGPU_JOB_QUEUED 0  /* submitted, not admitted yet */
#define MCGPU_JOB_RUNNING 1 /* resident on the device */
#define MCGPU_JOB_DONE 2    /* finished (status 0) or failed (status < 0) */

typedef struct mcgpu_job_result {
  int64_t job_id;
  double arrival_s;
  double admit_s;
  double finish_s;
  double final_energy; /* running energy (live while running) */
  double start_energy;
  int64_t steps;       /* requested moves; moves made once done */
  int64_t accepted;
  int32_t status; /* 0 ok, <0 error code */
  int32_t state;  /* MCGPU_JOB_* */
  int64_t steps_done;  /* moves made so far (updated after every tick) */
} mcgpu_job_result;

/* The serving loop runs on a host thread of the server, which alone drives
 * the device; every call below may come from any thread, also while the
 * loop runs.  The server clock starts at mcgpu_server_start. */
typedef struct mcgpu_server mcgpu_server;
int mcgpu_server_create(const mcgpu_server_config* cfg, mcgpu_server** out);
/* Stops the loop (see mcgpu_server_stop) and frees the server. */
int mcgpu_server_destroy(mcgpu_server* s);
/* Queues a job, before or after the start.  arrival_s: when the job becomes
 * visible on the server clock; < 0 means now.  The server reads `box` when
 * it admits the job: the caller keeps it alive until the job's state is no
 * longer MCGPU_JOB_QUEUED.  Refused (MCGPU_ERR_STATE) once stop was called. */
int mcgpu_server_submit(mcgpu_server* s, mcgpu_hostbox* box,
                        const mcgpu_run_params* params, double arrival_s,
                        int64_t* job_id);
/* Starts the clock and the continuous-batching loop on the server's thread;
 * returns at once. */
int mcgpu_server_start(mcgpu_server* s);
/* Blocks until job job_id (every submitted job when job_id < 0) is done,
 * or timeout_s passes (MCGPU_ERR_STATE; < 0 waits without limit).  Returns
 * the loop's error if the loop ended on one. */
int mcgpu_server_wait(mcgpu_server* s, int64_t job_id, double timeout_s);
/* Refuses further submissions, lets every submitted job finish, and joins
 * the loop; returns the loop's status. */
int mcgpu_server_stop(mcgpu_server* s);
/* Start, wait for every submitted job, stop: a replay of a job list whose
 * arrival times were given up front. */
int mcgpu_server_run(mcgpu_server* s);
/* One job's record as it stands (live progress while it runs). */
int mcgpu_server_poll(mcgpu_server* s, int64_t job_id, mcgpu_job_result* out);
int mcgpu_server_results(mcgpu_server* s, mcgpu_job_result* out, int64_t cap,
                         int64_t* n);
typedef struct mcgpu_server_stats {
  double wall_s;           /* server clock at the last completion */
  int64_t total_moves;     /* moves made so far by every job */
  int64_t ticks;
  double kernel_ms_total;  /* summed kernel time (complete once stopped) */
  int64_t peak_resident;
  size_t peak_bytes;       /* arena high-water mark (set once stopped) */
  int64_t work_candidates; /* summed over retired jobs (see mcgpu_box_stats) */
  int64_t work_sites;
  int32_t resident;        /* jobs on the device now */
  int32_t reserved;
  int64_t queued;          /* submitted, not admitted */
  int64_t completed;
  int64_t submitted;
  double clock_s;          /* server clock now */
} mcgpu_server_stats;
/* Live statistics (any thread, any time). */
int mcgpu_server_get_stats(mcgpu_server* s, mcgpu_server_stats* out);

#ifdef __cplusplus
}
#endif
#endif
