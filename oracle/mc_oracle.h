/*
 * CPU oracle for the MCGPU energy hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's serial ("-s", BruteForce) energy
 * path, used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg as the checker.  The product (mcgpu_amd/) never links or calls it.
 *
 * Every function cites the reference file:line it follows.  Units: kcal/mol,
 * Angstrom, degrees (move), K.
 *
 * Pinning (what checks each function besides the GPU comparison):
 *   orc_glibc_*            libc rand() itself, and the known-answer values
 *                          for srand(1) / srand(12345).
 *   orc_lj / orc_coulomb   closed forms (r = sigma -> 0, r = 2^(1/6) sigma ->
 *                          -eps, Coulomb at 1 A with q = +-1 -> -+332.06).
 *   orc_molecule_energy,   an independent double loop in the tests, the
 *   orc_system_energy      invariant sum_i E_i(0) = 2 sum_i E_i(i), and the
 *                          per-move trace of the reference's own CPU path
 *                          (oracle/_ref, committed as tests/golden fixtures).
 *   orc_move / rollback    rigid-body invariants (distances preserved) and
 *                          the reference trace (every accepted move feeds the
 *                          next energies).
 *   orc_run_nvt            the reference trace: identical molecule choice,
 *                          dE and accept/reject for every move.
 *   orc_lrc_reference      the textbook single-species tail correction.
 *   orc_lrc_species        the textbook single-species value for two species
 *                          with identical LJ sites (N_A + N_B molecules), and
 *                          a numerical quadrature of the LJ tail integral,
 *                          with orc_lj_params as the pair potential, for a
 *                          two-species box of different sites.
 *   orc_volume_accept_prob the reference's own SimCalcs::acceptVolumeMove
 *                          (oracle/_ref/ref_volume) on a grid of (N, V, V',
 *                          dE, P, kT, seed): identical decisions; V' = V
 *                          gives exp(-dE/kT); falls as P grows for V' > V.
 *   orc_volume_scale       closed forms: internal geometry preserved,
 *                          primary positions multiplied by f.
 *   orc_run_npt            the build's own NPT semantics (the reference's
 *                          volume move is defective): its decisions use the
 *                          pinned functions above; the GPU path must give
 *                          the same move and volume-move sequence.
 */
#ifndef MCGPU_ORACLE_H
#define MCGPU_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_box {
  int32_t n_mol;
  int32_t n_site;
  int32_t n_pidx;
  double L[3];
  double cutoff;
  double kT;
  double max_translate;
  double max_rotate; /* degrees */
  double *x, *y, *z; /* [n_site], mutable */
  const double *sigma, *epsilon, *charge; /* [n_site] */
  const int32_t *mol_start, *mol_len, *mol_pidx_start, *mol_pidx_count;
  const int32_t *pidx; /* [n_pidx], absolute site indices */
} orc_box;

/* glibc TYPE_3 additive generator, restated (random_r.c algorithm). */
typedef struct orc_rng {
  int32_t state[31];
  int32_t f; /* front index (separation 3 ahead of r) */
  int32_t r; /* rear index */
} orc_rng;

void orc_glibc_srand(orc_rng* g, uint32_t seed);
int32_t orc_glibc_rand(orc_rng* g);
/* libc pass-through, so Python tests can compare against the real thing */
void orc_libc_srand(uint32_t seed);
int32_t orc_libc_rand(void);

/* MathLibrary.cpp:31-33 */
double orc_random_real(orc_rng* g, double start, double end);

/* SimulationStep.cpp:377-385, 332-342, 344-375 */
double orc_make_periodic(double x, double L);
double orc_dist2(const orc_box* b, int32_t a1, int32_t a2);
double orc_lj(const orc_box* b, int32_t a1, int32_t a2, double r2);
double orc_coulomb(const orc_box* b, int32_t a1, int32_t a2, double r);
/* closed-form helpers on raw parameters (same arithmetic) */
double orc_lj_params(double s1, double e1, double s2, double e2, double r2);
double orc_coulomb_params(double q1, double q2, double r);

/* SimulationStep.cpp:317-330 */
int32_t orc_in_range(const orc_box* b, int32_t m1, int32_t m2);
/* BruteForceStep.cpp:48-79 */
double orc_pair_sum(const orc_box* b, int32_t m1, int32_t m2);
/* BruteForceStep.cpp:15-46 */
double orc_molecule_energy(const orc_box* b, int32_t cur, int32_t start);
/* SimulationStep.cpp:40-71 (without the printed VERBOSE duplicate) */
double orc_system_energy(const orc_box* b);
/* SerialCalcs.cpp:37-108: the reference's two-solvent formula (molecules 0/1) */
double orc_lrc_reference(const orc_box* b);

/* Box.cpp:84-109 / SimulationStep.cpp:710-729 */
void orc_keep_in_box(orc_box* b, int32_t mol);
/* SimulationStep.cpp:520-566 with the draws given explicitly;
 * rollback_xyz receives 3*len doubles (x block, y block, z block). */
void orc_apply_move(orc_box* b, int32_t mol, int32_t vertex, const double d[3],
                    const double rot_deg[3], double* rollback_xyz);
/* SimulationStep.cpp:731-749 */
void orc_rollback(orc_box* b, int32_t mol, const double* rollback_xyz);

/* Per-move trace of the NVT loop (Simulation.cpp:219-264). */
typedef struct orc_trace {
  int32_t* mol;
  int32_t* vertex;
  double* old_e;
  double* new_e;
  double* de;
  int8_t* accepted;
} orc_trace;

/* Runs the reference's start-up energy + NVT step loop on b (mutated).
 * use_libc != 0 draws from libc srand/rand (as the reference does);
 * otherwise from the restated generator.  trace arrays may be NULL.
 * Returns the final energy (E0 + sum of accepted dE); *e0_out gets E0
 * (pairs + reference LRC), *n_acc the accepted count. */
double orc_run_nvt(orc_box* b, uint32_t seed, int64_t steps, int use_libc,
                   orc_trace* trace, double* e0_out, int64_t* n_acc);

/* Enumerates the molecule pairs (i < j) with orc_in_range true.
 * Writes up to cap pairs as (i, j) into out; returns the total count. */
int64_t orc_inrange_pairs(const orc_box* b, int32_t* out, int64_t cap);

/* ---- NPT volume move (the build's own semantics; see header) ---- */
/* Multiplies every molecule's first-primary position by f (each molecule
 * translated rigidly), and the box edges by f. */
void orc_volume_scale(orc_box* b, double f);
/* Tail correction summed over every molecule pair species (i.e. the correct
 * K-species form of SerialCalcs.cpp:37-108). */
double orc_lrc_species(const orc_box* b);
/* Volume-move acceptance probability (SimulationStep.cpp:476-488 formula):
 * exp(N ln(V'/V) - (dE + P dV)/kT) */
double orc_volume_accept_prob(int32_t n_mol, double v_old, double v_new,
                              double de, double pressure, double kT);

/* The build's NPT loop: the NVT loop above with a volume move after every
 * volume_interval moves while moves remain.  Volume move: u1, u2 drawn from
 * the MC stream; V' = V + dv (2 u1 - 1) (fixed width dv, in A^3); molecules
 * translated so first primaries scale by (V'/V)^(1/3); V' taken as the
 * product of the new edges; LRC (K-species form) scaled as 1/V; accept iff
 * orc_volume_accept_prob(N, V, V', E' - E, P, kT) >= u2, with E the running
 * energy; a proposal V' <= 0 is rejected. */
typedef struct orc_npt {
  int32_t volume_interval;
  int32_t reserved;
  double pressure; /* kcal/mol/A^3 */
  double dv;       /* half-width of the volume step, A^3 */
} orc_npt;

/* Per volume move: moves done before it, V, V', running energy before it,
 * E' at V' (pairs + LRC), u2, accepted.  Arrays of length cap (NULL to
 * skip); n receives the number of volume moves made. */
typedef struct orc_vtrace {
  int64_t cap;
  int64_t n;
  int64_t* step;
  double *v_old, *v_new, *e_old, *e_new, *u;
  int8_t* accepted;
} orc_vtrace;

double orc_run_npt(orc_box* b, uint32_t seed, int64_t steps, const orc_npt* p,
                   int use_libc, orc_trace* t, orc_vtrace* vt, double* e0_out,
                   int64_t* n_acc);

#ifdef __cplusplus
}
#endif
#endif
