/*
 * CPU oracle for the MCGPU energy hot path -- TEST INFRASTRUCTURE ONLY.
 * See mc_oracle.h for scope and for what pins each function.
 *
 * The arithmetic follows the reference's double-precision expressions
 * operation by operation (same pow/sqrt calls, same evaluation order, no
 * FMA contraction), so on the same glibc it reproduces the reference's CPU
 * results bit for bit.
 */
#include "mc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ---------------------------------------------------------------- RNG */

/* glibc srandom_r for TYPE_3 (degree 31, separation 3): Park-Miller
 * multiplier 16807 via Schrage's decomposition, then 310 discarded outputs.
 * Reference use: MathLibrary.cpp:27-29 (srand). */
void orc_glibc_srand(orc_rng* g, uint32_t seed) {
  int32_t word;
  int i;
  if (seed == 0) seed = 1;
  g->state[0] = (int32_t)seed;
  word = (int32_t)seed;
  for (i = 1; i < 31; i++) {
    long hi = word / 127773;
    long lo = word % 127773;
    long w = 16807 * lo - 2836 * hi;
    if (w < 0) w += 2147483647;
    word = (int32_t)w;
    g->state[i] = word;
  }
  g->f = 3;
  g->r = 0;
  for (i = 0; i < 310; i++) (void)orc_glibc_rand(g);
}

/* glibc random_r for TYPE_3: state[f] += state[r]; output the sum >> 1.
 * Reference use: MathLibrary.cpp:32 (rand). */
int32_t orc_glibc_rand(orc_rng* g) {
  uint32_t v = (uint32_t)g->state[g->f] + (uint32_t)g->state[g->r];
  g->state[g->f] = (int32_t)v;
  g->f = (g->f + 1) % 31;
  g->r = (g->r + 1) % 31;
  return (int32_t)(v >> 1);
}

void orc_libc_srand(uint32_t seed) { srand(seed); }
int32_t orc_libc_rand(void) { return (int32_t)rand(); }

/* MathLibrary.cpp:31-33: (end-start) * ((Real)rand() / RAND_MAX) + start */
static double random_real_from(int32_t r, double start, double end) {
  return (end - start) * ((double)r / 2147483647) + start;
}

double orc_random_real(orc_rng* g, double start, double end) {
  return random_real_from(orc_glibc_rand(g), start, end);
}

/* ---------------------------------------------------------- pair math */

/* SimulationStep.cpp:377-385 (single wrap per dimension) */
double orc_make_periodic(double x, double L) {
  int lt = (x < -0.5 * L);
  x += lt * L;
  int gt = (x > 0.5 * L);
  x -= gt * L;
  return x;
}

/* SimulationStep.cpp:332-342: d = x[a2] - x[a1] */
double orc_dist2(const orc_box* b, int32_t a1, int32_t a2) {
  double dx = orc_make_periodic(b->x[a2] - b->x[a1], b->L[0]);
  double dy = orc_make_periodic(b->y[a2] - b->y[a1], b->L[1]);
  double dz = orc_make_periodic(b->z[a2] - b->z[a1], b->L[2]);
  return dx * dx + dy * dy + dz * dz;
}

/* SimulationStep.cpp:369-375 */
static double blend(double a, double c) {
  if (a * c >= 0) return sqrt(a * c);
  return sqrt(-1 * a * c);
}

/* SimulationStep.cpp:344-358 */
double orc_lj_params(double s1, double e1, double s2, double e2, double r2) {
  if (r2 == 0.0) return 0.0;
  const double sigma = blend(s1, s2);
  const double epsilon = blend(e1, e2);
  const double s2r2 = pow(sigma, 2) / r2;
  const double s6r6 = pow(s2r2, 3);
  const double s12r12 = pow(s6r6, 2);
  return 4.0 * epsilon * (s12r12 - s6r6);
}

double orc_lj(const orc_box* b, int32_t a1, int32_t a2, double r2) {
  return orc_lj_params(b->sigma[a1], b->epsilon[a1], b->sigma[a2],
                       b->epsilon[a2], r2);
}

/* SimulationStep.cpp:360-367 */
double orc_coulomb_params(double q1, double q2, double r) {
  if (r == 0.0) return 0.0;
  const double e = 332.06;
  return (q1 * q2 * e) / r;
}

double orc_coulomb(const orc_box* b, int32_t a1, int32_t a2, double r) {
  return orc_coulomb_params(b->charge[a1], b->charge[a2], r);
}

/* ------------------------------------------------------ molecule energy */

/* SimulationStep.cpp:317-330: any primary pair with r^2 <= rc^2 */
int32_t orc_in_range(const orc_box* b, int32_t m1, int32_t m2) {
  int out = 0;
  const int32_t p1s = b->mol_pidx_start[m1], p1e = p1s + b->mol_pidx_count[m1];
  const int32_t p2s = b->mol_pidx_start[m2], p2e = p2s + b->mol_pidx_count[m2];
  for (int32_t i = p1s; i < p1e; i++)
    for (int32_t j = p2s; j < p2e; j++)
      out |= (orc_dist2(b, b->pidx[i], b->pidx[j]) <= b->cutoff * b->cutoff);
  return out;
}

/* BruteForceStep.cpp:48-79 */
double orc_pair_sum(const orc_box* b, int32_t m1, int32_t m2) {
  double sum = 0;
  const int32_t s1 = b->mol_start[m1], e1 = s1 + b->mol_len[m1];
  const int32_t s2 = b->mol_start[m2], e2 = s2 + b->mol_len[m2];
  for (int32_t i = s1; i < e1; i++) {
    for (int32_t j = s2; j < e2; j++) {
      if (b->sigma[i] >= 0 && b->sigma[j] >= 0 && b->epsilon[i] >= 0 &&
          b->epsilon[j] >= 0) {
        const double r2 = orc_dist2(b, i, j);
        if (r2 == 0.0) {
          sum += 0.0;
        } else {
          sum += orc_lj(b, i, j, r2);
          sum += orc_coulomb(b, i, j, sqrt(r2));
        }
      }
    }
  }
  return sum;
}

/* BruteForceStep.cpp:15-46 */
double orc_molecule_energy(const orc_box* b, int32_t cur, int32_t start) {
  double total = 0;
  for (int32_t other = start; other < b->n_mol; other++) {
    if (other != cur && orc_in_range(b, cur, other))
      total += orc_pair_sum(b, cur, other);
  }
  return total;
}

/* SimulationStep.cpp:40-71: subLJ + subCharge (both 0 at start-up) plus
 * sum_mol calcMoleculeEnergy(mol, mol); intramolecular terms are disabled
 * (SimulationStep.h:15), so each term is the intermolecular contribution. */
double orc_system_energy(const orc_box* b) {
  double total = 0.0 + 0.0;
  for (int32_t m = 0; m < b->n_mol; m++) total += orc_molecule_energy(b, m, m);
  return total;
}

/* SerialCalcs.cpp:37-108, restated with the same operation order.  Uses the
 * first two molecules as "solvent 1" and "solvent 2" and N/2 of each. */
double orc_lrc_reference(const orc_box* b) {
  /* The pair loop below indexes each molecule's coefficients with the other
   * molecule's site index (SerialCalcs.cpp:98-104): defined only for equal
   * lengths of molecules 0 and 1.  The reference reads past its arrays
   * otherwise, which nothing can match; such boxes take the species form
   * (parity with the reference unpinned there). */
  if (b->mol_len[0] != b->mol_len[b->n_mol > 1 ? 1 : 0]) return orc_lrc_species(b);
  const double PI_ = 3.14159265; /* MathLibrary.h:51 */
  double ecut = 0.0;
  const double vnew = b->L[0] * b->L[1] * b->L[2];
  const double rc3 = 1.00 / pow(b->cutoff, 3);
  const double rc9 = pow(rc3, 3);
  const double nmol1 = (double)(b->n_mol / 2);
  const double nmol2 = (double)(b->n_mol / 2);
  const int32_t ma = 0, mb = b->n_mol > 1 ? 1 : 0;
  const int32_t na = b->mol_len[ma], nb = b->mol_len[mb];
  double* a6 = (double*)calloc((size_t)(na + nb) * 2, sizeof(double));
  double* a12 = a6 + na;
  double* b6 = a12 + na;
  double* b12 = b6 + nb;
  for (int32_t i = 0; i < na; i++) {
    const int32_t s = b->mol_start[ma] + i;
    double sg = b->sigma[s], ep = b->epsilon[s];
    if (sg < 0 || ep < 0) sg = ep = 0.0;
    const double sig2 = pow(sg, 2), sig6 = pow(sig2, 3), sig12 = pow(sig6, 2);
    a6[i] = sqrt(4 * ep * sig6);
    a12[i] = sqrt(4 * ep * sig12);
  }
  for (int32_t i = 0; i < nb; i++) {
    const int32_t s = b->mol_start[mb] + i;
    double sg = b->sigma[s], ep = b->epsilon[s];
    if (sg < 0 || ep < 0) sg = ep = 0.0;
    const double sig2 = pow(sg, 2), sig6 = pow(sig2, 3), sig12 = pow(sig6, 2);
    b6[i] = sqrt(4 * ep * sig6);
    b12[i] = sqrt(4 * ep * sig12);
  }
  for (int32_t i = 0; i < na; i++) {
    for (int32_t j = 0; j < nb; j++) {
      ecut += (2 * PI_ * nmol1 * nmol1 / (3.0 * vnew)) *
              (a12[i] * a12[j] * rc9 / 3.0 - a6[i] * a6[j] * rc3);
      ecut += (2 * PI_ * nmol2 * nmol2 / (3.0 * vnew)) *
              (b12[i] * b12[j] * rc9 / 3.0 - b6[i] * b6[j] * rc3);
      ecut += (4 * PI_ * nmol1 * nmol2 / (3.0 * vnew)) *
              (a12[i] * b12[j] * rc9 / 3.0 - a6[i] * b6[j] * rc3);
    }
  }
  free(a6);
  return ecut;
}

/* ------------------------------------------------------------- moves */

static double* coord(orc_box* b, int d) {
  return d == 0 ? b->x : (d == 1 ? b->y : b->z);
}

/* SimulationStep.cpp:710-729 (first primary of the molecule decides) */
void orc_keep_in_box(orc_box* b, int32_t mol) {
  const int32_t start = b->mol_start[mol], end = start + b->mol_len[mol];
  const int32_t p = b->pidx[b->mol_pidx_start[mol]];
  for (int d = 0; d < 3; d++) {
    double* c = coord(b, d);
    if (c[p] < 0) {
      for (int32_t j = start; j < end; j++) c[j] += b->L[d];
    } else if (c[p] > b->L[d]) {
      for (int32_t j = start; j < end; j++) c[j] -= b->L[d];
    }
  }
}

/* SimulationStep.cpp:400-422 */
static void rot_pair(double* u, double* v, double deg) {
  const double rad = deg * 3.14159265358979 / 180.0;
  const double ou = *u, ov = *v;
  *u = ou * cos(rad) + ov * sin(rad);
  *v = ov * cos(rad) - ou * sin(rad);
}

/* SimulationStep.cpp:520-566 (draws supplied by the caller) */
void orc_apply_move(orc_box* b, int32_t mol, int32_t vertex, const double d[3],
                    const double rot[3], double* rb) {
  const int32_t start = b->mol_start[mol], len = b->mol_len[mol];
  for (int32_t i = 0; i < len; i++) {
    const int32_t a = start + i, pv = start + vertex;
    rb[i] = b->x[a];
    rb[len + i] = b->y[a];
    rb[2 * len + i] = b->z[a];
    if (i == vertex) continue;
    /* rotateAtom (SimulationStep.cpp:387-398): pivot read live */
    const double px = b->x[pv], py = b->y[pv], pz = b->z[pv];
    b->x[a] += -px;
    b->y[a] += -py;
    b->z[a] += -pz;
    rot_pair(&b->y[a], &b->z[a], rot[0]); /* rotateX */
    rot_pair(&b->z[a], &b->x[a], rot[1]); /* rotateY */
    rot_pair(&b->x[a], &b->y[a], rot[2]); /* rotateZ */
    b->x[a] += px;
    b->y[a] += py;
    b->z[a] += pz;
    /* translateAtom */
    b->x[a] += d[0];
    b->y[a] += d[1];
    b->z[a] += d[2];
  }
  b->x[start + vertex] += d[0];
  b->y[start + vertex] += d[1];
  b->z[start + vertex] += d[2];
  orc_keep_in_box(b, mol);
}

/* SimulationStep.cpp:731-749 */
void orc_rollback(orc_box* b, int32_t mol, const double* rb) {
  const int32_t start = b->mol_start[mol], len = b->mol_len[mol];
  for (int32_t j = 0; j < len; j++) {
    b->x[start + j] = rb[j];
    b->y[start + j] = rb[len + j];
    b->z[start + j] = rb[2 * len + j];
  }
}

/* ---------------------------------------------------------- NVT loop */

static int32_t draw(orc_rng* g, int use_libc) {
  return use_libc ? (int32_t)rand() : orc_glibc_rand(g);
}

/* One NPT volume move (the build's own semantics; the reference's NPT path,
 * Ensemble.cpp:9-28, always grows the box by 1.1 and never recomputes the
 * energy).  u1, u2 ~ U(0,1) from the MC stream; V' = V + dv (2 u1 - 1), a
 * fixed-width (symmetric) step; molecules translated so first primaries
 * scale by f = (V'/V)^(1/3); the new volume is the product of the new edges,
 * as acceptVolumeMove reads it (SimulationStep.cpp:476-481); the LRC scales
 * as 1/V; accept iff acceptVolumeMove's probability (SimulationStep.cpp:
 * 482-487) >= u2.  A proposal V' <= 0 is rejected without a rescale. */
static int volume_move(orc_box* b, orc_rng* g, int use_libc, const orc_npt* p,
                       double* energy, double* lrc, double rec[4]) {
  const double u1 = random_real_from(draw(g, use_libc), 0.0, 1.0);
  const double u2 = random_real_from(draw(g, use_libc), 0.0, 1.0);
  const double v_old = b->L[0] * b->L[1] * b->L[2];
  const double v_prop = v_old + p->dv * (2.0 * u1 - 1.0);
  double v_new = v_prop, e_new = *energy;
  int accept = 0;
  if (v_prop > 0) {
    const size_t n = (size_t)b->n_site;
    double* bk = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    double L_old[3] = {b->L[0], b->L[1], b->L[2]};
    memcpy(bk, b->x, sizeof(double) * n);
    memcpy(bk + n, b->y, sizeof(double) * n);
    memcpy(bk + 2 * n, b->z, sizeof(double) * n);
    orc_volume_scale(b, cbrt(v_prop / v_old));
    v_new = b->L[0] * b->L[1] * b->L[2];
    const double lrc_new = *lrc * v_old / v_new;
    e_new = orc_system_energy(b) + lrc_new;
    const double prob = orc_volume_accept_prob(b->n_mol, v_old, v_new, e_new - *energy,
                                               p->pressure, b->kT);
    accept = prob >= u2;
    if (accept) {
      *energy = e_new;
      *lrc = lrc_new;
    } else {
      memcpy(b->x, bk, sizeof(double) * n);
      memcpy(b->y, bk + n, sizeof(double) * n);
      memcpy(b->z, bk + 2 * n, sizeof(double) * n);
      for (int d = 0; d < 3; d++) b->L[d] = L_old[d];
    }
    free(bk);
  }
  if (rec) {
    rec[0] = v_old;
    rec[1] = v_new;
    rec[2] = e_new;
    rec[3] = u2;
  }
  return accept;
}

/* Simulation.cpp:65-264 (BruteForce, serial), NVT or, with npt, the build's
 * NPT loop: a volume move after every volume_interval moves while moves
 * remain.  The molecule index is clamped to n_mol-1 for the
 * 2^-31-probability draw rand() == RAND_MAX, where the reference indexes one
 * past its arrays (SimulationStep.cpp:25-27). */
static double run_loop(orc_box* b, uint32_t seed, int64_t steps, int use_libc,
                       const orc_npt* npt, orc_trace* t, orc_vtrace* vt,
                       double* e0_out, int64_t* n_acc) {
  orc_rng g;
  int32_t maxlen = 0;
  for (int32_t m = 0; m < b->n_mol; m++) {
    orc_keep_in_box(b, m); /* Simulation.cpp:70-72 */
    if (b->mol_len[m] > maxlen) maxlen = b->mol_len[m];
  }
  if (use_libc)
    srand(seed);
  else
    orc_glibc_srand(&g, seed);
  /* NVT: the reference's LRC (Simulation.cpp:85); NPT: the K-species form,
   * which the volume moves rescale */
  double lrc = npt ? orc_lrc_species(b) : orc_lrc_reference(b);
  double energy = orc_system_energy(b) + lrc;
  if (e0_out) *e0_out = energy;
  double* rb = (double*)malloc(sizeof(double) * 3 * (size_t)(maxlen > 0 ? maxlen : 1));
  int64_t acc_count = 0, nv = 0;
  for (int64_t s = 0; s < steps; s++) {
    int32_t idx = (int32_t)random_real_from(draw(&g, use_libc), 0, b->n_mol);
    if (idx >= b->n_mol) idx = b->n_mol - 1;
    const double old_e = orc_molecule_energy(b, idx, 0);
    const int32_t len = b->mol_len[idx];
    int32_t vertex = (int32_t)random_real_from(draw(&g, use_libc), 0, len);
    if (vertex >= len) vertex = len - 1;
    const double mt = b->max_translate, mr = b->max_rotate;
    double d[3], r[3];
    d[0] = random_real_from(draw(&g, use_libc), -mt, mt);
    d[1] = random_real_from(draw(&g, use_libc), -mt, mt);
    d[2] = random_real_from(draw(&g, use_libc), -mt, mt);
    r[0] = random_real_from(draw(&g, use_libc), -mr, mr);
    r[1] = random_real_from(draw(&g, use_libc), -mr, mr);
    r[2] = random_real_from(draw(&g, use_libc), -mr, mr);
    orc_apply_move(b, idx, vertex, d, r, rb);
    const double new_e = orc_molecule_energy(b, idx, 0);
    const double de = new_e - old_e;
    /* SimulationStep.cpp:794-803 */
    int accept;
    if (de < 0) {
      accept = 1;
    } else {
      accept = exp(-(de) / b->kT) >= random_real_from(draw(&g, use_libc), 0.0, 1.0);
    }
    if (accept) {
      acc_count++;
      energy += new_e - old_e;
    } else {
      orc_rollback(b, idx, rb);
    }
    if (t) {
      if (t->mol) t->mol[s] = idx;
      if (t->vertex) t->vertex[s] = vertex;
      if (t->old_e) t->old_e[s] = old_e;
      if (t->new_e) t->new_e[s] = new_e;
      if (t->de) t->de[s] = de;
      if (t->accepted) t->accepted[s] = (int8_t)accept;
    }
    if (npt && (s + 1) % npt->volume_interval == 0 && s + 1 < steps) {
      double rec[4];
      const double e_before = energy;
      const int va = volume_move(b, &g, use_libc, npt, &energy, &lrc, rec);
      if (vt && nv < vt->cap) {
        if (vt->step) vt->step[nv] = s + 1;
        if (vt->v_old) vt->v_old[nv] = rec[0];
        if (vt->v_new) vt->v_new[nv] = rec[1];
        if (vt->e_old) vt->e_old[nv] = e_before;
        if (vt->e_new) vt->e_new[nv] = rec[2];
        if (vt->u) vt->u[nv] = rec[3];
        if (vt->accepted) vt->accepted[nv] = (int8_t)va;
      }
      nv++;
    }
  }
  free(rb);
  if (vt) vt->n = nv;
  if (n_acc) *n_acc = acc_count;
  return energy;
}

double orc_run_nvt(orc_box* b, uint32_t seed, int64_t steps, int use_libc,
                   orc_trace* t, double* e0_out, int64_t* n_acc) {
  return run_loop(b, seed, steps, use_libc, NULL, t, NULL, e0_out, n_acc);
}

double orc_run_npt(orc_box* b, uint32_t seed, int64_t steps, const orc_npt* p,
                   int use_libc, orc_trace* t, orc_vtrace* vt, double* e0_out,
                   int64_t* n_acc) {
  return run_loop(b, seed, steps, use_libc, p, t, vt, e0_out, n_acc);
}

int64_t orc_inrange_pairs(const orc_box* b, int32_t* out, int64_t cap) {
  int64_t n = 0;
  for (int32_t i = 0; i < b->n_mol; i++)
    for (int32_t j = i + 1; j < b->n_mol; j++)
      if (orc_in_range(b, i, j)) {
        if (out && n < cap) {
          out[2 * n] = i;
          out[2 * n + 1] = j;
        }
        n++;
      }
  return n;
}

/* ----------------------------------------------------------- NPT parts */

void orc_volume_scale(orc_box* b, double f) {
  for (int32_t m = 0; m < b->n_mol; m++) {
    const int32_t start = b->mol_start[m], end = start + b->mol_len[m];
    const int32_t p = b->pidx[b->mol_pidx_start[m]];
    for (int d = 0; d < 3; d++) {
      double* c = coord(b, d);
      const double shift = c[p] * f - c[p];
      for (int32_t j = start; j < end; j++) c[j] += shift;
    }
  }
  for (int d = 0; d < 3; d++) b->L[d] *= f;
}

/* Species = molecules with identical length and per-site (sigma, eps). */
static int same_species(const orc_box* b, int32_t m1, int32_t m2) {
  if (b->mol_len[m1] != b->mol_len[m2]) return 0;
  for (int32_t k = 0; k < b->mol_len[m1]; k++) {
    const int32_t a = b->mol_start[m1] + k, c = b->mol_start[m2] + k;
    if (b->sigma[a] != b->sigma[c] || b->epsilon[a] != b->epsilon[c]) return 0;
  }
  return 1;
}

double orc_lrc_species(const orc_box* b) {
  const double pi = 3.14159265358979323846;
  int32_t* rep = (int32_t*)malloc(sizeof(int32_t) * (size_t)b->n_mol);
  double* cnt = (double*)malloc(sizeof(double) * (size_t)b->n_mol);
  int32_t ns = 0;
  for (int32_t m = 0; m < b->n_mol; m++) {
    int32_t k = 0;
    for (; k < ns; k++)
      if (same_species(b, rep[k], m)) break;
    if (k == ns) {
      rep[ns] = m;
      cnt[ns] = 0;
      ns++;
    }
    cnt[k] += 1.0;
  }
  const double v = b->L[0] * b->L[1] * b->L[2];
  const double rc = b->cutoff, rc3 = rc * rc * rc, rc9 = rc3 * rc3 * rc3;
  double e = 0.0;
  for (int32_t A = 0; A < ns; A++)
    for (int32_t B = 0; B < ns; B++)
      for (int32_t i = 0; i < b->mol_len[rep[A]]; i++)
        for (int32_t j = 0; j < b->mol_len[rep[B]]; j++) {
          const int32_t a = b->mol_start[rep[A]] + i, c = b->mol_start[rep[B]] + j;
          if (b->sigma[a] < 0 || b->epsilon[a] < 0 || b->sigma[c] < 0 ||
              b->epsilon[c] < 0)
            continue;
          const double s2 = b->sigma[a] * b->sigma[c]; /* sigma_ab^2 */
          const double s6 = s2 * s2 * s2;
          const double eps = sqrt(b->epsilon[a] * b->epsilon[c]);
          e += 2.0 * pi * cnt[A] * cnt[B] / v * 4.0 * eps * s6 *
               (s6 / (9.0 * rc9) - 1.0 / (3.0 * rc3));
        }
  free(rep);
  free(cnt);
  return e;
}

/* SimulationStep.cpp:476-488 */
double orc_volume_accept_prob(int32_t n_mol, double v_old, double v_new,
                              double de, double pressure, double kT) {
  double prob = n_mol * (log(v_new) - log(v_old));
  prob -= (de + pressure * (v_new - v_old)) / kT;
  return exp(prob);
}
