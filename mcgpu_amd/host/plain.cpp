// Plain GPU path: host RNG and orchestration (Simulation.cpp:219-264) with
// one fused trial-move launch per move, plus the reference-shaped strategy
// (GPUStep) that issues the same calls the reference's SimulationStep makes.
//
// Floating-point note: this file is compiled without FMA contraction so the
// proposed coordinates are the same doubles the reference's CPU path makes.
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "gpu_step.h"
#include "mcgpu_host.h"

namespace mcgpu {

HostBox& hostbox_ref(mcgpu_hostbox* h);
double lrc_reference(const mcgpu_box_desc& d);

namespace {

// SimulationStep.cpp:400-422 (u' = u cos + v sin, v' = v cos - u sin)
inline void rot_pair(double& u, double& v, double deg) {
  const double rad = deg * 3.14159265358979 / 180.0;
  const double ou = u, ov = v;
  u = ou * std::cos(rad) + ov * std::sin(rad);
  v = ov * std::cos(rad) - ou * std::sin(rad);
}

}  // namespace

void propose_move(const HostBox& hb, int mol, int vertex, const double d[3],
                  const double rot[3], double* out) {
  const int start = hb.mol_start[mol], len = hb.mol_len[mol];
  double* X = out;
  double* Y = out + len;
  double* Z = out + 2 * len;
  for (int k = 0; k < len; k++) {
    X[k] = hb.x[start + k];
    Y[k] = hb.y[start + k];
    Z[k] = hb.z[start + k];
  }
  const double px = X[vertex], py = Y[vertex], pz = Z[vertex];
  for (int k = 0; k < len; k++) {
    if (k == vertex) continue;
    X[k] += -px;
    Y[k] += -py;
    Z[k] += -pz;
    rot_pair(Y[k], Z[k], rot[0]);  // rotateX
    rot_pair(Z[k], X[k], rot[1]);  // rotateY
    rot_pair(X[k], Y[k], rot[2]);  // rotateZ
    X[k] += px;
    Y[k] += py;
    Z[k] += pz;
    X[k] += d[0];
    Y[k] += d[1];
    Z[k] += d[2];
  }
  X[vertex] += d[0];
  Y[vertex] += d[1];
  Z[vertex] += d[2];
  // keepMoleculeInBox by the first primary site
  const int p = hb.pidx[hb.mol_pidx_start[mol]] - start;
  double* C[3] = {X, Y, Z};
  for (int dim = 0; dim < 3; dim++) {
    if (C[dim][p] < 0) {
      for (int k = 0; k < len; k++) C[dim][k] += hb.L[dim];
    } else if (C[dim][p] > hb.L[dim]) {
      for (int k = 0; k < len; k++) C[dim][k] -= hb.L[dim];
    }
  }
}

void keep_in_box(HostBox& hb, int mol) {
  const int start = hb.mol_start[mol], end = start + hb.mol_len[mol];
  const int p = hb.pidx[hb.mol_pidx_start[mol]];
  std::vector<double>* C[3] = {&hb.x, &hb.y, &hb.z};
  for (int dim = 0; dim < 3; dim++) {
    std::vector<double>& c = *C[dim];
    if (c[p] < 0) {
      for (int j = start; j < end; j++) c[j] += hb.L[dim];
    } else if (c[p] > hb.L[dim]) {
      for (int j = start; j < end; j++) c[j] -= hb.L[dim];
    }
  }
}

namespace {

// Draws of one trial move, in the reference's order (draws #1-#8).
struct MoveDraws {
  int mol;
  int vertex;
  double d[3];
  double r[3];
};

// The molecule index is clamped to N-1 when rand() returns RAND_MAX (a
// 2^-31 event where the reference indexes past its arrays,
// SimulationStep.cpp:25-27); likewise the pivot index.
inline int clamp_index(double v, int n) {
  int i = (int)v;
  return i >= n ? n - 1 : i;
}

MoveDraws draw_move(const HostBox& hb, Rng31& rng) {
  MoveDraws m;
  m.mol = clamp_index(random_real(rng31_next(rng), 0, hb.n_mol()), hb.n_mol());
  const int len = hb.mol_len[m.mol];
  m.vertex = clamp_index(random_real(rng31_next(rng), 0, len), len);
  const double mt = hb.max_translate, mr = hb.max_rotate;
  for (int k = 0; k < 3; k++) m.d[k] = random_real(rng31_next(rng), -mt, mt);
  for (int k = 0; k < 3; k++) m.r[k] = random_real(rng31_next(rng), -mr, mr);
  return m;
}

double seconds_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

// ------------------------------------------------------------ GPUStep

GPUStep::GPUStep(HostBox* box, int precision, int neighbor, double skin)
    : hb_(box) {
  const mcgpu_box_desc d = hb_->desc();
  status_ = mcgpu_box_create(&d, precision, &dev_);
  if (!(skin > 0)) skin = 0.2875 * hb_->cutoff;
  if (status_ == MCGPU_OK && neighbor != MCGPU_NEIGHBOR_BRUTE)
    status_ = mcgpu_neighbor_rebuild(dev_, neighbor, skin);
}

GPUStep::~GPUStep() {
  if (dev_) mcgpu_box_destroy(dev_);
}

int GPUStep::chooseMolecule(Rng31& rng) const {
  return clamp_index(random_real(rng31_next(rng), 0, hb_->n_mol()), hb_->n_mol());
}

double GPUStep::calcMoleculeEnergy(int currMol, int startMol) {
  return calcMolecularEnergyContribution(currMol, startMol) + 0;
}

double GPUStep::calcMolecularEnergyContribution(int currMol, int startMol) {
  double e = 0.0;
  const int rc = mcgpu_molecule_energy(dev_, currMol, startMol, &e);
  if (rc != MCGPU_OK) status_ = rc;
  return e;
}

void GPUStep::changeMolecule(int molIdx, Rng31& rng) {
  const int len = hb_->mol_len[molIdx], start = hb_->mol_start[molIdx];
  int vertex = clamp_index(random_real(rng31_next(rng), 0, len), len);
  const double mt = hb_->max_translate, mr = hb_->max_rotate;
  double d[3], r[3];
  for (int k = 0; k < 3; k++) d[k] = random_real(rng31_next(rng), -mt, mt);
  for (int k = 0; k < 3; k++) r[k] = random_real(rng31_next(rng), -mr, mr);
  rollback_.resize(3 * len);
  scratch_.resize(3 * len);
  for (int k = 0; k < len; k++) {
    rollback_[k] = hb_->x[start + k];
    rollback_[len + k] = hb_->y[start + k];
    rollback_[2 * len + k] = hb_->z[start + k];
  }
  propose_move(*hb_, molIdx, vertex, d, r, scratch_.data());
  for (int k = 0; k < len; k++) {
    hb_->x[start + k] = scratch_[k];
    hb_->y[start + k] = scratch_[len + k];
    hb_->z[start + k] = scratch_[2 * len + k];
  }
  const int rc = mcgpu_set_molecule_coords(dev_, molIdx, scratch_.data());
  if (rc != MCGPU_OK) status_ = rc;
}

void GPUStep::rollback(int molIdx) {
  const int len = hb_->mol_len[molIdx], start = hb_->mol_start[molIdx];
  for (int k = 0; k < len; k++) {
    hb_->x[start + k] = rollback_[k];
    hb_->y[start + k] = rollback_[len + k];
    hb_->z[start + k] = rollback_[2 * len + k];
  }
  const int rc = mcgpu_set_molecule_coords(dev_, molIdx, rollback_.data());
  if (rc != MCGPU_OK) status_ = rc;
}

double GPUStep::calcSystemEnergy(double& subLJ, double& subCharge,
                                 int numMolecules) {
  (void)numMolecules;
  double e = 0.0;
  const int rc = mcgpu_system_energy(dev_, &e);
  if (rc != MCGPU_OK) status_ = rc;
  return subLJ + subCharge + e;
}

}  // namespace mcgpu

using namespace mcgpu;

// Start-up common to both drivers: Simulation.cpp:70-72 (wrap every
// molecule by its primary) and :85 (LRC).
static double startup(HostBox& hb) {
  for (int m = 0; m < hb.n_mol(); m++) keep_in_box(hb, m);
  const mcgpu_box_desc d = hb.desc();
  return lrc_reference(d);
}

extern "C" int mcgpu_plain_run(mcgpu_hostbox* h, int32_t precision,
                               int32_t neighbor, int64_t steps,
                               int32_t* trace_mol, double* trace_de,
                               int8_t* trace_acc, mcgpu_plain_result* out) {
  if (!h || !out || steps < 0) return MCGPU_ERR_INVALID;
  HostBox& hb = hostbox_ref(h);
  std::memset(out, 0, sizeof(*out));
  const double lrc = startup(hb);
  GPUStep step(&hb, precision, neighbor);
  if (step.status() != MCGPU_OK) return step.status();
  mcgpu_box* dev = step.device_box();
  auto t0 = std::chrono::steady_clock::now();
  double lj = 0, ch = 0;
  double energy = step.calcSystemEnergy(lj, ch, hb.n_mol());
  if (step.status() != MCGPU_OK) return step.status();
  energy += lrc;
  out->startup_seconds = seconds_since(t0);
  out->lrc = lrc;
  out->start_energy = energy;

  Rng31 rng;
  rng31_seed(rng, hb.mc_seed);
  std::vector<double> prop(3 * MCGPU_MAX_MOL_SITES);
  int64_t acc = 0;
  t0 = std::chrono::steady_clock::now();
  for (int64_t s = 0; s < steps; s++) {
    const MoveDraws m = draw_move(hb, rng);
    propose_move(hb, m.mol, m.vertex, m.d, m.r, prop.data());
    // Draw #9 is consumed only when dE >= 0 (SimulationStep.cpp:794-803):
    // peek it from a copy of the stream and advance the stream afterwards.
    Rng31 peek = rng;
    const double u = random_real(rng31_next(peek), 0.0, 1.0);
    mcgpu_move_result r;
    const int rc = mcgpu_trial_move(dev, m.mol, prop.data(), u, &r);
    if (rc != MCGPU_OK) return rc;
    if (!(r.delta < 0)) rng = peek;
    if (r.accepted) {
      acc++;
      energy += r.new_energy - r.old_energy;
      const int len = hb.mol_len[m.mol], start = hb.mol_start[m.mol];
      for (int k = 0; k < len; k++) {
        hb.x[start + k] = prop[k];
        hb.y[start + k] = prop[len + k];
        hb.z[start + k] = prop[2 * len + k];
      }
    }
    if (trace_mol) trace_mol[s] = m.mol;
    if (trace_de) trace_de[s] = r.delta;
    if (trace_acc) trace_acc[s] = (int8_t)r.accepted;
  }
  out->loop_seconds = seconds_since(t0);
  out->final_energy = energy;
  out->accepted = acc;
  out->rejected = steps - acc;
  return mcgpu_box_neighbor_stats(dev, &out->neighbor_builds, &out->neighbor_fallbacks);
}

extern "C" int mcgpu_strategy_run(mcgpu_hostbox* h, int64_t steps,
                                  int32_t* trace_mol, double* trace_de,
                                  int8_t* trace_acc, mcgpu_plain_result* out) {
  if (!h || !out || steps < 0) return MCGPU_ERR_INVALID;
  HostBox& hb = hostbox_ref(h);
  std::memset(out, 0, sizeof(*out));
  const double lrc = startup(hb);
  GPUStep step(&hb, MCGPU_FP64, MCGPU_NEIGHBOR_BRUTE);
  if (step.status() != MCGPU_OK) return step.status();
  auto t0 = std::chrono::steady_clock::now();
  double lj = 0, ch = 0;
  double energy = step.calcSystemEnergy(lj, ch, hb.n_mol()) + lrc;
  out->startup_seconds = seconds_since(t0);
  out->lrc = lrc;
  out->start_energy = energy;
  Rng31 rng;
  rng31_seed(rng, hb.mc_seed);
  int64_t acc = 0;
  t0 = std::chrono::steady_clock::now();
  // Simulation.cpp:219-264
  for (int64_t s = 0; s < steps; s++) {
    const int idx = step.chooseMolecule(rng);
    const double oldE = step.calcMoleculeEnergy(idx, 0);
    step.changeMolecule(idx, rng);
    const double newE = step.calcMoleculeEnergy(idx, 0);
    const double de = newE - oldE;
    bool accept;
    if (de < 0) {
      accept = true;
    } else {
      accept = std::exp(-(de) / hb.kT) >= random_real(rng31_next(rng), 0.0, 1.0);
    }
    if (accept) {
      acc++;
      energy += newE - oldE;
    } else {
      step.rollback(idx);
    }
    if (step.status() != MCGPU_OK) return step.status();
    if (trace_mol) trace_mol[s] = idx;
    if (trace_de) trace_de[s] = de;
    if (trace_acc) trace_acc[s] = (int8_t)accept;
  }
  out->loop_seconds = seconds_since(t0);
  out->final_energy = energy;
  out->accepted = acc;
  out->rejected = steps - acc;
  return MCGPU_OK;
}
