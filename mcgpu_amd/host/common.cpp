// Shared host pieces: error reporting, descriptor validation, and the
// long-range (LJ tail) corrections.
#include <cmath>
#include <string>
#include <vector>

#include "hostbox.h"
#include "mcgpu_host.h"

namespace mcgpu {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

std::string validate_desc(const mcgpu_box_desc& d) {
  if (d.n_molecules < 1) return "n_molecules < 1";
  if (d.n_sites < d.n_molecules) return "n_sites < n_molecules";
  if (d.n_primary < d.n_molecules) return "n_primary < n_molecules";
  if (!d.x || !d.y || !d.z || !d.sigma || !d.epsilon || !d.charge)
    return "null site array";
  if (!d.mol_start || !d.mol_len || !d.mol_pidx_start || !d.mol_pidx_count ||
      !d.primary_index)
    return "null molecule array";
  for (int k = 0; k < 3; k++)
    if (!(d.box[k] > 0)) return "box edge <= 0";
  if (!(d.cutoff > 0)) return "cutoff <= 0";
  if (!(d.kT > 0)) return "kT <= 0";
  int32_t expect = 0;
  for (int32_t m = 0; m < d.n_molecules; m++) {
    if (d.mol_start[m] != expect) return "molecules must be contiguous and in order";
    if (d.mol_len[m] < 1 || d.mol_len[m] > MCGPU_MAX_MOL_SITES)
      return "molecule length outside [1, MCGPU_MAX_MOL_SITES]";
    expect += d.mol_len[m];
    if (d.mol_pidx_count[m] < 1) return "molecule without primary index";
    if (d.mol_pidx_count[m] > MCGPU_MAX_MOL_SITES)
      return "more than MCGPU_MAX_MOL_SITES primary indexes in a molecule";
    if (d.mol_pidx_start[m] < 0 ||
        d.mol_pidx_start[m] + d.mol_pidx_count[m] > d.n_primary)
      return "primary index range out of bounds";
    for (int32_t p = 0; p < d.mol_pidx_count[m]; p++) {
      const int32_t s = d.primary_index[d.mol_pidx_start[m] + p];
      if (s < d.mol_start[m] || s >= d.mol_start[m] + d.mol_len[m])
        return "primary index outside its molecule";
    }
  }
  if (expect != d.n_sites) return "sum of mol_len != n_sites";
  for (int32_t s = 0; s < d.n_sites; s++)
    if (!std::isfinite(d.x[s]) || !std::isfinite(d.y[s]) || !std::isfinite(d.z[s]))
      return "non-finite coordinate";
  return "";
}

HostBox HostBox::from_desc(const mcgpu_box_desc& d) {
  HostBox b;
  for (int k = 0; k < 3; k++) b.L[k] = d.box[k];
  b.cutoff = d.cutoff;
  b.kT = d.kT;
  b.temperature = d.kT / kBoltz;
  b.max_translate = d.max_translate;
  b.max_rotate = d.max_rotate;
  b.x.assign(d.x, d.x + d.n_sites);
  b.y.assign(d.y, d.y + d.n_sites);
  b.z.assign(d.z, d.z + d.n_sites);
  b.sigma.assign(d.sigma, d.sigma + d.n_sites);
  b.epsilon.assign(d.epsilon, d.epsilon + d.n_sites);
  b.charge.assign(d.charge, d.charge + d.n_sites);
  b.mol_start.assign(d.mol_start, d.mol_start + d.n_molecules);
  b.mol_len.assign(d.mol_len, d.mol_len + d.n_molecules);
  b.mol_pidx_start.assign(d.mol_pidx_start, d.mol_pidx_start + d.n_molecules);
  b.mol_pidx_count.assign(d.mol_pidx_count, d.mol_pidx_count + d.n_molecules);
  b.pidx.assign(d.primary_index, d.primary_index + d.n_primary);
  b.mol_type.assign(d.n_molecules, 0);
  return b;
}

double lrc_species(const mcgpu_box_desc& d);

// SerialCalcs.cpp:37-108 -- the reference's tail correction: "solvent 1"
// and "solvent 2" are molecules 0 and 1, each counted N/2 times.  The
// arithmetic order is kept so start energies match the reference exactly.
// Its pair loop indexes each molecule's coefficients with the other
// molecule's site index (SerialCalcs.cpp:98-104), so it is defined only when
// molecules 0 and 1 have the same length; other boxes (the reference reads
// past its arrays there) take the species form below.
double lrc_reference(const mcgpu_box_desc& d) {
  const int32_t ma = 0, mb = d.n_molecules > 1 ? 1 : 0;
  if (d.mol_len[ma] != d.mol_len[mb]) return lrc_species(d);
  const double PI_ = 3.14159265;  // MathLibrary.h:51
  double ecut = 0.0;
  const double vnew = d.box[0] * d.box[1] * d.box[2];
  const double rc3 = 1.00 / std::pow(d.cutoff, 3);
  const double rc9 = std::pow(rc3, 3);
  const double nmol1 = (double)(d.n_molecules / 2);
  const double nmol2 = (double)(d.n_molecules / 2);
  auto coeffs = [&](int32_t mol, std::vector<double>& c6, std::vector<double>& c12) {
    for (int32_t i = 0; i < d.mol_len[mol]; i++) {
      const int32_t s = d.mol_start[mol] + i;
      double sg = d.sigma[s], ep = d.epsilon[s];
      if (sg < 0 || ep < 0) sg = ep = 0.0;
      const double sig2 = std::pow(sg, 2), sig6 = std::pow(sig2, 3),
                   sig12 = std::pow(sig6, 2);
      c6.push_back(std::sqrt(4 * ep * sig6));
      c12.push_back(std::sqrt(4 * ep * sig12));
    }
  };
  std::vector<double> a6, a12, b6, b12;
  coeffs(ma, a6, a12);
  coeffs(mb, b6, b12);
  for (size_t i = 0; i < a6.size(); i++) {
    for (size_t j = 0; j < b6.size(); j++) {
      ecut += (2 * PI_ * nmol1 * nmol1 / (3.0 * vnew)) *
              (a12[i] * a12[j] * rc9 / 3.0 - a6[i] * a6[j] * rc3);
      ecut += (2 * PI_ * nmol2 * nmol2 / (3.0 * vnew)) *
              (b12[i] * b12[j] * rc9 / 3.0 - b6[i] * b6[j] * rc3);
      ecut += (4 * PI_ * nmol1 * nmol2 / (3.0 * vnew)) *
              (a12[i] * b12[j] * rc9 / 3.0 - a6[i] * b6[j] * rc3);
    }
  }
  return ecut;
}

// Tail correction over every pair of molecule species (molecules with the
// same length and per-site sigma/epsilon form a species):
//   E = (2 pi / V) sum_AB N_A N_B sum_ab 4 eps_ab sigma_ab^6
//       (sigma_ab^6 / (9 rc^9) - 1 / (3 rc^3))
double lrc_species(const mcgpu_box_desc& d) {
  std::vector<int32_t> rep;
  std::vector<double> cnt;
  auto same = [&](int32_t a, int32_t b) {
    if (d.mol_len[a] != d.mol_len[b]) return false;
    for (int32_t k = 0; k < d.mol_len[a]; k++) {
      const int32_t s = d.mol_start[a] + k, t = d.mol_start[b] + k;
      if (d.sigma[s] != d.sigma[t] || d.epsilon[s] != d.epsilon[t]) return false;
    }
    return true;
  };
  for (int32_t m = 0; m < d.n_molecules; m++) {
    size_t k = 0;
    for (; k < rep.size(); k++)
      if (same(rep[k], m)) break;
    if (k == rep.size()) {
      rep.push_back(m);
      cnt.push_back(0);
    }
    cnt[k] += 1.0;
  }
  const double v = d.box[0] * d.box[1] * d.box[2];
  const double rc = d.cutoff, rc3 = rc * rc * rc, rc9 = rc3 * rc3 * rc3;
  double e = 0.0;
  for (size_t A = 0; A < rep.size(); A++)
    for (size_t B = 0; B < rep.size(); B++)
      for (int32_t i = 0; i < d.mol_len[rep[A]]; i++)
        for (int32_t j = 0; j < d.mol_len[rep[B]]; j++) {
          const int32_t a = d.mol_start[rep[A]] + i, c = d.mol_start[rep[B]] + j;
          if (d.sigma[a] < 0 || d.epsilon[a] < 0 || d.sigma[c] < 0 || d.epsilon[c] < 0)
            continue;
          const double s2 = d.sigma[a] * d.sigma[c];
          const double s6 = s2 * s2 * s2;
          const double eps = std::sqrt(d.epsilon[a] * d.epsilon[c]);
          e += 2.0 * M_PI * cnt[A] * cnt[B] / v * 4.0 * eps * s6 *
               (s6 / (9.0 * rc9) - 1.0 / (3.0 * rc3));
        }
  return e;
}

}  // namespace mcgpu

extern "C" const char* mcgpu_last_error(void) {
  return mcgpu::g_last_error.c_str();
}

extern "C" int mcgpu_lrc_reference(const mcgpu_box_desc* d, double* e) {
  if (!d || !e) return MCGPU_ERR_INVALID;
  *e = mcgpu::lrc_reference(*d);
  return MCGPU_OK;
}

extern "C" int mcgpu_lrc_species(const mcgpu_box_desc* d, double* e) {
  if (!d || !e) return MCGPU_ERR_INVALID;
  *e = mcgpu::lrc_species(*d);
  return MCGPU_OK;
}

extern "C" int mcgpu_lrc_for_run(const mcgpu_box_desc* d, int32_t ensemble, double* e) {
  if (!d || !e || (ensemble != MCGPU_NVT && ensemble != MCGPU_NPT)) return MCGPU_ERR_INVALID;
  *e = ensemble == MCGPU_NPT ? mcgpu::lrc_species(*d) : mcgpu::lrc_reference(*d);
  return MCGPU_OK;
}
