// Synthetic workload generators for the energy path's benchmark and parity
// configurations: rigid TIP3P water, a rigid 16-site planar bicyclic
// molecule, and a water + methanol mixture.  Molecule geometries come from
// z-matrix rows (the reference's resources/exampleFiles/meoh.z and indole.z
// supply the bond lengths, angles and dihedrals used below); placement is a
// jittered cubic lattice with random orientations and overlap rejection.
// Everything is seeded; the same parameters give the same box bit for bit.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "hostbox.h"
#include "mcgpu_host.h"

namespace mcgpu {
namespace {

using Vec3 = std::array<double, 3>;

// splitmix64-seeded xoshiro256** generator.
class Gen {
 public:
  explicit Gen(uint64_t seed) {
    uint64_t s = seed;
    for (auto& w : st_) {
      s += 0x9E3779B97F4A7C15ull;
      uint64_t z = s;
      z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      w = z ^ (z >> 31);
    }
  }
  uint64_t next() {
    const uint64_t r = rotl(st_[1] * 5, 7) * 9;
    const uint64_t t = st_[1] << 17;
    st_[2] ^= st_[0];
    st_[3] ^= st_[1];
    st_[1] ^= st_[2];
    st_[0] ^= st_[3];
    st_[2] ^= t;
    st_[3] = rotl(st_[3], 45);
    return r;
  }
  double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
  double uniform(double a, double b) { return a + (b - a) * uniform(); }
  double normal() {
    double u1 = uniform();
    while (u1 <= 0.0) u1 = uniform();
    const double u2 = uniform();
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
  }
  // Uniform random rotation matrix (Shoemake's quaternion method).
  std::array<double, 9> rotation() {
    const double u1 = uniform(), u2 = uniform(), u3 = uniform();
    const double a = std::sqrt(1 - u1), b = std::sqrt(u1);
    const double qw = a * std::sin(2 * M_PI * u2), qx = a * std::cos(2 * M_PI * u2);
    const double qy = b * std::sin(2 * M_PI * u3), qz = b * std::cos(2 * M_PI * u3);
    return {1 - 2 * (qy * qy + qz * qz), 2 * (qx * qy - qz * qw), 2 * (qx * qz + qy * qw),
            2 * (qx * qy + qz * qw), 1 - 2 * (qx * qx + qz * qz), 2 * (qy * qz - qx * qw),
            2 * (qx * qz - qy * qw), 2 * (qy * qz + qx * qw), 1 - 2 * (qx * qx + qy * qy)};
  }

 private:
  static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
  uint64_t st_[4];
};

struct MolTemplate {
  std::vector<Vec3> pos;  // centred on the centroid
  std::vector<double> sigma, eps, q;
  std::vector<std::string> name;
  int primary = 0;  // 0-based
  double mass = 0;  // g/mol
};

Vec3 sub(const Vec3& a, const Vec3& b) { return {a[0] - b[0], a[1] - b[1], a[2] - b[2]}; }
Vec3 cross(const Vec3& a, const Vec3& b) {
  return {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
}
Vec3 unit(const Vec3& a) {
  const double n = std::sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
  return {a[0] / n, a[1] / n, a[2] / n};
}

// One z-matrix row: bond partner, bond length, angle partner, angle (deg),
// dihedral partner, dihedral (deg); partners are 1-based, 0 = none.
struct ZRow {
  int a;
  double r;
  int b;
  double theta;
  int c;
  double phi;
};

// Natural-extension reference frame placement of z-matrix rows.
std::vector<Vec3> zmatrix_to_xyz(const std::vector<ZRow>& rows) {
  std::vector<Vec3> p(rows.size(), Vec3{0, 0, 0});
  const double d2r = M_PI / 180.0;
  for (size_t k = 1; k < rows.size(); k++) {
    const ZRow& z = rows[k];
    const Vec3 A = p[z.a - 1];
    if (k == 1) {
      p[k] = {A[0] + z.r, A[1], A[2]};
      continue;
    }
    const Vec3 B = p[z.b - 1];
    Vec3 C;
    double phi = 0.0;
    if (z.c > 0) {
      C = p[z.c - 1];
      phi = z.phi * d2r;
    } else {
      C = {B[0], B[1] + 1.0, B[2]};  // any point off the A-B line, in-plane
    }
    const Vec3 bc = unit(sub(A, B));
    Vec3 n = cross(sub(B, C), bc);
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    if (nn < 1e-12) {
      C = {B[0] + 1.0, B[1], B[2]};
      n = cross(sub(B, C), bc);
    }
    n = unit(n);
    const Vec3 m = cross(n, bc);
    const double th = z.theta * d2r;
    const double d0 = -z.r * std::cos(th), d1 = z.r * std::sin(th) * std::cos(phi),
                 d2 = z.r * std::sin(th) * std::sin(phi);
    for (int d = 0; d < 3; d++) p[k][d] = A[d] + d0 * bc[d] + d1 * m[d] + d2 * n[d];
  }
  return p;
}

void center(MolTemplate& t) {
  Vec3 c{0, 0, 0};
  for (auto& v : t.pos)
    for (int d = 0; d < 3; d++) c[d] += v[d] / (double)t.pos.size();
  for (auto& v : t.pos)
    for (int d = 0; d < 3; d++) v[d] -= c[d];
}

// TIP3P water: O-H 0.9572 A, H-O-H 104.52 deg (oplsaa.par types 111/112).
MolTemplate water_template() {
  MolTemplate t;
  const double r = 0.9572, a = 104.52 * M_PI / 180.0;
  t.pos = {Vec3{0, 0, 0}, Vec3{r, 0, 0}, Vec3{r * std::cos(a), r * std::sin(a), 0}};
  t.sigma = {3.15061, 0.0, 0.0};
  t.eps = {0.1521, 0.0, 0.0};
  t.q = {-0.834, 0.417, 0.417};
  t.name = {"OW", "HW", "HW"};
  t.primary = 0;
  t.mass = 18.015;
  center(t);
  return t;
}

// Methanol, geometry of resources/exampleFiles/meoh.z (site order O, H(O),
// C, 3 x H(C)); OPLS-like parameters C 3.5/0.066 +0.145, O 3.12/0.17
// -0.683, H(O) 0/0 +0.418, H(C) 2.5/0.03 +0.04.
MolTemplate methanol_template() {
  MolTemplate t;
  std::vector<ZRow> rows = {{0, 0, 0, 0, 0, 0},
                            {1, 0.945698, 0, 0, 0, 0},
                            {1, 1.411870, 2, 108.989752, 0, 0},
                            {3, 1.090450, 1, 110.239001, 2, 179.999996},
                            {3, 1.090404, 1, 110.549523, 4, 119.850735},
                            {3, 1.090404, 1, 110.549526, 4, 240.149263}};
  t.pos = zmatrix_to_xyz(rows);
  t.sigma = {3.12, 0.0, 3.5, 2.5, 2.5, 2.5};
  t.eps = {0.17, 0.0, 0.066, 0.03, 0.03, 0.03};
  t.q = {-0.683, 0.418, 0.145, 0.04, 0.04, 0.04};
  t.name = {"O", "HO", "C", "HC", "HC", "HC"};
  t.primary = 0;
  t.mass = 32.042;
  center(t);
  return t;
}

// Planar bicyclic 16-site molecule: the indole skeleton of
// resources/exampleFiles/indole.z (9 heavy atoms, then 7 H), with a seeded
// in-plane perturbation of every site and seeded LJ / charge parameters:
// heavy sigma U[3.25,3.55], eps U[0.05,0.17]; H sigma 2.42, eps 0.03;
// q ~ N(0, 0.2) shifted so the molecule is neutral.  Primary = site 2.
MolTemplate bicyclic_template(Gen& g) {
  MolTemplate t;
  std::vector<ZRow> rows = {{0, 0, 0, 0, 0, 0},
                            {1, 1.397692, 0, 0, 0, 0},
                            {2, 1.396917, 1, 119.074387, 0, 0},
                            {1, 1.414955, 2, 123.498339, 3, 0.0},
                            {4, 1.407110, 1, 115.697714, 2, 0.0},
                            {5, 1.402345, 4, 121.754194, 1, 0.0},
                            {1, 1.380828, 2, 133.920701, 3, 180.0},
                            {4, 1.400742, 5, 133.107107, 6, 180.0},
                            {8, 1.369106, 4, 106.368117, 5, 180.0},
                            {2, 1.080000, 1, 120.462733, 4, 180.0},
                            {3, 1.080000, 2, 120.373225, 1, 180.0},
                            {5, 1.080000, 4, 119.122960, 1, 180.0},
                            {6, 1.080000, 5, 119.639092, 4, 180.0},
                            {7, 1.010000, 1, 125.392798, 2, 0.0},
                            {8, 1.080000, 4, 123.965934, 5, 0.0},
                            {9, 1.080000, 8, 131.479309, 4, 180.0}};
  t.pos = zmatrix_to_xyz(rows);
  // The rows are planar (dihedrals 0/180) in the plane of sites 1-3, which
  // zmatrix_to_xyz puts in z = 0; perturb in that plane only.
  for (auto& v : t.pos) {
    v[0] += g.uniform(-0.03, 0.03);
    v[1] += g.uniform(-0.03, 0.03);
    v[2] = 0.0;
  }
  double qsum = 0;
  for (int k = 0; k < 16; k++) {
    const bool heavy = k < 9;
    t.sigma.push_back(heavy ? g.uniform(3.25, 3.55) : 2.42);
    t.eps.push_back(heavy ? g.uniform(0.05, 0.17) : 0.03);
    t.q.push_back(0.2 * g.normal());
    qsum += t.q.back();
    t.name.push_back(heavy ? (k == 6 ? "N" : "C") : "H");
  }
  for (auto& q : t.q) q -= qsum / 16.0;
  t.primary = 1;
  t.mass = 117.151;
  center(t);
  return t;
}

// Grid of placed sites for the overlap test (periodic, cubic box).
class SiteGrid {
 public:
  SiteGrid(double L, double cell) : L_(L) {
    n_ = std::max(1, (int)std::floor(L / cell));
    w_ = L / n_;
    cells_.resize((size_t)n_ * n_ * n_);
  }
  // Minimum periodic distance from p to any stored site.
  double min_dist(const Vec3& p) const {
    int c[3];
    for (int d = 0; d < 3; d++) c[d] = cell_of(p[d]);
    double best = 1e300;
    const int r = n_ >= 3 ? 1 : 0;
    auto scan = [&](int idx) {
      for (const Vec3& q : cells_[idx]) {
        double s = 0;
        for (int d = 0; d < 3; d++) {
          double dd = q[d] - p[d];
          dd -= L_ * std::round(dd / L_);
          s += dd * dd;
        }
        best = std::min(best, s);
      }
    };
    if (r == 0) {
      for (size_t i = 0; i < cells_.size(); i++) scan((int)i);
    } else {
      for (int dx = -1; dx <= 1; dx++)
        for (int dy = -1; dy <= 1; dy++)
          for (int dz = -1; dz <= 1; dz++)
            scan(index(c[0] + dx, c[1] + dy, c[2] + dz));
    }
    return std::sqrt(best);
  }
  void add(const Vec3& p) {
    cells_[index(cell_of(p[0]), cell_of(p[1]), cell_of(p[2]))].push_back(p);
  }

 private:
  int cell_of(double v) const {
    double u = v - L_ * std::floor(v / L_);
    int c = (int)(u / w_);
    return c >= n_ ? n_ - 1 : c;
  }
  int index(int a, int b, int c) const {
    a = ((a % n_) + n_) % n_;
    b = ((b % n_) + n_) % n_;
    c = ((c % n_) + n_) % n_;
    return (a * n_ + b) * n_ + c;
  }
  double L_, w_;
  int n_;
  std::vector<std::vector<Vec3>> cells_;
};

// Places molecules (template index per molecule) on a jittered cubic lattice
// with random orientations; each molecule tries up to `tries` orientations
// and keeps the first whose sites are all >= dmin from every placed site
// (else the best one found).  The molecule is then wrapped so its first
// primary site lies in [0, L).
void place(HostBox& hb, const std::vector<const MolTemplate*>& mols, double L,
           double jitter_frac, double dmin, int tries, Gen& g) {
  const int n = (int)mols.size();
  int m = 1;
  while (m * m * m < n) m++;
  const double a = L / m;
  SiteGrid grid(L, std::max(dmin, 2.0));
  std::vector<Vec3> cand, best;
  for (int i = 0; i < n; i++) {
    const MolTemplate& t = *mols[i];
    const int ix = i / (m * m), iy = (i / m) % m, iz = i % m;
    const Vec3 c0 = {(ix + 0.5) * a, (iy + 0.5) * a, (iz + 0.5) * a};
    double best_d = -1;
    for (int k = 0; k < tries; k++) {
      const auto R = g.rotation();
      Vec3 c = c0;
      for (int d = 0; d < 3; d++) c[d] += g.uniform(-jitter_frac, jitter_frac) * a;
      cand.resize(t.pos.size());
      double md = 1e300;
      for (size_t s = 0; s < t.pos.size(); s++) {
        const Vec3& v = t.pos[s];
        for (int d = 0; d < 3; d++)
          cand[s][d] = c[d] + R[3 * d] * v[0] + R[3 * d + 1] * v[1] + R[3 * d + 2] * v[2];
        md = std::min(md, grid.min_dist(cand[s]));
      }
      if (md > best_d) {
        best_d = md;
        best = cand;
      }
      if (md >= dmin) break;
    }
    // Wrap by the primary site into [0, L).
    const Vec3 p = best[t.primary];
    Vec3 shift;
    for (int d = 0; d < 3; d++) shift[d] = -L * std::floor(p[d] / L);
    const int32_t start = hb.n_site();
    hb.mol_start.push_back(start);
    hb.mol_len.push_back((int32_t)t.pos.size());
    hb.mol_pidx_start.push_back((int32_t)hb.pidx.size());
    hb.mol_pidx_count.push_back(1);
    hb.pidx.push_back(start + t.primary);
    for (size_t s = 0; s < best.size(); s++) {
      Vec3 v = best[s];
      for (int d = 0; d < 3; d++) v[d] += shift[d];
      hb.x.push_back(v[0]);
      hb.y.push_back(v[1]);
      hb.z.push_back(v[2]);
      hb.sigma.push_back(t.sigma[s]);
      hb.epsilon.push_back(t.eps[s]);
      hb.charge.push_back(t.q[s]);
      hb.site_name.push_back(t.name[s]);
      grid.add(v);
    }
  }
}

}  // namespace

int generate(const mcgpu_workload_params& p, HostBox& hb) {
  Gen g(p.seed);
  const int n = p.n_molecules;
  if (n < 2) return MCGPU_ERR_INVALID;
  MolTemplate water = water_template();
  MolTemplate meoh = methanol_template();
  MolTemplate bic;
  std::vector<const MolTemplate*> mols(n);
  std::vector<int32_t> types(n, 0);
  double mass = 0, dmin = 1.6;
  if (p.kind == MCGPU_WL_WATER) {
    for (int i = 0; i < n; i++) mols[i] = &water;
    mass = n * water.mass;
    hb.primary_line = "1";
  } else if (p.kind == MCGPU_WL_BICYCLIC) {
    bic = bicyclic_template(g);
    for (int i = 0; i < n; i++) mols[i] = &bic;
    mass = n * bic.mass;
    hb.primary_line = "2";
    dmin = 2.2;
  } else if (p.kind == MCGPU_WL_MIXTURE) {
    if (p.n_second < 0 || p.n_second > n) return MCGPU_ERR_INVALID;
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    for (int i = n - 1; i > 0; i--) {  // Fisher-Yates shuffle of species
      int j = (int)(g.uniform() * (i + 1));
      if (j > i) j = i;
      std::swap(order[i], order[j]);
    }
    for (int i = 0; i < n; i++) {
      const bool is_meoh = order[i] < p.n_second;
      mols[i] = is_meoh ? &meoh : &water;
      types[i] = is_meoh ? 1 : 0;
    }
    mass = (n - p.n_second) * water.mass + p.n_second * meoh.mass;
    hb.primary_line = "1,1";
  } else {
    return MCGPU_ERR_INVALID;
  }
  double L = p.edge;
  if (L <= 0) {
    if (p.density <= 0) return MCGPU_ERR_INVALID;
    // rho [g/cm^3] = mass [g/mol] / (V [A^3] * 0.60221)
    L = std::cbrt(mass / (p.density * 0.60221));
  }
  hb.L[0] = hb.L[1] = hb.L[2] = L;
  hb.cutoff = std::min(p.cutoff, 0.5 * L);
  hb.temperature = p.temperature;
  hb.kT = p.temperature * kBoltz;
  hb.max_translate = p.max_translate;
  hb.max_rotate = p.max_rotate;
  hb.mc_seed = p.mc_seed;
  const double jit = (p.lattice_jitter_permille > 0 ? p.lattice_jitter_permille : 100) / 1000.0;
  // The bicyclic molecules jam at 1.1 g/cm3: most placements exhaust every
  // try, so fewer tries cost little quality and keep 16k-molecule boxes fast.
  place(hb, mols, L, jit, dmin, p.kind == MCGPU_WL_BICYCLIC ? 32 : 300, g);
  hb.mol_type = types;
  return MCGPU_OK;
}

}  // namespace mcgpu

using mcgpu::HostBox;

struct mcgpu_hostbox {
  HostBox b;
};

extern "C" int mcgpu_workload_generate(const mcgpu_workload_params* p,
                                       mcgpu_hostbox** out) {
  if (!p || !out) return MCGPU_ERR_INVALID;
  auto* h = new (std::nothrow) mcgpu_hostbox();
  if (!h) return MCGPU_ERR_NOMEM;
  int rc = mcgpu::generate(*p, h->b);
  if (rc != MCGPU_OK) {
    delete h;
    mcgpu::set_error("mcgpu_workload_generate: invalid parameters");
    return rc;
  }
  *out = h;
  return MCGPU_OK;
}

extern "C" int mcgpu_hostbox_free(mcgpu_hostbox* h) {
  delete h;
  return MCGPU_OK;
}

extern "C" int mcgpu_hostbox_desc(mcgpu_hostbox* h, mcgpu_box_desc* out) {
  if (!h || !out) return MCGPU_ERR_INVALID;
  *out = h->b.desc();
  return MCGPU_OK;
}

extern "C" int mcgpu_hostbox_info(mcgpu_hostbox* h, double* temperature,
                                  uint32_t* mc_seed) {
  if (!h) return MCGPU_ERR_INVALID;
  if (temperature) *temperature = h->b.temperature;
  if (mc_seed) *mc_seed = h->b.mc_seed;
  return MCGPU_OK;
}

extern "C" int mcgpu_hostbox_coords(mcgpu_hostbox* h, double** x, double** y,
                                    double** z) {
  if (!h || !x || !y || !z) return MCGPU_ERR_INVALID;
  *x = h->b.x.data();
  *y = h->b.y.data();
  *z = h->b.z.data();
  return MCGPU_OK;
}

extern "C" int mcgpu_hostbox_set_edges(mcgpu_hostbox* h, const double* edges3) {
  if (!h || !edges3) return MCGPU_ERR_INVALID;
  for (int k = 0; k < 3; k++)
    if (!(edges3[k] > 0)) {
      mcgpu::set_error("mcgpu_hostbox_set_edges: edge <= 0");
      return MCGPU_ERR_INVALID;
    }
  for (int k = 0; k < 3; k++) h->b.L[k] = edges3[k];
  return MCGPU_OK;
}

namespace mcgpu {
HostBox& hostbox_ref(mcgpu_hostbox* h) { return h->b; }
mcgpu_hostbox* hostbox_new(HostBox&& b) {
  auto* h = new (std::nothrow) mcgpu_hostbox();
  if (h) h->b = std::move(b);
  return h;
}
}  // namespace mcgpu
