// metrosim's .config input route: the key=value run file, the OPLS .par
// parameter table and the BOSS-style z-matrix, turned into the host box the
// energy path runs on.  The start coordinates are the ones the reference
// builds: each molecule type's internal coordinates from its z-matrix rows,
// replicated, placed on an FCC lattice and wrapped into the box by the first
// primary site.  Reference:
//   ConfigScanner::readInConfig, parsePrimaryIndexDefinitions
//                              ConfigScanner.cpp:20-265
//   OplsScanner::readInOpls, addLineToTable, checkFormat, getAtom
//                              OplsScanner.cpp:18-101, 165-172
//   ZmatrixScanner::readInZmatrix, parseLine, checkFormat
//                              ZmatrixScanner.cpp:21-215
//   buildMoleculeXYZ, cross    MathLibrary.cpp:751-902, 946-962
//   loadBoxData, buildBoxData, generatefccBox
//                              FileUtilities.cpp:19-110, 285-541, 543-642
//   keepMoleculeInBox          Box.cpp:84-109 (called per molecule by
//                              Simulation.cpp:70-72)
//   addPrimaryIndexes          SimBoxBuilder.cpp:282-299
// Bonds, angles and dihedrals are read only to place the sites: the
// reference's intramolecular terms are disabled (SimulationStep.h:15), so the
// box keeps rigid molecules.  The reference also requires an oplsaa.sb next
// to the .par file (FileUtilities.cpp:63-71); its bond and angle data feed
// only those disabled terms, so this loader does not read it.
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "hostbox.h"
#include "mcgpu_host.h"

namespace mcgpu {
HostBox& hostbox_ref(mcgpu_hostbox* h);
mcgpu_hostbox* hostbox_new(HostBox&& b);
std::vector<std::vector<int>> parse_primary_line(const std::string& s);

namespace {

// Degrees to radians as the reference converts them (MathLibrary.h:51).
constexpr double kPiRef = 3.14159265;

std::string trim_ws(const std::string& s) {
  const char* ws = " \n\r\t";
  const size_t b = s.find_first_not_of(ws);
  if (b == std::string::npos) return "";
  const size_t e = s.find_last_not_of(ws);
  return s.substr(b, e - b + 1);
}

// ---------------------------------------------------------------- .config --

struct RunConfig {
  double x = 0, y = 0, z = 0, temp = 0, max_t = 0, max_r = 0, cutoff = 0;
  double pressure = 0;
  int molecules = 0;
  long steps = 0;
  int seed = 0;
  int volume_interval = 0, ewald_radius = 0;
  std::string par, zmatrix, state_in, state_out, pdb_out, name, strategy,
      ensemble, primary_line;
};

bool read_run_config(const std::string& path, RunConfig& c, std::string& err) {
  std::ifstream in(path);
  if (!in.is_open()) {
    err = "unable to open configuration file (" + path + ")";
    return false;
  }
  std::string raw;
  while (std::getline(in, raw)) {
    const std::string line = trim_ws(raw);
    if (line.empty() || line[0] == '#' || line[0] == ';') continue;
    const size_t eq = line.find('=');
    if (eq == std::string::npos) {
      err = "configuration file not well formed: \"" + line + "\"";
      return false;
    }
    const std::string key = line.substr(0, eq);
    const std::string val = line.substr(eq + 1);
    const bool has = !val.empty();
    const char* v = val.c_str();
    auto missing = [&](const char* what) {
      err = std::string("configuration file: missing ") + what + " value";
      return false;
    };
    if (key == "x") {
      if (!has) return missing("x");
      c.x = std::atof(v);
    } else if (key == "y") {
      if (!has) return missing("y");
      c.y = std::atof(v);
    } else if (key == "z") {
      if (!has) return missing("z");
      c.z = std::atof(v);
    } else if (key == "temp") {
      if (!has) return missing("temperature");
      c.temp = std::atof(v);
    } else if (key == "max-translation") {
      if (!has) return missing("max translation");
      c.max_t = std::atof(v);
    } else if (key == "steps") {
      if (!has) return missing("steps");
      c.steps = std::atoi(v);
    } else if (key == "molecules") {
      if (!has) return missing("molecules");
      c.molecules = std::atoi(v);
    } else if (key == "opla.par") {
      if (!has) return missing("opla.par path");
      c.par = val;
    } else if (key == "z-matrix") {
      c.zmatrix = val;  // empty: the run relies on state-input
    } else if (key == "state-input") {
      c.state_in = val;
    } else if (key == "state-output") {
      if (!has) return missing("state-output path");
      c.state_out = val;
    } else if (key == "pdb-output") {
      if (!has) return missing("pdb-output path");
      c.pdb_out = val;
    } else if (key == "cutoff") {
      if (!has) return missing("cutoff");
      c.cutoff = std::atof(v);
    } else if (key == "max-rotation") {
      if (!has) return missing("max rotation");
      c.max_r = std::atof(v);
    } else if (key == "random-seed") {
      c.seed = has ? std::atoi(v) : (int)(unsigned)std::time(nullptr);
    } else if (key == "primary-atom") {
      if (!has) return missing("primary-atom");
      c.primary_line = val;
    } else if (key == "sim-name") {
      if (has) c.name = val;
    } else if (key == "strategy") {
      if (has) c.strategy = val;
    } else if (key == "max-bond-delta" || key == "max-angle-delta") {
      // flexible-molecule move sizes: not used by rigid molecules
    } else if (key == "ensemble") {
      if (has) c.ensemble = val;
    } else if (key == "pressure") {
      if (!has) return missing("pressure");
      c.pressure = std::atof(v);
    } else if (key == "ewald-radius") {
      if (has) c.ewald_radius = std::atoi(v);
    } else if (key == "volume-interval") {
      if (has) c.volume_interval = std::atoi(v);
    } else {
      err = "unexpected key encountered: " + key;
      return false;
    }
  }
  return true;
}

// ------------------------------------------------------------------- .par --

struct OplsAtom {
  double sigma, epsilon, charge;
  std::string name;
};

// Line kinds of an OPLS file: 2 = torsion (Fourier) row "id v0 v1 v2 v3",
// 1 = atom row "id an name charge sigma epsilon", -1 = anything else.  The
// torsion test comes first, as in the reference, so an atom row is one whose
// third field is not a number.
int opls_line_kind(const std::string& line) {
  {
    std::istringstream s(line);
    int id;
    double v0, v1, v2, v3;
    if (s >> id >> v0 >> v1 >> v2 >> v3) return 2;
  }
  std::istringstream s(line);
  int id, an;
  std::string name;
  double q, sg, ep;
  if (s >> id >> an >> name >> q >> sg >> ep) return 1;
  return -1;
}

bool read_opls(const std::string& path, std::map<std::string, OplsAtom>& table,
               std::string& err) {
  std::ifstream in(path);
  if (!in.is_open()) {
    err = "could not open OPLS file (" + path + ")";
    return false;
  }
  std::string line;
  int n = 0;
  while (std::getline(in, line)) {
    n++;
    if (n == 1 || line.empty() || line[0] == '#') continue;  // title row
    if (opls_line_kind(line) != 1) continue;
    std::istringstream s(line);
    std::string key, name;
    int an;
    OplsAtom a;
    s >> key >> an >> name >> a.charge >> a.sigma >> a.epsilon;
    a.name = name;
    table.emplace(key, a);  // the first row of a repeated id wins
  }
  return true;
}

// --------------------------------------------------------------- z-matrix --

struct ZAtom {
  long id;
  double x, y, z;
  double sigma, epsilon, charge;
  std::string name;
};
struct ZLink {
  long a1, a2;   // a1 = the row's own atom id, a2 = the referenced id
  double value;  // bond length (A) or angle (degrees)
};
struct ZMolecule {
  std::vector<ZAtom> atoms;
  std::vector<ZLink> bonds, angles, dihedrals;
};

// 1 = an atom row (11 numeric-shaped fields), 2 = TERZ (next molecule),
// 3 = the "Geometry Variations" block that ends the last molecule, 0 = other.
int zmatrix_line_kind(const std::string& line) {
  std::istringstream s(line);
  int id, oa, ob, bw, aw, dw, extra;
  std::string type;
  double bd, am, dm;
  if (s >> id >> type >> oa >> ob >> bw >> bd >> aw >> am >> dw >> dm >> extra)
    return 1;
  if (line.find("TERZ") != std::string::npos) return 2;
  if (line.find("Geometry Variations") != std::string::npos) return 3;
  return 0;
}

bool read_zmatrix(const std::string& path,
                  const std::map<std::string, OplsAtom>& opls,
                  std::vector<ZMolecule>& mols, std::string& err) {
  std::ifstream in(path);
  if (!in.is_open()) {
    err = "unable to open z-matrix file (" + path + ")";
    return false;
  }
  std::string line;
  int n = 0;
  ZMolecule cur;
  while (std::getline(in, line)) {
    n++;
    if (n == 1 || line.empty() || line[0] == '#') continue;  // title row
    const int kind = zmatrix_line_kind(line);
    if (kind == 2 || kind == 3) {
      mols.push_back(std::move(cur));
      cur = ZMolecule();
      continue;
    }
    if (kind != 1) continue;
    std::istringstream s(line);
    std::string id, type, oa, ob, bw, bd, aw, am, dw, dm;
    s >> id >> type >> oa >> ob >> bw >> bd >> aw >> am >> dw >> dm;
    ZAtom a;
    a.id = std::atoi(id.c_str());
    if (oa != "-1") {
      auto it = opls.find(oa);
      if (it == opls.end()) {
        err = "z-matrix atom " + id + ": OPLS type " + oa + " not in the .par file";
        return false;
      }
      a.sigma = it->second.sigma;
      a.epsilon = it->second.epsilon;
      a.charge = it->second.charge;
      a.name = it->second.name;
      a.x = a.y = a.z = 0.0;
    } else {
      // dummy site: geometry only (sigma = epsilon = -1, no charge)
      a.sigma = a.epsilon = -1.0;
      a.charge = 0.0;
      a.name = type;
      a.x = a.y = a.z = -1.0;
    }
    cur.atoms.push_back(a);
    // A reference field is the literal "0" when absent (string comparison).
    if (bw != "0") cur.bonds.push_back({a.id, std::atoi(bw.c_str()), std::atof(bd.c_str())});
    if (aw != "0") cur.angles.push_back({a.id, std::atoi(aw.c_str()), std::atof(am.c_str())});
    if (dw != "0")
      cur.dihedrals.push_back({a.id, std::atoi(dw.c_str()), std::atof(dm.c_str())});
  }
  return true;
}

// C = A x B, normalised (MathLibrary.cpp:946-962).
void cross_unit(const double* A, const double* B, double* C) {
  const double x = A[1] * B[2] - B[1] * A[2];
  const double y = B[0] * A[2] - A[0] * B[2];
  const double z = A[0] * B[1] - B[0] * A[1];
  double len = (x * x) + (y * y) + (z * z);
  if (len < 1.0e-36) len = 1.0e-36;
  const double r = std::sqrt(len);
  C[0] = x / r;
  C[1] = y / r;
  C[2] = z / r;
}

// First link whose own atom is `id`, searched over every molecule type in
// file order (the reference collects each type's rows into one list before
// searching).
const ZLink* first_link(const std::vector<ZMolecule>& mols,
                        std::vector<ZLink> ZMolecule::*which, long id) {
  for (const ZMolecule& m : mols)
    for (const ZLink& l : m.*which)
      if (l.a1 == id) return &l;
  return nullptr;
}

// Internal coordinates of every molecule type (buildMoleculeXYZ): atom 1 at
// the origin, atom 2 on +x, atom 3 in the xy plane, atoms 4.. from bond,
// angle and dihedral in the local frame of their bond/angle/dihedral
// partners.  Atom ids referenced by a later type are shifted back by the
// sizes of the types before it.  When a type has no bond for its second atom,
// no angle for its third, or fewer than four atoms, the reference stops
// building there and leaves every later type at its read-in coordinates;
// that behaviour is kept.
bool build_internal_coords(std::vector<ZMolecule>& mols, std::string& err) {
  unsigned long offset = 0;
  for (size_t m = 0; m < mols.size(); m++) {
    std::vector<ZAtom>& at = mols[m].atoms;
    const long len = (long)at.size();
    if (len < 1) {
      err = "z-matrix molecule without atoms";
      return false;
    }
    at[0].x = at[0].y = at[0].z = 0.0;
    if (len < 2) return true;
    const ZLink* b1 = first_link(mols, &ZMolecule::bonds, at[1].id);
    if (!b1) return true;
    at[1].x = b1->value;
    at[1].y = 0.0;
    at[1].z = 0.0;
    if (len < 3) return true;
    const ZLink* a2 = first_link(mols, &ZMolecule::angles, at[2].id);
    if (!a2) return true;
    const ZLink* b2 = first_link(mols, &ZMolecule::bonds, at[2].id);
    if (!b2) {
      err = "z-matrix atom " + std::to_string(at[2].id) + " has an angle but no bond";
      return false;
    }
    {
      const unsigned long to = (unsigned long)b2->a2 - offset;
      const double th = a2->value * kPiRef / 180.0;
      if (to == 1) {
        at[2].x = b2->value * std::cos(th);
      } else if (to == 2) {
        at[2].x = at[1].x - b2->value * std::cos(th);
      }
      at[2].y = b2->value * std::sin(th);
      at[2].z = 0.0;
    }
    if (len < 4) return true;
    for (long N = 3; N < len; N++) {
      const long id = at[N].id;
      const ZLink* b = first_link(mols, &ZMolecule::bonds, id);
      const ZLink* a = first_link(mols, &ZMolecule::angles, id);
      const ZLink* d = first_link(mols, &ZMolecule::dihedrals, id);
      if (!b || !a || !d) {
        err = "z-matrix atom " + std::to_string(id) +
              " needs a bond, an angle and a dihedral";
        return false;
      }
      const double th = a->value * kPiRef / 180.0;
      const double ph = d->value * kPiRef / 180.0;
      const double sinval = std::sin(th);
      const double xbs = b->value * sinval * std::cos(ph);
      const double ybs = b->value * sinval * std::sin(ph);
      const double zbs = -b->value * std::cos(th);
      const unsigned long bt = (unsigned long)b->a2 - offset;
      const unsigned long aw = (unsigned long)a->a2 - offset;
      const unsigned long dw = (unsigned long)d->a2 - offset;
      if (bt < 1 || bt > (unsigned long)len || aw < 1 || aw > (unsigned long)len ||
          dw < 1 || dw > (unsigned long)len) {
        err = "z-matrix atom " + std::to_string(id) +
              " refers to an atom outside its molecule";
        return false;
      }
      const double ia[3] = {at[dw - 1].x, at[dw - 1].y, at[dw - 1].z};
      const double ib[3] = {at[aw - 1].x, at[aw - 1].y, at[aw - 1].z};
      const double ic[3] = {at[bt - 1].x, at[bt - 1].y, at[bt - 1].z};
      double vad[3], vab[3], vx[3], vy[3], vz[3];
      for (int k = 0; k < 3; k++) {
        vad[k] = ia[k] - ic[k];
        vab[k] = ib[k] - ic[k];
      }
      cross_unit(vad, vab, vy);
      cross_unit(vab, vy, vx);
      cross_unit(vx, vy, vz);
      at[N].x = vx[0] * xbs + vy[0] * ybs + vz[0] * zbs + ic[0];
      at[N].y = vx[1] * xbs + vy[1] * ybs + vz[1] * zbs + ic[1];
      at[N].z = vx[2] * xbs + vy[2] * ybs + vz[2] * zbs + ic[2];
    }
    offset += (unsigned long)len;
  }
  return true;
}

// FCC start (generatefccBox): molecules 1..3 of the first unit cell are
// offset by half a cell; every molecule i then takes the site coordinates of
// unit-cell molecule a = i mod 4 (read site by site from a's first site on,
// as flat site indexes), shifted by whole cells; everything is finally moved
// back by half a cell.  `n` is the requested molecule count, which the loop
// covers; the box may hold a few more (rounded up to whole type sets).
void place_fcc(HostBox& b, int n) {
  const double dcells = std::pow(0.25 * (double)n, 1.0 / 3.0);
  double cells = (double)(int)(dcells + 0.5);
  while ((4 * cells * cells * cells) < n) cells++;
  const double cellL = b.L[0] / cells;
  const double half = 0.5 * cellL;
  const double off[4][3] = {{0.0, 0.0, 0.0}, {half, half, 0.0},
                            {0.0, half, half}, {half, 0.0, half}};
  const int nmol = b.n_mol();
  for (int a = 0; a < 4 && a < nmol; a++) {
    for (int j = 0; j < b.mol_len[a]; j++) {
      const int s = b.mol_start[a] + j;
      b.x[s] += off[a][0];
      b.y[s] += off[a][1];
      b.z[s] += off[a][2];
    }
  }
  const int nsite = b.n_site();
  int base = 0;
  for (int cz = 1; cz <= cells; cz++) {
    for (int cy = 1; cy <= cells; cy++) {
      for (int cx = 1; cx <= cells; cx++) {
        for (int a = 0; a < 4; a++) {
          const int i = a + base;
          if (i >= n || a >= nmol) continue;
          for (int j = 0; j < b.mol_len[i]; j++) {
            const int d = b.mol_start[i] + j;
            const int s = b.mol_start[a] + j;
            if (s >= nsite) continue;
            b.x[d] = b.x[s] + cellL * (cx - 1);
            b.y[d] = b.y[s] + cellL * (cy - 1);
            b.z[d] = b.z[s] + cellL * (cz - 1);
          }
        }
        base += 4;
      }
    }
  }
  for (int i = 0; i < n && i < nmol; i++) {
    for (int j = 0; j < b.mol_len[i]; j++) {
      const int s = b.mol_start[i] + j;
      b.x[s] -= half;
      b.y[s] -= half;
      b.z[s] -= half;
    }
  }
}

// Box::keepMoleculeInBox: one whole-molecule wrap per dimension, decided by
// the first primary site (below 0: +L, above L: -L).
void keep_in_box(HostBox& b, int m) {
  const int p = b.pidx[b.mol_pidx_start[m]];
  const double pc[3] = {b.x[p], b.y[p], b.z[p]};
  double* c[3] = {b.x.data(), b.y.data(), b.z.data()};
  for (int k = 0; k < 3; k++) {
    const int pos = pc[k] < 0 ? -1 : (pc[k] > b.L[k] ? 1 : 0);
    if (pos == 0) continue;
    for (int j = 0; j < b.mol_len[m]; j++) {
      double& v = c[k][b.mol_start[m] + j];
      if (pos > 0)
        v -= b.L[k];
      else
        v += b.L[k];
    }
  }
}

void copy_str(char* dst, size_t cap, const std::string& s) {
  std::snprintf(dst, cap, "%s", s.c_str());
}

}  // namespace

// Loads a .config run: returns the start box (coordinates as the reference
// builds them) and the run settings.
int config_load(const char* path, HostBox& out, mcgpu_config_info* info,
                std::string& err) {
  RunConfig c;
  if (!read_run_config(path, c, err)) return MCGPU_ERR_INVALID;
  if (c.zmatrix.empty()) {
    err = "configuration file names no z-matrix";
    return MCGPU_ERR_INVALID;
  }
  std::map<std::string, OplsAtom> opls;
  if (!read_opls(c.par, opls, err)) return MCGPU_ERR_INVALID;
  std::vector<ZMolecule> types;
  if (!read_zmatrix(c.zmatrix, opls, types, err)) return MCGPU_ERR_INVALID;
  if (types.empty()) {
    err = "z-matrix holds no complete molecule";
    return MCGPU_ERR_INVALID;
  }
  const std::vector<std::vector<int>> prim = parse_primary_line(c.primary_line);
  if (prim.size() != types.size()) {
    err = "the z-matrix holds " + std::to_string(types.size()) +
          " molecule types but primary-atom defines " + std::to_string(prim.size());
    return MCGPU_ERR_INVALID;
  }
  if (c.molecules < 1) {
    err = "the run needs at least one molecule";
    return MCGPU_ERR_INVALID;
  }
  if (!build_internal_coords(types, err)) return MCGPU_ERR_INVALID;

  const int ntype = (int)types.size();
  const int n = c.molecules;
  const int rounded = (n + ntype - 1) / ntype * ntype;
  HostBox b;
  b.L[0] = c.x;
  b.L[1] = c.y;
  b.L[2] = c.z;
  b.cutoff = c.cutoff;
  b.temperature = c.temp;
  b.kT = c.temp * kBoltz;
  b.max_translate = c.max_t;
  b.max_rotate = c.max_r;
  b.mc_seed = (uint32_t)c.seed;
  b.primary_line = c.primary_line;
  for (int i = 0; i < rounded; i++) {
    const ZMolecule& t = types[i % ntype];
    b.mol_start.push_back(b.n_site());
    b.mol_len.push_back((int32_t)t.atoms.size());
    b.mol_type.push_back(i % ntype);
    for (const ZAtom& a : t.atoms) {
      b.x.push_back(a.x);
      b.y.push_back(a.y);
      b.z.push_back(a.z);
      b.sigma.push_back(a.sigma);
      b.epsilon.push_back(a.epsilon);
      b.charge.push_back(a.charge);
      b.site_name.push_back(a.name);
    }
  }
  place_fcc(b, n);
  // The run covers the requested molecules only.
  const int nsite = b.mol_start[n - 1] + b.mol_len[n - 1];
  b.mol_start.resize(n);
  b.mol_len.resize(n);
  b.mol_type.resize(n);
  for (auto* v : {&b.x, &b.y, &b.z, &b.sigma, &b.epsilon, &b.charge}) v->resize(nsite);
  b.site_name.resize(nsite);
  for (int m = 0; m < n; m++) {
    const std::vector<int>& p = prim[b.mol_type[m]];
    if (p.empty()) {
      err = "molecule type without primary index";
      return MCGPU_ERR_INVALID;
    }
    b.mol_pidx_start.push_back((int32_t)b.pidx.size());
    b.mol_pidx_count.push_back((int32_t)p.size());
    for (int k : p) {
      if (k < 0 || k >= b.mol_len[m]) {
        err = "primary-atom index outside its molecule type";
        return MCGPU_ERR_INVALID;
      }
      b.pidx.push_back(b.mol_start[m] + k);
    }
  }
  for (int m = 0; m < n; m++) keep_in_box(b, m);

  if (info) {
    std::memset(info, 0, sizeof(*info));
    info->steps = c.steps;
    info->n_molecules = n;
    info->n_types = ntype;
    info->random_seed = c.seed;
    info->pressure = c.pressure;
    info->volume_interval = c.volume_interval;
    info->ewald_radius = c.ewald_radius;
    copy_str(info->sim_name, sizeof(info->sim_name), c.name);
    copy_str(info->strategy, sizeof(info->strategy), c.strategy);
    copy_str(info->ensemble, sizeof(info->ensemble), c.ensemble);
    copy_str(info->state_output, sizeof(info->state_output), c.state_out);
    copy_str(info->pdb_output, sizeof(info->pdb_output), c.pdb_out);
  }
  out = std::move(b);
  return MCGPU_OK;
}

}  // namespace mcgpu

extern "C" int mcgpu_config_read(const char* path, mcgpu_hostbox** out,
                                 mcgpu_config_info* info) {
  if (!path || !out) return MCGPU_ERR_INVALID;
  *out = nullptr;
  mcgpu::HostBox b;
  std::string err;
  const int rc = mcgpu::config_load(path, b, info, err);
  if (rc != MCGPU_OK) {
    mcgpu::set_error("mcgpu_config_read: " + err);
    return rc;
  }
  const std::string bad = mcgpu::validate_desc(b.desc());
  if (!bad.empty()) {
    mcgpu::set_error("mcgpu_config_read: " + bad);
    return MCGPU_ERR_INVALID;
  }
  mcgpu_hostbox* h = mcgpu::hostbox_new(std::move(b));
  if (!h) return MCGPU_ERR_NOMEM;
  *out = h;
  return MCGPU_OK;
}
