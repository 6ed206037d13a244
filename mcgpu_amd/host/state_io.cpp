// .state checkpoint files in the reference's layout
// (StateScanner.cpp:495-597 writes, :18-195/:235-281/:391-491 read):
//   line 1  "x y z nMol nAtoms T cutoff maxT maxR primaryLine seed"
//   line 2  step number, line 3 blank, then per molecule:
//   id / "= Type" / type / "= Atoms" / "id x y z sigma eps q name" ...
//   / "= Bonds" / "= Dihedrals" / "= Hops" / "= Angles" / "=="
// Numbers are written with %.17g (the reference writes 6 significant digits,
// which makes its resume lossy), so a reload reproduces every double.
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "hostbox.h"
#include "mcgpu_host.h"

namespace mcgpu {
HostBox& hostbox_ref(mcgpu_hostbox* h);
mcgpu_hostbox* hostbox_new(HostBox&& b);

// First token of the resume-record line of a .state file.
constexpr const char* kResumeTag = "#mcgpu-resume";

// Parses a primary definition line such as "1", "2,1", "[1,2],3"
// (StateScanner.cpp:437-491, 1-based) into per-type 0-based index lists.
std::vector<std::vector<int>> parse_primary_line(const std::string& s) {
  std::vector<std::vector<int>> out;
  size_t i = 0;
  while (i < s.size()) {
    if (s[i] == ',') {
      i++;
      continue;
    }
    std::vector<int> v;
    if (s[i] == '[') {
      size_t j = s.find(']', i);
      if (j == std::string::npos) return {};
      std::stringstream ss(s.substr(i + 1, j - i - 1));
      std::string tok;
      while (std::getline(ss, tok, ',')) v.push_back(std::atoi(tok.c_str()) - 1);
      i = j + 1;
    } else {
      size_t j = s.find(',', i);
      if (j == std::string::npos) j = s.size();
      v.push_back(std::atoi(s.substr(i, j - i).c_str()) - 1);
      i = j;
    }
    out.push_back(v);
  }
  return out;
}

}  // namespace mcgpu

using mcgpu::HostBox;
using mcgpu::kResumeTag;

extern "C" int mcgpu_state_write(mcgpu_hostbox* h, const char* path,
                                 int64_t step) {
  if (!h || !path) return MCGPU_ERR_INVALID;
  const HostBox& b = mcgpu::hostbox_ref(h);
  FILE* f = std::fopen(path, "w");
  if (!f) {
    mcgpu::set_error(std::string("mcgpu_state_write: cannot open ") + path);
    return MCGPU_ERR_INVALID;
  }
  std::fprintf(f, "%.17g %.17g %.17g %d %d %.17g %.17g %.17g %.17g %s %u \n",
               b.L[0], b.L[1], b.L[2], b.n_mol(), b.n_site(), b.temperature,
               b.cutoff, b.max_translate, b.max_rotate, b.primary_line.c_str(),
               b.mc_seed);
  std::fprintf(f, "%lld\n\n", (long long)step);
  int atom_id = 0;
  for (int m = 0; m < b.n_mol(); m++) {
    std::fprintf(f, "%d\n= Type\n%d\n= Atoms\n", m,
                 b.mol_type.empty() ? 0 : b.mol_type[m]);
    for (int k = 0; k < b.mol_len[m]; k++) {
      const int s = b.mol_start[m] + k;
      const char* nm = b.site_name.empty() ? "X" : b.site_name[s].c_str();
      std::fprintf(f, "%d %.17g %.17g %.17g %.17g %.17g %.17g %s \n", atom_id++,
                   b.x[s], b.y[s], b.z[s], b.sigma[s], b.epsilon[s], b.charge[s],
                   nm);
    }
    std::fprintf(f, "= Bonds\n= Dihedrals\n= Hops\n= Angles\n==\n");
  }
  if (b.has_resume) {
    // Resume record after the last molecule, on one line that opens with
    // '#': the reference's reader (StateScanner.cpp:54-195) takes a line
    // that does not open with "= " or "==" in its molecule-id state as an
    // id and keeps no molecule from it, so it loads such a file as it loads
    // the file without this line.
    const mcgpu_resume_state& r = b.resume;
    std::fprintf(f, "%s %lld %lld %lld %lld %d %.17g %.17g %.17g %d %d %.17g %d %d", kResumeTag,
                 (long long)r.steps_done, (long long)r.accepted, (long long)r.volume_moves,
                 (long long)r.volume_accepted, r.volume_pending, r.energy, r.lrc,
                 r.volume_step, r.ensemble, r.volume_interval, r.pressure, r.rng_front,
                 r.rng_rear);
    for (int k = 0; k < 31; k++) std::fprintf(f, " %d", r.rng_words[k]);
    std::fprintf(f, "\n");
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok) {
    mcgpu::set_error(std::string("mcgpu_state_write: write failed on ") + path);
    return MCGPU_ERR_INVALID;
  }
  return MCGPU_OK;
}

extern "C" int mcgpu_state_read(const char* path, mcgpu_hostbox** out,
                                int64_t* step) {
  if (!path || !out) return MCGPU_ERR_INVALID;
  std::ifstream in(path);
  if (!in) {
    mcgpu::set_error(std::string("mcgpu_state_read: cannot open ") + path);
    return MCGPU_ERR_INVALID;
  }
  HostBox b;
  std::string line;
  std::getline(in, line);
  {
    std::istringstream ss(line);
    int nmol = 0, nat = 0;
    long long seed = 0;
    if (!(ss >> b.L[0] >> b.L[1] >> b.L[2] >> nmol >> nat >> b.temperature >>
          b.cutoff >> b.max_translate >> b.max_rotate >> b.primary_line >> seed)) {
      mcgpu::set_error("mcgpu_state_read: bad environment line");
      return MCGPU_ERR_INVALID;
    }
    b.mc_seed = (uint32_t)seed;
    b.kT = b.temperature * mcgpu::kBoltz;
  }
  std::getline(in, line);
  if (step) *step = std::atoll(line.c_str());
  std::getline(in, line);  // blank
  const auto prim = mcgpu::parse_primary_line(b.primary_line);
  if (prim.empty()) {
    mcgpu::set_error("mcgpu_state_read: bad primary index definition");
    return MCGPU_ERR_INVALID;
  }
  int section = 0, type = 0;
  int32_t start = 0;
  while (std::getline(in, line)) {
    const std::string hold = line.substr(0, 2);
    if (section == 0) {  // molecule id
      if (hold == "= ") {
        section = 1;
      } else if (line.compare(0, std::strlen(kResumeTag), kResumeTag) == 0) {
        mcgpu_resume_state& r = b.resume;
        std::istringstream ss(line.substr(std::strlen(kResumeTag)));
        long long sd, ac, vm, va;
        ss >> sd >> ac >> vm >> va >> r.volume_pending >> r.energy >> r.lrc >>
            r.volume_step >> r.ensemble >> r.volume_interval >> r.pressure >> r.rng_front >>
            r.rng_rear;
        for (int k = 0; k < 31; k++) ss >> r.rng_words[k];
        if (!ss) {
          mcgpu::set_error("mcgpu_state_read: bad resume record");
          return MCGPU_ERR_INVALID;
        }
        r.steps_done = sd;
        r.accepted = ac;
        r.volume_moves = vm;
        r.volume_accepted = va;
        b.has_resume = true;
      } else if (line.find_first_not_of(" \t\r") == std::string::npos) {
        break;
      }
      continue;
    }
    if (section == 1) {  // type
      if (hold == "= ") {
        section = 2;
        start = b.n_site();
      } else {
        type = std::atoi(line.c_str());
      }
      continue;
    }
    if (section == 2) {  // atoms
      if (hold == "= ") {
        section = 3;
        continue;
      }
      std::istringstream ss(line);
      long id;
      double x, y, z, s, e, q;
      std::string nm = "X";
      if (!(ss >> id >> x >> y >> z >> s >> e >> q)) {
        mcgpu::set_error("mcgpu_state_read: bad atom line");
        return MCGPU_ERR_INVALID;
      }
      ss >> nm;
      b.x.push_back(x);
      b.y.push_back(y);
      b.z.push_back(z);
      b.sigma.push_back(s);
      b.epsilon.push_back(e);
      b.charge.push_back(q);
      b.site_name.push_back(nm);
      continue;
    }
    // bonds / dihedrals / hops / angles: skipped (rigid molecules)
    if (hold == "==") {
      if (type < 0 || type >= (int)prim.size()) {
        mcgpu::set_error("mcgpu_state_read: molecule type without primary index");
        return MCGPU_ERR_INVALID;
      }
      const int32_t len = b.n_site() - start;
      b.mol_start.push_back(start);
      b.mol_len.push_back(len);
      b.mol_type.push_back(type);
      b.mol_pidx_start.push_back((int32_t)b.pidx.size());
      b.mol_pidx_count.push_back((int32_t)prim[type].size());
      for (int p : prim[type]) {
        if (p < 0 || p >= len) {
          mcgpu::set_error("mcgpu_state_read: primary index outside molecule");
          return MCGPU_ERR_INVALID;
        }
        b.pidx.push_back(start + p);
      }
      section = 0;
    }
  }
  if (b.n_mol() < 1) {
    mcgpu::set_error("mcgpu_state_read: no molecules");
    return MCGPU_ERR_INVALID;
  }
  mcgpu_hostbox* h = mcgpu::hostbox_new(std::move(b));
  if (!h) return MCGPU_ERR_NOMEM;
  *out = h;
  return MCGPU_OK;
}

extern "C" int mcgpu_hostbox_set_resume(mcgpu_hostbox* h, const mcgpu_resume_state* r) {
  if (!h) return MCGPU_ERR_INVALID;
  HostBox& b = mcgpu::hostbox_ref(h);
  b.has_resume = r != nullptr;
  b.resume = r ? *r : mcgpu_resume_state{};
  return MCGPU_OK;
}

extern "C" int mcgpu_hostbox_get_resume(mcgpu_hostbox* h, mcgpu_resume_state* r,
                                        int32_t* present) {
  if (!h || !r || !present) return MCGPU_ERR_INVALID;
  const HostBox& b = mcgpu::hostbox_ref(h);
  *present = b.has_resume ? 1 : 0;
  *r = b.resume;
  return MCGPU_OK;
}

// PDB snapshot in the reference's column layout (Simulation.cpp:362-417):
// one ATOM row per site ("ATOM", site number, name, "UNK", molecule number,
// x y z with 3 decimals), TER after each molecule, END at the close.
extern "C" int mcgpu_pdb_write(mcgpu_hostbox* h, const char* path) {
  if (!h || !path) return MCGPU_ERR_INVALID;
  const HostBox& b = mcgpu::hostbox_ref(h);
  FILE* f = std::fopen(path, "w");
  if (!f) {
    mcgpu::set_error(std::string("mcgpu_pdb_write: cannot open ") + path);
    return MCGPU_ERR_INVALID;
  }
  std::fprintf(f, "REMARK Created by MCGPU\n");
  for (int m = 0; m < b.n_mol(); m++) {
    for (int k = 0; k < b.mol_len[m]; k++) {
      const int s = b.mol_start[m] + k;
      const char* nm = b.site_name.empty() ? "X" : b.site_name[s].c_str();
      std::fprintf(f, "%-6s%5d%3s%6s%6d%12.3f%8.3f%8.3f\n", "ATOM", s + 1, nm,
                   "UNK", m + 1, b.x[s], b.y[s], b.z[s]);
    }
    std::fprintf(f, "TER\n");
  }
  std::fprintf(f, "END\n");
  const bool ok = std::fclose(f) == 0;
  if (!ok) {
    mcgpu::set_error(std::string("mcgpu_pdb_write: write failed on ") + path);
    return MCGPU_ERR_INVALID;
  }
  return MCGPU_OK;
}
