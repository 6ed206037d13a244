// Work-group level evaluation of one molecule's intermolecular energy at its
// current and (optionally) proposed position, in one pass over candidates.
//
// Preparation (wave 0): the neighbour source -- every molecule, the occupied
// slots of the cells spanning the old and proposed cell (their occupancies
// prefix-summed by a DPP wave scan, with an item -> cell map in LDS), or the
// Verlet list.
// Candidate scan (all lanes): each lane takes one candidate j, tests the
// primary-site cutoff at the old and proposed position (SimulationStep.cpp:
// 317-330), and the hits are compacted into LDS in candidate order (ballot +
// mbcnt, one barrier per pass).
// Pair phase: lanes are grouped Aw = next_pow2(m_i) at a time; a group takes
// one (hit, site b) item and each lane of it one site a of the moving
// molecule, so every (a, b) pair of every hit is one lane-step
// (BruteForceStep.cpp:48-79).  Sums are FP64; per-wave DPP sums and a
// fixed-order sum over waves end the evaluation, so results are
// deterministic for a given work-group size.
#pragma once

#include "device_common.h"
#include "wave.h"

namespace mcgpu {

// Marks a staged site (in EvalShared::sp) of a hit whose site pairs may
// wrap differently: its pairs take wrap1 per axis.
constexpr int kPairWrap = 1 << 20;
// A staged site's hit is in range of the moving molecule's old (kPairOld)
// or proposed (kPairNew) position: its pairs enter e_old or e_new.
constexpr int kPairOld = 1 << 21;
constexpr int kPairNew = 1 << 22;

template <int NT>
struct EvalShared {
  static constexpr int kCap = NT * 2 < 256 ? 256 : NT * 2;
  static constexpr int kItemCap = 2048;  // items with a cell entry in item_cell
  static constexpr int kCellLds = 1024;  // cells whose counts a launch keeps in LDS
  // staged hit sites; 1024-lane groups keep their LDS under 64 KiB, which
  // LDS instructions address from one base with an immediate offset
  static constexpr int kSiteCap = NT >= 1024 ? 768 : (NT >= 512 ? 1536 : (NT >= 256 ? 768 : 192));
  static constexpr int kNW = NT / 64;
  // What the pair loop reads comes first, within the 64 KiB that LDS
  // instructions reach with an immediate offset from one base.
  // Sites of a chunk of hits, staged for the pair sums: coordinates and the
  // site-parameter entry (-1: no such site in that hit); coordinates are
  // shifted by the periodic image of their pairs with the moving molecule,
  // except for hits whose site pairs may wrap differently (kPairWrap in sp
  // marks those); sp also carries the hit's in-range flags (kPairOld,
  // kPairNew).
  double sx[kSiteCap], sy[kSiteCap], sz[kSiteCap];
  int32_t sp[kSiteCap];
  // the box's site-parameter table
  SiteParam64 prm64[kParamCap];
  SiteParam32 prm32[kParamCap];
  // in-range molecules of the current evaluation (hits), compacted:
  // site count | flags << 8 (bit0 old, bit1 new) | image code << 10 (see
  // image_code; the same for the old and the proposed position, else
  // kImageMixed)
  int32_t cand_lf[kCap];
  double ox[kMaxSites], oy[kMaxSites], oz[kMaxSites];
  double nx[kMaxSites], ny[kMaxSites], nz[kMaxSites];
  SiteParam64 q64[kMaxSites];
  SiteParam32 q32[kMaxSites];
  int32_t pofs[kMaxSites];   // primary offsets within the molecule
  int32_t cand_start[kCap];  // first site
  int32_t cand_param[kCap];  // first site-parameter entry
  int32_t cell_id[64];       // cells of the current range (at most 4x4x4)
  int32_t cell_pre[65];      // prefix of their occupancies
  uint8_t item_cell[kItemCap];  // range cell of each candidate item
  // copy of the box's cell occupancies, kept by launches that hold the box
  // for many moves (the batch kernel); prepare_cells_wave reads it when asked
  int32_t ccount[kCellLds];
  double red[kNW];
  double red2[kNW];
  unsigned long long redc[kNW];  // sites << 32 | candidates, per wave
  unsigned redp[kNW];            // site pairs at the old position, per wave
  int32_t scan[2][kNW];          // hit counts per wave, double-buffered
  int32_t npri;
  int32_t len;
  int32_t aw_log2;  // log2 of the lane-group width of the pair phase
  int32_t flag;
  int32_t fault;    // MCGPU_FAULT_* raised by the last preparation (0 = none)
  int32_t src;      // neighbour source of the current evaluation
  int32_t nitems;   // candidate items it holds
  int32_t ncells;   // cells of the range (cell source)
};

__device__ __forceinline__ int lane_group_log2(int len) {
  return len > 1 ? 32 - __clz(len - 1) : 0;
}

// Copies the box's site-parameter table into LDS (all threads; the caller
// places a barrier before it is read).  Tables are at most kParamCap entries
// (checked when a box is created).
template <int NT>
__device__ __forceinline__ void load_params(const BoxView& bv, EvalShared<NT>& sh) {
  for (int k = threadIdx.x; k < bv.n_param; k += NT) {
    sh.prm64[k] = bv.p64[k];
    sh.prm32[k] = bv.p32[k];
  }
}

// Loads molecule `mol`'s current sites (and parameters) into LDS.
template <int NT>
__device__ __forceinline__ void load_molecule(const BoxView& bv, int mol,
                                              EvalShared<NT>& sh) {
  const int t = threadIdx.x;
  const int start = bv.mol_start[mol];
  int len = bv.mol_len[mol];
  len = len < 1 ? 1 : (len > kMaxSites ? kMaxSites : len);
  const int pbase = bv.mol_param[mol];
  for (int k = t; k < len; k += NT) {
    sh.ox[k] = bv.x[start + k];
    sh.oy[k] = bv.y[start + k];
    sh.oz[k] = bv.z[start + k];
    sh.q64[k] = bv.p64[pbase + k];
    sh.q32[k] = bv.p32[pbase + k];
  }
  if (t == 0) {
    const int ps = bv.mol_pidx_start[mol];
    int pc = bv.mol_pidx_count[mol];
    pc = pc < 1 ? 1 : (pc > kMaxSites ? kMaxSites : pc);
    sh.fault = (bv.mol_len[mol] != len) ? MCGPU_FAULT_MOLLEN : 0;
    for (int k = 0; k < pc; k++) sh.pofs[k] = bv.pidx[ps + k] - start;
    sh.npri = pc;
    sh.len = len;
    sh.aw_log2 = lane_group_log2(len);
  }
}

// Any primary pair of (moving molecule at X/Y/Z, molecule with primaries
// ps .. ps+pc-1 in the mirror) within the cutoff (SimulationStep.cpp:317-330).
__device__ __forceinline__ bool primary_in_range(const BoxView& bv, int ps, int pc,
                                                 const double* X,
                                                 const double* Y,
                                                 const double* Z,
                                                 const int32_t* pofs,
                                                 int npri) {
  bool in = false;
  for (int q = 0; q < pc; q++) {
    const double qx = bv.px[ps + q], qy = bv.py[ps + q], qz = bv.pz[ps + q];
    for (int p = 0; p < npri; p++) {
      const int o = pofs[p];
      const double dx = wrap1(qx - X[o], bv.L[0]);
      const double dy = wrap1(qy - Y[o], bv.L[1]);
      const double dz = wrap1(qz - Z[o], bv.L[2]);
      in |= (dx * dx + dy * dy + dz * dz) <= bv.cutoff2;
    }
  }
  return in;
}

// floor(u / m) for u < 65536, 1 <= m <= 32, with magic = 2^32 / m + 1.
__device__ __forceinline__ int div_small(int u, unsigned long long magic) {
  return (int)(((unsigned long long)(unsigned)u * magic) >> 32);
}

// Sums the pair energies of the compacted hits into e_old / e_new, a chunk
// of hits at a time: all lanes first stage the chunk's sites from the box
// into LDS (one round trip to memory), then lanes are grouped Aw at a time,
// a group takes one (hit, site b) item and each lane of it one site a of the
// moving molecule.
template <int NT, bool MIXED>
__device__ __forceinline__ void pair_phase(const BoxView& bv, EvalShared<NT>& sh,
                                           int nhits, double& e_old, double& e_new,
                                           unsigned& npairs) {
  constexpr int CAP = EvalShared<NT>::kSiteCap;
  const int M = bv.max_len;
  const unsigned long long magic = (1ull << 32) / (unsigned)M + 1ull;
  const int len = sh.len;
  const int lg = sh.aw_log2;
  const int a = threadIdx.x & ((1 << lg) - 1);
  const int ac = a < len ? a : 0;
  const int G = NT >> lg;
  const bool lane_real = a < len && sh.q64[ac].real != 0.0;
  // this lane's site a: old and proposed position, parameters
  const double xo = sh.ox[ac], yo = sh.oy[ac], zo = sh.oz[ac];
  const double xn = sh.nx[ac], yn = sh.ny[ac], zn = sh.nz[ac];
  const double sa = sh.q64[ac].s, ea = sh.q64[ac].e, qa = sh.q64[ac].q;
  const float sa32 = sh.q32[ac].s, ea32 = sh.q32[ac].e;
  const double L0 = bv.L[0], L1 = bv.L[1], L2 = bv.L[2];
  const int chunk = CAP / M;
  for (int h0 = 0; h0 < nhits; h0 += chunk) {
    const int nh = nhits - h0 < chunk ? nhits - h0 : chunk;
    const int tot = nh * M;
    for (int u = threadIdx.x; u < tot; u += NT) {
      const int kk = div_small(u, magic);
      const int b = u - kk * M;
      const int k = h0 + kk;
      const int lf = sh.cand_lf[k];
      int pi = -1;
      if (b < (lf & 0xff)) {
        const int site = sh.cand_start[k] + b;
        const int img = (lf >> 10) & 0x7f;
        const bool mixed = img & kImageMixed;
        sh.sx[u] = bv.x[site] + (mixed ? 0.0 : image_shift(img & 3, L0));
        sh.sy[u] = bv.y[site] + (mixed ? 0.0 : image_shift((img >> 2) & 3, L1));
        sh.sz[u] = bv.z[site] + (mixed ? 0.0 : image_shift((img >> 4) & 3, L2));
        pi = (sh.cand_param[k] + b) | (mixed ? kPairWrap : 0) | ((lf >> 8) & 3) * kPairOld;
      }
      sh.sp[u] = pi;
    }
    __syncthreads();
    // two items per step for ILP; 1024-lane groups hide latency with
    // their 16 waves instead and need the registers
    constexpr int U = NT >= 1024 ? 1 : 2;
    for (int i0 = threadIdx.x >> lg; i0 < tot; i0 += U * G) {
      int pi[U];
      double xb[U], yb[U], zb[U];
#pragma unroll
      for (int v = 0; v < U; v++) {
        const int i = i0 + v * G;
        pi[v] = -1;
        if (i < tot) {
          pi[v] = sh.sp[i];
          xb[v] = sh.sx[i];
          yb[v] = sh.sy[i];
          zb[v] = sh.sz[i];
        }
        if (!lane_real) pi[v] = -1;
      }
#pragma unroll
      for (int v = 0; v < U; v++) {
        if (pi[v] < 0) continue;
        const int pk = pi[v] & (kPairWrap - 1);
        // displacements: the staged site carries its hit's periodic image,
        // or wrap1 per site pair when the hit's site pairs may wrap
        // differently
        double dxo = xb[v] - xo, dyo = yb[v] - yo, dzo = zb[v] - zo;
        double dxn = xb[v] - xn, dyn = yb[v] - yn, dzn = zb[v] - zn;
        if (pi[v] & kPairWrap) {
          dxo = wrap1(dxo, L0);
          dyo = wrap1(dyo, L1);
          dzo = wrap1(dzo, L2);
          dxn = wrap1(dxn, L0);
          dyn = wrap1(dyn, L1);
          dzn = wrap1(dzn, L2);
        }
        if (MIXED) {
          const SiteParam32 pb = sh.prm32[pk];
          const SiteParam64& pd = sh.prm64[pk];
          if (pi[v] & kPairOld) {
            e_old += pair_energy32(dxo, dyo, dzo, sa32, ea32, pb.s, pb.e, sa, ea, qa, pd.s,
                                   pd.e, pd.q);
            npairs++;
          }
          if (pi[v] & kPairNew)
            e_new += pair_energy32(dxn, dyn, dzn, sa32, ea32, pb.s, pb.e, sa, ea, qa, pd.s,
                                   pd.e, pd.q);
        } else {
          const SiteParam64 pb = sh.prm64[pk];
          if (pi[v] & kPairOld) {
            e_old += pair_energy64(dxo, dyo, dzo, sa, ea, qa, pb.s, pb.e, pb.q);
            npairs++;
          }
          if (pi[v] & kPairNew)
            e_new += pair_energy64(dxn, dyn, dzn, sa, ea, qa, pb.s, pb.e, pb.q);
        }
      }
    }
    __syncthreads();
  }
}

// Neighbour source actually usable for this evaluation (block-uniform).
__device__ __forceinline__ int pick_source(const BoxView& bv, const double* nx,
                                           const double* ny, const double* nz,
                                           int p0, int mol, bool has_new) {
  if (bv.mode == MCGPU_NB_CELL && bv.cells.overflow[0] == 0) return MCGPU_NB_CELL;
  if (bv.mode == MCGPU_NB_VERLET && bv.verlet.stale[0] == 0 &&
      bv.verlet.count[mol] <= bv.verlet.cap) {
    const double* r = bv.verlet.ref + 3 * mol;
    if (has_new) {
      const double dx = wrap1(nx[p0] - r[0], bv.L[0]);
      const double dy = wrap1(ny[p0] - r[1], bv.L[1]);
      const double dz = wrap1(nz[p0] - r[2], bv.L[2]);
      if (dx * dx + dy * dy + dz * dz > bv.verlet.skin_half2) return MCGPU_NB_BRUTE;
    }
    return MCGPU_NB_VERLET;
  }
  return MCGPU_NB_BRUTE;
}

// Wave 0 (all 64 lanes): the cells around old cell co and proposed cell cn,
// their occupancies (from sh.ccount when lds_counts), the prefix and, when
// fill, the item -> cell map (else the caller fills it with
// fill_item_cells).  Returns the cell count, or -1 when the range exceeds 64
// cells (the caller then scans every molecule); *nitems receives the
// candidate count.
template <int NT>
__device__ __forceinline__ int prepare_cells_wave(const BoxView& bv,
                                                  EvalShared<NT>& sh,
                                                  const int co[3], const int cn[3],
                                                  int* nitems, bool lds_counts = false,
                                                  bool fill = true) {
  const CellGrid& g = bv.cells;
  const CellRange range = make_range(g, co, cn);
  const int ncells = range.total();
  if (ncells > 64) return -1;
  const int lane = wlane();
  int n = 0, c = 0;
  if (lane < ncells) {
    c = range.cell(g, lane);
    n = lds_counts ? sh.ccount[c] : g.count[c];
    n = n < g.cap ? n : g.cap;
    sh.cell_id[lane] = c;
  }
  const int incl = wave_incl_scan_i32(n);
  sh.cell_pre[lane + 1] = incl;
  if (lane == 0) sh.cell_pre[0] = 0;
  if (fill) {
    const int end = incl < EvalShared<NT>::kItemCap ? incl : EvalShared<NT>::kItemCap;
    for (int k = incl - n; k < end; k++) sh.item_cell[k] = (uint8_t)lane;
  }
  *nitems = readlane_i32(incl, 63);
  return ncells;
}

// All NT threads: the item -> cell map of the ncells cells whose prefix
// prepare_cells_wave left in sh.cell_pre, NT / 64 threads per cell.  The
// caller places a barrier before and after it.
template <int NT>
__device__ __forceinline__ void fill_item_cells(EvalShared<NT>& sh, int ncells) {
  constexpr int kPer = NT / 64;
  const int c = threadIdx.x / kPer;
  if (c < ncells) {
    const int lim = EvalShared<NT>::kItemCap;
    int end = sh.cell_pre[c + 1];
    end = end < lim ? end : lim;
    for (int k = sh.cell_pre[c] + (int)threadIdx.x % kPer; k < end; k += kPer)
      sh.item_cell[k] = (uint8_t)c;
  }
}

// The range cell holding candidate `item` of the cell source.
template <int NT>
__device__ __forceinline__ int item_cell_of(const EvalShared<NT>& sh, int item,
                                            int ncells) {
  if (item < EvalShared<NT>::kItemCap) return sh.item_cell[item];
  int lo = 0, hi = ncells;  // last cell whose prefix <= item
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (sh.cell_pre[mid] <= item) lo = mid; else hi = mid;
  }
  return lo;
}

// Work counters of one evaluation (block-uniform): candidate molecules
// examined and sites of in-range molecules loaded; they give the kernel's
// algorithmic bytes (24 B per candidate primary, 24 B per site).
struct EvalWork {
  unsigned cand;
  unsigned sites;
};

// Candidate scan + pair phase over the prepared source (sh.src / nitems /
// ncells and the cell tables), for molecule `mol` at the LDS-held old
// position and, when has_new, the proposed one.  Called by all NT threads
// after a barrier that follows the preparation.  Leaves per-wave partial
// sums in sh.red / red2 / redc / redp and ends with a barrier.  The
// candidate and site counts are per-wave ballot counts; the site-pair count
// (redp) is summed only when count_pairs.
template <int NT, bool MIXED>
__device__ __forceinline__ void eval_core(const BoxView& bv, EvalShared<NT>& sh,
                                          int mol, int jlo, bool has_new,
                                          bool count_pairs = true) {
  const int t = threadIdx.x;
  const int src = __builtin_amdgcn_readfirstlane(sh.src);
  const int nitems = __builtin_amdgcn_readfirstlane(sh.nitems);
  const int ncells = __builtin_amdgcn_readfirstlane(sh.ncells);
  const int npri = __builtin_amdgcn_readfirstlane(sh.npri);
  const int uni_len = bv.uni_len;
  double eo = 0.0, en = 0.0;
  unsigned npairs = 0;
  unsigned ncand_w = 0, nsites_w = 0;  // this wave's candidates and hit sites
  int nhits = 0, par = 0;
  for (int base = 0; base < nitems; base += NT) {
    const int item = base + t;
    int j = -1;
    if (item < nitems) {
      if (src == MCGPU_NB_CELL) {
        const int lo = item_cell_of(sh, item, ncells);
        j = bv.cells.slots[sh.cell_id[lo] * bv.cells.cap + (item - sh.cell_pre[lo])];
      } else if (src == MCGPU_NB_VERLET) {
        j = bv.verlet.list[mol * bv.verlet.cap + item];
      } else {
        j = jlo + item;
      }
    }
    int f = 0;
    MolMeta mj{0, 0, 0, 0, 0};
    const bool cand = j >= jlo && j != mol;
    ncand_w += (unsigned)__popcll(__ballot(cand));
    if (cand) {
      if (uni_len > 0) {  // uniform box: the record follows from the index
        mj.start = j * uni_len;
        mj.len = (int16_t)uni_len;
        mj.param = bv.uni_param;
        mj.pstart = j;
      } else {
        mj = bv.meta[j];
      }
      int ps = j, pc = 1;  // one primary per molecule: no indirection
      if (!bv.single_primary) {
        ps = mj.pstart;
        pc = bv.mol_pidx_count[j];
      }
      f = primary_in_range(bv, ps, pc, sh.ox, sh.oy, sh.oz, sh.pofs, npri) ? 1 : 0;
      if (has_new)
        f |= primary_in_range(bv, ps, pc, sh.nx, sh.ny, sh.nz, sh.pofs, npri) ? 2 : 0;
      if (f) {
        // periodic image of the pair, from the first primaries
        const int o = sh.pofs[0];
        const double qx = bv.px[ps], qy = bv.py[ps], qz = bv.pz[ps];
        int img = image_code(qx - sh.ox[o], qy - sh.oy[o], qz - sh.oz[o], bv.L, bv.reach2);
        if (has_new &&
            image_code(qx - sh.nx[o], qy - sh.ny[o], qz - sh.nz[o], bv.L, bv.reach2) != img)
          img = kImageMixed;
        f |= img << 2;
      }
    }
    // compaction: rank within the wave, wave counts through LDS
    const unsigned long long bal = __ballot(f != 0);
    if (uni_len > 0)
      nsites_w += (unsigned)__popcll(bal) * (unsigned)uni_len;
    else
      nsites_w += wave_sum_u32(f ? (unsigned)mj.len : 0u);
    if (wlane() == 0) sh.scan[par][wave_id()] = __popcll(bal);
    __syncthreads();
    int pre = 0, total = 0;
#pragma unroll
    for (int w = 0; w < NT / 64; w++) {
      const int c = sh.scan[par][w];
      pre += (w < wave_id()) ? c : 0;
      total += c;
    }
    par ^= 1;
    if (nhits + total > EvalShared<NT>::kCap) {  // block-uniform
      // (ends with a barrier: every lane is done with the hits before reuse)
      pair_phase<NT, MIXED>(bv, sh, nhits, eo, en, npairs);
      nhits = 0;
    }
    if (f) {
      const int pos = nhits + pre + wave_rank(bal);
      sh.cand_start[pos] = mj.start;
      sh.cand_param[pos] = mj.param;
      sh.cand_lf[pos] = mj.len | (f << 8);
    }
    nhits += total;
  }
  __syncthreads();
  if (nhits > 0) pair_phase<NT, MIXED>(bv, sh, nhits, eo, en, npairs);
  eo = wave_sum_f64(eo);
  en = wave_sum_f64(en);
  const unsigned wpairs = count_pairs ? wave_sum_u32(npairs) : 0u;
  if (wlane() == 0) {
    sh.red[wave_id()] = eo;
    sh.red2[wave_id()] = en;
    sh.redc[wave_id()] = ((unsigned long long)nsites_w << 32) | ncand_w;
    sh.redp[wave_id()] = wpairs;
  }
  __syncthreads();
}

// Fixed-order totals of the per-wave partials left by eval_core.
template <int NT>
__device__ __forceinline__ void eval_totals(const EvalShared<NT>& sh, double& e_old,
                                            double& e_new, unsigned long long& counts,
                                            unsigned& pairs) {
  double so = 0.0, sn = 0.0;
  unsigned long long c = 0;
  unsigned p = 0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) {
    so += sh.red[w];
    sn += sh.red2[w];
    c += sh.redc[w];
    p += sh.redp[w];
  }
  e_old = so;
  e_new = sn;
  counts = c;
  pairs = p;
}

// Evaluates E_mol over candidates j >= jlo (j != mol) at the LDS-held old
// position, and at the proposed position when has_new.  Must be called by
// all NT threads after load_molecule (+ proposed coords) and a barrier.
// Results are block-uniform.
template <int NT, bool MIXED>
__device__ __forceinline__ void eval_molecule(const BoxView& bv, int mol, int jlo,
                                              bool has_new, EvalShared<NT>& sh,
                                              double& e_old, double& e_new,
                                              unsigned long long* pair_count,
                                              EvalWork* work = nullptr) {
  if (threadIdx.x < 64) {  // wave 0 prepares the neighbour source
    const int p0 = sh.pofs[0];
    int src = pick_source(bv, sh.nx, sh.ny, sh.nz, p0, mol, has_new);
    int nitems = 0, ncells = 0;
    if (src == MCGPU_NB_CELL) {
      const CellGrid& g = bv.cells;
      int co[3], cn[3];
      co[0] = cell_coord(sh.ox[p0], g.inv_w[0], g.n[0]);
      co[1] = cell_coord(sh.oy[p0], g.inv_w[1], g.n[1]);
      co[2] = cell_coord(sh.oz[p0], g.inv_w[2], g.n[2]);
      if (has_new) {
        cn[0] = cell_coord(sh.nx[p0], g.inv_w[0], g.n[0]);
        cn[1] = cell_coord(sh.ny[p0], g.inv_w[1], g.n[1]);
        cn[2] = cell_coord(sh.nz[p0], g.inv_w[2], g.n[2]);
      } else {
        cn[0] = co[0];
        cn[1] = co[1];
        cn[2] = co[2];
      }
      ncells = prepare_cells_wave(bv, sh, co, cn, &nitems);
      if (ncells < 0) {
        src = MCGPU_NB_BRUTE;
        ncells = 0;
      }
    }
    if (src == MCGPU_NB_VERLET) nitems = bv.verlet.count[mol];
    if (src == MCGPU_NB_BRUTE) nitems = bv.n_mol - jlo;
    // A neighbour source never holds more candidates than molecules; if it
    // claims so its lists are corrupt: record it and evaluate nothing.
    int fault = 0;
    if (nitems < 0 || nitems > bv.n_mol) {
      fault = MCGPU_FAULT_CANDIDATES;
      nitems = 0;
    }
    if (threadIdx.x == 0) {
      sh.src = src;
      sh.nitems = nitems;
      sh.ncells = ncells;
      if (fault) sh.fault = fault;
    }
  }
  __syncthreads();
  eval_core<NT, MIXED>(bv, sh, mol, jlo, has_new);
  unsigned long long c;
  unsigned p;
  eval_totals<NT>(sh, e_old, e_new, c, p);
  if (!has_new) e_new = 0.0;
  if (work) {
    work->cand = (unsigned)(c & 0xffffffffu);
    work->sites = (unsigned)(c >> 32);
  }
  if (pair_count) *pair_count = p;
}

// Moves molecule `mol` from cell cold to cell cnew in the cell lists (one
// lane): swap-remove from the old cell, append to the new one; a full cell
// marks the lists incomplete (scans then fall back to every molecule), and
// the call returns false.
__device__ __forceinline__ bool cell_move(const CellGrid& g, int mol, int cold,
                                          int cnew) {
  const int slot = g.mol_slot[mol];
  const int last = g.count[cold] - 1;
  const int lm = g.slots[cold * g.cap + last];
  g.slots[cold * g.cap + slot] = lm;
  g.mol_slot[lm] = slot;
  g.count[cold] = last;
  const int n = g.count[cnew];
  if (n < g.cap) {
    g.slots[cnew * g.cap + n] = mol;
    g.mol_slot[mol] = n;
    g.count[cnew] = n + 1;
    g.mol_cell[mol] = cnew;
    return true;
  }
  g.overflow[0] = 1;
  return false;
}

// Writes the LDS-held proposed coordinates of `mol` into the box and keeps
// the primary mirror, cell lists and Verlet bookkeeping consistent.
template <int NT>
__device__ void commit_molecule(const BoxView& bv, int mol,
                                const EvalShared<NT>& sh) {
  const int t = threadIdx.x;
  const int start = bv.mol_start[mol];
  for (int k = t; k < sh.len; k += NT) {
    bv.x[start + k] = sh.nx[k];
    bv.y[start + k] = sh.ny[k];
    bv.z[start + k] = sh.nz[k];
  }
  if (t == 0) {
    const int ps = bv.mol_pidx_start[mol];
    for (int k = 0; k < sh.npri; k++) {
      bv.px[ps + k] = sh.nx[sh.pofs[k]];
      bv.py[ps + k] = sh.ny[sh.pofs[k]];
      bv.pz[ps + k] = sh.nz[sh.pofs[k]];
    }
    const int p0 = sh.pofs[0];
    if (bv.mode == MCGPU_NB_CELL && bv.cells.overflow[0] == 0) {
      const CellGrid& g = bv.cells;
      const int cnew = cell_index(g, cell_coord(sh.nx[p0], g.inv_w[0], g.n[0]),
                                  cell_coord(sh.ny[p0], g.inv_w[1], g.n[1]),
                                  cell_coord(sh.nz[p0], g.inv_w[2], g.n[2]));
      const int cold = g.mol_cell[mol];
      if (cnew != cold) cell_move(g, mol, cold, cnew);
    }
    if (bv.mode == MCGPU_NB_VERLET) {
      const double* r = bv.verlet.ref + 3 * mol;
      const double dx = wrap1(sh.nx[p0] - r[0], bv.L[0]);
      const double dy = wrap1(sh.ny[p0] - r[1], bv.L[1]);
      const double dz = wrap1(sh.nz[p0] - r[2], bv.L[2]);
      if (dx * dx + dy * dy + dz * dz > bv.verlet.skin_half2) bv.verlet.stale[0] = 1;
    }
  }
}

}  // namespace mcgpu
