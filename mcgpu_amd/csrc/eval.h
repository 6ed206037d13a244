// Work-group level evaluation of one molecule's intermolecular energy at its
// current and (optionally) proposed position, in one pass over candidates.
//
// Phase 1 (candidate scan): lanes take candidate molecules j from the
// neighbour source -- every molecule, the occupied slots of the cells around
// the old and proposed cells (enumerated through a prefix over the cell
// counts, so no lane looks at an empty slot), or the Verlet list -- test the
// primary-site cutoff at the old and proposed position (SimulationStep.cpp:
// 317-330), and the hits are compacted into LDS in candidate order by a
// ballot/popcount block scan.
// Phase 2 (pair sum): lanes are grouped Aw = next_pow2(m_i) at a time; a
// group takes one (hit, site b) item and each lane of it one site a of the
// moving molecule, so every (a, b) pair of every hit is one lane-step
// (BruteForceStep.cpp:48-79).  Sums are FP64; one fixed-order block
// reduction ends the evaluation, so results are deterministic for a given
// work-group size.
#pragma once

#include "device_common.h"

namespace mcgpu {

template <int NT>
struct EvalShared {
  static constexpr int kCap = NT * 2 < 256 ? 256 : NT * 2;
  double ox[kMaxSites], oy[kMaxSites], oz[kMaxSites];
  double nx[kMaxSites], ny[kMaxSites], nz[kMaxSites];
  SiteParam64 q64[kMaxSites];
  SiteParam32 q32[kMaxSites];
  int32_t pofs[kMaxSites];  // primary offsets within the molecule
  // in-range molecules of the current evaluation (hits), compacted:
  int32_t cand_start[kCap];  // first site
  int32_t cand_param[kCap];  // first site-parameter entry
  int32_t cand_lf[kCap];     // site count | flags << 8 (bit0 old, bit1 new)
  int32_t cell_id[64];   // cells of the current range (at most 4x4x4)
  int32_t cell_pre[65];  // prefix of their occupancies
  double red[NT / 64];
  double red2[NT / 64];
  unsigned long long redc[NT / 64];
  int32_t scan[NT / 64];
  int32_t npri;
  int32_t len;
  int32_t nreal;
  int32_t aw_log2;  // log2 of the lane-group width of the pair phase
  int32_t flag;
  int32_t fault;    // MCGPU_FAULT_* raised by the last evaluation (0 = none)
};

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
    int nr = 0;
    for (int k = 0; k < len; k++) nr += bv.p64[pbase + k].real != 0.0;
    sh.nreal = nr;
    int lg = 0;
    while ((1 << lg) < len) lg++;
    sh.aw_log2 = lg;
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

// Sums the pair energies of the compacted hits into e_old / e_new.  Each
// lane handles two (hit, site b) items per iteration so the site loads of
// both are in flight together.
template <int NT, bool MIXED>
__device__ __forceinline__ void pair_phase(const BoxView& bv, EvalShared<NT>& sh,
                                           int nhits, double& e_old,
                                           double& e_new,
                                           unsigned long long& npairs,
                                           unsigned& nsites) {
  const int M = bv.max_len;
  const int len = sh.len;
  const int lg = sh.aw_log2;
  const int a = threadIdx.x & ((1 << lg) - 1);
  const int G = NT >> lg;
  const bool lane_real = a < len && sh.q64[a < len ? a : 0].real != 0.0;
  const int total = nhits * M;
  for (int i0 = threadIdx.x >> lg; i0 < total; i0 += 2 * G) {
    int site[2], pi[2], f[2];
    bool ok[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
      const int item = i0 + u * G;
      ok[u] = false;
      site[u] = pi[u] = f[u] = 0;
      if (item < total) {
        const int k = item / M;
        const int b = item - k * M;
        const int lf = sh.cand_lf[k];
        if (b < (lf & 0xff)) {
          ok[u] = true;
          site[u] = sh.cand_start[k] + b;
          pi[u] = sh.cand_param[k] + b;
          f[u] = lf >> 8;
        }
      }
      nsites += (ok[u] && a == 0);
      ok[u] = ok[u] && lane_real;
    }
    double xb[2], yb[2], zb[2];
    if (MIXED) {
      SiteParam32 pb[2];
#pragma unroll
      for (int u = 0; u < 2; u++)
        if (ok[u]) {
          xb[u] = bv.x[site[u]];
          yb[u] = bv.y[site[u]];
          zb[u] = bv.z[site[u]];
          pb[u] = bv.p32[pi[u]];
        }
#pragma unroll
      for (int u = 0; u < 2; u++) {
        if (!ok[u] || pb[u].real == 0.0f) continue;
        if (f[u] & 1) {
          e_old += pair_energy32(wrap1(xb[u] - sh.ox[a], bv.L[0]),
                                 wrap1(yb[u] - sh.oy[a], bv.L[1]),
                                 wrap1(zb[u] - sh.oz[a], bv.L[2]), sh.q32[a], pb[u],
                                 sh.q64[a], bv.p64 + pi[u]);
          npairs++;
        }
        if (f[u] & 2)
          e_new += pair_energy32(wrap1(xb[u] - sh.nx[a], bv.L[0]),
                                 wrap1(yb[u] - sh.ny[a], bv.L[1]),
                                 wrap1(zb[u] - sh.nz[a], bv.L[2]), sh.q32[a], pb[u],
                                 sh.q64[a], bv.p64 + pi[u]);
      }
    } else {
      SiteParam64 pb[2];
#pragma unroll
      for (int u = 0; u < 2; u++)
        if (ok[u]) {
          xb[u] = bv.x[site[u]];
          yb[u] = bv.y[site[u]];
          zb[u] = bv.z[site[u]];
          pb[u] = bv.p64[pi[u]];
        }
#pragma unroll
      for (int u = 0; u < 2; u++) {
        if (!ok[u] || pb[u].real == 0.0) continue;
        if (f[u] & 1) {
          e_old += pair_energy64(wrap1(xb[u] - sh.ox[a], bv.L[0]),
                                 wrap1(yb[u] - sh.oy[a], bv.L[1]),
                                 wrap1(zb[u] - sh.oz[a], bv.L[2]), sh.q64[a], pb[u]);
          npairs++;
        }
        if (f[u] & 2)
          e_new += pair_energy64(wrap1(xb[u] - sh.nx[a], bv.L[0]),
                                 wrap1(yb[u] - sh.ny[a], bv.L[1]),
                                 wrap1(zb[u] - sh.nz[a], bv.L[2]), sh.q64[a], pb[u]);
      }
    }
  }
  __syncthreads();
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

// Work counters of one evaluation (block-uniform): candidate molecules
// examined and sites of in-range molecules loaded; they give the kernel's
// algorithmic bytes (24 B per candidate primary, 24 B per site).
struct EvalWork {
  unsigned cand;
  unsigned sites;
};

// Evaluates E_mol over candidates j >= jlo (j != mol) at the LDS-held old
// position, and at the proposed position when has_new.  Must be called by
// all NT threads after load_molecule (+ proposed coords) and a barrier.
// Results are block-uniform.
template <int NT, bool MIXED>
__device__ void eval_molecule(const BoxView& bv, int mol, int jlo,
                              bool has_new, EvalShared<NT>& sh, double& e_old,
                              double& e_new, unsigned long long* pair_count,
                              EvalWork* work = nullptr) {
  const int t = threadIdx.x;
  const int p0 = sh.pofs[0];
  int src = pick_source(bv, sh.nx, sh.ny, sh.nz, p0, mol, has_new);
  int nitems, ncells = 0;
  CellRange range{};
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
    range = make_range(g, co, cn);
    ncells = range.total();
    if (ncells > 64) src = MCGPU_NB_BRUTE;  // see make_range (uniform)
  }
  if (src == MCGPU_NB_CELL) {
    const CellGrid& g = bv.cells;
    // ncells <= 64: one wave builds the prefix of the cell occupancies
    if (t < 64) {
      int n = 0;
      if (t < ncells) {
        const int c = range.cell(g, t);
        n = g.count[c];
        n = n < g.cap ? n : g.cap;
        sh.cell_id[t] = c;
      }
#pragma unroll
      for (int o = 1; o < 64; o <<= 1) {
        const int u = __shfl_up(n, o);
        if (t >= o) n += u;
      }
      sh.cell_pre[t + 1] = n;
      if (t == 0) sh.cell_pre[0] = 0;
    }
    __syncthreads();
    nitems = sh.cell_pre[ncells];
  } else if (src == MCGPU_NB_VERLET) {
    nitems = bv.verlet.count[mol];
  } else {
    nitems = bv.n_mol - jlo;
  }
  // A neighbour source never holds more candidates than molecules; if it
  // claims so its lists are corrupt: record it and evaluate nothing.
  if (nitems < 0 || nitems > bv.n_mol) {
    if (t == 0) sh.fault = MCGPU_FAULT_CANDIDATES;
    nitems = 0;
  }

  double eo = 0.0, en = 0.0;
  unsigned long long np = 0;
  unsigned ncand = 0, nsites = 0;
  int nhits = 0;
  for (int base = 0; base < nitems; base += NT) {
    const int item = base + t;
    int j = -1;
    if (item < nitems) {
      if (src == MCGPU_NB_CELL) {
        int lo = 0, hi = ncells;  // last cell whose prefix <= item
        while (hi - lo > 1) {
          const int mid = (lo + hi) >> 1;
          if (sh.cell_pre[mid] <= item) lo = mid; else hi = mid;
        }
        j = bv.cells.slots[sh.cell_id[lo] * bv.cells.cap + (item - sh.cell_pre[lo])];
      } else if (src == MCGPU_NB_VERLET) {
        j = bv.verlet.list[mol * bv.verlet.cap + item];
      } else {
        j = jlo + item;
      }
    }
    int f = 0;
    MolMeta mj{0, 0, 0, 0};
    if (j >= jlo && j != mol) {
      ncand++;
      mj = bv.meta[j];
      int ps = j, pc = 1;  // one primary per molecule: no indirection
      if (!bv.single_primary) {
        ps = mj.pstart;
        pc = bv.mol_pidx_count[j];
      }
      f = primary_in_range(bv, ps, pc, sh.ox, sh.oy, sh.oz, sh.pofs, sh.npri) ? 1 : 0;
      if (has_new)
        f |= primary_in_range(bv, ps, pc, sh.nx, sh.ny, sh.nz, sh.pofs, sh.npri) ? 2 : 0;
    }
    int total;
    const int pos = block_scan_flag<NT>(f != 0, sh.scan, total);
    if (nhits + total > EvalShared<NT>::kCap) {
      pair_phase<NT, MIXED>(bv, sh, nhits, eo, en, np, nsites);
      nhits = 0;
    }
    if (f) {
      sh.cand_start[nhits + pos] = mj.start;
      sh.cand_param[nhits + pos] = mj.param;
      sh.cand_lf[nhits + pos] = mj.len | (f << 8);
    }
    nhits += total;
    __syncthreads();
  }
  if (nhits > 0) pair_phase<NT, MIXED>(bv, sh, nhits, eo, en, np, nsites);
  // one fixed-order reduction of both energies and the work counters
  unsigned long long wc = ((unsigned long long)nsites << 32) | ncand;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    eo += __shfl_xor(eo, o);
    en += __shfl_xor(en, o);
    wc += __shfl_xor(wc, o);
  }
  if (lane_id() == 0) {
    sh.red[wave_id()] = eo;
    sh.red2[wave_id()] = en;
    sh.redc[wave_id()] = wc;
  }
  __syncthreads();
  const double so = wave_sum_partials<NT / 64>(sh.red);
  const double sn = wave_sum_partials<NT / 64>(sh.red2);
  const unsigned long long sc = wave_sum_partials<NT / 64>(sh.redc);
  __syncthreads();
  e_old = so;
  e_new = has_new ? sn : 0.0;
  if (work) {
    work->cand = (unsigned)(sc & 0xffffffffu);
    work->sites = (unsigned)(sc >> 32);
  }
  if (pair_count) {
    const double c = block_sum<NT>((double)np, sh.red);
    if (t == 0) *pair_count = (unsigned long long)c;
  }
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
      if (cnew != cold) {
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
        } else {
          g.overflow[0] = 1;  // lists incomplete from here: scans fall back
        }
      }
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
