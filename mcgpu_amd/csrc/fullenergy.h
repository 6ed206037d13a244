// Full-system energy on the device, a throughput kernel body shared by the
// single-box and batched paths.
#pragma once

#include "cells.h"
#include "device_common.h"
#include "eval.h"
#include "wave.h"

namespace mcgpu {

// Full-system energy (SimulationStep.cpp:40-71: every molecule pair within
// the primary cutoff of SimulationStep.cpp:317-330 once), one wave per
// molecule i and no barriers between waves:
//  * the lanes scan i's candidate partners -- with cell lists, i's own cell
//    (partners j > i) and the 13 neighbour cells on one side of it (any j),
//    so that every pair of cells is visited from one end only, or, with
//    fewer than 3 cells along an axis, every cell around i (j > i); without
//    cell lists, all j > i -- and compact the hits' ids into the wave's LDS
//    list by ballot;
//  * one lane per hit reads the partner's record and classifies the pair's
//    periodic image;
//  * the hits' real sites are staged densely in LDS (coordinates already
//    carrying the pair's periodic image), and each lane takes one staged
//    site b against every real site a of i, whose coordinates and
//    parameters the wave holds in LDS.
// partial[i] receives i's share; the shares are reduced in a fixed order
// elsewhere, so the total is deterministic.  Which molecule of a pair holds
// its terms differs from the reference's E_i(i) split (j > i); the set of
// pairs, and so the total up to summation order, is the same.
constexpr int kWaveHits = 192;   // hits buffered before their pairs are summed
constexpr int kWaveItems = 128;  // staged sites b per pass of the pair loop
constexpr int kScanItems = 1024; // scan items mapped per segment (cell lists)
constexpr int kScanCells = 32;   // cells of one scan (14 half-shell, <= 27 full)
// Boxes whose largest molecule has at least this many sites split the pairs
// by index parity over all 27 cells instead of the half shell: a molecule's
// share of the pair work then varies far less (per-SIMD maximum over mean
// 1.17 against 1.63 on a 4,000-molecule 16-site box), which pays for
// scanning twice the cells when each hit carries ~100+ site pairs.
constexpr int kParityLen = 8;
static_assert(kParamCap <= 256 && kMaxSites < 256, "hit_info packs both in 8 bits");

struct WaveEval {
  // real sites of molecule i (compacted): coordinates, FP64 and FP32 pair
  // parameters
  double ax[kMaxSites], ay[kMaxSites], az[kMaxSites];
  SiteParam64 ap[kMaxSites];
  SiteParam32 ap32[kMaxSites];
  int32_t hit_start[kWaveHits];  // partner id until wave_resolve_hits
  // parameter entry | site count << 8 | periodic image (image_code) << 16
  int32_t hit_info[kWaveHits];
  uint16_t hit_first[kWaveHits];  // first item of the hit's sites in the run
  // staged real sites b of the hits: coordinates and parameter entry, with
  // kPairWrap set when the hit's site pairs may wrap differently
  double bx[kWaveItems], by[kWaveItems], bz[kWaveItems];
  int32_t bp[kWaveItems];
  // the scan's non-empty cells in order: cell id (bit 31 set when partners
  // must exceed i) and first scan item; a bitmap over the current segment
  // marks the first item of every cell after the first
  int32_t scan_cell[kScanCells];
  int32_t scan_first[kScanCells];
  unsigned long long scan_marks[kScanItems / 64];
};

// Pairs of the staged sites b [0, n) with the na real sites of i.
template <bool MIXED>
__device__ __forceinline__ double wave_items_energy(const BoxView& bv, const WaveEval& W,
                                                    int na, int n, const SiteParam64* P64,
                                                    const SiteParam32* P32) {
  const double L0 = bv.L[0], L1 = bv.L[1], L2 = bv.L[2];
  double e = 0.0;
  for (int it = wlane(); it < n; it += 64) {
    const double xb = W.bx[it], yb = W.by[it], zb = W.bz[it];
    const int bp = W.bp[it];
    const int pb = bp & (kPairWrap - 1);
    const SiteParam64 B = P64[pb];
    const SiteParam32 B32 = P32[pb];
    // two site pairs per step: independent chains for the FP64 latency
    if (bp & kPairWrap) {
#pragma unroll 2
      for (int a = 0; a < na; a++) {
        const double dx = wrap1(xb - W.ax[a], L0);
        const double dy = wrap1(yb - W.ay[a], L1);
        const double dz = wrap1(zb - W.az[a], L2);
        const SiteParam64& A = W.ap[a];
        if (MIXED) {
          const SiteParam32& A32 = W.ap32[a];
          e += pair_energy32(dx, dy, dz, A32.s, A32.e, B32.s, B32.e, A.s, A.e, A.q, B.s, B.e,
                             B.q);
        } else {
          e += pair_energy64(dx, dy, dz, A.s, A.e, A.q, B.s, B.e, B.q);
        }
      }
    } else {
#pragma unroll 2
      for (int a = 0; a < na; a++) {
        const double dx = xb - W.ax[a], dy = yb - W.ay[a], dz = zb - W.az[a];
        const SiteParam64& A = W.ap[a];
        if (MIXED) {
          const SiteParam32& A32 = W.ap32[a];
          e += pair_energy32(dx, dy, dz, A32.s, A32.e, B32.s, B32.e, A.s, A.e, A.q, B.s, B.e,
                             B.q);
        } else {
          e += pair_energy64(dx, dy, dz, A.s, A.e, A.q, B.s, B.e, B.q);
        }
      }
    }
  }
  return e;
}

// The scan leaves partner ids in hit_start[0, nh); one lane per hit turns
// each into its record (first site, site count, parameter entry) and the
// pair's periodic image, classified from the two first primaries (p0 is
// i's).  Doing this once per hit rather than once per scanned candidate
// keeps the scan loop to loads and the cutoff test.  The hits' sites are
// numbered in one run (hit_first: prefix of the site counts); returns the
// run's length.
__device__ __forceinline__ int wave_resolve_hits(const BoxView& bv, WaveEval& W, int nh,
                                                 double p0x, double p0y, double p0z) {
  const int lane = wlane();
  int first = 0;
  for (int h0 = 0; h0 < nh; h0 += 64) {  // wave-uniform
    const int h = h0 + lane;
    int len = 0;
    if (h < nh) {
      const int j = W.hit_start[h];
      const int psj = bv.single_primary ? j : bv.mol_pidx_start[j];
      const int img = image_code(bv.px[psj] - p0x, bv.py[psj] - p0y, bv.pz[psj] - p0z,
                                 bv.L, bv.reach2);
      int param;
      if (bv.uni_len > 0) {
        W.hit_start[h] = j * bv.uni_len;
        len = bv.uni_len;
        param = bv.uni_param;
      } else {
        const MolMeta m = bv.meta[j];
        W.hit_start[h] = m.start;
        len = m.len;
        param = m.param;
      }
      W.hit_info[h] = param | len << 8 | img << 16;
    }
    const int incl = wave_incl_scan_i32(len);
    if (h < nh) W.hit_first[h] = (uint16_t)(first + incl - len);
    first += readlane_i32(incl, 63);
  }
  __builtin_amdgcn_wave_barrier();
  return first;
}

// Sums the pairs of hits [0, nh) with i.  The run of the hits' sites is
// staged 64 sites per pass: lane t takes site t of the run, the hit holding
// it found by counting, with one popcount, the marks of a bitmap of the
// hits' first sites (segments of kScanItems sites; it reuses the scan's
// bitmap, which the scan rebuilds afterwards).  Real sites are compacted by
// ballot; whenever the buffer could overflow, the pair loop takes the whole
// 64-site passes and the remainder moves to the front.  npairs counts the
// site pairs.
template <bool MIXED>
__device__ __forceinline__ double wave_hits_energy(const BoxView& bv, WaveEval& W, int na,
                                                   int nh, const SiteParam64* P64,
                                                   const SiteParam32* P32, unsigned& npairs,
                                                   double p0x, double p0y, double p0z) {
  const int nsite = wave_resolve_hits(bv, W, nh, p0x, p0y, p0z);
  const double L0 = bv.L[0], L1 = bv.L[1], L2 = bv.L[2];
  const int lane = wlane();
  double e = 0.0;
  int n = 0, seg = 0, seg_end = 0, before = 0;
  for (int base = 0; base < nsite; base += 64) {
    if (n + 64 > kWaveItems) {  // wave-uniform
      const int full = n & ~63;
      __builtin_amdgcn_wave_barrier();
      e += wave_items_energy<MIXED>(bv, W, na, full, P64, P32);
      npairs += (unsigned)(full * na);
      __builtin_amdgcn_wave_barrier();
      const int rest = n - full;
      if (lane < rest) {
        W.bx[lane] = W.bx[full + lane];
        W.by[lane] = W.by[full + lane];
        W.bz[lane] = W.bz[full + lane];
        W.bp[lane] = W.bp[full + lane];
      }
      __builtin_amdgcn_wave_barrier();
      n = rest;
    }
    if (base >= seg_end) {  // wave-uniform
      seg = base;
      seg_end = base + kScanItems;
      if (lane < kScanItems / 64) W.scan_marks[lane] = 0ull;
      __builtin_amdgcn_wave_barrier();
      before = 0;
      for (int h0 = 0; h0 < nh; h0 += 64) {
        const int h = h0 + lane;
        bool pre = false;
        if (h >= 1 && h < nh) {
          const int f = W.hit_first[h];
          pre = f < seg;
          if (!pre && f - seg < kScanItems)
            atomicOr(&W.scan_marks[(f - seg) >> 6], 1ull << ((f - seg) & 63));
        }
        before += __popcll(__ballot(pre));
      }
      __builtin_amdgcn_wave_barrier();
    }
    const unsigned long long word = W.scan_marks[(base - seg) >> 6];
    // marks at run positions base..base+lane (2 << 63 wraps to 0: all bits)
    const int k = before + __popcll(word & ((2ull << lane) - 1ull));
    before += __popcll(word);
    const int item = base + lane;
    bool ok = false;
    int pb = 0, site = 0, img = 0;
    if (item < nsite) {
      const int info = W.hit_info[k];
      const int b = item - W.hit_first[k];
      pb = (info & 0xff) + b;
      ok = P64[pb].real != 0.0;
      site = W.hit_start[k] + b;
      img = info >> 16;
    }
    const unsigned long long bal = __ballot(ok);
    if (ok) {
      const int pos = n + wave_rank(bal);
      const bool mixed = img & kImageMixed;
      W.bx[pos] = bv.x[site] + (mixed ? 0.0 : image_shift(img & 3, L0));
      W.by[pos] = bv.y[site] + (mixed ? 0.0 : image_shift((img >> 2) & 3, L1));
      W.bz[pos] = bv.z[site] + (mixed ? 0.0 : image_shift((img >> 4) & 3, L2));
      W.bp[pos] = pb | (mixed ? kPairWrap : 0);
    }
    n += __popcll(bal);
  }
  __builtin_amdgcn_wave_barrier();
  if (n > 0) {
    e += wave_items_energy<MIXED>(bv, W, na, n, P64, P32);
    npairs += (unsigned)(n * na);
  }
  __builtin_amdgcn_wave_barrier();
  return e;
}

// Returns molecule i's share of the full energy (wave-uniform); pairs
// receives the real site pairs summed.
template <bool MIXED>
__device__ __forceinline__ double wave_molecule_energy(const BoxView& bv, WaveEval& W, int i,
                                                       const SiteParam64* P64,
                                                       const SiteParam32* P32,
                                                       unsigned& pairs) {
  unsigned np = 0;
  const int lane = wlane();
  int start, len, param, psi, pci;
  if (bv.uni_len > 0) {
    len = bv.uni_len;
    start = i * len;
    param = bv.uni_param;
    psi = i;
    pci = 1;
  } else {
    const MolMeta m = bv.meta[i];
    start = m.start;
    len = m.len;
    param = m.param;
    psi = bv.single_primary ? i : m.pstart;
    pci = bv.single_primary ? 1 : bv.mol_pidx_count[i];
  }
  // i's real sites, compacted
  int na;
  {
    const bool real = lane < len && P64[param + lane].real != 0.0;
    const unsigned long long bal = __ballot(real);
    if (real) {
      const int pos = wave_rank(bal);
      W.ax[pos] = bv.x[start + lane];
      W.ay[pos] = bv.y[start + lane];
      W.az[pos] = bv.z[start + lane];
      W.ap[pos] = P64[param + lane];
      W.ap32[pos] = P32[param + lane];
    }
    na = __popcll(bal);
    __builtin_amdgcn_wave_barrier();
  }
  // the primaries of i (at most kMaxSites) as lane values
  double pix = 0.0, piy = 0.0, piz = 0.0;
  if (lane < pci) {
    pix = bv.px[psi + lane];
    piy = bv.py[psi + lane];
    piz = bv.pz[psi + lane];
  }
  const double p0x = readlane_f64(pix, 0), p0y = readlane_f64(piy, 0),
               p0z = readlane_f64(piz, 0);
  const double L0 = bv.L[0], L1 = bv.L[1], L2 = bv.L[2];
  // candidate source: lane c < ncell holds cell c of the scan and its
  // occupancy prefix; cell 0 is i's own (partners j > i)
  const CellGrid& g = bv.cells;
  const bool cells = bv.mode == MCGPU_NB_CELL && g.overflow[0] == 0;
  int ncell = 0, nfilled = 0, nitems = bv.n_mol - (i + 1), own_full = 0;
  int my_cell = 0;
  bool parity = false;
  if (cells) {
    int co[3];
    co[0] = cell_coord(p0x, g.inv_w[0], g.n[0]);
    co[1] = cell_coord(p0y, g.inv_w[1], g.n[1]);
    co[2] = cell_coord(p0z, g.inv_w[2], g.n[2]);
    parity = g.n[0] >= 3 && g.n[1] >= 3 && g.n[2] >= 3 && bv.max_len >= kParityLen;
    if (parity) {
      // all 27 cells around i; i holds the pair with j when i + j is even
      // and j > i, or odd and j < i (every pair once, from either end)
      ncell = 27;
      if (lane < 27) {
        int a = co[0] + lane / 9 - 1, b = co[1] + (lane / 3) % 3 - 1, c = co[2] + lane % 3 - 1;
        a += a < 0 ? g.n[0] : 0;
        a -= a >= g.n[0] ? g.n[0] : 0;
        b += b < 0 ? g.n[1] : 0;
        b -= b >= g.n[1] ? g.n[1] : 0;
        c += c < 0 ? g.n[2] : 0;
        c -= c >= g.n[2] ? g.n[2] : 0;
        my_cell = cell_index(g, a, b, c);
      }
    } else if (g.n[0] >= 3 && g.n[1] >= 3 && g.n[2] >= 3) {
      // own cell + the 13 offsets (dx, dy, dz) > (0, 0, 0) lexicographically
      ncell = 14;
      if (lane < 14) {
        int d0 = 0, d1 = 0, d2 = 0;
        if (lane > 0) {
          const int t = lane + 13;  // 14..26 -> offsets above the centre
          d0 = t / 9 - 1;
          d1 = (t / 3) % 3 - 1;
          d2 = t % 3 - 1;
        }
        int a = co[0] + d0, b = co[1] + d1, c = co[2] + d2;
        a += a < 0 ? g.n[0] : 0;
        a -= a >= g.n[0] ? g.n[0] : 0;
        b += b < 0 ? g.n[1] : 0;
        b -= b >= g.n[1] ? g.n[1] : 0;
        c += c < 0 ? g.n[2] : 0;
        c -= c >= g.n[2] ? g.n[2] : 0;
        my_cell = cell_index(g, a, b, c);
      }
    } else {
      // fewer than 3 cells along some axis: every cell around i, j > i
      const CellRange r = make_range(g, co, co);
      ncell = r.total();
      own_full = 1;
      if (lane < ncell) my_cell = r.cell(g, lane);
    }
    int cnt = 0;
    if (lane < ncell) {
      cnt = g.count[my_cell];
      cnt = cnt < g.cap ? cnt : g.cap;
    }
    const int incl = wave_incl_scan_i32(cnt);
    nitems = readlane_i32(incl, 63);
    const bool filled = lane < ncell && cnt > 0;
    const unsigned long long fb = __ballot(filled);
    nfilled = __popcll(fb);
    if (filled) {
      const int r = wave_rank(fb);
      W.scan_cell[r] = my_cell | ((!parity && (own_full || lane == 0)) ? (int)0x80000000 : 0);
      W.scan_first[r] = incl - cnt;
    }
    __builtin_amdgcn_wave_barrier();
  }
  // Scan item -> candidate.  Without cells, j = i + 1 + item.  With cells,
  // the items of the non-empty cells follow one another, so item t lies in
  // cell r = the number of cells after the first whose first item is <= t:
  // those first items are marked in a bitmap over a segment of kScanItems
  // items, and each lane counts the marks up to its own item.
  auto mark_segment = [&](int seg) -> int {
    if (lane < kScanItems / 64) W.scan_marks[lane] = 0ull;
    __builtin_amdgcn_wave_barrier();
    const int f = lane >= 1 && lane < nfilled ? W.scan_first[lane] : 0x7fffffff;
    if (f >= seg && f - seg < kScanItems)
      atomicOr(&W.scan_marks[(f - seg) >> 6], 1ull << ((f - seg) & 63));
    __builtin_amdgcn_wave_barrier();
    return __popcll(__ballot(f < seg));  // cells starting before the segment
  };
  double e = 0.0;
  int nh = 0;
  int base = 0, seg = 0, seg_end = 0, before = 0;
  do {
    // windows of 64 items while the hit buffer has room for a whole window
    for (; base < nitems && nh <= kWaveHits - 64; base += 64) {
      const int item = base + lane;
      int j = -1;
      bool greater = false;
      if (cells) {
        if (base >= seg_end) {  // wave-uniform
          seg = base;
          seg_end = base + kScanItems;
          before = mark_segment(seg);
        }
        const unsigned long long word = W.scan_marks[(base - seg) >> 6];
        // marks at window positions 0..lane (2 << 63 wraps to 0: all bits)
        const int r = before + __popcll(word & ((2ull << lane) - 1ull));
        before += __popcll(word);
        const int c = W.scan_cell[r];
        greater = c < 0;
        if (item < nitems) j = g.slots[(c & 0x7fffffff) * g.cap + (item - W.scan_first[r])];
      } else if (item < nitems) {
        j = i + 1 + item;
      }
      const bool ok = j >= 0 && j != i &&
                      (parity ? (((i ^ j) & 1) == 0) == (j > i) : (!greater || j > i));
      bool hit = false;
      if (bv.single_primary) {
        const int jj = ok ? j : i;  // every lane loads a valid entry
        const double dx = wrap1(bv.px[jj] - p0x, L0);
        const double dy = wrap1(bv.py[jj] - p0y, L1);
        const double dz = wrap1(bv.pz[jj] - p0z, L2);
        hit = ok && dx * dx + dy * dy + dz * dz <= bv.cutoff2;
      } else if (ok) {
        const int psj = bv.mol_pidx_start[j];
        const int pcj = bv.mol_pidx_count[j];
        for (int q = 0; q < pcj; q++) {
          const double qx = bv.px[psj + q], qy = bv.py[psj + q], qz = bv.pz[psj + q];
          for (int p = 0; p < pci; p++) {
            const double dx = wrap1(qx - readlane_f64(pix, p), L0);
            const double dy = wrap1(qy - readlane_f64(piy, p), L1);
            const double dz = wrap1(qz - readlane_f64(piz, p), L2);
            hit |= dx * dx + dy * dy + dz * dz <= bv.cutoff2;
          }
        }
      }
      const unsigned long long bal = __ballot(hit);
      if (hit) W.hit_start[nh + wave_rank(bal)] = j;
      nh += __popcll(bal);
      __builtin_amdgcn_wave_barrier();
    }
    if (nh > 0) {  // wave-uniform
      e += wave_hits_energy<MIXED>(bv, W, na, nh, P64, P32, np, p0x, p0y, p0z);
      nh = 0;
      seg_end = 0;  // the staging reused the scan's bitmap: map again
    }
  } while (base < nitems);
  pairs = np;  // wave-uniform count
  return wave_sum_f64(e);
}

// A 256-lane work-group's share of a box's full energy: molecules
// blockIdx.x * 4 + wave, + 4 gridDim.x, ...; partial[i] = i's share and,
// when pairs is given, pairs[i] = the real site pairs it summed.
template <bool MIXED>
__device__ __forceinline__ void full_energy_block(const BoxView& bv, double* partial,
                                                  double* pairs, WaveEval* ws,
                                                  SiteParam64* p64, SiteParam32* p32) {
  if ((int)blockIdx.x * 4 >= bv.n_mol) return;  // the whole block: no molecule
  for (int k = threadIdx.x; k < bv.n_param; k += 256) {
    p64[k] = bv.p64[k];
    p32[k] = bv.p32[k];
  }
  __syncthreads();
  const int w = threadIdx.x >> 6;
  for (int i = blockIdx.x * 4 + w; i < bv.n_mol; i += gridDim.x * 4) {
    unsigned np = 0;
    const double e = wave_molecule_energy<MIXED>(bv, ws[w], i, p64, p32, np);
    if (wlane() == 0) {
      partial[i] = e;
      if (pairs) pairs[i] = (double)np;
    }
  }
}

}  // namespace mcgpu
