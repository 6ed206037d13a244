// Full-system energy on the device, a throughput kernel body shared by the
// single-box and batched paths.
#pragma once

#include "cells.h"
#include "device_common.h"
#include "eval.h"
#include "wave.h"

namespace mcgpu {

// Full-system energy (SimulationStep.cpp:40-71: the sum over i of E_i(i),
// partners j > i) of every listed box, one wave per molecule i: the lanes
// scan the candidates (the cells around i's first primary, or all j > i),
// test the primary cutoff (SimulationStep.cpp:317-330), compact the hits
// into the wave's LDS list by ballot, then sum the site pairs of the hits
// (each lane one (hit, site b) item against every site a of i), and the
// wave sum goes to partial[i].  Waves never wait for each other: a
// throughput kernel that keeps many waves per CU in flight.
constexpr int kWaveHits = 256;

struct WaveEval {
  double xa[kMaxSites], ya[kMaxSites], za[kMaxSites];  // sites of molecule i
  int32_t pa[kMaxSites];                               // their parameter entries
  int32_t hit_start[kWaveHits];
  int32_t hit_param[kWaveHits];
  int32_t hit_len[kWaveHits];
  int32_t hit_img[kWaveHits];  // periodic image of the pair (image_code)
};

template <bool MIXED>
__device__ __forceinline__ double wave_hits_energy(const BoxView& bv, const WaveEval& W,
                                                   int len_i, int nh,
                                                   const SiteParam64* P64,
                                                   const SiteParam32* P32, unsigned& npairs) {
  const int M = bv.max_len;
  const unsigned long long magic = (1ull << 32) / (unsigned)M + 1ull;
  const double L0 = bv.L[0], L1 = bv.L[1], L2 = bv.L[2];
  double e = 0.0;
  for (int u = wlane(); u < nh * M; u += 64) {
    const int k = div_small(u, magic);
    const int b = u - k * M;
    if (b >= W.hit_len[k]) continue;
    const int sb = W.hit_start[k] + b;
    const int pb = W.hit_param[k] + b;
    const SiteParam64 B = P64[pb];
    if (B.real == 0.0) continue;
    // site b carries the pair's periodic image, unless the hit's site pairs
    // may wrap differently (then wrap1 per site pair)
    const int img = W.hit_img[k];
    const bool per_pair = img & kImageMixed;
    const double xb = bv.x[sb] + (per_pair ? 0.0 : image_shift(img & 3, L0));
    const double yb = bv.y[sb] + (per_pair ? 0.0 : image_shift((img >> 2) & 3, L1));
    const double zb = bv.z[sb] + (per_pair ? 0.0 : image_shift((img >> 4) & 3, L2));
    for (int a = 0; a < len_i; a++) {
      const int pa = W.pa[a];
      const SiteParam64 A = P64[pa];
      if (A.real == 0.0) continue;
      double dx = xb - W.xa[a], dy = yb - W.ya[a], dz = zb - W.za[a];
      if (per_pair) {
        dx = wrap1(dx, L0);
        dy = wrap1(dy, L1);
        dz = wrap1(dz, L2);
      }
      if (MIXED) {
        const SiteParam32 A32 = P32[pa], B32 = P32[pb];
        e += pair_energy32(dx, dy, dz, A32.s, A32.e, A32.q, B32.s, B32.e, B32.q, A.s, A.e, B.s,
                           B.e);
      } else {
        e += pair_energy64(dx, dy, dz, A.s, A.e, A.q, B.s, B.e, B.q);
      }
      npairs++;
    }
  }
  return e;
}

// Returns E_i(i) (wave-uniform); pairs receives the real site pairs summed.
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
  if (lane < len) {
    W.xa[lane] = bv.x[start + lane];
    W.ya[lane] = bv.y[start + lane];
    W.za[lane] = bv.z[start + lane];
    W.pa[lane] = param + lane;
  }
  // the primaries of i (at most kMaxSites) as lane values
  double pix = 0.0, piy = 0.0, piz = 0.0;
  if (lane < pci) {
    pix = bv.px[psi + lane];
    piy = bv.py[psi + lane];
    piz = bv.pz[psi + lane];
  }
  __builtin_amdgcn_wave_barrier();
  const double L0 = bv.L[0], L1 = bv.L[1], L2 = bv.L[2];
  // candidate source
  const CellGrid& g = bv.cells;
  const bool cells = bv.mode == MCGPU_NB_CELL && g.overflow[0] == 0;
  CellRange r{};
  int ncell = 0, incl = 0, nitems = bv.n_mol - (i + 1);
  if (cells) {
    int co[3];
    co[0] = cell_coord(readlane_f64(pix, 0), g.inv_w[0], g.n[0]);
    co[1] = cell_coord(readlane_f64(piy, 0), g.inv_w[1], g.n[1]);
    co[2] = cell_coord(readlane_f64(piz, 0), g.inv_w[2], g.n[2]);
    r = make_range(g, co, co);  // at most 3 x 3 x 3 cells
    ncell = r.total();
    int n = 0;
    if (lane < ncell) {
      n = g.count[r.cell(g, lane)];
      n = n < g.cap ? n : g.cap;
    }
    incl = wave_incl_scan_i32(n);
    nitems = readlane_i32(incl, 63);
  }
  double e = 0.0;
  int nh = 0;
  for (int base = 0; base < nitems; base += 64) {
    const int item = base + lane;
    int j = -1;
    if (cells) {
      // cell holding `item`: the number of cells whose inclusive prefix is
      // <= item (all lanes take part in the lane reads)
      int lo = 0;
      for (int step = 32; step > 0; step >>= 1) {
        const int probe = lo + step - 1;
        const int v = __shfl(incl, probe < 63 ? probe : 63);
        if (probe < ncell && v <= item) lo += step;
      }
      const int before = __shfl(incl, lo > 0 ? lo - 1 : 0);
      if (item < nitems) j = g.slots[r.cell(g, lo) * g.cap + (item - (lo > 0 ? before : 0))];
    } else if (item < nitems) {
      j = i + 1 + item;
    }
    bool hit = false;
    int hs = 0, hp = 0, hl = 0, himg = 0;
    if (j > i) {
      const int psj = bv.single_primary ? j : bv.mol_pidx_start[j];
      const int pcj = bv.single_primary ? 1 : bv.mol_pidx_count[j];
      for (int q = 0; q < pcj; q++) {
        const double qx = bv.px[psj + q], qy = bv.py[psj + q], qz = bv.pz[psj + q];
        for (int p = 0; p < pci; p++) {
          const double dx = wrap1(qx - readlane_f64(pix, p), L0);
          const double dy = wrap1(qy - readlane_f64(piy, p), L1);
          const double dz = wrap1(qz - readlane_f64(piz, p), L2);
          hit |= dx * dx + dy * dy + dz * dz <= bv.cutoff2;
        }
      }
      if (hit) {
        himg = image_code(bv.px[psj] - readlane_f64(pix, 0), bv.py[psj] - readlane_f64(piy, 0),
                          bv.pz[psj] - readlane_f64(piz, 0), bv.L, bv.reach2);
        if (bv.uni_len > 0) {
          hl = bv.uni_len;
          hs = j * hl;
          hp = bv.uni_param;
        } else {
          const MolMeta m = bv.meta[j];
          hs = m.start;
          hl = m.len;
          hp = m.param;
        }
      }
    }
    const unsigned long long bal = __ballot(hit);
    const int cnt = __popcll(bal);
    if (nh + cnt > kWaveHits) {  // wave-uniform
      e += wave_hits_energy<MIXED>(bv, W, len, nh, P64, P32, np);
      __builtin_amdgcn_wave_barrier();
      nh = 0;
    }
    if (hit) {
      const int pos = nh + wave_rank(bal);
      W.hit_start[pos] = hs;
      W.hit_param[pos] = hp;
      W.hit_len[pos] = hl;
      W.hit_img[pos] = himg;
    }
    nh += cnt;
    __builtin_amdgcn_wave_barrier();
  }
  if (nh > 0) e += wave_hits_energy<MIXED>(bv, W, len, nh, P64, P32, np);
  __builtin_amdgcn_wave_barrier();
  pairs = wave_sum_u32(np);
  return wave_sum_f64(e);
}

// A 256-lane work-group's share of a box's full energy: molecules
// blockIdx.x * 4 + wave, + 4 gridDim.x, ...; partial[i] = E_i(i) and, when
// pairs is given, pairs[i] = the real site pairs summed.
__device__ __forceinline__ void full_energy_block(const BoxView& bv, bool mixed,
                                                  double* partial, double* pairs,
                                                  WaveEval* ws, SiteParam64* p64,
                                                  SiteParam32* p32) {
  for (int k = threadIdx.x; k < bv.n_param; k += 256) {
    p64[k] = bv.p64[k];
    p32[k] = bv.p32[k];
  }
  __syncthreads();
  const int w = threadIdx.x >> 6;
  for (int i = blockIdx.x * 4 + w; i < bv.n_mol; i += gridDim.x * 4) {
    unsigned np = 0;
    const double e = mixed ? wave_molecule_energy<true>(bv, ws[w], i, p64, p32, np)
                           : wave_molecule_energy<false>(bv, ws[w], i, p64, p32, np);
    if (wlane() == 0) {
      partial[i] = e;
      if (pairs) pairs[i] = (double)np;
    }
  }
}

}  // namespace mcgpu
