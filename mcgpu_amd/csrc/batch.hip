// Batched boxes: hundreds of independent boxes advanced by one persistent
// launch.  One work-group owns one box for the whole launch and runs the
// reference's step loop (Simulation.cpp:219-264) on the device: its own
// glibc-TYPE_3 RNG stream (MathLibrary.cpp:27-33), the rigid trial move
// (SimulationStep.cpp:520-566), old+new energy in one pass, the Metropolis
// test (SimulationStep.cpp:794-803), commit or nothing (rollback is free:
// proposals never touch the box).  No grid-wide synchronisation: boxes are
// independent, so a launch is as long as its slowest box.
//
// Box memory is carved from one HBM arena (admission by footprint).  NPT
// volume moves run on a stream of their own beside a launch, decided on the
// device by a short sequence of multi-box kernels around the full-energy
// pass (see queue_volume_moves).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "box_internal.h"
#include "cells.h"
#include "eval.h"
#include "fastmath.h"
#include "fullenergy.h"
#include "mcgpu_host.h"
#include "wave.h"

namespace mcgpu {

// Per-box control/state, written by the kernel, mirrored to the host after
// every launch.
struct BoxCtl {
  int64_t steps_done;
  int64_t accepted;
  int64_t steps_target;
  int64_t next_stop;  // the kernel stops this box when steps_done reaches it
  int64_t work_cand;   // candidate molecules examined (sum over moves)
  int64_t work_sites;  // sites of in-range molecules loaded
  int64_t work_self;   // sites of the moving molecule (old + proposed)
  int64_t vol_moves;   // NPT volume moves tried / accepted
  int64_t vol_acc;
  int64_t fallbacks;   // moves of a cell-list box that scanned every molecule
  double energy;
  double start_energy;
  double L[3];         // current edges
  double lrc;          // current tail correction (scaled with 1/V)
  // the volume move in flight
  double v_old, v_new, u_accept;
  double L_old[3];
  int32_t status;  // kRunning / kDone / kVolumePending
  int32_t fault;   // MCGPU_FAULT_* that stopped the box (0 = none)
};
constexpr int32_t kRunning = 0, kDone = 1, kVolumePending = 2;

struct BatchBox {
  BoxView bv;
  Rng31* rng;
  int32_t* tr_mol;
  double* tr_de;
  int8_t* tr_acc;
  double* partial;  // [n_mol] per-molecule energies of a full-energy pass
  double* total;    // [4] full-energy totals
  double* bx;       // [n_site] x3: coordinates before an NPT rescale
  double* by;
  double* bz;
  mcgpu_volume_record* tr_vol;  // [tr_vol_cap] NPT volume moves (record_trace)
  double max_t;
  double max_r;
  int32_t precision;
  int32_t record;
  // NPT: pressure (kcal/mol/A^3), half-width of the volume step (A^3),
  // moves between volume moves; cell geometry inputs (cutoff, primary
  // spread, the cells the allocation holds; 0 = no cell lists)
  double pressure;
  double dv;
  double cutoff;
  double spread;
  int32_t vol_interval;
  int32_t ncell_cap;
  int64_t tr_vol_cap;
  // NPT cell boxes: the other set of cell-list arrays and its mode.  A
  // volume move builds the new geometry's lists in it and swaps it in; an
  // undone move swaps the untouched lists back.
  CellGrid cells_alt;
  int32_t mode_alt;
  int32_t pad_alt;
};

// One box of a batched full-energy / rescale / cell pass.
struct VolItem {
  int32_t slot;
  int32_t undo;  // rejected volume move: restore the coordinates
  double f;      // edge scale factor of the rescale
};

// ------------------------------------------------------------- kernel

// LDS state of one work-group's box for a launch.
template <int NT>
struct BatchShared {
  EvalShared<NT> ev;
  BoxCtl c;      // the box's control block (wave 0, lane 0 updates it)
  int32_t go;    // the current move runs (else the box stops this launch)
  int32_t mol, start, pstart;
  int32_t cells_live;  // cell lists in use and complete for this move
  int32_t co_cell, cn_cell;  // first-primary cell before / after the move
  // the cell lists' overflow flag, and whether ev.ccount holds the cell
  // occupancies, for this launch (only this launch's moves change them)
  int32_t ovf, cc_live;
};

// Wave 0: draws and builds the next trial move (Simulation.cpp:219-264 draw
// order: molecule, vertex, dX, dY, dZ, rX, rY, rZ; SimulationStep.cpp:
// 520-566 for the move itself), and prepares the neighbour source.
//
// The eight draws come from the wave-resident generator; lane k then forms
// U_k = (b - a) * (r_k / RAND_MAX) + a for its own draw (MathLibrary.cpp:
// 31-33) and lanes 5..7 take sin/cos of their angle, so the serial chain is
// eight register draws long.  Lane k < len holds site k: the pivot and the
// first primary are read across lanes, every site except the pivot is
// rotated about the pivot by the X, Y, Z Euler angles, all are translated,
// and the molecule is wrapped by its first primary (SimulationStep.cpp:
// 387-422, 545-565, 710-729).  FMA contraction is off so the doubles match
// the reference's.
template <int NT>
__device__ __forceinline__ void front_move(BatchShared<NT>& S, const BatchBox& B, WaveRng& g, int s,
                                           long long budget, long long t_start) {
#pragma clang fp contract(off)
  const BoxView& bv = B.bv;
  EvalShared<NT>& sh = S.ev;
  const int lane = wlane();
  bool go = S.c.steps_done < S.c.next_stop;
  if (go && budget > 0 && s > 0) go = (wall_clock64() - t_start) < budget;
  if (!go) {
    if (lane == 0) S.go = 0;
    return;
  }
  const int n = bv.n_mol;
  int R = 0;
#pragma unroll
  for (int k = 0; k < 8; k++) {
    const int r = wrng_next(g);
    R = lane == k ? r : R;
  }
  const double q = (double)R / 2147483647;
  double a = 0.0, b = 0.0;
  if (lane == 0) {
    b = (double)n;
  } else if (lane >= 2 && lane < 5) {
    a = -B.max_t;
    b = B.max_t;
  } else if (lane >= 5 && lane < 8) {
    a = -B.max_r;
    b = B.max_r;
  }
  const double U = (b - a) * q + a;
  int mol = (int)readlane_f64(U, 0);
  mol = mol >= n ? n - 1 : mol;  // rand() == RAND_MAX
  // the molecule's record
  int start, len, param, p0, pstart, npri;
  int fault = 0;
  if (bv.uni_len > 0) {
    len = bv.uni_len;
    start = mol * len;
    param = bv.uni_param;
    p0 = bv.uni_p0;
    pstart = mol;
    npri = 1;
  } else {
    const MolMeta m = bv.meta[mol];
    start = m.start;
    len = m.len;
    param = m.param;
    p0 = m.p0;
    pstart = m.pstart;
    npri = bv.single_primary ? 1 : bv.mol_pidx_count[mol];
    if (len < 1 || len > kMaxSites) {
      fault = MCGPU_FAULT_MOLLEN;
      len = len < 1 ? 1 : kMaxSites;
    }
    npri = npri < 1 ? 1 : (npri > kMaxSites ? kMaxSites : npri);
  }
  const int cells_live = bv.mode == MCGPU_NB_CELL && S.ovf == 0;
  // committed stores of the previous move (other lanes' cell counts) are
  // complete before this move's loads
  __threadfence_block();
  double xo = 0.0, yo = 0.0, zo = 0.0;
  int pof = p0;
  if (lane < len) {
    xo = bv.x[start + lane];
    yo = bv.y[start + lane];
    zo = bv.z[start + lane];
    sh.q64[lane] = sh.prm64[param + lane];
    sh.q32[lane] = sh.prm32[param + lane];
  }
  if (npri > 1 && lane < npri) pof = bv.pidx[pstart + lane] - start;
  // the rotation's draws while the sites load
  const double rad = U * 3.14159265358979 / 180.0;
  double sn, cs;
  sincos_fp64(rad, sn, cs);
  int vtx = (int)(((double)len - 0.0) * readlane_f64(q, 1) + 0.0);
  vtx = vtx >= len ? len - 1 : vtx;
  const double pvx = readlane_f64(xo, vtx), pvy = readlane_f64(yo, vtx),
               pvz = readlane_f64(zo, vtx);
  const double c0 = readlane_f64(cs, 5), s0 = readlane_f64(sn, 5);
  const double c1 = readlane_f64(cs, 6), s1 = readlane_f64(sn, 6);
  const double c2 = readlane_f64(cs, 7), s2 = readlane_f64(sn, 7);
  double X = xo, Y = yo, Z = zo;
  if (lane != vtx) {
    X += -pvx;
    Y += -pvy;
    Z += -pvz;
    double u = Y, v = Z;  // rotateX
    Y = u * c0 + v * s0;
    Z = v * c0 - u * s0;
    u = Z;
    v = X;  // rotateY
    Z = u * c1 + v * s1;
    X = v * c1 - u * s1;
    u = X;
    v = Y;  // rotateZ
    X = u * c2 + v * s2;
    Y = v * c2 - u * s2;
    X += pvx;
    Y += pvy;
    Z += pvz;
  }
  X += readlane_f64(U, 2);
  Y += readlane_f64(U, 3);
  Z += readlane_f64(U, 4);
  const int pf = readlane_i32(pof, 0);  // first primary
  {
    const double qx = readlane_f64(X, pf), qy = readlane_f64(Y, pf),
                 qz = readlane_f64(Z, pf);
    const double sx = qx < 0 ? bv.L[0] : (qx > bv.L[0] ? -bv.L[0] : 0.0);
    const double sy = qy < 0 ? bv.L[1] : (qy > bv.L[1] ? -bv.L[1] : 0.0);
    const double sz = qz < 0 ? bv.L[2] : (qz > bv.L[2] ? -bv.L[2] : 0.0);
    if (sx != 0.0) X += sx;
    if (sy != 0.0) Y += sy;
    if (sz != 0.0) Z += sz;
  }
  if (lane < len) {
    sh.ox[lane] = xo;
    sh.oy[lane] = yo;
    sh.oz[lane] = zo;
    sh.nx[lane] = X;
    sh.ny[lane] = Y;
    sh.nz[lane] = Z;
  }
  if (lane < npri) sh.pofs[lane] = pof;
  // neighbour source
  int src = MCGPU_NB_BRUTE, nitems = n, ncells = 0, cold = 0, cnew = 0;
  if (cells_live) {
    const CellGrid& cg = bv.cells;
    int co[3], cn[3];
    co[0] = cell_coord(readlane_f64(xo, pf), cg.inv_w[0], cg.n[0]);
    co[1] = cell_coord(readlane_f64(yo, pf), cg.inv_w[1], cg.n[1]);
    co[2] = cell_coord(readlane_f64(zo, pf), cg.inv_w[2], cg.n[2]);
    cn[0] = cell_coord(readlane_f64(X, pf), cg.inv_w[0], cg.n[0]);
    cn[1] = cell_coord(readlane_f64(Y, pf), cg.inv_w[1], cg.n[1]);
    cn[2] = cell_coord(readlane_f64(Z, pf), cg.inv_w[2], cg.n[2]);
    cold = cell_index(cg, co[0], co[1], co[2]);
    cnew = cell_index(cg, cn[0], cn[1], cn[2]);
    int ni = 0;
    // (the item -> cell map is filled by every wave after the barrier)
    const int nc = prepare_cells_wave(bv, sh, co, cn, &ni, S.cc_live != 0, false);
    if (nc >= 0) {
      src = MCGPU_NB_CELL;
      nitems = ni;
      ncells = nc;
    }
  }
  if (nitems < 0 || nitems > n) {  // corrupt lists: evaluate nothing, stop
    fault = MCGPU_FAULT_CANDIDATES;
    nitems = 0;
  }
  if (lane == 0) {
    // a cell-list box whose move scans every molecule
    if (bv.mode == MCGPU_NB_CELL && src != MCGPU_NB_CELL) S.c.fallbacks++;
    sh.src = src;
    sh.nitems = nitems;
    sh.ncells = ncells;
    sh.len = len;
    sh.npri = npri;
    sh.aw_log2 = lane_group_log2(len);
    sh.fault = fault;
    S.mol = mol;
    S.start = start;
    S.pstart = pstart;
    S.cells_live = cells_live;
    S.co_cell = cold;
    S.cn_cell = cnew;
    S.go = 1;
  }
}

// Wave 0, after the evaluation: the Metropolis test (SimulationStep.cpp:
// 794-803; the uniform is drawn only when dE >= 0), the per-box counters and
// trace, and on acceptance the commit of the proposed sites, primary mirror
// and cell lists.
template <int NT>
__device__ __forceinline__ void finish_move(BatchShared<NT>& S, const BatchBox& B, WaveRng& g) {
  const BoxView& bv = B.bv;
  const EvalShared<NT>& sh = S.ev;
  const int lane = wlane();
  double eo, en;
  unsigned long long wc;
  unsigned np;
  eval_totals<NT>(sh, eo, en, wc, np);
  const double de = readlane_f64(en - eo, 0);
  int acc;
  if (de < 0) {
    acc = 1;
  } else {
    acc = exp_fp64(-(de) / bv.kT) >= random_real(wrng_next(g), 0.0, 1.0);
  }
  const int mol = S.mol, start = S.start, len = sh.len;
  if (lane == 0) {
    const int64_t step = S.c.steps_done;
    S.c.work_cand += (unsigned)(wc & 0xffffffffu);
    S.c.work_sites += (unsigned)(wc >> 32);
    S.c.work_self += 2 * len;
    if (B.record) {
      B.tr_mol[step] = mol;
      B.tr_de[step] = de;
      B.tr_acc[step] = (int8_t)acc;
    }
    if (acc) {
      S.c.energy += de;
      S.c.accepted++;
    }
    S.c.steps_done = step + 1;
  }
  if (!acc) return;
  if (lane < len) {
    bv.x[start + lane] = sh.nx[lane];
    bv.y[start + lane] = sh.ny[lane];
    bv.z[start + lane] = sh.nz[lane];
  }
  if (lane < sh.npri) {
    const int o = sh.pofs[lane];
    bv.px[S.pstart + lane] = sh.nx[o];
    bv.py[S.pstart + lane] = sh.ny[o];
    bv.pz[S.pstart + lane] = sh.nz[o];
  }
  // the cell of the first primary changes for a few per cent of moves; the
  // old cell is the one its current position maps to (cells are assigned
  // from the same positions), checked against the list before it is used
  if (lane == 0 && S.cells_live && S.co_cell != S.cn_cell) {
    const CellGrid& cg = bv.cells;
    if (cg.mol_cell[mol] == S.co_cell) {
      const bool placed = cell_move(cg, mol, S.co_cell, S.cn_cell);
      if (S.cc_live) {  // the LDS occupancies follow the lists
        S.ev.ccount[S.co_cell] -= 1;
        if (placed) S.ev.ccount[S.cn_cell] += 1;
      }
      if (!placed) S.ovf = 1;
    } else {
      cg.overflow[0] = 1;  // lists disagree with positions: scan everything
      S.ovf = 1;
    }
  }
}

// The persistent launch: work-group g owns box active[1 + g] and runs up to
// `sweeps` trial moves on it, stopping early at the box's step target, at a
// pending NPT volume move, or when `budget` wall-clock ticks have passed
// since the work-group started (0 = no time limit).  Wave 0 draws, builds,
// decides and commits each move; all waves scan candidates and sum pairs.
// The LATE instance runs the boxes of a tick whose NPT volume move comes
// first (the volume pass queued just before it on its stream, beside the
// rest of the tick): their budget counts from *t0, the wall clock stamped on
// that stream as the tick began, so they stop with the rest of the tick.
template <int NT, bool LATE>
__global__ __launch_bounds__(NT) void k_batch(const BatchBox* __restrict__ table,
                                              BoxCtl* ctl,
                                              const int32_t* active,
                                              int sweeps, long long budget,
                                              const long long* t0) {
  __shared__ BatchShared<NT> S;
  const int t = threadIdx.x;
  const bool w0 = t < 64;
  // active = [count, slot 0, slot 1, ...]; the grid is rounded up
  if ((int)blockIdx.x >= __builtin_amdgcn_readfirstlane(active[0])) return;
  const int slot = __builtin_amdgcn_readfirstlane(active[1 + blockIdx.x]);
  const long long t_start = LATE ? *t0 : wall_clock64();
  // the descriptor is read in place (scalar loads from a table this launch
  // does not write)
  const BatchBox& B = table[slot];
  if (t == 0) S.c = ctl[slot];
  __syncthreads();
  WaveRng g{0, 0, 0};
  if (w0) {
    const Rng31* rs = B.rng;
    const int lane = wlane();
    g.s = lane < 31 ? rs->s[lane] : 0;
    g.f = __builtin_amdgcn_readfirstlane(rs->f);
    g.r = __builtin_amdgcn_readfirstlane(rs->r);
  }
  const BoxView& bv = B.bv;
  const bool mixed = B.precision == MCGPU_MIXED;
  load_params<NT>(bv, S.ev);
  // the cell occupancies into LDS, when the grid fits: only this launch's
  // commits change them until it ends
  const bool cell_box = bv.mode == MCGPU_NB_CELL;
  const int ncell = cell_box ? bv.cells.n[0] * bv.cells.n[1] * bv.cells.n[2] : 0;
  const bool cc = cell_box && ncell <= EvalShared<NT>::kCellLds;
  if (cc)
    for (int c = t; c < ncell; c += NT) S.ev.ccount[c] = bv.cells.count[c];
  if (t == 0) {
    S.cc_live = cc;
    S.ovf = cell_box ? bv.cells.overflow[0] : 0;
  }
  __syncthreads();
  for (int s = 0; s < sweeps; s++) {
    if (w0) front_move<NT>(S, B, g, s, budget, t_start);
    __syncthreads();
    if (!S.go || S.ev.fault) break;
    if (__builtin_amdgcn_readfirstlane(S.ev.src) == MCGPU_NB_CELL) {
      fill_item_cells<NT>(S.ev, __builtin_amdgcn_readfirstlane(S.ev.ncells));
      __syncthreads();
    }
    const int mol = __builtin_amdgcn_readfirstlane(S.mol);
    if (mixed)
      eval_core<NT, true>(bv, S.ev, mol, 0, true, false);
    else
      eval_core<NT, false>(bv, S.ev, mol, 0, true, false);
    if (w0) finish_move<NT>(S, B, g);
  }
  if (w0) {
    const int lane = wlane();
    Rng31* rs = B.rng;
    if (lane < 31) rs->s[lane] = g.s;
    if (lane == 0) {
      rs->f = g.f;
      rs->r = g.r;
      if (S.ev.fault && S.go) S.c.fault = S.ev.fault;
      if (S.c.steps_done >= S.c.steps_target || S.c.fault)
        S.c.status = kDone;
      else if (S.c.steps_done >= S.c.next_stop)
        S.c.status = kVolumePending;
      ctl[slot] = S.c;
    }
  }
}

// Start energy of a newly added box: pairs (reduced by k_reduce_multi into
// total[0]) plus the offset (the LRC), or, for a box that continues from a
// resume record, that record's running energy.
__global__ void k_set_start_energy(BoxCtl* c, const double* total, double offset,
                                   int resumed, double energy) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    c->energy = resumed ? energy : total[0] + offset;
    c->start_energy = c->energy;
  }
}

// ---------------------------------------- batched NPT passes (many boxes)
//
// One launch covers every box of a list (blockIdx.y = list entry), so the
// volume moves of all boxes due in a tick cost a handful of launches.

// Cell geometry of a box for its current edges (the host's cell_geometry):
// cells of width >= rc + 2 spread, cell mode while the grid fits the
// allocation, else the all-molecule scan.
__device__ void set_geometry_dev(BatchBox& B) {
  BoxView& v = B.bv;
  if (B.ncell_cap == 0) {
    v.mode = MCGPU_NB_BRUTE;
    return;
  }
  const double w = B.cutoff + 2.0 * B.spread + 1e-9;
  int n[3];
  for (int d = 0; d < 3; d++) {
    n[d] = (int)floor(v.L[d] / w);
    n[d] = n[d] < 1 ? 1 : n[d];
  }
  if (n[0] * n[1] * n[2] > B.ncell_cap) {
    v.mode = MCGPU_NB_BRUTE;
    return;
  }
  for (int d = 0; d < 3; d++) {
    v.cells.n[d] = n[d];
    v.cells.inv_w[d] = n[d] / v.L[d];
  }
  v.mode = MCGPU_NB_CELL;
}

// The wall clock into *t0: the start of a tick for the boxes that move after
// their volume move.
__global__ void k_stamp_clock(long long* t0) { *t0 = wall_clock64(); }

// NPT volume move, first half, one thread per listed box (semantics: see
// queue_volume_moves): u1 and u2 drawn from the box's own generator, the
// new edges and cell geometry set, the scale factor handed to the rescale.
// V' is the product of the new edges, as acceptVolumeMove reads it
// (SimulationStep.cpp:476-481); a proposal <= 0 keeps the box (f = 1) and
// is rejected by the decision.
__global__ void k_vol_begin(BatchBox* table, BoxCtl* ctl, VolItem* items, int k) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const int s = items[i].slot;
  BatchBox& B = table[s];
  BoxCtl& c = ctl[s];
  Rng31 g = *B.rng;
  const double u1 = random_real(rng31_next(g), 0.0, 1.0);
  c.u_accept = random_real(rng31_next(g), 0.0, 1.0);
  *B.rng = g;
  BoxView& v = B.bv;
  for (int d = 0; d < 3; d++) c.L_old[d] = v.L[d];
  c.v_old = v.L[0] * v.L[1] * v.L[2];
  const double v_prop = c.v_old + B.dv * (2.0 * u1 - 1.0);
  const double f = v_prop > 0.0 ? cbrt(v_prop / c.v_old) : 1.0;
  items[i].f = f;
  items[i].undo = 0;
  for (int d = 0; d < 3; d++) v.L[d] = c.L_old[d] * f;
  c.v_new = v_prop > 0.0 ? v.L[0] * v.L[1] * v.L[2] : v_prop;
  // the new geometry's cell lists go to the other set of arrays; the
  // current lists stay as they are for an undo
  const CellGrid cur = v.cells;
  v.cells = B.cells_alt;
  B.cells_alt = cur;
  B.mode_alt = v.mode;
  set_geometry_dev(B);
}

// Second half, after the full energy at the new volume (out[2 i + 1]):
// the acceptance test, the running energy and LRC on acceptance, the old
// edges and geometry (and the undo flag for the coordinates) on rejection,
// and the box's next volume stop.  e_pairs: the pair energy at the new
// volume.
__device__ void vol_decide(BatchBox& B, BoxCtl& c, VolItem& item, double e_pairs) {
#pragma clang fp contract(off)
  const bool valid = c.v_new > 0.0;
  const double lrc_new = valid ? c.lrc * c.v_old / c.v_new : c.lrc;
  const double e_new = valid ? e_pairs + lrc_new : c.energy;
  bool accept = false;
  if (valid) {
    // SimCalcs::acceptVolumeMove (SimulationStep.cpp:476-488)
    const double de = e_new - c.energy;
    double prob = B.bv.n_mol * (log(c.v_new) - log(c.v_old));
    prob -= (de + B.pressure * (c.v_new - c.v_old)) / B.bv.kT;
    prob = exp_fp64(prob);
    accept = prob >= c.u_accept;
  }
  if (B.tr_vol && c.vol_moves < B.tr_vol_cap) {
    mcgpu_volume_record r;
    r.step = c.steps_done;
    r.v_old = c.v_old;
    r.v_new = c.v_new;
    r.e_old = c.energy;
    r.e_new = e_new;
    r.u = c.u_accept;
    r.accepted = accept ? 1 : 0;
    r.reserved = 0;
    B.tr_vol[c.vol_moves] = r;
  }
  c.vol_moves++;
  if (accept) {
    c.vol_acc++;
    c.energy = e_new;
    c.lrc = lrc_new;
    for (int d = 0; d < 3; d++) c.L[d] = B.bv.L[d];
  } else {
    // the old edges, and the cell lists and mode from before the move
    for (int d = 0; d < 3; d++) B.bv.L[d] = c.L_old[d];
    const CellGrid cur = B.bv.cells;
    B.bv.cells = B.cells_alt;
    B.cells_alt = cur;
    B.bv.mode = B.mode_alt;
    item.undo = 1;
  }
  c.next_stop = c.steps_target < c.steps_done + B.vol_interval ? c.steps_target
                                                               : c.steps_done + B.vol_interval;
  c.status = c.steps_done >= c.steps_target ? kDone : kRunning;
}

// One instance per precision (admissions: the start energy of one box);
// blocks whose entry has the other precision return at once.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_energy_waves(const BatchBox* table,
                                                      const VolItem* items) {
  __shared__ WaveEval ws[4];
  __shared__ SiteParam64 p64[kParamCap];
  __shared__ SiteParam32 p32[kParamCap];
  const BatchBox& B = table[items[blockIdx.y].slot];
  if ((B.precision == MCGPU_MIXED) != MIXED) return;
  full_energy_block<MIXED>(B.bv, B.partial, nullptr, ws, p64, p32);
}

// Entry blockIdx.x of a volume pass: the fixed-order sum of its
// per-molecule partials at the new volume (as k_reduce_multi sums them),
// then the decision.
__global__ __launch_bounds__(256) void k_vol_settle(BatchBox* table, BoxCtl* ctl,
                                                    VolItem* items) {
  __shared__ double red[4];
  VolItem& item = items[blockIdx.x];
  BatchBox& B = table[item.slot];
  const int n = B.bv.n_mol;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += B.partial[i];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) {
    B.total[1] = s;
    vol_decide(B, ctl[item.slot], item, s);
  }
}

// Both precisions in one launch (a volume pass lists boxes of either): each
// block takes its entry's.
__global__ __launch_bounds__(256) void k_energy_waves_any(const BatchBox* table,
                                                          const VolItem* items) {
  __shared__ WaveEval ws[4];
  __shared__ SiteParam64 p64[kParamCap];
  __shared__ SiteParam32 p32[kParamCap];
  const BatchBox& B = table[items[blockIdx.y].slot];
  if (B.precision == MCGPU_MIXED)
    full_energy_block<true>(B.bv, B.partial, nullptr, ws, p64, p32);
  else
    full_energy_block<false>(B.bv, B.partial, nullptr, ws, p64, p32);
}

// Fixed-order sum of entry blockIdx.x's per-molecule partials into
// total[which] and out[2 * entry + which].
__global__ __launch_bounds__(256) void k_reduce_multi(const BatchBox* table,
                                                      const VolItem* items, int which,
                                                      double* out) {
  __shared__ double red[4];
  const BatchBox& B = table[items[blockIdx.x].slot];
  const int n = B.bv.n_mol;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) s += B.partial[i];
  s = block_sum<256>(s, red);
  if (threadIdx.x == 0) {
    B.total[which] = s;
    out[2 * blockIdx.x + which] = s;
  }
}

// Clears entry y's cell occupancies for its current geometry (block x's
// 256 cells), when the box is in cell mode.
__device__ __forceinline__ void clear_cells_block(const BoxView& bv) {
  if (bv.mode != MCGPU_NB_CELL) return;
  const int c = blockIdx.x * 256 + threadIdx.x;
  if (c < cells_total(bv.cells)) bv.cells.count[c] = 0;
  if (c == 0) bv.cells.overflow[0] = 0;
}

// NPT rescale of entry y: molecule m is translated so its first primary
// moves from p to f p (internal geometry untouched); its sites are backed up
// first and its primary mirror entries refreshed.  The entry's cell
// occupancies are cleared for the new geometry on the way (block x clears
// cells [256 x, 256 x + 256)).  A uniform box's block takes whole molecules
// (256 / len of them), one lane per site, so its loads and stores are
// contiguous: the first primaries' shifts go through LDS before any site is
// written.  Other boxes take one molecule per lane.
__global__ __launch_bounds__(256) void k_scale_multi(const BatchBox* table,
                                                     const VolItem* items) {
  __shared__ double shift[3][256];
  const VolItem it = items[blockIdx.y];
  const BatchBox& B = table[it.slot];
  const BoxView& bv = B.bv;
  clear_cells_block(bv);
  const double f = it.f;
  const int t = threadIdx.x;
  const int len = bv.uni_len;
  if (len > 0 && bv.single_primary) {
    const int per = 256 / len;  // molecules per block
    const int m0 = blockIdx.x * per;
    if (m0 >= bv.n_mol) return;
    const int nm = bv.n_mol - m0 < per ? bv.n_mol - m0 : per;
    if (t < nm) {
      const int m = m0 + t;
      const int p = m * len + bv.uni_p0;
      const double px = bv.x[p], py = bv.y[p], pz = bv.z[p];
      const double sx = px * f - px;
      const double sy = py * f - py;
      const double sz = pz * f - pz;
      shift[0][t] = sx;
      shift[1][t] = sy;
      shift[2][t] = sz;
      bv.px[m] = px + sx;
      bv.py[m] = py + sy;
      bv.pz[m] = pz + sz;
    }
    __syncthreads();
    if (t < nm * len) {
      const int s = m0 * len + t;
      const int k = t / len;
      const double x = bv.x[s], y = bv.y[s], z = bv.z[s];
      B.bx[s] = x;
      B.by[s] = y;
      B.bz[s] = z;
      bv.x[s] = x + shift[0][k];
      bv.y[s] = y + shift[1][k];
      bv.z[s] = z + shift[2][k];
    }
    return;
  }
  const int m = blockIdx.x * 256 + t;
  if (m >= bv.n_mol) return;
  const int p = bv.pidx[bv.mol_pidx_start[m]];
  const double sx = bv.x[p] * f - bv.x[p];
  const double sy = bv.y[p] * f - bv.y[p];
  const double sz = bv.z[p] * f - bv.z[p];
  const int s0 = bv.mol_start[m], s1 = s0 + bv.mol_len[m];
  for (int s = s0; s < s1; s++) {
    B.bx[s] = bv.x[s];
    B.by[s] = bv.y[s];
    B.bz[s] = bv.z[s];
    bv.x[s] += sx;
    bv.y[s] += sy;
    bv.z[s] += sz;
  }
  const int ps = bv.mol_pidx_start[m], pe = ps + bv.mol_pidx_count[m];
  for (int k = ps; k < pe; k++) {
    const int q = bv.pidx[k];
    bv.px[k] = bv.x[q];
    bv.py[k] = bv.y[q];
    bv.pz[k] = bv.z[q];
  }
}

// A rejected rescale undone: molecule m's sites and primaries restored (the
// decision put the cell lists from before the move back).  Blocks map to
// molecules as in k_scale_multi.
__global__ __launch_bounds__(256) void k_restore_multi(const BatchBox* table,
                                                       const VolItem* items) {
  if (!items[blockIdx.y].undo) return;
  const BatchBox& B = table[items[blockIdx.y].slot];
  const BoxView& bv = B.bv;
  const int t = threadIdx.x;
  const int len = bv.uni_len;
  if (len > 0 && bv.single_primary) {
    const int per = 256 / len;
    const int m0 = blockIdx.x * per;
    if (m0 >= bv.n_mol) return;
    const int nm = bv.n_mol - m0 < per ? bv.n_mol - m0 : per;
    if (t < nm * len) {
      const int s = m0 * len + t;
      bv.x[s] = B.bx[s];
      bv.y[s] = B.by[s];
      bv.z[s] = B.bz[s];
    }
    if (t < nm) {
      const int m = m0 + t;
      const int p = m * len + bv.uni_p0;
      bv.px[m] = B.bx[p];
      bv.py[m] = B.by[p];
      bv.pz[m] = B.bz[p];
    }
    return;
  }
  const int m = blockIdx.x * 256 + t;
  if (m >= bv.n_mol) return;
  const int s0 = bv.mol_start[m], s1 = s0 + bv.mol_len[m];
  for (int s = s0; s < s1; s++) {
    bv.x[s] = B.bx[s];
    bv.y[s] = B.by[s];
    bv.z[s] = B.bz[s];
  }
  const int ps = bv.mol_pidx_start[m], pe = ps + bv.mol_pidx_count[m];
  for (int k = ps; k < pe; k++) {
    const int q = bv.pidx[k];
    bv.px[k] = bv.x[q];
    bv.py[k] = bv.y[q];
    bv.pz[k] = bv.z[q];
  }
}

// Cell lists of every listed box in cell mode, built for the new geometry
// after k_scale_multi cleared the occupancies: assign, order (see cells.h).
__global__ __launch_bounds__(256) void k_cells_assign_multi(const BatchBox* table,
                                                            const VolItem* items) {
  const BoxView& bv = table[items[blockIdx.y].slot].bv;
  if (bv.mode != MCGPU_NB_CELL) return;
  const int m = blockIdx.x * 256 + threadIdx.x;
  if (m < bv.n_mol) cells_assign_one(bv, m);
}

__global__ __launch_bounds__(256) void k_cells_sort_multi(const BatchBox* table,
                                                          const VolItem* items) {
  const BoxView& bv = table[items[blockIdx.y].slot].bv;
  if (bv.mode != MCGPU_NB_CELL) return;
  const int c = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (c < cells_total(bv.cells)) cells_sort_wave(bv.cells, c);
}

// ------------------------------------------------------------ host side

// Best-fit allocator over one device chunk (offsets, 256-byte aligned).
class Arena {
 public:
  ~Arena() {
    if (base_) (void)hipFree(base_);
  }
  int init(size_t bytes) {
    bytes = (bytes + 255) & ~size_t(255);
    MCGPU_HIP_TRY(hipMalloc(&base_, bytes));
    cap_ = bytes;
    free_[0] = bytes;
    return MCGPU_OK;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = (bytes + 255) & ~size_t(255);
    auto best = free_.end();
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= bytes && (best == free_.end() || it->second < best->second))
        best = it;
    if (best == free_.end()) return false;
    *off = best->first;
    const size_t rest = best->second - bytes;
    const size_t roff = best->first + bytes;
    free_.erase(best);
    if (rest) free_[roff] = rest;
    live_[*off] = bytes;
    used_ += bytes;
    peak_ = std::max(peak_, used_);
    return true;
  }
  void release(size_t off) {
    auto it = live_.find(off);
    if (it == live_.end()) return;
    size_t size = it->second;
    live_.erase(it);
    used_ -= size;
    auto nx = free_.lower_bound(off);
    if (nx != free_.end() && off + size == nx->first) {
      size += nx->second;
      nx = free_.erase(nx);
    }
    if (nx != free_.begin()) {
      auto pv = std::prev(nx);
      if (pv->first + pv->second == off) {
        pv->second += size;
        return;
      }
    }
    free_[off] = size;
  }
  char* base() const { return (char*)base_; }
  size_t used() const { return used_; }
  size_t peak() const { return peak_; }
  size_t capacity() const { return cap_; }

 private:
  void* base_ = nullptr;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_, live_;
};

// Offsets of one box's arrays inside its arena block.
struct BoxLayout {
  size_t x, y, z, px, py, pz, ms, ml, mps, mpc, mp, meta, pidx, p64, p32;
  size_t ccount, cslots, mcell, mslot, covf, vstale, vcount, partial, pairs,
      total, bx, by, bz, trm, trd, tra, trv;
  size_t ccount2, cslots2, mcell2, mslot2, covf2;  // NPT cell boxes: second set
  bool alt_cells;
  int64_t trv_cap;
  size_t bytes;
  int ncell, cap;
};

BoxLayout plan_layout(const mcgpu_box_desc& d, const mcgpu_run_params& p,
                      const ParamTable& pt) {
  BoxLayout L{};
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += (std::max<size_t>(bytes, 8) + 255) & ~size_t(255);
    return o;
  };
  const size_t ns = d.n_sites, nm = d.n_molecules, np = d.n_primary;
  L.x = take(8 * ns);
  L.y = take(8 * ns);
  L.z = take(8 * ns);
  L.px = take(8 * np);
  L.py = take(8 * np);
  L.pz = take(8 * np);
  L.ms = take(4 * nm);
  L.ml = take(4 * nm);
  L.mps = take(4 * nm);
  L.mpc = take(4 * nm);
  L.mp = take(4 * nm);
  L.meta = take(sizeof(MolMeta) * nm);
  L.pidx = take(4 * np);
  L.p64 = take(sizeof(SiteParam64) * pt.p64.size());
  L.p32 = take(sizeof(SiteParam32) * pt.p32.size());
  L.ncell = 0;
  if (p.neighbor == MCGPU_NEIGHBOR_CELL) {
    // room for the grid of a box up to 26% larger in volume (NPT)
    double Lg[3];
    for (int k = 0; k < 3; k++) Lg[k] = d.box[k] * 1.08;
    int n[3], cn[3];
    double iw[3];
    cell_geometry(Lg, d.cutoff, pt.spread, n, iw);
    L.ncell = n[0] * n[1] * n[2];
    cell_geometry(d.box, d.cutoff, pt.spread, cn, iw);
    const int nc_now = cn[0] * cn[1] * cn[2];
    L.cap = ((int)std::ceil(3.0 * nm / nc_now) + 16 + 3) & ~3;
    L.ccount = take(4 * (size_t)L.ncell);
    L.cslots = take(4 * (size_t)L.ncell * L.cap);
  }
  L.mcell = take(4 * nm);
  L.mslot = take(4 * nm);
  L.covf = take(4);
  L.alt_cells = L.ncell > 0 && p.ensemble == MCGPU_NPT;
  if (L.alt_cells) {
    L.ccount2 = take(4 * (size_t)L.ncell);
    L.cslots2 = take(4 * (size_t)L.ncell * L.cap);
    L.mcell2 = take(4 * nm);
    L.mslot2 = take(4 * nm);
    L.covf2 = take(4);
  }
  L.vstale = take(4);
  L.vcount = take(4);
  L.partial = take(8 * std::max(ns, nm));  // per-site partial energies
  L.pairs = take(8 * nm);
  L.total = take(32);  // [E_pairs, pairs] twice (before / after a rescale)
  if (p.ensemble == MCGPU_NPT) {
    L.bx = take(8 * ns);
    L.by = take(8 * ns);
    L.bz = take(8 * ns);
  }
  if (p.record_trace) {
    L.trm = take(4 * (size_t)p.steps);
    L.trd = take(8 * (size_t)p.steps);
    L.tra = take((size_t)p.steps);
    if (p.ensemble == MCGPU_NPT) {
      L.trv_cap = p.steps / p.volume_interval + 1;
      L.trv = take(sizeof(mcgpu_volume_record) * (size_t)L.trv_cap);
    }
  }
  L.bytes = off;
  return L;
}

// Work-group size for a box, from the per-move work estimated with the
// molecule density inside the cutoff sphere: candidates scanned plus pair
// lane-steps.  256 lanes for light moves, 512 for medium, 1024 (16 waves,
// which hide the latency of the candidate loads and keep the FP64 units fed
// during the pair sums) for heavy ones; measured per-move latency on MI355X
// with one box per CU: water, 256 molecules, all-pairs scan 7.6 / 7.0 /
// 7.4 us (256 / 512 / 1024 lanes); water with cells 12.7 / 10.3 / 9.9 us;
// 16-site bicyclic with cells 23.8 / 17.9 / 16.6 us.
constexpr int kGroups = 3;
constexpr int kGroupNT[kGroups] = {256, 512, 1024};

double move_work(const mcgpu_box_desc& d, int neighbor, int max_len) {
  const double V = d.box[0] * d.box[1] * d.box[2];
  const double rho = d.n_molecules / V;
  const double rc = d.cutoff;
  const double hits = std::min<double>(d.n_molecules - 1, 4.18879 * rc * rc * rc * rho);
  double cand = d.n_molecules;
  if (neighbor == MCGPU_NEIGHBOR_CELL) cand = std::min<double>(cand, 27.0 * rc * rc * rc * rho);
  int aw = 1;
  while (aw < max_len) aw *= 2;
  return cand + hits * max_len * aw / 2;
}

int choose_group(double work) { return work > 1000.0 ? 2 : (work > 600.0 ? 1 : 0); }

}  // namespace mcgpu

using namespace mcgpu;

struct BatchSlot {
  bool used = false;
  size_t off = 0;
  BoxLayout lay{};
  mcgpu_run_params params{};
  int32_t n_pidx = 0;
  int32_t group = 0;   // preferred work-group size (estimate or tuner)
  double spread = 0.0; // largest primary-to-first-primary distance
  int fault = 0;
  int64_t tune_key = 0;    // box class for the work-group-size tuner
  int trial = -1;          // group being measured for the class (-1: none)
  int64_t steps_at_launch = 0;
  // counters before this run (a box added from a resume record)
  int64_t base_steps = 0, base_acc = 0, base_vol = 0, base_vacc = 0;
  bool dirty = false;      // the device's control block is newer than h_ctl
  bool late = false;       // runs after its volume move in the tick in flight
};

// One captured tick launch: the persistent kernel of each work-group-size
// group over its active list, side by side; replayed while the grid sizes,
// sweep count and time budget are unchanged.
struct TickGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int sweeps = -1;
  uint64_t version = 0;
  int grid[kGroups] = {};
};

// Host side of a batch.  Streams: `stream` carries the ticks and the
// read-back after them; `vstream`, beside each tick, the NPT volume moves of
// the boxes that reached a volume stop in the previous one (decided on the
// device) and then those boxes' own share of the tick (the late graph), which
// the read-back waits for (ev_vdone); `side[]` serve graph capture (the
// work-group-size groups of a tick run side by side); `sstream` admissions
// (their uploads, cells and start energy), which every tick waits for
// (ev_ready).
// Ownership: control blocks and generator states are the device's; the host
// mirrors them from each tick's read-back (merged for the boxes the tick
// ran) and writes a box's only when it sets the box up.
struct mcgpu_batch {
  hipStream_t stream = nullptr;
  hipStream_t side[kGroups - 1] = {};
  hipStream_t sstream = nullptr;
  hipStream_t vstream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, fork = nullptr;
  hipEvent_t join[kGroups - 1] = {};
  hipEvent_t ev_ready = nullptr;  // set-up work queued so far is done
  hipEvent_t ev_vdone = nullptr;  // the volume branch of the tick is done
  Arena arena;
  int32_t max_boxes = 0;
  int device = 0;  // the device current when the batch was created
  // Compute units, and the share of one a resident work-group of each size
  // takes (1 / the work-groups of that kernel one CU holds at once, from
  // its registers and LDS): a tick whose boxes need more than the device
  // holds runs its work-groups in rounds.
  int n_cu = 256;
  double cu_share[kGroups] = {1.0 / 3.0, 1.0, 1.0};
  std::vector<BatchSlot> slots;
  BatchBox* d_table = nullptr;
  BatchBox* h_table = nullptr;  // pinned mirror (the host's copy)
  BoxCtl* d_ctl = nullptr;
  BoxCtl* h_ctl = nullptr;      // pinned mirror
  BoxCtl* h_ctl_rb = nullptr;   // pinned read-back of a tick
  Rng31* d_rng = nullptr;
  Rng31* h_rng = nullptr;       // pinned mirror
  Rng31* h_rng_rb = nullptr;    // pinned read-back of a tick
  // [2 * kGroups * (max_boxes + 1)]: per group [count, slots...], the
  // main lists then the late ones
  int32_t* d_active = nullptr;
  int32_t* h_active = nullptr;  // pinned
  std::vector<int32_t> next_active;  // the lists of the tick being set up
  std::vector<int> eff_group;        // work-group size of each box this tick
  VolItem* d_items = nullptr;   // [max_boxes] list of a volume pass
  VolItem* h_items = nullptr;   // pinned
  VolItem* d_one = nullptr;     // [max_boxes] entry {s} of slot s (admission)
  long long* d_t0 = nullptr;    // wall clock at the start of the volume pass
  long long budget_ticks = 0;   // per-launch time limit (wall-clock ticks)
  double clock_khz = 100000.0;
  bool pending = false;         // a tick in flight
  std::vector<int> launched;    // slots of the tick in flight
  std::vector<int> late;        // of which after their volume move
  TickGraph main_graph, late_graph;
  uint64_t version = 1;  // bumped when the time budget changes
  // work-group-size tuner: per box class, moves per full tick at each size
  struct Tune {
    int best = -1;
    double moves[kGroups] = {};
    int samples[kGroups] = {};
    bool trialing = false;
  };
  std::map<int64_t, Tune> tune;
  bool autotune = true;
  // host seconds: waiting for ticks, volume moves, launching ticks
  double t_wait = 0.0, t_volume = 0.0, t_launch = 0.0;
  // timing
  double kernel_ms_sum = 0.0;
  int64_t kernel_launches = 0;
  bool timing_pending = false;
};

namespace {

double host_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

int upload_slot(mcgpu_batch* b, int s, hipStream_t st) {
  MCGPU_HIP_TRY(hipMemcpyAsync(b->d_table + s, b->h_table + s, sizeof(BatchBox),
                               hipMemcpyHostToDevice, st));
  return MCGPU_OK;
}

int upload_ctl(mcgpu_batch* b, int s, hipStream_t st) {
  MCGPU_HIP_TRY(hipMemcpyAsync(b->d_ctl + s, b->h_ctl + s, sizeof(BoxCtl),
                               hipMemcpyHostToDevice, st));
  return MCGPU_OK;
}

int upload_rng(mcgpu_batch* b, int s, hipStream_t st) {
  MCGPU_HIP_TRY(hipMemcpyAsync(b->d_rng + s, b->h_rng + s, sizeof(Rng31),
                               hipMemcpyHostToDevice, st));
  return MCGPU_OK;
}

char* slot_base(mcgpu_batch* b, int s) { return b->arena.base() + b->slots[s].off; }

// Cell geometry of slot s's current box into its descriptor (host copy):
// cell mode when the grid fits the allocation, else the all-molecule scan.
// Returns the cell count (0 without cells).
int slot_set_geometry(mcgpu_batch* b, int s) {
  BatchSlot& S = b->slots[s];
  BoxView& v = b->h_table[s].bv;
  if (S.lay.ncell == 0) {
    v.mode = MCGPU_NB_BRUTE;
    return 0;
  }
  int n[3];
  double iw[3];
  cell_geometry(v.L, b->h_table[s].cutoff, S.spread, n, iw);
  if (n[0] * n[1] * n[2] > S.lay.ncell) {
    v.mode = MCGPU_NB_BRUTE;
    return 0;
  }
  for (int k = 0; k < 3; k++) {
    v.cells.n[k] = n[k];
    v.cells.inv_w[k] = iw[k];
  }
  v.mode = MCGPU_NB_CELL;
  return n[0] * n[1] * n[2];
}

// The batched passes over the first k entries of a device list: per-molecule
// full-energy partials; cell lists for the current geometry (occupancies
// already cleared).
void launch_energy_partials(mcgpu_batch* b, const VolItem* items, int k, int max_mol,
                            hipStream_t st) {
  const dim3 grid((max_mol + 3) / 4, k);  // blocks past a box's molecules exit at once
  hipLaunchKernelGGL(k_energy_waves_any, grid, dim3(256), 0, st, b->d_table, items);
}

int launch_cells_pass(mcgpu_batch* b, const VolItem* items, int k, int max_mol,
                      int max_cells, hipStream_t st) {
  if (max_cells == 0) return MCGPU_OK;
  hipLaunchKernelGGL(k_cells_assign_multi, dim3((max_mol + 255) / 256, k), dim3(256), 0, st,
                     b->d_table, items);
  hipLaunchKernelGGL(k_cells_sort_multi, dim3((max_cells + 3) / 4, k), dim3(256), 0, st,
                     b->d_table, items);
  MCGPU_HIP_TRY(hipGetLastError());
  return MCGPU_OK;
}

// NPT volume moves of the boxes in `pend`, batched and decided on the
// device, queued on `st` at the start of the tick in which those boxes move
// on (beside the tick's other boxes): u1, u2 from each box's generator, the
// rescale (with backup), cells for the new geometry (in the box's second set
// of cell arrays, so that an undo only swaps the old lists back), the full
// energy, the decision, and the undo of rejected moves.  Semantics (the
// build's own; the reference's NPT path, Ensemble.cpp:9-28 /
// SimulationStep.cpp:476-518, always grows the box by 1.1 and never
// recomputes the energy):
//   V' = V + dv (2 u1 - 1), dv = delta V0 (V0 the volume at admission): a
//   fixed-width, symmetric step, so acceptVolumeMove's test is the whole
//   acceptance rule; molecules translated so first primaries scale by
//   f = (V'/V)^(1/3); V' then taken as the product of the new edges;
//   E' = E_pairs(V') + LRC V/V'; dE = E' - E, with E the box's running
//   energy (start energy plus accepted move dE); accept iff
//   exp(N ln(V'/V) - (dE + P (V' - V)) / kT) >= u2 (SimulationStep.cpp:
//   476-488), and then the running energy becomes E' exactly.  A proposal
//   V' <= 0 is rejected.
int queue_volume_moves(mcgpu_batch* b, const std::vector<int>& pend, hipStream_t st) {
  int rc;
  const int K = (int)pend.size();
  for (int k = 0; k < K; k++) b->h_items[k] = VolItem{pend[k], 0, 1.0};
  MCGPU_HIP_TRY(hipMemcpyAsync(b->d_items, b->h_items, sizeof(VolItem) * K,
                               hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_vol_begin, dim3((K + 63) / 64), dim3(64), 0, st, b->d_table, b->d_ctl,
                     b->d_items, K);
  int max_mol = 0, max_cells = 0, scale_blocks = 0;
  for (int s : pend) {
    const BoxView& v = b->h_table[s].bv;
    max_mol = std::max(max_mol, v.n_mol);
    max_cells = std::max(max_cells, b->slots[s].lay.ncell);
    // k_scale_multi and k_restore_multi: whole molecules per block for
    // uniform boxes
    const int per = v.uni_len > 0 && v.single_primary ? 256 / v.uni_len : 256;
    scale_blocks = std::max(scale_blocks, (v.n_mol + per - 1) / per);
  }
  // blocks for the occupancies too (at most the allocated cells)
  const int clear_blocks = (max_cells + 255) / 256;
  const int rescale_grid = std::max(scale_blocks, clear_blocks);
  hipLaunchKernelGGL(k_scale_multi, dim3(rescale_grid, K), dim3(256), 0, st, b->d_table,
                     b->d_items);
  MCGPU_HIP_TRY(hipGetLastError());
  if ((rc = launch_cells_pass(b, b->d_items, K, max_mol, max_cells, st))) return rc;
  launch_energy_partials(b, b->d_items, K, max_mol, st);
  hipLaunchKernelGGL(k_vol_settle, dim3(K), dim3(256), 0, st, b->d_table, b->d_ctl,
                     b->d_items);
  hipLaunchKernelGGL(k_restore_multi, dim3(scale_blocks, K), dim3(256), 0, st, b->d_table,
                     b->d_items);
  MCGPU_HIP_TRY(hipGetLastError());
  return MCGPU_OK;
}

// Work-group-size tuner.  Per box class (neighbour source, pair precision,
// largest molecule, molecule count in quarter-octaves) the first box of the
// class runs one tick at each work-group size; all boxes of a tick get the
// same time budget, so the moves it made in a full tick measure its speed
// at that size.  The fastest size is kept for the class; until then other
// boxes of the class use the density estimate (choose_group).
int64_t tune_key_of(const mcgpu_box_desc& d, const mcgpu_run_params& p, int max_len) {
  const int64_t nb = (int64_t)std::lround(4.0 * std::log2((double)std::max(d.n_molecules, 1)));
  return ((int64_t)p.neighbor * 4 + p.precision) * 1000000 + (int64_t)max_len * 1000 + nb;
}

void tune_sample(mcgpu_batch* b, int s) {
  BatchSlot& S = b->slots[s];
  // a box that took its volume move first ran a shortened tick
  if (S.trial < 0 || S.late) return;
  const BoxCtl& c = b->h_ctl[s];
  const int64_t moved = c.steps_done - S.steps_at_launch;
  // only full ticks measure speed: the box ran until the time budget
  if (c.status != kRunning || moved < 1 || c.steps_done >= c.next_stop) return;
  auto& T = b->tune[S.tune_key];
  T.moves[S.trial] += (double)moved;
  T.samples[S.trial]++;
  int next = -1;
  for (int g = 0; g < kGroups; g++)
    if (T.samples[g] == 0) {
      next = g;
      break;
    }
  if (next >= 0) {
    S.trial = next;
    S.group = next;
    return;
  }
  int best = 0;
  for (int g = 1; g < kGroups; g++)
    if (T.moves[g] / T.samples[g] > T.moves[best] / T.samples[best]) best = g;
  T.best = best;
  T.trialing = false;
  S.trial = -1;
  for (int k = 0; k < b->max_boxes; k++)
    if (b->slots[k].used && b->slots[k].tune_key == S.tune_key) b->slots[k].group = best;
}

// Tick boundary: consumes a finished tick (its read-back merged for the
// boxes it ran, its volume branch included).  Boxes that reached a volume
// stop in it stay at kVolumePending until the next tick takes their volume
// move.  Without a tick in flight it does nothing.
int batch_boundary(mcgpu_batch* b) {
  if (!b->pending) return MCGPU_OK;
  const double t0 = host_now();
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  b->t_wait += host_now() - t0;
  b->pending = false;
  for (int s : b->launched) {
    b->h_ctl[s] = b->h_ctl_rb[s];
    b->h_rng[s] = b->h_rng_rb[s];
    b->slots[s].dirty = false;
    if (b->budget_ticks > 0) tune_sample(b, s);
    b->slots[s].late = false;
  }
  if (b->timing_pending) {
    float ms = 0.f;
    MCGPU_HIP_TRY(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    b->kernel_ms_sum += ms;
    b->kernel_launches++;
    b->timing_pending = false;
  }
  b->launched.clear();
  b->late.clear();
  return MCGPU_OK;
}

// Whether the tick in flight runs slot s.
bool in_flight(const mcgpu_batch* b, int s) {
  return b->pending && std::find(b->launched.begin(), b->launched.end(), s) != b->launched.end();
}

// Control blocks of slots set up since their last tick, from the device.
int refresh_dirty(mcgpu_batch* b, int only_slot) {
  bool any = false;
  for (int s = 0; s < b->max_boxes; s++)
    if (b->slots[s].used && b->slots[s].dirty && (only_slot < 0 || s == only_slot)) {
      if (!any) MCGPU_HIP_TRY(hipStreamSynchronize(b->sstream));
      any = true;
      MCGPU_HIP_TRY(hipMemcpy(b->h_ctl + s, b->d_ctl + s, sizeof(BoxCtl),
                              hipMemcpyDeviceToHost));
      b->slots[s].dirty = false;
    }
  return MCGPU_OK;
}

// Everything queued is done and every mirror is current.
int batch_quiesce(mcgpu_batch* b) {
  int rc;
  if ((rc = batch_boundary(b))) return rc;
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  MCGPU_HIP_TRY(hipStreamSynchronize(b->vstream));
  MCGPU_HIP_TRY(hipStreamSynchronize(b->sstream));
  // nothing in flight: the mirrors take the device's state for every box
  MCGPU_HIP_TRY(hipMemcpy(b->h_ctl, b->d_ctl, sizeof(BoxCtl) * b->max_boxes,
                          hipMemcpyDeviceToHost));
  MCGPU_HIP_TRY(hipMemcpy(b->h_rng, b->d_rng, sizeof(Rng31) * b->max_boxes,
                          hipMemcpyDeviceToHost));
  for (BatchSlot& S : b->slots) S.dirty = false;
  return MCGPU_OK;
}

void free_slot(mcgpu_batch* b, int s) {
  if (!b->slots[s].used) return;
  if (b->slots[s].trial >= 0) {
    auto it = b->tune.find(b->slots[s].tune_key);
    if (it != b->tune.end()) it->second.trialing = false;
  }
  b->arena.release(b->slots[s].off);
  b->slots[s] = BatchSlot();
}

}  // namespace

extern "C" int mcgpu_batch_create(int32_t max_boxes, size_t budget,
                                  mcgpu_batch** out) {
  if (!out || max_boxes < 1) {
    set_error("mcgpu_batch_create: bad argument");
    return MCGPU_ERR_INVALID;
  }
  *out = nullptr;
  int rc = ensure_device();
  if (rc) return rc;
  auto* b = new (std::nothrow) mcgpu_batch();
  if (!b) return MCGPU_ERR_NOMEM;
  b->max_boxes = max_boxes;
  if (hipGetDevice(&b->device) != hipSuccess) b->device = 0;
  b->slots.resize(max_boxes);
  if (budget == 0) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = size_t(8) << 30;
    budget = std::min(fr / 4, size_t(64) << 30);
  }
  auto fail = [&](int code) {
    mcgpu_batch_destroy(b);
    return code;
  };
  if ((rc = b->arena.init(budget))) return fail(rc);
  {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device) ==
            hipSuccess &&
        cus > 0)
      b->n_cu = cus;
    const void* kern[kGroups] = {(const void*)k_batch<256, false>,
                                 (const void*)k_batch<512, false>,
                                 (const void*)k_batch<1024, false>};
    for (int g = 0; g < kGroups; g++) {
      int per_cu = 0;
      if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern[g], kGroupNT[g], 0) ==
              hipSuccess &&
          per_cu > 0)
        b->cu_share[g] = 1.0 / per_cu;
    }
  }
  int dev = 0, khz = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) == hipSuccess &&
      khz > 0)
    b->clock_khz = khz;
  const size_t M = max_boxes;
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&b->sstream, hipStreamNonBlocking) == hipSuccess &&
            hipStreamCreateWithFlags(&b->vstream, hipStreamNonBlocking) == hipSuccess &&
            hipEventCreate(&b->ev0) == hipSuccess && hipEventCreate(&b->ev1) == hipSuccess &&
            hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_ready, hipEventDisableTiming) == hipSuccess &&
            hipEventCreateWithFlags(&b->ev_vdone, hipEventDisableTiming) == hipSuccess;
  for (int g = 0; ok && g < kGroups - 1; g++)
    ok = hipStreamCreateWithFlags(&b->side[g], hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&b->join[g], hipEventDisableTiming) == hipSuccess;
  ok = ok && hipMalloc(&b->d_table, sizeof(BatchBox) * M) == hipSuccess &&
       hipMalloc(&b->d_ctl, sizeof(BoxCtl) * M) == hipSuccess &&
       hipMalloc(&b->d_rng, sizeof(Rng31) * M) == hipSuccess &&
       hipMalloc(&b->d_active, sizeof(int32_t) * 2 * kGroups * (M + 1)) == hipSuccess &&
       hipMalloc(&b->d_items, sizeof(VolItem) * M) == hipSuccess &&
       hipMalloc(&b->d_one, sizeof(VolItem) * M) == hipSuccess &&
       hipMalloc(&b->d_t0, sizeof(long long)) == hipSuccess &&
       hipHostMalloc((void**)&b->h_table, sizeof(BatchBox) * M, 0) == hipSuccess &&
       hipHostMalloc((void**)&b->h_ctl, sizeof(BoxCtl) * M, 0) == hipSuccess &&
       hipHostMalloc((void**)&b->h_ctl_rb, sizeof(BoxCtl) * M, 0) == hipSuccess &&
       hipHostMalloc((void**)&b->h_rng, sizeof(Rng31) * M, 0) == hipSuccess &&
       hipHostMalloc((void**)&b->h_rng_rb, sizeof(Rng31) * M, 0) == hipSuccess &&
       hipHostMalloc((void**)&b->h_active, sizeof(int32_t) * 2 * kGroups * (M + 1), 0) ==
           hipSuccess &&
       hipHostMalloc((void**)&b->h_items, sizeof(VolItem) * M, 0) == hipSuccess;
  if (!ok) {
    set_error("mcgpu_batch_create: allocation failed");
    return fail(MCGPU_ERR_NOMEM);
  }
  std::memset(b->h_table, 0, sizeof(BatchBox) * M);
  std::memset(b->h_ctl, 0, sizeof(BoxCtl) * M);
  // empty active lists on both sides: a tick uploads a group's list only
  // when it differs from what the device holds (h_active mirrors d_active)
  std::memset(b->h_active, 0, sizeof(int32_t) * 2 * kGroups * (M + 1));
  if (hipMemset(b->d_active, 0, sizeof(int32_t) * 2 * kGroups * (M + 1)) != hipSuccess) {
    set_error("mcgpu_batch_create: allocation failed");
    return fail(MCGPU_ERR_HIP);
  }
  std::memset(b->h_rng, 0, sizeof(Rng31) * M);
  for (size_t s = 0; s < M; s++) b->h_ctl[s].status = kDone;
  {
    std::vector<VolItem> one(M);
    for (size_t s = 0; s < M; s++) one[s] = VolItem{(int32_t)s, 0, 1.0};
    ok = hipMemcpy(b->d_one, one.data(), sizeof(VolItem) * M, hipMemcpyHostToDevice) ==
         hipSuccess;
  }
  if (!ok ||
      hipMemcpy(b->d_ctl, b->h_ctl, sizeof(BoxCtl) * M, hipMemcpyHostToDevice) != hipSuccess) {
    set_error("mcgpu_batch_create: upload failed");
    return fail(MCGPU_ERR_HIP);
  }
  *out = b;
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_destroy(mcgpu_batch* b) {
  if (!b) return MCGPU_OK;
  for (hipStream_t s : {b->stream, b->vstream, b->side[0], b->side[1], b->sstream})
    if (s) (void)hipStreamSynchronize(s);
  for (TickGraph* G : {&b->main_graph, &b->late_graph}) {
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
  }
  for (void* p : {(void*)b->d_table, (void*)b->d_ctl, (void*)b->d_rng, (void*)b->d_active,
                  (void*)b->d_items, (void*)b->d_one, (void*)b->d_t0})
    if (p) (void)hipFree(p);
  for (void* p : {(void*)b->h_table, (void*)b->h_ctl, (void*)b->h_ctl_rb, (void*)b->h_rng,
                  (void*)b->h_rng_rb, (void*)b->h_active, (void*)b->h_items})
    if (p) (void)hipHostFree(p);
  for (hipEvent_t e : {b->ev0, b->ev1, b->fork, b->ev_ready, b->ev_vdone})
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : b->join)
    if (e) (void)hipEventDestroy(e);
  for (hipStream_t s : {b->stream, b->vstream, b->side[0], b->side[1], b->sstream})
    if (s) (void)hipStreamDestroy(s);
  delete b;
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_bind_thread(mcgpu_batch* b) {
  if (!b) return MCGPU_ERR_INVALID;
  MCGPU_HIP_TRY(hipSetDevice(b->device));
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_set_tick_budget(mcgpu_batch* b, double microseconds) {
  if (!b || microseconds < 0) return MCGPU_ERR_INVALID;
  b->budget_ticks = (long long)(microseconds * b->clock_khz / 1000.0);
  b->version++;
  return MCGPU_OK;
}

extern "C" int mcgpu_box_footprint(const mcgpu_box_desc* d,
                                   const mcgpu_run_params* p, size_t* bytes) {
  if (!d || !p || !bytes) return MCGPU_ERR_INVALID;
  const std::string why = validate_desc(*d);
  if (!why.empty()) {
    set_error("mcgpu_box_footprint: " + why);
    return MCGPU_ERR_INVALID;
  }
  const ParamTable pt = build_param_table(*d);
  *bytes = plan_layout(*d, *p, pt).bytes;
  return MCGPU_OK;
}

// Admission: the box's arrays are uploaded and its cells, start energy and
// control block set up on the set-up stream, beside a tick in flight (a
// tick only touches the boxes it runs); the box joins the next tick.  With
// a resume record r the box continues that record's generator stream,
// running energy, LRC and NPT schedule.
static int batch_add_impl(mcgpu_batch* b, const mcgpu_box_desc* d, const mcgpu_run_params* p,
                          const mcgpu_resume_state* r, int32_t* slot_out) {
  if (!b || !d || !p || !slot_out) return MCGPU_ERR_INVALID;
  if (r && (r->rng_front < 0 || r->rng_front > 30 || r->rng_rear < 0 || r->rng_rear > 30 ||
            (r->rng_front - r->rng_rear + 31) % 31 != 3 || r->steps_done < 0)) {
    set_error("mcgpu_batch_add_resumed: bad resume record");
    return MCGPU_ERR_INVALID;
  }
  const std::string why = validate_desc(*d);
  if (!why.empty()) {
    set_error("mcgpu_batch_add: " + why);
    return MCGPU_ERR_INVALID;
  }
  if ((p->precision != MCGPU_FP64 && p->precision != MCGPU_MIXED) ||
      (p->ensemble != MCGPU_NVT && p->ensemble != MCGPU_NPT) || p->steps < 0 ||
      (p->neighbor != MCGPU_NEIGHBOR_BRUTE && p->neighbor != MCGPU_NEIGHBOR_CELL) ||
      (p->ensemble == MCGPU_NPT && (p->volume_interval < 1 || !(p->max_volume_delta > 0) ||
                                    !(p->max_volume_delta < 0.5)))) {
    set_error("mcgpu_batch_add: bad run parameters");
    return MCGPU_ERR_INVALID;
  }
  int rc;
  int s = -1;
  for (int k = 0; k < b->max_boxes; k++)
    if (!b->slots[k].used) {
      s = k;
      break;
    }
  if (s < 0) {
    set_error("mcgpu_batch_add: no free slot");
    return MCGPU_ERR_NOMEM;
  }
  const ParamTable pt = build_param_table(*d);
  if ((int)pt.p64.size() > kParamCap) {
    set_error("mcgpu_batch_add: more than " + std::to_string(kParamCap) +
              " distinct site-parameter entries");
    return MCGPU_ERR_INVALID;
  }
  const BoxLayout L = plan_layout(*d, *p, pt);
  size_t off = 0;
  if (!b->arena.alloc(L.bytes, &off)) {
    set_error("mcgpu_batch_add: HBM budget exhausted");
    return MCGPU_ERR_NOMEM;
  }
  BatchSlot& S = b->slots[s];
  S = BatchSlot();
  S.used = true;
  S.off = off;
  S.lay = L;
  S.params = *p;
  S.n_pidx = d->n_primary;
  S.spread = pt.spread;
  S.group = choose_group(move_work(*d, p->neighbor, pt.max_len));
  S.tune_key = tune_key_of(*d, *p, pt.max_len);
  if (b->autotune && b->budget_ticks > 0) {
    auto& T = b->tune[S.tune_key];
    if (T.best >= 0) {
      S.group = T.best;
    } else if (!T.trialing) {  // this box measures the class
      T.trialing = true;
      for (int g = 0; g < kGroups; g++)
        if (T.samples[g] == 0) {
          S.trial = S.group = g;
          break;
        }
    }
  }
  char* base = b->arena.base() + off;
  BatchBox& E = b->h_table[s];
  E = BatchBox{};
  BoxView& v = E.bv;
  v.x = gmem((double*)(base + L.x));
  v.y = gmem((double*)(base + L.y));
  v.z = gmem((double*)(base + L.z));
  v.px = gmem((double*)(base + L.px));
  v.py = gmem((double*)(base + L.py));
  v.pz = gmem((double*)(base + L.pz));
  v.mol_start = gmem((int32_t*)(base + L.ms));
  v.mol_len = gmem((int32_t*)(base + L.ml));
  v.mol_pidx_start = gmem((int32_t*)(base + L.mps));
  v.mol_pidx_count = gmem((int32_t*)(base + L.mpc));
  v.mol_param = gmem((int32_t*)(base + L.mp));
  v.meta = gmem((MolMeta*)(base + L.meta));
  v.single_primary = pt.single_primary;
  v.pidx = gmem((int32_t*)(base + L.pidx));
  v.p64 = gmem((SiteParam64*)(base + L.p64));
  v.p32 = gmem((SiteParam32*)(base + L.p32));
  v.n_mol = d->n_molecules;
  v.n_site = d->n_sites;
  v.max_len = pt.max_len;
  v.uni_len = pt.uni_len;
  v.uni_param = pt.uni_param;
  v.uni_p0 = pt.uni_p0;
  v.n_param = (int32_t)pt.p64.size();
  v.mode = MCGPU_NB_BRUTE;
  for (int k = 0; k < 3; k++) v.L[k] = d->box[k];
  v.cutoff2 = d->cutoff * d->cutoff;
  v.reach2 = reach2_of(pt.reach);
  v.kT = d->kT;
  v.cells.count = gmem((int32_t*)(base + L.ccount));
  v.cells.slots = gmem((int32_t*)(base + L.cslots));
  v.cells.mol_cell = gmem((int32_t*)(base + L.mcell));
  v.cells.mol_slot = gmem((int32_t*)(base + L.mslot));
  v.cells.overflow = gmem((int32_t*)(base + L.covf));
  v.cells.cap = L.cap;
  E.cells_alt = v.cells;
  if (L.alt_cells) {
    E.cells_alt.count = gmem((int32_t*)(base + L.ccount2));
    E.cells_alt.slots = gmem((int32_t*)(base + L.cslots2));
    E.cells_alt.mol_cell = gmem((int32_t*)(base + L.mcell2));
    E.cells_alt.mol_slot = gmem((int32_t*)(base + L.mslot2));
    E.cells_alt.overflow = gmem((int32_t*)(base + L.covf2));
  }
  E.mode_alt = MCGPU_NB_BRUTE;
  v.verlet.stale = gmem((int32_t*)(base + L.vstale));
  v.verlet.count = gmem((int32_t*)(base + L.vcount));
  E.rng = b->d_rng + s;
  E.max_t = d->max_translate;
  E.max_r = d->max_rotate;
  E.precision = p->precision;
  E.record = p->record_trace ? 1 : 0;
  E.partial = (double*)(base + L.partial);
  E.total = (double*)(base + L.total);
  E.bx = (double*)(base + L.bx);
  E.by = (double*)(base + L.by);
  E.bz = (double*)(base + L.bz);
  E.pressure = p->pressure;
  // fixed-width volume step: a symmetric proposal, so the acceptance test
  // needs no Hastings factor
  E.dv = p->max_volume_delta * (d->box[0] * d->box[1] * d->box[2]);
  if (r && r->volume_step > 0) E.dv = r->volume_step;
  E.cutoff = d->cutoff;
  E.spread = pt.spread;
  E.vol_interval = p->volume_interval;
  E.ncell_cap = L.ncell;
  E.tr_mol = (int32_t*)(base + L.trm);
  E.tr_de = (double*)(base + L.trd);
  E.tr_acc = (int8_t*)(base + L.tra);
  E.tr_vol = L.trv_cap > 0 ? (mcgpu_volume_record*)(base + L.trv) : nullptr;
  E.tr_vol_cap = L.trv_cap;
  if (r) {
    for (int k = 0; k < 31; k++) b->h_rng[s].s[k] = r->rng_words[k];
    b->h_rng[s].f = r->rng_front;
    b->h_rng[s].r = r->rng_rear;
    S.base_steps = r->steps_done;
    S.base_acc = r->accepted;
    S.base_vol = r->volume_moves;
    S.base_vacc = r->volume_accepted;
  } else {
    rng31_seed(b->h_rng[s], p->seed);
  }
  const hipStream_t st = b->sstream;
  const int32_t zero[3] = {0, 0, 0};
  struct Copy {
    void* dst;
    const void* src;
    size_t bytes;
  } copies[] = {{base + L.x, d->x, 8 * (size_t)d->n_sites},
                {base + L.y, d->y, 8 * (size_t)d->n_sites},
                {base + L.z, d->z, 8 * (size_t)d->n_sites},
                {base + L.ms, d->mol_start, 4 * (size_t)d->n_molecules},
                {base + L.ml, d->mol_len, 4 * (size_t)d->n_molecules},
                {base + L.mps, d->mol_pidx_start, 4 * (size_t)d->n_molecules},
                {base + L.mpc, d->mol_pidx_count, 4 * (size_t)d->n_molecules},
                {base + L.mp, pt.mol_param.data(), 4 * (size_t)d->n_molecules},
                {base + L.meta, pt.meta.data(), sizeof(MolMeta) * (size_t)d->n_molecules},
                {base + L.pidx, d->primary_index, 4 * (size_t)d->n_primary},
                {base + L.p64, pt.p64.data(), sizeof(SiteParam64) * pt.p64.size()},
                {base + L.p32, pt.p32.data(), sizeof(SiteParam32) * pt.p32.size()},
                {base + L.covf, zero, 4},
                {base + L.vstale, zero, 4},
                {base + L.vcount, zero, 4},
                {b->d_rng + s, b->h_rng + s, sizeof(Rng31)}};
  for (const Copy& c : copies) {
    if (hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyHostToDevice, st) != hipSuccess) {
      free_slot(b, s);
      set_error("mcgpu_batch_add: upload failed");
      return MCGPU_ERR_HIP;
    }
  }
  // The sources are the caller's (and this call's) host arrays: wait for
  // the copies (the set-up stream only); the set-up kernels stay queued.
  if (hipStreamSynchronize(st) != hipSuccess) {
    free_slot(b, s);
    set_error("mcgpu_batch_add: upload failed");
    return MCGPU_ERR_HIP;
  }
  int ncell = 0;
  if (p->neighbor == MCGPU_NEIGHBOR_CELL) ncell = slot_set_geometry(b, s);
  if ((rc = upload_slot(b, s, st)) || (rc = launch_refresh_primary(v, S.n_pidx, st)) ||
      (ncell > 0 && (rc = launch_cells_build(v, st)))) {
    free_slot(b, s);
    return rc;
  }
  // start energy: the full-energy pass over this one box
  if (p->precision == MCGPU_MIXED)
    hipLaunchKernelGGL(k_energy_waves<true>, dim3((v.n_mol + 3) / 4, 1),
                       dim3(256), 0, st, b->d_table, b->d_one + s);
  else
    hipLaunchKernelGGL(k_energy_waves<false>, dim3((v.n_mol + 3) / 4, 1),
                       dim3(256), 0, st, b->d_table, b->d_one + s);
  hipLaunchKernelGGL(k_reduce_multi, dim3(1), dim3(256), 0, st, b->d_table, b->d_one + s, 0,
                     E.total + 2);
  MCGPU_HIP_TRY(hipGetLastError());
  BoxCtl& c = b->h_ctl[s];
  c = BoxCtl{};
  for (int k = 0; k < 3; k++) c.L[k] = d->box[k];
  c.lrc = r ? r->lrc : p->energy_offset;
  c.steps_target = p->steps;
  c.next_stop = p->steps;
  if (p->ensemble == MCGPU_NPT) {
    // volume moves after every volume_interval moves since the first start;
    // one due where a previous run stopped comes first
    const int64_t prior = r ? r->steps_done : 0;
    const int64_t due = (r && r->volume_pending) ? 0 : p->volume_interval - prior % p->volume_interval;
    c.next_stop = std::min<int64_t>(p->steps, due);
  }
  c.status = p->steps > 0 ? kRunning : kDone;
  if ((rc = upload_ctl(b, s, st))) {
    free_slot(b, s);
    return rc;
  }
  hipLaunchKernelGGL(k_set_start_energy, dim3(1), dim3(64), 0, st, b->d_ctl + s,
                     (const double*)E.total, p->energy_offset, r ? 1 : 0,
                     r ? r->energy : 0.0);
  MCGPU_HIP_TRY(hipGetLastError());
  S.dirty = true;  // the device now holds the start energy
  *slot_out = s;
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_add(mcgpu_batch* b, const mcgpu_box_desc* d,
                               const mcgpu_run_params* p, int32_t* slot_out) {
  return batch_add_impl(b, d, p, nullptr, slot_out);
}

extern "C" int mcgpu_batch_add_resumed(mcgpu_batch* b, const mcgpu_box_desc* d,
                                       const mcgpu_run_params* p, const mcgpu_resume_state* r,
                                       int32_t* slot_out) {
  if (!r) return MCGPU_ERR_INVALID;
  return batch_add_impl(b, d, p, r, slot_out);
}

extern "C" int mcgpu_batch_get_resume(mcgpu_batch* b, int32_t s, mcgpu_resume_state* out) {
  if (!b || !out || s < 0 || s >= b->max_boxes || !b->slots[s].used) {
    set_error("mcgpu_batch_get_resume: bad slot");
    return MCGPU_ERR_RANGE;
  }
  int rc = batch_quiesce(b);
  if (rc) return rc;
  const BatchSlot& S = b->slots[s];
  const BoxCtl& c = b->h_ctl[s];
  const Rng31& g = b->h_rng[s];
  *out = mcgpu_resume_state{};
  for (int k = 0; k < 31; k++) out->rng_words[k] = g.s[k];
  out->rng_front = g.f;
  out->rng_rear = g.r;
  out->steps_done = S.base_steps + c.steps_done;
  out->accepted = S.base_acc + c.accepted;
  out->volume_moves = S.base_vol + c.vol_moves;
  out->volume_accepted = S.base_vacc + c.vol_acc;
  out->energy = c.energy;
  out->lrc = c.lrc;
  out->ensemble = S.params.ensemble;
  out->volume_interval = S.params.volume_interval;
  out->pressure = S.params.pressure;
  const int vi = S.params.volume_interval;
  if (S.params.ensemble == MCGPU_NPT && vi > 0) {
    out->volume_step = b->h_table[s].dv;
    out->volume_pending = out->steps_done > 0 && out->steps_done % vi == 0 &&
                          out->volume_moves < out->steps_done / vi;
  }
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_remove(mcgpu_batch* b, int32_t s) {
  if (!b || s < 0 || s >= b->max_boxes || !b->slots[s].used) {
    set_error("mcgpu_batch_remove: bad slot");
    return MCGPU_ERR_RANGE;
  }
  // a tick in flight without this box touches neither its control block nor
  // its arena block (its read-back copies the block but merges launched
  // slots only), so only a tick that runs it is waited for
  int rc;
  if (in_flight(b, s) && (rc = batch_boundary(b))) return rc;
  if (b->slots[s].dirty) MCGPU_HIP_TRY(hipStreamSynchronize(b->sstream));
  b->h_ctl[s] = BoxCtl{};
  b->h_ctl[s].status = kDone;
  if ((rc = upload_ctl(b, s, b->sstream))) return rc;
  free_slot(b, s);
  return MCGPU_OK;
}

namespace {

// Replays G on `st` (the late graph on the volume stream, after the volume
// pass), capturing it again when the grids, sweep count or time budget
// changed; e0 and e1 (optional) are recorded around the replay.  Every group
// is launched, with its list's count rounded up to a multiple of 32
// work-groups (32 at least; surplus work-groups exit at once), and a grid is
// kept while it is at most twice what the list needs, so a graph survives
// most admissions, retirements and volume stops: capturing and instantiating
// one takes the host a few hundred microseconds, during which the device
// would idle.
int launch_tick_graph(mcgpu_batch* b, TickGraph& G, bool late, hipStream_t st,
                      const int n[kGroups], int sweeps, hipEvent_t e0, hipEvent_t e1) {
  const int stride = b->max_boxes + 1;
  const int32_t* act = b->d_active + (late ? kGroups * stride : 0);
  int grid[kGroups];
  bool same = G.exec && G.version == b->version && G.sweeps == sweeps;
  for (int g = 0; g < kGroups; g++) {
    const int need = std::min(b->max_boxes, (std::max(n[g], 1) + 31) & ~31);
    const int cur = G.exec ? G.grid[g] : 0;
    grid[g] = cur >= need && cur <= 2 * need ? cur : need;
    same = same && G.grid[g] == grid[g];
  }
  if (!same) {
    if (G.exec) MCGPU_HIP_TRY(hipGraphExecDestroy(G.exec));
    if (G.graph) MCGPU_HIP_TRY(hipGraphDestroy(G.graph));
    G.exec = nullptr;
    G.graph = nullptr;
    using K = void (*)(const BatchBox*, BoxCtl*, const int32_t*, int, long long,
                       const long long*);
    const K kern[2][kGroups] = {
        {k_batch<256, false>, k_batch<512, false>, k_batch<1024, false>},
        {k_batch<256, true>, k_batch<512, true>, k_batch<1024, true>}};
    const hipStream_t on[kGroups] = {st, b->side[0], b->side[1]};
    MCGPU_HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    MCGPU_HIP_TRY(hipEventRecord(b->fork, st));
    for (int g = 1; g < kGroups; g++)
      MCGPU_HIP_TRY(hipStreamWaitEvent(b->side[g - 1], b->fork, 0));
    for (int g = 0; g < kGroups; g++)
      hipLaunchKernelGGL(kern[late][g], dim3(grid[g]), dim3(kGroupNT[g]), 0, on[g],
                         b->d_table, b->d_ctl, act + g * stride, sweeps, b->budget_ticks,
                         (const long long*)b->d_t0);
    for (int g = 1; g < kGroups; g++) {
      MCGPU_HIP_TRY(hipEventRecord(b->join[g - 1], b->side[g - 1]));
      MCGPU_HIP_TRY(hipStreamWaitEvent(st, b->join[g - 1], 0));
    }
    MCGPU_HIP_TRY(hipStreamEndCapture(st, &G.graph));
    MCGPU_HIP_TRY(hipGraphInstantiate(&G.exec, G.graph, nullptr, nullptr, 0));
    G.version = b->version;
    G.sweeps = sweeps;
    for (int g = 0; g < kGroups; g++) G.grid[g] = grid[g];
  }
  if (e0) MCGPU_HIP_TRY(hipEventRecord(e0, st));
  MCGPU_HIP_TRY(hipGraphLaunch(G.exec, st));
  if (e1) MCGPU_HIP_TRY(hipEventRecord(e1, st));
  return MCGPU_OK;
}

// Uploads the active lists of one side (main or late) that changed.  The
// previous tick, and so its uploads from h_active, has completed (the
// boundary waited for it).
int upload_lists(mcgpu_batch* b, bool late, hipStream_t st) {
  const int stride = b->max_boxes + 1;
  const size_t base = late ? (size_t)kGroups * stride : 0;
  for (int g = 0; g < kGroups; g++) {
    const int32_t* want = &b->next_active[base + g * stride];
    const size_t words = 1 + (size_t)want[0];
    int32_t* h = b->h_active + base + g * stride;
    if (std::memcmp(h, want, sizeof(int32_t) * words) == 0) continue;
    std::memcpy(h, want, sizeof(int32_t) * words);
    MCGPU_HIP_TRY(hipMemcpyAsync(b->d_active + base + g * stride, h, sizeof(int32_t) * words,
                                 hipMemcpyHostToDevice, st));
  }
  return MCGPU_OK;
}

}  // namespace

// One tick: every running box advances up to `sweeps` moves in one graph
// launch; beside it, on the volume stream, the boxes that reached a volume
// stop in the previous tick take their volume move and then advance in the
// late graph until the same deadline.  The previous tick is consumed first;
// with sync the call returns after this tick is consumed too.
extern "C" int mcgpu_batch_run(mcgpu_batch* b, int32_t sweeps, int32_t sync) {
  if (!b || sweeps < 1) return MCGPU_ERR_INVALID;
  int rc = batch_boundary(b);
  if (rc) return rc;
  const int stride = b->max_boxes + 1;
  b->launched.clear();
  b->late.clear();
  for (int s = 0; s < b->max_boxes; s++) {
    const BatchSlot& S = b->slots[s];
    const int32_t st = b->h_ctl[s].status;
    if (!S.used || (st != kRunning && st != kVolumePending)) continue;
    b->launched.push_back(s);
    b->slots[s].steps_at_launch = b->h_ctl[s].steps_done;
    if (st == kVolumePending) {
      b->late.push_back(s);
      b->slots[s].late = true;
    }
  }
  if (b->launched.empty()) return MCGPU_OK;
  // Work-group sizes of this tick: each box's preferred one, unless the
  // boxes need more compute units than the device has.  Then every box
  // takes 256 lanes, whose work-groups share CUs, so the whole tick is
  // resident at once: with 512- and 1024-lane groups oversubscribed, groups
  // wait for whole CUs and run in rounds (512 water boxes of 64-4,096
  // molecules: 1.9e7 moves/s with the preferred sizes, 2.3e7 all at 256
  // lanes; demoting only the lightest boxes until the CUs sufficed gave
  // 1.8e7-2.4e7, as the three group kernels then fragment the CUs).  Boxes
  // measured by the tuner keep their size.
  std::vector<int>& eff = b->eff_group;
  eff.assign(b->max_boxes, 0);
  double demand = 0.0;
  for (int s : b->launched) {
    eff[s] = b->slots[s].group;
    demand += b->cu_share[eff[s]];
  }
  if (demand > b->n_cu)
    for (int s : b->launched)
      if (b->slots[s].trial < 0) eff[s] = 0;
  std::vector<int32_t>& act = b->next_active;
  act.assign((size_t)2 * kGroups * stride, 0);
  int n[2][kGroups] = {};
  for (int s : b->launched) {
    const int side = b->slots[s].late ? 1 : 0;
    const size_t at = (size_t)(side * kGroups + eff[s]) * stride;
    act[at + 1 + n[side][eff[s]]++] = s;
  }
  for (int side = 0; side < 2; side++)
    for (int g = 0; g < kGroups; g++) act[(size_t)(side * kGroups + g) * stride] = n[side][g];
  const double tl = host_now();
  // set-up work queued so far (admissions) first
  MCGPU_HIP_TRY(hipEventRecord(b->ev_ready, b->sstream));
  const bool any_main = b->late.size() < b->launched.size();
  const bool any_late = !b->late.empty();
  double tv = 0.0;
  // Order: the volume branch's clock stamp (the late boxes' tick start), the
  // main graph, then the rest of the branch, so that both sides start
  // together and the main graph waits for as little host work as possible.
  if (any_late) {
    MCGPU_HIP_TRY(hipStreamWaitEvent(b->vstream, b->ev_ready, 0));
    hipLaunchKernelGGL(k_stamp_clock, dim3(1), dim3(1), 0, b->vstream, b->d_t0);
    MCGPU_HIP_TRY(hipGetLastError());
  }
  MCGPU_HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_ready, 0));
  if (any_main &&
      ((rc = upload_lists(b, false, b->stream)) ||
       (rc = launch_tick_graph(b, b->main_graph, false, b->stream, n[0], sweeps, b->ev0,
                               b->ev1))))
    return rc;
  if (any_late) {
    if ((rc = upload_lists(b, true, b->vstream))) return rc;
    const double t1 = host_now();
    if ((rc = queue_volume_moves(b, b->late, b->vstream))) return rc;
    tv = host_now() - t1;
    b->t_volume += tv;
    if ((rc = launch_tick_graph(b, b->late_graph, true, b->vstream, n[1], sweeps, nullptr,
                                nullptr)))
      return rc;
    MCGPU_HIP_TRY(hipEventRecord(b->ev_vdone, b->vstream));
    MCGPU_HIP_TRY(hipStreamWaitEvent(b->stream, b->ev_vdone, 0));
  }
  // read-back of the launched boxes' control blocks and generators into
  // staging (slots up to the highest launched one; the boundary merges the
  // launched slots only)
  const size_t nrb = (size_t)b->launched.back() + 1;
  MCGPU_HIP_TRY(hipMemcpyAsync(b->h_ctl_rb, b->d_ctl, sizeof(BoxCtl) * nrb,
                               hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(b->h_rng_rb, b->d_rng, sizeof(Rng31) * nrb,
                               hipMemcpyDeviceToHost, b->stream));
  b->pending = true;
  b->timing_pending = any_main;
  b->t_launch += host_now() - tl - tv;
  if (sync) return batch_boundary(b);
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_host_seconds(mcgpu_batch* b, double* wait, double* volume,
                                        double* launch) {
  if (!b) return MCGPU_ERR_INVALID;
  if (wait) *wait = b->t_wait;
  if (volume) *volume = b->t_volume;
  if (launch) *launch = b->t_launch;
  return MCGPU_OK;
}

namespace {

// A slot's statistics from its host mirror.
void stats_from_mirror(const mcgpu_batch* b, int s, mcgpu_box_stats* out) {
  const BoxCtl& c = b->h_ctl[s];
  out->steps_done = c.steps_done;
  out->accepted = c.accepted;
  out->volume_moves = c.vol_moves;
  out->volume_accepted = c.vol_acc;
  out->energy = c.energy;
  out->start_energy = c.start_energy;
  out->work_candidates = c.work_cand;
  out->work_sites = c.work_sites + c.work_self;
  out->neighbor_fallbacks = c.fallbacks;
  for (int k = 0; k < 3; k++) out->box[k] = c.L[k];
  out->done = c.status == kDone;
  out->fault = c.fault ? c.fault : b->slots[s].fault;
}

}  // namespace

extern "C" int mcgpu_batch_stats(mcgpu_batch* b, int32_t s, mcgpu_box_stats* out) {
  if (!b || !out || s < 0 || s >= b->max_boxes || !b->slots[s].used) {
    set_error("mcgpu_batch_stats: bad slot");
    return MCGPU_ERR_RANGE;
  }
  int rc = batch_boundary(b);
  if (rc || (rc = refresh_dirty(b, s))) return rc;
  stats_from_mirror(b, s, out);
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_peek(mcgpu_batch* b, int32_t s, mcgpu_box_stats* out) {
  if (!b || !out || s < 0 || s >= b->max_boxes || !b->slots[s].used) {
    set_error("mcgpu_batch_peek: bad slot");
    return MCGPU_ERR_RANGE;
  }
  if (b->slots[s].dirty) {
    if (!in_flight(b, s)) {
      int rc = refresh_dirty(b, s);
      if (rc) return rc;
    } else {
      // admitted just before the tick in flight, which writes its control
      // block: only the start energy is read (after the set-up stream, which
      // computes it; the tick stores it back unchanged), and the box is
      // reported as before its first move
      MCGPU_HIP_TRY(hipStreamSynchronize(b->sstream));
      MCGPU_HIP_TRY(hipMemcpy(&b->h_ctl[s].start_energy, &b->d_ctl[s].start_energy,
                              sizeof(double), hipMemcpyDeviceToHost));
      b->h_ctl[s].energy = b->h_ctl[s].start_energy;
    }
  }
  stats_from_mirror(b, s, out);
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_download(mcgpu_batch* b, int32_t s, double* x,
                                    double* y, double* z) {
  if (!b || s < 0 || s >= b->max_boxes || !b->slots[s].used || !x || !y || !z) {
    set_error("mcgpu_batch_download: bad argument");
    return MCGPU_ERR_RANGE;
  }
  int rc = batch_quiesce(b);
  if (rc) return rc;
  const BoxView& v = b->h_table[s].bv;
  const size_t nb = 8 * (size_t)v.n_site;
  MCGPU_HIP_TRY(hipMemcpyAsync(x, v.x, nb, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(y, v.y, nb, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(z, v.z, nb, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_trace(mcgpu_batch* b, int32_t s, int32_t* mol,
                                 double* delta, int8_t* acc, int64_t cap,
                                 int64_t* n) {
  if (!b || s < 0 || s >= b->max_boxes || !b->slots[s].used || !n) {
    set_error("mcgpu_batch_trace: bad argument");
    return MCGPU_ERR_RANGE;
  }
  int rc = batch_quiesce(b);
  if (rc) return rc;
  const BatchBox& E = b->h_table[s];
  if (!E.record) {
    set_error("mcgpu_batch_trace: box was added without record_trace");
    return MCGPU_ERR_STATE;
  }
  const int64_t k = std::min<int64_t>(cap, b->h_ctl[s].steps_done);
  if (k > 0) {
    if (mol) MCGPU_HIP_TRY(hipMemcpyAsync(mol, E.tr_mol, 4 * k, hipMemcpyDeviceToHost, b->stream));
    if (delta) MCGPU_HIP_TRY(hipMemcpyAsync(delta, E.tr_de, 8 * k, hipMemcpyDeviceToHost, b->stream));
    if (acc) MCGPU_HIP_TRY(hipMemcpyAsync(acc, E.tr_acc, k, hipMemcpyDeviceToHost, b->stream));
    MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  }
  *n = k;
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_volume_trace(mcgpu_batch* b, int32_t s,
                                        mcgpu_volume_record* out, int64_t cap,
                                        int64_t* n) {
  if (!b || s < 0 || s >= b->max_boxes || !b->slots[s].used || !n) {
    set_error("mcgpu_batch_volume_trace: bad argument");
    return MCGPU_ERR_RANGE;
  }
  int rc = batch_quiesce(b);
  if (rc) return rc;
  const BatchBox& E = b->h_table[s];
  if (!E.tr_vol) {
    set_error("mcgpu_batch_volume_trace: box is not NPT or was added without record_trace");
    return MCGPU_ERR_STATE;
  }
  const int64_t k = std::min<int64_t>(std::min<int64_t>(cap, E.tr_vol_cap),
                                      b->h_ctl[s].vol_moves);
  if (k > 0 && out) {
    MCGPU_HIP_TRY(hipMemcpyAsync(out, E.tr_vol, sizeof(mcgpu_volume_record) * k,
                                 hipMemcpyDeviceToHost, b->stream));
    MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  }
  *n = k;
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_kernel_ms(mcgpu_batch* b, double* avg_ms,
                                     int64_t* launches) {
  if (!b || !avg_ms || !launches) return MCGPU_ERR_INVALID;
  int rc = batch_boundary(b);
  if (rc) return rc;
  *launches = b->kernel_launches;
  *avg_ms = b->kernel_launches ? b->kernel_ms_sum / b->kernel_launches : 0.0;
  b->kernel_ms_sum = 0.0;
  b->kernel_launches = 0;
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_memory(mcgpu_batch* b, size_t* used, size_t* peak,
                                  size_t* capacity) {
  if (!b) return MCGPU_ERR_INVALID;
  if (used) *used = b->arena.used();
  if (peak) *peak = b->arena.peak();
  if (capacity) *capacity = b->arena.capacity();
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_set_autotune(mcgpu_batch* b, int32_t on) {
  if (!b) return MCGPU_ERR_INVALID;
  b->autotune = on != 0;
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_tuning(mcgpu_batch* b, int64_t* keys, int32_t* threads,
                                  double* moves_per_tick, int32_t cap, int32_t* n) {
  if (!b || !n) return MCGPU_ERR_INVALID;
  int k = 0;
  for (const auto& kv : b->tune) {
    if (kv.second.best < 0) continue;
    if (k < cap) {
      if (keys) keys[k] = kv.first;
      if (threads) threads[k] = kGroupNT[kv.second.best];
      for (int g = 0; moves_per_tick && g < kGroups; g++)
        moves_per_tick[kGroups * k + g] =
            kv.second.samples[g] ? kv.second.moves[g] / kv.second.samples[g] : 0.0;
    }
    k++;
  }
  *n = k;
  return MCGPU_OK;
}

extern "C" int mcgpu_batch_group(mcgpu_batch* b, int32_t s, int32_t* threads) {
  if (!b || !threads || s < 0 || s >= b->max_boxes || !b->slots[s].used)
    return MCGPU_ERR_RANGE;
  *threads = kGroupNT[b->slots[s].group];
  return MCGPU_OK;
}
