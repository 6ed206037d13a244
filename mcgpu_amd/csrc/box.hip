// Single device-resident box: residency (replaces GPUCopy, GPUCopy.cpp:97-222),
// the molecule-energy and fused trial-move kernels (replace the BruteForceStep
// OpenACC region, BruteForceStep.cpp:31-76, and the move/rollback regions,
// SimulationStep.cpp:545-565,738-744), the full-system energy
// (SimulationStep.cpp:40-71), neighbour rebuilds (NLCStep.cpp:125-253,
// VerletStep.cpp:119-207, done correctly) and the NPT box rescale.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "box_internal.h"
#include "cells.h"
#include "eval.h"
#include "fastmath.h"
#include "fullenergy.h"

namespace mcgpu {

// ------------------------------------------------------------- kernels

struct MoveArgs {
  int32_t mol;
  int32_t start_mol;
  int32_t has_new;
  int32_t metropolis;  // apply the Metropolis test with u and commit
  double u;
  double nxyz[3 * kMaxSites];  // x block, y block, z block
};

template <int NT, bool MIXED>
__global__ __launch_bounds__(NT) void k_move(BoxView bv, MoveArgs a,
                                             MoveOut* out) {
  __shared__ EvalShared<NT> sh;
  const int t = threadIdx.x;
  load_params<NT>(bv, sh);
  load_molecule<NT>(bv, a.mol, sh);
  const int len = bv.mol_len[a.mol];
  if (a.has_new) {
    for (int k = t; k < len; k += NT) {
      sh.nx[k] = a.nxyz[k];
      sh.ny[k] = a.nxyz[len + k];
      sh.nz[k] = a.nxyz[2 * len + k];
    }
  }
  __syncthreads();
  double eo, en;
  eval_molecule<NT, MIXED>(bv, a.mol, a.start_mol, a.has_new != 0, sh, eo, en,
                           nullptr);
  if (t == 0) {
    const double de = en - eo;
    int acc = 0;
    if (a.metropolis) acc = (de < 0) || (exp_fp64(-(de) / bv.kT) >= a.u);
    out->e_old = eo;
    out->e_new = en;
    out->de = de;
    out->accepted = acc;
    out->source = sh.src;
    sh.flag = acc;
  }
  __syncthreads();
  if (a.metropolis && sh.flag) commit_molecule<NT>(bv, a.mol, sh);
  // thread 0 made the commit (and its Verlet displacement test)
  if (t == 0) out->stale = bv.mode == MCGPU_NB_VERLET ? bv.verlet.stale[0] : 0;
}

// Overwrite one molecule (changeMolecule / rollback from the host).
__global__ __launch_bounds__(64) void k_set_mol(BoxView bv, MoveArgs a, MoveOut* out) {
  __shared__ EvalShared<64> sh;
  const int t = threadIdx.x;
  load_molecule<64>(bv, a.mol, sh);
  const int len = bv.mol_len[a.mol];
  for (int k = t; k < len; k += 64) {
    sh.nx[k] = a.nxyz[k];
    sh.ny[k] = a.nxyz[len + k];
    sh.nz[k] = a.nxyz[2 * len + k];
  }
  __syncthreads();
  commit_molecule<64>(bv, a.mol, sh);
  if (t == 0) out->stale = bv.mode == MCGPU_NB_VERLET ? bv.verlet.stale[0] : 0;
}

__global__ void k_refresh_primary(BoxView bv, int n_pidx) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n_pidx) {
    const int s = bv.pidx[k];
    bv.px[k] = bv.x[s];
    bv.py[k] = bv.y[s];
    bv.pz[k] = bv.z[s];
  }
}

__global__ void k_cells_clear(CellGrid g, int ncell) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncell) g.count[c] = 0;
  if (c == 0) g.overflow[0] = 0;
}

__global__ void k_cells_assign(BoxView bv) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m < bv.n_mol) cells_assign_one(bv, m);
}

// One wave per cell (4 cells per 256-lane work-group).
__global__ __launch_bounds__(256) void k_cells_sort(CellGrid g, int ncell) {
  const int c = blockIdx.x * 4 + (int)(threadIdx.x >> 6);
  if (c < ncell) cells_sort_wave(g, c);
}

// Verlet list of molecule i: every j whose first primary lies within rv of
// i's (minimum image), in increasing j.
template <int NT>
__global__ __launch_bounds__(NT) void k_verlet_build(BoxView bv, double rv2) {
  __shared__ int32_t scan[NT / 64];
  const VerletList& v = bv.verlet;
  for (int i = blockIdx.x; i < bv.n_mol; i += gridDim.x) {
    const int pi = bv.mol_pidx_start[i];
    const double xi = bv.px[pi], yi = bv.py[pi], zi = bv.pz[pi];
    if (threadIdx.x == 0) {
      v.ref[3 * i] = xi;
      v.ref[3 * i + 1] = yi;
      v.ref[3 * i + 2] = zi;
    }
    int n = 0;
    for (int base = 0; base < bv.n_mol; base += NT) {
      const int j = base + threadIdx.x;
      bool f = false;
      if (j < bv.n_mol && j != i) {
        const int pj = bv.mol_pidx_start[j];
        const double dx = wrap1(bv.px[pj] - xi, bv.L[0]);
        const double dy = wrap1(bv.py[pj] - yi, bv.L[1]);
        const double dz = wrap1(bv.pz[pj] - zi, bv.L[2]);
        f = dx * dx + dy * dy + dz * dz <= rv2;
      }
      int total;
      const int pos = block_scan_flag<NT>(f, scan, total);
      if (f && n + pos < v.cap) v.list[(size_t)i * v.cap + n + pos] = j;
      n += total;
    }
    if (threadIdx.x == 0) {
      v.count[i] = n;
      if (n > v.cap) atomicOr(v.stale, 2);  // capacity exceeded
    }
    __syncthreads();
  }
}

// Verlet list of molecule i built from a cell grid whose cells are at least
// rv wide (>= 3 per edge): the candidates are the molecules binned in the 27
// cells around i's first primary.  Hits are gathered in LDS and written in
// increasing j, the order the all-molecule build gives, so the list (and the
// summation order of every energy taken from it) does not depend on the
// grid.  More than kVerletGather hits for one molecule sets stale bit 4 (the
// host then uses the all-molecule build).
constexpr int kVerletGather = 2048;

template <int NT>
__global__ __launch_bounds__(NT) void k_verlet_build_cells(BoxView bv, double rv2) {
  __shared__ int32_t scan[NT / 64];
  __shared__ int32_t hits[kVerletGather];
  const VerletList& v = bv.verlet;
  const CellGrid& g = bv.cells;
  for (int i = blockIdx.x; i < bv.n_mol; i += gridDim.x) {
    const int pi = bv.mol_pidx_start[i];
    const double xi = bv.px[pi], yi = bv.py[pi], zi = bv.pz[pi];
    if (threadIdx.x == 0) {
      v.ref[3 * i] = xi;
      v.ref[3 * i + 1] = yi;
      v.ref[3 * i + 2] = zi;
    }
    const int cx = cell_coord(xi, g.inv_w[0], g.n[0]);
    const int cy = cell_coord(yi, g.inv_w[1], g.n[1]);
    const int cz = cell_coord(zi, g.inv_w[2], g.n[2]);
    int n = 0;
    for (int q = 0; q < 27; q++) {
      int ax = cx + q % 3 - 1, ay = cy + (q / 3) % 3 - 1, az = cz + q / 9 - 1;
      ax += ax < 0 ? g.n[0] : (ax >= g.n[0] ? -g.n[0] : 0);
      ay += ay < 0 ? g.n[1] : (ay >= g.n[1] ? -g.n[1] : 0);
      az += az < 0 ? g.n[2] : (az >= g.n[2] ? -g.n[2] : 0);
      const int c = cell_index(g, ax, ay, az);
      const int cnt = min(g.count[c], g.cap);
      for (int base = 0; base < cnt; base += NT) {
        const int k = base + (int)threadIdx.x;
        bool f = false;
        int j = -1;
        if (k < cnt) {
          j = g.slots[(size_t)c * g.cap + k];
          if (j != i) {
            const int pj = bv.mol_pidx_start[j];
            const double dx = wrap1(bv.px[pj] - xi, bv.L[0]);
            const double dy = wrap1(bv.py[pj] - yi, bv.L[1]);
            const double dz = wrap1(bv.pz[pj] - zi, bv.L[2]);
            f = dx * dx + dy * dy + dz * dz <= rv2;
          }
        }
        int total;
        const int pos = block_scan_flag<NT>(f, scan, total);
        if (f && n + pos < kVerletGather) hits[n + pos] = j;
        n += total;
      }
    }
    __syncthreads();
    const int m = min(n, kVerletGather);
    // rank of each hit among the gathered ones (ids are distinct)
    for (int k = threadIdx.x; k < m; k += NT) {
      const int j = hits[k];
      int rank = 0;
      for (int q = 0; q < m; q++) rank += hits[q] < j ? 1 : 0;
      if (rank < v.cap) v.list[(size_t)i * v.cap + rank] = j;
    }
    if (threadIdx.x == 0) {
      v.count[i] = n;
      if (n > v.cap) atomicOr(v.stale, 2);  // capacity exceeded
      if (n > kVerletGather) atomicOr(v.stale, 4);
    }
    __syncthreads();
  }
}

// Full-system energy of the box (SimulationStep.cpp:40-71), one wave per
// molecule (see fullenergy.h): partial[i] = the energy of the pairs molecule
// i holds (every pair is held by one of its molecules), pairs[i] = their
// site pairs.
template <bool MIXED>
__global__ __launch_bounds__(256) void k_system_waves(BoxView bv, double* partial,
                                                      double* pairs) {
  __shared__ WaveEval ws[4];
  __shared__ SiteParam64 p64[kParamCap];
  __shared__ SiteParam32 p32[kParamCap];
  full_energy_block<MIXED>(bv, partial, pairs, ws, p64, p32);
}

// Fixed-order sum of n partials into out[0] (and n counts into out[1]).
__global__ __launch_bounds__(256) void k_reduce2(const double* a,
                                                 const double* b, int n,
                                                 double* out) {
  __shared__ double red[4];
  double s = 0.0, c = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) {
    s += a[i];
    c += b[i];
  }
  s = block_sum<256>(s, red);
  c = block_sum<256>(c, red);
  if (threadIdx.x == 0) {
    out[0] = s;
    out[1] = c;
  }
}

// NPT rescale: every molecule translated so its first primary moves from p
// to f * p (internal geometry untouched).
__global__ void k_volume_scale(BoxView bv, double f) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= bv.n_mol) return;
  const int p = bv.pidx[bv.mol_pidx_start[m]];
  const double sx = bv.x[p] * f - bv.x[p];
  const double sy = bv.y[p] * f - bv.y[p];
  const double sz = bv.z[p] * f - bv.z[p];
  const int s0 = bv.mol_start[m], s1 = s0 + bv.mol_len[m];
  for (int s = s0; s < s1; s++) {
    bv.x[s] += sx;
    bv.y[s] += sy;
    bv.z[s] += sz;
  }
}

// Test hook: the device sin/cos/exp used by trial moves and Metropolis.
__global__ void k_debug_math(const double* x, int n, double* s, double* c,
                             double* e) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    sincos_fp64(x[i], s[i], c[i]);
    e[i] = exp_fp64(x[i]);
  }
}

// Test hook: the FP64 1/sqrt of the pair energies.
__global__ void k_debug_rsqrt(const double* x, int n, double* y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = rsqrt64(x[i]);
}

// ------------------------------------------------------ launch helpers

int launch_refresh_primary(const BoxView& bv, int n_pidx, hipStream_t s) {
  hipLaunchKernelGGL(k_refresh_primary, dim3((n_pidx + 255) / 256), dim3(256), 0,
                     s, bv, n_pidx);
  MCGPU_HIP_TRY(hipGetLastError());
  return MCGPU_OK;
}

int launch_cells_build(const BoxView& bv, hipStream_t s) {
  const CellGrid& g = bv.cells;
  const int ncell = g.n[0] * g.n[1] * g.n[2];
  hipLaunchKernelGGL(k_cells_clear, dim3((ncell + 255) / 256), dim3(256), 0, s,
                     g, ncell);
  hipLaunchKernelGGL(k_cells_assign, dim3((bv.n_mol + 255) / 256), dim3(256), 0,
                     s, bv);
  hipLaunchKernelGGL(k_cells_sort, dim3((ncell + 3) / 4), dim3(256), 0, s, g, ncell);
  MCGPU_HIP_TRY(hipGetLastError());
  return MCGPU_OK;
}

int launch_volume_scale(const BoxView& bv, double f, hipStream_t s) {
  hipLaunchKernelGGL(k_volume_scale, dim3((bv.n_mol + 255) / 256), dim3(256), 0,
                     s, bv, f);
  MCGPU_HIP_TRY(hipGetLastError());
  return MCGPU_OK;
}

// ------------------------------------------------------------ host side

int ensure_device() {
  static int state = -1;  // -1 unknown, 0 ok, >0 error
  if (state == 0) return MCGPU_OK;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess || n < 1) {
    set_error("no HIP device visible (the MI355X energy path needs a GPU)");
    return MCGPU_ERR_NODEVICE;
  }
  state = 0;
  return MCGPU_OK;
}

double molecule_reach(const double* x, const double* y, const double* z, int len,
                      int pf) {
  double r2 = 0.0;
  for (int k = 0; k < len; k++) {
    const double dx = x[k] - x[pf], dy = y[k] - y[pf], dz = z[k] - z[pf];
    r2 = std::max(r2, dx * dx + dy * dy + dz * dz);
  }
  return std::sqrt(r2);
}

ParamTable build_param_table(const mcgpu_box_desc& d) {
  ParamTable t;
  t.mol_param.resize(d.n_molecules);
  t.meta.resize(d.n_molecules);
  std::map<std::vector<uint64_t>, int32_t> sig;
  for (int32_t m = 0; m < d.n_molecules; m++) {
    const int32_t s0 = d.mol_start[m], len = d.mol_len[m];
    t.max_len = std::max(t.max_len, len);
    std::vector<uint64_t> key;
    key.reserve(1 + 3 * len);
    key.push_back((uint64_t)len);
    for (int32_t k = 0; k < len; k++) {
      uint64_t a, b, c;
      std::memcpy(&a, &d.sigma[s0 + k], 8);
      std::memcpy(&b, &d.epsilon[s0 + k], 8);
      std::memcpy(&c, &d.charge[s0 + k], 8);
      key.push_back(a);
      key.push_back(b);
      key.push_back(c);
    }
    auto it = sig.find(key);
    if (it == sig.end()) {
      const int32_t base = (int32_t)t.p64.size();
      for (int32_t k = 0; k < len; k++) {
        const double sg = d.sigma[s0 + k], ep = d.epsilon[s0 + k];
        const bool real = sg >= 0 && ep >= 0;  // BruteForceStep.cpp:64-65
        SiteParam64 p;
        p.s = real ? sg : 0.0;
        p.e = real ? 2.0 * std::sqrt(ep) : 0.0;
        p.q = real ? d.charge[s0 + k] * std::sqrt(kCoulombK) : 0.0;
        p.real = real ? 1.0 : 0.0;
        t.p64.push_back(p);
        t.p32.push_back(SiteParam32{(float)p.s, (float)p.e, (float)p.q, (float)p.real});
      }
      it = sig.emplace(key, base).first;
    }
    t.mol_param[m] = it->second;
    const int32_t pf = d.primary_index[d.mol_pidx_start[m]] - s0;
    t.meta[m] = MolMeta{s0, (int16_t)len, (int16_t)pf, it->second, d.mol_pidx_start[m]};
    if (d.mol_pidx_count[m] != 1 || d.mol_pidx_start[m] != m) t.single_primary = 0;
    if (m == 0) {
      t.uni_len = len;
      t.uni_param = it->second;
      t.uni_p0 = pf;
    }
    if (s0 != m * t.uni_len || len != t.uni_len || it->second != t.uni_param ||
        pf != t.uni_p0)
      t.uniform = false;
    // primary spread and site reach (rigid molecules: fixed)
    const int32_t p0 = d.primary_index[d.mol_pidx_start[m]];
    t.reach = std::max(t.reach, molecule_reach(d.x + s0, d.y + s0, d.z + s0, len, pf));
    for (int32_t k = 1; k < d.mol_pidx_count[m]; k++) {
      const int32_t p = d.primary_index[d.mol_pidx_start[m] + k];
      const double dx = d.x[p] - d.x[p0], dy = d.y[p] - d.y[p0], dz = d.z[p] - d.z[p0];
      t.spread = std::max(t.spread, std::sqrt(dx * dx + dy * dy + dz * dz));
    }
  }
  if (!t.single_primary) t.uniform = false;
  if (!t.uniform) t.uni_len = t.uni_param = t.uni_p0 = 0;
  return t;
}

void cell_geometry(const double L[3], double cutoff, double spread, int n[3],
                   double inv_w[3]) {
  const double w = cutoff + 2.0 * spread + 1e-9;
  for (int k = 0; k < 3; k++) {
    n[k] = std::max(1, (int)std::floor(L[k] / w));
    inv_w[k] = n[k] / L[k];
  }
}

}  // namespace mcgpu

using namespace mcgpu;

struct mcgpu_box {
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int precision = MCGPU_FP64;
  BoxView bv{};
  int32_t n_pidx = 0;
  double spread = 0.0;
  double skin = 0.0;
  std::vector<void*> allocs;  // device allocations owned by the box
  int32_t* cell_count = nullptr;
  int32_t* cell_slots = nullptr;
  int32_t ncell_alloc = 0;
  int32_t slots_alloc = 0;
  MoveOut* h_out = nullptr;  // pinned, mapped
  MoveOut* d_out = nullptr;
  double* d_partial = nullptr;
  double* d_pairs = nullptr;
  double* h_total = nullptr;  // pinned, mapped [2]
  double* d_total = nullptr;
  int64_t last_pairs = 0;
  float last_ms = 0.0f;
  std::vector<int32_t> h_mol_start, h_mol_len;
  std::vector<int16_t> h_pf;  // first primary's offset within each molecule
  double reach = 0.0;         // see ParamTable::reach (only ever grows)
  // Verlet mode: a commit moved a primary past skin/2 (rebuild before the
  // next evaluation), or a set-molecule call may have (read its flag first)
  bool verlet_dirty = false, verlet_check = false;
  std::vector<double> h_ref;  // first-primary positions of the last build
  int64_t verlet_builds = 0;  // list builds since the box was created
  int64_t fallbacks = 0;      // evaluations that scanned every molecule
                              // although cell or Verlet lists were selected
};

namespace {

template <class T>
int dalloc(mcgpu_box* b, T** p, size_t count) {
  void* q = nullptr;
  MCGPU_HIP_TRY(hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T)));
  b->allocs.push_back(q);
  *p = (T*)q;
  return MCGPU_OK;
}

void free_box(mcgpu_box* b) {
  if (!b) return;
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  for (void* p : b->allocs) (void)hipFree(p);
  if (b->cell_count) (void)hipFree(b->cell_count);
  if (b->cell_slots) (void)hipFree(b->cell_slots);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->h_total) (void)hipHostFree(b->h_total);
  if (b->ev0) (void)hipEventDestroy(b->ev0);
  if (b->ev1) (void)hipEventDestroy(b->ev1);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
}

int check_mol(mcgpu_box* b, int32_t mol) {
  if (!b) {
    set_error("null box");
    return MCGPU_ERR_INVALID;
  }
  if (mol < 0 || mol >= b->bv.n_mol) {
    set_error("molecule index out of range");
    return MCGPU_ERR_RANGE;
  }
  return MCGPU_OK;
}

int refresh_primary(mcgpu_box* b) {
  const int n = b->n_pidx;
  hipLaunchKernelGGL(k_refresh_primary, dim3((n + 255) / 256), dim3(256), 0,
                     b->stream, b->bv, n);
  MCGPU_HIP_TRY(hipGetLastError());
  return MCGPU_OK;
}

// Bins every molecule by its first primary into cells at least
// `reach + 2 spread` wide (cell_geometry), slots ordered by molecule id.
int bin_cells(mcgpu_box* b, double reach) {
  CellGrid& g = b->bv.cells;
  cell_geometry(b->bv.L, reach, b->spread, g.n, g.inv_w);
  const int ncell = g.n[0] * g.n[1] * g.n[2];
  if (ncell > b->ncell_alloc) {
    if (b->cell_count) MCGPU_HIP_TRY(hipFree(b->cell_count));
    b->cell_count = nullptr;
    MCGPU_HIP_TRY(hipMalloc(&b->cell_count, sizeof(int32_t) * ncell));
    b->ncell_alloc = ncell;
  }
  g.count = gmem(b->cell_count);
  int cap = std::max(8, (int)std::ceil(2.0 * b->bv.n_mol / ncell) + 8);
  cap = (cap + 3) & ~3;
  for (int attempt = 0; attempt < 8; attempt++) {
    if ((size_t)ncell * cap > (size_t)b->slots_alloc) {
      if (b->cell_slots) MCGPU_HIP_TRY(hipFree(b->cell_slots));
      b->cell_slots = nullptr;
      MCGPU_HIP_TRY(hipMalloc(&b->cell_slots, sizeof(int32_t) * (size_t)ncell * cap));
      b->slots_alloc = ncell * cap;
    }
    g.cap = cap;
    g.slots = gmem(b->cell_slots);
    hipLaunchKernelGGL(k_cells_clear, dim3((ncell + 255) / 256), dim3(256), 0,
                       b->stream, g, ncell);
    hipLaunchKernelGGL(k_cells_assign, dim3((b->bv.n_mol + 255) / 256), dim3(256),
                       0, b->stream, b->bv);
    MCGPU_HIP_TRY(hipGetLastError());
    int32_t ovf = 0;
    MCGPU_HIP_TRY(hipMemcpyAsync(&ovf, g.overflow, sizeof(int32_t),
                                 hipMemcpyDeviceToHost, b->stream));
    MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
    if (!ovf) {
      hipLaunchKernelGGL(k_cells_sort, dim3((ncell + 3) / 4), dim3(256), 0, b->stream,
                         g, ncell);
      MCGPU_HIP_TRY(hipGetLastError());
      return MCGPU_OK;
    }
    std::vector<int32_t> counts(ncell);
    MCGPU_HIP_TRY(hipMemcpy(counts.data(), g.count, sizeof(int32_t) * ncell,
                            hipMemcpyDeviceToHost));
    const int mx = *std::max_element(counts.begin(), counts.end());
    cap = ((int)(mx * 1.5) + 4 + 3) & ~3;
  }
  set_error("cell list build did not converge");
  return MCGPU_ERR_STATE;
}

int build_cells(mcgpu_box* b) { return bin_cells(b, std::sqrt(b->bv.cutoff2)); }

// Verlet lists (VerletCalcs::createVerlet/fillVerlet, VerletStep.cpp:119-184,
// done correctly): every j whose first primary lies within rv = rc +
// 2 spread + skin of i's.  Candidates come from cells of width >= rv when
// the box holds at least 3 per edge (O(N) work), else from every molecule.
int build_verlet(mcgpu_box* b) {
  VerletList& v = b->bv.verlet;
  const double rv = std::sqrt(b->bv.cutoff2) + 2.0 * b->spread + b->skin;
  v.skin_half2 = 0.25 * b->skin * b->skin;
  int n[3];
  double iw[3];
  cell_geometry(b->bv.L, std::sqrt(b->bv.cutoff2) + b->skin, b->spread, n, iw);
  bool from_cells = n[0] >= 3 && n[1] >= 3 && n[2] >= 3;
  if (from_cells) {
    const int rc = bin_cells(b, std::sqrt(b->bv.cutoff2) + b->skin);
    if (rc) return rc;
  }
  b->verlet_builds++;
  for (int attempt = 0; attempt < 6; attempt++) {
    MCGPU_HIP_TRY(hipMemsetAsync(v.stale, 0, sizeof(int32_t), b->stream));
    const int grid = std::min(b->bv.n_mol, 4096);
    if (from_cells)
      hipLaunchKernelGGL(k_verlet_build_cells<256>, dim3(grid), dim3(256), 0, b->stream,
                         b->bv, rv * rv);
    else
      hipLaunchKernelGGL(k_verlet_build<256>, dim3(grid), dim3(256), 0, b->stream,
                         b->bv, rv * rv);
    MCGPU_HIP_TRY(hipGetLastError());
    int32_t st = 0;
    MCGPU_HIP_TRY(hipMemcpyAsync(&st, v.stale, sizeof(int32_t),
                                 hipMemcpyDeviceToHost, b->stream));
    MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
    if (!st) {
      // the build positions, for the host's check of proposals (see
      // mcgpu_trial_move)
      b->h_ref.resize(3 * (size_t)b->bv.n_mol);
      MCGPU_HIP_TRY(hipMemcpy(b->h_ref.data(), v.ref, sizeof(double) * b->h_ref.size(),
                              hipMemcpyDeviceToHost));
      return MCGPU_OK;
    }
    if (st & 4) from_cells = false;  // a molecule gathered more than LDS holds
    std::vector<int32_t> cnt(b->bv.n_mol);
    MCGPU_HIP_TRY(hipMemcpy(cnt.data(), v.count, sizeof(int32_t) * cnt.size(),
                            hipMemcpyDeviceToHost));
    const int mx = *std::max_element(cnt.begin(), cnt.end());
    const int cap = ((int)(mx * 1.25) + 8 + 7) & ~7;
    // grow the list storage
    for (auto it = b->allocs.begin(); it != b->allocs.end(); ++it)
      if (*it == (void*)v.list) {
        b->allocs.erase(it);
        break;
      }
    MCGPU_HIP_TRY(hipFree(v.list));
    v.list = nullptr;
    if (dalloc(b, &v.list, (size_t)b->bv.n_mol * cap) != MCGPU_OK) return MCGPU_ERR_HIP;
    v.cap = cap;
  }
  set_error("Verlet list build did not converge");
  return MCGPU_ERR_STATE;
}

int rebuild_neighbors(mcgpu_box* b) {
  if (b->bv.mode == MCGPU_NB_CELL) return build_cells(b);
  if (b->bv.mode == MCGPU_NB_VERLET) return build_verlet(b);
  return MCGPU_OK;
}

int launch_move(mcgpu_box* b, const MoveArgs& a) {
  const bool big = (b->bv.mode == MCGPU_NB_BRUTE) && b->bv.n_mol > 2048;
  const bool mixed = b->precision == MCGPU_MIXED;
  if (b->bv.mode == MCGPU_NB_VERLET) {
    // VerletCalcs::updateVerlet (VerletStep.cpp:186-207): once a primary has
    // moved more than skin/2 since the build, rebuild before evaluating
    if (b->verlet_check) {
      MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
      if (b->h_out->stale) b->verlet_dirty = true;
      b->verlet_check = false;
    }
    if (b->verlet_dirty) {
      const int rc = build_verlet(b);
      if (rc) return rc;
      b->verlet_dirty = false;
    }
  }
  if (b->ev0) MCGPU_HIP_TRY(hipEventRecord(b->ev0, b->stream));
  if (big) {
    if (mixed)
      hipLaunchKernelGGL((k_move<1024, true>), dim3(1), dim3(1024), 0, b->stream, b->bv, a, b->d_out);
    else
      hipLaunchKernelGGL((k_move<1024, false>), dim3(1), dim3(1024), 0, b->stream, b->bv, a, b->d_out);
  } else {
    if (mixed)
      hipLaunchKernelGGL((k_move<256, true>), dim3(1), dim3(256), 0, b->stream, b->bv, a, b->d_out);
    else
      hipLaunchKernelGGL((k_move<256, false>), dim3(1), dim3(256), 0, b->stream, b->bv, a, b->d_out);
  }
  MCGPU_HIP_TRY(hipGetLastError());
  if (b->ev1) MCGPU_HIP_TRY(hipEventRecord(b->ev1, b->stream));
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->ev0) MCGPU_HIP_TRY(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  if (b->bv.mode != MCGPU_NB_BRUTE && b->h_out->source == MCGPU_NB_BRUTE) b->fallbacks++;
  if (b->bv.mode == MCGPU_NB_VERLET && b->h_out->stale) b->verlet_dirty = true;
  return MCGPU_OK;
}

// Caller-supplied coordinates may change a molecule's shape: keep the
// box's site reach an upper bound of every molecule's.
void note_reach(mcgpu_box* b, double reach) {
  if (reach > b->reach) {
    b->reach = reach;
    b->bv.reach2 = reach2_of(reach);
  }
}

int fill_coords(mcgpu_box* b, int32_t mol, const double* xyz, MoveArgs& a) {
  const int len = b->h_mol_len[mol];
  for (int k = 0; k < 3 * len; k++) {
    if (!std::isfinite(xyz[k])) {
      set_error("non-finite coordinate");
      return MCGPU_ERR_INVALID;
    }
    a.nxyz[k] = xyz[k];
  }
  note_reach(b, molecule_reach(xyz, xyz + len, xyz + 2 * len, len, b->h_pf[mol]));
  return MCGPU_OK;
}

}  // namespace

extern "C" int mcgpu_device_count(int* count) {
  if (!count) return MCGPU_ERR_INVALID;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return MCGPU_OK;
}

extern "C" int mcgpu_device_synchronize(void) {
  MCGPU_HIP_TRY(hipDeviceSynchronize());
  return MCGPU_OK;
}

extern "C" int mcgpu_box_create(const mcgpu_box_desc* d, int precision,
                                mcgpu_box** out) {
  if (!d || !out) {
    set_error("mcgpu_box_create: null argument");
    return MCGPU_ERR_INVALID;
  }
  *out = nullptr;
  const std::string why = validate_desc(*d);
  if (!why.empty()) {
    set_error("mcgpu_box_create: " + why);
    return MCGPU_ERR_INVALID;
  }
  if (precision != MCGPU_FP64 && precision != MCGPU_MIXED) {
    set_error("mcgpu_box_create: precision must be MCGPU_FP64 or MCGPU_MIXED");
    return MCGPU_ERR_INVALID;
  }
  const ParamTable pt = build_param_table(*d);
  if ((int)pt.p64.size() > kParamCap) {
    set_error("mcgpu_box_create: more than " + std::to_string(kParamCap) +
              " distinct site-parameter entries");
    return MCGPU_ERR_INVALID;
  }
  int rc = ensure_device();
  if (rc) return rc;
  auto* b = new (std::nothrow) mcgpu_box();
  if (!b) return MCGPU_ERR_NOMEM;
  b->precision = precision;
  b->spread = pt.spread;
  b->n_pidx = d->n_primary;
  b->h_mol_start.assign(d->mol_start, d->mol_start + d->n_molecules);
  b->h_mol_len.assign(d->mol_len, d->mol_len + d->n_molecules);
  b->h_pf.resize(d->n_molecules);
  for (int32_t m = 0; m < d->n_molecules; m++) b->h_pf[m] = pt.meta[m].p0;
  b->reach = pt.reach;
  BoxView& v = b->bv;
  v.reach2 = reach2_of(pt.reach);
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
  v.kT = d->kT;
  const size_t ns = d->n_sites, nm = d->n_molecules, np = d->n_primary;
#define ALLOC(p, n)                                           \
  if ((rc = dalloc(b, &p, n)) != MCGPU_OK) {                  \
    free_box(b);                                              \
    return rc;                                                \
  }
  double *x, *y, *z, *px, *py, *pz;
  int32_t *ms, *ml, *mps, *mpc, *mp, *pi;
  MolMeta* mm;
  SiteParam64* p64;
  SiteParam32* p32;
  ALLOC(x, ns);
  ALLOC(y, ns);
  ALLOC(z, ns);
  ALLOC(px, np);
  ALLOC(py, np);
  ALLOC(pz, np);
  ALLOC(ms, nm);
  ALLOC(ml, nm);
  ALLOC(mps, nm);
  ALLOC(mpc, nm);
  ALLOC(mp, nm);
  ALLOC(mm, nm);
  ALLOC(pi, np);
  ALLOC(p64, pt.p64.size());
  ALLOC(p32, pt.p32.size());
  ALLOC(v.cells.mol_cell, nm);
  ALLOC(v.cells.mol_slot, nm);
  ALLOC(v.cells.overflow, 1);
  ALLOC(v.verlet.count, nm);
  ALLOC(v.verlet.ref, 3 * nm);
  ALLOC(v.verlet.stale, 1);
  v.verlet.cap = 64;
  ALLOC(v.verlet.list, nm * 64);
  ALLOC(b->d_partial, nm);
  ALLOC(b->d_pairs, nm);
#undef ALLOC
  v.x = gmem(x);
  v.y = gmem(y);
  v.z = gmem(z);
  v.px = gmem(px);
  v.py = gmem(py);
  v.pz = gmem(pz);
  v.mol_start = gmem(ms);
  v.mol_len = gmem(ml);
  v.mol_pidx_start = gmem(mps);
  v.mol_pidx_count = gmem(mpc);
  v.mol_param = gmem(mp);
  v.meta = gmem(mm);
  v.single_primary = pt.single_primary;
  v.pidx = gmem(pi);
  v.p64 = gmem(p64);
  v.p32 = gmem(p32);
  auto fail = [&](int code) {
    free_box(b);
    return code;
  };
  if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    return fail(MCGPU_ERR_HIP);
  }
  if (hipHostMalloc((void**)&b->h_out, sizeof(MoveOut), hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&b->d_out, b->h_out, 0) != hipSuccess ||
      hipHostMalloc((void**)&b->h_total, 2 * sizeof(double), hipHostMallocMapped) != hipSuccess ||
      hipHostGetDevicePointer((void**)&b->d_total, b->h_total, 0) != hipSuccess) {
    set_error("pinned result buffer allocation failed");
    return fail(MCGPU_ERR_NOMEM);
  }
  struct Copy {
    void* dst;
    const void* src;
    size_t bytes;
  } copies[] = {{x, d->x, 8 * ns},
                {y, d->y, 8 * ns},
                {z, d->z, 8 * ns},
                {ms, d->mol_start, 4 * nm},
                {ml, d->mol_len, 4 * nm},
                {mps, d->mol_pidx_start, 4 * nm},
                {mpc, d->mol_pidx_count, 4 * nm},
                {mp, pt.mol_param.data(), 4 * nm},
                {mm, pt.meta.data(), sizeof(MolMeta) * nm},
                {pi, d->primary_index, 4 * np},
                {p64, pt.p64.data(), sizeof(SiteParam64) * pt.p64.size()},
                {p32, pt.p32.data(), sizeof(SiteParam32) * pt.p32.size()}};
  for (const Copy& c : copies) {
    if (hipMemcpy(c.dst, c.src, c.bytes, hipMemcpyHostToDevice) != hipSuccess) {
      set_error("mcgpu_box_create: upload failed");
      return fail(MCGPU_ERR_HIP);
    }
  }
  if (hipMemset(v.cells.overflow, 0, 4) != hipSuccess ||
      hipMemset(v.verlet.stale, 0, 4) != hipSuccess ||
      hipMemset(v.verlet.count, 0, 4 * nm) != hipSuccess) {
    set_error("mcgpu_box_create: memset failed");
    return fail(MCGPU_ERR_HIP);
  }
  if ((rc = refresh_primary(b)) != MCGPU_OK) return fail(rc);
  if (hipStreamSynchronize(b->stream) != hipSuccess) {
    set_error("mcgpu_box_create: synchronize failed");
    return fail(MCGPU_ERR_HIP);
  }
  *out = b;
  return MCGPU_OK;
}

extern "C" int mcgpu_box_destroy(mcgpu_box* b) {
  free_box(b);
  return MCGPU_OK;
}

extern "C" int mcgpu_box_set_precision(mcgpu_box* b, int precision) {
  if (!b || (precision != MCGPU_FP64 && precision != MCGPU_MIXED)) {
    set_error("mcgpu_box_set_precision: bad argument");
    return MCGPU_ERR_INVALID;
  }
  b->precision = precision;
  return MCGPU_OK;
}

extern "C" int mcgpu_box_enable_timing(mcgpu_box* b, int on) {
  if (!b) return MCGPU_ERR_INVALID;
  if (on && !b->ev0) {
    MCGPU_HIP_TRY(hipEventCreate(&b->ev0));
    MCGPU_HIP_TRY(hipEventCreate(&b->ev1));
  } else if (!on && b->ev0) {
    MCGPU_HIP_TRY(hipEventDestroy(b->ev0));
    MCGPU_HIP_TRY(hipEventDestroy(b->ev1));
    b->ev0 = b->ev1 = nullptr;
  }
  return MCGPU_OK;
}

extern "C" int mcgpu_last_kernel_ms(mcgpu_box* b, double* ms) {
  if (!b || !ms) return MCGPU_ERR_INVALID;
  *ms = b->last_ms;
  return MCGPU_OK;
}

extern "C" int mcgpu_neighbor_rebuild(mcgpu_box* b, int kind, double skin) {
  if (!b) return MCGPU_ERR_INVALID;
  if (kind != MCGPU_NEIGHBOR_BRUTE && kind != MCGPU_NEIGHBOR_CELL &&
      kind != MCGPU_NEIGHBOR_VERLET) {
    set_error("mcgpu_neighbor_rebuild: unknown kind");
    return MCGPU_ERR_INVALID;
  }
  if (kind == MCGPU_NEIGHBOR_VERLET && !(skin > 0)) {
    set_error("mcgpu_neighbor_rebuild: Verlet skin must be > 0");
    return MCGPU_ERR_INVALID;
  }
  b->bv.mode = kind;
  b->skin = skin;
  int rc = rebuild_neighbors(b);
  if (rc) return rc;
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  return MCGPU_OK;
}

extern "C" int mcgpu_upload_coords(mcgpu_box* b, const double* x,
                                   const double* y, const double* z) {
  if (!b || !x || !y || !z) return MCGPU_ERR_INVALID;
  const size_t n = b->bv.n_site;
  for (int32_t m = 0; m < b->bv.n_mol; m++) {
    const int32_t s0 = b->h_mol_start[m];
    note_reach(b, molecule_reach(x + s0, y + s0, z + s0, b->h_mol_len[m], b->h_pf[m]));
  }
  MCGPU_HIP_TRY(hipMemcpyAsync(b->bv.x, x, 8 * n, hipMemcpyHostToDevice, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(b->bv.y, y, 8 * n, hipMemcpyHostToDevice, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(b->bv.z, z, 8 * n, hipMemcpyHostToDevice, b->stream));
  int rc = refresh_primary(b);
  if (rc) return rc;
  rc = rebuild_neighbors(b);
  if (rc) return rc;
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  return MCGPU_OK;
}

extern "C" int mcgpu_download_coords(mcgpu_box* b, double* x, double* y,
                                     double* z) {
  if (!b || !x || !y || !z) return MCGPU_ERR_INVALID;
  const size_t n = b->bv.n_site;
  MCGPU_HIP_TRY(hipMemcpyAsync(x, b->bv.x, 8 * n, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(y, b->bv.y, 8 * n, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(z, b->bv.z, 8 * n, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  return MCGPU_OK;
}

extern "C" int mcgpu_set_molecule_coords(mcgpu_box* b, int32_t mol,
                                         const double* xyz) {
  int rc = check_mol(b, mol);
  if (rc) return rc;
  if (!xyz) return MCGPU_ERR_INVALID;
  MoveArgs a{};
  a.mol = mol;
  if ((rc = fill_coords(b, mol, xyz, a)) != MCGPU_OK) return rc;
  hipLaunchKernelGGL(k_set_mol, dim3(1), dim3(64), 0, b->stream, b->bv, a, b->d_out);
  MCGPU_HIP_TRY(hipGetLastError());
  // the commit's Verlet displacement test is read before the next evaluation
  if (b->bv.mode == MCGPU_NB_VERLET) b->verlet_check = true;
  return MCGPU_OK;
}

extern "C" int mcgpu_box_neighbor_stats(mcgpu_box* b, int64_t* builds, int64_t* fallbacks) {
  if (!b || !builds || !fallbacks) return MCGPU_ERR_INVALID;
  *builds = b->verlet_builds;
  *fallbacks = b->fallbacks;
  return MCGPU_OK;
}

extern "C" int mcgpu_get_molecule_coords(mcgpu_box* b, int32_t mol,
                                         double* xyz) {
  int rc = check_mol(b, mol);
  if (rc) return rc;
  if (!xyz) return MCGPU_ERR_INVALID;
  const int len = b->h_mol_len[mol], s = b->h_mol_start[mol];
  MCGPU_HIP_TRY(hipMemcpyAsync(xyz, b->bv.x + s, 8 * len, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(xyz + len, b->bv.y + s, 8 * len, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipMemcpyAsync(xyz + 2 * len, b->bv.z + s, 8 * len, hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  return MCGPU_OK;
}

extern "C" int mcgpu_molecule_energy(mcgpu_box* b, int32_t mol,
                                     int32_t start_mol, double* energy) {
  int rc = check_mol(b, mol);
  if (rc) return rc;
  if (!energy) return MCGPU_ERR_INVALID;
  if (start_mol < 0 || start_mol > b->bv.n_mol) {
    set_error("start_mol out of range");
    return MCGPU_ERR_RANGE;
  }
  MoveArgs a{};
  a.mol = mol;
  a.start_mol = start_mol;
  if ((rc = launch_move(b, a)) != MCGPU_OK) return rc;
  *energy = b->h_out->e_old;
  return MCGPU_OK;
}

extern "C" int mcgpu_trial_move(mcgpu_box* b, int32_t mol,
                                const double* new_xyz, double u,
                                mcgpu_move_result* res) {
  int rc = check_mol(b, mol);
  if (rc) return rc;
  if (!new_xyz || !res) return MCGPU_ERR_INVALID;
  MoveArgs a{};
  a.mol = mol;
  a.start_mol = 0;
  a.has_new = 1;
  a.metropolis = 1;
  a.u = u;
  if ((rc = fill_coords(b, mol, new_xyz, a)) != MCGPU_OK) return rc;
  if (b->bv.mode == MCGPU_NB_VERLET && b->h_ref.size() == 3 * (size_t)b->bv.n_mol) {
    // A proposal whose first primary lies more than skin/2 from its build
    // position is not covered by the lists: rebuild them first (from the
    // committed positions), so the evaluation can still use them.
    const int len = b->h_mol_len[mol], pf = b->h_pf[mol];
    const double* r = b->h_ref.data() + 3 * (size_t)mol;
    double d2 = 0.0;
    for (int k = 0; k < 3; k++) {
      double d = new_xyz[k * len + pf] - r[k];
      const double L = b->bv.L[k];
      d += d < -0.5 * L ? L : 0.0;  // SimulationStep.cpp:377-385
      d -= d > 0.5 * L ? L : 0.0;
      d2 += d * d;
    }
    if (d2 > b->bv.verlet.skin_half2) b->verlet_dirty = true;
  }
  if ((rc = launch_move(b, a)) != MCGPU_OK) return rc;
  res->old_energy = b->h_out->e_old;
  res->new_energy = b->h_out->e_new;
  res->delta = b->h_out->de;
  res->accepted = b->h_out->accepted;
  res->reserved = 0;
  return MCGPU_OK;
}

extern "C" int mcgpu_system_energy(mcgpu_box* b, double* energy) {
  if (!b || !energy) return MCGPU_ERR_INVALID;
  const int n = b->bv.n_mol;
  // one molecule per wave: the hardware hands out blocks as others finish,
  // which evens out molecules of unequal pair work
  const int grid = (n + 3) / 4;
  if (b->ev0) MCGPU_HIP_TRY(hipEventRecord(b->ev0, b->stream));
  if (b->precision == MCGPU_MIXED)
    hipLaunchKernelGGL(k_system_waves<true>, dim3(grid), dim3(256), 0, b->stream, b->bv,
                       b->d_partial, b->d_pairs);
  else
    hipLaunchKernelGGL(k_system_waves<false>, dim3(grid), dim3(256), 0, b->stream, b->bv,
                       b->d_partial, b->d_pairs);
  MCGPU_HIP_TRY(hipGetLastError());
  if (b->ev1) MCGPU_HIP_TRY(hipEventRecord(b->ev1, b->stream));
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, b->stream, b->d_partial,
                     b->d_pairs, n, b->d_total);
  MCGPU_HIP_TRY(hipGetLastError());
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  if (b->ev0) MCGPU_HIP_TRY(hipEventElapsedTime(&b->last_ms, b->ev0, b->ev1));
  *energy = b->h_total[0];
  b->last_pairs = (int64_t)b->h_total[1];
  return MCGPU_OK;
}

extern "C" int mcgpu_system_pair_count(mcgpu_box* b, int64_t* pairs) {
  if (!b || !pairs) return MCGPU_ERR_INVALID;
  *pairs = b->last_pairs;
  return MCGPU_OK;
}

// ---------------------------------------- full energy without allocation
//
// The per-molecule partial energies and pair counts of a full-energy pass,
// and its two totals, live in one file-scope device array instead of the
// box's own buffers: [0, K) partials, [K, 2K) pair counts, [2K, 2K + 2)
// totals, K = kScratchMolecules.  Every entry a pass reads it writes first
// (one wave per molecule writes that molecule's partial and count; the
// reduction writes both totals), so nothing needs clearing between calls.
// Each GPU holds its own copy until the process ends; calls that overlap on
// one GPU share it and race.
constexpr int kScratchMolecules = 1 << 17;
static __device__ __align__(256) double scratch[2 * kScratchMolecules + 2];

namespace {

constexpr int kMaxScratchDevices = 64;
double* g_scratch[kMaxScratchDevices] = {};
std::mutex g_scratch_mu;

// The device address of the current device's `scratch` (the name itself is
// only a host-side stand-in): looked up once per device with that device
// current, then reused.
int scratch_address(double** out) {
  int dev = 0;
  MCGPU_HIP_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxScratchDevices) {
    set_error("device index beyond the scratch table");
    return MCGPU_ERR_RANGE;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return MCGPU_ERR_HIP;
    }
    g_scratch[dev] = (double*)p;
  }
  *out = g_scratch[dev];
  return MCGPU_OK;
}

}  // namespace

extern "C" int mcgpu_system_energy_noalloc(mcgpu_box* b, double* energy) {
  if (!b || !energy) return MCGPU_ERR_INVALID;
  const int n = b->bv.n_mol;
  if (n > kScratchMolecules) {
    set_error("mcgpu_system_energy_noalloc: " + std::to_string(n) +
              " molecules exceed the scratch limit of " + std::to_string(kScratchMolecules));
    return MCGPU_ERR_RANGE;
  }
  double* ptr = nullptr;
  int rc = scratch_address(&ptr);
  if (rc) return rc;
  double* partial = ptr;
  double* pairs = ptr + kScratchMolecules;
  double* total = ptr + 2 * kScratchMolecules;
  const int grid = (n + 3) / 4;
  if (b->precision == MCGPU_MIXED)
    hipLaunchKernelGGL(k_system_waves<true>, dim3(grid), dim3(256), 0, b->stream, b->bv,
                       partial, pairs);
  else
    hipLaunchKernelGGL(k_system_waves<false>, dim3(grid), dim3(256), 0, b->stream, b->bv,
                       partial, pairs);
  MCGPU_HIP_TRY(hipGetLastError());
  hipLaunchKernelGGL(k_reduce2, dim3(1), dim3(256), 0, b->stream, partial, pairs, n, total);
  MCGPU_HIP_TRY(hipGetLastError());
  double t[2] = {0.0, 0.0};
  MCGPU_HIP_TRY(hipMemcpyAsync(t, total, sizeof(t), hipMemcpyDeviceToHost, b->stream));
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  *energy = t[0];
  b->last_pairs = (int64_t)t[1];
  return MCGPU_OK;
}

extern "C" int mcgpu_volume_scale(mcgpu_box* b, double factor) {
  if (!b || !(factor > 0) || !std::isfinite(factor)) {
    set_error("mcgpu_volume_scale: factor must be finite and > 0");
    return MCGPU_ERR_INVALID;
  }
  const int n = b->bv.n_mol;
  hipLaunchKernelGGL(k_volume_scale, dim3((n + 255) / 256), dim3(256), 0,
                     b->stream, b->bv, factor);
  MCGPU_HIP_TRY(hipGetLastError());
  for (int k = 0; k < 3; k++) b->bv.L[k] *= factor;
  int rc = refresh_primary(b);
  if (rc) return rc;
  if ((rc = rebuild_neighbors(b)) != MCGPU_OK) return rc;
  MCGPU_HIP_TRY(hipStreamSynchronize(b->stream));
  return MCGPU_OK;
}

extern "C" int mcgpu_box_size(mcgpu_box* b, double* edges3) {
  if (!b || !edges3) return MCGPU_ERR_INVALID;
  for (int k = 0; k < 3; k++) edges3[k] = b->bv.L[k];
  return MCGPU_OK;
}

extern "C" int mcgpu_debug_math(const double* x, int32_t n, double* s, double* c,
                                double* e) {
  if (!x || !s || !c || !e || n < 0) return MCGPU_ERR_INVALID;
  int rc = ensure_device();
  if (rc) return rc;
  if (n == 0) return MCGPU_OK;
  double* d = nullptr;
  MCGPU_HIP_TRY(hipMalloc(&d, sizeof(double) * 4 * (size_t)n));
  hipError_t err = hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(k_debug_math, dim3((n + 255) / 256), dim3(256), 0, 0, d, n,
                       d + n, d + 2 * (size_t)n, d + 3 * (size_t)n);
    err = hipDeviceSynchronize();
  }
  if (err == hipSuccess) err = hipMemcpy(s, d + n, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(c, d + 2 * (size_t)n, sizeof(double) * n, hipMemcpyDeviceToHost);
  if (err == hipSuccess) err = hipMemcpy(e, d + 3 * (size_t)n, sizeof(double) * n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (err != hipSuccess) {
    set_error(std::string("mcgpu_debug_math: ") + hipGetErrorString(err));
    return MCGPU_ERR_HIP;
  }
  return MCGPU_OK;
}

extern "C" int mcgpu_debug_rsqrt(const double* x, int32_t n, double* y) {
  if (!x || !y || n < 0) return MCGPU_ERR_INVALID;
  int rc = ensure_device();
  if (rc) return rc;
  if (n == 0) return MCGPU_OK;
  double* d = nullptr;
  MCGPU_HIP_TRY(hipMalloc(&d, sizeof(double) * 2 * (size_t)n));
  hipError_t err = hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice);
  if (err == hipSuccess) {
    hipLaunchKernelGGL(k_debug_rsqrt, dim3((n + 255) / 256), dim3(256), 0, 0, d, n, d + n);
    err = hipDeviceSynchronize();
  }
  if (err == hipSuccess) err = hipMemcpy(y, d + n, sizeof(double) * n, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (err != hipSuccess) {
    set_error(std::string("mcgpu_debug_rsqrt: ") + hipGetErrorString(err));
    return MCGPU_ERR_HIP;
  }
  return MCGPU_OK;
}
