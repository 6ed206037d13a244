// Device-side building blocks of the energy path for gfx950 (CDNA4, wave64).
//
// Data layout in HBM (one box):
//   x, y, z          [n_site]  FP64 master coordinates (SoA, coalesced)
//   px, py, pz       [n_pidx]  FP64 mirror of the primary sites, so the
//                              candidate scan reads 24 contiguous bytes per
//                              molecule instead of gathering from x/y/z
//   mol_start/len, mol_pidx_start/count, mol_param   [n_mol] int32
//   p64 / p32        per distinct molecule signature, per site:
//                    {sqrt(sigma), sqrt(eps), q, real}; sigma_ab =
//                    sqrt(sigma_a) sqrt(sigma_b), eps_ab likewise, which is
//                    the reference's geometric mixing (SimulationStep.cpp:
//                    369-375) without a per-pair square root
//   cell lists       count[ncell], slots[ncell*cap], mol_cell/mol_slot[n_mol]
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "rng31.h"

namespace mcgpu {

constexpr int kMaxSites = 32;        // MCGPU_MAX_MOL_SITES
constexpr double kCoulombK = 332.06; // SimulationStep.cpp:364

// Site parameters in the form the pair terms use: s = sigma, e =
// 2 sqrt(epsilon), q = charge * sqrt(332.06), real = 1 (0 for dummy sites,
// whose other fields are 0).  Then sigma_ab^2 = s_a s_b, 4 eps_ab = e_a e_b
// and 332.06 q_a q_b = q_a q_b (SimulationStep.cpp:344-375 with its
// geometric mixing sigma_ab = sqrt(|sigma_a sigma_b|), eps_ab likewise).
struct SiteParam64 {
  double s, e, q, real;
};
struct SiteParam32 {
  float s, e, q, real;
};
constexpr int kParamCap = 128;  // distinct site-parameter entries per box

// Per-molecule record read with one 16-byte load: first site, site count,
// offset of the first primary site within the molecule, first entry in the
// site-parameter table, first entry in the primary list.
struct alignas(16) MolMeta {
  int32_t start;
  int16_t len;
  int16_t p0;
  int32_t param;
  int32_t pstart;
};

// Box arrays are global memory.  Kernels read their box descriptors from
// memory (the batch table), and a pointer read back from memory has no
// known address space: the compiler would reach the arrays with flat
// instructions, which also count against the LDS wait counter (an LDS read
// after one waits for it).  Device code therefore declares the descriptor's
// pointers global.
#if defined(__HIP_DEVICE_COMPILE__)
#define MCGPU_GMEM __attribute__((address_space(1)))
#else
#define MCGPU_GMEM
#endif

// Host code fills the descriptors from plain device pointers.
template <class T>
inline MCGPU_GMEM T* gmem(T* p) {
  return (MCGPU_GMEM T*)p;
}

struct CellGrid {
  int32_t n[3];       // cells per dimension (>= 1)
  int32_t cap;        // slots per cell
  double inv_w[3];    // 1 / cell width
  MCGPU_GMEM int32_t* count;     // [ncell]
  MCGPU_GMEM int32_t* slots;     // [ncell * cap]
  MCGPU_GMEM int32_t* mol_cell;  // [n_mol]
  MCGPU_GMEM int32_t* mol_slot;  // [n_mol]
  MCGPU_GMEM int32_t* overflow;  // [1]: nonzero -> lists incomplete, scan all molecules
};

struct VerletList {
  int32_t cap;        // entries per molecule
  double skin_half2;  // (skin/2)^2
  MCGPU_GMEM int32_t* count;     // [n_mol]
  MCGPU_GMEM int32_t* list;      // [n_mol * cap]
  MCGPU_GMEM double* ref;        // [3*n_mol] first-primary position at build
  MCGPU_GMEM int32_t* stale;     // [1]: nonzero -> some molecule moved > skin/2
};

#define MCGPU_NB_BRUTE 0
#define MCGPU_NB_CELL 1
#define MCGPU_NB_VERLET 2

// Fault codes a kernel records (and then stops the box) when a bound that
// the data should always satisfy does not hold.
#define MCGPU_FAULT_CANDIDATES 1  // neighbour source claims more candidates than molecules
#define MCGPU_FAULT_MOLLEN 2      // molecule length outside [1, kMaxSites]

struct BoxView {
  MCGPU_GMEM double* x;
  MCGPU_GMEM double* y;
  MCGPU_GMEM double* z;
  MCGPU_GMEM double* px;
  MCGPU_GMEM double* py;
  MCGPU_GMEM double* pz;
  const MCGPU_GMEM int32_t* mol_start;
  const MCGPU_GMEM int32_t* mol_len;
  const MCGPU_GMEM int32_t* mol_pidx_start;
  const MCGPU_GMEM int32_t* mol_pidx_count;
  const MCGPU_GMEM int32_t* mol_param;
  const MCGPU_GMEM MolMeta* meta;  // [n_mol]
  const MCGPU_GMEM int32_t* pidx;
  const MCGPU_GMEM SiteParam64* p64;
  const MCGPU_GMEM SiteParam32* p32;
  int32_t n_mol;
  int32_t n_site;
  int32_t max_len;
  int32_t mode;  // MCGPU_NB_*
  int32_t single_primary;  // every molecule j has one primary, entry j
  // Uniform boxes (every molecule alike: uni_len sites starting at
  // m * uni_len, parameters from entry uni_param, one primary at offset
  // uni_p0): a molecule's record follows from its index without a load.
  int32_t uni_len;  // 0 when the box is not uniform
  int32_t uni_param;
  int32_t uni_p0;
  int32_t n_param;  // entries of the site-parameter tables
  int32_t pad0;
  double L[3];
  double cutoff2;
  // Twice the largest distance of a site from its molecule's first primary,
  // plus 1e-6 A for rounding: bounds, per axis, how far a site pair's
  // displacement lies from the displacement of the two first primaries.
  double reach2;
  double kT;
  CellGrid cells;
  VerletList verlet;
};

// ---------------------------------------------------------------- math

// SimulationStep.cpp:377-385: one wrap per dimension, d + L when d < -L/2,
// d - L when d > L/2 (the second test can only fire when the first did
// not), as one FMA with k in {1, 0, -1}: k L is exact, so the sum rounds
// exactly as the reference's addition.
__device__ __forceinline__ double wrap1(double d, double L) {
  const double h = 0.5 * L;
  const double k = d < -h ? 1.0 : (d > h ? -1.0 : 0.0);
  return fma(k, L, d);
}

// The wrap of wrap1 for every site pair of a molecule pair at once, from
// the raw displacement d of their first primaries along one axis (h = L/2,
// reach = BoxView::reach2): every site pair's displacement lies within
// reach of d, so it is 0 (no wrap) when d stays within h - reach of zero,
// 1 (+L) when d < -(h + reach), 2 (-L) when d > h + reach, and 3 (site
// pairs may wrap differently) otherwise.
__device__ __forceinline__ int image_class(double d, double h, double reach) {
  if (d > h + reach) return 2;
  if (d < -(h + reach)) return 1;
  return (d <= h - reach && d >= reach - h) ? 0 : 3;
}

// The three axes' classes packed 2 bits each (x lowest), or kImageMixed
// when some axis is 3.
constexpr int kImageMixed = 1 << 6;
__device__ __forceinline__ int image_code(double dx, double dy, double dz,
                                          const double L[3], double reach) {
  const int cx = image_class(dx, 0.5 * L[0], reach);
  const int cy = image_class(dy, 0.5 * L[1], reach);
  const int cz = image_class(dz, 0.5 * L[2], reach);
  return (cx == 3 || cy == 3 || cz == 3) ? kImageMixed : (cx | cy << 2 | cz << 4);
}

// The shift image_class stands for (applied to the far molecule's sites, so
// a site pair's displacement differs from wrap1's only in the rounding of
// coordinates that are shifted by L).
__device__ __forceinline__ double image_shift(int c, double L) {
  return c == 1 ? L : (c == 2 ? -L : 0.0);
}

// 1/sqrt(x) in FP64: the hardware estimate y (v_rsq_f64, relative error
// up to ~2^-22) refined by one third-order step of the binomial series
// 1/sqrt(x) = y (1 - e)^(-1/2) = y (1 + e/2 + 3e^2/8 + ...), e = 1 - x y^2
// computed by an FMA.  The dropped terms are below 2^-64; with every
// operation rounded once, start errors up to 2^-22 leave the result within
// 0.95 ulp of 1/sqrt(x) in an exact rational simulation of these steps, in
// five FP64 operations (two Newton steps take seven).
__device__ __forceinline__ double rsqrt64(double x) {
  const double y = __builtin_amdgcn_rsq(x);
  const double e = fma(-(x * y), y, 1.0);
  return fma(y * e, fma(e, 0.375, 0.5), y);
}

// FP64 LJ + Coulomb of one site pair from the wrapped displacement
// (SimulationStep.cpp:344-367; BruteForceStep.cpp:64-73): with s2 =
// sigma_ab^2 / r^2, LJ = 4 eps_ab (s2^6 - s2^3) and Coulomb = 332.06 q_a q_b
// / r.  r^2 == 0 contributes nothing; dummy sites (all-zero parameters)
// contribute 0.
__device__ __forceinline__ double pair_energy64(double dx, double dy, double dz,
                                                double sa, double ea, double qa,
                                                double sb, double eb, double qb) {
  const double r2 = dx * dx + dy * dy + dz * dz;
  const double rinv = rsqrt64(r2);
  const double s2 = (sa * sb) * (rinv * rinv);
  const double s6 = s2 * s2 * s2;
  const double e = (ea * eb) * fma(s6, s6, -s6) + (qa * qb) * rinv;
  return r2 == 0.0 ? 0.0 : e;
}

// Mixed precision: displacement wrapped in FP64; 1/r estimated in FP32
// (v_rsq_f32, ~1 ulp) and LJ computed from it in FP32; the result handed
// back for FP64 accumulation.
//  * Near-contact LJ pairs (s^6 > 4, r < 0.79 sigma) are re-evaluated in
//    FP64: there the FP32 rounding of s^12 (~20 ulp) would reach
//    1e-5..1e-3 kcal/mol per pair.
//  * Coulomb takes the FP32 1/r refined by one FP64 Newton step (relative
//    error ~1e-14) and an FP64 product.  Coulomb terms are ~10-100
//    kcal/mol per pair, so FP32 products (~1 ulp each) summed over the
//    ~2000 pairs of a water move put |dE| errors up to ~1.6e-4 kcal/mol;
//    with this step the per-move error is that of the FP32 LJ terms alone.
constexpr float kNearContactS6 = 4.0f;

__device__ __forceinline__ double pair_energy32(double dx, double dy, double dz,
                                                float sa, float ea, float sb, float eb,
                                                double sa64, double ea64, double qa64,
                                                double sb64, double eb64, double qb64) {
  const double r2d = dx * dx + dy * dy + dz * dz;
  const float r2 = (float)r2d;
  const float rinv = __builtin_amdgcn_rsqf(r2);
  const float s2 = (sa * sb) * (rinv * rinv);
  const float s6 = s2 * s2 * s2;
  const float eps = ea * eb;
  double rd = (double)rinv;
  rd = rd * fma(-0.5 * r2d * rd, rd, 1.5);
  const double coul = (qa64 * qb64) * rd;
  double lj;
  if (s6 > kNearContactS6 && eps != 0.0f) {
    const double s2d = (sa64 * sb64) / r2d;
    const double s6d = s2d * s2d * s2d;
    lj = (ea64 * eb64) * (s6d * s6d - s6d);
  } else {
    lj = (double)(eps * (s6 * s6 - s6));
  }
  return r2d == 0.0 ? 0.0 : lj + coul;
}

// ------------------------------------------------------- wave / block ops

__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }
__device__ __forceinline__ int wave_id() { return threadIdx.x >> 6; }

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Sum over the lanes of a wave of per-wave partials held in LDS (lane w
// reads partial w), in a fixed order; every lane gets the total.
template <int NW, class T>
__device__ __forceinline__ T wave_sum_partials(const T* lds) {
  T v = lane_id() < NW ? lds[lane_id()] : T(0);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  return v;
}

// Block sum in a fixed order (deterministic for a given NT).
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* lds) {
  v = wave_sum(v);
  if (lane_id() == 0) lds[wave_id()] = v;
  __syncthreads();
  const double t = wave_sum_partials<NT / 64>(lds);
  __syncthreads();
  return t;
}

// Exclusive prefix of a per-thread flag across the block, in thread order.
template <int NT>
__device__ __forceinline__ int block_scan_flag(bool f, int* lds, int& total) {
  const unsigned long long m = __ballot(f);
  const int lane = lane_id();
  const int below = __popcll(m & ((1ull << lane) - 1ull));
  if (lane == 0) lds[wave_id()] = __popcll(m);
  __syncthreads();
  int v = lane < NT / 64 ? lds[lane] : 0;  // inclusive scan of wave counts
#pragma unroll
  for (int o = 1; o < NT / 64; o <<= 1) {
    const int u = __shfl_up(v, o);
    if (lane >= o) v += u;
  }
  const int incl_prev = __shfl(v, wave_id() > 0 ? wave_id() - 1 : 0);
  total = __shfl(v, NT / 64 - 1);
  __syncthreads();
  return (wave_id() > 0 ? incl_prev : 0) + below;
}

// ------------------------------------------------------------ cell grid

// Cell of a first-primary coordinate along one axis (n cells of width
// 1/inv_w spanning the box edge).  Coordinates handed in through the ABI need
// not lie in [0, L): a position within one box length outside falls in the
// cell of its periodic image, so the 27 cells around it still hold every
// candidate in range.  The final clamp keeps any other input on the grid.
__device__ __forceinline__ int cell_coord(double p, double inv_w, int n) {
  int c = (int)floor(p * inv_w);
  c = c < 0 ? c + n : (c >= n ? c - n : c);
  return c < 0 ? 0 : (c >= n ? n - 1 : c);
}

__device__ __forceinline__ int cell_index(const CellGrid& g, int cx, int cy,
                                          int cz) {
  return (cx * g.n[1] + cy) * g.n[2] + cz;
}

// A range of cells per dimension covering the neighbourhoods of up to two
// positions (old and proposed); never wraps onto itself.
struct CellRange {
  int lo[3];
  int cnt[3];
  __device__ int total() const { return cnt[0] * cnt[1] * cnt[2]; }
  // k < 64 (callers hold at most 64 cells): the quotients by the small
  // counts come from the hardware reciprocal (v_rcp_f32, relative error
  // about 2^-23), (k + 0.5) / m being at least 0.5 / m from an integer, far
  // beyond that error
  __device__ int cell(const CellGrid& g, int k) const {
    const int t = (int)(((float)k + 0.5f) * __builtin_amdgcn_rcpf((float)cnt[2]));
    const int c2 = k - t * cnt[2];
    const int c0 = (int)(((float)t + 0.5f) * __builtin_amdgcn_rcpf((float)cnt[1]));
    const int c1 = t - c0 * cnt[1];
    int a = lo[0] + c0, b = lo[1] + c1, c = lo[2] + c2;
    a += (a < 0) ? g.n[0] : 0;
    a -= (a >= g.n[0]) ? g.n[0] : 0;
    b += (b < 0) ? g.n[1] : 0;
    b -= (b >= g.n[1]) ? g.n[1] : 0;
    c += (c < 0) ? g.n[2] : 0;
    c -= (c >= g.n[2]) ? g.n[2] : 0;
    return cell_index(g, a, b, c);
  }
};

// Cells around an old cell co and a proposed cell cn.  Per dimension the
// proposed cell is taken at its periodic (minimum-image) offset from co, so
// a molecule crossing the box boundary spans {co-1 .. cn+1} across the wrap
// (4 cells) rather than the whole row; when the span covers the row the row
// is taken once.  A span over 4 cells (a jump of 2+ cells) cannot occur for
// moves shorter than a cell; it is reported as a total over 64, and the
// caller then scans every molecule.
__device__ __forceinline__ CellRange make_range(const CellGrid& g,
                                                const int co[3],
                                                const int cn[3]) {
  CellRange r;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int n = g.n[d];
    int delta = cn[d] - co[d];
    if (2 * delta > n) delta -= n;
    if (2 * delta < -n) delta += n;
    const int lo = co[d] + (delta < 0 ? delta : 0) - 1;
    const int hi = co[d] + (delta > 0 ? delta : 0) + 1;
    const int c = hi - lo + 1;
    if (c >= n) {
      r.lo[d] = 0;
      r.cnt[d] = n;
    } else {
      r.lo[d] = lo;
      r.cnt[d] = c;
    }
  }
  return r;
}

// CellRange::cell wraps each coordinate once, which is exact because
// lo >= -n and lo + cnt - 1 < 2 n for every range make_range builds.

}  // namespace mcgpu
