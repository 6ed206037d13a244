// Wavefront primitives for CDNA (wave64): DPP reductions and scans, lane
// reads of 64-bit values, and the glibc TYPE_3 generator held across the
// lanes of one wave.
//
// DPP (data-parallel primitives) moves values between lanes inside the VALU
// pipeline, so a reduction step costs a few cycles instead of an LDS round
// trip (ds_bpermute, which __shfl_* compiles to).  Controls used (GFX9 DPP
// encoding): quad_perm [1,0,3,2] = 0xb1, quad_perm [2,3,0,1] = 0x4e,
// row_shr:n = 0x110 | n, row_ror:n = 0x120 | n.  Every DPP move here starts
// from 0 in the destination, so a lane whose source lies outside its row
// reads 0.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mcgpu {

__device__ __forceinline__ int wlane() { return threadIdx.x & 63; }

template <int CTRL>
__device__ __forceinline__ int dpp_i32(int v) {
  return __builtin_amdgcn_update_dpp(0, v, CTRL, 0xf, 0xf, false);
}

template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = dpp_i32<CTRL>((int)(b & 0xffffffffll));
  const int hi = dpp_i32<CTRL>((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Value of lane `lane` (wave-uniform index) for every lane.
__device__ __forceinline__ int readlane_i32(int v, int lane) {
  return __builtin_amdgcn_readlane(v, lane);
}
__device__ __forceinline__ double readlane_f64(double v, int lane) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}
__device__ __forceinline__ unsigned long long readlane_u64(unsigned long long v,
                                                           int lane) {
  const unsigned lo = __builtin_amdgcn_readlane((int)(v & 0xffffffffull), lane);
  const unsigned hi = __builtin_amdgcn_readlane((int)(v >> 32), lane);
  return ((unsigned long long)hi << 32) | lo;
}

// Sum over the 64 lanes in a fixed order; the result is wave-uniform.
// Within each 16-lane row: pairs, quads, then rotations by 4 and 8 leave the
// row sum in every lane of the row; the four row sums (read at lanes 0, 16,
// 32, 48) are added as (r0 + r1) + (r2 + r3).
__device__ __forceinline__ double wave_sum_f64(double v) {
  v += dpp_f64<0xb1>(v);
  v += dpp_f64<0x4e>(v);
  v += dpp_f64<0x124>(v);
  v += dpp_f64<0x128>(v);
  const double r0 = readlane_f64(v, 0), r1 = readlane_f64(v, 16);
  const double r2 = readlane_f64(v, 32), r3 = readlane_f64(v, 48);
  return (r0 + r1) + (r2 + r3);
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
  v += (unsigned)dpp_i32<0xb1>((int)v);
  v += (unsigned)dpp_i32<0x4e>((int)v);
  v += (unsigned)dpp_i32<0x124>((int)v);
  v += (unsigned)dpp_i32<0x128>((int)v);
  return (unsigned)readlane_i32((int)v, 0) + (unsigned)readlane_i32((int)v, 16) +
         (unsigned)readlane_i32((int)v, 32) + (unsigned)readlane_i32((int)v, 48);
}

// Inclusive prefix sum over the 64 lanes (lane order).  Rows are scanned by
// shifts of 1, 2, 4, 8 (Hillis-Steele); the totals of the preceding rows are
// then added from lanes 15, 31, 47.
__device__ __forceinline__ int wave_incl_scan_i32(int v) {
  v += dpp_i32<0x111>(v);
  v += dpp_i32<0x112>(v);
  v += dpp_i32<0x114>(v);
  v += dpp_i32<0x118>(v);
  const int t0 = readlane_i32(v, 15);
  const int t1 = readlane_i32(v, 31);
  const int t2 = readlane_i32(v, 47);
  const int row = wlane() >> 4;
  v += (row >= 1 ? t0 : 0) + (row >= 2 ? t1 : 0) + (row >= 3 ? t2 : 0);
  return v;
}

// Number of set flags in lanes below this one (ballot + mbcnt).
__device__ __forceinline__ int wave_rank(unsigned long long ballot) {
  return __builtin_amdgcn_mbcnt_hi((unsigned)(ballot >> 32),
                                   __builtin_amdgcn_mbcnt_lo((unsigned)ballot, 0u));
}

// glibc random_r TYPE_3 state spread over one wave: lane k (k < 31) holds
// state word k; f (front) and r (rear = f - 3 mod 31) are wave-uniform.
// A draw is two lane reads, an add and a lane-masked select: the same words
// and order as rng31_next, without a memory round trip.
struct WaveRng {
  int s;
  int f;
  int r;
};

__device__ __forceinline__ int wrng_next(WaveRng& g) {
  const unsigned a = (unsigned)__builtin_amdgcn_readlane(g.s, g.f);
  const unsigned b = (unsigned)__builtin_amdgcn_readlane(g.s, g.r);
  const unsigned v = a + b;
  g.s = wlane() == g.f ? (int)v : g.s;
  g.f = g.f == 30 ? 0 : g.f + 1;
  g.r = g.r == 30 ? 0 : g.r + 1;
  return (int)(v >> 1);
}

}  // namespace mcgpu
