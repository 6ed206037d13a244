// Small-register FP64 sin/cos/exp for the serial parts of the batched move
// (the rotation of a trial move and the Metropolis exp(-dE/kT)).  Standard
// methods: Cody-Waite argument reduction with the constant split into parts
// whose products with the reduction multiple are exact, then Taylor
// polynomials on the reduced interval evaluated by Horner's rule with FMA.
// Errors are about 1 ulp, like the host libm, so trial moves agree with the
// host's to the last bit or two.
#pragma once

#include <hip/hip_runtime.h>

namespace mcgpu {

// pi/2 = kPio2A + kPio2B + kPio2C; A and B carry 30 significant bits, so
// k * A and k * B are exact for |k| < 2^22.
constexpr double kPio2A = 1.5707963267341256;
constexpr double kPio2B = 6.07710050359346e-11;
constexpr double kPio2C = 2.912732056093356e-20;
constexpr double kTwoOverPi = 0.63661977236758134;
// ln 2 = kLn2A + kLn2B (A with 30 significant bits).
constexpr double kLn2A = 0.6931471801362932;
constexpr double kLn2B = 4.236521365809284e-10;
constexpr double kInvLn2 = 1.4426950408889634;

// Coefficient tables in constant memory.  The functions read them through a
// pointer the optimiser cannot see through, so the coefficients are loaded
// (scalar loads) where they are used rather than held in registers across
// the caller's loops, where they would crowd out the loop's own values.
__constant__ double kSinCosTab[24] = {
    kTwoOverPi, kPio2A, kPio2B, kPio2C,
    // sin: 1/17!, -1/15!, 1/13!, -1/11!, 1/9!, -1/7!, 1/5!, -1/3!
    2.8114572543455206e-15, -7.647163731819816e-13, 1.6059043836821613e-10,
    -2.505210838544172e-08, 2.7557319223985893e-06, -0.0001984126984126984,
    0.008333333333333333, -0.16666666666666666,
    // cos: -1/18!, 1/16!, -1/14!, 1/12!, -1/10!, 1/8!, -1/6!, 1/4!, -1/2!
    -1.5619206968586225e-16, 4.779477332387385e-14, -1.1470745597729725e-11,
    2.08767569878681e-09, -2.755731922398589e-07, 2.48015873015873e-05,
    -0.001388888888888889, 0.041666666666666664, -0.5, 0.0};
__constant__ double kExpTab[20] = {
    kInvLn2, kLn2A, kLn2B, -745.2, 709.8,
    // 1/13! ... 1/2!, 1, 1
    1.6059043836821613e-10, 2.08767569878681e-09, 2.505210838544172e-08,
    2.755731922398589e-07, 2.7557319223985893e-06, 2.48015873015873e-05,
    0.0001984126984126984, 0.001388888888888889, 0.008333333333333333,
    0.041666666666666664, 0.16666666666666666, 0.5, 1.0, 1.0, 0.0};

typedef const __attribute__((address_space(4))) double* ConstTab;

__device__ __forceinline__ ConstTab opaque_tab(ConstTab p) {
  asm volatile("" : "+s"(p));
  return p;
}

// sin and cos of x for |x| < 2^21 (trial-move angles are within +-2 pi).
__device__ __forceinline__ void sincos_fp64(double x, double& s, double& c) {
  const ConstTab T = opaque_tab((ConstTab)kSinCosTab);
  const double k = rint(x * T[0]);
  double r = fma(-k, T[1], x);
  r = fma(-k, T[2], r);
  r = fma(-k, T[3], r);  // |r| <= pi/4
  const double r2 = r * r;
  // sin r = r + r^3 P(r^2), Taylor terms to r^17
  double ps = T[4];
#pragma unroll
  for (int i = 5; i < 12; i++) ps = fma(ps, r2, T[i]);
  const double sr = fma(ps * r2, r, r);
  // cos r = 1 + r^2 Q(r^2), Taylor terms to r^18
  double pc = T[12];
#pragma unroll
  for (int i = 13; i < 21; i++) pc = fma(pc, r2, T[i]);
  const double cr = fma(pc, r2, 1.0);
  const int q = ((int)k) & 3;
  s = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
  c = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}

// e^x for finite x (NaN passes through, underflows to 0, overflows to inf).
__device__ __forceinline__ double exp_fp64(double x) {
  const ConstTab T = opaque_tab((ConstTab)kExpTab);
  if (!(x == x)) return x;
  if (x < T[3]) return 0.0;
  if (x > T[4]) return __builtin_huge_val();
  const double k = rint(x * T[0]);
  double r = fma(-k, T[1], x);
  r = fma(-k, T[2], r);  // |r| <= ln2/2
  double p = T[5];
#pragma unroll
  for (int i = 6; i < 19; i++) p = fma(p, r, T[i]);
  return ldexp(p, (int)k);
}

}  // namespace mcgpu
