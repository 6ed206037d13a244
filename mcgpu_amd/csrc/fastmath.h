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

// sin and cos of x for |x| < 2^21 (trial-move angles are within +-2 pi).
__device__ __forceinline__ void sincos_fp64(double x, double& s, double& c) {
  const double k = rint(x * kTwoOverPi);
  double r = fma(-k, kPio2A, x);
  r = fma(-k, kPio2B, r);
  r = fma(-k, kPio2C, r);  // |r| <= pi/4
  const double r2 = r * r;
  // sin r = r (1 - r^2/3! + r^4/5! - ... - r^16/17!)... as r + r^3 P(r^2)
  double ps = 2.8114572543455206e-15;        //  1/17!
  ps = fma(ps, r2, -7.647163731819816e-13);  // -1/15!
  ps = fma(ps, r2, 1.6059043836821613e-10);  //  1/13!
  ps = fma(ps, r2, -2.505210838544172e-08);  // -1/11!
  ps = fma(ps, r2, 2.7557319223985893e-06);  //  1/9!
  ps = fma(ps, r2, -0.0001984126984126984);  // -1/7!
  ps = fma(ps, r2, 0.008333333333333333);    //  1/5!
  ps = fma(ps, r2, -0.16666666666666666);    // -1/3!
  const double sr = fma(ps * r2, r, r);
  // cos r = 1 - r^2/2! + ... + r^18/18! as 1 + r^2 Q(r^2)
  double pc = -1.5619206968586225e-16;       // -1/18!
  pc = fma(pc, r2, 4.779477332387385e-14);   //  1/16!
  pc = fma(pc, r2, -1.1470745597729725e-11); // -1/14!
  pc = fma(pc, r2, 2.08767569878681e-09);    //  1/12!
  pc = fma(pc, r2, -2.755731922398589e-07);  // -1/10!
  pc = fma(pc, r2, 2.48015873015873e-05);    //  1/8!
  pc = fma(pc, r2, -0.001388888888888889);   // -1/6!
  pc = fma(pc, r2, 0.041666666666666664);    //  1/4!
  pc = fma(pc, r2, -0.5);                    // -1/2!
  const double cr = fma(pc, r2, 1.0);
  const int q = ((int)k) & 3;
  s = q == 0 ? sr : (q == 1 ? cr : (q == 2 ? -sr : -cr));
  c = q == 0 ? cr : (q == 1 ? -sr : (q == 2 ? -cr : sr));
}

// e^x for finite x (NaN passes through, underflows to 0, overflows to inf).
__device__ __forceinline__ double exp_fp64(double x) {
  if (!(x == x)) return x;
  if (x < -745.2) return 0.0;
  if (x > 709.8) return __builtin_huge_val();
  const double k = rint(x * kInvLn2);
  double r = fma(-k, kLn2A, x);
  r = fma(-k, kLn2B, r);  // |r| <= ln2/2
  double p = 1.6059043836821613e-10;  // 1/13!
  p = fma(p, r, 2.08767569878681e-09);   // 1/12!
  p = fma(p, r, 2.505210838544172e-08);  // 1/11!
  p = fma(p, r, 2.755731922398589e-07);  // 1/10!
  p = fma(p, r, 2.7557319223985893e-06); // 1/9!
  p = fma(p, r, 2.48015873015873e-05);   // 1/8!
  p = fma(p, r, 0.0001984126984126984);  // 1/7!
  p = fma(p, r, 0.001388888888888889);   // 1/6!
  p = fma(p, r, 0.008333333333333333);   // 1/5!
  p = fma(p, r, 0.041666666666666664);   // 1/4!
  p = fma(p, r, 0.16666666666666666);    // 1/3!
  p = fma(p, r, 0.5);                    // 1/2!
  p = fma(p, r, 1.0);
  p = fma(p, r, 1.0);
  return ldexp(p, (int)k);
}

}  // namespace mcgpu
