// nsm_fp64_math.h -- FP64 per-point arithmetic of the parity path.
//
// These device functions reproduce NimbleSM's operation order exactly (the
// file including this header is compiled with -ffp-contract=off), so that the
// GPU result equals the reference's FP64 result bit for bit.  Citations are
// NimbleSM file:line.
#pragma once

#include <hip/hip_runtime.h>

namespace nsm {
namespace fp64 {

// symmetric / full tensor storage, src/nimble_utils.h:91-109
enum { S_XX = 0, S_YY = 1, S_ZZ = 2, S_XY = 3, S_YZ = 4, S_ZX = 5 };
enum { F_XX = 0, F_YY = 1, F_ZZ = 2, F_XY = 3, F_YZ = 4, F_ZX = 5, F_YX = 6, F_ZY = 7, F_XZ = 8 };

// ---------------------------------------------------------------------------
// Quotients over one denominator.  The device's IEEE division of n by d
// (the compiler's sequence for gfx950) is
//   ds = div_scale(d), ns = div_scale(n) (with flag vcc), r = rcp(ds),
//   two Newton steps r = fma(r, fma(-ds, r, 1), r), q = ns r,
//   q = div_fmas(fma(-ds, q, ns), r, q, vcc), result = div_fixup(q, d, n).
// The scale steps are identities (ds = d, ns = n, vcc = 0) whenever d and n
// are finite, nonzero and normal with exponents far from the ends of the
// range -- here: frexp exponents in [-299, 300] -- and div_fixup settles
// zero, infinite and NaN operands from d and n alone, whatever quotient it
// is given.  So for such operands r depends on d only, and the quotients
// over one denominator share it: the same operations on the same values,
// hence the same bits as dividing one by one, with one reciprocal (a
// quarter-rate instruction) and its refinement per denominator instead of
// per quotient.  A wave whose operands are not all of this kind divides
// one by one.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double refined_rcp(double d)
{
  double r = __builtin_amdgcn_rcp(d);
  double t = __builtin_fma(-d, r, 1.0);
  r        = __builtin_fma(r, t, r);
  t        = __builtin_fma(-d, r, 1.0);
  return __builtin_fma(r, t, r);
}

// n / d given r = refined_rcp(d), for operands that pass plain_operand
__device__ __forceinline__ double div_shared(double n, double d, double r)
{
  const double q = n * r;
  const double e = __builtin_fma(-d, q, n);
  return __builtin_amdgcn_div_fixup(__builtin_amdgcn_div_fmas(e, r, q, false), d, n);
}

// zero, infinite, NaN, or |x| in [2^-300, 2^300) (frexp exponent of
// zero, infinity and NaN is 0)
__device__ __forceinline__ bool plain_operand(double x)
{
  return (unsigned)(__builtin_amdgcn_frexp_exp(x) + 299) < 600u;
}

// q[i] = n[i] / d for i < kN, bit for bit as kN separate divisions
template <int kN>
__device__ __forceinline__ void divide_by(const double (&n)[kN], double d, double (&q)[kN])
{
  bool plain = plain_operand(d);
#pragma unroll
  for (int i = 0; i < kN; i++) plain = plain && plain_operand(n[i]);
  if (__builtin_expect(__ballot(!plain) == 0, 1)) {  // uniform across the wave
    const double r = refined_rcp(d);
#pragma unroll
    for (int i = 0; i < kN; i++) q[i] = div_shared(n[i], d, r);
  } else {
#pragma unroll
    for (int i = 0; i < kN; i++) q[i] = n[i] / d;
  }
}

// Invert3x3 on a row-major 3x3 (src/nimble_utils.h:1229-1268); returns det.
// kShared = false divides one by one: the inverse of the reference Jacobian,
// formed while the current Jacobian is held, is cheaper in registers that way.
template <bool kShared = true>
__device__ __forceinline__ double invert3x3(const double m[9], double inv[9])
{
  const double minor0 = m[4] * m[8] - m[5] * m[7];
  const double minor1 = m[3] * m[8] - m[5] * m[6];
  const double minor2 = m[3] * m[7] - m[4] * m[6];
  const double minor3 = m[1] * m[8] - m[2] * m[7];
  const double minor4 = m[0] * m[8] - m[6] * m[2];
  const double minor5 = m[0] * m[7] - m[1] * m[6];
  const double minor6 = m[1] * m[5] - m[2] * m[4];
  const double minor7 = m[0] * m[5] - m[2] * m[3];
  const double minor8 = m[0] * m[4] - m[1] * m[3];
  const double det    = m[0] * minor0 - m[1] * minor1 + m[2] * minor2;
  const double n[9]   = {minor0,        -1.0 * minor3, minor6, -1.0 * minor1, minor4,
                         -1.0 * minor7, minor2,        -1.0 * minor5, minor8};
  if (kShared) {
    double q[9];
    divide_by(n, det, q);
#pragma unroll
    for (int i = 0; i < 9; i++) inv[i] = q[i];
  } else {
#pragma unroll
    for (int i = 0; i < 9; i++) inv[i] = n[i] / det;
  }
  return det;
}

// Invert_Full33 on K_F storage (src/nimble_utils.h:523-551).
__device__ __forceinline__ double invert_full33(const double* a, double* inv)
{
  const double minor0 = a[F_YY] * a[F_ZZ] - a[F_YZ] * a[F_ZY];
  const double minor1 = a[F_YX] * a[F_ZZ] - a[F_YZ] * a[F_ZX];
  const double minor2 = a[F_YX] * a[F_ZY] - a[F_YY] * a[F_ZX];
  const double minor3 = a[F_XY] * a[F_ZZ] - a[F_XZ] * a[F_ZY];
  const double minor4 = a[F_XX] * a[F_ZZ] - a[F_ZX] * a[F_XZ];
  const double minor5 = a[F_XX] * a[F_ZY] - a[F_XY] * a[F_ZX];
  const double minor6 = a[F_XY] * a[F_YZ] - a[F_XZ] * a[F_YY];
  const double minor7 = a[F_XX] * a[F_YZ] - a[F_XZ] * a[F_YX];
  const double minor8 = a[F_XX] * a[F_YY] - a[F_XY] * a[F_YX];
  const double det    = a[F_XX] * minor0 - a[F_XY] * minor1 + a[F_XZ] * minor2;
  const double n[9]   = {minor0,        -1.0 * minor3, minor6, -1.0 * minor1, minor4,
                         -1.0 * minor7, minor2,        -1.0 * minor5, minor8};
  double       q[9];
  divide_by(n, det, q);
  inv[F_XX] = q[0];
  inv[F_XY] = q[1];
  inv[F_XZ] = q[2];
  inv[F_YX] = q[3];
  inv[F_YY] = q[4];
  inv[F_YZ] = q[5];
  inv[F_ZX] = q[6];
  inv[F_ZY] = q[7];
  inv[F_ZZ] = q[8];
  return det;
}

// Cube root by the method of the platform libm the reference links (GNU libc
// dbl-64 cbrt, Alboth & Drepper 1997): reduce |x| = xm 2^xe with xm in
// [0.5, 1), a degree-6 polynomial start u(xm), one rational correction
// u (u^3 + 2 xm) / (2 u^3 + xm), a 2^{(xe mod 3)/3} factor and an exact
// scaling by 2^{xe/3}.  That libm cbrt is not correctly rounded (about 45% of
// arguments near 1 are 1-3 ulp off), and NimbleSM's stress calls it
// (src/nimble_material.cc:274), so the parity path evaluates the same steps
// with the same constants to reproduce the reference's FP64 bits.
__device__ __forceinline__ double cbrt_ref(double x)
{
  if (x == 0.0 || isnan(x) || isinf(x)) return x + x;
  int          xe;
  const double xm = frexp(fabs(x), &xe);
  const double u  = (0.354895765043919860 +
                    ((1.50819193781584896 +
                      ((-2.11499494167371287 +
                        ((2.44693122563534430 +
                          ((-1.83469277483613086 + (0.784932344976639262 - 0.145263899385486377 * xm) * xm) * xm)) *
                         xm)) *
                       xm)) *
                     xm));
  const double t2  = u * u * u;
  const double cb2 = 1.2599210498948731648;   // 2^(1/3)
  const double sq2 = 1.5874010519681994748;   // 2^(2/3)
  const int    r   = xe % 3;                  // C remainder: -2..2
  const double fac = r == -2 ? 1.0 / sq2 : r == -1 ? 1.0 / cb2 : r == 0 ? 1.0 : r == 1 ? cb2 : sq2;
  const double ym  = u * (t2 + 2.0 * xm) / (2.0 * t2 + xm) * fac;
  return ldexp(x > 0.0 ? ym : -ym, xe / 3);
}

// x times the sign of y (src/nimble_utils.h:131-136)
__device__ __forceinline__ double mult_sign(double x, double y) { return x * (double)(1 - 2 * (y < 0.0)); }

// 6/5 minimax rational approximation of cos(acos(x)/3) on [0,1]
// (coefficients as published in src/nimble_utils.h:650-665).
__device__ __forceinline__ double cos_acos_div3(double x)
{
  const double x2 = x * x;
  const double x4 = x2 * x2;
  const double num = 0.866025403784438713 + 2.12714890259493060 * x +
                     ((1.89202064815951569 + 0.739603278343401613 * x) * x2 +
                      (0.121973926953064794 + x * (0.00655637626263929360 + 0.0000390884982780803443 * x)) * x4);
  const double den = 1.0 + 2.26376989330935617 * x +
                     ((1.80461009751278976 + 0.603976798217196003 * x) * x2 +
                      (0.0783255761115461708 + 0.00268525944538021629 * x) * x4);
  return num / den;
}

// Closed-form eigen-decomposition of a symmetric 3x3 with column-pivoted QR
// and a Wilkinson shift (src/nimble_utils.h:667-857).  Outputs unnormalised
// eigenvectors e0, e1, e2 for eigenvalues ev[0..2].
__device__ __forceinline__ void eigen_sym33(const double t[6], double ev[3], double e0[3], double e1[3], double e2[3])
{
  double       cxx = t[S_XX], cyy = t[S_YY], czz = t[S_ZZ];
  const double cxy = t[S_XY], cyz = t[S_YZ], czx = t[S_ZX];

  const double c1 = (cxx + cyy + czz) / 3.0;
  cxx -= c1;
  cyy -= c1;
  czz -= c1;

  const double cxy_cxy = cxy * cxy;
  const double cyz_cyz = cyz * cyz;
  const double czx_czx = czx * czx;
  const double cxx_cyy = cxx * cyy;

  const double c2 = cxx_cyy + cyy * czz + czz * cxx - cxy_cxy - cyz_cyz - czx_czx;

  const double three_over_a      = -3.0 / c2;
  const double sqrt_three_over_a = sqrt(three_over_a);

  const double c3 = cxx * cyz_cyz + cyy * czx_czx - 2.0 * cxy * cyz * czx + czz * (cxy_cxy - cxx_cyy);
  const double rr = -0.5 * c3 * three_over_a * sqrt_three_over_a;

  const double arg  = fabs(rr) < 1.0 ? fabs(rr) : 1.0;
  const double cth  = cos_acos_div3(arg);
  const double tcth = 2.0 * mult_sign(cth, rr);

  double eval2 = tcth / sqrt_three_over_a;

  const double r0[3] = {cxx - eval2, cxy, czx};
  const double r1[3] = {cxy, cyy - eval2, cyz};
  const double r2[3] = {czx, cyz, czz - eval2};

  const double k0 = r0[0] * r0[0] + cxy_cxy + czx_czx;
  const double k1 = cxy_cxy + r1[1] * r1[1] + cyz_cyz;
  const double k2 = czx_czx + cyz_cyz + r2[2] * r2[2];

  const bool k0gk1 = k1 <= k0;
  const bool k0gk2 = k2 <= k0;
  const bool k1gk2 = k2 <= k1;
  const bool k0l   = k0gk1 && k0gk2;
  const bool k1l   = k1gk2 && !k0gk1;
  const bool k2l   = !(k0l || k1l);

  double kr[3], ra[3], rb[3];
  for (int i = 0; i < 3; i++) kr[i] = (k0l ? r0[i] : 0.0) + (k1l ? r1[i] : 0.0) + (k2l ? r2[i] : 0.0);
  for (int i = 0; i < 3; i++) ra[i] = k0l ? r1[i] : r0[i];
  for (int i = 0; i < 3; i++) rb[i] = k2l ? r1[i] : r2[i];

  const double ki_ki = 1.0 / ((k0l ? k0 : 0.0) + (k1l ? k1 : 0.0) + (k2l ? k2 : 0.0));

  const double ki_dpr1 = ki_ki * (kr[0] * ra[0] + kr[1] * ra[1] + kr[2] * ra[2]);
  const double ki_dpr2 = ki_ki * (kr[0] * rb[0] + kr[1] * rb[1] + kr[2] * rb[2]);
  for (int i = 0; i < 3; i++) ra[i] -= ki_dpr1 * kr[i];
  for (int i = 0; i < 3; i++) rb[i] -= ki_dpr2 * kr[i];

  const double a0 = ra[0] * ra[0] + ra[1] * ra[1] + ra[2] * ra[2];
  const double a1 = rb[0] * rb[0] + rb[1] * rb[1] + rb[2] * rb[2];

  const bool a0lea1 = a0 <= a1;
  double     ar[3];
  for (int i = 0; i < 3; i++) ar[i] = a0lea1 ? rb[i] : ra[i];
  const double ai_ai = 1.0 / (a0lea1 ? a1 : a0);

  e2[0] = kr[1] * ar[2] - kr[2] * ar[1];
  e2[1] = kr[2] * ar[0] - kr[0] * ar[2];
  e2[2] = kr[0] * ar[1] - kr[1] * ar[0];

  const double k_atr11 = cxx * kr[0] + cxy * kr[1] + czx * kr[2];
  const double k_atr21 = cxy * kr[0] + cyy * kr[1] + cyz * kr[2];
  const double k_atr31 = czx * kr[0] + cyz * kr[1] + czz * kr[2];
  const double a_atr12 = cxx * ar[0] + cxy * ar[1] + czx * ar[2];
  const double a_atr22 = cxy * ar[0] + cyy * ar[1] + cyz * ar[2];
  const double a_atr32 = czx * ar[0] + cyz * ar[1] + czz * ar[2];

  double       rm2xx       = (kr[0] * k_atr11 + kr[1] * k_atr21 + kr[2] * k_atr31) * ki_ki;
  const double k_a_rm2xy   = (kr[0] * a_atr12 + kr[1] * a_atr22 + kr[2] * a_atr32);
  double       rm2yy       = (ar[0] * a_atr12 + ar[1] * a_atr22 + ar[2] * a_atr32) * ai_ai;
  const double rm2xy_rm2xy = k_a_rm2xy * k_a_rm2xy * ai_ai * ki_ki;

  // Wilkinson shift
  const double b     = 0.5 * (rm2xx - rm2yy);
  const double sterm = mult_sign(sqrt(b * b + rm2xy_rm2xy), b);
  double       eval0 = rm2yy + b - sterm;
  double       eval1 = rm2xx + rm2yy - eval0;

  rm2xx -= eval0;
  rm2yy -= eval0;
  const double rm2xx2 = rm2xx * rm2xx;
  const double rm2yy2 = rm2yy * rm2yy;

  const double fac1 = rm2xx2 < rm2yy2 ? k_a_rm2xy * ai_ai : rm2xx;
  const double fac2 = rm2xx2 < rm2yy2 ? rm2yy : ki_ki * k_a_rm2xy;

  for (int i = 0; i < 3; i++) e0[i] = fac1 * ar[i] - fac2 * kr[i];
  const bool both_zero = (rm2xx2 == 0.0) && (rm2xy_rm2xy == 0.0);
  for (int i = 0; i < 3; i++) e0[i] = both_zero ? ar[i] : e0[i];

  e1[0] = e2[1] * e0[2] - e2[2] * e0[1];
  e1[1] = e2[2] * e0[0] - e2[0] * e0[2];
  e1[2] = e2[0] * e0[1] - e2[1] * e0[0];

  eval0 += c1;
  eval1 += c1;
  eval2 += c1;

  const double c2tol = (c1 * c1) * (-1.0e-30);
  const bool   neg   = c2 < c2tol;  // otherwise: triple eigenvalue
  ev[0] = neg ? eval0 : c1;
  ev[1] = neg ? eval1 : c1;
  ev[2] = neg ? eval2 : c1;
  e0[0] = neg ? e0[0] : 1.0;
  e0[1] = neg ? e0[1] : 0.0;
  e0[2] = neg ? e0[2] : 0.0;
  e1[0] = neg ? e1[0] : 0.0;
  e1[1] = neg ? e1[1] : 1.0;
  e1[2] = neg ? e1[2] : 0.0;
  e2[0] = neg ? e2[0] : 0.0;
  e2[1] = neg ? e2[1] : 0.0;
  e2[2] = neg ? e2[2] : 1.0;
}

// Left stretch V of F = V R from the eigen-decomposition of F^{-T} F^{-1}
// (Polar_Decomp, src/nimble_utils.h:859-908).  Only V is formed; the
// Neo-Hookean model does not use R.
__device__ __forceinline__ void left_stretch(const double* F, double V[6])
{
  double fi[9];
  invert_full33(F, fi);

  double c[6];  // fi^T fi, Square_Full33T_Full33 (src/nimble_utils.h:194-205)
  c[S_XX] = fi[F_XX] * fi[F_XX] + fi[F_YX] * fi[F_YX] + fi[F_ZX] * fi[F_ZX];
  c[S_YY] = fi[F_XY] * fi[F_XY] + fi[F_YY] * fi[F_YY] + fi[F_ZY] * fi[F_ZY];
  c[S_ZZ] = fi[F_XZ] * fi[F_XZ] + fi[F_YZ] * fi[F_YZ] + fi[F_ZZ] * fi[F_ZZ];
  c[S_XY] = fi[F_XX] * fi[F_XY] + fi[F_YX] * fi[F_YY] + fi[F_ZX] * fi[F_ZY];
  c[S_YZ] = fi[F_XY] * fi[F_XZ] + fi[F_YY] * fi[F_YZ] + fi[F_ZY] * fi[F_ZZ];
  c[S_ZX] = fi[F_XX] * fi[F_XZ] + fi[F_YX] * fi[F_YZ] + fi[F_ZX] * fi[F_ZZ];

  double ev[3], v1[3], v2[3], v3[3];
  eigen_sym33(c, ev, v1, v2, v3);
  ev[0] = ev[0] < 0.0 ? 0.0 : ev[0];
  ev[1] = ev[1] < 0.0 ? 0.0 : ev[1];
  ev[2] = ev[2] < 0.0 ? 0.0 : ev[2];

  const double l1 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
  const double l2 = v2[0] * v2[0] + v2[1] * v2[1] + v2[2] * v2[2];
  const double l3 = v3[0] * v3[0] + v3[1] * v3[1] + v3[2] * v3[2];

  const double xi = 1.0 / (sqrt(ev[0]) * l1);
  const double yi = 1.0 / (sqrt(ev[1]) * l2);
  const double zi = 1.0 / (sqrt(ev[2]) * l3);

  V[S_XX] = xi * v1[0] * v1[0] + yi * v2[0] * v2[0] + zi * v3[0] * v3[0];
  V[S_YY] = xi * v1[1] * v1[1] + yi * v2[1] * v2[1] + zi * v3[1] * v3[1];
  V[S_ZZ] = xi * v1[2] * v1[2] + yi * v2[2] * v2[2] + zi * v3[2] * v3[2];
  V[S_XY] = xi * v1[0] * v1[1] + yi * v2[0] * v2[1] + zi * v3[0] * v3[1];
  V[S_YZ] = xi * v1[1] * v1[2] + yi * v2[1] * v2[2] + zi * v3[1] * v3[2];
  V[S_ZX] = xi * v1[2] * v1[0] + yi * v2[2] * v2[0] + zi * v3[2] * v3[0];
}

// Neo-Hookean Cauchy stress from F (src/nimble_material.cc:252-310).
__device__ __forceinline__ void neohookean_stress(double K, double G, const double* F, double* sig)
{
  double v[6];
  left_stretch(F, v);

  const double xj = v[S_XX] * v[S_YY] * v[S_ZZ] + 2.0 * v[S_XY] * v[S_YZ] * v[S_ZX] - v[S_XX] * v[S_YZ] * v[S_YZ] -
                    v[S_YY] * v[S_ZX] * v[S_ZX] - v[S_ZZ] * v[S_XY] * v[S_XY];

  const double cj  = cbrt_ref(xj);
  const double fac = 1.0 / (cj * cj);

  double bxx = v[S_XX] * v[S_XX] + v[S_XY] * v[S_XY] + v[S_ZX] * v[S_ZX];
  double byy = v[S_XY] * v[S_XY] + v[S_YY] * v[S_YY] + v[S_YZ] * v[S_YZ];
  double bzz = v[S_ZX] * v[S_ZX] + v[S_YZ] * v[S_YZ] + v[S_ZZ] * v[S_ZZ];
  double bxy = v[S_XX] * v[S_XY] + v[S_XY] * v[S_YY] + v[S_ZX] * v[S_YZ];
  double byz = v[S_XY] * v[S_ZX] + v[S_YY] * v[S_YZ] + v[S_YZ] * v[S_ZZ];
  double bzx = v[S_ZX] * v[S_XX] + v[S_YZ] * v[S_XY] + v[S_ZZ] * v[S_ZX];

  bxx = fac * bxx;
  byy = fac * byy;
  bzz = fac * bzz;
  bxy = fac * bxy;
  byz = fac * byz;
  bzx = fac * bzx;

  const double trace = bxx + byy + bzz;
  bxx = bxx - trace / 3.0;
  byy = byy - trace / 3.0;
  bzz = bzz - trace / 3.0;

  // 1/xj of the pressure and the six quotients of the deviator share xj
  const double n[7] = {1.0, G * bxx, G * byy, G * bzz, G * bxy, G * byz, G * bzx};
  double       q[7];
  divide_by(n, xj, q);
  const double pressure = 0.5 * K * (xj - q[0]);
  sig[S_XX] = pressure + q[1];
  sig[S_YY] = pressure + q[2];
  sig[S_ZZ] = pressure + q[3];
  sig[S_XY] = q[4];
  sig[S_YZ] = q[5];
  sig[S_ZX] = q[6];
}

}  // namespace fp64
}  // namespace nsm
