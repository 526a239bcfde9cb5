// nsm_mixed_math.h -- per-element arithmetic of the mixed-precision path.
//
// One element per thread.  The trilinear field of a HEX8 is written in its
// Walsh (hourglass) modes: with node signs s_j = (sx, sy, sz) in
// {-1, +1}^3,
//   x(xi, eta, zeta) = 1/8 sum_modes W_mode * mode(xi, eta, zeta),
//   W_mode = sum_j mode(s_j) x_j,
// so the Jacobian columns at a Gauss point (xi_q, eta_q, zeta_q) in {+-g}^3 are
//   dx/dxi   = 1/8 (W_x + eta W_xy + zeta W_xz + eta zeta W_xyz), ...
// This replaces the 8x3 shape-derivative sums of
// src/nimble_element.h:442-482 (same values in exact arithmetic) by 7 modal
// sums per component.  Forces are accumulated in the same basis:
//   f_n = -sum_q sum_m dN_n/dxi_m(q) M_q[m][c],  M_q = adj(a_q) sigma_q
// (src/nimble_element.h:586-617 with det(a) a^{-1} = adj(a), w_q = 1) become
// 7 modal moments of M_q that one inverse transform maps back to the nodes.
//
// Precision: node coordinates and displacements arrive in FP64 and the modal
// sums are formed in FP64 (they are differences of nearby node values), then
// converted to FP32; everything else is FP32.  The stress uses the closed form
// on H = F - I (no cancellation at F ~ I):
//   B - I = H + H^T + H H^T,  J - 1 = tr H + ((tr H)^2 - tr H^2)/2 + det H,
//   sigma = K/2 (J - 1/J) I + G J^{-5/3} dev(B)
// which equals src/nimble_material.cc:261-309 (V^2 = F F^T, det V = J).
#pragma once

#include <hip/hip_runtime.h>

namespace nsm {
namespace mixed {

// Gauss coordinate as written in src/nimble_element.cc:58
constexpr float kG  = 0.577350269189626f;
constexpr float kG2 = (float)(0.577350269189626 * 0.577350269189626);

// Walsh modes of one scalar nodal field given in HEX8 node order
// (src/nimble_element.cc:100-121): 0(-,-,-) 1(+,-,-) 2(+,+,-) 3(-,+,-)
// 4(-,-,+) 5(+,-,+) 6(+,+,+) 7(-,+,+).  Output w[0..6] =
// W_x, W_y, W_z, W_xy, W_yz, W_xz, W_xyz.
__host__ __device__ __forceinline__ void walsh7(const double v[8], double w[7])
{
  // pairs in x at (y, z) = (-,-), (+,-), (-,+), (+,+)
  const double s00 = v[1] + v[0], d00 = v[1] - v[0];
  const double s10 = v[2] + v[3], d10 = v[2] - v[3];
  const double s01 = v[5] + v[4], d01 = v[5] - v[4];
  const double s11 = v[6] + v[7], d11 = v[6] - v[7];
  // y
  const double ss0 = s10 + s00, sd0 = s10 - s00, ss1 = s11 + s01, sd1 = s11 - s01;
  const double ds0 = d10 + d00, dd0 = d10 - d00, ds1 = d11 + d01, dd1 = d11 - d01;
  // z
  w[0] = ds1 + ds0;  // W_x
  w[1] = sd1 + sd0;  // W_y
  w[2] = ss1 - ss0;  // W_z
  w[3] = dd1 + dd0;  // W_xy
  w[4] = sd1 - sd0;  // W_yz
  w[5] = ds1 - ds0;  // W_xz
  w[6] = dd1 - dd0;  // W_xyz
}

// Modal coefficients scaled for evaluation at Gauss points: c[0..2] = W_x,y,z,
// c[3..5] = g W_xy,yz,xz, c[6] = g^2 W_xyz.  That is 8 times the Jacobian's
// coefficients (the 1/8 of the shape functions is left out): a power of
// two, so every quantity formed from them is exactly 8^k times its value
// and rounds the same -- H, F, J and sigma are unchanged, the force moments
// are 64 times theirs and forces_from_moments divides by 64.
struct Modes {
  float c[7][3];  // [mode][component]
};

__host__ __device__ __forceinline__ void make_modes(const double v[8][3], Modes& m)
{
  for (int i = 0; i < 3; i++) {
    const double col[8] = {v[0][i], v[1][i], v[2][i], v[3][i], v[4][i], v[5][i], v[6][i], v[7][i]};
    double       w[7];
    walsh7(col, w);
    m.c[0][i] = (float)w[0];
    m.c[1][i] = (float)w[1];
    m.c[2][i] = (float)w[2];
    m.c[3][i] = (float)(w[3] * 0.577350269189626);
    m.c[4][i] = (float)(w[4] * 0.577350269189626);
    m.c[5][i] = (float)(w[5] * 0.577350269189626);
    m.c[6][i] = (float)(w[6] * (0.577350269189626 * 0.577350269189626));
  }
}

// Jacobian J[i][d] = dx_i / dxi_d at the Gauss point with signs (sx, sy, sz):
//   J[i][0] = c0 + sy c3 + sz c5 + sy sz c6,
//   J[i][1] = c1 + sx c3 + sz c4 + sx sz c6,
//   J[i][2] = c2 + sx c5 + sy c4 + sx sy c6   (c = m.c[.][i]),
// evaluated for the two points (-g, SY g, SZ g) and (+g, SY g, SZ g): the
// d/dxi column is common, the others are a common part plus or minus their
// xi-dependent part.
template <int SY, int SZ>
__host__ __device__ __forceinline__ void jacobian_pair(const Modes& m, float Jm[9], float Jp[9])
{
  for (int i = 0; i < 3; i++) {
    const float j0 = m.c[0][i] + (float)SY * m.c[3][i] + (float)SZ * m.c[5][i] + (float)(SY * SZ) * m.c[6][i];
    const float a1 = m.c[1][i] + (float)SZ * m.c[4][i], b1 = m.c[3][i] + (float)SZ * m.c[6][i];
    const float a2 = m.c[2][i] + (float)SY * m.c[4][i], b2 = m.c[5][i] + (float)SY * m.c[6][i];
    Jm[3 * i + 0] = j0;
    Jp[3 * i + 0] = j0;
    Jm[3 * i + 1] = a1 - b1;
    Jp[3 * i + 1] = a1 + b1;
    Jm[3 * i + 2] = a2 - b2;
    Jp[3 * i + 2] = a2 + b2;
  }
}

__host__ __device__ __forceinline__ void adjugate(const float m[9], float adj[9])
{
  adj[0] = m[4] * m[8] - m[5] * m[7];
  adj[1] = m[2] * m[7] - m[1] * m[8];
  adj[2] = m[1] * m[5] - m[2] * m[4];
  adj[3] = m[5] * m[6] - m[3] * m[8];
  adj[4] = m[0] * m[8] - m[2] * m[6];
  adj[5] = m[2] * m[3] - m[0] * m[5];
  adj[6] = m[3] * m[7] - m[4] * m[6];
  adj[7] = m[1] * m[6] - m[0] * m[7];
  adj[8] = m[0] * m[4] - m[1] * m[3];
}

__host__ __device__ __forceinline__ float j_pow_m23(float J)
{
#if defined(__HIP_DEVICE_COMPILE__)
  // hardware log2 / exp2; J ~ 1, absolute error ~1e-7 on a factor that
  // multiplies the (small) deviator
  return __builtin_amdgcn_exp2f(-0.666666667f * __builtin_amdgcn_logf(J));
#else
  return powf(J, -2.0f / 3.0f);
#endif
}

__host__ __device__ __forceinline__ float rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}

// One Gauss point: F (K_F order), sigma (K_S order) and M = adj(a) sigma
// (row-major), from the reference Jacobian Jx and displacement Jacobian Ju.
// Returns false when det(Jx) or det(a) = J det(Jx) is not positive.
__host__ __device__ __forceinline__ bool point(const float Jx[9], const float Ju[9], float K, float G, float F[9],
                                               float S[6], float M[9])
{
  float adjb[9];
  adjugate(Jx, adjb);
  const float detb = Jx[0] * adjb[0] + Jx[1] * adjb[3] + Jx[2] * adjb[6];
  const float rdb  = rcp(detb);
  float       H[9];
  for (int i = 0; i < 3; i++)
    for (int k = 0; k < 3; k++)
      H[3 * i + k] = (Ju[3 * i] * adjb[k] + Ju[3 * i + 1] * adjb[3 + k] + Ju[3 * i + 2] * adjb[6 + k]) * rdb;

  F[0] = 1.0f + H[0];
  F[1] = 1.0f + H[4];
  F[2] = 1.0f + H[8];
  F[3] = H[1];
  F[4] = H[5];
  F[5] = H[6];
  F[6] = H[3];
  F[7] = H[7];
  F[8] = H[2];

  const float bxx  = 2.0f * H[0] + (H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
  const float byy  = 2.0f * H[4] + (H[3] * H[3] + H[4] * H[4] + H[5] * H[5]);
  const float bzz  = 2.0f * H[8] + (H[6] * H[6] + H[7] * H[7] + H[8] * H[8]);
  const float bxy  = (H[1] + H[3]) + (H[0] * H[3] + H[1] * H[4] + H[2] * H[5]);
  const float byz  = (H[5] + H[7]) + (H[3] * H[6] + H[4] * H[7] + H[5] * H[8]);
  const float bzx  = (H[6] + H[2]) + (H[6] * H[0] + H[7] * H[1] + H[8] * H[2]);
  const float trH  = H[0] + H[4] + H[8];
  const float trH2 = H[0] * H[0] + H[4] * H[4] + H[8] * H[8] + 2.0f * (H[1] * H[3] + H[2] * H[6] + H[5] * H[7]);
  const float detH = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                     H[2] * (H[3] * H[7] - H[4] * H[6]);
  const float dJ       = trH + 0.5f * (trH * trH - trH2) + detH;
  const float J        = 1.0f + dJ;
  const float rJ       = rcp(J);
  const float fac      = j_pow_m23(J);
  const float pressure = 0.5f * K * dJ * (2.0f + dJ) * rJ;
  const float tr3      = (bxx + byy + bzz) * (1.0f / 3.0f);
  const float coef     = G * fac * rJ;
  S[0]                 = pressure + coef * (bxx - tr3);
  S[1]                 = pressure + coef * (byy - tr3);
  S[2]                 = pressure + coef * (bzz - tr3);
  S[3]                 = coef * bxy;
  S[4]                 = coef * byz;
  S[5]                 = coef * bzx;

  float a[9], adja[9];
  for (int i = 0; i < 9; i++) a[i] = Jx[i] + Ju[i];
  adjugate(a, adja);
  const float Sm[9] = {S[0], S[3], S[5], S[3], S[1], S[4], S[5], S[4], S[2]};
  for (int m = 0; m < 3; m++)
    for (int c = 0; c < 3; c++) M[3 * m + c] = adja[3 * m] * Sm[c] + adja[3 * m + 1] * Sm[3 + c] + adja[3 * m + 2] * Sm[6 + c];
  // det(a) = J det(Jx): with det(Jx) > 0, a is inverted exactly when J <= 0
  return detb > 0.0f && J > 0.0f;
}

// Modal moments of M over the Gauss points (signs only; g factors applied at
// the end):  A[m][c] = sum_q M_q[m][c],  B*[m][c] = sum_q s_q^e M_q[m][c],
// C[m][c] = sum_q s_q^e s_q^f M_q[m][c], e, f the two directions other than m.
struct ForceMoments {
  float A[9], Be[9], Bf[9], C[9];
};

// Two Gauss points that differ only in the sign of xi (Mm at -g, Mp at +g),
// with the signs SY, SZ they share: their sum P and xi-weighted difference D
// enter the moments of the x-row (e = y, f = z: constant over the pair) and
// of the y- and z-rows (e = x: D; f = z or y: constant).
template <int SY, int SZ>
__host__ __device__ __forceinline__ void accumulate_pair(ForceMoments& fm, const float Mm[9], const float Mp[9])
{
  float P[9], D[9];
  for (int i = 0; i < 9; i++) P[i] = Mm[i] + Mp[i];
  for (int i = 3; i < 9; i++) D[i] = Mp[i] - Mm[i];
  for (int c = 0; c < 3; c++) {
    fm.A[c] += P[c];
    fm.Be[c] += (float)SY * P[c];
    fm.Bf[c] += (float)SZ * P[c];
    fm.C[c] += (float)(SY * SZ) * P[c];
    fm.A[3 + c] += P[3 + c];
    fm.Be[3 + c] += D[3 + c];
    fm.Bf[3 + c] += (float)SZ * P[3 + c];
    fm.C[3 + c] += (float)SZ * D[3 + c];
    fm.A[6 + c] += P[6 + c];
    fm.Be[6 + c] += D[6 + c];
    fm.Bf[6 + c] += (float)SY * P[6 + c];
    fm.C[6 + c] += (float)SY * D[6 + c];
  }
}

// Node forces f[n][c] (with the reference's minus sign) from the moments.
__host__ __device__ __forceinline__ void forces_from_moments(const ForceMoments& fm, float f[8][3])
{
  // -1/8 of the shape functions, and 1/64 for the moments' scale (Modes)
  const float s0 = -0.125f / 64.0f, s1 = s0 * kG, s2 = s0 * kG2;
  for (int c = 0; c < 3; c++) {
    // mode coefficients of the force field
    const float cx   = s0 * fm.A[0 + c];
    const float cy   = s0 * fm.A[3 + c];
    const float cz   = s0 * fm.A[6 + c];
    const float cxy  = s1 * (fm.Be[0 + c] + fm.Be[3 + c]);  // x-row paired with y, y-row paired with x
    const float cxz  = s1 * (fm.Bf[0 + c] + fm.Be[6 + c]);  // x-row with z, z-row with x
    const float cyz  = s1 * (fm.Bf[3 + c] + fm.Bf[6 + c]);  // y-row with z, z-row with y
    const float cxyz = s2 * (fm.C[0 + c] + fm.C[3 + c] + fm.C[6 + c]);
    // f(sx, sy, sz) = sx cx + sy cy + sz cz + sx sy cxy + sy sz cyz + sx sz cxz + sx sy sz cxyz
    //               = P(sx, sy) + sz Q(sx, sy)
    const float p1p = cx + cxy, p1m = cx - cxy;      // coefficient of sx for sy = +, -
    const float q0p = cz + cyz, q0m = cz - cyz;      // sz-part constant in sx
    const float q1p = cxz + cxyz, q1m = cxz - cxyz;  // sz-part coefficient of sx
    const float Ppp = cy + p1p, Pmp = cy - p1p, Ppm = -cy + p1m, Pmm = -cy - p1m;  // P(sx, sy)
    const float Qpp = q0p + q1p, Qmp = q0p - q1p, Qpm = q0m + q1m, Qmm = q0m - q1m;
    f[0][c] = Pmm - Qmm;  // (-,-,-)
    f[1][c] = Ppm - Qpm;  // (+,-,-)
    f[2][c] = Ppp - Qpp;  // (+,+,-)
    f[3][c] = Pmp - Qmp;  // (-,+,-)
    f[4][c] = Pmm + Qmm;  // (-,-,+)
    f[5][c] = Ppm + Qpm;  // (+,-,+)
    f[6][c] = Ppp + Qpp;  // (+,+,+)
    f[7][c] = Pmp + Qmp;  // (-,+,+)
  }
}

// Whole element.  Gauss point q has signs (sx, sy, sz) of node q
// (src/nimble_element.cc:60-90).  F and sigma of each pair of points
// (0,1), (3,2), (4,5), (7,6) go to store(q_minus, q_plus, F-, S-, F+, S+)
// (F in K_F order, sigma in K_S order); f [8][3].
template <typename Store>
__host__ __device__ __forceinline__ bool element(const double X[8][3], const double u[8][3], float K, float G,
                                                 Store&& store, float f[8][3])
{
  Modes mx, mu;
  make_modes(X, mx);
  make_modes(u, mu);
  ForceMoments fm;
  for (int i = 0; i < 9; i++) fm.A[i] = fm.Be[i] = fm.Bf[i] = fm.C[i] = 0.0f;
  bool ok = true;
  // points in pairs (xi = -g, +g) with the same eta, zeta signs; the pair's
  // F and sigma go to store together
#define NSM_MIXED_PAIR(QM, QP, SY, SZ)                                               \
  {                                                                                  \
    float Jxm[9], Jxp[9], Jum[9], Jup[9], Mm[9], Mp[9], Fm[9], Sm[6], Fp[9], Sp[6]; \
    jacobian_pair<SY, SZ>(mx, Jxm, Jxp);                                             \
    jacobian_pair<SY, SZ>(mu, Jum, Jup);                                             \
    ok = point(Jxm, Jum, K, G, Fm, Sm, Mm) && ok;                                    \
    ok = point(Jxp, Jup, K, G, Fp, Sp, Mp) && ok;                                    \
    store(QM, QP, Fm, Sm, Fp, Sp);                                                   \
    accumulate_pair<SY, SZ>(fm, Mm, Mp);                                             \
  }
  NSM_MIXED_PAIR(0, 1, -1, -1)
  NSM_MIXED_PAIR(3, 2, 1, -1)
  NSM_MIXED_PAIR(4, 5, -1, 1)
  NSM_MIXED_PAIR(7, 6, 1, 1)
#undef NSM_MIXED_PAIR
  forces_from_moments(fm, f);
  return ok;
}

}  // namespace mixed
}  // namespace nsm
