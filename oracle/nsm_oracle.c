/*
 * nsm_oracle.c -- CPU FP64 restatement of NimbleSM's explicit internal-force
 * path.  TEST INFRASTRUCTURE ONLY (the checker and the timed CPU baseline);
 * the HIP product never links, loads or calls this code.
 *
 * Build with -ffp-contract=off and without -ffast-math: every expression is
 * written in the reference's evaluation order so that, on x86-64 (SSE2, no
 * contraction), results are bit-identical to the reference's own sources.
 * Citations are NimbleSM file:line.
 */
#include "nsm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* tensor storage indices, src/nimble_utils.h:91-109 */
enum { S_XX = 0, S_YY = 1, S_ZZ = 2, S_XY = 3, S_YZ = 4, S_ZX = 5, S_YX = 3, S_ZY = 4, S_XZ = 5 };
enum { F_XX = 0, F_YY = 1, F_ZZ = 2, F_XY = 3, F_YZ = 4, F_ZX = 5, F_YX = 6, F_ZY = 7, F_XZ = 8 };

/* ------------------------------------------------------------------------ */
/* element tables: src/nimble_element.cc:55-171                              */
/* ------------------------------------------------------------------------ */
static double g_int_pts[24], g_N[64], g_dN[192];
static int    g_tables_ready = 0;

void orc_hex_tables(double int_pts[24], double N[64], double dN[192])
{
  /* Gauss constant as written in src/nimble_element.cc:58 (truncated 1/sqrt(3)) */
  const double val = 0.577350269189626;
  /* ipt sign pattern, src/nimble_element.cc:60-90 */
  static const int sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                               {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  for (int q = 0; q < 8; q++)
    for (int d = 0; d < 3; d++) int_pts[3 * q + d] = sg[q][d] > 0 ? val : -val;

  const double c = 1.0 / 8.0;
  for (int i = 0; i < 8; i++) { /* src/nimble_element.cc:100-121 */
    double r = int_pts[3 * i], s = int_pts[3 * i + 1], t = int_pts[3 * i + 2];
    N[8 * i + 0] = c * (1.0 - r) * (1.0 - s) * (1.0 - t);
    N[8 * i + 1] = c * (1.0 + r) * (1.0 - s) * (1.0 - t);
    N[8 * i + 2] = c * (1.0 + r) * (1.0 + s) * (1.0 - t);
    N[8 * i + 3] = c * (1.0 - r) * (1.0 + s) * (1.0 - t);
    N[8 * i + 4] = c * (1.0 - r) * (1.0 - s) * (1.0 + t);
    N[8 * i + 5] = c * (1.0 + r) * (1.0 - s) * (1.0 + t);
    N[8 * i + 6] = c * (1.0 + r) * (1.0 + s) * (1.0 + t);
    N[8 * i + 7] = c * (1.0 - r) * (1.0 + s) * (1.0 + t);
  }
  for (int q = 0; q < 8; q++) { /* src/nimble_element.cc:123-171 */
    double r = int_pts[3 * q], s = int_pts[3 * q + 1], t = int_pts[3 * q + 2];
    double *d = &dN[24 * q];
    d[0]  = -c * (1.0 - s) * (1.0 - t);
    d[1]  = -c * (1.0 - r) * (1.0 - t);
    d[2]  = -c * (1.0 - r) * (1.0 - s);
    d[3]  = c * (1.0 - s) * (1.0 - t);
    d[4]  = -c * (1.0 + r) * (1.0 - t);
    d[5]  = -c * (1.0 + r) * (1.0 - s);
    d[6]  = c * (1.0 + s) * (1.0 - t);
    d[7]  = c * (1.0 + r) * (1.0 - t);
    d[8]  = -c * (1.0 + r) * (1.0 + s);
    d[9]  = -c * (1.0 + s) * (1.0 - t);
    d[10] = c * (1.0 - r) * (1.0 - t);
    d[11] = -c * (1.0 - r) * (1.0 + s);
    d[12] = -c * (1.0 - s) * (1.0 + t);
    d[13] = -c * (1.0 - r) * (1.0 + t);
    d[14] = c * (1.0 - r) * (1.0 - s);
    d[15] = c * (1.0 - s) * (1.0 + t);
    d[16] = -c * (1.0 + r) * (1.0 + t);
    d[17] = c * (1.0 + r) * (1.0 - s);
    d[18] = c * (1.0 + s) * (1.0 + t);
    d[19] = c * (1.0 + r) * (1.0 + t);
    d[20] = c * (1.0 + r) * (1.0 + s);
    d[21] = -c * (1.0 + s) * (1.0 + t);
    d[22] = c * (1.0 - r) * (1.0 + t);
    d[23] = c * (1.0 - r) * (1.0 + s);
  }
}

static void ensure_tables(void)
{
  if (!g_tables_ready) {
    orc_hex_tables(g_int_pts, g_N, g_dN);
    g_tables_ready = 1;
  }
}

/* ------------------------------------------------------------------------ */
/* 3x3 kernels: src/nimble_utils.h                                           */
/* ------------------------------------------------------------------------ */

/* Invert3x3, src/nimble_utils.h:1229-1268.  m is row-major m[i][j].
 * status is set to 1 when det <= 0 (the reference asserts on the host). */
double orc_invert3x3(const double m[9], double inv[9], int *status)
{
#define M(i, j) m[3 * (i) + (j)]
  double minor0 = M(1, 1) * M(2, 2) - M(1, 2) * M(2, 1);
  double minor1 = M(1, 0) * M(2, 2) - M(1, 2) * M(2, 0);
  double minor2 = M(1, 0) * M(2, 1) - M(1, 1) * M(2, 0);
  double minor3 = M(0, 1) * M(2, 2) - M(0, 2) * M(2, 1);
  double minor4 = M(0, 0) * M(2, 2) - M(2, 0) * M(0, 2);
  double minor5 = M(0, 0) * M(2, 1) - M(0, 1) * M(2, 0);
  double minor6 = M(0, 1) * M(1, 2) - M(0, 2) * M(1, 1);
  double minor7 = M(0, 0) * M(1, 2) - M(0, 2) * M(1, 0);
  double minor8 = M(0, 0) * M(1, 1) - M(0, 1) * M(1, 0);
  double det    = M(0, 0) * minor0 - M(0, 1) * minor1 + M(0, 2) * minor2;
#undef M
  if (det <= 0.0 && status) *status = 1;
  inv[0] = minor0 / det;
  inv[1] = -1.0 * minor3 / det;
  inv[2] = minor6 / det;
  inv[3] = -1.0 * minor1 / det;
  inv[4] = minor4 / det;
  inv[5] = -1.0 * minor7 / det;
  inv[6] = minor2 / det;
  inv[7] = -1.0 * minor5 / det;
  inv[8] = minor8 / det;
  return det;
}

/* Invert_Full33, src/nimble_utils.h:523-551 */
double orc_invert_full33(const double mat[9], double inv[9])
{
  double minor0 = mat[F_YY] * mat[F_ZZ] - mat[F_YZ] * mat[F_ZY];
  double minor1 = mat[F_YX] * mat[F_ZZ] - mat[F_YZ] * mat[F_ZX];
  double minor2 = mat[F_YX] * mat[F_ZY] - mat[F_YY] * mat[F_ZX];
  double minor3 = mat[F_XY] * mat[F_ZZ] - mat[F_XZ] * mat[F_ZY];
  double minor4 = mat[F_XX] * mat[F_ZZ] - mat[F_ZX] * mat[F_XZ];
  double minor5 = mat[F_XX] * mat[F_ZY] - mat[F_XY] * mat[F_ZX];
  double minor6 = mat[F_XY] * mat[F_YZ] - mat[F_XZ] * mat[F_YY];
  double minor7 = mat[F_XX] * mat[F_YZ] - mat[F_XZ] * mat[F_YX];
  double minor8 = mat[F_XX] * mat[F_YY] - mat[F_XY] * mat[F_YX];
  double det    = mat[F_XX] * minor0 - mat[F_XY] * minor1 + mat[F_XZ] * minor2;
  inv[F_XX] = minor0 / det;
  inv[F_XY] = -1.0 * minor3 / det;
  inv[F_XZ] = minor6 / det;
  inv[F_YX] = -1.0 * minor1 / det;
  inv[F_YY] = minor4 / det;
  inv[F_YZ] = -1.0 * minor7 / det;
  inv[F_ZX] = minor2 / det;
  inv[F_ZY] = -1.0 * minor5 / det;
  inv[F_ZZ] = minor8 / det;
  return det;
}

/* Square_Full33T_Full33: result = mat^T mat, src/nimble_utils.h:194-205 */
static void square_full33T_full33(const double *mat, double *r)
{
  r[S_XX] = mat[F_XX] * mat[F_XX] + mat[F_YX] * mat[F_YX] + mat[F_ZX] * mat[F_ZX];
  r[S_YY] = mat[F_XY] * mat[F_XY] + mat[F_YY] * mat[F_YY] + mat[F_ZY] * mat[F_ZY];
  r[S_ZZ] = mat[F_XZ] * mat[F_XZ] + mat[F_YZ] * mat[F_YZ] + mat[F_ZZ] * mat[F_ZZ];
  r[S_XY] = mat[F_XX] * mat[F_XY] + mat[F_YX] * mat[F_YY] + mat[F_ZX] * mat[F_ZY];
  r[S_YZ] = mat[F_XY] * mat[F_XZ] + mat[F_YY] * mat[F_YZ] + mat[F_ZY] * mat[F_ZZ];
  r[S_ZX] = mat[F_XX] * mat[F_XZ] + mat[F_YX] * mat[F_YZ] + mat[F_ZX] * mat[F_ZZ];
}

/* Mult_Full33_Sym33_ReturnT, src/nimble_utils.h:324-339 */
static void mult_full33_sym33_returnT(const double *full, const double *sym, double *r)
{
  r[F_XX] = full[F_XX] * sym[S_XX] + full[F_XY] * sym[S_YX] + full[F_XZ] * sym[S_ZX];
  r[F_YX] = full[F_XX] * sym[S_XY] + full[F_XY] * sym[S_YY] + full[F_XZ] * sym[S_ZY];
  r[F_ZX] = full[F_XX] * sym[S_XZ] + full[F_XY] * sym[S_YZ] + full[F_XZ] * sym[S_ZZ];
  r[F_XY] = full[F_YX] * sym[S_XX] + full[F_YY] * sym[S_YX] + full[F_YZ] * sym[S_ZX];
  r[F_YY] = full[F_YX] * sym[S_XY] + full[F_YY] * sym[S_YY] + full[F_YZ] * sym[S_ZY];
  r[F_ZY] = full[F_YX] * sym[S_XZ] + full[F_YY] * sym[S_YZ] + full[F_YZ] * sym[S_ZZ];
  r[F_XZ] = full[F_ZX] * sym[S_XX] + full[F_ZY] * sym[S_YX] + full[F_ZZ] * sym[S_ZX];
  r[F_YZ] = full[F_ZX] * sym[S_XY] + full[F_ZY] * sym[S_YY] + full[F_ZZ] * sym[S_ZY];
  r[F_ZZ] = full[F_ZX] * sym[S_XZ] + full[F_ZY] * sym[S_YZ] + full[F_ZZ] * sym[S_ZZ];
}

/* MultiplySign (x times sign of y), src/nimble_utils.h:131-136 */
static inline double multiply_sign(double x, double y) { return x * (double)(1 - 2 * (y < 0)); }

/* 6/5 minimax rational approximation of cos(acos(x)/3) on [0,1];
 * coefficients as published in src/nimble_utils.h:650-665 */
double orc_cos_of_acos_div3(double x)
{
  const double x2 = x * x;
  const double x4 = x2 * x2;
  return (0.866025403784438713 + 2.12714890259493060 * x +
          ((1.89202064815951569 + 0.739603278343401613 * x) * x2 +
           (0.121973926953064794 + x * (0.00655637626263929360 + 0.0000390884982780803443 * x)) * x4)) /
         (1.0 + 2.26376989330935617 * x +
          ((1.80461009751278976 + 0.603976798217196003 * x) * x2 +
           (0.0783255761115461708 + 0.00268525944538021629 * x) * x4));
}

/* Closed-form symmetric 3x3 eigen-solver, src/nimble_utils.h:667-857 */
void orc_eigen_sym33_nonunit(const double t[6], double ev[3], double evec0[3], double evec1[3], double evec2[3])
{
  double       cxx = t[S_XX];
  double       cyy = t[S_YY];
  double       czz = t[S_ZZ];
  const double cxy = t[S_XY];
  const double cyz = t[S_YZ];
  const double czx = t[S_ZX];

  const double c1 = (cxx + cyy + czz) / 3.0;
  cxx -= c1;
  cyy -= c1;
  czz -= c1;

  const double cxy_cxy = cxy * cxy;
  const double cyz_cyz = cyz * cyz;
  const double czx_czx = czx * czx;
  const double cxx_cyy = cxx * cyy;

  const double c2 = cxx_cyy + cyy * czz + czz * cxx - cxy_cxy - cyz_cyz - czx_czx;

  const double ThreeOverA     = -3.0 / c2;
  const double sqrtThreeOverA = sqrt(ThreeOverA);

  const double c3 = cxx * cyz_cyz + cyy * czx_czx - 2.0 * cxy * cyz * czx + czz * (cxy_cxy - cxx_cyy);

  const double rr = -0.5 * c3 * ThreeOverA * sqrtThreeOverA;

  const double arg = fabs(rr) < 1.0 ? fabs(rr) : 1.0; /* Minimum(|rr|, 1) */

  const double cos_thd3     = orc_cos_of_acos_div3(arg);
  const double two_cos_thd3 = 2.0 * multiply_sign(cos_thd3, rr);

  double eval0, eval1, eval2 = two_cos_thd3 / sqrtThreeOverA;

  double crow0[3] = {cxx - eval2, cxy, czx};
  double crow1[3] = {cxy, cyy - eval2, cyz};
  double crow2[3] = {czx, cyz, czz - eval2};

  /* QR decomposition with column pivoting */
  const double k0 = crow0[0] * crow0[0] + cxy_cxy + czx_czx;
  const double k1 = cxy_cxy + crow1[1] * crow1[1] + cyz_cyz;
  const double k2 = czx_czx + cyz_cyz + crow2[2] * crow2[2];

  double k_row1[3], row2[3], row3[3];

  const int k0gk1 = k1 <= k0;
  const int k0gk2 = k2 <= k0;
  const int k1gk2 = k2 <= k1;

  const int k0_largest = k0gk1 && k0gk2;
  const int k1_largest = k1gk2 && !k0gk1;
  const int k2_largest = !(k0_largest || k1_largest);

  for (int i = 0; i < 3; i++)
    k_row1[i] = (k0_largest ? crow0[i] : 0.0) + (k1_largest ? crow1[i] : 0.0) + (k2_largest ? crow2[i] : 0.0);
  for (int i = 0; i < 3; i++) row2[i] = k0_largest ? crow1[i] : crow0[i];
  for (int i = 0; i < 3; i++) row3[i] = k2_largest ? crow1[i] : crow2[i];

  const double ki_ki = 1.0 / ((k0_largest ? k0 : 0.0) + (k1_largest ? k1 : 0.0) + (k2_largest ? k2 : 0.0));

  const double ki_dpr1 = ki_ki * (k_row1[0] * row2[0] + k_row1[1] * row2[1] + k_row1[2] * row2[2]);
  const double ki_dpr2 = ki_ki * (k_row1[0] * row3[0] + k_row1[1] * row3[1] + k_row1[2] * row3[2]);

  for (int i = 0; i < 3; i++) row2[i] -= ki_dpr1 * k_row1[i];
  for (int i = 0; i < 3; i++) row3[i] -= ki_dpr2 * k_row1[i];

  const double a0 = row2[0] * row2[0] + row2[1] * row2[1] + row2[2] * row2[2];
  const double a1 = row3[0] * row3[0] + row3[1] * row3[1] + row3[2] * row3[2];

  double    a_row2[3];
  const int a0lea1 = a0 <= a1;
  for (int i = 0; i < 3; i++) a_row2[i] = a0lea1 ? row3[i] : row2[i];
  const double ai_ai = 1.0 / (a0lea1 ? a1 : a0);

  evec2[0] = k_row1[1] * a_row2[2] - k_row1[2] * a_row2[1];
  evec2[1] = k_row1[2] * a_row2[0] - k_row1[0] * a_row2[2];
  evec2[2] = k_row1[0] * a_row2[1] - k_row1[1] * a_row2[0];

  const double k_atr11 = cxx * k_row1[0] + cxy * k_row1[1] + czx * k_row1[2];
  const double k_atr21 = cxy * k_row1[0] + cyy * k_row1[1] + cyz * k_row1[2];
  const double k_atr31 = czx * k_row1[0] + cyz * k_row1[1] + czz * k_row1[2];

  const double a_atr12 = cxx * a_row2[0] + cxy * a_row2[1] + czx * a_row2[2];
  const double a_atr22 = cxy * a_row2[0] + cyy * a_row2[1] + cyz * a_row2[2];
  const double a_atr32 = czx * a_row2[0] + cyz * a_row2[1] + czz * a_row2[2];

  double       rm2xx       = (k_row1[0] * k_atr11 + k_row1[1] * k_atr21 + k_row1[2] * k_atr31) * ki_ki;
  const double k_a_rm2xy   = (k_row1[0] * a_atr12 + k_row1[1] * a_atr22 + k_row1[2] * a_atr32);
  double       rm2yy       = (a_row2[0] * a_atr12 + a_row2[1] * a_atr22 + a_row2[2] * a_atr32) * ai_ai;
  const double rm2xy_rm2xy = k_a_rm2xy * k_a_rm2xy * ai_ai * ki_ki;

  /* Wilkinson shift */
  const double b        = 0.5 * (rm2xx - rm2yy);
  const double sqrtTerm = multiply_sign(sqrt(b * b + rm2xy_rm2xy), b);
  eval0                 = rm2yy + b - sqrtTerm;
  eval1                 = rm2xx + rm2yy - eval0;

  rm2xx -= eval0;
  rm2yy -= eval0;

  const double rm2xx2 = rm2xx * rm2xx;
  const double rm2yy2 = rm2yy * rm2yy;

  const double fac1 = rm2xx2 < rm2yy2 ? k_a_rm2xy * ai_ai : rm2xx;
  const double fac2 = rm2xx2 < rm2yy2 ? rm2yy : ki_ki * k_a_rm2xy;

  for (int i = 0; i < 3; i++) evec0[i] = fac1 * a_row2[i] - fac2 * k_row1[i];

  const int both_zero = (rm2xx2 == 0.0) && (rm2xy_rm2xy == 0.0);
  for (int i = 0; i < 3; i++) evec0[i] = both_zero ? a_row2[i] : evec0[i];

  evec1[0] = evec2[1] * evec0[2] - evec2[2] * evec0[1];
  evec1[1] = evec2[2] * evec0[0] - evec2[0] * evec0[2];
  evec1[2] = evec2[0] * evec0[1] - evec2[1] * evec0[0];

  eval0 += c1;
  eval1 += c1;
  eval2 += c1;

  const double c2tol        = (c1 * c1) * (-1.0e-30);
  const int    c2lsmall_neg = c2 < c2tol;

  ev[0] = c2lsmall_neg ? eval0 : c1;
  ev[1] = c2lsmall_neg ? eval1 : c1;
  ev[2] = c2lsmall_neg ? eval2 : c1;

  evec0[0] = c2lsmall_neg ? evec0[0] : 1.0;
  evec0[1] = c2lsmall_neg ? evec0[1] : 0.0;
  evec0[2] = c2lsmall_neg ? evec0[2] : 0.0;
  evec1[0] = c2lsmall_neg ? evec1[0] : 0.0;
  evec1[1] = c2lsmall_neg ? evec1[1] : 1.0;
  evec1[2] = c2lsmall_neg ? evec1[2] : 0.0;
  evec2[0] = c2lsmall_neg ? evec2[0] : 0.0;
  evec2[1] = c2lsmall_neg ? evec2[1] : 0.0;
  evec2[2] = c2lsmall_neg ? evec2[2] : 1.0;
}

/* Polar_Decomp (left stretch V, rotation R), src/nimble_utils.h:859-908 */
void orc_polar_decomp(const double mat[9], double V[6], double R[9])
{
  double mat_inv[9];
  orc_invert_full33(mat, mat_inv);

  double c[6];
  square_full33T_full33(mat_inv, c);

  double vec1[3] = {0.0, 0.0, 0.0}, vec2[3] = {0.0, 0.0, 0.0}, vec3[3] = {0.0, 0.0, 0.0};
  double ev[3] = {0.0, 0.0, 0.0};
  orc_eigen_sym33_nonunit(c, ev, vec1, vec2, vec3);

  ev[0] = ev[0] < 0.0 ? 0.0 : ev[0];
  ev[1] = ev[1] < 0.0 ? 0.0 : ev[1];
  ev[2] = ev[2] < 0.0 ? 0.0 : ev[2];

  const double len1_sq = vec1[0] * vec1[0] + vec1[1] * vec1[1] + vec1[2] * vec1[2];
  const double len2_sq = vec2[0] * vec2[0] + vec2[1] * vec2[1] + vec2[2] * vec2[2];
  const double len3_sq = vec3[0] * vec3[0] + vec3[1] * vec3[1] + vec3[2] * vec3[2];

  const double xlx = sqrt(ev[0]);
  const double xly = sqrt(ev[1]);
  const double xlz = sqrt(ev[2]);

  const double xlxi = 1.0 / (xlx * len1_sq);
  const double xlyi = 1.0 / (xly * len2_sq);
  const double xlzi = 1.0 / (xlz * len3_sq);

  V[S_XX] = xlxi * vec1[0] * vec1[0] + xlyi * vec2[0] * vec2[0] + xlzi * vec3[0] * vec3[0];
  V[S_YY] = xlxi * vec1[1] * vec1[1] + xlyi * vec2[1] * vec2[1] + xlzi * vec3[1] * vec3[1];
  V[S_ZZ] = xlxi * vec1[2] * vec1[2] + xlyi * vec2[2] * vec2[2] + xlzi * vec3[2] * vec3[2];
  V[S_XY] = xlxi * vec1[0] * vec1[1] + xlyi * vec2[0] * vec2[1] + xlzi * vec3[0] * vec3[1];
  V[S_YZ] = xlxi * vec1[1] * vec1[2] + xlyi * vec2[1] * vec2[2] + xlzi * vec3[1] * vec3[2];
  V[S_ZX] = xlxi * vec1[2] * vec1[0] + xlyi * vec2[2] * vec2[0] + xlzi * vec3[2] * vec3[0];

  if (R) mult_full33_sym33_returnT(mat_inv, V, R);
}

/* ------------------------------------------------------------------------ */
/* Neo-Hookean Cauchy stress: src/nimble_material.cc:252-310                 */
/* ------------------------------------------------------------------------ */
void orc_neohookean_stress(double K, double G, const double F[9], double sig[6])
{
  double v[6], r[9];
  orc_polar_decomp(F, v, r);

  double xj = v[S_XX] * v[S_YY] * v[S_ZZ] + 2.0 * v[S_XY] * v[S_YZ] * v[S_ZX] - v[S_XX] * v[S_YZ] * v[S_YZ] -
              v[S_YY] * v[S_ZX] * v[S_ZX] - v[S_ZZ] * v[S_XY] * v[S_XY];

  double cbrt_xj  = cbrt(xj);
  double fac      = 1.0 / (cbrt_xj * cbrt_xj);
  double pressure = 0.5 * K * (xj - 1.0 / xj);

  double bxx = v[S_XX] * v[S_XX] + v[S_XY] * v[S_YX] + v[S_XZ] * v[S_ZX];
  double byy = v[S_YX] * v[S_XY] + v[S_YY] * v[S_YY] + v[S_YZ] * v[S_ZY];
  double bzz = v[S_ZX] * v[S_XZ] + v[S_ZY] * v[S_YZ] + v[S_ZZ] * v[S_ZZ];
  double bxy = v[S_XX] * v[S_XY] + v[S_XY] * v[S_YY] + v[S_XZ] * v[S_ZY];
  double byz = v[S_YX] * v[S_XZ] + v[S_YY] * v[S_YZ] + v[S_YZ] * v[S_ZZ];
  double bzx = v[S_ZX] * v[S_XX] + v[S_ZY] * v[S_YX] + v[S_ZZ] * v[S_ZX];

  bxx = fac * bxx;
  byy = fac * byy;
  bzz = fac * bzz;
  bxy = fac * bxy;
  byz = fac * byz;
  bzx = fac * bzx;

  double trace = bxx + byy + bzz;
  bxx          = bxx - trace / 3.0;
  byy          = byy - trace / 3.0;
  bzz          = bzz - trace / 3.0;

  sig[S_XX] = pressure + G * bxx / xj;
  sig[S_YY] = pressure + G * byy / xj;
  sig[S_ZZ] = pressure + G * bzz / xj;
  sig[S_XY] = G * bxy / xj;
  sig[S_YZ] = G * byz / xj;
  sig[S_ZX] = G * bzx / xj;
}

/* ------------------------------------------------------------------------ */
/* HexElement kernels (Kokkos form: current coords x = X + u)                */
/* ------------------------------------------------------------------------ */

/* ComputeDeformationGradients_impl, src/nimble_element.h:430-502 */
int orc_hex_deformation_gradients(const double X[24], const double u[24], double Fout[72])
{
  ensure_tables();
  int    status = 0;
  double b_inv[9];
  for (int q = 0; q < 8; q++) {
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    double b[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int j = 0; j < 8; j++) {
      double rc1 = X[3 * j], rc2 = X[3 * j + 1], rc3 = X[3 * j + 2];
      double cc1 = X[3 * j] + u[3 * j];
      double cc2 = X[3 * j + 1] + u[3 * j + 1];
      double cc3 = X[3 * j + 2] + u[3 * j + 2];
      double sfd1 = g_dN[24 * q + 3 * j], sfd2 = g_dN[24 * q + 3 * j + 1], sfd3 = g_dN[24 * q + 3 * j + 2];
      a[0] += cc1 * sfd1;
      a[1] += cc1 * sfd2;
      a[2] += cc1 * sfd3;
      a[3] += cc2 * sfd1;
      a[4] += cc2 * sfd2;
      a[5] += cc2 * sfd3;
      a[6] += cc3 * sfd1;
      a[7] += cc3 * sfd2;
      a[8] += cc3 * sfd3;
      b[0] += rc1 * sfd1;
      b[1] += rc1 * sfd2;
      b[2] += rc1 * sfd3;
      b[3] += rc2 * sfd1;
      b[4] += rc2 * sfd2;
      b[5] += rc2 * sfd3;
      b[6] += rc3 * sfd1;
      b[7] += rc3 * sfd2;
      b[8] += rc3 * sfd3;
    }
    orc_invert3x3(b, b_inv, &status);
    double dg[9];
    for (int j = 0; j < 3; j++)
      for (int k = 0; k < 3; k++)
        dg[3 * j + k] = a[3 * j + 0] * b_inv[0 + k] + a[3 * j + 1] * b_inv[3 + k] + a[3 * j + 2] * b_inv[6 + k];
    double *F = &Fout[9 * q];
    F[F_XX] = dg[0];
    F[F_XY] = dg[1];
    F[F_XZ] = dg[2];
    F[F_YX] = dg[3];
    F[F_YY] = dg[4];
    F[F_YZ] = dg[5];
    F[F_ZX] = dg[6];
    F[F_ZY] = dg[7];
    F[F_ZZ] = dg[8];
  }
  return status;
}

/* ComputeNodalForces_impl, src/nimble_element.h:539-625 */
int orc_hex_nodal_forces(const double X[24], const double u[24], const double sig[48], double fout[24])
{
  ensure_tables();
  int    status = 0;
  double force[24];
  double a_inv[9];
  for (int i = 0; i < 24; i++) force[i] = 0.0;
  for (int q = 0; q < 8; q++) {
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int n = 0; n < 8; n++) {
      double cc1 = X[3 * n] + u[3 * n];
      double cc2 = X[3 * n + 1] + u[3 * n + 1];
      double cc3 = X[3 * n + 2] + u[3 * n + 2];
      double sfd1 = g_dN[24 * q + 3 * n], sfd2 = g_dN[24 * q + 3 * n + 1], sfd3 = g_dN[24 * q + 3 * n + 2];
      a[0] += cc1 * sfd1;
      a[1] += cc1 * sfd2;
      a[2] += cc1 * sfd3;
      a[3] += cc2 * sfd1;
      a[4] += cc2 * sfd2;
      a[5] += cc2 * sfd3;
      a[6] += cc3 * sfd1;
      a[7] += cc3 * sfd2;
      a[8] += cc3 * sfd3;
    }
    double jac_det = orc_invert3x3(a, a_inv, &status);
    const double *s = &sig[6 * q];
    for (int node = 0; node < 8; node++) {
      const double d0 = g_dN[24 * q + 3 * node], d1 = g_dN[24 * q + 3 * node + 1], d2 = g_dN[24 * q + 3 * node + 2];
      double dN_dx1 = d0 * a_inv[0] + d1 * a_inv[3] + d2 * a_inv[6];
      double dN_dx2 = d0 * a_inv[1] + d1 * a_inv[4] + d2 * a_inv[7];
      double dN_dx3 = d0 * a_inv[2] + d1 * a_inv[5] + d2 * a_inv[8];
      double f1 = dN_dx1 * s[S_XX] + dN_dx2 * s[S_YX] + dN_dx3 * s[S_ZX];
      double f2 = dN_dx1 * s[S_XY] + dN_dx2 * s[S_YY] + dN_dx3 * s[S_ZY];
      double f3 = dN_dx1 * s[S_XZ] + dN_dx2 * s[S_YZ] + dN_dx3 * s[S_ZZ];
      f1 *= jac_det * 1.0;
      f2 *= jac_det * 1.0;
      f3 *= jac_det * 1.0;
      force[3 * node] -= f1;
      force[3 * node + 1] -= f2;
      force[3 * node + 2] -= f3;
    }
  }
  for (int i = 0; i < 24; i++) fout[i] = force[i];
  return status;
}

/* ComputeLumpedMass + ComputeConsistentMass_impl,
 * src/nimble_element.cc:173-193, src/nimble_element.h:266-309 */
int orc_hex_lumped_mass(double density, const double X[24], double lumped[8])
{
  ensure_tables();
  int    status = 0;
  double jac_det[8], a_inv[9];
  for (int q = 0; q < 8; q++) {
    double a[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int n = 0; n < 8; n++) {
      double rc1 = X[3 * n], rc2 = X[3 * n + 1], rc3 = X[3 * n + 2];
      double sfd1 = g_dN[24 * q + 3 * n], sfd2 = g_dN[24 * q + 3 * n + 1], sfd3 = g_dN[24 * q + 3 * n + 2];
      a[0] += rc1 * sfd1;
      a[1] += rc1 * sfd2;
      a[2] += rc1 * sfd3;
      a[3] += rc2 * sfd1;
      a[4] += rc2 * sfd2;
      a[5] += rc2 * sfd3;
      a[6] += rc3 * sfd1;
      a[7] += rc3 * sfd2;
      a[8] += rc3 * sfd3;
    }
    jac_det[q] = orc_invert3x3(a, a_inv, &status);
  }
  double cm[8][8];
  for (int i = 0; i < 8; i++)
    for (int j = 0; j < 8; j++) {
      cm[i][j] = 0.0;
      for (int q = 0; q < 8; q++) cm[i][j] += 1.0 * density * g_N[q * 8 + i] * g_N[q * 8 + j] * jac_det[q];
    }
  for (int i = 0; i < 8; i++) {
    lumped[i] = 0.0;
    for (int j = 0; j < 8; j++) lumped[i] += cm[i][j];
  }
  return status;
}

/* ComputeCharacteristicLength, src/nimble_element.cc:221-260 */
double orc_hex_char_length(const double x[24])
{
  double x_min, x_max, y_min, y_max, z_min, z_max, min_d2;
  x_max = y_max = z_max = 0.0;
  min_d2 = x_min = y_min = z_min = DBL_MAX;
  for (int n = 0; n < 8; n++) {
    double nx = x[3 * n], ny = x[3 * n + 1], nz = x[3 * n + 2];
    if (nx < x_min) x_min = nx;
    if (nx > x_max) x_max = nx;
    if (ny < y_min) y_min = ny;
    if (ny > y_max) y_max = ny;
    if (nz < z_min) z_min = nz;
    if (nz > z_max) z_max = nz;
    for (int m = n + 1; m < 8; m++) {
      double mx = x[3 * m], my = x[3 * m + 1], mz = x[3 * m + 2];
      double d2 = (nx - mx) * (nx - mx) + (ny - my) * (ny - my) + (nz - mz) * (nz - mz);
      if (d2 < min_d2) min_d2 = d2;
    }
  }
  double cl      = sqrt(min_d2);
  double min_box = x_max - x_min;
  if (y_max - y_min < min_box) min_box = y_max - y_min;
  if (z_max - z_min < min_box) min_box = z_max - z_min;
  if (min_box < cl) cl = min_box;
  return cl;
}

/* ------------------------------------------------------------------------ */
/* batched seams                                                             */
/* ------------------------------------------------------------------------ */
int orc_batch_deformation_gradients(long E, const double *Xg, const double *ug, double *F)
{
  int st = 0;
  for (long e = 0; e < E; e++) st |= orc_hex_deformation_gradients(Xg + 24 * e, ug + 24 * e, F + 72 * e);
  return st;
}

void orc_batch_stress(long P, double K, double G, const double *F, double *sig)
{
  for (long p = 0; p < P; p++) orc_neohookean_stress(K, G, F + 9 * p, sig + 6 * p);
}

int orc_batch_nodal_forces(long E, const double *Xg, const double *ug, const double *sig, double *f)
{
  int st = 0;
  for (long e = 0; e < E; e++) st |= orc_hex_nodal_forces(Xg + 24 * e, ug + 24 * e, sig + 48 * e, f + 24 * e);
  return st;
}

/* ------------------------------------------------------------------------ */
/* model level                                                               */
/* ------------------------------------------------------------------------ */
static void gather_elem(const orc_mesh *mesh, const double *field, long e, double out[24])
{
  for (int n = 0; n < 8; n++) {
    long id = mesh->conn[8 * e + n];
    out[3 * n]     = field[3 * id];
    out[3 * n + 1] = field[3 * id + 1];
    out[3 * n + 2] = field[3 * id + 2];
  }
}

/* Block::ComputeLumpedMassMatrix over blocks, src/nimble_block.cc:110-146 and
 * src/nimble_model_data.cc:495-530 (serial element-order scatter). */
void orc_model_lumped_mass(const orc_mesh *mesh, double *m)
{
  for (long i = 0; i < mesh->num_nodes; i++) m[i] = 0.0;
  for (int b = 0; b < mesh->num_blocks; b++) {
    for (long e = mesh->blk_off[b]; e < mesh->blk_off[b + 1]; e++) {
      double X[24], mass[8];
      gather_elem(mesh, mesh->X, e, X);
      orc_hex_lumped_mass(mesh->rho[b], X, mass);
      for (int n = 0; n < 8; n++) m[mesh->conn[8 * e + n]] += mass[n];
    }
  }
}

/* BlockBase::ComputeCriticalTimeStep per block, minimum over blocks,
 * src/nimble_block_base.cc:52-84, src/nimble_kokkos_model_data.cc:1139-1141 */
double orc_model_critical_dt(const orc_mesh *mesh, const double *u)
{
  double dt = DBL_MAX;
  for (int b = 0; b < mesh->num_blocks; b++) {
    double sound_speed = sqrt(mesh->K[b] / mesh->rho[b]);
    double bdt         = DBL_MAX;
    for (long e = mesh->blk_off[b]; e < mesh->blk_off[b + 1]; e++) {
      double x[24];
      for (int n = 0; n < 8; n++) {
        long id = mesh->conn[8 * e + n];
        for (int i = 0; i < 3; i++) x[3 * n + i] = mesh->X[3 * id + i] + (u ? u[3 * id + i] : 0.0);
      }
      double edt = orc_hex_char_length(x) / sound_speed;
      if (edt < bdt) bdt = edt;
    }
    if (bdt < dt) dt = bdt;
  }
  return dt;
}

static int element_force(const orc_mesh *mesh, int b, long e, const double *u, double fe[24], double *Fip,
                         double *sip)
{
  double X[24], ue[24], F[72], s[48];
  gather_elem(mesh, mesh->X, e, X);
  gather_elem(mesh, u, e, ue);
  int st = orc_hex_deformation_gradients(X, ue, F);
  for (int q = 0; q < 8; q++) orc_neohookean_stress(mesh->K[b], mesh->G[b], F + 9 * q, s + 6 * q);
  st |= orc_hex_nodal_forces(X, ue, s, fe);
  if (Fip) memcpy(Fip + 72 * e, F, sizeof(F));
  if (sip) memcpy(sip + 48 * e, s, sizeof(s));
  return st;
}

/* ModelData::ComputeInternalForce, serial, src/nimble_model_data.cc:620-667 with
 * Block functor src/nimble_block.cc:209-386 (x = X + u as the Kokkos build). */
int orc_model_internal_force(const orc_mesh *mesh, const double *u, double *f, double *Fip, double *sip)
{
  int st = 0;
  for (long i = 0; i < 3 * mesh->num_nodes; i++) f[i] = 0.0;
  for (int b = 0; b < mesh->num_blocks; b++) {
    for (long e = mesh->blk_off[b]; e < mesh->blk_off[b + 1]; e++) {
      double fe[24];
      st |= element_force(mesh, b, e, u, fe, Fip, sip);
      for (int n = 0; n < 8; n++) {
        long id = mesh->conn[8 * e + n];
        for (int i = 0; i < 3; i++) f[3 * id + i] += fe[3 * n + i];
      }
    }
  }
  return st;
}

void orc_build_node_csr(const orc_mesh *mesh, long *node_ptr, long *node_ent)
{
  long N = mesh->num_nodes, E = mesh->num_elem;
  for (long i = 0; i <= N; i++) node_ptr[i] = 0;
  for (long k = 0; k < 8 * E; k++) node_ptr[mesh->conn[k] + 1]++;
  for (long i = 0; i < N; i++) node_ptr[i + 1] += node_ptr[i];
  long *fill = (long *)malloc(sizeof(long) * (size_t)(N > 0 ? N : 1));
  for (long i = 0; i < N; i++) fill[i] = node_ptr[i];
  for (long k = 0; k < 8 * E; k++) node_ent[fill[mesh->conn[k]]++] = k; /* ascending k => element order */
  free(fill);
}

int orc_model_internal_force_omp(const orc_mesh *mesh, const double *u, double *f, double *Fip, double *sip,
                                 double *slots, const long *node_ptr, const long *node_ent)
{
  int st = 0;
  for (int b = 0; b < mesh->num_blocks; b++) {
    long e0 = mesh->blk_off[b], e1 = mesh->blk_off[b + 1];
#pragma omp parallel for schedule(static) reduction(| : st)
    for (long e = e0; e < e1; e++) st |= element_force(mesh, b, e, u, slots + 24 * e, Fip, sip);
  }
#pragma omp parallel for schedule(static)
  for (long i = 0; i < mesh->num_nodes; i++) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    for (long k = node_ptr[i]; k < node_ptr[i + 1]; k++) {
      long ent = node_ent[k];
      s0 += slots[3 * ent];
      s1 += slots[3 * ent + 1];
      s2 += slots[3 * ent + 2];
    }
    f[3 * i] = s0;
    f[3 * i + 1] = s1;
    f[3 * i + 2] = s2;
  }
  return st;
}

/* BoundaryConditionManager::ApplyInitialConditions,
 * src/nimble_boundary_condition_manager.h:93-134 (constant magnitudes) */
void orc_apply_initial_conditions(const orc_bcs *bcs, double *v)
{
  for (long k = 0; k < bcs->count; k++)
    if (bcs->kind[k] == 0) v[3 * bcs->node[k] + bcs->coord[k]] = bcs->value[k];
}

/* BoundaryConditionManager::ApplyKinematicBC (explicit scheme),
 * src/nimble_boundary_condition_manager.h:136-217 */
void orc_apply_kinematic_bc(const orc_bcs *bcs, double t_cur, double t_prev, const double *u, double *v)
{
  double delta_t = t_cur - t_prev;
  for (long k = 0; k < bcs->count; k++) {
    long i = 3 * (long)bcs->node[k] + bcs->coord[k];
    if (bcs->kind[k] == 1) {
      v[i] = bcs->value[k];
    } else if (bcs->kind[k] == 2 && delta_t > 0.0) {
      v[i] = (bcs->value[k] - u[i]) / delta_t;
    }
  }
}

int orc_explicit_steps(const orc_mesh *mesh, const orc_bcs *bcs, orc_state *st, int step_begin, int step_end,
                       int num_load_steps, int output_frequency, double dt_user, const int *reaction_nodes,
                       long num_reaction_nodes, int reaction_coord, double *reaction, int threads)
{
  const long N3 = 3 * mesh->num_nodes;
  int        status = 0;
  long      *node_ptr = NULL, *node_ent = NULL;
  double    *slots = NULL;
  if (threads > 1) {
    node_ptr = (long *)malloc(sizeof(long) * (size_t)(mesh->num_nodes + 1));
    node_ent = (long *)malloc(sizeof(long) * (size_t)(8 * mesh->num_elem));
    slots    = (double *)malloc(sizeof(double) * (size_t)(24 * mesh->num_elem));
    orc_build_node_csr(mesh, node_ptr, node_ent);
#ifdef _OPENMP
    omp_set_num_threads(threads);
#endif
  }
  double time_current = st->time_current, time_previous;
  for (int step = step_begin; step < step_end; step++) {
    int is_output_step = 0;
    if (output_frequency != 0)
      if (step % output_frequency == 0 || step == num_load_steps - 1) is_output_step = 1;

    time_previous          = time_current;
    time_current          += dt_user;
    double delta_time      = time_current - time_previous;
    double half_delta_time = 0.5 * delta_time;

    /* velocity += half_delta_time * acceleration (src/nimble_view.h:181-195) */
    double prod = half_delta_time * 1.0;
    for (long i = 0; i < N3; i++) st->v[i] += prod * st->a[i];
    orc_apply_kinematic_bc(bcs, time_current, time_previous, st->u, st->v);
    prod = delta_time * 1.0;
    for (long i = 0; i < N3; i++) st->u[i] += prod * st->v[i];
    orc_apply_kinematic_bc(bcs, time_current, time_previous, st->u, st->v);

    if (threads > 1)
      status |= orc_model_internal_force_omp(mesh, st->u, st->f, st->Fip, st->sip, slots, node_ptr, node_ent);
    else
      status |= orc_model_internal_force(mesh, st->u, st->f, st->Fip, st->sip);

    /* explicit_time_integrator.cc:250-256, external force is zero */
    for (long i = 0; i < mesh->num_nodes; i++) {
      const double oneOverM = 1.0 / st->m[i];
      st->a[3 * i]     = oneOverM * (st->f[3 * i] + 0.0);
      st->a[3 * i + 1] = oneOverM * (st->f[3 * i + 1] + 0.0);
      st->a[3 * i + 2] = oneOverM * (st->f[3 * i + 2] + 0.0);
    }
    prod = half_delta_time * 1.0;
    for (long i = 0; i < N3; i++) st->v[i] += prod * st->a[i];
    if (is_output_step) orc_apply_kinematic_bc(bcs, time_current, time_previous, st->u, st->v);

    if (reaction) {
      double s = 0.0;
      for (long k = 0; k < num_reaction_nodes; k++) s += st->f[3 * (long)reaction_nodes[k] + reaction_coord];
      reaction[step] = -s;
    }
  }
  st->time_current = time_current;
  free(node_ptr);
  free(node_ent);
  free(slots);
  return status;
}

/* ------------------------------------------------------------------------ */
/* synthetic jittered cube                                                   */
/* ------------------------------------------------------------------------ */
unsigned long long orc_splitmix64(unsigned long long x)
{
  unsigned long long z = x + 0x9E3779B97F4A7C15ULL;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

void orc_jittered_cube(int n, unsigned long long seed, double jitter_frac, double *X, int *conn)
{
  const long n1 = n + 1;
  const unsigned long long key = orc_splitmix64(seed);
  for (long k = 0; k <= n; k++)
    for (long j = 0; j <= n; j++)
      for (long i = 0; i <= n; i++) {
        long id = i + n1 * (j + n1 * k);
        X[3 * id]     = (double)i / (double)n;
        X[3 * id + 1] = (double)j / (double)n;
        X[3 * id + 2] = (double)k / (double)n;
        if (i > 0 && i < n && j > 0 && j < n && k > 0 && k < n) {
          for (int d = 0; d < 3; d++) {
            unsigned long long r = orc_splitmix64(key ^ (unsigned long long)(3 * id + d));
            double U             = (double)(r >> 11) * (1.0 / 9007199254740992.0);
            X[3 * id + d] += (2.0 * U - 1.0) * jitter_frac / (double)n;
          }
        }
      }
  for (long k = 0; k < n; k++)
    for (long j = 0; j < n; j++)
      for (long i = 0; i < n; i++) {
        long e   = i + n * (j + n * k);
        long n0  = i + n1 * (j + n1 * k);
        int *c   = conn + 8 * e;
        c[0] = (int)n0;
        c[1] = (int)(n0 + 1);
        c[2] = (int)(n0 + 1 + n1);
        c[3] = (int)(n0 + n1);
        c[4] = (int)(n0 + n1 * n1);
        c[5] = (int)(n0 + 1 + n1 * n1);
        c[6] = (int)(n0 + 1 + n1 + n1 * n1);
        c[7] = (int)(n0 + n1 + n1 * n1);
      }
}
