// nsm_mixed.hip -- mixed-precision fused HEX8 element kernel.
//
// F, Cauchy stress and B^T.sigma are computed in FP32; node coordinates and
// displacements stay FP64 in HBM and are turned into FP32 only after the
// element's node 0 has been subtracted in FP64, so the small displacement
// gradient does not cancel against the large absolute position.  The
// Neo-Hookean stress uses the closed form on the displacement gradient
// H = F - I:
//   B - I = H + H^T + H H^T,  J - 1 = tr H + ((tr H)^2 - tr H^2)/2 + det H,
//   sigma = K/2 (J - 1/J) I + G J^{-5/3} dev(B)
// which equals NimbleSM's Polar_Decomp route (V^2 = F F^T, det V = det F,
// src/nimble_material.cc:261-309) without its cancellation at F ~ I.  The
// force uses det(a) a^{-1} = adj(a), so no division is needed:
//   f_n = - sum_q sum_m dN_n/dxi_m(q) [adj(a_q) sigma_q]_{m,c}
// (src/nimble_element.h:586-617 with the weights w_q = 1).
//
// Mapping as in the FP64 kernel: 8 lanes per element; lane k gathers node k,
// evaluates Gauss point k, and finally accumulates the force of node k.
#include "nsm_kernels.h"

namespace nsm {
namespace {

struct HexTablesF {
  float dN[192];  // [ipt][node][dir]
};

constexpr HexTablesF make_tables_f()
{
  HexTablesF   t{};
  const double val      = 0.577350269189626;
  const int    sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  for (int q = 0; q < 8; q++)
    for (int n = 0; n < 8; n++) {
      // dN_n/dxi_d = s_n^d / 8 * prod_{e != d} (1 + s_n^e g_q^e)
      for (int d = 0; d < 3; d++) {
        double v = 0.125 * sg[n][d];
        for (int e2 = 0; e2 < 3; e2++)
          if (e2 != d) v *= (1.0 + sg[n][e2] * (sg[q][e2] > 0 ? val : -val));
        t.dN[24 * q + 3 * n + d] = (float)v;
      }
    }
  return t;
}

__constant__ HexTablesF c_hexf = make_tables_f();

constexpr int kEPW     = 32;
constexpr int kThreads = 8 * kEPW;

__device__ __forceinline__ void adjugate(const float m[9], float adj[9])
{
  // adj = det(m) m^{-1}, row-major
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

struct LdsMixed {
  float d[kEPW][8][8];   // per node: dX xyz, du xyz (relative to node 0), pad
  float M[kEPW][8][12];  // per point: adj(a) sigma (9), pad
};

__global__ __launch_bounds__(kThreads) void k_element_mixed(int64_t e_begin, int64_t e_end,
                                                            const int* __restrict__ conn,
                                                            const double* __restrict__ X,
                                                            const double* __restrict__ u, float K, float G,
                                                            float* __restrict__ Fip, float* __restrict__ sip,
                                                            float* __restrict__ slots, int* __restrict__ status)
{
  __shared__ LdsMixed lds;
  const int     lane  = threadIdx.x & 63;
  const int     el    = threadIdx.x >> 3;
  const int     k     = threadIdx.x & 7;
  const int64_t e     = e_begin + (int64_t)blockIdx.x * kEPW + el;
  const bool    valid = e < e_end;

  double x0 = 0.0, x1 = 0.0, x2 = 0.0, u0 = 0.0, u1 = 0.0, u2 = 0.0;
  if (valid) {
    const int64_t nd = conn[e * 8 + k];
    x0 = X[3 * nd];
    x1 = X[3 * nd + 1];
    x2 = X[3 * nd + 2];
    u0 = u[3 * nd];
    u1 = u[3 * nd + 1];
    u2 = u[3 * nd + 2];
  }
  // node 0 of the element sits in lane (lane & ~7) of the same wavefront
  const int src = lane & ~7;
  const double bx0 = __shfl(x0, src, 64), bx1 = __shfl(x1, src, 64), bx2 = __shfl(x2, src, 64);
  const double bu0 = __shfl(u0, src, 64), bu1 = __shfl(u1, src, 64), bu2 = __shfl(u2, src, 64);
  {
    float* d = lds.d[el][k];
    d[0]     = (float)(x0 - bx0);
    d[1]     = (float)(x1 - bx1);
    d[2]     = (float)(x2 - bx2);
    d[3]     = (float)(u0 - bu0);
    d[4]     = (float)(u1 - bu1);
    d[5]     = (float)(u2 - bu2);
  }
  __syncthreads();

  int bad = 0;
  if (valid) {
    const int q = k;
    float     b[9], g[9];
#pragma unroll
    for (int i = 0; i < 9; i++) b[i] = g[i] = 0.0f;
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const float* d  = lds.d[el][j];
      const float  s1 = c_hexf.dN[24 * q + 3 * j], s2 = c_hexf.dN[24 * q + 3 * j + 1],
                  s3 = c_hexf.dN[24 * q + 3 * j + 2];
#pragma unroll
      for (int i = 0; i < 3; i++) {
        b[3 * i]     = fmaf(d[i], s1, b[3 * i]);
        b[3 * i + 1] = fmaf(d[i], s2, b[3 * i + 1]);
        b[3 * i + 2] = fmaf(d[i], s3, b[3 * i + 2]);
        g[3 * i]     = fmaf(d[3 + i], s1, g[3 * i]);
        g[3 * i + 1] = fmaf(d[3 + i], s2, g[3 * i + 1]);
        g[3 * i + 2] = fmaf(d[3 + i], s3, g[3 * i + 2]);
      }
    }
    // H = g b^{-1}
    float adjb[9];
    adjugate(b, adjb);
    const float detb = b[0] * adjb[0] + b[1] * adjb[3] + b[2] * adjb[6];
    bad |= !(detb > 0.0f);
    const float rdb = 1.0f / detb;
    float       H[9];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int m = 0; m < 3; m++)
        H[3 * i + m] = (g[3 * i] * adjb[m] + g[3 * i + 1] * adjb[3 + m] + g[3 * i + 2] * adjb[6 + m]) * rdb;

    float* Fo = Fip + (e * 8 + q) * 9;  // K_F order XX,YY,ZZ,XY,YZ,ZX,YX,ZY,XZ
    Fo[0]     = 1.0f + H[0];
    Fo[1]     = 1.0f + H[4];
    Fo[2]     = 1.0f + H[8];
    Fo[3]     = H[1];
    Fo[4]     = H[5];
    Fo[5]     = H[6];
    Fo[6]     = H[3];
    Fo[7]     = H[7];
    Fo[8]     = H[2];

    // B - I = H + H^T + H H^T
    const float bxx = 2.0f * H[0] + (H[0] * H[0] + H[1] * H[1] + H[2] * H[2]);
    const float byy = 2.0f * H[4] + (H[3] * H[3] + H[4] * H[4] + H[5] * H[5]);
    const float bzz = 2.0f * H[8] + (H[6] * H[6] + H[7] * H[7] + H[8] * H[8]);
    const float bxy = H[1] + H[3] + (H[0] * H[3] + H[1] * H[4] + H[2] * H[5]);
    const float byz = H[5] + H[7] + (H[3] * H[6] + H[4] * H[7] + H[5] * H[8]);
    const float bzx = H[6] + H[2] + (H[6] * H[0] + H[7] * H[1] + H[8] * H[2]);
    // J - 1 without cancellation
    const float trH  = H[0] + H[4] + H[8];
    const float trH2 = H[0] * H[0] + H[4] * H[4] + H[8] * H[8] + 2.0f * (H[1] * H[3] + H[2] * H[6] + H[5] * H[7]);
    const float detH = H[0] * (H[4] * H[8] - H[5] * H[7]) - H[1] * (H[3] * H[8] - H[5] * H[6]) +
                       H[2] * (H[3] * H[7] - H[4] * H[6]);
    const float dJ   = trH + 0.5f * (trH * trH - trH2) + detH;
    const float J    = 1.0f + dJ;
    const float rJ   = 1.0f / J;
    // J^{-2/3} via the hardware log2/exp2 (J ~ 1: absolute error ~1e-7 on a
    // factor that multiplies the small deviator)
    const float fac      = __builtin_amdgcn_exp2f(-0.666666667f * __builtin_amdgcn_logf(J));
    const float pressure = 0.5f * K * dJ * (2.0f + dJ) * rJ;
    const float tr3      = (bxx + byy + bzz) * (1.0f / 3.0f);
    const float coef     = G * fac * rJ;
    float       sg[6];
    sg[0] = pressure + coef * (bxx - tr3);
    sg[1] = pressure + coef * (byy - tr3);
    sg[2] = pressure + coef * (bzz - tr3);
    sg[3] = coef * bxy;
    sg[4] = coef * byz;
    sg[5] = coef * bzx;
    float* so = sip + (e * 8 + q) * 6;
#pragma unroll
    for (int i = 0; i < 6; i++) so[i] = sg[i];

    // M = adj(a) sigma with a = b + g the current Jacobian
    float a[9], adja[9];
#pragma unroll
    for (int i = 0; i < 9; i++) a[i] = b[i] + g[i];
    adjugate(a, adja);
    const float deta = a[0] * adja[0] + a[1] * adja[3] + a[2] * adja[6];
    bad |= !(deta > 0.0f);
    // sigma as full symmetric matrix rows: [xx xy zx; xy yy yz; zx yz zz]
    const float S[9] = {sg[0], sg[3], sg[5], sg[3], sg[1], sg[4], sg[5], sg[4], sg[2]};
    float*      M    = lds.M[el][q];
#pragma unroll
    for (int m = 0; m < 3; m++)
#pragma unroll
      for (int c = 0; c < 3; c++) M[3 * m + c] = adja[3 * m] * S[c] + adja[3 * m + 1] * S[3 + c] + adja[3 * m + 2] * S[6 + c];
  }
  __syncthreads();
  if (valid) {
    const int n  = k;
    float     f0 = 0.0f, f1 = 0.0f, f2 = 0.0f;
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const float* M  = lds.M[el][q];
      const float  d0 = c_hexf.dN[24 * q + 3 * n], d1 = c_hexf.dN[24 * q + 3 * n + 1],
                  d2 = c_hexf.dN[24 * q + 3 * n + 2];
      f0 += d0 * M[0] + d1 * M[3] + d2 * M[6];
      f1 += d0 * M[1] + d1 * M[4] + d2 * M[7];
      f2 += d0 * M[2] + d1 * M[5] + d2 * M[8];
    }
    float* so = slots + (e * 8 + n) * 3;
    so[0]     = -f0;
    so[1]     = -f1;
    so[2]     = -f2;
  }
  if (bad) atomicOr(status, 1);
}

inline unsigned grid_of(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void launch_element_mixed(int64_t e_begin, int64_t e_end, const int* conn, const double* X, const double* u, float K,
                          float G, float* Fip, float* sip, float* slots, int* status, hipStream_t s)
{
  if (e_end <= e_begin) return;
  hipLaunchKernelGGL(k_element_mixed, dim3(grid_of(e_end - e_begin, kEPW)), dim3(kThreads), 0, s, e_begin, e_end,
                     conn, X, u, K, G, Fip, sip, slots, status);
}

}  // namespace nsm
