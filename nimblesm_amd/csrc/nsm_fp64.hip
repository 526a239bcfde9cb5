// nsm_fp64.hip -- FP64 parity path: fused HEX8 element kernel (F -> Neo-Hookean
// stress -> B^T.sigma into per-element force slots), the element / material
// seam kernels and the lumped-mass / critical-time-step setup kernels.
//
// Compiled with -ffp-contract=off: every FP64 expression keeps NimbleSM's
// evaluation order so results match the reference's FP64 run bit for bit.
// Work mapping: one 64-wide wavefront holds 8 elements, lane
// k of an element serves as node k (gather, force output) and as Gauss point
// k (F, stress); LDS carries the per-element node data and per-point
// inverse Jacobians between the two roles.
#include "nsm_ensemble.h"
#include "nsm_fp64_math.h"
#include "nsm_kernels.h"

namespace nsm {
namespace {

// ---------------------------------------------------------------------------
// HEX8 tables, evaluated at compile time with IEEE double arithmetic in the
// order of src/nimble_element.cc:55-171 (Gauss constant written as in :58).
// ---------------------------------------------------------------------------
struct HexTables {
  double N[64];    // [ipt][node]
  double dN[192];  // [ipt][node][dir]
};

constexpr HexTables make_hex_tables()
{
  HexTables    t{};
  const double val     = 0.577350269189626;
  const int    sg[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                           {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};
  const double c       = 1.0 / 8.0;
  for (int q = 0; q < 8; q++) {
    const double r = sg[q][0] > 0 ? val : -val;
    const double s = sg[q][1] > 0 ? val : -val;
    const double u = sg[q][2] > 0 ? val : -val;
    t.N[8 * q + 0] = c * (1.0 - r) * (1.0 - s) * (1.0 - u);
    t.N[8 * q + 1] = c * (1.0 + r) * (1.0 - s) * (1.0 - u);
    t.N[8 * q + 2] = c * (1.0 + r) * (1.0 + s) * (1.0 - u);
    t.N[8 * q + 3] = c * (1.0 - r) * (1.0 + s) * (1.0 - u);
    t.N[8 * q + 4] = c * (1.0 - r) * (1.0 - s) * (1.0 + u);
    t.N[8 * q + 5] = c * (1.0 + r) * (1.0 - s) * (1.0 + u);
    t.N[8 * q + 6] = c * (1.0 + r) * (1.0 + s) * (1.0 + u);
    t.N[8 * q + 7] = c * (1.0 - r) * (1.0 + s) * (1.0 + u);
    double* d = &t.dN[24 * q];
    d[0]  = -c * (1.0 - s) * (1.0 - u);
    d[1]  = -c * (1.0 - r) * (1.0 - u);
    d[2]  = -c * (1.0 - r) * (1.0 - s);
    d[3]  = c * (1.0 - s) * (1.0 - u);
    d[4]  = -c * (1.0 + r) * (1.0 - u);
    d[5]  = -c * (1.0 + r) * (1.0 - s);
    d[6]  = c * (1.0 + s) * (1.0 - u);
    d[7]  = c * (1.0 + r) * (1.0 - u);
    d[8]  = -c * (1.0 + r) * (1.0 + s);
    d[9]  = -c * (1.0 + s) * (1.0 - u);
    d[10] = c * (1.0 - r) * (1.0 - u);
    d[11] = -c * (1.0 - r) * (1.0 + s);
    d[12] = -c * (1.0 - s) * (1.0 + u);
    d[13] = -c * (1.0 - r) * (1.0 + u);
    d[14] = c * (1.0 - r) * (1.0 - s);
    d[15] = c * (1.0 - s) * (1.0 + u);
    d[16] = -c * (1.0 + r) * (1.0 + u);
    d[17] = c * (1.0 + r) * (1.0 - s);
    d[18] = c * (1.0 + s) * (1.0 + u);
    d[19] = c * (1.0 + r) * (1.0 + u);
    d[20] = c * (1.0 + r) * (1.0 + s);
    d[21] = -c * (1.0 + s) * (1.0 + u);
    d[22] = c * (1.0 - r) * (1.0 + u);
    d[23] = c * (1.0 - r) * (1.0 + s);
  }
  return t;
}

__constant__ HexTables c_hex = make_hex_tables();

constexpr int kEPW     = 32;        // elements per 256-thread workgroup
constexpr int kThreads = 8 * kEPW;  // 8 lanes per element

using namespace fp64;

// a = sum_j x_j (x) dN_j/dxi at ipt q with x = X + u, and b = sum_j X_j (x)
// dN_j/dxi (src/nimble_element.h:442-482, which forms both in one loop; each
// entry is its own chain of products and sums, so forming them apart gives
// the same bits).  xs/us are [8][3]: b alone, then both.
__device__ __forceinline__ void reference_jacobian(int q, const double (*xs)[3], double b[9])
{
  for (int i = 0; i < 9; i++) b[i] = 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const double rc1 = xs[j][0], rc2 = xs[j][1], rc3 = xs[j][2];
    const double s1 = c_hex.dN[24 * q + 3 * j], s2 = c_hex.dN[24 * q + 3 * j + 1], s3 = c_hex.dN[24 * q + 3 * j + 2];
    b[0] += rc1 * s1;
    b[1] += rc1 * s2;
    b[2] += rc1 * s3;
    b[3] += rc2 * s1;
    b[4] += rc2 * s2;
    b[5] += rc2 * s3;
    b[6] += rc3 * s1;
    b[7] += rc3 * s2;
    b[8] += rc3 * s3;
  }
}

// both in one pass over the nodes
__device__ __forceinline__ void jacobians(int q, const double (*xs)[3], const double (*us)[3], double a[9],
                                          double b[9])
{
  for (int i = 0; i < 9; i++) a[i] = b[i] = 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const double rc1 = xs[j][0], rc2 = xs[j][1], rc3 = xs[j][2];
    const double cc1 = rc1 + us[j][0];
    const double cc2 = rc2 + us[j][1];
    const double cc3 = rc3 + us[j][2];
    const double s1 = c_hex.dN[24 * q + 3 * j], s2 = c_hex.dN[24 * q + 3 * j + 1], s3 = c_hex.dN[24 * q + 3 * j + 2];
    a[0] += cc1 * s1;
    a[1] += cc1 * s2;
    a[2] += cc1 * s3;
    a[3] += cc2 * s1;
    a[4] += cc2 * s2;
    a[5] += cc2 * s3;
    a[6] += cc3 * s1;
    a[7] += cc3 * s2;
    a[8] += cc3 * s3;
    b[0] += rc1 * s1;
    b[1] += rc1 * s2;
    b[2] += rc1 * s3;
    b[3] += rc2 * s1;
    b[4] += rc2 * s2;
    b[5] += rc2 * s3;
    b[6] += rc3 * s1;
    b[7] += rc3 * s2;
    b[8] += rc3 * s3;
  }
}

// The same two Jacobians from the current coordinates x = X + u formed once
// per node (xc), as the element kernels stage them: each node's sum is the
// one the functions above form in every lane, so the bits are the same.
__device__ __forceinline__ void current_jacobian_x(int q, const double (*xc)[3], double a[9])
{
  for (int i = 0; i < 9; i++) a[i] = 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const double cc1 = xc[j][0], cc2 = xc[j][1], cc3 = xc[j][2];
    const double s1 = c_hex.dN[24 * q + 3 * j], s2 = c_hex.dN[24 * q + 3 * j + 1], s3 = c_hex.dN[24 * q + 3 * j + 2];
    a[0] += cc1 * s1;
    a[1] += cc1 * s2;
    a[2] += cc1 * s3;
    a[3] += cc2 * s1;
    a[4] += cc2 * s2;
    a[5] += cc2 * s3;
    a[6] += cc3 * s1;
    a[7] += cc3 * s2;
    a[8] += cc3 * s3;
  }
}

__device__ __forceinline__ void jacobians_x(int q, const double (*xs)[3], const double (*xc)[3], double a[9],
                                            double b[9])
{
  for (int i = 0; i < 9; i++) a[i] = b[i] = 0.0;
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const double rc1 = xs[j][0], rc2 = xs[j][1], rc3 = xs[j][2];
    const double cc1 = xc[j][0], cc2 = xc[j][1], cc3 = xc[j][2];
    const double s1 = c_hex.dN[24 * q + 3 * j], s2 = c_hex.dN[24 * q + 3 * j + 1], s3 = c_hex.dN[24 * q + 3 * j + 2];
    a[0] += cc1 * s1;
    a[1] += cc1 * s2;
    a[2] += cc1 * s3;
    a[3] += cc2 * s1;
    a[4] += cc2 * s2;
    a[5] += cc2 * s3;
    a[6] += cc3 * s1;
    a[7] += cc3 * s2;
    a[8] += cc3 * s3;
    b[0] += rc1 * s1;
    b[1] += rc1 * s2;
    b[2] += rc1 * s3;
    b[3] += rc2 * s1;
    b[4] += rc2 * s2;
    b[5] += rc2 * s3;
    b[6] += rc3 * s1;
    b[7] += rc3 * s2;
    b[8] += rc3 * s3;
  }
}

// F = a b^{-1} in K_F storage (src/nimble_element.h:484-500)
__device__ __forceinline__ void def_grad(const double a[9], const double bi[9], double F[9])
{
  double dg[9];
#pragma unroll
  for (int j = 0; j < 3; j++)
#pragma unroll
    for (int k = 0; k < 3; k++) dg[3 * j + k] = a[3 * j] * bi[k] + a[3 * j + 1] * bi[3 + k] + a[3 * j + 2] * bi[6 + k];
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

// Force on node n from all 8 points, in ipt order, from the per-point inverse
// Jacobian / det (ai[q][0..8], ai[q][9]) and stress sg[q][0..5]
// (src/nimble_element.h:586-617).
__device__ __forceinline__ void node_force(int n, const double (*ai)[10], const double (*sg)[6], double f[3])
{
  f[0] = f[1] = f[2] = 0.0;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    const double d0 = c_hex.dN[24 * q + 3 * n], d1 = c_hex.dN[24 * q + 3 * n + 1], d2 = c_hex.dN[24 * q + 3 * n + 2];
    const double* A  = ai[q];
    const double* s  = sg[q];
    const double  x1 = d0 * A[0] + d1 * A[3] + d2 * A[6];
    const double  x2 = d0 * A[1] + d1 * A[4] + d2 * A[7];
    const double  x3 = d0 * A[2] + d1 * A[5] + d2 * A[8];
    double        f1 = x1 * s[S_XX] + x2 * s[S_XY] + x3 * s[S_ZX];
    double        f2 = x1 * s[S_XY] + x2 * s[S_YY] + x3 * s[S_YZ];
    double        f3 = x1 * s[S_ZX] + x2 * s[S_YZ] + x3 * s[S_ZZ];
    const double  jw = A[9] * 1.0;
    f1 *= jw;
    f2 *= jw;
    f3 *= jw;
    f[0] -= f1;
    f[1] -= f2;
    f[2] -= f3;
  }
}

struct Lds {
  union {
    struct {
      double x[kEPW][8][3];   // reference coordinates X
      double xc[kEPW][8][3];  // current coordinates X + u
    } nodes;
    struct {
      double ai[kEPW][8][10];
      double sg[kEPW][8][6];
    } pts;
  };
};

// kMode: kElementFull (F, stress and forces), kElementDefGrad (F only, the
// first kernel of src/nimble_kokkos_model_data.cc:1196-1227) or
// kElementForcesFromStress (forces from the stress stored in sip, the
// kernel after a block material interface's ComputeStress,
// src/nimble_kokkos_model_data.cc:1229-1257).
//
// kRefTable: each point's reference inverse Jacobian b^{-1} and det b are
// read from rjac (k_reference_jacobians; b depends on X only), else formed
// here as the reference does in every call (src/nimble_element.h:442-500);
// the same operations either way, so the same bits.
template <int kMode, bool kRefTable>
__device__ __forceinline__ void element_fp64(Lds& lds, int64_t e, bool valid, double K, double G, int64_t E_total,
                                             const int* __restrict__ conn, const double* __restrict__ X,
                                             const double* __restrict__ u, const double* __restrict__ rjac,
                                             double* __restrict__ Fip, double* __restrict__ sip,
                                             double* __restrict__ slots, int* __restrict__ status)
{
  const int el = threadIdx.x >> 3;
  const int k  = threadIdx.x & 7;
  double    bi[9], detb = 1.0;
  if (kRefTable && kMode != kElementForcesFromStress && valid) {
    const double2* r = (const double2*)(rjac + (e * 8 + k) * 10);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const double2 v = r[i];
      bi[2 * i]       = v.x;
      bi[2 * i + 1]   = v.y;
    }
    const double2 v = r[4];
    bi[8]           = v.x;
    detb            = v.y;
  }
  if (valid) {
    const int64_t nd = conn[e * 8 + k];
    lds.nodes.x[el][k][0] = X[3 * nd];
    lds.nodes.x[el][k][1] = X[3 * nd + 1];
    lds.nodes.x[el][k][2] = X[3 * nd + 2];
    lds.nodes.xc[el][k][0] = lds.nodes.x[el][k][0] + u[3 * nd];
    lds.nodes.xc[el][k][1] = lds.nodes.x[el][k][1] + u[3 * nd + 1];
    lds.nodes.xc[el][k][2] = lds.nodes.x[el][k][2] + u[3 * nd + 2];
  }
  __syncthreads();

  double ainv[10], sig[6];
  int    bad = 0;
  if (valid) {
    double a[9];
    if (kRefTable || kMode == kElementForcesFromStress) {
      current_jacobian_x(k, lds.nodes.xc[el], a);
    } else {
      double b[9];
      jacobians_x(k, lds.nodes.x[el], lds.nodes.xc[el], a, b);
      detb = invert3x3<false>(b, bi);
    }
    if (kMode != kElementForcesFromStress) {
      double F[9];
      bad |= detb <= 0.0;
      def_grad(a, bi, F);
      // groups of 8 elements (fp64_ip_index; no stores when Fip is null)
      if (Fip)
#pragma unroll
        for (int i = 0; i < 9; i++) Fip[fp64_ip_index(e, 9, i, k)] = F[i];
      if (kMode == kElementFull) {
        neohookean_stress(K, G, F, sig);
        if (sip)
#pragma unroll
          for (int i = 0; i < 6; i++) sip[fp64_ip_index(e, 6, i, k)] = sig[i];
      }
    } else {
#pragma unroll
      for (int i = 0; i < 6; i++) sig[i] = sip[fp64_ip_index(e, 6, i, k)];
    }
    // the force routine rebuilds a from the same x = X + u in the same order
    if (kMode != kElementDefGrad) {
      ainv[9] = invert3x3(a, ainv);
      bad |= ainv[9] <= 0.0;
    }
  }
  if (kMode == kElementDefGrad) {
    if (bad) atomicOr(status, 1);
    return;
  }
  __syncthreads();  // node data no longer needed
  if (valid) {
#pragma unroll
    for (int i = 0; i < 10; i++) lds.pts.ai[el][k][i] = ainv[i];
#pragma unroll
    for (int i = 0; i < 6; i++) lds.pts.sg[el][k][i] = sig[i];
  }
  __syncthreads();
  if (valid) {
    double f[3];
    node_force(k, lds.pts.ai[el], lds.pts.sg[el], f);
    const int64_t ss = 8 * round_up8(E_total), sl = fp64_slot(e, k);  // [comp][slot]
    slots[sl]          = f[0];
    slots[ss + sl]     = f[1];
    slots[2 * ss + sl] = f[2];
  }
  if (bad) atomicOr(status, 1);
}

// The full evaluation is held to 5 waves per SIMD (at most 96 registers):
// more would cost occupancy, fewer spill.
template <int kMode, bool kRefTable>
__global__ __launch_bounds__(kThreads, kMode == kElementFull && kRefTable ? 5 : 1) void k_element_fp64(
    const BlockTable<double> bt, int64_t E_total, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ u, const double* __restrict__ rjac, double* __restrict__ Fip,
    double* __restrict__ sip, double* __restrict__ slots, int* __restrict__ status)
{
  __shared__ Lds lds;
  const int     el    = threadIdx.x >> 3;
  const int64_t e     = bt.off[0] + (int64_t)xcd_tile(blockIdx.x, gridDim.x) * kEPW + el;
  const bool    valid = e < bt.off[bt.count];
  const int     blk   = valid ? block_of(bt, e) : 0;
  element_fp64<kMode, kRefTable>(lds, e, valid, bt.K[blk], bt.G[blk], E_total, conn, X, u, rjac, Fip, sip, slots,
                                 status);
}

// Ensemble: workgroup g takes chunk chunks[g] (at most kEPW elements of one
// block of one sample); the sample's slots are [3][8][E_mesh] at slot_base,
// its mesh's reference Jacobians [E_mesh][8][10] at 80 elem_base.
__global__ __launch_bounds__(kThreads, 5) void k_ens_element_fp64(EnsNodes en, const EnsRange* __restrict__ chunks,
                                                               const int* __restrict__ conn,
                                                               const double* __restrict__ rjac,
                                                               const double* __restrict__ Kd,
                                                               const double* __restrict__ Gd,
                                                               double* __restrict__ slots, int* __restrict__ status)
{
  __shared__ Lds  lds;
  const EnsRange  ch = ens_load(chunks, (int64_t)xcd_tile(blockIdx.x, gridDim.x));
  const EnsSample sm = ens_load(en.samples, (int64_t)ch.s);
  const EnsMesh   ms = ens_load(en.meshes, (int64_t)sm.mesh);
  const int       el = threadIdx.x >> 3;
  const int       kb = sm.kg_base + ch.block;
  element_fp64<kElementFull, true>(lds, (int64_t)ch.first + el, el < ch.count, ens_load(Kd, (int64_t)kb),
                             ens_load(Gd, (int64_t)kb), ms.num_elem, conn + 8 * (int64_t)ms.elem_base,
                             en.X + 3 * (int64_t)sm.x_base, en.u + 3 * (int64_t)sm.node_base,
                             rjac + 80 * (int64_t)ms.elem_base, nullptr, nullptr, slots + sm.slot_base, status);
}

// Reference inverse Jacobian and its determinant at every point of elements
// [0, E) (Invert3x3 of b = sum_j X_j (x) dN_j/dxi, src/nimble_element.h:442-500):
// out[(8 e + q) 10 + i] = b^{-1}[i] (row-major) for i < 9, det b for i = 9.
// The same operations as the element kernels' would be, so the same bits.
__global__ __launch_bounds__(kThreads) void k_reference_jacobians(int64_t E, const int* __restrict__ conn,
                                                                  const double* __restrict__ X,
                                                                  double* __restrict__ out)
{
  __shared__ double xs[kEPW][8][3];
  const int     el = threadIdx.x >> 3, k = threadIdx.x & 7;
  const int64_t e  = (int64_t)blockIdx.x * kEPW + el;
  if (e < E) {
    const int64_t nd = conn[e * 8 + k];
    xs[el][k][0]     = X[3 * nd];
    xs[el][k][1]     = X[3 * nd + 1];
    xs[el][k][2]     = X[3 * nd + 2];
  }
  __syncthreads();
  if (e < E) {
    double b[9], bi[9];
    reference_jacobian(k, xs[el], b);
    const double det = invert3x3(b, bi);
#pragma unroll
    for (int i = 0; i < 9; i++) out[(e * 8 + k) * 10 + i] = bi[i];
    out[(e * 8 + k) * 10 + 9] = det;
  }
}

// ---------------------------------------------------------------------------
// Seams on gathered arrays
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void k_seam_defgrad(int64_t E, const double* __restrict__ Xg,
                                                           const double* __restrict__ ug, double* __restrict__ F,
                                                           int* __restrict__ status)
{
  __shared__ double xs[kEPW][8][3], us[kEPW][8][3];
  const int     el = threadIdx.x >> 3, k = threadIdx.x & 7;
  const int64_t e  = (int64_t)blockIdx.x * kEPW + el;
  if (e < E)
    for (int c = 0; c < 3; c++) {
      xs[el][k][c] = Xg[(e * 8 + k) * 3 + c];
      us[el][k][c] = ug[(e * 8 + k) * 3 + c];
    }
  __syncthreads();
  if (e < E) {
    double a[9], b[9], bi[9], Fq[9];
    jacobians(k, xs[el], us[el], a, b);
    if (invert3x3<false>(b, bi) <= 0.0) atomicOr(status, 1);
    def_grad(a, bi, Fq);
    for (int i = 0; i < 9; i++) F[(e * 8 + k) * 9 + i] = Fq[i];
  }
}

__global__ __launch_bounds__(kThreads) void k_seam_forces(int64_t E, const double* __restrict__ Xg,
                                                          const double* __restrict__ ug,
                                                          const double* __restrict__ sig, double* __restrict__ f,
                                                          int* __restrict__ status)
{
  __shared__ Lds lds;
  const int     el = threadIdx.x >> 3, k = threadIdx.x & 7;
  const int64_t e  = (int64_t)blockIdx.x * kEPW + el;
  if (e < E)
    for (int c = 0; c < 3; c++) {
      lds.nodes.x[el][k][c]  = Xg[(e * 8 + k) * 3 + c];
      lds.nodes.xc[el][k][c] = Xg[(e * 8 + k) * 3 + c] + ug[(e * 8 + k) * 3 + c];
    }
  __syncthreads();
  double ainv[10];
  if (e < E) {
    double a[9], b[9];
    jacobians_x(k, lds.nodes.x[el], lds.nodes.xc[el], a, b);
    ainv[9] = invert3x3(a, ainv);
    if (ainv[9] <= 0.0) atomicOr(status, 1);
  }
  __syncthreads();
  if (e < E) {
    for (int i = 0; i < 10; i++) lds.pts.ai[el][k][i] = ainv[i];
    for (int i = 0; i < 6; i++) lds.pts.sg[el][k][i] = sig[(e * 8 + k) * 6 + i];
  }
  __syncthreads();
  if (e < E) {
    double fn[3];
    node_force(k, lds.pts.ai[el], lds.pts.sg[el], fn);
    for (int c = 0; c < 3; c++) f[(e * 8 + k) * 3 + c] = fn[c];
  }
}

__global__ __launch_bounds__(256) void k_seam_stress(int64_t P, double K, double G, const double* __restrict__ F,
                                                     double* __restrict__ sig)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= P) return;
  double Fp[9], s[6];
  for (int i = 0; i < 9; i++) Fp[i] = F[p * 9 + i];
  neohookean_stress(K, G, Fp, s);
  for (int i = 0; i < 6; i++) sig[p * 6 + i] = s[i];
}

// The same per point over num_elem x num_ipt points at arbitrary strides
// (point p = (e - first) * num_ipt + q), so a block's STEP_NP1 views are used
// where they lie: Kokkos LayoutLeft [E][8][9], LayoutRight, or a model's
// FP64 arrays (elements in groups of 8).
__global__ __launch_bounds__(256) void k_seam_stress_strided(int64_t first, int64_t num_elem, int num_ipt, double K,
                                                             double G, const double* __restrict__ F, IptStrides fs,
                                                             double* __restrict__ sig, IptStrides ss)
{
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= num_elem * num_ipt) return;
  const int64_t e = first + p / num_ipt, q = p % num_ipt;
  double        Fp[9], s[6];
  const double* Fe = F + (e >> 3) * fs.group + (e & 7) * fs.elem + q * fs.ipt;
  for (int i = 0; i < 9; i++) Fp[i] = Fe[i * fs.comp];
  neohookean_stress(K, G, Fp, s);
  double* Se = sig + (e >> 3) * ss.group + (e & 7) * ss.elem + q * ss.ipt;
  for (int i = 0; i < 6; i++) Se[i * ss.comp] = s[i];
}

// Lumped mass: lane q forms det J_q from reference coordinates, lane i then
// sums row i of the consistent mass (src/nimble_element.h:266-309 and
// src/nimble_element.cc:189-192).
__device__ __forceinline__ void lumped_row(int i, double rho, const double* detJ, double* out)
{
  double lm = 0.0;
  for (int j = 0; j < 8; j++) {
    double cm = 0.0;
    for (int q = 0; q < 8; q++) cm += 1.0 * rho * c_hex.N[q * 8 + i] * c_hex.N[q * 8 + j] * detJ[q];
    lm += cm;
  }
  *out = lm;
}

__global__ __launch_bounds__(kThreads) void k_seam_lumped_mass(int64_t E, double rho, const double* __restrict__ Xg,
                                                               double* __restrict__ m, int* __restrict__ status)
{
  __shared__ double xs[kEPW][8][3], zs[kEPW][8][3], det[kEPW][8];
  const int     el = threadIdx.x >> 3, k = threadIdx.x & 7;
  const int64_t e  = (int64_t)blockIdx.x * kEPW + el;
  if (e < E)
    for (int c = 0; c < 3; c++) {
      xs[el][k][c] = Xg[(e * 8 + k) * 3 + c];
      zs[el][k][c] = 0.0;
    }
  __syncthreads();
  if (e < E) {
    double a[9], b[9], bi[9];
    jacobians(k, xs[el], zs[el], a, b);
    det[el][k] = invert3x3(b, bi);
    if (det[el][k] <= 0.0) atomicOr(status, 1);
  }
  __syncthreads();
  if (e < E) lumped_row(k, rho, det[el], &m[e * 8 + k]);
}

__device__ __forceinline__ double char_length(const double* x)
{
  // src/nimble_element.cc:221-260 (x_max starts at 0.0 as in the reference)
  double x_min, x_max, y_min, y_max, z_min, z_max, min_d2;
  x_max = y_max = z_max = 0.0;
  min_d2 = x_min = y_min = z_min = 1.7976931348623157e308;
  for (int n = 0; n < 8; n++) {
    const double nx = x[3 * n], ny = x[3 * n + 1], nz = x[3 * n + 2];
    if (nx < x_min) x_min = nx;
    if (nx > x_max) x_max = nx;
    if (ny < y_min) y_min = ny;
    if (ny > y_max) y_max = ny;
    if (nz < z_min) z_min = nz;
    if (nz > z_max) z_max = nz;
    for (int m = n + 1; m < 8; m++) {
      const double mx = x[3 * m], my = x[3 * m + 1], mz = x[3 * m + 2];
      const double d2 = (nx - mx) * (nx - mx) + (ny - my) * (ny - my) + (nz - mz) * (nz - mz);
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

__global__ __launch_bounds__(256) void k_seam_char_length(int64_t E, const double* __restrict__ xg,
                                                          double* __restrict__ out)
{
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= E) return;
  double x[24];
  for (int i = 0; i < 24; i++) x[i] = xg[e * 24 + i];
  out[e] = char_length(x);
}

// ---------------------------------------------------------------------------
// Model setup
// ---------------------------------------------------------------------------
// Element masses of elements [e_begin, e_end): element e's node j at
// fp64_slot(e, j) (grouped) or j E_total + e.
__global__ __launch_bounds__(kThreads) void k_element_mass(int64_t e_begin, int64_t e_end, int64_t E_total,
                                                           int grouped, const int* __restrict__ conn,
                                                           const double* __restrict__ X, double rho,
                                                           double* __restrict__ elem_mass, int* __restrict__ status)
{
  __shared__ double xs[kEPW][8][3], zs[kEPW][8][3], det[kEPW][8];
  const int     el = threadIdx.x >> 3, k = threadIdx.x & 7;
  const int64_t e  = e_begin + (int64_t)blockIdx.x * kEPW + el;
  if (e < e_end) {
    const int64_t nd = conn[e * 8 + k];
    for (int c = 0; c < 3; c++) {
      xs[el][k][c] = X[3 * nd + c];
      zs[el][k][c] = 0.0;
    }
  }
  __syncthreads();
  if (e < e_end) {
    double a[9], b[9], bi[9];
    jacobians(k, xs[el], zs[el], a, b);
    det[el][k] = invert3x3(b, bi);
    if (det[el][k] <= 0.0) atomicOr(status, 1);
  }
  __syncthreads();
  if (e < e_end) lumped_row(k, rho, det[el], &elem_mass[grouped ? fp64_slot(e, k) : (int64_t)k * E_total + e]);
}

// Ensemble: element masses of chunk chunks[g] with its sample's block
// density, element e's node j at slot_base / 3 + fp64_slot(e, j).
__global__ __launch_bounds__(kThreads) void k_ens_element_mass(EnsNodes en, const EnsRange* __restrict__ chunks,
                                                               const int* __restrict__ conn,
                                                               const double* __restrict__ rho,
                                                               double* __restrict__ elem_mass,
                                                               int* __restrict__ status)
{
  __shared__ double xs[kEPW][8][3], zs[kEPW][8][3], det[kEPW][8];
  const EnsRange    ch = ens_load(chunks, (int64_t)blockIdx.x);
  const EnsSample   sm = ens_load(en.samples, (int64_t)ch.s);
  const EnsMesh     ms = ens_load(en.meshes, (int64_t)sm.mesh);
  const int         el = threadIdx.x >> 3, k = threadIdx.x & 7;
  const int64_t     e  = (int64_t)ch.first + el;
  const bool        ok = el < ch.count;
  const double*     X  = en.X + 3 * (int64_t)sm.x_base;
  if (ok) {
    const int64_t nd = conn[(ms.elem_base + e) * 8 + k];
    for (int c = 0; c < 3; c++) {
      xs[el][k][c] = X[3 * nd + c];
      zs[el][k][c] = 0.0;
    }
  }
  __syncthreads();
  if (ok) {
    double a[9], b[9], bi[9];
    jacobians(k, xs[el], zs[el], a, b);
    det[el][k] = invert3x3(b, bi);
    if (det[el][k] <= 0.0) atomicOr(status, 1);
  }
  __syncthreads();
  if (ok)
    lumped_row(k, ens_load(rho, (int64_t)(sm.kg_base + ch.block)), det[el],
               &elem_mass[sm.slot_base / 3 + fp64_slot(e, k)]);
}

// Minimum over the workgroup's elements of charlen / sound_speed on current
// coordinates X + u (src/nimble_block_base.cc:52-84); wave-level shuffles
// then LDS across the 4 waves.  min is exact in any order.
__global__ __launch_bounds__(256) void k_element_dt(int64_t e_begin, int64_t e_end, const int* __restrict__ conn,
                                                    const double* __restrict__ X, const double* __restrict__ u,
                                                    double sound_speed, double* __restrict__ wg_min)
{
  __shared__ double wmin[4];
  const int64_t     e  = e_begin + (int64_t)blockIdx.x * 256 + threadIdx.x;
  double            dt = 1.7976931348623157e308;
  if (e < e_end) {
    double x[24];
    for (int n = 0; n < 8; n++) {
      const int64_t nd = conn[e * 8 + n];
      for (int c = 0; c < 3; c++) x[3 * n + c] = X[3 * nd + c] + u[3 * nd + c];
    }
    dt = char_length(x) / sound_speed;
  }
  for (int off = 32; off > 0; off >>= 1) dt = fmin(dt, __shfl_xor(dt, off, 64));
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = dt;
  __syncthreads();
  if (threadIdx.x == 0) wg_min[blockIdx.x] = fmin(fmin(wmin[0], wmin[1]), fmin(wmin[2], wmin[3]));
}

__global__ __launch_bounds__(256) void k_min_reduce(const double* __restrict__ in, int64_t n, double* __restrict__ out)
{
  __shared__ double wmin[4];
  double            m = 1.7976931348623157e308;
  for (int64_t i = threadIdx.x; i < n; i += 256) m = fmin(m, in[i]);
  for (int off = 32; off > 0; off >>= 1) m = fmin(m, __shfl_xor(m, off, 64));
  if ((threadIdx.x & 63) == 0) wmin[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) *out = fmin(fmin(wmin[0], wmin[1]), fmin(wmin[2], wmin[3]));
}

inline unsigned grid_of(int64_t n, int64_t per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

void launch_element_fp64(const BlockTable<double>& bt, int64_t E_total, const int* conn, const double* X,
                         const double* u, const double* rjac, double* Fip, double* sip, double* slots, int* status,
                         hipStream_t s, int mode, const LaunchTimer& timer)
{
  const int64_t n = bt.off[bt.count] - bt.off[0];
  if (n <= 0) return complete_timer(timer, s);
  const dim3 grid(grid_of(n, kEPW)), block(kThreads);
#define NSM_FP64_LAUNCH(MODE)                                                                                   \
  do {                                                                                                          \
    if (rjac)                                                                                                   \
      hipExtLaunchKernelGGL((k_element_fp64<MODE, true>), grid, block, 0, s, timer.start, timer.stop, 0, bt,    \
                            E_total, conn, X, u, rjac, Fip, sip, slots, status);                                \
    else                                                                                                        \
      hipExtLaunchKernelGGL((k_element_fp64<MODE, false>), grid, block, 0, s, timer.start, timer.stop, 0, bt,   \
                            E_total, conn, X, u, rjac, Fip, sip, slots, status);                                \
  } while (0)
  if (mode == kElementDefGrad)
    NSM_FP64_LAUNCH(kElementDefGrad);
  else if (mode == kElementForcesFromStress)
    NSM_FP64_LAUNCH(kElementForcesFromStress);
  else
    NSM_FP64_LAUNCH(kElementFull);
#undef NSM_FP64_LAUNCH
}

void launch_reference_jacobians(int64_t E, const int* conn, const double* X, double* out, hipStream_t s)
{
  if (E <= 0) return;
  hipLaunchKernelGGL(k_reference_jacobians, dim3(grid_of(E, kEPW)), dim3(kThreads), 0, s, E, conn, X, out);
}

void launch_ens_element_fp64(const EnsNodes& en, const EnsRange* chunks, int num_chunks, const int* conn,
                             const double* rjac, const double* K, const double* G, double* slots, int* status,
                             hipStream_t s)
{
  if (num_chunks <= 0) return;
  hipLaunchKernelGGL(k_ens_element_fp64, dim3((unsigned)num_chunks), dim3(kThreads), 0, s, en, chunks, conn, rjac, K,
                     G, slots, status);
}

void launch_ens_element_mass(const EnsNodes& en, const EnsRange* chunks, int num_chunks, const int* conn,
                             const double* rho, double* elem_mass, int* status, hipStream_t s)
{
  if (num_chunks <= 0) return;
  hipLaunchKernelGGL(k_ens_element_mass, dim3((unsigned)num_chunks), dim3(kThreads), 0, s, en, chunks, conn, rho,
                     elem_mass, status);
}

void launch_seam_defgrad(int64_t E, const double* Xg, const double* ug, double* F, int* status, hipStream_t s)
{
  if (E <= 0) return;
  hipLaunchKernelGGL(k_seam_defgrad, dim3(grid_of(E, kEPW)), dim3(kThreads), 0, s, E, Xg, ug, F, status);
}

void launch_seam_forces(int64_t E, const double* Xg, const double* ug, const double* sig, double* f, int* status,
                        hipStream_t s)
{
  if (E <= 0) return;
  hipLaunchKernelGGL(k_seam_forces, dim3(grid_of(E, kEPW)), dim3(kThreads), 0, s, E, Xg, ug, sig, f, status);
}

void launch_seam_stress_strided(int64_t first, int64_t num_elem, int num_ipt, double K, double G, const double* F,
                                IptStrides fs, double* sig, IptStrides ss, hipStream_t s)
{
  const int64_t P = num_elem * num_ipt;
  if (P <= 0) return;
  hipLaunchKernelGGL(k_seam_stress_strided, dim3(grid_of(P, 256)), dim3(256), 0, s, first, num_elem, num_ipt, K, G,
                     F, fs, sig, ss);
}

void launch_seam_stress(int64_t P, double K, double G, const double* F, double* sig, hipStream_t s)
{
  if (P <= 0) return;
  hipLaunchKernelGGL(k_seam_stress, dim3(grid_of(P, 256)), dim3(256), 0, s, P, K, G, F, sig);
}

void launch_seam_lumped_mass(int64_t E, double rho, const double* Xg, double* m, int* status, hipStream_t s)
{
  if (E <= 0) return;
  hipLaunchKernelGGL(k_seam_lumped_mass, dim3(grid_of(E, kEPW)), dim3(kThreads), 0, s, E, rho, Xg, m, status);
}

void launch_seam_char_length(int64_t E, const double* xg, double* out, hipStream_t s)
{
  if (E <= 0) return;
  hipLaunchKernelGGL(k_seam_char_length, dim3(grid_of(E, 256)), dim3(256), 0, s, E, xg, out);
}

void launch_element_mass(int64_t e_begin, int64_t e_end, int64_t E_total, bool grouped, const int* conn,
                         const double* X, double rho, double* elem_mass, int* status, hipStream_t s)
{
  if (e_end <= e_begin) return;
  hipLaunchKernelGGL(k_element_mass, dim3(grid_of(e_end - e_begin, kEPW)), dim3(kThreads), 0, s, e_begin, e_end,
                     E_total, (int)grouped, conn, X, rho, elem_mass, status);
}

int launch_element_dt(int64_t e_begin, int64_t e_end, const int* conn, const double* X, const double* u,
                      double sound_speed, double* wg_min, hipStream_t s)
{
  if (e_end <= e_begin) return 0;
  const unsigned g = grid_of(e_end - e_begin, 256);
  hipLaunchKernelGGL(k_element_dt, dim3(g), dim3(256), 0, s, e_begin, e_end, conn, X, u, sound_speed, wg_min);
  return (int)g;
}

void launch_min_reduce(const double* in, int64_t n, double* out, hipStream_t s)
{
  hipLaunchKernelGGL(k_min_reduce, dim3(1), dim3(256), 0, s, in, n, out);
}

}  // namespace nsm
