// nsm_mixed.hip -- mixed-precision fused HEX8 element kernel.
//
// One thread per element (the per-element arithmetic is in nsm_mixed_math.h):
// gather 8 nodes of X and u (FP64), form the Walsh modes, evaluate F, the
// Neo-Hookean stress and adj(a) sigma at the 8 Gauss points in FP32, and
// write F and sigma (element-fastest "LayoutLeft" [comp][ipt][E], the layout
// Kokkos gives the reference's integration-point views on GPUs,
// src/nimble_kokkos_defs.h:86) plus the element's 24 force slots
// ([comp][node][E]).  Every store of a wavefront is one contiguous 256-byte
// segment.  Workgroups are mapped so that each XCD processes one contiguous
// range of elements, keeping shared nodes in that XCD's L2.
#include "nsm_kernels.h"
#include "nsm_mixed_math.h"

namespace nsm {
namespace {

constexpr int kThreads = 256;

// Optional occupancy floor (waves per SIMD) for tuning builds; the default
// build lets the compiler use the registers it needs (2 waves per SIMD).
#ifdef NSM_MIXED_MIN_WAVES
#define NSM_MIXED_WAVES_ATTR __attribute__((amdgpu_waves_per_eu(NSM_MIXED_MIN_WAVES, 8)))
#else
#define NSM_MIXED_WAVES_ATTR
#endif

__global__ __launch_bounds__(kThreads) NSM_MIXED_WAVES_ATTR void k_element_mixed(
    const BlockTable<float> bt, int64_t E_total, const int* __restrict__ conn, const double* __restrict__ X,
    const double* __restrict__ u, float* __restrict__ Fip, float* __restrict__ sip, float* __restrict__ slots,
    int* __restrict__ status)
{
  const int64_t e = bt.off[0] + (int64_t)xcd_tile(blockIdx.x, gridDim.x) * kThreads + threadIdx.x;
  if (e >= bt.off[bt.count]) return;
  const int   b = block_of(bt, e);
  const float K = bt.K[b], G = bt.G[b];

  const int4 c0 = *reinterpret_cast<const int4*>(conn + 8 * e);
  const int4 c1 = *reinterpret_cast<const int4*>(conn + 8 * e + 4);
  const int  nd[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
  double     Xe[8][3], ue[8][3];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int64_t n = nd[j];
    Xe[j][0]        = X[3 * n];
    Xe[j][1]        = X[3 * n + 1];
    Xe[j][2]        = X[3 * n + 2];
    ue[j][0]        = u[3 * n];
    ue[j][1]        = u[3 * n + 1];
    ue[j][2]        = u[3 * n + 2];
  }
  // F and sigma are not read again by the step: non-temporal stores (none
  // when Fip is null)
  auto store_point = [&](int q, const float* Fq, const float* Sq) {
    if (!Fip) return;
#pragma unroll
    for (int c = 0; c < 9; c++) __builtin_nontemporal_store(Fq[c], Fip + (int64_t)(c * 8 + q) * E_total + e);
#pragma unroll
    for (int c = 0; c < 6; c++) __builtin_nontemporal_store(Sq[c], sip + (int64_t)(c * 8 + q) * E_total + e);
  };
  float      f[8][3];
  auto store = [&](int qm, int qp, const float* Fm, const float* Sm, const float* Fp, const float* Sp) {
    store_point(qm, Fm, Sm);
    store_point(qp, Fp, Sp);
  };
  const bool ok = mixed::element(Xe, ue, K, G, store, f);
#pragma unroll
  for (int n = 0; n < 8; n++)
#pragma unroll
    for (int c = 0; c < 3; c++) slots[(int64_t)(c * 8 + n) * E_total + e] = f[n][c];
  if (!ok) atomicOr(status, 1);
}

}  // namespace

void launch_element_mixed(const BlockTable<float>& bt, int64_t E_total, const int* conn, const double* X,
                          const double* u, float* Fip, float* sip, float* slots, int* status, hipStream_t s,
                          const LaunchTimer& timer)
{
  const int64_t n = bt.off[bt.count] - bt.off[0];
  if (n <= 0) return complete_timer(timer, s);
  const unsigned grid = (unsigned)((n + kThreads - 1) / kThreads);
  hipExtLaunchKernelGGL(k_element_mixed, dim3(grid), dim3(kThreads), 0, s, timer.start, timer.stop, 0, bt, E_total,
                        conn, X, u, Fip, sip, slots, status);
}

}  // namespace nsm
