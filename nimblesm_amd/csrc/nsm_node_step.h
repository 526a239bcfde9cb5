// nsm_node_step.h -- the per-node part of a central-difference step, shared
// by the node kernel and the tiled mixed-precision element kernel.
//
// Arithmetic keeps the reference's separate multiply and add
// (src/nimble_view.h:181-195: data += (alpha * 1.0) * rhs;
// src/integrators/explicit_time_integrator.cc:197-270), whatever the
// contraction setting of the including file.
#pragma once

#include "nsm_kernels.h"

namespace nsm {

struct StepArgs {
  StepTimes cur, nxt;
  int       output_bc;  // apply kinematic conditions after the second half kick
  int       fuse;       // also perform the next step's first half
  int       write_a;    // store the acceleration
  int       update_v;   // 0: only store the internal force
  int       write_f;    // store the internal force (always when update_v is 0)
};

// ApplyKinematicBC for the explicit scheme
// (src/nimble_boundary_condition_manager.h:136-217), resolved per node.
__device__ __forceinline__ void apply_bc(const BcEntry* table, int code, double v[3], const double u[3],
                                         double delta_t)
{
#pragma clang fp contract(off)
  if (code == 0) return;
  const BcEntry& b = table[code - 1];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (b.kind[c] == kBcVelocity)
      v[c] = b.value[c];
    else if (b.kind[c] >= kBcDisplacement && delta_t > 0.0)
      v[c] = (b.value[c] - u[c]) / delta_t;
    else if (b.kind[c] == kBcDisplacementOverVelocity)
      v[c] = b.fallback[c];
  }
}

// kick: v += (h * 1.0) * a  (also the drift u += (dt * 1.0) * v)
__device__ __forceinline__ void kick(double v[3], const double a[3], double h)
{
#pragma clang fp contract(off)
  const double prod = h * 1.0;
  v[0] += prod * a[0];
  v[1] += prod * a[1];
  v[2] += prod * a[2];
}

__device__ __forceinline__ void load3(const double* p, int64_t i, double r[3])
{
  r[0] = p[3 * i];
  r[1] = p[3 * i + 1];
  r[2] = p[3 * i + 2];
}

__device__ __forceinline__ void store3(double* p, int64_t i, const double r[3])
{
  p[3 * i]     = r[0];
  p[3 * i + 1] = r[1];
  p[3 * i + 2] = r[2];
}

// store of a value the step does not read back (the internal force):
// non-temporal, so it does not displace node data other launches reuse
__device__ __forceinline__ void store3_stream(double* p, int64_t i, const double r[3])
{
  __builtin_nontemporal_store(r[0], p + 3 * i);
  __builtin_nontemporal_store(r[1], p + 3 * i + 1);
  __builtin_nontemporal_store(r[2], p + 3 * i + 2);
}

// Node i with assembled internal force f, current displacement u, velocity v
// and lumped mass m: store f, then a = (1/m)(f + f_ext) with f_ext = 0, second
// half kick, BCs on output steps, and (fused) the next step's half kick, BC,
// drift, BC.  Performs stores only (table may live in LDS).
__device__ __forceinline__ void node_update(const NodeArrays& na, const BcEntry* table, int64_t i, int flags,
                                            const double f[3], double u[3], double v[3], double m,
                                            const StepArgs& sa)
{
#pragma clang fp contract(off)
  if (sa.write_f || !sa.update_v) store3_stream(na.f, i, f);
  if (!sa.update_v) return;
  const int    code     = flags & 0x7f;
  const double oneOverM = 1.0 / m;
  double       a[3];
  a[0] = oneOverM * (f[0] + 0.0);
  a[1] = oneOverM * (f[1] + 0.0);
  a[2] = oneOverM * (f[2] + 0.0);
  kick(v, a, sa.cur.half_dt);
  if (sa.output_bc) apply_bc(table, code, v, u, sa.cur.t_cur - sa.cur.t_prev);
  if (sa.fuse) {
    const double delta_t = sa.nxt.t_cur - sa.nxt.t_prev;
    kick(v, a, sa.nxt.half_dt);
    apply_bc(table, code, v, u, delta_t);
    kick(u, v, sa.nxt.dt);
    apply_bc(table, code, v, u, delta_t);
    store3(na.u, i, u);
  }
  store3(na.v, i, v);
  if (sa.write_a) store3(na.a, i, a);
}

// Reaction total -sum_k partials[tiles[k]] of one step, summed in 256
// interleaved lanes (lane v takes k = v, v + 256, ...) and then a fixed tree,
// so the result is the same for any workgroup size kThreads dividing 256 or
// divisible by it (threads past 256 then only take part in the barriers).
// red: 256 doubles of LDS.  Valid in thread 0.
template <int kThreads>
__device__ __forceinline__ double reaction_total(const double* partials, const int* tiles, int n, double* red)
{
  constexpr int kLanes = 256;
  static_assert(kLanes % kThreads == 0 || kThreads % kLanes == 0, "workgroup size and 256 must divide");
  for (int v = threadIdx.x; v < kLanes; v += kThreads) {
    double s = 0.0;
    for (int k = v; k < n; k += kLanes) s += partials[tiles[k]];
    red[v] = s;
  }
  __syncthreads();
  for (int w = kLanes / 2; w > 0; w >>= 1) {
    for (int v = threadIdx.x; v < w; v += kThreads) red[v] += red[v + w];
    __syncthreads();
  }
  return -red[0];
}

// reaction_total computed by one 64-lane wave, with the same additions in the
// same order: lane l holds the 256-lane sums of lanes l, l + 64, l + 128,
// l + 192, the first two tree levels (w = 128, 64) are added in registers
// and the rest (w = 32 .. 1, red[v] += red[v + w]) by shuffles down.  Valid
// in lane 0; no barrier, so any wave can fold any sample.
__device__ __forceinline__ double wave_reaction_total(const double* partials, const int* idx, int n)
{
  const int lane = (int)(threadIdx.x & 63);
  double    r[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    double s = 0.0;
    for (int k = lane + 64 * j; k < n; k += 256) s += partials[idx[k]];
    r[j] = s;
  }
  double v = (r[0] + r[2]) + (r[1] + r[3]);
#pragma unroll
  for (int w = 32; w > 0; w >>= 1) v += __shfl_down(v, w, 64);
  return -v;
}

// node_update with v and m loaded from the node arrays.
__device__ __forceinline__ void node_step(const NodeArrays& na, int64_t i, int flags, const double f[3],
                                          double u[3], const StepArgs& sa)
{
  if (!sa.update_v) {
    store3_stream(na.f, i, f);
    return;
  }
  double v[3];
  load3(na.v, i, v);
  node_update(na, na.bc_table, i, flags, f, u, v, na.mass[i], sa);
}

}  // namespace nsm
