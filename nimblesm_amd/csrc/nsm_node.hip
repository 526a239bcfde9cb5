// nsm_node.hip -- node-side kernels: deterministic CSR gather of the element
// force slots fused with the central-difference update and the reaction-force
// sum, kinematic conditions, and the lumped-mass gather.
//
// Compiled with -ffp-contract=off: the velocity / displacement updates keep
// the reference's separate multiply and add
// (src/nimble_view.h:181-195: data += (alpha * 1.0) * rhs) and the gather
// adds element contributions in ascending element order starting from 0.0,
// which is the serial scatter order of src/nimble_block.cc:374-383 and
// src/nimble_model_data.cc:636-659.
#include "nsm_kernels.h"

namespace nsm {
namespace {

constexpr int kNodeThreads = 256;

// ApplyKinematicBC for the explicit scheme
// (src/nimble_boundary_condition_manager.h:136-217), resolved per node.
__device__ __forceinline__ void apply_bc(const NodeArrays& na, int code, double v[3], const double u[3],
                                         double delta_t)
{
  if (code == 0) return;
  const BcEntry& b = na.bc_table[code - 1];
#pragma unroll
  for (int c = 0; c < 3; c++) {
    if (b.kind[c] == 1)
      v[c] = b.value[c];
    else if (b.kind[c] == 2 && delta_t > 0.0)
      v[c] = (b.value[c] - u[c]) / delta_t;
  }
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

// kick: v += (h * 1.0) * a  (also the drift u += (dt * 1.0) * v)
__device__ __forceinline__ void kick(double v[3], const double a[3], double h)
{
  const double prod = h * 1.0;
  v[0] += prod * a[0];
  v[1] += prod * a[1];
  v[2] += prod * a[2];
}

__global__ __launch_bounds__(kNodeThreads) void k_apply_bc(NodeArrays na, StepTimes t)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na.num_nodes) return;
  const int code = na.flags[i] & 0x7f;
  if (code == 0) return;
  double u[3], v[3];
  load3(na.u, i, u);
  load3(na.v, i, v);
  apply_bc(na, code, v, u, t.t_cur - t.t_prev);
  store3(na.v, i, v);
}

// First half of a step (src/integrators/explicit_time_integrator.cc:197-216)
__global__ __launch_bounds__(kNodeThreads) void k_prekick(NodeArrays na, StepTimes t)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na.num_nodes) return;
  const int code = na.flags[i] & 0x7f;
  double    u[3], v[3], a[3];
  load3(na.u, i, u);
  load3(na.v, i, v);
  load3(na.a, i, a);
  const double delta_t = t.t_cur - t.t_prev;
  kick(v, a, t.half_dt);
  apply_bc(na, code, v, u, delta_t);
  kick(u, v, t.dt);
  apply_bc(na, code, v, u, delta_t);
  store3(na.u, i, u);
  store3(na.v, i, v);
}

// Internal-force gather and the second half of step n
// (src/integrators/explicit_time_integrator.cc:250-270), optionally fused
// with the first half of step n+1, plus per-tile partial reaction sums over
// flagged nodes.
template <typename SlotT, bool kFuse>
__global__ __launch_bounds__(kNodeThreads) void k_node_update(NodeArrays na, const SlotT* __restrict__ slots,
                                                              StepTimes cur, int output_bc, StepTimes nxt,
                                                              int write_a, int update_v, ReactionArgs rx)
{
  const unsigned tile    = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t  i       = (int64_t)tile * kNodeThreads + threadIdx.x;
  double         contrib = 0.0;
  int            member  = 0;  // node belongs to the reaction set
  if (i < na.num_nodes) {
    // gather: ascending element order, starting from 0.0; incidences are
    // fetched 8 at a time so their loads overlap, then added in order
    const uint32_t p0 = na.node_ptr[i], p1 = na.node_ptr[i + 1];
    double         f[3] = {0.0, 0.0, 0.0};
    for (uint32_t p = p0; p < p1; p += 8) {
      const uint32_t n = p1 - p < 8u ? p1 - p : 8u;
      uint32_t       ent[8];
      SlotT          s0[8], s1[8], s2[8];
#pragma unroll
      for (uint32_t k = 0; k < 8; k++)
        if (k < n) ent[k] = na.node_ent[p + k];
#pragma unroll
      for (uint32_t k = 0; k < 8; k++)
        if (k < n) {
          s0[k] = slots[ent[k]];
          s1[k] = slots[na.slot_stride + ent[k]];
          s2[k] = slots[2 * na.slot_stride + ent[k]];
        }
#pragma unroll
      for (uint32_t k = 0; k < 8; k++)
        if (k < n) {
          f[0] += (double)s0[k];
          f[1] += (double)s1[k];
          f[2] += (double)s2[k];
        }
    }
    store3(na.f, i, f);
    const int flags = na.flags[i];
    if (rx.coord >= 0 && (flags & kReactionFlag)) {
      member  = 1;
      contrib = f[rx.coord];
    }
    if (update_v) {
      const int    code     = flags & 0x7f;
      const double oneOverM = 1.0 / na.mass[i];
      double       a[3], u[3], v[3];
      a[0] = oneOverM * (f[0] + 0.0);
      a[1] = oneOverM * (f[1] + 0.0);
      a[2] = oneOverM * (f[2] + 0.0);
      load3(na.u, i, u);
      load3(na.v, i, v);
      kick(v, a, cur.half_dt);
      if (output_bc) apply_bc(na, code, v, u, cur.t_cur - cur.t_prev);
      if (kFuse) {
        const double delta_t = nxt.t_cur - nxt.t_prev;
        kick(v, a, nxt.half_dt);
        apply_bc(na, code, v, u, delta_t);
        kick(u, v, nxt.dt);
        apply_bc(na, code, v, u, delta_t);
        store3(na.u, i, u);
      }
      store3(na.v, i, v);
      if (write_a) store3(na.a, i, a);
    }
  }
  if (rx.coord < 0) return;  // uniform across the grid

  // partial of this node tile (fixed tree) when it holds reaction nodes
  if (!__syncthreads_or(member)) return;
  __shared__ double red[kNodeThreads];
  red[threadIdx.x] = contrib;
  __syncthreads();
  for (int w = kNodeThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) rx.partials[tile] = red[0];
}

__global__ __launch_bounds__(kNodeThreads) void k_reaction_finalize(const double* __restrict__ partials,
                                                                    const int* __restrict__ tiles, int num_tiles,
                                                                    double* __restrict__ out)
{
  __shared__ double red[kNodeThreads];
  double            s = 0.0;
  for (int k = threadIdx.x; k < num_tiles; k += kNodeThreads) s += partials[tiles[k]];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = kNodeThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = -red[0];
}

__global__ __launch_bounds__(kNodeThreads) void k_mass_gather(int64_t N, const uint32_t* __restrict__ node_ptr,
                                                              const uint32_t* __restrict__ node_ent,
                                                              const double* __restrict__ elem_mass,
                                                              double* __restrict__ m)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double s = 0.0;
  for (uint32_t p = node_ptr[i]; p < node_ptr[i + 1]; ++p) s += elem_mass[node_ent[p]];
  m[i] = s;
}

}  // namespace

unsigned node_grid(int64_t num_nodes) { return (unsigned)((num_nodes + kNodeThreads - 1) / kNodeThreads); }

void launch_apply_bc(const NodeArrays& na, StepTimes t, hipStream_t s)
{
  if (na.num_nodes <= 0) return;
  hipLaunchKernelGGL(k_apply_bc, dim3(node_grid(na.num_nodes)), dim3(kNodeThreads), 0, s, na, t);
}

void launch_prekick(const NodeArrays& na, StepTimes t, hipStream_t s)
{
  if (na.num_nodes <= 0) return;
  hipLaunchKernelGGL(k_prekick, dim3(node_grid(na.num_nodes)), dim3(kNodeThreads), 0, s, na, t);
}

template <typename SlotT>
void launch_node_update(const NodeArrays& na, const SlotT* slots, StepTimes cur, bool output_bc, bool fuse_next,
                        StepTimes next, bool write_a, bool update_v, ReactionArgs rx, hipStream_t s)
{
  if (na.num_nodes <= 0) return;
  if (fuse_next)
    hipLaunchKernelGGL((k_node_update<SlotT, true>), dim3(node_grid(na.num_nodes)), dim3(kNodeThreads), 0, s, na,
                       slots, cur, (int)output_bc, next, (int)write_a, (int)update_v, rx);
  else
    hipLaunchKernelGGL((k_node_update<SlotT, false>), dim3(node_grid(na.num_nodes)), dim3(kNodeThreads), 0, s, na,
                       slots, cur, (int)output_bc, next, (int)write_a, (int)update_v, rx);
}

template void launch_node_update<double>(const NodeArrays&, const double*, StepTimes, bool, bool, StepTimes, bool,
                                         bool, ReactionArgs, hipStream_t);
template void launch_node_update<float>(const NodeArrays&, const float*, StepTimes, bool, bool, StepTimes, bool, bool,
                                        ReactionArgs, hipStream_t);

void launch_reaction_finalize(const double* partials, const int* tiles, int num_tiles, double* out, hipStream_t s)
{
  hipLaunchKernelGGL(k_reaction_finalize, dim3(1), dim3(kNodeThreads), 0, s, partials, tiles, num_tiles, out);
}

void launch_mass_gather(int64_t N, const uint32_t* node_ptr, const uint32_t* node_ent, const double* elem_mass,
                        double* m, hipStream_t s)
{
  if (N <= 0) return;
  hipLaunchKernelGGL(k_mass_gather, dim3(node_grid(N)), dim3(kNodeThreads), 0, s, N, node_ptr, node_ent, elem_mass,
                     m);
}

}  // namespace nsm
