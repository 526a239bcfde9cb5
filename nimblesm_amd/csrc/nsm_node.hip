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
#include "nsm_ensemble.h"
#include "nsm_node_step.h"

namespace nsm {
namespace {

constexpr int kNodeThreads = 256;

__global__ __launch_bounds__(kNodeThreads) void k_apply_bc(NodeArrays na, StepTimes t)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= na.num_nodes) return;
  const int code = na.flags[i] & 0x7f;
  if (code == 0) return;
  double u[3], v[3];
  load3(na.u, i, u);
  load3(na.v, i, v);
  apply_bc(na.bc_table, code, v, u, t.t_cur - t.t_prev);
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
  apply_bc(na.bc_table, code, v, u, delta_t);
  kick(u, v, t.dt);
  apply_bc(na.bc_table, code, v, u, delta_t);
  store3(na.u, i, u);
  store3(na.v, i, v);
}

// Internal force of a node from the element force slots of its incidences
// [ptr[0], ptr[1]): ascending element order, starting from 0.0; incidences
// are fetched 8 at a time so their loads overlap, then added in order.
template <typename SlotT>
__device__ __forceinline__ void gather_force(const uint32_t* ptr, const uint32_t* __restrict__ node_ent,
                                             const SlotT* __restrict__ slots, int64_t stride, double f[3])
{
  const uint32_t p0 = ptr[0], p1 = ptr[1];
  f[0] = f[1] = f[2] = 0.0;
  for (uint32_t p = p0; p < p1; p += 8) {
    const uint32_t n = p1 - p < 8u ? p1 - p : 8u;
    uint32_t       ent[8];
    SlotT          s0[8], s1[8], s2[8];
#pragma unroll
    for (uint32_t k = 0; k < 8; k++)
      if (k < n) ent[k] = node_ent[p + k];
#pragma unroll
    for (uint32_t k = 0; k < 8; k++)
      if (k < n) {
        s0[k] = slots[ent[k]];
        s1[k] = slots[stride + ent[k]];
        s2[k] = slots[2 * stride + ent[k]];
      }
#pragma unroll
    for (uint32_t k = 0; k < 8; k++)
      if (k < n) {
        f[0] += (double)s0[k];
        f[1] += (double)s1[k];
        f[2] += (double)s2[k];
      }
  }
}

// Partial reaction sum of a 256-node workgroup (fixed tree) when it holds
// reaction nodes.
__device__ __forceinline__ void node_tile_partial(double contrib, int member, double* out)
{
  if (!__syncthreads_or(member)) return;
  __shared__ double red[kNodeThreads];
  red[threadIdx.x] = contrib;
  __syncthreads();
  for (int w = kNodeThreads / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = red[0];
}

// Internal-force gather and the second half of step n
// (src/integrators/explicit_time_integrator.cc:250-270), optionally fused
// with the first half of step n+1, plus per-tile partial reaction sums over
// flagged nodes.  Workgroup 0 first writes the folded reaction total of a
// previous step.
template <typename SlotT>
__global__ __launch_bounds__(kNodeThreads) void k_node_update(NodeArrays na, const SlotT* __restrict__ slots,
                                                              StepArgs sa, ReactionArgs rx, ReactionFold fold)
{
  if (fold.n > 0 && blockIdx.x == 0) {  // uniform
    __shared__ double fold_red[256];
    const double      total = reaction_total<kNodeThreads>(fold.partials, fold.tiles, fold.n, fold_red);
    if (threadIdx.x == 0) *fold.out = total;
  }
  const unsigned tile    = xcd_tile(blockIdx.x, gridDim.x);
  const int64_t  i       = (int64_t)tile * kNodeThreads + threadIdx.x;
  double         contrib = 0.0;
  int            member  = 0;  // node belongs to the reaction set
  if (i < na.num_nodes) {
    double f[3];
    gather_force(na.node_ptr + i, na.node_ent, slots, na.slot_stride, f);
    const int flags = na.flags[i];
    if (rx.coord >= 0 && (flags & kReactionFlag)) {
      member  = 1;
      contrib = f[rx.coord];
    }
    double u[3];
    if (sa.update_v) load3(na.u, i, u);
    node_step(na, i, flags, f, u, sa);
  }
  if (rx.coord < 0) return;  // uniform across the grid

  node_tile_partial(contrib, member, rx.partials + tile);
}

__global__ __launch_bounds__(kNodeThreads) void k_reaction_finalize(const double* __restrict__ partials,
                                                                    const int* __restrict__ tiles, int num_tiles,
                                                                    double* __restrict__ out)
{
  __shared__ double red[256];
  const double      total = reaction_total<kNodeThreads>(partials, tiles, num_tiles, red);
  if (threadIdx.x == 0) *out = total;
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

// Reaction history sampled at num_samples equally spaced times: sample k at
// p = (k+1) n / num_samples steps, linear between integer steps, where step
// j >= 1 holds hist[j-1] (the total after step j-1) and step 0 is 0 (a model
// at rest).
__device__ __forceinline__ double sampled_history(const double* __restrict__ hist, int64_t n, int k, int num_samples)
{
  const double  p  = (double)(k + 1) * (double)n / (double)num_samples;
  int64_t       j  = (int64_t)p;  // p >= 0
  if (j > n) j = n;
  const double  w  = p - (double)j;
  const double  r0 = j == 0 ? 0.0 : hist[j - 1];
  const double  r1 = j < n ? hist[j] : r0;
  return (1.0 - w) * r0 + w * r1;
}

__global__ __launch_bounds__(kNodeThreads) void k_sample_history(const double* __restrict__ hist, int64_t n,
                                                                 int num_samples, double* __restrict__ out)
{
  const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (k >= num_samples) return;
  out[k] = sampled_history(hist, n, k, num_samples);
}

// One node per thread; its w values move as one row.
__global__ __launch_bounds__(kNodeThreads) void k_permute_nodes(const double* __restrict__ src,
                                                                const int32_t* __restrict__ map, int64_t N, int w,
                                                                int to_device, double* __restrict__ dst)
{
  const int64_t n = (int64_t)blockIdx.x * kNodeThreads + threadIdx.x;
  if (n >= N) return;
  const int64_t d = map[n];
  if (to_device)
    for (int c = 0; c < w; c++) dst[d * w + c] = src[n * w + c];
  else
    for (int c = 0; c < w; c++) dst[n * w + c] = src[d * w + c];
}

// One value per thread, written in output order.
__global__ __launch_bounds__(kNodeThreads) void k_ip_gather(const void* __restrict__ src, IpLayout L,
                                                            int64_t e_begin, int64_t count, double* __restrict__ dst)
{
  const int64_t v   = (int64_t)blockIdx.x * kNodeThreads + threadIdx.x;
  const int     per = 8 * L.nc;
  if (v >= count * per) return;
  const int64_t e = e_begin + v / per;
  const int     r = (int)(v % per), q = r / L.nc, c = r % L.nc;
  double        x;
  if (L.kind == 0) {
    x = static_cast<const double*>(src)[fp64_ip_index(e, L.nc, c, q)];
  } else if (L.kind == 1) {
    x = static_cast<const float*>(src)[(int64_t)(c * 8 + q) * L.E + e];
  } else {
    const int64_t d = L.dev_elem[e];
    int           lo = 0, hi = L.num_tiles;  // elem_off[lo] <= d < elem_off[hi]
    while (hi - lo > 1) {
      const int mid = (lo + hi) >> 1;
      if (L.elem_off[mid] <= d)
        lo = mid;
      else
        hi = mid;
    }
    const int64_t e0 = L.elem_off[lo], n = L.elem_off[lo + 1] - e0;
    x = static_cast<const float*>(src)[8 * L.nc * e0 + (c * 4 + q / 2) * 2 * n + 2 * (d - e0) + (q & 1)];
  }
  dst[v] = x;
}

// UQ closure (UqModel::ApplyClosure, src/uq/nimble_uq.cc:327-346) fused with
// the approximate samples' node update: per node, f_s = f_nominal * c0, then
// f_s += f_exact_k * c_{k+1} for k ascending, then node_step with the
// sample's u, v, a, f.  The first kCachedExact exact forces of the node stay
// in registers across the samples.
constexpr int kCachedExact = 8;

__global__ __launch_bounds__(kNodeThreads) void k_uq_closure(NodeArrays na, UqClosure uc, StepArgs sa)
{
  const int64_t i = (int64_t)xcd_tile(blockIdx.x, gridDim.x) * kNodeThreads + threadIdx.x;
  if (i >= na.num_nodes) return;
  const int ne = uc.num_exact;
  double    fn[3], fe[kCachedExact][3];
  load3(uc.f_nominal, i, fn);
#pragma unroll
  for (int k = 0; k < kCachedExact; k++)
    if (k < ne) load3(uc.state + (4 * (int64_t)k + 3) * uc.stride, i, fe[k]);
  const int flags = na.flags[i];
  for (int s = 0; s < uc.num_approx; s++) {
    const double* c  = uc.coeff + (int64_t)s * (ne + 1);
    double        f[3] = {fn[0] * c[0], fn[1] * c[0], fn[2] * c[0]};
#pragma unroll
    for (int k = 0; k < kCachedExact; k++)
      if (k < ne) {
        const double ck = c[k + 1];
        f[0] += fe[k][0] * ck;
        f[1] += fe[k][1] * ck;
        f[2] += fe[k][2] * ck;
      }
    for (int k = kCachedExact; k < ne; k++) {
      double g[3];
      load3(uc.state + (4 * (int64_t)k + 3) * uc.stride, i, g);
      const double ck = c[k + 1];
      f[0] += g[0] * ck;
      f[1] += g[1] * ck;
      f[2] += g[2] * ck;
    }
    NodeArrays ns = na;
    double*    b  = uc.state + 4 * (int64_t)(ne + s) * uc.stride;
    ns.u          = b;
    ns.v          = b + uc.stride;
    ns.a          = b + 2 * uc.stride;
    ns.f          = b + 3 * uc.stride;
    double u[3];
    if (sa.update_v) load3(ns.u, i, u);
    node_step(ns, i, flags, f, u, sa);
  }
}

// ---------------------------------------------------------------------------
// Ensemble node kernels: workgroup g takes chunks[g], at most kNodeThreads
// consecutive nodes (rows local to the sample's mesh) of one sample.
// ---------------------------------------------------------------------------
struct EnsNode {
  int64_t row, xrow;  // row of u, v, a, f, mass; row of X and flags
  bool    valid;
};

__device__ __forceinline__ EnsNode ens_node(const EnsNodes& en, const EnsRange& ch, const EnsSample& sm)
{
  const int     l = ch.first + (int)threadIdx.x;
  EnsNode       n;
  n.valid = (int)threadIdx.x < ch.count;
  n.row   = (int64_t)sm.node_base + l;
  n.xrow  = (int64_t)sm.x_base + l;
  return n;
}

__device__ __forceinline__ NodeArrays ens_arrays(const EnsNodes& en, const EnsSample& sm)
{
  NodeArrays na;
  na.num_nodes   = 0;
  na.slot_stride = 0;
  na.node_ptr    = nullptr;
  na.node_ent    = nullptr;
  na.mass        = en.mass;
  na.u           = en.u;
  na.v           = en.v;
  na.a           = en.a;
  na.f           = en.f;
  na.flags       = en.flags;
  na.bc_table    = en.bc + sm.bc_base;
  na.num_bc      = 0;
  return na;
}

// every sample at rest: u = v = a = f = 0, then its initial velocities
// (ModelDataBase::ApplyInitialConditions) and its kinematic conditions at
// the initial time (ApplyKinematicConditions(t0, t0): delta t = 0)
__global__ __launch_bounds__(kNodeThreads) void k_ens_reset(EnsNodes en, const EnsRange* __restrict__ chunks)
{
  const EnsRange  ch = ens_load(chunks, (int64_t)blockIdx.x);
  const EnsSample sm = ens_load(en.samples, (int64_t)ch.s);
  const EnsNode   n  = ens_node(en, ch, sm);
  if (!n.valid) return;
  double    u[3] = {0.0, 0.0, 0.0}, v[3] = {0.0, 0.0, 0.0};
  const int code = en.flags[n.xrow] & 0x7f;
  if (code) {
    const BcEntry& ini = en.init[sm.bc_base + code - 1];
#pragma unroll
    for (int c = 0; c < 3; c++)
      if (ini.kind[c]) v[c] = ini.value[c];
    apply_bc(en.bc + sm.bc_base, code, v, u, 0.0);
  }
  store3(en.u, n.row, u);
  store3(en.v, n.row, v);
  store3(en.a, n.row, u);
  store3(en.f, n.row, u);
}

// first half of step es.step (src/integrators/explicit_time_integrator.cc:197-216)
__global__ __launch_bounds__(kNodeThreads) void k_ens_prekick(EnsNodes en, const EnsRange* __restrict__ chunks,
                                                              EnsStep es)
{
  const EnsRange  ch = ens_load(chunks, (int64_t)blockIdx.x);
  const EnsSample sm = ens_load(en.samples, (int64_t)ch.s);
  const StepTimes t  = ens_step_args(es, ch.s).cur;
  const EnsNode   n  = ens_node(en, ch, sm);
  if (!n.valid) return;
  const int code = en.flags[n.xrow] & 0x7f;
  double    u[3], v[3], a[3];
  load3(en.u, n.row, u);
  load3(en.v, n.row, v);
  load3(en.a, n.row, a);
  const double delta_t = t.t_cur - t.t_prev;
  kick(v, a, t.half_dt);
  apply_bc(en.bc + sm.bc_base, code, v, u, delta_t);
  kick(u, v, t.dt);
  apply_bc(en.bc + sm.bc_base, code, v, u, delta_t);
  store3(en.u, n.row, u);
  store3(en.v, n.row, v);
}

// every sample's reaction total of fold.step, one wave per sample
__device__ __forceinline__ void ens_fold_waves(const EnsFold& f, int first_wave, int wave_stride)
{
  if (f.step < 0) return;
  for (int s = first_wave + (int)(threadIdx.x >> 6); s < f.num_samples; s += wave_stride) {
    const int    p0    = f.ptr[s];
    const double total = wave_reaction_total(f.partials, f.idx + p0, f.ptr[s + 1] - p0);
    if ((threadIdx.x & 63) == 0) f.hist[(int64_t)s * f.hist_stride + f.step] = total;
  }
}

__global__ __launch_bounds__(kNodeThreads) void k_ens_fold(EnsFold fold)
{
  constexpr int kWaves = kNodeThreads / 64;
  ens_fold_waves(fold, (int)blockIdx.x * kWaves, (int)gridDim.x * kWaves);
}

// FP64 step of one chunk of nodes: the gather of k_node_update over the
// sample's slots and its node update; a previous step's totals are folded
// by the first waves of the launch.
__global__ __launch_bounds__(kNodeThreads) void k_ens_node_fp64(EnsNodes en, const EnsRange* __restrict__ chunks,
                                                                const uint32_t* __restrict__ node_ptr,
                                                                const uint32_t* __restrict__ node_ent,
                                                                const double* __restrict__ slots, EnsStep es,
                                                                EnsReaction rx, EnsFold fold)
{
  constexpr int kWaves = kNodeThreads / 64;
  ens_fold_waves(fold, (int)blockIdx.x * kWaves, (int)gridDim.x * kWaves);
  const unsigned  g  = xcd_tile(blockIdx.x, gridDim.x);
  const EnsRange  ch = ens_load(chunks, (int64_t)g);
  const EnsSample sm = ens_load(en.samples, (int64_t)ch.s);
  const EnsMesh   ms = ens_load(en.meshes, (int64_t)sm.mesh);
  const StepArgs  sa = ens_step_args(es, ch.s);
  const EnsNode   n  = ens_node(en, ch, sm);
  double          contrib = 0.0;
  int             member  = 0;
  if (n.valid) {
    double f[3];
    gather_force(node_ptr + ms.ptr_base + ch.first + threadIdx.x, node_ent, slots + sm.slot_base,
                 8 * round_up8(ms.num_elem), f);
    const int flags = en.flags[n.xrow];
    if (rx.coord >= 0 && (flags & kReactionFlag)) {
      member  = 1;
      contrib = f[rx.coord];
    }
    const NodeArrays na = ens_arrays(en, sm);
    double           u[3];
    if (sa.update_v) load3(na.u, n.row, u);
    node_step(na, n.row, flags, f, u, sa);
  }
  if (rx.coord < 0) return;  // uniform across the grid
  node_tile_partial(contrib, member, rx.partials + g);
}

__global__ __launch_bounds__(kNodeThreads) void k_ens_mass_gather(EnsNodes en, const EnsRange* __restrict__ chunks,
                                                                  const uint32_t* __restrict__ node_ptr,
                                                                  const uint32_t* __restrict__ node_ent,
                                                                  const double* __restrict__ elem_mass,
                                                                  double* __restrict__ m)
{
  const EnsRange  ch = ens_load(chunks, (int64_t)blockIdx.x);
  const EnsSample sm = ens_load(en.samples, (int64_t)ch.s);
  const EnsMesh   ms = ens_load(en.meshes, (int64_t)sm.mesh);
  const EnsNode   n  = ens_node(en, ch, sm);
  if (!n.valid) return;
  const uint32_t* ptr = node_ptr + ms.ptr_base + ch.first + threadIdx.x;
  const double*   em  = elem_mass + sm.slot_base / 3;
  double          s   = 0.0;
  for (uint32_t p = ptr[0]; p < ptr[1]; ++p) s += em[node_ent[p]];
  m[n.row] = s;
}

__global__ __launch_bounds__(kNodeThreads) void k_ens_sample_history(const double* __restrict__ hist,
                                                                     int64_t hist_stride, int64_t n, int num_samples,
                                                                     int num_points, double* __restrict__ out)
{
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= (int64_t)num_samples * num_points) return;
  const int s = (int)(k / num_points), p = (int)(k % num_points);
  out[k] = sampled_history(hist + (int64_t)s * hist_stride, n, p, num_points);
}

}  // namespace

void launch_ens_reset(const EnsNodes& en, const EnsRange* chunks, int num_chunks, EnsStep, hipStream_t s)
{
  if (num_chunks <= 0) return;
  hipLaunchKernelGGL(k_ens_reset, dim3((unsigned)num_chunks), dim3(kNodeThreads), 0, s, en, chunks);
}

void launch_ens_prekick(const EnsNodes& en, const EnsRange* chunks, int num_chunks, EnsStep es, hipStream_t s)
{
  if (num_chunks <= 0) return;
  hipLaunchKernelGGL(k_ens_prekick, dim3((unsigned)num_chunks), dim3(kNodeThreads), 0, s, en, chunks, es);
}

void launch_ens_node_fp64(const EnsNodes& en, const EnsRange* chunks, int num_chunks, const uint32_t* node_ptr,
                          const uint32_t* node_ent, const double* slots, EnsStep es, EnsReaction rx, EnsFold fold,
                          hipStream_t s)
{
  if (num_chunks <= 0) return;
  hipLaunchKernelGGL(k_ens_node_fp64, dim3((unsigned)num_chunks), dim3(kNodeThreads), 0, s, en, chunks, node_ptr,
                     node_ent, slots, es, rx, fold);
}

void launch_ip_gather(const void* src, const IpLayout& layout, int64_t e_begin, int64_t count, double* dst,
                      hipStream_t s)
{
  const int64_t values = count * 8 * layout.nc;
  if (values <= 0) return;
  hipLaunchKernelGGL(k_ip_gather, dim3((unsigned)((values + kNodeThreads - 1) / kNodeThreads)), dim3(kNodeThreads), 0,
                     s, src, layout, e_begin, count, dst);
}

void launch_permute_nodes(const double* src, const int32_t* map, int64_t N, int w, bool to_device, double* dst,
                          hipStream_t s)
{
  if (N <= 0) return;
  hipLaunchKernelGGL(k_permute_nodes, dim3((unsigned)((N + kNodeThreads - 1) / kNodeThreads)), dim3(kNodeThreads), 0,
                     s, src, map, N, w, (int)to_device, dst);
}

void launch_ens_fold(EnsFold fold, hipStream_t s)
{
  if (fold.step < 0 || fold.num_samples <= 0) return;
  constexpr int kWaves = kNodeThreads / 64;
  hipLaunchKernelGGL(k_ens_fold, dim3((unsigned)((fold.num_samples + kWaves - 1) / kWaves)), dim3(kNodeThreads), 0,
                     s, fold);
}

void launch_ens_mass_gather(const EnsNodes& en, const EnsRange* chunks, int num_chunks, const uint32_t* node_ptr,
                            const uint32_t* node_ent, const double* elem_mass, double* mass, hipStream_t s)
{
  if (num_chunks <= 0) return;
  hipLaunchKernelGGL(k_ens_mass_gather, dim3((unsigned)num_chunks), dim3(kNodeThreads), 0, s, en, chunks, node_ptr,
                     node_ent, elem_mass, mass);
}

void launch_ens_sample_history(const double* hist, int64_t hist_stride, int64_t n, int num_samples, int num_points,
                               double* out, hipStream_t s)
{
  const int64_t total = (int64_t)num_samples * num_points;
  if (total <= 0) return;
  hipLaunchKernelGGL(k_ens_sample_history, dim3((unsigned)((total + kNodeThreads - 1) / kNodeThreads)),
                     dim3(kNodeThreads), 0, s, hist, hist_stride, n, num_samples, num_points, out);
}


void launch_sample_history(const double* hist, int64_t n, int num_samples, double* out, hipStream_t s)
{
  if (num_samples <= 0) return;
  hipLaunchKernelGGL(k_sample_history, dim3((unsigned)((num_samples + kNodeThreads - 1) / kNodeThreads)),
                     dim3(kNodeThreads), 0, s, hist, n, num_samples, out);
}

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
void launch_node_update(const NodeArrays& na, const SlotT* slots, const StepArgs& sa, ReactionArgs rx,
                        ReactionFold fold, hipStream_t s, const LaunchTimer& timer)
{
  if (na.num_nodes <= 0) return complete_timer(timer, s);
  hipExtLaunchKernelGGL((k_node_update<SlotT>), dim3(node_grid(na.num_nodes)), dim3(kNodeThreads), 0, s, timer.start,
                        timer.stop, 0, na, slots, sa, rx, fold);
}

template void launch_node_update<double>(const NodeArrays&, const double*, const StepArgs&, ReactionArgs,
                                         ReactionFold, hipStream_t, const LaunchTimer&);
template void launch_node_update<float>(const NodeArrays&, const float*, const StepArgs&, ReactionArgs, ReactionFold,
                                        hipStream_t, const LaunchTimer&);

void launch_uq_closure(const NodeArrays& na, const UqClosure& uc, const StepArgs& sa, hipStream_t s)
{
  if (na.num_nodes <= 0 || uc.num_approx <= 0) return;
  hipLaunchKernelGGL(k_uq_closure, dim3(node_grid(na.num_nodes)), dim3(kNodeThreads), 0, s, na, uc, sa);
}

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
