// nsm_kernels.h -- host-side launchers of the HIP kernels (internal).
//
// Device layouts:
//   node vectors          [N][3] double (the reference's AoS node layout)
//   integration points    FP64: groups of 8 elements (fp64_ip_index); FP32:
//                         [c][8][E] (component, Gauss point, element), or
//                         tile-blocked in tiled models (launch_element_tiled)
//   element force slots   [3][slots] (component, slot): slot fp64_slot(e, k)
//                         in FP64 models (slots = 8 E, E rounded up to 8),
//                         k*E + e in the untiled mixed path; tiled models
//                         keep none
//   node -> element CSR   node_ent[p] = slot of the p-th incidence (element
//                         e, local node k), ascending e per node
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nsm {

// Per-node resolved kinematic condition, the net effect of the condition
// list applied in order (src/nimble_boundary_condition_manager.h:136-217).
// kind: 0 none, 1 prescribed velocity, 2 prescribed displacement, 3
// prescribed displacement listed after a prescribed velocity: with
// delta_t <= 0 the displacement condition does nothing, so the velocity
// condition's value (fallback) is what remains.
constexpr int32_t kBcVelocity = 1, kBcDisplacement = 2, kBcDisplacementOverVelocity = 3;
struct BcEntry {
  int32_t kind[3];
  int32_t pad;
  double  value[3];
  double  fallback[3];
};

// node_flags: low 7 bits = bc_table index + 1 (0: free), bit 7 = node belongs
// to the reaction-force set
constexpr uint8_t kReactionFlag = 0x80;
constexpr int     kMaxBcEntries = 127;

// Events a timed launch records itself (hipExtLaunchKernelGGL): the start
// and end of the kernel's own execution, without the stream markers that an
// event recorded before and after the launch would add.  Null: not timed.
struct LaunchTimer {
  hipEvent_t start = nullptr, stop = nullptr;
};
// a timed launch that has no work still completes its events (zero time)
inline void complete_timer(const LaunchTimer& t, hipStream_t s)
{
  if (t.start) {
    (void)hipEventRecord(t.start, s);
    (void)hipEventRecord(t.stop, s);
  }
}

// Times of one explicit step, computed on the host exactly as
// src/integrators/explicit_time_integrator.cc:192-195 does.
struct StepTimes {
  double t_prev, t_cur, dt, half_dt;
};

struct NodeArrays {
  int64_t         num_nodes;
  int64_t         slot_stride;  // 8*E: distance between force components in the slots
  const uint32_t* node_ptr;     // [N+1] CSR offsets into node_ent
  const uint32_t* node_ent;
  const double*   mass;         // [N]
  double*         u;            // [N][3]
  double*         v;            // [N][3]
  double*         a;            // [N][3]
  double*         f;            // [N][3] internal force
  const uint8_t*  flags;        // [N]
  const BcEntry*  bc_table;
  int             num_bc;       // entries of bc_table
};

// Reaction-force sum: the node kernel writes one partial per node tile
// (256 nodes) that holds reaction nodes; launch_reaction_finalize then sums
// the partials of those tiles in ascending tile order.
struct ReactionArgs {
  int     coord;     // -1: no reaction sum
  double* partials;  // [node tiles]
};

// A previous step's reaction total, summed by workgroup 0 of the next launch
// instead of a separate finalize launch (n = 0: none).
struct ReactionFold {
  const double* partials;
  const int*    tiles;
  int           n;
  double*       out;
};

// Workgroup b of nb -> tile index so that XCD (b mod 8) processes one
// contiguous range of tiles (workgroups are dealt to the 8 XCDs round-robin).
__device__ __forceinline__ unsigned xcd_tile(unsigned b, unsigned nb)
{
  const unsigned x = b & 7u, k = b >> 3, q = nb >> 3, r = nb & 7u;
  return x * q + (x < r ? x : r) + k;
}

// Up to kMaxLaunchBlocks consecutive element blocks handled by one launch:
// block b owns elements [off[b], off[b+1]) with moduli K[b], G[b].
constexpr int kMaxLaunchBlocks = 8;
template <typename Real>
struct BlockTable {
  int     count;
  int64_t off[kMaxLaunchBlocks + 1];
  Real    K[kMaxLaunchBlocks];
  Real    G[kMaxLaunchBlocks];
};

// block of element e (blocks ascending, e inside the table's range)
template <typename Real>
__device__ __forceinline__ int block_of(const BlockTable<Real>& bt, int64_t e)
{
  int b = 0;
  while (b + 1 < bt.count && e >= bt.off[b + 1]) b++;
  return b;
}

// ---- FP64 element-data layout ----
// The FP64 element kernel's wavefront holds 8 consecutive elements, so its
// element data is stored in groups of 8 elements, one group contiguous:
// value (c, q) of element e (nc components, 8 points) at
//   64 nc (e / 8) + 64 c + 8 q + e % 8,
// and the force slot of element e's node j (the node gather's index; three
// components at slot_stride = 8 E8 apart, E8 = E rounded up to 8) at
//   64 (e / 8) + 8 j + e % 8.
// A store of one component then covers 512 contiguous bytes per wavefront.
__host__ __device__ __forceinline__ int64_t fp64_ip_index(int64_t e, int nc, int c, int q)
{
  return 64 * (int64_t)nc * (e >> 3) + 64 * c + 8 * q + (e & 7);
}
__host__ __device__ __forceinline__ int64_t fp64_slot(int64_t e, int j) { return 64 * (e >> 3) + 8 * j + (e & 7); }
__host__ __device__ __forceinline__ int64_t round_up8(int64_t n) { return (n + 7) & ~(int64_t)7; }

// ---- element kernels over the elements of a block table ----
// FP64 element kernel modes: F, stress and forces; F only; forces from the
// stress already in sip
constexpr int kElementFull = 0, kElementDefGrad = 1, kElementForcesFromStress = 2;
// rjac: the reference inverse Jacobians of launch_reference_jacobians
void launch_element_fp64(const BlockTable<double>& bt, int64_t E_total, const int* conn, const double* X,
                         const double* u, const double* rjac, double* Fip, double* sip, double* slots, int* status,
                         hipStream_t s, int mode = kElementFull, const LaunchTimer& timer = LaunchTimer{});
// b^{-1} (row-major) and det b of the reference Jacobian b at every point of
// elements [0, E): out[(8 e + q) 10 + i], i < 9 the inverse, i = 9 det b
// ([E][8][10] doubles).  A function of X alone, formed once per mesh.
void launch_reference_jacobians(int64_t E, const int* conn, const double* X, double* out, hipStream_t s);
void launch_element_mixed(const BlockTable<float>& bt, int64_t E_total, const int* conn, const double* X,
                          const double* u, float* Fip, float* sip, float* slots, int* status, hipStream_t s,
                          const LaunchTimer& timer = LaunchTimer{});

// ---- element / material seams on gathered arrays ----
void launch_seam_defgrad(int64_t E, const double* Xg, const double* ug, double* F, int* status, hipStream_t s);
void launch_seam_forces(int64_t E, const double* Xg, const double* ug, const double* sig, double* f, int* status,
                        hipStream_t s);
void launch_seam_stress(int64_t P, double K, double G, const double* F, double* sig, hipStream_t s);
// stress of the points of elements e in [first, first + num_elem): F value
// (e, q, c) at F[(e / 8) fg + (e % 8) fe + q fq + c fc], sigma value at the
// same form with the sigma strides (plain strides: g = 8 e, first = 0)
struct IptStrides {
  int64_t group, elem, ipt, comp;
};
void launch_seam_stress_strided(int64_t first, int64_t num_elem, int num_ipt, double K, double G, const double* F,
                                IptStrides fs, double* sig, IptStrides ss, hipStream_t s);
void launch_seam_lumped_mass(int64_t E, double rho, const double* Xg, double* m, int* status, hipStream_t s);
void launch_seam_char_length(int64_t E, const double* xg, double* out, hipStream_t s);

// ---- model setup ----
// element masses in the node gather's slot order: element e's node k at
// elem_mass[fp64_slot(e, k)] (grouped) or elem_mass[k*E_total + e]
void launch_element_mass(int64_t e_begin, int64_t e_end, int64_t E_total, bool grouped, const int* conn,
                         const double* X, double rho, double* elem_mass, int* status, hipStream_t s);
// per-workgroup minimum of charlen/c over [e_begin, e_end) into wg_min[blockIdx]
int  launch_element_dt(int64_t e_begin, int64_t e_end, const int* conn, const double* X, const double* u,
                       double sound_speed, double* wg_min, hipStream_t s);
void launch_min_reduce(const double* in, int64_t n, double* out, hipStream_t s);
void launch_mass_gather(int64_t N, const uint32_t* node_ptr, const uint32_t* node_ent, const double* elem_mass,
                        double* m, hipStream_t s);

// ---- node kernels ----
unsigned node_grid(int64_t num_nodes);
// Kinematic conditions on v: the explicit scheme's ApplyKinematicBC.
void launch_apply_bc(const NodeArrays& na, StepTimes t, hipStream_t s);
// first half kick + BC + drift + BC of one step
void launch_prekick(const NodeArrays& na, StepTimes t, hipStream_t s);
struct StepArgs;
// gather f from element slots, a = f/m, second half kick (+BC on output
// steps); if fused, also the next step's first half kick, BC, drift, BC;
// optional per-tile reaction partials; workgroup 0 first sums the fold.
template <typename SlotT>
void launch_node_update(const NodeArrays& na, const SlotT* slots, const StepArgs& sa, ReactionArgs rx,
                        ReactionFold fold, hipStream_t s, const LaunchTimer& timer = LaunchTimer{});
// Off-nominal ensemble: sample s owns u, v, a, f at state + (4 s + j) stride,
// j = 0..3 (exact samples first).  The closure kernel forms each approximate
// sample's force f_s = c[s][0] f_nominal + sum_k c[s][k+1] f_exact_k
// (src/uq/nimble_uq.cc:327-346, same operation order) and applies the node
// update of node_step to it; na carries the samples' mass, flags and
// kinematic-condition table.
struct UqClosure {
  int           num_exact, num_approx;
  int64_t       stride;     // 3 N
  const double* f_nominal;  // [N][3]
  double*       state;
  const double* coeff;      // [num_approx][num_exact + 1]
};
void launch_uq_closure(const NodeArrays& na, const UqClosure& uc, const StepArgs& sa, hipStream_t s);
// reaction history sampled at num_samples equally spaced times (device out)
void launch_sample_history(const double* hist, int64_t n, int num_samples, double* out, hipStream_t s);
// out = -sum_k partials[tiles[k]], k ascending (fixed tree, reproducible)
void launch_reaction_finalize(const double* partials, const int* tiles, int num_tiles, double* out, hipStream_t s);
// Integration-point values of mesh elements [e_begin, e_begin + count) in
// the reference's element-data order (dst[(e - e_begin) 8 + q][c], FP64)
// from a model's device layout:
//   kind 0: FP64 in groups of 8 elements (fp64_ip_index);
//   kind 1: FP32 [nc][8][E];
//   kind 2: FP32 tile-blocked, tile t holding device elements
//           [elem_off[t], elem_off[t + 1]); mesh element e at device
//           position dev_elem[e].
struct IpLayout {
  int            kind, nc;
  int64_t        E;
  const int32_t* dev_elem;
  const int32_t* elem_off;
  int            num_tiles;
};
void launch_ip_gather(const void* src, const IpLayout& layout, int64_t e_begin, int64_t count, double* dst,
                      hipStream_t s);
// node field of width w between mesh order and a model's device order:
// to_device: dst[map[n] w + c] = src[n w + c]; else dst[n w + c] = src[map[n] w + c]
// (map[n] = device index of mesh node n, a permutation of [0, N))
void launch_permute_nodes(const double* src, const int32_t* map, int64_t N, int w, bool to_device, double* dst,
                          hipStream_t s);

}  // namespace nsm
