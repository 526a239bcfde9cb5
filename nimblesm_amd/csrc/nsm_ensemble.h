// nsm_ensemble.h -- device tables and launchers of the batched ensemble
// (internal).
//
// An ensemble holds meshes and samples; sample s is a copy of mesh
// EnsSample::mesh with its own node state.  Node arrays of all samples are
// concatenated (sample s owns rows [node_base, node_base + N_mesh) of u, v,
// a, f and the lumped mass, in its mesh's device node order); X and the node
// flags are stored once per mesh (rows [x_base, x_base + N_mesh)).  Work is
// described by tables whose entries each name one sample, so every
// workgroup (or element tile) works on one sample and reads that sample's
// record by scalar loads.
#pragma once

#include "nsm_kernels.h"
#include "nsm_node_step.h"

namespace nsm {

// kinematic-condition codes (distinct per-node condition combinations) per
// mesh; sample s's entries are bc[bc_base, bc_base + codes)
constexpr int kEnsMaxBcCodes = 8;
// words (8 bytes) of one BcEntry
constexpr int kBcWords = (int)(sizeof(BcEntry) / 8);
static_assert(sizeof(BcEntry) % 8 == 0, "BcEntry is staged in 8-byte words");

struct EnsSample {
  int32_t mesh;
  int32_t node_base;     // first row of u, v, a, f, mass
  int32_t x_base;        // first row of the mesh's X and flags
  int32_t partial_base;  // mixed: first tile partial of the sample
  int32_t bc_base;       // first kinematic-condition entry
  int32_t kg_base;       // first block of the sample's moduli
  int64_t slot_base;     // FP64: first element force slot ([3][8 E8] per sample, E8: E_mesh rounded up to 8)
};

struct EnsMesh {
  int32_t num_nodes, num_elem;
  int32_t elem_base;     // first element of the mesh in the concatenated connectivity
  int32_t ptr_base;      // first node_ptr entry of the mesh's node->element CSR
  int32_t num_interior;  // mixed: device id of the mesh's boundary node 0
  int32_t bnd_pos;       // mixed: first bnd_ptr entry of the mesh
};

// mixed element work: tile t (of the concatenated tile tables) of sample s
struct EnsTile {
  int32_t t, s;
};

// a range of one sample: nodes [first, first + count) (node launches,
// boundary nodes of the mixed kernel) or elements of block `block`
// (FP64 element chunks), local to the sample's mesh
struct EnsRange {
  int32_t s, first, count, block;
};

// Read-only table entry at a uniform index, through the constant address
// space: it comes in by scalar loads, which do not count in the
// vector-memory counter.
template <typename T>
__device__ __forceinline__ T ens_load(const T* p, int64_t i)
{
  static_assert(sizeof(T) % 4 == 0, "loaded in 4-byte words");
  typedef const __attribute__((address_space(4))) uint32_t* cword;
  const cword w = (cword)(p + i);
  uint32_t    buf[sizeof(T) / 4];
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); k++) buf[k] = w[k];
  T r;
  __builtin_memcpy(&r, buf, sizeof(T));
  return r;
}

// One step of every sample: sample s's step n runs with times
// times[s * times_stride + n] (computed on the host exactly as the
// reference's integrator does, src/integrators/explicit_time_integrator.cc:192-195).
struct EnsStep {
  const StepTimes* times;
  int64_t          times_stride;
  int              step;
  int              output_bc, fuse, write_a, update_v;
  int              write_f;  // store the internal force (a run's last step)
};

// StepArgs of step es.step for sample s (s uniform).
__device__ __forceinline__ StepArgs ens_step_args(const EnsStep& es, int s)
{
  const int64_t row = (int64_t)s * es.times_stride + es.step;
  StepArgs      sa;
  sa.cur       = ens_load(es.times, row);
  sa.nxt       = es.fuse ? ens_load(es.times, row + 1) : sa.cur;
  sa.output_bc = es.output_bc;
  sa.fuse      = es.fuse;
  sa.write_a   = es.write_a;
  sa.write_f   = es.write_f;
  sa.update_v  = es.update_v;
  return sa;
}

// Shared node-side data of an ensemble.
struct EnsNodes {
  const EnsSample* samples;
  const EnsMesh*   meshes;
  const double*    X;      // [sum N_mesh][3]
  const uint8_t*   flags;  // [sum N_mesh]
  const double*    mass;   // [sum N_sample]
  double *         u, *v, *a, *f;
  const BcEntry*   bc;     // [sum codes]
  const BcEntry*   init;   // initial velocities per code (kind[c] = 1: v[c] = value[c])
};

// Reaction partials of an ensemble step: one per work item of the kernels
// that hold reaction nodes; sample s's total is -sum of partials[idx[k]],
// k in [ptr[s], ptr[s+1]).
struct EnsReaction {
  int      coord;     // -1: none
  double*  partials;  // this step's partials
};
struct EnsFold {
  const double* partials;  // a previous step's partials
  const int*    ptr;       // [S+1]
  const int*    idx;
  int           num_samples;
  double*       hist;      // sample s's total at hist[s * hist_stride + step]
  int64_t       hist_stride;
  int64_t       step;      // -1: nothing to fold
};

// ---- launchers ----
// u = v = a = f = 0, initial velocities, kinematic conditions at t = t_0
void launch_ens_reset(const EnsNodes& en, const EnsRange* chunks, int num_chunks, EnsStep es, hipStream_t s);
// first half of step es.step (kick, BC, drift, BC)
void launch_ens_prekick(const EnsNodes& en, const EnsRange* chunks, int num_chunks, EnsStep es, hipStream_t s);
// FP64: element forces into each sample's slots (rjac: each mesh's reference
// inverse Jacobians, launch_reference_jacobians, at 80 elem_base)
void launch_ens_element_fp64(const EnsNodes& en, const EnsRange* chunks, int num_chunks, const int* conn,
                             const double* rjac, const double* K, const double* G, double* slots, int* status,
                             hipStream_t s);
// FP64: CSR gather + node update, reaction partial per chunk at
// rx.partials[chunk]; the first waves first fold a previous step's totals
void launch_ens_node_fp64(const EnsNodes& en, const EnsRange* chunks, int num_chunks, const uint32_t* node_ptr,
                          const uint32_t* node_ent, const double* slots, EnsStep es, EnsReaction rx, EnsFold fold,
                          hipStream_t s);
// sample totals of a step's partials (after a run's last step)
void launch_ens_fold(EnsFold fold, hipStream_t s);
// lumped mass: element masses per sample (FP64 setup), then the CSR gather
void launch_ens_element_mass(const EnsNodes& en, const EnsRange* chunks, int num_chunks, const int* conn,
                             const double* rho, double* elem_mass, int* status, hipStream_t s);
void launch_ens_mass_gather(const EnsNodes& en, const EnsRange* chunks, int num_chunks, const uint32_t* node_ptr,
                            const uint32_t* node_ent, const double* elem_mass, double* mass, hipStream_t s);
// reaction histories sampled at num_points times, out [S][num_points]
void launch_ens_sample_history(const double* hist, int64_t hist_stride, int64_t n, int num_samples, int num_points,
                               double* out, hipStream_t s);

}  // namespace nsm
