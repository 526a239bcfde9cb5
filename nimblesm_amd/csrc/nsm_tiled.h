// nsm_tiled.h -- launchers of the tile-local mixed-precision step (internal).
//
// Device arrays of a TileSet (nsm_tiles.h) plus the partial forces of the
// tile-node slots that are not interior to their tile, stored by node (the
// partials of one node are contiguous, in ascending tile order).  A partial
// is summed in FP64 inside its tile and stored in FP32, the precision of the
// element forces it adds up.
#pragma once

#include "nsm_ensemble.h"
#include "nsm_kernels.h"
#include "nsm_node_step.h"

namespace nsm {

struct TileArgs {
  int             tile_elems;       // 256 or 128: threads per workgroup, nodes per tile <= 2*tile_elems
  int             resident_groups;  // element-kernel workgroups resident on the device at once
  int             num_tiles;
  int             num_bnd;
  const int32_t*  elem_off;    // [T+1]
  const int32_t*  node_off;    // [T+1]
  const int32_t*  nodes;       // [slots]
  const int32_t*  slot_dst;    // [slots] -1: interior node, else index of its partial
  const uint16_t* lconn;       // [E][8], internal element order
  const uint8_t*  color;       // [E]
  const int32_t*  num_colors;  // [T]
  const int32_t*  gather;      // [T] nonzero: per-corner gather assembly (TileSet::gather)
  int             bnd_base;    // device id of boundary node 0 (boundary node b is bnd_base + b)
  const int32_t*  bnd_ptr;     // [B+1]
  float*          partials;    // [partials][3]; node b's are [bnd_ptr[b], bnd_ptr[b+1])
};

// What one tile of an element launch works on: tile t of the tile tables,
// for sample s (ensembles; 0 otherwise).  The tables' node ids and partial
// indices are local to the tile's mesh; the offsets turn them into rows of
// the launch's arrays.
struct TileCtx {
  int t, s;
  int node_base;     // added to node ids: rows of u, v, a, f and the mass
  int x_delta;       // added to those rows: rows of X and the node flags
  int partial_base;  // added to partial indices
  int bc_base;       // first kinematic-condition entry of the sample
  int kg_base;       // first block of the sample's moduli
};

// The tiles [t_begin, t_begin + t_count) of one model, moduli from a block
// table, one step for all; workgroup 0 first writes the folded reaction
// total of a previous step.
struct ModelTileWork {
  BlockTable<float> bt;
  int               t_begin, t_count;
  StepArgs          sa;
  ReactionFold      fold;
};

// Ensemble: work item i is tile tiles[i].t of sample tiles[i].s, with the
// sample's moduli, kinematic-condition entries and step times; the waves of
// the launch first fold every sample's reaction total of the previous step.
struct EnsTileWork {
  const EnsTile*   tiles;
  int              count;
  const EnsSample* samples;
  const float*     K;           // sample s, block b at [samples[s].kg_base + b]
  const float*     G;
  const int32_t*   tile_block;  // block (of its mesh) of each tile
  const BcEntry*   bc;          // kEnsMaxBcCodes entries per sample at samples[s].bc_base
  EnsStep          es;
  EnsFold          fold;
};

// boundary nodes per workgroup of the boundary kernel (one per thread)
constexpr int kBoundaryNodesPerGroup = 256;
unsigned      boundary_grid(int num_bnd);

// Tiles [t_begin, t_begin + t_count) (all inside the blocks of bt): element
// arithmetic, force assembly in LDS, the step of the tile-interior nodes,
// partials of the others.  Reaction partials go to rx.partials[tile].
// Workgroup 0 first writes the folded reaction total of a previous step.
// F and sigma are tile-blocked: component c, point q of local element l of
// tile t (elements [e0, e0 + n)) at Fip[72 e0 + (4 c + q/2) 2n + 2l + q%2]
// and sip[48 e0 + (4 c + q/2) 2n + 2l + q%2]; with Fip = sip = nullptr the
// launch keeps neither.
void launch_element_tiled(const BlockTable<float>& bt, const TileArgs& ta, int t_begin, int t_count,
                          const double* X, const NodeArrays& na, float* Fip, float* sip, const StepArgs& sa,
                          ReactionArgs rx, ReactionFold fold, int* status, hipStream_t s,
                          const LaunchTimer& timer = LaunchTimer{});
// The same kernel over the tiles of every sample of an ensemble (ta: the
// concatenated tile tables of its meshes; na: all samples' node rows, X and
// flags of all meshes); it keeps no F or sigma.
void launch_ens_element_tiled(const EnsTileWork& work, const TileArgs& ta, const double* X, const NodeArrays& na,
                              ReactionArgs rx, int* status, hipStream_t s);
// Boundary nodes of every sample: group g (ranges[g]: sample, first local
// boundary node, count) updates count <= kBoundaryNodesPerGroup nodes of one
// sample; reaction partial of group g at rx.partials[rx_base + g].
void launch_ens_boundary_update(const TileArgs& ta, const EnsNodes& en, const EnsRange* groups, int num_groups,
                                EnsStep es, ReactionArgs rx, int rx_base, hipStream_t s);
// Element-kernel workgroups one CU holds at once (occupancy).
int element_tiled_groups_per_cu(int tile_elems);
// Step of the remaining nodes from their partials (ascending tile order).
// Reaction partials go to rx.partials[num_tiles + workgroup].
void launch_boundary_update(const TileArgs& ta, const NodeArrays& na, const StepArgs& sa, ReactionArgs rx,
                            hipStream_t s, const LaunchTimer& timer = LaunchTimer{});

}  // namespace nsm
