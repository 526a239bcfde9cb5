// nsm_tiled.hip -- mixed-precision step with tile-local force assembly.
//
// k_element_tiled: one workgroup per element tile (nsm_tiles.h), one thread
// per element.  The tile's nodes (X, u in FP64) are staged in LDS; each
// thread runs the mixed-precision element arithmetic (nsm_mixed_math.h),
// writes F and sigma, and its 24 node forces are summed per node in a fixed
// order: gathered per corner index after one barrier where no node repeats
// a corner index in the tile, else added into FP64 LDS accumulators colour
// by colour.  Tile-interior nodes then take their whole central-difference
// update here; every other node gets this tile's partial sum (FP64, stored
// in FP32).  k_boundary_update adds the partials of the remaining nodes in
// ascending tile order and updates them.  Each kernel also leaves one
// reaction partial per workgroup.
//
// The loads a tile needs are issued before its element arithmetic (one tile
// ahead for X, u and element data, two for node ids; at the tile's start for
// the flags, mass and v of its interior nodes) and the kinematic-condition
// table sits in LDS: the F and sigma stores issued during the arithmetic
// count in the same in-order vector-memory counter as loads, so a load
// issued after them would wait until all of them had completed.
#include "nsm_mixed_math.h"
#include "nsm_node_step.h"
#include "nsm_tiled.h"
#include "nsm_tiles.h"

#include <algorithm>

namespace nsm {
namespace {

// Sum over the workgroup in a fixed order (butterfly inside each 64-lane
// wave, then the wave sums in ascending wave order); valid in thread 0.
template <int kThreads>
__device__ __forceinline__ double block_sum(double v, double* red)
{
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
#pragma unroll
    for (int w = 0; w < kThreads / 64; w++) s += red[w];
  return s;
}

// Tile indices are uniform; readfirstlane keeps them in scalar registers, and
// the per-tile tables (read-only for the whole launch) are read through the
// constant address space, so they come in by scalar loads, which do not
// count in the vector-memory counter.
__device__ __forceinline__ int uniform(int t) { return __builtin_amdgcn_readfirstlane(t); }

template <typename T>
__device__ __forceinline__ T tile_table(const T* p, int t)
{
  return ((const __attribute__((address_space(4))) T*)p)[t];
}

// F / sigma rows go out as buffer stores: the resource covers one tile's
// block (base and size uniform), the row offset is a scalar operand and the
// lane's element offset one vector register shared by all 60 stores, so
// the stores need no per-store address arithmetic.  A store past the block
// is dropped by the range check.  0x00020000: gfx9 resource word 3 (32-bit
// data format, which untyped buffer accesses need).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t tile_rsrc(float* p, int bytes)
{
  return __builtin_amdgcn_make_buffer_rsrc(p, 0, bytes, 0x00020000);
}

constexpr int kNonTemporal = 2;  // cache-policy operand: nt

// Values of points 2p and 2p+1 of one component, side by side: one 8-byte
// store per lane.
template <int kAux>
__device__ __forceinline__ void store_pair(float even, float odd, __amdgpu_buffer_rsrc_t r, int lane_off,
                                           int row_off)
{
  typedef unsigned v2u __attribute__((ext_vector_type(2)));
  const v2u d = {__builtin_bit_cast(uint32_t, even), __builtin_bit_cast(uint32_t, odd)};
  __builtin_amdgcn_raw_buffer_store_b64(d, r, lane_off, row_off, kAux);
}

// Node ids of the (at most two) tile-node slots a thread stages, with the
// slot's partial destination (kInterior for a tile-interior node).  A thread
// past the tile's last slot gets that last slot's node (so every load stays
// unconditional) with destination kAbsent.
constexpr int kInterior = -1;
constexpr int kAbsent   = -2;
struct SlotIds {
  int gid[2];
  int dst[2];
};

template <int kThreads>
__device__ __forceinline__ SlotIds load_ids(const TileArgs& ta, int t)
{
  t            = uniform(t);
  const int n0 = tile_table(ta.node_off, t), nn = tile_table(ta.node_off, t + 1) - n0;
  SlotIds   r;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int ln = (int)threadIdx.x + k * kThreads;
    const int sl = n0 + min(ln, nn - 1);
    r.gid[k]     = ta.nodes[sl];
    r.dst[k]     = ln < nn ? ta.slot_dst[sl] : kAbsent;
  }
  return r;
}

// Element data of one thread: local connectivity (8 x uint16) and colour
// byte.  A thread past the tile's last element takes that last element (it
// recomputes it and rewrites identical F and sigma) with colour 0xff, so it
// adds no force.
struct ElemData {
  uint4 lq;
  int   cinfo;
};

template <int kThreads>
__device__ __forceinline__ ElemData load_elem(const TileArgs& ta, int t)
{
  t               = uniform(t);
  const int e0    = tile_table(ta.elem_off, t), ne = tile_table(ta.elem_off, t + 1) - e0;
  const int64_t e = e0 + (int64_t)min((int)threadIdx.x, ne - 1);
  ElemData      d;
  d.lq    = *reinterpret_cast<const uint4*>(ta.lconn + 8 * e);
  d.cinfo = (int)threadIdx.x < ne ? (int)ta.color[e] : 0xff;
  return d;
}

// LDS of the force assembly.  Colour tiles: one FP64 accumulator per node
// slot.  Gather tiles: per node slot, the element holding it at each corner
// index (kNoElement: none), and the elements' FP32 forces component-major
// (ef[3 j + c][element]: neighbouring elements in neighbouring banks, for
// the writes and for the gathers of neighbouring slots), so a slot sums its
// at most 8 contributions in ascending corner order after one barrier.
// Column kThreads of ef is zero (absent corners read it); staging a gather
// tile rewrites it, since a colour tile's accumulators overlap it.
constexpr uint16_t kNoElement = 0xffff;
template <int kThreads>
union Assembly {
  double acc[2 * kThreads][3];
  struct {
    uint16_t inv[2 * kThreads][8];
    float    ef[24][kThreads + 1];
  } g;
};

// Staging of one tile: its nodes' X and u into LDS, cleared assembly slots,
// and the loads of its interior nodes' flags (and v, m when the step updates
// velocities).  Values of non-interior slots are left unset: they are never
// read, and no merge of a default forces an early wait.
struct NodeState {
  int    flg[2];
  double v[2][3], m[2];
};

template <int kThreads>
__device__ __forceinline__ void stage(const SlotIds& ids, const double (&xr)[2][3], const double (&ur)[2][3],
                                      const NodeArrays& na, const StepArgs& sa, double (*sxu)[6],
                                      Assembly<kThreads>& as, bool gather, NodeState& ns)
{
  if (gather && threadIdx.x < 24) as.g.ef[threadIdx.x][kThreads] = 0.0f;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int ln = (int)threadIdx.x + k * kThreads;
    if (ids.dst[k] != kAbsent) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        sxu[ln][c]     = xr[k][c];
        sxu[ln][3 + c] = ur[k][c];
      }
      if (gather)
        *reinterpret_cast<uint4*>(as.g.inv[ln]) = make_uint4(~0u, ~0u, ~0u, ~0u);
      else
#pragma unroll
        for (int c = 0; c < 3; c++) as.acc[ln][c] = 0.0;
    }
    if (ids.dst[k] == kInterior) {
      ns.flg[k] = na.flags[ids.gid[k]];
      if (sa.update_v) {
        load3(na.v, ids.gid[k], ns.v[k]);
        ns.m[k] = na.mass[ids.gid[k]];
      }
    }
  }
}

// Persistent workgroups, each walking every G-th tile.  kThreads elements
// and at most 2*kThreads nodes per tile.  While tile t computes, X, u and
// element data of its next tile and node ids of the one after are in flight;
// the next tile is staged into LDS at the end of the iteration.
// Inside the loop every vector load and every F / sigma store is executed on
// every path (indices clamped), so at least the 120 stores of the element
// arithmetic always separate a load from its first use and the waits the
// compiler places never wait for those stores.  kStoreIp = false (off-nominal
// UQ samples, which keep no element data) sends the same stores to a small
// sink that stays in L2 instead of F and sigma: without them the compiler
// schedules the element arithmetic with more registers than a wave pair
// has, and the kernel is no faster for the bytes it saves.  The launch bound
// holds the kernel to two waves per SIMD (at most 256 registers): left free,
// the scheduler takes more for the gather assembly and halves occupancy.
template <int kThreads, bool kStoreIp>
__global__ __launch_bounds__(kThreads, 2) void k_element_tiled(const BlockTable<float> bt,
                                                            TileArgs ta, int t_begin, int t_count,
                                                            const double* __restrict__ X, NodeArrays na,
                                                            float* __restrict__ Fip, float* __restrict__ sip,
                                                            StepArgs sa, ReactionArgs rx, ReactionFold fold,
                                                            int* __restrict__ status)
{
  constexpr int kNodes = 2 * kThreads;
  // X and u of a node slot side by side (48 bytes): an element reads each
  // of its nodes with three 16-byte LDS reads
  __shared__ __attribute__((aligned(16))) double sxu[kNodes][6];
  __shared__ __attribute__((aligned(16))) Assembly<kThreads> as;
  __shared__ double  red[kThreads / 64];
  __shared__ BcEntry sbc[kMaxBcEntries];
  __shared__ int     bad;
  __shared__ double  fold_red[256];

  if (fold.n > 0 && blockIdx.x == 0) {  // uniform
    const double total = reaction_total<kThreads>(fold.partials, fold.tiles, fold.n, fold_red);
    if (threadIdx.x == 0) *fold.out = total;
  }

  // Workgroup w (numbered so that each XCD holds a contiguous run of w)
  // processes tiles t_begin + w, + G, + 2G, ... (G workgroups): at any time
  // the grid works on one window of G consecutive tiles, so the F and sigma
  // rows are written in one narrow address range and each XCD's L2 sees
  // neighbouring tiles.
  const int G    = (int)gridDim.x;
  const int w    = (int)xcd_tile(blockIdx.x, gridDim.x);
  const int tend = t_begin + t_count;
  const int tb   = uniform(t_begin + w);
  if (tb >= tend) return;  // uniform across the workgroup
  // last tile of this workgroup, and the tile after t (t itself if none)
  const int  te   = uniform(tb + ((tend - 1 - tb) / G) * G + 1);
  const auto next = [&](int t) { return t + G < te ? t + G : t; };
  const int  tid  = (int)threadIdx.x;

  // ---- prologue: stage tile tb, load tile tb+1's node ids
  if (tid == 0) bad = 0;
  if (sa.update_v)
    for (int b = tid; b < na.num_bc; b += kThreads) sbc[b] = na.bc_table[b];
  SlotIds ids      = load_ids<kThreads>(ta, tb);
  SlotIds ids_next = load_ids<kThreads>(ta, next(tb));
  double  xr[2][3], ur[2][3];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    load3(X, ids.gid[k], xr[k]);
    load3(na.u, ids.gid[k], ur[k]);
  }
  ElemData  cur = load_elem<kThreads>(ta, tb);
  NodeState ns;
  stage<kThreads>(ids, xr, ur, na, sa, sxu, as, tile_table(ta.gather, tb) != 0, ns);
  // every prologue load complete (vmcnt(0); lgkmcnt and expcnt at their
  // maximum, gfx9 encoding): the loop's waits then depend on its body only
  __builtin_amdgcn_s_waitcnt(0x0F70);

  for (int tl = tb; tl < te; tl += G) {
    const int t    = uniform(tl);
    const int e0   = tile_table(ta.elem_off, t);
    const int ncol = tile_table(ta.num_colors, t);
    __syncthreads();  // tile t staged

    // ---- prefetch (clamped to this workgroup's last tile): the next tile's
    // nodes and elements, the node ids of the one after
#pragma unroll
    for (int k = 0; k < 2; k++) {
      load3(X, ids_next.gid[k], xr[k]);
      load3(na.u, ids_next.gid[k], ur[k]);
    }
    const ElemData el_next   = load_elem<kThreads>(ta, next(t));
    const SlotIds  ids_after = load_ids<kThreads>(ta, next(next(t)));

    // ---- element arithmetic (every thread; see load_elem)
    const int ntl = tile_table(ta.elem_off, t + 1) - e0;  // elements of tile t
    const int lcl = min(tid, ntl - 1);
    float         f[8][3];
    uint16_t      lc[8];
    lc[0] = (uint16_t)(cur.lq.x & 0xffff);
    lc[1] = (uint16_t)(cur.lq.x >> 16);
    lc[2] = (uint16_t)(cur.lq.y & 0xffff);
    lc[3] = (uint16_t)(cur.lq.y >> 16);
    lc[4] = (uint16_t)(cur.lq.z & 0xffff);
    lc[5] = (uint16_t)(cur.lq.z >> 16);
    lc[6] = (uint16_t)(cur.lq.w & 0xffff);
    lc[7] = (uint16_t)(cur.lq.w >> 16);
    {
      double Xe[8][3], ue[8][3];
#pragma unroll
      for (int j = 0; j < 8; j++)
#pragma unroll
        for (int c = 0; c < 3; c++) {
          Xe[j][c] = sxu[lc[j]][c];
          ue[j][c] = sxu[lc[j]][3 + c];
        }
      const int b = uniform(block_of(bt, (int64_t)e0));  // tiles never straddle blocks
      // tile-blocked F and sigma: tile t owns [72 e0, 72 (e0 + n)) and
      // [48 e0, 48 (e0 + n)), in rows (component c, point pair p) of 2n
      // values, element l's points 2p and 2p+1 side by side, so each tile
      // writes two contiguous blocks with 8-byte stores; non-temporal stores
      // (the step never reads them back) keep node data and partials in cache
      float* const Ft   = kStoreIp ? Fip + (int64_t)e0 * 72 : ta.ip_sink + (t & 7) * 120 * ntl;
      float* const St   = kStoreIp ? sip + (int64_t)e0 * 48 : Ft + 72 * ntl;
      const auto   rF   = tile_rsrc(Ft, 72 * 4 * ntl);
      const auto   rS   = tile_rsrc(St, 48 * 4 * ntl);
      const int    loff = 8 * lcl;
      // the sink is rewritten every 8 tiles: keep it in L2
      constexpr int kAux  = kStoreIp ? kNonTemporal : 0;
      auto store = [&](int qm, int qp, const float* Fm, const float* Sm, const float* Fp, const float* Sp) {
        const bool   m_even = qm < qp;  // qm, qp: constants after inlining
        const int    p      = (m_even ? qm : qp) / 2;
        const float* Fe     = m_even ? Fm : Fp;
        const float* Fo     = m_even ? Fp : Fm;
        const float* Se     = m_even ? Sm : Sp;
        const float* So     = m_even ? Sp : Sm;
#pragma unroll
        for (int c = 0; c < 9; c++) store_pair<kAux>(Fe[c], Fo[c], rF, loff, 8 * (c * 4 + p) * ntl);
#pragma unroll
        for (int c = 0; c < 6; c++) store_pair<kAux>(Se[c], So[c], rS, loff, 8 * (c * 4 + p) * ntl);
      };
      if (!mixed::element(Xe, ue, bt.K[b], bt.G[b], store, f)) bad = 1;
    }

    const bool gather = tile_table(ta.gather, t) != 0;  // uniform
    if (gather) {
      // ---- gather tiles: every element publishes its forces and, for each
      // corner j, itself as the slot's corner-j element (no other element of
      // the tile holds that slot at corner j)
      if (tid < ntl) {
#pragma unroll
        for (int j = 0; j < 8; j++) as.g.inv[lc[j]][j] = (uint16_t)tid;
#pragma unroll
        for (int j = 0; j < 8; j++)
#pragma unroll
          for (int c = 0; c < 3; c++) as.g.ef[3 * j + c][tid] = f[j][c];
      }
      __syncthreads();
    } else {
      // ---- colour-ordered assembly: the elements of one colour touch
      // distinct nodes.  An element's 8 nodes are distinct unless it is
      // flagged kRepeatedNode, so its 24 updates are done as one batch of
      // reads, adds and writes; a flagged element adds node by node.
      const int  col      = cur.cinfo == 0xff ? -1 : (cur.cinfo & kColorMask);
      const bool repeated = cur.cinfo != 0xff && (cur.cinfo & kRepeatedNode) != 0;
      for (int c = 0; c < ncol; c++) {
        if (col == c) {
          if (!repeated) {
            double a[8][3];
#pragma unroll
            for (int j = 0; j < 8; j++)
#pragma unroll
              for (int k = 0; k < 3; k++) a[j][k] = as.acc[lc[j]][k];
#pragma unroll
            for (int j = 0; j < 8; j++)
#pragma unroll
              for (int k = 0; k < 3; k++) as.acc[lc[j]][k] = a[j][k] + (double)f[j][k];
          } else {
            for (int j = 0; j < 8; j++)
              for (int k = 0; k < 3; k++) as.acc[lc[j]][k] += (double)f[j][k];
          }
        }
        __syncthreads();
      }
    }

    // ---- interior nodes: complete step; others: this tile's partial
    double rxs = 0.0;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int ln = tid + k * kThreads;
      if (ids.dst[k] == kAbsent) continue;
      double fn[3];
      if (gather) {
        // the slot's contributions in ascending corner order, from 0.0
        const uint4    iv   = *reinterpret_cast<const uint4*>(as.g.inv[ln]);
        const unsigned w[4] = {iv.x, iv.y, iv.z, iv.w};
        fn[0] = fn[1] = fn[2] = 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const unsigned e  = (w[j >> 1] >> (16 * (j & 1))) & 0xffffu;
          const int      ec = e == kNoElement ? kThreads : (int)e;
#pragma unroll
          for (int c = 0; c < 3; c++) fn[c] += (double)as.g.ef[3 * j + c][ec];
        }
      } else {
#pragma unroll
        for (int c = 0; c < 3; c++) fn[c] = as.acc[ln][c];
      }
      if (ids.dst[k] == kInterior) {
        if (rx.coord >= 0 && (ns.flg[k] & kReactionFlag)) rxs += fn[rx.coord];
        double un[3] = {sxu[ln][3], sxu[ln][4], sxu[ln][5]};
        node_update(na, sbc, ids.gid[k], ns.flg[k], fn, un, ns.v[k], ns.m[k], sa);
      } else {
        float* const pp = ta.partials + 3 * (int64_t)ids.dst[k];
        pp[0]           = (float)fn[0];
        pp[1]           = (float)fn[1];
        pp[2]           = (float)fn[2];
      }
    }
    if (rx.coord >= 0) {  // uniform across the grid
      const double s = block_sum<kThreads>(rxs, red);
      if (threadIdx.x == 0) rx.partials[t] = s;
    }
    __syncthreads();  // tile t's LDS reads done

    // ---- stage the next tile
    ids      = ids_next;
    ids_next = ids_after;
    cur      = el_next;
    if (t + G < te) stage<kThreads>(ids, xr, ur, na, sa, sxu, as, tile_table(ta.gather, t + G) != 0, ns);
  }
  if (tid == 0 && bad) atomicOr(status, 1);
}

constexpr int kBndThreads = kBoundaryNodesPerGroup;

__global__ __launch_bounds__(kBndThreads) void k_boundary_update(TileArgs ta, NodeArrays na, StepArgs sa,
                                                                 ReactionArgs rx)
{
  __shared__ double red[kBndThreads / 64];
  const unsigned    g   = xcd_tile(blockIdx.x, gridDim.x);
  const int         bi  = (int)g * kBndThreads + threadIdx.x;
  double            rxs = 0.0;
  if (bi < ta.num_bnd) {
    // the node's own state is loaded first, so its latency overlaps the
    // dependent partial-sum loads
    const int64_t i     = (int64_t)ta.bnd_base + bi;
    const int     flags = na.flags[i];
    double        u[3], v[3], m = 1.0;
    if (sa.update_v) {
      load3(na.u, i, u);
      load3(na.v, i, v);
      m = na.mass[i];
    }
    const int p0 = ta.bnd_ptr[bi], p1 = ta.bnd_ptr[bi + 1];
    // up to 8 partials are loaded at once, then added in ascending order
    // (the same order as one at a time); more than 8 continue one by one
    constexpr int kBatch = 8;
    float         pv[kBatch][3];
#pragma unroll
    for (int i = 0; i < kBatch; i++)
      if (p0 + i < p1)
#pragma unroll
        for (int c = 0; c < 3; c++) pv[i][c] = ta.partials[3 * ((int64_t)p0 + i) + c];
    double f[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < kBatch; i++)
      if (p0 + i < p1) {
        f[0] += (double)pv[i][0];
        f[1] += (double)pv[i][1];
        f[2] += (double)pv[i][2];
      }
    for (int64_t p = (int64_t)p0 + kBatch; p < p1; p++) {
      f[0] += (double)ta.partials[3 * p];
      f[1] += (double)ta.partials[3 * p + 1];
      f[2] += (double)ta.partials[3 * p + 2];
    }
    if (rx.coord >= 0 && (flags & kReactionFlag)) rxs = f[rx.coord];
    node_update(na, na.bc_table, i, flags, f, u, v, m, sa);
  }
  if (rx.coord < 0) return;
  const double s = block_sum<kBndThreads>(rxs, red);
  if (threadIdx.x == 0) rx.partials[ta.num_tiles + g] = s;
}

}  // namespace

unsigned boundary_grid(int num_bnd) { return (unsigned)((num_bnd + kBndThreads - 1) / kBndThreads); }

template <int kThreads>
void launch_tiles(unsigned grid, const BlockTable<float>& bt, const TileArgs& ta, int t_begin, int t_count,
                  const double* X, const NodeArrays& na, float* Fip, float* sip, const StepArgs& sa, ReactionArgs rx,
                  ReactionFold fold, int* status, hipStream_t s)
{
  if (Fip)
    hipLaunchKernelGGL((k_element_tiled<kThreads, true>), dim3(grid), dim3(kThreads), 0, s, bt, ta, t_begin,
                       t_count, X, na, Fip, sip, sa, rx, fold, status);
  else
    hipLaunchKernelGGL((k_element_tiled<kThreads, false>), dim3(grid), dim3(kThreads), 0, s, bt, ta, t_begin,
                       t_count, X, na, Fip, sip, sa, rx, fold, status);
}

void launch_element_tiled(const BlockTable<float>& bt, const TileArgs& ta, int t_begin, int t_count,
                          const double* X, const NodeArrays& na, float* Fip, float* sip, const StepArgs& sa,
                          ReactionArgs rx, ReactionFold fold, int* status, hipStream_t s)
{
  if (t_count <= 0) return;
  const unsigned grid = (unsigned)std::min(t_count, std::max(ta.resident_groups, 1));
  switch (ta.tile_elems) {
    case 64: launch_tiles<64>(grid, bt, ta, t_begin, t_count, X, na, Fip, sip, sa, rx, fold, status, s); break;
    case 128: launch_tiles<128>(grid, bt, ta, t_begin, t_count, X, na, Fip, sip, sa, rx, fold, status, s); break;
    case 512: launch_tiles<512>(grid, bt, ta, t_begin, t_count, X, na, Fip, sip, sa, rx, fold, status, s); break;
    default: launch_tiles<256>(grid, bt, ta, t_begin, t_count, X, na, Fip, sip, sa, rx, fold, status, s);
  }
}

int element_tiled_groups_per_cu(int tile_elems)
{
  int n = 0;
  switch (tile_elems) {
    case 64: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_element_tiled<64, true>, 64, 0); break;
    case 128: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_element_tiled<128, true>, 128, 0); break;
    case 512: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_element_tiled<512, true>, 512, 0); break;
    default: (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_element_tiled<256, true>, 256, 0);
  }
  return n;
}

void launch_boundary_update(const TileArgs& ta, const NodeArrays& na, const StepArgs& sa, ReactionArgs rx,
                            hipStream_t s)
{
  if (ta.num_bnd <= 0) return;
  hipLaunchKernelGGL(k_boundary_update, dim3(boundary_grid(ta.num_bnd)), dim3(kBndThreads), 0, s, ta, na, sa, rx);
}

}  // namespace nsm
