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
#include <type_traits>

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
__device__ __forceinline__ SlotIds load_ids(const TileArgs& ta, const TileCtx& c)
{
  const int t  = uniform(c.t);
  const int n0 = tile_table(ta.node_off, t), nn = tile_table(ta.node_off, t + 1) - n0;
  SlotIds   r;
#pragma unroll
  for (int k = 0; k < 2; k++) {
    const int ln = (int)threadIdx.x + k * kThreads;
    const int sl = n0 + min(ln, nn - 1);
    r.gid[k]     = ta.nodes[sl] + c.node_base;
    const int d  = ta.slot_dst[sl];
    r.dst[k]     = ln < nn ? (d >= 0 ? d + c.partial_base : d) : kAbsent;
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
__device__ __forceinline__ void stage(const SlotIds& ids, int x_delta, const double (&xr)[2][3],
                                      const double (&ur)[2][3], const NodeArrays& na, int update_v,
                                      double (*sxu)[6], Assembly<kThreads>& as, bool gather, NodeState& ns)
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
      ns.flg[k] = na.flags[ids.gid[k] + x_delta];
      if (update_v) {
        load3(na.v, ids.gid[k], ns.v[k]);
        ns.m[k] = na.mass[ids.gid[k]];
      }
    }
  }
}

// ---- the two kinds of element-launch work (nsm_tiled.h) ----
__device__ __forceinline__ int work_count(const ModelTileWork& w) { return w.t_count; }
__device__ __forceinline__ int work_count(const EnsTileWork& w) { return w.count; }
__device__ __forceinline__ int work_update_v(const ModelTileWork& w) { return w.sa.update_v; }
__device__ __forceinline__ int work_update_v(const EnsTileWork& w) { return w.es.update_v; }

__device__ __forceinline__ TileCtx tile_ctx(const ModelTileWork& w, int i)
{
  return TileCtx{w.t_begin + i, 0, 0, 0, 0, 0, 0};
}
__device__ __forceinline__ TileCtx tile_ctx(const EnsTileWork& w, int i)
{
  const EnsTile   et = ens_load(w.tiles, (int64_t)uniform(i));
  const EnsSample sm = ens_load(w.samples, (int64_t)uniform(et.s));
  return TileCtx{uniform(et.t), uniform(et.s), uniform(sm.node_base), uniform(sm.x_base - sm.node_base),
                 uniform(sm.partial_base), uniform(sm.bc_base), uniform(sm.kg_base)};
}

__device__ __forceinline__ const BcEntry* work_bc(const ModelTileWork&) { return nullptr; }
__device__ __forceinline__ const BcEntry* work_bc(const EnsTileWork& w) { return w.bc; }

// moduli of the tile's block (tiles never straddle blocks)
__device__ __forceinline__ void tile_moduli(const ModelTileWork& w, const TileCtx&, int e0, float& K, float& G)
{
  const int b = uniform(block_of(w.bt, (int64_t)e0));
  K           = w.bt.K[b];
  G           = w.bt.G[b];
}
__device__ __forceinline__ void tile_moduli(const EnsTileWork& w, const TileCtx& c, int, float& K, float& G)
{
  const int b = c.kg_base + tile_table(w.tile_block, c.t);
  K           = tile_table(w.K, b);
  G           = tile_table(w.G, b);
}

__device__ __forceinline__ StepArgs tile_step(const ModelTileWork& w, const TileCtx&) { return w.sa; }
__device__ __forceinline__ StepArgs tile_step(const EnsTileWork& w, const TileCtx& c)
{
  return ens_step_args(w.es, c.s);
}

// index of the tile's reaction partial
__device__ __forceinline__ int tile_rx_slot(const ModelTileWork&, const TileCtx& c, int) { return c.t; }
__device__ __forceinline__ int tile_rx_slot(const EnsTileWork&, const TileCtx&, int i) { return i; }

// A previous step's reaction total(s): workgroup 0 sums the model's
// partials; in an ensemble every wave of the launch takes samples in turn.
template <int kThreads>
__device__ __forceinline__ void fold_previous(const ModelTileWork& w, double* fold_red)
{
  if (w.fold.n > 0 && blockIdx.x == 0) {  // uniform
    const double total = reaction_total<kThreads>(w.fold.partials, w.fold.tiles, w.fold.n, fold_red);
    if (threadIdx.x == 0) *w.fold.out = total;
  }
}
template <int kThreads>
__device__ __forceinline__ void fold_previous(const EnsTileWork& w, double*)
{
  const EnsFold& f = w.fold;
  if (f.step < 0) return;
  constexpr int kWaves = kThreads / 64;
  for (int s = (int)blockIdx.x * kWaves + (int)(threadIdx.x >> 6); s < f.num_samples;
       s += (int)gridDim.x * kWaves) {
    const int    p0    = f.ptr[s];
    const double total = wave_reaction_total(f.partials, f.idx + p0, f.ptr[s + 1] - p0);
    if ((threadIdx.x & 63) == 0) f.hist[(int64_t)s * f.hist_stride + f.step] = total;
  }
}

// Persistent workgroups, each walking every G-th tile.  kThreads elements
// and at most 2*kThreads nodes per tile.  While tile t computes, X, u and
// element data of its next tile and node ids of the one after are in flight;
// the next tile is staged into LDS at the end of the iteration.
// Inside the loop every vector load and every F / sigma store is executed on
// every path (indices clamped), so at least the 120 stores of the element
// arithmetic always separate a load from its first use and the waits the
// compiler places never wait for those stores.  kStoreIp = false (ensembles
// and off-nominal UQ samples, which keep no element data) stores nothing,
// but each Gauss-point pair's stress still passes through an empty asm
// statement where its stores would be: with no such point of use the
// compiler schedules the element arithmetic with more registers than a wave
// pair has and spills.  The launch bound
// holds the kernel to two waves per SIMD (at most 256 registers): left free,
// the scheduler takes more for the gather assembly and halves occupancy.
template <int kThreads, bool kStoreIp, typename Work>
__global__ __launch_bounds__(kThreads, 2) void k_element_tiled(const Work work, TileArgs ta,
                                                            const double* __restrict__ X, NodeArrays na,
                                                            float* __restrict__ Fip, float* __restrict__ sip,
                                                            ReactionArgs rx, int* __restrict__ status)
{
  constexpr int  kNodes      = 2 * kThreads;
  constexpr bool kEnsemble   = std::is_same<Work, EnsTileWork>::value;
  // ensembles stage each tile's sample's kinematic-condition entries (one
  // 8-byte word per thread)
  constexpr int  kTileBcWords = kBcWords * kEnsMaxBcCodes;
  static_assert(!kEnsemble || kTileBcWords <= kThreads, "one condition word per thread");
  // X and u of a node slot side by side (48 bytes): an element reads each
  // of its nodes with three 16-byte LDS reads
  __shared__ __attribute__((aligned(16))) double sxu[kNodes][6];
  __shared__ __attribute__((aligned(16))) Assembly<kThreads> as;
  __shared__ double  red[kThreads / 64];
  __shared__ BcEntry sbc[kMaxBcEntries];
  __shared__ int     bad;
  __shared__ double  fold_red[256];

  fold_previous<kThreads>(work, fold_red);

  // Workgroup w (numbered so that each XCD holds a contiguous run of w)
  // processes work items w, w + G, w + 2G, ... (G workgroups): at any time
  // the grid works on one window of G consecutive tiles, so the F and sigma
  // rows are written in one narrow address range and each XCD's L2 sees
  // neighbouring tiles.
  const int G        = (int)gridDim.x;
  const int w        = (int)xcd_tile(blockIdx.x, gridDim.x);
  const int count    = work_count(work);
  const int update_v = work_update_v(work);
  const int ib       = uniform(w);
  if (ib >= count) return;  // uniform across the workgroup
  // last work item of this workgroup, and the item after i (i itself if none)
  const int  ie   = uniform(ib + ((count - 1 - ib) / G) * G + 1);
  const auto next = [&](int i) { return i + G < ie ? i + G : i; };
  const int  tid  = (int)threadIdx.x;

  // ---- prologue: stage item ib, load item ib+G's node ids
  if (tid == 0) bad = 0;
  TileCtx c_first = tile_ctx(work, ib);
  if (!kEnsemble && update_v)
    for (int b = tid; b < na.num_bc; b += kThreads) sbc[b] = na.bc_table[b];
  if (kEnsemble && tid < kTileBcWords)
    reinterpret_cast<uint64_t*>(sbc)[tid] =
        reinterpret_cast<const uint64_t*>(work_bc(work) + c_first.bc_base)[tid];
  SlotIds ids      = load_ids<kThreads>(ta, c_first);
  SlotIds ids_next = load_ids<kThreads>(ta, tile_ctx(work, next(ib)));
  double  xr[2][3], ur[2][3];
#pragma unroll
  for (int k = 0; k < 2; k++) {
    load3(X, ids.gid[k] + c_first.x_delta, xr[k]);
    load3(na.u, ids.gid[k], ur[k]);
  }
  ElemData  cur = load_elem<kThreads>(ta, c_first.t);
  NodeState ns;
  stage<kThreads>(ids, c_first.x_delta, xr, ur, na, update_v, sxu, as, tile_table(ta.gather, c_first.t) != 0, ns);
  // every prologue load complete (vmcnt(0); lgkmcnt and expcnt at their
  // maximum, gfx9 encoding): the loop's waits then depend on its body only
  __builtin_amdgcn_s_waitcnt(0x0F70);

  for (int il = ib; il < ie; il += G) {
    const int     i    = uniform(il);
    const TileCtx c    = tile_ctx(work, i);
    const TileCtx cn   = tile_ctx(work, next(i));
    const int     t    = c.t;
    const int     e0   = tile_table(ta.elem_off, t);
    const int     ncol = tile_table(ta.num_colors, t);
    __syncthreads();  // tile t staged

    // ---- prefetch (clamped to this workgroup's last item): the next
    // item's nodes and elements, the node ids of the one after, and (in
    // ensembles) the next item's sample's condition entries
#pragma unroll
    for (int k = 0; k < 2; k++) {
      load3(X, ids_next.gid[k] + cn.x_delta, xr[k]);
      load3(na.u, ids_next.gid[k], ur[k]);
    }
    const ElemData el_next   = load_elem<kThreads>(ta, cn.t);
    const SlotIds  ids_after = load_ids<kThreads>(ta, tile_ctx(work, next(next(i))));
    uint64_t       bc_next   = 0;
    if (kEnsemble)
      bc_next = reinterpret_cast<const uint64_t*>(work_bc(work) + cn.bc_base)[min(tid, kTileBcWords - 1)];

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
      float Kt, Gt;
      tile_moduli(work, c, e0, Kt, Gt);
      // tile-blocked F and sigma: tile t owns [72 e0, 72 (e0 + n)) and
      // [48 e0, 48 (e0 + n)), in rows (component c, point pair p) of 2n
      // values, element l's points 2p and 2p+1 side by side, so each tile
      // writes two contiguous blocks with 8-byte stores; non-temporal stores
      // (the step never reads them back) keep node data and partials in cache
      float* const Ft   = kStoreIp ? Fip + (int64_t)e0 * 72 : nullptr;
      float* const St   = kStoreIp ? sip + (int64_t)e0 * 48 : nullptr;
      const auto   rF   = tile_rsrc(Ft, 72 * 4 * ntl);
      const auto   rS   = tile_rsrc(St, 48 * 4 * ntl);
      const int    loff = 8 * lcl;
      auto store = [&](int qm, int qp, const float* Fm, const float* Sm, const float* Fp, const float* Sp) {
        const bool   m_even = qm < qp;  // qm, qp: constants after inlining
        const int    p      = (m_even ? qm : qp) / 2;
        const float* Fe     = m_even ? Fm : Fp;
        const float* Fo     = m_even ? Fp : Fm;
        const float* Se     = m_even ? Sm : Sp;
        const float* So     = m_even ? Sp : Sm;
        if (kStoreIp) {
#pragma unroll
          for (int c = 0; c < 9; c++) store_pair<kNonTemporal>(Fe[c], Fo[c], rF, loff, 8 * (c * 4 + p) * ntl);
#pragma unroll
          for (int c = 0; c < 6; c++) store_pair<kNonTemporal>(Se[c], So[c], rS, loff, 8 * (c * 4 + p) * ntl);
        } else {
          // the pair's point of use (see above); emits no instruction
#pragma unroll
          for (int c = 0; c < 6; c++) asm volatile("" ::"v"(Se[c]), "v"(So[c]));
          __builtin_amdgcn_sched_barrier(0);
        }
      };
      if (!mixed::element(Xe, ue, Kt, Gt, store, f)) bad = 1;
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
    const StepArgs sa  = tile_step(work, c);
    double         rxs = 0.0;
#pragma unroll
    for (int k = 0; k < 2; k++) {
      const int ln = tid + k * kThreads;
      if (ids.dst[k] == kAbsent) continue;
      double fn[3];
      if (gather) {
        // the slot's contributions in ascending corner order, from 0.0
        const uint4    iv   = *reinterpret_cast<const uint4*>(as.g.inv[ln]);
        const unsigned wd[4] = {iv.x, iv.y, iv.z, iv.w};
        fn[0] = fn[1] = fn[2] = 0.0;
#pragma unroll
        for (int j = 0; j < 8; j++) {
          const unsigned e  = (wd[j >> 1] >> (16 * (j & 1))) & 0xffffu;
          const int      ec = e == kNoElement ? kThreads : (int)e;
#pragma unroll
          for (int cc = 0; cc < 3; cc++) fn[cc] += (double)as.g.ef[3 * j + cc][ec];
        }
      } else {
#pragma unroll
        for (int cc = 0; cc < 3; cc++) fn[cc] = as.acc[ln][cc];
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
      if (threadIdx.x == 0) rx.partials[tile_rx_slot(work, c, i)] = s;
    }
    __syncthreads();  // tile t's LDS reads done

    // ---- stage the next item
    ids      = ids_next;
    ids_next = ids_after;
    cur      = el_next;
    if (i + G < ie) {
      if (kEnsemble && tid < kTileBcWords) reinterpret_cast<uint64_t*>(sbc)[tid] = bc_next;
      stage<kThreads>(ids, cn.x_delta, xr, ur, na, update_v, sxu, as, tile_table(ta.gather, cn.t) != 0, ns);
    }
  }
  if (tid == 0 && bad) atomicOr(status, 1);
}

constexpr int kBndThreads = kBoundaryNodesPerGroup;

// Boundary node bi (partials [p0, p1)) at row i, flags at row xi: sums its
// partials in ascending tile order from 0.0 in FP64, then the node update.
// Returns its reaction contribution.
__device__ __forceinline__ double boundary_node(const TileArgs& ta, const NodeArrays& na, const BcEntry* table,
                                                int64_t i, int64_t xi, int64_t p0, int64_t p1, const StepArgs& sa,
                                                int rx_coord)
{
  // the node's own state is loaded first, so its latency overlaps the
  // dependent partial-sum loads
  const int flags = na.flags[xi];
  double    u[3], v[3], m = 1.0;
  if (sa.update_v) {
    load3(na.u, i, u);
    load3(na.v, i, v);
    m = na.mass[i];
  }
  // up to 8 partials are loaded at once, then added in ascending order
  // (the same order as one at a time); more than 8 continue one by one
  constexpr int kBatch = 8;
  float         pv[kBatch][3];
#pragma unroll
  for (int k = 0; k < kBatch; k++)
    if (p0 + k < p1)
#pragma unroll
      for (int c = 0; c < 3; c++) pv[k][c] = ta.partials[3 * (p0 + k) + c];
  double f[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int k = 0; k < kBatch; k++)
    if (p0 + k < p1) {
      f[0] += (double)pv[k][0];
      f[1] += (double)pv[k][1];
      f[2] += (double)pv[k][2];
    }
  for (int64_t p = p0 + kBatch; p < p1; p++) {
    f[0] += (double)ta.partials[3 * p];
    f[1] += (double)ta.partials[3 * p + 1];
    f[2] += (double)ta.partials[3 * p + 2];
  }
  const double rxs = rx_coord >= 0 && (flags & kReactionFlag) ? f[rx_coord] : 0.0;
  node_update(na, table, i, flags, f, u, v, m, sa);
  return rxs;
}

__global__ __launch_bounds__(kBndThreads) void k_boundary_update(TileArgs ta, NodeArrays na, StepArgs sa,
                                                                 ReactionArgs rx)
{
  __shared__ double red[kBndThreads / 64];
  const unsigned    g   = xcd_tile(blockIdx.x, gridDim.x);
  const int         bi  = (int)g * kBndThreads + threadIdx.x;
  double            rxs = 0.0;
  if (bi < ta.num_bnd) {
    const int64_t i = (int64_t)ta.bnd_base + bi;
    rxs = boundary_node(ta, na, na.bc_table, i, i, ta.bnd_ptr[bi], ta.bnd_ptr[bi + 1], sa, rx.coord);
  }
  if (rx.coord < 0) return;
  const double s = block_sum<kBndThreads>(rxs, red);
  if (threadIdx.x == 0) rx.partials[ta.num_tiles + g] = s;
}

// Ensemble boundary nodes: group g is nodes [first, first + count) of the
// boundary nodes of sample groups[g].s.
__global__ __launch_bounds__(kBndThreads) void k_ens_boundary_update(TileArgs ta, EnsNodes en,
                                                                     const EnsRange* __restrict__ groups, EnsStep es,
                                                                     ReactionArgs rx, int rx_base)
{
  __shared__ double red[kBndThreads / 64];
  const unsigned    g   = xcd_tile(blockIdx.x, gridDim.x);
  const EnsRange    gr  = ens_load(groups, (int64_t)g);
  const EnsSample   sm  = ens_load(en.samples, (int64_t)gr.s);
  const EnsMesh     ms  = ens_load(en.meshes, (int64_t)sm.mesh);
  const StepArgs    sa  = ens_step_args(es, gr.s);
  double            rxs = 0.0;
  if ((int)threadIdx.x < gr.count) {
    const int     bi  = gr.first + (int)threadIdx.x;
    const int64_t loc = (int64_t)ms.num_interior + bi;  // mesh-local device row
    NodeArrays    na;
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
    na.bc_table    = en.bc;
    na.num_bc      = 0;
    const int64_t p0 = (int64_t)sm.partial_base + ta.bnd_ptr[ms.bnd_pos + bi];
    const int64_t p1 = (int64_t)sm.partial_base + ta.bnd_ptr[ms.bnd_pos + bi + 1];
    rxs = boundary_node(ta, na, en.bc + sm.bc_base, sm.node_base + loc, sm.x_base + loc, p0, p1, sa, rx.coord);
  }
  if (rx.coord < 0) return;
  const double s = block_sum<kBndThreads>(rxs, red);
  if (threadIdx.x == 0) rx.partials[rx_base + (int)g] = s;
}

}  // namespace

unsigned boundary_grid(int num_bnd) { return (unsigned)((num_bnd + kBndThreads - 1) / kBndThreads); }

template <int kThreads, typename Work>
void launch_tiles(unsigned grid, const Work& work, const TileArgs& ta, const double* X, const NodeArrays& na,
                  float* Fip, float* sip, ReactionArgs rx, int* status, hipStream_t s, const LaunchTimer& t)
{
  if (Fip)
    hipExtLaunchKernelGGL((k_element_tiled<kThreads, true, Work>), dim3(grid), dim3(kThreads), 0, s, t.start, t.stop,
                          0, work, ta, X, na, Fip, sip, rx, status);
  else
    hipExtLaunchKernelGGL((k_element_tiled<kThreads, false, Work>), dim3(grid), dim3(kThreads), 0, s, t.start,
                          t.stop, 0, work, ta, X, na, Fip, sip, rx, status);
}

template <typename Work>
void launch_work(const Work& work, int count, const TileArgs& ta, const double* X, const NodeArrays& na, float* Fip,
                 float* sip, ReactionArgs rx, int* status, hipStream_t s, const LaunchTimer& t = LaunchTimer{})
{
  if (count <= 0) return complete_timer(t, s);
  const unsigned grid = (unsigned)std::min(count, std::max(ta.resident_groups, 1));
  switch (ta.tile_elems) {
    case 64: launch_tiles<64>(grid, work, ta, X, na, Fip, sip, rx, status, s, t); break;
    case 128: launch_tiles<128>(grid, work, ta, X, na, Fip, sip, rx, status, s, t); break;
    case 512: launch_tiles<512>(grid, work, ta, X, na, Fip, sip, rx, status, s, t); break;
    default: launch_tiles<256>(grid, work, ta, X, na, Fip, sip, rx, status, s, t);
  }
}

void launch_element_tiled(const BlockTable<float>& bt, const TileArgs& ta, int t_begin, int t_count,
                          const double* X, const NodeArrays& na, float* Fip, float* sip, const StepArgs& sa,
                          ReactionArgs rx, ReactionFold fold, int* status, hipStream_t s, const LaunchTimer& timer)
{
  const ModelTileWork work{bt, t_begin, t_count, sa, fold};
  launch_work(work, t_count, ta, X, na, Fip, sip, rx, status, s, timer);
}

void launch_ens_element_tiled(const EnsTileWork& work, const TileArgs& ta, const double* X, const NodeArrays& na,
                              ReactionArgs rx, int* status, hipStream_t s)
{
  launch_work(work, work.count, ta, X, na, nullptr, nullptr, rx, status, s);
}

int element_tiled_groups_per_cu(int tile_elems)
{
  int n = 0;
  switch (tile_elems) {
    case 64:
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_element_tiled<64, true, ModelTileWork>, 64, 0);
      break;
    case 128:
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_element_tiled<128, true, ModelTileWork>, 128, 0);
      break;
    case 512:
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_element_tiled<512, true, ModelTileWork>, 512, 0);
      break;
    default:
      (void)hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_element_tiled<256, true, ModelTileWork>, 256, 0);
  }
  return n;
}

void launch_boundary_update(const TileArgs& ta, const NodeArrays& na, const StepArgs& sa, ReactionArgs rx,
                            hipStream_t s, const LaunchTimer& timer)
{
  if (ta.num_bnd <= 0) return complete_timer(timer, s);
  hipExtLaunchKernelGGL(k_boundary_update, dim3(boundary_grid(ta.num_bnd)), dim3(kBndThreads), 0, s, timer.start,
                        timer.stop, 0, ta, na, sa, rx);
}

void launch_ens_boundary_update(const TileArgs& ta, const EnsNodes& en, const EnsRange* groups, int num_groups,
                                EnsStep es, ReactionArgs rx, int rx_base, hipStream_t s)
{
  if (num_groups <= 0) return;
  hipLaunchKernelGGL(k_ens_boundary_update, dim3((unsigned)num_groups), dim3(kBndThreads), 0, s, ta, en, groups, es,
                     rx, rx_base);
}

}  // namespace nsm
