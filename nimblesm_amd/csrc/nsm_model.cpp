// nsm_model.cpp -- host side of the C ABI: the device-resident ModelData,
// the explicit central-difference driver and the element / material seams.
//
// The model keeps every node and integration-point array in HBM for its whole
// life; a step is a short sequence of stream-ordered launches with no host
// round trip: FP64 runs the element kernel (force slots) then the CSR node
// kernel; mixed precision runs the tiled element kernel (LDS assembly and the
// update of tile-interior nodes) then the boundary-node kernel.  A step's
// reaction total is summed by workgroup 0 of the next step's first launch
// (or by a separate launch after a call's last step).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/nimblesm_amd.h"
#include "nsm_autotune.h"
#include "nsm_kernels.h"
#include "nsm_node_step.h"
#include "nsm_tiled.h"
#include "nsm_tiles.h"

namespace {

thread_local std::string g_last_error;

int fail(int status, const char* fmt, ...)
{
  char    buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return status;
}

}  // namespace

namespace nsm {
// the calling thread's error message, shared with the other host files
int set_last_error(int status, const char* msg)
{
  g_last_error = msg;
  return status;
}
}  // namespace nsm

namespace {

#define NSM_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t _e = (call);                                                                        \
    if (_e != hipSuccess) return fail(NSM_ERR_HIP, "%s: %s", #call, hipGetErrorString(_e));        \
  } while (0)

#define NSM_LAUNCHED()                                                                             \
  do {                                                                                             \
    hipError_t _e = hipGetLastError();                                                             \
    if (_e != hipSuccess) return fail(NSM_ERR_HIP, "kernel launch: %s", hipGetErrorString(_e));    \
  } while (0)

// Small device scratch shared by the stateless seam entry points: one status
// word per device, allocated on first use and kept for the process.
std::mutex       g_seam_mutex;
std::map<int, int*> g_seam_status;

int seam_status_buffer(int** out)
{
  int dev = 0;
  NSM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_seam_mutex);
  auto                        it = g_seam_status.find(dev);
  if (it == g_seam_status.end()) {
    int* p = nullptr;
    NSM_HIP(hipMalloc(&p, sizeof(int)));
    it = g_seam_status.emplace(dev, p).first;
  }
  *out = it->second;
  return NSM_OK;
}

template <typename T>
int dalloc(T** p, size_t count, size_t* total)
{
  *p = nullptr;
  if (count == 0) return NSM_OK;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(NSM_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
  }
  *total += count * sizeof(T);
  return NSM_OK;
}

template <typename T>
void free_dev(T*& p)
{
  if (p) (void)hipFree((void*)p);
  p = nullptr;
}

struct BcRecord {
  int              kind, coord;
  double           value;
  std::vector<int> nodes;
};

struct EventPair {
  hipEvent_t start, stop;
};

}  // namespace

struct nsm_model {
  int         device    = 0;
  hipStream_t stream    = nullptr;
  int         precision = NSM_FP64;

  int64_t              N = 0, E = 0;
  int                  num_blocks = 0;
  std::vector<int>     block_ids;
  std::vector<int64_t> blk_off;
  std::vector<double>  rho, K, G;

  // device arrays
  double*   X = nullptr;
  double*   u = nullptr;
  double*   v = nullptr;
  double*   a = nullptr;
  double*   f = nullptr;
  double*   m = nullptr;
  int*      conn = nullptr;
  uint32_t* node_ptr = nullptr;
  uint32_t* node_ent = nullptr;
  void*     Fip = nullptr;       // double or float [9][8][E] (component, ipt, element); tiled: per tile
  void*     sip = nullptr;       // double or float [6][8][E]; tiled: per tile (see ip_index)
  void*     slots = nullptr;     // double or float [3][8][E] element force slots
  double*   rjac = nullptr;      // FP64: reference inverse Jacobians and det, [E][8][10] (small meshes, see use_reference_table)
  double*   elem_tmp = nullptr;  // [8][E] setup scratch (element masses)
  double*   wg_min = nullptr;
  int64_t   wg_min_len = 0;
  int*      status = nullptr;
  uint8_t*  flags = nullptr;  // [N] bc table index + 1 | reaction bit
  nsm::BcEntry* bc_table = nullptr;
  int64_t   num_reaction_nodes = 0;
  int       reaction_coord = 2;
  double*   reaction_hist = nullptr;
  int64_t   reaction_cap = 0;
  double*   reaction_partials = nullptr;  // 2 x reaction_stride: one per workgroup of the kernels that
                                          // sum reactions, for even and for odd steps
  int64_t   reaction_stride = 0;
  int*      reaction_tiles = nullptr;     // partials holding reaction nodes, ascending
  int       num_reaction_tiles = 0;
  size_t    device_bytes = 0;

  // tile-local assembly (mixed precision): element tiles, device copies of
  // their arrays; integration-point arrays are stored in tile element order
  bool                 tiled = false;
  // element masses (and FP64 force slots) in the grouped slot order
  // fp64_slot(e, j): FP64 and tiled models; untiled mixed ones keep j E + e,
  // the order of their element kernel's FP32 slots
  bool grouped() const { return precision == NSM_FP64 || tiled; }
  nsm::TileSet         tiles;
  std::vector<int>     blk_tile_off;  // first tile of each block, [num_blocks + 1]
  nsm::TileArgs        tile_args{};
  std::vector<void*>   tile_bufs;
  // host transfers of node fields: mesh -> device node map and a staging
  // array in mesh order (allocated at the first transfer)
  const int32_t* d_node_new   = nullptr;  // in tile_bufs
  double*        d_stage      = nullptr;
  size_t         stage_values = 0;
  const int32_t* d_elem_pos   = nullptr;  // mesh element -> device element (tiled; in tile_bufs)

  // off-nominal ensemble (UQ): uq_exact exact samples, then uq_approx
  // approximate ones; sample s owns u, v, a, f at uq_state + (4 s + j) 3N
  int                 uq_exact = 0, uq_approx = 0;
  std::vector<double> uq_K, uq_G;           // [uq_exact][num_blocks]
  double*             uq_state    = nullptr;
  double*             uq_coeff    = nullptr;  // [uq_approx][uq_exact + 1]
  nsm::BcEntry*       uq_bc_table = nullptr;  // bc_table without prescribed displacements

  std::vector<BcRecord> bcs;
  std::vector<uint8_t>  reaction_mask;  // host copy of the reaction bit per node
  bool                  bc_dirty = true;
  int                   num_bc_entries = 0;
  bool                  have_mass = false;
  double                critical_dt = 0.0;

  bool                                timing       = false;
  int                                 timing_every = 1;  // time every n-th launch of a kernel
  int64_t                             timing_count[4] = {0, 0, 0, 0};
  std::vector<EventPair>              event_pool;
  size_t                              event_next = 0;
  std::vector<std::pair<int, size_t>> timed;  // (kernel, event index)

  int uq_samples() const { return uq_exact + uq_approx; }

  // node arrays of the nominal trajectory (sample < 0) or of UQ sample
  // `sample`, which shares mass and flags but not the kinematic-condition
  // table
  nsm::NodeArrays node_arrays(int sample = -1) const
  {
    nsm::NodeArrays na;
    na.num_nodes   = N;
    na.slot_stride = grouped() ? 8 * nsm::round_up8(E) : 8 * E;
    na.node_ptr    = node_ptr;
    na.node_ent  = node_ent;
    na.mass      = m;
    na.u         = u;
    na.v         = v;
    na.a         = a;
    na.f         = f;
    na.flags     = flags;
    na.bc_table  = bc_table;
    na.num_bc    = num_bc_entries;
    if (sample >= 0) {
      const int64_t N3 = 3 * N;
      double*       b  = uq_state + 4 * (int64_t)sample * N3;
      na.u             = b;
      na.v             = b + N3;
      na.a             = b + 2 * N3;
      na.f             = b + 3 * N3;
      na.bc_table      = uq_bc_table;
    }
    return na;
  }

  // moduli of block b for the nominal trajectory or an exact sample
  double bulk(int sample, int b) const { return sample < 0 ? K[b] : uq_K[(size_t)sample * num_blocks + b]; }
  double shear(int sample, int b) const { return sample < 0 ? G[b] : uq_G[(size_t)sample * num_blocks + b]; }
};

namespace {

// original element (conn order) stored at device element position e
inline int64_t element_of(const nsm_model* md, int64_t e)
{
  return md->tiled ? md->tiles.perm[(size_t)e] : e;
}

// device index of mesh node n, and mesh node at device index i
inline int64_t dev_node(const nsm_model* md, int64_t n) { return md->tiled ? md->tiles.node_new[(size_t)n] : n; }
inline int64_t mesh_node(const nsm_model* md, int64_t i) { return md->tiled ? md->tiles.node_old[(size_t)i] : i; }

// Calls f(device index, device element, component, point) for every value of
// an integration-point field with nc components: groups of 8 elements
// (fp64_ip_index) in FP64 models; element-fastest rows [c][q][E] in untiled
// mixed ones; in tiled ones each tile t (device elements [e0, e0 + n)) owns
// [8 nc e0, 8 nc (e0 + n)) with rows (4 c + q/2) of length 2n, element l's
// points q and q+1 (q even) side by side.
template <typename Fn>
void for_each_ip(const nsm_model* md, int nc, Fn&& fn)
{
  if (md->precision == NSM_FP64) {
    for (int64_t e = 0; e < md->E; e++)
      for (int c = 0; c < nc; c++)
        for (int q = 0; q < 8; q++) fn(nsm::fp64_ip_index(e, nc, c, q), e, c, q);
    return;
  }
  if (!md->tiled) {
    for (int c = 0; c < nc; c++)
      for (int q = 0; q < 8; q++)
        for (int64_t e = 0; e < md->E; e++) fn((int64_t)(c * 8 + q) * md->E + e, e, c, q);
    return;
  }
  const auto& eo = md->tiles.elem_off;
  for (int t = 0; t < md->tiles.num_tiles; t++) {
    const int64_t e0 = eo[(size_t)t], n = eo[(size_t)t + 1] - e0, base = 8 * nc * e0;
    for (int c = 0; c < nc; c++)
      for (int q = 0; q < 8; q++)
        for (int64_t l = 0; l < n; l++) fn(base + (c * 4 + q / 2) * 2 * n + 2 * l + (q & 1), e0 + l, c, q);
  }
}

int ensure_device(nsm_model* md)
{
  NSM_HIP(hipSetDevice(md->device));
  return NSM_OK;
}

// Every timing_every-th launch of each kernel is timed by an event pair.
// With `timer` the launch records the pair itself at the kernel's start and
// end (time_end then has nothing to do); without, time_begin and time_end
// record it on the stream around the launch.
int time_begin(nsm_model* md, int kernel, size_t* idx, nsm::LaunchTimer* timer = nullptr)
{
  *idx = (size_t)-1;
  if (timer) *timer = nsm::LaunchTimer{};
  if (!md->timing) return NSM_OK;
  if (kernel >= 0 && kernel < 4 && (md->timing_count[kernel]++ % md->timing_every) != 0) return NSM_OK;
  if (md->event_next == md->event_pool.size()) {
    EventPair p;
    NSM_HIP(hipEventCreate(&p.start));
    NSM_HIP(hipEventCreate(&p.stop));
    md->event_pool.push_back(p);
  }
  const size_t i = md->event_next++;
  md->timed.emplace_back(kernel, i);
  if (timer) {
    *timer = nsm::LaunchTimer{md->event_pool[i].start, md->event_pool[i].stop};
    return NSM_OK;
  }
  *idx = i;
  NSM_HIP(hipEventRecord(md->event_pool[i].start, md->stream));
  return NSM_OK;
}

int time_end(nsm_model* md, size_t idx)
{
  if (idx == (size_t)-1) return NSM_OK;
  NSM_HIP(hipEventRecord(md->event_pool[idx].stop, md->stream));
  return NSM_OK;
}

// Resolve the prescribed conditions, in insertion order, to their net effect
// per node and coordinate, as the reference's sequential loops leave it
// (src/nimble_boundary_condition_manager.h:136-217): the last condition
// decides, except that a prescribed displacement does nothing while
// delta_t <= 0, where the last prescribed velocity before it still holds
// (nsm::kBcDisplacementOverVelocity).
int rebuild_bcs(nsm_model* md)
{
  if (!md->bc_dirty) return NSM_OK;
  std::vector<nsm::BcEntry> per_node;
  std::vector<int>          touched;
  std::map<int, size_t>     slot_of;
  for (const auto& b : md->bcs) {
    if (b.kind == NSM_BC_INITIAL_VELOCITY) continue;
    for (int n : b.nodes) {
      auto it = slot_of.find(n);
      if (it == slot_of.end()) {
        nsm::BcEntry z;
        std::memset(&z, 0, sizeof(z));
        it = slot_of.emplace(n, per_node.size()).first;
        per_node.push_back(z);
        touched.push_back(n);
      }
      nsm::BcEntry& e = per_node[it->second];
      const int     c = b.coord;
      if (b.kind == NSM_BC_PRESCRIBED_VELOCITY) {
        e.kind[c]     = nsm::kBcVelocity;
        e.value[c]    = b.value;
        e.fallback[c] = 0.0;
      } else {  // prescribed displacement: over an earlier velocity condition?
        const bool over_v = e.kind[c] == nsm::kBcVelocity || e.kind[c] == nsm::kBcDisplacementOverVelocity;
        const double v    = e.kind[c] == nsm::kBcVelocity ? e.value[c] : e.fallback[c];
        e.kind[c]         = over_v ? nsm::kBcDisplacementOverVelocity : nsm::kBcDisplacement;
        e.value[c]        = b.value;
        e.fallback[c]     = over_v ? v : 0.0;
      }
    }
  }
  // distinct combinations -> table
  std::vector<nsm::BcEntry> table;
  std::vector<uint8_t>      code((size_t)md->N, 0);
  for (size_t t = 0; t < touched.size(); t++) {
    const nsm::BcEntry& e   = per_node[t];
    int                 idx = -1;
    for (size_t j = 0; j < table.size(); j++)
      if (std::memcmp(&table[j], &e, sizeof(e)) == 0) {
        idx = (int)j;
        break;
      }
    if (idx < 0) {
      if ((int)table.size() >= nsm::kMaxBcEntries)
        return fail(NSM_ERR_ARG, "more than %d distinct kinematic-condition combinations are not supported",
                    nsm::kMaxBcEntries);
      idx = (int)table.size();
      table.push_back(e);
    }
    code[(size_t)dev_node(md, touched[t])] = (uint8_t)(idx + 1);
  }
  if (md->bc_table) {
    NSM_HIP(hipFree(md->bc_table));
    md->bc_table = nullptr;
  }
  size_t dummy = 0;
  int    st    = dalloc(&md->bc_table, std::max<size_t>(table.size(), 1), &dummy);
  if (st) return st;
  if (!table.empty())
    NSM_HIP(hipMemcpy(md->bc_table, table.data(), table.size() * sizeof(nsm::BcEntry), hipMemcpyHostToDevice));
  // UQ samples take initial and prescribed velocities but not prescribed
  // displacements (src/nimble_boundary_condition_manager.h:184-215 sets the
  // nominal velocity only), so a sample keeps the last prescribed velocity:
  // the value of a velocity entry, the fallback of a displacement entry
  // listed after one, nothing for a displacement entry without one
  free_dev(md->uq_bc_table);
  if (md->uq_samples() > 0) {
    std::vector<nsm::BcEntry> sample_table(table);
    for (auto& e : sample_table)
      for (int c = 0; c < 3; c++) {
        if (e.kind[c] == nsm::kBcDisplacementOverVelocity) {
          e.kind[c]  = nsm::kBcVelocity;
          e.value[c] = e.fallback[c];
        } else if (e.kind[c] == nsm::kBcDisplacement) {
          e.kind[c]  = 0;
          e.value[c] = 0.0;
        }
        e.fallback[c] = 0.0;
      }
    if ((st = dalloc(&md->uq_bc_table, std::max<size_t>(sample_table.size(), 1), &dummy))) return st;
    if (!sample_table.empty())
      NSM_HIP(hipMemcpy(md->uq_bc_table, sample_table.data(), sample_table.size() * sizeof(nsm::BcEntry),
                        hipMemcpyHostToDevice));
  }
  std::vector<int> tiles;
  const bool       have_rx = md->reaction_mask.size() == code.size();
  for (size_t i = 0; i < code.size(); i++)  // device node order
    if (have_rx && md->reaction_mask[(size_t)mesh_node(md, (int64_t)i)]) {
      code[i] |= nsm::kReactionFlag;
      const int t = (int)(i / 256);  // node tile of the node kernel
      if (!md->tiled && (tiles.empty() || tiles.back() != t)) tiles.push_back(t);
    }
  if (md->tiled && have_rx) {
    // element tiles whose interior nodes hold reaction nodes, then the
    // workgroups of the boundary kernel that do
    const nsm::TileSet& ts = md->tiles;
    for (int t = 0; t < ts.num_tiles; t++)
      for (int s = ts.node_off[(size_t)t]; s < ts.node_off[(size_t)t + 1]; s++)
        if ((ts.node_info[(size_t)s] & nsm::kInteriorNode) &&
            md->reaction_mask[(size_t)mesh_node(md, ts.nodes[(size_t)s])]) {
          tiles.push_back(t);
          break;
        }
    for (size_t b = 0; b < ts.bnd_nodes.size(); b++)
      if (md->reaction_mask[(size_t)mesh_node(md, ts.bnd_nodes[b])]) {
        const int g = ts.num_tiles + (int)(b / nsm::kBoundaryNodesPerGroup);
        if (tiles.empty() || tiles.back() != g) tiles.push_back(g);
      }
  }
  NSM_HIP(hipMemcpy(md->flags, code.data(), code.size(), hipMemcpyHostToDevice));
  free_dev(md->reaction_tiles);
  md->num_reaction_tiles = 0;
  if (!tiles.empty()) {
    size_t dummy2 = 0;
    int    st2    = dalloc(&md->reaction_tiles, tiles.size(), &dummy2);
    if (st2) return st2;
    NSM_HIP(hipMemcpy(md->reaction_tiles, tiles.data(), tiles.size() * sizeof(int), hipMemcpyHostToDevice));
    md->num_reaction_tiles = (int)tiles.size();
  }
  md->num_bc_entries = (int)table.size();
  md->bc_dirty       = false;
  return NSM_OK;
}

int check_status(nsm_model* md)
{
  int h = 0;
  NSM_HIP(hipMemcpyAsync(&h, md->status, sizeof(int), hipMemcpyDeviceToHost, md->stream));
  NSM_HIP(hipStreamSynchronize(md->stream));
  if (h) {
    NSM_HIP(hipMemsetAsync(md->status, 0, sizeof(int), md->stream));
    NSM_HIP(hipStreamSynchronize(md->stream));
    return fail(NSM_ERR_NEG_JACOBIAN, "non-positive Jacobian determinant in an element (Invert3x3)");
  }
  return NSM_OK;
}

// F / sigma destination of an element launch: the model's arrays for the
// nominal trajectory, none for UQ samples (the reference keeps no element
// data for them, src/uq/nimble_uq_block.cc:101-163)
void* ip_out(const nsm_model*, void* p, int sample) { return sample < 0 ? p : nullptr; }

nsm::BlockTable<float> float_table(const nsm_model* md, int b0, int nb, int sample)
{
  nsm::BlockTable<float> bt{};
  bt.count = nb;
  for (int b = 0; b <= nb; b++) bt.off[b] = md->blk_off[b0 + b];
  for (int b = 0; b < nb; b++) {
    bt.K[b] = (float)md->bulk(sample, b0 + b);
    bt.G[b] = (float)md->shear(sample, b0 + b);
  }
  return bt;
}

// All blocks in one launch when they fit one block table, else groups of
// kMaxLaunchBlocks consecutive blocks; displacement and moduli of the nominal
// trajectory (sample < 0) or of exact UQ sample `sample`.
int launch_elements(nsm_model* md, int sample, int fp64_mode = nsm::kElementFull)
{
  const double* u = md->node_arrays(sample).u;
  for (int b0 = 0; b0 < md->num_blocks; b0 += nsm::kMaxLaunchBlocks) {
    const int nb = std::min(nsm::kMaxLaunchBlocks, md->num_blocks - b0);
    size_t           ti;
    nsm::LaunchTimer timer;
    int              st = time_begin(md, NSM_KERNEL_ELEMENT, &ti, &timer);
    if (st) return st;
    if (md->precision == NSM_FP64) {
      nsm::BlockTable<double> bt{};
      bt.count = nb;
      for (int b = 0; b <= nb; b++) bt.off[b] = md->blk_off[b0 + b];
      for (int b = 0; b < nb; b++) {
        bt.K[b] = md->bulk(sample, b0 + b);
        bt.G[b] = md->shear(sample, b0 + b);
      }
      nsm::launch_element_fp64(bt, md->E, md->conn, md->X, u, md->rjac, (double*)ip_out(md, md->Fip, sample),
                               (double*)ip_out(md, md->sip, sample), (double*)md->slots, md->status, md->stream,
                               fp64_mode, timer);
    } else {
      nsm::launch_element_mixed(float_table(md, b0, nb, sample), md->E, md->conn, md->X, u,
                                (float*)ip_out(md, md->Fip, sample), (float*)ip_out(md, md->sip, sample),
                                (float*)md->slots, md->status, md->stream, timer);
    }
    NSM_LAUNCHED();
    if ((st = time_end(md, ti))) return st;
  }
  return NSM_OK;
}

// Reaction partials of `step`: two buffers used by alternate steps, so a
// step's partials can still be summed while the next step writes its own.
double* reaction_partials_of(const nsm_model* md, int64_t step)
{
  return md->reaction_partials + (step & 1) * md->reaction_stride;
}

bool wants_reaction(const nsm_model* md, int64_t step)
{
  return md->num_reaction_tiles > 0 && step >= 0 && step < md->reaction_cap;
}

// Reaction sum of `step` (if a reaction set exists and the step is recorded):
// the arguments the kernels fill their partials with.
bool reaction_args(const nsm_model* md, int64_t step, nsm::ReactionArgs* rx)
{
  *rx                = nsm::ReactionArgs{-1, nullptr};
  const bool want_rx = wants_reaction(md, step);
  if (want_rx) *rx = nsm::ReactionArgs{md->reaction_coord, reaction_partials_of(md, step)};
  return want_rx;
}

int launch_reaction(nsm_model* md, int64_t step)
{
  size_t ti;
  int    st = time_begin(md, NSM_KERNEL_REACTION, &ti);
  if (st) return st;
  nsm::launch_reaction_finalize(reaction_partials_of(md, step), md->reaction_tiles, md->num_reaction_tiles,
                                md->reaction_hist + step, md->stream);
  NSM_LAUNCHED();
  return time_end(md, ti);
}

// The reaction total of `fold_step` (a previous step, or -1), summed by
// workgroup 0 of the next launch that takes it.
nsm::ReactionFold reaction_fold(const nsm_model* md, int64_t fold_step)
{
  if (fold_step < 0) return nsm::ReactionFold{nullptr, nullptr, 0, nullptr};
  return nsm::ReactionFold{reaction_partials_of(md, fold_step), md->reaction_tiles, md->num_reaction_tiles,
                           md->reaction_hist + fold_step};
}

// Node kernel of the slot path: it writes the reaction partials of `step`
// (nominal trajectory only; left pending, `*pending` = step when there are
// any) and sums those of `fold_step` first.
int launch_node(nsm_model* md, const nsm::StepArgs& sa, int64_t step, int64_t fold_step, int64_t* pending,
                int sample)
{
  size_t           ti;
  nsm::LaunchTimer timer;
  int              st = time_begin(md, NSM_KERNEL_NODE, &ti, &timer);
  if (st) return st;
  nsm::ReactionArgs       rx;
  const bool              want_rx = reaction_args(md, sample < 0 ? step : -1, &rx);
  const nsm::NodeArrays   na      = md->node_arrays(sample);
  const nsm::ReactionFold fold    = reaction_fold(md, fold_step);
  if (md->precision == NSM_FP64)
    nsm::launch_node_update<double>(na, (const double*)md->slots, sa, rx, fold, md->stream, timer);
  else
    nsm::launch_node_update<float>(na, (const float*)md->slots, sa, rx, fold, md->stream, timer);
  NSM_LAUNCHED();
  if ((st = time_end(md, ti))) return st;
  *pending = want_rx ? step : -1;
  return NSM_OK;
}

// Tile-local step: element tiles (with the interior-node update), then the
// boundary-node update.  The reaction total of `step` is left to the caller
// (`*pending` = step when it has one); the total of `fold_step` (a previous
// step, or -1) is summed by the first element launch.
int launch_tiled(nsm_model* md, const nsm::StepArgs& sa, int64_t step, int64_t fold_step, int64_t* pending,
                 int sample)
{
  nsm::ReactionArgs     rx;
  const bool            want_rx = reaction_args(md, step, &rx);
  const nsm::NodeArrays na      = md->node_arrays(sample);
  nsm::ReactionFold     fold    = reaction_fold(md, fold_step);
  int st = NSM_OK;
  for (int b0 = 0; b0 < md->num_blocks; b0 += nsm::kMaxLaunchBlocks) {
    const int nb = std::min(nsm::kMaxLaunchBlocks, md->num_blocks - b0);
    const int t0 = md->blk_tile_off[(size_t)b0], t1 = md->blk_tile_off[(size_t)(b0 + nb)];
    if (t1 <= t0) continue;  // no launch: the fold goes to the next one
    size_t           ti;
    nsm::LaunchTimer timer;
    if ((st = time_begin(md, NSM_KERNEL_ELEMENT, &ti, &timer))) return st;
    nsm::launch_element_tiled(float_table(md, b0, nb, sample), md->tile_args, t0, t1 - t0, md->X, na,
                              (float*)ip_out(md, md->Fip, sample), (float*)ip_out(md, md->sip, sample), sa, rx,
                              fold, md->status, md->stream, timer);
    NSM_LAUNCHED();
    if ((st = time_end(md, ti))) return st;
    fold.n = 0;
  }
  size_t           ti;
  nsm::LaunchTimer timer;
  if ((st = time_begin(md, NSM_KERNEL_NODE, &ti, &timer))) return st;
  nsm::launch_boundary_update(md->tile_args, na, sa, rx, md->stream, timer);
  NSM_LAUNCHED();
  if ((st = time_end(md, ti))) return st;
  *pending = want_rx ? step : -1;
  return NSM_OK;
}

// One trajectory's force evaluation and node update (update_v = 0: f_int
// only).  The reaction total of `step` is left pending (`*pending` = step
// when there is one) and that of `fold_step` is summed by the first launch
// that can take it.  UQ samples (sample >= 0) pass step = fold_step = -1:
// no reaction.
int launch_trajectory(nsm_model* md, const nsm::StepArgs& sa, int64_t step, int64_t fold_step, int64_t* pending,
                      int sample)
{
  *pending = -1;
  if (md->tiled) return launch_tiled(md, sa, step, fold_step, pending, sample);
  int st = launch_elements(md, sample);
  return st ? st : launch_node(md, sa, step, fold_step, pending, sample);
}

// Closure forces of the approximate UQ samples from this step's nominal and
// exact forces, with their node update.
int launch_closure(nsm_model* md, const nsm::StepArgs& sa)
{
  size_t ti;
  int    st = time_begin(md, NSM_KERNEL_CLOSURE, &ti);
  if (st) return st;
  const nsm::UqClosure uc{md->uq_exact, md->uq_approx, 3 * md->N, md->f, md->uq_state, md->uq_coeff};
  nsm::launch_uq_closure(md->node_arrays(0), uc, sa, md->stream);
  NSM_LAUNCHED();
  return time_end(md, ti);
}

// One step of every trajectory: exact UQ samples (their element launches
// store no F or sigma), the nominal trajectory, then the closure of the
// approximate samples, which reads this step's nominal and exact forces.
// Samples are independent of each other and of the nominal trajectory until
// the closure.
int launch_step(nsm_model* md, const nsm::StepArgs& sa, int64_t step, int64_t fold_step, int64_t* pending)
{
  int     st     = NSM_OK;
  int64_t unused = -1;
  for (int k = 0; k < md->uq_exact; k++)
    if ((st = launch_trajectory(md, sa, -1, -1, &unused, k))) return st;
  if ((st = launch_trajectory(md, sa, step, fold_step, pending, -1))) return st;
  return md->uq_approx > 0 ? launch_closure(md, sa) : NSM_OK;
}

// Device copies of the tile arrays.
template <typename T>
int upload(nsm_model* md, const std::vector<T>& h, const T** out)
{
  T*  p  = nullptr;
  int st = dalloc(&p, std::max<size_t>(h.size(), 1), &md->device_bytes);
  if (st) return st;
  md->tile_bufs.push_back((void*)p);
  if (!h.empty()) NSM_HIP(hipMemcpy(p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = p;
  return NSM_OK;
}

int setup_tiles(nsm_model* md)
{
  const nsm::TileSet& ts = md->tiles;
  nsm::TileArgs&      ta = md->tile_args;
  // the element kernel runs one thread per tile element and stages at most
  // two tile nodes per thread
  if (ts.tile_elems != 64 && ts.tile_elems != 128 && ts.tile_elems != 256 && ts.tile_elems != 512)
    return fail(NSM_ERR_ARG, "unsupported tile size");
  for (int t = 0; t < ts.num_tiles; t++)
    if (ts.elem_off[(size_t)t + 1] - ts.elem_off[(size_t)t] > ts.tile_elems ||
        ts.node_off[(size_t)t + 1] - ts.node_off[(size_t)t] > 2 * ts.tile_elems)
      return fail(NSM_ERR_ARG, "element tile %d exceeds the kernel's tile size", t);
  ta.tile_elems          = ts.tile_elems;
  ta.bnd_base            = (int)ts.num_interior;
  ta.num_tiles           = ts.num_tiles;
  int cus                = 0;
  NSM_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, md->device));
  ta.resident_groups = std::max(cus, 1) * std::max(nsm::element_tiled_groups_per_cu(ts.tile_elems), 1);
  ta.num_bnd             = (int)ts.bnd_nodes.size();
  int st                 = 0;
  if ((st = upload(md, ts.elem_off, &ta.elem_off)) || (st = upload(md, ts.node_off, &ta.node_off)) ||
      (st = upload(md, ts.nodes, &ta.nodes)) || (st = upload(md, ts.slot_dst, &ta.slot_dst)) ||
      (st = upload(md, ts.lconn, &ta.lconn)) || (st = upload(md, ts.color, &ta.color)) ||
      (st = upload(md, std::vector<int32_t>(ts.num_colors.begin(), ts.num_colors.end()), &ta.num_colors)) ||
      (st = upload(md, std::vector<int32_t>(ts.gather.begin(), ts.gather.end()), &ta.gather)) ||
      (st = upload(md, ts.bnd_ptr, &ta.bnd_ptr)))
    return st;
  float* partials = nullptr;
  if ((st = dalloc(&partials, std::max<size_t>(3 * ts.bnd_ent.size(), 1), &md->device_bytes))) return st;
  md->tile_bufs.push_back((void*)partials);
  ta.partials = partials;
  md->blk_tile_off.assign((size_t)md->num_blocks + 1, 0);
  for (int b = 0; b <= md->num_blocks; b++)
    md->blk_tile_off[(size_t)b] =
        (int)(std::lower_bound(ts.elem_off.begin(), ts.elem_off.end(), (int32_t)md->blk_off[(size_t)b]) -
              ts.elem_off.begin());
  return NSM_OK;
}


nsm::StepTimes step_times(double* time_current, double dt_user)
{
  // src/integrators/explicit_time_integrator.cc:192-195
  nsm::StepTimes t;
  t.t_prev  = *time_current;
  *time_current += dt_user;
  t.t_cur   = *time_current;
  t.dt      = t.t_cur - t.t_prev;
  t.half_dt = 0.5 * t.dt;
  return t;
}

bool is_output(int64_t step, int64_t num_load_steps, int64_t output_frequency)
{
  if (output_frequency == 0) return false;
  return step % output_frequency == 0 || step == num_load_steps - 1;
}

// device node array with w values per node -> host array in mesh node order
// device staging array of at least `need` values
int ensure_stage(nsm_model* md, size_t need)
{
  if (md->stage_values < need) {
    NSM_HIP(hipStreamSynchronize(md->stream));
    free_dev(md->d_stage);
    md->device_bytes -= std::min(md->device_bytes, md->stage_values * sizeof(double));
    md->stage_values = 0;
    int st = dalloc(&md->d_stage, need, &md->device_bytes);
    if (st) return st;
    md->stage_values = need;
  }
  return NSM_OK;
}

// device map and staging array for node fields of width up to w (tiled models)
int ensure_node_staging(nsm_model* md, size_t w)
{
  if (!md->d_node_new) {
    int st = upload(md, md->tiles.node_new, &md->d_node_new);
    if (st) return st;
  }
  return ensure_stage(md, w * (size_t)md->N);
}

// Element data (F or sigma, nc components) of every element into `out` in
// the reference's [element][point][component] order: gathered on the device
// into the staging array a chunk of elements at a time, then copied.
constexpr size_t kIpStageValues = (size_t)1 << 23;  // 64 MB of FP64 per chunk

int ip_field_out(nsm_model* md, const void* src, int nc, double* out)
{
  const int64_t E = md->E;
  if (E == 0) return NSM_OK;
  nsm::IpLayout L{md->precision == NSM_FP64 ? 0 : (md->tiled ? 2 : 1), nc, E, nullptr, nullptr, 0};
  if (md->tiled) {
    if (!md->d_elem_pos) {
      std::vector<int32_t> pos((size_t)E);
      for (int64_t d = 0; d < E; d++) pos[(size_t)md->tiles.perm[(size_t)d]] = (int32_t)d;
      int st = upload(md, pos, &md->d_elem_pos);
      if (st) return st;
    }
    L.dev_elem  = md->d_elem_pos;
    L.elem_off  = md->tile_args.elem_off;
    L.num_tiles = md->tiles.num_tiles;
  }
  const int64_t per   = 8 * (int64_t)nc;
  const int64_t chunk = std::min<int64_t>(E, std::max<int64_t>(1, (int64_t)kIpStageValues / per));
  int           st    = ensure_stage(md, (size_t)(chunk * per));
  if (st) return st;
  for (int64_t e0 = 0; e0 < E; e0 += chunk) {
    const int64_t n = std::min(chunk, E - e0);
    nsm::launch_ip_gather(src, L, e0, n, md->d_stage, md->stream);
    NSM_LAUNCHED();
    NSM_HIP(hipMemcpyAsync(out + e0 * per, md->d_stage, (size_t)(n * per) * sizeof(double), hipMemcpyDeviceToHost,
                           md->stream));
    NSM_HIP(hipStreamSynchronize(md->stream));
  }
  return NSM_OK;
}

// device node array -> host array in mesh node order; a tiled model's rows
// are permuted on the device into the staging array first
int node_field_out(nsm_model* md, const double* src, size_t w, double* out)
{
  const size_t cnt = w * (size_t)md->N;
  if (md->tiled) {
    int st = ensure_node_staging(md, w);
    if (st) return st;
    nsm::launch_permute_nodes(src, md->d_node_new, md->N, (int)w, false, md->d_stage, md->stream);
    NSM_LAUNCHED();
    src = md->d_stage;
  }
  NSM_HIP(hipMemcpyAsync(out, src, cnt * sizeof(double), hipMemcpyDeviceToHost, md->stream));
  NSM_HIP(hipStreamSynchronize(md->stream));
  return NSM_OK;
}

// host array in mesh node order -> device node array (through the staging
// array and a device permutation for a tiled model)
int node_field_in(nsm_model* md, double* dst, size_t w, const double* in)
{
  const size_t cnt = w * (size_t)md->N;
  if (!md->tiled) {
    NSM_HIP(hipMemcpyAsync(dst, in, cnt * sizeof(double), hipMemcpyHostToDevice, md->stream));
    NSM_HIP(hipStreamSynchronize(md->stream));
    return NSM_OK;
  }
  int st = ensure_node_staging(md, w);
  if (st) return st;
  NSM_HIP(hipMemcpyAsync(md->d_stage, in, cnt * sizeof(double), hipMemcpyHostToDevice, md->stream));
  nsm::launch_permute_nodes(md->d_stage, md->d_node_new, md->N, (int)w, true, dst, md->stream);
  NSM_LAUNCHED();
  NSM_HIP(hipStreamSynchronize(md->stream));
  return NSM_OK;
}

}  // namespace

extern "C" {

const char* nsm_status_string(int status)
{
  switch (status) {
    case NSM_OK: return "ok";
    case NSM_ERR_ARG: return "invalid argument";
    case NSM_ERR_HIP: return "HIP runtime error";
    case NSM_ERR_NEG_JACOBIAN: return "non-positive Jacobian";
    case NSM_ERR_STATE: return "invalid call order";
    case NSM_ERR_NOMEM: return "out of device memory";
    default: return "unknown status";
  }
}

const char* nsm_last_error(void) { return g_last_error.c_str(); }

const char* nsm_version(void) { return "nimblesm_amd 0.1 gfx950"; }

// ------------------------------------------------------------------- seams
#define NSM_SEAM_PROLOGUE(n)                                                          \
  if ((n) < 0) return fail(NSM_ERR_ARG, "negative size");                             \
  if ((n) == 0) return NSM_OK;                                                        \
  int* status = nullptr;                                                              \
  {                                                                                   \
    int _st = seam_status_buffer(&status);                                            \
    if (_st) return _st;                                                              \
  }                                                                                   \
  hipStream_t s = (hipStream_t)stream;                                                \
  NSM_HIP(hipMemsetAsync(status, 0, sizeof(int), s));

static int seam_epilogue(int* status, hipStream_t s)
{
  NSM_LAUNCHED();
  int h = 0;
  NSM_HIP(hipMemcpyAsync(&h, status, sizeof(int), hipMemcpyDeviceToHost, s));
  NSM_HIP(hipStreamSynchronize(s));
  if (h) return fail(NSM_ERR_NEG_JACOBIAN, "non-positive Jacobian determinant (Invert3x3)");
  return NSM_OK;
}

int nsm_hex_compute_deformation_gradients(int64_t num_elem, const double* X_gathered, const double* u_gathered,
                                          double* F, void* stream)
{
  if (num_elem > 0 && (!X_gathered || !u_gathered || !F)) return fail(NSM_ERR_ARG, "null pointer");
  NSM_SEAM_PROLOGUE(num_elem);
  nsm::launch_seam_defgrad(num_elem, X_gathered, u_gathered, F, status, s);
  return seam_epilogue(status, s);
}

int nsm_hex_compute_nodal_forces(int64_t num_elem, const double* X_gathered, const double* u_gathered,
                                 const double* sigma, double* f, void* stream)
{
  if (num_elem > 0 && (!X_gathered || !u_gathered || !sigma || !f)) return fail(NSM_ERR_ARG, "null pointer");
  NSM_SEAM_PROLOGUE(num_elem);
  nsm::launch_seam_forces(num_elem, X_gathered, u_gathered, sigma, f, status, s);
  return seam_epilogue(status, s);
}

int nsm_hex_compute_lumped_mass(int64_t num_elem, double density, const double* X_gathered, double* m, void* stream)
{
  if (num_elem > 0 && (!X_gathered || !m)) return fail(NSM_ERR_ARG, "null pointer");
  NSM_SEAM_PROLOGUE(num_elem);
  nsm::launch_seam_lumped_mass(num_elem, density, X_gathered, m, status, s);
  return seam_epilogue(status, s);
}

int nsm_hex_compute_characteristic_length(int64_t num_elem, const double* x_gathered, double* out, void* stream)
{
  if (num_elem > 0 && (!x_gathered || !out)) return fail(NSM_ERR_ARG, "null pointer");
  if (num_elem < 0) return fail(NSM_ERR_ARG, "negative size");
  if (num_elem == 0) return NSM_OK;
  nsm::launch_seam_char_length(num_elem, x_gathered, out, (hipStream_t)stream);
  NSM_LAUNCHED();
  return NSM_OK;
}

int nsm_neohookean_get_stress(int64_t num_points, double bulk_modulus, double shear_modulus, const double* F_np1,
                              double* sigma_np1, void* stream)
{
  if (num_points < 0) return fail(NSM_ERR_ARG, "negative size");
  if (num_points == 0) return NSM_OK;
  if (!F_np1 || !sigma_np1) return fail(NSM_ERR_ARG, "null pointer");
  nsm::launch_seam_stress(num_points, bulk_modulus, shear_modulus, F_np1, sigma_np1, (hipStream_t)stream);
  NSM_LAUNCHED();
  return NSM_OK;
}

// ------------------------------------------------------------------- model
}  // extern "C"

namespace {

int model_create(int64_t num_nodes, const double* reference_coordinates, int num_blocks, const int* block_ids,
                 const int64_t* block_num_elem, const int* conn, const double* density, const double* bulk_modulus,
                 const double* shear_modulus, int precision, int tile_override, nsm_model** out);

// Whether an FP64 model of E elements reads its reference inverse Jacobians
// from a table formed once (80 doubles per element) instead of forming them
// in every element launch: NSM_FP64_REF_TABLE=0/1 when set, else when the
// table takes at most kRefTableBytes.  Reading the table saves 13% of the
// element kernel's instructions, but it is a further 640 bytes per element
// and launch: it pays while the table stays in the device's caches between
// launches, and costs time once it streams from HBM.
constexpr int64_t kRefTableBytes = 96ll << 20;
bool use_reference_table(int64_t E)
{
  const char* v = std::getenv("NSM_FP64_REF_TABLE");
  if (v && *v) return std::strcmp(v, "0") != 0;
  return E > 0 && 640 * E <= kRefTableBytes;
}

// Element-tile size of a mixed model: NSM_TILE_ELEMS when it names one;
// 512 when that already gives every CU two tiles or more; otherwise the
// autotuner's choice from a few timed steps of a model built with each
// candidate (cached per mesh shape).
int model_tile_elems(int64_t num_nodes, const double* X, int num_blocks, const int* block_ids,
                     const int64_t* block_num_elem, const int* conn, const double* density, const double* K,
                     const double* G, int64_t E)
{
  const char* te = std::getenv("NSM_TILE_ELEMS");
  if (te && std::strcmp(te, "auto") != 0) {
    for (int c : nsm::kTileCandidates)
      if (std::atoi(te) == c) return c;
    return nsm::kTileElems;
  }
  int dev = 0, cus = 256;
  if (hipGetDevice(&dev) == hipSuccess)
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
  // 512-element tiles when every CU gets at least 8 of them (a 128^3 cube
  // gets 16 on 256 CUs); with 2 per CU (64^3) 256-element tiles were 7% faster
  constexpr int64_t kFixedTilesPerCu = 8;
  if (E >= kFixedTilesPerCu * cus * nsm::kTileElems) return nsm::kTileElems;
  char key[160];
  snprintf(key, sizeof(key), "model:mixed:N%lld:E%lld:B%d", (long long)num_nodes, (long long)E, num_blocks);
  auto time_of = [&](int c) -> double {
    nsm_model* tmp = nullptr;
    if (model_create(num_nodes, X, num_blocks, block_ids, block_num_elem, conn, density, K, G, NSM_MIXED, c, &tmp))
      return -1.0;
    double dt = 0.0, t = 0.0, sec = -1.0;
    if (tmp->tiled && !nsm_model_compute_lumped_mass(tmp) && !nsm_model_get_critical_time_step(tmp, &dt) &&
        !nsm_model_explicit_steps(tmp, 0, 2, 8, 0, 0.5 * dt, &t)) {
      const auto t0 = std::chrono::steady_clock::now();
      if (!nsm_model_explicit_steps(tmp, 2, 8, 8, 0, 0.5 * dt, &t))
        sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 6.0;
    }
    nsm_model_destroy(tmp);
    return sec;
  };
  const int c = nsm::autotune_tiles(key, time_of, nullptr);
  return c ? c : nsm::kTileElems;
}

}  // namespace

extern "C" {

int nsm_model_create(int64_t num_nodes, const double* reference_coordinates, int num_blocks, const int* block_ids,
                     const int64_t* block_num_elem, const int* conn, const double* density,
                     const double* bulk_modulus, const double* shear_modulus, int precision, nsm_model** out)
{
  return model_create(num_nodes, reference_coordinates, num_blocks, block_ids, block_num_elem, conn, density,
                      bulk_modulus, shear_modulus, precision, 0, out);
}

int nsm_model_tile_elems(nsm_model* md, int* tile_elems)
{
  if (!md || !tile_elems) return fail(NSM_ERR_ARG, "null argument");
  *tile_elems = md->tiled ? md->tiles.tile_elems : 0;
  return NSM_OK;
}

}  // extern "C"

namespace {

int model_create(int64_t num_nodes, const double* reference_coordinates, int num_blocks, const int* block_ids,
                 const int64_t* block_num_elem, const int* conn, const double* density, const double* bulk_modulus,
                 const double* shear_modulus, int precision, int tile_override, nsm_model** out)
{
  if (!out) return fail(NSM_ERR_ARG, "out is null");
  *out = nullptr;
  if (num_nodes <= 0 || !reference_coordinates) return fail(NSM_ERR_ARG, "need at least one node");
  if (num_blocks <= 0 || !block_ids || !block_num_elem || !density || !bulk_modulus || !shear_modulus)
    return fail(NSM_ERR_ARG, "need at least one block with material parameters");
  if (precision != NSM_FP64 && precision != NSM_MIXED) return fail(NSM_ERR_ARG, "unknown precision %d", precision);
  if (num_nodes > (int64_t)INT32_MAX) return fail(NSM_ERR_ARG, "node ids are int32");

  std::vector<int64_t> off(1, 0);
  for (int b = 0; b < num_blocks; b++) {
    if (block_num_elem[b] < 0) return fail(NSM_ERR_ARG, "negative element count in block %d", b);
    if (!(density[b] > 0.0) || !(bulk_modulus[b] > 0.0) || !(shear_modulus[b] >= 0.0))
      return fail(NSM_ERR_ARG, "invalid material parameters in block %d", b);
    off.push_back(off.back() + block_num_elem[b]);
  }
  const int64_t E = off.back();
  if (E > 0 && !conn) return fail(NSM_ERR_ARG, "conn is null");
  if (8 * nsm::round_up8(E) >= (int64_t)UINT32_MAX) return fail(NSM_ERR_ARG, "too many elements for 32-bit CSR entries");
  for (int64_t k = 0; k < 8 * E; k++)
    if (conn[k] < 0 || conn[k] >= num_nodes)
      return fail(NSM_ERR_ARG, "connectivity entry %lld = %d out of range", (long long)k, conn[k]);

  nsm_model* md  = new nsm_model();
  md->precision  = precision;
  md->N          = num_nodes;
  md->E          = E;
  md->num_blocks = num_blocks;
  md->block_ids.assign(block_ids, block_ids + num_blocks);
  md->blk_off = off;
  md->rho.assign(density, density + num_blocks);
  md->K.assign(bulk_modulus, bulk_modulus + num_blocks);
  md->G.assign(shear_modulus, shear_modulus + num_blocks);

  auto cleanup = [&](int st) {
    nsm_model_destroy(md);
    return st;
  };
  if (hipGetDevice(&md->device) != hipSuccess) return cleanup(fail(NSM_ERR_HIP, "no HIP device"));
  if (hipStreamCreateWithFlags(&md->stream, hipStreamNonBlocking) != hipSuccess)
    return cleanup(fail(NSM_ERR_HIP, "hipStreamCreate failed"));

  // Mixed precision assembles forces tile-locally unless the mesh cannot be
  // tiled (an element tile needing more than kMaxColors colours) or
  // NSM_MIXED_ASSEMBLY=slots asks for the element-slot + node-gather path.
  if (precision == NSM_MIXED && E > 0) {
    const char* mode = std::getenv("NSM_MIXED_ASSEMBLY");
    if (!(mode && std::strcmp(mode, "slots") == 0)) {
      // 64 to 512 elements per tile (one to eight waves per workgroup):
      // NSM_TILE_ELEMS, else 512 on large meshes, else the autotuner's choice
      const int tile_elems = tile_override > 0 ? tile_override
                                               : model_tile_elems(num_nodes, reference_coordinates, num_blocks,
                                                                  block_ids, block_num_elem, conn, density,
                                                                  bulk_modulus, shear_modulus, E);
      std::string err;
      md->tiled = nsm::build_tiles(num_nodes, reference_coordinates, conn, off, tile_elems, md->tiles, &err);
      if (!md->tiled) md->tiles = nsm::TileSet();
      // NSM_TILE_ASSEMBLY=colors: colour-ordered assembly in every tile
      // (otherwise tiles that allow it gather per corner)
      const char* ta = std::getenv("NSM_TILE_ASSEMBLY");
      if (ta && std::strcmp(ta, "colors") == 0) std::fill(md->tiles.gather.begin(), md->tiles.gather.end(), 0);
    }
  }

  // Device node numbering: the mesh's, or the tiled model's (tile-interior
  // nodes contiguous per tile, then the nodes shared between tiles).
  std::vector<double> X_dev;
  std::vector<int>    conn_dev;
  const double*       Xd = reference_coordinates;
  const int*          cd = conn;
  if (md->tiled) {
    const auto& old_of = md->tiles.node_old;
    const auto& new_of = md->tiles.node_new;
    X_dev.resize(3 * (size_t)num_nodes);
    for (int64_t i = 0; i < num_nodes; i++)
      for (int c = 0; c < 3; c++) X_dev[(size_t)(3 * i + c)] = reference_coordinates[3 * (int64_t)old_of[(size_t)i] + c];
    conn_dev.resize((size_t)(8 * E));
    for (int64_t k = 0; k < 8 * E; k++) conn_dev[(size_t)k] = new_of[(size_t)conn[k]];
    Xd = X_dev.data();
    cd = conn_dev.data();
  }

  // node -> (element e, local node j) incidence in ascending element order,
  // stored as the slot index of (e, j): fp64_slot(e, j) in grouped models,
  // else j*E + e (the element-fastest FP32 slots of the untiled mixed kernel)
  std::vector<uint32_t> ptr((size_t)num_nodes + 1, 0), ent((size_t)(8 * E));
  for (int64_t k = 0; k < 8 * E; k++) ptr[(size_t)cd[k] + 1]++;
  for (int64_t i = 0; i < num_nodes; i++) ptr[(size_t)i + 1] += ptr[(size_t)i];
  {
    std::vector<uint32_t> fill(ptr.begin(), ptr.end() - 1);
    const bool            grouped = md->grouped();
    for (int64_t k = 0; k < 8 * E; k++)
      ent[fill[(size_t)cd[k]]++] =
          (uint32_t)(grouped ? nsm::fp64_slot(k >> 3, (int)(k & 7)) : (k & 7) * E + (k >> 3));
  }

  const size_t N3 = 3 * (size_t)num_nodes;
  const size_t ipw = precision == NSM_FP64 ? sizeof(double) : sizeof(float);
  size_t&      tot = md->device_bytes;
  int          st  = 0;
  // FP64 element data in groups of 8 elements (fp64_ip_index, fp64_slot)
  const size_t E_ip         = precision == NSM_FP64 ? (size_t)nsm::round_up8(E) : (size_t)E;
  const size_t num_slots    = md->tiled ? 0 : 24 * E_ip;
  const size_t num_partials = md->tiled ? (size_t)md->tiles.num_tiles +
                                              nsm::boundary_grid((int)md->tiles.bnd_nodes.size())
                                        : (size_t)nsm::node_grid(num_nodes);
  md->reaction_stride = (int64_t)std::max<size_t>(num_partials, 1);
  if ((st = dalloc(&md->X, N3, &tot)) || (st = dalloc(&md->u, N3, &tot)) || (st = dalloc(&md->v, N3, &tot)) ||
      (st = dalloc(&md->a, N3, &tot)) || (st = dalloc(&md->f, N3, &tot)) ||
      (st = dalloc(&md->m, (size_t)num_nodes, &tot)) || (st = dalloc(&md->conn, (size_t)(8 * E), &tot)) ||
      (st = dalloc(&md->node_ptr, ptr.size(), &tot)) || (st = dalloc(&md->node_ent, ent.size(), &tot)) ||
      (st = dalloc((char**)&md->Fip, 72 * E_ip * ipw, &tot)) ||
      (st = dalloc((char**)&md->sip, 48 * E_ip * ipw, &tot)) ||
      (st = dalloc((char**)&md->slots, num_slots * ipw, &tot)) ||
      (st = dalloc(&md->rjac, precision == NSM_FP64 && use_reference_table(E) ? (size_t)(80 * E) : 0, &tot)) ||
      (st = dalloc(&md->status, 1, &tot)) || (st = dalloc(&md->flags, (size_t)num_nodes, &tot)) ||
      (st = dalloc(&md->reaction_partials, 2 * std::max<size_t>(num_partials, 1), &tot)) ||
      (md->tiled && (st = setup_tiles(md))))
    return cleanup(st);

  hipStream_t s = md->stream;
#define NSM_HIP_C(call)                                                                            \
  do {                                                                                             \
    hipError_t _e = (call);                                                                        \
    if (_e != hipSuccess) return cleanup(fail(NSM_ERR_HIP, "%s: %s", #call, hipGetErrorString(_e))); \
  } while (0)
  NSM_HIP_C(hipMemcpyAsync(md->X, Xd, N3 * sizeof(double), hipMemcpyHostToDevice, s));
  if (E > 0) {
    NSM_HIP_C(hipMemcpyAsync(md->conn, cd, (size_t)(8 * E) * sizeof(int), hipMemcpyHostToDevice, s));
    NSM_HIP_C(hipMemcpyAsync(md->node_ent, ent.data(), ent.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  }
  NSM_HIP_C(hipMemcpyAsync(md->node_ptr, ptr.data(), ptr.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
  NSM_HIP_C(hipMemsetAsync(md->u, 0, N3 * sizeof(double), s));
  NSM_HIP_C(hipMemsetAsync(md->v, 0, N3 * sizeof(double), s));
  NSM_HIP_C(hipMemsetAsync(md->a, 0, N3 * sizeof(double), s));
  NSM_HIP_C(hipMemsetAsync(md->f, 0, N3 * sizeof(double), s));
  NSM_HIP_C(hipMemsetAsync(md->m, 0, (size_t)num_nodes * sizeof(double), s));
  NSM_HIP_C(hipMemsetAsync(md->status, 0, sizeof(int), s));
  NSM_HIP_C(hipMemsetAsync(md->flags, 0, (size_t)num_nodes, s));
  if (md->rjac) {
    nsm::launch_reference_jacobians(E, md->conn, md->X, md->rjac, s);
    NSM_HIP_C(hipGetLastError());
  }
  if (E > 0) {
    // STEP_NP1 deformation gradient starts at the identity, stress at zero
    // (src/nimble_kokkos_model_data.cc:921-924); the diagonal components XX,
    // YY, ZZ are components 0-2
    std::vector<char> ident(72 * E_ip * ipw, 0);
    for_each_ip(md, 9, [&](int64_t p, int64_t, int c, int) {
      if (c >= 3) return;
      if (precision == NSM_FP64)
        ((double*)ident.data())[p] = 1.0;
      else
        ((float*)ident.data())[p] = 1.0f;
    });
    NSM_HIP_C(hipMemcpy(md->Fip, ident.data(), ident.size(), hipMemcpyHostToDevice));
    NSM_HIP_C(hipMemsetAsync(md->sip, 0, 48 * E_ip * ipw, s));
  }
  NSM_HIP_C(hipStreamSynchronize(s));
#undef NSM_HIP_C
  *out = md;
  return NSM_OK;
}

}  // namespace

extern "C" {

int nsm_model_destroy(nsm_model* md)
{
  if (!md) return NSM_OK;
  (void)hipSetDevice(md->device);
  if (md->stream) (void)hipStreamSynchronize(md->stream);
  free_dev(md->X);
  free_dev(md->u);
  free_dev(md->v);
  free_dev(md->a);
  free_dev(md->f);
  free_dev(md->m);
  free_dev(md->conn);
  free_dev(md->node_ptr);
  free_dev(md->node_ent);
  free_dev(md->Fip);
  free_dev(md->sip);
  free_dev(md->slots);
  free_dev(md->rjac);
  free_dev(md->elem_tmp);
  free_dev(md->wg_min);
  free_dev(md->status);
  free_dev(md->flags);
  free_dev(md->bc_table);
  free_dev(md->reaction_hist);
  free_dev(md->reaction_partials);
  free_dev(md->reaction_tiles);
  free_dev(md->uq_state);
  free_dev(md->uq_coeff);
  free_dev(md->uq_bc_table);
  free_dev(md->d_stage);
  for (void* p : md->tile_bufs) (void)hipFree(p);
  for (auto& p : md->event_pool) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  if (md->stream) (void)hipStreamDestroy(md->stream);
  delete md;
  return NSM_OK;
}

int nsm_model_add_boundary_condition(nsm_model* md, int kind, int coordinate, double magnitude, int64_t num_nodes,
                                     const int* node_ids)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  if (kind < NSM_BC_INITIAL_VELOCITY || kind > NSM_BC_PRESCRIBED_DISPLACEMENT)
    return fail(NSM_ERR_ARG, "unknown boundary-condition kind %d", kind);
  if (coordinate < 0 || coordinate > 2) return fail(NSM_ERR_ARG, "coordinate must be 0, 1 or 2");
  if (num_nodes < 0 || (num_nodes > 0 && !node_ids)) return fail(NSM_ERR_ARG, "invalid node list");
  BcRecord r;
  r.kind  = kind;
  r.coord = coordinate;
  r.value = magnitude;
  r.nodes.assign(node_ids, node_ids + num_nodes);
  for (int n : r.nodes)
    if (n < 0 || n >= md->N) return fail(NSM_ERR_ARG, "node id %d out of range", n);
  md->bcs.push_back(std::move(r));
  md->bc_dirty = true;
  return NSM_OK;
}

int nsm_model_apply_initial_conditions(nsm_model* md)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  int st = ensure_device(md);
  if (st) return st;
  // the nominal velocity and every UQ sample's
  // (src/nimble_boundary_condition_manager.h:93-134)
  std::vector<double> v((size_t)(3 * md->N));
  for (int s = -1; s < md->uq_samples(); s++) {
    double* dv = md->node_arrays(s).v;
    NSM_HIP(hipMemcpyAsync(v.data(), dv, v.size() * sizeof(double), hipMemcpyDeviceToHost, md->stream));
    NSM_HIP(hipStreamSynchronize(md->stream));
    for (const auto& b : md->bcs)
      if (b.kind == NSM_BC_INITIAL_VELOCITY)
        for (int n : b.nodes) v[(size_t)(3 * dev_node(md, n) + b.coord)] = b.value;
    NSM_HIP(hipMemcpyAsync(dv, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice, md->stream));
    NSM_HIP(hipStreamSynchronize(md->stream));
  }
  return NSM_OK;
}

int nsm_model_apply_kinematic_conditions(nsm_model* md, double time_current, double time_previous)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  int st = ensure_device(md);
  if (st || (st = rebuild_bcs(md))) return st;
  nsm::StepTimes t{time_previous, time_current, time_current - time_previous, 0.0};
  for (int s = -1; s < md->uq_samples(); s++) {
    nsm::launch_apply_bc(md->node_arrays(s), t, md->stream);
    NSM_LAUNCHED();
  }
  return NSM_OK;
}

int nsm_model_compute_lumped_mass(nsm_model* md)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  int st = ensure_device(md);
  if (st) return st;
  size_t tmp = 0;
  if (!md->elem_tmp && (st = dalloc(&md->elem_tmp, std::max<size_t>((size_t)(8 * nsm::round_up8(md->E)), 1), &tmp)))
    return st;
  md->device_bytes += tmp;
  for (int b = 0; b < md->num_blocks; b++) {
    nsm::launch_element_mass(md->blk_off[b], md->blk_off[b + 1], md->E, md->grouped(), md->conn, md->X, md->rho[b],
                             md->elem_tmp, md->status, md->stream);
    NSM_LAUNCHED();
  }
  nsm::launch_mass_gather(md->N, md->node_ptr, md->node_ent, md->elem_tmp, md->m, md->stream);
  NSM_LAUNCHED();

  // critical time step, minimum over blocks of min_e charlen_e / sqrt(K/rho)
  int64_t need = 0;
  for (int b = 0; b < md->num_blocks; b++)
    need = std::max<int64_t>(need, (md->blk_off[b + 1] - md->blk_off[b] + 255) / 256 + 1);
  if (need > md->wg_min_len) {
    free_dev(md->wg_min);
    tmp = 0;
    if ((st = dalloc(&md->wg_min, (size_t)need, &tmp))) return st;
    md->wg_min_len = need;
  }
  double dt = 1.7976931348623157e308;
  for (int b = 0; b < md->num_blocks; b++) {
    const double c  = std::sqrt(md->K[b] / md->rho[b]);
    const int    ng = nsm::launch_element_dt(md->blk_off[b], md->blk_off[b + 1], md->conn, md->X, md->u, c,
                                             md->wg_min, md->stream);
    NSM_LAUNCHED();
    if (ng == 0) continue;
    nsm::launch_min_reduce(md->wg_min, ng, md->wg_min + md->wg_min_len - 1, md->stream);
    NSM_LAUNCHED();
    double bdt = 0.0;
    NSM_HIP(hipMemcpyAsync(&bdt, md->wg_min + md->wg_min_len - 1, sizeof(double), hipMemcpyDeviceToHost,
                           md->stream));
    NSM_HIP(hipStreamSynchronize(md->stream));
    if (bdt < dt) dt = bdt;
  }
  md->critical_dt = dt;
  if ((st = check_status(md))) return st;
  md->have_mass = true;
  return NSM_OK;
}

int nsm_model_get_critical_time_step(nsm_model* md, double* dt)
{
  if (!md || !dt) return fail(NSM_ERR_ARG, "null argument");
  if (!md->have_mass) return fail(NSM_ERR_STATE, "call nsm_model_compute_lumped_mass first");
  *dt = md->critical_dt;
  return NSM_OK;
}

int nsm_model_compute_internal_force(nsm_model* md, double time_previous, double time_current, int is_output_step)
{
  (void)time_previous;
  (void)time_current;
  (void)is_output_step;
  if (!md) return fail(NSM_ERR_ARG, "null model");
  int st = ensure_device(md);
  if (st) return st;
  const nsm::StepTimes none{0.0, 0.0, 0.0, 0.0};
  const nsm::StepArgs  sa{none, none, 0, 0, 0, 0, 1};
  int64_t              pending = -1;
  if ((st = launch_step(md, sa, -1, -1, &pending))) return st;
  return check_status(md);
}

// The split form of ComputeInternalForce that lets a block material
// interface compute the stress (src/nimble_kokkos_model_data.cc:1196-1257):
// F_np1 first, then the plug-in's stress into the model's sigma_np1, then
// the forces from it.  FP64 models; the nominal trajectory only.
static int check_split(nsm_model* md)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  if (md->precision != NSM_FP64)
    return fail(NSM_ERR_STATE, "the split force evaluation needs an FP64 model (mixed precision evaluates the "
                               "stress inside its element kernel)");
  if (md->uq_samples() > 0) return fail(NSM_ERR_STATE, "the split force evaluation does not advance UQ samples");
  return ensure_device(md);
}

int nsm_model_compute_deformation_gradients(nsm_model* md)
{
  int st = check_split(md);
  if (st || (st = launch_elements(md, -1, nsm::kElementDefGrad))) return st;
  return check_status(md);
}

int nsm_model_compute_internal_force_from_stress(nsm_model* md)
{
  int st = check_split(md);
  if (st || (st = launch_elements(md, -1, nsm::kElementForcesFromStress))) return st;
  const nsm::StepTimes none{0.0, 0.0, 0.0, 0.0};
  const nsm::StepArgs  sa{none, none, 0, 0, 0, 0, 1};
  int64_t              pending = -1;
  if ((st = launch_node(md, sa, -1, -1, &pending, -1))) return st;
  return check_status(md);
}

int nsm_neohookean_get_stress_strided(int64_t num_elem, int num_ipt, double bulk_modulus, double shear_modulus,
                                      const double* F_np1, int64_t F_elem_stride, int64_t F_ipt_stride,
                                      int64_t F_comp_stride, double* sigma_np1, int64_t sigma_elem_stride,
                                      int64_t sigma_ipt_stride, int64_t sigma_comp_stride, void* stream)
{
  if (num_elem < 0 || num_ipt < 0) return fail(NSM_ERR_ARG, "negative size");
  if (num_elem == 0 || num_ipt == 0) return NSM_OK;
  if (!F_np1 || !sigma_np1) return fail(NSM_ERR_ARG, "null pointer");
  // plain strides: a group of 8 elements is 8 element strides
  nsm::launch_seam_stress_strided(0, num_elem, num_ipt, bulk_modulus, shear_modulus, F_np1,
                                  nsm::IptStrides{8 * F_elem_stride, F_elem_stride, F_ipt_stride, F_comp_stride},
                                  sigma_np1,
                                  nsm::IptStrides{8 * sigma_elem_stride, sigma_elem_stride, sigma_ipt_stride,
                                                  sigma_comp_stride},
                                  (hipStream_t)stream);
  NSM_LAUNCHED();
  return NSM_OK;
}

int nsm_neohookean_get_stress_grouped(int64_t first_elem, int64_t num_elem, int num_ipt, double bulk_modulus,
                                      double shear_modulus, const double* F_np1, const int64_t* F_strides,
                                      double* sigma_np1, const int64_t* sigma_strides, void* stream)
{
  if (first_elem < 0 || num_elem < 0 || num_ipt < 0) return fail(NSM_ERR_ARG, "negative size");
  if (num_elem == 0 || num_ipt == 0) return NSM_OK;
  if (!F_np1 || !sigma_np1 || !F_strides || !sigma_strides) return fail(NSM_ERR_ARG, "null pointer");
  nsm::launch_seam_stress_strided(first_elem, num_elem, num_ipt, bulk_modulus, shear_modulus, F_np1,
                                  nsm::IptStrides{F_strides[0], F_strides[1], F_strides[2], F_strides[3]}, sigma_np1,
                                  nsm::IptStrides{sigma_strides[0], sigma_strides[1], sigma_strides[2],
                                                  sigma_strides[3]},
                                  (hipStream_t)stream);
  NSM_LAUNCHED();
  return NSM_OK;
}

int nsm_model_get_stream(nsm_model* md, void** stream)
{
  if (!md || !stream) return fail(NSM_ERR_ARG, "null argument");
  *stream = (void*)md->stream;
  return NSM_OK;
}

int nsm_model_update_states(nsm_model* md)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  return NSM_OK;
}

int nsm_model_set_reaction_nodes(nsm_model* md, int64_t num_nodes, const int* node_ids, int coordinate)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  if (coordinate < 0 || coordinate > 2) return fail(NSM_ERR_ARG, "coordinate must be 0, 1 or 2");
  if (num_nodes < 0 || (num_nodes > 0 && !node_ids)) return fail(NSM_ERR_ARG, "invalid node list");
  for (int64_t k = 0; k < num_nodes; k++)
    if (node_ids[k] < 0 || node_ids[k] >= md->N) return fail(NSM_ERR_ARG, "node id %d out of range", node_ids[k]);
  // the set is kept as one flag bit per node (the node kernel sums flagged
  // nodes); a node listed twice counts once
  md->reaction_mask.assign((size_t)md->N, 0);
  for (int64_t k = 0; k < num_nodes; k++) md->reaction_mask[(size_t)node_ids[k]] = 1;
  md->num_reaction_nodes = num_nodes;
  md->reaction_coord     = coordinate;
  md->bc_dirty           = true;
  return NSM_OK;
}

int nsm_model_get_reaction_history(nsm_model* md, int64_t first_step, int64_t count, double* out)
{
  if (!md || (count > 0 && !out)) return fail(NSM_ERR_ARG, "null argument");
  if (first_step < 0 || count < 0 || first_step + count > md->reaction_cap)
    return fail(NSM_ERR_ARG, "reaction history range [%lld, %lld) not recorded", (long long)first_step,
                (long long)(first_step + count));
  if (count == 0) return NSM_OK;
  int st = ensure_device(md);
  if (st) return st;
  NSM_HIP(hipMemcpyAsync(out, md->reaction_hist + first_step, (size_t)count * sizeof(double), hipMemcpyDeviceToHost,
                         md->stream));
  NSM_HIP(hipStreamSynchronize(md->stream));
  return NSM_OK;
}

int nsm_model_sample_reaction_history(nsm_model* md, int64_t num_steps, int num_samples, double* out)
{
  if (!md || (num_samples > 0 && !out)) return fail(NSM_ERR_ARG, "null argument");
  if (num_steps < 1 || num_steps > md->reaction_cap)
    return fail(NSM_ERR_ARG, "reaction history of %lld steps not recorded", (long long)num_steps);
  if (num_samples < 0) return fail(NSM_ERR_ARG, "negative sample count");
  if (num_samples == 0) return NSM_OK;
  int st = ensure_device(md);
  if (st) return st;
  double* d_out = nullptr;
  size_t  tmp   = 0;
  if ((st = dalloc(&d_out, (size_t)num_samples, &tmp))) return st;
  nsm::launch_sample_history(md->reaction_hist, num_steps, num_samples, d_out, md->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)num_samples * sizeof(double), hipMemcpyDeviceToHost,
                                          md->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(md->stream);
  free_dev(d_out);
  if (e != hipSuccess) return fail(NSM_ERR_HIP, "sampling the reaction history: %s", hipGetErrorString(e));
  return NSM_OK;
}

}  // extern "C"

namespace {

// Staging of nsm_model_sample_reaction_history_noalloc: the sampled values
// on their way to the host.  Every GPU the library runs on holds its own copy
// (NSM_NOALLOC_MAX_SAMPLES doubles, 32 KiB) until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: the
// device address comes from hipGetSymbolAddress, once per device.
__device__ __align__(256) double scratch[NSM_NOALLOC_MAX_SAMPLES];

std::mutex            g_scratch_mutex;
std::map<int, double*> g_scratch_ptr;  // device -> address of that device's scratch

// Device address of `scratch` on the current device (looked up on the first
// call on that device, with it current).
int scratch_address(double** out)
{
  int dev = 0;
  NSM_HIP(hipGetDevice(&dev));
  std::lock_guard<std::mutex> lk(g_scratch_mutex);
  auto                        it = g_scratch_ptr.find(dev);
  if (it == g_scratch_ptr.end()) {
    double* ptr = nullptr;
    NSM_HIP(hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) return fail(NSM_ERR_HIP, "hipGetSymbolAddress returned no address for the sampling scratch");
    it = g_scratch_ptr.emplace(dev, ptr).first;
  }
  *out = it->second;
  return NSM_OK;
}

}  // namespace

extern "C" {

int nsm_model_sample_reaction_history_noalloc(nsm_model* md, int64_t num_steps, int num_samples, double* out)
{
  if (!md || (num_samples > 0 && !out)) return fail(NSM_ERR_ARG, "null argument");
  if (num_steps < 1 || num_steps > md->reaction_cap)
    return fail(NSM_ERR_ARG, "reaction history of %lld steps not recorded", (long long)num_steps);
  if (num_samples < 0) return fail(NSM_ERR_ARG, "negative sample count");
  if (num_samples > NSM_NOALLOC_MAX_SAMPLES)
    return fail(NSM_ERR_ARG, "%d samples exceed the allocation-free maximum of %d", num_samples,
                NSM_NOALLOC_MAX_SAMPLES);
  if (num_samples == 0) return NSM_OK;
  int st = ensure_device(md);
  if (st) return st;
  double* d_out = nullptr;
  if ((st = scratch_address(&d_out))) return st;
  // the kernel writes every one of the num_samples values: nothing to clear
  nsm::launch_sample_history(md->reaction_hist, num_steps, num_samples, d_out, md->stream);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, (size_t)num_samples * sizeof(double), hipMemcpyDeviceToHost,
                                          md->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(md->stream);
  if (e != hipSuccess) return fail(NSM_ERR_HIP, "sampling the reaction history: %s", hipGetErrorString(e));
  return NSM_OK;
}

int nsm_model_explicit_steps(nsm_model* md, int64_t step_begin, int64_t step_end, int64_t num_load_steps,
                             int64_t output_frequency, double dt_user, double* time_current)
{
  if (!md || !time_current) return fail(NSM_ERR_ARG, "null argument");
  if (step_begin < 0 || step_end < step_begin || step_end > num_load_steps)
    return fail(NSM_ERR_ARG, "invalid step range");
  if (!md->have_mass) return fail(NSM_ERR_STATE, "call nsm_model_compute_lumped_mass first");
  if (step_end == step_begin) return NSM_OK;
  int st = ensure_device(md);
  if (st || (st = rebuild_bcs(md))) return st;
  if (md->num_reaction_nodes > 0 && num_load_steps > md->reaction_cap) {
    double* h   = nullptr;
    size_t  tmp = 0;
    if ((st = dalloc(&h, (size_t)num_load_steps, &tmp))) return st;
    NSM_HIP(hipMemsetAsync(h, 0, (size_t)num_load_steps * sizeof(double), md->stream));
    if (md->reaction_hist && md->reaction_cap > 0)
      NSM_HIP(hipMemcpyAsync(h, md->reaction_hist, (size_t)md->reaction_cap * sizeof(double),
                             hipMemcpyDeviceToDevice, md->stream));
    NSM_HIP(hipStreamSynchronize(md->stream));
    free_dev(md->reaction_hist);
    md->reaction_hist = h;
    md->reaction_cap  = num_load_steps;
  }

  double         t   = *time_current;
  nsm::StepTimes cur = step_times(&t, dt_user);
  for (int s = -1; s < md->uq_samples(); s++) {
    nsm::launch_prekick(md->node_arrays(s), cur, md->stream);
    NSM_LAUNCHED();
  }
  int64_t pending = -1;  // step whose reaction total is not yet summed
  for (int64_t step = step_begin; step < step_end; step++) {
    const bool     last = step + 1 == step_end;
    double         tn   = t;
    nsm::StepTimes nxt  = cur;
    if (!last) nxt = step_times(&tn, dt_user);
    // the internal force is the step's entry-point output: stored every step
    const nsm::StepArgs sa{cur, nxt, (int)is_output(step, num_load_steps, output_frequency), (int)!last, (int)last, 1, 1};
    // the previous step's pending reaction total is summed inside this launch
    if ((st = launch_step(md, sa, step, pending, &pending))) return st;
    if (!last) {
      t   = tn;
      cur = nxt;
    }
  }
  if (pending >= 0 && (st = launch_reaction(md, pending))) return st;
  *time_current = t;
  return check_status(md);
}

int nsm_model_get_field(nsm_model* md, int field, double* out)
{
  if (!md || !out) return fail(NSM_ERR_ARG, "null argument");
  int st = ensure_device(md);
  if (st) return st;
  const size_t N3 = 3 * (size_t)md->N;
  const double* src = nullptr;
  size_t        cnt = 0;
  switch (field) {
    case NSM_FIELD_REFERENCE_COORDINATE: src = md->X, cnt = N3; break;
    case NSM_FIELD_DISPLACEMENT: src = md->u, cnt = N3; break;
    case NSM_FIELD_VELOCITY: src = md->v, cnt = N3; break;
    case NSM_FIELD_ACCELERATION: src = md->a, cnt = N3; break;
    case NSM_FIELD_INTERNAL_FORCE: src = md->f, cnt = N3; break;
    case NSM_FIELD_LUMPED_MASS: src = md->m, cnt = (size_t)md->N; break;
    case NSM_FIELD_DEFORMATION_GRADIENT:
    case NSM_FIELD_STRESS: {
      // device layout (for_each_ip) -> host [E][ipt][comp] (the reference's element-data order)
      const int   nc = field == NSM_FIELD_DEFORMATION_GRADIENT ? 9 : 6;
      const void* p  = field == NSM_FIELD_DEFORMATION_GRADIENT ? md->Fip : md->sip;
      return ip_field_out(md, p, nc, out);
    }
    default: return fail(NSM_ERR_ARG, "unknown field %d", field);
  }
  return node_field_out(md, src, cnt / (size_t)md->N, out);
}

int nsm_model_set_field(nsm_model* md, int field, const double* in)
{
  if (!md || !in) return fail(NSM_ERR_ARG, "null argument");
  int st = ensure_device(md);
  if (st) return st;
  const size_t N3  = 3 * (size_t)md->N;
  double*      dst = nullptr;
  size_t       cnt = N3;
  switch (field) {
    case NSM_FIELD_DISPLACEMENT: dst = md->u; break;
    case NSM_FIELD_VELOCITY: dst = md->v; break;
    case NSM_FIELD_ACCELERATION: dst = md->a; break;
    case NSM_FIELD_INTERNAL_FORCE: dst = md->f; break;
    case NSM_FIELD_LUMPED_MASS:
      dst           = md->m;
      cnt           = (size_t)md->N;
      md->have_mass = true;
      break;
    default: return fail(NSM_ERR_ARG, "field %d cannot be set", field);
  }
  return node_field_in(md, dst, cnt / (size_t)md->N, in);
}

int nsm_model_set_uq_samples(nsm_model* md, int num_exact, const double* bulk_modulus, const double* shear_modulus,
                             int num_approx, const double* coefficients)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  if (num_exact < 0 || num_approx < 0) return fail(NSM_ERR_ARG, "negative sample count");
  if (num_exact > 0 && (!bulk_modulus || !shear_modulus)) return fail(NSM_ERR_ARG, "null sample moduli");
  if (num_approx > 0 && !coefficients) return fail(NSM_ERR_ARG, "null closure coefficients");
  const size_t nk = (size_t)num_exact * (size_t)md->num_blocks;
  const size_t nc = (size_t)num_approx * (size_t)(num_exact + 1);
  for (size_t i = 0; i < nk; i++)
    if (!std::isfinite(bulk_modulus[i]) || !std::isfinite(shear_modulus[i]))
      return fail(NSM_ERR_ARG, "non-finite modulus for exact sample %zu", i / (size_t)md->num_blocks);
  for (size_t i = 0; i < nc; i++)
    if (!std::isfinite(coefficients[i]))
      return fail(NSM_ERR_ARG, "non-finite closure coefficient for sample %zu", i / (size_t)(num_exact + 1));
  int st = ensure_device(md);
  if (st) return st;
  NSM_HIP(hipStreamSynchronize(md->stream));
  const size_t state_doubles = 12 * (size_t)md->N * (size_t)(num_exact + num_approx);
  const size_t old_bytes     = (12 * (size_t)md->N * (size_t)md->uq_samples() +
                            (size_t)md->uq_approx * (size_t)(md->uq_exact + 1)) * sizeof(double);
  free_dev(md->uq_state);
  free_dev(md->uq_coeff);
  free_dev(md->uq_bc_table);
  md->device_bytes -= std::min(md->device_bytes, old_bytes);
  md->uq_exact = md->uq_approx = 0;
  md->uq_K.clear();
  md->uq_G.clear();
  md->bc_dirty = true;
  if (num_exact + num_approx == 0) return NSM_OK;
  if ((st = dalloc(&md->uq_state, state_doubles, &md->device_bytes)) ||
      (st = dalloc(&md->uq_coeff, nc, &md->device_bytes)))
    return st;
  NSM_HIP(hipMemsetAsync(md->uq_state, 0, state_doubles * sizeof(double), md->stream));
  if (nc) NSM_HIP(hipMemcpyAsync(md->uq_coeff, coefficients, nc * sizeof(double), hipMemcpyHostToDevice, md->stream));
  NSM_HIP(hipStreamSynchronize(md->stream));
  md->uq_K.assign(bulk_modulus, bulk_modulus + nk);
  md->uq_G.assign(shear_modulus, shear_modulus + nk);
  md->uq_exact  = num_exact;
  md->uq_approx = num_approx;
  return NSM_OK;
}

// node array of field `field` of UQ sample `sample`
static int uq_field(nsm_model* md, int sample, int field, double** p)
{
  if (sample < 0 || sample >= md->uq_samples())
    return fail(NSM_ERR_ARG, "UQ sample %d out of range (%d samples)", sample, md->uq_samples());
  const nsm::NodeArrays na = md->node_arrays(sample);
  switch (field) {
    case NSM_FIELD_DISPLACEMENT: *p = na.u; break;
    case NSM_FIELD_VELOCITY: *p = na.v; break;
    case NSM_FIELD_ACCELERATION: *p = na.a; break;
    case NSM_FIELD_INTERNAL_FORCE: *p = na.f; break;
    default: return fail(NSM_ERR_ARG, "field %d is not a UQ sample field", field);
  }
  return NSM_OK;
}

int nsm_model_get_uq_field(nsm_model* md, int sample, int field, double* out)
{
  if (!md || !out) return fail(NSM_ERR_ARG, "null argument");
  double* p  = nullptr;
  int     st = ensure_device(md);
  if (st || (st = uq_field(md, sample, field, &p))) return st;
  return node_field_out(md, p, 3, out);
}

int nsm_model_set_uq_field(nsm_model* md, int sample, int field, const double* in)
{
  if (!md || !in) return fail(NSM_ERR_ARG, "null argument");
  double* p  = nullptr;
  int     st = ensure_device(md);
  if (st || (st = uq_field(md, sample, field, &p))) return st;
  return node_field_in(md, p, 3, in);
}

int nsm_model_field_device_pointer(nsm_model* md, int field, void** ptr, int* elem_bytes)
{
  if (!md || !ptr) return fail(NSM_ERR_ARG, "null argument");
  int eb = 8;
  switch (field) {
    case NSM_FIELD_REFERENCE_COORDINATE: *ptr = md->X; break;
    case NSM_FIELD_DISPLACEMENT: *ptr = md->u; break;
    case NSM_FIELD_VELOCITY: *ptr = md->v; break;
    case NSM_FIELD_ACCELERATION: *ptr = md->a; break;
    case NSM_FIELD_INTERNAL_FORCE: *ptr = md->f; break;
    case NSM_FIELD_LUMPED_MASS: *ptr = md->m; break;
    case NSM_FIELD_DEFORMATION_GRADIENT:
      *ptr = md->Fip;
      eb   = md->precision == NSM_FP64 ? 8 : 4;
      break;
    case NSM_FIELD_STRESS:
      *ptr = md->sip;
      eb   = md->precision == NSM_FP64 ? 8 : 4;
      break;
    default: return fail(NSM_ERR_ARG, "unknown field %d", field);
  }
  if (elem_bytes) *elem_bytes = eb;
  return NSM_OK;
}

int nsm_model_synchronize(nsm_model* md)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  NSM_HIP(hipSetDevice(md->device));
  NSM_HIP(hipStreamSynchronize(md->stream));
  return NSM_OK;
}

int nsm_model_enable_kernel_timing(nsm_model* md, int enable)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  NSM_HIP(hipStreamSynchronize(md->stream));
  if (enable < 0) return fail(NSM_ERR_ARG, "negative timing interval");
  md->timing       = enable != 0;
  md->timing_every = enable > 0 ? enable : 1;
  for (auto& c : md->timing_count) c = 0;
  md->event_next = 0;
  md->timed.clear();
  return NSM_OK;
}

int nsm_model_kernel_time(nsm_model* md, int kernel, double* avg_ms, int64_t* launches)
{
  if (!md || !avg_ms) return fail(NSM_ERR_ARG, "null argument");
  NSM_HIP(hipStreamSynchronize(md->stream));
  double  sum = 0.0;
  int64_t n   = 0;
  for (const auto& t : md->timed) {
    if (t.first != kernel) continue;
    float ms = 0.f;
    NSM_HIP(hipEventElapsedTime(&ms, md->event_pool[t.second].start, md->event_pool[t.second].stop));
    sum += ms;
    n++;
  }
  *avg_ms = n ? sum / (double)n : 0.0;
  if (launches) *launches = n;
  return NSM_OK;
}

int nsm_model_get_node_order(nsm_model* md, int64_t* out)
{
  if (!md || !out) return fail(NSM_ERR_ARG, "null argument");
  for (int64_t i = 0; i < md->N; i++) out[i] = mesh_node(md, i);
  return NSM_OK;
}

int nsm_model_get_element_order(nsm_model* md, int64_t* out)
{
  if (!md || (md->E > 0 && !out)) return fail(NSM_ERR_ARG, "null argument");
  for (int64_t e = 0; e < md->E; e++) out[e] = element_of(md, e);
  return NSM_OK;
}

int nsm_model_assembly_info(nsm_model* md, int* tiled, int64_t* num_tiles, int64_t* interior_nodes,
                            int64_t* boundary_nodes, int64_t* gather_tiles)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  if (tiled) *tiled = md->tiled ? 1 : 0;
  if (gather_tiles) *gather_tiles = std::count(md->tiles.gather.begin(), md->tiles.gather.end(), (uint8_t)1);
  if (num_tiles) *num_tiles = md->tiles.num_tiles;
  if (interior_nodes) *interior_nodes = md->tiles.num_interior;
  if (boundary_nodes) *boundary_nodes = md->tiled ? (int64_t)md->tiles.bnd_nodes.size() : md->N;
  return NSM_OK;
}

int nsm_model_info(nsm_model* md, int64_t* num_nodes, int64_t* num_elem, int64_t* csr_entries, int64_t* device_bytes)
{
  if (!md) return fail(NSM_ERR_ARG, "null model");
  if (num_nodes) *num_nodes = md->N;
  if (num_elem) *num_elem = md->E;
  if (csr_entries) *csr_entries = 8 * md->E;
  if (device_bytes) *device_bytes = (int64_t)md->device_bytes;
  return NSM_OK;
}

// ------------------------------------------------------------ synthetic mesh
static uint64_t splitmix64(uint64_t x)
{
  uint64_t z = x + 0x9E3779B97F4A7C15ULL;
  z          = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z          = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int nsm_jittered_cube(int n, uint64_t seed, double jitter_fraction, double* X, int* conn)
{
  if (n <= 0 || !X || !conn) return fail(NSM_ERR_ARG, "invalid cube arguments");
  const int64_t n1 = n + 1;
  if (n1 * n1 * n1 > (int64_t)INT32_MAX) return fail(NSM_ERR_ARG, "cube too large for int32 node ids");
  const uint64_t key = splitmix64(seed);
  for (int64_t k = 0; k <= n; k++)
    for (int64_t j = 0; j <= n; j++)
      for (int64_t i = 0; i <= n; i++) {
        const int64_t id = i + n1 * (j + n1 * k);
        X[3 * id]        = (double)i / (double)n;
        X[3 * id + 1]    = (double)j / (double)n;
        X[3 * id + 2]    = (double)k / (double)n;
        if (i > 0 && i < n && j > 0 && j < n && k > 0 && k < n)
          for (int d = 0; d < 3; d++) {
            const uint64_t r = splitmix64(key ^ (uint64_t)(3 * id + d));
            const double   U = (double)(r >> 11) * (1.0 / 9007199254740992.0);
            X[3 * id + d] += (2.0 * U - 1.0) * jitter_fraction / (double)n;
          }
      }
  for (int64_t k = 0; k < n; k++)
    for (int64_t j = 0; j < n; j++)
      for (int64_t i = 0; i < n; i++) {
        const int64_t e  = i + n * (j + n * k);
        const int64_t b  = i + n1 * (j + n1 * k);
        int*          c  = conn + 8 * e;
        c[0]             = (int)b;
        c[1]             = (int)(b + 1);
        c[2]             = (int)(b + 1 + n1);
        c[3]             = (int)(b + n1);
        c[4]             = (int)(b + n1 * n1);
        c[5]             = (int)(b + 1 + n1 * n1);
        c[6]             = (int)(b + 1 + n1 + n1 * n1);
        c[7]             = (int)(b + n1 + n1 * n1);
      }
  return NSM_OK;
}

}  // extern "C"
