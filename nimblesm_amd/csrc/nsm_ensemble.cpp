// nsm_ensemble.cpp -- host side of the batched ensemble (nsm_ensemble_* of
// the C ABI).
//
// Samples of all meshes are laid out back to back (nsm_ensemble.h) and every
// step of every sample is two launches: the element launch (mixed: the
// persistent tiled kernel over the element tiles of all samples, which also
// completes tile-interior nodes; FP64: element chunks into per-sample force
// slots) and the node launch (mixed: boundary nodes; FP64: the CSR gather and
// update of all nodes).  Each sample's arithmetic, summation orders and
// reaction sum are those of a model holding that sample alone, so a sample's
// trajectory does not depend on the rest of the ensemble.
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
#include "nsm_ensemble.h"
#include "nsm_kernels.h"
#include "nsm_tiled.h"
#include "nsm_tiles.h"

namespace nsm {
int set_last_error(int status, const char* msg);
}

namespace {

int fail(int status, const char* fmt, ...)
{
  char    buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  return nsm::set_last_error(status, buf);
}

#define ENS_HIP(call)                                                                              \
  do {                                                                                             \
    hipError_t _e = (call);                                                                        \
    if (_e != hipSuccess) return fail(NSM_ERR_HIP, "%s: %s", #call, hipGetErrorString(_e));        \
  } while (0)

#define ENS_LAUNCHED()                                                                             \
  do {                                                                                             \
    hipError_t _e = hipGetLastError();                                                             \
    if (_e != hipSuccess) return fail(NSM_ERR_HIP, "kernel launch: %s", hipGetErrorString(_e));    \
  } while (0)

constexpr int kElemChunk = 32;   // elements per FP64 element workgroup
constexpr int kNodeChunk = 256;  // nodes per node workgroup (the model's node tiles)
constexpr int kMaxBlocksPerMesh = 8;

struct EnsBc {
  int              kind, coord;
  std::vector<int> nodes;
};

struct HostMesh {
  int64_t              N = 0, E = 0;
  int                  nb = 0;
  std::vector<int64_t> blk_off;
  std::vector<double>  X;     // mesh order
  std::vector<int>     conn;  // mesh order
  std::vector<EnsBc>   bcs;
  std::vector<int>     rx_nodes;
  int                  rx_coord = -1;
  std::vector<int>     samples;
  // derived at setup
  nsm::TileSet                  tiles;     // mixed
  std::vector<int32_t>          dev_of;    // mesh node -> device row
  std::vector<uint8_t>          flags;     // device order
  std::vector<std::vector<int>> code_bcs;  // code c (1-based) -> conditions applying, in order
  std::vector<double>           min_charlen;  // per block
  int                           tile_base = 0, bnd_pos = 0, ptr_base = 0, elem_base = 0, x_base = 0;
  int                           num_partials = 0;
};

struct HostSample {
  int                 mesh = 0;
  std::vector<double> rho, K, G, mag;
};

struct EventPair {
  hipEvent_t start, stop;
};

template <typename T>
int dalloc(std::vector<void*>& bufs, T** p, size_t count, size_t* total)
{
  *p = nullptr;
  if (count == 0) count = 1;
  hipError_t e = hipMalloc((void**)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(NSM_ERR_NOMEM, "hipMalloc(%zu bytes): %s", count * sizeof(T), hipGetErrorString(e));
  }
  bufs.push_back((void*)*p);
  *total += count * sizeof(T);
  return NSM_OK;
}

template <typename T>
int upload(std::vector<void*>& bufs, const std::vector<T>& h, T** out, size_t* total, hipStream_t s)
{
  int st = dalloc(bufs, out, h.size(), total);
  if (st) return st;
  if (!h.empty()) ENS_HIP(hipMemcpyAsync(*out, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, s));
  return NSM_OK;
}

nsm::StepTimes step_times(double* time_current, double dt_user)
{
  // src/integrators/explicit_time_integrator.cc:192-195
  nsm::StepTimes t;
  t.t_prev = *time_current;
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

}  // namespace

struct nsm_ensemble {
  int         device = 0;
  hipStream_t stream = nullptr;
  int         precision = NSM_FP64;
  int         tile_request = 0;  // 0: autotune
  int         tile_elems = 0;
  bool        ready = false;

  std::vector<HostMesh>   meshes;
  std::vector<HostSample> samples;
  std::vector<double>     critical_dt;

  // device data
  std::vector<void*> bufs;
  size_t             device_bytes = 0;
  nsm::EnsNodes      en{};
  nsm::EnsSample*    d_samples = nullptr;
  nsm::EnsMesh*      d_meshes  = nullptr;
  double *           d_K = nullptr, *d_G = nullptr, *d_rho = nullptr;
  float *            d_Kf = nullptr, *d_Gf = nullptr;
  nsm::BcEntry *     d_bc = nullptr, *d_init = nullptr;
  double*            d_mass = nullptr;
  int*               d_conn = nullptr;
  uint32_t *         d_ptr = nullptr, *d_ent = nullptr;
  double*            d_slots = nullptr;      // FP64
  double*            d_rjac = nullptr;       // FP64: reference inverse Jacobians, [E_mesh][8][10] per mesh
  double*            d_elem_mass = nullptr;  // setup
  double*            d_zero = nullptr;       // zero displacement of the largest mesh
  double*            d_wg_min = nullptr;
  int*               d_status = nullptr;
  nsm::EnsRange *    d_elem_chunks = nullptr, *d_node_chunks = nullptr, *d_bnd_groups = nullptr;
  nsm::EnsTile*      d_tiles = nullptr;
  int32_t*           d_tile_block = nullptr;
  int                num_elem_chunks = 0, num_node_chunks = 0, num_bnd_groups = 0, num_tile_work = 0;
  nsm::TileArgs      ta{};
  double*            d_rx_partials = nullptr;  // 2 x rx_stride
  int64_t            rx_stride = 0;
  int *              d_fold_ptr = nullptr, *d_fold_idx = nullptr;
  bool               have_reaction = false;
  int                rx_coord      = -1;  // one coordinate for every mesh's reaction sum
  int64_t            N_total = 0, E_total = 0;
  int64_t            slot_elems = 0;  // sum over samples of E_mesh rounded up to 8 (slot rows)

  // per run
  nsm::StepTimes* d_times = nullptr;
  int64_t         times_cap = 0;
  double*         d_hist = nullptr;
  int64_t         hist_cap = 0, hist_steps = 0;
  double*         d_qoi = nullptr;
  int64_t         qoi_cap = 0;
  std::map<std::pair<int64_t, int64_t>, hipGraphExec_t> graphs;
  std::vector<std::pair<int, double>>                   tune_timings;  // autotuner: (tile size, s per step)

  bool                                timing = false;
  std::vector<EventPair>              event_pool;
  size_t                              event_next = 0;
  std::vector<std::pair<int, size_t>> timed;

  int64_t S() const { return (int64_t)samples.size(); }

  void drop_graphs()
  {
    for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
    graphs.clear();
  }
};

namespace {

int time_begin(nsm_ensemble* es, int kernel, size_t* idx)
{
  *idx = (size_t)-1;
  if (!es->timing) return NSM_OK;
  if (es->event_next == es->event_pool.size()) {
    EventPair p;
    ENS_HIP(hipEventCreate(&p.start));
    ENS_HIP(hipEventCreate(&p.stop));
    es->event_pool.push_back(p);
  }
  *idx = es->event_next++;
  ENS_HIP(hipEventRecord(es->event_pool[*idx].start, es->stream));
  es->timed.emplace_back(kernel, *idx);
  return NSM_OK;
}

int time_end(nsm_ensemble* es, size_t idx)
{
  if (idx == (size_t)-1) return NSM_OK;
  ENS_HIP(hipEventRecord(es->event_pool[idx].stop, es->stream));
  return NSM_OK;
}

// Kinematic-condition entry of one code for one sample: the conditions of
// the code applied in order (as nsm_model's per-node resolution), and the
// initial velocity the code's initial-velocity conditions leave.
void resolve_code(const HostMesh& m, const std::vector<int>& conds, const std::vector<double>& mag,
                  nsm::BcEntry& e, nsm::BcEntry& init)
{
  std::memset(&e, 0, sizeof(e));
  std::memset(&init, 0, sizeof(init));
  for (int bi : conds) {
    const EnsBc& b = m.bcs[(size_t)bi];
    const int    c = b.coord;
    const double v = mag[(size_t)bi];
    if (b.kind == NSM_BC_INITIAL_VELOCITY) {
      init.kind[c]  = 1;
      init.value[c] = v;
    } else if (b.kind == NSM_BC_PRESCRIBED_VELOCITY) {
      e.kind[c]     = nsm::kBcVelocity;
      e.value[c]    = v;
      e.fallback[c] = 0.0;
    } else {  // prescribed displacement: over an earlier velocity condition?
      const bool   over_v = e.kind[c] == nsm::kBcVelocity || e.kind[c] == nsm::kBcDisplacementOverVelocity;
      const double fb     = e.kind[c] == nsm::kBcVelocity ? e.value[c] : e.fallback[c];
      e.kind[c]           = over_v ? nsm::kBcDisplacementOverVelocity : nsm::kBcDisplacement;
      e.value[c]          = v;
      e.fallback[c]       = over_v ? fb : 0.0;
    }
  }
}

// Per-mesh derived data: device node order, condition codes and flags.
int prepare_mesh(nsm_ensemble* es, HostMesh& m)
{
  m.dev_of.assign((size_t)m.N, 0);
  if (es->precision == NSM_MIXED && m.E > 0) {
    std::string err;
    if (!nsm::build_tiles(m.N, m.X.data(), m.conn.data(), m.blk_off, es->tile_elems, m.tiles, &err))
      return fail(NSM_ERR_ARG, "mesh cannot be tiled: %s", err.c_str());
    for (int64_t n = 0; n < m.N; n++) m.dev_of[(size_t)n] = m.tiles.node_new[(size_t)n];
  } else {
    m.tiles = nsm::TileSet();
    for (int64_t n = 0; n < m.N; n++) m.dev_of[(size_t)n] = (int32_t)n;
  }
  // condition lists per node -> codes
  std::vector<std::vector<int>> per_node((size_t)m.N);
  for (size_t bi = 0; bi < m.bcs.size(); bi++)
    for (int n : m.bcs[bi].nodes) per_node[(size_t)n].push_back((int)bi);
  std::map<std::vector<int>, int> code_of;
  m.code_bcs.clear();
  m.flags.assign((size_t)m.N, 0);
  for (int64_t n = 0; n < m.N; n++) {
    const auto& l = per_node[(size_t)n];
    if (l.empty()) continue;
    auto it = code_of.find(l);
    if (it == code_of.end()) {
      if ((int)m.code_bcs.size() >= nsm::kEnsMaxBcCodes)
        return fail(NSM_ERR_ARG, "a mesh's conditions form more than %d distinct node combinations",
                    nsm::kEnsMaxBcCodes);
      m.code_bcs.push_back(l);
      it = code_of.emplace(l, (int)m.code_bcs.size()).first;
    }
    m.flags[(size_t)m.dev_of[(size_t)n]] = (uint8_t)it->second;
  }
  for (int n : m.rx_nodes) m.flags[(size_t)m.dev_of[(size_t)n]] |= nsm::kReactionFlag;
  return NSM_OK;
}

// Moduli, densities and condition tables of every sample to the device.
int upload_parameters(nsm_ensemble* es)
{
  std::vector<double>       K, G, rho;
  std::vector<float>        Kf, Gf;
  std::vector<nsm::BcEntry> bc((size_t)es->S() * nsm::kEnsMaxBcCodes), init(bc.size());
  std::memset(bc.data(), 0, bc.size() * sizeof(nsm::BcEntry));
  std::memset(init.data(), 0, init.size() * sizeof(nsm::BcEntry));
  for (int64_t s = 0; s < es->S(); s++) {
    const HostSample& hs = es->samples[(size_t)s];
    const HostMesh&   m  = es->meshes[(size_t)hs.mesh];
    for (int b = 0; b < m.nb; b++) {
      K.push_back(hs.K[(size_t)b]);
      G.push_back(hs.G[(size_t)b]);
      rho.push_back(hs.rho[(size_t)b]);
      Kf.push_back((float)hs.K[(size_t)b]);
      Gf.push_back((float)hs.G[(size_t)b]);
    }
    for (size_t c = 0; c < m.code_bcs.size(); c++)
      resolve_code(m, m.code_bcs[c], hs.mag, bc[(size_t)s * nsm::kEnsMaxBcCodes + c],
                   init[(size_t)s * nsm::kEnsMaxBcCodes + c]);
  }
  hipStream_t st = es->stream;
  auto        h2d = [&](void* d, const void* h, size_t bytes) -> int {
    if (bytes) ENS_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
    return NSM_OK;
  };
  int r;
  if ((r = h2d(es->d_K, K.data(), K.size() * 8)) || (r = h2d(es->d_G, G.data(), G.size() * 8)) ||
      (r = h2d(es->d_rho, rho.data(), rho.size() * 8)) || (r = h2d(es->d_Kf, Kf.data(), Kf.size() * 4)) ||
      (r = h2d(es->d_Gf, Gf.data(), Gf.size() * 4)) ||
      (r = h2d(es->d_bc, bc.data(), bc.size() * sizeof(nsm::BcEntry))) ||
      (r = h2d(es->d_init, init.data(), init.size() * sizeof(nsm::BcEntry))))
    return r;
  ENS_HIP(hipStreamSynchronize(st));
  return NSM_OK;
}

int check_status(nsm_ensemble* es)
{
  int h = 0;
  ENS_HIP(hipMemcpyAsync(&h, es->d_status, sizeof(int), hipMemcpyDeviceToHost, es->stream));
  ENS_HIP(hipStreamSynchronize(es->stream));
  if (h) {
    ENS_HIP(hipMemsetAsync(es->d_status, 0, sizeof(int), es->stream));
    ENS_HIP(hipStreamSynchronize(es->stream));
    return fail(NSM_ERR_NEG_JACOBIAN, "non-positive Jacobian determinant in an element (Invert3x3)");
  }
  return NSM_OK;
}

// Lumped mass of every sample (its block densities inside the element sums,
// then the ordered node gather), and each sample's critical time step from
// its mesh's per-block minimum characteristic length: min_e (l_e / c) =
// (min_e l_e) / c exactly, since a correctly rounded division by c > 0 is
// monotone.
int compute_mass_and_dt(nsm_ensemble* es)
{
  nsm::launch_ens_element_mass(es->en, es->d_elem_chunks, es->num_elem_chunks, es->d_conn, es->d_rho,
                               es->d_elem_mass, es->d_status, es->stream);
  ENS_LAUNCHED();
  nsm::launch_ens_mass_gather(es->en, es->d_node_chunks, es->num_node_chunks, es->d_ptr, es->d_ent,
                              es->d_elem_mass, es->d_mass, es->stream);
  ENS_LAUNCHED();
  int st = check_status(es);
  if (st) return st;
  es->critical_dt.assign((size_t)es->S(), 0.0);
  for (int64_t s = 0; s < es->S(); s++) {
    const HostSample& hs = es->samples[(size_t)s];
    const HostMesh&   m  = es->meshes[(size_t)hs.mesh];
    double            dt = 1.7976931348623157e308;
    for (int b = 0; b < m.nb; b++) {
      if (m.blk_off[(size_t)b + 1] == m.blk_off[(size_t)b]) continue;
      const double c  = std::sqrt(hs.K[(size_t)b] / hs.rho[(size_t)b]);
      const double bd = m.min_charlen[(size_t)b] / c;
      if (bd < dt) dt = bd;
    }
    es->critical_dt[(size_t)s] = dt;
  }
  return NSM_OK;
}

int min_charlen(nsm_ensemble* es)
{
  for (auto& m : es->meshes) {
    m.min_charlen.assign((size_t)m.nb, 1.7976931348623157e308);
    for (int b = 0; b < m.nb; b++) {
      const int64_t e0 = m.blk_off[(size_t)b], e1 = m.blk_off[(size_t)b + 1];
      if (e1 <= e0) continue;
      const int ng = nsm::launch_element_dt(e0, e1, es->d_conn + 8 * (int64_t)m.elem_base,
                                            es->en.X + 3 * (int64_t)m.x_base, es->d_zero, 1.0, es->d_wg_min,
                                            es->stream);
      ENS_LAUNCHED();
      nsm::launch_min_reduce(es->d_wg_min, ng, es->d_wg_min + ng, es->stream);
      ENS_LAUNCHED();
      double v = 0.0;
      ENS_HIP(hipMemcpyAsync(&v, es->d_wg_min + ng, sizeof(double), hipMemcpyDeviceToHost, es->stream));
      ENS_HIP(hipStreamSynchronize(es->stream));
      m.min_charlen[(size_t)b] = v;
    }
  }
  return NSM_OK;
}

int build_device(nsm_ensemble* es)
{
  hipStream_t st  = es->stream;
  size_t&     tot = es->device_bytes;
  auto&       B   = es->bufs;
  int         r   = 0;
  const bool  mixed = es->precision == NSM_MIXED;

  // ---- meshes: X, flags, connectivity, CSR, tile tables
  std::vector<double>   X;
  std::vector<uint8_t>  flags;
  std::vector<int>      conn;
  std::vector<uint32_t> ptr, ent;
  std::vector<int32_t>  elem_off, node_off, nodes, slot_dst, num_colors, gather, bnd_ptr, tile_block;
  std::vector<uint16_t> lconn;
  std::vector<uint8_t>  color;
  std::vector<nsm::EnsMesh> dm;
  int64_t               max_nodes = 1, max_elem_dt = 1;
  for (auto& m : es->meshes) {
    m.x_base    = (int)(X.size() / 3);
    m.elem_base = (int)(conn.size() / 8);
    m.ptr_base  = (int)ptr.size();
    const size_t xb = X.size();
    X.resize(xb + 3 * (size_t)m.N);
    for (int64_t n = 0; n < m.N; n++)
      for (int c = 0; c < 3; c++) X[xb + 3 * (size_t)m.dev_of[(size_t)n] + c] = m.X[3 * (size_t)n + c];
    flags.insert(flags.end(), m.flags.begin(), m.flags.end());
    for (int64_t k = 0; k < 8 * m.E; k++) conn.push_back(m.dev_of[(size_t)m.conn[(size_t)k]]);
    // node -> (element, corner) incidences in ascending element order, as
    // the slot index fp64_slot(element, corner) of the FP64 force slots and
    // element masses; node_ptr values index the concatenated entries
    std::vector<uint32_t> cnt((size_t)m.N + 1, 0);
    const int*            cd = conn.data() + 8 * (size_t)m.elem_base;
    for (int64_t k = 0; k < 8 * m.E; k++) cnt[(size_t)cd[k] + 1]++;
    for (int64_t i = 0; i < m.N; i++) cnt[(size_t)i + 1] += cnt[(size_t)i];
    const uint32_t eb = (uint32_t)ent.size();
    for (int64_t i = 0; i <= m.N; i++) ptr.push_back(eb + cnt[(size_t)i]);
    ent.resize(eb + (size_t)(8 * m.E));
    {
      std::vector<uint32_t> fill(cnt.begin(), cnt.end() - 1);
      for (int64_t k = 0; k < 8 * m.E; k++)
        ent[eb + fill[(size_t)cd[k]]++] = (uint32_t)nsm::fp64_slot(k >> 3, (int)(k & 7));
    }
    nsm::EnsMesh d{};
    d.num_nodes = (int32_t)m.N;
    d.num_elem  = (int32_t)m.E;
    d.elem_base = m.elem_base;
    d.ptr_base  = m.ptr_base;
    if (mixed && m.E > 0) {
      const nsm::TileSet& ts = m.tiles;
      m.tile_base    = (int)num_colors.size();
      m.bnd_pos      = (int)bnd_ptr.size();
      m.num_partials = (int)ts.bnd_ent.size();
      const int32_t e_shift = (int32_t)lconn.size() / 8, s_shift = (int32_t)nodes.size();
      for (int t = 0; t < ts.num_tiles; t++) {
        elem_off.push_back(ts.elem_off[(size_t)t] + e_shift);
        node_off.push_back(ts.node_off[(size_t)t] + s_shift);
        num_colors.push_back(ts.num_colors[(size_t)t]);
        gather.push_back(ts.gather[(size_t)t]);
        const int64_t e0 = ts.elem_off[(size_t)t];
        tile_block.push_back((int32_t)(std::upper_bound(m.blk_off.begin(), m.blk_off.end(), e0) -
                                       m.blk_off.begin() - 1));
      }
      nodes.insert(nodes.end(), ts.nodes.begin(), ts.nodes.end());
      slot_dst.insert(slot_dst.end(), ts.slot_dst.begin(), ts.slot_dst.end());
      lconn.insert(lconn.end(), ts.lconn.begin(), ts.lconn.end());
      color.insert(color.end(), ts.color.begin(), ts.color.end());
      bnd_ptr.insert(bnd_ptr.end(), ts.bnd_ptr.begin(), ts.bnd_ptr.end());
      d.num_interior = (int32_t)ts.num_interior;
      d.bnd_pos      = m.bnd_pos;
    }
    dm.push_back(d);
    max_nodes = std::max<int64_t>(max_nodes, m.N);
    for (int b = 0; b < m.nb; b++)
      max_elem_dt = std::max<int64_t>(max_elem_dt, (m.blk_off[(size_t)b + 1] - m.blk_off[(size_t)b] + 255) / 256);
  }
  if (mixed) {
    // closing entries of the concatenated tile ranges
    elem_off.push_back((int32_t)lconn.size() / 8);
    node_off.push_back((int32_t)nodes.size());
  }

  // ---- samples: offsets, work tables, reaction folds
  std::vector<nsm::EnsSample> ds;
  std::vector<nsm::EnsRange>  elem_chunks, node_chunks, bnd_groups;
  std::vector<nsm::EnsTile>   tile_work;
  int64_t                     node_base = 0, partial_base = 0, elem_prefix = 0, slot_elems = 0, kg = 0;
  for (int64_t s = 0; s < es->S(); s++) {
    const HostMesh& m = es->meshes[(size_t)es->samples[(size_t)s].mesh];
    nsm::EnsSample  d{};
    d.mesh         = es->samples[(size_t)s].mesh;
    d.node_base    = (int32_t)node_base;
    d.x_base       = m.x_base;
    d.partial_base = (int32_t)partial_base;
    d.bc_base      = (int32_t)(s * nsm::kEnsMaxBcCodes);
    d.kg_base      = (int32_t)kg;
    d.slot_base    = 24 * slot_elems;  // [3][8 E8] per sample (E8: E_mesh rounded up to 8)
    ds.push_back(d);
    for (int b = 0; b < m.nb; b++)
      for (int64_t e = m.blk_off[(size_t)b]; e < m.blk_off[(size_t)b + 1]; e += kElemChunk)
        elem_chunks.push_back(nsm::EnsRange{(int32_t)s, (int32_t)e,
                                            (int32_t)std::min<int64_t>(kElemChunk, m.blk_off[(size_t)b + 1] - e), b});
    for (int64_t n = 0; n < m.N; n += kNodeChunk)
      node_chunks.push_back(nsm::EnsRange{(int32_t)s, (int32_t)n, (int32_t)std::min<int64_t>(kNodeChunk, m.N - n), 0});
    if (mixed && m.E > 0) {
      for (int t = 0; t < m.tiles.num_tiles; t++) tile_work.push_back(nsm::EnsTile{m.tile_base + t, (int32_t)s});
      const int64_t nb = (int64_t)m.tiles.bnd_nodes.size();
      for (int64_t b = 0; b < nb; b += nsm::kBoundaryNodesPerGroup)
        bnd_groups.push_back(nsm::EnsRange{(int32_t)s, (int32_t)b,
                                           (int32_t)std::min<int64_t>(nsm::kBoundaryNodesPerGroup, nb - b), 0});
    }
    node_base += m.N;
    partial_base += m.num_partials;
    elem_prefix += m.E;
    slot_elems += nsm::round_up8(m.E);
    kg += m.nb;
  }
  if (node_base > (int64_t)INT32_MAX / 3 || partial_base > (int64_t)INT32_MAX / 3 ||
      8 * slot_elems > (int64_t)INT32_MAX)
    return fail(NSM_ERR_ARG, "ensemble too large for 32-bit row indices");
  es->N_total         = node_base;
  es->E_total         = elem_prefix;
  es->slot_elems      = slot_elems;
  es->num_elem_chunks = (int)elem_chunks.size();
  es->num_node_chunks = (int)node_chunks.size();
  es->num_bnd_groups  = (int)bnd_groups.size();
  es->num_tile_work   = (int)tile_work.size();

  // reaction folds: per sample, the partials holding its reaction nodes in
  // the order a model of the sample alone sums them
  std::vector<int> fold_ptr(1, 0), fold_idx;
  es->have_reaction = false;
  {
    size_t tw = 0, bg = 0, nc = 0;
    for (int64_t s = 0; s < es->S(); s++) {
      const HostMesh&   m = es->meshes[(size_t)es->samples[(size_t)s].mesh];
      std::vector<char> is_rx((size_t)m.N, 0);  // device rows
      for (int n : m.rx_nodes) is_rx[(size_t)m.dev_of[(size_t)n]] = 1;
      if (!m.rx_nodes.empty()) es->have_reaction = true;
      if (mixed && m.E > 0) {
        const nsm::TileSet& ts = m.tiles;
        for (int t = 0; t < ts.num_tiles; t++, tw++)
          for (int sl = ts.node_off[(size_t)t]; sl < ts.node_off[(size_t)t + 1]; sl++)
            if ((ts.node_info[(size_t)sl] & nsm::kInteriorNode) && is_rx[(size_t)ts.nodes[(size_t)sl]]) {
              fold_idx.push_back((int)tw);
              break;
            }
        const int64_t nb = (int64_t)ts.bnd_nodes.size();
        for (int64_t b0 = 0; b0 < nb; b0 += nsm::kBoundaryNodesPerGroup, bg++)
          for (int64_t b = b0; b < std::min<int64_t>(nb, b0 + nsm::kBoundaryNodesPerGroup); b++)
            if (is_rx[(size_t)(ts.num_interior + b)]) {
              fold_idx.push_back((int)(tile_work.size() + bg));
              break;
            }
      } else {
        for (int64_t n0 = 0; n0 < m.N; n0 += kNodeChunk, nc++)
          for (int64_t n = n0; n < std::min<int64_t>(m.N, n0 + kNodeChunk); n++)
            if (is_rx[(size_t)n]) {
              fold_idx.push_back((int)nc);
              break;
            }
      }
      fold_ptr.push_back((int)fold_idx.size());
    }
  }
  // Work of the samples of one mesh side by side: element chunks by (mesh,
  // block, chunk, sample), tile work items by (mesh, tile, sample).  The
  // workgroups resident at one time then share the chunk's or tile's
  // connectivity, coordinates and (FP64) reference Jacobians in L2.  Each
  // item still writes only its own sample's rows, so the order changes no
  // result; the reaction folds follow the tile items to their new positions.
  // NSM_ENS_CHUNK_ORDER=sample keeps the sample-major order, for comparison.
  {
    const char* ord          = std::getenv("NSM_ENS_CHUNK_ORDER");
    const bool  sample_major = ord && std::strcmp(ord, "sample") == 0;
    if (!sample_major) {
      auto mesh_of = [&](int32_t s) { return es->samples[(size_t)s].mesh; };
      std::stable_sort(elem_chunks.begin(), elem_chunks.end(), [&](const nsm::EnsRange& x, const nsm::EnsRange& y) {
        if (mesh_of(x.s) != mesh_of(y.s)) return mesh_of(x.s) < mesh_of(y.s);
        return x.first < y.first;  // blocks are ascending element ranges
      });
      std::vector<int> perm(tile_work.size());
      for (size_t i = 0; i < perm.size(); i++) perm[i] = (int)i;
      std::stable_sort(perm.begin(), perm.end(), [&](int x, int y) {
        const nsm::EnsTile &a = tile_work[(size_t)x], &b = tile_work[(size_t)y];
        if (mesh_of(a.s) != mesh_of(b.s)) return mesh_of(a.s) < mesh_of(b.s);
        return a.t < b.t;
      });
      std::vector<nsm::EnsTile> reordered(tile_work.size());
      std::vector<int>          pos(tile_work.size());
      for (size_t i = 0; i < perm.size(); i++) {
        reordered[i]         = tile_work[(size_t)perm[i]];
        pos[(size_t)perm[i]] = (int)i;
      }
      tile_work.swap(reordered);
      for (int& k : fold_idx)
        if (mixed && k < (int)tile_work.size()) k = pos[(size_t)k];
    }
  }
  es->rx_stride = mixed ? (int64_t)tile_work.size() + (int64_t)bnd_groups.size() : (int64_t)node_chunks.size();

  // ---- allocations and uploads
  const int64_t S  = es->S();
  const int64_t NB = kg;
  if ((r = upload(B, X, (double**)&es->en.X, &tot, st)) || (r = upload(B, flags, (uint8_t**)&es->en.flags, &tot, st)) ||
      (r = upload(B, conn, &es->d_conn, &tot, st)) || (r = upload(B, ptr, &es->d_ptr, &tot, st)) ||
      (r = upload(B, ent, &es->d_ent, &tot, st)) || (r = upload(B, dm, &es->d_meshes, &tot, st)) ||
      (r = upload(B, ds, &es->d_samples, &tot, st)) || (r = upload(B, elem_chunks, &es->d_elem_chunks, &tot, st)) ||
      (r = upload(B, node_chunks, &es->d_node_chunks, &tot, st)) || (r = upload(B, fold_ptr, &es->d_fold_ptr, &tot, st)) ||
      (r = upload(B, fold_idx, &es->d_fold_idx, &tot, st)))
    return r;
  const size_t N3 = 3 * (size_t)es->N_total;
  double *     u, *v, *a, *f;
  if ((r = dalloc(B, &u, N3, &tot)) || (r = dalloc(B, &v, N3, &tot)) || (r = dalloc(B, &a, N3, &tot)) ||
      (r = dalloc(B, &f, N3, &tot)) || (r = dalloc(B, &es->d_mass, (size_t)es->N_total, &tot)) ||
      (r = dalloc(B, &es->d_K, (size_t)NB, &tot)) || (r = dalloc(B, &es->d_G, (size_t)NB, &tot)) ||
      (r = dalloc(B, &es->d_rho, (size_t)NB, &tot)) || (r = dalloc(B, &es->d_Kf, (size_t)NB, &tot)) ||
      (r = dalloc(B, &es->d_Gf, (size_t)NB, &tot)) ||
      (r = dalloc(B, &es->d_bc, (size_t)(S * nsm::kEnsMaxBcCodes), &tot)) ||
      (r = dalloc(B, &es->d_init, (size_t)(S * nsm::kEnsMaxBcCodes), &tot)) ||
      (r = dalloc(B, &es->d_elem_mass, (size_t)(8 * es->slot_elems), &tot)) ||
      (r = dalloc(B, &es->d_zero, 3 * (size_t)max_nodes, &tot)) ||
      (r = dalloc(B, &es->d_wg_min, (size_t)max_elem_dt + 1, &tot)) || (r = dalloc(B, &es->d_status, 1, &tot)) ||
      (r = dalloc(B, &es->d_rx_partials, 2 * (size_t)std::max<int64_t>(es->rx_stride, 1), &tot)))
    return r;
  es->en.u = u;
  es->en.v = v;
  es->en.a = a;
  es->en.f = f;
  ENS_HIP(hipMemsetAsync(u, 0, N3 * sizeof(double), st));
  ENS_HIP(hipMemsetAsync(v, 0, N3 * sizeof(double), st));
  ENS_HIP(hipMemsetAsync(a, 0, N3 * sizeof(double), st));
  ENS_HIP(hipMemsetAsync(f, 0, N3 * sizeof(double), st));
  ENS_HIP(hipMemsetAsync(es->d_zero, 0, 3 * (size_t)max_nodes * sizeof(double), st));
  ENS_HIP(hipMemsetAsync(es->d_status, 0, sizeof(int), st));
  es->en.samples = es->d_samples;
  es->en.meshes  = es->d_meshes;
  es->en.mass    = es->d_mass;
  es->en.bc      = es->d_bc;
  es->en.init    = es->d_init;
  if (!mixed) {
    if ((r = dalloc(B, &es->d_slots, (size_t)(24 * es->slot_elems), &tot)) ||
        (r = dalloc(B, &es->d_rjac, conn.size() * 10, &tot)))
      return r;
    for (const HostMesh& m : es->meshes) {
      nsm::launch_reference_jacobians(m.E, es->d_conn + 8 * (int64_t)m.elem_base, es->en.X + 3 * (int64_t)m.x_base,
                                      es->d_rjac + 80 * (int64_t)m.elem_base, st);
      ENS_HIP(hipGetLastError());
    }
  } else {
    nsm::TileArgs& ta = es->ta;
    ta.tile_elems     = es->tile_elems;
    int cus           = 0;
    ENS_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, es->device));
    ta.resident_groups = std::max(cus, 1) * std::max(nsm::element_tiled_groups_per_cu(es->tile_elems), 1);
    ta.num_tiles       = (int)num_colors.size();
    ta.num_bnd         = 0;
    ta.bnd_base        = 0;
    float* partials = nullptr;
    if ((r = upload(B, elem_off, (int32_t**)&ta.elem_off, &tot, st)) ||
        (r = upload(B, node_off, (int32_t**)&ta.node_off, &tot, st)) ||
        (r = upload(B, nodes, (int32_t**)&ta.nodes, &tot, st)) ||
        (r = upload(B, slot_dst, (int32_t**)&ta.slot_dst, &tot, st)) ||
        (r = upload(B, lconn, (uint16_t**)&ta.lconn, &tot, st)) ||
        (r = upload(B, color, (uint8_t**)&ta.color, &tot, st)) ||
        (r = upload(B, num_colors, (int32_t**)&ta.num_colors, &tot, st)) ||
        (r = upload(B, gather, (int32_t**)&ta.gather, &tot, st)) ||
        (r = upload(B, bnd_ptr, (int32_t**)&ta.bnd_ptr, &tot, st)) ||
        (r = upload(B, tile_block, &es->d_tile_block, &tot, st)) ||
        (r = upload(B, tile_work, &es->d_tiles, &tot, st)) ||
        (r = upload(B, bnd_groups, &es->d_bnd_groups, &tot, st)) ||
        (r = dalloc(B, &partials, 3 * (size_t)std::max<int64_t>(partial_base, 1), &tot)))
      return r;
    ta.partials = partials;
  }
  ENS_HIP(hipStreamSynchronize(st));
  if ((r = upload_parameters(es))) return r;
  if ((r = min_charlen(es))) return r;
  return compute_mass_and_dt(es);
}

void free_device(nsm_ensemble* es)
{
  es->drop_graphs();
  for (void* p : es->bufs) (void)hipFree(p);
  es->bufs.clear();
  es->device_bytes = 0;
  if (es->d_times) (void)hipFree(es->d_times);
  if (es->d_hist) (void)hipFree(es->d_hist);
  if (es->d_qoi) (void)hipFree(es->d_qoi);
  es->d_times   = nullptr;
  es->d_hist    = nullptr;
  es->d_qoi     = nullptr;
  es->times_cap = es->hist_cap = es->qoi_cap = 0;
}

// The launches of a whole run from rest (captured as they are issued when
// a graph is being recorded).
int issue_run(nsm_ensemble* es, int64_t n, int64_t output_frequency)
{
  const bool     mixed = es->precision == NSM_MIXED;
  nsm::EnsStep   st0{es->d_times, n + 1, 0, 0, 0, 0, 1, 1};
  nsm::launch_ens_reset(es->en, es->d_node_chunks, es->num_node_chunks, st0, es->stream);
  ENS_LAUNCHED();
  nsm::launch_ens_prekick(es->en, es->d_node_chunks, es->num_node_chunks, st0, es->stream);
  ENS_LAUNCHED();
  nsm::NodeArrays na{};
  na.mass  = es->en.mass;
  na.u     = es->en.u;
  na.v     = es->en.v;
  na.a     = es->en.a;
  na.f     = es->en.f;
  na.flags = es->en.flags;
  for (int64_t step = 0; step < n; step++) {
    const bool   last = step + 1 == n;
    // a sample's internal force is read only after the run: stored on the
    // last step (as the acceleration), not on every step
    nsm::EnsStep es_step{es->d_times, n + 1, (int)step, (int)is_output(step, n, output_frequency), (int)!last,
                         (int)last, 1, (int)last};
    nsm::ReactionArgs rx{-1, nullptr};
    if (es->have_reaction) rx = nsm::ReactionArgs{es->rx_coord, es->d_rx_partials + (step & 1) * es->rx_stride};
    nsm::EnsFold fold{es->d_rx_partials + ((step + 1) & 1) * es->rx_stride, es->d_fold_ptr, es->d_fold_idx,
                      (int)es->S(), es->d_hist, n, es->have_reaction && step > 0 ? step - 1 : -1};
    size_t ti;
    int    r = time_begin(es, NSM_ENSEMBLE_KERNEL_ELEMENT, &ti);
    if (r) return r;
    if (mixed) {
      nsm::EnsTileWork w{es->d_tiles, es->num_tile_work, es->d_samples, es->d_Kf, es->d_Gf, es->d_tile_block,
                         es->d_bc, es_step, fold};
      nsm::launch_ens_element_tiled(w, es->ta, es->en.X, na, rx, es->d_status, es->stream);
    } else {
      nsm::launch_ens_element_fp64(es->en, es->d_elem_chunks, es->num_elem_chunks, es->d_conn, es->d_rjac, es->d_K,
                                   es->d_G, es->d_slots, es->d_status, es->stream);
    }
    ENS_LAUNCHED();
    if ((r = time_end(es, ti))) return r;
    if ((r = time_begin(es, NSM_ENSEMBLE_KERNEL_NODE, &ti))) return r;
    if (mixed) {
      nsm::launch_ens_boundary_update(es->ta, es->en, es->d_bnd_groups, es->num_bnd_groups, es_step, rx,
                                      es->num_tile_work, es->stream);
    } else {
      nsm::EnsReaction erx{rx.coord, rx.partials};
      nsm::launch_ens_node_fp64(es->en, es->d_node_chunks, es->num_node_chunks, es->d_ptr, es->d_ent, es->d_slots,
                                es_step, erx, fold, es->stream);
    }
    ENS_LAUNCHED();
    if ((r = time_end(es, ti))) return r;
  }
  if (es->have_reaction) {
    nsm::EnsFold fold{es->d_rx_partials + ((n - 1) & 1) * es->rx_stride, es->d_fold_ptr, es->d_fold_idx,
                      (int)es->S(), es->d_hist, n, n - 1};
    nsm::launch_ens_fold(fold, es->stream);
    ENS_LAUNCHED();
  }
  return NSM_OK;
}

}  // namespace

extern "C" {

int nsm_ensemble_create(int precision, int tile_elems, nsm_ensemble** out)
{
  if (!out) return fail(NSM_ERR_ARG, "out is null");
  *out = nullptr;
  if (precision != NSM_FP64 && precision != NSM_MIXED) return fail(NSM_ERR_ARG, "unknown precision %d", precision);
  if (tile_elems != 0 && tile_elems != 64 && tile_elems != 128 && tile_elems != 256 && tile_elems != 512)
    return fail(NSM_ERR_ARG, "tile_elems must be 0 (autotune), 64, 128, 256 or 512");
  nsm_ensemble* es = new nsm_ensemble();
  es->precision    = precision;
  es->tile_request = tile_elems;
  if (hipGetDevice(&es->device) != hipSuccess) {
    delete es;
    return fail(NSM_ERR_HIP, "no HIP device");
  }
  if (hipStreamCreateWithFlags(&es->stream, hipStreamNonBlocking) != hipSuccess) {
    delete es;
    return fail(NSM_ERR_HIP, "hipStreamCreate failed");
  }
  *out = es;
  return NSM_OK;
}

int nsm_ensemble_destroy(nsm_ensemble* es)
{
  if (!es) return NSM_OK;
  (void)hipSetDevice(es->device);
  if (es->stream) (void)hipStreamSynchronize(es->stream);
  free_device(es);
  for (auto& p : es->event_pool) {
    (void)hipEventDestroy(p.start);
    (void)hipEventDestroy(p.stop);
  }
  if (es->stream) (void)hipStreamDestroy(es->stream);
  delete es;
  return NSM_OK;
}

int nsm_ensemble_add_mesh(nsm_ensemble* es, int64_t num_nodes, const double* X, int num_blocks,
                          const int64_t* block_num_elem, const int* conn, int* mesh_id)
{
  if (!es || !X || !block_num_elem || !mesh_id) return fail(NSM_ERR_ARG, "null argument");
  if (es->ready) return fail(NSM_ERR_STATE, "the ensemble is already set up");
  if (num_nodes <= 0 || num_nodes > (int64_t)INT32_MAX) return fail(NSM_ERR_ARG, "invalid node count");
  if (num_blocks <= 0 || num_blocks > kMaxBlocksPerMesh)
    return fail(NSM_ERR_ARG, "a mesh has 1 to %d blocks", kMaxBlocksPerMesh);
  HostMesh m;
  m.N  = num_nodes;
  m.nb = num_blocks;
  m.blk_off.assign(1, 0);
  for (int b = 0; b < num_blocks; b++) {
    if (block_num_elem[b] < 0) return fail(NSM_ERR_ARG, "negative element count in block %d", b);
    m.blk_off.push_back(m.blk_off.back() + block_num_elem[b]);
  }
  m.E = m.blk_off.back();
  // every node is updated by the element or node launches of its sample,
  // which a mesh without elements has none of
  if (m.E <= 0) return fail(NSM_ERR_ARG, "an ensemble mesh needs at least one element");
  if (!conn) return fail(NSM_ERR_ARG, "conn is null");
  for (int64_t k = 0; k < 8 * m.E; k++)
    if (conn[k] < 0 || conn[k] >= num_nodes)
      return fail(NSM_ERR_ARG, "connectivity entry %lld = %d out of range", (long long)k, conn[k]);
  m.X.assign(X, X + 3 * num_nodes);
  m.conn.assign(conn, conn + 8 * m.E);
  *mesh_id = (int)es->meshes.size();
  es->meshes.push_back(std::move(m));
  return NSM_OK;
}

int nsm_ensemble_add_boundary_condition(nsm_ensemble* es, int mesh_id, int kind, int coordinate, int64_t num_nodes,
                                        const int* node_ids)
{
  if (!es || (num_nodes > 0 && !node_ids)) return fail(NSM_ERR_ARG, "null argument");
  if (es->ready) return fail(NSM_ERR_STATE, "the ensemble is already set up");
  if (mesh_id < 0 || mesh_id >= (int)es->meshes.size()) return fail(NSM_ERR_ARG, "no mesh %d", mesh_id);
  HostMesh& m = es->meshes[(size_t)mesh_id];
  if (!m.samples.empty()) return fail(NSM_ERR_STATE, "conditions must be added before the mesh's samples");
  if (kind < NSM_BC_INITIAL_VELOCITY || kind > NSM_BC_PRESCRIBED_DISPLACEMENT)
    return fail(NSM_ERR_ARG, "unknown condition kind %d", kind);
  if (coordinate < 0 || coordinate > 2) return fail(NSM_ERR_ARG, "coordinate must be 0, 1 or 2");
  EnsBc b;
  b.kind  = kind;
  b.coord = coordinate;
  for (int64_t k = 0; k < num_nodes; k++) {
    if (node_ids[k] < 0 || node_ids[k] >= m.N) return fail(NSM_ERR_ARG, "node id %d out of range", node_ids[k]);
    b.nodes.push_back(node_ids[k]);
  }
  m.bcs.push_back(std::move(b));
  return NSM_OK;
}

int nsm_ensemble_set_reaction_nodes(nsm_ensemble* es, int mesh_id, int64_t num_nodes, const int* node_ids,
                                    int coordinate)
{
  if (!es || (num_nodes > 0 && !node_ids)) return fail(NSM_ERR_ARG, "null argument");
  if (es->ready) return fail(NSM_ERR_STATE, "the ensemble is already set up");
  if (mesh_id < 0 || mesh_id >= (int)es->meshes.size()) return fail(NSM_ERR_ARG, "no mesh %d", mesh_id);
  if (coordinate < 0 || coordinate > 2) return fail(NSM_ERR_ARG, "coordinate must be 0, 1 or 2");
  HostMesh&         m = es->meshes[(size_t)mesh_id];
  std::vector<char> seen((size_t)m.N, 0);
  m.rx_nodes.clear();
  for (int64_t k = 0; k < num_nodes; k++) {
    if (node_ids[k] < 0 || node_ids[k] >= m.N) return fail(NSM_ERR_ARG, "node id %d out of range", node_ids[k]);
    if (!seen[(size_t)node_ids[k]]) m.rx_nodes.push_back(node_ids[k]);
    seen[(size_t)node_ids[k]] = 1;
  }
  m.rx_coord = coordinate;
  return NSM_OK;
}

static int sample_params(const HostMesh& m, int k, const double* density, const double* bulk, const double* shear,
                         const double* mag, HostSample& hs)
{
  hs.rho.assign(density + (size_t)k * m.nb, density + (size_t)(k + 1) * m.nb);
  hs.K.assign(bulk + (size_t)k * m.nb, bulk + (size_t)(k + 1) * m.nb);
  hs.G.assign(shear + (size_t)k * m.nb, shear + (size_t)(k + 1) * m.nb);
  for (int b = 0; b < m.nb; b++)
    if (!(hs.rho[(size_t)b] > 0.0) || !(hs.K[(size_t)b] > 0.0) || !(hs.G[(size_t)b] >= 0.0))
      return fail(NSM_ERR_ARG, "invalid material parameters in block %d of a sample", b);
  const size_t nc = m.bcs.size();
  if (nc > 0 && !mag) return fail(NSM_ERR_ARG, "bc_magnitude is null but the mesh has conditions");
  hs.mag.assign(nc, 0.0);
  for (size_t c = 0; c < nc; c++) hs.mag[c] = mag[(size_t)k * nc + c];
  return NSM_OK;
}

int nsm_ensemble_add_samples(nsm_ensemble* es, int mesh_id, int num_samples, const double* density,
                             const double* bulk_modulus, const double* shear_modulus, const double* bc_magnitude,
                             int* first_sample)
{
  if (!es || !first_sample) return fail(NSM_ERR_ARG, "null argument");
  if (es->ready) return fail(NSM_ERR_STATE, "the ensemble is already set up");
  if (mesh_id < 0 || mesh_id >= (int)es->meshes.size()) return fail(NSM_ERR_ARG, "no mesh %d", mesh_id);
  if (num_samples <= 0) return fail(NSM_ERR_ARG, "need at least one sample");
  if (!density || !bulk_modulus || !shear_modulus) return fail(NSM_ERR_ARG, "null material arrays");
  HostMesh&               m = es->meshes[(size_t)mesh_id];
  std::vector<HostSample> add((size_t)num_samples);
  for (int k = 0; k < num_samples; k++) {
    add[(size_t)k].mesh = mesh_id;
    int st = sample_params(m, k, density, bulk_modulus, shear_modulus, bc_magnitude, add[(size_t)k]);
    if (st) return st;
  }
  *first_sample = (int)es->samples.size();
  for (int k = 0; k < num_samples; k++) m.samples.push_back(*first_sample + k);
  es->samples.insert(es->samples.end(), add.begin(), add.end());
  return NSM_OK;
}

int nsm_ensemble_set_samples(nsm_ensemble* es, int first_sample, int num_samples, const double* density,
                             const double* bulk_modulus, const double* shear_modulus, const double* bc_magnitude)
{
  if (!es || !density || !bulk_modulus || !shear_modulus) return fail(NSM_ERR_ARG, "null argument");
  if (num_samples <= 0 || first_sample < 0 || first_sample + (int64_t)num_samples > es->S())
    return fail(NSM_ERR_ARG, "sample range out of bounds");
  const int mesh = es->samples[(size_t)first_sample].mesh;
  for (int k = 0; k < num_samples; k++)
    if (es->samples[(size_t)(first_sample + k)].mesh != mesh)
      return fail(NSM_ERR_ARG, "the samples of one call must share a mesh");
  const HostMesh&         m = es->meshes[(size_t)mesh];
  std::vector<HostSample> upd((size_t)num_samples);
  for (int k = 0; k < num_samples; k++) {
    upd[(size_t)k].mesh = mesh;
    int st = sample_params(m, k, density, bulk_modulus, shear_modulus, bc_magnitude, upd[(size_t)k]);
    if (st) return st;
  }
  for (int k = 0; k < num_samples; k++) es->samples[(size_t)(first_sample + k)] = upd[(size_t)k];
  if (!es->ready) return NSM_OK;
  ENS_HIP(hipSetDevice(es->device));
  int st = upload_parameters(es);
  return st ? st : compute_mass_and_dt(es);
}

}  // extern "C"

namespace {

constexpr int kTuneSamplesPerMesh = 256;

// Element-tile size of a mixed ensemble chosen by the autotuner from timed
// steps of the same meshes with at most kTuneSamplesPerMesh samples each
// (enough to fill the device), cached per ensemble shape.
int ensemble_tile_elems(nsm_ensemble* es, std::vector<std::pair<int, double>>* timings)
{
  std::string key = "ensemble:mixed";
  for (const auto& m : es->meshes) {
    char b[96];
    snprintf(b, sizeof(b), ":N%lld,E%lld,B%d,S%d", (long long)m.N, (long long)m.E, m.nb,
             (int)std::min<size_t>(m.samples.size(), kTuneSamplesPerMesh));
    key += b;
  }
  auto time_of = [&](int c) -> double {
    nsm_ensemble* tmp = nullptr;
    if (nsm_ensemble_create(NSM_MIXED, c, &tmp)) return -1.0;
    for (const auto& m : es->meshes) {
      HostMesh d;
      d.N        = m.N;
      d.E        = m.E;
      d.nb       = m.nb;
      d.blk_off  = m.blk_off;
      d.X        = m.X;
      d.conn     = m.conn;
      d.bcs      = m.bcs;
      d.rx_nodes = m.rx_nodes;
      d.rx_coord = m.rx_coord;
      tmp->meshes.push_back(std::move(d));
    }
    for (size_t mi = 0; mi < es->meshes.size(); mi++) {
      const auto& ids = es->meshes[mi].samples;
      for (size_t k = 0; k < std::min<size_t>(ids.size(), kTuneSamplesPerMesh); k++) {
        tmp->meshes[mi].samples.push_back((int)tmp->samples.size());
        tmp->samples.push_back(es->samples[(size_t)ids[k]]);
      }
    }
    double sec = -1.0;
    if (!nsm_ensemble_setup(tmp)) {
      std::vector<double> ft(tmp->critical_dt.size());
      for (size_t s = 0; s < ft.size(); s++) ft[s] = 6.0 * 0.5 * tmp->critical_dt[s];
      if (!nsm_ensemble_run(tmp, 2, 0, ft.data(), 0)) {
        const auto t0 = std::chrono::steady_clock::now();
        if (!nsm_ensemble_run(tmp, 6, 0, ft.data(), 0))
          sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() / 6.0;
      }
    }
    nsm_ensemble_destroy(tmp);
    return sec;
  };
  const int c = nsm::autotune_tiles(key, time_of, timings);
  return c ? c : 512;
}

}  // namespace

extern "C" {

int nsm_ensemble_setup(nsm_ensemble* es)
{
  if (!es) return fail(NSM_ERR_ARG, "null ensemble");
  if (es->ready) return fail(NSM_ERR_STATE, "the ensemble is already set up");
  if (es->samples.empty()) return fail(NSM_ERR_STATE, "the ensemble has no samples");
  int rx_coord = -1;
  for (const auto& m : es->meshes)
    if (!m.rx_nodes.empty()) {
      if (rx_coord >= 0 && m.rx_coord != rx_coord)
        return fail(NSM_ERR_ARG, "all meshes' reaction sums must use one coordinate");
      rx_coord = m.rx_coord;
    }
  es->rx_coord = rx_coord;
  ENS_HIP(hipSetDevice(es->device));
  es->tile_elems = es->tile_request                ? es->tile_request
                   : es->precision == NSM_MIXED ? ensemble_tile_elems(es, &es->tune_timings)
                                                  : 0;
  int st = NSM_OK;
  for (auto& m : es->meshes)
    if ((st = prepare_mesh(es, m))) return st;
  if ((st = build_device(es))) {
    free_device(es);
    return st;
  }
  es->ready = true;
  return NSM_OK;
}

int nsm_ensemble_get_critical_time_steps(nsm_ensemble* es, double* dt)
{
  if (!es || !dt) return fail(NSM_ERR_ARG, "null argument");
  if (!es->ready) return fail(NSM_ERR_STATE, "call nsm_ensemble_setup first");
  std::copy(es->critical_dt.begin(), es->critical_dt.end(), dt);
  return NSM_OK;
}

int nsm_ensemble_run(nsm_ensemble* es, int64_t num_steps, int64_t output_frequency, const double* final_time,
                     int use_graph)
{
  if (!es || !final_time) return fail(NSM_ERR_ARG, "null argument");
  if (!es->ready) return fail(NSM_ERR_STATE, "call nsm_ensemble_setup first");
  if (num_steps < 1 || num_steps > (int64_t)INT32_MAX - 2) return fail(NSM_ERR_ARG, "invalid step count");
  if (output_frequency < 0) return fail(NSM_ERR_ARG, "negative output frequency");
  ENS_HIP(hipSetDevice(es->device));
  const int64_t S = es->S();
  // times of every step, as the integrator forms them
  std::vector<nsm::StepTimes> times((size_t)(S * (num_steps + 1)));
  for (int64_t s = 0; s < S; s++) {
    if (!(final_time[s] >= 0.0)) return fail(NSM_ERR_ARG, "invalid final time of sample %lld", (long long)s);
    const double dt_user = (final_time[s] - 0.0) / (double)num_steps;
    double       t       = 0.0;
    for (int64_t n = 0; n <= num_steps; n++) times[(size_t)(s * (num_steps + 1) + n)] = step_times(&t, dt_user);
  }
  bool realloc = false;
  if ((int64_t)times.size() > es->times_cap) {
    if (es->d_times) ENS_HIP(hipFree(es->d_times));
    es->d_times = nullptr;
    ENS_HIP(hipMalloc((void**)&es->d_times, times.size() * sizeof(nsm::StepTimes)));
    es->times_cap = (int64_t)times.size();
    realloc       = true;
  }
  if (S * num_steps > es->hist_cap) {
    if (es->d_hist) ENS_HIP(hipFree(es->d_hist));
    es->d_hist = nullptr;
    ENS_HIP(hipMalloc((void**)&es->d_hist, (size_t)(S * num_steps) * sizeof(double)));
    es->hist_cap = S * num_steps;
    realloc      = true;
  }
  if (realloc) es->drop_graphs();
  ENS_HIP(hipMemcpyAsync(es->d_times, times.data(), times.size() * sizeof(nsm::StepTimes), hipMemcpyHostToDevice,
                         es->stream));
  es->hist_steps = num_steps;
  int st         = NSM_OK;
  if (use_graph && !es->timing) {
    const auto key = std::make_pair(num_steps, output_frequency);
    auto       it  = es->graphs.find(key);
    if (it == es->graphs.end()) {
      hipGraph_t graph = nullptr;
      ENS_HIP(hipStreamBeginCapture(es->stream, hipStreamCaptureModeThreadLocal));
      st = issue_run(es, num_steps, output_frequency);
      hipError_t e = hipStreamEndCapture(es->stream, &graph);
      if (st) {
        if (graph) (void)hipGraphDestroy(graph);
        return st;
      }
      if (e != hipSuccess) return fail(NSM_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
      hipGraphExec_t exec = nullptr;
      e                   = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
      (void)hipGraphDestroy(graph);
      if (e != hipSuccess) return fail(NSM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
      it = es->graphs.emplace(key, exec).first;
    }
    ENS_HIP(hipGraphLaunch(it->second, es->stream));
  } else {
    if ((st = issue_run(es, num_steps, output_frequency))) return st;
  }
  return check_status(es);
}

int nsm_ensemble_sample_reaction_history(nsm_ensemble* es, int num_points, double* out)
{
  if (!es || (num_points > 0 && !out)) return fail(NSM_ERR_ARG, "null argument");
  if (es->hist_steps < 1) return fail(NSM_ERR_STATE, "no run to sample");
  if (num_points < 0) return fail(NSM_ERR_ARG, "negative point count");
  if (num_points == 0) return NSM_OK;
  ENS_HIP(hipSetDevice(es->device));
  const int64_t total = es->S() * num_points;
  if (total > es->qoi_cap) {
    if (es->d_qoi) ENS_HIP(hipFree(es->d_qoi));
    es->d_qoi = nullptr;
    ENS_HIP(hipMalloc((void**)&es->d_qoi, (size_t)total * sizeof(double)));
    es->qoi_cap = total;
  }
  nsm::launch_ens_sample_history(es->d_hist, es->hist_steps, es->hist_steps, (int)es->S(), num_points, es->d_qoi,
                                 es->stream);
  ENS_LAUNCHED();
  ENS_HIP(hipMemcpyAsync(out, es->d_qoi, (size_t)total * sizeof(double), hipMemcpyDeviceToHost, es->stream));
  ENS_HIP(hipStreamSynchronize(es->stream));
  return NSM_OK;
}

int nsm_ensemble_get_reaction_history(nsm_ensemble* es, int sample, int64_t first_step, int64_t count, double* out)
{
  if (!es || (count > 0 && !out)) return fail(NSM_ERR_ARG, "null argument");
  if (sample < 0 || sample >= es->S()) return fail(NSM_ERR_ARG, "no sample %d", sample);
  if (first_step < 0 || count < 0 || first_step + count > es->hist_steps)
    return fail(NSM_ERR_ARG, "steps [%lld, %lld) not recorded", (long long)first_step, (long long)(first_step + count));
  if (count == 0) return NSM_OK;
  ENS_HIP(hipSetDevice(es->device));
  ENS_HIP(hipMemcpyAsync(out, es->d_hist + (int64_t)sample * es->hist_steps + first_step,
                         (size_t)count * sizeof(double), hipMemcpyDeviceToHost, es->stream));
  ENS_HIP(hipStreamSynchronize(es->stream));
  return NSM_OK;
}

int nsm_ensemble_get_field(nsm_ensemble* es, int sample, int field, double* out)
{
  if (!es || !out) return fail(NSM_ERR_ARG, "null argument");
  if (!es->ready) return fail(NSM_ERR_STATE, "call nsm_ensemble_setup first");
  if (sample < 0 || sample >= es->S()) return fail(NSM_ERR_ARG, "no sample %d", sample);
  const double* src = nullptr;
  size_t        w   = 3;
  switch (field) {
    case NSM_FIELD_DISPLACEMENT: src = es->en.u; break;
    case NSM_FIELD_VELOCITY: src = es->en.v; break;
    case NSM_FIELD_ACCELERATION: src = es->en.a; break;
    case NSM_FIELD_INTERNAL_FORCE: src = es->en.f; break;
    case NSM_FIELD_LUMPED_MASS:
      src = es->d_mass;
      w   = 1;
      break;
    default: return fail(NSM_ERR_ARG, "field %d is not an ensemble node field", field);
  }
  ENS_HIP(hipSetDevice(es->device));
  const HostMesh& m    = es->meshes[(size_t)es->samples[(size_t)sample].mesh];
  int64_t         base = 0;
  for (int s = 0; s < sample; s++) base += es->meshes[(size_t)es->samples[(size_t)s].mesh].N;
  std::vector<double> tmp(w * (size_t)m.N);
  ENS_HIP(hipMemcpyAsync(tmp.data(), src + w * (size_t)base, tmp.size() * sizeof(double), hipMemcpyDeviceToHost,
                         es->stream));
  ENS_HIP(hipStreamSynchronize(es->stream));
  for (int64_t n = 0; n < m.N; n++)
    for (size_t c = 0; c < w; c++) out[(size_t)n * w + c] = tmp[(size_t)m.dev_of[(size_t)n] * w + c];
  return NSM_OK;
}

int nsm_ensemble_info(nsm_ensemble* es, int64_t* num_samples, int64_t* num_nodes, int64_t* num_elem,
                      int64_t* work_items, int* tile_elems, int64_t* device_bytes)
{
  if (!es) return fail(NSM_ERR_ARG, "null ensemble");
  if (num_samples) *num_samples = es->S();
  if (num_nodes) *num_nodes = es->N_total;
  if (num_elem) *num_elem = es->E_total;
  if (work_items) *work_items = es->precision == NSM_MIXED ? es->num_tile_work : es->num_elem_chunks;
  if (tile_elems) *tile_elems = es->precision == NSM_MIXED ? es->tile_elems : 0;
  if (device_bytes) *device_bytes = (int64_t)es->device_bytes;
  return NSM_OK;
}

int nsm_ensemble_autotune_timings(nsm_ensemble* es, int capacity, int* count, int* tile_elems,
                                  double* seconds_per_step)
{
  if (!es || !count) return fail(NSM_ERR_ARG, "null argument");
  *count = (int)es->tune_timings.size();
  for (int k = 0; k < std::min(capacity, *count); k++) {
    if (tile_elems) tile_elems[k] = es->tune_timings[(size_t)k].first;
    if (seconds_per_step) seconds_per_step[k] = es->tune_timings[(size_t)k].second;
  }
  return NSM_OK;
}

int nsm_ensemble_enable_kernel_timing(nsm_ensemble* es, int enable)
{
  if (!es) return fail(NSM_ERR_ARG, "null ensemble");
  ENS_HIP(hipSetDevice(es->device));
  ENS_HIP(hipStreamSynchronize(es->stream));
  es->timing     = enable != 0;
  es->event_next = 0;
  es->timed.clear();
  return NSM_OK;
}

int nsm_ensemble_kernel_time(nsm_ensemble* es, int kernel, double* avg_ms, int64_t* launches)
{
  if (!es || !avg_ms || !launches) return fail(NSM_ERR_ARG, "null argument");
  ENS_HIP(hipSetDevice(es->device));
  ENS_HIP(hipStreamSynchronize(es->stream));
  double  sum = 0.0;
  int64_t n   = 0;
  for (const auto& t : es->timed)
    if (t.first == kernel) {
      float ms = 0.f;
      ENS_HIP(hipEventElapsedTime(&ms, es->event_pool[t.second].start, es->event_pool[t.second].stop));
      sum += ms;
      n++;
    }
  *avg_ms   = n ? sum / (double)n : 0.0;
  *launches = n;
  return NSM_OK;
}

}  // extern "C"
