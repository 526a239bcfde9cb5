// nsm_tiles.cpp -- host construction of the element tiles (see nsm_tiles.h).
#include "nsm_tiles.h"

#include <algorithm>
#include <cmath>
#include <string>

namespace nsm {
namespace {

// 21-bit coordinate -> every third bit of a 63-bit Morton code
uint64_t spread3(uint64_t v)
{
  v &= 0x1fffffULL;
  v = (v | v << 32) & 0x1f00000000ffffULL;
  v = (v | v << 16) & 0x1f0000ff0000ffULL;
  v = (v | v << 8) & 0x100f00f00f00f00fULL;
  v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
  v = (v | v << 2) & 0x1249249249249249ULL;
  return v;
}

}  // namespace

bool build_tiles(int64_t num_nodes, const double* X, const int* conn, const std::vector<int64_t>& blk_off,
                 int tile_elems, TileSet& ts, std::string* err)
{
  const int64_t E = blk_off.back();
  ts              = TileSet();
  ts.tile_elems   = tile_elems;
  const int tile_nodes = 2 * tile_elems;
  ts.perm.reserve((size_t)E);

  // Morton order of element centroids within each block.  Centroids are
  // quantized to cells of the mean element size h = (bounding-box volume /
  // element count)^(1/3) measured from the block's node bounding box, so on a
  // (jittered) structured grid every element gets its own lattice cell and
  // tiles of 2^k consecutive codes are aligned boxes.
  for (size_t b = 0; b + 1 < blk_off.size(); b++) {
    const int64_t e0 = blk_off[b], e1 = blk_off[b + 1];
    if (e1 <= e0) continue;
    std::vector<double> cen((size_t)(e1 - e0) * 3);
    double              lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t e = e0; e < e1; e++)
      for (int d = 0; d < 3; d++) {
        double s = 0.0;
        for (int j = 0; j < 8; j++) {
          const double x = X[3 * (int64_t)conn[8 * e + j] + d];
          s += x;
          lo[d] = std::min(lo[d], x);
          hi[d] = std::max(hi[d], x);
        }
        cen[(size_t)(e - e0) * 3 + d] = s * 0.125;
      }
    const double ext[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    const double vol    = ext[0] * ext[1] * ext[2];
    double       h      = vol > 0.0 ? std::cbrt(vol / (double)(e1 - e0)) : 0.0;
    if (!(h > 0.0)) h = std::max(std::max(ext[0], ext[1]), std::max(ext[2], 1e-300)) / std::cbrt((double)(e1 - e0));
    std::vector<std::pair<uint64_t, int64_t>> key((size_t)(e1 - e0));
    for (int64_t e = e0; e < e1; e++) {
      uint64_t code = 0;
      for (int d = 0; d < 3; d++) {
        const double   x = (cen[(size_t)(e - e0) * 3 + d] - lo[d]) / h;
        const uint64_t q = (uint64_t)std::min(std::max(std::floor(x), 0.0), 2097151.0);
        code |= spread3(q) << d;
      }
      key[(size_t)(e - e0)] = {code, e};
    }
    std::sort(key.begin(), key.end());
    for (const auto& k : key) ts.perm.push_back(k.second);
  }

  // greedy tiles (never crossing a block boundary), local node slots, colours
  std::vector<int>      slot_of((size_t)num_nodes, -1);
  std::vector<int32_t>  deg((size_t)num_nodes, 0);
  for (int64_t k = 0; k < 8 * E; k++) deg[(size_t)conn[k]]++;
  std::vector<int32_t>  in_tile_cnt;  // incidences of each tile-node slot inside its tile
  ts.elem_off.push_back(0);
  ts.node_off.push_back(0);
  ts.lconn.resize((size_t)(8 * E));
  ts.color.resize((size_t)E);
  int64_t pos = 0;
  for (size_t b = 0; b + 1 < blk_off.size(); b++) {
    const int64_t e1 = blk_off[b + 1];
    while (pos < e1) {
      std::vector<int32_t> tnodes;
      const int64_t        tstart = pos;
      while (pos < e1 && pos - tstart < tile_elems) {
        const int64_t e     = ts.perm[(size_t)pos];
        int           fresh = 0;
        for (int j = 0; j < 8; j++) {
          const int n = conn[8 * e + j];
          if (slot_of[(size_t)n] < 0) {
            bool dup = false;  // a node repeated within the element counts once
            for (int i = 0; i < j; i++) dup = dup || conn[8 * e + i] == n;
            fresh += !dup;
          }
        }
        if ((int64_t)tnodes.size() + fresh > tile_nodes) break;
        for (int j = 0; j < 8; j++) {
          const int n = conn[8 * e + j];
          if (slot_of[(size_t)n] < 0) {
            slot_of[(size_t)n] = 0;
            tnodes.push_back(n);
          }
        }
        pos++;
      }
      std::sort(tnodes.begin(), tnodes.end());
      for (size_t i = 0; i < tnodes.size(); i++) slot_of[(size_t)tnodes[i]] = (int)i;
      std::vector<uint32_t> used(tnodes.size(), 0u);  // colours used at each node slot
      std::vector<int32_t>  cnt(tnodes.size(), 0);
      int                   ncol = 0;
      for (int64_t p = tstart; p < pos; p++) {
        const int64_t e    = ts.perm[(size_t)p];
        uint32_t      busy = 0u;
        for (int j = 0; j < 8; j++) {
          const int s = slot_of[(size_t)conn[8 * e + j]];
          ts.lconn[(size_t)(8 * p + j)] = (uint16_t)s;
          busy |= used[(size_t)s];
          cnt[(size_t)s]++;
        }
        int c = 0;
        while (c < kMaxColors && (busy >> c & 1u)) c++;
        if (c == kMaxColors) {
          if (err) *err = "element tile needs more than 32 colours";
          return false;
        }
        bool repeated = false;
        for (int j = 0; j < 8; j++) {
          used[(size_t)slot_of[(size_t)conn[8 * e + j]]] |= 1u << c;
          for (int i = 0; i < j; i++) repeated = repeated || conn[8 * e + i] == conn[8 * e + j];
        }
        ts.color[(size_t)p] = (uint8_t)(c | (repeated ? kRepeatedNode : 0));
        ncol                = std::max(ncol, c + 1);
      }
      for (size_t i = 0; i < tnodes.size(); i++) {
        ts.nodes.push_back(tnodes[i]);
        in_tile_cnt.push_back(cnt[i]);
        slot_of[(size_t)tnodes[i]] = -1;
      }
      ts.num_colors.push_back((uint8_t)ncol);
      ts.elem_off.push_back((int32_t)pos);
      ts.node_off.push_back((int32_t)ts.nodes.size());
    }
  }
  ts.num_tiles = (int)ts.num_colors.size();

  // interior: every incidence of the node lies in the tile
  ts.node_info.assign(ts.nodes.size(), 0);
  std::vector<uint8_t> interior_somewhere((size_t)num_nodes, 0);
  for (size_t s = 0; s < ts.nodes.size(); s++)
    if (in_tile_cnt[s] == deg[(size_t)ts.nodes[s]]) {
      ts.node_info[s]                         = kInteriorNode;
      interior_somewhere[(size_t)ts.nodes[s]] = 1;
      ts.num_interior++;
    }
  // boundary nodes (including nodes no element touches) and their partials,
  // in ascending tile order
  std::vector<int32_t> nparts((size_t)num_nodes, 0);
  for (size_t s = 0; s < ts.nodes.size(); s++)
    if (!ts.node_info[s]) nparts[(size_t)ts.nodes[s]]++;
  std::vector<int32_t> first((size_t)num_nodes, -1);
  ts.bnd_ptr.push_back(0);
  for (int64_t n = 0; n < num_nodes; n++) {
    if (interior_somewhere[(size_t)n]) continue;
    first[(size_t)n] = (int32_t)ts.bnd_nodes.size();
    ts.bnd_nodes.push_back((int32_t)n);
    ts.bnd_ptr.push_back(ts.bnd_ptr.back() + nparts[(size_t)n]);
  }
  ts.bnd_ent.assign((size_t)ts.bnd_ptr.back(), 0);
  std::vector<int32_t> fill(ts.bnd_ptr.begin(), ts.bnd_ptr.end() - 1);
  for (size_t s = 0; s < ts.nodes.size(); s++)  // slots are in ascending tile order
    if (!ts.node_info[s]) {
      const int32_t bi         = first[(size_t)ts.nodes[s]];
      ts.bnd_ent[(size_t)fill[(size_t)bi]++] = (int32_t)s;
    }
  return true;
}

}  // namespace nsm
