// nsm_tiles.cpp -- host construction of the element tiles (see nsm_tiles.h).
#include "nsm_tiles.h"

#include <algorithm>
#include <cmath>
#include <cstddef>
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

void renumber_nodes(int64_t num_nodes, TileSet& ts);

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
      std::vector<uint8_t>  corners(tnodes.size(), 0);  // corner indices at which a slot occurs
      std::vector<int32_t>  cnt(tnodes.size(), 0);
      int                   ncol   = 0;
      bool                  gather = true;
      for (int64_t p = tstart; p < pos; p++) {
        const int64_t e    = ts.perm[(size_t)p];
        uint32_t      busy = 0u;
        for (int j = 0; j < 8; j++) {
          const int s = slot_of[(size_t)conn[8 * e + j]];
          ts.lconn[(size_t)(8 * p + j)] = (uint16_t)s;
          busy |= used[(size_t)s];
          cnt[(size_t)s]++;
          gather = gather && !(corners[(size_t)s] >> j & 1);
          corners[(size_t)s] |= (uint8_t)(1u << j);
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
      // Elements of the tile in colour order (stable): a 64-lane wave then
      // holds few colours and skips the assembly passes of the others.
      {
        const size_t           n0 = (size_t)tstart, n = (size_t)(pos - tstart);
        std::vector<size_t>    idx(n);
        for (size_t i = 0; i < n; i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) {
          return (ts.color[n0 + a] & kColorMask) < (ts.color[n0 + b] & kColorMask);
        });
        std::vector<int64_t>  perm2(n);
        std::vector<uint8_t>  color2(n);
        std::vector<uint16_t> lconn2(8 * n);
        for (size_t i = 0; i < n; i++) {
          perm2[i]  = ts.perm[n0 + idx[i]];
          color2[i] = ts.color[n0 + idx[i]];
          for (int j = 0; j < 8; j++) lconn2[8 * i + j] = ts.lconn[8 * (n0 + idx[i]) + j];
        }
        std::copy(perm2.begin(), perm2.end(), ts.perm.begin() + (std::ptrdiff_t)n0);
        std::copy(color2.begin(), color2.end(), ts.color.begin() + (std::ptrdiff_t)n0);
        std::copy(lconn2.begin(), lconn2.end(), ts.lconn.begin() + (std::ptrdiff_t)(8 * n0));
      }
      for (size_t i = 0; i < tnodes.size(); i++) {
        ts.nodes.push_back(tnodes[i]);
        in_tile_cnt.push_back(cnt[i]);
        slot_of[(size_t)tnodes[i]] = -1;
      }
      ts.num_colors.push_back((uint8_t)ncol);
      ts.gather.push_back((uint8_t)gather);
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
  renumber_nodes(num_nodes, ts);
  return true;
}

// Node numbering of the tiled model: the interior nodes of tile 0, of tile 1,
// ... (each tile's in slot order), then the boundary nodes in ascending
// original id.  Within every tile the interior slots move to the front.  All
// node ids of the TileSet become new ids; node_new / node_old map between
// the two numberings.
void renumber_nodes(int64_t num_nodes, TileSet& ts)
{
  ts.node_new.assign((size_t)num_nodes, -1);
  int32_t next = 0;
  ts.int_off.assign((size_t)ts.num_tiles + 1, 0);
  for (int t = 0; t < ts.num_tiles; t++) {
    ts.int_off[(size_t)t] = next;
    for (int32_t sl = ts.node_off[(size_t)t]; sl < ts.node_off[(size_t)t + 1]; sl++)
      if (ts.node_info[(size_t)sl] & kInteriorNode) ts.node_new[(size_t)ts.nodes[(size_t)sl]] = next++;
  }
  ts.int_off[(size_t)ts.num_tiles] = next;
  for (size_t b = 0; b < ts.bnd_nodes.size(); b++) ts.node_new[(size_t)ts.bnd_nodes[b]] = next++;
  ts.node_old.assign((size_t)num_nodes, 0);
  for (int64_t n = 0; n < num_nodes; n++) ts.node_old[(size_t)ts.node_new[(size_t)n]] = (int32_t)n;

  // interior slots first inside each tile; remap the local connectivity
  std::vector<int32_t> nodes2(ts.nodes.size());
  std::vector<uint8_t> info2(ts.nodes.size());
  std::vector<int32_t> global_slot(ts.nodes.size());  // old global slot -> new global slot
  for (int t = 0; t < ts.num_tiles; t++) {
    const int32_t s0 = ts.node_off[(size_t)t], s1 = ts.node_off[(size_t)t + 1];
    int32_t       d  = s0;
    for (int pass = 0; pass < 2; pass++)
      for (int32_t sl = s0; sl < s1; sl++)
        if (((ts.node_info[(size_t)sl] & kInteriorNode) != 0) == (pass == 0)) {
          global_slot[(size_t)sl] = d;
          nodes2[(size_t)d]       = ts.node_new[(size_t)ts.nodes[(size_t)sl]];
          info2[(size_t)d]        = ts.node_info[(size_t)sl];
          d++;
        }
    for (int32_t e = ts.elem_off[(size_t)t]; e < ts.elem_off[(size_t)t + 1]; e++)
      for (int j = 0; j < 8; j++) {
        uint16_t& l = ts.lconn[(size_t)(8 * e + j)];
        l           = (uint16_t)(global_slot[(size_t)(s0 + l)] - s0);
      }
  }
  ts.nodes.swap(nodes2);
  ts.node_info.swap(info2);
  for (auto& s : ts.bnd_ent) s = global_slot[(size_t)s];  // still ascending tile order per node
  for (auto& n : ts.bnd_nodes) n = ts.node_new[(size_t)n];
  ts.slot_dst.assign(ts.nodes.size(), -1);
  for (size_t p = 0; p < ts.bnd_ent.size(); p++) ts.slot_dst[(size_t)ts.bnd_ent[p]] = (int32_t)p;
}

}  // namespace nsm
