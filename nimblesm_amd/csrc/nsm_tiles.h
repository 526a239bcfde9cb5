// nsm_tiles.h -- element tiles with tile-local force assembly (mixed path).
//
// Elements of each block are put in Morton order of their centroids and cut
// into tiles of at most tile_elems elements (512 by default; 64, 128, 256) touching
// at most 2*tile_elems distinct nodes.  A node whose incident elements all lie in one tile is
// "interior" to it: that tile sums its force completely and performs its
// central-difference update in the element kernel.  Every other node gets one
// partial sum per tile that touches it; the boundary kernel adds those
// partials in ascending tile order.  Inside a tile, a node's force is the sum
// of its element contributions in ascending corner index when no two elements
// of the tile hold that node at the same corner index (every tile of a
// structured mesh; TileSet::gather), otherwise the sum over element colours
// in ascending colour (no two elements of a colour share a node).  Either
// way the summation order is fixed and runs are reproducible.
#pragma once

#include <stdint.h>

#include <string>
#include <vector>

namespace nsm {

constexpr int     kTileElems    = 512;  // default elements per tile (one thread each)
constexpr int     kMaxColors    = 32;
constexpr uint8_t kColorMask    = 0x1f;  // colour bits of TileSet::color
constexpr uint8_t kRepeatedNode = 0x80;  // element lists one node more than once
constexpr uint8_t kInteriorNode = 1;

struct TileSet {
  int                   tile_elems = kTileElems;  // nodes per tile <= 2 * tile_elems
  int                   num_tiles = 0;
  std::vector<int64_t>  perm;           // internal element -> original element
  std::vector<int32_t>  elem_off;       // [T+1] internal element range of each tile
  std::vector<int32_t>  node_off;       // [T+1] range into nodes
  std::vector<int32_t>  nodes;          // global node id per tile-node slot (ascending per tile)
  std::vector<uint8_t>  node_info;      // kInteriorNode per tile-node slot
  std::vector<uint16_t> lconn;          // [E][8] local node slot of each element node (internal order)
  std::vector<uint8_t>  color;          // [E] colour | kRepeatedNode (internal order)
  std::vector<uint8_t>  num_colors;     // [T]
  std::vector<uint8_t>  gather;         // [T] 1: no two elements of the tile share a node at the same
                                        // corner index, so each node's forces can be gathered per corner
  std::vector<int32_t>  bnd_nodes;      // boundary nodes, ascending
  std::vector<int32_t>  bnd_ptr;        // [B+1]
  std::vector<int32_t>  bnd_ent;        // tile-node slot of each partial, ascending tile
  std::vector<int32_t>  slot_dst;       // per tile-node slot: -1 interior, else its partial's index p
                                        // (bnd_ent[p] == slot; partials of a node are contiguous)
  int64_t               num_interior = 0;
  // Node numbering of the tiled model (every node id above is a new id):
  // tile t's interior nodes are [int_off[t], int_off[t+1]); boundary node b
  // (bnd_nodes[b]) is num_interior + b.  Inside a tile the interior slots
  // come first.
  std::vector<int32_t>  node_new;       // [N] original id -> new id
  std::vector<int32_t>  node_old;       // [N] new id -> original id
  std::vector<int32_t>  int_off;        // [T+1]
};

// Builds the tiles of blocks given by element offsets blk_off (conn in
// original element and node numbering, [E][8]) and renumbers the nodes (see
// TileSet::node_new).  Returns false with a message when a tile
// cannot be coloured with kMaxColors colours.
bool build_tiles(int64_t num_nodes, const double* X, const int* conn, const std::vector<int64_t>& blk_off,
                 int tile_elems, TileSet& ts, std::string* err);

}  // namespace nsm
