// C entry points over the host tile builder (nimblesm_amd/csrc/nsm_tiles.cpp)
// so the CPU tests can check its invariants without a GPU.  Built by
// tests/test_tiles.py with g++ into a temporary directory.
#include <cstring>

#include "../nimblesm_amd/csrc/nsm_tiles.h"

template <typename T>
static void put(T* dst, const std::vector<T>& v)
{
  if (!v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(T));
}

extern "C" {

void* tiles_build(int64_t num_nodes, const double* X, const int* conn, int num_blocks, const int64_t* blk_off,
                  int tile_elems)
{
  auto*                ts = new nsm::TileSet();
  std::vector<int64_t> off(blk_off, blk_off + num_blocks + 1);
  std::string          err;
  if (!nsm::build_tiles(num_nodes, X, conn, off, tile_elems, *ts, &err)) {
    delete ts;
    return nullptr;
  }
  return ts;
}

// num_tiles, tile-node slots, boundary nodes, boundary partial entries, interior nodes
void tiles_sizes(const void* h, int64_t out[5])
{
  const auto* ts = static_cast<const nsm::TileSet*>(h);
  out[0]         = ts->num_tiles;
  out[1]         = (int64_t)ts->nodes.size();
  out[2]         = (int64_t)ts->bnd_nodes.size();
  out[3]         = (int64_t)ts->bnd_ent.size();
  out[4]         = ts->num_interior;
}

void tiles_copy(const void* h, int64_t* perm, int32_t* elem_off, int32_t* node_off, int32_t* nodes,
                uint8_t* node_info, uint16_t* lconn, uint8_t* color, uint8_t* num_colors, int32_t* bnd_nodes,
                int32_t* bnd_ptr, int32_t* bnd_ent, int32_t* slot_dst, int32_t* node_new, int32_t* int_off,
                uint8_t* gather)
{
  const auto* ts = static_cast<const nsm::TileSet*>(h);
  put(perm, ts->perm);
  put(elem_off, ts->elem_off);
  put(node_off, ts->node_off);
  put(nodes, ts->nodes);
  put(node_info, ts->node_info);
  put(lconn, ts->lconn);
  put(color, ts->color);
  put(num_colors, ts->num_colors);
  put(bnd_nodes, ts->bnd_nodes);
  put(bnd_ptr, ts->bnd_ptr);
  put(bnd_ent, ts->bnd_ent);
  put(slot_dst, ts->slot_dst);
  put(node_new, ts->node_new);
  put(int_off, ts->int_off);
  put(gather, ts->gather);
}

void tiles_free(void* h) { delete static_cast<nsm::TileSet*>(h); }
}
