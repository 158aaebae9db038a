// bl_tiles.h -- the node tiling of the tile step (bl_tiles.hip, bl_fused.hip).
#pragma once
#include <stdint.h>

namespace bl {

struct Handle;

constexpr int kTileNodes = 255;   // nodes per tile (an 8-bit slot, kNoSlot free)
constexpr int kTileEls = 256;     // elements around a tile's nodes: one per thread
constexpr int kTileTarget = 180;  // the element count the bisection aims at
constexpr uint32_t kNoSlot = 0xFF;
// an entry's flag: one of the tile's own elements (the tile completes it and
// writes its outputs); the other entries are the ring around its nodes
constexpr uint32_t kTileOwn = 1;

// Cut the nodes of h's mesh into tiles (h->tile_state 1), or find that the
// mesh has a node that is the same corner of two elements (h->tile_state
// -1, no tiling); ielnd and the node positions come from the registry.
int build_tiles(Handle* h);
void free_tiles(Handle* h);

}  // namespace bl
