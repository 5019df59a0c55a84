// wf_tiles.hpp — grid description and LDS-tile neighbour layout shared by the
// search (wf_nsearch.hip), the tile build (wf_tiles.hip) and the pair kernels.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/wf_abi.h"

namespace wf {

constexpr int kTileP = 128;         // particles per tile = threads per pair workgroup
constexpr int kTileMaxCells = 1024; // occupied neighbour cells per tile
// quads of 16-bit slots per particle (list capacity of either rebuild policy)
constexpr int kSlotQuads = (WF_MAX_NEIGHBORS_SKIN + 3) / 4;

// Uniform search grid: cell (cx,cy,cz) = floor((p - org) * inv_cs); cell key
// = 8^3 meta-block index * 512 + Morton code inside the block.
struct GridDesc {
  double org[3];
  double inv_cs;
  int dim[3];
  int mdim[3];
};

// What the pair kernels read.  slotq holds four 16-bit staged-set positions
// per 8 bytes: entry k of particle i is 16-bit lane (k&3) of
// slotq[(k>>2)*npad + i]; rows are in ascending staged (= particle) order.
struct TileLists {
  const int* count;
  const uint2* slotq;
  const int* uoff;  // [ntiles+1] offsets of each tile's region of tj
  const int* ucnt;  // [ntiles] size of each tile's staged set (a prefix of its region)
  const int* tj;    // staged particle indices
  int64_t npad;
};

void launch_tile_cells(hipStream_t s, const float* const pf[3], int n, const GridDesc& g,
                       const int* cell_start, int* tile_keys, int* tile_nc, int* tile_ucnt,
                       int* overflow);
void launch_tile_lists(hipStream_t s, const float* const pf[3], int n, const GridDesc& g,
                       const int* cell_start, const int* tile_keys, const int* tile_nc,
                       const int* tile_uoff, float r2, int cap, int64_t npad, int* count,
                       unsigned short* slot, int* tile_j, int* tile_uc, int* over_cap);

}  // namespace wf
