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
// mixed precision packs a particle's cell in 11/11/10 bits (pack_cell,
// wf_device.hpp), so its grids hold at most these cells per axis
constexpr int kPackedCells[3] = {2048, 2048, 1024};

// The cell and key functions below are the one definition the search's
// counting sort (wf_nsearch.hip) and the tile builder (wf_tiles.hip) share:
// the lists are right only if both bin a position into the same cell and
// order cells by the same key (the 8^3 meta-block layout of
// cuNSearchKernels.cu:10-21, the bits interleaved x, y, z from bit 0).
__device__ __forceinline__ int morton3(int x, int y, int z) {
  int m = 0;
#pragma unroll
  for (int b = 0; b < 3; b++)
    m |= (((x >> b) & 1) << (3 * b)) | (((y >> b) & 1) << (3 * b + 1)) |
         (((z >> b) & 1) << (3 * b + 2));
  return m;
}
__device__ __forceinline__ int cell_key(const GridDesc& g, int cx, int cy, int cz) {
  int meta = ((cz >> 3) * g.mdim[1] + (cy >> 3)) * g.mdim[0] + (cx >> 3);
  return meta * 512 + morton3(cx & 7, cy & 7, cz & 7);
}
// cell of an FP32 position (FP64 arithmetic), clamped into the grid
__device__ __forceinline__ void cell_of(const GridDesc& g, float px, float py, float pz, int& cx,
                                        int& cy, int& cz) {
  cx = (int)floor(((double)px - g.org[0]) * g.inv_cs);
  cy = (int)floor(((double)py - g.org[1]) * g.inv_cs);
  cz = (int)floor(((double)pz - g.org[2]) * g.inv_cs);
  cx = min(max(cx, 0), g.dim[0] - 1);
  cy = min(max(cy, 0), g.dim[1] - 1);
  cz = min(max(cz, 0), g.dim[2] - 1);
}

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

// tiles/ntiles: the tiles to build (tiles null: tiles 0..ntiles-1);
// sort_own: the tiles' particles are not in cell order (no reorder since the
// last build)
void launch_tile_cells(hipStream_t s, const float* const pf[3], int n, const GridDesc& g,
                       const int* cell_start, const int* tiles, int ntiles, bool sort_own,
                       int* tile_keys, int* tile_nc, int* tile_ucnt, int* overflow);
// pf: cell-sorted positions; own (may be null): positions by storage index
// and map (may be null) the storage index of each sorted position, for a
// build without a reorder; a list holds at most cap entries, and tile_over[t]
// gets the number of the tile's particles with more than rcap neighbours;
// a tile's adjacent-cell set is compacted to the particles its lists
// reference when it holds at most compact_max (0: 4096) candidates
void launch_tile_lists(hipStream_t s, const float* const pf[3], const float* const own[3],
                       const int* map, const int* tiles, int ntiles, int n, const GridDesc& g,
                       const int* cell_start, const int* tile_keys, const int* tile_nc,
                       const int* tile_uoff, float r2, int cap, int rcap, int64_t npad,
                       int* count, unsigned short* slot, int* tile_j, int* tile_uc,
                       int* tile_over, int* overflow, int compact_max);

}  // namespace wf
