// wf_tiles.hip — LDS-staged neighbour tiles.
//
// A tile is kTileP consecutive particles of the Morton order (one workgroup
// of the pair kernels).  At each rebuild the tile gets:
//   - the sorted set of occupied grid cells adjacent to any of its particles
//     (tile_keys); the particles of those cells (in cell order, ascending
//     particle index within a cell) that are a neighbour of at least one of
//     the tile's particles form the tile's staged set (tile_j, tile_uc
//     entries) — on a cubic lattice with h = 1.2 dx about 740 of the ~1,230
//     particles of the adjacent cells;
//   - for each of its particles, the neighbour list in ascending staged
//     order as 16-bit positions in the staged set, four per 8-byte word,
//     interleaved [(k/4)*npad + i] so a wavefront reads contiguous words (the
//     particle indices are tile_j of those positions).
// The pair kernels write the staged set to LDS (position + index per slot)
// and read a neighbour's position from LDS through its slot.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>

#include "wf_device.hpp"
#include "wf_tiles.hpp"

namespace wf {

constexpr int kSortN = 4096;  // >= kTileP * 27, power of two
static_assert(kTileP * 27 <= kSortN, "sort buffer too small");

// The cell coordinates of a cell key (the inverse of cell_key).
__device__ __forceinline__ void key_cell(const GridDesc& g, int key, int& cx, int& cy, int& cz) {
  const int meta = key >> 9, m = key & 511;
  int x = 0, y = 0, z = 0;
#pragma unroll
  for (int b = 0; b < 3; b++) {
    x |= ((m >> (3 * b)) & 1) << b;
    y |= ((m >> (3 * b + 1)) & 1) << b;
    z |= ((m >> (3 * b + 2)) & 1) << b;
  }
  const int mx = meta % g.mdim[0], myz = meta / g.mdim[0];
  cx = 8 * mx + x;
  cy = 8 * (myz % g.mdim[1]) + y;
  cz = 8 * (myz / g.mdim[1]) + z;
}

// Per tile: sorted unique occupied neighbour cells and the staged count.
// After a reorder the tile's particles are in cell order, so its own cells
// are the runs of equal keys; a tile re-listed without a reorder
// (SORT_OWN) first sorts its particles' keys.  Only the tile's few distinct
// cells (typically 10-20) are expanded to their 27 neighbours, and the
// candidate keys are sorted in the smallest power of two that holds them.
// tiles: the tiles to build (null: block b builds tile b).
template <bool SORT_OWN>
__global__ __launch_bounds__(kTileP) void k_tile_cells(const float* __restrict__ p0,
                                                       const float* __restrict__ p1,
                                                       const float* __restrict__ p2, int n,
                                                       GridDesc g,
                                                       const int* __restrict__ cell_start,
                                                       const int* __restrict__ tiles,
                                                       int* tile_keys, int* tile_nc,
                                                       int* tile_ucnt, int* overflow) {
  __shared__ int keys[kSortN];
  __shared__ int segcnt[kTileP];
  __shared__ int own[kTileP];   // the key of each particle's cell
  __shared__ int wcnt[kTileP / 64 + 1];
  const int t = tiles ? tiles[blockIdx.x] : blockIdx.x;
  const int tid = threadIdx.x;
  const int i = t * kTileP + tid;
  int cx = 0, cy = 0, cz = 0, kown = INT_MAX;
  if (i < n) {
    cell_of(g, p0[i], p1[i], p2[i], cx, cy, cz);
    kown = cell_key(g, cx, cy, cz);
  }
  own[tid] = kown;
  __syncthreads();
  if constexpr (SORT_OWN) {
    for (int size = 2; size <= kTileP; size <<= 1) {
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        if (tid < kTileP / 2) {
          const int lo = 2 * tid - (tid & (stride - 1));
          const int hi = lo + stride;
          const bool up = (lo & size) == 0;
          const int a = own[lo], b = own[hi];
          if ((a > b) == up) {
            own[lo] = b;
            own[hi] = a;
          }
        }
        __syncthreads();
      }
    }
    kown = own[tid];
    if (kown != INT_MAX) key_cell(g, kown, cx, cy, cz);
  }
  // first particle of each run of equal cells (keys ascend along the tile)
  const bool first = kown != INT_MAX && (tid == 0 || own[tid - 1] != kown);
  const unsigned long long bal = __ballot(first);
  const int lane = lane_id(), w = tid >> 6;
  if (lane == 0) wcnt[w] = __popcll(bal);
  __syncthreads();
  int wbase = 0;
  for (int u = 0; u < w; u++) wbase += wcnt[u];
  int nown = 0;
  for (int u = 0; u < kTileP / 64; u++) nown += wcnt[u];
  const int rank = wbase + __popcll(bal & ((1ull << lane) - 1ull));
  // the coordinates of the tile's unique cells, in key order
  __shared__ int ocx[kTileP], ocy[kTileP], ocz[kTileP];
  if (first) {
    ocx[rank] = cx;
    ocy[rank] = cy;
    ocz[rank] = cz;
  }
  __syncthreads();
  const int total = 27 * nown;
  int S = 2;
  while (S < total) S <<= 1;
  for (int q = tid; q < S; q += kTileP) {
    int k = INT_MAX;
    if (q < total) {
      const int o = q / 27, r = q - 27 * o;
      const int x = ocx[o] + (r % 3) - 1, y = ocy[o] + ((r / 3) % 3) - 1, z = ocz[o] + r / 9 - 1;
      if (x >= 0 && y >= 0 && z >= 0 && x < g.dim[0] && y < g.dim[1] && z < g.dim[2]) {
        const int kk = cell_key(g, x, y, z);
        if (cell_start[kk + 1] > cell_start[kk]) k = kk;
      }
    }
    keys[q] = k;
  }
  __syncthreads();
  // bitonic sort of the S candidate keys
  for (int size = 2; size <= S; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int s2 = tid; s2 < S / 2; s2 += kTileP) {
        int lo = 2 * s2 - (s2 & (stride - 1));
        int hi = lo + stride;
        bool up = ((lo & size) == 0);
        int a = keys[lo], b = keys[hi];
        if ((a > b) == up) {
          keys[lo] = b;
          keys[hi] = a;
        }
      }
      __syncthreads();
    }
  }
  // unique: each thread owns a segment of ceil(S/kTileP) entries
  const int seg = (S + kTileP - 1) / kTileP;
  const int sb0 = min(tid * seg, S), sb1 = min(sb0 + seg, S);
  int c = 0;
  for (int s = sb0; s < sb1; s++) {
    int k = keys[s];
    if (k != INT_MAX && (s == 0 || keys[s - 1] != k)) c++;
  }
  segcnt[tid] = c;
  __syncthreads();
  if (tid == 0) {
    int run = 0;
    for (int u = 0; u < kTileP; u++) {
      int v = segcnt[u];
      segcnt[u] = run;
      run += v;
    }
    tile_nc[t] = run;
    if (run > kTileMaxCells) atomicAdd(overflow, 1);
  }
  __syncthreads();
  int pos = segcnt[tid];
  int* out = tile_keys + (int64_t)t * kTileMaxCells;
  for (int s = sb0; s < sb1; s++) {
    int k = keys[s];
    if (k != INT_MAX && (s == 0 || keys[s - 1] != k)) {
      if (pos < kTileMaxCells) out[pos] = k;
      pos++;
    }
  }
  __syncthreads();
  // staged particle count
  int nc = min(tile_nc[t], kTileMaxCells);
  int cnt = 0;
  for (int p = tid; p < nc; p += kTileP) {
    int k = out[p];
    cnt += cell_start[k + 1] - cell_start[k];
  }
  segcnt[tid] = cnt;
  __syncthreads();
  if (tid == 0) {
    int tot = 0;
    for (int u = 0; u < kTileP; u++) tot += segcnt[u];
    tile_ucnt[t] = tot;
    if (tot > 65535) atomicAdd(overflow, 1);
  }
}

// adjacent-cell sets up to compact_max (at most this, the LDS flag array)
// are compacted to the particles that some list references (larger ones are
// staged whole)
constexpr int kCompactMax = 4096;
// candidate positions held in LDS at a time by the list builder
constexpr int kListStage = 512;

// Neighbour lists in staged order (16-bit slots), and the staged set
// tile_j / tile_uc.  Same float criterion as the reference
// search: 0 < d2 < r2, d2 = dx*dx + dy*dy + dz*dz without contraction, at
// most cap entries (70 as in the reference, 96 with a list skin).  Every
// thread tests every particle of the tile's adjacent-cell set, in set order,
// from positions staged in LDS: all lanes read the same candidate (an LDS
// broadcast) and loop the same count.  A candidate outside the cells around
// the particle's own cell is more than the 1.01 r cell width away, so it
// never passes the test, and the lists are those of a scan over the 27 cells
// around each particle.  The compaction keeps the order of the set, so
// every list stays in ascending staged order.
constexpr int kListThreads = 2 * kTileP;  // two threads per particle

// Exclusive prefix of v over the kListThreads threads of the block (wave
// prefixes by shuffles, then the wave totals through LDS); `total` receives
// the block sum.  Every thread of the block must call it.
__device__ __forceinline__ int block_excl_scan(int v, int* wsum, int& total) {
  constexpr int NW = kListThreads / 64;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int incl = v;
  for (int d = 1; d < 64; d <<= 1) {
    const int o = __shfl_up(incl, d);
    if (lane >= d) incl += o;
  }
  if (lane == 63) wsum[w] = incl;
  __syncthreads();
  int off = 0;
  total = 0;
  for (int u = 0; u < NW; u++) {
    const int t = wsum[u];
    off += u < w ? t : 0;
    total += t;
  }
  __syncthreads();  // wsum may be written again
  return off + incl - v;
}

// The cell holding staged position q: the last p with spre[p] <= q (an empty
// cell shares its prefix with the next one, so the cell found is never empty).
__device__ __forceinline__ int staged_cell(const int* spre, int nc, int q) {
  int lo = 0, hi = nc - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (spre[mid] <= q) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// p0..p2: the positions of the cell-sorted order (candidates); o0..o2: the
// positions of the tiles' own particles by storage index, and map[s] the
// storage index at sorted position s (both null right after a reorder,
// where the two orders are one).  tiles: the tiles to build (null: all).
// With map, each tile's staged set must fit the region it was given at the
// last reorder; a tile whose set does not raises *overflow.
__global__ __launch_bounds__(kListThreads) void k_tile_lists(
    const float* __restrict__ p0, const float* __restrict__ p1, const float* __restrict__ p2,
    const float* __restrict__ o0, const float* __restrict__ o1, const float* __restrict__ o2,
    const int* __restrict__ map, const int* __restrict__ tiles,
    int n, GridDesc g, const int* __restrict__ cell_start, const int* __restrict__ tile_keys,
    const int* __restrict__ tile_nc, const int* __restrict__ tile_uoff, float r2, int cap,
    int rcap, int64_t npad, int* __restrict__ count, unsigned short* __restrict__ slot,
    int* __restrict__ tile_j, int* __restrict__ tile_uc, int* __restrict__ tile_over,
    int* overflow, int compact_max) {
#pragma clang fp contract(off)
  constexpr int NT = kListThreads;
  __shared__ int sb[kTileMaxCells], sn[kTileMaxCells], spre[kTileMaxCells];
  __shared__ float4 spos[kListStage];  // candidate position, particle index in w
  __shared__ short sidx[kCompactMax];  // referenced flag, then compacted position (or -1)
  __shared__ int wsum[NT / 64];
  const int t = tiles ? tiles[blockIdx.x] : blockIdx.x;
  const int tid = threadIdx.x;
  const int sub = tid & 1;  // lanes 2p and 2p+1 hold particle p
  const int nc = min(tile_nc[t], kTileMaxCells);
  const int* keys = tile_keys + (int64_t)t * kTileMaxCells;
  // cell ranges and their prefix: thread tid takes cells [tid*K, tid*K+K)
  const int K = (nc + NT - 1) / NT;  // at most kTileMaxCells / NT
  const int p0c = min(tid * K, nc), p1c = min(p0c + K, nc);
  int mine = 0;
  for (int p = p0c; p < p1c; p++) {
    const int k = keys[p];
    const int b = cell_start[k];
    sb[p] = b;
    sn[p] = cell_start[k + 1] - b;
    mine += sn[p];
  }
  int U;
  int run0 = block_excl_scan(mine, wsum, U);
  for (int p = p0c; p < p1c; p++) {
    spre[p] = run0;
    run0 += sn[p];
  }
  const bool compact = U <= compact_max;
  if (compact)
    for (int s = tid; s < U; s += NT) sidx[s] = 0;
  __syncthreads();
  const int uoff = tile_uoff[t];
  const int i = t * kTileP + (tid >> 1);
  const bool act = i < n;
  int c = 0;  // the particle's list length so far (the same in both lanes)
  float xi = 0.f, yi = 0.f, zi = 0.f;
  if (act) {
    xi = o0 ? o0[i] : p0[i];
    yi = o1 ? o1[i] : p1[i];
    zi = o2 ? o2[i] : p2[i];
  }
  for (int c0 = 0; c0 < U; c0 += kListStage) {
    const int m = min(kListStage, U - c0);
    __syncthreads();  // the previous chunk is consumed (and spre is complete)
    for (int q = tid; q < m; q += NT) {
      const int p = staged_cell(spre, nc, c0 + q);
      const int j = sb[p] + (c0 + q - spre[p]);
      spos[q] = make_float4(p0[j], p1[j], p2[j], __int_as_float(j));
    }
    __syncthreads();
    if (act) {
      // sixteen candidates at a time: each lane tests eight into a bit mask,
      // the two masks are swapped, and each lane appends its own hits at the
      // positions the combined mask gives them, in candidate order
      for (int q0 = 0; q0 < m; q0 += 16) {
        const int qb = q0 + 8 * sub;
        unsigned mine = 0u;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const float4 pj = spos[min(qb + u, m - 1)];
          float ddx = pj.x - xi, ddy = pj.y - yi, ddz = pj.z - zi;
          float d2 = __fadd_rn(__fadd_rn(__fmul_rn(ddx, ddx), __fmul_rn(ddy, ddy)),
                               __fmul_rn(ddz, ddz));
          if (qb + u < m && d2 < r2 && d2 > 0.0f) mine |= 1u << u;
        }
        const unsigned other = (unsigned)__shfl_xor((int)mine, 1);
        const int nlo = __popc(sub ? other : mine);  // hits among the first eight
        int k = c + (sub ? nlo : 0);
        unsigned h = mine;
        while (h) {
          const int q = qb + __builtin_ctz(h);
          h &= h - 1u;
          if (k < cap) {
            const int sl = c0 + q;
            slot[(((int64_t)(k >> 2)) * npad + i) * 4 + (k & 3)] = (unsigned short)sl;
            if (compact) sidx[sl] = 1;
          }
          k++;
        }
        c += __popc(mine) + __popc(other);
      }
    }
  }
  if (act && sub == 0) count[i] = c < cap ? c : cap;
  __syncthreads();  // every slot write (either lane's) is visible
  int used = U;
  if (compact) {
    // exclusive scan of the flags, one contiguous segment per thread
    const int seg = (U + NT - 1) / NT;
    const int s0 = min(tid * seg, U), s1 = min(s0 + seg, U);
    int sum = 0;
    for (int s = s0; s < s1; s++) sum += sidx[s];
    int run = block_excl_scan(sum, wsum, used);
    for (int s = s0; s < s1; s++) {
      const int fl = sidx[s];
      sidx[s] = fl ? run : -1;
      run += fl;
    }
    __syncthreads();
    // the particle's slots to compacted positions; lane `sub` takes the
    // quads q with q % 2 == sub, four of them in flight at a time
    const int m = act ? min(c, cap) : 0;
    const int nq = (m + 3) >> 2;
    uint2* qrow = reinterpret_cast<uint2*>(slot) + i;
    for (int q0 = sub; q0 < nq; q0 += 8) {
      uint2 w[4];
#pragma unroll
      for (int u = 0; u < 4; u++)
        if (q0 + 2 * u < nq) w[u] = qrow[(int64_t)(q0 + 2 * u) * npad];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int qq = q0 + 2 * u;
        if (qq >= nq) continue;
        const int k = 4 * qq;  // slots past m keep what they hold
        unsigned lo = w[u].x, hi = w[u].y;
        if (k < m) lo = (lo & 0xFFFF0000u) | (unsigned short)sidx[lo & 0xFFFFu];
        if (k + 1 < m) lo = (lo & 0xFFFFu) | ((unsigned)(unsigned short)sidx[lo >> 16] << 16);
        if (k + 2 < m) hi = (hi & 0xFFFF0000u) | (unsigned short)sidx[hi & 0xFFFFu];
        if (k + 3 < m) hi = (hi & 0xFFFFu) | ((unsigned)(unsigned short)sidx[hi >> 16] << 16);
        qrow[(int64_t)qq * npad] = make_uint2(lo, hi);
      }
    }
  }
  // the region this tile was given at the last reorder
  const int room = map ? tile_uoff[t + 1] - uoff : INT_MAX;
  if (used > room) {
    if (tid == 0) atomicAdd(overflow, 1);
  } else {
    for (int s = tid; s < U; s += NT) {
      const int d = compact ? sidx[s] : s;
      if (d >= 0) {
        const int p = staged_cell(spre, nc, s);
        const int js = sb[p] + (s - spre[p]);
        tile_j[uoff + d] = map ? map[js] : js;
      }
    }
  }
  if (tid == 0) tile_uc[t] = used;
  // particles of the tile over the reported cap, one count each
  const int over = act && sub == 0 && c > rcap;
  const unsigned long long o = wave_sum_u64((unsigned long long)over);
  if (lane_id() == 0) wsum[threadIdx.x >> 6] = (int)o;
  __syncthreads();
  if (tid == 0) {
    int tot = 0;
    for (int u = 0; u < NT / 64; u++) tot += wsum[u];
    tile_over[t] = tot;
  }
}

void launch_tile_cells(hipStream_t s, const float* const pf[3], int n, const GridDesc& g,
                       const int* cell_start, const int* tiles, int ntiles, bool sort_own,
                       int* tile_keys, int* tile_nc, int* tile_ucnt, int* overflow) {
  if (ntiles <= 0) return;
  if (sort_own)
    hipLaunchKernelGGL(k_tile_cells<true>, dim3(ntiles), dim3(kTileP), 0, s, pf[0], pf[1], pf[2],
                       n, g, cell_start, tiles, tile_keys, tile_nc, tile_ucnt, overflow);
  else
    hipLaunchKernelGGL(k_tile_cells<false>, dim3(ntiles), dim3(kTileP), 0, s, pf[0], pf[1],
                       pf[2], n, g, cell_start, tiles, tile_keys, tile_nc, tile_ucnt, overflow);
}

void launch_tile_lists(hipStream_t s, const float* const pf[3], const float* const own[3],
                       const int* map, const int* tiles, int ntiles, int n, const GridDesc& g,
                       const int* cell_start, const int* tile_keys, const int* tile_nc,
                       const int* tile_uoff, float r2, int cap, int rcap, int64_t npad,
                       int* count, unsigned short* slot, int* tile_j, int* tile_uc,
                       int* tile_over, int* overflow, int compact_max) {
  if (ntiles <= 0) return;
  compact_max = compact_max > 0 ? std::min(compact_max, kCompactMax) : kCompactMax;
  const float* o0 = own ? own[0] : nullptr;
  const float* o1 = own ? own[1] : nullptr;
  const float* o2 = own ? own[2] : nullptr;
  hipLaunchKernelGGL(k_tile_lists, dim3(ntiles), dim3(kListThreads), 0, s, pf[0], pf[1], pf[2],
                     o0, o1, o2, map, tiles, n, g, cell_start, tile_keys, tile_nc, tile_uoff, r2,
                     cap, rcap, npad, count, slot, tile_j, tile_uc, tile_over, overflow,
                     compact_max);
}

}  // namespace wf
