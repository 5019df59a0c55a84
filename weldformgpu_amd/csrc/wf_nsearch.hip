// wf_nsearch.hip — fixed-radius neighbour search with Morton-sorted
// particle storage (replaces cuNSearch, src/cuNSearch/src/cuNSearch.cu:111-137).
//
//  1. positions -> float (the reference searches in FP32: Common.h:11-15),
//     bounding box by a two-level reduction (one 24-byte D2H per rebuild);
//  2. uniform grid of cell size slightly above the radius, padded by a cell;
//     cell key = 8^3 meta-block index * 512 + Morton code inside the block,
//     so consecutive keys are spatially compact (the same layout family as
//     cuNSearch's ToCellIndex_MortonMetaGrid, cuNSearchKernels.cu:10-21);
//  3. counting sort by key (atomic histogram + exclusive scan), then each
//     cell's particles ordered by original index so the permutation is
//     deterministic (the reference's atomic ranks are not,
//     cuNSearchKernels.cu:104);
//  4. the caller gathers every field into the new order;
//  5. per tile of the new order, the occupied neighbour cells, the staged
//     set and the neighbour lists (wf_tiles.hip): j is a neighbour of i iff
//     0 < d2 < r2 in float, d2 = dx*dx + dy*dy + dz*dz without contraction
//     (cuNSearchKernels.cu:160-165), at most 70 kept (:169).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "wf_device.hpp"
#include "wf_tiles.hpp"

namespace wf {

struct NsWork {
  int compact_max = 0;  // launch_tile_lists' compaction limit (0: its default)
  int n_cap = 0;
  int64_t key_cap = 0;
  float* pf[3] = {nullptr, nullptr, nullptr};   // float positions, old order
  float* pfn[3] = {nullptr, nullptr, nullptr};  // float positions, new order
  int* key = nullptr;
  int* rank = nullptr;
  int* slot_orig = nullptr;  // key-sorted: original index (for ordering)
  int* cell_count = nullptr;
  int* cell_start = nullptr;  // key_cap + 1
  int* scan_part = nullptr;
  float* mm_part = nullptr;  // per block 6 floats
  float* mm = nullptr;       // 6 floats
  float* mm_host = nullptr;  // pinned
  // tiles
  int nt_cap = 0;
  int* tile_keys = nullptr;
  int* tile_nc = nullptr;
  int* tile_ucnt = nullptr;
  int* tile_uoff = nullptr;
  int* tile_uc = nullptr;  // compacted staged-set sizes
  int* tile_scan = nullptr;
  int* tile_ovf = nullptr;
  int* tile_host = nullptr;  // pinned: [total staged, overflow]
  int64_t tj_cap = 0;
  int* tile_j = nullptr;
  unsigned short* slot = nullptr;
  int* tile_over = nullptr;  // particles over the list cap, per tile
  // re-listing without a reorder: the cell sort of the current positions
  // (sorted position -> storage index), the cells near the plastic zone,
  // the tiles they touch and the list of those tiles
  int* lgather = nullptr;
  int* lperm = nullptr;
  int* cell_flag = nullptr;
  int* tile_aff = nullptr;
  int* tile_pos = nullptr;
  int* tile_list = nullptr;
};

constexpr int kScanBlock = 1024;
constexpr int kMmBlocks = 512;

__global__ void k_ns_to_float(const double* __restrict__ x0, const double* __restrict__ x1,
                              const double* __restrict__ x2, float* p0, float* p1, float* p2,
                              int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    p0[i] = (float)x0[i];
    p1[i] = (float)x1[i];
    p2[i] = (float)x2[i];
  }
}

__global__ __launch_bounds__(256) void k_ns_minmax_part(const float* p0, const float* p1,
                                                        const float* p2, int n, float* part) {
  __shared__ float sh[6][256];
  float mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    float v[3] = {p0[i], p1[i], p2[i]};
    for (int c = 0; c < 3; c++) {
      mn[c] = fminf(mn[c], v[c]);
      mx[c] = fmaxf(mx[c], v[c]);
    }
  }
  for (int c = 0; c < 3; c++) {
    sh[c][threadIdx.x] = mn[c];
    sh[3 + c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 3; c++) {
        sh[c][threadIdx.x] = fminf(sh[c][threadIdx.x], sh[c][threadIdx.x + s]);
        sh[3 + c][threadIdx.x] = fmaxf(sh[3 + c][threadIdx.x], sh[3 + c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_ns_minmax_final(const float* part, int nb, float* mm) {
  __shared__ float sh[6][256];
  float v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
  for (int b = threadIdx.x; b < nb; b += 256)
    for (int c = 0; c < 3; c++) {
      v[c] = fminf(v[c], part[b * 6 + c]);
      v[3 + c] = fmaxf(v[3 + c], part[b * 6 + 3 + c]);
    }
  for (int c = 0; c < 6; c++) sh[c][threadIdx.x] = v[c];
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st)
      for (int c = 0; c < 3; c++) {
        sh[c][threadIdx.x] = fminf(sh[c][threadIdx.x], sh[c][threadIdx.x + st]);
        sh[3 + c][threadIdx.x] = fmaxf(sh[3 + c][threadIdx.x], sh[3 + c][threadIdx.x + st]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) mm[threadIdx.x] = sh[threadIdx.x][0];
}

using Grid = GridDesc;

__global__ void k_ns_fill_i32(int* a, int v, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}

// Cell key of each particle and a rank unique within its cell.  The
// particles are still in the last rebuild's cell order, so neighbouring lanes
// mostly share a cell: each run of equal keys within a wave (active lanes are
// a prefix of the wave) takes its ranks with one atomic by its first lane.
// Which rank a particle gets does not matter: k_ns_order orders each cell by
// original index.
__global__ void k_ns_keys(const float* p0, const float* p1, const float* p2, int n, Grid g,
                          int* key, int* rank, int* cell_count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = i < n;
  int k = -1;
  if (act) {
    int cx, cy, cz;
    cell_of(g, p0[i], p1[i], p2[i], cx, cy, cz);
    k = cell_key(g, cx, cy, cz);
    key[i] = k;
  }
  const int lane = lane_id();
  const int kp = __shfl_up(k, 1);
  const bool head = act && (lane == 0 || kp != k);
  const unsigned long long H = __ballot(head);
  const unsigned long long A = __ballot(act);
  const unsigned long long upto = lane == 63 ? ~0ull : (2ull << lane) - 1ull;
  const unsigned long long hb = H & upto;
  const int h = hb ? 63 - __clzll((long long)hb) : 0;  // my run's first lane
  int base = 0;
  if (head) {
    const unsigned long long above = lane == 63 ? 0ull : H >> (lane + 1);
    const int next = above ? lane + 1 + __builtin_ctzll(above) : __popcll(A);
    base = atomicAdd(&cell_count[k], next - lane);
  }
  base = __shfl(base, h);
  if (act) rank[i] = base + (lane - h);
}

// exclusive scan over int (length m), three phases
__global__ __launch_bounds__(kScanBlock) void k_scan_reduce(const int* in, int64_t m, int* part) {
  __shared__ int sh[kScanBlock];
  int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  sh[threadIdx.x] = i < m ? in[i] : 0;
  __syncthreads();
  for (int s = kScanBlock / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ void k_scan_parts(int* part, int nb) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    int run = 0;
    for (int b = 0; b < nb; b++) {
      int t = part[b];
      part[b] = run;
      run += t;
    }
  }
}
__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const int* in, int64_t m,
                                                           const int* part, int* out) {
  __shared__ int sh[kScanBlock];
  int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
  int v = i < m ? in[i] : 0;
  sh[threadIdx.x] = v;
  __syncthreads();
  // Hillis-Steele inclusive scan
  for (int off = 1; off < kScanBlock; off <<= 1) {
    int t = threadIdx.x >= off ? sh[threadIdx.x - off] : 0;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  if (i < m) out[i] = part[blockIdx.x] + sh[threadIdx.x] - v;
  if (i == m - 1) out[m] = part[blockIdx.x] + sh[threadIdx.x];
}

__global__ void k_ns_place(const int* key, const int* rank, const int* cell_start,
                           const int* perm, int n, int* slot_orig) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  slot_orig[cell_start[key[i]] + rank[i]] = perm[i];
}

// Final slot of each particle: its cell's start plus the number of particles
// of the cell with a smaller original index, so a cell lists its particles by
// original index and the order is deterministic.  Writes the new permutation
// (slot -> original index) and the gather map (slot -> current slot).
__global__ void k_ns_order(const int* key, const int* cell_start, const int* perm,
                           const int* slot_orig, int n, int* perm_new, int* gather) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int k = key[i];
  const int b = cell_start[k], e = cell_start[k + 1];
  const int o = perm[i];
  int r = 0;
  for (int t = b; t < e; t++) r += slot_orig[t] < o;
  perm_new[b + r] = o;
  gather[b + r] = i;
}

__global__ void k_ns_gather_f32(float* dst, const float* src, const int* idx, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

// *out = sum of v[0..n) (one block, a fixed order: deterministic)
__global__ __launch_bounds__(256) void k_sum_tiles(const int* __restrict__ v, int n, int* out) {
  __shared__ int sh[256];
  int a = 0;
  for (int i = threadIdx.x; i < n; i += 256) a += v[i];
  sh[threadIdx.x] = a;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// Flag the cells within two cells (on each axis) of a particle with
// eps_p > ep_hot: with cells of 1.01 r, r = 2h + skin, every particle that
// such a particle can reach before it has moved skin/2 lies in them.
__global__ void k_hot_cells(const float* p0, const float* p1, const float* p2, int n, Grid g,
                            const double* __restrict__ pl_strain, double ep_hot, int* flag) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || !(pl_strain[i] > ep_hot)) return;
  int cx, cy, cz;
  cell_of(g, p0[i], p1[i], p2[i], cx, cy, cz);
  for (int z = max(cz - 2, 0); z <= min(cz + 2, g.dim[2] - 1); z++)
    for (int y = max(cy - 2, 0); y <= min(cy + 2, g.dim[1] - 1); y++)
      for (int x = max(cx - 2, 0); x <= min(cx + 2, g.dim[0] - 1); x++)
        flag[cell_key(g, x, y, z)] = 1;
}

// aff[t] = 1 when a particle of tile t lies in a flagged cell
__global__ __launch_bounds__(kTileP) void k_tiles_touched(const float* p0, const float* p1,
                                                           const float* p2, int n, Grid g,
                                                           const int* __restrict__ flag,
                                                           int* aff) {
  const int t = blockIdx.x, i = t * kTileP + threadIdx.x;
  int hit = 0;
  if (i < n) {
    int cx, cy, cz;
    cell_of(g, p0[i], p1[i], p2[i], cx, cy, cz);
    hit = flag[cell_key(g, cx, cy, cz)];
  }
  hit = __syncthreads_or(hit);
  if (threadIdx.x == 0) aff[t] = hit ? 1 : 0;
}

// list of the flagged tiles, ascending
__global__ void k_tile_compact(const int* aff, const int* pos, int nt, int* list) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nt && aff[t]) list[pos[t]] = t;
}

// the build positions of the re-listed tiles' particles
__global__ __launch_bounds__(kTileP) void k_tiles_xb(const double* x0, const double* x1,
                                                     const double* x2, int n,
                                                     const int* __restrict__ list, double* b0,
                                                     double* b1, double* b2) {
  const int i = list[blockIdx.x] * kTileP + threadIdx.x;
  if (i >= n) return;
  b0[i] = x0[i];
  b1[i] = x1[i];
  b2[i] = x2[i];
}

static int grid1(int64_t n, int b) { return (int)((n + b - 1) / b); }

void launch_scan_exclusive(hipStream_t s, const int* in, int64_t m, int* part, int* out) {
  const int nb = grid1(m, kScanBlock);
  if (nb <= 0) return;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nb), dim3(kScanBlock), 0, s, in, m, part);
  hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(64), 0, s, part, nb);
  hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, s, in, m, part, out);
}

static int ensure(NsWork* w, int n, int64_t nkeys) {
  if (n > w->n_cap) {
    for (int c = 0; c < 3; c++) {
      hipFree(w->pf[c]);
      hipFree(w->pfn[c]);
    }
    hipFree(w->key);
    hipFree(w->rank);
    hipFree(w->slot_orig);
    int cap = std::max(n, 64);
    for (int c = 0; c < 3; c++) {
      if (hipMalloc(&w->pf[c], sizeof(float) * cap) != hipSuccess) return WF_E_NOMEM;
      if (hipMalloc(&w->pfn[c], sizeof(float) * cap) != hipSuccess) return WF_E_NOMEM;
    }
    if (hipMalloc(&w->key, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->rank, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->slot_orig, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    hipFree(w->lgather);
    hipFree(w->lperm);
    if (hipMalloc(&w->lgather, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->lperm, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    w->n_cap = cap;
  }
  if (nkeys > w->key_cap) {
    hipFree(w->cell_count);
    hipFree(w->cell_start);
    hipFree(w->scan_part);
    hipFree(w->cell_flag);
    int64_t cap = nkeys + nkeys / 4 + 1024;
    if (hipMalloc(&w->cell_flag, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->cell_count, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->cell_start, sizeof(int) * (cap + 1)) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->scan_part, sizeof(int) * (grid1(cap, kScanBlock) + 1)) != hipSuccess)
      return WF_E_NOMEM;
    w->key_cap = cap;
  }
  int nt = (n + kTileP - 1) / kTileP;
  if (nt > w->nt_cap) {
    hipFree(w->tile_keys);
    hipFree(w->tile_nc);
    hipFree(w->tile_ucnt);
    hipFree(w->tile_uoff);
    hipFree(w->tile_uc);
    hipFree(w->tile_scan);
    hipFree(w->slot);
    hipFree(w->tile_over);
    hipFree(w->tile_aff);
    hipFree(w->tile_pos);
    hipFree(w->tile_list);
    int cap = std::max(nt, 1);
    int64_t npad = ((int64_t)cap * kTileP);
    if (hipMalloc(&w->tile_keys, sizeof(int) * (int64_t)cap * kTileMaxCells) != hipSuccess)
      return WF_E_NOMEM;
    if (hipMalloc(&w->tile_nc, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->tile_ucnt, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->tile_uoff, sizeof(int) * (cap + 1)) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->tile_uc, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->tile_scan, sizeof(int) * (grid1(cap, kScanBlock) + 1)) != hipSuccess)
      return WF_E_NOMEM;
    if (hipMalloc(&w->slot, sizeof(unsigned short) * 4 * kSlotQuads * npad) != hipSuccess)
      return WF_E_NOMEM;
    if (hipMalloc(&w->tile_over, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->tile_aff, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->tile_pos, sizeof(int) * (cap + 1)) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->tile_list, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    w->nt_cap = cap;
  }
  if (!w->tile_ovf) {
    if (hipMalloc(&w->tile_ovf, sizeof(int)) != hipSuccess) return WF_E_NOMEM;
    if (hipHostMalloc(&w->tile_host, sizeof(int) * 2) != hipSuccess) return WF_E_NOMEM;
  }
  if (!w->mm_part) {
    if (hipMalloc(&w->mm_part, sizeof(float) * 6 * kMmBlocks) != hipSuccess) return WF_E_NOMEM;
    if (hipMalloc(&w->mm, sizeof(float) * 6) != hipSuccess) return WF_E_NOMEM;
    if (hipHostMalloc(&w->mm_host, sizeof(float) * 6) != hipSuccess) return WF_E_NOMEM;
  }
  return WF_OK;
}

void ns_set_compact_max(NsWork** work, int compact_max) {
  if (!*work) *work = new NsWork();
  (*work)->compact_max = compact_max;
}

void ns_free(NsWork* w) {
  if (!w) return;
  for (int c = 0; c < 3; c++) {
    hipFree(w->pf[c]);
    hipFree(w->pfn[c]);
  }
  hipFree(w->key);
  hipFree(w->rank);
  hipFree(w->slot_orig);
  hipFree(w->cell_count);
  hipFree(w->cell_start);
  hipFree(w->scan_part);
  hipFree(w->mm_part);
  hipFree(w->mm);
  if (w->mm_host) hipHostFree(w->mm_host);
  hipFree(w->tile_keys);
  hipFree(w->tile_nc);
  hipFree(w->tile_ucnt);
  hipFree(w->tile_uoff);
  hipFree(w->tile_uc);
  hipFree(w->tile_scan);
  hipFree(w->tile_ovf);
  if (w->tile_host) hipHostFree(w->tile_host);
  hipFree(w->tile_j);
  hipFree(w->tile_over);
  hipFree(w->lgather);
  hipFree(w->lperm);
  hipFree(w->cell_flag);
  hipFree(w->tile_aff);
  hipFree(w->tile_pos);
  hipFree(w->tile_list);
  hipFree(w->slot);
  delete w;
}

// gather: out slot -> old sorted index; perm (in/out): sorted slot -> original
int ns_build(hipStream_t s, NsWork** work, const double* const pos_src[3], int n, int64_t npad,
             double radius, int cap, int* gather, int* count, Counters* ctr,
             int* perm, int* perm_tmp, TileLists* tiles, GridDesc* grid_out,
             const int* max_dim) {
  if (!*work) *work = new NsWork();
  NsWork* w = *work;
  int rc = ensure(w, n, 0);
  if (rc) return rc;
  const float rf = (float)radius;
  const float r2 = rf * rf;
  hipLaunchKernelGGL(k_ns_to_float, dim3(grid1(n, 256)), dim3(256), 0, s, pos_src[0], pos_src[1],
                     pos_src[2], w->pf[0], w->pf[1], w->pf[2], n);
  int nbm = std::min(kMmBlocks, std::max(1, grid1(n, 256)));
  hipLaunchKernelGGL(k_ns_minmax_part, dim3(nbm), dim3(256), 0, s, w->pf[0], w->pf[1], w->pf[2],
                     n, w->mm_part);
  hipLaunchKernelGGL(k_ns_minmax_final, dim3(1), dim3(256), 0, s, w->mm_part, nbm, w->mm);
  if (hipMemcpyAsync(w->mm_host, w->mm, 6 * sizeof(float), hipMemcpyDeviceToHost, s) != hipSuccess)
    return WF_E_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return WF_E_HIP;
  Grid g;
  double cs = (double)rf * 1.01;
  g.inv_cs = 1.0 / cs;
  int64_t nkeys = 1;
  for (int c = 0; c < 3; c++) {
    double lo = w->mm_host[c], hi = w->mm_host[3 + c];
    if (!std::isfinite(lo) || !std::isfinite(hi)) return WF_E_NONFINITE;
    g.org[c] = lo - cs;
    double ext = (hi - g.org[c]) / cs;
    if (ext > 1.0e6) return WF_E_ARG;  // domain too sparse for a uniform grid
    g.dim[c] = (int)std::floor(ext) + 2;
    if (max_dim && g.dim[c] > max_dim[c]) return WF_E_UNSUPPORTED;
    g.mdim[c] = (g.dim[c] + 7) / 8;
    nkeys *= g.mdim[c];
  }
  nkeys *= 512;
  if (nkeys > (int64_t)1 << 30) return WF_E_ARG;
  rc = ensure(w, n, nkeys);
  if (rc) return rc;
  hipLaunchKernelGGL(k_ns_fill_i32, dim3(grid1(nkeys + 1, 256)), dim3(256), 0, s, w->cell_count, 0,
                     nkeys + 1);
  hipLaunchKernelGGL(k_ns_keys, dim3(grid1(n, 256)), dim3(256), 0, s, w->pf[0], w->pf[1],
                     w->pf[2], n, g, w->key, w->rank, w->cell_count);
  int nsb = grid1(nkeys, kScanBlock);
  hipLaunchKernelGGL(k_scan_reduce, dim3(nsb), dim3(kScanBlock), 0, s, w->cell_count, nkeys,
                     w->scan_part);
  hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(64), 0, s, w->scan_part, nsb);
  hipLaunchKernelGGL(k_scan_apply, dim3(nsb), dim3(kScanBlock), 0, s, w->cell_count, nkeys,
                     w->scan_part, w->cell_start);
  hipLaunchKernelGGL(k_ns_place, dim3(grid1(n, 256)), dim3(256), 0, s, w->key, w->rank,
                     w->cell_start, perm, n, w->slot_orig);
  // new permutation and gather map
  hipLaunchKernelGGL(k_ns_order, dim3(grid1(n, 256)), dim3(256), 0, s, w->key, w->cell_start,
                     perm, w->slot_orig, n, perm_tmp, gather);
  for (int c = 0; c < 3; c++)
    hipLaunchKernelGGL(k_ns_gather_f32, dim3(grid1(n, 256)), dim3(256), 0, s, w->pfn[c], w->pf[c],
                       gather, n);
  hipLaunchKernelGGL(k_ns_fill_i32, dim3(1), dim3(64), 0, s, &ctr->nb_over_cap, 0, (int64_t)1);
  hipLaunchKernelGGL(k_ns_fill_i32, dim3(1), dim3(64), 0, s, w->tile_ovf, 0, (int64_t)1);
  // tiles of the new order: occupied neighbour cells and staged counts
  const int nt = (n + kTileP - 1) / kTileP;
  const float* pfn[3] = {w->pfn[0], w->pfn[1], w->pfn[2]};
  launch_tile_cells(s, pfn, n, g, w->cell_start, nullptr, nt, false, w->tile_keys, w->tile_nc,
                    w->tile_ucnt, w->tile_ovf);
  int ntb = grid1(nt, kScanBlock);
  hipLaunchKernelGGL(k_scan_reduce, dim3(ntb), dim3(kScanBlock), 0, s, w->tile_ucnt,
                     (int64_t)nt, w->tile_scan);
  hipLaunchKernelGGL(k_scan_parts, dim3(1), dim3(64), 0, s, w->tile_scan, ntb);
  hipLaunchKernelGGL(k_scan_apply, dim3(ntb), dim3(kScanBlock), 0, s, w->tile_ucnt, (int64_t)nt,
                     w->tile_scan, w->tile_uoff);
  if (hipMemcpyAsync(w->tile_host, w->tile_uoff + nt, sizeof(int), hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return WF_E_HIP;
  if (hipMemcpyAsync(w->tile_host + 1, w->tile_ovf, sizeof(int), hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return WF_E_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return WF_E_HIP;
  if (w->tile_host[1] != 0) return WF_E_NBCAP;  // tile outgrew its cell or slot budget
  const int64_t staged = w->tile_host[0];
  if (staged > w->tj_cap) {
    hipFree(w->tile_j);
    int64_t cap = staged + staged / 4 + 1024;
    if (hipMalloc(&w->tile_j, sizeof(int) * cap) != hipSuccess) return WF_E_NOMEM;
    w->tj_cap = cap;
  }
  launch_tile_lists(s, pfn, nullptr, nullptr, nullptr, nt, n, g, w->cell_start, w->tile_keys,
                    w->tile_nc, w->tile_uoff, r2, kListCap, cap, npad, count, w->slot, w->tile_j,
                    w->tile_uc, w->tile_over, w->tile_ovf, w->compact_max);
  hipLaunchKernelGGL(k_sum_tiles, dim3(1), dim3(256), 0, s, w->tile_over, nt, &ctr->nb_over_cap);
  if (hipGetLastError() != hipSuccess) return WF_E_HIP;
  *grid_out = g;
  tiles->count = count;
  tiles->slotq = (const uint2*)w->slot;
  tiles->uoff = w->tile_uoff;
  tiles->ucnt = w->tile_uc;
  tiles->tj = w->tile_j;
  tiles->npad = npad;
  return WF_OK;
}

int ns_relist_local(hipStream_t s, NsWork* w, const double* const x[3], const double* pl_strain,
                    double ep_hot, int n, int64_t npad, const GridDesc& g, double radius, int cap,
                    const int* perm, int* count, Counters* ctr, double* const xb[3],
                    int* relisted) {
  *relisted = 0;
  if (!w || w->n_cap < n) return WF_E_STATE;
  const int64_t nkeys = (int64_t)g.mdim[0] * g.mdim[1] * g.mdim[2] * 512;
  if (nkeys > w->key_cap) return WF_E_STATE;
  const float rf = (float)radius;
  const float r2 = rf * rf;
  const int nt = (n + kTileP - 1) / kTileP;
  // the cell sort of the current positions, on the last reorder's grid;
  // the particles stay where they are (the sorted order only indexes them)
  hipLaunchKernelGGL(k_ns_to_float, dim3(grid1(n, 256)), dim3(256), 0, s, x[0], x[1], x[2],
                     w->pf[0], w->pf[1], w->pf[2], n);
  hipLaunchKernelGGL(k_ns_fill_i32, dim3(grid1(nkeys + 1, 256)), dim3(256), 0, s, w->cell_count, 0,
                     nkeys + 1);
  hipLaunchKernelGGL(k_ns_fill_i32, dim3(grid1(nkeys + 1, 256)), dim3(256), 0, s, w->cell_flag, 0,
                     nkeys + 1);
  hipLaunchKernelGGL(k_ns_keys, dim3(grid1(n, 256)), dim3(256), 0, s, w->pf[0], w->pf[1],
                     w->pf[2], n, g, w->key, w->rank, w->cell_count);
  launch_scan_exclusive(s, w->cell_count, nkeys, w->scan_part, w->cell_start);
  hipLaunchKernelGGL(k_ns_place, dim3(grid1(n, 256)), dim3(256), 0, s, w->key, w->rank,
                     w->cell_start, perm, n, w->slot_orig);
  hipLaunchKernelGGL(k_ns_order, dim3(grid1(n, 256)), dim3(256), 0, s, w->key, w->cell_start,
                     perm, w->slot_orig, n, w->lperm, w->lgather);
  for (int c = 0; c < 3; c++)
    hipLaunchKernelGGL(k_ns_gather_f32, dim3(grid1(n, 256)), dim3(256), 0, s, w->pfn[c], w->pf[c],
                       w->lgather, n);
  // the tiles within reach of the plastic zone
  hipLaunchKernelGGL(k_hot_cells, dim3(grid1(n, 256)), dim3(256), 0, s, w->pf[0], w->pf[1],
                     w->pf[2], n, g, pl_strain, ep_hot, w->cell_flag);
  hipLaunchKernelGGL(k_tiles_touched, dim3(nt), dim3(kTileP), 0, s, w->pf[0], w->pf[1], w->pf[2],
                     n, g, w->cell_flag, w->tile_aff);
  launch_scan_exclusive(s, w->tile_aff, nt, w->tile_scan, w->tile_pos);
  hipLaunchKernelGGL(k_tile_compact, dim3(grid1(nt, 256)), dim3(256), 0, s, w->tile_aff,
                     w->tile_pos, nt, w->tile_list);
  if (hipMemcpyAsync(w->tile_host, w->tile_pos + nt, sizeof(int), hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return WF_E_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return WF_E_HIP;
  const int ntl = w->tile_host[0];
  *relisted = ntl;
  if (ntl == 0) return WF_OK;
  hipLaunchKernelGGL(k_ns_fill_i32, dim3(1), dim3(64), 0, s, w->tile_ovf, 0, (int64_t)1);
  const float* pfn[3] = {w->pfn[0], w->pfn[1], w->pfn[2]};
  const float* own[3] = {w->pf[0], w->pf[1], w->pf[2]};
  launch_tile_cells(s, own, n, g, w->cell_start, w->tile_list, ntl, true, w->tile_keys,
                    w->tile_nc, w->tile_ucnt, w->tile_ovf);
  launch_tile_lists(s, pfn, own, w->lgather, w->tile_list, ntl, n, g, w->cell_start,
                    w->tile_keys, w->tile_nc, w->tile_uoff, r2, kListCap, cap, npad, count,
                    w->slot,
                    w->tile_j, w->tile_uc, w->tile_over, w->tile_ovf, w->compact_max);
  if (hipMemcpyAsync(w->tile_host + 1, w->tile_ovf, sizeof(int), hipMemcpyDeviceToHost, s) !=
      hipSuccess)
    return WF_E_HIP;
  if (hipStreamSynchronize(s) != hipSuccess) return WF_E_HIP;
  if (w->tile_host[1] != 0) return WF_E_NBCAP;  // a tile outgrew its region: reorder
  hipLaunchKernelGGL(k_sum_tiles, dim3(1), dim3(256), 0, s, w->tile_over, nt, &ctr->nb_over_cap);
  hipLaunchKernelGGL(k_tiles_xb, dim3(ntl), dim3(kTileP), 0, s, x[0], x[1], x[2], n,
                     w->tile_list, xb[0], xb[1], xb[2]);
  if (hipGetLastError() != hipSuccess) return WF_E_HIP;
  return WF_OK;
}

}  // namespace wf
