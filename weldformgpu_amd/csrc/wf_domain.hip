// wf_domain.hip — the C-ABI of libwfhip (include/wf_abi.h): the device domain
// handle, field transfers, neighbour rebuilds with physical reordering, the
// per-kernel entry points and the fused Fraser step.
//
// One handle owns one HIP stream (default priority) and plain hipMalloc'd
// planes; no managed memory (the reference keeps Domain_d in managed memory,
// WeldFormGPU.cu:170).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "wf_internal.hpp"

using namespace wf;

namespace {

thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

#define HIPC(expr)                                                                      \
  do {                                                                                  \
    hipError_t e_ = (expr);                                                             \
    if (e_ != hipSuccess)                                                               \
      return fail(WF_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));         \
  } while (0)

// ---- per-field plane access -------------------------------------------------
struct PlaneRef {
  double** d = nullptr;
  int** i = nullptr;
  float** f = nullptr;  // an FP32 plane behind a double field (mixed stress state)
};

}  // namespace

constexpr int kGatherBatch = 8;  // planes per reorder gather launch

struct wf_domain {
  wf_config cfg;
  int n = 0, solid = 0;
  int64_t npad = 0;
  hipStream_t stream = nullptr;
  Fields f{};
  std::vector<void*> owned;  // every allocation, freed in wf_destroy
  int* perm = nullptr;       // sorted slot -> original index
  int* perm_tmp = nullptr;
  int* gather = nullptr;
  int* nb_count = nullptr;
  double* spare_d[kGatherBatch] = {};  // reorder scratch planes
  float* spare_f[kGatherBatch] = {};   // the same for the FP32 planes
  int* spare_i = nullptr;
  void* stage = nullptr;      // transfer staging (6*n doubles)
  Counters* ctr = nullptr;
  Counters* ctr_host = nullptr;  // pinned
  double* sum_part = nullptr;
  double* sum_out = nullptr;
  NsWork* ns = nullptr;
  TileLists tl{};
  GridDesc grid{};
  Frame32 frame{};
  // return-map statistics and surrogate
  BlockStats* bstats = nullptr;
  int nblk = 0;  // blocks of the per-particle stress kernel
  int64_t mlp_cap = 0;
  bool jc_newton = false;  // Johnson-Cook under the Newton return map
  // WF_REBUILD_PLASTIC_ZONE: positions at the last build, the trigger's
  // largest squared displacement (device, pinned host copy), and whether
  // the next step starts with a rebuild
  double* xb[3] = {nullptr, nullptr, nullptr};
  unsigned long long *pz_d2 = nullptr, *pz_d2_host = nullptr;
  bool pz_pending = false;
  bool plastic_zone() const {
    return cfg.rebuild_policy == WF_REBUILD_PLASTIC_ZONE ||
           cfg.rebuild_policy == WF_REBUILD_PLASTIC_ZONE_LOCAL;
  }
  // WF_REBUILD_PLASTIC_ZONE_LOCAL: re-lists since the last reorder, and the
  // totals of re-lists and re-listed tiles
  int local_since_reorder = 0;
  int ns_local_builds = 0;
  int64_t ns_local_tiles = 0;
  // with a list skin the surface pass counts only pairs within 2 h_glob
  double surface_r2cut() const { return plastic_zone() ? 4.0 * cfg.h_glob * cfg.h_glob : 0.0; }
  // the free-surface pass in the domain's precision
  void launch_surface() {
    const double ntot = (double)n + (double)mesh.ne;
    if (mixed) {
      refresh_mirrors();
      launch_surface_mixed(stream, f, tl, solid, ntot, totmass, cfg.id_free_surf,
                           surface_r2cut(), uniform_m, m_value, frame.cs32);
    } else {
      launch_surface_tiled(stream, f, tl, solid, ntot, totmass, cfg.id_free_surf,
                           surface_r2cut(), uniform_m, m_value);
    }
  }
  // the stress pass's return-map arguments; the packed weights are brought
  // up to date (queued on the stream ahead of the pass) when they are used
  RetArgs ret_args() {
    const bool use = cfg.surrogate && mlp_trained;
    if (use) pack_mlp();
    return RetArgs{mlp_pack, use ? 1 : 0, (long long)step};
  }
  void pack_mlp() {
    if (!mlp_pack_stale) return;
    launch_mlp_pack(stream, mlp_w, mlp_pack);
    mlp_pack_stale = false;
  }
  float *mlp_w = nullptr, *mlp_m = nullptr, *mlp_v = nullptr, *mlp_grad = nullptr;
  void* mlp_pack = nullptr;     // per-lane MFMA operands of mlp_w
  bool mlp_pack_stale = true;   // mlp_w changed since the last pack
  float *mlp_part = nullptr, *mlp_loss_part = nullptr, *mlp_loss = nullptr;
  float *mlp_X = nullptr, *mlp_Y = nullptr;
  int *mlp_flags = nullptr, *mlp_scan = nullptr, *mlp_scan_part = nullptr, *mlp_count = nullptr;
  int* host_i = nullptr;    // pinned scratch
  float* host_f = nullptr;  // pinned scratch
  int adam_t = 0;
  bool mlp_trained = false;
  double mlp_last_loss = 0.0;
  double mlp_start_loss = 0.0;  // loss of the network on a round's samples before its first step
  int64_t mlp_last_samples = 0;
  int mlp_interval = 0;         // current training interval in steps (0: not started)
  int64_t mlp_next_train = 0;   // the step count at which the next round is due
  bool mixed = false;
  bool uniform_h = true;     // every particle has the same h (FP32 loops skip h_j)
  bool uniform_m = true;     // every particle has the same mass m_value (the
  double m_value = 0.0;      // surface pass skips the neighbour-mass gather)
  bool mirror_dirty = true;  // FP32 mirrors out of date with the FP64 state
  bool nb_built = false;
  bool snapshot_set = false;
  // the search reads the snapshot x_h under WF_Q_STALE_NS_POS: a rebuild
  // from an unchanged snapshot reproduces the current order and lists, so
  // it is skipped (snap_gen counts snapshot writes, built_gen is the one the
  // lists were built from)
  int64_t snap_gen = 0, built_gen = -1;
  BcTable bcs{};
  // rigid meshes (all concatenated) and contact state
  MeshDev mesh{};
  MeshMotions motions{};
  int n_meshes = 0;
  int mesh_node_off[kMaxMeshes + 1] = {0}, mesh_el_off[kMaxMeshes + 1] = {0};
  int mesh_zone[kMaxMeshes] = {0};
  std::vector<void*> mesh_owned;
  ContactLists cl{};
  int* cl_list_alt = nullptr;  // the candidate rows of the next order (reorder)
  int* cnb_act = nullptr;  // particles needing candidates: length, then indices
  bool cont_built = false;
  bool surf_init = false;  // not_write_surf_ID and total mass taken
  double totmass = 0.0;
  int *tri_cnt = nullptr, *tri_start = nullptr, *tri_rank = nullptr, *tri_sorted = nullptr;
  int* tri_scan_part = nullptr;
  int64_t tri_cells_cap = 0;
  double* cscratch = nullptr;   // device: centroid bbox, max h, block partials
  double* cscratch_h = nullptr; // pinned
  // run state (SolverFraser.cu:30-409)
  double time = 0.0;
  int64_t step = 0;
  int ts_i = 0;
  bool is_yielding = false;
  bool isfirst_step = true;
  int ns_builds = 0;
  int mlp_trainings = 0;
  // the spare plane the mixed acceleration pass writes the integrated
  // position mirror to (the velocity mirror goes to v4_alt)
  float4* p4_spare = nullptr;
  // timing: ev0/ev1 bracket wf_time_pass; each timed kernel launch gets an
  // event pair of the pool (no system-scope fence, no wait per launch), read
  // when the pool is full or the times are asked for
  bool timing = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  static constexpr int kTimerPool = 256;
  hipEvent_t tev[2][kTimerPool] = {};
  int tev_k[kTimerPool] = {0};
  int tev_n = 0;
  bool tev_pass = false;
  // before yield: the step's max eps_p, copied to ctr_host behind the stress
  // pass and read when the step's later passes are queued
  hipEvent_t ev_yield = nullptr;
  bool yield_pending = false;
  // yielding, under the plastic-zone trigger: the step's largest squared hot
  // displacement, copied to pz_d2_host behind the hot-displacement pass and
  // read once the step's conduction pass is queued
  hipEvent_t ev_pz = nullptr;
  bool pz_wait = false;
  double kt_ms[WF_K_COUNT_] = {0};
  int64_t kt_n[WF_K_COUNT_] = {0};

  StepParams params() const;
  PlaneRef plane(int field, int c);
  std::vector<PlaneRef> all_planes();
  int alloc_plane_d(double** p);
  int alloc_plane_i(int** p);
  int alloc_plane_f(float** p);
  void refresh_mirrors();
  int mlp_train_round(int iters, float lr);
  int relist_local();
  int mlp_fit(const float* X, const float* Y, int n, int stride, int iters, float lr);
  int dev_alloc(void** p, size_t bytes);
  int surface_init();
  int contact_nb();
  ContactParams contact_params() const;
};

int wf_domain::dev_alloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes) != hipSuccess) return WF_E_NOMEM;
  owned.push_back(*p);
  return WF_OK;
}

namespace {

int field_info(int field, int* comps, int* elem) {
  if (field >= WF_F_X && field <= WF_F_CONTFORCE) {
    *comps = 3;
    *elem = 8;
  } else if (field >= WF_F_RHO && field <= WF_F_Q_FRICTION) {
    *comps = 1;
    *elem = 8;
  } else if (field >= WF_F_SIGMA && field <= WF_F_ROTRATE) {
    *comps = 6;
    *elem = 8;
  } else if (field >= WF_F_ID && field <= WF_F_CONT_NB_COUNT) {
    *comps = 1;
    *elem = 4;
  } else {
    return WF_E_ARG;
  }
  return WF_OK;
}

// interleaved host order (original index) <-> planes (sorted order)
struct Planes6 {
  double* p[6];
};

__global__ void k_upload_f64(Planes6 planes, int comps, const double* stage, const int* perm,
                             int n) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int o = perm[s];
  for (int c = 0; c < comps; c++) planes.p[c][s] = stage[(int64_t)o * comps + c];
}
__global__ void k_download_f64(Planes6 planes, int comps, double* stage, const int* perm, int n) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int o = perm[s];
  for (int c = 0; c < comps; c++) stage[(int64_t)o * comps + c] = planes.p[c][s];
}
__global__ void k_upload_i32(int* plane, const int* stage, const int* perm, int n) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) plane[s] = stage[perm[s]];
}
// the same for a field held in FP32 planes (rounded on upload)
struct Planes6f {
  float* p[6];
};
__global__ void k_upload_f64_f32(Planes6f planes, int comps, const double* stage,
                                 const int* perm, int n) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int o = perm[s];
  for (int c = 0; c < comps; c++) planes.p[c][s] = (float)stage[(int64_t)o * comps + c];
}
__global__ void k_download_f32_f64(Planes6f planes, int comps, double* stage, const int* perm,
                                   int n) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  int o = perm[s];
  for (int c = 0; c < comps; c++) stage[(int64_t)o * comps + c] = (double)planes.p[c][s];
}
__global__ void k_fill_f32v(float* a, float v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
__global__ void k_download_i32(const int* plane, int* stage, const int* perm, int n) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < n) stage[perm[s]] = plane[s];
}
__global__ void k_iota(int* a, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = i;
}
__global__ void k_copy_f64(double* dst, const double* src, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}
__global__ void k_zero_bytes(unsigned char* p, int64_t nbytes) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nbytes) p[i] = 0;
}

int gridn(int64_t n) { return (int)((n + 255) / 256); }

}  // namespace

StepParams wf_domain::params() const {
  StepParams p{};
  p.n = n;
  p.solid = solid;
  p.dt = cfg.dt;
  p.alpha = cfg.alpha;
  p.beta = cfg.beta;
  p.quirks = cfg.quirks;
  p.return_map = cfg.return_map;
  p.taylor_quinney = cfg.taylor_quinney;
  p.T_inf = cfg.T_inf;
  p.h_conv = cfg.h_conv;
  p.newton_tol = cfg.newton_tol;
  p.newton_max_iter = cfg.newton_max_iter;
  const wf_material_params& m = cfg.mat;
  p.mat.model = m.model;
  p.mat.E = m.E;
  p.mat.Et = m.Et;
  // Ep = E*Et/(E-Et) (WeldFormGPU.cu:755), E if negative (Mechanical.cu:527)
  double Ep = (m.E != m.Et) ? m.E * m.Et / (m.E - m.Et) : m.E;
  if (Ep < 0) Ep = 1.0 * m.E;
  p.mat.Ep = Ep;
  p.mat.sy0 = m.sy0;
  p.mat.K_h = m.K_h;
  p.mat.m_h = m.m_h;
  p.mat.eps0_h = (m.E != 0) ? m.sy0 / m.E : 0.0;  // Material.cuh:57-59
  p.mat.eps1_h = (m.K_h > 0 && m.m_h != 0) ? std::pow(m.sy0 / m.K_h, 1. / m.m_h) : 0.0;
  p.mat.A = m.A;
  p.mat.B = m.B;
  p.mat.n = m.n;
  p.mat.C = m.C;
  p.mat.eps0 = m.eps0;
  p.mat.m = m.m;
  p.mat.Tm = m.Tm;
  p.mat.Tt = m.Tt;
  return p;
}

PlaneRef wf_domain::plane(int field, int c) {
  PlaneRef r;
  switch (field) {
    case WF_F_X: r.d = &f.x[c]; break;
    case WF_F_V: r.d = &f.v[c]; break;
    case WF_F_A: r.d = &f.a[c]; break;
    case WF_F_U: r.d = &f.u[c]; break;
    case WF_F_VA: r.d = &f.va[c]; break;
    case WF_F_X_SNAPSHOT: r.d = &f.xs[c]; break;
    case WF_F_NORMAL: r.d = &f.nrm[c]; break;
    case WF_F_CONTFORCE: r.d = &f.cf[c]; break;
    case WF_F_RHO: r.d = &f.rho; break;
    case WF_F_RHO0: r.d = &f.rho0; break;
    case WF_F_DRHO: r.d = &f.drho; break;
    case WF_F_M: r.d = &f.m; break;
    case WF_F_H: r.d = &f.h; break;
    case WF_F_P: r.d = &f.p; break;
    case WF_F_CS: r.d = &f.Cs; break;
    case WF_F_G: r.d = &f.G; break;
    case WF_F_SIGMA_Y: r.d = &f.sigma_y; break;
    case WF_F_PL_STRAIN: r.d = &f.pl_strain; break;
    case WF_F_SIGMA_EQ: r.d = &f.sigma_eq; break;
    case WF_F_EFF_STRAIN_RATE: r.d = &f.esr; break;
    case WF_F_INT_ENERGY: r.d = &f.eint; break;
    case WF_F_T: r.d = &f.T; break;
    case WF_F_TA: r.d = &f.Ta; break;
    case WF_F_TB: r.d = &f.Tb; break;
    case WF_F_DTDT: r.d = &f.dTdt; break;
    case WF_F_K_T: r.d = &f.k; break;
    case WF_F_CP_T: r.d = &f.cp; break;
    case WF_F_Q_PLASTIC: r.d = &f.qpl; break;
    case WF_F_Q_FRICTION: r.d = &f.qfric; break;
    // sigma, the deviatoric stress and the rates: FP32 planes in mixed precision
    case WF_F_SIGMA:
      if (mixed) r.f = &f.sigma32[c];
      else r.d = &f.sigma[c];
      break;
    case WF_F_SHEARSTRESS:
      if (mixed) r.f = &f.ss32[c];
      else r.d = &f.ss[c];
      break;
    case WF_F_STRAIN: r.d = &f.strain[c]; break;
    case WF_F_STRRATE:
      if (mixed) r.f = &f.strrate32[c];
      else r.d = &f.strrate[c];
      break;
    case WF_F_ROTRATE:
      if (mixed) r.f = &f.rotrate32[c];
      else r.d = &f.rotrate[c];
      break;
    case WF_F_ID: r.i = &f.ID; break;
    case WF_F_BC_T: r.i = &f.BC_T; break;
    case WF_F_NB_COUNT: r.i = &nb_count; break;
    case WF_F_CONT_NB_COUNT: r.i = &cl.count; break;
    default: break;
  }
  return r;
}

std::vector<PlaneRef> wf_domain::all_planes() {
  std::vector<PlaneRef> v;
  for (int fld = 0; fld < WF_F_COUNT_; fld++) {
    if (fld == WF_F_NB_COUNT || fld == WF_F_CONT_NB_COUNT) continue;  // derived, read-only
    int comps, elem;
    if (field_info(fld, &comps, &elem) != WF_OK) continue;
    for (int c = 0; c < comps; c++) {
      PlaneRef r = plane(fld, c);
      if (r.d || r.i || r.f) v.push_back(r);
    }
  }
  return v;
}

namespace {
__global__ void k_fill_i32v(int* a, int v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) a[i] = v;
}
template <class T>
struct GatherSet {
  const T* src[kGatherBatch];
  T* dst[kGatherBatch];
  int count;
};
// up to kGatherBatch planes into the new order in one pass: the map is read
// once and a particle's loads from the planes are independent
template <class T>
__global__ void k_gather_batch(GatherSet<T> g, const int* __restrict__ idx, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = idx[i];
  T v[kGatherBatch];
#pragma unroll
  for (int q = 0; q < kGatherBatch; q++)
    if (q < g.count) v[q] = g.src[q][j];
#pragma unroll
  for (int q = 0; q < kGatherBatch; q++)
    if (q < g.count) g.dst[q][i] = v[q];
}
// the contact candidate rows into the new order, only the entries below each
// particle's count (the rest are never read)
__global__ void k_gather_cont_rows(int* __restrict__ dst, const int* __restrict__ src,
                                   const int* __restrict__ cnt, const int* __restrict__ idx,
                                   int n, int64_t npad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int j = idx[i];
  const int c = cnt[j];
  for (int k = 0; k < c; k++) dst[k * npad + i] = src[k * npad + j];
}
// splitmix64, for the deterministic initial weights
uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
}  // namespace

int wf_domain::alloc_plane_d(double** p) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(double) * npad) != hipSuccess) return WF_E_NOMEM;
  owned.push_back(q);
  *p = (double*)q;
  hipLaunchKernelGGL(k_zero_bytes, dim3(gridn(sizeof(double) * npad)), dim3(256), 0, stream,
                     (unsigned char*)q, (int64_t)(sizeof(double) * npad));
  return WF_OK;
}
int wf_domain::alloc_plane_f(float** p) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(float) * npad) != hipSuccess) return WF_E_NOMEM;
  owned.push_back(q);
  *p = (float*)q;
  hipLaunchKernelGGL(k_zero_bytes, dim3(gridn(sizeof(float) * npad)), dim3(256), 0, stream,
                     (unsigned char*)q, (int64_t)(sizeof(float) * npad));
  return WF_OK;
}
// rebuild every FP32 mirror (mixed) or FP64 neighbour record when the FP64
// state changed outside the step's own producers (uploads, rebuilds,
// single-kernel calls)
void wf_domain::refresh_mirrors() {
  if (!mirror_dirty || !nb_built) return;
  if (mixed) launch_mirror32(stream, f, n, frame);
  else launch_records64(stream, f, n);
  mirror_dirty = false;
}
int wf_domain::alloc_plane_i(int** p) {
  void* q = nullptr;
  if (hipMalloc(&q, sizeof(int) * npad) != hipSuccess) return WF_E_NOMEM;
  owned.push_back(q);
  *p = (int*)q;
  hipLaunchKernelGGL(k_zero_bytes, dim3(gridn(sizeof(int) * npad)), dim3(256), 0, stream,
                     (unsigned char*)q, (int64_t)(sizeof(int) * npad));
  return WF_OK;
}

// one online training round on the logged samples (ascending particle index,
// at most 16384 taken with a fixed stride)
int wf_domain::mlp_train_round(int iters, float lr) {
  launch_mlp_collect(stream, f, n, mlp_X, mlp_Y, mlp_flags, mlp_scan, mlp_scan_part, mlp_count);
  HIPC(hipMemcpyAsync(host_i, mlp_count, sizeof(int), hipMemcpyDeviceToHost, stream));
  HIPC(hipStreamSynchronize(stream));
  const int count = host_i[0];
  mlp_last_samples = count;
  if (count < 64) return WF_OK;
  const int ntrain = std::min(count, 16384);
  const int stride = count / ntrain;
  int rc = mlp_fit(mlp_X, mlp_Y, ntrain, stride, iters, lr);
  if (rc) return rc;
  mlp_trained = true;
  mlp_trainings++;
  return WF_OK;
}

// Adam iterations on device-resident samples X[s*stride], Y[s*stride]
int wf_domain::mlp_fit(const float* X, const float* Y, int nsamp, int stride, int iters,
                       float lr) {
  const int nblocks = mlp_grad_blocks(nsamp);
  for (int it = 0; it < iters; it++) {
    launch_mlp_train_iter(stream, mlp_w, mlp_m, mlp_v, X, Y, nsamp, stride, mlp_part,
                          mlp_loss_part, nblocks, ++adam_t, lr, mlp_loss);
    mlp_pack_stale = true;
    // the first iteration's loss is that of the weights the round started from
    if (it == 0)
      HIPC(hipMemcpyAsync(mlp_loss + 1, mlp_loss, sizeof(float), hipMemcpyDeviceToDevice,
                          stream));
  }
  HIPC(hipMemcpyAsync(host_f, mlp_loss, 2 * sizeof(float), hipMemcpyDeviceToHost, stream));
  HIPC(hipStreamSynchronize(stream));
  mlp_last_loss = host_f[0];
  mlp_start_loss = iters > 0 ? host_f[1] : host_f[0];
  return WF_OK;
}

// not_write_surf_ID (ID != 0) and the total mass, taken at the first surface
// pass after IDs or masses were set (WeldFormGPU.cu:808-858)
int wf_domain::surface_init() {
  if (surf_init) return WF_OK;
  launch_surface_lock(stream, f, n);
  const int nb = std::min(1024, std::max(1, gridn(n)));
  launch_sum(stream, f.m, n, sum_part, nb, sum_out);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(cscratch_h, sum_out, sizeof(double), hipMemcpyDeviceToHost, stream));
  HIPC(hipStreamSynchronize(stream));
  totmass = cscratch_h[0];
  surf_init = true;
  return WF_OK;
}

// contact candidates: bin the triangle centroids on a grid of edge >= 2 max h,
// then one thread per particle scans its 3x3x3 block of cells
int wf_domain::contact_nb() {
  constexpr int64_t kMaxTriCells = 1 << 22;
  launch_tri_bbox(stream, mesh, cscratch);
  launch_max_f64(stream, f.h, solid, cscratch + 6, cscratch + 8);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(cscratch_h, cscratch, sizeof(double) * 7, hipMemcpyDeviceToHost, stream));
  HIPC(hipStreamSynchronize(stream));
  const double* b = cscratch_h;
  const double hmax = b[6];
  if (!(hmax > 0.0) || !std::isfinite(hmax)) return fail(WF_E_ARG, "smoothing length h must be positive");
  for (int c = 0; c < 6; c++)
    if (!std::isfinite(b[c])) return fail(WF_E_NONFINITE, "non-finite rigid mesh node");
  TriGrid g{};
  double cs = 2.0 * hmax * (1.0 + 1e-6);
  int64_t cells = 0;
  for (;;) {
    cells = 1;
    for (int c = 0; c < 3; c++) {
      const double ext = b[3 + c] - b[c];
      const double k = std::floor(ext / cs) + 1.0;
      g.dim[c] = k > 1e6 ? 1000001 : (int)k;
      cells *= g.dim[c];
    }
    if (cells <= kMaxTriCells) break;
    cs *= 1.26;
  }
  for (int c = 0; c < 3; c++) g.org[c] = b[c];
  g.inv_cs = 1.0 / cs;
  if (cells + 1 > tri_cells_cap) {
    const int64_t cap = std::max<int64_t>(cells + 1, 4096);
    int *cnt = nullptr, *start = nullptr, *part = nullptr;
    if (hipMalloc(&cnt, sizeof(int) * cap) != hipSuccess ||
        hipMalloc(&start, sizeof(int) * (cap + 1)) != hipSuccess ||
        hipMalloc(&part, sizeof(int) * (cap / 1024 + 4)) != hipSuccess) {
      hipFree(cnt);
      hipFree(start);
      return fail(WF_E_NOMEM, "contact grid allocation failed");
    }
    for (void* q : {(void*)cnt, (void*)start, (void*)part}) mesh_owned.push_back(q);
    tri_cnt = cnt;
    tri_start = start;
    tri_scan_part = part;
    tri_cells_cap = cap;
  }
  hipLaunchKernelGGL(k_fill_i32v, dim3(gridn(cells)), dim3(256), 0, stream, tri_cnt, 0,
                     (int)cells);
  launch_tri_bin(stream, mesh, g, tri_cnt, tri_rank, tri_start, tri_sorted, tri_scan_part);
  hipLaunchKernelGGL(k_fill_i32v, dim3(1), dim3(64), 0, stream, cnb_act, 0, 1);
  launch_contact_nb(stream, f, solid, mesh, g, tri_start, tri_sorted, cl, cfg.id_free_surf,
                    cnb_act, ctr);
  HIPC(hipGetLastError());
  cont_built = true;
  return WF_OK;
}

ContactParams wf_domain::contact_params() const {
  ContactParams cp{};
  cp.dt = cfg.dt;
  cp.kappa_E_dx = cfg.contact_kappa * cfg.mat.E * cfg.contact_dx;
  cp.mu_s = cfg.fric_sta;
  cp.mu_d = cfg.fric_dyn;
  cp.fric_heat = cfg.fric_heat;
  cp.law = cfg.contact_law;
  return cp;
}

// ============================================================================
// the planes into the new order, kGatherBatch per launch; each gathered plane
// swaps places with a scratch plane
template <class T>
static void gather_planes(wf_domain* d, const std::vector<T**>& planes, T** spares) {
  for (size_t b = 0; b < planes.size(); b += kGatherBatch) {
    GatherSet<T> g{};
    g.count = (int)std::min<size_t>(kGatherBatch, planes.size() - b);
    for (int q = 0; q < g.count; q++) {
      g.src[q] = *planes[b + q];
      g.dst[q] = spares[q];
    }
    hipLaunchKernelGGL(k_gather_batch<T>, dim3(gridn(d->n)), dim3(256), 0, d->stream, g,
                       d->gather, d->n);
    for (int q = 0; q < g.count; q++) std::swap(*planes[b + q], spares[q]);
  }
}

extern "C" {

void wf_config_default(wf_config* c) {
  if (!c) return;
  std::memset(c, 0, sizeof(*c));
  c->abi_version = WF_ABI_VERSION;
  c->precision = WF_FP64;
  c->alpha = 1.0;  // Domain_d.cu:191-192
  c->beta = 0.0;
  c->return_map = WF_RETURN_REFERENCE;
  c->quirks = WF_Q_ALL;
  c->mechanics = 1;
  c->thermal = 0;
  c->T_inf = 500.0;  // Thermal.cu:203-204
  c->h_conv = 100.0;
  c->ns_interval = 5;        // SolverFraser.cu:91
  c->ns_min_plastic = 1e-6;  // SolverFraser.cu:17
  c->newton_max_iter = 50;
  c->newton_tol = 1e-11;
  c->mlp_train_interval = 200;
  c->mlp_train_iters = 250;
  c->mlp_lr = 2e-3;
  c->mlp_interval_max = 0;
  c->mlp_refit_ratio = 2.0;
  c->contact = 0;
  c->contact_law = WF_CONTACT_REFERENCE;
  c->contact_kappa = 0.1;
  c->id_free_surf = 1;  // SolverFraser.cu:120
  c->rebuild_policy = WF_REBUILD_EVERY_N;
  c->fric_heat = 0.0;
  c->ns_skin = 0.0;
  c->ep_hot = 0.05;
  c->ns_global_every = 32;
  c->fuse_integrate = 1;
}

int wf_field_shape(int field, int* components, int* elem_bytes) {
  int c = 0, e = 0;
  if (field_info(field, &c, &e) != WF_OK) return fail(WF_E_ARG, "unknown field");
  if (components) *components = c;
  if (elem_bytes) *elem_bytes = e;
  return WF_OK;
}

const char* wf_last_error(void) { return g_err.c_str(); }

int wf_create(const wf_config* cfg, wf_domain** out) {
  if (!cfg || !out) return fail(WF_E_ARG, "null argument");
  if (cfg->abi_version != WF_ABI_VERSION) return fail(WF_E_ARG, "abi_version mismatch");
  if (cfg->n_particles <= 0) return fail(WF_E_ARG, "n_particles must be positive");
  // the pair loops address per-particle records (up to 96 B) with 32-bit
  // byte offsets
  if (cfg->n_particles > kMaxParticles)
    return fail(WF_E_ARG, "n_particles exceeds 40,000,000");
  if (cfg->ns_compact_max < 0) return fail(WF_E_ARG, "ns_compact_max must be >= 0");
  if (cfg->solid_count < 0 || cfg->solid_count > cfg->n_particles)
    return fail(WF_E_ARG, "solid_count out of range");
  if (cfg->precision != WF_FP64 && cfg->precision != WF_MIXED)
    return fail(WF_E_ARG, "unknown precision mode");
  if (cfg->ns_interval <= 0) return fail(WF_E_ARG, "ns_interval must be positive");
  if (cfg->contact_law != WF_CONTACT_REFERENCE && cfg->contact_law != WF_CONTACT_STIFFNESS)
    return fail(WF_E_ARG, "unknown contact law");
  if (!(cfg->fric_heat >= 0.0 && cfg->fric_heat <= 1.0))
    return fail(WF_E_ARG, "fric_heat must lie in [0, 1]");
  if (cfg->rebuild_policy != WF_REBUILD_EVERY_N && cfg->rebuild_policy != WF_REBUILD_PLASTIC_ZONE &&
      cfg->rebuild_policy != WF_REBUILD_PLASTIC_ZONE_LOCAL)
    return fail(WF_E_ARG, "unknown rebuild policy");
  if (cfg->rebuild_policy != WF_REBUILD_EVERY_N && !(cfg->ns_skin > 0.0 && cfg->ep_hot >= 0.0))
    return fail(WF_E_ARG, "the plastic-zone rebuild needs ns_skin > 0 and ep_hot >= 0");
  if (cfg->rebuild_policy == WF_REBUILD_PLASTIC_ZONE_LOCAL && cfg->ns_skin > 2.0 * cfg->h_glob)
    return fail(WF_E_ARG, "the local re-list reaches two cells: ns_skin must not exceed 2 h_glob");
  if (cfg->mlp_train_iters < 0 || !(cfg->mlp_lr > 0.0) || cfg->mlp_interval_max < 0 ||
      !(cfg->mlp_refit_ratio >= 0.0))
    return fail(WF_E_ARG, "surrogate training: iters >= 0, lr > 0, interval cap >= 0, ratio >= 0");
  if (cfg->contact && cfg->contact_law == WF_CONTACT_STIFFNESS &&
      !(cfg->contact_kappa > 0.0 && cfg->contact_dx > 0.0))
    return fail(WF_E_ARG, "the stiffness contact law needs contact_kappa > 0 and contact_dx > 0");
  int dev_count = 0;
  if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0)
    return fail(WF_E_HIP, "no HIP device");
  wf_domain* d = new wf_domain();
  d->cfg = *cfg;
  d->n = cfg->n_particles;
  d->solid = cfg->solid_count ? cfg->solid_count : cfg->n_particles;
  d->npad = ((int64_t)d->n + 63) / 64 * 64;
  if (hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking) != hipSuccess) {
    delete d;
    return fail(WF_E_HIP, "hipStreamCreate failed");
  }
  int rc = WF_OK;
  d->mixed = cfg->precision == WF_MIXED;  // before the planes: it selects their precision
  for (PlaneRef r : d->all_planes()) {
    rc = r.d ? d->alloc_plane_d(r.d) : r.f ? d->alloc_plane_f(r.f) : d->alloc_plane_i(r.i);
    if (rc) break;
  }
  if (d->mixed) {
    if (!rc) rc = d->alloc_plane_f(&d->f.h32);
    if (!rc) rc = d->dev_alloc((void**)&d->f.p4, sizeof(float4) * d->npad);
    if (!rc) rc = d->dev_alloc((void**)&d->p4_spare, sizeof(float4) * d->npad);
    float4** f4[] = {&d->f.v4, &d->f.v4_alt};
    for (float4** q : f4)
      if (!rc) rc = d->dev_alloc((void**)q, sizeof(float4) * d->npad);
    if (!rc) rc = d->dev_alloc((void**)&d->f.arec, 2 * sizeof(float4) * d->npad);
    if (!rc) rc = d->dev_alloc((void**)&d->f.tk4, sizeof(float4) * d->npad);
    if (!rc) rc = d->dev_alloc((void**)&d->f.tk4_alt, sizeof(float4) * d->npad);
  } else {
    double4** d4[] = {&d->f.kv, &d->f.kv_alt, &d->f.tk64, &d->f.tk64_alt};
    for (double4** q : d4)
      if (!rc) rc = d->dev_alloc((void**)q, sizeof(double4) * d->npad);
    if (!rc) rc = d->dev_alloc((void**)&d->f.arec64, 3 * sizeof(double4) * d->npad);
  }
  // return-map statistics, surrogate samples and weights
  d->jc_newton = cfg->return_map == WF_RETURN_NEWTON && cfg->mat.model == WF_MAT_JOHNSON_COOK;
  d->nblk = (int)((d->npad + kBlock - 1) / kBlock);
  if (!rc) rc = d->dev_alloc((void**)&d->bstats, sizeof(BlockStats) * (d->nblk + 1));
  for (int k = 0; k < 4 && !rc; k++) rc = d->alloc_plane_f(&d->f.mlx[k]);
  if (!rc) rc = d->alloc_plane_f(&d->f.mly);
  if (!rc) rc = d->alloc_plane_i(&d->f.mlstep);
  if (!rc) {
    const size_t P = sizeof(float) * kMlpParams;
    rc = d->dev_alloc((void**)&d->mlp_w, P);
    if (!rc) rc = d->dev_alloc(&d->mlp_pack, (size_t)16 * 64 * kMlpPackWords);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_m, P);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_v, P);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_grad, P);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_part, P * kMlpMaxBlocks);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_loss_part, sizeof(float) * kMlpMaxBlocks);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_loss, sizeof(float) * 4);
    d->mlp_cap = std::max<int64_t>(d->npad, 65536);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_X, sizeof(float) * 4 * d->mlp_cap);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_Y, sizeof(float) * d->mlp_cap);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_flags, sizeof(int) * d->npad);
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_scan, sizeof(int) * (d->npad + 1));
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_scan_part, sizeof(int) * (d->npad / 1024 + 4));
    if (!rc) rc = d->dev_alloc((void**)&d->mlp_count, sizeof(int) * 4);
  }
  if (!rc && hipHostMalloc(&d->host_i, sizeof(int) * 16) != hipSuccess) rc = WF_E_NOMEM;
  if (!rc && hipHostMalloc(&d->host_f, sizeof(float) * 16) != hipSuccess) rc = WF_E_NOMEM;
  if (!rc) {
    // Xavier-uniform weights from a fixed seed, zero biases, zero Adam moments
    std::vector<float> w(kMlpParams, 0.f);
    uint64_t seed = 0x5EEDF00DULL;
    auto uni = [&](float lim) {
      return (float)(((splitmix64(seed) >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * lim;
    };
    for (int k = 0; k < 128; k++) w[kMlpOffW1 + k] = uni(std::sqrt(6.f / (4 + 32)));
    for (int k = 0; k < 1024; k++) w[kMlpOffW2 + k] = uni(std::sqrt(6.f / (32 + 32)));
    for (int k = 0; k < 32; k++) w[kMlpOffW3 + k] = uni(std::sqrt(6.f / (32 + 1)));
    std::vector<float> z(kMlpParams, 0.f);
    if (hipMemcpy(d->mlp_w, w.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) ||
        hipMemcpy(d->mlp_m, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice) ||
        hipMemcpy(d->mlp_v, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice))
      rc = WF_E_HIP;
    hipLaunchKernelGGL(k_fill_i32v, dim3(gridn(d->npad)), dim3(256), 0, d->stream, d->f.mlstep, -1,
                       (int)d->npad);
  }
  if (!rc) rc = d->alloc_plane_d(&d->f.rho_alt);
  if (!rc) rc = d->alloc_plane_d(&d->f.T_alt);
  for (int q = 0; q < kGatherBatch && !rc; q++) rc = d->alloc_plane_d(&d->spare_d[q]);
  for (int q = 0; q < kGatherBatch && !rc; q++) rc = d->alloc_plane_f(&d->spare_f[q]);
  if (!rc) rc = d->alloc_plane_i(&d->spare_i);
  if (!rc) rc = d->alloc_plane_i(&d->perm);
  if (!rc) rc = d->alloc_plane_i(&d->perm_tmp);
  if (!rc) rc = d->alloc_plane_i(&d->gather);
  if (!rc) rc = d->alloc_plane_i(&d->nb_count);
  if (!rc) rc = d->alloc_plane_i(&d->f.idlock);
  if (!rc) rc = d->dev_alloc((void**)&d->cscratch, sizeof(double) * (8 + kMaxParts));
  if (!rc && hipHostMalloc(&d->cscratch_h, sizeof(double) * 8) != hipSuccess) rc = WF_E_NOMEM;
  if (d->plastic_zone()) {
    for (int c = 0; c < 3 && !rc; c++) rc = d->alloc_plane_d(&d->xb[c]);
    if (!rc) rc = d->dev_alloc((void**)&d->pz_d2, sizeof(unsigned long long));
    if (!rc && hipHostMalloc(&d->pz_d2_host, sizeof(unsigned long long)) != hipSuccess)
      rc = WF_E_NOMEM;
  }
  if (!rc) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(double) * 6 * (int64_t)d->n) != hipSuccess) rc = WF_E_NOMEM;
    else {
      d->owned.push_back(q);
      d->stage = q;
    }
  }
  if (!rc) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(Counters)) != hipSuccess) rc = WF_E_NOMEM;
    else {
      d->owned.push_back(q);
      d->ctr = (Counters*)q;
      hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(256), 0, d->stream, (unsigned char*)q,
                         (int64_t)sizeof(Counters));
    }
  }
  if (!rc && hipHostMalloc(&d->ctr_host, sizeof(Counters)) != hipSuccess) rc = WF_E_NOMEM;
  if (!rc) {
    void* q = nullptr;
    if (hipMalloc(&q, sizeof(double) * 1025) != hipSuccess) rc = WF_E_NOMEM;
    else {
      d->owned.push_back(q);
      d->sum_part = (double*)q;
      d->sum_out = d->sum_part + 1024;
    }
  }
  if (!rc && hipEventCreateWithFlags(&d->ev_yield, hipEventDisableTiming) != hipSuccess)
    rc = WF_E_HIP;
  if (!rc && hipEventCreateWithFlags(&d->ev_pz, hipEventDisableTiming) != hipSuccess)
    rc = WF_E_HIP;
  if (!rc && (hipEventCreate(&d->ev0) != hipSuccess || hipEventCreate(&d->ev1) != hipSuccess))
    rc = WF_E_HIP;
  for (int b = 0; b < 2 && !rc; b++)
    for (int e = 0; e < wf_domain::kTimerPool && !rc; e++)
      if (hipEventCreateWithFlags(&d->tev[b][e], hipEventDisableSystemFence) != hipSuccess)
        rc = WF_E_HIP;
  if (rc) {
    wf_destroy(d);
    return fail(rc, "device allocation failed");
  }
  hipLaunchKernelGGL(k_iota, dim3(gridn(d->n)), dim3(256), 0, d->stream, d->perm, d->n);
  // thermal defaults so an unused T never poisons the JC law: T = 0
  if (hipStreamSynchronize(d->stream) != hipSuccess) {
    wf_destroy(d);
    return fail(WF_E_HIP, "initialisation failed");
  }
  *out = d;
  return WF_OK;
}

int wf_destroy(wf_domain* d) {
  if (!d) return WF_OK;
  if (d->stream) hipStreamSynchronize(d->stream);
  for (void* p : d->owned) hipFree(p);
  if (d->ctr_host) hipHostFree(d->ctr_host);
  if (d->host_i) hipHostFree(d->host_i);
  if (d->host_f) hipHostFree(d->host_f);
  if (d->cscratch_h) hipHostFree(d->cscratch_h);
  if (d->pz_d2_host) hipHostFree(d->pz_d2_host);
  for (void* p : d->mesh_owned) hipFree(p);
  ns_free(d->ns);
  if (d->ev0) hipEventDestroy(d->ev0);
  if (d->ev1) hipEventDestroy(d->ev1);
  for (int b = 0; b < 2; b++)
    for (int e = 0; e < wf_domain::kTimerPool; e++)
      if (d->tev[b][e]) hipEventDestroy(d->tev[b][e]);
  if (d->ev_yield) hipEventDestroy(d->ev_yield);
  if (d->ev_pz) hipEventDestroy(d->ev_pz);
  if (d->stream) hipStreamDestroy(d->stream);
  delete d;
  return WF_OK;
}

int wf_sync(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  HIPC(hipStreamSynchronize(d->stream));
  return WF_OK;
}

int wf_set_field(wf_domain* d, int field, const void* host, size_t bytes) {
  if (!d || !host) return fail(WF_E_ARG, "null argument");
  int comps, elem;
  if (field_info(field, &comps, &elem)) return fail(WF_E_ARG, "unknown field");
  if (bytes != (size_t)d->n * comps * elem) return fail(WF_E_ARG, "size mismatch");
  if (field == WF_F_NB_COUNT || field == WF_F_CONT_NB_COUNT)
    return fail(WF_E_ARG, "read-only field");
  if (field == WF_F_ID || field == WF_F_M) d->surf_init = false;
  if (field == WF_F_H) {
    const double* hv = (const double*)host;
    d->uniform_h = true;
    for (int i = 1; i < d->n && d->uniform_h; i++) d->uniform_h = hv[i] == hv[0];
  }
  if (field == WF_F_M) {
    const double* mv = (const double*)host;
    d->uniform_m = true;
    d->m_value = mv[0];
    for (int i = 1; i < d->n && d->uniform_m; i++) d->uniform_m = mv[i] == mv[0];
  }
  HIPC(hipMemcpyAsync(d->stage, host, bytes, hipMemcpyHostToDevice, d->stream));
  if (elem == 8 && d->plane(field, 0).f) {
    Planes6f planes{};
    for (int c = 0; c < comps; c++) planes.p[c] = *d->plane(field, c).f;
    hipLaunchKernelGGL(k_upload_f64_f32, dim3(gridn(d->n)), dim3(256), 0, d->stream, planes,
                       comps, (const double*)d->stage, d->perm, d->n);
    if (field == WF_F_ROTRATE)  // the antisymmetric W's diagonal stays zero
      for (int c = 0; c < 3; c++)
        hipLaunchKernelGGL(k_zero_bytes, dim3(gridn(sizeof(float) * d->npad)), dim3(256), 0,
                           d->stream, (unsigned char*)d->f.rotrate32[c],
                           (int64_t)sizeof(float) * d->npad);
  } else if (elem == 8) {
    Planes6 planes{};
    for (int c = 0; c < comps; c++) planes.p[c] = *d->plane(field, c).d;
    hipLaunchKernelGGL(k_upload_f64, dim3(gridn(d->n)), dim3(256), 0, d->stream, planes, comps,
                       (const double*)d->stage, d->perm, d->n);
    // the rotation rate is antisymmetric: its diagonal planes stay zero (the
    // pair and stress kernels neither write nor read them)
    if (field == WF_F_ROTRATE)
      for (int c = 0; c < 3; c++)
        hipLaunchKernelGGL(k_zero_bytes, dim3(gridn(sizeof(double) * d->npad)), dim3(256), 0,
                           d->stream, (unsigned char*)d->f.rotrate[c],
                           (int64_t)sizeof(double) * d->npad);
  } else {
    hipLaunchKernelGGL(k_upload_i32, dim3(gridn(d->n)), dim3(256), 0, d->stream,
                       *d->plane(field, 0).i, (const int*)d->stage, d->perm, d->n);
  }
  if (field == WF_F_X_SNAPSHOT) d->snap_gen++;
  // x also initialises the search snapshot x_h (SolverFraser.cu:56)
  if (field == WF_F_X) {
    for (int c = 0; c < 3; c++)
      hipLaunchKernelGGL(k_copy_f64, dim3(gridn(d->n)), dim3(256), 0, d->stream, d->f.xs[c],
                         d->f.x[c], d->n);
    d->snapshot_set = true;
    d->snap_gen++;
  }
  d->mirror_dirty = true;
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(d->stream));
  return WF_OK;
}

int wf_get_field(wf_domain* d, int field, void* host, size_t bytes) {
  if (!d || !host) return fail(WF_E_ARG, "null argument");
  int comps, elem;
  if (field_info(field, &comps, &elem)) return fail(WF_E_ARG, "unknown field");
  if (bytes != (size_t)d->n * comps * elem) return fail(WF_E_ARG, "size mismatch");
  if (elem == 4 && !*d->plane(field, 0).i) {  // contact lists not allocated: none
    std::memset(host, 0, bytes);
    return WF_OK;
  }
  if (elem == 8 && d->plane(field, 0).f) {
    Planes6f planes{};
    for (int c = 0; c < comps; c++) planes.p[c] = *d->plane(field, c).f;
    hipLaunchKernelGGL(k_download_f32_f64, dim3(gridn(d->n)), dim3(256), 0, d->stream, planes,
                       comps, (double*)d->stage, d->perm, d->n);
  } else if (elem == 8) {
    Planes6 planes{};
    for (int c = 0; c < comps; c++) planes.p[c] = *d->plane(field, c).d;
    hipLaunchKernelGGL(k_download_f64, dim3(gridn(d->n)), dim3(256), 0, d->stream, planes, comps,
                       (double*)d->stage, d->perm, d->n);
  } else {
    hipLaunchKernelGGL(k_download_i32, dim3(gridn(d->n)), dim3(256), 0, d->stream,
                       *d->plane(field, 0).i, (int*)d->stage, d->perm, d->n);
  }
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(host, d->stage, bytes, hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  return WF_OK;
}

int wf_fill_field(wf_domain* d, int field, double value) {
  if (!d) return fail(WF_E_ARG, "null domain");
  int comps, elem;
  if (field_info(field, &comps, &elem) || elem != 8) return fail(WF_E_ARG, "not a double field");
  for (int c = 0; c < comps; c++) {
    const PlaneRef r = d->plane(field, c);
    if (r.f)
      hipLaunchKernelGGL(k_fill_f32v, dim3(gridn(d->n)), dim3(256), 0, d->stream, *r.f,
                         (float)value, d->n);
    else
      launch_fill_f64(d->stream, *r.d, value, d->n);
  }
  if (field == WF_F_ROTRATE)  // the antisymmetric W's diagonal stays zero
    for (int c = 0; c < 3; c++) {
      const PlaneRef r = d->plane(field, c);
      if (r.f)
        hipLaunchKernelGGL(k_fill_f32v, dim3(gridn(d->n)), dim3(256), 0, d->stream, *r.f, 0.f,
                           d->n);
      else
        launch_fill_f64(d->stream, *r.d, 0.0, d->n);
    }
  if (field == WF_F_M) {
    d->surf_init = false;
    d->uniform_m = true;
    d->m_value = value;
  }
  if (field == WF_F_H) d->uniform_h = true;
  if (field == WF_F_X) {
    for (int c = 0; c < 3; c++) launch_fill_f64(d->stream, d->f.xs[c], value, d->n);
    d->snap_gen++;
  }
  if (field == WF_F_X_SNAPSHOT) d->snap_gen++;
  d->mirror_dirty = true;
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(d->stream));
  return WF_OK;
}

int wf_snapshot_positions(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  for (int c = 0; c < 3; c++)
    hipLaunchKernelGGL(k_copy_f64, dim3(gridn(d->n)), dim3(256), 0, d->stream, d->f.xs[c],
                       d->f.x[c], d->n);
  d->snapshot_set = true;
  d->snap_gen++;
  HIPC(hipGetLastError());
  return WF_OK;
}

int wf_set_bcs(wf_domain* d, int count, const int* zone_ids, const double* values) {
  if (!d) return fail(WF_E_ARG, "null domain");
  if (count < 0 || count > WF_MAX_BCS) return fail(WF_E_ARG, "too many boundary conditions");
  if (count > 0 && (!zone_ids || !values)) return fail(WF_E_ARG, "null argument");
  d->bcs.count = count;
  for (int b = 0; b < count; b++) {
    d->bcs.zone[b] = zone_ids[b];
    for (int c = 0; c < 3; c++) d->bcs.val[b][c] = values[3 * b + c];
  }
  return WF_OK;
}

// add the recorded event pairs to the kernel totals (waits for the last)
static int timer_resolve(wf_domain* d) {
  if (d->tev_n == 0) return WF_OK;
  HIPC(hipEventSynchronize(d->tev[1][d->tev_n - 1]));
  for (int e = 0; e < d->tev_n; e++) {
    float ms = 0.f;
    HIPC(hipEventElapsedTime(&ms, d->tev[0][e], d->tev[1][e]));
    d->kt_ms[d->tev_k[e]] += ms;
    d->kt_n[d->tev_k[e]] += 1;
  }
  d->tev_n = 0;
  return WF_OK;
}
// pass: the region is one pair pass, timed by its own dispatch timestamps;
// otherwise by events recorded before and after the region's launches
static void timer_begin(wf_domain* d, bool pass = false) {
  if (!d->timing) return;
  if (d->tev_n == wf_domain::kTimerPool) timer_resolve(d);
  d->tev_pass = pass;
  if (pass) g_next_launch = LaunchEvents{d->tev[0][d->tev_n], d->tev[1][d->tev_n]};
  else hipEventRecord(d->tev[0][d->tev_n], d->stream);
}
static int timer_end(wf_domain* d, int k) {
  if (!d->timing) return WF_OK;
  if (d->tev_pass) {
    // nothing was launched (an empty pass): nothing to count
    const bool launched = g_next_launch.start == nullptr;
    g_next_launch = LaunchEvents{nullptr, nullptr};
    if (!launched) return WF_OK;
  } else {
    HIPC(hipEventRecord(d->tev[1][d->tev_n], d->stream));
  }
  d->tev_k[d->tev_n++] = k;
  return WF_OK;
}

// WF_REBUILD_PLASTIC_ZONE_LOCAL: instead of a rebuild (cell sort, reorder of
// every field, every tile's lists), re-list in place the tiles within two
// search cells of a particle with eps_p > ep_hot, from a cell sort of the
// current positions that only indexes the particles.  Their build
// positions restart.  Every hot particle lies in such a tile, so after a
// re-list every hot particle's displacement counts from its new lists, and
// a tile left alone is farther than 2h + skin from any hot particle.
// Returns WF_E_STATE when a full rebuild is due instead: under the other
// policies, before the first build, after ns_global_every re-lists (the
// tiles drift out of their cells as the particles stay in place), or when a
// re-listed tile outgrows the staging region its last rebuild gave it.
int wf_domain::relist_local() {
  if (cfg.rebuild_policy != WF_REBUILD_PLASTIC_ZONE_LOCAL || !nb_built) return WF_E_STATE;
  const int every = cfg.ns_global_every > 0 ? cfg.ns_global_every : 32;
  if (local_since_reorder >= every) return WF_E_STATE;
  timer_begin(this);
  const double* xs[3] = {f.x[0], f.x[1], f.x[2]};
  const double radius = 2.0 * cfg.h_glob + cfg.ns_skin;
  int tiles = 0;
  int rc = ns_relist_local(stream, ns, xs, f.pl_strain, cfg.ep_hot, n, npad, grid, radius,
                           kListCap, perm, nb_count, ctr, xb, &tiles);
  if (rc == WF_E_NBCAP) return WF_E_STATE;  // a tile outgrew its region: full rebuild
  if (rc) return fail(rc, "local neighbour re-list failed");
  pz_pending = false;
  local_since_reorder++;
  ns_local_builds++;
  ns_local_tiles += tiles;
  return timer_end(this, WF_K_NSEARCH);
}

int wf_nsearch_build(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  // the plastic-zone policy builds from the current positions with a skin
  const bool pz = d->plastic_zone();
  const bool stale = !pz && (d->cfg.quirks & WF_Q_STALE_NS_POS) != 0;
  if (stale && d->nb_built && d->built_gen == d->snap_gen) {
    d->ns_builds++;
    return WF_OK;
  }
  timer_begin(d);
  const double* src[3];
  for (int c = 0; c < 3; c++) src[c] = stale ? d->f.xs[c] : d->f.x[c];
  const double radius = 2.0 * d->cfg.h_glob + (pz ? d->cfg.ns_skin : 0.0);
  ns_set_compact_max(&d->ns, d->cfg.ns_compact_max);
  // mixed precision keeps each particle's cell packed in 11/11/10 bits
  int rc = ns_build(d->stream, &d->ns, src, d->n, d->npad, radius, pz ? kListCap : kMaxNb,
                    d->gather, d->nb_count, d->ctr, d->perm, d->perm_tmp, &d->tl,
                    &d->grid, d->mixed ? kPackedCells : nullptr);
  if (rc)
    return fail(rc, rc == WF_E_NONFINITE  ? "non-finite particle position"
                    : rc == WF_E_NBCAP    ? "a neighbour tile exceeds its cell/slot budget"
                    : rc == WF_E_UNSUPPORTED
                        ? "mixed precision packs a particle's search cell in 11/11/10 bits: "
                          "the domain spans more than 2048 x 2048 x 1024 cells of 2h"
                        : "neighbour search failed");
  // physically reorder every field into the new order
  std::vector<double**> pd;
  std::vector<float**> pf;
  for (PlaneRef r : d->all_planes()) {
    if (r.d) {
      pd.push_back(r.d);
    } else if (r.f) {
      pf.push_back(r.f);
    } else {
      launch_gather_i32(d->stream, d->spare_i, *r.i, d->gather, d->n);
      std::swap(*r.i, d->spare_i);
    }
  }
  for (int k = 0; k < 4; k++) pf.push_back(&d->f.mlx[k]);
  pf.push_back(&d->f.mly);
  gather_planes(d, pd, d->spare_d);
  gather_planes(d, pf, d->spare_f);
  launch_gather_i32(d->stream, d->spare_i, d->f.mlstep, d->gather, d->n);
  std::swap(d->f.mlstep, d->spare_i);
  launch_gather_i32(d->stream, d->spare_i, d->f.idlock, d->gather, d->n);
  std::swap(d->f.idlock, d->spare_i);
  if (d->cl.count) {  // contact candidates follow their particles
    hipLaunchKernelGGL(k_gather_cont_rows, dim3(gridn(d->n)), dim3(256), 0, d->stream,
                       d->cl_list_alt, d->cl.list, d->cl.count, d->gather, d->n, d->npad);
    std::swap(d->cl.list, d->cl_list_alt);
    launch_gather_i32(d->stream, d->spare_i, d->cl.count, d->gather, d->n);
    std::swap(d->cl.count, d->spare_i);
  }
  std::swap(d->perm, d->perm_tmp);
  if (pz)
    for (int c = 0; c < 3; c++)
      HIPC(hipMemcpyAsync(d->xb[c], d->f.x[c], sizeof(double) * d->n, hipMemcpyDeviceToDevice,
                          d->stream));
  d->pz_pending = false;
  d->local_since_reorder = 0;
  HIPC(hipGetLastError());
  d->nb_built = true;
  d->built_gen = d->snap_gen;
  for (int c = 0; c < 3; c++) d->frame.org[c] = d->grid.org[c];
  d->frame.cs = 1.0 / d->grid.inv_cs;
  d->frame.cs32 = (float)d->frame.cs;
  d->mirror_dirty = true;
  d->ns_builds++;
  return timer_end(d, WF_K_NSEARCH);
}

int wf_nsearch_stats(wf_domain* d, int64_t* total, int* over_cap) {
  if (!d) return fail(WF_E_ARG, "null domain");
  if (!d->nb_built) return fail(WF_E_STATE, "neighbour lists not built");
  std::vector<int> cnt(d->n);
  HIPC(hipMemcpyAsync(cnt.data(), d->nb_count, sizeof(int) * d->n, hipMemcpyDeviceToHost,
                      d->stream));
  HIPC(hipMemcpyAsync(d->ctr_host, d->ctr, sizeof(Counters), hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  int64_t t = 0;
  for (int c : cnt) t += c;
  if (total) *total = t;
  if (over_cap) *over_cap = d->ctr_host->nb_over_cap;
  return WF_OK;
}

int wf_get_neighbors(wf_domain* d, int* counts, int* offsets, int* list, int64_t list_cap) {
  if (!d || !counts || !offsets) return fail(WF_E_ARG, "null argument");
  if (!d->nb_built) return fail(WF_E_STATE, "neighbour lists not built");
  int n = d->n;
  // rows from the tile lists: slot k of sorted particle s is a position in
  // its tile's staged set tj[uoff[t] ..]
  const int nt = (n + kTileP - 1) / kTileP;
  std::vector<int> cnt(n), perm(n), uoff(nt + 1);
  HIPC(hipMemcpyAsync(cnt.data(), d->nb_count, sizeof(int) * n, hipMemcpyDeviceToHost, d->stream));
  HIPC(hipMemcpyAsync(perm.data(), d->perm, sizeof(int) * n, hipMemcpyDeviceToHost, d->stream));
  HIPC(hipMemcpyAsync(uoff.data(), d->tl.uoff, sizeof(int) * (nt + 1), hipMemcpyDeviceToHost,
                      d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  std::vector<int> tj(uoff[nt]);
  std::vector<uint2> slotq((size_t)d->tl.npad * kSlotQuads);
  HIPC(hipMemcpyAsync(tj.data(), d->tl.tj, sizeof(int) * tj.size(), hipMemcpyDeviceToHost,
                      d->stream));
  HIPC(hipMemcpyAsync(slotq.data(), d->tl.slotq, sizeof(uint2) * slotq.size(),
                      hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  std::vector<std::vector<int>> rows(n);
  for (int s = 0; s < n; s++) {
    auto& row = rows[perm[s]];
    const int base = uoff[s / kTileP];
    for (int k = 0; k < cnt[s]; k++) {
      const uint2 w = slotq[(size_t)(k >> 2) * d->tl.npad + s];
      const unsigned h = (k & 2) ? w.y : w.x;
      const int sl = (int)((k & 1) ? h >> 16 : h & 0xFFFFu);
      row.push_back(perm[tj[base + sl]]);
    }
    std::sort(row.begin(), row.end());
  }
  int64_t t = 0;
  for (int o = 0; o < n; o++) {
    offsets[o] = (int)t;
    counts[o] = (int)rows[o].size();
    for (int j : rows[o]) {
      if (list && t < list_cap) list[t] = j;
      t++;
    }
  }
  offsets[n] = (int)t;
  if (list && t > list_cap) return fail(WF_E_ARG, "list_cap too small");
  return WF_OK;
}


#define NEED_NB(d) \
  if (!(d)->nb_built) return fail(WF_E_STATE, "neighbour lists not built (call wf_nsearch_build)")

int wf_calc_dens_inc(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  NEED_NB(d);
  d->refresh_mirrors();
  if (d->mixed)
    launch_dens_mixed(d->stream, d->f, d->tl, d->n, d->solid, d->cfg.dt, false, d->frame.cs32,
                      d->uniform_h);
  else
    launch_dens_tiled(d->stream, d->f, d->tl, d->n, d->solid, d->cfg.dt, false, d->uniform_h);
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_update_density(wf_domain* d, double dt) {
  if (!d) return fail(WF_E_ARG, "null domain");
  launch_update_density(d->stream, d->f, d->params(), dt);
  d->mirror_dirty = true;
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_calc_rate_tensors(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  NEED_NB(d);
  d->refresh_mirrors();
  if (d->mixed)
    launch_rates_mixed(d->stream, d->f, d->tl, d->solid, d->cfg.quirks, d->frame.cs32,
                       d->uniform_h);
  else
    launch_rates_tiled(d->stream, d->f, d->tl, d->solid, d->cfg.quirks, d->uniform_h);
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_pressure(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  launch_pressure(d->stream, d->f, d->params());
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_stress_strain(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  const StepParams p = d->params();
  launch_stress(d->stream, d->f, p, d->bstats, false, true, kRecNone, d->mixed, d->ret_args());
  launch_fold_stats(d->stream, d->bstats, (p.solid + kBlock - 1) / kBlock, d->ctr);
  d->mirror_dirty = true;
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_calc_accel(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  NEED_NB(d);
  d->refresh_mirrors();
  if (d->mixed)
    launch_accel_mixed(d->stream, d->f, d->tl, d->solid, d->cfg.alpha, d->cfg.beta,
                       d->frame.cs32, d->uniform_h);
  else
    launch_accel_tiled(d->stream, d->f, d->tl, d->solid, d->cfg.alpha, d->cfg.beta,
                       d->uniform_h);
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_apply_axisymm_bc(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  launch_axisymm_bc(d->stream, d->f, d->n);
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_apply_bc_vel(wf_domain* d, int zone, double vx, double vy, double vz) {
  if (!d) return fail(WF_E_ARG, "null domain");
  launch_apply_bc(d->stream, d->f, d->n, zone, vx, vy, vz);
  d->mirror_dirty = true;
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_update_pos_fraser(wf_domain* d, double dt) {
  if (!d) return fail(WF_E_ARG, "null domain");
  launch_update_pos(d->stream, d->f, d->n, dt);
  d->mirror_dirty = true;
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_update_vel(wf_domain* d, double dt) {
  if (!d) return fail(WF_E_ARG, "null domain");
  launch_update_vel(d->stream, d->f, d->n, dt);
  d->mirror_dirty = true;
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_calc_temp_inc(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  NEED_NB(d);
  d->refresh_mirrors();
  if (d->mixed)
    launch_cond_mixed(d->stream, d->f, d->tl, d->n, d->cfg.T_inf, d->cfg.h_conv, d->cfg.dt,
                      false, false, d->frame.cs32, d->uniform_h);
  else
    launch_cond_tiled(d->stream, d->f, d->tl, d->n, d->cfg.T_inf, d->cfg.h_conv, d->cfg.dt, false,
                      false, d->uniform_h);
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_thermal_calcs(wf_domain* d, double dt) {
  if (!d) return fail(WF_E_ARG, "null domain");
  NEED_NB(d);
  d->refresh_mirrors();
  if (d->mixed)
    launch_cond_mixed(d->stream, d->f, d->tl, d->n, d->cfg.T_inf, d->cfg.h_conv, dt, false,
                      false, d->frame.cs32, d->uniform_h);
  else
    launch_cond_tiled(d->stream, d->f, d->tl, d->n, d->cfg.T_inf, d->cfg.h_conv, dt, false,
                      false, d->uniform_h);
  launch_leapfrog(d->stream, d->f, d->n, dt, d->isfirst_step);
  d->mirror_dirty = true;
  d->isfirst_step = false;
  HIPC(hipGetLastError());
  return WF_OK;
}
int wf_int_energy(wf_domain* d, double* sum) {
  if (!d || !sum) return fail(WF_E_ARG, "null argument");
  int nb = std::min(1024, std::max(1, gridn(d->n)));
  launch_sum(d->stream, d->f.eint, d->n, d->sum_part, nb, d->sum_out);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(sum, d->sum_out, sizeof(double), hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  return WF_OK;
}

// ---- rigid meshes and contact ---------------------------------------------
int wf_add_mesh(wf_domain* d, const double* nodes, int nn, const int* el, int ne, int zone,
                int* mesh_out) {
  if (!d || !nodes || !el) return fail(WF_E_ARG, "null argument");
  if (nn <= 0 || ne <= 0) return fail(WF_E_ARG, "empty mesh");
  if (d->n_meshes >= kMaxMeshes) return fail(WF_E_ARG, "too many meshes");
  for (int k = 0; k < 3 * ne; k++)
    if (el[k] < 0 || el[k] >= nn) return fail(WF_E_ARG, "element node index out of range");
  for (int k = 0; k < 3 * nn; k++)
    if (!std::isfinite(nodes[k])) return fail(WF_E_ARG, "non-finite node coordinate");
  const int nn0 = d->mesh.nn, ne0 = d->mesh.ne;
  if ((int64_t)nn0 + nn > INT32_MAX / 4 || (int64_t)ne0 + ne > INT32_MAX / 4)
    return fail(WF_E_ARG, "mesh too large");
  HIPC(hipStreamSynchronize(d->stream));  // mesh moves queued on the handle's stream
  const int NN = nn0 + nn, NE = ne0 + ne;
  // current device state of the meshes already added, plus the new one
  std::vector<double> hx[3], hv[3];
  for (int c = 0; c < 3; c++) {
    hx[c].resize(NN);
    hv[c].assign(NN, 0.0);
    if (nn0) {
      HIPC(hipMemcpy(hx[c].data(), d->mesh.x[c], sizeof(double) * nn0, hipMemcpyDeviceToHost));
      HIPC(hipMemcpy(hv[c].data(), d->mesh.v[c], sizeof(double) * nn0, hipMemcpyDeviceToHost));
    }
    for (int q = 0; q < nn; q++) hx[c][nn0 + q] = nodes[3 * q + c];
  }
  std::vector<int> nmesh(NN), hel(3 * (size_t)NE);
  if (nn0) {
    HIPC(hipMemcpy(nmesh.data(), d->mesh.node_mesh, sizeof(int) * nn0, hipMemcpyDeviceToHost));
    HIPC(hipMemcpy(hel.data(), d->mesh.el, sizeof(int) * 3 * ne0, hipMemcpyDeviceToHost));
  }
  const int mi = d->n_meshes;
  for (int q = 0; q < nn; q++) nmesh[nn0 + q] = mi;
  for (int k = 0; k < 3 * ne; k++) hel[3 * (size_t)ne0 + k] = el[k] + nn0;
  // fresh device arrays
  std::vector<void*> fresh;
  auto grab = [&](void** p, size_t bytes) {
    if (hipMalloc(p, bytes) != hipSuccess) return false;
    fresh.push_back(*p);
    return true;
  };
  MeshDev m{};
  bool ok = true;
  for (int c = 0; c < 3 && ok; c++)
    ok = grab((void**)&m.x[c], sizeof(double) * NN) && grab((void**)&m.v[c], sizeof(double) * NN) &&
         grab((void**)&m.c[c], sizeof(double) * NE) && grab((void**)&m.nrm[c], sizeof(double) * NE);
  int *rank = nullptr, *sorted = nullptr;
  ok = ok && grab((void**)&m.node_mesh, sizeof(int) * NN) &&
       grab((void**)&m.el, sizeof(int) * 3 * NE) && grab((void**)&m.pp, sizeof(double) * NE) &&
       grab((void**)&m.cell, sizeof(int) * NE) && grab((void**)&rank, sizeof(int) * NE) &&
       grab((void**)&sorted, sizeof(int) * NE);
  if (!ok) {
    for (void* q : fresh) hipFree(q);
    return fail(WF_E_NOMEM, "mesh allocation failed");
  }
  m.nn = NN;
  m.ne = NE;
  for (int c = 0; c < 3; c++) {
    HIPC(hipMemcpy(m.x[c], hx[c].data(), sizeof(double) * NN, hipMemcpyHostToDevice));
    HIPC(hipMemcpy(m.v[c], hv[c].data(), sizeof(double) * NN, hipMemcpyHostToDevice));
  }
  HIPC(hipMemcpy(m.node_mesh, nmesh.data(), sizeof(int) * NN, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(m.el, hel.data(), sizeof(int) * 3 * NE, hipMemcpyHostToDevice));
  // contact lists on first use
  if (!d->cl.count) {
    int rc = d->alloc_plane_i(&d->cl.count);
    if (!rc) rc = d->dev_alloc((void**)&d->cl.list, sizeof(int) * d->npad * kMaxContNb);
    if (!rc) rc = d->dev_alloc((void**)&d->cl_list_alt, sizeof(int) * d->npad * kMaxContNb);
    if (!rc) rc = d->dev_alloc((void**)&d->cnb_act, sizeof(int) * (d->npad + 1));
    if (rc) {
      // all or none (what was allocated is freed with the domain)
      d->cl.count = d->cl.list = d->cl_list_alt = d->cnb_act = nullptr;
      for (void* q : fresh) hipFree(q);
      return fail(rc, "contact list allocation failed");
    }
    d->cl.npad = d->npad;
  }
  // swap in; the old arrays (and the grid scratch) go
  HIPC(hipStreamSynchronize(d->stream));
  for (void* q : d->mesh_owned) hipFree(q);
  d->mesh_owned = fresh;
  d->tri_cnt = d->tri_start = d->tri_scan_part = nullptr;
  d->tri_cells_cap = 0;
  d->tri_rank = rank;
  d->tri_sorted = sorted;
  d->mesh = m;
  d->mesh_node_off[mi + 1] = NN;
  d->mesh_el_off[mi + 1] = NE;
  d->mesh_zone[mi] = zone;
  d->motions.m[mi] = MeshMotion{};
  d->n_meshes++;
  d->cont_built = false;
  launch_mesh_geom(d->stream, d->mesh);
  HIPC(hipGetLastError());
  HIPC(hipStreamSynchronize(d->stream));
  if (mesh_out) *mesh_out = mi;
  return WF_OK;
}

int wf_set_mesh_motion(wf_domain* d, int mesh, const double v[3], const double w[3],
                       const double center[3]) {
  if (!d) return fail(WF_E_ARG, "null domain");
  if (mesh < 0 || mesh >= d->n_meshes) return fail(WF_E_ARG, "no such mesh");
  MeshMotion& M = d->motions.m[mesh];
  for (int c = 0; c < 3; c++) {
    M.vel[c] = v ? v[c] : 0.0;
    M.omega[c] = w ? w[c] : 0.0;
    M.center[c] = center ? center[c] : 0.0;
  }
  return WF_OK;
}

int wf_mesh_info(wf_domain* d, int mesh, int* n_nodes, int* n_elems, int* zone_id) {
  if (!d) return fail(WF_E_ARG, "null domain");
  if (mesh == -1) {
    if (n_nodes) *n_nodes = d->mesh.nn;
    if (n_elems) *n_elems = d->mesh.ne;
    if (zone_id) *zone_id = d->n_meshes;
    return WF_OK;
  }
  if (mesh < 0 || mesh >= d->n_meshes) return fail(WF_E_ARG, "no such mesh");
  if (n_nodes) *n_nodes = d->mesh_node_off[mesh + 1] - d->mesh_node_off[mesh];
  if (n_elems) *n_elems = d->mesh_el_off[mesh + 1] - d->mesh_el_off[mesh];
  if (zone_id) *zone_id = d->mesh_zone[mesh];
  return WF_OK;
}

int wf_get_mesh(wf_domain* d, int mesh, double* nodes, double* node_v, double* centroids,
                double* normals, double* pplane) {
  if (!d) return fail(WF_E_ARG, "null domain");
  if (mesh < 0 || mesh >= d->n_meshes) return fail(WF_E_ARG, "no such mesh");
  HIPC(hipStreamSynchronize(d->stream));
  const int n0 = d->mesh_node_off[mesh], nn = d->mesh_node_off[mesh + 1] - n0;
  const int e0 = d->mesh_el_off[mesh], ne = d->mesh_el_off[mesh + 1] - e0;
  std::vector<double> buf(std::max(nn, ne));
  auto get3 = [&](double* out, double* const* planes, int off, int cnt) -> int {
    if (!out) return WF_OK;
    for (int c = 0; c < 3; c++) {
      HIPC(hipMemcpy(buf.data(), planes[c] + off, sizeof(double) * cnt, hipMemcpyDeviceToHost));
      for (int k = 0; k < cnt; k++) out[3 * k + c] = buf[k];
    }
    return WF_OK;
  };
  int rc = get3(nodes, d->mesh.x, n0, nn);
  if (!rc) rc = get3(node_v, d->mesh.v, n0, nn);
  if (!rc) rc = get3(centroids, d->mesh.c, e0, ne);
  if (!rc) rc = get3(normals, d->mesh.nrm, e0, ne);
  if (rc) return rc;
  if (pplane) HIPC(hipMemcpy(pplane, d->mesh.pp + e0, sizeof(double) * ne, hipMemcpyDeviceToHost));
  return WF_OK;
}

int wf_calc_surface(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  NEED_NB(d);
  int rc = d->surface_init();
  if (rc) return rc;
  d->launch_surface();
  HIPC(hipGetLastError());
  return WF_OK;
}

int wf_calc_contact_nb(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  if (d->n_meshes == 0) return fail(WF_E_STATE, "no rigid mesh (call wf_add_mesh)");
  return d->contact_nb();
}

int wf_calc_contact_forces(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  if (!d->cont_built) return fail(WF_E_STATE, "contact neighbours not built");
  launch_contact_force(d->stream, d->f, d->solid, d->mesh, d->cl, d->contact_params());
  HIPC(hipGetLastError());
  return WF_OK;
}

// the rotation centre of each mesh travels with its velocity (a tool spins
// about its own moving axis)
static void advance_mesh_centres(wf_domain* d, double dt) {
  for (int k = 0; k < d->n_meshes; k++)
    for (int c = 0; c < 3; c++) d->motions.m[k].center[c] += d->motions.m[k].vel[c] * dt;
}

int wf_mesh_update(wf_domain* d, double dt) {
  if (!d) return fail(WF_E_ARG, "null domain");
  launch_mesh_move(d->stream, d->mesh, d->motions, dt);
  advance_mesh_centres(d, dt);
  HIPC(hipGetLastError());
  return WF_OK;
}

int wf_get_contact_neighbors(wf_domain* d, int* counts, int* list) {
  if (!d || !counts) return fail(WF_E_ARG, "null argument");
  const int n = d->n;
  if (!d->cl.count || !d->cont_built) {
    std::memset(counts, 0, sizeof(int) * n);
    if (list)
      for (int64_t k = 0; k < (int64_t)n * kMaxContNb; k++) list[k] = -1;
    return WF_OK;
  }
  std::vector<int> cnt(n), perm(n), lst((size_t)d->npad * kMaxContNb);
  HIPC(hipMemcpyAsync(cnt.data(), d->cl.count, sizeof(int) * n, hipMemcpyDeviceToHost, d->stream));
  HIPC(hipMemcpyAsync(perm.data(), d->perm, sizeof(int) * n, hipMemcpyDeviceToHost, d->stream));
  HIPC(hipMemcpyAsync(lst.data(), d->cl.list, sizeof(int) * lst.size(), hipMemcpyDeviceToHost,
                      d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  for (int s = 0; s < n; s++) {
    const int o = perm[s];
    counts[o] = cnt[s];
    if (list)
      for (int k = 0; k < kMaxContNb; k++)
        list[(int64_t)o * kMaxContNb + k] = k < cnt[s] ? lst[(size_t)k * d->npad + s] : -1;
  }
  return WF_OK;
}

int wf_contact_force_sum(wf_domain* d, double* sum) {
  if (!d || !sum) return fail(WF_E_ARG, "null argument");
  launch_contact_norm(d->stream, d->f, d->n, d->cfg.id_free_surf, d->spare_d[0]);
  const int nb = std::min(1024, std::max(1, gridn(d->n)));
  launch_sum(d->stream, d->spare_d[0], d->n, d->sum_part, nb, d->sum_out);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(sum, d->sum_out, sizeof(double), hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  return WF_OK;
}

// Whether the next step starts with a neighbour rebuild that actually runs
// (SolverFraser.cu:166; under the stale-snapshot quirk a rebuild from an
// unchanged snapshot is skipped, see wf_nsearch_build).
// Before yield each step's max eps_p is copied back behind its stress pass;
// the host reads it (SolverFraser.cu:260-271) only when it must, by then
// with the step's later passes queued, so the device does not wait for it.
static int resolve_yield(wf_domain* d) {
  if (!d->yield_pending) return WF_OK;
  HIPC(hipEventSynchronize(d->ev_yield));
  d->yield_pending = false;
  double mxp;
  const unsigned long long b = d->ctr_host->max_pl_strain_bits;
  std::memcpy(&mxp, &b, sizeof(mxp));
  if (mxp > d->cfg.ns_min_plastic) d->is_yielding = true;
  return WF_OK;
}

// The plastic-zone trigger (SolverFraser.cu:166 under the skin policy): a
// rebuild starts the next step when a hot particle moved more than skin/2
// since the last build.  The host reads the step's largest squared
// displacement once the conduction pass behind it is queued.
static int resolve_pz(wf_domain* d) {
  if (!d->pz_wait) return WF_OK;
  HIPC(hipEventSynchronize(d->ev_pz));
  d->pz_wait = false;
  double d2;
  std::memcpy(&d2, d->pz_d2_host, sizeof(d2));
  d->pz_pending = 4.0 * d2 > d->cfg.ns_skin * d->cfg.ns_skin;
  return WF_OK;
}

// whether the next step starts with a rebuild that changes the lists; the
// yield read-back is waited for only when the answer depends on it
static int rebuild_next_step(wf_domain* d, bool* out) {
  *out = false;
  if (d->plastic_zone()) {
    int rc = resolve_pz(d);
    if (rc) return rc;
    *out = d->is_yielding && d->pz_pending;
    return WF_OK;
  }
  const int ts_next = d->ts_i + 1 > d->cfg.ns_interval - 1 ? 0 : d->ts_i + 1;
  if (ts_next != 0) return WF_OK;
  const bool stale = (d->cfg.quirks & WF_Q_STALE_NS_POS) != 0;
  if (stale && d->nb_built && d->built_gen == d->snap_gen) return WF_OK;  // a no-op rebuild
  int rc = resolve_yield(d);
  if (rc) return rc;
  *out = d->is_yielding;
  return WF_OK;
}

// Domain_d::MechFraserSolve loop body, SolverFraser.cu:164-399.  Inside one
// call, a step's conduction pass and the next step's density pass share one
// traversal of the neighbour lists when no rebuild runs between them (both
// read the positions after integration; see CondDens in wf_pair.hip).
int wf_step(wf_domain* d, int n_steps) {
  if (!d) return fail(WF_E_ARG, "null domain");
  if (n_steps < 0) return fail(WF_E_ARG, "negative step count");
  int rc;
  if (!d->nb_built) {  // initial search, SolverFraser.cu:109
    rc = wf_nsearch_build(d);
    if (rc) return rc;
  }
  bool dens_done = false;  // this step's density pass ran with the last conduction pass
  for (int s = 0; s < n_steps; s++) {
    StepParams p = d->params();
    if (d->cfg.mechanics) {
      if ((rc = resolve_pz(d))) return rc;
      if (d->plastic_zone() ? (d->is_yielding && d->pz_pending)
                            : (d->ts_i == 0 && d->is_yielding)) {
        rc = d->relist_local();  // WF_E_STATE: a full rebuild is due
        if (rc == WF_E_STATE) rc = wf_nsearch_build(d);
        if (rc) return rc;
      }
      d->refresh_mirrors();
      const bool mx = d->mixed;
      const float cs32 = d->frame.cs32;
      const bool uh = d->uniform_h;
      if (!dens_done) {
        timer_begin(d, true);
        if (mx) {
          launch_dens_mixed(d->stream, d->f, d->tl, d->n, d->solid, p.dt, true, cs32, uh);
          std::swap(d->f.v4, d->f.v4_alt);
        } else {
          launch_dens_tiled(d->stream, d->f, d->tl, d->n, d->solid, p.dt, true, uh);
          std::swap(d->f.kv, d->f.kv_alt);
        }
        std::swap(d->f.rho, d->f.rho_alt);
        if ((rc = timer_end(d, WF_K_DENS))) return rc;
      }
      dens_done = false;
      timer_begin(d, true);
      if (mx) launch_rates_mixed(d->stream, d->f, d->tl, d->solid, p.quirks, cs32, uh);
      else launch_rates_tiled(d->stream, d->f, d->tl, d->solid, p.quirks, uh);
      if ((rc = timer_end(d, WF_K_RATES))) return rc;
      timer_begin(d);
      const int rec = mx ? kRec32 : kRec64;
      launch_stress(d->stream, d->f, p, d->bstats, true, true, rec, mx, d->ret_args());
      if ((rc = timer_end(d, WF_K_STRESS))) return rc;
      // the acceleration pass's first workgroup folds the stress statistics
      const StatsFold fold{d->bstats, (p.solid + kBlock - 1) / kBlock, d->ctr};
      // Without contact forces or the axisymmetric clamp between them, the
      // mixed acceleration pass integrates each particle as soon as its sum
      // is done (the reference's BC + UpdatePosFraser + UpdateVel + BC,
      // SolverFraser.cu:273-281): the new FP32 mirrors go to spare planes,
      // swapped in after the pass.  (The FP64 pass, at lower occupancy, was
      // slower fused: 0.396 against 0.332 + 0.032 ms on a 759k-particle
      // cylinder.)
      const bool fuse_int = mx && d->cfg.fuse_integrate && !d->cfg.axisymm &&
                            !d->cfg.contact && d->solid == d->n;
      timer_begin(d, true);
      if (fuse_int) {
        // the FP32 pass reads no FP64 position, so x is updated in place
        IntegrateOut out{};
        for (int c = 0; c < 3; c++) out.x[c] = d->f.x[c];
        out.p4 = d->p4_spare;
        out.v4 = d->f.v4_alt;
        launch_accel_integrate_mixed(d->stream, d->f, d->tl, d->solid, p.alpha, p.beta, cs32, uh,
                                     d->cfg.dt, d->bcs, d->frame, out, fold);
        std::swap(d->f.p4, d->p4_spare);
        std::swap(d->f.v4, d->f.v4_alt);
        if ((rc = timer_end(d, WF_K_ACCEL_INTEGRATE))) return rc;
      } else {
        if (mx)
          launch_accel_mixed(d->stream, d->f, d->tl, d->solid, p.alpha, p.beta, cs32, uh, fold);
        else launch_accel_tiled(d->stream, d->f, d->tl, d->solid, p.alpha, p.beta, uh, fold);
        if (d->cfg.axisymm) launch_axisymm_bc(d->stream, d->f, d->n);  // SolverFraser.cu:229-231
        if ((rc = timer_end(d, WF_K_ACCEL))) return rc;
      }
      if (d->cfg.contact) {  // SolverFraser.cu:233-257
        timer_begin(d);
        if ((rc = d->surface_init())) return rc;
        d->launch_surface();
        if (d->n_meshes > 0) {
          if (d->ts_i == 0 && (rc = d->contact_nb())) return rc;
          launch_contact_force(d->stream, d->f, d->solid, d->mesh, d->cl, d->contact_params());
        }
        if ((rc = timer_end(d, WF_K_CONTACT))) return rc;
      }
      if (!d->is_yielding) {  // SolverFraser.cu:260-271, on the device-side max
        HIPC(hipMemcpyAsync(d->ctr_host, d->ctr, sizeof(Counters), hipMemcpyDeviceToHost,
                            d->stream));
        HIPC(hipEventRecord(d->ev_yield, d->stream));
        d->yield_pending = true;
      }
      if (!fuse_int) {
        timer_begin(d);
        launch_integrate(d->stream, d->f, d->n, d->cfg.dt, d->bcs, rec, d->frame);
        if ((rc = timer_end(d, WF_K_INTEGRATE))) return rc;
      }
      if (d->plastic_zone() && (rc = resolve_yield(d))) return rc;
      if (d->plastic_zone() && d->is_yielding) {  // the rebuild trigger, read back each step
        hipLaunchKernelGGL(k_zero_bytes, dim3(1), dim3(256), 0, d->stream,
                           (unsigned char*)d->pz_d2, (int64_t)sizeof(unsigned long long));
        launch_hot_disp(d->stream, d->f, d->xb, d->n, d->cfg.ep_hot, d->pz_d2);
        HIPC(hipMemcpyAsync(d->pz_d2_host, d->pz_d2, sizeof(unsigned long long),
                            hipMemcpyDeviceToHost, d->stream));
        HIPC(hipEventRecord(d->ev_pz, d->stream));
        d->pz_wait = true;
      }
    }
    if (d->cfg.thermal) {
      d->refresh_mirrors();
      // With the trigger's answer still on its way, the fused pass is queued
      // anyway and takes the answer from device memory: when a rebuild is
      // due its density half writes nothing, and the host, reading the same
      // value behind it, runs the density pass after the rebuild as usual.
      const bool pz_open = d->pz_wait;
      bool rebuild_next = false;
      if (d->cfg.mechanics && s + 1 < n_steps && !pz_open &&
          (rc = rebuild_next_step(d, &rebuild_next)))
        return rc;
      const bool fuse = d->cfg.mechanics && s + 1 < n_steps && !rebuild_next;
      const unsigned long long* pz_d2 = pz_open ? d->pz_d2 : nullptr;
      const double skin2 = d->cfg.ns_skin * d->cfg.ns_skin;
      timer_begin(d, true);
      if (fuse && d->mixed) {
        launch_cond_dens_mixed(d->stream, d->f, d->tl, d->n, d->solid, p.T_inf, p.h_conv, p.dt,
                               true, d->isfirst_step, d->frame.cs32, d->uniform_h, pz_d2, skin2);
      } else if (fuse) {
        launch_cond_dens_tiled(d->stream, d->f, d->tl, d->n, d->solid, p.T_inf, p.h_conv,
                               p.dt, true, d->isfirst_step, d->uniform_h, pz_d2, skin2);
      } else if (d->mixed) {
        launch_cond_mixed(d->stream, d->f, d->tl, d->n, p.T_inf, p.h_conv, p.dt, true,
                          d->isfirst_step, d->frame.cs32, d->uniform_h);
      } else {
        launch_cond_tiled(d->stream, d->f, d->tl, d->n, p.T_inf, p.h_conv, p.dt, true,
                          d->isfirst_step, d->uniform_h);
      }
      if (d->mixed) std::swap(d->f.tk4, d->f.tk4_alt);
      else std::swap(d->f.tk64, d->f.tk64_alt);
      std::swap(d->f.T, d->f.T_alt);
      if ((rc = timer_end(d, fuse ? WF_K_COND_DENS : WF_K_THERMAL))) return rc;
      if ((rc = resolve_pz(d))) return rc;  // the conduction pass is queued
      // whether the density half wrote the next step's density
      if (fuse && !(pz_open && d->is_yielding && d->pz_pending)) {
        if (d->mixed) std::swap(d->f.v4, d->f.v4_alt);
        else std::swap(d->f.kv, d->f.kv_alt);
        std::swap(d->f.rho, d->f.rho_alt);
        dens_done = true;
      }
    }
    if (d->cfg.mechanics && d->cfg.contact && d->n_meshes > 0) {  // SolverFraser.cu:290-318
      timer_begin(d);
      launch_mesh_move(d->stream, d->mesh, d->motions, p.dt);
      advance_mesh_centres(d, p.dt);
      if ((rc = timer_end(d, WF_K_CONTACT))) return rc;
    }
    if ((rc = resolve_pz(d))) return rc;
    if ((rc = resolve_yield(d))) return rc;  // the step's later passes are queued
    if (d->cfg.surrogate && d->jc_newton && d->cfg.mlp_train_interval > 0) {
      // rounds are due every mlp_interval steps, taken while yielding; the
      // interval stays mlp_train_interval, or with mlp_interval_max > 0 it
      // doubles (up to that cap) after a round that finds the network still
      // fitting the new samples: its starting loss within mlp_refit_ratio of
      // the previous round's final loss.  A round that does not resets it.
      if (d->mlp_interval == 0) {
        d->mlp_interval = d->cfg.mlp_train_interval;
        d->mlp_next_train = d->mlp_interval;
      }
      if (d->step + 1 >= d->mlp_next_train) {
        if (d->is_yielding) {
          const int done = d->mlp_trainings;
          const double prev = d->mlp_last_loss;
          timer_begin(d);
          rc = d->mlp_train_round(d->cfg.mlp_train_iters, (float)d->cfg.mlp_lr);
          if (rc) return rc;
          if ((rc = timer_end(d, WF_K_MLP))) return rc;
          if (d->cfg.mlp_interval_max > 0 && d->mlp_trainings > done) {
            const bool fits = done > 0 && d->mlp_start_loss <= d->cfg.mlp_refit_ratio * prev;
            d->mlp_interval = fits ? std::min(2 * d->mlp_interval, std::max(
                                                  d->cfg.mlp_interval_max, d->cfg.mlp_train_interval))
                                   : d->cfg.mlp_train_interval;
          }
        }
        d->mlp_next_train = d->step + 1 + d->mlp_interval;
      }
    }
    d->isfirst_step = false;
    d->time += d->cfg.dt;
    d->step++;
    d->ts_i++;
    if (d->ts_i > d->cfg.ns_interval - 1) d->ts_i = 0;
  }
  HIPC(hipGetLastError());
  return WF_OK;
}

int wf_get_stats(wf_domain* d, wf_stats* out) {
  if (!d || !out) return fail(WF_E_ARG, "null argument");
  HIPC(hipMemcpyAsync(d->ctr_host, d->ctr, sizeof(Counters), hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  std::memset(out, 0, sizeof(*out));
  out->step = d->step;
  out->time = d->time;
  out->is_yielding = d->is_yielding;
  out->ns_builds = d->ns_builds;
  out->newton_iters = (int64_t)d->ctr_host->newton_iters;
  out->newton_calls = (int64_t)d->ctr_host->newton_calls;
  out->newton_iters_wave = (int64_t)d->ctr_host->newton_iters_wave;
  out->mlp_trainings = d->mlp_trainings;
  out->nb_over_cap = d->ctr_host->nb_over_cap;
  out->cont_over_cap = d->ctr_host->cont_over_cap;
  out->mlp_loss = d->mlp_last_loss;
  out->mlp_samples = d->mlp_last_samples;
  out->mlp_interval = d->mlp_interval > 0 ? d->mlp_interval : d->cfg.mlp_train_interval;
  out->mlp_start_loss = d->mlp_start_loss;
  out->ns_local_builds = d->ns_local_builds;
  out->ns_local_tiles = d->ns_local_tiles;
  return WF_OK;
}

int wf_mlp_get_params(wf_domain* d, float* out, int n) {
  if (!d || !out || n != kMlpParams) return fail(WF_E_ARG, "need WF_MLP_PARAMS floats");
  HIPC(hipMemcpyAsync(out, d->mlp_w, sizeof(float) * n, hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  return WF_OK;
}

int wf_mlp_set_params(wf_domain* d, const float* in, int n) {
  if (!d || !in || n != kMlpParams) return fail(WF_E_ARG, "need WF_MLP_PARAMS floats");
  HIPC(hipMemcpyAsync(d->mlp_w, in, sizeof(float) * n, hipMemcpyHostToDevice, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  d->mlp_trained = true;
  d->mlp_pack_stale = true;
  return WF_OK;
}

int wf_mlp_reset_adam(wf_domain* d) {
  if (!d) return fail(WF_E_ARG, "null domain");
  std::vector<float> z(kMlpParams, 0.f);
  HIPC(hipMemcpy(d->mlp_m, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  HIPC(hipMemcpy(d->mlp_v, z.data(), sizeof(float) * kMlpParams, hipMemcpyHostToDevice));
  d->adam_t = 0;
  return WF_OK;
}

static int upload_samples(wf_domain* d, const float* X, const float* Y, int n) {
  if (n <= 0 || n > d->mlp_cap) return fail(WF_E_ARG, "sample count out of range");
  HIPC(hipMemcpyAsync(d->mlp_X, X, sizeof(float) * 4 * n, hipMemcpyHostToDevice, d->stream));
  if (Y) HIPC(hipMemcpyAsync(d->mlp_Y, Y, sizeof(float) * n, hipMemcpyHostToDevice, d->stream));
  return WF_OK;
}

int wf_mlp_infer(wf_domain* d, const float* X, int n, float* y) {
  if (!d || !X || !y) return fail(WF_E_ARG, "null argument");
  int rc = upload_samples(d, X, nullptr, n);
  if (rc) return rc;
  d->pack_mlp();
  launch_mlp_infer(d->stream, d->mlp_pack, d->mlp_X, n, d->mlp_Y);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(y, d->mlp_Y, sizeof(float) * n, hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  return WF_OK;
}

int wf_mlp_grad(wf_domain* d, const float* X, const float* Y, int n, float* grad, float* loss) {
  if (!d || !X || !Y || !grad) return fail(WF_E_ARG, "null argument");
  int rc = upload_samples(d, X, Y, n);
  if (rc) return rc;
  const int nblocks = mlp_grad_blocks(n);
  launch_mlp_grad(d->stream, d->mlp_w, d->mlp_X, d->mlp_Y, n, 1, d->mlp_part, d->mlp_loss_part,
                  nblocks, d->mlp_grad, d->mlp_loss);
  HIPC(hipGetLastError());
  HIPC(hipMemcpyAsync(grad, d->mlp_grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost,
                      d->stream));
  HIPC(hipMemcpyAsync(d->host_f, d->mlp_loss, sizeof(float), hipMemcpyDeviceToHost, d->stream));
  HIPC(hipStreamSynchronize(d->stream));
  if (loss) *loss = d->host_f[0];
  return WF_OK;
}

int wf_mlp_fit(wf_domain* d, const float* X, const float* Y, int n, int iters, float lr,
               float* loss) {
  if (!d || !X || !Y || iters < 0) return fail(WF_E_ARG, "bad argument");
  int rc = upload_samples(d, X, Y, n);
  if (rc) return rc;
  rc = d->mlp_fit(d->mlp_X, d->mlp_Y, n, 1, iters, lr);
  if (rc) return rc;
  d->mlp_trained = true;
  if (loss) *loss = (float)d->mlp_last_loss;
  return WF_OK;
}

int wf_mlp_train(wf_domain* d, int iters, float lr, int64_t* samples, float* loss) {
  if (!d || iters < 0) return fail(WF_E_ARG, "bad argument");
  int rc = d->mlp_train_round(iters, lr);
  if (rc) return rc;
  if (samples) *samples = d->mlp_last_samples;
  if (loss) *loss = (float)d->mlp_last_loss;
  return WF_OK;
}

int wf_set_timing(wf_domain* d, int on) {
  if (!d) return fail(WF_E_ARG, "null domain");
  HIPC(hipStreamSynchronize(d->stream));
  d->tev_n = 0;
  d->timing = on != 0;
  for (int k = 0; k < WF_K_COUNT_; k++) {
    d->kt_ms[k] = 0.0;
    d->kt_n[k] = 0;
  }
  return WF_OK;
}

int wf_kernel_time(wf_domain* d, int k, double* total_ms, int64_t* launches) {
  if (!d || k < 0 || k >= WF_K_COUNT_) return fail(WF_E_ARG, "bad argument");
  if (int rc = timer_resolve(d)) return rc;
  if (total_ms) *total_ms = d->kt_ms[k];
  if (launches) *launches = d->kt_n[k];
  return WF_OK;
}

int wf_time_pass(wf_domain* d, int k, int reps, double* ms_per_launch) {
  if (!d || !ms_per_launch || reps < 1 || (k != WF_K_RATES && k != WF_K_ACCEL))
    return fail(WF_E_ARG, "wf_time_pass: WF_K_RATES or WF_K_ACCEL, reps >= 1");
  NEED_NB(d);
  d->refresh_mirrors();
  const bool mx = d->mixed, uh = d->uniform_h;
  const float cs32 = d->frame.cs32;
  auto launch = [&]() {
    if (k == WF_K_RATES) {
      if (mx) launch_rates_mixed(d->stream, d->f, d->tl, d->solid, d->cfg.quirks, cs32, uh);
      else launch_rates_tiled(d->stream, d->f, d->tl, d->solid, d->cfg.quirks, uh);
    } else {
      if (mx) launch_accel_mixed(d->stream, d->f, d->tl, d->solid, d->cfg.alpha, d->cfg.beta,
                                 cs32, uh);
      else launch_accel_tiled(d->stream, d->f, d->tl, d->solid, d->cfg.alpha, d->cfg.beta, uh);
    }
  };
  launch();  // warm
  HIPC(hipEventRecord(d->ev0, d->stream));
  for (int r = 0; r < reps; r++) launch();
  HIPC(hipEventRecord(d->ev1, d->stream));
  HIPC(hipEventSynchronize(d->ev1));
  float ms = 0.f;
  HIPC(hipEventElapsedTime(&ms, d->ev0, d->ev1));
  HIPC(hipGetLastError());
  *ms_per_launch = (double)ms / reps;
  return WF_OK;
}

}  // extern "C"

int wf::set_last_error(int code, const char* msg) { return fail(code, msg ? msg : ""); }
