// wf_internal.hpp — device state layout and kernel launchers of libwfhip.
//
// Particle data lives in HBM as structure-of-arrays in Morton-sorted order
// (see wf_nsearch.hip): one contiguous block per field, components of a
// vector/tensor field stored as separate planes of npad elements, so a
// wavefront's 64 consecutive particles read 64 consecutive elements.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/wf_abi.h"
#include "wf_tiles.hpp"

namespace wf {

constexpr int kBlock = 256;       // threads per block of the per-particle kernels
constexpr int kMaxNb = WF_MAX_NEIGHBORS;          // the reference's cap
constexpr int kListCap = WF_MAX_NEIGHBORS_SKIN;  // list storage (either policy)

// Plane pointers of every particle field (device addresses, sorted order).
// Passed by value to kernels.
struct Fields {
  double *x[3], *v[3], *a[3], *u[3], *va[3], *xs[3];
  double *rho, *rho_alt, *rho0, *drho, *m, *h, *p, *Cs, *G;
  double *sigma_y, *pl_strain, *sigma_eq, *esr, *eint;
  double *T, *T_alt, *Ta, *Tb, *dTdt, *k, *cp, *qpl, *qfric;
  // FP64 stress state, strain and rates (WF_FP64; in WF_MIXED only strain)
  double *sigma[6], *ss[6], *strain[6], *strrate[6], *rotrate[6];
  // WF_MIXED keeps sigma, the deviatoric stress and the rates in FP32 (the
  // stress pass computes in FP64 from them); the FP64 planes above are not
  // allocated then
  float *sigma32[6], *ss32[6], *strrate32[6], *rotrate32[6];
  int *ID, *BC_T;
  // FP32 mirrors read by the mixed-precision pair loops (WF_MIXED): p4 =
  // {position relative to the origin of an integer grid cell, the cell packed
  // as cx | cy << 11 | cz << 22 in the bits of w} (one 16-byte load stages a
  // neighbour) and h; the neighbour fields a
  // pair gathers, packed per particle: v4 = {v, m/rho} (double-buffered like
  // rho), tk4 = {T, k, m/rho, 0} (double-buffered like T; m/rho refreshed by
  // the density update), and the acceleration loop's
  // 32-byte record per particle, written by the stress pass: arec[2i] =
  // sigma/rho^2 {xx, yy, zz, xy}, arec[2i+1] = {sigma/rho^2 yz, xz, rho, Cs}
  // (one aligned half of a 64-byte line; the neighbour's v4 is staged in LDS
  // with its position).
  float4* p4;
  float* h32;
  float4 *v4, *v4_alt, *arec;
  float4 *tk4, *tk4_alt;
  // the same packed neighbour records in FP64 for the FP64 pair loops
  // (WF_FP64): kv = {v, m/rho} (double-buffered), arec64[3i..3i+2] = {v, m},
  // sigma/rho^2 {xx, yy, zz, xy}, {sigma/rho^2 yz, xz, rho, Cs}, and
  // tk64 = {T, k, m/rho, 0} (double-buffered)
  double4 *kv, *kv_alt, *arec64, *tk64, *tk64_alt;
  // return-map surrogate training samples, one slot per particle: features
  // (s_trial/A, 10 eps_p, log10(max(rate,1e-3))/3, theta), target dg/dg_max,
  // step of the sample (-1: none)
  float *mlx[4], *mly;
  int* mlstep;
  // contact: surface normal, contact force, not_write_surf_ID (ID != 0 at
  // the first surface pass)
  double *nrm[3], *cf[3];
  int* idlock;
};

// the most particles a domain holds: the pair loops address the per-particle
// records (at most 96 B each) with 32-bit byte offsets
constexpr int kMaxParticles = 40000000;

// element c (0..2) of particle i's packed neighbour record (records are
// contiguous per particle: one record spans one or two cache lines, where
// three separate arrays would touch three)
__device__ __forceinline__ size_t rec_at(const Fields&, int i, int c) {
  return 3 * (size_t)i + c;
}
// element c (0..1) of particle i's FP32 acceleration record
__device__ __forceinline__ size_t rec32_at(int i, int c) { return 2 * (size_t)i + c; }

// which packed neighbour records a producer kernel also writes
enum { kRecNone = 0, kRec32 = 1, kRec64 = 2 };

// ---- rigid triangle meshes (wf_contact.hip) --------------------------------
struct Counters;
constexpr int kMaxContNb = WF_MAX_CONT_NB;
constexpr int kMaxMeshes = WF_MAX_MESHES;

// all meshes concatenated: nodes (position, velocity, owning mesh),
// connectivity with global node indices, per-element centroid, unit normal
// and plane constant n . x_node0
struct MeshDev {
  double *x[3], *v[3];
  int* node_mesh;
  int* el;
  double *c[3], *nrm[3], *pp;
  int* cell;  // centroid grid cell of each element (contact candidate build)
  int nn, ne;
};
struct MeshMotion {
  double vel[3], omega[3], center[3];
};
struct MeshMotions {
  MeshMotion m[kMaxMeshes];
};
// uniform grid over triangle centroids, cell edge >= 2 max h
struct TriGrid {
  double org[3];
  double inv_cs;
  int dim[3];
};
// contact candidates: entry k of particle i at list[k*npad + i]
struct ContactLists {
  int* count;
  int* list;
  int64_t npad;
};
struct ContactParams {
  double dt, kappa_E_dx, mu_s, mu_d;
  double fric_heat;  // fraction of the sliding-friction power into the particle
  int law;
};

void launch_mesh_geom(hipStream_t s, const MeshDev& m);
void launch_mesh_move(hipStream_t s, const MeshDev& m, const MeshMotions& mo, double dt);
// centroid bounding box (6 doubles: min xyz, max xyz) and max of h over n
void launch_tri_bbox(hipStream_t s, const MeshDev& m, double* out);
// max of v[0..n) into *out; part holds kMaxParts doubles of block partials
constexpr int kMaxParts = 1024;
void launch_max_f64(hipStream_t s, const double* v, int n, double* out, double* part);
void launch_tri_bin(hipStream_t s, const MeshDev& m, const TriGrid& g, int* cnt, int* rank,
                    int* start, int* sorted, int* scan_part);
void launch_surface_lock(hipStream_t s, const Fields& f, int n);
// r2cut > 0: only pairs with |x_ij|^2 < r2cut count (lists built with a skin);
// um: every particle's mass is m_value (no neighbour-mass gather)
void launch_surface_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                          double scale_num, double totmass, int id_free_surf, double r2cut,
                          bool um, double m_value);
// mixed precision: FP32 tile-frame positions (cs32 the FP32 cell size)
void launch_surface_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                          double scale_num, double totmass, int id_free_surf, double r2cut,
                          bool um, double m_value, float cs32);
void launch_contact_nb(hipStream_t s, const Fields& f, int solid, const MeshDev& m,
                       const TriGrid& g, const int* start, const int* sorted,
                       const ContactLists& cl, int id_free_surf, int* act, Counters* ctr);
void launch_contact_force(hipStream_t s, const Fields& f, int solid, const MeshDev& m,
                          const ContactLists& cl, const ContactParams& cp);
// |cf_i| where ID_i == id, else 0
void launch_contact_norm(hipStream_t s, const Fields& f, int n, int id, double* out);

// What the stress kernel needs for the Johnson-Cook Newton return map: the
// surrogate's weights, whether to use them, and the step number stamped on
// the logged training samples.
struct RetArgs {
  const void* mlp_pack;  // the weights as per-lane MFMA operands (launch_mlp_pack)
  int use_mlp;
  long long step;
};

// per-block statistics of the stress kernel, folded into Counters
// once per step: max eps_p bits, Newton iterations, Newton calls, sum over
// wavefronts of the wave's largest iteration count
struct BlockStats {
  unsigned long long ep_bits, iters, calls, wave_iters;
};
// the fold of a stress pass's nb BlockStats into ctr, done by the first
// workgroup of the acceleration pass that follows it (nb = 0: none)
struct StatsFold {
  const BlockStats* bs;
  int nb;
  Counters* ctr;
};

constexpr int kMlpParams = 1249;  // 4x32 + 32 + 32x32 + 32 + 32 + 1
constexpr int kMlpMaxBlocks = 256;  // gradient blocks (one 64-sample chunk each)
constexpr int kMlpOffW1 = 0, kMlpOffB1 = 128, kMlpOffW2 = 160, kMlpOffB2 = 1184,
              kMlpOffW3 = 1216, kMlpOffB3 = 1248;

// frame of the FP32 cell-relative coordinates
struct Frame32 {
  double org[3];
  double cs;      // cell edge (FP64) used to form the relative coordinates
  float cs32;     // the same edge in FP32, used inside the pair loops
};

// Neighbour lists: interleaved, entry k of particle i at list[k*npad + i]
// (a wavefront reads 64 consecutive ints per k), count[i] <= the build's cap
// (kMaxNb, or kListCap with a list skin).
struct NbLists {
  const int* count;
  const int* list;
  int64_t npad;
};

struct MatParams {
  int model;
  double E, Et, Ep, sy0, K_h, m_h, eps0_h, eps1_h;
  double A, B, n, C, eps0, m, Tm, Tt;
};

struct StepParams {
  int n, solid;
  double dt;
  double alpha, beta;
  int quirks;
  int return_map;
  double taylor_quinney;
  double T_inf, h_conv;
  double newton_tol;
  int newton_max_iter;
  MatParams mat;
};

struct BcTable {
  int count;
  int zone[WF_MAX_BCS];
  double val[WF_MAX_BCS][3];
};

// Where the integration writes what other threads of the same kernel may
// still read: the new positions (x) and FP32 mirrors (p4, v4).  The
// integration kernel writes them in place; the acceleration pass that
// integrates its own particles writes the spare planes, swapped in after.
struct IntegrateOut {
  double* x[3];
  float4 *p4, *v4;
};

// device-side counters written by the kernels (one allocation)
struct Counters {
  unsigned long long max_pl_strain_bits;  // max eps_p as ordered bits (eps_p >= 0)
  unsigned long long newton_iters;
  unsigned long long newton_calls;
  unsigned long long newton_iters_wave;
  int nb_over_cap;
  int nonfinite;
  int cont_over_cap;
  int _pad;
};

// ---- LDS-tiled pair loops (wf_pair.hip) ----------------------------------
// Timing aid: when start is set, the next pair-pass launch (dens, rates,
// accel, cond, fused cond + dens) records its own start and end in these two
// events (the kernel's dispatch timestamps, as a kernel trace reports them)
// and clears them.  One process-wide pair, for a host thread timing one
// domain at a time.
struct LaunchEvents {
  hipEvent_t start, stop;
};
extern LaunchEvents g_next_launch;

// FP64 pair loops (uh: every particle has the same h)
void launch_dens_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int n, int solid,
                       double dt, bool fuse, bool uh);
void launch_rates_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                        int quirks, bool uh);
void launch_accel_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                        double alpha, double beta, bool uh, const StatsFold& fold = {});
void launch_cond_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int n, double T_inf,
                       double h_conv, double dt, bool leap, bool first, bool uh);

// the mixed acceleration pass with each particle's BC + position / velocity
// update done as soon as its acceleration is summed (contact-free,
// non-axisymmetric steps with every particle solid); the new positions and
// FP32 mirrors go to out's planes (IntegrateOut, above)
void launch_accel_integrate_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                                  double alpha, double beta, float cs32, bool uh, double dt,
                                  const BcTable& bc, const Frame32& fr, const IntegrateOut& out,
                                  const StatsFold& fold = {});

// mixed precision (FP32 pair loops, FP32 terms summed in FP64)
// uh: every particle has the same h
void launch_dens_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int n, int solid,
                       double dt, bool fuse, float cs32, bool uh);
void launch_rates_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                        int quirks, float cs32, bool uh);
void launch_accel_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int solid,
                        double alpha, double beta, float cs32, bool uh, const StatsFold& fold = {});
void launch_cond_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int n, double T_inf,
                       double h_conv, double dt, bool leap, bool first, float cs32, bool uh);
// conduction of this step fused with the density pass (+ update) of the
// next; with pz_d2 (device, FP64 bits of the plastic-zone trigger's largest
// squared hot displacement) the density half writes nothing when
// 4 * d2 > pz_skin2, the host's own test for a rebuild at the next step
void launch_cond_dens_tiled(hipStream_t s, const Fields& f, const TileLists& tl, int n,
                            int solid, double T_inf, double h_conv, double dt, bool leap,
                            bool first, bool uh, const unsigned long long* pz_d2,
                            double pz_skin2);
void launch_cond_dens_mixed(hipStream_t s, const Fields& f, const TileLists& tl, int n,
                            int solid, double T_inf, double h_conv, double dt, bool leap,
                            bool first, float cs32, bool uh, const unsigned long long* pz_d2,
                            double pz_skin2);
// refresh every FP32 mirror / FP64 record from the FP64 state (after uploads
// and rebuilds)
void launch_mirror32(hipStream_t s, const Fields& f, int n, const Frame32& fr);
void launch_records64(hipStream_t s, const Fields& f, int n);

// ---- per-particle kernels (wf_kernels.hip) --------------------------------
void launch_update_density(hipStream_t s, const Fields& f, const StepParams& p, double dt);
void launch_pressure(hipStream_t s, const Fields& f, const StepParams& p);
// s32: the FP32 stress-state planes (mixed precision) instead of the FP64 ones
void launch_stress(hipStream_t s, const Fields& f, const StepParams& p, BlockStats* bs,
                   bool fuse_pressure, bool energy, int rec, bool s32, const RetArgs& ra);
// fold the per-block statistics of the last stress (+ return) pass into ctr
// largest squared displacement since the last build (x - xb) over the
// particles with eps_p > ep_hot, as the bits of a non-negative double
// max-combined into *d2_bits (which the caller clears)
void launch_hot_disp(hipStream_t s, const Fields& f, const double* const xb[3], int n,
                     double ep_hot, unsigned long long* d2_bits);
void launch_fold_stats(hipStream_t s, const BlockStats* bs, int nblocks, Counters* ctr);
// ---- return-map surrogate (wf_mlp.hip) -------------------------------------
// MFMA inference of the surrogate on given feature rows (test hook)
// the surrogate's per-lane MFMA operands, kMlpPackWords 16-byte words a lane
// (64 lanes), written from the weights after every weight update
constexpr int kMlpPackWords = 11;
void launch_mlp_pack(hipStream_t s, const float* w, void* pack);
// inference on n explicit feature rows from the packed weights
void launch_mlp_infer(hipStream_t s, const void* pack, const float* X, int n, float* y);
// one gradient of the mean squared error over n samples (deterministic)
void launch_mlp_grad(hipStream_t s, const float* w, const float* X, const float* Y, int n,
                     int stride, float* part, float* loss_part, int nblocks, float* grad,
                     float* loss);
// one Adam iteration: gradient block partials over samples s*stride, s < n,
// then their sum in block order and the Adam step t (both kernels)
void launch_mlp_train_iter(hipStream_t s, float* w, float* m, float* v, const float* X,
                           const float* Y, int n, int stride, float* part, float* loss_part,
                           int nblocks, int t, float lr, float* loss);
// gradient blocks for n samples: one 64-sample chunk each, at most kMlpMaxBlocks
int mlp_grad_blocks(int n);
// record msg as the calling thread's wf_last_error() text; returns code
int set_last_error(int code, const char* msg);
// gather the logged samples (ascending particle index) into X (n x 4), Y
void launch_mlp_collect(hipStream_t s, const Fields& f, int n, float* X, float* Y, int* flags,
                        int* scan, int* scan_part, int* count);
void launch_axisymm_bc(hipStream_t s, const Fields& f, int n);
void launch_apply_bc(hipStream_t s, const Fields& f, int n, int zone, double vx, double vy,
                     double vz);
void launch_update_pos(hipStream_t s, const Fields& f, int n, double dt);
void launch_update_vel(hipStream_t s, const Fields& f, int n, double dt);
// rec kRec32: also refresh xr and v4 (fr is the FP32 frame); kRec64: kv
void launch_integrate(hipStream_t s, const Fields& f, int n, double dt, const BcTable& bc,
                      int rec, const Frame32& fr);
void launch_leapfrog(hipStream_t s, const Fields& f, int n, double dt, bool first);
void launch_sum(hipStream_t s, const double* v, int n, double* partial, int nblocks,
                double* out);

// ---- neighbour search (wf_nsearch.hip) -----------------------------------
struct NsWork;  // opaque scratch of the search
struct NsResult {
  int over_cap;
};
// Build order + lists.  pos_src: 3 planes (sorted order) the search reads.
// Outputs: new_from_old gather map (slot -> old sorted index) in `gather`,
// lists in list/count (new order).  Returns WF_OK or an error code.
// perm: sorted slot -> original index of the current order; perm_tmp receives
// the new one.
// Re-list, without reordering, the tiles within two cells of a particle with
// eps_p > ep_hot, on the grid g of the last ns_build, and set their
// particles' build positions xb to x; *relisted gets the tile count.
// WF_E_NBCAP: a tile outgrew its staging region (a full build is needed).
int ns_relist_local(hipStream_t s, NsWork* w, const double* const x[3], const double* pl_strain,
                    double ep_hot, int n, int64_t npad, const GridDesc& g, double radius, int cap,
                    const int* perm, int* count, Counters* ctr, double* const xb[3],
                    int* relisted);
// max_dim (may be null): the largest grid, in cells per axis, the caller can
// represent; a larger one returns WF_E_UNSUPPORTED before anything changes
int ns_build(hipStream_t s, NsWork** work, const double* const pos_src[3], int n, int64_t npad,
             double radius, int cap, int* gather, int* count, Counters* ctr,
             int* perm, int* perm_tmp, TileLists* tiles, GridDesc* grid_out,
             const int* max_dim);
void ns_free(NsWork* w);
// the list builder's compaction limit (launch_tile_lists; 0: 4096)
void ns_set_compact_max(NsWork** work, int compact_max);
// exclusive scan of m ints into out[0..m] (out[m] = total); part needs
// ceil(m/1024)+1 ints
void launch_scan_exclusive(hipStream_t s, const int* in, int64_t m, int* part, int* out);

// generic gathers used to reorder fields: dst[i] = src[idx[i]]
void launch_gather_f64(hipStream_t s, double* dst, const double* src, const int* idx, int n);
void launch_gather_i32(hipStream_t s, int* dst, const int* src, const int* idx, int n);
// scatter for downloads: dst[idx[i]] = src[i]
void launch_scatter_f64(hipStream_t s, double* dst, const double* src, const int* idx, int n);
void launch_scatter_i32(hipStream_t s, int* dst, const int* src, const int* idx, int n);
// interleave components: dst[i*comps + c] = planes[c][i] and back
void launch_fill_f64(hipStream_t s, double* dst, double v, int n);

}  // namespace wf
