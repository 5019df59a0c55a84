// wf_contact.hip — rigid triangle-mesh penalty contact (gfx950).
//
// The reference turns every rigid triangle into a zero-mass particle at its
// centroid and finds contact pairs through the SPH neighbour search
// (Domain.cpp:967-983, Contact.cu:29-62).  Here the rigid meshes live in
// their own node/element arrays and a uniform grid over triangle centroids
// (cell edge >= 2 max h) gives each free-surface particle the triangles whose
// centroid lies within 2 h_i — the reference's contact criterion
// (Contact.cu:48) — in ascending element order, at most kMaxContNb
// (MAX_NB_COUNT, Contact.cu:7).
//
// Forces follow CalcContactForcesWang (Contact.cu:134-321): per candidate,
// the predicted position (with the acceleration already raised by earlier
// candidates), its distance to the element plane, the inside-triangle test
// by edge cross products, the penalty normal force (which replaces the
// stored contact force, while the acceleration accumulates) and static or
// dynamic Coulomb friction from the predicted tangential slip.  Mesh motion,
// centroids, normals and plane constants follow TriMesh_d::Move /
// CalcCentroids / CalcNormals / UpdatePlaneCoeff (Mesh.cu:223-276).
#include <algorithm>
#include <climits>

#include "wf_device.hpp"

namespace wf {

__global__ __launch_bounds__(256) void k_tri_bbox(MeshDev m, double* out) {
  __shared__ double sh[6][256];
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int e = threadIdx.x; e < m.ne; e += 256)
    for (int c = 0; c < 3; c++) {
      mn[c] = fmin(mn[c], m.c[c][e]);
      mx[c] = fmax(mx[c], m.c[c][e]);
    }
  for (int c = 0; c < 3; c++) {
    sh[c][threadIdx.x] = mn[c];
    sh[3 + c][threadIdx.x] = mx[c];
  }
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int c = 0; c < 3; c++) {
        sh[c][threadIdx.x] = fmin(sh[c][threadIdx.x], sh[c][threadIdx.x + s]);
        sh[3 + c][threadIdx.x] = fmax(sh[3 + c][threadIdx.x], sh[3 + c][threadIdx.x + s]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) out[threadIdx.x] = sh[threadIdx.x][0];
}

// max of v[0..n) in two passes: each block reduces a grid-stride share into
// part[blockIdx.x], then one block reduces the parts (max is exact, so the
// result does not depend on the split)
__global__ __launch_bounds__(256) void k_max_f64(const double* v, int64_t n, double* out) {
  __shared__ double sh[256];
  double mx = -1e300;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    mx = fmax(mx, v[i]);
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = sh[0];
}

// grid cell of a point; false when it lies beyond one cell outside the grid
// (or is not finite), so that no cell of its 3x3x3 block exists
__device__ __forceinline__ bool tri_cell(const TriGrid& g, double x, double y, double z, int& cx,
                                         int& cy, int& cz) {
  const double fx = floor((x - g.org[0]) * g.inv_cs);
  const double fy = floor((y - g.org[1]) * g.inv_cs);
  const double fz = floor((z - g.org[2]) * g.inv_cs);
  if (!(fx >= -1.0 && fy >= -1.0 && fz >= -1.0 && fx <= (double)g.dim[0] &&
        fy <= (double)g.dim[1] && fz <= (double)g.dim[2]))
    return false;
  cx = (int)fx;
  cy = (int)fy;
  cz = (int)fz;
  return true;
}

__global__ void k_tri_count(MeshDev m, TriGrid g, int* cnt, int* rank) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  int cx = 0, cy = 0, cz = 0;
  tri_cell(g, m.c[0][e], m.c[1][e], m.c[2][e], cx, cy, cz);
  cx = min(max(cx, 0), g.dim[0] - 1);
  cy = min(max(cy, 0), g.dim[1] - 1);
  cz = min(max(cz, 0), g.dim[2] - 1);
  const int c = (cz * g.dim[1] + cy) * g.dim[0] + cx;
  rank[e] = atomicAdd(&cnt[c], 1);
  m.cell[e] = c;
}

__global__ void k_tri_place(MeshDev m, const int* start, const int* rank, int* sorted) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e < m.ne) sorted[start[m.cell[e]] + rank[e]] = e;
}

// contact candidates of free-surface particles: triangles whose centroid is
// within 2 h_i (Contact.cu:40-55), ascending element index; candidates past
// kMaxContNb are counted and the smallest indices kept (the reference writes
// slot 20 before its overflow check)
// Contact candidates in two passes.  Pass 1 clears every particle's count and
// appends the particles that need candidates (free-surface ID, within a cell
// of the triangle grid) to `act` (act[0] the length, the indices from
// act[1]), in no particular order: pass 2 writes only each particle's own row.
__global__ __launch_bounds__(256) void k_contact_active(Fields f, int solid, TriGrid g,
                                                       ContactLists cl, int id_free_surf,
                                                       int* __restrict__ act) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool need = false;
  if (i < solid) {
    int cx, cy, cz;
    need = f.ID[i] == id_free_surf && tri_cell(g, f.x[0][i], f.x[1][i], f.x[2][i], cx, cy, cz);
    cl.count[i] = 0;
  }
  const unsigned long long b = __ballot(need);
  if (!b) return;
  const int lane = lane_id();
  const int lead = __ffsll((long long)b) - 1;
  int base = 0;
  if (lane == lead) base = atomicAdd(act, __popcll(b));
  base = __shfl(base, lead);
  if (need) act[1 + base + __popcll(b & ((1ull << lane) - 1ull))] = i;
}

// Orders this wave's LDS writes before its following LDS reads (the wave's
// own operations only; other waves of the block are not waited for).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// Pass 2 (CalcContactNb, Contact.cu:29-62): one wave per listed particle.
// Lanes 0..26 read the 3x3x3 cells around the particle's cell; the wave then
// tests their triangles 64 at a time (lane k of a batch takes the k-th
// triangle of the cells' concatenation) against sqrt(d2) < 2 h in FP64, and
// merges the hits into the particle's list, kept as the kMaxContNb smallest
// element indices in ascending order (lanes 0..L-1 hold it in a register).
// A merged value's position is its rank in the union of the list and the
// batch's hits (indices are distinct: a triangle lies in one cell).  A
// particle with more than kMaxContNb hits is counted once in cont_over_cap.
constexpr int kCnbWaves = 4;
__global__ __launch_bounds__(64 * kCnbWaves) void k_contact_nb(Fields f, MeshDev m, TriGrid g,
                                                              const int* __restrict__ start,
                                                              const int* __restrict__ sorted,
                                                              ContactLists cl,
                                                              const int* __restrict__ act,
                                                              Counters* ctr) {
  __shared__ int s_pre[kCnbWaves][28];  // first flat position of each cell
  __shared__ int s_beg[kCnbWaves][27];  // its first entry in `sorted`
  __shared__ int s_top[kCnbWaves][kMaxContNb];
  const int w = threadIdx.x >> 6, lane = lane_id();
  const int nact = act[0];
  const int nw = gridDim.x * kCnbWaves;
  for (int a = blockIdx.x * kCnbWaves + w; a < nact; a += nw) {
    const int i = act[1 + a];
    const double xi = f.x[0][i], yi = f.x[1][i], zi = f.x[2][i], r = 2. * f.h[i];
    int cx, cy, cz;
    tri_cell(g, xi, yi, zi, cx, cy, cz);  // inside (pass 1 listed it)
    int beg = 0, cnt = 0;
    if (lane < 27) {
      const int x = cx + lane % 3 - 1, y = cy + (lane / 3) % 3 - 1, z = cz + lane / 9 - 1;
      if (x >= 0 && y >= 0 && z >= 0 && x < g.dim[0] && y < g.dim[1] && z < g.dim[2]) {
        const int cell = (z * g.dim[1] + y) * g.dim[0] + x;
        beg = start[cell];
        cnt = start[cell + 1] - beg;
      }
    }
    int incl = cnt;  // inclusive prefix over lanes 0..26 (lanes above 26 add 0)
    for (int d = 1; d < 32; d <<= 1) {
      const int o = __shfl_up(incl, d);
      if (lane >= d) incl += o;
    }
    if (lane < 27) {
      s_pre[w][lane + 1] = incl;
      s_beg[w][lane] = beg;
    }
    if (lane == 0) s_pre[w][0] = 0;
    const int total = __shfl(incl, 26);
    wave_lds_sync();
    int L = 0, hits = 0;
    int tq = INT_MAX;  // lane q < L: the list's q-th entry
    for (int base = 0; base < total; base += 64) {
      const int k = base + lane;
      int e = INT_MAX;
      bool hit = false;
      if (k < total) {
        int sg = 0;
        for (int q = 1; q < 27; q++) sg += k >= s_pre[w][q];
        e = sorted[s_beg[w][sg] + (k - s_pre[w][sg])];
        const double ex = xi - m.c[0][e], ey = yi - m.c[1][e], ez = zi - m.c[2][e];
        hit = sqrt(ex * ex + ey * ey + ez * ez) < r;
      }
      const unsigned long long hm = __ballot(hit);
      if (!hm) continue;
      int below_e = 0, below_t = 0;  // hits smaller than my hit / my list entry
      for (unsigned long long mm = hm; mm; mm &= mm - 1ull) {
        const int v = __shfl(e, __ffsll((long long)mm) - 1);
        below_e += v < e;
        below_t += v < tq;
      }
      for (int q = 0; q < L; q++) below_e += __shfl(tq, q) < e;
      if (hit && below_e < kMaxContNb) s_top[w][below_e] = e;
      if (lane < L && lane + below_t < kMaxContNb) s_top[w][lane + below_t] = tq;
      wave_lds_sync();
      const int nh = __popcll(hm);
      L = min(L + nh, kMaxContNb);
      hits += nh;
      tq = lane < L ? s_top[w][lane] : INT_MAX;
      wave_lds_sync();
    }
    if (lane < L) cl.list[(int64_t)lane * cl.npad + i] = tq;
    if (lane == 0) {
      cl.count[i] = L;
      if (hits > kMaxContNb) atomicAdd(&ctr->cont_over_cap, 1);
    }
  }
}

// penalty contact + Coulomb friction (Contact.cu:134-321)
__global__ __launch_bounds__(256) void k_contact_force(Fields f, int solid, MeshDev m,
                                                       ContactLists cl, ContactParams cp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= solid) return;
  const int cnt = cl.count[i];
  double cf[3] = {0.0, 0.0, 0.0};
  double qf = 0.0;  // sliding-friction heat power [W]
  if (cnt > 0) {
    const double dt = cp.dt;
    const double mi = f.m[i], hi = f.h[i];
    const double x[3] = {f.x[0][i], f.x[1][i], f.x[2][i]};
    const double v[3] = {f.v[0][i], f.v[1][i], f.v[2][i]};
    double a[3] = {f.a[0][i], f.a[1][i], f.a[2][i]};
    // normal stiffness and its factor: 0.4 * 2 m/dt^2 (Contact.cu:247,266)
    // or kappa E dx
    const double kij = cp.law == WF_CONTACT_REFERENCE ? 2.0 * mi / (dt * dt) : cp.kappa_E_dx;
    const double fn = cp.law == WF_CONTACT_REFERENCE ? 0.4 : 1.0;
    for (int k = 0; k < cnt; k++) {
      const int e = cl.list[(int64_t)k * cl.npad + i];
      double xp[3];
      for (int c = 0; c < 3; c++) xp[c] = x[c] + v[c] * dt + a[c] * dt * dt / 2.;
      const double n[3] = {m.nrm[0][e], m.nrm[1][e], m.nrm[2][e]};
      const double dist = n[0] * xp[0] + n[1] * xp[1] + n[2] * xp[2] - m.pp[e];
      if (!(dist < hi)) continue;
      const double Q[3] = {x[0] - dist * n[0], x[1] - dist * n[1], x[2] - dist * n[2]};
      bool inside = true;
      for (int l = 0; l < 3 && inside; l++) {
        const int nl = m.el[3 * e + l], nn = m.el[3 * e + (l == 2 ? 0 : l + 1)];
        const double u[3] = {m.x[0][nn] - m.x[0][nl], m.x[1][nn] - m.x[1][nl],
                             m.x[2][nn] - m.x[2][nl]};
        const double w[3] = {Q[0] - m.x[0][nl], Q[1] - m.x[1][nl], Q[2] - m.x[2][nl]};
        const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                              u[0] * w[1] - u[1] * w[0]};
        if (cr[0] * n[0] + cr[1] * n[1] + cr[2] * n[2] < 0.0) inside = false;
      }
      if (!inside) continue;
      const double delta = hi - dist;
      for (int c = 0; c < 3; c++) {
        cf[c] = fn * kij * delta * n[c];
        a[c] += cf[c] / mi;
      }
      if (cp.mu_s > 0.0) {
        // element velocity: mean of its node velocities (Contact.cu:85-112)
        double du[3];
        for (int c = 0; c < 3; c++) {
          const double ve =
              (m.v[c][m.el[3 * e]] + m.v[c][m.el[3 * e + 1]] + m.v[c][m.el[3 * e + 2]]) / 3.;
          du[c] = xp[c] - x[c] - ve * dt;
        }
        const double dn = du[0] * n[0] + du[1] * n[1] + du[2] * n[2];
        double tg[3];
        for (int c = 0; c < 3; c++) tg[c] = kij * (du[c] - dn * n[c]);
        const double ltg = sqrt(tg[0] * tg[0] + tg[1] * tg[1] + tg[2] * tg[2]);
        const double lcf = sqrt(cf[0] * cf[0] + cf[1] * cf[1] + cf[2] * cf[2]);
        if (ltg < cp.mu_s * lcf) {  // static: no slip
          for (int c = 0; c < 3; c++) a[c] -= tg[c] / mi;
        } else {
          for (int c = 0; c < 3; c++) {
            const double fd = cp.mu_d * lcf * tg[c] / ltg;
            cf[c] -= fd;
            a[c] -= fd / mi;
          }
          // dissipated power: |F_friction| x slip speed, |du_t| = |tg| / k
          qf += cp.fric_heat * (cp.mu_d * lcf) * (ltg / kij) / dt;
        }
      }
    }
    for (int c = 0; c < 3; c++) f.a[c][i] = a[c];
  }
  for (int c = 0; c < 3; c++) f.cf[c][i] = cf[c];
  f.qfric[i] = qf;
}

// rigid motion of the nodes: v = v_m + w x (x - c), x += v dt (Mesh.cu:260-276)
__global__ void k_mesh_move(MeshDev m, MeshMotions mo, double dt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= m.nn) return;
  const MeshMotion& M = mo.m[m.node_mesh[q]];
  const double r[3] = {m.x[0][q] - M.center[0], m.x[1][q] - M.center[1],
                       m.x[2][q] - M.center[2]};
  const double v[3] = {M.vel[0] + (M.omega[1] * r[2] - M.omega[2] * r[1]),
                       M.vel[1] + (M.omega[2] * r[0] - M.omega[0] * r[2]),
                       M.vel[2] + (M.omega[0] * r[1] - M.omega[1] * r[0])};
  for (int c = 0; c < 3; c++) {
    m.v[c][q] = v[c];
    m.x[c][q] += v[c] * dt;
  }
}

// centroids, unit normals (n0n1 x n0n2) and plane constants n . x_node0
// (Mesh.cu:223-258).  The reference's UpdatePlaneCoeff reads node
// elnode[3e]+2 instead of a node of the element; on a plate all nodes share
// the plane, so the constant is the same.
__global__ void k_mesh_geom(MeshDev m) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m.ne) return;
  const int a = m.el[3 * e], b = m.el[3 * e + 1], c = m.el[3 * e + 2];
  double u[3], w[3];
  for (int k = 0; k < 3; k++) {
    m.c[k][e] = (m.x[k][a] + m.x[k][b] + m.x[k][c]) / 3.;
    u[k] = m.x[k][b] - m.x[k][a];
    w[k] = m.x[k][c] - m.x[k][a];
  }
  const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2],
                        u[0] * w[1] - u[1] * w[0]};
  const double l = sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
  double nr[3];
  for (int k = 0; k < 3; k++) {
    nr[k] = cr[k] / l;
    m.nrm[k][e] = nr[k];
  }
  m.pp[e] = m.x[0][a] * nr[0] + m.x[1][a] * nr[1] + m.x[2][a] * nr[2];
}

// not_write_surf_ID: particles whose ID is non-zero keep it
// (WeldFormGPU.cu:845-858)
__global__ void k_surface_lock(Fields f, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f.idlock[i] = f.ID[i] != 0;
}

__global__ void k_contact_norm(Fields f, int n, int id, double* out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x = f.cf[0][i], y = f.cf[1][i], z = f.cf[2][i];
  out[i] = f.ID[i] == id ? sqrt(x * x + y * y + z * z) : 0.0;
}

static int gblk(int n) { return (n + 255) / 256; }

void launch_mesh_geom(hipStream_t s, const MeshDev& m) {
  if (m.ne > 0) hipLaunchKernelGGL(k_mesh_geom, dim3(gblk(m.ne)), dim3(256), 0, s, m);
}
void launch_mesh_move(hipStream_t s, const MeshDev& m, const MeshMotions& mo, double dt) {
  if (m.nn > 0) hipLaunchKernelGGL(k_mesh_move, dim3(gblk(m.nn)), dim3(256), 0, s, m, mo, dt);
  launch_mesh_geom(s, m);
}
void launch_tri_bbox(hipStream_t s, const MeshDev& m, double* out) {
  hipLaunchKernelGGL(k_tri_bbox, dim3(1), dim3(256), 0, s, m, out);
}
void launch_max_f64(hipStream_t s, const double* v, int n, double* out, double* part) {
  const int nb = std::max(1, std::min(kMaxParts, (int)(((int64_t)n + 255) / 256)));
  hipLaunchKernelGGL(k_max_f64, dim3(nb), dim3(256), 0, s, v, (int64_t)n, part);
  hipLaunchKernelGGL(k_max_f64, dim3(1), dim3(256), 0, s, (const double*)part, (int64_t)nb, out);
}
void launch_tri_bin(hipStream_t s, const MeshDev& m, const TriGrid& g, int* cnt, int* rank,
                    int* start, int* sorted, int* scan_part) {
  if (m.ne <= 0) return;
  hipLaunchKernelGGL(k_tri_count, dim3(gblk(m.ne)), dim3(256), 0, s, m, g, cnt, rank);
  launch_scan_exclusive(s, cnt, (int64_t)g.dim[0] * g.dim[1] * g.dim[2], scan_part, start);
  hipLaunchKernelGGL(k_tri_place, dim3(gblk(m.ne)), dim3(256), 0, s, m, start, rank, sorted);
}
void launch_surface_lock(hipStream_t s, const Fields& f, int n) {
  hipLaunchKernelGGL(k_surface_lock, dim3(gblk(n)), dim3(256), 0, s, f, n);
}
void launch_contact_nb(hipStream_t s, const Fields& f, int solid, const MeshDev& m,
                       const TriGrid& g, const int* start, const int* sorted,
                       const ContactLists& cl, int id_free_surf, int* act, Counters* ctr) {
  if (solid <= 0) return;  // act[0] was cleared by the caller
  hipLaunchKernelGGL(k_contact_active, dim3(gblk(solid)), dim3(256), 0, s, f, solid, g, cl,
                     id_free_surf, act);
  // enough waves to fill the chip; each loops over the listed particles
  const int blocks = std::max(1, std::min(2048, gblk(solid)));
  hipLaunchKernelGGL(k_contact_nb, dim3(blocks), dim3(64 * kCnbWaves), 0, s, f, m, g, start,
                     sorted, cl, (const int*)act, ctr);
}
void launch_contact_force(hipStream_t s, const Fields& f, int solid, const MeshDev& m,
                          const ContactLists& cl, const ContactParams& cp) {
  if (solid <= 0) return;
  hipLaunchKernelGGL(k_contact_force, dim3(gblk(solid)), dim3(256), 0, s, f, solid, m, cl, cp);
}
void launch_contact_norm(hipStream_t s, const Fields& f, int n, int id, double* out) {
  hipLaunchKernelGGL(k_contact_norm, dim3(gblk(n)), dim3(256), 0, s, f, n, id, out);
}

}  // namespace wf
