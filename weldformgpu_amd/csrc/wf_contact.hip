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
__global__ __launch_bounds__(256) void k_contact_nb(Fields f, int solid, MeshDev m, TriGrid g,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ sorted,
                                                   ContactLists cl, int id_free_surf,
                                                   Counters* ctr) {
  // the thread's candidates, kept sorted in LDS while they are collected
  __shared__ int L[kMaxContNb * 256];
  const int tl = threadIdx.x;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= solid) return;
  int c = 0, over = 0;
  int cx, cy, cz;
  if (f.ID[i] == id_free_surf && tri_cell(g, f.x[0][i], f.x[1][i], f.x[2][i], cx, cy, cz)) {
    const double xi = f.x[0][i], yi = f.x[1][i], zi = f.x[2][i], r = 2. * f.h[i];
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          const int x = cx + dx, y = cy + dy, z = cz + dz;
          if (x < 0 || y < 0 || z < 0 || x >= g.dim[0] || y >= g.dim[1] || z >= g.dim[2])
            continue;
          const int cell = (z * g.dim[1] + y) * g.dim[0] + x;
          for (int s = start[cell]; s < start[cell + 1]; s++) {
            const int e = sorted[s];
            const double ex = xi - m.c[0][e], ey = yi - m.c[1][e], ez = zi - m.c[2][e];
            if (!(sqrt(ex * ex + ey * ey + ez * ez) < r)) continue;
            if (c == kMaxContNb) {
              over++;
              if (e > L[(c - 1) * 256 + tl]) continue;
              c--;  // drop the largest index
            }
            int p = c;
            while (p > 0 && L[(p - 1) * 256 + tl] > e) {
              L[p * 256 + tl] = L[(p - 1) * 256 + tl];
              p--;
            }
            L[p * 256 + tl] = e;
            c++;
          }
        }
  }
  for (int k = 0; k < c; k++) cl.list[(int64_t)k * cl.npad + i] = L[k * 256 + tl];
  cl.count[i] = c;
  if (over) atomicAdd(&ctr->cont_over_cap, 1);
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
                       const ContactLists& cl, int id_free_surf, Counters* ctr) {
  if (solid <= 0) return;
  hipLaunchKernelGGL(k_contact_nb, dim3(gblk(solid)), dim3(256), 0, s, f, solid, m, g, start,
                     sorted, cl, id_free_surf, ctr);
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
