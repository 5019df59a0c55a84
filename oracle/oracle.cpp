// oracle.cpp — FP64 CPU restatement of the WeldFormGPU SPH step.
//
// TEST INFRASTRUCTURE ONLY (the checker, never the product).  Built with
// -O2 -ffp-contract=off so every FP64 expression rounds in source order, the
// way the reference's --fmad=false build does (CMakeLists.txt:29).  Each
// function cites the reference lines it restates.
#include <omp.h>
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr double kPi = 3.14159265358979323846;

struct V3 {
  double x, y, z;
};
inline V3 ld3(const double* a, int i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
inline void st3(double* a, int i, V3 v) {
  a[3 * i] = v.x;
  a[3 * i + 1] = v.y;
  a[3 * i + 2] = v.z;
}
inline V3 sub(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 scl(V3 a, double s) { return {a.x * s, a.y * s, a.z * s}; }
// vector_math.h:805-827: dot = a.x*b.x + a.y*b.y + a.z*b.z, length = sqrt(dot)
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline double length(V3 a) { return std::sqrt(dot(a, a)); }

// 9-component tensor, tensor.cuh:52-62
struct T9 {
  double xx, xy, xz, yx, yy, yz, zx, zy, zz;
};
inline T9 zero9() { return {0, 0, 0, 0, 0, 0, 0, 0, 0}; }
// FromFlatSym, tensor3.cu:62-70
inline T9 from_flat_sym(const double* f) {
  return {f[0], f[3], f[5], f[3], f[1], f[4], f[5], f[4], f[2]};
}
// FromFlatAntiSym, tensor3.cu:72-85
inline T9 from_flat_antisym(const double* f) {
  return {f[0], f[3], f[5], -f[3], f[1], f[4], -f[5], -f[4], f[2]};
}
// ToFlatSymPtr, tensor3.cu:95-100 (upper triangle)
inline void to_flat(const T9& t, double* f) {
  f[0] = t.xx;
  f[1] = t.yy;
  f[2] = t.zz;
  f[3] = t.xy;
  f[4] = t.yz;
  f[5] = t.xz;
}
// Trans, tensor3.cu:285-292
inline T9 trans(const T9& m) { return {m.xx, m.yx, m.zx, m.xy, m.yy, m.zy, m.xz, m.yz, m.zz}; }
// operator*(double, tensor3) and operator*(tensor3, double), tensor3.cu:294-308
inline T9 smul(const T9& m, double f) {
  return {m.xx * f, m.xy * f, m.xz * f, m.yx * f, m.yy * f, m.yz * f, m.zx * f, m.zy * f, m.zz * f};
}
// operator+, tensor3.cu:310-328
inline T9 tadd(const T9& a, const T9& b) {
  return {a.xx + b.xx, a.xy + b.xy, a.xz + b.xz, a.yx + b.yx, a.yy + b.yy,
          a.yz + b.yz, a.zx + b.zx, a.zy + b.zy, a.zz + b.zz};
}
// operator-(tensor, tensor) (component-wise, same file)
inline T9 tsub(const T9& a, const T9& b) {
  return {a.xx - b.xx, a.xy - b.xy, a.xz - b.xz, a.yx - b.yx, a.yy - b.yy,
          a.yz - b.yz, a.zx - b.zx, a.zy - b.zy, a.zz - b.zz};
}
// Identity, tensor3.cu:102-108
inline T9 ident() { return {1, 0, 0, 0, 1, 0, 0, 0, 1}; }

// tensor3 operator*(tensor3, tensor3), tensor3.cu:241-255, as written: the
// first column is (A*B)_ix, the second and third are (A*B^T)_ij.
inline T9 tmul_reference(const T9& a, const T9& b) {
  T9 r;
  r.xx = a.xx * b.xx + a.xy * b.yx + a.xz * b.zx;
  r.xy = a.xx * b.yx + a.xy * b.yy + a.xz * b.yz;
  r.xz = a.xx * b.zx + a.xy * b.zy + a.xz * b.zz;
  r.yx = a.yx * b.xx + a.yy * b.yx + a.yz * b.zx;
  r.yy = a.yx * b.yx + a.yy * b.yy + a.yz * b.yz;
  r.yz = a.yx * b.zx + a.yy * b.zy + a.yz * b.zz;
  r.zx = a.zx * b.xx + a.zy * b.yx + a.zz * b.zx;
  r.zy = a.zx * b.yx + a.zy * b.yy + a.zz * b.yz;
  r.zz = a.zx * b.zx + a.zy * b.zy + a.zz * b.zz;
  return r;
}
// the matrix product the reference intended
inline T9 tmul_true(const T9& a, const T9& b) {
  T9 r;
  r.xx = a.xx * b.xx + a.xy * b.yx + a.xz * b.zx;
  r.xy = a.xx * b.xy + a.xy * b.yy + a.xz * b.zy;
  r.xz = a.xx * b.xz + a.xy * b.yz + a.xz * b.zz;
  r.yx = a.yx * b.xx + a.yy * b.yx + a.yz * b.zx;
  r.yy = a.yx * b.xy + a.yy * b.yy + a.yz * b.zy;
  r.yz = a.yx * b.xz + a.yy * b.yz + a.yz * b.zz;
  r.zx = a.zx * b.xx + a.zy * b.yx + a.zz * b.zx;
  r.zy = a.zx * b.xy + a.zy * b.yy + a.zz * b.zy;
  r.zz = a.zx * b.xz + a.zy * b.yz + a.zz * b.zz;
  return r;
}
// operator*(tensor3, double3), tensor3.cu:225-231
inline V3 tvec(const T9& m, V3 v) {
  return {m.xx * v.x + m.xy * v.y + m.xz * v.z, m.yx * v.x + m.yy * v.y + m.yz * v.z,
          m.zx * v.x + m.zy * v.y + m.zz * v.z};
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------------------
// Cubic spline kernel, Functions.cu:27-39 (KT 0, Dim 3)
double or_kernel(double q, double h) {
  double C = 1.0 / (h * h * h * M_PI);
  if (q < 1.0) return C * (1.0 - (3.0 / 2.0) * q * q + (3.0 / 4.0) * q * q * q);
  if (q < 2.0) return C * ((1.0 / 4.0) * (2.0 - q) * (2.0 - q) * (2.0 - q));
  return 0.0;
}

// GradKernel, Functions.cu:76-88 (KT 0, Dim 3): (1/r) dW/dr
double or_grad_kernel(double q, double h) {
  double C = 1.0 / (h * h * h * h * M_PI);
  if (q == 0.0) return C / h * (-3.0 + (9.0 / 2.0) * q);
  if (q < 1.0) return C / (q * h) * (-3.0 * q + (9.0 / 4.0) * q * q);
  if (q < 2.0) return C / (q * h) * ((-3.0 / 4.0) * (2.0 - q) * (2.0 - q));
  return 0.0;
}

// ---------------------------------------------------------------------------
// Domain::AddBoxLength, cubic packing (type 1), Domain.cpp:270-314 and the
// 3D mass rule Domain.cpp:316-329.
int or_gen_box(const double V[3], double Lx, double Ly, double Lz, double r, double* xyz,
               int cap, double density, double* mass_out) {
  int count = 0;
  double zp = V[2];
  size_t k = 0;
  while (zp <= (V[2] + Lz - r)) {
    size_t j = 0;
    double yp = V[1];
    while (yp <= (V[1] + Ly - r)) {
      size_t i = 0;
      double xp = V[0];
      while (xp <= (V[0] + Lx - r)) {
        double x = V[0] + (2.0 * i + 1) * r;
        double y = V[1] + (2.0 * j + 1) * r;
        double z = V[2] + (2.0 * k + 1) * r;
        if (xyz && count < cap) {
          xyz[3 * count] = x;
          xyz[3 * count + 1] = y;
          xyz[3 * count + 2] = z;
        }
        count++;
        i++;
        xp = V[0] + (2 * i + 1) * r;
      }
      j++;
      yp = V[1] + (2.0 * j + 1) * r;
    }
    k++;
    zp = V[2] + (2.0 * k + 1) * r;
  }
  if (mass_out && xyz && count <= cap && count > 0) {
    double mx = V[0], my = V[1], mz = V[2];
    for (int a = 0; a < count; a++) {
      if (xyz[3 * a] > mx) mx = xyz[3 * a];
      if (xyz[3 * a + 1] > my) my = xyz[3 * a + 1];
      if (xyz[3 * a + 2] > mz) mz = xyz[3 * a + 2];
    }
    mx += r;
    my += r;
    mz += r;
    double tx = mx - V[0], ty = my - V[1], tz = mz - V[2];
    *mass_out = tx * ty * tz * density / count;
  }
  return count;
}

// calcHalfPartCount, Domain.cpp:473-488
int or_half_part_count(double r, double R, int xinc) {
  int ypartcount = -1;
  if (xinc > 0) {
    ypartcount = 1;
    double yp = r;
    double xp = r + (double)(xinc - 1) * 2. * r;
    double rad = std::sqrt(yp * yp + xp * xp);
    while (rad <= R - r) {
      yp += 2. * r;
      rad = std::sqrt(yp * yp + xp * xp);
      ypartcount++;
    }
    ypartcount -= 1;
  }
  return ypartcount;
}

// Domain::AddCylinderLength, Domain.cpp:490-574 (Dim 3)
int or_gen_cylinder(const double V[3], double Rxy, double Lz, double r, double* xyz, int cap,
                    double density, double* mass_out) {
  int count = 0;
  int numpartxy = or_half_part_count(r, Rxy, 1);
  double zp = V[2] + r;
  while (zp <= (V[2] + Lz - r)) {
    double yp = V[1] - r - (2. * r * (numpartxy - 1));
    int numypart = 2 * numpartxy;
    int yinc = numpartxy;
    int yinc_sign = -1;
    for (int j = 0; j < numypart; j++) {
      int numxpart = or_half_part_count(r, Rxy, yinc);
      double xp = V[0] - r - (2. * r * (numxpart - 1));
      for (int i = 0; i < 2 * numxpart; i++) {
        if (xyz && count < cap) {
          xyz[3 * count] = xp;
          xyz[3 * count + 1] = yp;
          xyz[3 * count + 2] = zp;
        }
        count++;
        xp += 2. * r;
      }
      yp += 2. * r;
      yinc += yinc_sign;
      if (yinc < 1) {
        yinc = 1;
        yinc_sign = 1;
      }
    }
    zp += 2. * r;
  }
  if (mass_out && count > 0) {
    double Vol = M_PI * Rxy * Rxy * Lz;
    *mass_out = Vol * density / count;
  }
  return count;
}

// Domain::AddCylUniformLength, Domain.cpp:1071-1193 (Dimension 3): radial
// rings per layer, z0 = r, the start point unused; rows/ang mark slice edges
int or_gen_cylinder_polar(int tag, double Rxy, double Lz, double r, double density, double ang,
                          int rows, double r_i, double* xyz, int* ID, int cap,
                          double* mass_out) {
  int count = 0;
  const double z0 = r;
  double zp = z0;
  while (zp <= (z0 + Lz - r)) {
    for (double ri = r_i; ri < Rxy; ri += 2. * r) {
      int tgcount;
      double dalpha;
      if (ri == 0.) {
        tgcount = 1;
        dalpha = 0.;
      } else {
        tgcount = (int)ceil((ang * ri) / (2. * r));
        dalpha = ang / (tgcount);
      }
      for (int alphai = 0; alphai < tgcount; alphai++) {
        int id = tag;
        if (ang < 2.0 * M_PI) {
          if (alphai < rows) id = 2;
          else if (alphai > tgcount - 1 - rows) id = 3;
        }
        const double xp = ri * cos(alphai * dalpha);
        const double yp = ri * sin(alphai * dalpha);
        if ((fabs(xp) < r / 10) && (fabs(yp) < r / 10)) id = 4;
        if (count < cap) {
          if (xyz) {
            xyz[3 * count] = xp;
            xyz[3 * count + 1] = yp;
            xyz[3 * count + 2] = zp;
          }
          if (ID) ID[count] = id;
        }
        count++;
      }
    }
    zp += 2.0 * r;
  }
  if (mass_out && count > 0) *mass_out = M_PI * Rxy * Rxy * Lz * density / count;
  return count;
}

// Domain::AssignZone, Domain.cpp:989-1005
int or_assign_zone(const double* xyz, int n, const double start[3], const double end[3], int id,
                   int* ID) {
  int partcount = 0;
  for (int a = 0; a < n; a++) {
    bool included = true;
    for (int i = 0; i < 3; i++)
      if (xyz[3 * a + i] < start[i] || xyz[3 * a + i] > end[i]) included = false;
    if (included) {
      ID[a] = id;
      partcount++;
    }
  }
  return partcount;
}

// ---------------------------------------------------------------------------
// Neighbour criterion, cuNSearchKernels.cu:160-165: positions and radius in
// float, diff = pj - pi, d2 = dx*dx + dy*dy + dz*dz (no contraction),
// neighbour iff d2 < r*r and d2 > 0.
static inline bool nb_test(const float* p, int i, int j, float r2) {
  float dx = p[3 * j] - p[3 * i];
  float dy = p[3 * j + 1] - p[3 * i + 1];
  float dz = p[3 * j + 2] - p[3 * i + 2];
  volatile float a = dx * dx;
  volatile float b = dy * dy;
  volatile float c = dz * dz;
  float d2 = (a + b) + c;
  return d2 < r2 && d2 > 0.0f;
}

int64_t or_neighbors_bruteforce(const double* xyz, int n, double radius, int* counts, int* offsets,
                                int* list, int64_t cap) {
  std::vector<float> p(3 * (size_t)n);
  for (size_t k = 0; k < 3 * (size_t)n; k++) p[k] = (float)xyz[k];
  float rf = (float)radius;
  float r2 = rf * rf;
  int64_t total = 0;
  for (int i = 0; i < n; i++) {
    offsets[i] = (int)total;
    int c = 0;
    for (int j = 0; j < n; j++)
      if (nb_test(p.data(), i, j, r2)) {
        if (total < cap) list[total] = j;
        total++;
        c++;
      }
    counts[i] = c;
  }
  offsets[n] = (int)total;
  return total;
}

// Cell-list search over a uniform grid of cell size > radius; the set it
// returns is exactly the criterion above (rows sorted ascending).  The cap
// of 70 (cuNSearchKernels.cu:169) is reported through over_cap, not applied.
int64_t or_neighbors(const double* xyz, int n, double radius, int* counts, int* offsets, int* list,
                     int64_t cap, int* over_cap) {
  std::vector<float> p(3 * (size_t)n);
  for (size_t k = 0; k < 3 * (size_t)n; k++) p[k] = (float)xyz[k];
  float rf = (float)radius;
  float r2 = rf * rf;
  if (n == 0) {
    offsets[0] = 0;
    if (over_cap) *over_cap = 0;
    return 0;
  }
  double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < n; i++)
    for (int d = 0; d < 3; d++) {
      mn[d] = std::min(mn[d], (double)p[3 * i + d]);
      mx[d] = std::max(mx[d], (double)p[3 * i + d]);
    }
  double cs = (double)rf * 1.0001;
  int dim[3];
  for (int d = 0; d < 3; d++) dim[d] = (int)std::floor((mx[d] - mn[d]) / cs) + 1;
  size_t ncell = (size_t)dim[0] * dim[1] * dim[2];
  std::vector<int> cell(n), cstart(ncell + 1, 0), order(n);
  for (int i = 0; i < n; i++) {
    int c[3];
    for (int d = 0; d < 3; d++) {
      c[d] = (int)std::floor(((double)p[3 * i + d] - mn[d]) / cs);
      c[d] = std::min(std::max(c[d], 0), dim[d] - 1);
    }
    cell[i] = (c[2] * dim[1] + c[1]) * dim[0] + c[0];
    cstart[cell[i] + 1]++;
  }
  for (size_t c = 0; c < ncell; c++) cstart[c + 1] += cstart[c];
  {
    std::vector<int> fill(cstart.begin(), cstart.end() - 1);
    for (int i = 0; i < n; i++) order[fill[cell[i]]++] = i;
  }
  std::vector<std::vector<int>> rows(n);
#pragma omp parallel for schedule(dynamic, 256)
  for (int i = 0; i < n; i++) {
    int ci = cell[i];
    int cx = ci % dim[0], cy = (ci / dim[0]) % dim[1], cz = ci / (dim[0] * dim[1]);
    std::vector<int>& row = rows[i];
    for (int dz = -1; dz <= 1; dz++)
      for (int dy = -1; dy <= 1; dy++)
        for (int dx = -1; dx <= 1; dx++) {
          int x = cx + dx, y = cy + dy, z = cz + dz;
          if (x < 0 || y < 0 || z < 0 || x >= dim[0] || y >= dim[1] || z >= dim[2]) continue;
          int c = (z * dim[1] + y) * dim[0] + x;
          for (int k = cstart[c]; k < cstart[c + 1]; k++) {
            int j = order[k];
            if (nb_test(p.data(), i, j, r2)) row.push_back(j);
          }
        }
    std::sort(row.begin(), row.end());
  }
  int64_t total = 0;
  int oc = 0;
  for (int i = 0; i < n; i++) {
    offsets[i] = (int)total;
    counts[i] = (int)rows[i].size();
    if (counts[i] > WF_MAX_NEIGHBORS) oc++;
    for (int j : rows[i]) {
      if (total < cap) list[total] = j;
      total++;
    }
  }
  offsets[n] = (int)total;
  if (over_cap) *over_cap = oc;
  return total;
}

// ---------------------------------------------------------------------------
// Material laws, Material.cuh:192-242.  Johnson-Cook homologous temperature
// is clamped to [0,1] (the reference takes pow() of a negative base when
// T < Tt, Material.cuh:211,216, which is NaN).
double or_hollomon_yield(const wf_material_params* mp, double strain) {
  // Material.cuh:51-59 (eps0 = sy0/E, eps1 = (sy0/K)^(1/m)) and :192-201
  double eps0 = mp->sy0 / mp->E;
  double eps1 = std::pow(mp->sy0 / mp->K_h, 1. / mp->m_h);
  if (strain + eps0 > eps1) return mp->K_h * std::pow(strain + eps0, mp->m_h);
  return mp->sy0;
}

static inline double jc_theta(const wf_material_params* mp, double T) {
  double T_h = (T - mp->Tt) / (mp->Tm - mp->Tt);
  if (!(T_h > 0.0)) T_h = 0.0;
  if (T_h > 1.0) T_h = 1.0;
  return T_h;
}

double or_jc_yield(const wf_material_params* mp, double strain, double strain_rate, double temp) {
  double T_h = jc_theta(mp, temp);
  double sr = strain_rate;
  if (strain_rate == 0.0) sr = 1.e-5;
  double sy = (mp->A + mp->B * std::pow(strain, mp->n)) * (1.0 + mp->C * std::log(sr / mp->eps0)) *
              (1.0 - std::pow(T_h, mp->m));
  return sy;
}

// d sigma_y / d eps_p of the Johnson-Cook law (the reference's tangent
// modulus, Material.cuh:229-242, without its arbitrary 0.1E at eps_p = 0)
double or_jc_hardening(const wf_material_params* mp, double ep, double strain_rate, double temp) {
  double T_h = jc_theta(mp, temp);
  double sr = strain_rate;
  if (strain_rate == 0.0) sr = 1.e-5;
  if (!(ep > 0.0)) return HUGE_VAL;
  return mp->n * mp->B * std::pow(ep, mp->n - 1.) * (1.0 + mp->C * std::log(sr / mp->eps0)) *
         (1.0 - std::pow(T_h, mp->m));
}

// Safeguarded Newton on r(g) = s_trial - 3 G g - s_y(ep + g), bracketed in
// [0, (s_trial - s_y(ep))/(3G)] where r changes sign (s_y is nondecreasing in
// eps_p).  Steps leaving the bracket are replaced by bisection.
double or_jc_newton(const wf_material_params* mp, double s_trial, double G, double ep,
                    double eps_rate, double T, double g0, double tol, int max_iter, int* iters) {
  double sy0 = or_jc_yield(mp, ep, eps_rate, T);
  double lo = 0.0, hi = (s_trial - sy0) / (3.0 * G);
  if (!(hi > 0.0)) {
    if (iters) *iters = 0;
    return 0.0;
  }
  double g = g0;
  if (!(g > lo) || !(g < hi)) g = (g0 >= hi) ? hi : lo;
  int it = 0;
  for (; it < max_iter; it++) {
    double sy = or_jc_yield(mp, ep + g, eps_rate, T);
    double r = s_trial - 3.0 * G * g - sy;
    if (std::fabs(r) <= tol * s_trial) break;
    if (r > 0.0) lo = g;
    else hi = g;
    double H = or_jc_hardening(mp, ep + g, eps_rate, T);
    double gn = g + r / (3.0 * G + H);
    if (!(gn > lo && gn < hi)) gn = 0.5 * (lo + hi);
    g = gn;
  }
  if (iters) *iters = it;
  return g;
}

// ---------------------------------------------------------------------------
// Pair loops.  Only free particles exist on the production path (SetFreePart
// sets IsFree = true, WeldFormGPU.cu:838), so di = rho_i, dj = rho_j,
// mj = m_j (InteractionAlt.cu:62-75), and PresEq = 0 (never uploaded,
// Domain_d.cu:105) selects EOS / SoundSpeed case 0.

// CalcDensInc, InteractionAlt.cu:166-246
void or_calc_dens_inc(or_state* s, const wf_config*) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->solid; i++) {
    double acc = 0.0;
    V3 xi = ld3(s->x, i), vi = ld3(s->v, i);
    for (int k = 0; k < s->nb_count[i]; k++) {
      int j = s->nb_list[s->nb_offset[i] + k];
      V3 xij = sub(xi, ld3(s->x, j));
      double rij = length(xij);
      double di = s->rho[i], dj = s->rho[j], mj = s->m[j];
      V3 vij = sub(vi, ld3(s->v, j));
      double h_ = (s->h[i] + s->h[j]) / 2.0;
      double GK = or_grad_kernel(rij / h_, h_);
      double temp1 = dot(vij, scl(xij, GK));
      acc += mj * (di / dj) * temp1;
    }
    s->drho[i] = acc;
  }
}

// UpdateDensity, Mechanical.cu:198-203
void or_update_density(or_state* s, double dt) {
  for (int i = 0; i < s->solid; i++) s->rho[i] += dt * s->drho[i];
}

// CalcRateTensors, InteractionAlt.cu:262-383
void or_calc_rate_tensors(or_state* s, const wf_config* c) {
  const bool q_xz = (c->quirks & WF_Q_ROT_XZ_SIGN) != 0;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->solid; i++) {
    T9 SR = zero9(), RR = zero9();
    V3 xi = ld3(s->x, i), vi = ld3(s->v, i);
    for (int k = 0; k < s->nb_count[i]; k++) {
      int j = s->nb_list[s->nb_offset[i] + k];
      V3 xij = sub(xi, ld3(s->x, j));
      double rij = length(xij);
      double dj = s->rho[j], mj = s->m[j];
      V3 vab = sub(vi, ld3(s->v, j));
      double h_ = (s->h[i] + s->h[j]) / 2.0;
      double GK = or_grad_kernel(rij / h_, h_);
      T9 E;
      E.xx = 2.0 * vab.x * xij.x;
      E.xy = vab.x * xij.y + vab.y * xij.x;
      E.xz = vab.x * xij.z + vab.z * xij.x;
      E.yx = E.xy;
      E.yy = 2.0 * vab.y * xij.y;
      E.yz = vab.y * xij.z + vab.z * xij.y;
      E.zx = E.xz;
      E.zy = E.yz;
      E.zz = 2.0 * vab.z * xij.z;
      E = smul(E, (-0.5) * GK);
      T9 W = zero9();
      W.xy = vab.x * xij.y - vab.y * xij.x;
      W.xz = vab.x * xij.z - vab.z * xij.x;
      W.yz = vab.y * xij.z - vab.z * xij.y;
      W.yx = -W.xy;
      if (q_xz) W.xz = -W.xz;  // InteractionAlt.cu:362 (zx stays 0)
      else W.zx = -W.xz;       // Interaction.cu:236
      W.zy = -W.yz;
      W = smul(W, -0.5 * GK);
      double mj_dj = mj / dj;
      SR = tadd(SR, smul(E, mj_dj));
      RR = tadd(RR, smul(W, mj_dj));
    }
    to_flat(RR, s->rotrate + 6 * i);
    to_flat(SR, s->strrate + 6 * i);
  }
}

// PressureKernelExt -> EOS case 0, Mechanical.cu:289-298, Functions.cu:211-213
void or_pressure(or_state* s) {
  for (int i = 0; i < s->n; i++) s->p[i] = 0.0 + (s->Cs[i] * s->Cs[i]) * (s->rho[i] - s->rho0[i]);
}

// StressStrainOne, Mechanical.cu:458-554, with the return-map choice of the
// config.  Reference mode (Mechanical.cu:504-535): scale back with the old
// sigma_y, re-evaluate Hollomon/JC sigma_y, dep = (s_trial - s_y)/(3G),
// bilinear s_y += dep*Ep.  Newton mode: radial return to the consistent
// plastic increment (bilinear closed form, JC safeguarded Newton).
void or_stress_strain(or_state* s, const wf_config* c) {
  const wf_material_params* mp = &c->mat;
  const bool q_prod = (c->quirks & WF_Q_TENSOR_PRODUCT) != 0;
  const double dt = c->dt;
  int64_t iters_total = 0, calls = 0;
#pragma omp parallel for schedule(static) reduction(+ : iters_total, calls)
  for (int i = 0; i < s->solid; i++) {
    T9 S = from_flat_sym(s->shearstress + 6 * i);
    T9 D = from_flat_sym(s->strrate + 6 * i);
    T9 W = from_flat_antisym(s->rotrate + 6 * i);
    T9 WT = trans(W);
    T9 SRT = q_prod ? tmul_reference(S, WT) : tmul_true(S, WT);
    T9 RS = q_prod ? tmul_reference(W, S) : tmul_true(W, S);
    double G = s->G[i];
    T9 dev = tsub(D, smul(ident(), 1.0 / 3.0 * (D.xx + D.yy + D.zz)));
    S = tadd(smul(tadd(tadd(smul(dev, 2.0 * G), SRT), RS), dt), S);

    double esr = std::sqrt(0.5 * ((D.xx - D.yy) * (D.xx - D.yy) + (D.yy - D.zz) * (D.yy - D.zz) +
                                  (D.zz - D.xx) * (D.zz - D.xx)) +
                           3.0 * (D.xy * D.xy + D.yz * D.yz + D.zx * D.zx));
    s->eff_strain_rate[i] = esr;
    double J2 = 0.5 * (S.xx * S.xx + 2.0 * S.xy * S.yx + 2.0 * S.xz * S.zx + S.yy * S.yy +
                       2.0 * S.yz * S.zy + S.zz * S.zz);
    double sig_trial = std::sqrt(3.0 * J2);
    double dep = 0.0;
    double T = s->T ? s->T[i] : 0.0;

    if (c->return_map == WF_RETURN_REFERENCE) {
      if (s->sigma_y[i] < sig_trial) S = smul(S, s->sigma_y[i] / sig_trial);
      if (mp->model == WF_MAT_HOLLOMON) s->sigma_y[i] = or_hollomon_yield(mp, s->pl_strain[i]);
      else if (mp->model == WF_MAT_JOHNSON_COOK)
        s->sigma_y[i] = or_jc_yield(mp, s->pl_strain[i], esr, T);
      s->sigma_eq[i] = sig_trial;
      if (sig_trial > s->sigma_y[i]) {
        double Ep = mp->E * mp->Et / (mp->E - mp->Et);
        if (Ep < 0) Ep = 1.0 * mp->E;
        dep = (sig_trial - s->sigma_y[i]) / (3. * G);
        s->pl_strain[i] += dep;
        s->sigma_eq[i] = s->sigma_y[i];
        if (mp->model == WF_MAT_BILINEAR) s->sigma_y[i] += dep * Ep;
      }
    } else {
      double ep = s->pl_strain[i];
      double sy;
      if (mp->model == WF_MAT_HOLLOMON) sy = or_hollomon_yield(mp, ep);
      else if (mp->model == WF_MAT_JOHNSON_COOK) sy = or_jc_yield(mp, ep, esr, T);
      else sy = s->sigma_y[i];
      s->sigma_eq[i] = sig_trial;
      if (sig_trial > sy) {
        double g = 0.0;
        if (mp->model == WF_MAT_BILINEAR) {
          double Ep = mp->E * mp->Et / (mp->E - mp->Et);
          g = (sig_trial - sy) / (3.0 * G + Ep);
          sy = sy + Ep * g;
        } else if (mp->model == WF_MAT_JOHNSON_COOK) {
          int it = 0;
          g = or_jc_newton(mp, sig_trial, G, ep, esr, T, 0.0, c->newton_tol, c->newton_max_iter,
                           &it);
          iters_total += it;
          calls += 1;
          sy = or_jc_yield(mp, ep + g, esr, T);
        } else {  // Hollomon: bisection to the consistent increment
          double lo = 0.0, hi = (sig_trial - sy) / (3.0 * G);
          for (int it = 0; it < 200; it++) {
            double mid = 0.5 * (lo + hi);
            double r = sig_trial - 3.0 * G * mid - or_hollomon_yield(mp, ep + mid);
            if (r > 0) lo = mid;
            else hi = mid;
          }
          g = 0.5 * (lo + hi);
          sy = or_hollomon_yield(mp, ep + g);
        }
        double scale = (sig_trial - 3.0 * G * g) / sig_trial;
        S = smul(S, scale);
        dep = g;
        s->pl_strain[i] = ep + g;
        s->sigma_eq[i] = sig_trial - 3.0 * G * g;
      }
      s->sigma_y[i] = sy;
    }
    if (s->q_plastic) s->q_plastic[i] = c->taylor_quinney * s->sigma_eq[i] * dep / dt;

    double pi = s->p[i];
    T9 Sig = tadd(smul(ident(), -pi), S);
    // Strain = deltat*StrainRate + Strain (Mechanical.cu:542); the reference
    // never loads Strain from the array (an uninitialised local), here it
    // accumulates the stored strain.
    T9 St = tadd(smul(D, dt), from_flat_sym(s->strain + 6 * i));
    to_flat(Sig, s->sigma + 6 * i);
    to_flat(St, s->strain + 6 * i);
    to_flat(S, s->shearstress + 6 * i);
  }
  s->newton_iters = iters_total;
  s->newton_calls = calls;
}

// CalcAccel, InteractionAlt.cu:18-150 (KernelType 0, XSPH 0)
void or_calc_accel(or_state* s, const wf_config* c) {
  const double Alpha = c->alpha, Beta = c->beta;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->solid; i++) {
    V3 ai = {0.0, 0.0, 0.0};
    V3 xi = ld3(s->x, i), vi = ld3(s->v, i);
    T9 Si = from_flat_sym(s->sigma + 6 * i);
    for (int k = 0; k < s->nb_count[i]; k++) {
      int j = s->nb_list[s->nb_offset[i] + k];
      V3 xij = sub(xi, ld3(s->x, j));
      double rij = length(xij);
      double di = s->rho[i], dj = s->rho[j], mj = s->m[j];
      V3 vij = sub(vi, ld3(s->v, j));
      double h_ = (s->h[i] + s->h[j]) / 2.0;
      double GK = or_grad_kernel(rij / h_, h_);
      T9 PI = zero9();
      if (Alpha != 0.0 || Beta != 0.0) {
        double MUij = h_ * dot(vij, xij) / (rij * rij + 0.01 * h_ * h_);
        double Ci = s->Cs[i], Cj = s->Cs[j];
        double Cij = 0.5 * (Ci + Cj);
        if (dot(vij, xij) < 0)
          PI = smul(ident(), (Alpha * Cij * MUij + Beta * MUij * MUij) / (0.5 * (di + dj)));
      }
      T9 Sj = from_flat_sym(s->sigma + 6 * j);
      T9 M = tadd(tadd(smul(Si, 1.0 / (di * di)), smul(Sj, 1.0 / (dj * dj))), PI);
      ai = add(ai, tvec(smul(M, mj), scl(xij, GK)));
    }
    st3(s->a, i, ai);
  }
}

// ApplyBCVel, Mechanical.cu:572-584
void or_apply_bc_vel(or_state* s, int zone, double vx, double vy, double vz) {
  for (int i = 0; i < s->n; i++)
    if (s->ID[i] == zone) {
      st3(s->a, i, {0.0, 0.0, 0.0});
      st3(s->v, i, {vx, vy, vz});
      if (s->va) st3(s->va, i, {vx, vy, vz});
    }
}

// UpdatePosFraser, Mechanical.cu:179-186: u_ = dt*v + 0.5*a*dt*dt
void or_update_pos_fraser(or_state* s, double dt) {
  for (int i = 0; i < s->n; i++) {
    V3 v = ld3(s->v, i), a = ld3(s->a, i);
    V3 u_ = add(scl(v, dt), scl(scl(scl(a, 0.5), dt), dt));
    st3(s->x, i, add(ld3(s->x, i), u_));
    st3(s->u, i, add(ld3(s->u, i), u_));
  }
}

// UpdateVel, Mechanical.cu:142-148
void or_update_vel(or_state* s, double dt) {
  for (int i = 0; i < s->n; i++) st3(s->v, i, add(ld3(s->v, i), scl(ld3(s->a, i), dt)));
}

// ThermalSolveKernel (Thermal.cu:101-142, Fraser eq. 3.99) on the mechanical
// neighbour lists, then CalcConvHeatKernel (Thermal.cu:36-64); plastic work
// heat q_plastic [W/m^3] and sliding-friction heat q_friction [W] (an
// extension: the reference has no friction heat) enter as sources;
// BC_T == WF_BCT_FIXED holds T.
void or_calc_temp_inc(or_state* s, const wf_config* c) {
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->n; i++) {
    double acc = 0.0;
    V3 xi = ld3(s->x, i);
    for (int k = 0; k < s->nb_count[i]; k++) {
      int j = s->nb_list[s->nb_offset[i] + k];
      V3 xij = sub(xi, ld3(s->x, j));
      double h_ = (s->h[i] + s->h[j]) / 2.0;
      double nxij = length(xij);
      double GK = or_grad_kernel(nxij / h_, h_);
      acc += s->m[j] / s->rho[j] *
             (4.0 * s->k_T[i] * s->k_T[j] / (s->k_T[i] + s->k_T[j]) * (s->T[i] - s->T[j])) *
             dot(xij, scl(xij, GK)) / (nxij * nxij);
    }
    acc *= 1. / (s->rho[i] * s->cp_T[i]);
    int bc = s->BC_T ? s->BC_T[i] : 0;
    if (bc == WF_BCT_CONVECTION) {
      double dS2 = std::pow(s->m[i] / s->rho[i], 0.666666666);
      double q_conv = c->h_conv * dS2 * (c->T_inf - s->T[i]) / s->m[i];
      acc += 1. / (s->cp_T[i]) * q_conv;
    }
    if (s->q_plastic && s->q_plastic[i] != 0.0) acc += s->q_plastic[i] * (1. / (s->rho[i] * s->cp_T[i]));
    if (s->q_friction && s->q_friction[i] != 0.0) acc += s->q_friction[i] / (s->m[i] * s->cp_T[i]);
    if (bc == WF_BCT_FIXED) acc = 0.0;
    s->dTdt[i] = acc;
  }
}

// ThermalCalcs, Thermal.cu:68-99 with TempCalcLeapfrogFirst/TempCalcLeapfrog
// (Thermal.cu:16-31)
void or_thermal_calcs(or_state* s, const wf_config* c, int* isfirst_step) {
  or_calc_temp_inc(s, c);
  double dt = c->dt;
  if (*isfirst_step) {
    for (int i = 0; i < s->n; i++) s->Ta[i] = s->T[i] - dt / 2.0 * s->dTdt[i];
    *isfirst_step = 0;
  } else {
    for (int i = 0; i < s->n; i++) {
      s->Tb[i] = s->Ta[i];
      s->Ta[i] += s->dTdt[i] * dt;
      s->T[i] = (s->Ta[i] + s->Tb[i]) / 2.;
    }
  }
}

// CalcIntEnergy, Energy.cu:5-19, per particle (the reference reads particle
// 0's tensors for every particle and sums racily; this is the intended form)
void or_int_energy(or_state* s, const wf_config* c) {
  for (int i = 0; i < s->n; i++) {
    const double* ss = s->shearstress + 6 * i;
    const double* sr = s->strrate + 6 * i;
    double de = s->m[i] / s->rho[i] *
                (ss[0] * sr[0] + ss[1] * sr[1] + ss[2] * sr[2] +
                 2.0 * (ss[3] * sr[3] + ss[4] * sr[4] + ss[5] * sr[5]));
    s->int_energy[i] += de * c->dt;
  }
}

// serial Kahan-compensated sum
double or_int_energy_sum(const or_state* s) {
  double sum = 0.0, comp = 0.0;
  for (int i = 0; i < s->n; i++) {
    double y = s->int_energy[i] - comp;
    double t = sum + y;
    comp = (t - sum) - y;
    sum = t;
  }
  return sum;
}

// ---------------------------------------------------------------------------
// surrogate MLP in FP64: y = W3 . relu(W2^T relu(W1^T x + b1) + b2) + b3
static void mlp_fwd(const double* w, const double* x, double* h1, double* h2, double* y) {
  for (int u = 0; u < 32; u++) {
    double a = w[128 + u];
    for (int k = 0; k < 4; k++) a += w[k * 32 + u] * x[k];
    h1[u] = a > 0.0 ? a : 0.0;
  }
  double out = w[1248];
  for (int u = 0; u < 32; u++) {
    double a = w[1184 + u];
    for (int k = 0; k < 32; k++) a += w[160 + k * 32 + u] * h1[k];
    h2[u] = a > 0.0 ? a : 0.0;
    out += w[1216 + u] * h2[u];
  }
  *y = out;
}

double or_mlp_forward(const double* w, const double* x) {
  double h1[32], h2[32], y;
  mlp_fwd(w, x, h1, h2, &y);
  return y;
}

double or_mlp_loss_grad(const double* w, const double* X, const double* Y, int n, double* g) {
  for (int p = 0; p < 1249; p++) g[p] = 0.0;
  double loss = 0.0;
  for (int s = 0; s < n; s++) {
    const double* x = X + 4 * s;
    double h1[32], h2[32], y;
    mlp_fwd(w, x, h1, h2, &y);
    const double e = y - Y[s];
    loss += e * e / n;
    const double dy = 2.0 * e / n;
    double d2[32], d1[32];
    for (int u = 0; u < 32; u++) d2[u] = h2[u] > 0.0 ? dy * w[1216 + u] : 0.0;
    for (int k = 0; k < 32; k++) {
      double a = 0.0;
      for (int u = 0; u < 32; u++) a += w[160 + k * 32 + u] * d2[u];
      d1[k] = h1[k] > 0.0 ? a : 0.0;
    }
    for (int k = 0; k < 4; k++)
      for (int u = 0; u < 32; u++) g[k * 32 + u] += d1[u] * x[k];
    for (int u = 0; u < 32; u++) g[128 + u] += d1[u];
    for (int k = 0; k < 32; k++)
      for (int u = 0; u < 32; u++) g[160 + k * 32 + u] += d2[u] * h1[k];
    for (int u = 0; u < 32; u++) g[1184 + u] += d2[u];
    for (int u = 0; u < 32; u++) g[1216 + u] += dy * h2[u];
    g[1248] += dy;
  }
  return loss;
}

// cuNSearch find_neighbors (SolverFraser.cu:72,171-175): radius 2 h_glob
// from the last output snapshot under the stale-position quirk.  Under
// WF_REBUILD_PLASTIC_ZONE (an extension) the radius is 2 h_glob + ns_skin,
// the positions are the current ones, and they are kept in xb.
int or_nsearch_build(or_state* s, const wf_config* c, int* over_cap) {
  const bool pz = c->rebuild_policy == WF_REBUILD_PLASTIC_ZONE;
  const double* src = (!pz && (c->quirks & WF_Q_STALE_NS_POS)) ? s->xs : s->x;
  const double radius = 2.0 * c->h_glob + (pz ? c->ns_skin : 0.0);
  int64_t tot = or_neighbors(src, s->n, radius, s->nb_count, s->nb_offset, s->nb_list,
                             s->nb_cap, over_cap);
  if (tot > s->nb_cap) return WF_E_NBCAP;
  if (pz && over_cap) {  // skin lists hold up to WF_MAX_NEIGHBORS_SKIN entries
    *over_cap = 0;
    for (int i = 0; i < s->n; i++) *over_cap += s->nb_count[i] > WF_MAX_NEIGHBORS_SKIN;
  }
  if (pz && s->xb)
    for (int64_t q = 0; q < 3 * (int64_t)s->n; q++) s->xb[q] = s->x[q];
  return WF_OK;
}

// Domain_d::ApplyAxiSymmBC, Domain_d.cu:665-689, over every particle: the
// projection of a_xy on the radial direction that it computes is then
// overwritten by a_x = a_y = 0, so only the condition matters (printf output
// omitted)
void or_apply_axisymm_bc(or_state* s) {
  for (int i = 0; i < s->n; i++) {
    const double ax = s->a[3 * i], ay = s->a[3 * i + 1];
    const double abs_a = sqrt(ax * ax + ay * ay);
    if (abs_a > 1.0e-3 && s->x[3 * i] > s->h[i] / 2.0 && ax > 1.0e-2) {
      s->a[3 * i] = 0.0;
      s->a[3 * i + 1] = 0.0;
    }
  }
}

// WF_REBUILD_PLASTIC_ZONE trigger: some particle with eps_p > ep_hot has
// moved more than ns_skin / 2 since the last build (squared distances).
int or_plastic_zone_moved(const or_state* s, const wf_config* c) {
  double d2max = 0.0;
  for (int i = 0; i < s->n; i++) {
    if (!(s->pl_strain[i] > c->ep_hot)) continue;
    const V3 d = sub(ld3(s->x, i), ld3(s->xb, i));
    const double d2 = dot(d, d);
    if (d2 > d2max) d2max = d2;
  }
  return 4.0 * d2max > c->ns_skin * c->ns_skin;
}

// One pass of the MechFraserSolve loop body, SolverFraser.cu:164-399.
// ---- rigid-mesh contact ----------------------------------------------------
namespace {
// vector_math.h:811-814
inline V3 cross(V3 a, V3 b) {
  return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
}  // namespace

// TriMesh_d::AxisPlaneMesh(2, false, p1, p2, dens), Mesh.cu:48-194: nodes row
// by row with the spacing dl = (p2-p1).x/dens accumulated in x and y,
// triangles [n0,n2,n1], [n1,n2,n3] per quad
int or_gen_plane_mesh(const double p1[3], const double p2[3], int dens, double* nodes,
                      int* elnode) {
  const double dl = (p2[0] - p1[0]) / dens;
  if (nodes) {
    int vi = 0;
    double x2 = p1[1];
    for (int j = 0; j < dens + 1; j++) {
      double x1 = p1[0];
      for (int i = 0; i < dens + 1; i++) {
        nodes[3 * vi] = x1;
        nodes[3 * vi + 1] = x2;
        nodes[3 * vi + 2] = p1[2];
        vi++;
        x1 += dl;
      }
      x2 += dl;
    }
  }
  if (elnode) {
    int el = 0;
    for (int j = 0; j < dens; j++)
      for (int i = 0; i < dens; i++) {
        int n0 = (dens + 1) * j + i, n1 = n0 + 1, n2 = (dens + 1) * (j + 1) + i, n3 = n2 + 1;
        int con[2][3] = {{n0, n2, n1}, {n1, n2, n3}};
        for (int e = 0; e < 2; e++, el++)
          for (int k = 0; k < 3; k++) elnode[3 * el + k] = con[e][k];
      }
  }
  return 2 * dens * dens;
}

// CalcCentroids, CalcNormals, UpdatePlaneCoeff (Mesh.cu:223-258); the plane
// constant from the element's first node (the reference reads node
// elnode[3e]+2, which lies in the same plane on a plate)
void or_mesh_geom(or_mesh* m) {
  for (int e = 0; e < m->ne; e++) {
    V3 a = ld3(m->x, m->el[3 * e]), b = ld3(m->x, m->el[3 * e + 1]),
       c = ld3(m->x, m->el[3 * e + 2]);
    V3 sum = add(add(a, b), c);  // (n0 + n1 + n2) / 3.
    st3(m->c, e, V3{sum.x / 3., sum.y / 3., sum.z / 3.});
    V3 w = cross(sub(b, a), sub(c, a));
    double l = length(w);
    V3 n{w.x / l, w.y / l, w.z / l};
    st3(m->nrm, e, n);
    m->pp[e] = dot(a, n);
  }
}

// TriMesh_d::Move (Mesh.cu:260-276) with the rotation about a centre
void or_mesh_move(or_mesh* m, double dt) {
  for (int q = 0; q < m->nn; q++) {
    const double* mo = m->motion + 9 * m->node_mesh[q];
    V3 vel{mo[0], mo[1], mo[2]}, w{mo[3], mo[4], mo[5]}, ctr{mo[6], mo[7], mo[8]};
    V3 v = add(vel, cross(w, sub(ld3(m->x, q), ctr)));
    st3(m->v, q, v);
    st3(m->x, q, add(ld3(m->x, q), scl(v, dt)));
  }
  // the rotation centre travels with the mesh velocity (extension: the
  // reference's TriMesh_d::Move rotates about the fixed origin)
  for (int k = 0; k < m->n_meshes; k++)
    for (int c = 0; c < 3; c++) m->motion[9 * k + 6 + c] += m->motion[9 * k + c] * dt;
  or_mesh_geom(m);
}

// CalculateSurface, Geometry.cu:26-88: normal_i = sum_j m_j x_ij, times
// N/(M nb_i) with N = particles + rigid triangles and M the total mass;
// |normal| >= 0.25 h and nb <= 46 marks the free surface unless the
// particle's ID was non-zero at the first pass (WeldFormGPU.cu:845-858)
void or_calc_surface(or_state* s, const wf_config* c) {
  or_contact* ct = s->contact;
  if (!ct->surf_init) {
    double tm = 0.0;
    for (int i = 0; i < s->n; i++) tm += s->m[i];  // WeldFormGPU.cu:808-814
    ct->totmass = tm;
    for (int i = 0; i < s->n; i++) ct->idlock[i] = s->ID[i] != 0;
    ct->surf_init = 1;
  }
  const double ntot = (double)s->n + (double)(ct->mesh ? ct->mesh->ne : 0);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->solid; i++) {
    V3 n{0.0, 0.0, 0.0};
    V3 xi = ld3(s->x, i);
    // with a list skin only the pairs within the search radius count
    const double rc2 = c->rebuild_policy == WF_REBUILD_PLASTIC_ZONE
                           ? 4.0 * c->h_glob * c->h_glob : 0.0;
    int cnt = 0;
    for (int k = 0; k < s->nb_count[i]; k++) {
      int j = s->nb_list[s->nb_offset[i] + k];
      const V3 d = sub(xi, ld3(s->x, j));
      if (rc2 > 0.0 && !(dot(d, d) < rc2)) continue;
      n = add(n, scl(d, s->m[j]));
      cnt++;
    }
    n = scl(n, ntot / (ct->totmass * (double)cnt));
    st3(ct->normal, i, n);
    if (length(n) >= 0.25 * s->h[i] && cnt <= 46 && !ct->idlock[i]) s->ID[i] = c->id_free_surf;
  }
}

// CalcContactNb, Contact.cu:29-62: free-surface particles against every
// triangle, centroid closer than 2 h_i, ascending element index, at most 20
void or_contact_nb(or_state* s, const wf_config* c) {
  or_contact* ct = s->contact;
  const or_mesh* m = ct->mesh;
  int over = 0;
  for (int i = 0; i < s->solid; i++) {
    int cnt = 0;
    if (s->ID[i] == c->id_free_surf) {
      V3 xi = ld3(s->x, i);
      bool ov = false;
      for (int e = 0; e < m->ne; e++) {
        if (!(length(sub(xi, ld3(m->c, e))) < 2. * s->h[i])) continue;
        if (cnt < WF_MAX_CONT_NB) ct->cnb_list[WF_MAX_CONT_NB * i + cnt++] = e;
        else ov = true;
      }
      over += ov;
    }
    ct->cnb_count[i] = cnt;
  }
  ct->cont_over_cap += over;
}

// CalcContactForcesWang, Contact.cu:134-321; in the sliding branch the
// dissipated power fric_heat x mu_d |F_n| |slip| / dt goes to q_friction
// (an extension).
void or_contact_forces(or_state* s, const wf_config* c) {
  or_contact* ct = s->contact;
  const or_mesh* m = ct->mesh;
  const double dt = c->dt;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < s->solid; i++) {
    V3 cf{0.0, 0.0, 0.0};
    double qf = 0.0;  // sliding-friction heat power [W] (an extension)
    const V3 xi = ld3(s->x, i), vi = ld3(s->v, i);
    V3 ai = ld3(s->a, i);
    const double mi = s->m[i], hi = s->h[i];
    for (int k = 0; k < ct->cnb_count[i]; k++) {
      const int e = ct->cnb_list[WF_MAX_CONT_NB * i + k];
      // x + v dt + a dt dt / 2. with the acceleration of earlier contacts
      const V3 t = scl(scl(ai, dt), dt);
      const V3 x_pred = add(add(xi, scl(vi, dt)), V3{t.x / 2., t.y / 2., t.z / 2.});
      const V3 n = ld3(m->nrm, e);
      const double dist = dot(n, x_pred) - m->pp[e];
      if (!(dist < hi)) continue;
      const V3 Qj = sub(xi, scl(n, dist));
      bool inside = true;
      for (int l = 0; l < 3 && inside; l++) {
        int nn = l + 1;
        if (nn > 2) nn = 0;
        V3 pl = ld3(m->x, m->el[3 * e + l]), pn = ld3(m->x, m->el[3 * e + nn]);
        if (dot(cross(sub(pn, pl), sub(Qj, pl)), n) < 0.0) inside = false;
      }
      if (!inside) continue;
      const double delta = hi - dist;
      double kij, fac;
      if (c->contact_law == WF_CONTACT_REFERENCE) {
        kij = 2.0 * mi / (dt * dt);
        fac = 0.4;
      } else {
        kij = c->contact_kappa * c->mat.E * c->contact_dx;
        fac = 1.0;
      }
      cf = scl(n, fac * kij * delta);
      ai = add(ai, V3{cf.x / mi, cf.y / mi, cf.z / mi});
      if (c->fric_sta > 0.) {
        V3 vv{0.0, 0.0, 0.0};
        for (int en = 0; en < 3; en++) vv = add(vv, ld3(m->v, m->el[3 * e + en]));
        V3 ve{vv.x / 3., vv.y / 3., vv.z / 3.};
        V3 du = sub(sub(x_pred, xi), scl(ve, dt));
        V3 dtg = sub(du, scl(n, dot(du, n)));
        V3 tg = scl(dtg, kij);
        const double ltg = length(tg), lcf = length(cf);
        if (ltg < c->fric_sta * lcf) {
          ai = sub(ai, V3{tg.x / mi, tg.y / mi, tg.z / mi});
        } else {
          V3 fd{c->fric_dyn * lcf * tg.x / ltg, c->fric_dyn * lcf * tg.y / ltg,
                c->fric_dyn * lcf * tg.z / ltg};
          cf = sub(cf, fd);
          ai = sub(ai, V3{fd.x / mi, fd.y / mi, fd.z / mi});
          // |F_friction| times the slip speed |dtg| / dt
          qf += c->fric_heat * (c->fric_dyn * lcf) * (ltg / kij) / dt;
        }
      }
    }
    st3(s->a, i, ai);
    st3(ct->cf, i, cf);
    if (s->q_friction) s->q_friction[i] = qf;
  }
}

int or_step(or_state* s, const wf_config* c, or_run* run, int nbc, const int* bc_zone,
            const double* bc_val) {
  if (c->mechanics) {
    const bool pz = c->rebuild_policy == WF_REBUILD_PLASTIC_ZONE;
    if (pz ? (run->is_yielding && run->rebuild_pending) : (run->ts_i == 0 && run->is_yielding)) {
      run->rebuild_pending = 0;
      int oc = 0;
      int rc = or_nsearch_build(s, c, &oc);
      if (rc) return rc;
      run->nb_over_cap = oc;
      run->ns_builds++;
    }
    or_calc_dens_inc(s, c);
    or_update_density(s, c->dt);
    or_calc_rate_tensors(s, c);
    or_pressure(s);
    or_stress_strain(s, c);
    or_calc_accel(s, c);
    if (c->axisymm) or_apply_axisymm_bc(s);  // SolverFraser.cu:229-231
    if (c->contact && s->contact) {  // SolverFraser.cu:233-257
      or_calc_surface(s, c);
      if (s->contact->mesh && s->contact->mesh->ne > 0) {
        if (run->ts_i == 0) or_contact_nb(s, c);
        or_contact_forces(s, c);
      }
    }
    if (!run->is_yielding) {
      for (int i = 0; i < s->n; i++)
        if (s->pl_strain[i] > run->max_pl_strain) run->max_pl_strain = s->pl_strain[i];
      if (run->max_pl_strain > c->ns_min_plastic) run->is_yielding = 1;
    }
    for (int b = 0; b < nbc; b++)
      or_apply_bc_vel(s, bc_zone[b], bc_val[3 * b], bc_val[3 * b + 1], bc_val[3 * b + 2]);
    or_update_pos_fraser(s, c->dt);
    or_update_vel(s, c->dt);
    for (int b = 0; b < nbc; b++)
      or_apply_bc_vel(s, bc_zone[b], bc_val[3 * b], bc_val[3 * b + 1], bc_val[3 * b + 2]);
    if (pz && run->is_yielding && s->xb) run->rebuild_pending = or_plastic_zone_moved(s, c);
  }
  if (c->thermal) or_thermal_calcs(s, c, &run->isfirst_step);
  if (c->mechanics) or_int_energy(s, c);
  if (c->mechanics && c->contact && s->contact && s->contact->mesh)  // SolverFraser.cu:290-318
    or_mesh_move(s->contact->mesh, c->dt);
  run->isfirst_step = 0;
  run->time += c->dt;
  run->step++;
  run->ts_i++;
  if (run->ts_i > (c->ns_interval - 1)) run->ts_i = 0;
  return WF_OK;
}

void or_set_threads(int n) { omp_set_num_threads(n > 0 ? n : 1); }
int or_get_threads(void) { return omp_get_max_threads(); }

}  // extern "C"
