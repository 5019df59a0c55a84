// wf_device.hpp — device helpers: cubic-spline kernel, material laws,
// wavefront reductions.  FP64 unless the name says otherwise.
#pragma once
#include <hip/hip_runtime.h>

#include "wf_internal.hpp"

namespace wf {

constexpr double kPi = 3.14159265358979323846;

// Cubic spline gradient kernel (1/r) dW/dr, 3D: Functions.cu:76-88
__device__ __forceinline__ double grad_kernel(double q, double h) {
  double C = 1.0 / (h * h * h * h * kPi);
  if (q == 0.0) return C / h * (-3.0 + (9.0 / 2.0) * q);
  if (q < 1.0) return C / (q * h) * (-3.0 * q + (9.0 / 4.0) * q * q);
  if (q < 2.0) return C / (q * h) * ((-3.0 / 4.0) * (2.0 - q) * (2.0 - q));
  return 0.0;
}

// Johnson-Cook flow stress, Material.cuh:209-219, homologous temperature
// clamped to [0,1] (the reference's pow() of a negative base is NaN).
__device__ __forceinline__ double jc_theta(const MatParams& mp, double T) {
  double t = (T - mp.Tt) / (mp.Tm - mp.Tt);
  if (!(t > 0.0)) t = 0.0;
  if (t > 1.0) t = 1.0;
  return t;
}
__device__ __forceinline__ double jc_rate_factor(const MatParams& mp, double sr) {
  if (sr == 0.0) sr = 1.e-5;
  return 1.0 + mp.C * log(sr / mp.eps0);
}
__device__ __forceinline__ double jc_yield(const MatParams& mp, double ep, double rate_f,
                                           double therm_f) {
  return (mp.A + mp.B * pow(ep, mp.n)) * rate_f * therm_f;
}
// d sigma_y / d eps_p
__device__ __forceinline__ double jc_hardening(const MatParams& mp, double ep, double rate_f,
                                               double therm_f) {
  if (!(ep > 0.0)) return HUGE_VAL;
  return mp.n * mp.B * pow(ep, mp.n - 1.) * rate_f * therm_f;
}
// Hollomon, Material.cuh:192-201 with eps0 = sy0/E, eps1 = (sy0/K)^(1/m)
__device__ __forceinline__ double hollomon_yield(const MatParams& mp, double ep) {
  if (ep + mp.eps0_h > mp.eps1_h) return mp.K_h * pow(ep + mp.eps0_h, mp.m_h);
  return mp.sy0;
}

// packed integer cell of the FP32 frame: cx | cy << 11 | cz << 22 (a grid
// of at most kPackedCells cells, wf_tiles.hpp)
__device__ __forceinline__ int pack_cell(int cx, int cy, int cz) {
  return cx | (cy << 11) | (cz << 22);
}
__device__ __forceinline__ void unpack_cell(int c, int& cx, int& cy, int& cz) {
  cx = c & 2047;
  cy = (c >> 11) & 2047;
  cz = (c >> 22) & 1023;
}

// ---- integration of one particle -------------------------------------------
// The state of particle i that integrate_particle reads (integrate_load);
// the acceleration pass that integrates its own particles loads it before
// its pair loop, so that the loads are in flight while the pairs are summed.
struct IntegrateIn {
  double x[3], v[3];
  int id;
  float cell, vw;  // the FP32 mirrors' kept cell and m/rho (kRec32)
};
template <int REC>
__device__ __forceinline__ void integrate_load(const Fields& f, int i, IntegrateIn& in) {
  in.id = f.ID[i];
  for (int c = 0; c < 3; c++) {
    in.x[c] = f.x[c][i];
    in.v[c] = f.v[c][i];
  }
  if (REC == kRec32) {
    in.cell = f.p4[i].w;
    in.vw = f.v4[i].w;
  }
}

// BC, position and velocity update of particle i from its acceleration a:
// ApplyBCVelKernel (Mechanical.cu:572-589) before and after
// UpdatePosFraserKernel (:179-191, u = v dt + a dt^2/2) and UpdateVelKernel
// (:142-152), plus the record the next passes read (REC), from the state in
template <int REC>
__device__ __forceinline__ void integrate_apply(const Fields& f, int i, double a0, double a1,
                                                double a2, double dt, const BcTable& bc,
                                                const Frame32& fr, const IntegrateOut& o,
                                                const IntegrateIn& in) {
  // no FMA contraction: the separate integration pass and the acceleration
  // pass that integrates its own particles round identically (and as the
  // FP64 oracle does)
#pragma clang fp contract(off)
  double v0 = in.v[0], v1 = in.v[1], v2 = in.v[2];
  int hit = -1;
  for (int b = 0; b < bc.count; b++)
    if (in.id == bc.zone[b]) hit = b;
  if (hit >= 0) {
    a0 = a1 = a2 = 0.0;
    v0 = bc.val[hit][0];
    v1 = bc.val[hit][1];
    v2 = bc.val[hit][2];
  }
  const double u0 = v0 * dt + ((a0 * 0.5) * dt) * dt;
  const double u1 = v1 * dt + ((a1 * 0.5) * dt) * dt;
  const double u2 = v2 * dt + ((a2 * 0.5) * dt) * dt;
  const double x0 = in.x[0] + u0, x1 = in.x[1] + u1, x2 = in.x[2] + u2;
  o.x[0][i] = x0;
  o.x[1][i] = x1;
  o.x[2][i] = x2;
  f.u[0][i] += u0;
  f.u[1][i] += u1;
  f.u[2][i] += u2;
  v0 += a0 * dt;
  v1 += a1 * dt;
  v2 += a2 * dt;
  if (hit >= 0) {
    a0 = a1 = a2 = 0.0;
    v0 = bc.val[hit][0];
    v1 = bc.val[hit][1];
    v2 = bc.val[hit][2];
    f.va[0][i] = v0;
    f.va[1][i] = v1;
    f.va[2][i] = v2;
  }
  f.v[0][i] = v0;
  f.v[1][i] = v1;
  f.v[2][i] = v2;
  f.a[0][i] = a0;
  f.a[1][i] = a1;
  f.a[2][i] = a2;
  if (REC == kRec64) f.kv[i] = make_double4(v0, v1, v2, f.kv[i].w);
  if (REC == kRec32) {
    int cx, cy, cz;
    unpack_cell(__float_as_int(in.cell), cx, cy, cz);  // the cell kept since the last refresh
    o.p4[i] = make_float4((float)(x0 - (fr.org[0] + cx * fr.cs)),
                          (float)(x1 - (fr.org[1] + cy * fr.cs)),
                          (float)(x2 - (fr.org[2] + cz * fr.cs)), in.cell);
    o.v4[i] = make_float4((float)v0, (float)v1, (float)v2, in.vw);
  }
}

template <int REC>
__device__ __forceinline__ void integrate_particle(const Fields& f, int i, double a0, double a1,
                                                   double a2, double dt, const BcTable& bc,
                                                   const Frame32& fr, const IntegrateOut& o) {
  IntegrateIn in;
  integrate_load<REC>(f, i, in);
  integrate_apply<REC>(f, i, a0, a1, a2, dt, bc, fr, o, in);
}

// ---- wavefront (64-lane) reductions ---------------------------------------
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    unsigned long long o = __shfl_xor(v, off, 64);
    v = o > v ? o : v;
  }
  return v;
}
__device__ __forceinline__ unsigned long long wave_sum_u64(unsigned long long v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ int lane_id() { return threadIdx.x & 63; }

// Fold a stress pass's per-block statistics into the run counters with the
// whole workgroup (blockDim.x a multiple of 64, at most 1,024): four entries
// in flight per thread, then wavefront and block reductions.  sh: LDS for
// 4 x blockDim.x / 64 values, free for the call.  Integer max and sums, so
// the result does not depend on the order.
__device__ __forceinline__ void block_fold_stats(const StatsFold& fd, unsigned long long* sh) {
  const int nt = blockDim.x, nw = nt >> 6;
  unsigned long long m = 0ull, it = 0ull, ca = 0ull, wi = 0ull;
#pragma unroll 4
  for (int b = threadIdx.x; b < fd.nb; b += nt) {
    const BlockStats x = fd.bs[b];
    m = x.ep_bits > m ? x.ep_bits : m;
    it += x.iters;
    ca += x.calls;
    wi += x.wave_iters;
  }
  m = wave_max_u64(m);
  it = wave_sum_u64(it);
  ca = wave_sum_u64(ca);
  wi = wave_sum_u64(wi);
  const int w = threadIdx.x >> 6;
  if (lane_id() == 0) {
    sh[w] = m;
    sh[nw + w] = it;
    sh[2 * nw + w] = ca;
    sh[3 * nw + w] = wi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int u = 1; u < nw; u++) {
      m = sh[u] > m ? sh[u] : m;
      it += sh[nw + u];
      ca += sh[2 * nw + u];
      wi += sh[3 * nw + u];
    }
    Counters* c = fd.ctr;
    if (m > c->max_pl_strain_bits) c->max_pl_strain_bits = m;
    c->newton_iters += it;
    c->newton_calls += ca;
    c->newton_iters_wave += wi;
  }
}

}  // namespace wf
