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

}  // namespace wf
