// wf_kernels.hip — FP64 per-particle kernels of the SPH pipeline (gfx950):
// pressure + Jaumann stress + return map + energy, BC + integration,
// leapfrog, reductions and reorder helpers.  One thread per particle over
// Morton-sorted SoA planes, so every load and store is coalesced.  The pair
// loops are in wf_pair.hip.
#include "wf_mlp.hpp"

namespace wf {

static inline int grid_for(int n) { return (n + kBlock - 1) / kBlock; }

__global__ __launch_bounds__(kBlock) void k_update_density(Fields f, int solid, double dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < solid) f.rho[i] += dt * f.drho[i];
}

// ---------------------------------------------------------------------------
// Pressure, PressureKernelExt + EOS case 0 (Mechanical.cu:289-298,
// Functions.cu:211-213) with PresEq = P0 = 0.
__global__ __launch_bounds__(kBlock) void k_pressure(Fields f, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) f.p[i] = 0.0 + (f.Cs[i] * f.Cs[i]) * (f.rho[i] - f.rho0[i]);
}

struct T9 {
  double xx, xy, xz, yx, yy, yz, zx, zy, zz;
};

// tensor3 operator*, tensor3.cu:241-255 as written (first column A*B, the
// others A*B^T) or the true product.
__device__ __forceinline__ T9 tmul(const T9& a, const T9& b, bool quirk) {
  T9 r;
  r.xx = a.xx * b.xx + a.xy * b.yx + a.xz * b.zx;
  r.yx = a.yx * b.xx + a.yy * b.yx + a.yz * b.zx;
  r.zx = a.zx * b.xx + a.zy * b.yx + a.zz * b.zx;
  if (quirk) {
    r.xy = a.xx * b.yx + a.xy * b.yy + a.xz * b.yz;
    r.xz = a.xx * b.zx + a.xy * b.zy + a.xz * b.zz;
    r.yy = a.yx * b.yx + a.yy * b.yy + a.yz * b.yz;
    r.yz = a.yx * b.zx + a.yy * b.zy + a.yz * b.zz;
    r.zy = a.zx * b.yx + a.zy * b.yy + a.zz * b.yz;
    r.zz = a.zx * b.zx + a.zy * b.zy + a.zz * b.zz;
  } else {
    r.xy = a.xx * b.xy + a.xy * b.yy + a.xz * b.zy;
    r.xz = a.xx * b.xz + a.xy * b.yz + a.xz * b.zz;
    r.yy = a.yx * b.xy + a.yy * b.yy + a.yz * b.zy;
    r.yz = a.yx * b.xz + a.yy * b.yz + a.yz * b.zz;
    r.zy = a.zx * b.xy + a.zy * b.yy + a.zz * b.zy;
    r.zz = a.zx * b.xz + a.zy * b.yz + a.zz * b.zz;
  }
  return r;
}

// Safeguarded Newton for the Johnson-Cook radial return,
// r(g) = s_trial - 3G g - (A + B (ep+g)^n) rt = 0 on [0, hi], from g0, with
// rt = rate factor x thermal factor and hi = (s_trial - s_y(ep)) / (3G).
// On convergence *xn holds (ep+g)^n of the returned g (the final yield
// stress needs it), else it is left alone.
__device__ __forceinline__ double jc_newton(const MatParams& mp, double s_trial, double G,
                                            double ep, double rt, double hi, double g0,
                                            double tol, int max_iter, int* iters, double* xn_out) {
  double lo = 0.0;
  if (!(hi > 0.0)) {
    *iters = 0;
    return 0.0;
  }
  double g = g0;
  if (!(g > lo) || !(g < hi)) g = (g0 >= hi) ? hi : lo;
  int it = 0;
  for (; it < max_iter; it++) {
    // one pow per iterate: s_y = (A + B x^n) rt, H = n B x^n / x rt
    const double x = ep + g;
    const double xn = pow(x, mp.n);
    const double r = s_trial - 3.0 * G * g - (mp.A + mp.B * xn) * rt;
    if (fabs(r) <= tol * s_trial) {
      *xn_out = xn;
      break;
    }
    if (r > 0.0) lo = g;
    else hi = g;
    const double H = x > 0.0 ? mp.n * mp.B * (xn / x) * rt : HUGE_VAL;
    double gn = g + r / (3.0 * G + H);
    if (!(gn > lo && gn < hi)) gn = 0.5 * (lo + hi);
    g = gn;
  }
  *iters = it;
  return g;
}

// ---------------------------------------------------------------------------
// C: Jaumann stress update + return map, StressStrainOne (Mechanical.cu:458-554),
// optionally with the pressure (FUSE_P), the per-particle internal-energy
// increment (ENERGY, Energy.cu:5-19 with per-particle indexing) and the
// packed FP32 or FP64 neighbour records of the pair loops (REC).
//
// The Johnson-Cook Newton return map (an extension: the reference re-evaluates
// the yield stress explicitly, Mechanical.cu:504-508) is solved in place, in
// three phases per wavefront: every lane forms its elastic predictor; if any
// lane yields, the whole wavefront evaluates the surrogate MLP on its 64
// feature rows (BF16 MFMA, wf_mlp.hpp) for an initial plastic increment and
// the yielding lanes run the FP64 Newton from it (from 0 without the
// surrogate) to the same residual; then every lane finishes and writes its
// state once.  The solved increment is logged as a training sample.
// S32: sigma, the deviatoric stress and the rates are stored in FP32
// (mixed precision); the arithmetic is FP64 either way.
template <bool FUSE_P, bool ENERGY, int REC, bool S32>
__global__ __launch_bounds__(kBlock) void k_stress(Fields f, StepParams p, BlockStats* bs,
                                                   RetArgs ra) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = i < p.solid;
  const MatParams& mp = p.mat;
  const double dt = p.dt;
  unsigned long long ep_bits = 0ull, nit = 0ull, ncall = 0ull;
  T9 S{}, D{};
  double G = 0.0, sig_trial = 0.0, sigy = 0.0, ep = 0.0, sigeq = 0.0, dep = 0.0, scale = 1.0;
  double esr = 0.0, T = 0.0, rt = 1.0, gmax = 0.0;
  // the particle's density, sound speed and pressure, read once (the
  // stores below could alias them, so the compiler would read them again)
  double rho_i = 0.0, cs_i = 0.0, p_i = 0.0;
  bool newton = false;  // a yielding Johnson-Cook particle under the Newton return map
  if (active) {
    const bool qp = (p.quirks & WF_Q_TENSOR_PRODUCT) != 0;
    rho_i = f.rho[i];
    cs_i = f.Cs[i];
    if (FUSE_P) {
      p_i = 0.0 + (cs_i * cs_i) * (rho_i - f.rho0[i]);
      f.p[i] = p_i;
    } else {
      p_i = f.p[i];
    }
    auto ss = [&](int k) { return S32 ? (double)f.ss32[k][i] : f.ss[k][i]; };
    auto dd = [&](int k) { return S32 ? (double)f.strrate32[k][i] : f.strrate[k][i]; };
    auto ww = [&](int k) { return S32 ? (double)f.rotrate32[k][i] : f.rotrate[k][i]; };
    S.xx = ss(0); S.yy = ss(1); S.zz = ss(2);
    S.xy = S.yx = ss(3); S.yz = S.zy = ss(4); S.xz = S.zx = ss(5);
    D.xx = dd(0); D.yy = dd(1); D.zz = dd(2);
    D.xy = D.yx = dd(3); D.yz = D.zy = dd(4); D.xz = D.zx = dd(5);
    T9 W;
    W.xx = W.yy = W.zz = 0.0;  // antisymmetric: the diagonal planes hold zeros
    W.xy = ww(3); W.yz = ww(4); W.xz = ww(5);
    W.yx = -W.xy; W.zy = -W.yz; W.zx = -W.xz;
    T9 WT = {W.xx, W.yx, W.zx, W.xy, W.yy, W.zy, W.xz, W.yz, W.zz};
    T9 SRT = tmul(S, WT, qp);
    T9 RS = tmul(W, S, qp);
    if (REC == kRec64)
      f.arec64[rec_at(f, i, 0)] = make_double4(f.v[0][i], f.v[1][i], f.v[2][i], f.m[i]);
    G = f.G[i];
    const double tr3 = 1.0 / 3.0 * (D.xx + D.yy + D.zz);
    const double g2 = 2.0 * G;
    // S = dt*(2G(D - tr/3 I) + SRT + RS) + S   (Mechanical.cu:490)
    S.xx = ((D.xx - tr3) * g2 + SRT.xx + RS.xx) * dt + S.xx;
    S.yy = ((D.yy - tr3) * g2 + SRT.yy + RS.yy) * dt + S.yy;
    S.zz = ((D.zz - tr3) * g2 + SRT.zz + RS.zz) * dt + S.zz;
    S.xy = (D.xy * g2 + SRT.xy + RS.xy) * dt + S.xy;
    S.yx = (D.yx * g2 + SRT.yx + RS.yx) * dt + S.yx;
    S.xz = (D.xz * g2 + SRT.xz + RS.xz) * dt + S.xz;
    S.zx = (D.zx * g2 + SRT.zx + RS.zx) * dt + S.zx;
    S.yz = (D.yz * g2 + SRT.yz + RS.yz) * dt + S.yz;
    S.zy = (D.zy * g2 + SRT.zy + RS.zy) * dt + S.zy;

    esr = sqrt(0.5 * ((D.xx - D.yy) * (D.xx - D.yy) + (D.yy - D.zz) * (D.yy - D.zz) +
                      (D.zz - D.xx) * (D.zz - D.xx)) +
               3.0 * (D.xy * D.xy + D.yz * D.yz + D.zx * D.zx));
    f.esr[i] = esr;
    const double J2 = 0.5 * (S.xx * S.xx + 2.0 * S.xy * S.yx + 2.0 * S.xz * S.zx +
                             S.yy * S.yy + 2.0 * S.yz * S.zy + S.zz * S.zz);
    sig_trial = sqrt(3.0 * J2);
    T = f.T[i];
    sigy = f.sigma_y[i];
    ep = f.pl_strain[i];
    sigeq = sig_trial;
    if (p.return_map == WF_RETURN_REFERENCE) {
      if (sigy < sig_trial) scale = sigy / sig_trial;
      if (mp.model == WF_MAT_HOLLOMON) sigy = hollomon_yield(mp, ep);
      else if (mp.model == WF_MAT_JOHNSON_COOK)
        sigy = jc_yield(mp, ep, jc_rate_factor(mp, esr), 1.0 - pow(jc_theta(mp, T), mp.m));
      if (sig_trial > sigy) {
        dep = (sig_trial - sigy) / (3. * G);
        ep += dep;
        sigeq = sigy;
        if (mp.model == WF_MAT_BILINEAR) sigy += dep * mp.Ep;
      }
    } else {
      double rate_f = 1.0, therm_f = 1.0;
      if (mp.model == WF_MAT_HOLLOMON) sigy = hollomon_yield(mp, ep);
      else if (mp.model == WF_MAT_JOHNSON_COOK) {
        rate_f = jc_rate_factor(mp, esr);
        therm_f = 1.0 - pow(jc_theta(mp, T), mp.m);
        sigy = jc_yield(mp, ep, rate_f, therm_f);
      }
      if (sig_trial > sigy) {
        double g = 0.0;
        if (mp.model == WF_MAT_BILINEAR) {
          g = (sig_trial - sigy) / (3.0 * G + mp.Ep);
          sigy = sigy + mp.Ep * g;
        } else if (mp.model == WF_MAT_JOHNSON_COOK) {
          newton = true;  // solved below, with the rest of the wavefront
          rt = rate_f * therm_f;
          gmax = (sig_trial - sigy) / (3.0 * G);
        } else {
          double lo = 0.0, hi = (sig_trial - sigy) / (3.0 * G);
          for (int it = 0; it < 200; it++) {
            double mid = 0.5 * (lo + hi);
            double r = sig_trial - 3.0 * G * mid - hollomon_yield(mp, ep + mid);
            if (r > 0) lo = mid;
            else hi = mid;
          }
          g = 0.5 * (lo + hi);
          sigy = hollomon_yield(mp, ep + g);
        }
        if (!newton) {
          scale = (sig_trial - 3.0 * G * g) / sig_trial;
          dep = g;
          ep += g;
          sigeq = sig_trial - 3.0 * G * g;
        }
      }
    }
  }
  // wave-uniform: the MFMA needs all 64 lanes
  if (__ballot(newton) != 0ull) {
    float f0, f1, f2, f3;
    mlp_features(newton ? sig_trial : 1.0, mp.A, ep, esr, newton ? jc_theta(mp, T) : 0.0, f0,
                 f1, f2, f3);
    double g0 = 0.0;
    if (ra.use_mlp) {
      MlpWave mw;
      mlp_wave_load_packed((const uint4*)ra.mlp_pack, mw);
      const float y = mlp_wave_eval(mw, f0, f1, f2, f3);
      if (isfinite(y)) g0 = fmin(fmax((double)y, 0.0), 1.0) * gmax;
    }
    if (newton) {
      int it = 0;
      double xn = -1.0;  // (ep+g)^n >= 0 once the Newton has converged
      const double g = jc_newton(mp, sig_trial, G, ep, rt, gmax, g0, p.newton_tol,
                                 p.newton_max_iter, &it, &xn);
      nit = (unsigned long long)it;
      ncall = 1ull;
      scale = (sig_trial - 3.0 * G * g) / sig_trial;
      dep = g;
      sigeq = sig_trial - 3.0 * G * g;
      // training sample (one slot per particle: deterministic)
      f.mlx[0][i] = f0;
      f.mlx[1][i] = f1;
      f.mlx[2][i] = f2;
      f.mlx[3][i] = f3;
      f.mly[i] = gmax > 0.0 ? (float)(g / gmax) : 0.f;
      f.mlstep[i] = (int)ra.step;
      ep += g;  // the same rounding as the Newton's ep + g
      sigy = (mp.A + mp.B * (xn >= 0.0 ? xn : pow(ep, mp.n))) * rt;
    }
  }
  if (active) {
    if (scale != 1.0) {
      S.xx *= scale; S.yy *= scale; S.zz *= scale;
      S.xy *= scale; S.yz *= scale; S.xz *= scale;
    }
    const double sv[6] = {S.xx, S.yy, S.zz, S.xy, S.yz, S.xz};
#pragma unroll
    for (int k = 0; k < 6; k++) {
      if (S32) f.ss32[k][i] = (float)sv[k];
      else f.ss[k][i] = sv[k];
    }
    f.strain[0][i] += D.xx * dt; f.strain[1][i] += D.yy * dt; f.strain[2][i] += D.zz * dt;
    f.strain[3][i] += D.xy * dt; f.strain[4][i] += D.yz * dt; f.strain[5][i] += D.xz * dt;
    f.sigma_y[i] = sigy;
    f.pl_strain[i] = ep;
    f.sigma_eq[i] = sigeq;
    f.qpl[i] = p.taylor_quinney * sigeq * dep / dt;
    const double pi = p_i;
    // sigma = -p I + S (Mechanical.cu:540); store upper triangle
    const double sg[6] = {-pi + S.xx, -pi + S.yy, -pi + S.zz, S.xy, S.yz, S.xz};
#pragma unroll
    for (int k = 0; k < 6; k++) {
      if (S32) f.sigma32[k][i] = (float)sg[k];
      else f.sigma[k][i] = sg[k];
    }
    if (REC == kRec32) {  // sigma/rho^2, rho and Cs for the FP32 acceleration loop
      const double rho = rho_i, ir2 = 1.0 / (rho * rho);
      f.arec[rec32_at(i, 0)] = make_float4((float)(sg[0] * ir2), (float)(sg[1] * ir2),
                                           (float)(sg[2] * ir2), (float)(sg[3] * ir2));
      f.arec[rec32_at(i, 1)] =
          make_float4((float)(sg[4] * ir2), (float)(sg[5] * ir2), (float)rho, (float)cs_i);
    }
    if (REC == kRec64) {  // the same in FP64 for the FP64 acceleration loop
      const double rho = rho_i, ir2 = 1.0 / (rho * rho);
      f.arec64[rec_at(f, i, 1)] =
          make_double4(sg[0] * ir2, sg[1] * ir2, sg[2] * ir2, sg[3] * ir2);
      f.arec64[rec_at(f, i, 2)] = make_double4(sg[4] * ir2, sg[5] * ir2, rho, cs_i);
    }
    if (ENERGY) {
      double de = f.m[i] / rho_i *
                  (S.xx * D.xx + S.yy * D.yy + S.zz * D.zz +
                   2.0 * (S.xy * D.xy + S.yz * D.yz + S.xz * D.xz));
      f.eint[i] += de * dt;
    }
    ep_bits = (unsigned long long)__double_as_longlong(ep > 0.0 ? ep : 0.0);
  }
  // block statistics
  __shared__ unsigned long long wst[kBlock / 64][4];
  const int lane = lane_id(), w = threadIdx.x >> 6;
  const unsigned long long mx = wave_max_u64(ep_bits);
  const unsigned long long sit = wave_sum_u64(nit);
  const unsigned long long sc = wave_sum_u64(ncall);
  const unsigned long long wmax = wave_max_u64(nit);
  if (lane == 0) {
    wst[w][0] = mx;
    wst[w][1] = sit;
    wst[w][2] = sc;
    wst[w][3] = wmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    BlockStats b{0ull, 0ull, 0ull, 0ull};
    for (int u = 0; u < kBlock / 64; u++) {
      b.ep_bits = wst[u][0] > b.ep_bits ? wst[u][0] : b.ep_bits;
      b.iters += wst[u][1];
      b.calls += wst[u][2];
      b.wave_iters += wst[u][3];
    }
    bs[blockIdx.x] = b;
  }
}

// Rebuild trigger of WF_REBUILD_PLASTIC_ZONE: the largest squared
// displacement from the last build over the particles with eps_p > ep_hot.
// A max of non-negative doubles is the max of their bit patterns, so the
// result does not depend on the order of the atomics.
__global__ __launch_bounds__(kBlock) void k_hot_disp(Fields f, const double* __restrict__ xb0,
                                                     const double* __restrict__ xb1,
                                                     const double* __restrict__ xb2, int n,
                                                     double ep_hot,
                                                     unsigned long long* d2_bits) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long b = 0ull;
  if (i < n && f.pl_strain[i] > ep_hot) {
    const double dx = f.x[0][i] - xb0[i], dy = f.x[1][i] - xb1[i], dz = f.x[2][i] - xb2[i];
    b = (unsigned long long)__double_as_longlong(dx * dx + dy * dy + dz * dz);
  }
  b = wave_max_u64(b);
  if (lane_id() == 0 && b) atomicMax(d2_bits, b);
}

// one block folds the per-block statistics into the run counters (used
// where no acceleration pass follows the stress pass to do it); one
// 256-thread block reading one entry at a time took 20 us for the 10.7k
// blocks of a 2.7M-particle step
constexpr int kFoldThreads = 1024;
__global__ __launch_bounds__(kFoldThreads) void k_fold_stats(const BlockStats* bs, int nb,
                                                             Counters* ctr) {
  __shared__ unsigned long long sh[4 * kFoldThreads / 64];
  block_fold_stats(StatsFold{bs, nb, ctr}, sh);
}

// ---------------------------------------------------------------------------
// F: velocity BCs + UpdatePosFraser + UpdateVel + BCs in one pass
// (SolverFraser.cu:273-281; Mechanical.cu:142-191,572-584)
template <int REC>
__global__ __launch_bounds__(kBlock) void k_integrate(Fields f, int n, double dt, BcTable bc,
                                                     Frame32 fr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // in place: each thread reads and writes only its own particle
  const IntegrateOut o{{f.x[0], f.x[1], f.x[2]}, f.p4, f.v4};
  integrate_particle<REC>(f, i, f.a[0][i], f.a[1][i], f.a[2][i], dt, bc, fr, o);
}

// every FP32 mirror from the FP64 state; the integer cell of each particle is
// (re)chosen here and kept until the next refresh
__global__ __launch_bounds__(kBlock) void k_mirror32(Fields f, int n, Frame32 fr) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int c[3];
  float xr[3];
  for (int a = 0; a < 3; a++) {
    double u = floor((f.x[a][i] - fr.org[a]) / fr.cs);
    int lim = a < 2 ? 2047 : 1023;
    c[a] = (int)fmin(fmax(u, 0.0), (double)lim);
    xr[a] = (float)(f.x[a][i] - (fr.org[a] + c[a] * fr.cs));
  }
  f.p4[i] = make_float4(xr[0], xr[1], xr[2], __int_as_float(pack_cell(c[0], c[1], c[2])));
  const double rho = f.rho[i];
  f.h32[i] = (float)f.h[i];
  const float4 v4 = make_float4((float)f.v[0][i], (float)f.v[1][i], (float)f.v[2][i],
                                (float)(f.m[i] / rho));
  f.v4[i] = v4;
  f.v4_alt[i] = v4;
  const double ir2 = 1.0 / (rho * rho);
  f.arec[rec32_at(i, 0)] =
      make_float4((float)(f.sigma32[0][i] * ir2), (float)(f.sigma32[1][i] * ir2),
                  (float)(f.sigma32[2][i] * ir2), (float)(f.sigma32[3][i] * ir2));
  f.arec[rec32_at(i, 1)] = make_float4((float)(f.sigma32[4][i] * ir2),
                                          (float)(f.sigma32[5][i] * ir2), (float)rho,
                                          (float)f.Cs[i]);
  const float4 tk = make_float4((float)f.T[i], (float)f.k[i], v4.w, 0.f);
  f.tk4[i] = tk;
  f.tk4_alt[i] = tk;
}

// every FP64 neighbour record from the FP64 state
__global__ __launch_bounds__(kBlock) void k_records64(Fields f, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double rho = f.rho[i], mdj = f.m[i] / rho, ir2 = 1.0 / (rho * rho);
  const double4 kv = make_double4(f.v[0][i], f.v[1][i], f.v[2][i], mdj);
  f.kv[i] = kv;
  f.kv_alt[i] = kv;
  f.arec64[rec_at(f, i, 0)] = make_double4(f.v[0][i], f.v[1][i], f.v[2][i], f.m[i]);
  f.arec64[rec_at(f, i, 1)] = make_double4(f.sigma[0][i] * ir2, f.sigma[1][i] * ir2,
                                           f.sigma[2][i] * ir2, f.sigma[3][i] * ir2);
  f.arec64[rec_at(f, i, 2)] = make_double4(f.sigma[4][i] * ir2, f.sigma[5][i] * ir2, rho, f.Cs[i]);
  const double4 tk = make_double4(f.T[i], f.k[i], mdj, 0.0);
  f.tk64[i] = tk;
  f.tk64_alt[i] = tk;
}

__global__ __launch_bounds__(kBlock) void k_apply_bc(Fields f, int n, int zone, double vx,
                                                     double vy, double vz) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || f.ID[i] != zone) return;
  f.a[0][i] = 0.0; f.a[1][i] = 0.0; f.a[2][i] = 0.0;
  f.v[0][i] = vx; f.v[1][i] = vy; f.v[2][i] = vz;
  f.va[0][i] = vx; f.va[1][i] = vy; f.va[2][i] = vz;
}

// ApplyAxiSymmBC (Domain_d.cu:665-689): the radial projection computed there
// is overwritten by zeros, so a particle meeting the condition loses its
// in-plane acceleration
__global__ __launch_bounds__(kBlock) void k_axisymm_bc(Fields f, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ax = f.a[0][i], ay = f.a[1][i];
  if (sqrt(ax * ax + ay * ay) > 1.0e-3 && f.x[0][i] > f.h[i] / 2.0 && ax > 1.0e-2) {
    f.a[0][i] = 0.0;
    f.a[1][i] = 0.0;
  }
}

__global__ __launch_bounds__(kBlock) void k_update_pos(Fields f, int n, double dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 3; c++) {
    double u_ = f.v[c][i] * dt + ((f.a[c][i] * 0.5) * dt) * dt;
    f.x[c][i] += u_;
    f.u[c][i] += u_;
  }
}

__global__ __launch_bounds__(kBlock) void k_update_vel(Fields f, int n, double dt) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < 3; c++) f.v[c][i] += f.a[c][i] * dt;
}

__global__ __launch_bounds__(kBlock) void k_leapfrog(Fields f, int n, double dt, int first) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (first) {
    f.Ta[i] = f.T[i] - dt / 2.0 * f.dTdt[i];
  } else {
    f.Tb[i] = f.Ta[i];
    f.Ta[i] += f.dTdt[i] * dt;
    f.T[i] = (f.Ta[i] + f.Tb[i]) / 2.;
  }
}

// Deterministic FP64 sum: fixed-shape block partials, then one block.
__global__ __launch_bounds__(kBlock) void k_sum_partial(const double* v, int n, double* partial) {
  __shared__ double sh[kBlock / 64];
  double acc = 0.0;
  for (int i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) acc += v[i];
  acc = wave_sum_f64(acc);
  if (lane_id() == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kBlock / 64; w++) t += sh[w];
    partial[blockIdx.x] = t;
  }
}
__global__ void k_sum_final(const double* partial, int nb, double* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    double t = 0.0;
    for (int b = 0; b < nb; b++) t += partial[b];
    *out = t;
  }
}

// ---- generic reorder helpers ----------------------------------------------
__global__ void k_gather_f64(double* dst, const double* src, const int* idx, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_gather_i32(int* dst, const int* src, const int* idx, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}
__global__ void k_scatter_f64(double* dst, const double* src, const int* idx, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}
__global__ void k_scatter_i32(int* dst, const int* src, const int* idx, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[idx[i]] = src[i];
}
__global__ void k_fill_f64(double* dst, double v, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = v;
}

// ---- launchers --------------------------------------------------------------
void launch_update_density(hipStream_t s, const Fields& f, const StepParams& p, double dt) {
  if (p.solid <= 0) return;
  hipLaunchKernelGGL(k_update_density, dim3(grid_for(p.solid)), dim3(kBlock), 0, s, f, p.solid,
                     dt);
}
void launch_pressure(hipStream_t s, const Fields& f, const StepParams& p) {
  if (p.n <= 0) return;
  hipLaunchKernelGGL(k_pressure, dim3(grid_for(p.n)), dim3(kBlock), 0, s, f, p.n);
}
template <bool S32>
static void launch_stress_t(hipStream_t s, const Fields& f, const StepParams& p,
                            BlockStats* bs, bool fuse_pressure, bool energy, int rec,
                            const RetArgs& ra) {
  dim3 g(grid_for(p.solid)), b(kBlock);
  if (rec == kRec32)
    hipLaunchKernelGGL((k_stress<true, true, kRec32, true>), g, b, 0, s, f, p, bs, ra);
  else if (rec == kRec64)
    hipLaunchKernelGGL((k_stress<true, true, kRec64, false>), g, b, 0, s, f, p, bs, ra);
  else if (fuse_pressure && energy)
    hipLaunchKernelGGL((k_stress<true, true, kRecNone, S32>), g, b, 0, s, f, p, bs, ra);
  else if (fuse_pressure)
    hipLaunchKernelGGL((k_stress<true, false, kRecNone, S32>), g, b, 0, s, f, p, bs, ra);
  else if (energy)
    hipLaunchKernelGGL((k_stress<false, true, kRecNone, S32>), g, b, 0, s, f, p, bs, ra);
  else
    hipLaunchKernelGGL((k_stress<false, false, kRecNone, S32>), g, b, 0, s, f, p, bs, ra);
}
void launch_stress(hipStream_t s, const Fields& f, const StepParams& p, BlockStats* bs,
                   bool fuse_pressure, bool energy, int rec, bool s32, const RetArgs& ra) {
  if (p.solid <= 0) return;
  if (s32) launch_stress_t<true>(s, f, p, bs, fuse_pressure, energy, rec, ra);
  else launch_stress_t<false>(s, f, p, bs, fuse_pressure, energy, rec, ra);
}
void launch_hot_disp(hipStream_t s, const Fields& f, const double* const xb[3], int n,
                     double ep_hot, unsigned long long* d2_bits) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_hot_disp, dim3(grid_for(n)), dim3(kBlock), 0, s, f, xb[0], xb[1], xb[2],
                     n, ep_hot, d2_bits);
}
void launch_fold_stats(hipStream_t s, const BlockStats* bs, int nblocks, Counters* ctr) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(k_fold_stats, dim3(1), dim3(kFoldThreads), 0, s, bs, nblocks, ctr);
}
void launch_apply_bc(hipStream_t s, const Fields& f, int n, int zone, double vx, double vy,
                     double vz) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_apply_bc, dim3(grid_for(n)), dim3(kBlock), 0, s, f, n, zone, vx, vy, vz);
}
void launch_axisymm_bc(hipStream_t s, const Fields& f, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_axisymm_bc, dim3(grid_for(n)), dim3(kBlock), 0, s, f, n);
}
void launch_update_pos(hipStream_t s, const Fields& f, int n, double dt) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_update_pos, dim3(grid_for(n)), dim3(kBlock), 0, s, f, n, dt);
}
void launch_update_vel(hipStream_t s, const Fields& f, int n, double dt) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_update_vel, dim3(grid_for(n)), dim3(kBlock), 0, s, f, n, dt);
}
void launch_integrate(hipStream_t s, const Fields& f, int n, double dt, const BcTable& bc,
                      int rec, const Frame32& fr) {
  if (n <= 0) return;
  const dim3 g(grid_for(n)), b(kBlock);
  if (rec == kRec32)
    hipLaunchKernelGGL(k_integrate<kRec32>, g, b, 0, s, f, n, dt, bc, fr);
  else if (rec == kRec64)
    hipLaunchKernelGGL(k_integrate<kRec64>, g, b, 0, s, f, n, dt, bc, fr);
  else
    hipLaunchKernelGGL(k_integrate<kRecNone>, g, b, 0, s, f, n, dt, bc, fr);
}
void launch_records64(hipStream_t s, const Fields& f, int n) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_records64, dim3(grid_for(n)), dim3(kBlock), 0, s, f, n);
}
void launch_mirror32(hipStream_t s, const Fields& f, int n, const Frame32& fr) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mirror32, dim3(grid_for(n)), dim3(kBlock), 0, s, f, n, fr);
}
void launch_leapfrog(hipStream_t s, const Fields& f, int n, double dt, bool first) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_leapfrog, dim3(grid_for(n)), dim3(kBlock), 0, s, f, n, dt, first ? 1 : 0);
}
void launch_sum(hipStream_t s, const double* v, int n, double* partial, int nblocks,
                double* out) {
  hipLaunchKernelGGL(k_sum_partial, dim3(nblocks), dim3(kBlock), 0, s, v, n, partial);
  hipLaunchKernelGGL(k_sum_final, dim3(1), dim3(64), 0, s, partial, nblocks, out);
}
void launch_gather_f64(hipStream_t s, double* dst, const double* src, const int* idx, int n) {
  if (n > 0) hipLaunchKernelGGL(k_gather_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, idx, n);
}
void launch_gather_i32(hipStream_t s, int* dst, const int* src, const int* idx, int n) {
  if (n > 0) hipLaunchKernelGGL(k_gather_i32, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, idx, n);
}
void launch_scatter_f64(hipStream_t s, double* dst, const double* src, const int* idx, int n) {
  if (n > 0) hipLaunchKernelGGL(k_scatter_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, idx, n);
}
void launch_scatter_i32(hipStream_t s, int* dst, const int* src, const int* idx, int n) {
  if (n > 0) hipLaunchKernelGGL(k_scatter_i32, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, src, idx, n);
}
void launch_fill_f64(hipStream_t s, double* dst, double v, int n) {
  if (n > 0) hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(n)), dim3(kBlock), 0, s, dst, v, n);
}

}  // namespace wf
