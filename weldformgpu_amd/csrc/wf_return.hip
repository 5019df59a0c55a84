// wf_return.hip — batched Johnson-Cook radial return over the queue of
// yielding particles filled by k_stress.
//
// Each wavefront takes 64 queue entries: the surrogate MLP gives every entry
// an initial plastic increment (BF16 MFMA, wf_mlp.hpp), then an FP64
// safeguarded Newton solves s_trial - 3G dg - s_y(eps_p + dg) = 0 to the
// same residual as the plain return map (which starts from dg = 0 and is
// what runs when the surrogate is off or not yet trained).  The particle's
// update is then completed as in StressStrainOne (Mechanical.cu:504-551):
// deviatoric scale-back, sigma = -p I + S, eps_p, sigma_y, sigma_eq, plastic
// work heat, internal energy; and the solved sample is logged for training.
#include "wf_mlp.hpp"

namespace wf {

// Newton on [0, hi] from g0 with rt = rate factor x thermal factor
__device__ __forceinline__ double jc_newton_q(const MatParams& mp, double s_trial, double G,
                                              double ep, double rt, double hi, double g0,
                                              double tol, int max_iter, int* iters) {
  double lo = 0.0;
  if (!(hi > 0.0)) {
    *iters = 0;
    return 0.0;
  }
  double g = g0;
  if (!(g > lo) || !(g < hi)) g = (g0 >= hi) ? hi : lo;
  int it = 0;
  for (; it < max_iter; it++) {
    const double x = ep + g;
    const double xn = pow(x, mp.n);
    const double r = s_trial - 3.0 * G * g - (mp.A + mp.B * xn) * rt;
    if (fabs(r) <= tol * s_trial) break;
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

template <int REC>
__global__ __launch_bounds__(kBlock) void k_return(Fields f, StepParams p, BlockStats* bs,
                                                   RetQueue q, const float* __restrict__ w,
                                                   int use_mlp, long long step) {
  // block b solves the queue segment its stress-kernel block filled
  const int b = blockIdx.x;
  const int cnt = q.count[b];
  const int lane = lane_id();
  const int wv = threadIdx.x >> 6;
  const MatParams& mp = p.mat;
  unsigned long long ep_bits = 0ull, it_sum = 0ull, calls = 0ull, wmax = 0ull;
  if (wv * 64 < cnt) {  // wave-uniform: the whole wavefront takes part in the MFMA
    const int e = wv * 64 + lane;
    const bool valid = e < cnt;
    const int qi = b * kBlock + e;
    const int i = valid ? q.idx[qi] : 0;
    const double st = valid ? q.strial[qi] : 1.0;
    const double rt = valid ? q.rt[qi] : 1.0;
    const double gmax = valid ? q.gmax[qi] : 0.0;
    const double ep = valid ? f.pl_strain[i] : 0.0;
    const double rate = valid ? f.esr[i] : 0.0;
    const double theta = valid ? jc_theta(mp, f.T[i]) : 0.0;
    float f0, f1, f2, f3;
    mlp_features(st, mp.A, ep, rate, theta, f0, f1, f2, f3);
    double g0 = 0.0;
    if (use_mlp) {
      MlpWave mw;
      mlp_wave_load(w, mw);
      const float y = mlp_wave_eval(mw, f0, f1, f2, f3);
      if (isfinite(y)) g0 = fmin(fmax((double)y, 0.0), 1.0) * gmax;
    }
    int it = 0;
    if (valid) {
      const double G = f.G[i];
      const double g = jc_newton_q(mp, st, G, ep, rt, gmax, g0, p.newton_tol,
                                   p.newton_max_iter, &it);
      const double scale = (st - 3.0 * G * g) / st;
      double S[6];
#pragma unroll
      for (int k = 0; k < 6; k++) {
        S[k] = f.ss[k][i] * scale;
        f.ss[k][i] = S[k];
      }
      const double epn = ep + g;
      const double sigeq = st - 3.0 * G * g;
      f.pl_strain[i] = epn;
      f.sigma_y[i] = (mp.A + mp.B * pow(epn, mp.n)) * rt;
      f.sigma_eq[i] = sigeq;
      f.qpl[i] = p.taylor_quinney * sigeq * g / p.dt;
      const double pi = f.p[i];
      const double sg[6] = {-pi + S[0], -pi + S[1], -pi + S[2], S[3], S[4], S[5]};
#pragma unroll
      for (int k = 0; k < 6; k++) f.sigma[k][i] = sg[k];
      const double rho = f.rho[i];
      if (REC == kRec64) {
        const double ir2 = 1.0 / (rho * rho);
        f.arec64[rec_at(f, i, 1)] =
            make_double4(sg[0] * ir2, sg[1] * ir2, sg[2] * ir2, sg[3] * ir2);
        f.arec64[rec_at(f, i, 2)] = make_double4(sg[4] * ir2, sg[5] * ir2, rho, f.Cs[i]);
      }
      if (REC == kRec32) {
        const double ir2 = 1.0 / (rho * rho);
        f.arec[rec_at(f, i, 1)] = make_float4((float)(sg[0] * ir2), (float)(sg[1] * ir2),
                                                (float)(sg[2] * ir2), (float)(sg[3] * ir2));
        f.arec[rec_at(f, i, 2)] = make_float4((float)(sg[4] * ir2), (float)(sg[5] * ir2),
                                                (float)rho, (float)f.Cs[i]);
      }
      const double de = f.m[i] / rho *
                        (S[0] * f.strrate[0][i] + S[1] * f.strrate[1][i] +
                         S[2] * f.strrate[2][i] +
                         2.0 * (S[3] * f.strrate[3][i] + S[4] * f.strrate[4][i] +
                                S[5] * f.strrate[5][i]));
      f.eint[i] += de * p.dt;
      // training sample (one slot per particle: deterministic)
      f.mlx[0][i] = f0;
      f.mlx[1][i] = f1;
      f.mlx[2][i] = f2;
      f.mlx[3][i] = f3;
      f.mly[i] = gmax > 0.0 ? (float)(g / gmax) : 0.f;
      f.mlstep[i] = (int)step;
      ep_bits = (unsigned long long)__double_as_longlong(epn > 0.0 ? epn : 0.0);
      it_sum = (unsigned long long)it;
      calls = 1ull;
    }
    wmax = wave_max_u64((unsigned long long)it);
    ep_bits = wave_max_u64(ep_bits);
    it_sum = wave_sum_u64(it_sum);
    calls = wave_sum_u64(calls);
  }
  __shared__ unsigned long long wst[kBlock / 64][4];
  if (lane == 0) {
    wst[wv][0] = ep_bits;
    wst[wv][1] = it_sum;
    wst[wv][2] = calls;
    wst[wv][3] = wmax;
  }
  __syncthreads();
  if (threadIdx.x == 0 && cnt > 0) {
    BlockStats x = bs[b];
    for (int u = 0; u < kBlock / 64; u++) {
      x.ep_bits = wst[u][0] > x.ep_bits ? wst[u][0] : x.ep_bits;
      x.iters += wst[u][1];
      x.calls += wst[u][2];
      x.wave_iters += wst[u][3];
    }
    bs[b] = x;
  }
}

void launch_return(hipStream_t s, const Fields& f, const StepParams& p, BlockStats* bs,
                   const RetQueue& q, const float* mlp_w, bool use_mlp, int rec,
                   int64_t step) {
  const int blocks = (p.solid + kBlock - 1) / kBlock;
  if (blocks <= 0) return;
  const int um = use_mlp ? 1 : 0;
  if (rec == kRec32)
    hipLaunchKernelGGL(k_return<kRec32>, dim3(blocks), dim3(kBlock), 0, s, f, p, bs, q, mlp_w,
                       um, (long long)step);
  else if (rec == kRec64)
    hipLaunchKernelGGL(k_return<kRec64>, dim3(blocks), dim3(kBlock), 0, s, f, p, bs, q, mlp_w,
                       um, (long long)step);
  else
    hipLaunchKernelGGL(k_return<kRecNone>, dim3(blocks), dim3(kBlock), 0, s, f, p, bs, q, mlp_w,
                       um, (long long)step);
}

}  // namespace wf
