// wf_mlp.hip — online training of the return-map surrogate (wf_mlp.hpp):
// deterministic FP32 forward/backward over logged samples and Adam
// (Kingma & Ba 2015), plus the inference test hook and sample collection.
//
// Gradients are summed in a fixed order: a block owns a contiguous range of
// samples, stages 64 samples' activations and back-propagated errors in LDS,
// and each thread sums its parameters over those samples in sample order;
// block partials are then added in block order by the Adam kernel.  The
// same inputs therefore always give bit-identical weights.
#include "wf_mlp.hpp"

namespace wf {

constexpr int kGradThreads = 256;
constexpr int kGradChunk = 64;
constexpr int kParamsPerThread = (kMlpParams + kGradThreads - 1) / kGradThreads;

// the per-lane operands of the 64 lanes of a wavefront (one wavefront)
__global__ __launch_bounds__(64) void k_mlp_pack(const float* __restrict__ w,
                                                 uint4* __restrict__ pack) {
  MlpWave mw;
  mlp_wave_load(w, mw);
  MlpLanePack lp;
  lp.a1 = mw.a1;
#pragma unroll
  for (int kb = 0; kb < 4; kb++) lp.a2[kb] = mw.a2[kb];
  lp.b3 = mw.b3;
  lp.pad = 0.f;
#pragma unroll
  for (int r = 0; r < 16; r++) {
    lp.b2r[r] = mw.b2r[r];
    lp.w3r[r] = mw.w3r[r];
  }
  uint4 q[kMlpPackWords];
  __builtin_memcpy(q, &lp, sizeof(lp));
  const int l = lane_id();
#pragma unroll
  for (int k = 0; k < kMlpPackWords; k++) pack[64 * k + l] = q[k];
}

// MFMA inference on explicit rows (n x 4): the test hook of mlp_wave_eval,
// loading the weights as the stress kernel does
__global__ __launch_bounds__(256) void k_mlp_infer(const uint4* __restrict__ pack,
                                                   const float* __restrict__ X, int n,
                                                   float* __restrict__ y) {
  MlpWave mw;
  mlp_wave_load_packed(pack, mw);
  const int lane = lane_id();
  const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int nwaves = (gridDim.x * blockDim.x) >> 6;
  for (int base = wave * 64; base < n; base += nwaves * 64) {
    const int r = base + lane;
    const bool valid = r < n;
    const float f0 = valid ? X[4 * r] : 0.f, f1 = valid ? X[4 * r + 1] : 0.f;
    const float f2 = valid ? X[4 * r + 2] : 0.f, f3 = valid ? X[4 * r + 3] : 0.f;
    const float out = mlp_wave_eval(mw, f0, f1, f2, f3);
    if (valid) y[r] = out;
  }
}

// gradient of L = 1/n sum (y_s - t_s)^2 over samples s*stride, s < n
__global__ __launch_bounds__(kGradThreads) void k_mlp_grad(const float* __restrict__ w,
                                                           const float* __restrict__ X,
                                                           const float* __restrict__ Y, int n,
                                                           int stride, int per_block,
                                                           float* __restrict__ part,
                                                           float* __restrict__ loss_part) {
  __shared__ float sw[kMlpParams];
  __shared__ float sx[kGradChunk][4];
  __shared__ float sh1[kGradChunk][32], sh2[kGradChunk][32];
  __shared__ float sd1[kGradChunk][32], sd2[kGradChunk][32];
  __shared__ float sdy[kGradChunk];
  __shared__ float sloss[kGradChunk];
  const int tid = threadIdx.x;
  for (int k = tid; k < kMlpParams; k += kGradThreads) sw[k] = w[k];
  float acc[kParamsPerThread];
#pragma unroll
  for (int u = 0; u < kParamsPerThread; u++) acc[u] = 0.f;
  float loss = 0.f;
  const int s0 = blockIdx.x * per_block;
  const int s1 = min(n, s0 + per_block);
  const float inv_n = 1.0f / (float)n;
  __syncthreads();
  for (int c0 = s0; c0 < s1; c0 += kGradChunk) {
    const int m = min(kGradChunk, s1 - c0);
    if (tid < m) {
      const int s = (c0 + tid) * stride;
      float x[4];
      for (int k = 0; k < 4; k++) x[k] = X[4 * (int64_t)s + k];
      const float t = Y[s];
      float h1[32], h2[32];
      for (int u = 0; u < 32; u++) {
        float a = sw[kMlpOffB1 + u];
        for (int k = 0; k < 4; k++) a += sw[kMlpOffW1 + k * 32 + u] * x[k];
        h1[u] = fmaxf(a, 0.f);
      }
      float y = sw[kMlpOffB3];
      for (int u = 0; u < 32; u++) {
        float a = sw[kMlpOffB2 + u];
        for (int k = 0; k < 32; k++) a += sw[kMlpOffW2 + k * 32 + u] * h1[k];
        h2[u] = fmaxf(a, 0.f);
        y += sw[kMlpOffW3 + u] * h2[u];
      }
      const float e = y - t;
      const float dy = 2.f * e * inv_n;
      float d2[32];
      for (int u = 0; u < 32; u++) d2[u] = h2[u] > 0.f ? dy * sw[kMlpOffW3 + u] : 0.f;
      for (int k = 0; k < 32; k++) {
        float a = 0.f;
        for (int u = 0; u < 32; u++) a += sw[kMlpOffW2 + k * 32 + u] * d2[u];
        sd1[tid][k] = h1[k] > 0.f ? a : 0.f;
        sh1[tid][k] = h1[k];
        sh2[tid][k] = h2[k];
        sd2[tid][k] = d2[k];
      }
      for (int k = 0; k < 4; k++) sx[tid][k] = x[k];
      sdy[tid] = dy;
      sloss[tid] = e * e * inv_n;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < kParamsPerThread; u++) {
      const int pidx = tid + u * kGradThreads;
      if (pidx >= kMlpParams) break;
      float a = 0.f;
      if (pidx < kMlpOffB1) {
        const int k = pidx / 32, o = pidx % 32;
        for (int s = 0; s < m; s++) a += sd1[s][o] * sx[s][k];
      } else if (pidx < kMlpOffW2) {
        const int o = pidx - kMlpOffB1;
        for (int s = 0; s < m; s++) a += sd1[s][o];
      } else if (pidx < kMlpOffB2) {
        const int k = (pidx - kMlpOffW2) / 32, o = (pidx - kMlpOffW2) % 32;
        for (int s = 0; s < m; s++) a += sd2[s][o] * sh1[s][k];
      } else if (pidx < kMlpOffW3) {
        const int o = pidx - kMlpOffB2;
        for (int s = 0; s < m; s++) a += sd2[s][o];
      } else if (pidx < kMlpOffB3) {
        const int o = pidx - kMlpOffW3;
        for (int s = 0; s < m; s++) a += sdy[s] * sh2[s][o];
      } else {
        for (int s = 0; s < m; s++) a += sdy[s];
      }
      acc[u] += a;
    }
    if (tid == 0)
      for (int s = 0; s < m; s++) loss += sloss[s];
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < kParamsPerThread; u++) {
    const int pidx = tid + u * kGradThreads;
    if (pidx < kMlpParams) part[(int64_t)blockIdx.x * kMlpParams + pidx] = acc[u];
  }
  if (tid == 0) loss_part[blockIdx.x] = loss;
}

// sum block partials in block order
__global__ void k_mlp_reduce(const float* __restrict__ part, const float* __restrict__ loss_part,
                             int nblocks, float* __restrict__ grad, float* __restrict__ loss) {
  const int pidx = blockIdx.x * blockDim.x + threadIdx.x;
  if (pidx < kMlpParams) {
    float a = 0.f;
    for (int b = 0; b < nblocks; b++) a += part[(int64_t)b * kMlpParams + pidx];
    grad[pidx] = a;
  }
  if (pidx == 0) {
    float l = 0.f;
    for (int b = 0; b < nblocks; b++) l += loss_part[b];
    *loss = l;
  }
}

// one training iteration's second half: the block partials summed in block
// order (as k_mlp_reduce) and the Adam update of the same parameter (Kingma &
// Ba 2015, Algorithm 1: beta1 0.9, beta2 0.999, eps 1e-8)
__global__ __launch_bounds__(256) void k_mlp_reduce_adam(const float* __restrict__ part,
                                                         const float* __restrict__ loss_part,
                                                         int nblocks, float* __restrict__ w,
                                                         float* __restrict__ m,
                                                         float* __restrict__ v, int t, float lr,
                                                         float* __restrict__ loss) {
  // one wavefront per parameter (four per block); lane l sums partials
  // l, l+64, ... in order, then a fixed butterfly adds the lanes, so the
  // sum does not depend on timing
  const int pidx = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (pidx == 0) {  // the loss, the same way
    float l = 0.f;
    for (int b = lane; b < nblocks; b += 64) l += loss_part[b];
    for (int o = 32; o > 0; o >>= 1) l += __shfl_xor(l, o);
    if (lane == 0) *loss = l;
  }
  if (pidx >= kMlpParams) return;
  float g = 0.f;
  for (int b = lane; b < nblocks; b += 64) g += part[(int64_t)b * kMlpParams + pidx];
  for (int o = 32; o > 0; o >>= 1) g += __shfl_xor(g, o);
  if (lane != 0) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float mn = b1 * m[pidx] + (1.f - b1) * g;
  const float vn = b2 * v[pidx] + (1.f - b2) * g * g;
  m[pidx] = mn;
  v[pidx] = vn;
  const float mhat = mn / (1.f - powf(b1, (float)t));
  const float vhat = vn / (1.f - powf(b2, (float)t));
  w[pidx] -= lr * mhat / (sqrtf(vhat) + eps);
}


// sample collection: flags, exclusive scan (caller), scatter
__global__ void k_mlp_flags(const int* __restrict__ mlstep, int n, int* __restrict__ flags) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) flags[i] = mlstep[i] >= 0 ? 1 : 0;
}
__global__ void k_mlp_scatter(Fields f, int n, const int* __restrict__ pos, float* __restrict__ X,
                              float* __restrict__ Y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n || f.mlstep[i] < 0) return;
  const int s = pos[i];
  for (int k = 0; k < 4; k++) X[4 * (int64_t)s + k] = f.mlx[k][i];
  Y[s] = f.mly[i];
}

void launch_mlp_pack(hipStream_t s, const float* w, void* pack) {
  hipLaunchKernelGGL(k_mlp_pack, dim3(1), dim3(64), 0, s, w, (uint4*)pack);
}

void launch_mlp_infer(hipStream_t s, const void* pack, const float* X, int n, float* y) {
  if (n <= 0) return;
  const int blocks = std::min(1024, (n + 255) / 256);
  hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, s, (const uint4*)pack, X, n, y);
}

void launch_mlp_grad(hipStream_t s, const float* w, const float* X, const float* Y, int n,
                     int stride, float* part, float* loss_part, int nblocks, float* grad,
                     float* loss) {
  const int per_block = (n + nblocks - 1) / nblocks;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(kGradThreads), 0, s, w, X, Y, n, stride,
                     per_block, part, loss_part);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((kMlpParams + 255) / 256), dim3(256), 0, s, part,
                     loss_part, nblocks, grad, loss);
}

void launch_mlp_train_iter(hipStream_t s, float* w, float* m, float* v, const float* X,
                           const float* Y, int n, int stride, float* part, float* loss_part,
                           int nblocks, int t, float lr, float* loss) {
  const int per_block = (n + nblocks - 1) / nblocks;
  hipLaunchKernelGGL(k_mlp_grad, dim3(nblocks), dim3(kGradThreads), 0, s, w, X, Y, n, stride,
                     per_block, part, loss_part);
  hipLaunchKernelGGL(k_mlp_reduce_adam, dim3((kMlpParams + 3) / 4), dim3(256), 0, s, part,
                     loss_part, nblocks, w, m, v, t, lr, loss);
}
int mlp_grad_blocks(int n) { return std::max(1, std::min(kMlpMaxBlocks, (n + kGradChunk - 1) / kGradChunk)); }


void launch_mlp_collect(hipStream_t s, const Fields& f, int n, float* X, float* Y, int* flags,
                        int* scan, int* scan_part, int* count) {
  const int g = (n + 255) / 256;
  hipLaunchKernelGGL(k_mlp_flags, dim3(g), dim3(256), 0, s, f.mlstep, n, flags);
  launch_scan_exclusive(s, flags, n, scan_part, scan);
  hipLaunchKernelGGL(k_mlp_scatter, dim3(g), dim3(256), 0, s, f, n, scan, X, Y);
  hipMemcpyAsync(count, scan + n, sizeof(int), hipMemcpyDeviceToDevice, s);
}

}  // namespace wf
