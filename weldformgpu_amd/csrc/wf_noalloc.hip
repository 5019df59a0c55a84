// wf_noalloc.hip — wf_mlp_grad_noalloc: the surrogate's loss gradient
// (the computation of wf_mlp_grad) on caller-provided device arrays, for
// callers that cannot allocate device memory (HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).
//
// The gradient kernels of wf_mlp_grad write per-block partial gradients and
// loss partials to scratch memory and sum them in block order.  Here that
// scratch is a file-scope device array instead of a library allocation, so
// a call allocates nothing: the caller passes the parameters, samples and
// output arrays, all on the current device.
//
// Each GPU the library runs on holds its own copy of `scratch` (1.22 MiB)
// until the process ends.  Its address on a device is looked up once, with
// that device current, at the entry point's first call there, and kept; the
// kernels take that address as an argument (the array's name in host code
// is not a device address).  Every call on a device shares that device's
// array: calls that overlap on one GPU (two streams, or two host threads)
// race on it.  Nothing in it needs clearing between calls: a call's
// gradient blocks each write all of their partials before the reduction
// reads exactly those blocks.
#include <atomic>
#include <mutex>
#include <string>

#include "wf_internal.hpp"

namespace wf {

// partial gradients of kMlpMaxBlocks blocks, then their loss partials
constexpr size_t kScratchFloats = (size_t)kMlpMaxBlocks * kMlpParams + kMlpMaxBlocks;
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

constexpr int kMaxDevices = 64;
std::atomic<float*> g_scratch[kMaxDevices];
std::mutex g_scratch_mu;

// this device's address of `scratch` (looked up once per device)
int scratch_on_device(float** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess)
    return set_last_error(WF_E_HIP, (std::string("hipGetDevice: ") + hipGetErrorString(e)).c_str());
  if (dev < 0 || dev >= kMaxDevices)
    return set_last_error(WF_E_UNSUPPORTED, "device index beyond the scratch table");
  float* p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (!p) {
      void* q = nullptr;
      e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
      if (e != hipSuccess || !q)
        return set_last_error(
            WF_E_HIP, (std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e)).c_str());
      p = static_cast<float*>(q);
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return WF_OK;
}

}  // namespace
}  // namespace wf

extern "C" int wf_mlp_grad_noalloc(const float* params, const float* X, const float* Y, int n,
                                   float* grad, float* loss, void* stream) {
  using namespace wf;
  if (!params || !X || !Y || !grad || !loss)
    return set_last_error(WF_E_ARG, "wf_mlp_grad_noalloc: null device pointer");
  if (n <= 0 || n > WF_MLP_NOALLOC_MAX_SAMPLES)
    return set_last_error(WF_E_ARG,
                          "wf_mlp_grad_noalloc: sample count must be in 1..WF_MLP_NOALLOC_MAX_SAMPLES");
  float* sp = nullptr;
  int rc = scratch_on_device(&sp);
  if (rc) return rc;
  const int nblocks = mlp_grad_blocks(n);  // <= kMlpMaxBlocks: the partials fit
  float* part = sp;
  float* loss_part = sp + (size_t)kMlpMaxBlocks * kMlpParams;
  launch_mlp_grad(static_cast<hipStream_t>(stream), params, X, Y, n, 1, part, loss_part, nblocks,
                  grad, loss);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    return set_last_error(WF_E_HIP,
                          (std::string("wf_mlp_grad_noalloc launch: ") + hipGetErrorString(e)).c_str());
  return WF_OK;
}
