// Allocation-free form of the batched, HBM-resident Redheffer product
// (rcwa_star_product_batched_noalloc, include/rcwa_amd.h).  It has no
// counterpart in the reference; it serves callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops, hosts
// that forbid hipMalloc).  Its work slots, Gauss-Jordan workspace and
// singularity flags live in the file-scope device array `scratch` below:
// every call on a GPU shares that GPU's copy, which keeps the last call's
// contents, so calls that overlap on one GPU race (one call at a time per
// device, or one context's stream).  The array's device address is taken
// once per device with hipGetSymbolAddress, with that device current, and
// only that address is handed to kernels and copies.
#include <mutex>

#include "rcwa_amd.h"
#include "common.hpp"
#include "kernels.hpp"
#include "runtime.hpp"

namespace {

// Limits: n <= 1024, batch <= 64 and n * n * batch <= 2^19 complex values per
// work slot (8 MiB); five slots, the Gauss-Jordan workspace (at most 12 MiB
// for these limits: batch * (2 n 64 * 16 + 4 n) + 256 bytes <= 5760 * 2052 +
// 256), and 64 flags.
constexpr int kMaxN = 1024;
constexpr int kMaxBatch = 64;
constexpr size_t kSlotElems = size_t(1) << 19;
constexpr size_t kSlotBytes = kSlotElems * sizeof(zc);
constexpr size_t kGjBytes = size_t(12) << 20;
constexpr size_t kInfoOff = 5 * kSlotBytes + kGjBytes;
constexpr size_t kScratchBytes = kInfoOff + kMaxBatch * sizeof(int);
constexpr int kMaxDevices = 64;

std::mutex g_mu;
char* g_ptr[kMaxDevices] = {};

__global__ void clear_flags_kernel(int* p, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) p[i] = 0;
}

}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

// The scratch array's address on `device` (current), taken on first use.
rcwa_status scratch_of(int device, char** out) {
  if (device < 0 || device >= kMaxDevices)
    return rcwa_set_error(RCWA_E_ARG, "device index beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_ptr[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return rcwa_set_error(RCWA_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") +
                                            hipGetErrorString(e));
    g_ptr[device] = (char*)p;
  }
  *out = g_ptr[device];
  return RCWA_OK;
}

}  // namespace

extern "C" rcwa_status rcwa_star_product_batched_noalloc(rcwa_ctx* ctx, int n, int batch,
                                                         const rcwa_c128* const* A_dev,
                                                         const rcwa_c128* const* B_dev,
                                                         rcwa_c128* const* S_dev) {
  if (!ctx || n <= 0 || batch <= 0 || !A_dev || !B_dev || !S_dev)
    return rcwa_set_error(RCWA_E_ARG, "bad argument");
  for (int k = 0; k < 4; ++k)
    if (!A_dev[k] || !B_dev[k] || !S_dev[k]) return rcwa_set_error(RCWA_E_ARG, "null block");
  if (n > kMaxN || batch > kMaxBatch || (size_t)n * n * batch > kSlotElems ||
      gj_workspace_bytes(n, batch) > kGjBytes)
    return rcwa_set_error(RCWA_E_ARG,
                          "rcwa_star_product_batched_noalloc: n <= 1024, batch <= 64 and "
                          "n * n * batch <= 524288 required (its fixed scratch is " +
                              std::to_string(kScratchBytes >> 20) + " MiB)");
  Ctx& c = *reinterpret_cast<Ctx*>(ctx);
  hipError_t e = hipSetDevice(c.device);
  if (e != hipSuccess) return rcwa_set_error(RCWA_E_HIP, hipGetErrorString(e));
  char* base = nullptr;
  if (rcwa_status st = scratch_of(c.device, &base)) return st;
  // FP64 products only: the split form would allocate its slice workspace
  c.precision = RCWA_FP64;
  zc* t[5];
  for (int k = 0; k < 5; ++k) t[k] = (zc*)(base + k * kSlotBytes);
  void* gj = base + 5 * kSlotBytes;
  int* info = (int*)(base + kInfoOff);
  hipLaunchKernelGGL(clear_flags_kernel, dim3(1), dim3(kMaxBatch), 0, c.stream, info, batch);
  if ((e = hipGetLastError()) != hipSuccess)
    return rcwa_set_error(RCWA_E_HIP, hipGetErrorString(e));
  if (rcwa_status st = star_product_device(c, n, batch, A_dev, B_dev, S_dev, t, gj, info))
    return st;
  int h[kMaxBatch];
  if ((e = hipMemcpyAsync(h, info, batch * sizeof(int), hipMemcpyDeviceToHost, c.stream)) !=
          hipSuccess ||
      (e = hipStreamSynchronize(c.stream)) != hipSuccess)
    return rcwa_set_error(RCWA_E_HIP, hipGetErrorString(e));
  for (int b = 0; b < batch; ++b)
    if (h[b]) return rcwa_set_error(RCWA_E_SINGULAR, "exactly singular matrix inverted");
  return RCWA_OK;
}
