// The learned network's training gradient without a caller workspace, for
// callers that cannot allocate device memory (code under HIP graph capture,
// serving loops, harnesses that forbid hipMalloc).  No counterpart in the
// reference.
//
// paliwa_learned_gradient sums per-workgroup partial gradients that it keeps
// in a workspace the caller allocates.  Here they live in `scratch`, a static
// device array of the library sized for PALIWA_LEARNED_STATIC_MAX_ROWS rows;
// every GPU the library runs on holds its own copy until the process ends.
// The array's name is only a host-side handle: its device address on the
// current GPU comes from hipGetSymbolAddress, taken once per device on the
// first call there and kept.  Each call writes every partial sum it reads, so
// nothing has to be cleared between calls; calls that overlap on one GPU share
// the array and race.
#include "common.hpp"

#include <mutex>
#include <string>

namespace paliwa_hip {
// learned.hip
int64_t learned_gradient_floats(int32_t count);
hipError_t launch_learned_gradient(float const *pairs, int32_t stride,
                                   int32_t const *index, int32_t count,
                                   float const *weights, float *grad,
                                   float *loss_sum, float *partials,
                                   hipStream_t stream);
} // namespace paliwa_hip

namespace {

using namespace paliwa_hip;

// One partial row of 1378 floats (1377 parameters and the loss) per 64 rows.
constexpr int64_t kScratchFloats = (PALIWA_LEARNED_STATIC_MAX_ROWS / 64) * 1378;
constexpr int kMaxDevices = 64;

std::mutex g_mutex;
float *g_scratch[kMaxDevices] = {};

} // namespace

static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

// The device address of `scratch` on the current GPU (set up once per device).
int scratch_on_current_device(float **out) {
  int device = 0;
  hipError_t err = hipGetDevice(&device);
  if (err != hipSuccess)
    return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
  if (device < 0 || device >= kMaxDevices)
    return fail(PALIWA_ERR_UNSUPPORTED, "device index beyond 63");
  std::lock_guard<std::mutex> lock(g_mutex);
  if (g_scratch[device] == nullptr) {
    void *ptr = nullptr;
    err = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (err != hipSuccess || ptr == nullptr)
      return fail(PALIWA_ERR_HIP, std::string("HIP: hipGetSymbolAddress: ") +
                                      hipGetErrorString(err));
    g_scratch[device] = static_cast<float *>(ptr);
  }
  *out = g_scratch[device];
  return PALIWA_OK;
}

} // namespace

extern "C" {

int paliwa_learned_gradient_static(const float *pairs, int32_t stride,
                                   const int32_t *index, int32_t count,
                                   const float *weights, float *grad,
                                   float *loss_sum, void *stream) {
  if (pairs == nullptr || weights == nullptr || grad == nullptr ||
      loss_sum == nullptr || count < 1 || stride < 9)
    return fail(PALIWA_ERR_INVALID_ARGUMENT, "bad arguments");
  if (count > PALIWA_LEARNED_STATIC_MAX_ROWS || learned_gradient_floats(count) > kScratchFloats)
    return fail(PALIWA_ERR_UNSUPPORTED,
                "at most " + std::to_string(PALIWA_LEARNED_STATIC_MAX_ROWS) +
                    " rows per call without a workspace (paliwa_learned_gradient takes "
                    "more)");
  float *partials = nullptr;
  if (int rc = scratch_on_current_device(&partials); rc != PALIWA_OK)
    return rc;
  hipError_t const err =
      launch_learned_gradient(pairs, stride, index, count, weights, grad, loss_sum, partials,
                              static_cast<hipStream_t>(stream));
  if (err != hipSuccess)
    return fail(PALIWA_ERR_HIP, std::string("HIP: ") + hipGetErrorString(err));
  return PALIWA_OK;
}

} // extern "C"
