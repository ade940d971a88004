// sctl_amd_eval_noalloc: the dense Eval of sctl_amd_eval_ctx for callers
// that cannot allocate device memory (code under HIP graph capture, serving
// loops, hosts that forbid hipMalloc).  Its scratch -- the mixed path's
// Hilbert orders, packed source tiles, tile centres, sort workspace and split
// partial sums; the FP64 path's split partial sums -- comes from one
// file-scope device array, `scratch`, of NOALLOC_SCRATCH_BYTES per device.
//
// Every call on a device shares that device's array, which keeps the last
// call's contents: calls that overlap on one device (other streams, other
// host threads) race and must be serialised by the caller.  Nothing in the
// array has to start at zero: every byte a call reads was written earlier in
// the same call.  The split partial sums fit in what the rest leaves over:
// the splits run in groups of successive launches that each fit (down to one
// split per launch, which accumulates straight into v_trg).  Sizes whose
// orders, tiles and sort workspace alone exceed the array are rejected.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "internal.hpp"

namespace sctl_amd {

namespace {

constexpr size_t NOALLOC_SCRATCH_BYTES = (size_t)128 << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

// Device address of `scratch` on each device, looked up once per device
// (with that device current) by hipGetSymbolAddress: in host code `scratch`
// names a host-side stand-in, not device memory.
constexpr int MAX_DEV = 64;
std::mutex g_mu;
void* g_scratch[MAX_DEV] = {};

int scratch_ptr(const DeviceInfo& dev, char** out) {
  SCTL_CHECK_ARG(dev.device >= 0 && dev.device < MAX_DEV, "device index out of range");
  std::lock_guard<std::mutex> lock(g_mu);
  if (!g_scratch[dev.device]) {
    void* p = nullptr;
    SCTL_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    SCTL_CHECK_ARG(p, "no device address for the scratch array");
    g_scratch[dev.device] = p;
  }
  *out = static_cast<char*>(g_scratch[dev.device]);
  return SCTL_AMD_SUCCESS;
}

}  // namespace

}  // namespace sctl_amd

using namespace sctl_amd;

extern "C" {

size_t sctl_amd_eval_noalloc_scratch_bytes(void) { return NOALLOC_SCRATCH_BYTES; }

int sctl_amd_eval_noalloc(int kernel, const double* ctx, int precision, int digits, const double* r_trg, int64_t nt,
                          const double* r_src, const double* n_src, const double* v_src, int64_t ns, double* v_trg,
                          uint32_t flags, void* stream) {
  SCTL_TRY(eval_args(kernel, ctx, &precision, digits, r_trg, nt, r_src, n_src, v_src, ns, v_trg, flags,
                     SCTL_AMD_MIXED_NEWTON));
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  char* w = nullptr;
  SCTL_TRY(scratch_ptr(dev, &w));
  if (nt == 0 || ns == 0) return SCTL_AMD_SUCCESS;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  const int newton =
      precision == SCTL_AMD_FP64 ? newton_steps_f64(digits) : ((flags & SCTL_AMD_MIXED_NEWTON) ? 1 : 0);
  const DenseShape shape = pick_dense_shape(kernel, precision, newton, nt, ns);
  const double mu = eval_ctx_value(kernel, ctx);
  if (precision == SCTL_AMD_FP64) {
    // Split partial sums only; with the whole array as their budget.
    return f64_apply(kernel, newton, r_trg, nt, r_src, n_src, v_src, ns, v_trg, shape, mu,
                     reinterpret_cast<double*>(w), s, NOALLOC_SCRATCH_BYTES);
  }
  MixedGeometry g;
  SCTL_TRY(mixed_geometry(kernel, nt, ns, shape, &g));
  const size_t fixed = mixed_fixed_bytes(g, nt, ns);
  if (fixed > NOALLOC_SCRATCH_BYTES) {
    set_error("sctl_amd_eval_noalloc: " + std::to_string(nt) + " targets x " + std::to_string(ns) +
              " sources need " + std::to_string(fixed) + " bytes of scratch for the orders and source tiles, more " +
              "than its " + std::to_string(NOALLOC_SCRATCH_BYTES) + " bytes");
    return SCTL_AMD_ERROR_INVALID_VALUE;
  }
  SCTL_TRY(mixed_geometry(kernel, nt, ns, shape, &g, NOALLOC_SCRATCH_BYTES - fixed));
  return mixed_eval_in(kernel, newton, g, shape, r_trg, nt, r_src, n_src, v_src, ns, v_trg, mu, w, s);
}

}  // extern "C"
