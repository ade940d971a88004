// C-ABI entry points (include/sctl_amd.h): argument checking, host staging,
// precision/Newton selection, and dispatch to the HIP launchers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"

namespace sctl_amd {

namespace {
thread_local std::string g_error;

struct KernelMeta {
  int K0, K1, ND, flops;
  const char* name;
};
// GenericKernel dims (generic-kernel.txx:48-62) and FLOPS()/Name()
// (kernel_functions.hpp:33-39,62-67,82-88,126-131,181-186,221-227,265-271,
// 326-332,367-373,393-399,417-423,452-458,492-498,530-536,574-580,628-634).
constexpr KernelMeta META[NUM_KERNELS] = {
    {1, 1, 0, 6, "Laplace3D-FxU"},      {1, 1, 3, 14, "Laplace3D-DxU"},    {3, 3, 0, 23, "Stokes3D-FxU"},
    {3, 3, 3, 26, "Stokes3D-DxU"},      {1, 3, 0, 11, "Laplace3D-FxdU"},   {1, 3, 3, 22, "Laplace3D-DxdU"},
    {1, 9, 0, 25, "Laplace3D-Fxd2U"},   {3, 9, 0, 39, "Stokes3D-FxT"},     {4, 3, 0, 26, "Stokes3D-FSxU"},
    {3, 4, 0, 26, "Stokes3D-FxUP"},     {3, 3, 0, 15, "BiotSavart3D-FxU"}, {3, 9, 0, 51, "BiotSavart3D-FxdU"},
    {2, 2, 0, 20, "Helmholtz3D-FxU"},   {2, 2, 3, 29, "Helmholtz3D-DxU"},  {2, 6, 0, 35, "Helmholtz3D-FxdU"},
    {2, 6, 0, 41, "HelmholtzDiff3D-FxdU"},
};
static_assert(Ker<LAP_FXD2U>::K1 == 9 && Ker<STK_FSXU>::K0 == 4 && Ker<LAP_DXDU>::ND == 3, "tables disagree");

// const_pi<double>() of math_utils.hpp:43, so the scale factors round as the
// reference's uKerScaleFactor<double>() does.
double const_pi() {
  return (double)14488038916154245684ull / (double)(1ull << 62) +
         (double)7089564414062235240ull / (double)(1ull << 62) / (double)(1ull << 63);
}

std::mutex g_dev_mutex;
constexpr int MAX_DEV = 64;
DeviceInfo g_dev[MAX_DEV];

}  // namespace

void set_error(const std::string& msg) { g_error = msg; }
const char* last_error() { return g_error.c_str(); }

int kernel_dims(int ker, int* K0, int* K1, int* ND) {
  if (ker < 0 || ker >= NUM_KERNELS) return SCTL_AMD_ERROR_UNKNOWN_KERNEL;
  *K0 = META[ker].K0;
  *K1 = META[ker].K1;
  *ND = META[ker].ND;
  return SCTL_AMD_SUCCESS;
}

// uKerScaleFactor<double>() (kernel_functions.hpp:40,69,89,133,188,228,272,
// 333,374,400,424,459,499,537,581,635).
double kernel_scale(int ker) {
  switch (ker) {
    case LAP_FXU:
    case LAP_DXU:
    case LAP_DXDU:
    case LAP_FXD2U:
    case BS_FXU:
    case HLM_FXU:
    case HLM_FXDU:
    case HLMD_FXDU: return 1 / (4 * const_pi());
    case STK_FXU:
    case STK_FSXU:
    case STK_FXUP: return 1 / (8 * const_pi());
    case STK_DXU: return 3 / (4 * const_pi());
    case LAP_FXDU:
    case BS_FXDU:
    case HLM_DXU: return -1 / (4 * const_pi());
    case STK_FXT: return -3 / (4 * const_pi());
  }
  return 0;
}

bool kernel_needs_ctx(int ker) { return needs_ctx(ker); }

int kernel_max_tpl(int ker) {
  if (ker < 0 || ker >= NUM_KERNELS) return 1;
  return with_kernel(ker, [](auto kc) { return max_tpl<decltype(kc)::value>(); });
}

// v_rsq_f64 is good to about 2^-24 (7 digits); each Newton step doubles the
// bits.  digits follows the reference's knob: -1 or >= 16 means machine
// precision (generic-kernel.txx:97,103).
int newton_steps_f64(int digits) {
  const int d = (digits < 0 || digits >= 16) ? 15 : digits;
  return d <= 7 ? 0 : (d <= 14 ? 1 : 2);
}

int device_info(DeviceInfo* out) {
  // The current device's cached facts; the device count (a slow call) is
  // asked only the first time a device is seen.
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0) {
    set_error("no HIP device visible");
    return SCTL_AMD_ERROR_NO_DEVICE;
  }
  if (dev >= MAX_DEV) {
    set_error("device index out of range");
    return SCTL_AMD_ERROR_NO_DEVICE;
  }
  std::lock_guard<std::mutex> lock(g_dev_mutex);
  DeviceInfo& d = g_dev[dev];
  if (!d.ok) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
      set_error("no HIP device visible");
      return SCTL_AMD_ERROR_NO_DEVICE;
    }
    hipDeviceProp_t p;
    SCTL_HIP_TRY(hipGetDeviceProperties(&p, dev));
    d.device = dev;
    d.num_cus = p.multiProcessorCount;
    d.max_threads = p.maxThreadsPerBlock;
    // The library's own stream-ordered pool on this device (the process-wide
    // default pool keeps its settings): freed scratch stays reserved for
    // reuse up to POOL_KEEP_BYTES, since returning it to the driver and
    // reserving it again costs tens of microseconds per call.
    hipMemPoolProps props = {};
    props.allocType = hipMemAllocationTypePinned;
    props.handleTypes = hipMemHandleTypeNone;
    props.location.type = hipMemLocationTypeDevice;
    props.location.id = dev;
    SCTL_HIP_TRY(hipMemPoolCreate(&d.pool, &props));
    uint64_t threshold = POOL_KEEP_BYTES;
    SCTL_HIP_TRY(hipMemPoolSetAttribute(d.pool, hipMemPoolAttrReleaseThreshold, &threshold));
    d.ok = true;
  }
  *out = d;
  return SCTL_AMD_SUCCESS;
}

namespace {

hipStream_t as_stream(void* s) { return s ? static_cast<hipStream_t>(s) : hipStreamPerThread; }

// Copies `n` doubles from host to a stream-ordered device buffer.
int stage_in(Scratch& buf, const double* host, int64_t n, hipStream_t s, const double** dev) {
  if (!host || n <= 0) {
    *dev = nullptr;
    return SCTL_AMD_SUCCESS;
  }
  SCTL_TRY(buf.alloc(sizeof(double) * n, s));
  SCTL_HIP_TRY(hipMemcpyAsync(buf.as<double>(), host, sizeof(double) * n, hipMemcpyHostToDevice, s));
  *dev = buf.as<double>();
  return SCTL_AMD_SUCCESS;
}

// Pinned host staging for large device-to-host results (a host KernelMatrix):
// chunks of PIN_CHUNK bytes taken from a process-wide pool of page-locked
// buffers (allocated with hipHostMalloc on first use, reused, never returned
// to the driver), two per call so that the copy of one chunk into the
// caller's memory overlaps the DMA of the next.
constexpr size_t PIN_CHUNK = (size_t)8 << 20;
std::mutex g_pin_mutex;
std::vector<void*> g_pin_free;

int pin_take(void** p) {
  {
    std::lock_guard<std::mutex> lock(g_pin_mutex);
    if (!g_pin_free.empty()) {
      *p = g_pin_free.back();
      g_pin_free.pop_back();
      return SCTL_AMD_SUCCESS;
    }
  }
  SCTL_HIP_TRY(hipHostMalloc(p, PIN_CHUNK, hipHostMallocDefault));
  return SCTL_AMD_SUCCESS;
}
void pin_give(void* p) {
  if (!p) return;
  std::lock_guard<std::mutex> lock(g_pin_mutex);
  g_pin_free.push_back(p);
}

// host[0, bytes) = dev[0, bytes), through two pinned chunks on stream s
// (synchronous on return).
int copy_to_host_pinned(void* host, const void* dev, size_t bytes, hipStream_t s) {
  void* pin[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  int rc = SCTL_AMD_SUCCESS;
  for (int i = 0; i < 2 && rc == SCTL_AMD_SUCCESS; i++) {
    rc = pin_take(&pin[i]);
    if (rc == SCTL_AMD_SUCCESS && hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess) {
      set_error("hipEventCreateWithFlags failed");
      rc = SCTL_AMD_ERROR_HIP;
    }
  }
  size_t prev_off = 0, prev_len = 0;
  for (size_t off = 0, i = 0; rc == SCTL_AMD_SUCCESS && off < bytes + PIN_CHUNK; off += PIN_CHUNK, i++) {
    const size_t len = off < bytes ? std::min(PIN_CHUNK, bytes - off) : 0;
    if (len) {
      hipError_t e = hipMemcpyAsync(pin[i % 2], static_cast<const char*>(dev) + off, len, hipMemcpyDeviceToHost, s);
      if (e == hipSuccess) e = hipEventRecord(ev[i % 2], s);
      if (e != hipSuccess) {
        set_error(std::string("pinned staging copy: ") + hipGetErrorString(e));
        rc = SCTL_AMD_ERROR_HIP;
        break;
      }
    }
    if (prev_len) {  // the previous chunk: wait for its DMA, copy it out
      const hipError_t e = hipEventSynchronize(ev[(i + 1) % 2]);
      if (e != hipSuccess) {
        set_error(std::string("pinned staging wait: ") + hipGetErrorString(e));
        rc = SCTL_AMD_ERROR_HIP;
        break;
      }
      std::memcpy(static_cast<char*>(host) + prev_off, pin[(i + 1) % 2], prev_len);
    }
    prev_off = off;
    prev_len = len;
    if (!len) break;
  }
  if (rc != SCTL_AMD_SUCCESS) (void)hipStreamSynchronize(s);
  for (int i = 0; i < 2; i++) {
    if (ev[i]) (void)hipEventDestroy(ev[i]);
    pin_give(pin[i]);
  }
  return rc;
}

int check_common(int kernel, const double* r_trg, int64_t nt, const double* r_src, const double* n_src, int64_t ns,
                 int digits) {
  if (kernel < 0 || kernel >= NUM_KERNELS) {
    set_error("unknown kernel id " + std::to_string(kernel));
    return SCTL_AMD_ERROR_UNKNOWN_KERNEL;
  }
  SCTL_CHECK_ARG(nt >= 0 && ns >= 0, "negative target or source count");
  SCTL_CHECK_ARG(digits >= -1, "digits must be -1 or >= 0");
  SCTL_CHECK_ARG(nt == 0 || r_trg, "r_trg is NULL");
  SCTL_CHECK_ARG(ns == 0 || r_src, "r_src is NULL");
  SCTL_CHECK_ARG(ns == 0 || META[kernel].ND == 0 || n_src, "n_src is NULL for a kernel with a normal");
  return SCTL_AMD_SUCCESS;
}

// The Helmholtz kernels read their wavenumber through ctx (a pointer to one
// double, as the reference's ctx_ptr points to one Real); the others ignore it.
int check_ctx(int kernel, const double* ctx) {
  if (!needs_ctx(kernel)) return SCTL_AMD_SUCCESS;
  SCTL_CHECK_ARG(ctx, std::string(META[kernel].name) + " needs its wavenumber (ctx is NULL)");
  SCTL_CHECK_ARG(std::isfinite(*ctx), "the wavenumber is not finite");
  return SCTL_AMD_SUCCESS;
}
double ctx_value(int kernel, const double* ctx) { return needs_ctx(kernel) && ctx ? *ctx : 0.0; }

}  // namespace

int eval_args(int kernel, const double* ctx, int* precision, int digits, const double* r_trg, int64_t nt,
              const double* r_src, const double* n_src, const double* v_src, int64_t ns, const double* v_trg,
              uint32_t flags, uint32_t allowed_flags) {
  SCTL_TRY(check_common(kernel, r_trg, nt, r_src, n_src, ns, digits));
  SCTL_TRY(check_ctx(kernel, ctx));
  SCTL_CHECK_ARG(ns == 0 || v_src, "v_src is NULL");
  SCTL_CHECK_ARG(nt == 0 || v_trg, "v_trg is NULL");
  SCTL_CHECK_ARG((flags & ~allowed_flags) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(*precision >= SCTL_AMD_FP64 && *precision <= SCTL_AMD_AUTO, "unknown precision");
  if (*precision == SCTL_AMD_AUTO) *precision = (digits >= 0 && digits <= 7) ? SCTL_AMD_MIXED : SCTL_AMD_FP64;
  return SCTL_AMD_SUCCESS;
}

double eval_ctx_value(int kernel, const double* ctx) { return ctx_value(kernel, ctx); }

}  // namespace sctl_amd

using namespace sctl_amd;

extern "C" {

const char* sctl_amd_last_error(void) { return last_error(); }

const char* sctl_amd_version(void) { return "sctl_amd 0.1 (gfx950)"; }

int sctl_amd_init(int device) {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    set_error("no HIP device visible");
    return SCTL_AMD_ERROR_NO_DEVICE;
  }
  SCTL_CHECK_ARG(device >= 0 && device < count, "device index out of range");
  SCTL_HIP_TRY(hipSetDevice(device));
  DeviceInfo d;
  return device_info(&d);
}

int sctl_amd_kernel_info(int kernel, int* coord_dim, int* src_dim, int* trg_dim, int* normal_dim, int* flops,
                         double* scale, const char** name) {
  if (kernel < 0 || kernel >= NUM_KERNELS) {
    set_error("unknown kernel id " + std::to_string(kernel));
    return SCTL_AMD_ERROR_UNKNOWN_KERNEL;
  }
  const KernelMeta& m = META[kernel];
  if (coord_dim) *coord_dim = 3;
  if (src_dim) *src_dim = m.K0;
  if (trg_dim) *trg_dim = m.K1;
  if (normal_dim) *normal_dim = m.ND;
  if (flops) *flops = m.flops;
  if (scale) *scale = kernel_scale(kernel);
  if (name) *name = m.name;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_eval(int kernel, int precision, int digits, const double* r_trg, int64_t nt, const double* r_src,
                  const double* n_src, const double* v_src, int64_t ns, double* v_trg, uint32_t flags,
                  void* stream) {
  return sctl_amd_eval_ctx(kernel, nullptr, precision, digits, r_trg, nt, r_src, n_src, v_src, ns, v_trg, flags,
                           stream);
}

int sctl_amd_eval_ctx(int kernel, const double* ctx, int precision, int digits, const double* r_trg, int64_t nt,
                      const double* r_src, const double* n_src, const double* v_src, int64_t ns, double* v_trg,
                      uint32_t flags, void* stream) {
  SCTL_TRY(eval_args(kernel, ctx, &precision, digits, r_trg, nt, r_src, n_src, v_src, ns, v_trg, flags,
                     SCTL_AMD_HOST_PTRS | SCTL_AMD_MIXED_NEWTON));
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  if (nt == 0 || ns == 0) return SCTL_AMD_SUCCESS;

  const KernelMeta& m = META[kernel];
  hipStream_t s = as_stream(stream);
  const bool host = flags & SCTL_AMD_HOST_PTRS;
  Scratch b_rt, b_rs, b_ns, b_vs, b_vt;
  const double *rt = r_trg, *rs = r_src, *nsrc = m.ND ? n_src : nullptr, *vs = v_src;
  double* vt = v_trg;
  if (host) {
    SCTL_TRY(stage_in(b_rt, r_trg, nt * 3, s, &rt));
    SCTL_TRY(stage_in(b_rs, r_src, ns * 3, s, &rs));
    if (m.ND) SCTL_TRY(stage_in(b_ns, n_src, ns * 3, s, &nsrc));
    SCTL_TRY(stage_in(b_vs, v_src, ns * m.K0, s, &vs));
    const double* vt_c = nullptr;
    SCTL_TRY(stage_in(b_vt, v_trg, nt * m.K1, s, &vt_c));
    vt = b_vt.as<double>();
  }
  const int newton = precision == SCTL_AMD_FP64 ? newton_steps_f64(digits) : ((flags & SCTL_AMD_MIXED_NEWTON) ? 1 : 0);
  const DenseShape shape = pick_dense_shape(kernel, precision, newton, nt, ns);
  if (precision == SCTL_AMD_FP64) {
    SCTL_TRY(launch_eval_f64(kernel, newton, rt, nt, rs, nsrc, vs, ns, vt, shape, ctx_value(kernel, ctx), s));
  } else {
    SCTL_TRY(launch_eval_mixed(kernel, newton, rt, nt, rs, nsrc, vs, ns, vt, shape, ctx_value(kernel, ctx), s));
  }
  if (host) {
    SCTL_HIP_TRY(hipMemcpyAsync(v_trg, vt, sizeof(double) * nt * m.K1, hipMemcpyDeviceToHost, s));
    b_rt.release(); b_rs.release(); b_ns.release(); b_vs.release(); b_vt.release();
    SCTL_HIP_TRY(hipStreamSynchronize(s));
  }
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_kernel_matrix(int kernel, int digits, const double* r_trg, int64_t nt, const double* r_src,
                           const double* n_src, int64_t ns, double* M, uint32_t flags, void* stream) {
  return sctl_amd_kernel_matrix_ctx(kernel, nullptr, digits, r_trg, nt, r_src, n_src, ns, M, flags, stream);
}

int sctl_amd_kernel_matrix_ctx(int kernel, const double* ctx, int digits, const double* r_trg, int64_t nt,
                               const double* r_src, const double* n_src, int64_t ns, double* M, uint32_t flags,
                               void* stream) {
  SCTL_TRY(check_common(kernel, r_trg, nt, r_src, n_src, ns, digits));
  SCTL_TRY(check_ctx(kernel, ctx));
  SCTL_CHECK_ARG(nt == 0 || ns == 0 || M, "M is NULL");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  if (nt == 0 || ns == 0) return SCTL_AMD_SUCCESS;
  const KernelMeta& m = META[kernel];
  hipStream_t s = as_stream(stream);
  const bool host = flags & SCTL_AMD_HOST_PTRS;
  Scratch b_rt, b_rs, b_ns, b_M;
  const double *rt = r_trg, *rs = r_src, *nsrc = m.ND ? n_src : nullptr;
  double* Md = M;
  const int64_t nM = ns * m.K0 * nt * m.K1;
  if (host) {
    SCTL_TRY(stage_in(b_rt, r_trg, nt * 3, s, &rt));
    SCTL_TRY(stage_in(b_rs, r_src, ns * 3, s, &rs));
    if (m.ND) SCTL_TRY(stage_in(b_ns, n_src, ns * 3, s, &nsrc));
    SCTL_TRY(b_M.alloc(sizeof(double) * nM, s));
    Md = b_M.as<double>();
  }
  SCTL_TRY(launch_kernel_matrix(kernel, newton_steps_f64(digits), rt, nt, rs, nsrc, ns, Md, ctx_value(kernel, ctx),
                                s));
  if (host) {
    SCTL_TRY(copy_to_host_pinned(M, Md, sizeof(double) * nM, s));
    b_rt.release(); b_rs.release(); b_ns.release(); b_M.release();
  }
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_malloc(void** ptr, size_t bytes) {
  SCTL_CHECK_ARG(ptr, "ptr is NULL");
  *ptr = nullptr;
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  if (!bytes) return SCTL_AMD_SUCCESS;
  SCTL_HIP_TRY(hipMalloc(ptr, bytes));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_free(void* ptr) {
  if (ptr) SCTL_HIP_TRY(hipFree(ptr));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream) {
  SCTL_CHECK_ARG(kind >= 0 && kind <= 2, "kind must be 0 (h2d), 1 (d2h) or 2 (d2d)");
  if (!bytes) return SCTL_AMD_SUCCESS;
  SCTL_CHECK_ARG(dst && src, "NULL pointer");
  const hipMemcpyKind k = kind == 0 ? hipMemcpyHostToDevice : (kind == 1 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
  hipStream_t s = as_stream(stream);
  SCTL_HIP_TRY(hipMemcpyAsync(dst, src, bytes, k, s));
  if (kind != 2) SCTL_HIP_TRY(hipStreamSynchronize(s));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_stream_create(void** stream) {
  SCTL_CHECK_ARG(stream, "stream is NULL");
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  hipStream_t s;
  SCTL_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_stream_destroy(void* stream) {
  if (stream) SCTL_HIP_TRY(hipStreamDestroy(static_cast<hipStream_t>(stream)));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_stream_synchronize(void* stream) {
  SCTL_HIP_TRY(hipStreamSynchronize(as_stream(stream)));
  return SCTL_AMD_SUCCESS;
}

}  // extern "C"
