// Dense plans: one (target set, source set) Eval whose geometry stays on the
// device across applications -- the far field of BoundaryIntegralOp and
// ParticleFMM::EvalDirect, where the same targets and sources are summed once
// per GMRES iteration with new densities (boundary_integral.txx:1248-1309,
// fmm-wrapper.txx:551-623).
//
// Creation copies the positions (and normals) to plan-owned device memory,
// resolves the launch shape and, for mixed precision, computes the Hilbert
// orders of targets and sources once; it allocates every buffer an apply
// needs.  An apply is then the dense Eval without its per-call set-up: for
// mixed precision the tile packing (the densities change), the pairwise
// kernel and the split reduction; for FP64 the kernel and the reduction.  It
// allocates nothing and issues no memset, so applies can be captured in a
// hipGraph (sctl_amd_dense_plan_graph_create).  Results equal those of
// sctl_amd_eval with the same shape, bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "internal.hpp"
#include "kernels.hpp"

using namespace sctl_amd;

struct sctl_amd_dense_plan_s {
  int kernel = 0, precision = SCTL_AMD_FP64, newton = 2, K0 = 1, K1 = 1, ND = 0, device = 0;
  int64_t nt = 0, ns = 0;
  double ctx = 0.0;
  DenseShape shape;
  MixedGeometry geo;
  double *d_xt = nullptr, *d_xs = nullptr, *d_xn = nullptr, *d_part = nullptr;
  uint32_t *d_tperm = nullptr, *d_sperm = nullptr;
  float4* d_packed = nullptr;
  double4* d_meta = nullptr;

  ~sctl_amd_dense_plan_s() {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != device) (void)hipSetDevice(device);
    void* ptrs[] = {d_xt, d_xs, d_xn, d_part, d_tperm, d_sperm, d_packed, d_meta};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
  }
};

namespace {

int dev_alloc(void** p, size_t bytes) {
  *p = nullptr;
  if (!bytes) return SCTL_AMD_SUCCESS;
  SCTL_HIP_TRY(hipMalloc(p, bytes));
  return SCTL_AMD_SUCCESS;
}

int dense_plan_apply_dev(sctl_amd_dense_plan_s* p, const double* v_src, double* v_trg, hipStream_t s) {
  if (p->nt == 0 || p->ns == 0) return SCTL_AMD_SUCCESS;
  if (p->precision == SCTL_AMD_FP64)
    return f64_apply(p->kernel, p->newton, p->d_xt, p->nt, p->d_xs, p->d_xn, v_src, p->ns, v_trg, p->shape, p->ctx,
                     p->d_part, s);
  return mixed_apply(p->kernel, p->newton, p->geo, p->shape, p->d_xt, p->nt, p->d_tperm, p->d_xs, p->d_xn, v_src,
                     p->ns, p->d_sperm, p->d_packed, p->d_meta, p->d_part, v_trg, p->ctx, s);
}

}  // namespace

// Used by the graph module: sizes and the allocation-free apply.
namespace sctl_amd {
int dense_plan_sizes(sctl_amd_dense_plan p, int64_t* ns, int64_t* nt, int* K0, int* K1, int* kernel) {
  SCTL_CHECK_ARG(p, "dense plan is NULL");
  *ns = p->ns;
  *nt = p->nt;
  *K0 = p->K0;
  *K1 = p->K1;
  *kernel = p->kernel;
  return SCTL_AMD_SUCCESS;
}
int dense_plan_apply_device(sctl_amd_dense_plan p, const double* v_src, double* v_trg, hipStream_t s) {
  SCTL_CHECK_ARG(p, "dense plan is NULL");
  return dense_plan_apply_dev(p, v_src, v_trg, s);
}
}  // namespace sctl_amd

extern "C" {

int sctl_amd_dense_plan_create(sctl_amd_dense_plan* out, int kernel, int precision, int digits, uint32_t flags,
                               const double* r_trg, int64_t nt, const double* r_src, const double* n_src, int64_t ns,
                               const double* ctx, void* stream) {
  SCTL_CHECK_ARG(out, "plan is NULL");
  *out = nullptr;
  int K0, K1, ND;
  if (kernel_dims(kernel, &K0, &K1, &ND) != SCTL_AMD_SUCCESS) {
    set_error("unknown kernel id " + std::to_string(kernel));
    return SCTL_AMD_ERROR_UNKNOWN_KERNEL;
  }
  SCTL_CHECK_ARG(precision >= SCTL_AMD_FP64 && precision <= SCTL_AMD_AUTO, "unknown precision");
  SCTL_CHECK_ARG(digits >= -1, "digits must be -1 or >= 0");
  SCTL_CHECK_ARG((flags & ~(SCTL_AMD_HOST_PTRS | SCTL_AMD_MIXED_NEWTON)) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(nt >= 0 && ns >= 0, "negative target or source count");
  SCTL_CHECK_ARG(nt == 0 || r_trg, "r_trg is NULL");
  SCTL_CHECK_ARG(ns == 0 || r_src, "r_src is NULL");
  SCTL_CHECK_ARG(ns == 0 || ND == 0 || n_src, "n_src is NULL for a kernel with a normal");
  SCTL_CHECK_ARG(!kernel_needs_ctx(kernel) || ctx, "the kernel needs its wavenumber (ctx is NULL)");
  SCTL_CHECK_ARG(!kernel_needs_ctx(kernel) || std::isfinite(*ctx), "the wavenumber is not finite");
  if (precision == SCTL_AMD_AUTO) precision = (digits >= 0 && digits <= 7) ? SCTL_AMD_MIXED : SCTL_AMD_FP64;
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;

  auto* p = new sctl_amd_dense_plan_s();
  p->kernel = kernel;
  p->precision = precision;
  p->newton = precision == SCTL_AMD_FP64 ? newton_steps_f64(digits) : ((flags & SCTL_AMD_MIXED_NEWTON) ? 1 : 0);
  p->K0 = K0;
  p->K1 = K1;
  p->ND = ND;
  p->device = dev.device;
  p->nt = nt;
  p->ns = ns;
  p->ctx = kernel_needs_ctx(kernel) ? *ctx : 0.0;
  p->shape = pick_dense_shape(kernel, precision, p->newton, nt, ns);
  const hipMemcpyKind kind = (flags & SCTL_AMD_HOST_PTRS) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  auto copy = [&](double** dst, const double* src, int64_t n) -> int {
    if (n <= 0) return SCTL_AMD_SUCCESS;
    SCTL_TRY(dev_alloc(reinterpret_cast<void**>(dst), sizeof(double) * n));
    SCTL_HIP_TRY(hipMemcpyAsync(*dst, src, sizeof(double) * n, kind, s));
    return SCTL_AMD_SUCCESS;
  };
  int rc = copy(&p->d_xt, r_trg, nt * 3);
  if (!rc) rc = copy(&p->d_xs, r_src, ns * 3);
  if (!rc && ND) rc = copy(&p->d_xn, n_src, ns * 3);
  if (!rc && nt > 0 && ns > 0) {
    size_t part = 0;
    if (precision == SCTL_AMD_FP64) {
      rc = f64_part_bytes(kernel, nt, ns, p->shape, &part);
    } else {
      rc = mixed_geometry(kernel, nt, ns, p->shape, &p->geo);
      part = p->geo.part_bytes;
      if (!rc) rc = dev_alloc(reinterpret_cast<void**>(&p->d_tperm), sizeof(uint32_t) * nt);
      if (!rc) rc = dev_alloc(reinterpret_cast<void**>(&p->d_sperm), sizeof(uint32_t) * ns);
      if (!rc) rc = dev_alloc(reinterpret_cast<void**>(&p->d_packed), p->geo.packed_bytes);
      if (!rc) rc = dev_alloc(reinterpret_cast<void**>(&p->d_meta), p->geo.meta_bytes);
      if (!rc) rc = mixed_order(p->d_xt, nt, p->d_xs, ns, p->d_tperm, p->d_sperm, nullptr, s);
    }
    if (!rc) rc = dev_alloc(reinterpret_cast<void**>(&p->d_part), part);
  }
  if (!rc && hipStreamSynchronize(s) != hipSuccess) {
    set_error("hipStreamSynchronize failed while creating a dense plan");
    rc = SCTL_AMD_ERROR_HIP;
  }
  if (rc) {
    delete p;
    return rc;
  }
  *out = p;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_dense_plan_apply(sctl_amd_dense_plan p, const double* v_src, double* v_trg, uint32_t flags,
                              void* stream) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(p->ns == 0 || v_src, "v_src is NULL");
  SCTL_CHECK_ARG(p->nt == 0 || v_trg, "v_trg is NULL");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  if (!(flags & SCTL_AMD_HOST_PTRS)) return dense_plan_apply_dev(p, v_src, v_trg, s);
  Scratch bs, bt;
  SCTL_TRY(bs.alloc(sizeof(double) * p->ns * p->K0, s));
  SCTL_TRY(bt.alloc(sizeof(double) * p->nt * p->K1, s));
  if (bs.bytes()) SCTL_HIP_TRY(hipMemcpyAsync(bs.as<double>(), v_src, bs.bytes(), hipMemcpyHostToDevice, s));
  if (bt.bytes()) SCTL_HIP_TRY(hipMemcpyAsync(bt.as<double>(), v_trg, bt.bytes(), hipMemcpyHostToDevice, s));
  SCTL_TRY(dense_plan_apply_dev(p, bs.as<double>(), bt.as<double>(), s));
  if (bt.bytes()) SCTL_HIP_TRY(hipMemcpyAsync(v_trg, bt.as<double>(), bt.bytes(), hipMemcpyDeviceToHost, s));
  bs.release();
  bt.release();
  SCTL_HIP_TRY(hipStreamSynchronize(s));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_dense_plan_shape(sctl_amd_dense_plan p, int* precision, int* targets_per_lane, int* wg_per_cu_cap) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  if (precision) *precision = p->precision;
  if (targets_per_lane) *targets_per_lane = p->shape.tpl;
  if (wg_per_cu_cap) *wg_per_cu_cap = p->shape.occ;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_dense_plan_destroy(sctl_amd_dense_plan p) {
  delete p;
  return SCTL_AMD_SUCCESS;
}

}  // extern "C"
