// Host-side internals shared by the C-ABI and the kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "../../include/sctl_amd.h"

namespace sctl_amd {

// Per-thread message of the last failing call (sctl_amd_last_error()).
void set_error(const std::string& msg);
const char* last_error();

// Status helpers: every HIP call on the product path is checked.
#define SCTL_HIP_TRY(expr)                                                                      \
  do {                                                                                          \
    hipError_t e_ = (expr);                                                                     \
    if (e_ != hipSuccess) {                                                                     \
      ::sctl_amd::set_error(std::string(#expr) + " failed: " + hipGetErrorString(e_));          \
      return e_ == hipErrorOutOfMemory ? SCTL_AMD_ERROR_OUT_OF_MEMORY : SCTL_AMD_ERROR_HIP;    \
    }                                                                                           \
  } while (0)

#define SCTL_TRY(expr)                   \
  do {                                   \
    int s_ = (expr);                     \
    if (s_ != SCTL_AMD_SUCCESS) return s_; \
  } while (0)

#define SCTL_CHECK_ARG(cond, msg)                       \
  do {                                                  \
    if (!(cond)) {                                      \
      ::sctl_amd::set_error(msg);                       \
      return SCTL_AMD_ERROR_INVALID_VALUE;              \
    }                                                   \
  } while (0)

// Device facts, looked up once per device, and the library's memory pool on it.
struct DeviceInfo {
  int device = -1;
  int num_cus = 0;
  int max_threads = 0;
  hipMemPool_t pool = nullptr;
  bool ok = false;
};
int device_info(DeviceInfo* out);  // for the current device
// Freed scratch the library's pool keeps reserved for reuse (its release
// threshold); beyond it, memory goes back to the driver at synchronisation.
constexpr uint64_t POOL_KEEP_BYTES = (uint64_t)4 << 30;

// Stream-ordered device scratch: allocated from the library's pool on the
// current device, freed on the same stream.
class Scratch {
 public:
  Scratch() = default;
  Scratch(const Scratch&) = delete;
  Scratch& operator=(const Scratch&) = delete;
  ~Scratch() { release(); }
  int alloc(size_t bytes, hipStream_t s) {
    release();
    stream_ = s;
    if (!bytes) return SCTL_AMD_SUCCESS;
    DeviceInfo dev;
    SCTL_TRY(device_info(&dev));
    SCTL_HIP_TRY(hipMallocFromPoolAsync(&ptr_, bytes, dev.pool, s));
    bytes_ = bytes;
    return SCTL_AMD_SUCCESS;
  }
  void release() {
    if (ptr_) (void)hipFreeAsync(ptr_, stream_);
    ptr_ = nullptr;
    bytes_ = 0;
  }
  template <class T> T* as() const { return static_cast<T*>(ptr_); }
  size_t bytes() const { return bytes_; }

 private:
  void* ptr_ = nullptr;
  size_t bytes_ = 0;
  hipStream_t stream_ = nullptr;
};

// Offsets of sub-buffers carved from one device allocation (256-byte aligned),
// so that a call's scratch costs one stream-ordered allocation and free.
struct Carve {
  size_t bytes = 0;
  size_t take(size_t n) {
    const size_t off = (bytes + 255) & ~(size_t)255;
    bytes = off + n;
    return off;
  }
};

// Launch-shape knobs of the dense kernels (chosen by the tuner or defaults).
struct DenseShape {
  int tpl = 0;  // targets per lane: 1, 2 or 4 (0: default for the size)
  int occ = 0;  // cap on resident workgroups per CU (0: no cap); enforced with dynamic LDS
};

// Per-split partial sums of a dense Eval whose sources are split over
// gridDim.y take nsplit * nt * K1 doubles when every split runs in one
// launch.  Above `budget` bytes the splits run in groups of successive
// launches over one group's worth of partials, each group reduced into
// v_trg before the next; a group of one split accumulates straight into
// v_trg and needs no partials at all.  Returns the splits per launch.
constexpr size_t PART_BUDGET = (size_t)256 << 20;
inline int64_t split_group(int64_t nsplit, size_t split_bytes, size_t budget) {
  if (nsplit <= 1) return 1;
  if (split_bytes == 0 || (size_t)nsplit * split_bytes <= budget) return nsplit;
  const int64_t g = (int64_t)(budget / split_bytes);
  return g < 2 ? 1 : g;
}

// Dynamic LDS bytes that cap a kernel with `static_lds` bytes of LDS at `occ`
// resident workgroups per CU (0 when occ <= 0).  Sets the kernel's dynamic-LDS
// limit when the request exceeds the default.
int occupancy_lds(const void* kernel, int occ, size_t static_lds, size_t* dyn_bytes);

// Dense Eval launchers (device pointers, stream-ordered).  v_trg is
// accumulated: v_trg[t*K1+k] += scale * sum.  ctx: the wavenumber of the
// Helmholtz kernels (ignored by the others).
int launch_eval_f64(int ker, int newton, const double* r_trg, int64_t nt, const double* r_src,
                    const double* n_src, const double* v_src, int64_t ns, double* v_trg, const DenseShape& shape,
                    double ctx, hipStream_t stream);
int launch_eval_mixed(int ker, int newton, const double* r_trg, int64_t nt, const double* r_src,
                      const double* n_src, const double* v_src, int64_t ns, double* v_trg, const DenseShape& shape,
                      double ctx, hipStream_t stream);
int launch_kernel_matrix(int ker, int newton, const double* r_trg, int64_t nt, const double* r_src,
                         const double* n_src, int64_t ns, double* M, double ctx, hipStream_t stream);

// The dense Eval launchers in two parts, for callers that keep the geometry
// (dense plans): a one-time part (sizes; for mixed precision the Hilbert
// orders) and an apply that allocates nothing and issues no memset.
// Both take the same part_budget (split_group); part holds f64_part_bytes.
int f64_part_bytes(int ker, int64_t nt, int64_t ns, const DenseShape& shape, size_t* bytes,
                   size_t part_budget = PART_BUDGET);
int f64_apply(int ker, int newton, const double* r_trg, int64_t nt, const double* r_src, const double* n_src,
              const double* v_src, int64_t ns, double* v_trg, const DenseShape& shape, double ctx, double* part,
              hipStream_t stream, size_t part_budget = PART_BUDGET);
struct MixedGeometry {
  int ti = 0, Q = 0;
  int64_t tb = 0, ntiles = 0, nsplit = 1, tps = 0;
  int64_t group = 1;  // splits per launch (split_group)
  size_t packed_bytes = 0, meta_bytes = 0, part_bytes = 0;
};
// part_budget bounds part_bytes (split_group).
int mixed_geometry(int ker, int64_t nt, int64_t ns, const DenseShape& shape, MixedGeometry* g,
                   size_t part_budget = PART_BUDGET);
// Hilbert orders of targets and sources; `work` (device) holds
// mixed_order_bytes(nt, ns) bytes, or is null to allocate it per call.
size_t mixed_order_bytes(int64_t nt, int64_t ns);
int mixed_order(const double* r_trg, int64_t nt, const double* r_src, int64_t ns, uint32_t* tperm, uint32_t* sperm,
                void* work, hipStream_t stream);
int mixed_apply(int ker, int newton, const MixedGeometry& g, const DenseShape& shape, const double* r_trg, int64_t nt,
                const uint32_t* tperm, const double* r_src, const double* n_src, const double* v_src, int64_t ns,
                const uint32_t* sperm, float4* packed, double4* meta, double* part, double* v_trg, double ctx,
                hipStream_t stream);
// A whole mixed Eval in a caller-provided workspace `work` of
// mixed_fixed_bytes(g, nt, ns) + g.part_bytes bytes (256-byte aligned):
// orders, packing, summation; allocates nothing.
size_t mixed_fixed_bytes(const MixedGeometry& g, int64_t nt, int64_t ns);
int mixed_eval_in(int ker, int newton, const MixedGeometry& g, const DenseShape& shape, const double* r_trg,
                  int64_t nt, const double* r_src, const double* n_src, const double* v_src, int64_t ns,
                  double* v_trg, double ctx, char* work, hipStream_t stream);
// True while `s` is being captured into a graph (no host-side counting then).
bool stream_capturing(hipStream_t s);

// Hilbert-curve orderings of two point sets (keys over each set's bounding
// box) by one sort: permutations of [0, na) and [0, nb).  `work` (device)
// holds hilbert_order_bytes(na, nb) bytes.  Implemented in sort.hip.
size_t hilbert_order_bytes(int64_t na, int64_t nb);
int hilbert_order(const double* a, int64_t na, uint32_t* perm_a, const double* b, int64_t nb, uint32_t* perm_b,
                 void* work, hipStream_t stream);
// Axis-aligned bounding box of two point sets (either may be empty):
// box6 (device) = {min x, min y, min z, max x, max y, max z}.  sort.hip.
int bounding_box(const double* a, int64_t na, const double* b, int64_t nb, double* box6, hipStream_t stream);

// Shape defaults and tuning-database lookup (tune.cpp).  `newton` is the
// Newton count of the path (part of the tuning key).
DenseShape pick_dense_shape(int ker, int precision, int newton, int64_t nt, int64_t ns);

// Batched plan knobs (tune.cpp / batched.hip).
struct PlanShape {
  int spl = 0;  // sources per lane per LDS stage: 1, 2 or 4 (stage = 64*spl sources)
  int tpl = 0;  // targets per lane: 1 or 2
  int occ = 0;  // cap on resident workgroups per CU (0: no cap)
};
bool lookup_plan_shape(int ker, int precision, int64_t interactions, int64_t nitems, PlanShape* out);

// Profiling (profile.cpp): count(...) is always on; a ProfileScope records
// HIP events around one launch while timing is enabled.
void profile_count(int ker, int64_t interactions);
class ProfileScope {
 public:
  explicit ProfileScope(hipStream_t s);
  ~ProfileScope();

 private:
  hipEvent_t a_ = nullptr;
  hipStream_t s_ = nullptr;
};

// Dense plans (dense_plan.cpp): sizes and the allocation-free apply.
int dense_plan_sizes(sctl_amd_dense_plan p, int64_t* ns, int64_t* nt, int* K0, int* K1, int* kernel);
int dense_plan_apply_device(sctl_amd_dense_plan p, const double* v_src, double* v_trg, hipStream_t s);

// Plan sizes (batched.hip): source/target totals, K0/K1, interactions per apply.
int plan_sizes(sctl_amd_plan p, int64_t* ns, int64_t* nt, int* K0, int* K1, int64_t* interactions);
int plan_kernel(sctl_amd_plan p);
// Changes whenever the plan's launch state changes (shape, tuning, context).
uint64_t plan_generation(sctl_amd_plan p);

// Argument checks of an Eval call (capi.cpp); resolves SCTL_AMD_AUTO in
// *precision.  Flag bits outside allowed_flags are an error.
int eval_args(int kernel, const double* ctx, int* precision, int digits, const double* r_trg, int64_t nt,
              const double* r_src, const double* n_src, const double* v_src, int64_t ns, const double* v_trg,
              uint32_t flags, uint32_t allowed_flags);
// The kernel's context value as the launchers take it (0 when unused).
double eval_ctx_value(int kernel, const double* ctx);

int kernel_dims(int ker, int* K0, int* K1, int* ND);
// True for the kernels that need a context value (the Helmholtz wavenumber).
bool kernel_needs_ctx(int ker);
double kernel_scale(int ker);
// Largest targets-per-lane count the dense and batched kernels take for a
// kernel (larger requests run with this many).
int kernel_max_tpl(int ker);
int newton_steps_f64(int digits);

}  // namespace sctl_amd
