// Repeated applications of a plan (batched or dense) captured as one hipGraph, and the device
// side of the counter-based random number generator.
//
// A Krylov solve applies the same boundary-integral operator once per
// iteration with new densities (lin-solve.txx:146-156 calls the operator,
// boundary_integral.txx:716-722 ComputePotential).  The graph built here
// records, for step s = 0..nsteps-1,
//     densities  <- N(0,1) draws of (seed + s, stream)       (k_gen_normal)
//     v_trg      += plan apply(densities)                    (two kernels)
// so a whole sequence of applies is one hipGraphLaunch.  Nothing in the
// graph allocates memory or issues a memset.
//
// Generator: SplitMix64 (Steele, Lea, Flood, OOPSLA 2014) used as a
// counter-based generator, element i of (seed, stream) is
// mix64(mix64(seed ^ stream*K) + (i+1)*golden); N(0,1) pairs by Box-Muller
// from consecutive draws.  The same definition is implemented on the host in
// sctl_amd/workloads.py.
#include <hip/hip_runtime.h>

#include <cmath>

#include "internal.hpp"

namespace sctl_amd {

namespace {

constexpr int GWG = 256;

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__global__ void __launch_bounds__(GWG) k_gen_normal(uint64_t seed, uint64_t stream, int64_t n, double* __restrict__ out) {
  const int64_t j = (int64_t)blockIdx.x * GWG + threadIdx.x;  // pair index
  if (2 * j >= n) return;
  const uint64_t key = mix64(seed ^ (stream * 0xD1B54A32D192ED03ull));
  const uint64_t x0 = mix64(key + (uint64_t)(2 * j + 1) * 0x9E3779B97F4A7C15ull);
  const uint64_t x1 = mix64(key + (uint64_t)(2 * j + 2) * 0x9E3779B97F4A7C15ull);
  const double u1 = ((double)(x0 >> 11) + 1.0) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(x1 >> 11) * (1.0 / 9007199254740992.0);
  const double rad = sqrt(-2.0 * log(u1));
  const double ang = 2.0 * M_PI * u2;
  out[2 * j] = rad * cos(ang);
  if (2 * j + 1 < n) out[2 * j + 1] = rad * sin(ang);
}

}  // namespace

int launch_gen_normal(uint64_t seed, uint64_t stream_id, int64_t n, double* out, hipStream_t s) {
  if (n <= 0) return SCTL_AMD_SUCCESS;
  const int64_t pairs = (n + 1) / 2;
  hipLaunchKernelGGL(k_gen_normal, dim3((unsigned)((pairs + GWG - 1) / GWG)), dim3(GWG), 0, s, seed, stream_id, n, out);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

}  // namespace sctl_amd

using namespace sctl_amd;

struct sctl_amd_graph_s {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int nsteps = 0, kernel = 0;
  int64_t interactions = 0;
  sctl_amd_plan plan = nullptr;  // batched plan captured (null for dense plans)
  uint64_t generation = 0;       // its generation at capture
  ~sctl_amd_graph_s() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
  }
};

extern "C" {

int sctl_amd_gen_normal(uint64_t seed, uint64_t stream_id, int64_t n, double* out, void* stream) {
  SCTL_CHECK_ARG(n >= 0, "negative size");
  SCTL_CHECK_ARG(n == 0 || out, "out is NULL");
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  return launch_gen_normal(seed, stream_id, n, out, stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread);
}

}  // extern "C"

namespace {

// Captures nsteps x (regenerate v_src, apply) on a private stream and
// instantiates the graph.
template <class Apply>
int capture_applies(sctl_amd_graph* out, int kernel, int64_t n_dens, int64_t interactions, int nsteps, uint64_t seed,
                    uint64_t stream_id, double* v_src, Apply apply) {
  hipStream_t cap;
  SCTL_HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
  auto* g = new sctl_amd_graph_s();
  g->nsteps = nsteps;
  g->kernel = kernel;
  g->interactions = interactions * nsteps;
  int rc = SCTL_AMD_SUCCESS;
  hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) {
    set_error(std::string("hipStreamBeginCapture: ") + hipGetErrorString(e));
    rc = SCTL_AMD_ERROR_HIP;
  }
  for (int step = 0; step < nsteps && rc == SCTL_AMD_SUCCESS; step++) {
    rc = launch_gen_normal(seed + (uint64_t)step, stream_id, n_dens, v_src, cap);
    if (rc == SCTL_AMD_SUCCESS) rc = apply(cap);
  }
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(cap, &graph);
  if (rc == SCTL_AMD_SUCCESS && e != hipSuccess) {
    set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    rc = SCTL_AMD_ERROR_HIP;
  }
  g->graph = graph;
  if (rc == SCTL_AMD_SUCCESS) {
    e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
      rc = SCTL_AMD_ERROR_HIP;
    }
  }
  (void)hipStreamDestroy(cap);
  if (rc != SCTL_AMD_SUCCESS) {
    delete g;
    return rc;
  }
  *out = g;
  return SCTL_AMD_SUCCESS;
}

}  // namespace

extern "C" {

int sctl_amd_plan_graph_create(sctl_amd_graph* out, sctl_amd_plan plan, int nsteps, uint64_t seed,
                               uint64_t stream_id, double* v_src, double* v_trg) {
  SCTL_CHECK_ARG(out, "graph is NULL");
  *out = nullptr;
  SCTL_CHECK_ARG(plan, "plan is NULL");
  SCTL_CHECK_ARG(nsteps >= 1, "nsteps must be >= 1");
  int64_t ns = 0, nt = 0, inter = 0;
  int k0 = 0, k1 = 0;
  SCTL_TRY(plan_sizes(plan, &ns, &nt, &k0, &k1, &inter));
  SCTL_CHECK_ARG(ns == 0 || v_src, "v_src is NULL");
  SCTL_CHECK_ARG(nt == 0 || v_trg, "v_trg is NULL");
  SCTL_TRY(capture_applies(out, plan_kernel(plan), ns * k0, inter, nsteps, seed, stream_id, v_src,
                           [&](hipStream_t cap) { return sctl_amd_plan_apply(plan, v_src, v_trg, 0, cap); }));
  (*out)->plan = plan;
  (*out)->generation = plan_generation(plan);
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_dense_plan_graph_create(sctl_amd_graph* out, sctl_amd_dense_plan plan, int nsteps, uint64_t seed,
                                     uint64_t stream_id, double* v_src, double* v_trg) {
  SCTL_CHECK_ARG(out, "graph is NULL");
  *out = nullptr;
  SCTL_CHECK_ARG(plan, "plan is NULL");
  SCTL_CHECK_ARG(nsteps >= 1, "nsteps must be >= 1");
  int64_t ns = 0, nt = 0;
  int k0 = 0, k1 = 0, kernel = 0;
  SCTL_TRY(dense_plan_sizes(plan, &ns, &nt, &k0, &k1, &kernel));
  SCTL_CHECK_ARG(ns == 0 || v_src, "v_src is NULL");
  SCTL_CHECK_ARG(nt == 0 || v_trg, "v_trg is NULL");
  return capture_applies(out, kernel, ns * k0, ns * nt, nsteps, seed, stream_id, v_src,
                         [&](hipStream_t cap) { return dense_plan_apply_device(plan, v_src, v_trg, cap); });
}

int sctl_amd_graph_launch(sctl_amd_graph g, void* stream) {
  SCTL_CHECK_ARG(g && g->exec, "graph is NULL");
  SCTL_CHECK_ARG(!g->plan || plan_generation(g->plan) == g->generation,
                 "the plan's launch shape or context changed after this graph was captured "
                 "(sctl_amd_plan_set_shape, sctl_amd_plan_tune, sctl_amd_plan_set_ctx): capture it again");
  SCTL_HIP_TRY(hipGraphLaunch(g->exec, stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread));
  profile_count(g->kernel, g->interactions);
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_graph_interactions(sctl_amd_graph g, int64_t* interactions) {
  SCTL_CHECK_ARG(g && interactions, "NULL argument");
  *interactions = g->interactions;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_graph_destroy(sctl_amd_graph g) {
  delete g;
  return SCTL_AMD_SUCCESS;
}

}  // extern "C"
