// Profile counters and kernel timing (sctl_amd_profile_*).
#include <hip/hip_runtime.h>

#include <atomic>
#include <mutex>
#include <vector>

#include "internal.hpp"

namespace sctl_amd {

namespace {
// FLOPS() of kernel_functions.hpp:37,66,86,130,185,225,269,330,371,397,421,456,
// 496,534,578,632 (kernel ids 0..15).
constexpr int64_t REF_FLOPS[] = {6, 14, 23, 26, 11, 22, 25, 39, 26, 26, 15, 51, 20, 29, 35, 41};
constexpr int N_REF_FLOPS = (int)(sizeof(REF_FLOPS) / sizeof(REF_FLOPS[0]));
std::atomic<int64_t> g_inter{0}, g_flops{0};
std::atomic<bool> g_timing{false};
std::mutex g_mu;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_pending;
int64_t g_launches = 0;
double g_ms = 0;
}  // namespace

bool stream_capturing(hipStream_t s) {
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}

void profile_count(int ker, int64_t interactions) {
  g_inter += interactions;
  if (ker >= 0 && ker < N_REF_FLOPS) g_flops += interactions * REF_FLOPS[ker];
}

ProfileScope::ProfileScope(hipStream_t s) : s_(s) {
  if (!g_timing.load()) return;
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s_, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return;
  if (hipEventCreate(&a_) != hipSuccess) {
    a_ = nullptr;
    return;
  }
  if (hipEventRecord(a_, s_) != hipSuccess) {
    (void)hipEventDestroy(a_);
    a_ = nullptr;
  }
}

ProfileScope::~ProfileScope() {
  if (!a_) return;
  hipEvent_t b = nullptr;
  if (hipEventCreate(&b) == hipSuccess && hipEventRecord(b, s_) == hipSuccess) {
    std::lock_guard<std::mutex> lock(g_mu);
    g_pending.emplace_back(a_, b);
    return;
  }
  if (b) (void)hipEventDestroy(b);
  (void)hipEventDestroy(a_);
}

}  // namespace sctl_amd

using namespace sctl_amd;

extern "C" {

int sctl_amd_profile_enable(int enable) {
  g_timing = enable != 0;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_profile_read(int64_t* interactions, int64_t* flops, int64_t* timed_launches, double* kernel_ms) {
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pend;
  {
    std::lock_guard<std::mutex> lock(g_mu);
    pend.swap(g_pending);
  }
  double ms_sum = 0;
  int64_t n = 0;
  int rc = SCTL_AMD_SUCCESS;
  for (auto& ab : pend) {
    float ms = 0;
    hipError_t e = hipEventSynchronize(ab.second);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, ab.first, ab.second);
    if (e != hipSuccess) {
      set_error(std::string("profile event: ") + hipGetErrorString(e));
      rc = SCTL_AMD_ERROR_HIP;
    } else {
      ms_sum += ms;
      n++;
    }
    (void)hipEventDestroy(ab.first);
    (void)hipEventDestroy(ab.second);
  }
  std::lock_guard<std::mutex> lock(g_mu);
  g_ms += ms_sum;
  g_launches += n;
  if (interactions) *interactions = g_inter.load();
  if (flops) *flops = g_flops.load();
  if (timed_launches) *timed_launches = g_launches;
  if (kernel_ms) *kernel_ms = g_ms;
  return rc;
}

int sctl_amd_profile_reset(void) {
  int64_t a;
  sctl_amd_profile_read(&a, &a, &a, nullptr);
  std::lock_guard<std::mutex> lock(g_mu);
  g_inter = 0;
  g_flops = 0;
  g_launches = 0;
  g_ms = 0;
  return SCTL_AMD_SUCCESS;
}

}  // extern "C"
