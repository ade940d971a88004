// Launch-shape autotuner.
//
// Knobs:
//   dense Eval   TPL (targets per lane: 1, 2, 4) and a cap on resident
//                workgroups per CU (waves per CU = 4 x cap; 0 = no cap)
//   plans        SPL (LDS stage of 64*SPL sources: 1, 2, 4), TPL (1, 2) and
//                the workgroups-per-CU cap
// The tuner times every candidate on the caller's own data with HIP events on
// the caller's stream and keeps the fastest.  Results live in an in-memory
// database keyed by (operation, kernel, precision, Newton count, log2 size
// classes); the database is read from and written to the file named by
// sctl_amd_tune_db_path() or the SCTL_AMD_TUNE_DB environment variable, one
// entry per line.  Lookups that miss fall back to size-based defaults.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <string>
#include <vector>

#include "internal.hpp"

struct sctl_amd_plan_s;

namespace sctl_amd {

int plan_configure(sctl_amd_plan_s* p, PlanShape s);
int plan_precision(sctl_amd_plan p);
int64_t plan_nonempty_blocks(sctl_amd_plan p);
PlanShape plan_shape(sctl_amd_plan p);

namespace {

std::mutex g_mu;
std::map<std::string, std::vector<double>> g_db;
std::string g_path;
bool g_loaded = false;

int log2_bucket(int64_t n) {
  int b = 0;
  while (n > 1) {
    n >>= 1;
    b++;
  }
  return b;
}

std::string dense_key(int ker, int precision, int newton, int64_t nt, int64_t ns) {
  char buf[96];
  std::snprintf(buf, sizeof(buf), "dense k%d p%d n%d t%d s%d", ker, precision, newton, log2_bucket(nt), log2_bucket(ns));
  return buf;
}
std::string plan_key(int ker, int precision, int64_t inter, int64_t blocks) {
  char buf[96];
  std::snprintf(buf, sizeof(buf), "plan k%d p%d i%d b%d", ker, precision, log2_bucket(inter), log2_bucket(blocks));
  return buf;
}

void load_locked() {
  if (g_loaded) return;
  g_loaded = true;
  if (g_path.empty()) {
    const char* env = std::getenv("SCTL_AMD_TUNE_DB");
    if (env) g_path = env;
  }
  if (g_path.empty()) return;
  std::ifstream in(g_path);
  std::string line;
  while (std::getline(in, line)) {
    const size_t bar = line.find('|');
    if (bar == std::string::npos || line[0] == '#') continue;
    std::istringstream vs(line.substr(bar + 1));
    std::vector<double> v;
    double x;
    while (vs >> x) v.push_back(x);
    g_db[line.substr(0, bar)] = v;
  }
}

void store_locked(const std::string& key, const std::vector<double>& v) {
  g_db[key] = v;
  if (g_path.empty()) return;
  std::ofstream out(g_path, std::ios::trunc);
  out << "# sctl_amd tuning database: key | values\n";
  for (auto& kv : g_db) {
    out << kv.first << " |";
    for (double x : kv.second) out << " " << x;
    out << "\n";
  }
}

bool lookup(const std::string& key, std::vector<double>* v) {
  std::lock_guard<std::mutex> lock(g_mu);
  load_locked();
  auto it = g_db.find(key);
  if (it == g_db.end()) return false;
  *v = it->second;
  return true;
}

void store(const std::string& key, const std::vector<double>& v) {
  std::lock_guard<std::mutex> lock(g_mu);
  load_locked();
  store_locked(key, v);
}

// Milliseconds of one run of fn() on stream s; a run shorter than 20 ms is
// repeated at least `reps` more times, and as often as fills about 20 ms (at
// most 50), and the mean of those is returned instead.
template <class F> int time_ms(F fn, hipStream_t s, int reps, double* ms) {
  hipEvent_t a, b;
  SCTL_HIP_TRY(hipEventCreate(&a));
  SCTL_HIP_TRY(hipEventCreate(&b));
  int rc = SCTL_AMD_SUCCESS;
  float t = 0;
  for (int pass = 0; pass < 2 && !rc; pass++) {
    const int n = pass == 0 ? 1 : std::min(50, std::max(reps, (int)std::ceil(20.0 / std::max(t, 1e-3f))));
    if (hipEventRecord(a, s) != hipSuccess) rc = SCTL_AMD_ERROR_HIP;
    for (int r = 0; r < n && !rc; r++) rc = fn();
    if (!rc && hipEventRecord(b, s) != hipSuccess) rc = SCTL_AMD_ERROR_HIP;
    if (!rc && (hipEventSynchronize(b) != hipSuccess || hipEventElapsedTime(&t, a, b) != hipSuccess))
      rc = SCTL_AMD_ERROR_HIP;
    t /= n;
    if (t >= 20.0f) break;
  }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  if (rc == SCTL_AMD_ERROR_HIP) set_error("event timing failed while tuning");
  *ms = t;
  return rc;
}

}  // namespace

int occupancy_lds(const void* kernel, int occ, size_t static_lds, size_t* dyn_bytes) {
  *dyn_bytes = 0;
  if (occ <= 0) return SCTL_AMD_SUCCESS;
  int dev = 0;
  SCTL_HIP_TRY(hipGetDevice(&dev));
  int per_cu = 0;
  SCTL_HIP_TRY(hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
  if (per_cu <= 0) per_cu = 160 * 1024;
  // The kernel's own static LDS as the runtime reports it; the caller's
  // figure is a lower bound.
  hipFuncAttributes attr;
  SCTL_HIP_TRY(hipFuncGetAttributes(&attr, kernel));
  static_lds = std::max(static_lds, (size_t)attr.sharedSizeBytes);
  size_t want = (size_t)per_cu / (size_t)occ;
  want -= want % 512;
  if (want <= static_lds + 512) return SCTL_AMD_SUCCESS;  // the kernel alone already fits no more than occ
  const size_t dyn = want - static_lds - 512;
  if (dyn > 64 * 1024) {
    SCTL_HIP_TRY(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)dyn));
  }
  *dyn_bytes = dyn;
  return SCTL_AMD_SUCCESS;
}

DenseShape pick_dense_shape(int ker, int precision, int newton, int64_t nt, int64_t ns) {
  DenseShape s;
  std::vector<double> v;
  if (lookup(dense_key(ker, precision, newton, nt, ns), &v) && v.size() >= 2) {
    s.tpl = (int)v[0];
    s.occ = (int)v[1];
    return s;
  }
  // FP64 on 2^16 or more targets: four targets per lane (clamped per kernel;
  // the tuner's choice on 2^20 x 2^20 and 7-16 % faster than one on 2^16-2^18
  // cubes for the Laplace and Stokes single and double layers), except for
  // the Helmholtz kernels, whose phase terms make four targets per lane up to
  // a third slower.  Otherwise enough target blocks to fill the device with
  // two targets per lane, else one target per lane and the source range split
  // over the grid.
  s.occ = 0;
  if (precision == SCTL_AMD_FP64)
    s.tpl = (!kernel_needs_ctx(ker) && nt >= (1 << 16)) ? 4 : 1;
  else
    s.tpl = nt >= 256LL * 2 * 2048 ? 2 : 1;
  return s;
}

bool lookup_plan_shape(int ker, int precision, int64_t interactions, int64_t nitems, PlanShape* out) {
  std::vector<double> v;
  if (!lookup(plan_key(ker, precision, interactions, nitems), &v) || v.size() < 3) return false;
  out->spl = (int)v[0];
  out->tpl = (int)v[1];
  out->occ = (int)v[2];
  return true;
}

}  // namespace sctl_amd

using namespace sctl_amd;

extern "C" {

int sctl_amd_tune_db_path(const char* path) {
  std::lock_guard<std::mutex> lock(g_mu);
  g_path = path ? path : "";
  g_loaded = false;
  g_db.clear();
  load_locked();
  return SCTL_AMD_SUCCESS;
}

namespace {

// Resolves precision AUTO and the Newton count exactly as sctl_amd_eval does.
int resolve_dense(int kernel, int* precision, int digits, uint32_t flags, int* newton) {
  int K0, K1, ND;
  if (kernel_dims(kernel, &K0, &K1, &ND) != SCTL_AMD_SUCCESS) {
    set_error("unknown kernel id " + std::to_string(kernel));
    return SCTL_AMD_ERROR_UNKNOWN_KERNEL;
  }
  SCTL_CHECK_ARG(digits >= -1, "digits must be -1 or >= 0");
  SCTL_CHECK_ARG(*precision >= SCTL_AMD_FP64 && *precision <= SCTL_AMD_AUTO, "unknown precision");
  if (*precision == SCTL_AMD_AUTO) *precision = (digits >= 0 && digits <= 7) ? SCTL_AMD_MIXED : SCTL_AMD_FP64;
  *newton = *precision == SCTL_AMD_FP64 ? newton_steps_f64(digits) : ((flags & SCTL_AMD_MIXED_NEWTON) ? 1 : 0);
  return SCTL_AMD_SUCCESS;
}

}  // namespace

int sctl_amd_dense_shape_get(int kernel, int precision, int digits, uint32_t flags, int64_t nt, int64_t ns,
                             int* targets_per_lane, int* wg_per_cu_cap) {
  int newton = 0;
  SCTL_TRY(resolve_dense(kernel, &precision, digits, flags, &newton));
  SCTL_CHECK_ARG(nt >= 0 && ns >= 0, "negative size");
  const DenseShape s = pick_dense_shape(kernel, precision, newton, nt, ns);
  if (targets_per_lane) *targets_per_lane = s.tpl;
  if (wg_per_cu_cap) *wg_per_cu_cap = s.occ;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_dense_shape_set(int kernel, int precision, int digits, uint32_t flags, int64_t nt, int64_t ns,
                             int targets_per_lane, int wg_per_cu_cap) {
  int newton = 0;
  SCTL_TRY(resolve_dense(kernel, &precision, digits, flags, &newton));
  SCTL_CHECK_ARG(nt >= 0 && ns >= 0, "negative size");
  SCTL_CHECK_ARG(targets_per_lane == 1 || targets_per_lane == 2 || targets_per_lane == 4,
                 "targets_per_lane must be 1, 2 or 4");
  SCTL_CHECK_ARG(wg_per_cu_cap >= 0, "wg_per_cu_cap must be >= 0");
  store(dense_key(kernel, precision, newton, nt, ns), {(double)targets_per_lane, (double)wg_per_cu_cap, 0.0});
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_tune_eval(int kernel, int precision, int digits, const double* r_trg, int64_t nt, const double* r_src,
                       const double* n_src, const double* v_src, int64_t ns, double* v_scratch, uint32_t flags,
                       void* stream, int* best_tpl, int* best_occ, double* best_ms) {
  int K0, K1, ND;
  if (kernel_dims(kernel, &K0, &K1, &ND) != SCTL_AMD_SUCCESS) {
    set_error("unknown kernel id " + std::to_string(kernel));
    return SCTL_AMD_ERROR_UNKNOWN_KERNEL;
  }
  SCTL_CHECK_ARG(!(flags & SCTL_AMD_HOST_PTRS), "the tuner takes device pointers");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_MIXED_NEWTON) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(nt > 0 && ns > 0 && r_trg && r_src && v_src && v_scratch && (!ND || n_src), "bad arguments");
  SCTL_CHECK_ARG(digits >= -1, "digits must be -1 or >= 0");
  SCTL_CHECK_ARG(precision >= SCTL_AMD_FP64 && precision <= SCTL_AMD_AUTO, "unknown precision");
  if (precision == SCTL_AMD_AUTO) precision = (digits >= 0 && digits <= 7) ? SCTL_AMD_MIXED : SCTL_AMD_FP64;
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  const int newton = precision == SCTL_AMD_FP64 ? newton_steps_f64(digits) : ((flags & SCTL_AMD_MIXED_NEWTON) ? 1 : 0);
  double best = 1e300;
  DenseShape bs;
  // The Helmholtz kernels are timed at wavenumber 1 (the shape does not depend on it).
  const double mu = kernel_needs_ctx(kernel) ? 1.0 : 0.0;
  for (int tpl : {1, 2, 4}) {
    if (tpl > kernel_max_tpl(kernel)) continue;
    for (int occ : {0, 3, 2}) {
      DenseShape sh;
      sh.tpl = tpl;
      sh.occ = occ;
      double ms = 0;
      auto run = [&]() {
        return precision == SCTL_AMD_FP64
                   ? launch_eval_f64(kernel, newton, r_trg, nt, r_src, n_src, v_src, ns, v_scratch, sh, mu, s)
                   : launch_eval_mixed(kernel, newton, r_trg, nt, r_src, n_src, v_src, ns, v_scratch, sh, mu, s);
      };
      SCTL_TRY(time_ms(run, s, 2, &ms));
      if (ms < best) {
        best = ms;
        bs = sh;
      }
    }
  }
  store(dense_key(kernel, precision, newton, nt, ns), {(double)bs.tpl, (double)bs.occ, best});
  if (best_tpl) *best_tpl = bs.tpl;
  if (best_occ) *best_occ = bs.occ;
  if (best_ms) *best_ms = best;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_tune(sctl_amd_plan plan, const double* v_src, double* v_scratch, void* stream, int* best_spl,
                       int* best_tpl, int* best_occ, double* best_ms) {
  SCTL_CHECK_ARG(plan, "plan is NULL");
  int64_t ns = 0, nt = 0, inter = 0;
  int k0 = 0, k1 = 0;
  SCTL_TRY(plan_sizes(plan, &ns, &nt, &k0, &k1, &inter));
  SCTL_CHECK_ARG(ns == 0 || v_src, "v_src is NULL");
  SCTL_CHECK_ARG(nt == 0 || v_scratch, "v_scratch is NULL");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  // Every candidate once (as configured: a kernel without two targets per
  // lane runs one), then the three fastest again, the best of that second
  // pass winning, so that a clock change during the search does not pick.
  struct Cand {
    double ms;
    PlanShape sh;
  };
  std::vector<Cand> cands;
  auto timed = [&](double* ms) {
    return time_ms([&]() { return sctl_amd_plan_apply(plan, v_src, v_scratch, 0, s); }, s, 3, ms);
  };
  for (int spl : {1, 2, 4}) {
    for (int tpl : {1, 2}) {
      for (int occ : {0, 4, 3, 2, 1}) {
        PlanShape sh;
        sh.spl = spl;
        sh.tpl = tpl;
        sh.occ = occ;
        SCTL_TRY(plan_configure(plan, sh));
        const PlanShape act = plan_shape(plan);
        bool seen = false;
        for (const Cand& c : cands) seen |= c.sh.spl == act.spl && c.sh.tpl == act.tpl && c.sh.occ == act.occ;
        if (seen) continue;
        double ms = 0;
        SCTL_TRY(timed(&ms));
        cands.push_back({ms, act});
      }
    }
  }
  std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.ms < b.ms; });
  const size_t top = std::min<size_t>(3, cands.size());
  double best = 1e300;
  PlanShape bs = cands[0].sh;
  for (size_t i = 0; i < top; i++) {
    SCTL_TRY(plan_configure(plan, cands[i].sh));
    double ms = 0;
    SCTL_TRY(timed(&ms));
    if (ms < best) {
      best = ms;
      bs = cands[i].sh;
    }
  }
  SCTL_TRY(plan_configure(plan, bs));
  store(plan_key(plan_kernel(plan), plan_precision(plan), inter, plan_nonempty_blocks(plan)),
        {(double)bs.spl, (double)bs.tpl, (double)bs.occ, best});
  if (best_spl) *best_spl = bs.spl;
  if (best_tpl) *best_tpl = bs.tpl;
  if (best_occ) *best_occ = bs.occ;
  if (best_ms) *best_ms = best;
  return SCTL_AMD_SUCCESS;
}

}  // extern "C"
