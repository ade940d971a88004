// Batched near-field blocks: the plan API of include/sctl_amd.h.
//
// BoundaryIntegralOp evaluates thousands of small, uneven (target-set,
// source-set) blocks (boundary_integral.txx:1009-1032 builds them, :1169-1238
// and :1342-1364 evaluate them one OpenMP task each).  Here all blocks of a
// plan run in ONE persistent launch:
//
//  * Work items are (block, chunk of <= 64*TPL targets).  Items are ordered by
//    cost Ns*Nt, largest first (longest-processing-time order), and every
//    wave of the persistent grid takes the next item from a global atomic
//    counter, so the tail of the launch is made of the smallest items.
//  * A wave owns an item; each lane holds TPL targets.  Chunks of <= 16*TPL or
//    <= 32*TPL targets use G = 4 or 2 lane groups (64/G lanes each): group g
//    takes every G-th source and the groups' sums are combined with
//    cross-lane shuffles, so a 16-target block still keeps 64 lanes busy.
//  * The block's sources are staged through a wave-private LDS region,
//    64*SPL records at a time (no workgroup barrier: waves run independently).
//  * Mixed precision: each pair is differenced in FP64 and the difference
//    rounded to FP32 (the Laplace single and double layers round r^2 and
//    f n.r, formed in FP64, instead), the rest of the interaction is FP32;
//    each stage's FP32 partial is added into FP64 sums with Kahan
//    compensation.  FP64: plain FP64 records, as in the dense path.
//  * A stage's source loop stops at the first multiple of 4 sources per
//    lane group past its real sources, so a block's last, partial stage
//    costs what it holds.
//
// The knobs (SPL: LDS stage size, TPL: targets per lane, a cap on resident
// workgroups per CU) are chosen per plan by the tuner (tune.cpp) or defaults.
// Every wave leaves the loop when the counter passes the number of items, so
// the grid always drains.  An apply allocates nothing and issues no memset
// (the counter is cleared by a one-lane kernel), so it can be captured in a
// hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <numeric>
#include <string>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace sctl_amd {

namespace {

constexpr int BWG = 256;  // 4 waves per workgroup, each wave independent

// Padding of a partial source stage: the slots past the block's last source
// repeat the stage's first source with zero density and normal.  Every
// kernel is linear in the density, so such a slot adds exactly 0 wherever
// the first source's own pairs are finite; a far-away padding point would
// instead overflow the FP32 differences of the mixed path (0 * inf = NaN,
// sending the stage down the masked retry) and send Helmholtz phases down
// the large-argument sincos branch.

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ int next_item(int* counter, int lane) {
  int it = 0;
  if (lane == 0) it = atomicAdd(counter, 1);
  return __builtin_amdgcn_readfirstlane(it);
}

__device__ __forceinline__ double shfl_xor_d(double v, int m) {
  const int lane = threadIdx.x & 63;
  return __shfl(v, lane ^ m, 64);
}

// Calls body(i) for the sources i = g, g + G, ... of lane group g, in steps
// of 4 per group, up to the first multiple of 4G at or past `len` (the
// stage's real sources; the staged padding covers the rest, and STAGE is a
// multiple of 4G, so every i < STAGE).  The group count G is a compile-time
// constant (the item's lane-group count is wave-uniform): the 4 unrolled
// calls address LDS with immediate offsets from one base.
template <int G, int STAGE, class Body> __device__ __forceinline__ void group_sources(int g, int len, Body&& body) {
  static_assert(STAGE % (4 * G) == 0, "the stage must hold whole steps");
  const int steps = (len + 4 * G - 1) / (4 * G);
  for (int k = 0; k < steps; k++) {
#pragma unroll
    for (int u = 0; u < 4; u++) body(g + (4 * k + u) * G);
  }
}
template <int STAGE, class Body> __device__ __forceinline__ void group_sources(int lg, int g, int len, Body&& body) {
  switch (lg) {
    case 0: group_sources<1, STAGE>(g, len, body); break;
    case 1: group_sources<2, STAGE>(g, len, body); break;
    default: group_sources<4, STAGE>(g, len, body); break;
  }
}

__global__ void k_reset_counter(int* counter) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *counter = 0;
}

// ------------------------------------------------------------- FP64 blocks
// The Laplace double layer with f n staged per source: z^3 (f n).r, one
// multiplication per pair fewer than n.r z^3 f.
template <int NEWTON, bool MASK>
__device__ __forceinline__ void lap_dxu_fn(double& acc, double rx, double ry, double rz, const double* fn) {
  const double r2 = fmad(rz, rz, fmad(ry, ry, rx * rx));
  double z = rsq_deferred<NEWTON>(r2);
  if constexpr (MASK) z = r2 > 0.0 ? z : 0.0;
  const double nd = fmad(rz, fn[2], fmad(ry, fn[1], rx * fn[0]));
  acc = fmad(nd, z * z * z, acc);
}

template <int KER, int NEWTON, int SPL, int TPL>
__global__ void __launch_bounds__(BWG)
k_blocks_f64(const Item* __restrict__ items, int nitems, int* __restrict__ counter,
             const int64_t* __restrict__ trg_dsp, const int64_t* __restrict__ src_dsp,
             const int32_t* __restrict__ src_cnt, const double* __restrict__ xt, const double* __restrict__ xs,
             const double* __restrict__ xn, const double* __restrict__ fs, double* __restrict__ out, double scale,
             double mu) {
  using K = Ker<KER>;
  constexpr int K0 = K::K0, K1 = K::K1, ND = K::ND, R = rec_f64<KER>(), STAGE = 64 * SPL;
  __shared__ double lds_all[(BWG / 64) * STAGE * R];
  const int lane = threadIdx.x & 63;
  double* lds = lds_all + (threadIdx.x >> 6) * (STAGE * R);

  for (int it = next_item(counter, lane); it < nitems; it = next_item(counter, lane)) {
    const Item I = items[it];
    const int lg = I.lg, G = 1 << lg, L = 64 >> lg;
    const int g = lane >> (6 - lg), tl = lane & (L - 1);
    const int64_t tb = trg_dsp[I.block] + I.t_off;
    bool tv[TPL];
    int64_t t[TPL];
    double tx[TPL], ty[TPL], tz[TPL], acc[TPL][K1];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
      tv[j] = tl + j * L < I.t_cnt;
      t[j] = tb + (tv[j] ? tl + j * L : 0);
      tx[j] = xt[t[j] * 3 + 0];
      ty[j] = xt[t[j] * 3 + 1];
      tz[j] = xt[t[j] * 3 + 2];
#pragma unroll
      for (int k = 0; k < K1; k++) acc[j][k] = 0.0;
    }
    const int64_t sb = src_dsp[I.block];
    const int nsb = src_cnt[I.block];

    for (int s0 = 0; s0 < nsb; s0 += STAGE) {
      const int len = min(STAGE, nsb - s0);
      wave_sync();
#pragma unroll
      for (int q = 0; q < SPL; q++) {
        const int i = lane + 64 * q;
        double* rec = lds + i * R;
        const bool real = i < len;
        const int64_t s = sb + s0 + (real ? i : 0);
        rec[0] = xs[s * 3 + 0];
        rec[1] = xs[s * 3 + 1];
        rec[2] = xs[s * 3 + 2];
        const double fn = KER == LAP_DXU && real ? fs[s] : 1.0;  // f n staged for the double layer
#pragma unroll
        for (int k = 0; k < ND; k++) rec[3 + k] = real ? fn * xn[s * 3 + k] : 0.0;
#pragma unroll
        for (int k = 0; k < K0; k++) rec[3 + ND + k] = real ? fs[s * K0 + k] : 0.0;
#pragma unroll
        for (int k = 3 + ND + K0; k < R; k++) rec[k] = 0.0;
      }
      wave_sync();
      double saved[TPL][K1];
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < K1; k++) saved[j][k] = acc[j][k];
      group_sources<STAGE>(lg, g, len, [&](int i) {
        const double* rec = lds + i * R;
        if constexpr (KER == LAP_DXU) {
#pragma unroll
          for (int j = 0; j < TPL; j++)
            lap_dxu_fn<NEWTON, false>(acc[j][0], tx[j] - rec[0], ty[j] - rec[1], tz[j] - rec[2], rec + 3);
        } else {
          double n[3] = {0, 0, 0}, f[K0];
#pragma unroll
          for (int k = 0; k < ND; k++) n[k] = rec[3 + k];
#pragma unroll
          for (int k = 0; k < K0; k++) f[k] = rec[3 + ND + k];
#pragma unroll
          for (int j = 0; j < TPL; j++)
            interact<KER, NEWTON, false>(acc[j], tx[j] - rec[0], ty[j] - rec[1], tz[j] - rec[2], n, f, mu);
        }
      });
      bool bad = false;
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < K1; k++) bad |= !__builtin_isfinite(acc[j][k]);
      if (bad) {
#pragma unroll
        for (int j = 0; j < TPL; j++)
#pragma unroll
          for (int k = 0; k < K1; k++) acc[j][k] = saved[j][k];
        for (int i = g; i < len; i += G) {
          const double* rec = lds + i * R;
          if constexpr (KER == LAP_DXU) {
#pragma unroll
            for (int j = 0; j < TPL; j++)
              lap_dxu_fn<NEWTON, true>(acc[j][0], tx[j] - rec[0], ty[j] - rec[1], tz[j] - rec[2], rec + 3);
          } else {
            double n[3] = {0, 0, 0}, f[K0];
            for (int k = 0; k < ND; k++) n[k] = rec[3 + k];
            for (int k = 0; k < K0; k++) f[k] = rec[3 + ND + k];
#pragma unroll
            for (int j = 0; j < TPL; j++)
              interact<KER, NEWTON, true>(acc[j], tx[j] - rec[0], ty[j] - rec[1], tz[j] - rec[2], n, f, mu);
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
      for (int m = L; m < 64; m <<= 1)
#pragma unroll
        for (int k = 0; k < K1; k++) acc[j][k] += shfl_xor_d(acc[j][k], m);
      if (g == 0 && tv[j]) {
#pragma unroll
        for (int k = 0; k < K1; k++) out[t[j] * K1 + k] += acc[j][k] * scale;
      }
    }
  }
}

// ------------------------------------------------------------ mixed blocks
// Near-field blocks put most targets close to their block's sources, so the
// mixed path differences every pair in FP64 and rounds the difference to
// FP32 (accurate to 2^-24 relative to r however small r is); the rest of the
// interaction is FP32 with v_rsq_f32.  The stage's FP32 partial is added into
// FP64 sums with Kahan compensation; a non-finite partial (r == 0 in the
// stage) is redone with the r^2 > 0 mask.
//
// The Laplace single and double layers need only r^2 and f n.r of each pair.
// The single layer forms r^2 in FP64 from the FP64 differences (FP64 FMAs
// issue at the FP32 rate on CDNA4) and rounds it once: 6 FP64 operations and
// 1 conversion.  The double layer rounds its three FP64 differences and forms
// r^2 and f n.r in FP32 from f n staged in FP32: 3 FP64 operations, 3
// conversions and 6 FP32 operations, which can pair with other waves' FP32
// work where FP64 cannot (4 % faster on the 4096 torus blocks than forming
// both in FP64; the single layer is 3.5 % slower that way).
template <int KER> constexpr bool lap_r2d() { return KER == LAP_FXU || KER == LAP_DXU; }

// One Laplace pair: pr = the source position (FP64), ar = {f} (FxU) or
// {f n_x, f n_y, f n_z} (DxU) in FP32.  z carries the same power as in
// interact(), so the deferred-Newton constant folds the same way.
template <int KER, int NEWTON, bool MASK>
__device__ __forceinline__ void lap_pair_r2d(float& p, double tx, double ty, double tz, const double* pr,
                                             const float* ar) {
  if constexpr (KER == LAP_FXU) {
    const double dx = tx - pr[0], dy = ty - pr[1], dz = tz - pr[2];
    const float r2 = (float)fmad(dz, dz, fmad(dy, dy, dx * dx));
    float z = rsq_deferred<NEWTON>(r2);
    if constexpr (MASK) z = r2 > 0.0f ? z : 0.0f;
    p = fmad(z, ar[0], p);
  } else {
    const float dx = (float)(tx - pr[0]), dy = (float)(ty - pr[1]), dz = (float)(tz - pr[2]);
    const float r2 = fmad(dz, dz, fmad(dy, dy, dx * dx));
    float z = rsq_deferred<NEWTON>(r2);
    if constexpr (MASK) z = r2 > 0.0f ? z : 0.0f;
    const float nd = fmad(dz, ar[2], fmad(dy, ar[1], dx * ar[0]));
    p = fmad(nd * z, z * z, p);
  }
}

template <int KER, int NEWTON, int SPL, int TPL>
__global__ void __launch_bounds__(BWG)
k_blocks_mixed(const Item* __restrict__ items, int nitems, int* __restrict__ counter,
               const int64_t* __restrict__ trg_dsp, const int64_t* __restrict__ src_dsp,
               const int32_t* __restrict__ src_cnt, const double* __restrict__ xt, const double* __restrict__ xs,
               const double* __restrict__ xn, const double* __restrict__ fs, double* __restrict__ out,
               double scale, double ctx) {
  const float mu = (float)ctx;
  using K = Ker<KER>;
  constexpr int K0 = K::K0, K1 = K::K1, ND = K::ND, STAGE = 64 * SPL;
  constexpr bool R2D = lap_r2d<KER>();
  constexpr int PW = 4;                                          // doubles per source: s, padded
  constexpr int AW = KER == LAP_DXU ? 4 : ((ND + K0) + 3) & ~3;  // float words: f n (DxU), else normal, density
  __shared__ double pos_all[(BWG / 64) * STAGE * PW];
  __shared__ float aux_all[(BWG / 64) * STAGE * AW];
  const int lane = threadIdx.x & 63;
  double* pos = pos_all + (threadIdx.x >> 6) * (STAGE * PW);
  float* aux = aux_all + (threadIdx.x >> 6) * (STAGE * AW);

  for (int it = next_item(counter, lane); it < nitems; it = next_item(counter, lane)) {
    const Item I = items[it];
    const int lg = I.lg, G = 1 << lg, L = 64 >> lg;
    const int g = lane >> (6 - lg), tl = lane & (L - 1);
    const int64_t tb = trg_dsp[I.block] + I.t_off;
    bool tv[TPL];
    int64_t t[TPL];
    double tx[TPL], ty[TPL], tz[TPL], acc[TPL][K1], comp[TPL][K1];
#pragma unroll
    for (int j = 0; j < TPL; j++) {
      tv[j] = tl + j * L < I.t_cnt;
      t[j] = tb + (tv[j] ? tl + j * L : 0);
      tx[j] = xt[t[j] * 3 + 0];
      ty[j] = xt[t[j] * 3 + 1];
      tz[j] = xt[t[j] * 3 + 2];
#pragma unroll
      for (int k = 0; k < K1; k++) acc[j][k] = comp[j][k] = 0.0;
    }
    const int64_t sb = src_dsp[I.block];
    const int nsb = src_cnt[I.block];

    for (int s0 = 0; s0 < nsb; s0 += STAGE) {
      const int len = min(STAGE, nsb - s0);
      wave_sync();
#pragma unroll
      for (int q = 0; q < SPL; q++) {
        const int i = lane + 64 * q;
        double* pr = pos + i * PW;
        float* ar = aux + i * AW;
        const bool real = i < len;
        const int64_t s = sb + s0 + (real ? i : 0);
        pr[0] = xs[s * 3 + 0];
        pr[1] = xs[s * 3 + 1];
        pr[2] = xs[s * 3 + 2];
        if constexpr (KER == LAP_DXU) {
          const double f = real ? fs[s] : 0.0;
#pragma unroll
          for (int k = 0; k < 3; k++) ar[k] = (float)(f * xn[s * 3 + k]);
          ar[3] = 0.0f;
        } else {
#pragma unroll
          for (int k = 0; k < ND; k++) ar[k] = real ? (float)xn[s * 3 + k] : 0.0f;
#pragma unroll
          for (int k = 0; k < K0; k++) ar[ND + k] = real ? (float)fs[s * K0 + k] : 0.0f;
#pragma unroll
          for (int k = ND + K0; k < AW; k++) ar[k] = 0.0f;
        }
      }
      wave_sync();
      float p[TPL][K1];
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < K1; k++) p[j][k] = 0.0f;
      group_sources<STAGE>(lg, g, len, [&](int i) {
        const double* pr = pos + i * PW;
        const float* ar = aux + i * AW;
        if constexpr (R2D) {
#pragma unroll
          for (int j = 0; j < TPL; j++) lap_pair_r2d<KER, NEWTON, false>(p[j][0], tx[j], ty[j], tz[j], pr, ar);
        } else {
          float n[3] = {0, 0, 0}, f[K0];
#pragma unroll
          for (int k = 0; k < ND; k++) n[k] = ar[k];
#pragma unroll
          for (int k = 0; k < K0; k++) f[k] = ar[ND + k];
#pragma unroll
          for (int j = 0; j < TPL; j++)
            interact<KER, NEWTON, false>(p[j], (float)(tx[j] - pr[0]), (float)(ty[j] - pr[1]),
                                         (float)(tz[j] - pr[2]), n, f, mu);
        }
      });
      bool bad = false;
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < K1; k++) bad |= !__builtin_isfinite(p[j][k]);
      if (bad) {
#pragma unroll
        for (int j = 0; j < TPL; j++)
#pragma unroll
          for (int k = 0; k < K1; k++) p[j][k] = 0.0f;
        for (int i = g; i < len; i += G) {
          const double* pr = pos + i * PW;
          const float* ar = aux + i * AW;
          if constexpr (R2D) {
#pragma unroll
            for (int j = 0; j < TPL; j++) lap_pair_r2d<KER, NEWTON, true>(p[j][0], tx[j], ty[j], tz[j], pr, ar);
          } else {
            float n[3] = {0, 0, 0}, f[K0];
            for (int k = 0; k < ND; k++) n[k] = ar[k];
            for (int k = 0; k < K0; k++) f[k] = ar[ND + k];
#pragma unroll
            for (int j = 0; j < TPL; j++)
              interact<KER, NEWTON, true>(p[j], (float)(tx[j] - pr[0]), (float)(ty[j] - pr[1]),
                                          (float)(tz[j] - pr[2]), n, f, mu);
          }
        }
      }
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < K1; k++) {  // Kahan: acc += p
          const double y = (double)p[j][k] - comp[j][k];
          const double s = acc[j][k] + y;
          comp[j][k] = (s - acc[j][k]) - y;
          acc[j][k] = s;
        }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++) {
      double v[K1];
#pragma unroll
      for (int k = 0; k < K1; k++) v[k] = acc[j][k] - comp[j][k];
      for (int m = L; m < 64; m <<= 1)
#pragma unroll
        for (int k = 0; k < K1; k++) v[k] += shfl_xor_d(v[k], m);
      if (g == 0 && tv[j]) {
#pragma unroll
        for (int k = 0; k < K1; k++) out[t[j] * K1 + k] += v[k] * scale;
      }
    }
  }
}

using BlocksFn = void (*)(const Item*, int, int*, const int64_t*, const int64_t*, const int32_t*, const double*,
                          const double*, const double*, const double*, double*, double, double);

// Targets per lane: 1 or 2 (2 only where max_tpl allows it).
template <int KER, int NEWTON, int SPL> BlocksFn f64_blocks_t(int tpl) {
  if constexpr (max_tpl<KER>() >= 2) {
    if (tpl >= 2) return k_blocks_f64<KER, NEWTON, SPL, 2>;
  }
  return k_blocks_f64<KER, NEWTON, SPL, 1>;
}
template <int KER, int NEWTON> BlocksFn f64_blocks_s(int spl, int tpl) {
  return spl >= 4 ? f64_blocks_t<KER, NEWTON, 4>(tpl)
                  : (spl >= 2 ? f64_blocks_t<KER, NEWTON, 2>(tpl) : f64_blocks_t<KER, NEWTON, 1>(tpl));
}
BlocksFn f64_blocks(int ker, int newton, int spl, int tpl) {
  return with_kernel(ker, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    return newton == 0 ? f64_blocks_s<K, 0>(spl, tpl)
                       : (newton == 1 ? f64_blocks_s<K, 1>(spl, tpl) : f64_blocks_s<K, 2>(spl, tpl));
  });
}
template <int KER, int NEWTON, int SPL> BlocksFn mix_blocks_t(int tpl) {
  if constexpr (max_tpl<KER>() >= 2) {
    if (tpl >= 2) return k_blocks_mixed<KER, NEWTON, SPL, 2>;
  }
  return k_blocks_mixed<KER, NEWTON, SPL, 1>;
}
template <int KER, int NEWTON> BlocksFn mix_blocks_s(int spl, int tpl) {
  return spl >= 4 ? mix_blocks_t<KER, NEWTON, 4>(tpl)
                  : (spl >= 2 ? mix_blocks_t<KER, NEWTON, 2>(tpl) : mix_blocks_t<KER, NEWTON, 1>(tpl));
}
BlocksFn mix_blocks(int ker, int newton, int spl, int tpl) {
  return with_kernel(ker, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    return newton ? mix_blocks_s<K, 1>(spl, tpl) : mix_blocks_s<K, 0>(spl, tpl);
  });
}
int plan_max_tpl(int ker) {
  return with_kernel(ker, [](auto kc) { return max_tpl<decltype(kc)::value>(); });
}
// Static LDS bytes of one workgroup of a blocks kernel.
size_t blocks_lds(int ker, int precision, int spl) {
  int K0, K1, ND;
  kernel_dims(ker, &K0, &K1, &ND);
  const size_t stage = 64 * (size_t)spl * (BWG / 64);
  if (precision == SCTL_AMD_MIXED) {
    if (ker == LAP_DXU) return stage * (6 * sizeof(double) + sizeof(float));  // s, f n; one unused float
    return stage * (4 * sizeof(double) + (((ND + K0) + 3) & ~3) * sizeof(float));
  }
  return stage * (((3 + ND + K0) + 1) & ~1) * sizeof(double);
}
}  // namespace
}  // namespace sctl_amd

using namespace sctl_amd;

namespace sctl_amd {

namespace {

// Work items for `tpl` targets per lane: chunks of <= 64*tpl targets, lane
// groups for small chunks, longest (Ns x lanes) first.
std::vector<Item> build_items(const sctl_amd_plan_s* p, int tpl) {
  std::vector<Item> items;
  std::vector<int64_t> cost;
  const int64_t chunk = 64LL * tpl;
  for (int64_t b = 0; b < p->nblocks; b++) {
    const int64_t tc = p->trg_cnt[b], sc = p->src_cnt[b];
    if (!tc || !sc) continue;
    for (int64_t t0 = 0; t0 < tc; t0 += chunk) {
      const int cnt = (int)std::min<int64_t>(chunk, tc - t0);
      const int lg = cnt <= 16 * tpl ? 2 : (cnt <= 32 * tpl ? 1 : 0);
      items.push_back(Item{(int32_t)b, (int32_t)t0, cnt, lg});
      cost.push_back((int64_t)(64 >> lg) * tpl * sc);
    }
  }
  std::vector<size_t> idx(items.size());
  std::iota(idx.begin(), idx.end(), 0);
  std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
  std::vector<Item> sorted(items.size());
  for (size_t i = 0; i < idx.size(); i++) sorted[i] = items[idx[i]];
  return sorted;
}

BlocksFn plan_fn(const sctl_amd_plan_s* p, const PlanShape& s) {
  return p->precision == SCTL_AMD_MIXED ? mix_blocks(p->kernel, p->newton, s.spl, s.tpl)
                                        : f64_blocks(p->kernel, p->newton, s.spl, s.tpl);
}

}  // namespace

// Applies knobs to a plan: rebuilds the work items for the targets per lane
// and sizes the persistent grid from the occupancy calculator (with the cap).
int plan_configure(sctl_amd_plan_s* p, PlanShape s) {
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  s.spl = s.spl >= 4 ? 4 : (s.spl >= 2 ? 2 : 1);
  s.tpl = s.tpl >= 2 && plan_max_tpl(p->kernel) >= 2 ? 2 : 1;
  s.occ = std::max(0, s.occ);
  if (!p->d_items || s.tpl != p->shape.tpl) {
    const std::vector<Item> items = build_items(p, s.tpl);
    SCTL_CHECK_ARG(items.size() < 0x7fffffffULL, "too many work items");
    if (p->d_items) SCTL_HIP_TRY(hipFree(p->d_items));
    p->d_items = nullptr;
    SCTL_TRY(dev_upload(&p->d_items, items));
    p->nitems = (int)items.size();
  }
  const BlocksFn fn = plan_fn(p, s);
  size_t dyn = 0;
  SCTL_TRY(occupancy_lds((const void*)fn, s.occ, blocks_lds(p->kernel, p->precision, s.spl), &dyn));
  int per_cu = 0;
  SCTL_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)fn, BWG, dyn));
  p->grid = std::max(1, per_cu) * dev.num_cus;
  p->dyn_lds = dyn;
  p->shape = s;
  p->generation++;
  return SCTL_AMD_SUCCESS;
}

int plan_sizes(sctl_amd_plan p, int64_t* ns, int64_t* nt, int* K0, int* K1, int64_t* interactions) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  *ns = p->ns_total;
  *nt = p->nt_total;
  *K0 = p->K0;
  *K1 = p->K1;
  *interactions = p->interactions;
  return SCTL_AMD_SUCCESS;
}
int plan_kernel(sctl_amd_plan p) { return p ? p->kernel : -1; }
uint64_t plan_generation(sctl_amd_plan p) { return p ? p->generation : 0; }
int plan_precision(sctl_amd_plan p) { return p ? p->precision : -1; }
int64_t plan_nonempty_blocks(sctl_amd_plan p) { return p ? p->nonempty : 0; }
PlanShape plan_shape(sctl_amd_plan p) { return p ? p->shape : PlanShape(); }

int plan_launch(sctl_amd_plan_s* p, const double* v_src, double* v_trg, hipStream_t s) {
  if (p->nitems == 0) return SCTL_AMD_SUCCESS;
  hipLaunchKernelGGL(k_reset_counter, dim3(1), dim3(64), 0, s, p->d_counter);
  SCTL_HIP_TRY(hipGetLastError());
  const double scale = kernel_scale(p->kernel) * fold_factor(p->kernel, p->newton);
  const int grid = (int)std::min<int64_t>(p->grid, ((int64_t)p->nitems + 3) / 4);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  SCTL_HIP_TRY(hipStreamIsCapturing(s, &cs));
  if (cs == hipStreamCaptureStatusNone) profile_count(p->kernel, p->interactions);  // graphs count per launch
  ProfileScope prof(s);
  hipLaunchKernelGGL(plan_fn(p, p->shape), dim3(grid), dim3(BWG), p->dyn_lds, s, p->d_items, p->nitems, p->d_counter,
                     p->d_trg_dsp, p->d_src_dsp, p->d_src_cnt, p->d_xt, p->d_xs, p->d_xn, v_src, v_trg, scale, p->ctx);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

}  // namespace sctl_amd

extern "C" {

int sctl_amd_plan_create(sctl_amd_plan* plan, int kernel, int precision, int digits, int64_t nblocks,
                         const int64_t* trg_dsp, const int64_t* trg_cnt, const int64_t* src_dsp,
                         const int64_t* src_cnt, int64_t nt_total, int64_t ns_total) {
  SCTL_CHECK_ARG(plan, "plan is NULL");
  *plan = nullptr;
  int K0, K1, ND;
  if (kernel_dims(kernel, &K0, &K1, &ND) != SCTL_AMD_SUCCESS) {
    set_error("unknown kernel id " + std::to_string(kernel));
    return SCTL_AMD_ERROR_UNKNOWN_KERNEL;
  }
  SCTL_CHECK_ARG(precision >= SCTL_AMD_FP64 && precision <= SCTL_AMD_AUTO, "unknown precision");
  SCTL_CHECK_ARG(digits >= -1, "digits must be -1 or >= 0");
  SCTL_CHECK_ARG(nblocks >= 0 && nt_total >= 0 && ns_total >= 0, "negative size");
  SCTL_CHECK_ARG(nblocks == 0 || (trg_dsp && trg_cnt && src_dsp && src_cnt), "NULL block array");
  SCTL_CHECK_ARG(nblocks < 0x7fffffffLL, "too many blocks");
  if (precision == SCTL_AMD_AUTO) precision = (digits >= 0 && digits <= 7) ? SCTL_AMD_MIXED : SCTL_AMD_FP64;

  std::vector<int64_t> order(nblocks);
  std::iota(order.begin(), order.end(), 0);
  for (int64_t b = 0; b < nblocks; b++) {
    SCTL_CHECK_ARG(trg_cnt[b] >= 0 && src_cnt[b] >= 0 && trg_dsp[b] >= 0 && src_dsp[b] >= 0,
                   "negative block offset or count (block " + std::to_string(b) + ")");
    SCTL_CHECK_ARG(trg_dsp[b] + trg_cnt[b] <= nt_total,
                   "block target range exceeds nt_total (block " + std::to_string(b) + ")");
    SCTL_CHECK_ARG(src_dsp[b] + src_cnt[b] <= ns_total,
                   "block source range exceeds ns_total (block " + std::to_string(b) + ")");
    SCTL_CHECK_ARG(src_cnt[b] <= 0x7fffffffLL && trg_cnt[b] <= 0x7fffffffLL, "block too large");
  }
  std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return trg_dsp[a] < trg_dsp[b]; });
  for (size_t i = 1; i < order.size(); i++) {
    const int64_t a = order[i - 1], b = order[i];
    if (trg_cnt[a] && trg_cnt[b])
      SCTL_CHECK_ARG(trg_dsp[a] + trg_cnt[a] <= trg_dsp[b], "target ranges of blocks overlap");
  }

  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  auto* p = new sctl_amd_plan_s();
  p->kernel = kernel;
  p->precision = precision;
  p->newton = precision == SCTL_AMD_MIXED ? 0 : newton_steps_f64(digits);
  p->K0 = K0;
  p->K1 = K1;
  p->ND = ND;
  p->device = dev.device;
  p->nblocks = nblocks;
  p->nt_total = nt_total;
  p->ns_total = ns_total;
  p->trg_dsp.assign(trg_dsp, trg_dsp + nblocks);
  p->trg_cnt.assign(trg_cnt, trg_cnt + nblocks);
  p->src_dsp.assign(src_dsp, src_dsp + nblocks);
  p->src_cnt.assign(src_cnt, src_cnt + nblocks);
  for (int64_t b = 0; b < nblocks; b++) {
    p->interactions += trg_cnt[b] * src_cnt[b];
    p->nonempty += (trg_cnt[b] && src_cnt[b]) ? 1 : 0;
  }
  std::vector<int64_t> v_td(trg_dsp, trg_dsp + nblocks), v_sd(src_dsp, src_dsp + nblocks);
  std::vector<int32_t> v_sc(nblocks), v_tc(nblocks);
  for (int64_t b = 0; b < nblocks; b++) {
    v_sc[b] = (int32_t)src_cnt[b];
    v_tc[b] = (int32_t)trg_cnt[b];
  }
  int rc = SCTL_AMD_SUCCESS;
  if (!rc) rc = dev_upload(&p->d_trg_dsp, v_td);
  if (!rc) rc = dev_upload(&p->d_src_dsp, v_sd);
  if (!rc) rc = dev_upload(&p->d_src_cnt, v_sc);
  if (!rc) rc = dev_upload(&p->d_trg_cnt, v_tc);
  if (!rc) {
    hipError_t e = hipMalloc(&p->d_counter, sizeof(int));
    if (e != hipSuccess) {
      set_error(std::string("hipMalloc(counter): ") + hipGetErrorString(e));
      rc = SCTL_AMD_ERROR_OUT_OF_MEMORY;
    }
  }
  if (!rc) {
    PlanShape s;
    if (!lookup_plan_shape(kernel, precision, p->interactions, p->nonempty, &s)) {
      // Untuned default, from sweeps of 4096- and 2048-block surface
      // batches (Laplace and Stokes double layers) on MI355X: 256-source
      // stages and two targets per lane; at most two workgroups per CU for
      // mixed scalar outputs, four otherwise.
      s.spl = 4;
      s.tpl = 2;
      s.occ = (precision == SCTL_AMD_MIXED && K1 == 1) ? 2 : 4;
    }
    rc = plan_configure(p, s);
  }
  if (rc) {
    delete p;
    return rc;
  }
  *plan = p;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_set_geometry(sctl_amd_plan p, const double* r_trg, const double* r_src, const double* n_src,
                               uint32_t flags, void* stream) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(p->nt_total == 0 || r_trg, "r_trg is NULL");
  SCTL_CHECK_ARG(p->ns_total == 0 || r_src, "r_src is NULL");
  SCTL_CHECK_ARG(p->ns_total == 0 || !p->ND || n_src, "n_src is NULL for a kernel with a normal");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  const hipMemcpyKind kind = (flags & SCTL_AMD_HOST_PTRS) ? hipMemcpyHostToDevice : hipMemcpyDeviceToDevice;
  auto copy = [&](double** dst, const double* src, int64_t n) -> int {
    if (n <= 0) return SCTL_AMD_SUCCESS;
    if (!*dst) SCTL_HIP_TRY(hipMalloc(dst, sizeof(double) * n));
    SCTL_HIP_TRY(hipMemcpyAsync(*dst, src, sizeof(double) * n, kind, s));
    return SCTL_AMD_SUCCESS;
  };
  SCTL_TRY(copy(&p->d_xt, r_trg, p->nt_total * 3));
  SCTL_TRY(copy(&p->d_xs, r_src, p->ns_total * 3));
  if (p->ND) SCTL_TRY(copy(&p->d_xn, n_src, p->ns_total * 3));
  if (flags & SCTL_AMD_HOST_PTRS) SCTL_HIP_TRY(hipStreamSynchronize(s));
  p->has_geometry = true;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_set_ctx(sctl_amd_plan p, const double* ctx) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  SCTL_CHECK_ARG(!ctx || std::isfinite(*ctx), "the context value is not finite");
  p->has_ctx = ctx != nullptr;
  p->ctx = ctx ? *ctx : 0.0;
  p->generation++;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_apply(sctl_amd_plan p, const double* v_src, double* v_trg, uint32_t flags, void* stream) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  SCTL_CHECK_ARG(p->has_geometry, "plan has no geometry (call sctl_amd_plan_set_geometry first)");
  SCTL_CHECK_ARG(!needs_ctx(p->kernel) || p->has_ctx,
                 "the kernel needs a wavenumber (call sctl_amd_plan_set_ctx first)");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(p->ns_total == 0 || v_src, "v_src is NULL");
  SCTL_CHECK_ARG(p->nt_total == 0 || v_trg, "v_trg is NULL");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  if (!(flags & SCTL_AMD_HOST_PTRS)) return plan_launch(p, v_src, v_trg, s);
  Scratch bs, bt;
  SCTL_TRY(bs.alloc(sizeof(double) * p->ns_total * p->K0, s));
  SCTL_TRY(bt.alloc(sizeof(double) * p->nt_total * p->K1, s));
  if (p->ns_total) SCTL_HIP_TRY(hipMemcpyAsync(bs.as<double>(), v_src, bs.bytes(), hipMemcpyHostToDevice, s));
  if (p->nt_total) SCTL_HIP_TRY(hipMemcpyAsync(bt.as<double>(), v_trg, bt.bytes(), hipMemcpyHostToDevice, s));
  SCTL_TRY(plan_launch(p, bs.as<double>(), bt.as<double>(), s));
  if (p->nt_total) SCTL_HIP_TRY(hipMemcpyAsync(v_trg, bt.as<double>(), bt.bytes(), hipMemcpyDeviceToHost, s));
  bs.release();
  bt.release();
  SCTL_HIP_TRY(hipStreamSynchronize(s));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_interactions(sctl_amd_plan p, int64_t* interactions) {
  SCTL_CHECK_ARG(p && interactions, "NULL argument");
  *interactions = p->interactions;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_get_shape(sctl_amd_plan p, int* sources_per_lane, int* targets_per_lane, int* wg_per_cu_cap) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  if (sources_per_lane) *sources_per_lane = p->shape.spl;
  if (targets_per_lane) *targets_per_lane = p->shape.tpl;
  if (wg_per_cu_cap) *wg_per_cu_cap = p->shape.occ;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_set_shape(sctl_amd_plan p, int sources_per_lane, int targets_per_lane, int wg_per_cu_cap) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  SCTL_CHECK_ARG(sources_per_lane == 1 || sources_per_lane == 2 || sources_per_lane == 4,
                 "sources_per_lane must be 1, 2 or 4");
  SCTL_CHECK_ARG(targets_per_lane == 1 || targets_per_lane == 2, "targets_per_lane must be 1 or 2");
  SCTL_CHECK_ARG(wg_per_cu_cap >= 0, "wg_per_cu_cap must be >= 0");
  PlanShape s;
  s.spl = sources_per_lane;
  s.tpl = targets_per_lane;
  s.occ = wg_per_cu_cap;
  return plan_configure(p, s);
}

int sctl_amd_plan_destroy(sctl_amd_plan p) {
  delete p;
  return SCTL_AMD_SUCCESS;
}

}  // extern "C"
