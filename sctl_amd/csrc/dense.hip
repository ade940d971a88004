// Dense direct summation on gfx950: FP64 Eval and KernelMatrix.
//
// FP64 Eval (reference: generic-kernel.txx:102-211).  A workgroup of 256 lanes
// owns 256*TPL targets (TPL targets per lane, held in registers) and walks a
// contiguous range of sources in tiles of TILE sources staged through LDS;
// every lane of a wave reads the same source record (an LDS broadcast).  When
// there are too few target blocks to fill the 256 CUs, the source range is
// split over gridDim.y and the per-split sums are reduced in split order by a
// second kernel.  No MFMA: a pairwise sum is not a contraction.
//
// FP64 path: the accumulator runs over the sources in index order, as the
// reference does; per tile the accumulator is saved, and if the tile made it
// non-finite (some r2 == 0 in that tile) it is restored and the tile redone
// with the r2 > 0 mask.
//
// The mixed-precision Eval is in mixed.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "kernels.hpp"

namespace sctl_amd {

constexpr int WG = 256;
constexpr int TILE = 256;

// Far-away padding for the unused slots of a partial tile: the contribution is
// a finite, vanishing kernel value times a zero density.
constexpr double PAD_F64 = 1e100;


// ------------------------------------------------------------------ FP64 Eval
template <int KER, int TPL, int NEWTON>
__global__ void __launch_bounds__(WG)
k_eval_f64(const double* __restrict__ xt, int64_t nt, const double* __restrict__ xs, const double* __restrict__ xn,
           const double* __restrict__ fs, int64_t ns, int64_t chunk, double* __restrict__ out, int direct,
           double scale, double mu) {
  using K = Ker<KER>;
  constexpr int K0 = K::K0, K1 = K::K1, ND = K::ND, R = rec_f64<KER>();
  __shared__ double lds[TILE * R];

  const int tid = threadIdx.x;
  const SplitCoord wg = split_coord((nt + WG * TPL - 1) / (WG * TPL));
  const int64_t s_begin = wg.split * chunk;
  const int64_t s_end = s_begin + chunk < ns ? s_begin + chunk : ns;
  const int64_t t0 = wg.block * (WG * TPL) + tid;

  double tx[TPL], ty[TPL], tz[TPL], acc[TPL][K1];
#pragma unroll
  for (int j = 0; j < TPL; j++) {
    const int64_t t = t0 + (int64_t)j * WG;
    const bool ok = t < nt;
    tx[j] = ok ? xt[t * 3 + 0] : 0.0;
    ty[j] = ok ? xt[t * 3 + 1] : 0.0;
    tz[j] = ok ? xt[t * 3 + 2] : 0.0;
#pragma unroll
    for (int k = 0; k < K1; k++) acc[j][k] = 0.0;
  }

  for (int64_t sb = s_begin; sb < s_end; sb += TILE) {
    const int len = (int)(s_end - sb < TILE ? s_end - sb : TILE);
    __syncthreads();
    for (int i = tid; i < TILE; i += WG) {
      double* rec = lds + i * R;
      if (i < len) {
        const int64_t s = sb + i;
        rec[0] = xs[s * 3 + 0];
        rec[1] = xs[s * 3 + 1];
        rec[2] = xs[s * 3 + 2];
#pragma unroll
        for (int k = 0; k < ND; k++) rec[3 + k] = xn[s * 3 + k];
#pragma unroll
        for (int k = 0; k < K0; k++) rec[3 + ND + k] = fs[s * K0 + k];
      } else {
        rec[0] = rec[1] = rec[2] = PAD_F64;
#pragma unroll
        for (int k = 3; k < R; k++) rec[k] = 0.0;
      }
    }
    __syncthreads();

    double saved[TPL][K1];
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
      for (int k = 0; k < K1; k++) saved[j][k] = acc[j][k];

#pragma unroll 2
    for (int i = 0; i < TILE; i++) {
      const double* rec = lds + i * R;
      const double sx = rec[0], sy = rec[1], sz = rec[2];
      double n[3] = {0, 0, 0}, f[K0];
#pragma unroll
      for (int k = 0; k < ND; k++) n[k] = rec[3 + k];
#pragma unroll
      for (int k = 0; k < K0; k++) f[k] = rec[3 + ND + k];
#pragma unroll
      for (int j = 0; j < TPL; j++) interact<KER, NEWTON, false>(acc[j], tx[j] - sx, ty[j] - sy, tz[j] - sz, n, f, mu);
    }

    bool bad = false;
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
      for (int k = 0; k < K1; k++) bad |= !__builtin_isfinite(acc[j][k]);
    if (bad) {  // some r2 == 0 in this tile (or non-finite input): redo it masked
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < K1; k++) acc[j][k] = saved[j][k];
      for (int i = 0; i < len; i++) {
        const double* rec = lds + i * R;
        double n[3] = {0, 0, 0}, f[K0];
        for (int k = 0; k < ND; k++) n[k] = rec[3 + k];
        for (int k = 0; k < K0; k++) f[k] = rec[3 + ND + k];
#pragma unroll
        for (int j = 0; j < TPL; j++)
          interact<KER, NEWTON, true>(acc[j], tx[j] - rec[0], ty[j] - rec[1], tz[j] - rec[2], n, f, mu);
      }
    }
  }

#pragma unroll
  for (int j = 0; j < TPL; j++) {
    const int64_t t = t0 + (int64_t)j * WG;
    if (t < nt) {
#pragma unroll
      for (int k = 0; k < K1; k++) {
        if (direct) out[t * K1 + k] += acc[j][k] * scale;
        else out[wg.split * nt * K1 + t * K1 + k] = acc[j][k];
      }
    }
  }
}

// v_trg[i] += scale * sum_{y < nsplit} part[y][i]; one lane quad per element
// (sum_splits_quad).
__global__ void __launch_bounds__(WG)
k_reduce_splits(const double* __restrict__ part, int64_t n, int nsplit, double* __restrict__ out, double scale) {
  const int64_t i = ((int64_t)blockIdx.x * WG + threadIdx.x) >> 2;
  const double s = sum_splits_quad(part, n, nsplit, i, i < n);
  if (i < n && (threadIdx.x & 3) == 0) out[i] += s * scale;
}

// ------------------------------------------------------------- KernelMatrix
// M[(s*K0+k0)*ld + t*K1+k1] = scale * u[k0][k1]; one lane per target, the
// workgroup's gridDim.y slices stride over the sources.
template <int KER, int NEWTON>
__global__ void __launch_bounds__(WG)
k_kernel_matrix(const double* __restrict__ xt, int64_t nt, const double* __restrict__ xs,
                const double* __restrict__ xn, int64_t ns, double* __restrict__ M, double scale, double mu) {
  using K = Ker<KER>;
  constexpr int K0 = K::K0, K1 = K::K1, ND = K::ND;
  const int64_t t = (int64_t)blockIdx.x * WG + threadIdx.x;
  if (t >= nt) return;
  const int64_t ld = nt * K1;
  const double tx = xt[t * 3 + 0], ty = xt[t * 3 + 1], tz = xt[t * 3 + 2];
  for (int64_t s = blockIdx.y; s < ns; s += gridDim.y) {
    double n[3] = {0, 0, 0};
#pragma unroll
    for (int k = 0; k < ND; k++) n[k] = xn[s * 3 + k];
    double u[K0][K1];
    kmatrix<KER, NEWTON>(u, tx - xs[s * 3 + 0], ty - xs[s * 3 + 1], tz - xs[s * 3 + 2], n, mu);
#pragma unroll
    for (int k0 = 0; k0 < K0; k0++)
#pragma unroll
      for (int k1 = 0; k1 < K1; k1++) M[(s * K0 + k0) * ld + t * K1 + k1] = u[k0][k1] * scale;
  }
}

// --------------------------------------------------------------- dispatch
namespace {

using EvalF64Fn = void (*)(const double*, int64_t, const double*, const double*, const double*, int64_t, int64_t,
                           double*, int, double, double);
using KMatFn = void (*)(const double*, int64_t, const double*, const double*, int64_t, double*, double, double);

constexpr int TPLS[3] = {1, 2, 4};
int tpl_index(int tpl) { return tpl >= 4 ? 2 : (tpl >= 2 ? 1 : 0); }

template <int KER, int NEWTON> EvalF64Fn f64_fn(int ti) {
  constexpr int MT = max_tpl<KER>();
  if (ti >= 2 && MT >= 4) return k_eval_f64<KER, (MT >= 4 ? 4 : 1), NEWTON>;
  if (ti >= 1 && MT >= 2) return k_eval_f64<KER, (MT >= 2 ? 2 : 1), NEWTON>;
  return k_eval_f64<KER, 1, NEWTON>;
}
EvalF64Fn f64_fn(int ker, int newton, int ti) {
  return with_kernel(ker, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    return newton == 0 ? f64_fn<K, 0>(ti) : (newton == 1 ? f64_fn<K, 1>(ti) : f64_fn<K, 2>(ti));
  });
}
int rec_words_f64(int ker) {
  return with_kernel(ker, [](auto kc) { return rec_f64<decltype(kc)::value>(); });
}
int max_tpl_index(int ker) {
  return with_kernel(ker, [](auto kc) { return tpl_index(max_tpl<decltype(kc)::value>()); });
}
KMatFn kmat_fn(int ker, int newton) {
  return with_kernel(ker, [&](auto kc) -> KMatFn {
    constexpr int K = decltype(kc)::value;
    return newton == 0 ? k_kernel_matrix<K, 0> : (newton == 1 ? k_kernel_matrix<K, 1> : k_kernel_matrix<K, 2>);
  });
}

// Splits the source range over gridDim.y when the target blocks alone cannot
// fill the device (aim for >= 8 workgroups per CU, each split >= 4 tiles),
// and so that no split exceeds SPLIT_TILES tiles (2 MiB of Laplace sources):
// the workgroups resident at a time (dispatched in linear-index order) then
// share one split in L2 rather than each streaming the whole source array.
void split_sources(int64_t target_blocks, int64_t ntiles, int num_cus, int64_t* nsplit, int64_t* tiles_per_split) {
  constexpr int64_t SPLIT_TILES = 256;
  const int64_t want = 8LL * std::max(num_cus, 1);
  int64_t ns_ = 1;
  if (target_blocks < want) ns_ = (want + target_blocks - 1) / target_blocks;
  ns_ = std::min<int64_t>(ns_, std::max<int64_t>(1, ntiles / 4));
  ns_ = std::max<int64_t>(ns_, (ntiles + SPLIT_TILES - 1) / SPLIT_TILES);
  ns_ = std::min<int64_t>(ns_, 65535);
  int64_t tps = (ntiles + ns_ - 1) / ns_;
  *tiles_per_split = tps;
  *nsplit = (ntiles + tps - 1) / tps;
}

}  // namespace

// Bytes of the per-split partial sums an FP64 Eval of this size and shape
// needs, bounded by part_budget (split_group).
int f64_part_bytes(int ker, int64_t nt, int64_t ns, const DenseShape& shape, size_t* bytes, size_t part_budget) {
  *bytes = 0;
  if (nt <= 0 || ns <= 0) return SCTL_AMD_SUCCESS;
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  int K0, K1, ND;
  kernel_dims(ker, &K0, &K1, &ND);
  const int tpl = TPLS[std::min(tpl_index(shape.tpl), max_tpl_index(ker))];
  const int64_t tb = (nt + (int64_t)WG * tpl - 1) / ((int64_t)WG * tpl);
  int64_t nsplit, tps;
  split_sources(tb, (ns + TILE - 1) / TILE, dev.num_cus, &nsplit, &tps);
  const size_t split_bytes = sizeof(double) * (size_t)nt * K1;
  const int64_t group = split_group(nsplit, split_bytes, part_budget);
  *bytes = group > 1 ? split_bytes * group : 0;
  return SCTL_AMD_SUCCESS;
}

int f64_apply(int ker, int newton, const double* r_trg, int64_t nt, const double* r_src, const double* n_src,
              const double* v_src, int64_t ns, double* v_trg, const DenseShape& shape, double ctx, double* part,
              hipStream_t stream, size_t part_budget) {
  if (nt <= 0 || ns <= 0) return SCTL_AMD_SUCCESS;
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  int K0, K1, ND;
  kernel_dims(ker, &K0, &K1, &ND);
  const int ti = std::min(tpl_index(shape.tpl), max_tpl_index(ker)), tpl = TPLS[ti];
  const int64_t tb = (nt + (int64_t)WG * tpl - 1) / ((int64_t)WG * tpl);
  SCTL_CHECK_ARG(tb <= 0x7fffffffLL, "too many targets for one launch");
  const int64_t ntiles = (ns + TILE - 1) / TILE;
  int64_t nsplit, tps;
  split_sources(tb, ntiles, dev.num_cus, &nsplit, &tps);
  const int64_t group = split_group(nsplit, sizeof(double) * (size_t)nt * K1, part_budget);
  SCTL_CHECK_ARG(group <= 1 || part, "no buffer for the split partial sums");
  const double scale = kernel_scale(ker) * fold_factor(ker, newton);
  const EvalF64Fn fn = f64_fn(ker, newton, ti);
  size_t dyn = 0;
  SCTL_TRY(occupancy_lds((const void*)fn, shape.occ, sizeof(double) * TILE * rec_words_f64(ker), &dyn));
  if (!stream_capturing(stream)) profile_count(ker, nt * ns);
  const int64_t chunk = tps * TILE;
  // Groups of `group` splits: sources [y0*chunk, (y0+gy)*chunk) per launch.
  for (int64_t y0 = 0; y0 < nsplit; y0 += group) {
    const int64_t gy = std::min(group, nsplit - y0);
    const int64_t s0 = y0 * chunk, ns_g = std::min(ns - s0, gy * chunk);
    const bool direct = gy == 1;
    {
      ProfileScope prof(stream);
      hipLaunchKernelGGL(fn, split_grid(tb, gy, WG), dim3(WG), dyn, stream, r_trg, nt, r_src + s0 * 3,
                         ND ? n_src + s0 * 3 : n_src, v_src + s0 * K0, ns_g, chunk, direct ? v_trg : part,
                         direct ? 1 : 0, scale, ctx);
    }
    SCTL_HIP_TRY(hipGetLastError());
    if (!direct) {
      const int64_t n = nt * K1;
      hipLaunchKernelGGL(k_reduce_splits, dim3((unsigned)((4 * n + WG - 1) / WG)), dim3(WG), 0, stream, part, n,
                         (int)gy, v_trg, scale);
      SCTL_HIP_TRY(hipGetLastError());
    }
  }
  return SCTL_AMD_SUCCESS;
}

int launch_eval_f64(int ker, int newton, const double* r_trg, int64_t nt, const double* r_src, const double* n_src,
                    const double* v_src, int64_t ns, double* v_trg, const DenseShape& shape, double ctx,
                    hipStream_t stream) {
  if (nt <= 0 || ns <= 0) return SCTL_AMD_SUCCESS;
  size_t bytes = 0;
  SCTL_TRY(f64_part_bytes(ker, nt, ns, shape, &bytes));
  Scratch part;
  SCTL_TRY(part.alloc(bytes, stream));
  return f64_apply(ker, newton, r_trg, nt, r_src, n_src, v_src, ns, v_trg, shape, ctx, part.as<double>(), stream);
}

int launch_kernel_matrix(int ker, int newton, const double* r_trg, int64_t nt, const double* r_src,
                         const double* n_src, int64_t ns, double* M, double ctx, hipStream_t stream) {
  if (nt <= 0 || ns <= 0) return SCTL_AMD_SUCCESS;
  const int64_t tb = (nt + WG - 1) / WG;
  SCTL_CHECK_ARG(tb <= 0x7fffffffLL, "too many targets for one launch");
  const int64_t gy = std::min<int64_t>(ns, 8192);
  const double scale = kernel_scale(ker);
  profile_count(ker, nt * ns);
  ProfileScope prof(stream);
  hipLaunchKernelGGL(kmat_fn(ker, newton), dim3((unsigned)tb, (unsigned)gy), dim3(WG), 0, stream, r_trg, nt, r_src,
                     n_src, ns, M, scale, ctx);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

}  // namespace sctl_amd
