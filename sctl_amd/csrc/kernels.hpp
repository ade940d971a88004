// Device-side micro-kernels for the direct summation on gfx950.
//
// Each micro-kernel is the per-interaction body of one of SCTL's real-valued
// uKernels (kernel_functions.hpp:32-79 Laplace3D_FxU/DxU, :81-123 Laplace3D_FxdU,
// :125-178 Stokes3D_FxU, :180-218 Stokes3D_DxU, :220-263 Stokes3D_FxT,
// :264-324 Stokes3D_FSxU, :325-365 Stokes3D_FxUP, :366-390 Laplace3D_Fxd2U,
// :392-414 Laplace3D_DxdU, :416-450 BiotSavart3D_FxU, :451-489
// BiotSavart3D_FxdU, :491-527 Helmholtz3D_FxU, :529-571 Helmholtz3D_DxU,
// :573-620 Helmholtz3D_FxdU, :627-659 HelmholtzDiff3D_FxdU).  The Helmholtz
// kernels take the wavenumber mu (the value the reference reads through
// ctx_ptr) as an argument; their complex values are (re, im) pairs.  Eval uses a fused apply form (K * f accumulated without
// forming K; the reference's own for the kernels that have one, e.g.
// kernel_functions.hpp:165-176,204-217, and an algebraically equal one written
// here for Laplace DxdU / Fxd2U and BiotSavart FxdU, whose reference Eval
// forms the matrix); KernelMatrix uses the matrix form.  One lane owns one
// target; the source record is the same for every lane of the wave
// (broadcast from LDS).
//
// Reciprocal square root.  The hardware estimate (v_rsq_f64 / v_rsq_f32) is
// refined by one deferred-scaling Newton step, the scheme of
// intrin-wrapper.hpp:540-559 (z <- z * (x*z*z - 3), the constant -1/2 applied
// once), for about 14 digits, or by one third-order step for full FP64
// precision (rsq_deferred).  The Newton constant is folded into the
// per-target scale wherever the kernel value is a single power of 1/r
// (Laplace FxU: r^-1, Laplace DxU and FxdU, BiotSavart FxU: r^-3, Stokes DxU
// and FxT: r^-5); the other kernels scale 1/r once per pair.
//
// Masking.  The reference masks r2 > 0 so that coincident points contribute
// exactly zero (kernel_functions.hpp:45, vec-test.hpp:449-463).  The fast
// loops here do not test r2 per interaction: r2 == 0 makes the estimate +inf
// and the running sum non-finite, which the caller detects once per source
// tile and then recomputes that tile with the masked body (MASK = true).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <type_traits>

namespace sctl_amd {

enum KernelId {
  LAP_FXU = 0, LAP_DXU = 1, STK_FXU = 2, STK_DXU = 3,
  LAP_FXDU = 4, LAP_DXDU = 5, LAP_FXD2U = 6, STK_FXT = 7, STK_FSXU = 8, STK_FXUP = 9, BS_FXU = 10, BS_FXDU = 11,
  HLM_FXU = 12, HLM_DXU = 13, HLM_FXDU = 14, HLMD_FXDU = 15,
  NUM_KERNELS = 16
};

// K0 densities per source, K1 outputs per target, ND normal words per source
// (generic-kernel.txx:48-62), POW: the single power of 1/r of the kernel
// value (0 if it mixes powers).
template <int KER> struct Ker;
template <> struct Ker<LAP_FXU> { static constexpr int K0 = 1, K1 = 1, ND = 0, POW = 1; };
template <> struct Ker<LAP_DXU> { static constexpr int K0 = 1, K1 = 1, ND = 3, POW = 3; };
template <> struct Ker<STK_FXU> { static constexpr int K0 = 3, K1 = 3, ND = 0, POW = 0; };
template <> struct Ker<STK_DXU> { static constexpr int K0 = 3, K1 = 3, ND = 3, POW = 5; };
template <> struct Ker<LAP_FXDU> { static constexpr int K0 = 1, K1 = 3, ND = 0, POW = 3; };
template <> struct Ker<LAP_DXDU> { static constexpr int K0 = 1, K1 = 3, ND = 3, POW = 0; };
template <> struct Ker<LAP_FXD2U> { static constexpr int K0 = 1, K1 = 9, ND = 0, POW = 0; };
template <> struct Ker<STK_FXT> { static constexpr int K0 = 3, K1 = 9, ND = 0, POW = 5; };
template <> struct Ker<STK_FSXU> { static constexpr int K0 = 4, K1 = 3, ND = 0, POW = 0; };
template <> struct Ker<STK_FXUP> { static constexpr int K0 = 3, K1 = 4, ND = 0, POW = 0; };
template <> struct Ker<BS_FXU> { static constexpr int K0 = 3, K1 = 3, ND = 0, POW = 3; };
template <> struct Ker<BS_FXDU> { static constexpr int K0 = 3, K1 = 9, ND = 0, POW = 0; };
template <> struct Ker<HLM_FXU> { static constexpr int K0 = 2, K1 = 2, ND = 0, POW = 0; };
template <> struct Ker<HLM_DXU> { static constexpr int K0 = 2, K1 = 2, ND = 3, POW = 0; };
template <> struct Ker<HLM_FXDU> { static constexpr int K0 = 2, K1 = 6, ND = 0, POW = 0; };
template <> struct Ker<HLMD_FXDU> { static constexpr int K0 = 2, K1 = 6, ND = 0, POW = 0; };
constexpr int KMAX0 = 4, KMAX1 = 9;
// Kernels that need the wavenumber (the reference's ctx_ptr).
constexpr bool needs_ctx(int ker) { return ker >= HLM_FXU && ker <= HLMD_FXDU; }

// Calls f(std::integral_constant<int, KER>()) for a kernel id (host side
// dispatch to the kernel templates; ids are checked by the C-ABI first).
template <class F> inline auto with_kernel(int ker, F&& f) {
  switch (ker) {
    case LAP_FXU: return f(std::integral_constant<int, LAP_FXU>());
    case LAP_DXU: return f(std::integral_constant<int, LAP_DXU>());
    case STK_FXU: return f(std::integral_constant<int, STK_FXU>());
    case STK_DXU: return f(std::integral_constant<int, STK_DXU>());
    case LAP_FXDU: return f(std::integral_constant<int, LAP_FXDU>());
    case LAP_DXDU: return f(std::integral_constant<int, LAP_DXDU>());
    case LAP_FXD2U: return f(std::integral_constant<int, LAP_FXD2U>());
    case STK_FXT: return f(std::integral_constant<int, STK_FXT>());
    case STK_FSXU: return f(std::integral_constant<int, STK_FSXU>());
    case STK_FXUP: return f(std::integral_constant<int, STK_FXUP>());
    case BS_FXU: return f(std::integral_constant<int, BS_FXU>());
    case BS_FXDU: return f(std::integral_constant<int, BS_FXDU>());
    case HLM_FXU: return f(std::integral_constant<int, HLM_FXU>());
    case HLM_DXU: return f(std::integral_constant<int, HLM_DXU>());
    case HLM_FXDU: return f(std::integral_constant<int, HLM_FXDU>());
    default: return f(std::integral_constant<int, HLMD_FXDU>());
  }
}

// Largest targets-per-lane count instantiated for a kernel: TPL * K1
// accumulators (plus their saved copies) stay in registers.
template <int KER> constexpr int max_tpl() { return Ker<KER>::K1 <= 3 ? 4 : (Ker<KER>::K1 <= 4 ? 2 : 1); }

// Words per packed source record: position, normal (if any), density.
template <int KER> constexpr int src_words() { return 3 + Ker<KER>::ND + Ker<KER>::K0; }
// Record widths padded to 16 bytes (LDS reads of 128 bits).
template <int KER> constexpr int rec_f64() { return (src_words<KER>() + 1) & ~1; }
template <int KER> constexpr int rec_f32() { return (src_words<KER>() + 3) & ~3; }

__device__ __forceinline__ double hw_rsq(double x) { return __builtin_amdgcn_rsq(x); }
__device__ __forceinline__ float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Launch of tb target blocks x nsplit source splits: a one-dimensional grid of
// tb * nsplit workgroups (target block fastest) where that fits the 32-bit
// grid, else a tb x nsplit grid.  On gfx950 a 79 x 79 grid of 256-thread
// workgroups with ~7 us of arithmetic each took 95 us, the same workgroups in
// a flat grid 55 us; the pairwise kernels, whose workgroups run longer, gain
// 1-3 %.  split_coord returns the workgroup's (target block, split) for
// either form.
inline dim3 split_grid(int64_t tb, int64_t nsplit, int wg) {
  return tb * nsplit * wg <= 0xffffffffLL ? dim3((unsigned)(tb * nsplit)) : dim3((unsigned)tb, (unsigned)nsplit);
}
struct SplitCoord {
  int64_t block, split;
};
__device__ __forceinline__ SplitCoord split_coord(int64_t tb) {
  if (gridDim.y > 1) return {(int64_t)blockIdx.x, (int64_t)blockIdx.y};
  const uint32_t b = (uint32_t)tb;
  return {(int64_t)(blockIdx.x % b), (int64_t)(blockIdx.x / b)};
}

// sum_{y < nsplit} part[y * n + i] by the four lanes of a lane quad (lanes
// 4m .. 4m+3 of a wave, all with the same i): lane j adds the splits
// [j*q, min((j+1)*q, nsplit)) in split order, q = ceil(nsplit / 4), issuing
// the reads of sixteen splits before their additions, and the quad returns
// (s0 + s1) + (s2 + s3) on every lane.  A reduction of many splits over few
// targets thus keeps 64 reads per target in flight.  `ok` lanes only read;
// the others take part in the exchange with zeros.
__device__ __forceinline__ double sum_splits_quad(const double* __restrict__ part, int64_t n, int nsplit, int64_t i,
                                                  bool ok) {
  const int j = threadIdx.x & 3, q = (nsplit + 3) >> 2;
  const int y0 = j * q, y1 = y0 + q < nsplit ? y0 + q : nsplit;
  double s = 0.0;
  if (ok) {
    for (int y = y0; y < y1; y += 16) {
      double v[16];
#pragma unroll
      for (int k = 0; k < 16; k++) v[k] = y + k < y1 ? part[(int64_t)(y + k) * n + i] : 0.0;
#pragma unroll
      for (int k = 0; k < 16; k++)
        if (y + k < y1) s += v[k];
    }
  }
  s += __shfl_xor(s, 1, 64);
  s += __shfl_xor(s, 2, 64);
  return s;
}
// sin and cos of one argument, for the Helmholtz phase mu*r.  Cody and
// Waite's reduction x = k*pi/2 + r (W. J. Cody, W. Waite, Software Manual for
// the Elementary Functions, 1980), with pi/2 split into three FP64 (two FP32)
// parts, each step one FMA, then the Taylor series of sin and cos on
// |r| <= pi/4 (truncated below 2^-55 resp. 2^-28 relative) and the quadrant
// k mod 4.  Arguments beyond 2^20 (FP64) / 2^16 (FP32), where the split
// would lose accuracy, take the device math library's sincos instead.
template <class T> struct SinCosConst;
template <> struct SinCosConst<double> {
  static constexpr double TWO_OVER_PI = 0.6366197723675814;
  static constexpr double P1 = 1.5707963267948966, P2 = 6.123233995736766e-17, P3 = -1.4973849048591698e-33;
  static constexpr double LIMIT = 1048576.0;
};
template <> struct SinCosConst<float> {
  static constexpr float TWO_OVER_PI = 0.63661975f;
  static constexpr float P1 = 1.57079637f, P2 = -4.37113883e-08f;
  static constexpr float LIMIT = 65536.0f;
};

__host__ __device__ __forceinline__ double round_even(double x) { return __builtin_rint(x); }
__host__ __device__ __forceinline__ float round_even(float x) { return __builtin_rintf(x); }

template <class T> __host__ __device__ __forceinline__ void sincos_reduced(T x, T* s, T* c) {
  using C = SinCosConst<T>;
  const T k = round_even(x * C::TWO_OVER_PI);
  T r = __builtin_fma(-k, C::P1, x);
  r = __builtin_fma(-k, C::P2, r);
  if constexpr (std::is_same<T, double>::value) r = __builtin_fma(-k, C::P3, r);
  const T r2 = r * r;
  T ps, pc;  // sin r = r + r^3 ps(r^2), cos r = 1 + r^2 pc(r^2)
  if constexpr (std::is_same<T, double>::value) {
    ps = __builtin_fma(r2, 1.0 / 355687428096000.0, -1.0 / 1307674368000.0);
    ps = __builtin_fma(r2, ps, 1.0 / 6227020800.0);
    ps = __builtin_fma(r2, ps, -1.0 / 39916800.0);
    ps = __builtin_fma(r2, ps, 1.0 / 362880.0);
    ps = __builtin_fma(r2, ps, -1.0 / 5040.0);
    ps = __builtin_fma(r2, ps, 1.0 / 120.0);
    ps = __builtin_fma(r2, ps, -1.0 / 6.0);
    pc = __builtin_fma(r2, 1.0 / 20922789888000.0, -1.0 / 87178291200.0);
    pc = __builtin_fma(r2, pc, 1.0 / 479001600.0);
    pc = __builtin_fma(r2, pc, -1.0 / 3628800.0);
    pc = __builtin_fma(r2, pc, 1.0 / 40320.0);
    pc = __builtin_fma(r2, pc, -1.0 / 720.0);
    pc = __builtin_fma(r2, pc, 1.0 / 24.0);
    pc = __builtin_fma(r2, pc, -0.5);
  } else {
    ps = __builtin_fmaf(r2, 1.0f / 362880.0f, -1.0f / 5040.0f);
    ps = __builtin_fmaf(r2, ps, 1.0f / 120.0f);
    ps = __builtin_fmaf(r2, ps, -1.0f / 6.0f);
    pc = __builtin_fmaf(r2, -1.0f / 3628800.0f, 1.0f / 40320.0f);
    pc = __builtin_fmaf(r2, pc, -1.0f / 720.0f);
    pc = __builtin_fmaf(r2, pc, 1.0f / 24.0f);
    pc = __builtin_fmaf(r2, pc, -0.5f);
  }
  const T sr = __builtin_fma(r * r2, ps, r), cr = __builtin_fma(r2, pc, (T)1);
  const int q = (int)k & 3;  // quadrant (k mod 4, also for negative k)
  const T sq = (q & 1) ? cr : sr, cq = (q & 1) ? sr : cr;
  *s = (q & 2) ? -sq : sq;
  *c = ((q + 1) & 2) ? -cq : cq;
}

__device__ __forceinline__ void dev_sincos(double x, double* s, double* c) {
  if (__builtin_fabs(x) > SinCosConst<double>::LIMIT) sincos(x, s, c);
  else sincos_reduced(x, s, c);
}
__device__ __forceinline__ void dev_sincos(float x, float* s, float* c) {
  if (__builtin_fabsf(x) > SinCosConst<float>::LIMIT) sincosf(x, s, c);
  else sincos_reduced(x, s, c);
}

// x cos x - sin x without cancellation for |x| < 1/2 (its Taylor series
// -x^3/3 + x^5/30 - x^7/840 + x^9/45360, next term < 2^-28 relative there),
// directly otherwise.  For FP32, where the direct form loses about
// 3 / x^2 ulp as x -> 0.
__device__ __forceinline__ float x_cos_minus_sin(float x, float sn, float cs) {
  const float x2 = x * x;
  const float series = x * x2 * fmad(x2, fmad(x2, fmad(x2, 1.0f / 45360, -1.0f / 840), 1.0f / 30), -1.0f / 3);
  return fabsf(x) < 0.5f ? series : fmad(x, cs, -sn);
}

// Helmholtz: (G0, G1) = the complex factor of the kernel for z = 1/r
// (already scaled), as the reference forms it (kernel_functions.hpp:505-510,
// 544-549, 587-592, 641-654):
//   FxU         e^{i mu r} / r
//   DxU, FxdU   (-mu sin - cos / r, mu cos - sin / r) / r^2   (DxU: times r.n)
//   Diff FxdU   the same with cos - 1 = -2 sin^2(mu r / 2) for cos
// In FP32 (the mixed path) the imaginary part mu cos - sin / r, which
// cancels to -mu^3 r^2 / 3 as mu r -> 0, is formed as z (x cos x - sin x)
// with x = mu r (x_cos_minus_sin); FP64 keeps the reference's form.
template <int KER, class T>
__device__ __forceinline__ void helmholtz_factor(T r2, T z, T mu, T* G0, T* G1) {
  constexpr bool F32 = std::is_same<T, float>::value;
  if constexpr (KER == HLMD_FXDU) {
    const T xh = r2 * z * mu * (T)0.5;
    T sh, ch;
    dev_sincos(xh, &sh, &ch);
    const T sn = (T)2 * sh * ch, cm1 = (T)-2 * sh * sh;
    const T z2 = z * z;
    *G0 = fmad(-mu, sn, -cm1 * z) * z2;
    if constexpr (F32) *G1 = x_cos_minus_sin((T)2 * xh, sn, (T)1 + cm1) * (z2 * z);
    else *G1 = fmad(mu, (T)1 + cm1, -sn * z) * z2;
  } else {
    const T x = r2 * z * mu;
    T sn, cs;
    dev_sincos(x, &sn, &cs);
    if constexpr (KER == HLM_FXU) {
      *G0 = cs * z;
      *G1 = sn * z;
    } else {
      const T z2 = z * z;
      *G0 = fmad(-mu, sn, -cs * z) * z2;
      if constexpr (F32) *G1 = x_cos_minus_sin(x, sn, cs) * (z2 * z);
      else *G1 = fmad(mu, cs, -sn * z) * z2;
    }
  }
}

// Deferred-scaling Newton constants (intrin-wrapper.hpp:540-559).  NEWTON = 2
// (full FP64 precision) is not two deferred steps here but one third-order
// step, which needs no final constant (newton_c2(2) = 1; see rsq_deferred).
__host__ __device__ constexpr double newton_c1(int j) { return j == 0 ? -3.0 : (j == 1 ? -12.0 : -768.0); }
__host__ __device__ constexpr double newton_c2(int n) { return n == 1 ? -0.5 : 1.0; }
__host__ __device__ constexpr double ipow(double x, int p) { return p == 0 ? 1.0 : x * ipow(x, p - 1); }

// Unscaled rsqrt: equals newton_c2(NEWTON)^-1 / sqrt(x) up to rounding.
//   NEWTON 0: the hardware estimate (v_rsq_f64 / v_rsq_f32, ~2^-23 relative);
//   NEWTON 1: one deferred Newton step z (x z^2 - 3) (= -2 / sqrt(x), ~2^-47);
//   NEWTON 2: the third-order step z (1 + e/2 + 3 e^2 / 8), e = 1 - x z^2,
//             the first terms of (1 - e)^(-1/2) (Householder's method of
//             order 3): error ~ (5/16) e^3 ~ 2^-70, i.e. rounding only, in 5
//             operations instead of the 6 of two deferred Newton steps.
template <int NEWTON, class T> __device__ __forceinline__ T rsq_deferred(T x) {
  T z = hw_rsq(x);
  if constexpr (NEWTON == 1) z = z * fmad(x, z * z, (T)newton_c1(0));
  if constexpr (NEWTON >= 2) {
    const T e = fmad(-(x * z), z, (T)1);
    z = fmad(z * e, fmad(e, (T)0.375, (T)0.5), z);
  }
  return z;
}

// Factor that turns the accumulated (unscaled) sum into the kernel value:
// newton_c2^POW for single-power kernels, 1 for the others (scaled in-loop).
template <int KER, int NEWTON> __host__ __device__ constexpr double fold_factor() {
  return Ker<KER>::POW ? ipow(newton_c2(NEWTON), Ker<KER>::POW) : 1.0;
}
inline double fold_factor(int ker, int newton) {
  return with_kernel(ker, [&](auto kc) {
    constexpr int K = decltype(kc)::value;
    return newton == 0 ? fold_factor<K, 0>() : (newton == 1 ? fold_factor<K, 1>() : fold_factor<K, 2>());
  });
}

// acc[K1] += K(r, n) * f  (fused apply, unscaled).  r = x_t - x_s.
// n: source normal (ignored when ND == 0); f: K0 densities; mu: the
// wavenumber (Helmholtz kernels only).
template <int KER, int NEWTON, bool MASK, class T>
__device__ __forceinline__ void interact(T* acc, T rx, T ry, T rz, const T* n, const T* f, T mu = (T)0) {
  const T r2 = fmad(rz, rz, fmad(ry, ry, rx * rx));
  T z = rsq_deferred<NEWTON>(r2);
  if constexpr (MASK) z = (r2 > (T)0) ? z : (T)0;
  if constexpr (KER == LAP_FXU) {
    acc[0] = fmad(z, f[0], acc[0]);
  } else if constexpr (KER == LAP_DXU) {
    const T rdotn = fmad(rz, n[2], fmad(ry, n[1], rx * n[0]));
    const T z3 = z * z * z;
    acc[0] = fmad(rdotn * z3, f[0], acc[0]);
  } else if constexpr (KER == STK_FXU) {  // z (f + r (f.r) z^2): 11 operations after z
    if constexpr (NEWTON > 0) z = z * (T)newton_c2(NEWTON);
    const T w = fmad(f[2], rz, fmad(f[1], ry, f[0] * rx)) * (z * z);
    acc[0] = fmad(fmad(rx, w, f[0]), z, acc[0]);
    acc[1] = fmad(fmad(ry, w, f[1]), z, acc[1]);
    acc[2] = fmad(fmad(rz, w, f[2]), z, acc[2]);
  } else if constexpr (KER == STK_DXU) {
    const T z2 = z * z;
    const T z5 = z2 * z2 * z;
    const T rn5 = fmad(rz, n[2], fmad(ry, n[1], rx * n[0])) * z5;
    const T q = fmad(f[2], rz, fmad(f[1], ry, f[0] * rx)) * rn5;
    acc[0] = fmad(rx, q, acc[0]);
    acc[1] = fmad(ry, q, acc[1]);
    acc[2] = fmad(rz, q, acc[2]);
  } else {
    static_assert(KER >= LAP_FXDU && KER < NUM_KERNELS, "unknown kernel");
    (void)mu;
    const T r[3] = {rx, ry, rz};
    if constexpr (Ker<KER>::POW == 0 && NEWTON > 0) z = z * (T)newton_c2(NEWTON);
    if constexpr (KER == LAP_FXDU) {  // f r / r^3
      const T g = f[0] * (z * z * z);
#pragma unroll
      for (int i = 0; i < 3; i++) acc[i] = fmad(r[i], g, acc[i]);
    } else if constexpr (KER == LAP_DXDU) {  // f (n / r^3 - 3 (r.n) r / r^5)
      const T z2 = z * z;
      const T g = f[0] * (z2 * z);
      const T t = (T)3 * fmad(rz, n[2], fmad(ry, n[1], rx * n[0])) * z2;
#pragma unroll
      for (int i = 0; i < 3; i++) acc[i] = fmad(g, fmad(-t, r[i], n[i]), acc[i]);
    } else if constexpr (KER == LAP_FXD2U) {  // f (3 r r^T / r^5 - I / r^3)
      const T z2 = z * z;
      const T g = f[0] * (z2 * z);
      const T t = (T)3 * g * z2;
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const T a = t * r[i];
#pragma unroll
        for (int j = 0; j < 3; j++) acc[i * 3 + j] = fmad(a, r[j], acc[i * 3 + j]);
        acc[i * 3 + i] -= g;
      }
    } else if constexpr (KER == STK_FXT) {  // r (x) r (r.f) / r^5
      const T z2 = z * z;
      const T q = fmad(f[2], rz, fmad(f[1], ry, f[0] * rx)) * (z2 * z2 * z);
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const T a = r[i] * q;
#pragma unroll
        for (int j = 0; j < 3; j++) acc[i * 3 + j] = fmad(a, r[j], acc[i * 3 + j]);
      }
    } else if constexpr (KER == STK_FSXU) {  // f/r + r (r.f + f3) / r^3
      const T t = (fmad(f[2], rz, fmad(f[1], ry, f[0] * rx)) + f[3]) * (z * z * z);
#pragma unroll
      for (int i = 0; i < 3; i++) acc[i] = fmad(f[i], z, fmad(r[i], t, acc[i]));
    } else if constexpr (KER == STK_FXUP) {  // [f/r + r (r.f) / r^3, (r.f) / r^3]
      const T t = fmad(f[2], rz, fmad(f[1], ry, f[0] * rx)) * (z * z * z);
#pragma unroll
      for (int i = 0; i < 3; i++) acc[i] = fmad(f[i], z, fmad(r[i], t, acc[i]));
      acc[3] += t;
    } else if constexpr (KER == BS_FXU) {  // (f x r) / r^3
      const T z3 = z * z * z;
      acc[0] = fmad(fmad(f[1], rz, -ry * f[2]), z3, acc[0]);
      acc[1] = fmad(fmad(f[2], rx, -rz * f[0]), z3, acc[1]);
      acc[2] = fmad(fmad(f[0], ry, -rx * f[1]), z3, acc[2]);
    } else if constexpr (KER == BS_FXDU) {  // output i*3+j: 3 (f x r)_i r_j / r^5 - (f x e_j)_i / r^3
      const T z2 = z * z;
      const T z3 = z2 * z;
      const T t = (T)3 * z3 * z2;
      const T c[3] = {fmad(f[1], rz, -ry * f[2]), fmad(f[2], rx, -rz * f[0]), fmad(f[0], ry, -rx * f[1])};
#pragma unroll
      for (int i = 0; i < 3; i++) {
        const T a = t * c[i];
#pragma unroll
        for (int j = 0; j < 3; j++) acc[i * 3 + j] = fmad(a, r[j], acc[i * 3 + j]);
      }
      acc[1] = fmad(z3, f[2], acc[1]);   // -(f x e_1)_0 =  f2
      acc[2] = fmad(-z3, f[1], acc[2]);  // -(f x e_2)_0 = -f1
      acc[3] = fmad(-z3, f[2], acc[3]);  // -(f x e_0)_1 = -f2
      acc[5] = fmad(z3, f[0], acc[5]);   // -(f x e_2)_1 =  f0
      acc[6] = fmad(z3, f[1], acc[6]);   // -(f x e_0)_2 =  f1
      acc[7] = fmad(-z3, f[0], acc[7]);  // -(f x e_1)_2 = -f0
    } else {  // Helmholtz: complex factor times the complex density (f0 + i f1)
      T G0, G1;
      helmholtz_factor<KER>(r2, z, mu, &G0, &G1);
      if constexpr (KER == HLM_DXU) {
        const T rn = fmad(rz, n[2], fmad(ry, n[1], rx * n[0]));
        G0 = G0 * rn;
        G1 = G1 * rn;
      }
      const T fG0 = fmad(f[0], G0, -f[1] * G1), fG1 = fmad(f[0], G1, f[1] * G0);
      if constexpr (KER == HLM_FXU || KER == HLM_DXU) {
        acc[0] += fG0;
        acc[1] += fG1;
      } else {
#pragma unroll
        for (int i = 0; i < 3; i++) {
          acc[i * 2 + 0] = fmad(fG0, r[i], acc[i * 2 + 0]);
          acc[i * 2 + 1] = fmad(fG1, r[i], acc[i * 2 + 1]);
        }
      }
    }
  }
}

// Matrix form (uKerMatrix): u[k0][k1], unscaled, masked; kernel_functions.hpp
// :43-47, :72-78, :92-99, :136-145, :191-201, :231-243, :275-288, :336-348,
// :377-389, :403-413, :427-434, :462-486.  Used by KernelMatrix.
template <int KER, int NEWTON>
__device__ __forceinline__ void kmatrix(double (&u)[Ker<KER>::K0][Ker<KER>::K1], double rx, double ry, double rz,
                                        const double* n, double mu = 0.0) {
  const double r2 = fmad(rz, rz, fmad(ry, ry, rx * rx));
  double z = rsq_deferred<NEWTON>(r2) * newton_c2(NEWTON);
  z = (r2 > 0.0) ? z : 0.0;
  const double r[3] = {rx, ry, rz};
  const double z2 = z * z, z3 = z2 * z, z5 = z3 * z2;
  if constexpr (KER == LAP_FXU) {
    u[0][0] = z;
  } else if constexpr (KER == LAP_DXU) {
    const double rdotn = fmad(rz, n[2], fmad(ry, n[1], rx * n[0]));
    u[0][0] = rdotn * (z * z * z);
  } else if constexpr (KER == STK_FXU) {
    const double z3 = z * z * z;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) u[i][j] = (i == j ? z : 0.0) + r[i] * r[j] * z3;
  } else if constexpr (KER == STK_DXU) {
    const double rn5 = fmad(rz, n[2], fmad(ry, n[1], rx * n[0])) * (z2 * z2 * z);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) u[i][j] = r[i] * r[j] * rn5;
  } else if constexpr (KER == LAP_FXDU) {
#pragma unroll
    for (int i = 0; i < 3; i++) u[0][i] = r[i] * z3;
  } else if constexpr (KER == LAP_DXDU) {
    const double t5 = 3.0 * fmad(rz, n[2], fmad(ry, n[1], rx * n[0])) * z5;
#pragma unroll
    for (int i = 0; i < 3; i++) u[0][i] = fmad(-r[i], t5, n[i] * z3);
  } else if constexpr (KER == LAP_FXD2U) {
    const double t5 = 3.0 * z5;
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) u[0][i * 3 + j] = t5 * r[i] * r[j] - (i == j ? z3 : 0.0);
  } else if constexpr (KER == STK_FXT) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) u[i][j * 3 + k] = r[i] * r[j] * r[k] * z5;
  } else if constexpr (KER == STK_FSXU || KER == STK_FXUP) {
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) u[i][j] = (i == j ? z : 0.0) + r[i] * r[j] * z3;
#pragma unroll
    for (int j = 0; j < 3; j++) {
      if constexpr (KER == STK_FSXU) u[3][j] = r[j] * z3;
      else u[j][3] = r[j] * z3;
    }
  } else if constexpr (KER == BS_FXU) {
    u[0][0] = 0.0;
    u[0][1] = -r[2] * z3;
    u[0][2] = r[1] * z3;
    u[1][0] = r[2] * z3;
    u[1][1] = 0.0;
    u[1][2] = -r[0] * z3;
    u[2][0] = -r[1] * z3;
    u[2][1] = r[0] * z3;
    u[2][2] = 0.0;
  } else if constexpr (KER >= HLM_FXU) {
    double G0, G1;
    helmholtz_factor<KER>(r2, z, mu, &G0, &G1);
    if constexpr (KER == HLM_DXU) {
      const double rn = fmad(rz, n[2], fmad(ry, n[1], rx * n[0]));
      G0 *= rn;
      G1 *= rn;
    }
    if constexpr (KER == HLM_FXU || KER == HLM_DXU) {
      u[0][0] = G0;
      u[0][1] = G1;
      u[1][0] = -G1;
      u[1][1] = G0;
    } else {
#pragma unroll
      for (int i = 0; i < 3; i++) {
        u[0][i * 2 + 0] = G0 * r[i];
        u[0][i * 2 + 1] = G1 * r[i];
        u[1][i * 2 + 0] = -G1 * r[i];
        u[1][i * 2 + 1] = G0 * r[i];
      }
    }
  } else {  // BS_FXDU: source component k, column i*3+j = d/dr_j of (e_k x r)_i / r^3
    const double t5 = 3.0 * z5;
#pragma unroll
    for (int k = 0; k < 3; k++)
#pragma unroll
      for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
          // (e_k x r)_i = eps(k, m, i) r_m with m the third index
          const int m = 3 - k - i;
          const double eps = (k == i) ? 0.0 : (((k + 1) % 3 == m) ? 1.0 : -1.0);
          u[k][i * 3 + j] = (k == i) ? 0.0 : eps * (t5 * r[m] * r[j] - (m == j ? z3 : 0.0));
        }
  }
}

}  // namespace sctl_amd
