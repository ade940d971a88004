// Device-side micro-kernels for the direct summation on gfx950.
//
// Each micro-kernel is the per-interaction body of one of SCTL's uKernels
// (kernel_functions.hpp:32-79 Laplace3D_FxU/DxU, :125-178 Stokes3D_FxU,
// :180-218 Stokes3D_DxU).  Eval uses the fused apply form (K * f accumulated
// without forming K, kernel_functions.hpp:165-176,204-217); KernelMatrix uses
// the matrix form.  One lane owns one target; the source record is the same
// for every lane of the wave (broadcast from LDS).
//
// Reciprocal square root.  The hardware estimate (v_rsq_f64 / v_rsq_f32) is
// refined by N deferred-scaling Newton steps, the scheme of
// intrin-wrapper.hpp:540-559: z <- z * (x*z*z + c1_j) with c1_j =
// -3*2^(3^j-1), and the constant (-1/2)^((3^(N-1)-1)*3/2+1) applied once.
// Here that constant is folded into the per-target scale wherever the kernel
// value is a single power of 1/r (Laplace FxU: r^-1, Laplace DxU: r^-3,
// Stokes DxU: r^-5), so a Newton step costs three VALU operations.
//
// Masking.  The reference masks r2 > 0 so that coincident points contribute
// exactly zero (kernel_functions.hpp:45, vec-test.hpp:449-463).  The fast
// loops here do not test r2 per interaction: r2 == 0 makes the estimate +inf
// and the running sum non-finite, which the caller detects once per source
// tile and then recomputes that tile with the masked body (MASK = true).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace sctl_amd {

enum KernelId { LAP_FXU = 0, LAP_DXU = 1, STK_FXU = 2, STK_DXU = 3, NUM_KERNELS = 4 };

template <int KER> struct Ker;
template <> struct Ker<LAP_FXU> { static constexpr int K0 = 1, K1 = 1, ND = 0, POW = 1; };
template <> struct Ker<LAP_DXU> { static constexpr int K0 = 1, K1 = 1, ND = 3, POW = 3; };
template <> struct Ker<STK_FXU> { static constexpr int K0 = 3, K1 = 3, ND = 0, POW = 0; };
template <> struct Ker<STK_DXU> { static constexpr int K0 = 3, K1 = 3, ND = 3, POW = 5; };

// Words per packed source record: position, normal (if any), density.
template <int KER> constexpr int src_words() { return 3 + Ker<KER>::ND + Ker<KER>::K0; }
// Record widths padded to 16 bytes (LDS reads of 128 bits).
template <int KER> constexpr int rec_f64() { return (src_words<KER>() + 1) & ~1; }
template <int KER> constexpr int rec_f32() { return (src_words<KER>() + 3) & ~3; }

__device__ __forceinline__ double hw_rsq(double x) { return __builtin_amdgcn_rsq(x); }
__device__ __forceinline__ float hw_rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double fmad(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmad(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Deferred-scaling Newton constants (intrin-wrapper.hpp:540-559).
__host__ __device__ constexpr double newton_c1(int j) { return j == 0 ? -3.0 : (j == 1 ? -12.0 : -768.0); }
__host__ __device__ constexpr double newton_c2(int n) { return n == 0 ? 1.0 : (n == 1 ? -0.5 : (n == 2 ? 1.0 / 16 : -1.0 / 8192)); }
__host__ __device__ constexpr double ipow(double x, int p) { return p == 0 ? 1.0 : x * ipow(x, p - 1); }

// Unscaled rsqrt estimate after NEWTON deferred steps: equals
// newton_c2(NEWTON)^-1 / sqrt(x) up to rounding.
template <int NEWTON, class T> __device__ __forceinline__ T rsq_deferred(T x) {
  T z = hw_rsq(x);
  if constexpr (NEWTON >= 1) z = z * fmad(x, z * z, (T)newton_c1(0));
  if constexpr (NEWTON >= 2) z = z * fmad(x, z * z, (T)newton_c1(1));
  return z;
}

// Factor that turns the accumulated (unscaled) sum into the kernel value:
// newton_c2^POW for single-power kernels, 1 for Stokes FxU (scaled in-loop).
template <int KER, int NEWTON> __host__ __device__ constexpr double fold_factor() {
  return Ker<KER>::POW ? ipow(newton_c2(NEWTON), Ker<KER>::POW) : 1.0;
}

// acc[K1] += K(r, n) * f  (fused apply, unscaled).  r = x_t - x_s.
// n: source normal (ignored when ND == 0); f: K0 densities.
template <int KER, int NEWTON, bool MASK, class T>
__device__ __forceinline__ void interact(T* acc, T rx, T ry, T rz, const T* n, const T* f) {
  const T r2 = fmad(rz, rz, fmad(ry, ry, rx * rx));
  T z = rsq_deferred<NEWTON>(r2);
  if constexpr (MASK) z = (r2 > (T)0) ? z : (T)0;
  if constexpr (KER == LAP_FXU) {
    acc[0] = fmad(z, f[0], acc[0]);
  } else if constexpr (KER == LAP_DXU) {
    const T rdotn = fmad(rz, n[2], fmad(ry, n[1], rx * n[0]));
    const T z3 = z * z * z;
    acc[0] = fmad(rdotn * z3, f[0], acc[0]);
  } else if constexpr (KER == STK_FXU) {
    if constexpr (NEWTON > 0) z = z * (T)newton_c2(NEWTON);
    const T z3 = z * z * z;
    const T fr = fmad(f[2], rz, fmad(f[1], ry, f[0] * rx)) * z3;
    acc[0] = fmad(f[0], z, fmad(rx, fr, acc[0]));
    acc[1] = fmad(f[1], z, fmad(ry, fr, acc[1]));
    acc[2] = fmad(f[2], z, fmad(rz, fr, acc[2]));
  } else {  // STK_DXU
    const T z2 = z * z;
    const T z5 = z2 * z2 * z;
    const T rn5 = fmad(rz, n[2], fmad(ry, n[1], rx * n[0])) * z5;
    const T q = fmad(f[2], rz, fmad(f[1], ry, f[0] * rx)) * rn5;
    acc[0] = fmad(rx, q, acc[0]);
    acc[1] = fmad(ry, q, acc[1]);
    acc[2] = fmad(rz, q, acc[2]);
  }
}

// Matrix form (uKerMatrix): u[k0][k1], unscaled, masked; kernel_functions.hpp
// :43-47, :72-78, :136-145, :191-201.  Used by KernelMatrix.
template <int KER, int NEWTON>
__device__ __forceinline__ void kmatrix(double (&u)[3][3], double rx, double ry, double rz, const double* n) {
  const double r2 = fmad(rz, rz, fmad(ry, ry, rx * rx));
  double z = rsq_deferred<NEWTON>(r2) * newton_c2(NEWTON);
  z = (r2 > 0.0) ? z : 0.0;
  const double r[3] = {rx, ry, rz};
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
  } else {
    const double z2 = z * z;
    const double rn5 = fmad(rz, n[2], fmad(ry, n[1], rx * n[0])) * (z2 * z2 * z);
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
      for (int j = 0; j < 3; j++) u[i][j] = r[i] * r[j] * rn5;
  }
}

}  // namespace sctl_amd
