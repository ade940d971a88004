// Wave-wide reductions by DPP lane permutations: four register-to-register
// steps make every 16-lane row uniform, then the four row values are read
// into scalar registers.  A chain of LDS-routed shuffles (ds_bpermute) takes
// about a hundred cycles per step; these take a few.  All 64 lanes of the wave
// must be active where they are called.
#pragma once
#include <hip/hip_runtime.h>

namespace sctl_amd {

// DPP controls (GCN/CDNA): quad_perm [1,0,3,2] and [2,3,0,1] (lane xor 1,
// xor 2), row_half_mirror (lane i <- 7 - i within 8), row_mirror (lane i <-
// 15 - i within 16).
constexpr int DPP_QUAD_XOR1 = 0xB1, DPP_QUAD_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

template <int CTRL> __device__ __forceinline__ float dpp_mov(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), CTRL, 0xf, 0xf, false));
}
template <int CTRL> __device__ __forceinline__ double dpp_mov(double v) {
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, false);
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, false);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ double lane_value(double v, int lane) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                          __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// op over the 64 lanes, returned on every lane.  op must be associative and
// commutative; for min and max the result does not depend on the order.
template <class T, class Op> __device__ __forceinline__ T wave_reduce(T v, Op op) {
  v = op(v, dpp_mov<DPP_QUAD_XOR1>(v));
  v = op(v, dpp_mov<DPP_QUAD_XOR2>(v));
  v = op(v, dpp_mov<DPP_HALF_MIRROR>(v));
  v = op(v, dpp_mov<DPP_MIRROR>(v));
  return op(op(lane_value(v, 0), lane_value(v, 16)), op(lane_value(v, 32), lane_value(v, 48)));
}

struct MinOp {
  __device__ float operator()(float a, float b) const { return fminf(a, b); }
  __device__ double operator()(double a, double b) const { return fmin(a, b); }
};
struct MaxOp {
  __device__ float operator()(float a, float b) const { return fmaxf(a, b); }
  __device__ double operator()(double a, double b) const { return fmax(a, b); }
};
template <class T> __device__ __forceinline__ T wave_min(T v) { return wave_reduce(v, MinOp{}); }
template <class T> __device__ __forceinline__ T wave_max(T v) { return wave_reduce(v, MaxOp{}); }

}  // namespace sctl_amd
