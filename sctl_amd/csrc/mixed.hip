// Mixed-precision dense Eval on gfx950: FP64 tile-centred differencing, FP32
// pairwise arithmetic, Kahan-compensated FP64 accumulation.
//
// Sources and targets are both put in Hilbert-curve order.  The sources are packed,
// TILE at a time, relative to each tile's bounding-box centre c: for every
// source the FP64 offset s - c is split into an FP32 value s' and its FP32
// residual s'' (s - c = s' + s'' to about 2^-48 relative), and the tile keeps
// H^2 = max |s'|^2.  A target (held in FP64) forms t - c once per tile the same
// way, t' + t''.
//
// Per (wave, tile) one of two loops runs:
//   far  -- every target of the wave has |t'| >= 3H, so every pair has
//           r >= 2H.  Laplace kernels use r^2 = |t'|^2 + |s'|^2 - 2 t'.s'
//           (one add and three FMAs; the rounding error is at most a few ulp
//           of r^2 because r^2 >= (|t'| - H)^2 is never small against the
//           terms).  Stokes kernels use dx = t' - s'.
//   near -- some target is within 3H of the tile centre: positions are
//           differenced from both halves, dx = (t' - s') + (t'' - s''), which
//           is accurate to about 2^-23 relative to dx however small r is, and
//           r^2 == 0 is masked as in the reference (kernel_functions.hpp:45).
// Targets in Hilbert order keep a wave's targets together, so only tiles next
// to the wave take the near loop.  The pairs of every 64 sources are summed
// in FP32 and added into an FP64 sum of the tile, so that a tile's largest
// term does not take up the rounding of 255 FP32 additions (this halved the
// headline's error against FP64, 5.4e-7 -> 2.7e-7, at no cost in time).  A
// far tile whose sum is not finite is redone by the near loop.  Each tile's
// sum is added into the FP64 per-target sum with Kahan compensation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "internal.hpp"
#include "kernels.hpp"
#include "wave.hpp"

namespace sctl_amd {

namespace {

constexpr int WG = 256;
constexpr int TILE = 256;
constexpr float PADF = 1e18f;

// Record layout (floats) per source; far-path words first.
//   LAP_FXU  [s'x s'y s'z S | q s''x s''y s''z]                         S = |s'|^2
//   LAP_DXU  [s'x s'y s'z S | nq_x nq_y nq_z -s'.nq | s''x s''y s''z 0]   nq = q n
//   others   [s'x s'y s'z | f_0 .. f_{K0-1} | n (if any) | s''x s''y s''z], padded
//            to whole float4s, e.g.
//   STK_FXU  [s'x s'y s'z f0 | f1 f2 s''x s''y | s''z 0 0 0]
//   STK_DXU  [s'x s'y s'z f0 | f1 f2 n_x n_y | n_z s''x s''y s''z]
// FOFF / NOFF / LO: word offsets of the densities, the normal and s''.
template <int KER> struct MixLayout {
  static constexpr int FOFF = 3, NOFF = 3 + Ker<KER>::K0, LO = NOFF + Ker<KER>::ND, Q = (LO + 3 + 3) / 4;
};
template <> struct MixLayout<LAP_FXU> { static constexpr int Q = 2, LO = 5; };
template <> struct MixLayout<LAP_DXU> { static constexpr int Q = 3, LO = 8; };

template <int KER>
__device__ __forceinline__ void fill_record(float* r, const double* d, const double* n, const double* f) {
  float hi[3], lo[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    hi[k] = (float)d[k];
    lo[k] = (float)(d[k] - (double)hi[k]);
  }
  const double S = (double)hi[0] * hi[0] + (double)hi[1] * hi[1] + (double)hi[2] * hi[2];
  constexpr int LO = MixLayout<KER>::LO;
  r[0] = hi[0];
  r[1] = hi[1];
  r[2] = hi[2];
  if constexpr (KER == LAP_FXU) {
    r[3] = (float)S;
    r[4] = (float)f[0];
  } else if constexpr (KER == LAP_DXU) {
    const float nq[3] = {(float)(n[0] * f[0]), (float)(n[1] * f[0]), (float)(n[2] * f[0])};
    r[3] = (float)S;
    r[4] = nq[0];
    r[5] = nq[1];
    r[6] = nq[2];
    r[7] = (float)(-((double)hi[0] * nq[0] + (double)hi[1] * nq[1] + (double)hi[2] * nq[2]));
    r[11] = 0.0f;
  } else {
    using L = MixLayout<KER>;
#pragma unroll
    for (int k = 0; k < Ker<KER>::K0; k++) r[L::FOFF + k] = (float)f[k];
#pragma unroll
    for (int k = 0; k < Ker<KER>::ND; k++) r[L::NOFF + k] = (float)n[k];
  }
  r[LO + 0] = lo[0];
  r[LO + 1] = lo[1];
  r[LO + 2] = lo[2];
}

// Far-field interaction.  th: t' ; a: -2 t' ; T: |t'|^2 ; r: the record.
template <int KER, int NEWTON>
__device__ __forceinline__ void mix_far(float* p, const float* th, const float* a, float T, const float* r, float mu) {
  if constexpr (KER == LAP_FXU) {
    const float r2 = fmad(a[2], r[2], fmad(a[1], r[1], fmad(a[0], r[0], T + r[3])));
    p[0] = fmad(rsq_deferred<NEWTON>(r2), r[4], p[0]);
  } else if constexpr (KER == LAP_DXU) {
    const float r2 = fmad(a[2], r[2], fmad(a[1], r[1], fmad(a[0], r[0], T + r[3])));
    const float rnq = fmad(th[2], r[6], fmad(th[1], r[5], fmad(th[0], r[4], r[7])));
    const float z = rsq_deferred<NEWTON>(r2);
    p[0] = fmad(rnq, z * z * z, p[0]);
  } else {
    using L = MixLayout<KER>;
    interact<KER, NEWTON, false>(p, th[0] - r[0], th[1] - r[1], th[2] - r[2], r + L::NOFF, r + L::FOFF, mu);
  }
}

// Mid-range pair of the Laplace kernels (every target of the wave at least
// 2H from the tile centre, so r >= H): one-part FP32 differences, accurate
// to a few ulp of r there, and no mask.
template <int KER, int NEWTON>
__device__ __forceinline__ void mix_mid(float* p, const float* th, const float* r) {
  const float dx = th[0] - r[0], dy = th[1] - r[1], dz = th[2] - r[2];
  if constexpr (KER == LAP_FXU) {
    const float f[1] = {r[4]};
    interact<LAP_FXU, NEWTON, false>(p, dx, dy, dz, f, f);
  } else {
    const float one[1] = {1.0f}, nq[3] = {r[4], r[5], r[6]};
    interact<LAP_DXU, NEWTON, false>(p, dx, dy, dz, nq, one);
  }
}

// Near-field interaction: two-part differencing, r^2 == 0 masked.
template <int KER, int NEWTON>
__device__ __forceinline__ void mix_near(float* p, const float* th, const float* tl, const float* r, float mu) {
  constexpr int LO = MixLayout<KER>::LO;
  const float dx = (th[0] - r[0]) + (tl[0] - r[LO + 0]);
  const float dy = (th[1] - r[1]) + (tl[1] - r[LO + 1]);
  const float dz = (th[2] - r[2]) + (tl[2] - r[LO + 2]);
  if constexpr (KER == LAP_FXU) {
    const float f[1] = {r[4]};
    interact<LAP_FXU, NEWTON, true>(p, dx, dy, dz, f, f);
  } else if constexpr (KER == LAP_DXU) {
    const float one[1] = {1.0f}, nq[3] = {r[4], r[5], r[6]};
    interact<LAP_DXU, NEWTON, true>(p, dx, dy, dz, nq, one);
  } else {
    using L = MixLayout<KER>;
    interact<KER, NEWTON, true>(p, dx, dy, dz, r + L::NOFF, r + L::FOFF, mu);
  }
}

// One workgroup per tile: tile centre and H^2, then the records, sources in
// the order perm[].  meta[tile] = {c_x, c_y, c_z, H^2}; sub[4 tile + q] =
// {b_x, b_y, b_z, h^2} of sub-tile q (sources 64q .. 64q + 63, one wave):
// the centre of its FP32 offsets' bounding box, relative to the tile centre,
// and the squared radius of its offsets about that centre ({0, 0, 0, 0} for
// a sub-tile of padding only).
template <int KER>
__global__ void __launch_bounds__(WG)
k_pack_tiles(const double* __restrict__ xs, const double* __restrict__ xn, const double* __restrict__ fs, int64_t ns,
             const uint32_t* __restrict__ perm, float4* __restrict__ packed, double4* __restrict__ meta,
             float4* __restrict__ sub, int64_t ntiles, const double* __restrict__ xt, int64_t nt,
             const uint32_t* __restrict__ tperm, double* __restrict__ xt_sorted) {
  using K = Ker<KER>;
  constexpr int K0 = K::K0, ND = K::ND, Q = MixLayout<KER>::Q;
  static_assert(TILE == WG, "one source per lane");
  __shared__ double red[WG / 64][6];
  __shared__ float hred[WG / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  if ((int64_t)blockIdx.x >= ntiles) {  // the workgroups after the tiles copy the targets into their order
    const int64_t q = ((int64_t)blockIdx.x - ntiles) * WG + tid;
    if (q < nt) {
      const int64_t t = tperm[q];
#pragma unroll
      for (int k = 0; k < 3; k++) xt_sorted[q * 3 + k] = xt[t * 3 + k];
    }
    return;
  }
  const int64_t s = (int64_t)blockIdx.x * TILE + tid;
  const bool ok = s < ns;
  const int64_t src = ok ? (int64_t)perm[s] : 0;
  // every read of the source issued at once
  double p[3] = {0, 0, 0}, n[3] = {0, 0, 0}, f[KMAX0] = {0, 0, 0, 0};
  if (ok) {
#pragma unroll
    for (int k = 0; k < 3; k++) p[k] = xs[src * 3 + k];
#pragma unroll
    for (int k = 0; k < ND; k++) n[k] = xn[src * 3 + k];
#pragma unroll
    for (int k = 0; k < K0; k++) f[k] = fs[src * K0 + k];
  }
  // tile bounding box: per wave by exchanges, then over the four waves
#pragma unroll
  for (int k = 0; k < 3; k++) {
    const double lo = wave_min(ok ? p[k] : INFINITY), hi = wave_max(ok ? p[k] : -INFINITY);
    if (lane == 0) {
      red[wv][k] = lo;
      red[wv][3 + k] = hi;
    }
  }
  __syncthreads();
  double bb[6];
#pragma unroll
  for (int k = 0; k < 6; k++) {
    bb[k] = red[0][k];
#pragma unroll
    for (int w = 1; w < WG / 64; w++) bb[k] = k < 3 ? fmin(bb[k], red[w][k]) : fmax(bb[k], red[w][k]);
  }
  const double c[3] = {0.5 * (bb[0] + bb[3]), 0.5 * (bb[1] + bb[4]), 0.5 * (bb[2] + bb[5])};
  float rec[4 * Q];
#pragma unroll
  for (int k = 0; k < 4 * Q; k++) rec[k] = 0.0f;
  float h2 = 0.0f;
  if (ok) {
    const double d[3] = {p[0] - c[0], p[1] - c[1], p[2] - c[2]};
    fill_record<KER>(rec, d, n, f);
    h2 = rec[0] * rec[0] + rec[1] * rec[1] + rec[2] * rec[2];
  } else {
    rec[0] = rec[1] = rec[2] = PADF;
    if constexpr (KER == LAP_FXU || KER == LAP_DXU) rec[3] = 3.0f * PADF * PADF;
  }
  h2 = wave_max(h2);
  if (lane == 0) hred[wv] = h2;
  __syncthreads();
  if (tid == 0) {
    float hm = hred[0];
#pragma unroll
    for (int w = 1; w < WG / 64; w++) hm = fmaxf(hm, hred[w]);
    meta[blockIdx.x] = make_double4(c[0], c[1], c[2], (double)hm * 1.0000001);
  }
  {
    float b[3], bh2 = 0.0f;
    const bool any = __builtin_amdgcn_ballot_w64(ok) != 0;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      const float lo = wave_min(ok ? rec[k] : INFINITY), hi = wave_max(ok ? rec[k] : -INFINITY);
      b[k] = any ? 0.5f * (lo + hi) : 0.0f;
    }
    if (ok) {
      const float dx = rec[0] - b[0], dy = rec[1] - b[1], dz = rec[2] - b[2];
      bh2 = dx * dx + dy * dy + dz * dz;
    }
    bh2 = wave_max(bh2);
    if ((tid & 63) == 0) sub[(int64_t)blockIdx.x * (WG / 64) + (tid >> 6)] = make_float4(b[0], b[1], b[2], bh2 * 1.0001f);
  }
  float4* dst = packed + ((int64_t)blockIdx.x * TILE + tid) * Q;
#pragma unroll
  for (int q = 0; q < Q; q++) dst[q] = make_float4(rec[4 * q], rec[4 * q + 1], rec[4 * q + 2], rec[4 * q + 3]);
}

// Waves per SIMD that the register allocation must leave room for: the
// Helmholtz gradient kernels (6 outputs per target) need 140-143 VGPRs at one
// target per lane, which allows 3 waves; holding them to 128 allows 4.
template <int KER, int TPL> constexpr int mixed_min_waves() {
  return (KER == HLM_FXDU || KER == HLMD_FXDU) && TPL == 1 ? 4 : 1;
}

template <int KER, int TPL, int NEWTON>
__global__ void __launch_bounds__(WG, (mixed_min_waves<KER, TPL>()))
k_eval_mixed(const double* __restrict__ xt, int64_t nt, const uint32_t* __restrict__ tperm,
             const float4* __restrict__ packed, const double4* __restrict__ meta, const float4* __restrict__ sub,
             int64_t ntiles, int64_t tiles_per_split, double* __restrict__ out, int direct, double scale, double ctx) {
  using K = Ker<KER>;
  constexpr int K1 = K::K1, Q = MixLayout<KER>::Q, R = 4 * Q;
  __shared__ float4 lds[TILE * Q];
  const float mu = (float)ctx;

  const int tid = threadIdx.x;
  const SplitCoord wg = split_coord((nt + WG * TPL - 1) / (WG * TPL));
  const int64_t tile_begin = wg.split * tiles_per_split;
  const int64_t tile_end = tile_begin + tiles_per_split < ntiles ? tile_begin + tiles_per_split : ntiles;
  const int64_t q0 = wg.block * (WG * TPL) + tid;

  double t64[TPL][3], acc[TPL][K1], comp[TPL][K1];
#pragma unroll
  for (int j = 0; j < TPL; j++) {
    const int64_t q = q0 + (int64_t)j * WG;
    const bool ok = q < nt;
#pragma unroll
    for (int k = 0; k < 3; k++) t64[j][k] = ok ? xt[q * 3 + k] : 0.0;  // xt in the order tperm
#pragma unroll
    for (int k = 0; k < K1; k++) acc[j][k] = comp[j][k] = 0.0;
  }

  for (int64_t tile = tile_begin; tile < tile_end; tile++) {
    __syncthreads();
    const float4* src = packed + tile * (TILE * Q);
#pragma unroll
    for (int q = 0; q < Q; q++) lds[q * WG + tid] = src[q * WG + tid];
    const double4 m = meta[tile];
    __syncthreads();

    // A wave whose targets are all |t'| >= 3H from the tile centre takes the
    // far form over the whole tile (the expanded r^2 of the Laplace kernels,
    // one-part differences for the others; the latter from |t'| >= 2H).
    // Otherwise each 64-source sub-tile q (centre b_q, radius h_q about it)
    // is classified on its own, with X = |t'| + H the scale of the rounding
    // of t' and s' and r >= |t' - b_q| - h_q: far where that bound is >= X/2
    // (relative error of the expanded r^2 <= 4 eps), mid (one-part
    // differences, relative error <= 3 eps) where it is >= X/3, else near
    // (two-part differences, masked).  A non-finite partial redoes the tile
    // in the near form.
    constexpr bool EXPANDED = KER == LAP_FXU || KER == LAP_DXU;
    const float far_r2 = (EXPANDED ? 9.0f : 4.0f) * (float)m.w;
    // p: FP32 sums of at most 64 pairs, added into the FP64 tile sums pt
    // after every 64 sources (sub64), so that a tile's dominant term does
    // not absorb the rounding of 255 FP32 additions.
    float th[TPL][3], T[TPL], p[TPL][K1];
    double pt[TPL][K1];
    bool notfar = false;
#pragma unroll
    for (int j = 0; j < TPL; j++) {
      th[j][0] = (float)(t64[j][0] - m.x);
      th[j][1] = (float)(t64[j][1] - m.y);
      th[j][2] = (float)(t64[j][2] - m.z);
      T[j] = fmad(th[j][2], th[j][2], fmad(th[j][1], th[j][1], th[j][0] * th[j][0]));
      notfar |= T[j] < far_r2;
#pragma unroll
      for (int k = 0; k < K1; k++) {
        p[j][k] = 0.0f;
        pt[j][k] = 0.0;
      }
    }
    auto sub64 = [&]() {
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < K1; k++) {
          pt[j][k] += (double)p[j][k];
          p[j][k] = 0.0f;
        }
    };
    const float* recs = reinterpret_cast<const float*>(lds);
    bool use_near = false;
    // -2 t', formed where a far loop needs it (not held across the sub-tiles)
    auto minus2 = [&](float(&a)[TPL][3]) {
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < 3; k++) a[j][k] = -2.0f * th[j][k];
    };
    // At more than one target per lane a wave's targets span more space and
    // the sub-tile classes would raise the kernel's registers (two targets
    // per lane: 70 -> 93 VGPRs); such shapes classify whole tiles only, with
    // the mid class at |t'| >= 2H.
    constexpr bool SUBTILES = TPL == 1;
    bool tile_mid = false;
    if constexpr (!SUBTILES) {
      const float mid_r2 = 4.0f * (float)m.w;
      bool near = false;
#pragma unroll
      for (int j = 0; j < TPL; j++) near |= T[j] < mid_r2;
      use_near = __builtin_amdgcn_ballot_w64(near) != 0;
      tile_mid = !use_near && EXPANDED && __builtin_amdgcn_ballot_w64(notfar) != 0;
    }
    if (use_near) {
      // the near loop below
    } else if (tile_mid) {
      if constexpr (EXPANDED) {
        for (int i0 = 0; i0 < TILE; i0 += 64) {
#pragma unroll 4
          for (int i = i0; i < i0 + 64; i++) {
            const float* rec = recs + i * R;
#pragma unroll
            for (int j = 0; j < TPL; j++) mix_mid<KER, NEWTON>(p[j], th[j], rec);
          }
          sub64();
        }
      }
    } else if (!SUBTILES || __builtin_amdgcn_ballot_w64(notfar) == 0) {
      float a[TPL][3];
      minus2(a);
      for (int i0 = 0; i0 < TILE; i0 += 64) {
#pragma unroll 4
        for (int i = i0; i < i0 + 64; i++) {
          const float* rec = recs + i * R;
#pragma unroll
          for (int j = 0; j < TPL; j++) mix_far<KER, NEWTON>(p[j], th[j], a[j], T[j], rec, mu);
        }
        sub64();
      }
    } else if constexpr (SUBTILES) {
      const float H = sqrtf((float)m.w);
#pragma unroll 1
      for (int q = 0; q < WG / 64; q++) {
        const float4 b = sub[tile * (WG / 64) + q];
        const float h = sqrtf(b.w);
        bool qnotfar = false, qnear = false;
#pragma unroll
        for (int j = 0; j < TPL; j++) {
          const float X = sqrtf(T[j]) + H;
          const float dx = th[j][0] - b.x, dy = th[j][1] - b.y, dz = th[j][2] - b.z;
          const float d2 = fmad(dz, dz, fmad(dy, dy, dx * dx));
          const float ef = fmad(0.5f, X, h), em = fmad(1.0f / 3.0f, X, h);
          qnotfar |= d2 < ef * ef;
          qnear |= d2 < em * em;
        }
        const float* qrec = recs + q * 64 * R;
        if (__builtin_amdgcn_ballot_w64(qnear) != 0) {
          float tl[TPL][3];
#pragma unroll
          for (int j = 0; j < TPL; j++) {
            tl[j][0] = (float)((t64[j][0] - m.x) - (double)th[j][0]);
            tl[j][1] = (float)((t64[j][1] - m.y) - (double)th[j][1]);
            tl[j][2] = (float)((t64[j][2] - m.z) - (double)th[j][2]);
          }
#pragma unroll 2
          for (int i = 0; i < 64; i++) {
#pragma unroll
            for (int j = 0; j < TPL; j++) mix_near<KER, NEWTON>(p[j], th[j], tl[j], qrec + i * R, mu);
          }
        } else if (EXPANDED && __builtin_amdgcn_ballot_w64(qnotfar) != 0) {
          if constexpr (EXPANDED) {
#pragma unroll 4
            for (int i = 0; i < 64; i++) {
#pragma unroll
              for (int j = 0; j < TPL; j++) mix_mid<KER, NEWTON>(p[j], th[j], qrec + i * R);
            }
          }
        } else {
          float a[TPL][3];
          minus2(a);
#pragma unroll 4
          for (int i = 0; i < 64; i++) {
#pragma unroll
            for (int j = 0; j < TPL; j++) mix_far<KER, NEWTON>(p[j], th[j], a[j], T[j], qrec + i * R, mu);
          }
        }
        sub64();
      }
    }
    if (!use_near) {
      bool bad = false;
#pragma unroll
      for (int j = 0; j < TPL; j++)
#pragma unroll
        for (int k = 0; k < K1; k++) bad |= !__builtin_isfinite(pt[j][k]);
      use_near = __builtin_amdgcn_ballot_w64(bad) != 0;
    }
    if (use_near) {
      float tl[TPL][3];
#pragma unroll
      for (int j = 0; j < TPL; j++) {
        tl[j][0] = (float)((t64[j][0] - m.x) - (double)th[j][0]);
        tl[j][1] = (float)((t64[j][1] - m.y) - (double)th[j][1]);
        tl[j][2] = (float)((t64[j][2] - m.z) - (double)th[j][2]);
#pragma unroll
        for (int k = 0; k < K1; k++) {
          p[j][k] = 0.0f;
          pt[j][k] = 0.0;
        }
      }
      for (int i0 = 0; i0 < TILE; i0 += 64) {
#pragma unroll 2
        for (int i = i0; i < i0 + 64; i++) {
          const float* rec = recs + i * R;
#pragma unroll
          for (int j = 0; j < TPL; j++) mix_near<KER, NEWTON>(p[j], th[j], tl[j], rec, mu);
        }
        sub64();
      }
    }
#pragma unroll
    for (int j = 0; j < TPL; j++)
#pragma unroll
      for (int k = 0; k < K1; k++) {  // Kahan: acc += pt
        const double y = pt[j][k] - comp[j][k];
        const double s = acc[j][k] + y;
        comp[j][k] = (s - acc[j][k]) - y;
        acc[j][k] = s;
      }
  }

#pragma unroll
  for (int j = 0; j < TPL; j++) {
    const int64_t q = q0 + (int64_t)j * WG;
    if (q < nt) {
      // the target's own index is read only here, so it holds no register
      // through the tile loop
      const int64_t t = direct ? (int64_t)tperm[q] : 0;
#pragma unroll
      for (int k = 0; k < K1; k++) {
        const double v = acc[j][k] - comp[j][k];
        if (direct) out[t * K1 + k] += v * scale;
        else out[wg.split * nt * K1 + q * K1 + k] = v;
      }
    }
  }
}

// v_trg[tperm[q]*K1+k] += scale * sum_y part[y][q*K1+k]; one lane quad per
// element (sum_splits_quad).
__global__ void __launch_bounds__(WG)
k_reduce_splits_perm(const double* __restrict__ part, int64_t nt, int K1, int nsplit, const uint32_t* __restrict__ tperm,
                     double* __restrict__ out, double scale) {
  const int64_t i = ((int64_t)blockIdx.x * WG + threadIdx.x) >> 2;
  const int64_t n = nt * K1;
  const double s = sum_splits_quad(part, n, nsplit, i, i < n);
  if (i >= n || (threadIdx.x & 3) != 0) return;
  const int64_t q = i / K1, k = i - q * K1;
  out[(int64_t)tperm[q] * K1 + k] += s * scale;
}

using EvalMixFn = void (*)(const double*, int64_t, const uint32_t*, const float4*, const double4*, const float4*, int64_t,
                           int64_t, double*, int, double, double);
using PackFn = void (*)(const double*, const double*, const double*, int64_t, const uint32_t*, float4*, double4*,
                        float4*, int64_t, const double*, int64_t, const uint32_t*, double*);

constexpr int TPLS[3] = {1, 2, 4};
int tpl_index(int tpl) { return tpl >= 4 ? 2 : (tpl >= 2 ? 1 : 0); }

template <int KER, int NEWTON> EvalMixFn mix_fn(int ti) {
  constexpr int MT = max_tpl<KER>();
  if (ti >= 2 && MT >= 4) return k_eval_mixed<KER, (MT >= 4 ? 4 : 1), NEWTON>;
  if (ti >= 1 && MT >= 2) return k_eval_mixed<KER, (MT >= 2 ? 2 : 1), NEWTON>;
  return k_eval_mixed<KER, 1, NEWTON>;
}
EvalMixFn mix_fn(int ker, int newton, int ti) {
  return with_kernel(ker, [&](auto kc) -> EvalMixFn {
    constexpr int K = decltype(kc)::value;
    return newton ? mix_fn<K, 1>(ti) : mix_fn<K, 0>(ti);
  });
}
PackFn pack_fn(int ker) {
  return with_kernel(ker, [](auto kc) -> PackFn { return k_pack_tiles<decltype(kc)::value>; });
}
int layout_q(int ker) {
  return with_kernel(ker, [](auto kc) { return MixLayout<decltype(kc)::value>::Q; });
}
int max_tpl_index(int ker) {
  return with_kernel(ker, [](auto kc) { return tpl_index(max_tpl<decltype(kc)::value>()); });
}

}  // namespace

// Launch geometry of one mixed Eval: target blocks and the source splits.
// Split the sources over gridDim.y when the target blocks cannot fill the
// device (>= 8 workgroups per CU, each split >= 4 tiles), and so that no
// split exceeds SPLIT_TILES tiles: workgroups are dispatched in order of
// their linear index (blockIdx.x fastest), so the ones resident at a time
// share one split's records (2 MiB for Laplace FxU) in their XCD's 4 MiB L2
// instead of drifting apart over the whole source array.  The splits' partial
// sums stay within part_budget bytes (split_group).
int mixed_geometry(int ker, int64_t nt, int64_t ns, const DenseShape& shape, MixedGeometry* g, size_t part_budget) {
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  SCTL_CHECK_ARG(ns <= 0xffffffffLL && nt <= 0xffffffffLL, "mixed precision supports at most 2^32-1 points");
  g->ti = std::min(tpl_index(shape.tpl), max_tpl_index(ker));
  const int tpl = TPLS[g->ti];
  g->tb = (nt + (int64_t)WG * tpl - 1) / ((int64_t)WG * tpl);
  SCTL_CHECK_ARG(g->tb <= 0x7fffffffLL, "too many targets for one launch");
  g->ntiles = (ns + TILE - 1) / TILE;
  SCTL_CHECK_ARG(g->ntiles <= 0x7fffffffLL, "too many sources for one launch");
  g->Q = layout_q(ker);
  // Enough workgroups for 32 per CU (one tile per split at the least): a
  // small problem (a few hundred tiles) still fills the GPU.
  constexpr int64_t SPLIT_TILES = 256;
  const int64_t want = 32LL * std::max(dev.num_cus, 1);
  int64_t nsplit = g->tb < want ? (want + g->tb - 1) / g->tb : 1;
  nsplit = std::min<int64_t>(std::min<int64_t>(nsplit, g->ntiles), 65535);
  nsplit = std::min<int64_t>(std::max<int64_t>(nsplit, (g->ntiles + SPLIT_TILES - 1) / SPLIT_TILES), 65535);
  g->tps = (g->ntiles + nsplit - 1) / nsplit;
  g->nsplit = (g->ntiles + g->tps - 1) / g->tps;
  int K0, K1, ND;
  kernel_dims(ker, &K0, &K1, &ND);
  g->packed_bytes = sizeof(float4) * g->Q * TILE * g->ntiles;
  // tile centres, then sub-tiles, then the targets in their order
  g->meta_bytes = (sizeof(double4) + sizeof(float4) * (WG / 64)) * g->ntiles + sizeof(double) * 3 * nt;
  const size_t split_bytes = sizeof(double) * (size_t)nt * K1;
  g->group = split_group(g->nsplit, split_bytes, part_budget);
  g->part_bytes = g->group > 1 ? split_bytes * g->group : 0;
  return SCTL_AMD_SUCCESS;
}

size_t mixed_order_bytes(int64_t nt, int64_t ns) { return hilbert_order_bytes(ns, nt); }

int mixed_order(const double* r_trg, int64_t nt, const double* r_src, int64_t ns, uint32_t* tperm, uint32_t* sperm,
                void* work, hipStream_t stream) {
  Scratch own;
  if (!work) {
    SCTL_TRY(own.alloc(mixed_order_bytes(nt, ns), stream));
    work = own.as<void>();
  }
  return hilbert_order(r_src, ns, sperm, r_trg, nt, tperm, work, stream);
}

int mixed_apply(int ker, int newton, const MixedGeometry& g, const DenseShape& shape, const double* r_trg, int64_t nt,
                const uint32_t* tperm, const double* r_src, const double* n_src, const double* v_src, int64_t ns,
                const uint32_t* sperm, float4* packed, double4* meta, double* part, double* v_trg, double ctx,
                hipStream_t stream) {
  if (nt <= 0 || ns <= 0) return SCTL_AMD_SUCCESS;
  SCTL_CHECK_ARG(g.group <= 1 || part, "no buffer for the split partial sums");
  int K0, K1, ND;
  kernel_dims(ker, &K0, &K1, &ND);
  // The sub-tile records follow the tile centres in the same buffer, then
  // the targets in the order tperm (written by the pack launch's last
  // workgroups), which the pairwise kernel reads once per split.
  float4* sub = reinterpret_cast<float4*>(meta + g.ntiles);
  double* xt_sorted = reinterpret_cast<double*>(sub + g.ntiles * (WG / 64));
  hipLaunchKernelGGL(pack_fn(ker), dim3((unsigned)(g.ntiles + (nt + WG - 1) / WG)), dim3(WG), 0, stream, r_src, n_src,
                     v_src, ns, sperm, packed, meta, sub, g.ntiles, r_trg, nt, tperm, xt_sorted);
  SCTL_HIP_TRY(hipGetLastError());
  const double scale = kernel_scale(ker) * fold_factor(ker, newton ? 1 : 0);
  const EvalMixFn fn = mix_fn(ker, newton, g.ti);
  size_t dyn = 0;
  SCTL_TRY(occupancy_lds((const void*)fn, shape.occ, sizeof(float4) * TILE * g.Q, &dyn));
  if (!stream_capturing(stream)) profile_count(ker, nt * ns);
  // Groups of g.group splits: tiles [y0*tps, (y0+gy)*tps) per launch.
  for (int64_t y0 = 0; y0 < g.nsplit; y0 += g.group) {
    const int64_t gy = std::min(g.group, g.nsplit - y0);
    const int64_t tile0 = y0 * g.tps, ntiles = std::min(g.ntiles - tile0, gy * g.tps);
    const bool direct = gy == 1;
    {
      ProfileScope prof(stream);
      hipLaunchKernelGGL(fn, split_grid(g.tb, gy, WG), dim3(WG), dyn, stream, xt_sorted, nt, tperm,
                         packed + tile0 * (int64_t)(TILE * g.Q), meta + tile0, sub + tile0 * (WG / 64), ntiles, g.tps,
                         direct ? v_trg : part,
                         direct ? 1 : 0, scale, ctx);
    }
    SCTL_HIP_TRY(hipGetLastError());
    if (!direct) {
      const int64_t n = nt * K1;
      hipLaunchKernelGGL(k_reduce_splits_perm, dim3((unsigned)((4 * n + WG - 1) / WG)), dim3(WG), 0, stream, part, nt, K1,
                         (int)gy, tperm, v_trg, scale);
      SCTL_HIP_TRY(hipGetLastError());
    }
  }
  return SCTL_AMD_SUCCESS;
}

namespace {
// Workspace of one mixed Eval: orders, packed tiles, tile centres and the
// sorts' workspace, then the split partial sums last.
struct MixedCarve {
  size_t sperm, tperm, packed, meta, work, part, fixed_bytes;
};
MixedCarve mixed_carve(const MixedGeometry& g, int64_t nt, int64_t ns) {
  MixedCarve m;
  Carve c;
  m.sperm = c.take(sizeof(uint32_t) * ns);
  m.tperm = c.take(sizeof(uint32_t) * nt);
  m.packed = c.take(g.packed_bytes);
  m.meta = c.take(g.meta_bytes);
  m.work = c.take(mixed_order_bytes(nt, ns));
  m.part = c.take(0);
  m.fixed_bytes = m.part;
  return m;
}
}  // namespace

size_t mixed_fixed_bytes(const MixedGeometry& g, int64_t nt, int64_t ns) { return mixed_carve(g, nt, ns).fixed_bytes; }

int mixed_eval_in(int ker, int newton, const MixedGeometry& g, const DenseShape& shape, const double* r_trg,
                  int64_t nt, const double* r_src, const double* n_src, const double* v_src, int64_t ns,
                  double* v_trg, double ctx, char* w, hipStream_t stream) {
  if (nt <= 0 || ns <= 0) return SCTL_AMD_SUCCESS;
  const MixedCarve m = mixed_carve(g, nt, ns);
  uint32_t* sperm = reinterpret_cast<uint32_t*>(w + m.sperm);
  uint32_t* tperm = reinterpret_cast<uint32_t*>(w + m.tperm);
  SCTL_TRY(mixed_order(r_trg, nt, r_src, ns, tperm, sperm, w + m.work, stream));
  return mixed_apply(ker, newton, g, shape, r_trg, nt, tperm, r_src, n_src, v_src, ns, sperm,
                     reinterpret_cast<float4*>(w + m.packed), reinterpret_cast<double4*>(w + m.meta),
                     g.part_bytes ? reinterpret_cast<double*>(w + m.part) : nullptr, v_trg, ctx, stream);
}

int launch_eval_mixed(int ker, int newton, const double* r_trg, int64_t nt, const double* r_src,
                      const double* n_src, const double* v_src, int64_t ns, double* v_trg, const DenseShape& shape,
                      double ctx, hipStream_t stream) {
  if (nt <= 0 || ns <= 0) return SCTL_AMD_SUCCESS;
  MixedGeometry g;
  SCTL_TRY(mixed_geometry(ker, nt, ns, shape, &g));
  // One allocation for the call: orders, packed tiles, the sorts' workspace
  // and the partial sums.
  Scratch buf;
  SCTL_TRY(buf.alloc(mixed_fixed_bytes(g, nt, ns) + g.part_bytes, stream));
  return mixed_eval_in(ker, newton, g, shape, r_trg, nt, r_src, n_src, v_src, ns, v_trg, ctx, buf.as<char>(),
                       stream);
}

}  // namespace sctl_amd
