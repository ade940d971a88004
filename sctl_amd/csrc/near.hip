// Near-field operators of BoundaryIntegralOp on the device.
//
// BoundaryIntegralOp keeps, per element, a target-major matrix K_near of
// near-interaction corrections (boundary_integral.txx:1034-1058):
// K_near_e has near_elem_cnt[e]*KDIM1_ rows and elem_nds_cnt[e]*KDIM0
// columns, KDIM1_ = KDIM1 (or KDIM1/3 when the output is contracted with the
// target normal).  Three steps of the reference run here:
//
//  * near-block partition (:1009-1032)            sctl_amd_near_blocks (host)
//  * KNearSubtract (:1165-1240): per block
//      K_near_[t*KDIM1_+k1][s*KDIM0+k0] -= w_s * M[s*KDIM0+k0][t*KDIM1+k1]
//    with M the scaled KernelMatrix of the block (optionally contracted with
//    the target normal, :1193-1209)              k_near_matrix
//  * ComputeNearInterac, matrix route (:1324-1340): per block
//      U_near[block rows] = K_near_ * F[element's nodes]      k_near_gemv
//    then ScatterForward and the per-target sum (:1368-1379,
//    comm.txx:1497-1508)                                      k_near_accumulate
//
// The blocks are the plan's blocks (target range = the block's near targets,
// source range = the element's far nodes, which are the element's nodes when
// FarFieldDensityOperatorTranspose is the identity, boundary_integral.txx:531).
//
// All three kernels are bound by HBM traffic (8 bytes of K_near per matrix
// entry, read and written by the subtract, read by the GEMV); the matrix
// entries are FP64 whatever the plan's precision.  Work is cut into wave
// tiles listed once per layout (tile -> block, no search at run time), each
// a contiguous range of K_near so that no cache line is shared between
// waves (or XCDs) except at tile ends, blocks longest first:
//   near matrix  a stripe of 16 targets x all the block's sources: one
//                source per lane, 64 at a time, the target loop
//                wave-uniform (targets broadcast from lanes 0..15), so every
//                row segment written is 64*KDIM0 consecutive doubles;
//   GEMV         16 rows, 8 at a time, the 64 lanes along the rows (64
//                consecutive entries of a row per load), the 8 row sums
//                reduced across the wave by a halving butterfly.
// Nothing here allocates or issues a memset with device pointers, so the
// calls can be captured in a hipGraph.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "internal.hpp"
#include "kernels.hpp"
#include "plan.hpp"

namespace sctl_amd {

namespace {

constexpr int NWG = 256;        // 4 independent waves per workgroup
constexpr int MAT_ST = 16;      // targets per near-matrix stripe (one wave)
constexpr int GEMV_ROWS = 16;   // rows per GEMV wave tile
constexpr int GEMV_GR = 8;      // GEMV rows walked together
constexpr int NEAR_NEWTON = 2;  // KernelMatrix at digits -1 (boundary_integral.txx:1197,1211)

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double readlane_d(double v, int l) {
  const long long i = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(i & 0xffffffffLL), l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(i >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

typedef unsigned int v2u __attribute__((__vector_size__(8)));

// A wave-uniform pointer kept in scalar registers (the buffer resources built
// from it below must be).
__device__ __forceinline__ double* uniform_ptr(double* p) {
  const uint64_t a = reinterpret_cast<uint64_t>(p);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)a);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(a >> 32));
  return reinterpret_cast<double*>(((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ double bpermute_d(double v, int addr) {
  const long long i = __double_as_longlong(v);
  const unsigned lo = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(i & 0xffffffffLL));
  const unsigned hi = (unsigned)__builtin_amdgcn_ds_bpermute(addr, (int)(i >> 32));
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// A row of K_near as a raw buffer (gfx9 resource word 3: 32-bit data format):
// accesses at byte offsets outside [0, extent) are dropped (stores) or read 0
// (loads) by the hardware's range check.  a = the row start's offset in
// entries past a 128-byte line boundary.
struct Row {
  __amdgpu_buffer_rsrc_t rsrc;
  int a;
};

// Row k of target t (dof rows per target) of a stripe of nts targets; the
// rows of targets past the stripe get extent 0.
__device__ __forceinline__ Row row_of(double* Ks, int t, int nts, int64_t ld, int dof = 1, int k = 0) {
  // Built where it is used: kept opaque so that the resources of all rows
  // are not hoisted out of the caller's loop (64 scalar registers).
  uint64_t base = reinterpret_cast<uint64_t>(Ks);
  asm volatile("" : "+s"(base), "+s"(ld), "+s"(nts));
  double* p = reinterpret_cast<double*>(base) + ((int64_t)min(t, nts - 1) * dof + k) * ld;
  Row R;
  R.a = (int)((reinterpret_cast<uintptr_t>(p) >> 3) & 15);
  R.rsrc = __builtin_amdgcn_make_buffer_rsrc(p, 0, t < nts ? (int)(ld * 8) : 0, 0x00020000);
  return R;
}

// K_near is written once and not read again before it leaves the caches
// (13.5 GB and more), so the stores are nontemporal.
__device__ __forceinline__ void buffer_store_d(double v, __amdgpu_buffer_rsrc_t r, int off) {
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, off, 0, 2);
}

__device__ __forceinline__ double buffer_load_d(__amdgpu_buffer_rsrc_t r, int off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, off, 0, 0));
}

// KNearSubtract / store of one layout's blocks.  NDOT: contract the KDIM1
// outputs with the target normal (KDIM1 == 3 -> one row per target).
// A wave owns one stripe (MAT_ST targets x all sources of a block), a
// contiguous range of K_near: lanes 0..MAT_ST-1 hold the stripe's targets,
// the wave walks the sources 64 at a time (one per lane) and, per chunk,
// writes every target row's 64*KDIM0 entries.
//
// Rows are written as whole 128-byte lines wherever the LDS allows it: a
// row segment that starts inside a line would leave that line half-written
// until the next chunk, by when it has usually left L2 and the memory must
// read it back to merge the halves.  Rows of one entry per source (the
// Laplace kernels, and kernels contracted to one output per target) rotate
// the chunk across lanes (the W == 1 branch); rows of K0 entries per source
// with at most 3 rows per target go through an LDS stage (ALIGNED_WIDE).
// Kernels with more rows per target (9 outputs, uncontracted) keep per-lane
// stores of their K0 entries.
//
// Stores and loads share one completion counter on this GPU, so a wait for
// a chunk's source records would also wait for the previous chunk's stores.
// The kernels therefore prefetch the next chunk's records before the target
// loop and issue a fixed number of stores per chunk, so that the wait for
// the prefetched records can leave the chunk's stores in flight: the aligned
// branches through the range check of their buffer stores (rows past the
// stripe have extent 0), the per-lane store mode (at most 3 entries per
// source) by repeating the stripe's last target and the block's last source
// (a repeat writes the identical value to the same entry).  The per-lane
// subtract mode cannot repeat an update and keeps its stores exact; the
// W == 1 subtract loads the old entries of all MAT_ST rows before rewriting
// any of them, where a row-by-row update would wait for the previous row's
// store before each load.
template <int KER, bool NDOT, bool SUB>
__global__ void __launch_bounds__(NWG)
k_near_matrix(const int32_t* __restrict__ tile_blk, const int64_t* __restrict__ tile_dsp, int64_t ntiles,
              const int64_t* __restrict__ trg_dsp, const int32_t* __restrict__ trg_cnt,
              const int64_t* __restrict__ src_dsp, const int32_t* __restrict__ src_cnt,
              const int64_t* __restrict__ mat_off, const double* __restrict__ xt, const double* __restrict__ n_trg,
              const double* __restrict__ xs, const double* __restrict__ xn, const double* __restrict__ wts,
              double* __restrict__ K, double scale, double mu) {
  using Kt = Ker<KER>;
  constexpr int K0 = Kt::K0, K1 = Kt::K1, ND = Kt::ND;
  constexpr int DOF = NDOT ? K1 / 3 : K1;
  // W: entries per (target, source).  The pipelined store (16 unrolled target
  // rows) of the wider rows only up to 3 entries: more unrolled rows would
  // cost the kernel its occupancy.
  constexpr int W = K0 * DOF;
  constexpr bool PIPE = !SUB && W <= 3;
  // Rows of more than one entry per source, at most 3 rows per target (the
  // LDS carries of a stripe: MAT_ST * DOF * 128 bytes per wave).
  constexpr bool ALIGNED_WIDE = W > 1 && DOF <= 3;
  static_assert(!NDOT || K1 % 3 == 0, "the target-normal contraction needs 3 outputs per row");
  const int lane = threadIdx.x & 63;
  __shared__ double tslot_all[(NWG / 64) * MAT_ST * 6];
  double* tslot = tslot_all + (threadIdx.x >> 6) * (MAT_ST * 6);
  const int64_t nwaves = (int64_t)gridDim.x * (NWG / 64);
  const int64_t first = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (NWG / 64) + (threadIdx.x >> 6)));
  for (int64_t tile = first; tile < ntiles; tile += nwaves) {
    const int b = __builtin_amdgcn_readfirstlane(tile_blk[tile]);
    const int t0 = (int)(tile - tile_dsp[b]) * MAT_ST;
    const int sc = src_cnt[b];
    const int nts = min(MAT_ST, trg_cnt[b] - t0);
    // The stripe's targets (and normals) in the wave's LDS slot, read back per
    // row by broadcast (one address for the whole wave): no VALU work per row
    // and no registers held across the chunk loop.
    wave_sync();
    if (lane < MAT_ST) {
      const int64_t tg = trg_dsp[b] + t0 + min(lane, nts - 1);
#pragma unroll
      for (int k = 0; k < 3; k++) tslot[lane * 6 + k] = xt[tg * 3 + k];
      if constexpr (NDOT) {
#pragma unroll
        for (int k = 0; k < 3; k++) tslot[lane * 6 + 3 + k] = n_trg[tg * 3 + k];
      }
    }
    wave_sync();
    const double* tx = tslot;
    const int64_t ld = (int64_t)sc * K0;
    double* Ks = uniform_ptr(K + mat_off[b] + (int64_t)t0 * DOF * ld);
    const int64_t sb = src_dsp[b];
    // This lane's source record of the chunk at s0: position, normal, weight
    // (the column clamped to the block's last source).
    auto load = [&](int s0, double(&r)[7]) {
      const int64_t gs = sb + min(s0 + lane, sc - 1);
      r[0] = xs[gs * 3 + 0];
      r[1] = xs[gs * 3 + 1];
      r[2] = xs[gs * 3 + 2];
#pragma unroll
      for (int k = 0; k < 3; k++) r[3 + k] = k < ND ? xn[gs * 3 + k] : 0.0;
      r[6] = wts ? wts[gs] : 1.0;
    };
    // Entries of target tt (wave-uniform) and this lane's source, written at
    // this lane's column offset col (entries) of each row; ws = the weight
    // times the kernel's scale factor (rows of more than one entry per source).
    auto entries = [&](int tt, const double(&r)[7], double ws, int col, bool active) {
      double u[K0][K1];
      const double n[3] = {r[3], r[4], r[5]};
      kmatrix<KER, NEAR_NEWTON>(u, tx[tt * 6 + 0] - r[0], tx[tt * 6 + 1] - r[1], tx[tt * 6 + 2] - r[2], n, mu);
      if constexpr (NDOT) {
        const double n0 = tx[tt * 6 + 3], n1 = tx[tt * 6 + 4], n2 = tx[tt * 6 + 5];
#pragma unroll
        for (int k1 = 0; k1 < DOF; k1++) {
          double* row = Ks + (int64_t)(tt * DOF + k1) * ld;
#pragma unroll
          for (int k0 = 0; k0 < K0; k0++) {
            double v = 0.0;
            v += (u[k0][k1 * 3 + 0] * ws) * n0;
            v += (u[k0][k1 * 3 + 1] * ws) * n1;
            v += (u[k0][k1 * 3 + 2] * ws) * n2;
            if constexpr (SUB) {
              if (active) row[col + k0] = row[col + k0] - v;
            } else {
              if (active) row[col + k0] = v;
            }
          }
        }
      } else {
#pragma unroll
        for (int k1 = 0; k1 < K1; k1++) {
          double* row = Ks + (int64_t)(tt * K1 + k1) * ld;
#pragma unroll
          for (int k0 = 0; k0 < K0; k0++) {
            const double v = u[k0][k1] * ws;
            if constexpr (SUB) {
              if (active) row[col + k0] = row[col + k0] - v;
            } else {
              if (active) row[col + k0] = v;
            }
          }
        }
      }
    };
    // The entry of target tt (wave-uniform) and this lane's source when a row
    // holds one entry per source (W == 1), computed as entries() does.
    auto value1 = [&](int tt, const double(&r)[7], double ws) -> double {
      double u[K0][K1];
      const double n[3] = {r[3], r[4], r[5]};
      kmatrix<KER, NEAR_NEWTON>(u, tx[tt * 6 + 0] - r[0], tx[tt * 6 + 1] - r[1], tx[tt * 6 + 2] - r[2], n, mu);
      if constexpr (NDOT) {
        double v = 0.0;
        v += (u[0][0] * ws) * tx[tt * 6 + 3];
        v += (u[0][1] * ws) * tx[tt * 6 + 4];
        v += (u[0][2] * ws) * tx[tt * 6 + 5];
        return v;
      } else {
        return u[0][0] * ws;
      }
    };
    double cur[7], nxt[7];
    load(0, cur);
    if constexpr (W == 1) {
      // One entry per source and row: the row segments are written as whole
      // 128-byte lines.  Row t starts a_t entries past a line boundary, so
      // in the chunk at s0 lane l writes column s0 - a_t + l: the chunk's
      // entries rotated by a_t lanes (ds_bpermute), the first a_t lanes
      // taking the previous chunk's last a_t entries from the wave's LDS
      // carry slots, and a final pass writes the last carry.  The stores go
      // through a buffer resource of the row's extent, so columns before
      // the row start (first chunk), past its end and rows past the stripe
      // (extent 0) are dropped by the hardware's range check: each chunk
      // issues exactly MAT_ST row stores with no branch.
      __shared__ double carry_all[(NWG / 64) * MAT_ST * 16];
      double* cs = carry_all + (threadIdx.x >> 6) * (MAT_ST * 16);
      const int nch = (sc + 63) >> 6;
      for (int c = 0; c <= nch; c++) {
        const int s0 = c * 64;
        asm volatile("" ::: "memory");
        const bool tail = c == nch;  // wave-uniform: the carry flush only
        const double ws = cur[6] * scale;
        if (!tail) load(min(s0 + 64, (nch - 1) * 64), nxt);
        double old[MAT_ST];
        if constexpr (SUB) {
#pragma unroll
          for (int t = 0; t < MAT_ST; t++) {
            const Row R = row_of(Ks, t, nts, ld);
            old[t] = buffer_load_d(R.rsrc, (s0 - R.a + lane) * 8);
          }
        }
#pragma unroll
        for (int t = 0; t < MAT_ST; t++) {
          // per-row values derived here, not kept across the chunk loop
          int nr = nts;
          asm volatile("" : "+s"(nr));
          const Row R = row_of(Ks, t, nr, ld);
          const double v = tail ? 0.0 : value1(min(t, nr - 1), cur, ws);
          const double rot = bpermute_d(v, ((lane - R.a) & 63) << 2);
          const double prev = cs[t * 16 + (lane & 15)];
          double out = lane < R.a ? prev : rot;
          if (lane < 16) cs[t * 16 + lane] = rot;
          if constexpr (SUB) out = old[t] - out;
          buffer_store_d(out, R.rsrc, (s0 - R.a + lane) * 8);
        }
#pragma unroll
        for (int i = 0; i < 7; i++) cur[i] = nxt[i];
      }
      continue;
    }
    if constexpr (ALIGNED_WIDE) {
      // K0 entries per source and up to 3 rows per target: each row's chunk
      // segment (64 K0 consecutive entries) goes through the wave's LDS
      // stage, lane l holding entries l K0 .. l K0 + K0 - 1, and is written
      // back as whole 128-byte lines: entry s0 K0 - a + 64 i + l by lane l in
      // store i, the first a entries from the row's carry (the previous
      // chunk's last 16 entries); a final pass writes the last carry.
      __shared__ double stage_all[(NWG / 64) * 64 * K0];
      __shared__ double carry_w[(NWG / 64) * MAT_ST * DOF * 16];
      double* stg = stage_all + (threadIdx.x >> 6) * (64 * K0);
      double* cw = carry_w + (threadIdx.x >> 6) * (MAT_ST * DOF * 16);
      const int nch = (sc + 63) >> 6;
      for (int c = 0; c <= nch; c++) {
        const int s0 = c * 64;
        asm volatile("" ::: "memory");
        const bool tail = c == nch;  // wave-uniform: the carry flush only
        const double ws = cur[6] * scale;
        if (!tail) load(min(s0 + 64, (nch - 1) * 64), nxt);
#pragma unroll
        for (int t = 0; t < MAT_ST; t++) {
          int nr = nts;
          asm volatile("" : "+s"(nr));
          double e[DOF][K0];
          if (!tail) {
            const int tt = min(t, nr - 1);
            double u[K0][K1];
            const double n[3] = {cur[3], cur[4], cur[5]};
            kmatrix<KER, NEAR_NEWTON>(u, tx[tt * 6 + 0] - cur[0], tx[tt * 6 + 1] - cur[1], tx[tt * 6 + 2] - cur[2], n,
                                     mu);
#pragma unroll
            for (int k1 = 0; k1 < DOF; k1++)
#pragma unroll
              for (int k0 = 0; k0 < K0; k0++) {
                if constexpr (NDOT) {
                  double v = 0.0;
                  v += (u[k0][k1 * 3 + 0] * ws) * tx[tt * 6 + 3];
                  v += (u[k0][k1 * 3 + 1] * ws) * tx[tt * 6 + 4];
                  v += (u[k0][k1 * 3 + 2] * ws) * tx[tt * 6 + 5];
                  e[k1][k0] = v;
                } else {
                  e[k1][k0] = u[k0][k1] * ws;
                }
              }
          } else {
#pragma unroll
            for (int k1 = 0; k1 < DOF; k1++)
#pragma unroll
              for (int k0 = 0; k0 < K0; k0++) e[k1][k0] = 0.0;
          }
#pragma unroll
          for (int k1 = 0; k1 < DOF; k1++) {
            const Row R = row_of(Ks, t, nr, ld, DOF, k1);
            double* cr = cw + (t * DOF + k1) * 16;
#pragma unroll
            for (int k0 = 0; k0 < K0; k0++) stg[lane * K0 + k0] = e[k1][k0];
            wave_sync();
            double out[K0];
#pragma unroll
            for (int i = 0; i < K0; i++) {
              const int idx = i * 64 + lane - R.a;
              out[i] = idx >= 0 ? stg[idx] : cr[16 + idx];
            }
            const double last = stg[64 * K0 - 16 + (lane & 15)];
            wave_sync();
            if (lane < 16) cr[lane] = last;
#pragma unroll
            for (int i = 0; i < K0; i++) {
              const int off = (s0 * K0 - R.a + i * 64 + lane) * 8;
              double v = out[i];
              if constexpr (SUB) v = buffer_load_d(R.rsrc, off) - v;
              buffer_store_d(v, R.rsrc, off);
            }
          }
        }
#pragma unroll
        for (int i = 0; i < 7; i++) cur[i] = nxt[i];
      }
      continue;
    }
    for (int s0 = 0; s0 < sc; s0 += 64) {
      // keep the per-row target reads inside the chunk loop (no registers
      // held across chunks); no wait is issued for it
      asm volatile("" ::: "memory");
      const int col = min(s0 + lane, sc - 1) * K0;
      const double ws = cur[6] * scale;  // loaded a chunk ago
      load(s0 + 64, nxt);  // clamped: in range past the last chunk too
      if constexpr (PIPE) {
#pragma unroll
        for (int t = 0; t < MAT_ST; t++) entries(min(t, nts - 1), cur, ws, col, true);
      } else {
        const bool active = s0 + lane < sc;
        for (int t = 0; t < nts; t++) entries(t, cur, ws, col, active);
      }
#pragma unroll
      for (int i = 0; i < 7; i++) cur[i] = nxt[i];
    }
  }
}

// U_near[block rows] = K_near_ * F_element.  Blocks without a matrix
// (offset < 0) or without sources get zero rows, as the reference's GEMV loop
// writes them (boundary_integral.txx:1329-1332).  A wave owns GEMV_ROWS
// consecutive rows (a contiguous range of K_near) and walks them GEMV_GR at a
// time with its 64 lanes along the rows: every load is 64 consecutive
// entries of one row, each lane's F entries serve all GEMV_GR rows, and two
// column chunks per step keep 2*GEMV_GR loads in flight.  Rows past the
// block's last (a tail tile) re-read its last row and columns past the row
// end re-read its last entry, with F taken as 0, so every load stays inside
// the block and no load is predicated.  The GEMV_GR row sums are reduced
// across the wave by a halving butterfly: each exchange hands half of the
// rows still held to the partner lane (4 + 2 + 1 + 3 exchanges for 8 rows).
__global__ void __launch_bounds__(NWG)
k_near_gemv(const int32_t* __restrict__ tile_blk, const int64_t* __restrict__ tile_dsp, int64_t ntiles,
            const int64_t* __restrict__ trg_dsp, const int32_t* __restrict__ trg_cnt,
            const int64_t* __restrict__ src_dsp, const int32_t* __restrict__ src_cnt,
            const int64_t* __restrict__ mat_off, int K0, int dof, const double* __restrict__ K,
            const double* __restrict__ F, double* __restrict__ U) {
  static_assert(GEMV_GR == 8 && GEMV_ROWS % GEMV_GR == 0, "the butterfly below reduces 8 rows");
  const int lane = threadIdx.x & 63;
  const int64_t nwaves = (int64_t)gridDim.x * (NWG / 64);
  const int64_t first = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (NWG / 64) + (threadIdx.x >> 6)));
  for (int64_t tile = first; tile < ntiles; tile += nwaves) {
    const int b = __builtin_amdgcn_readfirstlane(tile_blk[tile]);
    const int64_t local = tile - tile_dsp[b];
    const int64_t rows = (int64_t)trg_cnt[b] * dof;
    const int64_t ld = (int64_t)src_cnt[b] * K0;
    const int64_t off = mat_off[b];
    const double* Fb = F + src_dsp[b] * K0;
    double* Ub = U + trg_dsp[b] * dof;
    for (int rr = 0; rr < GEMV_ROWS; rr += GEMV_GR) {
      const int64_t row0 = local * GEMV_ROWS + rr;
      if (row0 >= rows) break;
      const int64_t nr = min((int64_t)GEMV_GR, rows - row0);
      double acc[GEMV_GR];
#pragma unroll
      for (int r = 0; r < GEMV_GR; r++) acc[r] = 0.0;
      if (off >= 0) {
        const double* Kr = K + off + row0 * ld;
        for (int64_t j0 = 0; j0 < ld; j0 += 128) {
          const int64_t ja = j0 + lane, jb = ja + 64;
          const int64_t ca = min(ja, ld - 1), cb = min(jb, ld - 1);
          const double fa = ja < ld ? Fb[ca] : 0.0, fb = jb < ld ? Fb[cb] : 0.0;
          double ka[GEMV_GR], kb[GEMV_GR];
#pragma unroll
          for (int r = 0; r < GEMV_GR; r++) {
            const double* Kq = Kr + min((int64_t)r, nr - 1) * ld;
            ka[r] = Kq[ca];
            kb[r] = Kq[cb];
          }
#pragma unroll
          for (int r = 0; r < GEMV_GR; r++) {
            ka[r] = ja < ld ? ka[r] : 0.0;
            kb[r] = jb < ld ? kb[r] : 0.0;
            acc[r] = fmad(kb[r], fb, fmad(ka[r], fa, acc[r]));
          }
        }
      }
#pragma unroll
      for (int h = GEMV_GR / 2, m = 32; h >= 1; h >>= 1, m >>= 1) {
        const bool up = lane & m;
#pragma unroll
        for (int i = 0; i < h; i++) {
          const double send = up ? acc[i] : acc[i + h];
          const double keep = up ? acc[i + h] : acc[i];
          acc[i] = keep + __shfl_xor(send, m, 64);
        }
      }
#pragma unroll
      for (int m = 4; m >= 1; m >>= 1) acc[0] += __shfl_xor(acc[0], m, 64);
      const int rl = lane >> 3;  // row of this lane's sum: bits 5, 4, 3 of the lane
      if ((lane & 7) == 0 && rl < nr) Ub[row0 + rl] = acc[0];
    }
  }
}

// ScatterForward + per-target accumulation: for every target i and every
// component k, U[i*DOF+k] += U_near[scatter[dsp_i + j]*DOF + k] for j in
// order, the reference's order of additions.  Indices outside the arrays are
// skipped (the host checks them when it has them).
template <int DOF>
__global__ void __launch_bounds__(NWG)
k_near_accumulate(int64_t ntrg, const int64_t* __restrict__ cnt, const int64_t* __restrict__ dsp, int64_t n_scatter,
                  const int64_t* __restrict__ scatter, int64_t n_near, const double* __restrict__ U_near,
                  double* __restrict__ U) {
  const int64_t i = (int64_t)blockIdx.x * NWG + threadIdx.x;
  if (i >= ntrg) return;
  const int64_t c = cnt[i], d = dsp[i];
  double acc[DOF];
#pragma unroll
  for (int k = 0; k < DOF; k++) acc[k] = U[i * DOF + k];
  for (int64_t j = 0; j < c; j++) {
    const int64_t q = d + j;
    if (q < 0 || q >= n_scatter) continue;
    const int64_t src = scatter[q];
    if (src < 0 || src >= n_near) continue;
#pragma unroll
    for (int k = 0; k < DOF; k++) acc[k] += U_near[src * DOF + k];
  }
#pragma unroll
  for (int k = 0; k < DOF; k++) U[i * DOF + k] = acc[k];
}

using NearMatFn = void (*)(const int32_t*, const int64_t*, int64_t, const int64_t*, const int32_t*, const int64_t*,
                           const int32_t*, const int64_t*, const double*, const double*, const double*,
                           const double*, const double*, double*, double, double);

template <int KER, bool NDOT> NearMatFn near_mat_fn(bool sub) {
  return sub ? k_near_matrix<KER, NDOT, true> : k_near_matrix<KER, NDOT, false>;
}
NearMatFn near_mat_fn(int ker, bool ndot, bool sub) {
  return with_kernel(ker, [&](auto kc) -> NearMatFn {
    constexpr int K = decltype(kc)::value;
    if constexpr (Ker<K>::K1 % 3 == 0) {
      if (ndot) return near_mat_fn<K, true>(sub);
    }
    return near_mat_fn<K, false>(sub);
  });
}

// Flattens per-block tile counts into a device tile list, blocks in
// decreasing order of cost per tile (longest-processing-time first; a
// block's tiles stay consecutive).
int build_tiles(const std::vector<int64_t>& per_block, const std::vector<int64_t>& cost, TileList* out) {
  const size_t nb = per_block.size();
  std::vector<size_t> order(nb);
  std::iota(order.begin(), order.end(), 0);
  std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost[a] > cost[b]; });
  std::vector<int64_t> dsp(nb);
  int64_t total = 0;
  for (size_t b : order) {
    dsp[b] = total;
    total += per_block[b];
  }
  SCTL_CHECK_ARG(total < (int64_t)1 << 34, "near-field layout has too many tiles");
  std::vector<int32_t> blk((size_t)total);
  for (size_t b = 0; b < nb; b++)
    for (int64_t i = 0; i < per_block[b]; i++) blk[(size_t)(dsp[b] + i)] = (int32_t)b;
  out->ntiles = total;
  SCTL_TRY(dev_upload(&out->d_tile_blk, blk));
  SCTL_TRY(dev_upload(&out->d_tile_dsp, dsp));
  return SCTL_AMD_SUCCESS;
}

int tile_grid(int64_t ntiles) {
  DeviceInfo dev;
  if (device_info(&dev) != SCTL_AMD_SUCCESS) return 1;
  const int64_t want = (ntiles + NWG / 64 - 1) / (NWG / 64);
  return (int)std::max<int64_t>(1, std::min<int64_t>(want, (int64_t)dev.num_cus * 64));
}


int near_matrix_launch(sctl_amd_plan_s* p, bool sub, const double* wts, const double* n_trg, double* K,
                       hipStream_t s) {
  if (p->mat_tiles.ntiles == 0) return SCTL_AMD_SUCCESS;
  if (!stream_capturing(s)) {
    int64_t inter = 0;
    for (int64_t b = 0; b < p->nblocks; b++)
      if (p->near_off[b] >= 0) inter += p->trg_cnt[b] * p->src_cnt[b];
    profile_count(p->kernel, inter);
  }
  ProfileScope prof(s);
  hipLaunchKernelGGL(near_mat_fn(p->kernel, p->near_ndot, sub), dim3(tile_grid(p->mat_tiles.ntiles)), dim3(NWG), 0,
                     s, p->mat_tiles.d_tile_blk, p->mat_tiles.d_tile_dsp, p->mat_tiles.ntiles, p->d_trg_dsp,
                     p->d_trg_cnt, p->d_src_dsp, p->d_src_cnt, p->d_near_off, p->d_xt, n_trg, p->d_xs, p->d_xn, wts,
                     K, kernel_scale(p->kernel), p->ctx);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

int near_gemv_launch(sctl_amd_plan_s* p, const double* K, const double* F, double* U, hipStream_t s) {
  if (p->gemv_tiles.ntiles == 0) return SCTL_AMD_SUCCESS;
  ProfileScope prof(s);
  hipLaunchKernelGGL(k_near_gemv, dim3(tile_grid(p->gemv_tiles.ntiles)), dim3(NWG), 0, s, p->gemv_tiles.d_tile_blk,
                     p->gemv_tiles.d_tile_dsp, p->gemv_tiles.ntiles, p->d_trg_dsp, p->d_trg_cnt, p->d_src_dsp,
                     p->d_src_cnt, p->d_near_off, p->K0, p->near_dof, K, F, U);
  SCTL_HIP_TRY(hipGetLastError());
  return SCTL_AMD_SUCCESS;
}

// Host staging of the HOST_PTRS mode: copies in, runs, copies `out` back.
struct Staged {
  Scratch buf;
  const void* host = nullptr;
  int64_t bytes = 0;
  int in(const void* h, int64_t b, hipStream_t s) {
    host = h;
    bytes = b;
    if (!h || b <= 0) return SCTL_AMD_SUCCESS;
    SCTL_TRY(buf.alloc((size_t)b, s));
    SCTL_HIP_TRY(hipMemcpyAsync(buf.as<void>(), h, (size_t)b, hipMemcpyHostToDevice, s));
    return SCTL_AMD_SUCCESS;
  }
  int out(void* h, hipStream_t s) {
    if (!h || bytes <= 0) return SCTL_AMD_SUCCESS;
    SCTL_HIP_TRY(hipMemcpyAsync(h, buf.as<void>(), (size_t)bytes, hipMemcpyDeviceToHost, s));
    return SCTL_AMD_SUCCESS;
  }
  template <class T> T* dev() const { return host && bytes > 0 ? buf.as<T>() : nullptr; }
};

}  // namespace
}  // namespace sctl_amd

using namespace sctl_amd;

extern "C" {

int sctl_amd_near_blocks(int64_t nelem, const int64_t* near_elem_cnt, const int64_t* near_elem_dsp, int nthreads,
                         int64_t* nblk, int64_t* blk_elem, int64_t* blk_t0, int64_t* blk_cnt) {
  SCTL_CHECK_ARG(nblk, "nblk is NULL");
  SCTL_CHECK_ARG(nelem >= 0, "negative element count");
  SCTL_CHECK_ARG(nelem == 0 || (near_elem_cnt && near_elem_dsp), "NULL near-list array");
  SCTL_CHECK_ARG(nthreads >= 1, "nthreads must be >= 1");
  for (int64_t e = 0; e < nelem; e++)
    SCTL_CHECK_ARG(near_elem_cnt[e] >= 0, "negative near-target count (element " + std::to_string(e) + ")");
  const int64_t n_near = nelem ? near_elem_dsp[nelem - 1] + near_elem_cnt[nelem - 1] : 0;
  const int64_t grain = std::max<int64_t>(1, n_near / (4 * (int64_t)nthreads));
  int64_t count = 0;
  for (int64_t e = 0; e < nelem; e++) count += (near_elem_cnt[e] + grain - 1) / grain;
  const bool fill = blk_elem || blk_t0 || blk_cnt;
  SCTL_CHECK_ARG(!fill || (blk_elem && blk_t0 && blk_cnt), "give all three block arrays or none");
  if (fill) {
    int64_t b = 0;
    for (int64_t e = 0; e < nelem; e++) {
      for (int64_t t0 = 0; t0 < near_elem_cnt[e]; t0 += grain, b++) {
        blk_elem[b] = e;
        blk_t0[b] = t0;
        blk_cnt[b] = std::min(grain, near_elem_cnt[e] - t0);
      }
    }
  }
  *nblk = count;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_near_layout(sctl_amd_plan p, const int64_t* blk_off, int trg_normal_dot, int64_t* k_len) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  SCTL_CHECK_ARG(!trg_normal_dot || p->K1 % 3 == 0,
                 "the target-normal contraction needs a kernel whose outputs per target are a multiple of 3");
  const int dof = trg_normal_dot ? p->K1 / 3 : p->K1;
  const int64_t nb = p->nblocks;
  std::vector<int64_t> off(nb), size(nb);
  int64_t len = 0;
  for (int64_t b = 0; b < nb; b++) {
    // rows are addressed with 32-bit byte offsets
    SCTL_CHECK_ARG((int64_t)p->src_cnt[b] * p->K0 < ((int64_t)1 << 28),
                   "block " + std::to_string(b) + " has too many sources for a near-field matrix row");
    size[b] = p->trg_cnt[b] * dof * p->src_cnt[b] * p->K0;
    if (blk_off) {
      off[b] = blk_off[b] < 0 ? -1 : blk_off[b];
      if (off[b] >= 0) {
        SCTL_CHECK_ARG(off[b] <= ((int64_t)1 << 62) - size[b], "block matrix offset too large");
        len = std::max(len, off[b] + size[b]);
      }
    } else {
      off[b] = len;
      len += size[b];
    }
  }
  if (blk_off) {  // block matrices must not overlap (their writes would race)
    std::vector<int64_t> order;
    for (int64_t b = 0; b < nb; b++)
      if (off[b] >= 0 && size[b] > 0) order.push_back(b);
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return off[a] < off[b]; });
    for (size_t i = 1; i < order.size(); i++)
      SCTL_CHECK_ARG(off[order[i - 1]] + size[order[i - 1]] <= off[order[i]],
                     "block matrices overlap (blocks " + std::to_string(order[i - 1]) + " and " +
                         std::to_string(order[i]) + ")");
  }
  std::vector<int64_t> mat(nb), gemv(nb), cost(nb);
  for (int64_t b = 0; b < nb; b++) {
    const int64_t tc = p->trg_cnt[b], sc = p->src_cnt[b];
    mat[b] = (off[b] >= 0 && tc > 0 && sc > 0) ? (tc + MAT_ST - 1) / MAT_ST : 0;
    gemv[b] = tc > 0 ? (tc * dof + GEMV_ROWS - 1) / GEMV_ROWS : 0;
    cost[b] = off[b] >= 0 ? sc : 0;
  }
  p->free_near();
  int rc = dev_upload(&p->d_near_off, off);
  if (!rc) rc = build_tiles(mat, cost, &p->mat_tiles);
  if (!rc) rc = build_tiles(gemv, cost, &p->gemv_tiles);
  if (rc) {
    p->free_near();
    return rc;
  }
  p->near_off = off;
  p->near_dof = dof;
  p->near_ndot = trg_normal_dot != 0;
  p->near_len = len;
  p->has_near = true;
  if (k_len) *k_len = len;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_near_matrix(sctl_amd_plan p, int mode, const double* wts, const double* n_trg, double* K,
                              uint32_t flags, void* stream) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  SCTL_CHECK_ARG(p->has_geometry, "plan has no geometry (call sctl_amd_plan_set_geometry first)");
  SCTL_CHECK_ARG(p->has_near, "plan has no near-field layout (call sctl_amd_plan_near_layout first)");
  SCTL_CHECK_ARG(!needs_ctx(p->kernel) || p->has_ctx,
                 "the kernel needs a wavenumber (call sctl_amd_plan_set_ctx first)");
  SCTL_CHECK_ARG(mode == SCTL_AMD_NEAR_SUBTRACT || mode == SCTL_AMD_NEAR_STORE, "unknown near-matrix mode");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(p->near_len == 0 || K, "K is NULL");
  SCTL_CHECK_ARG(!p->near_ndot || p->nt_total == 0 || n_trg, "n_trg is NULL for a target-normal layout");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  const bool sub = mode == SCTL_AMD_NEAR_SUBTRACT;
  const double* nt = p->near_ndot ? n_trg : nullptr;
  if (!(flags & SCTL_AMD_HOST_PTRS)) return near_matrix_launch(p, sub, wts, nt, K, s);
  Staged sw, sn, sk;
  SCTL_TRY(sw.in(wts, (int64_t)sizeof(double) * p->ns_total, s));
  SCTL_TRY(sn.in(nt, (int64_t)sizeof(double) * p->nt_total * 3, s));
  // K is copied in for both modes: the store mode leaves the entries outside
  // the block matrices as the caller had them.
  SCTL_TRY(sk.in(K, (int64_t)sizeof(double) * p->near_len, s));
  SCTL_TRY(near_matrix_launch(p, sub, sw.dev<double>(), sn.dev<double>(), sk.dev<double>(), s));
  SCTL_TRY(sk.out(K, s));
  SCTL_HIP_TRY(hipStreamSynchronize(s));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_plan_near_gemv(sctl_amd_plan p, const double* K, const double* F, double* U_near, uint32_t flags,
                            void* stream) {
  SCTL_CHECK_ARG(p, "plan is NULL");
  SCTL_CHECK_ARG(p->has_near, "plan has no near-field layout (call sctl_amd_plan_near_layout first)");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(p->near_len == 0 || K, "K is NULL");
  SCTL_CHECK_ARG(p->ns_total == 0 || F, "F is NULL");
  SCTL_CHECK_ARG(p->nt_total == 0 || U_near, "U_near is NULL");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  if (!(flags & SCTL_AMD_HOST_PTRS)) return near_gemv_launch(p, K, F, U_near, s);
  Staged sk, sf, su;
  SCTL_TRY(sk.in(K, (int64_t)sizeof(double) * p->near_len, s));
  SCTL_TRY(sf.in(F, (int64_t)sizeof(double) * p->ns_total * p->K0, s));
  SCTL_TRY(su.in(U_near, (int64_t)sizeof(double) * p->nt_total * p->near_dof, s));
  SCTL_TRY(near_gemv_launch(p, sk.dev<double>(), sf.dev<double>(), su.dev<double>(), s));
  SCTL_TRY(su.out(U_near, s));
  SCTL_HIP_TRY(hipStreamSynchronize(s));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_near_accumulate(int64_t ntrg, int dof, const int64_t* near_trg_cnt, const int64_t* near_trg_dsp,
                             int64_t n_scatter, const int64_t* scatter_index, int64_t n_near, const double* U_near,
                             double* U, uint32_t flags, void* stream) {
  SCTL_CHECK_ARG(dof == 1 || dof == 3 || dof == 4 || dof == 9, "dof must be 1, 3, 4 or 9");
  SCTL_CHECK_ARG(ntrg >= 0 && n_scatter >= 0 && n_near >= 0, "negative size");
  SCTL_CHECK_ARG(ntrg == 0 || (near_trg_cnt && near_trg_dsp && U), "NULL target array");
  SCTL_CHECK_ARG(n_scatter == 0 || scatter_index, "scatter_index is NULL");
  SCTL_CHECK_ARG(n_near == 0 || U_near, "U_near is NULL");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  if (ntrg == 0) return SCTL_AMD_SUCCESS;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  const int64_t grid = (ntrg + NWG - 1) / NWG;
  SCTL_CHECK_ARG(grid <= 0x7fffffffLL, "too many targets for one launch");
  auto launch = [&](const int64_t* cnt, const int64_t* dsp, const int64_t* sc, const double* un, double* u) -> int {
    ProfileScope prof(s);
    auto fn = dof == 1 ? k_near_accumulate<1>
                       : (dof == 3 ? k_near_accumulate<3> : (dof == 4 ? k_near_accumulate<4> : k_near_accumulate<9>));
    hipLaunchKernelGGL(fn, dim3((unsigned)grid), dim3(NWG), 0, s, ntrg, cnt, dsp, n_scatter, sc, n_near, un, u);
    SCTL_HIP_TRY(hipGetLastError());
    return SCTL_AMD_SUCCESS;
  };
  if (!(flags & SCTL_AMD_HOST_PTRS)) return launch(near_trg_cnt, near_trg_dsp, scatter_index, U_near, U);
  for (int64_t i = 0; i < ntrg; i++) {
    SCTL_CHECK_ARG(near_trg_cnt[i] >= 0, "negative near count (target " + std::to_string(i) + ")");
    if (near_trg_cnt[i] == 0) continue;
    SCTL_CHECK_ARG(near_trg_dsp[i] >= 0 && near_trg_dsp[i] + near_trg_cnt[i] <= n_scatter,
                   "near range outside scatter_index (target " + std::to_string(i) + ")");
  }
  for (int64_t q = 0; q < n_scatter; q++)
    SCTL_CHECK_ARG(scatter_index[q] >= 0 && scatter_index[q] < n_near,
                   "scatter_index out of range (entry " + std::to_string(q) + ")");
  Staged sc, sd, ss, sun, su;
  SCTL_TRY(sc.in(near_trg_cnt, (int64_t)sizeof(int64_t) * ntrg, s));
  SCTL_TRY(sd.in(near_trg_dsp, (int64_t)sizeof(int64_t) * ntrg, s));
  SCTL_TRY(ss.in(scatter_index, (int64_t)sizeof(int64_t) * n_scatter, s));
  SCTL_TRY(sun.in(U_near, (int64_t)sizeof(double) * n_near * dof, s));
  SCTL_TRY(su.in(U, (int64_t)sizeof(double) * ntrg * dof, s));
  SCTL_TRY(launch(sc.dev<int64_t>(), sd.dev<int64_t>(), ss.dev<int64_t>(), sun.dev<double>(), su.dev<double>()));
  SCTL_TRY(su.out(U, s));
  SCTL_HIP_TRY(hipStreamSynchronize(s));
  return SCTL_AMD_SUCCESS;
}

}  // extern "C"
