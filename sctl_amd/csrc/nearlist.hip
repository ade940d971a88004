// Near-interaction lists of BoundaryIntegralOp on the device: the
// single-process output of BuildNearList (boundary_integral.txx:50-520).
//
// Definition (the reference's, :390-401): target t is near element e when
// some node s of e has |x_t - x_s|^2 < r_s^2.  The near pairs are listed by
// element, then target index (:418-428 sort by (elem_idx, idx)); per pair the
// target's coordinates (and normal) are gathered (:430-447); near_elem_cnt /
// near_elem_dsp count pairs per element (:449-476); near_scatter_index orders
// the pairs by target (SortScatterIndex, :483, comm.txx:1393) and
// near_trg_cnt / near_trg_dsp count them per target, with dsp 0 for a target
// without pairs (:485-514).  The reference sorts the pairs of one target by
// key only with std::sort, which leaves their order unspecified; here they
// follow element order.
//
// The reference finds candidates through a Morton tree of the nodes.  Here:
//   1. targets get 63-bit Morton keys over the reference's bounding cube
//      (all points, grown by 5 % per side, :104-154) and are radix sorted;
//   2. one wave per element (its nodes in the lanes) takes the bounding box
//      of its nodes grown by the largest radius, picks the finest Morton
//      depth at which that box spans at most 64 cells, finds each cell's
//      targets by binary search in the sorted keys, queues the candidates
//      inside the grown box and tests them 64 at a time exactly against the
//      element's nodes (one ballot per candidate and 64 nodes, each node
//      chunk loaded once per batch, first hit wins);
//   3. a first pass counts the candidates inside the grown boxes only (an
//      upper bound of the pairs, without the node tests), the second tests
//      them and appends the (element, target) keys of the near ones through
//      an atomic counter; the keys are radix sorted (so the list does not
//      depend on the order of the appends), each element's range is found
//      by binary search, and a second sort by (target, pair) gives the
//      scatter index.
// Each element is one workgroup of 4 waves that take turns over its
// 64-candidate groups; elements are launched largest first.
// Every target within a node's radius lies in the grown box, hence in one of
// the cells (the cell range is computed with the same quantisation as the
// target keys), so the list is exact.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>
#include <rocprim/device/device_scan.hpp>

#include <algorithm>
#include <string>

#include "internal.hpp"

struct sctl_amd_near_list_s {
  int device = 0;
  int64_t ntrg = 0, nelem = 0, n_near = 0;
  bool has_normal = false;
  double *d_xtn = nullptr, *d_xntn = nullptr;
  int64_t *d_ecnt = nullptr, *d_edsp = nullptr, *d_scatter = nullptr, *d_tcnt = nullptr, *d_tdsp = nullptr;
  ~sctl_amd_near_list_s() {
    int cur = -1;
    (void)hipGetDevice(&cur);
    if (cur != device) (void)hipSetDevice(device);
    void* ptrs[] = {d_xtn, d_xntn, d_ecnt, d_edsp, d_scatter, d_tcnt, d_tdsp};
    for (void* p : ptrs)
      if (p) (void)hipFree(p);
    if (cur >= 0 && cur != device) (void)hipSetDevice(cur);
  }
};

namespace sctl_amd {
namespace {

constexpr int NL_WG = 256;
constexpr int MAX_DEPTH = 21;    // 21 bits per axis, 63-bit keys
constexpr int MAX_CELLS = 64;    // cells per element at the chosen depth

__device__ __forceinline__ uint64_t spread21(uint64_t v) {  // 21 bits -> every third bit
  v &= 0x1fffffULL;
  v = (v | (v << 32)) & 0x001f00000000ffffULL;
  v = (v | (v << 16)) & 0x001f0000ff0000ffULL;
  v = (v | (v << 8)) & 0x100f00f00f00f00fULL;
  v = (v | (v << 4)) & 0x10c30c30c30c30c3ULL;
  v = (v | (v << 2)) & 0x1249249249249249ULL;
  return v;
}
__device__ __forceinline__ uint64_t interleave(uint64_t x, uint64_t y, uint64_t z) {
  return spread21(x) | (spread21(y) << 1) | (spread21(z) << 2);
}

// The reference's bounding cube (:104-154), computed once so that the key
// and the search kernels quantise with the same bits: cube = {origin x, y,
// z, 2^21 / edge length}.
__global__ void k_nl_cube(const double* box, double* cube) {  // in place: cube may alias box
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double bl = 0;
  for (int k = 0; k < 3; k++) bl = fmax(bl, box[3 + k] - box[k]);
  const double L = bl > 0 ? bl : 1.0;
  for (int k = 0; k < 3; k++) cube[k] = box[k] - 0.05 * L;
  cube[3] = 2097152.0 / (1.1 * L);
}
__device__ __forceinline__ uint32_t quant(double x, double x0, double scale) {
  double u = (x - x0) * scale;
  u = fmin(fmax(u, 0.0), 2097151.0);  // fmax returns 0 for a NaN u
  return (uint32_t)u;
}

__global__ void __launch_bounds__(NL_WG)
k_nl_keys(const double* __restrict__ xt, int64_t nt, const double* __restrict__ box, uint64_t* __restrict__ keys,
          uint32_t* __restrict__ idx) {
  const int64_t i = (int64_t)blockIdx.x * NL_WG + threadIdx.x;
  if (i >= nt) return;
  const double scale = box[3];
  keys[i] = interleave(quant(xt[i * 3 + 0], box[0], scale), quant(xt[i * 3 + 1], box[1], scale),
                       quant(xt[i * 3 + 2], box[2], scale));
  idx[i] = (uint32_t)i;
}

__device__ __forceinline__ int64_t lower_bound_u64(const uint64_t* __restrict__ a, int64_t n, uint64_t key) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) >> 1;
    if (a[mid] < key) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = fmin(v, __shfl_xor(v, m, 64));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) v = fmax(v, __shfl_xor(v, m, 64));
  return v;
}


// |d|^2 rounded as node_dist2 sums it (boundary_integral.txx:72-79), without
// fused multiply-adds, so the radius test matches the reference's bit for bit.
__device__ __forceinline__ double dist2(double dx, double dy, double dz) {
#pragma clang fp contract(off)
  return (dx * dx + dy * dy) + dz * dz;
}

constexpr int NODE_REGS = 4;  // node chunks of 64 kept in registers (larger elements load the rest per batch)
constexpr int QCAP = 128;     // per-wave queue of candidates inside the grown box
constexpr int NWAVE = NL_WG / 64;  // waves sharing one element's candidates

struct Cand {
  double x, y, z;
  uint64_t t;
};

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One workgroup per element.  FILL = false: counts[e * NWAVE + w] = the
// candidates inside the element's grown box that wave w fetched (no node
// tests); FILL = true: appends (e << 32 | t) of each near target at
// out[atomicAdd(counter)], cap entries at most.
// Lanes hold the element's nodes (64 per chunk, the first NODE_REGS chunks in
// registers).  Candidate targets are fetched 64 at a time; those inside the
// grown box are appended to a per-wave LDS queue, and every 64 queued
// candidates are tested as a batch: node chunk by node chunk (each chunk
// loaded once per batch), each undecided candidate broadcast from LDS and
// tested against the chunk's 64 nodes with one ballot, decided at its first
// hit.
template <bool FILL>
__global__ void __launch_bounds__(NL_WG)
k_nl_elements(const uint32_t* __restrict__ eorder, const int64_t* __restrict__ ecnt,
              const int64_t* __restrict__ edsp, int64_t nsrc, const double* __restrict__ xs,
              const double* __restrict__ rad, int64_t nt, const double* __restrict__ xt,
              const uint64_t* __restrict__ skeys, const uint32_t* __restrict__ sidx, const double* __restrict__ box,
              int64_t* __restrict__ counts, unsigned long long* __restrict__ counter, int64_t cap,
              uint64_t* __restrict__ out) {
  __shared__ Cand queue_all[NWAVE][QCAP];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  Cand* queue = queue_all[w];
  const double x0[3] = {box[0], box[1], box[2]}, scale = box[3];
  {
    const int64_t e = eorder[blockIdx.x];
    const int64_t s0 = edsp[e], sc = ecnt[e];
    int64_t count = 0;
    if (s0 >= 0 && sc > 0 && s0 <= nsrc && sc <= nsrc - s0 && nt > 0) {
      const int nch = (int)((sc + 63) / 64);
      // nodes of the first chunks in registers; r2 = -1 marks a lane without a node
      double nx[NODE_REGS], ny[NODE_REGS], nz[NODE_REGS], nr2[NODE_REGS];
      double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, rmax = 0;
#pragma unroll
      for (int c = 0; c < NODE_REGS; c++) {
        nx[c] = ny[c] = nz[c] = 0.0;
        nr2[c] = -1.0;
      }
      for (int64_t s = lane; s < sc; s += 64) {
        const double x = xs[(s0 + s) * 3 + 0], y = xs[(s0 + s) * 3 + 1], z = xs[(s0 + s) * 3 + 2];
        const double r = rad[s0 + s];
        lo[0] = fmin(lo[0], x);
        lo[1] = fmin(lo[1], y);
        lo[2] = fmin(lo[2], z);
        hi[0] = fmax(hi[0], x);
        hi[1] = fmax(hi[1], y);
        hi[2] = fmax(hi[2], z);
        rmax = fmax(rmax, fabs(r));
        const int c = (int)(s >> 6);
#pragma unroll
        for (int k = 0; k < NODE_REGS; k++)
          if (c == k) {
            nx[k] = x;
            ny[k] = y;
            nz[k] = z;
            nr2[k] = r * r;
          }
      }
#pragma unroll
      for (int k = 0; k < 3; k++) {
        lo[k] = wave_min(lo[k]);
        hi[k] = wave_max(hi[k]);
      }
      rmax = wave_max(rmax);

      // Tests queue entries [0, n) (n <= 64); returns the mask of near ones
      // and, with FILL, writes them.
      auto batch = [&](int n) {
        uint64_t rem = n >= 64 ? ~0ULL : ((1ULL << n) - 1);
        uint64_t nearm = 0;
        for (int k = 0; k < nch && rem; k++) {
          double bx = 0, by = 0, bz = 0, br2 = -1.0;
          if (k < NODE_REGS) {
#pragma unroll
            for (int kk = 0; kk < NODE_REGS; kk++)
              if (kk == k) {
                bx = nx[kk];
                by = ny[kk];
                bz = nz[kk];
                br2 = nr2[kk];
              }
          } else {
            const int64_t s = (int64_t)k * 64 + lane;
            if (s < sc) {
              bx = xs[(s0 + s) * 3 + 0];
              by = xs[(s0 + s) * 3 + 1];
              bz = xs[(s0 + s) * 3 + 2];
              const double r = rad[s0 + s];
              br2 = r * r;
            }
          }
          uint64_t todo = rem;
          while (todo) {
            const int j = __builtin_ctzll(todo);
            todo &= todo - 1;
            const double cx = queue[j].x, cy = queue[j].y, cz = queue[j].z;
            if (__builtin_amdgcn_ballot_w64(dist2(bx - cx, by - cy, bz - cz) < br2) != 0) {
              nearm |= 1ULL << j;
              rem &= ~(1ULL << j);
            }
          }
        }
        if constexpr (FILL) {
          const int nn = __builtin_popcountll(nearm);
          if (nn) {
            unsigned long long base = 0;
            if (lane == 0) base = atomicAdd(counter, (unsigned long long)nn);
            base = __shfl(base, 0, 64);
            if ((nearm >> lane) & 1ULL) {
              const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(nearm >> 32),
                                                           __builtin_amdgcn_mbcnt_lo((uint32_t)nearm, 0));
              const int64_t pos = (int64_t)base + below;
              if (pos < cap) out[pos] = ((uint64_t)e << 32) | queue[lane].t;
            }
          }
        }
      };

      if (rmax > 0) {
        // Grown box, with a margin above the rounding of the two additions.
        const double grow = rmax * (1.0 + 1e-9);
        double glo[3], ghi[3];
        uint32_t qlo[3], qhi[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
          glo[k] = lo[k] - grow - 1e-14 * fabs(lo[k]);
          ghi[k] = hi[k] + grow + 1e-14 * fabs(hi[k]);
          qlo[k] = quant(glo[k], x0[k], scale);
          qhi[k] = quant(ghi[k], x0[k], scale);
        }
        int shift = 0;
        for (int d = MAX_DEPTH; d >= 0; d--) {
          shift = MAX_DEPTH - d;
          int64_t cells = 1;
#pragma unroll
          for (int k = 0; k < 3; k++) cells *= (int64_t)(qhi[k] >> shift) - (qlo[k] >> shift) + 1;
          if (cells <= MAX_CELLS) break;
        }
        const uint64_t span = 1ULL << (3 * shift);
        int qn = 0;        // queued candidates (< 64 between batches)
        int64_t group = 0; // 64-candidate groups fetched so far; wave w takes every NWAVE-th
        for (uint32_t bz = qlo[2] >> shift; bz <= (qhi[2] >> shift); bz++)
          for (uint32_t by = qlo[1] >> shift; by <= (qhi[1] >> shift); by++)
            for (uint32_t bx = qlo[0] >> shift; bx <= (qhi[0] >> shift); bx++) {
              const uint64_t klo = interleave(bx, by, bz) << (3 * shift);
              const int64_t i0 = lower_bound_u64(skeys, nt, klo);
              const int64_t i1 = lower_bound_u64(skeys, nt, klo + span);
              for (int64_t c0 = i0; c0 < i1; c0 += 64) {
                if ((group++ % NWAVE) != w) continue;
                const int64_t c = c0 + lane;
                bool inside = false;
                double mx = 0, my = 0, mz = 0;
                uint64_t mt = 0;
                if (c < i1) {
                  mt = sidx[c];
                  mx = xt[mt * 3 + 0];
                  my = xt[mt * 3 + 1];
                  mz = xt[mt * 3 + 2];
                  inside = mx >= glo[0] && mx <= ghi[0] && my >= glo[1] && my <= ghi[1] && mz >= glo[2] &&
                           mz <= ghi[2];
                }
                const uint64_t in = __builtin_amdgcn_ballot_w64(inside);
                if constexpr (!FILL) {
                  count += __builtin_popcountll(in);
                  continue;
                }
                if (inside) {
                  const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(in >> 32),
                                                               __builtin_amdgcn_mbcnt_lo((uint32_t)in, 0));
                  Cand& q = queue[qn + below];  // qn < 64, so qn + below < QCAP
                  q.x = mx;
                  q.y = my;
                  q.z = mz;
                  q.t = mt;
                }
                qn += __builtin_popcountll(in);
                wave_sync();
                if (qn >= 64) {
                  batch(64);
                  wave_sync();
                  const int from = 64 + (lane < qn - 64 ? lane : 0);
                  const double rx = queue[from].x, ry = queue[from].y, rz = queue[from].z;
                  const uint64_t rt = queue[from].t;
                  wave_sync();
                  if (lane < qn - 64) {
                    queue[lane].x = rx;
                    queue[lane].y = ry;
                    queue[lane].z = rz;
                    queue[lane].t = rt;
                  }
                  qn -= 64;
                  wave_sync();
                }
              }
            }
        if (qn > 0) batch(qn);
        wave_sync();
      }
    }
    if constexpr (!FILL) {
      if (lane == 0) counts[e * NWAVE + w] = count;
    }
  }
}

// near_elem_dsp[e] / near_elem_cnt[e]: the range of element e in the sorted
// keys (e << 32 | t), found by binary search (dsp = the pairs of the
// elements before e, also for an element without pairs).
__global__ void __launch_bounds__(NL_WG)
k_nl_elem_ranges(int64_t nelem, const uint64_t* __restrict__ keys, int64_t n, int64_t* __restrict__ ecnt_out,
                 int64_t* __restrict__ edsp_out) {
  const int64_t e = (int64_t)blockIdx.x * NL_WG + threadIdx.x;
  if (e >= nelem) return;
  const int64_t lo = lower_bound_u64(keys, n, (uint64_t)e << 32);
  const int64_t hi = lower_bound_u64(keys, n, (uint64_t)(e + 1) << 32);
  ecnt_out[e] = hi - lo;
  edsp_out[e] = lo;
}

// Elements ordered by node count, largest first (the order only schedules).
__global__ void __launch_bounds__(NL_WG)
k_nl_order_keys(int64_t nelem, const int64_t* __restrict__ ecnt, uint32_t* __restrict__ keys,
                uint32_t* __restrict__ idx) {
  const int64_t e = (int64_t)blockIdx.x * NL_WG + threadIdx.x;
  if (e >= nelem) return;
  const int64_t c = ecnt[e];
  keys[e] = 0xffffffffu - (uint32_t)(c < 0 ? 0 : (c > 0xffffffffLL ? 0xffffffffLL : c));
  idx[e] = (uint32_t)e;
}

__global__ void __launch_bounds__(NL_WG) k_nl_zero(int64_t* __restrict__ a, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * NL_WG + threadIdx.x;
  if (i < n) a[i] = 0;
}

// Pair i (element-major): gather the target's coordinates / normal, count it
// for its target, and form the (target, pair) key of the target-major sort.
__global__ void __launch_bounds__(NL_WG)
k_nl_unpack(int64_t n, const uint64_t* __restrict__ pairs, const double* __restrict__ xt,
            const double* __restrict__ xn, double* __restrict__ xtn, double* __restrict__ xntn,
            unsigned long long* __restrict__ tcnt, uint64_t* __restrict__ key2) {
  const int64_t i = (int64_t)blockIdx.x * NL_WG + threadIdx.x;
  if (i >= n) return;
  const uint64_t t = pairs[i] & 0xffffffffULL;
#pragma unroll
  for (int k = 0; k < 3; k++) xtn[i * 3 + k] = xt[t * 3 + k];
  if (xn) {
#pragma unroll
    for (int k = 0; k < 3; k++) xntn[i * 3 + k] = xn[t * 3 + k];
  }
  atomicAdd(&tcnt[t], 1ULL);
  key2[i] = (t << 32) | (uint64_t)i;
}

// Target-major position q: scatter_index[q] = the pair; near_trg_dsp of a
// target = its first position.
__global__ void __launch_bounds__(NL_WG)
k_nl_scatter(int64_t n, const uint64_t* __restrict__ key2s, int64_t* __restrict__ scatter, int64_t* __restrict__ tdsp) {
  const int64_t q = (int64_t)blockIdx.x * NL_WG + threadIdx.x;
  if (q >= n) return;
  const uint64_t k = key2s[q];
  scatter[q] = (int64_t)(k & 0xffffffffULL);
  const uint64_t t = k >> 32;
  if (q == 0 || (key2s[q - 1] >> 32) != t) tdsp[t] = q;
}

int grid_of(int64_t n) { return (int)((n + NL_WG - 1) / NL_WG); }

template <class T> int dev_alloc(T** p, int64_t n) {
  *p = nullptr;
  if (n <= 0) return SCTL_AMD_SUCCESS;
  SCTL_HIP_TRY(hipMalloc(p, sizeof(T) * n));
  return SCTL_AMD_SUCCESS;
}

int bits_for(int64_t n) {
  int b = 0;
  while (b < 63 && ((int64_t)1 << b) < n) b++;
  return b;
}

// The build on device arrays.
int build_list(sctl_amd_near_list_s* L, const double* xt, int64_t nt, const double* xn, const double* xs, int64_t ns,
               const double* rad, int64_t nelem, const int64_t* ecnt, const int64_t* edsp, hipStream_t s) {
  SCTL_TRY(dev_alloc(&L->d_ecnt, nelem));
  SCTL_TRY(dev_alloc(&L->d_edsp, nelem));
  SCTL_TRY(dev_alloc(&L->d_tcnt, nt));
  SCTL_TRY(dev_alloc(&L->d_tdsp, nt));
  if (nt > 0) {
    hipLaunchKernelGGL(k_nl_zero, dim3(grid_of(nt)), dim3(NL_WG), 0, s, L->d_tcnt, nt);
    hipLaunchKernelGGL(k_nl_zero, dim3(grid_of(nt)), dim3(NL_WG), 0, s, L->d_tdsp, nt);
    SCTL_HIP_TRY(hipGetLastError());
  }
  if (nelem > 0) {
    hipLaunchKernelGGL(k_nl_zero, dim3(grid_of(nelem)), dim3(NL_WG), 0, s, L->d_ecnt, nelem);
    hipLaunchKernelGGL(k_nl_zero, dim3(grid_of(nelem)), dim3(NL_WG), 0, s, L->d_edsp, nelem);
    SCTL_HIP_TRY(hipGetLastError());
  }
  if (nt == 0 || ns == 0 || nelem == 0) return SCTL_AMD_SUCCESS;

  Scratch box, keys_in, keys, idx_in, idx, tmp;
  SCTL_TRY(box.alloc(sizeof(double) * 6, s));
  SCTL_TRY(bounding_box(xt, nt, xs, ns, box.as<double>(), s));
  hipLaunchKernelGGL(k_nl_cube, dim3(1), dim3(64), 0, s, box.as<double>(), box.as<double>());
  SCTL_HIP_TRY(hipGetLastError());
  SCTL_TRY(keys_in.alloc(sizeof(uint64_t) * nt, s));
  SCTL_TRY(keys.alloc(sizeof(uint64_t) * nt, s));
  SCTL_TRY(idx_in.alloc(sizeof(uint32_t) * nt, s));
  SCTL_TRY(idx.alloc(sizeof(uint32_t) * nt, s));
  hipLaunchKernelGGL(k_nl_keys, dim3(grid_of(nt)), dim3(NL_WG), 0, s, xt, nt, box.as<double>(), keys_in.as<uint64_t>(),
                     idx_in.as<uint32_t>());
  SCTL_HIP_TRY(hipGetLastError());
  size_t bytes = 0;
  SCTL_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, keys_in.as<uint64_t>(), keys.as<uint64_t>(),
                                         idx_in.as<uint32_t>(), idx.as<uint32_t>(), (size_t)nt, 0, 63, s));
  SCTL_TRY(tmp.alloc(bytes, s));
  SCTL_HIP_TRY(rocprim::radix_sort_pairs(tmp.as<void>(), bytes, keys_in.as<uint64_t>(), keys.as<uint64_t>(),
                                         idx_in.as<uint32_t>(), idx.as<uint32_t>(), (size_t)nt, 0, 63, s));
  tmp.release();
  keys_in.release();
  idx_in.release();

  // elements largest first (scheduling only)
  Scratch okeys_in, okeys, oidx_in, oidx;
  SCTL_TRY(okeys_in.alloc(sizeof(uint32_t) * nelem, s));
  SCTL_TRY(okeys.alloc(sizeof(uint32_t) * nelem, s));
  SCTL_TRY(oidx_in.alloc(sizeof(uint32_t) * nelem, s));
  SCTL_TRY(oidx.alloc(sizeof(uint32_t) * nelem, s));
  hipLaunchKernelGGL(k_nl_order_keys, dim3(grid_of(nelem)), dim3(NL_WG), 0, s, nelem, ecnt, okeys_in.as<uint32_t>(),
                     oidx_in.as<uint32_t>());
  SCTL_HIP_TRY(hipGetLastError());
  SCTL_HIP_TRY(rocprim::radix_sort_pairs(nullptr, bytes, okeys_in.as<uint32_t>(), okeys.as<uint32_t>(),
                                         oidx_in.as<uint32_t>(), oidx.as<uint32_t>(), (size_t)nelem, 0, 32, s));
  SCTL_TRY(tmp.alloc(bytes, s));
  SCTL_HIP_TRY(rocprim::radix_sort_pairs(tmp.as<void>(), bytes, okeys_in.as<uint32_t>(), okeys.as<uint32_t>(),
                                         oidx_in.as<uint32_t>(), oidx.as<uint32_t>(), (size_t)nelem, 0, 32, s));
  tmp.release();

  // Count pass: candidates inside the grown boxes, per (element, wave): an
  // upper bound of the near pairs.
  const int64_t nw = nelem * NWAVE;
  Scratch cnt_w, dsp_w;
  SCTL_TRY(cnt_w.alloc(sizeof(int64_t) * nw, s));
  SCTL_TRY(dsp_w.alloc(sizeof(int64_t) * nw, s));
  const dim3 egrid((unsigned)nelem);
  hipLaunchKernelGGL(k_nl_elements<false>, egrid, dim3(NL_WG), 0, s, oidx.as<uint32_t>(), ecnt, edsp, ns, xs, rad,
                     nt, xt, keys.as<uint64_t>(), idx.as<uint32_t>(), box.as<double>(), cnt_w.as<int64_t>(), nullptr,
                     (int64_t)0, nullptr);
  SCTL_HIP_TRY(hipGetLastError());
  SCTL_HIP_TRY(rocprim::exclusive_scan(nullptr, bytes, cnt_w.as<int64_t>(), dsp_w.as<int64_t>(), (int64_t)0,
                                       (size_t)nw, rocprim::plus<int64_t>(), s));
  SCTL_TRY(tmp.alloc(bytes, s));
  SCTL_HIP_TRY(rocprim::exclusive_scan(tmp.as<void>(), bytes, cnt_w.as<int64_t>(), dsp_w.as<int64_t>(), (int64_t)0,
                                       (size_t)nw, rocprim::plus<int64_t>(), s));
  tmp.release();
  int64_t last[2] = {0, 0};
  SCTL_HIP_TRY(hipMemcpyAsync(&last[0], cnt_w.as<int64_t>() + nw - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  SCTL_HIP_TRY(hipMemcpyAsync(&last[1], dsp_w.as<int64_t>() + nw - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  SCTL_HIP_TRY(hipStreamSynchronize(s));
  cnt_w.release();
  dsp_w.release();
  const int64_t cap = last[0] + last[1];

  // Fill pass: the node tests, near pairs appended in any order.
  int64_t n = 0;
  Scratch pairs_in, pairs, key2_in, key2, counter;
  if (cap > 0) {
    SCTL_TRY(pairs_in.alloc(sizeof(uint64_t) * cap, s));
    SCTL_TRY(counter.alloc(sizeof(int64_t), s));
    hipLaunchKernelGGL(k_nl_zero, dim3(1), dim3(NL_WG), 0, s, counter.as<int64_t>(), (int64_t)1);
    hipLaunchKernelGGL(k_nl_elements<true>, egrid, dim3(NL_WG), 0, s, oidx.as<uint32_t>(), ecnt, edsp, ns, xs, rad,
                       nt, xt, keys.as<uint64_t>(), idx.as<uint32_t>(), box.as<double>(), nullptr,
                       reinterpret_cast<unsigned long long*>(counter.as<int64_t>()), cap, pairs_in.as<uint64_t>());
    SCTL_HIP_TRY(hipGetLastError());
    SCTL_HIP_TRY(hipMemcpyAsync(&n, counter.as<int64_t>(), sizeof(int64_t), hipMemcpyDeviceToHost, s));
    SCTL_HIP_TRY(hipStreamSynchronize(s));
  }
  SCTL_CHECK_ARG(n <= cap, "near-pair count exceeds its bound (internal error)");
  keys.release();
  idx.release();
  L->n_near = n;
  if (n == 0) return SCTL_AMD_SUCCESS;  // d_ecnt, d_edsp were zeroed
  SCTL_CHECK_ARG(n < ((int64_t)1 << 32), "more than 2^32-1 near pairs");

  SCTL_TRY(pairs.alloc(sizeof(uint64_t) * n, s));
  const int ebits = 32 + bits_for(nelem);
  SCTL_HIP_TRY(rocprim::radix_sort_keys(nullptr, bytes, pairs_in.as<uint64_t>(), pairs.as<uint64_t>(), (size_t)n, 0,
                                        ebits, s));
  SCTL_TRY(tmp.alloc(bytes, s));
  SCTL_HIP_TRY(rocprim::radix_sort_keys(tmp.as<void>(), bytes, pairs_in.as<uint64_t>(), pairs.as<uint64_t>(),
                                        (size_t)n, 0, ebits, s));
  tmp.release();
  pairs_in.release();
  hipLaunchKernelGGL(k_nl_elem_ranges, dim3(grid_of(nelem)), dim3(NL_WG), 0, s, nelem, pairs.as<uint64_t>(), n,
                     L->d_ecnt, L->d_edsp);
  SCTL_HIP_TRY(hipGetLastError());

  SCTL_TRY(dev_alloc(&L->d_xtn, n * 3));
  if (xn) SCTL_TRY(dev_alloc(&L->d_xntn, n * 3));
  SCTL_TRY(dev_alloc(&L->d_scatter, n));
  SCTL_TRY(key2_in.alloc(sizeof(uint64_t) * n, s));
  SCTL_TRY(key2.alloc(sizeof(uint64_t) * n, s));
  hipLaunchKernelGGL(k_nl_unpack, dim3(grid_of(n)), dim3(NL_WG), 0, s, n, pairs.as<uint64_t>(), xt, xn, L->d_xtn,
                     L->d_xntn, reinterpret_cast<unsigned long long*>(L->d_tcnt), key2_in.as<uint64_t>());
  SCTL_HIP_TRY(hipGetLastError());
  const int tbits = 32 + bits_for(nt);
  SCTL_HIP_TRY(rocprim::radix_sort_keys(nullptr, bytes, key2_in.as<uint64_t>(), key2.as<uint64_t>(), (size_t)n, 0,
                                        tbits, s));
  SCTL_TRY(tmp.alloc(bytes, s));
  SCTL_HIP_TRY(rocprim::radix_sort_keys(tmp.as<void>(), bytes, key2_in.as<uint64_t>(), key2.as<uint64_t>(),
                                        (size_t)n, 0, tbits, s));
  hipLaunchKernelGGL(k_nl_scatter, dim3(grid_of(n)), dim3(NL_WG), 0, s, n, key2.as<uint64_t>(), L->d_scatter,
                     L->d_tdsp);
  SCTL_HIP_TRY(hipGetLastError());
  SCTL_HIP_TRY(hipStreamSynchronize(s));
  return SCTL_AMD_SUCCESS;
}

}  // namespace
}  // namespace sctl_amd

using namespace sctl_amd;

extern "C" {

int sctl_amd_near_list_build(sctl_amd_near_list* list, const double* Xtrg, int64_t ntrg, const double* Xn_trg,
                             const double* Xsrc, int64_t nsrc, const double* src_radius, int64_t nelem,
                             const int64_t* elem_nds_cnt, const int64_t* elem_nds_dsp, uint32_t flags,
                             void* stream) {
  SCTL_CHECK_ARG(list, "list is NULL");
  *list = nullptr;
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(ntrg >= 0 && nsrc >= 0 && nelem >= 0, "negative size");
  SCTL_CHECK_ARG(ntrg < ((int64_t)1 << 32) && nelem < ((int64_t)1 << 31), "too many targets or elements");
  SCTL_CHECK_ARG(ntrg == 0 || Xtrg, "Xtrg is NULL");
  SCTL_CHECK_ARG(nsrc == 0 || (Xsrc && src_radius), "Xsrc or src_radius is NULL");
  SCTL_CHECK_ARG(nelem == 0 || (elem_nds_cnt && elem_nds_dsp), "NULL element array");
  const bool host = flags & SCTL_AMD_HOST_PTRS;
  if (host) {
    for (int64_t e = 0; e < nelem; e++)
      SCTL_CHECK_ARG(elem_nds_cnt[e] >= 0 && elem_nds_dsp[e] >= 0 && elem_nds_dsp[e] + elem_nds_cnt[e] <= nsrc,
                     "element node range outside the sources (element " + std::to_string(e) + ")");
  }
  DeviceInfo dev;
  SCTL_TRY(device_info(&dev));
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  auto* L = new sctl_amd_near_list_s();
  L->device = dev.device;
  L->ntrg = ntrg;
  L->nelem = nelem;
  L->has_normal = Xn_trg != nullptr;
  int rc = SCTL_AMD_SUCCESS;
  if (!host) {
    rc = build_list(L, Xtrg, ntrg, Xn_trg, Xsrc, nsrc, src_radius, nelem, elem_nds_cnt, elem_nds_dsp, s);
  } else {
    Scratch a, b, c, d, e, f;
    auto up = [&](Scratch& buf, const void* src, size_t bytes) -> int {
      if (!src || !bytes) return SCTL_AMD_SUCCESS;
      SCTL_TRY(buf.alloc(bytes, s));
      SCTL_HIP_TRY(hipMemcpyAsync(buf.as<void>(), src, bytes, hipMemcpyHostToDevice, s));
      return SCTL_AMD_SUCCESS;
    };
    if (!rc) rc = up(a, Xtrg, sizeof(double) * 3 * ntrg);
    if (!rc) rc = up(b, Xn_trg, sizeof(double) * 3 * ntrg);
    if (!rc) rc = up(c, Xsrc, sizeof(double) * 3 * nsrc);
    if (!rc) rc = up(d, src_radius, sizeof(double) * nsrc);
    if (!rc) rc = up(e, elem_nds_cnt, sizeof(int64_t) * nelem);
    if (!rc) rc = up(f, elem_nds_dsp, sizeof(int64_t) * nelem);
    if (!rc)
      rc = build_list(L, a.as<double>(), ntrg, Xn_trg ? b.as<double>() : nullptr, c.as<double>(), nsrc,
                      d.as<double>(), nelem, e.as<int64_t>(), f.as<int64_t>(), s);
    if (!rc && hipStreamSynchronize(s) != hipSuccess) rc = SCTL_AMD_ERROR_HIP;
  }
  if (rc) {
    delete L;
    return rc;
  }
  *list = L;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_near_list_size(sctl_amd_near_list list, int64_t* n_near) {
  SCTL_CHECK_ARG(list && n_near, "NULL argument");
  *n_near = list->n_near;
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_near_list_get(sctl_amd_near_list list, double* Xtrg_near, double* Xn_trg_near, int64_t* near_elem_cnt,
                           int64_t* near_elem_dsp, int64_t* near_scatter_index, int64_t* near_trg_cnt,
                           int64_t* near_trg_dsp, uint32_t flags, void* stream) {
  SCTL_CHECK_ARG(list, "list is NULL");
  SCTL_CHECK_ARG((flags & ~SCTL_AMD_HOST_PTRS) == 0, "unknown flag bits");
  SCTL_CHECK_ARG(!Xn_trg_near || list->has_normal || list->n_near == 0, "the list was built without target normals");
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : hipStreamPerThread;
  const hipMemcpyKind kind = (flags & SCTL_AMD_HOST_PTRS) ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
  auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
    if (!dst || !src || !bytes) return SCTL_AMD_SUCCESS;
    SCTL_HIP_TRY(hipMemcpyAsync(dst, src, bytes, kind, s));
    return SCTL_AMD_SUCCESS;
  };
  const int64_t n = list->n_near;
  SCTL_TRY(cp(Xtrg_near, list->d_xtn, sizeof(double) * 3 * n));
  SCTL_TRY(cp(Xn_trg_near, list->d_xntn, sizeof(double) * 3 * n));
  SCTL_TRY(cp(near_elem_cnt, list->d_ecnt, sizeof(int64_t) * list->nelem));
  SCTL_TRY(cp(near_elem_dsp, list->d_edsp, sizeof(int64_t) * list->nelem));
  SCTL_TRY(cp(near_scatter_index, list->d_scatter, sizeof(int64_t) * n));
  SCTL_TRY(cp(near_trg_cnt, list->d_tcnt, sizeof(int64_t) * list->ntrg));
  SCTL_TRY(cp(near_trg_dsp, list->d_tdsp, sizeof(int64_t) * list->ntrg));
  if (flags & SCTL_AMD_HOST_PTRS) SCTL_HIP_TRY(hipStreamSynchronize(s));
  return SCTL_AMD_SUCCESS;
}

int sctl_amd_near_list_destroy(sctl_amd_near_list list) {
  delete list;
  return SCTL_AMD_SUCCESS;
}

}  // extern "C"
