/*
 * ORACLE -- test infrastructure only.
 *
 * Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py
 * may load this library, and only as the checker.  The product (libsctl_amd)
 * never links, loads or calls it.
 *
 * A CPU restatement of BoundaryIntegralOp's near-field steps, on the
 * reference's own arrays (one process, one element list):
 *
 *   near-block partition      boundary_integral.txx:1009-1032
 *   K_near_cnt / K_near_dsp   boundary_integral.txx:1034-1057
 *   KNearSubtract             boundary_integral.txx:1165-1240, with the
 *                             default (identity) FarFieldDensityOperator-
 *                             Transpose, boundary_integral.txx:531-533
 *   ComputeNearInterac GEMV   boundary_integral.txx:1324-1340 through
 *                             Matrix::GEMM (matrix.txx:272-279,
 *                             mat_utils.txx:30-41)
 *   ScatterForward + sum      boundary_integral.txx:1368-1379,
 *                             comm.txx:1497-1508 (one process)
 *
 * KernelMatrix itself is oracle_kernel_matrix (direct_sum.c, exact mode).
 * The pins for these functions are in tests/pins.py (near_* checks).
 */
#include <stdlib.h>
#include <string.h>

#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

int oracle_kernel_matrix(int ker, const double* r_trg, long nt, const double* r_src, const double* n_src,
                         long ns, double* M, int nthreads);
int oracle_kernel_matrix_ctx(int ker, const double* r_trg, long nt, const double* r_src, const double* n_src,
                             long ns, double* M, int nthreads, const double* ctx);
int oracle_dims(int ker, int* dims);

static long min_l(long a, long b) { return a < b ? a : b; }

/* boundary_integral.txx:1009-1032.  Returns the number of blocks; fills the
 * three arrays when blk_elem is not NULL. */
long oracle_near_blocks(long nelem, const long* near_elem_cnt, const long* near_elem_dsp, int nthreads,
                        long* blk_elem, long* blk_t0, long* blk_cnt) {
  const long n_near = nelem ? near_elem_dsp[nelem - 1] + near_elem_cnt[nelem - 1] : 0;
  long grain = n_near / ((ORACLE_MUTANT == 10 ? 1L : 4L) * nthreads);
  if (grain < 1) grain = 1;
  long b = 0;
  for (long e = 0; e < nelem; e++) {
    for (long t0 = 0; t0 < near_elem_cnt[e]; t0 += grain, b++) {
      if (blk_elem) {
        blk_elem[b] = e;
        blk_t0[b] = t0;
        blk_cnt[b] = min_l(grain, near_elem_cnt[e] - t0);
      }
    }
  }
  return b;
}

/* boundary_integral.txx:1039-1053: K_near_cnt[e] = elem_nds_cnt[e] *
 * near_elem_cnt[e] (0 for a matrix-free element list), K_near_dsp its
 * exclusive scan.  Returns the length of K_near in doubles
 * ((K_near_dsp[last] + K_near_cnt[last]) * KDIM0 * KDIM1_). */
long oracle_near_layout(int ker, int normal_dot, long nelem, const long* near_elem_cnt, const long* elem_nds_cnt,
                        const int* matrix_free, long* K_near_cnt, long* K_near_dsp) {
  int d[5];
  if (oracle_dims(ker, d) != 0) return -1;
  const int K0 = d[1], K1 = d[2];
  if (normal_dot && K1 % 3) return -1;
  const int K1_ = normal_dot ? K1 / 3 : K1;
  long sum = 0;
  for (long e = 0; e < nelem; e++) {
    K_near_cnt[e] = (matrix_free && matrix_free[e]) ? 0 : elem_nds_cnt[e] * near_elem_cnt[e];
    K_near_dsp[e] = sum;
    sum += K_near_cnt[e];
  }
  return sum * K0 * K1_;
}

/* KNearSubtract, boundary_integral.txx:1165-1240. */
int oracle_near_subtract(int ker, int normal_dot, long nblk, const long* near_blk_elem, const long* near_blk_t0,
                         const long* near_blk_cnt, const long* near_elem_dsp, const long* elem_nds_cnt,
                         const long* elem_nds_cnt_far, const long* elem_nds_dsp_far, const long* K_near_cnt,
                         const long* K_near_dsp, const int* matrix_free, const double* Xtrg_near,
                         const double* Xn_trg_near, const double* X_far, const double* Xn_far, const double* wts_far,
                         double* K_near, const double* ctx) {
  int d[5];
  if (oracle_dims(ker, d) != 0) return -1;
  const int K0 = d[1], K1 = d[2], ND = d[3];
  if (normal_dot && K1 % 3) return -1;
  const int K1_ = normal_dot ? K1 / 3 : K1;
  for (long blk = 0; blk < nblk; blk++) {
    const long e = near_blk_elem[blk], t0 = near_blk_t0[blk];
    if (matrix_free && matrix_free[e]) continue; /* :1168 */
    const long trg_cnt = near_blk_cnt[blk];
    const long trg_dsp = near_elem_dsp[e] + t0;
    if (!trg_cnt) continue;
    const long far_src_cnt = elem_nds_cnt_far[e];
    const long far_src_dsp = elem_nds_dsp_far[e];
    const long src_dof = elem_nds_cnt[e] * K0;
    if (K_near_cnt[e] == 0 && src_dof) return -2;
    double* Kb = K_near + K_near_dsp[e] * K0 * K1_ + t0 * K1_ * src_dof; /* :1189 */
    const long nrow_m = far_src_cnt * K0;
    double* Mk = (double*)calloc((size_t)(nrow_m * trg_cnt * K1 + 1), sizeof(double));
    double* Mker = (double*)calloc((size_t)(nrow_m * trg_cnt * K1_ + 1), sizeof(double));
    if (!Mk || !Mker) {
      free(Mk);
      free(Mker);
      return -3;
    }
    if (oracle_kernel_matrix_ctx(ker, Xtrg_near + trg_dsp * 3, trg_cnt, X_far + far_src_dsp * 3,
                                 ND ? Xn_far + far_src_dsp * 3 : NULL, far_src_cnt, Mk, 1, ctx) != 0) {
      free(Mk);
      free(Mker);
      return -1;
    }
    const long ldk = trg_cnt * K1, ldm = trg_cnt * K1_;
    for (long s = 0; s < far_src_cnt; s++) {
      const double w = ORACLE_MUTANT == 8 ? 1.0 : wts_far[far_src_dsp + s];
      for (long k0 = 0; k0 < K0; k0++) {
        const long r = s * K0 + k0;
        if (normal_dot) { /* :1193-1209 */
          for (long t = 0; t < trg_cnt; t++)
            for (long k1 = 0; k1 < K1_; k1++) {
              double acc = 0;
              for (long l = 0; l < 3; l++) {
                const double nl = ORACLE_MUTANT == 7 ? Xn_far[(far_src_dsp + s) * 3 + l]
                                                     : Xn_trg_near[(trg_dsp + t) * 3 + l];
                acc += Mk[r * ldk + (t * K1_ + k1) * 3 + l] * w * nl;
              }
              Mker[r * ldm + t * K1_ + k1] = acc;
            }
        } else { /* :1210-1219 */
          for (long t = 0; t < trg_cnt * K1; t++) Mker[r * ldm + t] = Mk[r * ldk + t] * w;
        }
      }
    }
    /* :1220-1236, identity FarFieldDensityOperatorTranspose: Msub = Mker,
     * (src_dof x trg_cnt*KDIM1_), subtracted transposed. */
    const long nr = trg_cnt * K1_, nc = src_dof;
    if (nrow_m != nc) {
      free(Mk);
      free(Mker);
      return -4;
    }
    for (long i = 0; i < nr; i++)
      for (long j = 0; j < nc; j++) Kb[i * nc + j] -= Mker[j * ldm + i];
    free(Mk);
    free(Mker);
  }
  return 0;
}

/* ComputeNearInterac's matrix route, boundary_integral.txx:1324-1340:
 * U_near[(near_elem_dsp[e]+t0)*KDIM1_ + r] = sum_k K_[r][k] F_[k], summed in
 * k order as mat::gemm does (mat_utils.txx:30-41, alpha 1, beta 0). */
int oracle_near_gemv(int ker, int normal_dot, long nblk, const long* near_blk_elem, const long* near_blk_t0,
                     const long* near_blk_cnt, const long* near_elem_dsp, const long* elem_nds_cnt,
                     const long* elem_nds_dsp, const long* K_near_cnt, const long* K_near_dsp, const double* K_near,
                     const double* F, double* U_near) {
  int d[5];
  if (oracle_dims(ker, d) != 0) return -1;
  const int K0 = d[1], K1 = d[2];
  if (normal_dot && K1 % 3) return -1;
  const int K1_ = normal_dot ? K1 / 3 : K1;
  for (long blk = 0; blk < nblk; blk++) {
    const long e = near_blk_elem[blk], t0 = near_blk_t0[blk], nt = near_blk_cnt[blk];
    const long src_dof = elem_nds_cnt[e] * K0;
    const long trg_dof = nt * K1_;
    double* U_ = U_near + (near_elem_dsp[e] + t0) * K1_;
    if (src_dof == 0 || trg_dof == 0 || K_near_cnt[e] == 0) {
      for (long r = 0; r < trg_dof; r++) U_[r] = 0;
      continue;
    }
    const double* K_ = K_near + K_near_dsp[e] * K0 * K1_ + t0 * K1_ * src_dof;
    const double* F_ = F + elem_nds_dsp[e] * (ORACLE_MUTANT == 11 ? 1 : K0);
    for (long r = 0; r < trg_dof; r++) {
      double acc = 0;
      for (long k = 0; k < src_dof; k++) acc += K_[r * src_dof + k] * F_[k];
      U_[r] = 1.0 * acc + 0.0;
    }
  }
  return 0;
}

/* ScatterForward (comm.txx:1497-1508, one process) then the per-target sum
 * of boundary_integral.txx:1370-1379. */
int oracle_near_accumulate(long ntrg, int dof, const long* near_trg_cnt, const long* near_trg_dsp, long n_scatter,
                           const long* scatter_index, const double* U_near, double* U) {
  double* data = (double*)malloc(sizeof(double) * (size_t)(n_scatter * dof + 1));
  if (!data) return -3;
  for (long i = 0; i < n_scatter; i++)
    for (long j = 0; j < dof; j++) data[i * dof + j] = U_near[scatter_index[i] * dof + j];
  for (long i = 0; i < ntrg; i++) {
    const long cnt = near_trg_cnt[i], dsp = near_trg_dsp[i];
    for (long j = (ORACLE_MUTANT == 9 ? 1 : 0); j < cnt; j++)
      for (long k = 0; k < dof; k++) U[i * dof + k] += data[(dsp + j) * dof + k];
  }
  free(data);
  return 0;
}

/* BuildNearList (boundary_integral.txx:50-520) on one process, by brute
 * force over the output it defines: target t is near element e when some node
 * s of e has dist2(s, t) < rad_s^2 (:390-401, dist2 summed over the axes as
 * node_dist2, :72-79); pairs ordered by (element, target index) (:418-428);
 * Xtrg_near / Xn_trg_near gathered per pair (:430-447); near_elem_cnt and
 * its exclusive scan (:449-476); the pairs in target order (SortScatterIndex,
 * :483 -- ties in element order here: the reference's std::sort on the key
 * alone leaves them unspecified); near_trg_cnt, and near_trg_dsp = first
 * target-major position for targets with pairs, 0 otherwise (:485-514).
 * Returns the number of pairs; fills the outputs when max_near >= it. */
long oracle_near_list(const double* Xtrg, long ntrg, const double* Xn_trg, const double* Xsrc, long nsrc,
                      const double* rad, long nelem, const long* ecnt, const long* edsp, long max_near,
                      double* Xtrg_near, double* Xn_trg_near, long* near_elem_cnt, long* near_elem_dsp,
                      long* scatter, long* near_trg_cnt, long* near_trg_dsp) {
  (void)nsrc;
  long n = 0;
  long* pair_trg = NULL;
  for (int pass = 0; pass < 2; pass++) {
    n = 0;
    for (long e = 0; e < nelem; e++) {
      const long cnt0 = n;
      for (long t = 0; t < ntrg; t++) {
        int near = 0;
        for (long s = edsp[e]; s < edsp[e] + ecnt[e] && !near; s++) {
          double dist2 = 0;
          for (int k = 0; k < 3; k++) {
            const double dX = Xsrc[s * 3 + k] - Xtrg[t * 3 + k];
            dist2 += dX * dX;
          }
          near = ORACLE_MUTANT == 12 ? (dist2 <= rad[s] * rad[s]) : (dist2 < rad[s] * rad[s]);
        }
        if (near) {
          if (pass == 1) {
            for (int k = 0; k < 3; k++) {
              Xtrg_near[n * 3 + k] = Xtrg[t * 3 + k];
              if (Xn_trg && Xn_trg_near) Xn_trg_near[n * 3 + k] = Xn_trg[t * 3 + k];
            }
            pair_trg[n] = t;
          }
          n++;
        }
      }
      if (pass == 1) {
        near_elem_cnt[e] = n - cnt0;
        near_elem_dsp[e] = cnt0;
      }
    }
    if (pass == 0) {
      if (max_near < n) return n;
      pair_trg = (long*)malloc(sizeof(long) * (size_t)(n + 1));
      if (!pair_trg) return -3;
    }
  }
  /* target order, ties in pair (= element) order: a stable counting sort */
  for (long t = 0; t < ntrg; t++) near_trg_cnt[t] = 0;
  for (long i = 0; i < n; i++) near_trg_cnt[pair_trg[i]]++;
  long* next = (long*)malloc(sizeof(long) * (size_t)(ntrg + 1));
  if (!next) {
    free(pair_trg);
    return -3;
  }
  long run = 0;
  for (long t = 0; t < ntrg; t++) {
    next[t] = run;
    near_trg_dsp[t] = (ORACLE_MUTANT == 13 || near_trg_cnt[t]) ? run : 0;
    run += near_trg_cnt[t];
  }
  for (long i = 0; i < n; i++) scatter[next[pair_trg[i]]++] = i;
  free(next);
  free(pair_trg);
  return n;
}
