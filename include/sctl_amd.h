/*
 * sctl_amd.h -- C-ABI of the MI355X (gfx950) direct kernel summation.
 *
 * Drop-in device backend for SCTL's GenericKernel<uKernel>::Eval /
 * KernelMatrix (include/sctl/generic-kernel.hpp:99,112,124) for the sixteen
 * micro-kernels of include/sctl/kernel_functions.hpp (Laplace3D FxU, DxU,
 * FxdU, DxdU, Fxd2U; Stokes3D FxU, DxU, FxT, FSxU, FxUP; BiotSavart3D FxU,
 * FxdU; Helmholtz3D FxU, DxU, FxdU; HelmholtzDiff3D FxdU; aliases at
 * :669-684), and for the batches of near-field blocks that
 * BoundaryIntegralOp evaluates through them
 * (include/sctl/boundary_integral.txx:1009-1032,1165-1240,1311-1381).
 *
 * Conventions
 *  - Plain pointers and sizes only; every function returns an sctl_amd_status
 *    and never aborts.  sctl_amd_last_error() gives the message of the last
 *    failing call on the calling thread.
 *  - Coordinates, normals and densities are array-of-structures doubles exactly
 *    as in SCTL's Vector arguments: r[i*3+k], n[i*3+k], v[i*K0+k0],
 *    v_trg[t*K1+k1].
 *  - Array arguments are device pointers unless SCTL_AMD_HOST_PTRS is set in
 *    `flags`; then they are host pointers, the library stages them and the
 *    call returns after the result has been copied back.
 *  - `stream` is a hipStream_t (NULL: the calling thread's per-thread default
 *    stream).  Device-pointer calls are asynchronous on that stream.
 *  - Thread-safe: calls from different host threads on different streams may
 *    overlap.  Plans (sctl_amd_plan) are not shared between threads.
 */
#ifndef SCTL_AMD_H_
#define SCTL_AMD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  SCTL_AMD_SUCCESS = 0,
  SCTL_AMD_ERROR_INVALID_VALUE = 1,  /* bad size, pointer, flag, digits or layout */
  SCTL_AMD_ERROR_UNKNOWN_KERNEL = 2, /* kernel id outside the sixteen micro-kernels */
  SCTL_AMD_ERROR_UNSUPPORTED = 3,    /* precision/option combination not provided */
  SCTL_AMD_ERROR_NO_DEVICE = 4,      /* no HIP device visible */
  SCTL_AMD_ERROR_OUT_OF_MEMORY = 5,  /* device allocation failed */
  SCTL_AMD_ERROR_HIP = 6,            /* a HIP runtime call failed */
  SCTL_AMD_ERROR_QUEUE = 7           /* batched work queue did not drain as planned */
} sctl_amd_status;

/* Micro-kernel ids: the aliases of kernel_functions.hpp:669-684.  r = x_t - x_s;
 * K0 densities per source, K1 outputs per target.  The Helmholtz kernels
 * (ids 12..15) take the wavenumber mu through ctx (the *_ctx entry points,
 * sctl_amd_plan_set_ctx), as the reference reads it through ctx_ptr
 * (kernel_functions.hpp:485-489); their complex densities and outputs are
 * (re, im) pairs. */
enum {
  SCTL_AMD_LAPLACE3D_FXU = 0,    /* K = 1/r,                  scale  1/(4 pi), K0=1, K1=1, no normal */
  SCTL_AMD_LAPLACE3D_DXU = 1,    /* K = (r.n)/r^3,            scale  1/(4 pi), K0=1, K1=1, normal    */
  SCTL_AMD_STOKES3D_FXU = 2,     /* K = I/r + r r^T/r^3,      scale  1/(8 pi), K0=3, K1=3, no normal */
  SCTL_AMD_STOKES3D_DXU = 3,     /* K = r r^T (r.n)/r^5,      scale  3/(4 pi), K0=3, K1=3, normal    */
  SCTL_AMD_LAPLACE3D_FXDU = 4,   /* K = r^T/r^3,              scale -1/(4 pi), K0=1, K1=3, no normal */
  SCTL_AMD_LAPLACE3D_DXDU = 5,   /* K = n^T/r^3 - 3(r.n)r^T/r^5, scale 1/(4 pi), K0=1, K1=3, normal  */
  SCTL_AMD_LAPLACE3D_FXD2U = 6,  /* K = 3 r r^T/r^5 - I/r^3 (row-major 9), 1/(4 pi), K0=1, K1=9      */
  SCTL_AMD_STOKES3D_FXT = 7,     /* K[i][j*3+k] = r_i r_j r_k/r^5, scale -3/(4 pi), K0=3, K1=9      */
  SCTL_AMD_STOKES3D_FSXU = 8,    /* [f, q] -> f/r + r (r.f + q)/r^3, scale 1/(8 pi), K0=4, K1=3     */
  SCTL_AMD_STOKES3D_FXUP = 9,    /* f -> [Stokes FxU f, (r.f)/r^3], scale 1/(8 pi), K0=3, K1=4      */
  SCTL_AMD_BIOTSAVART3D_FXU = 10, /* f -> (f x r)/r^3,         scale  1/(4 pi), K0=3, K1=3           */
  SCTL_AMD_BIOTSAVART3D_FXDU = 11, /* gradient of BiotSavart FxU, scale -1/(4 pi), K0=3, K1=9         */
  SCTL_AMD_HELMHOLTZ3D_FXU = 12,  /* K = e^{i mu r}/r,           scale  1/(4 pi), K0=2, K1=2           */
  SCTL_AMD_HELMHOLTZ3D_DXU = 13,  /* K = (i mu r - 1) e^{i mu r} (r.n)/r^3, scale -1/(4 pi), K0=2, K1=2, normal */
  SCTL_AMD_HELMHOLTZ3D_FXDU = 14, /* K = (i mu r - 1) e^{i mu r} r/r^3, scale 1/(4 pi), K0=2, K1=6      */
  SCTL_AMD_HELMHOLTZDIFF3D_FXDU = 15 /* Helmholtz FxdU minus Laplace's, scale 1/(4 pi), K0=2, K1=6     */
};

/* Arithmetic of an evaluation. */
enum {
  SCTL_AMD_FP64 = 0,  /* FP64 throughout; digits selects the Newton steps after v_rsq_f64 */
  SCTL_AMD_MIXED = 1, /* FP64 tile-centred differencing, FP32 pairwise, Kahan FP64 accumulation */
  SCTL_AMD_AUTO = 2   /* MIXED when 0 <= digits <= 7, else FP64 */
};

/* flags */
#define SCTL_AMD_HOST_PTRS 0x1u     /* array arguments are host memory; call is synchronous */
#define SCTL_AMD_MIXED_NEWTON 0x2u  /* MIXED: one FP32 Newton step after v_rsq_f32 */

/* ------------------------------------------------------------------ setup */

/* Selects and initialises `device` for the calling thread (optional: every
 * entry point initialises the current device on first use). */
int sctl_amd_init(int device);

/* Message of the last failing call on the calling thread ("" if none). */
const char* sctl_amd_last_error(void);

/* Library version string. */
const char* sctl_amd_version(void);

/* Kernel metadata: CoordDim/SrcDim/TrgDim/NormalDim (generic-kernel.hpp:48-62),
 * the reference's FLOPS() Profile count and uKerScaleFactor<double>()
 * (kernel_functions.hpp:37-41,66-70,130-134,185-189 and the other kernels'
 * FLOPS()/uKerScaleFactor()), and Name().
 * Any output pointer may be NULL. */
int sctl_amd_kernel_info(int kernel, int* coord_dim, int* src_dim, int* trg_dim, int* normal_dim, int* flops,
                         double* scale, const char** name);

/* ------------------------------------------------------------------ Eval
 * Replaces GenericKernel<uKernel>::Eval<double,omp,digits>
 * (generic-kernel.hpp:112, generic-kernel.txx:102-211) and the type-erased
 * static Eval(..., digits, self) that ParticleFMM stores as a function pointer
 * (generic-kernel.hpp:99, generic-kernel.txx:72-100, fmm-wrapper.txx:153-154,618).
 *
 *   v_trg[t*K1+k1] += s * sum_{s} sum_{k0} K(r_trg[t]-r_src[s], n_src[s])[k0][k1] * v_src[s*K0+k0]
 *
 * v_trg holds nt*K1 values and is accumulated into (the caller resizes and
 * zeroes it first, as generic-kernel.txx:129-132 does on a size mismatch).
 * n_src may be NULL iff the kernel has no normal.  ns == 0 or nt == 0 leaves
 * v_trg unchanged.  digits: -1 (machine precision) or 0..15; >= 16 means -1.
 * Coincident points (r == 0) contribute exactly zero. */
int sctl_amd_eval(int kernel, int precision, int digits, const double* r_trg, int64_t nt, const double* r_src,
                  const double* n_src, const double* v_src, int64_t ns, double* v_trg, uint32_t flags, void* stream);

/* sctl_amd_eval with the kernel's context (GenericKernel::SetCtxPtr,
 * generic-kernel.txx:64-66): ctx points to the wavenumber (one double, read
 * at the call) for the Helmholtz kernels, which fail without it; the other
 * kernels ignore ctx.  sctl_amd_eval(...) is sctl_amd_eval_ctx(kernel, NULL, ...). */
int sctl_amd_eval_ctx(int kernel, const double* ctx, int precision, int digits, const double* r_trg, int64_t nt,
                      const double* r_src, const double* n_src, const double* v_src, int64_t ns, double* v_trg,
                      uint32_t flags, void* stream);

/* No counterpart in the reference: sctl_amd_eval_ctx for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * hosts that forbid hipMalloc).  Same arguments and results, device pointers
 * only (SCTL_AMD_HOST_PTRS is refused; flags: SCTL_AMD_MIXED_NEWTON).  It
 * allocates nothing: its scratch is one fixed device array per GPU of
 * sctl_amd_eval_noalloc_scratch_bytes() bytes (128 MiB), held by the library
 * until the process ends.  FP64 Evals of any size fit (their split partial
 * sums run in as many launches as the array needs); a mixed-precision Eval
 * whose Hilbert orders, packed source tiles and sort workspace exceed the
 * array (about 2^20 + 2^20 Laplace points is the largest that fits) fails
 * with SCTL_AMD_ERROR_INVALID_VALUE.  Every call on a GPU uses that GPU's one
 * array: calls that overlap on a GPU (other streams or host threads) race,
 * and the caller must serialise them. */
int sctl_amd_eval_noalloc(int kernel, const double* ctx, int precision, int digits, const double* r_trg, int64_t nt,
                          const double* r_src, const double* n_src, const double* v_src, int64_t ns, double* v_trg,
                          uint32_t flags, void* stream);
size_t sctl_amd_eval_noalloc_scratch_bytes(void);

/* ------------------------------------------------------------ KernelMatrix
 * Replaces GenericKernel<uKernel>::KernelMatrix<double,omp,digits>
 * (generic-kernel.hpp:124, generic-kernel.txx:213-329):
 *
 *   M[(s*K0+k0)*(nt*K1) + t*K1+k1] = s * K(r_trg[t]-r_src[s], n_src[s])[k0][k1]
 *
 * M is row-major (ns*K0) x (nt*K1) and is overwritten. FP64. */
int sctl_amd_kernel_matrix(int kernel, int digits, const double* r_trg, int64_t nt, const double* r_src,
                           const double* n_src, int64_t ns, double* M, uint32_t flags, void* stream);

/* sctl_amd_kernel_matrix with the kernel's context (as sctl_amd_eval_ctx). */
int sctl_amd_kernel_matrix_ctx(int kernel, const double* ctx, int digits, const double* r_trg, int64_t nt,
                               const double* r_src, const double* n_src, int64_t ns, double* M, uint32_t flags,
                               void* stream);

/* -------------------------------------------------------------- dense plans
 * One (target set, source set) Eval whose geometry stays on the device across
 * applications: the far field that BoundaryIntegralOp::ComputeFarField
 * (boundary_integral.txx:1248-1309) and ParticleFMM::EvalDirect
 * (fmm-wrapper.txx:551-623) sum once per GMRES iteration with new densities.
 * Creation copies r_trg (nt x 3), r_src and n_src (ns x 3; n_src NULL iff the
 * kernel has no normal) to plan-owned device memory (HOST_PTRS: from host
 * memory), resolves precision and launch shape as sctl_amd_eval does and, for
 * mixed precision, orders targets and sources once; ctx is the Helmholtz
 * wavenumber (read now).  flags: HOST_PTRS, MIXED_NEWTON. */
typedef struct sctl_amd_dense_plan_s* sctl_amd_dense_plan;

int sctl_amd_dense_plan_create(sctl_amd_dense_plan* plan, int kernel, int precision, int digits, uint32_t flags,
                               const double* r_trg, int64_t nt, const double* r_src, const double* n_src, int64_t ns,
                               const double* ctx, void* stream);

/* v_trg (nt*K1) += Eval(densities v_src (ns*K0)), the same values as
 * sctl_amd_eval with the plan's shape.  With device pointers it allocates
 * nothing and issues no memset (capturable in a hipGraph). */
int sctl_amd_dense_plan_apply(sctl_amd_dense_plan plan, const double* v_src, double* v_trg, uint32_t flags,
                              void* stream);

/* The resolved precision and launch shape (any output may be NULL). */
int sctl_amd_dense_plan_shape(sctl_amd_dense_plan plan, int* precision, int* targets_per_lane, int* wg_per_cu_cap);

int sctl_amd_dense_plan_destroy(sctl_amd_dense_plan plan);

/* ------------------------------------------------------------ batched blocks
 * A plan holds nblocks (target-set, source-set) blocks over one target array
 * and one source array -- the near-field blocks (elem, t0, cnt) that
 * BoundaryIntegralOp builds (boundary_integral.txx:1009-1032) and evaluates
 * block by block (:1169-1238 KernelMatrix, :1342-1364 matrix-free), and that
 * ParticleFMM::EvalDirect sums directly (fmm-wrapper.txx:551-623).
 *
 * Block b covers targets [trg_dsp[b], trg_dsp[b]+trg_cnt[b]) and sources
 * [src_dsp[b], src_dsp[b]+src_cnt[b]).  Target ranges of different blocks must
 * not overlap.  sctl_amd_plan_apply computes, for every block, the same
 * accumulation as sctl_amd_eval restricted to that block, in one persistent
 * launch fed by a work queue.
 *
 * The offset/count arrays are host memory, copied at creation. */
typedef struct sctl_amd_plan_s* sctl_amd_plan;

int sctl_amd_plan_create(sctl_amd_plan* plan, int kernel, int precision, int digits, int64_t nblocks,
                         const int64_t* trg_dsp, const int64_t* trg_cnt, const int64_t* src_dsp,
                         const int64_t* src_cnt, int64_t nt_total, int64_t ns_total);

/* Uploads (HOST_PTRS) or copies (device pointers) the geometry: r_trg
 * nt_total x 3, r_src ns_total x 3, n_src ns_total x 3 (NULL iff no normal).
 * The plan keeps its own device copy, so the caller's buffers may be reused. */
int sctl_amd_plan_set_geometry(sctl_amd_plan plan, const double* r_trg, const double* r_src, const double* n_src,
                               uint32_t flags, void* stream);

/* The kernel's context for later applies and near-field matrices: the
 * wavenumber (*ctx, copied now) of the Helmholtz kernels, which need it
 * (GenericKernel::SetCtxPtr, generic-kernel.txx:64-66).  NULL clears it. */
int sctl_amd_plan_set_ctx(sctl_amd_plan plan, const double* ctx);

/* v_trg (nt_total*K1) += block sums for densities v_src (ns_total*K0).
 * Allocates nothing; may be captured into a hipGraph. */
int sctl_amd_plan_apply(sctl_amd_plan plan, const double* v_src, double* v_trg, uint32_t flags, void* stream);

/* Number of pairwise interactions one apply evaluates (sum of Ns*Nt). */
int sctl_amd_plan_interactions(sctl_amd_plan plan, int64_t* interactions);

int sctl_amd_plan_destroy(sctl_amd_plan plan);

/* Launch knobs of a plan: sources staged through LDS per wave-stage
 * (64 * sources_per_lane), targets per lane (1 or 2; kernels with more than
 * 4 outputs per target always run with 1), and a cap on resident workgroups
 * of 4 waves per CU (0 = no cap).  Chosen at creation from the tuning
 * database, else defaults (256-source stages, 2 targets per lane, cap 2 for
 * mixed precision with one output per target, else 4); sctl_amd_plan_tune
 * searches them. */
int sctl_amd_plan_get_shape(sctl_amd_plan plan, int* sources_per_lane, int* targets_per_lane, int* wg_per_cu_cap);
int sctl_amd_plan_set_shape(sctl_amd_plan plan, int sources_per_lane, int targets_per_lane, int wg_per_cu_cap);

/* ------------------------------------------------------ near-field operator
 * BoundaryIntegralOp's near interactions (boundary_integral.txx:1009-1058,
 * :1165-1240, :1311-1381).  Per element e the reference keeps a target-major
 * matrix K_near_e of near_elem_cnt[e]*KDIM1_ rows by elem_nds_cnt[e]*KDIM0
 * columns (KDIM1_ = KDIM1, or KDIM1/3 when the output is contracted with the
 * target normal, trg_normal_dot_prod_), all elements' matrices concatenated
 * in element order (K_near_dsp, :1034-1058).  A plan whose block b is the
 * near block (elem, t0, cnt) -- targets [near_elem_dsp[elem]+t0, +cnt) of
 * Xtrg_near, sources [elem_nds_dsp_far[elem], +elem_nds_cnt_far[elem]) of
 * X_far -- addresses block b's rows of K_near_elem as one row-major
 * (cnt*KDIM1_) x (src_cnt*KDIM0) matrix. */

/* Replaces the near-block partition of BoundaryIntegralOp::SetupNear
 * (boundary_integral.txx:1009-1032): grain = max(1, N_near/(4*nthreads)),
 * N_near = near_elem_dsp[nelem-1] + near_elem_cnt[nelem-1]; element e's
 * near targets are cut into blocks (e, t0, min(grain, cnt-t0)) for
 * t0 = 0, grain, ...; blocks are listed in element order.  Pass the
 * reference's thread count (SCTL_GET_MAX_THREADS()) for identical blocks.
 * *nblk receives the block count; the three block arrays (host, nblk
 * entries) are filled when given -- call once with NULLs to size them.
 * Host-only; touches no device. */
int sctl_amd_near_blocks(int64_t nelem, const int64_t* near_elem_cnt, const int64_t* near_elem_dsp, int nthreads,
                         int64_t* nblk, int64_t* blk_elem, int64_t* blk_t0, int64_t* blk_cnt);

/* Sets where each block's matrix lives in K_near: blk_off[b] is the offset
 * (in doubles) of block b's first entry, -1 for a block without a matrix
 * (an element list with MatrixFree(), :1047-1048); NULL packs the blocks in
 * block order, which is the reference's layout when the blocks are
 * sctl_amd_near_blocks' blocks of matrix elements
 * (K_near_dsp[e]*KDIM0*KDIM1_ + t0*KDIM1_*src_dof, :1187-1189).
 * trg_normal_dot selects the target-normal contraction (kernels with 3 or 9
 * outputs per target; KDIM1_ = KDIM1/3).
 * *k_len receives the length K_near must have.  Block matrices must not
 * overlap. */
int sctl_amd_plan_near_layout(sctl_amd_plan plan, const int64_t* blk_off, int trg_normal_dot, int64_t* k_len);

#define SCTL_AMD_NEAR_SUBTRACT 0 /* K_near_ -= w * M^T  (KNearSubtract, :1165-1240) */
#define SCTL_AMD_NEAR_STORE 1    /* K_near_  = w * M^T */

/* Replaces the KNearSubtract loop of BoundaryIntegralOp::SetupNear
 * (boundary_integral.txx:1165-1240) with the identity
 * FarFieldDensityOperatorTranspose (:531-533): for every block with a matrix,
 *   K_near_[t*KDIM1_+k1][s*KDIM0+k0] -= wts[s] * M[s*KDIM0+k0][t*KDIM1+k1]
 * with M the block's KernelMatrix at full precision (:1211), or with the
 * target-normal layout
 *   K_near_[t][s*KDIM0+k0] -= sum_l M[s*KDIM0+k0][t*3+l] * wts[s] * n_trg[t][l]
 * (:1193-1209).  wts (ns_total; NULL = all ones) is indexed like the plan's
 * sources, n_trg (nt_total x 3) like its targets.  FP64. */
int sctl_amd_plan_near_matrix(sctl_amd_plan plan, int mode, const double* wts, const double* n_trg, double* K,
                              uint32_t flags, void* stream);

/* Replaces the matrix route of BoundaryIntegralOp::ComputeNearInterac
 * (boundary_integral.txx:1324-1340): for every block,
 *   U_near[(trg_dsp[b])*KDIM1_ + r] = sum_j K_near_[r][j] * F[src_dsp[b]*KDIM0 + j]
 * (overwrites; zero rows for blocks without a matrix or without sources).
 * F (ns_total*KDIM0) holds the element-node densities, U_near
 * (nt_total*KDIM1_) the near targets' values.  FP64. */
int sctl_amd_plan_near_gemv(sctl_amd_plan plan, const double* K, const double* F, double* U_near, uint32_t flags,
                            void* stream);

/* Replaces ScatterForward + the per-target sum of ComputeNearInterac
 * (boundary_integral.txx:1368-1379; comm.txx:1497-1508 on one process):
 *   U[i*dof+k] += sum_{j < near_trg_cnt[i]} U_near[scatter_index[near_trg_dsp[i]+j]*dof + k]
 * with the additions in the reference's order (bit-identical results).
 * dof is 1, 3, 4 or 9; U_near has n_near*dof entries, scatter_index n_scatter.
 * With HOST_PTRS out-of-range indices are an error; with device pointers
 * they are skipped. */
int sctl_amd_near_accumulate(int64_t ntrg, int dof, const int64_t* near_trg_cnt, const int64_t* near_trg_dsp,
                             int64_t n_scatter, const int64_t* scatter_index, int64_t n_near, const double* U_near,
                             double* U, uint32_t flags, void* stream);

/* ------------------------------------------------------------ near lists
 * Replaces BuildNearList (boundary_integral.txx:50-520) on one process:
 * target t is near element e when some node s of e (nodes
 * [elem_nds_dsp[e], +elem_nds_cnt[e]) of Xsrc) has |x_t - x_s|^2 <
 * src_radius[s]^2.  The near pairs are listed by element, then target index;
 * Xtrg_near / Xn_trg_near hold the pairs' target coordinates / normals,
 * near_elem_cnt / near_elem_dsp count them per element, near_scatter_index
 * lists them in target order (pairs of one target in element order; the
 * reference leaves that order unspecified, comm.txx:1393 sorts by key only),
 * near_trg_cnt / near_trg_dsp count them per target (dsp 0 for a target with
 * no pair, as :485-514).  The build synchronises `stream` (it sizes the
 * output); the list keeps its results on the device until destroyed. */
typedef struct sctl_amd_near_list_s* sctl_amd_near_list;

int sctl_amd_near_list_build(sctl_amd_near_list* list, const double* Xtrg, int64_t ntrg, const double* Xn_trg,
                             const double* Xsrc, int64_t nsrc, const double* src_radius, int64_t nelem,
                             const int64_t* elem_nds_cnt, const int64_t* elem_nds_dsp, uint32_t flags,
                             void* stream);
/* Number of near pairs (length of near_scatter_index; Xtrg_near has 3x). */
int sctl_amd_near_list_size(sctl_amd_near_list list, int64_t* n_near);
/* Copies the results out (host with SCTL_AMD_HOST_PTRS, else device); any
 * output may be NULL.  Xn_trg_near needs a list built with Xn_trg. */
int sctl_amd_near_list_get(sctl_amd_near_list list, double* Xtrg_near, double* Xn_trg_near, int64_t* near_elem_cnt,
                           int64_t* near_elem_dsp, int64_t* near_scatter_index, int64_t* near_trg_cnt,
                           int64_t* near_trg_dsp, uint32_t flags, void* stream);
int sctl_amd_near_list_destroy(sctl_amd_near_list list);

/* ------------------------------------------------------ repeated applies
 * A Krylov solve applies the same operator once per iteration with new
 * densities (lin-solve.txx:146-156 calling BoundaryIntegralOp::
 * ComputePotential, boundary_integral.txx:716-722).  This captures nsteps
 * applications of a plan as one hipGraph: for step s, v_src is overwritten
 * with N(0,1) draws of the counter-based generator for (seed + s, stream_id)
 * (as sctl_amd_gen_normal), then v_trg += apply(v_src).  v_src (ns_total*K0)
 * and v_trg (nt_total*K1) are device buffers that the graph keeps using on
 * every launch.  The graph allocates nothing and contains no memset.
 * A graph bakes in its plan's launch shape, work items and context value:
 * after sctl_amd_plan_set_shape, sctl_amd_plan_tune or sctl_amd_plan_set_ctx
 * on that plan, sctl_amd_graph_launch refuses it (SCTL_AMD_ERROR_INVALID_VALUE)
 * until it is captured again.  A graph must not outlive its plan or buffers. */
typedef struct sctl_amd_graph_s* sctl_amd_graph;

int sctl_amd_plan_graph_create(sctl_amd_graph* graph, sctl_amd_plan plan, int nsteps, uint64_t seed,
                               uint64_t stream_id, double* v_src, double* v_trg);
/* The same for a dense plan: v_src (ns*K0), v_trg (nt*K1). */
int sctl_amd_dense_plan_graph_create(sctl_amd_graph* graph, sctl_amd_dense_plan plan, int nsteps, uint64_t seed,
                                     uint64_t stream_id, double* v_src, double* v_trg);
int sctl_amd_graph_launch(sctl_amd_graph graph, void* stream);
int sctl_amd_graph_interactions(sctl_amd_graph graph, int64_t* interactions);
int sctl_amd_graph_destroy(sctl_amd_graph graph);

/* Device counter-based generator: out[i] = i-th N(0,1) draw of
 * (seed, stream_id) -- SplitMix64 counters, Box-Muller pairs. */
int sctl_amd_gen_normal(uint64_t seed, uint64_t stream_id, int64_t n, double* out, void* stream);

/* The Hilbert-curve orders that the mixed-precision path applies to its sources
 * (set a) and targets (set b) before summing: perm_a is the order of the
 * points of r_a (na x 3, device) by their Hilbert key over r_a's bounding
 * cube, ties kept in input order, and likewise for set b.  Keys take 5 bits
 * per axis when na + nb <= 65536, else 10.  No counterpart in the reference
 * (whose direct Eval sums in input order); exposed so that callers and tests
 * can see the order a mixed Eval sums in. */
int sctl_amd_hilbert_order(const double* r_a, int64_t na, uint32_t* perm_a, const double* r_b, int64_t nb,
                           uint32_t* perm_b, void* stream);

/* ----------------------------------------------------------------- tuning
 * The autotuner times candidate launch shapes on the caller's own data (HIP
 * events on `stream`) and records the fastest in a database keyed by
 * operation, kernel, precision and log2 size classes.  The database is kept in
 * memory and, when a path is set (here or through the SCTL_AMD_TUNE_DB
 * environment variable), read from and written to that file.  Later calls of
 * sctl_amd_eval / sctl_amd_plan_create with a matching key use the result. */
int sctl_amd_tune_db_path(const char* path);

/* Dense Eval: targets per lane {1,2,4} x workgroups-per-CU cap {none,3,2}.
 * Device pointers only; v_scratch (nt*K1) receives garbage. */
int sctl_amd_tune_eval(int kernel, int precision, int digits, const double* r_trg, int64_t nt, const double* r_src,
                       const double* n_src, const double* v_src, int64_t ns, double* v_scratch, uint32_t flags,
                       void* stream, int* best_tpl, int* best_occ, double* best_ms);

/* The launch shape sctl_amd_eval uses for this (kernel, precision, digits,
 * flags, nt, ns) size class: get resolves it (tuning database, else
 * defaults); set records one in the database (targets_per_lane 1, 2 or 4,
 * capped at 2 for 4 outputs per target and at 1 for 9; wg_per_cu_cap >= 0).
 * Host-only. */
int sctl_amd_dense_shape_get(int kernel, int precision, int digits, uint32_t flags, int64_t nt, int64_t ns,
                             int* targets_per_lane, int* wg_per_cu_cap);
int sctl_amd_dense_shape_set(int kernel, int precision, int digits, uint32_t flags, int64_t nt, int64_t ns,
                             int targets_per_lane, int wg_per_cu_cap);

/* Plan: sources per lane {1,2,4} x targets per lane {1,2} x cap {none,4,3,2,1},
 * each timed once, the three fastest timed again; leaves the plan configured
 * with the fastest of that second pass.  v_scratch (nt_total*K1,
 * device) receives garbage. */
int sctl_amd_plan_tune(sctl_amd_plan plan, const double* v_src, double* v_scratch, void* stream, int* best_spl,
                       int* best_tpl, int* best_occ, double* best_ms);

/* -------------------------------------------------------------- profiling
 * Counterpart of the reference's Profile FLOP counter, which Eval and
 * KernelMatrix increment by Ns*Nt*FLOPS() on every call
 * (generic-kernel.txx:210,224; profile.txx:466-473).  The library always
 * counts the interactions and the reference-convention FLOPs of the work it
 * launches.  While timing is enabled it also records a pair of HIP events, on
 * the caller's stream, around every main summation kernel (dense Eval,
 * KernelMatrix, plan apply); sctl_amd_profile_read waits for them and adds
 * their elapsed times. */
int sctl_amd_profile_enable(int enable);
int sctl_amd_profile_read(int64_t* interactions, int64_t* flops, int64_t* timed_launches, double* kernel_ms);
int sctl_amd_profile_reset(void);

/* -------------------------------------------------------- device plumbing
 * Minimal device-memory and stream helpers for hosts that have no HIP
 * allocator of their own (bindings, tests).  kind: 0 host->device,
 * 1 device->host, 2 device->device. */
int sctl_amd_malloc(void** ptr, size_t bytes);
int sctl_amd_free(void* ptr);
int sctl_amd_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream);
int sctl_amd_stream_create(void** stream);
int sctl_amd_stream_destroy(void* stream);
int sctl_amd_stream_synchronize(void* stream);

#ifdef __cplusplus
}
#endif

#endif /* SCTL_AMD_H_ */
