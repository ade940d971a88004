// Near-field set-up of BoundaryIntegralOp from element lists, on the device.
//
// sctl_amd::NearFieldOp<Real, uKernel, VectorT> takes the reference's element lists as they
// are -- any type with the ElementListBase<Real> interface
// (boundary_integral.hpp:65-240): GetNodeCoord, GetFarFieldNodes, MatrixFree
// -- and does on the device what BoundaryIntegralOp's set-up and apply do for
// the near field:
//
//   Setup(tol)       SetupFar (boundary_integral.txx:852-890): every list's
//                    far nodes, normals, weights, dist_far and per-element
//                    node counts, concatenated in list order; then the near
//                    list (BuildNearList, :50-520, with dist_far as each
//                    node's radius, :948), the block partition (:1009-1032),
//                    one plan over (block targets, element far nodes) and
//                    K_near's layout (:1034-1058; no matrix for the elements
//                    of MatrixFree() lists).
//   KNearSubtract    K_near -= the direct part of every block (:1165-1240),
//                    with the identity FarFieldDensityOperatorTranspose
//                    (:531-533), the only one the reference's lists use.
//   NearInterac      ComputeNearInterac's matrix route (:1311-1381): the
//                    GEMV of K_near with the element-node densities, the
//                    ScatterForward to target order and the per-target sum.
//   NearDirect       the direct part itself, sum over near elements of
//                    w F K (the matrix-free route's direct sum), through
//                    the plan's batched kernel.
//
// The element-specific singular quadrature that fills K_near before the
// subtraction (ElementListBase::NearInterac, SelfInterac) stays with the
// caller.  Vectors are the caller's type (sctl::Vector or any type with its
// interface, VectorT<Real> / VectorT<Long>, the types the lists' methods
// take); K_near is a host array or a device pointer.  Errors abort with the
// library's message, as the rest of the drop-in does.
#ifndef SCTL_AMD_NEAR_FIELD_HPP_
#define SCTL_AMD_NEAR_FIELD_HPP_

#include <algorithm>
#include <cstdint>
#include <memory>
#include <vector>

#include "generic_kernel.hpp"

namespace sctl_amd {

template <class Real, class uKernel, template <class> class VectorT> class NearFieldOp {
  static constexpr int KER = uKernel::KERNEL_ID;
  static constexpr Long DIM = 3, K0 = uKernel::SrcDim(), K1 = uKernel::TrgDim(), ND = uKernel::NormalDim();

 public:
  NearFieldOp() = default;
  NearFieldOp(const NearFieldOp&) = delete;
  NearFieldOp& operator=(const NearFieldOp&) = delete;
  ~NearFieldOp() { Clear(); }

  // Element lists are kept by reference (as BoundaryIntegralOp::AddElemList
  // keeps its lists) and read at Setup.
  template <class ElemList> void AddElemList(const ElemList& lst) {
    lists_.push_back(std::unique_ptr<ListRef>(new ListRefT<ElemList>(lst)));
  }

  // Targets (and their normals when the output is contracted with them).
  void SetTargets(const VectorT<Real>& Xtrg, const VectorT<Real>& Xn_trg, bool trg_normal_dot_prod = false) {
    Xtrg_.assign(Xtrg.Dim(), 0.0);
    for (Long i = 0; i < (Long)Xtrg.Dim(); i++) Xtrg_[i] = (double)Xtrg[i];
    Xn_trg_.assign(Xn_trg.Dim(), 0.0);
    for (Long i = 0; i < (Long)Xn_trg.Dim(); i++) Xn_trg_[i] = (double)Xn_trg[i];
    ndot_ = trg_normal_dot_prod;
  }

  void SetCtxPtr(const double* ctx) { ctx_ = ctx; }

  // SetupFar + the near-list, block and K_near-layout part of SetupNear.
  void Setup(Real tol, int nthreads = 1) {
    Clear();
    // SetupFar: concatenation in list order (boundary_integral.txx:863-883).
    std::vector<double> wts;
    for (auto& l : lists_) {
      l->far_nodes(tol, X_far_, Xn_far_, wts, dist_far_, far_cnt_);
      l->node_counts(nds_cnt_);
      SCTL_AMD_ASSERT((Long)far_cnt_.size() == (Long)nds_cnt_.size());
      matrix_free_.resize(far_cnt_.size(), l->matrix_free());
    }
    wts_far_ = wts;
    const Long nelem = (Long)far_cnt_.size();
    far_dsp_.assign(nelem, 0);
    for (Long e = 1; e < nelem; e++) far_dsp_[e] = far_dsp_[e - 1] + far_cnt_[e - 1];
    const Long ns = (Long)X_far_.size() / DIM, nt = (Long)Xtrg_.size() / DIM;
    SCTL_AMD_ASSERT(!ND || (Long)Xn_far_.size() == ns * DIM);
    SCTL_AMD_ASSERT((Long)dist_far_.size() == ns && (Long)wts_far_.size() == ns);

    // BuildNearList with dist_far as each far node's radius (:948).
    sctl_amd_near_list nl = nullptr;
    const bool with_n = ndot_ && !Xn_trg_.empty();
    SCTL_AMD_CHECK(sctl_amd_near_list_build(&nl, Xtrg_.data(), nt, with_n ? Xn_trg_.data() : nullptr, X_far_.data(),
                                            ns, dist_far_.data(), nelem, far_cnt_.data(), far_dsp_.data(),
                                            SCTL_AMD_HOST_PTRS, nullptr));
    int64_t n_near = 0;
    SCTL_AMD_CHECK(sctl_amd_near_list_size(nl, &n_near));
    Xtrg_near_.assign(3 * n_near, 0.0);
    Xn_trg_near_.assign(with_n ? 3 * n_near : 0, 0.0);
    near_elem_cnt_.assign(nelem, 0);
    near_elem_dsp_.assign(nelem, 0);
    scatter_.assign(n_near, 0);
    near_trg_cnt_.assign(nt, 0);
    near_trg_dsp_.assign(nt, 0);
    const int rc = sctl_amd_near_list_get(nl, Xtrg_near_.data(), with_n ? Xn_trg_near_.data() : nullptr,
                                          near_elem_cnt_.data(), near_elem_dsp_.data(), scatter_.data(),
                                          near_trg_cnt_.data(), near_trg_dsp_.data(), SCTL_AMD_HOST_PTRS, nullptr);
    sctl_amd_near_list_destroy(nl);
    SCTL_AMD_CHECK(rc);

    // Block partition (:1009-1032) and the plan over the blocks.
    int64_t nblk = 0;
    SCTL_AMD_CHECK(sctl_amd_near_blocks(nelem, near_elem_cnt_.data(), near_elem_dsp_.data(), nthreads, &nblk,
                                        nullptr, nullptr, nullptr));
    blk_elem_.assign(nblk, 0);
    blk_t0_.assign(nblk, 0);
    blk_cnt_.assign(nblk, 0);
    SCTL_AMD_CHECK(sctl_amd_near_blocks(nelem, near_elem_cnt_.data(), near_elem_dsp_.data(), nthreads, &nblk,
                                        blk_elem_.data(), blk_t0_.data(), blk_cnt_.data()));
    std::vector<int64_t> td(nblk), sd(nblk), sc(nblk);
    for (Long b = 0; b < nblk; b++) {
      td[b] = near_elem_dsp_[blk_elem_[b]] + blk_t0_[b];
      sd[b] = far_dsp_[blk_elem_[b]];
      sc[b] = far_cnt_[blk_elem_[b]];
    }
    SCTL_AMD_CHECK(sctl_amd_plan_create(&plan_, KER, SCTL_AMD_FP64, -1, nblk, td.data(), blk_cnt_.data(), sd.data(),
                                        sc.data(), n_near, ns));
    if (ctx_) SCTL_AMD_CHECK(sctl_amd_plan_set_ctx(plan_, ctx_));
    SCTL_AMD_CHECK(sctl_amd_plan_set_geometry(plan_, Xtrg_near_.data(), X_far_.data(), ND ? Xn_far_.data() : nullptr,
                                              SCTL_AMD_HOST_PTRS, nullptr));

    // K_near layout (:1039-1058; :1187-1189 per block): no matrix for the
    // elements of matrix-free lists.
    const Long dof = ndot_ ? K1 / DIM : K1;
    std::vector<int64_t> kcnt(nelem), kdsp(nelem, 0), off(nblk);
    for (Long e = 0; e < nelem; e++) {
      // the identity FarFieldDensityOperatorTranspose: far nodes are the element's nodes
      SCTL_AMD_ASSERT(matrix_free_[e] || far_cnt_[e] == nds_cnt_[e]);
      kcnt[e] = matrix_free_[e] ? 0 : nds_cnt_[e] * near_elem_cnt_[e];
    }
    for (Long e = 1; e < nelem; e++) kdsp[e] = kdsp[e - 1] + kcnt[e - 1];
    for (Long b = 0; b < nblk; b++) {
      const Long e = blk_elem_[b];
      off[b] = matrix_free_[e] ? -1 : kdsp[e] * K0 * dof + blk_t0_[b] * dof * (nds_cnt_[e] * K0);
    }
    SCTL_AMD_CHECK(sctl_amd_plan_near_layout(plan_, off.data(), ndot_ ? 1 : 0, &k_len_));
    const int64_t ref_len = nelem ? (kdsp[nelem - 1] + kcnt[nelem - 1]) * K0 * dof : 0;
    SCTL_AMD_ASSERT(k_len_ <= ref_len);
    k_len_ = ref_len;
    // device copies of the weights and target normals for device-resident K_near
    upload(&d_wts_, wts_far_);
    upload(&d_ntrg_, Xn_trg_near_);
  }

  // Length of K_near (doubles) in the reference's layout.
  int64_t KNearLength() const { return k_len_; }
  // Near targets (rows of K_near), per element: the reference's
  // near_elem_cnt / near_elem_dsp and Xtrg_near.
  const std::vector<int64_t>& NearElemCnt() const { return near_elem_cnt_; }
  const std::vector<int64_t>& NearElemDsp() const { return near_elem_dsp_; }
  const std::vector<double>& XtrgNear() const { return Xtrg_near_; }
  const std::vector<int64_t>& FarNodeCnt() const { return far_cnt_; }
  const std::vector<int64_t>& FarNodeDsp() const { return far_dsp_; }

  // K_near -= direct part: K_near of KNearLength() doubles in host memory
  // (staged through the device) ...
  void KNearSubtract(std::vector<double>& K_near) {
    SCTL_AMD_ASSERT(plan_ && (int64_t)K_near.size() >= k_len_);
    const bool with_n = ndot_ && !Xn_trg_near_.empty();
    SCTL_AMD_CHECK(sctl_amd_plan_near_matrix(plan_, SCTL_AMD_NEAR_SUBTRACT, wts_far_.data(),
                                             with_n ? Xn_trg_near_.data() : nullptr, K_near.data(),
                                             SCTL_AMD_HOST_PTRS, nullptr));
  }
  // ... or device memory (stream-ordered on `stream`, NULL: per-thread stream).
  void KNearSubtractDevice(double* K_near_dev, void* stream = nullptr) {
    SCTL_AMD_ASSERT(plan_);
    const bool with_n = ndot_ && !Xn_trg_near_.empty();
    SCTL_AMD_CHECK(sctl_amd_plan_near_matrix(plan_, SCTL_AMD_NEAR_SUBTRACT, d_wts_, with_n ? d_ntrg_ : nullptr,
                                             K_near_dev, 0, stream));
  }

  // U (Ntrg*KDIM1_, host) += near interactions from K_near (host) and the
  // element-node densities F (host): ComputeNearInterac's matrix route.
  void NearInterac(std::vector<double>& U, const std::vector<double>& K_near, const std::vector<double>& F) {
    SCTL_AMD_ASSERT(plan_);
    const Long dof = ndot_ ? K1 / DIM : K1, nt = (Long)Xtrg_.size() / DIM;
    if ((Long)U.size() != nt * dof) U.assign(nt * dof, 0.0);
    std::vector<double> U_near(Xtrg_near_.size() / DIM * dof, 0.0);
    SCTL_AMD_CHECK(sctl_amd_plan_near_gemv(plan_, K_near.data(), F.data(), U_near.data(), SCTL_AMD_HOST_PTRS,
                                           nullptr));
    SCTL_AMD_CHECK(sctl_amd_near_accumulate(nt, (int)dof, near_trg_cnt_.data(), near_trg_dsp_.data(),
                                            (int64_t)scatter_.size(), scatter_.data(), (int64_t)U_near.size() / dof,
                                            U_near.data(), U.data(), SCTL_AMD_HOST_PTRS, nullptr));
  }

  // U_near (rows of Xtrg_near, host) = sum over each near target's element
  // of K(x_t - y) w F: the direct part of the near interactions.
  void NearDirect(std::vector<double>& U_near, const std::vector<double>& F) {
    SCTL_AMD_ASSERT(plan_);
    std::vector<double> wF(F.size());
    for (size_t i = 0; i < F.size(); i++) wF[i] = wts_far_[i / K0] * F[i];
    U_near.assign(Xtrg_near_.size() / DIM * K1, 0.0);
    SCTL_AMD_CHECK(sctl_amd_plan_apply(plan_, wF.data(), U_near.data(), SCTL_AMD_HOST_PTRS, nullptr));
  }

  void Clear() {
    if (plan_) sctl_amd_plan_destroy(plan_);
    plan_ = nullptr;
    if (d_wts_) sctl_amd_free(d_wts_);
    if (d_ntrg_) sctl_amd_free(d_ntrg_);
    d_wts_ = d_ntrg_ = nullptr;
    X_far_.clear();
    Xn_far_.clear();
    wts_far_.clear();
    dist_far_.clear();
    far_cnt_.clear();
    far_dsp_.clear();
    nds_cnt_.clear();
    matrix_free_.clear();
  }

 private:
  static void upload(double** dst, const std::vector<double>& v) {
    *dst = nullptr;
    if (v.empty()) return;
    SCTL_AMD_CHECK(sctl_amd_malloc(reinterpret_cast<void**>(dst), sizeof(double) * v.size()));
    SCTL_AMD_CHECK(sctl_amd_memcpy(*dst, v.data(), sizeof(double) * v.size(), 0, nullptr));
  }

  // Type-erased access to one element list through the ElementListBase
  // interface, with the caller's vector types.
  struct ListRef {
    virtual ~ListRef() = default;
    virtual void far_nodes(Real tol, std::vector<double>& X, std::vector<double>& Xn, std::vector<double>& wts,
                           std::vector<double>& dist, std::vector<int64_t>& cnt) const = 0;
    virtual void node_counts(std::vector<int64_t>& cnt) const = 0;
    virtual bool matrix_free() const = 0;
  };
  template <class ElemList> struct ListRefT : ListRef {
    explicit ListRefT(const ElemList& l) : lst(l) {}
    template <class V, class T> static void append(const V& v, std::vector<T>& out) {
      for (Long i = 0; i < (Long)v.Dim(); i++) out.push_back((T)v[i]);
    }
    void far_nodes(Real tol, std::vector<double>& X, std::vector<double>& Xn, std::vector<double>& wts,
                   std::vector<double>& dist, std::vector<int64_t>& cnt) const override {
      VectorT<Real> X_, Xn_, w_, d_;
      VectorT<Long> c_;
      lst.GetFarFieldNodes(X_, Xn_, w_, d_, c_, tol);  // boundary_integral.txx:870
      append(X_, X);
      append(Xn_, Xn);
      append(w_, wts);
      append(d_, dist);
      append(c_, cnt);
    }
    void node_counts(std::vector<int64_t>& cnt) const override {
      VectorT<Long> c_;
      lst.GetNodeCoord(nullptr, nullptr, &c_);  // boundary_integral.txx:790-850 (SetupBasic)
      append(c_, cnt);
    }
    bool matrix_free() const override { return lst.MatrixFree(); }
    const ElemList& lst;
  };

  std::vector<std::unique_ptr<ListRef>> lists_;
  std::vector<double> Xtrg_, Xn_trg_, X_far_, Xn_far_, wts_far_, dist_far_, Xtrg_near_, Xn_trg_near_;
  std::vector<int64_t> far_cnt_, far_dsp_, nds_cnt_, near_elem_cnt_, near_elem_dsp_, scatter_, near_trg_cnt_,
      near_trg_dsp_, blk_elem_, blk_t0_, blk_cnt_;
  std::vector<bool> matrix_free_;
  bool ndot_ = false;
  const double* ctx_ = nullptr;
  double *d_wts_ = nullptr, *d_ntrg_ = nullptr;
  sctl_amd_plan plan_ = nullptr;
  int64_t k_len_ = 0;
};

}  // namespace sctl_amd

#endif  // SCTL_AMD_NEAR_FIELD_HPP_
