// C++ drop-in for SCTL's GenericKernel on MI355X.
//
// sctl_amd::GenericKernel<uKernel> keeps the interface of
// sctl::GenericKernel<uKernel> (include/sctl/generic-kernel.hpp:20-140):
//
//   CoordDim / NormalDim / SrcDim / TrgDim / SetCtxPtr / GetCtxPtr / Name /
//   FLOPS / uKerScaleFactor,
//   member   Eval<Real, enable_openmp, digits>(v_trg, r_trg, r_src, n_src, v_src)
//   static   Eval<Real, enable_openmp>(v_trg, r_trg, r_src, n_src, v_src, digits, self)
//   member   KernelMatrix<Real, enable_openmp, digits>(M, Xt, Xs, Xn)
//
// with the same semantics: sizes are asserted (abort on mismatch, as
// SCTL_ASSERT does, generic-kernel.txx:125-128); v_trg is resized and zeroed
// when its length is not Nt*TrgDim and accumulated into otherwise (:129-132);
// KernelMatrix resizes and zeroes M on a dims mismatch and overwrites it
// (:220-223); digits follows the reference's knob (-1: machine precision).
// The work runs on the current HIP device through the C-ABI of
// include/sctl_amd.h; a failing call aborts with the library's message.
//
// The vector and matrix types are template parameters: any type with the
// interface of sctl::Vector / sctl::Matrix works (Dim(), ReInit(), SetZero(),
// begin(), operator[]), so sctl::Vector<double> arguments are accepted as
// they are.  sctl_amd/containers.hpp provides minimal owning containers for
// hosts without SCTL.
//
// All sixteen micro-kernels of kernel_functions.hpp exist here (Laplace3D
// FxU/DxU/FxdU/DxdU/Fxd2U, Stokes3D FxU/DxU/FxT/FSxU/FxUP, BiotSavart3D
// FxU/FxdU, Helmholtz3D FxU/DxU/FxdU, HelmholtzDiff3D FxdU; the Helmholtz
// kernels read their wavenumber through SetCtxPtr); any other uKernel is a
// compile-time error.
//
// Precision: Real = double runs FP64 unless 0 <= digits <= 7, which selects
// the mixed-precision form (FP64 differencing, FP32 pairwise, Kahan FP64
// accumulation).  Real = float is evaluated by the mixed-precision form and
// rounded to float.
//
// Profiling: Eval and KernelMatrix add Ns*Nt*FLOPS() to the reference
// profiler's FLOP counter (generic-kernel.txx:210,224) when sctl/profile.hpp
// (e.g. through <sctl.hpp>) is included before this header.
#ifndef SCTL_AMD_GENERIC_KERNEL_HPP_
#define SCTL_AMD_GENERIC_KERNEL_HPP_

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <type_traits>
#include <vector>

#include "../sctl_amd.h"

namespace sctl_amd {

using Integer = long;   // sctl::Integer (common.hpp:50)
using Long = int64_t;   // sctl::Long (common.hpp:51)

namespace detail {

[[noreturn]] inline void fail(const char* what, const char* file, int line) {
  std::fprintf(stderr, "\n%s:%d: sctl_amd: %s\n", file, line, what);
  std::abort();
}

#define SCTL_AMD_ASSERT(cond)                                                          \
  do {                                                                                 \
    if (!(cond)) ::sctl_amd::detail::fail("Assertion `" #cond "' failed.", __FILE__, __LINE__); \
  } while (0)

inline void check(int status, const char* file, int line) {
  if (status != SCTL_AMD_SUCCESS) fail(sctl_amd_last_error(), file, line);
}

// Ns*Nt*FLOPS() into the reference profiler's FLOP counter, as its Eval and
// KernelMatrix do (generic-kernel.txx:210,224), when that profiler is part of
// the translation unit (sctl/profile.hpp included before this header).
inline void count_flops(int64_t n) {
#if defined(_SCTL_PROFILE_HPP_)
  sctl::Profile::IncrementCounter(sctl::ProfileCounter::FLOP, (sctl::Long)n);
#else
  (void)n;
#endif
}
#define SCTL_AMD_CHECK(expr) ::sctl_amd::detail::check((expr), __FILE__, __LINE__)

}  // namespace detail

namespace kernel_impl {

// The real-valued micro-kernels: Name() and FLOPS() as in
// kernel_functions.hpp, uKerScaleFactor<Real>() with the reference's const_pi
// (math_utils.hpp:43).
// True for the micro-kernel types of this header only.
template <class T> struct is_device_kernel : std::false_type {};

template <class Real> constexpr Real const_pi() {
  return (Real)14488038916154245684ull / (Real)(1ull << 62) +
         (Real)7089564414062235240ull / (Real)(1ull << 62) / (Real)(1ull << 63);
}

struct Laplace3D_FxU {  // kernel_functions.hpp:32-59
  static constexpr int KERNEL_ID = SCTL_AMD_LAPLACE3D_FXU;
  static constexpr Integer DIM = 3, KDIM0 = 1, KDIM1 = 1, N_DIM = 0;
  static const std::string& Name() {
    static const std::string name = "Laplace3D-FxU";
    return name;
  }
  static constexpr Integer FLOPS() { return 6; }
  template <class Real> static constexpr Real uKerScaleFactor() { return 1 / (4 * const_pi<Real>()); }
};

struct Laplace3D_DxU {  // kernel_functions.hpp:61-79
  static constexpr int KERNEL_ID = SCTL_AMD_LAPLACE3D_DXU;
  static constexpr Integer DIM = 3, KDIM0 = 1, KDIM1 = 1, N_DIM = 3;
  static const std::string& Name() {
    static const std::string name = "Laplace3D-DxU";
    return name;
  }
  static constexpr Integer FLOPS() { return 14; }
  template <class Real> static constexpr Real uKerScaleFactor() { return 1 / (4 * const_pi<Real>()); }
};

struct Stokes3D_FxU {  // kernel_functions.hpp:125-178
  static constexpr int KERNEL_ID = SCTL_AMD_STOKES3D_FXU;
  static constexpr Integer DIM = 3, KDIM0 = 3, KDIM1 = 3, N_DIM = 0;
  static const std::string& Name() {
    static const std::string name = "Stokes3D-FxU";
    return name;
  }
  static constexpr Integer FLOPS() { return 23; }
  template <class Real> static constexpr Real uKerScaleFactor() { return 1 / (8 * const_pi<Real>()); }
};

struct Stokes3D_DxU {  // kernel_functions.hpp:180-218
  static constexpr int KERNEL_ID = SCTL_AMD_STOKES3D_DXU;
  static constexpr Integer DIM = 3, KDIM0 = 3, KDIM1 = 3, N_DIM = 3;
  static const std::string& Name() {
    static const std::string name = "Stokes3D-DxU";
    return name;
  }
  static constexpr Integer FLOPS() { return 26; }
  template <class Real> static constexpr Real uKerScaleFactor() { return 3 / (4 * const_pi<Real>()); }
};

#define SCTL_AMD_UKERNEL(NAME, ID, K0, K1, ND, STR, FLOPS_, NUM, DEN)                    \
  struct NAME {                                                                       \
    static constexpr int KERNEL_ID = ID;                                              \
    static constexpr Integer DIM = 3, KDIM0 = K0, KDIM1 = K1, N_DIM = ND;             \
    static const std::string& Name() {                                                \
      static const std::string name = STR;                                            \
      return name;                                                                    \
    }                                                                                 \
    static constexpr Integer FLOPS() { return FLOPS_; }                               \
    template <class Real> static constexpr Real uKerScaleFactor() {                   \
      return NUM / (DEN * const_pi<Real>());                                          \
    }                                                                                 \
  };                                                                                  \
  template <> struct is_device_kernel<NAME> : std::true_type {}

SCTL_AMD_UKERNEL(Laplace3D_FxdU, SCTL_AMD_LAPLACE3D_FXDU, 1, 3, 0, "Laplace3D-FxdU", 11, -1, 4);      // :81-123
SCTL_AMD_UKERNEL(Laplace3D_DxdU, SCTL_AMD_LAPLACE3D_DXDU, 1, 3, 3, "Laplace3D-DxdU", 22, 1, 4);       // :392-414
SCTL_AMD_UKERNEL(Laplace3D_Fxd2U, SCTL_AMD_LAPLACE3D_FXD2U, 1, 9, 0, "Laplace3D-Fxd2U", 25, 1, 4);    // :366-390
SCTL_AMD_UKERNEL(Stokes3D_FxT, SCTL_AMD_STOKES3D_FXT, 3, 9, 0, "Stokes3D-FxT", 39, -3, 4);            // :220-263
SCTL_AMD_UKERNEL(Stokes3D_FSxU, SCTL_AMD_STOKES3D_FSXU, 4, 3, 0, "Stokes3D-FSxU", 26, 1, 8);          // :264-324
SCTL_AMD_UKERNEL(Stokes3D_FxUP, SCTL_AMD_STOKES3D_FXUP, 3, 4, 0, "Stokes3D-FxUP", 26, 1, 8);          // :325-365
SCTL_AMD_UKERNEL(BiotSavart3D_FxU, SCTL_AMD_BIOTSAVART3D_FXU, 3, 3, 0, "BiotSavart3D-FxU", 15, 1, 4); // :416-450
SCTL_AMD_UKERNEL(BiotSavart3D_FxdU, SCTL_AMD_BIOTSAVART3D_FXDU, 3, 9, 0, "BiotSavart3D-FxdU", 51, -1, 4);  // :451-489
// The Helmholtz kernels read the wavenumber through SetCtxPtr (a pointer to one Real).
SCTL_AMD_UKERNEL(Helmholtz3D_FxU, SCTL_AMD_HELMHOLTZ3D_FXU, 2, 2, 0, "Helmholtz3D-FxU", 20, 1, 4);    // :491-527
SCTL_AMD_UKERNEL(Helmholtz3D_DxU, SCTL_AMD_HELMHOLTZ3D_DXU, 2, 2, 3, "Helmholtz3D-DxU", 29, -1, 4);   // :529-571
SCTL_AMD_UKERNEL(Helmholtz3D_FxdU, SCTL_AMD_HELMHOLTZ3D_FXDU, 2, 6, 0, "Helmholtz3D-FxdU", 35, 1, 4);  // :573-620
SCTL_AMD_UKERNEL(HelmholtzDiff3D_FxdU, SCTL_AMD_HELMHOLTZDIFF3D_FXDU, 2, 6, 0, "HelmholtzDiff3D-FxdU", 41, 1,
                 4);  // :627-659
#undef SCTL_AMD_UKERNEL
template <> struct is_device_kernel<Laplace3D_FxU> : std::true_type {};
template <> struct is_device_kernel<Laplace3D_DxU> : std::true_type {};
template <> struct is_device_kernel<Stokes3D_FxU> : std::true_type {};
template <> struct is_device_kernel<Stokes3D_DxU> : std::true_type {};

}  // namespace kernel_impl

template <class uKernel> class GenericKernel : public uKernel {
  // Only the kernel_impl types above (which carry a device kernel id) run on
  // the device; sctl::kernel_impl types fail here rather than fall back to
  // the CPU.
  static_assert(kernel_impl::is_device_kernel<uKernel>::value,
                "sctl_amd::GenericKernel: only the sctl_amd::kernel_impl micro-kernels run on the device");
  static constexpr Integer DIM = uKernel::DIM, KDIM0 = uKernel::KDIM0, KDIM1 = uKernel::KDIM1,
                           N_DIM = uKernel::N_DIM;

 public:
  GenericKernel() : ctx_ptr(nullptr) {}

  static constexpr Integer CoordDim() { return DIM; }
  static constexpr Integer NormalDim() { return N_DIM; }
  static constexpr Integer SrcDim() { return KDIM0; }
  static constexpr Integer TrgDim() { return KDIM1; }

  void SetCtxPtr(void* ctx) { ctx_ptr = ctx; }
  const void* GetCtxPtr() const { return ctx_ptr; }

  // Type-erased form stored as a function pointer by ParticleFMM
  // (generic-kernel.txx:72-100, fmm-wrapper.txx:153-154,450-451).  The
  // reference types `self` as sctl::ConstIterator<char>: `const char*` in
  // default builds, a checked iterator class under SCTL_MEMDEBUG
  // (common.hpp:110-118).  SelfT is deduced from the function-pointer type
  // the caller assigns to, so `KerS2T::template Eval<Real,false>` binds in
  // both builds.
  template <class Real, bool enable_openmp, class VectorT, class SelfT>
  static void Eval(VectorT& v_trg, const VectorT& r_trg, const VectorT& r_src, const VectorT& n_src,
                   const VectorT& v_src, Integer digits, SelfT self) {
    // self is the kernel object (generic-kernel.txx:75): its context travels along.
    const char* p = address_of(self);
    const void* ctx = p ? reinterpret_cast<const GenericKernel*>(p)->ctx_ptr : nullptr;
    EvalImpl<Real>(v_trg, r_trg, r_src, n_src, v_src, (digits >= 16 ? -1 : digits), ctx);
  }

  template <class Real, bool enable_openmp = false, Integer digits = -1, class VectorT>
  void Eval(VectorT& v_trg, const VectorT& r_trg, const VectorT& r_src, const VectorT& n_src,
            const VectorT& v_src) const {
    EvalImpl<Real>(v_trg, r_trg, r_src, n_src, v_src, digits, ctx_ptr);
  }

  template <class Real, bool enable_openmp = false, Integer digits = -1, class MatrixT, class VectorT>
  void KernelMatrix(MatrixT& M, const VectorT& Xt, const VectorT& Xs, const VectorT& Xn) const {
    const Long Ns = Xs.Dim() / DIM;
    const Long Nt = Xt.Dim() / DIM;
    if (M.Dim(0) != Ns * KDIM0 || M.Dim(1) != Nt * KDIM1) {
      M.ReInit(Ns * KDIM0, Nt * KDIM1);
      M.SetZero();
    }
    detail::count_flops(Ns * Nt * uKernel::FLOPS());
    if (!Ns || !Nt) return;
    SCTL_AMD_ASSERT(!N_DIM || Xn.Dim() >= Ns * N_DIM);
    std::vector<double> xt, xs, xn, m;
    const double* pt = as_double(Xt, xt);
    const double* ps = as_double(Xs, xs);
    const double* pn = N_DIM ? as_double(Xn, xn) : nullptr;
    const Long n = Ns * KDIM0 * Nt * KDIM1;
    double mu;
    const double* ctx = ctx_value<Real>(ctx_ptr, &mu);
    if (std::is_same<Real, double>::value) {
      SCTL_AMD_CHECK(sctl_amd_kernel_matrix_ctx(uKernel::KERNEL_ID, ctx, (int)(digits >= 16 ? -1 : digits), pt, Nt,
                                                ps, pn, Ns, reinterpret_cast<double*>(&M[0][0]), SCTL_AMD_HOST_PTRS,
                                                nullptr));
    } else {
      m.resize(n);
      SCTL_AMD_CHECK(sctl_amd_kernel_matrix_ctx(uKernel::KERNEL_ID, ctx, (int)(digits >= 16 ? -1 : digits), pt, Nt,
                                                ps, pn, Ns, m.data(), SCTL_AMD_HOST_PTRS, nullptr));
      Real* out = &M[0][0];
      for (Long i = 0; i < n; i++) out[i] = (Real)m[i];
    }
  }

 private:
  // Address held by `self`: a plain pointer, or an iterator class whose
  // operator-> yields it (sctl::ConstIterator under SCTL_MEMDEBUG,
  // iterator.hpp:19-78).
  template <class T> static const char* address_of(T* p) { return reinterpret_cast<const char*>(p); }
  static const char* address_of(std::nullptr_t) { return nullptr; }
  template <class It> static const char* address_of(const It& it) {
    return reinterpret_cast<const char*>(it.operator->());
  }

  // The kernels read their context as *(const Real*)ctx_ptr (kernel_functions.hpp:485-489);
  // the library takes it as a double.
  template <class Real> static const double* ctx_value(const void* ctx, double* tmp) {
    if (!ctx) return nullptr;
    *tmp = (double)*static_cast<const Real*>(ctx);
    return tmp;
  }

  template <class VectorT> static const double* as_double(const VectorT& v, std::vector<double>& tmp) {
    using T = typename std::remove_cv<typename std::remove_reference<decltype(v[0])>::type>::type;
    if (std::is_same<T, double>::value) return v.Dim() ? reinterpret_cast<const double*>(&v[0]) : nullptr;
    tmp.resize(v.Dim());
    for (Long i = 0; i < (Long)v.Dim(); i++) tmp[i] = (double)v[i];
    return tmp.data();
  }

  template <class Real, class VectorT>
  static void EvalImpl(VectorT& v_trg, const VectorT& r_trg, const VectorT& r_src, const VectorT& n_src,
                       const VectorT& v_src, Integer digits, const void* ctx_in) {
    const Long Ns = r_src.Dim() / DIM;
    const Long Nt = r_trg.Dim() / DIM;
    SCTL_AMD_ASSERT(r_trg.Dim() == Nt * DIM);
    SCTL_AMD_ASSERT(r_src.Dim() == Ns * DIM);
    SCTL_AMD_ASSERT(v_src.Dim() == Ns * KDIM0);
    SCTL_AMD_ASSERT(n_src.Dim() == Ns * N_DIM || !N_DIM);
    if (v_trg.Dim() != Nt * KDIM1) {
      v_trg.ReInit(Nt * KDIM1);
      v_trg.SetZero();
    }
    detail::count_flops(Ns * Nt * uKernel::FLOPS());
    if (!Ns || !Nt) return;
    const bool is_double = std::is_same<Real, double>::value;
    const int precision = is_double ? SCTL_AMD_AUTO : SCTL_AMD_MIXED;
    std::vector<double> xt, xs, xn, vs, vt;
    const double* pt = as_double(r_trg, xt);
    const double* ps = as_double(r_src, xs);
    const double* pn = N_DIM ? as_double(n_src, xn) : nullptr;
    const double* pv = as_double(v_src, vs);
    double mu;
    const double* ctx = ctx_value<Real>(ctx_in, &mu);
    if (is_double) {
      SCTL_AMD_CHECK(sctl_amd_eval_ctx(uKernel::KERNEL_ID, ctx, precision, (int)digits, pt, Nt, ps, pn, pv, Ns,
                                       reinterpret_cast<double*>(&v_trg[0]), SCTL_AMD_HOST_PTRS, nullptr));
    } else {
      vt.assign(Nt * KDIM1, 0.0);
      SCTL_AMD_CHECK(sctl_amd_eval_ctx(uKernel::KERNEL_ID, ctx, precision, (int)digits, pt, Nt, ps, pn, pv, Ns,
                                       vt.data(), SCTL_AMD_HOST_PTRS, nullptr));
      for (Long i = 0; i < Nt * KDIM1; i++) v_trg[i] += (Real)vt[i];
    }
  }

  void* ctx_ptr;
};

// kernel_functions.hpp:669-684 aliases for the four device kernels.
using Laplace3D_FxU = GenericKernel<kernel_impl::Laplace3D_FxU>;
using Laplace3D_DxU = GenericKernel<kernel_impl::Laplace3D_DxU>;
using Stokes3D_FxU = GenericKernel<kernel_impl::Stokes3D_FxU>;
using Stokes3D_DxU = GenericKernel<kernel_impl::Stokes3D_DxU>;
using Laplace3D_FxdU = GenericKernel<kernel_impl::Laplace3D_FxdU>;
using Laplace3D_DxdU = GenericKernel<kernel_impl::Laplace3D_DxdU>;
using Laplace3D_Fxd2U = GenericKernel<kernel_impl::Laplace3D_Fxd2U>;
using Stokes3D_FxT = GenericKernel<kernel_impl::Stokes3D_FxT>;
using Stokes3D_FSxU = GenericKernel<kernel_impl::Stokes3D_FSxU>;
using Stokes3D_FxUP = GenericKernel<kernel_impl::Stokes3D_FxUP>;
using BiotSavart3D_FxU = GenericKernel<kernel_impl::BiotSavart3D_FxU>;
using BiotSavart3D_FxdU = GenericKernel<kernel_impl::BiotSavart3D_FxdU>;
using Helmholtz3D_FxU = GenericKernel<kernel_impl::Helmholtz3D_FxU>;
using Helmholtz3D_DxU = GenericKernel<kernel_impl::Helmholtz3D_DxU>;
using Helmholtz3D_FxdU = GenericKernel<kernel_impl::Helmholtz3D_FxdU>;
using HelmholtzDiff3D_FxdU = GenericKernel<kernel_impl::HelmholtzDiff3D_FxdU>;

}  // namespace sctl_amd

#endif  // SCTL_AMD_GENERIC_KERNEL_HPP_
