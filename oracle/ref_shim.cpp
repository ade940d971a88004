// Test infrastructure only: a C-ABI wrapper that instantiates the reference
// library's own GenericKernel code so that it can be called from the test
// suite and from the CPU-baseline timer.  It is compiled from the reference
// headers in place (see oracle/Makefile) into oracle/_ref/, never shipped as
// part of the product and never called by it.
//
// Wrapped reference entry points:
//   GenericKernel<uKernel>::Eval<Real,omp>(..., digits, self)   generic-kernel.txx:72-100
//   GenericKernel<uKernel>::KernelMatrix<Real,false,digits>     generic-kernel.txx:213-329
//   approx_rsqrt<digits>(x, mask)                               vec.txx:379-382
// for the micro-kernels of kernel_functions.hpp:32-659 (the Helmholtz kernels
// with their wavenumber set through GenericKernel::SetCtxPtr,
// generic-kernel.txx:64-66).
#include <cstring>
#include <string>

#include "sctl/common.hpp"
#include "sctl/iterator.hpp"
#include "sctl/iterator.txx"
#include "sctl/vector.hpp"
#include "sctl/vector.txx"
#include "sctl/matrix.hpp"
#include "sctl/matrix.txx"
#include "sctl/generic-kernel.hpp"
#include "sctl/generic-kernel.txx"
#include "sctl/kernel_functions.hpp"

using sctl::Long;
using sctl::Integer;

namespace {

template <class Real> using Vec = sctl::Vector<Real>;

template <class K, class Real>
void eval_one(int omp, int digits, const Real* r_trg, long nt, const Real* r_src, const Real* n_src,
              const Real* v_src, long ns, Real* v_trg, const double* ctx) {
  K ker;
  Real mu = ctx ? (Real)*ctx : (Real)0;  // the kernels read *(const Real*)ctx_ptr
  if (ctx) ker.SetCtxPtr(&mu);
  const Long nd = K::NormalDim();
  Vec<Real> Xt(nt * 3, (sctl::Iterator<Real>)r_trg, false);
  Vec<Real> Xs(ns * 3, (sctl::Iterator<Real>)r_src, false);
  Vec<Real> Xn(nd ? ns * nd : 0, (sctl::Iterator<Real>)(nd ? n_src : nullptr), false);
  Vec<Real> Vs(ns * K::SrcDim(), (sctl::Iterator<Real>)v_src, false);
  Vec<Real> Vt(nt * K::TrgDim(), v_trg, false);  // pre-sized: Eval accumulates into it
  const sctl::ConstIterator<char> self = (sctl::ConstIterator<char>)(const char*)&ker;
  if (omp) K::template Eval<Real, true>(Vt, Xt, Xs, Xn, Vs, digits, self);
  else     K::template Eval<Real, false>(Vt, Xt, Xs, Xn, Vs, digits, self);
}

template <class K>
void kmat_one(int digits, const double* r_trg, long nt, const double* r_src, const double* n_src,
              long ns, double* M_out, const double* ctx) {
  K ker;
  double mu = ctx ? *ctx : 0.0;
  if (ctx) ker.SetCtxPtr(&mu);
  const Long nd = K::NormalDim();
  Vec<double> Xt(nt * 3, (sctl::Iterator<double>)r_trg, false);
  Vec<double> Xs(ns * 3, (sctl::Iterator<double>)r_src, false);
  Vec<double> Xn(nd ? ns * nd : 0, (sctl::Iterator<double>)(nd ? n_src : nullptr), false);
  sctl::Matrix<double> M;
  switch (digits) {
    case 3:  ker.template KernelMatrix<double, false, 3>(M, Xt, Xs, Xn); break;
    case 7:  ker.template KernelMatrix<double, false, 7>(M, Xt, Xs, Xn); break;
    case 11: ker.template KernelMatrix<double, false, 11>(M, Xt, Xs, Xn); break;
    default: ker.template KernelMatrix<double, false, -1>(M, Xt, Xs, Xn); break;
  }
  const long n = (long)M.Dim(0) * (long)M.Dim(1);
  if (n) std::memcpy(M_out, &M[0][0], sizeof(double) * n);
}

template <Integer digits> void rsqrt_digits(const double* x, long n, double* y) {
  constexpr Integer VL = sctl::DefaultVecLen<double>();
  using V = sctl::Vec<double, VL>;
  for (long i = 0; i < n; i += VL) {
    alignas(64) double buf[VL], out[VL];
    for (Integer k = 0; k < VL; k++) buf[k] = (i + k < n ? x[i + k] : 1.0);
    V xv = V::LoadAligned(buf);
    V yv = sctl::approx_rsqrt<digits>(xv, xv > V::Zero());
    yv.StoreAligned(out);
    for (Integer k = 0; k < VL && i + k < n; k++) y[i + k] = out[k];
  }
}

// Calls f(GenericKernel object, kernel_impl tag) for a kernel id:
//   0 Laplace3D_FxU  1 Laplace3D_DxU  2 Stokes3D_FxU  3 Stokes3D_DxU  4 Laplace3D_FxdU
//   5 Laplace3D_DxdU 6 Laplace3D_Fxd2U 7 Stokes3D_FxT 8 Stokes3D_FSxU 9 Stokes3D_FxUP
//   10 BiotSavart3D_FxU 11 BiotSavart3D_FxdU 12 Helmholtz3D_FxU 13 Helmholtz3D_DxU
//   14 Helmholtz3D_FxdU 15 HelmholtzDiff3D_FxdU    (kernel_functions.hpp:669-684)
template <class F> int with_kernel(int kernel, F&& f) {
  namespace ki = sctl::kernel_impl;
  switch (kernel) {
    case 0: f(sctl::Laplace3D_FxU(), ki::Laplace3D_FxU()); return 0;
    case 1: f(sctl::Laplace3D_DxU(), ki::Laplace3D_DxU()); return 0;
    case 2: f(sctl::Stokes3D_FxU(), ki::Stokes3D_FxU()); return 0;
    case 3: f(sctl::Stokes3D_DxU(), ki::Stokes3D_DxU()); return 0;
    case 4: f(sctl::Laplace3D_FxdU(), ki::Laplace3D_FxdU()); return 0;
    case 5: f(sctl::Laplace3D_DxdU(), ki::Laplace3D_DxdU()); return 0;
    case 6: f(sctl::Laplace3D_Fxd2U(), ki::Laplace3D_Fxd2U()); return 0;
    case 7: f(sctl::Stokes3D_FxT(), ki::Stokes3D_FxT()); return 0;
    case 8: f(sctl::Stokes3D_FSxU(), ki::Stokes3D_FSxU()); return 0;
    case 9: f(sctl::Stokes3D_FxUP(), ki::Stokes3D_FxUP()); return 0;
    case 10: f(sctl::BiotSavart3D_FxU(), ki::BiotSavart3D_FxU()); return 0;
    case 11: f(sctl::BiotSavart3D_FxdU(), ki::BiotSavart3D_FxdU()); return 0;
    case 12: f(sctl::Helmholtz3D_FxU(), ki::Helmholtz3D_FxU()); return 0;
    case 13: f(sctl::Helmholtz3D_DxU(), ki::Helmholtz3D_DxU()); return 0;
    case 14: f(sctl::Helmholtz3D_FxdU(), ki::Helmholtz3D_FxdU()); return 0;
    case 15: f(sctl::HelmholtzDiff3D_FxdU(), ki::HelmholtzDiff3D_FxdU()); return 0;
  }
  return -1;
}

}  // namespace

extern "C" {

int sctlref_info(int kernel, int* dims, double* scale, char* name, int name_len) {
  return with_kernel(kernel, [&](auto ker, auto impl) {
    using K = decltype(ker);
    dims[0] = K::CoordDim(); dims[1] = K::SrcDim(); dims[2] = K::TrgDim(); dims[3] = K::NormalDim();
    dims[4] = (int)K::FLOPS();
    *scale = decltype(impl)::template uKerScaleFactor<double>();
    std::strncpy(name, ker.Name().c_str(), name_len);
  });
}

int sctlref_veclen(void) { return (int)sctl::DefaultVecLen<double>(); }

// ctx: the wavenumber for the Helmholtz kernels (required by them), else NULL.
int sctlref_eval(int kernel, int digits, int omp, const double* r_trg, long nt, const double* r_src,
                 const double* n_src, const double* v_src, long ns, double* v_trg, const double* ctx) {
  if (kernel >= 12 && !ctx) return -1;
  return with_kernel(kernel, [&](auto ker, auto) {
    eval_one<decltype(ker), double>(omp, digits, r_trg, nt, r_src, n_src, v_src, ns, v_trg, ctx);
  });
}

int sctlref_eval_f32(int kernel, int digits, int omp, const float* r_trg, long nt, const float* r_src,
                     const float* n_src, const float* v_src, long ns, float* v_trg, const double* ctx) {
  if (kernel >= 12 && !ctx) return -1;
  return with_kernel(kernel, [&](auto ker, auto) {
    eval_one<decltype(ker), float>(omp, digits, r_trg, nt, r_src, n_src, v_src, ns, v_trg, ctx);
  });
}

int sctlref_kernel_matrix(int kernel, int digits, const double* r_trg, long nt, const double* r_src,
                          const double* n_src, long ns, double* M, const double* ctx) {
  if (kernel >= 12 && !ctx) return -1;
  return with_kernel(kernel, [&](auto ker, auto) {
    kmat_one<decltype(ker)>(digits, r_trg, nt, r_src, n_src, ns, M, ctx);
  });
}

// BoundaryIntegralOp's per-block loop (boundary_integral.txx:1165-1211):
// `#pragma omp parallel for schedule(dynamic)` over the blocks, each block
// (targets [trg_dsp, +trg_cnt), sources [src_dsp, +src_cnt)) evaluated by
// GenericKernel::Eval<double,false> into v_trg (mode 0), or by
// KernelMatrix<double,false> into a per-block matrix whose entries are summed
// into *checksum (mode 1; the matrices themselves are not kept).
int sctlref_blocks(int kernel, int mode, long nblocks, const long* trg_dsp, const long* trg_cnt, const long* src_dsp,
                   const long* src_cnt, const double* r_trg, const double* r_src, const double* n_src,
                   const double* v_src, double* v_trg, double* checksum) {
  if (kernel >= 12 || mode < 0 || mode > 1) return -1;
  return with_kernel(kernel, [&](auto ker, auto) {
    using K = decltype(ker);
    const Long K0 = K::SrcDim(), K1 = K::TrgDim(), nd = K::NormalDim();
    double sum = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : sum)
    for (long b = 0; b < nblocks; b++) {
      const long nt = trg_cnt[b], ns = src_cnt[b];
      if (!nt || !ns) continue;
      Vec<double> Xt(nt * 3, (sctl::Iterator<double>)(r_trg + trg_dsp[b] * 3), false);
      Vec<double> Xs(ns * 3, (sctl::Iterator<double>)(r_src + src_dsp[b] * 3), false);
      Vec<double> Xn(nd ? ns * 3 : 0, (sctl::Iterator<double>)(nd ? n_src + src_dsp[b] * 3 : nullptr), false);
      if (mode == 0) {
        Vec<double> Vs(ns * K0, (sctl::Iterator<double>)(v_src + src_dsp[b] * K0), false);
        Vec<double> Vt(nt * K1, v_trg + trg_dsp[b] * K1, false);
        ker.template Eval<double, false>(Vt, Xt, Xs, Xn, Vs);
      } else {
        sctl::Matrix<double> M;
        ker.template KernelMatrix<double, false>(M, Xt, Xs, Xn);
        sum += M[0][0] + M[M.Dim(0) - 1][M.Dim(1) - 1];
      }
    }
    if (checksum) *checksum = sum;
  });
}

int sctlref_rsqrt(int digits, const double* x, long n, double* y) {
  switch (digits) {
    case 4: rsqrt_digits<4>(x, n, y); return 0;
    case 7: rsqrt_digits<7>(x, n, y); return 0;
    case 11: rsqrt_digits<11>(x, n, y); return 0;
    case -1: rsqrt_digits<-1>(x, n, y); return 0;
  }
  return -1;
}

}  // extern "C"
