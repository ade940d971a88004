// Compile-only check: the drop-in accepts SCTL's own containers and its
// static Eval has the type ParticleFMM stores as a function pointer
// (fmm-wrapper.txx:153-154).  Built against the reference headers where they
// are present; never linked or run.
#include "sctl/common.hpp"
#include "sctl/profile.hpp"  // the drop-in then feeds the profiler's FLOP counter
#include "sctl/iterator.hpp"
#include "sctl/iterator.txx"
#include "sctl/vector.hpp"
#include "sctl/vector.txx"
#include "sctl/matrix.hpp"
#include "sctl/matrix.txx"

#include "sctl_amd/generic_kernel.hpp"

template <class Real>
using EvalFn = void (*)(sctl::Vector<Real>& v_trg, const sctl::Vector<Real>& r_trg, const sctl::Vector<Real>& r_src,
                        const sctl::Vector<Real>& n_src, const sctl::Vector<Real>& v_src, sctl::Integer digits,
                        sctl::ConstIterator<char> self);

template <class KerS2T> void use(const KerS2T& ker) {
  EvalFn<double> fn = KerS2T::template Eval<double, false, sctl::Vector<double>>;
  EvalFn<double> fn_omp = KerS2T::template Eval<double, true, sctl::Vector<double>>;
  (void)fn;
  (void)fn_omp;
  sctl::Vector<double> vt, xt, xs, xn, vs;
  ker.template Eval<double, true>(vt, xt, xs, xn, vs);
  ker.template Eval<double, false, 7>(vt, xt, xs, xn, vs);
  sctl::Matrix<double> M;
  ker.template KernelMatrix<double, false>(M, xt, xs, xn);
  sctl::Vector<float> ft, fx, fs, fn_, fv;
  ker.template Eval<float, false>(ft, fx, fs, fn_, fv);
}

void instantiate() {
  use(sctl_amd::Laplace3D_FxU());
  use(sctl_amd::Laplace3D_DxU());
  use(sctl_amd::Stokes3D_FxU());
  use(sctl_amd::Stokes3D_DxU());
}
