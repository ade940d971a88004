// The C++ drop-in (include/sctl_amd/generic_kernel.hpp) exercised the way the
// reference's own tests exercise sctl::GenericKernel:
//   src/test-generic-kernel.cpp:38-103   accessors, Name, scale factor,
//                                        1x1 KernelMatrix, Eval at distances
//                                        1,2,3,5, superposition
//   src/test-kernel_functions.cpp:35-131 closed forms per kernel
//   src/test-kernel_functions.cpp:487-526 Eval vs KernelMatrix * v (Ns=7, Nt=5)
// plus the static type-erased Eval (the function pointer ParticleFMM stores),
// accumulation into a pre-sized v_trg, digits and Real = float.
// Runs on the GPU; prints PASS/FAIL like the reference's test-utils.hpp.
#include <cmath>
#include <cstdio>
#include <string>

#include "sctl_amd/containers.hpp"
#include "sctl_amd/generic_kernel.hpp"

using sctl_amd::Long;
using sctl_amd::Matrix;
using sctl_amd::Vector;

static int failures = 0;
#define CHECK(...)                                                   \
  do {                                                               \
    if (!(__VA_ARGS__)) {                                            \
      std::printf("  FAIL @%d: %s\n", __LINE__, #__VA_ARGS__);       \
      ++failures;                                                    \
    }                                                                \
  } while (0)

template <class T> static bool approx_eq(T a, T b, T tol) {
  const T d = std::fabs(a - b), m = std::fmax(std::fabs(a), std::fabs(b));
  return d <= tol || d <= tol * m;
}

template <class K> static void eval_vs_matrix(Long K0, Long K1, Long ND, double mu = 0.0) {
  const Long Ns = 7, Nt = 5;
  Vector<double> Xs(Ns * 3), Xn(Ns * ND), Vs(Ns * K0), Xt(Nt * 3);
  for (Long i = 0; i < Xs.Dim(); i++) Xs[i] = 0.1 * (i + 1) - 0.35;
  for (Long i = 0; i < Vs.Dim(); i++) Vs[i] = 0.3 * (i % 5) - 0.4;
  for (Long i = 0; i < Xt.Dim(); i++) Xt[i] = 1.7 + 0.2 * i;
  for (Long i = 0; i < Ns && ND; i++) {
    Xn[i * 3 + 0] = 0;
    Xn[i * 3 + 1] = 1;
    Xn[i * 3 + 2] = 0;
  }
  K ker;
  if (mu != 0.0) ker.SetCtxPtr(&mu);
  Vector<double> Vt;
  ker.template Eval<double, false>(Vt, Xt, Xs, Xn, Vs);
  Matrix<double> M;
  ker.template KernelMatrix<double, false>(M, Xt, Xs, Xn);
  CHECK(M.Dim(0) == Ns * K0 && M.Dim(1) == Nt * K1);
  for (Long t = 0; t < Nt; t++)
    for (Long k1 = 0; k1 < K1; k1++) {
      double acc = 0;
      for (Long s = 0; s < Ns; s++)
        for (Long k0 = 0; k0 < K0; k0++) acc += M[s * K0 + k0][t * K1 + k1] * Vs[s * K0 + k0];
      CHECK(approx_eq(Vt[t * K1 + k1], acc, 1e-10));
    }
}

int main() {
  const double pi = 3.14159265358979323846, tol = 1e-9;
  // ---- test-generic-kernel.cpp: Laplace3D_FxU
  {
    sctl_amd::Laplace3D_FxU K;
    CHECK(K.CoordDim() == 3 && K.SrcDim() == 1 && K.TrgDim() == 1 && K.NormalDim() == 0);
    CHECK(K.Name() == std::string("Laplace3D-FxU"));
    const double sf = sctl_amd::kernel_impl::Laplace3D_FxU::uKerScaleFactor<double>();
    CHECK(approx_eq(sf, 1.0 / (4.0 * pi), 1e-12));
    Vector<double> Xt({1.0, 0.0, 0.0}), Xs({0.0, 0.0, 0.0}), Xn;
    Matrix<double> M;
    K.KernelMatrix<double, false>(M, Xt, Xs, Xn);
    CHECK(M.Dim(0) == 1 && M.Dim(1) == 1);
    CHECK(approx_eq(M(0, 0), 1.0 / (4.0 * pi), tol));
    Vector<double> v_s({1.0}), Xt4({1, 0, 0, 2, 0, 0, 3, 0, 0, 5, 0, 0}), v_t;
    K.Eval<double, false>(v_t, Xt4, Xs, Xn, v_s);
    CHECK(v_t.Dim() == 4);
    CHECK(approx_eq(v_t[0], sf / 1.0, tol) && approx_eq(v_t[1], sf / 2.0, tol));
    CHECK(approx_eq(v_t[2], sf / 3.0, tol) && approx_eq(v_t[3], sf / 5.0, tol));
    Vector<double> Xs2({0, 0, 0, 4, 0, 0}), v_s2({1.0, 1.0}), Xt1({2, 0, 0}), v_t1;
    K.Eval<double, false>(v_t1, Xt1, Xs2, Xn, v_s2);
    CHECK(v_t1.Dim() == 1 && approx_eq(v_t1[0], sf / 2.0 + sf / 2.0, tol));
    // accumulation into a pre-sized v_trg (generic-kernel.txx:129-132)
    K.Eval<double, false>(v_t1, Xt1, Xs2, Xn, v_s2);
    CHECK(approx_eq(v_t1[0], 2 * sf, tol));
    // the type-erased static Eval, as stored by ParticleFMM
    void (*fn)(Vector<double>&, const Vector<double>&, const Vector<double>&, const Vector<double>&,
               const Vector<double>&, sctl_amd::Integer, const char*) =
        sctl_amd::Laplace3D_FxU::Eval<double, true, Vector<double>>;
    Vector<double> v_t2;
    fn(v_t2, Xt4, Xs, Xn, v_s, 11, (const char*)&K);
    CHECK(v_t2.Dim() == 4 && approx_eq(v_t2[3], sf / 5.0, tol));
    // Real = float (mixed precision) and digits <= 7 (mixed precision)
    Vector<float> fXt({1, 0, 0, 2, 0, 0}), fXs({0, 0, 0}), fXn, fv({1.0f}), fo;
    K.Eval<float, false>(fo, fXt, fXs, fXn, fv);
    CHECK(fo.Dim() == 2 && approx_eq<float>(fo[1], (float)(sf / 2.0), 1e-6f));
    Vector<double> v_t3;
    K.Eval<double, false, 6>(v_t3, Xt4, Xs, Xn, v_s);
    CHECK(approx_eq(v_t3[1], sf / 2.0, 1e-6));
  }
  // ---- test-kernel_functions.cpp:35-131 closed forms
  {
    sctl_amd::Laplace3D_FxU K;
    Vector<double> Xs({0, 0, 0}), Xt({2, 0, 0}), Xn, v_s({3.0}), v_t;
    K.Eval<double, false>(v_t, Xt, Xs, Xn, v_s);
    CHECK(approx_eq(v_t[0], 3.0 / (4 * pi * 2.0), tol));
  }
  {
    sctl_amd::Laplace3D_DxU K;
    CHECK(K.NormalDim() == 3 && K.Name() == std::string("Laplace3D-DxU"));
    Vector<double> Xs({0, 0, 0}), Xt({2, 0, 0}), Xn({1, 0, 0}), v_s({1.0}), v_t;
    K.Eval<double, false>(v_t, Xt, Xs, Xn, v_s);
    CHECK(approx_eq(v_t[0], 1.0 / (16 * pi), tol));
  }
  {
    sctl_amd::Stokes3D_FxU K;
    CHECK(K.SrcDim() == 3 && K.TrgDim() == 3 && K.NormalDim() == 0 && K.Name() == std::string("Stokes3D-FxU"));
    Vector<double> Xs({0, 0, 0}), Xn, Xt({2.0, 0.0, 0.0}), v_s({1.0, 0.0, 0.0}), v_t;
    K.Eval<double, false>(v_t, Xt, Xs, Xn, v_s);
    CHECK(v_t.Dim() == 3 && approx_eq(v_t[0], 1.0 / (4 * pi * 2.0), tol));
    CHECK(std::fabs(v_t[1]) < 1e-12 && std::fabs(v_t[2]) < 1e-12);
  }
  {
    sctl_amd::Stokes3D_DxU K;
    CHECK(K.SrcDim() == 3 && K.TrgDim() == 3 && K.NormalDim() == 3 && K.Name() == std::string("Stokes3D-DxU"));
    Vector<double> Xs({0, 0, 0}), Xt({3, 0, 0}), Xn({0, 1, 0}), v_s({1.0, 0.0, 0.0}), v_t;
    K.Eval<double, false>(v_t, Xt, Xs, Xn, v_s);
    CHECK(v_t.Dim() == 3);
  }
  // ---- test-kernel_functions.cpp:487-526 Eval vs KernelMatrix
  eval_vs_matrix<sctl_amd::Laplace3D_FxU>(1, 1, 0);
  eval_vs_matrix<sctl_amd::Laplace3D_DxU>(1, 1, 3);
  eval_vs_matrix<sctl_amd::Stokes3D_FxU>(3, 3, 0);
  eval_vs_matrix<sctl_amd::Stokes3D_DxU>(3, 3, 3);
  eval_vs_matrix<sctl_amd::Laplace3D_FxdU>(1, 3, 0);
  eval_vs_matrix<sctl_amd::Laplace3D_DxdU>(1, 3, 3);
  eval_vs_matrix<sctl_amd::Laplace3D_Fxd2U>(1, 9, 0);
  eval_vs_matrix<sctl_amd::Stokes3D_FxT>(3, 9, 0);
  eval_vs_matrix<sctl_amd::Stokes3D_FSxU>(4, 3, 0);
  eval_vs_matrix<sctl_amd::Stokes3D_FxUP>(3, 4, 0);
  eval_vs_matrix<sctl_amd::BiotSavart3D_FxU>(3, 3, 0);
  eval_vs_matrix<sctl_amd::BiotSavart3D_FxdU>(3, 9, 0);
  eval_vs_matrix<sctl_amd::Helmholtz3D_FxU>(2, 2, 0, 2.5);
  eval_vs_matrix<sctl_amd::Helmholtz3D_DxU>(2, 2, 3, 2.5);
  eval_vs_matrix<sctl_amd::Helmholtz3D_FxdU>(2, 6, 0, 2.5);
  eval_vs_matrix<sctl_amd::HelmholtzDiff3D_FxdU>(2, 6, 0, 2.5);
  {  // Helmholtz3D_FxU of one unit source: e^{i mu r} / (4 pi r); through the static Eval and `self`
    sctl_amd::Helmholtz3D_FxU K;
    double mu = 1.3;
    K.SetCtxPtr(&mu);
    Vector<double> Xs({0, 0, 0}), Xt({2, 0, 0}), Xn, v_s({1.0, 0.0}), v_t;
    sctl_amd::Helmholtz3D_FxU::Eval<double, false, Vector<double>>(v_t, Xt, Xs, Xn, v_s, -1, (const char*)&K);
    CHECK(v_t.Dim() == 2 && approx_eq(v_t[0], std::cos(2 * mu) / (8 * pi), 1e-13) &&
          approx_eq(v_t[1], std::sin(2 * mu) / (8 * pi), 1e-13));
  }

  std::printf("\n%s (%d check(s) failed)\n", failures == 0 ? "PASS" : "FAIL", failures);
  return failures == 0 ? 0 : 1;
}
