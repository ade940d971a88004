// sctl_amd::NearFieldOp built from element lists with the ElementListBase
// interface (boundary_integral.hpp:65-240): GetNodeCoord, GetFarFieldNodes,
// MatrixFree.  Two lists of torus patches (one with near matrices, one
// matrix-free), targets near the surface, Laplace DxU and Stokes DxU.
// Checks, against sums worked out here in long double from the lists'
// own nodes:
//   * the near list: a target is near an element iff some node s of the
//     element has |x_t - x_s| < dist_far[s] (BuildNearList's criterion,
//     boundary_integral.txx:390-401);
//   * KNearSubtract on a zero K_near followed by NearInterac (GEMV,
//     ScatterForward, per-target sum): U[t] = -sum over the matrix elements
//     near t of sum_s K(x_t - x_s) w_s F_s;
//   * NearDirect: every near (element, target) row = sum_s K w_s F_s, the
//     matrix-free elements included.
// All to 1e-12 relative to the largest entry.  Runs on the GPU.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <vector>

#include "sctl_amd/containers.hpp"
#include "sctl_amd/near_field.hpp"

using sctl_amd::Long;
using sctl_amd::Vector;

static int failures = 0;
#define CHECK(cond, ...)                   \
  do {                                     \
    if (!(cond)) {                         \
      std::printf("  FAIL: " __VA_ARGS__); \
      std::printf("\n");                   \
      ++failures;                          \
    }                                      \
  } while (0)

// Torus patches (R = 1, r = 0.3), p x p nodes per patch at cell midpoints,
// outward normals, area weights; every node's far-field radius `rad`.
class PatchList {
 public:
  PatchList(int nu, int nv, int u0, int nu_all, int p, double rad, bool matrix_free)
      : nu_(nu), nv_(nv), u0_(u0), nu_all_(nu_all), p_(p), rad_(rad), mf_(matrix_free) {}

  Long Size() const { return (Long)nu_ * nv_; }

  void GetNodeCoord(Vector<double>* X, Vector<double>* Xn, Vector<Long>* cnt) const {
    Vector<double> w, d;
    Vector<Long> c;
    nodes(X, Xn, &w, &d, &c);
    if (cnt) *cnt = c;
  }
  void GetFarFieldNodes(Vector<double>& X, Vector<double>& Xn, Vector<double>& wts, Vector<double>& dist_far,
                        Vector<Long>& cnt, const double tol) const {
    (void)tol;  // the far-field quadrature is the element's own (identity FarFieldDensityOperatorTranspose)
    nodes(&X, &Xn, &wts, &dist_far, &cnt);
  }
  bool MatrixFree() const { return mf_; }

 private:
  void nodes(Vector<double>* X, Vector<double>* Xn, Vector<double>* W, Vector<double>* D, Vector<Long>* C) const {
    const double R = 1.0, r = 0.3, pi = 3.14159265358979323846;
    const double du = 2 * pi / nu_all_, dv = 2 * pi / nv_;
    const Long n = Size() * p_ * p_;
    if (X) X->ReInit(3 * n);
    if (Xn) Xn->ReInit(3 * n);
    if (W) W->ReInit(n);
    if (D) D->ReInit(n);
    if (C) C->ReInit(Size());
    Long k = 0;
    for (int iu = 0; iu < nu_; iu++)
      for (int iv = 0; iv < nv_; iv++) {
        if (C) (*C)[iu * nv_ + iv] = (Long)p_ * p_;
        for (int a = 0; a < p_; a++)
          for (int b = 0; b < p_; b++, k++) {
            const double u = (u0_ + iu + (a + 0.5) / p_) * du, v = (iv + (b + 0.5) / p_) * dv;
            const double nx = std::cos(v) * std::cos(u), ny = std::cos(v) * std::sin(u), nz = std::sin(v);
            const double x[3] = {(R + r * std::cos(v)) * std::cos(u), (R + r * std::cos(v)) * std::sin(u),
                                 r * std::sin(v)};
            for (int q = 0; q < 3; q++) {
              if (X) (*X)[3 * k + q] = x[q];
            }
            if (Xn) {
              (*Xn)[3 * k + 0] = nx;
              (*Xn)[3 * k + 1] = ny;
              (*Xn)[3 * k + 2] = nz;
            }
            if (W) (*W)[k] = r * (R + r * std::cos(v)) * du * dv / (p_ * p_);
            if (D) (*D)[k] = rad_;
          }
      }
  }
  int nu_, nv_, u0_, nu_all_, p_;
  double rad_;
  bool mf_;
};

struct Nodes {
  Vector<double> X, Xn, W, D;
  Vector<Long> C;
};

// Laplace DxU / Stokes DxU pair K(r, n) f, long double, unscaled by the
// kernel's scale factor (applied by the caller).
static void pair(int ker, const long double* r, const long double* n, const long double* f, long double* out) {
  const long double r2 = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
  if (r2 == 0) return;
  const long double ir = 1 / std::sqrt(r2);
  const long double rn = r[0] * n[0] + r[1] * n[1] + r[2] * n[2];
  if (ker == SCTL_AMD_LAPLACE3D_DXU) {
    out[0] += rn * ir * ir * ir * f[0];
  } else {  // Stokes DxU: r (r.f)(r.n) / r^5
    const long double rf = r[0] * f[0] + r[1] * f[1] + r[2] * f[2];
    const long double c = rf * rn * ir * ir * ir * ir * ir;
    for (int k = 0; k < 3; k++) out[k] += r[k] * c;
  }
}

template <class K> static void check(const char* name) {
  const int ker = K::KERNEL_ID;
  const Long K0 = K::KDIM0, K1 = K::KDIM1;
  const double scale = K::template uKerScaleFactor<double>();
  const int NU = 16, NV = 12, P = 6;
  const double rad = 0.06;
  PatchList A(NU / 2, NV, 0, NU, P, rad, false), B(NU / 2, NV, NU / 2, NU, P, rad, true);
  Nodes na, nb;
  A.GetFarFieldNodes(na.X, na.Xn, na.W, na.D, na.C, 1e-10);
  B.GetFarFieldNodes(nb.X, nb.Xn, nb.W, nb.D, nb.C, 1e-10);
  const Long nelemA = A.Size(), nelem = A.Size() + B.Size();
  const Long nper = (Long)P * P;
  auto node = [&](Long e, Long j, int q) { return e < nelemA ? na.X[3 * (e * nper + j) + q] : nb.X[3 * ((e - nelemA) * nper + j) + q]; };
  auto nrm = [&](Long e, Long j, int q) { return e < nelemA ? na.Xn[3 * (e * nper + j) + q] : nb.Xn[3 * ((e - nelemA) * nper + j) + q]; };
  auto wt = [&](Long e, Long j) { return e < nelemA ? na.W[e * nper + j] : nb.W[(e - nelemA) * nper + j]; };

  // targets: points near the torus surface
  const Long nt = 1500;
  Vector<double> Xt(3 * nt), Xn_empty;
  uint64_t st = 12345;
  auto uni = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return (double)(st >> 11) * (1.0 / 9007199254740992.0);
  };
  for (Long t = 0; t < nt; t++) {
    const double pi = 3.14159265358979323846, u = 2 * pi * uni(), v = 2 * pi * uni(), h = 0.1 * (uni() - 0.5);
    const double rr = 0.3 + h;
    Xt[3 * t + 0] = (1 + rr * std::cos(v)) * std::cos(u);
    Xt[3 * t + 1] = (1 + rr * std::cos(v)) * std::sin(u);
    Xt[3 * t + 2] = rr * std::sin(v);
  }
  // densities F on all element nodes (K0 per node, lists in order)
  const Long ns = nelem * nper;
  std::vector<double> F(ns * K0);
  for (auto& x : F) x = 2 * uni() - 1;

  sctl_amd::NearFieldOp<double, sctl_amd::GenericKernel<K>, Vector> op;
  op.AddElemList(A);
  op.AddElemList(B);
  op.SetTargets(Xt, Xn_empty);
  op.Setup(1e-10, 4);

  // brute-force near list: element e near target t
  std::vector<std::vector<Long>> near_t(nt);
  Long pairs = 0;
  for (Long t = 0; t < nt; t++)
    for (Long e = 0; e < nelem; e++) {
      bool is_near = false;
      for (Long j = 0; j < nper && !is_near; j++) {
        double d2 = 0;
        for (int q = 0; q < 3; q++) d2 += (Xt[3 * t + q] - node(e, j, q)) * (Xt[3 * t + q] - node(e, j, q));
        is_near = d2 < rad * rad;
      }
      if (is_near) near_t[t].push_back(e), pairs++;
    }
  const auto& ecnt = op.NearElemCnt();
  const auto& edsp = op.NearElemDsp();
  Long total = 0;
  for (Long e = 0; e < nelem; e++) total += ecnt[e];
  CHECK(total == pairs, "%s: near pairs %lld, expected %lld", name, (long long)total, (long long)pairs);

  // expected sums, long double
  auto elem_sum = [&](const double* x, Long e, long double* out) {
    for (Long k = 0; k < K1; k++) out[k] = 0;
    for (Long j = 0; j < nper; j++) {
      long double r[3], n[3], f[3] = {0, 0, 0};
      for (int q = 0; q < 3; q++) {
        r[q] = (long double)x[q] - node(e, j, q);
        n[q] = nrm(e, j, q);
      }
      for (Long k = 0; k < K0; k++) f[k] = (long double)wt(e, j) * F[(e * nper + j) * K0 + k];
      pair(ker, r, n, f, out);
    }
    for (Long k = 0; k < K1; k++) out[k] *= scale;
  };

  // KNearSubtract on zero K_near, then NearInterac
  std::vector<double> Kn(op.KNearLength(), 0.0), U;
  op.KNearSubtract(Kn);
  op.NearInterac(U, Kn, F);
  double err = 0, mag = 0;
  for (Long t = 0; t < nt; t++) {
    long double want[3] = {0, 0, 0};
    for (Long e : near_t[t]) {
      if (e >= nelemA) continue;  // matrix-free list: no K_near
      long double s[3];
      elem_sum(&Xt[3 * t], e, s);
      for (Long k = 0; k < K1; k++) want[k] -= s[k];
    }
    for (Long k = 0; k < K1; k++) {
      err = std::fmax(err, std::fabs((double)(U[t * K1 + k] - want[k])));
      mag = std::fmax(mag, std::fabs((double)want[k]));
    }
  }
  std::printf("%-14s near pairs %lld  K_near %lld entries  KNearSubtract + NearInterac rel err %.2e\n", name,
              (long long)pairs, (long long)op.KNearLength(), err / mag);
  CHECK(mag > 0 && err <= 1e-12 * mag, "%s: KNearSubtract/NearInterac error %.3e of %.3e", name, err, mag);

  // NearDirect: every near row, matrix-free elements included
  std::vector<double> Un;
  op.NearDirect(Un, F);
  const auto& Xnear = op.XtrgNear();
  err = mag = 0;
  for (Long e = 0; e < nelem; e++)
    for (Long i = 0; i < ecnt[e]; i++) {
      const Long row = edsp[e] + i;
      long double s[3];
      elem_sum(&Xnear[3 * row], e, s);
      for (Long k = 0; k < K1; k++) {
        err = std::fmax(err, std::fabs((double)(Un[row * K1 + k] - s[k])));
        mag = std::fmax(mag, std::fabs((double)s[k]));
      }
    }
  std::printf("%-14s NearDirect rel err %.2e\n", name, err / mag);
  CHECK(mag > 0 && err <= 1e-12 * mag, "%s: NearDirect error %.3e of %.3e", name, err, mag);
}

int main() {
  check<sctl_amd::kernel_impl::Laplace3D_DxU>("Laplace3D-DxU");
  check<sctl_amd::kernel_impl::Stokes3D_DxU>("Stokes3D-DxU");
  std::printf("\n%s (%d check(s) failed)\n", failures == 0 ? "PASS" : "FAIL", failures);
  return failures == 0 ? 0 : 1;
}
