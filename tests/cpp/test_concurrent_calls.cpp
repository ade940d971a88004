// Concurrent calls of the C++ drop-in from OpenMP threads, the way
// BoundaryIntegralOp makes them: KNearSubtract runs
// `#pragma omp parallel for schedule(dynamic)` over near blocks and calls
// `ker_.template KernelMatrix<Real,false>(...)` inside the loop
// (boundary_integral.txx:1165-1211); ParticleFMM's direct path calls the
// static Eval of each kernel the same way from caller threads.
//
// Each block (one element's far nodes x up to `grain` of its near targets,
// as the partition of boundary_integral.txx:1009-1032 cuts them) is
// evaluated by KernelMatrix<double,false>, Eval<double,false> (FP64) and
// Eval<double,false,6> (mixed precision) on 8 OpenMP threads at once, three
// times over, and once more on one thread.  Checks:
//   * every concurrent result is bit-identical to the single-thread result;
//   * FP64 KernelMatrix and Eval agree with the oracle (exact mode) to 1e-13
//     (max |diff| / max |ref| per block); mixed Eval to 1e-6 of
//     sum_s |K| |f| per target.
// Runs on the GPU; links the oracle as the checker.
#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <vector>

#include "sctl_amd/containers.hpp"
#include "sctl_amd/generic_kernel.hpp"

extern "C" {
int oracle_eval(int ker, int mode, int digits, const double* r_trg, long nt, const double* r_src,
                const double* n_src, const double* v_src, long ns, double* v_trg, int nthreads);
int oracle_kernel_matrix(int ker, const double* r_trg, long nt, const double* r_src, const double* n_src, long ns,
                         double* M, int nthreads);
}

using sctl_amd::Long;
using sctl_amd::Matrix;
using sctl_amd::Vector;

static int failures = 0;
#define CHECK(cond, ...)                  \
  do {                                    \
    if (!(cond)) {                        \
      std::printf("  FAIL: " __VA_ARGS__); \
      std::printf("\n");                  \
      ++failures;                         \
    }                                     \
  } while (0)

// SplitMix64 (Steele, Lea and Flood 2014) for the synthetic geometry.
struct Rng {
  uint64_t s;
  uint64_t next() {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
  }
  double uni() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct Block {
  Long elem, t0, cnt;
};

struct Surface {
  std::vector<Vector<double>> X, Xn, F;  // per element: far nodes, normals, densities (K0 = 3 storage)
  std::vector<Vector<double>> Xt;        // per element: near targets
  std::vector<Block> blocks;
};

static Surface make_surface(int nelem, Long grain, uint64_t seed) {
  Rng r{seed};
  Surface S;
  for (int e = 0; e < nelem; e++) {
    const Long ns = 16 + (Long)(r.uni() * 290), nt = 4 + (Long)(r.uni() * 150);
    const double cx = 4 * r.uni(), cy = 4 * r.uni(), cz = 4 * r.uni();
    Vector<double> X(ns * 3), Xn(ns * 3), F(ns * 3), Xt(nt * 3);
    for (Long s = 0; s < ns; s++) {
      X[s * 3 + 0] = cx + 0.2 * r.uni();
      X[s * 3 + 1] = cy + 0.2 * r.uni();
      X[s * 3 + 2] = cz + 0.01 * r.uni();
      double n[3] = {r.uni() - 0.5, r.uni() - 0.5, 1.0}, l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
      for (int k = 0; k < 3; k++) Xn[s * 3 + k] = n[k] / l;
      for (int k = 0; k < 3; k++) F[s * 3 + k] = 2 * r.uni() - 1;
    }
    for (Long t = 0; t < nt; t++) {
      Xt[t * 3 + 0] = cx + 0.2 * r.uni();
      Xt[t * 3 + 1] = cy + 0.2 * r.uni();
      Xt[t * 3 + 2] = cz + 0.005 + 0.05 * (r.uni() - 0.5);
    }
    S.X.push_back(X);
    S.Xn.push_back(Xn);
    S.F.push_back(F);
    S.Xt.push_back(Xt);
    for (Long t0 = 0; t0 < nt; t0 += grain) S.blocks.push_back({e, t0, std::min(grain, nt - t0)});
  }
  return S;
}

struct Result {
  std::vector<Matrix<double>> M;
  std::vector<Vector<double>> U, Um;
};

template <class K> static void run_block(const K& ker, const Surface& S, const Block& b, Result& R, size_t i) {
  const Long K0 = K::SrcDim(), ND = K::NormalDim();
  const Vector<double>& X = S.X[b.elem];
  const Long ns = X.Dim() / 3;
  Vector<double> Xt(b.cnt * 3);
  for (Long j = 0; j < b.cnt * 3; j++) Xt[j] = S.Xt[b.elem][b.t0 * 3 + j];
  Vector<double> Xn(ND ? ns * 3 : 0), F(ns * K0);
  for (Long j = 0; j < Xn.Dim(); j++) Xn[j] = S.Xn[b.elem][j];
  for (Long s = 0; s < ns; s++)
    for (Long k = 0; k < K0; k++) F[s * K0 + k] = S.F[b.elem][s * 3 + k];
  ker.template KernelMatrix<double, false>(R.M[i], Xt, X, Xn);
  ker.template Eval<double, false>(R.U[i], Xt, X, Xn, F);
  ker.template Eval<double, false, 6>(R.Um[i], Xt, X, Xn, F);
}

template <class V> static bool same_bits(const V& a, const V& b, Long n) {
  return std::memcmp(&a[0], &b[0], sizeof(double) * n) == 0;
}

template <class K> static void check_kernel(int id, const Surface& S) {
  K ker;
  const Long K0 = K::SrcDim(), K1 = K::TrgDim(), ND = K::NormalDim();
  const size_t nb = S.blocks.size();
  Result serial{std::vector<Matrix<double>>(nb), std::vector<Vector<double>>(nb), std::vector<Vector<double>>(nb)};
  for (size_t i = 0; i < nb; i++) run_block(ker, S, S.blocks[i], serial, i);
  int mismatches = 0;
  for (int rep = 0; rep < 3; rep++) {
    Result conc{std::vector<Matrix<double>>(nb), std::vector<Vector<double>>(nb), std::vector<Vector<double>>(nb)};
#pragma omp parallel for schedule(dynamic) num_threads(8)
    for (Long i = 0; i < (Long)nb; i++) run_block(ker, S, S.blocks[i], conc, (size_t)i);
    for (size_t i = 0; i < nb; i++) {
      const Block& b = S.blocks[i];
      const Long ns = S.X[b.elem].Dim() / 3;
      const Long nM = ns * K0 * b.cnt * K1;
      bool ok = conc.M[i].Dim(0) == ns * K0 && conc.M[i].Dim(1) == b.cnt * K1 && same_bits(conc.M[i].begin(), serial.M[i].begin(), nM) &&
                conc.U[i].Dim() == b.cnt * K1 && same_bits(conc.U[i], serial.U[i], b.cnt * K1) &&
                conc.Um[i].Dim() == b.cnt * K1 && same_bits(conc.Um[i], serial.Um[i], b.cnt * K1);
      if (!ok) mismatches++;
    }
  }
  CHECK(mismatches == 0, "%s: %d concurrent block result(s) differ from the single-thread run", K::Name().c_str(),
        mismatches);

  // the oracle, per block
  double worst_m = 0, worst_u = 0, worst_mixed = 0;
  for (size_t i = 0; i < nb; i++) {
    const Block& b = S.blocks[i];
    const Long ns = S.X[b.elem].Dim() / 3;
    const double* xt = &S.Xt[b.elem][b.t0 * 3];
    std::vector<double> xn(ND ? ns * 3 : 0), f(ns * K0);
    for (Long j = 0; j < (Long)xn.size(); j++) xn[j] = S.Xn[b.elem][j];
    for (Long s = 0; s < ns; s++)
      for (Long k = 0; k < K0; k++) f[s * K0 + k] = S.F[b.elem][s * 3 + k];
    const Long nM = ns * K0 * b.cnt * K1;
    std::vector<double> M(nM), u(b.cnt * K1, 0.0);
    oracle_kernel_matrix(id, xt, b.cnt, &S.X[b.elem][0], ND ? xn.data() : nullptr, ns, M.data(), 1);
    oracle_eval(id, 0, -1, xt, b.cnt, &S.X[b.elem][0], ND ? xn.data() : nullptr, f.data(), ns, u.data(), 1);
    double dm = 0, am = 0, du = 0, au = 0;
    for (Long j = 0; j < nM; j++) {
      dm = std::fmax(dm, std::fabs(serial.M[i].begin()[j] - M[j]));
      am = std::fmax(am, std::fabs(M[j]));
    }
    for (Long j = 0; j < b.cnt * K1; j++) {
      du = std::fmax(du, std::fabs(serial.U[i][j] - u[j]));
      au = std::fmax(au, std::fabs(u[j]));
    }
    worst_m = std::fmax(worst_m, dm / am);
    worst_u = std::fmax(worst_u, du / au);
    // mixed: per target output, against sum_s |K| |f| from the oracle's matrix
    for (Long t = 0; t < b.cnt; t++)
      for (Long k1 = 0; k1 < K1; k1++) {
        double mag = 0;
        for (Long s = 0; s < ns * K0; s++) mag += std::fabs(M[s * b.cnt * K1 + t * K1 + k1] * f[s]);
        worst_mixed = std::fmax(worst_mixed, std::fabs(serial.Um[i][t * K1 + k1] - u[t * K1 + k1]) / mag);
      }
  }
  std::printf("%-14s blocks %zu  KernelMatrix vs oracle %.2e  Eval %.2e  mixed Eval (of sum |K||f|) %.2e\n",
              K::Name().c_str(), nb, worst_m, worst_u, worst_mixed);
  CHECK(worst_m <= 1e-13, "%s KernelMatrix error %.3e", K::Name().c_str(), worst_m);
  CHECK(worst_u <= 1e-13, "%s Eval error %.3e", K::Name().c_str(), worst_u);
  CHECK(worst_mixed <= 1e-6, "%s mixed Eval error %.3e", K::Name().c_str(), worst_mixed);
}

int main() {
  const Surface S = make_surface(48, 40, 7);
  std::printf("OpenMP threads: 8 (of %d available)\n", omp_get_num_procs());
  check_kernel<sctl_amd::Laplace3D_FxU>(SCTL_AMD_LAPLACE3D_FXU, S);
  check_kernel<sctl_amd::Laplace3D_DxU>(SCTL_AMD_LAPLACE3D_DXU, S);
  check_kernel<sctl_amd::Stokes3D_FxU>(SCTL_AMD_STOKES3D_FXU, S);
  check_kernel<sctl_amd::Stokes3D_DxU>(SCTL_AMD_STOKES3D_DXU, S);
  std::printf("\n%s (%d check(s) failed)\n", failures == 0 ? "PASS" : "FAIL", failures);
  return failures == 0 ? 0 : 1;
}
