// Test infrastructure only: a C-ABI wrapper around the reference library's own
// BuildNearList (boundary_integral.txx:50-520), compiled from the reference
// headers in place (see oracle/Makefile) into oracle/_ref/.  Used to write
// golden fixtures of the near-list builder; never shipped with or called by
// the product.
#include <cstring>

#include "sctl.hpp"

using sctl::Long;

extern "C" {

// Runs BuildNearList on one process.  Outputs are returned through buffers
// the caller sizes from a first call with max_near = 0 (only *n_near is set).
// Returns 0 on success, 1 if max_near is too small.
int sctlref_build_near_list(const double* Xtrg, long ntrg, const double* Xn_trg, const double* Xsrc, long nsrc,
                            const double* src_radius, long nelem, const long* elem_nds_cnt, const long* elem_nds_dsp,
                            long max_near, long* n_near, double* Xtrg_near, double* Xn_trg_near, long* near_elem_cnt,
                            long* near_elem_dsp, long* near_scatter_index, long* near_trg_cnt, long* near_trg_dsp) {
  sctl::Vector<double> Xt(ntrg * 3, (sctl::Iterator<double>)Xtrg, false);
  sctl::Vector<double> Xnt(Xn_trg ? ntrg * 3 : 0, (sctl::Iterator<double>)Xn_trg, false);
  sctl::Vector<double> Xs(nsrc * 3, (sctl::Iterator<double>)Xsrc, false);
  sctl::Vector<double> rad(nsrc, (sctl::Iterator<double>)src_radius, false);
  sctl::Vector<Long> ecnt(nelem), edsp(nelem);
  for (long i = 0; i < nelem; i++) {
    ecnt[i] = elem_nds_cnt[i];
    edsp[i] = elem_nds_dsp[i];
  }
  sctl::Vector<double> Xtn, Xntn;
  sctl::Vector<Long> nec, ned, nsi, ntc, ntd;
  sctl::BuildNearList<double, 3>(Xtn, Xntn, nec, ned, nsi, ntc, ntd, Xt, Xnt, Xs, rad, ecnt, edsp, sctl::Comm::Self());
  const long n = nsi.Dim();
  *n_near = n;
  if (max_near < n) return 1;
  for (long i = 0; i < n * 3; i++) Xtrg_near[i] = Xtn[i];
  if (Xn_trg_near) for (long i = 0; i < Xntn.Dim(); i++) Xn_trg_near[i] = Xntn[i];
  for (long i = 0; i < nelem; i++) {
    near_elem_cnt[i] = nec[i];
    near_elem_dsp[i] = ned[i];
  }
  for (long i = 0; i < n; i++) near_scatter_index[i] = nsi[i];
  for (long i = 0; i < ntrg; i++) {
    near_trg_cnt[i] = ntc[i];
    near_trg_dsp[i] = ntd[i];
  }
  return 0;
}

}  // extern "C"
