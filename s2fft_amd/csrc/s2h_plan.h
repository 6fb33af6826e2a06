// Host-side plan: constant tables for one (sampling, L, nside, spin, ...) set.
#pragma once

#include <vector>

#include "s2h_kernels.h"

struct s2h_plan {
  s2h_desc d{};
  int device = 0;
  int m_lo = 0;      // first m computed (0 with reality, else -(L-1))
  int nm = 0;        // number of m computed
  int Mc = 0;        // columns of the internal ftm (m in [m_lo, L))
  double sign = 1;   // global sign folded into the recursion

  // forward
  int fwd_kind = 0;  // 0 direct ring FFT, 1 MW resample, 2 MWSS resample
  int fwd_rows = 0;  // rows of the internal forward ftm
  int fwd_in_rows = 0;  // rows of the phi-spectrum before resampling (kinds 1, 2)
  s2h::JobSet fwd_set;
  s2h::LatRing* fwd_rings = nullptr;
  int n_fwd_rings = 0;
  bool fwd_paired = false;
  int fwd_row_south = -1, fwd_row_north = -1;
  s2h::MStart* fwd_mstart = nullptr;
  s2h::Coef* fwd_coef = nullptr;
  s2h::Coef* fwd_ucoef = nullptr;  // unit-c form (make_unit_coefs)
  double* fwd_norm = nullptr;
  double* fwd_weights = nullptr;  // MWSS-at-2L weights (kinds 1, 2)
  int rs_info[8] = {0};           // resampling DFT lengths / conv lengths
  int64_t rs_chirp[4] = {0};

  // inverse
  int inv_rows = 0;
  s2h::JobSet inv_set;
  s2h::LatRing* inv_rings = nullptr;
  int n_inv_rings = 0;
  bool inv_paired = false;
  int inv_row_south = -1, inv_row_north = -1;
  s2h::MStart* inv_mstart = nullptr;
  s2h::Coef* inv_coef = nullptr;
  s2h::Coef* inv_ucoef = nullptr;
  double* inv_norm = nullptr;
  bool inv_out_complex = true;

  s2h::cplx* chirp = nullptr;
  int64_t npix = 0;   // pixels per map (forward input / inverse output)
  std::vector<void*> allocations;
};

namespace s2h {
// Ring-job tables for the reference-layout HEALPix FFTs (cached per device).
struct HpTables {
  JobSet set;
  cplx* chirp = nullptr;
};
// Builds and caches the tables of (current device, nside, L) once
// (allocates, synchronises); healpix_tables only looks them up and returns
// S2H_ERR_INVALID for tables that were never prepared.
int healpix_tables_prepare(int nside, int L);
int healpix_tables(int nside, int L, HpTables* out);
}  // namespace s2h
