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

  // ring spectra of the C ABI: fwd_rows rounded up to a whole block of
  // kSpecBlock rows, and the forward ring of each of those rows (nphi 0 for
  // rows without one) for the tiled alias kernel
  int spec_rows = 0;
  s2h::RingJob* fwd_rowjobs = nullptr;
  // row blocks without aliasing (every ring n >= 2L - 1) and the others
  int* spec_belt_blocks = nullptr;
  int n_belt_blocks = 0;
  int* spec_polar_blocks = nullptr;
  int n_polar_blocks = 0;
  // the belt blocks' rows [spec_belt_row_lo, spec_belt_row_hi) when they are
  // contiguous and share one w n (HEALPix: spec_belt_wn), else spec_belt_wn 0
  double spec_belt_wn = 0.0;
  int spec_belt_row_lo = 0, spec_belt_row_hi = 0;

  // per-row tables of the forward grid (fwd_rows) and the inverse grid
  // (inv_rows), for the precompute kernels: colatitudes, the forward
  // quadrature weight of each forward row, the HEALPix first-pixel azimuth
  double* fwd_theta = nullptr;
  double* inv_theta = nullptr;
  double* fwd_row_w = nullptr;
  double* row_phi0 = nullptr;  // HEALPix only (forward and inverse grids agree)
  std::vector<void*> allocations;
};

namespace s2h {
// A multi-plane launch: map b of a batch is the n-plane with its own
// recursion tables (fwd[b] / inv[b], device arrays) and spin spin0 +
// spin_step b; every other table comes from the geometry plan, whose rings
// must be unpaired (a spin != 0 plan).  The Wigner transforms run all their
// n-planes as one batch this way.
struct PlaneSet {
  const LatPlane* fwd = nullptr;
  const LatPlane* inv = nullptr;
  int spin0 = 0;
  int spin_step = 0;
};

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

// Wigner (SO(3)) transform plan: one spin plan per n (spin -n, Wigner
// normalisation, sign (-1)^n), with their recursion tables gathered into
// per-plane arrays so that the n-planes run as one batch (s2h_capi.cpp).
struct s2h_wigner_plan {
  s2h_wigner_desc d{};
  int device = 0;
  int G = 0;                       // 2N - 1 gamma samples / n-planes
  int64_t npix = 0;                // (beta, alpha) samples per gamma
  std::vector<s2h_plan*> plans;    // plans[N - 1 + n], complex, n in [-(N-1), N)
  s2h_plan* real0 = nullptr;       // reality: the n = 0 plane's forward (spin 0, real)
  s2h_plan* geom = nullptr;        // an unpaired plan for multi-plane launches (N > 1)
  // per-plane tables, n = -(N-1) .. N-1 (index N - 1 + n)
  s2h::LatPlane* fwd_planes = nullptr;
  s2h::LatPlane* inv_planes = nullptr;
  std::vector<void*> allocations;
};
