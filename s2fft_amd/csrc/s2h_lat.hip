// Latitudinal (Wigner-d) steps: the ell-recursion per (ring, m) in registers,
// fused with the synthesis (inverse) or the projection (forward).
//
// Replaces s2fft/transforms/otf_recursions.py, which recurses over m at fixed
// (theta, ell) with log-space renormalisation (price_mcewen.py:118-255).  Here
// each lane owns one ring (or a ring pair theta, pi-theta for spin 0) and walks
//   d^{l+1} = (a_l x - b_l) d^l - c_l d^{l-1}
// from l0 = max(|m|, |spin|).  The coefficients (a, b, c) fold the
// normalisation sqrt((2l+1)/4pi) (or the Wigner one) into the recursion and
// are wave-uniform, so they arrive through scalar loads.  Underflow of the
// boundary value sin^a cos^b at high m is handled with an integer scale
// index: true value = stored * 2^(512 * lscale), lscale <= 0; a lane only
// contributes once lscale reaches 0 (earlier terms are below 2^-256 of the
// O(1) ones).
#include "s2h_fft_lds.h"
#include "s2h_kernels.h"
#include "s2h_lat_common.h"

namespace s2h {

namespace {

constexpr int kC = lat::kChunk;  // 32 steps per staged chunk (and per rescale check)
constexpr int kGroup = 4;        // steps whose coefficients are read from LDS together

using lat::m_of_block;
using lat::rescale;
using lat::start_value;

}  // namespace

// ------------------------------------------------------------------ inverse
// One lane per ring (or ring pair), NB maps per workgroup.  The recursion
// coefficients and the NB maps' flm of each 32-l chunk are loaded one chunk
// ahead (one element per thread) and staged in a double-buffered LDS array,
// one barrier per chunk; every step reads them with broadcast LDS loads, each
// kGroup-step group's coefficients while the previous group steps.
template <int NB, bool PAIRED>
__global__ void __launch_bounds__(kThreads) lat_inverse_kernel(LatArgs a) {
  static_assert(NB * kC <= kThreads, "one staged flm element per thread");
  __shared__ lat::ChunkCoef ccs[2][kC];
  __shared__ cplx fls[2][NB][kC];
  const int m = m_of_block(blockIdx.x, a.nm, a.m_lo);
  const int am = m < 0 ? -m : m;
  const int tid = threadIdx.x;
  const int p = blockIdx.y * kThreads + tid;
  const bool valid = p < a.nrings;
  const int b0 = blockIdx.z * NB;
  // multi-plane launches (NB = 1): this map's own recursion tables
  const LatPlane* pl = a.planes ? a.planes + b0 : nullptr;
  const MStart st = (pl ? pl->mstart : a.mstart)[m - a.m_lo];
  const int lmin = pl ? pl->lmin_out : a.lmin_out;
  const double bsign = m < 0 ? -1.0 : 1.0;
  const Coef* __restrict__ cf = (pl ? pl->coef : a.coef) + (int64_t)am * a.L;
  const int L = a.L;
  const int nb = min(NB, a.batch - b0);
  const int64_t frow = flm_row(a);
  const cplx* __restrict__ flm = a.flm + (int64_t)b0 * a.flm_map_stride + flm_col(a, m);

  LatRing rg = {0, -1, 0.0, 1.0, 0.0};
  double d = 0.0, dm1 = 0.0;
  int lscale = 0;
  if (valid) {
    rg = a.rings[p];
    start_value(st, rg.c2, rg.s2, d, lscale);
  }
  const double x = rg.x;

  // staging: thread t < kC loads the coefficient of l = blk + t; thread t <
  // NB kC loads flm of map t / kC at l = blk + t % kC (zero outside [lmin, L))
  const int fj = tid / kC, fi = tid % kC;
  auto load_f = [&](int blk) {
    cplx v = {0.0, 0.0};
    const int l = blk + fi;
    if (tid < NB * kC && fj < nb && l >= lmin && l < L)
      v = flm[(int64_t)fj * a.flm_map_stride + (int64_t)l * frow];
    return v;
  };
  Coef cn = {0.0, 0.0, 0.0, 0.0};
  cplx fn = {0.0, 0.0};
  const int l0 = st.ell0;
  if (tid < kC) ccs[0][tid] = lat::finish_coef(lat::load_coef_raw(cf, l0, tid, L), bsign);
  if (tid < NB * kC) fls[0][fj][fi] = load_f(l0);
  __syncthreads();

  // sums over l by the parity of l - ell0 (chunks and groups start at even offsets)
  cplx acc[2][NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) acc[0][j] = acc[1][j] = {0.0, 0.0};
  int buf = 0;
  for (int blk = l0; blk < L; blk += kC, buf ^= 1) {
    const bool more = blk + kC < L;
    if (more) {
      if (tid < kC) cn = lat::load_coef_raw(cf, blk + kC, tid, L);
      if (tid < NB * kC) fn = load_f(blk + kC);
    }
    const bool live = lscale == 0;
    const lat::ChunkCoef* cs = ccs[buf];
    cplx part[2][NB];
#pragma unroll
    for (int j = 0; j < NB; ++j) part[0][j] = part[1][j] = {0.0, 0.0};
    // each group's coefficients and flm are read from LDS while the previous
    // group steps (two groups in registers at a time)
    lat::ChunkCoef cg[kGroup], ng[kGroup];
    cplx fg[kGroup][NB], nf[kGroup][NB];
#pragma unroll
    for (int e = 0; e < kGroup; ++e) {
      cg[e] = cs[e];
#pragma unroll
      for (int j = 0; j < NB; ++j) fg[e][j] = fls[buf][j][e];
    }
#pragma unroll 2
    for (int g = 0; g < kC / kGroup; ++g) {
      const int gn = g + 1 < kC / kGroup ? g + 1 : g;
#pragma unroll
      for (int e = 0; e < kGroup; ++e) {
        ng[e] = cs[kGroup * gn + e];
#pragma unroll
        for (int j = 0; j < NB; ++j) nf[e][j] = fls[buf][j][kGroup * gn + e];
      }
#pragma unroll
      for (int e = 0; e < kGroup; ++e) {
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          part[e & 1][j].x = fma(d, fg[e][j].x, part[e & 1][j].x);
          part[e & 1][j].y = fma(d, fg[e][j].y, part[e & 1][j].y);
        }
        const double dn = fma(fma(cg[e].a, x, cg[e].nb), d, -cg[e].c * dm1);
        dm1 = d;
        d = dn;
      }
#pragma unroll
      for (int e = 0; e < kGroup; ++e) {
        cg[e] = ng[e];
#pragma unroll
        for (int j = 0; j < NB; ++j) fg[e][j] = nf[e][j];
      }
    }
    // a lane contributes once its scale index is 0 (earlier terms are below
    // 2^-256 of the O(1) ones)
    if (live) {
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        acc[0][j] = cadd(acc[0][j], part[0][j]);
        acc[1][j] = cadd(acc[1][j], part[1][j]);
      }
    }
    rescale(d, dm1, lscale);
    if (more) {
      if (tid < kC) ccs[buf ^ 1][tid] = lat::finish_coef(cn, bsign);
      if (tid < NB * kC) fls[buf ^ 1][fj][fi] = fn;
    }
    __syncthreads();
  }
  if (!valid) return;

  // paired rings: ring b takes (-1)^(l + m) d, i.e. the parity of l - ell0,
  // with the roles swapped when ell0 + m is odd
  const bool flip = ((l0 + m) & 1) != 0;
  cplx* out = a.ftm + (int64_t)b0 * a.ftm_map_stride + (int64_t)(m - a.ftm_m_first) * spec_cstride(a.ftm_tiled);
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < nb) {
      cplx* o = out + (int64_t)j * a.ftm_map_stride;
      const cplx ae = flip ? acc[1][j] : acc[0][j], ao = flip ? acc[0][j] : acc[1][j];
      o[spec_row(rg.ta, a.ftm_ncol, a.ftm_tiled)] = cadd(ae, ao);
      if (PAIRED && rg.tb >= 0 && rg.tb != rg.ta) o[spec_row(rg.tb, a.ftm_ncol, a.ftm_tiled)] = csub(ae, ao);
    }
  }
}

// ------------------------------------------------------------------ forward
// Each lane owns R ring entries (pairs) of one map.  Every kSub steps the
// lane's sums for those kSub l (2 kSub values) are summed across the wave by
// an LDS transpose in the wave's own LDS region (no barrier: each wave
// reduces on its own) and added to flm with FP64 atomics (the callers zero
// flm first or accumulate into it), so the waves and workgroups of one m do
// not synchronise beyond the coefficient staging (one barrier per 32 l).
constexpr int kSub = 8;
constexpr int kRedStride = 2 * kSub + 1;  // doubles per lane row (+1 pad: conflict-free columns)

template <int R, bool PAIRED>
__global__ void __launch_bounds__(kThreads) lat_forward_kernel(LatArgs a) {
  static_assert(kC % kSub == 0 && kSub % kGroup == 0, "step blocking");
  __shared__ lat::ChunkCoef ccs[2][kC];
  __shared__ double red[kThreads / 64][64 * kRedStride];
  const int m = m_of_block(blockIdx.x, a.nm, a.m_lo);
  const int am = m < 0 ? -m : m;
  const int64_t b = blockIdx.z;
  const int g = blockIdx.y;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  // multi-plane launches: this map's own recursion tables
  const LatPlane* pl = a.planes ? a.planes + b : nullptr;
  const MStart st = (pl ? pl->mstart : a.mstart)[m - a.m_lo];
  const int lmin = pl ? pl->lmin_out : a.lmin_out;
  const double bsign = m < 0 ? -1.0 : 1.0;
  const Coef* __restrict__ cf = (pl ? pl->coef : a.coef) + (int64_t)am * a.L;
  const int L = a.L;
  const cplx* ftm = a.ftm + b * a.ftm_map_stride + (int64_t)(m - a.ftm_m_first) * spec_cstride(a.ftm_tiled);
  double* rw = red[wave];

  double x[R], d[R], dm1[R];
  int ls[R];
  cplx se[R], so[R];  // ring-side data (sum / difference of the pair)
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int p = (g * R + r) * kThreads + tid;
    dm1[r] = 0.0;
    x[r] = 0.0;
    d[r] = 0.0;
    ls[r] = 0;
    se[r] = so[r] = {0.0, 0.0};
    if (p < a.nrings) {
      const LatRing rg = a.rings[p];
      x[r] = rg.x;
      start_value(st, rg.c2, rg.s2, d[r], ls[r]);
      const cplx fa = ftm[spec_row(rg.ta, a.ftm_ncol, a.ftm_tiled)];
      cplx fb = {0.0, 0.0};
      if (PAIRED && rg.tb >= 0 && rg.tb != rg.ta) fb = ftm[spec_row(rg.tb, a.ftm_ncol, a.ftm_tiled)];
      se[r] = cadd(fa, fb);
      so[r] = csub(fa, fb);
    }
  }
  // l - ell0 even: the pair's sum where l + m is even, else its difference
  const bool flip = PAIRED && ((st.ell0 + m) & 1) != 0;

  cplx* out = a.flm + b * a.flm_map_stride + flm_col(a, m);
  const int64_t frow = flm_row(a);
  const int l0 = st.ell0;
  Coef cn = {0.0, 0.0, 0.0, 0.0};
  if (tid < kC) ccs[0][tid] = lat::finish_coef(lat::load_coef_raw(cf, l0, tid, L), bsign);
  __syncthreads();
  int buf = 0;
  for (int blk = l0; blk < L; blk += kC, buf ^= 1) {
    const bool more = blk + kC < L;
    if (more && tid < kC) cn = lat::load_coef_raw(cf, blk + kC, tid, L);
    const lat::ChunkCoef* cs = ccs[buf];
    cplx sA[R], sB[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool live = ls[r] == 0;
      const cplx ev = flip ? so[r] : se[r];
      const cplx od = PAIRED ? (flip ? se[r] : so[r]) : se[r];
      sA[r] = live ? ev : cplx{0.0, 0.0};
      sB[r] = live ? od : cplx{0.0, 0.0};
    }
#pragma unroll 1
    for (int sb = 0; sb < kC; sb += kSub) {
      double part[2 * kSub];
#pragma unroll
      for (int i = 0; i < kSub; ++i) {
        const lat::ChunkCoef c = cs[sb + i];  // broadcast LDS read
        double px = 0.0, py = 0.0;
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const cplx u = (i & 1) ? sB[r] : sA[r];
          px = fma(d[r], u.x, px);
          py = fma(d[r], u.y, py);
          const double dn = fma(fma(c.a, x[r], c.nb), d[r], -c.c * dm1[r]);
          dm1[r] = d[r];
          d[r] = dn;
        }
        part[2 * i] = px;
        part[2 * i + 1] = py;
      }
      // sum the wave's 64 rows: lane v < 2 kSub adds column v of rows 0..31,
      // lane v + 32 rows 32..63, and the two halves meet by a swizzle
#pragma unroll
      for (int v = 0; v < 2 * kSub; ++v) rw[lane * kRedStride + v] = part[v];
      lat::wave_sync();
      const int v = lane & 31, half = lane >> 5;
      double t = 0.0;
      if (v < 2 * kSub) {
#pragma unroll 8
        for (int rr = 0; rr < 32; ++rr) t += rw[(32 * half + rr) * kRedStride + v];
      }
      t += __shfl_xor(t, 32);
      lat::wave_sync();
      const int l = blk + sb + (v >> 1);
      if (half == 0 && v < 2 * kSub && l < L && l >= lmin)
        atomicAdd(reinterpret_cast<double*>(out + (int64_t)l * frow) + (v & 1), t);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) rescale(d[r], dm1[r], ls[r]);
    if (more && tid < kC) ccs[buf ^ 1][tid] = lat::finish_coef(cn, bsign);
    __syncthreads();
  }
}

// ------------------------------------------------------------------ launchers
template <int NB>
static hipError_t lat_inv_nb(const LatArgs& a, bool paired, hipStream_t s) {
  dim3 grid(a.nm, (a.nrings + kThreads - 1) / kThreads, (a.batch + NB - 1) / NB);
  if (paired)
    hipLaunchKernelGGL((lat_inverse_kernel<NB, true>), grid, dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL((lat_inverse_kernel<NB, false>), grid, dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_lat_inverse(const LatArgs& a, bool paired, hipStream_t s) {
  if (a.nm == 0 || a.nrings == 0 || a.batch == 0) return hipSuccess;
  if (a.planes) return lat_inv_nb<1>(a, paired, s);  // one map (plane) per block
  if (a.batch >= 4) return lat_inv_nb<4>(a, paired, s);
  if (a.batch >= 2) return lat_inv_nb<2>(a, paired, s);
  return lat_inv_nb<1>(a, paired, s);
}

template <int R>
static hipError_t lat_fwd_r(const LatArgs& a, bool paired, hipStream_t s) {
  dim3 grid(a.nm, a.groups, a.batch);
  if (paired)
    hipLaunchKernelGGL((lat_forward_kernel<R, true>), grid, dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL((lat_forward_kernel<R, false>), grid, dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_lat_forward(const LatArgs& a0, bool paired, hipStream_t s) {
  if (a0.nm == 0 || a0.nrings == 0 || a0.batch == 0) return hipSuccess;
  LatArgs a = a0;
  const int per = (a.nrings + kThreads - 1) / kThreads;  // entries per lane if one group
  int R = per <= 1 ? 1 : per <= 2 ? 2 : per <= 4 ? 4 : 8;
  a.groups = (a.nrings + R * kThreads - 1) / (R * kThreads);
  switch (R) {
    case 1: return lat_fwd_r<1>(a, paired, s);
    case 2: return lat_fwd_r<2>(a, paired, s);
    case 4: return lat_fwd_r<4>(a, paired, s);
    default: return lat_fwd_r<8>(a, paired, s);
  }
}

}  // namespace s2h
