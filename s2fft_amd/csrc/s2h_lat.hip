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

constexpr int kCheck = 8;  // steps between rescale checks (growth < 2^100 per check)

using lat::m_of_block;
using lat::rescale;
using lat::start_value;

}  // namespace

// ------------------------------------------------------------------ inverse
template <int NB, bool PAIRED>
__global__ void __launch_bounds__(kThreads) lat_inverse_kernel(LatArgs a) {
  const int m = m_of_block(blockIdx.x, a.nm, a.m_lo);
  const int am = m < 0 ? -m : m;
  const int p = blockIdx.y * kThreads + threadIdx.x;
  if (p >= a.nrings) return;
  const int b0 = blockIdx.z * NB;
  // multi-plane launches (NB = 1): this map's own recursion tables
  const LatPlane* pl = a.planes ? a.planes + b0 : nullptr;
  const MStart st = (pl ? pl->mstart : a.mstart)[m - a.m_lo];
  const int lmin = pl ? pl->lmin_out : a.lmin_out;
  const LatRing rg = a.rings[p];
  const double x = rg.x;
  const double bsign = m < 0 ? -1.0 : 1.0;
  const Coef* __restrict__ cf = (pl ? pl->coef : a.coef) + (int64_t)am * a.L;
  const cplx* __restrict__ flm = a.flm + (int64_t)b0 * a.flm_map_stride + flm_col(a, m);

  double d, dm1 = 0.0;
  int lscale;
  start_value(st, rg.c2, rg.s2, d, lscale);

  cplx ae[NB], ao[NB];
#pragma unroll
  for (int j = 0; j < NB; ++j) ae[j] = ao[j] = {0.0, 0.0};

  const int L = a.L;
  const int nb = min(NB, a.batch - b0);
  for (int blk = st.ell0; blk < L; blk += kCheck) {
    const bool live = lscale == 0;
    // the block's coefficients are loaded together (one memory latency per
    // kCheck steps instead of one per step)
    lat::ChunkCoef cs[kCheck];
#pragma unroll
    for (int i = 0; i < kCheck; ++i) cs[i] = lat::load_coef(cf, blk, i, L, bsign);
#pragma unroll
    for (int i = 0; i < kCheck; ++i) {
      const int l = blk + i;
      if (l < L && l >= lmin) {
        const bool odd = PAIRED && ((l + m) & 1);
#pragma unroll
        for (int j = 0; j < NB; ++j) {
          if (j < nb) {
            const cplx f = flm[(int64_t)j * a.flm_map_stride + (int64_t)l * flm_row(a)];
            if (odd) {
              ao[j].x = fma(d, f.x, ao[j].x);
              ao[j].y = fma(d, f.y, ao[j].y);
            } else {
              ae[j].x = fma(d, f.x, ae[j].x);
              ae[j].y = fma(d, f.y, ae[j].y);
            }
          }
        }
      }
      const double dn = lat::rec_step(cs[i], x, d, dm1);
      dm1 = d;
      d = dn;
    }
    if (!live) {
#pragma unroll
      for (int j = 0; j < NB; ++j) ae[j] = ao[j] = {0.0, 0.0};
    }
    rescale(d, dm1, lscale);
  }

  cplx* out = a.ftm + (int64_t)b0 * a.ftm_map_stride + (m - a.ftm_m_first);
#pragma unroll
  for (int j = 0; j < NB; ++j) {
    if (j < nb) {
      cplx* o = out + (int64_t)j * a.ftm_map_stride;
      o[(int64_t)rg.ta * a.ftm_ncol] = cadd(ae[j], ao[j]);
      if (PAIRED && rg.tb >= 0 && rg.tb != rg.ta) o[(int64_t)rg.tb * a.ftm_ncol] = csub(ae[j], ao[j]);
    }
  }
}

// ------------------------------------------------------------------ forward
// Each lane owns R ring entries; a chunk of kChunk consecutive l is reduced
// across the wave through an LDS transpose, then across the 4 waves.
constexpr int kChunk = 8;

template <int R, bool PAIRED>
__global__ void __launch_bounds__(kThreads) lat_forward_kernel(LatArgs a) {
  __shared__ cplx red[kThreads / 64][kChunk][64];
  __shared__ cplx wsum[kThreads / 64][kChunk];
  const int m = m_of_block(blockIdx.x, a.nm, a.m_lo);
  const int am = m < 0 ? -m : m;
  const int64_t b = blockIdx.z;
  const int g = blockIdx.y;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  // multi-plane launches: this map's own recursion tables
  const LatPlane* pl = a.planes ? a.planes + b : nullptr;
  const MStart st = (pl ? pl->mstart : a.mstart)[m - a.m_lo];
  const int lmin = pl ? pl->lmin_out : a.lmin_out;
  const double bsign = m < 0 ? -1.0 : 1.0;
  const Coef* __restrict__ cf = (pl ? pl->coef : a.coef) + (int64_t)am * a.L;
  const int L = a.L;
  const cplx* ftm = a.ftm + b * a.ftm_map_stride + (m - a.ftm_m_first);

  double x[R], d[R], dm1[R];
  int ls[R];
  cplx se[R], so[R];      // ring-side data (sum / difference of the pair)
  cplx ue[R], uo[R];      // what is used: zero until the lane is live
#pragma unroll
  for (int r = 0; r < R; ++r) {
    const int p = (g * R + r) * kThreads + threadIdx.x;
    dm1[r] = 0.0;
    if (p < a.nrings) {
      const LatRing rg = a.rings[p];
      x[r] = rg.x;
      start_value(st, rg.c2, rg.s2, d[r], ls[r]);
      const cplx fa = ftm[(int64_t)rg.ta * a.ftm_ncol];
      cplx fb = {0.0, 0.0};
      if (PAIRED && rg.tb >= 0 && rg.tb != rg.ta) fb = ftm[(int64_t)rg.tb * a.ftm_ncol];
      se[r] = cadd(fa, fb);
      so[r] = csub(fa, fb);
    } else {
      x[r] = 0.0;
      d[r] = 0.0;
      ls[r] = 0;
      se[r] = so[r] = {0.0, 0.0};
    }
  }

  cplx* out = a.flm + b * a.flm_map_stride + flm_col(a, m);
  for (int blk = st.ell0; blk < L; blk += kChunk) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const bool live = ls[r] == 0;
      ue[r] = live ? se[r] : cplx{0.0, 0.0};
      uo[r] = live ? so[r] : cplx{0.0, 0.0};
    }
    cplx part[kChunk];
    lat::ChunkCoef cs[kChunk];
#pragma unroll
    for (int i = 0; i < kChunk; ++i) cs[i] = lat::load_coef(cf, blk, i, L, bsign);
#pragma unroll
    for (int i = 0; i < kChunk; ++i) {
      const int l = blk + i;
      const bool odd = PAIRED && ((l + m) & 1);
      cplx acc = {0.0, 0.0};
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const cplx u = odd ? uo[r] : ue[r];
        acc.x = fma(d[r], u.x, acc.x);
        acc.y = fma(d[r], u.y, acc.y);
      }
      part[i] = acc;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const double dn = lat::rec_step(cs[i], x[r], d[r], dm1[r]);
        dm1[r] = d[r];
        d[r] = dn;
      }
    }
#pragma unroll
    for (int i = 0; i < kChunk; ++i) red[wave][i][lane] = part[i];
    __syncthreads();
    {
      const int i = lane >> 3, seg = lane & 7;
      cplx s = {0.0, 0.0};
#pragma unroll
      for (int j = 0; j < 8; ++j) s = cadd(s, red[wave][i][seg * 8 + j]);
      s.x += __shfl_xor(s.x, 1);
      s.y += __shfl_xor(s.y, 1);
      s.x += __shfl_xor(s.x, 2);
      s.y += __shfl_xor(s.y, 2);
      s.x += __shfl_xor(s.x, 4);
      s.y += __shfl_xor(s.y, 4);
      if (seg == 0) wsum[wave][i] = s;
    }
    __syncthreads();
    if (threadIdx.x < kChunk) {
      const int l = blk + threadIdx.x;
      if (l < L && l >= lmin) {
        cplx t = wsum[0][threadIdx.x];
#pragma unroll
        for (int w = 1; w < kThreads / 64; ++w) t = cadd(t, wsum[w][threadIdx.x]);
        cplx* o = out + (int64_t)l * flm_row(a);
        if (a.groups > 1 || a.accumulate) {
          atomicAdd(&o->x, t.x);
          atomicAdd(&o->y, t.y);
        } else {
          *o = t;
        }
      }
    }
#pragma unroll
    for (int r = 0; r < R; ++r) rescale(d[r], dm1[r], ls[r]);
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
  if (a.batch >= 8) return lat_inv_nb<8>(a, paired, s);
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
