// Device helpers shared by the latitudinal (Wigner-d) kernels.
//
// Start values: d^{l0}_{m,n}(theta) = K cos(theta/2)^ecos sin(theta/2)^esin
// evaluated as (mantissa, exponent) so that sin^a cos^b cannot underflow at
// high m; the result is split into a stored double and an integer scale index
// lscale <= 0 with true value = stored * 2^(512 lscale).  A lane contributes
// only while lscale == 0 (earlier terms are below 2^-256 of the O(1) ones).
#pragma once

#include "s2h_internal.h"
#include "s2h_kernels.h"

namespace s2h {
namespace lat {

struct Ext {
  double m;
  int e;
};

__device__ __forceinline__ Ext ext_of(double v) {
  int e;
  const double m = frexp(v, &e);
  return {m, e};
}
__device__ __forceinline__ Ext ext_mul(Ext a, Ext b) {
  int e;
  const double m = frexp(a.m * b.m, &e);
  return {m, a.e + b.e + e};
}
__device__ __forceinline__ Ext ext_pow(double base, int k) {
  Ext r = {0.5, 1};
  Ext x = ext_of(base);
  while (k) {
    if (k & 1) r = ext_mul(r, x);
    k >>= 1;
    if (k) x = ext_mul(x, x);
  }
  return r;
}

// Start value d^{l0}_{m,n}(theta) split into (stored, lscale).
__device__ __forceinline__ void start_value(const MStart& st, double c2, double s2, double& d,
                                            int& lscale) {
  Ext v = {st.kmant, st.kexp};
  if (st.ecos) v = ext_mul(v, ext_pow(c2, st.ecos));
  if (st.esin) v = ext_mul(v, ext_pow(s2, st.esin));
  if (v.m == 0.0) {
    d = 0.0;
    lscale = 0;
    return;
  }
  int ls = -((256 - v.e) >> 9);
  if (ls > 0) ls = 0;
  lscale = ls;
  d = ldexp(v.m, v.e - 512 * ls);
}

// Moves (d, dm1) down by 2^512 once |d| passes 2^256.  The recursion grows
// by less than 2.5 per step, so checks every 32 steps cannot overflow.
__device__ __forceinline__ void rescale(double& d, double& dm1, int& lscale) {
  const bool big = fabs(d) > 0x1p256;
  const double f = big ? 0x1p-512 : 1.0;
  d *= f;
  dm1 *= f;
  lscale += big ? 1 : 0;
}

// blockIdx.x -> m: 4 adjacent m per XCD in a row (workgroups are dealt to the
// 8 XCDs round-robin), heaviest (smallest |m|) first.  For m_lo < 0 the order
// is 0, -1, 1, -2, 2, ...
__device__ __forceinline__ int m_of_block(int i, int nm, int m_lo) {
  const int nfull = nm & ~31;  // the remap is a bijection on [0, nfull)
  int idx = i;
  if (i < nfull) {
    const int xcd = i & 7, k = i >> 3;
    idx = ((k >> 2) << 5) + (xcd << 2) + (k & 3);
  }
  if (m_lo == 0) return idx;
  return (idx & 1) ? -((idx + 1) >> 1) : (idx >> 1);
}

// Block i of a launch over nm m-values x G ring groups (grid.x = nm G) ->
// (m, ring group g).  Blocks are dispatched round-robin over the 8 XCDs by
// linear index, so i & 7 is the XCD and i >> 3 the order on it: the G groups
// of one m are consecutive blocks of one XCD.  They run together and read the
// m's harmonic and recursion coefficients through that XCD's L2, not HBM.  The m-values
// follow m_of_block.  A bijection of [0, nm G) onto [0, nm) x [0, G).
__device__ __forceinline__ int m_group_of_block(int i, int nm, int G, int m_lo, int& g) {
  const int nm8 = nm & ~7;
  int ip;
  if (i < nm8 * G) {
    const int xcd = i & 7, j = i >> 3;
    const int k = j / G;
    g = j - k * G;
    ip = 8 * k + xcd;
  } else {
    const int r = i - nm8 * G;
    const int k = r / G;
    g = r - k * G;
    ip = nm8 + k;
  }
  return m_of_block(ip, nm, m_lo);
}

// Recursion coefficients of one 32-step chunk, staged in LDS so that every
// step reads them with a broadcast LDS load instead of a dependent global or
// scalar load.  (a, -b * bsign, c) for l = blk + i; zero from l = L - 1 on,
// which makes the steps past the band-limit produce zeros.
struct ChunkCoef {
  double a, nb, c, pad;
};
constexpr int kChunk = 32;

// Chunk `blk`'s coefficient for lane i (< 32) as stored (zero from l = L - 1
// on).  A prefetch keeps it in this form until it is written to LDS
// (finish_coef there): arithmetic on the loaded value right after the load
// would make the wave wait for the memory latency at once.
__device__ __forceinline__ Coef load_coef_raw(const Coef* __restrict__ cf, int blk, int i, int L) {
  const int l = blk + i;
  Coef v = {0.0, 0.0, 0.0, 0.0};
  if (l + 1 < L) v = cf[l];
  return v;
}

__device__ __forceinline__ ChunkCoef finish_coef(const Coef& c, double bsign) {
  return {c.a, -c.b * bsign, c.c, 0.0};
}

// Loads chunk `blk`'s coefficient for lane i (< 32) into registers.
__device__ __forceinline__ ChunkCoef load_coef(const Coef* __restrict__ cf, int blk, int i, int L,
                                               double bsign) {
  return finish_coef(load_coef_raw(cf, blk, i, L), bsign);
}

// Unit-c coefficients {a', b', sigma, 0} (the plan's make_unit_coefs) of
// chunk `blk` for lane i, read as stored up to l = L - 1 (a' and b' are
// already zero there; sigma_{L-1} is needed), zero past it.
__device__ __forceinline__ Coef load_ucoef_raw(const Coef* __restrict__ cf, int blk, int i, int L) {
  const int l = blk + i;
  Coef v = {0.0, 0.0, 0.0, 0.0};
  if (l < L) v = cf[l];
  return v;
}

// One recursion step with staged coefficients: returns d_{l+1}.
__device__ __forceinline__ double rec_step(const ChunkCoef& c, double x, double d, double dm1) {
  return fma(fma(c.a, x, c.nb), d, -c.c * dm1);
}
// The same step in unit-c form (unit coefficients staged): e_{l+1}.
__device__ __forceinline__ double rec_step_unit(const ChunkCoef& c, double x, double e, double em1) {
  return fma(fma(c.a, x, c.nb), e, -em1);
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// The part of the conjugate-gradient dot <Y, A(Y)> over the ring rows
// [a.dot_row_lo, a.dot_row_hi), fused into a synthesis epilogue: those rows
// have no aliasing and one w n = a.dot_wn, so A(Y) = w n Y and a stored value
// v adds c_m w n v^2 (c_m = 2 for m > 0 of real maps).  dacc[t]: the lane's
// sum of v^2 for its map tile t (map 8 t + (lane % 16) / 2 of the workgroup,
// both components).  The workgroup adds the lanes up in a fixed order
// (component, lane group, wave) into a.dot_part[map * a.dot_part_stride +
// blockIdx.x].  red: 2 NW 64 doubles of LDS no longer read by any thread;
// every thread calls this.
template <int NW>
__device__ void synth_dot_partials(const LatArgs& a, int m, int b0, const double (&dacc)[2],
                                   double* red) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  red[(wave * 64 + lane) * 2] = dacc[0];
  red[(wave * 64 + lane) * 2 + 1] = dacc[1];
  __syncthreads();
  if (threadIdx.x < 16) {
    const int mi = threadIdx.x, t = mi >> 3, j = mi & 7;
    double s = 0.0;
    for (int w = 0; w < NW; ++w)
      for (int kg = 0; kg < 4; ++kg)
        for (int c = 0; c < 2; ++c) s += red[(w * 64 + 16 * kg + 2 * j + c) * 2 + t];
    if (b0 + mi < a.batch)
      a.dot_part[(int64_t)(b0 + mi) * a.dot_part_stride + blockIdx.x] =
          (m > 0 ? 2.0 : 1.0) * a.dot_wn * s;
  }
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16x16x4 matrix-core products.  Operand layouts (both types): A[i][k] in
// lane i + 16k, B[k][j] in lane j + 16k.  The result layouts differ:
// FP64 holds C[i][j] in lane j + 16 (i%4), register i/4; FP32 holds it in
// lane j + 16 (i/4), register i%4.  row_of(r, kq) is the row held in
// register r by lanes 16 kq .. 16 kq + 15.
template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef v4d acc;
  __device__ static acc run(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  __device__ static int row_of(int r, int kq) { return 4 * r + kq; }
};
template <>
struct Mma<float> {
  typedef v4f acc;
  __device__ static acc run(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  __device__ static int row_of(int r, int kq) { return 4 * kq + r; }
};

}  // namespace lat
}  // namespace s2h
