// Latitudinal steps with the projection / synthesis on the BF16 matrix cores
// (split-BF16 operands, FP32 accumulation); the ell-recursion stays FP64.
//
// Every FP64 operand v (a d-value or a ring / flm value) is split into
// hi = bf16(v) and lo = bf16(v - hi), so v = hi + lo to 2^-17 relative, and
// each product is formed as hi*hi + hi*lo + lo*hi on
// V_MFMA_F32_16X16X32_BF16 (A[i][k]: lane i + 16 (k/8), element k%8; B[k][j]:
// lane j + 16 (k/8), element k%8; C[i][j]: lane j + 16 (i/4), register i%4).
// Products carry ~1e-5 relative error: this is the fast inner transform of
// a refinement whose residual is FP64.
//
// Forward:  flm[b, l, m] = sum_p S_par(l)[b, p] d_l(p)   (K = ring pairs)
//   A = S (16 rows = 8 maps x re/im, per 8-map tile), B = d (K = 32 pairs x
//   16 l of one parity).  Each lane runs the recursion of one ring pair and
//   writes its split d-values to LDS as [l][pair], the B layout.
// Inverse:  ftm[b, p, m] = sum_l d_l(p) flm[b, l, m]      (K = l)
//   A = d (16 pairs x K = 32 l), B = flm (32 l x 16 columns = 8 maps x re/im).
//   Each lane writes its d-values as [pair][l], the A layout; the flm chunk
//   is staged once per workgroup as [column][l], the B layout.  Even and odd
//   l (relative to m) accumulate separately to serve theta and pi - theta.
//
// Both run the recursion in its unit-c form (the plan's unit coefficients):
// e_{l+1} = (a'_l x - b'_l) e_l - e_{l-1}, two FP64 operations per step
// instead of three, with d_l = sigma_l e_l.  sigma_l depends on l only, so
// the forward scales its FP64 sums by it and the inverse its staged flm.
#include "s2h_fft_lds.h"
#include "s2h_kernels.h"
#include "s2h_lat_common.h"

namespace s2h {

namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

using lat::m_of_block;
using lat::start_value;

// FP64 -> FP32 -> BF16, each step rounded to nearest.  The empty asm keeps the
// compiler from fusing the two conversions into one correctly rounded FP64 ->
// BF16 conversion, which it expands to about ten instructions (round-to-odd
// fix-up with two FP64 compares) per value; the double rounding it avoids is
// far below the 2^-8 of a BF16 operand.
__device__ __forceinline__ float to_f32(double v) {
  float f = (float)v;
  asm("" : "+v"(f));
  return f;
}

__device__ __forceinline__ void split2f(float f, __bf16& hi, __bf16& lo) {
  hi = (__bf16)f;
  lo = (__bf16)(f - (float)hi);
}

__device__ __forceinline__ void split2(double v, __bf16& hi, __bf16& lo) {
  split2f(to_f32(v), hi, lo);
}

typedef float f32x2 __attribute__((ext_vector_type(2)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));

// split2f of two values into elements e, e + 1 of hi / lo: one paired
// conversion for the high parts, their exact FP32 values from the bits, one
// paired conversion of the remainders
__device__ __forceinline__ void split_pair(float a, float b, bf16x8& hi, bf16x8& lo, int e) {
  const bf16x2 h = __builtin_convertvector(f32x2{a, b}, bf16x2);
  const unsigned bits = __builtin_bit_cast(unsigned, h);
  const float ha = __builtin_bit_cast(float, bits << 16);
  const float hb = __builtin_bit_cast(float, bits & 0xffff0000u);
  const bf16x2 l = __builtin_convertvector(f32x2{a - ha, b - hb}, bf16x2);
  hi[e] = h[0];
  hi[e + 1] = h[1];
  lo[e] = l[0];
  lo[e + 1] = l[1];
}

// The value of lane i ^ 1 (DPP quad permutation [1, 0, 3, 2]).
__device__ __forceinline__ double swap_pair_lanes(double v) {
  const unsigned long long u = __builtin_bit_cast(unsigned long long, v);
  const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), 0xB1, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ f32x4 mma(const bf16x8& a, const bf16x8& b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// hi*hi + hi*lo + lo*hi
__device__ __forceinline__ f32x4 mma3(const bf16x8& ah, const bf16x8& al, const bf16x8& bh,
                                      const bf16x8& bl, f32x4 c) {
  c = mma(al, bh, c);
  c = mma(ah, bl, c);
  return mma(ah, bh, c);
}

constexpr int kFW = 8;          // forward: waves per workgroup
constexpr int kFT = kFW * 64;
constexpr int kFTiles = 2;      // 8-map tiles per workgroup (16 maps)

}  // namespace

// ------------------------------------------------------------------ forward
// NP: adjacent ring pairs per lane.  A wave covers 64 NP consecutive pairs
// (lane i: pairs NP i .. NP i + NP - 1), a workgroup kFW * 64 NP.  At each
// recursion step a lane holds the d-values of its NP pairs at one l, which
// are adjacent in the tile row of that l: one packed LDS store per step
// (8 bytes for NP = 4, 4 bytes for NP = 2; the lo parts a second one).
// SPLIT: 3 = hi*hi + hi*lo + lo*hi, 1 = hi*hi only (BF16 operands, ~4e-3
// relative: the correction transform of a refinement whose residual is FP64).
// TILED: the spectra layout (spec_row) as a compile-time choice.
template <bool PAIRED, int NP, int SPLIT, bool TILED>
__global__ void __launch_bounds__(kFT) lat_forward_bf16_kernel(LatArgs a) {
  constexpr bool LO = SPLIT == 3;
  static_assert(NP == 2 || NP == 4, "pairs per lane");
  constexpr int KB = 2 * NP;                      // 32-pair K blocks per wave
  constexpr int kRow = 64 * NP + 8;               // tile row stride (BF16): 8 pad, so the
                                                  // 16-byte B reads of 8 rows hit distinct banks
  constexpr int kTileE = 32 * kRow;               // [l_local 32 (even 16 | odd 16)][pair]
  constexpr int kTileB = kTileE * 2;
  constexpr int kWaveB = (LO ? 2 : 1) * kTileB;
  static_assert(kWaveB >= kFTiles * 2 * 256 * 4, "result tile fits the wave's d tiles");
  // per wave: the hi d tile (+ the lo tile); after the chunk's products the
  // same bytes hold the wave's results [tile][parity][row 16][col 16] (FP32)
  __shared__ __attribute__((aligned(16))) unsigned char wbuf[kFW][kWaveB];
  __shared__ lat::ChunkCoef ccs[2][lat::kChunk];  // recursion coefficients, double-buffered
  const int m = m_of_block(blockIdx.x, a.nm, a.m_lo);
  const int am = m < 0 ? -m : m;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const int L = a.L;
  const int b0 = blockIdx.z * (8 * kFTiles);
  const MStart st = a.mstart[m - a.m_lo];
  const double bsign = m < 0 ? -1.0 : 1.0;
  const Coef* __restrict__ cf = a.ucoef + (int64_t)am * L;  // unit-c: d_l = sigma_l e_l
  __bf16* dh = reinterpret_cast<__bf16*>(wbuf[wave]);
  __bf16* dl = reinterpret_cast<__bf16*>(wbuf[wave] + kTileB);
  float* ctw = reinterpret_cast<float*>(wbuf[wave]);
  const int pbase = (blockIdx.y * kFW + wave) * 64 * NP;

  // recursion state of this lane's pairs
  double x[NP], d[NP], dm1[NP];
  int ls[NP];
#pragma unroll
  for (int j = 0; j < NP; ++j) {
    const int p = pbase + NP * lane + j;
    dm1[j] = 0.0;
    x[j] = 0.0;
    d[j] = 0.0;
    ls[j] = 0;
    if (p < a.nrings) {
      const LatRing rg = a.rings[p];
      x[j] = rg.x;
      start_value(st, rg.c2, rg.s2, d[j], ls[j]);
    }
  }
  // A operands: row = (map row/2 of the tile, re/im row%2), K = pairs
  // 32 kb + 8 kg + e of the wave.  Each ring value is loaded once: of the two
  // lanes of a map (rows 2r: re, 2r + 1: im) the first loads ring ta, the
  // second ring tb, and they swap the component the other needs.
  bf16x8 seh[KB][kFTiles], soh[KB][kFTiles];
  bf16x8 sel[LO ? KB : 1][kFTiles], sol[LO ? KB : 1][kFTiles];
  const int comp = row & 1;
#pragma unroll
  for (int t = 0; t < kFTiles; ++t) {
    const int amap = b0 + 8 * t + (row >> 1);
    const bool have = amap < a.batch;
    const cplx* ftm = a.ftm + (int64_t)amap * a.ftm_map_stride + (int64_t)(m - a.ftm_m_first) * spec_cstride(TILED);
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const int q = pbase + 32 * kb + 8 * kg + e;
        cplx v = {0.0, 0.0};
        if (have && q < a.nrings) {
          const int2 tt = *reinterpret_cast<const int2*>(&a.rings[q].ta);
          const int tr = comp ? tt.y : tt.x;
          if (comp == 0 || (PAIRED && tt.y >= 0 && tt.y != tt.x))
            v = ftm[spec_row(tr, a.ftm_ncol, TILED)];
        }
        // lane 2r keeps fa.x and receives fb.x; lane 2r + 1 keeps fb.y and
        // receives fa.y (a zero fb where the ring is unpaired)
        const double give = comp ? v.x : v.y;
        const double got = swap_pair_lanes(give);
        const double ca = comp ? got : v.x;
        const double cb = comp ? v.y : got;
        const double ve = ca + cb;
        const double vo = PAIRED ? ca - cb : ca;
        __bf16 h, l;
        split2(ve, h, l);
        seh[kb][t][e] = h;
        if constexpr (LO) sel[kb][t][e] = l;
        split2(vo, h, l);
        soh[kb][t][e] = h;
        if constexpr (LO) sol[kb][t][e] = l;
      }
    }
  }

  cplx* out = a.flm + (int64_t)b0 * a.flm_map_stride + flm_col(a, m);
  const bool mirror = a.mirror && m > 0;
  if (a.fill && !a.accumulate) {
    // structural zeros of columns m (and -m): l < ell0 (from l = |m| when packed)
    const int l_beg = a.flm_packed ? am : 0;
    const int l_end = st.ell0 < L ? st.ell0 : L;
    const int nl = l_end > l_beg ? l_end - l_beg : 0;
    for (int e = threadIdx.x; e < nl * 8 * kFTiles; e += kFT) {
      const int j = e / nl, l = l_beg + e - j * nl;
      if (b0 + j >= a.batch) continue;
      cplx* o = out + (int64_t)j * a.flm_map_stride + (int64_t)l * flm_row(a);
      *o = cplx{0.0, 0.0};
      if (mirror) o[-2 * m] = cplx{0.0, 0.0};
    }
  }
  Coef cn = {0.0, 0.0, 0.0, 0.0};
  if (threadIdx.x < lat::kChunk)
    ccs[0][threadIdx.x] = lat::finish_coef(lat::load_ucoef_raw(cf, st.ell0, threadIdx.x, L), bsign);
  __syncthreads();
  int buf = 0;
  for (int blk = st.ell0; blk < L; blk += 32, buf ^= 1) {
    const bool more = blk + 32 < L;
    if (more && threadIdx.x < lat::kChunk) cn = lat::load_ucoef_raw(cf, blk + 32, threadIdx.x, L);
    const lat::ChunkCoef* cs = ccs[buf];
    f32x4 acc[kFTiles][2];
#pragma unroll
    for (int t = 0; t < kFTiles; ++t) {
      acc[t][0] = f32x4{0.f, 0.f, 0.f, 0.f};
      acc[t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    }
    // a pair contributes only while its scale index is 0 (it changes only in
    // the rescale after the chunk): one mask per two pairs, applied to their
    // packed BF16 values
    unsigned mk[NP / 2];
#pragma unroll
    for (int h = 0; h < NP / 2; ++h)
      mk[h] = (ls[2 * h] == 0 ? 0x0000ffffu : 0u) | (ls[2 * h + 1] == 0 ? 0xffff0000u : 0u);
    // 32 unit-c steps e_{l+1} = (a' x - b') e_l - e_{l-1}; the values past
    // the band-limit are zero (zero coefficients from l = L - 1 on).  Each
    // GS-step group's coefficients are read from LDS while the previous group
    // steps.  Step i goes to tile row (i odd ? 16 : 0) + i / 2.
    constexpr int GS = LO ? 4 : 2;
    double ca[GS], cb[GS], na[GS], nb[GS];
#pragma unroll
    for (int e = 0; e < GS; ++e) {
      ca[e] = cs[e].a;
      cb[e] = cs[e].nb;
    }
#pragma unroll
    for (int g = 0; g < 32 / GS; ++g) {
      if (g + 1 < 32 / GS) {
#pragma unroll
        for (int e = 0; e < GS; ++e) {
          na[e] = cs[GS * (g + 1) + e].a;
          nb[e] = cs[GS * (g + 1) + e].nb;
        }
      }
#pragma unroll
      for (int e = 0; e < GS; ++e) {
        const int i = GS * g + e;
        const int trow = (i & 1) * 16 + (i >> 1);
        float fv[NP];
#pragma unroll
        for (int j = 0; j < NP; ++j) {
          fv[j] = to_f32(d[j]);
          const double dn = fma(fma(ca[e], x[j], cb[e]), d[j], -dm1[j]);
          dm1[j] = d[j];
          d[j] = dn;
        }
        unsigned hw[NP / 2], lw[NP / 2];
#pragma unroll
        for (int h = 0; h < NP / 2; ++h) {
          const bf16x2 hh = __builtin_convertvector(f32x2{fv[2 * h], fv[2 * h + 1]}, bf16x2);
          const unsigned bits = __builtin_bit_cast(unsigned, hh);
          hw[h] = bits & mk[h];
          if constexpr (LO) {
            const float h0 = __builtin_bit_cast(float, bits << 16);
            const float h1 = __builtin_bit_cast(float, bits & 0xffff0000u);
            const bf16x2 lo = __builtin_convertvector(f32x2{fv[2 * h] - h0, fv[2 * h + 1] - h1},
                                                      bf16x2);
            lw[h] = __builtin_bit_cast(unsigned, lo) & mk[h];
          }
        }
        const int off = trow * kRow + NP * lane;
        if constexpr (NP == 4) {
          *reinterpret_cast<uint2*>(dh + off) = uint2{hw[0], hw[1]};
          if constexpr (LO) *reinterpret_cast<uint2*>(dl + off) = uint2{lw[0], lw[1]};
        } else {
          *reinterpret_cast<unsigned*>(dh + off) = hw[0];
          if constexpr (LO) *reinterpret_cast<unsigned*>(dl + off) = lw[0];
        }
      }
#pragma unroll
      for (int e = 0; e < GS; ++e) {
        ca[e] = na[e];
        cb[e] = nb[e];
      }
    }
#pragma unroll
    for (int j = 0; j < NP; ++j) lat::rescale(d[j], dm1[j], ls[j]);
    lat::wave_sync();
#pragma unroll
    for (int kb = 0; kb < KB; ++kb) {
#pragma unroll
      for (int par = 0; par < 2; ++par) {
        // B[k][j]: d of pair 32 kb + 8 kg + e at column j = row (parity par)
        const int off = (par * 16 + row) * kRow + 32 * kb + 8 * kg;
        const bf16x8 bh = *reinterpret_cast<const bf16x8*>(dh + off);
        if constexpr (LO) {
          const bf16x8 bl = *reinterpret_cast<const bf16x8*>(dl + off);
#pragma unroll
          for (int t = 0; t < kFTiles; ++t) {
            if (par == 0)
              acc[t][0] = mma3(seh[kb][t], sel[kb][t], bh, bl, acc[t][0]);
            else
              acc[t][1] = mma3(soh[kb][t], sol[kb][t], bh, bl, acc[t][1]);
          }
        } else {
#pragma unroll
          for (int t = 0; t < kFTiles; ++t) {
            if (par == 0)
              acc[t][0] = mma(seh[kb][t], bh, acc[t][0]);
            else
              acc[t][1] = mma(soh[kb][t], bh, acc[t][1]);
          }
        }
      }
    }
    lat::wave_sync();
    // C[i][j]: lane j + 16 (i/4), register i%4 -> this wave's result tile,
    // over its own d tile (its last reads are ordered before by wave_sync)
#pragma unroll
    for (int t = 0; t < kFTiles; ++t)
#pragma unroll
      for (int par = 0; par < 2; ++par)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          ctw[(t * 2 + par) * 256 + (4 * kg + r) * 16 + row] = acc[t][par][r];
    if (more && threadIdx.x < lat::kChunk) ccs[buf ^ 1][threadIdx.x] = lat::finish_coef(cn, bsign);
    __syncthreads();
    // element (t, parity, row i, col j) -> l = blk + 2 j + parity, map 8t + i/2;
    // sigma_l (still in this chunk's coefficient buffer) turns e_l into d_l
    for (int e = threadIdx.x; e < kFTiles * 2 * 256; e += kFT) {
      const int t = e >> 9, par = (e >> 8) & 1, ri = (e >> 4) & 15, cj = e & 15;
      const int l = blk + 2 * cj + par;
      const int bm = 8 * t + (ri >> 1);
      double v = 0.0;
#pragma unroll
      for (int w = 0; w < kFW; ++w) v += (double)reinterpret_cast<const float*>(wbuf[w])[e];
      v *= cs[2 * cj + par].c;
      if (l < L && b0 + bm < a.batch && (l >= a.lmin_out || a.fill)) {
        if (l < a.lmin_out) v = 0.0;
        double* o = reinterpret_cast<double*>(out + (int64_t)bm * a.flm_map_stride +
                                              (int64_t)l * flm_row(a)) + (ri & 1);
        if (a.groups > 1 || a.accumulate)
          atomicAdd(o, v);
        else
          *o = v;
        if (mirror) {
          // column -m: (-1)^m conj, i.e. the imaginary part negated once more
          const double mv = ((m & 1) ? -v : v) * ((ri & 1) ? -1.0 : 1.0);
          double* om = o - 4 * m;
          if (a.accumulate)
            atomicAdd(om, mv);
          else
            *om = mv;
        }
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------------ inverse
namespace {
constexpr int kIW = 4;          // inverse: waves per workgroup
constexpr int kIRow = 40;       // inverse tile row stride (BF16): 32 l + 8 pad, so the 16-byte
                                // operand reads and writes of 16 rows hit distinct banks
constexpr int kIT = kIW * 64;
constexpr int kIMaps = 16;      // maps per workgroup (two column tiles)
}  // namespace

// LO: split operands (hi*hi + hi*lo + lo*hi, ~1e-5 relative); else one BF16
// product (~4e-3 relative) and no lo tiles.  Both run 3 waves per SIMD (168
// registers; LDS 27 or 53 KB per workgroup): the split one at 2 waves took
// 3.57 instead of 3.43 ms at L 2048, 16 maps.
template <bool PAIRED, bool LO, bool TILED>
__global__ void __launch_bounds__(kIT, 3) lat_inverse_bf16_kernel(LatArgs a) {
  constexpr int kLoW = LO ? kIW : 1, kLoE = LO ? 64 * kIRow : 8, kLoB = LO ? 32 * kIRow : 8;
  // d tile per wave: [pair 64][l_local 32] hi / lo (A layout)
  __shared__ __attribute__((aligned(16))) __bf16 dth[kIW][64 * kIRow];
  __shared__ __attribute__((aligned(16))) __bf16 dtl[kLoW][kLoE];
  // flm chunk: [parity][column 32 = 16 maps x re/im][l 32] hi / lo (B layout),
  // entries of the other parity zero
  __shared__ __attribute__((aligned(16))) __bf16 bth[2][32 * kIRow];
  __shared__ __attribute__((aligned(16))) __bf16 btl[2][kLoB];
  __shared__ lat::ChunkCoef ccs[2][lat::kChunk];  // recursion coefficients, double-buffered
  int grp;
  const int m = lat::m_group_of_block(blockIdx.x, a.nm, (a.nrings + kIT - 1) / kIT, a.m_lo, grp);
  const int am = m < 0 ? -m : m;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int row = lane & 15, kg = lane >> 4;
  const int L = a.L;
  const int b0 = blockIdx.z * kIMaps;
  const MStart st = a.mstart[m - a.m_lo];
  const double bsign = m < 0 ? -1.0 : 1.0;
  const Coef* __restrict__ cf = a.ucoef + (int64_t)am * L;  // unit-c: d_l = sigma_l e_l
  __bf16* dh = dth[wave];
  __bf16* dl = dtl[LO ? wave : 0];

  const int pbase = (grp * kIW + wave) * 64;
  const int p = pbase + lane;
  double x = 0.0, d = 0.0, dm1 = 0.0;
  int ls = 0;
  if (p < a.nrings) {
    const LatRing rg = a.rings[p];
    x = rg.x;
    start_value(st, rg.c2, rg.s2, d, ls);
  }
  // accumulators [q: 16-pair tile][column tile][parity]
  f32x4 acc[4][2][2];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      acc[q][t][0] = f32x4{0.f, 0.f, 0.f, 0.f};
      acc[q][t][1] = f32x4{0.f, 0.f, 0.f, 0.f};
    }

  const cplx* flm = a.flm + (int64_t)b0 * a.flm_map_stride + flm_col(a, m);
  // flm chunk staging: thread t holds elements e = t + kIT * j (j < 4) of
  // [column 32][l_local 32]; the next chunk is loaded while this one runs
  constexpr int kStage = 32 * 32 / kIT;
  auto load_flm = [&](int blk, double* v) {
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
      const int e = threadIdx.x + kIT * j;
      const int col = e >> 5, i = e & 31;
      const int l = blk + i, bm = col >> 1;
      double x0 = 0.0;
      if (l < L && l >= a.lmin_out && b0 + bm < a.batch) {
        const double* f = reinterpret_cast<const double*>(
            flm + (int64_t)bm * a.flm_map_stride + (int64_t)l * flm_row(a));
        x0 = f[col & 1];
      }
      v[j] = x0;
    }
  };
  // by parity of i, the other parity's entry zero; scaled by sigma_l (from
  // the chunk's staged coefficients), since the A tile holds e_l = d_l / sigma_l
  auto store_flm = [&](const double* v, const lat::ChunkCoef* cc) {
#pragma unroll
    for (int j = 0; j < kStage; ++j) {
      const int e = threadIdx.x + kIT * j;
      const int col = e >> 5, i = e & 31;
      __bf16 h, lo;
      split2(v[j] * cc[i].c, h, lo);
      const int par = i & 1;
      bth[par][col * kIRow + i] = h;
      bth[par ^ 1][col * kIRow + i] = (__bf16)0.0f;
      if constexpr (LO) {
        btl[par][col * kIRow + i] = lo;
        btl[par ^ 1][col * kIRow + i] = (__bf16)0.0f;
      }
    }
  };
  double fv[kStage];
  Coef cn = {0.0, 0.0, 0.0, 0.0};
  load_flm(st.ell0, fv);
  if (threadIdx.x < lat::kChunk)
    ccs[0][threadIdx.x] = lat::finish_coef(lat::load_ucoef_raw(cf, st.ell0, threadIdx.x, L), bsign);
  __syncthreads();
  store_flm(fv, ccs[0]);
  __syncthreads();
  int buf = 0;
  for (int blk = st.ell0; blk < L; blk += 32, buf ^= 1) {
    const int nxt = blk + 32;
    const bool more = nxt < L;
    if (more) {
      load_flm(nxt, fv);
      if (threadIdx.x < lat::kChunk) cn = lat::load_ucoef_raw(cf, nxt, threadIdx.x, L);
    }
    // 32 unit-c recursion steps (e_l = d_l / sigma_l) into the A tile, 8 l
    // per 16-byte store.  The coefficients from l = L - 1 on are zero, so e
    // vanishes past the band-limit.  Each 4-step group's coefficients are read from LDS while
    // the previous group steps, keeping the LDS latency off the FP64 chain.
    const bool live = ls == 0;
    const lat::ChunkCoef* cs = ccs[buf];
    double ca[4], cb[4], na[4], nb[4];
#pragma unroll
    for (int e = 0; e < 4; ++e) {
      ca[e] = cs[e].a;
      cb[e] = cs[e].nb;
    }
#pragma unroll
    for (int g8 = 0; g8 < 4; ++g8) {
      float v[8];
#pragma unroll
      for (int h4 = 0; h4 < 2; ++h4) {
        const int g = 2 * g8 + h4;
        if (g < 7) {
#pragma unroll
          for (int e = 0; e < 4; ++e) {
            na[e] = cs[4 * (g + 1) + e].a;
            nb[e] = cs[4 * (g + 1) + e].nb;
          }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float f = to_f32(d);
          v[4 * h4 + e] = live ? f : 0.0f;
          const double dn = fma(fma(ca[e], x, cb[e]), d, -dm1);
          dm1 = d;
          d = dn;
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          ca[e] = na[e];
          cb[e] = nb[e];
        }
      }
      bf16x8 vh, vl;
      if constexpr (LO) {
#pragma unroll
        for (int e = 0; e < 8; e += 2) split_pair(v[e], v[e + 1], vh, vl, e);
        *reinterpret_cast<bf16x8*>(dl + lane * kIRow + 8 * g8) = vl;
      } else {
#pragma unroll
        for (int e = 0; e < 8; e += 2) {
          const bf16x2 h = __builtin_convertvector(f32x2{v[e], v[e + 1]}, bf16x2);
          vh[e] = h[0];
          vh[e + 1] = h[1];
        }
      }
      *reinterpret_cast<bf16x8*>(dh + lane * kIRow + 8 * g8) = vh;
    }
    lat::rescale(d, dm1, ls);
    if (more && threadIdx.x < lat::kChunk) ccs[buf ^ 1][threadIdx.x] = lat::finish_coef(cn, bsign);
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      // A[i][k]: pair 16 q + row, l = 8 kg + e
      const int aoff = (16 * q + row) * kIRow + 8 * kg;
      const bf16x8 ah = *reinterpret_cast<const bf16x8*>(dh + aoff);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int par = 0; par < 2; ++par) {
          // B[k][j]: column 16 t + row, l = 8 kg + e
          const int boff = (16 * t + row) * kIRow + 8 * kg;
          const bf16x8 bh = *reinterpret_cast<const bf16x8*>(&bth[par][boff]);
          if constexpr (LO) {
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(dl + aoff);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(&btl[par][boff]);
            acc[q][t][par] = mma3(ah, al, bh, bl, acc[q][t][par]);
          } else {
            acc[q][t][par] = mma(ah, bh, acc[q][t][par]);
          }
        }
    }
    __syncthreads();
    if (more) store_flm(fv, ccs[buf ^ 1]);
  }
  // C[i][j]: pair 16 q + 4 kg + r, column 16 t + row -> map t*8 + row/2
  // Parity is relative to l = blk offsets, i.e. to l0 = ell0, whose parity is
  // that of |m| for spin 0 (the paired case).
  // ring rows of the 16 pairs this lane writes, loaded before any store (the
  // stores could alias the ring table, which would serialise the loads)
  int ta[16], tb[16];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int pp = pbase + 16 * q + 4 * kg + r;
      int ra = -1, rb = -1;
      if (pp < a.nrings) {
        const LatRing rg = a.rings[pp];
        ra = rg.ta;
        if (PAIRED && rg.tb >= 0 && rg.tb != rg.ta) rb = rg.tb;
      }
      ta[4 * q + r] = ra;
      tb[4 * q + r] = rb;
    }
  const bool dot = a.dot_part != nullptr;
  double dacc[2] = {0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int bm = b0 + t * 8 + (row >> 1);
    if (bm >= a.batch) continue;
    double* out = reinterpret_cast<double*>(a.ftm + (int64_t)bm * a.ftm_map_stride +
                                            (int64_t)(m - a.ftm_m_first) * spec_cstride(TILED)) + (row & 1);
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int ra = ta[4 * q + r], rb = tb[4 * q + r];
        const double e = (double)acc[q][t][0][r], o = (double)acc[q][t][1][r];
        if (ra >= 0) {
          out[2 * spec_row(ra, a.ftm_ncol, TILED)] = e + o;
          if (dot && ra >= a.dot_row_lo && ra < a.dot_row_hi) dacc[t] = fma(e + o, e + o, dacc[t]);
        }
        if (rb >= 0) {
          out[2 * spec_row(rb, a.ftm_ncol, TILED)] = e - o;
          if (dot && rb >= a.dot_row_lo && rb < a.dot_row_hi) dacc[t] = fma(e - o, e - o, dacc[t]);
        }
      }
  }
  // the d tiles are no longer read (the loop ended on a barrier)
  if (dot) lat::synth_dot_partials<kIW>(a, m, b0, dacc, reinterpret_cast<double*>(dth));
}

// ------------------------------------------------------------------ launchers
template <int NP, int SPLIT>
static hipError_t lat_fwd_bf16(const LatArgs& a0, bool paired, hipStream_t s) {
  LatArgs a = a0;
  const int per_wg = kFW * NP * 64;
  a.groups = (a.nrings + per_wg - 1) / per_wg;
  dim3 grid(a.nm, a.groups, (a.batch + 8 * kFTiles - 1) / (8 * kFTiles));
  if (paired && a.ftm_tiled)
    hipLaunchKernelGGL((lat_forward_bf16_kernel<true, NP, SPLIT, true>), grid, dim3(kFT), 0, s, a);
  else if (paired)
    hipLaunchKernelGGL((lat_forward_bf16_kernel<true, NP, SPLIT, false>), grid, dim3(kFT), 0, s, a);
  else if (a.ftm_tiled)
    hipLaunchKernelGGL((lat_forward_bf16_kernel<false, NP, SPLIT, true>), grid, dim3(kFT), 0, s, a);
  else
    hipLaunchKernelGGL((lat_forward_bf16_kernel<false, NP, SPLIT, false>), grid, dim3(kFT), 0, s,
                       a);
  return hipGetLastError();
}

int lat_forward_bf16_groups(const LatArgs& a, int split) {
  const int per_wg = kFW * (split == 1 ? 4 : 2) * 64;
  return (a.nrings + per_wg - 1) / per_wg;
}

hipError_t launch_lat_forward_bf16(const LatArgs& a, bool paired, int split, hipStream_t s) {
  if (a.nm == 0 || a.nrings == 0 || a.batch == 0) return hipSuccess;
  return split == 1 ? lat_fwd_bf16<4, 1>(a, paired, s) : lat_fwd_bf16<2, 3>(a, paired, s);
}

template <bool LO>
static hipError_t lat_inv_bf16(const LatArgs& a, bool paired, hipStream_t s) {
  dim3 grid(a.nm * ((a.nrings + kIT - 1) / kIT), 1, (a.batch + kIMaps - 1) / kIMaps);
  if (paired && a.ftm_tiled)
    hipLaunchKernelGGL((lat_inverse_bf16_kernel<true, LO, true>), grid, dim3(kIT), 0, s, a);
  else if (paired)
    hipLaunchKernelGGL((lat_inverse_bf16_kernel<true, LO, false>), grid, dim3(kIT), 0, s, a);
  else if (a.ftm_tiled)
    hipLaunchKernelGGL((lat_inverse_bf16_kernel<false, LO, true>), grid, dim3(kIT), 0, s, a);
  else
    hipLaunchKernelGGL((lat_inverse_bf16_kernel<false, LO, false>), grid, dim3(kIT), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_lat_inverse_bf16(const LatArgs& a, bool paired, int split, hipStream_t s) {
  if (a.nm == 0 || a.nrings == 0 || a.batch == 0) return hipSuccess;
  return split == 1 ? lat_inv_bf16<false>(a, paired, s) : lat_inv_bf16<true>(a, paired, s);
}

}  // namespace s2h
