// Learned preconditioner for the iterative forward transform: a per-(l, m)
// MLP  (re, im, l/L, m/L) -> correction (re, im),  two hidden layers of kH
// with ReLU, FP16 weights, features and activations, FP32 accumulation:
//
//   z1 = W1h h(x) + b1,  a1 = h(relu(z1)),  z2 = W2h a1 + b2,
//   a2 = h(relu(z2)),    y  = W3h a2 + b3,   Wh = h(W), h = round to FP16
//
// x = (h_re / s, h_im / s, l / L, m / L) for the harmonic residual h of a map
// whose RMS over its coefficients is s; the correction is s * y.  Inference
// runs all three layers on the FP16 matrix cores with the activations on
// packed FP16; training (forward + backward, FP32 on the vector ALUs)
// accumulates per-workgroup gradient partials that a second kernel reduces
// in a fixed order, and Adam updates the FP32 master weights.  The backward
// treats the roundings as the identity (straight-through).
#include "s2h_internal.h"
#include "s2h_kernels.h"

namespace s2h {

namespace {

constexpr int kH = kMlpHidden;

// master parameter offsets (floats)
constexpr int oW1 = 0;               // [kH][4]
constexpr int oB1 = oW1 + kH * 4;    // [kH]
constexpr int oW2 = oB1 + kH;        // [kH out][kH in]
constexpr int oB2 = oW2 + kH * kH;   // [kH]
constexpr int oW3 = oB2 + kH;        // [2][kH]
constexpr int oB3 = oW3 + 2 * kH;    // [2]
static_assert(oB3 + 2 == kMlpParams, "parameter layout");

__device__ __forceinline__ float bf16_round(float v) {
  return (float)(__bf16)v;
}
__device__ __forceinline__ float f16_round(float v) {
  return (float)(_Float16)v;
}

// Activation: ReLU (derivative 1 for x > 0, else 0).
__device__ __forceinline__ float act(float x) { return fmaxf(x, 0.0f); }
__device__ __forceinline__ float act_grad(float x) { return x > 0.0f ? 1.0f : 0.0f; }

typedef _Float16 h2 __attribute__((ext_vector_type(2)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// The MFMA operand of eight hidden units held as two accumulator tiles:
// elements 0-3 from t0, 4-7 from t1, through ReLU on packed FP16 (one
// conversion and one maximum per pair; the rounding commutes with ReLU).
__device__ __forceinline__ f16x8 act_operand(const f32x4& t0, const f32x4& t1) {
  const h2 z = {(_Float16)0.0f, (_Float16)0.0f};
  const h2 g0 = __builtin_elementwise_max(h2{(_Float16)t0[0], (_Float16)t0[1]}, z);
  const h2 g1 = __builtin_elementwise_max(h2{(_Float16)t0[2], (_Float16)t0[3]}, z);
  const h2 g2 = __builtin_elementwise_max(h2{(_Float16)t1[0], (_Float16)t1[1]}, z);
  const h2 g3 = __builtin_elementwise_max(h2{(_Float16)t1[2], (_Float16)t1[3]}, z);
  return f16x8{g0[0], g0[1], g1[0], g1[1], g2[0], g2[1], g3[0], g3[1]};
}

// packed (l, m) enumeration: reality -> m in [0, l], P = L(L+1)/2;
// else m in [-l, l], P = L^2
__device__ __forceinline__ void lm_of(int64_t p, int reality, int& l, int& m) {
  if (reality) {
    int ll = (int)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
    while ((int64_t)(ll + 1) * (ll + 2) / 2 <= p) ++ll;
    while ((int64_t)ll * (ll + 1) / 2 > p) --ll;
    l = ll;
    m = (int)(p - (int64_t)ll * (ll + 1) / 2);
  } else {
    int ll = (int)sqrt((double)p);
    while ((int64_t)(ll + 1) * (ll + 1) <= p) ++ll;
    while ((int64_t)ll * ll > p) --ll;
    l = ll;
    m = (int)(p - (int64_t)ll * ll - ll);
  }
}

// (l, m) of packed row p < 2^31 in 32-bit arithmetic: the FP32 root of
// moff(m) = p (moff as packed_moff) lands within a few m of the answer and
// the two searches settle it.
__device__ __forceinline__ void packed_lm_i32(int p, int L, int& l, int& m) {
  const float b = 2.0f * (float)L + 1.0f;
  const float d = fmaxf(fmaf(b, b, -8.0f * (float)p), 0.0f);
  int mm = (int)((b - __builtin_amdgcn_sqrtf(d)) * 0.5f);
  mm = min(max(mm, 0), L - 1);
  while (mm > 0 && mm * L - (mm * (mm - 1)) / 2 > p) --mm;
  while (mm + 1 < L && (mm + 1) * L - ((mm + 1) * mm) / 2 <= p) ++mm;
  m = mm;
  l = p - (mm * L - (mm * (mm - 1)) / 2) + mm;
}

// row p -> (l, m) and the element index of row p in h / out
__device__ __forceinline__ int64_t row_index(const MlpRowArgs& a, int64_t p, int& l, int& m) {
  if (a.packed) {
    packed_lm(p, a.L, l, m);
    return p;
  }
  lm_of(p, a.reality, l, m);
  return (int64_t)l * (2 * a.L - 1) + a.L - 1 + m;
}

}  // namespace

// ---------------------------------------------------------------- statistics
// Per map, kMlpSumsqParts partial sums of |h|^2 over the rows with l >= lmin
// (workgroup w: rows w kThreads + j kMlpSumsqParts kThreads + thread), into
// part[b kMlpSumsqParts + w]; mlp_sigma adds them up in a fixed order, so the
// RMS, and with it the network's output, is the same on every run.
__global__ void __launch_bounds__(kThreads) mlp_sumsq_kernel(MlpRowArgs a, double* part) {
  const int b = blockIdx.y;
  __shared__ double red[kThreads / 64];
  double acc = 0.0;
  for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < a.P;
       p += (int64_t)gridDim.x * kThreads) {
    int l, m;
    const int64_t idx = row_index(a, p, l, m);
    if (l < a.lmin) continue;
    const cplx v = a.h[(int64_t)b * a.map_stride + idx];
    acc += v.x * v.x + v.y * v.y;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) t += red[w];
    part[(int64_t)b * kMlpSumsqParts + blockIdx.x] = t;
  }
}

// RMS of map b's residual from its partial sums (fixed order: kMlpSumsqParts
// / 64 adjacent ones per lane of the calling wave, then a butterfly); every
// lane of the wave gets it
__device__ __forceinline__ double mlp_sigma(const double* part, int b, int64_t count) {
  constexpr int kPer = kMlpSumsqParts / 64;
  static_assert(kPer * 64 == kMlpSumsqParts, "whole partial sums per lane");
  const int lane = threadIdx.x & 63;
  const double* p = part + (int64_t)b * kMlpSumsqParts + kPer * lane;
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) t += p[i];
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
  return t > 0.0 ? sqrt(t / (double)count) : 1.0;
}

// ---------------------------------------------------------------- inference
// One wave handles TILES tiles of 16 rows per pass, all three layers on the
// FP16 matrix cores with the activations in registers (operand layouts of
// V_MFMA_F32_16X16X{16,32}_F16: A[i][k] in lane i + 16 (k/E % 4), element
// k % E, B[k][j] in lane j + 16 (k/E % 4), element k % E, E = 4 or 8 halves
// per lane; result C[i][j] in lane j + 16 (i/4), register i % 4):
//  * layer 1 as kH/16 V_MFMA_F32_16X16X16_F16, one per 16 hidden units: A = W1h
//    (lanes 0-15 hold the unit's 4 weights, the rest zero), B = the rows'
//    4 features (lane r < 16 holds row r's), bias preloaded in the
//    accumulator.  The result holds unit 16 nt + 4 (lane/16) + i of row
//    lane%16 in register i of tile nt.
//  * layer 2 as kH/16 output tiles x kH/32 k-blocks of V_MFMA_F32_16X16X32_F16 with
//    the weights as A and the activations as B.  B is taken straight from
//    layer 1's registers (through ReLU), which fixes the order of the
//    hidden units along k: k = 32 kb + 8 q + e is unit
//    32 kb + 16 (e/4) + 4 q + e%4.  W2h is staged in LDS in the A layout
//    with its columns in that order.  The result again holds unit
//    16 ct + 4 (lane/16) + i of row lane%16 in register i of tile ct.
//  * layer 3 the same way as kH/32 MFMAs with W3h padded to 16 outputs as A:
//    lanes 0-15 end with the row's two outputs in registers 0 and 1.
// Each weight fragment read from LDS serves every tile of the pass.
typedef _Float16 f16x4 __attribute__((ext_vector_type(4)));

constexpr int kInfWaves = 4;
constexpr int kInfTiles = 1;  // 16-row tiles per wave per pass
constexpr int kNT = kH / 16;  // 16-unit tiles of a hidden layer
constexpr int kKB = kH / 32;  // 32-unit k-blocks of a hidden layer
static_assert(kKB * 32 == kH && kKB >= 1, "hidden width: a multiple of 32");

template <int TILES>
__global__ void __launch_bounds__(kInfWaves * 64) mlp_apply_kernel(MlpRowArgs a,
                                                                   const float* __restrict__ w,
                                                                   const double* sumsq,
                                                                   double* dot_part) {
  // A operand of layer-2 tile ct, k-block kb for lane l:
  // W2h[16 ct + l%16][unit(32 kb + 8 (l/16) + e)], unit() as above
  __shared__ f16x8 w2a[kNT][kKB][64];
  // A operand of layer 3, k-block kb: W3h[l%16][unit(...)], zero for l%16 >= 2
  __shared__ f16x8 w3a[kKB][64];
  // A operand of layer-1 tile nt for lane l < 16: W1h[16 nt + l][0..3]
  __shared__ f16x4 w1a[kNT][16];
  __shared__ __attribute__((aligned(16))) float b1[kH];
  __shared__ __attribute__((aligned(16))) float b2[kH];
  __shared__ float b3[2];
  for (int e = threadIdx.x; e < kNT * kKB * 64; e += kInfWaves * 64) {
    const int ct = e / (kKB * 64), kb = (e >> 6) % kKB, l = e & 63;
    const int j = 16 * ct + (l & 15), q4 = 4 * (l >> 4);
    f16x8 v;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      v[q] = (_Float16)w[oW2 + j * kH + 32 * kb + 16 * (q >> 2) + q4 + (q & 3)];
    w2a[ct][kb][l] = v;
  }
  for (int e = threadIdx.x; e < kKB * 64; e += kInfWaves * 64) {
    const int kb = e >> 6, l = e & 63;
    const int o = l & 15, q4 = 4 * (l >> 4);
    f16x8 v;
#pragma unroll
    for (int q = 0; q < 8; ++q)
      v[q] = o < 2 ? (_Float16)w[oW3 + o * kH + 32 * kb + 16 * (q >> 2) + q4 + (q & 3)]
                   : (_Float16)0.0f;
    w3a[kb][l] = v;
  }
  for (int e = threadIdx.x; e < kNT * 16; e += kInfWaves * 64) {
    const int u = e;  // unit 16 nt + l
    f16x4 v;
#pragma unroll
    for (int f = 0; f < 4; ++f) v[f] = (_Float16)w[oW1 + 4 * u + f];
    w1a[u >> 4][u & 15] = v;
  }
  for (int e = threadIdx.x; e < kH; e += kInfWaves * 64) {
    b1[e] = w[oB1 + e];
    b2[e] = w[oB2 + e];
  }
  if (threadIdx.x < 2) b3[threadIdx.x] = w[oB3 + threadIdx.x];
  __shared__ double sig_s;
  if (threadIdx.x < 64) {
    const double sg = mlp_sigma(sumsq, blockIdx.y, a.count);
    if (threadIdx.x == 0) sig_s = sg;
  }
  __syncthreads();

  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int b = blockIdx.y;
  const double sig = sig_s;
  const float inv_sig = (float)(1.0 / sig);
  const float inv_L = 1.0f / (float)a.L;
  const cplx* __restrict__ h = a.h + (int64_t)b * a.map_stride;
  cplx* __restrict__ out = a.out + (int64_t)b * a.map_stride;
  const int row = lane & 15, kq = lane >> 4;
  const f16x4 zero4 = {(_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f, (_Float16)0.0f};

  const int64_t ntile = (a.P + 15) / 16;
  const int64_t npass = (ntile + TILES - 1) / TILES;
  const int64_t stride = (int64_t)gridDim.x * kInfWaves;
  // dot_part: this lane's part of <delta, h> (c_m = 2 for the m > 0 rows of
  // real maps: their conjugate rows), summed in a fixed order below
  double dsum = 0.0;
  // Row `row` of tile `tile`: (l, m), its element index and (lanes 0-15) its
  // residual, loaded one pass ahead so that the load's latency overlaps the
  // pass before.  Packed: (l, m) from the tile's first row, stepping over the
  // m-blocks (block m holds l = m .. L - 1).
  auto locate = [&](int64_t tile, int& l, int& m, int64_t& idx, cplx& v) {
    const int64_t p = tile * 16 + row;
    l = 0;
    m = 0;
    idx = p;
    v = cplx{0.0, 0.0};
    if (p >= a.P) return;
    if (a.packed) {
      int l0, m0;
      packed_lm_i32((int)(tile * 16), a.L, l0, m0);
      l = l0 + row;
      m = m0;
      while (l >= a.L && m + 1 < a.L) {
        l -= a.L - m - 1;
        ++m;
      }
    } else {
      idx = row_index(a, p, l, m);
    }
    if (kq == 0 && l >= a.lmin) v = h[idx];
  };
  int lr[TILES], mr[TILES];
  int64_t ir[TILES];
  cplx vr[TILES];
  int64_t pass = (int64_t)blockIdx.x * kInfWaves + wave;
#pragma unroll
  for (int s = 0; s < TILES; ++s) locate(pass * TILES + s, lr[s], mr[s], ir[s], vr[s]);
  for (; pass < npass; pass += stride) {
    // The weights are re-read from LDS every pass: kept in registers across
    // passes (the compiler's choice without this barrier) they take ~200
    // registers and leave one wave per SIMD to hide every latency.
    asm volatile("" ::: "memory");
    int ln[TILES], mn[TILES];
    int64_t in[TILES];
    cplx vn[TILES];
#pragma unroll
    for (int s = 0; s < TILES; ++s) locate((pass + stride) * TILES + s, ln[s], mn[s], in[s], vn[s]);
    f16x8 av[TILES][kKB];
#pragma unroll
    for (int s = 0; s < TILES; ++s) {
      // lanes 0-15: the features (h_re / s, h_im / s, l / L, m / L) of row `row`
      f16x4 xf = zero4;
      if (kq == 0 && (pass * TILES + s) * 16 + row < a.P && lr[s] >= a.lmin)
        xf = f16x4{(_Float16)((float)vr[s].x * inv_sig), (_Float16)((float)vr[s].y * inv_sig),
                   (_Float16)((float)lr[s] * inv_L), (_Float16)((float)mr[s] * inv_L)};
      // layer 1: z of unit 16 nt + 4 kq + i, row `row`, in register i of z1[nt]
      f32x4 z1[kNT];
#pragma unroll
      for (int nt = 0; nt < kNT; ++nt) {
        z1[nt] = *reinterpret_cast<const f32x4*>(&b1[16 * nt + 4 * kq]);
        const f16x4 wv = kq == 0 ? w1a[nt][row] : zero4;
        z1[nt] = __builtin_amdgcn_mfma_f32_16x16x16f16(wv, xf, z1[nt], 0, 0, 0);
      }
#pragma unroll
      for (int kb = 0; kb < kKB; ++kb) av[s][kb] = act_operand(z1[2 * kb], z1[2 * kb + 1]);
    }
    // layer 2: unit 16 ct + 4 kq + i of row `row` in register i of acc[s][ct]
    // (each tile's bias read from LDS by itself: one read instead of four moves)
    f32x4 acc[TILES][kNT];
#pragma unroll
    for (int s = 0; s < TILES; ++s) {
#pragma unroll
      for (int ct = 0; ct < kNT; ++ct)
        acc[s][ct] = *reinterpret_cast<const volatile f32x4*>(&b2[16 * ct + 4 * kq]);
    }
#pragma unroll
    for (int kb = 0; kb < kKB; ++kb) {
#pragma unroll
      for (int ct = 0; ct < kNT; ++ct) {
        const f16x8 wa = w2a[ct][kb][lane];
#pragma unroll
        for (int s = 0; s < TILES; ++s)
          acc[s][ct] = __builtin_amdgcn_mfma_f32_16x16x32_f16(wa, av[s][kb], acc[s][ct], 0, 0, 0);
      }
    }
    // layer 3 and the write-out: lane `row` (kq = 0) holds the row's outputs
#pragma unroll
    for (int s = 0; s < TILES; ++s) {
      f32x4 y = {b3[0], b3[1], 0.f, 0.f};
#pragma unroll
      for (int kb = 0; kb < kKB; ++kb)
        y = __builtin_amdgcn_mfma_f32_16x16x32_f16(
            w3a[kb][lane], act_operand(acc[s][2 * kb], acc[s][2 * kb + 1]), y, 0, 0, 0);
      const int64_t q = (pass * TILES + s) * 16 + row;
      if (kq == 0 && q < a.P) {
        const int lq = lr[s], mq = mr[s];
        const int64_t idx = ir[s];
        if (lq < a.lmin) {
          out[idx] = cplx{0.0, 0.0};
        } else {
          const double re = sig * (double)y[0];
          double im = sig * (double)y[1];
          if (a.reality && mq == 0) im = 0.0;
          out[idx] = cplx{re, im};
          if (dot_part) {
            const double c = a.reality && mq != 0 ? 2.0 : 1.0;
            dsum = fma(c, fma(re, vr[s].x, im * vr[s].y), dsum);
          }
          if (a.reality && !a.packed && mq > 0) {
            const double sg = (mq & 1) ? -1.0 : 1.0;
            out[(int64_t)lq * (2 * a.L - 1) + a.L - 1 - mq] = cplx{sg * re, -sg * im};
          }
        }
      }
    }
#pragma unroll
    for (int s = 0; s < TILES; ++s) {
      lr[s] = ln[s];
      mr[s] = mn[s];
      ir[s] = in[s];
      vr[s] = vn[s];
    }
  }
  if (dot_part) {
    // lanes 16-63 hold 0; butterfly per wave, then the waves in order
    __shared__ double wsum[kInfWaves];
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o);
    if (lane == 0) wsum[wave] = dsum;
    __syncthreads();
    if (threadIdx.x == 0) {
      double t = 0.0;
      for (int q = 0; q < kInfWaves; ++q) t += wsum[q];
      dot_part[(int64_t)b * kMlpDotParts + blockIdx.x] = t;
    }
  }
}

// dot[b] = the fixed-order sum of map b's nparts partial sums
__global__ void __launch_bounds__(64) mlp_dot_final_kernel(const double* part, int nparts,
                                                           double* dot) {
  constexpr int kPer = kMlpDotParts / 64;
  const int b = blockIdx.x, lane = threadIdx.x;
  double t = 0.0;
#pragma unroll
  for (int i = 0; i < kPer; ++i) {
    const int k = kPer * lane + i;
    if (k < nparts) t += part[(int64_t)b * kMlpDotParts + k];
  }
  for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
  if (lane == 0) dot[b] = t;
}

// ---------------------------------------------------------------- features
// Training rows from (h, e): X = (h / s, l / L, m / L), Y = e / s, one row
// per packed coefficient with l >= lmin (others get a zero row and weight 0).
__global__ void __launch_bounds__(kThreads) mlp_features_kernel(MlpRowArgs a, const cplx* e,
                                                                const double* sumsq, float* X,
                                                                float* Y) {
  const int b = blockIdx.y;
  __shared__ double sig_s;
  if (threadIdx.x < 64) {
    const double sg = mlp_sigma(sumsq, b, a.count);
    if (threadIdx.x == 0) sig_s = sg;
  }
  __syncthreads();
  const double sig = sig_s;
  for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < a.P;
       p += (int64_t)gridDim.x * kThreads) {
    int l, m;
    const int64_t idx0 = row_index(a, p, l, m);
    const int64_t row = (int64_t)b * a.P + p;
    float4 xv = {0.f, 0.f, 0.f, 0.f};
    float2 yv = {0.f, 0.f};
    if (l >= a.lmin) {
      const int64_t idx = (int64_t)b * a.map_stride + idx0;
      const cplx hv = a.h[idx], ev = e[idx];
      xv = float4{(float)(hv.x / sig), (float)(hv.y / sig), (float)l / a.L, (float)m / a.L};
      yv = float2{(float)(ev.x / sig), (float)(ev.y / sig)};
    }
    reinterpret_cast<float4*>(X)[row] = xv;
    reinterpret_cast<float2*>(Y)[row] = yv;
  }
}

// ---------------------------------------------------------------- training
// 256 threads, rows in blocks of 16.  Thread t owns hidden unit k = t % kH
// for rows r = kG i + g (g = t / kH, i < kTR / kG) in layers 1 and 2, and the
// kH kH / 256 entries (j, k) = (kG i + g, k) of dW2.
constexpr int kTR = 16;            // rows per block
constexpr int kTT = 256;           // threads
constexpr int kW2S = kH + 1;       // padded row stride of W2 in LDS
constexpr int kG = kTT / kH;       // thread groups per hidden unit
constexpr int kRG = kTR / kG;      // rows per thread in a block
constexpr int kNW2 = kH * kH / kTT;  // dW2 entries per thread
static_assert(kG * kH == kTT && kRG * kG == kTR, "training layout");

__global__ void __launch_bounds__(kTT) mlp_grad_kernel(const float* __restrict__ w,
                                                       const float* __restrict__ X,
                                                       const float* __restrict__ Y,
                                                       const float* __restrict__ rw,
                                                       int64_t n, float loss_scale,
                                                       float* __restrict__ partial,
                                                       double* __restrict__ loss_partial) {
  __shared__ float W2s[kH * kW2S];  // W2h[j][k] at j * kW2S + k
  __shared__ float W1s[kH * 4], B1s[kH], B2s[kH], W3s[2 * kH], B3s[2];
  __shared__ float xs[kTR][4], a1s[kTR][kW2S], a2s[kTR][kW2S], dz2s[kTR][kW2S];
  __shared__ float douts[kTR][2];
  __shared__ double lred[kTT / 64];
  const int t = threadIdx.x;
  for (int e = t; e < kH * kH; e += kTT) W2s[(e / kH) * kW2S + (e % kH)] = f16_round(w[oW2 + e]);
  for (int e = t; e < kH * 4; e += kTT) W1s[e] = f16_round(w[oW1 + e]);
  for (int e = t; e < kH; e += kTT) {
    B1s[e] = w[oB1 + e];
    B2s[e] = w[oB2 + e];
  }
  for (int e = t; e < 2 * kH; e += kTT) W3s[e] = f16_round(w[oW3 + e]);
  if (t < 2) B3s[t] = w[oB3 + t];
  __syncthreads();

  const int k = t % kH, grp = t / kH;
  float gW2[kNW2];
#pragma unroll
  for (int i = 0; i < kNW2; ++i) gW2[i] = 0.f;
  float gW1[4] = {0.f, 0.f, 0.f, 0.f}, gB1 = 0.f, gB2 = 0.f, gW3[2] = {0.f, 0.f};
  float gB3[2] = {0.f, 0.f};
  double lsum = 0.0;

  const int64_t nblk = (n + kTR - 1) / kTR;
  for (int64_t blk = blockIdx.x; blk < nblk; blk += gridDim.x) {
    const int64_t r0 = blk * kTR;
    if (t < kTR * 4) {
      const int r = t / 4, f = t % 4;
      xs[r][f] = (r0 + r < n) ? f16_round(X[(r0 + r) * 4 + f]) : 0.f;
    }
    __syncthreads();
    // layer 1 (unit k, rows kG i + grp)
    float z1[kRG];
#pragma unroll
    for (int i = 0; i < kRG; ++i) {
      const int r = kG * i + grp;
      z1[i] = fmaf(W1s[4 * k + 3], xs[r][3],
                   fmaf(W1s[4 * k + 2], xs[r][2],
                        fmaf(W1s[4 * k + 1], xs[r][1], fmaf(W1s[4 * k], xs[r][0], B1s[k]))));
      a1s[r][k] = f16_round(act(z1[i]));
    }
    __syncthreads();
    // layer 2 (unit j = k)
    float z2[kRG];
#pragma unroll
    for (int i = 0; i < kRG; ++i) {
      const int r = kG * i + grp;
      float s = B2s[k];
#pragma unroll 8
      for (int q = 0; q < kH; ++q) s = fmaf(W2s[k * kW2S + q], a1s[r][q], s);
      z2[i] = s;
      a2s[r][k] = f16_round(act(s));
    }
    __syncthreads();
    // layer 3 and the loss: threads 0..31 -> (row t/2, output t%2)
    if (t < 2 * kTR) {
      const int r = t >> 1, o = t & 1;
      float s = B3s[o];
      for (int q = 0; q < kH; ++q) s = fmaf(W3s[o * kH + q], a2s[r][q], s);
      float d = 0.f;
      if (r0 + r < n) {
        const float wt = rw ? rw[r0 + r] : 1.0f;
        const float diff = s - Y[(r0 + r) * 2 + o];
        d = 2.0f * diff * wt * loss_scale;
        lsum += (double)(diff * diff * wt);
      }
      douts[r][o] = d;
    }
    __syncthreads();
    // backward through layer 3 and relu(z2)
#pragma unroll
    for (int i = 0; i < kRG; ++i) {
      const int r = kG * i + grp;
      const float d0 = douts[r][0], d1 = douts[r][1];
      const float a2 = a2s[r][k];
      gW3[0] = fmaf(d0, a2, gW3[0]);
      gW3[1] = fmaf(d1, a2, gW3[1]);
      const float dz = (d0 * W3s[k] + d1 * W3s[kH + k]) * act_grad(z2[i]);
      dz2s[r][k] = dz;
      gB2 += dz;
    }
    if (t < 2) {
#pragma unroll
      for (int r = 0; r < kTR; ++r) gB3[t] += douts[r][t];
    }
    __syncthreads();
    // dW2[j][k] += dz2[r][j] a1q[r][k] for the owned (j, k) = (kG i + grp, k)
#pragma unroll 1
    for (int r = 0; r < kTR; ++r) {
      const float a1 = a1s[r][k];
#pragma unroll
      for (int i = 0; i < kNW2; ++i) gW2[i] = fmaf(dz2s[r][kG * i + grp], a1, gW2[i]);
    }
    // da1[r][k] = sum_j dz2[r][j] W2h[j][k]; back through relu(z1)
#pragma unroll
    for (int i = 0; i < kRG; ++i) {
      const int r = kG * i + grp;
      float s = 0.f;
#pragma unroll 8
      for (int j = 0; j < kH; ++j) s = fmaf(dz2s[r][j], W2s[j * kW2S + k], s);
      const float dz1 = s * act_grad(z1[i]);
      gB1 += dz1;
#pragma unroll
      for (int f = 0; f < 4; ++f) gW1[f] = fmaf(dz1, xs[r][f], gW1[f]);
    }
    __syncthreads();
  }

  // per-workgroup partials: the kG groups of a unit combine through LDS, in
  // group order
  float* pw = partial + (int64_t)blockIdx.x * kMlpParams;
#pragma unroll
  for (int i = 0; i < kNW2; ++i) pw[oW2 + (kG * i + grp) * kH + k] = gW2[i];
  __shared__ float comb[kG][kH][8];
  comb[grp][k][0] = gW1[0];
  comb[grp][k][1] = gW1[1];
  comb[grp][k][2] = gW1[2];
  comb[grp][k][3] = gW1[3];
  comb[grp][k][4] = gB1;
  comb[grp][k][5] = gB2;
  comb[grp][k][6] = gW3[0];
  comb[grp][k][7] = gW3[1];
  __syncthreads();
  if (grp == 0) {
    float c[8];
#pragma unroll
    for (int f = 0; f < 8; ++f) {
      c[f] = comb[0][k][f];
      for (int g = 1; g < kG; ++g) c[f] += comb[g][k][f];
    }
#pragma unroll
    for (int f = 0; f < 4; ++f) pw[oW1 + 4 * k + f] = c[f];
    pw[oB1 + k] = c[4];
    pw[oB2 + k] = c[5];
    pw[oW3 + k] = c[6];
    pw[oW3 + kH + k] = c[7];
  }
  if (t < 2) pw[oB3 + t] = gB3[t];
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o);
  if ((t & 63) == 0) lred[t >> 6] = lsum;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int q = 0; q < kTT / 64; ++q) s += lred[q];
    loss_partial[blockIdx.x] = s;
  }
}

// grad[p] = sum over workgroups of partial[wg][p] (fixed order, FP64 sum)
__global__ void __launch_bounds__(kThreads) mlp_grad_reduce_kernel(const float* partial, int nwg,
                                                                   float* grad,
                                                                   const double* loss_partial,
                                                                   double* loss) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p < kMlpParams) {
    double s = 0.0;
    for (int g = 0; g < nwg; ++g) s += (double)partial[(int64_t)g * kMlpParams + p];
    grad[p] = (float)s;
  }
  if (p == 0 && loss) {
    double s = 0.0;
    for (int g = 0; g < nwg; ++g) s += loss_partial[g];
    *loss = s;
  }
}

// Adam (Kingma & Ba 2015) on the FP32 master weights.
__global__ void __launch_bounds__(kThreads) mlp_adam_kernel(float* w, const float* grad, float* m1,
                                                            float* m2, int step, float lr,
                                                            float beta1, float beta2, float eps) {
  const int p = blockIdx.x * kThreads + threadIdx.x;
  if (p >= kMlpParams) return;
  const float g = grad[p];
  const float a = beta1 * m1[p] + (1.0f - beta1) * g;
  const float v = beta2 * m2[p] + (1.0f - beta2) * g * g;
  m1[p] = a;
  m2[p] = v;
  const float c1 = 1.0f - powf(beta1, (float)step);
  const float c2 = 1.0f - powf(beta2, (float)step);
  w[p] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
}

// ---------------------------------------------------------------- launchers
hipError_t launch_mlp_sumsq(const MlpRowArgs& a, double* part, hipStream_t s) {
  if (a.batch == 0 || a.P == 0) return hipSuccess;
  // every partial is written (a workgroup past the rows writes 0)
  hipLaunchKernelGGL(mlp_sumsq_kernel, dim3(kMlpSumsqParts, a.batch), dim3(kThreads), 0, s, a,
                     part);
  return hipGetLastError();
}

hipError_t launch_mlp_apply(const MlpRowArgs& a, const float* w, const double* sumsq,
                            double* dot_part, double* dot, hipStream_t s) {
  if (a.batch == 0 || a.P == 0) return hipSuccess;
  if (a.packed && a.P + 32 > INT32_MAX) return hipErrorInvalidValue;
  const int64_t ntile = (a.P + 15) / 16;
  const int64_t npass = (ntile + kInfTiles - 1) / kInfTiles;
  const int64_t want = (npass + kInfWaves - 1) / kInfWaves;
  // each workgroup stages the weights once (16K conversions); 256 per map
  // (64 passes per wave at L = 2048) keep that a small part of its work
  static_assert(kMlpDotParts >= 256 && kMlpDotParts % 64 == 0, "dot partial sums");
  dim3 grid((unsigned)(want < 256 ? want : 256), a.batch);
  hipLaunchKernelGGL(mlp_apply_kernel<kInfTiles>, grid, dim3(kInfWaves * 64), 0, s, a, w, sumsq,
                     dot ? dot_part : nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !dot) return e;
  hipLaunchKernelGGL(mlp_dot_final_kernel, dim3(a.batch), dim3(64), 0, s, dot_part,
                     (int)grid.x, dot);
  return hipGetLastError();
}

hipError_t launch_mlp_features(const MlpRowArgs& a, const cplx* e, const double* sumsq, float* X,
                               float* Y, hipStream_t s) {
  if (a.batch == 0 || a.P == 0) return hipSuccess;
  const int64_t want = (a.P + kThreads - 1) / kThreads;
  dim3 grid((unsigned)(want < 4096 ? want : 4096), a.batch);
  hipLaunchKernelGGL(mlp_features_kernel, grid, dim3(kThreads), 0, s, a, e, sumsq, X, Y);
  return hipGetLastError();
}

hipError_t launch_mlp_grad(const float* w, const float* X, const float* Y, const float* rw,
                           int64_t n, float loss_scale, int nwg, float* partial,
                           double* loss_partial, float* grad, double* loss, hipStream_t s) {
  if (nwg < 1) return hipErrorInvalidValue;
  hipLaunchKernelGGL(mlp_grad_kernel, dim3(nwg), dim3(kTT), 0, s, w, X, Y, rw, n, loss_scale,
                     partial, loss_partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(mlp_grad_reduce_kernel, dim3((kMlpParams + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, s, partial, nwg, grad, loss_partial, loss);
  return hipGetLastError();
}

hipError_t launch_mlp_adam(float* w, const float* grad, float* m1, float* m2, int step, float lr,
                           float beta1, float beta2, float eps, hipStream_t s) {
  hipLaunchKernelGGL(mlp_adam_kernel, dim3((kMlpParams + kThreads - 1) / kThreads),
                     dim3(kThreads), 0, s, w, grad, m1, m2, step, lr, beta1, beta2, eps);
  return hipGetLastError();
}

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// One V_MFMA_F32_16X16X32_BF16 on caller operands (64 x 8 bf16 bit patterns
// each), for checking the lane layouts: out[4 lane + r] = register r.
__global__ void mfma_bf16_probe_kernel(const uint16_t* a, const uint16_t* b, float* out) {
  const int lane = threadIdx.x;
  bf16x8 av, bv;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    uint16_t ua = a[lane * 8 + q], ub = b[lane * 8 + q];
    av[q] = __builtin_bit_cast(__bf16, ua);
    bv[q] = __builtin_bit_cast(__bf16, ub);
  }
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc, 0, 0, 0);
  for (int r = 0; r < 4; ++r) out[lane * 4 + r] = acc[r];
}

hipError_t launch_mfma_bf16_probe(const uint16_t* a, const uint16_t* b, float* out,
                                  hipStream_t s) {
  hipLaunchKernelGGL(mfma_bf16_probe_kernel, dim3(1), dim3(64), 0, s, a, b, out);
  return hipGetLastError();
}

}  // namespace s2h
