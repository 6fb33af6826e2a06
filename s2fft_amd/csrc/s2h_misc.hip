// Pole rows, conjugate-symmetric fill and the MW/MWSS theta resampling.
#include "s2h_fft_lds.h"
#include "s2h_kernels.h"

namespace s2h {

// ------------------------------------------------------------------ poles
// Inverse: s2fft/transforms/spherical.py:166-177.  The south-pole row holds
// only m = spin: sum_l (-1)^|l-spin| flm[l, spin]; the north-pole row only
// m = -spin: sum_l flm[l, -spin] (both with the normalisation folded in).
__global__ void __launch_bounds__(kThreads) pole_inverse_kernel(PoleArgs a) {
  const int64_t b = blockIdx.y;
  const bool south = blockIdx.x == 0;
  const int row = south ? a.row_south : a.row_north;
  if (row < 0) return;
  const int spin = pole_spin(a, b);
  const int aspin = spin < 0 ? -spin : spin;
  const int lmin = a.lmin_spin && aspin > a.lmin ? aspin : a.lmin;
  cplx* r = a.ftm + b * a.ftm_map_stride + (int64_t)row * a.ftm_ncol;
  for (int c = threadIdx.x; c < a.ftm_ncol; c += kThreads) r[c] = {0.0, 0.0};
  __syncthreads();
  if (threadIdx.x == 0) {
    const int m = south ? spin : -spin;
    const cplx* f = a.flm + b * a.flm_map_stride + (a.L - 1 + m);
    cplx acc = {0.0, 0.0};
    for (int l = lmin; l < a.L; ++l) {
      double w = a.norm[l];
      if (south && ((l - spin) & 1)) w = -w;
      const cplx v = f[(int64_t)l * (2 * a.L - 1)];
      acc.x += w * v.x;
      acc.y += w * v.y;
    }
    const int c = m - a.ftm_m_first;
    if (c >= 0 && c < a.ftm_ncol) r[c] = acc;
  }
}

// Forward: s2fft/transforms/spherical.py:562-571.  flm[l, spin] gets
// (-1)^|l-spin| ftm[south, spin]; flm[l, -spin] gets ftm[north, -spin].
__global__ void __launch_bounds__(kThreads) pole_forward_kernel(PoleArgs a) {
  const int64_t b = blockIdx.y;
  const int l = blockIdx.x * kThreads + threadIdx.x;
  const int spin = pole_spin(a, b);
  const int aspin = spin < 0 ? -spin : spin;
  const int lmin = a.lmin_spin && aspin > a.lmin ? aspin : a.lmin;
  if (l < lmin || l >= a.L) return;
  const cplx* base = a.ftm + b * a.ftm_map_stride;
  cplx* f = a.flm + b * a.flm_map_stride + (int64_t)l * (2 * a.L - 1);
  const double w = a.norm[l];
  if (a.row_south >= 0) {
    const int c = spin - a.ftm_m_first;
    if (c >= 0 && c < a.ftm_ncol) {
      const cplx v = base[(int64_t)a.row_south * a.ftm_ncol + c];
      const double s = ((l - spin) & 1) ? -w : w;
      cplx* o = f + (a.L - 1 + spin);
      o->x += s * v.x;
      o->y += s * v.y;
    }
  }
  if (a.row_north >= 0) {
    const int c = -spin - a.ftm_m_first;
    if (c >= 0 && c < a.ftm_ncol) {
      const cplx v = base[(int64_t)a.row_north * a.ftm_ncol + c];
      cplx* o = f + (a.L - 1 - spin);
      o->x += w * v.x;
      o->y += w * v.y;
    }
  }
}

// flm[l, -m] = (-1)^m conj(flm[l, m]) for m >= 1.  spherical.py:579-583.
__global__ void __launch_bounds__(kThreads) reality_fill_kernel(cplx* flm, int64_t map_stride,
                                                                int L) {
  const int64_t b = blockIdx.z;
  const int l = blockIdx.y;
  const int m = blockIdx.x * kThreads + threadIdx.x + 1;
  if (m >= L) return;
  cplx* row = flm + b * map_stride + (int64_t)l * (2 * L - 1);
  const cplx v = row[L - 1 + m];
  const double s = (m & 1) ? -1.0 : 1.0;
  row[L - 1 - m] = {s * v.x, -s * v.y};
}

hipError_t launch_pole_inverse(const PoleArgs& a, hipStream_t s) {
  if (a.batch == 0) return hipSuccess;
  hipLaunchKernelGGL(pole_inverse_kernel, dim3(2, a.batch), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_pole_forward(const PoleArgs& a, hipStream_t s) {
  if (a.batch == 0) return hipSuccess;
  hipLaunchKernelGGL(pole_forward_kernel, dim3((a.L + kThreads - 1) / kThreads, a.batch),
                     dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_reality_fill(cplx* flm, int64_t map_stride, int L, int batch, hipStream_t s) {
  if (L < 2 || batch == 0) return hipSuccess;
  hipLaunchKernelGGL(reality_fill_kernel, dim3((L - 1 + kThreads - 1) / kThreads, L, batch),
                     dim3(kThreads), 0, s, flm, map_stride, L);
  return hipGetLastError();
}

}  // namespace s2h

namespace s2h {

// Zero a buffer of n doubles with a kernel (no memset nodes, so the launch is
// safe inside a captured HIP graph).
__global__ void __launch_bounds__(kThreads) zero_kernel(double* p, int64_t n) {
  const int64_t stride = (int64_t)gridDim.x * kThreads;
  for (int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

// Zeros of the padding rows [row0, rows) of ring spectra (either layout):
// the spectra producers write ring rows only.
__global__ void __launch_bounds__(kThreads) spec_pad_zero_kernel(cplx* spec, int64_t map_stride,
                                                                 int ncol, int row0, int rows,
                                                                 int tiled) {
  cplx* base = spec + (int64_t)blockIdx.x * map_stride;
  const int n = (rows - row0) * ncol;
  const int cs = spec_cstride(tiled);
  for (int i = threadIdx.x; i < n; i += kThreads) {
    const int t = row0 + i / ncol, c = i % ncol;
    base[spec_row(t, ncol, tiled) + (int64_t)c * cs] = cplx{0.0, 0.0};
  }
}

hipError_t launch_spec_pad_zero(cplx* spec, int64_t map_stride, int ncol, int row0, int rows,
                                int tiled, int batch, hipStream_t s) {
  if (batch <= 0 || rows <= row0 || ncol <= 0) return hipSuccess;
  hipLaunchKernelGGL(spec_pad_zero_kernel, dim3(batch), dim3(kThreads), 0, s, spec, map_stride,
                     ncol, row0, rows, tiled);
  return hipGetLastError();
}

hipError_t launch_zero(void* p, int64_t ndoubles, hipStream_t s) {
  if (ndoubles <= 0) return hipSuccess;
  int64_t blocks = (ndoubles + kThreads - 1) / kThreads;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s,
                     static_cast<double*>(p), ndoubles);
  return hipGetLastError();
}

// Reference layout flm[b][l][L-1+m] <-> packed m-major layout of real maps
// (flm_p[b][moff(m) + l - m], l >= m >= 0).  Unpacking writes every entry:
// zeros where |m| > l and (-1)^m conj(flm[l, -m]) for m < 0.
// One 32 x 32 tile of (l, m >= 0) per workgroup through LDS: the packed
// reads run along l (contiguous for one m), the reference-layout writes
// along m, for both signs of m.
__global__ void __launch_bounds__(kThreads) flm_unpack_kernel(int L, const cplx* src, cplx* dst) {
  __shared__ cplx tile[32][33];  // [m_local][l_local]
  const int l0 = blockIdx.x * 32, m0 = blockIdx.y * 32;
  const int64_t b = blockIdx.z;
  const cplx* sp = src + b * ((int64_t)L * (L + 1) / 2);
  cplx* dp = dst + b * ((int64_t)L * (2 * L - 1));
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int m = m0 + ty + 8 * k, l = l0 + tx;
    cplx v = {0.0, 0.0};
    if (l < L && m <= l) v = sp[packed_moff(L, m) + l - m];
    tile[ty + 8 * k][tx] = v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int l = l0 + ty + 8 * k, m = m0 + tx;
    if (l >= L || m >= L) continue;
    const cplx t = tile[tx][ty + 8 * k];
    cplx* row = dp + (int64_t)l * (2 * L - 1) + (L - 1);
    row[m] = t;
    if (m > 0) {
      const double sg = (m & 1) ? -1.0 : 1.0;
      row[-m] = cplx{sg * t.x, -sg * t.y};
    }
  }
}

__global__ void __launch_bounds__(kThreads) flm_pack_kernel(int L, const cplx* src, cplx* dst) {
  const int64_t b = blockIdx.y;
  const int64_t P = (int64_t)L * (L + 1) / 2;
  const cplx* sp = src + b * (int64_t)L * (2 * L - 1);
  cplx* dp = dst + b * P;
  for (int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x; p < P;
       p += (int64_t)gridDim.x * kThreads) {
    int l, m;
    packed_lm(p, L, l, m);
    dp[p] = sp[(int64_t)l * (2 * L - 1) + L - 1 + m];
  }
}

hipError_t launch_flm_unpack(int L, const cplx* src, cplx* dst, int batch, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const unsigned t = (unsigned)((L + 31) / 32);
  hipLaunchKernelGGL(flm_unpack_kernel, dim3(t, t, batch), dim3(kThreads), 0, s, L, src, dst);
  return hipGetLastError();
}

hipError_t launch_flm_pack(int L, const cplx* src, cplx* dst, int batch, hipStream_t s) {
  if (batch == 0) return hipSuccess;
  const int64_t P = (int64_t)L * (L + 1) / 2;
  const int64_t blocks = (P + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(flm_pack_kernel, dim3((unsigned)(blocks < 4096 ? blocks : 4096), batch),
                     dim3(kThreads), 0, s, L, src, dst);
  return hipGetLastError();
}

// Per-map sum of squares out[b] = sum_{i < n} x[b * stride + i]^2 in two
// passes with a fixed summation order (the result does not depend on the
// launch): kSumsqBlocks contiguous slices per map, four interleaved FP64
// accumulators per thread, then the slices in order.
__device__ __forceinline__ double block_sum(double v, double* red) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kThreads / 64; ++w) t += red[w];
  return t;
}

__global__ void __launch_bounds__(kThreads) batch_sumsq_partial_kernel(const double* __restrict__ x,
                                                                        int64_t n, int64_t stride,
                                                                        double* partial) {
  __shared__ double red[kThreads / 64];
  const double* xb = x + (int64_t)blockIdx.y * stride;
  const int64_t per = (n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < n ? lo + per : n;
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int64_t i = lo + threadIdx.x;
  for (; i + 3 * kThreads < hi; i += 4 * kThreads) {
    const double v0 = xb[i], v1 = xb[i + kThreads], v2 = xb[i + 2 * kThreads],
                 v3 = xb[i + 3 * kThreads];
    a0 = fma(v0, v0, a0);
    a1 = fma(v1, v1, a1);
    a2 = fma(v2, v2, a2);
    a3 = fma(v3, v3, a3);
  }
  for (; i < hi; i += kThreads) a0 = fma(xb[i], xb[i], a0);
  const double t = block_sum((a0 + a1) + (a2 + a3), red);
  if (threadIdx.x == 0) partial[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = t;
}

__global__ void __launch_bounds__(kThreads) batch_sumsq_final_kernel(const double* partial,
                                                                      int nblk, double* out) {
  __shared__ double red[kThreads / 64];
  double a = 0.0;
  for (int k = threadIdx.x; k < nblk; k += kThreads) a += partial[(int64_t)blockIdx.x * nblk + k];
  const double t = block_sum(a, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

hipError_t launch_batch_sumsq(const double* x, int64_t n, int64_t stride, int batch,
                              double* partial, double* out, hipStream_t s) {
  if (batch <= 0) return hipSuccess;
  hipLaunchKernelGGL(batch_sumsq_partial_kernel, dim3(kSumsqBlocks, batch), dim3(kThreads), 0, s,
                     x, n, stride, partial);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(batch_sumsq_final_kernel, dim3(batch), dim3(kThreads), 0, s, partial,
                     kSumsqBlocks, out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ Krylov helpers
// Per-map inner products <z, v_i> = sum_j w_j Re(conj(z_j) v_{i,j}) of
// coefficient vectors, w_j = 1 for j < packed_L (the m = 0 block of the
// packed m-major layout) and 2 beyond it, or 1 everywhere when packed_L = 0
// (full layouts hold both signs of m).  Slices of n per workgroup, then a
// fixed-order sum of the slices: the result does not depend on scheduling.
__global__ void __launch_bounds__(kThreads) flm_dots_partial_kernel(FlmDotArgs a) {
  __shared__ double red[kThreads / 64];
  const int64_t base = (int64_t)blockIdx.y * a.stride;
  const cplx* __restrict__ z = a.z + base;
  const int64_t per = (a.n + gridDim.x - 1) / gridDim.x;
  const int64_t lo = (int64_t)blockIdx.x * per;
  const int64_t hi = lo + per < a.n ? lo + per : a.n;
  double acc[kFlmDotMax];
#pragma unroll
  for (int i = 0; i < kFlmDotMax; ++i) acc[i] = 0.0;
  for (int64_t j = lo + threadIdx.x; j < hi; j += kThreads) {
    const cplx zv = z[j];
    const double w = (a.packed_L > 0 && j >= a.packed_L) ? 2.0 : 1.0;
#pragma unroll
    for (int i = 0; i < kFlmDotMax; ++i) {
      if (i < a.nv) {
        const cplx v = a.v[i][base + j];
        acc[i] = fma(w, fma(zv.x, v.x, zv.y * v.y), acc[i]);
      }
    }
  }
  for (int i = 0; i < a.nv; ++i) {
    const double t = block_sum(acc[i], red);
    if (threadIdx.x == 0)
      a.partial[((int64_t)i * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x] = t;
  }
}

__global__ void __launch_bounds__(kThreads) flm_dots_final_kernel(const double* partial, int nblk,
                                                                  double* out) {
  __shared__ double red[kThreads / 64];
  double a = 0.0;
  for (int k = threadIdx.x; k < nblk; k += kThreads) a += partial[(int64_t)blockIdx.x * nblk + k];
  const double t = block_sum(a, red);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

hipError_t launch_flm_dots(const FlmDotArgs& a, int batch, hipStream_t s) {
  if (batch <= 0 || a.nv <= 0) return hipSuccess;
  hipLaunchKernelGGL(flm_dots_partial_kernel, dim3(kFlmDotBlocks, batch), dim3(kThreads), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(flm_dots_final_kernel, dim3(a.nv * batch), dim3(kThreads), 0, s, a.partial,
                     kFlmDotBlocks, a.out);
  return hipGetLastError();
}

// x += alpha p (x optional); p_new = z - g1 p - g2 p_prev (p_new may be z;
// p_prev, g2 optional; p_new null: only x).  A thread takes kCgU elements
// kThreads apart per pass and issues all their loads before the first store.
constexpr int kCgU = 4;

__global__ void __launch_bounds__(kThreads) cg_update_kernel(CgUpdateArgs a) {
  const int b = blockIdx.y;
  const int64_t base = (int64_t)b * a.stride;
  const double al = a.x ? a.alpha[b] : 0.0;
  const double g1 = a.p_new ? a.g1[b] : 0.0;
  const double g2 = a.p_prev ? a.g2[b] : 0.0;
  const cplx zero = {0.0, 0.0};
  for (int64_t j0 = (int64_t)blockIdx.x * (kThreads * kCgU) + threadIdx.x; j0 < a.n;
       j0 += (int64_t)gridDim.x * (kThreads * kCgU)) {
    cplx pv[kCgU], xv[kCgU], zv[kCgU], qv[kCgU];
#pragma unroll
    for (int u = 0; u < kCgU; ++u) {
      const int64_t j = j0 + u * kThreads;
      const bool in = j < a.n;
      pv[u] = in ? a.p[base + j] : zero;
      xv[u] = in && a.x ? a.x[base + j] : zero;
      zv[u] = in && a.p_new ? a.z[base + j] : zero;
      qv[u] = in && a.p_prev ? a.p_prev[base + j] : zero;
    }
#pragma unroll
    for (int u = 0; u < kCgU; ++u) {
      const int64_t j = j0 + u * kThreads;
      if (j >= a.n) continue;
      if (a.x)
        a.x[base + j] = cplx{fma(al, pv[u].x, xv[u].x), fma(al, pv[u].y, xv[u].y)};
      if (a.p_new) {
        cplx nv = {fma(-g1, pv[u].x, zv[u].x), fma(-g1, pv[u].y, zv[u].y)};
        if (a.p_prev) {
          nv.x = fma(-g2, qv[u].x, nv.x);
          nv.y = fma(-g2, qv[u].y, nv.y);
        }
        a.p_new[base + j] = nv;
      }
    }
  }
}

hipError_t launch_cg_update(const CgUpdateArgs& a, int batch, hipStream_t s) {
  if (batch <= 0 || a.n <= 0) return hipSuccess;
  const int64_t want = (a.n + kThreads * kCgU - 1) / (kThreads * kCgU);
  const unsigned gx = (unsigned)(want < 2048 ? want : 2048);
  hipLaunchKernelGGL(cg_update_kernel, dim3(gx, batch), dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

// x + sum_j alpha_j[b] p_j, the terms added in the order j = 0, 1, ... (the
// same roundings as one x += alpha_j p_j per step)
__device__ __forceinline__ cplx cg_sum(const CgFinishArgs& a, int b, int64_t idx) {
  cplx v = a.x[idx];
  for (int j = 0; j < a.nd; ++j) {
    const double al = a.alpha[j][b];
    const cplx pv = a.p[j][idx];
    v.x = fma(al, pv.x, v.x);
    v.y = fma(al, pv.y, v.y);
  }
  return v;
}

__global__ void __launch_bounds__(kThreads) cg_finish_kernel(CgFinishArgs a) {
  const int b = blockIdx.y;
  const int64_t base = (int64_t)b * a.n;
  for (int64_t j = (int64_t)blockIdx.x * kThreads + threadIdx.x; j < a.n;
       j += (int64_t)gridDim.x * kThreads)
    a.x[base + j] = cg_sum(a, b, base + j);
}

// the tiles of flm_unpack_kernel, reading x + sum alpha_j p_j
__global__ void __launch_bounds__(kThreads) cg_finish_unpack_kernel(CgFinishArgs a) {
  __shared__ cplx tile[32][33];  // [m_local][l_local]
  const int L = a.L;
  const int l0 = blockIdx.x * 32, m0 = blockIdx.y * 32;
  const int b = blockIdx.z;
  const int64_t base = (int64_t)b * a.n;
  cplx* dp = a.dst + (int64_t)b * ((int64_t)L * (2 * L - 1));
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int m = m0 + ty + 8 * k, l = l0 + tx;
    cplx v = {0.0, 0.0};
    if (l < L && m <= l) v = cg_sum(a, b, base + packed_moff(L, m) + l - m);
    tile[ty + 8 * k][tx] = v;
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int l = l0 + ty + 8 * k, m = m0 + tx;
    if (l >= L || m >= L) continue;
    const cplx t = tile[tx][ty + 8 * k];
    cplx* row = dp + (int64_t)l * (2 * L - 1) + (L - 1);
    row[m] = t;
    if (m > 0) {
      const double sg = (m & 1) ? -1.0 : 1.0;
      row[-m] = cplx{sg * t.x, -sg * t.y};
    }
  }
}

hipError_t launch_cg_finish(const CgFinishArgs& a, int batch, hipStream_t s) {
  if (batch <= 0 || a.n <= 0) return hipSuccess;
  if (a.dst) {
    const unsigned t = (unsigned)((a.L + 31) / 32);
    hipLaunchKernelGGL(cg_finish_unpack_kernel, dim3(t, t, batch), dim3(kThreads), 0, s, a);
  } else {
    const int64_t want = (a.n + kThreads * 4 - 1) / (kThreads * 4);
    const unsigned gx = (unsigned)(want < 1024 ? want : 1024);
    hipLaunchKernelGGL(cg_finish_kernel, dim3(gx, batch), dim3(kThreads), 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace s2h

namespace s2h {

// ------------------------------------------------------------------ resampling
// MW / MWSS forward resampling in (theta, m) space, one workgroup per (m, map).
// s2fft/utils/resampling.py: mw_to_mwss (248-355) followed by
// upsample_by_two_mwss (64-160).  Every step there is linear along theta per
// phi-Fourier column m (the phi-rolls of the mirrored rows are (-1)^m), so
// the whole chain runs on one column in LDS:
//   MW:   extend L rows to 2L-1 with (-1)^(m+s) mirror, DFT_{2L-1}/(2L-1),
//         phase e^{-i p pi/(2L-1)}, IDFT_{2L} -> rows 0..L of MWSS
//   MWSS: extend L+1 rows to 2L with (-1)^(m+s) mirror, DFT_{2L}/(2L),
//         zero-pad, IDFT_{4L} -> rows 0..2L of MWSS at 2L
// then the quadrature weight of each output row and, for MW, the 2L/(2L-1)
// of the phi zero-padding.
__device__ __forceinline__ void col_dft(double* re, double* im, int n, int conv, int64_t off,
                                        const cplx* pool, int sign) {
  if (conv == 0)
    fft_mixed<kMaxConvLen>(re, im, n, sign);
  else
    fft_bluestein<kMaxConvLen>(re, im, n, conv, pool + off, sign);
}

__global__ void __launch_bounds__(kThreads) resample_kernel(ResampleArgs a) {
  extern __shared__ double smem[];
  const int L = a.L;
  const int col = blockIdx.x;
  const int m = a.m_lo + col;
  const int64_t b = blockIdx.y;
  double* re = smem;
  double* im = smem + a.lds;
  const int spin_b = a.spin + a.spin_step * (int)b;
  const double msgn = (((m + spin_b) % 2) + 2) % 2 ? -1.0 : 1.0;
  const cplx* g = a.in + b * a.in_map_stride + col;

  int nrow;  // rows of the MWSS column at L now held in LDS
  if (a.mw) {
    const int n1 = 2 * L - 1;
    for (int t = threadIdx.x; t < n1; t += kThreads) {
      cplx v;
      if (t < L) {
        v = g[(int64_t)t * a.ncol];
      } else {
        const cplx u = g[(int64_t)(2 * L - 2 - t) * a.ncol];
        v = {msgn * u.x, msgn * u.y};
      }
      re[t] = v.x;
      im[t] = v.y;
    }
    __syncthreads();
    col_dft(re, im, n1, a.conv[0], a.chirp_off[0], a.pool, -1);
    // scale, phase and move bins p<0 from index n1+p to 2L+p (register staged)
    constexpr int kPer = kMaxConvLen / kThreads;
    cplx v[kPer];
    int idx[kPer];
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int k = threadIdx.x + c * kThreads;
      idx[c] = -1;
      if (k < n1) {
        const int p = k < L ? k : k - n1;
        cplx x = {re[k] / n1, im[k] / n1};
        x = cmul(x, expi_pi(-(int64_t)p, n1));
        v[c] = x;
        idx[c] = p >= 0 ? p : 2 * L + p;
      }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 2 * L; k += kThreads) {
      re[k] = 0.0;
      im[k] = 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      if (idx[c] >= 0) {
        re[idx[c]] = v[c].x;
        im[idx[c]] = v[c].y;
      }
    }
    __syncthreads();
    col_dft(re, im, 2 * L, a.conv[1], a.chirp_off[1], a.pool, +1);
    nrow = L + 1;
  } else {
    for (int t = threadIdx.x; t <= L; t += kThreads) {
      const cplx v = g[(int64_t)t * a.ncol];
      re[t] = v.x;
      im[t] = v.y;
    }
    __syncthreads();
    nrow = L + 1;
  }
  (void)nrow;
  // mirror rows L+1..2L-1 from rows L-1..1
  {
    constexpr int kPer = kMaxConvLen / kThreads;
    cplx v[kPer];
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int t = L + 1 + threadIdx.x + c * kThreads;
      if (t < 2 * L) v[c] = {msgn * re[2 * L - t], msgn * im[2 * L - t]};
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int t = L + 1 + threadIdx.x + c * kThreads;
      if (t < 2 * L) {
        re[t] = v[c].x;
        im[t] = v[c].y;
      }
    }
    __syncthreads();
  }
  col_dft(re, im, 2 * L, a.conv[1], a.chirp_off[1], a.pool, -1);
  // scale by 1/(2L), move q<0 from 2L+q to 4L+q, zero the middle
  {
    constexpr int kPer = kMaxConvLen / kThreads;
    cplx v[kPer];
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int k = threadIdx.x + c * kThreads;
      if (k < 2 * L) v[c] = {re[k] / (2 * L), im[k] / (2 * L)};
    }
    __syncthreads();
    for (int k = threadIdx.x; k < 4 * L; k += kThreads) {
      re[k] = 0.0;
      im[k] = 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int k = threadIdx.x + c * kThreads;
      if (k < 2 * L) {
        const int dst = k < L ? k : k + 2 * L;
        if (a.nyq_split && k == L) {
          // real signals: the reference upsamples the complex map and then
          // takes its real part (spherical.py:528-534); on one column that is
          // the Nyquist bin shared half/half between q = -L and q = +L
          re[3 * L] = 0.5 * v[c].x;
          im[3 * L] = 0.5 * v[c].y;
          re[L] = 0.5 * v[c].x;
          im[L] = 0.5 * v[c].y;
        } else {
          re[dst] = v[c].x;
          im[dst] = v[c].y;
        }
      }
    }
    __syncthreads();
  }
  col_dft(re, im, 4 * L, a.conv[2], a.chirp_off[2], a.pool, +1);
  cplx* out = a.out + b * a.out_map_stride + col;
  for (int t = threadIdx.x; t <= 2 * L; t += kThreads) {
    const double w = (a.weights ? a.weights[t] : 1.0) * a.phi_factor;
    out[(int64_t)t * a.ncol] = {re[t] * w, im[t] * w};
  }
}

hipError_t launch_resample(const ResampleArgs& a, int batch, hipStream_t s) {
  if (batch == 0 || a.nm == 0) return hipSuccess;
  const size_t lds = 2 * sizeof(double) * (size_t)a.lds;
  hipLaunchKernelGGL(resample_kernel, dim3(a.nm, batch), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t allow_large_lds_resample() {
  return hipFuncSetAttribute((const void*)resample_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(2 * sizeof(double) * kMaxConvLen));
}

}  // namespace s2h

namespace s2h {

// Adjoint of the conjugate fill (reality): for m >= 1
// out[l, m] = (g[l, m] + (-1)^m conj(g[l, -m])) / 2, out[l, 0] = g[l, 0].
// The 1/2 pairs with the Hermitian completion of the inverse ring kernel.
__global__ void __launch_bounds__(kThreads) reality_fold_kernel(const cplx* g, cplx* out,
                                                                int64_t map_stride, int L) {
  const int64_t b = blockIdx.z;
  const int l = blockIdx.y;
  const int m = blockIdx.x * kThreads + threadIdx.x;
  if (m >= L) return;
  const cplx* row = g + b * map_stride + (int64_t)l * (2 * L - 1);
  cplx* o = out + b * map_stride + (int64_t)l * (2 * L - 1);
  if (m == 0) {
    o[L - 1] = row[L - 1];
    return;
  }
  const cplx a = row[L - 1 + m], c = row[L - 1 - m];
  const double s = (m & 1) ? -1.0 : 1.0;
  o[L - 1 + m] = {0.5 * (a.x + s * c.x), 0.5 * (a.y - s * c.y)};
}

// flm[l, m] *= factor for m >= 1 (adjoint of the Hermitian completion).
__global__ void __launch_bounds__(kThreads) scale_pos_m_kernel(cplx* flm, int64_t map_stride,
                                                               int L, double factor) {
  const int64_t b = blockIdx.z;
  const int l = blockIdx.y;
  const int m = blockIdx.x * kThreads + threadIdx.x + 1;
  if (m >= L) return;
  cplx* o = flm + b * map_stride + (int64_t)l * (2 * L - 1) + (L - 1 + m);
  o->x *= factor;
  o->y *= factor;
}

hipError_t launch_reality_fold(const cplx* g, cplx* out, int64_t map_stride, int L, int batch,
                               hipStream_t s) {
  if (batch == 0) return hipSuccess;
  hipLaunchKernelGGL(reality_fold_kernel, dim3((L + kThreads - 1) / kThreads, L, batch),
                     dim3(kThreads), 0, s, g, out, map_stride, L);
  return hipGetLastError();
}

hipError_t launch_scale_pos_m(cplx* flm, int64_t map_stride, int L, int batch, double factor,
                              hipStream_t s) {
  if (batch == 0 || L < 2) return hipSuccess;
  hipLaunchKernelGGL(scale_pos_m_kernel, dim3((L - 1 + kThreads - 1) / kThreads, L, batch),
                     dim3(kThreads), 0, s, flm, map_stride, L, factor);
  return hipGetLastError();
}

}  // namespace s2h

namespace s2h {

// Adjoint of resample_kernel (same column-per-workgroup layout, steps in
// reverse with each step's adjoint): weights, zero-extension of the kept
// rows, DFT for every IDFT, bin extraction for every zero-pad, IDFT/n for
// every DFT/n, and folding of the mirrored rows.
__global__ void __launch_bounds__(kThreads) resample_adjoint_kernel(ResampleArgs a) {
  extern __shared__ double smem[];
  const int L = a.L;
  const int col = blockIdx.x;
  const int m = a.m_lo + col;
  const int64_t b = blockIdx.y;
  double* re = smem;
  double* im = smem + a.lds;
  const int spin_b = a.spin + a.spin_step * (int)b;
  const double msgn = (((m + spin_b) % 2) + 2) % 2 ? -1.0 : 1.0;
  constexpr int kPer = kMaxConvLen / kThreads;

  // rows 0..2L of the upsampled grid, weighted, zero-extended to 4L
  const cplx* g = a.out + b * a.out_map_stride + col;
  for (int t = threadIdx.x; t < 4 * L; t += kThreads) {
    cplx v = {0.0, 0.0};
    if (t <= 2 * L) {
      const cplx u = g[(int64_t)t * a.ncol];
      const double w = a.weights[t] * a.phi_factor;
      v = {u.x * w, u.y * w};
    }
    re[t] = v.x;
    im[t] = v.y;
  }
  __syncthreads();
  col_dft(re, im, 4 * L, a.conv[2], a.chirp_off[2], a.pool, -1);
  // keep bins q in [-L, L): index q (q >= 0) and 4L+q (q < 0) -> 2L layout
  {
    cplx v[kPer];
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int k = threadIdx.x + c * kThreads;
      if (k < 2 * L) {
        const int src = k < L ? k : k + 2 * L;
        v[c] = {re[src], im[src]};
        if (a.nyq_split && k == L) v[c] = {0.5 * (re[3 * L] + re[L]), 0.5 * (im[3 * L] + im[L])};
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int k = threadIdx.x + c * kThreads;
      if (k < 2 * L) {
        re[k] = v[c].x;
        im[k] = v[c].y;
      }
    }
    __syncthreads();
  }
  col_dft(re, im, 2 * L, a.conv[1], a.chirp_off[1], a.pool, +1);
  // 1/(2L) and fold rows L+1..2L-1 onto L-1..1
  {
    cplx v[kPer];
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int t = threadIdx.x + c * kThreads;
      if (t >= 1 && t <= L - 1) v[c] = {re[t] + msgn * re[2 * L - t], im[t] + msgn * im[2 * L - t]};
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int t = threadIdx.x + c * kThreads;
      if (t >= 1 && t <= L - 1) {
        re[t] = v[c].x;
        im[t] = v[c].y;
      }
    }
    __syncthreads();
    for (int t = threadIdx.x; t <= L; t += kThreads) {
      re[t] /= (2 * L);
      im[t] /= (2 * L);
    }
    __syncthreads();
  }
  cplx* dst = const_cast<cplx*>(a.in) + b * a.in_map_stride + col;
  if (!a.mw) {
    for (int t = threadIdx.x; t <= L; t += kThreads) dst[(int64_t)t * a.ncol] = {re[t], im[t]};
    return;
  }
  // MW part: rows 0..L zero-extended to 2L, DFT_{2L}, bins p in [-(L-1), L-1]
  for (int t = L + 1 + threadIdx.x; t < 2 * L; t += kThreads) {
    re[t] = 0.0;
    im[t] = 0.0;
  }
  __syncthreads();
  col_dft(re, im, 2 * L, a.conv[1], a.chirp_off[1], a.pool, -1);
  const int n1 = 2 * L - 1;
  {
    cplx v[kPer];
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int k = threadIdx.x + c * kThreads;
      if (k < n1) {
        const int p = k < L ? k : k - n1;
        const int src = p >= 0 ? p : 2 * L + p;
        cplx x = {re[src], im[src]};
        x = cmul(x, expi_pi((int64_t)p, n1));
        v[c] = {x.x / n1, x.y / n1};
      }
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < kPer; ++c) {
      const int k = threadIdx.x + c * kThreads;
      if (k < n1) {
        re[k] = v[c].x;
        im[k] = v[c].y;
      }
    }
    __syncthreads();
  }
  col_dft(re, im, n1, a.conv[0], a.chirp_off[0], a.pool, +1);
  for (int t = threadIdx.x; t < L; t += kThreads) {
    cplx v = {re[t], im[t]};
    if (t <= L - 2) {
      v.x += msgn * re[2 * L - 2 - t];
      v.y += msgn * im[2 * L - 2 - t];
    }
    dst[(int64_t)t * a.ncol] = v;
  }
}

hipError_t launch_resample_adjoint(const ResampleArgs& a, int batch, hipStream_t s) {
  if (batch == 0 || a.nm == 0) return hipSuccess;
  const size_t lds = 2 * sizeof(double) * (size_t)a.lds;
  hipLaunchKernelGGL(resample_adjoint_kernel, dim3(a.nm, batch), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t allow_large_lds_resample_adjoint() {
  return hipFuncSetAttribute((const void*)resample_adjoint_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize,
                             (int)(2 * sizeof(double) * kMaxConvLen));
}

}  // namespace s2h
