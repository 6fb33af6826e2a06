// Ragged ring FFT kernels: one workgroup per (ring, map), the whole ring in LDS.
//
// Forward  (s2fft/utils/healpix_ffts.py:187-226 + spherical.py:533-546):
//   out[row, col] = w_t * exp(-i m phi0_t) * F_t[m mod nphi_t],  m = m_first + col
//   with F_t the unnormalised DFT of the ring; the periodic extension, the
//   quadrature weight and the ring phase are fused into the store.
// Inverse  (healpix_ffts.py:335-371 + spherical.py:183-200):
//   F_t[q] = sum_{m = q mod nphi} exp(+i m phi0_t) c(m), f_t[p] = sum_q F_t[q] e^{2 pi i q p/nphi}
//   the folding and the phase are fused into the LDS load.
#include "s2h_fft_lds.h"
#include "s2h_kernels.h"

namespace s2h {

constexpr int kFftMax = kMaxConvLen;

__device__ __forceinline__ void ring_dft(double* re, double* im, const RingJob& job,
                                         const cplx* chirp_pool, int sign) {
  if (job.conv_len == 0)
    fft_mixed<kFftMax>(re, im, job.nphi, sign);
  else
    fft_bluestein<kFftMax>(re, im, job.nphi, job.conv_len, chirp_pool + job.chirp_off, sign);
}

__global__ void __launch_bounds__(kThreads) ring_fft_forward_kernel(RingFwdArgs a) {
  extern __shared__ double smem[];
  const RingJob job = a.jobs[blockIdx.x];
  const int64_t b = blockIdx.y;
  const int n = job.nphi;
  const int cap = job.conv_len ? job.conv_len : n;
  double* re = smem;
  double* im = smem + cap;

  const bool real_only = !a.f_complex || a.real_only_all || (a.real_only_2L && n == 2 * a.L);
  if (a.f_complex) {
    const cplx* src = reinterpret_cast<const cplx*>(a.f) + b * a.map_stride + job.pix_offset;
    for (int p = threadIdx.x; p < n; p += kThreads) {
      const cplx v = src[p];
      re[p] = v.x;
      im[p] = real_only ? 0.0 : v.y;
    }
  } else {
    const double* src = reinterpret_cast<const double*>(a.f) + b * a.map_stride + job.pix_offset;
    for (int p = threadIdx.x; p < n; p += kThreads) {
      re[p] = src[p];
      im[p] = 0.0;
    }
  }
  __syncthreads();
  ring_dft(re, im, job, a.chirp_pool, -1);

  cplx* dst = a.out + b * a.out_map_stride + (int64_t)job.row * a.ncol;
  const double w = a.apply_weight ? job.weight : 1.0;
  for (int col = threadIdx.x; col < a.ncol; col += kThreads) {
    const int m = a.m_first + col;
    int k = m % n;
    if (k < 0) k += n;
    cplx v = {re[k] * w, im[k] * w};
    if (a.apply_phase) {
      double s, c;
      sincospi(-(double)m * job.phi0_over_pi, &s, &c);
      v = cmul(v, cplx{c, s});
    }
    if (a.zero_neg_2L && n == 2 * a.L && m < 0) v = {0.0, 0.0};
    dst[col] = v;
  }
}

__global__ void __launch_bounds__(kThreads) ring_fft_inverse_kernel(RingInvArgs a) {
  extern __shared__ double smem[];
  const RingJob job = a.jobs[blockIdx.x];
  const int64_t b = blockIdx.y;
  const int n = job.nphi;
  const int cap = job.conv_len ? job.conv_len : n;
  double* re = smem;
  double* im = smem + cap;
  const cplx* src = a.in + b * a.in_map_stride + (int64_t)job.row * a.ncol;
  const bool herm = a.herm_mode == 1 || (a.herm_mode == 2 && n == 2 * a.L);
  // folded range of m
  const int lo = herm ? -(a.L - 1) : a.m_lo;
  const int hi = herm ? (a.L - 1) : a.m_hi;
  const double w = a.apply_weight ? job.weight : 1.0;
  for (int q = threadIdx.x; q < n; q += kThreads) {
    int m = lo + (((q - lo) % n) + n) % n;
    cplx acc = {0.0, 0.0};
    for (; m <= hi; m += n) {
      cplx v;
      if (herm && m < 0) {
        const cplx t = src[-m - a.m_first];
        v = {t.x, -t.y};
      } else {
        v = src[m - a.m_first];
      }
      if (a.apply_phase) {
        double s, c;
        sincospi((double)m * job.phi0_over_pi, &s, &c);
        v = cmul(v, cplx{c, s});
      }
      acc = cadd(acc, v);
    }
    re[q] = acc.x * w;
    im[q] = acc.y * w;
  }
  __syncthreads();
  ring_dft(re, im, job, a.chirp_pool, +1);

  if (a.f_complex) {
    cplx* dst = reinterpret_cast<cplx*>(a.f) + b * a.map_stride + job.pix_offset;
    for (int p = threadIdx.x; p < n; p += kThreads) dst[p] = {re[p], herm ? 0.0 : im[p]};
  } else {
    double* dst = reinterpret_cast<double*>(a.f) + b * a.map_stride + job.pix_offset;
    for (int p = threadIdx.x; p < n; p += kThreads) dst[p] = re[p];
  }
}

// FFT_M(b)/M for Bluestein, b_j = exp(+i pi j^2 / n) for |j| < n, wrapped.
__global__ void __launch_bounds__(kThreads) chirp_setup_kernel(const int* lengths,
                                                               const int* conv_lens,
                                                               const int64_t* offs,
                                                               cplx* pool) {
  extern __shared__ double smem[];
  const int n = lengths[blockIdx.x];
  const int M = conv_lens[blockIdx.x];
  double* re = smem;
  double* im = smem + M;
  for (int j = threadIdx.x; j < M; j += kThreads) {
    cplx v = {0.0, 0.0};
    if (j < n)
      v = expi_pi((int64_t)j * j, n);
    else if (j > M - n)
      v = expi_pi((int64_t)(M - j) * (M - j), n);
    re[j] = v.x;
    im[j] = v.y;
  }
  __syncthreads();
  fft_mixed<kFftMax>(re, im, M, -1);
  cplx* dst = pool + offs[blockIdx.x];
  const double s = 1.0 / M;
  for (int j = threadIdx.x; j < M; j += kThreads) dst[j] = {re[j] * s, im[j] * s};
}

// Direct DFT over gamma (G = 2N-1 <= 15 points) for every (beta, alpha)
// sample.  Modes: 0 forward from complex f, 2 forward from real f (both give
// fban[k], k = n + N-1, n in [-(N-1), N-1]), 1 inverse to complex f from all
// 2N-1 rows, 3 inverse to real f from the rows n >= 0 only (Hermitian).
__global__ void __launch_bounds__(kThreads) gamma_dft_kernel(int N, int64_t npix, int mode,
                                                             double scale, const void* in_,
                                                             void* out_) {
  const int64_t p = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (p >= npix) return;
  const int G = 2 * N - 1;
  cplx x[15];
  if (mode == 0 || mode == 1) {
    const cplx* in = static_cast<const cplx*>(in_);
    for (int g = 0; g < G; ++g) x[g] = in[(int64_t)g * npix + p];
  } else if (mode == 2) {
    const double* in = static_cast<const double*>(in_);
    for (int g = 0; g < G; ++g) x[g] = {in[(int64_t)g * npix + p], 0.0};
  } else {
    const cplx* in = static_cast<const cplx*>(in_);
    for (int k = 0; k < N; ++k) x[k] = in[(int64_t)k * npix + p];
  }
  if (mode == 0 || mode == 2) {
    cplx* out = static_cast<cplx*>(out_);
    for (int k = 0; k < G; ++k) {
      const int n = k - (N - 1);
      cplx acc = {0.0, 0.0};
      for (int g = 0; g < G; ++g) acc = cadd(acc, cmul(x[g], expi_pi(-2 * (int64_t)n * g, G)));
      out[(int64_t)k * npix + p] = {acc.x * scale, acc.y * scale};
    }
  } else if (mode == 1) {
    cplx* out = static_cast<cplx*>(out_);
    for (int g = 0; g < G; ++g) {
      cplx acc = {0.0, 0.0};
      for (int k = 0; k < G; ++k) {
        const int n = k - (N - 1);
        acc = cadd(acc, cmul(x[k], expi_pi(2 * (int64_t)n * g, G)));
      }
      out[(int64_t)g * npix + p] = {acc.x * scale, acc.y * scale};
    }
  } else {
    double* out = static_cast<double*>(out_);
    for (int g = 0; g < G; ++g) {
      double acc = x[0].x;
      for (int n = 1; n < N; ++n) acc += 2.0 * cmul(x[n], expi_pi(2 * (int64_t)n * g, G)).x;
      out[(int64_t)g * npix + p] = acc * scale;
    }
  }
}

// flmn[N-1-n][l][L-1+m] = (-1)^(m+n) conj(flmn[N-1+n][l][L-1-m]) for n >= 1
// (s2fft/transforms/wigner.py:523-526).
__global__ void __launch_bounds__(kThreads) wigner_fill_kernel(cplx* flmn, int L, int N) {
  const int n = blockIdx.z + 1;
  const int l = blockIdx.y;
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= 2 * L - 1) return;
  const int m = c - (L - 1);
  const int64_t plane = (int64_t)L * (2 * L - 1);
  const cplx v = flmn[(N - 1 + n) * plane + (int64_t)l * (2 * L - 1) + (L - 1 - m)];
  const double s = ((m + n) & 1) ? -1.0 : 1.0;
  flmn[(N - 1 - n) * plane + (int64_t)l * (2 * L - 1) + c] = {s * v.x, -s * v.y};
}

// ------------------------------------------------------------ launchers
hipError_t launch_ring_fft_forward(const RingFwdArgs& a, int njobs, int batch, int lds_elems,
                                   hipStream_t s) {
  if (njobs == 0 || batch == 0) return hipSuccess;
  const size_t lds = 2 * sizeof(double) * (size_t)lds_elems;
  hipLaunchKernelGGL(ring_fft_forward_kernel, dim3(njobs, batch), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_ring_fft_inverse(const RingInvArgs& a, int njobs, int batch, int lds_elems,
                                   hipStream_t s) {
  if (njobs == 0 || batch == 0) return hipSuccess;
  const size_t lds = 2 * sizeof(double) * (size_t)lds_elems;
  hipLaunchKernelGGL(ring_fft_inverse_kernel, dim3(njobs, batch), dim3(kThreads), lds, s, a);
  return hipGetLastError();
}

hipError_t launch_chirp_setup(const int* lengths, const int* conv_lens, const int64_t* offs,
                              cplx* pool, int count, int max_conv, hipStream_t s) {
  if (count == 0) return hipSuccess;
  const size_t lds = 2 * sizeof(double) * (size_t)max_conv;
  hipLaunchKernelGGL(chirp_setup_kernel, dim3(count), dim3(kThreads), lds, s, lengths, conv_lens,
                     offs, pool);
  return hipGetLastError();
}

hipError_t launch_gamma_dft(int N, int64_t npix, int mode, double scale, const void* in,
                            void* out, hipStream_t s) {
  if (npix == 0) return hipSuccess;
  const int64_t blocks = (npix + kThreads - 1) / kThreads;
  hipLaunchKernelGGL(gamma_dft_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s, N, npix,
                     mode, scale, in, out);
  return hipGetLastError();
}

hipError_t launch_wigner_fill(cplx* flmn, int L, int N, hipStream_t s) {
  if (N < 2) return hipSuccess;
  hipLaunchKernelGGL(wigner_fill_kernel, dim3((2 * L - 1 + kThreads - 1) / kThreads, L, N - 1),
                     dim3(kThreads), 0, s, flmn, L, N);
  return hipGetLastError();
}

hipError_t allow_large_lds() {
  const int bytes = 2 * sizeof(double) * kFftMax;
  hipError_t e = hipFuncSetAttribute((const void*)ring_fft_forward_kernel,
                                     hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  e = hipFuncSetAttribute((const void*)ring_fft_inverse_kernel,
                          hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return e;
  return hipFuncSetAttribute((const void*)chirp_setup_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace s2h
