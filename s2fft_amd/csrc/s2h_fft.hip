// Bluestein chirp tables, the DFT over gamma of the Wigner transforms and the
// reality fill of Wigner coefficients.  (The ring FFTs live in
// s2h_ring_fft.hip.)
#include "s2h_fft_lds.h"
#include "s2h_kernels.h"

namespace s2h {

constexpr int kFftMax = kMaxConvLen;

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
// 2N-1 rows, 3 inverse to real f from the rows n >= 0 only (Hermitian),
// 4 Re of the inverse sum over the rows n >= 0 only (the adjoint of mode 2
// restricted to n >= 0: Re sum_{n >= 0} x_n e^{+2 pi i n g / G}), 5 the
// adjoint of mode 3: from real f, the N rows n >= 0 of the forward sum, the
// rows n >= 1 doubled.
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
  } else if (mode == 2 || mode == 5) {
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
  } else if (mode == 5) {
    cplx* out = static_cast<cplx*>(out_);
    for (int n = 0; n < N; ++n) {
      cplx acc = {0.0, 0.0};
      for (int g = 0; g < G; ++g) acc = cadd(acc, cmul(x[g], expi_pi(-2 * (int64_t)n * g, G)));
      const double w = (n ? 2.0 : 1.0) * scale;
      out[(int64_t)n * npix + p] = {acc.x * w, acc.y * w};
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
    const double wn = mode == 3 ? 2.0 : 1.0;
    for (int g = 0; g < G; ++g) {
      double acc = x[0].x;
      for (int n = 1; n < N; ++n) acc += wn * cmul(x[n], expi_pi(2 * (int64_t)n * g, G)).x;
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

// Adjoint of wigner_fill_kernel for a real-valued loss: the cotangent of
// each filled entry flmn[N-1-n][l][L-1+m] flows back to its source
// flmn[N-1+n][l][L-1-m] as (-1)^(m+n) conj(.): g[N-1+n][l][L-1+m] +=
// (-1)^(m+n) conj(g[N-1-n][l][L-1-m]) for n >= 1 (in place).
__global__ void __launch_bounds__(kThreads) wigner_fill_adjoint_kernel(cplx* g, int L, int N) {
  const int n = blockIdx.z + 1;
  const int l = blockIdx.y;
  const int c = blockIdx.x * kThreads + threadIdx.x;
  if (c >= 2 * L - 1) return;
  const int m = c - (L - 1);
  const int64_t plane = (int64_t)L * (2 * L - 1);
  const cplx v = g[(N - 1 - n) * plane + (int64_t)l * (2 * L - 1) + (L - 1 - m)];
  const double s = ((m + n) & 1) ? -1.0 : 1.0;
  cplx* o = g + (N - 1 + n) * plane + (int64_t)l * (2 * L - 1) + c;
  o->x += s * v.x;
  o->y -= s * v.y;
}

// ------------------------------------------------------------ launchers
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
  hipError_t e = allow_large_lds_rings();
  if (e != hipSuccess) return e;
  const int bytes = 2 * sizeof(double) * kFftMax;
  return hipFuncSetAttribute((const void*)chirp_setup_kernel,
                             hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace s2h

namespace s2h {
__global__ void __launch_bounds__(kThreads) real_to_complex_kernel(const double* in, cplx* out,
                                                                   int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (i < n) out[i] = {in[i], 0.0};
}

hipError_t launch_real_to_complex(const double* in, cplx* out, int64_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(real_to_complex_kernel, dim3((unsigned)((n + kThreads - 1) / kThreads)),
                     dim3(kThreads), 0, s, in, out, n);
  return hipGetLastError();
}

hipError_t launch_wigner_fill_adjoint(cplx* g, int L, int N, hipStream_t s) {
  if (N < 2) return hipSuccess;
  hipLaunchKernelGGL(wigner_fill_adjoint_kernel,
                     dim3((2 * L - 1 + kThreads - 1) / kThreads, L, N - 1), dim3(kThreads), 0, s,
                     g, L, N);
  return hipGetLastError();
}
}  // namespace s2h
