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
  cplx* r = a.ftm + b * a.ftm_map_stride + (int64_t)row * a.ftm_ncol;
  for (int c = threadIdx.x; c < a.ftm_ncol; c += kThreads) r[c] = {0.0, 0.0};
  __syncthreads();
  if (threadIdx.x == 0) {
    const int m = south ? a.spin : -a.spin;
    const cplx* f = a.flm + b * a.flm_map_stride + (a.L - 1 + m);
    cplx acc = {0.0, 0.0};
    for (int l = a.lmin; l < a.L; ++l) {
      double w = a.norm[l];
      if (south && ((l - a.spin) & 1)) w = -w;
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
  if (l < a.lmin || l >= a.L) return;
  const cplx* base = a.ftm + b * a.ftm_map_stride;
  cplx* f = a.flm + b * a.flm_map_stride + (int64_t)l * (2 * a.L - 1);
  const double w = a.norm[l];
  if (a.row_south >= 0) {
    const int c = a.spin - a.ftm_m_first;
    if (c >= 0 && c < a.ftm_ncol) {
      const cplx v = base[(int64_t)a.row_south * a.ftm_ncol + c];
      const double s = ((l - a.spin) & 1) ? -w : w;
      cplx* o = f + (a.L - 1 + a.spin);
      o->x += s * v.x;
      o->y += s * v.y;
    }
  }
  if (a.row_north >= 0) {
    const int c = -a.spin - a.ftm_m_first;
    if (c >= 0 && c < a.ftm_ncol) {
      const cplx v = base[(int64_t)a.row_north * a.ftm_ncol + c];
      cplx* o = f + (a.L - 1 - a.spin);
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

hipError_t launch_zero(void* p, int64_t ndoubles, hipStream_t s) {
  if (ndoubles <= 0) return hipSuccess;
  int64_t blocks = (ndoubles + kThreads - 1) / kThreads;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(zero_kernel, dim3((unsigned)blocks), dim3(kThreads), 0, s,
                     static_cast<double*>(p), ndoubles);
  return hipGetLastError();
}

}  // namespace s2h
