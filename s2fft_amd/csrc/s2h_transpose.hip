// Harmonic-coefficient layout changes between the reference layout
// flm[b][l][L-1+m] and the m-major working layout flmT[b][m - m_lo][l].
//
// The latitudinal kernels walk l for a fixed m, so in the reference layout
// consecutive l of one m are (2L-1) * 16 bytes apart and every access is a
// separate cache line.  In the m-major layout a 32-l chunk of one m is 512
// contiguous bytes.  Both transposes go through a padded 32 x 32 LDS tile so
// that reads and writes are coalesced.
#include "s2h_kernels.h"

namespace s2h {

namespace {
constexpr int kTile = 32;
constexpr int kTT = 256;  // threads: 32 x 8, four elements each
}  // namespace

// flmT[b][mi][l] = flm[b][l][L-1+m_lo+mi] for mi < nm, l < L.
__global__ void __launch_bounds__(kTT) flm_to_mmajor_kernel(const cplx* flm, cplx* flmT, int L,
                                                           int m_lo, int nm) {
  __shared__ cplx tile[kTile][kTile + 1];
  const int64_t b = blockIdx.z;
  const int m0 = blockIdx.x * kTile, l0 = blockIdx.y * kTile;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const cplx* src = flm + b * (int64_t)L * (2 * L - 1);
#pragma unroll
  for (int k = 0; k < kTile; k += 8) {
    const int l = l0 + ty + k, mi = m0 + tx;
    if (l < L && mi < nm) tile[ty + k][tx] = src[(int64_t)l * (2 * L - 1) + (L - 1 + m_lo + mi)];
  }
  __syncthreads();
  cplx* dst = flmT + b * (int64_t)nm * L;
#pragma unroll
  for (int k = 0; k < kTile; k += 8) {
    const int mi = m0 + ty + k, l = l0 + tx;
    if (l < L && mi < nm) dst[(int64_t)mi * L + l] = tile[tx][ty + k];
  }
}

// flm[b][l][L-1+m_lo+mi] = flmT[b][mi][l] for mi < nm, l < L (other columns
// untouched).
__global__ void __launch_bounds__(kTT) mmajor_to_flm_kernel(const cplx* flmT, cplx* flm, int L,
                                                           int m_lo, int nm) {
  __shared__ cplx tile[kTile][kTile + 1];
  const int64_t b = blockIdx.z;
  const int m0 = blockIdx.x * kTile, l0 = blockIdx.y * kTile;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
  const cplx* src = flmT + b * (int64_t)nm * L;
#pragma unroll
  for (int k = 0; k < kTile; k += 8) {
    const int mi = m0 + ty + k, l = l0 + tx;
    if (l < L && mi < nm) tile[ty + k][tx] = src[(int64_t)mi * L + l];
  }
  __syncthreads();
  cplx* dst = flm + b * (int64_t)L * (2 * L - 1);
#pragma unroll
  for (int k = 0; k < kTile; k += 8) {
    const int l = l0 + ty + k, mi = m0 + tx;
    if (l < L && mi < nm) dst[(int64_t)l * (2 * L - 1) + (L - 1 + m_lo + mi)] = tile[tx][ty + k];
  }
}

hipError_t launch_flm_to_mmajor(const cplx* flm, cplx* flmT, int L, int m_lo, int nm, int batch,
                                hipStream_t s) {
  if (batch == 0 || nm == 0) return hipSuccess;
  dim3 grid((nm + kTile - 1) / kTile, (L + kTile - 1) / kTile, batch);
  hipLaunchKernelGGL(flm_to_mmajor_kernel, grid, dim3(kTT), 0, s, flm, flmT, L, m_lo, nm);
  return hipGetLastError();
}

hipError_t launch_mmajor_to_flm(const cplx* flmT, cplx* flm, int L, int m_lo, int nm, int batch,
                                hipStream_t s) {
  if (batch == 0 || nm == 0) return hipSuccess;
  dim3 grid((nm + kTile - 1) / kTile, (L + kTile - 1) / kTile, batch);
  hipLaunchKernelGGL(mmajor_to_flm_kernel, grid, dim3(kTT), 0, s, flmT, flm, L, m_lo, nm);
  return hipGetLastError();
}

}  // namespace s2h
