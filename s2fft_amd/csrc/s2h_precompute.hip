// Stored Wigner-d tables and the dense contractions of the precompute path.
//
// The reference's precompute path builds a kernel K[t][l][m] =
// d^l_{m,-spin}(theta_t) sqrt((2l+1)/4pi) (x forward quadrature weights, x
// HEALPix ring phases) once (s2fft/precompute_transforms/construct.py:16-161)
// and applies it as einsum "tlm,tm->lm" (forward) or "tlm,lm->tm" (inverse)
// (s2fft/precompute_transforms/spherical.py:87-149, 314-380); its
// compute_all_slices (s2fft/recursions/price_mcewen.py:377-643) returns the
// bare d-values dl[L-1+m][t][l].  Both tables come from one builder here: a
// lane per (m, theta) walks the l-recursion of the latitudinal kernels with a
// plan's own coefficient tables and writes every l of its (m, theta) through
// arbitrary strides; for the kernel layout [t][l][m] consecutive lanes (m)
// store consecutive addresses.  The contractions are FP64 VALU, one lane per
// output (l or theta, m) over a chunk of maps: HBM-bound on the table reads,
// which is what this path trades for its O(L^3) memory.
#include "s2h_kernels.h"
#include "s2h_lat_common.h"

namespace s2h {

namespace {

constexpr int kPcMaps = 8;   // maps per lane in the contractions
constexpr int kRescale = 8;  // steps between rescale checks

__global__ void __launch_bounds__(kThreads) dtable_kernel(DTableArgs a) {
  const int mi = blockIdx.x * kThreads + threadIdx.x;
  const int t = blockIdx.y;
  if (mi >= a.nm) return;
  const int m = a.m_lo + mi;
  const int am = m < 0 ? -m : m;
  const double th = a.theta[t];
  double x = cos(th), c2 = cos(0.5 * th), s2 = sin(0.5 * th);
  // the poles exactly (cos(pi/2) rounds to 6e-17, not 0): d^l_{mn}(0) and
  // d^l_{mn}(pi) vanish unless m = n or m = -n
  if (fabs(th) < 1e-14) {
    x = 1.0;
    c2 = 1.0;
    s2 = 0.0;
  } else if (fabs(th - M_PI) < 1e-14) {
    x = -1.0;
    c2 = 0.0;
    s2 = 1.0;
  }
  const MStart st = a.mstart[mi];
  const Coef* __restrict__ cf = a.coef + (int64_t)am * a.L;
  const double bsign = m < 0 ? -1.0 : 1.0;
  double fac = a.scale * (a.wt ? a.wt[t] : 1.0);
  double pre = 0.0, pim = 0.0;  // e^{i phase_sign m phi0}
  if (a.cplx) {
    const double ang = a.phi0 ? a.phase_sign * (double)m * a.phi0[t] : 0.0;
    sincos(ang, &pim, &pre);
  }
  const int64_t base = (int64_t)mi * a.sm + (int64_t)t * a.st;
  const auto put = [&](int l, double v) {
    const int64_t o = base + (int64_t)l * a.sl;
    if (a.cplx) {
      cplx* out = static_cast<cplx*>(a.out);
      out[o] = {v * pre, v * pim};
    } else {
      static_cast<double*>(a.out)[o] = v;
    }
  };
  const int l0 = st.ell0 < a.L ? st.ell0 : a.L;
  for (int l = 0; l < l0; ++l) put(l, 0.0);
  double d, dm1 = 0.0;
  int lscale;
  lat::start_value(st, c2, s2, d, lscale);
  for (int l = l0; l < a.L; ++l) {
    put(l, lscale == 0 ? d * fac * (a.lw ? a.lw[l] : 1.0) : 0.0);
    const lat::ChunkCoef c = lat::load_coef(cf, l, 0, a.L, bsign);
    const double dn = lat::rec_step(c, x, d, dm1);
    dm1 = d;
    d = dn;
    if (((l - l0) % kRescale) == kRescale - 1) lat::rescale(d, dm1, lscale);
  }
}

template <typename K>
__device__ __forceinline__ cplx kval(const K* k, int64_t i);
template <>
__device__ __forceinline__ cplx kval<double>(const double* k, int64_t i) {
  return {k[i], 0.0};
}
template <>
__device__ __forceinline__ cplx kval<cplx>(const cplx* k, int64_t i) {
  return k[i];
}

// flm[b][l][L-1+m] = scale sum_t K[t][l][mi] Y[b][t][mi]
template <typename K>
__global__ void __launch_bounds__(kThreads) pc_forward_kernel(PcArgs a) {
  const int mi = blockIdx.x * kThreads + threadIdx.x;
  const int l = blockIdx.y;
  const int b0 = blockIdx.z * kPcMaps;
  if (mi >= a.nm) return;
  const K* __restrict__ kk = static_cast<const K*>(a.kernel);
  cplx acc[kPcMaps];
#pragma unroll
  for (int j = 0; j < kPcMaps; ++j) acc[j] = {0.0, 0.0};
  const int nb = min(kPcMaps, a.batch - b0);
  for (int t = 0; t < a.T; ++t) {
    const cplx k = kval<K>(kk, ((int64_t)t * a.L + l) * a.nm + mi);
#pragma unroll
    for (int j = 0; j < kPcMaps; ++j) {
      if (j < nb) {
        const cplx y = a.y[(int64_t)(b0 + j) * a.y_map_stride + (int64_t)t * a.nm + mi];
        acc[j].x += k.x * y.x - k.y * y.y;
        acc[j].y += k.x * y.y + k.y * y.x;
      }
    }
  }
  const int m = a.m_lo + mi;
#pragma unroll
  for (int j = 0; j < kPcMaps; ++j) {
    if (j < nb)
      a.flm[(int64_t)(b0 + j) * a.flm_map_stride + (int64_t)l * (2 * a.L - 1) + (a.L - 1 + m)] =
          {a.scale * acc[j].x, a.scale * acc[j].y};
  }
}

// Y[b][t][mi] = scale sum_l K[t][l][mi] flm[b][l][L-1+m]
template <typename K>
__global__ void __launch_bounds__(kThreads) pc_inverse_kernel(PcArgs a) {
  const int mi = blockIdx.x * kThreads + threadIdx.x;
  const int t = blockIdx.y;
  const int b0 = blockIdx.z * kPcMaps;
  if (mi >= a.nm) return;
  const int m = a.m_lo + mi;
  const K* __restrict__ kk = static_cast<const K*>(a.kernel);
  cplx acc[kPcMaps];
#pragma unroll
  for (int j = 0; j < kPcMaps; ++j) acc[j] = {0.0, 0.0};
  const int nb = min(kPcMaps, a.batch - b0);
  for (int l = 0; l < a.L; ++l) {
    const cplx k = kval<K>(kk, ((int64_t)t * a.L + l) * a.nm + mi);
#pragma unroll
    for (int j = 0; j < kPcMaps; ++j) {
      if (j < nb) {
        const cplx f = a.flm[(int64_t)(b0 + j) * a.flm_map_stride + (int64_t)l * (2 * a.L - 1) +
                             (a.L - 1 + m)];
        acc[j].x += k.x * f.x - k.y * f.y;
        acc[j].y += k.x * f.y + k.y * f.x;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < kPcMaps; ++j) {
    if (j < nb)
      a.y[(int64_t)(b0 + j) * a.y_map_stride + (int64_t)t * a.nm + mi] =
          {a.scale * acc[j].x, a.scale * acc[j].y};
  }
}

}  // namespace

hipError_t launch_dtable(const DTableArgs& a, hipStream_t s) {
  if (a.nm == 0 || a.T == 0 || a.L == 0) return hipSuccess;
  hipLaunchKernelGGL(dtable_kernel, dim3((a.nm + kThreads - 1) / kThreads, a.T), dim3(kThreads),
                     0, s, a);
  return hipGetLastError();
}

hipError_t launch_pc_forward(const PcArgs& a, hipStream_t s) {
  if (a.batch == 0 || a.nm == 0) return hipSuccess;
  const dim3 grid((a.nm + kThreads - 1) / kThreads, a.L, (a.batch + kPcMaps - 1) / kPcMaps);
  if (a.kernel_complex)
    hipLaunchKernelGGL(pc_forward_kernel<cplx>, grid, dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL(pc_forward_kernel<double>, grid, dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_pc_inverse(const PcArgs& a, hipStream_t s) {
  if (a.batch == 0 || a.nm == 0) return hipSuccess;
  const dim3 grid((a.nm + kThreads - 1) / kThreads, a.T, (a.batch + kPcMaps - 1) / kPcMaps);
  if (a.kernel_complex)
    hipLaunchKernelGGL(pc_inverse_kernel<cplx>, grid, dim3(kThreads), 0, s, a);
  else
    hipLaunchKernelGGL(pc_inverse_kernel<double>, grid, dim3(kThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace s2h
