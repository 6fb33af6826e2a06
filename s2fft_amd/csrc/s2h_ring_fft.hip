// Ragged ring FFT kernels: one workgroup per (ring, map or map pair), the
// whole ring in LDS.
//
// Forward  (s2fft/utils/healpix_ffts.py:187-226 + spherical.py:533-546):
//   out[row, col] = w_t * exp(-i m phi0_t) * F_t[m mod nphi_t],  m = m_first + col
//   with F_t the unnormalised DFT of the ring; the periodic extension, the
//   quadrature weight and the ring phase are fused into the store.
// Inverse  (healpix_ffts.py:335-371 + spherical.py:183-200):
//   F_t[q] = sum_{m = q mod nphi} exp(+i m phi0_t) c(m), f_t[p] = sum_q F_t[q] e^{2 pi i q p/nphi}
//   the folding and the phase are fused into the LDS load.
//
// Rings fall in three classes launched separately so that each gets the LDS
// and registers it needs: powers of two (radix-8 Stockham passes and one
// radix 4 or 2), other lengths whose prime factors are all <= 13 (radices
// 3, 5, 7, 11, 13 as well) and the rest (Bluestein's chirp-z over a
// power-of-two convolution).  Twiddles, chirps and ring
// phases come from two-level tables in LDS (e^{i a k} = hi[k / 64] lo[k % 64])
// built once per workgroup, so no transcendental is evaluated per point.
// Two real maps share one complex transform (z = x + i y, split by the
// conjugate symmetry of the spectra), halving the transforms for real data.
#include "s2h_fft_lds.h"
#include "s2h_kernels.h"

namespace s2h {

namespace {

constexpr int kTabLo = 64;

// LDS word index of data point p: one pad word per 32 breaks the power-of-two
// strides of the early Stockham passes across the LDS banks (their writes go
// to stride R, then to runs of R at stride R^2).
__device__ __forceinline__ int lx(int p) { return p + (p >> 5); }
__host__ __device__ constexpr int lx_len(int n) { return n + (n >> 5) + 1; }

// e^{i pi num/den}-based two-level table: value(k) = hi[k >> 6] * lo[k & 63],
// value(k) = e^{i pi s k / d} for 0 <= k < count.
struct Tab2 {
  const cplx* hi;
  const cplx* lo;
  __device__ cplx at(int k) const { return cmul(hi[k >> 6], lo[k & 63]); }
};

// Builds hi[0..ceil(count/64)) and lo[0..64) for e^{i pi s k / d}.
template <int TPB>
__device__ void build_tab(cplx* hi, cplx* lo, int count, int s, int d) {
  const int nhi = (count + kTabLo - 1) / kTabLo;
  for (int k = threadIdx.x; k < nhi + kTabLo; k += TPB) {
    if (k < nhi)
      hi[k] = expi_pi((int64_t)s * kTabLo * k, d);
    else
      lo[k - nhi] = expi_pi((int64_t)s * (k - nhi), d);
  }
}

// Phase table e^{i sgn (m0 + c) phi0} for c in [0, count): double-precision
// angles (the reference forms m * phi0 the same way).
template <int TPB>
__device__ void build_phase(cplx* hi, cplx* lo, int count, int m0, double phi0_over_pi,
                            double sgn) {
  const int nhi = (count + kTabLo - 1) / kTabLo;
  for (int k = threadIdx.x; k < nhi + kTabLo; k += TPB) {
    double s, c;
    if (k < nhi) {
      sincospi(sgn * (double)(m0 + kTabLo * k) * phi0_over_pi, &s, &c);
      hi[k] = {c, s};
    } else {
      sincospi(sgn * (double)(k - nhi) * phi0_over_pi, &s, &c);
      lo[k - nhi] = {c, s};
    }
  }
}

__device__ __forceinline__ cplx conj_if(cplx v, bool c) { return c ? cplx{v.x, -v.y} : v; }

// for (p = tid; p < n; p += TPB) use(p, load(p)), with the loads of U
// consecutive iterations issued before any of their uses.  Without this each
// iteration waits out one global-memory latency, and a ring workgroup has too
// few waves beside it to hide that.
template <int TPB, int U, typename T, typename Load, typename Use>
__device__ __forceinline__ void batched_for(int n, Load load, Use use) {
  for (int p0 = threadIdx.x; p0 < n; p0 += U * TPB) {
    T v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p = p0 + u * TPB;
      v[u] = p < n ? load(p) : T{};
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const int p = p0 + u * TPB;
      if (p < n) use(p, v[u]);
    }
  }
}

// Radix-R DFT in place.  Odd R uses the symmetric form: with s_k = v_k + v_{R-k},
// d_k = v_k - v_{R-k} (k = 1..H, H = (R-1)/2) and c_m + i t_m = e^{sign 2 pi i m/R},
//   X_j = v_0 + sum_k c_{jk} s_k + i t_{jk} d_k,  X_{R-j} = the same with -i,
// so only R + 2H complex values are live.  cs[m] = (c_m, t_m), m = 1..H.
template <int R>
__device__ __forceinline__ void radix_dft(cplx* v, int sign, const cplx* cs) {
  if constexpr (R == 2 || R == 4 || R == 8) {
    SmallDft<R>::run(v, sign, nullptr);
  } else {
    constexpr int H = (R - 1) / 2;
    cplx sm[H], df[H];
    const cplx x0 = v[0];
    cplx tot = x0;
#pragma unroll
    for (int k = 1; k <= H; ++k) {
      sm[k - 1] = cadd(v[k], v[R - k]);
      df[k - 1] = csub(v[k], v[R - k]);
      tot = cadd(tot, sm[k - 1]);
    }
    v[0] = tot;
#pragma unroll
    for (int j = 1; j <= H; ++j) {
      cplx a = x0, b = {0.0, 0.0};
#pragma unroll
      for (int k = 1; k <= H; ++k) {
        const int m = (j * k) % R;
        const double c = m <= H ? cs[m].x : cs[R - m].x;
        const double t = m <= H ? cs[m].y : -cs[R - m].y;
        a.x = fma(c, sm[k - 1].x, a.x);
        a.y = fma(c, sm[k - 1].y, a.y);
        b.x = fma(t, df[k - 1].x, b.x);
        b.y = fma(t, df[k - 1].y, b.y);
      }
      v[j] = {a.x - b.y, a.y + b.x};
      v[R - j] = {a.x + b.y, a.y - b.x};
    }
  }
}

// One radix-R Stockham pass over n points (Ns = product of the radices done).
// tw.at(k) = e^{-2 pi i k / n}; the inverse (sign > 0) uses its conjugate.
// The last pass (Ns R = n) writes exactly the points each butterfly read, so
// it streams one butterfly at a time; earlier passes hold their butterflies
// in registers across the barrier between the reads and the writes.
// P2: Ns is a power of two (every pass of a power-of-two length).
template <int R, bool P2>
__device__ __forceinline__ void butterfly(cplx* v, const double* re, const double* im, int j,
                                          int nb, int Ns, int step, bool cj, int sign,
                                          const Tab2& tw, const cplx* cs) {
#pragma unroll
  for (int r = 0; r < R; ++r) v[r] = {re[lx(j + r * nb)], im[lx(j + r * nb)]};
  const int jl = P2 ? (j & (Ns - 1)) : j % Ns;
  if (jl != 0) {
    const cplx w1 = conj_if(tw.at(jl * step), cj);
    cplx wr = w1;
#pragma unroll
    for (int r = 1; r < R; ++r) {
      v[r] = cmul(v[r], wr);
      if (r + 1 < R) wr = cmul(wr, w1);
    }
  }
  radix_dft<R>(v, sign, cs);
}

template <int R, int TPB, int MAXN, bool P2>
__device__ __forceinline__ void pass_t(double* re, double* im, int n, int Ns, int sign, const Tab2& tw) {
  constexpr int MAXB = (MAXN / R + TPB - 1) / TPB;
  constexpr int H = R / 2 + 1;
  const int nb = n / R;
  const bool cj = sign > 0;
  const int step = n / (Ns * R);
  cplx cs[H];
  if constexpr (R != 2 && R != 4 && R != 8) {
#pragma unroll
    for (int k = 1; k < H; ++k) cs[k] = conj_if(tw.at(k * (n / R)), cj);
  }
  // radices 2 and 4 only ever run as the last pass (see fft_t)
  if (R == 2 || R == 4 || Ns * R == n) {
    for (int j = threadIdx.x; j < nb; j += TPB) {
      cplx v[R];
      butterfly<R, P2>(v, re, im, j, nb, Ns, step, cj, sign, tw, cs);
#pragma unroll
      for (int k = 0; k < R; ++k) {
        re[lx(j + k * Ns)] = v[k].x;
        im[lx(j + k * Ns)] = v[k].y;
      }
    }
    __syncthreads();
    return;
  }
  if constexpr (R != 2 && R != 4) {
    cplx v[MAXB][R];
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
      const int j = threadIdx.x + c * TPB;
      if (j < nb) butterfly<R, P2>(v[c], re, im, j, nb, Ns, step, cj, sign, tw, cs);
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < MAXB; ++c) {
      const int j = threadIdx.x + c * TPB;
      if (j < nb) {
        const int jl = P2 ? (j & (Ns - 1)) : j % Ns;
        const int base = (j - jl) * R + jl;
#pragma unroll
        for (int k = 0; k < R; ++k) {
          re[lx(base + k * Ns)] = v[c][k].x;
          im[lx(base + k * Ns)] = v[c][k].y;
        }
      }
    }
    __syncthreads();
  }
}

// Out-of-line pass: the smooth and Bluestein kernels call their passes so
// that each pass gets its own register allocation.
template <int R, int TPB, int MAXN, bool P2>
__device__ __noinline__ void pass_call(double* re, double* im, int n, int Ns, int sign,
                                       const Tab2 tw) {
  pass_t<R, TPB, MAXN, P2>(re, im, n, Ns, sign, tw);
}

template <int R, int TPB, int MAXN, bool INL, bool P2>
__device__ __forceinline__ void pass_any(double* re, double* im, int n, int Ns, int sign,
                                         const Tab2& tw) {
  if constexpr (INL)
    pass_t<R, TPB, MAXN, P2>(re, im, n, Ns, sign, tw);
  else
    pass_call<R, TPB, MAXN, P2>(re, im, n, Ns, sign, tw);
}

template <int TPB, int MAXN, bool POW2, bool INL>
__device__ __forceinline__ void run_pass(int R, double* re, double* im, int n, int Ns, int sign,
                                         const Tab2& tw) {
  switch (R) {
    case 8: pass_any<8, TPB, MAXN, INL, POW2>(re, im, n, Ns, sign, tw); break;
    case 4: pass_any<4, TPB, MAXN, INL, POW2>(re, im, n, Ns, sign, tw); break;
    case 2: pass_any<2, TPB, MAXN, INL, POW2>(re, im, n, Ns, sign, tw); break;
    default:
      if constexpr (!POW2) {
        switch (R) {
          case 3: pass_any<3, TPB, MAXN, INL, false>(re, im, n, Ns, sign, tw); break;
          case 5: pass_any<5, TPB, MAXN, INL, false>(re, im, n, Ns, sign, tw); break;
          case 7: pass_any<7, TPB, MAXN, INL, false>(re, im, n, Ns, sign, tw); break;
          case 11: pass_any<11, TPB, MAXN, INL, false>(re, im, n, Ns, sign, tw); break;
          case 13: pass_any<13, TPB, MAXN, INL, false>(re, im, n, Ns, sign, tw); break;
          default: break;
        }
      }
  }
}

// Mixed-radix FFT in place (prime factors <= 13, or powers of two only when
// POW2); sign -1 forward, +1 inverse.  Radices in the order: odd primes
// ascending, 8s, then the one 4 or 2 -- so a 4 or 2 is always the last
// (streamed) pass.
template <int TPB, int MAXN, bool POW2, bool INL>
__device__ void fft_t(double* re, double* im, int n, int sign, const Tab2& tw) {
  int fac[16];
  int nf = 0, rem = n;
  if (!POW2) {
    const int odd[5] = {3, 5, 7, 11, 13};
    for (int q = 0; q < 5; ++q)
      while (rem % odd[q] == 0 && nf < 16) { fac[nf++] = odd[q]; rem /= odd[q]; }
  }
  while (rem % 8 == 0 && nf < 16) { fac[nf++] = 8; rem /= 8; }
  if (rem % 4 == 0) { fac[nf++] = 4; rem /= 4; }
  if (rem % 2 == 0) { fac[nf++] = 2; rem /= 2; }
  if (rem != 1) return;  // unreachable: the host routes other lengths to Bluestein
  int Ns = 1;
  for (int i = 0; i < nf; ++i) {
    run_pass<TPB, MAXN, POW2, INL>(fac[i], re, im, n, Ns, sign, tw);
    Ns *= fac[i];
  }
}

// LDS layout of a ring workgroup: tables (576 complex) | re im, each
// lx_len(N) words (points at lx(p)); the data region is at least 2 TPB
// doubles (the inverse fold's partial sums).
// Tables: FFT twiddles (hi <= 128, lo 64), chirp e^{-i pi k/n} for k < 2n
// (hi <= 128, lo 64; Bluestein only), phase over <= 8192 values of m
// (hi <= 128, lo 64).
constexpr int kTabCplx = 576;

struct RingLds {
  double* re;
  double* im;
  cplx* tw_hi;
  cplx* tw_lo;
  cplx* ch_hi;
  cplx* ch_lo;
  cplx* ph_hi;
  cplx* ph_lo;
};

__device__ __forceinline__ RingLds ring_lds(double* smem, int N) {
  RingLds s;
  s.re = smem + 2 * kTabCplx;
  s.im = s.re + lx_len(N);
  cplx* t = reinterpret_cast<cplx*>(smem);
  s.tw_hi = t;
  s.tw_lo = t + 128;
  s.ch_hi = t + 192;
  s.ch_lo = t + 320;
  s.ph_hi = t + 384;
  s.ph_lo = t + 512;
  return s;
}

// The DFT of the n points in re/im (sign -1 forward, +1 inverse, unnormalised),
// direct or by Bluestein; tables must be built.
// KIND: kPow2 (power-of-two ring), kSmooth (prime factors <= 13), kBlue.
constexpr int kPow2 = 0, kSmooth = 1, kBlue = 2;

template <int TPB, int MAXN, int KIND>
__device__ void ring_transform(const RingLds& s, int n, int M, const cplx* chirp_hat, int sign) {
  const Tab2 tw = {s.tw_hi, s.tw_lo};
  if constexpr (KIND != kBlue) {
    fft_t<TPB, MAXN, KIND == kPow2, KIND == kPow2>(s.re, s.im, n, sign, tw);
  } else {
  const Tab2 ch = {s.ch_hi, s.ch_lo};  // e^{-i pi k / n}, k < 2n (n <= 4096: k^2 < 2^31)
  const bool cj = sign > 0;
  for (int p = threadIdx.x; p < M; p += TPB) {
    if (p < n) {
      cplx x = {s.re[lx(p)], cj ? -s.im[lx(p)] : s.im[lx(p)]};
      x = cmul(x, ch.at((int)(((unsigned)p * (unsigned)p) % (unsigned)(2 * n))));
      s.re[lx(p)] = x.x;
      s.im[lx(p)] = x.y;
    } else {
      s.re[lx(p)] = 0.0;
      s.im[lx(p)] = 0.0;
    }
  }
  __syncthreads();
  fft_t<TPB, MAXN, true, false>(s.re, s.im, M, -1, tw);
  batched_for<TPB, 8, cplx>(
      M, [&](int k) { return chirp_hat[k]; },
      [&](int k, cplx c) {
        const cplx a = cmul(cplx{s.re[lx(k)], s.im[lx(k)]}, c);
        s.re[lx(k)] = a.x;
        s.im[lx(k)] = a.y;
      });
  __syncthreads();
  fft_t<TPB, MAXN, true, false>(s.re, s.im, M, +1, tw);
  for (int k = threadIdx.x; k < n; k += TPB) {
    const cplx a = cmul(cplx{s.re[lx(k)], s.im[lx(k)]}, ch.at((int)(((unsigned)k * (unsigned)k) % (unsigned)(2 * n))));
    s.re[lx(k)] = a.x;
    s.im[lx(k)] = cj ? -a.y : a.y;
  }
  __syncthreads();
  }
}

template <int TPB, bool BLUE>
__device__ void build_fft_tables(const RingLds& s, int n, int M) {
  const int N = BLUE ? M : n;
  build_tab<TPB>(s.tw_hi, s.tw_lo, N, -2, N);        // e^{-2 pi i k / N}
  if (BLUE) build_tab<TPB>(s.ch_hi, s.ch_lo, 2 * n, -1, n);  // e^{-i pi k / n}
}

}  // namespace

// Extra scale of one ring's transform: 1, 1/n or 1/sqrt(n) (the reference's
// fft_norm modes, lib/src/s2fft_callbacks.cu:256-308).
__device__ __forceinline__ double norm_scale(int mode, int n) {
  return mode == 1 ? 1.0 / (double)n : mode == 2 ? 1.0 / sqrt((double)n) : 1.0;
}

// ------------------------------------------------------------------ forward
// gridDim.y: maps (complex input) or map pairs (real input, `paired`).
template <int TPB, int MAXN, int KIND>
__global__ void __launch_bounds__(TPB) ring_fwd_kernel(RingFwdArgs a, int job0, int paired) {
  constexpr bool BLUE = KIND == kBlue;
  extern __shared__ double smem[];
  const RingJob job = a.jobs[job0 + blockIdx.x];
  const int n = job.nphi;
  const int M = BLUE ? job.conv_len : n;
  const RingLds s = ring_lds(smem, M);
  const int64_t b0 = paired ? 2 * (int64_t)blockIdx.y : (int64_t)blockIdx.y;
  const bool two = paired && b0 + 1 < a.batch;

  build_fft_tables<TPB, BLUE>(s, n, M);
  if (a.apply_phase) build_phase<TPB>(s.ph_hi, s.ph_lo, a.ncol, a.m_first, job.phi0_over_pi, -1.0);

  const auto put = [&](int p, cplx v) {
    s.re[lx(p)] = v.x;
    s.im[lx(p)] = v.y;
  };
  if (paired) {
    // real data: z = x_b0 + i x_b0+1
    if (a.f_complex) {
      const cplx* s0 = reinterpret_cast<const cplx*>(a.f) + b0 * a.map_stride + job.pix_offset;
      batched_for<TPB, 8, cplx>(
          n, [&](int p) { return cplx{s0[p].x, two ? s0[a.map_stride + p].x : 0.0}; }, put);
    } else {
      const double* s0 = reinterpret_cast<const double*>(a.f) + b0 * a.map_stride + job.pix_offset;
      batched_for<TPB, 8, cplx>(
          n, [&](int p) { return cplx{s0[p], two ? s0[a.map_stride + p] : 0.0}; }, put);
    }
  } else if (a.in_single) {
    const bool real_only = a.real_only_all || (a.real_only_2L && n == 2 * a.L);
    const float2* src = reinterpret_cast<const float2*>(a.f) + b0 * a.map_stride + job.pix_offset;
    batched_for<TPB, 8, cplx>(
        n,
        [&](int p) {
          const float2 v = src[p];
          return cplx{(double)v.x, real_only ? 0.0 : (double)v.y};
        },
        put);
  } else {
    const bool real_only = a.real_only_all || (a.real_only_2L && n == 2 * a.L);
    const cplx* src = reinterpret_cast<const cplx*>(a.f) + b0 * a.map_stride + job.pix_offset;
    batched_for<TPB, 8, cplx>(
        n,
        [&](int p) {
          const cplx v = src[p];
          return cplx{v.x, real_only ? 0.0 : v.y};
        },
        put);
  }
  __syncthreads();
  ring_transform<TPB, MAXN, KIND>(s, n, M, a.chirp_pool + job.chirp_off, -1);

  const double w = (a.apply_weight ? job.weight : 1.0) * norm_scale(a.norm_mode, n);
  const Tab2 ph = {s.ph_hi, s.ph_lo};
  const int64_t o0 = b0 * a.out_map_stride + (int64_t)job.row * a.ncol;
  cplx* d0 = a.out + o0;
  float2* d0s = reinterpret_cast<float2*>(a.out) + o0;
  for (int col = threadIdx.x; col < a.ncol; col += TPB) {
    const int m = a.m_first + col;
    int k = m % n;
    if (k < 0) k += n;
    const bool zero = a.zero_neg_2L && n == 2 * a.L && m < 0;
    const cplx f = a.apply_phase ? ph.at(col) : cplx{1.0, 0.0};
    const cplx z = {s.re[lx(k)], s.im[lx(k)]};
    if (paired) {
      const int k2 = k == 0 ? 0 : n - k;
      const cplx zc = {s.re[lx(k2)], -s.im[lx(k2)]};
      // X = (z + zc)/2, Y = (z - zc)/(2i)
      const cplx X = {0.5 * (z.x + zc.x), 0.5 * (z.y + zc.y)};
      const cplx Y = {0.5 * (z.y - zc.y), -0.5 * (z.x - zc.x)};
      const cplx vx = cmul(cplx{X.x * w, X.y * w}, f);
      d0[col] = zero ? cplx{0.0, 0.0} : vx;
      if (two) {
        const cplx vy = cmul(cplx{Y.x * w, Y.y * w}, f);
        d0[a.out_map_stride + col] = zero ? cplx{0.0, 0.0} : vy;
      }
    } else {
      const cplx v = zero ? cplx{0.0, 0.0} : cmul(cplx{z.x * w, z.y * w}, f);
      if (a.out_single)
        d0s[col] = make_float2((float)v.x, (float)v.y);
      else
        d0[col] = v;
    }
  }
}

// ------------------------------------------------------------------ inverse
// The folded spectrum of one ring, F[q] = w sum_{m = q mod n} e^{i m phi0} c(m),
// for a map (z = c) or a pair of real maps (z = c_x + i c_y, whose inverse is
// x + i y), m over [lo, hi].
struct FoldSrc {
  const cplx* x;  // column 0 of the ring's row (m = m_first)
  const cplx* y;  // second map of a pair, or null
  const float2* xs;  // complex64 row instead of x (unpaired maps), or null
  int m_first;
  bool herm;      // m < 0 read as conj(c(-m))
  bool phase;
  Tab2 ph;        // e^{i (lo + k) phi0}
  int lo;

  __device__ __forceinline__ cplx one(const cplx* src, int m) const {
    if (herm && m < 0) {
      const cplx t = src[-m - m_first];
      return {t.x, -t.y};
    }
    return src[m - m_first];
  }
  __device__ __forceinline__ cplx one_single(int m) const {
    if (herm && m < 0) {
      const float2 t = xs[-m - m_first];
      return {(double)t.x, -(double)t.y};
    }
    const float2 t = xs[m - m_first];
    return {(double)t.x, (double)t.y};
  }
  // z(m) with the phase applied
  __device__ __forceinline__ cplx at(int m) const {
    cplx z = xs ? one_single(m) : one(x, m);
    if (y) {
      const cplx v = one(y, m);
      z = {z.x - v.y, z.y + v.x};
    }
    return phase ? cmul(z, ph.at(m - lo)) : z;
  }
};

// Writes F[q] (times w) to re/im[q], q < n: the sum over m = q (mod n), i.e.
// m = lo + r + c n with r = (q - lo) mod n, c = 0, 1, ...  The terms of one q
// are summed in increasing m, so the result does not depend on the launch.
// Two layouts:
//  n >= TPB: thread t owns q = t, t + TPB, ... and adds their terms in LDS;
//  n <  TPB: G = TPB / n threads share a q (thread t: q = t % n, g = t / n
//            takes the chunks c = g, g + G, ...), their partial sums meet in
//            LDS (part, 2 TPB doubles) in order of g.
// Both issue U loads before using them: the loads are independent, and a
// ring workgroup has too few waves to hide one latency per term.
template <int TPB>
__device__ void fold_ring(const FoldSrc& src, int n, int hi, double w, double* re, double* im,
                          double* part) {
  constexpr int U = 8;
  const int t = threadIdx.x;
  const int lo = src.lo;
  const int nm = hi - lo + 1;
  const int nc = (nm + n - 1) / n;  // chunks of n consecutive m
  const int lo_mod = ((-lo) % n + n) % n;
  const auto r_of = [&](int q) { return q + lo_mod >= n ? q + lo_mod - n : q + lo_mod; };
  if (n >= TPB) {
    for (int q = t; q < n; q += TPB) {
      re[lx(q)] = 0.0;
      im[lx(q)] = 0.0;
    }
    const int nq = (n + TPB - 1) / TPB;
    const int nslot = nq * nc;  // slot = (qi, c): q = t + qi TPB, m = lo + q + c n
    for (int s0 = 0; s0 < nslot; s0 += U) {
      cplx v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int sl = s0 + u, qi = sl / nc, q = t + qi * TPB;
        const int m = lo + r_of(q) + (sl - qi * nc) * n;
        v[u] = (sl < nslot && q < n && m <= hi) ? src.at(m) : cplx{0.0, 0.0};
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int sl = s0 + u, qi = sl / nc, q = t + qi * TPB;
        const int m = lo + r_of(q) + (sl - qi * nc) * n;
        if (sl < nslot && q < n && m <= hi) {
          re[lx(q)] += v[u].x;
          im[lx(q)] += v[u].y;
        }
      }
    }
    for (int q = t; q < n; q += TPB) {
      re[lx(q)] *= w;
      im[lx(q)] *= w;
    }
    return;
  }
  const int G = TPB / n;
  const int q = t % n, g = t / n, r = r_of(q);
  const int ns = (nc + G - 1) / G;  // chunks per thread
  cplx acc = {0.0, 0.0};
  if (g < G) {
    for (int s0 = 0; s0 < ns; s0 += U) {
      cplx v[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int m = lo + r + (g + (s0 + u) * G) * n;
        v[u] = (s0 + u < ns && m <= hi) ? src.at(m) : cplx{0.0, 0.0};
      }
#pragma unroll
      for (int u = 0; u < U; ++u) acc = cadd(acc, v[u]);
    }
    part[2 * t] = acc.x;
    part[2 * t + 1] = acc.y;
  }
  __syncthreads();
  cplx tot = {0.0, 0.0};
  if (t < n) {
    for (int gg = 0; gg < G; ++gg) tot = cadd(tot, cplx{part[2 * (gg * n + t)], part[2 * (gg * n + t) + 1]});
  }
  __syncthreads();  // part overlaps re/im
  if (t < n) {
    re[lx(t)] = tot.x * w;
    im[lx(t)] = tot.y * w;
  }
}

template <int TPB, int MAXN, int KIND>
__global__ void __launch_bounds__(TPB) ring_inv_kernel(RingInvArgs a, int job0, int paired) {
  constexpr bool BLUE = KIND == kBlue;
  extern __shared__ double smem[];
  const RingJob job = a.jobs[job0 + blockIdx.x];
  const int n = job.nphi;
  const int M = BLUE ? job.conv_len : n;
  const RingLds s = ring_lds(smem, M);
  const int64_t b0 = paired ? 2 * (int64_t)blockIdx.y : (int64_t)blockIdx.y;
  const bool two = paired && b0 + 1 < a.batch;
  const bool herm = a.herm_mode == 1 || (a.herm_mode == 2 && n == 2 * a.L);
  const int lo = herm ? -(a.L - 1) : a.m_lo;
  const int hi = herm ? (a.L - 1) : a.m_hi;

  build_fft_tables<TPB, BLUE>(s, n, M);
  if (a.apply_phase) build_phase<TPB>(s.ph_hi, s.ph_lo, hi - lo + 1, lo, job.phi0_over_pi, 1.0);
  __syncthreads();

  const double w = (a.apply_weight ? job.weight : 1.0) * norm_scale(a.norm_mode, n);
  FoldSrc fs;
  fs.x = a.in + b0 * a.in_map_stride + (int64_t)job.row * a.ncol;
  fs.y = two ? fs.x + a.in_map_stride : nullptr;
  fs.xs = a.in_single ? reinterpret_cast<const float2*>(a.in) + b0 * a.in_map_stride +
                            (int64_t)job.row * a.ncol
                      : nullptr;
  fs.m_first = a.m_first;
  fs.herm = herm;
  fs.phase = a.apply_phase;
  fs.ph = {s.ph_hi, s.ph_lo};
  fs.lo = lo;
  fold_ring<TPB>(fs, n, hi, w, s.re, s.im, s.re);
  __syncthreads();
  ring_transform<TPB, MAXN, KIND>(s, n, M, a.chirp_pool + job.chirp_off, +1);

  // store; with sub_from the result is subtracted from it (a residual)
  const int64_t off = b0 * a.map_stride + job.pix_offset;
  if (paired || !a.f_complex) {
    double* d0 = reinterpret_cast<double*>(a.f) + off;
    if (a.sub_from) {
      const double* u0 = reinterpret_cast<const double*>(a.sub_from) + off;
      batched_for<TPB, 8, cplx>(
          n, [&](int p) { return cplx{u0[p], two ? u0[a.map_stride + p] : 0.0}; },
          [&](int p, cplx u) {
            d0[p] = u.x - s.re[lx(p)];
            if (two) d0[a.map_stride + p] = u.y - s.im[lx(p)];
          });
    } else {
      for (int p = threadIdx.x; p < n; p += TPB) {
        d0[p] = s.re[lx(p)];
        if (two) d0[a.map_stride + p] = s.im[lx(p)];
      }
    }
  } else if (a.f_single) {
    float2* dst = reinterpret_cast<float2*>(a.f) + off;
    for (int p = threadIdx.x; p < n; p += TPB)
      dst[p] = make_float2((float)s.re[lx(p)], herm ? 0.0f : (float)s.im[lx(p)]);
  } else {
    cplx* dst = reinterpret_cast<cplx*>(a.f) + off;
    const cplx* u0 = a.sub_from ? reinterpret_cast<const cplx*>(a.sub_from) + off : nullptr;
    batched_for<TPB, 8, cplx>(
        n, [&](int p) { return u0 ? u0[p] : cplx{0.0, 0.0}; },
        [&](int p, cplx u) {
          const cplx v = {s.re[lx(p)], herm ? 0.0 : s.im[lx(p)]};
          dst[p] = u0 ? csub(u, v) : v;
        });
  }
}

// ------------------------------------------------------------------ aliasing
// R = G - A(Y) for one (ring, map), A = (ring FFT) . (ring iFFT) with the
// FFTs cancelled: with m = q + j_m n (q = m mod n) and w = e^{i n phi0},
//   A(Y)[m] = w_t n e^{-i m phi0} sum_{m' = q mod n} e^{i m' phi0} c(m')
//           = w_t n w^{-j_m} T[q],   T[q] = sum_j w^j c(q + j n),
// m' = q + j n over [-(L-1), L-1] (c(m') = conj c(-m') for Hermitian spectra).
// On the supported grids n phi0 / pi is an integer, so w = +-1 (sgn).
// Rings with n >= 2L - 1 have one m' per q: R = G - w_t n Y elementwise.
// Otherwise thread t owns the residue q = t (mod n): with n >= TPB it sums
// T[q] itself; with n < TPB the G = TPB / n threads of a residue split the j
// (j = g, g + G, ... from the first) and meet in LDS in order of g.  Each
// thread then writes the outputs of its residue (the same split over j).
__device__ __forceinline__ cplx alias_term(const cplx* y, int m, int m_first, bool herm) {
  if (herm && m < 0) {
    const cplx t = y[-m - m_first];
    return {t.x, -t.y};
  }
  return y[m - m_first];
}

// floor(a / n) by a multiply-high with mag = floor((2^32 - 1) / n) + 1: exact
// for 0 <= a < 2^16 and 0 < n < 2^16 (the error a (n mag - 2^32) / 2^32 n stays
// below 1/n), so one division per thread replaces one per residue
struct FastDiv {
  int n;
  unsigned mag;
  __device__ explicit FastDiv(int n_) : n(n_), mag(n_ > 1 ? 0xFFFFFFFFu / (unsigned)n_ + 1u : 0u) {}
  __device__ __forceinline__ int floor(int a) const {
    if (n == 1) return a;
    return a >= 0 ? (int)__umulhi((unsigned)a, mag) : -(int)__umulhi((unsigned)(n - 1 - a), mag);
  }
};

// DOT: instead of writing R, the workgroup adds up
//   sum_m c_m Re(conj(Y[m]) A(Y)[m]),   c_m = 2 for m > 0 of Hermitian spectra, else 1,
// and stores it in a.part[map * jobs + job]: summed over rings this is
// <p, Lat_fwd(A(Lat_inv p))> for Y = Lat_inv(p) (the latitudinal transforms
// are transposes of each other), the curvature a conjugate-gradient step needs
// without projecting A(Y).  With a.alpha the written residual is
// R = G - alpha[map] A(Y); with a.part in that mode the workgroup also stores
// the sum of squares of the R it wrote (re^2 + im^2 over its columns).
template <int TPB, bool DOT>
__global__ void __launch_bounds__(TPB) ring_alias_kernel(RingAliasArgs a) {
  __shared__ double part[2 * TPB];
  const RingJob job = a.jobs[blockIdx.x];
  const int n = job.nphi;
  const int64_t off = (int64_t)blockIdx.y * a.map_stride + (int64_t)job.row * a.ncol;
  const cplx* __restrict__ y = a.y + off;
  const cplx* g = a.g + off;
  cplx* r = a.r + off;
  const double wn = job.weight * n * (a.alpha && !DOT ? a.alpha[blockIdx.y] : 1.0);
  const int lo = -(a.L - 1), hi = a.L - 1;
  const bool herm = a.herm != 0;
  const int t = threadIdx.x;
  double dsum = 0.0;
  if (n > hi - lo) {
    // no aliasing: the column's own m is the only term
    if constexpr (DOT) {
      batched_for<TPB, 8, cplx>(
          a.ncol, [&](int col) { return y[col]; },
          [&](int col, cplx yv) {
            const int m = a.m_first + col;
            if (m >= lo && m <= hi)
              dsum += (herm && m > 0 ? 2.0 * wn : wn) * (yv.x * yv.x + yv.y * yv.y);
          });
    } else {
      batched_for<TPB, 8, cplx>(
          a.ncol, [&](int col) { return g[col]; },
          [&](int col, cplx gv) {
            const int m = a.m_first + col;
            cplx v = {0.0, 0.0};
            if (m >= lo && m <= hi) v = alias_term(y, m, a.m_first, herm);
            const cplx rv = {gv.x - wn * v.x, gv.y - wn * v.y};
            r[col] = rv;
            dsum = fma(rv.x, rv.x, fma(rv.y, rv.y, dsum));
          });
      if (!a.part) return;
    }
  } else if (herm && a.m_first == 0 && n >= a.L) {
    // Hermitian spectra stored for m in [0, L) on a ring with L <= n < 2L - 1:
    // the only alias of m is m - n, whose coefficient is conj Y[n - m], so
    // A(Y)[m] = w n (Y[m] + w^{-1} conj Y[n - m]) for m >= n - L + 1, else w n Y[m]
    const int e = (int)rint(n * job.phi0_over_pi);
    const double sgn = (a.apply_phase && (e & 1)) ? -1.0 : 1.0;
    const int m_pair = n - a.L + 1;
    for (int m = t; m < a.ncol; m += TPB) {
      const cplx yv = y[m];
      cplx av = yv;
      if (m >= m_pair) {
        const cplx yp = y[n - m];
        av = {fma(sgn, yp.x, av.x), fma(-sgn, yp.y, av.y)};
      }
      if constexpr (DOT) {
        dsum += (m > 0 ? 2.0 * wn : wn) * (yv.x * av.x + yv.y * av.y);
      } else {
        const cplx gv = g[m];
        const cplx rv = {gv.x - wn * av.x, gv.y - wn * av.y};
        r[m] = rv;
        dsum = fma(rv.x, rv.x, fma(rv.y, rv.y, dsum));
      }
    }
    if (!DOT && !a.part) return;
  } else {
    const int e = (int)rint(n * job.phi0_over_pi);
    const double sgn = (a.apply_phase && (e & 1)) ? -1.0 : 1.0;
    // the division's operands: |q - lo|, |hi - q|, |q - m_first| and
    // |m_last - q| are below n + 2L (the host keeps n + 2L < 2^16)
    const FastDiv dv(n);
    const int G = n >= TPB ? 1 : TPB / n;
    const int nq = n >= TPB ? (n + TPB - 1) / TPB : 1;  // residues per thread
    const int gi = n >= TPB ? 0 : t / n;
    const bool active = gi < G;
    const int m_last = a.m_first + a.ncol - 1;
    for (int qi = 0; qi < nq; ++qi) {
      const int q = n >= TPB ? t + qi * TPB : t % n;
      const bool own = active && q < n;
      // T[q]: j from j0 = ceil((lo - q) / n) to floor((hi - q) / n), this
      // thread's j = j0 + gi + k G
      cplx acc = {0.0, 0.0};
      int j0 = 0;
      if (own) {
        j0 = -dv.floor(q - lo);
        const int j1 = dv.floor(hi - q);
        const int cnt = j1 - j0 + 1;
        for (int k0 = gi; k0 < cnt; k0 += 4 * G) {
          cplx v[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * G;
            v[u] = k < cnt ? alias_term(y, q + (j0 + k) * n, a.m_first, herm) : cplx{0.0, 0.0};
          }
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const int k = k0 + u * G;
            const double sj = ((j0 + k) & 1) && sgn < 0 ? -1.0 : 1.0;
            acc = {fma(sj, v[u].x, acc.x), fma(sj, v[u].y, acc.y)};
          }
        }
      }
      if (G > 1) {
        if (own) {
          part[2 * t] = acc.x;
          part[2 * t + 1] = acc.y;
        }
        __syncthreads();
        cplx tot = {0.0, 0.0};
        if (own)
          for (int gg = 0; gg < G; ++gg)
            tot = cadd(tot, cplx{part[2 * (gg * n + q)], part[2 * (gg * n + q) + 1]});
        __syncthreads();
        acc = tot;
      }
      if (!own) continue;
      // outputs m = q + j n in [m_first, m_last], this thread's j = jo + gi + k G
      const int jo = -dv.floor(q - a.m_first);
      const int cnt = dv.floor(m_last - q) - jo + 1;
      const cplx* src = DOT ? y : g;
      for (int k0 = gi; k0 < cnt; k0 += 4 * G) {
        cplx gv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int k = k0 + u * G;
          gv[u] = k < cnt ? src[q + (jo + k) * n - a.m_first] : cplx{0.0, 0.0};
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const int k = k0 + u * G;
          if (k < cnt) {
            const double c = (((jo + k) & 1) && sgn < 0 ? -wn : wn);
            if constexpr (DOT) {
              const int m = q + (jo + k) * n;
              dsum += (herm && m > 0 ? 2.0 * c : c) * (gv[u].x * acc.x + gv[u].y * acc.y);
            } else {
              const cplx rv = {gv[u].x - c * acc.x, gv[u].y - c * acc.y};
              r[q + (jo + k) * n - a.m_first] = rv;
              dsum = fma(rv.x, rv.x, fma(rv.y, rv.y, dsum));
            }
          }
        }
      }
    }
  }
  // residue-owned columns: the written values of R in the no-alias branch
  // when a.part is set (write mode) fall through to the same reduction
  if (DOT || a.part) {
    // fixed-order workgroup sum: waves by shuffles, then wave 0 over the waves
    for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o);
    __syncthreads();
    if ((t & 63) == 0) part[t >> 6] = dsum;
    __syncthreads();
    if (t == 0) {
      double tot = 0.0;
      for (int w = 0; w < TPB / 64; ++w) tot += part[w];
      a.part[(int64_t)blockIdx.y * gridDim.x + blockIdx.x] = tot;
    }
  }
}

// pq[b] = sum over the rings' partial sums in ring order; alpha[b] = num[b] / pq[b]
__global__ void __launch_bounds__(256) ring_dot_final_kernel(const double* part, int njobs,
                                                             const double* num, double* pq,
                                                             double* alpha) {
  __shared__ double red[4];
  const int b = blockIdx.x;
  double acc = 0.0;
  for (int k = threadIdx.x; k < njobs; k += 256) acc += part[(int64_t)b * njobs + k];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double tot = (red[0] + red[1]) + (red[2] + red[3]);
    if (pq) pq[b] = tot;
    if (alpha) alpha[b] = (num && tot != 0.0) ? num[b] / tot : 0.0;
  }
}

// ------------------------------------------------------------------ aliasing, tiled
// ring_alias_kernel's cases for real maps (Hermitian spectra stored for m in
// [0, L)) on the tiled spectra (spec_row), per row block of 8 rows:
//  * belt blocks (every row n >= 2L - 1, no alias: R = G - w n Y) by
//    ring_alias_belt_kernel: kAliasSplit workgroups per block, lane 8 cg + r
//    on row r and column group v = split NG + cg, so that a wave's access to
//    8 columns of the 8 rows is one contiguous kilobyte;
//  * the other blocks by ring_alias_polar_kernel: RW rows of the block per
//    workgroup, their spectra staged in LDS by whole columns (RW 16 coalesced
//    bytes each), then per row as in ring_alias_kernel:
//      L <= n < 2L - 1: the one alias m - n of m >= n - L + 1, whose
//        coefficient is conj Y[n - m]: A(Y)[m] = w n (Y[m] + sgn conj Y[n - m]);
//      n < L: residue sums T[q] = sum_j s_j c(q + j n) over m' in
//        [-(L-1), L-1] (c(m') = conj Y[-m'] for m' < 0, s_j = -1 for odd j
//        when sgn < 0), then A(Y)[q + j n] = s_j w n T[q];  residues of rows
//        with n below the lanes per row split their j over groups that meet
//        in LDS in group order,
//    A(Y) / (w n) written over the staged Y, and R = G - w n A written back
//    by whole columns.
// Every per-row sum runs in a fixed order, and each workgroup's partial sum
// (dot or |R|^2) is reduced in a fixed order into part[map][block][slot]
// (kAliasSlots slots per block; the unused ones are zeroed): deterministic.
constexpr int kAliasSplit = 4;
constexpr int kAliasSlots = 8;
// polar workgroups: TPB threads staging up to LDSB bytes of spectra (8
// columns per lane): 1024 threads and 128 KB (L <= 8192); the dot up to
// L = 4096 as two workgroups of 512 threads and 64 KB per CU, so that one
// stages while the other sums
constexpr int kAliasPolarTpb = 1024;
constexpr int kAliasPolarLds = 128 * 1024;
constexpr int kAliasDotTpb = 512;
constexpr int kAliasDotLds = 64 * 1024;

// rows of a block per polar workgroup: the largest power of two <= 8 with
// RW L 16 bytes within the LDS budget ldsb
__host__ __device__ inline int alias_polar_rows(int L, int ldsb) {
  int rw = kSpecBlock;
  while (rw > 1 && (int64_t)rw * L * 16 > ldsb) rw >>= 1;
  return rw;
}

// fixed-order workgroup sum of dsum into part[slot]; slot 0's workgroup
// also zeroes the slots >= nslots
template <int TPB>
__device__ __forceinline__ void alias_block_sum(double dsum, double* red, double* part, int slot,
                                                int nslots) {
  const int t = threadIdx.x;
  for (int o = 32; o > 0; o >>= 1) dsum += __shfl_xor(dsum, o);
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = dsum;
  __syncthreads();
  if (t == 0) {
    double tot = 0.0;
    for (int w = 0; w < TPB / 64; ++w) tot += red[w];
    part[slot] = tot;
    if (slot == 0)
      for (int k = nslots; k < kAliasSlots; ++k) part[k] = 0.0;
  }
}

template <int TPB, bool DOT>
__global__ void __launch_bounds__(TPB) ring_alias_belt_kernel(RingAliasTiledArgs a) {
  constexpr int NG = TPB / kSpecBlock;  // column groups per workgroup
  constexpr int NV = kAliasSplit * NG;  // column groups per row block
  constexpr int U = 4;                  // loads in flight
  __shared__ double red[TPB / 64];
  const int t = threadIdx.x, r = t & 7, cg = t >> 3;
  const int blk = a.belt_blocks[blockIdx.x], b = blockIdx.y, sp = blockIdx.z;
  const int v = sp * NG + cg;
  const int row = kSpecBlock * blk + r;
  const RingJob job = a.rowjobs[row];
  const int n = job.nphi;
  const int L = a.L;
  const int64_t tile = (int64_t)b * a.map_stride + (int64_t)blk * kSpecBlock * a.ncol;
  const cplx* __restrict__ y = a.y + tile + r;  // column m at y[8 m]
  cplx* rr = a.r + tile + r;
  const cplx* g = a.g_tiled ? a.g + tile + r
                            : a.g + (int64_t)b * a.map_stride + (int64_t)row * a.ncol;
  const int gs = a.g_tiled ? kSpecBlock : 1;
  const double wn = job.weight * n * (a.alpha && !DOT ? a.alpha[b] : 1.0);
  double dsum = 0.0;
  if (n > 0) {
    for (int m0 = v; m0 < L; m0 += U * NV) {
      cplx yv[U], gv[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int m = m0 + u * NV;
        yv[u] = gv[u] = cplx{0.0, 0.0};
        if (m < L) {
          yv[u] = y[kSpecBlock * m];
          if (!DOT) gv[u] = g[gs * m];
        }
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int m = m0 + u * NV;
        if (m >= L) continue;
        if constexpr (DOT) {
          dsum += (m > 0 ? 2.0 * wn : wn) * (yv[u].x * yv[u].x + yv[u].y * yv[u].y);
        } else {
          const cplx rv = {gv[u].x - wn * yv[u].x, gv[u].y - wn * yv[u].y};
          rr[kSpecBlock * m] = rv;
          dsum = fma(rv.x, rv.x, fma(rv.y, rv.y, dsum));
        }
      }
    }
  }
  if (DOT || a.part)
    alias_block_sum<TPB>(dsum, red,
                         a.part + (int64_t)b * a.part_map_stride + a.belt_part_off +
                             (int64_t)blockIdx.x * kAliasSlots,
                         sp, kAliasSplit);
}

template <bool DOT, int TPB, int LDSB>
__global__ void __launch_bounds__(TPB) ring_alias_polar_kernel(RingAliasTiledArgs a) {
  constexpr int kAliasMaxCols = LDSB / 16 / TPB;  // staged columns per lane
  extern __shared__ __attribute__((aligned(16))) unsigned char lds_raw[];
  cplx* ys = reinterpret_cast<cplx*>(lds_raw);  // [column m][row of the workgroup]
  __shared__ double xch[2 * TPB];                // residue groups' partial sums
  __shared__ double red[TPB / 64];
  const int L = a.L;
  const int RW = alias_polar_rows(L, LDSB);
  const int H = kSpecBlock / RW;                 // workgroups per block
  // blockIdx.x -> (block, part z): the H parts of a block on one XCD
  // (workgroups are dealt to the 8 XCDs round-robin), so that they read each
  // 128-byte line of the block through that XCD's L2 at about the same time
  const int i = blockIdx.x;
  const int z = (i >> 3) % H;
  const int bi = ((i >> 3) / H) * 8 + (i & 7);
  if (bi >= a.n_polar) return;
  const int blk = a.polar_blocks[bi], b = blockIdx.y;
  const int t = threadIdx.x;
  const int r0 = z * RW;
  const int64_t tile = (int64_t)b * a.map_stride + (int64_t)blk * kSpecBlock * a.ncol;
  const int ntot = L * RW;  // <= kAliasMaxCols TPB
  // stage Y: element (m, q) of the workgroup's rows at ys[m RW + q]; every
  // load of the thread in flight before the first LDS store (q = t % RW is
  // the thread's own: TPB is a multiple of RW)
  const int qt = t % RW;
  {
    cplx v[kAliasMaxCols];
#pragma unroll
    for (int k = 0; k < kAliasMaxCols; ++k) {
      const int idx = t + k * TPB;
      v[k] = idx < ntot ? a.y[tile + (int64_t)(idx / RW) * kSpecBlock + r0 + qt]
                        : cplx{0.0, 0.0};
    }
#pragma unroll
    for (int k = 0; k < kAliasMaxCols; ++k) {
      const int idx = t + k * TPB;
      if (idx < ntot) ys[idx] = v[k];
    }
  }
  __syncthreads();

  const int LPR = TPB / RW;                      // lanes per row
  const int rq = t / LPR, li = t - rq * LPR;
  const int row = kSpecBlock * blk + r0 + rq;
  const RingJob job = a.rowjobs[row];
  const int n = job.nphi;
  const double wn = job.weight * n * (a.alpha && !DOT ? a.alpha[b] : 1.0);
  const int e = (int)rint(n * job.phi0_over_pi);
  const double sgn = (a.apply_phase && (e & 1)) ? -1.0 : 1.0;
  auto Y = [&](int m) { return ys[m * RW + rq]; };
  const int kind = n == 0 ? 0 : n > 2 * L - 2 ? 1 : n >= L ? 2 : 3;  // none, no alias, pair, fold
  double dsum = 0.0;

  // pair rows: A(Y) / (w n) of columns li + k LPR, fold rows: T of residues
  // q = li + k LPR (n >= LPR) or of the one residue of a split, in registers
  cplx av[kAliasMaxCols];
  cplx* tq = av;
  const bool splitf = kind == 3 && n < LPR;
  const int G = splitf ? LPR / n : 1;
  const int gi = splitf ? li / n : 0;
  const int qs = splitf ? li % n : li;
  const bool owns = kind == 3 && gi < G;
  FastDiv dv(kind == 3 ? n : 1);
#pragma unroll
  for (int k = 0; k < kAliasMaxCols; ++k) av[k] = cplx{0.0, 0.0};
  if (kind == 2) {
    const int m_pair = n - L + 1;
#pragma unroll
    for (int k = 0; k < kAliasMaxCols; ++k) {
      const int m = li + k * LPR;
      if (m < L) {
        const cplx y0 = Y(m);
        cplx v = y0;
        if (m >= m_pair) {
          const cplx yp = Y(n - m);
          v = {fma(sgn, yp.x, v.x), fma(-sgn, yp.y, v.y)};
        }
        av[k] = v;
        if constexpr (DOT) dsum += (m > 0 ? 2.0 * wn : wn) * (y0.x * v.x + y0.y * v.y);
      }
    }
  } else if (owns) {
#pragma unroll
    for (int k = 0; k < kAliasMaxCols; ++k) {
      const int q = qs + k * (splitf ? 0 : LPR);
      if ((splitf && k > 0) || q >= n) continue;
      const int j0 = -dv.floor(q + L - 1);
      const int cnt = dv.floor(L - 1 - q) - j0 + 1;
      cplx acc = {0.0, 0.0};
      for (int kk = gi; kk < cnt; kk += G) {
        const int mp = q + (j0 + kk) * n;
        cplx v = Y(mp < 0 ? -mp : mp);
        if (mp < 0) v.y = -v.y;
        const double sj = ((j0 + kk) & 1) && sgn < 0 ? -1.0 : 1.0;
        acc = {fma(sj, v.x, acc.x), fma(sj, v.y, acc.y)};
      }
      tq[k] = acc;
    }
  }
  // split residues: the groups' partial sums meet in LDS in group order
  xch[2 * t] = tq[0].x;
  xch[2 * t + 1] = tq[0].y;
  __syncthreads();
  if (splitf && owns) {
    cplx T = {0.0, 0.0};
    for (int gg = 0; gg < G; ++gg) {
      const int src = rq * LPR + gg * n + qs;
      T = {T.x + xch[2 * src], T.y + xch[2 * src + 1]};
    }
    tq[0] = T;
  }
  if constexpr (DOT) {
    // a fold row's dot from its residues alone: the sum over m >= 0 of
    // c_m Re(conj Y[m] A[m]) is the sum over all m' in (-L, L) of
    // conj(c(m')) s_j T[q] (A(-m) = conj A(m) for a real ring), i.e.
    // sum_q T[q] conj(sum_j s_j c(q + j n)) = sum_q |T[q]|^2
    if (owns && (!splitf || gi == 0)) {
#pragma unroll
      for (int k = 0; k < kAliasMaxCols; ++k) {
        const int q = qs + k * (splitf ? 0 : LPR);
        if ((splitf && k > 0) || q >= n) continue;
        dsum = fma(wn, tq[k].x * tq[k].x + tq[k].y * tq[k].y, dsum);
      }
    }
  } else {
    __syncthreads();  // every read of the staged Y is done
  }

  // columns of the row's outputs: A / (w n) over the staged Y (write mode),
  // or the dot's terms c_m w n Re(conj Y[m] A[m]) (dot mode)
  auto out = [&](int m, cplx A) {
    if constexpr (DOT) {
      const cplx yv = Y(m);
      dsum += (m > 0 ? 2.0 * wn : wn) * (yv.x * A.x + yv.y * A.y);
    } else {
      ys[m * RW + rq] = A;
    }
  };
  if (kind == 1) {
    if constexpr (DOT)
      for (int m = li; m < L; m += LPR) out(m, Y(m));
  } else if (DOT) {
    // pair and fold rows: summed above
  } else if (kind == 2) {
#pragma unroll
    for (int k = 0; k < kAliasMaxCols; ++k) {
      const int m = li + k * LPR;
      if (m < L) out(m, av[k]);
    }
  } else if (owns) {
#pragma unroll
    for (int k = 0; k < kAliasMaxCols; ++k) {
      const int q = qs + k * (splitf ? 0 : LPR);
      if ((splitf && k > 0) || q >= n) continue;
      const int cnt = dv.floor(L - 1 - q) + 1;
      for (int kk = gi; kk < cnt; kk += G) {
        const double sj = ((kk & 1) && sgn < 0) ? -1.0 : 1.0;
        out(q + kk * n, cplx{sj * tq[k].x, sj * tq[k].y});
      }
    }
  }

  if constexpr (!DOT) {
    __syncthreads();
    // R = G - w n A by whole columns (this thread: row r0 + qt), the G loads
    // in flight together
    const int rw_row = kSpecBlock * blk + r0 + qt;
    const RingJob jq = a.rowjobs[rw_row];
    if (jq.nphi > 0) {
      const double c = jq.weight * jq.nphi * (a.alpha ? a.alpha[b] : 1.0);
      cplx gv[kAliasMaxCols];
#pragma unroll
      for (int k = 0; k < kAliasMaxCols; ++k) {
        const int idx = t + k * TPB;
        const int m = idx / RW;
        gv[k] = idx >= ntot ? cplx{0.0, 0.0}
                : a.g_tiled ? a.g[tile + (int64_t)m * kSpecBlock + r0 + qt]
                            : a.g[(int64_t)b * a.map_stride + (int64_t)rw_row * a.ncol + m];
      }
#pragma unroll
      for (int k = 0; k < kAliasMaxCols; ++k) {
        const int idx = t + k * TPB;
        if (idx >= ntot) continue;
        const cplx A = ys[idx];
        const cplx rv = {gv[k].x - c * A.x, gv[k].y - c * A.y};
        a.r[tile + (int64_t)(idx / RW) * kSpecBlock + r0 + qt] = rv;
        dsum = fma(rv.x, rv.x, fma(rv.y, rv.y, dsum));
      }
    }
  }
  if (DOT || a.part)
    alias_block_sum<TPB>(dsum, red,
                         a.part + (int64_t)b * a.part_map_stride + a.polar_part_off +
                             (int64_t)bi * kAliasSlots,
                         z, H);
}

// ------------------------------------------------------------------ launchers
namespace {

constexpr int kTpbPow2 = 512;
constexpr int kTpbSmooth = 256;
constexpr int kTpbBlue = 512;
constexpr int kTpbAlias = 256;
constexpr int kTabBytes = kTabCplx * sizeof(cplx);

// tables + data region (2 lx_len(N) doubles, at least 2 TPB for the fold's partials)
size_t ring_lds_bytes(int N, int tpb) {
  const int w = lx_len(N);
  const size_t nd = 2 * (size_t)(w > tpb ? w : tpb);
  return nd * sizeof(double) + kTabBytes;
}

bool fwd_paired(const RingFwdArgs& a) { return !a.f_complex || a.real_only_all; }
bool inv_paired(const RingInvArgs& a) { return a.herm_mode == 1 && !a.f_complex; }

}  // namespace

hipError_t launch_ring_fft_forward(const RingFwdArgs& a0, const JobSet& js, int batch,
                                   hipStream_t s) {
  if (batch == 0 || js.n_pow2 + js.n_smooth + js.n_blue == 0) return hipSuccess;
  RingFwdArgs a = a0;
  a.jobs = js.jobs;
  a.batch = batch;
  const int paired = fwd_paired(a) ? 1 : 0;
  const unsigned gy = paired ? (unsigned)((batch + 1) / 2) : (unsigned)batch;
  if (js.n_pow2)
    hipLaunchKernelGGL((ring_fwd_kernel<kTpbPow2, kMaxRingLen, kPow2>), dim3(js.n_pow2, gy),
                       dim3(kTpbPow2), ring_lds_bytes(js.max_pow2, kTpbPow2), s, a, 0, paired);
  if (js.n_smooth)
    hipLaunchKernelGGL((ring_fwd_kernel<kTpbSmooth, kMaxRingLen, kSmooth>),
                       dim3(js.n_smooth, gy), dim3(kTpbSmooth), ring_lds_bytes(js.max_smooth, kTpbSmooth), s,
                       a, js.n_pow2, paired);
  // Bluestein: the long convolutions (one workgroup per CU), then the short
  // ones with their own, smaller LDS size
  const int b0 = js.n_pow2 + js.n_smooth, nshort = js.n_blue - js.n_blue_long;
  if (js.n_blue_long)
    hipLaunchKernelGGL((ring_fwd_kernel<kTpbBlue, kMaxConvLen, kBlue>), dim3(js.n_blue_long, gy),
                       dim3(kTpbBlue), ring_lds_bytes(js.max_conv, kTpbBlue), s, a, b0, paired);
  if (nshort)
    hipLaunchKernelGGL((ring_fwd_kernel<kTpbBlue, kMaxConvLen, kBlue>), dim3(nshort, gy),
                       dim3(kTpbBlue), ring_lds_bytes(js.max_conv_short, kTpbBlue), s, a,
                       b0 + js.n_blue_long, paired);
  return hipGetLastError();
}

hipError_t launch_ring_fft_inverse(const RingInvArgs& a0, const JobSet& js, int batch,
                                   hipStream_t s) {
  if (batch == 0 || js.n_pow2 + js.n_smooth + js.n_blue == 0) return hipSuccess;
  RingInvArgs a = a0;
  a.jobs = js.jobs;
  a.batch = batch;
  const int paired = inv_paired(a) ? 1 : 0;
  const unsigned gy = paired ? (unsigned)((batch + 1) / 2) : (unsigned)batch;
  if (js.n_pow2)
    hipLaunchKernelGGL((ring_inv_kernel<kTpbPow2, kMaxRingLen, kPow2>), dim3(js.n_pow2, gy),
                       dim3(kTpbPow2), ring_lds_bytes(js.max_pow2, kTpbPow2), s, a, 0, paired);
  if (js.n_smooth)
    hipLaunchKernelGGL((ring_inv_kernel<kTpbSmooth, kMaxRingLen, kSmooth>),
                       dim3(js.n_smooth, gy), dim3(kTpbSmooth), ring_lds_bytes(js.max_smooth, kTpbSmooth), s,
                       a, js.n_pow2, paired);
  const int b0 = js.n_pow2 + js.n_smooth, nshort = js.n_blue - js.n_blue_long;
  if (js.n_blue_long)
    hipLaunchKernelGGL((ring_inv_kernel<kTpbBlue, kMaxConvLen, kBlue>), dim3(js.n_blue_long, gy),
                       dim3(kTpbBlue), ring_lds_bytes(js.max_conv, kTpbBlue), s, a, b0, paired);
  if (nshort)
    hipLaunchKernelGGL((ring_inv_kernel<kTpbBlue, kMaxConvLen, kBlue>), dim3(nshort, gy),
                       dim3(kTpbBlue), ring_lds_bytes(js.max_conv_short, kTpbBlue), s, a,
                       b0 + js.n_blue_long, paired);
  return hipGetLastError();
}

hipError_t launch_ring_alias(const RingAliasArgs& a, const JobSet& js, int batch, hipStream_t s) {
  const int njobs = js.n_pow2 + js.n_smooth + js.n_blue;
  if (batch == 0 || njobs == 0) return hipSuccess;
  RingAliasArgs b = a;
  b.jobs = js.jobs;
  hipLaunchKernelGGL((ring_alias_kernel<kTpbAlias, false>), dim3(njobs, batch), dim3(kTpbAlias), 0,
                     s, b);
  return hipGetLastError();
}

hipError_t launch_ring_alias_sumsq(const RingAliasArgs& a, const JobSet& js, int batch,
                                   double* part, double* sumsq, hipStream_t s) {
  const int njobs = ring_dot_jobs(js);
  if (batch == 0 || njobs == 0) return hipSuccess;
  RingAliasArgs b = a;
  b.jobs = js.jobs;
  b.part = part;
  hipLaunchKernelGGL((ring_alias_kernel<kTpbAlias, false>), dim3(njobs, batch), dim3(kTpbAlias), 0,
                     s, b);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ring_dot_final_kernel, dim3(batch), dim3(256), 0, s, part, njobs, nullptr,
                     sumsq, nullptr);
  return hipGetLastError();
}

int ring_dot_jobs(const JobSet& js) { return js.n_pow2 + js.n_smooth + js.n_blue; }
int ring_dot_tiled_parts(int nblocks) { return nblocks * kAliasSlots; }
namespace {
// both tiled alias kernels over their blocks (part: null, or the per-block slots)
hipError_t launch_alias_tiled_pair(const RingAliasTiledArgs& a, int batch, bool dot,
                                   hipStream_t s) {
  if (a.n_belt > 0) {
    const dim3 grid(a.n_belt, batch, kAliasSplit);
    if (dot)
      hipLaunchKernelGGL((ring_alias_belt_kernel<kTpbAlias, true>), grid, dim3(kTpbAlias), 0, s,
                         a);
    else
      hipLaunchKernelGGL((ring_alias_belt_kernel<kTpbAlias, false>), grid, dim3(kTpbAlias), 0,
                         s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (a.n_polar > 0) {
    if ((int64_t)a.L * 16 > kAliasPolarLds) return hipErrorInvalidValue;
    const bool small = dot && (int64_t)a.L * 16 <= kAliasDotLds;
    const int RW = alias_polar_rows(a.L, small ? kAliasDotLds : kAliasPolarLds);
    const int H = kSpecBlock / RW;
    const dim3 grid((unsigned)((a.n_polar + 7) / 8 * 8 * H), batch);
    const size_t lds = (size_t)RW * a.L * sizeof(cplx);
    constexpr int T = kAliasPolarTpb, TD = kAliasDotTpb;
    constexpr int B = kAliasPolarLds, BD = kAliasDotLds;
    if (small)
      hipLaunchKernelGGL((ring_alias_polar_kernel<true, TD, BD>), grid, dim3(TD), lds, s, a);
    else if (dot)
      hipLaunchKernelGGL((ring_alias_polar_kernel<true, T, B>), grid, dim3(T), lds, s, a);
    else
      hipLaunchKernelGGL((ring_alias_polar_kernel<false, T, B>), grid, dim3(T), lds, s, a);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

// every row block's slots, polar blocks first
void full_part_layout(RingAliasTiledArgs& b) {
  b.part_map_stride = (int64_t)b.nblocks * kAliasSlots;
  b.polar_part_off = 0;
  b.belt_part_off = b.n_polar * kAliasSlots;
}
}  // namespace

hipError_t launch_ring_alias_tiled(const RingAliasTiledArgs& a, int batch, hipStream_t s) {
  if (batch == 0 || a.nblocks == 0) return hipSuccess;
  RingAliasTiledArgs b = a;
  b.part = nullptr;
  return launch_alias_tiled_pair(b, batch, false, s);
}

hipError_t launch_ring_alias_tiled_sumsq(const RingAliasTiledArgs& a, int batch, double* part,
                                         double* sumsq, hipStream_t s) {
  if (batch == 0 || a.nblocks == 0) return hipSuccess;
  RingAliasTiledArgs b = a;
  b.part = part;
  full_part_layout(b);
  hipError_t e = launch_alias_tiled_pair(b, batch, false, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ring_dot_final_kernel, dim3(batch), dim3(256), 0, s, part,
                     a.nblocks * kAliasSlots, nullptr, sumsq, nullptr);
  return hipGetLastError();
}

hipError_t launch_ring_dot_tiled(const RingAliasTiledArgs& a, int batch, double* part,
                                 const double* num, double* pq, double* alpha, hipStream_t s) {
  if (batch == 0 || a.nblocks == 0) return hipSuccess;
  RingAliasTiledArgs b = a;
  b.part = part;
  b.alpha = nullptr;
  full_part_layout(b);
  hipError_t e = launch_alias_tiled_pair(b, batch, true, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ring_dot_final_kernel, dim3(batch), dim3(256), 0, s, part,
                     a.nblocks * kAliasSlots, num, pq, alpha);
  return hipGetLastError();
}

hipError_t launch_ring_dot_tiled_polar(const RingAliasTiledArgs& a, int batch, double* part,
                                       int n_head, const double* num, double* pq, double* alpha,
                                       hipStream_t s) {
  if (batch == 0) return hipSuccess;
  RingAliasTiledArgs b = a;
  b.part = part;
  b.alpha = nullptr;
  b.n_belt = 0;
  b.part_map_stride = n_head + (int64_t)b.n_polar * kAliasSlots;
  b.polar_part_off = n_head;
  b.belt_part_off = 0;
  hipError_t e = launch_alias_tiled_pair(b, batch, true, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ring_dot_final_kernel, dim3(batch), dim3(256), 0, s, part,
                     (int)b.part_map_stride, num, pq, alpha);
  return hipGetLastError();
}

hipError_t launch_ring_dot(const RingAliasArgs& a, const JobSet& js, int batch, double* part,
                           const double* num, double* pq, double* alpha, hipStream_t s) {
  const int njobs = ring_dot_jobs(js);
  if (batch == 0 || njobs == 0) return hipSuccess;
  RingAliasArgs b = a;
  b.jobs = js.jobs;
  b.part = part;
  b.alpha = nullptr;
  hipLaunchKernelGGL((ring_alias_kernel<kTpbAlias, true>), dim3(njobs, batch), dim3(kTpbAlias), 0,
                     s, b);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(ring_dot_final_kernel, dim3(batch), dim3(256), 0, s, part, njobs, num, pq,
                     alpha);
  return hipGetLastError();
}

hipError_t allow_large_lds_rings() {
  const int bytes = (int)ring_lds_bytes(kMaxConvLen, kTpbBlue);
  const void* fns[] = {(const void*)ring_fwd_kernel<kTpbPow2, kMaxRingLen, kPow2>,
                       (const void*)ring_fwd_kernel<kTpbSmooth, kMaxRingLen, kSmooth>,
                       (const void*)ring_fwd_kernel<kTpbBlue, kMaxConvLen, kBlue>,
                       (const void*)ring_inv_kernel<kTpbPow2, kMaxRingLen, kPow2>,
                       (const void*)ring_inv_kernel<kTpbSmooth, kMaxRingLen, kSmooth>,
                       (const void*)ring_inv_kernel<kTpbBlue, kMaxConvLen, kBlue>};
  for (const void* f : fns) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e != hipSuccess) return e;
  }
  constexpr int T = kAliasPolarTpb, TD = kAliasDotTpb;
  constexpr int B = kAliasPolarLds, BD = kAliasDotLds;
  const void* alias[] = {(const void*)ring_alias_polar_kernel<true, T, B>,
                         (const void*)ring_alias_polar_kernel<false, T, B>,
                         (const void*)ring_alias_polar_kernel<true, TD, BD>};
  const int lds[] = {B, B, BD};
  for (int i = 0; i < 3; ++i) {
    const hipError_t e =
        hipFuncSetAttribute(alias[i], hipFuncAttributeMaxDynamicSharedMemorySize, lds[i]);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace s2h
