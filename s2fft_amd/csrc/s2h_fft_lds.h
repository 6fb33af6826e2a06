// In-LDS FFT building blocks for one workgroup of s2h::kThreads threads.
//
// Data live in two LDS arrays (real and imaginary parts).  A transform is a
// sequence of self-sorting Stockham passes; each pass stages the radix-R
// butterflies of every thread in registers between two barriers, so one LDS
// buffer suffices.  Lengths whose prime factors are all <= 13 use these
// passes directly; other lengths use Bluestein's chirp-z algorithm (Bluestein
// 1970) over a power-of-two convolution length.
#pragma once

#include "s2h_internal.h"

namespace s2h {

__device__ __forceinline__ cplx cmul(cplx a, cplx b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
__device__ __forceinline__ cplx cadd(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__device__ __forceinline__ cplx csub(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
// multiply by sign * i
__device__ __forceinline__ cplx cmul_si(cplx a, int sign) {
  return sign > 0 ? cplx{-a.y, a.x} : cplx{a.y, -a.x};
}

// exp(i * pi * num / den) with an exactly reduced argument.
__device__ __forceinline__ cplx expi_pi(int64_t num, int64_t den) {
  int64_t r = num % (2 * den);
  if (r < 0) r += 2 * den;
  double s, c;
  sincospi((double)r / (double)den, &s, &c);
  return {c, s};
}

// ------------------------------------------------------------ small DFTs
template <int R>
struct SmallDft;

template <>
struct SmallDft<2> {
  __device__ static void run(cplx* v, int, const cplx*) {
    cplx a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};

template <>
struct SmallDft<4> {
  __device__ static void run(cplx* v, int sign, const cplx*) {
    cplx s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    cplx s13 = cadd(v[1], v[3]), d13 = cmul_si(csub(v[1], v[3]), sign);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};

template <>
struct SmallDft<8> {
  __device__ static void run(cplx* v, int sign, const cplx*) {
    cplx e[4] = {v[0], v[2], v[4], v[6]};
    cplx o[4] = {v[1], v[3], v[5], v[7]};
    SmallDft<4>::run(e, sign, nullptr);
    SmallDft<4>::run(o, sign, nullptr);
    const double h = 0.70710678118654752440;
    // W8^k = exp(sign * 2 pi i k / 8)
    cplx w1 = {h, sign * h}, w3 = {-h, sign * h};
    cplx t0 = o[0];
    cplx t1 = cmul(w1, o[1]);
    cplx t2 = cmul_si(o[2], sign);
    cplx t3 = cmul(w3, o[3]);
    v[0] = cadd(e[0], t0);
    v[4] = csub(e[0], t0);
    v[1] = cadd(e[1], t1);
    v[5] = csub(e[1], t1);
    v[2] = cadd(e[2], t2);
    v[6] = csub(e[2], t2);
    v[3] = cadd(e[3], t3);
    v[7] = csub(e[3], t3);
  }
};

// Direct DFT for the odd prime radices; w[k] = exp(sign 2 pi i k / R).
template <int R>
struct SmallDft {
  __device__ static void run(cplx* v, int, const cplx* w) {
    cplx out[R];
#pragma unroll
    for (int k = 0; k < R; ++k) {
      cplx acc = v[0];
#pragma unroll
      for (int r = 1; r < R; ++r) acc = cadd(acc, cmul(v[r], w[(r * k) % R]));
      out[k] = acc;
    }
#pragma unroll
    for (int k = 0; k < R; ++k) v[k] = out[k];
  }
};

// ------------------------------------------------------------ Stockham pass
// One radix-R pass over n points; Ns = product of the radices already done.
// Butterfly j reads in[j + r n/R], twists by exp(sign 2 pi i r (j mod Ns)/(Ns R)),
// and writes out[(j div Ns) Ns R + (j mod Ns) + k Ns].
template <int R, int MAXN>
__device__ void stockham_pass(double* re, double* im, int n, int Ns, int sign) {
  constexpr int MAXB = (MAXN / R + kThreads - 1) / kThreads;
  const int nb = n / R;
  cplx w[R];
  if (R != 2 && R != 4 && R != 8) {
#pragma unroll
    for (int k = 0; k < R; ++k) {
      double s, c;
      sincospi(2.0 * sign * k / R, &s, &c);
      w[k] = {c, s};
    }
  }
  cplx v[MAXB][R];
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    const int j = threadIdx.x + c * kThreads;
    if (j < nb) {
#pragma unroll
      for (int r = 0; r < R; ++r) v[c][r] = {re[j + r * nb], im[j + r * nb]};
      const int jl = j % Ns;
      if (jl != 0) {
        double s, cc;
        sincospi(2.0 * sign * (double)jl / (double)(Ns * R), &s, &cc);
        const cplx w1 = {cc, s};
        cplx wr = w1;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          v[c][r] = cmul(v[c][r], wr);
          if (r + 1 < R) wr = cmul(wr, w1);
        }
      }
      SmallDft<R>::run(v[c], sign, w);
    }
  }
  __syncthreads();
#pragma unroll
  for (int c = 0; c < MAXB; ++c) {
    const int j = threadIdx.x + c * kThreads;
    if (j < nb) {
      const int jl = j % Ns;
      const int base = (j - jl) * R + jl;
#pragma unroll
      for (int k = 0; k < R; ++k) {
        re[base + k * Ns] = v[c][k].x;
        im[base + k * Ns] = v[c][k].y;
      }
    }
  }
  __syncthreads();
}


// Mixed-radix FFT in place, all prime factors of n must be <= 13.
// sign = -1: X[k] = sum_p x[p] e^{-2 pi i k p/n};  sign = +1: unnormalised inverse.
template <int MAXN>
__device__ void fft_mixed(double* re, double* im, int n, int sign) {
  int rem = n, Ns = 1;
  while (rem > 1) {
    if (rem % 8 == 0) {
      stockham_pass<8, MAXN>(re, im, n, Ns, sign);
      Ns *= 8; rem /= 8;
    } else if (rem % 4 == 0) {
      stockham_pass<4, MAXN>(re, im, n, Ns, sign);
      Ns *= 4; rem /= 4;
    } else if (rem % 2 == 0) {
      stockham_pass<2, MAXN>(re, im, n, Ns, sign);
      Ns *= 2; rem /= 2;
    } else if (rem % 3 == 0) {
      stockham_pass<3, MAXN>(re, im, n, Ns, sign);
      Ns *= 3; rem /= 3;
    } else if (rem % 5 == 0) {
      stockham_pass<5, MAXN>(re, im, n, Ns, sign);
      Ns *= 5; rem /= 5;
    } else if (rem % 7 == 0) {
      stockham_pass<7, MAXN>(re, im, n, Ns, sign);
      Ns *= 7; rem /= 7;
    } else if (rem % 11 == 0) {
      stockham_pass<11, MAXN>(re, im, n, Ns, sign);
      Ns *= 11; rem /= 11;
    } else if (rem % 13 == 0) {
      stockham_pass<13, MAXN>(re, im, n, Ns, sign);
      Ns *= 13; rem /= 13;
    } else {
      return;  // unreachable for supported lengths (checked on the host)
    }
  }
}

// Bluestein DFT of the n points already in re/im[0..n), convolution length M
// (power of two, M >= 2n-1).  chirp_hat = FFT_M(b)/M with b_j = e^{+i pi j^2/n}
// wrapped cyclically.  sign as in fft_mixed (the inverse uses conjugation).
template <int MAXM>
__device__ void fft_bluestein(double* re, double* im, int n, int M, const cplx* chirp_hat,
                              int sign) {
  for (int p = threadIdx.x; p < M; p += kThreads) {
    if (p < n) {
      cplx x = {re[p], sign > 0 ? -im[p] : im[p]};
      x = cmul(x, expi_pi(-(int64_t)p * p, n));
      re[p] = x.x;
      im[p] = x.y;
    } else {
      re[p] = 0.0;
      im[p] = 0.0;
    }
  }
  __syncthreads();
  fft_mixed<MAXM>(re, im, M, -1);
  for (int k = threadIdx.x; k < M; k += kThreads) {
    cplx a = {re[k], im[k]};
    a = cmul(a, chirp_hat[k]);
    re[k] = a.x;
    im[k] = a.y;
  }
  __syncthreads();
  fft_mixed<MAXM>(re, im, M, +1);
  for (int k = threadIdx.x; k < n; k += kThreads) {
    cplx a = {re[k], im[k]};
    a = cmul(a, expi_pi(-(int64_t)k * k, n));
    re[k] = a.x;
    im[k] = sign > 0 ? -a.y : a.y;
  }
  __syncthreads();
}

}  // namespace s2h
