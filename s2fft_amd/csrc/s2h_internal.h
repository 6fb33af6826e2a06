// Internal types shared by the host plan code and the HIP kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/s2fft_amd.h"

namespace s2h {

constexpr int kThreads = 256;        // workgroup size of every kernel here
constexpr int kMaxRingLen = 4096;    // longest ring FFT held in LDS (Stockham)
constexpr int kMaxConvLen = 8192;    // longest Bluestein convolution held in LDS
constexpr int kMaxL = 4096;           // ring phase tables cover |m| < kMaxL
constexpr int kMlpHidden = 64;        // learned preconditioner: hidden width
constexpr int kMlpParams = kMlpHidden * 4 + kMlpHidden + kMlpHidden * kMlpHidden + kMlpHidden +
                           2 * kMlpHidden + 2;  // W1 b1 W2 b2 W3 b3

// One ring FFT job: a contiguous run of nphi pixels.
struct RingJob {
  int64_t pix_offset;   // first pixel within one map
  int32_t nphi;         // ring length
  int32_t row;          // row of the ftm array this ring maps to
  double phi0_over_pi;  // azimuth of the first pixel / pi
  double weight;        // quadrature weight applied by the forward FFT
  int32_t conv_len;     // 0: mixed-radix; else Bluestein length M (power of 2)
  int32_t pad;
  int64_t chirp_off;    // offset (complex elements) of FFT(chirp)/M in the chirp pool
};

// Ring jobs of one transform on the device, in three consecutive classes:
// n_pow2 power-of-two lengths, n_smooth other lengths with prime factors
// <= 13, n_blue Bluestein rings; each class sorted by decreasing length.
// Bluestein rings whose convolution is longer than this launch separately:
// the shorter ones then fit two workgroups' LDS per CU.
constexpr int kBlueSplit = 4096;

struct JobSet {
  RingJob* jobs = nullptr;
  int n_pow2 = 0, n_smooth = 0, n_blue = 0;
  int n_blue_long = 0;  // Bluestein rings with convolution > kBlueSplit (listed first)
  int max_pow2 = 0;     // longest power-of-two ring
  int max_smooth = 0;   // longest other mixed-radix ring
  int max_conv = 0;     // longest Bluestein convolution
  int max_conv_short = 0;  // longest Bluestein convolution <= kBlueSplit
  int max_nphi = 0;     // longest ring
};

// Start of the ell-recursion for one signed m.
struct MStart {
  int32_t ell0;         // first ell with d^ell_{m,n} != 0, max(|m|,|n|)
  int32_t ecos;         // exponent of cos(theta/2)
  int32_t esin;         // exponent of sin(theta/2)
  int32_t pad;
  double kmant;         // signed mantissa of the theta-independent factor
  int32_t kexp;         // its binary exponent
  int32_t pad2;
};

// d^{l+1} = (a x - b) d^l - c d^{l-1} for the normalised d (per |m|, l).
struct Coef {
  double a, b, c, pad;
};

// A ring of the latitudinal step.  Paired rings share |d| through
// d^l_{m,0}(pi - theta) = (-1)^{l+m} d^l_{m,0}(theta) (spin 0 only).
struct LatRing {
  int32_t ta;           // ftm row of this ring
  int32_t tb;           // ftm row of the mirrored ring, -1 if unpaired, ta if self
  double x;             // cos(theta_a)
  double c2;            // cos(theta_a / 2)
  double s2;            // sin(theta_a / 2)
};

struct cplx {
  double x, y;
};

// Largest prime factor of n (n >= 1).
__host__ __device__ inline int largest_prime_factor(int n) {
  int best = 1;
  for (int p = 2; p * p <= n; ++p)
    while (n % p == 0) {
      best = p;
      n /= p;
    }
  return n > 1 ? (n > best ? n : best) : best;
}

}  // namespace s2h
