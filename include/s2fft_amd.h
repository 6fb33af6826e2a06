/*
 * s2fft_amd: C ABI of the MI355X (gfx950) spin spherical-harmonic transforms.
 *
 * Every entry point takes plain device pointers (FP64 / complex128 data laid
 * out exactly as the reference's NumPy arrays), a batch count and a HIP
 * stream.  Nothing on a launch path allocates or synchronises (only plan
 * creation and the explicit prepare / descriptor calls do): the caller owns
 * every input, output and workspace buffer; a plan owns only its constant
 * tables.  Plans are immutable after creation and may be shared by threads.
 * Every function returns an s2h status (0 = ok, negative = error; see
 * s2h_status_string); none of them calls exit().
 *
 * Reference interfaces replaced (paths relative to the s2fft repository):
 *   s2h_forward / s2h_inverse      s2fft/transforms/spherical.py:352-464 and
 *                                  21-109 (forward_numpy 467-588,
 *                                  inverse_numpy 112-208) for one spin
 *   s2h_lat_forward / s2h_lat_inverse
 *                                  s2fft/transforms/otf_recursions.py:432-588
 *                                  and 15-171 (latitudinal Wigner-d steps)
 *   s2h_healpix_fft / s2h_healpix_ifft
 *                                  s2fft/utils/healpix_ffts.py:150-226 and
 *                                  297-371 (healpix_fft / healpix_ifft)
 *   healpix_fft_hip                lib/src/extensions.cc:26-100 (the XLA
 *                                  custom call healpix_fft_cuda) with the
 *                                  descriptor of lib/include/s2fft.h:29-59
 *   s2h_build_healpix_fft_descriptor
 *                                  lib/src/extensions.cc:115-149
 *   s2h_compute_all_slices         s2fft/recursions/price_mcewen.py:377-643
 *   s2h_precompute_kernel / _forward / _inverse
 *                                  s2fft/precompute_transforms/construct.py:
 *                                  16-161, spherical.py:87-149,314-380
 *   s2h_wigner_*                   s2fft/transforms/wigner.py:13-107,346-439
 *   s2h_gamma_dft                  the FFT over gamma of
 *                                  s2fft/transforms/wigner.py:183-195,495-502
 *   s2h_wigner_reality_fill        s2fft/transforms/wigner.py:523-526
 */
#ifndef S2FFT_AMD_H
#define S2FFT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct s2h_plan s2h_plan;
typedef void *s2h_stream; /* a hipStream_t; NULL = the default stream */

enum s2h_sampling { S2H_MW = 0, S2H_MWSS = 1, S2H_DH = 2, S2H_GL = 3, S2H_HEALPIX = 4 };

enum s2h_status {
  S2H_OK = 0,
  S2H_ERR_INVALID = -1,     /* bad argument (the reference raises ValueError) */
  S2H_ERR_UNSUPPORTED = -2, /* valid but outside what this build supports */
  S2H_ERR_HIP = -3,         /* a HIP runtime call failed */
  S2H_ERR_NOMEM = -4,       /* device or host allocation failed at plan creation */
  S2H_ERR_SPIN = -5,        /* |spin| >= 8 (the reference raises Warning) */
  S2H_ERR_BANDLIMIT = -6,   /* L < 2 nside for HEALPix (reference AssertionError) */
  S2H_ERR_SCRATCH = -7      /* s2h_forward_noalloc: workspace beyond its fixed scratch */
};

/* Normalisation folded into a plan's Wigner-d recursion. */
enum s2h_norm {
  S2H_NORM_SPIN = 0,    /* sqrt((2l+1)/4pi) and (-1)^spin: spherical.py */
  S2H_NORM_WIGNER = 1,  /* forward 1, inverse (2l+1)/(8 pi^2 sqrt(pi))... wigner.py */
  S2H_NORM_NONE = 2     /* bare d^l_{m,-spin}: the latitudinal steps */
};

typedef struct s2h_desc {
  int32_t sampling;   /* enum s2h_sampling */
  int32_t L;          /* harmonic band-limit */
  int32_t nside;      /* HEALPix resolution (ignored otherwise) */
  int32_t spin;       /* |spin| < 8 */
  int32_t L_lower;    /* transform only L_lower <= l < L */
  int32_t reality;    /* real signal: m >= 0 computed, conjugate symmetry */
  int32_t norm;       /* enum s2h_norm */
  int32_t extra_sign; /* +1 or -1, multiplies every output (Wigner (-1)^n) */
} s2h_desc;

const char *s2h_status_string(int status);
int s2h_version(void);

int s2h_plan_create(s2h_plan **plan, const s2h_desc *desc);
int s2h_plan_destroy(s2h_plan *plan);

/* Bytes of device workspace s2h_forward / s2h_inverse need for `batch` maps. */
size_t s2h_workspace_bytes(const s2h_plan *plan, int batch);
/* 1 if s2h_inverse writes complex128 pixels, 0 if float64. */
int s2h_inverse_output_is_complex(const s2h_plan *plan);

/* f: batch maps, float64 (f_is_complex = 0) or complex128 (1), layout of
 * s2fft.sampling.s2_samples.f_shape.  flm: batch x (L, 2L-1) complex128.  */
int s2h_forward(const s2h_plan *plan, const void *f, int f_is_complex, void *flm,
                void *work, int batch, s2h_stream stream);
int s2h_inverse(const s2h_plan *plan, const void *flm, void *f, void *work, int batch,
                s2h_stream stream);

/* Arithmetic of the latitudinal projection / synthesis (no counterpart in
 * the reference, which is FP64 throughout).  S2H_PREC_FP64: FP64 everywhere.
 * In the other modes the Wigner-d recursion, its renormalisation, the ring
 * FFTs and every sum across waves stay FP64:
 * S2H_PREC_MIXED: the d-values and the ring (or flm) data are rounded to FP32
 *   and multiplied on the FP32 matrix cores with FP32 accumulation over one
 *   wave's 128 rings (forward) or one 32-l chunk (inverse); relative error
 *   ~1e-7 of max |flm| for band-limited maps at L = 2048.
 * S2H_PREC_BF16: each operand split into two BF16 (hi + lo, 16 significant
 *   bits), products hi*hi + hi*lo + lo*hi on the BF16 matrix cores with FP32
 *   accumulation (forward: over one wave's ring pairs; inverse: over all l);
 *   relative error ~1e-5.
 * S2H_PREC_BF16X1: BF16 operands and one product (~4e-3 relative): the
 *   correction transform of a refinement, whose contraction it changes by
 *   its own relative error (or, for the synthesis, a step length).
 * All are meant for iterative refinement whose residual is FP64. */
enum s2h_precision {
  S2H_PREC_FP64 = 0,
  S2H_PREC_MIXED = 1,
  S2H_PREC_BF16 = 2,
  S2H_PREC_BF16X1 = 3
};
/* flags of s2h_forward_ex: S2H_FWD_ACCUMULATE adds F(f) into flm (which must
 * hold coefficients with the plan's symmetry) instead of overwriting it.
 * flags of s2h_inverse_ex: S2H_INV_REAL_OUTPUT writes float64 pixels for a
 * reality plan even where s2h_inverse_output_is_complex says complex (its
 * imaginary part is zero there); S2H_INV_SUBTRACT writes sub_from - I(flm)
 * (sub_from in the output layout, may alias f): a residual in one pass. */
enum { S2H_FWD_ACCUMULATE = 1 };
enum { S2H_INV_REAL_OUTPUT = 1, S2H_INV_SUBTRACT = 2 };
int s2h_forward_ex(const s2h_plan *plan, const void *f, int f_is_complex, void *flm,
                   void *work, int batch, int precision, int flags, s2h_stream stream);
int s2h_inverse_ex(const s2h_plan *plan, const void *flm, void *f, void *work, int batch,
                   int precision, int flags, const void *sub_from, s2h_stream stream);

/* Adjoints (vector-Jacobian products) of s2h_forward and s2h_inverse for
 * the torch complex-gradient convention: grad_f = F^H g, grad_flm = I^H g.
 * Replace the custom VJPs of s2fft/transforms/spherical.py:291-325 and
 * 695-730 (whose bodies are the opposite latitudinal step) together with the
 * JAX autodiff of the FFT, weight and phase stages around them.
 * s2h_forward_adjoint: g is batch x (L, 2L-1) complex128; grad_f is float64
 * with reality, else complex128 (f layout); for MW/MWSS it includes the
 * adjoint of the theta resampling.
 * s2h_inverse_adjoint: g in the f layout (complex128 if g_is_complex, else
 * float64); grad_flm is batch x (L, 2L-1) complex128. */
int s2h_forward_adjoint(const s2h_plan *plan, const void *g, void *grad_f, void *work,
                        int batch, s2h_stream stream);
int s2h_inverse_adjoint(const s2h_plan *plan, const void *g, int g_is_complex, void *grad_flm,
                        void *work, int batch, s2h_stream stream);

/* Latitudinal steps on the plan's theta grid (plan created with
 * S2H_NORM_NONE).  ftm: batch x (ntheta, 2L or 2L-1) complex128 in the
 * reference ftm layout; flm: batch x (L, 2L-1) complex128.  */
int s2h_lat_forward(const s2h_plan *plan, const void *ftm, void *flm, int batch,
                    s2h_stream stream);
int s2h_lat_inverse(const s2h_plan *plan, const void *flm, void *ftm, int batch,
                    s2h_stream stream);

/* HEALPix ring FFTs with the reference layouts: f (12 nside^2,) complex,
 * ftm (4 nside - 1, 2L) complex; batch arrays back to back.
 * s2h_healpix_fft_prepare builds the ring tables of (current device, nside,
 * L) once -- it allocates and synchronises -- and every launch below only
 * looks them up: S2H_ERR_INVALID if they were never prepared.
 * s2h_healpix_fft / s2h_healpix_ifft: complex128, unnormalised in both
 * directions (healpix_ffts.py:150-226 / 297-371, norm "backward" forward and
 * "forward" inverse).  The _ex forms take the reference's fft_norm
 * (lib/include/s2fft.h:3-6: the forward divides each ring by its length n
 * under FORWARD, the inverse under BACKWARD, both by sqrt(n) under ORTHO,
 * neither under NONE) and double_precision = 0 for complex64 buffers (the
 * arithmetic stays FP64).  With reality the rings of length 2L use Re(f) and
 * zero m < 0, as the NumPy/JAX paths do. */
enum s2h_fft_norm {
  S2H_FFT_NORM_FORWARD = 1,
  S2H_FFT_NORM_BACKWARD = 2,
  S2H_FFT_NORM_ORTHO = 3,
  S2H_FFT_NORM_NONE = 4
};
int s2h_healpix_fft_prepare(int nside, int L);
int s2h_healpix_fft(int nside, int L, int reality, const void *f, void *ftm, int batch,
                    s2h_stream stream);
int s2h_healpix_ifft(int nside, int L, int reality, const void *ftm, void *f, int batch,
                     s2h_stream stream);
int s2h_healpix_fft_ex(int nside, int L, int reality, int norm, int double_precision,
                       const void *f, void *ftm, int batch, s2h_stream stream);
int s2h_healpix_ifft_ex(int nside, int L, int reality, int norm, int double_precision,
                        const void *ftm, void *f, int batch, s2h_stream stream);

/* XLA custom-call entry (replaces healpix_fft_cuda, lib/src/extensions.cc:90):
 * buffers[0] = input, buffers[1] = output, opaque = the bytes of an
 * s2h_healpix_fft_descriptor made by s2h_build_healpix_fft_descriptor (which
 * maps (forward, normalize) to a norm as build_healpix_fft_descriptor does,
 * extensions.cc:115-149, and prepares the tables, so the call itself neither
 * allocates nor synchronises).  The fields are those of the reference's
 * s2fftDescriptor (lib/include/s2fft.h:29-59) as int32 plus a batch count
 * (0 or 1: one array).  The input is never modified (the reference FFTs in
 * place).  shift must be 1 (the reference always shifts). */
typedef struct s2h_healpix_fft_descriptor {
  int32_t nside, L;          /* nside, harmonic_band_limit */
  int32_t reality, forward;
  int32_t norm;              /* enum s2h_fft_norm */
  int32_t shift;
  int32_t double_precision;  /* 1 complex128, 0 complex64 */
  int32_t batch;
} s2h_healpix_fft_descriptor;
int s2h_build_healpix_fft_descriptor(int nside, int L, int reality, int forward, int normalize,
                                     int double_precision, s2h_healpix_fft_descriptor *out);
void healpix_fft_hip(void *stream, void **buffers, const char *opaque, size_t opaque_len);
/* Status of the last healpix_fft_hip call on this thread (XLA's signature has
 * no return value). */
int healpix_fft_hip_last_status(void);

/* DFT over the first (gamma) axis of an SO(3) signal with G = 2N-1 gamma
 * samples, for each of the npix (beta, alpha) samples (N <= 8):
 *   mode 0: complex128 f[g]        -> fban[k] = scale sum_g f[g] e^{-2 pi i n g/G}
 *   mode 2: float64 f[g]           -> same, for a real signal
 *   mode 1: complex128 fban[k]     -> f[g] = scale sum_n fban[k] e^{+2 pi i n g/G}
 *   mode 3: complex128 fban[n >= 0] (N rows) -> float64 f[g], Hermitian in n
 *   mode 4: complex128 x[n >= 0] (N rows) -> float64 f[g] = scale Re sum_{n >= 0}
 *           x[n] e^{+2 pi i n g/G} (the adjoint of mode 2 kept to n >= 0)
 *   mode 5: float64 f[g] -> complex128 y[n >= 0] (N rows) = w_n scale
 *           sum_g f[g] e^{-2 pi i n g/G}, w_0 = 1, w_n = 2 (the adjoint of mode 3)
 * with n = k - (N-1).  in and out must not alias. */
int s2h_gamma_dft(int N, int64_t npix, int mode, double scale, const void *in, void *out,
                  s2h_stream stream);
/* Negative-n planes of a real signal's flmn (batch of one):
 * flmn[N-1-n, l, L-1+m] = (-1)^(m+n) conj(flmn[N-1+n, l, L-1-m]), n = 1..N-1
 * (s2fft/transforms/wigner.py:523-526). */
int s2h_wigner_reality_fill(void *flmn, int L, int N, s2h_stream stream);

/* Wigner (SO(3)) transforms with every n-plane in one launch per stage
 * (s2fft/transforms/wigner.py:13-107 inverse, 346-439 forward; the reference
 * maps a spin transform over n with shared precomputes, price_mcewen.py:
 * 309-374).  f: [gamma (2N-1)][beta, alpha] float64 (reality) or complex128,
 * or [gamma][pixel] for HEALPix; flmn: (2N-1, L, 2L-1) complex128.  The
 * gamma DFT, the ring FFTs, the theta resampling and the latitudinal steps
 * each run once over all planes (each plane with its own spin -n recursion
 * tables); with reality the n = 0 plane of the forward runs as its own real
 * transform (the reference's reality semantics) and the n < 0 planes are
 * filled by symmetry.  N <= 7 (N >= 8 returns S2H_ERR_SPIN).
 * s2h_wigner_forward_adjoint / s2h_wigner_inverse_adjoint: the VJPs in the
 * torch convention (grad_f = F^H g, grad_flmn = I^H g; float64 grad_f with
 * reality), for autograd; no counterpart beyond the reference's JAX autodiff
 * of wigner.py:536-645 through the spin custom VJPs. */
typedef struct s2h_wigner_plan s2h_wigner_plan;
typedef struct s2h_wigner_desc {
  int32_t sampling, L, N, nside, L_lower, reality;
} s2h_wigner_desc;
int s2h_wigner_plan_create(s2h_wigner_plan **plan, const s2h_wigner_desc *desc);
int s2h_wigner_plan_destroy(s2h_wigner_plan *plan);
size_t s2h_wigner_workspace_bytes(const s2h_wigner_plan *plan);
int s2h_wigner_forward(const s2h_wigner_plan *plan, const void *f, int f_is_complex, void *flmn,
                       void *work, s2h_stream stream);
int s2h_wigner_inverse(const s2h_wigner_plan *plan, const void *flmn, void *f, void *work,
                       s2h_stream stream);
int s2h_wigner_forward_adjoint(const s2h_wigner_plan *plan, const void *g, void *grad_f,
                               void *work, s2h_stream stream);
int s2h_wigner_inverse_adjoint(const s2h_wigner_plan *plan, const void *g, int g_is_complex,
                               void *grad_flmn, void *work, s2h_stream stream);

/* Stored Wigner-d tables.
 * s2h_compute_all_slices replaces compute_all_slices
 * (s2fft/recursions/price_mcewen.py:377-643): for a plan created with
 * S2H_NORM_NONE (reality 0; its sampling is irrelevant), beta: ntheta device
 * doubles, dl: (2L-1, ntheta, L) float64 with dl[L-1+m][t][l] =
 * d^l_{m,-spin}(beta_t), zero for l < max(|m|, |spin|).
 * The precompute path (s2fft/precompute_transforms/construct.py:16-161 and
 * spherical.py:87-149, 314-380): s2h_precompute_kernel writes the reference's
 * kernel K[t][l][m - m_start] = scale d^l_{m,-spin}(theta_t) x the plan's
 * normalisation (x the forward quadrature weight, x the HEALPix ring phase
 * e^{-+ i m phi0}) on the forward grid (MW/MWSS: the MWSS grid at 2L, poles
 * included) or the inverse grid, float64 or complex128 (HEALPix),
 * s2h_precompute_kernel_bytes(plan, forward) bytes; m from 0 with reality,
 * else from -(L-1).  s2h_precompute_forward: the ring FFTs (+ resampling)
 * without weights or phases, then flm = scale sum_t K ftm, conjugate fill
 * with reality (the reference's scale is (-1)^spin).  s2h_precompute_inverse:
 * ftm = scale sum_l K flm, then the ring iFFTs without phases.  A kernel the
 * reference built (same layout) may be passed.  work: s2h_workspace_bytes. */
int s2h_compute_all_slices(const s2h_plan *plan, const double *beta, int ntheta, double *dl,
                           s2h_stream stream);
size_t s2h_precompute_kernel_bytes(const s2h_plan *plan, int forward);
int s2h_precompute_kernel(const s2h_plan *plan, int forward, double scale, void *kernel,
                          s2h_stream stream);
int s2h_precompute_forward(const s2h_plan *plan, const void *kernel, int kernel_is_complex,
                           double scale, const void *f, int f_is_complex, void *flm, void *work,
                           int batch, s2h_stream stream);
int s2h_precompute_inverse(const s2h_plan *plan, const void *kernel, int kernel_is_complex,
                           double scale, const void *flm, void *f, void *work, int batch,
                           s2h_stream stream);

/* The transforms split at the ring spectra (no counterpart in the reference;
 * for refinement in spectral space).  For plans whose forward and inverse
 * share one ring grid with neither pole rows nor resampling (HEALPix, GL;
 * others return S2H_ERR_UNSUPPORTED):
 *   s2h_forward = s2h_forward_rings . s2h_ring_analysis,
 *   F(f - I(x)) = s2h_forward_rings(s2h_ring_residual(G, s2h_inverse_rings(x)))
 * with G = s2h_ring_analysis(f).  Spectra: batch x (rows, Mc) complex128,
 * s2h_ring_spectra_bytes(plan, batch) bytes.  s2h_ring_residual writes
 * R = G - A(Y), A = ring FFT . ring iFFT evaluated without the FFTs (a fold of
 * each ring's m-spectrum onto its n_t bins and the periodic extension back,
 * with weights and ring phases); R may be G, not Y.  s2h_forward_rings takes
 * S2H_FWD_ACCUMULATE; precision as in s2h_forward_ex. */
size_t s2h_ring_spectra_bytes(const s2h_plan *plan, int batch);
int s2h_ring_analysis(const s2h_plan *plan, const void *f, int f_is_complex, void *G, int batch,
                      s2h_stream stream);
int s2h_forward_rings(const s2h_plan *plan, const void *G, void *flm, int batch, int precision,
                      int flags, s2h_stream stream);
int s2h_inverse_rings(const s2h_plan *plan, const void *flm, void *Y, int batch, int precision,
                      int flags, s2h_stream stream);
/* S2H_FLM_PACKED (flags of s2h_forward_rings / s2h_inverse_rings, real maps
 * only): flm in the packed m-major layout, batch x L(L+1)/2 complex128 with
 * (l, m), l >= m >= 0, at m L - m(m-1)/2 + l - m -- the l of one m contiguous,
 * so the latitudinal kernels read and write whole cache lines.  Converters
 * to and from the reference layout (unpack writes the zeros and the m < 0
 * half, (-1)^m conj). */
enum { S2H_FLM_PACKED = 4 };
/* S2H_SPECTRA_TILED (flags of s2h_forward_rings, s2h_inverse_rings,
 * s2h_ring_residual_ex, s2h_ring_dot_ex): the spectra in blocks of 8 rows,
 * element (row t, column c) of a map at ((t / 8) Mc + c) 8 + t % 8 complex128
 * entries, so that the 8 rings of a block share a 128-byte line per column
 * (the latitudinal kernels gather and scatter one column over every ring).
 * Same size as the ring-major spectra: s2h_ring_spectra_bytes rounds the
 * rows up to a multiple of 8 for both.  s2h_ring_residual_ex and
 * s2h_ring_dot_ex take it for real maps only.  S2H_SPECTRA_G_RING_MAJOR
 * (s2h_ring_residual_ex, with S2H_SPECTRA_TILED): G ring-major (as
 * s2h_ring_analysis writes it), Y and R tiled; R may not be G. */
enum { S2H_SPECTRA_TILED = 8, S2H_SPECTRA_G_RING_MAJOR = 16 };
int s2h_flm_pack(int L, const void *flm, void *flm_packed, int batch, s2h_stream stream);
int s2h_flm_unpack(int L, const void *flm_packed, void *flm, int batch, s2h_stream stream);
int s2h_ring_residual(const s2h_plan *plan, const void *G, const void *Y, void *R, int batch,
                      s2h_stream stream);

/* Conjugate-gradient refinement on the ring spectra (no counterpart in the
 * reference, whose refinement is Jacobi, s2fft/utils/iterative_refinement.py:9-46).
 * F I is Hermitian positive definite on the coefficients (F = w I^H with the
 * HEALPix / GL quadrature), so the forward is refined by CG on F I x = F f.
 *   s2h_ring_residual_scaled: R = G - alpha[b] A(Y) (alpha: batch doubles on
 *     the device; null = 1; R may be G, not Y); with sumsq (batch doubles)
 *     also sumsq[b] = sum |R|^2 over the map's spectra, fixed order
 *     (work: s2h_ring_dot_workspace_bytes; sumsq and work may be null).
 *   s2h_ring_dot: pq[b] = <p, F A(Y)> for Y = s2h_inverse_rings(p), i.e.
 *     sum over rings and m of c_m Re(conj(Y) A(Y)), c_m = 2 for m > 0 of real
 *     maps, else 1 (the CG curvature, without projecting A(Y)); with alpha,
 *     alpha[b] = num[b] / pq[b].  pq or alpha may be null (not both).
 *     work: s2h_ring_dot_workspace_bytes(plan, batch).
 *   s2h_flm_dots: out[i * batch + b] = <z_b, v_i,b> = sum_j w_j Re(conj(z) v_i)
 *     over n coefficients per map (maps contiguous), 1 <= nv <= 4 (v: host
 *     array of device pointers); w_j = 1 for j < packed_L and 2 beyond (the
 *     packed layout of real maps: the m = 0 block first), all 1 when
 *     packed_L = 0.  Fixed summation order.  work: s2h_flm_dots_workspace_bytes.
 *   s2h_cg_update: x += alpha[b] p (x may be null); then, if p_new is
 *     given, p_new = z - g1[b] p - g2[b] p_prev (p_prev and g2 optional;
 *     p_new may be z, not p or x).
 *   s2h_cg_finish: x + sum_{j < nd} alpha_j[b] p_j (nd <= 4, the terms in
 *     the order j = 0, 1, ...), into x, or with flm_out (packed real maps,
 *     n = L(L+1)/2) unpacked to the reference layout batch x (L, 2L-1) as
 *     s2h_flm_unpack does: the deferred coefficient updates and the unpack
 *     in one pass. */
int s2h_ring_residual_scaled(const s2h_plan *plan, const void *G, const void *Y,
                             const double *alpha, void *R, double *sumsq, void *work, int batch,
                             s2h_stream stream);
size_t s2h_ring_dot_workspace_bytes(const s2h_plan *plan, int batch);
int s2h_ring_dot(const s2h_plan *plan, const void *Y, const double *num, double *pq,
                 double *alpha, void *work, int batch, s2h_stream stream);
/* The same with layout flags (S2H_SPECTRA_TILED, S2H_SPECTRA_G_RING_MAJOR);
 * flags 0 is s2h_ring_residual_scaled / s2h_ring_dot. */
int s2h_ring_residual_ex(const s2h_plan *plan, const void *G, const void *Y, const double *alpha,
                         void *R, double *sumsq, void *work, int batch, int flags,
                         s2h_stream stream);
int s2h_ring_dot_ex(const s2h_plan *plan, const void *Y, const double *num, double *pq,
                    double *alpha, void *work, int batch, int flags, s2h_stream stream);
/* s2h_inverse_rings_dot: Y = s2h_inverse_rings(flm) and the dot of
 * s2h_ring_dot_ex on it in one call, flags S2H_SPECTRA_TILED (required) and
 * S2H_FLM_PACKED, real maps.  The matrix-core syntheses add up the dot over
 * the rows without aliasing as they store them, so only the aliased rows are
 * read again.  work: s2h_inverse_rings_dot_workspace_bytes(plan, batch). */
size_t s2h_inverse_rings_dot_workspace_bytes(const s2h_plan *plan, int batch);
int s2h_inverse_rings_dot(const s2h_plan *plan, const void *flm, void *Y, int batch, int precision,
                          int flags, const double *num, double *pq, double *alpha, void *work,
                          s2h_stream stream);
size_t s2h_flm_dots_workspace_bytes(int nv, int batch);
int s2h_flm_dots(const void *z, const void *const *v, int nv, int64_t n, int packed_L, int batch,
                 double *out, void *work, s2h_stream stream);
int s2h_cg_update(void *x, const void *p, const double *alpha, void *p_new, const void *z,
                  const double *g1, const void *p_prev, const double *g2, int64_t n, int batch,
                  s2h_stream stream);
int s2h_cg_finish(void *x, const void *const *p, const double *const *alpha, int nd, int64_t n,
                  int L, void *flm_out, int batch, s2h_stream stream);

/* Per-map sums of squares out[b] = sum_{i < n} x[b * stride + i]^2 (doubles;
 * a complex map is 2n doubles), with a fixed summation order.  No
 * counterpart in the reference: the residual norms that judge the learned
 * preconditioner's steps.  work: s2h_batch_sumsq_workspace_bytes(batch). */
size_t s2h_batch_sumsq_workspace_bytes(int batch);
int s2h_batch_sumsq(const double *x, int64_t n, int64_t stride, int batch, double *out,
                    void *work, s2h_stream stream);

/* Kernel timing for benchmarks (no counterpart in the reference).  While
 * enabled, s2h_forward / s2h_inverse record a HIP event pair on the launch
 * stream around each kernel; s2h_profile_query synchronises on them and
 * returns the summed milliseconds and the launch count of one kernel kind.
 * s2h_profile_enable(0/1) clears the records.  Off by default.  The
 * latitudinal kinds count FP64 launches; the _FP32 / _BF16 kinds count the
 * reduced-precision ones (S2H_PREC_MIXED, and S2H_PREC_BF16 / _BF16X1). */
enum s2h_kernel_kind {
  S2H_KERNEL_LAT_FORWARD = 0,
  S2H_KERNEL_LAT_INVERSE = 1,
  S2H_KERNEL_RING_FFT_FORWARD = 2,
  S2H_KERNEL_RING_FFT_INVERSE = 3,
  S2H_KERNEL_RESAMPLE = 4,
  S2H_KERNEL_LAT_FORWARD_FP32 = 5,
  S2H_KERNEL_LAT_INVERSE_FP32 = 6,
  S2H_KERNEL_LAT_FORWARD_BF16 = 7,
  S2H_KERNEL_LAT_INVERSE_BF16 = 8,
  S2H_KERNEL_RING_RESIDUAL = 9
};
int s2h_profile_enable(int on);
int s2h_profile_query(int kernel, double *total_ms, int *count);

/* Learned preconditioner for the iterative forward transform (no
 * counterpart in the reference, whose refinement is plain Jacobi,
 * s2fft/utils/iterative_refinement.py:9-46).  A per-(l, m) MLP
 *   (re, im, l/L, m/L) -> correction (re, im)
 * with two hidden layers of H = 64 (ReLU), FP16-rounded weights, features and
 * activations and FP32 accumulation; (re, im) is the harmonic residual
 * divided by its RMS over the map's coefficients (summed in a fixed order:
 * the same output on every run), and the correction is scaled back by it.
 * params: s2h_mlp_param_count() = H^2 + 8 H + 2 FP32 values, layout
 *   W1[H][4] b1[H] W2[H][H] b2[H] W3[2][H] b3[2]. */
int s2h_mlp_param_count(void);
/* delta = correction for every packed coefficient (m in [0, l] with reality,
 * whose m < 0 entries are filled by conjugate symmetry; else m in [-l, l]),
 * zero for l < lmin.  h, delta: batch x (L, 2L-1) complex128, distinct;
 * reality = 2: real maps with h and delta in the S2H_FLM_PACKED layout
 * (batch x L(L+1)/2).  work: s2h_mlp_workspace_bytes(batch) bytes. */
size_t s2h_mlp_workspace_bytes(int batch);
int s2h_mlp_precondition(const float *params, const void *h, void *delta, int L, int reality,
                         int lmin, int batch, void *work, s2h_stream stream);
/* The same, and dot[b] (device) = <delta, h> per map: sum over the stored
 * rows of c Re(conj(delta) h), c = 2 for the m > 0 rows of real maps (their
 * m < 0 rows) and 1 otherwise; the value s2h_flm_dots gives, summed in
 * another fixed order.  For the conjugate-gradient step that starts from a
 * preconditioned residual.  work: s2h_mlp_workspace_bytes(batch). */
int s2h_mlp_precondition_dot(const float *params, const void *h, void *delta, int L, int reality,
                             int lmin, int batch, double *dot, void *work, s2h_stream stream);
/* Training rows from a residual h and the error e it should predict:
 * X[b*P + p] = (h/s, l/L, m/L) (4 floats), Y = e/s (2 floats), with P the
 * packed rows per map; rows with l < lmin are zero.
 * work: s2h_mlp_workspace_bytes(batch) bytes. */
int s2h_mlp_features(const void *h, const void *e, int L, int reality, int lmin, int batch,
                     float *X, float *Y, void *work, s2h_stream stream);
/* Gradient of sum_rows w_r |y(X_r) - Y_r|^2 * loss_scale over n rows
 * (row_weight may be NULL = 1) into grad (param layout); *loss (device
 * double, may be NULL) = sum_rows w_r |y - Y|^2.  nwg workgroups; work:
 * s2h_mlp_grad_workspace_bytes(nwg) bytes.  Deterministic for fixed nwg. */
size_t s2h_mlp_grad_workspace_bytes(int nwg);
int s2h_mlp_grad(const float *params, const float *X, const float *Y, const float *row_weight,
                 int64_t n, float loss_scale, float *grad, double *loss, void *work, int nwg,
                 s2h_stream stream);
/* One Adam step (step >= 1) on the FP32 master parameters. */
int s2h_mlp_adam(float *params, const float *grad, float *m1, float *m2, int step, float lr,
                 float beta1, float beta2, float eps, s2h_stream stream);

/* s2h_forward_noalloc: s2h_forward (FP64) with its workspace in a fixed
 * device array of S2H_NOALLOC_SCRATCH_BYTES per device, held from the first
 * call until the process ends, so that the call allocates nothing (no
 * hipMalloc, no hipFree, no allocating library routine).  No counterpart in
 * the reference: it is for callers that cannot allocate (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  Returns
 * S2H_ERR_SCRATCH when s2h_workspace_bytes(plan, batch) exceeds the scratch
 * (real HEALPix transforms: 4 maps at nside 256, 1 at nside 512).  Every call on a
 * device shares its scratch: calls that overlap on one device race.
 * s2h_noalloc_scratch_bytes returns the scratch size. */
#define S2H_NOALLOC_SCRATCH_BYTES (64u << 20)
size_t s2h_noalloc_scratch_bytes(void);
int s2h_forward_noalloc(const s2h_plan *plan, const void *f, int f_is_complex, void *flm,
                        int batch, s2h_stream stream);

/* Diagnostics (no counterpart in the reference): one V_MFMA_F64_16X16X4F64
 * with a[lane], b[lane] (64 doubles each) as the A and B operands and a zero
 * accumulator; out[4*lane + r] = result register r of lane (256 doubles).
 * Lets a test confirm the operand / result lane layouts the kernels assume. */
int s2h_probe_mfma_f64(const double *a, const double *b, double *out, s2h_stream stream);
/* Sustained FP64 matrix-core rate (no counterpart in the reference): blocks x
 * 256 threads, each wave issuing 8 independent V_MFMA_F64_16X16X4F64 (2048
 * flops each) per iteration, iters iterations; out: blocks * 256 doubles. */
int s2h_probe_mfma_f64_rate(int blocks, int iters, double *out, s2h_stream stream);
/* The same for V_MFMA_F32_16X16X4F32 (64 floats in each of a, b; 256 out). */
int s2h_probe_mfma_f32(const float *a, const float *b, float *out, s2h_stream stream);
/* The same for V_MFMA_F32_16X16X32_BF16: a, b hold 8 bf16 bit patterns per
 * lane (lane-major, 512 each); out 256 floats. */
int s2h_probe_mfma_bf16(const uint16_t *a, const uint16_t *b, float *out, s2h_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* S2FFT_AMD_H */
