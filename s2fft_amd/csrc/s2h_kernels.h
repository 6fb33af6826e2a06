// Kernel argument blocks and launcher declarations.
#pragma once

#include "s2h_internal.h"

namespace s2h {

struct RingFwdArgs {
  const RingJob* jobs;
  const void* f;            // float64 or complex128 maps
  int f_complex;
  int real_only_2L;         // reference reality: rings with nphi == 2L use Re(f)
  int real_only_all;        // every ring uses Re(f)
  int64_t map_stride;       // pixels per map
  cplx* out;
  int64_t out_map_stride;   // complex elements per map of the output
  int ncol;                 // output columns per row
  int m_first;              // m of column 0
  int apply_weight;
  int apply_phase;
  int zero_neg_2L;          // reference reality: m < 0 zero on rings with nphi == 2L
  int L;
  const cplx* chirp_pool;
  int batch;                // set by the launcher
  int in_single;            // complex64 input (complex, unpaired maps only)
  int out_single;           // complex64 output
  int norm_mode;            // extra per-ring scale: 0 none, 1: 1/nphi, 2: 1/sqrt(nphi)
};

struct RingInvArgs {
  const RingJob* jobs;
  const cplx* in;
  int64_t in_map_stride;
  int ncol;
  int m_first;              // m of input column 0
  int m_lo, m_hi;           // folded m range (non-Hermitian rings)
  int herm_mode;            // 0 none, 1 every ring, 2 rings with nphi == 2L
  int apply_phase;
  int apply_weight;
  void* f;
  int f_complex;
  int64_t map_stride;
  const cplx* chirp_pool;
  int L;
  int batch;                // set by the launcher
  const void* sub_from;     // non-null: f = sub_from - (transform), same layout as f
  int in_single;            // complex64 input spectra (unpaired maps only)
  int f_single;             // complex64 output pixels (complex output only)
  int norm_mode;            // extra per-ring scale: 0 none, 1: 1/nphi, 2: 1/sqrt(nphi)
};

// R = G - A(Y) on ring spectra, A = (ring FFT) . (ring iFFT): per ring, the
// fold of Y's m-spectrum onto n_t bins with e^{+i m phi0} (as the inverse
// FFT reads it), then the periodic extension times n_t w_t e^{-i m phi0} (as
// the forward FFT writes it).  Same layouts as RingInvArgs.in / RingFwdArgs.out.
// Ring-spectra addressing within one map.  Ring-major: element (row t,
// column c) at t ncol + c.  Tiled: rows in blocks of 8, at
// ((t / 8) ncol + c) 8 + t % 8, so that for every column the 8 rings of a
// block share one 128-byte line (the latitudinal kernels gather and scatter
// one column over all rings).
constexpr int kSpecBlock = 8;
__host__ __device__ __forceinline__ int64_t spec_row(int t, int ncol, int tiled) {
  return tiled ? (((int64_t)(t >> 3) * ncol) << 3) + (t & 7) : (int64_t)t * ncol;
}
__host__ __device__ __forceinline__ int spec_cstride(int tiled) { return tiled ? kSpecBlock : 1; }

// R = G - alpha A(Y) / the CG dot on tiled spectra of real maps (Hermitian,
// m in [0, L)): one workgroup per block of 8 rows and map.
struct RingAliasTiledArgs {
  const RingJob* rowjobs;   // ring of each spectra row (nphi 0: padding row)
  int nblocks;              // rows / 8
  const int* belt_blocks;   // blocks whose rows all have n >= 2L - 1 (no alias)
  int n_belt;
  const int* polar_blocks;  // the other blocks
  int n_polar;
  const cplx* y;            // tiled
  const cplx* g;            // tiled, or ring-major (g_tiled 0); may equal r when tiled
  cplx* r;                  // tiled
  int g_tiled;
  int64_t map_stride;       // rows * ncol, all three arrays
  int ncol;                 // L
  int apply_phase;
  int L;
  const double* alpha;      // per-map scale (null: 1)
  double* part;             // partial sums (the dot in dot mode, else |R|^2): map b's
                            // at part + b part_map_stride, kAliasSlots per row block,
                            // the polar blocks' from polar_part_off, the belt
                            // blocks' from belt_part_off (in list order)
  int64_t part_map_stride;
  int polar_part_off;
  int belt_part_off;
};

struct RingAliasArgs {
  const RingJob* jobs;
  const cplx* y;            // spectra to alias (m in [-(L-1), L-1], Hermitian if herm)
  const cplx* g;            // subtrahend G (may equal r)
  cplx* r;
  int64_t map_stride;       // rows * ncol
  int ncol;                 // columns of every spectrum, m = m_first + col
  int m_first;
  int herm;                 // y stores m >= 0 of Hermitian spectra
  int apply_phase;
  int L;
  const double* alpha;      // per-map scale of A(Y) in R = G - alpha A(Y) (null: 1)
  double* part;             // [map][ring job] partial sums: the dot (dot mode) or |R|^2
};

// Recursion tables of one map of a multi-plane launch (the n-planes of a
// Wigner transform: one spin -n per map, every other table shared).
struct LatPlane {
  const MStart* mstart;     // indexed by m - m_lo
  const Coef* coef;         // [|m|][l]
  int lmin_out;             // first l written / read
  int pad;
};

struct LatArgs {
  const LatRing* rings;
  int nrings;               // entries of rings (pairs count as one)
  const MStart* mstart;     // indexed by m - m_lo
  const Coef* coef;         // [|m|][l]
  const Coef* ucoef;        // [|m|][l], unit-c form {a', b', sigma, 0} (BF16 kernels)
  int L;
  int m_lo;                 // first m handled (0 or -(L-1))
  int nm;                   // number of m values handled
  int lmin_out;             // first l written / read (L_lower)
  int batch;
  // harmonic side: flm[b][l][L-1+m]
  cplx* flm;
  int64_t flm_map_stride;
  // ring side: ftm[b] + spec_row(row) + (m - ftm_m_first) spec_cstride
  cplx* ftm;
  int64_t ftm_map_stride;
  int ftm_ncol;
  int ftm_m_first;
  int ftm_tiled;            // ring spectra in blocks of 8 rings (spec_row)
  int groups;               // ring groups per m (forward); >1 accumulates atomically
  int accumulate;           // forward: add into flm instead of overwriting it
  int fill;                 // forward, one ring group per m: also write zeros below
                            // max(ell0, lmin_out), so flm needs no zeroing pass
  int mirror;               // forward with fill, real maps: also write
                            // flm[l, -m] = (-1)^m conj(flm[l, m]) (no conjugate fill pass)
  int flm_packed;           // flm in the packed m-major layout (real maps, m >= 0 only)
  const LatPlane* planes;   // non-null: map b uses planes[b] in place of mstart, coef and
                            // lmin_out (VALU kernels, unpaired rings)
  // matrix-core syntheses of real maps: with dot_part, the part of the CG dot
  // <Y, A(Y)> over the rows [dot_row_lo, dot_row_hi), which have no aliasing
  // and the one w n = dot_wn, into dot_part[map * dot_part_stride + blockIdx.x]
  double dot_wn;
  int dot_row_lo, dot_row_hi;
  double* dot_part;
  int64_t dot_part_stride;
};

// Harmonic-side addressing.  Reference layout: flm[b][l][L-1+m].  Packed
// m-major layout (real maps): the L(L+1)/2 coefficients l >= m >= 0 of a map
// stored as m-blocks, flm_p[b][moff(m) + l - m], moff(m) = m L - m(m-1)/2, so
// the l of one m are contiguous.  Element (l, m) sits at flm_col(m) + l * flm_row.
__host__ __device__ __forceinline__ int64_t packed_moff(int L, int m) {
  return (int64_t)m * L - (int64_t)m * (m - 1) / 2;
}
__host__ __device__ __forceinline__ int64_t flm_col(const LatArgs& a, int m) {
  return a.flm_packed ? packed_moff(a.L, m) - m : (int64_t)(a.L - 1 + m);
}
__host__ __device__ __forceinline__ int64_t flm_row(const LatArgs& a) {
  return a.flm_packed ? 1 : 2 * a.L - 1;
}
// (l, m) of packed row p
__device__ __forceinline__ void packed_lm(int64_t p, int L, int& l, int& m) {
  const double b = 2.0 * L + 1.0;
  int mm = (int)((b - sqrt(b * b - 8.0 * (double)p)) * 0.5);
  if (mm < 0) mm = 0;
  while (mm > 0 && packed_moff(L, mm) > p) --mm;
  while (mm + 1 < L && packed_moff(L, mm + 1) <= p) ++mm;
  m = mm;
  l = (int)(p - packed_moff(L, mm)) + mm;
}

hipError_t launch_ring_fft_forward(const RingFwdArgs& a, const JobSet& js, int batch,
                                   hipStream_t s);
hipError_t launch_ring_fft_inverse(const RingInvArgs& a, const JobSet& js, int batch,
                                   hipStream_t s);
hipError_t allow_large_lds_rings();
hipError_t launch_ring_alias(const RingAliasArgs& a, const JobSet& js, int batch, hipStream_t s);
// pq[b] = sum_m c_m Re(conj(Y) A(Y)) over every ring of map b (part: batch x
// ring_dot_jobs(js) doubles); alpha[b] = num[b] / pq[b] when alpha is given
int ring_dot_jobs(const JobSet& js);
// launch_ring_alias that also writes sumsq[b] = sum |R|^2 of map b (part as
// for launch_ring_dot)
hipError_t launch_ring_alias_tiled(const RingAliasTiledArgs& a, int batch, hipStream_t s);
// partial sums per map of the tiled alias kernel (its workspace)
int ring_dot_tiled_parts(int nblocks);
hipError_t launch_ring_alias_tiled_sumsq(const RingAliasTiledArgs& a, int batch, double* part,
                                         double* sumsq, hipStream_t s);
hipError_t launch_ring_dot_tiled(const RingAliasTiledArgs& a, int batch, double* part,
                                 const double* num, double* pq, double* alpha, hipStream_t s);
// The dot's polar row blocks only, after n_head partial sums per map already
// in part (the synthesis-fused part over the other rows): pq and alpha over
// both, part laid out per map as [n_head][polar blocks x slots].
hipError_t launch_ring_dot_tiled_polar(const RingAliasTiledArgs& a, int batch, double* part,
                                       int n_head, const double* num, double* pq, double* alpha,
                                       hipStream_t s);
hipError_t launch_ring_alias_sumsq(const RingAliasArgs& a, const JobSet& js, int batch,
                                   double* part, double* sumsq, hipStream_t s);
hipError_t launch_ring_dot(const RingAliasArgs& a, const JobSet& js, int batch, double* part,
                           const double* num, double* pq, double* alpha, hipStream_t s);
hipError_t launch_chirp_setup(const int* lengths, const int* conv_lens, const int64_t* offs,
                              cplx* pool, int count, int max_conv, hipStream_t s);
hipError_t launch_gamma_dft(int N, int64_t npix, int mode, double scale, const void* in,
                            void* out, hipStream_t s);
hipError_t launch_wigner_fill(cplx* flmn, int L, int N, hipStream_t s);
hipError_t launch_wigner_fill_adjoint(cplx* g, int L, int N, hipStream_t s);
hipError_t launch_real_to_complex(const double* in, cplx* out, int64_t n, hipStream_t s);
hipError_t allow_large_lds();

hipError_t launch_lat_inverse(const LatArgs& a, bool paired, hipStream_t s);
hipError_t launch_lat_forward(const LatArgs& a, bool paired, hipStream_t s);
hipError_t launch_lat_forward_mfma(const LatArgs& a, bool paired, bool fp32, hipStream_t s);
hipError_t launch_mfma_f64_probe(const double* a, const double* b, double* out, hipStream_t s);
// out: blocks * 256 doubles; 8 * iters * blocks * 4 V_MFMA_F64_16X16X4F64
hipError_t launch_mfma_f64_rate(int blocks, int iters, double* out, hipStream_t s);
hipError_t launch_lat_inverse_mma(const LatArgs& a, bool paired, bool fp32, hipStream_t s);
hipError_t launch_lat_forward_bf16(const LatArgs& a, bool paired, int split, hipStream_t s);
// ring groups per m of launch_lat_forward_bf16 (fill and mirror need 1)
int lat_forward_bf16_groups(const LatArgs& a, int split);
hipError_t launch_flm_to_mmajor(const cplx* flm, cplx* flmT, int L, int m_lo, int nm, int batch,
                                hipStream_t s);
hipError_t launch_mmajor_to_flm(const cplx* flmT, cplx* flm, int L, int m_lo, int nm, int batch,
                                hipStream_t s);
// split: 3 = hi/lo operands (three products), 1 = single BF16 operands
hipError_t launch_lat_inverse_bf16(const LatArgs& a, bool paired, int split, hipStream_t s);
hipError_t launch_mfma_f32_probe(const float* a, const float* b, float* out, hipStream_t s);

// learned preconditioner (per-(l, m) rows of one or more flm arrays)
struct MlpRowArgs {
  const cplx* h;       // batch x (L, 2L-1), or batch x L(L+1)/2 packed m-major
  cplx* out;           // same layout as h (apply)
  int L, reality, lmin, batch;
  int packed;          // h / out in the packed m-major layout (reality)
  int64_t P;           // rows per map: L(L+1)/2 (reality) or L^2
  int64_t count;       // rows with l >= lmin (for the RMS)
  int64_t map_stride;  // L (2L-1), or L(L+1)/2 packed
};
hipError_t launch_flm_pack(int L, const cplx* src, cplx* dst, int batch, hipStream_t s);
hipError_t launch_flm_unpack(int L, const cplx* src, cplx* dst, int batch, hipStream_t s);
// partial sums of the RMS pass per map (work of s2h_mlp_precondition / _features)
constexpr int kMlpSumsqParts = 1024;
hipError_t launch_mlp_sumsq(const MlpRowArgs& a, double* part, hipStream_t s);
// partial sums of <delta, h> per map (s2h_mlp_precondition_dot)
constexpr int kMlpDotParts = 256;
// dot (optional, with dot_part): dot[b] = sum over the rows of c_m Re(conj(delta) h)
hipError_t launch_mlp_apply(const MlpRowArgs& a, const float* w, const double* sumsq,
                            double* dot_part, double* dot, hipStream_t s);
hipError_t launch_mlp_features(const MlpRowArgs& a, const cplx* e, const double* sumsq, float* X,
                               float* Y, hipStream_t s);
hipError_t launch_mlp_grad(const float* w, const float* X, const float* Y, const float* rw,
                           int64_t n, float loss_scale, int nwg, float* partial,
                           double* loss_partial, float* grad, double* loss, hipStream_t s);
hipError_t launch_mlp_adam(float* w, const float* grad, float* m1, float* m2, int step, float lr,
                           float beta1, float beta2, float eps, hipStream_t s);
hipError_t launch_mfma_bf16_probe(const uint16_t* a, const uint16_t* b, float* out,
                                  hipStream_t s);

// Stored d-tables (s2h_precompute.hip): out[mi * sm + t * st + l * sl] =
// scale wt[t] lw[l] d^l_{m,n}(theta_t) (times e^{i phase_sign m phi0[t]} as
// complex128 when cplx), m = m_lo + mi, from a plan's recursion tables.
struct DTableArgs {
  const MStart* mstart;     // indexed by mi
  const Coef* coef;         // [|m|][l]
  const double* theta;      // T colatitudes
  int T, L, m_lo, nm;
  const double* wt;         // per-theta factor, or null
  const double* lw;         // per-l factor, or null
  const double* phi0;       // per-theta azimuth (cplx), or null
  double phase_sign;
  double scale;
  int cplx;
  void* out;
  int64_t sm, st, sl;
};
hipError_t launch_dtable(const DTableArgs& a, hipStream_t s);

// Contractions with a stored kernel K[t][l][mi] (float64 or complex128):
// forward flm[b][l][L-1+m] = scale sum_t K Y[b][t][mi]; inverse
// Y[b][t][mi] = scale sum_l K flm[b][l][L-1+m]; m = m_lo + mi.
struct PcArgs {
  const void* kernel;
  int kernel_complex;
  int T, L, m_lo, nm;
  cplx* y;                  // [b][t][mi], y_map_stride per map
  int64_t y_map_stride;
  cplx* flm;                // [b][l][L-1+m], flm_map_stride per map
  int64_t flm_map_stride;
  double scale;
  int batch;
};
hipError_t launch_pc_forward(const PcArgs& a, hipStream_t s);
hipError_t launch_pc_inverse(const PcArgs& a, hipStream_t s);

// Small fix-up kernels (s2h_misc.hip).
struct PoleArgs {
  // inverse: set ftm rows of the poles; forward: add pole terms into flm
  cplx* flm;
  int64_t flm_map_stride;
  cplx* ftm;
  int64_t ftm_map_stride;
  int ftm_ncol;
  int ftm_m_first;
  int L;
  int spin;
  int lmin;                 // L_lower (inverse) or max(|spin|, L_lower) (forward)
  int row_south;            // -1 if absent
  int row_north;            // -1 if absent
  const double* norm;       // per-l normalisation (times the global sign)
  int batch;
  int spin_step;            // map b has spin + spin_step b (multi-plane launches)
  int lmin_spin;            // lmin of map b is max(lmin, |its spin|)
};
__host__ __device__ __forceinline__ int pole_spin(const PoleArgs& a, int64_t b) {
  return a.spin + a.spin_step * (int)b;
}
hipError_t launch_pole_inverse(const PoleArgs& a, hipStream_t s);
hipError_t launch_pole_forward(const PoleArgs& a, hipStream_t s);
hipError_t launch_zero(void* p, int64_t ndoubles, hipStream_t s);
hipError_t launch_spec_pad_zero(cplx* spec, int64_t map_stride, int ncol, int row0, int rows,
                                int tiled, int batch, hipStream_t s);
constexpr int kFlmDotMax = 4;      // vectors per launch_flm_dots
constexpr int kFlmDotBlocks = 256; // slices per map of launch_flm_dots
struct FlmDotArgs {
  const cplx* z;
  const cplx* v[kFlmDotMax];
  int nv;
  int64_t n;                // coefficients per map
  int64_t stride;           // between maps
  int packed_L;             // > 0: packed m-major layout of real maps (weights 1 / 2)
  double* partial;          // nv * batch * kFlmDotBlocks
  double* out;              // [i][b]
};
hipError_t launch_flm_dots(const FlmDotArgs& a, int batch, hipStream_t s);
struct CgUpdateArgs {
  cplx* x;                  // optional
  const cplx* p;
  const double* alpha;
  cplx* p_new;              // null: x only
  const cplx* z;
  const double* g1;
  const cplx* p_prev;       // optional second direction
  const double* g2;
  int64_t n;
  int64_t stride;
};
hipError_t launch_cg_update(const CgUpdateArgs& a, int batch, hipStream_t s);
constexpr int kCgFinishMax = 4;    // directions per launch_cg_finish
struct CgFinishArgs {
  cplx* x;                  // n per map; updated in place when dst is null
  const cplx* p[kCgFinishMax];
  const double* alpha[kCgFinishMax];
  int nd;
  int64_t n;
  int L;                    // dst: packed real maps of band-limit L, n = L(L+1)/2
  cplx* dst;                // optional: x + sum alpha p unpacked to (L, 2L-1)
};
hipError_t launch_cg_finish(const CgFinishArgs& a, int batch, hipStream_t s);
constexpr int kSumsqBlocks = 256;  // slices per map of launch_batch_sumsq
// out[b] = sum_{i < n} x[b * stride + i]^2; partial holds kSumsqBlocks * batch doubles
hipError_t launch_batch_sumsq(const double* x, int64_t n, int64_t stride, int batch,
                              double* partial, double* out, hipStream_t s);
hipError_t launch_reality_fill(cplx* flm, int64_t map_stride, int L, int batch, hipStream_t s);
hipError_t launch_reality_fold(const cplx* g, cplx* out, int64_t map_stride, int L, int batch,
                               hipStream_t s);
hipError_t launch_scale_pos_m(cplx* flm, int64_t map_stride, int L, int batch, double factor,
                              hipStream_t s);

struct ResampleArgs {
  const cplx* in;           // phi-spectrum [b][row][col], col <-> m = m_lo + col
  int64_t in_map_stride;
  cplx* out;                // [b][2L+1][col]
  int64_t out_map_stride;
  int ncol;                 // columns of both arrays
  int m_lo, nm;
  int L, spin;
  int mw;                   // 1: MW input (L rows), 0: MWSS input (L+1 rows)
  int nyq_split;            // real signals: split the theta-Nyquist bin (see kernel)
  const double* weights;    // 2L+1 quadrature weights of MWSS at 2L (null: none)
  double phi_factor;        // 2L/(2L-1) for MW, 1 for MWSS
  int lds;                  // LDS elements (complex) per workgroup
  int conv[3];              // Bluestein lengths of the DFTs 2L-1, 2L, 4L (0 = mixed)
  int64_t chirp_off[3];
  const cplx* pool;
  int spin_step;            // map b has spin + spin_step b (multi-plane launches)
};
hipError_t launch_resample(const ResampleArgs& a, int batch, hipStream_t s);
hipError_t allow_large_lds_resample();
hipError_t launch_resample_adjoint(const ResampleArgs& a, int batch, hipStream_t s);
hipError_t allow_large_lds_resample_adjoint();

}  // namespace s2h
