// extern "C" entry points: argument checks, kernel sequencing, status codes.
#include <cstring>
#include <mutex>
#include <vector>

#include "s2h_plan.h"

using namespace s2h;

namespace {

thread_local int g_last_xla_status = S2H_OK;

int hip_status(hipError_t e) { return e == hipSuccess ? S2H_OK : S2H_ERR_HIP; }

constexpr int kMfmaMinBatch = 4;

int check_device(const s2h_plan* p) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return S2H_ERR_HIP;
  return dev == p->device ? S2H_OK : S2H_ERR_INVALID;
}

int64_t flm_stride(const s2h_plan* p) { return (int64_t)p->d.L * (2 * p->d.L - 1); }

LatArgs lat_args(const s2h_plan* p, bool forward, int batch) {
  LatArgs a{};
  a.rings = forward ? p->fwd_rings : p->inv_rings;
  a.nrings = forward ? p->n_fwd_rings : p->n_inv_rings;
  a.mstart = forward ? p->fwd_mstart : p->inv_mstart;
  a.coef = forward ? p->fwd_coef : p->inv_coef;
  a.ucoef = forward ? p->fwd_ucoef : p->inv_ucoef;
  a.L = p->d.L;
  a.m_lo = p->m_lo;
  a.nm = p->nm;
  a.lmin_out = p->d.L_lower;
  a.batch = batch;
  a.flm_map_stride = flm_stride(p);
  a.groups = 1;
  return a;
}

// Optional kernel timing: HIP events recorded on the launch stream around
// each kernel while enabled (s2h_profile_enable).  Off by default.
struct ProfRecord {
  int kernel;
  hipEvent_t start, stop;
};
std::mutex g_prof_mu;
bool g_prof_on = false;
std::vector<ProfRecord> g_prof_records;
std::vector<std::pair<hipEvent_t, hipEvent_t>> g_prof_pool;

struct ProfScope {
  int kernel;
  hipStream_t s;
  ProfRecord rec{};
  bool active = false;
  ProfScope(int k, hipStream_t st) : kernel(k), s(st) {
    std::lock_guard<std::mutex> lock(g_prof_mu);
    if (!g_prof_on) return;
    if (g_prof_pool.empty()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess) return;
      if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return;
      }
      g_prof_pool.push_back({a, b});
    }
    rec.kernel = kernel;
    rec.start = g_prof_pool.back().first;
    rec.stop = g_prof_pool.back().second;
    g_prof_pool.pop_back();
    active = hipEventRecord(rec.start, s) == hipSuccess;
  }
  ~ProfScope() {
    if (!active) return;
    (void)hipEventRecord(rec.stop, s);
    std::lock_guard<std::mutex> lock(g_prof_mu);
    g_prof_records.push_back(rec);
  }
};

}  // namespace

extern "C" int s2h_profile_enable(int on) {
  std::lock_guard<std::mutex> lock(g_prof_mu);
  for (auto& r : g_prof_records) g_prof_pool.push_back({r.start, r.stop});
  g_prof_records.clear();
  g_prof_on = on != 0;
  return S2H_OK;
}

extern "C" int s2h_profile_query(int kernel, double* total_ms, int* count) {
  if (!total_ms || !count) return S2H_ERR_INVALID;
  std::lock_guard<std::mutex> lock(g_prof_mu);
  double tot = 0.0;
  int n = 0;
  for (auto& r : g_prof_records) {
    if (r.kernel != kernel) continue;
    if (hipEventSynchronize(r.stop) != hipSuccess) return S2H_ERR_HIP;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, r.start, r.stop) != hipSuccess) return S2H_ERR_HIP;
    tot += ms;
    ++n;
  }
  *total_ms = tot;
  *count = n;
  return S2H_OK;
}

extern "C" const char* s2h_status_string(int s) {
  switch (s) {
    case S2H_OK: return "ok";
    case S2H_ERR_INVALID: return "invalid argument";
    case S2H_ERR_UNSUPPORTED: return "unsupported configuration";
    case S2H_ERR_HIP: return "HIP runtime error";
    case S2H_ERR_NOMEM: return "out of memory";
    case S2H_ERR_SPIN: return "Recursive transform may provide lower precision beyond spin ~ 8";
    case S2H_ERR_BANDLIMIT: return "HEALPix requires L >= 2 nside";
    case S2H_ERR_SCRATCH:
      return "workspace exceeds the fixed scratch of s2h_forward_noalloc "
             "(S2H_NOALLOC_SCRATCH_BYTES); use s2h_forward with a workspace";
    default: return "unknown status";
  }
}

extern "C" int s2h_version(void) { return 1; }

extern "C" size_t s2h_workspace_bytes(const s2h_plan* p, int batch) {
  if (!p || batch < 0) return 0;
  size_t rows = (size_t)std::max(p->fwd_rows, p->inv_rows);
  if (p->fwd_kind != 0) rows += (size_t)p->fwd_in_rows;
  // + one flm-sized buffer for the adjoint of the conjugate fill
  const size_t flm = (size_t)p->d.L * (2 * p->d.L - 1);
  return (rows * (size_t)p->Mc + flm) * sizeof(cplx) * (size_t)batch;
}

extern "C" int s2h_inverse_output_is_complex(const s2h_plan* p) {
  return p && p->inv_out_complex ? 1 : 0;
}

extern "C" int s2h_forward(const s2h_plan* p, const void* f, int f_is_complex, void* flm,
                           void* work, int batch, s2h_stream stream) {
  return s2h_forward_ex(p, f, f_is_complex, flm, work, batch, S2H_PREC_FP64, 0, stream);
}

namespace {

bool valid_precision(int precision) {
  return precision >= S2H_PREC_FP64 && precision <= S2H_PREC_BF16X1;
}

// The ring FFTs of f (weights and ring phases fused) into `out`: the plan's
// forward grid, or its input rows when the forward resamples.
int ring_analysis_stage(const s2h_plan* p, const void* f, int f_is_complex, cplx* out, int batch,
                        hipStream_t s, bool raw = false, int64_t out_map_stride = 0) {
  const bool real = p->d.reality != 0;
  const bool resample = p->fwd_kind != 0;
  RingFwdArgs r{};
  r.f = f;
  r.f_complex = f_is_complex;
  r.real_only_2L = 0;
  r.real_only_all = 0;
  if (real && f_is_complex) {
    // reference reality semantics: HEALPix rings of length 2L and every
    // equiangular ring use Re(f) (healpix_ffts.py:211-215, spherical.py:528-534)
    if (p->d.sampling == S2H_HEALPIX)
      r.real_only_2L = 1;
    else
      r.real_only_all = 1;
  }
  r.map_stride = p->npix;
  r.out = out;
  r.out_map_stride = out_map_stride > 0
                         ? out_map_stride
                         : (int64_t)(resample ? p->fwd_in_rows : p->fwd_rows) * p->Mc;
  r.ncol = p->Mc;
  r.m_first = p->m_lo;
  // raw: no quadrature weight and no ring phase (a stored precompute kernel
  // carries both)
  r.apply_weight = resample || raw ? 0 : 1;
  r.apply_phase = !raw && p->d.sampling == S2H_HEALPIX;
  r.zero_neg_2L = 0;
  r.L = p->d.L;
  r.chirp_pool = p->chirp;
  ProfScope ps(S2H_KERNEL_RING_FFT_FORWARD, s);
  return hip_status(launch_ring_fft_forward(r, p->fwd_set, batch, s));
}

// Latitudinal projection of forward-grid ring spectra (+ pole rows and the
// conjugate fill): flm = or += the forward transform of ftm.  The BF16
// kernels with one ring group per m write every entry themselves (zeros and
// the m < 0 mirror); otherwise flm is zeroed first and filled after.
// The ring spectra of the C ABI (s2h_ring_*): map stride and layout.  The
// internal stages default to the forward grid's rows, ring-major.
struct SpecView {
  int64_t map_stride;
  int tiled;
};

int lat_forward_stage(const s2h_plan* p, const cplx* ftm, void* flm, int batch, int precision,
                      bool accumulate, hipStream_t s, bool packed = false,
                      const PlaneSet* ps = nullptr, const SpecView* sv = nullptr) {
  const bool mixed = precision == S2H_PREC_MIXED;
  const bool bf16 = precision == S2H_PREC_BF16 || precision == S2H_PREC_BF16X1;
  const int split = precision == S2H_PREC_BF16X1 ? 1 : 3;
  const int L = p->d.L;
  const int64_t fs = packed ? (int64_t)L * (L + 1) / 2 : flm_stride(p);
  const int64_t ftm_stride = (int64_t)p->fwd_rows * p->Mc;
  int st = S2H_OK;
  LatArgs a = lat_args(p, true, batch);
  a.flm = static_cast<cplx*>(flm);
  a.flm_packed = packed ? 1 : 0;
  a.flm_map_stride = fs;
  a.ftm = const_cast<cplx*>(ftm);
  a.ftm_map_stride = sv ? sv->map_stride : ftm_stride;
  a.ftm_ncol = p->Mc;
  a.ftm_m_first = p->m_lo;
  a.ftm_tiled = sv ? sv->tiled : 0;
  a.lmin_out = std::max(std::abs(p->d.spin), p->d.L_lower);
  a.accumulate = accumulate ? 1 : 0;
  if (ps) a.planes = ps->fwd;
  const bool poles = p->fwd_row_south >= 0 || p->fwd_row_north >= 0;
  // the FP64 matrix-core projection (one workgroup per m) and the BF16 ones
  // with one ring group per m write every entry of their m columns
  const bool mfma64 = !ps && !bf16 && !mixed && batch >= kMfmaMinBatch;
  const bool fill =
      mfma64 || (!ps && bf16 && !poles && lat_forward_bf16_groups(a, split) == 1);
  a.fill = fill ? 1 : 0;
  a.mirror = fill && !poles && p->d.reality && !packed ? 1 : 0;
  if (!accumulate && !fill && (st = hip_status(launch_zero(flm, 2 * fs * batch, s)))) return st;
  {
    ProfScope prof(bf16 ? S2H_KERNEL_LAT_FORWARD_BF16
                        : mixed ? S2H_KERNEL_LAT_FORWARD_FP32 : S2H_KERNEL_LAT_FORWARD,
                   s);
    // batches of 4 or more maps share each d-value across an 8-map MFMA tile;
    // the mixed-precision projection always runs on the FP32 matrix cores
    // multi-plane launches: the VALU kernel reads each map's own tables
    if (ps)
      st = hip_status(launch_lat_forward(a, false, s));
    else if (bf16)
      st = hip_status(launch_lat_forward_bf16(a, p->fwd_paired, split, s));
    else if (mixed || batch >= kMfmaMinBatch)
      st = hip_status(launch_lat_forward_mfma(a, p->fwd_paired, mixed, s));
    else
      st = hip_status(launch_lat_forward(a, p->fwd_paired, s));
    if (st) return st;
  }

  if (p->fwd_row_south >= 0 || p->fwd_row_north >= 0) {
    PoleArgs pa{};
    pa.flm = static_cast<cplx*>(flm);
    pa.flm_map_stride = fs;
    pa.ftm = const_cast<cplx*>(ftm);
    pa.ftm_map_stride = ftm_stride;
    pa.ftm_ncol = p->Mc;
    pa.ftm_m_first = p->m_lo;
    pa.L = L;
    pa.spin = ps ? ps->spin0 : p->d.spin;
    pa.spin_step = ps ? ps->spin_step : 0;
    pa.lmin = ps ? p->d.L_lower : a.lmin_out;
    pa.lmin_spin = ps ? 1 : 0;
    pa.row_south = p->fwd_row_south;
    pa.row_north = p->fwd_row_north;
    pa.norm = p->fwd_norm;
    pa.batch = batch;
    if ((st = hip_status(launch_pole_forward(pa, s)))) return st;
  }

  if (p->d.reality && !a.mirror && !packed) {
    if ((st = hip_status(launch_reality_fill(static_cast<cplx*>(flm), fs, L, batch, s)))) return st;
  }
  return S2H_OK;
}

// Latitudinal synthesis (+ pole rows) into inverse-grid ring spectra ftm.
int lat_inverse_stage(const s2h_plan* p, const void* flm, cplx* ftm, int batch, int precision,
                      hipStream_t s, bool packed = false, const PlaneSet* ps = nullptr,
                      const SpecView* sv = nullptr, double* dot_part = nullptr,
                      int64_t dot_part_stride = 0) {
  const bool mixed = precision == S2H_PREC_MIXED;
  const bool bf16 = precision == S2H_PREC_BF16 || precision == S2H_PREC_BF16X1;
  const int64_t ftm_stride = (int64_t)p->inv_rows * p->Mc;
  int st = S2H_OK;
  LatArgs a = lat_args(p, false, batch);
  a.flm = const_cast<cplx*>(static_cast<const cplx*>(flm));
  a.flm_packed = packed ? 1 : 0;
  if (packed) a.flm_map_stride = (int64_t)p->d.L * (p->d.L + 1) / 2;
  a.ftm = ftm;
  a.ftm_map_stride = sv ? sv->map_stride : ftm_stride;
  a.ftm_ncol = p->Mc;
  a.ftm_m_first = p->m_lo;
  a.ftm_tiled = sv ? sv->tiled : 0;
  if (ps) a.planes = ps->inv;
  if (dot_part) {
    a.dot_wn = p->spec_belt_wn;
    a.dot_row_lo = p->spec_belt_row_lo;
    a.dot_row_hi = p->spec_belt_row_hi;
    a.dot_part = dot_part;
    a.dot_part_stride = dot_part_stride;
  }
  {
    ProfScope prof(bf16 ? S2H_KERNEL_LAT_INVERSE_BF16
                        : mixed ? S2H_KERNEL_LAT_INVERSE_FP32 : S2H_KERNEL_LAT_INVERSE,
                   s);
    // batches of 4 or more maps: each d-value feeds 16 maps through the MFMA;
    // multi-plane launches: the VALU kernel reads each map's own tables
    if (ps)
      st = hip_status(launch_lat_inverse(a, false, s));
    else if (bf16)
      st = hip_status(launch_lat_inverse_bf16(a, p->inv_paired,
                                              precision == S2H_PREC_BF16X1 ? 1 : 3, s));
    else if (mixed || batch >= kMfmaMinBatch)
      st = hip_status(launch_lat_inverse_mma(a, p->inv_paired, mixed, s));
    else
      st = hip_status(launch_lat_inverse(a, p->inv_paired, s));
    if (st) return st;
  }

  if (p->inv_row_south >= 0 || p->inv_row_north >= 0) {
    PoleArgs pa{};
    pa.flm = a.flm;
    pa.flm_map_stride = flm_stride(p);
    pa.ftm = ftm;
    pa.ftm_map_stride = ftm_stride;
    pa.ftm_ncol = p->Mc;
    pa.ftm_m_first = p->m_lo;
    pa.L = p->d.L;
    pa.spin = ps ? ps->spin0 : p->d.spin;
    pa.spin_step = ps ? ps->spin_step : 0;
    pa.lmin = p->d.L_lower;
    pa.row_south = p->inv_row_south;
    pa.row_north = p->inv_row_north;
    pa.norm = p->inv_norm;
    pa.batch = batch;
    if ((st = hip_status(launch_pole_inverse(pa, s)))) return st;
  }
  return S2H_OK;
}

// Resampling arguments (MW / MWSS forward): phi-spectra of the input rows in
// `spec`, the MWSS-at-2L grid into `ftm`.
ResampleArgs resample_args(const s2h_plan* p, const cplx* spec, cplx* ftm, const PlaneSet* ps) {
  const int L = p->d.L;
  ResampleArgs ra{};
  ra.in = spec;
  ra.in_map_stride = (int64_t)p->fwd_in_rows * p->Mc;
  ra.out = ftm;
  ra.out_map_stride = (int64_t)p->fwd_rows * p->Mc;
  ra.ncol = p->Mc;
  ra.m_lo = p->m_lo;
  ra.nm = p->nm;
  ra.L = L;
  ra.spin = ps ? ps->spin0 : p->d.spin;
  ra.spin_step = ps ? ps->spin_step : 0;
  ra.mw = p->fwd_kind == 1;
  ra.nyq_split = p->d.reality ? 1 : 0;
  ra.weights = p->fwd_weights;
  ra.phi_factor = ra.mw ? (2.0 * L) / (2.0 * L - 1.0) : 1.0;
  ra.lds = p->rs_info[3];
  for (int i = 0; i < 3; ++i) {
    ra.conv[i] = p->rs_info[i];
    ra.chirp_off[i] = p->rs_chirp[i];
  }
  ra.pool = p->chirp;
  return ra;
}

// Plans whose forward and inverse share one ring grid without pole rows or
// resampling (HEALPix, GL): their ring spectra compose directly.
// The alias kernel's residue arithmetic needs ring length + 2L < 2^16.
bool spectral_ok(const s2h_plan* p) {
  return p->fwd_kind == 0 && p->fwd_rows == p->inv_rows && p->fwd_row_south < 0 &&
         p->fwd_row_north < 0 && p->inv_row_south < 0 && p->inv_row_north < 0 &&
         p->d.norm != S2H_NORM_NONE && p->fwd_set.max_nphi + 2 * p->d.L < (1 << 16);
}

}  // namespace

namespace {

// The forward pipeline: ring FFTs (+ MW / MWSS resampling) and the
// latitudinal projection; with ps, the batch is a set of n-planes.
int forward_impl(const s2h_plan* p, const PlaneSet* ps, const void* f, int f_is_complex,
                 void* flm, void* work, int batch, int precision, bool accumulate,
                 hipStream_t s) {
  int st;
  cplx* ftm = static_cast<cplx*>(work);
  const bool resample = p->fwd_kind != 0;
  // resampled forward: phi-spectrum of the input rows first, then the
  // resampled MWSS-at-2L ftm after it in the workspace
  cplx* spec = ftm;
  if (resample) ftm = spec + (int64_t)batch * p->fwd_in_rows * p->Mc;
  if ((st = ring_analysis_stage(p, f, f_is_complex, resample ? spec : ftm, batch, s))) return st;
  if (resample) {
    const ResampleArgs ra = resample_args(p, spec, ftm, ps);
    ProfScope prof(S2H_KERNEL_RESAMPLE, s);
    if ((st = hip_status(launch_resample(ra, batch, s)))) return st;
  }
  return lat_forward_stage(p, ftm, flm, batch, precision, accumulate, s, false, ps);
}

}  // namespace

extern "C" int s2h_forward_ex(const s2h_plan* p, const void* f, int f_is_complex, void* flm,
                              void* work, int batch, int precision, int flags,
                              s2h_stream stream) {
  if (!p || !f || !flm || batch < 0 || (batch > 0 && !work)) return S2H_ERR_INVALID;
  if (!valid_precision(precision)) return S2H_ERR_INVALID;
  if (flags & ~S2H_FWD_ACCUMULATE) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  if (p->d.norm == S2H_NORM_NONE) return S2H_ERR_INVALID;
  return forward_impl(p, nullptr, f, f_is_complex, flm, work, batch, precision,
                      (flags & S2H_FWD_ACCUMULATE) != 0, static_cast<hipStream_t>(stream));
}

extern "C" int s2h_inverse(const s2h_plan* p, const void* flm, void* f, void* work, int batch,
                           s2h_stream stream) {
  return s2h_inverse_ex(p, flm, f, work, batch, S2H_PREC_FP64, 0, nullptr, stream);
}

namespace {

// The inverse pipeline: latitudinal synthesis (+ poles) and the ring iFFTs;
// with ps, the batch is a set of n-planes (complex output).
int inverse_impl(const s2h_plan* p, const PlaneSet* ps, const void* flm, void* f, void* work,
                 int batch, int precision, bool real_out, const void* sub_from, hipStream_t s) {
  int st;
  const int L = p->d.L;
  cplx* ftm = static_cast<cplx*>(work);
  const int64_t ftm_stride = (int64_t)p->inv_rows * p->Mc;
  if ((st = lat_inverse_stage(p, flm, ftm, batch, precision, s, false, ps))) return st;
  RingInvArgs r{};
  r.in = ftm;
  r.in_map_stride = ftm_stride;
  r.ncol = p->Mc;
  r.m_first = p->m_lo;
  r.m_lo = -(L - 1);
  r.m_hi = L - 1;
  r.herm_mode = p->d.reality ? 1 : 0;
  r.apply_phase = p->d.sampling == S2H_HEALPIX;
  r.apply_weight = 0;
  r.f = f;
  r.f_complex = p->inv_out_complex && !real_out ? 1 : 0;
  r.map_stride = p->npix;
  r.chirp_pool = p->chirp;
  r.L = L;
  r.sub_from = sub_from;
  ProfScope prof(S2H_KERNEL_RING_FFT_INVERSE, s);
  return hip_status(launch_ring_fft_inverse(r, p->inv_set, batch, s));
}

}  // namespace

extern "C" int s2h_inverse_ex(const s2h_plan* p, const void* flm, void* f, void* work, int batch,
                              int precision, int flags, const void* sub_from,
                              s2h_stream stream) {
  if (!p || !flm || !f || batch < 0 || (batch > 0 && !work)) return S2H_ERR_INVALID;
  if (flags & ~(S2H_INV_REAL_OUTPUT | S2H_INV_SUBTRACT)) return S2H_ERR_INVALID;
  const bool real_out = (flags & S2H_INV_REAL_OUTPUT) != 0;
  if (real_out && !p->d.reality) return S2H_ERR_INVALID;
  if ((flags & S2H_INV_SUBTRACT) && !sub_from) return S2H_ERR_INVALID;
  if (!valid_precision(precision)) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  if (p->d.norm == S2H_NORM_NONE) return S2H_ERR_INVALID;
  return inverse_impl(p, nullptr, flm, f, work, batch, precision, real_out,
                      (flags & S2H_INV_SUBTRACT) ? sub_from : nullptr,
                      static_cast<hipStream_t>(stream));
}

// ---------------------------------------------------------------- ring spectra
// The transforms split at the ring spectra, for refinement in spectral space:
// F(f - I(x)) = Lat_fwd(G - A(Lat_inv(x))) with G the ring spectra of f and A
// the ring FFT of the ring iFFT, applied without either FFT.
namespace {
// map stride of the C ABI's ring spectra (rows padded to whole row blocks)
int64_t spec_stride(const s2h_plan* p) { return (int64_t)p->spec_rows * p->Mc; }
SpecView spec_view(const s2h_plan* p, int flags) {
  return SpecView{spec_stride(p), (flags & S2H_SPECTRA_TILED) ? 1 : 0};
}
// zeros into the padding rows of C-ABI spectra the call wrote
int zero_spec_pad(const s2h_plan* p, void* spec, int tiled, int batch, hipStream_t s) {
  return hip_status(launch_spec_pad_zero(static_cast<cplx*>(spec), spec_stride(p), p->Mc,
                                         p->fwd_rows, p->spec_rows, tiled, batch, s));
}
}  // namespace

extern "C" size_t s2h_ring_spectra_bytes(const s2h_plan* p, int batch) {
  if (!p || batch < 0 || !spectral_ok(p)) return 0;
  return (size_t)batch * (size_t)spec_stride(p) * sizeof(cplx);
}

extern "C" int s2h_ring_analysis(const s2h_plan* p, const void* f, int f_is_complex, void* G,
                                 int batch, s2h_stream stream) {
  if (!p || !f || !G || batch < 0) return S2H_ERR_INVALID;
  if (!spectral_ok(p)) return S2H_ERR_UNSUPPORTED;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  if ((st = ring_analysis_stage(p, f, f_is_complex, static_cast<cplx*>(G), batch,
                                static_cast<hipStream_t>(stream), false, spec_stride(p))))
    return st;
  return zero_spec_pad(p, G, 0, batch, static_cast<hipStream_t>(stream));
}

extern "C" int s2h_forward_rings(const s2h_plan* p, const void* G, void* flm, int batch,
                                 int precision, int flags, s2h_stream stream) {
  if (!p || !G || !flm || batch < 0) return S2H_ERR_INVALID;
  if (!valid_precision(precision) ||
      (flags & ~(S2H_FWD_ACCUMULATE | S2H_FLM_PACKED | S2H_SPECTRA_TILED)))
    return S2H_ERR_INVALID;
  const bool packed = (flags & S2H_FLM_PACKED) != 0;
  if (packed && !p->d.reality) return S2H_ERR_INVALID;
  if (!spectral_ok(p)) return S2H_ERR_UNSUPPORTED;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const bool accumulate = (flags & S2H_FWD_ACCUMULATE) != 0;
  const SpecView sv = spec_view(p, flags);
  return lat_forward_stage(p, static_cast<const cplx*>(G), flm, batch, precision, accumulate, s,
                           packed, nullptr, &sv);
}

extern "C" int s2h_inverse_rings(const s2h_plan* p, const void* flm, void* Y, int batch,
                                 int precision, int flags, s2h_stream stream) {
  if (!p || !flm || !Y || batch < 0 || !valid_precision(precision)) return S2H_ERR_INVALID;
  if (flags & ~(S2H_FLM_PACKED | S2H_SPECTRA_TILED)) return S2H_ERR_INVALID;
  const bool packed = (flags & S2H_FLM_PACKED) != 0;
  if (packed && !p->d.reality) return S2H_ERR_INVALID;
  if (!spectral_ok(p)) return S2H_ERR_UNSUPPORTED;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  const SpecView sv = spec_view(p, flags);
  if ((st = lat_inverse_stage(p, flm, static_cast<cplx*>(Y), batch, precision,
                              static_cast<hipStream_t>(stream), packed, nullptr, &sv)))
    return st;
  return zero_spec_pad(p, Y, sv.tiled, batch, static_cast<hipStream_t>(stream));
}

extern "C" int s2h_flm_pack(int L, const void* flm, void* flm_p, int batch, s2h_stream stream) {
  if (L < 1 || batch < 0 || (batch > 0 && (!flm || !flm_p)) || flm == flm_p)
    return S2H_ERR_INVALID;
  return hip_status(launch_flm_pack(L, static_cast<const cplx*>(flm), static_cast<cplx*>(flm_p),
                                    batch, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_flm_unpack(int L, const void* flm_p, void* flm, int batch, s2h_stream stream) {
  if (L < 1 || batch < 0 || (batch > 0 && (!flm || !flm_p)) || flm == flm_p)
    return S2H_ERR_INVALID;
  return hip_status(launch_flm_unpack(L, static_cast<const cplx*>(flm_p), static_cast<cplx*>(flm),
                                      batch, static_cast<hipStream_t>(stream)));
}

namespace {

RingAliasArgs alias_args(const s2h_plan* p, const void* G, const void* Y, void* R) {
  RingAliasArgs a{};
  a.y = static_cast<const cplx*>(Y);
  a.g = static_cast<const cplx*>(G);
  a.r = static_cast<cplx*>(R);
  a.map_stride = spec_stride(p);
  a.ncol = p->Mc;
  a.m_first = p->m_lo;
  a.herm = p->d.reality ? 1 : 0;
  a.apply_phase = p->d.sampling == S2H_HEALPIX;
  a.L = p->d.L;
  return a;
}

}  // namespace

namespace {

// Tiled alias arguments (real maps only: Hermitian spectra, m in [0, L)).
RingAliasTiledArgs alias_tiled_args(const s2h_plan* p, const void* G, const void* Y, void* R,
                                    bool g_tiled) {
  RingAliasTiledArgs a{};
  a.rowjobs = p->fwd_rowjobs;
  a.nblocks = p->spec_rows / kSpecBlock;
  a.belt_blocks = p->spec_belt_blocks;
  a.n_belt = p->n_belt_blocks;
  a.polar_blocks = p->spec_polar_blocks;
  a.n_polar = p->n_polar_blocks;
  a.y = static_cast<const cplx*>(Y);
  a.g = static_cast<const cplx*>(G);
  a.r = static_cast<cplx*>(R);
  a.g_tiled = g_tiled ? 1 : 0;
  a.map_stride = spec_stride(p);
  a.ncol = p->Mc;
  a.apply_phase = p->d.sampling == S2H_HEALPIX;
  a.L = p->d.L;
  return a;
}

}  // namespace

extern "C" int s2h_ring_residual_ex(const s2h_plan* p, const void* G, const void* Y,
                                    const double* alpha, void* R, double* sumsq, void* work,
                                    int batch, int flags, s2h_stream stream) {
  if (!p || !G || !Y || !R || batch < 0 || Y == R) return S2H_ERR_INVALID;
  if (flags & ~(S2H_SPECTRA_TILED | S2H_SPECTRA_G_RING_MAJOR)) return S2H_ERR_INVALID;
  const bool tiled = (flags & S2H_SPECTRA_TILED) != 0;
  const bool g_rm = (flags & S2H_SPECTRA_G_RING_MAJOR) != 0;
  if (g_rm && (!tiled || G == R)) return S2H_ERR_INVALID;
  if (!spectral_ok(p)) return S2H_ERR_UNSUPPORTED;
  if (tiled && (!p->d.reality || p->m_lo != 0)) return S2H_ERR_UNSUPPORTED;
  if (batch == 0) return S2H_OK;
  if (sumsq && !work) return S2H_ERR_INVALID;
  int st = check_device(p);
  if (st) return st;
  ProfScope ps(S2H_KERNEL_RING_RESIDUAL, static_cast<hipStream_t>(stream));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (tiled) {
    RingAliasTiledArgs a = alias_tiled_args(p, G, Y, R, !g_rm);
    a.alpha = alpha;
    st = sumsq ? hip_status(
                     launch_ring_alias_tiled_sumsq(a, batch, static_cast<double*>(work), sumsq, s))
               : hip_status(launch_ring_alias_tiled(a, batch, s));
  } else {
    RingAliasArgs a = alias_args(p, G, Y, R);
    a.alpha = alpha;
    st = sumsq ? hip_status(launch_ring_alias_sumsq(a, p->fwd_set, batch,
                                                    static_cast<double*>(work), sumsq, s))
               : hip_status(launch_ring_alias(a, p->fwd_set, batch, s));
  }
  if (st || R == G) return st;
  return zero_spec_pad(p, R, tiled ? 1 : 0, batch, s);
}

extern "C" int s2h_ring_residual_scaled(const s2h_plan* p, const void* G, const void* Y,
                                        const double* alpha, void* R, double* sumsq, void* work,
                                        int batch, s2h_stream stream) {
  return s2h_ring_residual_ex(p, G, Y, alpha, R, sumsq, work, batch, 0, stream);
}

extern "C" int s2h_ring_residual(const s2h_plan* p, const void* G, const void* Y, void* R,
                                 int batch, s2h_stream stream) {
  return s2h_ring_residual_scaled(p, G, Y, nullptr, R, nullptr, nullptr, batch, stream);
}

extern "C" size_t s2h_ring_dot_workspace_bytes(const s2h_plan* p, int batch) {
  if (!p || batch <= 0 || !spectral_ok(p)) return 0;
  // one partial per ring job (ring-major) or per row block (tiled)
  const int parts =
      std::max(ring_dot_jobs(p->fwd_set), ring_dot_tiled_parts(p->spec_rows / kSpecBlock));
  return (size_t)batch * (size_t)parts * sizeof(double);
}

extern "C" int s2h_ring_dot_ex(const s2h_plan* p, const void* Y, const double* num, double* pq,
                               double* alpha, void* work, int batch, int flags,
                               s2h_stream stream) {
  if (!p || !Y || batch < 0 || (!pq && !alpha) || (alpha && !num)) return S2H_ERR_INVALID;
  if (flags & ~S2H_SPECTRA_TILED) return S2H_ERR_INVALID;
  const bool tiled = (flags & S2H_SPECTRA_TILED) != 0;
  if (!spectral_ok(p)) return S2H_ERR_UNSUPPORTED;
  if (tiled && (!p->d.reality || p->m_lo != 0)) return S2H_ERR_UNSUPPORTED;
  if (batch == 0) return S2H_OK;
  if (!work) return S2H_ERR_INVALID;
  int st = check_device(p);
  if (st) return st;
  ProfScope ps(S2H_KERNEL_RING_RESIDUAL, static_cast<hipStream_t>(stream));
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (tiled) {
    const RingAliasTiledArgs a = alias_tiled_args(p, Y, Y, nullptr, true);
    return hip_status(
        launch_ring_dot_tiled(a, batch, static_cast<double*>(work), num, pq, alpha, s));
  }
  const RingAliasArgs a = alias_args(p, Y, Y, nullptr);
  return hip_status(launch_ring_dot(a, p->fwd_set, batch, static_cast<double*>(work), num, pq,
                                    alpha, s));
}

extern "C" int s2h_ring_dot(const s2h_plan* p, const void* Y, const double* num, double* pq,
                            double* alpha, void* work, int batch, s2h_stream stream) {
  return s2h_ring_dot_ex(p, Y, num, pq, alpha, work, batch, 0, stream);
}

namespace {
// Whether the synthesis of this call runs on the matrix-core kernels, whose
// epilogue adds up the dot over the rows without aliasing; their grid.x.
int synth_dot_workgroups(const s2h_plan* p, int batch, int precision) {
  const bool bf16 = precision == S2H_PREC_BF16 || precision == S2H_PREC_BF16X1;
  const bool mma = precision == S2H_PREC_MIXED || batch >= kMfmaMinBatch;
  if ((!bf16 && !mma) || p->spec_belt_wn == 0.0) return 0;
  return p->nm * ((p->n_inv_rings + 255) / 256);
}
}  // namespace

extern "C" size_t s2h_inverse_rings_dot_workspace_bytes(const s2h_plan* p, int batch) {
  if (!p || batch <= 0 || !spectral_ok(p)) return 0;
  const int64_t head = (int64_t)p->nm * ((p->n_inv_rings + 255) / 256);
  const int64_t parts =
      std::max<int64_t>(head + (int64_t)p->n_polar_blocks * ring_dot_tiled_parts(1),
                        (int64_t)ring_dot_tiled_parts(p->spec_rows / kSpecBlock));
  return (size_t)batch * (size_t)parts * sizeof(double);
}

extern "C" int s2h_inverse_rings_dot(const s2h_plan* p, const void* flm, void* Y, int batch,
                                     int precision, int flags, const double* num, double* pq,
                                     double* alpha, void* work, s2h_stream stream) {
  if (!p || !flm || !Y || batch < 0 || !valid_precision(precision)) return S2H_ERR_INVALID;
  if (flags & ~(S2H_FLM_PACKED | S2H_SPECTRA_TILED)) return S2H_ERR_INVALID;
  if (!(flags & S2H_SPECTRA_TILED) || (!pq && !alpha) || (alpha && !num)) return S2H_ERR_INVALID;
  const bool packed = (flags & S2H_FLM_PACKED) != 0;
  if (!spectral_ok(p)) return S2H_ERR_UNSUPPORTED;
  if (!p->d.reality || p->m_lo != 0) return S2H_ERR_UNSUPPORTED;
  if (batch == 0) return S2H_OK;
  if (!work) return S2H_ERR_INVALID;
  int st = check_device(p);
  if (st) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const SpecView sv = spec_view(p, flags);
  double* part = static_cast<double*>(work);
  const int head = synth_dot_workgroups(p, batch, precision);
  const int64_t stride = head + (int64_t)p->n_polar_blocks * ring_dot_tiled_parts(1);
  if ((st = lat_inverse_stage(p, flm, static_cast<cplx*>(Y), batch, precision, s, packed, nullptr,
                              &sv, head > 0 ? part : nullptr, stride)))
    return st;
  if ((st = zero_spec_pad(p, Y, 1, batch, s))) return st;
  ProfScope ps(S2H_KERNEL_RING_RESIDUAL, s);
  const RingAliasTiledArgs a = alias_tiled_args(p, Y, Y, nullptr, true);
  if (head == 0)  // vector-ALU synthesis: the whole dot over the spectra
    return hip_status(launch_ring_dot_tiled(a, batch, part, num, pq, alpha, s));
  return hip_status(launch_ring_dot_tiled_polar(a, batch, part, head, num, pq, alpha, s));
}

namespace {

int forward_adjoint_impl(const s2h_plan* p, const PlaneSet* ps, const void* g, void* grad_f,
                         void* work, int batch, hipStream_t s) {
  int st;
  const int L = p->d.L;
  const bool real = p->d.reality != 0;
  const int64_t fs = flm_stride(p);
  const bool resample = p->fwd_kind != 0;
  cplx* spec = static_cast<cplx*>(work);
  cplx* ftm = resample ? spec + (int64_t)batch * p->fwd_in_rows * p->Mc : spec;
  const int64_t ftm_stride = (int64_t)p->fwd_rows * p->Mc;
  const cplx* gin = static_cast<const cplx*>(g);
  if (real) {
    size_t rows = (size_t)std::max(p->fwd_rows, p->inv_rows) + (resample ? p->fwd_in_rows : 0);
    cplx* folded = spec + (int64_t)batch * rows * p->Mc;
    if ((st = hip_status(launch_reality_fold(gin, folded, fs, L, batch, s)))) return st;
    gin = folded;
  }
  LatArgs a = lat_args(p, true, batch);
  a.flm = const_cast<cplx*>(gin);
  a.ftm = ftm;
  a.ftm_map_stride = ftm_stride;
  a.ftm_ncol = p->Mc;
  a.ftm_m_first = p->m_lo;
  a.lmin_out = std::max(std::abs(p->d.spin), p->d.L_lower);
  if (ps) a.planes = ps->fwd;
  if ((st = hip_status(launch_lat_inverse(a, ps ? false : p->fwd_paired, s)))) return st;
  if (p->fwd_row_south >= 0 || p->fwd_row_north >= 0) {
    // adjoint of the forward pole terms: the pole rows of the synthesis
    PoleArgs pa{};
    pa.flm = const_cast<cplx*>(gin);
    pa.flm_map_stride = fs;
    pa.ftm = ftm;
    pa.ftm_map_stride = ftm_stride;
    pa.ftm_ncol = p->Mc;
    pa.ftm_m_first = p->m_lo;
    pa.L = L;
    pa.spin = ps ? ps->spin0 : p->d.spin;
    pa.spin_step = ps ? ps->spin_step : 0;
    pa.lmin = ps ? p->d.L_lower : a.lmin_out;
    pa.lmin_spin = ps ? 1 : 0;
    pa.row_south = p->fwd_row_south;
    pa.row_north = p->fwd_row_north;
    pa.norm = p->fwd_norm;
    pa.batch = batch;
    if ((st = hip_status(launch_pole_inverse(pa, s)))) return st;
  }
  int64_t in_stride = ftm_stride;
  if (resample) {
    const ResampleArgs ra = resample_args(p, spec, ftm, ps);
    if ((st = hip_status(launch_resample_adjoint(ra, batch, s)))) return st;
    in_stride = ra.in_map_stride;
  }

  RingInvArgs r{};
  r.in = resample ? spec : ftm;
  r.in_map_stride = in_stride;
  r.ncol = p->Mc;
  r.m_first = p->m_lo;
  r.m_lo = -(L - 1);
  r.m_hi = L - 1;
  r.herm_mode = real ? 1 : 0;
  r.apply_phase = p->d.sampling == S2H_HEALPIX;
  r.apply_weight = resample ? 0 : 1;
  r.f = grad_f;
  r.f_complex = real ? 0 : 1;
  r.map_stride = p->npix;
  r.chirp_pool = p->chirp;
  r.L = L;
  return hip_status(launch_ring_fft_inverse(r, p->fwd_set, batch, s));
}

int inverse_adjoint_impl(const s2h_plan* p, const PlaneSet* ps, const void* g, int g_is_complex,
                         void* grad_flm, void* work, int batch, hipStream_t s) {
  int st;
  const int L = p->d.L;
  const bool real = p->d.reality != 0;
  const int64_t fs = flm_stride(p);
  cplx* ftm = static_cast<cplx*>(work);
  const int64_t ftm_stride = (int64_t)p->inv_rows * p->Mc;
  if ((st = hip_status(launch_zero(grad_flm, 2 * fs * batch, s)))) return st;

  RingFwdArgs r{};
  r.f = g;
  r.f_complex = g_is_complex;
  r.real_only_all = (real && g_is_complex) ? 1 : 0;
  r.map_stride = p->npix;
  r.out = ftm;
  r.out_map_stride = ftm_stride;
  r.ncol = p->Mc;
  r.m_first = p->m_lo;
  r.apply_weight = 0;
  r.apply_phase = p->d.sampling == S2H_HEALPIX;
  r.L = L;
  r.chirp_pool = p->chirp;
  if ((st = hip_status(launch_ring_fft_forward(r, p->inv_set, batch, s)))) return st;

  LatArgs a = lat_args(p, false, batch);
  a.flm = static_cast<cplx*>(grad_flm);
  a.ftm = ftm;
  a.ftm_map_stride = ftm_stride;
  a.ftm_ncol = p->Mc;
  a.ftm_m_first = p->m_lo;
  a.lmin_out = p->d.L_lower;
  if (ps) a.planes = ps->inv;
  if ((st = hip_status(launch_lat_forward(a, ps ? false : p->inv_paired, s)))) return st;
  if (p->inv_row_south >= 0 || p->inv_row_north >= 0) {
    PoleArgs pa{};
    pa.flm = static_cast<cplx*>(grad_flm);
    pa.flm_map_stride = fs;
    pa.ftm = ftm;
    pa.ftm_map_stride = ftm_stride;
    pa.ftm_ncol = p->Mc;
    pa.ftm_m_first = p->m_lo;
    pa.L = L;
    pa.spin = ps ? ps->spin0 : p->d.spin;
    pa.spin_step = ps ? ps->spin_step : 0;
    pa.lmin = p->d.L_lower;
    pa.row_south = p->inv_row_south;
    pa.row_north = p->inv_row_north;
    pa.norm = p->inv_norm;
    pa.batch = batch;
    if ((st = hip_status(launch_pole_forward(pa, s)))) return st;
  }
  if (real) {
    if ((st = hip_status(launch_scale_pos_m(static_cast<cplx*>(grad_flm), fs, L, batch, 2.0, s))))
      return st;
  }
  return S2H_OK;
}

}  // namespace

// Adjoint of s2h_forward (the VJP of a complex-linear map, torch convention):
// grad_f = F^H g.  F = sign N_l Lat_fwd . (e^{-im phi0} w_t) . DFT_ext (+ poles),
// so F^H = DFT_ext^H . (w_t e^{+im phi0}) . Lat_inv with the forward plan's
// tables: the ring iFFT with folding and weights after the latitudinal
// synthesis on the forward grid.  With reality the fill's adjoint folds the
// m < 0 cotangents onto m > 0 and the result is real.
extern "C" int s2h_forward_adjoint(const s2h_plan* p, const void* g, void* grad_f, void* work,
                                   int batch, s2h_stream stream) {
  if (!p || !g || !grad_f || batch < 0 || (batch > 0 && !work)) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  if (p->d.norm == S2H_NORM_NONE) return S2H_ERR_INVALID;
  return forward_adjoint_impl(p, nullptr, g, grad_f, work, batch,
                              static_cast<hipStream_t>(stream));
}

// Adjoint of s2h_inverse: grad_flm = I^H g = sign N_l Lat_fwd . (e^{-im phi0})
// . DFT_ext on the inverse grid, plus the pole rows' adjoint.  With reality
// the inverse reads only m >= 0 and completes Hermitian-symmetrically, so the
// m > 0 cotangents double and m < 0 stay zero.
extern "C" int s2h_inverse_adjoint(const s2h_plan* p, const void* g, int g_is_complex,
                                   void* grad_flm, void* work, int batch, s2h_stream stream) {
  if (!p || !g || !grad_flm || batch < 0 || (batch > 0 && !work)) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  if (p->d.norm == S2H_NORM_NONE) return S2H_ERR_INVALID;
  return inverse_adjoint_impl(p, nullptr, g, g_is_complex, grad_flm, work, batch,
                              static_cast<hipStream_t>(stream));
}

extern "C" int s2h_lat_forward(const s2h_plan* p, const void* ftm, void* flm, int batch,
                               s2h_stream stream) {
  if (!p || !ftm || !flm || batch < 0) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  if (p->d.norm != S2H_NORM_NONE) return S2H_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int L = p->d.L;
  const bool wide = p->d.sampling == S2H_MWSS || p->d.sampling == S2H_HEALPIX;
  const int ncol = wide ? 2 * L : 2 * L - 1;
  if ((st = hip_status(launch_zero(flm, 2 * flm_stride(p) * batch, s)))) return st;
  LatArgs a = lat_args(p, true, batch);
  a.rings = p->inv_rings;
  a.nrings = p->n_inv_rings;
  a.flm = static_cast<cplx*>(flm);
  a.ftm = const_cast<cplx*>(static_cast<const cplx*>(ftm));
  a.ftm_map_stride = (int64_t)p->inv_rows * ncol;
  a.ftm_ncol = ncol;
  a.ftm_m_first = wide ? -L : -(L - 1);
  return hip_status(launch_lat_forward(a, p->inv_paired, s));
}

extern "C" int s2h_lat_inverse(const s2h_plan* p, const void* flm, void* ftm, int batch,
                               s2h_stream stream) {
  if (!p || !ftm || !flm || batch < 0) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  if (p->d.norm != S2H_NORM_NONE) return S2H_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int L = p->d.L;
  const bool wide = p->d.sampling == S2H_MWSS || p->d.sampling == S2H_HEALPIX;
  const int ncol = wide ? 2 * L : 2 * L - 1;
  const int64_t stride = (int64_t)p->inv_rows * ncol;
  if ((st = hip_status(launch_zero(ftm, 2 * stride * batch, s)))) return st;
  LatArgs a = lat_args(p, false, batch);
  a.flm = const_cast<cplx*>(static_cast<const cplx*>(flm));
  a.ftm = static_cast<cplx*>(ftm);
  a.ftm_map_stride = stride;
  a.ftm_ncol = ncol;
  a.ftm_m_first = wide ? -L : -(L - 1);
  return hip_status(launch_lat_inverse(a, p->inv_paired, s));
}

namespace {

int healpix_fft_checks(int nside, int L, int norm, int batch) {
  if (nside < 1 || batch < 0) return S2H_ERR_INVALID;
  if (norm < S2H_FFT_NORM_FORWARD || norm > S2H_FFT_NORM_NONE) return S2H_ERR_INVALID;
  if (L < 2 * nside) return S2H_ERR_BANDLIMIT;
  if (4 * nside > kMaxRingLen || L > kMaxL) return S2H_ERR_UNSUPPORTED;
  return S2H_OK;
}

// The per-ring scale of a transform direction under a norm (the reference's
// cuFFT store callbacks, lib/src/s2fft_callbacks.cu:256-308): the forward
// divides by n under FORWARD, the backward under BACKWARD, both by sqrt(n)
// under ORTHO; NONE scales neither.  0 none, 1 1/n, 2 1/sqrt(n).
int norm_mode(int norm, bool forward) {
  if (norm == S2H_FFT_NORM_ORTHO) return 2;
  if (norm == (forward ? S2H_FFT_NORM_FORWARD : S2H_FFT_NORM_BACKWARD)) return 1;
  return 0;
}

}  // namespace

extern "C" int s2h_healpix_fft_prepare(int nside, int L) {
  const int st = healpix_fft_checks(nside, L, S2H_FFT_NORM_NONE, 0);
  if (st) return st;
  return healpix_tables_prepare(nside, L);
}

extern "C" int s2h_healpix_fft_ex(int nside, int L, int reality, int norm, int double_precision,
                                  const void* f, void* ftm, int batch, s2h_stream stream) {
  int st = healpix_fft_checks(nside, L, norm, batch);
  if (st) return st;
  if (batch > 0 && (!f || !ftm)) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  HpTables tb;
  if ((st = healpix_tables(nside, L, &tb))) return st;
  RingFwdArgs r{};
  r.f = f;
  r.f_complex = 1;
  r.in_single = double_precision ? 0 : 1;
  r.out_single = double_precision ? 0 : 1;
  r.norm_mode = norm_mode(norm, true);
  r.real_only_2L = reality ? 1 : 0;
  r.map_stride = 12LL * nside * nside;
  r.out = static_cast<cplx*>(ftm);
  r.out_map_stride = (int64_t)(4 * nside - 1) * 2 * L;
  r.ncol = 2 * L;
  r.m_first = -L;
  r.zero_neg_2L = reality ? 1 : 0;
  r.L = L;
  r.chirp_pool = tb.chirp;
  return hip_status(
      launch_ring_fft_forward(r, tb.set, batch, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_healpix_ifft_ex(int nside, int L, int reality, int norm, int double_precision,
                                   const void* ftm, void* f, int batch, s2h_stream stream) {
  int st = healpix_fft_checks(nside, L, norm, batch);
  if (st) return st;
  if (batch > 0 && (!f || !ftm)) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  HpTables tb;
  if ((st = healpix_tables(nside, L, &tb))) return st;
  RingInvArgs r{};
  r.in = static_cast<const cplx*>(ftm);
  r.in_single = double_precision ? 0 : 1;
  r.f_single = double_precision ? 0 : 1;
  r.norm_mode = norm_mode(norm, false);
  r.in_map_stride = (int64_t)(4 * nside - 1) * 2 * L;
  r.ncol = 2 * L;
  r.m_first = -L;
  r.m_lo = -L;
  r.m_hi = L - 1;
  r.herm_mode = reality ? 2 : 0;
  r.f = f;
  r.f_complex = 1;
  r.map_stride = 12LL * nside * nside;
  r.chirp_pool = tb.chirp;
  r.L = L;
  return hip_status(
      launch_ring_fft_inverse(r, tb.set, batch, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_healpix_fft(int nside, int L, int reality, const void* f, void* ftm, int batch,
                               s2h_stream stream) {
  return s2h_healpix_fft_ex(nside, L, reality, S2H_FFT_NORM_BACKWARD, 1, f, ftm, batch, stream);
}

extern "C" int s2h_healpix_ifft(int nside, int L, int reality, const void* ftm, void* f,
                                int batch, s2h_stream stream) {
  return s2h_healpix_ifft_ex(nside, L, reality, S2H_FFT_NORM_FORWARD, 1, ftm, f, batch, stream);
}

// (forward, normalize) -> norm as build_healpix_fft_descriptor does
// (lib/src/extensions.cc:115-149); shift is always on; the tables are built
// here, at descriptor time, so that the launches never allocate.
extern "C" int s2h_build_healpix_fft_descriptor(int nside, int L, int reality, int forward,
                                                int normalize, int double_precision,
                                                s2h_healpix_fft_descriptor* out) {
  if (!out) return S2H_ERR_INVALID;
  s2h_healpix_fft_descriptor d{};
  d.nside = nside;
  d.L = L;
  d.reality = reality ? 1 : 0;
  d.forward = forward ? 1 : 0;
  if (forward)
    d.norm = normalize ? S2H_FFT_NORM_FORWARD : S2H_FFT_NORM_BACKWARD;
  else
    d.norm = normalize ? S2H_FFT_NORM_BACKWARD : S2H_FFT_NORM_FORWARD;
  d.shift = 1;
  d.double_precision = double_precision ? 1 : 0;
  d.batch = 1;
  const int st = s2h_healpix_fft_prepare(nside, L);
  if (st) return st;
  *out = d;
  return S2H_OK;
}

extern "C" void healpix_fft_hip(void* stream, void** buffers, const char* opaque,
                                size_t opaque_len) {
  if (!buffers || !opaque || opaque_len != sizeof(s2h_healpix_fft_descriptor)) {
    g_last_xla_status = S2H_ERR_INVALID;
    return;
  }
  s2h_healpix_fft_descriptor d;
  std::memcpy(&d, opaque, sizeof(d));
  if (d.shift != 1 || d.batch < 0 || (d.double_precision != 0 && d.double_precision != 1)) {
    // the reference always shifts (extensions.cc:131-132)
    g_last_xla_status = d.shift != 1 ? S2H_ERR_UNSUPPORTED : S2H_ERR_INVALID;
    return;
  }
  const int batch = d.batch > 0 ? d.batch : 1;
  g_last_xla_status =
      d.forward ? s2h_healpix_fft_ex(d.nside, d.L, d.reality, d.norm, d.double_precision,
                                     buffers[0], buffers[1], batch, stream)
                : s2h_healpix_ifft_ex(d.nside, d.L, d.reality, d.norm, d.double_precision,
                                      buffers[0], buffers[1], batch, stream);
}

extern "C" int healpix_fft_hip_last_status(void) { return g_last_xla_status; }

extern "C" int s2h_gamma_dft(int N, int64_t npix, int mode, double scale, const void* in,
                             void* out, s2h_stream stream) {
  if (N < 1 || N > 8 || npix < 0 || mode < 0 || mode > 5 || !in || !out || in == out)
    return S2H_ERR_INVALID;
  return hip_status(
      launch_gamma_dft(N, npix, mode, scale, in, out, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_wigner_reality_fill(void* flmn, int L, int N, s2h_stream stream) {
  if (!flmn || L < 1 || N < 1) return S2H_ERR_INVALID;
  return hip_status(launch_wigner_fill(static_cast<cplx*>(flmn), L, N,
                                       static_cast<hipStream_t>(stream)));
}

extern "C" size_t s2h_batch_sumsq_workspace_bytes(int batch) {
  return batch > 0 ? (size_t)kSumsqBlocks * (size_t)batch * sizeof(double) : 0;
}

extern "C" int s2h_batch_sumsq(const double* x, int64_t n, int64_t stride, int batch, double* out,
                               void* work, s2h_stream stream) {
  if (batch < 0 || n < 0 || (batch > 0 && (!out || !work || (n > 0 && !x) || stride < n)))
    return S2H_ERR_INVALID;
  return hip_status(launch_batch_sumsq(x, n, stride, batch, static_cast<double*>(work), out,
                                       static_cast<hipStream_t>(stream)));
}

extern "C" size_t s2h_flm_dots_workspace_bytes(int nv, int batch) {
  if (nv < 1 || nv > kFlmDotMax || batch <= 0) return 0;
  return (size_t)nv * (size_t)batch * kFlmDotBlocks * sizeof(double);
}

extern "C" int s2h_flm_dots(const void* z, const void* const* v, int nv, int64_t n, int packed_L,
                            int batch, double* out, void* work, s2h_stream stream) {
  if (batch < 0 || n < 0 || nv < 1 || nv > kFlmDotMax || packed_L < 0) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  if (!out || !work || !v || (n > 0 && !z)) return S2H_ERR_INVALID;
  FlmDotArgs a{};
  a.z = static_cast<const cplx*>(z);
  for (int i = 0; i < nv; ++i) {
    if (n > 0 && !v[i]) return S2H_ERR_INVALID;
    a.v[i] = static_cast<const cplx*>(v[i]);
  }
  a.nv = nv;
  a.n = n;
  a.stride = n;
  a.packed_L = packed_L;
  a.partial = static_cast<double*>(work);
  a.out = out;
  return hip_status(launch_flm_dots(a, batch, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_cg_update(void* x, const void* p, const double* alpha, void* p_new,
                             const void* z, const double* g1, const void* p_prev,
                             const double* g2, int64_t n, int batch, s2h_stream stream) {
  if (batch < 0 || n < 0) return S2H_ERR_INVALID;
  if (batch == 0 || n == 0) return S2H_OK;
  if (!p || (!x && !p_new) || (x && !alpha) || (p_new && (!z || !g1)) ||
      (p_prev && (!g2 || !p_new)) || (p_new && (p_new == p || p_new == x)))
    return S2H_ERR_INVALID;
  CgUpdateArgs a{};
  a.x = static_cast<cplx*>(x);
  a.p = static_cast<const cplx*>(p);
  a.alpha = alpha;
  a.p_new = static_cast<cplx*>(p_new);
  a.z = static_cast<const cplx*>(z);
  a.g1 = g1;
  a.p_prev = static_cast<const cplx*>(p_prev);
  a.g2 = g2;
  a.n = n;
  a.stride = n;
  return hip_status(launch_cg_update(a, batch, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_cg_finish(void* x, const void* const* p, const double* const* alpha, int nd,
                             int64_t n, int L, void* flm_out, int batch, s2h_stream stream) {
  if (batch < 0 || n < 0 || nd < 0 || nd > kCgFinishMax) return S2H_ERR_INVALID;
  if (flm_out && (L < 1 || n != (int64_t)L * (L + 1) / 2)) return S2H_ERR_INVALID;
  if (batch == 0 || n == 0) return S2H_OK;
  if (!x || (nd > 0 && (!p || !alpha)) || flm_out == x) return S2H_ERR_INVALID;
  CgFinishArgs a{};
  a.x = static_cast<cplx*>(x);
  for (int j = 0; j < nd; ++j) {
    if (!p[j] || !alpha[j]) return S2H_ERR_INVALID;
    a.p[j] = static_cast<const cplx*>(p[j]);
    a.alpha[j] = alpha[j];
  }
  a.nd = nd;
  a.n = n;
  a.L = L;
  a.dst = static_cast<cplx*>(flm_out);
  return hip_status(launch_cg_finish(a, batch, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_probe_mfma_f64(const double* a, const double* b, double* out,
                                  s2h_stream stream) {
  if (!a || !b || !out) return S2H_ERR_INVALID;
  return hip_status(launch_mfma_f64_probe(a, b, out, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_probe_mfma_f64_rate(int blocks, int iters, double* out, s2h_stream stream) {
  if (blocks < 1 || iters < 1 || !out) return S2H_ERR_INVALID;
  return hip_status(launch_mfma_f64_rate(blocks, iters, out, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_probe_mfma_f32(const float* a, const float* b, float* out, s2h_stream stream) {
  if (!a || !b || !out) return S2H_ERR_INVALID;
  return hip_status(launch_mfma_f32_probe(a, b, out, static_cast<hipStream_t>(stream)));
}

// ---------------------------------------------------------------- Wigner
// SO(3) transforms with the n-planes as one batch per stage (the reference
// loops, or vmaps, a spin transform over n: s2fft/transforms/wigner.py:110-196,
// 442-533, 536-645).  Planes are indexed k = N - 1 + n.
namespace {

int64_t wig_flm_stride(const s2h_wigner_plan* w) { return (int64_t)w->d.L * (2 * w->d.L - 1); }

// The planes k0 .. k0 + count - 1 as one multi-plane batch (count >= 2), or
// a plain transform of plane k0's own plan (count == 1).
const s2h_plan* plane_plan(const s2h_wigner_plan* w, int k0, int count, PlaneSet* ps) {
  if (count == 1) return w->plans[k0];
  ps->fwd = w->fwd_planes + k0;
  ps->inv = w->inv_planes + k0;
  ps->spin0 = (w->d.N - 1) - k0;  // spin of plane k is -(k - (N - 1))
  ps->spin_step = -1;
  return w->geom;
}

// Workspace layout: [fban: G npix complex][flmn copy: G L (2L-1) complex]
// [one real plane: npix doubles, padded to complex][spin work]
struct WigWork {
  cplx* fban;
  cplx* gcopy;
  double* real_plane;
  void* spin;
};

WigWork wig_work(const s2h_wigner_plan* w, void* work) {
  WigWork ww;
  ww.fban = static_cast<cplx*>(work);
  ww.gcopy = ww.fban + (int64_t)w->G * w->npix;
  ww.real_plane = reinterpret_cast<double*>(ww.gcopy + (int64_t)w->G * wig_flm_stride(w));
  ww.spin = reinterpret_cast<cplx*>(ww.real_plane) + (w->npix + 1) / 2;
  return ww;
}

int wig_check(const s2h_wigner_plan* w, const void* a, const void* b, const void* work) {
  if (!w || !a || !b || !work) return S2H_ERR_INVALID;
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess) return S2H_ERR_HIP;
  return dev == w->device ? S2H_OK : S2H_ERR_INVALID;
}

}  // namespace

extern "C" size_t s2h_wigner_workspace_bytes(const s2h_wigner_plan* w) {
  if (!w) return 0;
  const s2h_plan* p = w->plans[w->d.N - 1];
  return ((size_t)w->G * ((size_t)w->npix + (size_t)wig_flm_stride(w)) +
          (size_t)(w->npix + 1) / 2) * sizeof(cplx) +
         s2h_workspace_bytes(p, w->G);
}

extern "C" int s2h_wigner_forward(const s2h_wigner_plan* w, const void* f, int f_is_complex,
                                  void* flmn, void* work, s2h_stream stream) {
  int st = wig_check(w, f, flmn, work);
  if (st) return st;
  if (w->d.reality && f_is_complex) return S2H_ERR_INVALID;  // a real signal: float64 f
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int N = w->d.N, G = w->G;
  const WigWork ww = wig_work(w, work);
  const int64_t fs = wig_flm_stride(w);
  cplx* out = static_cast<cplx*>(flmn);
  // fban[k] = (2 pi / G) sum_g f[g] e^{-2 pi i n g / G} (wigner.py:495-502)
  if ((st = hip_status(launch_gamma_dft(N, w->npix, f_is_complex ? 0 : 2, 2.0 * M_PI / G, f,
                                        ww.fban, s))))
    return st;
  PlaneSet ps;
  if (!w->d.reality) {
    const s2h_plan* p = plane_plan(w, 0, G, &ps);
    return forward_impl(p, G > 1 ? &ps : nullptr, ww.fban, 1, out, ww.spin, G, S2H_PREC_FP64,
                        false, s);
  }
  // reality: the n = 0 plane as a real transform, the n >= 1 planes as one
  // batch, the n < 0 planes by symmetry (wigner.py:506-526)
  if ((st = forward_impl(w->real0, nullptr, ww.fban + (int64_t)(N - 1) * w->npix, 1,
                         out + (N - 1) * fs, ww.spin, 1, S2H_PREC_FP64, false, s)))
    return st;
  if (N > 1) {
    const s2h_plan* p = plane_plan(w, N, N - 1, &ps);
    if ((st = forward_impl(p, N - 1 > 1 ? &ps : nullptr, ww.fban + (int64_t)N * w->npix, 1,
                           out + (int64_t)N * fs, ww.spin, N - 1, S2H_PREC_FP64, false, s)))
      return st;
  }
  return hip_status(launch_wigner_fill(out, w->d.L, N, s));
}

extern "C" int s2h_wigner_inverse(const s2h_wigner_plan* w, const void* flmn, void* f,
                                  void* work, s2h_stream stream) {
  int st = wig_check(w, flmn, f, work);
  if (st) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int N = w->d.N, G = w->G;
  const WigWork ww = wig_work(w, work);
  const int64_t fs = wig_flm_stride(w);
  const cplx* in = static_cast<const cplx*>(flmn);
  PlaneSet ps;
  if (!w->d.reality) {
    const s2h_plan* p = plane_plan(w, 0, G, &ps);
    if ((st = inverse_impl(p, G > 1 ? &ps : nullptr, in, ww.fban, ww.spin, G, S2H_PREC_FP64,
                           false, nullptr, s)))
      return st;
    return hip_status(launch_gamma_dft(N, w->npix, 1, 1.0, ww.fban, f, s));
  }
  // a real signal uses the planes n >= 0 (wigner.py:179-195): n = 0 as a real
  // transform (its m >= 0 coefficients, Hermitian completion), the n >= 1
  // planes as one batch
  if ((st = inverse_impl(w->real0, nullptr, in + (N - 1) * fs, ww.real_plane, ww.spin, 1,
                         S2H_PREC_FP64, true, nullptr, s)))
    return st;
  if ((st = hip_status(launch_real_to_complex(ww.real_plane, ww.fban, w->npix, s)))) return st;
  if (N > 1) {
    const s2h_plan* p = plane_plan(w, N, N - 1, &ps);
    if ((st = inverse_impl(p, N - 1 > 1 ? &ps : nullptr, in + N * fs, ww.fban + w->npix, ww.spin,
                           N - 1, S2H_PREC_FP64, false, nullptr, s)))
      return st;
  }
  return hip_status(launch_gamma_dft(N, w->npix, 3, 1.0, ww.fban, f, s));
}

// grad_f = F^H g: the planes' spin adjoints, then the adjoint of the gamma DFT.
extern "C" int s2h_wigner_forward_adjoint(const s2h_wigner_plan* w, const void* g, void* grad_f,
                                          void* work, s2h_stream stream) {
  int st = wig_check(w, g, grad_f, work);
  if (st) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int N = w->d.N, G = w->G;
  const WigWork ww = wig_work(w, work);
  const int64_t fs = wig_flm_stride(w);
  PlaneSet ps;
  if (!w->d.reality) {
    const s2h_plan* p = plane_plan(w, 0, G, &ps);
    if ((st = forward_adjoint_impl(p, G > 1 ? &ps : nullptr, g, ww.fban, ww.spin, G, s)))
      return st;
    return hip_status(launch_gamma_dft(N, w->npix, 1, 2.0 * M_PI / G, ww.fban, grad_f, s));
  }
  // reality: the n < 0 planes' cotangents fold back onto n > 0 (adjoint of the
  // symmetric fill), the n = 0 plane goes through the real plan's adjoint
  if (hipMemcpyAsync(ww.gcopy, g, (size_t)G * fs * sizeof(cplx), hipMemcpyDeviceToDevice, s) !=
      hipSuccess)
    return S2H_ERR_HIP;
  if ((st = hip_status(launch_wigner_fill_adjoint(ww.gcopy, w->d.L, N, s)))) return st;
  // the real plan's adjoint writes float64, widened into plane n = 0 of fban
  if ((st = forward_adjoint_impl(w->real0, nullptr, ww.gcopy + (N - 1) * fs, ww.real_plane,
                                 ww.spin, 1, s)))
    return st;
  if ((st = hip_status(launch_real_to_complex(ww.real_plane, ww.fban, w->npix, s)))) return st;
  if (N > 1) {
    const s2h_plan* p = plane_plan(w, N, N - 1, &ps);
    if ((st = forward_adjoint_impl(p, N - 1 > 1 ? &ps : nullptr, ww.gcopy + (int64_t)N * fs,
                                   ww.fban + w->npix, ww.spin, N - 1, s)))
      return st;
  }
  return hip_status(launch_gamma_dft(N, w->npix, 4, 2.0 * M_PI / G, ww.fban, grad_f, s));
}

// grad_flmn = I^H g: the adjoint of the gamma DFT, then the planes' spin
// adjoints (zero for the n < 0 planes of a real signal, which the inverse
// does not read; its n = 0 plane through the real plan's adjoint).
extern "C" int s2h_wigner_inverse_adjoint(const s2h_wigner_plan* w, const void* g,
                                          int g_is_complex, void* grad_flmn, void* work,
                                          s2h_stream stream) {
  int st = wig_check(w, g, grad_flmn, work);
  if (st) return st;
  if (w->d.reality && g_is_complex) return S2H_ERR_INVALID;  // real output: real cotangent
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int N = w->d.N, G = w->G;
  const WigWork ww = wig_work(w, work);
  const int64_t fs = wig_flm_stride(w);
  cplx* out = static_cast<cplx*>(grad_flmn);
  const int mode = w->d.reality ? 5 : (g_is_complex ? 0 : 2);
  if ((st = hip_status(launch_gamma_dft(N, w->npix, mode, 1.0, g, ww.fban, s)))) return st;
  PlaneSet ps;
  if (!w->d.reality) {
    const s2h_plan* p = plane_plan(w, 0, G, &ps);
    return inverse_adjoint_impl(p, G > 1 ? &ps : nullptr, ww.fban, 1, out, ww.spin, G, s);
  }
  if (N > 1 && (st = hip_status(launch_zero(out, 2 * (int64_t)(N - 1) * fs, s)))) return st;
  // plane n = 0 (row 0 of fban, real-valued): the real plan reads its real part
  if ((st = inverse_adjoint_impl(w->real0, nullptr, ww.fban, 1, out + (N - 1) * fs, ww.spin, 1,
                                 s)))
    return st;
  if (N == 1) return S2H_OK;
  const s2h_plan* p = plane_plan(w, N, N - 1, &ps);
  return inverse_adjoint_impl(p, N - 1 > 1 ? &ps : nullptr, ww.fban + w->npix, 1, out + N * fs,
                              ww.spin, N - 1, s);
}
// ---------------------------------------------------------------- d-tables
// compute_all_slices (s2fft/recursions/price_mcewen.py:377-643) and the
// precompute path (s2fft/precompute_transforms/construct.py:16-161,
// spherical.py:87-149, 314-380) on the device.
extern "C" int s2h_compute_all_slices(const s2h_plan* p, const double* beta, int ntheta,
                                      double* dl, s2h_stream stream) {
  if (!p || ntheta < 0 || (ntheta > 0 && (!beta || !dl))) return S2H_ERR_INVALID;
  if (p->d.norm != S2H_NORM_NONE || p->d.reality) return S2H_ERR_INVALID;
  if (ntheta == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  DTableArgs a{};
  a.mstart = p->inv_mstart;
  a.coef = p->inv_coef;
  a.theta = beta;
  a.T = ntheta;
  a.L = p->d.L;
  a.m_lo = p->m_lo;
  a.nm = p->nm;
  a.scale = p->sign;  // undo the plan's global sign: bare d
  a.out = dl;
  a.sm = (int64_t)ntheta * p->d.L;
  a.st = p->d.L;
  a.sl = 1;
  return hip_status(launch_dtable(a, static_cast<hipStream_t>(stream)));
}

extern "C" size_t s2h_precompute_kernel_bytes(const s2h_plan* p, int forward) {
  if (!p || p->d.norm == S2H_NORM_NONE) return 0;
  const size_t T = (size_t)(forward ? p->fwd_rows : p->inv_rows);
  const size_t el = p->d.sampling == S2H_HEALPIX ? sizeof(cplx) : sizeof(double);
  return T * (size_t)p->d.L * (size_t)p->nm * el;
}

extern "C" int s2h_precompute_kernel(const s2h_plan* p, int forward, double scale, void* kernel,
                                     s2h_stream stream) {
  if (!p || !kernel || p->d.norm == S2H_NORM_NONE) return S2H_ERR_INVALID;
  int st = check_device(p);
  if (st) return st;
  const bool hp = p->d.sampling == S2H_HEALPIX;
  DTableArgs a{};
  a.mstart = forward ? p->fwd_mstart : p->inv_mstart;
  a.coef = forward ? p->fwd_coef : p->inv_coef;
  a.theta = forward ? p->fwd_theta : p->inv_theta;
  a.T = forward ? p->fwd_rows : p->inv_rows;
  a.L = p->d.L;
  a.m_lo = p->m_lo;
  a.nm = p->nm;
  a.wt = forward ? p->fwd_row_w : nullptr;
  a.phi0 = hp ? p->row_phi0 : nullptr;
  a.phase_sign = forward ? -1.0 : 1.0;
  a.scale = scale * p->sign;  // the plan's tables carry its global sign
  a.cplx = hp ? 1 : 0;
  a.out = kernel;
  a.sm = 1;
  a.st = (int64_t)p->d.L * p->nm;
  a.sl = p->nm;
  return hip_status(launch_dtable(a, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_precompute_forward(const s2h_plan* p, const void* kernel, int kernel_is_complex,
                                      double scale, const void* f, int f_is_complex, void* flm,
                                      void* work, int batch, s2h_stream stream) {
  if (!p || !kernel || !f || !flm || batch < 0 || (batch > 0 && !work)) return S2H_ERR_INVALID;
  if (p->d.norm == S2H_NORM_NONE) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  cplx* ftm = static_cast<cplx*>(work);
  const bool resample = p->fwd_kind != 0;
  cplx* spec = ftm;
  if (resample) ftm = spec + (int64_t)batch * p->fwd_in_rows * p->Mc;
  if ((st = ring_analysis_stage(p, f, f_is_complex, resample ? spec : ftm, batch, s, true)))
    return st;
  if (resample) {
    ResampleArgs ra = resample_args(p, spec, ftm, nullptr);
    ra.weights = nullptr;  // the kernel carries the quadrature weights
    if ((st = hip_status(launch_resample(ra, batch, s)))) return st;
  }
  PcArgs a{};
  a.kernel = kernel;
  a.kernel_complex = kernel_is_complex ? 1 : 0;
  a.T = p->fwd_rows;
  a.L = p->d.L;
  a.m_lo = p->m_lo;
  a.nm = p->nm;
  a.y = ftm;
  a.y_map_stride = (int64_t)p->fwd_rows * p->Mc;
  a.flm = static_cast<cplx*>(flm);
  a.flm_map_stride = flm_stride(p);
  a.scale = scale;
  a.batch = batch;
  if ((st = hip_status(launch_pc_forward(a, s)))) return st;
  if (p->d.reality)
    return hip_status(launch_reality_fill(static_cast<cplx*>(flm), flm_stride(p), p->d.L, batch, s));
  return S2H_OK;
}

extern "C" int s2h_precompute_inverse(const s2h_plan* p, const void* kernel, int kernel_is_complex,
                                      double scale, const void* flm, void* f, void* work,
                                      int batch, s2h_stream stream) {
  if (!p || !kernel || !f || !flm || batch < 0 || (batch > 0 && !work)) return S2H_ERR_INVALID;
  if (p->d.norm == S2H_NORM_NONE) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  int st = check_device(p);
  if (st) return st;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const int L = p->d.L;
  cplx* ftm = static_cast<cplx*>(work);
  PcArgs a{};
  a.kernel = kernel;
  a.kernel_complex = kernel_is_complex ? 1 : 0;
  a.T = p->inv_rows;
  a.L = L;
  a.m_lo = p->m_lo;
  a.nm = p->nm;
  a.y = ftm;
  a.y_map_stride = (int64_t)p->inv_rows * p->Mc;
  a.flm = const_cast<cplx*>(static_cast<const cplx*>(flm));
  a.flm_map_stride = flm_stride(p);
  a.scale = scale;
  a.batch = batch;
  if ((st = hip_status(launch_pc_inverse(a, s)))) return st;
  RingInvArgs r{};
  r.in = ftm;
  r.in_map_stride = a.y_map_stride;
  r.ncol = p->Mc;
  r.m_first = p->m_lo;
  r.m_lo = -(L - 1);
  r.m_hi = L - 1;
  r.herm_mode = p->d.reality ? 1 : 0;
  r.apply_phase = 0;  // the kernel carries the ring phases
  r.apply_weight = 0;
  r.f = f;
  r.f_complex = p->inv_out_complex ? 1 : 0;
  r.map_stride = p->npix;
  r.chirp_pool = p->chirp;
  r.L = L;
  return hip_status(launch_ring_fft_inverse(r, p->inv_set, batch, s));
}
