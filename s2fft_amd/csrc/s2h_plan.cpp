// Plan construction: geometry, quadrature, recursion tables, FFT job tables.
//
// Geometry and weights follow s2fft/sampling/s2_samples.py:203-308 (thetas),
// s2fft/utils/quadrature.py:99-288 (weights) and
// s2fft/utils/healpix_ffts.py:429-452 (first-pixel azimuths).
#include "s2h_plan.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>

namespace s2h {

namespace {

constexpr long double kPi = 3.141592653589793238462643383279502884L;

template <typename T>
int upload(s2h_plan* p, const std::vector<T>& v, T** out) {
  *out = nullptr;
  if (v.empty()) return S2H_OK;
  void* ptr = nullptr;
  if (hipMalloc(&ptr, sizeof(T) * v.size()) != hipSuccess) return S2H_ERR_NOMEM;
  if (p) p->allocations.push_back(ptr);
  if (hipMemcpy(ptr, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) != hipSuccess)
    return S2H_ERR_HIP;
  *out = static_cast<T*>(ptr);
  return S2H_OK;
}

// ---------------------------------------------------------------- geometry
double healpix_z(int t, int nside) {
  const double ns = nside;
  if (t < nside - 1) return 1.0 - (t + 1.0) * (t + 1.0) / (3.0 * ns * ns);
  if (t <= 3 * nside - 1) return 4.0 / 3.0 - 2.0 * (t + 1.0) / (3.0 * ns);
  const double u = 4.0 * ns - 1.0 - t;
  return u * u / (3.0 * ns * ns) - 1.0;
}

int healpix_nphi(int t, int nside) {
  if (t < nside - 1) return 4 * (t + 1);
  if (t <= 3 * nside - 1) return 4 * nside;
  return 4 * (4 * nside - t - 1);
}

double healpix_phi0_over_pi(int t, int nside) {
  if (t + 1 < nside) return 0.25 / (t + 1);
  if (t + 1 <= 3 * nside) return 0.25 * ((t - nside + 2) % 2) / nside;
  return 0.25 / (4 * nside - t - 1);
}

// Gauss-Legendre nodes (cos theta, descending) and weights, Newton on P_L.
void gauss_legendre(int L, std::vector<long double>& z, std::vector<long double>& w) {
  z.assign(L, 0.0L);
  w.assign(L, 0.0L);
  const int half = (L + 1) / 2;
  for (int i = 1; i <= half; ++i) {
    long double x = std::cos(kPi * (i - 0.25L) / (L + 0.5L));
    long double pp = 1.0L;
    for (int it = 0; it < 100; ++it) {
      long double p1 = 1.0L, p2 = 0.0L;
      for (int j = 1; j <= L; ++j) {
        const long double p3 = p2;
        p2 = p1;
        p1 = ((2.0L * j - 1.0L) * x * p2 - (j - 1.0L) * p3) / j;
      }
      pp = L * (x * p1 - p2) / (x * x - 1.0L);
      const long double dx = p1 / pp;
      x -= dx;
      if (std::fabs(dx) < 1e-19L) break;
    }
    {
      long double p1 = 1.0L, p2 = 0.0L;
      for (int j = 1; j <= L; ++j) {
        const long double p3 = p2;
        p2 = p1;
        p1 = ((2.0L * j - 1.0L) * x * p2 - (j - 1.0L) * p3) / j;
      }
      pp = L * (x * p1 - p2) / (x * x - 1.0L);
    }
    z[i - 1] = x;
    z[L - i] = -x;
    w[i - 1] = 2.0L / ((1.0L - x * x) * pp * pp);
    w[L - i] = w[i - 1];
  }
}

std::vector<double> thetas(int sampling, int L, int nside) {
  std::vector<double> th;
  switch (sampling) {
    case S2H_MW:
      for (int t = 0; t < L; ++t) th.push_back((2.0 * t + 1.0) * M_PI / (2.0 * L - 1.0));
      break;
    case S2H_MWSS:
      for (int t = 0; t <= L; ++t) th.push_back(2.0 * t * M_PI / (2.0 * L));
      break;
    case S2H_DH:
      for (int t = 0; t < 2 * L; ++t) th.push_back((2.0 * t + 1.0) * M_PI / (4.0 * L));
      break;
    case S2H_GL: {
      std::vector<long double> z, w;
      gauss_legendre(L, z, w);
      for (int t = 0; t < L; ++t) th.push_back((double)std::acos((double)z[t]));
      break;
    }
    case S2H_HEALPIX:
      for (int t = 0; t < 4 * nside - 1; ++t) th.push_back(std::acos(healpix_z(t, nside)));
      break;
  }
  return th;
}

// Weights the forward transform applies per ring (quadrature.py:7-50).
std::vector<double> forward_weights(int sampling, int L, int nside) {
  std::vector<double> w;
  if (sampling == S2H_HEALPIX) {
    w.assign(4 * nside - 1, (double)(4.0L * kPi / (12.0L * nside * nside)));
  } else if (sampling == S2H_GL) {
    std::vector<long double> z, ww;
    gauss_legendre(L, z, ww);
    for (int t = 0; t < L; ++t) w.push_back((double)(ww[t] * 2.0L * kPi / (2.0L * L - 1.0L)));
  } else if (sampling == S2H_DH) {
    for (int t = 0; t < 2 * L; ++t) {
      const long double th = (2.0L * t + 1.0L) * kPi / (4.0L * L);
      long double s = 0.0L;
      for (int k = 0; k < L; ++k) s += std::sin((2.0L * k + 1.0L) * th) / (2.0L * k + 1.0L);
      s *= 2.0L / L * std::sin(th);
      w.push_back((double)(s * 2.0L * kPi / (2.0L * L - 1.0L)));
    }
  } else {
    // MWSS at 2L (used after the resampling of MW / MWSS): quadrature.py:238-262
    const int Lp = 2 * L;
    std::vector<long double> wr(2 * Lp);
    for (int t = 0; t < 2 * Lp; ++t) {
      const long double a = kPi * t / Lp;
      long double s = 2.0L + kPi * std::sin(a);
      for (int p = 2; p <= Lp - 1; p += 2) s += 4.0L * std::cos(a * p) / (1.0L - (long double)p * p);
      wr[t] = s / (2.0L * Lp);
    }
    for (int t = 0; t <= Lp; ++t) {
      long double q = wr[t];
      if (t >= 1 && t <= Lp - 1) q += wr[2 * Lp - t];
      w.push_back((double)(q * 2.0L * kPi / (2.0L * L)));
    }
  }
  return w;
}

// ---------------------------------------------------------------- recursion
long double ell_norm(int kind, int l, bool inverse) {
  const long double n = std::sqrt((2.0L * l + 1.0L) / (4.0L * kPi));
  if (kind == S2H_NORM_NONE) return 1.0L;
  if (kind == S2H_NORM_SPIN) return n;
  // Wigner: wigner.py:637-644 (forward), 259-266 (inverse)
  if (!inverse) return 1.0L;
  return n * std::sqrt((2.0L * l + 1.0L) / (16.0L * kPi * kPi * kPi));
}

// sqrt(binomial(2j, j+k)) as (mantissa in [0.5,1), exponent).
void sqrt_binom(int j, int k, long double& mant, int& ex) {
  long double v = 1.0L;
  int e = 0;
  const int lo = j + k, cnt = j - k;
  for (int i = 1; i <= cnt; ++i) {
    v *= (long double)(lo + i) / (long double)i;
    int ee;
    v = std::frexp(v, &ee);
    e += ee;
  }
  // sqrt(v 2^e)
  if (e & 1) {
    v *= 2.0L;
    e -= 1;
  }
  v = std::sqrt(v);
  int ee;
  v = std::frexp(v, &ee);
  mant = v;
  ex = e / 2 + ee;
}

// Start value data for signed m; n = -spin.  See d^j formulas (Wigner's sum
// with one index at +/-j).
MStart make_mstart(int m, int n, int kind, bool inverse, double sign) {
  MStart st{};
  const int am = std::abs(m), an = std::abs(n);
  const int j = std::max(am, an);
  st.ell0 = j;
  int k, sgn = 1;
  if (am >= an) {
    k = n;
    if (m >= 0) {
      st.ecos = j + n;
      st.esin = j - n;
      sgn = ((j - n) & 1) ? -1 : 1;
    } else {
      st.ecos = j - n;
      st.esin = j + n;
    }
  } else {
    k = m;
    if (n > 0) {
      st.ecos = j + m;
      st.esin = j - m;
    } else {
      st.ecos = j - m;
      st.esin = j + m;
      sgn = ((j + m) & 1) ? -1 : 1;
    }
  }
  long double mant;
  int ex;
  sqrt_binom(j, std::abs(k), mant, ex);
  mant *= ell_norm(kind, j, inverse) * sgn * sign;
  int ee;
  const double mm = (double)std::frexp(mant, &ee);
  st.kmant = mm;
  st.kexp = ex + ee;
  return st;
}

// Normalised three-term coefficients for |m| (b for m = +|m|).
void make_coefs(int L, int n, int kind, bool inverse, std::vector<Coef>& out) {
  out.assign((size_t)L * L, Coef{0, 0, 0, 0});
  std::vector<long double> N(L + 1);
  for (int l = 0; l <= L; ++l) N[l] = ell_norm(kind, l, inverse);
  const long double nn = n;
  for (int am = 0; am < L; ++am) {
    const long double mm = am;
    const int l0 = std::max(am, std::abs(n));
    for (int l = l0; l < L; ++l) {
      Coef c{0, 0, 0, 0};
      if (l == 0) {
        c.a = (double)(N[1] / N[0]);
      } else {
        const long double L1 = l + 1.0L, Ll = l;
        const long double D =
            Ll * std::sqrt((L1 * L1 - mm * mm) * (L1 * L1 - nn * nn));
        const long double A = (2.0L * Ll + 1.0L) * Ll * L1 / D;
        const long double B = (2.0L * Ll + 1.0L) * mm * nn / D;
        const long double C =
            L1 * std::sqrt(std::max(0.0L, (Ll * Ll - mm * mm) * (Ll * Ll - nn * nn))) / D;
        c.a = (double)(A * N[l + 1] / N[l]);
        c.b = (double)(B * N[l + 1] / N[l]);
        c.c = (double)(C * N[l + 1] / N[l - 1]);
      }
      out[(size_t)am * L + l] = c;
    }
  }
}

// The same recursion with unit c: d_l = sigma_l e_l, where sigma_l0 =
// sigma_{l0+1} = 1 and sigma_{l+1} = c_l sigma_{l-1}, turns
//   d_{l+1} = (a_l x - b_l) d_l - c_l d_{l-1}
// into e_{l+1} = (a'_l x - b'_l) e_l - e_{l-1} with (a', b') = (a, b) sigma_l /
// sigma_{l+1}: one FP64 multiply less per step.  Entry l holds
// {a'_l, b'_l, sigma_l, 0}; a' and b' are zero from l = L - 1 on (the steps
// past the band-limit give zeros), sigma is kept up to l = L - 1.  c_l > 0
// for l > l0, so sigma never vanishes; it varies slowly with l (the c_l are
// near 1 away from l = |m|, |n|).
void make_unit_coefs(int L, int n, const std::vector<Coef>& c, std::vector<Coef>& out) {
  out.assign((size_t)L * L, Coef{0, 0, 0, 0});
  for (int am = 0; am < L; ++am) {
    const int l0 = std::max(am, std::abs(n));
    if (l0 >= L) continue;
    const Coef* cm = c.data() + (size_t)am * L;
    Coef* um = out.data() + (size_t)am * L;
    long double s_prev = 1.0L, s_cur = 1.0L;  // sigma_{l-1}, sigma_l at l = l0
    for (int l = l0; l < L; ++l) {
      const long double s_next = l == l0 ? 1.0L : (long double)cm[l].c * s_prev;
      Coef u{0, 0, (double)s_cur, 0};
      if (l + 1 < L) {
        u.a = (double)((long double)cm[l].a * s_cur / s_next);
        u.b = (double)((long double)cm[l].b * s_cur / s_next);
      }
      um[l] = u;
      s_prev = s_cur;
      s_cur = s_next;
    }
  }
}

// ---------------------------------------------------------------- FFT jobs
int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

struct ChirpBuilder {
  std::map<int, std::pair<int, int64_t>> table;  // n -> (M, offset)
  int64_t total = 0;
  int max_conv = 0;
  // Returns conv length (0 = mixed radix) and sets the chirp offset.
  int add(int n, int64_t* off) {
    *off = 0;
    if (largest_prime_factor(n) <= 13) return 0;
    auto it = table.find(n);
    if (it != table.end()) {
      *off = it->second.second;
      return it->second.first;
    }
    const int M = next_pow2(2 * n - 1);
    table[n] = {M, total};
    *off = total;
    total += M;
    max_conv = std::max(max_conv, M);
    return M;
  }
  // Allocates the chirp pool and fills it on the device (synchronously: plan
  // creation only).  On success the pool is recorded in allocs (freed with the
  // plan) or, with allocs null, owned by the caller; on failure nothing leaks.
  int build(cplx** pool, std::vector<void*>* allocs) {
    *pool = nullptr;
    if (table.empty()) return S2H_OK;
    std::vector<int> ns, Ms;
    std::vector<int64_t> offs;
    for (auto& kv : table) {
      ns.push_back(kv.first);
      Ms.push_back(kv.second.first);
      offs.push_back(kv.second.second);
    }
    // every temporary is freed on every path
    struct Tmp {
      void* p[4] = {nullptr, nullptr, nullptr, nullptr};
      ~Tmp() {
        for (void* q : p)
          if (q) (void)hipFree(q);
      }
    } t;
    const size_t cnt = ns.size();
    if (hipMalloc(&t.p[0], sizeof(cplx) * total) != hipSuccess) return S2H_ERR_NOMEM;
    if (hipMalloc(&t.p[1], sizeof(int) * cnt) != hipSuccess) return S2H_ERR_NOMEM;
    if (hipMalloc(&t.p[2], sizeof(int) * cnt) != hipSuccess) return S2H_ERR_NOMEM;
    if (hipMalloc(&t.p[3], sizeof(int64_t) * cnt) != hipSuccess) return S2H_ERR_NOMEM;
    if (hipMemcpy(t.p[1], ns.data(), sizeof(int) * cnt, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t.p[2], Ms.data(), sizeof(int) * cnt, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t.p[3], offs.data(), sizeof(int64_t) * cnt, hipMemcpyHostToDevice) != hipSuccess)
      return S2H_ERR_HIP;
    hipError_t e = launch_chirp_setup(static_cast<int*>(t.p[1]), static_cast<int*>(t.p[2]),
                                      static_cast<int64_t*>(t.p[3]), static_cast<cplx*>(t.p[0]),
                                      (int)cnt, max_conv, 0);
    if (e == hipSuccess) e = hipDeviceSynchronize();
    if (e != hipSuccess) return S2H_ERR_HIP;
    *pool = static_cast<cplx*>(t.p[0]);
    t.p[0] = nullptr;  // kept
    if (allocs) allocs->push_back(*pool);
    return S2H_OK;
  }
};

RingJob make_job(int64_t off, int nphi, int row, double phi0, double w, ChirpBuilder& cb) {
  RingJob j{};
  j.pix_offset = off;
  j.nphi = nphi;
  j.row = row;
  j.phi0_over_pi = phi0;
  j.weight = w;
  j.conv_len = cb.add(nphi, &j.chirp_off);
  return j;
}

// Orders the jobs (powers of two, other mixed-radix lengths, Bluestein;
// longest first within a class so that the heavy workgroups start early) and
// uploads them.
int ring_class(const RingJob& j) {
  if (j.conv_len) return 2;
  return (j.nphi & (j.nphi - 1)) == 0 ? 0 : 1;
}

int make_set(s2h_plan* p, std::vector<RingJob> jobs, JobSet* out) {
  std::stable_sort(jobs.begin(), jobs.end(), [](const RingJob& x, const RingJob& y) {
    const int cx = ring_class(x), cy = ring_class(y);
    if (cx != cy) return cx < cy;
    const int kx = cx == 2 ? x.conv_len : x.nphi, ky = cy == 2 ? y.conv_len : y.nphi;
    return kx > ky;
  });
  JobSet js;
  for (auto& j : jobs) {
    js.max_nphi = std::max(js.max_nphi, j.nphi);
    switch (ring_class(j)) {
      case 0:
        ++js.n_pow2;
        js.max_pow2 = std::max(js.max_pow2, j.nphi);
        break;
      case 1:
        ++js.n_smooth;
        js.max_smooth = std::max(js.max_smooth, j.nphi);
        break;
      default:
        ++js.n_blue;
        js.max_conv = std::max(js.max_conv, j.conv_len);
        if (j.conv_len > kBlueSplit)
          ++js.n_blue_long;
        else
          js.max_conv_short = std::max(js.max_conv_short, j.conv_len);
    }
  }
  const int st = upload<RingJob>(p, jobs, &js.jobs);
  if (st) return st;
  *out = js;
  return S2H_OK;
}

// Latitudinal ring list over rows [r0, r1] of a theta grid.
std::vector<LatRing> lat_rings(const std::vector<double>& th, int r0, int r1, bool paired) {
  std::vector<LatRing> out;
  auto mk = [&](int ta, int tb) {
    LatRing r{};
    r.ta = ta;
    r.tb = tb;
    r.x = std::cos(th[ta]);
    r.c2 = std::cos(th[ta] / 2.0);
    r.s2 = std::sin(th[ta] / 2.0);
    return r;
  };
  if (!paired) {
    for (int t = r0; t <= r1; ++t) out.push_back(mk(t, -1));
  } else {
    for (int t = r0, u = r1; t <= u; ++t, --u) out.push_back(mk(t, u));
  }
  return out;
}

}  // namespace

// Reference-layout HEALPix ring tables, cached per (device, nside, L).  Built
// by healpix_tables_prepare (allocates and synchronises once); the launch
// paths only look them up.
namespace {
std::mutex g_hp_mu;
std::map<std::tuple<int, int, int>, HpTables> g_hp_cache;
}  // namespace

int healpix_tables_prepare(int nside, int L) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return S2H_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_hp_mu);
  const auto key = std::make_tuple(dev, nside, L);
  if (g_hp_cache.count(key)) return S2H_OK;
  if (allow_large_lds() != hipSuccess) return S2H_ERR_HIP;
  ChirpBuilder cb;
  std::vector<RingJob> jobs;
  int64_t off = 0;
  for (int t = 0; t < 4 * nside - 1; ++t) {
    const int n = healpix_nphi(t, nside);
    jobs.push_back(make_job(off, n, t, healpix_phi0_over_pi(t, nside), 1.0, cb));
    off += n;
  }
  HpTables tb;
  std::vector<void*> own;  // the job table and the chirp pool (kept for the process)
  s2h_plan tmp;            // make_set records its upload here
  int st = make_set(&tmp, jobs, &tb.set);
  if (st == S2H_OK) st = cb.build(&tb.chirp, &tmp.allocations);
  if (st != S2H_OK) {
    for (void* q : tmp.allocations) (void)hipFree(q);
    return st;
  }
  g_hp_cache[key] = tb;
  return S2H_OK;
}

int healpix_tables(int nside, int L, HpTables* out) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return S2H_ERR_HIP;
  std::lock_guard<std::mutex> lock(g_hp_mu);
  const auto it = g_hp_cache.find(std::make_tuple(dev, nside, L));
  if (it == g_hp_cache.end()) return S2H_ERR_INVALID;  // not prepared
  *out = it->second;
  return S2H_OK;
}

}  // namespace s2h

using namespace s2h;

static int plan_build(s2h_plan* p) {
  const s2h_desc& d = p->d;
  const int L = d.L, nside = d.nside, spin = d.spin;
  const bool real = d.reality != 0;
  const bool hp = d.sampling == S2H_HEALPIX;
  const int n = -spin;
  if (allow_large_lds() != hipSuccess) return S2H_ERR_HIP;

  p->m_lo = real ? 0 : -(L - 1);
  p->nm = L - p->m_lo;
  p->Mc = p->nm;
  p->sign = (double)d.extra_sign * ((d.norm != S2H_NORM_NONE && (spin & 1)) ? -1.0 : 1.0);
  const bool pair_ok = spin == 0;

  ChirpBuilder cb;
  std::vector<double> thi = thetas(d.sampling, L, nside);

  // ---------------- inverse geometry
  std::vector<LatRing> inv_r;
  std::vector<RingJob> inv_j;
  switch (d.sampling) {
    case S2H_HEALPIX: {
      p->inv_rows = 4 * nside - 1;
      inv_r = lat_rings(thi, 0, p->inv_rows - 1, pair_ok);
      int64_t off = 0;
      for (int t = 0; t < p->inv_rows; ++t) {
        const int np = healpix_nphi(t, nside);
        inv_j.push_back(make_job(off, np, t, healpix_phi0_over_pi(t, nside), 1.0, cb));
        off += np;
      }
      p->npix = off;
      break;
    }
    case S2H_DH:
    case S2H_GL: {
      p->inv_rows = (int)thi.size();
      inv_r = lat_rings(thi, 0, p->inv_rows - 1, pair_ok);
      const int np = 2 * L - 1;
      for (int t = 0; t < p->inv_rows; ++t) inv_j.push_back(make_job((int64_t)t * np, np, t, 0.0, 1.0, cb));
      p->npix = (int64_t)p->inv_rows * np;
      break;
    }
    case S2H_MW: {
      p->inv_rows = L;
      inv_r = lat_rings(thi, 0, L - 2, false);  // theta_{L-1} = pi is the pole row
      p->inv_row_south = L - 1;
      const int np = 2 * L - 1;
      for (int t = 0; t < L; ++t) inv_j.push_back(make_job((int64_t)t * np, np, t, 0.0, 1.0, cb));
      p->npix = (int64_t)L * np;
      break;
    }
    case S2H_MWSS: {
      p->inv_rows = L + 1;
      inv_r = lat_rings(thi, 1, L - 1, pair_ok);
      p->inv_row_north = 0;
      p->inv_row_south = L;
      const int np = 2 * L;
      for (int t = 0; t <= L; ++t) inv_j.push_back(make_job((int64_t)t * np, np, t, 0.0, 1.0, cb));
      p->npix = (int64_t)(L + 1) * np;
      break;
    }
    default:
      return S2H_ERR_INVALID;
  }
  p->inv_paired = pair_ok && d.sampling != S2H_MW;
  p->inv_out_complex = hp || !real;

  // ---------------- forward geometry
  std::vector<LatRing> fwd_r;
  std::vector<RingJob> fwd_j;
  std::vector<double> wf;
  if (d.sampling == S2H_HEALPIX || d.sampling == S2H_DH || d.sampling == S2H_GL) {
    p->fwd_kind = 0;
    p->fwd_rows = p->inv_rows;
    fwd_r = inv_r;
    p->fwd_paired = p->inv_paired;
    wf = forward_weights(d.sampling, L, nside);
    fwd_j = inv_j;
    for (size_t t = 0; t < fwd_j.size(); ++t) fwd_j[t].weight = wf[t];
  } else {
    // resampled to MWSS at 2L: rows 0..2L, poles at 0 and 2L
    p->fwd_kind = d.sampling == S2H_MW ? 1 : 2;
    p->fwd_rows = 2 * L + 1;
    std::vector<double> thf;
    for (int t = 0; t <= 2 * L; ++t) thf.push_back(2.0 * t * M_PI / (4.0 * L));
    fwd_r = lat_rings(thf, 1, 2 * L - 1, pair_ok);
    p->fwd_paired = pair_ok;
    p->fwd_row_north = 0;
    p->fwd_row_south = 2 * L;
    wf = forward_weights(d.sampling, L, nside);
    p->fwd_in_rows = p->inv_rows;
    fwd_j = inv_j;  // phi-DFT of the input rows, weight 1
    const int lens[3] = {2 * L - 1, 2 * L, 4 * L};
    int need = 4 * L;
    for (int i = 0; i < 3; ++i) {
      p->rs_info[i] = cb.add(lens[i], &p->rs_chirp[i]);
      need = std::max(need, p->rs_info[i] ? p->rs_info[i] : lens[i]);
    }
    if (need > kMaxConvLen) return S2H_ERR_UNSUPPORTED;
    p->rs_info[3] = need;
    if (allow_large_lds_resample() != hipSuccess) return S2H_ERR_HIP;
    if (allow_large_lds_resample_adjoint() != hipSuccess) return S2H_ERR_HIP;
  }

  // ---------------- recursion tables
  std::vector<MStart> msf, msi;
  for (int m = p->m_lo; m < L; ++m) {
    msf.push_back(make_mstart(m, n, d.norm, false, p->sign));
    msi.push_back(make_mstart(m, n, d.norm, true, p->sign));
  }
  std::vector<Coef> cf, ci;
  make_coefs(L, n, d.norm, false, cf);
  const bool same = d.norm != S2H_NORM_WIGNER;
  if (!same) make_coefs(L, n, d.norm, true, ci);
  std::vector<double> nf(L), ni(L);
  for (int l = 0; l < L; ++l) {
    nf[l] = (double)(ell_norm(d.norm, l, false) * p->sign);
    ni[l] = (double)(ell_norm(d.norm, l, true) * p->sign);
  }

  int st;
  if ((st = upload(p, inv_r, &p->inv_rings))) return st;
  p->n_inv_rings = (int)inv_r.size();
  if ((st = upload(p, fwd_r, &p->fwd_rings))) return st;
  p->n_fwd_rings = (int)fwd_r.size();
  if ((st = make_set(p, inv_j, &p->inv_set))) return st;
  if ((st = make_set(p, fwd_j, &p->fwd_set))) return st;
  {
    // forward ring of each ring-spectra row, padded to whole row blocks
    p->spec_rows = (p->fwd_rows + kSpecBlock - 1) / kSpecBlock * kSpecBlock;
    std::vector<RingJob> rj(p->spec_rows, RingJob{});
    for (const RingJob& j : fwd_j)
      if (j.row >= 0 && j.row < p->spec_rows) rj[j.row] = j;
    if ((st = upload(p, rj, &p->fwd_rowjobs))) return st;
    std::vector<int> belt, polar;
    for (int blk = 0; blk < p->spec_rows / kSpecBlock; ++blk) {
      bool no_alias = true;
      for (int q = 0; q < kSpecBlock; ++q) {
        const int n = rj[blk * kSpecBlock + q].nphi;
        if (n > 0 && n < 2 * L - 1) no_alias = false;
      }
      (no_alias ? belt : polar).push_back(blk);
    }
    // the belt rows for the synthesis-fused part of the CG dot: one w n over
    // a contiguous run of belt blocks (else left to the alias kernels)
    p->spec_belt_wn = 0.0;
    if (!belt.empty() && belt.back() - belt.front() + 1 == (int)belt.size()) {
      const RingJob& j0 = rj[belt.front() * kSpecBlock];
      bool uniform = j0.nphi > 0;
      for (int blk : belt)
        for (int q = 0; q < kSpecBlock; ++q) {
          const RingJob& j = rj[blk * kSpecBlock + q];
          uniform = uniform && j.nphi == j0.nphi && j.weight == j0.weight;
        }
      if (uniform) {
        p->spec_belt_wn = j0.weight * j0.nphi;
        p->spec_belt_row_lo = belt.front() * kSpecBlock;
        p->spec_belt_row_hi = (belt.back() + 1) * kSpecBlock;
      }
    }
    p->n_belt_blocks = (int)belt.size();
    p->n_polar_blocks = (int)polar.size();
    if (!belt.empty() && (st = upload(p, belt, &p->spec_belt_blocks))) return st;
    if (!polar.empty() && (st = upload(p, polar, &p->spec_polar_blocks))) return st;
  }
  if ((st = upload(p, msf, &p->fwd_mstart))) return st;
  if ((st = upload(p, msi, &p->inv_mstart))) return st;
  if ((st = upload(p, cf, &p->fwd_coef))) return st;
  if (same) {
    p->inv_coef = p->fwd_coef;
  } else if ((st = upload(p, ci, &p->inv_coef))) {
    return st;
  }
  {
    std::vector<Coef> u;
    make_unit_coefs(L, n, cf, u);
    if ((st = upload(p, u, &p->fwd_ucoef))) return st;
    if (same) {
      p->inv_ucoef = p->fwd_ucoef;
    } else {
      make_unit_coefs(L, n, ci, u);
      if ((st = upload(p, u, &p->inv_ucoef))) return st;
    }
  }
  if ((st = upload(p, nf, &p->fwd_norm))) return st;
  if ((st = upload(p, ni, &p->inv_norm))) return st;
  if (!wf.empty() && p->fwd_kind != 0) {
    if ((st = upload(p, wf, &p->fwd_weights))) return st;
  }
  {
    std::vector<double> thf_rows;
    if (p->fwd_kind == 0) {
      thf_rows = thi;
    } else {
      for (int t = 0; t <= 2 * L; ++t) thf_rows.push_back(2.0 * t * M_PI / (4.0 * L));
    }
    if ((st = upload(p, thf_rows, &p->fwd_theta))) return st;
    if ((st = upload(p, thi, &p->inv_theta))) return st;
    if ((st = upload(p, wf, &p->fwd_row_w))) return st;
    if (hp) {
      std::vector<double> ph;
      for (int t = 0; t < p->inv_rows; ++t) ph.push_back(M_PI * healpix_phi0_over_pi(t, nside));
      if ((st = upload(p, ph, &p->row_phi0))) return st;
    }
  }
  if ((st = cb.build(&p->chirp, &p->allocations))) return st;
  return S2H_OK;
}

extern "C" int s2h_plan_create(s2h_plan** out, const s2h_desc* desc) {
  if (!out || !desc) return S2H_ERR_INVALID;
  *out = nullptr;
  const s2h_desc& d = *desc;
  if (d.sampling < S2H_MW || d.sampling > S2H_HEALPIX) return S2H_ERR_INVALID;
  if (d.L < 2) return S2H_ERR_INVALID;
  if (d.spin >= 8 || d.spin <= -8) return S2H_ERR_SPIN;
  if (std::abs(d.spin) >= d.L) return S2H_ERR_INVALID;
  if (d.L_lower < 0 || d.L_lower >= d.L) return S2H_ERR_INVALID;
  if (d.reality && d.spin != 0) return S2H_ERR_INVALID;
  if (d.norm < S2H_NORM_SPIN || d.norm > S2H_NORM_NONE) return S2H_ERR_INVALID;
  if (d.extra_sign != 1 && d.extra_sign != -1) return S2H_ERR_INVALID;
  if (d.sampling == S2H_HEALPIX) {
    if (d.nside < 1) return S2H_ERR_INVALID;
    if (d.L < 2 * d.nside) return S2H_ERR_BANDLIMIT;
    if (4 * d.nside > kMaxRingLen || d.L > kMaxL) return S2H_ERR_UNSUPPORTED;
  } else {
    if (2 * d.L > kMaxRingLen) return S2H_ERR_UNSUPPORTED;
  }
  s2h_plan* p = new (std::nothrow) s2h_plan();
  if (!p) return S2H_ERR_NOMEM;
  p->d = d;
  if (hipGetDevice(&p->device) != hipSuccess) {
    delete p;
    return S2H_ERR_HIP;
  }
  const int st = plan_build(p);
  if (st != S2H_OK) {
    s2h_plan_destroy(p);
    return st;
  }
  *out = p;
  return S2H_OK;
}

extern "C" int s2h_plan_destroy(s2h_plan* p) {
  if (!p) return S2H_OK;
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(p->device);
  for (void* ptr : p->allocations) (void)hipFree(ptr);
  (void)hipSetDevice(cur);
  delete p;
  return S2H_OK;
}

// ---------------------------------------------------------------- Wigner
// One spin plan per n (spin -n, Wigner normalisation, sign (-1)^n: the
// reference's per-n precomputes of price_mcewen.py:258-374 with the
// normalisation of wigner.py:179-195 folded in), and their recursion tables
// gathered into per-plane arrays for multi-plane launches.
extern "C" int s2h_wigner_plan_destroy(s2h_wigner_plan* w) {
  if (!w) return S2H_OK;
  for (s2h_plan* p : w->plans) s2h_plan_destroy(p);
  s2h_plan_destroy(w->real0);
  int cur = 0;
  (void)hipGetDevice(&cur);
  (void)hipSetDevice(w->device);
  for (void* ptr : w->allocations) (void)hipFree(ptr);
  (void)hipSetDevice(cur);
  delete w;
  return S2H_OK;
}

extern "C" int s2h_wigner_plan_create(s2h_wigner_plan** out, const s2h_wigner_desc* desc) {
  if (!out || !desc) return S2H_ERR_INVALID;
  *out = nullptr;
  const s2h_wigner_desc& d = *desc;
  if (d.N < 1) return S2H_ERR_INVALID;
  if (d.N >= 8) return S2H_ERR_SPIN;  // wigner.py:87-88, 419-420
  if (d.N - 1 >= d.L) return S2H_ERR_INVALID;
  s2h_wigner_plan* w = new (std::nothrow) s2h_wigner_plan();
  if (!w) return S2H_ERR_NOMEM;
  w->d = d;
  w->G = 2 * d.N - 1;
  if (hipGetDevice(&w->device) != hipSuccess) {
    delete w;
    return S2H_ERR_HIP;
  }
  int st = S2H_OK;
  for (int n = -(d.N - 1); n < d.N && st == S2H_OK; ++n) {
    s2h_desc sd{d.sampling, d.L, d.nside, -n, d.L_lower, 0, S2H_NORM_WIGNER, (n & 1) ? -1 : 1};
    s2h_plan* p = nullptr;
    st = s2h_plan_create(&p, &sd);
    if (st == S2H_OK) w->plans.push_back(p);
  }
  if (st == S2H_OK && d.reality) {
    s2h_desc sd{d.sampling, d.L, d.nside, 0, d.L_lower, 1, S2H_NORM_WIGNER, 1};
    st = s2h_plan_create(&w->real0, &sd);
  }
  if (st == S2H_OK) {
    w->npix = w->plans[d.N - 1]->npix;
    if (d.N > 1) w->geom = w->plans[d.N];  // n = 1: spin -1, unpaired rings
    std::vector<LatPlane> fp, ip;
    for (int k = 0; k < w->G; ++k) {
      const s2h_plan* p = w->plans[k];
      const int an = std::abs(k - (d.N - 1));
      fp.push_back(LatPlane{p->fwd_mstart, p->fwd_coef, std::max(an, d.L_lower), 0});
      ip.push_back(LatPlane{p->inv_mstart, p->inv_coef, d.L_lower, 0});
    }
    s2h_plan holder;  // upload() records allocations on a plan
    st = upload(&holder, fp, &w->fwd_planes);
    if (st == S2H_OK) st = upload(&holder, ip, &w->inv_planes);
    w->allocations = holder.allocations;
  }
  if (st != S2H_OK) {
    s2h_wigner_plan_destroy(w);
    return st;
  }
  *out = w;
  return S2H_OK;
}
