// extern "C" entry points of the learned preconditioner.
#include "s2h_plan.h"

using namespace s2h;

namespace {

int hip_status(hipError_t e) { return e == hipSuccess ? S2H_OK : S2H_ERR_HIP; }

// reality: 0 complex maps, 1 real maps, 2 real maps in the packed m-major layout
int row_args(const void* h, void* out, int L, int reality, int lmin, int batch, MlpRowArgs& a) {
  if (!h || L < 1 || batch < 0 || lmin < 0 || lmin >= L || reality < 0 || reality > 2)
    return S2H_ERR_INVALID;
  a = MlpRowArgs{};
  a.h = static_cast<const cplx*>(h);
  a.out = static_cast<cplx*>(out);
  a.L = L;
  a.reality = reality ? 1 : 0;
  a.packed = reality == 2 ? 1 : 0;
  a.lmin = lmin;
  a.batch = batch;
  a.P = reality ? (int64_t)L * (L + 1) / 2 : (int64_t)L * L;
  const int64_t below = reality ? (int64_t)lmin * (lmin + 1) / 2 : (int64_t)lmin * lmin;
  a.count = a.P - below;
  a.map_stride = a.packed ? a.P : (int64_t)L * (2 * L - 1);
  return S2H_OK;
}

}  // namespace

extern "C" int s2h_mlp_param_count(void) { return kMlpParams; }

extern "C" size_t s2h_mlp_workspace_bytes(int batch) {
  return batch > 0 ? (size_t)batch * (kMlpSumsqParts + kMlpDotParts) * sizeof(double) : 0;
}

namespace {

int precondition(const float* params, const void* h, void* delta, int L, int reality, int lmin,
                 int batch, double* dot, void* work, s2h_stream stream) {
  MlpRowArgs a;
  int st = row_args(h, delta, L, reality, lmin, batch, a);
  if (st) return st;
  if (!params || !delta || (batch > 0 && !work)) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  if (h == delta) return S2H_ERR_INVALID;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* sumsq = static_cast<double*>(work);
  double* dot_part = sumsq + (size_t)batch * kMlpSumsqParts;
  // entries outside the rows (m > l) stay zero; the packed layout has none
  if (!a.packed && (st = hip_status(launch_zero(delta, 2 * a.map_stride * batch, s)))) return st;
  if ((st = hip_status(launch_mlp_sumsq(a, sumsq, s)))) return st;
  return hip_status(launch_mlp_apply(a, params, sumsq, dot_part, dot, s));
}

}  // namespace

extern "C" int s2h_mlp_precondition(const float* params, const void* h, void* delta, int L,
                                    int reality, int lmin, int batch, void* work,
                                    s2h_stream stream) {
  return precondition(params, h, delta, L, reality, lmin, batch, nullptr, work, stream);
}

extern "C" int s2h_mlp_precondition_dot(const float* params, const void* h, void* delta, int L,
                                        int reality, int lmin, int batch, double* dot,
                                        void* work, s2h_stream stream) {
  if (!dot && batch > 0) return S2H_ERR_INVALID;
  return precondition(params, h, delta, L, reality, lmin, batch, dot, work, stream);
}

extern "C" int s2h_mlp_features(const void* h, const void* e, int L, int reality, int lmin,
                                int batch, float* X, float* Y, void* work, s2h_stream stream) {
  MlpRowArgs a;
  int st = row_args(h, nullptr, L, reality, lmin, batch, a);
  if (st) return st;
  if (!e || !X || !Y || (batch > 0 && !work)) return S2H_ERR_INVALID;
  if (batch == 0) return S2H_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* sumsq = static_cast<double*>(work);
  if ((st = hip_status(launch_mlp_sumsq(a, sumsq, s)))) return st;
  return hip_status(launch_mlp_features(a, static_cast<const cplx*>(e), sumsq, X, Y, s));
}

extern "C" size_t s2h_mlp_grad_workspace_bytes(int nwg) {
  if (nwg < 1) return 0;
  return (size_t)nwg * kMlpParams * sizeof(float) + (size_t)nwg * sizeof(double);
}

extern "C" int s2h_mlp_grad(const float* params, const float* X, const float* Y,
                            const float* row_weight, int64_t n, float loss_scale, float* grad,
                            double* loss, void* work, int nwg, s2h_stream stream) {
  if (!params || !grad || n < 0 || nwg < 1 || nwg > 65535 || !work) return S2H_ERR_INVALID;
  if (n > 0 && (!X || !Y)) return S2H_ERR_INVALID;
  float* partial = static_cast<float*>(work);
  double* lpart = reinterpret_cast<double*>(partial + (size_t)nwg * kMlpParams);
  return hip_status(launch_mlp_grad(params, X, Y, row_weight, n, loss_scale, nwg, partial, lpart,
                                    grad, loss, static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_mlp_adam(float* params, const float* grad, float* m1, float* m2, int step,
                            float lr, float beta1, float beta2, float eps, s2h_stream stream) {
  if (!params || !grad || !m1 || !m2 || step < 1) return S2H_ERR_INVALID;
  return hip_status(launch_mlp_adam(params, grad, m1, m2, step, lr, beta1, beta2, eps,
                                    static_cast<hipStream_t>(stream)));
}

extern "C" int s2h_probe_mfma_bf16(const uint16_t* a, const uint16_t* b, float* out,
                                   s2h_stream stream) {
  if (!a || !b || !out) return S2H_ERR_INVALID;
  return hip_status(launch_mfma_bf16_probe(a, b, out, static_cast<hipStream_t>(stream)));
}
