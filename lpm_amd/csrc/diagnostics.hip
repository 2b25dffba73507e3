// Diagnostics on device fields: the relative error norms of ErrNorms
// (src/lpm_error.hpp:81-131, src/lpm_error_impl.hpp:61-182) and the
// tangency dot product of SphereTangentFunctor
// (src/lpm_sphere_functions.hpp:62-78).  O(N), HBM-bound.
//
// Error norms: per element e_i = (mask_i ? 0 : appx_i - exact_i) with
// |.| the Euclidean magnitude over the ncomp components, and
//   l1 = sum |e_i| w_i / sum |exact_i| w_i,
//   l2 = sqrt(sum |e_i|^2 w_i / sum |exact_i|^2 w_i),
//   linf = max |e_i| / max |exact_i|,
// every quotient taken as num * den / (den^2 + eps^2) with eps the FP64
// machine epsilon (the reference's safe denominator).  Masked elements keep
// their exact magnitude in the denominators, as in the reference.  The six
// sums are reduced in a fixed order (per-block partials, then one block),
// so the result does not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "internal.hpp"

namespace lpm_amd {
namespace {

constexpr int kDiagBlock = 256;
constexpr int kDiagMaxBlocks = 4096;

struct NormSums {
  double l1n, l1d, l2n, l2d, lin, lid;
};

__device__ __forceinline__ void wave_reduce(NormSums& a) {
  for (int off = 32; off > 0; off >>= 1) {
    a.l1n += __shfl_down(a.l1n, off, 64);
    a.l1d += __shfl_down(a.l1d, off, 64);
    a.l2n += __shfl_down(a.l2n, off, 64);
    a.l2d += __shfl_down(a.l2d, off, 64);
    a.lin = fmax(a.lin, __shfl_down(a.lin, off, 64));
    a.lid = fmax(a.lid, __shfl_down(a.lid, off, 64));
  }
}

// Block partials of the six sums over a grid-stride range; writes the error
// field when err is given.
__global__ __launch_bounds__(kDiagBlock) void err_partial_kernel(
    const double* __restrict__ appx, int64_t aes, int64_t acs,
    const double* __restrict__ exact, int64_t xes, int64_t xcs,
    const double* __restrict__ wt, const uint8_t* __restrict__ mask,
    int64_t n, int ncomp, double* __restrict__ err, int64_t ees, int64_t ecs,
    double* __restrict__ partial) {
  NormSums a{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const int64_t stride = static_cast<int64_t>(gridDim.x) * kDiagBlock;
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kDiagBlock + threadIdx.x;
       i < n; i += stride) {
    const bool masked = mask && mask[i];
    double em = 0.0, xm = 0.0;
    for (int k = 0; k < ncomp; ++k) {
      const double xv = exact[i * xes + k * xcs];
      const double e = masked ? 0.0 : appx[i * aes + k * acs] - xv;
      if (err) err[i * ees + k * ecs] = e;
      em += e * e;
      xm += xv * xv;
    }
    em = sqrt(em);
    xm = sqrt(xm);
    const double w = wt ? wt[i] : 1.0;
    a.l1n += em * w;
    a.l1d += xm * w;
    a.l2n += em * em * w;
    a.l2d += xm * xm * w;
    a.lin = fmax(a.lin, em);
    a.lid = fmax(a.lid, xm);
  }
  wave_reduce(a);
  __shared__ NormSums red[kDiagBlock / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    NormSums t = red[0];
    for (int w = 1; w < kDiagBlock / 64; ++w) {
      t.l1n += red[w].l1n;
      t.l1d += red[w].l1d;
      t.l2n += red[w].l2n;
      t.l2d += red[w].l2d;
      t.lin = fmax(t.lin, red[w].lin);
      t.lid = fmax(t.lid, red[w].lid);
    }
    double* p = partial + 6 * static_cast<int64_t>(blockIdx.x);
    p[0] = t.l1n;
    p[1] = t.l1d;
    p[2] = t.l2n;
    p[3] = t.l2d;
    p[4] = t.lin;
    p[5] = t.lid;
  }
}

__device__ __forceinline__ double safe_quotient(double num, double den) {
  return num * (den / (den * den + DBL_EPSILON * DBL_EPSILON));
}

// One block: the partials of nb blocks in block order, then the norms.
__global__ __launch_bounds__(kDiagBlock) void err_final_kernel(
    const double* __restrict__ partial, int nb, double* __restrict__ out) {
  NormSums a{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nb; b += kDiagBlock) {
    const double* p = partial + 6 * b;
    a.l1n += p[0];
    a.l1d += p[1];
    a.l2n += p[2];
    a.l2d += p[3];
    a.lin = fmax(a.lin, p[4]);
    a.lid = fmax(a.lid, p[5]);
  }
  wave_reduce(a);
  __shared__ NormSums red[kDiagBlock / 64];
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a;
  __syncthreads();
  if (threadIdx.x == 0) {
    NormSums t = red[0];
    for (int w = 1; w < kDiagBlock / 64; ++w) {
      t.l1n += red[w].l1n;
      t.l1d += red[w].l1d;
      t.l2n += red[w].l2n;
      t.l2d += red[w].l2d;
      t.lin = fmax(t.lin, red[w].lin);
      t.lid = fmax(t.lid, red[w].lid);
    }
    out[0] = safe_quotient(t.l1n, t.l1d);
    out[1] = sqrt(safe_quotient(t.l2n, t.l2d));
    out[2] = safe_quotient(t.lin, t.lid);
  }
}

__global__ __launch_bounds__(256) void tangency_kernel(
    const double* __restrict__ u, int64_t ues, int64_t ucs,
    const double* __restrict__ x, int64_t xes, int64_t xcs, int64_t n,
    double* __restrict__ out) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  if (i >= n) return;
  out[i] = x[i * xes] * u[i * ues] + x[i * xes + xcs] * u[i * ues + ucs] +
           x[i * xes + 2 * xcs] * u[i * ues + 2 * ucs];
}

}  // namespace
}  // namespace lpm_amd

using namespace lpm_amd;

extern "C" {

int lpm_err_norms(lpm_ctx* ctx, const double* appx, int64_t appx_es,
                  int64_t appx_cs, const double* exact, int64_t exact_es,
                  int64_t exact_cs, const double* weight, const uint8_t* mask,
                  int64_t n, int ncomp, double* err, int64_t err_es,
                  int64_t err_cs, double* norms, void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(norms, "lpm_err_norms: norms is NULL");
    LPM_ARG_CHECK(n >= 0, "lpm_err_norms: negative size");
    LPM_ARG_CHECK(ncomp >= 1 && ncomp <= 3, "lpm_err_norms: ncomp must be 1..3");
    LPM_ARG_CHECK(n == 0 || (appx && exact), "lpm_err_norms: appx/exact NULL");
    LPM_ARG_CHECK(appx_es > 0 && appx_cs > 0 && exact_es > 0 && exact_cs > 0,
                  "lpm_err_norms: strides must be > 0");
    LPM_ARG_CHECK(!err || (err_es > 0 && err_cs > 0),
                  "lpm_err_norms: err strides must be > 0");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const int64_t want = (n + kDiagBlock - 1) / kDiagBlock;
    const int nb = static_cast<int>(
        std::max<int64_t>(1, std::min<int64_t>(want, kDiagMaxBlocks)));
    double* part = static_cast<double*>(
        ctx->diag.ensure(sizeof(double) * (6 * static_cast<size_t>(nb) + 3)));
    double* res = part + 6 * nb;
    err_partial_kernel<<<nb, kDiagBlock, 0, s>>>(
        appx, appx_es, appx_cs, exact, exact_es, exact_cs, weight, mask, n,
        ncomp, err, err_es, err_cs, part);
    LPM_HIP_CHECK(hipGetLastError());
    err_final_kernel<<<1, kDiagBlock, 0, s>>>(part, nb, res);
    LPM_HIP_CHECK(hipGetLastError());
    LPM_HIP_CHECK(hipMemcpyAsync(norms, res, 3 * sizeof(double),
                                 hipMemcpyDeviceToHost, s));
    LPM_HIP_CHECK(hipStreamSynchronize(s));
    return LPM_OK;
  } catch (const Error& e) {
    set_error(ctx, e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_error(ctx, e.what());
    return LPM_ERR_INTERNAL;
  }
}

int lpm_sphere_tangency(lpm_ctx* ctx, const double* u, int64_t u_es,
                        int64_t u_cs, const double* x, int64_t x_es,
                        int64_t x_cs, int64_t n, double* out, void* stream) {
  try {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(n >= 0, "lpm_sphere_tangency: negative size");
    LPM_ARG_CHECK(n == 0 || (u && x && out),
                  "lpm_sphere_tangency: NULL argument");
    LPM_ARG_CHECK(u_es > 0 && u_cs > 0 && x_es > 0 && x_cs > 0,
                  "lpm_sphere_tangency: strides must be > 0");
    if (n == 0) return LPM_OK;
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    tangency_kernel<<<static_cast<int>((n + 255) / 256), 256, 0, s>>>(
        u, u_es, u_cs, x, x_es, x_cs, n, out);
    LPM_HIP_CHECK(hipGetLastError());
    return LPM_OK;
  } catch (const Error& e) {
    set_error(ctx, e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_error(ctx, e.what());
    return LPM_ERR_INTERNAL;
  }
}

}  // extern "C"
