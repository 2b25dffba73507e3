// Tiled all-pairs Biot-Savart / Green's-function sums on the unit sphere for
// gfx950 (CDNA4), replacing LPM's one-team-per-target Kokkos reductions
// (src/lpm_bve_sphere_kernels.hpp:20-394).
//
// Prologue: undivided faces are compacted (mask scan) and packed as
//   FP64:  double4 {y0, y1, y2, s},  s = -zeta A / (4 pi)
//   mixed: float4  {y0-c0, y1-c1, y2-c2, 2 s} relative to the FP64 centre c
//          of each 256-source packing tile, plus double4 centres.
// Main loop: each lane owns TPL targets held in registers; a block stages
// TILE packed sources in LDS and every lane reads them as wave-uniform
// broadcasts.  Per pair the FP64 kernel evaluates t = x.y, w = s/(1-t+eps^2)
// and accumulates S += w y; the velocity is x cross S, which equals
// sum_j w_j (x cross y_j) because x is fixed per target.  The mixed kernel
// evaluates d = y - x in tile-relative FP32, r2 = |d|^2 (= 2(1 - x.y) on the
// unit sphere), w = 2s / r2 and S += w d, carrying per-tile FP32 partials
// into FP64 sums; x cross sum w (y - x) = x cross sum w y.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace lpm_amd {

namespace {

constexpr double kInv4Pi = 0.079577471545947667884441881686257181;
constexpr int kPackTile = 256;  // mixed-precision packing tile (sources)
constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;  // elements per thread in the mask scan
constexpr int kScanChunk = kScanBlock * kScanItems;

// ------------------------------------------------------------ prologue ---
__global__ __launch_bounds__(kScanBlock) void mask_count_kernel(
    const uint8_t* __restrict__ mask, int64_t n, int32_t* __restrict__ counts) {
  __shared__ int32_t red[kScanBlock / 64];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk;
  int c = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + threadIdx.x * kScanItems + k;
    if (i < n && !mask[i]) ++c;
  }
  // wave reduction then across waves
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) t += red[w];
    counts[blockIdx.x] = t;
  }
}

// Exclusive scan of nb block counts in place; writes the total to counts[nb].
__global__ __launch_bounds__(1024) void scan_counts_kernel(
    int32_t* __restrict__ counts, int nb) {
  __shared__ int32_t buf[1024];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b0 = 0; b0 < nb; b0 += 1024) {
    const int i = b0 + threadIdx.x;
    const int v = i < nb ? counts[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < nb) counts[i] = carry + buf[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) counts[nb] = carry;
}

__global__ __launch_bounds__(kScanBlock) void mask_rank_kernel(
    const uint8_t* __restrict__ mask, int64_t n,
    const int32_t* __restrict__ offsets, int32_t* __restrict__ rank,
    int32_t* __restrict__ inv) {
  __shared__ int32_t buf[kScanBlock];
  const int64_t base = static_cast<int64_t>(blockIdx.x) * kScanChunk +
                       static_cast<int64_t>(threadIdx.x) * kScanItems;
  int c = 0;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n && !mask[i]) ++c;
  }
  buf[threadIdx.x] = c;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    const int add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
    __syncthreads();
    buf[threadIdx.x] += add;
    __syncthreads();
  }
  int r = offsets[blockIdx.x] + buf[threadIdx.x] - c;
  for (int k = 0; k < kScanItems; ++k) {
    const int64_t i = base + k;
    if (i < n) {
      if (mask[i]) {
        rank[i] = -1;
      } else {
        rank[i] = r;
        inv[r] = static_cast<int32_t>(i);
        ++r;
      }
    }
  }
}

__global__ void set_count_kernel(int32_t* __restrict__ count, int32_t value) {
  *count = value;
}

// Gathers the packed sources in packed order (coalesced writes).
__global__ __launch_bounds__(256) void pack_f64_kernel(
    const double* __restrict__ sx, int64_t es, int64_t cs,
    const double* __restrict__ zeta, const double* __restrict__ area,
    const int32_t* __restrict__ inv, const int32_t* __restrict__ count,
    double4* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= *count) return;
  const int64_t i = inv ? inv[r] : r;
  const double* p = sx + i * es;
  out[r] = make_double4(p[0], p[cs], p[2 * cs], -zeta[i] * area[i] * kInv4Pi);
}

// One block per packing tile: FP64 centre = first source of the tile.
__global__ __launch_bounds__(kPackTile) void pack_mixed_kernel(
    const double4* __restrict__ p64, const int32_t* __restrict__ count,
    float4* __restrict__ out, double4* __restrict__ centers) {
  const int n = *count;
  const int j0 = blockIdx.x * kPackTile;
  if (j0 >= n) return;
  const double4 c = p64[j0];
  if (threadIdx.x == 0) centers[blockIdx.x] = c;
  const int j = j0 + threadIdx.x;
  if (j < n) {
    const double4 y = p64[j];
    out[j] = make_float4(static_cast<float>(y.x - c.x),
                         static_cast<float>(y.y - c.y),
                         static_cast<float>(y.z - c.z),
                         static_cast<float>(2.0 * y.w));
  }
}

// ----------------------------------------------------------- arithmetic ---
// Reciprocal: hardware estimate refined by two Newton-Raphson steps
// (r <- r + r (1 - d r)), which restores full FP64 accuracy.
__device__ __forceinline__ double rcp_f64(double d) {
  double r = __builtin_amdgcn_rcp(d);
  double e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-d, r, 1.0);
  r = __builtin_fma(r, e, r);
  return r;
}

// ------------------------------------------------------- kernel argument ---
struct SegArgs {
  const double* x;
  int64_t es, cs, n;
  const int32_t* index;
  const int32_t* d_n;
  const int32_t* self_rank;
  int self_identity;
  int nblocks;
  const uint8_t* skip;
  double* out;
  int64_t oes, ocs;
  double* out_scalar;
};

__device__ __forceinline__ const SegArgs& pick(const SegArgs& a,
                                               const SegArgs& b, bool first) {
  return first ? a : b;
}

// --------------------------------------------------------- pair bodies ---
// Reciprocal with NR Newton-Raphson refinements of the hardware estimate.
template <int NR>
__device__ __forceinline__ double rcp_nr(double d) {
  double r = __builtin_amdgcn_rcp(d);
#pragma unroll
  for (int it = 0; it < NR; ++it) {
    const double e = __builtin_fma(-d, r, 1.0);
    r = __builtin_fma(r, e, r);
  }
  return r;
}

// One source: S += s / (1 - x.y [+ eps^2]) * y.
template <bool EPS, int NR>
__device__ __forceinline__ void f64_one(const double* x, const double4& y,
                                        double eps2, double* S) {
  const double t = x[0] * y.x + x[1] * y.y + x[2] * y.z;
  double d = 1.0 - t;
  if (EPS) d += eps2;
  const double w = y.w * rcp_nr<NR>(d);
  S[0] = __builtin_fma(w, y.x, S[0]);
  S[1] = __builtin_fma(w, y.y, S[1]);
  S[2] = __builtin_fma(w, y.z, S[2]);
}

// Two sources sharing one reciprocal: r = 1/(da db), 1/da = db r,
// 1/db = da r.  Terms are accumulated in source order.
template <bool EPS, int NR>
__device__ __forceinline__ void f64_two(const double* x, const double4& ya,
                                        const double4& yb, double eps2,
                                        double* S) {
  const double ta = x[0] * ya.x + x[1] * ya.y + x[2] * ya.z;
  const double tb = x[0] * yb.x + x[1] * yb.y + x[2] * yb.z;
  double da = 1.0 - ta, db = 1.0 - tb;
  if (EPS) {
    da += eps2;
    db += eps2;
  }
  const double r = rcp_nr<NR>(da * db);
  const double wa = ya.w * (db * r);
  const double wb = yb.w * (da * r);
  S[0] = __builtin_fma(wa, ya.x, S[0]);
  S[1] = __builtin_fma(wa, ya.y, S[1]);
  S[2] = __builtin_fma(wa, ya.z, S[2]);
  S[0] = __builtin_fma(wb, yb.x, S[0]);
  S[1] = __builtin_fma(wb, yb.y, S[1]);
  S[2] = __builtin_fma(wb, yb.z, S[2]);
}

// Mixed: d = y' - x' (tile-relative FP32), r2 = |d|^2 + 2 eps^2,
// T += (2 s / r2) d.
template <bool EPS>
__device__ __forceinline__ void f32_one(const float* x, const float4& y,
                                        float e2, float* T) {
  const float d0 = y.x - x[0], d1 = y.y - x[1], d2 = y.z - x[2];
  float r2 = d0 * d0 + d1 * d1 + d2 * d2;
  if (EPS) r2 += e2;
  const float w = y.w * __builtin_amdgcn_rcpf(r2);
  T[0] = __builtin_fmaf(w, d0, T[0]);
  T[1] = __builtin_fmaf(w, d1, T[1]);
  T[2] = __builtin_fmaf(w, d2, T[2]);
}

template <bool EPS>
__device__ __forceinline__ void f32_two(const float* x, const float4& ya,
                                        const float4& yb, float e2, float* T) {
  const float a0 = ya.x - x[0], a1 = ya.y - x[1], a2 = ya.z - x[2];
  const float b0 = yb.x - x[0], b1 = yb.y - x[1], b2 = yb.z - x[2];
  float ra = a0 * a0 + a1 * a1 + a2 * a2;
  float rb = b0 * b0 + b1 * b1 + b2 * b2;
  if (EPS) {
    ra += e2;
    rb += e2;
  }
  const float r = __builtin_amdgcn_rcpf(ra * rb);
  const float wa = ya.w * (rb * r);
  const float wb = yb.w * (ra * r);
  T[0] = __builtin_fmaf(wa, a0, T[0]);
  T[1] = __builtin_fmaf(wa, a1, T[1]);
  T[2] = __builtin_fmaf(wa, a2, T[2]);
  T[0] = __builtin_fmaf(wb, b0, T[0]);
  T[1] = __builtin_fmaf(wb, b1, T[1]);
  T[2] = __builtin_fmaf(wb, b2, T[2]);
}

// Loads the TPL targets of this thread; returns false if the whole block
// is idle.  Slot c maps to target i = index ? index[c] : c.
template <int BLOCK, int TPL>
__device__ __forceinline__ bool load_targets(const SegArgs& sg, int blk,
                                             double (*x)[3], int* self,
                                             bool* live, int64_t* tgt) {
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  const int64_t slot0 = static_cast<int64_t>(blk) * (BLOCK * TPL);
  if (slot0 >= n_live) return false;
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int64_t c = slot0 + k * BLOCK + threadIdx.x;
    const bool in = c < n_live;
    const int64_t i = in ? (sg.index ? sg.index[c] : c) : 0;
    tgt[k] = i;
    live[k] = in && !(sg.skip && sg.skip[i]);
    const double* p = sg.x + i * sg.es;
    x[k][0] = p[0];
    x[k][1] = p[sg.cs];
    x[k][2] = p[2 * sg.cs];
    self[k] = !in ? -1
                  : (sg.self_rank ? sg.self_rank[i]
                                  : (sg.self_identity ? static_cast<int>(c) : -1));
  }
  return true;
}

// u = x cross S at each live target.
template <int TPL>
__device__ __forceinline__ void store_cross(const SegArgs& sg,
                                            const double (*x)[3],
                                            const double (*S)[3],
                                            const bool* live,
                                            const int64_t* tgt) {
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    if (!live[k]) continue;
    double* o = sg.out + tgt[k] * sg.oes;
    o[0] = x[k][1] * S[k][2] - x[k][2] * S[k][1];
    o[sg.ocs] = x[k][2] * S[k][0] - x[k][0] * S[k][2];
    o[2 * sg.ocs] = x[k][0] * S[k][1] - x[k][1] * S[k][0];
  }
}

// ---------------------------------------------------------- FP64 kernel ---
// PAIR: share one reciprocal between two sources in the inner loop.
template <int BLOCK, int TPL, int TILE, bool EPS, int NR, bool PAIR>
__global__ __launch_bounds__(BLOCK) void bve_vel_f64_kernel(
    const double4* __restrict__ src, const int32_t* __restrict__ count,
    SegArgs s0, SegArgs s1, double eps2) {
  static_assert(TILE % 8 == 0, "tile must be a multiple of the unroll");
  __shared__ double4 tile[TILE];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  const int nsrc = *count;

  double x[TPL][3], S[TPL][3];
  int self[TPL];
  bool live[TPL];
  int64_t tgt[TPL];
  if (!load_targets<BLOCK, TPL>(sg, blk, x, self, live, tgt)) return;
#pragma unroll
  for (int k = 0; k < TPL; ++k) S[k][0] = S[k][1] = S[k][2] = 0.0;

  for (int j0 = 0; j0 < nsrc; j0 += TILE) {
    __syncthreads();
    for (int t = threadIdx.x; t < TILE; t += BLOCK) {
      const int j = j0 + t;
      // padding: zero strength at the origin -> 1 - x.y = 1, w = 0
      tile[t] = j < nsrc ? src[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    __syncthreads();
    bool diag = false;
#pragma unroll
    for (int k = 0; k < TPL; ++k)
      diag |= static_cast<unsigned>(self[k] - j0) < static_cast<unsigned>(TILE);
    if (__ballot(diag)) {
      // tile holds some target's own source: per-pair exclusion
      for (int jj = 0; jj < TILE; ++jj) {
        const double4 y = tile[jj];
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
          const double t = x[k][0] * y.x + x[k][1] * y.y + x[k][2] * y.z;
          double d = 1.0 - t;
          if (EPS) d += eps2;
          double w = y.w * rcp_nr<2>(d);
          w = (j0 + jj == self[k]) ? 0.0 : w;
          S[k][0] = __builtin_fma(w, y.x, S[k][0]);
          S[k][1] = __builtin_fma(w, y.y, S[k][1]);
          S[k][2] = __builtin_fma(w, y.z, S[k][2]);
        }
      }
    } else if (PAIR) {
#pragma unroll 2
      for (int jj = 0; jj < TILE; jj += 2) {
        const double4 ya = tile[jj];
        const double4 yb = tile[jj + 1];
#pragma unroll
        for (int k = 0; k < TPL; ++k) f64_two<EPS, NR>(x[k], ya, yb, eps2, S[k]);
      }
    } else {
#pragma unroll 4
      for (int jj = 0; jj < TILE; ++jj) {
        const double4 y = tile[jj];
#pragma unroll
        for (int k = 0; k < TPL; ++k) f64_one<EPS, NR>(x[k], y, eps2, S[k]);
      }
    }
  }
  store_cross<TPL>(sg, x, S, live, tgt);
}

// --------------------------------------------------------- mixed kernel ---
// TILE is a multiple of the packing tile so that one LDS tile spans whole
// packing tiles; the target offset x - c is formed in FP64 once per packing
// tile and rounded to FP32.  Per packing tile the FP32 partial is carried
// into FP64 accumulators.
template <int BLOCK, int TPL, int TILE, bool EPS, bool PAIR>
__global__ __launch_bounds__(BLOCK) void bve_vel_mixed_kernel(
    const float4* __restrict__ src, const double4* __restrict__ centers,
    const int32_t* __restrict__ count, SegArgs s0, SegArgs s1, float eps2x2) {
  static_assert(TILE % kPackTile == 0, "tile must cover whole packing tiles");
  __shared__ float4 tile[TILE];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  const int nsrc = *count;

  double x[TPL][3], S[TPL][3];
  int self[TPL];
  bool live[TPL];
  int64_t tgt[TPL];
  if (!load_targets<BLOCK, TPL>(sg, blk, x, self, live, tgt)) return;
#pragma unroll
  for (int k = 0; k < TPL; ++k) S[k][0] = S[k][1] = S[k][2] = 0.0;

  for (int j0 = 0; j0 < nsrc; j0 += TILE) {
    __syncthreads();
    for (int t = threadIdx.x; t < TILE; t += BLOCK) {
      const int j = j0 + t;
      // padding: far away, zero strength -> contributes exactly 0
      tile[t] = j < nsrc ? src[j] : make_float4(8.0f, 8.0f, 8.0f, 0.0f);
    }
    __syncthreads();
    bool diag = false;
#pragma unroll
    for (int k = 0; k < TPL; ++k)
      diag |= static_cast<unsigned>(self[k] - j0) < static_cast<unsigned>(TILE);
    const bool any_diag = __ballot(diag) != 0;

    for (int p0 = 0; p0 < TILE && j0 + p0 < nsrc; p0 += kPackTile) {
      const double4 c = centers[(j0 + p0) / kPackTile];
      float xr[TPL][3], T[TPL][3];
#pragma unroll
      for (int k = 0; k < TPL; ++k) {
        xr[k][0] = static_cast<float>(x[k][0] - c.x);
        xr[k][1] = static_cast<float>(x[k][1] - c.y);
        xr[k][2] = static_cast<float>(x[k][2] - c.z);
        T[k][0] = T[k][1] = T[k][2] = 0.0f;
      }
      if (any_diag) {
        for (int jj = 0; jj < kPackTile; ++jj) {
          const float4 y = tile[p0 + jj];
#pragma unroll
          for (int k = 0; k < TPL; ++k) {
            const float d0 = y.x - xr[k][0];
            const float d1 = y.y - xr[k][1];
            const float d2 = y.z - xr[k][2];
            float r2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (EPS) r2 += eps2x2;
            float w = y.w * __builtin_amdgcn_rcpf(r2);
            w = (j0 + p0 + jj == self[k]) ? 0.0f : w;
            T[k][0] = __builtin_fmaf(w, d0, T[k][0]);
            T[k][1] = __builtin_fmaf(w, d1, T[k][1]);
            T[k][2] = __builtin_fmaf(w, d2, T[k][2]);
          }
        }
      } else if (PAIR) {
#pragma unroll 4
        for (int jj = 0; jj < kPackTile; jj += 2) {
          const float4 ya = tile[p0 + jj];
          const float4 yb = tile[p0 + jj + 1];
#pragma unroll
          for (int k = 0; k < TPL; ++k) f32_two<EPS>(xr[k], ya, yb, eps2x2, T[k]);
        }
      } else {
#pragma unroll 8
        for (int jj = 0; jj < kPackTile; ++jj) {
          const float4 y = tile[p0 + jj];
#pragma unroll
          for (int k = 0; k < TPL; ++k) f32_one<EPS>(xr[k], y, eps2x2, T[k]);
        }
      }
#pragma unroll
      for (int k = 0; k < TPL; ++k) {
        S[k][0] += static_cast<double>(T[k][0]);
        S[k][1] += static_cast<double>(T[k][1]);
        S[k][2] += static_cast<double>(T[k][2]);
      }
    }
  }
  store_cross<TPL>(sg, x, S, live, tgt);
}

// ------------------------------------------------------ stream function ---
// psi_i = sum_j s_j log(1 - x_i.y_j); FP64 log from the device math library.
template <int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void bve_psi_f64_kernel(
    const double4* __restrict__ src, const int32_t* __restrict__ count,
    SegArgs s0, SegArgs s1) {
  __shared__ double4 tile[TILE];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  const int nsrc = *count;
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  if (static_cast<int64_t>(blk) * BLOCK >= n_live) return;
  const int64_t c = static_cast<int64_t>(blk) * BLOCK + threadIdx.x;
  const bool in = c < n_live;
  const int64_t i = in ? (sg.index ? sg.index[c] : c) : 0;
  const bool live = in && !(sg.skip && sg.skip[i]);
  const double* p = sg.x + i * sg.es;
  const double x0 = p[0], x1 = p[sg.cs], x2 = p[2 * sg.cs];
  const int self = !in ? -1
                       : (sg.self_rank ? sg.self_rank[i]
                                       : (sg.self_identity ? static_cast<int>(c) : -1));
  double psi = 0.0;
  for (int j0 = 0; j0 < nsrc; j0 += TILE) {
    __syncthreads();
    for (int t = threadIdx.x; t < TILE; t += BLOCK) {
      const int j = j0 + t;
      tile[t] = j < nsrc ? src[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    __syncthreads();
    for (int jj = 0; jj < TILE; ++jj) {
      const double4 y = tile[jj];
      const double t = x0 * y.x + x1 * y.y + x2 * y.z;
      const double v = y.w * log(1.0 - t);
      psi += (j0 + jj == self) ? 0.0 : v;
    }
  }
  if (live) sg.out_scalar[i] = psi;
}

__global__ __launch_bounds__(256) void finite_kernel(
    const double* __restrict__ base, int64_t n, int64_t es, int64_t cs,
    int ncomp, int32_t* __restrict__ flag) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool bad = false;
  for (int k = 0; k < ncomp; ++k) bad |= !isfinite(base[i * es + k * cs]);
  if (bad) atomicOr(flag, 1);
}

__global__ void zero_int_kernel(int32_t* p) { *p = 0; }

// ------------------------------------------------------ kernel registry ---
using F64Kernel = void (*)(const double4*, const int32_t*, SegArgs, SegArgs,
                           double);
using MixedKernel = void (*)(const float4*, const double4*, const int32_t*,
                             SegArgs, SegArgs, float);

struct F64Variant {
  F64Kernel fn, fn_eps;
  int block, tpl;
  const char* name;
};
struct MixedVariant {
  MixedKernel fn, fn_eps;
  int block, tpl;
  const char* name;
};

#define LPM_F64_VARIANT(B, T, NR, PAIR)                                  \
  {bve_vel_f64_kernel<B, T, 256, false, NR, PAIR>,                      \
   bve_vel_f64_kernel<B, T, 256, true, NR, PAIR>, B, T,                 \
   "block" #B "_tpl" #T "_nr" #NR "_pair" #PAIR}
#define LPM_MIXED_VARIANT(B, T, PAIR)                                     \
  {bve_vel_mixed_kernel<B, T, 256, false, PAIR>,                         \
   bve_vel_mixed_kernel<B, T, 256, true, PAIR>, B, T,                    \
   "block" #B "_tpl" #T "_pair" #PAIR}

const F64Variant kF64Variants[] = {
    LPM_F64_VARIANT(256, 2, 2, false),  // 0: baseline loop
    LPM_F64_VARIANT(256, 2, 2, true),   // 1
    LPM_F64_VARIANT(256, 1, 2, true),   // 2
    LPM_F64_VARIANT(256, 4, 2, true),   // 3
    LPM_F64_VARIANT(128, 2, 2, true),   // 4
    LPM_F64_VARIANT(256, 2, 1, true),   // 5: one Newton step
    LPM_F64_VARIANT(256, 4, 1, true),   // 6
    LPM_F64_VARIANT(256, 1, 1, true),   // 7
};
const MixedVariant kMixedVariants[] = {
    LPM_MIXED_VARIANT(256, 2, false),  // 0
    LPM_MIXED_VARIANT(256, 2, true),   // 1
    LPM_MIXED_VARIANT(256, 1, true),   // 2
    LPM_MIXED_VARIANT(256, 4, true),   // 3
    LPM_MIXED_VARIANT(128, 2, true),   // 4
    LPM_MIXED_VARIANT(256, 1, false),  // 5
    LPM_MIXED_VARIANT(256, 4, false),  // 6
};
constexpr int kNumF64 = sizeof(kF64Variants) / sizeof(kF64Variants[0]);
constexpr int kNumMixed = sizeof(kMixedVariants) / sizeof(kMixedVariants[0]);
// Variants used until the autotuner (or the caller) picks another one.
constexpr int kDefaultF64 = 1;
constexpr int kDefaultMixed = 0;

}  // namespace

int variant_count(int family) {
  if (family == KID_BVE_VEL_F64) return kNumF64;
  if (family == KID_BVE_VEL_MIXED) return kNumMixed;
  return 0;
}

const char* variant_name(int family, int v) {
  if (v < 0 || v >= variant_count(family)) return nullptr;
  return family == KID_BVE_VEL_F64 ? kF64Variants[v].name
                                   : kMixedVariants[v].name;
}

int default_variant(int family) {
  return family == KID_BVE_VEL_F64 ? kDefaultF64 : kDefaultMixed;
}

namespace {

SegArgs to_args(const TargetSeg& t, int per_block) {
  SegArgs a{};
  a.x = t.x.base;
  a.es = t.x.es;
  a.cs = t.x.cs;
  a.n = t.n;
  a.index = t.index;
  a.d_n = t.d_n;
  a.self_rank = t.self_rank;
  a.self_identity = t.self_identity ? 1 : 0;
  a.nblocks = static_cast<int>((t.n + per_block - 1) / per_block);
  a.skip = t.skip;
  a.out = t.out.base;
  a.oes = t.out.es;
  a.ocs = t.out.cs;
  a.out_scalar = t.out_scalar;
  return a;
}

double count_pairs(const TargetSeg* segs, int nseg, int64_t nsrc_upper) {
  double p = 0;
  for (int s = 0; s < nseg; ++s) p += static_cast<double>(segs[s].n);
  return p * static_cast<double>(nsrc_upper);
}

}  // namespace

// --------------------------------------------------------------- hosts ---
RankMap build_rank(lpm_ctx* ctx, const uint8_t* mask, int64_t n,
                   hipStream_t s) {
  LPM_ARG_CHECK(n >= 0 && n < (int64_t(1) << 31), "source count out of range");
  RankMap map;
  map.n = n;
  if (!mask) {
    int32_t* c = static_cast<int32_t*>(ctx->count_ident.ensure(sizeof(int32_t)));
    set_count_kernel<<<1, 1, 0, s>>>(c, static_cast<int32_t>(n));
    LPM_HIP_CHECK(hipGetLastError());
    map.d_count = c;
    return map;
  }
  const int nb = static_cast<int>((n + kScanChunk - 1) / kScanChunk);
  int32_t* counts = static_cast<int32_t*>(
      ctx->block_counts.ensure(sizeof(int32_t) * (nb + 1)));
  const size_t nn = sizeof(int32_t) * std::max<int64_t>(n, 1);
  int32_t* rank = static_cast<int32_t*>(ctx->rank.ensure(nn));
  int32_t* inv = static_cast<int32_t*>(ctx->inv.ensure(nn));
  if (nb > 0) {
    mask_count_kernel<<<nb, kScanBlock, 0, s>>>(mask, n, counts);
    LPM_HIP_CHECK(hipGetLastError());
  }
  scan_counts_kernel<<<1, 1024, 0, s>>>(counts, nb);
  LPM_HIP_CHECK(hipGetLastError());
  if (nb > 0) {
    mask_rank_kernel<<<nb, kScanBlock, 0, s>>>(mask, n, counts, rank, inv);
    LPM_HIP_CHECK(hipGetLastError());
  }
  map.rank = rank;
  map.inv = inv;
  map.d_count = counts + nb;
  return map;
}

TargetSeg face_targets(CrdIn x, const RankMap& map, const uint8_t* mask,
                       bool skip_masked, CrdOut out, double* out_scalar) {
  TargetSeg t;
  t.x = x;
  t.n = map.n;
  t.out = out;
  t.out_scalar = out_scalar;
  if (skip_masked && mask) {
    // live slots are the packed faces; slot c is its own source c
    t.index = map.inv;
    t.d_n = map.d_count;
    t.self_identity = true;
  } else if (map.rank) {
    t.self_rank = map.rank;
  } else {
    t.self_identity = true;
  }
  return t;
}

PackedSources pack_sources(lpm_ctx* ctx, CrdIn sx, const double* zeta,
                           const double* area, const RankMap& map,
                           int precision, hipStream_t s) {
  PackedSources ps;
  const int64_t n = map.n;
  ps.count = map.d_count;
  ps.capacity = n;
  const int64_t cap = std::max<int64_t>(n, 1);
  double4* p64 =
      static_cast<double4*>(ctx->packed64.ensure(sizeof(double4) * cap));
  if (n > 0) {
    pack_f64_kernel<<<static_cast<int>((n + 255) / 256), 256, 0, s>>>(
        sx.base, sx.es, sx.cs, zeta, area, map.inv, map.d_count, p64);
    LPM_HIP_CHECK(hipGetLastError());
  }
  ps.data = p64;
  if (precision == LPM_MIXED) {
    const int ntile = static_cast<int>((cap + kPackTile - 1) / kPackTile);
    float4* p32 =
        static_cast<float4*>(ctx->packed32.ensure(sizeof(float4) * cap));
    double4* cen = static_cast<double4*>(
        ctx->centers.ensure(sizeof(double4) * ntile));
    pack_mixed_kernel<<<ntile, kPackTile, 0, s>>>(p64, map.d_count, p32, cen);
    LPM_HIP_CHECK(hipGetLastError());
    ps.data = p32;
    ps.centers = cen;
  }
  return ps;
}

void launch_bve_velocity(lpm_ctx* ctx, const PackedSources& src,
                         TargetSeg* segs, int nseg, double eps2, int precision,
                         hipStream_t s) {
  LPM_ARG_CHECK(nseg >= 1 && nseg <= 2, "one or two target segments");
  TargetSeg empty;
  const TargetSeg& t0 = segs[0];
  const TargetSeg& t1 = nseg > 1 ? segs[1] : empty;
  for (int k = 0; k < nseg; ++k)
    LPM_ARG_CHECK(segs[k].n >= 0 && segs[k].n < (int64_t(1) << 31),
                  "target count out of range");
  if (t0.n + t1.n == 0) return;
  const double pairs = count_pairs(segs, nseg, src.capacity);
  if (precision == LPM_FP64) {
    const F64Variant& v =
        kF64Variants[pick_variant(ctx, KID_BVE_VEL_F64, t0.n + t1.n)];
    const int per = v.block * v.tpl;
    SegArgs a0 = to_args(t0, per), a1 = to_args(t1, per);
    const int grid = a0.nblocks + a1.nblocks;
    TimedLaunch tl(ctx, KID_BVE_VEL_F64, s, pairs);
    hipLaunchKernelGGL(eps2 > 0 ? v.fn_eps : v.fn, dim3(grid), dim3(v.block), 0, s,
                       static_cast<const double4*>(src.data), src.count, a0,
                       a1, eps2);
    LPM_HIP_CHECK(hipGetLastError());
  } else if (precision == LPM_MIXED) {
    const MixedVariant& v =
        kMixedVariants[pick_variant(ctx, KID_BVE_VEL_MIXED, t0.n + t1.n)];
    const int per = v.block * v.tpl;
    SegArgs a0 = to_args(t0, per), a1 = to_args(t1, per);
    const int grid = a0.nblocks + a1.nblocks;
    TimedLaunch tl(ctx, KID_BVE_VEL_MIXED, s, pairs);
    hipLaunchKernelGGL(eps2 > 0 ? v.fn_eps : v.fn, dim3(grid), dim3(v.block), 0, s,
                       static_cast<const float4*>(src.data), src.centers,
                       src.count, a0, a1, static_cast<float>(2.0 * eps2));
    LPM_HIP_CHECK(hipGetLastError());
  } else {
    throw Error(LPM_ERR_ARG, "unknown precision");
  }
}

void launch_bve_stream_fn(lpm_ctx* ctx, const PackedSources& src,
                          TargetSeg* segs, int nseg, hipStream_t s) {
  TargetSeg empty;
  const TargetSeg& t0 = segs[0];
  const TargetSeg& t1 = nseg > 1 ? segs[1] : empty;
  if (t0.n + t1.n == 0) return;
  constexpr int kBlock = 256;
  SegArgs a0 = to_args(t0, kBlock), a1 = to_args(t1, kBlock);
  TimedLaunch tl(ctx, KID_BVE_PSI, s, count_pairs(segs, nseg, src.capacity));
  hipLaunchKernelGGL((bve_psi_f64_kernel<kBlock, 256>),
                     dim3(a0.nblocks + a1.nblocks), dim3(kBlock), 0, s,
                     static_cast<const double4*>(src.data), src.count, a0, a1);
  LPM_HIP_CHECK(hipGetLastError());
}

void check_finite(lpm_ctx* ctx, const double* base, int64_t n, int64_t es,
                  int64_t cs, int ncomp, hipStream_t s) {
  if (n == 0) return;
  int32_t* flag = static_cast<int32_t*>(ctx->flag.ensure(sizeof(int32_t)));
  zero_int_kernel<<<1, 1, 0, s>>>(flag);
  finite_kernel<<<static_cast<int>((n + 255) / 256), 256, 0, s>>>(
      base, n, es, cs, ncomp, flag);
  LPM_HIP_CHECK(hipGetLastError());
  int32_t h = 0;
  LPM_HIP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int32_t),
                               hipMemcpyDeviceToHost, s));
  LPM_HIP_CHECK(hipStreamSynchronize(s));
  if (h) throw Error(LPM_ERR_NONFINITE, "non-finite value in direct-sum output");
}

}  // namespace lpm_amd
