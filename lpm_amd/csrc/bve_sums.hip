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

// ---------------------------------------------------------- FP64 kernel ---
template <int BLOCK, int TPL, int TILE>
__global__ __launch_bounds__(BLOCK) void bve_vel_f64_kernel(
    const double4* __restrict__ src, const int32_t* __restrict__ count,
    SegArgs s0, SegArgs s1, double eps2) {
  __shared__ double4 tile[TILE];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  const int nsrc = *count;

  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  const int64_t slot0 = static_cast<int64_t>(blk) * (BLOCK * TPL);
  if (slot0 >= n_live) return;  // whole block idle (uniform exit)

  double x[TPL][3], S[TPL][3];
  int self[TPL];
  bool live[TPL];
  int64_t tgt[TPL];
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
    S[k][0] = S[k][1] = S[k][2] = 0.0;
  }

  for (int j0 = 0; j0 < nsrc; j0 += TILE) {
    __syncthreads();
    for (int t = threadIdx.x; t < TILE; t += BLOCK) {
      const int j = j0 + t;
      tile[t] = j < nsrc ? src[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    __syncthreads();
    bool diag = false;
#pragma unroll
    for (int k = 0; k < TPL; ++k)
      diag |= static_cast<unsigned>(self[k] - j0) < static_cast<unsigned>(TILE);
    if (__ballot(diag)) {
      for (int jj = 0; jj < TILE; ++jj) {
        const double4 y = tile[jj];
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
          const double t = x[k][0] * y.x + x[k][1] * y.y + x[k][2] * y.z;
          double w = y.w * rcp_f64((1.0 - t) + eps2);
          w = (j0 + jj == self[k]) ? 0.0 : w;
          S[k][0] = __builtin_fma(w, y.x, S[k][0]);
          S[k][1] = __builtin_fma(w, y.y, S[k][1]);
          S[k][2] = __builtin_fma(w, y.z, S[k][2]);
        }
      }
    } else {
#pragma unroll 4
      for (int jj = 0; jj < TILE; ++jj) {
        const double4 y = tile[jj];
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
          const double t = x[k][0] * y.x + x[k][1] * y.y + x[k][2] * y.z;
          const double w = y.w * rcp_f64((1.0 - t) + eps2);
          S[k][0] = __builtin_fma(w, y.x, S[k][0]);
          S[k][1] = __builtin_fma(w, y.y, S[k][1]);
          S[k][2] = __builtin_fma(w, y.z, S[k][2]);
        }
      }
    }
  }

#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    if (!live[k]) continue;
    double* o = sg.out + tgt[k] * sg.oes;
    o[0] = x[k][1] * S[k][2] - x[k][2] * S[k][1];
    o[sg.ocs] = x[k][2] * S[k][0] - x[k][0] * S[k][2];
    o[2 * sg.ocs] = x[k][0] * S[k][1] - x[k][1] * S[k][0];
  }
}

// --------------------------------------------------------- mixed kernel ---
// TILE must be a multiple of the packing tile so that one LDS tile spans
// whole packing tiles; the target offset x - c is formed in FP64 once per
// packing tile and rounded to FP32.
template <int BLOCK, int TPL, int TILE>
__global__ __launch_bounds__(BLOCK) void bve_vel_mixed_kernel(
    const float4* __restrict__ src, const double4* __restrict__ centers,
    const int32_t* __restrict__ count, SegArgs s0, SegArgs s1, float eps2x2) {
  static_assert(TILE % kPackTile == 0, "tile must cover whole packing tiles");
  __shared__ float4 tile[TILE];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  const int nsrc = *count;

  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  const int64_t slot0 = static_cast<int64_t>(blk) * (BLOCK * TPL);
  if (slot0 >= n_live) return;  // whole block idle (uniform exit)

  double x[TPL][3], S[TPL][3];
  int self[TPL];
  bool live[TPL];
  int64_t tgt[TPL];
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
    S[k][0] = S[k][1] = S[k][2] = 0.0;
  }

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
            const float r2 = d0 * d0 + d1 * d1 + d2 * d2 + eps2x2;
            float w = y.w * __builtin_amdgcn_rcpf(r2);
            w = (j0 + p0 + jj == self[k]) ? 0.0f : w;
            T[k][0] = __builtin_fmaf(w, d0, T[k][0]);
            T[k][1] = __builtin_fmaf(w, d1, T[k][1]);
            T[k][2] = __builtin_fmaf(w, d2, T[k][2]);
          }
        }
      } else {
#pragma unroll 8
        for (int jj = 0; jj < kPackTile; ++jj) {
          const float4 y = tile[p0 + jj];
#pragma unroll
          for (int k = 0; k < TPL; ++k) {
            const float d0 = y.x - xr[k][0];
            const float d1 = y.y - xr[k][1];
            const float d2 = y.z - xr[k][2];
            const float r2 = d0 * d0 + d1 * d1 + d2 * d2 + eps2x2;
            const float w = y.w * __builtin_amdgcn_rcpf(r2);
            T[k][0] = __builtin_fmaf(w, d0, T[k][0]);
            T[k][1] = __builtin_fmaf(w, d1, T[k][1]);
            T[k][2] = __builtin_fmaf(w, d2, T[k][2]);
          }
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

#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    if (!live[k]) continue;
    double* o = sg.out + tgt[k] * sg.oes;
    o[0] = x[k][1] * S[k][2] - x[k][2] * S[k][1];
    o[sg.ocs] = x[k][2] * S[k][0] - x[k][0] * S[k][2];
    o[2 * sg.ocs] = x[k][0] * S[k][1] - x[k][1] * S[k][0];
  }
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
  F64Kernel fn;
  int block, tpl;
};
struct MixedVariant {
  MixedKernel fn;
  int block, tpl;
};

const F64Variant kF64Variants[] = {
    {bve_vel_f64_kernel<256, 1, 256>, 256, 1},
    {bve_vel_f64_kernel<256, 2, 256>, 256, 2},
    {bve_vel_f64_kernel<128, 2, 256>, 128, 2},
    {bve_vel_f64_kernel<256, 4, 256>, 256, 4},
};
const MixedVariant kMixedVariants[] = {
    {bve_vel_mixed_kernel<256, 1, 256>, 256, 1},
    {bve_vel_mixed_kernel<256, 2, 256>, 256, 2},
    {bve_vel_mixed_kernel<128, 2, 256>, 128, 2},
    {bve_vel_mixed_kernel<256, 4, 256>, 256, 4},
};

int default_variant(lpm_ctx* ctx, const char* kind) {
  auto it = ctx->tuned.find(kind);
  return it == ctx->tuned.end() ? 1 : it->second;
}

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
    const F64Variant& v = kF64Variants[default_variant(ctx, "bve_f64")];
    const int per = v.block * v.tpl;
    SegArgs a0 = to_args(t0, per), a1 = to_args(t1, per);
    const int grid = a0.nblocks + a1.nblocks;
    TimedLaunch tl(ctx, KID_BVE_VEL_F64, s, pairs);
    hipLaunchKernelGGL(v.fn, dim3(grid), dim3(v.block), 0, s,
                       static_cast<const double4*>(src.data), src.count, a0,
                       a1, eps2);
    LPM_HIP_CHECK(hipGetLastError());
  } else if (precision == LPM_MIXED) {
    const MixedVariant& v = kMixedVariants[default_variant(ctx, "bve_mixed")];
    const int per = v.block * v.tpl;
    SegArgs a0 = to_args(t0, per), a1 = to_args(t1, per);
    const int grid = a0.nblocks + a1.nblocks;
    TimedLaunch tl(ctx, KID_BVE_VEL_MIXED, s, pairs);
    hipLaunchKernelGGL(v.fn, dim3(grid), dim3(v.block), 0, s,
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
