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

#include "bve_device.hpp"
#include "internal.hpp"

namespace lpm_amd {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;  // elements per thread in the mask scan
constexpr int kScanChunk = kScanBlock * kScanItems;
static_assert(kScanChunk == kMaskScanChunk, "scan chunk must match the header");

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

// Gathers the packed sources in packed order (coalesced writes); out_sy
// (optional) receives {s y, 0} for the kernels that take the weight from it.
__global__ __launch_bounds__(256) void pack_f64_kernel(
    const double* __restrict__ sx, int64_t es, int64_t cs,
    const double* __restrict__ zeta, const double* __restrict__ area,
    const int32_t* __restrict__ inv, const int32_t* __restrict__ count,
    double4* __restrict__ out, double4* __restrict__ out_sy) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= *count) return;
  const int64_t i = inv ? inv[r] : r;
  const double* p = sx + i * es;
  const double y0 = p[0], y1 = p[cs], y2 = p[2 * cs];
  const double sv = -zeta[i] * area[i] * kInv4Pi;
  out[r] = make_double4(y0, y1, y2, sv);
  if (out_sy) out_sy[r] = make_double4(sv * y0, sv * y1, sv * y2, 0.0);
}

// ------------------------------------------------------- velocity sums ---
// One launch covers two target segments (vertices, then faces): blocks
// [0, s0.nblocks) take s0, the rest s1.
// SY: the weights come from the packed s y (src_sy), one multiply less per
// pair; otherwise src_sy is not read.
template <int BLOCK, int TPL, int TILE, bool EPS, int NR, bool PAIR,
          bool FUSE = false, bool QUAD = false, bool SY = false, int QU = 1>
__global__ __launch_bounds__(BLOCK) void bve_vel_f64_kernel(
    const double4* __restrict__ src, const double4* __restrict__ src_sy,
    const int32_t* __restrict__ count, SegArgs s0, SegArgs s1, double eps2) {
  __shared__ double4 tile[TILE];
  __shared__ double4 tile_sy[SY ? TILE : 1];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  f64_block<BLOCK, TPL, TILE, EPS, NR, PAIR, FUSE, QUAD, SY, QU>(
      src, src_sy, *count, sg, blk, eps2, tile, tile_sy);
}

// Source-split form: block b * KS + k sums source chunk k of target block
// b into sg.part; bve_split_reduce_kernel then adds the KS chunks in order
// and writes x cross S.  Equal-sized work items finer than a whole target
// block let every CU finish at nearly the same time.
template <int BLOCK, int TPL, int TILE, bool EPS, int NR, bool PAIR,
          bool FUSE, bool QUAD, bool SY, int QU, int KS>
__global__ __launch_bounds__(BLOCK) void bve_vel_f64_split_kernel(
    const double4* __restrict__ src, const double4* __restrict__ src_sy,
    const int32_t* __restrict__ count, SegArgs s0, SegArgs s1, double eps2) {
  __shared__ double4 tile[TILE];
  __shared__ double4 tile_sy[SY ? TILE : 1];
  const int b = static_cast<int>(blockIdx.x) / KS;
  const int chunk = static_cast<int>(blockIdx.x) % KS;
  const bool first = b < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? b : b - s0.nblocks;
  f64_block<BLOCK, TPL, TILE, EPS, NR, PAIR, FUSE, QUAD, SY, QU, KS>(
      src, src_sy, *count, sg, blk, eps2, tile, tile_sy, chunk);
}

// u = x cross (sum over the ks chunks of S, in chunk order) at every live
// slot of both segments.
__global__ __launch_bounds__(256) void bve_split_reduce_kernel(SegArgs s0,
                                                               SegArgs s1,
                                                               int ks) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const bool first = g < s0.n;
  const SegArgs& sg = pick(s0, s1, first);
  const int64_t c = first ? g : g - s0.n;
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  if (c >= n_live) return;
  const int64_t i = sg.index ? sg.index[c] : sg.slot_base + c;
  if (sg.skip && sg.skip[i]) return;
  split_reduce_one(sg.part + c, sg.pstride, ks, sg.x + i * sg.es, sg.cs,
                   sg.out + i * sg.oes, sg.ocs);
}

template <int BLOCK, int TPL, int TILE, bool EPS, bool PAIR>
__global__ __launch_bounds__(BLOCK) void bve_vel_mixed_kernel(
    const float4* __restrict__ src, const double4* __restrict__ centers,
    const int32_t* __restrict__ count, SegArgs s0, SegArgs s1, float eps2x2) {
  __shared__ float4 tile[TILE];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  mixed_block<BLOCK, TPL, TILE, EPS, PAIR>(src, centers, *count, sg, blk,
                                           eps2x2, tile);
}

template <bool EPS>
__global__ __launch_bounds__(256) void bve_vel_mfma_kernel(
    MfmaSrc src, const int32_t* __restrict__ count, SegArgs s0, SegArgs s1,
    float eps2x2) {
  __shared__ float4 ly[kPackTile];
  __shared__ float lq[kPackTile];
  __shared__ float ls[kPackTile];
  __shared__ float4 lt[kPackTile];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  mfma_block<EPS>(src, *count, sg, blk, eps2x2, ly, lq, ls, lt);
}

// MINW: minimum waves per SIMD asked of the register allocator (1: none);
// RS: distances per reciprocal (1, 2 or 4).
template <bool EPS, int MINW = 1, int NACC = 1, int RS = 1>
__global__ __launch_bounds__(256, MINW) void bve_vel_mfma16_kernel(
    Mfma16Src src, const int32_t* __restrict__ count, SegArgs s0, SegArgs s1,
    float eps2x2) {
  __shared__ float4 ly[kPackTile];
  __shared__ float ls[kPackTile];
  __shared__ uint4 la[2 * kPackTile];
  __shared__ float4 lt[kPackTile];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  mfma16_block<EPS, NACC, 1, RS>(src, *count, sg, blk, eps2x2, ly, ls, la,
                                 lt);
}

// Source-split form of bve_vel_mfma16_kernel (block b * KS + k: chunk k of
// target block b, partial sums to sg.part; bve_split_reduce_kernel adds
// them): enough equal work items that the last round of waves is full.
template <bool EPS, int MINW, int NACC, int KS, int RS = 1>
__global__ __launch_bounds__(256, MINW) void bve_vel_mfma16_split_kernel(
    Mfma16Src src, const int32_t* __restrict__ count, SegArgs s0, SegArgs s1,
    float eps2x2) {
  __shared__ float4 ly[kPackTile];
  __shared__ float ls[kPackTile];
  __shared__ uint4 la[2 * kPackTile];
  __shared__ float4 lt[kPackTile];
  const int b = static_cast<int>(blockIdx.x) / KS;
  const int chunk = static_cast<int>(blockIdx.x) % KS;
  const bool first = b < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? b : b - s0.nblocks;
  mfma16_block<EPS, NACC, KS, RS>(src, *count, sg, blk, eps2x2, ly, ls, la,
                                  lt, chunk);
}

// ------------------------------------------------------ stream function ---
// psi_i = sum_j s_j log(1 - x_i.y_j) with the table-driven FP64 log.
// CHECK drops the pair jj == self (tile-relative); only tiles holding some
// lane's own source of the wave need it.  SPECIAL evaluates log's zero /
// negative cases; without it the smallest argument is returned in amin.
template <bool CHECK, bool SPECIAL, int TILE>
__device__ __forceinline__ double psi_tile_f64(const double4* tile, double x0,
                                               double x1, double x2, int self,
                                               const double2* ltab,
                                               double& amin) {
  double psi = 0.0;
  for (int jj = 0; jj < TILE; ++jj) {
    const double4 y = tile[jj];
    const double d = __builtin_fma(
        -x2, y.z, __builtin_fma(-x1, y.y, __builtin_fma(-x0, y.x, 1.0)));
    if (!SPECIAL) amin = fmin(amin, d);
    const double v = y.w * log_f64<SPECIAL>(d, ltab);
    psi += (CHECK && jj == self) ? 0.0 : v;
  }
  return psi;
}

// Redone with log's special cases in the rare waves where some argument
// was <= 0 (coincident points).
template <bool CHECK, int TILE>
__device__ __forceinline__ double psi_tile_sum(const double4* tile, double x0,
                                               double x1, double x2, int self,
                                               const double2* ltab) {
  double amin = 1.0;
  double p = psi_tile_f64<CHECK, false, TILE>(tile, x0, x1, x2, self, ltab, amin);
  if (__ballot(amin <= 0.0))
    p = psi_tile_f64<CHECK, true, TILE>(tile, x0, x1, x2, self, ltab, amin);
  return p;
}

// Block b * ks + k sums source chunk k (of ks equal runs of tiles) for
// target block b: into sg.part[k pstride + c] when split, else psi itself.
template <int BLOCK, int TILE>
__global__ __launch_bounds__(BLOCK) void bve_psi_f64_kernel(
    const double4* __restrict__ src, const int32_t* __restrict__ count,
    SegArgs s0, SegArgs s1, int ks) {
  __shared__ double4 tile[TILE];
  __shared__ double2 ltab[kLogTable];
  const int b = static_cast<int>(blockIdx.x) / ks;
  const int chunk = static_cast<int>(blockIdx.x) % ks;
  const bool first = b < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? b : b - s0.nblocks;
  const int nsrc = *count;
  const int ntile = (nsrc + TILE - 1) / TILE;
  const int jbeg = (ntile * chunk) / ks * TILE;
  const int jend = (ntile * (chunk + 1)) / ks * TILE;
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  if (static_cast<int64_t>(blk) * BLOCK >= n_live) return;
  const int64_t c = static_cast<int64_t>(blk) * BLOCK + threadIdx.x;
  const bool in = c < n_live;
  const int64_t i =
      in ? (sg.index ? sg.index[c] : sg.slot_base + c) : sg.slot_base;
  const bool live = in && !(sg.skip && sg.skip[i]);
  const double* p = sg.x + i * sg.es;
  const double x0 = p[0], x1 = p[sg.cs], x2 = p[2 * sg.cs];
  int self = -1;
  if (in) {
    if (sg.self_rank) {
      const int r = sg.self_rank[i];
      self = r < 0 ? -1 : r - sg.rank_base;
    } else if (sg.self_identity) {
      self = static_cast<int>(c);
    }
  }
  double psi = 0.0;
  fill_log_table(ltab);
  for (int j0 = jbeg; j0 < jend; j0 += TILE) {
    __syncthreads();
    for (int t = threadIdx.x; t < TILE; t += BLOCK) {
      const int j = j0 + t;
      tile[t] = j < nsrc ? src[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    __syncthreads();
    const int so = self < 0 ? -1 : self - j0;
    if (__ballot(static_cast<unsigned>(so) < static_cast<unsigned>(TILE)))
      psi += psi_tile_sum<true, TILE>(tile, x0, x1, x2, so, ltab);
    else
      psi += psi_tile_sum<false, TILE>(tile, x0, x1, x2, so, ltab);
  }
  if (sg.part) {
    if (in) sg.part[chunk * sg.pstride + c] = psi;
  } else if (live) {
    sg.out_scalar[i] = psi;
  }
}

// psi = the ks chunk partials added in chunk order, at every live slot.
__global__ __launch_bounds__(256) void bve_psi_reduce_kernel(SegArgs s0,
                                                             SegArgs s1,
                                                             int ks) {
  const int64_t g = static_cast<int64_t>(blockIdx.x) * 256 + threadIdx.x;
  const bool first = g < s0.n;
  const SegArgs& sg = pick(s0, s1, first);
  const int64_t c = first ? g : g - s0.n;
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  if (c >= n_live) return;
  const int64_t i = sg.index ? sg.index[c] : sg.slot_base + c;
  if (sg.skip && sg.skip[i]) return;
  double psi = 0.0;
  for (int k = 0; k < ks; ++k) psi += sg.part[k * sg.pstride + c];
  sg.out_scalar[i] = psi;
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
using F64Kernel = void (*)(const double4*, const double4*, const int32_t*,
                          SegArgs, SegArgs, double);
using MixedKernel = void (*)(const float4*, const double4*, const int32_t*,
                             SegArgs, SegArgs, float);
using MfmaKernel = void (*)(MfmaSrc, const int32_t*, SegArgs, SegArgs, float);
using Mfma16Kernel = void (*)(Mfma16Src, const int32_t*, SegArgs, SegArgs,
                              float);

struct F64Variant {
  F64Kernel fn, fn_eps;
  int block, tpl;
  const char* name;
  bool sy = false;  // reads the packed s y
  int ksplit = 1;   // > 1: source-split kernel (needs the partial sums
  int twin = -1;    // workspace); twin: the same shape unsplit
};
struct MixedVariant {
  MixedKernel fn, fn_eps;  // vector-ALU kernels (or null)
  MfmaKernel mfn, mfn_eps; // matrix-core kernels (or null)
  int block, tpl;
  const char* name;
  Mfma16Kernel m16, m16_eps;  // f16-split matrix-core kernels (or null)
  int ksplit = 1;  // > 1: source-split kernel (partial sums workspace);
  int twin = -1;   // twin: the same shape unsplit
};

#define LPM_F64_VARIANT(B, T, NR, PAIR)                                  \
  {bve_vel_f64_kernel<B, T, 256, false, NR, PAIR>,                      \
   bve_vel_f64_kernel<B, T, 256, true, NR, PAIR>, B, T,                 \
   "block" #B "_tpl" #T "_nr" #NR "_pair" #PAIR}
#define LPM_F64_FUSED(B, T, NR)                                          \
  {bve_vel_f64_kernel<B, T, 256, false, NR, true, true>,                \
   bve_vel_f64_kernel<B, T, 256, true, NR, true, true>, B, T,           \
   "block" #B "_tpl" #T "_nr" #NR "_pairtrue_fused"}
#define LPM_F64_QUAD(B, T, NR)                                           \
  {bve_vel_f64_kernel<B, T, 256, false, NR, true, true, true>,          \
   bve_vel_f64_kernel<B, T, 256, true, NR, true, true, true>, B, T,     \
   "block" #B "_tpl" #T "_nr" #NR "_quad_fused"}
#define LPM_F64_QUAD_SY(B, T, NR)                                        \
  {bve_vel_f64_kernel<B, T, 256, false, NR, true, true, true, true>,    \
   bve_vel_f64_kernel<B, T, 256, true, NR, true, true, true, true>, B, T, \
   "block" #B "_tpl" #T "_nr" #NR "_quad_fused_sy", true}
#define LPM_F64_QUAD_SY_U(B, T, NR, QU)                                   \
  {bve_vel_f64_kernel<B, T, 256, false, NR, true, true, true, true, QU>,  \
   bve_vel_f64_kernel<B, T, 256, true, NR, true, true, true, true, QU>,   \
   B, T, "block" #B "_tpl" #T "_nr" #NR "_quad_fused_sy_u" #QU, true}
#define LPM_F64_QUAD_SY_TILE(B, T, NR, TL, QU)                                 \
  {bve_vel_f64_kernel<B, T, TL, false, NR, true, true, true, true, QU>,        \
   bve_vel_f64_kernel<B, T, TL, true, NR, true, true, true, true, QU>, B, T,   \
   "block" #B "_tpl" #T "_nr" #NR "_quad_fused_sy_u" #QU "_tile" #TL, true}
// named as the twin (TAG: its tile suffix) + "_split" K
#define LPM_F64_QUAD_SY_SPLIT(B, T, NR, TL, TAG, QU, KS, TWIN)                  \
  {bve_vel_f64_split_kernel<B, T, TL, false, NR, true, true, true, true, QU, KS>, \
   bve_vel_f64_split_kernel<B, T, TL, true, NR, true, true, true, true, QU, KS>,  \
   B, T, "block" #B "_tpl" #T "_nr" #NR "_quad_fused_sy_u" #QU TAG "_split" #KS, \
   true, KS, TWIN}
#define LPM_MIXED_VARIANT(B, T, PAIR)                                     \
  {bve_vel_mixed_kernel<B, T, 256, false, PAIR>,                         \
   bve_vel_mixed_kernel<B, T, 256, true, PAIR>, nullptr, nullptr, B, T,  \
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
    LPM_F64_FUSED(256, 2, 1),           // 8: 1 - x.y as one FMA chain
    LPM_F64_FUSED(256, 4, 1),           // 9
    LPM_F64_FUSED(256, 1, 1),           // 10
    LPM_F64_QUAD(256, 2, 1),            // 11: one reciprocal per 4 pairs
    LPM_F64_QUAD(256, 4, 1),            // 12
    LPM_F64_QUAD_SY(256, 2, 1),         // 13: weights from the packed s y
    LPM_F64_QUAD_SY(256, 4, 1),         // 14
    LPM_F64_QUAD_SY_U(256, 4, 1, 2),    // 15: two groups per iteration
    LPM_F64_QUAD_SY_U(256, 2, 1, 2),    // 16
    LPM_F64_QUAD_SY_U(256, 2, 1, 4),    // 17
    LPM_F64_QUAD_SY_TILE(256, 4, 1, 512, 2),   // 18: 512-source LDS tiles
    LPM_F64_QUAD_SY_TILE(256, 4, 1, 1024, 2),  // 19
    LPM_F64_QUAD_SY_TILE(256, 2, 1, 512, 2),   // 20
    // 21-25: the sources in 8 to 32 chunks per target block, partial
    // sums reduced in a second kernel (steps and init_velocity only; other
    // calls run the twin)
    LPM_F64_QUAD_SY_SPLIT(256, 4, 1, 1024, "_tile1024", 2, 8, 19),
    LPM_F64_QUAD_SY_SPLIT(256, 4, 1, 512, "_tile512", 2, 8, 18),
    LPM_F64_QUAD_SY_SPLIT(256, 4, 1, 512, "_tile512", 2, 16, 18),
    LPM_F64_QUAD_SY_SPLIT(256, 4, 1, 256, "", 2, 8, 15),
    LPM_F64_QUAD_SY_SPLIT(256, 4, 1, 512, "_tile512", 2, 32, 18),  // 25
};
const MixedVariant kMixedVariants[] = {
    LPM_MIXED_VARIANT(256, 2, false),  // 0
    LPM_MIXED_VARIANT(256, 2, true),   // 1
    LPM_MIXED_VARIANT(256, 1, true),   // 2
    LPM_MIXED_VARIANT(256, 4, true),   // 3
    LPM_MIXED_VARIANT(128, 2, true),   // 4
    LPM_MIXED_VARIANT(256, 1, false),  // 5
    LPM_MIXED_VARIANT(256, 4, false),  // 6
    {nullptr, nullptr, bve_vel_mfma_kernel<false>, bve_vel_mfma_kernel<true>,
     256, 1, "mfma_block256_near3"},   // 7: matrix-core far field
    {nullptr, nullptr, nullptr, nullptr, 256, 1, "mfma16_block256_near3",
     bve_vel_mfma16_kernel<false>, bve_vel_mfma16_kernel<true>},  // 8: f16 split
    {nullptr, nullptr, nullptr, nullptr, 256, 1, "mfma16_occ5_block256_near3",
     bve_vel_mfma16_kernel<false, 5>, bve_vel_mfma16_kernel<true, 5>},  // 9
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_acc2_block256_near3", bve_vel_mfma16_kernel<false, 5, 2>,
     bve_vel_mfma16_kernel<true, 5, 2>},                                 // 10
    // 11-13: variant 10 with the sources in 2 to 8 chunks per target
    // block (steps and init_velocity only; other calls run variant 10)
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_acc2_block256_near3_split2",
     bve_vel_mfma16_split_kernel<false, 5, 2, 2>,
     bve_vel_mfma16_split_kernel<true, 5, 2, 2>, 2, 10},                // 11
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_acc2_block256_near3_split4",
     bve_vel_mfma16_split_kernel<false, 5, 2, 4>,
     bve_vel_mfma16_split_kernel<true, 5, 2, 4>, 4, 10},                // 12
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_acc2_block256_near3_split8",
     bve_vel_mfma16_split_kernel<false, 5, 2, 8>,
     bve_vel_mfma16_split_kernel<true, 5, 2, 8>, 8, 10},                // 13
    // 14-18: variants 10 (and its 8-chunk split) and 9 with one reciprocal
    // per two (prcp) or four (qrcp) distances
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_acc2_prcp_block256_near3",
     bve_vel_mfma16_kernel<false, 5, 2, 2>,
     bve_vel_mfma16_kernel<true, 5, 2, 2>},                              // 14
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_acc2_prcp_block256_near3_split8",
     bve_vel_mfma16_split_kernel<false, 5, 2, 8, 2>,
     bve_vel_mfma16_split_kernel<true, 5, 2, 8, 2>, 8, 14},             // 15
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_acc2_qrcp_block256_near3",
     bve_vel_mfma16_kernel<false, 5, 2, 4>,
     bve_vel_mfma16_kernel<true, 5, 2, 4>},                              // 16
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_acc2_qrcp_block256_near3_split8",
     bve_vel_mfma16_split_kernel<false, 5, 2, 8, 4>,
     bve_vel_mfma16_split_kernel<true, 5, 2, 8, 4>, 8, 16},             // 17
    {nullptr, nullptr, nullptr, nullptr, 256, 1,
     "mfma16_occ5_qrcp_block256_near3",
     bve_vel_mfma16_kernel<false, 5, 1, 4>,
     bve_vel_mfma16_kernel<true, 5, 1, 4>},                              // 18
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
  if (family == KID_BVE_ENS_F64 || family == KID_BVE_ENS_MIXED)
    return ens_variant_count(family);
  if (family == KID_SWE_F64 || family == KID_SWE_MIXED)
    return swe_variant_count(family);
  return 0;
}

const char* variant_name(int family, int v) {
  if (v < 0 || v >= variant_count(family)) return nullptr;
  if (family == KID_BVE_ENS_F64 || family == KID_BVE_ENS_MIXED)
    return ens_variant_name(family, v);
  if (family == KID_SWE_F64 || family == KID_SWE_MIXED)
    return swe_variant_name(family, v);
  return family == KID_BVE_VEL_F64 ? kF64Variants[v].name
                                   : kMixedVariants[v].name;
}

int default_variant(int family) {
  if (family == KID_BVE_ENS_F64 || family == KID_BVE_ENS_MIXED)
    return ens_default_variant(family);
  if (family == KID_SWE_F64 || family == KID_SWE_MIXED) return 0;
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
  a.pstride = t.n;
  return a;
}

double count_pairs(const TargetSeg* segs, int nseg, int64_t nsrc_upper) {
  double p = 0;
  for (int s = 0; s < nseg; ++s) p += static_cast<double>(segs[s].n);
  return p * static_cast<double>(nsrc_upper);
}

}  // namespace

// --------------------------------------------------------------- hosts ---
RankMap build_rank_into(const RankWs& ws, const uint8_t* mask, int64_t n,
                        hipStream_t s) {
  LPM_ARG_CHECK(n >= 0 && n < (int64_t(1) << 31), "source count out of range");
  RankMap map;
  map.n = n;
  if (!mask) {
    set_count_kernel<<<1, 1, 0, s>>>(ws.count_ident, static_cast<int32_t>(n));
    LPM_HIP_CHECK(hipGetLastError());
    map.d_count = ws.count_ident;
    return map;
  }
  const int nb = static_cast<int>((n + kScanChunk - 1) / kScanChunk);
  if (nb > 0) {
    mask_count_kernel<<<nb, kScanBlock, 0, s>>>(mask, n, ws.counts);
    LPM_HIP_CHECK(hipGetLastError());
  }
  scan_counts_kernel<<<1, 1024, 0, s>>>(ws.counts, nb);
  LPM_HIP_CHECK(hipGetLastError());
  if (nb > 0) {
    mask_rank_kernel<<<nb, kScanBlock, 0, s>>>(mask, n, ws.counts, ws.rank,
                                               ws.inv);
    LPM_HIP_CHECK(hipGetLastError());
  }
  map.rank = ws.rank;
  map.inv = ws.inv;
  map.d_count = ws.counts + nb;
  map.block_offsets = ws.counts;
  return map;
}

RankMap build_rank(lpm_ctx* ctx, const uint8_t* mask, int64_t n,
                   hipStream_t s) {
  LPM_ARG_CHECK(n >= 0 && n < (int64_t(1) << 31), "source count out of range");
  RankWs ws;
  if (!mask) {
    ws.count_ident =
        static_cast<int32_t*>(ctx->count_ident.ensure(sizeof(int32_t)));
  } else {
    const int nb = static_cast<int>((n + kScanChunk - 1) / kScanChunk);
    ws.counts = static_cast<int32_t*>(
        ctx->block_counts.ensure(sizeof(int32_t) * (nb + 1)));
    const size_t nn = sizeof(int32_t) * std::max<int64_t>(n, 1);
    ws.rank = static_cast<int32_t*>(ctx->rank.ensure(nn));
    ws.inv = static_cast<int32_t*>(ctx->inv.ensure(nn));
  }
  return build_rank_into(ws, mask, n, s);
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

PackedSources pack_sources_into(const PackWs& ws, unsigned layouts, CrdIn sx,
                                const double* zeta, const double* area,
                                const RankMap& map, int precision,
                                hipStream_t s) {
  PackedSources ps;
  const int64_t n = map.n;
  ps.count = map.d_count;
  ps.capacity = n;
  const int64_t cap = std::max<int64_t>(n, 1);
  const bool sy = precision == LPM_FP64 && (layouts & kLayoutSy);
  LPM_ARG_CHECK(ws.p64 && (!sy || ws.psy), "packing buffers missing");
  if (n > 0) {
    pack_f64_kernel<<<static_cast<int>((n + 255) / 256), 256, 0, s>>>(
        sx.base, sx.es, sx.cs, zeta, area, map.inv, map.d_count, ws.p64,
        sy ? ws.psy : nullptr);
    LPM_HIP_CHECK(hipGetLastError());
  }
  ps.data = ws.p64;
  ps.sy = sy ? ws.psy : nullptr;
  if (precision != LPM_MIXED) return ps;
  const int ntile = static_cast<int>((cap + kPackTile - 1) / kPackTile);
  if (layouts & kLayoutVec32) {
    LPM_ARG_CHECK(ws.p32 && ws.centers, "packing buffers missing");
    pack_mixed_kernel<<<ntile, kPackTile, 0, s>>>(ws.p64, map.d_count, ws.p32,
                                                  ws.centers);
    LPM_HIP_CHECK(hipGetLastError());
    ps.data = ws.p32;
    ps.centers = ws.centers;
  }
  if (layouts & (kLayoutMfma | kLayoutMfma16)) {
    const bool full = (layouts & kLayoutMfma) != 0;
    LPM_ARG_CHECK(ws.mf_y && ws.mf_s && ws.mf_c &&
                      (!full || (ws.mf_q && ws.mf_t)),
                  "packing buffers missing");
    pack_mfma_kernel<<<ntile, kPackTile, 0, s>>>(
        ws.p64, map.d_count, ws.mf_y, full ? ws.mf_q : nullptr, ws.mf_s,
        full ? ws.mf_t : nullptr, ws.mf_c);
    LPM_HIP_CHECK(hipGetLastError());
    ps.mf_y = ws.mf_y;
    ps.mf_s = ws.mf_s;
    ps.mf_c = ws.mf_c;
    if (full) {
      ps.mf_q = ws.mf_q;
      ps.mf_t = ws.mf_t;
    }
  }
  if (layouts & kLayoutMfma16) {
    LPM_ARG_CHECK(ws.mf_a && ws.mf_t32 && ws.mf_is, "packing buffers missing");
    pack_mfma16_kernel<<<ntile, kPackTile, 0, s>>>(ws.p64, map.d_count,
                                                   ws.mf_c, ws.mf_a, ws.mf_t32,
                                                   ws.mf_is);
    LPM_HIP_CHECK(hipGetLastError());
    ps.mf_a = ws.mf_a;
    ps.mf_t32 = ws.mf_t32;
    ps.mf_is = ws.mf_is;
  }
  return ps;
}

// Context-owned buffers for every layout, so that any variant (tuned, forced
// or default) finds its sources packed.
PackedSources pack_sources(lpm_ctx* ctx, CrdIn sx, const double* zeta,
                           const double* area, const RankMap& map,
                           int precision, hipStream_t s) {
  const int64_t cap = std::max<int64_t>(map.n, 1);
  PackWs ws;
  ws.p64 = static_cast<double4*>(ctx->packed64.ensure(sizeof(double4) * cap));
  unsigned layouts = 0;
  if (precision == LPM_FP64) {
    ws.psy = static_cast<double4*>(ctx->packed64sy.ensure(sizeof(double4) * cap));
    layouts = kLayoutSy;
  } else if (precision == LPM_MIXED) {
    const int ntile = static_cast<int>((cap + kPackTile - 1) / kPackTile);
    const int64_t tcap = static_cast<int64_t>(ntile) * kPackTile;
    ws.p32 = static_cast<float4*>(ctx->packed32.ensure(sizeof(float4) * cap));
    ws.centers =
        static_cast<double4*>(ctx->centers.ensure(sizeof(double4) * ntile));
    ws.mf_y = static_cast<float4*>(ctx->mf_y.ensure(sizeof(float4) * tcap));
    ws.mf_q = static_cast<float*>(ctx->mf_q.ensure(sizeof(float) * tcap));
    ws.mf_s = static_cast<float*>(ctx->mf_s.ensure(sizeof(float) * tcap));
    ws.mf_t = static_cast<float*>(ctx->mf_t.ensure(sizeof(float) * 4 * tcap));
    ws.mf_c = static_cast<double4*>(ctx->mf_c.ensure(sizeof(double4) * ntile));
    ws.mf_a = static_cast<uint4*>(ctx->mf_a.ensure(sizeof(uint4) * 2 * tcap));
    ws.mf_t32 =
        static_cast<float*>(ctx->mf_t32.ensure(sizeof(float) * 4 * tcap));
    ws.mf_is = static_cast<double*>(ctx->mf_is.ensure(sizeof(double) * ntile));
    layouts = kLayoutVec32 | kLayoutMfma | kLayoutMfma16;
  }
  return pack_sources_into(ws, layouts, sx, zeta, area, map, precision, s);
}

int bve_velocity_variant(lpm_ctx* ctx, int precision, int64_t n_targets) {
  return pick_variant(
      ctx, precision == LPM_MIXED ? KID_BVE_VEL_MIXED : KID_BVE_VEL_F64,
      n_targets);
}

unsigned bve_variant_layouts(int precision, int variant) {
  if (precision == LPM_FP64) {
    LPM_ARG_CHECK(variant >= 0 && variant < kNumF64, "variant out of range");
    return kF64Variants[variant].sy ? kLayoutSy : 0u;
  }
  LPM_ARG_CHECK(variant >= 0 && variant < kNumMixed, "variant out of range");
  const MixedVariant& v = kMixedVariants[variant];
  if (v.m16) return kLayoutMfma16;
  if (v.mfn) return kLayoutMfma;
  return kLayoutVec32;
}

namespace {

// Chunks per target block of a variant (1: not source-split) and its
// unsplit twin.
int split_of(int precision, int variant, int* twin) {
  if (precision == LPM_FP64 && variant >= 0 && variant < kNumF64) {
    *twin = kF64Variants[variant].twin;
    return kF64Variants[variant].ksplit;
  }
  if (precision == LPM_MIXED && variant >= 0 && variant < kNumMixed) {
    *twin = kMixedVariants[variant].twin;
    return kMixedVariants[variant].ksplit;
  }
  *twin = -1;
  return 1;
}

}  // namespace

size_t bve_split_bytes(int precision, int variant, int64_t n_targets) {
  int twin;
  const int ks = split_of(precision, variant, &twin);
  if (ks <= 1) return 0;
  return sizeof(double) * 3 * static_cast<size_t>(ks) *
         static_cast<size_t>(std::max<int64_t>(n_targets, 1));
}

double* split_workspace(lpm_ctx* ctx, size_t need, hipStream_t s) {
  if (ctx->split_part.bytes >= need) return ctx->split_part.as<double>();
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  LPM_HIP_CHECK(hipStreamIsCapturing(s, &cs));
  if (cs != hipStreamCaptureStatusNone) return nullptr;
  try {
    return static_cast<double*>(ctx->split_part.ensure(need));
  } catch (const Error& e) {
    if (e.code != LPM_ERR_NOMEM) throw;
    (void)hipGetLastError();  // the failed hipMalloc's error is handled
    return nullptr;
  }
}

namespace {

void launch_split_reduce(const SegArgs& a0, const SegArgs& a1, int ks,
                         hipStream_t s) {
  const int64_t nt = a0.n + a1.n;
  hipLaunchKernelGGL(bve_split_reduce_kernel,
                     dim3(static_cast<unsigned>((nt + 255) / 256)), dim3(256),
                     0, s, a0, a1, ks);
  LPM_HIP_CHECK(hipGetLastError());
}

}  // namespace

void launch_bve_velocity(lpm_ctx* ctx, const PackedSources& src,
                         TargetSeg* segs, int nseg, double eps2, int precision,
                         hipStream_t s, int variant, bool timed,
                         bool allow_split) {
  LPM_ARG_CHECK(nseg >= 1 && nseg <= 2, "one or two target segments");
  TargetSeg empty;
  const TargetSeg& t0 = segs[0];
  const TargetSeg& t1 = nseg > 1 ? segs[1] : empty;
  for (int k = 0; k < nseg; ++k)
    LPM_ARG_CHECK(segs[k].n >= 0 && segs[k].n < (int64_t(1) << 31),
                  "target count out of range");
  if (t0.n + t1.n == 0) return;
  const double pairs = count_pairs(segs, nseg, src.capacity);
  LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                "unknown precision");
  if (variant < 0) variant = bve_velocity_variant(ctx, precision, t0.n + t1.n);
  LPM_ARG_CHECK(variant >= 0 && variant < (precision == LPM_FP64 ? kNumF64
                                                                  : kNumMixed),
                "variant out of range");
  double* part = nullptr;
  int twin = -1;
  if (split_of(precision, variant, &twin) > 1) {
    // the partial sums live in a context workspace: a step's prepare sizes
    // it before capture; a call that may not allocate it (no permission,
    // capture with it too small, or no memory for it) runs the unsplit twin
    if (allow_split)
      part = split_workspace(
          ctx, bve_split_bytes(precision, variant, t0.n + t1.n), s);
    if (!part) variant = twin;
  }
  if (precision == LPM_FP64) {
    const F64Variant& v = kF64Variants[variant];
    const int per = v.block * v.tpl;
    SegArgs a0 = to_args(t0, per), a1 = to_args(t1, per);
    const int grid = a0.nblocks + a1.nblocks;
    LPM_ARG_CHECK(!v.sy || src.sy, "packed s y missing for the FP64 variant");
    LPM_ARG_CHECK(v.ksplit == 1 || part, "split workspace missing");
    TimedLaunch tl(ctx, timed ? KID_BVE_VEL_F64 : -1, s, pairs);
    if (part) {
      a0.part = part;
      a1.part = part + 3 * static_cast<int64_t>(v.ksplit) * t0.n;
    }
    hipLaunchKernelGGL(eps2 > 0 ? v.fn_eps : v.fn,
                       dim3(static_cast<unsigned>(grid) * v.ksplit),
                       dim3(v.block), 0, s,
                       static_cast<const double4*>(src.data), src.sy, src.count,
                       a0, a1, eps2);
    LPM_HIP_CHECK(hipGetLastError());
    if (part) launch_split_reduce(a0, a1, v.ksplit, s);
  } else if (precision == LPM_MIXED) {
    const MixedVariant& v = kMixedVariants[variant];
    const int per = v.block * v.tpl;
    SegArgs a0 = to_args(t0, per), a1 = to_args(t1, per);
    const int grid = a0.nblocks + a1.nblocks;
    TimedLaunch tl(ctx, timed ? KID_BVE_VEL_MIXED : -1, s, pairs);
    if (v.m16) {
      LPM_ARG_CHECK(src.mf_a, "matrix-core layout not packed");
      LPM_ARG_CHECK(v.ksplit == 1 || part, "split workspace missing");
      Mfma16Src m{src.mf_y, src.mf_s, src.mf_a, src.mf_t32, src.mf_c, src.mf_is};
      if (part) {
        a0.part = part;
        a1.part = part + 3 * static_cast<int64_t>(v.ksplit) * t0.n;
      }
      hipLaunchKernelGGL(eps2 > 0 ? v.m16_eps : v.m16,
                         dim3(static_cast<unsigned>(grid) * v.ksplit),
                         dim3(v.block), 0, s, m, src.count, a0, a1,
                         static_cast<float>(2.0 * eps2));
      LPM_HIP_CHECK(hipGetLastError());
      if (part) launch_split_reduce(a0, a1, v.ksplit, s);
    } else if (v.mfn) {
      LPM_ARG_CHECK(src.mf_y, "matrix-core layout not packed");
      MfmaSrc m{src.mf_y, src.mf_q, src.mf_s, src.mf_t, src.mf_c};
      hipLaunchKernelGGL(eps2 > 0 ? v.mfn_eps : v.mfn, dim3(grid),
                         dim3(v.block), 0, s, m, src.count, a0, a1,
                         static_cast<float>(2.0 * eps2));
    } else {
      LPM_ARG_CHECK(src.centers, "mixed vector layout not packed");
      hipLaunchKernelGGL(eps2 > 0 ? v.fn_eps : v.fn, dim3(grid), dim3(v.block),
                         0, s, static_cast<const float4*>(src.data),
                         src.centers, src.count, a0, a1,
                         static_cast<float>(2.0 * eps2));
    }
    LPM_HIP_CHECK(hipGetLastError());
  } else {
    throw Error(LPM_ERR_ARG, "unknown precision");
  }
}

int split_chunks(lpm_ctx* ctx, const void* kernel, int block, int64_t blocks,
                 int64_t nsrc_upper, int tile) {
  int per_cu = 0;
  LPM_HIP_CHECK(
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
  const int64_t want = 8 * static_cast<int64_t>(std::max(per_cu, 1)) *
                       std::max(ctx->num_cus, 1);
  const int64_t ntile = (nsrc_upper + tile - 1) / tile;
  if (blocks <= 0 || blocks >= want || ntile < 16) return 1;
  const int64_t ks = std::min((want + blocks - 1) / blocks, ntile / 8);
  return static_cast<int>(std::max<int64_t>(ks, 1));
}

void launch_bve_stream_fn(lpm_ctx* ctx, const PackedSources& src,
                          TargetSeg* segs, int nseg, hipStream_t s) {
  TargetSeg empty;
  const TargetSeg& t0 = segs[0];
  const TargetSeg& t1 = nseg > 1 ? segs[1] : empty;
  if (t0.n + t1.n == 0) return;
  constexpr int kBlock = 256, kTile = 256;
  SegArgs a0 = to_args(t0, kBlock), a1 = to_args(t1, kBlock);
  const int blocks = a0.nblocks + a1.nblocks;
  int ks = split_chunks(
      ctx, reinterpret_cast<const void*>(&bve_psi_f64_kernel<kBlock, kTile>),
      kBlock, blocks, src.capacity, kTile);
  if (ks > 1) {
    // no workspace (a caller's capture, no memory): one chunk
    double* part = split_workspace(
        ctx, sizeof(double) * static_cast<size_t>(ks) * (t0.n + t1.n), s);
    if (part) {
      a0.part = part;
      a1.part = part + static_cast<int64_t>(ks) * t0.n;
    } else {
      ks = 1;
    }
  }
  TimedLaunch tl(ctx, KID_BVE_PSI, s, count_pairs(segs, nseg, src.capacity));
  hipLaunchKernelGGL((bve_psi_f64_kernel<kBlock, kTile>),
                     dim3(static_cast<unsigned>(blocks) * ks), dim3(kBlock), 0,
                     s, static_cast<const double4*>(src.data), src.count, a0,
                     a1, ks);
  LPM_HIP_CHECK(hipGetLastError());
  if (ks > 1) {
    const int64_t nt = t0.n + t1.n;
    hipLaunchKernelGGL(bve_psi_reduce_kernel,
                       dim3(static_cast<unsigned>((nt + 255) / 256)),
                       dim3(256), 0, s, a0, a1, ks);
    LPM_HIP_CHECK(hipGetLastError());
  }
}

void check_finite(lpm_ctx* ctx, const double* base, int64_t n, int64_t es,
                  int64_t cs, int ncomp, hipStream_t s) {
  if (n == 0) return;
  check_finite_into(static_cast<int32_t*>(ctx->flag.ensure(sizeof(int32_t))),
                    base, n, es, cs, ncomp, s);
}

void check_finite_into(int32_t* flag, const double* base, int64_t n,
                       int64_t es, int64_t cs, int ncomp, hipStream_t s) {
  if (n == 0) return;
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
