// Device building blocks of the BVE direct sums, shared by the one-problem
// kernels (bve_sums.hip) and the ensemble kernels (bve_ensemble.hip): target
// segment arguments, the pair bodies, and the per-block loops over LDS tiles
// of packed sources.  The loops replace the per-target team reductions of
// BVEVertexVelocity / BVEFaceVelocity (src/lpm_bve_sphere_kernels.hpp:
// 53-82, 179-211, 249-274, 365-394).
//
// Packed source formats:
//   FP64:  double4 {y0, y1, y2, s},  s = -zeta A / (4 pi)
//   mixed: float4  {y0-c0, y1-c1, y2-c2, 2 s} relative to the FP64 centre c
//          of each 256-source packing tile, plus double4 centres.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace lpm_amd {
namespace {

constexpr double kInv4Pi = 0.079577471545947667884441881686257181;
constexpr int kPackTile = 256;  // mixed-precision packing tile (sources)

// One segment of targets.  Slot c of the segment is target
//   i = index ? index[c] : slot_base + c
// of the arrays x / out.  A target's own source (excluded from its sum) is
//   self_rank ? self_rank[i] - rank_base  (negative self_rank: none)
//   : self_identity ? c : none,
// counted from the first packed source the kernel is given.
struct SegArgs {
  const double* x;
  int64_t es, cs, n;
  const int32_t* index;
  const int32_t* d_n;  // device-side live slot count (overrides n)
  const int32_t* self_rank;
  int self_identity;
  int nblocks;
  const uint8_t* skip;
  double* out;
  int64_t oes, ocs;
  double* out_scalar;
  int64_t slot_base;
  int32_t rank_base;
  // source-split launches: chunk k of slot c adds its three sums at
  // part[(3 k + comp) * pstride + c]; a reduction kernel forms x cross S
  double* part;
  int64_t pstride;
};

__device__ __forceinline__ const SegArgs& pick(const SegArgs& a,
                                               const SegArgs& b, bool first) {
  return first ? a : b;
}

// --------------------------------------------------------- pair bodies ---
// Reciprocal with NR Newton-Raphson refinements (r <- r + r (1 - d r)) of
// the hardware estimate.
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

// FP64 natural logarithm, table-driven after P. T. P. Tang, "Table-driven
// implementation of the logarithm function in IEEE floating-point
// arithmetic", ACM TOMS 16(4), 1990.  With a = m 2^e, m in [1/2, 1) and j
// the seven fraction bits of m below its leading one, c_j = 2 / (1 + (j +
// 1/2) / 128) brings m near 1:
//   r = m c_j - 1 (one FMA, |r| < 2^-7),
//   log a = k ln 2 + t_j + log1p(r),  t_j = -log(c_j) + (e - k) ln 2,
// log1p(r) by its Taylor series to r^8 (truncation below 2^-56 |r|).
// k = e - 1 for m < 3/4 (j < 64) and k = e above, so that a just below a
// power of two does not cancel k ln 2 against t_j; the end intervals use
// c_0 = 2 and c_127 = 1 (t = 0), exact on either side of 1.  ln 2 =
// kLn2Hi + kLn2Lo with kLn2Hi carrying 42 significant bits, so k kLn2Hi
// is exact for |k| < 2^11.  Error within 1.3 ulp against an extended-
// precision log over the double range; a <= 0, inf and NaN give log's
// special values.  The table lives in LDS (2 KiB per block).
constexpr int kLogTable = 128;
constexpr double kLn2Hi = 0x1.62e42fefa3800p-1;
constexpr double kLn2Lo = 0x1.ef35793c76730p-45;

// Fills tab[j] = {c_j, t_j} with the block's threads; the caller
// synchronises before the first log_f64.
__device__ __forceinline__ void fill_log_table(double2* tab) {
  for (int j = threadIdx.x; j < kLogTable; j += blockDim.x) {
    double c = 2.0 / (1.0 + (j + 0.5) / kLogTable);
    if (j == 0) c = 2.0;
    if (j == kLogTable - 1) c = 1.0;
    tab[j] = make_double2(c, -log(j < kLogTable / 2 ? 0.5 * c : c));
  }
}

// SPECIAL = false leaves out the zero / negative / infinite cases (their
// results are then finite garbage) for callers that screen arguments
// themselves; NaN still propagates.
template <bool SPECIAL = true>
__device__ __forceinline__ double log_f64(double a, const double2* tab) {
  const double m = __builtin_amdgcn_frexp_mant(a);
  const int e = __builtin_amdgcn_frexp_exp(a);
  const int j = (__double2hiint(m) >> 13) & (kLogTable - 1);
  const double2 ct = tab[j];
  const double k = static_cast<double>(e - 1 + (j >> 6));
  const double r = __builtin_fma(m, ct.x, -1.0);
  double q = -1.0 / 8.0;
  q = __builtin_fma(q, r, 1.0 / 7.0);
  q = __builtin_fma(q, r, -1.0 / 6.0);
  q = __builtin_fma(q, r, 1.0 / 5.0);
  q = __builtin_fma(q, r, -1.0 / 4.0);
  q = __builtin_fma(q, r, 1.0 / 3.0);
  q = __builtin_fma(q, r, -1.0 / 2.0);
  const double hi = __builtin_fma(k, kLn2Hi, ct.y);
  const double lo = __builtin_fma(r * r, q, k * kLn2Lo);
  double v = hi + (r + lo);
  if (SPECIAL && !(a > 0.0 && a < __builtin_inf()))
    v = a == 0.0 ? -__builtin_inf()
                 : (a == __builtin_inf() ? a : __builtin_nan(""));
  return v;
}

// One source: S += s / (1 - x.y [+ eps^2]) * y.
// FUSE: 1 - x.y as one FMA chain from 1 (three operations instead of four;
// the same O(1e-16) absolute rounding as the dot-then-subtract form).
template <bool FUSE>
__device__ __forceinline__ double one_minus_dot(const double* x,
                                                const double4& y) {
  if (FUSE)
    return __builtin_fma(-x[2], y.z,
                         __builtin_fma(-x[1], y.y, __builtin_fma(-x[0], y.x, 1.0)));
  const double t = x[0] * y.x + x[1] * y.y + x[2] * y.z;
  return 1.0 - t;
}

// S += (s / d) y for one source from its reciprocal distance inv_d = 1 / d.
// SY: the source's s y is packed (sy), so that the weight needs no multiply;
// otherwise w = s inv_d scales y.
template <bool SY>
__device__ __forceinline__ void acc_w(double* S, double inv_d, const double4& y,
                                      const double4& sy) {
  if (SY) {
    S[0] = __builtin_fma(inv_d, sy.x, S[0]);
    S[1] = __builtin_fma(inv_d, sy.y, S[1]);
    S[2] = __builtin_fma(inv_d, sy.z, S[2]);
  } else {
    const double w = y.w * inv_d;
    S[0] = __builtin_fma(w, y.x, S[0]);
    S[1] = __builtin_fma(w, y.y, S[1]);
    S[2] = __builtin_fma(w, y.z, S[2]);
  }
}

template <bool EPS, int NR, bool FUSE = false, bool SY = false>
__device__ __forceinline__ void f64_one(const double* x, const double4& y,
                                        const double4& sy, double eps2,
                                        double* S) {
  double d = one_minus_dot<FUSE>(x, y);
  if (EPS) d += eps2;
  acc_w<SY>(S, rcp_nr<NR>(d), y, sy);
}

// Two sources sharing one reciprocal: r = 1/(da db), 1/da = db r,
// 1/db = da r.  Terms are accumulated in source order.
template <bool EPS, int NR, bool FUSE = false, bool SY = false>
__device__ __forceinline__ void f64_two(const double* x, const double4& ya,
                                        const double4& yb, const double4& sya,
                                        const double4& syb, double eps2,
                                        double* S) {
  double da = one_minus_dot<FUSE>(x, ya), db = one_minus_dot<FUSE>(x, yb);
  if (EPS) {
    da += eps2;
    db += eps2;
  }
  const double r = rcp_nr<NR>(da * db);
  acc_w<SY>(S, db * r, ya, sya);
  acc_w<SY>(S, da * r, yb, syb);
}

// Four sources sharing one reciprocal: r = 1/(da db dc dd),
// 1/da = db (r dc dd), 1/db = da (r dc dd), and so on; terms in source order.
template <bool EPS, int NR, bool FUSE, bool SY = false>
__device__ __forceinline__ void f64_four(const double* x, const double4& ya,
                                         const double4& yb, const double4& yc,
                                         const double4& yd, const double4& sya,
                                         const double4& syb, const double4& syc,
                                         const double4& syd, double eps2,
                                         double* S) {
  double da = one_minus_dot<FUSE>(x, ya), db = one_minus_dot<FUSE>(x, yb);
  double dc = one_minus_dot<FUSE>(x, yc), dd = one_minus_dot<FUSE>(x, yd);
  if (EPS) {
    da += eps2;
    db += eps2;
    dc += eps2;
    dd += eps2;
  }
  const double pab = da * db, pcd = dc * dd;
  const double r = rcp_nr<NR>(pab * pcd);
  const double rab = r * pcd, rcd = r * pab;
  acc_w<SY>(S, db * rab, ya, sya);
  acc_w<SY>(S, da * rab, yb, syb);
  acc_w<SY>(S, dd * rcd, yc, syc);
  acc_w<SY>(S, dc * rcd, yd, syd);
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
// is idle (block-uniform).
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
    const int64_t i = in ? (sg.index ? sg.index[c] : sg.slot_base + c)
                         : sg.slot_base;
    tgt[k] = i;
    live[k] = in && !(sg.skip && sg.skip[i]);
    const double* p = sg.x + i * sg.es;
    x[k][0] = p[0];
    x[k][1] = p[sg.cs];
    x[k][2] = p[2 * sg.cs];
    int sf = -1;
    if (in) {
      if (sg.self_rank) {
        const int r = sg.self_rank[i];
        sf = r < 0 ? -1 : r - sg.rank_base;
      } else if (sg.self_identity) {
        sf = static_cast<int>(c);
      }
    }
    self[k] = sf;
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

// Source-split launches: the sums S of chunk `chunk` at slot c go to
// sg.part[(3 chunk + comp) pstride + c] for every slot below the live
// count; split_reduce_one adds the chunks and forms x cross S.
template <int BLOCK, int TPL>
__device__ __forceinline__ void store_partial(const SegArgs& sg, int blk,
                                              int chunk,
                                              const double (*S)[3]) {
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  const int64_t slot0 = static_cast<int64_t>(blk) * (BLOCK * TPL);
#pragma unroll
  for (int k = 0; k < TPL; ++k) {
    const int64_t c = slot0 + k * BLOCK + threadIdx.x;
    if (c >= n_live) continue;
#pragma unroll
    for (int comp = 0; comp < 3; ++comp)
      sg.part[(3 * static_cast<int64_t>(chunk) + comp) * sg.pstride + c] =
          S[k][comp];
  }
}

// One target of a source-split sum: S = the ks chunk partials (chunk k's
// component comp at part[(nper k + comp) pstride], nper values per chunk)
// added in chunk order, and o = x cross S (x at p, its components cs apart;
// o's ocs apart).  Every split reduction goes through this function, so
// equal partials give equal bits.
__device__ __forceinline__ void split_reduce_one(const double* part,
                                                 int64_t pstride, int ks,
                                                 const double* p, int64_t cs,
                                                 double* o, int64_t ocs,
                                                 int nper = 3) {
  double S[3] = {0.0, 0.0, 0.0};
  for (int k = 0; k < ks; ++k)
#pragma unroll
    for (int comp = 0; comp < 3; ++comp)
      S[comp] += part[(static_cast<int64_t>(nper) * k + comp) * pstride];
  const double x0 = p[0], x1 = p[cs], x2 = p[2 * cs];
  o[0] = x1 * S[2] - x2 * S[1];
  o[ocs] = x2 * S[0] - x0 * S[2];
  o[2 * ocs] = x0 * S[1] - x1 * S[0];
}

// The packed s y of source j (SY), or any source record the caller passes
// where acc_w does not read it.
template <bool SY>
__device__ __forceinline__ const double4& sy_of(const double4* tsy, int j,
                                                const double4& y) {
  if constexpr (SY) return tsy[j];
  else return y;
}

// ---------------------------------------------------------- FP64 block ---
// Block `blk` of segment `sg` against nsrc packed FP64 sources.  Every
// thread of the block must call it (it synchronises); `tile` is the
// block's LDS staging array of TILE sources (and `tile_sy` of their s y
// when SY).  PAIR shares one reciprocal between two sources, QUAD between
// four; QU unrolls the four-source loop (the next group's LDS reads then
// overlap the current group's arithmetic).  Each tile is summed into its own partial T, then added to S: two
// levels of summation keep the rounding of S (whose component along x
// cancels in x cross S) from growing with every source.  KS > 1: the block
// sums only chunk `chunk` of KS equal runs of source tiles and writes S to
// sg.part (a third summation level, in chunk order, in the reduction).
template <int BLOCK, int TPL, int TILE, bool EPS, int NR, bool PAIR,
          bool FUSE = false, bool QUAD = false, bool SY = false, int QU = 1,
          int KS = 1>
__device__ __forceinline__ void f64_block(const double4* __restrict__ src,
                                          const double4* __restrict__ src_sy,
                                          int nsrc, const SegArgs& sg, int blk,
                                          double eps2, double4* tile,
                                          double4* tile_sy, int chunk = 0) {
  static_assert(TILE % 8 == 0, "tile must be a multiple of the unroll");
  double x[TPL][3], S[TPL][3];
  int self[TPL];
  bool live[TPL];
  int64_t tgt[TPL];
  if (!load_targets<BLOCK, TPL>(sg, blk, x, self, live, tgt)) return;
#pragma unroll
  for (int k = 0; k < TPL; ++k) S[k][0] = S[k][1] = S[k][2] = 0.0;
  const int ntile = (nsrc + TILE - 1) / TILE;
  const int jbeg = KS > 1 ? (ntile * chunk) / KS * TILE : 0;
  const int jend = KS > 1 ? (ntile * (chunk + 1)) / KS * TILE : nsrc;
  for (int j0 = jbeg; j0 < jend; j0 += TILE) {
    double T[TPL][3];
#pragma unroll
    for (int k = 0; k < TPL; ++k) T[k][0] = T[k][1] = T[k][2] = 0.0;
    __syncthreads();
    for (int t = threadIdx.x; t < TILE; t += BLOCK) {
      const int j = j0 + t;
      // padding: zero strength at the origin -> 1 - x.y = 1, w = 0
      tile[t] = j < nsrc ? src[j] : make_double4(0.0, 0.0, 0.0, 0.0);
      if (SY)
        tile_sy[t] = j < nsrc ? src_sy[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    __syncthreads();
    const double4* tsy = SY ? tile_sy : tile;  // unread without SY
    bool diag = false;
#pragma unroll
    for (int k = 0; k < TPL; ++k)
      diag |= static_cast<unsigned>(self[k] - j0) < static_cast<unsigned>(TILE);
    if (__ballot(diag)) {
      // tile holds some target's own source: per-pair exclusion
      for (int jj = 0; jj < TILE; ++jj) {
        const double4 y = tile[jj];
        const double4& sy = sy_of<SY>(tsy, jj, y);
#pragma unroll
        for (int k = 0; k < TPL; ++k) {
          const double t = x[k][0] * y.x + x[k][1] * y.y + x[k][2] * y.z;
          double d = 1.0 - t;
          if (EPS) d += eps2;
          double r = rcp_nr<2>(d);
          r = (j0 + jj == self[k]) ? 0.0 : r;
          acc_w<SY>(T[k], r, y, sy);
        }
      }
    } else if (QUAD) {
#pragma unroll QU
      for (int jj = 0; jj < TILE; jj += 4) {
        const double4 ya = tile[jj], yb = tile[jj + 1];
        const double4 yc = tile[jj + 2], yd = tile[jj + 3];
        const double4& sya = sy_of<SY>(tsy, jj, ya);
        const double4& syb = sy_of<SY>(tsy, jj + 1, yb);
        const double4& syc = sy_of<SY>(tsy, jj + 2, yc);
        const double4& syd = sy_of<SY>(tsy, jj + 3, yd);
#pragma unroll
        for (int k = 0; k < TPL; ++k)
          f64_four<EPS, NR, FUSE, SY>(x[k], ya, yb, yc, yd, sya, syb, syc, syd,
                                      eps2, T[k]);
      }
    } else if (PAIR) {
#pragma unroll 2
      for (int jj = 0; jj < TILE; jj += 2) {
        const double4 ya = tile[jj];
        const double4 yb = tile[jj + 1];
        const double4& sya = sy_of<SY>(tsy, jj, ya);
        const double4& syb = sy_of<SY>(tsy, jj + 1, yb);
#pragma unroll
        for (int k = 0; k < TPL; ++k)
          f64_two<EPS, NR, FUSE, SY>(x[k], ya, yb, sya, syb, eps2, T[k]);
      }
    } else {
#pragma unroll 4
      for (int jj = 0; jj < TILE; ++jj) {
        const double4 y = tile[jj];
        const double4& sy = sy_of<SY>(tsy, jj, y);
#pragma unroll
        for (int k = 0; k < TPL; ++k)
          f64_one<EPS, NR, FUSE, SY>(x[k], y, sy, eps2, T[k]);
      }
    }
#pragma unroll
    for (int k = 0; k < TPL; ++k) {
      S[k][0] += T[k][0];
      S[k][1] += T[k][1];
      S[k][2] += T[k][2];
    }
  }
  if constexpr (KS > 1)
    store_partial<BLOCK, TPL>(sg, blk, chunk, S);
  else
    store_cross<TPL>(sg, x, S, live, tgt);
}

// --------------------------------------------------------- mixed block ---
// As f64_block over FP32 tile-relative sources.  TILE is a multiple of the
// packing tile so that one LDS tile spans whole packing tiles; the target
// offset x - c is formed in FP64 once per packing tile and rounded to FP32.
// Per packing tile the FP32 partial is carried into FP64 accumulators.
// centers[t] is the centre of sources [256 t, 256 t + 256).
template <int BLOCK, int TPL, int TILE, bool EPS, bool PAIR>
__device__ __forceinline__ void mixed_block(const float4* __restrict__ src,
                                            const double4* __restrict__ centers,
                                            int nsrc, const SegArgs& sg,
                                            int blk, float eps2x2,
                                            float4* tile) {
  static_assert(TILE % kPackTile == 0, "tile must cover whole packing tiles");
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

// One block per packing tile of the FP64 packed array (*count entries):
// FP64 centre = first source of the tile, FP32 offsets and 2 s.
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


// ------------------------------------------------ matrix-core mixed path ---
// Far-field pairs on the matrix cores, per 256-source tile with FP64 centre
// c (centroid) and radius R, in coordinates relative to c rounded to FP32:
//   GEMM1 (v_mfma_f32_16x16x4_f32, 16 sources x 16 targets x K = 4):
//     r2 = |x'|^2 - 2 x'.y' + |y'|^2  with A_j = (y', 1),
//     B_i = (-2 x', |x'|^2), C_ji = |y'_j|^2;
//   VALU: 1 / r2;
//   GEMM2 (v_mfma_f32_4x4x1_16b_f32): sum_j (1 / r2) 2 s_j (y', 1) per
//     target (the weight 2 s is folded into the packed operand), i.e. the
//     tile's sum of w y' and of w = 2 s / r2, carried into FP64 as
//     S += S' + c W.
// A wave (64 targets) uses the matrix path for a tile only when none of its
// targets lies within kMfmaNear R of c: then r2 >= (kMfmaNear - 1)^2 R^2
// while its FP32 terms are at most (kMfmaNear + 1)^2 R^2, which bounds the
// relative rounding of r2 by about 12 ulp for kMfmaNear = 3.  Tiles near
// some target of the wave (including every tile holding a target's own
// source) take the difference form d = y' - x' on the vector ALUs, as the
// mixed block does.
//
// Packed matrix-core layout (arrays of ntile * 256 entries, padded tiles):
//   mf_y[j] = (y'_0, y'_1, y'_2, 1)       GEMM1 A and near-path positions
//   mf_q[j] = |y'|^2                       GEMM1 C
//   mf_s[j] = 2 s                          weights (0 in padding)
//   mf_t[t 1024 + ((ch 4 + g) 4 + c) 4 + r] = 2 s times component c of
//            source 256 t + 16 ch + 4 g + r (c = 3: 2 s), the GEMM2 A
//            operand
//   mf_c[t] = (c, R^2)
// Padding sources sit at y' = (8, 8, 8) with zero weight.
constexpr double kMfmaNear = 3.0;
constexpr float kMfmaPad = 8.0f;

typedef float mf4 __attribute__((ext_vector_type(4)));

struct MfmaSrc {
  const float4* y;
  const float* q;
  const float* s;
  const float* t;
  const double4* c;
};

// One block per 256-source tile of the FP64 packed array (count entries):
// centroid, radius and the matrix-core layout.  A source whose FP64 record
// is all zero is padding (real sources lie on the unit sphere).
__global__ __launch_bounds__(kPackTile) void pack_mfma_kernel(
    const double4* __restrict__ p64, const int32_t* __restrict__ count,
    float4* __restrict__ mf_y, float* __restrict__ mf_q,
    float* __restrict__ mf_s, float* __restrict__ mf_t,
    double4* __restrict__ mf_c) {
  __shared__ double red[4][kPackTile / 64];
  __shared__ double4 cen;
  const int n = *count;
  const int j0 = blockIdx.x * kPackTile;
  if (j0 >= n) return;
  const int j = j0 + threadIdx.x;
  double4 y = make_double4(0.0, 0.0, 0.0, 0.0);
  if (j < n) y = p64[j];
  const bool real = (j < n) && !(y.x == 0.0 && y.y == 0.0 && y.z == 0.0 &&
                                 y.w == 0.0);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  // centroid of the real sources
  double v[4] = {real ? y.x : 0.0, real ? y.y : 0.0, real ? y.z : 0.0,
                 real ? 1.0 : 0.0};
  for (int k = 0; k < 4; ++k) {
    double a = v[k];
    for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
    if (lane == 0) red[k][wv] = a;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum[4] = {0, 0, 0, 0};
    for (int k = 0; k < 4; ++k)
      for (int w = 0; w < kPackTile / 64; ++w) sum[k] += red[k][w];
    const double m = sum[3] > 0 ? sum[3] : 1.0;
    cen = make_double4(sum[0] / m, sum[1] / m, sum[2] / m, 0.0);
  }
  __syncthreads();
  const double4 c = cen;
  const double e0 = y.x - c.x, e1 = y.y - c.y, e2 = y.z - c.z;
  double r2 = real ? e0 * e0 + e1 * e1 + e2 * e2 : 0.0;
  for (int off = 32; off > 0; off >>= 1) r2 = fmax(r2, __shfl_down(r2, off, 64));
  __syncthreads();
  if (lane == 0) red[0][wv] = r2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kPackTile / 64; ++w) m = fmax(m, red[0][w]);
    mf_c[blockIdx.x] = make_double4(c.x, c.y, c.z, m);
  }
  float y0 = kMfmaPad, y1 = kMfmaPad, y2 = kMfmaPad, ws = 0.0f;
  if (real) {
    y0 = static_cast<float>(e0);
    y1 = static_cast<float>(e1);
    y2 = static_cast<float>(e2);
    ws = static_cast<float>(2.0 * y.w);
  }
  mf_y[j] = make_float4(y0, y1, y2, 1.0f);
  mf_s[j] = ws;
  // q and the GEMM2 operand are read by the f32 matrix-core kernel only
  // (null: not packed)
  if (mf_q) mf_q[j] = y0 * y0 + y1 * y1 + y2 * y2;
  if (mf_t) {
    // transposed GEMM2 operand: local index jj = 16 ch + 4 g + r
    const int jj = threadIdx.x, ch = jj >> 4, g = (jj >> 2) & 3, r = jj & 3;
    float* tt = mf_t + static_cast<int64_t>(blockIdx.x) * 1024 + (ch * 4 + g) * 16 + r;
    tt[0] = ws * y0;
    tt[4] = ws * y1;
    tt[8] = ws * y2;
    tt[12] = ws;
  }
}

// Block `blk` (256 targets, one per lane) of segment sg against nsrc
// sources in the matrix-core layout.  Every thread of the block calls it.
template <bool EPS>
__device__ __forceinline__ void mfma_block(const MfmaSrc& src, int nsrc,
                                           const SegArgs& sg, int blk,
                                           float eps2x2, float4* ly, float* lq,
                                           float* ls, float4* lt) {
  double x[1][3];
  int self[1];
  bool live[1];
  int64_t tgt[1];
  if (!load_targets<256, 1>(sg, blk, x, self, live, tgt)) return;
  const int lane = threadIdx.x & 63;
  const int k = lane >> 4;   // MFMA K row of this lane
  const int col = lane & 15; // MFMA target column
  // coordinates of the 4 x 16 targets of this wave's matrix columns
  double tx[4][3];
#pragma unroll
  for (int tb = 0; tb < 4; ++tb)
#pragma unroll
    for (int a = 0; a < 3; ++a) tx[tb][a] = __shfl(x[0][a], 16 * tb + col, 64);
  double Sf[4][3];  // far-field sums per (target column, source group)
  double Sn[3] = {0.0, 0.0, 0.0};  // near-field sums of this lane's target
#pragma unroll
  for (int tb = 0; tb < 4; ++tb) Sf[tb][0] = Sf[tb][1] = Sf[tb][2] = 0.0;

  for (int j0 = 0; j0 < nsrc; j0 += kPackTile) {
    const int t = j0 / kPackTile;
    __syncthreads();
    {
      const int64_t base = static_cast<int64_t>(t) * kPackTile + threadIdx.x;
      ly[threadIdx.x] = src.y[base];
      lq[threadIdx.x] = src.q[base];
      ls[threadIdx.x] = src.s[base];
      lt[threadIdx.x] = reinterpret_cast<const float4*>(src.t)[base];
    }
    __syncthreads();
    const double4 c = src.c[t];
    const double e0 = x[0][0] - c.x, e1 = x[0][1] - c.y, e2 = x[0][2] - c.z;
    const bool near =
        e0 * e0 + e1 * e1 + e2 * e2 <= kMfmaNear * kMfmaNear * c.w;
    if (__ballot(near)) {
      // difference form on the vector ALUs, one target per lane
      const float xr0 = static_cast<float>(e0), xr1 = static_cast<float>(e1),
                  xr2 = static_cast<float>(e2);
      const bool diag = static_cast<unsigned>(self[0] - j0) <
                        static_cast<unsigned>(kPackTile);
      float T0 = 0.0f, T1 = 0.0f, T2 = 0.0f;
      for (int jj = 0; jj < kPackTile; ++jj) {
        const float4 y = ly[jj];
        const float d0 = y.x - xr0, d1 = y.y - xr1, d2 = y.z - xr2;
        float r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (EPS) r2 += eps2x2;
        float w = ls[jj] * __builtin_amdgcn_rcpf(r2);
        if (diag && j0 + jj == self[0]) w = 0.0f;
        T0 = __builtin_fmaf(w, d0, T0);
        T1 = __builtin_fmaf(w, d1, T1);
        T2 = __builtin_fmaf(w, d2, T2);
      }
      Sn[0] += static_cast<double>(T0);
      Sn[1] += static_cast<double>(T1);
      Sn[2] += static_cast<double>(T2);
      continue;
    }
    // matrix-core far field
    float B[4];
#pragma unroll
    for (int tb = 0; tb < 4; ++tb) {
      const float a0 = static_cast<float>(tx[tb][0] - c.x);
      const float a1 = static_cast<float>(tx[tb][1] - c.y);
      const float a2 = static_cast<float>(tx[tb][2] - c.z);
      const float n2 = a0 * a0 + a1 * a1 + a2 * a2;
      B[tb] = k == 0 ? -2.0f * a0 : k == 1 ? -2.0f * a1 : k == 2 ? -2.0f * a2 : n2;
    }
    mf4 D2[4];
#pragma unroll
    for (int tb = 0; tb < 4; ++tb) D2[tb] = mf4{0.0f, 0.0f, 0.0f, 0.0f};
    const float* lyf = reinterpret_cast<const float*>(ly);
#pragma unroll 2
    for (int ch = 0; ch < kPackTile / 16; ++ch) {
      const float A1 = lyf[(16 * ch + col) * 4 + k];
      const float4 qv = reinterpret_cast<const float4*>(lq)[4 * ch + k];
      const float4 tv = lt[(ch * 4 + k) * 4 + (lane & 3)];
      mf4 Cy = {qv.x, qv.y, qv.z, qv.w};
      if (EPS) Cy += mf4{eps2x2, eps2x2, eps2x2, eps2x2};
#pragma unroll
      for (int tb = 0; tb < 4; ++tb) {
        const mf4 r2 =
            __builtin_amdgcn_mfma_f32_16x16x4f32(A1, B[tb], Cy, 0, 0, 0);
        const float w0 = __builtin_amdgcn_rcpf(r2[0]);
        const float w1 = __builtin_amdgcn_rcpf(r2[1]);
        const float w2 = __builtin_amdgcn_rcpf(r2[2]);
        const float w3 = __builtin_amdgcn_rcpf(r2[3]);
        D2[tb] = __builtin_amdgcn_mfma_f32_4x4x1f32(tv.x, w0, D2[tb], 0, 0, 0);
        D2[tb] = __builtin_amdgcn_mfma_f32_4x4x1f32(tv.y, w1, D2[tb], 0, 0, 0);
        D2[tb] = __builtin_amdgcn_mfma_f32_4x4x1f32(tv.z, w2, D2[tb], 0, 0, 0);
        D2[tb] = __builtin_amdgcn_mfma_f32_4x4x1f32(tv.w, w3, D2[tb], 0, 0, 0);
      }
    }
#pragma unroll
    for (int tb = 0; tb < 4; ++tb) {
      const double W = static_cast<double>(D2[tb][3]);
      Sf[tb][0] += static_cast<double>(D2[tb][0]) + c.x * W;
      Sf[tb][1] += static_cast<double>(D2[tb][1]) + c.y * W;
      Sf[tb][2] += static_cast<double>(D2[tb][2]) + c.z * W;
    }
  }
  // sum the four source groups (lanes col, col + 16, col + 32, col + 48)
#pragma unroll
  for (int tb = 0; tb < 4; ++tb)
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      double v = Sf[tb][a];
      v += __shfl_xor(v, 16, 64);
      v += __shfl_xor(v, 32, 64);
      Sf[tb][a] = v;
    }
  // lane l is target 16 (l / 16) + l % 16 = l: take its column block k
  double S[1][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double f = k == 0 ? Sf[0][a] : k == 1 ? Sf[1][a] : k == 2 ? Sf[2][a] : Sf[3][a];
    S[0][a] = f + Sn[a];
  }
  store_cross<1>(sg, x, S, live, tgt);
}

// ---------------------------------------------- f16 hi/lo distance GEMM ---
// Variant of the matrix-core far field whose distance product runs on the
// f16 matrix cores (v_mfma_f32_32x32x16_f16: 32 sources x 32 targets per
// instruction, f32 accumulation).  Per tile the offsets are scaled by
// 1/S, S = max(R, kMfma16MinScale), so |u| = |y'|/S <= 1 and
// |v| = |x'|/S <= 2/S, and every FP32 coordinate is split into
// f16 hi + lo parts (22 significant bits).  With the K slots
//   A_j = (uh, ul, uh, 1, 1, 1, qh, ql, 0, 0)      (q = |u|^2)
//   B_i = (-2vh, -2vh, -2vl, ph, pl, pll, 1, 1, 0, 0)  (p = |v|^2 [+ eps])
// (vectors count 3 slots, 16 in all) one instruction gives
//   r^2 / S^2 = p + q - 2 (uh.vh + ul.vh + uh.vl),
// missing only ul.vl (2^-22 |u||v|).  p, the dominant term of a far pair,
// is split three ways so its own rounding stays near 2^-33 p.  The weight
// product stays in FP32 (v_mfma_f32_4x4x1_16b_f32) against
//   T_j = 2 s_j / S^2 (y'_j, 1)
// stored for the 32x32 output layout: lane l, register v of the distance
// product holds source 8 (v / 4) + 4 (l / 32) + v % 4, target l % 32.
constexpr double kMfma16MinScale = 0.01;

typedef _Float16 mh8 __attribute__((ext_vector_type(8)));
typedef float mf16 __attribute__((ext_vector_type(16)));

struct Mfma16Src {
  const float4* y;   // (y', 1) FP32 for the near path
  const float* s;    // 2 s
  const uint4* a;    // two uint4 (16 f16) per source
  const float* t;    // 1024 floats per tile
  const double4* c;  // (c, R^2) per tile
  const double* is;  // 1 / S per tile
};

__device__ __forceinline__ uint32_t pack_h2(_Float16 lo, _Float16 hi) {
  return static_cast<uint32_t>(__builtin_bit_cast(uint16_t, lo)) |
         (static_cast<uint32_t>(__builtin_bit_cast(uint16_t, hi)) << 16);
}

__device__ __forceinline__ void split_h(float x, _Float16& hi, _Float16& lo) {
  hi = static_cast<_Float16>(x);
  lo = static_cast<_Float16>(x - static_cast<float>(hi));
}

// One block per 256-source tile, after pack_mfma_kernel (reads its
// centroids and radii): the f16 records, 1/S and the weight operand.
__global__ __launch_bounds__(kPackTile) void pack_mfma16_kernel(
    const double4* __restrict__ p64, const int32_t* __restrict__ count,
    const double4* __restrict__ mf_c, uint4* __restrict__ mf_a,
    float* __restrict__ mf_t32, double* __restrict__ mf_is) {
  const int n = *count;
  const int j0 = blockIdx.x * kPackTile;
  if (j0 >= n) return;
  const int j = j0 + threadIdx.x;
  double4 y = make_double4(0.0, 0.0, 0.0, 0.0);
  if (j < n) y = p64[j];
  const bool real = (j < n) && !(y.x == 0.0 && y.y == 0.0 && y.z == 0.0 &&
                                 y.w == 0.0);
  const double4 c = mf_c[blockIdx.x];
  const double S = fmax(sqrt(c.w), kMfma16MinScale);
  const double inv = 1.0 / S;
  if (threadIdx.x == 0) mf_is[blockIdx.x] = inv;
  const double e0 = y.x - c.x, e1 = y.y - c.y, e2 = y.z - c.z;
  float u0 = 0.0f, u1 = 0.0f, u2 = 0.0f;
  double ws = 0.0;
  if (real) {
    u0 = static_cast<float>(e0 * inv);
    u1 = static_cast<float>(e1 * inv);
    u2 = static_cast<float>(e2 * inv);
    ws = 2.0 * y.w * inv * inv;
  }
  const float q = u0 * u0 + u1 * u1 + u2 * u2;
  _Float16 h0, l0, h1, l1, h2, l2, qh, ql;
  split_h(u0, h0, l0);
  split_h(u1, h1, l1);
  split_h(u2, h2, l2);
  split_h(q, qh, ql);
  const _Float16 one = static_cast<_Float16>(1.0f), zero = static_cast<_Float16>(0.0f);
  // K 0..7: uh0 uh1 uh2 ul0 ul1 ul2 uh0 uh1; K 8..15: uh2 1 1 1 qh ql 0 0
  mf_a[2 * static_cast<int64_t>(j)] =
      make_uint4(pack_h2(h0, h1), pack_h2(h2, l0), pack_h2(l1, l2), pack_h2(h0, h1));
  mf_a[2 * static_cast<int64_t>(j) + 1] =
      make_uint4(pack_h2(h2, one), pack_h2(one, one), pack_h2(qh, ql),
                 pack_h2(zero, zero));
  // weight operand: local source i = 32 ch + 8 (v / 4) + 4 h + v % 4
  const int jj = threadIdx.x, ch = jj >> 5, i = jj & 31;
  const int h = (i >> 2) & 1, v = 4 * (i >> 3) + (i & 3);
  float* tt = mf_t32 + static_cast<int64_t>(blockIdx.x) * 1024 +
              ((ch * 2 + h) * 4) * 16 + v;
  tt[0] = static_cast<float>(ws * e0);
  tt[16] = static_cast<float>(ws * e1);
  tt[32] = static_cast<float>(ws * e2);
  tt[48] = static_cast<float>(ws);
}

// Difference form d = y' - x' on the vector ALUs over n sources of a tile
// (near tiles), skipping local source `self` (any value outside [0, n):
// none).
template <bool EPS>
__device__ __forceinline__ void mfma16_near(const float4* ly, const float* ls,
                                            int n, float xr0, float xr1,
                                            float xr2, float eps2x2, int self,
                                            float& T0, float& T1, float& T2) {
  for (int jj = 0; jj < n; ++jj) {
    const float4 y = ly[jj];
    const float d0 = y.x - xr0, d1 = y.y - xr1, d2 = y.z - xr2;
    float r2 = d0 * d0 + d1 * d1 + d2 * d2;
    if (EPS) r2 += eps2x2;
    float w = ls[jj] * __builtin_amdgcn_rcpf(r2);
    if (jj == self) w = 0.0f;
    T0 = __builtin_fmaf(w, d0, T0);
    T1 = __builtin_fmaf(w, d1, T1);
    T2 = __builtin_fmaf(w, d2, T2);
  }
}

// B operands of the distance product for this wave's two 32-target blocks:
// the lane's target in tile coordinates (offset e from the centre, scale
// inv), split into f16 hi/lo, gathered so that lane l holds K half h of
// target column col of each block.
template <bool EPS>
__device__ __forceinline__ void mfma16_targets(double e0, double e1, double e2,
                                               double inv, float eps2x2,
                                               int col, int h, mh8* B) {
  const float v0 = static_cast<float>(e0 * inv), v1 = static_cast<float>(e1 * inv),
              v2 = static_cast<float>(e2 * inv);
  float p = v0 * v0 + v1 * v1 + v2 * v2;
  if (EPS) p += eps2x2 * static_cast<float>(inv * inv);
  _Float16 a0, b0, a1, b1, a2, b2, ph, pl, pll;
  split_h(-2.0f * v0, a0, b0);
  split_h(-2.0f * v1, a1, b1);
  split_h(-2.0f * v2, a2, b2);
  split_h(p, ph, pl);
  pll = static_cast<_Float16>(p - static_cast<float>(ph) - static_cast<float>(pl));
  const _Float16 one = static_cast<_Float16>(1.0f), zero = static_cast<_Float16>(0.0f);
  // K 0..7: -2vh0 -2vh1 -2vh2 -2vh0 -2vh1 -2vh2 -2vl0 -2vl1
  // K 8..15: -2vl2 ph pl pll 1 1 0 0
  const uint32_t own[8] = {pack_h2(a0, a1), pack_h2(a2, a0), pack_h2(a1, a2),
                           pack_h2(b0, b1), pack_h2(b2, ph), pack_h2(pl, pll),
                           pack_h2(one, one), pack_h2(zero, zero)};
#pragma unroll
  for (int tb = 0; tb < 2; ++tb) {
    const int from = 32 * tb + col;
    uint32_t g[8];
#pragma unroll
    for (int e = 0; e < 8; ++e)
      g[e] = static_cast<uint32_t>(__shfl(static_cast<int>(own[e]), from, 64));
    const uint4 sel = h ? make_uint4(g[4], g[5], g[6], g[7])
                        : make_uint4(g[0], g[1], g[2], g[3]);
    B[tb] = __builtin_bit_cast(mh8, sel);
  }
}

// One 32-source chunk: the distance products of both target blocks, their
// reciprocals and the weight products into D2.  la / lt address the chunk's
// sources (32 sources: 64 uint4 of A, 32 float4 of T).  RS: distances per
// reciprocal: 2 gives 1/Da = Db / (Da Db), 4 shares one reciprocal of
// (Da Db)(Dc Dd) the same way (multiplies for quarter-rate reciprocals).
// Far-tile distances are scaled by at most 1/kMfma16MinScale^2 and lie in
// about [0.2, 4e4] even on very fine meshes, so a product of four stays far
// inside the FP32 range.
template <int NACC, int RS = 1>
__device__ __forceinline__ void mfma16_chunk(const uint4* la, const float4* lt,
                                             int col, int h, int c4,
                                             const mh8* B, mf4 (*D2)[NACC]) {
  const mh8 A = __builtin_bit_cast(mh8, la[col * 2 + h]);
  const float4* tp = lt + (h * 4 + c4) * 4;
  const float4 t0 = tp[0], t1 = tp[1], t2 = tp[2], t3 = tp[3];
  const float tv[16] = {t0.x, t0.y, t0.z, t0.w, t1.x, t1.y, t1.z, t1.w,
                        t2.x, t2.y, t2.z, t2.w, t3.x, t3.y, t3.z, t3.w};
#pragma unroll
  for (int tb = 0; tb < 2; ++tb) {
    const mf16 D = __builtin_amdgcn_mfma_f32_32x32x16_f16(A, B[tb], mf16{}, 0, 0, 0);
    float w[16];
    if constexpr (RS == 4) {
#pragma unroll
      for (int v = 0; v < 16; v += 4) {
        const float pab = D[v] * D[v + 1], pcd = D[v + 2] * D[v + 3];
        const float r = __builtin_amdgcn_rcpf(pab * pcd);
        const float rab = r * pcd, rcd = r * pab;
        w[v] = D[v + 1] * rab;
        w[v + 1] = D[v] * rab;
        w[v + 2] = D[v + 3] * rcd;
        w[v + 3] = D[v + 2] * rcd;
      }
    } else if constexpr (RS == 2) {
#pragma unroll
      for (int v = 0; v < 16; v += 2) {
        const float r = __builtin_amdgcn_rcpf(D[v] * D[v + 1]);
        w[v] = D[v + 1] * r;
        w[v + 1] = D[v] * r;
      }
    } else {
#pragma unroll
      for (int v = 0; v < 16; ++v) w[v] = __builtin_amdgcn_rcpf(D[v]);
    }
#pragma unroll
    for (int v = 0; v < 16; ++v)
      D2[tb][v % NACC] = __builtin_amdgcn_mfma_f32_4x4x1f32(
          tv[v], w[v], D2[tb][v % NACC], 0, 0, 0);
  }
}

// The tile's (sum w y', sum w) for this lane's target from D2: the two
// source halves (lanes l, l ^ 32) of each column; lane l keeps column block
// h, which holds its own target 32 h + col = l.
template <int NACC>
__device__ __forceinline__ void mfma16_reduce(mf4 (*D2)[NACC], int h, float* o) {
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    float v0s = D2[0][0][a], v1s = D2[1][0][a];
#pragma unroll
    for (int q = 1; q < NACC; ++q) {
      v0s += D2[0][q][a];
      v1s += D2[1][q][a];
    }
    v0s += __shfl_xor(v0s, 32, 64);
    v1s += __shfl_xor(v1s, 32, 64);
    o[a] = h ? v1s : v0s;
  }
}

// NACC: independent weight-product accumulators per target block (1 or 2;
// two shorten the dependent MFMA chain).  KS > 1: only source chunk `chunk`
// of KS equal runs of packing tiles, sums to sg.part (as f64_block).
template <bool EPS, int NACC = 1, int KS = 1, int RS = 1>
__device__ __forceinline__ void mfma16_block(const Mfma16Src& src, int nsrc,
                                             const SegArgs& sg, int blk,
                                             float eps2x2, float4* ly,
                                             float* ls, uint4* la, float4* lt,
                                             int chunk = 0) {
  double x[1][3];
  int self[1];
  bool live[1];
  int64_t tgt[1];
  if (!load_targets<256, 1>(sg, blk, x, self, live, tgt)) return;
  const int lane = threadIdx.x & 63;
  const int h = lane >> 5;    // K half of the operands, source half of D
  const int col = lane & 31;  // target column
  const int c4 = lane & 3;
  double S0 = 0.0, S1 = 0.0, S2 = 0.0;
  const int ntile = (nsrc + kPackTile - 1) / kPackTile;
  const int jbeg = KS > 1 ? (ntile * chunk) / KS * kPackTile : 0;
  const int jend = KS > 1 ? (ntile * (chunk + 1)) / KS * kPackTile : nsrc;
  for (int j0 = jbeg; j0 < jend; j0 += kPackTile) {
    const int t = j0 / kPackTile;
    __syncthreads();
    {
      const int64_t base = static_cast<int64_t>(t) * kPackTile + threadIdx.x;
      ly[threadIdx.x] = src.y[base];
      ls[threadIdx.x] = src.s[base];
      la[2 * threadIdx.x] = src.a[2 * base];
      la[2 * threadIdx.x + 1] = src.a[2 * base + 1];
      lt[threadIdx.x] = reinterpret_cast<const float4*>(src.t)[base];
    }
    __syncthreads();
    const double4 c = src.c[t];
    const double e0 = x[0][0] - c.x, e1 = x[0][1] - c.y, e2 = x[0][2] - c.z;
    const bool near =
        e0 * e0 + e1 * e1 + e2 * e2 <= kMfmaNear * kMfmaNear * c.w;
    if (__ballot(near)) {
      float T0 = 0.0f, T1 = 0.0f, T2 = 0.0f;
      mfma16_near<EPS>(ly, ls, kPackTile, static_cast<float>(e0),
                       static_cast<float>(e1), static_cast<float>(e2), eps2x2,
                       self[0] - j0, T0, T1, T2);
      S0 += static_cast<double>(T0);
      S1 += static_cast<double>(T1);
      S2 += static_cast<double>(T2);
      continue;
    }
    mh8 B[2];
    mfma16_targets<EPS>(e0, e1, e2, src.is[t], eps2x2, col, h, B);
    mf4 D2[2][NACC];
#pragma unroll
    for (int tb = 0; tb < 2; ++tb)
#pragma unroll
      for (int q = 0; q < NACC; ++q) D2[tb][q] = mf4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 1
    for (int ch = 0; ch < kPackTile / 32; ++ch)
      mfma16_chunk<NACC, RS>(la + 64 * ch, lt + 32 * ch, col, h, c4, B, D2);
    float o[4];
    mfma16_reduce<NACC>(D2, h, o);
    const double W = static_cast<double>(o[3]);
    S0 += static_cast<double>(o[0]) + c.x * W;
    S1 += static_cast<double>(o[1]) + c.y * W;
    S2 += static_cast<double>(o[2]) + c.z * W;
  }
  double S[1][3] = {{S0, S1, S2}};
  if constexpr (KS > 1)
    store_partial<256, 1>(sg, blk, chunk, S);
  else
    store_cross<1>(sg, x, S, live, tgt);
}

}  // namespace
}  // namespace lpm_amd
