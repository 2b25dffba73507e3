// Sphere shallow-water direct sums for gfx950: velocity from relative
// vorticity and divergence, the velocity-gradient double dot
// grad(u):grad(u)^T, and the PSE surface Laplacian, per target.  Replaces
// SphereVertexSums / SphereFaceSums with SphereSweDirectSumReducer and
// sphere_swe_velocity_sums (src/lpm_swe_kernels.hpp:334-362, 578-619,
// 723-780, 877-930) plus the Gaussian PSE Laplacian of LaplacianReducer
// (src/lpm_pse.hpp:157-203) on chord distances.
//
// Per pair (target x, source y) with d = y - x, delta = 1 - x.y,
// D = delta + eps^2, a = zeta A / 4pi, b = sigma A / 4pi:
//   alpha = a / D, beta = b / D, a' = alpha / D, b' = beta / D.
// The reference's 12-tuple sums factor into per-target accumulators
//   A0 = sum alpha, Vd = sum alpha d, Yd = sum a' d d^T, Wd = sum a' d,
//   Bs = sum beta, Vs = sum beta d, Sbd = sum beta delta,
//   Ys = sum b' d d^T, Ws = sum b' delta d, Sd2 = sum b' delta^2,
// and the finalisation applies the target-only factors once:
//   u  = -x cross Vd - Vs - x Sbd
//   Mz = E(x) A0 + E(Vd) + [x]x (Yd - eps^2 Wd x^T)       (E_ab = e_abc v_c)
//   Ms = (Bs - Sbd) P(x) + x (Vs + Sbd x)^T
//        - (Ys + Ws x^T + x Ws^T + Sd2 x x^T)
//   M  = Mz + Ms  (the reference's packed gradient, whose vorticity part
//                  carries grad_kzeta's opposite sign convention)
//   ddot = sum_ab M_ab M_ba.
// The PSE Laplacian sum_j (s_j - s_i) A_j (4/pi) exp(-|x-y|^2/e^2) / e^4 over
// x.y > 0 is accumulated as L = sum A eta (s_j - s_i) in FP64 by both
// kernels (a second difference: its first-order terms cancel by symmetry,
// which FP32 sums cannot resolve).  Source tiles whose bounding sphere lies
// beyond sqrt(kPseCut) e of every target of a wave are skipped (their terms
// are below exp(-kPseCut) relative), so the PSE work covers a few percent of
// the pairs.
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace lpm_amd {

namespace {

constexpr double kInv4Pi = 0.079577471545947667884441881686257181;
constexpr double k4OverPi = 1.2732395447351626861510701069801148;
constexpr int kSweTile = 256;
// mixed kernel: source tiles this many bounding radii from a target of the
// wave (centre = the tile's first source) take the FP64 pair loop
constexpr double kSweNearFp64 = 2.0;
constexpr double kPseCut = 50.0;  // exp(-50) ~ 2e-22

struct SweSrc {
  double4 p;  // y0, y1, y2, a = zeta A / 4pi
  double4 q;  // b = sigma A / 4pi, A, s, 0
};

__global__ __launch_bounds__(256) void swe_pack_kernel(
    const double* __restrict__ sx, int64_t es, int64_t cs,
    const double* __restrict__ zeta, const double* __restrict__ sigma,
    const double* __restrict__ area, const double* __restrict__ surf,
    const int32_t* __restrict__ inv, const int32_t* __restrict__ count,
    SweSrc* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= *count) return;
  const int64_t i = inv ? inv[r] : r;
  const double* p = sx + i * es;
  const double A = area[i];
  SweSrc s;
  s.p = make_double4(p[0], p[cs], p[2 * cs], zeta[i] * A * kInv4Pi);
  s.q = make_double4(sigma[i] * A * kInv4Pi, A, surf ? surf[i] : 0.0, 0.0);
  out[r] = s;
}

// Bounding sphere of each packing tile: centre = first source, radius^2 =
// max |y - c|^2 (for PSE culling); w holds radius^2.
__global__ __launch_bounds__(kSweTile) void swe_tile_bounds_kernel(
    const SweSrc* __restrict__ src, const int32_t* __restrict__ count,
    double4* __restrict__ bounds) {
  __shared__ double red[kSweTile / 64];
  const int n = *count;
  const int j0 = blockIdx.x * kSweTile;
  if (j0 >= n) return;
  const double4 c = src[j0].p;
  const int j = j0 + threadIdx.x;
  double r2 = 0.0;
  if (j < n) {
    const double4 y = src[j].p;
    const double a = y.x - c.x, b = y.y - c.y, e = y.z - c.z;
    r2 = a * a + b * b + e * e;
  }
  for (int off = 32; off > 0; off >>= 1) r2 = fmax(r2, __shfl_down(r2, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kSweTile / 64; ++w) m = fmax(m, red[w]);
    bounds[blockIdx.x] = make_double4(c.x, c.y, c.z, m);
  }
}

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

// x.y > 0 evaluated with the reference's rounding, ((x0 y0 + x1 y1) + x2 y2)
// without contraction, so that points on the cut (x.y = 0 in exact
// arithmetic, common on the cubed sphere) are classified identically.
// __dmul_rn / __dadd_rn are plain * and + in this toolchain and would still
// be fused into FMAs; the pragma keeps three multiplies and two adds.
__device__ __forceinline__ bool hemisphere(const double* x, const double4& y) {
#pragma clang fp contract(off)
  const double t = (x[0] * y.x + x[1] * y.y) + x[2] * y.z;
  return t > 0.0;
}

struct SweSeg {
  const double* x;
  int64_t es, cs, n;
  const int32_t* index;
  const int32_t* d_n;
  const int32_t* self_rank;
  int self_identity;
  int nblocks;
  const double* surf;  // target surface (PSE)
  double* out_u;
  int64_t oes, ocs;
  double* out_ddot;
  double* out_laps;
  int64_t slot_base;  // slot c is target index[c] or slot_base + c
  int32_t rank_base;  // self = self_rank[i] - rank_base
};

// Target of slot c and its own source (-1: none) as in the BVE kernels.
__device__ __forceinline__ int64_t swe_target(const SweSeg& sg, int64_t c,
                                              bool in) {
  return in ? (sg.index ? sg.index[c] : sg.slot_base + c) : sg.slot_base;
}
__device__ __forceinline__ int swe_self(const SweSeg& sg, int64_t c,
                                        int64_t i, bool in, int skip_self) {
  if (!in || !skip_self) return -1;
  if (sg.self_rank) {
    const int r = sg.self_rank[i];
    return r < 0 ? -1 : r - sg.rank_base;
  }
  return sg.self_identity ? static_cast<int>(c) : -1;
}

// Accumulators of one target (see the file comment).
template <typename T>
struct SweAcc {
  T A0, Vd[3], Yd[6], Wd[3];  // Yd: xx, xy, xz, yy, yz, zz
  T Bs, Vs[3], Sbd, Ys[6], Ws[3], Sd2;
  T L1, L0;  // L1: PSE sum (FP64 accumulators only); L0 unused
  __device__ void zero() {
    A0 = Bs = Sbd = Sd2 = L1 = L0 = T(0);
    for (int k = 0; k < 3; ++k) Vd[k] = Wd[k] = Vs[k] = Ws[k] = T(0);
    for (int k = 0; k < 6; ++k) Yd[k] = Ys[k] = T(0);
  }
};

// Velocity/gradient terms of one pair.  d: y - x; delta: 1 - x.y.
template <typename T>
__device__ __forceinline__ void swe_pair(SweAcc<T>& acc, const T d0,
                                         const T d1, const T d2,
                                         const T delta, const T r,
                                         const T a, const T b) {
  const T al = a * r, be = b * r;
  const T ap = al * r, bp = be * r;
  acc.A0 += al;
  acc.Vd[0] = fma(al, d0, acc.Vd[0]);
  acc.Vd[1] = fma(al, d1, acc.Vd[1]);
  acc.Vd[2] = fma(al, d2, acc.Vd[2]);
  const T e0 = ap * d0, e1 = ap * d1, e2 = ap * d2;
  acc.Wd[0] += e0;
  acc.Wd[1] += e1;
  acc.Wd[2] += e2;
  acc.Yd[0] = fma(e0, d0, acc.Yd[0]);
  acc.Yd[1] = fma(e0, d1, acc.Yd[1]);
  acc.Yd[2] = fma(e0, d2, acc.Yd[2]);
  acc.Yd[3] = fma(e1, d1, acc.Yd[3]);
  acc.Yd[4] = fma(e1, d2, acc.Yd[4]);
  acc.Yd[5] = fma(e2, d2, acc.Yd[5]);
  acc.Bs += be;
  acc.Vs[0] = fma(be, d0, acc.Vs[0]);
  acc.Vs[1] = fma(be, d1, acc.Vs[1]);
  acc.Vs[2] = fma(be, d2, acc.Vs[2]);
  acc.Sbd = fma(be, delta, acc.Sbd);
  const T f0 = bp * d0, f1 = bp * d1, f2 = bp * d2;
  acc.Ys[0] = fma(f0, d0, acc.Ys[0]);
  acc.Ys[1] = fma(f0, d1, acc.Ys[1]);
  acc.Ys[2] = fma(f0, d2, acc.Ys[2]);
  acc.Ys[3] = fma(f1, d1, acc.Ys[3]);
  acc.Ys[4] = fma(f1, d2, acc.Ys[4]);
  acc.Ys[5] = fma(f2, d2, acc.Ys[5]);
  const T bd = bp * delta;
  acc.Ws[0] = fma(bd, d0, acc.Ws[0]);
  acc.Ws[1] = fma(bd, d1, acc.Ws[1]);
  acc.Ws[2] = fma(bd, d2, acc.Ws[2]);
  acc.Sd2 = fma(bd, delta, acc.Sd2);
}

__device__ __forceinline__ void sym(const double* Y, double (*m)[3]) {
  m[0][0] = Y[0];
  m[0][1] = m[1][0] = Y[1];
  m[0][2] = m[2][0] = Y[2];
  m[1][1] = Y[3];
  m[1][2] = m[2][1] = Y[4];
  m[2][2] = Y[5];
}

// u, ddot and the Laplacian from the accumulators of target x.
__device__ __forceinline__ void swe_finalize(const SweAcc<double>& A, const double* x,
                             double s_i, double eps2, double pse_scale,
                             double* u, double* ddot, double* laps) {
  u[0] = -(x[1] * A.Vd[2] - x[2] * A.Vd[1]) - A.Vs[0] - x[0] * A.Sbd;
  u[1] = -(x[2] * A.Vd[0] - x[0] * A.Vd[2]) - A.Vs[1] - x[1] * A.Sbd;
  u[2] = -(x[0] * A.Vd[1] - x[1] * A.Vd[0]) - A.Vs[2] - x[2] * A.Sbd;

  double Yd[3][3], Ys[3][3];
  sym(A.Yd, Yd);
  sym(A.Ys, Ys);
  // G = Yd - eps^2 Wd x^T ; Mz = E(x) A0 + E(Vd) + [x]x G
  double G[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) G[a][b] = Yd[a][b] - eps2 * A.Wd[a] * x[b];
  // [x]x G: row a = (x cross column)
  double M[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    M[0][b] = x[1] * G[2][b] - x[2] * G[1][b];
    M[1][b] = x[2] * G[0][b] - x[0] * G[2][b];
    M[2][b] = x[0] * G[1][b] - x[1] * G[0][b];
  }
  const double v0 = x[0] * A.A0 + A.Vd[0], v1 = x[1] * A.A0 + A.Vd[1],
               v2 = x[2] * A.A0 + A.Vd[2];
  // E(v): [[0, v2, -v1], [-v2, 0, v0], [v1, -v0, 0]]
  M[0][1] += v2;
  M[0][2] -= v1;
  M[1][0] -= v2;
  M[1][2] += v0;
  M[2][0] += v1;
  M[2][1] -= v0;
  const double s1 = A.Bs - A.Sbd;
  const double w[3] = {A.Vs[0] + A.Sbd * x[0], A.Vs[1] + A.Sbd * x[1],
                       A.Vs[2] + A.Sbd * x[2]};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double P = (a == b ? 1.0 : 0.0) - x[a] * x[b];
      M[a][b] += s1 * P + x[a] * w[b] -
                 (Ys[a][b] + A.Ws[a] * x[b] + x[a] * A.Ws[b] +
                  A.Sd2 * x[a] * x[b]);
    }
  double dd = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) dd += M[a][b] * M[b][a];
  *ddot = dd;
  *laps = pse_scale * A.L1;
}

// FP64 pair terms of one LDS tile (cnt sources, zero padding up to an even
// count) for target x; `self` is the target's own source in tile numbering
// (outside [0, cnt): none) and any_self says whether some lane of the wave
// has one here.  Without, two sources share one reciprocal,
// 1/D_a = D_b / (D_a D_b).
__device__ __forceinline__ void swe_tile_f64(SweAcc<double>& acc,
                                             const SweSrc* tile, int cnt,
                                             const double* x, double eps2,
                                             int self, bool any_self) {
  if (any_self) {
    for (int jj = 0; jj < cnt; ++jj) {
      const double4 yp = tile[jj].p;
      const double4 yq = tile[jj].q;
      const double d0 = yp.x - x[0], d1 = yp.y - x[1], d2 = yp.z - x[2];
      const double t = x[0] * yp.x + x[1] * yp.y + x[2] * yp.z;
      const double delta = 1.0 - t;
      double r = rcp_nr<2>(delta + eps2);
      if (jj == self) r = 0.0;
      swe_pair<double>(acc, d0, d1, d2, delta, r, yp.w, yq.x);
    }
  } else {
    for (int jj = 0; jj < cnt; jj += 2) {
      const double4 ya = tile[jj].p, qa = tile[jj].q;
      const double4 yb = tile[jj + 1].p, qb = tile[jj + 1].q;
      const double a0 = ya.x - x[0], a1 = ya.y - x[1], a2 = ya.z - x[2];
      const double b0 = yb.x - x[0], b1 = yb.y - x[1], b2 = yb.z - x[2];
      const double da = 1.0 - (x[0] * ya.x + x[1] * ya.y + x[2] * ya.z);
      const double db = 1.0 - (x[0] * yb.x + x[1] * yb.y + x[2] * yb.z);
      const double Da = da + eps2, Db = db + eps2;
      const double r = rcp_nr<2>(Da * Db);
      swe_pair<double>(acc, a0, a1, a2, da, Db * r, ya.w, qa.x);
      swe_pair<double>(acc, b0, b1, b2, db, Da * r, yb.w, qb.x);
    }
  }
}

// ----------------------------------------------------------- FP64 kernel --
// One target per lane; sources staged through LDS.  PSE terms are evaluated
// only for tiles near some target of the wave.
// Parameters shared by every target of a launch.
struct SweParams {
  double eps2, inv_pse_e2, pse_cut2;
  int do_pse, skip_self;
};

// Block `blk` of segment `sg` against nsrc packed sources (bounds: their
// tile bounding spheres).  Every thread of the block calls it.
template <int BLOCK>
__device__ __forceinline__ void swe_f64_block(
    const SweSrc* __restrict__ src, const double4* __restrict__ bounds,
    int nsrc, const SweSeg& sg, int blk, const SweParams& P, SweSrc* tile) {
  const double eps2 = P.eps2, inv_pse_e2 = P.inv_pse_e2;
  const double pse_cut2 = P.pse_cut2;
  const int do_pse = P.do_pse;
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  const int64_t c = static_cast<int64_t>(blk) * BLOCK + threadIdx.x;
  if (static_cast<int64_t>(blk) * BLOCK >= n_live) return;
  const bool in = c < n_live;
  const int64_t i = swe_target(sg, c, in);
  const double* px = sg.x + i * sg.es;
  const double x[3] = {px[0], px[sg.cs], px[2 * sg.cs]};
  const double s_i = (do_pse && sg.surf) ? sg.surf[i] : 0.0;
  const int self = swe_self(sg, c, i, in, P.skip_self);

  SweAcc<double> acc;
  acc.zero();
  for (int j0 = 0; j0 < nsrc; j0 += kSweTile) {
    __syncthreads();
    for (int t = threadIdx.x; t < kSweTile; t += BLOCK) {
      const int j = j0 + t;
      if (j < nsrc) {
        tile[t] = src[j];
      } else {
        tile[t].p = make_double4(0.0, 0.0, 0.0, 0.0);
        tile[t].q = make_double4(0.0, 0.0, 0.0, 0.0);
      }
    }
    __syncthreads();
    const int cnt = min(kSweTile, nsrc - j0);
    const bool diag = static_cast<unsigned>(self - j0) <
                      static_cast<unsigned>(cnt);
    swe_tile_f64(acc, tile, cnt, x, eps2, self - j0, __ballot(diag) != 0);
    if (do_pse) {
      // wave-uniform test: is any target of the wave near this tile?
      const double4 bd = bounds[j0 / kSweTile];
      const double e0 = x[0] - bd.x, e1 = x[1] - bd.y, e2 = x[2] - bd.z;
      const double dist2 = e0 * e0 + e1 * e1 + e2 * e2;
      const double rr = sqrt(bd.w) + sqrt(pse_cut2);
      const bool near = in && dist2 <= rr * rr;
      if (__ballot(near)) {
        for (int jj = 0; jj < cnt; ++jj) {
          const double4 yp = tile[jj].p;
          const double4 yq = tile[jj].q;
          const double d0 = yp.x - x[0], d1 = yp.y - x[1], d2 = yp.z - x[2];
          const double rho2 = d0 * d0 + d1 * d1 + d2 * d2;
          double eta = exp(-rho2 * inv_pse_e2) * yq.y;
          if (!hemisphere(x, yp)) eta = 0.0;
          acc.L1 = fma(eta, yq.z - s_i, acc.L1);
        }
      }
    }
  }
  if (!in) return;
  double u[3], dd, lap;
  swe_finalize(acc, x, s_i, eps2, k4OverPi * inv_pse_e2 * inv_pse_e2, u, &dd,
               &lap);
  if (sg.out_u) {
    double* o = sg.out_u + i * sg.oes;
    o[0] = u[0];
    o[sg.ocs] = u[1];
    o[2 * sg.ocs] = u[2];
  }
  if (sg.out_ddot) sg.out_ddot[i] = dd;
  if (sg.out_laps && do_pse) sg.out_laps[i] = lap;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void swe_f64_kernel(
    const SweSrc* __restrict__ src, const double4* __restrict__ bounds,
    const int32_t* __restrict__ count, SweSeg s0, SweSeg s1, SweParams P) {
  __shared__ SweSrc tile[kSweTile];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SweSeg& sg = first ? s0 : s1;
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  swe_f64_block<BLOCK>(src, bounds, *count, sg, blk, P, tile);
}

// ---------------------------------------------------------- mixed kernel --
// FP32 pair arithmetic on tile-relative coordinates (d formed from offsets
// to the tile's FP64 centre, delta = |d|^2 / 2); FP32 partial sums per
// source tile are carried into FP64 accumulators.
struct F4x2 {
  float4 p;  // y - c (3), a
  float4 q;  // b, A, 0, 0
};

// As swe_f64_block with FP32 pair arithmetic; tile64 is the FP64 copy of
// the LDS tile for the PSE terms.
template <int BLOCK>
__device__ __forceinline__ void swe_mixed_block(
    const SweSrc* __restrict__ src, const double4* __restrict__ bounds,
    int nsrc, const SweSeg& sg, int blk, const SweParams& P, F4x2* tile,
    SweSrc* tile64) {
  const double eps2 = P.eps2, inv_pse_e2 = P.inv_pse_e2;
  const double pse_cut2 = P.pse_cut2;
  const int do_pse = P.do_pse;
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  const int64_t c = static_cast<int64_t>(blk) * BLOCK + threadIdx.x;
  if (static_cast<int64_t>(blk) * BLOCK >= n_live) return;
  const bool in = c < n_live;
  const int64_t i = swe_target(sg, c, in);
  const double* px = sg.x + i * sg.es;
  const double x[3] = {px[0], px[sg.cs], px[2 * sg.cs]};
  const double s_i = (do_pse && sg.surf) ? sg.surf[i] : 0.0;
  const int self = swe_self(sg, c, i, in, P.skip_self);
  const float e2f = static_cast<float>(eps2);

  SweAcc<double> acc;
  acc.zero();
  for (int j0 = 0; j0 < nsrc; j0 += kSweTile) {
    const double4 bd = bounds[j0 / kSweTile];
    __syncthreads();
    for (int t = threadIdx.x; t < kSweTile; t += BLOCK) {
      const int j = j0 + t;
      F4x2 z;
      if (j < nsrc) {
        const SweSrc sj = src[j];
        z.p = make_float4(static_cast<float>(sj.p.x - bd.x),
                          static_cast<float>(sj.p.y - bd.y),
                          static_cast<float>(sj.p.z - bd.z),
                          static_cast<float>(sj.p.w));
        z.q = make_float4(static_cast<float>(sj.q.x),
                          static_cast<float>(sj.q.y), 0.0f, 0.0f);
        tile64[t] = sj;
      } else {
        z.p = make_float4(8.0f, 8.0f, 8.0f, 0.0f);
        z.q = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        tile64[t].p = make_double4(0.0, 0.0, 0.0, 0.0);
        tile64[t].q = make_double4(0.0, 0.0, 0.0, 0.0);
      }
      tile[t] = z;
    }
    __syncthreads();
    const float xr[3] = {static_cast<float>(x[0] - bd.x),
                         static_cast<float>(x[1] - bd.y),
                         static_cast<float>(x[2] - bd.z)};
    const int cnt = min(kSweTile, nsrc - j0);
    const bool diag = static_cast<unsigned>(self - j0) <
                      static_cast<unsigned>(cnt);
    const double g0 = x[0] - bd.x, g1 = x[1] - bd.y, g2 = x[2] - bd.z;
    // tiles within kSweNearFp64 bounding radii of some target of the wave
    // hold the largest pair terms of grad u, whose assembly cancels: those
    // run in FP64 straight into the FP64 accumulators
    if (__ballot(in && g0 * g0 + g1 * g1 + g2 * g2 <=
                           kSweNearFp64 * kSweNearFp64 * bd.w)) {
      swe_tile_f64(acc, tile64, cnt, x, eps2, self - j0, __ballot(diag) != 0);
    } else {
      SweAcc<float> p;
      p.zero();
      // the per-pair self test only in tiles holding some lane's own source
      auto pairs = [&](auto check) {
        for (int jj = 0; jj < cnt; ++jj) {
          const float4 yp = tile[jj].p;
          const float4 yq = tile[jj].q;
          const float d0 = yp.x - xr[0], d1 = yp.y - xr[1], d2 = yp.z - xr[2];
          const float rho2 = d0 * d0 + d1 * d1 + d2 * d2;
          const float delta = 0.5f * rho2;
          float r = __builtin_amdgcn_rcpf(delta + e2f);
          if (decltype(check)::value && j0 + jj == self) r = 0.0f;
          swe_pair<float>(p, d0, d1, d2, delta, r, yp.w, yq.x);
        }
      };
      if (__ballot(diag))
        pairs(std::true_type{});
      else
        pairs(std::false_type{});
      // carry the tile's FP32 partials into the FP64 accumulators
      acc.A0 += p.A0;
      acc.Bs += p.Bs;
      acc.Sbd += p.Sbd;
      acc.Sd2 += p.Sd2;

      for (int k = 0; k < 3; ++k) {
        acc.Vd[k] += p.Vd[k];
        acc.Wd[k] += p.Wd[k];
        acc.Vs[k] += p.Vs[k];
        acc.Ws[k] += p.Ws[k];
      }
      for (int k = 0; k < 6; ++k) {
        acc.Yd[k] += p.Yd[k];
        acc.Ys[k] += p.Ys[k];
      }
    }
    if (do_pse) {
      const double e0 = x[0] - bd.x, e1 = x[1] - bd.y, e2 = x[2] - bd.z;
      const double dist2 = e0 * e0 + e1 * e1 + e2 * e2;
      const double rr = sqrt(bd.w) + sqrt(pse_cut2);
      const bool near = in && dist2 <= rr * rr;
      if (__ballot(near)) {
        // the PSE sum is a second difference whose first-order terms cancel
        // by symmetry; it covers few tiles, so it runs in FP64
        double L = 0.0;
        for (int jj = 0; jj < cnt; ++jj) {
          const double4 yp = tile64[jj].p;
          const double4 yq = tile64[jj].q;
          const double d0 = yp.x - x[0], d1 = yp.y - x[1], d2 = yp.z - x[2];
          const double rho2 = d0 * d0 + d1 * d1 + d2 * d2;
          double eta = exp(-rho2 * inv_pse_e2) * yq.y;
          if (!hemisphere(x, yp)) eta = 0.0;
          L = fma(eta, yq.z - s_i, L);
        }
        acc.L1 += L;
      }
    }
  }
  if (!in) return;
  double u[3], dd, lap;
  swe_finalize(acc, x, s_i, eps2, k4OverPi * inv_pse_e2 * inv_pse_e2, u, &dd,
               &lap);
  if (sg.out_u) {
    double* o = sg.out_u + i * sg.oes;
    o[0] = u[0];
    o[sg.ocs] = u[1];
    o[2 * sg.ocs] = u[2];
  }
  if (sg.out_ddot) sg.out_ddot[i] = dd;
  if (sg.out_laps && do_pse) sg.out_laps[i] = lap;
}

template <int BLOCK>
// at least 3 waves per SIMD: the FP64 accumulators, the FP32 tile partials
// and the PSE loop otherwise hold 198 registers (2 waves); the few spills
// this costs sit outside the pair loop
__global__ __launch_bounds__(BLOCK, 3) void swe_mixed_kernel(
    const SweSrc* __restrict__ src, const double4* __restrict__ bounds,
    const int32_t* __restrict__ count, SweSeg s0, SweSeg s1, SweParams P) {
  __shared__ F4x2 tile[kSweTile];
  __shared__ SweSrc tile64[kSweTile];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SweSeg& sg = first ? s0 : s1;
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  swe_mixed_block<BLOCK>(src, bounds, *count, sg, blk, P, tile, tile64);
}

// ------------------------------------------------------------ ensembles --
// grid (ceil(max member faces / 256), n_members): member m's c-th undivided
// face at packed slot P_m + c; zero sources pad the member's last tile.
__global__ __launch_bounds__(256) void swe_ens_pack_kernel(
    const double* __restrict__ sx, int64_t es, int64_t cs,
    const double* __restrict__ zeta, const double* __restrict__ sigma,
    const double* __restrict__ area, const double* __restrict__ surf,
    const int32_t* __restrict__ inv, const EnsMember* __restrict__ mem,
    SweSrc* __restrict__ out) {
  const EnsMember M = mem[blockIdx.y];
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int padded = (M.cnt + kSweTile - 1) / kSweTile * kSweTile;
  if (c >= padded) return;
  SweSrc v;
  v.p = make_double4(0.0, 0.0, 0.0, 0.0);
  v.q = make_double4(0.0, 0.0, 0.0, 0.0);
  if (c < M.cnt) {
    const int64_t r = static_cast<int64_t>(M.R) + c;
    const int64_t i = inv ? inv[r] : r;
    const double* p = sx + i * es;
    const double A = area[i];
    v.p = make_double4(p[0], p[cs], p[2 * cs], zeta[i] * A * kInv4Pi);
    v.q = make_double4(sigma[i] * A * kInv4Pi, A, surf ? surf[i] : 0.0, 0.0);
  }
  out[M.P + c] = v;
}

// grid (ceil(max member faces / 256), n_members): bounding sphere of tile t
// of member m over its real sources only.
__global__ __launch_bounds__(kSweTile) void swe_ens_bounds_kernel(
    const SweSrc* __restrict__ src, const EnsMember* __restrict__ mem,
    double4* __restrict__ bounds) {
  __shared__ double red[kSweTile / 64];
  const EnsMember M = mem[blockIdx.y];
  const int j0 = blockIdx.x * kSweTile;
  if (j0 >= M.cnt) return;
  const double4 c = src[M.P + j0].p;
  const int j = j0 + threadIdx.x;
  double r2 = 0.0;
  if (j < M.cnt) {
    const double4 y = src[M.P + j].p;
    const double a = y.x - c.x, b = y.y - c.y, e = y.z - c.z;
    r2 = a * a + b * b + e * e;
  }
  for (int off = 32; off > 0; off >>= 1) r2 = fmax(r2, __shfl_down(r2, off, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = r2;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0.0;
    for (int w = 0; w < kSweTile / 64; ++w) m = fmax(m, red[w]);
    bounds[(M.P + j0) / kSweTile] = make_double4(c.x, c.y, c.z, m);
  }
}

struct SweEnsArgs {
  const EnsMember* mem;
  const int4* items;
  int n_items;
  int32_t* queue;
  SweSeg vert;  // arrays of the vertex targets (x, surf, outputs)
  SweSeg face;  // arrays of the face targets
  const int32_t* rank;
  const int32_t* inv;
  int skip_masked;
  SweParams P;
};

__device__ __forceinline__ SweSeg swe_ens_segment(const SweEnsArgs& A,
                                                  const EnsMember& M, int seg) {
  if (seg == 0) {
    SweSeg sg = A.vert;
    sg.n = M.nv;
    sg.slot_base = M.v0;
    return sg;
  }
  SweSeg sg = A.face;
  if (A.skip_masked && A.inv) {
    sg.index = A.inv + M.R;
    sg.n = M.cnt;
    sg.self_identity = 1;
  } else {
    sg.n = M.nf;
    sg.slot_base = M.f0;
    if (A.rank) {
      sg.self_rank = A.rank;
      sg.rank_base = M.R;
    } else {
      sg.self_identity = 1;
    }
  }
  return sg;
}

__device__ __forceinline__ int swe_ens_next(const SweEnsArgs& A, int* slot) {
  if (threadIdx.x == 0) *slot = atomicAdd(A.queue, 1);
  __syncthreads();
  const int it = *slot;
  __syncthreads();
  return it;
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void swe_ens_f64_kernel(
    const SweSrc* __restrict__ src, const double4* __restrict__ bounds,
    SweEnsArgs A) {
  __shared__ SweSrc tile[kSweTile];
  __shared__ int slot;
  for (;;) {
    const int it = swe_ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SweSeg sg = swe_ens_segment(A, M, w.y);
    swe_f64_block<BLOCK>(src + M.P, bounds + M.P / kSweTile, M.cnt, sg, w.z,
                         A.P, tile);
  }
}

template <int BLOCK>
// at least 3 waves per SIMD: the FP64 accumulators, the FP32 tile partials
// and the PSE loop otherwise hold 198 registers (2 waves); the few spills
// this costs sit outside the pair loop
__global__ __launch_bounds__(BLOCK, 3) void swe_ens_mixed_kernel(
    const SweSrc* __restrict__ src, const double4* __restrict__ bounds,
    SweEnsArgs A) {
  __shared__ F4x2 tile[kSweTile];
  __shared__ SweSrc tile64[kSweTile];
  __shared__ int slot;
  for (;;) {
    const int it = swe_ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SweSeg sg = swe_ens_segment(A, M, w.y);
    swe_mixed_block<BLOCK>(src + M.P, bounds + M.P / kSweTile, M.cnt, sg, w.z,
                           A.P, tile, tile64);
  }
}

SweSeg to_seg(const SweTargets& t, int per_block) {
  SweSeg a{};
  a.x = t.base.x.base;
  a.es = t.base.x.es;
  a.cs = t.base.x.cs;
  a.n = t.base.n;
  a.index = t.base.index;
  a.d_n = t.base.d_n;
  a.self_rank = t.base.self_rank;
  a.self_identity = t.base.self_identity ? 1 : 0;
  a.nblocks = static_cast<int>((t.base.n + per_block - 1) / per_block);
  a.surf = t.surf;
  a.out_u = t.base.out.base;
  a.oes = t.base.out.es;
  a.ocs = t.base.out.cs;
  a.out_ddot = t.ddot;
  a.out_laps = t.laps;
  return a;
}

}  // namespace

SwePacked swe_pack_sources(lpm_ctx* ctx, CrdIn sx, const double* zeta,
                           const double* sigma, const double* area,
                           const double* surf, const RankMap& map,
                           hipStream_t s) {
  SwePacked ps;
  const int64_t n = map.n;
  const int64_t cap = std::max<int64_t>(n, 1);
  SweSrc* p = static_cast<SweSrc*>(ctx->swe_packed.ensure(sizeof(SweSrc) * cap));
  const int ntile = static_cast<int>((cap + kSweTile - 1) / kSweTile);
  double4* b = static_cast<double4*>(ctx->swe_bounds.ensure(sizeof(double4) * ntile));
  if (n > 0) {
    swe_pack_kernel<<<static_cast<int>((n + 255) / 256), 256, 0, s>>>(
        sx.base, sx.es, sx.cs, zeta, sigma, area, surf, map.inv, map.d_count, p);
    LPM_HIP_CHECK(hipGetLastError());
    swe_tile_bounds_kernel<<<ntile, kSweTile, 0, s>>>(p, map.d_count, b);
    LPM_HIP_CHECK(hipGetLastError());
  }
  ps.data = p;
  ps.bounds = b;
  ps.count = map.d_count;
  ps.capacity = n;
  return ps;
}

void launch_swe_sums(lpm_ctx* ctx, const SwePacked& src, SweTargets* segs,
                     int nseg, double eps, double pse_eps, int precision,
                     hipStream_t s) {
  LPM_ARG_CHECK(nseg >= 1 && nseg <= 2, "one or two target segments");
  SweTargets empty;
  const SweTargets& t0 = segs[0];
  const SweTargets& t1 = nseg > 1 ? segs[1] : empty;
  if (t0.base.n + t1.base.n == 0) return;
  constexpr int kBlock = 256;
  SweSeg a0 = to_seg(t0, kBlock), a1 = to_seg(t1, kBlock);
  const int grid = a0.nblocks + a1.nblocks;
  const int do_pse = pse_eps > 0 ? 1 : 0;
  const double inv_e2 = do_pse ? 1.0 / (pse_eps * pse_eps) : 0.0;
  const double cut2 = do_pse ? kPseCut * pse_eps * pse_eps : 0.0;
  // SphereFaceSums skips j == i only for the singular kernel (eps == 0)
  const int skip_self = eps == 0.0 ? 1 : 0;
  double pairs = static_cast<double>(t0.base.n + t1.base.n) * src.capacity;
  const SweParams P{eps * eps, inv_e2, cut2, do_pse, skip_self};
  if (precision == LPM_FP64) {
    TimedLaunch tl(ctx, KID_SWE_F64, s, pairs);
    hipLaunchKernelGGL((swe_f64_kernel<kBlock>), dim3(grid), dim3(kBlock), 0, s,
                       static_cast<const SweSrc*>(src.data), src.bounds,
                       src.count, a0, a1, P);
  } else {
    TimedLaunch tl(ctx, KID_SWE_MIXED, s, pairs);
    hipLaunchKernelGGL((swe_mixed_kernel<kBlock>), dim3(grid), dim3(kBlock), 0,
                       s, static_cast<const SweSrc*>(src.data), src.bounds,
                       src.count, a0, a1, P);
  }
  LPM_HIP_CHECK(hipGetLastError());
}

uint64_t swe_ensemble_prepare(lpm_ctx* ctx, const EnsembleDesc& d) {
  const EnsLayout L = ensemble_plan(ctx, d, 256);
  const int64_t cap = std::max<int64_t>(d.total_padded_upper, kSweTile);
  ctx->swe_packed.ensure(sizeof(SweSrc) * cap);
  ctx->swe_bounds.ensure(sizeof(double4) * (cap / kSweTile));
  return L.serial;
}

void swe_ensemble_sums(lpm_ctx* ctx, const EnsembleDesc& d,
                       const lpm_swe_fields& f, CrdIn vx, CrdIn fx,
                       const double* fzeta, const double* fsigma,
                       const double* farea, const RankMap& map, double eps,
                       double pse_eps, int precision, unsigned flags,
                       hipStream_t s) {
  constexpr int kBlock = 256;
  swe_ensemble_prepare(ctx, d);
  const EnsLayout L =
      ensemble_layout(ctx, d, map, f.face_mask, kBlock, s);
  SweSrc* src = ctx->swe_packed.as<SweSrc>();
  double4* bounds = ctx->swe_bounds.as<double4>();
  const dim3 grid2(static_cast<unsigned>((d.max_member_faces + 255) / 256),
                   static_cast<unsigned>(d.n_members));
  swe_ens_pack_kernel<<<grid2, 256, 0, s>>>(fx.base, fx.es, fx.cs, fzeta,
                                             fsigma, farea, f.face_surf,
                                             map.inv, L.mem, src);
  LPM_HIP_CHECK(hipGetLastError());
  swe_ens_bounds_kernel<<<grid2, kSweTile, 0, s>>>(src, L.mem, bounds);
  LPM_HIP_CHECK(hipGetLastError());
  if (L.n_items == 0) return;

  SweEnsArgs A{};
  A.mem = L.mem;
  A.items = L.items;
  A.n_items = L.n_items;
  A.queue = L.queue;
  A.vert.x = vx.base;
  A.vert.es = vx.es;
  A.vert.cs = vx.cs;
  A.vert.surf = f.vert_surf;
  A.vert.out_u = f.vert_u;
  A.vert.oes = f.vert_u_es;
  A.vert.ocs = f.vert_u_cs;
  A.vert.out_ddot = f.vert_ddot;
  A.vert.out_laps = f.vert_laps;
  A.face.x = fx.base;
  A.face.es = fx.es;
  A.face.cs = fx.cs;
  A.face.surf = f.face_surf;
  A.face.out_u = f.face_u;
  A.face.oes = f.face_u_es;
  A.face.ocs = f.face_u_cs;
  A.face.out_ddot = f.face_ddot;
  A.face.out_laps = f.face_laps;
  A.rank = map.rank;
  A.inv = map.inv;
  A.skip_masked = (flags & LPM_FLAG_SKIP_MASKED_TARGETS) ? 1 : 0;
  const int do_pse = pse_eps > 0 ? 1 : 0;
  A.P = SweParams{eps * eps, do_pse ? 1.0 / (pse_eps * pse_eps) : 0.0,
                  do_pse ? kPseCut * pse_eps * pse_eps : 0.0, do_pse,
                  eps == 0.0 ? 1 : 0};
  ensemble_reset_queue(L, s);
  if (precision == LPM_FP64) {
    const int grid = ensemble_grid(
        ctx, reinterpret_cast<const void*>(&swe_ens_f64_kernel<kBlock>),
        kBlock, L.n_items);
    TimedLaunch tl(ctx, KID_SWE_F64, s, L.pairs);
    hipLaunchKernelGGL((swe_ens_f64_kernel<kBlock>), dim3(grid), dim3(kBlock),
                       0, s, static_cast<const SweSrc*>(src),
                       static_cast<const double4*>(bounds), A);
  } else {
    const int grid = ensemble_grid(
        ctx, reinterpret_cast<const void*>(&swe_ens_mixed_kernel<kBlock>),
        kBlock, L.n_items);
    TimedLaunch tl(ctx, KID_SWE_MIXED, s, L.pairs);
    hipLaunchKernelGGL((swe_ens_mixed_kernel<kBlock>), dim3(grid),
                       dim3(kBlock), 0, s, static_cast<const SweSrc*>(src),
                       static_cast<const double4*>(bounds), A);
  }
  LPM_HIP_CHECK(hipGetLastError());
}

}  // namespace lpm_amd
