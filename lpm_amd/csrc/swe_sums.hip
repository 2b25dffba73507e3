// Sphere shallow-water direct sums for gfx950: velocity from relative
// vorticity and divergence, the velocity-gradient double dot
// grad(u):grad(u)^T, and the PSE surface Laplacian, per target.  Replaces
// SphereVertexSums / SphereFaceSums with SphereSweDirectSumReducer and
// sphere_swe_velocity_sums (src/lpm_swe_kernels.hpp:334-362, 578-619,
// 723-780, 877-930) plus the Gaussian PSE Laplacian of LaplacianReducer
// (src/lpm_pse.hpp:157-203) on chord distances.
//
// Per pair (target x, source y) with t = x.y, D = 1 - t + eps^2,
// a = zeta A / 4pi, b = sigma A / 4pi and g = y - t x (= (I - x x^T) y):
//   alpha = a / D, beta = b / D, a' = alpha / D, b' = beta / D.
// The reference's 12-tuple sums factor into per-target accumulators
//   Vy = sum alpha y,  Gd = sum a' g g^T,  Vg = sum beta g,
//   Bt = sum beta t,   Gs = sum b' g g^T,
// and the finalisation applies the target-only factors once:
//   u  = -x cross Vy - Vg
//   Mz = E(Vy) + [x]x Gd - (x cross Vy) x^T        (E_ab = e_abc v_c)
//   Ms = Bt (I - x x^T) + x Vg^T - Gs
//   M  = Mz + Ms  (the reference's packed gradient, whose vorticity part
//                  carries grad_kzeta's opposite sign convention)
//   ddot = sum_ab M_ab M_ba.
// These are identities of the pair sums for any x (no |x| = 1 needed):
// with d = y - x and delta = 1 - t, d + delta x = g, x cross d = x cross g,
// d - eps^2 x = g - D x and a' D = alpha, so the vorticity gradient
// sum a' (x cross d)(d - eps^2 x)^T is [x]x Gd - (x cross Vy) x^T, and the
// divergence terms sum b' (d + delta x)(d + delta x)^T, sum beta (d + delta
// x) and sum beta (1 - delta) are Gs, Vg and Bt.  Per pair that is 17 FMAs
// of accumulation and one reciprocal (shared by four pairs in FP64).
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

// FP64 accumulators of one target (see the file comment); Gd and Gs hold
// the symmetric matrices as xx, xy, xz, yy, yz, zz.
struct SweAcc {
  double Vy[3], Gd[6], Vg[3], Bt, Gs[6];
  double L1;  // PSE sum
  __device__ void zero() {
    Bt = L1 = 0.0;
    for (int k = 0; k < 3; ++k) Vy[k] = Vg[k] = 0.0;
    for (int k = 0; k < 6; ++k) Gd[k] = Gs[k] = 0.0;
  }
};

// FP32 partials of one source tile: Vy is carried as A0 = sum alpha and
// Vd = sum alpha d with d = y - x (Vy = x A0 + Vd), which keeps the
// tile-relative precision of d.
struct SweAcc32 {
  float A0, Vd[3], Gd[6], Vg[3], Bt, Gs[6];
  __device__ void zero() {
    A0 = Bt = 0.0f;
    for (int k = 0; k < 3; ++k) Vd[k] = Vg[k] = 0.0f;
    for (int k = 0; k < 6; ++k) Gd[k] = Gs[k] = 0.0f;
  }
};

// Gd += ap g g^T, Vg += be g, Bt += be t, Gs += bp g g^T.
template <typename T, typename Acc>
__device__ __forceinline__ void swe_pair_g(Acc& acc, const T g0, const T g1,
                                           const T g2, const T t, const T be,
                                           const T ap, const T bp) {
  const T e0 = ap * g0, e1 = ap * g1, e2 = ap * g2;
  acc.Gd[0] = fma(e0, g0, acc.Gd[0]);
  acc.Gd[1] = fma(e0, g1, acc.Gd[1]);
  acc.Gd[2] = fma(e0, g2, acc.Gd[2]);
  acc.Gd[3] = fma(e1, g1, acc.Gd[3]);
  acc.Gd[4] = fma(e1, g2, acc.Gd[4]);
  acc.Gd[5] = fma(e2, g2, acc.Gd[5]);
  acc.Vg[0] = fma(be, g0, acc.Vg[0]);
  acc.Vg[1] = fma(be, g1, acc.Vg[1]);
  acc.Vg[2] = fma(be, g2, acc.Vg[2]);
  acc.Bt = fma(be, t, acc.Bt);
  const T f0 = bp * g0, f1 = bp * g1, f2 = bp * g2;
  acc.Gs[0] = fma(f0, g0, acc.Gs[0]);
  acc.Gs[1] = fma(f0, g1, acc.Gs[1]);
  acc.Gs[2] = fma(f0, g2, acc.Gs[2]);
  acc.Gs[3] = fma(f1, g1, acc.Gs[3]);
  acc.Gs[4] = fma(f1, g2, acc.Gs[4]);
  acc.Gs[5] = fma(f2, g2, acc.Gs[5]);
}

// FP64 pair: source y, t = x.y, r = 1 / D.
__device__ __forceinline__ void swe_pair_f64(SweAcc& acc, const double4& y,
                                             const double t, const double r,
                                             const double b, const double* x) {
  const double al = y.w * r, be = b * r;
  const double ap = al * r, bp = be * r;
  acc.Vy[0] = fma(al, y.x, acc.Vy[0]);
  acc.Vy[1] = fma(al, y.y, acc.Vy[1]);
  acc.Vy[2] = fma(al, y.z, acc.Vy[2]);
  const double g0 = fma(-t, x[0], y.x), g1 = fma(-t, x[1], y.y),
               g2 = fma(-t, x[2], y.z);
  swe_pair_g<double>(acc, g0, g1, g2, t, be, ap, bp);
}

// FP32 pair on tile-relative offsets: d = y - x, rho2 = |d|^2 (= 2 delta on
// the unit sphere), r = 1 / D; xh = x / 2, so g = d + delta x = d + rho2 xh
// and t = 1 - delta.
__device__ __forceinline__ void swe_pair_f32(SweAcc32& acc, const float d0,
                                             const float d1, const float d2,
                                             const float rho2, const float r,
                                             const float a, const float b,
                                             const float* xh) {
  const float al = a * r, be = b * r;
  const float ap = al * r, bp = be * r;
  acc.A0 += al;
  acc.Vd[0] = fmaf(al, d0, acc.Vd[0]);
  acc.Vd[1] = fmaf(al, d1, acc.Vd[1]);
  acc.Vd[2] = fmaf(al, d2, acc.Vd[2]);
  const float g0 = fmaf(rho2, xh[0], d0), g1 = fmaf(rho2, xh[1], d1),
              g2 = fmaf(rho2, xh[2], d2);
  const float t = fmaf(-0.5f, rho2, 1.0f);
  swe_pair_g<float>(acc, g0, g1, g2, t, be, ap, bp);
}

// Adds one tile's FP32 partials to the FP64 accumulators of target x.
__device__ __forceinline__ void swe_carry(SweAcc& acc, const SweAcc32& p,
                                          const double* x) {
  for (int k = 0; k < 3; ++k) {
    acc.Vy[k] += fma(x[k], static_cast<double>(p.A0),
                     static_cast<double>(p.Vd[k]));
    acc.Vg[k] += p.Vg[k];
  }
  acc.Bt += p.Bt;
  for (int k = 0; k < 6; ++k) {
    acc.Gd[k] += p.Gd[k];
    acc.Gs[k] += p.Gs[k];
  }
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
__device__ __forceinline__ void swe_finalize(const SweAcc& A, const double* x,
                                             double pse_scale, double* u,
                                             double* ddot, double* laps) {
  const double* V = A.Vy;
  const double xv[3] = {x[1] * V[2] - x[2] * V[1], x[2] * V[0] - x[0] * V[2],
                        x[0] * V[1] - x[1] * V[0]};
  u[0] = -xv[0] - A.Vg[0];
  u[1] = -xv[1] - A.Vg[1];
  u[2] = -xv[2] - A.Vg[2];

  double Gd[3][3], Gs[3][3];
  sym(A.Gd, Gd);
  sym(A.Gs, Gs);
  // [x]x Gd: row a = (x cross column)_a
  double M[3][3];
#pragma unroll
  for (int b = 0; b < 3; ++b) {
    M[0][b] = x[1] * Gd[2][b] - x[2] * Gd[1][b];
    M[1][b] = x[2] * Gd[0][b] - x[0] * Gd[2][b];
    M[2][b] = x[0] * Gd[1][b] - x[1] * Gd[0][b];
  }
  // E(Vy): [[0, v2, -v1], [-v2, 0, v0], [v1, -v0, 0]]
  M[0][1] += V[2];
  M[0][2] -= V[1];
  M[1][0] -= V[2];
  M[1][2] += V[0];
  M[2][0] += V[1];
  M[2][1] -= V[0];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double P = (a == b ? 1.0 : 0.0) - x[a] * x[b];
      M[a][b] += A.Bt * P + x[a] * A.Vg[b] - Gs[a][b] - xv[a] * x[b];
    }
  double dd = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) dd += M[a][b] * M[b][a];
  *ddot = dd;
  *laps = pse_scale * A.L1;
}

// FP64 pair terms of one LDS tile (cnt sources, zero padding up to a
// multiple of SHARE) for target x; `self` is the target's own source in tile
// numbering (outside [0, cnt): none) and any_self says whether some lane of
// the wave has one here.  Without, SHARE sources share one reciprocal,
// e.g. 1/D_a = D_b (D_c D_d) / (D_a D_b D_c D_d).
__device__ __forceinline__ double swe_t(const double* x, const double4& y) {
  return fma(x[2], y.z, fma(x[1], y.y, x[0] * y.x));
}

template <int SHARE = 4>
__device__ __forceinline__ void swe_tile_f64(SweAcc& acc, const SweSrc* tile,
                                             int cnt, const double* x,
                                             double eps2, int self,
                                             bool any_self) {
  static_assert(SHARE == 2 || SHARE == 4, "two or four pairs per reciprocal");
  if (any_self) {
    for (int jj = 0; jj < cnt; ++jj) {
      const double4 yp = tile[jj].p;
      const double t = swe_t(x, yp);
      double r = rcp_nr<1>((1.0 - t) + eps2);
      if (jj == self) r = 0.0;
      swe_pair_f64(acc, yp, t, r, tile[jj].q.x, x);
    }
  } else if (SHARE == 2) {
    for (int jj = 0; jj < cnt; jj += 2) {
      const double4 ya = tile[jj].p, yb = tile[jj + 1].p;
      const double ta = swe_t(x, ya), tb = swe_t(x, yb);
      const double Da = (1.0 - ta) + eps2, Db = (1.0 - tb) + eps2;
      const double r = rcp_nr<1>(Da * Db);
      swe_pair_f64(acc, ya, ta, Db * r, tile[jj].q.x, x);
      swe_pair_f64(acc, yb, tb, Da * r, tile[jj + 1].q.x, x);
    }
  } else {
    for (int jj = 0; jj < cnt; jj += 4) {
      const double4 ya = tile[jj].p, yb = tile[jj + 1].p;
      const double4 yc = tile[jj + 2].p, yd = tile[jj + 3].p;
      const double ta = swe_t(x, ya), tb = swe_t(x, yb);
      const double tc = swe_t(x, yc), td = swe_t(x, yd);
      const double Da = (1.0 - ta) + eps2, Db = (1.0 - tb) + eps2;
      const double Dc = (1.0 - tc) + eps2, Dd = (1.0 - td) + eps2;
      const double pab = Da * Db, pcd = Dc * Dd;
      const double r = rcp_nr<1>(pab * pcd);
      const double rab = pcd * r, rcd = pab * r;
      swe_pair_f64(acc, ya, ta, Db * rab, tile[jj].q.x, x);
      swe_pair_f64(acc, yb, tb, Da * rab, tile[jj + 1].q.x, x);
      swe_pair_f64(acc, yc, tc, Dd * rcd, tile[jj + 2].q.x, x);
      swe_pair_f64(acc, yd, td, Dc * rcd, tile[jj + 3].q.x, x);
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
template <int BLOCK, int SHARE>
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

  SweAcc acc;
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
    swe_tile_f64<SHARE>(acc, tile, cnt, x, eps2, self - j0,
                        __ballot(diag) != 0);
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
  swe_finalize(acc, x, k4OverPi * inv_pse_e2 * inv_pse_e2, u, &dd, &lap);
  if (sg.out_u) {
    double* o = sg.out_u + i * sg.oes;
    o[0] = u[0];
    o[sg.ocs] = u[1];
    o[2 * sg.ocs] = u[2];
  }
  if (sg.out_ddot) sg.out_ddot[i] = dd;
  if (sg.out_laps && do_pse) sg.out_laps[i] = lap;
}

// SHARE: pairs per FP64 reciprocal; MINW: minimum waves per SIMD asked of
// the register allocator.
template <int BLOCK, int SHARE, int MINW>
__global__ __launch_bounds__(BLOCK, MINW) void swe_f64_kernel(
    const SweSrc* __restrict__ src, const double4* __restrict__ bounds,
    const int32_t* __restrict__ count, SweSeg s0, SweSeg s1, SweParams P) {
  __shared__ SweSrc tile[kSweTile];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SweSeg& sg = first ? s0 : s1;
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  swe_f64_block<BLOCK, SHARE>(src, bounds, *count, sg, blk, P, tile);
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

  SweAcc acc;
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
      swe_tile_f64<2>(acc, tile64, cnt, x, eps2, self - j0,
                      __ballot(diag) != 0);
    } else {
      SweAcc32 p;
      p.zero();
      const float xh[3] = {0.5f * static_cast<float>(x[0]),
                           0.5f * static_cast<float>(x[1]),
                           0.5f * static_cast<float>(x[2])};
      // the per-pair self test only in tiles holding some lane's own source;
      // the padding sources (far away, zero strengths) add nothing, so the
      // loop runs to a multiple of four and unrolls
      const int cnt4 = (cnt + 3) & ~3;
      auto pairs = [&](auto check) {
#pragma unroll 4
        for (int jj = 0; jj < cnt4; ++jj) {
          const float4 yp = tile[jj].p;
          const float4 yq = tile[jj].q;
          const float d0 = yp.x - xr[0], d1 = yp.y - xr[1], d2 = yp.z - xr[2];
          const float rho2 = fmaf(d2, d2, fmaf(d1, d1, d0 * d0));
          float r = __builtin_amdgcn_rcpf(fmaf(0.5f, rho2, e2f));
          if (decltype(check)::value && j0 + jj == self) r = 0.0f;
          swe_pair_f32(p, d0, d1, d2, rho2, r, yp.w, yq.x, xh);
        }
      };
      if (__ballot(diag))
        pairs(std::true_type{});
      else
        pairs(std::false_type{});
      // carry the tile's FP32 partials into the FP64 accumulators
      swe_carry(acc, p, x);
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
  swe_finalize(acc, x, k4OverPi * inv_pse_e2 * inv_pse_e2, u, &dd, &lap);
  if (sg.out_u) {
    double* o = sg.out_u + i * sg.oes;
    o[0] = u[0];
    o[sg.ocs] = u[1];
    o[2 * sg.ocs] = u[2];
  }
  if (sg.out_ddot) sg.out_ddot[i] = dd;
  if (sg.out_laps && do_pse) sg.out_laps[i] = lap;
}

// MINW: minimum waves per SIMD asked of the register allocator (the FP64
// accumulators, the FP32 tile partials and the FP64 near-tile and PSE loops
// hold about 170 registers at 3 waves).
template <int BLOCK, int MINW>
__global__ __launch_bounds__(BLOCK, MINW) void swe_mixed_kernel(
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

template <int BLOCK, int SHARE, int MINW>
__global__ __launch_bounds__(BLOCK, MINW) void swe_ens_f64_kernel(
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
    swe_f64_block<BLOCK, SHARE>(src + M.P, bounds + M.P / kSweTile, M.cnt,
                                sg, w.z, A.P, tile);
  }
}

template <int BLOCK, int MINW>
__global__ __launch_bounds__(BLOCK, MINW) void swe_ens_mixed_kernel(
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

// ------------------------------------------------------ kernel registry ---
using SweKernel = void (*)(const SweSrc*, const double4*, const int32_t*,
                           SweSeg, SweSeg, SweParams);
using SweEnsKernel = void (*)(const SweSrc*, const double4*, SweEnsArgs);
struct SweVariant {
  SweKernel fn;
  SweEnsKernel ens;
  const char* name;
};
constexpr int kSweBlock = 256;
const SweVariant kSweF64Variants[] = {
    {swe_f64_kernel<kSweBlock, 4, 3>, swe_ens_f64_kernel<kSweBlock, 4, 3>,
     "share4_occ3"},
    {swe_f64_kernel<kSweBlock, 4, 4>, swe_ens_f64_kernel<kSweBlock, 4, 4>,
     "share4_occ4"},
    {swe_f64_kernel<kSweBlock, 2, 4>, swe_ens_f64_kernel<kSweBlock, 2, 4>,
     "share2_occ4"},
    {swe_f64_kernel<kSweBlock, 2, 3>, swe_ens_f64_kernel<kSweBlock, 2, 3>,
     "share2_occ3"},
};
const SweVariant kSweMixedVariants[] = {
    {swe_mixed_kernel<kSweBlock, 4>, swe_ens_mixed_kernel<kSweBlock, 4>,
     "occ4"},
    {swe_mixed_kernel<kSweBlock, 3>, swe_ens_mixed_kernel<kSweBlock, 3>,
     "occ3"},
    {swe_mixed_kernel<kSweBlock, 2>, swe_ens_mixed_kernel<kSweBlock, 2>,
     "occ2"},
};
constexpr int kNumSweF64 = sizeof(kSweF64Variants) / sizeof(kSweF64Variants[0]);
constexpr int kNumSweMixed =
    sizeof(kSweMixedVariants) / sizeof(kSweMixedVariants[0]);

const SweVariant& swe_variant(lpm_ctx* ctx, int precision, int64_t n_targets) {
  const int family = precision == LPM_FP64 ? KID_SWE_F64 : KID_SWE_MIXED;
  const int v = pick_variant(ctx, family, n_targets);
  if (precision == LPM_FP64) {
    LPM_ARG_CHECK(v >= 0 && v < kNumSweF64, "variant out of range");
    return kSweF64Variants[v];
  }
  LPM_ARG_CHECK(v >= 0 && v < kNumSweMixed, "variant out of range");
  return kSweMixedVariants[v];
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

int swe_variant_count(int family) {
  return family == KID_SWE_F64 ? kNumSweF64
                               : (family == KID_SWE_MIXED ? kNumSweMixed : 0);
}

const char* swe_variant_name(int family, int v) {
  if (v < 0 || v >= swe_variant_count(family)) return nullptr;
  return family == KID_SWE_F64 ? kSweF64Variants[v].name
                               : kSweMixedVariants[v].name;
}

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
  const SweVariant& v = swe_variant(ctx, precision, t0.base.n + t1.base.n);
  {
    TimedLaunch tl(ctx, precision == LPM_FP64 ? KID_SWE_F64 : KID_SWE_MIXED, s,
                   pairs);
    hipLaunchKernelGGL(v.fn, dim3(grid), dim3(kBlock), 0, s,
                       static_cast<const SweSrc*>(src.data), src.bounds,
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
  const SweVariant& v = swe_variant(ctx, precision, d.n_verts + d.n_faces);
  const int grid = ensemble_grid(ctx, reinterpret_cast<const void*>(v.ens),
                                 kBlock, L.n_items);
  {
    TimedLaunch tl(ctx, precision == LPM_FP64 ? KID_SWE_F64 : KID_SWE_MIXED, s,
                   L.pairs);
    hipLaunchKernelGGL(v.ens, dim3(grid), dim3(kBlock), 0, s,
                       static_cast<const SweSrc*>(src),
                       static_cast<const double4*>(bounds), A);
  }
  LPM_HIP_CHECK(hipGetLastError());
}

}  // namespace lpm_amd
