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
// beyond sqrt(kPseCut) e of every target of a wave are skipped, and so are
// the pairs beyond that radius in the others (their terms are below
// exp(-kPseCut) relative), so the PSE work covers a few percent of the pairs.
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
    int64_t n_pad, SweSrc* __restrict__ out) {
  const int64_t r = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  if (r >= n_pad) return;
  SweSrc s;
  s.p = make_double4(0.0, 0.0, 0.0, 0.0);
  s.q = make_double4(0.0, 0.0, 0.0, 0.0);
  if (r < *count) {
    // zero records pad the last tile: kernels may read whole tiles
    const int64_t i = inv ? inv[r] : r;
    const double* p = sx + i * es;
    const double A = area[i];
    s.p = make_double4(p[0], p[cs], p[2 * cs], zeta[i] * A * kInv4Pi);
    s.q = make_double4(sigma[i] * A * kInv4Pi, A, surf ? surf[i] : 0.0, 0.0);
  }
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

// PSE terms of one tile (cnt sources, zero records up to a multiple of 4)
// for target x with surface s_i, added to L: only pairs within the
// Gaussian's cut-off radius (rho^2 < pse_cut2); the others are below
// exp(-kPseCut) relative.  Positions are read four at a time.
__device__ __forceinline__ void swe_pse_tile(const SweSrc* tile, int cnt,
                                             const double* x, double s_i,
                                             double inv_pse_e2,
                                             double pse_cut2, double& L) {
  for (int j = 0; j < cnt; j += 4) {
    double4 yp[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) yp[k] = tile[j + k].p;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const double d0 = yp[k].x - x[0], d1 = yp[k].y - x[1],
                   d2 = yp[k].z - x[2];
      const double rho2 = d0 * d0 + d1 * d1 + d2 * d2;
      if (j + k < cnt && rho2 < pse_cut2) {
        const double4 yq = tile[j + k].q;
        double eta = exp(-rho2 * inv_pse_e2) * yq.y;
        if (!hemisphere(x, yp[k])) eta = 0.0;
        L = fma(eta, yq.z - s_i, L);
      }
    }
  }
}

// p[t] through vector loads (the index is held in a vector register), so
// that a load issued one iteration ahead is waited for in order.
__device__ __forceinline__ double4 vload(const double4* p, int t) {
  int v = t;
  asm volatile("" : "+v"(v));
  return p[v];
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
  double* far;         // two-pass matrix-core sums: row r of slot c at
  int64_t far_stride;  // far[r * far_stride + c]
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

constexpr double kSweMfNear = 3.0;
constexpr int kSweMfSub = 64;                     // sources per sub-tile
constexpr int kSweMfNsub = kSweTile / kSweMfSub;  // sub-tiles per tile

// Whether target x lies within kSweMfNear radii of the centroid of a tile
// or sub-tile (c.w = radius^2): the near / far split of the matrix-core
// kernels.  Contraction is off so that every kernel rounds it alike: both
// passes classify every tile and sub-tile, and must agree.
__device__ __forceinline__ bool swe_mf_near(const double* x, const double4& c) {
#pragma clang fp contract(off)
  const double e0 = x[0] - c.x, e1 = x[1] - c.y, e2 = x[2] - c.z;
  return (e0 * e0 + e1 * e1) + e2 * e2 <= (kSweMfNear * kSweMfNear) * c.w;
}
constexpr int kSweMfR1 = 6;   // set-1 feature rows
constexpr int kSweMfR2 = 12;  // set-2 feature rows
constexpr int kSweMfKb = kSweTile / 32;
constexpr int kSweMfFr = kSweMfKb * 2 * 4 * kSweMfR1;   // uint4 per tile
constexpr int kSweMfFr2 = kSweMfKb * 2 * 4 * kSweMfR2;  // uint4 per tile
constexpr int64_t kSweMfTileBytes = 32 + 32 * kSweMfNsub + 32 * 8 +
                                    2 * kSweTile * 16 +
                                    (kSweMfFr + kSweMfFr2) * 16;

// Matrix-core records of consecutive tiles (see the matrix-core section).
struct SweMfTiles {
  double4* cen;   // per tile: centroid, radius^2
  double4* scen;  // per sub-tile: centroid, radius^2 (-1: no sources)
  double* fs;     // 32 per tile
  float4* ya;     // per source, relative to the tile's centroid
  float4* yas;    // per source, relative to its sub-tile's centroid
  uint4* fr;
  uint4* fr2;
  // the records of tiles t0 and later
  __host__ __device__ SweMfTiles from(int64_t t0) const {
    return SweMfTiles{cen + t0,           scen + kSweMfNsub * t0,
                      fs + 32 * t0,       ya + kSweTile * t0,
                      yas + kSweTile * t0, fr + kSweMfFr * t0,
                      fr2 + kSweMfFr2 * t0};
  }
};

// ----------------------------------------------------------- FP64 kernel --
// One target per lane; sources staged through LDS.  PSE terms are evaluated
// only for tiles near some target of the wave.
// Parameters shared by every target of a launch.
struct SweParams {
  double eps2, inv_pse_e2, pse_cut2;
  int do_pse, skip_self;
  SweMfTiles mf{};  // matrix-core variant: the tile records
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
        swe_pse_tile(tile, cnt, x, s_i, inv_pse_e2, pse_cut2, acc.L1);
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
        swe_pse_tile(tile64, cnt, x, s_i, inv_pse_e2, pse_cut2, L);
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

// ------------------------------------------------ matrix-core far field --
// Mixed precision on the matrix cores, in two passes.  A tile or 64-source
// sub-tile is far from a wave when no target of the wave lies within
// kSweMfNear radii of its centroid c (swe_mf_near).  The first pass sums
// every far tile, and the far sub-tiles of the other tiles, per 32 sources
// and 16 targets in two stages:
//   stage 1 (v_mfma_f32_16x16x4_f32): sigma D for 16 sources x 16 targets,
//     D = |x' - y'|^2 / 2 + eps^2 = |x'|^2 / 2 - x'.y' + |y'|^2 / 2 + eps^2
//     on FP32 offsets x' = x - c, y' = y - c:  A_j = (y'_j, |y'_j|^2 / 2),
//     B_i = sigma (-x'_i, 1), C_i = sigma (|x'_i|^2 / 2 + eps^2);
//   vector ALUs: w = 1 / (sigma D) and w^2, each split into f16 hi + lo;
//   stage 2 (v_mfma_f32_16x16x32_f16): per target, the sums over 32 sources
//     of w times six source features and of w^2 times twelve, as
//     F_hi w_hi + F_lo w_hi + F_hi w_lo (22 significant bits).
// With a = zeta A / 4pi, b = sigma A / 4pi and absolute coordinates y the
// features are
//   set 1 (weight 1 / D):    a y0, a y1, a y2, b y0, b y1, b y2
//   set 2 (weight 1 / D^2):  a y y^T (xx, xy, xz, yy, yz, zz), b y y^T
// and the far sums follow from them by target-only projections
// (g = Q y, Q = I - x x^T):
//   Vy += set 1 rows 0..2;  Vg += Q (rows 3..5);  Bt += x . (rows 3..5);
//   Gd += Q (set 2 rows 0..5) Q;  Gs += Q (rows 6..11) Q.
// The projections are linear, so each lane carries the FP32 sums into FP64
// per tile or sub-tile and the 18 raw sums of every target go to a
// workspace; the second pass sums the near sub-tiles in FP64 (and the PSE
// terms), adds the projected far sums and writes the outputs.  Every feature
// row of a tile is scaled by a power of two to a maximum in [2^13, 2^14), and
// sigma, a power of two per 16 targets, bounds w by 2^7 and w^2 by 2^14 from
// the group's smallest D over the (sub-)tile's bounding sphere, so no f16
// operand overflows.  Offsets are taken from the centroid of the tile or
// sub-tile being summed, so the FP32 distance form cancels by at most
// (kSweMfNear + 1)^2 / (kSweMfNear - 1)^2 against its terms.
//
// Tile records (one 256-source tile each):
//   cen        double4  (c, R^2)
//   scen[4]    double4  sub-tile (c_s, r_s^2), r_s^2 = -1 without sources
//   fs[32]     double   inverse row scales: set 1 at 0..5, set 2 at 16..27
//                       (zero for the unused rows)
//   ya[256]    float4   (y - c, |y - c|^2 / 2); padding (8, 8, 8, 96)
//   yas[256]   float4   the same relative to the source's sub-tile centroid
//   fr[384]    uint4    set 1, [k-block 8][hi/lo][g 4][row 6] x 8 f16
//   fr2[768]   uint4    set 2, [k-block 8][hi/lo][g 4][row 12] x 8 f16
// Slot i of K-group g in a 32-source k-block is source 4 g + i (i < 4) or
// 16 + 4 g + i - 4: the rows that lane (g, target) holds after stage 1.

// Records of ntile tiles carved from one buffer (kSweMfTileBytes each).
SweMfTiles swe_mf_tiles(void* base, int64_t ntile) {
  unsigned char* p = static_cast<unsigned char*>(base);
  SweMfTiles T;
  T.cen = reinterpret_cast<double4*>(p);
  p += 32 * ntile;
  T.scen = reinterpret_cast<double4*>(p);
  p += 32 * kSweMfNsub * ntile;
  T.fs = reinterpret_cast<double*>(p);
  p += 256 * ntile;
  T.ya = reinterpret_cast<float4*>(p);
  p += 16 * kSweTile * ntile;
  T.yas = reinterpret_cast<float4*>(p);
  p += 16 * kSweTile * ntile;
  T.fr = reinterpret_cast<uint4*>(p);
  p += 16 * kSweMfFr * ntile;
  T.fr2 = reinterpret_cast<uint4*>(p);
  return T;
}

// Block-wide max / sum of v over 256 threads (red: 4 doubles of LDS); every
// thread gets the result.
__device__ __forceinline__ double block_max256(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
}

__device__ __forceinline__ double block_sum256(double v, double* red) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  return (red[0] + red[1]) + (red[2] + red[3]);
}

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// (y', |y'|^2 / 2) in FP32 for offset e (FP64); q from the rounded offset.
__device__ __forceinline__ float4 swe_mf_offset(double e0, double e1,
                                                double e2) {
  const float f0 = static_cast<float>(e0), f1 = static_cast<float>(e1),
              f2 = static_cast<float>(e2);
  const double q = static_cast<double>(f0) * f0 +
                   static_cast<double>(f1) * f1 + static_cast<double>(f2) * f2;
  return make_float4(f0, f1, f2, static_cast<float>(0.5 * q));
}

// The records of one tile from its cnt (> 0) packed sources at src; a block
// of 256 threads, thread jj takes source jj (wave s: sub-tile s).
__device__ __forceinline__ void swe_mf_pack_tile(const SweSrc* __restrict__ src,
                                                 int cnt, const SweMfTiles& T) {
  __shared__ double red[4];
  __shared__ double scale[32];
  const int jj = threadIdx.x;
  const bool real = jj < cnt;
  SweSrc s;
  s.p = make_double4(0.0, 0.0, 0.0, 0.0);
  s.q = make_double4(0.0, 0.0, 0.0, 0.0);
  if (real) s = src[jj];
  const double m = static_cast<double>(cnt);
  const double c0 = block_sum256(s.p.x, red) / m;
  const double c1 = block_sum256(s.p.y, red) / m;
  const double c2 = block_sum256(s.p.z, red) / m;
  const double e0 = s.p.x - c0, e1 = s.p.y - c1, e2 = s.p.z - c2;
  const double R2 = block_max256(real ? e0 * e0 + e1 * e1 + e2 * e2 : 0.0, red);
  if (jj == 0) T.cen[0] = make_double4(c0, c1, c2, R2);
  T.ya[jj] = real ? swe_mf_offset(e0, e1, e2)
                  : make_float4(8.0f, 8.0f, 8.0f, 96.0f);

  // this wave's sub-tile
  const int sub = jj / kSweMfSub;
  const int ns = min(max(cnt - sub * kSweMfSub, 0), kSweMfSub);
  if (ns > 0) {
    const double ms = static_cast<double>(ns);
    const double s0 = wave_sum(s.p.x) / ms, s1 = wave_sum(s.p.y) / ms,
                 s2 = wave_sum(s.p.z) / ms;
    const double f0 = s.p.x - s0, f1 = s.p.y - s1, f2 = s.p.z - s2;
    double r2 = real ? f0 * f0 + f1 * f1 + f2 * f2 : 0.0;
    for (int off = 32; off > 0; off >>= 1) r2 = fmax(r2, __shfl_xor(r2, off, 64));
    if ((jj & 63) == 0) T.scen[sub] = make_double4(s0, s1, s2, r2);
    T.yas[jj] = real ? swe_mf_offset(f0, f1, f2)
                     : make_float4(8.0f, 8.0f, 8.0f, 96.0f);
  } else {
    if ((jj & 63) == 0) T.scen[sub] = make_double4(0.0, 0.0, 0.0, -1.0);
    T.yas[jj] = make_float4(8.0f, 8.0f, 8.0f, 96.0f);
  }

  const double y0 = s.p.x, y1 = s.p.y, y2 = s.p.z, a = s.p.w, b = s.q.x;
  double F[kSweMfR1 + kSweMfR2] = {
      a * y0,      a * y1,      a * y2,      b * y0,      b * y1,
      b * y2,      a * y0 * y0, a * y0 * y1, a * y0 * y2, a * y1 * y1,
      a * y1 * y2, a * y2 * y2, b * y0 * y0, b * y0 * y1, b * y0 * y2,
      b * y1 * y1, b * y1 * y2, b * y2 * y2};
  // per-row power-of-two scales: row maximum into [2^13, 2^14)
#pragma unroll
  for (int r = 0; r < kSweMfR1 + kSweMfR2; ++r) {
    const double mx = block_max256(fabs(F[r]), red);
    const int k = mx > 0.0 ? 14 - ilogb(mx) - 1 : 0;
    if (jj == 0) scale[r] = mx > 0.0 ? ldexp(1.0, -k) : 0.0;
    F[r] = ldexp(F[r], k);
  }
  __syncthreads();
  if (jj < 32) {
    double v = 0.0;
    if (jj < kSweMfR1) v = scale[jj];
    if (jj >= 16 && jj < 16 + kSweMfR2) v = scale[kSweMfR1 + jj - 16];
    T.fs[jj] = v;
  }
  // slot (g, i) of k-block kb that holds source jj
  const int kb = jj >> 5, sl = jj & 31;
  const int g = (sl & 15) >> 2, i = (sl & 3) + 4 * (sl >> 4);
  _Float16* h1 = reinterpret_cast<_Float16*>(T.fr);
  _Float16* h2 = reinterpret_cast<_Float16*>(T.fr2);
#pragma unroll
  for (int r = 0; r < kSweMfR1 + kSweMfR2; ++r) {
    const _Float16 hi = static_cast<_Float16>(static_cast<float>(F[r]));
    const _Float16 lo =
        static_cast<_Float16>(static_cast<float>(F[r] - static_cast<double>(hi)));
    if (r < kSweMfR1) {
      h1[((((kb * 2 + 0) * 4 + g) * kSweMfR1) + r) * 8 + i] = hi;
      h1[((((kb * 2 + 1) * 4 + g) * kSweMfR1) + r) * 8 + i] = lo;
    } else {
      const int r2 = r - kSweMfR1;
      h2[((((kb * 2 + 0) * 4 + g) * kSweMfR2) + r2) * 8 + i] = hi;
      h2[((((kb * 2 + 1) * 4 + g) * kSweMfR2) + r2) * 8 + i] = lo;
    }
  }
}

// One block per tile of the packed sources j < *count.
__global__ __launch_bounds__(256) void swe_mf_pack_kernel(
    const SweSrc* __restrict__ src, const int32_t* __restrict__ count,
    SweMfTiles T) {
  const int n = *count;
  const int j0 = blockIdx.x * kSweTile;
  if (j0 >= n) return;
  swe_mf_pack_tile(src + j0, min(kSweTile, n - j0), T.from(blockIdx.x));
}

typedef float swe_mf4 __attribute__((ext_vector_type(4)));
typedef _Float16 swe_mh8 __attribute__((ext_vector_type(8)));
typedef _Float16 swe_mh2 __attribute__((ext_vector_type(2)));
typedef float swe_mf2 __attribute__((ext_vector_type(2)));

// Two FP32 values into packed f16 hi parts and packed f16 lo parts.
__device__ __forceinline__ void split_pair(float a, float b, uint32_t& hi,
                                           uint32_t& lo) {
  const swe_mh2 h = __builtin_convertvector(swe_mf2{a, b}, swe_mh2);
  const swe_mf2 r = {a - static_cast<float>(h[0]), b - static_cast<float>(h[1])};
  hi = __builtin_bit_cast(uint32_t, h);
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, swe_mh2));
}

// x minus the f16 in half H of h, in one v_fma_mix_f32 (x - hi is exact
// when hi is x rounded to f16).
template <int H>
__device__ __forceinline__ float minus_half(float x, uint32_t h) {
  float r;
  if constexpr (H == 0)
    asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel_hi:[0,0,1]"
        : "=v"(r) : "v"(x), "v"(h));
  else
    asm("v_fma_mix_f32 %0, %1, 1.0, -%2 op_sel:[0,0,1] op_sel_hi:[0,0,1]"
        : "=v"(r) : "v"(x), "v"(h));
  return r;
}

// split_pair with the differences in v_fma_mix_f32.
__device__ __forceinline__ void split_pair_mix(float a, float b, uint32_t& hi,
                                               uint32_t& lo) {
  hi = __builtin_bit_cast(
      uint32_t, __builtin_convertvector(swe_mf2{a, b}, swe_mh2));
  const swe_mf2 r = {minus_half<0>(a, hi), minus_half<1>(b, hi)};
  lo = __builtin_bit_cast(uint32_t, __builtin_convertvector(r, swe_mh2));
}

// Adds Q Y Q (Q = I - x x^T, Y symmetric: xx, xy, xz, yy, yz, zz) to G.
__device__ __forceinline__ void add_qyq(double* G, const double* Y,
                                        const double* x) {
  const double v0 = Y[0] * x[0] + Y[1] * x[1] + Y[2] * x[2];
  const double v1 = Y[1] * x[0] + Y[3] * x[1] + Y[4] * x[2];
  const double v2 = Y[2] * x[0] + Y[4] * x[1] + Y[5] * x[2];
  const double sv = x[0] * v0 + x[1] * v1 + x[2] * v2;
  const double v[3] = {v0, v1, v2};
  const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int e = 0; e < 6; ++e) {
    const int p = ia[e], q = ib[e];
    G[e] += Y[e] - x[p] * v[q] - v[p] * x[q] + sv * x[p] * x[q];
  }
}

// Far sums of k-blocks kb0 .. kb1 - 1 of one tile (records ya, fr, fr2, fs
// in global memory; ya relative to the centroid c of the tile or sub-tile
// summed) for the wave's groups G0 .. G0 + NG - 1 of 16 targets, carried
// into F1 / F2: lane (g, col) holds rows 4 g + v of target 16 G + col.  e is
// this lane's target minus c, dist2 = |e|^2, R2 = the radius^2 around c.
template <int NG>
__device__ __forceinline__ void swe_mf_far_tile(
    const float4* __restrict__ ya, const uint4* __restrict__ lfr,
    const uint4* __restrict__ lfr2, const double* __restrict__ lfs, int kb0,
    int kb1, bool in, double e0, double e1, double e2, double dist2, double R2,
    double eps2, int col, int g, int G0, double (*F1)[4], double (*F2)[4]) {
  const float e2f = static_cast<float>(eps2);
  // this lane's target as a stage-1 operand: (-x', 1) and |x'|^2 / 2 +
  // eps^2 (lanes past the end: D = 4 for every source)
  const float xp0 = static_cast<float>(e0), xp1 = static_cast<float>(e1),
              xp2 = static_cast<float>(e2);
  const float own[5] = {
      in ? -xp0 : 0.0f, in ? -xp1 : 0.0f, in ? -xp2 : 0.0f, in ? 1.0f : 0.0f,
      in ? fmaf(0.5f, fmaf(xp2, xp2, fmaf(xp1, xp1, xp0 * xp0)), e2f) : 4.0f};
  // smallest D over the sphere around c in each group of 16 targets
  const double gap = sqrt(dist2) - sqrt(R2);
  float dl = in ? static_cast<float>(0.5 * gap * gap + eps2) : 4.0f;
  dl = fmaxf(fminf(dl, 4.0f), 1e-30f);
#pragma unroll
  for (int off = 1; off < 16; off <<= 1) dl = fminf(dl, __shfl_xor(dl, off, 64));
  float sig[NG], Bv[NG], Cv[NG];
#pragma unroll
  for (int G = 0; G < NG; ++G) {
    const float dG = __shfl(dl, 16 * (G0 + G), 64);
    // sigma dl in [2^-7, 2^-6): w = 1 / (sigma D) <= 2^7
    sig[G] = __builtin_ldexpf(1.0f, -__builtin_amdgcn_frexp_expf(dG) - 6);
    const int from = 16 * (G0 + G) + col;
    const float b0 = __shfl(own[0], from, 64), b1 = __shfl(own[1], from, 64);
    const float b2 = __shfl(own[2], from, 64), b3 = __shfl(own[3], from, 64);
    const float pq = __shfl(own[4], from, 64);
    Bv[G] = sig[G] * (g == 0 ? b0 : g == 1 ? b1 : g == 2 ? b2 : b3);
    Cv[G] = sig[G] * pq;
  }
  swe_mf4 S1[NG], S2[NG];
#pragma unroll
  for (int G = 0; G < NG; ++G) S1[G] = S2[G] = swe_mf4{0.0f, 0.0f, 0.0f, 0.0f};
  const float* lyaf = reinterpret_cast<const float*>(ya);
  const int f1 = min(col, kSweMfR1 - 1), f2 = min(col, kSweMfR2 - 1);
  // operands of k-block kb; the next block's are fetched before the current
  // one is computed
  struct KbOps {
    float a0, a1;
    uint4 u1h, u1l, u2h, u2l;
  };
  auto fetch = [&](int kb) {
    KbOps o;
    o.a0 = lyaf[(32 * kb + col) * 4 + g];
    o.a1 = lyaf[(32 * kb + 16 + col) * 4 + g];
    o.u1h = lfr[((kb * 2 + 0) * 4 + g) * kSweMfR1 + f1];
    o.u1l = lfr[((kb * 2 + 1) * 4 + g) * kSweMfR1 + f1];
    o.u2h = lfr2[((kb * 2 + 0) * 4 + g) * kSweMfR2 + f2];
    o.u2l = lfr2[((kb * 2 + 1) * 4 + g) * kSweMfR2 + f2];
    return o;
  };
  KbOps cur = fetch(kb0);
#pragma unroll 1
  for (int kb = kb0; kb < kb1; ++kb) {
    const KbOps nxt = fetch(min(kb + 1, kb1 - 1));
    const float a0 = cur.a0, a1 = cur.a1;
    uint4 u1h = cur.u1h, u1l = cur.u1l, u2h = cur.u2h, u2l = cur.u2l;
    const uint4 z = make_uint4(0u, 0u, 0u, 0u);
    if (col >= kSweMfR1) u1h = u1l = z;
    if (col >= kSweMfR2) u2h = u2l = z;
    const swe_mh8 A1h = __builtin_bit_cast(swe_mh8, u1h);
    const swe_mh8 A1l = __builtin_bit_cast(swe_mh8, u1l);
    const swe_mh8 A2h = __builtin_bit_cast(swe_mh8, u2h);
    const swe_mh8 A2l = __builtin_bit_cast(swe_mh8, u2l);
    cur = nxt;
#pragma unroll
    for (int G = 0; G < NG; ++G) {
      const swe_mf4 C = {Cv[G], Cv[G], Cv[G], Cv[G]};
      const swe_mf4 o0 = __builtin_amdgcn_mfma_f32_16x16x4f32(a0, Bv[G], C, 0, 0, 0);
      const swe_mf4 o1 = __builtin_amdgcn_mfma_f32_16x16x4f32(a1, Bv[G], C, 0, 0, 0);
      float w[8];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        w[v] = __builtin_amdgcn_rcpf(o0[v]);
        w[4 + v] = __builtin_amdgcn_rcpf(o1[v]);
      }
      uint32_t wh[4], wl[4], vh[4], vl[4];
#pragma unroll
      for (int p = 0; p < 4; ++p) {
        split_pair_mix(w[2 * p], w[2 * p + 1], wh[p], wl[p]);
        split_pair(w[2 * p] * w[2 * p], w[2 * p + 1] * w[2 * p + 1], vh[p],
                   vl[p]);
      }
      const swe_mh8 Wh = __builtin_bit_cast(swe_mh8, make_uint4(wh[0], wh[1], wh[2], wh[3]));
      const swe_mh8 Wl = __builtin_bit_cast(swe_mh8, make_uint4(wl[0], wl[1], wl[2], wl[3]));
      const swe_mh8 Vh = __builtin_bit_cast(swe_mh8, make_uint4(vh[0], vh[1], vh[2], vh[3]));
      const swe_mh8 Vl = __builtin_bit_cast(swe_mh8, make_uint4(vl[0], vl[1], vl[2], vl[3]));
      S1[G] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A1h, Wh, S1[G], 0, 0, 0);
      S2[G] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A2h, Vh, S2[G], 0, 0, 0);
      S1[G] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A1l, Wh, S1[G], 0, 0, 0);
      S2[G] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A2l, Vh, S2[G], 0, 0, 0);
      S1[G] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A1h, Wl, S1[G], 0, 0, 0);
      S2[G] = __builtin_amdgcn_mfma_f32_16x16x32_f16(A2h, Vl, S2[G], 0, 0, 0);
    }
  }
  // undo the scales into the FP64 sums: w = 1 / (sigma D)
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const double s1 = lfs[4 * g + v], s2 = lfs[16 + 4 * g + v];
#pragma unroll
    for (int G = 0; G < NG; ++G) {
      const double sg1 = static_cast<double>(sig[G]);
      F1[G][v] = fma(static_cast<double>(S1[G][v]), sg1 * s1, F1[G][v]);
      F2[G][v] = fma(static_cast<double>(S2[G][v]), sg1 * sg1 * s2, F2[G][v]);
    }
  }
}

constexpr int kSweMfRows = kSweMfR1 + kSweMfR2;

// Row r of this lane's target's far sums (the target is 16 g + col; row
// 4 q + v of it sits in lane col + 16 q, group g), for lanes whose group g
// is one of G0 .. G0 + NG - 1 held in F1 / F2.  Every lane calls it.
template <int NG>
__device__ __forceinline__ void swe_mf_gather(double (*F1)[4], double (*F2)[4],
                                              int col, int g, int G0,
                                              double* R) {
#pragma unroll
  for (int r = 0; r < kSweMfR2; ++r) {
    const int from = col + 16 * (r >> 2), v = r & 3;
    double a1 = 0.0, a2 = 0.0;
#pragma unroll
    for (int G = 0; G < NG; ++G) {
      const double t2 = __shfl(F2[G][v], from, 64);
      if (g == G0 + G) a2 = t2;
      if (r < kSweMfR1) {
        const double t1 = __shfl(F1[G][v], from, 64);
        if (g == G0 + G) a1 = t1;
      }
    }
    R[kSweMfR1 + r] = a2;
    if (r < kSweMfR1) R[r] = a1;
  }
}

// Adds the projected far sums R (set 1 rows 0..5, set 2 rows 6..17) of
// target x to its accumulators.
__device__ __forceinline__ void swe_mf_add_far(SweAcc& acc, const double* R,
                                               const double* x) {
  acc.Vy[0] += R[0];
  acc.Vy[1] += R[1];
  acc.Vy[2] += R[2];
  const double xs = x[0] * R[3] + x[1] * R[4] + x[2] * R[5];
  acc.Vg[0] += R[3] - xs * x[0];
  acc.Vg[1] += R[4] - xs * x[1];
  acc.Vg[2] += R[5] - xs * x[2];
  acc.Bt += xs;
  add_qyq(acc.Gd, R + kSweMfR1, x);
  add_qyq(acc.Gs, R + kSweMfR1 + 6, x);
}

// First pass: the far tiles, and the far sub-tiles of the others, of each
// wave, NG groups of 16 targets per sweep over the tiles (4 / NG sweeps);
// the 18 raw far sums of slot c go to sg.far[r * sg.far_stride + c].  Tile
// records are read from global memory (L2-resident), the next tile's
// centroid one tile ahead.
template <int BLOCK, int NG>
__device__ __forceinline__ void swe_mf_far_block(const SweMfTiles& mt,
                                                 int nsrc, const SweSeg& sg,
                                                 int blk, const SweParams& P) {
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  const int64_t c = static_cast<int64_t>(blk) * BLOCK + threadIdx.x;
  const int64_t wave0 = c - (threadIdx.x & 63);
  if (wave0 >= n_live) return;
  const bool in = c < n_live;
  const int64_t i = swe_target(sg, c, in);
  const double* px = sg.x + i * sg.es;
  const double x[3] = {px[0], px[sg.cs], px[2 * sg.cs]};
  const int lane = threadIdx.x & 63;
  const int col = lane & 15, g = lane >> 4;
  const int last = max((nsrc + kSweTile - 1) / kSweTile - 1, 0);
#pragma unroll 1
  for (int G0 = 0; G0 < 4; G0 += NG) {
    double F1[NG][4], F2[NG][4];
#pragma unroll
    for (int G = 0; G < NG; ++G)
#pragma unroll
      for (int v = 0; v < 4; ++v) F1[G][v] = F2[G][v] = 0.0;
    double4 cc_next = vload(mt.cen, 0);
    for (int t = 0; t <= last && t * kSweTile < nsrc; ++t) {
      const double4 cc = cc_next;
      cc_next = vload(mt.cen, min(t + 1, last));
      const uint4* fr = mt.fr + static_cast<int64_t>(t) * kSweMfFr;
      const uint4* fr2 = mt.fr2 + static_cast<int64_t>(t) * kSweMfFr2;
      const double* fs = mt.fs + 32 * static_cast<int64_t>(t);
      // a far tile is one range of k-blocks; a near tile, its far sub-tiles
      // with offsets from their own centroids
      const bool tile_far = !__ballot(in && swe_mf_near(x, cc));
      const int nrange = tile_far ? 1 : kSweMfNsub;
#pragma unroll 1
      for (int sub = 0; sub < nrange; ++sub) {
        double4 cs = cc;
        const float4* ya = mt.ya + static_cast<int64_t>(t) * kSweTile;
        int kb0 = 0, kb1 = kSweMfKb;
        if (!tile_far) {
          cs = mt.scen[kSweMfNsub * static_cast<int64_t>(t) + sub];
          if (cs.w < 0.0 || __ballot(in && swe_mf_near(x, cs))) continue;
          ya = mt.yas + static_cast<int64_t>(t) * kSweTile;
          kb0 = sub * (kSweMfSub / 32);
          kb1 = kb0 + kSweMfSub / 32;
        }
        const double e0 = x[0] - cs.x, e1 = x[1] - cs.y, e2 = x[2] - cs.z;
        swe_mf_far_tile<NG>(ya, fr, fr2, fs, kb0, kb1, in, e0, e1, e2,
                            e0 * e0 + e1 * e1 + e2 * e2, cs.w, P.eps2, col,
                            g, G0, F1, F2);
      }
    }
    double R[kSweMfRows];
    swe_mf_gather<NG>(F1, F2, col, g, G0, R);
    if (in && g >= G0 && g < G0 + NG)
      for (int r = 0; r < kSweMfRows; ++r) sg.far[r * sg.far_stride + c] = R[r];
  }
}

// Second pass, each wave on its own (no block barriers, so a wave's near
// work never idles the others): the wave scans the tiles (the first pass's
// classification: tile near, then sub-tile near) and the PSE range of its
// targets, stages each sub-tile it needs into its own LDS slice (one record
// per lane) and sums it in FP64.  Then the first pass's far sums are added
// and the outputs written.  SHARE: pairs per FP64 reciprocal without a self
// pair.  stage: kSweMfSub records of this wave.
template <int BLOCK, int SHARE>
__device__ __forceinline__ void swe_mf_near_block(
    const SweSrc* __restrict__ src, const SweMfTiles& mt, int nsrc,
    const SweSeg& sg, int blk, const SweParams& P, SweSrc* stage) {
  const double eps2 = P.eps2, inv_pse_e2 = P.inv_pse_e2;
  const double pse_cut2 = P.pse_cut2;
  const int do_pse = P.do_pse;
  const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
  const int64_t c = static_cast<int64_t>(blk) * BLOCK + threadIdx.x;
  const int64_t wave0 = c - (threadIdx.x & 63);
  if (wave0 >= n_live) return;
  const bool in = c < n_live;
  const int64_t i = swe_target(sg, c, in);
  const double* px = sg.x + i * sg.es;
  const double x[3] = {px[0], px[sg.cs], px[2 * sg.cs]};
  const double s_i = (do_pse && sg.surf) ? sg.surf[i] : 0.0;
  const int self = swe_self(sg, c, i, in, P.skip_self);
  const double rcut = sqrt(pse_cut2);
  const int lane = threadIdx.x & 63;
  const int last = max((nsrc + kSweTile - 1) / kSweTile - 1, 0);
  // is x within PSE range of some source in the sphere (c, c.w = r^2)?
  auto pse_near = [&](const double4& s) {
    const double p0 = x[0] - s.x, p1 = x[1] - s.y, p2 = x[2] - s.z;
    const double rr = sqrt(s.w) + rcut;
    return in && p0 * p0 + p1 * p1 + p2 * p2 <= rr * rr;
  };
  SweAcc acc;
  acc.zero();
  double4 cc_next = vload(mt.cen, 0);
  for (int t = 0; t <= last && t * kSweTile < nsrc; ++t) {
    const double4 cc = cc_next;
    cc_next = vload(mt.cen, min(t + 1, last));
    const bool near = __ballot(in && swe_mf_near(x, cc)) != 0;
    const bool pse = do_pse && __ballot(pse_near(cc)) != 0;
    if (!near && !pse) continue;
#pragma unroll 1
    for (int sub = 0; sub < kSweMfNsub; ++sub) {
      const double4 cs = mt.scen[kSweMfNsub * static_cast<int64_t>(t) + sub];
      if (cs.w < 0.0) continue;
      const bool nsub = near && __ballot(in && swe_mf_near(x, cs)) != 0;
      const bool psub = pse && __ballot(pse_near(cs)) != 0;
      if (!nsub && !psub) continue;
      const int s0 = t * kSweTile + sub * kSweMfSub;
      const int cnt = min(kSweMfSub, nsrc - s0);
      // the previous sub-tile's reads are done before it is overwritten;
      // the writes land before any lane reads (LDS is in order per wave)
      __builtin_amdgcn_wave_barrier();
      stage[lane] = src[s0 + lane];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (nsub) {
        const bool diag = static_cast<unsigned>(self - s0) <
                          static_cast<unsigned>(cnt);
        if (__ballot(diag))
          swe_tile_f64<2>(acc, stage, cnt, x, eps2, self - s0, true);
        else
          swe_tile_f64<SHARE>(acc, stage, cnt, x, eps2, -1, false);
      }
      if (psub) {
        double L = 0.0;
        swe_pse_tile(stage, cnt, x, s_i, inv_pse_e2, pse_cut2, L);
        acc.L1 += L;
      }
    }
  }
  if (!in) return;
  double R[kSweMfRows];
  for (int r = 0; r < kSweMfRows; ++r) R[r] = sg.far[r * sg.far_stride + c];
  swe_mf_add_far(acc, R, x);
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

// MINW: minimum waves per SIMD asked of the register allocator; NG: groups
// of 16 targets per wave and sweep of the first pass (4 or 2).
template <int BLOCK, int MINW, int NG>
__global__ __launch_bounds__(BLOCK, MINW) void swe_mf_far_kernel(
    const SweSrc* __restrict__, const double4* __restrict__,
    const int32_t* __restrict__ count, SweSeg s0, SweSeg s1, SweParams P) {
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SweSeg& sg = first ? s0 : s1;
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  swe_mf_far_block<BLOCK, NG>(P.mf, *count, sg, blk, P);
}

// SHARE: pairs per FP64 reciprocal in near sub-tiles (2 or 4).
template <int BLOCK, int MINW, int SHARE>
__global__ __launch_bounds__(BLOCK, MINW) void swe_mf_near_kernel(
    const SweSrc* __restrict__ src, const double4* __restrict__,
    const int32_t* __restrict__ count, SweSeg s0, SweSeg s1, SweParams P) {
  __shared__ SweSrc stage[BLOCK];
  const bool first = static_cast<int>(blockIdx.x) < s0.nblocks;
  const SweSeg& sg = first ? s0 : s1;
  const int blk = first ? blockIdx.x : blockIdx.x - s0.nblocks;
  swe_mf_near_block<BLOCK, SHARE>(src, P.mf, *count, sg, blk, P,
                                  stage + (threadIdx.x & ~63));
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

// grid (ceil(max member faces / 256), n_members): the tile records of
// member m (its tiles start at packed tile P_m / 256).
__global__ __launch_bounds__(256) void swe_ens_mf_pack_kernel(
    const SweSrc* __restrict__ src, const EnsMember* __restrict__ mem,
    SweMfTiles T) {
  const EnsMember M = mem[blockIdx.y];
  const int j0 = blockIdx.x * kSweTile;
  if (j0 >= M.cnt) return;
  swe_mf_pack_tile(src + M.P + j0, min(kSweTile, M.cnt - j0),
                   T.from((static_cast<int64_t>(M.P) + j0) / kSweTile));
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
  double* far;          // two-pass matrix-core sums (rows of far_stride):
  int64_t far_stride;   // vertex v at v, face slot f at far_face0 + f
  int64_t far_face0;
};

__device__ __forceinline__ SweSeg swe_ens_segment(const SweEnsArgs& A,
                                                  const EnsMember& M, int seg) {
  if (seg == 0) {
    SweSeg sg = A.vert;
    sg.n = M.nv;
    sg.slot_base = M.v0;
    if (A.vert.index) sg.index = A.vert.index + M.v0;
    sg.far = A.far ? A.far + M.v0 : nullptr;
    sg.far_stride = A.far_stride;
    return sg;
  }
  SweSeg sg = A.face;
  int64_t f0;
  if (A.skip_masked && A.inv) {
    sg.index = A.inv + M.R;
    sg.n = M.cnt;
    sg.self_identity = 1;
    f0 = M.R;
  } else {
    sg.n = M.nf;
    sg.slot_base = M.f0;
    if (A.rank) {
      sg.self_rank = A.rank;
      sg.rank_base = M.R;
    } else {
      sg.self_identity = 1;
    }
    f0 = M.f0;
  }
  sg.far = A.far ? A.far + A.far_face0 + f0 : nullptr;
  sg.far_stride = A.far_stride;
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

template <int BLOCK, int MINW, int NG>
__global__ __launch_bounds__(BLOCK, MINW) void swe_ens_mf_far_kernel(
    const SweSrc* __restrict__, const double4* __restrict__, SweEnsArgs A) {
  __shared__ int slot;
  for (;;) {
    const int it = swe_ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SweSeg sg = swe_ens_segment(A, M, w.y);
    swe_mf_far_block<BLOCK, NG>(A.P.mf.from(M.P / kSweTile), M.cnt, sg, w.z,
                                A.P);
  }
}

template <int BLOCK, int MINW, int SHARE>
__global__ __launch_bounds__(BLOCK, MINW) void swe_ens_mf_near_kernel(
    const SweSrc* __restrict__ src, const double4* __restrict__,
    SweEnsArgs A) {
  __shared__ SweSrc stage[BLOCK];
  __shared__ int slot;
  for (;;) {
    const int it = swe_ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SweSeg sg = swe_ens_segment(A, M, w.y);
    swe_mf_near_block<BLOCK, SHARE>(src + M.P, A.P.mf.from(M.P / kSweTile),
                                    M.cnt, sg, w.z, A.P,
                                    stage + (threadIdx.x & ~63));
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
  bool mf = false;  // reads the matrix-core tile records
  // second pass (near tiles, PSE and outputs) after fn / ens wrote the far
  // sums; null: one pass
  SweKernel fn2 = nullptr;
  SweEnsKernel ens2 = nullptr;
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
    // two passes: far tiles and sub-tiles on the matrix cores, then near
    // sub-tiles and PSE terms in FP64 and the outputs
    {swe_mf_far_kernel<kSweBlock, 2, 4>, swe_ens_mf_far_kernel<kSweBlock, 2, 4>,
     "mfma_f2n3", true, swe_mf_near_kernel<kSweBlock, 3, 4>,
     swe_ens_mf_near_kernel<kSweBlock, 3, 4>},
    {swe_mf_far_kernel<kSweBlock, 2, 4>, swe_ens_mf_far_kernel<kSweBlock, 2, 4>,
     "mfma_f2n4", true, swe_mf_near_kernel<kSweBlock, 4, 4>,
     swe_ens_mf_near_kernel<kSweBlock, 4, 4>},
    // FP32 pair arithmetic on the vector ALUs
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

void swe_feature_workspace(lpm_ctx* ctx, int64_t n_src, int64_t n_targets,
                           int precision) {
  if (precision != LPM_MIXED) return;
  const SweVariant& v = swe_variant(ctx, precision, n_targets);
  if (!v.mf) return;
  const int64_t ntile = (std::max<int64_t>(n_src, 1) + kSweTile - 1) / kSweTile;
  ctx->swe_feat.ensure(kSweMfTileBytes * ntile);
  if (v.fn2)
    ctx->swe_far.ensure(sizeof(double) * kSweMfRows *
                        std::max<int64_t>(n_targets, 1));
}

SwePacked swe_pack_sources(lpm_ctx* ctx, CrdIn sx, const double* zeta,
                           const double* sigma, const double* area,
                           const double* surf, const RankMap& map,
                           hipStream_t s) {
  SwePacked ps;
  const int64_t n = map.n;
  const int ntile =
      static_cast<int>((std::max<int64_t>(n, 1) + kSweTile - 1) / kSweTile);
  const int64_t cap = static_cast<int64_t>(ntile) * kSweTile;
  SweSrc* p = static_cast<SweSrc*>(ctx->swe_packed.ensure(sizeof(SweSrc) * cap));
  double4* b = static_cast<double4*>(ctx->swe_bounds.ensure(sizeof(double4) * ntile));
  if (n > 0) {
    swe_pack_kernel<<<ntile, 256, 0, s>>>(sx.base, sx.es, sx.cs, zeta, sigma,
                                           area, surf, map.inv, map.d_count,
                                           cap, p);
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
  SweParams P{eps * eps, inv_e2, cut2, do_pse, skip_self};
  const SweVariant& v = swe_variant(ctx, precision, t0.base.n + t1.base.n);
  if (v.mf) {
    // the matrix-core records of every source tile
    const int64_t ntile =
        (std::max<int64_t>(src.capacity, 1) + kSweTile - 1) / kSweTile;
    P.mf = swe_mf_tiles(ctx->swe_feat.ensure(kSweMfTileBytes * ntile), ntile);
    swe_mf_pack_kernel<<<static_cast<int>(ntile), 256, 0, s>>>(
        static_cast<const SweSrc*>(src.data), src.count, P.mf);
    LPM_HIP_CHECK(hipGetLastError());
  }
  if (v.fn2) {
    // far sums of every slot of both segments, rows of a0.n + a1.n
    const int64_t total = a0.n + a1.n;
    double* far = static_cast<double*>(
        ctx->swe_far.ensure(sizeof(double) * kSweMfRows * total));
    a0.far = far;
    a1.far = far + a0.n;
    a0.far_stride = a1.far_stride = total;
  }
  {
    TimedLaunch tl(ctx, precision == LPM_FP64 ? KID_SWE_F64 : KID_SWE_MIXED, s,
                   pairs);
    hipLaunchKernelGGL(v.fn, dim3(grid), dim3(kBlock), 0, s,
                       static_cast<const SweSrc*>(src.data), src.bounds,
                       src.count, a0, a1, P);
    if (v.fn2) {
      LPM_HIP_CHECK(hipGetLastError());
      hipLaunchKernelGGL(v.fn2, dim3(grid), dim3(kBlock), 0, s,
                         static_cast<const SweSrc*>(src.data), src.bounds,
                         src.count, a0, a1, P);
    }
  }
  LPM_HIP_CHECK(hipGetLastError());
}

// Vertex slot order of the mixed sums: from the members' state positions
// (also for the stage evaluations of a step, whose positions are close).
const int32_t* swe_ens_vertex_order(lpm_ctx* ctx, const EnsembleDesc& d,
                                    const lpm_swe_fields& f, hipStream_t s) {
  return target_order(ctx, CrdIn{f.vert_x, f.vert_x_es, f.vert_x_cs},
                      d.n_verts, d.voff, d.n_members, s);
}

uint64_t swe_ensemble_prepare(lpm_ctx* ctx, const EnsembleDesc& d,
                              const lpm_swe_fields& f, int precision,
                              hipStream_t s) {
  const EnsLayout L = ensemble_plan(ctx, d, 256);
  const int64_t cap = std::max<int64_t>(d.total_padded_upper, kSweTile);
  ctx->swe_packed.ensure(sizeof(SweSrc) * cap);
  ctx->swe_bounds.ensure(sizeof(double4) * (cap / kSweTile));
  if (precision == LPM_MIXED) {
    const SweVariant& v = swe_variant(ctx, precision, d.n_verts + d.n_faces);
    if (v.mf) ctx->swe_feat.ensure(kSweMfTileBytes * (cap / kSweTile));
    if (v.ens2)
      ctx->swe_far.ensure(sizeof(double) * kSweMfRows *
                          std::max<int64_t>(d.n_verts + d.n_faces, 1));
    swe_ens_vertex_order(ctx, d, f, s);
  }
  return L.serial;
}

void swe_ensemble_sums(lpm_ctx* ctx, const EnsembleDesc& d,
                       const lpm_swe_fields& f, CrdIn vx, CrdIn fx,
                       const double* fzeta, const double* fsigma,
                       const double* farea, const RankMap& map, double eps,
                       double pse_eps, int precision, unsigned flags,
                       hipStream_t s) {
  constexpr int kBlock = 256;
  swe_ensemble_prepare(ctx, d, f, precision, s);
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
  if (precision == LPM_MIXED) A.vert.index = swe_ens_vertex_order(ctx, d, f, s);
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
  const SweVariant& v = swe_variant(ctx, precision, d.n_verts + d.n_faces);
  if (v.mf) {
    const int64_t cap = std::max<int64_t>(d.total_padded_upper, kSweTile);
    A.P.mf = swe_mf_tiles(ctx->swe_feat.as<void>(), cap / kSweTile);
    swe_ens_mf_pack_kernel<<<grid2, 256, 0, s>>>(src, L.mem, A.P.mf);
    LPM_HIP_CHECK(hipGetLastError());
  }
  if (v.ens2) {
    A.far = ctx->swe_far.as<double>();
    A.far_stride = d.n_verts + d.n_faces;
    A.far_face0 = d.n_verts;
  }
  ensemble_reset_queue(L, s);
  const int grid = ensemble_grid(ctx, reinterpret_cast<const void*>(v.ens),
                                 kBlock, L.n_items);
  {
    TimedLaunch tl(ctx, precision == LPM_FP64 ? KID_SWE_F64 : KID_SWE_MIXED, s,
                   L.pairs);
    hipLaunchKernelGGL(v.ens, dim3(grid), dim3(kBlock), 0, s,
                       static_cast<const SweSrc*>(src),
                       static_cast<const double4*>(bounds), A);
    if (v.ens2) {
      LPM_HIP_CHECK(hipGetLastError());
      ensemble_reset_queue(L, s);
      const int grid2b = ensemble_grid(
          ctx, reinterpret_cast<const void*>(v.ens2), kBlock, L.n_items);
      hipLaunchKernelGGL(v.ens2, dim3(grid2b), dim3(kBlock), 0, s,
                         static_cast<const SweSrc*>(src),
                         static_cast<const double4*>(bounds), A);
    }
  }
  LPM_HIP_CHECK(hipGetLastError());
}

}  // namespace lpm_amd
