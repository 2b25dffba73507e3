// 2-D incompressible flow on the sphere: the fused velocity + stream
// function direct sum of Incompressible2DKernels<SphereGeometry>::kernel_vals
// (src/lpm_incompressible2d_kernels.hpp:35-47) and the Heun RK2 stepper of
// Incompressible2DRK2 (src/lpm_incompressible2d_rk2_impl.hpp:75-172).
//
// The sum shares the BVE prologue (mask scan, packed FP64 sources
// {y, s = -zeta A / 4 pi}) and target segments.  Per pair:
//   arg = 1 - x.y + eps^2,  S += (s / arg) y,  psi += s log(arg),
// and u = x cross S at the end (x cross y / (-4 pi arg) Gamma = x cross
// (s / arg) y).  The table-driven FP64 log (bve_device.hpp) dominates the
// pair cost.
//
// Mixed precision: the pair runs in FP32 on tile-relative offsets
// d = y' - x' (as the BVE mixed kernel), with 1 - x.y = |d|^2 / 2 on the
// sphere, so arg = |d|^2 / 2 + eps^2, the weight 2 s / (|d|^2 + 2 eps^2)
// and psi += s log(arg) with the native FP32 log; FP32 partials per
// 256-source tile are carried into FP64.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <functional>
#include <sstream>

#include "bve_device.hpp"
#include "internal.hpp"

namespace lpm_amd {
namespace {

// One LDS tile of FP64 sources against this lane's target, summed into
// S / psi.  CHECK drops the pair jj == self (tile-relative); only tiles
// holding some lane's own source of the wave need it.  SPECIAL evaluates
// log's zero / negative cases; without it the smallest argument is
// returned so that the caller can redo the tile when one was <= 0.
template <bool CHECK, bool SPECIAL>
__device__ __forceinline__ double ic2d_tile_f64(const double4* tile,
                                                const double* x, double one_e,
                                                int self, const double2* ltab,
                                                double* S, double& psi) {
  double amin = 1.0;
  for (int jj = 0; jj < kPackTile; ++jj) {
    const double4 y = tile[jj];
    // 1 - x.y + eps^2 as one FMA chain from 1 + eps^2
    const double arg = __builtin_fma(
        -x[2], y.z, __builtin_fma(-x[1], y.y, __builtin_fma(-x[0], y.x, one_e)));
    if (!SPECIAL) amin = fmin(amin, arg);
    double w = y.w * rcp_nr<2>(arg);
    double p = y.w * log_f64<SPECIAL>(arg, ltab);
    if (CHECK && jj == self) {
      w = 0.0;
      p = 0.0;
    }
    S[0] = __builtin_fma(w, y.x, S[0]);
    S[1] = __builtin_fma(w, y.y, S[1]);
    S[2] = __builtin_fma(w, y.z, S[2]);
    psi += p;
  }
  return amin;
}

// Per-tile partial sums, redone with log's special cases in the rare waves
// where some argument was <= 0 (coincident points without smoothing).
template <bool CHECK>
__device__ __forceinline__ void ic2d_tile_sum(const double4* tile,
                                              const double* x, double one_e,
                                              int self, const double2* ltab,
                                              double* S, double& psi) {
  double T[3] = {0.0, 0.0, 0.0}, P = 0.0;
  const double amin =
      ic2d_tile_f64<CHECK, false>(tile, x, one_e, self, ltab, T, P);
  if (__ballot(amin <= 0.0)) {
    T[0] = T[1] = T[2] = P = 0.0;
    ic2d_tile_f64<CHECK, true>(tile, x, one_e, self, ltab, T, P);
  }
  S[0] += T[0];
  S[1] += T[1];
  S[2] += T[2];
  psi += P;
}

// Source chunk `chunk` of ks equal runs of packing tiles: [*jbeg, *jend).
__device__ __forceinline__ void ic2d_chunk(int nsrc, int ks, int chunk,
                                           int* jbeg, int* jend) {
  const int ntile = (nsrc + kPackTile - 1) / kPackTile;
  *jbeg = (ntile * chunk) / ks * kPackTile;
  *jend = (ntile * (chunk + 1)) / ks * kPackTile;
}

// Stores S and psi of slot blk * BLOCK + threadIdx.x: to the chunk's
// partials when the launch is source-split (sg.part), else u = x cross S
// and psi at a live target.
template <int BLOCK>
__device__ __forceinline__ void ic2d_store(const SegArgs& sg, int blk,
                                           int chunk, const double (*x)[3],
                                           const double (*S)[3], double psi,
                                           const bool* live,
                                           const int64_t* tgt) {
  if (sg.part) {
    const int64_t n_live = sg.d_n ? *sg.d_n : sg.n;
    const int64_t c = static_cast<int64_t>(blk) * BLOCK + threadIdx.x;
    if (c >= n_live) return;
    double* q = sg.part + c;
    const int64_t k4 = 4 * static_cast<int64_t>(chunk);
    q[(k4 + 0) * sg.pstride] = S[0][0];
    q[(k4 + 1) * sg.pstride] = S[0][1];
    q[(k4 + 2) * sg.pstride] = S[0][2];
    q[(k4 + 3) * sg.pstride] = psi;
    return;
  }
  if (sg.out) store_cross<1>(sg, x, S, live, tgt);
  if (sg.out_scalar && live[0]) sg.out_scalar[tgt[0]] = psi;
}

// Block b * ks + k sums source chunk k for target block b (ks = 1: all
// sources, results stored directly).
template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void ic2d_f64_kernel(
    const double4* __restrict__ src, const int32_t* __restrict__ count,
    SegArgs s0, SegArgs s1, double eps2, int skip_self, int ks) {
  __shared__ double4 tile[kPackTile];
  __shared__ double2 ltab[kLogTable];
  const int b = static_cast<int>(blockIdx.x) / ks;
  const int chunk = static_cast<int>(blockIdx.x) % ks;
  const bool first = b < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? b : b - s0.nblocks;
  const int nsrc = *count;
  int jbeg, jend;
  ic2d_chunk(nsrc, ks, chunk, &jbeg, &jend);
  double x[1][3];
  int self[1];
  bool live[1];
  int64_t tgt[1];
  if (!load_targets<BLOCK, 1>(sg, blk, x, self, live, tgt)) return;
  if (!skip_self) self[0] = -1;
  double S[1][3] = {{0.0, 0.0, 0.0}};
  double psi = 0.0;
  fill_log_table(ltab);
  for (int j0 = jbeg; j0 < jend; j0 += kPackTile) {
    __syncthreads();
    for (int t = threadIdx.x; t < kPackTile; t += BLOCK) {
      const int j = j0 + t;
      tile[t] = j < nsrc ? src[j] : make_double4(0.0, 0.0, 0.0, 0.0);
    }
    __syncthreads();
    const int so = self[0] < 0 ? -1 : self[0] - j0;
    if (__ballot(static_cast<unsigned>(so) < static_cast<unsigned>(kPackTile)))
      ic2d_tile_sum<true>(tile, x[0], 1.0 + eps2, so, ltab, S[0], psi);
    else
      ic2d_tile_sum<false>(tile, x[0], 1.0 + eps2, so, ltab, S[0], psi);
  }
  ic2d_store<BLOCK>(sg, blk, chunk, x, S, psi, live, tgt);
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) void ic2d_mixed_kernel(
    const float4* __restrict__ src, const double4* __restrict__ centers,
    const int32_t* __restrict__ count, SegArgs s0, SegArgs s1, float e2x2,
    float e2, int skip_self, int ks) {
  __shared__ float4 tile[kPackTile];
  const int b = static_cast<int>(blockIdx.x) / ks;
  const int chunk = static_cast<int>(blockIdx.x) % ks;
  const bool first = b < s0.nblocks;
  const SegArgs& sg = pick(s0, s1, first);
  const int blk = first ? b : b - s0.nblocks;
  const int nsrc = *count;
  int jbeg, jend;
  ic2d_chunk(nsrc, ks, chunk, &jbeg, &jend);
  double x[1][3];
  int self[1];
  bool live[1];
  int64_t tgt[1];
  if (!load_targets<BLOCK, 1>(sg, blk, x, self, live, tgt)) return;
  if (!skip_self) self[0] = -1;
  double S[1][3] = {{0.0, 0.0, 0.0}};
  double psi = 0.0;
  constexpr float kLn2 = 0.693147180559945309f;
  for (int j0 = jbeg; j0 < jend; j0 += kPackTile) {
    __syncthreads();
    for (int t = threadIdx.x; t < kPackTile; t += BLOCK) {
      const int j = j0 + t;
      // padding: far away, zero strength -> contributes exactly 0
      tile[t] = j < nsrc ? src[j] : make_float4(8.0f, 8.0f, 8.0f, 0.0f);
    }
    __syncthreads();
    const double4 c = centers[j0 / kPackTile];
    const float xr0 = static_cast<float>(x[0][0] - c.x);
    const float xr1 = static_cast<float>(x[0][1] - c.y);
    const float xr2 = static_cast<float>(x[0][2] - c.z);
    const bool diag = static_cast<unsigned>(self[0] - j0) <
                      static_cast<unsigned>(kPackTile);
    float T0 = 0.0f, T1 = 0.0f, T2 = 0.0f, P = 0.0f;
    for (int jj = 0; jj < kPackTile; ++jj) {
      const float4 y = tile[jj];
      const float d0 = y.x - xr0, d1 = y.y - xr1, d2 = y.z - xr2;
      const float r2 = d0 * d0 + d1 * d1 + d2 * d2;
      float w = y.w * __builtin_amdgcn_rcpf(r2 + e2x2);
      // y.w = 2 s: psi term s log(arg) = y.w log2(arg) (ln 2 / 2)
      float p = y.w * __builtin_amdgcn_logf(__builtin_fmaf(0.5f, r2, e2));
      if (diag && j0 + jj == self[0]) {
        w = 0.0f;
        p = 0.0f;
      }
      T0 = __builtin_fmaf(w, d0, T0);
      T1 = __builtin_fmaf(w, d1, T1);
      T2 = __builtin_fmaf(w, d2, T2);
      P += p;
    }
    S[0][0] += static_cast<double>(T0);
    S[0][1] += static_cast<double>(T1);
    S[0][2] += static_cast<double>(T2);
    psi += static_cast<double>(P) * (0.5 * static_cast<double>(kLn2));
  }
  ic2d_store<BLOCK>(sg, blk, chunk, x, S, psi, live, tgt);
}

// u = x cross (sum of the ks chunk partials of S, in chunk order) and psi
// = the sum of the chunk partials of psi, at every live slot.
__global__ __launch_bounds__(256) void ic2d_split_reduce_kernel(SegArgs s0,
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
  const double* q = sg.part + c;
  if (sg.out)
    split_reduce_one(q, sg.pstride, ks, sg.x + i * sg.es, sg.cs,
                     sg.out + i * sg.oes, sg.ocs, 4);
  if (sg.out_scalar) {
    double psi = 0.0;
    for (int k = 0; k < ks; ++k)
      psi += q[(4 * static_cast<int64_t>(k) + 3) * sg.pstride];
    sg.out_scalar[i] = psi;
  }
}

// Stage storage of one particle kind (SoA, leading dimension n): x1 (3n),
// z1 (n), xw (3n), zw (n).
struct Ic2dStage {
  double* x1;
  double* z1;
  double* xw;
  double* zw;
  int64_t n;
};

Ic2dStage carve(DevBuf& buf, int64_t n) {
  const int64_t nn = std::max<int64_t>(n, 1);
  double* p = static_cast<double*>(buf.ensure(sizeof(double) * 8 * nn));
  return Ic2dStage{p, p + 3 * nn, p + 4 * nn, p + 7 * nn, n};
}

struct Ic2dArgs {
  double* x;  // state positions (strided)
  int64_t xes, xcs;
  double* zeta;
  const double* u;  // velocity (strided)
  int64_t ues, ucs;
  const uint8_t* skip;
  Ic2dStage st;
  double dt, Omega;
  int renorm;
};

__device__ __forceinline__ void project(double* v) {
  const double r = 1.0 / sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
  v[0] *= r;
  v[1] *= r;
  v[2] *= r;
}

// Stage 1: x1 = dt u, z1 = -2 Omega u_z (a rate), xw = x + dt u,
// zw = zeta + dt z1 (KokkosBlas::scal / update and
// Incompressible2DTendencies of the reference's first stage).
__global__ __launch_bounds__(256) void ic2d_stage1_kernel(Ic2dArgs A) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = A.st.n;
  if (i >= n) return;
  if (A.skip && A.skip[i]) return;
  const double* ui = A.u + i * A.ues;
  const double* xi = A.x + i * A.xes;
  const double u[3] = {ui[0], ui[A.ucs], ui[2 * A.ucs]};
  double w[3];
  for (int k = 0; k < 3; ++k) {
    A.st.x1[i + k * n] = A.dt * u[k];
    w[k] = xi[k * A.xcs] + A.dt * u[k];
  }
  if (A.renorm) project(w);
  for (int k = 0; k < 3; ++k) A.st.xw[i + k * n] = w[k];
  const double z1 = -2.0 * A.Omega * u[2];
  A.st.z1[i] = z1;
  A.st.zw[i] = A.zeta[i] + A.dt * z1;
}

// Stage 2 and the Heun update: z2 = -2 Omega u2_z;
// zeta += dt/2 z1 + dt/2 z2, x += x1/2 + (dt u2)/2.
__global__ __launch_bounds__(256) void ic2d_update_kernel(Ic2dArgs A) {
  const int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
  const int64_t n = A.st.n;
  if (i >= n) return;
  if (A.skip && A.skip[i]) return;
  const double* ui = A.u + i * A.ues;
  double* xi = A.x + i * A.xes;
  const double u[3] = {ui[0], ui[A.ucs], ui[2 * A.ucs]};
  const double z2 = -2.0 * A.Omega * u[2];
  A.zeta[i] += 0.5 * A.dt * A.st.z1[i] + 0.5 * A.dt * z2;
  double v[3];
  for (int k = 0; k < 3; ++k)
    v[k] = xi[k * A.xcs] + (0.5 * A.st.x1[i + k * n] + 0.5 * (A.dt * u[k]));
  if (A.renorm) project(v);
  for (int k = 0; k < 3; ++k) xi[k * A.xcs] = v[k];
}

inline int nblk(int64_t n) { return static_cast<int>((n + 255) / 256); }

constexpr int kIc2dBlock = 256;

// Source chunks per target block (split_chunks): with the level-6 sums'
// 588 target blocks far below the blocks the chip holds at once, each
// block's sources are split so that every CU holds enough waves to hide
// the log's latency to the end of the launch (level 6: 2, 4 and 8 rounds of
// resident blocks measured 823, 898 and 913 Gpairs/s in FP64, 746 unsplit).
int ic2d_chunks(lpm_ctx* ctx, int precision, int64_t blocks,
                int64_t nsrc_upper) {
  const void* fn =
      precision == LPM_MIXED
          ? reinterpret_cast<const void*>(&ic2d_mixed_kernel<kIc2dBlock>)
          : reinterpret_cast<const void*>(&ic2d_f64_kernel<kIc2dBlock>);
  return split_chunks(ctx, fn, kIc2dBlock, blocks, nsrc_upper, kPackTile);
}

size_t ic2d_split_bytes(int ks, int64_t n_targets) {
  return ks > 1 ? sizeof(double) * 4 * static_cast<size_t>(ks) *
                      static_cast<size_t>(std::max<int64_t>(n_targets, 1))
                : 0;
}

void launch_ic2d(lpm_ctx* ctx, const PackedSources& src, TargetSeg* segs,
                 int nseg, double eps, bool collocated_skip, int precision,
                 hipStream_t s) {
  TargetSeg empty;
  const TargetSeg& t0 = segs[0];
  const TargetSeg& t1 = nseg > 1 ? segs[1] : empty;
  if (t0.n + t1.n == 0) return;
  constexpr int kBlock = kIc2dBlock;
  auto to_args = [](const TargetSeg& t) {
    SegArgs a{};
    a.x = t.x.base;
    a.es = t.x.es;
    a.cs = t.x.cs;
    a.n = t.n;
    a.index = t.index;
    a.d_n = t.d_n;
    a.self_rank = t.self_rank;
    a.self_identity = t.self_identity ? 1 : 0;
    a.nblocks = static_cast<int>((t.n + kBlock - 1) / kBlock);
    a.skip = t.skip;
    a.out = t.out.base;
    a.oes = t.out.es;
    a.ocs = t.out.cs;
    a.out_scalar = t.out_scalar;
    a.pstride = t.n;
    return a;
  };
  SegArgs a0 = to_args(t0), a1 = to_args(t1);
  const int blocks = a0.nblocks + a1.nblocks;
  int ks = ic2d_chunks(ctx, precision, blocks, src.capacity);
  if (ks > 1) {
    // no workspace (a caller's capture, no memory): one chunk
    double* part = split_workspace(ctx, ic2d_split_bytes(ks, t0.n + t1.n), s);
    if (part) {
      a0.part = part;
      a1.part = part + 4 * static_cast<int64_t>(ks) * t0.n;
    } else {
      ks = 1;
    }
  }
  const double pairs = static_cast<double>(t0.n + t1.n) * src.capacity;
  TimedLaunch tl(ctx, KID_IC2D, s, pairs);
  const dim3 grid(static_cast<unsigned>(blocks) * ks);
  if (precision == LPM_MIXED) {
    hipLaunchKernelGGL((ic2d_mixed_kernel<kBlock>), grid, dim3(kBlock), 0, s,
                       static_cast<const float4*>(src.data), src.centers,
                       src.count, a0, a1, static_cast<float>(2.0 * eps * eps),
                       static_cast<float>(eps * eps), collocated_skip ? 1 : 0,
                       ks);
  } else {
    hipLaunchKernelGGL((ic2d_f64_kernel<kBlock>), grid, dim3(kBlock), 0, s,
                       static_cast<const double4*>(src.data), src.count, a0,
                       a1, eps * eps, collocated_skip ? 1 : 0, ks);
  }
  LPM_HIP_CHECK(hipGetLastError());
  if (ks > 1) {
    const int64_t nt = t0.n + t1.n;
    hipLaunchKernelGGL(ic2d_split_reduce_kernel,
                       dim3(static_cast<unsigned>((nt + 255) / 256)),
                       dim3(256), 0, s, a0, a1, ks);
    LPM_HIP_CHECK(hipGetLastError());
  }
}

// Sums of the fields at positions (vx, fx) with face vorticity fzeta.
void ic2d_eval(lpm_ctx* ctx, const lpm_ic2d_fields& f, CrdIn vx, CrdIn fx,
               const double* fzeta, const RankMap& map, double eps,
               int precision, unsigned flags, hipStream_t s) {
  PackedSources ps = pack_sources(ctx, fx, fzeta, f.face_area, map, precision, s);
  TargetSeg segs[2];
  segs[0].x = vx;
  segs[0].n = f.n_verts;
  segs[0].out = CrdOut{f.vert_u, f.vert_u_es, f.vert_u_cs};
  segs[0].out_scalar = f.vert_psi;
  segs[1] = face_targets(fx, map, f.face_mask,
                         (flags & LPM_FLAG_SKIP_MASKED_TARGETS) != 0,
                         CrdOut{f.face_u, f.face_u_es, f.face_u_cs},
                         f.face_psi);
  // the self-exclusion of segment 1 applies only when eps == 0
  launch_ic2d(ctx, ps, segs, 2, eps, eps == 0.0, precision, s);
}

void validate(const lpm_ic2d_fields* f) {
  LPM_ARG_CHECK(f, "null fields");
  LPM_ARG_CHECK(f->n_verts >= 0 && f->n_faces > 0, "bad sizes");
  if (f->n_verts > 0)
    LPM_ARG_CHECK(f->vert_x && f->vert_zeta && f->vert_u && f->vert_x_es > 0 &&
                      f->vert_x_cs > 0 && f->vert_u_es > 0 && f->vert_u_cs > 0,
                  "vertex fields incomplete");
  LPM_ARG_CHECK(f->face_x && f->face_zeta && f->face_u && f->face_area &&
                    f->face_x_es > 0 && f->face_x_cs > 0 && f->face_u_es > 0 &&
                    f->face_u_cs > 0,
                "face fields incomplete");
}

void issue_rk2(lpm_ctx* ctx, const lpm_ic2d_fields& f, double dt, double Omega,
               double eps, int precision, unsigned flags, Ic2dStage& V,
               Ic2dStage& F, hipStream_t s) {
  const RankMap map = build_rank(ctx, f.face_mask, f.n_faces, s);
  const uint8_t* fskip =
      (flags & LPM_FLAG_SKIP_MASKED_TARGETS) ? f.face_mask : nullptr;
  const int renorm = (flags & LPM_RK4_RENORMALIZE) ? 1 : 0;
  auto args = [&](bool face) {
    Ic2dArgs a{};
    a.x = face ? f.face_x : f.vert_x;
    a.xes = face ? f.face_x_es : f.vert_x_es;
    a.xcs = face ? f.face_x_cs : f.vert_x_cs;
    a.zeta = face ? f.face_zeta : f.vert_zeta;
    a.u = face ? f.face_u : f.vert_u;
    a.ues = face ? f.face_u_es : f.vert_u_es;
    a.ucs = face ? f.face_u_cs : f.vert_u_cs;
    a.skip = face ? fskip : nullptr;
    a.st = face ? F : V;
    a.dt = dt;
    a.Omega = Omega;
    a.renorm = renorm;
    return a;
  };
  if (f.n_verts > 0) {
    ic2d_stage1_kernel<<<nblk(f.n_verts), 256, 0, s>>>(args(false));
    LPM_HIP_CHECK(hipGetLastError());
  }
  ic2d_stage1_kernel<<<nblk(f.n_faces), 256, 0, s>>>(args(true));
  LPM_HIP_CHECK(hipGetLastError());
  ic2d_eval(ctx, f, CrdIn{V.xw, 1, V.n}, CrdIn{F.xw, 1, F.n}, F.zw, map, eps,
            precision, flags, s);
  if (f.n_verts > 0) {
    ic2d_update_kernel<<<nblk(f.n_verts), 256, 0, s>>>(args(false));
    LPM_HIP_CHECK(hipGetLastError());
  }
  ic2d_update_kernel<<<nblk(f.n_faces), 256, 0, s>>>(args(true));
  LPM_HIP_CHECK(hipGetLastError());
  ic2d_eval(ctx, f, CrdIn{f.vert_x, f.vert_x_es, f.vert_x_cs},
            CrdIn{f.face_x, f.face_x_es, f.face_x_cs}, f.face_zeta, map, eps,
            precision, flags, s);
}

std::string ic2d_key(const lpm_ic2d_fields& f, double dt, double Omega,
                     double eps, int precision, unsigned flags,
                     hipStream_t s) {
  std::ostringstream k;
  k << "ic2d|" << f.n_verts << '|' << f.vert_x << '|' << f.vert_x_es << '|'
    << f.vert_x_cs << '|' << f.vert_zeta << '|' << f.vert_u << '|'
    << f.vert_u_es << '|' << f.vert_u_cs << '|' << f.vert_psi << '|'
    << f.n_faces << '|' << f.face_x << '|' << f.face_x_es << '|'
    << f.face_x_cs << '|' << f.face_zeta << '|' << f.face_u << '|'
    << f.face_u_es << '|' << f.face_u_cs << '|' << f.face_psi << '|'
    << f.face_area << '|' << static_cast<const void*>(f.face_mask) << '|';
  k.precision(17);
  k << dt << '|' << Omega << '|' << eps << '|' << precision << '|' << flags
    << '|' << s;
  return k.str();
}

template <typename F>
int guarded(lpm_ctx* ctx, F&& fn) {
  try {
    fn();
    return LPM_OK;
  } catch (const Error& e) {
    set_error(ctx, e.what());
    return e.code;
  } catch (const std::exception& e) {
    set_error(ctx, e.what());
    return LPM_ERR_INTERNAL;
  }
}

}  // namespace
}  // namespace lpm_amd

using namespace lpm_amd;

extern "C" {

int lpm_ic2d_sums(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                  int64_t tgt_cs, int64_t n_tgt, const double* src_x,
                  int64_t src_es, int64_t src_cs, const double* src_zeta,
                  const double* src_area, const uint8_t* src_mask,
                  int64_t n_src, int collocated, double eps, int precision,
                  unsigned flags, double* out_u, int64_t out_es,
                  int64_t out_cs, double* out_psi, void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    LPM_ARG_CHECK(n_tgt >= 0 && n_src >= 0, "negative size");
    LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                  "unknown precision");
    LPM_ARG_CHECK(eps >= 0 && std::isfinite(eps), "eps must be >= 0");
    if (n_tgt > 0) {
      LPM_ARG_CHECK(tgt_x && tgt_es > 0 && tgt_cs > 0, "bad tgt_x");
      LPM_ARG_CHECK(!out_u || (out_es > 0 && out_cs > 0), "bad out_u strides");
    }
    if (n_src > 0) {
      LPM_ARG_CHECK(src_x && src_es > 0 && src_cs > 0, "bad src_x");
      LPM_ARG_CHECK(src_zeta && src_area, "src_zeta/src_area are NULL");
    }
    if (collocated)
      LPM_ARG_CHECK(n_tgt == n_src,
                    "collocated: targets must be the sources (n_tgt == n_src)");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const RankMap map = build_rank(ctx, src_mask, n_src, s);
    PackedSources ps = pack_sources(ctx, CrdIn{src_x, src_es, src_cs}, src_zeta,
                                    src_area, map, precision, s);
    TargetSeg t;
    if (collocated) {
      t = face_targets(CrdIn{tgt_x, tgt_es, tgt_cs}, map, src_mask,
                       (flags & LPM_FLAG_SKIP_MASKED_TARGETS) != 0,
                       CrdOut{out_u, out_es, out_cs}, out_psi);
    } else {
      t.x = CrdIn{tgt_x, tgt_es, tgt_cs};
      t.n = n_tgt;
      t.out = CrdOut{out_u, out_es, out_cs};
      t.out_scalar = out_psi;
    }
    launch_ic2d(ctx, ps, &t, 1, eps, collocated && eps == 0.0, precision, s);
    if ((flags & LPM_FLAG_CHECK_FINITE) && out_u)
      check_finite(ctx, out_u, n_tgt, out_es, out_cs, 3, s);
  });
}

int lpm_bve_solve(lpm_ctx* ctx, const double* tgt_x, int64_t tgt_es,
                  int64_t tgt_cs, int64_t n_tgt, const double* src_x,
                  int64_t src_es, int64_t src_cs, const double* src_zeta,
                  const double* src_area, const uint8_t* src_mask,
                  int64_t n_src, int collocated, int precision, unsigned flags,
                  double* out_u, int64_t out_es, int64_t out_cs,
                  double* out_psi, void* stream) {
  // BVEVertexSolve / BVEFaceSolve are the kernel_vals sums without
  // smoothing: one fused pass gives both of their reductions
  return lpm_ic2d_sums(ctx, tgt_x, tgt_es, tgt_cs, n_tgt, src_x, src_es,
                       src_cs, src_zeta, src_area, src_mask, n_src, collocated,
                       0.0, precision, flags, out_u, out_es, out_cs, out_psi,
                       stream);
}

int lpm_ic2d_init_direct_sums(lpm_ctx* ctx, const lpm_ic2d_fields* f,
                              double eps, int precision, unsigned flags,
                              void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    validate(f);
    LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                  "unknown precision");
    LPM_ARG_CHECK(eps >= 0 && std::isfinite(eps), "eps must be >= 0");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    const RankMap map = build_rank(ctx, f->face_mask, f->n_faces, s);
    ic2d_eval(ctx, *f, CrdIn{f->vert_x, f->vert_x_es, f->vert_x_cs},
              CrdIn{f->face_x, f->face_x_es, f->face_x_cs}, f->face_zeta, map,
              eps, precision, flags, s);
  });
}

int lpm_ic2d_rk2_step(lpm_ctx* ctx, const lpm_ic2d_fields* f, double dt,
                      double Omega, double eps, int precision, unsigned flags,
                      void* stream) {
  return guarded(ctx, [&] {
    LPM_ARG_CHECK(ctx, "null context");
    validate(f);
    LPM_ARG_CHECK(precision == LPM_FP64 || precision == LPM_MIXED,
                  "unknown precision");
    LPM_ARG_CHECK(std::isfinite(dt) && std::isfinite(Omega),
                  "dt and Omega must be finite");
    LPM_ARG_CHECK(eps >= 0 && std::isfinite(eps), "eps must be >= 0");
    LPM_HIP_CHECK(hipSetDevice(ctx->device));
    hipStream_t s = as_stream(stream);
    order_stream(ctx, s);
    Ic2dStage V = carve(ctx->ic2d_v, f->n_verts);
    Ic2dStage F = carve(ctx->ic2d_f, f->n_faces);
    run_maybe_graph(
        ctx, (flags & LPM_RK4_USE_GRAPH) != 0, s,
        [&]() {
          const RankMap map = build_rank(ctx, f->face_mask, f->n_faces, s);
          pack_sources(ctx, CrdIn{f->face_x, f->face_x_es, f->face_x_cs},
                       f->face_zeta, f->face_area, map, precision, s);
          // the split workspace of the two evaluations, before any capture
          const int ks = ic2d_chunks(
              ctx, precision, nblk(f->n_verts) + nblk(f->n_faces), f->n_faces);
          if (ks > 1)
            split_workspace(ctx, ic2d_split_bytes(ks, f->n_verts + f->n_faces),
                            s);
          return ic2d_key(*f, dt, Omega, eps, precision, flags, s);
        },
        [&]() {
          issue_rk2(ctx, *f, dt, Omega, eps, precision, flags, V, F, s);
        });
    if (flags & LPM_FLAG_CHECK_FINITE) {
      check_finite(ctx, f->vert_u, f->n_verts, f->vert_u_es, f->vert_u_cs, 3, s);
      check_finite(ctx, f->face_u, f->n_faces, f->face_u_es, f->face_u_cs, 3, s);
    }
  });
}

}  // extern "C"
