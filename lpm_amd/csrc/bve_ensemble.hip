// BVE ensembles: every member's velocity from one persistent launch.  The
// member layout and work plan come from ensemble.hip; a work item runs the
// same block loop (bve_device.hpp) on the same target slots and source tiles
// as a one-member launch, so results agree bit for bit with
// lpm_bve_init_velocity on each member alone.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "bve_device.hpp"
#include "internal.hpp"

namespace lpm_amd {
namespace {

struct EnsArgs {
  const EnsMember* mem;
  const int4* items;  // {member, segment (0 vertices, 1 faces), block, 0}
  int n_items;
  int32_t* queue;
  const double* vx;
  int64_t vxes, vxcs;
  double* vu;
  int64_t vues, vucs;
  const double* fx;
  int64_t fxes, fxcs;
  double* fu;
  int64_t fues, fucs;
  const int32_t* rank;  // compacted rank per face (nullptr: no mask)
  const int32_t* inv;   // face per compacted rank (nullptr: no mask)
  int skip_masked;
  const int32_t* vorder;  // vertex slot -> vertex (nullptr: identity)
};

// grid (ceil(max member faces / 256), n_members): packed slot P_m + c holds
// member m's c-th undivided face, or a zero source for the tile padding.
__global__ __launch_bounds__(256) void ens_pack_f64_kernel(
    const double* __restrict__ sx, int64_t es, int64_t cs,
    const double* __restrict__ zeta, const double* __restrict__ area,
    const int32_t* __restrict__ inv, const EnsMember* __restrict__ mem,
    double4* __restrict__ out, double4* __restrict__ out_sy) {
  const EnsMember M = mem[blockIdx.y];
  const int c = blockIdx.x * 256 + threadIdx.x;
  const int padded = (M.cnt + kPackTile - 1) / kPackTile * kPackTile;
  if (c >= padded) return;
  double4 v = make_double4(0.0, 0.0, 0.0, 0.0);
  if (c < M.cnt) {
    const int64_t r = static_cast<int64_t>(M.R) + c;
    const int64_t i = inv ? inv[r] : r;
    const double* p = sx + i * es;
    v = make_double4(p[0], p[cs], p[2 * cs], -zeta[i] * area[i] * kInv4Pi);
  }
  out[M.P + c] = v;
  // {s y, 0} for the kernels that take the weight from it
  if (out_sy) out_sy[M.P + c] = make_double4(v.w * v.x, v.w * v.y, v.w * v.z, 0.0);
}

// Targets of work item segment `seg` of member M, as a SegArgs whose self
// indices count from the member's first packed source.
__device__ __forceinline__ SegArgs ens_segment(const EnsArgs& A,
                                               const EnsMember& M, int seg) {
  SegArgs sg{};
  if (seg == 0) {
    sg.x = A.vx;
    sg.es = A.vxes;
    sg.cs = A.vxcs;
    sg.n = M.nv;
    sg.slot_base = M.v0;
    if (A.vorder) sg.index = A.vorder + M.v0;
    sg.out = A.vu;
    sg.oes = A.vues;
    sg.ocs = A.vucs;
    return sg;
  }
  sg.x = A.fx;
  sg.es = A.fxes;
  sg.cs = A.fxcs;
  sg.out = A.fu;
  sg.oes = A.fues;
  sg.ocs = A.fucs;
  if (A.skip_masked && A.inv) {
    // compacted targets: slot c is the member's c-th undivided face, which
    // is also its c-th packed source
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

// Takes the next work item for the block; every block leaves the loop when
// the queue passes n_items.
__device__ __forceinline__ int ens_next(const EnsArgs& A, int* slot) {
  if (threadIdx.x == 0) *slot = atomicAdd(A.queue, 1);
  __syncthreads();
  const int it = *slot;
  __syncthreads();
  return it;
}

template <int BLOCK, int TPL, int NR, bool PAIR, bool FUSE = false,
          bool QUAD = false, bool SY = false>
__global__ __launch_bounds__(BLOCK) void ens_f64_kernel(
    const double4* __restrict__ src, const double4* __restrict__ src_sy,
    EnsArgs A) {
  constexpr int TILE = 256;
  __shared__ double4 tile[TILE];
  __shared__ double4 tile_sy[SY ? TILE : 1];
  __shared__ int slot;
  for (;;) {
    const int it = ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SegArgs sg = ens_segment(A, M, w.y);
    f64_block<BLOCK, TPL, TILE, false, NR, PAIR, FUSE, QUAD, SY>(
        src + M.P, SY ? src_sy + M.P : nullptr, M.cnt, sg, w.z, 0.0, tile,
        tile_sy);
  }
}

template <int BLOCK, int TPL, bool PAIR>
__global__ __launch_bounds__(BLOCK) void ens_mixed_kernel(
    const float4* __restrict__ src, const double4* __restrict__ centers,
    EnsArgs A) {
  constexpr int TILE = 256;
  __shared__ float4 tile[TILE];
  __shared__ int slot;
  for (;;) {
    const int it = ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SegArgs sg = ens_segment(A, M, w.y);
    mixed_block<BLOCK, TPL, TILE, false, PAIR>(
        src + M.P, centers + M.P / kPackTile, M.cnt, sg, w.z, 0.0f, tile);
  }
}

__global__ __launch_bounds__(256) void ens_mfma_kernel(MfmaSrc src, EnsArgs A) {
  __shared__ float4 ly[kPackTile];
  __shared__ float lq[kPackTile];
  __shared__ float ls[kPackTile];
  __shared__ float4 lt[kPackTile];
  __shared__ int slot;
  for (;;) {
    const int it = ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SegArgs sg = ens_segment(A, M, w.y);
    const int t0 = M.P / kPackTile;
    MfmaSrc ms{src.y + M.P, src.q + M.P, src.s + M.P,
               src.t + static_cast<int64_t>(t0) * 1024, src.c + t0};
    mfma_block<false>(ms, M.cnt, sg, w.z, 0.0f, ly, lq, ls, lt);
  }
}

template <int MINW, int NACC = 1, int RS = 1>
__global__ __launch_bounds__(256, MINW) void ens_mfma16_kernel(Mfma16Src src,
                                                               EnsArgs A) {
  __shared__ float4 ly[kPackTile];
  __shared__ float ls[kPackTile];
  __shared__ uint4 la[2 * kPackTile];
  __shared__ float4 lt[kPackTile];
  __shared__ int slot;
  for (;;) {
    const int it = ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SegArgs sg = ens_segment(A, M, w.y);
    const int t0 = M.P / kPackTile;
    Mfma16Src ms{src.y + M.P, src.s + M.P, src.a + 2 * static_cast<int64_t>(M.P),
                 src.t + static_cast<int64_t>(t0) * 1024, src.c + t0,
                 src.is + t0};
    mfma16_block<false, NACC, 1, RS>(ms, M.cnt, sg, w.z, 0.0f, ly, ls, la,
                                     lt);
  }
}

// ------------------------------------------------------ kernel registry ---
// Every variant has a one-problem twin of the same shape in bve_sums.hip.
using EnsF64Fn = void (*)(const double4*, const double4*, EnsArgs);
using EnsMixedFn = void (*)(const float4*, const double4*, EnsArgs);
using EnsMfmaFn = void (*)(MfmaSrc, EnsArgs);
using EnsMfma16Fn = void (*)(Mfma16Src, EnsArgs);

struct EnsVariant {
  EnsF64Fn f64;
  EnsMixedFn mixed;
  int block, tpl;
  const char* name;
  EnsMfmaFn mfma;
  EnsMfma16Fn mfma16;
  const void* fn() const {
    return f64 ? reinterpret_cast<const void*>(f64)
           : mixed ? reinterpret_cast<const void*>(mixed)
           : mfma ? reinterpret_cast<const void*>(mfma)
                  : reinterpret_cast<const void*>(mfma16);
  }
};

#define LPM_ENS_F64(B, T, NR, PAIR)                  \
  {ens_f64_kernel<B, T, NR, PAIR>, nullptr, B, T,    \
   "persistent_block" #B "_tpl" #T "_nr" #NR "_pair" #PAIR, nullptr}
#define LPM_ENS_MIXED(B, T, PAIR)                 \
  {nullptr, ens_mixed_kernel<B, T, PAIR>, B, T,   \
   "persistent_block" #B "_tpl" #T "_pair" #PAIR, nullptr}

const EnsVariant kEnsF64[] = {
    LPM_ENS_F64(256, 2, 2, true),  // 0: fixed default
    LPM_ENS_F64(256, 2, 1, true),  // 1
    LPM_ENS_F64(256, 1, 1, true),  // 2
    LPM_ENS_F64(256, 4, 1, true),  // 3
    LPM_ENS_F64(128, 2, 2, true),  // 4
    {ens_f64_kernel<256, 2, 1, true, true>, nullptr, 256, 2,
     "persistent_block256_tpl2_nr1_pairtrue_fused", nullptr},  // 5
    {ens_f64_kernel<256, 4, 1, true, true>, nullptr, 256, 4,
     "persistent_block256_tpl4_nr1_pairtrue_fused", nullptr},  // 6
    {ens_f64_kernel<256, 4, 1, true, true, true>, nullptr, 256, 4,
     "persistent_block256_tpl4_nr1_quad_fused", nullptr},      // 7
    {ens_f64_kernel<256, 4, 1, true, true, true, true>, nullptr, 256, 4,
     "persistent_block256_tpl4_nr1_quad_fused_sy", nullptr},   // 8: s y packed
};
const EnsVariant kEnsMixed[] = {
    LPM_ENS_MIXED(256, 2, false),  // 0: fixed default
    LPM_ENS_MIXED(256, 1, true),   // 1
    LPM_ENS_MIXED(256, 1, false),  // 2
    LPM_ENS_MIXED(256, 2, true),   // 3
    LPM_ENS_MIXED(256, 4, false),  // 4
    LPM_ENS_MIXED(128, 2, true),   // 5
    {nullptr, nullptr, 256, 1, "persistent_mfma_block256_near3",
     ens_mfma_kernel},             // 6: matrix-core far field
    {nullptr, nullptr, 256, 1, "persistent_mfma16_block256_near3", nullptr,
     ens_mfma16_kernel<1>},        // 7: f16-split distance product
    {nullptr, nullptr, 256, 1, "persistent_mfma16_occ5_block256_near3", nullptr,
     ens_mfma16_kernel<5>},        // 8: register budget for 5 waves per SIMD
    {nullptr, nullptr, 256, 1, "persistent_mfma16_occ5_acc2_block256_near3",
     nullptr, ens_mfma16_kernel<5, 2>},  // 9: two weight-product accumulators
    // 10-11: one reciprocal per four distances
    {nullptr, nullptr, 256, 1, "persistent_mfma16_occ5_qrcp_block256_near3",
     nullptr, ens_mfma16_kernel<5, 1, 4>},
    {nullptr, nullptr, 256, 1,
     "persistent_mfma16_occ5_acc2_qrcp_block256_near3", nullptr,
     ens_mfma16_kernel<5, 2, 4>},
};
constexpr int kNumEnsF64 = sizeof(kEnsF64) / sizeof(kEnsF64[0]);
constexpr int kNumEnsMixed = sizeof(kEnsMixed) / sizeof(kEnsMixed[0]);

const EnsVariant& ens_variant(int family, int v) {
  return family == KID_BVE_ENS_F64 ? kEnsF64[v] : kEnsMixed[v];
}

}  // namespace

int ens_variant_count(int family) {
  return family == KID_BVE_ENS_F64 ? kNumEnsF64 : kNumEnsMixed;
}
const char* ens_variant_name(int family, int v) {
  return ens_variant(family, v).name;
}
int ens_default_variant(int) { return 0; }

}  // namespace lpm_amd

namespace lpm_amd {

namespace {

const EnsVariant& pick_ens_variant(lpm_ctx* ctx, const EnsembleDesc& d,
                                   int precision) {
  const int family =
      precision == LPM_MIXED ? KID_BVE_ENS_MIXED : KID_BVE_ENS_F64;
  return ens_variant(family,
                     pick_variant(ctx, family, d.n_verts + d.n_faces));
}

void ensure_sources(lpm_ctx* ctx, const EnsembleDesc& d, int precision) {
  const int64_t cap = std::max<int64_t>(d.total_padded_upper, kPackTile);
  ctx->packed64.ensure(sizeof(double4) * cap);
  if (precision == LPM_FP64) ctx->packed64sy.ensure(sizeof(double4) * cap);
  if (precision == LPM_MIXED) {
    ctx->packed32.ensure(sizeof(float4) * cap);
    ctx->centers.ensure(sizeof(double4) * (cap / kPackTile));
    ctx->mf_y.ensure(sizeof(float4) * cap);
    ctx->mf_q.ensure(sizeof(float) * cap);
    ctx->mf_s.ensure(sizeof(float) * cap);
    ctx->mf_t.ensure(sizeof(float) * 4 * cap);
    ctx->mf_c.ensure(sizeof(double4) * (cap / kPackTile));
    ctx->mf_a.ensure(sizeof(uint4) * 2 * cap);
    ctx->mf_t32.ensure(sizeof(float) * 4 * cap);
    ctx->mf_is.ensure(sizeof(double) * (cap / kPackTile));
  }
}

}  // namespace

// Vertex slot order of the mixed sums, from the members' state positions
// (also used for the stage evaluations of a step, whose positions are close).
const int32_t* bve_ens_vertex_order(lpm_ctx* ctx, const EnsembleDesc& d,
                                    const lpm_bve_fields& f, hipStream_t s) {
  return target_order(ctx, CrdIn{f.vert_x, f.vert_x_es, f.vert_x_cs},
                      d.n_verts, d.voff, d.n_members, s);
}

uint64_t ensemble_prepare(lpm_ctx* ctx, const EnsembleDesc& d,
                          const lpm_bve_fields& f, int precision,
                          hipStream_t s) {
  const EnsVariant& v = pick_ens_variant(ctx, d, precision);
  const EnsLayout L = ensemble_plan(ctx, d, v.block * v.tpl);
  ensure_sources(ctx, d, precision);
  if (precision == LPM_MIXED) bve_ens_vertex_order(ctx, d, f, s);
  return L.serial;
}

void ensemble_velocity(lpm_ctx* ctx, const EnsembleDesc& d,
                       const lpm_bve_fields& f, CrdIn vx, CrdIn fx,
                       const double* fzeta, const RankMap& map, int precision,
                       unsigned flags, hipStream_t s) {
  const int M = d.n_members;
  const int family =
      precision == LPM_MIXED ? KID_BVE_ENS_MIXED : KID_BVE_ENS_F64;
  const EnsVariant& v = pick_ens_variant(ctx, d, precision);
  ensure_sources(ctx, d, precision);
  const EnsLayout L =
      ensemble_layout(ctx, d, map, f.face_mask, v.block * v.tpl, s);

  double4* p64 = ctx->packed64.as<double4>();
  double4* psy =
      precision == LPM_FP64 ? ctx->packed64sy.as<double4>() : nullptr;
  const dim3 pgrid(static_cast<unsigned>((d.max_member_faces + 255) / 256),
                   static_cast<unsigned>(M));
  ens_pack_f64_kernel<<<pgrid, 256, 0, s>>>(fx.base, fx.es, fx.cs, fzeta,
                                             f.face_area, map.inv, L.mem, p64,
                                             psy);
  LPM_HIP_CHECK(hipGetLastError());

  EnsArgs A{};
  A.mem = L.mem;
  A.items = L.items;
  A.n_items = L.n_items;
  A.queue = L.queue;
  A.vx = vx.base;
  A.vxes = vx.es;
  A.vxcs = vx.cs;
  A.vu = f.vert_u;
  A.vues = f.vert_u_es;
  A.vucs = f.vert_u_cs;
  A.fx = fx.base;
  A.fxes = fx.es;
  A.fxcs = fx.cs;
  A.fu = f.face_u;
  A.fues = f.face_u_es;
  A.fucs = f.face_u_cs;
  A.rank = map.rank;
  A.inv = map.inv;
  A.skip_masked = (flags & LPM_FLAG_SKIP_MASKED_TARGETS) ? 1 : 0;
  if (precision == LPM_MIXED) A.vorder = bve_ens_vertex_order(ctx, d, f, s);
  if (A.n_items == 0) return;
  const float4* p32 = nullptr;
  const double4* cen = nullptr;
  MfmaSrc ms{};
  Mfma16Src ms16{};
  if (precision == LPM_MIXED) {
    const int ntile = static_cast<int>(d.total_padded_upper / kPackTile);
    p32 = ctx->packed32.as<float4>();
    cen = ctx->centers.as<double4>();
    if (v.mfma || v.mfma16) {
      ms = MfmaSrc{ctx->mf_y.as<float4>(), ctx->mf_q.as<float>(),
                   ctx->mf_s.as<float>(), ctx->mf_t.as<float>(),
                   ctx->mf_c.as<double4>()};
      // member padding is all-zero records, which the packer skips
      pack_mfma_kernel<<<ntile, kPackTile, 0, s>>>(
          p64, L.padded_total, ctx->mf_y.as<float4>(), ctx->mf_q.as<float>(),
          ctx->mf_s.as<float>(), ctx->mf_t.as<float>(),
          ctx->mf_c.as<double4>());
      if (v.mfma16) {
        LPM_HIP_CHECK(hipGetLastError());
        pack_mfma16_kernel<<<ntile, kPackTile, 0, s>>>(
            p64, L.padded_total, ctx->mf_c.as<double4>(), ctx->mf_a.as<uint4>(),
            ctx->mf_t32.as<float>(), ctx->mf_is.as<double>());
        ms16 = Mfma16Src{ctx->mf_y.as<float4>(), ctx->mf_s.as<float>(),
                         ctx->mf_a.as<uint4>(), ctx->mf_t32.as<float>(),
                         ctx->mf_c.as<double4>(), ctx->mf_is.as<double>()};
      }
    } else {
      pack_mixed_kernel<<<ntile, kPackTile, 0, s>>>(
          p64, L.padded_total, ctx->packed32.as<float4>(),
          ctx->centers.as<double4>());
    }
    LPM_HIP_CHECK(hipGetLastError());
  }
  ensemble_reset_queue(L, s);
  const int grid = ensemble_grid(ctx, v.fn(), v.block, A.n_items);
  TimedLaunch tl(ctx, family, s, L.pairs);
  if (v.mfma16)
    hipLaunchKernelGGL(v.mfma16, dim3(grid), dim3(v.block), 0, s, ms16, A);
  else if (v.mfma)
    hipLaunchKernelGGL(v.mfma, dim3(grid), dim3(v.block), 0, s, ms, A);
  else if (precision == LPM_MIXED)
    hipLaunchKernelGGL(v.mixed, dim3(grid), dim3(v.block), 0, s, p32, cen, A);
  else
    hipLaunchKernelGGL(v.f64, dim3(grid), dim3(v.block), 0, s,
                       static_cast<const double4*>(p64),
                       static_cast<const double4*>(psy), A);
  LPM_HIP_CHECK(hipGetLastError());
}

}  // namespace lpm_amd
