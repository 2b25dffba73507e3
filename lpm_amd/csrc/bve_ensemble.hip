// Ensembles of independent BVE problems evaluated by one persistent launch.
//
// Members are stored back to back in one set of arrays (lpm_bve_ensemble).
// Per evaluation:
//   1. the face mask of the whole ensemble is compacted by the same scan as
//      a single problem (build_rank); member m's undivided faces are then
//      the compacted ranks [R_m, R_m + cnt_m);
//   2. each member's sources are packed into its own region of the packed
//      array, starting at P_m, a multiple of the 256-source packing tile, and
//      padded with zero-strength sources to a whole tile, so that every
//      member's tiles (and mixed-precision tile centres) are exactly those a
//      one-member call would build;
//   3. a persistent kernel (grid = resident blocks of the device) takes
//      (member, segment, target block) work items from a queue through one
//      atomic counter; items are ordered by decreasing source count, so the
//      largest members start first and small ones fill the tail.
// A work item runs the same block loop (bve_device.hpp) on the same target
// slots and source tiles as a one-member launch, so results agree bit for
// bit with lpm_bve_init_velocity on each member alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "bve_device.hpp"
#include "internal.hpp"

namespace lpm_amd {
namespace {

// Device member table.  v0/nv/f0/nf come from the host offsets; R (first
// compacted rank), cnt (undivided faces) and P (first packed slot) are
// computed on the device from the mask at every evaluation.
struct EnsMember {
  int64_t v0, nv, f0, nf;
  int32_t R, cnt, P, pad;
};

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
};

// R_b = undivided faces before member b's first face (b = n_members: all).
__global__ __launch_bounds__(64) void ens_bounds_kernel(
    const uint8_t* __restrict__ mask, const int32_t* __restrict__ block_offsets,
    int64_t n_faces, const EnsMember* __restrict__ mem, int n_members,
    int32_t* __restrict__ bounds) {
  const int b = blockIdx.x;
  const int64_t pos = b < n_members ? mem[b].f0 : n_faces;
  if (!mask) {
    if (threadIdx.x == 0) bounds[b] = static_cast<int32_t>(pos);
    return;
  }
  const int64_t chunk = pos / kMaskScanChunk;
  int c = 0;
  for (int64_t i = chunk * kMaskScanChunk + threadIdx.x; i < pos; i += 64)
    c += mask[i] ? 0 : 1;
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if (threadIdx.x == 0) bounds[b] = block_offsets[chunk] + c;
}

// cnt_m, R_m and P_m = sum over earlier members of cnt rounded up to the
// packing tile; *padded_total = P after the last member.
__global__ __launch_bounds__(1024) void ens_layout_kernel(
    EnsMember* __restrict__ mem, int n_members,
    const int32_t* __restrict__ bounds, int32_t* __restrict__ padded_total) {
  __shared__ int32_t buf[1024];
  __shared__ int32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int m0 = 0; m0 < n_members; m0 += 1024) {
    const int m = m0 + threadIdx.x;
    const int cnt = m < n_members ? bounds[m + 1] - bounds[m] : 0;
    const int padded = (cnt + kPackTile - 1) / kPackTile * kPackTile;
    buf[threadIdx.x] = padded;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const int add = threadIdx.x >= off ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    if (m < n_members) {
      mem[m].R = bounds[m];
      mem[m].cnt = cnt;
      mem[m].P = carry + buf[threadIdx.x] - padded;
    }
    __syncthreads();
    if (threadIdx.x == 1023) carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *padded_total = carry;
}

// grid (ceil(max member faces / 256), n_members): packed slot P_m + c holds
// member m's c-th undivided face, or a zero source for the tile padding.
__global__ __launch_bounds__(256) void ens_pack_f64_kernel(
    const double* __restrict__ sx, int64_t es, int64_t cs,
    const double* __restrict__ zeta, const double* __restrict__ area,
    const int32_t* __restrict__ inv, const EnsMember* __restrict__ mem,
    double4* __restrict__ out) {
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
}

__global__ void ens_zero_kernel(int32_t* p) { *p = 0; }

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

template <int BLOCK, int TPL, int NR, bool PAIR>
__global__ __launch_bounds__(BLOCK) void ens_f64_kernel(
    const double4* __restrict__ src, EnsArgs A) {
  constexpr int TILE = 256;
  __shared__ double4 tile[TILE];
  __shared__ int slot;
  for (;;) {
    const int it = ens_next(A, &slot);
    if (it >= A.n_items) break;
    const int4 w = A.items[it];
    const EnsMember M = A.mem[w.x];
    const SegArgs sg = ens_segment(A, M, w.y);
    f64_block<BLOCK, TPL, TILE, false, NR, PAIR>(src + M.P, M.cnt, sg, w.z,
                                                 0.0, tile);
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

// ------------------------------------------------------ kernel registry ---
// Every variant has a one-problem twin of the same shape in bve_sums.hip.
using EnsF64Fn = void (*)(const double4*, EnsArgs);
using EnsMixedFn = void (*)(const float4*, const double4*, EnsArgs);

struct EnsVariant {
  EnsF64Fn f64;
  EnsMixedFn mixed;
  int block, tpl;
  const char* name;
  const void* fn() const {
    return f64 ? reinterpret_cast<const void*>(f64)
               : reinterpret_cast<const void*>(mixed);
  }
};

#define LPM_ENS_F64(B, T, NR, PAIR)                  \
  {ens_f64_kernel<B, T, NR, PAIR>, nullptr, B, T,    \
   "persistent_block" #B "_tpl" #T "_nr" #NR "_pair" #PAIR}
#define LPM_ENS_MIXED(B, T, PAIR)                 \
  {nullptr, ens_mixed_kernel<B, T, PAIR>, B, T,   \
   "persistent_block" #B "_tpl" #T "_pair" #PAIR}

const EnsVariant kEnsF64[] = {
    LPM_ENS_F64(256, 2, 2, true),  // 0: fixed default
    LPM_ENS_F64(256, 2, 1, true),  // 1
    LPM_ENS_F64(256, 1, 1, true),  // 2
    LPM_ENS_F64(256, 4, 1, true),  // 3
    LPM_ENS_F64(128, 2, 2, true),  // 4
};
const EnsVariant kEnsMixed[] = {
    LPM_ENS_MIXED(256, 2, false),  // 0: fixed default
    LPM_ENS_MIXED(256, 1, true),   // 1
    LPM_ENS_MIXED(256, 1, false),  // 2
    LPM_ENS_MIXED(256, 2, true),   // 3
    LPM_ENS_MIXED(256, 4, false),  // 4
    LPM_ENS_MIXED(128, 2, true),   // 5
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

// Work list and member table of one offset layout, for one target-block
// size.  Rebuilt (and uploaded synchronously) when either changes.
struct lpm_ctx::EnsemblePlan {
  std::vector<int64_t> voff, foff;
  int per = 0;
  int n_items = 0;
  double pairs = 0;  // sum of (member targets) x (member faces)
  uint64_t serial = 0;
};

namespace lpm_amd {

EnsembleDesc validate_ensemble(const lpm_bve_ensemble* e) {
  LPM_ARG_CHECK(e, "null ensemble");
  const int M = e->n_members;
  LPM_ARG_CHECK(M >= 1 && M <= 65535, "n_members must be in [1, 65535]");
  LPM_ARG_CHECK(e->vert_offsets && e->face_offsets, "null member offsets");
  const lpm_bve_fields& f = e->fields;
  LPM_ARG_CHECK(e->vert_offsets[0] == 0 && e->face_offsets[0] == 0,
                "member offsets must start at 0");
  LPM_ARG_CHECK(e->vert_offsets[M] == f.n_verts,
                "vert_offsets[n_members] must equal n_verts");
  LPM_ARG_CHECK(e->face_offsets[M] == f.n_faces,
                "face_offsets[n_members] must equal n_faces");
  EnsembleDesc d;
  d.e = e;
  for (int m = 0; m < M; ++m) {
    const int64_t nv = e->vert_offsets[m + 1] - e->vert_offsets[m];
    const int64_t nf = e->face_offsets[m + 1] - e->face_offsets[m];
    LPM_ARG_CHECK(nv >= 0, "vert_offsets must be nondecreasing");
    LPM_ARG_CHECK(nf >= 1, "every member needs at least one face");
    d.max_member_faces = std::max(d.max_member_faces, nf);
    d.total_padded_upper += (nf + kPackTile - 1) / kPackTile * kPackTile;
  }
  LPM_ARG_CHECK(d.total_padded_upper < (int64_t(1) << 31),
                "ensemble too large for 32-bit packed indices");
  return d;
}

namespace {

std::shared_ptr<lpm_ctx::EnsemblePlan> ensure_plan(lpm_ctx* ctx,
                                                   const lpm_bve_ensemble& e,
                                                   int per) {
  const int M = e.n_members;
  auto& p = ctx->ens_plan;
  if (p && p->per == per && static_cast<int>(p->voff.size()) == M + 1 &&
      std::equal(p->voff.begin(), p->voff.end(), e.vert_offsets) &&
      std::equal(p->foff.begin(), p->foff.end(), e.face_offsets))
    return p;
  auto np = std::make_shared<lpm_ctx::EnsemblePlan>();
  np->voff.assign(e.vert_offsets, e.vert_offsets + M + 1);
  np->foff.assign(e.face_offsets, e.face_offsets + M + 1);
  np->per = per;
  std::vector<EnsMember> mem(M);
  struct Item {
    int64_t cost;
    int4 w;
  };
  std::vector<Item> items;
  for (int m = 0; m < M; ++m) {
    EnsMember& x = mem[m];
    x.v0 = np->voff[m];
    x.nv = np->voff[m + 1] - np->voff[m];
    x.f0 = np->foff[m];
    x.nf = np->foff[m + 1] - np->foff[m];
    x.R = x.cnt = x.P = x.pad = 0;
    np->pairs += static_cast<double>(x.nv + x.nf) * static_cast<double>(x.nf);
    const int64_t vb = (x.nv + per - 1) / per, fb = (x.nf + per - 1) / per;
    for (int64_t b = 0; b < vb; ++b)
      items.push_back({x.nf, make_int4(m, 0, static_cast<int>(b), 0)});
    for (int64_t b = 0; b < fb; ++b)
      items.push_back({x.nf, make_int4(m, 1, static_cast<int>(b), 0)});
  }
  LPM_ARG_CHECK(items.size() < (size_t(1) << 31), "too many work items");
  std::stable_sort(items.begin(), items.end(),
                   [](const Item& a, const Item& b) { return a.cost > b.cost; });
  std::vector<int4> w(items.size());
  for (size_t k = 0; k < items.size(); ++k) w[k] = items[k].w;
  np->n_items = static_cast<int>(w.size());
  void* dm = ctx->ens_members.ensure(sizeof(EnsMember) * M);
  void* di = ctx->ens_items.ensure(sizeof(int4) * std::max<size_t>(w.size(), 1));
  LPM_HIP_CHECK(hipMemcpy(dm, mem.data(), sizeof(EnsMember) * M,
                          hipMemcpyHostToDevice));
  if (!w.empty())
    LPM_HIP_CHECK(hipMemcpy(di, w.data(), sizeof(int4) * w.size(),
                            hipMemcpyHostToDevice));
  np->serial = ++ctx->ens_plan_serial;
  p = np;
  return p;
}

int resident_blocks(lpm_ctx* ctx, const EnsVariant& v) {
  int per_cu = 0;
  LPM_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, v.fn(),
                                                             v.block, 0));
  return std::max(1, per_cu) * std::max(1, ctx->num_cus);
}

}  // namespace

namespace {

// Allocates every workspace an evaluation of this ensemble uses (so that a
// stream capture that follows allocates nothing) and returns the plan.
std::shared_ptr<lpm_ctx::EnsemblePlan> prepare(lpm_ctx* ctx,
                                               const EnsembleDesc& d,
                                               int precision) {
  const lpm_bve_ensemble& e = *d.e;
  const int family =
      precision == LPM_MIXED ? KID_BVE_ENS_MIXED : KID_BVE_ENS_F64;
  const EnsVariant& v = ens_variant(
      family, pick_variant(ctx, family, e.fields.n_verts + e.fields.n_faces));
  auto plan = ensure_plan(ctx, e, v.block * v.tpl);
  ctx->ens_bounds.ensure(sizeof(int32_t) * (e.n_members + 1));
  ctx->ens_queue.ensure(sizeof(int32_t) * 2);
  const int64_t cap = std::max<int64_t>(d.total_padded_upper, kPackTile);
  ctx->packed64.ensure(sizeof(double4) * cap);
  if (precision == LPM_MIXED) {
    ctx->packed32.ensure(sizeof(float4) * cap);
    ctx->centers.ensure(sizeof(double4) * (cap / kPackTile));
  }
  return plan;
}

}  // namespace

uint64_t ensemble_prepare(lpm_ctx* ctx, const EnsembleDesc& d, int precision) {
  return prepare(ctx, d, precision)->serial;
}

void ensemble_velocity(lpm_ctx* ctx, const EnsembleDesc& d, CrdIn vx,
                       CrdIn fx, const double* fzeta, const RankMap& map,
                       int precision, unsigned flags, hipStream_t s) {
  const lpm_bve_ensemble& e = *d.e;
  const lpm_bve_fields& f = e.fields;
  const int M = e.n_members;
  const int family =
      precision == LPM_MIXED ? KID_BVE_ENS_MIXED : KID_BVE_ENS_F64;
  const EnsVariant& v =
      ens_variant(family, pick_variant(ctx, family, f.n_verts + f.n_faces));
  auto plan = prepare(ctx, d, precision);
  LPM_ARG_CHECK(plan->per == v.block * v.tpl, "ensemble plan mismatch");

  EnsMember* mem = ctx->ens_members.as<EnsMember>();
  int32_t* bounds = ctx->ens_bounds.as<int32_t>();
  int32_t* queue = ctx->ens_queue.as<int32_t>();
  int32_t* padded_total = queue + 1;
  ens_bounds_kernel<<<M + 1, 64, 0, s>>>(f.face_mask, map.block_offsets,
                                          f.n_faces, mem, M, bounds);
  LPM_HIP_CHECK(hipGetLastError());
  ens_layout_kernel<<<1, 1024, 0, s>>>(mem, M, bounds, padded_total);
  LPM_HIP_CHECK(hipGetLastError());
  double4* p64 = ctx->packed64.as<double4>();
  const dim3 pgrid(static_cast<unsigned>((d.max_member_faces + 255) / 256),
                   static_cast<unsigned>(M));
  ens_pack_f64_kernel<<<pgrid, 256, 0, s>>>(fx.base, fx.es, fx.cs, fzeta,
                                             f.face_area, map.inv, mem, p64);
  LPM_HIP_CHECK(hipGetLastError());

  EnsArgs A{};
  A.mem = mem;
  A.items = ctx->ens_items.as<int4>();
  A.n_items = plan->n_items;
  A.queue = queue;
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
  if (A.n_items == 0) return;
  const float4* p32 = nullptr;
  const double4* cen = nullptr;
  if (precision == LPM_MIXED) {
    const int ntile = static_cast<int>(d.total_padded_upper / kPackTile);
    p32 = ctx->packed32.as<float4>();
    cen = ctx->centers.as<double4>();
    pack_mixed_kernel<<<ntile, kPackTile, 0, s>>>(
        p64, padded_total, ctx->packed32.as<float4>(),
        ctx->centers.as<double4>());
    LPM_HIP_CHECK(hipGetLastError());
  }
  ens_zero_kernel<<<1, 1, 0, s>>>(queue);
  LPM_HIP_CHECK(hipGetLastError());
  const int grid = std::min(resident_blocks(ctx, v), A.n_items);
  TimedLaunch tl(ctx, family, s, plan->pairs);
  if (precision == LPM_MIXED)
    hipLaunchKernelGGL(v.mixed, dim3(grid), dim3(v.block), 0, s, p32, cen, A);
  else
    hipLaunchKernelGGL(v.f64, dim3(grid), dim3(v.block), 0, s,
                       static_cast<const double4*>(p64), A);
  LPM_HIP_CHECK(hipGetLastError());
}

}  // namespace lpm_amd
