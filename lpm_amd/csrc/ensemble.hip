// Ensembles of independent problems stored back to back: member offsets,
// work plans and the device member table, shared by the BVE and SWE
// ensemble kernels.
//
// Per evaluation the face mask of the whole ensemble is compacted by the
// scan of a single problem (build_rank); member m's undivided faces are then
// the compacted ranks [R_m, R_m + cnt_m).  Each member's sources are packed
// into their own region starting at P_m, a multiple of the 256-source
// packing tile, so that every member's tiles are exactly those a
// one-member call would build.  A persistent kernel takes (member, segment,
// target block) work items from a queue through one atomic counter; items
// are ordered by decreasing member face count, so the largest members start
// first and small ones fill the tail.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <vector>

#include "internal.hpp"

// Work list and member table of one offset layout for one target-block
// size.  Rebuilt (and uploaded synchronously) when the offsets change.
struct lpm_ctx::EnsemblePlan {
  std::vector<int64_t> voff, foff;
  int per = 0;
  int n_items = 0;
  double pairs = 0;
  uint64_t serial = 0;
  lpm_amd::DevBuf members;  // EnsMember per member
  lpm_amd::DevBuf items;    // int4 per work item
};

namespace lpm_amd {
namespace {

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
    const int padded = (cnt + kEnsTile - 1) / kEnsTile * kEnsTile;
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

__global__ void ens_zero_kernel(int32_t* p) { *p = 0; }

std::shared_ptr<lpm_ctx::EnsemblePlan> plan_for(lpm_ctx* ctx,
                                                const EnsembleDesc& d,
                                                int per) {
  const int M = d.n_members;
  auto& slot = ctx->ens_plans[per];
  if (slot && static_cast<int>(slot->voff.size()) == M + 1 &&
      std::equal(slot->voff.begin(), slot->voff.end(), d.voff) &&
      std::equal(slot->foff.begin(), slot->foff.end(), d.foff))
    return slot;
  auto np = std::make_shared<lpm_ctx::EnsemblePlan>();
  np->voff.assign(d.voff, d.voff + M + 1);
  np->foff.assign(d.foff, d.foff + M + 1);
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
  void* dm = np->members.ensure(sizeof(EnsMember) * M);
  void* di = np->items.ensure(sizeof(int4) * std::max<size_t>(w.size(), 1));
  LPM_HIP_CHECK(hipMemcpy(dm, mem.data(), sizeof(EnsMember) * M,
                          hipMemcpyHostToDevice));
  if (!w.empty())
    LPM_HIP_CHECK(hipMemcpy(di, w.data(), sizeof(int4) * w.size(),
                            hipMemcpyHostToDevice));
  np->serial = ++ctx->ens_plan_serial;
  slot = np;
  return slot;
}

}  // namespace

EnsembleDesc validate_members(int32_t n_members, const int64_t* voff,
                              const int64_t* foff, int64_t n_verts,
                              int64_t n_faces) {
  const int M = n_members;
  LPM_ARG_CHECK(M >= 1 && M <= 65535, "n_members must be in [1, 65535]");
  LPM_ARG_CHECK(voff && foff, "null member offsets");
  LPM_ARG_CHECK(voff[0] == 0 && foff[0] == 0, "member offsets must start at 0");
  LPM_ARG_CHECK(voff[M] == n_verts,
                "vert_offsets[n_members] must equal n_verts");
  LPM_ARG_CHECK(foff[M] == n_faces,
                "face_offsets[n_members] must equal n_faces");
  EnsembleDesc d;
  d.n_members = M;
  d.voff = voff;
  d.foff = foff;
  d.n_verts = n_verts;
  d.n_faces = n_faces;
  for (int m = 0; m < M; ++m) {
    const int64_t nv = voff[m + 1] - voff[m];
    const int64_t nf = foff[m + 1] - foff[m];
    LPM_ARG_CHECK(nv >= 0, "vert_offsets must be nondecreasing");
    LPM_ARG_CHECK(nf >= 1, "every member needs at least one face");
    d.max_member_faces = std::max(d.max_member_faces, nf);
    d.total_padded_upper += (nf + kEnsTile - 1) / kEnsTile * kEnsTile;
  }
  LPM_ARG_CHECK(d.total_padded_upper < (int64_t(1) << 31),
                "ensemble too large for 32-bit packed indices");
  return d;
}

EnsLayout ensemble_plan(lpm_ctx* ctx, const EnsembleDesc& d, int per) {
  auto p = plan_for(ctx, d, per);
  ctx->ens_bounds.ensure(sizeof(int32_t) * (d.n_members + 1));
  int32_t* q = static_cast<int32_t*>(ctx->ens_queue.ensure(sizeof(int32_t) * 2));
  EnsLayout L;
  L.mem = p->members.as<EnsMember>();
  L.items = p->items.as<int4>();
  L.n_items = p->n_items;
  L.queue = q;
  L.padded_total = q + 1;
  L.pairs = p->pairs;
  L.serial = p->serial;
  return L;
}

EnsLayout ensemble_layout(lpm_ctx* ctx, const EnsembleDesc& d,
                          const RankMap& map, const uint8_t* face_mask,
                          int per, hipStream_t s) {
  EnsLayout L = ensemble_plan(ctx, d, per);
  const int M = d.n_members;
  int32_t* bounds = ctx->ens_bounds.as<int32_t>();
  ens_bounds_kernel<<<M + 1, 64, 0, s>>>(face_mask, map.block_offsets,
                                          d.n_faces, L.mem, M, bounds);
  LPM_HIP_CHECK(hipGetLastError());
  ens_layout_kernel<<<1, 1024, 0, s>>>(L.mem, M, bounds, L.padded_total);
  LPM_HIP_CHECK(hipGetLastError());
  return L;
}

void ensemble_reset_queue(const EnsLayout& L, hipStream_t s) {
  ens_zero_kernel<<<1, 1, 0, s>>>(L.queue);
  LPM_HIP_CHECK(hipGetLastError());
}

int ensemble_grid(lpm_ctx* ctx, const void* kernel, int block, int n_items) {
  int per_cu = 0;
  LPM_HIP_CHECK(
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, block, 0));
  const int resident = std::max(1, per_cu) * std::max(1, ctx->num_cus);
  return std::max(1, std::min(resident, n_items));
}

}  // namespace lpm_amd
