// Spatially compact slot order of target sets (host side).
//
// The matrix-core and mixed-precision sums decide per wave (64 targets)
// which source tiles are near and take an exact path for those; a wave whose
// targets are spread over the sphere finds nearly every tile near.  Mesh
// vertices come in refinement order (the 12 seed vertices, then each
// level's edge midpoints), so the first vertex waves span the whole sphere.
// Ordering the targets by the Morton (Z-order) key of their position keeps
// each wave's targets together.  Any order gives the same sums up to
// rounding: it only regroups targets into waves.
#include <algorithm>
#include <cmath>
#include <vector>

#include "internal.hpp"
#include "order_keys.hpp"

namespace lpm_amd {

namespace {

uint64_t hash_ranges(const int64_t* off, int n_ranges) {
  uint64_t h = 1469598103934665603ull;  // FNV-1a
  for (int m = 0; off && m <= n_ranges; ++m) {
    h ^= static_cast<uint64_t>(off[m]);
    h *= 1099511628211ull;
  }
  return h ^ static_cast<uint64_t>(n_ranges);
}

}  // namespace

const int32_t* target_order(lpm_ctx* ctx, CrdIn x, int64_t n,
                            const int64_t* off, int n_ranges, hipStream_t s) {
  if (n <= 0) return nullptr;
  LPM_ARG_CHECK(n < (int64_t(1) << 31), "target count out of range");
  const uint64_t h = hash_ranges(off, n_ranges);
  for (auto& e : ctx->orders)
    if (e->x == x.base && e->es == x.es && e->cs == x.cs && e->n == n &&
        e->ranges == h) {
      e->last_use = ++ctx->order_clock;
      return e->perm.as<int32_t>();
    }
  // under a capture the caller did not prepare for, keep the given order
  // (the step functions compute theirs before they capture)
  hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
  LPM_HIP_CHECK(hipStreamIsCapturing(s, &cap));
  if (cap != hipStreamCaptureStatusNone) return nullptr;
  // positions as they are once the stream's earlier work has finished
  LPM_HIP_CHECK(hipStreamSynchronize(s));
  const int64_t span = (n - 1) * x.es + 2 * x.cs + 1;
  std::vector<double> hx(static_cast<size_t>(span));
  LPM_HIP_CHECK(hipMemcpy(hx.data(), x.base, sizeof(double) * span,
                          hipMemcpyDeviceToHost));
  std::vector<std::pair<uint64_t, int32_t>> key(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) {
    const double* p = hx.data() + i * x.es;
    key[i] = {morton_key(p[0], p[x.cs], p[2 * x.cs]), static_cast<int32_t>(i)};
  }
  // sort within each range; ties keep index order
  std::vector<int64_t> bounds;
  if (off && n_ranges > 0) {
    bounds.assign(off, off + n_ranges + 1);
  } else {
    bounds = {0, n};
  }
  for (size_t m = 0; m + 1 < bounds.size(); ++m)
    std::sort(key.begin() + bounds[m], key.begin() + bounds[m + 1]);
  std::vector<int32_t> perm(static_cast<size_t>(n));
  for (int64_t i = 0; i < n; ++i) perm[i] = key[i].second;
  // bounded cache: retire the least recently used order (a graph that
  // recorded it is re-captured: the epoch changes)
  while (ctx->orders.size() >= static_cast<size_t>(kOrderCacheMax)) {
    auto old = ctx->orders.begin();
    for (auto it = ctx->orders.begin(); it != ctx->orders.end(); ++it)
      if ((*it)->last_use < (*old)->last_use) old = it;
    ctx->orders.erase(old);
    g_workspace_epoch.fetch_add(1);
  }
  auto e = std::make_unique<lpm_ctx::TargetOrder>();
  e->x = x.base;
  e->es = x.es;
  e->cs = x.cs;
  e->n = n;
  e->ranges = h;
  e->last_use = ++ctx->order_clock;
  int32_t* d = static_cast<int32_t*>(e->perm.ensure(sizeof(int32_t) * n));
  LPM_HIP_CHECK(hipMemcpy(d, perm.data(), sizeof(int32_t) * n,
                          hipMemcpyHostToDevice));
  ctx->orders.push_back(std::move(e));
  return d;
}

}  // namespace lpm_amd
